"""numpy emulator of qj_apply_tile_pass, decoding ops and data exactly as
include/qjhip.h documents them; used to check the tile planner/encoder on
the CPU (no GPU)."""

import numpy as np

from qibojit_amd.tiles import TILE_DENSE, TILE_PHASE


def apply_tile_pass(state, nqubits, low, high, ops, data, npdtype):
    state = np.asarray(state)
    tb = low + len(high)
    la = tb // 2
    pos = list(range(low)) + list(high)  # tile-local bit -> global bit
    idx = np.arange(state.size, dtype=np.int64)
    # tile-local coordinate of every amplitude
    local = np.zeros(state.size, dtype=np.int64)
    for lb, gb in enumerate(pos):
        local |= ((idx >> gb) & 1) << lb
    itemsize = np.dtype(npdtype).itemsize
    out = state.astype(np.complex128).copy()
    for op in ops:
        gmask = int(op["global_ctrl"])
        lmask = int(op["local_ctrl"])
        ok = ((idx & gmask) == gmask) & ((local & lmask) == lmask)
        off = int(op["offset"])
        if op["kind"] == TILE_DENSE:
            k = int(op["k"])
            n = 2**k
            m = np.frombuffer(data, dtype=npdtype, count=n * n, offset=off).reshape(n, n).astype(np.complex128)
            tg = [pos[int(t)] for t in op["target"][:k]]
            tmask = sum(1 << b for b in tg)
            new = out.copy()
            for i in np.nonzero(ok & ((idx & tmask) == 0))[0]:
                rows = [i | sum(((r >> (k - 1 - u)) & 1) << tg[u] for u in range(k)) for r in range(n)]
                new[rows] = m @ out[rows]
            out = new
        else:
            assert op["kind"] == TILE_PHASE
            c0 = np.frombuffer(data, dtype="<f8", count=1, offset=off)[0]
            nout = int(np.frombuffer(data, dtype="<i4", count=1, offset=off + 8)[0])
            terms = np.frombuffer(data, dtype=[("bit", "<i4"), ("pad", "<i4"), ("theta", "<f8")], count=nout, offset=off + 16)
            toff = off + 16 + 16 * nout
            ta = np.frombuffer(data, dtype=npdtype, count=2**la, offset=toff).astype(np.complex128)
            tbl = np.frombuffer(data, dtype=npdtype, count=2 ** (tb - la), offset=toff + itemsize * 2**la).astype(np.complex128)
            ang = np.full(state.size, c0)
            for t in terms:
                ang += t["theta"] * ((idx >> int(t["bit"])) & 1)
            f = np.exp(1j * ang) * ta[local & (2**la - 1)] * tbl[local >> la]
            out = np.where(ok, out * f, out)
    return out.astype(state.dtype)
