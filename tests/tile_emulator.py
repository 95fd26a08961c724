"""numpy emulator of qj_apply_tile_pass, decoding ops and data exactly as
include/qjhip.h documents them; used to check the tile planner/encoder on
the CPU (no GPU)."""

import numpy as np

from qibojit_amd.tiles import TILE_DENSE, TILE_PHASE, TILE_SWAP


def apply_tile_pass(state, nqubits, low, high, ops, data, npdtype):
    state = np.asarray(state)
    tb = low + len(high)
    pos = list(range(low)) + list(high)  # tile-local bit -> global bit
    outside = [b for b in range(nqubits) if b not in pos]
    idx = np.arange(state.size, dtype=np.int64)
    # tile-local coordinate and tile index of every amplitude
    local = np.zeros(state.size, dtype=np.int64)
    for lb, gb in enumerate(pos):
        local |= ((idx >> gb) & 1) << lb
    tindex = np.zeros(state.size, dtype=np.int64)
    for r, gb in enumerate(outside):
        tindex |= ((idx >> gb) & 1) << r
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
        elif op["kind"] == TILE_SWAP:
            a, b = (pos[int(t)] for t in op["target"][:2])
            flip = ((idx >> a) & 1) != ((idx >> b) & 1)
            partner = idx ^ ((1 << a) | (1 << b))
            out = np.where(flip, out[partner], out)
        else:
            assert op["kind"] == TILE_PHASE
            ga, gb, tbh, _ = (int(v) for v in np.frombuffer(data, dtype="<i4", count=4, offset=off))
            assert ga + gb == len(outside) and ga == (len(outside) + 1) // 2 and tbh == tb
            tabs, o = [], off + 16
            for size in (2**ga, 2**gb, tb):
                tabs.append(np.frombuffer(data, dtype=npdtype, count=size, offset=o).astype(np.complex128))
                o += size * itemsize
            f = tabs[0][tindex & (2**ga - 1)] * tabs[1][tindex >> ga]
            for b in range(tb):
                f = np.where((local >> b) & 1, f * tabs[2][b], f)
            out = np.where(ok, out * f, out)
    return out.astype(state.dtype)


def apply_bit_swaps(state, nqubits, pairs):
    """qj_apply_bit_swaps: psi'(i) = psi(i with the bits of each pair exchanged)."""
    idx = np.arange(state.size, dtype=np.int64)
    src = idx.copy()
    for a, b in pairs:
        flip = ((idx >> a) & 1) != ((idx >> b) & 1)
        src = np.where(flip, src ^ ((1 << a) | (1 << b)), src)
    return np.asarray(state)[src]
