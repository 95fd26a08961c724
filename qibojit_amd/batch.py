"""Energies of many hardware-efficient-ansatz circuits of uneven shape in one
captured HIP graph.

Circuits are sorted by qubit count.  Each qubit count n runs on one of two
paths, chosen per n by :func:`autotune` (or given as a plan):

* ``("lds",)``: n <= 13 (complex64) / 12 (complex128).  ``qj_ansatz_lds``,
  a persistent kernel, runs whole circuits inside one workgroup's LDS and
  writes only their energies; one launch serves every LDS-path circuit.
* ``("pass", k)``: the states of all circuits with this n sit back to back
  in two ping-pong buffers and every layer is ceil(n/k) launches of
  ``qj_rot_pass`` over the circuits still running at that layer (the ladder
  is folded into the next layer's first pass), then one ``qj_tfim_batch``
  and ``qj_reduce_rows`` give the FP64 energies.
* ``("tile",)``: as ``pass`` with LDS tile passes (``qj_rot_tile_pass``,
  12 bits each, complex64, n >= 12).

All inputs live in static device buffers, so :meth:`AnsatzBatch.capture`
records the whole sequence once into a ``torch.cuda.CUDAGraph`` (a hipGraph
on ROCm) and :meth:`AnsatzBatch.replay` runs it as one graph launch;
:meth:`AnsatzBatch.set_parameters` refreshes the matrices in place.

``plain_energies`` is the per-circuit baseline: HipBackend.apply_gate per
gate and ``expectation_tfim`` per circuit.
"""

import ctypes
import time

import numpy as np
import torch

from qibojit_amd import _native
from qibojit_amd._native import call, i32_array
from qibojit_amd.ansatz import (
    LOAD_LADDER,
    LOAD_PLAIN,
    LOAD_ZERO,
    ROT_BLOCK,
    TILE_BITS,
    HardwareEfficientAnsatz,
    register_passes,
    ry_matrices,
    rz_matrices,
    tile_passes,
)

_CODES = {torch.complex64: _native.QJ_COMPLEX64, torch.complex128: _native.QJ_COMPLEX128}
_NP = {torch.complex64: np.complex64, torch.complex128: np.complex128}
LDS_MAX_QUBITS = {torch.complex64: 13, torch.complex128: 12}
LDS_MIN_QUBITS = 2


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _fused_layers(n, nlayers, params):
    """(L, n_bits, 4) complex128: RZ @ RY of each layer per bit p = qubit n-1-p."""
    p = np.asarray(params, dtype=np.float64).reshape(nlayers, 2, n)
    m = rz_matrices(p[:, 1]) @ ry_matrices(p[:, 0])  # (L, n_qubits, 2, 2)
    return m[:, ::-1].reshape(nlayers, n, 4)


class _PassGroup:
    """Device buffers and launch list of the pass path for one qubit count:
    register passes of k bits, or (k = None) LDS tile passes."""

    def __init__(self, n, members, layers, k, tdtype, device, tfim_cols=64):
        self.n, self.members, self.k = n, members, k
        self.layers = np.asarray(layers)
        cnt = len(members)
        self.cnt = cnt
        self.dim = 1 << n
        self.lmax = int(self.layers.max())
        self.passes = tile_passes(n) if k is None else register_passes(n, k)
        self.width = max(len(ps.bits) for ps in self.passes)
        self.bits = [i32_array(ps.bits) for ps in self.passes]
        self.rotmask = [sum(1 << a for a, b in enumerate(ps.bits) if b in ps.rot) for ps in self.passes]
        self.psi = torch.empty((2, cnt, self.dim), dtype=tdtype, device=device)
        self.mats = torch.zeros((self.lmax, len(self.passes), cnt, self.width, 1, 4), dtype=tdtype, device=device)
        self.sel = []
        for layer in range(self.lmax):
            active = np.nonzero(self.layers > layer)[0]
            self.sel.append(None if len(active) == cnt else torch.tensor(active, dtype=torch.int32, device=device))
        self.nsel = [int(np.sum(self.layers > layer)) for layer in range(self.lmax)]
        final_buf = (self.layers - 1) % 2
        self.tfim_sel = torch.tensor(final_buf * cnt + np.arange(cnt), dtype=torch.int32, device=device)
        want = -(-self.dim // ROT_BLOCK)
        self.tfim_cols = int(min(tfim_cols, want))
        self.partials = torch.zeros((cnt, self.tfim_cols), dtype=torch.float64, device=device)
        self.cols = [int(min(self.dim >> TILE_BITS, 65535)) if ps.kind == "tile"
                     else int(min(-(-(self.dim >> len(ps.bits)) // ROT_BLOCK), 65535)) for ps in self.passes]

    def pack(self, fused):
        """fused: list over members of (L_c, n, 4) -> device matrices."""
        host = np.zeros(tuple(self.mats.shape), dtype=_NP[self.mats.dtype])
        for s, m in enumerate(fused):
            for i, ps in enumerate(self.passes):
                for a, bit in enumerate(ps.bits):
                    if bit in ps.rot:
                        host[: m.shape[0], i, s, a, 0] = m[:, bit]
        self.mats.copy_(torch.from_numpy(host))

    def launch(self, code, g, stream, energies_out):
        psi = self.psi
        for layer in range(self.lmax):
            sel = self.sel[layer]
            for i, (bits_ptr, _keep, nbits) in enumerate(self.bits):
                if layer == 0 and i == 0:
                    src, dst, load = None, psi[0], LOAD_ZERO
                elif i == 0:
                    src, dst, load = psi[(layer - 1) % 2], psi[layer % 2], LOAD_LADDER
                else:
                    src, dst, load = psi[layer % 2], psi[layer % 2], LOAD_PLAIN
                if self.passes[i].kind == "tile":
                    call("qj_rot_tile_pass", _vp(src), _vp(dst), None, None, code, self.n, self.dim, _vp(sel),
                         self.nsel[layer], bits_ptr, self.rotmask[i], 1, None, load, _vp(self.mats[layer, i]),
                         self.width * 4, None, self.cols[i], stream)
                else:
                    call("qj_rot_pass", _vp(src), _vp(dst), None, None, code, self.n, self.dim, _vp(sel),
                         self.nsel[layer], bits_ptr, nbits, 1, None, load, _vp(self.mats[layer, i]),
                         self.width * 4, None, self.cols[i], stream)
        call("qj_tfim_batch", _vp(psi), None, code, self.n, self.dim, _vp(self.tfim_sel), self.cnt, g, 1,
             _vp(self.partials), self.tfim_cols, stream)
        call("qj_reduce_rows", _vp(self.partials), self.cnt, self.tfim_cols, _vp(energies_out), stream)


class AnsatzBatch:
    """TFIM energies of ansatz circuits ``shapes[c] = (nqubits, nlayers)``."""

    def __init__(self, shapes, dtype="complex64", g=1.0, plan=None, device=None, lds_blocks=None):
        self.shapes = [(int(n), int(nl)) for n, nl in shapes]
        for n, nl in self.shapes:
            if n < 2 or nl < 1:
                raise ValueError(f"Bad circuit shape {(n, nl)}.")
        self.tdtype = {"complex64": torch.complex64, "complex128": torch.complex128}[str(dtype)]
        self.code = _CODES[self.tdtype]
        self.g = float(g)
        self.device = torch.device("cuda", 0) if device is None else torch.device(device)
        ncirc = len(self.shapes)
        nq = np.array([s[0] for s in self.shapes])
        self.order = np.argsort(nq, kind="stable")  # sorted position -> circuit
        self.inv = torch.from_numpy(np.argsort(self.order)).to(self.device)
        self.plan = dict(plan) if plan is not None else default_plan(sorted(set(nq.tolist())), self.tdtype)
        for n in set(nq.tolist()):
            if n not in self.plan:
                raise ValueError(f"No plan for {n} qubits.")
            if self.plan[n][0] == "lds" and not LDS_MIN_QUBITS <= n <= LDS_MAX_QUBITS[self.tdtype]:
                raise ValueError(f"{n} qubits do not fit the LDS path.")
            if self.plan[n][0] == "tile" and (n < TILE_BITS or self.tdtype != torch.complex64):
                raise ValueError(f"The tile path needs complex64 and at least {TILE_BITS} qubits (got {n}).")
        self.energies_sorted = torch.zeros(ncirc, dtype=torch.float64, device=self.device)
        # LDS path: sorted positions 0..nlds-1 (smallest n first)
        sorted_n = nq[self.order]
        lds_mask = np.array([self.plan[int(n)][0] == "lds" for n in sorted_n], dtype=bool)
        self.lds_pos = np.nonzero(lds_mask)[0]
        if len(self.lds_pos) and self.lds_pos[-1] != len(self.lds_pos) - 1:
            # keep LDS circuits first in sorted order so energies land contiguously
            self.order = np.concatenate([self.order[lds_mask], self.order[~lds_mask]])
            self.inv = torch.from_numpy(np.argsort(self.order)).to(self.device)
            sorted_n = nq[self.order]
            self.lds_pos = np.arange(int(lds_mask.sum()))
        self.nlds = len(self.lds_pos)
        if self.nlds:
            lds_circ = self.order[: self.nlds]
            sizes = [self.shapes[c][1] * self.shapes[c][0] * 4 for c in lds_circ]
            offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
            self.lds_offsets = offs
            self.lds_mats = torch.zeros(int(sum(sizes)), dtype=self.tdtype, device=self.device)
            desc = np.zeros((self.nlds, 4), dtype=np.int64)
            for j, c in enumerate(lds_circ):
                desc[j] = (self.shapes[c][0], self.shapes[c][1], offs[j] & 0xFFFFFFFF, offs[j] >> 32)
            self.lds_desc = torch.from_numpy(desc.astype(np.uint32).view(np.int32)).to(self.device)
            self.lds_nqmax = int(max(self.shapes[c][0] for c in lds_circ))
            self.lds_blocks = int(lds_blocks or min(self.nlds, 1024))
        # pass path, one group per qubit count
        self.pass_groups = []
        pos = self.nlds
        for n in sorted({int(v) for v in sorted_n[self.nlds:]}):
            members = [c for c in self.order[self.nlds:] if self.shapes[c][0] == n]
            k = None if self.plan[n][0] == "tile" else self.plan[n][1]
            grp = _PassGroup(n, members, [self.shapes[c][1] for c in members], k, self.tdtype, self.device)
            grp.offset = pos
            pos += len(members)
            self.pass_groups.append(grp)
        self.graph = None

    # -------------------------------------------------------------- data --
    def set_parameters(self, params):
        """params[c]: the 2 * n * L angles of circuit c (original order)."""
        if len(params) != len(self.shapes):
            raise ValueError("One parameter vector per circuit is required.")
        fused = []
        for c, (n, nl) in enumerate(self.shapes):
            p = np.asarray(params[c], dtype=np.float64)
            if p.size != 2 * n * nl:
                raise ValueError(f"Circuit {c} needs {2 * n * nl} parameters, got {p.size}.")
            fused.append(_fused_layers(n, nl, p))
        if self.nlds:
            host = np.concatenate([fused[c].reshape(-1) for c in self.order[: self.nlds]]).astype(_NP[self.tdtype])
            self.lds_mats.copy_(torch.from_numpy(host))
        for grp in self.pass_groups:
            grp.pack([fused[c] for c in grp.members])

    # --------------------------------------------------------- execution --
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def run(self):
        """Enqueue every circuit on the current stream (eager)."""
        stream = self._stream()
        if self.nlds:
            call("qj_ansatz_lds", self.code, _vp(self.lds_desc), self.nlds, _vp(self.lds_mats), self.lds_mats.numel(),
                 self.g, _vp(self.energies_sorted), self.lds_nqmax, self.lds_blocks, stream)
        for grp in self.pass_groups:
            out = self.energies_sorted[grp.offset: grp.offset + grp.cnt]
            grp.launch(self.code, self.g, stream, out)

    def capture(self):
        """Record ``run`` into one graph (after an eager warm-up)."""
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            self.run()
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.run()
        return self.graph

    def replay(self):
        if self.graph is None:
            raise RuntimeError("capture() first.")
        self.graph.replay()

    def energies(self):
        """FP64 energies in the original circuit order (device tensor)."""
        return self.energies_sorted[self.inv]


def default_plan(qubit_counts, tdtype=torch.complex64, k=5):
    lds_max = LDS_MAX_QUBITS[tdtype] if isinstance(tdtype, torch.dtype) else LDS_MAX_QUBITS[torch.complex64]
    return {int(n): ("lds",) if n <= lds_max else ("pass", min(k, int(n))) for n in qubit_counts}


def autotune(shapes, params, dtype="complex64", g=1.0, reps=3, ks=(3, 4, 5)):
    """Per qubit count, time each feasible path on that count's circuits and
    keep the fastest.  Returns (plan, timings) with timings[n][choice] in s."""
    tdtype = {"complex64": torch.complex64, "complex128": torch.complex128}[str(dtype)]
    by_n = {}
    for c, (n, nl) in enumerate(shapes):
        by_n.setdefault(int(n), []).append(c)
    plan, timings = {}, {}
    for n, members in sorted(by_n.items()):
        sub_shapes = [shapes[c] for c in members]
        sub_params = [params[c] for c in members]
        choices = [("pass", k) for k in ks if k <= n]
        if LDS_MIN_QUBITS <= n <= LDS_MAX_QUBITS[tdtype]:
            choices.insert(0, ("lds",))
        if n >= TILE_BITS and tdtype == torch.complex64:
            choices.append(("tile",))
        timings[n] = {}
        for choice in choices:
            b = AnsatzBatch(sub_shapes, dtype=dtype, g=g, plan={n: choice})
            b.set_parameters(sub_params)
            b.run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                b.run()
            torch.cuda.synchronize()
            timings[n][choice] = (time.perf_counter() - t0) / reps
            del b
        plan[n] = min(timings[n], key=timings[n].get)
    torch.cuda.empty_cache()
    return plan, timings


def plain_energies(backend, shapes, params, g=1.0):
    """One kernel launch per gate and one expectation per circuit."""
    out = np.empty(len(shapes))
    for c, (n, nl) in enumerate(shapes):
        circuit = HardwareEfficientAnsatz(n, nl).circuit(params[c])
        state = backend.zero_state(n)
        for gate in circuit.queue:
            state = backend.apply_gate(gate, state, n)
        out[c] = backend.expectation_tfim(state, n, g)
    return out
