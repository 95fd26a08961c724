"""Hardware-efficient ansatz: batched energies and adjoint gradients on gfx950.

A hardware-efficient ansatz (HEA) on n qubits with L layers applies, per
layer, RY(theta) on every qubit, then RZ(theta) on every qubit, then the CNOT
ladder CNOT(0,1), CNOT(1,2), ..., CNOT(n-2,n-1).  Its loss is the energy of
the open-chain transverse-field Ising model
H = -sum_i Z_i Z_{i+1} - g sum_i X_i.

Nothing here has a counterpart in qibojit (it has no batching and no
differentiation); the gates are qibo's RY/RZ/CNOT with qibo's conventions
(RY(t) = exp(-i t Y / 2), RZ(t) = exp(-i t Z / 2)) and qubit q is bit n-1-q
of the state index (src/qibojit/backends/cpu.py:565-569).

``AnsatzEngine`` runs a batch of parameter sets as a few launches per layer
of ``qj_rot_pass``: each pass holds k bits of every state in registers and
applies their rotations, and the ladder is folded into the next pass's loads
(it is the permutation k -> P(k), P^-1(k) = k ^ (k >> 1)).  Gradients use the
adjoint method: lambda = H psi_L, then both states walk back through the
circuit while each rotation adds Im<lambda|G|psi> (FP64) to its gradient.

``plain_adjoint`` is the same method on one circuit with one kernel launch
per gate (HipBackend.apply_gate, vdot) -- the per-circuit baseline.
``AnsatzEnergy`` wraps the engine as a ``torch.autograd.Function``.
"""

import ctypes

import numpy as np
import torch

from qibojit_amd import _native
from qibojit_amd._native import call, i32_array
from qibojit_amd._qibo import Circuit, gates

LOAD_PLAIN, LOAD_LADDER, LOAD_UNLADDER, LOAD_ZERO = 0, 1, 2, 3
GEN_X, GEN_Y, GEN_Z = 1, 2, 3
ROT_MAX_BITS = 5
ROT_MAX_ADJ_BITS = {torch.complex64: 5, torch.complex128: 4}
ROT_BLOCK = 256
MAX_COLS = 65535

_CODES = {torch.complex64: _native.QJ_COMPLEX64, torch.complex128: _native.QJ_COMPLEX128}
_NP = {torch.complex64: np.complex64, torch.complex128: np.complex128}


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def ry_matrices(theta):
    """RY(theta) = [[c, -s], [s, c]], c = cos(theta/2), s = sin(theta/2); shape (..., 2, 2)."""
    theta = np.asarray(theta, dtype=np.float64)
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    m = np.empty(theta.shape + (2, 2), dtype=np.complex128)
    m[..., 0, 0], m[..., 0, 1], m[..., 1, 0], m[..., 1, 1] = c, -s, s, c
    return m


def rz_matrices(theta):
    """RZ(theta) = diag(exp(-i theta/2), exp(i theta/2)); shape (..., 2, 2)."""
    theta = np.asarray(theta, dtype=np.float64)
    m = np.zeros(theta.shape + (2, 2), dtype=np.complex128)
    m[..., 0, 0] = np.exp(-0.5j * theta)
    m[..., 1, 1] = np.exp(0.5j * theta)
    return m


class HardwareEfficientAnsatz:
    """Layout of an HEA: parameter (layer l, kind j, qubit q) sits at
    ``l * 2n + j * n + q`` with j = 0 for RY and 1 for RZ."""

    def __init__(self, nqubits, nlayers):
        if nqubits < 2:
            raise ValueError("The ansatz needs at least two qubits.")
        if nlayers < 1:
            raise ValueError("The ansatz needs at least one layer.")
        self.nqubits = int(nqubits)
        self.nlayers = int(nlayers)
        self.nparams = 2 * self.nqubits * self.nlayers

    def index(self, layer, kind, qubit):
        return layer * 2 * self.nqubits + kind * self.nqubits + qubit

    def random_parameters(self, rng, batch=None):
        shape = (self.nparams,) if batch is None else (batch, self.nparams)
        return rng.uniform(-np.pi, np.pi, size=shape)

    def angles(self, params):
        """(batch, L, 2, n) view of a (batch, P) parameter array."""
        params = np.asarray(params, dtype=np.float64)
        return params.reshape(params.shape[0], self.nlayers, 2, self.nqubits)

    def circuit(self, params):
        """The ansatz as a gate list (qibo gates, or the local stand-ins)."""
        params = np.asarray(params, dtype=np.float64).reshape(self.nlayers, 2, self.nqubits)
        c = Circuit(self.nqubits)
        for layer in range(self.nlayers):
            for q in range(self.nqubits):
                c.add(gates.RY(q, float(params[layer, 0, q])))
            for q in range(self.nqubits):
                c.add(gates.RZ(q, float(params[layer, 1, q])))
            for q in range(self.nqubits - 1):
                c.add(gates.CNOT(q, q + 1))
        return c


def interleaved_groups(nbits, k):
    """Split bit positions 0..nbits-1 into ceil(nbits/k) ascending groups of
    at most k bits, group i = {i, i + G, i + 2G, ...}: every group holds at
    most one of the lowest G bits, so a pass's loads stay line-coalesced."""
    ngroups = -(-nbits // k)
    return [tuple(range(i, nbits, ngroups)) for i in range(ngroups)]


TILE_BITS = 12
TILE_LOW = 4  # at least bits 0..3 in every tile: 16 complex64 = one 128-byte line
TILE_LOW_MAX = 6  # bits 0..5: a wave's load is one 512-byte run


class PassSpec:
    """One pass of a layer: ``kind`` "rot" (register group, bits = rotated
    bits) or "tile" (12 LDS tile bits, of which ``rot`` are rotated)."""

    __slots__ = ("kind", "bits", "rot")

    def __init__(self, kind, bits, rot):
        self.kind, self.bits, self.rot = kind, tuple(bits), tuple(rot)

    def __repr__(self):
        return f"PassSpec({self.kind!r}, {self.bits}, rot={self.rot})"


def register_passes(nbits, k):
    return [PassSpec("rot", g, g) for g in interleaved_groups(nbits, k)]


def tile_passes(nbits, tile=TILE_BITS, low=None):
    """Cover bits 0..nbits-1 with LDS tile passes: the first tile is bits
    0..tile-1, every later one keeps bits 0..low-1 (contiguous loads) and
    rotates up to tile-low new bits, padded with already-rotated bits.
    Without ``low``, the widest low block in [TILE_LOW, TILE_LOW_MAX] that
    needs no extra pass."""
    if nbits < tile:
        raise ValueError(f"Tile passes need at least {tile} qubits.")
    if low is None:
        count = {w: 1 + -(-(nbits - tile) // (tile - w)) for w in range(TILE_LOW, TILE_LOW_MAX + 1)}
        best = min(count.values())
        low = max(w for w, c in count.items() if c == best)
    passes = [PassSpec("tile", range(tile), range(tile))]
    done = tile
    while done < nbits:
        new = list(range(done, min(nbits, done + tile - low)))
        fill = [b for b in range(low, tile)][: tile - low - len(new)]
        passes.append(PassSpec("tile", sorted(list(range(low)) + fill + new), new))
        done += len(new)
    return passes


def _cols(ngroups, limit):
    """Workgroups per state: one group per thread, at most ``limit``."""
    want = max(1, -(-ngroups // ROT_BLOCK))
    return int(min(want, limit, MAX_COLS))


def _tfim_cols(dim, tdtype):
    """Workgroups per state of qj_tfim_batch: one 32-byte group (4 complex64
    / 2 complex128 amplitudes) per thread.  At 24 qubits, batch 16 this full
    grid took 3.2 ms (complex64) / 6.5 ms (complex128) against 4.2 / 9.2 ms
    with 256 grid-stride workgroups per state."""
    return _cols(dim // (4 if tdtype == torch.complex64 else 2), MAX_COLS)


class AnsatzEngine:
    """Batched energies and adjoint gradients of one ansatz shape.

    States of the batch live back to back in two ping-pong buffers (a pass
    that loads through the ladder permutation writes the other buffer);
    gradients add a second pair for lambda.  With ``tiles`` (default for
    complex64 and n >= 12) every forward pass is an LDS tile pass of 12 bits,
    else a register pass of ``k_forward`` bits; ``tiles_adjoint`` (same
    default) does the same for the adjoint passes, else register passes of
    ``k_adjoint`` bits.  At 24 qubits, batch 16: 0.106 s per gradient batch
    with tile adjoint passes, 0.112 s with 4-bit register passes.
    """

    def __init__(self, ansatz, batch, dtype="complex64", g=1.0, k_forward=5, k_adjoint=4, device=None,
                 adjoint_cols=256, tfim_cols=None, stream=None, tiles=None, tiles_adjoint=None):
        self.ansatz = ansatz
        self.batch = int(batch)
        if not 1 <= self.batch <= 65535:
            raise ValueError("batch must be in [1, 65535].")
        self.tdtype = {"complex64": torch.complex64, "complex128": torch.complex128}[str(dtype)]
        self.code = _CODES[self.tdtype]
        self.g = float(g)
        n = ansatz.nqubits
        self.k_forward = int(min(k_forward, ROT_MAX_BITS, n))
        self.k_adjoint = int(min(k_adjoint, ROT_MAX_ADJ_BITS[self.tdtype], n))
        if tiles is None:
            tiles = self.tdtype == torch.complex64 and n >= TILE_BITS
        if tiles_adjoint is None:
            tiles_adjoint = self.tdtype == torch.complex64 and n >= TILE_BITS
        if (tiles or tiles_adjoint) and (self.tdtype != torch.complex64 or n < TILE_BITS):
            raise ValueError("Tile passes need complex64 and at least 12 qubits.")
        self.tiles = bool(tiles)
        self.tiles_adjoint = bool(tiles_adjoint)
        self.device = torch.device("cuda", 0) if device is None else torch.device(device)
        self.stream = stream
        self.dim = 1 << n
        self.passes_f = tile_passes(n) if self.tiles else register_passes(n, self.k_forward)
        self.passes_a = tile_passes(n) if self.tiles_adjoint else register_passes(n, self.k_adjoint)
        self.width_f = max(len(ps.bits) for ps in self.passes_f)
        self.width_a = max(len(ps.bits) for ps in self.passes_a)
        self.cols_f = [self._pass_cols(ps, 1 << 23) for ps in self.passes_f]
        # one column count for every adjoint pass keeps each pass's partials contiguous
        self.cols_a = min(self._pass_cols(ps, adjoint_cols) for ps in self.passes_a)
        self.tfim_cols = _tfim_cols(self.dim, self.tdtype) if tfim_cols is None else _cols(self.dim, tfim_cols)
        self.psi = torch.empty((2, self.batch, self.dim), dtype=self.tdtype, device=self.device)
        self.lam = None
        self.cur = 0
        self._forward_key = None
        self._bits_f = [i32_array(ps.bits) for ps in self.passes_f]
        self._bits_a = [i32_array(ps.bits) for ps in self.passes_a]
        self._gens = [i32_array([g for b in ps.bits for g in ((GEN_Y, GEN_Z) if b in ps.rot else (0, 0))])
                      for ps in self.passes_a]
        self.tfim_partials = torch.zeros((self.batch, self.tfim_cols), dtype=torch.float64, device=self.device)
        self.energies_dev = torch.zeros(self.batch, dtype=torch.float64, device=self.device)

    # kept for callers that inspect the register grouping
    @property
    def groups_f(self):
        return [ps.bits for ps in self.passes_f]

    @property
    def groups_a(self):
        return [ps.bits for ps in self.passes_a]

    # ---------------------------------------------------------- helpers --
    def _pass_cols(self, ps, limit):
        if ps.kind == "tile":
            return int(min(self.dim >> TILE_BITS, limit, MAX_COLS))
        return _cols(self.dim >> len(ps.bits), limit)

    def _stream(self):
        s = self.stream if self.stream is not None else torch.cuda.current_stream(self.device)
        return ctypes.c_void_p(s.cuda_stream)

    def _qubit_of(self, bit):
        return self.ansatz.nqubits - 1 - bit

    def _params(self, params):
        p = params.detach().cpu().numpy() if isinstance(params, torch.Tensor) else np.asarray(params)
        p = np.asarray(p, dtype=np.float64)
        if p.shape != (self.batch, self.ansatz.nparams):
            raise ValueError(f"Expected parameters of shape {(self.batch, self.ansatz.nparams)}, got {p.shape}.")
        return p

    def _forward_mats(self, p):
        """[L, P, B, W, 1, 4]: RZ @ RY per rotated bit at its place in the pass."""
        ang = self.ansatz.angles(p)  # (B, L, 2, n)
        fused = rz_matrices(ang[:, :, 1]) @ ry_matrices(ang[:, :, 0])  # (B, L, n, 2, 2)
        return self._pack(fused[:, :, :, None], self.passes_f, self.width_f)

    def _adjoint_mats(self, p):
        """[L, P, B, W, 2, 4]: step 0 RY, step 1 RZ per rotated bit."""
        ang = self.ansatz.angles(p)
        steps = np.stack([ry_matrices(ang[:, :, 0]), rz_matrices(ang[:, :, 1])], axis=3)  # (B, L, n, 2, 2, 2)
        return self._pack(steps, self.passes_a, self.width_a)

    def _pack(self, per_qubit, passes, width):
        b, nl, _n, nsteps = per_qubit.shape[:4]
        out = np.zeros((nl, len(passes), b, width, nsteps, 4), dtype=_NP[self.tdtype])
        for i, ps in enumerate(passes):
            for a, bit in enumerate(ps.bits):
                if bit in ps.rot:
                    q = self._qubit_of(bit)
                    out[:, i, :, a] = per_qubit[:, :, q].reshape(b, nl, nsteps, 4).transpose(1, 0, 2, 3)
        return torch.from_numpy(out).to(self.device, non_blocking=False)

    def _run_pass(self, ps, src, dst, lsrc, ldst, bits, nrot, gens, load, mats, mstride, partials, ncols):
        bits_ptr, _keep, nbits = bits
        gens_ptr = gens[0] if gens is not None else None
        if ps.kind == "tile":
            rotmask = sum(1 << a for a, b in enumerate(ps.bits) if b in ps.rot)
            call(
                "qj_rot_tile_pass", _vp(src), _vp(dst), _vp(lsrc), _vp(ldst), self.code, self.ansatz.nqubits,
                self.dim, None, self.batch, bits_ptr, rotmask, nrot, gens_ptr, load, _vp(mats), mstride,
                _vp(partials), ncols, self._stream(),
            )
        else:
            call(
                "qj_rot_pass", _vp(src), _vp(dst), _vp(lsrc), _vp(ldst), self.code, self.ansatz.nqubits,
                self.dim, None, self.batch, bits_ptr, nbits, nrot, gens_ptr, load, _vp(mats), mstride,
                _vp(partials), ncols, self._stream(),
            )

    def _reduce(self, partials, out):
        rows = partials.numel() // partials.shape[-1]
        call("qj_reduce_rows", _vp(partials), rows, partials.shape[-1], _vp(out), self._stream())

    # ---------------------------------------------------------- forward --
    def run_forward(self, params):
        """Evolve |0...0> through the ansatz for every parameter set; the
        states are left before the last ladder (``self.psi[self.cur]``)."""
        p = self._params(params)
        mats = self._forward_mats(p)
        psi = self.psi
        cur = 0
        for layer in range(self.ansatz.nlayers):
            for i, ps in enumerate(self.passes_f):
                if layer == 0 and i == 0:
                    src, dst, load = None, psi[cur], LOAD_ZERO
                elif i == 0:
                    src, dst, load = psi[cur], psi[1 - cur], LOAD_LADDER
                    cur = 1 - cur
                else:
                    src, dst, load = psi[cur], psi[cur], LOAD_PLAIN
                self._run_pass(ps, src, dst, None, None, self._bits_f[i], 1, None, load, mats[layer, i],
                               self.width_f * 4, None, self.cols_f[i])
        self.cur = cur
        self._forward_key = p.tobytes()
        self._mats_keepalive = mats
        return p

    def energies(self, params):
        """<H> per parameter set as a float64 device tensor of shape (batch,)."""
        self.run_forward(params)
        self._tfim(None)
        return self.energies_dev.clone()

    def _tfim(self, out):
        call(
            "qj_tfim_batch", _vp(self.psi[self.cur]), _vp(out), self.code, self.ansatz.nqubits, self.dim, None,
            self.batch, self.g, 1, _vp(self.tfim_partials), self.tfim_cols, self._stream(),
        )
        self._reduce(self.tfim_partials, self.energies_dev)

    def final_states(self):
        """States after the last ladder, (batch, 2^n), as a new tensor."""
        k = torch.arange(self.dim, device=self.device, dtype=torch.int64)
        return self.psi[self.cur][:, k ^ (k >> 1)].clone()

    # ---------------------------------------------------------- adjoint --
    def gradients(self, params):
        """(energies (B,), gradients (B, P)) as float64 device tensors.

        Reuses the forward states when ``params`` equals the last forward
        call's parameters."""
        p = self._params(params)
        if self._forward_key != p.tobytes():
            self.run_forward(p)
        if self.lam is None:
            self.lam = torch.empty_like(self.psi)
        mats = self._adjoint_mats(p)
        nl, npass, width = self.ansatz.nlayers, len(self.passes_a), self.width_a
        ncols = self.cols_a
        partials = torch.zeros((nl, npass, width * 2, self.batch, ncols), dtype=torch.float64, device=self.device)
        # lambda = L^+ H L psi on the pre-ladder states
        lcur = 0
        self._tfim(self.lam[lcur])
        psi, lam = self.psi, self.lam
        cur = self.cur
        for layer in reversed(range(nl)):
            for i, ps in enumerate(self.passes_a):
                if i == 0 and layer < nl - 1:
                    load = LOAD_UNLADDER
                    src, dst, lsrc, ldst = psi[cur], psi[1 - cur], lam[lcur], lam[1 - lcur]
                    cur, lcur = 1 - cur, 1 - lcur
                else:
                    load = LOAD_PLAIN
                    src, dst, lsrc, ldst = psi[cur], psi[cur], lam[lcur], lam[lcur]
                self._run_pass(ps, src, dst, lsrc, ldst, self._bits_a[i], 2, self._gens[i], load, mats[layer, i],
                               width * 2 * 4, partials[layer, i], ncols)
        self.cur = cur
        self._forward_key = None  # psi now holds the initial states again
        sums = torch.empty(partials.shape[:-1], dtype=torch.float64, device=self.device)
        self._reduce(partials, sums)
        return self.energies_dev.clone(), self._scatter(sums)

    def _scatter(self, sums):
        """(L, P, W*2, B) slot sums -> (B, P) gradients."""
        n = self.ansatz.nqubits
        index = np.full((self.ansatz.nlayers, len(self.passes_a), self.width_a * 2), -1, dtype=np.int64)
        for layer in range(self.ansatz.nlayers):
            for i, ps in enumerate(self.passes_a):
                for a, bit in enumerate(ps.bits):
                    if bit not in ps.rot:
                        continue
                    q = n - 1 - bit
                    for j in range(2):
                        index[layer, i, a * 2 + j] = self.ansatz.index(layer, j, q)
        flat = index.reshape(-1)
        valid = torch.from_numpy(np.nonzero(flat >= 0)[0]).to(self.device)
        dest = torch.from_numpy(flat[flat >= 0]).to(self.device)
        rows = sums.reshape(-1, self.batch)[valid]  # (P, B)
        grads = torch.empty((self.ansatz.nparams, self.batch), dtype=torch.float64, device=self.device)
        grads[dest] = rows
        return grads.T.contiguous()


class AnsatzEnergy(torch.autograd.Function):
    """energies = AnsatzEnergy.apply(params, engine): (B, P) -> (B,) float64,
    differentiable in ``params`` through the engine's adjoint pass."""

    @staticmethod
    def forward(ctx, params, engine):  # pylint: disable=arguments-differ
        energies = engine.energies(params)
        ctx.engine = engine
        ctx.save_for_backward(params)
        return energies.to(params.device)

    @staticmethod
    def backward(ctx, grad_energies):  # pylint: disable=arguments-differ
        (params,) = ctx.saved_tensors
        _, grads = ctx.engine.gradients(params)
        grads = grads.to(device=grad_energies.device, dtype=grad_energies.dtype)
        return (grad_energies[:, None] * grads).to(params.dtype), None


_GENERATORS = {"RX": ("X", GEN_X), "RY": ("Y", GEN_Y), "RZ": ("Z", GEN_Z)}
_SELF_INVERSE = ("X", "Y", "Z", "H", "CNOT", "CZ", "SWAP", "TOFFOLI")


def plain_adjoint(backend, circuit, g=1.0):
    """Energy and gradients of <psi|H_TFIM|psi> for any circuit of RX/RY/RZ
    and self-inverse gates, one kernel launch per gate on ``backend``.

    Returns (energy, gradients) with one gradient per parametrized gate in
    queue order.  This is the adjoint method without fusion or batching.
    """
    n = circuit.nqubits
    queue = list(circuit.queue)
    for gate in queue:
        name = gate.__class__.__name__
        if name not in _GENERATORS and name not in _SELF_INVERSE:
            raise NotImplementedError(f"plain_adjoint does not handle {name}.")
    psi = backend.zero_state(n)
    for gate in queue:
        psi = backend.apply_gate(gate, psi, n)
    lam = torch.empty_like(psi)
    cols = _tfim_cols(psi.numel(), psi.dtype)
    partials = torch.zeros((1, cols), dtype=torch.float64, device=psi.device)
    energy = torch.empty(1, dtype=torch.float64, device=psi.device)
    stream = backend._stream()  # pylint: disable=protected-access
    call("qj_tfim_batch", _vp(psi), _vp(lam), _CODES[psi.dtype], n, psi.numel(), None, 1, float(g), 0,
         _vp(partials), cols, stream)
    call("qj_reduce_rows", _vp(partials), 1, cols, _vp(energy), stream)
    tmp = torch.empty_like(psi)
    grads = []
    for gate in reversed(queue):
        name = gate.__class__.__name__
        if name in _GENERATORS:
            pauli = getattr(gates, _GENERATORS[name][0])(*gate.target_qubits)
            tmp.copy_(psi)
            tmp = backend.apply_gate(pauli, tmp, n)
            grads.append(backend.vdot(lam, tmp).imag)
            inverse = getattr(gates, name)(*gate.target_qubits, -gate.parameters[0])
            psi = backend.apply_gate(inverse, psi, n)
            lam = backend.apply_gate(inverse, lam, n)
        else:
            psi = backend.apply_gate(gate, psi, n)
            lam = backend.apply_gate(gate, lam, n)
    return float(energy.item()), np.array(grads[::-1], dtype=np.float64)


def ansatz_gradient_order(ansatz):
    """Parameter index of each parametrized gate of ``ansatz.circuit`` in
    queue order (to compare ``plain_adjoint`` with the engine)."""
    order = []
    for layer in range(ansatz.nlayers):
        for j in range(2):
            for q in range(ansatz.nqubits):
                order.append(ansatz.index(layer, j, q))
    return np.array(order, dtype=np.int64)
