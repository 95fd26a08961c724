"""Benchmark: effective HBM bandwidth of the 30-qubit gate sweep on one MI355X.

One step applies the sweep to a 30-qubit complex128 state of seeded complex
Gaussian amplitudes resident in HBM: H, X, Z, RZ, CNOT, SWAP, fSim, CU3 and a
seeded 3-qubit Haar unitary, each on target qubits 0, 15 and 29 (bit
strides 2^29, 2^14, 1) -- 27 kernel launches.  `value` is the sweep's
algorithmic bytes divided by the step time (GB/s).

Algorithmic bytes of a gate: 2 x 128 B x (number of 128-byte lines of the
state holding an amplitude the gate can change) -- read + write of every
touched line, counted from the gate's index set, not measured.

Also reported (not `value`): the same sweep in complex64, per-gate GB/s, the
dominant kernel's roofline, the 20-qubit depth-20 random circuit (seconds),
QFT on 32 qubits complex64 (seconds), collapse /
probabilities / 10^6-shot sampling on qubits {0, 7, 29}, a 33-qubit
complex128 capacity run, and the CPU baseline: the C/OpenMP restatement of
the reference's numba kernels on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras] [--no-cpu]
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# HBM bytes per launch of the dominant kernel, measured with rocprofv3
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (scale from a plain 4 GiB copy):
# profiles/sweep_30q_c128_hbm_traffic.md.  Keyed by (kernel, nqubits).
PROFILED_TRAFFIC = {("k_wave<double,1,1>", 30): 34.360e9, ("k_one_qubit<double,0,1,256>", 30): 28.633e9}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--nqubits", type=int, default=30)
    p.add_argument("--no-extras", action="store_true")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--only-sweep", action="store_true", help="sweep only (for profiler runs)")
    p.add_argument("--copy-probe", action="store_true", help="plain 4 GiB copy (counter scale reference)")
    p.add_argument("--only-adjoint", action="store_true", help="adjoint-gradient benchmark only")
    p.add_argument("--only-batch", action="store_true", help="many-circuit ansatz benchmark only")
    p.add_argument("--only-precision", action="store_true", help="complex64 vs complex128 runs and the MFMA kernel")
    p.add_argument("--only-clifford", action="store_true", help="stabilizer-tableau benchmark only")
    p.add_argument("--only-mfma", action="store_true", help="the 5-qubit block kernels alone (for counter runs)")
    p.add_argument("--only-measure", action="store_true", help="probabilities, sampling and collapse only")
    p.add_argument("--only-qft", action="store_true", help="32-qubit QFT through the tile executor only")
    p.add_argument("--only-random", action="store_true", help="20-qubit depth-20 random circuit only")
    p.add_argument("--blocks", default="low,mid,high,low_ctrl",
                   help="placements of --only-mfma (comma-separated: low, mid, high, low_ctrl)")
    p.add_argument("--block-paths", default="mfma,valu", help="kernels of --only-mfma: mfma, valu or both")
    return p.parse_args()


# ------------------------------------------------------------ sweep spec --
def sweep_gates(n, rng):
    """(label, gate, selecting qubits, predicate on their bits) for the sweep."""
    from qibojit_amd._qibo import gates

    haar = []
    for _ in range(3):
        z = (rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))) / np.sqrt(2)
        q, r = np.linalg.qr(z)
        haar.append(q * (np.diag(r) / np.abs(np.diag(r))))
    out = []
    theta = 0.731
    for i, t in enumerate((0, n // 2, n - 1)):
        p2 = (t, t + 1) if t < n - 1 else (t - 1, t)
        ctrl = t + 1 if t < n - 1 else t - 1
        h3 = (t, t + 1, t + 2) if t + 2 < n else (n - 3, n - 2, n - 1)
        always = lambda bits: True  # noqa: E731
        out += [
            (f"H({t})", gates.H(t), (t,), always),
            (f"X({t})", gates.X(t), (t,), always),
            (f"Z({t})", gates.Z(t), (t,), lambda b, t=t: b[t] == 1),
            (f"RZ({t})", gates.RZ(t, theta), (t,), always),
            (f"CNOT({ctrl},{t})", gates.CNOT(ctrl, t), (ctrl, t), lambda b, c=ctrl: b[c] == 1),
            (f"SWAP{p2}", gates.SWAP(*p2), p2, lambda b, p=p2: b[p[0]] != b[p[1]]),
            (f"fSim{p2}", gates.fSim(*p2, 0.3, 0.4), p2, lambda b, p=p2: b[p[0]] == 1 or b[p[1]] == 1),
            (f"CU3({ctrl},{t})", gates.CU3(ctrl, t, 0.1, 0.2, 0.3), (ctrl, t), lambda b, c=ctrl: b[c] == 1),
            (f"Haar3{h3}", gates.Unitary(haar[i], *h3), h3, always),
        ]
    return out


def changed_fraction(n, qubits, pred, amp_bytes):
    """(line fraction, element fraction) of amplitudes the gate can change."""
    line_bits = int(math.log2(128 // amp_bytes))
    bits = {q: n - 1 - q for q in qubits}
    high = [q for q in qubits if bits[q] >= line_bits]
    low = [q for q in qubits if bits[q] < line_bits]
    lines = 0
    for hv in range(2 ** len(high)):
        assign = {q: (hv >> i) & 1 for i, q in enumerate(high)}
        for lv in range(2 ** len(low)):
            assign.update({q: (lv >> i) & 1 for i, q in enumerate(low)})
            if pred(assign):
                lines += 1
                break
    elems = 0
    for v in range(2 ** len(qubits)):
        if pred({q: (v >> i) & 1 for i, q in enumerate(qubits)}):
            elems += 1
    return lines / 2 ** len(high), elems / 2 ** len(qubits)


def kernel_name(backend, gate, n, dtype):
    """Kernel template instance the dispatch selects (for grouping launches)."""
    from qibojit_amd.backends.hip import GATE_OPS

    r = "float" if dtype == "complex64" else "double"
    qubits = backend._create_qubits_tensor(gate, n)  # pylint: disable=protected-access
    v = 2 if dtype == "complex64" and int(qubits.min()) >= 1 else 1
    k = len(gate.target_qubits)
    name = gate.__class__.__name__
    if k <= 2 and n >= 6 and all(n - 1 - t < 6 for t in gate.target_qubits):
        ilp = os.environ.get("QJ_WAVE_ILP", "")
        if ilp not in ("1", "2", "4"):
            ilp = "2" if dtype == "complex64" else "1"
        if k == 2 and dtype == "complex128" and GATE_OPS.get(name) in ("apply_swap", "apply_fsim"):
            return f"k_wave_pair<{r},{ilp}>"  # matrix coupling r and r ^ 3 only: one shuffle
        return f"k_wave<{r},{k},{ilp}>"  # targets inside a 64-amplitude chunk: in-wave exchange
    if k == 1:
        kind = {"apply_x": 1, "apply_y": 2, "apply_z": 3, "apply_z_pow": 4}.get(GATE_OPS.get(name), 0)
        m = n - 1 - gate.target_qubits[0]
        if v == 2 and kind in (0, 1, 2) and (m >= 22 or m <= 10):
            v = 1  # one complex64 amplitude per lane at these target bits (qj_gates.hip)
        nt = 128 if dtype == "complex64" and v == 1 and m >= 22 else 256  # workgroup size
        return f"k_one_qubit<{r},{kind},{v},{nt}>"
    if k == 2:
        kind = {"apply_swap": 1, "apply_fsim": 2}.get(GATE_OPS.get(name), 0)
        return f"k_two_qubit<{r},{kind},{v},256>"
    if k == 3 and len(gate.control_qubits) == 0 and n >= 11 and all(n - 1 - t < 3 for t in gate.target_qubits):
        chunk = 1024 if dtype == "complex128" else 2048
        return f"k_multi_low3<{r},{chunk}>"  # the three lowest bits: LDS-staged chunks
    if k == 3:
        m = min(n - 1 - t for t in gate.target_qubits)
        if dtype == "complex128":
            nt = 256 if 9 <= m <= 21 else 128
        else:
            nt = 128 if m >= 9 else 256
        return f"k_multi_reg<{r},3,{nt}>"
    return f"k_multi_reg<{r},{k}>"


# ----------------------------------------------------------------- sweep --
def run_sweep(backend, n, dtype, steps, warmup, seed=1234):
    import torch

    rng = np.random.default_rng(seed)
    spec = sweep_gates(n, rng)
    amp = 8 if dtype == "complex64" else 16
    state_bytes = amp * 2**n
    gen = torch.Generator(device="cuda").manual_seed(seed)
    tdt = torch.complex64 if dtype == "complex64" else torch.complex128
    state = torch.randn(2**n, dtype=tdt, device="cuda", generator=gen)
    state /= torch.linalg.vector_norm(state)
    backend.set_dtype(dtype)
    torch.cuda.synchronize()

    for _ in range(warmup):
        for _, g, _, _ in spec:
            state = backend.apply_gate(g, state, n)
    torch.cuda.synchronize()

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps * (len(spec) + 1))]
    t0 = time.perf_counter()
    k = 0
    for _ in range(steps):
        ev[k].record()
        k += 1
        for _, g, _, _ in spec:
            state = backend.apply_gate(g, state, n)
            ev[k].record()
            k += 1
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    per_gate = {lbl: [] for lbl, _, _, _ in spec}
    step_ms = []
    k = 0
    for _ in range(steps):
        start = k
        for lbl, _, _, _ in spec:
            per_gate[lbl].append(ev[k].elapsed_time(ev[k + 1]))
            k += 1
        step_ms.append(ev[start].elapsed_time(ev[k]))
        k += 1

    rows = []
    total_bytes = 0.0
    for lbl, g, qs, pred in spec:
        fl, fe = changed_fraction(n, qs, pred, amp)
        b = 2 * state_bytes * fl
        total_bytes += b
        ms = float(np.median(per_gate[lbl]))
        rows.append({
            "gate": lbl,
            "kernel": kernel_name(backend, g, n, dtype),
            "bytes": b,
            "elem_bytes": 2 * state_bytes * fe,
            "ms_median": ms,
            "ms_mean": float(np.mean(per_gate[lbl])),
            "GBps": b / (ms * 1e-3) / 1e9,
        })
    ms_per_step = float(np.mean(step_ms))
    norm = backend.norm_squared(state)
    del state
    torch.cuda.empty_cache()
    return {
        "rows": rows,
        "bytes_per_step": total_bytes,
        "ms_per_step": ms_per_step,
        "GBps": total_bytes / (ms_per_step * 1e-3) / 1e9,
        "wall_s": wall,
        "final_norm2": norm,
    }


def dominant_kernel(rows):
    by = {}
    for r in rows:
        d = by.setdefault(r["kernel"], {"ms": 0.0, "bytes": 0.0, "launches": 0})
        d["ms"] += r["ms_mean"]
        d["bytes"] += r["bytes"]
        d["launches"] += 1
    name = max(by, key=lambda k: by[k]["ms"])
    d = by[name]
    avg_ms = d["ms"] / d["launches"]
    avg_bytes = d["bytes"] / d["launches"]
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
    return name, avg_ms, avg_bytes, achieved, d["launches"]


# ---------------------------------------------------------------- extras --
def run_qft(backend, n, dtype, fused=False, reps=1):
    """QFT (H + controlled-phase ladder + final swaps) from a seeded state,
    gate by gate or through the fused tile executor.  With reps > 1 the
    circuit runs that many times on the evolving state (every run moves the
    same bytes) and `seconds` is the median run, host planning included."""
    import torch

    from qibojit_amd._qibo import Circuit, gates

    c = Circuit(n)
    for i in range(n):
        c.add(gates.H(i))
        for j in range(i + 1, n):
            c.add(gates.CU1(j, i, np.pi / 2 ** (j - i)))
    for i in range(n // 2):
        c.add(gates.SWAP(i, n - i - 1))
    backend.set_dtype(dtype)
    tdt = torch.complex64 if dtype == "complex64" else torch.complex128
    gen = torch.Generator(device="cuda").manual_seed(7)
    state = torch.randn(2**n, dtype=tdt, device="cuda", generator=gen)
    state /= torch.linalg.vector_norm(state)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if fused:
        from qibojit_amd import tiles

        low, max_high = backend.tile_geometry(n, tdt)
        ops = [tiles.gate_to_op(backend._as_custom_matrix, g, n) for g in c.queue]  # pylint: disable=protected-access
        npasses = len(tiles.plan(ops, n, low, max_high))
    runs = []
    for _ in range(reps):
        e0.record()
        if fused:
            state = backend.apply_gates_fused(c.queue, state, n)
        else:
            for g in c.queue:
                state = backend.apply_gate(g, state, n)
        e1.record()
        torch.cuda.synchronize()
        runs.append(e0.elapsed_time(e1) / 1e3)
    sec = float(np.median(runs))
    amp = 8 if dtype == "complex64" else 16
    # algorithmic bytes: H full pass; CU1 touches control=1 & target=1 (1/4);
    # SWAP touches bit1 != bit2 (1/2); line rounding for low bits.
    total = 0.0
    for g in c.queue:
        name = g.__class__.__name__
        if name == "H":
            fl = 1.0
        elif name == "CU1":
            fl, _ = changed_fraction(n, g.qubits, lambda b, q=g.qubits: b[q[0]] == 1 and b[q[1]] == 1, amp)
        else:
            fl, _ = changed_fraction(n, g.qubits, lambda b, q=g.qubits: b[q[0]] != b[q[1]], amp)
        total += 2 * amp * 2**n * fl
    norm = backend.norm_squared(state)
    del state
    torch.cuda.empty_cache()
    out = {"nqubits": n, "dtype": dtype, "gates": len(c.queue), "seconds": sec, "runs": runs, "final_norm2": norm}
    if fused:
        # one read + one write of the state per pass
        passes_bytes = npasses * 2 * amp * 2**n
        out.update({"fused": True, "passes": npasses, "pass_GB": passes_bytes / 1e9,
                    "GBps_passes": passes_bytes / sec / 1e9})
    else:
        out.update({"fused": False, "algorithmic_GB": total / 1e9, "GBps": total / sec / 1e9})
    return out


MFMA_FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix
MFMA_FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix


def brickwork(n, depth, seed):
    """Depth-`depth` brickwork of seeded 2-qubit Haar unitaries: layer d acts
    on pairs (q, q+1) with q = d mod 2, d mod 2 + 2, ..."""
    from qibojit_amd._qibo import Circuit, gates

    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for d in range(depth):
        for q in range(d % 2, n - 1, 2):
            z = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) / np.sqrt(2)
            qm, r = np.linalg.qr(z)
            c.add(gates.Unitary(qm * (np.diag(r) / np.abs(np.diag(r))), q, q + 1))
    return c


def random_circuit_20q(n=20, depth=20, seed=1234):
    """Depth-`depth` layers of U3 (angles U[0, 2pi)) on every qubit, then CZ
    on the pairs of a seeded random permutation redrawn per layer."""
    from qibojit_amd._qibo import Circuit, gates

    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for _ in range(depth):
        for q in range(n):
            c.add(gates.U3(q, *rng.uniform(0, 2 * np.pi, 3)))
        perm = rng.permutation(n)
        for j in range(n // 2):
            c.add(gates.CZ(int(perm[2 * j]), int(perm[2 * j + 1])))
    return c


def run_random_circuit(backend, n=20, depth=20, reps=5, cpu=True):
    """The 20-qubit complex128 random circuit from |0...0>: the state is
    16.8 MB and stays in the L2/MALL caches, so this is a time, not a
    bandwidth, figure.  Times execute_circuit with the fused executor
    (host planning included) and the gate-by-gate apply_gate loop (median of
    `reps` runs each, wall clock after a device synchronisation), and the
    C/OpenMP restatement of the numba kernels on the host cores once."""
    import torch

    c = random_circuit_20q(n, depth)
    backend.set_dtype("complex128")
    out = {"nqubits": n, "depth": depth, "gates": len(c.queue)}

    def fused():
        return backend.execute_circuit(c).state()

    def plain():
        st = backend.zero_state(n)
        for g in c.queue:
            st = backend.apply_gate(g, st, n)
        return st

    results = {}
    for label, fn in (("fused", fused), ("gate_by_gate", plain)):
        fn()  # first call builds plans and warms the allocator
        runs = []
        for _ in range(reps):
            st, sec = _timed(fn)
            runs.append(sec)
        results[label] = st
        med = float(np.median(runs))
        out[f"{label}_s"] = med
        out[f"{label}_gates_per_s"] = len(c.queue) / med
    got_f = backend.to_numpy(results["fused"])
    got_p = backend.to_numpy(results["gate_by_gate"])
    out["max_abs_diff_fused_vs_gate_by_gate"] = float(np.abs(got_f - got_p).max())
    if cpu:
        from oracle import port

        threads = len(os.sched_getaffinity(0))
        pb = port.PortBackend(nthreads=threads)
        t = time.perf_counter()
        want = pb.execute_circuit(c)
        cpu_s = time.perf_counter() - t
        out.update({"cpu_port_s": cpu_s, "cpu_port_threads": threads,
                    "max_abs_err_vs_cpu_port": float(np.abs(got_f - want).max()),
                    "speedup_fused_vs_cpu_port": cpu_s / out["fused_s"]})
    del results
    torch.cuda.empty_cache()
    return out


def _timed(fn):
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def run_block_kernels(backend, launches=5, seed=11, placements=("low", "mid", "high", "low_ctrl"),
                      paths=(True, False)):
    """Dense 4- and 5-target blocks on the 32-qubit complex64 and 30-qubit
    complex128 states (34.36 / 17.18 GB), matrix-core 3M kernel against the
    VALU register kernel, targets on the lowest, middle and highest qubit
    bits, plus the low block controlled by qubit 0.  Algorithmic bytes per
    launch: one read and one write of the state (of its 2^-c share when
    controlled); flops: 8 NR^2 per column (the complex product), of which the
    3M form executes 6 NR^2 on the matrix cores."""
    import torch

    from qibojit_amd._qibo import gates

    rng = np.random.default_rng(seed)
    res = {}
    for dtype, n in (("complex64", 32), ("complex128", 30)):
        backend.set_dtype(dtype)
        tdt = torch.complex64 if dtype == "complex64" else torch.complex128
        amp = 8 if dtype == "complex64" else 16
        peak = MFMA_FP32_PEAK_TFLOPS if dtype == "complex64" else MFMA_FP64_PEAK_TFLOPS
        state = torch.full((2**n,), 2 ** (-n / 2), dtype=tdt, device="cuda")
        for k in (4, 5):
            nr = 2**k
            z = (rng.standard_normal((nr, nr)) + 1j * rng.standard_normal((nr, nr))) / np.sqrt(2)
            u, _ = np.linalg.qr(z)
            placements_all = (("low", tuple(range(n - k, n)), ()), ("mid", tuple(range(10, 10 + k)), ()),
                              ("high", tuple(range(k)), ()), ("low_ctrl", tuple(range(n - k, n)), (0,)))
            for label, targets, controls in [p for p in placements_all if p[0] in placements]:
                gate = gates.Unitary(u, *targets)
                if controls:
                    gate = gate.controlled_by(*controls)
                share = 2.0 ** -len(controls)
                for mfma in paths:
                    backend.mfma = mfma
                    backend.apply_gate(gate, state, n)  # warm-up
                    ev = [torch.cuda.Event(enable_timing=True) for _ in range(launches + 1)]
                    ev[0].record()
                    for j in range(launches):
                        backend.apply_gate(gate, state, n)
                        ev[j + 1].record()
                    torch.cuda.synchronize()
                    ms = float(np.mean([ev[j].elapsed_time(ev[j + 1]) for j in range(launches)]))
                    nbytes = 2 * amp * 2**n * share
                    cols = 2 ** (n - k) * share
                    res[f"{dtype}_k{k}_{label}_{'mfma' if mfma else 'valu'}"] = {
                        "nqubits": n, "targets": list(targets), "controls": list(controls), "ms": round(ms, 4),
                        "GBps": round(nbytes / ms / 1e6, 1), "hbm_frac": round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4),
                        "mfma_util": round(6 * nr * nr * cols / ms / 1e9 / peak, 4) if mfma else None,
                    }
        backend.mfma = True
        del state
        torch.cuda.empty_cache()
    backend.set_dtype("complex128")
    return res


def run_precision(backend, n=32, depth=40, seed=5):
    """complex64 results against complex128 runs of the same circuits on the
    GPU: (a) QFT from a seeded state (fused tile executor in both
    precisions), (b) depth-`depth` brickwork of 2-qubit Haar unitaries from
    |0...0>: complex64 fused into 5-qubit blocks on matrix cores, complex128
    gate by gate."""
    import torch

    from qibojit_amd.fusion import fuse_blocks

    out = {}
    # (a) QFT
    from qibojit_amd._qibo import Circuit, gates

    qft = Circuit(n)
    for i in range(n):
        qft.add(gates.H(i))
        for j in range(i + 1, n):
            qft.add(gates.CU1(j, i, np.pi / 2 ** (j - i)))
    for i in range(n // 2):
        qft.add(gates.SWAP(i, n - i - 1))
    gen = torch.Generator(device="cuda").manual_seed(7)
    ref = torch.randn(2**n, dtype=torch.complex128, device="cuda", generator=gen)
    ref /= torch.linalg.vector_norm(ref)
    st = ref.to(torch.complex64)
    backend.set_dtype("complex128")
    ref, t128 = _timed(lambda: backend.apply_gates_fused(qft.queue, ref, n))
    backend.set_dtype("complex64")
    st, t64 = _timed(lambda: backend.apply_gates_fused(qft.queue, st, n))
    cmp = backend.compare_states(st, ref)
    out["qft"] = {"gates": len(qft.queue), "seconds_c64_fused": t64, "seconds_c128_fused": t128,
                  "infidelity": cmp["infidelity"], "max_abs_error": cmp["max_abs_diff"],
                  "norm2_c64": cmp["norm2_a"]}
    del st, ref
    torch.cuda.empty_cache()
    # (b) brickwork
    circ = brickwork(n, depth, seed)
    fused = fuse_blocks(circ, 5)
    sizes = {}
    for g in fused.queue:
        sizes[len(g.qubits)] = sizes.get(len(g.qubits), 0) + 1
    backend.set_dtype("complex64")

    def run(queue):
        # the state is allocated and initialised outside the timed region
        s = backend.zero_state(n)
        torch.cuda.synchronize()

        def apply():
            out = s
            for g in queue:
                out = backend.apply_gate(g, out, n)
            return out

        return _timed(apply)

    st, t_mfma = run(fused.queue)
    backend.set_dtype("complex128")
    ref, t_ref = run(circ.queue)
    cmp = backend.compare_states(st, ref)
    del st
    torch.cuda.empty_cache()
    # complex128 fused blocks on FP64 matrix cores against the gate-by-gate run
    st128, t_mfma128 = run(fused.queue)
    cmp128 = backend.compare_states(st128, ref)
    del st128
    torch.cuda.empty_cache()
    backend.set_dtype("complex64")
    backend.mfma = False
    st_valu, t_valu = run(fused.queue)
    backend.mfma = True
    del st_valu
    torch.cuda.empty_cache()
    out["brickwork"] = {
        "nqubits": n, "depth": depth, "gates": len(circ.queue), "fused_blocks": len(fused.queue),
        "block_sizes": {str(k): v for k, v in sorted(sizes.items())},
        "seconds_c64_fused_mfma": t_mfma, "seconds_c64_fused_valu": t_valu, "seconds_c128_unfused": t_ref,
        "seconds_c128_fused_mfma": t_mfma128, "c128_fused_max_abs_error": cmp128["max_abs_diff"],
        "infidelity": cmp["infidelity"], "max_abs_error": cmp["max_abs_diff"], "norm2_c64": cmp["norm2_a"],
    }
    del ref
    torch.cuda.empty_cache()
    return out


def _event_ms(fn, reps):
    """Median milliseconds of fn() over `reps` calls, each bracketed by HIP
    events on the current stream (the stream the library launches on)."""
    import torch

    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    return float(np.median(times))


def run_measurement_ops(backend, n, dtype, reps=5):
    """Marginal probabilities, 10^6-shot Metropolis sampling and a normalised
    collapse on qubits {0, 7, n-1}.  Collapse is timed on the collapsed state
    again and again: every repetition zeroes the same 7/8 of the state and
    rescales the same kept 1/8, so each moves the same bytes."""
    import torch

    backend.set_dtype(dtype)
    tdt = torch.complex64 if dtype == "complex64" else torch.complex128
    gen = torch.Generator(device="cuda").manual_seed(11)
    state = torch.randn(2**n, dtype=tdt, device="cuda", generator=gen)
    state /= torch.linalg.vector_norm(state)
    qubits = [0, 7, n - 1]
    amp = 8 if dtype == "complex64" else 16
    out = {}
    probs = backend.calculate_probabilities(state, qubits, n)
    ms = _event_ms(lambda: backend.calculate_probabilities(state, qubits, n), reps)
    out["probabilities"] = {"ms": ms, "GBps": amp * 2**n / ms / 1e6}
    torch.cuda.synchronize()
    t = time.perf_counter()
    freqs = backend.measure_frequencies_op(
        torch.zeros(8, dtype=torch.int64, device="cuda"), probs, 10**6, 3, 1234, backend.nthreads
    )
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    out["metropolis_1e6_shots"] = {"ms": dt * 1e3, "chains": backend.nthreads, "total": int(freqs.sum())}
    backend.collapse_state(state, qubits, 5, n)
    ms = _event_ms(lambda: backend.collapse_state(state, qubits, 5, n), reps)
    # 1.125 x state: 7/8 written as zeros, the kept 1/8 read and rewritten.
    # Line level (qubit n-1 is bit 0, so a kept 128-byte line holds kept and
    # dropped amplitudes): 3/4 of the lines written, 1/4 read twice and
    # written once = 1.5 x state.
    out["collapse_normalized"] = {"ms": ms, "GBps_at_1.125x": amp * 2**n * 1.125 / ms / 1e6,
                                  "GBps_line_level_1.5x": amp * 2**n * 1.5 / ms / 1e6}
    del state
    torch.cuda.empty_cache()
    return out


def run_capacity(backend, n=33, dtype="complex128"):
    """One dense 1-qubit gate on an n-qubit state held whole on the device
    (137 GB at 33 qubits complex128 or 34 qubits complex64)."""
    import torch

    from qibojit_amd._qibo import gates

    backend.set_dtype(dtype)
    amp = 8 if dtype == "complex64" else 16
    torch.cuda.empty_cache()
    try:
        state = backend.zero_state(n)  # values do not affect the timing
    except MemoryError as exc:
        return {"nqubits": n, "dtype": dtype, "error": str(exc)[:200]}
    g = gates.H(n // 2)
    state = backend.apply_gate(g, state, n)  # warm
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    state = backend.apply_gate(g, state, n)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    b = 2 * amp * 2**n
    norm = backend.norm_squared(state)
    del state
    torch.cuda.empty_cache()
    return {"nqubits": n, "dtype": dtype, "state_GB": amp * 2**n / 1e9, "gate": f"H({n // 2})", "ms": ms,
            "GBps": b / ms * 1e3 / 1e9, "final_norm2": norm}


def _shifted_energies(ansatz, base, shifts, chunk=64):
    """Energies (complex128 engine) of ``base`` with each parameter moved by
    each of ``shifts``: returns array (len(shifts), P)."""
    import torch

    from qibojit_amd.ansatz import AnsatzEngine

    nparams = ansatz.nparams
    sets = []
    for sh in shifts:
        for i in range(nparams):
            p = base.copy()
            p[i] += sh
            sets.append(p)
    sets = np.array(sets)
    out = np.empty(len(sets))
    eng = AnsatzEngine(ansatz, chunk, dtype="complex128")
    for s in range(0, len(sets), chunk):
        part = sets[s:s + chunk]
        if part.shape[0] < chunk:
            part = np.concatenate([part, np.repeat(part[-1:], chunk - part.shape[0], axis=0)])
        out[s:s + chunk] = eng.energies(part).cpu().numpy()[: min(chunk, len(sets) - s)]
    del eng
    torch.cuda.empty_cache()
    return out.reshape(len(shifts), nparams)


def run_adjoint(backend, n=24, layers=8, batch=16, steps=3, g=1.0, seed=2024):
    """Adjoint gradients of <H_TFIM> for an n-qubit, `layers`-layer
    hardware-efficient ansatz, `batch` parameter sets: the batched engine
    (complex64 and complex128) against the plain per-gate complex128 path,
    with gradient errors against complex128 central differences (h = 1e-4)
    and parameter shift for parameter set 0."""
    import torch

    from qibojit_amd.ansatz import AnsatzEngine, HardwareEfficientAnsatz, ansatz_gradient_order, plain_adjoint

    ansatz = HardwareEfficientAnsatz(n, layers)
    rng = np.random.default_rng(seed)
    params = ansatz.random_parameters(rng, batch)
    res = {"nqubits": n, "layers": layers, "params": ansatz.nparams, "batch": batch, "g": g}
    grads = {}
    for dtype in ("complex64", "complex128"):
        eng = AnsatzEngine(ansatz, batch, dtype=dtype, g=g)
        eng.gradients(params)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            energies, gr = eng.gradients(params)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        grads[dtype] = gr.cpu().numpy()
        res[f"batched_{dtype}"] = {
            "s_per_batch": dt,
            "grads_per_s": batch * ansatz.nparams / dt,
            "k_forward": eng.k_forward,
            "k_adjoint": eng.k_adjoint,
            "passes_per_batch": ansatz.nlayers * (len(eng.groups_f) + len(eng.groups_a)) + 1,
        }
        res[f"energy0_{dtype}"] = float(energies[0].item())
        del eng
        torch.cuda.empty_cache()
    # plain per-gate complex128 path on parameter set 0
    backend.set_dtype("complex128")
    circuit = ansatz.circuit(params[0])
    plain_adjoint(backend, HardwareEfficientAnsatz(6, 1).circuit(params[0][:12]), g=g)  # warm-up kernels
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e_plain, g_plain = plain_adjoint(backend, circuit, g=g)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    plain = np.empty(ansatz.nparams)
    plain[ansatz_gradient_order(ansatz)] = g_plain
    res["plain_complex128"] = {"s_per_circuit": dt, "grads_per_s": ansatz.nparams / dt, "energy": e_plain}
    # references for parameter set 0
    t0 = time.perf_counter()
    h = 1e-4
    e_fd = _shifted_energies(ansatz, params[0], [h, -h])
    fd = (e_fd[0] - e_fd[1]) / (2 * h)
    e_ps = _shifted_energies(ansatz, params[0], [np.pi / 2, -np.pi / 2])
    ps = 0.5 * (e_ps[0] - e_ps[1])
    res["reference_seconds"] = time.perf_counter() - t0
    err = lambda a, b: float(np.max(np.abs(a - b)))  # noqa: E731
    res["max_grad_error"] = {
        "batched_c64_vs_fd": err(grads["complex64"][0], fd),
        "batched_c64_vs_ps": err(grads["complex64"][0], ps),
        "batched_c128_vs_fd": err(grads["complex128"][0], fd),
        "batched_c128_vs_ps": err(grads["complex128"][0], ps),
        "plain_c128_vs_ps": err(plain, ps),
        "fd_vs_ps": err(fd, ps),
        "batched_c64_vs_batched_c128_all_sets": err(grads["complex64"], grads["complex128"]),
    }
    res["max_abs_grad"] = float(np.max(np.abs(ps)))
    res["speedup_batched_c64_vs_plain_c128"] = res["batched_complex64"]["grads_per_s"] / res["plain_complex128"]["grads_per_s"]
    res["speedup_batched_c128_vs_plain_c128"] = res["batched_complex128"]["grads_per_s"] / res["plain_complex128"]["grads_per_s"]
    return res


def run_batch(backend, ncirc=1024, steps=5, g=1.0, seed=4096):
    """TFIM energies of `ncirc` ansatz circuits, qubits U[10,22], layers
    U[4,12], angles U[-pi, pi), complex64: autotuned plan, one captured graph
    per batch, against a loop of plain per-circuit launches."""
    import torch

    from qibojit_amd.batch import AnsatzBatch, autotune, plain_energies

    rng = np.random.default_rng(seed)
    shapes = [(int(rng.integers(10, 23)), int(rng.integers(4, 13))) for _ in range(ncirc)]
    params = [rng.uniform(-np.pi, np.pi, 2 * n * nl) for n, nl in shapes]
    res = {"circuits": ncirc, "qubits": [10, 22], "layers": [4, 12], "dtype": "complex64",
           "gates": int(sum(nl * (3 * n - 1) for n, nl in shapes))}
    t0 = time.perf_counter()
    plan, timings = autotune(shapes, params, reps=2)
    res["autotune_seconds"] = time.perf_counter() - t0
    res["plan"] = {str(n): list(c) for n, c in sorted(plan.items())}
    res["autotune_ms"] = {str(n): {"/".join(map(str, c)): round(t * 1e3, 3) for c, t in tv.items()}
                          for n, tv in sorted(timings.items())}
    batch = AnsatzBatch(shapes, dtype="complex64", g=g, plan=plan)
    batch.set_parameters(params)
    batch.capture()
    batch.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    e64 = batch.energies().cpu().numpy()
    t0 = time.perf_counter()
    batch.run()
    torch.cuda.synchronize()
    res["eager_seconds"] = time.perf_counter() - t0
    res["graph_seconds_per_batch"] = dt
    res["graph_circuits_per_s"] = ncirc / dt
    del batch
    torch.cuda.empty_cache()
    # FP64-state reference of every circuit (same runner, complex128)
    ref = AnsatzBatch(shapes, dtype="complex128", g=g)
    ref.set_parameters(params)
    ref.run()
    e128 = ref.energies().cpu().numpy()
    del ref
    torch.cuda.empty_cache()
    # plain loop: one launch per gate, one expectation per circuit
    backend.set_dtype("complex64")
    plain_energies(backend, shapes[:2], params[:2], g)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e_plain = plain_energies(backend, shapes, params, g)
    torch.cuda.synchronize()
    res["plain_loop_seconds"] = time.perf_counter() - t0
    res["plain_circuits_per_s"] = ncirc / res["plain_loop_seconds"]
    backend.set_dtype("complex128")
    sub = list(range(0, ncirc, 16))
    e_plain128 = plain_energies(backend, [shapes[i] for i in sub], [params[i] for i in sub], g)
    res["speedup_graph_vs_plain_loop"] = res["plain_loop_seconds"] / dt
    res["max_energy_error"] = {
        "batched_c64_vs_batched_c128_all": float(np.max(np.abs(e64 - e128))),
        "batched_c64_vs_plain_c64_all": float(np.max(np.abs(e64 - e_plain))),
        "batched_c128_vs_plain_c128_every_16th": float(np.max(np.abs(e128[sub] - e_plain128))),
    }
    return res


def _clifford_gates(n, ngates, seed):
    """Seeded gate list: 60% one-qubit gates (10 kinds), 40% two-qubit (5)."""
    from qibojit_amd.backends import clifford_operations_hip as ch

    rng = np.random.default_rng(seed)
    two = rng.random(ngates) < 0.4
    kinds = np.where(two, rng.integers(10, 15, ngates), rng.integers(0, 10, ngates)).astype(np.int32)
    a = rng.integers(0, n, ngates)
    b = (a + rng.integers(1, n, ngates)) % n if n > 1 else a
    gates = np.stack([kinds, a, np.where(two, b, -1)], axis=1).astype(np.int32)
    names = {v: k for k, v in ch.KINDS.items()}
    return gates, names


def run_clifford(sizes=((100, 100000), (1000, 100000), (10000, 20000)), reps=3, nmeas=64, cpu_seconds=2.0,
                 seed=77, nshots=100000, loop_shots=5):
    """Stabilizer tableaux: a seeded random Clifford gate list from |0...0>
    applied in one launch (k_cliff_circuit; schedule compiled once, launches
    timed with HIP events on the launch stream), the same list one launch per
    gate (k_cliff_gate, as the reference launches), the oracle's per-gate
    NumPy rules on the host (bounded sample), bit-exact check of the launch
    against the oracle, collapsing measurements of `nmeas` qubits, and
    `nshots` shots of 32 qubits (affine-map sampling) against the
    shot-by-shot measurement loop."""
    import torch

    from oracle import clifford_ref as cr
    from qibojit_amd.backends import clifford_operations_hip as ch

    out = []
    for n, ngates in sizes:
        gates, names = _clifford_gates(n, ngates, seed + n)
        rec = {"nqubits": n, "gates": ngates, "two_qubit_fraction": float(np.mean(gates[:, 0] >= 10))}
        tab = ch.zero_state(n)
        ch.apply_circuit(tab, gates, n)  # warm-up + parity run
        torch.cuda.synchronize()
        # oracle on the same list (bounded by time: checks the prefix it reaches)
        sm = cr.zero_tableau(n)
        t0 = time.perf_counter()
        done = 0
        for kind, a, b in gates:
            cr.apply(sm, names[int(kind)], (int(a),) if kind < 10 else (int(a), int(b)), n)
            done += 1
            if time.perf_counter() - t0 > 4 * cpu_seconds and done < ngates:
                break
        cpu_t = time.perf_counter() - t0
        rec["cpu_oracle_gates_per_s"] = done / cpu_t
        rec["cpu_oracle_sample"] = f"first {done} gates, {cpu_t:.2f} s, NumPy rules on the dense uint8 tableau, 1 thread"
        if done == ngates:
            rec["bit_exact_vs_oracle"] = bool(np.array_equal(ch._clifford_post_execution_reshape(tab, n).cpu().numpy(), sm))
        else:
            check = ch.zero_state(n)
            ch.apply_circuit(check, gates[:done], n)
            rec["bit_exact_vs_oracle"] = bool(
                np.array_equal(ch._clifford_post_execution_reshape(check, n).cpu().numpy(), sm))
            rec["bit_exact_prefix_gates"] = done
            del check
        # one launch for the whole list (schedule compiled once on the host)
        t0 = time.perf_counter()
        sched = ch.compile_circuit(gates, n)
        torch.cuda.synchronize()
        rec["schedule_seconds"] = time.perf_counter() - t0
        rec["groups"] = sched.ngroups
        stream = torch.cuda.current_stream()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(stream)
        for _ in range(reps):
            ch.apply_schedule(tab, sched, n)
        ev[1].record(stream)
        torch.cuda.synchronize()
        t_launch = ev[0].elapsed_time(ev[1]) / 1e3 / reps
        rec["launch_seconds"] = t_launch
        rec["launch_gates_per_s"] = ngates / t_launch
        rec["ns_per_gate"] = t_launch / ngates * 1e9
        # apply_circuit: host scheduling + descriptor upload + launch
        t0 = time.perf_counter()
        ch.apply_circuit(tab, gates, n)
        torch.cuda.synchronize()
        rec["call_seconds_incl_host_schedule"] = time.perf_counter() - t0
        # one launch per gate
        sub = gates[: min(ngates, 20000)]
        ev[0].record(stream)
        for kind, a, b in sub:
            ch._apply(int(kind), tab, int(a), int(b), n)
        ev[1].record(stream)
        torch.cuda.synchronize()
        t_pg = ev[0].elapsed_time(ev[1]) / 1e3
        rec["per_gate_launch_gates_per_s"] = len(sub) / t_pg
        rec["speedup_launch_vs_per_gate"] = rec["launch_gates_per_s"] / rec["per_gate_launch_gates_per_s"]
        rec["speedup_launch_vs_cpu_oracle"] = rec["launch_gates_per_s"] / rec["cpu_oracle_gates_per_s"]
        # the reference's traffic model: each gate reads and writes its 3 (5) columns
        col_bytes = (2 * n + 1) / 8
        rec["per_gate_model_GBps"] = float(np.mean(np.where(gates[:, 0] >= 10, 5, 3)) * 2 * col_bytes
                                           * rec["launch_gates_per_s"] / 1e9)
        rec["tableau_MB"] = (2 * n + 1) * ch.column_words(n) * 4 / 1e6
        # collapsing measurements
        rng = np.random.default_rng(seed)
        qubits = [int(v) for v in rng.choice(n, size=min(nmeas, n), replace=False)]
        np.random.seed(seed)
        ch.M(tab, qubits[:2], n, collapse=True)
        tab_host_loop = tab.clone()
        torch.cuda.synchronize()
        state = np.random.get_state()
        t0 = time.perf_counter()
        sample = ch.M(tab, qubits, n, collapse=True)  # one call: the loop enqueued on the device
        torch.cuda.synchronize()
        rec["measure_seconds_per_qubit"] = (time.perf_counter() - t0) / len(qubits)
        rec["measured_ones"] = int(sum(sample))
        # the same measurements through the per-qubit entry points (two host
        # round trips per qubit, as qibo's loop over the outcome functions)
        np.random.set_state(state)
        t0 = time.perf_counter()
        sample_host_loop = []
        for q in qubits:
            p = ch.first_stabilizer(tab_host_loop, q, n)
            if p >= 0:
                _, o = ch._random_outcome(tab_host_loop, [p], q, n)
            else:
                _, o = ch._determined_outcome(tab_host_loop, q, n)
            sample_host_loop.append(int(o))
        torch.cuda.synchronize()
        rec["measure_seconds_per_qubit_host_loop"] = (time.perf_counter() - t0) / len(qubits)
        rec["measure_loops_agree"] = bool(sample_host_loop == sample and torch.equal(tab_host_loop, tab))
        del tab_host_loop
        # many shots of 32 qubits not measured above: affine-map sampling vs
        # the shot-by-shot loop
        measured = set(qubits)
        sq = [q for q in range(n) if q not in measured][:32]
        np.random.seed(seed)
        t0 = time.perf_counter()
        shots = ch.sample_shots(tab, sq, n, nshots)
        torch.cuda.synchronize()
        rec["sample_shots"] = {"qubits": len(sq), "shots": nshots, "seconds": time.perf_counter() - t0}
        t0 = time.perf_counter()
        for _ in range(loop_shots):
            ch.M(tab, sq, n)
        torch.cuda.synchronize()
        per_shot = (time.perf_counter() - t0) / loop_shots
        rec["sample_shots"]["shot_by_shot_seconds_per_shot"] = per_shot
        rec["sample_shots"]["speedup_vs_shot_by_shot"] = per_shot * nshots / rec["sample_shots"]["seconds"]
        rec["sample_shots"]["mean"] = float(shots.float().mean())
        # every qubit up to 1000 of them, 10000 shots: one symbolic pass
        wide = list(range(min(n, 1000)))
        np.random.seed(seed)
        t0 = time.perf_counter()
        forms, k = ch.measurement_forms(tab, wide, n)
        t_forms = time.perf_counter() - t0
        np.random.seed(seed)
        t0 = time.perf_counter()
        shots = ch.sample_shots(tab, wide, n, 10000)
        torch.cuda.synchronize()
        rec["sample_shots_wide"] = {"qubits": len(wide), "random_outcomes": k, "shots": 10000,
                                    "forms_seconds": t_forms, "seconds": time.perf_counter() - t0}
        out.append(rec)
        del tab
        torch.cuda.empty_cache()
    return out


def cpu_baseline(n, min_seconds=10.0, max_seconds=30.0):
    """The reference's kernels restated in C/OpenMP, on this host's cores.

    Sample: the 27-gate sweep (targets 0, n/2, n-1) on an n-qubit complex128
    state, repeated gate by gate until at least `min_seconds` of CPU work
    (stopping early past `max_seconds`); GB/s = algorithmic bytes / time,
    the same byte count as the GPU sweep.
    """
    from oracle import port

    threads = len(os.sched_getaffinity(0))
    pb = port.PortBackend(nthreads=threads)
    rng = np.random.default_rng(1234)
    spec = sweep_gates(n, rng)
    state = np.empty(2**n, dtype=np.complex128)
    state.fill(1 / math.sqrt(2**n))  # values do not affect the timing
    total_b, total_t, applied = 0.0, 0.0, 0
    while total_t < min_seconds:
        for lbl, g, qs, pred in spec:
            fl, _ = changed_fraction(n, qs, pred, 16)
            t = time.perf_counter()
            state = pb.apply_gate(g, state, n)
            total_t += time.perf_counter() - t
            total_b += 2 * 16 * 2**n * fl
            applied += 1
            if total_t > max_seconds:
                break
        if total_t > max_seconds:
            break
    cpu = "unknown"
    try:
        with open("/proc/cpuinfo", encoding="utf-8") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": total_b / total_t / 1e9,
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{applied} gate applications of the {len(spec)}-gate sweep (targets 0, {n // 2}, {n - 1}) on a "
                  f"{n}-qubit complex128 state, {total_t:.1f} s of CPU time; C/OpenMP restatement of the numba "
                  f"kernels, {threads} threads; CPU: {cpu}",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: every box here has a single MI355X (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import torch

    if args.copy_probe:
        # A plain device copy of exactly 4 GiB each way: the reference for
        # scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run.
        a = torch.ones(2**28, dtype=torch.complex128, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": 16 * 2**28, "launches": 3}), flush=True)
        return

    from qibojit_amd.backends.hip import HipBackend

    backend = HipBackend()
    if args.only_adjoint:
        print(json.dumps({"adjoint_24q": run_adjoint(backend)}), flush=True)
        return
    if args.only_batch:
        print(json.dumps({"batch_1024": run_batch(backend)}), flush=True)
        return
    if args.only_mfma:
        paths = tuple(p == "mfma" for p in args.block_paths.split(","))
        print(json.dumps({"block_kernels": run_block_kernels(backend, placements=tuple(args.blocks.split(",")),
                                                             paths=paths)}), flush=True)
        return
    if args.only_qft:
        for dt in ("complex64", "complex128"):
            print(json.dumps({f"qft_32q_{dt}_fused": run_qft(backend, 32, dt, fused=True, reps=3)}), flush=True)
        return
    if args.only_random:
        print(json.dumps({"random_circuit_20q_complex128": run_random_circuit(backend, cpu=not args.no_cpu)}),
              flush=True)
        return
    if args.only_measure:
        for dt in ("complex128", "complex64"):
            print(json.dumps({f"measurement_{dt}": run_measurement_ops(backend, args.nqubits, dt)}), flush=True)
        return
    if args.only_clifford:
        print(json.dumps({"clifford": run_clifford()}), flush=True)
        return
    if args.only_precision:
        print(json.dumps({"block_kernels": run_block_kernels(backend), "precision_32q": run_precision(backend)},
                         default=str), flush=True)
        return
    n = args.nqubits
    print(f"# sweep complex128 n={n} steps={args.steps} warmup={args.warmup}", file=sys.stderr, flush=True)
    main_res = run_sweep(backend, n, "complex128", args.steps, args.warmup)
    # Detail lines first (one JSON object per line); the last stdout line is
    # the compact headline.
    emit("per_gate_complex128", [_rounded(r) for r in main_res["rows"]])
    summary = {}
    if not args.only_sweep:
        print("# sweep complex64", file=sys.stderr, flush=True)
        c64 = run_sweep(backend, n, "complex64", args.steps, args.warmup)
        emit("sweep_complex64", {"GBps": c64["GBps"], "ms_per_step": c64["ms_per_step"],
                                 "frac": c64["GBps"] / HBM_PEAK_GBS, "rows": [_rounded(r) for r in c64["rows"]]})
        summary["sweep_c64_GBps"] = round(c64["GBps"], 1)
    if not args.no_extras and not args.only_sweep:
        print("# measurement ops", file=sys.stderr, flush=True)
        meas = run_measurement_ops(backend, n, "complex128")
        emit("measurement_complex128", meas)
        summary["collapse_c128_ms"] = round(meas["collapse_normalized"]["ms"], 3)
        print("# random circuit 20q complex128", file=sys.stderr, flush=True)
        rc = run_random_circuit(backend, cpu=not args.no_cpu)
        emit("random_circuit_20q_complex128", rc)
        summary["random_20q_c128_fused_ms"] = round(rc["fused_s"] * 1e3, 3)
        print("# qft 32q complex64", file=sys.stderr, flush=True)
        qft = run_qft(backend, 32, "complex64", fused=True, reps=3)
        emit("qft_32q_complex64_fused", qft)
        summary["qft_32q_c64_s"] = round(qft["seconds"], 4)
        emit("qft_32q_complex64", run_qft(backend, 32, "complex64"))
        print("# capacity 33q complex128, 34q complex64", file=sys.stderr, flush=True)
        cap = run_capacity(backend, 33, "complex128")
        emit("capacity_33q_complex128", cap)
        cap64 = run_capacity(backend, 34, "complex64")
        emit("capacity_34q_complex64", cap64)
        summary["capacity_GBps"] = {"33q_c128": round(cap.get("GBps", float("nan")), 1),
                                    "34q_c64": round(cap64.get("GBps", float("nan")), 1)}
        print("# precision 32q complex64 vs complex128", file=sys.stderr, flush=True)
        blocks = run_block_kernels(backend)
        emit("block_kernels_k4_k5", blocks)
        summary["block5_bits0_4_hbm_frac"] = {"c128": blocks["complex128_k5_low_mfma"]["hbm_frac"],
                                              "c64": blocks["complex64_k5_low_mfma"]["hbm_frac"]}
        emit("precision_32q_complex64", run_precision(backend))
        print("# adjoint gradients 24q", file=sys.stderr, flush=True)
        adj = run_adjoint(backend)
        emit("adjoint_24q_8layers_batch16", adj)
        summary["adjoint_24q_c64_grads_per_s"] = round(adj["batched_complex64"]["grads_per_s"], 1)
        summary["adjoint_vs_plain_c128"] = round(adj["speedup_batched_c64_vs_plain_c128"], 2)
        summary["adjoint_c64_max_grad_err_vs_fd"] = adj["max_grad_error"]["batched_c64_vs_fd"]
        print("# 1024 ansatz circuits", file=sys.stderr, flush=True)
        bat = run_batch(backend)
        emit("batch_1024_ansatz_complex64", bat)
        print("# clifford tableaux", file=sys.stderr, flush=True)
        emit("clifford_tableau", run_clifford())

    name, avg_ms, avg_bytes, achieved, launches = dominant_kernel(main_res["rows"])
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": PROFILED_TRAFFIC.get((name, n)),
        "kernel": name,
        "launches_per_step": launches,
        "avg_launch_ms": round(avg_ms, 4),
        "algorithmic_bytes_per_launch": avg_bytes,
    }
    cpu = None
    if not args.no_cpu and not args.only_sweep:
        print("# cpu baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(n)

    line = {
        "metric": "eff. HBM GB/s (30q gate sweep); QFT circuit s; adjoint grads/s vs plain @1e-6",
        "value": round(main_res["GBps"], 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_res["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded complex Gaussian state, normalised)",
        "config": {
            "workload": f"{n}-qubit gate sweep: H, X, Z, RZ, CNOT, SWAP, fSim, CU3, 3q Haar on targets 0, {n // 2}, {n - 1}",
            "nqubits": n,
            "gates_per_step": len(main_res["rows"]),
            "algorithmic_GB_per_step": round(main_res["bytes_per_step"] / 1e9, 3),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "summary": summary,
        "device": torch.cuda.get_device_name(0),
    }
    print(json.dumps(line, separators=(",", ":")), flush=True)


def _rounded(row):
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in row.items()}


def emit(name, payload):
    """One detail record per stdout line, ahead of the headline."""
    print(json.dumps({"detail": name, "result": payload}, default=str, separators=(",", ":")), flush=True)


if __name__ == "__main__":
    main()
