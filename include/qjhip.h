/*
 * qjhip.h -- C-ABI of the MI355X (gfx950) state-vector kernels.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns an int status (QJ_OK == 0).  On failure the thread-local message of
 * qj_last_error() says what went wrong; the Python host maps QJ_EVALUE to
 * ValueError, QJ_EMEMORY to MemoryError and QJ_EHIP to RuntimeError.
 *
 * State vectors live in device memory owned by the caller (any allocator:
 * qj_malloc below, or a framework tensor's data pointer).  All calls that take
 * a `stream` (a hipStream_t, NULL = the legacy default stream) are
 * stream-ordered and do not synchronise the host, unless documented otherwise
 * (functions returning a host value synchronise that stream).
 *
 * Bit convention (reference: src/qibojit/backends/cpu.py:565-569, :610):
 * qubit q of an n-qubit register is bit (n - 1 - q) of the amplitude index,
 * so qubit 0 is the most significant bit.
 *
 * Gate matrices are host pointers to row-major complex numbers of the state's
 * precision (complex64 = 2 x float, complex128 = 2 x double).  They are copied
 * into kernel arguments or a stream-ordered staging ring; the caller may reuse
 * the host buffer as soon as the call returns.
 *
 * Concurrency: the library keeps per-device scratch (matrix staging ring,
 * reduction partials).  Calls on one device must be issued on one stream at a
 * time; calls issued concurrently on different streams of one device race.
 */
#ifndef QJHIP_H
#define QJHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define QJ_OK 0
#define QJ_EVALUE 1  /* bad argument (shape, qubit index, too many targets) */
#define QJ_EMEMORY 2 /* device or pinned-host allocation failed */
#define QJ_EHIP 3    /* HIP runtime error (launch failure, no device, ...) */

/* ---- element types ---- */
#define QJ_COMPLEX64 0
#define QJ_COMPLEX128 1
#define QJ_FLOAT32 2
#define QJ_FLOAT64 3

/* ---- one-qubit kernel kinds (reference names: GATE_OPS, cpu.py:23-36) ---- */
#define QJ_APPLY_GATE 0  /* dense 2x2 */
#define QJ_APPLY_X 1     /* swap of the pair */
#define QJ_APPLY_Y 2     /* (a, b) -> (-i b, i a) */
#define QJ_APPLY_Z 3     /* b -> -b */
#define QJ_APPLY_Z_POW 4 /* b -> gate[0] * b */

/* ---- two-qubit kernel kinds ---- */
#define QJ_APPLY_TWO_QUBIT_GATE 0 /* dense 4x4 */
#define QJ_APPLY_SWAP 1           /* |01> <-> |10> */
#define QJ_APPLY_FSIM 2           /* 2x2 on {|01>,|10>} (gate[0..3]), phase gate[4] on |11> */

/* ---- runtime ---- */

/* Thread-local text of the last error ("" when none). */
const char* qj_last_error(void);
/* Version of this ABI; bumped on any signature change. */
int qj_abi_version(void);
/* Number of visible HIP devices (0 without a GPU; no error). */
int qj_device_count(int* count);
/* Make `device` current for the calling host thread. */
int qj_set_device(int device);
/* Device allocation of `nbytes` (QJ_EMEMORY on failure). */
int qj_malloc(void** ptr, uint64_t nbytes);
int qj_free(void* ptr);
/* Stream-ordered copies host <-> device.  Pageable host memory makes these
 * effectively synchronous for the host. */
int qj_memcpy_htod(void* dst, const void* src, uint64_t nbytes, void* stream);
int qj_memcpy_dtoh(void* dst, const void* src, uint64_t nbytes, void* stream);
int qj_memcpy_dtod(void* dst, const void* src, uint64_t nbytes, void* stream);
int qj_stream_synchronize(void* stream);

/* ---- state preparation ----
 * Replaces ops.initial_state_vector (src/qibojit/custom_operators/ops.py:10-18)
 * and the CupyBackend zeros + initial_state_kernel pair
 * (src/qibojit/backends/gpu.py:584-600, raw_kernels.py:548-558).
 * Writes |0...0> into `nelements` amplitudes (a density matrix of dimension d
 * is the d*d vector with element 0 set: ops.py:21-30). */
int qj_initial_state(void* state, int dtype, uint64_t nelements, void* stream);

/* ---- gate application ----
 * `nqubits` is the length exponent of the vector (2*n for a density matrix).
 * `qubits` is the reference's "qubits tensor" (cpu.py:565-569): the bit
 * positions of controls and targets, sorted ascending, `nqubits_sorted` long
 * (pass NULL/0 for a gate without controls: the targets are then implied).
 * Bit positions of controls are the entries of `qubits` that are not targets.
 */

/* One-qubit gate on bit `m` (= nqubits - target - 1).
 * Replaces NumbaBackend._one_qubit_base (cpu.py:606-616) with the kernels
 * apply_gate/x/y/z/z_pow_kernel and their multicontrol_ variants
 * (src/qibojit/custom_operators/gates.py:15-114; raw_kernels.py:128-221,
 * 293-396).  `gate`: 4 complex (dense, row-major) for QJ_APPLY_GATE, 1 complex
 * for QJ_APPLY_Z_POW, ignored (may be NULL) for X/Y/Z. */
int qj_apply_one_qubit(void* state, int dtype, int nqubits, int kind, int m,
                       const int32_t* qubits, int nqubits_sorted,
                       const void* gate, void* stream);

/* Two-qubit gate.  `bit1`/`bit2` are the bit positions of the FIRST and
 * SECOND listed target: matrix row index = 2*bit(target1) + bit(target2).
 * Replaces NumbaBackend._two_qubit_base (cpu.py:618-635), which encodes the
 * same pair as (m1 < m2, swap_targets), and the kernels
 * apply_two_qubit_gate/swap/fsim_kernel + multicontrol_ variants
 * (gates.py:117-254; raw_kernels.py:224-290, 399-477).  `gate`: 16 complex
 * (dense), 5 complex (fsim: gates.py:230-234), ignored for swap. */
int qj_apply_two_qubit(void* state, int dtype, int nqubits, int kind, int bit1,
                       int bit2, const int32_t* qubits, int nqubits_sorted,
                       const void* gate, void* stream);

/* Dense k-qubit gate, 1 <= k <= 12.  `target_bits[u]` is the bit position of
 * the u-th listed target; target 0 is the most significant bit of the matrix
 * index (equivalent to the reference's reversed int64 strides,
 * cpu.py:594-596 and gates.py:257-262).  `gate`: 4^k complex, row-major.
 * Replaces NumbaBackend._multi_qubit_base (cpu.py:581-604) with
 * apply_three/four/five_qubit_gate_kernel and apply_multi_qubit_gate_kernel
 * (gates.py:265-424; raw_kernels.py:480-518). */
int qj_apply_multi_qubit(void* state, int dtype, int nqubits,
                         const int32_t* target_bits, int ntargets,
                         const int32_t* qubits, int nqubits_sorted,
                         const void* gate, void* stream);

/* A dense 4- or 5-target block (2^k x 2^k, row-major, complex64 or
 * complex128), optionally controlled, applied on matrix cores: the state
 * viewed as 2^k x 2^(n-k-c) is multiplied by the block as three real GEMMs
 * (the 3M form) with v_mfma_f32_16x16x4f32 (complex64) or
 * v_mfma_f64_16x16x4f64 (complex128).  Same conventions as
 * qj_apply_multi_qubit: target_bits[0] is the row index's most significant
 * bit; `qubits` (NULL for no controls) lists every bit position, targets and
 * controls, ascending, and controls must be 1.  At least 4 bits must be left
 * free (16-column tiles).  It is the matrix-core form of
 * apply_four/five_qubit_gate_kernel and their multicontrol variants
 * (gates.py:265-411) for the fused blocks of the fusion executor
 * (cpu.py:535-537). */
int qj_apply_block_mfma(void* state, int dtype, int nqubits,
                        const int32_t* target_bits, int ntargets,
                        const int32_t* qubits, int nqubits_sorted,
                        const void* gate, void* stream);

/* ---- fused passes (tiles) ----
 * One pass streams the state through tiles of 2^(low_bits + nhigh)
 * amplitudes: a tile holds every combination of the `low_bits` lowest bits
 * (contiguous in memory) and of the `nhigh` bit positions `high_bits`
 * (ascending, all >= low_bits), for one assignment of the other bits.  The
 * `nops` operations are applied to each tile in order and the tile is
 * written back: the whole list costs one read and one write of the state.
 * This is the gate-fusion executor; it has no single counterpart in the
 * reference, which applies a fused block (qibo FusedGate via
 * matrix_fused, cpu.py:535-537) as one dense gate per pass.
 *
 * Tile-local bit b means bit b for b < low_bits and high_bits[b - low_bits]
 * otherwise.  Op kinds:
 *   QJ_TILE_DENSE: 2^k x 2^k row-major matrix (k <= 5) at `data + offset` on
 *     the tile-local target bits target[0..k-1] (target[0] = matrix MSB);
 *   QJ_TILE_PHASE: multiplies amplitude i by exp(i*angle(i)) with
 *     angle(i) = c0 + sum_j theta_j * bit_j(i).  The block at
 *     `data + offset` is {int32 ga, gb, tb, pad; OA[2^ga], OB[2^gb], W[tb]}
 *     (complex, the state's precision).  With t the tile index (the
 *     out-of-tile bits of i, ascending, packed; ga = (nqubits - tb + 1) / 2,
 *     gb = nqubits - tb - ga) and e the tile-local index (tb = low_bits +
 *     nhigh), the factor is OA[t mod 2^ga] * OB[t >> ga] * prod of W[b] over
 *     the bits b set in e: OA and OB hold exp(i * sum of the thetas of the
 *     out-of-tile bits set in their index), c0 folded into OA, and
 *     W[b] = exp(i * theta of tile bit b);
 *   QJ_TILE_SWAP: exchanges tile-local bits target[0] and target[1] (k = 2,
 *     no controls, no data), the reference's apply_swap_kernel
 *     (custom_operators/gates.py:196-216) inside the tile.
 * Dense and phase ops apply only where every bit of `local_ctrl` (tile-local
 * mask) and of `global_ctrl` (mask of out-of-tile bit positions) is 1. */
#define QJ_TILE_DENSE 0
#define QJ_TILE_PHASE 1
#define QJ_TILE_SWAP 2
#define QJ_TILE_MAX_K 5
#define QJ_TILE_MAX_BITS 13

typedef struct {
  int32_t kind;
  int32_t k;
  int32_t target[QJ_TILE_MAX_K + 1];
  uint64_t local_ctrl;
  uint64_t global_ctrl;
  uint64_t offset; /* byte offset of the op's data in `data` (16-aligned) */
  uint64_t reserved;
} qj_tile_op; /* 64 bytes */

int qj_apply_tile_pass(void* state, int dtype, int nqubits, int low_bits,
                       const int32_t* high_bits, int nhigh,
                       const qj_tile_op* ops, int nops, const void* data,
                       uint64_t data_bytes, void* stream);

/* Swaps of disjoint bit pairs (pairs[2j], pairs[2j+1]), j < npairs, all in one
 * read and one write of the state: the permutation of a run of SWAP gates
 * on distinct qubits, such as the final swaps of a QFT.  Each pair is the
 * reference's apply_swap_kernel (custom_operators/gates.py:196-216);
 * the reference runs one pass per pair. */
int qj_apply_bit_swaps(void* state, int dtype, int nqubits, const int32_t* pairs, int npairs, void* stream);

/* ---- layered variational circuits (batched forward, TFIM, adjoint) ----
 * No counterpart in the reference (qibojit has no differentiation or
 * batching).  These serve circuits of single-qubit rotation layers and CNOT
 * ladders CNOT(q, q+1), q = 0..n-2, applied to many states at once.
 *
 * States: launch row b (< nrows) works on state s = sel[b] (sel is a device
 * array of int32, or NULL for s = b) stored at s * stride amplitudes from the
 * buffer start.  Matrices: a device array, state s's at mats + s * mstride
 * complex entries (the state's precision), laid out [bit j][step][2x2
 * row-major].
 *
 * qj_rot_pass owns the nbits (1..QJ_ROT_MAX_BITS, adjoint mode
 * 1..QJ_ROT_MAX_ADJ_BITS) ascending bit positions
 * bits[] and applies to each of them its nrot (<= QJ_ROT_MAX_STEPS) matrices
 * in order, after loading through `load`:
 *   QJ_LOAD_PLAIN    x = in (in == out allowed)
 *   QJ_LOAD_LADDER   x = CNOT ladder applied to in, x[k] = in[k ^ (k >> 1)]
 *   QJ_LOAD_UNLADDER x = ladder undone, x[k] = in[P(k)], bit j of P(k) the
 *                    parity of the bits >= j of k
 *   QJ_LOAD_ZERO     x = |0...0> (in not read)
 * Permuted loads need in != out.  With lam_in/lam_out set the pass runs
 * backwards (adjoint method) on psi and lam: for step j = nrot-1..0 of each
 * bit it adds Im<lam|G|psi> (G = X 1, Y 2, Z 3 from gens[bit * nrot + j],
 * 0 = no gradient) into partials[((bit * nrot + j) * nrows + b) * ncols + c]
 * (c < ncols, the workgroups per state) in FP64, then applies the step's
 * U^+ to both.  For exp(-i theta G / 2) and lam = U_{>k}^+ H psi_L this is
 * d<psi_L|H|psi_L>/d theta.  A nonzero gens entry also declares the step's
 * matrix to be exp(-i theta G / 2) (any theta, per state), and the pass
 * applies it in that form: Y real, Z diagonal, X real on the diagonal and
 * imaginary off it; the matrix entries outside that form are not read.
 * gens may be given without lam (forward passes use the forms too). */
#define QJ_ROT_MAX_BITS 5
#define QJ_ROT_MAX_ADJ_BITS 5 /* with lam (psi and lam in registers): 5 for complex64, 4 for complex128 */
#define QJ_ROT_MAX_STEPS 3
#define QJ_LOAD_PLAIN 0
#define QJ_LOAD_LADDER 1
#define QJ_LOAD_UNLADDER 2
#define QJ_LOAD_ZERO 3

int qj_rot_pass(const void* psi_in, void* psi_out, const void* lam_in,
                void* lam_out, int dtype, int nqubits, uint64_t stride,
                const int32_t* sel, int nrows, const int32_t* bits, int nbits,
                int nrot, const int32_t* gens, int load, const void* mats,
                uint64_t mstride, double* partials, int ncols, void* stream);
/* The same pass on LDS tiles (complex64): 12 tile bits tile_bits[] (ascending;
 * the lowest should be bits 0..3 so that loads are whole 128-byte lines) are
 * held per workgroup, and the rotations of the tile bits set in rotmask run
 * in three register phases with LDS transposes in between -- up to 12 bits
 * per pass instead of QJ_ROT_MAX_BITS.  Arguments as qj_rot_pass, with
 * matrices laid out [tile bit j][step][2x2] (mstride >= 12 * nrot * 4) and
 * gradient slot j * nrot + step; needs 12 <= nqubits. */
int qj_rot_tile_pass(const void* psi_in, void* psi_out, const void* lam_in,
                     void* lam_out, int dtype, int nqubits, uint64_t stride,
                     const int32_t* sel, int nrows, const int32_t* tile_bits,
                     int rotmask, int nrot, const int32_t* gens, int load,
                     const void* mats, uint64_t mstride, double* partials,
                     int ncols, void* stream);
/* Open-chain transverse-field Ising model over the bits,
 * H = -sum_j Z_j Z_{j+1} - g sum_q X_q.  For each row b: out = H' psi (out
 * may be NULL: energy only; it must not alias psi) and
 * partials[b * ncols + c] summing to Re<psi|H' psi>, where H' = H or, with
 * `ladder` set, L^+ H L for the CNOT ladder L (psi is the state before it). */
int qj_tfim_batch(const void* psi, void* out, int dtype, int nqubits,
                  uint64_t stride, const int32_t* sel, int nrows, double g,
                  int ladder, double* partials, int ncols, void* stream);
/* Whole ansatz circuits in LDS, one workgroup per circuit at a time
 * (persistent: workgroup w takes circuits w, w + nblocks, ...).  desc is a
 * device array of ncirc x {nqubits, nlayers, offset lo, offset hi}: circuit c
 * has nlayers layers of [RZ RY fused per bit][ladder] with matrices at
 * mats + offset, laid out [layer][bit][2x2] (bit p = qubit n-1-p), and
 * energies[c] receives its TFIM energy <H> (FP64).  nqmax bounds the qubit
 * count (2^nqmax amplitudes of LDS, at most 64 KiB: 13 qubits complex64,
 * 12 complex128); circuits outside the bounds get NaN. */
int qj_ansatz_lds(int dtype, const int32_t* desc, int ncirc, const void* mats,
                  uint64_t mats_len, double g, double* energies, int nqmax,
                  int nblocks, void* stream);
/* out[r] = sum_{c < ncols} in[r * ncols + c] in FP64, fixed order. */
int qj_reduce_rows(const double* in, uint64_t nrows, uint64_t ncols,
                   double* out, void* stream);

/* ---- measurement ---- */

/* Every one of the n amplitudes set to re + i im (stream-ordered): the plus
 * state 2^(-n/2) and its density matrix 2^-n.  Replaces qibo's plus_state
 * (NumpyBackend), which NumbaBackend inherits. */
int qj_fill_state(void* state, int dtype, uint64_t n, double re, double im,
                  void* stream);

/* Collapse onto outcome `result` of the measured bits `qubits_kernel`
 * (= [n - q - 1 for q in reversed(measured qubits)], cpu.py:550-552); bit j of
 * `result` is the value of bit qubits_kernel[j] (ops.py:33-40; for ascending
 * qubits_kernel this is the reference's collapse_index set exactly, and it
 * stays well defined for any order).  Zeroes every
 * other amplitude and, when `normalize`, divides the kept ones by the FP64
 * norm sqrt(sum |kept|^2).  Replaces ops.collapse_state /
 * collapse_state_normalized (ops.py:43-79) and CupyBackend.collapse_state
 * (gpu.py:616-644).  `norm_out` (host, may be NULL) receives that norm; when
 * it is non-NULL the call synchronises `stream`. */
int qj_collapse_state(void* state, int dtype, int nqubits,
                      const int32_t* qubits_kernel, int nmeasured,
                      int64_t result, int normalize, double* norm_out,
                      void* stream);

/* Marginal probabilities of `qubits` (qubit ids, in list order; the first is
 * the most significant bit of the output index), accumulated in FP64 and
 * written to `probs` (device, 2^nmeasured elements of `prob_dtype`
 * QJ_FLOAT32/QJ_FLOAT64).  Replaces the qibo calculate_probabilities used by
 * CupyBackend.calculate_probabilities (gpu.py:751-768). */
int qj_calculate_probabilities(const void* state, int dtype, int nqubits,
                               const int32_t* qubits, int nmeasured,
                               void* probs, int prob_dtype, void* stream);
/* Marginal probabilities of a density matrix (2^n x 2^n, row-major): the
 * diagonal Re rho_kk summed over the unmeasured qubits, same output order
 * and dtypes as qj_calculate_probabilities.  Replaces qibo's density-matrix
 * branch of calculate_probabilities, which CupyBackend.calculate_probabilities
 * forwards to (gpu.py:751-768). */
int qj_calculate_probabilities_dm(const void* state, int dtype, int nqubits,
                                  const int32_t* qubits, int nmeasured,
                                  void* probs, int prob_dtype, void* stream);

/* Metropolis frequency sampler (ops.measure_frequencies, ops.py:82-108):
 * `nchains` chains; chain c re-seeds an MT19937 generator with
 * `chain_seeds[c]` (host array; the reference draws them with
 * randint(0, 1e8) after seed(seed)), starts at argmax(probs), and for each of
 * its shots proposes (shot + randint(0, nstates)) % nstates with NumPy-legacy
 * masked randint, accepts when probs[new]/probs[shot] (computed in the probs
 * type, compared in double) > random(), then counts the current shot.  Chain
 * c runs nshots/nchains shots; the last chain also runs the remainder.
 * `frequencies` (device int64, nstates) is ADDED to, as in the reference. */
int qj_measure_frequencies(int64_t* frequencies, const void* probs,
                           int prob_dtype, int64_t nstates, int64_t nshots,
                           const uint32_t* chain_seeds, int nchains,
                           void* stream);

/* Shot sampler below the Metropolis threshold: NumPy's
 * RandomState.choice(nstates, nshots, p=probs), which qibo's sample_shots
 * runs on the host and CupyBackend.sample_shots (gpu.py:770-773) copies the
 * probabilities back for.  cdf = cumsum(probs) in FP64, cdf /= cdf[-1],
 * samples[j] = first index i with cdf[i] > uniforms[j] (searchsorted,
 * side='right').  `probs` (QJ_FLOAT32/QJ_FLOAT64), `uniforms` (double,
 * nshots values in [0, 1), drawn by the caller with the generator choice()
 * would use) and `samples` (int64, nshots) are device arrays.  Up to 2^16
 * states the cumsum is sequential in index order, as NumPy's; above, a
 * blocked FP64 scan. */
int qj_sample_shots(const void* probs, int prob_dtype, int64_t nstates,
                    const double* uniforms, int64_t nshots, int64_t* samples,
                    void* stream);
/* No counterpart in the reference: qj_sample_shots for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, semantics and results
 * as qj_sample_shots, but the CDF lives in a fixed per-GPU device array of
 * 2^20 + 64 doubles (8 MiB, held until the process ends) instead of the
 * library's growable work buffer, and the call makes no hipMalloc, hipFree
 * or other allocating call.  At most 2^20 states: larger `nstates` is
 * rejected with QJ_EVALUE before anything is enqueued.  Every call on one
 * GPU shares that array, so calls that overlap on one GPU (two streams, two
 * threads) race; calls ordered on one stream do not. */
int qj_sample_shots_noalloc(const void* probs, int prob_dtype, int64_t nstates,
                            const double* uniforms, int64_t nshots, int64_t* samples,
                            void* stream);
/* Counts of samples: frequencies[v] += 1 for each samples[j] = v in
 * [0, nstates) (device int64 arrays); samples outside that range add to
 * *out_of_range (device int64) instead.  The device side of
 * calculate_frequencies (gpu.py:744-748 counts on the host with np.unique). */
int qj_count_samples(const int64_t* samples, int64_t nshots,
                     int64_t* frequencies, int64_t nstates,
                     int64_t* out_of_range, void* stream);

/* ---- reductions (FP64 accumulation; synchronise `stream`) ---- */

/* Compare two states of n amplitudes, each complex64 or complex128 (FP64
 * sums, host output): out = {Re <a|b>, Im <a|b>, <a|a>, <b|b>,
 * max_i |a_i - b_i|}.  For precision reports (infidelity of a complex64 run
 * against a complex128 one); no reference counterpart. */
int qj_compare_states(const void* a, int dtype_a, const void* b, int dtype_b,
                      uint64_t n, double* out, void* stream);
/* sum |state_i|^2 over `nelements` amplitudes. */
int qj_norm_squared(const void* state, int dtype, uint64_t nelements,
                    double* out, void* stream);
/* <a|b> = sum conj(a_i) b_i ; out = {re, im}. */
int qj_vdot(const void* a, const void* b, int dtype, uint64_t nelements,
            double* out, void* stream);
/* <psi| H |psi> for the open-chain transverse-field Ising model
 * H = -sum_{i<n-1} Z_i Z_{i+1} - g sum_i X_i. */
int qj_expectation_tfim(const void* state, int dtype, int nqubits, double g,
                        double* out, void* stream);

/* ---- elementwise helpers (stream-ordered, no host sync) ---- */

/* y = alpha * x + beta * y  (alpha, beta complex: {re, im}); with beta == 0
 * y is write-only (it may hold uninitialised memory). */
int qj_axpby(const void* x, void* y, int dtype, uint64_t nelements,
             const double* alpha, const double* beta, void* stream);

/* ---- stabilizer tableaux (Clifford circuits) ----
 *
 * Replaces the module functions of src/qibojit/backends/
 * clifford_operations_gpu.py (CUDA RawKernels on a row-packed uint8 matrix)
 * and clifford_operations_cpu.py (numba, unpacked uint8 matrix).
 *
 * The reference's tableau of n qubits is a (2n+1) x (2n+1) matrix of 0/1
 * bytes: rows 0..n-1 destabilizers, n..2n-1 stabilizers, 2n scratch;
 * columns 0..n-1 x bits, n..2n-1 z bits, 2n the sign r.  Here the device
 * tableau (`tab`, uint32 words, device memory) stores it by column: column
 * c occupies words [c * ws, (c + 1) * ws) with ws =
 * qj_clifford_column_words(n), row j is bit j % 32 of word j / 32, and bits
 * of rows >= 2n+1 are zero.  A tableau is (2n+1) * ws words. */
#define QJ_CLIFF_H 0
#define QJ_CLIFF_S 1
#define QJ_CLIFF_Z 2
#define QJ_CLIFF_X 3
#define QJ_CLIFF_Y 4
#define QJ_CLIFF_SX 5
#define QJ_CLIFF_SDG 6
#define QJ_CLIFF_SXDG 7
#define QJ_CLIFF_RY_PI 8    /* RY(pi/2): the reference's RY_pi */
#define QJ_CLIFF_RY_3PI_2 9 /* RY(3pi/2) */
#define QJ_CLIFF_CNOT 10    /* first two-qubit kind: (control, target) */
#define QJ_CLIFF_CZ 11
#define QJ_CLIFF_SWAP 12
#define QJ_CLIFF_ISWAP 13
#define QJ_CLIFF_CY 14
#define QJ_CLIFF_MAX_QUBITS (1 << 19)

/* Words per column (ws) for n qubits: ceil((2n+1)/32) rounded up to a
 * multiple of 4; -1 for n outside [1, QJ_CLIFF_MAX_QUBITS]. */
int64_t qj_clifford_column_words(int nqubits);
/* Dense (2n+1) x (2n+1) row-major byte matrix (device; nonzero = 1) ->
 * device tableau, and back (bytes 0/1).  The layout change the reference
 * makes with _clifford_pre/post_execution_reshape
 * (clifford_operations_gpu.py:694-700). */
int qj_clifford_pack(const uint8_t* dense, int nqubits, uint32_t* tab, void* stream);
int qj_clifford_unpack(const uint32_t* tab, int nqubits, uint8_t* dense, void* stream);
/* One gate (the kernels apply_H ... apply_CY, clifford_operations_gpu.py:
 * 49-594; rules of clifford_operations_cpu.py:18-248) on every row.  kind is
 * a QJ_CLIFF_* value; q1 is the target of two-qubit kinds (q0 the control)
 * and ignored otherwise. */
int qj_clifford_apply_gate(uint32_t* tab, int nqubits, int kind, int q0, int q1, void* stream);
/* A gate list in one launch: `gates` is a host array of ngates triples
 * (kind, q0, q1).  The result equals the gates applied one by one in list
 * order.  No reference counterpart (the reference launches per gate). */
int qj_clifford_apply_circuit(uint32_t* tab, int nqubits, const int32_t* gates, int64_t ngates, void* stream);
/* The two halves of qj_clifford_apply_circuit, for a gate list applied many
 * times.  qj_clifford_schedule orders the list on the host into groups of
 * gates of one kind on disjoint qubits and writes *ngroups groups of
 * QJ_CLIFF_GROUP_WORDS uint32 words to the host array `desc` (desc == NULL:
 * only *ngroups; QJ_EVALUE if capacity, in groups, is smaller).
 * qj_clifford_apply_schedule runs such an array, copied unchanged to device
 * memory, on a tableau of the same nqubits.  No reference counterpart. */
#define QJ_CLIFF_GROUP_WORDS 32
int qj_clifford_schedule(const int32_t* gates, int64_t ngates, int nqubits, uint32_t* desc, int64_t capacity,
                         int64_t* ngroups);
int qj_clifford_apply_schedule(uint32_t* tab, int nqubits, const uint32_t* desc, int64_t ngroups, void* stream);
/* *p_out (device int32) = first stabilizer p (0-based: row n+p) with x bit
 * 1 on qubit q, or -1: the test of qibo's measurement loop that chooses
 * between the random and the determined outcome. */
int qj_clifford_first_stabilizer(const uint32_t* tab, int nqubits, int q, int32_t* p_out, void* stream);
/* Random outcome on qubit q with first anticommuting stabilizer p
 * (_random_outcome, clifford_operations_gpu.py:606-625): every row h < 2n,
 * h != n+p, with x bit 1 on q becomes P_{n+p} P_h (rowsum); row p becomes
 * the destabilizer row p; row n+p becomes (-1)^outcome Z_q.  The caller draws
 * `outcome` (0 or 1). */
int qj_clifford_random_outcome(uint32_t* tab, int nqubits, int p, int q, int outcome, void* stream);
/* Determined outcome on qubit q (_determined_outcome,
 * clifford_operations_gpu.py:628-640): the scratch row becomes the XOR of
 * the stabilizer rows n+i with destabilizer x bit 1 on q, with the XOR of
 * their signs (the reference's determined rowsum); *outcome_out (device
 * int32) = that sign. */
int qj_clifford_determined_outcome(uint32_t* tab, int nqubits, int q, int32_t* outcome_out, void* stream);
/* qj_clifford_determined_outcome, then the sign of the selected stabilizers'
 * product with its phases accumulated (the sequential rowsum of Aaronson and
 * Gottesman): outcome = (sum r_j + sum_{j<l} z_j.x_l + sum |x_j & z_j| / 2)
 * mod 2 over the selected rows in row order, into the scratch row's sign and
 * *outcome_out.  Always the state's certain outcome; the reference's XOR rule
 * can differ from it.  No reference counterpart (an opt-in beside
 * clifford_operations_gpu.py:628-640). */
int qj_clifford_determined_outcome_exact(uint32_t* tab, int nqubits, int q, int32_t* outcome_out, void* stream);
/* qibo's Clifford measurement loop (the loop of _random_outcome /
 * _determined_outcome, clifford_operations_gpu.py:606-640, over `qubits`)
 * enqueued in one call with no host round trip: for each qubits[i] (host
 * int32, m entries) the first anticommuting stabilizer is found on the
 * device; a random outcome takes draws[t] (device uint8, m entries, t = the
 * random outcomes taken before it), a determined one the reference's XOR
 * rule (exact != 0: the rule of qj_clifford_determined_outcome_exact).
 * sample_out (device int32, m) gets the outcomes, *nrandom_out (device
 * int32) the number of random ones.  Same tableau and outcomes as the
 * per-qubit entry points. */
int qj_clifford_measure(uint32_t* tab, int nqubits, const int32_t* qubits, int m, const uint8_t* draws, int exact,
                        int32_t* sample_out, int32_t* nrandom_out, void* stream);
/* rowsum of clifford_operations_cpu.py:337-385 on host arrays h, i of
 * nrows row indices: row h[j] <- P_{i[j]} P_{h[j]} (x, z XORed; r = (2 r_h +
 * 2 r_i + sum of the exponents of i) mod 4 != 0), every product from the
 * state before the call.  The h rows must be distinct and no i row may be an
 * h row.  determined != 0: every h[j] equals h[0], and row h[0] is XORed
 * with every (row i[j] XOR row h[0]) and its sign with every new sign. */
int qj_clifford_rowsum(uint32_t* tab, int nqubits, const int64_t* h, const int64_t* i, int64_t nrows,
                       int determined, void* stream);
/* Symbolic measurement pass (the map of qj_gf2_samples in one pass).
 * `signs` (device, (nvars + 1) * ws words) holds sign columns: column j of
 * row r = the coefficient of term j in row r's sign as an affine form over
 * GF(2) of the random outcomes (j = 0 the constant, j >= 1 variable j).
 * qj_clifford_sym_init sets column 0 to the tableau's sign column and the
 * others to 0.  qj_clifford_sym_random is qj_clifford_random_outcome on the
 * x/z bits of `tab`, with the signs carried as forms and row n+p's sign set
 * to variable `var` (1 <= var <= nvars < 65535).  qj_clifford_sym_determined
 * writes the form of qibo's determined outcome on q (the XOR of the selected
 * stabilizers' signs) to form[0..nvars] (device bytes).  The tableau's own
 * sign column is not maintained.  No reference counterpart. */
int qj_clifford_sym_init(const uint32_t* tab, int nqubits, uint32_t* signs, int nvars, void* stream);
int qj_clifford_sym_random(uint32_t* tab, int nqubits, uint32_t* signs, int nvars, int p, int q, int var,
                           void* stream);
int qj_clifford_sym_determined(const uint32_t* tab, int nqubits, const uint32_t* signs, int nvars, int q,
                               uint8_t* form, void* stream);
/* qj_clifford_sym_determined for the exact outcome of
 * qj_clifford_determined_outcome_exact: the phase term (set by the x/z bits
 * only) XORed into form[0].  No reference counterpart. */
int qj_clifford_sym_determined_exact(const uint32_t* tab, int nqubits, const uint32_t* signs, int nvars, int q,
                                     uint8_t* form, void* stream);
/* Many measurement shots of one tableau: for a fixed state the measured
 * bits are affine over GF(2) in the random outcomes drawn on the way, so
 * shot s is s0 ^ (u[s] . D).  out[s * m + q] = s0[q] ^ parity(u[s] & dt[q])
 * with u (nshots x kw uint64 words: shot s's random outcome bits, bit j of
 * the row = j-th random outcome), dt (m x kw words: row q = the outcomes
 * whose flip flips measured qubit q), s0 (m bytes: the sample with every
 * random outcome 0); all device arrays, out uint8 (nshots x m).  No
 * reference counterpart (qibo samples a Clifford state shot by shot). */
int qj_gf2_samples(const uint64_t* u, const uint64_t* dt, const uint8_t* s0, int64_t nshots, int64_t m, int64_t kw,
                   uint8_t* out, void* stream);
/* numpy.packbits / numpy.unpackbits(count=...) along the last axis of a
 * row-major byte matrix (device), big bit order: the reference's _packbits /
 * _unpackbits (clifford_operations_cpu.py:256-327). */
int qj_packbits_rows(const uint8_t* in, int64_t rows, int64_t cols, uint8_t* out, void* stream);
int qj_unpackbits_rows(const uint8_t* in, int64_t rows, int64_t packed_cols, int64_t count, uint8_t* out,
                       void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QJHIP_H */
