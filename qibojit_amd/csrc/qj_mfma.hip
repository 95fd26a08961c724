// Dense 4- and 5-target blocks as matrix-core GEMMs: FP32 MFMA on complex64
// states, FP64 MFMA on complex128 states.
//
// A K-target block multiplies every group of 2^K amplitudes (the targets'
// combinations for one assignment of the other bits) by the 2^K x 2^K complex
// matrix U: the state is a 2^K x ncols complex matrix X (rows = target index,
// columns = groups) and the block computes U X.  The default kernel
// (k_block_mfma3m below) uses three real products (the 3M form); two real
// 64x64 forms for uncontrolled complex64 5-target blocks are kept as
// measured alternatives: [Yr; Yi] = [[Ur, -Ui], [Ui, Ur]] [Xr; Xi], run with
// v_mfma_f32_32x32x2f32 (32x32 outputs, K = 2 per instruction) as 2 row
// tiles x 32 K-steps per 32 columns, or with v_mfma_f32_16x16x4f32.
//
// Lane layout of v_mfma_f32_32x32x2f32 for lane l, i = l % 32, h = l / 32:
//   A: A[i][h]    B: B[h][i]    D: register r holds D[8(r/4) + 4h + r%4][i].
// Each wave owns 32 columns: lane (i, h) loads the amplitudes of rows
// c = 2j + h (j < 16) of column i, which are exactly its B operands (K-step
// s < 16 reads Re X[2s + h], s >= 16 reads Im X[2(s-16) + h]); the real
// 64x64 matrix stays in registers as the A operands for the whole kernel.
// The outputs of row tile 0 are Re Y and of row tile 1 Im Y for the same
// rows, so each lane stores 16 complete amplitudes of its column.  Reads of
// a column finish before its writes (the MFMAs consume every lane's loads),
// so the update is in place.
//
// Matrix convention as qj_apply_multi_qubit: bit (K - 1 - u) of the row
// index is target target_bits[u] (the first target is the most significant
// bit).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qj_common.h"

namespace qj {

typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kMfmaBlock = 256;  // 4 waves, 32 columns each per step
constexpr u64 kMfmaMaxGrid = 8192;

// Workgroups of the 3M kernel: complex64 K = 4 131072 and K = 5 32768,
// complex128 K = 4 524288 and K = 5 8192 (the complex128 5-target kernel,
// 2 waves per SIMD, lost 15-60% on larger grids; the others gained 2-9% on
// the lowest bits over 8192 and stayed within noise elsewhere).
static u64 mfma_grid_cap(bool f64, int K) {
  if (K == 4) return f64 ? 524288 : 131072;
  return f64 ? 8192 : 32768;
}

__global__ void __launch_bounds__(kMfmaBlock)
    k_block5_mfma(cplx<float>* __restrict__ st, u64 ncols, Ins ins, Offs32 off, const float* __restrict__ w) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, i = lane & 31;
  float a0[32], a1[32];  // A operands of row tiles 0 (Re rows) and 1 (Im rows)
#pragma unroll
  for (int s = 0; s < 32; ++s) {
    a0[s] = w[i * 64 + 2 * s + h];
    a1[s] = w[(32 + i) * 64 + 2 * s + h];
  }
  const u64 step = (u64)gridDim.x * (kMfmaBlock / 64) * 32;
  for (u64 c0 = ((u64)blockIdx.x * (kMfmaBlock / 64) + wave) * 32; c0 < ncols; c0 += step) {
    const u64 base = insert_bits(c0 + i, ins);
    float xr[16], xi[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const cplx<float> v = st[base + off.o[2 * j + h]];
      xr[j] = v.re;
      xi[j] = v.im;
    }
    v16f d0 = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    v16f d1 = d0;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[s], xr[s], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[s], xr[s], d1, 0, 0, 0);
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      d0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[16 + s], xi[s], d0, 0, 0, 0);
      d1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[16 + s], xi[s], d1, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int m = 8 * (r >> 2) + 4 * h + (r & 3);
      st[base + off.o[m]] = mk<float>(d0[r], d1[r]);
    }
  }
}

// The same GEMM with v_mfma_f32_16x16x4f32 (16x16 outputs, K = 4): lane l,
// n = l % 16, q = l / 16 supplies A[n][q], B[q][n] and holds D[4q + r][n] in
// register r.  A wave owns 16 columns; lane (n, q) loads rows c = 4j + q
// (j < 8) of column n -- its B operands for K-steps j (Re) and 8 + j (Im).
// Four row tiles of 16: tiles 0/1 give Re Y of rows 0..31, tiles 2/3 Im Y of
// the same rows.  Half the operand registers of the 32x32x2 form.
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kMfmaBlock)
    k_block5_mfma16(cplx<float>* __restrict__ st, u64 ncols, Ins ins, Offs32 off, const float* __restrict__ w) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, n = lane & 15;
  float a[4][16];  // a[t][s] = W[16t + n][4s + q]
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int s = 0; s < 16; ++s) a[t][s] = w[(16 * t + n) * 64 + 4 * s + q];
  const u64 step = (u64)gridDim.x * (kMfmaBlock / 64) * 16;
  for (u64 c0 = ((u64)blockIdx.x * (kMfmaBlock / 64) + wave) * 16; c0 < ncols; c0 += step) {
    const u64 base = insert_bits(c0 + n, ins);
    float xr[8], xi[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const cplx<float> v = st[base + off.o[4 * j + q]];
      xr[j] = v.re;
      xi[j] = v.im;
    }
    v4f d[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) d[t] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int t = 0; t < 4; ++t) d[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][s], xr[s], d[t], 0, 0, 0);
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int t = 0; t < 4; ++t) d[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][8 + s], xi[s], d[t], 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      st[base + off.o[4 * q + r]] = mk<float>(d[0][r], d[2][r]);
      st[base + off.o[16 + 4 * q + r]] = mk<float>(d[1][r], d[3][r]);
    }
  }
}

// 3M form (default), for K = 4 or 5 targets (NR = 2^K rows) on complex64
// (v_mfma_f32_16x16x4f32) or complex128 (v_mfma_f64_16x16x4f64) states:
// three real NR x NR products instead of the 2NR x 2NR one,
//   P1 = Ur Xr,  P2 = Ui Xi,  P3 = (Ur + Ui)(Xr + Xi),
//   Re Y = P1 - P2,  Im Y = P3 - P1 - P2,
// 3 products x NR/16 row tiles x NR/4 K-steps MFMAs per 16 columns (48 for
// K = 5, 12 for K = 4).  Lane l, n = l % 16, q = l / 16 supplies A[n][q]
// and B[q][n] in both instructions; the results sit in register r as row
// 4q + r (f32) or q + 4r (f64) of column n.  A operands per lane:
// ur[t][s] = Ur[16t + n][4s + q] (and ui, us = ur + ui); lane (n, q) loads
// rows 4j + q of column n (its B operands for K-step j) and stores its
// output rows 16t + out_row(q, r) of column n.
// Columns are the assignments of the bits outside targets and controls
// (controls inserted as ones), so a controlled block runs on its 2^-c share.
// STAGE (targets all below bit K + 4, no control there): a wave's 16 columns
// are one contiguous run of 16 NR amplitudes, moved through the wave's own
// LDS region with coalesced loads and stores (direct loads of low targets
// are strided); slot a sits at a + a / NR so the column-wise reads spread
// over the banks.  Only the wave's lanes touch its region, so a wave barrier
// orders the LDS traffic (WAVE_SYNC; a workgroup barrier otherwise, which
// then needs ncols to be a multiple of 64 so that the loop is uniform per
// workgroup).
// The two instructions read A and B alike but place D differently: register
// r of lane n + 16q holds row 4q + r (f32) or row q + 4r (f64) of column n
// (measured with one-hot operands on gfx950).
template <typename R>
struct Mfma16;
template <>
struct Mfma16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int out_row(int q, int r) { return 4 * q + r; }
};
template <>
struct Mfma16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int out_row(int q, int r) { return q + 4 * r; }
};

template <int K>
constexpr int stage_amps() {
  return 16 << K;
}
template <int K>
constexpr int stage_slots() {
  return stage_amps<K>() + 16;
}

// One pad slot per column of 2^K amplitudes: the 16 columns a K-step reads
// start on 16 different banks.
template <int K>
__device__ __forceinline__ unsigned stage_slot(unsigned a) {
  return a + (a >> K);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename R, int K, bool STAGE, bool WAVE_SYNC>
__global__ void __launch_bounds__(kMfmaBlock)
    k_block_mfma3m(cplx<R>* __restrict__ st, u64 ncols, Ins ins, Offs32 off, u64 lowmask, const R* __restrict__ u3) {
  typedef typename Mfma16<R>::v4 v4;
  constexpr int NR = 1 << K, T = NR / 16, S = NR / 4;
  constexpr int AMPS = stage_amps<K>();
  __shared__ cplx<R> stage[STAGE ? kMfmaBlock / 64 : 1][STAGE ? stage_slots<K>() : 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, n = lane & 15;
  // the three real NR x NR blocks are read from global memory once per
  // workgroup (not once per wave), then each wave takes its A operands
  __shared__ R umat[3 * NR * NR];
  for (int i = threadIdx.x; i < 3 * NR * NR; i += kMfmaBlock) umat[i] = u3[i];
  __syncthreads();
  R ur[T][S], ui[T][S], us[T][S];
#pragma unroll
  for (int t = 0; t < T; ++t)
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int e = (16 * t + n) * NR + 4 * s + q;
      ur[t][s] = umat[e];
      ui[t][s] = umat[NR * NR + e];
      us[t][s] = umat[2 * NR * NR + e];
    }
  // chunk-local index of this lane's column (targets inserted as zeros)
  const unsigned colbase = STAGE ? (unsigned)(insert_bits((u64)n, ins) & lowmask) : 0u;
  const u64 step = (u64)gridDim.x * (kMfmaBlock / 64) * 16;
  for (u64 c0 = ((u64)blockIdx.x * (kMfmaBlock / 64) + wave) * 16; c0 < ncols; c0 += step) {
    const u64 base = insert_bits(c0 + (STAGE ? 0 : n), ins);
    R xr[S], xi[S], xs[S];
    if (STAGE) {
#pragma unroll
      for (int k = 0; k < AMPS / 64; ++k) stage[wave][stage_slot<K>(64 * k + lane)] = st[base + 64 * k + lane];
      if (WAVE_SYNC)
        wave_lds_sync();
      else
        __syncthreads();
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const cplx<R> v = stage[wave][stage_slot<K>(colbase + (unsigned)off.o[4 * j + q])];
        xr[j] = v.re;
        xi[j] = v.im;
        xs[j] = v.re + v.im;
      }
    } else {
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const cplx<R> v = st[base + off.o[4 * j + q]];
        xr[j] = v.re;
        xi[j] = v.im;
        xs[j] = v.re + v.im;
      }
    }
    v4 p1[T], p2[T], p3[T];
#pragma unroll
    for (int t = 0; t < T; ++t) p1[t] = p2[t] = p3[t] = v4{R(0), R(0), R(0), R(0)};
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int t = 0; t < T; ++t) {
        p1[t] = Mfma16<R>::mma(ur[t][s], xr[s], p1[t]);
        p2[t] = Mfma16<R>::mma(ui[t][s], xi[s], p2[t]);
        p3[t] = Mfma16<R>::mma(us[t][s], xs[s], p3[t]);
      }
#pragma unroll
    for (int t = 0; t < T; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const cplx<R> y = mk<R>(p1[t][r] - p2[t][r], p3[t][r] - p1[t][r] - p2[t][r]);
        const unsigned row = (unsigned)off.o[16 * t + Mfma16<R>::out_row(q, r)];
        if (STAGE)
          stage[wave][stage_slot<K>(colbase + row)] = y;
        else
          st[base + row] = y;
      }
    if (STAGE) {
      if (WAVE_SYNC)
        wave_lds_sync();
      else
        __syncthreads();
#pragma unroll
      for (int k = 0; k < AMPS / 64; ++k) st[base + 64 * k + lane] = stage[wave][stage_slot<K>(64 * k + lane)];
      if (WAVE_SYNC)
        wave_lds_sync();
      else
        __syncthreads();
    }
  }
}

// QJ_MFMA_SHAPE: 3 (3M form, default), 16 (real 64x64 form, 16x16x4) or 32
// (real 64x64 form, 32x32x2); the real forms cover uncontrolled complex64
// 5-target blocks only.
static int mfma_shape() {
  static const int shape = [] {
    const char* e = std::getenv("QJ_MFMA_SHAPE");
    const int v = e ? std::atoi(e) : 3;
    return (v == 32 || v == 16) ? v : 3;  // 3: the 3M form
  }();
  return shape;
}

// QJ_MFMA_SYNC: "wave" (default) orders a wave's staged LDS traffic with a
// wave barrier, "block" with a workgroup barrier.
static bool mfma_wave_sync() {
  static const bool wave = [] {
    const char* e = std::getenv("QJ_MFMA_SYNC");
    return !(e && std::strcmp(e, "block") == 0);
  }();
  return wave;
}

template <typename R, int K, bool STAGE>
static void launch_3m(bool wave_sync, unsigned grid, cplx<R>* st, u64 ncols, const Ins& ins, const Offs32& off,
                      u64 lowmask, const R* u3, hipStream_t s) {
  if (wave_sync)
    k_block_mfma3m<R, K, STAGE, true><<<dim3(grid), dim3(kMfmaBlock), 0, s>>>(st, ncols, ins, off, lowmask, u3);
  else
    k_block_mfma3m<R, K, STAGE, false><<<dim3(grid), dim3(kMfmaBlock), 0, s>>>(st, ncols, ins, off, lowmask, u3);
}

template <typename R>
static int block_impl(void* state, int nqubits, const int32_t* target_bits, int K, const int32_t* qubits, int nq,
                      const void* gate, hipStream_t s) {
  if (K != 4 && K != 5) return set_error(QJ_EVALUE, "apply_block_mfma: 4 or 5 targets, got %d", K);
  if (nqubits < 1 || nqubits > 62) return set_error(QJ_EVALUE, "apply_block_mfma: bad nqubits %d", nqubits);
  for (int u = 0; u < K; ++u) {
    if (target_bits[u] < 0 || target_bits[u] >= nqubits)
      return set_error(QJ_EVALUE, "apply_block_mfma: target bit %d out of range", target_bits[u]);
    for (int v = 0; v < u; ++v)
      if (target_bits[v] == target_bits[u]) return set_error(QJ_EVALUE, "apply_block_mfma: repeated target");
  }
  // every bit position (targets and controls), ascending; controls are the
  // positions of `qubits` that are not targets
  std::vector<int> pos(target_bits, target_bits + K);
  u64 ones = 0;
  if (qubits) {
    if (nq < K || nq > kMaxPos) return set_error(QJ_EVALUE, "apply_block_mfma: bad number of qubits %d", nq);
    int ntarget_seen = 0;
    for (int j = 0; j < nq; ++j) {
      const int b = qubits[j];
      if (b < 0 || b >= nqubits) return set_error(QJ_EVALUE, "apply_block_mfma: qubit %d out of range", b);
      if (j > 0 && qubits[j - 1] >= b) return set_error(QJ_EVALUE, "apply_block_mfma: qubits must be ascending");
      const bool is_target = std::find(target_bits, target_bits + K, b) != target_bits + K;
      if (is_target) {
        ++ntarget_seen;
      } else {
        pos.push_back(b);
        ones |= u64(1) << b;
      }
    }
    if (ntarget_seen != K) return set_error(QJ_EVALUE, "apply_block_mfma: qubits must contain every target");
  }
  std::sort(pos.begin(), pos.end());
  const int npos = (int)pos.size();
  if (nqubits - npos < 4)
    return set_error(QJ_EVALUE, "apply_block_mfma: needs at least 4 free qubits (16-column tiles), got %d",
                     nqubits - npos);
  Ins ins;
  std::memset(&ins, 0, sizeof(ins));
  ins.npos = npos;
  for (int j = 0; j < npos; ++j) ins.pos[j] = pos[j];
  ins.ones = ones;
  Offs32 off;
  std::memset(&off, 0, sizeof(off));
  const int NR = 1 << K;
  for (int r = 0; r < NR; ++r) {
    u64 o = 0;
    for (int u = 0; u < K; ++u)
      if ((r >> (K - 1 - u)) & 1) o |= u64(1) << target_bits[u];
    off.o[r] = o;
  }
  const cplx<R>* u = static_cast<const cplx<R>*>(gate);
  const u64 ncols = u64(1) << (nqubits - npos);
  cplx<R>* st = static_cast<cplx<R>*>(state);
  const int shape = mfma_shape();
  if (shape != 3 && sizeof(R) == 4 && K == 5 && npos == 5 && ncols >= 32) {
    // real 64x64 form [[Ur, -Ui], [Ui, Ur]], row-major
    std::vector<float> w(64 * 64);
    for (int r = 0; r < 32; ++r)
      for (int c = 0; c < 32; ++c) {
        const cplx<R> z = u[r * 32 + c];
        w[r * 64 + c] = (float)z.re;
        w[r * 64 + 32 + c] = (float)-z.im;
        w[(32 + r) * 64 + c] = (float)z.im;
        w[(32 + r) * 64 + 32 + c] = (float)z.re;
      }
    void* dw = nullptr;
    int ticket = -1;
    int rc = ring_upload(w.data(), w.size() * sizeof(float), s, &dw, &ticket);
    if (rc) return rc;
    const int cols_per_wave = shape == 32 ? 32 : 16;
    u64 grid = (ncols / cols_per_wave + (kMfmaBlock / 64) - 1) / (kMfmaBlock / 64);
    if (grid > kMfmaMaxGrid) grid = kMfmaMaxGrid;
    cplx<float>* sf = reinterpret_cast<cplx<float>*>(state);
    if (shape == 32)
      k_block5_mfma<<<dim3((unsigned)grid), dim3(kMfmaBlock), 0, s>>>(sf, ncols, ins, off, static_cast<const float*>(dw));
    else
      k_block5_mfma16<<<dim3((unsigned)grid), dim3(kMfmaBlock), 0, s>>>(sf, ncols, ins, off,
                                                                       static_cast<const float*>(dw));
    rc = check_launch("apply_block_mfma launch");
    ring_release(ticket, s);
    return rc;
  }
  // Ur, Ui, Ur + Ui as three NR x NR row-major blocks
  std::vector<R> w(3 * NR * NR);
  for (int e = 0; e < NR * NR; ++e) {
    w[e] = u[e].re;
    w[NR * NR + e] = u[e].im;
    w[2 * NR * NR + e] = u[e].re + u[e].im;
  }
  void* dw = nullptr;
  int ticket = -1;
  int rc = ring_upload(w.data(), w.size() * sizeof(R), s, &dw, &ticket);
  if (rc) return rc;
  const R* du = static_cast<const R*>(dw);
  u64 grid = (ncols / 16 + (kMfmaBlock / 64) - 1) / (kMfmaBlock / 64);
  // grid cap per precision and block width (each workgroup loads the
  // matrix once: 3 / 6 KiB for K = 4, 12 / 24 KiB for K = 5), measured at
  // 32 qubits complex64 and 30 qubits complex128 over caps of 8192..2^21
  const u64 cap = mfma_grid_cap(sizeof(R) == 8, K);
  if (grid > cap) grid = cap;
  // at least two 16-column groups per wave: one group per wave reloads the
  // matrix for every 2^K x 16 amplitudes (complex128 K = 4 with a control:
  // 58% of the HBM peak at one group per wave, 68% at two or more)
  const u64 two_per_wave = ncols / (2 * 16 * (kMfmaBlock / 64));
  if (two_per_wave >= 1 && grid > two_per_wave) grid = two_per_wave;
  // The lowest K + 4 bits hold exactly the targets and 4 column bits: a
  // wave's 16 columns are one contiguous run (staged through LDS).
  const int low = K + 4;
  bool stage = nqubits >= low && (ones & ((u64(1) << low) - 1)) == 0;
  for (int u2 = 0; u2 < K; ++u2) stage = stage && target_bits[u2] < low;
  const bool wave_sync = mfma_wave_sync();
  // a workgroup barrier needs every wave of a workgroup in the loop alike
  if (stage && !wave_sync && ncols % ((kMfmaBlock / 64) * 16) != 0) stage = false;
  const u64 lowmask = (u64(1) << low) - 1;
  const unsigned g = (unsigned)grid;
  if (K == 5) {
    if (stage) launch_3m<R, 5, true>(wave_sync, g, st, ncols, ins, off, lowmask, du, s);
    else launch_3m<R, 5, false>(wave_sync, g, st, ncols, ins, off, lowmask, du, s);
  } else {
    if (stage) launch_3m<R, 4, true>(wave_sync, g, st, ncols, ins, off, lowmask, du, s);
    else launch_3m<R, 4, false>(wave_sync, g, st, ncols, ins, off, lowmask, du, s);
  }
  rc = check_launch("apply_block_mfma launch");
  ring_release(ticket, s);
  return rc;
}

}  // namespace qj

using namespace qj;

extern "C" {

int qj_apply_block_mfma(void* state, int dtype, int nqubits, const int32_t* target_bits, int ntargets,
                        const int32_t* qubits, int nqubits_list, const void* gate, void* stream) {
  if (!state || !target_bits || !gate) return set_error(QJ_EVALUE, "apply_block_mfma: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64)
    return block_impl<float>(state, nqubits, target_bits, ntargets, qubits, nqubits_list, gate, s);
  if (dtype == QJ_COMPLEX128)
    return block_impl<double>(state, nqubits, target_bits, ntargets, qubits, nqubits_list, gate, s);
  return set_error(QJ_EVALUE, "apply_block_mfma: dtype must be complex64/complex128");
}

}  // extern "C"
