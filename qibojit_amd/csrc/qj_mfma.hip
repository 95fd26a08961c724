// Fused 5-qubit blocks on complex64 states as FP32 matrix-core GEMMs.
//
// A 5-target block multiplies every group of 32 amplitudes (the targets'
// combinations for one assignment of the other bits) by the 32x32 complex
// matrix U: the state is a 32 x 2^(n-5) complex matrix X (rows = target
// index, columns = groups) and the block computes U X.  Written over the
// reals, [Yr; Yi] = [[Ur, -Ui], [Ui, Ur]] [Xr; Xi]: a 64x64 by 64 x ncols
// GEMM, which v_mfma_f32_32x32x2f32 (32x32 outputs, K = 2 per instruction,
// FP32 products and sums) runs as 2 row tiles x 32 K-steps per 32 columns.
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
// Matrix convention as qj_apply_multi_qubit: bit (4 - u) of the row index is
// target target_bits[u] (the first target is the most significant bit).
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qj_common.h"

namespace qj {

typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int kMfmaBlock = 256;  // 4 waves, 32 columns each per step
constexpr u64 kMfmaMaxGrid = 8192;

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

// 3M form (default): three real 32x32 products instead of the 64x64 one,
//   P1 = Ur Xr,  P2 = Ui Xi,  P3 = (Ur + Ui)(Xr + Xi),
//   Re Y = P1 - P2,  Im Y = P3 - P1 - P2,
// with v_mfma_f32_16x16x4f32: 3 products x 2 row tiles x 8 K-steps = 48
// MFMAs per 16 columns (64 for the real 64x64 form).  A operands per lane:
// ur[t][s] = Ur[16t + n][4s + q] (and ui, us = ur + ui); lane (n, q) loads
// rows 4j + q of column n (its B operands for K-step j) and holds outputs
// rows 16t + 4q + r of column n.
// STAGE (every target below bit 9): a wave's 16 columns are one contiguous
// run of 512 amplitudes, moved through LDS with coalesced loads and stores
// (the direct loads of low targets are strided); slot a sits at a + a / 32
// so the column-wise reads spread over the banks.  The loop is uniform per
// block (ncols is a multiple of 64 there) so the block barriers are safe.
constexpr int kStageAmps = 512;
constexpr int kStageSlots = kStageAmps + kStageAmps / 32;

__device__ __forceinline__ unsigned stage_slot(unsigned a) { return a + (a >> 5); }

template <bool STAGE>
__global__ void __launch_bounds__(kMfmaBlock)
    k_block5_mfma3m(cplx<float>* __restrict__ st, u64 ncols, Ins ins, Offs32 off, const float* __restrict__ u3) {
  __shared__ cplx<float> stage[STAGE ? kMfmaBlock / 64 : 1][STAGE ? kStageSlots : 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, n = lane & 15;
  float ur[2][8], ui[2][8], us[2][8];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int e = (16 * t + n) * 32 + 4 * s + q;
      ur[t][s] = u3[e];
      ui[t][s] = u3[1024 + e];
      us[t][s] = u3[2048 + e];
    }
  // chunk-local index of this lane's column (targets inserted as zeros)
  const unsigned colbase = STAGE ? (unsigned)insert_bits((u64)n, ins) : 0u;
  const u64 step = (u64)gridDim.x * (kMfmaBlock / 64) * 16;
  for (u64 c0 = ((u64)blockIdx.x * (kMfmaBlock / 64) + wave) * 16; c0 < ncols; c0 += step) {
    const u64 base = insert_bits(c0 + (STAGE ? 0 : n), ins);
    float xr[8], xi[8], xs[8];
    if (STAGE) {
#pragma unroll
      for (int k = 0; k < kStageAmps / 64; ++k) stage[wave][stage_slot(64 * k + lane)] = st[base + 64 * k + lane];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const cplx<float> v = stage[wave][stage_slot(colbase + (unsigned)off.o[4 * j + q])];
        xr[j] = v.re;
        xi[j] = v.im;
        xs[j] = v.re + v.im;
      }
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const cplx<float> v = st[base + off.o[4 * j + q]];
        xr[j] = v.re;
        xi[j] = v.im;
        xs[j] = v.re + v.im;
      }
    }
    v4f p1[2], p2[2], p3[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) p1[t] = p2[t] = p3[t] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        p1[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ur[t][s], xr[s], p1[t], 0, 0, 0);
        p2[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ui[t][s], xi[s], p2[t], 0, 0, 0);
        p3[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(us[t][s], xs[s], p3[t], 0, 0, 0);
      }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const cplx<float> y = mk<float>(p1[t][r] - p2[t][r], p3[t][r] - p1[t][r] - p2[t][r]);
        const unsigned row = (unsigned)off.o[16 * t + 4 * q + r];
        if (STAGE)
          stage[wave][stage_slot(colbase + row)] = y;
        else
          st[base + row] = y;
      }
    if (STAGE) {
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kStageAmps / 64; ++k) st[base + 64 * k + lane] = stage[wave][stage_slot(64 * k + lane)];
      __syncthreads();
    }
  }
}

// QJ_MFMA_SHAPE: 3 (3M form, default), 16 (real 64x64 form, 16x16x4) or 32
// (real 64x64 form, 32x32x2).
static int mfma_shape() {
  static const int shape = [] {
    const char* e = std::getenv("QJ_MFMA_SHAPE");
    const int v = e ? std::atoi(e) : 3;
    return (v == 32 || v == 16) ? v : 3;  // 3: the 3M form
  }();
  return shape;
}

static int block5_impl(void* state, int nqubits, const int32_t* target_bits, const void* gate, hipStream_t s) {
  if (nqubits < 10 || nqubits > 40)
    return set_error(QJ_EVALUE, "apply_block_mfma: needs 10..40 qubits (32-column tiles), got %d", nqubits);
  Ins ins;
  std::memset(&ins, 0, sizeof(ins));
  int pos[5];
  for (int u = 0; u < 5; ++u) {
    if (target_bits[u] < 0 || target_bits[u] >= nqubits)
      return set_error(QJ_EVALUE, "apply_block_mfma: target bit %d out of range", target_bits[u]);
    for (int v = 0; v < u; ++v)
      if (target_bits[v] == target_bits[u]) return set_error(QJ_EVALUE, "apply_block_mfma: repeated target");
    pos[u] = target_bits[u];
  }
  std::sort(pos, pos + 5);
  ins.npos = 5;
  for (int u = 0; u < 5; ++u) ins.pos[u] = pos[u];
  Offs32 off;
  for (int r = 0; r < 32; ++r) {
    u64 o = 0;
    for (int u = 0; u < 5; ++u)
      if ((r >> (4 - u)) & 1) o |= u64(1) << target_bits[u];
    off.o[r] = o;
  }
  const cplx<float>* u = static_cast<const cplx<float>*>(gate);
  const int shape = mfma_shape();
  std::vector<float> w;
  if (shape == 3) {
    // Ur, Ui, Ur + Ui as three 32x32 row-major blocks
    w.resize(3 * 1024);
    for (int e = 0; e < 1024; ++e) {
      w[e] = u[e].re;
      w[1024 + e] = u[e].im;
      w[2048 + e] = u[e].re + u[e].im;
    }
  } else {
    // real 64x64 form [[Ur, -Ui], [Ui, Ur]], row-major
    w.resize(64 * 64);
    for (int r = 0; r < 32; ++r)
      for (int c = 0; c < 32; ++c) {
        const cplx<float> z = u[r * 32 + c];
        w[r * 64 + c] = z.re;
        w[r * 64 + 32 + c] = -z.im;
        w[(32 + r) * 64 + c] = z.im;
        w[(32 + r) * 64 + 32 + c] = z.re;
      }
  }
  void* dw = nullptr;
  int ticket = -1;
  int rc = ring_upload(w.data(), w.size() * sizeof(float), s, &dw, &ticket);
  if (rc) return rc;
  const u64 ncols = u64(1) << (nqubits - 5);
  const int cols_per_wave = shape == 32 ? 32 : 16;
  const u64 tasks = ncols / cols_per_wave;  // wave steps
  u64 grid = (tasks + (kMfmaBlock / 64) - 1) / (kMfmaBlock / 64);
  if (grid > kMfmaMaxGrid) grid = kMfmaMaxGrid;
  // targets below bit 9: a wave's columns form one contiguous run (staged)
  const bool stage = shape == 3 && pos[4] < 9 && nqubits >= 11;
  if (stage)
    k_block5_mfma3m<true><<<dim3((unsigned)grid), dim3(kMfmaBlock), 0, s>>>(
        static_cast<cplx<float>*>(state), ncols, ins, off, static_cast<const float*>(dw));
  else if (shape == 3)
    k_block5_mfma3m<false><<<dim3((unsigned)grid), dim3(kMfmaBlock), 0, s>>>(
        static_cast<cplx<float>*>(state), ncols, ins, off, static_cast<const float*>(dw));
  else if (cols_per_wave == 32)
    k_block5_mfma<<<dim3((unsigned)grid), dim3(kMfmaBlock), 0, s>>>(static_cast<cplx<float>*>(state), ncols, ins,
                                                                   off, static_cast<const float*>(dw));
  else
    k_block5_mfma16<<<dim3((unsigned)grid), dim3(kMfmaBlock), 0, s>>>(static_cast<cplx<float>*>(state), ncols,
                                                                     ins, off, static_cast<const float*>(dw));
  rc = check_launch("apply_block_mfma launch");
  ring_release(ticket, s);
  return rc;
}

}  // namespace qj

using namespace qj;

extern "C" {

int qj_apply_block_mfma(void* state, int dtype, int nqubits, const int32_t* target_bits, int ntargets,
                        const void* gate, void* stream) {
  if (!state || !target_bits || !gate) return set_error(QJ_EVALUE, "apply_block_mfma: null pointer");
  if (dtype != QJ_COMPLEX64) return set_error(QJ_EVALUE, "apply_block_mfma: complex64 only (FP32 matrix cores)");
  if (ntargets != 5) return set_error(QJ_EVALUE, "apply_block_mfma: exactly 5 targets, got %d", ntargets);
  return block5_impl(state, nqubits, target_bits, gate, (hipStream_t)stream);
}

}  // extern "C"
