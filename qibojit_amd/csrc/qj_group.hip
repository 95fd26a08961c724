// Layered variational circuits: batched forward evolution, TFIM energies and
// adjoint-method gradients for circuits made of single-qubit rotation layers
// and CNOT ladders (the hardware-efficient ansatz).
//
// Rotation pass.  A pass owns k <= 5 bit positions (ascending).  Each thread
// holds the 2^k amplitudes of one group (the other bits fixed) in registers
// and applies, on every local bit, a short sequence of 2x2 matrices; local
// bits are always compile-time indices, so the register arrays never spill.
// The ladder CNOT(q, q+1) for q = 0..n-2 is the GF(2)-linear permutation
// k -> P(k) with bit j of P(k) = parity of the bits >= j of k (qubit q is
// bit n-1-q); a pass can apply it, or its inverse, while loading:
//   QJ_LOAD_LADDER:   x[k] = in[k ^ (k >> 1)]            (= in[P^-1(k)])
//   QJ_LOAD_UNLADDER: x[k] = in[P(k)]
//   QJ_LOAD_ZERO:     x = |0...0>, `in` is not read.
// Both permutations keep 64-amplitude aligned blocks together, so permuted
// loads stay line-coalesced.
//
// Adjoint mode (lam != NULL) walks the sequence backwards on psi and lam: for
// step j = nrot-1..0 on each local bit it first accumulates, in FP64,
// Im<lam|G|psi> (G = X, Y or Z of that step) -- the derivative of
// <psi_L|H|psi_L> by the angle of a rotation exp(-i theta G / 2) when lam is
// U_{>k}^+ H psi_L -- then applies the step's U^+ to both states.
//
// Batches: launch row b works on state s = sel ? sel[b] : b at s * stride,
// with its matrices at mats + s * mstride.
#include <algorithm>
#include <cstring>
#include <vector>

#include "qj_common.h"

namespace qj {

constexpr int kRotBlock = 256;
constexpr int kRotMaxCols = 65535;

struct RotArgs {
  Ins ins;          // group positions, ascending, for insert_bits
  u64 off[32];      // offset of register r (bit A of r <-> position pos[A])
  int gen[QJ_ROT_MAX_BITS][QJ_ROT_MAX_STEPS];  // generator per step, 0 = none
  int nrot;         // steps per bit
  int load;         // QJ_LOAD_*
  u64 stride;       // amplitudes between states
  u64 mstride;      // complex matrix entries between states' matrices
  int nsel;         // launch rows
};

__device__ __forceinline__ u64 suffix_parity(u64 k) {
  k ^= k >> 1;
  k ^= k >> 2;
  k ^= k >> 4;
  k ^= k >> 8;
  k ^= k >> 16;
  k ^= k >> 32;
  return k;
}

__device__ __forceinline__ u64 load_index(u64 k, int load) {
  if (load == QJ_LOAD_LADDER) return k ^ (k >> 1);
  if (load == QJ_LOAD_UNLADDER) return suffix_parity(k);
  return k;
}

template <typename R>
__device__ __forceinline__ cplx<R> cconj(cplx<R> a) {
  return mk(a.re, -a.im);
}

__device__ __forceinline__ float rfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double rfma(double a, double b, double c) { return fma(a, b, c); }

// Applies the 2x2 matrix m (row-major; its conjugate transpose if DAG) to
// local bit A of the 2^K registers x.
template <typename R, int K, int A, bool DAG>
__device__ __forceinline__ void rot_apply(cplx<R> (&x)[1 << K], const cplx<R>* m) {
  const cplx<R> m00 = DAG ? cconj(m[0]) : m[0];
  const cplx<R> m01 = DAG ? cconj(m[2]) : m[1];
  const cplx<R> m10 = DAG ? cconj(m[1]) : m[2];
  const cplx<R> m11 = DAG ? cconj(m[3]) : m[3];
#pragma unroll
  for (int r = 0; r < (1 << K); ++r) {
    if (r & (1 << A)) continue;
    const cplx<R> a = x[r], b = x[r | (1 << A)];
    x[r] = mac2(m00, a, m01, b);
    x[r | (1 << A)] = mac2(m10, a, m11, b);
  }
}

// The same for a step whose matrix is exp(-i theta G / 2) of generator g
// (1 X, 2 Y, 3 Z; 0: any 2x2), using the form g fixes: Y real, Z diagonal,
// X real on the diagonal and imaginary off it (4 multiply-adds per output
// amplitude instead of 8 for a general complex matrix).
template <typename R, int K, int A, bool DAG>
__device__ __forceinline__ void rot_apply_g(cplx<R> (&x)[1 << K], const cplx<R>* m, int g) {
  if (g == 2) {
    const R m00 = m[0].re, m01 = DAG ? m[2].re : m[1].re, m10 = DAG ? m[1].re : m[2].re, m11 = m[3].re;
#pragma unroll
    for (int r = 0; r < (1 << K); ++r) {
      if (r & (1 << A)) continue;
      const cplx<R> a = x[r], b = x[r | (1 << A)];
      x[r] = mk(rfma(m01, b.re, m00 * a.re), rfma(m01, b.im, m00 * a.im));
      x[r | (1 << A)] = mk(rfma(m11, b.re, m10 * a.re), rfma(m11, b.im, m10 * a.im));
    }
  } else if (g == 3) {
    const cplx<R> d0 = DAG ? cconj(m[0]) : m[0], d1 = DAG ? cconj(m[3]) : m[3];
#pragma unroll
    for (int r = 0; r < (1 << K); ++r) {
      if (r & (1 << A)) continue;
      x[r] = cmul(d0, x[r]);
      x[r | (1 << A)] = cmul(d1, x[r | (1 << A)]);
    }
  } else if (g == 1) {
    // m = [[c0, i s01], [i s10, c1]] (conjugate transpose: s01 <- -s10, s10 <- -s01)
    const R c0 = m[0].re, c1 = m[3].re;
    const R s01 = DAG ? -m[2].im : m[1].im, s10 = DAG ? -m[1].im : m[2].im;
#pragma unroll
    for (int r = 0; r < (1 << K); ++r) {
      if (r & (1 << A)) continue;
      const cplx<R> a = x[r], b = x[r | (1 << A)];
      x[r] = mk(rfma(-s01, b.im, c0 * a.re), rfma(s01, b.re, c0 * a.im));
      x[r | (1 << A)] = mk(rfma(c1, b.re, -s10 * a.im), rfma(c1, b.im, s10 * a.re));
    }
  } else {
    rot_apply<R, K, A, DAG>(x, m);
  }
}

// Im <lam|G|psi> over the group for G = X (1), Y (2) or Z (3) on local bit A:
// each pair's four-term sum in the state's precision (the products of two
// amplitudes), accumulated in FP64.
template <typename R, int K, int A>
__device__ __forceinline__ double rot_gen(const cplx<R> (&p)[1 << K], const cplx<R> (&l)[1 << K], int gen) {
  double acc = 0.0;
#pragma unroll
  for (int r = 0; r < (1 << K); ++r) {
    if (r & (1 << A)) continue;
    const int s = r | (1 << A);
    const R p0r = p[r].re, p0i = p[r].im, p1r = p[s].re, p1i = p[s].im;
    const R l0r = l[r].re, l0i = l[r].im, l1r = l[s].re, l1i = l[s].im;
    // Im(conj(l0) g0 + conj(l1) g1), (g0, g1) = G (p0, p1)
    R t;
    if (gen == 1)  // X: (p1, p0)
      t = rfma(l0r, p1i, rfma(-l0i, p1r, rfma(l1r, p0i, -l1i * p0r)));
    else if (gen == 2)  // Y: (-i p1, i p0)
      t = rfma(-l0r, p1r, rfma(-l0i, p1i, rfma(l1r, p0r, l1i * p0i)));
    else  // Z: (p0, -p1)
      t = rfma(l0r, p0i, rfma(-l0i, p0r, rfma(-l1r, p1i, l1i * p1r)));
    acc += (double)t;
  }
  return acc;
}

template <typename R, int K, bool ADJ>
struct RotBit {
  // Runs the step sequence of local bit A (and recurses to A + 1).
  template <int A>
  static __device__ __forceinline__ void run(cplx<R> (&x)[1 << K], cplx<R> (&y)[1 << K], const cplx<R>* m,
                                             const RotArgs& a, double (&acc)[K * QJ_ROT_MAX_STEPS]) {
    if constexpr (A < K) {
      if (!ADJ) {
#pragma unroll
        for (int j = 0; j < QJ_ROT_MAX_STEPS; ++j)
          if (j < a.nrot) rot_apply_g<R, K, A, false>(x, m + (A * a.nrot + j) * 4, a.gen[A][j]);
      } else {
#pragma unroll
        for (int jj = QJ_ROT_MAX_STEPS - 1; jj >= 0; --jj) {
          if (jj < a.nrot) {
            const int g = a.gen[A][jj];
            if (g) acc[A * QJ_ROT_MAX_STEPS + jj] += rot_gen<R, K, A>(x, y, g);
            rot_apply_g<R, K, A, true>(x, m + (A * a.nrot + jj) * 4, g);
            rot_apply_g<R, K, A, true>(y, m + (A * a.nrot + jj) * 4, g);
          }
        }
      }
      run<A + 1>(x, y, m, a, acc);
    }
  }
};

__device__ __forceinline__ double rot_block_sum(double v, double* wsum) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < kRotBlock / 64; ++i) t += wsum[i];
  __syncthreads();
  return t;
}

// Register r of the group whose first index is `base`: load from `in`
// through the pass's load permutation.  V = 2 pairs (r, r+1) differ in
// position 0; both permutations map an aligned pair onto an aligned pair,
// swapped when the image of the even index is odd.
template <typename R, int K, int V>
__device__ __forceinline__ void rot_load(cplx<R> (&x)[1 << K], const cplx<R>* in, u64 base, const RotArgs& a) {
  if (a.load == QJ_LOAD_ZERO) {
#pragma unroll
    for (int r = 0; r < (1 << K); ++r) x[r] = mk<R>(R(base + a.off[r] == 0 ? 1 : 0), R(0));
    return;
  }
#pragma unroll
  for (int r = 0; r < (1 << K); r += V) {
    const u64 src = load_index(base + a.off[r], a.load);
    if (V == 2) {
      const cvec<R, 2> v = ldv<R, 2>(in + (src & ~u64(1)));
      const bool sw = src & 1;
      x[r] = sw ? v.v[1] : v.v[0];
      x[r + 1] = sw ? v.v[0] : v.v[1];
    } else {
      x[r] = in[src];
    }
  }
}

template <typename R, int K, int V>
__device__ __forceinline__ void rot_store(const cplx<R> (&x)[1 << K], cplx<R>* out, u64 base, const RotArgs& a) {
#pragma unroll
  for (int r = 0; r < (1 << K); r += V) {
    if (V == 2) {
      cvec<R, 2> v;
      v.v[0] = x[r];
      v.v[1] = x[r + 1];
      stv<R, 2>(out + base + a.off[r], v);
    } else {
      out[base + a.off[r]] = x[r];
    }
  }
}

template <typename R, int K, bool ADJ, int V>
__global__ void __launch_bounds__(kRotBlock)
    k_rot(const cplx<R>* psi_in, cplx<R>* psi_out, const cplx<R>* lam_in, cplx<R>* lam_out,
          const int32_t* __restrict__ sel, const cplx<R>* __restrict__ mats, double* __restrict__ partials,
          u64 ngroups, RotArgs a) {
  __shared__ double wsum[kRotBlock / 64];
  const int b = blockIdx.y;
  const u64 s = sel ? (u64)sel[b] : (u64)b;
  const cplx<R>* pin = psi_in + s * a.stride;
  cplx<R>* pout = psi_out + s * a.stride;
  const cplx<R>* lin = ADJ ? lam_in + s * a.stride : nullptr;
  cplx<R>* lout = ADJ ? lam_out + s * a.stride : nullptr;
  const cplx<R>* m = mats + s * a.mstride;
  double acc[K * QJ_ROT_MAX_STEPS];
#pragma unroll
  for (int i = 0; i < K * QJ_ROT_MAX_STEPS; ++i) acc[i] = 0.0;
  const u64 step = (u64)gridDim.x * kRotBlock;
  for (u64 g = (u64)blockIdx.x * kRotBlock + threadIdx.x; g < ngroups; g += step) {
    const u64 base = insert_bits(g, a.ins);
    cplx<R> x[1 << K], y[1 << K];
    rot_load<R, K, V>(x, pin, base, a);
    if (ADJ) rot_load<R, K, V>(y, lin, base, a);
    RotBit<R, K, ADJ>::template run<0>(x, y, m, a, acc);
    rot_store<R, K, V>(x, pout, base, a);
    if (ADJ) rot_store<R, K, V>(y, lout, base, a);
  }
  if (ADJ) {
    const u64 cols = gridDim.x, rows = (u64)a.nsel;
#pragma unroll
    for (int A = 0; A < K; ++A)
#pragma unroll
      for (int j = 0; j < QJ_ROT_MAX_STEPS; ++j)
        if (j < a.nrot && a.gen[A][j]) {
          const double t = rot_block_sum(acc[A * QJ_ROT_MAX_STEPS + j], wsum);
          if (threadIdx.x == 0) partials[((u64)(A * a.nrot + j) * rows + b) * cols + blockIdx.x] = t;
        }
  }
}

// ---------------------------------------------------------------- TFIM --
// For H = -sum_{j} Z_j Z_{j+1} - g sum_q X_q (open chain over bits), row b:
// out = H' psi and partials[b][block] of Re<psi|H' psi>, where H' = H, or
// with `ladder` set H' = L^+ H L for the CNOT ladder L (psi is the state
// before the ladder): there the diagonal reads bits of P(k) and X on bit q
// flips bits q and q-1.  out == NULL computes the energy only.
// Each thread takes V consecutive amplitudes: a flip of bits at or above
// log2(V) reads the partner group with one vector load, a flip inside the
// group pairs amplitudes already in registers.
template <typename R, int V>
__global__ void __launch_bounds__(kRotBlock)
    k_tfim_batch(const cplx<R>* psi, cplx<R>* out, const int32_t* __restrict__ sel, u64 stride, int nq, double g,
                 int ladder, double* __restrict__ partials) {
  __shared__ double wsum[kRotBlock / 64];
  const u64 n = u64(1) << nq;
  const int b = blockIdx.y;
  const u64 s = sel ? (u64)sel[b] : (u64)b;
  const cplx<R>* p = psi + s * stride;
  cplx<R>* o = out ? out + s * stride : nullptr;
  const u64 zzmask = (u64(1) << (nq - 1)) - 1;
  double acc = 0.0;
  const u64 step = (u64)gridDim.x * kRotBlock * V;
  for (u64 k0 = ((u64)blockIdx.x * kRotBlock + threadIdx.x) * V; k0 < n; k0 += step) {
    const cvec<R, V> xv = ldv<R, V>(p + k0);
    double zz[V];
#pragma unroll
    for (int j = 0; j < V; ++j) {
      const u64 kd = ladder ? suffix_parity(k0 + j) : k0 + j;
      zz[j] = (double)(nq - 1) - 2.0 * (double)__popcll((kd ^ (kd >> 1)) & zzmask);
    }
    if (o) {
      double re[V], im[V];
#pragma unroll
      for (int j = 0; j < V; ++j) {
        re[j] = -zz[j] * (double)xv.v[j].re;
        im[j] = -zz[j] * (double)xv.v[j].im;
      }
      for (int q = 0; q < nq; ++q) {
        const u64 f = (u64(1) << q) | (ladder && q ? u64(1) << (q - 1) : 0);
        const u64 fo = f & ~u64(V - 1);
        const int fi = (int)(f & (V - 1));
        const cvec<R, V> yv = fo ? ldv<R, V>(p + (k0 ^ fo)) : xv;
#pragma unroll
        for (int j = 0; j < V; ++j) {
          re[j] -= g * (double)yv.v[j ^ fi].re;
          im[j] -= g * (double)yv.v[j ^ fi].im;
        }
      }
      cvec<R, V> ov;
#pragma unroll
      for (int j = 0; j < V; ++j) {
        ov.v[j] = mk<R>((R)re[j], (R)im[j]);
        acc += (double)xv.v[j].re * re[j] + (double)xv.v[j].im * im[j];
      }
      stv<R, V>(o + k0, ov);
    } else {
      // each unordered pair {k, k^f} once, from its lower index
      double xs[V];
#pragma unroll
      for (int j = 0; j < V; ++j) xs[j] = 0.0;
      for (int q = 0; q < nq; ++q) {
        const u64 f = (u64(1) << q) | (ladder && q ? u64(1) << (q - 1) : 0);
        const u64 fo = f & ~u64(V - 1);
        const int fi = (int)(f & (V - 1));
        if (fo && ((k0 >> q) & 1)) continue;  // bit q (>= log2 V) is 1 for the whole group
        const cvec<R, V> yv = fo ? ldv<R, V>(p + (k0 ^ fo)) : xv;
#pragma unroll
        for (int j = 0; j < V; ++j)
          if (fo || !((j >> q) & 1))
            xs[j] += (double)xv.v[j].re * (double)yv.v[j ^ fi].re + (double)xv.v[j].im * (double)yv.v[j ^ fi].im;
      }
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const double xr = xv.v[j].re, xi = xv.v[j].im;
        acc += -zz[j] * (xr * xr + xi * xi) - 2.0 * g * xs[j];
      }
    }
  }
  acc = rot_block_sum(acc, wsum);
  if (threadIdx.x == 0) partials[(u64)b * gridDim.x + blockIdx.x] = acc;
}

// out[r] = sum_c in[r * ncols + c], one workgroup per row, fixed order.
__global__ void __launch_bounds__(kRotBlock) k_reduce_rows(const double* __restrict__ in, u64 ncols,
                                                           double* __restrict__ out) {
  __shared__ double wsum[kRotBlock / 64];
  const u64 r = blockIdx.x;
  double acc = 0.0;
  for (u64 c = threadIdx.x; c < ncols; c += kRotBlock) acc += in[r * ncols + c];
  acc = rot_block_sum(acc, wsum);
  if (threadIdx.x == 0) out[r] = acc;
}

// -------------------------------------------------- LDS-tiled passes ----
// complex64 only.  A workgroup of 256 threads owns tiles of 2^12 amplitudes
// (tile bit j = global bit tpos[j], ascending; the lowest tile bits should be
// the lowest global bits so that loads are whole lines).  Each thread holds
// 16 amplitudes; three register phases cover the tile bits 8..11, 4..7 and
// 0..3, with LDS transposes in between:
//   phase A: thread bits t[0..7] = tid,          registers t[8..11]
//   phase B: thread bits t[0..3], t[8..11],      registers t[4..7]
//   phase C: thread bits t[4..11],               registers t[0..3]
// Loads and stores use the phase-A layout (lanes = t[0..5]).  LDS address
// p(t) = t ^ ((t >> 4) & 31) spreads every phase's 64 lanes over all banks.
// Tile bits outside `rotmask` are carried but not rotated (they were done by
// an earlier pass of the layer).
constexpr int kTileBits = 12;
constexpr int kTileBlock = 256;

struct RotTileArgs {
  Ins ins;                 // tile positions, for insert_bits of the tile id
  u64 offr[16];            // offset of register r in phase A (tile bits 8..11)
  int tpos[kTileBits];     // global bit of tile bit j
  int rotmask;             // tile bits rotated by this pass
  int gen[kTileBits][QJ_ROT_MAX_STEPS];
  int nrot, load;
  u64 stride, mstride;
  int nsel;
};

__device__ __forceinline__ unsigned tile_swz(unsigned t) { return t ^ ((t >> 4) & 31u); }

template <int PH>
__device__ __forceinline__ unsigned tile_t(unsigned tid, int r) {
  if (PH == 0) return tid | ((unsigned)r << 8);
  if (PH == 1) return (tid & 15u) | ((unsigned)r << 4) | ((tid >> 4) << 8);
  return (unsigned)r | (tid << 4);
}

// Rotations of the phase's four register bits (tile bits B0..B0+3); NROT
// steps per bit (compile time, so the FP64 accumulators stay in registers).
// ZL: the last step of every rotated bit is a Z rotation done for the
// whole tile by tile_zlast (before these steps in adjoint mode, after them
// forward), so it is skipped here.
template <bool ADJ, int NROT, int B0, int A, bool ZL = false>
__device__ __forceinline__ void tile_rot(cplx<float> (&x)[16], cplx<float> (&y)[16], const cplx<float>* m,
                                         const RotTileArgs& a, double (&acc)[kTileBits * NROT]) {
  if constexpr (A < 4) {
    constexpr int j = B0 + A;
    if ((a.rotmask >> j) & 1) {
      if (!ADJ) {
#pragma unroll
        for (int s = 0; s < NROT - (ZL ? 1 : 0); ++s)
          rot_apply_g<float, 4, A, false>(x, m + (j * NROT + s) * 4, a.gen[j][s]);
      } else {
#pragma unroll
        for (int s = NROT - 1 - (ZL ? 1 : 0); s >= 0; --s) {
          const int g = a.gen[j][s];
          if (g) acc[j * NROT + s] += rot_gen<float, 4, A>(x, y, g);
          rot_apply_g<float, 4, A, true>(x, m + (j * NROT + s) * 4, g);
          rot_apply_g<float, 4, A, true>(y, m + (j * NROT + s) * 4, g);
        }
      }
    }
    tile_rot<ADJ, NROT, B0, A + 1, ZL>(x, y, m, a, acc);
  }
}

// Gradients Im<lam|Z_j|psi> = sum over k of (-1)^(k_j) Im(conj(lam_k) psi_k)
// of the last (Z) steps of every rotated bit, from x = psi and y = lam in
// the phase-A layout (thread bits t[0..7] = tid, register bits t[8..11] = r).
template <int NROT>
__device__ __forceinline__ void tile_zgrad(const cplx<float> (&x)[16], const cplx<float> (&y)[16],
                                           const RotTileArgs& a, double (&acc)[kTileBits * NROT]) {
  float c[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) c[r] = rfma(y[r].re, x[r].im, -y[r].im * x[r].re);
  // thread bits: the thread's total with its bit's sign
  float tot = c[0];
#pragma unroll
  for (int r = 1; r < 16; ++r) tot += c[r];
#pragma unroll
  for (int b = 0; b < 8; ++b)
    if (((a.rotmask >> b) & 1) && a.gen[b][NROT - 1]) acc[b * NROT + NROT - 1] += (double)(((threadIdx.x >> b) & 1) ? -tot : tot);
  // register bits 8..11
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int j = 8 + u;
    if (((a.rotmask >> j) & 1) && a.gen[j][NROT - 1]) {
      float sg = 0.f;
#pragma unroll
      for (int r = 0; r < 16; ++r) sg += ((r >> u) & 1) ? -c[r] : c[r];
      acc[j * NROT + NROT - 1] += (double)sg;
    }
  }
}

// Steps that are Z rotations (diagonal) and come last for every rotated
// bit: rotations on different bits commute, so these steps of all 12 bits
// are one diagonal (the product over the bits of d_j[k_j], conjugated in
// adjoint mode) applied once per amplitude in the phase-A layout instead
// of per bit -- after the other steps forward, before them in adjoint mode,
// where the gradients of these steps are taken first (tile_zgrad).  gq: the
// thread's factor times the register bits' bit-0 factors; q[u]: the ratio
// a set register bit u adds (bit-1 over bit-0 factor).
template <bool ADJ, int NROT>
__device__ __forceinline__ void tile_zlast(cplx<float> (&x)[16], cplx<float> (&y)[16], const RotTileArgs& a,
                                           double (&acc)[kTileBits * NROT], cplx<float> gq, const cplx<float> (&q)[4]) {
  if constexpr (ADJ) tile_zgrad<NROT>(x, y, a, acc);
  // registers 0..7 (bit 3 clear) from a product tree over bits 0..2, then
  // the same factors times q[3] for registers 8..15
  cplx<float> h[8];
  h[0] = gq;
  h[1] = cmul(gq, q[0]);
  h[2] = cmul(gq, q[1]);
  h[3] = cmul(h[1], q[1]);
#pragma unroll
  for (int r = 0; r < 4; ++r) h[r + 4] = cmul(h[r], q[2]);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    x[r] = cmul(h[r], x[r]);
    if (ADJ) y[r] = cmul(h[r], y[r]);
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const cplx<float> hh = cmul(h[r], q[3]);
    x[r + 8] = cmul(hh, x[r + 8]);
    if (ADJ) y[r + 8] = cmul(hh, y[r + 8]);
  }
}

template <int PH>
__device__ __forceinline__ void lds_put(cplx<float>* sm, const cplx<float> (&x)[16]) {
#pragma unroll
  for (int r = 0; r < 16; ++r) sm[tile_swz(tile_t<PH>(threadIdx.x, r))] = x[r];
}
template <int PH>
__device__ __forceinline__ void lds_get(const cplx<float>* sm, cplx<float> (&x)[16]) {
#pragma unroll
  for (int r = 0; r < 16; ++r) x[r] = sm[tile_swz(tile_t<PH>(threadIdx.x, r))];
}

// Moves x from layout P to layout Q through the LDS buffer (and y at the
// same time through a second tile buffer, in adjoint mode: two barriers per
// transpose for both states).
template <bool ADJ, int P, int Q>
__device__ __forceinline__ void tile_transpose(cplx<float>* sm, cplx<float> (&x)[16], cplx<float> (&y)[16]) {
  lds_put<P>(sm, x);
  if (ADJ) lds_put<P>(sm + (1 << kTileBits), y);
  __syncthreads();
  lds_get<Q>(sm, x);
  if (ADJ) lds_get<Q>(sm + (1 << kTileBits), y);
  __syncthreads();  // the buffers are free again
}

template <bool ADJ, int NROT, bool ZL = false>
__global__ void __launch_bounds__(kTileBlock)
    k_rot_tile(const cplx<float>* psi_in, cplx<float>* psi_out, const cplx<float>* lam_in, cplx<float>* lam_out,
               const int32_t* __restrict__ sel, const cplx<float>* __restrict__ mats,
               double* __restrict__ partials, u64 ntiles, RotTileArgs a) {
  // one 32 KiB tile buffer per state (the kernel's only LDS); the gradient
  // reduction reuses it after the last tile
  extern __shared__ __align__(16) unsigned char tile_raw[];
  cplx<float>* sm = reinterpret_cast<cplx<float>*>(tile_raw);
  double* wsum = reinterpret_cast<double*>(tile_raw);
  const int b = blockIdx.y;
  const u64 s = sel ? (u64)sel[b] : (u64)b;
  const cplx<float>* pin = psi_in + s * a.stride;
  cplx<float>* pout = psi_out + s * a.stride;
  const cplx<float>* lin = ADJ ? lam_in + s * a.stride : nullptr;
  cplx<float>* lout = ADJ ? lam_out + s * a.stride : nullptr;
  const cplx<float>* m = mats + s * a.mstride;
  double acc[kTileBits * NROT];
#pragma unroll
  for (int i = 0; i < kTileBits * NROT; ++i) acc[i] = 0.0;
  // offset of this thread's tile bits 0..7 (phase-A thread bits)
  u64 offt = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if ((threadIdx.x >> j) & 1) offt |= u64(1) << a.tpos[j];
  // ZL: the diagonal undo's thread factor (times the register bits'
  // bit-0 factors) and the register bits' ratios, for every tile
  // (products taken in FP64, rounded once)
  cplx<float> gq = mk<float>(1.f, 0.f), q[4];
  if constexpr (ZL) {
    cplx<double> g = mk<double>(1.0, 0.0);
#pragma unroll
    for (int j = 0; j < kTileBits; ++j) {
      if (!((a.rotmask >> j) & 1)) {
        if (j >= 8) q[j - 8] = mk<float>(1.f, 0.f);
        continue;
      }
      const cplx<float>* d = m + (j * NROT + NROT - 1) * 4;  // diag(d[0], d[3])
      const cplx<double> d0 = mk<double>(d[0].re, d[0].im), d1 = mk<double>(d[3].re, d[3].im);
      const cplx<double> e0 = ADJ ? cconj(d0) : d0, e1 = ADJ ? cconj(d1) : d1;
      if (j < 8) {
        g = cmul(g, ((threadIdx.x >> j) & 1) ? e1 : e0);
      } else {
        g = cmul(g, e0);
        const cplx<double> r = cmul(e1, cconj(e0));  // e1 / e0 (|e0| = 1)
        q[j - 8] = mk<float>((float)r.re, (float)r.im);
      }
    }
    gq = mk<float>((float)g.re, (float)g.im);
  }
  for (u64 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const u64 base = insert_bits(tile, a.ins) + offt;
    cplx<float> x[16], y[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const u64 gidx = base + a.offr[r];
      if (a.load == QJ_LOAD_ZERO) {
        x[r] = mk<float>(gidx == 0 ? 1.f : 0.f, 0.f);
      } else {
        const u64 src = load_index(gidx, a.load);
        x[r] = pin[src];
        if (ADJ) y[r] = lin[src];
      }
    }
    if constexpr (ZL && ADJ) tile_zlast<ADJ, NROT>(x, y, a, acc, gq, q);
    tile_rot<ADJ, NROT, 8, 0, ZL>(x, y, m, a, acc);
    tile_transpose<ADJ, 0, 1>(sm, x, y);
    tile_rot<ADJ, NROT, 4, 0, ZL>(x, y, m, a, acc);
    tile_transpose<ADJ, 1, 2>(sm, x, y);
    tile_rot<ADJ, NROT, 0, 0, ZL>(x, y, m, a, acc);
    tile_transpose<ADJ, 2, 0>(sm, x, y);
    if constexpr (ZL && !ADJ) tile_zlast<ADJ, NROT>(x, y, a, acc, gq, q);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      pout[base + a.offr[r]] = x[r];
      if (ADJ) lout[base + a.offr[r]] = y[r];
    }
  }
  if (ADJ) {
    const u64 cols = gridDim.x, rows = (u64)a.nsel;
#pragma unroll
    for (int j = 0; j < kTileBits; ++j)
#pragma unroll
      for (int st = 0; st < NROT; ++st)
        if (((a.rotmask >> j) & 1) && a.gen[j][st]) {
          const double t = rot_block_sum(acc[j * NROT + st], wsum);
          if (threadIdx.x == 0) partials[((u64)(j * NROT + st) * rows + b) * cols + blockIdx.x] = t;
        }
  }
}

// ------------------------------------------------ whole circuit in LDS --
// Persistent kernel for small ansatz circuits: workgroup w runs circuits
// w, w + gridDim.x, ... entirely in LDS (2^n amplitudes, n <= nqmax) and
// writes only their energies.  Per layer: rotations on interleaved groups
// of up to 4 bits (group i = bits i, i+G, i+2G, i+3G), then the ladder as an
// in-place gather through registers, then after the last layer the TFIM
// energy in FP64.  desc[c] = {nqubits, nlayers, matrix offset lo, hi}; the
// circuit's matrices are [layer][bit][2x2] (bit p = qubit n-1-p).  A circuit
// whose descriptor does not fit (n outside [2, nqmax], no layers, matrices
// past mats_len) gets a NaN energy and is not run.
constexpr int kLdsBlock = 512;
// amplitudes per thread in the ladder gather: 64 KiB of LDS / 512 threads
template <typename R>
constexpr int lds_per_thread() {
  return 65536 / (int)sizeof(cplx<R>) / kLdsBlock;
}

template <typename R, int K>
__device__ __forceinline__ void lds_group(cplx<R>* st, int n, int i, int G, const cplx<R>* ml) {
  int p[K];
#pragma unroll
  for (int A = 0; A < K; ++A) p[A] = i + A * G;
  const unsigned ngroups = 1u << (n - K);
  for (unsigned g = threadIdx.x; g < ngroups; g += kLdsBlock) {
    unsigned idx = g;
#pragma unroll
    for (int A = 0; A < K; ++A) {
      const unsigned lo = idx & ((1u << p[A]) - 1);
      idx = ((idx ^ lo) << 1) | lo;
    }
    cplx<R> x[1 << K];
#pragma unroll
    for (int r = 0; r < (1 << K); ++r) {
      unsigned off = 0;
#pragma unroll
      for (int A = 0; A < K; ++A)
        if (r & (1 << A)) off |= 1u << p[A];
      x[r] = st[idx + off];
    }
    if constexpr (K >= 1) rot_apply<R, K, 0, false>(x, ml + p[0] * 4);
    if constexpr (K >= 2) rot_apply<R, K, (K >= 2 ? 1 : 0), false>(x, ml + p[K >= 2 ? 1 : 0] * 4);
    if constexpr (K >= 3) rot_apply<R, K, (K >= 3 ? 2 : 0), false>(x, ml + p[K >= 3 ? 2 : 0] * 4);
    if constexpr (K >= 4) rot_apply<R, K, (K >= 4 ? 3 : 0), false>(x, ml + p[K >= 4 ? 3 : 0] * 4);
#pragma unroll
    for (int r = 0; r < (1 << K); ++r) {
      unsigned off = 0;
#pragma unroll
      for (int A = 0; A < K; ++A)
        if (r & (1 << A)) off |= 1u << p[A];
      st[idx + off] = x[r];
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(kLdsBlock)
    k_ansatz_lds(const int32_t* __restrict__ desc, int ncirc, const cplx<R>* __restrict__ mats, u64 mats_len,
                 double g, double* __restrict__ energies, int nqmax) {
  // the state is the only LDS (dynamic, exactly 2^nqmax amplitudes); the
  // energy's wave partials reuse its first bytes once it is no longer read
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cplx<R>* st = reinterpret_cast<cplx<R>*>(lds_raw);
  double* wsum = reinterpret_cast<double*>(lds_raw);
  for (int c = blockIdx.x; c < ncirc; c += gridDim.x) {
    const int n = desc[4 * c], L = desc[4 * c + 1];
    const u64 moff = (u64)(uint32_t)desc[4 * c + 2] | ((u64)(uint32_t)desc[4 * c + 3] << 32);
    const bool ok = n >= 2 && n <= nqmax && L >= 1 && moff <= mats_len && (u64)L * n * 4 <= mats_len - moff;
    if (!ok) {
      if (threadIdx.x == 0) energies[c] = __builtin_nan("");
      continue;  // uniform per workgroup
    }
    const cplx<R>* m = mats + moff;
    const unsigned dim = 1u << n;
    for (unsigned k = threadIdx.x; k < dim; k += kLdsBlock) st[k] = mk<R>(R(k == 0 ? 1 : 0), R(0));
    __syncthreads();
    const int G = (n + 3) / 4;
    for (int l = 0; l < L; ++l) {
      const cplx<R>* ml = m + (u64)l * n * 4;
      for (int i = 0; i < G; ++i) {
        const int K = (n - 1 - i) / G + 1;  // bits i + A*G < n
        switch (K) {
          case 1: lds_group<R, 1>(st, n, i, G, ml); break;
          case 2: lds_group<R, 2>(st, n, i, G, ml); break;
          case 3: lds_group<R, 3>(st, n, i, G, ml); break;
          default: lds_group<R, 4>(st, n, i, G, ml); break;
        }
        __syncthreads();
      }
      // ladder: x[k] = st[k ^ (k >> 1)], gathered to registers first
      constexpr int E = lds_per_thread<R>();
      R xr[E], xi[E];  // split re/im: whole-struct copies out of LDS stay in scratch
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const unsigned k = threadIdx.x + j * kLdsBlock;
        const unsigned kk = k < dim ? k : 0;  // in-bounds read; only k < dim is stored
        const unsigned src = kk ^ (kk >> 1);
        xr[j] = st[src].re;
        xi[j] = st[src].im;
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < E; ++j) {
        const unsigned k = threadIdx.x + j * kLdsBlock;
        if (k < dim) st[k] = mk<R>(xr[j], xi[j]);
      }
      __syncthreads();
    }
    const unsigned zzmask = (1u << (n - 1)) - 1;
    double acc = 0.0;
    for (unsigned k = threadIdx.x; k < dim; k += kLdsBlock) {
      const cplx<R> a = st[k];
      const double ar = a.re, ai = a.im;
      const double zz = (double)(n - 1) - 2.0 * (double)__popc((k ^ (k >> 1)) & zzmask);
      double xs = 0.0;
      for (int q = 0; q < n; ++q)
        if (!(k & (1u << q))) {
          const cplx<R> b = st[k ^ (1u << q)];
          xs += ar * (double)b.re + ai * (double)b.im;
        }
      acc += -zz * (ar * ar + ai * ai) - 2.0 * g * xs;
    }
    // block sum over 512 threads
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    __syncthreads();  // every thread is done reading st
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < kLdsBlock / 64; ++w) t += wsum[w];
      energies[c] = t;
    }
    __syncthreads();
  }
}

template <typename R, int K, bool ADJ>
static void launch_rot(bool vec, dim3 gd, hipStream_t st, const cplx<R>* pin, cplx<R>* pout, const cplx<R>* lin,
                       cplx<R>* lout, const int32_t* sel, const cplx<R>* mats, double* partials, u64 ngroups,
                       const RotArgs& a) {
  if (vec)
    k_rot<R, K, ADJ, 2><<<gd, dim3(kRotBlock), 0, st>>>(pin, pout, lin, lout, sel, mats, partials, ngroups, a);
  else
    k_rot<R, K, ADJ, 1><<<gd, dim3(kRotBlock), 0, st>>>(pin, pout, lin, lout, sel, mats, partials, ngroups, a);
}

template <typename R, bool ADJ>
static void launch_rot_k(int k, bool vec, dim3 gd, hipStream_t st, const cplx<R>* pin, cplx<R>* pout,
                         const cplx<R>* lin, cplx<R>* lout, const int32_t* sel, const cplx<R>* mats,
                         double* partials, u64 ngroups, const RotArgs& a) {
  switch (k) {
    case 1: launch_rot<R, 1, ADJ>(vec, gd, st, pin, pout, lin, lout, sel, mats, partials, ngroups, a); break;
    case 2: launch_rot<R, 2, ADJ>(vec, gd, st, pin, pout, lin, lout, sel, mats, partials, ngroups, a); break;
    case 3: launch_rot<R, 3, ADJ>(vec, gd, st, pin, pout, lin, lout, sel, mats, partials, ngroups, a); break;
    case 4: launch_rot<R, 4, ADJ>(vec, gd, st, pin, pout, lin, lout, sel, mats, partials, ngroups, a); break;
    default:
      // five bits of psi and lam: complex64 only, and only the 8-byte form
      // (the 16-byte form spills)
      if constexpr (!ADJ)
        launch_rot<R, 5, ADJ>(vec, gd, st, pin, pout, lin, lout, sel, mats, partials, ngroups, a);
      else if constexpr (sizeof(R) == 4)
        k_rot<R, 5, true, 1><<<gd, dim3(kRotBlock), 0, st>>>(pin, pout, lin, lout, sel, mats, partials, ngroups, a);
      break;
  }
}

static int check_rows(int nrows, const int32_t* sel, uint64_t stride, int nqubits, const char* what) {
  if (nrows < 1 || nrows > 65535) return set_error(QJ_EVALUE, "%s: bad number of states %d", what, nrows);
  if (stride < (u64(1) << nqubits) && !(sel == nullptr && nrows == 1))
    return set_error(QJ_EVALUE, "%s: state stride %llu below 2^nqubits", what, (unsigned long long)stride);
  return QJ_OK;
}

template <typename R>
static int rot_pass_impl(const void* psi_in, void* psi_out, const void* lam_in, void* lam_out, int nqubits,
                         uint64_t stride, const int32_t* sel, int nrows, const int32_t* bits, int k, int nrot,
                         const int32_t* gens, int load, const void* mats, uint64_t mstride, double* partials,
                         int ncols, hipStream_t st) {
  const char* what = "rot_pass";
  if (nqubits < 1 || nqubits > 40) return set_error(QJ_EVALUE, "%s: bad nqubits %d", what, nqubits);
  int rc = check_rows(nrows, sel, stride, nqubits, what);
  if (rc) return rc;
  if (k < 1 || k > QJ_ROT_MAX_BITS || k > nqubits) return set_error(QJ_EVALUE, "%s: bad group size %d", what, k);
  if (nrot < 1 || nrot > QJ_ROT_MAX_STEPS) return set_error(QJ_EVALUE, "%s: bad step count %d", what, nrot);
  if (mstride < (u64)(k * nrot * 4) && !(sel == nullptr && nrows == 1))
    return set_error(QJ_EVALUE, "%s: matrix stride below %d", what, k * nrot * 4);
  if (load < QJ_LOAD_PLAIN || load > QJ_LOAD_ZERO) return set_error(QJ_EVALUE, "%s: bad load mode %d", what, load);
  if (ncols < 1 || ncols > kRotMaxCols) return set_error(QJ_EVALUE, "%s: bad column count %d", what, ncols);
  const bool adj = lam_in != nullptr;
  if (adj != (lam_out != nullptr)) return set_error(QJ_EVALUE, "%s: lam_in and lam_out go together", what);
  const int max_adj = sizeof(R) == 4 ? QJ_ROT_MAX_ADJ_BITS : QJ_ROT_MAX_ADJ_BITS - 1;
  if (adj && k > max_adj) return set_error(QJ_EVALUE, "%s: adjoint groups hold at most %d bits", what, max_adj);
  if (load == QJ_LOAD_ZERO && adj) return set_error(QJ_EVALUE, "%s: zero load is forward only", what);
  if ((load == QJ_LOAD_LADDER || load == QJ_LOAD_UNLADDER) && (psi_in == psi_out || (adj && lam_in == lam_out)))
    return set_error(QJ_EVALUE, "%s: a permuted load needs distinct input and output", what);
  for (int j = 0; j < k; ++j) {
    if (bits[j] < 0 || bits[j] >= nqubits) return set_error(QJ_EVALUE, "%s: bit %d out of range", what, bits[j]);
    if (j && bits[j] <= bits[j - 1]) return set_error(QJ_EVALUE, "%s: bits must be ascending and unique", what);
  }
  RotArgs a;
  std::memset(&a, 0, sizeof(a));
  bool any_gen = false;
  if (gens) {
    for (int A = 0; A < k; ++A)
      for (int j = 0; j < nrot; ++j) {
        const int g = gens[A * nrot + j];
        if (g < 0 || g > 3) return set_error(QJ_EVALUE, "%s: bad generator %d", what, g);
        a.gen[A][j] = g;
        any_gen |= g != 0;
      }
  }
  if (adj && any_gen && !partials) return set_error(QJ_EVALUE, "%s: gradients need a partials buffer", what);
  a.ins.npos = k;
  for (int j = 0; j < k; ++j) a.ins.pos[j] = bits[j];
  for (int r = 0; r < (1 << k); ++r) {
    u64 o = 0;
    for (int j = 0; j < k; ++j)
      if ((r >> j) & 1) o |= u64(1) << bits[j];
    a.off[r] = o;
  }
  a.nrot = nrot;
  a.load = load;
  a.stride = stride;
  a.mstride = mstride;
  a.nsel = nrows;
  const u64 ngroups = (u64(1) << nqubits) >> k;
  dim3 gd((unsigned)ncols, (unsigned)nrows);
  const bool vec = sizeof(R) == 4 && bits[0] == 0;
  const cplx<R>* m = static_cast<const cplx<R>*>(mats);
  if (adj)
    launch_rot_k<R, true>(k, vec, gd, st, (const cplx<R>*)psi_in, (cplx<R>*)psi_out, (const cplx<R>*)lam_in,
                          (cplx<R>*)lam_out, sel, m, partials, ngroups, a);
  else
    launch_rot_k<R, false>(k, vec, gd, st, (const cplx<R>*)psi_in, (cplx<R>*)psi_out, nullptr, nullptr, sel, m,
                           nullptr, ngroups, a);
  return check_launch("rot_pass launch");
}

}  // namespace qj

using namespace qj;

extern "C" {

int qj_rot_pass(const void* psi_in, void* psi_out, const void* lam_in, void* lam_out, int dtype, int nqubits,
                uint64_t stride, const int32_t* sel, int nrows, const int32_t* bits, int nbits, int nrot,
                const int32_t* gens, int load, const void* mats, uint64_t mstride, double* partials, int ncols,
                void* stream) {
  if ((!psi_in && load != QJ_LOAD_ZERO) || !psi_out || !bits || !mats)
    return set_error(QJ_EVALUE, "rot_pass: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64)
    return rot_pass_impl<float>(psi_in, psi_out, lam_in, lam_out, nqubits, stride, sel, nrows, bits, nbits, nrot,
                                gens, load, mats, mstride, partials, ncols, st);
  if (dtype == QJ_COMPLEX128)
    return rot_pass_impl<double>(psi_in, psi_out, lam_in, lam_out, nqubits, stride, sel, nrows, bits, nbits, nrot,
                                 gens, load, mats, mstride, partials, ncols, st);
  return set_error(QJ_EVALUE, "rot_pass: dtype must be complex64/complex128");
}

int qj_rot_tile_pass(const void* psi_in, void* psi_out, const void* lam_in, void* lam_out, int dtype, int nqubits,
                     uint64_t stride, const int32_t* sel, int nrows, const int32_t* tile_bits, int rotmask, int nrot,
                     const int32_t* gens, int load, const void* mats, uint64_t mstride, double* partials, int ncols,
                     void* stream) {
  const char* what = "rot_tile_pass";
  if ((!psi_in && load != QJ_LOAD_ZERO) || !psi_out || !tile_bits || !mats)
    return set_error(QJ_EVALUE, "%s: null pointer", what);
  if (dtype != QJ_COMPLEX64) return set_error(QJ_EVALUE, "%s: complex64 only", what);
  if (nqubits < kTileBits || nqubits > 40) return set_error(QJ_EVALUE, "%s: needs 12..40 qubits, got %d", what, nqubits);
  int rc = check_rows(nrows, sel, stride, nqubits, what);
  if (rc) return rc;
  if (nrot < 1 || nrot > QJ_ROT_MAX_STEPS) return set_error(QJ_EVALUE, "%s: bad step count %d", what, nrot);
  if (rotmask <= 0 || rotmask >= (1 << kTileBits)) return set_error(QJ_EVALUE, "%s: bad rotation mask", what);
  if (mstride < (u64)(kTileBits * nrot * 4) && !(sel == nullptr && nrows == 1))
    return set_error(QJ_EVALUE, "%s: matrix stride below %d", what, kTileBits * nrot * 4);
  if (load < QJ_LOAD_PLAIN || load > QJ_LOAD_ZERO) return set_error(QJ_EVALUE, "%s: bad load mode %d", what, load);
  if (ncols < 1 || ncols > kRotMaxCols) return set_error(QJ_EVALUE, "%s: bad column count %d", what, ncols);
  const bool adj = lam_in != nullptr;
  if (adj != (lam_out != nullptr)) return set_error(QJ_EVALUE, "%s: lam_in and lam_out go together", what);
  if (load == QJ_LOAD_ZERO && adj) return set_error(QJ_EVALUE, "%s: zero load is forward only", what);
  if ((load == QJ_LOAD_LADDER || load == QJ_LOAD_UNLADDER) && (psi_in == psi_out || (adj && lam_in == lam_out)))
    return set_error(QJ_EVALUE, "%s: a permuted load needs distinct input and output", what);
  for (int j = 0; j < kTileBits; ++j) {
    if (tile_bits[j] < 0 || tile_bits[j] >= nqubits) return set_error(QJ_EVALUE, "%s: bit %d out of range", what, tile_bits[j]);
    if (j && tile_bits[j] <= tile_bits[j - 1]) return set_error(QJ_EVALUE, "%s: tile bits must be ascending and unique", what);
  }
  RotTileArgs a;
  std::memset(&a, 0, sizeof(a));
  bool any_gen = false;
  if (gens) {
    for (int j = 0; j < kTileBits; ++j)
      for (int st = 0; st < nrot; ++st) {
        const int g = gens[j * nrot + st];
        if (g < 0 || g > 3) return set_error(QJ_EVALUE, "%s: bad generator %d", what, g);
        a.gen[j][st] = g;
        any_gen |= g != 0 && ((rotmask >> j) & 1);
      }
  }
  if (adj && any_gen && !partials) return set_error(QJ_EVALUE, "%s: gradients need a partials buffer", what);
  a.ins.npos = kTileBits;
  for (int j = 0; j < kTileBits; ++j) {
    a.ins.pos[j] = tile_bits[j];
    a.tpos[j] = tile_bits[j];
  }
  for (int r = 0; r < 16; ++r) {
    u64 o = 0;
    for (int u = 0; u < 4; ++u)
      if ((r >> u) & 1) o |= u64(1) << tile_bits[8 + u];
    a.offr[r] = o;
  }
  a.rotmask = rotmask;
  a.nrot = nrot;
  a.load = load;
  a.stride = stride;
  a.mstride = mstride;
  a.nsel = nrows;
  const u64 ntiles = (u64(1) << nqubits) >> kTileBits;
  dim3 gd((unsigned)ncols, (unsigned)nrows);
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = (size_t)(adj ? 2 : 1) * (8 << kTileBits);  // one complex64 tile per state
  const cplx<float>* m = static_cast<const cplx<float>*>(mats);
  const cplx<float>* pi = (const cplx<float>*)psi_in;
  cplx<float>* po = (cplx<float>*)psi_out;
  const cplx<float>* li = (const cplx<float>*)lam_in;
  cplx<float>* lo = (cplx<float>*)lam_out;
  // passes whose last step is a Z rotation on every rotated bit (RY then
  // RZ layers) run those steps for the whole tile at once
  bool zlast = nrot == 2;
  for (int j = 0; zlast && j < kTileBits; ++j)
    if (((rotmask >> j) & 1) && a.gen[j][nrot - 1] != 3) zlast = false;
#define QJ_TILE(ADJ, NR) k_rot_tile<ADJ, NR><<<gd, dim3(kTileBlock), lds, st>>>(pi, po, li, lo, sel, m, partials, ntiles, a)
  if (adj) {
    if (nrot == 1) QJ_TILE(true, 1);
    else if (nrot == 2 && zlast)
      k_rot_tile<true, 2, true><<<gd, dim3(kTileBlock), lds, st>>>(pi, po, li, lo, sel, m, partials, ntiles, a);
    else if (nrot == 2) QJ_TILE(true, 2);
    else QJ_TILE(true, 3);
  } else {
    if (nrot == 1) QJ_TILE(false, 1);
    else if (nrot == 2 && zlast)
      k_rot_tile<false, 2, true><<<gd, dim3(kTileBlock), lds, st>>>(pi, po, li, lo, sel, m, partials, ntiles, a);
    else if (nrot == 2) QJ_TILE(false, 2);
    else QJ_TILE(false, 3);
  }
#undef QJ_TILE
  return check_launch("rot_tile_pass launch");
}

int qj_tfim_batch(const void* psi, void* out, int dtype, int nqubits, uint64_t stride, const int32_t* sel,
                  int nrows, double g, int ladder, double* partials, int ncols, void* stream) {
  if (!psi || !partials) return set_error(QJ_EVALUE, "tfim_batch: null pointer");
  if (nqubits < 2 || nqubits > 40) return set_error(QJ_EVALUE, "tfim_batch: nqubits %d out of range", nqubits);
  int rc = check_rows(nrows, sel, stride, nqubits, "tfim_batch");
  if (rc) return rc;
  if (ncols < 1 || ncols > kRotMaxCols) return set_error(QJ_EVALUE, "tfim_batch: bad column count %d", ncols);
  if (psi == out) return set_error(QJ_EVALUE, "tfim_batch: out must not alias psi");
  hipStream_t st = (hipStream_t)stream;
  dim3 gd((unsigned)ncols, (unsigned)nrows);
  // vector groups of 32 bytes when every state starts on one (the state
  // pointers are at least 256-byte aligned)
  if (dtype == QJ_COMPLEX64) {
    if (stride % 4 == 0 && (reinterpret_cast<uintptr_t>(psi) & 31) == 0 &&
        (!out || (reinterpret_cast<uintptr_t>(out) & 31) == 0))
      k_tfim_batch<float, 4><<<gd, dim3(kRotBlock), 0, st>>>((const cplx<float>*)psi, (cplx<float>*)out, sel,
                                                             stride, nqubits, g, ladder != 0, partials);
    else
      k_tfim_batch<float, 1><<<gd, dim3(kRotBlock), 0, st>>>((const cplx<float>*)psi, (cplx<float>*)out, sel,
                                                             stride, nqubits, g, ladder != 0, partials);
  } else if (dtype == QJ_COMPLEX128) {
    if (stride % 2 == 0 && (reinterpret_cast<uintptr_t>(psi) & 31) == 0 &&
        (!out || (reinterpret_cast<uintptr_t>(out) & 31) == 0))
      k_tfim_batch<double, 2><<<gd, dim3(kRotBlock), 0, st>>>((const cplx<double>*)psi, (cplx<double>*)out, sel,
                                                              stride, nqubits, g, ladder != 0, partials);
    else
      k_tfim_batch<double, 1><<<gd, dim3(kRotBlock), 0, st>>>((const cplx<double>*)psi, (cplx<double>*)out, sel,
                                                              stride, nqubits, g, ladder != 0, partials);
  } else {
    return set_error(QJ_EVALUE, "tfim_batch: dtype must be complex64/complex128");
  }
  return check_launch("tfim_batch launch");
}

int qj_ansatz_lds(int dtype, const int32_t* desc, int ncirc, const void* mats, uint64_t mats_len, double g,
                  double* energies, int nqmax, int nblocks, void* stream) {
  if (!desc || !mats || !energies) return set_error(QJ_EVALUE, "ansatz_lds: null pointer");
  if (ncirc < 0) return set_error(QJ_EVALUE, "ansatz_lds: negative circuit count");
  if (ncirc == 0) return QJ_OK;
  if (nblocks < 1 || nblocks > (1 << 20)) return set_error(QJ_EVALUE, "ansatz_lds: bad workgroup count %d", nblocks);
  const size_t amp = dtype == QJ_COMPLEX64 ? 8 : dtype == QJ_COMPLEX128 ? 16 : 0;
  if (!amp) return set_error(QJ_EVALUE, "ansatz_lds: dtype must be complex64/complex128");
  if (nqmax < 2 || nqmax > 13 || (amp << nqmax) > 65536)
    return set_error(QJ_EVALUE, "ansatz_lds: %d qubits do not fit one workgroup's LDS", nqmax);
  const size_t lds = amp << nqmax;
  hipStream_t st = (hipStream_t)stream;
  dim3 gd((unsigned)nblocks);
  if (dtype == QJ_COMPLEX64)
    k_ansatz_lds<float><<<gd, dim3(kLdsBlock), lds, st>>>(desc, ncirc, (const cplx<float>*)mats, mats_len, g,
                                                         energies, nqmax);
  else
    k_ansatz_lds<double><<<gd, dim3(kLdsBlock), lds, st>>>(desc, ncirc, (const cplx<double>*)mats, mats_len, g,
                                                          energies, nqmax);
  return check_launch("ansatz_lds launch");
}

int qj_reduce_rows(const double* in, uint64_t nrows, uint64_t ncols, double* out, void* stream) {
  if (!in || !out) return set_error(QJ_EVALUE, "reduce_rows: null pointer");
  if (nrows == 0) return QJ_OK;
  if (nrows > 0x7fffffffull) return set_error(QJ_EVALUE, "reduce_rows: too many rows");
  if (ncols == 0) return set_error(QJ_EVALUE, "reduce_rows: no columns");
  k_reduce_rows<<<dim3((unsigned)nrows), dim3(kRotBlock), 0, (hipStream_t)stream>>>(in, ncols, out);
  return check_launch("reduce_rows launch");
}

}  // extern "C"
