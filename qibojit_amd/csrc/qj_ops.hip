// State preparation, collapse, marginal probabilities, Metropolis sampling
// and FP64 reductions for gfx950, with their C-ABI entry points.
//
// Reductions are deterministic: each workgroup reduces in a fixed order
// (wave64 butterfly, then the 4 waves) into a partials array, and a second
// kernel sums the partials in a fixed order.  Nothing is accumulated with
// floating-point atomics, so repeated calls give bit-identical results.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "qj_common.h"

namespace qj {

constexpr int kRedBlocks = 1024;  // workgroups of a grid-stride reduction
// Workgroups of the marginal-probability, norm and overlap kernels: 32 per
// CU, so the loads of a full chip are in flight (at most 32 partials each:
// 256 Ki doubles).
constexpr u64 kProbsMaxBlocks = 8192;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of v over the workgroup, valid on thread 0.
__device__ __forceinline__ double block_sum(double v, double* wsum) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / 64; ++i) t += wsum[i];
  __syncthreads();
  return t;
}

template <typename R>
__device__ __forceinline__ double abs2(cplx<R> x) {
  const double re = (double)x.re, im = (double)x.im;
  return re * re + im * im;
}

// out[v] = (T) sum_b partials[v*sv + b*sb], b < np; one workgroup per value.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_finalize(const double* __restrict__ partials, u64 np, u64 sv, u64 sb, T* __restrict__ out) {
  __shared__ double wsum[kBlock / 64];
  const u64 v = blockIdx.x;
  double acc = 0.0;
  for (u64 b = threadIdx.x; b < np; b += kBlock) acc += partials[v * sv + b * sb];
  acc = block_sum(acc, wsum);
  if (threadIdx.x == 0) out[v] = (T)acc;
}

// ---------------------------------------------------------- initial state --
template <typename R>
__global__ void __launch_bounds__(kBlock) k_initial_state(cplx<R>* __restrict__ st, u64 n) {
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    st[i] = mk<R>(i == 0 ? R(1) : R(0), R(0));
}

// Every amplitude set to `value` (the plus state |+...+> and its density
// matrix are uniform).
template <typename R>
__global__ void __launch_bounds__(kBlock) k_fill(cplx<R>* __restrict__ st, u64 n, cplx<R> value) {
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) st[i] = value;
}

// --------------------------------------------------------------- collapse --
// Amplitude i is kept iff (i & mask) == keep.  The state is walked in 16-byte
// chunks (one complex128 or two complex64 amplitudes) grouped into 128-byte
// lines; the measured bits split into line bits (hi_*) and bits inside a line
// (lo_*).  A line whose line bits miss `keep` holds no kept amplitude and is
// overwritten with zeros without being read.
//
// Pass 1 zeroes those lines and, when the norm is wanted, reads the kept
// amplitudes of the other lines for FP64 partial sums of |kept|^2 (one per
// workgroup).  With `zero_in_kept` (no normalisation pass follows) it also
// zeroes the dropped amplitudes inside kept lines.  Pass 2 (normalisation)
// rewrites each kept line once: kept amplitudes divided by the norm, the
// others zero.  Each workgroup walks tiles of kBlock x kCollapseUnroll chunks,
// every unrolled step one contiguous 4 KiB span.
constexpr int kCollapseUnroll = 4;
constexpr u64 kCollapseMaxBlocks = u64(1) << 16;  // partials per pass-1 launch

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_collapse_lines(cplx<R>* __restrict__ st, u64 nchunks, u64 hi_mask, u64 hi_keep, u64 lo_mask, u64 lo_keep,
                     int want_norm, int zero_in_kept, double* __restrict__ partials) {
  constexpr int V = 16 / (int)sizeof(cplx<R>);
  constexpr int U = kCollapseUnroll;
  __shared__ double wsum[kBlock / 64];
  const u64 tile = (u64)kBlock * U;
  const u64 ntiles = (nchunks + tile - 1) / tile;
  cvec<R, V> zero;
#pragma unroll
  for (int j = 0; j < V; ++j) zero.v[j] = mk<R>(R(0), R(0));
  double acc = 0.0;
  for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    cvec<R, V> x[U];
    bool in[U], kept_line[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const u64 c = t * tile + (u64)u * kBlock + threadIdx.x;
      in[u] = c < nchunks;
      kept_line[u] = in[u] && ((c * V) & hi_mask) == hi_keep;
      if (kept_line[u] && want_norm) x[u] = ldv<R, V>(st + c * V);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!in[u]) continue;
      const u64 i0 = (t * tile + (u64)u * kBlock + threadIdx.x) * V;
      if (!kept_line[u]) {
        stv<R, V>(st + i0, zero);
        continue;
      }
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const bool kept = ((i0 + j) & lo_mask) == lo_keep;
        if (kept) {
          if (want_norm) acc += abs2(x[u].v[j]);
        } else if (zero_in_kept) {
          st[i0 + j] = mk<R>(R(0), R(0));
        }
      }
    }
  }
  acc = block_sum(acc, wsum);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// Pass 2 over the kept lines only: kept chunk k is chunk (k % cpl) of kept
// line k / cpl, whose line number comes from inserting the measured line bits
// (`lines`, positions relative to the line size 2^lb, `ones` = outcome).
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_collapse_rescale(cplx<R>* __restrict__ st, u64 nkept_chunks, int lb, int cpl_log2, Ins lines, u64 lo_mask,
                       u64 lo_keep, const double* __restrict__ norm2) {
  constexpr int V = 16 / (int)sizeof(cplx<R>);
  constexpr int U = kCollapseUnroll;
  const double norm = sqrt(*norm2);
  const u64 tile = (u64)kBlock * U;
  const u64 ntiles = (nkept_chunks + tile - 1) / tile;
  const u64 cpl_mask = (u64(1) << cpl_log2) - 1;
  for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    cvec<R, V> x[U];
    u64 i0[U];
    bool in[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const u64 k = t * tile + (u64)u * kBlock + threadIdx.x;
      in[u] = k < nkept_chunks;
      i0[u] = (insert_bits(k >> cpl_log2, lines) << lb) + (k & cpl_mask) * V;
      if (in[u]) x[u] = ldv<R, V>(st + i0[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (!in[u]) continue;
#pragma unroll
      for (int j = 0; j < V; ++j) {
        const bool kept = ((i0[u] + j) & lo_mask) == lo_keep;
        const cplx<R> a = x[u].v[j];
        x[u].v[j] = kept ? mk<R>((R)((double)a.re / norm), (R)((double)a.im / norm)) : mk<R>(R(0), R(0));
      }
      stv<R, V>(st + i0[u], x[u]);
    }
  }
}

// ---------------------------------------------------------- probabilities --
// Few outcomes (2^k <= 32): each thread keeps the 2^k partial sums in
// registers while it walks the free index; workgroup partials are then
// summed by k_finalize.
// Weight of basis index i: |psi_i|^2 for a state vector (dstride = 0), or
// Re rho_ii of a density matrix stored row-major (dstride = 2^n + 1).
template <typename R>
__device__ __forceinline__ double prob_at(const cplx<R>* st, u64 i, u64 dstride) {
  return dstride ? (double)st[i * dstride].re : abs2(st[i]);
}

template <typename R, int NB>
__global__ void __launch_bounds__(kBlock)
    k_probs_bins(const cplx<R>* __restrict__ st, u64 nfree, Ins fins, Offs32 hoff, u64 dstride,
                 double* __restrict__ partials) {
  __shared__ double wsum[kBlock / 64];
  double acc[NB];
#pragma unroll
  for (int h = 0; h < NB; ++h) acc[h] = 0.0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 g = (u64)blockIdx.x * kBlock + threadIdx.x; g < nfree; g += stride) {
    const u64 base = insert_bits(g, fins);
#pragma unroll
    for (int h = 0; h < NB; ++h) acc[h] += prob_at(st, base + hoff.o[h], dstride);
  }
#pragma unroll
  for (int h = 0; h < NB; ++h) {
    const double t = block_sum(acc[h], wsum);
    if (threadIdx.x == 0) partials[(u64)h * gridDim.x + blockIdx.x] = t;
  }
}

// Bit 0 measured on a complex64 state: the two values of bit 0 are adjacent
// amplitudes, so a thread reads them as one 16-byte pair (consecutive
// threads then read consecutive pairs) and adds them to the pair's two
// outcomes.  Pair p starts at offset poff.o[p] (bit 0 clear); its outcomes
// are bins.even[p] and bins.odd[p].
struct PairBins {
  int even[16], odd[16];
};

template <int NP>
__global__ void __launch_bounds__(kBlock)
    k_probs_pairs(const cplx<float>* __restrict__ st, u64 nfree, Ins fins, Offs32 poff, PairBins bins,
                  double* __restrict__ partials) {
  __shared__ double wsum[kBlock / 64];
  double a0[NP], a1[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) a0[p] = a1[p] = 0.0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 g = (u64)blockIdx.x * kBlock + threadIdx.x; g < nfree; g += stride) {
    const u64 base = insert_bits(g, fins);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const cvec<float, 2> v = ldv<float, 2>(st + base + poff.o[p]);
      a0[p] += abs2(v.v[0]);
      a1[p] += abs2(v.v[1]);
    }
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const double t0 = block_sum(a0[p], wsum);
    const double t1 = block_sum(a1[p], wsum);
    if (threadIdx.x == 0) {
      partials[(u64)bins.even[p] * gridDim.x + blockIdx.x] = t0;
      partials[(u64)bins.odd[p] * gridDim.x + blockIdx.x] = t1;
    }
  }
}

// Many outcomes: one work item per (chunk of the free index, outcome h).
template <typename R, typename T>
__global__ void __launch_bounds__(kBlock)
    k_probs_wide(const cplx<R>* __restrict__ st, u64 nbins, u64 nfree, u64 nchunks, Ins fins,
                 Ins obits, u64 dstride, double* __restrict__ partials, T* __restrict__ out) {
  const u64 total = nbins * nchunks;
  const u64 per = (nfree + nchunks - 1) / nchunks;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 w = (u64)blockIdx.x * kBlock + threadIdx.x; w < total; w += stride) {
    const u64 h = w % nbins, c = w / nbins;
    // output bit (k-1-j) of h is the value of bit position obits.pos[j]
    u64 off = 0;
    for (int j = 0; j < obits.npos; ++j)
      if ((h >> (obits.npos - 1 - j)) & 1) off |= u64(1) << obits.pos[j];
    const u64 g0 = c * per, g1 = (g0 + per < nfree) ? g0 + per : nfree;
    double acc = 0.0;
    for (u64 g = g0; g < g1; ++g) acc += prob_at(st, insert_bits(g, fins) | off, dstride);
    if (nchunks == 1) out[h] = (T)acc;
    else partials[c * nbins + h] = acc;
  }
}

// ------------------------------------------------------ Metropolis sampler --
// MT19937 with NumPy's legacy RandomState semantics: seed(int) is the
// Knuth/Matsumoto init_genrand recurrence; outputs are tempered key words.
constexpr int kMtN = 624, kMtM = 397;

__device__ __forceinline__ uint32_t mt_temper(uint32_t y) {
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}
__device__ __forceinline__ uint32_t mt_mix(uint32_t a, uint32_t b, uint32_t c) {
  const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
  return c ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
}

// Regenerates key[0..623] with the whole workgroup.  Each index range reads
// old values into registers before the barrier that precedes its writes;
// the ranges are ordered by their dependencies: [0,227) uses old words,
// [227,454) uses [0,227), [454,623) uses [227,396), 623 uses key[0], key[396].
__device__ void mt_twist_block(uint32_t* key) {
  const int t = threadIdx.x;
  uint32_t v = 0;
  // [0, 227)
  for (int i0 = 0; i0 < kMtN - kMtM; i0 += kBlock) {
    const int i = i0 + t;
    if (i < kMtN - kMtM) v = mt_mix(key[i], key[i + 1], key[i + kMtM]);
    __syncthreads();
    if (i < kMtN - kMtM) key[i] = v;
    __syncthreads();
  }
  // [227, 454) then [454, 623)
  for (int lo = kMtN - kMtM; lo < kMtN - 1; lo += kMtN - kMtM) {
    const int hi = min(lo + (kMtN - kMtM), kMtN - 1);
    for (int i0 = lo; i0 < hi; i0 += kBlock) {
      const int i = i0 + t;
      if (i < hi) v = mt_mix(key[i], key[i + 1], key[i - (kMtN - kMtM)]);
      __syncthreads();
      if (i < hi) key[i] = v;
      __syncthreads();
    }
  }
  if (t == 0) key[kMtN - 1] = mt_mix(key[kMtN - 1], key[0], key[kMtM - 1]);
  __syncthreads();
}

__device__ void mt_seed_block(uint32_t* key, uint32_t seed) {
  if (threadIdx.x == 0) {
    uint32_t s = seed;
    for (int i = 0; i < kMtN; ++i) {
      key[i] = s;
      s = 1812433253u * (s ^ (s >> 30)) + (uint32_t)(i + 1);
    }
  }
  __syncthreads();
}

// random_sample(): 53-bit double from two consecutive words.
__device__ __forceinline__ double mt_double(uint32_t w1, uint32_t w2) {
  const double a = (double)(w1 >> 5), b = (double)(w2 >> 6);
  return (a * 67108864.0 + b) / 9007199254740992.0;
}

// One Metropolis step: propose (x + off) mod N, accept when the probability
// ratio, computed in the probs type, exceeds u (compared in double).
template <typename T>
__device__ __forceinline__ int mh_step(int x, uint32_t off, double u, const T* p, int nmask) {
  const int y = (int)((x + off) & (uint32_t)nmask);
  const T ratio = p[y] / p[x];
  return ((double)ratio > u) ? y : x;
}

constexpr int kSegSteps = 65536;  // steps per segment of a chain
constexpr int kMaxScanStates = 64;

// Power-of-two N <= 64: one workgroup per chain.  Per segment: (1) the
// workgroup generates the chain's words (3 per step: randint then the two
// halves of random()); (2) thread t composes the transition functions of its
// run of steps into a table T_t; (3) thread 0 chains the tables to find each
// run's start state; (4) every thread replays its run from that state and
// counts visits.  The visited sequence is identical to the sequential chain.
template <typename T, int N>
__global__ void __launch_bounds__(kBlock)
    k_metropolis_scan(int64_t* __restrict__ freqs, const T* __restrict__ probs,
                      const uint32_t* __restrict__ seeds, const int64_t* __restrict__ chain_shots,
                      uint32_t* __restrict__ words_all, int start_state) {
  __shared__ uint32_t key[kMtN];
  __shared__ T p[N];
  __shared__ unsigned char tables[kBlock][N];
  __shared__ int run_start[kBlock];
  __shared__ unsigned int counts[N];
  __shared__ int carried;
  const int chain = blockIdx.x;
  uint32_t* words = words_all + (size_t)chain * (3 * (size_t)kSegSteps);
  for (int i = threadIdx.x; i < N; i += kBlock) {
    p[i] = probs[i];
    counts[i] = 0;
  }
  mt_seed_block(key, seeds[chain]);
  int pos = kMtN;  // words of the current key batch already used
  int state = start_state;
  const int64_t total = chain_shots[chain];
  for (int64_t done = 0; done < total; done += kSegSteps) {
    const int steps = (int)((total - done < kSegSteps) ? (total - done) : kSegSteps);
    const int need = 3 * steps;
    // (1) words
    for (int filled = 0; filled < need;) {
      if (pos == kMtN) {
        mt_twist_block(key);
        pos = 0;
      }
      const int take = min(kMtN - pos, need - filled);
      for (int i = threadIdx.x; i < take; i += kBlock) words[filled + i] = mt_temper(key[pos + i]);
      filled += take;
      pos += take;
    }
    __syncthreads();
    __threadfence_block();
    // (2) per-thread composition over a run of steps
    const int per = (steps + kBlock - 1) / kBlock;
    const int s0 = min(steps, threadIdx.x * per), s1 = min(steps, s0 + per);
    int tbl[N];
#pragma unroll
    for (int x = 0; x < N; ++x) tbl[x] = x;
    for (int s = s0; s < s1; ++s) {
      const uint32_t off = words[3 * s] & (uint32_t)(N - 1);
      const double u = mt_double(words[3 * s + 1], words[3 * s + 2]);
#pragma unroll
      for (int x = 0; x < N; ++x) tbl[x] = mh_step<T>(tbl[x], off, u, p, N - 1);
    }
#pragma unroll
    for (int x = 0; x < N; ++x) tables[threadIdx.x][x] = (unsigned char)tbl[x];
    __syncthreads();
    // (3) run start states
    if (threadIdx.x == 0) {
      int x = state;
      for (int t = 0; t < kBlock; ++t) {
        run_start[t] = x;
        x = tables[t][x];
      }
      carried = x;  // start state of the next segment
    }
    __syncthreads();
    // (4) replay and count
    int x = run_start[threadIdx.x];
    for (int s = s0; s < s1; ++s) {
      const uint32_t off = words[3 * s] & (uint32_t)(N - 1);
      const double u = mt_double(words[3 * s + 1], words[3 * s + 2]);
      x = mh_step<T>(x, off, u, p, N - 1);
      atomicAdd(&counts[x], 1u);
    }
    state = carried;
    __syncthreads();  // words/tables are rewritten by the next segment
  }
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += kBlock)
    if (counts[i]) atomicAdd((unsigned long long*)&freqs[i], (unsigned long long)counts[i]);
}

// General N (any size, rejection possible): one thread per chain walks the
// chain exactly as the reference loop does, drawing words on demand.
template <typename T>
__global__ void __launch_bounds__(64)
    k_metropolis_seq(int64_t* __restrict__ freqs, const T* __restrict__ probs, int64_t nstates,
                     const uint32_t* __restrict__ seeds, const int64_t* __restrict__ chain_shots,
                     int64_t start_state) {
  __shared__ uint32_t key[kMtN];
  if (threadIdx.x != 0) return;
  const int chain = blockIdx.x;
  {
    uint32_t s = seeds[chain];
    for (int i = 0; i < kMtN; ++i) {
      key[i] = s;
      s = 1812433253u * (s ^ (s >> 30)) + (uint32_t)(i + 1);
    }
  }
  int pos = kMtN;
  auto next32 = [&]() -> uint32_t {
    if (pos == kMtN) {
      for (int i = 0; i < kMtN - kMtM; ++i) key[i] = mt_mix(key[i], key[i + 1], key[i + kMtM]);
      for (int i = kMtN - kMtM; i < kMtN - 1; ++i)
        key[i] = mt_mix(key[i], key[i + 1], key[i - (kMtN - kMtM)]);
      key[kMtN - 1] = mt_mix(key[kMtN - 1], key[0], key[kMtM - 1]);
      pos = 0;
    }
    return mt_temper(key[pos++]);
  };
  const uint64_t rng = (uint64_t)nstates - 1;
  uint32_t mask = (uint32_t)rng;
  mask |= mask >> 1; mask |= mask >> 2; mask |= mask >> 4; mask |= mask >> 8; mask |= mask >> 16;
  int64_t x = start_state;
  const int64_t total = chain_shots[chain];
  for (int64_t i = 0; i < total; ++i) {
    uint64_t off = 0;
    if (rng != 0) {
      uint32_t v;
      if (rng == 0xFFFFFFFFull) v = next32();
      else do { v = next32() & mask; } while (v > rng);
      off = v;
    }
    const uint32_t w1 = next32(), w2 = next32();
    const double u = mt_double(w1, w2);
    const int64_t y = (int64_t)(((uint64_t)x + off) % (uint64_t)nstates);
    const T ratio = probs[y] / probs[x];
    if ((double)ratio > u) x = y;
    atomicAdd((unsigned long long*)&freqs[x], 1ull);
  }
}

// ------------------------------------------------------- FP64 reductions --
// Norm and overlap read V amplitudes per lane and step (V = 2 for 16-byte
// aligned complex64 states: 16-byte loads); a tail of n mod V amplitudes is
// taken by thread 0 of workgroup 0.
template <typename R, int V>
__global__ void __launch_bounds__(kBlock)
    k_norm2(const cplx<R>* __restrict__ st, u64 n, double* __restrict__ partials) {
  __shared__ double wsum[kBlock / 64];
  double acc = 0.0;
  const u64 stride = (u64)gridDim.x * kBlock, nv = n / V;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < nv; i += stride) {
    const cvec<R, V> x = ldv<R, V>(st + i * V);
#pragma unroll
    for (int k = 0; k < V; ++k) acc += abs2(x.v[k]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (u64 i = nv * V; i < n; ++i) acc += abs2(st[i]);
  acc = block_sum(acc, wsum);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

template <typename R, int V>
__global__ void __launch_bounds__(kBlock)
    k_vdot(const cplx<R>* __restrict__ a, const cplx<R>* __restrict__ b, u64 n,
           double* __restrict__ partials) {
  __shared__ double wsum[kBlock / 64];
  double re = 0.0, im = 0.0;
  auto add = [&](cplx<R> x, cplx<R> y) {
    const double xr = x.re, xi = x.im, yr = y.re, yi = y.im;
    re += xr * yr + xi * yi;
    im += xr * yi - xi * yr;
  };
  const u64 stride = (u64)gridDim.x * kBlock, nv = n / V;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < nv; i += stride) {
    const cvec<R, V> x = ldv<R, V>(a + i * V), y = ldv<R, V>(b + i * V);
#pragma unroll
    for (int k = 0; k < V; ++k) add(x.v[k], y.v[k]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (u64 i = nv * V; i < n; ++i) add(a[i], b[i]);
  re = block_sum(re, wsum);
  im = block_sum(im, wsum);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = re;
    partials[gridDim.x + blockIdx.x] = im;
  }
}

// Two states of possibly different precision: partial sums of Re/Im <a|b>,
// |a|^2, |b|^2 and the largest |a_i - b_i| per workgroup, in FP64.
template <typename RA, typename RB>
__global__ void __launch_bounds__(kBlock)
    k_compare(const cplx<RA>* __restrict__ a, const cplx<RB>* __restrict__ b, u64 n, double* __restrict__ partials) {
  __shared__ double wsum[kBlock / 64];
  double re = 0.0, im = 0.0, na = 0.0, nb = 0.0, dmax = 0.0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const double xr = a[i].re, xi = a[i].im, yr = b[i].re, yi = b[i].im;
    re += xr * yr + xi * yi;
    im += xr * yi - xi * yr;
    na += xr * xr + xi * xi;
    nb += yr * yr + yi * yi;
    const double dr = xr - yr, di = xi - yi;
    dmax = fmax(dmax, dr * dr + di * di);
  }
  re = block_sum(re, wsum);
  im = block_sum(im, wsum);
  na = block_sum(na, wsum);
  nb = block_sum(nb, wsum);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = dmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) m = fmax(m, wsum[w]);
    const u64 g = gridDim.x, k = blockIdx.x;
    partials[k] = re;
    partials[g + k] = im;
    partials[2 * g + k] = na;
    partials[3 * g + k] = nb;
    partials[4 * g + k] = sqrt(m);
  }
}

// <psi|H|psi>, H = -sum Z_i Z_{i+1} - g sum X_i on an open chain of n qubits.
// Diagonal: sum_i z_i z_{i+1} = (n-1) - 2 popcount((k ^ (k >> 1)) & (2^(n-1)-1)).
// X_b term: sum_k Re(conj(psi_k) psi_{k^2^b}) = 2 sum_{bit b of k = 0} Re(...).
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_tfim(const cplx<R>* __restrict__ st, int nq, double g, double* __restrict__ partials) {
  __shared__ double wsum[kBlock / 64];
  const u64 n = u64(1) << nq;
  const u64 zzmask = (u64(1) << (nq - 1)) - 1;
  double acc = 0.0;
  const u64 stride = (u64)gridDim.x * kBlock;
  for (u64 k = (u64)blockIdx.x * kBlock + threadIdx.x; k < n; k += stride) {
    const cplx<R> x = st[k];
    const int pc = __popcll((k ^ (k >> 1)) & zzmask);
    const double zz = (double)(nq - 1) - 2.0 * (double)pc;
    acc -= zz * abs2(x);
    double xs = 0.0;
    for (int b = 0; b < nq; ++b) {
      const u64 bit = u64(1) << b;
      if (!(k & bit)) {
        const cplx<R> y = st[k | bit];
        xs += (double)x.re * (double)y.re + (double)x.im * (double)y.im;
      }
    }
    acc -= 2.0 * g * xs;
  }
  acc = block_sum(acc, wsum);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_axpby(const cplx<R>* __restrict__ x, cplx<R>* __restrict__ y, u64 n, cplx<R> a, cplx<R> b) {
  const u64 stride = (u64)gridDim.x * kBlock;
  // beta == 0 never reads y, so y may start uninitialised
  const bool read_y = (b.re != R(0)) || (b.im != R(0));
  for (u64 i = (u64)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    y[i] = read_y ? cadd(cmul(a, x[i]), cmul(b, y[i])) : cmul(a, x[i]);
}

// ------------------------------------------------------------ host side --

static int reduce_to_host(int nvals, u64 np, u64 sv, u64 sb, double* out, hipStream_t s,
                          DeviceScratch* sc) {
  k_finalize<double><<<dim3(nvals), dim3(kBlock), 0, s>>>(sc->partials, np, sv, sb, sc->scalars);
  int rc = check_launch("finalize launch");
  if (rc) return rc;
  rc = hip_status(hipMemcpyAsync(out, sc->scalars, nvals * sizeof(double), hipMemcpyDeviceToHost, s),
                  "reduction readback");
  if (rc) return rc;
  return hip_status(hipStreamSynchronize(s), "reduction sync");
}

template <typename R>
static int collapse_impl(void* state, int nqubits, const int32_t* qk, int nm, int64_t result,
                         int normalize, double* norm_out, hipStream_t s) {
  if (nqubits < 1 || nqubits > 62) return set_error(QJ_EVALUE, "collapse_state: bad nqubits %d", nqubits);
  if (nm < 1 || nm > nqubits) return set_error(QJ_EVALUE, "collapse_state: bad number of measured qubits %d", nm);
  if (result < 0 || (nm < 63 && result >= (int64_t(1) << nm)))
    return set_error(QJ_EVALUE, "collapse_state: result %lld out of range", (long long)result);
  for (int j = 0; j < nm; ++j) {
    if (qk[j] < 0 || qk[j] >= nqubits)
      return set_error(QJ_EVALUE, "collapse_state: bit %d out of range", qk[j]);
    for (int i = 0; i < j; ++i)
      if (qk[i] == qk[j]) return set_error(QJ_EVALUE, "collapse_state: repeated qubit");
  }
  // Bit j of `result` is the value of bit position qk[j].  For ascending qk
  // (what the callers pass: qubits sorted, then reversed into bit
  // positions) this is exactly the index set of collapse_index
  // (ops.py:33-40), which inserts bit j of h at qk[j] in list order; for a
  // non-ascending list that insertion shifts earlier bits past the vector,
  // so positions are used directly instead.
  u64 mask = 0, keep = 0;
  for (int j = 0; j < nm; ++j) {
    mask |= u64(1) << qk[j];
    keep |= (u64)((result >> j) & 1) << qk[j];
  }
  DeviceScratch* sc = nullptr;
  int rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  constexpr int V = 16 / (int)sizeof(cplx<R>);
  constexpr int kLineLog2 = sizeof(R) == 8 ? 3 : 4;  // amplitudes per 128-byte line
  const int lb = std::min(kLineLog2, nqubits);
  const u64 n = u64(1) << nqubits;
  const u64 lo_bits = (u64(1) << lb) - 1;
  const u64 hi_mask = mask & ~lo_bits, hi_keep = keep & ~lo_bits;
  const u64 lo_mask = mask & lo_bits, lo_keep = keep & lo_bits;
  cplx<R>* st = static_cast<cplx<R>*>(state);
  const bool want_norm = normalize || norm_out;
  const u64 nchunks = n / V;
  const u64 tile = (u64)kBlock * kCollapseUnroll;
  const u64 grid = std::min<u64>(kCollapseMaxBlocks, (nchunks + tile - 1) / tile);
  if (grid > sc->partials_cap) return set_error(QJ_EHIP, "collapse_state: partials buffer too small");
  k_collapse_lines<R><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(st, nchunks, hi_mask, hi_keep, lo_mask, lo_keep,
                                                                   want_norm ? 1 : 0, normalize ? 0 : 1, sc->partials);
  rc = check_launch("collapse launch");
  if (rc) return rc;
  if (!want_norm) return QJ_OK;
  k_finalize<double><<<dim3(1), dim3(kBlock), 0, s>>>(sc->partials, grid, 0, 1, sc->scalars);
  rc = check_launch("collapse finalize launch");
  if (rc) return rc;
  if (normalize) {
    Ins lines;
    lines.npos = 0;
    lines.ones = 0;
    for (int b = lb; b < nqubits; ++b)
      if ((mask >> b) & 1) {
        lines.pos[lines.npos++] = b - lb;
        lines.ones |= ((keep >> b) & 1) << (b - lb);
      }
    for (int j = lines.npos; j < kMaxPos; ++j) lines.pos[j] = 0;
    const int cpl_log2 = lb - (V == 2 ? 1 : 0);  // chunks per line
    const u64 nkept_chunks = (n >> lines.npos) / V;
    const u64 grid2 = grid_for((nkept_chunks + tile - 1) / tile, 1);
    k_collapse_rescale<R><<<dim3((unsigned)grid2), dim3(kBlock), 0, s>>>(st, nkept_chunks, lb, cpl_log2, lines,
                                                                        lo_mask, lo_keep, sc->scalars);
    rc = check_launch("collapse rescale launch");
    if (rc) return rc;
  }
  if (norm_out) {
    double n2 = 0;
    rc = hip_status(hipMemcpyAsync(&n2, sc->scalars, sizeof(double), hipMemcpyDeviceToHost, s), "norm readback");
    if (rc) return rc;
    rc = hip_status(hipStreamSynchronize(s), "collapse sync");
    if (rc) return rc;
    *norm_out = std::sqrt(n2);
  }
  return QJ_OK;
}

template <typename R, typename T>
static int probs_impl(const void* state, int nqubits, const int32_t* qubits, int nm, T* out,
                      hipStream_t s, bool density_matrix = false) {
  if (density_matrix && (nqubits < 1 || nqubits > 31))
    return set_error(QJ_EVALUE, "calculate_probabilities: bad density-matrix nqubits %d", nqubits);
  const u64 dstride = density_matrix ? (u64(1) << nqubits) + 1 : 0;
  if (nqubits < 1 || nqubits > 62) return set_error(QJ_EVALUE, "calculate_probabilities: bad nqubits");
  if (nm < 1 || nm > nqubits || nm > kMaxPos)
    return set_error(QJ_EVALUE, "calculate_probabilities: bad number of qubits %d", nm);
  Ins obits;  // bit positions in output order (first qubit = MSB)
  obits.npos = nm;
  std::vector<int> sorted;
  for (int j = 0; j < nm; ++j) {
    if (qubits[j] < 0 || qubits[j] >= nqubits)
      return set_error(QJ_EVALUE, "calculate_probabilities: qubit %d out of range", qubits[j]);
    obits.pos[j] = nqubits - 1 - qubits[j];
    sorted.push_back(obits.pos[j]);
  }
  for (int j = nm; j < kMaxPos; ++j) obits.pos[j] = 0;
  obits.ones = 0;
  std::sort(sorted.begin(), sorted.end());
  for (size_t j = 1; j < sorted.size(); ++j)
    if (sorted[j] == sorted[j - 1]) return set_error(QJ_EVALUE, "calculate_probabilities: repeated qubit");
  Ins fins;
  fins.npos = nm;
  for (int j = 0; j < nm; ++j) fins.pos[j] = sorted[j];
  for (int j = nm; j < kMaxPos; ++j) fins.pos[j] = 0;
  fins.ones = 0;
  DeviceScratch* sc = nullptr;
  int rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  const cplx<R>* st = static_cast<const cplx<R>*>(state);
  const u64 nfree = u64(1) << (nqubits - nm);
  const u64 nbins = u64(1) << nm;
  if (nm <= 5) {
    Offs32 hoff;
    for (u64 h = 0; h < 32; ++h) {
      u64 o = 0;
      if (h < nbins)
        for (int j = 0; j < nm; ++j)
          if ((h >> (nm - 1 - j)) & 1) o |= u64(1) << obits.pos[j];
      hoff.o[h] = o;
    }
    const u64 grid = std::min<u64>(kProbsMaxBlocks, grid_for(nfree, kBlock));
    dim3 gd((unsigned)grid), bd(kBlock);
    if (sizeof(R) == 4 && dstride == 0 && fins.pos[0] == 0) {
      // bit 0 measured, complex64: 16-byte pairs
      Offs32 poff;
      PairBins bins;
      std::memset(&poff, 0, sizeof(poff));
      std::memset(&bins, 0, sizeof(bins));
      int np = 0;
      for (u64 h = 0; h < nbins; ++h)
        if (!(hoff.o[h] & 1)) {
          poff.o[np] = hoff.o[h];
          bins.even[np] = (int)h;
          for (u64 h1 = 0; h1 < nbins; ++h1)
            if (hoff.o[h1] == (hoff.o[h] | 1)) bins.odd[np] = (int)h1;
          ++np;
        }
      const cplx<float>* sf = reinterpret_cast<const cplx<float>*>(st);
      switch (np) {
        case 1: k_probs_pairs<1><<<gd, bd, 0, s>>>(sf, nfree, fins, poff, bins, sc->partials); break;
        case 2: k_probs_pairs<2><<<gd, bd, 0, s>>>(sf, nfree, fins, poff, bins, sc->partials); break;
        case 4: k_probs_pairs<4><<<gd, bd, 0, s>>>(sf, nfree, fins, poff, bins, sc->partials); break;
        case 8: k_probs_pairs<8><<<gd, bd, 0, s>>>(sf, nfree, fins, poff, bins, sc->partials); break;
        default: k_probs_pairs<16><<<gd, bd, 0, s>>>(sf, nfree, fins, poff, bins, sc->partials); break;
      }
      rc = check_launch("probabilities launch");
      if (rc) return rc;
      k_finalize<T><<<dim3((unsigned)nbins), dim3(kBlock), 0, s>>>(sc->partials, grid, grid, 1, out);
      return check_launch("probabilities finalize launch");
    }
    switch (nm) {
      case 1: k_probs_bins<R, 2><<<gd, bd, 0, s>>>(st, nfree, fins, hoff, dstride, sc->partials); break;
      case 2: k_probs_bins<R, 4><<<gd, bd, 0, s>>>(st, nfree, fins, hoff, dstride, sc->partials); break;
      case 3: k_probs_bins<R, 8><<<gd, bd, 0, s>>>(st, nfree, fins, hoff, dstride, sc->partials); break;
      case 4: k_probs_bins<R, 16><<<gd, bd, 0, s>>>(st, nfree, fins, hoff, dstride, sc->partials); break;
      default: k_probs_bins<R, 32><<<gd, bd, 0, s>>>(st, nfree, fins, hoff, dstride, sc->partials); break;
    }
    rc = check_launch("probabilities launch");
    if (rc) return rc;
    k_finalize<T><<<dim3((unsigned)nbins), dim3(kBlock), 0, s>>>(sc->partials, grid, grid, 1, out);
    return check_launch("probabilities finalize launch");
  }
  const u64 target_items = u64(1) << 20;
  u64 nchunks = 1;
  if (nbins < target_items) nchunks = std::min<u64>(nfree, target_items / nbins);
  if (nchunks < 1) nchunks = 1;
  const u64 items = nbins * nchunks;
  k_probs_wide<R, T><<<dim3((unsigned)grid_for(items, kBlock)), dim3(kBlock), 0, s>>>(
      st, nbins, nfree, nchunks, fins, obits, dstride, sc->partials, out);
  rc = check_launch("probabilities launch");
  if (rc || nchunks == 1) return rc;
  k_finalize<T><<<dim3((unsigned)nbins), dim3(kBlock), 0, s>>>(sc->partials, nchunks, 1, nbins, out);
  return check_launch("probabilities finalize launch");
}

template <typename T>
static int metropolis_impl(int64_t* freqs, const T* probs, int64_t nstates, int64_t nshots,
                           const uint32_t* seeds, int nchains, hipStream_t s) {
  if (nstates < 1) return set_error(QJ_EVALUE, "measure_frequencies: empty probabilities");
  if (nchains < 1) return set_error(QJ_EVALUE, "measure_frequencies: nthreads must be >= 1");
  if (nshots < 0) return set_error(QJ_EVALUE, "measure_frequencies: negative nshots");
  if (nstates > (int64_t(1) << 32))
    return set_error(QJ_EVALUE, "measure_frequencies: more than 2^32 states");
  // argmax (first maximum) of the probabilities, as probs.argmax() does.
  std::vector<T> hp(nstates);
  int rc = hip_status(hipMemcpyAsync(hp.data(), probs, nstates * sizeof(T), hipMemcpyDeviceToHost, s),
                      "probabilities readback");
  if (rc) return rc;
  rc = hip_status(hipStreamSynchronize(s), "metropolis sync");
  if (rc) return rc;
  int64_t start = 0;
  for (int64_t i = 1; i < nstates; ++i)
    if (hp[i] > hp[start]) start = i;
  std::vector<int64_t> shots(nchains, nshots / nchains);
  shots[nchains - 1] += nshots % nchains;
  // seeds + per-chain shot counts travel in one staging upload
  std::vector<char> buf(nchains * (sizeof(uint32_t) + sizeof(int64_t)) + 8);
  std::memcpy(buf.data(), shots.data(), nchains * sizeof(int64_t));
  std::memcpy(buf.data() + nchains * sizeof(int64_t), seeds, nchains * sizeof(uint32_t));
  void* dbuf = nullptr;
  int ticket = -1;
  rc = ring_upload(buf.data(), buf.size(), s, &dbuf, &ticket);
  if (rc) return rc;
  const int64_t* dshots = static_cast<const int64_t*>(dbuf);
  const uint32_t* dseeds =
      reinterpret_cast<const uint32_t*>(static_cast<const char*>(dbuf) + nchains * sizeof(int64_t));
  const bool pow2 = (nstates & (nstates - 1)) == 0;
  if (pow2 && nstates >= 2 && nstates <= kMaxScanStates) {
    DeviceScratch* sc = nullptr;
    rc = scratch_for_current_device(&sc);
    if (rc) return rc;
    // words: 3 per step per chain, one segment at a time
    const size_t need = (size_t)nchains * 3 * kSegSteps * sizeof(uint32_t);
    rc = scratch_reserve_big(sc, need, s);
    if (rc) return rc;
    uint32_t* words = static_cast<uint32_t*>(sc->big);
    dim3 gd(nchains), bd(kBlock);
    switch (nstates) {
      case 2: k_metropolis_scan<T, 2><<<gd, bd, 0, s>>>(freqs, probs, dseeds, dshots, words, (int)start); break;
      case 4: k_metropolis_scan<T, 4><<<gd, bd, 0, s>>>(freqs, probs, dseeds, dshots, words, (int)start); break;
      case 8: k_metropolis_scan<T, 8><<<gd, bd, 0, s>>>(freqs, probs, dseeds, dshots, words, (int)start); break;
      case 16: k_metropolis_scan<T, 16><<<gd, bd, 0, s>>>(freqs, probs, dseeds, dshots, words, (int)start); break;
      case 32: k_metropolis_scan<T, 32><<<gd, bd, 0, s>>>(freqs, probs, dseeds, dshots, words, (int)start); break;
      default: k_metropolis_scan<T, 64><<<gd, bd, 0, s>>>(freqs, probs, dseeds, dshots, words, (int)start); break;
    }
    rc = check_launch("metropolis launch");
  } else {
    k_metropolis_seq<T><<<dim3(nchains), dim3(64), 0, s>>>(freqs, probs, nstates, dseeds, dshots, start);
    rc = check_launch("metropolis launch");
  }
  ring_release(ticket, s);
  return rc;
}

// ------------------------------------------------- shot sampling (CDF) --
// NumPy's RandomState.choice(n, nshots, p=probs): cdf = cumsum(p) in FP64,
// cdf /= cdf[-1], sample = searchsorted(cdf, u, side='right') for uniforms u
// drawn by the caller.  Up to kSeqCdfStates the cumsum runs in one thread in
// index order (the sum NumPy computes); above, a blocked scan (per-thread runs
// of kCdfRun, block offsets from a sequential pass over the block sums).
constexpr int64_t kSeqCdfStates = int64_t(1) << 16;
constexpr int kCdfRun = 64;

template <typename T>
__global__ void __launch_bounds__(64) k_cdf_seq(const T* __restrict__ p, int64_t n, double* __restrict__ cdf) {
  if (threadIdx.x != 0) return;
  double c = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    c += (double)p[i];
    cdf[i] = c;
  }
}

// Pass 1: sum of each block's chunk (kBlock threads x kCdfRun elements).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cdf_block_sums(const T* __restrict__ p, int64_t n,
                                                           double* __restrict__ bsum) {
  __shared__ double part[kBlock];
  const int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kCdfRun;
  double c = 0.0;
  for (int k = 0; k < kCdfRun; ++k)
    if (base + k < n) c += (double)p[base + k];
  part[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kBlock; ++i) t += part[i];
    bsum[blockIdx.x] = t;
  }
}

// Pass 2 (one thread): exclusive prefix of the block sums, in place.
__global__ void __launch_bounds__(64) k_cdf_block_offsets(double* __restrict__ bsum, int64_t nblocks) {
  if (threadIdx.x != 0) return;
  double c = 0.0;
  for (int64_t b = 0; b < nblocks; ++b) {
    const double v = bsum[b];
    bsum[b] = c;
    c += v;
  }
}

// Pass 3: each thread writes the running sums of its run, offset by the
// block's prefix and the sums of the block's earlier threads.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_cdf_block_scan(const T* __restrict__ p, int64_t n,
                                                           const double* __restrict__ boff,
                                                           double* __restrict__ cdf) {
  __shared__ double part[kBlock];
  const int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kCdfRun;
  double c = 0.0;
  for (int k = 0; k < kCdfRun; ++k)
    if (base + k < n) c += (double)p[base + k];
  part[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = boff[blockIdx.x];
    for (int i = 0; i < kBlock; ++i) {
      const double v = part[i];
      part[i] = t;
      t += v;
    }
  }
  __syncthreads();
  c = part[threadIdx.x];
  for (int k = 0; k < kCdfRun; ++k)
    if (base + k < n) {
      c += (double)p[base + k];
      cdf[base + k] = c;
    }
}

// Each uniform's first index with cdf[i] / cdf[n-1] > u.  The quotient is
// NumPy's `cdf /= cdf[-1]` element (one correctly rounded FP64 division), taken
// at the compared index only, so the CDF buffer is never rewritten: an
// in-place normalisation would race with workgroups that still read cdf[n-1].
__global__ void __launch_bounds__(kBlock) k_cdf_search(const double* __restrict__ cdf, int64_t n,
                                                       const double* __restrict__ u, int64_t nshots,
                                                       int64_t* __restrict__ samples) {
  const double total = cdf[n - 1];
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < nshots; j += (int64_t)gridDim.x * kBlock) {
    const double x = u[j];
    int64_t lo = 0, hi = n;
    while (lo < hi) {
      const int64_t mid = lo + (hi - lo) / 2;
      if (__ddiv_rn(cdf[mid], total) > x)
        hi = mid;
      else
        lo = mid + 1;
    }
    // cdf[n-1] / total == 1 > u for u in [0, 1): lo < n; kept in range regardless
    samples[j] = lo < n ? lo : n - 1;
  }
}

// frequencies[s] += 1 for every sample s in [0, nstates); others are counted
// in *bad (device int64) and left out.
__global__ void __launch_bounds__(kBlock) k_count_samples(const int64_t* __restrict__ samples, int64_t nshots,
                                                          unsigned long long* __restrict__ freqs, int64_t nstates,
                                                          unsigned long long* __restrict__ bad) {
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < nshots; j += (int64_t)gridDim.x * kBlock) {
    const int64_t v = samples[j];
    if (v >= 0 && v < nstates)
      atomicAdd(freqs + v, 1ull);
    else
      atomicAdd(bad, 1ull);
  }
}

// Doubles of work memory a sample of `nstates` needs: the CDF, then the
// block sums of the blocked scan.
static constexpr int64_t cdf_work_doubles(int64_t nstates) {
  const int64_t per_block = (int64_t)kBlock * kCdfRun;
  return nstates + (nstates + per_block - 1) / per_block;
}

// The CDF pipeline and the search on caller-provided work memory `cdf`
// (cdf_work_doubles(nstates) doubles).  Every element it reads was written by
// an earlier kernel of the same call, so the memory needs no clearing.
template <typename T>
static int sample_shots_on(const T* probs, int64_t nstates, const double* u, int64_t nshots, int64_t* samples,
                           double* cdf, hipStream_t s) {
  const int64_t per_block = (int64_t)kBlock * kCdfRun;
  const int64_t nblocks = (nstates + per_block - 1) / per_block;
  double* bsum = cdf + nstates;
  if (nstates <= kSeqCdfStates) {
    k_cdf_seq<T><<<dim3(1), dim3(64), 0, s>>>(probs, nstates, cdf);
  } else {
    k_cdf_block_sums<T><<<dim3((unsigned)nblocks), dim3(kBlock), 0, s>>>(probs, nstates, bsum);
    k_cdf_block_offsets<<<dim3(1), dim3(64), 0, s>>>(bsum, nblocks);
    k_cdf_block_scan<T><<<dim3((unsigned)nblocks), dim3(kBlock), 0, s>>>(probs, nstates, bsum, cdf);
  }
  k_cdf_search<<<dim3((unsigned)grid_for(nshots, kBlock)), dim3(kBlock), 0, s>>>(cdf, nstates, u, nshots, samples);
  return check_launch("sample_shots launch");
}

template <typename T>
static int sample_shots_impl(const T* probs, int64_t nstates, const double* u, int64_t nshots, int64_t* samples,
                             hipStream_t s) {
  if (nstates < 1) return set_error(QJ_EVALUE, "sample_shots: nstates must be >= 1");
  if (nshots < 0) return set_error(QJ_EVALUE, "sample_shots: nshots must be >= 0");
  if (nshots == 0) return QJ_OK;
  DeviceScratch* sc = nullptr;
  int rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  rc = scratch_reserve_big(sc, (size_t)cdf_work_doubles(nstates) * sizeof(double), s);
  if (rc) return rc;
  return sample_shots_on<T>(probs, nstates, u, nshots, samples, static_cast<double*>(sc->big), s);
}

// ------------------------------------- shot sampling without allocation --
// Work memory of qj_sample_shots_noalloc: a file-scope device array, one copy
// per GPU, alive until the process ends.  Host code never uses the name
// `scratch` as an address (on the host it is only a stand-in for the device
// symbol); it uses the per-device address looked up once with
// hipGetSymbolAddress.
constexpr int64_t kNoAllocMaxStates = int64_t(1) << 20;
constexpr int64_t kNoAllocScratchDoubles = kNoAllocMaxStates + kNoAllocMaxStates / ((int64_t)kBlock * kCdfRun);
static_assert(cdf_work_doubles(kNoAllocMaxStates) <= kNoAllocScratchDoubles, "noalloc work buffer too small");
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mutex;
static double* g_noalloc_ptr[kNoAllocMaxDevices] = {nullptr};

// Device address of `scratch` on the current device (looked up on the first
// call on that device, with that device current).
static int noalloc_scratch(double** out) {
  int dev = 0;
  int rc = hip_status(hipGetDevice(&dev), "hipGetDevice");
  if (rc) return rc;
  if (dev < 0 || dev >= kNoAllocMaxDevices) return set_error(QJ_EHIP, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  if (!g_noalloc_ptr[dev]) {
    double* ptr = nullptr;
    rc = hip_status(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    if (!ptr) return set_error(QJ_EHIP, "hipGetSymbolAddress(scratch) returned a null address");
    g_noalloc_ptr[dev] = ptr;
  }
  *out = g_noalloc_ptr[dev];
  return QJ_OK;
}

template <typename T>
static int sample_shots_noalloc_impl(const T* probs, int64_t nstates, const double* u, int64_t nshots,
                                     int64_t* samples, hipStream_t s) {
  if (nstates < 1) return set_error(QJ_EVALUE, "sample_shots_noalloc: nstates must be >= 1");
  if (nstates > kNoAllocMaxStates)
    return set_error(QJ_EVALUE, "sample_shots_noalloc: %lld states exceed the fixed work buffer (at most %lld states)",
                     (long long)nstates, (long long)kNoAllocMaxStates);
  if (nshots < 0) return set_error(QJ_EVALUE, "sample_shots_noalloc: nshots must be >= 0");
  if (nshots == 0) return QJ_OK;
  double* cdf = nullptr;
  int rc = noalloc_scratch(&cdf);
  if (rc) return rc;
  return sample_shots_on<T>(probs, nstates, u, nshots, samples, cdf, s);
}

}  // namespace qj

using namespace qj;

extern "C" {

int qj_initial_state(void* state, int dtype, uint64_t n, void* stream) {
  if (!state) return set_error(QJ_EVALUE, "initial_state: null state");
  hipStream_t s = (hipStream_t)stream;
  const u64 grid = grid_for(n, kBlock);
  if (dtype == QJ_COMPLEX64)
    k_initial_state<float><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>((cplx<float>*)state, n);
  else if (dtype == QJ_COMPLEX128)
    k_initial_state<double><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>((cplx<double>*)state, n);
  else
    return set_error(QJ_EVALUE, "initial_state: dtype must be complex64/complex128");
  return check_launch("initial_state launch");
}

int qj_collapse_state(void* state, int dtype, int nqubits, const int32_t* qubits_kernel,
                      int nmeasured, int64_t result, int normalize, double* norm_out,
                      void* stream) {
  if (!state || !qubits_kernel) return set_error(QJ_EVALUE, "collapse_state: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64)
    return collapse_impl<float>(state, nqubits, qubits_kernel, nmeasured, result, normalize, norm_out, s);
  if (dtype == QJ_COMPLEX128)
    return collapse_impl<double>(state, nqubits, qubits_kernel, nmeasured, result, normalize, norm_out, s);
  return set_error(QJ_EVALUE, "collapse_state: dtype must be complex64/complex128");
}

int qj_calculate_probabilities_dm(const void* state, int dtype, int nqubits, const int32_t* qubits,
                                  int nmeasured, void* probs, int prob_dtype, void* stream) {
  if (!state || !qubits || !probs) return set_error(QJ_EVALUE, "calculate_probabilities: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64 && prob_dtype == QJ_FLOAT32)
    return probs_impl<float, float>(state, nqubits, qubits, nmeasured, (float*)probs, s, true);
  if (dtype == QJ_COMPLEX64 && prob_dtype == QJ_FLOAT64)
    return probs_impl<float, double>(state, nqubits, qubits, nmeasured, (double*)probs, s, true);
  if (dtype == QJ_COMPLEX128 && prob_dtype == QJ_FLOAT32)
    return probs_impl<double, float>(state, nqubits, qubits, nmeasured, (float*)probs, s, true);
  if (dtype == QJ_COMPLEX128 && prob_dtype == QJ_FLOAT64)
    return probs_impl<double, double>(state, nqubits, qubits, nmeasured, (double*)probs, s, true);
  return set_error(QJ_EVALUE, "calculate_probabilities: unsupported dtype pair");
}

int qj_calculate_probabilities(const void* state, int dtype, int nqubits, const int32_t* qubits,
                               int nmeasured, void* probs, int prob_dtype, void* stream) {
  if (!state || !qubits || !probs) return set_error(QJ_EVALUE, "calculate_probabilities: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64 && prob_dtype == QJ_FLOAT32)
    return probs_impl<float, float>(state, nqubits, qubits, nmeasured, (float*)probs, s);
  if (dtype == QJ_COMPLEX64 && prob_dtype == QJ_FLOAT64)
    return probs_impl<float, double>(state, nqubits, qubits, nmeasured, (double*)probs, s);
  if (dtype == QJ_COMPLEX128 && prob_dtype == QJ_FLOAT32)
    return probs_impl<double, float>(state, nqubits, qubits, nmeasured, (float*)probs, s);
  if (dtype == QJ_COMPLEX128 && prob_dtype == QJ_FLOAT64)
    return probs_impl<double, double>(state, nqubits, qubits, nmeasured, (double*)probs, s);
  return set_error(QJ_EVALUE, "calculate_probabilities: unsupported dtype pair");
}

int qj_measure_frequencies(int64_t* frequencies, const void* probs, int prob_dtype, int64_t nstates,
                           int64_t nshots, const uint32_t* chain_seeds, int nchains, void* stream) {
  if (!frequencies || !probs || !chain_seeds)
    return set_error(QJ_EVALUE, "measure_frequencies: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (prob_dtype == QJ_FLOAT32)
    return metropolis_impl<float>(frequencies, (const float*)probs, nstates, nshots, chain_seeds, nchains, s);
  if (prob_dtype == QJ_FLOAT64)
    return metropolis_impl<double>(frequencies, (const double*)probs, nstates, nshots, chain_seeds, nchains, s);
  return set_error(QJ_EVALUE, "measure_frequencies: probs must be float32/float64");
}

int qj_sample_shots(const void* probs, int prob_dtype, int64_t nstates, const double* uniforms, int64_t nshots,
                    int64_t* samples, void* stream) {
  if (!probs || (nshots > 0 && (!uniforms || !samples))) return set_error(QJ_EVALUE, "sample_shots: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (prob_dtype == QJ_FLOAT32)
    return sample_shots_impl<float>((const float*)probs, nstates, uniforms, nshots, samples, s);
  if (prob_dtype == QJ_FLOAT64)
    return sample_shots_impl<double>((const double*)probs, nstates, uniforms, nshots, samples, s);
  return set_error(QJ_EVALUE, "sample_shots: probs must be float32/float64");
}

int qj_sample_shots_noalloc(const void* probs, int prob_dtype, int64_t nstates, const double* uniforms,
                            int64_t nshots, int64_t* samples, void* stream) {
  if (!probs || (nshots > 0 && (!uniforms || !samples)))
    return set_error(QJ_EVALUE, "sample_shots_noalloc: null pointer");
  hipStream_t s = (hipStream_t)stream;
  if (prob_dtype == QJ_FLOAT32)
    return sample_shots_noalloc_impl<float>((const float*)probs, nstates, uniforms, nshots, samples, s);
  if (prob_dtype == QJ_FLOAT64)
    return sample_shots_noalloc_impl<double>((const double*)probs, nstates, uniforms, nshots, samples, s);
  return set_error(QJ_EVALUE, "sample_shots_noalloc: probs must be float32/float64");
}

int qj_count_samples(const int64_t* samples, int64_t nshots, int64_t* frequencies, int64_t nstates,
                     int64_t* out_of_range, void* stream) {
  if (nshots < 0 || nstates < 1) return set_error(QJ_EVALUE, "count_samples: bad sizes");
  if (!frequencies || !out_of_range || (nshots > 0 && !samples))
    return set_error(QJ_EVALUE, "count_samples: null pointer");
  if (nshots == 0) return QJ_OK;
  hipStream_t s = (hipStream_t)stream;
  k_count_samples<<<dim3((unsigned)grid_for(nshots, kBlock)), dim3(kBlock), 0, s>>>(
      samples, nshots, reinterpret_cast<unsigned long long*>(frequencies), nstates,
      reinterpret_cast<unsigned long long*>(out_of_range));
  return check_launch("count_samples launch");
}

int qj_norm_squared(const void* state, int dtype, uint64_t n, double* out, void* stream) {
  if (!state || !out) return set_error(QJ_EVALUE, "norm_squared: null pointer");
  hipStream_t s = (hipStream_t)stream;
  DeviceScratch* sc = nullptr;
  int rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  const u64 grid = std::min<u64>(kProbsMaxBlocks, grid_for(n / 2 + 1, kBlock));
  const bool a16 = ((uintptr_t)state & 15) == 0;
  if (dtype == QJ_COMPLEX64 && a16)
    k_norm2<float, 2><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>((const cplx<float>*)state, n, sc->partials);
  else if (dtype == QJ_COMPLEX64)
    k_norm2<float, 1><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>((const cplx<float>*)state, n, sc->partials);
  else if (dtype == QJ_COMPLEX128)
    k_norm2<double, 1><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>((const cplx<double>*)state, n, sc->partials);
  else
    return set_error(QJ_EVALUE, "norm_squared: dtype must be complex");
  rc = check_launch("norm launch");
  if (rc) return rc;
  return reduce_to_host(1, grid, 0, 1, out, s, sc);
}

int qj_vdot(const void* a, const void* b, int dtype, uint64_t n, double* out, void* stream) {
  if (!a || !b || !out) return set_error(QJ_EVALUE, "vdot: null pointer");
  hipStream_t s = (hipStream_t)stream;
  DeviceScratch* sc = nullptr;
  int rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  const u64 grid = std::min<u64>(kProbsMaxBlocks, grid_for(n / 2 + 1, kBlock));
  const bool a16 = (((uintptr_t)a | (uintptr_t)b) & 15) == 0;
  const dim3 gd((unsigned)grid), bd(kBlock);
  if (dtype == QJ_COMPLEX64 && a16)
    k_vdot<float, 2><<<gd, bd, 0, s>>>((const cplx<float>*)a, (const cplx<float>*)b, n, sc->partials);
  else if (dtype == QJ_COMPLEX64)
    k_vdot<float, 1><<<gd, bd, 0, s>>>((const cplx<float>*)a, (const cplx<float>*)b, n, sc->partials);
  else if (dtype == QJ_COMPLEX128)
    k_vdot<double, 1><<<gd, bd, 0, s>>>((const cplx<double>*)a, (const cplx<double>*)b, n, sc->partials);
  else
    return set_error(QJ_EVALUE, "vdot: dtype must be complex");
  rc = check_launch("vdot launch");
  if (rc) return rc;
  return reduce_to_host(2, grid, grid, 1, out, s, sc);
}

int qj_fill_state(void* state, int dtype, uint64_t n, double re, double im, void* stream) {
  if (!state) return set_error(QJ_EVALUE, "fill_state: null state");
  if (n == 0) return QJ_OK;
  hipStream_t s = (hipStream_t)stream;
  const u64 grid = grid_for(n, kBlock);
  if (dtype == QJ_COMPLEX64)
    k_fill<float><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>((cplx<float>*)state, n, mk<float>((float)re, (float)im));
  else if (dtype == QJ_COMPLEX128)
    k_fill<double><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>((cplx<double>*)state, n, mk<double>(re, im));
  else
    return set_error(QJ_EVALUE, "fill_state: dtype must be complex");
  return check_launch("fill launch");
}

int qj_compare_states(const void* a, int dtype_a, const void* b, int dtype_b, uint64_t n, double* out,
                      void* stream) {
  if (!a || !b || !out) return set_error(QJ_EVALUE, "compare_states: null pointer");
  if (!is_complex(dtype_a) || !is_complex(dtype_b)) return set_error(QJ_EVALUE, "compare_states: complex states only");
  if (n == 0) return set_error(QJ_EVALUE, "compare_states: empty states");
  hipStream_t s = (hipStream_t)stream;
  DeviceScratch* sc = nullptr;
  int rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  const u64 grid = std::min<u64>(kRedBlocks, grid_for(n, kBlock));
  const dim3 gd((unsigned)grid), bd(kBlock);
  const bool a64 = dtype_a == QJ_COMPLEX64, b64 = dtype_b == QJ_COMPLEX64;
  if (a64 && b64)
    k_compare<float, float><<<gd, bd, 0, s>>>((const cplx<float>*)a, (const cplx<float>*)b, n, sc->partials);
  else if (a64)
    k_compare<float, double><<<gd, bd, 0, s>>>((const cplx<float>*)a, (const cplx<double>*)b, n, sc->partials);
  else if (b64)
    k_compare<double, float><<<gd, bd, 0, s>>>((const cplx<double>*)a, (const cplx<float>*)b, n, sc->partials);
  else
    k_compare<double, double><<<gd, bd, 0, s>>>((const cplx<double>*)a, (const cplx<double>*)b, n, sc->partials);
  rc = check_launch("compare launch");
  if (rc) return rc;
  std::vector<double> host(5 * grid);
  rc = hip_status(hipMemcpyAsync(host.data(), sc->partials, host.size() * sizeof(double), hipMemcpyDeviceToHost, s),
                  "compare readback");
  if (rc) return rc;
  rc = hip_status(hipStreamSynchronize(s), "compare sync");
  if (rc) return rc;
  double v[5] = {0, 0, 0, 0, 0};
  for (u64 k = 0; k < grid; ++k) {
    for (int j = 0; j < 4; ++j) v[j] += host[j * grid + k];
    v[4] = std::max(v[4], host[4 * grid + k]);
  }
  out[0] = v[0];
  out[1] = v[1];
  out[2] = v[2];
  out[3] = v[3];
  out[4] = v[4];
  return QJ_OK;
}

int qj_expectation_tfim(const void* state, int dtype, int nqubits, double g, double* out,
                        void* stream) {
  if (!state || !out) return set_error(QJ_EVALUE, "expectation_tfim: null pointer");
  if (nqubits < 2 || nqubits > 40) return set_error(QJ_EVALUE, "expectation_tfim: nqubits %d out of range", nqubits);
  hipStream_t s = (hipStream_t)stream;
  DeviceScratch* sc = nullptr;
  int rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  const u64 n = u64(1) << nqubits;
  const u64 grid = std::min<u64>(kRedBlocks, grid_for(n, kBlock));
  if (dtype == QJ_COMPLEX64)
    k_tfim<float><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>((const cplx<float>*)state, nqubits, g, sc->partials);
  else if (dtype == QJ_COMPLEX128)
    k_tfim<double><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>((const cplx<double>*)state, nqubits, g, sc->partials);
  else
    return set_error(QJ_EVALUE, "expectation_tfim: dtype must be complex");
  rc = check_launch("tfim launch");
  if (rc) return rc;
  return reduce_to_host(1, grid, 0, 1, out, s, sc);
}

int qj_axpby(const void* x, void* y, int dtype, uint64_t n, const double* alpha, const double* beta,
             void* stream) {
  if (!x || !y || !alpha || !beta) return set_error(QJ_EVALUE, "axpby: null pointer");
  hipStream_t s = (hipStream_t)stream;
  const u64 grid = grid_for(n, kBlock);
  if (dtype == QJ_COMPLEX64)
    k_axpby<float><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(
        (const cplx<float>*)x, (cplx<float>*)y, n, mk<float>((float)alpha[0], (float)alpha[1]),
        mk<float>((float)beta[0], (float)beta[1]));
  else if (dtype == QJ_COMPLEX128)
    k_axpby<double><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(
        (const cplx<double>*)x, (cplx<double>*)y, n, mk<double>(alpha[0], alpha[1]), mk<double>(beta[0], beta[1]));
  else
    return set_error(QJ_EVALUE, "axpby: dtype must be complex");
  return check_launch("axpby launch");
}

}  // extern "C"
