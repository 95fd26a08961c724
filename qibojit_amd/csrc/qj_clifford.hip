// Stabilizer-tableau kernels on a column-major, bit-packed tableau.
//
// The tableau of n qubits has C = 2n+1 rows (destabilizers 0..n-1,
// stabilizers n..2n-1, scratch 2n) and C columns (x bits 0..n-1, z bits
// n..2n-1, sign r 2n).  Here it is stored by column: column c is a bit
// vector over the rows, W = ceil(C/32) 32-bit words, row j in bit j%32 of
// word j/32, and columns follow each other at a stride of ws = W rounded up
// to a multiple of 4 words.  Bits of rows >= C are zero and stay zero under
// every rule.
//
// Every Clifford rule rewrites a row from that row's own bits, so a gate on
// qubit q touches only the columns x_q, z_q (x_t, z_t) and r, and one
// bitwise operation on a word updates 32 rows.  A gate list runs in one
// launch: each lane owns one word of every column (32 rows), keeps r in a
// register and, where the workgroup's slice fits, works out of LDS; the host
// groups gates on disjoint qubits so that the loads of a group issue
// together.  Measurement (rowsum) works on the same layout: its phase sums
// run bit-sliced over 32 rows at a time with a 2-bit counter per row.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qj_common.h"

namespace qj {
namespace {

enum : int {
  kcH = QJ_CLIFF_H, kcS = QJ_CLIFF_S, kcZ = QJ_CLIFF_Z, kcX = QJ_CLIFF_X, kcY = QJ_CLIFF_Y,
  kcSX = QJ_CLIFF_SX, kcSDG = QJ_CLIFF_SDG, kcSXDG = QJ_CLIFF_SXDG, kcRYpi = QJ_CLIFF_RY_PI,
  kcRY3pi2 = QJ_CLIFF_RY_3PI_2, kcCNOT = QJ_CLIFF_CNOT, kcCZ = QJ_CLIFF_CZ, kcSWAP = QJ_CLIFF_SWAP,
  kciSWAP = QJ_CLIFF_ISWAP, kcCY = QJ_CLIFF_CY, kcNop = 15
};
constexpr int kcFirstTwo = kcCNOT;
// Gate lists only: a composed one-qubit Clifford, slot word 1 holding its
// code (bits 0-3: m00 m01 m10 m11 of x' = m00 x ^ m01 z, z' = m10 x ^ m11 z;
// bits 4-6: a, b, c of the sign update a x ^ b z ^ c xz).
constexpr int kcOne = 15;
constexpr int kCliffGroup = 16;         // gates per group (pairwise disjoint qubits)
constexpr int kCliffMaxQubits = 1 << 19;
constexpr int kMaxColumnChunks = 64;   // column chunks of a sliced rowsum
constexpr size_t kCliffLdsPerCu = 160 * 1024;

inline int64_t cliff_words(int n) { return (2 * (int64_t)n + 1 + 31) / 32; }
inline int64_t cliff_stride(int n) { return (cliff_words(n) + 3) & ~(int64_t)3; }

// The Clifford rules (src/qibojit/backends/clifford_operations_cpu.py:
// 18-248) on the words of one row block: a, b = x, z of qubit qa (the
// control), c, d = x, z of qubit qb (the target), all read before the gate.
// The columns a gate changes are stored through col (column c of the block
// at col[c * stride]; pa = qa * stride, pb = qb * stride, zoff = n * stride)
// and the sign update is returned, to be XORed into r.  Each sign update is
// the reference's Boolean function of the row's bits in a shorter form with
// the same truth table: (x&z)^(x&(z^x)) = x for Z, z for X, x^z for Y,
// z&~x for SX and RY(3pi/2), x&~z for SDG and RY(pi/2); of the two-qubit
// rules SWAP's is identically 0, CZ's a&c&(b^d), iSWAP's
// (~a&c&d)^(a&b&~c), CY's a&(b^c)&~(b^d) (the 16-entry tables of the
// reference's expressions, checked against the oracle bit for bit).
template <typename P>
__device__ __forceinline__ uint32_t cliff_apply(int kind, P* col, int64_t pa, int64_t pb, int64_t zoff, uint32_t a,
                                                uint32_t b, uint32_t c, uint32_t d) {
  switch (kind) {
    case kcH: col[pa] = b; col[zoff + pa] = a; return a & b;
    case kcS: col[zoff + pa] = a ^ b; return a & b;
    case kcZ: return a;
    case kcX: return b;
    case kcY: return a ^ b;
    case kcSX: col[pa] = a ^ b; return b & ~a;
    case kcSDG: col[zoff + pa] = a ^ b; return a & ~b;
    case kcSXDG: col[pa] = a ^ b; return a & b;
    case kcRYpi: col[pa] = b; col[zoff + pa] = a; return a & ~b;
    case kcRY3pi2: col[pa] = b; col[zoff + pa] = a; return b & ~a;
    case kcCNOT: col[pb] = c ^ a; col[zoff + pa] = b ^ d; return a & d & ~(b ^ c);
    case kcCZ: col[zoff + pa] = b ^ c; col[zoff + pb] = d ^ a; return a & c & (b ^ d);
    case kcSWAP: col[pa] = c; col[pb] = a; col[zoff + pa] = d; col[zoff + pb] = b; return 0u;
    case kciSWAP:
      col[pa] = c;
      col[pb] = a;
      col[zoff + pa] = c ^ d ^ a;
      col[zoff + pb] = c ^ b ^ a;
      return (~a & c & d) ^ (a & b & ~c);
    case kcCY: col[pb] = a ^ c; col[zoff + pa] = b ^ d ^ c; col[zoff + pb] = d ^ a; return a & (b ^ c) & ~(b ^ d);
    default: return 0u;  // empty slot
  }
}

__device__ __forceinline__ uint32_t bit_of(const uint32_t* col, int64_t row) {
  return (col[row >> 5] >> (row & 31)) & 1u;
}

// Mask of the rows of word w that are below `limit`.
__device__ __forceinline__ uint32_t rows_below(int64_t w, int64_t limit) {
  const int64_t lo = w * 32;
  if (limit <= lo) return 0u;
  if (limit >= lo + 32) return ~0u;
  return (1u << (limit - lo)) - 1u;
}

// ---- layout conversion ----

// dense: C x C bytes, row-major, nonzero = 1.  Thread per (column, word).
__global__ void k_cliff_pack(const uint8_t* __restrict__ dense, int64_t C, int64_t W, int64_t ws,
                             uint32_t* __restrict__ tab) {
  const int64_t total = C * ws;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = e % C, w = e / C;
    uint32_t v = 0;
    if (w < W) {
      const int64_t r0 = w * 32;
      const int nb = (int)std::min<int64_t>(32, C - r0);
      for (int b = 0; b < nb; ++b) v |= (dense[(r0 + b) * C + c] != 0 ? 1u : 0u) << b;
    }
    tab[c * ws + w] = v;
  }
}

// Thread per byte of the dense C x C matrix.
__global__ void k_cliff_unpack(const uint32_t* __restrict__ tab, int64_t C, int64_t ws, uint8_t* __restrict__ dense) {
  const int64_t total = C * C;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = e / C, c = e - row * C;
    dense[e] = (uint8_t)bit_of(tab + c * ws, row);
  }
}

// ---- gates ----

// One gate, thread per word.
__global__ void k_cliff_gate(uint32_t* __restrict__ tab, int n, int64_t ws, int64_t W, int kind, int qa, int qb) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= W) return;
  uint32_t* col = tab + w;
  const int64_t pa = (int64_t)qa * ws, zoff = (int64_t)n * ws;
  const int64_t pb = kind >= kcFirstTwo ? (int64_t)qb * ws : pa;
  const uint32_t a = col[pa], b = col[zoff + pa], c = col[pb], d = col[zoff + pb];
  const uint32_t t = cliff_apply(kind, col, pa, pb, zoff, a, b, c, d);
  col[2 * zoff] ^= t;
}

// A gate list, one wave per workgroup, lane l owning word w0 + l of every
// column (lanes past the slice repeat its last word, so every lane runs the
// same code and stores the same values).  LDS = true stages the slice
// (C x T words, column-major, lanes adjacent) in LDS for the whole list.
// desc: ngroups x kCliffGroup slots of two words {qa, qb}; slot 0 carries
// the group's kind in bits 24.. of its first word and its gate count in
// bits 24.. of its second.  The
// descriptors pass through an LDS ring of two chunks of kDescChunk groups:
// the next chunk is loaded from global memory while the current one runs,
// and the next group's words are read while the current group's columns
// are in flight, so no group waits on global memory for its descriptors.
constexpr int kDescWords = 2 * kCliffGroup;            // words per group
static_assert(kDescWords == QJ_CLIFF_GROUP_WORDS, "descriptor words per group");
constexpr int kDescChunk = 1024 / kDescWords;            // groups per chunk
constexpr int kDescChunkWords = kDescChunk * kDescWords;  // 1024
constexpr int kDescPerLane = kDescChunkWords / 64;      // 16
constexpr size_t kDescRingBytes = 2 * kDescChunkWords * sizeof(uint32_t);
constexpr uint32_t kSlotMask = (1u << 24) - 1;  // qubit field of a slot word

__device__ __forceinline__ void read_group(const uint32_t* p, uint32_t (&dw)[kDescWords]) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
  for (int k = 0; k < kDescWords / 4; ++k) {
    const uint4 v = q[k];
    dw[4 * k] = v.x;
    dw[4 * k + 1] = v.y;
    dw[4 * k + 2] = v.z;
    dw[4 * k + 3] = v.w;
  }
}

// One group of up to kCliffGroup gates of one kind on pairwise disjoint
// qubits, run with S >= cnt slots: every load first, then the rule slot by
// slot.  Slots past `cnt` repeat slot 0 (same loads, same values stored) and
// their sign updates are masked off, so the group runs without branches.
// The next group's descriptor words are read from LDS while this group's
// columns load.
template <int KIND, int S>
__device__ __forceinline__ uint32_t run_group(uint32_t* col0, int64_t zoff, int64_t cs, uint32_t qmax, int cnt,
                                              const uint32_t* cur, int g, int ng, uint32_t (&dw)[kDescWords]) {
  constexpr bool two = KIND >= kcFirstTwo && KIND != kcOne;
  int64_t oa[S], ob[S];
  uint32_t code[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    // clamped: a descriptor array from outside stays inside the tableau
    oa[j] = (int64_t)std::min(__builtin_amdgcn_readfirstlane(dw[2 * j]) & kSlotMask, qmax) * cs;
    code[j] = __builtin_amdgcn_readfirstlane(dw[2 * j + 1]) & kSlotMask;
    ob[j] = (int64_t)std::min(code[j], qmax) * cs;
  }
  uint32_t a[S], b[S], c[S], d[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    a[j] = col0[oa[j]];
    b[j] = col0[zoff + oa[j]];
    if (two) {
      c[j] = col0[ob[j]];
      d[j] = col0[zoff + ob[j]];
    } else {
      c[j] = a[j];
      d[j] = b[j];
    }
  }
  if (g + 1 < ng) read_group(cur + (g + 1) * kDescWords, dw);
  uint32_t t = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    uint32_t tj;
    if constexpr (KIND == kcOne) {
      const uint32_t k = code[j];
      const uint32_t m00 = 0u - (k & 1u), m01 = 0u - ((k >> 1) & 1u), m10 = 0u - ((k >> 2) & 1u),
                     m11 = 0u - ((k >> 3) & 1u), sa = 0u - ((k >> 4) & 1u), sb = 0u - ((k >> 5) & 1u),
                     sc = 0u - ((k >> 6) & 1u);
      col0[oa[j]] = (a[j] & m00) ^ (b[j] & m01);
      col0[zoff + oa[j]] = (a[j] & m10) ^ (b[j] & m11);
      tj = (a[j] & sa) ^ (b[j] & sb) ^ (a[j] & b[j] & sc);
    } else {
      tj = cliff_apply(KIND, col0, oa[j], ob[j], zoff, a[j], b[j], c[j], d[j]);
    }
    t ^= j < cnt ? tj : 0u;
  }
  return t;
}

// Smallest slot count that holds the group (2, 4, 8 or kCliffGroup).
template <int KIND>
__device__ __forceinline__ uint32_t run_sized(uint32_t* col0, int64_t zoff, int64_t cs, uint32_t qmax, int cnt,
                                              const uint32_t* cur, int g, int ng, uint32_t (&dw)[kDescWords]) {
  if (cnt <= 2) return run_group<KIND, 2>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw);
  if (cnt <= 4) return run_group<KIND, 4>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw);
  if (cnt <= 8) return run_group<KIND, 8>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw);
  return run_group<KIND, kCliffGroup>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw);
}

template <bool LDS>
__global__ __launch_bounds__(64) void k_cliff_circuit(uint32_t* __restrict__ tab, int n, int64_t ws, int64_t W,
                                                      int T, const uint32_t* __restrict__ desc, int64_t ngroups) {
  extern __shared__ __align__(16) uint32_t lds_words[];
  uint32_t* ring = lds_words;
  uint32_t* slice = lds_words + 2 * kDescChunkWords;
  const int lane = threadIdx.x;
  const int64_t w0 = (int64_t)blockIdx.x * T;
  const int nact = (int)std::min<int64_t>(T, W - w0);
  const int my = lane < nact ? lane : nact - 1;
  const int64_t C = 2 * (int64_t)n + 1;
  uint32_t* col0;
  int64_t cs;
  if (LDS) {
    const int64_t total = C * T;
    for (int64_t e = lane; e < total; e += 64) {
      const int64_t c = e / T;
      const int tt = (int)(e - c * T);
      slice[e] = tt < nact ? tab[c * ws + w0 + tt] : 0u;
    }
    col0 = slice + my;
    cs = T;
  } else {
    col0 = tab + w0 + my;
    cs = ws;
  }
  const int64_t total_words = ngroups * kDescWords;
  const int64_t nchunks = (ngroups + kDescChunk - 1) / kDescChunk;
  // chunk 0 into ring buffer 0
#pragma unroll
  for (int k = 0; k < kDescPerLane; ++k) {
    const int64_t e = (int64_t)k * 64 + lane;
    ring[e] = e < total_words ? desc[e] : 0u;
  }
  __syncthreads();
  const int64_t zoff = (int64_t)n * cs;
  const int64_t roff = 2 * (int64_t)n * cs;
  const uint32_t qmax = (uint32_t)(n - 1);
  uint32_t r = col0[roff];
  for (int64_t ch = 0; ch < nchunks; ++ch) {
    const uint32_t* cur = ring + (ch & 1) * kDescChunkWords;
    uint32_t nx[kDescPerLane];
    const bool more = ch + 1 < nchunks;
    if (more) {
#pragma unroll
      for (int k = 0; k < kDescPerLane; ++k) {
        const int64_t e = (ch + 1) * kDescChunkWords + (int64_t)k * 64 + lane;
        nx[k] = e < total_words ? desc[e] : 0u;
      }
    }
    const int ng = (int)std::min<int64_t>(kDescChunk, ngroups - ch * kDescChunk);
    uint32_t dw[kDescWords];
    read_group(cur, dw);
    for (int g = 0; g < ng; ++g) {
      const int kind = (int)(__builtin_amdgcn_readfirstlane(dw[0]) >> 24);
      const int cnt = (int)(__builtin_amdgcn_readfirstlane(dw[1]) >> 24);
      uint32_t t;
      switch (kind) {
        case kcH: t = run_sized<kcH>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcS: t = run_sized<kcS>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcZ: t = run_sized<kcZ>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcX: t = run_sized<kcX>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcY: t = run_sized<kcY>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcSX: t = run_sized<kcSX>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcSDG: t = run_sized<kcSDG>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcSXDG: t = run_sized<kcSXDG>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcRYpi: t = run_sized<kcRYpi>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcRY3pi2: t = run_sized<kcRY3pi2>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcCNOT: t = run_sized<kcCNOT>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcCZ: t = run_sized<kcCZ>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcSWAP: t = run_sized<kcSWAP>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kciSWAP: t = run_sized<kciSWAP>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        case kcOne: t = run_sized<kcOne>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
        default: t = run_sized<kcCY>(col0, zoff, cs, qmax, cnt, cur, g, ng, dw); break;
      }
      r ^= t;
    }
    if (more) {
      uint32_t* nxt = ring + ((ch + 1) & 1) * kDescChunkWords;
#pragma unroll
      for (int k = 0; k < kDescPerLane; ++k) nxt[k * 64 + lane] = nx[k];
      __syncthreads();
    }
  }
  col0[roff] = r;
  if (LDS) {
    __syncthreads();
    const int64_t total = C * T;
    for (int64_t e = lane; e < total; e += 64) {
      const int64_t c = e / T;
      const int tt = (int)(e - c * T);
      if (tt < nact) tab[c * ws + w0 + tt] = slice[e];
    }
  }
}

// ---- measurement ----
//
// The kernels of one measured qubit take an optional control block `ctl`
// (device int32: [0] the first stabilizer p - n or -1, [1] the random
// outcomes taken so far, [2] the determined outcome).  With ctl null they use
// their scalar arguments; with ctl set, p comes from ctl[0] and each kernel
// returns at once when its branch (random: p >= 0, determined: p < 0) is not
// the one taken, so a whole measurement loop is enqueued without a host
// round trip.

// First stabilizer row (n <= row < 2n) with x_q = 1, as row - n, or -1.
__global__ __launch_bounds__(256) void k_cliff_first_stab(const uint32_t* __restrict__ tab, int n, int64_t ws, int q,
                                                          int32_t* __restrict__ out) {
  __shared__ int best;
  if (threadIdx.x == 0) best = 0x7fffffff;
  __syncthreads();
  const uint32_t* col = tab + (int64_t)q * ws;
  const int64_t wlo = (int64_t)n >> 5, whi = (2 * (int64_t)n - 1) >> 5;
  int mine = 0x7fffffff;
  for (int64_t w = wlo + threadIdx.x; w <= whi; w += blockDim.x) {
    const uint32_t m = col[w] & rows_below(w, 2 * (int64_t)n) & ~rows_below(w, n);
    if (m) {
      mine = (int)(w * 32 + __builtin_ctz(m));
      break;  // words are visited in increasing order per thread
    }
  }
  if (mine != 0x7fffffff) atomicMin(&best, mine);
  __syncthreads();
  if (threadIdx.x == 0) out[0] = best == 0x7fffffff ? -1 : best - n;
}

// hm[w] = x_q bits of rows < 2n (the rows a random outcome multiplies).
__global__ void k_cliff_hmask(const uint32_t* __restrict__ tab, int64_t ws, int64_t W, int n, int q,
                              uint32_t* __restrict__ hm, const int32_t* __restrict__ ctl) {
  if (ctl && ctl[0] < 0) return;
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w < W) hm[w] = tab[(int64_t)q * ws + w] & rows_below(w, 2 * (int64_t)n);
}

// Rows h in hm (minus p) <- P_p P_h, bit-sliced: thread (w, chunk) updates
// the x/z words of columns [k0, k1) and sums the exponents of i of those
// columns into a 2-bit counter per row (c0 + 2 c1 mod 4), written to
// part[(chunk * W + w) * 2 + {0, 1}].  Row p is not rewritten, so its bits
// read by every thread stay those of the state before the call.
__global__ void k_cliff_rowsum_sliced(uint32_t* __restrict__ tab, int n, int64_t ws, int64_t W, int64_t p,
                                      const uint32_t* __restrict__ hm, int kchunk, uint32_t* __restrict__ part,
                                      const int32_t* __restrict__ ctl) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= W) return;
  if (ctl) {
    if (ctl[0] < 0) return;
    p = (int64_t)n + ctl[0];
  }
  uint32_t h = hm[w];
  if (w == (p >> 5)) h &= ~(1u << (p & 31));
  uint32_t c0 = 0, c1 = 0;
  const int k0 = blockIdx.y * kchunk;
  const int k1 = std::min(n, k0 + kchunk);
  if (h) {
    for (int k = k0; k < k1; ++k) {
      uint32_t* xc = tab + (int64_t)k * ws;
      uint32_t* zc = tab + ((int64_t)n + k) * ws;
      const uint32_t xp = bit_of(xc, p), zp = bit_of(zc, p);
      if (!(xp | zp)) continue;
      const uint32_t X = xc[w], Z = zc[w];
      uint32_t P, M;  // rows whose exponent is +1 / -1 (clifford_operations_cpu.py:361-365)
      if (xp && !zp) {
        P = X & Z;
        M = ~X & Z;
      } else if (!xp) {
        P = X & ~Z;
        M = X & Z;
      } else {
        P = ~X & Z;
        M = X & ~Z;
      }
      c1 ^= c0 & P;
      c0 ^= P;
      c1 ^= ~c0 & M;
      c0 ^= M;
      if (xp) xc[w] = X ^ h;
      if (zp) zc[w] = Z ^ h;
    }
  }
  part[(blockIdx.y * W + w) * 2] = c0;
  part[(blockIdx.y * W + w) * 2 + 1] = c1;
}

// r_h = (2 r_h + 2 r_p + sum) mod 4 != 0 for the rows of hm (minus p).
__global__ void k_cliff_rowsum_sign(uint32_t* __restrict__ tab, int n, int64_t ws, int64_t W, int64_t p,
                                    const uint32_t* __restrict__ hm, int nchunks, const uint32_t* __restrict__ part,
                                    const int32_t* __restrict__ ctl) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= W) return;
  if (ctl) {
    if (ctl[0] < 0) return;
    p = (int64_t)n + ctl[0];
  }
  uint32_t h = hm[w];
  if (w == (p >> 5)) h &= ~(1u << (p & 31));
  if (!h) return;
  uint32_t c0 = 0, c1 = 0;
#pragma unroll 8
  for (int y = 0; y < nchunks; ++y) {
    const uint32_t a0 = part[(y * W + w) * 2], a1 = part[(y * W + w) * 2 + 1];
    c1 ^= a1 ^ (c0 & a0);
    c0 ^= a0;
  }
  uint32_t* rc = tab + 2 * (int64_t)n * ws;
  const uint32_t r = rc[w];
  const uint32_t rp = bit_of(rc, p) ? ~0u : 0u;
  const uint32_t rnew = c0 | (c1 ^ r ^ rp);
  rc[w] = (r & ~h) | (rnew & h);
}

// Row p - n <- row p; row p <- (r = outcome, z_q = 1, zero elsewhere).
// Thread per column.  With ctl, the outcome is draws[ctl[1]].
__global__ void k_cliff_set_rows(uint32_t* __restrict__ tab, int n, int64_t ws, int64_t p, int q, int outcome,
                                 const int32_t* __restrict__ ctl, const uint8_t* __restrict__ draws) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t C = 2 * (int64_t)n + 1;
  if (c >= C) return;
  if (ctl) {
    if (ctl[0] < 0) return;
    p = (int64_t)n + ctl[0];
    outcome = draws[ctl[1]] & 1;
  }
  uint32_t* col = tab + c * ws;
  const int64_t d = p - n;
  const uint32_t b = bit_of(col, p);
  col[d >> 5] = (col[d >> 5] & ~(1u << (d & 31))) | (b << (d & 31));
  const uint32_t nb = c == (int64_t)n + q ? 1u : (c == C - 1 ? (uint32_t)(outcome & 1) : 0u);
  col[p >> 5] = (col[p >> 5] & ~(1u << (p & 31))) | (nb << (p & 31));
}

// Determined outcome (clifford_operations_gpu.py:628-640 with the determined
// rowsum of clifford_operations_cpu.py:337-385): the scratch row becomes the
// XOR of the stabilizer rows n+i with destabilizer bit x_q(i) = 1, its sign
// the XOR of their signs.  Workgroup (one wave) per column; the scratch bit
// of column c is written by its own wave only.
__global__ __launch_bounds__(64) void k_cliff_determined(uint32_t* __restrict__ tab, int n, int64_t ws, int q,
                                                         int32_t* out, const int32_t* ctl) {
  if (ctl && ctl[0] >= 0) return;
  const int64_t c = blockIdx.x;
  uint32_t* col = tab + c * ws;
  const uint32_t* xq = tab + (int64_t)q * ws;
  const int64_t nw = ((int64_t)n + 31) >> 5;
  const int sh = n & 31;
  uint32_t acc = 0;
  for (int64_t j = threadIdx.x; j < nw; j += 64) {
    const uint32_t m = xq[j] & rows_below(j, n);
    if (!m) continue;
    const int64_t a = (n + j * 32) >> 5;  // bits n+32j .. n+32j+31 of the column
    uint32_t v = col[a] >> sh;
    if (sh && a + 1 < ws) v |= col[a + 1] << (32 - sh);
    acc ^= v & m;
  }
  const uint64_t odd = __ballot(__builtin_popcount(acc) & 1);
  const uint32_t parity = (uint32_t)(__builtin_popcountll(odd) & 1);
  if (threadIdx.x == 0) {
    const int64_t s = 2 * (int64_t)n;
    col[s >> 5] = (col[s >> 5] & ~(1u << (s & 31))) | (parity << (s & 31));
    if (c == s) out[0] = (int32_t)parity;
  }
}

// Accumulated phase of the product of the stabilizers a determined outcome
// selects (rows n+i with destabilizer bit x_q(i) = 1, in row order j < l):
// with Y = iXZ per qubit, prod_j P_j = (-1)^(sum r_j + sum_{j<l} z_j.x_l)
// i^(sum |x_j & z_j|) Z_q (the sequential rowsum of Aaronson and Gottesman,
// Phys. Rev. A 70, 052328).  Workgroup (one wave) per qubit column k; lane
// per 32-row word, 64 words per pass.  The cross term takes an exclusive
// prefix parity of the selected z_k bits: within a word by a shift-XOR scan,
// across the wave's words by a ballot of the word parities, across passes by
// a carried bit.  The column adds (2 cross_k + Y_k) mod 4 to acc[0]; bit 1 of
// the total is then (sum cross_k + floor(sum Y_k / 2)) mod 2.
__global__ __launch_bounds__(64) void k_cliff_exact_terms(const uint32_t* __restrict__ tab, int n, int64_t ws, int q,
                                                          uint32_t* __restrict__ acc, const int32_t* __restrict__ ctl) {
  if (ctl && ctl[0] >= 0) return;
  const int64_t k = blockIdx.x;
  const int lane = (int)threadIdx.x;
  const uint32_t* xc = tab + k * ws;
  const uint32_t* zc = tab + ((int64_t)n + k) * ws;
  const uint32_t* xq = tab + (int64_t)q * ws;
  const int64_t nw = ((int64_t)n + 31) >> 5;
  const int sh = n & 31;
  const uint64_t below = (1ull << lane) - 1ull;
  uint32_t carry = 0, cross = 0, ycount = 0;
  for (int64_t base = 0; base < nw; base += 64) {
    const int64_t j = base + lane;
    uint32_t X = 0, Z = 0;
    if (j < nw) {
      const uint32_t m = xq[j] & rows_below(j, n);
      if (m) {
        const int64_t a = (n + j * 32) >> 5;  // bits n+32j .. n+32j+31 of the columns
        X = xc[a] >> sh;
        Z = zc[a] >> sh;
        if (sh && a + 1 < ws) {
          X |= xc[a + 1] << (32 - sh);
          Z |= zc[a + 1] << (32 - sh);
        }
        X &= m;
        Z &= m;
      }
    }
    uint32_t pz = Z;  // inclusive prefix parity within the word
    pz ^= pz << 1;
    pz ^= pz << 2;
    pz ^= pz << 4;
    pz ^= pz << 8;
    pz ^= pz << 16;
    const uint64_t odd = __ballot(__builtin_popcount(Z) & 1);
    const uint32_t before = carry ^ (uint32_t)(__builtin_popcountll(odd & below) & 1);
    pz = (pz << 1) ^ (before ? ~0u : 0u);  // exclusive, with every earlier row
    cross ^= (uint32_t)__builtin_popcount(X & pz) & 1u;
    ycount += (uint32_t)__builtin_popcount(X & Z);
    carry ^= (uint32_t)(__builtin_popcountll(odd) & 1);
  }
  const uint32_t c = (uint32_t)(__builtin_popcountll(__ballot(cross)) & 1);
  const uint32_t y = (uint32_t)__builtin_popcountll(__ballot(ycount & 1u)) +
                     2u * (uint32_t)__builtin_popcountll(__ballot(ycount & 2u));
  const uint32_t add = (2u * c + y) & 3u;
  if (lane == 0 && add) atomicAdd(acc, add);
}

// Exact outcome = (XOR of the selected signs, left in the scratch row's sign
// by k_cliff_determined) ^ bit 1 of acc[0]: into the scratch row's sign and
// *out; or, with `form` set, into the constant term of a symbolic outcome.
__global__ void k_cliff_exact_finish(uint32_t* __restrict__ tab, int n, int64_t ws, const uint32_t* __restrict__ acc,
                                     int32_t* out, uint8_t* __restrict__ form, const int32_t* ctl) {
  if (threadIdx.x || blockIdx.x) return;
  if (ctl && ctl[0] >= 0) return;
  const uint32_t c = (acc[0] >> 1) & 1u;
  if (form) {
    form[0] ^= (uint8_t)c;
    return;
  }
  const int64_t s = 2 * (int64_t)n;
  uint32_t* rc = tab + s * ws;
  const uint32_t o = ((rc[s >> 5] >> (s & 31)) & 1u) ^ c;
  rc[s >> 5] = (rc[s >> 5] & ~(1u << (s & 31))) | (o << (s & 31));
  out[0] = (int32_t)o;
}

// End of one measured qubit of a device loop: sample[i] = the random outcome
// draws[ctl[1]] (then ctl[1] + 1) or the determined outcome ctl[2];
// *nrandom = the random outcomes taken so far.
__global__ void k_cliff_meas_finish(int32_t* __restrict__ ctl, const uint8_t* __restrict__ draws,
                                    int32_t* __restrict__ sample, int i, int32_t* __restrict__ nrandom) {
  if (threadIdx.x || blockIdx.x) return;
  int32_t count = ctl[1];
  if (ctl[0] >= 0) {
    sample[i] = draws[count] & 1;
    ctl[1] = ++count;
  } else {
    sample[i] = ctl[2];
  }
  nrandom[0] = count;
}

// The whole measurement loop of a small tableau (n <= kMeasureSmallCap) in one
// workgroup of 1024 threads, the phases of the multi-kernel loop separated by
// barriers: first stabilizer (LDS atomicMin); random outcome as the sliced
// rowsum (thread = (word, column chunk), counters combined in LDS), sign
// update, row moves; determined outcome (wave per column, ballot parity) and
// its exact phase (wave per qubit column, LDS sum).  Every thread reaches
// every barrier; the branch taken is uniform (read from LDS).
constexpr int kMeasureSmallMax = 256;   // default: one workgroup up to here
constexpr int kMeasureSmallCap = 1024;  // largest n the kernel takes
constexpr int kMeasureThreads = 1024;

__global__ __launch_bounds__(1024) void k_cliff_measure_small(uint32_t* __restrict__ tab, int n, int64_t ws,
                                                              const int32_t* __restrict__ qubits, int m,
                                                              const uint8_t* __restrict__ draws, int exact,
                                                              int32_t* __restrict__ sample,
                                                              int32_t* __restrict__ nrandom) {
  __shared__ int s_best;
  __shared__ uint32_t s_acc;
  __shared__ uint32_t s_part[2 * kMeasureThreads];
  __shared__ uint32_t s_hm[(2 * kMeasureSmallCap + 1 + 31) / 32];
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6, nwaves = kMeasureThreads >> 6;
  const int64_t C = 2 * (int64_t)n + 1;
  const int W = (int)((C + 31) >> 5);
  const int nchunks = kMeasureThreads / W;  // >= 15 for n <= 1024
  const int kchunk = (n + nchunks - 1) / nchunks;
  const int nw = (n + 31) >> 5;
  const int sh = n & 31;
  uint32_t* rcol = tab + 2 * (int64_t)n * ws;
  int count = 0;
  for (int i = 0; i < m; ++i) {
    const int q = qubits[i];
    const uint32_t* xq = tab + (int64_t)q * ws;
    if (tid == 0) {
      s_best = 0x7fffffff;
      s_acc = 0;
    }
    __syncthreads();
    // first stabilizer row with x_q = 1; the x_q bits of rows < 2n kept in
    // LDS (the rowsum below rewrites column q)
    if (tid < W) {
      const uint32_t xw = xq[tid] & rows_below(tid, 2 * (int64_t)n);
      s_hm[tid] = xw;
      const uint32_t mk = xw & ~rows_below(tid, n);
      if (mk) atomicMin(&s_best, tid * 32 + __builtin_ctz(mk));
    }
    __syncthreads();
    const int best = s_best;
    int outcome;
    if (best != 0x7fffffff) {
      const int64_t p = best;  // row n + (p - n)
      const int w = tid % W, chunk = tid / W;
      uint32_t c0 = 0, c1 = 0;
      if (chunk < nchunks) {
        uint32_t h = s_hm[w];
        if (w == (p >> 5)) h &= ~(1u << (p & 31));
        const int k0 = chunk * kchunk, k1 = min(n, k0 + kchunk);
        if (h) {
          for (int k = k0; k < k1; ++k) {
            uint32_t* xc = tab + (int64_t)k * ws;
            uint32_t* zc = tab + ((int64_t)n + k) * ws;
            const uint32_t xp = bit_of(xc, p), zp = bit_of(zc, p);
            if (!(xp | zp)) continue;
            const uint32_t X = xc[w], Z = zc[w];
            uint32_t P, M;
            if (xp && !zp) {
              P = X & Z;
              M = ~X & Z;
            } else if (!xp) {
              P = X & ~Z;
              M = X & Z;
            } else {
              P = ~X & Z;
              M = X & ~Z;
            }
            c1 ^= c0 & P;
            c0 ^= P;
            c1 ^= ~c0 & M;
            c0 ^= M;
            if (xp) xc[w] = X ^ h;
            if (zp) zc[w] = Z ^ h;
          }
        }
      }
      s_part[2 * tid] = c0;
      s_part[2 * tid + 1] = c1;
      __syncthreads();
      if (tid < W) {
        uint32_t h = s_hm[tid];
        if (tid == (p >> 5)) h &= ~(1u << (p & 31));
        uint32_t a0 = 0, a1 = 0;
        for (int y = 0; y < nchunks; ++y) {
          const uint32_t b0 = s_part[2 * (y * W + tid)], b1 = s_part[2 * (y * W + tid) + 1];
          a1 ^= b1 ^ (a0 & b0);
          a0 ^= b0;
        }
        const uint32_t r = rcol[tid];
        const uint32_t rp = bit_of(rcol, p) ? ~0u : 0u;
        const uint32_t rnew = a0 | (a1 ^ r ^ rp);
        if (h) rcol[tid] = (r & ~h) | (rnew & h);
      }
      __syncthreads();
      outcome = draws[count] & 1;
      for (int64_t c = tid; c < C; c += kMeasureThreads) {
        uint32_t* col = tab + c * ws;
        const int64_t d = p - n;
        const uint32_t b = bit_of(col, p);
        col[d >> 5] = (col[d >> 5] & ~(1u << (d & 31))) | (b << (d & 31));
        const uint32_t nb = c == (int64_t)n + q ? 1u : (c == C - 1 ? (uint32_t)outcome : 0u);
        col[p >> 5] = (col[p >> 5] & ~(1u << (p & 31))) | (nb << (p & 31));
      }
      ++count;
    } else {
      // scratch row <- XOR of the selected stabilizer rows, wave per column
      for (int64_t c = wave; c < C; c += nwaves) {
        uint32_t* col = tab + c * ws;
        uint32_t acc = 0;
        if (lane < nw) {
          const uint32_t mk = xq[lane] & rows_below(lane, n);
          if (mk) {
            const int64_t a = ((int64_t)n + lane * 32) >> 5;
            uint32_t v = col[a] >> sh;
            if (sh && a + 1 < ws) v |= col[a + 1] << (32 - sh);
            acc = v & mk;
          }
        }
        const uint32_t parity = (uint32_t)(__builtin_popcountll(__ballot(__builtin_popcount(acc) & 1)) & 1);
        if (lane == 0) {
          const int64_t s2 = 2 * (int64_t)n;
          col[s2 >> 5] = (col[s2 >> 5] & ~(1u << (s2 & 31))) | (parity << (s2 & 31));
        }
      }
      if (exact) {
        __syncthreads();
        for (int k = wave; k < n; k += nwaves) {
          const uint32_t* xc = tab + (int64_t)k * ws;
          const uint32_t* zc = tab + ((int64_t)n + k) * ws;
          uint32_t X = 0, Z = 0;
          if (lane < nw) {
            const uint32_t mk = xq[lane] & rows_below(lane, n);
            if (mk) {
              const int64_t a = ((int64_t)n + lane * 32) >> 5;
              X = xc[a] >> sh;
              Z = zc[a] >> sh;
              if (sh && a + 1 < ws) {
                X |= xc[a + 1] << (32 - sh);
                Z |= zc[a + 1] << (32 - sh);
              }
              X &= mk;
              Z &= mk;
            }
          }
          uint32_t pz = Z;
          pz ^= pz << 1;
          pz ^= pz << 2;
          pz ^= pz << 4;
          pz ^= pz << 8;
          pz ^= pz << 16;
          const uint64_t odd = __ballot(__builtin_popcount(Z) & 1);
          const uint32_t before = (uint32_t)(__builtin_popcountll(odd & ((1ull << lane) - 1ull)) & 1);
          pz = (pz << 1) ^ (before ? ~0u : 0u);
          const uint32_t cr = (uint32_t)__builtin_popcount(X & pz) & 1u;
          const uint32_t yc = (uint32_t)__builtin_popcount(X & Z);
          const uint32_t cb = (uint32_t)(__builtin_popcountll(__ballot(cr)) & 1);
          const uint32_t y = (uint32_t)__builtin_popcountll(__ballot(yc & 1u)) +
                             2u * (uint32_t)__builtin_popcountll(__ballot(yc & 2u));
          const uint32_t add = (2u * cb + y) & 3u;
          if (lane == 0 && add) atomicAdd(&s_acc, add);
        }
      }
      __syncthreads();
      const int64_t s2 = 2 * (int64_t)n;
      outcome = (int)(((rcol[s2 >> 5] >> (s2 & 31)) & 1u) ^ (exact ? ((s_acc >> 1) & 1u) : 0u));
    }
    __syncthreads();  // every thread has read the scratch sign
    if (tid == 0) {
      sample[i] = outcome;
      if (best == 0x7fffffff && exact) {
        const int64_t s2 = 2 * (int64_t)n;
        rcol[s2 >> 5] = (rcol[s2 >> 5] & ~(1u << (s2 & 31))) | ((uint32_t)outcome << (s2 & 31));
      }
    }
    __syncthreads();
  }
  if (tid == 0) nrandom[0] = count;
}

// ---- general rowsum (clifford_operations_cpu.py:337-385) ----

// New sign of every pair from the state before the call: rnew[j], and
// flip[j] = rnew[j] ^ r_h[j].  Thread per pair.
__global__ void k_cliff_pair_sign(const uint32_t* __restrict__ tab, int n, int64_t ws, const int64_t* __restrict__ h,
                                  const int64_t* __restrict__ i, int64_t m, uint8_t* __restrict__ rnew,
                                  uint8_t* __restrict__ flip) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int64_t hj = h[j], ij = i[j];
  int e = 0;
  for (int k = 0; k < n; ++k) {
    const uint32_t* xc = tab + (int64_t)k * ws;
    const uint32_t* zc = tab + ((int64_t)n + k) * ws;
    const int xi = (int)bit_of(xc, ij), zi = (int)bit_of(zc, ij);
    const int xh = (int)bit_of(xc, hj), zh = (int)bit_of(zc, hj);
    e += 2 * (xi * xh * (zh - zi) + zi * zh * (xi - xh)) - xi * zh + xh * zi;
  }
  const uint32_t* rc = tab + 2 * (int64_t)n * ws;
  const int rh = (int)bit_of(rc, hj), ri = (int)bit_of(rc, ij);
  const int v = ((2 * rh + 2 * ri + e) % 4 + 4) % 4;
  rnew[j] = v != 0;
  flip[j] = (uint8_t)((v != 0) ^ rh);
}

// Rows h[j] <- x/z of i[j] XOR h[j], sign rnew[j]: thread per (pair,
// column), XOR atomics on the words of the rows h (pairwise distinct, none of
// them an i row, so the bits read are never the ones written).
__global__ void k_cliff_pair_apply(uint32_t* __restrict__ tab, int n, int64_t ws, const int64_t* __restrict__ h,
                                   const int64_t* __restrict__ i, int64_t m, const uint8_t* __restrict__ flip) {
  const int64_t C = 2 * (int64_t)n + 1;
  const int64_t total = m * C;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / C, c = e - j * C;
    uint32_t* col = tab + c * ws;
    const int64_t hj = h[j];
    const uint32_t t = c < C - 1 ? bit_of(col, i[j]) : (uint32_t)flip[j];
    if (t) atomicXor(col + (hj >> 5), 1u << (hj & 31));
  }
}

// Determined mode: row h0 ^= XOR_j (row i[j] ^ row h0) on x/z, sign ^=
// XOR_j rnew[j].  Thread per column.
__global__ void k_cliff_pair_accumulate(uint32_t* __restrict__ tab, int n, int64_t ws, int64_t h0,
                                        const int64_t* __restrict__ i, int64_t m, const uint8_t* __restrict__ rnew) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t C = 2 * (int64_t)n + 1;
  if (c >= C) return;
  uint32_t* col = tab + c * ws;
  const uint32_t orig = bit_of(col, h0);
  uint32_t acc = 0;
  if (c < C - 1) {
    for (int64_t j = 0; j < m; ++j) acc ^= bit_of(col, i[j]) ^ orig;
  } else {
    for (int64_t j = 0; j < m; ++j) acc ^= rnew[j];
  }
  if (acc) col[h0 >> 5] ^= 1u << (h0 & 31);
}

// ---- numpy-format bit packing of 0/1 byte matrices ----

// out[r, b] = sum_k (in[r, 8b + k] != 0) << (7 - k)   (numpy.packbits, big bit order)
__global__ void k_packbits_rows(const uint8_t* __restrict__ in, int64_t rows, int64_t cols, uint8_t* __restrict__ out) {
  const int64_t pc = (cols + 7) >> 3;
  const int64_t total = rows * pc;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / pc, b = e - r * pc;
    uint32_t v = 0;
    for (int k = 0; k < 8; ++k) {
      const int64_t c = 8 * b + k;
      if (c < cols && in[r * cols + c]) v |= 1u << (7 - k);
    }
    out[e] = (uint8_t)v;
  }
}

// out[r, c] = bit 7 - c%8 of in[r, c/8] for c < count   (numpy.unpackbits)
__global__ void k_unpackbits_rows(const uint8_t* __restrict__ in, int64_t rows, int64_t pcols, int64_t count,
                                  uint8_t* __restrict__ out) {
  const int64_t total = rows * count;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / count, c = e - r * count;
    out[e] = (uint8_t)((in[r * pcols + (c >> 3)] >> (7 - (c & 7))) & 1u);
  }
}

__global__ void k_fill_words(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = v;
}

// ---- symbolic signs: the measured bits as affine forms over GF(2) ----
//
// For a fixed state the x/z bits met by a sequence of measurements do not
// depend on the outcomes; the signs do, affinely in the random outcomes
// u_1..u_k.  Sign column j (j = 0: the constant term, j >= 1: variable j)
// holds, for every row, the coefficient of that term in the row's sign.

// Combined 2-bit exponent counters of a sliced rowsum: cc[2w], cc[2w+1].
__global__ void k_sym_counters(int64_t W, int nchunks, const uint32_t* __restrict__ part, uint32_t* __restrict__ cc) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= W) return;
  uint32_t c0 = 0, c1 = 0;
#pragma unroll 8
  for (int y = 0; y < nchunks; ++y) {
    const uint32_t a0 = part[(y * W + w) * 2], a1 = part[(y * W + w) * 2 + 1];
    c1 ^= a1 ^ (c0 & a0);
    c0 ^= a0;
  }
  cc[2 * w] = c0;
  cc[2 * w + 1] = c1;
}

// Sign forms of the rows h (hm minus p) after h <- P_p P_h: an odd exponent
// sum makes the sign the constant 1, an even one r_h ^ r_p ^ c1.  Thread
// per (word, sign column); bit p of every column is read, never written.
__global__ void k_sym_rowsum_signs(uint32_t* __restrict__ signs, int64_t ws, int64_t W, int64_t p,
                                   const uint32_t* __restrict__ hm, const uint32_t* __restrict__ cc) {
  const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (w >= W) return;
  uint32_t h = hm[w];
  if (w == (p >> 5)) h &= ~(1u << (p & 31));
  if (!h) return;
  const int64_t j = blockIdx.y;
  uint32_t* col = signs + j * ws;
  const uint32_t c0 = cc[2 * w], c1 = cc[2 * w + 1];
  const uint32_t old = col[w];
  const uint32_t rp = bit_of(col, p) ? ~0u : 0u;
  const uint32_t even = old ^ rp ^ (j == 0 ? c1 : 0u);
  const uint32_t odd = j == 0 ? ~0u : 0u;
  col[w] = (old & ~h) | (h & ~c0 & even) | (h & c0 & odd);
}

// Sign columns of rows p - n and p after a random outcome that becomes
// variable `var`: row p - n <- row p, row p <- u_var.  Thread per column.
__global__ void k_sym_set_rows(uint32_t* __restrict__ signs, int64_t ws, int nterms, int64_t p, int64_t d, int var) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nterms) return;
  uint32_t* col = signs + (int64_t)j * ws;
  const uint32_t b = bit_of(col, p);
  col[d >> 5] = (col[d >> 5] & ~(1u << (d & 31))) | (b << (d & 31));
  const uint32_t nb = j == var ? 1u : 0u;
  col[p >> 5] = (col[p >> 5] & ~(1u << (p & 31))) | (nb << (p & 31));
}

// Form of a determined outcome: for every sign column, the parity over the
// stabilizer rows n+i with destabilizer bit x_q(i) = 1 (the XOR rule of the
// determined rowsum).  Workgroup (one wave) per sign column.
__global__ __launch_bounds__(64) void k_sym_determined(const uint32_t* __restrict__ tab, const uint32_t* __restrict__ signs,
                                                       int n, int64_t ws, int q, uint8_t* __restrict__ form) {
  const int64_t j = blockIdx.x;
  const uint32_t* col = signs + j * ws;
  const uint32_t* xq = tab + (int64_t)q * ws;
  const int64_t nw = ((int64_t)n + 31) >> 5;
  const int sh = n & 31;
  uint32_t acc = 0;
  for (int64_t i = threadIdx.x; i < nw; i += 64) {
    const uint32_t m = xq[i] & rows_below(i, n);
    if (!m) continue;
    const int64_t a = (n + i * 32) >> 5;
    uint32_t v = col[a] >> sh;
    if (sh && a + 1 < ws) v |= col[a + 1] << (32 - sh);
    acc ^= v & m;
  }
  const uint64_t odd = __ballot(__builtin_popcount(acc) & 1);
  if (threadIdx.x == 0) form[j] = (uint8_t)(__builtin_popcountll(odd) & 1);
}

// ---- many shots: samples as affine functions of the random outcomes ----

// out[s, q] = s0[q] ^ parity(u[s] & dt[q]): u holds each shot's random
// outcome bits (kw 64-bit words), dt row q the outcomes that flip measured
// qubit q.  Thread per (shot, measured qubit).
__global__ void k_gf2_samples(const uint64_t* __restrict__ u, const uint64_t* __restrict__ dt,
                              const uint8_t* __restrict__ s0, int64_t nshots, int64_t m, int64_t kw,
                              uint8_t* __restrict__ out) {
  const int64_t total = nshots * m;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = e / m, q = e - s * m;
    const uint64_t* us = u + s * kw;
    const uint64_t* dq = dt + q * kw;
    uint64_t acc = 0;
    for (int64_t w = 0; w < kw; ++w) acc ^= us[w] & dq[w];
    out[e] = (uint8_t)((s0[q] ^ __builtin_popcountll(acc)) & 1u);
  }
}

// ---- host side ----

int check_tab(const void* tab, int n, const char* what) {
  if (!tab) return set_error(QJ_EVALUE, "%s: null tableau", what);
  if (n < 1 || n > kCliffMaxQubits) return set_error(QJ_EVALUE, "%s: nqubits %d out of range", what, n);
  return QJ_OK;
}

unsigned blocks_for(int64_t items, int per) {
  return (unsigned)grid_for((u64)std::max<int64_t>(items, 1), (u64)per);
}

void device_lds(int* cus, size_t* max_lds) {
  static int ncu[64] = {0};
  static int mlds[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = -1;
  if (dev >= 0 && !ncu[dev]) {
    int a = 0, b = 0;
    if (hipDeviceGetAttribute(&a, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || a <= 0) a = 256;
    if (hipDeviceGetAttribute(&b, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || b <= 0) b = 65536;
    ncu[dev] = a;
    mlds[dev] = b;
  }
  *cus = dev >= 0 ? ncu[dev] : 256;
  *max_lds = dev >= 0 ? (size_t)mlds[dev] : 65536;
}

// Column chunks of a sliced rowsum: enough (word, chunk) threads to keep
// every CU busy, at most kMaxColumnChunks partial counters per word.
void column_chunks(int n, int64_t W, int* nchunks_out, int* kchunk_out) {
  int cus = 0;
  size_t max_lds = 0;
  device_lds(&cus, &max_lds);
  int64_t nchunks = std::max<int64_t>(1, ((int64_t)cus * 1024) / std::max<int64_t>(W, 1));
  nchunks = std::min<int64_t>(nchunks, std::min<int64_t>(n, kMaxColumnChunks));
  const int kchunk = (int)((n + nchunks - 1) / nchunks);
  *kchunk_out = kchunk;
  *nchunks_out = (int)((n + kchunk - 1) / kchunk);
}

// One-qubit rules as (m, s): m = bits m00 m01 m10 m11 of the GF(2) map of
// (x, z), s = bits a b c of the sign update a x ^ b z ^ c xz (the same
// functions as cliff_apply's one-qubit cases).
constexpr uint8_t kOneM[10] = {0x6, 0xD, 0x9, 0x9, 0x9, 0xB, 0xD, 0xB, 0x6, 0x6};
constexpr uint8_t kOneS[10] = {0x4, 0x4, 0x1, 0x2, 0x3, 0x6, 0x5, 0x4, 0x5, 0x6};

inline int bit_of_host(uint32_t v, int i) { return (v >> i) & 1; }

// Rule (m2, s2) after rule (m1, s1), as one rule.  With x1 = m00 x ^ m01 z,
// z1 = m10 x ^ m11 z: x1 z1 = m00 m10 x ^ m01 m11 z ^ (m00 m11 ^ m01 m10) xz
// for bits, so the second sign update stays of the form a x ^ b z ^ c xz.
inline void compose_one(uint8_t m1, uint8_t s1, uint8_t m2, uint8_t s2, uint8_t* m, uint8_t* sgn) {
  const int m00 = bit_of_host(m1, 0), m01 = bit_of_host(m1, 1), m10 = bit_of_host(m1, 2), m11 = bit_of_host(m1, 3);
  const int n00 = bit_of_host(m2, 0), n01 = bit_of_host(m2, 1), n10 = bit_of_host(m2, 2), n11 = bit_of_host(m2, 3);
  const int a2 = bit_of_host(s2, 0), b2 = bit_of_host(s2, 1), c2 = bit_of_host(s2, 2);
  const int ax = (a2 & m00) ^ (b2 & m10) ^ (c2 & m00 & m10);
  const int bz = (a2 & m01) ^ (b2 & m11) ^ (c2 & m01 & m11);
  const int cxz = c2 & ((m00 & m11) ^ (m01 & m10));
  *sgn = (uint8_t)(s1 ^ (ax | (bz << 1) | (cxz << 2)));
  const int p00 = (n00 & m00) ^ (n01 & m10), p01 = (n00 & m01) ^ (n01 & m11);
  const int p10 = (n10 & m00) ^ (n11 & m10), p11 = (n10 & m01) ^ (n11 & m11);
  *m = (uint8_t)(p00 | (p01 << 1) | (p10 << 2) | (p11 << 3));
}

inline bool is_two(int kind) { return kind >= kcFirstTwo && kind != kcOne; }

// Validates the list and composes every qubit's runs of one-qubit gates
// (the gates between two gates that touch the qubit otherwise) into one
// kcOne gate (kind, qubit, code), dropping identities.  A run is emitted
// just before the next two-qubit gate on its qubit, or at the end: it only
// moves past gates on other qubits.
int compose_runs(const int32_t* gates, int64_t ngates, int n, std::vector<int32_t>& out) {
  std::vector<uint8_t> pm(n, 0x9), ps(n, 0), live(n, 0);
  out.clear();
  out.reserve((size_t)ngates * 3);
  auto flush = [&](int q) {
    if (live[q] && !(pm[q] == 0x9 && ps[q] == 0)) {
      int kind = kcOne;  // a named rule when the composite is one (its code is shorter)
      for (int k = 0; k < kcFirstTwo; ++k)
        if (kOneM[k] == pm[q] && kOneS[k] == ps[q]) kind = k;
      out.push_back(kind);
      out.push_back(q);
      out.push_back(kind == kcOne ? (pm[q] | (ps[q] << 4)) : q);
    }
    pm[q] = 0x9;
    ps[q] = 0;
    live[q] = 0;
  };
  for (int64_t g = 0; g < ngates; ++g) {
    const int kind = gates[3 * g], a = gates[3 * g + 1], b = gates[3 * g + 2];
    if (kind < 0 || kind >= kcNop)
      return set_error(QJ_EVALUE, "clifford circuit: gate %lld has bad kind %d", (long long)g, kind);
    if (a < 0 || a >= n) return set_error(QJ_EVALUE, "clifford circuit: gate %lld qubit %d out of range", (long long)g, a);
    if (kind < kcFirstTwo) {
      uint8_t m, sg;
      compose_one(pm[a], ps[a], kOneM[kind], kOneS[kind], &m, &sg);
      pm[a] = m;
      ps[a] = sg;
      live[a] = 1;
      continue;
    }
    if (b < 0 || b >= n) return set_error(QJ_EVALUE, "clifford circuit: gate %lld qubit %d out of range", (long long)g, b);
    if (b == a) return set_error(QJ_EVALUE, "clifford circuit: gate %lld acts twice on qubit %d", (long long)g, a);
    flush(a);
    flush(b);
    out.push_back(kind);
    out.push_back(a);
    out.push_back(b);
  }
  for (int q = 0; q < n; ++q) flush(q);
  return QJ_OK;
}

// Groups of up to kCliffGroup gates of one kind on pairwise disjoint
// qubits, by list scheduling: a gate is ready once every earlier gate that
// shares a qubit with it has been placed in a group, ready gates act on
// disjoint qubits (of two gates on one qubit only the earlier can be ready),
// and each group takes the ready gates of the kind with the most of them.
// This only moves gates past gates on other qubits: those commute, and so do
// their XORs into r, so the tableau is the same bit for bit.  One-qubit
// gates are composed per run first (compose_runs), so they share one kind.
// Returns the descriptor words (ngroups x kCliffGroup x 2).
int schedule_gates(const int32_t* in_gates, int64_t in_ngates, int n, std::vector<uint32_t>& desc) {
  std::vector<int32_t> composed;
  int rc = compose_runs(in_gates, in_ngates, n, composed);
  if (rc) return rc;
  const int32_t* gates = composed.data();
  const int64_t ngates = (int64_t)composed.size() / 3;
  constexpr int kKinds = kcOne + 1;
  std::vector<int64_t> count(n + 1, 0);
  for (int64_t g = 0; g < ngates; ++g) {
    ++count[gates[3 * g + 1] + 1];
    if (is_two(gates[3 * g])) ++count[gates[3 * g + 2] + 1];
  }
  // per-qubit gate lists in circuit order (CSR)
  for (int q = 0; q < n; ++q) count[q + 1] += count[q];
  std::vector<int64_t> lists(count[n]);
  std::vector<int64_t> fillp(count.begin(), count.end() - 1);
  for (int64_t g = 0; g < ngates; ++g) {
    lists[fillp[gates[3 * g + 1]]++] = g;
    if (is_two(gates[3 * g])) lists[fillp[gates[3 * g + 2]]++] = g;
  }
  std::vector<int64_t> head(count.begin(), count.end() - 1);
  std::vector<uint8_t> hits(ngates, 0);
  std::vector<std::vector<int64_t>> ready(kKinds);
  std::vector<size_t> rpos(kKinds, 0);
  int64_t nready = 0;
  auto hit = [&](int64_t g) {
    const int need = is_two(gates[3 * g]) ? 2 : 1;
    if (++hits[g] == need) {
      ready[gates[3 * g]].push_back(g);
      ++nready;
    }
  };
  for (int q = 0; q < n; ++q)
    if (head[q] < count[q + 1]) hit(lists[head[q]]);
  desc.clear();
  desc.reserve((size_t)(ngates / kCliffGroup + 16) * 2 * kCliffGroup);
  int64_t placed = 0;
  while (nready > 0) {
    int best = 0;
    size_t most = 0;
    for (int k = 0; k < kKinds; ++k) {
      const size_t avail = ready[k].size() - rpos[k];
      if (avail > most) {
        most = avail;
        best = k;
      }
    }
    const int cnt = (int)std::min<size_t>(most, kCliffGroup);
    uint32_t slot[2 * kCliffGroup];
    std::vector<int64_t> taken(ready[best].begin() + rpos[best], ready[best].begin() + rpos[best] + cnt);
    rpos[best] += cnt;
    nready -= cnt;
    if (rpos[best] == ready[best].size()) {
      ready[best].clear();
      rpos[best] = 0;
    }
    for (int j = 0; j < cnt; ++j) {
      const int64_t g = taken[j];
      const int a = gates[3 * g + 1];
      // second word: the target of a two-qubit gate, the code of a composed
      // one-qubit gate
      slot[2 * j] = (uint32_t)a;
      slot[2 * j + 1] = (uint32_t)gates[3 * g + 2];
    }
    for (int j = cnt; j < kCliffGroup; ++j) {  // padding: copies of slot 0
      slot[2 * j] = slot[0];
      slot[2 * j + 1] = slot[1];
    }
    slot[0] |= (uint32_t)best << 24;
    slot[1] |= (uint32_t)cnt << 24;
    desc.insert(desc.end(), slot, slot + 2 * kCliffGroup);
    placed += cnt;
    // advance the qubits of the placed gates; their next gates get a hit
    for (int j = 0; j < cnt; ++j) {
      const int64_t g = taken[j];
      const int qs[2] = {gates[3 * g + 1], gates[3 * g + 2]};
      for (int u = 0; u < (is_two(best) ? 2 : 1); ++u) {
        const int q = qs[u];
        ++head[q];
        if (head[q] < count[q + 1]) hit(lists[head[q]]);
      }
    }
  }
  if (placed != ngates) return set_error(QJ_EVALUE, "clifford circuit: scheduling placed %lld of %lld gates",
                                         (long long)placed, (long long)ngates);
  return QJ_OK;
}

// Multi-wave form of the LDS gate list: a workgroup of NW waves shares one
// slice; the slots of a group (gates on disjoint columns) are dealt to the
// waves, slot j to wave j mod NW, and a barrier separates groups.  A wave's
// padding repeats its own first slot, so no two waves touch one column in a
// group.  The sign column is an XOR of per-slot terms: every wave keeps a
// partial, combined through LDS at the end.  Slot descriptors are read from
// the LDS ring per group.
template <int KIND, int S, int NW>
__device__ __forceinline__ uint32_t run_group_waves(uint32_t* col0, int64_t zoff, int64_t cs, uint32_t qmax, int sv,
                                                    const uint32_t* gd, int v) {
  constexpr bool two = KIND >= kcFirstTwo && KIND != kcOne;
  int64_t oa[S], ob[S];
  uint32_t code[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int jj = v + NW * (j < sv ? j : 0);
    const uint2 dd = *reinterpret_cast<const uint2*>(gd + 2 * jj);
    oa[j] = (int64_t)std::min(__builtin_amdgcn_readfirstlane(dd.x) & kSlotMask, qmax) * cs;
    code[j] = __builtin_amdgcn_readfirstlane(dd.y) & kSlotMask;
    ob[j] = (int64_t)std::min(code[j], qmax) * cs;
  }
  uint32_t a[S], b[S], c[S], d[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    a[j] = col0[oa[j]];
    b[j] = col0[zoff + oa[j]];
    if (two) {
      c[j] = col0[ob[j]];
      d[j] = col0[zoff + ob[j]];
    } else {
      c[j] = a[j];
      d[j] = b[j];
    }
  }
  uint32_t t = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    uint32_t tj;
    if constexpr (KIND == kcOne) {
      const uint32_t k = code[j];
      const uint32_t m00 = 0u - (k & 1u), m01 = 0u - ((k >> 1) & 1u), m10 = 0u - ((k >> 2) & 1u),
                     m11 = 0u - ((k >> 3) & 1u), sa = 0u - ((k >> 4) & 1u), sb = 0u - ((k >> 5) & 1u),
                     sc = 0u - ((k >> 6) & 1u);
      col0[oa[j]] = (a[j] & m00) ^ (b[j] & m01);
      col0[zoff + oa[j]] = (a[j] & m10) ^ (b[j] & m11);
      tj = (a[j] & sa) ^ (b[j] & sb) ^ (a[j] & b[j] & sc);
    } else {
      tj = cliff_apply(KIND, col0, oa[j], ob[j], zoff, a[j], b[j], c[j], d[j]);
    }
    t ^= j < sv ? tj : 0u;
  }
  return t;
}

template <int KIND, int NW>
__device__ __forceinline__ uint32_t run_waves_sized(uint32_t* col0, int64_t zoff, int64_t cs, uint32_t qmax, int sv,
                                                    const uint32_t* gd, int v) {
  if (sv <= 0) return 0u;
  if (sv == 1) return run_group_waves<KIND, 1, NW>(col0, zoff, cs, qmax, sv, gd, v);
  if (NW >= 16) return 0u;  // one slot per wave at most
  if (sv == 2 || NW >= 8) return run_group_waves<KIND, 2, NW>(col0, zoff, cs, qmax, sv, gd, v);
  return run_group_waves<KIND, kCliffGroup / NW, NW>(col0, zoff, cs, qmax, sv, gd, v);
}

template <int NW, bool LDS>
__global__ __launch_bounds__(64 * NW) void k_cliff_circuit_waves(uint32_t* __restrict__ tab, int n, int64_t ws,
                                                                 int64_t W, int T, const uint32_t* __restrict__ desc,
                                                                 int64_t ngroups) {
  extern __shared__ __align__(16) uint32_t lds_words[];
  uint32_t* ring = lds_words;
  uint32_t* slice = lds_words + 2 * kDescChunkWords;
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  const int v = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t w0 = (int64_t)blockIdx.x * T;
  const int nact = (int)std::min<int64_t>(T, W - w0);
  const int my = lane < nact ? lane : nact - 1;
  const int64_t C = 2 * (int64_t)n + 1;
  const int64_t total = C * T;
  uint32_t* col0;
  int64_t cs;
  if (LDS) {
    for (int64_t e = tid; e < total; e += 64 * NW) {
      const int64_t c = e / T;
      const int tt = (int)(e - c * T);
      slice[e] = tt < nact ? tab[c * ws + w0 + tt] : 0u;
    }
    col0 = slice + my;
    cs = T;
  } else {
    col0 = tab + w0 + my;
    cs = ws;
  }
  const int64_t total_words = ngroups * kDescWords;
  const int64_t nchunks = (ngroups + kDescChunk - 1) / kDescChunk;
  for (int64_t e = tid; e < kDescChunkWords; e += 64 * NW) ring[e] = e < total_words ? desc[e] : 0u;
  __syncthreads();
  const int64_t zoff = (int64_t)n * cs;
  const int64_t roff = 2 * (int64_t)n * cs;
  const uint32_t qmax = (uint32_t)(n - 1);
  uint32_t t_acc = 0;
  constexpr int kPer = kDescChunkWords / (64 * NW);
  for (int64_t ch = 0; ch < nchunks; ++ch) {
    const uint32_t* cur = ring + (ch & 1) * kDescChunkWords;
    uint32_t nx[kPer];
    const bool more = ch + 1 < nchunks;
    if (more) {
#pragma unroll
      for (int k = 0; k < kPer; ++k) {
        const int64_t e = (ch + 1) * kDescChunkWords + (int64_t)k * 64 * NW + tid;
        nx[k] = e < total_words ? desc[e] : 0u;
      }
    }
    const int ng = (int)std::min<int64_t>(kDescChunk, ngroups - ch * kDescChunk);
    for (int g = 0; g < ng; ++g) {
      const uint32_t* gd = cur + g * kDescWords;
      const uint2 hd = *reinterpret_cast<const uint2*>(gd);
      const int kind = (int)(__builtin_amdgcn_readfirstlane(hd.x) >> 24);
      const int cnt = (int)(__builtin_amdgcn_readfirstlane(hd.y) >> 24);
      const int sv = cnt > v ? (cnt - v + NW - 1) / NW : 0;
      uint32_t t;
      switch (kind) {
        case kcH: t = run_waves_sized<kcH, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcS: t = run_waves_sized<kcS, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcZ: t = run_waves_sized<kcZ, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcX: t = run_waves_sized<kcX, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcY: t = run_waves_sized<kcY, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcSX: t = run_waves_sized<kcSX, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcSDG: t = run_waves_sized<kcSDG, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcSXDG: t = run_waves_sized<kcSXDG, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcRYpi: t = run_waves_sized<kcRYpi, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcRY3pi2: t = run_waves_sized<kcRY3pi2, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcCNOT: t = run_waves_sized<kcCNOT, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcCZ: t = run_waves_sized<kcCZ, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcSWAP: t = run_waves_sized<kcSWAP, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kciSWAP: t = run_waves_sized<kciSWAP, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        case kcOne: t = run_waves_sized<kcOne, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
        default: t = run_waves_sized<kcCY, NW>(col0, zoff, cs, qmax, sv, gd, v); break;
      }
      t_acc ^= t;
      __syncthreads();
    }
    if (more) {
      uint32_t* nxt = ring + ((ch + 1) & 1) * kDescChunkWords;
#pragma unroll
      for (int k = 0; k < kPer; ++k) nxt[k * 64 * NW + tid] = nx[k];
      __syncthreads();
    }
  }
  // partial signs through the ring (every descriptor read is done)
  ring[tid] = t_acc;
  __syncthreads();
  if (v == 0) {
    uint32_t r = col0[roff];
#pragma unroll
    for (int u = 0; u < NW; ++u) r ^= ring[u * 64 + lane];
    col0[roff] = r;
  }
  if (LDS) {
    __syncthreads();
    for (int64_t e = tid; e < total; e += 64 * NW) {
      const int64_t c = e / T;
      const int tt = (int)(e - c * T);
      if (tt < nact) tab[c * ws + w0 + tt] = slice[e];
    }
  }
}

// Launches k_cliff_circuit on a descriptor array in device memory: LDS
// slices when every workgroup is resident at once; otherwise the global
// variant (its columns stay in L2), which measures as fast as the LDS one
// from 2000 qubits up and needs no second round of workgroups.
void launch_circuit(uint32_t* tab, int nqubits, const uint32_t* ddesc, int64_t ngroups, hipStream_t s) {
  const int64_t W = cliff_words(nqubits), ws = cliff_stride(nqubits);
  const int64_t C = 2 * (int64_t)nqubits + 1;
  int cus = 0;
  size_t max_lds = 0;
  device_lds(&cus, &max_lds);
  const size_t cap = std::min(kCliffLdsPerCu, std::max<size_t>(max_lds, 65536)) - kDescRingBytes;
  int64_t T = (int64_t)(cap / (4 * (size_t)C));
  const char* kern = std::getenv("QJ_CLIFFORD_KERNEL");
  const bool force_global = kern && std::strcmp(kern, "global") == 0;
  bool use_lds = T >= 1 && !force_global;
  int64_t nwg = 0;
  if (use_lds) {
    T = std::min<int64_t>({T, 64, W});
    nwg = (W + T - 1) / T;
    T = (W + nwg - 1) / nwg;  // even split
    const size_t bytes = kDescRingBytes + (size_t)T * C * 4;
    const int64_t per_cu = std::max<int64_t>(1, (int64_t)(kCliffLdsPerCu / bytes));
    use_lds = nwg <= (int64_t)cus * per_cu;
  }
  const char* waves = std::getenv("QJ_CLIFFORD_WAVES");
  // waves per workgroup of the LDS variant (8 measured fastest: 56 M gates/s
  // at 1000 qubits against 42 M with 4 and 20 M with 1)
  int nwaves = 8;
  if (waves && std::strcmp(waves, "1") == 0) nwaves = 1;
  if (waves && std::strcmp(waves, "4") == 0) nwaves = 4;
  if (waves && std::strcmp(waves, "16") == 0) nwaves = 16;
  if (use_lds && nwaves == 4) {
    const size_t lds = kDescRingBytes + (size_t)T * C * 4;
    if (lds > 65536)
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_cliff_circuit_waves<4, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_cliff_circuit_waves<4, true><<<dim3((unsigned)nwg), dim3(256), lds, s>>>(tab, nqubits, ws, W, (int)T, ddesc,
                                                                         ngroups);
  } else if (use_lds && nwaves == 16) {
    const size_t lds = kDescRingBytes + (size_t)T * C * 4;
    if (lds > 65536)
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_cliff_circuit_waves<16, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_cliff_circuit_waves<16, true><<<dim3((unsigned)nwg), dim3(1024), lds, s>>>(tab, nqubits, ws, W, (int)T, ddesc,
                                                                             ngroups);
  } else if (use_lds && nwaves == 8) {
    const size_t lds = kDescRingBytes + (size_t)T * C * 4;
    if (lds > 65536)
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_cliff_circuit_waves<8, true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_cliff_circuit_waves<8, true><<<dim3((unsigned)nwg), dim3(512), lds, s>>>(tab, nqubits, ws, W, (int)T, ddesc,
                                                                         ngroups);
  } else if (use_lds) {
    const size_t lds = kDescRingBytes + (size_t)T * C * 4;
    if (lds > 65536)
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_cliff_circuit<true>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    k_cliff_circuit<true><<<dim3((unsigned)nwg), dim3(64), lds, s>>>(tab, nqubits, ws, W, (int)T, ddesc, ngroups);
  } else {
    const int64_t gwg = (W + 63) / 64;
    if (nwaves == 16)
      k_cliff_circuit_waves<16, false><<<dim3((unsigned)gwg), dim3(1024), kDescRingBytes, s>>>(tab, nqubits, ws, W,
                                                                                             64, ddesc, ngroups);
    else if (nwaves == 8)
      k_cliff_circuit_waves<8, false><<<dim3((unsigned)gwg), dim3(512), kDescRingBytes, s>>>(tab, nqubits, ws, W, 64,
                                                                                           ddesc, ngroups);
    else if (nwaves == 4)
      k_cliff_circuit_waves<4, false><<<dim3((unsigned)gwg), dim3(256), kDescRingBytes, s>>>(tab, nqubits, ws, W, 64,
                                                                                           ddesc, ngroups);
    else
      k_cliff_circuit<false><<<dim3((unsigned)gwg), dim3(64), kDescRingBytes, s>>>(tab, nqubits, ws, W, 64, ddesc,
                                                                                  ngroups);
  }
}

}  // namespace
}  // namespace qj

using namespace qj;

extern "C" {

int64_t qj_clifford_column_words(int nqubits) {
  if (nqubits < 1 || nqubits > kCliffMaxQubits) return -1;
  return cliff_stride(nqubits);
}

int qj_clifford_pack(const uint8_t* dense, int nqubits, uint32_t* tab, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_pack");
  if (rc) return rc;
  if (!dense) return set_error(QJ_EVALUE, "clifford_pack: null matrix");
  const int64_t C = 2 * (int64_t)nqubits + 1;
  k_cliff_pack<<<dim3(blocks_for(C * cliff_stride(nqubits), kBlock)), dim3(kBlock), 0, (hipStream_t)stream>>>(
      dense, C, cliff_words(nqubits), cliff_stride(nqubits), tab);
  return check_launch("clifford_pack launch");
}

int qj_clifford_unpack(const uint32_t* tab, int nqubits, uint8_t* dense, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_unpack");
  if (rc) return rc;
  if (!dense) return set_error(QJ_EVALUE, "clifford_unpack: null matrix");
  const int64_t C = 2 * (int64_t)nqubits + 1;
  k_cliff_unpack<<<dim3(blocks_for(C * C, kBlock)), dim3(kBlock), 0, (hipStream_t)stream>>>(
      tab, C, cliff_stride(nqubits), dense);
  return check_launch("clifford_unpack launch");
}

int qj_clifford_apply_gate(uint32_t* tab, int nqubits, int kind, int q0, int q1, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_apply_gate");
  if (rc) return rc;
  if (kind < 0 || kind >= kcNop) return set_error(QJ_EVALUE, "clifford_apply_gate: bad kind %d", kind);
  if (q0 < 0 || q0 >= nqubits) return set_error(QJ_EVALUE, "clifford_apply_gate: qubit %d out of range", q0);
  if (kind >= kcFirstTwo) {
    if (q1 < 0 || q1 >= nqubits) return set_error(QJ_EVALUE, "clifford_apply_gate: qubit %d out of range", q1);
    if (q1 == q0) return set_error(QJ_EVALUE, "clifford_apply_gate: control and target are both %d", q0);
  }
  const int64_t W = cliff_words(nqubits);
  k_cliff_gate<<<dim3(blocks_for(W, kBlock)), dim3(kBlock), 0, (hipStream_t)stream>>>(
      tab, nqubits, cliff_stride(nqubits), W, kind, q0, q1);
  return check_launch("clifford_apply_gate launch");
}

int qj_clifford_schedule(const int32_t* gates, int64_t ngates, int nqubits, uint32_t* desc, int64_t capacity,
                         int64_t* ngroups) {
  if (nqubits < 1 || nqubits > kCliffMaxQubits)
    return set_error(QJ_EVALUE, "clifford_schedule: nqubits %d out of range", nqubits);
  if (!ngroups) return set_error(QJ_EVALUE, "clifford_schedule: null group count");
  if (ngates < 0) return set_error(QJ_EVALUE, "clifford_schedule: negative gate count");
  if (ngates > 0 && !gates) return set_error(QJ_EVALUE, "clifford_schedule: null gate list");
  std::vector<uint32_t> words;
  if (ngates > 0) {
    const int rc = schedule_gates(gates, ngates, nqubits, words);
    if (rc) return rc;
  }
  *ngroups = (int64_t)words.size() / kDescWords;
  if (!desc) return QJ_OK;  // size query
  if (capacity < *ngroups)
    return set_error(QJ_EVALUE, "clifford_schedule: %lld groups do not fit a capacity of %lld", (long long)*ngroups,
                     (long long)capacity);
  std::memcpy(desc, words.data(), words.size() * sizeof(uint32_t));
  return QJ_OK;
}

int qj_clifford_apply_schedule(uint32_t* tab, int nqubits, const uint32_t* desc, int64_t ngroups, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_apply_schedule");
  if (rc) return rc;
  if (ngroups < 0) return set_error(QJ_EVALUE, "clifford_apply_schedule: negative group count");
  if (ngroups == 0) return QJ_OK;
  if (!desc) return set_error(QJ_EVALUE, "clifford_apply_schedule: null schedule");
  launch_circuit(tab, nqubits, desc, ngroups, (hipStream_t)stream);
  return check_launch("clifford_apply_schedule launch");
}

int qj_clifford_apply_circuit(uint32_t* tab, int nqubits, const int32_t* gates, int64_t ngates, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_apply_circuit");
  if (rc) return rc;
  if (ngates < 0) return set_error(QJ_EVALUE, "clifford_apply_circuit: negative gate count");
  if (ngates == 0) return QJ_OK;
  if (!gates) return set_error(QJ_EVALUE, "clifford_apply_circuit: null gate list");
  std::vector<uint32_t> desc;
  rc = schedule_gates(gates, ngates, nqubits, desc);
  if (rc) return rc;
  const int64_t ngroups = (int64_t)desc.size() / kDescWords;
  hipStream_t s = (hipStream_t)stream;
  void* ddesc = nullptr;
  int ticket = -1;
  rc = ring_upload(desc.data(), desc.size() * sizeof(uint32_t), s, &ddesc, &ticket);
  if (rc) return rc;
  launch_circuit(tab, nqubits, static_cast<const uint32_t*>(ddesc), ngroups, s);
  rc = check_launch("clifford_apply_circuit launch");
  ring_release(ticket, s);
  return rc;
}

int qj_clifford_first_stabilizer(const uint32_t* tab, int nqubits, int q, int32_t* p_out, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_first_stabilizer");
  if (rc) return rc;
  if (!p_out) return set_error(QJ_EVALUE, "clifford_first_stabilizer: null output");
  if (q < 0 || q >= nqubits) return set_error(QJ_EVALUE, "clifford_first_stabilizer: qubit %d out of range", q);
  k_cliff_first_stab<<<dim3(1), dim3(256), 0, (hipStream_t)stream>>>(tab, nqubits, cliff_stride(nqubits), q, p_out);
  return check_launch("clifford_first_stabilizer launch");
}

int qj_clifford_random_outcome(uint32_t* tab, int nqubits, int p, int q, int outcome, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_random_outcome");
  if (rc) return rc;
  if (q < 0 || q >= nqubits) return set_error(QJ_EVALUE, "clifford_random_outcome: qubit %d out of range", q);
  if (p < 0 || p >= nqubits) return set_error(QJ_EVALUE, "clifford_random_outcome: stabilizer %d out of range", p);
  if (outcome != 0 && outcome != 1) return set_error(QJ_EVALUE, "clifford_random_outcome: outcome must be 0 or 1");
  hipStream_t s = (hipStream_t)stream;
  const int64_t W = cliff_words(nqubits), ws = cliff_stride(nqubits);
  const int64_t prow = (int64_t)nqubits + p;
  // column chunks: enough (word, chunk) threads to keep every CU busy
  int nchunks = 0, kchunk = 0;
  column_chunks(nqubits, W, &nchunks, &kchunk);
  DeviceScratch* sc = nullptr;
  rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  rc = scratch_reserve_big(sc, (size_t)(W + 2 * nchunks * W) * sizeof(uint32_t), s);
  if (rc) return rc;
  uint32_t* hm = static_cast<uint32_t*>(sc->big);
  uint32_t* part = hm + W;
  k_cliff_hmask<<<dim3(blocks_for(W, kBlock)), dim3(kBlock), 0, s>>>(tab, ws, W, nqubits, q, hm, nullptr);
  k_cliff_rowsum_sliced<<<dim3(blocks_for(W, kBlock), (unsigned)nchunks), dim3(kBlock), 0, s>>>(
      tab, nqubits, ws, W, prow, hm, kchunk, part, nullptr);
  k_cliff_rowsum_sign<<<dim3(blocks_for(W, kBlock)), dim3(kBlock), 0, s>>>(tab, nqubits, ws, W, prow, hm,
                                                                           (int)nchunks, part, nullptr);
  const int64_t C = 2 * (int64_t)nqubits + 1;
  k_cliff_set_rows<<<dim3(blocks_for(C, kBlock)), dim3(kBlock), 0, s>>>(tab, nqubits, ws, prow, q, outcome, nullptr,
                                                                        nullptr);
  return check_launch("clifford_random_outcome launch");
}

int qj_clifford_determined_outcome(uint32_t* tab, int nqubits, int q, int32_t* outcome_out, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_determined_outcome");
  if (rc) return rc;
  if (!outcome_out) return set_error(QJ_EVALUE, "clifford_determined_outcome: null output");
  if (q < 0 || q >= nqubits) return set_error(QJ_EVALUE, "clifford_determined_outcome: qubit %d out of range", q);
  const int64_t C = 2 * (int64_t)nqubits + 1;
  k_cliff_determined<<<dim3((unsigned)C), dim3(64), 0, (hipStream_t)stream>>>(tab, nqubits, cliff_stride(nqubits), q,
                                                                              outcome_out, nullptr);
  return check_launch("clifford_determined_outcome launch");
}

// Phase terms of a determined outcome into a zeroed device word, then the
// finish on the scratch row (form == nullptr) or a symbolic form.
static int exact_phase(uint32_t* tab, int nqubits, int q, int32_t* out, uint8_t* form, hipStream_t s,
                       const char* what) {
  DeviceScratch* sc = nullptr;
  int rc = scratch_for_current_device(&sc);
  if (!rc) rc = scratch_reserve_big(sc, sizeof(uint32_t), s);
  if (rc) return rc;
  uint32_t* acc = static_cast<uint32_t*>(sc->big);
  const int64_t ws = cliff_stride(nqubits);
  k_fill_words<<<dim3(1), dim3(64), 0, s>>>(acc, 1, 0u);
  k_cliff_exact_terms<<<dim3((unsigned)nqubits), dim3(64), 0, s>>>(tab, nqubits, ws, q, acc, nullptr);
  k_cliff_exact_finish<<<dim3(1), dim3(64), 0, s>>>(tab, nqubits, ws, acc, out, form, nullptr);
  return check_launch(what);
}

int qj_clifford_determined_outcome_exact(uint32_t* tab, int nqubits, int q, int32_t* outcome_out, void* stream) {
  int rc = qj_clifford_determined_outcome(tab, nqubits, q, outcome_out, stream);
  if (rc) return rc;
  return exact_phase(tab, nqubits, q, outcome_out, nullptr, (hipStream_t)stream,
                     "clifford_determined_outcome_exact launch");
}

int qj_clifford_measure(uint32_t* tab, int nqubits, const int32_t* qubits, int m, const uint8_t* draws, int exact,
                        int32_t* sample_out, int32_t* nrandom_out, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_measure");
  if (rc) return rc;
  if (m < 0) return set_error(QJ_EVALUE, "clifford_measure: negative qubit count");
  if (m == 0) return QJ_OK;
  if (!qubits || !draws || !sample_out || !nrandom_out) return set_error(QJ_EVALUE, "clifford_measure: null pointer");
  for (int i = 0; i < m; ++i) {
    if (qubits[i] < 0 || qubits[i] >= nqubits)
      return set_error(QJ_EVALUE, "clifford_measure: qubit %d out of range at %d", qubits[i], i);
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t W = cliff_words(nqubits), ws = cliff_stride(nqubits);
  const int64_t C = 2 * (int64_t)nqubits + 1;
  const char* mode = std::getenv("QJ_CLIFFORD_MEASURE");
  const bool force_kernels = mode && std::strcmp(mode, "kernels") == 0;
  const bool force_small = mode && std::strcmp(mode, "small") == 0;
  if (!force_kernels && (nqubits <= kMeasureSmallMax || (force_small && nqubits <= kMeasureSmallCap))) {
    void* dq = nullptr;
    int ticket = -1;
    rc = ring_upload(qubits, (size_t)m * sizeof(int32_t), s, &dq, &ticket);
    if (rc) return rc;
    k_cliff_measure_small<<<dim3(1), dim3(kMeasureThreads), 0, s>>>(
        tab, nqubits, ws, static_cast<const int32_t*>(dq), m, draws, exact, sample_out, nrandom_out);
    rc = check_launch("clifford_measure launch");
    ring_release(ticket, s);
    return rc;
  }
  int nchunks = 0, kchunk = 0;
  column_chunks(nqubits, W, &nchunks, &kchunk);
  DeviceScratch* sc = nullptr;
  rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  rc = scratch_reserve_big(sc, (size_t)(W + 2 * nchunks * W + 8) * sizeof(uint32_t), s);
  if (rc) return rc;
  uint32_t* hm = static_cast<uint32_t*>(sc->big);
  uint32_t* part = hm + W;
  int32_t* ctl = reinterpret_cast<int32_t*>(part + 2 * nchunks * W);
  uint32_t* acc = reinterpret_cast<uint32_t*>(ctl + 4);
  k_fill_words<<<dim3(1), dim3(64), 0, s>>>(reinterpret_cast<uint32_t*>(ctl), 4, 0u);
  for (int i = 0; i < m; ++i) {
    const int q = qubits[i];
    k_cliff_first_stab<<<dim3(1), dim3(256), 0, s>>>(tab, nqubits, ws, q, ctl);
    k_cliff_hmask<<<dim3(blocks_for(W, kBlock)), dim3(kBlock), 0, s>>>(tab, ws, W, nqubits, q, hm, ctl);
    k_cliff_rowsum_sliced<<<dim3(blocks_for(W, kBlock), (unsigned)nchunks), dim3(kBlock), 0, s>>>(
        tab, nqubits, ws, W, 0, hm, kchunk, part, ctl);
    k_cliff_rowsum_sign<<<dim3(blocks_for(W, kBlock)), dim3(kBlock), 0, s>>>(tab, nqubits, ws, W, 0, hm,
                                                                             (int)nchunks, part, ctl);
    k_cliff_set_rows<<<dim3(blocks_for(C, kBlock)), dim3(kBlock), 0, s>>>(tab, nqubits, ws, 0, q, 0, ctl, draws);
    k_cliff_determined<<<dim3((unsigned)C), dim3(64), 0, s>>>(tab, nqubits, ws, q, ctl + 2, ctl);
    if (exact) {
      k_fill_words<<<dim3(1), dim3(64), 0, s>>>(acc, 1, 0u);
      k_cliff_exact_terms<<<dim3((unsigned)nqubits), dim3(64), 0, s>>>(tab, nqubits, ws, q, acc, ctl);
      k_cliff_exact_finish<<<dim3(1), dim3(64), 0, s>>>(tab, nqubits, ws, acc, ctl + 2, nullptr, ctl);
    }
    k_cliff_meas_finish<<<dim3(1), dim3(64), 0, s>>>(ctl, draws, sample_out, i, nrandom_out);
  }
  return check_launch("clifford_measure launch");
}

int qj_clifford_rowsum(uint32_t* tab, int nqubits, const int64_t* h, const int64_t* i, int64_t nrows, int determined,
                       void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_rowsum");
  if (rc) return rc;
  if (nrows < 0) return set_error(QJ_EVALUE, "clifford_rowsum: negative row count");
  if (nrows == 0) return QJ_OK;
  if (!h || !i) return set_error(QJ_EVALUE, "clifford_rowsum: null row list");
  const int64_t C = 2 * (int64_t)nqubits + 1;
  for (int64_t j = 0; j < nrows; ++j) {
    if (h[j] < 0 || h[j] >= C || i[j] < 0 || i[j] >= C)
      return set_error(QJ_EVALUE, "clifford_rowsum: row index out of range at %lld", (long long)j);
  }
  if (determined) {
    for (int64_t j = 0; j < nrows; ++j) {
      if (h[j] != h[0]) return set_error(QJ_EVALUE, "clifford_rowsum: determined mode needs a single target row");
      if (i[j] == h[0]) return set_error(QJ_EVALUE, "clifford_rowsum: source row equals the target row");
    }
  } else {
    std::vector<uint8_t> seen(C, 0);
    for (int64_t j = 0; j < nrows; ++j) {
      if (seen[h[j]] & 1) return set_error(QJ_EVALUE, "clifford_rowsum: target row %lld repeated", (long long)h[j]);
      seen[h[j]] |= 1;
    }
    for (int64_t j = 0; j < nrows; ++j) {
      if (seen[i[j]] & 1) return set_error(QJ_EVALUE, "clifford_rowsum: row %lld is both a source and a target",
                                           (long long)i[j]);
    }
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<int64_t> hi(2 * nrows);
  std::memcpy(hi.data(), h, nrows * sizeof(int64_t));
  std::memcpy(hi.data() + nrows, i, nrows * sizeof(int64_t));
  void* dhi = nullptr;
  int ticket = -1;
  rc = ring_upload(hi.data(), hi.size() * sizeof(int64_t), s, &dhi, &ticket);
  if (rc) return rc;
  DeviceScratch* sc = nullptr;
  rc = scratch_for_current_device(&sc);
  if (!rc) rc = scratch_reserve_big(sc, (size_t)2 * nrows, s);
  if (rc) {
    ring_release(ticket, s);
    return rc;
  }
  const int64_t* dh = static_cast<const int64_t*>(dhi);
  const int64_t* di = dh + nrows;
  uint8_t* rnew = static_cast<uint8_t*>(sc->big);
  uint8_t* flip = rnew + nrows;
  const int64_t ws = cliff_stride(nqubits);
  k_cliff_pair_sign<<<dim3(blocks_for(nrows, kBlock)), dim3(kBlock), 0, s>>>(tab, nqubits, ws, dh, di, nrows, rnew,
                                                                             flip);
  if (determined)
    k_cliff_pair_accumulate<<<dim3(blocks_for(C, kBlock)), dim3(kBlock), 0, s>>>(tab, nqubits, ws, h[0], di, nrows,
                                                                                 rnew);
  else
    k_cliff_pair_apply<<<dim3(blocks_for(nrows * C, kBlock)), dim3(kBlock), 0, s>>>(tab, nqubits, ws, dh, di, nrows,
                                                                                    flip);
  rc = check_launch("clifford_rowsum launch");
  ring_release(ticket, s);
  return rc;
}

int qj_clifford_sym_init(const uint32_t* tab, int nqubits, uint32_t* signs, int nvars, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_sym_init");
  if (rc) return rc;
  if (!signs) return set_error(QJ_EVALUE, "clifford_sym_init: null sign columns");
  if (nvars < 0) return set_error(QJ_EVALUE, "clifford_sym_init: negative variable count");
  hipStream_t s = (hipStream_t)stream;
  const int64_t ws = cliff_stride(nqubits);
  rc = hip_status(hipMemcpyAsync(signs, tab + 2 * (int64_t)nqubits * ws, ws * sizeof(uint32_t),
                                 hipMemcpyDeviceToDevice, s), "clifford_sym_init copy");
  if (rc) return rc;
  if (nvars > 0) {
    k_fill_words<<<dim3(blocks_for(nvars * ws, kBlock)), dim3(kBlock), 0, s>>>(signs + ws, (int64_t)nvars * ws, 0u);
    return check_launch("clifford_sym_init launch");
  }
  return QJ_OK;
}

int qj_clifford_sym_random(uint32_t* tab, int nqubits, uint32_t* signs, int nvars, int p, int q, int var,
                           void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_sym_random");
  if (rc) return rc;
  if (!signs) return set_error(QJ_EVALUE, "clifford_sym_random: null sign columns");
  if (q < 0 || q >= nqubits) return set_error(QJ_EVALUE, "clifford_sym_random: qubit %d out of range", q);
  if (p < 0 || p >= nqubits) return set_error(QJ_EVALUE, "clifford_sym_random: stabilizer %d out of range", p);
  if (var < 1 || var > nvars) return set_error(QJ_EVALUE, "clifford_sym_random: variable %d out of range", var);
  if (nvars >= 65535) return set_error(QJ_EVALUE, "clifford_sym_random: at most 65534 variables");
  hipStream_t s = (hipStream_t)stream;
  const int64_t W = cliff_words(nqubits), ws = cliff_stride(nqubits);
  const int64_t prow = (int64_t)nqubits + p;
  int nchunks = 0, kchunk = 0;
  column_chunks(nqubits, W, &nchunks, &kchunk);
  DeviceScratch* sc = nullptr;
  rc = scratch_for_current_device(&sc);
  if (rc) return rc;
  rc = scratch_reserve_big(sc, (size_t)(3 * W + 2 * (int64_t)nchunks * W) * sizeof(uint32_t), s);
  if (rc) return rc;
  uint32_t* hm = static_cast<uint32_t*>(sc->big);
  uint32_t* cc = hm + W;
  uint32_t* part = cc + 2 * W;
  k_cliff_hmask<<<dim3(blocks_for(W, kBlock)), dim3(kBlock), 0, s>>>(tab, ws, W, nqubits, q, hm, nullptr);
  k_cliff_rowsum_sliced<<<dim3(blocks_for(W, kBlock), (unsigned)nchunks), dim3(kBlock), 0, s>>>(
      tab, nqubits, ws, W, prow, hm, kchunk, part, nullptr);
  k_sym_counters<<<dim3(blocks_for(W, kBlock)), dim3(kBlock), 0, s>>>(W, nchunks, part, cc);
  k_sym_rowsum_signs<<<dim3(blocks_for(W, kBlock), (unsigned)(nvars + 1)), dim3(kBlock), 0, s>>>(signs, ws, W, prow,
                                                                                                  hm, cc);
  const int64_t C = 2 * (int64_t)nqubits + 1;
  k_cliff_set_rows<<<dim3(blocks_for(C, kBlock)), dim3(kBlock), 0, s>>>(tab, nqubits, ws, prow, q, 0, nullptr,
                                                                        nullptr);
  k_sym_set_rows<<<dim3(blocks_for(nvars + 1, kBlock)), dim3(kBlock), 0, s>>>(signs, ws, nvars + 1, prow,
                                                                             prow - nqubits, var);
  return check_launch("clifford_sym_random launch");
}

int qj_clifford_sym_determined(const uint32_t* tab, int nqubits, const uint32_t* signs, int nvars, int q,
                               uint8_t* form, void* stream) {
  int rc = check_tab(tab, nqubits, "clifford_sym_determined");
  if (rc) return rc;
  if (!signs || !form) return set_error(QJ_EVALUE, "clifford_sym_determined: null pointer");
  if (q < 0 || q >= nqubits) return set_error(QJ_EVALUE, "clifford_sym_determined: qubit %d out of range", q);
  if (nvars < 0) return set_error(QJ_EVALUE, "clifford_sym_determined: negative variable count");
  k_sym_determined<<<dim3((unsigned)(nvars + 1)), dim3(64), 0, (hipStream_t)stream>>>(tab, signs, nqubits,
                                                                                    cliff_stride(nqubits), q, form);
  return check_launch("clifford_sym_determined launch");
}

int qj_clifford_sym_determined_exact(const uint32_t* tab, int nqubits, const uint32_t* signs, int nvars, int q,
                                     uint8_t* form, void* stream) {
  int rc = qj_clifford_sym_determined(tab, nqubits, signs, nvars, q, form, stream);
  if (rc) return rc;
  return exact_phase(const_cast<uint32_t*>(tab), nqubits, q, nullptr, form, (hipStream_t)stream,
                     "clifford_sym_determined_exact launch");
}

int qj_gf2_samples(const uint64_t* u, const uint64_t* dt, const uint8_t* s0, int64_t nshots, int64_t m, int64_t kw,
                   uint8_t* out, void* stream) {
  if (nshots < 0 || m < 0 || kw < 0) return set_error(QJ_EVALUE, "gf2_samples: negative size");
  if (nshots == 0 || m == 0) return QJ_OK;
  if (!s0 || !out || (kw > 0 && (!u || !dt))) return set_error(QJ_EVALUE, "gf2_samples: null pointer");
  k_gf2_samples<<<dim3(blocks_for(nshots * m, kBlock)), dim3(kBlock), 0, (hipStream_t)stream>>>(u, dt, s0, nshots, m,
                                                                                              kw, out);
  return check_launch("gf2_samples launch");
}

int qj_packbits_rows(const uint8_t* in, int64_t rows, int64_t cols, uint8_t* out, void* stream) {
  if (rows < 0 || cols < 0) return set_error(QJ_EVALUE, "packbits_rows: negative size");
  if (rows == 0 || cols == 0) return QJ_OK;
  if (!in || !out) return set_error(QJ_EVALUE, "packbits_rows: null pointer");
  k_packbits_rows<<<dim3(blocks_for(rows * ((cols + 7) >> 3), kBlock)), dim3(kBlock), 0, (hipStream_t)stream>>>(
      in, rows, cols, out);
  return check_launch("packbits_rows launch");
}

int qj_unpackbits_rows(const uint8_t* in, int64_t rows, int64_t packed_cols, int64_t count, uint8_t* out,
                       void* stream) {
  if (rows < 0 || packed_cols < 0 || count < 0) return set_error(QJ_EVALUE, "unpackbits_rows: negative size");
  if (count > 8 * packed_cols) return set_error(QJ_EVALUE, "unpackbits_rows: count exceeds the packed bits");
  if (rows == 0 || count == 0) return QJ_OK;
  if (!in || !out) return set_error(QJ_EVALUE, "unpackbits_rows: null pointer");
  k_unpackbits_rows<<<dim3(blocks_for(rows * count, kBlock)), dim3(kBlock), 0, (hipStream_t)stream>>>(
      in, rows, packed_cols, count, out);
  return check_launch("unpackbits_rows launch");
}

}  // extern "C"
