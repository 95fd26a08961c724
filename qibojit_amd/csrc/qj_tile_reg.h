// Register-resident tile passes: the host planner and the per-thread steps
// of the kernel (host-callable too, so the planner can be checked on a CPU).
//
// A pass applies a list of tile ops to tiles of 2^TB amplitudes.  Each of
// the 2^NT threads of a workgroup holds 2^NRB amplitudes of its tile in
// registers (TB = NT + NRB).  A layout says which tile bit indexes what:
// slot s < NT is bit s of the thread index, slot NT + j is bit j of the
// register index, and slot[s] names the tile bit.  The op list is cut into
// segments; in a segment every dense op has all its targets on register
// bits, so it runs as a register mat-vec with compile-time register
// indices.  Between segments the tile is re-laid out through LDS (one write
// and one read of the tile).  Phase ops run in any layout.  A SWAP of two
// tile bits moves no data: it relabels the bits of every later op and of the
// store (bit b of the tile after the swaps lives in physical bit phys[b]).
//
// Segment 0's layout has the six lowest tile bits on the 64 lanes of a wave
// (coalesced 512-byte loads); the last segment's layout has the physical
// bits of the six lowest logical bits there (coalesced stores).
//
// Plan blob: RegHead, then the part a workgroup copies into LDS once
// (segments, ops, dense matrices with the register-order permutation
// applied, per-op phase factors W in slot order, and the global offsets of
// the phase ops' out-of-tile tables), then the out-of-tile tables OA/OB,
// read from global memory once per tile and phase op.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "qj_common.h"

namespace qj {

constexpr int kRegMaxSeg = 64;
constexpr int kRegMaxBits = 16;
constexpr int kRegLanes = 6;  // tile bits on the lanes of one wave

struct RegHead {
  int32_t TB, NT, NRB, nseg, nops, nphase, ga, gb;
  u64 lds_off, lds_bytes;      // LDS part: blob[lds_off, lds_off + lds_bytes)
  u64 segs_off, ops_off, ftab_off, ldata_off;  // offsets inside the LDS part
  // Optional per-segment register offsets (2^NRB uint16 per segment, 16-byte
  // aligned): the LDS byte offset of register r's part of the tile element
  // in segment s's layout (the swizzle of the register bits alone, times the
  // amplitude size).  A re-layout address is then the thread part's offset
  // XOR this uniform value, instead of a per-lane swizzle.  Left out
  // (ub_off = 0) when the plan would not fit the LDS budget with it.
  u64 ub_off;
  // Hadamard-form ops run as butterflies (a + b, a - b); the product of
  // their factors h is applied once to every amplitude at the store.
  double had_scale;              // 1 when nhad = 0
  int32_t nhad;
  int32_t segtab;                // set by the launcher: per-thread table of segment layouts in LDS
  uint8_t gbit_ld[kRegMaxBits];  // physical tile bit -> state bit (load)
  uint8_t gbit_st[kRegMaxBits];  // physical tile bit -> state bit (store)
};

struct RegSeg {
  int32_t o0, o1;                // ops [o0, o1) run in this layout
  int8_t slot[kRegMaxBits];      // slot -> physical tile bit
};

constexpr int kRegDense = 0;      // complex matrix
constexpr int kRegPhase = 1;
constexpr int kRegDenseReal = 2;  // 2x2 matrix with zero imaginary parts
constexpr int kRegHad = 3;        // h * [[1, 1], [1, -1]], no controls: factor h deferred
// Deferred factors per plan: the unscaled butterflies grow amplitudes by
// at most sqrt(2) each, 2^16 for 32 of them, far from overflow in FP32.
constexpr int kRegMaxHad = 32;

struct RegOp {
  int32_t kind;     // kRegDense, kRegPhase, kRegDenseReal or kRegHad
  int32_t mask;     // dense: register-index bits of the targets
  uint32_t lct;     // local controls on thread bits (mask of tile bits)
  int32_t rcm;      // local controls on register bits (register-index bits)
  int32_t fidx;     // phase: index of its out-of-tile factor
  int32_t pad;
  u64 gctrl;        // out-of-tile control mask (state bit positions)
  u64 loff;         // LDS-part offset: dense matrix, or phase tables
};

// LDS index of tile element e: bits >= S are folded into the S lowest bits
// (S = 5 for 8-byte, 4 for 16-byte amplitudes), so lanes that differ in
// tile bits of distinct classes hit distinct banks.  A bijection on
// [0, 2^TB).
template <typename R>
__host__ __device__ __forceinline__ int lds_swz(int e) {
  if (sizeof(R) == 4) return e ^ ((e >> 5) & 31) ^ ((e >> 10) & 31);
  return e ^ ((e >> 4) & 15) ^ ((e >> 8) & 15);
}
inline int swz_class(int b, int S) { return b < S ? b : (b - S) % S; }

__host__ __device__ constexpr int pdep_c(int v, int mask) {
  int r = 0, i = 0;
  for (int b = 0; b < 16; ++b)
    if ((mask >> b) & 1) {
      if ((v >> i) & 1) r |= 1 << b;
      ++i;
    }
  return r;
}
__host__ __device__ constexpr int popc_c(int v) { return v ? (v & 1) + popc_c(v >> 1) : 0; }
__host__ __device__ constexpr int topbit_c(int v) { return v > 1 ? 1 + topbit_c(v >> 1) : 0; }

// ---------------------------------------------------------------- layout --
template <int NRB>
struct RegLay {
  int ethr;       // tile bits set by the thread index
  int rb[NRB];    // tile bit of register-index bit j
};

// LDS slot of the element in register r: the swizzle is linear over GF(2),
// so it is the thread part's swizzle XOR the (wave-uniform) register part's.
template <typename R, int NRB>
__host__ __device__ __forceinline__ int reg_lds(int sw_thr, const RegLay<NRB>& L, int r) {
  int e = 0;
#pragma unroll
  for (int j = 0; j < NRB; ++j)
    if ((r >> j) & 1) e |= L.rb[j];
  return sw_thr ^ lds_swz<R>(e);
}

template <int NRB>
__host__ __device__ __forceinline__ RegLay<NRB> reg_layout(const RegSeg& sg, int NT, int tid) {
  RegLay<NRB> L;
  L.ethr = 0;
#pragma unroll
  for (int s = 0; s < kRegMaxBits; ++s)
    if (s < NT && ((tid >> s) & 1)) L.ethr |= 1 << sg.slot[s];
#pragma unroll
  for (int j = 0; j < NRB; ++j) L.rb[j] = 1 << sg.slot[NT + j];
  return L;
}

// Tile element held in register r.
template <int NRB>
__host__ __device__ __forceinline__ int reg_elem(const RegLay<NRB>& L, int r) {
  int e = L.ethr;
#pragma unroll
  for (int j = 0; j < NRB; ++j)
    if ((r >> j) & 1) e |= L.rb[j];
  return e;
}

// State offsets of a layout: thread part and one bit per register-index bit.
template <int NRB>
__host__ __device__ __forceinline__ void reg_state_offsets(const RegSeg& sg, const uint8_t* gbit, int NT,
                                                           int tid, u64& thr, u64 (&rb)[NRB]) {
  thr = 0;
#pragma unroll
  for (int s = 0; s < kRegMaxBits; ++s)
    if (s < NT && ((tid >> s) & 1)) thr |= u64(1) << gbit[sg.slot[s]];
#pragma unroll
  for (int j = 0; j < NRB; ++j) rb[j] = u64(1) << gbit[sg.slot[NT + j]];
}

template <int NRB>
__host__ __device__ __forceinline__ u64 reg_state_offset(const u64 (&rb)[NRB], int r) {
  u64 o = 0;
#pragma unroll
  for (int j = 0; j < NRB; ++j)
    if ((r >> j) & 1) o |= rb[j];
  return o;
}

// ------------------------------------------------------------------- ops --
__host__ __device__ __forceinline__ float qfma(float a, float b, float c) { return fmaf(a, b, c); }
__host__ __device__ __forceinline__ double qfma(double a, double b, double c) { return fma(a, b, c); }

// acc + m * x as four fused multiply-adds.
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cfma(cplx<R> m, cplx<R> x, cplx<R> acc) {
  acc.re = qfma(m.re, x.re, acc.re);
  acc.re = qfma(-m.im, x.im, acc.re);
  acc.im = qfma(m.re, x.im, acc.im);
  acc.im = qfma(m.im, x.re, acc.im);
  return acc;
}

// Dense op on the register bits MASK: row index of register g | pdep(c, MASK)
// is c (the lowest mask bit is the matrix index LSB).  Groups whose register
// bits fail the controls `rcm` are skipped; the caller masks the threads
// whose thread bits fail theirs.  REAL: a 2x2 matrix with real entries.
template <typename R, int NRB, int MASK, bool REAL>
__host__ __device__ __forceinline__ void reg_dense(cplx<R> (&x)[1 << NRB], const cplx<R>* __restrict__ m,
                                                   int rcm) {
  constexpr int NR = 1 << NRB;
  constexpr int N = 1 << popc_c(MASK);
  if constexpr (N == 2) {
    // 2x2: partial sums first, then each output as an FMA that is the last
    // use of its own input, so the result can stay in that input's register
    const cplx<R> m00 = m[0], m01 = m[1], m10 = m[2], m11 = m[3];
#pragma unroll
    for (int g = 0; g < NR; ++g) {
      if (g & MASK) continue;
      if ((g & rcm) != rcm) continue;
      cplx<R>& a = x[g];
      cplx<R>& b = x[g | MASK];
      if (REAL) {
        const R pa_re = m01.re * b.re, pa_im = m01.re * b.im;
        const R pb_re = m10.re * a.re, pb_im = m10.re * a.im;
        a.re = qfma(m00.re, a.re, pa_re);
        a.im = qfma(m00.re, a.im, pa_im);
        b.re = qfma(m11.re, b.re, pb_re);
        b.im = qfma(m11.re, b.im, pb_im);
      } else {
        R pa_re = qfma(m01.re, b.re, -m00.im * a.im);
        pa_re = qfma(-m01.im, b.im, pa_re);
        R pa_im = qfma(m01.re, b.im, m00.im * a.re);
        pa_im = qfma(m01.im, b.re, pa_im);
        R pb_re = qfma(m10.re, a.re, -m11.im * b.im);
        pb_re = qfma(-m10.im, a.im, pb_re);
        R pb_im = qfma(m10.re, a.im, m11.im * b.re);
        pb_im = qfma(m10.im, a.re, pb_im);
        a.re = qfma(m00.re, a.re, pa_re);
        a.im = qfma(m00.re, a.im, pa_im);
        b.re = qfma(m11.re, b.re, pb_re);
        b.im = qfma(m11.re, b.im, pb_im);
      }
    }
  } else {
#pragma unroll
    for (int g = 0; g < NR; ++g) {
      if (g & MASK) continue;
      if ((g & rcm) != rcm) continue;
      cplx<R> in[N];
#pragma unroll
      for (int c = 0; c < N; ++c) in[c] = x[g | pdep_c(c, MASK)];
#pragma unroll
      for (int r = 0; r < N; ++r) {
        cplx<R> acc = cmul(m[r * N], in[0]);
#pragma unroll
        for (int c = 1; c < N; ++c) acc = cfma(m[r * N + c], in[c], acc);
        x[g | pdep_c(r, MASK)] = acc;
      }
    }
  }
}

// Unscaled Hadamard butterfly on the register bit MASK: (a, b) -> (a + b, a - b).
template <typename R, int NRB, int MASK>
__host__ __device__ __forceinline__ void reg_had(cplx<R> (&x)[1 << NRB]) {
  constexpr int NR = 1 << NRB;
#pragma unroll
  for (int g = 0; g < NR; ++g) {
    if (g & MASK) continue;
    cplx<R>& a = x[g];
    cplx<R>& b = x[g | MASK];
    const R s_re = a.re + b.re, s_im = a.im + b.im;
    b.re = a.re - b.re;
    b.im = a.im - b.im;
    a.re = s_re;
    a.im = s_im;
  }
}

template <typename R, int NRB>
__host__ __device__ __forceinline__ void reg_dense_any(cplx<R> (&x)[1 << NRB], int kind, int mask,
                                                       const cplx<R>* __restrict__ m, int rcm) {
  if (kind == kRegHad) {
    switch (mask) {
#define QJ_REG_HAD(M)                                                          \
  case M:                                                                      \
    if constexpr (M < (1 << NRB)) reg_had<R, NRB, M>(x);                       \
    break;
      QJ_REG_HAD(1) QJ_REG_HAD(2) QJ_REG_HAD(4) QJ_REG_HAD(8) QJ_REG_HAD(16)
#undef QJ_REG_HAD
      default: break;
    }
    return;
  }
  if (kind == kRegDenseReal) {
    switch (mask) {
#define QJ_REG_REAL(M)                                                         \
  case M:                                                                      \
    if constexpr (M < (1 << NRB)) reg_dense<R, NRB, M, true>(x, m, rcm);       \
    break;
      QJ_REG_REAL(1) QJ_REG_REAL(2) QJ_REG_REAL(4) QJ_REG_REAL(8) QJ_REG_REAL(16)
#undef QJ_REG_REAL
      default: break;
    }
    return;
  }
  switch (mask) {
#define QJ_REG_CASE(M)                                                         \
  case M:                                                                      \
    if constexpr (M < (1 << NRB)) reg_dense<R, NRB, M, false>(x, m, rcm);      \
    break;
    QJ_REG_CASE(1) QJ_REG_CASE(2) QJ_REG_CASE(3) QJ_REG_CASE(4) QJ_REG_CASE(5) QJ_REG_CASE(6)
    QJ_REG_CASE(7) QJ_REG_CASE(8) QJ_REG_CASE(9) QJ_REG_CASE(10) QJ_REG_CASE(11) QJ_REG_CASE(12)
    QJ_REG_CASE(13) QJ_REG_CASE(14) QJ_REG_CASE(15) QJ_REG_CASE(16) QJ_REG_CASE(17) QJ_REG_CASE(18)
    QJ_REG_CASE(19) QJ_REG_CASE(20) QJ_REG_CASE(21) QJ_REG_CASE(22) QJ_REG_CASE(23) QJ_REG_CASE(24)
    QJ_REG_CASE(25) QJ_REG_CASE(26) QJ_REG_CASE(27) QJ_REG_CASE(28) QJ_REG_CASE(29) QJ_REG_CASE(30)
    QJ_REG_CASE(31)
#undef QJ_REG_CASE
    default: break;
  }
}

// Phase op in a layout: the factor of element (thread tid, register r) is
// f * TL[tid mod 64] * TW[tid / 64] * TR[r] -- the products of the per-bit
// factors over the set lane, wave and register slots (tables built by the
// planner).  Registers whose bits fail the controls `rcm` are skipped; the
// caller masks the threads whose thread bits fail theirs.
// RCM >= 0: the register controls as a compile-time mask (no per-register
// branches); RCM < 0: the runtime mask `rcm`.
template <typename R, int NRB, int RCM>
__host__ __device__ __forceinline__ void reg_phase_m(cplx<R> (&x)[1 << NRB], cplx<R> f,
                                                     const cplx<R>* __restrict__ tab, int NT, int tid, int rcm) {
  constexpr int NR = 1 << NRB;
  // register factors read ahead in groups (their LDS reads issued together,
  // not one read and one wait per register)
  constexpr int G = (sizeof(R) == 8 ? 4 : 8) < NR ? (sizeof(R) == 8 ? 4 : 8) : NR;
  const cplx<R>* tl = tab;
  const cplx<R>* tw = tl + 64;
  const cplx<R>* tr = tw + (1 << (NT - 6));
  const cplx<R> g = cmul(cmul(f, tl[tid & 63]), tw[tid >> 6]);
#pragma unroll
  for (int r0 = 0; r0 < NR; r0 += G) {
    cplx<R> t[G];
#pragma unroll
    for (int k = 0; k < G; ++k)
      if (RCM < 0 || ((r0 + k) & RCM) == RCM) t[k] = tr[r0 + k];
#pragma unroll
    for (int k = 0; k < G; ++k) {
      const int r = r0 + k;
      if (RCM >= 0 ? (r & RCM) != RCM : (r & rcm) != rcm) continue;
      const cplx<R> h = cmul(g, t[k]);
      // x *= h with the final FMAs as the last uses of x.re and x.im
      const R t_re = x[r].im * h.im, t_im = x[r].re * h.im;
      x[r].re = qfma(x[r].re, h.re, -t_re);
      x[r].im = qfma(x[r].im, h.re, t_im);
    }
  }
}

template <typename R, int NRB>
__host__ __device__ __forceinline__ void reg_phase(cplx<R> (&x)[1 << NRB], cplx<R> f,
                                                   const cplx<R>* __restrict__ tab, int NT, int tid, int rcm) {
  switch (rcm) {
    case 0: reg_phase_m<R, NRB, 0>(x, f, tab, NT, tid, rcm); break;
    case 1: reg_phase_m<R, NRB, 1>(x, f, tab, NT, tid, rcm); break;
    case 2: reg_phase_m<R, NRB, 2>(x, f, tab, NT, tid, rcm); break;
    case 4: reg_phase_m<R, NRB, 4>(x, f, tab, NT, tid, rcm); break;
    case 8: reg_phase_m<R, NRB, 8>(x, f, tab, NT, tid, rcm); break;
    default: reg_phase_m<R, NRB, -1>(x, f, tab, NT, tid, rcm); break;
  }
}

// Out-of-tile factor of a phase op for tile t (tables at `tab`).
template <typename R>
__host__ __device__ __forceinline__ cplx<R> reg_phase_outside(const cplx<R>* __restrict__ tab, int ga, u64 t) {
  return cmul(tab[t & ((u64(1) << ga) - 1)], tab[(size_t(1) << ga) + (t >> ga)]);
}

// -------------------------------------------------------------- planner --
struct RegPlanIn {
  int TB, NRB, S;              // tile bits, register bits, swizzle class bits
  int esize;                   // bytes per amplitude (8 or 16)
  int ga, gb;                  // phase table split of the out-of-tile bits
  int gpos[kRegMaxBits];       // tile bit -> state bit
  const qj_tile_op* ops;
  int nops;
  const char* data;
  u64 lds_cap;                 // largest LDS part a workgroup can hold
  bool reg_offsets = true;     // add the per-segment register offsets when they fit
};

inline int popc32(uint32_t v) { return __builtin_popcount(v); }

// Builds the plan blob (RegHead, RegSeg[], RegOp[], data) for the ops.
// Returns false when the ops need the LDS kernel (a dense op wider than the
// register bits, too few thread bits, or too many segments).
inline bool reg_plan(const RegPlanIn& in, std::vector<char>& blob) {
  const int TB = in.TB, NRB = in.NRB, NT = TB - NRB;
  if (NT < kRegLanes || TB > kRegMaxBits || TB > 30) return false;
  const uint32_t all = (uint32_t(1) << TB) - 1;

  // 1. relabel through the swaps
  struct P {
    int op, k;
    int tg[QJ_TILE_MAX_K];
    uint32_t lc;
    int8_t phys[kRegMaxBits];
  };
  int8_t phys[kRegMaxBits];
  for (int b = 0; b < kRegMaxBits; ++b) phys[b] = (int8_t)b;
  std::vector<P> pl;
  for (int o = 0; o < in.nops; ++o) {
    const qj_tile_op& op = in.ops[o];
    if (op.kind == QJ_TILE_SWAP) {
      std::swap(phys[op.target[0]], phys[op.target[1]]);
      continue;
    }
    P p;
    p.op = o;
    p.k = op.kind == QJ_TILE_DENSE ? op.k : 0;
    if (p.k > NRB) return false;
    for (int u = 0; u < p.k; ++u) p.tg[u] = phys[op.target[u]];
    p.lc = 0;
    for (int b = 0; b < TB; ++b)
      if ((op.local_ctrl >> b) & 1) p.lc |= uint32_t(1) << phys[b];
    std::memcpy(p.phys, phys, sizeof(phys));
    pl.push_back(p);
  }
  int inv[kRegMaxBits];
  for (int b = 0; b < TB; ++b) inv[phys[b]] = b;
  bool relabelled = false;
  for (int b = 0; b < TB; ++b) relabelled |= phys[b] != b;

  // 2. segments: register bits for the dense ops from i on, among `allowed`
  auto tmask = [&](size_t i) {
    uint32_t m = 0;
    for (int u = 0; u < pl[i].k; ++u) m |= uint32_t(1) << pl[i].tg[u];
    return m;
  };
  auto choose = [&](size_t i, uint32_t allowed) {
    uint32_t u = 0;
    for (size_t j = i; j < pl.size(); ++j) {
      if (!pl[j].k) continue;
      const uint32_t m = tmask(j);
      if ((m & ~allowed) || popc32(u | m) > NRB) break;
      u |= m;
    }
    for (int b = TB - 1; b >= 0 && popc32(u) < NRB; --b)
      if (((allowed >> b) & 1) && !((u >> b) & 1)) u |= uint32_t(1) << b;
    return u;
  };
  const uint32_t ld_lanes = (uint32_t(1) << kRegLanes) - 1;
  uint32_t st_lanes = 0;
  for (int b = 0; b < kRegLanes; ++b) st_lanes |= uint32_t(1) << phys[b];
  std::vector<uint32_t> regs;
  std::vector<int> first;
  uint32_t cur = choose(0, all & ~ld_lanes);
  regs.push_back(cur);
  first.push_back(0);
  for (size_t i = 0; i < pl.size(); ++i) {
    if (!pl[i].k) continue;
    const uint32_t m = tmask(i);
    if ((m & cur) == m) continue;
    cur = choose(i, all);
    regs.push_back(cur);
    first.push_back((int)i);
  }
  // a store layout of its own when the last layout cannot put the store bits
  // on the lanes, or when the swaps move amplitudes between threads (the
  // re-layout's barriers then order every load before every store)
  if ((regs.size() == 1 && (st_lanes != ld_lanes || relabelled)) || (cur & st_lanes)) {
    cur = choose(pl.size(), all & ~st_lanes);
    regs.push_back(cur);
    first.push_back((int)pl.size());
  }
  const int nseg = (int)regs.size();
  if (nseg > kRegMaxSeg) return false;

  std::vector<RegSeg> segs(nseg);
  for (int s = 0; s < nseg; ++s) {
    RegSeg& sg = segs[s];
    std::memset(&sg, 0, sizeof(sg));
    sg.o0 = first[s];
    sg.o1 = s + 1 < nseg ? first[s + 1] : (int)pl.size();
    const uint32_t thr = all & ~regs[s];
    std::vector<int> lanes;
    if (s == 0) {
      for (int b = 0; b < kRegLanes; ++b) lanes.push_back(b);
    } else if (s == nseg - 1) {
      for (int b = 0; b < kRegLanes; ++b) lanes.push_back(phys[b]);
    } else {
      // lanes 0..S-1 on bits of distinct swizzle classes where possible
      uint32_t used = 0;
      for (int c = 0; c < in.S && (int)lanes.size() < kRegLanes; ++c)
        for (int b = 0; b < TB; ++b)
          if (((thr >> b) & 1) && !((used >> b) & 1) && swz_class(b, in.S) == c) {
            lanes.push_back(b);
            used |= uint32_t(1) << b;
            break;
          }
      for (int b = 0; b < TB && (int)lanes.size() < kRegLanes; ++b)
        if (((thr >> b) & 1) && !((used >> b) & 1)) {
          lanes.push_back(b);
          used |= uint32_t(1) << b;
        }
    }
    uint32_t lane_mask = 0;
    int slot = 0;
    for (int b : lanes) {
      sg.slot[slot++] = (int8_t)b;
      lane_mask |= uint32_t(1) << b;
    }
    for (int b = 0; b < TB; ++b)
      if (((thr >> b) & 1) && !((lane_mask >> b) & 1)) sg.slot[slot++] = (int8_t)b;
    for (int b = 0; b < TB; ++b)
      if ((regs[s] >> b) & 1) sg.slot[slot++] = (int8_t)b;
    if (slot != TB) return false;
  }
  // per-segment register offsets (see RegHead::ub_off)
  std::vector<uint16_t> ub((size_t)nseg << NRB);
  for (int s = 0; s < nseg; ++s)
    for (int r = 0; r < (1 << NRB); ++r) {
      int e = 0;
      for (int j = 0; j < NRB; ++j)
        if ((r >> j) & 1) e |= 1 << segs[s].slot[NT + j];
      ub[((size_t)s << NRB) + r] = (uint16_t)((in.esize == 8 ? lds_swz<float>(e) : lds_swz<double>(e)) * in.esize);
    }

  // 3. ops, LDS data (matrices, W in slot order) and out-of-tile tables
  std::vector<RegOp> rops(pl.size());
  std::vector<char> ldata, gtab;
  std::vector<u64> ftab;
  double had_scale = 1.0;
  int nhad = 0;
  auto put = [](std::vector<char>& d, const void* src, size_t bytes) {
    const u64 off = d.size();
    d.resize(off + ((bytes + 15) & ~size_t(15)), 0);
    std::memcpy(d.data() + off, src, bytes);
    return off;
  };
  for (int s = 0; s < nseg; ++s) {
    const RegSeg& sg = segs[s];
    for (int i = sg.o0; i < sg.o1; ++i) {
      const P& p = pl[i];
      const qj_tile_op& op = in.ops[p.op];
      RegOp& ro = rops[i];
      std::memset(&ro, 0, sizeof(ro));
      ro.kind = op.kind == QJ_TILE_PHASE ? kRegPhase : kRegDense;
      ro.gctrl = op.global_ctrl;
      ro.fidx = -1;
      // local controls: thread-bit part (tile bits) and register-index part
      uint32_t thr_bits = 0;
      for (int q = 0; q < NT; ++q) thr_bits |= uint32_t(1) << sg.slot[q];
      ro.lct = p.lc & thr_bits;
      for (int j = 0; j < NRB; ++j)
        if ((p.lc >> sg.slot[NT + j]) & 1) ro.rcm |= 1 << j;
      if (op.kind == QJ_TILE_DENSE) {
        const int k = p.k, N = 1 << k;
        int j[QJ_TILE_MAX_K];
        for (int u = 0; u < k; ++u) {
          j[u] = -1;
          for (int q = 0; q < NRB; ++q)
            if (sg.slot[NT + q] == p.tg[u]) j[u] = q;
          if (j[u] < 0) return false;
          ro.mask |= 1 << j[u];
        }
        // kernel index bit q_u (rank of j[u] in the mask) <-> matrix bit k-1-u
        int perm[1 << QJ_TILE_MAX_K];
        for (int c = 0; c < N; ++c) {
          int a = 0;
          for (int u = 0; u < k; ++u) {
            const int q = popc32((uint32_t)ro.mask & ((uint32_t(1) << j[u]) - 1));
            if ((c >> q) & 1) a |= 1 << (k - 1 - u);
          }
          perm[c] = a;
        }
        std::vector<char> mat((size_t)N * N * in.esize);
        const char* src = in.data + op.offset;
        bool real = k == 1;
        for (int r = 0; r < N; ++r)
          for (int c = 0; c < N; ++c) {
            char* dst = mat.data() + ((size_t)r * N + c) * in.esize;
            std::memcpy(dst, src + ((size_t)perm[r] * N + perm[c]) * in.esize, in.esize);
            const bool zero_im = in.esize == 8 ? *reinterpret_cast<const float*>(dst + 4) == 0.0f
                                               : *reinterpret_cast<const double*>(dst + 8) == 0.0;
            real = real && zero_im;
          }
        if (real) ro.kind = kRegDenseReal;
        if (real && k == 1 && !ro.gctrl && !ro.lct && !ro.rcm && nhad < kRegMaxHad) {
          // Hadamard form: m00 = m01 = m10 = -m11 exactly, in the state's precision
          double e[4];
          for (int q = 0; q < 4; ++q)
            e[q] = in.esize == 8 ? (double)*reinterpret_cast<const float*>(mat.data() + q * 8)
                                 : *reinterpret_cast<const double*>(mat.data() + q * 16);
          if (e[0] != 0.0 && e[1] == e[0] && e[2] == e[0] && e[3] == -e[0]) {
            ro.kind = kRegHad;
            had_scale *= e[0];
            ++nhad;
          }
        }
        ro.loff = put(ldata, mat.data(), mat.size());
      } else {
        // ABI block: int32 {ga, gb, tb, pad}, OA[2^ga], OB[2^gb], W[tb] (logical bits)
        const char* src = in.data + op.offset + 16;
        const size_t tabs = ((size_t(1) << in.ga) + (size_t(1) << in.gb)) * in.esize;
        std::vector<double> wre(TB), wim(TB);  // per-slot factors
        for (int b = 0; b < TB; ++b) {
          // logical bit b lives in physical bit p.phys[b], held by the slot naming it
          int slot = -1;
          for (int q = 0; q < TB; ++q)
            if (sg.slot[q] == p.phys[b]) slot = q;
          if (slot < 0) return false;
          const char* wp = src + tabs + (size_t)b * in.esize;
          if (in.esize == 8) {
            float v[2];
            std::memcpy(v, wp, 8);
            wre[slot] = v[0];
            wim[slot] = v[1];
          } else {
            double v[2];
            std::memcpy(v, wp, 16);
            wre[slot] = v[0];
            wim[slot] = v[1];
          }
        }
        // TL over slots 0..5, TW over slots 6..NT-1, TR over the register slots
        const int nw = NT - kRegLanes;
        std::vector<char> tab(((size_t)64 + ((size_t)1 << nw) + ((size_t)1 << NRB)) * in.esize);
        size_t at = 0;
        auto table = [&](int s0, int nbits) {
          for (int e = 0; e < (1 << nbits); ++e) {
            double re = 1.0, im = 0.0;
            for (int j = 0; j < nbits; ++j)
              if ((e >> j) & 1) {
                const double nr = re * wre[s0 + j] - im * wim[s0 + j];
                im = re * wim[s0 + j] + im * wre[s0 + j];
                re = nr;
              }
            if (in.esize == 8) {
              const float v[2] = {(float)re, (float)im};
              std::memcpy(tab.data() + at, v, 8);
            } else {
              const double v[2] = {re, im};
              std::memcpy(tab.data() + at, v, 16);
            }
            at += in.esize;
          }
        };
        table(0, kRegLanes);
        table(kRegLanes, nw);
        table(NT, NRB);
        ro.loff = put(ldata, tab.data(), tab.size());
        ro.fidx = (int)ftab.size();
        ftab.push_back(put(gtab, src, tabs));
      }
    }
  }
  if ((int)ftab.size() > (1 << NT)) return false;  // one thread per out-of-tile factor

  RegHead h;
  std::memset(&h, 0, sizeof(h));
  h.TB = TB;
  h.NT = NT;
  h.NRB = NRB;
  h.nseg = nseg;
  h.nops = (int)rops.size();
  h.nphase = (int)ftab.size();
  h.ga = in.ga;
  h.gb = in.gb;
  h.had_scale = had_scale;
  h.nhad = nhad;
  for (int b = 0; b < TB; ++b) {
    h.gbit_ld[b] = (uint8_t)in.gpos[b];
    h.gbit_st[b] = (uint8_t)in.gpos[inv[b]];
  }

  auto align16 = [](u64 v) { return (v + 15) & ~u64(15); };
  h.segs_off = 0;
  h.ops_off = align16((u64)nseg * sizeof(RegSeg));
  h.ftab_off = align16(h.ops_off + (u64)rops.size() * sizeof(RegOp));
  h.ldata_off = align16(h.ftab_off + (u64)ftab.size() * sizeof(u64));
  h.lds_bytes = align16(h.ldata_off + ldata.size());
  if (h.lds_bytes > in.lds_cap) return false;
  const u64 ub_bytes = align16(ub.size() * sizeof(uint16_t));
  if (in.reg_offsets && NRB >= 3 && h.lds_bytes + ub_bytes <= in.lds_cap) {
    h.ub_off = h.lds_bytes;  // > 0: segments come first
    h.lds_bytes += ub_bytes;
  }
  h.lds_off = align16(sizeof(RegHead));
  const u64 gtab_off = h.lds_off + h.lds_bytes;
  for (u64& f : ftab) f += gtab_off;  // out-of-tile tables: offsets from the blob start
  blob.assign(gtab_off + gtab.size(), 0);
  char* lp = blob.data() + h.lds_off;
  std::memcpy(blob.data(), &h, sizeof(h));
  std::memcpy(lp + h.segs_off, segs.data(), (size_t)nseg * sizeof(RegSeg));
  if (!rops.empty()) std::memcpy(lp + h.ops_off, rops.data(), rops.size() * sizeof(RegOp));
  if (!ftab.empty()) std::memcpy(lp + h.ftab_off, ftab.data(), ftab.size() * sizeof(u64));
  if (!ldata.empty()) std::memcpy(lp + h.ldata_off, ldata.data(), ldata.size());
  if (h.ub_off) std::memcpy(lp + h.ub_off, ub.data(), ub.size() * sizeof(uint16_t));
  if (!gtab.empty()) std::memcpy(blob.data() + gtab_off, gtab.data(), gtab.size());
  return true;
}

}  // namespace qj
