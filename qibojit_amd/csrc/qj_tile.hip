// Fused tile passes: a list of gate operations applied to tiles of the
// state, one HBM read and one HBM write per pass.
//
// A tile is 2^TB amplitudes: all combinations of the L lowest bits (one
// contiguous run of 2^L amplitudes in memory) and of H = TB - L chosen high
// bit positions, for one assignment of the remaining bits (the tile base,
// whose compact index is the tile number t).  Workgroups take tiles in a
// grid-stride loop.
//
// Two kernels run a pass:
//  * k_tile_reg (qj_tile_reg.h): the tile lives in registers, 2^NRB
//    amplitudes per thread; dense ops are register mat-vecs, phase ops are
//    per-amplitude products, swaps relabel bits, and LDS only re-lays the
//    tile out between segments of ops.  Used whenever the widest dense op
//    fits the register bits (k <= 5 for complex64, k <= 4 for complex128).
//  * k_tile_pass: the tile lives in LDS and every op is one sweep over it
//    with a barrier after it.  It runs the passes the register kernel cannot
//    (or every pass, with QJ_TILE_KERNEL=lds).
//
// LDS kernel thread layout: thread `tid` of a BS-thread workgroup owns tile
// elements e = tid + j*BS.  The low 8 bits of e are then fixed per thread and
// the remaining bits are the same for all 64 lanes of a wave.
//  * Dense ops run as register mat-vecs on GB independent groups at a time
//    (all loads, then all arithmetic, then all stores), so LDS latency of
//    one group hides behind the others.
//  * Phase ops multiply element e by F(t) * A(e & 255) * B(e >> 8): F from
//    two tables over the bits of t (per tile, wave-uniform), A per thread,
//    B wave-uniform -- one complex product per element, no trigonometry.
//    The host expands the per-bit phase factors of the op into these tables.
// Kernel instances are specialised on the widest dense op of the pass
// (register budget) and on the block size.
//
// k_bit_swaps applies disjoint bit swaps (a qubit permutation that is its
// own inverse, such as the final swaps of a QFT) in one read and one write:
// a workgroup exchanges a tile with its partner tile.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qj_common.h"
#include "qj_tile_reg.h"

namespace qj {

constexpr int kTileMaxHighBits = 12;  // QJ_TILE_MAX_BITS - 1 (one low bit at least)
constexpr int kTileLA = 8;  // tile bits indexing phase table A

struct TileSpec {
  int L;                       // low contiguous bits
  int TB;                      // tile bits
  int H;                       // high bits
  int nout;                    // bits outside the tile (bits of the tile index t)
  int hb[kTileMaxHighBits];    // high bit positions, ascending
};

// Global offset of the high-bit combination h (bit j of h -> position hb[j]).
__device__ __forceinline__ u64 high_offset(int h, const TileSpec& ts) {
  u64 o = 0;
#pragma unroll
  for (int j = 0; j < kTileMaxHighBits; ++j)
    if (j < ts.H && ((h >> j) & 1)) o |= u64(1) << ts.hb[j];
  return o;
}

// LDS-kernel phase data at op.offset (built by the host from the per-bit
// factors of the op): int32 ga, int32 gb, int32 la, int32 pad; then complex
// tables OA[2^ga], OB[2^gb] over the tile index t (t & (2^ga-1), t >> ga),
// A[2^la] over tile bits < la, B[2^(TB-la)] over tile bits >= la.
struct PhaseHeader {
  int32_t ga, gb, la, pad;
};

template <typename R, int V, int BS>
__device__ __forceinline__ void tile_load(cplx<R>* __restrict__ tile, const cplx<R>* __restrict__ st,
                                          u64 base, const TileSpec& ts) {
  const int n = 1 << ts.TB;
  const int lmask = (1 << ts.L) - 1;
  for (int e = threadIdx.x * V; e < n; e += BS * V) {
    const u64 gi = base | (u64)(e & lmask) | high_offset(e >> ts.L, ts);
    const cvec<R, V> v = ldv<R, V>(st + gi);
    *reinterpret_cast<cvec<R, V>*>(tile + e) = v;
  }
}

template <typename R, int V, int BS>
__device__ __forceinline__ void tile_store(const cplx<R>* __restrict__ tile, cplx<R>* __restrict__ st,
                                           u64 base, const TileSpec& ts) {
  const int n = 1 << ts.TB;
  const int lmask = (1 << ts.L) - 1;
  for (int e = threadIdx.x * V; e < n; e += BS * V) {
    const u64 gi = base | (u64)(e & lmask) | high_offset(e >> ts.L, ts);
    stv<R, V>(st + gi, *reinterpret_cast<const cvec<R, V>*>(tile + e));
  }
}

// A phase op folded into the store of the dense op before it: element e is
// multiplied by f * A[e & amask] * B[e >> la] where its local controls hold.
template <typename R>
struct PhaseCtx {
  const cplx<R>* ta;
  const cplx<R>* tb;
  cplx<R> f;
  int amask, la, lc;
  bool on;
};

template <typename R>
__device__ __forceinline__ PhaseCtx<R> phase_ctx(const qj_tile_op& op, const char* __restrict__ data, u64 t) {
  PhaseCtx<R> c;
  const PhaseHeader h = *reinterpret_cast<const PhaseHeader*>(data + op.offset);
  const cplx<R>* oa = reinterpret_cast<const cplx<R>*>(data + op.offset + sizeof(PhaseHeader));
  const cplx<R>* ob = oa + (1 << h.ga);
  c.ta = ob + (1 << h.gb);
  c.tb = c.ta + (1 << h.la);
  c.f = cmul(oa[t & ((u64(1) << h.ga) - 1)], ob[t >> h.ga]);
  c.amask = (1 << h.la) - 1;
  c.la = h.la;
  c.lc = (int)op.local_ctrl;
  c.on = true;
  return c;
}

template <typename R, int K, int GB, int BS>
__device__ __forceinline__ void tile_dense(cplx<R>* __restrict__ tile, const qj_tile_op& op,
                                           const cplx<R>* __restrict__ m, int TB, const PhaseCtx<R>& ph) {
  constexpr int N = 1 << K;
  int off[N];
#pragma unroll
  for (int r = 0; r < N; ++r) {
    int o = 0;
#pragma unroll
    for (int u = 0; u < K; ++u)
      if ((r >> (K - 1 - u)) & 1) o |= 1 << op.target[u];
    off[r] = o;
  }
  int srt[K];  // ascending target bits for the zero insertion
#pragma unroll
  for (int u = 0; u < K; ++u) srt[u] = op.target[u];
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = a + 1; b < K; ++b)
      if (srt[b] < srt[a]) {
        const int t = srt[a];
        srt[a] = srt[b];
        srt[b] = t;
      }
  const int lc = (int)op.local_ctrl;
  const int ngr = 1 << (TB - K);
  for (int g0 = threadIdx.x; g0 < ngr; g0 += BS * GB) {
    int e0[GB];
    bool act[GB];
    cplx<R> x[GB][N];
#pragma unroll
    for (int b = 0; b < GB; ++b) {
      int e = g0 + b * BS;
      act[b] = e < ngr;
#pragma unroll
      for (int u = 0; u < K; ++u) {
        const int lo = e & ((1 << srt[u]) - 1);
        e = ((e ^ lo) << 1) | lo;
      }
      e0[b] = e;
      act[b] = act[b] && ((e & lc) == lc);
      if (act[b]) {
#pragma unroll
        for (int c = 0; c < N; ++c) x[b][c] = tile[e + off[c]];
      }
    }
#pragma unroll
    for (int b = 0; b < GB; ++b) {
      if (!act[b]) continue;
#pragma unroll
      for (int r = 0; r < N; ++r) {
        cplx<R> acc = cmul(m[r * N], x[b][0]);
#pragma unroll
        for (int c = 1; c < N; ++c) acc = cadd(acc, cmul(m[r * N + c], x[b][c]));
        const int e = e0[b] + off[r];
        if (ph.on && (e & ph.lc) == ph.lc)
          acc = cmul(acc, cmul(cmul(ph.f, ph.ta[e & ph.amask]), ph.tb[e >> ph.la]));
        tile[e] = acc;
      }
    }
  }
}

template <typename R, int BS>
__device__ __forceinline__ void tile_phase(cplx<R>* __restrict__ tile, const qj_tile_op& op,
                                           const char* __restrict__ data, u64 t, int TB) {
  const PhaseHeader h = *reinterpret_cast<const PhaseHeader*>(data + op.offset);
  const cplx<R>* oa = reinterpret_cast<const cplx<R>*>(data + op.offset + sizeof(PhaseHeader));
  const cplx<R>* ob = oa + (1 << h.ga);
  const cplx<R>* ta = ob + (1 << h.gb);
  const cplx<R>* tb = ta + (1 << h.la);
  const cplx<R> f = cmul(oa[t & ((u64(1) << h.ga) - 1)], ob[t >> h.ga]);
  const int amask = (1 << h.la) - 1;
  const int lc = (int)op.local_ctrl;
  const int n = 1 << TB;
  // BS is a multiple of 2^la (la <= 8), so e & amask == threadIdx.x & amask
  // for every element of this thread.
  const cplx<R> af = cmul(f, ta[threadIdx.x & amask]);
#pragma unroll 4
  for (int e = threadIdx.x; e < n; e += BS) {
    if ((e & lc) != lc) continue;
    // e >> la is the same for the 64 lanes of a wave: a scalar load
    const int bi = __builtin_amdgcn_readfirstlane(e >> h.la);
    tile[e] = cmul(tile[e], cmul(af, tb[bi]));
  }
}

// Swap of tile bits a and b: exchanges the elements with (bit a, bit b) =
// (1, 0) and (0, 1).
template <typename R, int BS>
__device__ __forceinline__ void tile_swap(cplx<R>* __restrict__ tile, int a, int b, int TB) {
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  const int ngr = 1 << (TB - 2);
  for (int g = threadIdx.x; g < ngr; g += BS) {
    int e = g;
    int l = e & ((1 << lo) - 1);
    e = ((e ^ l) << 1) | l;
    l = e & ((1 << hi) - 1);
    e = ((e ^ l) << 1) | l;
    const int e1 = e | (1 << a), e2 = e | (1 << b);
    const cplx<R> v = tile[e1];
    tile[e1] = tile[e2];
    tile[e2] = v;
  }
}

template <typename R, int V, int MAXK, int BS>
__global__ void __launch_bounds__(BS)
    k_tile_pass(cplx<R>* __restrict__ st, u64 ntiles, Ins tins, TileSpec ts,
                const qj_tile_op* __restrict__ ops, int nops, const char* __restrict__ data) {
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  cplx<R>* tile = reinterpret_cast<cplx<R>*>(smem_raw);
  for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const u64 base = insert_bits(t, tins);
    tile_load<R, V, BS>(tile, st, base, ts);
    __syncthreads();
    for (int o = 0; o < nops; ++o) {
      const qj_tile_op& op = ops[o];
      if ((base & op.global_ctrl) != op.global_ctrl) continue;  // uniform
      if (op.kind == QJ_TILE_PHASE) {
        tile_phase<R, BS>(tile, op, data, t, ts.TB);
      } else if (op.kind == QJ_TILE_SWAP) {
        tile_swap<R, BS>(tile, op.target[0], op.target[1], ts.TB);
      } else {
        const cplx<R>* m = reinterpret_cast<const cplx<R>*>(data + op.offset);
        // A phase op right after a dense op without local controls (which
        // then writes every element once) is applied in its store.
        PhaseCtx<R> ph;
        ph.on = false;
        if (op.local_ctrl == 0 && o + 1 < nops && ops[o + 1].kind == QJ_TILE_PHASE) {
          const qj_tile_op& nx = ops[o + 1];
          if ((base & nx.global_ctrl) == nx.global_ctrl) ph = phase_ctx<R>(nx, data, t);
          ++o;  // consumed (or skipped: its global controls do not hold here)
        }
        switch (op.k) {
          case 1: tile_dense<R, 1, 4, BS>(tile, op, m, ts.TB, ph); break;
          case 2: if (MAXK >= 2) tile_dense<R, 2, 2, BS>(tile, op, m, ts.TB, ph); break;
          case 3: if (MAXK >= 3) tile_dense<R, 3, 1, BS>(tile, op, m, ts.TB, ph); break;
          case 4: if (MAXK >= 4) tile_dense<R, 4, 1, BS>(tile, op, m, ts.TB, ph); break;
          default: if (MAXK >= 5) tile_dense<R, 5, 1, BS>(tile, op, m, ts.TB, ph); break;
        }
      }
      __syncthreads();
    }
    tile_store<R, V, BS>(tile, st, base, ts);
    __syncthreads();
  }
}

// Block size of the tile kernel: 512 threads (8 waves) by default, measured
// best on MI355X for 32-64 KiB tiles (two per CU); QJ_TILE_BLOCK
// (256/512/1024) overrides it for tuning.
static int tile_block_size() {
  static int bs = [] {
    const char* s = std::getenv("QJ_TILE_BLOCK");
    const int v = s ? std::atoi(s) : 512;
    return (v == 256 || v == 512 || v == 1024) ? v : 512;
  }();
  return bs;
}

template <typename R, int V, int MAXK>
static void launch_tile(int bs, u64 grid, size_t lds, hipStream_t s, cplx<R>* st, u64 ntiles, const Ins& tins,
                        const TileSpec& ts, const qj_tile_op* ops, int nops, const char* data) {
  dim3 gd((unsigned)grid);
  // 1024-thread blocks cap a thread at 128 VGPRs, which the complex128
  // 5-qubit mat-vec exceeds: that instance runs with 512 threads.
  constexpr bool wide_regs = sizeof(R) == 8 && MAXK >= 4;
  if (bs == 256) {
    k_tile_pass<R, V, MAXK, 256><<<gd, dim3(256), lds, s>>>(st, ntiles, tins, ts, ops, nops, data);
  } else if constexpr (wide_regs) {
    k_tile_pass<R, V, MAXK, 512><<<gd, dim3(512), lds, s>>>(st, ntiles, tins, ts, ops, nops, data);
  } else {
    if (bs == 512)
      k_tile_pass<R, V, MAXK, 512><<<gd, dim3(512), lds, s>>>(st, ntiles, tins, ts, ops, nops, data);
    else
      k_tile_pass<R, V, MAXK, 1024><<<gd, dim3(1024), lds, s>>>(st, ntiles, tins, ts, ops, nops, data);
  }
}

// ------------------------------------------------------ register kernel --
typedef __attribute__((address_space(3))) char lds_byte;
// Read-only plan data at wave-uniform addresses, read through the constant
// address space so that the loads are scalar.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(4))) const u32x4 const_u32x4;
// One amplitude as a two-element vector (8- or 16-byte LDS accesses at an
// integer LDS address).
template <typename R>
using amp2 = R __attribute__((ext_vector_type(2)));
template <typename R>
using lds_amp2 = __attribute__((address_space(3))) amp2<R>;

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ u64 uni64(u64 v) {
  return ((u64)(uint32_t)uni((int)(v >> 32)) << 32) | (uint32_t)uni((int)v);
}

// A persistent workgroup copies the LDS part of the plan once, then loops
// over tiles: the out-of-tile factors of the phase ops for the tile (one
// thread each, double-buffered by tile parity), the loads, and the ops of
// every segment; plan values read from LDS are made wave-uniform so the op
// dispatch is scalar.  WPE: waves per SIMD the register budget must allow.
template <typename R, int NRB, int BSMAX, int WPE>
__global__ void __launch_bounds__(BSMAX, WPE)
    k_tile_reg(cplx<R>* __restrict__ st, u64 ntiles, Ins tins, const char* __restrict__ plan) {
  constexpr int NR = 1 << NRB;
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  const RegHead& h = *reinterpret_cast<const RegHead*>(plan);
  const int TB = h.TB, NT = h.NT, nseg = h.nseg, nphase = h.nphase, ga = h.ga;
  cplx<R>* lds = reinterpret_cast<cplx<R>*>(smem_raw);
  char* lp = smem_raw + (sizeof(cplx<R>) << TB);
  cplx<R>* fv = reinterpret_cast<cplx<R>*>(lp + h.lds_bytes);
  const int tid = threadIdx.x;
  {
    const uint4* src = reinterpret_cast<const uint4*>(plan + h.lds_off);
    uint4* dst = reinterpret_cast<uint4*>(lp);
    const int n16 = (int)(h.lds_bytes / 16);
    for (int i = tid; i < n16; i += blockDim.x) dst[i] = src[i];
  }
  __syncthreads();
  const RegSeg* segs = reinterpret_cast<const RegSeg*>(lp + h.segs_off);
  // LDS byte address of the tile buffer: re-layout addresses are formed as
  // (base | thread part) ^ register part, which needs a 64 KiB-aligned base
  // (the buffer is the kernel's only LDS, so it starts at 0)
  const uint32_t lds_base = (uint32_t)(uintptr_t)(lds_byte*)smem_raw;
  // (the register parts are read from the plan in global memory: uniform
  // addresses, so they arrive in scalar registers)
  const uint16_t* ub_seg = h.ub_off && (lds_base & 0xffffu) == 0
                               ? reinterpret_cast<const uint16_t*>(plan + h.lds_off + h.ub_off)
                               : nullptr;
  const bool had = h.nhad != 0;
  const R had_scale = (R)h.had_scale;
  const RegOp* ops = reinterpret_cast<const RegOp*>(lp + h.ops_off);
  const u64* ftab = reinterpret_cast<const u64*>(lp + h.ftab_off);
  const char* ldata = lp + h.ldata_off;
  constexpr int AB = (int)sizeof(cplx<R>);
  u64 ld_thr, st_thr, ld_rb[NRB], st_rb[NRB];
  reg_state_offsets<NRB>(segs[0], h.gbit_ld, NT, tid, ld_thr, ld_rb);
  reg_state_offsets<NRB>(segs[nseg - 1], h.gbit_st, NT, tid, st_thr, st_rb);
#pragma unroll
  for (int j = 0; j < NRB; ++j) {
    ld_rb[j] = uni64(ld_rb[j]);
    st_rb[j] = uni64(st_rb[j]);
  }
  // amplitude addresses: a wave-uniform part (the tile base and the register
  // bits, scalar arithmetic) plus the lane's thread part in bytes
  const u64 ld_thr_b = ld_thr * AB, st_thr_b = st_thr * AB;
  // each thread's layout of segment 0 (the same for every tile) and, with
  // the segment table, its tile bits and swizzled LDS offset in segments
  // 1.. (written and read by that thread only)
  RegLay<NRB> lay0 = reg_layout<NRB>(segs[0], NT, tid);
#pragma unroll
  for (int j = 0; j < NRB; ++j) lay0.rb[j] = uni(lay0.rb[j]);
  const uint32_t swb0 = lds_base | (uint32_t)(lds_swz<R>(lay0.ethr) * AB);
  const bool segtab = ub_seg && h.segtab;
  uint32_t* stab = reinterpret_cast<uint32_t*>(fv + 2 * nphase) + tid;
  const int nthr = 1 << NT;
  if (segtab)
    for (int s = 1; s < nseg; ++s) {
      const int e = reg_layout<NRB>(segs[s], NT, tid).ethr;
      stab[(s - 1) * nthr] = (uint32_t)e | ((uint32_t)(lds_swz<R>(e) * AB) << 16);
    }
  int parity = 0;
  for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x, parity ^= 1) {
    const u64 base = insert_bits(t, tins);
    cplx<R>* f = fv + parity * nphase;
    if (tid < nphase) f[tid] = reg_phase_outside<R>(reinterpret_cast<const cplx<R>*>(plan + ftab[tid]), ga, t);
    cplx<R> x[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const char* p = reinterpret_cast<const char*>(st + (base | reg_state_offset<NRB>(ld_rb, r)));
      x[r] = *reinterpret_cast<const cplx<R>*>(p + ld_thr_b);
    }
    // the factors are written; every thread is done with the previous tile
    __syncthreads();
    RegLay<NRB> cur = lay0;
    uint32_t swb_cur = swb0;  // LDS byte address of the thread part in the current layout
    for (int s = 0; s < nseg; ++s) {
      const RegSeg& sg = segs[s];
      if (s > 0) {
        // re-layout through LDS; the first barrier keeps the previous
        // re-layout's reads ahead of these writes.  LDS byte address of
        // register r = the thread part's offset XOR the segment's
        // precomputed register part (the swizzle is linear); plans without
        // the offsets compute the register part per lane
        if (ub_seg) {
          RegLay<NRB> nxt;
          uint32_t swb_nxt;
          if (segtab) {
            const uint32_t v = stab[(s - 1) * nthr];
            nxt.ethr = (int)(v & 0xffffu);
            swb_nxt = lds_base | (v >> 16);
          } else {
            nxt.ethr = reg_layout<NRB>(sg, NT, tid).ethr;
            swb_nxt = lds_base | (uint32_t)(lds_swz<R>(nxt.ethr) * AB);
          }
          auto at = [](uint32_t a) { return (lds_amp2<R>*)(uintptr_t)a; };
          constexpr int NU = NR < 8 ? 1 : NR / 8;  // 16-byte groups of 8 offsets
          u32x4 u[NU];
          auto off = [&](int r) {
            const uint32_t w = u[r >> 3][(r >> 1) & 3];
            return (r & 1) ? (w >> 16) : (w & 0xffffu);
          };
#pragma unroll
          for (int i = 0; i < NU; ++i) u[i] = ((const_u32x4*)(ub_seg + (s - 1) * NR))[i];
          __syncthreads();
#pragma unroll
          for (int r = 0; r < NR; ++r) *at(swb_cur ^ off(r)) = amp2<R>{x[r].re, x[r].im};
#pragma unroll
          for (int i = 0; i < NU; ++i) u[i] = ((const_u32x4*)(ub_seg + s * NR))[i];
          __syncthreads();
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const amp2<R> v = *at(swb_nxt ^ off(r));
            x[r] = mk(v.x, v.y);
          }
          cur.ethr = nxt.ethr;
          swb_cur = swb_nxt;
        } else {
          RegLay<NRB> nxt = reg_layout<NRB>(sg, NT, tid);
#pragma unroll
          for (int j = 0; j < NRB; ++j) nxt.rb[j] = uni(nxt.rb[j]);
          const int sw_cur = lds_swz<R>(cur.ethr), sw_nxt = lds_swz<R>(nxt.ethr);
          __syncthreads();
#pragma unroll
          for (int r = 0; r < NR; ++r) lds[reg_lds<R, NRB>(sw_cur, cur, r)] = x[r];
          __syncthreads();
#pragma unroll
          for (int r = 0; r < NR; ++r) x[r] = lds[reg_lds<R, NRB>(sw_nxt, nxt, r)];
          cur = nxt;
        }
      }
      const int o1 = uni(sg.o1);
      for (int o = uni(sg.o0); o < o1; ++o) {
        const RegOp& op = ops[o];
        const u64 gc = uni64(op.gctrl);
        if ((base & gc) != gc) continue;
        const uint32_t lct = (uint32_t)uni((int)op.lct);
        if (((uint32_t)cur.ethr & lct) != lct) continue;  // threads failing thread-bit controls
        const int kind = uni(op.kind), rcm = uni(op.rcm);
        const cplx<R>* d = reinterpret_cast<const cplx<R>*>(ldata + uni((int)op.loff));
        if (kind == kRegPhase)
          reg_phase<R, NRB>(x, f[uni(op.fidx)], d, NT, tid, rcm);
        else
          reg_dense_any<R, NRB>(x, kind, uni(op.mask), d, rcm);
      }
    }
    if (had) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        x[r].re *= had_scale;
        x[r].im *= had_scale;
      }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      char* p = reinterpret_cast<char*>(st + (base | reg_state_offset<NRB>(st_rb, r)));
      *reinterpret_cast<cplx<R>*>(p + st_thr_b) = x[r];
    }
  }
}

// ----------------------------------------------------------- bit swaps --
struct SwapSpec {
  TileSpec ts;             // tile geometry (low bits + high positions)
  int8_t pl[16];           // tile bit -> tile bit of its partner
  int nbytes;              // bytes of the tile index looked up in `tabs`
};

// Tile element e = (tid + i * kSwapBlock) * V + k splits into a thread part, an
// iteration part and k, so its state offset and its permuted index are ORs
// of per-thread and per-iteration values.  The state offsets of tile t and
// of its partner tile come from byte tables (tabs[2*byte] for the tile,
// tabs[2*byte+1] for the partner, 256 entries each).  Every load of a tile
// is issued before any LDS write.
constexpr int kSwapBlock = 512;
constexpr int kSwapIters = 4;  // (2^12 amplitudes) / (512 * 2), (2^11) / (512 * 1)

template <typename R, int V>
__global__ void __launch_bounds__(kSwapBlock)
    k_bit_swaps(cplx<R>* __restrict__ st, u64 ntiles, SwapSpec sp, const u64* __restrict__ tabs) {
  extern __shared__ __attribute__((aligned(16))) char smem_raw[];
  const int TB = sp.ts.TB, L = sp.ts.L, n = 1 << TB;
  cplx<R>* a = reinterpret_cast<cplx<R>*>(smem_raw);
  cplx<R>* b = a + n;
  const int lmask = (1 << L) - 1;
  auto state_off = [&](int e) { return (u64)(e & lmask) | high_offset(e >> L, sp.ts); };
  auto perm = [&](int e) {
    int p = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (j < TB && ((e >> j) & 1)) p |= 1 << sp.pl[j];
    return p;
  };
  const int e_thr = threadIdx.x * V;
  const bool active = e_thr < n;  // states smaller than one block's share
  const u64 off_thr = state_off(e_thr);
  const int p_thr = perm(e_thr);
  const int p_one = perm(1);
  u64 off_it[kSwapIters];
  int p_it[kSwapIters];
#pragma unroll
  for (int i = 0; i < kSwapIters; ++i) {
    off_it[i] = off_thr | state_off(i * kSwapBlock * V);
    p_it[i] = p_thr | perm(i * kSwapBlock * V);
  }
  const int nbytes = sp.nbytes;
  for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    u64 base1 = 0, base2 = 0;
    for (int k = 0; k < nbytes; ++k) {
      const int byte = (int)((t >> (8 * k)) & 255);
      base1 |= tabs[(2 * k) * 256 + byte];
      base2 |= tabs[(2 * k + 1) * 256 + byte];
    }
    if (base2 < base1) continue;  // the partner tile's workgroup moves both (uniform)
    const bool pair = base2 != base1;
    cvec<R, V> v1[kSwapIters], v2[kSwapIters];
#pragma unroll
    for (int i = 0; i < kSwapIters; ++i) {
      if (i * kSwapBlock * V >= n || !active) continue;
      v1[i] = ldv<R, V>(st + (base1 | off_it[i]));
      if (pair) v2[i] = ldv<R, V>(st + (base2 | off_it[i]));
    }
#pragma unroll
    for (int i = 0; i < kSwapIters; ++i) {
      const int e = e_thr + i * kSwapBlock * V;
      if (i * kSwapBlock * V >= n || !active) continue;
#pragma unroll
      for (int k = 0; k < V; ++k) a[lds_swz<R>(e + k)] = v1[i].v[k];
      if (pair) {
#pragma unroll
        for (int k = 0; k < V; ++k) b[lds_swz<R>(e + k)] = v2[i].v[k];
      }
    }
    __syncthreads();
    // element e of the partner tile <- element pi(e) of this tile, and the reverse
#pragma unroll
    for (int i = 0; i < kSwapIters; ++i) {
      if (i * kSwapBlock * V >= n || !active) continue;
      cvec<R, V> o1, o2;
#pragma unroll
      for (int k = 0; k < V; ++k) {
        const int p = k ? (p_it[i] | p_one) : p_it[i];
        o1.v[k] = a[lds_swz<R>(p)];
        o2.v[k] = b[lds_swz<R>(p)];
      }
      stv<R, V>(st + (base2 | off_it[i]), o1);
      if (pair) stv<R, V>(st + (base1 | off_it[i]), o2);
    }
    __syncthreads();
  }
}


// ---------------------------------------------------------------- host --
// Per-bit phase blocks -> the LDS kernel's four-table blocks; dense matrices
// copied.  `ops_out` is `ops` with new offsets into `data_out`.
template <typename R>
static void lds_kernel_data(const qj_tile_op* ops, int nops, const char* data, int TB, int ga, int gb,
                            std::vector<qj_tile_op>& ops_out, std::vector<char>& data_out) {
  const int la = TB < kTileLA ? TB : kTileLA;
  ops_out.assign(ops, ops + nops);
  data_out.clear();
  auto put = [&](const void* src, size_t bytes) {
    const u64 off = data_out.size();
    data_out.resize(off + ((bytes + 15) & ~size_t(15)), 0);
    std::memcpy(data_out.data() + off, src, bytes);
    return off;
  };
  for (int o = 0; o < nops; ++o) {
    const qj_tile_op& op = ops[o];
    if (op.kind == QJ_TILE_DENSE) {
      ops_out[o].offset = put(data + op.offset, ((size_t)1 << (2 * op.k)) * sizeof(cplx<R>));
    } else if (op.kind == QJ_TILE_PHASE) {
      const cplx<R>* src = reinterpret_cast<const cplx<R>*>(data + op.offset + 16);
      const size_t na = size_t(1) << ga, nb = size_t(1) << gb;
      const cplx<R>* w = src + na + nb;
      std::vector<cplx<R>> blk(na + nb + (size_t(1) << la) + (size_t(1) << (TB - la)));
      std::memcpy(blk.data(), src, (na + nb) * sizeof(cplx<R>));
      auto prod = [&](int lo, int nbits, cplx<R>* out) {
        for (int e = 0; e < (1 << nbits); ++e) {
          double re = 1.0, im = 0.0;
          for (int j = 0; j < nbits; ++j)
            if ((e >> j) & 1) {
              const double wr = (double)w[lo + j].re, wi = (double)w[lo + j].im;
              const double nr = re * wr - im * wi;
              im = re * wi + im * wr;
              re = nr;
            }
          out[e] = mk((R)re, (R)im);
        }
      };
      prod(0, la, blk.data() + na + nb);
      prod(la, TB - la, blk.data() + na + nb + (size_t(1) << la));
      const int32_t head[4] = {ga, gb, la, 0};
      const u64 off = put(head, sizeof(head));
      put(blk.data(), blk.size() * sizeof(cplx<R>));
      ops_out[o].offset = off;
    }
  }
}

// LDS of one CU, and the share of it a register-kernel plan may take next to
// a 64 KiB tile (two workgroups per CU).
constexpr size_t kLdsPerCu = 160 * 1024;
constexpr u64 kRegPlanLds = 15 * 1024;

// Compute units and the largest LDS allocation of one workgroup, per device.
static void device_limits(int* cus, size_t* max_lds) {
  static int ncu[64] = {0};
  static int mlds[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = -1;
  if (dev >= 0 && !ncu[dev]) {
    int n = 0, m = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    if (hipDeviceGetAttribute(&m, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || m <= 0) m = 65536;
    ncu[dev] = n;
    mlds[dev] = m;
  }
  *cus = dev >= 0 ? ncu[dev] : 256;
  *max_lds = dev >= 0 ? (size_t)mlds[dev] : 65536;
}

// Register bits of the register kernel: 4 by default (QJ_TILE_NRB overrides:
// 4/5 for complex64, 3/4 for complex128), at least the widest dense op.
// Returns 0 when the pass needs the LDS kernel.
static int tile_register_bits(bool dbl, int maxk) {
  const char* kern = std::getenv("QJ_TILE_KERNEL");
  if (kern && std::strcmp(kern, "lds") == 0) return 0;
  const char* e = std::getenv("QJ_TILE_NRB");
  const int v = e ? std::atoi(e) : 0;
  const int lo = dbl ? 3 : 4, hi = dbl ? 4 : 5;
  int nrb = (v >= lo && v <= hi) ? v : 4;
  if (maxk > nrb) nrb = maxk;
  return nrb <= hi ? nrb : 0;
}

template <typename R, int NRB, int BSMAX, int WPE>
static void launch_reg(dim3 gd, dim3 bd, size_t lds, hipStream_t s, cplx<R>* st, u64 ntiles, const Ins& tins,
                       const char* plan) {
  // dynamic LDS above 64 KiB needs the kernel's limit raised first
  if (lds > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile_reg<R, NRB, BSMAX, WPE>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  k_tile_reg<R, NRB, BSMAX, WPE><<<gd, bd, lds, s>>>(st, ntiles, tins, plan);
}

template <typename R>
static void launch_tile_reg(int nrb, int NT, u64 grid, size_t lds, hipStream_t s, cplx<R>* st, u64 ntiles,
                            const Ins& tins, const char* plan) {
  const dim3 gd((unsigned)grid), bd(1u << NT);
  if constexpr (sizeof(R) == 4) {
    if (nrb == 4)
      launch_reg<R, 4, 512, 4>(gd, bd, lds, s, st, ntiles, tins, plan);
    else
      launch_reg<R, 5, 256, 2>(gd, bd, lds, s, st, ntiles, tins, plan);
  } else {
    if (nrb == 3)
      launch_reg<R, 3, 512, 4>(gd, bd, lds, s, st, ntiles, tins, plan);
    else
      launch_reg<R, 4, 256, 2>(gd, bd, lds, s, st, ntiles, tins, plan);
  }
}

template <typename R>
static int tile_pass_impl(void* state, int nqubits, int L, const int32_t* high, int nhigh,
                          const qj_tile_op* ops, int nops, const void* data, u64 data_bytes,
                          hipStream_t s) {
  const int TB = L + nhigh;
  if (nqubits < 1 || nqubits > 62) return set_error(QJ_EVALUE, "apply_tile_pass: bad nqubits %d", nqubits);
  if (L < 1 || nhigh < 0 || nhigh > kTileMaxHighBits || TB > nqubits || TB > QJ_TILE_MAX_BITS)
    return set_error(QJ_EVALUE, "apply_tile_pass: bad tile (low_bits=%d, nhigh=%d)", L, nhigh);
  if (((size_t)1 << TB) * sizeof(cplx<R>) > 65536)
    return set_error(QJ_EVALUE, "apply_tile_pass: tile of 2^%d amplitudes exceeds 64 KiB of LDS", TB);
  const int V = 16 / (int)sizeof(cplx<R>);
  if ((1 << L) < V) return set_error(QJ_EVALUE, "apply_tile_pass: low_bits too small");
  TileSpec ts;
  std::memset(&ts, 0, sizeof(ts));
  ts.L = L;
  ts.TB = TB;
  ts.H = nhigh;
  ts.nout = nqubits - TB;
  Ins tins;
  std::memset(&tins, 0, sizeof(tins));
  tins.npos = TB;
  for (int b = 0; b < L; ++b) tins.pos[b] = b;
  for (int j = 0; j < nhigh; ++j) {
    if (high[j] < L || high[j] >= nqubits || (j && high[j] <= high[j - 1]))
      return set_error(QJ_EVALUE, "apply_tile_pass: high bits must be ascending, >= low_bits, < nqubits");
    tins.pos[L + j] = high[j];
    ts.hb[j] = high[j];
  }
  u64 tile_mask = (u64(1) << L) - 1;
  for (int j = 0; j < nhigh; ++j) tile_mask |= u64(1) << high[j];
  const int ga = (ts.nout + 1) / 2, gb = ts.nout - ga;
  int maxk = 1;
  for (int o = 0; o < nops; ++o) {
    const qj_tile_op& op = ops[o];
    if (op.global_ctrl & tile_mask)
      return set_error(QJ_EVALUE, "apply_tile_pass: op %d has a global control inside the tile", o);
    if (op.global_ctrl >> nqubits)
      return set_error(QJ_EVALUE, "apply_tile_pass: op %d global control beyond the state", o);
    if (op.local_ctrl >> TB) return set_error(QJ_EVALUE, "apply_tile_pass: op %d local control out of tile", o);
    if (op.kind == QJ_TILE_DENSE || op.kind == QJ_TILE_SWAP) {
      const int k = op.kind == QJ_TILE_SWAP ? 2 : op.k;
      if (op.kind == QJ_TILE_SWAP && (op.k != 2 || op.local_ctrl || op.global_ctrl))
        return set_error(QJ_EVALUE, "apply_tile_pass: swap op %d needs k=2 and no controls", o);
      if (k < 1 || k > QJ_TILE_MAX_K) return set_error(QJ_EVALUE, "apply_tile_pass: op %d has k=%d", o, k);
      u64 seen = 0;
      for (int u = 0; u < k; ++u) {
        const int b = op.target[u];
        if (b < 0 || b >= TB || ((seen >> b) & 1) || ((op.local_ctrl >> b) & 1))
          return set_error(QJ_EVALUE, "apply_tile_pass: op %d target %d invalid", o, b);
        seen |= u64(1) << b;
      }
      if (op.kind == QJ_TILE_DENSE) {
        if (op.offset % 16) return set_error(QJ_EVALUE, "apply_tile_pass: op %d data offset not 16-aligned", o);
        const u64 need = op.offset + ((u64(1) << (2 * k)) * sizeof(cplx<R>));
        if (need > data_bytes) return set_error(QJ_EVALUE, "apply_tile_pass: op %d matrix outside data", o);
        if (k > maxk) maxk = k;
      }
    } else if (op.kind == QJ_TILE_PHASE) {
      if (op.offset % 16) return set_error(QJ_EVALUE, "apply_tile_pass: op %d data offset not 16-aligned", o);
      if (op.offset + 16 > data_bytes)
        return set_error(QJ_EVALUE, "apply_tile_pass: op %d phase data outside data", o);
      int32_t hh[4];
      std::memcpy(hh, static_cast<const char*>(data) + op.offset, sizeof(hh));
      if (hh[0] != ga || hh[1] != gb || hh[2] != TB)
        return set_error(QJ_EVALUE,
                         "apply_tile_pass: op %d phase block shape (%d,%d,%d) does not match the tile (%d,%d,%d)",
                         o, hh[0], hh[1], hh[2], ga, gb, TB);
      const u64 need = op.offset + 16 + ((u64(1) << ga) + (u64(1) << gb) + (u64)TB) * sizeof(cplx<R>);
      if (need > data_bytes) return set_error(QJ_EVALUE, "apply_tile_pass: op %d phase tables outside data", o);
    } else {
      return set_error(QJ_EVALUE, "apply_tile_pass: op %d has unknown kind %d", o, op.kind);
    }
  }
  const u64 ntiles = u64(1) << (nqubits - TB);
  const u64 grid = ntiles < kMaxGrid ? ntiles : kMaxGrid;
  const size_t lds = ((size_t)1 << TB) * sizeof(cplx<R>);
  cplx<R>* st = static_cast<cplx<R>*>(state);
  const char* cdata = static_cast<const char*>(data);

  const int nrb = tile_register_bits(sizeof(R) == 8, maxk);
  if (nrb) {
    RegPlanIn in;
    std::memset(&in, 0, sizeof(in));
    in.TB = TB;
    in.NRB = nrb;
    in.S = sizeof(R) == 4 ? 5 : 4;
    in.esize = (int)sizeof(cplx<R>);
    in.ga = ga;
    in.gb = gb;
    for (int b = 0; b < TB; ++b) in.gpos[b] = tins.pos[b];
    in.ops = ops;
    in.nops = nops;
    in.data = cdata;
    in.lds_cap = kRegPlanLds;
    // QJ_TILE_OFFSETS=0: re-layouts compute every register's swizzle per lane
    static const bool offsets = [] {
      const char* e = std::getenv("QJ_TILE_OFFSETS");
      return !(e && std::strcmp(e, "0") == 0);
    }();
    in.reg_offsets = offsets;
    std::vector<char> plan;
    RegHead hh;
    size_t lds_reg = 0;
    int cus = 0;
    size_t max_lds = 0;
    device_limits(&cus, &max_lds);
    const bool planned = reg_plan(in, plan);
    if (planned) {
      std::memcpy(&hh, plan.data(), sizeof(hh));
      lds_reg = lds + hh.lds_bytes + 2 * (size_t)hh.nphase * sizeof(cplx<R>);
    }
    if (planned && lds_reg <= max_lds) {
      const int bs = 1 << (TB - nrb);
      // per-thread table of the later segments' layouts, when it keeps two
      // workgroups per CU
      const size_t tab = hh.ub_off ? (size_t)(hh.nseg - 1) * (size_t)bs * 4 : 0;
      if (tab && (lds_reg + tab) * 2 <= kLdsPerCu && lds_reg + tab <= max_lds) {
        lds_reg += tab;
        reinterpret_cast<RegHead*>(plan.data())->segtab = 1;
      }
      int per_cu = (int)(kLdsPerCu / lds_reg);
      per_cu = std::min(per_cu, 2048 / bs);
      per_cu = std::max(per_cu, 1);
      const u64 rgrid = std::min<u64>(ntiles, (u64)cus * per_cu);
      void* dplan = nullptr;
      int ticket = -1;
      int rc = ring_upload(plan.data(), plan.size(), s, &dplan, &ticket);
      if (rc) return rc;
      launch_tile_reg<R>(nrb, TB - nrb, rgrid, lds_reg, s, st, ntiles, tins, static_cast<const char*>(dplan));
      rc = check_launch("apply_tile_pass launch");
      ring_release(ticket, s);
      return rc;
    }
  }

  // LDS kernel: ops and data travel in one staging upload, ops first (64 B
  // each, so the data block starts 16-byte aligned)
  std::vector<qj_tile_op> lops;
  std::vector<char> ldata;
  lds_kernel_data<R>(ops, nops, cdata, TB, ga, gb, lops, ldata);
  const size_t ops_bytes = (size_t)nops * sizeof(qj_tile_op);
  // never empty: an empty vector's data() is null and the ring copies from it
  std::vector<char> buf(ops_bytes + ldata.size() > 16 ? ops_bytes + ldata.size() : 16, 0);
  if (ops_bytes) std::memcpy(buf.data(), lops.data(), ops_bytes);
  if (!ldata.empty()) std::memcpy(buf.data() + ops_bytes, ldata.data(), ldata.size());
  void* dbuf = nullptr;
  int ticket = -1;
  int rc = ring_upload(buf.data(), buf.size(), s, &dbuf, &ticket);
  if (rc) return rc;
  const qj_tile_op* dops = static_cast<const qj_tile_op*>(dbuf);
  const char* ddata = static_cast<const char*>(dbuf) + ops_bytes;
  const int bs = tile_block_size();
  constexpr int VV = 16 / (int)sizeof(cplx<R>);
  if (maxk <= 2)
    launch_tile<R, VV, 2>(bs, grid, lds, s, st, ntiles, tins, ts, dops, nops, ddata);
  else if (maxk == 3)
    launch_tile<R, VV, 3>(bs, grid, lds, s, st, ntiles, tins, ts, dops, nops, ddata);
  else
    launch_tile<R, VV, 5>(bs, grid, lds, s, st, ntiles, tins, ts, dops, nops, ddata);
  rc = check_launch("apply_tile_pass launch");
  ring_release(ticket, s);
  return rc;
}

template <typename R>
static int bit_swaps_impl(void* state, int nqubits, const int32_t* pairs, int npairs, hipStream_t s) {
  if (nqubits < 1 || nqubits > 62) return set_error(QJ_EVALUE, "apply_bit_swaps: bad nqubits %d", nqubits);
  if (npairs < 0 || npairs > nqubits / 2) return set_error(QJ_EVALUE, "apply_bit_swaps: bad npairs %d", npairs);
  int partner[64];
  for (int b = 0; b < 64; ++b) partner[b] = b;
  for (int j = 0; j < npairs; ++j) {
    const int a = pairs[2 * j], c = pairs[2 * j + 1];
    if (a < 0 || c < 0 || a >= nqubits || c >= nqubits || a == c || partner[a] != a || partner[c] != c)
      return set_error(QJ_EVALUE, "apply_bit_swaps: pair %d (%d, %d) invalid or not disjoint", j, a, c);
    partner[a] = c;
    partner[c] = a;
  }
  if (npairs == 0) return QJ_OK;
  constexpr int V = 16 / (int)sizeof(cplx<R>);
  // tile: the 512-byte run of low bits, their partners, and further bits
  // (unswapped ones, or both bits of a pair) up to 32 KiB of amplitudes, so
  // that the tile set is closed under the permutation
  const int L = std::min(sizeof(R) == 4 ? 6 : 5, nqubits);
  const int tb_max = sizeof(R) == 4 ? 12 : 11;
  u64 in_tile = (u64(1) << L) - 1;
  for (int b = 0; b < L; ++b) in_tile |= u64(1) << partner[b];
  int tb = __builtin_popcountll(in_tile);
  for (int b = L; b < nqubits && tb < tb_max; ++b) {
    if ((in_tile >> b) & 1) continue;
    const int add = partner[b] == b ? 1 : 2;
    if (tb + add > tb_max || tb + add - L > kTileMaxHighBits) continue;
    in_tile |= (u64(1) << b) | (u64(1) << partner[b]);
    tb += add;
  }
  SwapSpec sp;
  std::memset(&sp, 0, sizeof(sp));
  TileSpec& ts = sp.ts;
  ts.L = L;
  ts.TB = tb;
  ts.H = tb - L;
  ts.nout = nqubits - tb;
  int local_of[64], gpos[16];
  int nl = 0;
  for (int b = 0; b < nqubits; ++b)
    if ((in_tile >> b) & 1) {
      if (b >= L) ts.hb[nl - L] = b;
      gpos[nl] = b;
      local_of[b] = nl++;
    }
  for (int i = 0; i < tb; ++i) sp.pl[i] = (int8_t)local_of[partner[gpos[i]]];
  int outside[64];
  int no = 0;
  for (int b = 0; b < nqubits; ++b)
    if (!((in_tile >> b) & 1)) outside[no++] = b;
  // The tile index walks the outside bits in ascending order of
  // max(bit, partner bit): tiles in flight together then differ in bits that
  // are low on both sides, so they and their partner tiles stay within a few
  // memory pages (for a bit reversal this is the middle bits first:
  // 18.9 -> 14.0 ms for the 32-qubit complex64 reversal, 37.8 -> 23.7 ms in
  // complex128).  QJ_SWAP_ORDER=ascending restores plain ascending order.
  const char* order = std::getenv("QJ_SWAP_ORDER");
  if (!(order && std::strcmp(order, "ascending") == 0))
    std::stable_sort(outside, outside + no, [&partner](int x, int y) {
      return std::max(x, partner[x]) < std::max(y, partner[y]);
    });
  // byte tables: state offset of tile t (and of its partner) = OR over the
  // bytes of t of one entry each
  const int nbytes = (no + 7) / 8;
  std::vector<u64> tabs((size_t)std::max(1, 2 * nbytes) * 256, 0);
  for (int k = 0; k < nbytes; ++k)
    for (int v = 0; v < 256; ++v) {
      u64 o1 = 0, o2 = 0;
      for (int j = 0; j < 8; ++j) {
        const int r = 8 * k + j;
        if (r < no && ((v >> j) & 1)) {
          o1 |= u64(1) << outside[r];
          o2 |= u64(1) << partner[outside[r]];
        }
      }
      tabs[(size_t)(2 * k) * 256 + v] = o1;
      tabs[(size_t)(2 * k + 1) * 256 + v] = o2;
    }
  sp.nbytes = nbytes;
  void* dtabs = nullptr;
  int ticket = -1;
  int rc = ring_upload(tabs.data(), tabs.size() * sizeof(u64), s, &dtabs, &ticket);
  if (rc) return rc;
  const u64 ntiles = u64(1) << ts.nout;
  int cus = 0;
  size_t max_lds = 0;
  device_limits(&cus, &max_lds);
  const size_t lds = 2 * ((size_t)1 << tb) * sizeof(cplx<R>);
  const u64 per_cu = std::max<u64>(1, std::min<u64>(kLdsPerCu / lds, 2048 / kSwapBlock));
  const u64 grid = std::min<u64>(ntiles, (u64)cus * per_cu);
  k_bit_swaps<R, V><<<dim3((unsigned)grid), dim3(kSwapBlock), lds, s>>>(static_cast<cplx<R>*>(state), ntiles, sp,
                                                                       static_cast<const u64*>(dtabs));
  rc = check_launch("apply_bit_swaps launch");
  ring_release(ticket, s);
  return rc;
}

}  // namespace qj

using namespace qj;

extern "C" int qj_apply_tile_pass(void* state, int dtype, int nqubits, int low_bits, const int32_t* high_bits,
                                  int nhigh, const qj_tile_op* ops, int nops, const void* data,
                                  uint64_t data_bytes, void* stream) {
  if (!state) return set_error(QJ_EVALUE, "apply_tile_pass: null state");
  if (nops < 0 || (nops > 0 && !ops)) return set_error(QJ_EVALUE, "apply_tile_pass: bad ops");
  if (nhigh > 0 && !high_bits) return set_error(QJ_EVALUE, "apply_tile_pass: null high bits");
  if (data_bytes > 0 && !data) return set_error(QJ_EVALUE, "apply_tile_pass: null data");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64)
    return tile_pass_impl<float>(state, nqubits, low_bits, high_bits, nhigh, ops, nops, data, data_bytes, s);
  if (dtype == QJ_COMPLEX128)
    return tile_pass_impl<double>(state, nqubits, low_bits, high_bits, nhigh, ops, nops, data, data_bytes, s);
  return set_error(QJ_EVALUE, "apply_tile_pass: dtype must be complex64/complex128");
}

extern "C" int qj_apply_bit_swaps(void* state, int dtype, int nqubits, const int32_t* pairs, int npairs,
                                  void* stream) {
  if (!state) return set_error(QJ_EVALUE, "apply_bit_swaps: null state");
  if (npairs > 0 && !pairs) return set_error(QJ_EVALUE, "apply_bit_swaps: null pairs");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64) return bit_swaps_impl<float>(state, nqubits, pairs, npairs, s);
  if (dtype == QJ_COMPLEX128) return bit_swaps_impl<double>(state, nqubits, pairs, npairs, s);
  return set_error(QJ_EVALUE, "apply_bit_swaps: dtype must be complex64/complex128");
}
