// Fused tile passes: a list of gate operations applied to LDS-resident tiles
// of the state, one HBM read and one HBM write per pass.
//
// A tile is 2^TB amplitudes: all combinations of the L lowest bits (one
// contiguous run of 2^L amplitudes in memory) and of H = TB - L chosen high
// bit positions, for one assignment of the remaining bits (the tile base,
// whose compact index is the tile number t).  Workgroups take tiles in a
// grid-stride loop: load (a wave reads 64 consecutive 16-byte vectors of one
// run), apply every op with a barrier between ops, store.
//
// Thread layout: thread `tid` of a BS-thread workgroup owns tile elements
// e = tid + j*BS.  The low 8 bits of e are then fixed per thread and the
// remaining bits are the same for all 64 lanes of a wave.
//  * Dense ops run as register mat-vecs on GB independent groups at a time
//    (all loads, then all arithmetic, then all stores), so LDS latency of
//    one group hides behind the others.
//  * Phase ops multiply element e by F(t) * A(e & 255) * B(e >> 8): F from
//    two tables over the bits of t (per tile, wave-uniform), A per thread,
//    B wave-uniform -- one complex product per element, no trigonometry.
// Kernel instances are specialised on the widest dense op of the pass
// (register budget) and on the block size.
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qj_common.h"

namespace qj {

constexpr int kTileMaxHighBits = 8;
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

// Phase data at op.offset: int32 ga, int32 gb, int32 la, int32 pad; then
// complex tables OA[2^ga], OB[2^gb] over the tile index t (t & (2^ga-1), t >> ga),
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
  int maxk = 1;
  const int la = TB < kTileLA ? TB : kTileLA;
  for (int o = 0; o < nops; ++o) {
    const qj_tile_op& op = ops[o];
    if (op.global_ctrl & tile_mask)
      return set_error(QJ_EVALUE, "apply_tile_pass: op %d has a global control inside the tile", o);
    if (op.local_ctrl >> TB) return set_error(QJ_EVALUE, "apply_tile_pass: op %d local control out of tile", o);
    if (op.offset % 16) return set_error(QJ_EVALUE, "apply_tile_pass: op %d data offset not 16-aligned", o);
    if (op.kind == QJ_TILE_DENSE) {
      if (op.k < 1 || op.k > QJ_TILE_MAX_K) return set_error(QJ_EVALUE, "apply_tile_pass: op %d has k=%d", o, op.k);
      u64 seen = 0;
      for (int u = 0; u < op.k; ++u) {
        const int b = op.target[u];
        if (b < 0 || b >= TB || ((seen >> b) & 1) || ((op.local_ctrl >> b) & 1))
          return set_error(QJ_EVALUE, "apply_tile_pass: op %d target %d invalid", o, b);
        seen |= u64(1) << b;
      }
      const u64 need = op.offset + ((u64(1) << (2 * op.k)) * sizeof(cplx<R>));
      if (need > data_bytes) return set_error(QJ_EVALUE, "apply_tile_pass: op %d matrix outside data", o);
      if (op.k > maxk) maxk = op.k;
    } else if (op.kind == QJ_TILE_PHASE) {
      if (op.offset + sizeof(PhaseHeader) > data_bytes)
        return set_error(QJ_EVALUE, "apply_tile_pass: op %d phase data outside data", o);
      PhaseHeader hh;
      std::memcpy(&hh, static_cast<const char*>(data) + op.offset, sizeof(hh));
      if (hh.ga < 0 || hh.gb < 0 || hh.ga + hh.gb != ts.nout || hh.la != la)
        return set_error(QJ_EVALUE, "apply_tile_pass: op %d phase table shape (%d,%d,%d) does not match the tile",
                         o, hh.ga, hh.gb, hh.la);
      const u64 need = op.offset + sizeof(PhaseHeader) +
                       ((u64(1) << hh.ga) + (u64(1) << hh.gb) + (u64(1) << la) + (u64(1) << (TB - la))) *
                           sizeof(cplx<R>);
      if (need > data_bytes) return set_error(QJ_EVALUE, "apply_tile_pass: op %d phase tables outside data", o);
    } else {
      return set_error(QJ_EVALUE, "apply_tile_pass: op %d has unknown kind %d", o, op.kind);
    }
  }
  // ops and data travel in one staging upload, ops first (64 B each, so the
  // data block starts 16-byte aligned)
  const size_t ops_bytes = (size_t)nops * sizeof(qj_tile_op);
  // never empty: an empty vector's data() is null and the ring copies from it
  std::vector<char> buf(ops_bytes + data_bytes > 16 ? ops_bytes + data_bytes : 16, 0);
  if (ops_bytes) std::memcpy(buf.data(), ops, ops_bytes);
  if (data_bytes) std::memcpy(buf.data() + ops_bytes, data, data_bytes);
  void* dbuf = nullptr;
  int ticket = -1;
  int rc = ring_upload(buf.data(), buf.size(), s, &dbuf, &ticket);
  if (rc) return rc;
  const qj_tile_op* dops = static_cast<const qj_tile_op*>(dbuf);
  const char* ddata = static_cast<const char*>(dbuf) + ops_bytes;
  const u64 ntiles = u64(1) << (nqubits - TB);
  const u64 grid = ntiles < kMaxGrid ? ntiles : kMaxGrid;
  const size_t lds = ((size_t)1 << TB) * sizeof(cplx<R>);
  cplx<R>* st = static_cast<cplx<R>*>(state);
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
