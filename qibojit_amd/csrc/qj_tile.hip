// Fused tile passes: a list of gate operations applied to LDS-resident tiles
// of the state, one HBM read and one HBM write per pass.
//
// A tile is 2^TB amplitudes: all combinations of the L lowest bits (one
// contiguous run of 2^L amplitudes in memory) and of H = TB - L chosen high
// bit positions, for one assignment of the remaining bits (the tile base).
// Workgroups take tiles in a grid-stride loop: load (coalesced: a wave reads
// 64 consecutive vectors of one run), apply every op with a barrier between
// ops, store.  Dense ops on tile-local targets run as per-group register
// mat-vecs; diagonal (phase) ops multiply each element by
// exp(i*angle) built from two per-pass lookup tables and one per-tile
// factor for the out-of-tile bits, so no trigonometry runs per element.
#include <cstring>
#include <vector>

#include "qj_common.h"

namespace qj {

constexpr int kTileMaxHighBits = 8;

struct TileSpec {
  int L;                       // low contiguous bits
  int TB;                      // tile bits
  int LA;                      // phase table split: tile bits < LA index table A
  int H;                       // high bits
  int hb[kTileMaxHighBits];    // their global positions, ascending
};

// Global offset of the high-bit combination h (bit j of h -> position hb[j]).
__device__ __forceinline__ u64 high_offset(int h, const TileSpec& ts) {
  u64 o = 0;
#pragma unroll
  for (int j = 0; j < kTileMaxHighBits; ++j)
    if (j < ts.H && ((h >> j) & 1)) o |= u64(1) << ts.hb[j];
  return o;
}

struct PhaseHeader {
  double c0;
  int32_t nout;
  int32_t pad;
};
struct PhaseTerm {
  int32_t bit;
  int32_t pad;
  double theta;
};

template <typename R, int V>
__device__ __forceinline__ void tile_load(cplx<R>* __restrict__ tile, const cplx<R>* __restrict__ st,
                                          u64 base, const TileSpec& ts) {
  const int n = 1 << ts.TB;
  const int lmask = (1 << ts.L) - 1;
  for (int e = threadIdx.x * V; e < n; e += kBlock * V) {
    const u64 gi = base | (u64)(e & lmask) | high_offset(e >> ts.L, ts);
    const cvec<R, V> v = ldv<R, V>(st + gi);
#pragma unroll
    for (int j = 0; j < V; ++j) tile[e + j] = v.v[j];
  }
}

template <typename R, int V>
__device__ __forceinline__ void tile_store(const cplx<R>* __restrict__ tile, cplx<R>* __restrict__ st,
                                           u64 base, const TileSpec& ts) {
  const int n = 1 << ts.TB;
  const int lmask = (1 << ts.L) - 1;
  for (int e = threadIdx.x * V; e < n; e += kBlock * V) {
    const u64 gi = base | (u64)(e & lmask) | high_offset(e >> ts.L, ts);
    cvec<R, V> v;
#pragma unroll
    for (int j = 0; j < V; ++j) v.v[j] = tile[e + j];
    stv<R, V>(st + gi, v);
  }
}

template <typename R, int K>
__device__ __forceinline__ void tile_dense(cplx<R>* __restrict__ tile, const qj_tile_op& op,
                                           const cplx<R>* __restrict__ m, int TB) {
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
  // ascending target bits for the zero insertion
  int srt[K];
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
  for (int g = threadIdx.x; g < ngr; g += kBlock) {
    int e0 = g;
#pragma unroll
    for (int u = 0; u < K; ++u) {
      const int lo = e0 & ((1 << srt[u]) - 1);
      e0 = ((e0 ^ lo) << 1) | lo;
    }
    if ((e0 & lc) != lc) continue;
    cplx<R> x[N];
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] = tile[e0 + off[c]];
#pragma unroll
    for (int r = 0; r < N; ++r) {
      cplx<R> acc = cmul(m[r * N], x[0]);
#pragma unroll
      for (int c = 1; c < N; ++c) acc = cadd(acc, cmul(m[r * N + c], x[c]));
      tile[e0 + off[r]] = acc;
    }
  }
}

template <typename R>
__device__ __forceinline__ void tile_phase(cplx<R>* __restrict__ tile, const qj_tile_op& op,
                                           const char* __restrict__ data, u64 base, const TileSpec& ts) {
  const PhaseHeader* h = reinterpret_cast<const PhaseHeader*>(data + op.offset);
  const PhaseTerm* terms = reinterpret_cast<const PhaseTerm*>(h + 1);
  double ang = h->c0;
  for (int j = 0; j < h->nout; ++j)
    if ((base >> terms[j].bit) & 1) ang += terms[j].theta;
  double sn, cs;
  sincos(ang, &sn, &cs);
  const cplx<R> f = mk<R>((R)cs, (R)sn);
  const cplx<R>* ta = reinterpret_cast<const cplx<R>*>(terms + h->nout);
  const cplx<R>* tb = ta + (1 << ts.LA);
  const int amask = (1 << ts.LA) - 1;
  const int lc = (int)op.local_ctrl;
  const int n = 1 << ts.TB;
  for (int e = threadIdx.x; e < n; e += kBlock) {
    if ((e & lc) != lc) continue;
    const cplx<R> ph = cmul(cmul(f, ta[e & amask]), tb[e >> ts.LA]);
    tile[e] = cmul(tile[e], ph);
  }
}

template <typename R, int V>
__global__ void __launch_bounds__(kBlock)
    k_tile_pass(cplx<R>* __restrict__ st, u64 ntiles, Ins tins, TileSpec ts,
                const qj_tile_op* __restrict__ ops, int nops, const char* __restrict__ data) {
  extern __shared__ char smem_raw[];
  cplx<R>* tile = reinterpret_cast<cplx<R>*>(smem_raw);
  for (u64 t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const u64 base = insert_bits(t, tins);
    tile_load<R, V>(tile, st, base, ts);
    __syncthreads();
    for (int o = 0; o < nops; ++o) {
      const qj_tile_op& op = ops[o];
      if ((base & op.global_ctrl) != op.global_ctrl) continue;  // uniform
      if (op.kind == QJ_TILE_PHASE) {
        tile_phase<R>(tile, op, data, base, ts);
      } else {
        const cplx<R>* m = reinterpret_cast<const cplx<R>*>(data + op.offset);
        switch (op.k) {
          case 1: tile_dense<R, 1>(tile, op, m, ts.TB); break;
          case 2: tile_dense<R, 2>(tile, op, m, ts.TB); break;
          case 3: tile_dense<R, 3>(tile, op, m, ts.TB); break;
          case 4: tile_dense<R, 4>(tile, op, m, ts.TB); break;
          default: tile_dense<R, 5>(tile, op, m, ts.TB); break;
        }
      }
      __syncthreads();
    }
    tile_store<R, V>(tile, st, base, ts);
    __syncthreads();
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
  ts.LA = TB / 2;
  ts.H = nhigh;
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
  for (int o = 0; o < nops; ++o) {
    const qj_tile_op& op = ops[o];
    if (op.global_ctrl & tile_mask)
      return set_error(QJ_EVALUE, "apply_tile_pass: op %d has a global control inside the tile", o);
    if (op.local_ctrl >> TB) return set_error(QJ_EVALUE, "apply_tile_pass: op %d local control out of tile", o);
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
    } else if (op.kind == QJ_TILE_PHASE) {
      if (op.offset + sizeof(PhaseHeader) > data_bytes)
        return set_error(QJ_EVALUE, "apply_tile_pass: op %d phase data outside data", o);
      PhaseHeader hh;
      std::memcpy(&hh, static_cast<const char*>(data) + op.offset, sizeof(hh));
      const u64 need = op.offset + sizeof(PhaseHeader) + (u64)hh.nout * sizeof(PhaseTerm) +
                       ((u64(1) << ts.LA) + (u64(1) << (TB - ts.LA))) * sizeof(cplx<R>);
      if (hh.nout < 0 || need > data_bytes)
        return set_error(QJ_EVALUE, "apply_tile_pass: op %d phase tables outside data", o);
      const PhaseTerm* terms = reinterpret_cast<const PhaseTerm*>(static_cast<const char*>(data) + op.offset +
                                                                  sizeof(PhaseHeader));
      for (int j = 0; j < hh.nout; ++j) {
        PhaseTerm tt;
        std::memcpy(&tt, terms + j, sizeof(tt));
        if (tt.bit < 0 || tt.bit >= nqubits || ((tile_mask >> tt.bit) & 1))
          return set_error(QJ_EVALUE, "apply_tile_pass: op %d phase bit %d invalid", o, tt.bit);
      }
    } else {
      return set_error(QJ_EVALUE, "apply_tile_pass: op %d has unknown kind %d", o, op.kind);
    }
  }
  // ops and data travel in one staging upload, ops first (64 B each, so the
  // data block starts 16-byte aligned)
  const size_t ops_bytes = (size_t)nops * sizeof(qj_tile_op);
  std::vector<char> buf(ops_bytes + data_bytes);
  std::memcpy(buf.data(), ops, ops_bytes);
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
  if (V == 2)
    k_tile_pass<R, 2><<<dim3((unsigned)grid), dim3(kBlock), lds, s>>>(st, ntiles, tins, ts, dops, nops, ddata);
  else
    k_tile_pass<R, 1><<<dim3((unsigned)grid), dim3(kBlock), lds, s>>>(st, ntiles, tins, ts, dops, nops, ddata);
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
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64)
    return tile_pass_impl<float>(state, nqubits, low_bits, high_bits, nhigh, ops, nops, data, data_bytes, s);
  if (dtype == QJ_COMPLEX128)
    return tile_pass_impl<double>(state, nqubits, low_bits, high_bits, nhigh, ops, nops, data, data_bytes, s);
  return set_error(QJ_EVALUE, "apply_tile_pass: dtype must be complex64/complex128");
}
