// Gate-application kernels for gfx950 and their C-ABI entry points.
//
// Every kernel is one in-place pass over the amplitudes a gate can change:
// the free index g (the bits that are neither targets nor controls) is mapped
// to a base index by inserting the gate's bit positions (insert_bits), and
// each work item owns the 2^k amplitudes base + offset(r) -- a bijection, so
// work items never overlap and no synchronisation is needed
// (custom_operators/gates.py:5-12 is the same bijection).
//
// Memory shape: consecutive lanes take consecutive g, so the loads of a wave
// are contiguous runs of the state for every target position; when every
// gate bit is >= 1 the lowest free bit is bit 0 and a lane moves V = 2
// consecutive amplitudes per access (16 B for complex64, 32 B for
// complex128; only complex64 uses it, complex128 moves one 16 B amplitude
// per access).  Z-like kernels (z, z_pow) read and write only the half they
// change; swap only touches |01>,|10>; fsim skips |00>.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qj_common.h"

namespace qj {

template <typename R>
struct M1 {
  cplx<R> m[4];
};
template <typename R>
struct M2 {
  cplx<R> m[16];
};

// One work item per thread and a full grid: measured fastest on gfx950 for
// every target stride (more items per thread or a persistent grid lowered
// the in-place rate by 5-10%).
constexpr int kIlp = 1;

// --------------------------------------------------------------- 1 qubit --
template <typename R, int KIND, int V, int NT = kBlock>
__global__ void __launch_bounds__(NT)
    k_one_qubit(cplx<R>* __restrict__ st, u64 ngroups, u64 tmask, Ins ins, M1<R> g) {
  const u64 stride = (u64)gridDim.x * (NT * kIlp);
  for (u64 b = (u64)blockIdx.x * (NT * kIlp) + threadIdx.x; b < ngroups; b += stride) {
    u64 i0[kIlp];
    bool ok[kIlp];
    cvec<R, V> a[kIlp], c[kIlp];
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      const u64 gv = b + (u64)u * NT;
      ok[u] = gv < ngroups;
      i0[u] = ok[u] ? insert_bits(gv * V, ins) : 0;
    }
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      if (!ok[u]) continue;
      if (KIND != QJ_APPLY_Z && KIND != QJ_APPLY_Z_POW) a[u] = ldv<R, V>(st + i0[u]);
      c[u] = ldv<R, V>(st + i0[u] + tmask);
    }
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      if (!ok[u]) continue;
      cvec<R, V> na, nc;
#pragma unroll
      for (int e = 0; e < V; ++e) {
        const cplx<R> x = a[u].v[e], y = c[u].v[e];
        if (KIND == QJ_APPLY_GATE) {
          na.v[e] = mac2(g.m[0], x, g.m[1], y);
          nc.v[e] = mac2(g.m[2], x, g.m[3], y);
        } else if (KIND == QJ_APPLY_X) {
          na.v[e] = y;
          nc.v[e] = x;
        } else if (KIND == QJ_APPLY_Y) {
          na.v[e] = mk(y.im, -y.re);   // -i * y
          nc.v[e] = mk(-x.im, x.re);   //  i * x
        } else if (KIND == QJ_APPLY_Z) {
          nc.v[e] = mk(-y.re, -y.im);
        } else {  // QJ_APPLY_Z_POW
          nc.v[e] = cmul(g.m[0], y);
        }
      }
      if (KIND != QJ_APPLY_Z && KIND != QJ_APPLY_Z_POW) stv<R, V>(st + i0[u], na);
      stv<R, V>(st + i0[u] + tmask, nc);
    }
  }
}

// --------------------------------------------------------------- 2 qubits --
// Amplitude r of a quad sits at base + (r>>1)*q1 + (r&1)*q2, q1/q2 = the
// strides of the first/second listed target (matrix index 2*t1 + t2).
template <typename R, int KIND, int V, int NT = kBlock>
__global__ void __launch_bounds__(NT)
    k_two_qubit(cplx<R>* __restrict__ st, u64 ngroups, u64 q1, u64 q2, Ins ins, M2<R> g) {
  const u64 stride = (u64)gridDim.x * (NT * kIlp);
  for (u64 b = (u64)blockIdx.x * (NT * kIlp) + threadIdx.x; b < ngroups; b += stride) {
    u64 i0[kIlp];
    bool ok[kIlp];
    cvec<R, V> a0[kIlp], a1[kIlp], a2[kIlp], a3[kIlp];
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      const u64 gv = b + (u64)u * NT;
      ok[u] = gv < ngroups;
      i0[u] = ok[u] ? insert_bits(gv * V, ins) : 0;
    }
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      if (!ok[u]) continue;
      if (KIND == QJ_APPLY_TWO_QUBIT_GATE) a0[u] = ldv<R, V>(st + i0[u]);
      a1[u] = ldv<R, V>(st + i0[u] + q2);
      a2[u] = ldv<R, V>(st + i0[u] + q1);
      if (KIND != QJ_APPLY_SWAP) a3[u] = ldv<R, V>(st + i0[u] + q1 + q2);
    }
#pragma unroll
    for (int u = 0; u < kIlp; ++u) {
      if (!ok[u]) continue;
      cvec<R, V> n0, n1, n2, n3;
#pragma unroll
      for (int e = 0; e < V; ++e) {
        if (KIND == QJ_APPLY_TWO_QUBIT_GATE) {
          const cplx<R> x0 = a0[u].v[e], x1 = a1[u].v[e], x2 = a2[u].v[e], x3 = a3[u].v[e];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            cplx<R> acc = cadd(cadd(cmul(g.m[4 * r + 0], x0), cmul(g.m[4 * r + 1], x1)),
                               cadd(cmul(g.m[4 * r + 2], x2), cmul(g.m[4 * r + 3], x3)));
            if (r == 0) n0.v[e] = acc;
            if (r == 1) n1.v[e] = acc;
            if (r == 2) n2.v[e] = acc;
            if (r == 3) n3.v[e] = acc;
          }
        } else if (KIND == QJ_APPLY_SWAP) {
          n1.v[e] = a2[u].v[e];
          n2.v[e] = a1[u].v[e];
        } else {  // QJ_APPLY_FSIM
          const cplx<R> x1 = a1[u].v[e], x2 = a2[u].v[e];
          n1.v[e] = mac2(g.m[0], x1, g.m[1], x2);
          n2.v[e] = mac2(g.m[2], x1, g.m[3], x2);
          n3.v[e] = cmul(g.m[4], a3[u].v[e]);
        }
      }
      if (KIND == QJ_APPLY_TWO_QUBIT_GATE) stv<R, V>(st + i0[u], n0);
      stv<R, V>(st + i0[u] + q2, n1);
      stv<R, V>(st + i0[u] + q1, n2);
      if (KIND != QJ_APPLY_SWAP) stv<R, V>(st + i0[u] + q1 + q2, n3);
    }
  }
}

// ------------------------------------------------------ 3..5 qubits (regs) --
// The 2^K amplitudes of a group are held in registers; the matrix is staged
// once per workgroup in LDS and read as wave-uniform broadcasts.
template <int K>
struct Offs {
  u64 o[1 << K];
};

template <typename R, int K, int NT = kBlock>
__global__ void __launch_bounds__(NT)
    k_multi_reg(cplx<R>* __restrict__ st, u64 ngroups, Ins ins, Offs<K> off,
                const cplx<R>* __restrict__ mat) {
  constexpr int N = 1 << K;
  __shared__ cplx<R> sm[N * N];
  const u64 stride = (u64)gridDim.x * NT;
  u64 g = (u64)blockIdx.x * NT + threadIdx.x;
  // the first group's loads are issued before the matrix is staged, so the
  // workgroup barrier does not delay them (one group per thread is the
  // usual launch)
  u64 base = 0;
  cplx<R> x[N];
  if (g < ngroups) {
    base = insert_bits(g, ins);
#pragma unroll
    for (int c = 0; c < N; ++c) x[c] = st[base + off.o[c]];
  }
  for (int t = threadIdx.x; t < N * N; t += NT) sm[t] = mat[t];
  __syncthreads();
  for (; g < ngroups; g += stride) {
#pragma unroll 1
    for (int r = 0; r < N; ++r) {
      cplx<R> acc = cmul(sm[r * N], x[0]);
#pragma unroll
      for (int c = 1; c < N; ++c) acc = cadd(acc, cmul(sm[r * N + c], x[c]));
      st[base + off.o[r]] = acc;
    }
    // the next group's amplitudes (every row of this one is stored: the
    // update stays in place)
    if (g + stride < ngroups) {
      base = insert_bits(g + stride, ins);
#pragma unroll
      for (int c = 0; c < N; ++c) x[c] = st[base + off.o[c]];
    }
  }
}

// ---------------------------------------------------- any k (LDS-staged) --
// GP groups per iteration are gathered into LDS, then each thread produces
// output rows; used for k > 5 (and as the reference path for any k).
template <typename R>
__global__ void __launch_bounds__(kBlock)
    k_multi_lds(cplx<R>* __restrict__ st, u64 ngroups, Ins ins, int K,
                const u64* __restrict__ offs, const cplx<R>* __restrict__ mat) {
  extern __shared__ char smem_raw[];
  cplx<R>* xs = reinterpret_cast<cplx<R>*>(smem_raw);
  const int N = 1 << K;
  const int GP = N >= kBlock ? 1 : kBlock / N;
  for (u64 g0 = (u64)blockIdx.x * GP; g0 < ngroups; g0 += (u64)gridDim.x * GP) {
    for (int t = threadIdx.x; t < GP * N; t += kBlock) {
      const int gi = t / N, c = t % N;
      const u64 g = g0 + gi;
      if (g < ngroups) xs[t] = st[insert_bits(g, ins) + offs[c]];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < GP * N; t += kBlock) {
      const int gi = t / N, r = t % N;
      const u64 g = g0 + gi;
      if (g < ngroups) {
        const cplx<R>* row = mat + (size_t)r * N;
        const cplx<R>* xv = xs + gi * N;
        cplx<R> acc = cmul(row[0], xv[0]);
        for (int c = 1; c < N; ++c) acc = cadd(acc, cmul(row[c], xv[c]));
        st[insert_bits(g, ins) + offs[r]] = acc;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ host side --

// Validates the sorted qubit list and builds the insertion spec: every
// position gets a 0 inserted, control positions (those not in `targets`) are
// then set to 1.
// ---- gates whose targets all lie in the 6 lowest bits: in-wave exchange --
// A wave owns a chunk of 64 consecutive amplitudes, one per lane, so every
// load and store of the wave is one contiguous run (the pair kernels above
// read such gates with strided per-lane accesses).  Lane l holds the
// amplitude of chunk-local index l; its group partners (other values of the
// target bits) live in lanes l ^ xmask[d] and arrive by cross-lane shuffles.
// A butterfly of conditional swaps orders the group by row index
// (x[c] = y[c ^ r]) so that the new amplitude sum_c U[r][c] x[c] is summed in
// ascending c, the reference's order.  Controls above bit 5 select the
// chunks (insert_bits on the chunk index); lanes whose control bits below
// bit 6 are not all 1 store their amplitude unchanged (every lane joins the
// shuffles, and every store covers whole lines: masked stores leave partial
// lines, measured 6% slower for CNOT(28, 29) at 30 qubits).
constexpr int kWaveBits = 6;

template <int K>
struct WaveSpec {
  int tb[K];           // target bit of row-index bit K-1-u (first target = MSB)
  int xmask[1 << K];   // lane xor of group member d
  int cmask, cones;    // low control bits of the lane index
};

template <typename R>
__device__ __forceinline__ cplx<R> shfl_xor_c(cplx<R> a, int mask) {
  return mk<R>(__shfl_xor(a.re, mask, 64), __shfl_xor(a.im, mask, 64));
}

// ILP chunks per wave and loop step (QJ_WAVE_ILP = 1 (default), 2, 4).
template <typename R, int K, int ILP>
__global__ void __launch_bounds__(kBlock)
    k_wave(cplx<R>* __restrict__ st, u64 nchunks, Ins ins, WaveSpec<K> ws, const cplx<R>* __restrict__ mat) {
  constexpr int N = 1 << K;
  const int lane = threadIdx.x & 63;
  int r = 0;
#pragma unroll
  for (int u = 0; u < K; ++u) r |= ((lane >> ws.tb[u]) & 1) << (K - 1 - u);
  cplx<R> row[N];
#pragma unroll
  for (int c = 0; c < N; ++c) row[c] = mat[r * N + c];
  const bool store = (lane & ws.cmask) == ws.cones;
  const u64 wave = (u64)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const u64 nwaves = (u64)gridDim.x * (kBlock / 64);
  for (u64 c0 = wave * ILP; c0 < nchunks; c0 += nwaves * ILP) {  // uniform per wave
    u64 idx[ILP];
    cplx<R> a[ILP];
#pragma unroll
    for (int i = 0; i < ILP; ++i) {
      const u64 c = c0 + i < nchunks ? c0 + i : c0;
      idx[i] = (insert_bits(c, ins) << kWaveBits) | (u64)lane;
      a[i] = st[idx[i]];
    }
#pragma unroll
    for (int i = 0; i < ILP; ++i) {
      cplx<R> y[N];
      y[0] = a[i];
#pragma unroll
      for (int d = 1; d < N; ++d) y[d] = shfl_xor_c(a[i], ws.xmask[d]);
      // x[c] = y[c ^ r]: swap across row bit b wherever r has it set
#pragma unroll
      for (int b = 0; b < K; ++b) {
        const bool sw = (r >> b) & 1;
#pragma unroll
        for (int d = 0; d < N; ++d) {
          if (d & (1 << b)) continue;
          const cplx<R> lo = y[d], hi = y[d | (1 << b)];
          y[d] = sw ? hi : lo;
          y[d | (1 << b)] = sw ? lo : hi;
        }
      }
      cplx<R> acc = cmul(row[0], y[0]);
#pragma unroll
      for (int c = 1; c < N; ++c) acc = cadd(acc, cmul(row[c], y[c]));
      // lanes failing a control below bit 6 write their amplitude back
      // (selected per component: a select of the struct puts a[] in scratch)
      const cplx<R> out = mk<R>(store ? acc.re : a[i].re, store ? acc.im : a[i].im);
      if (c0 + i < nchunks) st[idx[i]] = out;
    }
  }
}

// Two targets whose 4x4 matrix couples each basis state r only with r ^ 3
// (both target bits flipped): SWAP, fSim, GeneralizedfSim, XX-type
// rotations.  One shuffle per amplitude (partner lane l ^ xmask[3]) instead
// of three; each lane sums its two terms in ascending column order (the
// zero terms of the dense row change nothing but the sign of a zero sum).
template <typename R, int ILP>
__global__ void __launch_bounds__(kBlock)
    k_wave_pair(cplx<R>* __restrict__ st, u64 nchunks, Ins ins, WaveSpec<2> ws, const cplx<R>* __restrict__ mat) {
  const int lane = threadIdx.x & 63;
  const int r = (((lane >> ws.tb[0]) & 1) << 1) | ((lane >> ws.tb[1]) & 1);
  const int rp = r ^ 3;
  const cplx<R> cs = mat[r * 4 + r], cp = mat[r * 4 + rp];
  const bool self_first = r < rp;
  const int xm = ws.xmask[3];
  const bool store = (lane & ws.cmask) == ws.cones;
  const u64 wave = (u64)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const u64 nwaves = (u64)gridDim.x * (kBlock / 64);
  for (u64 c0 = wave * ILP; c0 < nchunks; c0 += nwaves * ILP) {  // uniform per wave
    u64 idx[ILP];
    cplx<R> a[ILP];
#pragma unroll
    for (int i = 0; i < ILP; ++i) {
      const u64 c = c0 + i < nchunks ? c0 + i : c0;
      idx[i] = (insert_bits(c, ins) << kWaveBits) | (u64)lane;
      a[i] = st[idx[i]];
    }
#pragma unroll
    for (int i = 0; i < ILP; ++i) {
      const cplx<R> p = shfl_xor_c(a[i], xm);
      const cplx<R> ts = cmul(cs, a[i]), tp = cmul(cp, p);
      const cplx<R> acc = self_first ? cadd(ts, tp) : cadd(tp, ts);
      const cplx<R> out = mk<R>(store ? acc.re : a[i].re, store ? acc.im : a[i].im);
      if (c0 + i < nchunks) st[idx[i]] = out;
    }
  }
}

// ---- three targets on bits {0, 1, 2}, uncontrolled: LDS-staged chunks --
// Each thread of the register kernel would own 8 consecutive amplitudes, so
// its loads stride 128 B across a wave.  Here a block moves CHUNK (1024 for
// complex128, 2048 for complex64) consecutive amplitudes with coalesced
// loads into LDS, each of its CHUNK / 8 threads then takes one
// 8-amplitude group, applies the 8x8 matrix in
// registers (rows summed in ascending column order), writes the group back
// to LDS, and the block stores the chunk coalesced.  LDS slot s lives at
// s ^ ((s >> 3) & 7): both the coalesced and the per-group accesses then hit
// distinct 16-byte bank groups within every 8 lanes.
constexpr int kLow3Chunk = 2048;

template <typename R>
constexpr size_t mat_bytes_of(int n) {
  return (size_t)n * n * sizeof(cplx<R>);
}

__device__ __forceinline__ unsigned low3_swz(unsigned s) { return s ^ ((s >> 3) & 7u); }

template <typename R, int CHUNK>
__global__ void __launch_bounds__(CHUNK / 8)
    k_multi_low3(cplx<R>* __restrict__ st, u64 nchunks, Offs32 roff, const cplx<R>* __restrict__ mat) {
  constexpr int NT = CHUNK / 8;  // one 8-amplitude group per thread
  __shared__ cplx<R> buf[CHUNK];
  const unsigned t = threadIdx.x;
  for (u64 c = blockIdx.x; c < nchunks; c += gridDim.x) {  // uniform per block
    cplx<R>* p = st + c * CHUNK;
#pragma unroll
    for (int i = 0; i < 8; ++i) buf[low3_swz(i * NT + t)] = p[i * NT + t];
    __syncthreads();
    cplx<R> x[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) x[r] = buf[low3_swz(8 * t + (unsigned)roff.o[r])];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      cplx<R> acc = cmul(mat[r * 8], x[0]);
#pragma unroll
      for (int k = 1; k < 8; ++k) acc = cadd(acc, cmul(mat[r * 8 + k], x[k]));
      buf[low3_swz(8 * t + (unsigned)roff.o[r])] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) p[i * NT + t] = buf[low3_swz(i * NT + t)];
    __syncthreads();
  }
}

// Dense target block of a kernel family (the reference's encodings).
template <typename R>
static void dense_one(int kind, const cplx<R>* g, cplx<R>* d) {
  const cplx<R> z = mk<R>(0, 0), o = mk<R>(1, 0);
  switch (kind) {
    case QJ_APPLY_GATE: for (int i = 0; i < 4; ++i) d[i] = g[i]; break;
    case QJ_APPLY_X: d[0] = z; d[1] = o; d[2] = o; d[3] = z; break;
    case QJ_APPLY_Y: d[0] = z; d[1] = mk<R>(0, -1); d[2] = mk<R>(0, 1); d[3] = z; break;
    case QJ_APPLY_Z: d[0] = o; d[1] = z; d[2] = z; d[3] = mk<R>(-1, 0); break;
    default: d[0] = o; d[1] = z; d[2] = z; d[3] = g[0]; break;  // z_pow
  }
}

template <typename R>
static void dense_two(int kind, const cplx<R>* g, cplx<R>* d) {
  for (int i = 0; i < 16; ++i) d[i] = mk<R>(0, 0);
  if (kind == QJ_APPLY_TWO_QUBIT_GATE) {
    for (int i = 0; i < 16; ++i) d[i] = g[i];
  } else if (kind == QJ_APPLY_SWAP) {
    d[0] = d[6] = d[9] = d[15] = mk<R>(1, 0);
  } else {  // fsim: 2x2 on |01>,|10>, phase on |11>
    d[0] = mk<R>(1, 0);
    d[5] = g[0];
    d[6] = g[1];
    d[9] = g[2];
    d[10] = g[3];
    d[15] = g[4];
  }
}

// Applies a dense 2^k block with the in-wave kernel when every target is
// below bit 6 and the state has at least one chunk; returns 1 if handled.
template <typename R>
static int try_wave(cplx<R>* st, int nqubits, const int* tb, int k, const Ins& gate_ins, const cplx<R>* dense,
                    hipStream_t s, int* rc) {
  // three targets: seven shuffles per amplitude cost more than the strided
  // register kernel loses, so only one and two targets come here
  if (k < 1 || k > 2 || nqubits < kWaveBits) return 0;
  for (int u = 0; u < k; ++u)
    if (tb[u] >= kWaveBits) return 0;
  Ins ins;
  std::memset(&ins, 0, sizeof(ins));
  int cmask = 0, cones = 0, nh = 0;
  for (int j = 0; j < gate_ins.npos; ++j) {
    const int p = gate_ins.pos[j];
    bool target = false;
    for (int u = 0; u < k; ++u) target |= (tb[u] == p);
    if (target) continue;
    if (p < kWaveBits) {
      cmask |= 1 << p;
      cones |= 1 << p;
    } else {
      ins.pos[nh++] = p - kWaveBits;
      ins.ones |= u64(1) << (p - kWaveBits);
    }
  }
  ins.npos = nh;
  const u64 nchunks = u64(1) << (nqubits - kWaveBits - nh);
  const int N = 1 << k;
  void* dmat = nullptr;
  int ticket = -1;
  *rc = ring_upload(dense, (size_t)N * N * sizeof(cplx<R>), s, &dmat, &ticket);
  if (*rc) return 1;
  // chunks per wave step: QJ_WAVE_ILP overrides; by default 1 for complex128
  // (16-byte lanes; 5.9 vs 5.6 TB/s with 2 at 30 qubits) and 2 for complex64
  // (8-byte lanes need more loads in flight)
  static const int ilp_env = [] {
    const char* e = std::getenv("QJ_WAVE_ILP");
    const int v = e ? std::atoi(e) : 0;
    return (v == 1 || v == 2 || v == 4) ? v : 0;
  }();
  const int ilp = ilp_env ? ilp_env : (sizeof(R) == 4 ? 2 : 1);
  const u64 waves = (nchunks + ilp - 1) / ilp;
  const u64 grid = grid_for(waves, kBlock / 64);
  const cplx<R>* dm = static_cast<const cplx<R>*>(dmat);
#define QJ_WAVE(KK)                                                                 \
  {                                                                                 \
    WaveSpec<KK> ws;                                                                \
    for (int u = 0; u < KK; ++u) ws.tb[u] = tb[u];                                  \
    for (int d = 0; d < (1 << KK); ++d) {                                           \
      int x = 0;                                                                    \
      for (int u = 0; u < KK; ++u)                                                  \
        if ((d >> (KK - 1 - u)) & 1) x |= 1 << tb[u];                               \
      ws.xmask[d] = x;                                                              \
    }                                                                               \
    ws.cmask = cmask;                                                               \
    ws.cones = cones;                                                               \
    if (ilp == 1) k_wave<R, KK, 1><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(st, nchunks, ins, ws, dm); \
    else if (ilp == 4) k_wave<R, KK, 4><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(st, nchunks, ins, ws, dm); \
    else k_wave<R, KK, 2><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(st, nchunks, ins, ws, dm); \
  }
  // two targets with a matrix that only couples r and r ^ 3: one shuffle
  // (complex128: SWAP / fSim(28, 29) at 30 qubits 5.6 -> 5.9 TB/s; complex64
  // keeps the general kernel, 1% faster there)
  bool pair = k == 2 && sizeof(R) == 8;
  for (int r = 0; pair && r < 4; ++r)
    for (int c = 0; c < 4; ++c)
      if (c != r && c != (r ^ 3) && (dense[r * 4 + c].re != 0 || dense[r * 4 + c].im != 0)) pair = false;
  if (pair) {
    WaveSpec<2> ws;
    ws.tb[0] = tb[0];
    ws.tb[1] = tb[1];
    for (int d = 0; d < 4; ++d) ws.xmask[d] = ((d >> 1) & 1 ? 1 << tb[0] : 0) | (d & 1 ? 1 << tb[1] : 0);
    ws.cmask = cmask;
    ws.cones = cones;
    if (ilp == 1) k_wave_pair<R, 1><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(st, nchunks, ins, ws, dm);
    else if (ilp == 4) k_wave_pair<R, 4><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(st, nchunks, ins, ws, dm);
    else k_wave_pair<R, 2><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(st, nchunks, ins, ws, dm);
  } else if (k == 1) QJ_WAVE(1) else QJ_WAVE(2)
#undef QJ_WAVE
  *rc = check_launch("wave gate launch");
  ring_release(ticket, s);
  return 1;
}

static int build_spec(int nqubits, const int* targets, int ntargets, const int32_t* qubits,
                      int nq, Ins* ins, const char* who) {
  if (nqubits < 1 || nqubits > 62)
    return set_error(QJ_EVALUE, "%s: nqubits=%d out of range", who, nqubits);
  std::vector<int> pos;
  if (qubits && nq > 0) {
    pos.assign(qubits, qubits + nq);
  } else {
    pos.assign(targets, targets + ntargets);
    std::sort(pos.begin(), pos.end());
  }
  if ((int)pos.size() > kMaxPos || (int)pos.size() > nqubits)
    return set_error(QJ_EVALUE, "%s: %zu gate qubits for a %d-qubit vector", who, pos.size(), nqubits);
  for (size_t j = 0; j < pos.size(); ++j) {
    if (pos[j] < 0 || pos[j] >= nqubits)
      return set_error(QJ_EVALUE, "%s: bit position %d out of range [0, %d)", who, pos[j], nqubits);
    if (j && pos[j] <= pos[j - 1])
      return set_error(QJ_EVALUE, "%s: qubits tensor must be sorted and unique", who);
  }
  for (int t = 0; t < ntargets; ++t) {
    if (targets[t] < 0 || targets[t] >= nqubits)
      return set_error(QJ_EVALUE, "%s: target bit %d out of range", who, targets[t]);
    for (int s = 0; s < t; ++s)
      if (targets[s] == targets[t]) return set_error(QJ_EVALUE, "%s: repeated target", who);
    if (std::find(pos.begin(), pos.end(), targets[t]) == pos.end())
      return set_error(QJ_EVALUE, "%s: target bit %d missing from qubits tensor", who, targets[t]);
  }
  ins->npos = (int)pos.size();
  ins->ones = 0;
  for (size_t j = 0; j < pos.size(); ++j) {
    ins->pos[j] = pos[j];
    bool is_target = false;
    for (int t = 0; t < ntargets; ++t) is_target |= (targets[t] == pos[j]);
    if (!is_target) ins->ones |= u64(1) << pos[j];
  }
  for (int j = ins->npos; j < kMaxPos; ++j) ins->pos[j] = 0;
  return QJ_OK;
}

// complex64: amplitudes per lane access of the one- and two-qubit kernels;
// QJ_C64_LANES=1 or 2 forces one for every target position.
// QJ_GATE_BLOCK=128|256|512: workgroup size of the one- and two-target
// strided kernels (default 256; 128 for complex64 one-target gates with one
// amplitude per lane and the target at bit >= 22).
static int gate_block_env() {
  static const int v = [] {
    const char* e = std::getenv("QJ_GATE_BLOCK");
    const int x = e ? std::atoi(e) : 0;
    return (x == 128 || x == 512) ? x : kBlock;
  }();
  return v;
}

static int c64_lanes_env() {
  static const int v = [] {
    const char* e = std::getenv("QJ_C64_LANES");
    const int x = e ? std::atoi(e) : 0;
    return (x == 1 || x == 2) ? x : 0;
  }();
  return v;
}

template <typename R, int KIND, int V>
static void launch_one(cplx<R>* st, u64 ngroups, u64 tmask, const Ins& ins, const M1<R>& g,
                       hipStream_t s, int nt) {
  const u64 grid = grid_for(ngroups, (u64)nt * kIlp);
  if (nt == 128)
    k_one_qubit<R, KIND, V, 128><<<dim3((unsigned)grid), dim3(128), 0, s>>>(st, ngroups, tmask, ins, g);
  else if (nt == 512)
    k_one_qubit<R, KIND, V, 512><<<dim3((unsigned)grid), dim3(512), 0, s>>>(st, ngroups, tmask, ins, g);
  else
    k_one_qubit<R, KIND, V><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(st, ngroups, tmask, ins, g);
}

template <typename R>
static int one_qubit_impl(void* state, int nqubits, int kind, int m, const int32_t* qubits,
                          int nq, const void* gate, hipStream_t s) {
  Ins ins;
  int rc = build_spec(nqubits, &m, 1, qubits, nq, &ins, "apply_one_qubit");
  if (rc) return rc;
  M1<R> g;
  std::memset(&g, 0, sizeof(g));
  if (kind == QJ_APPLY_GATE || kind == QJ_APPLY_Z_POW) {
    if (!gate) return set_error(QJ_EVALUE, "apply_one_qubit: gate matrix required");
    std::memcpy(g.m, gate, (kind == QJ_APPLY_GATE ? 4 : 1) * sizeof(cplx<R>));
  } else if (kind != QJ_APPLY_X && kind != QJ_APPLY_Y && kind != QJ_APPLY_Z) {
    return set_error(QJ_EVALUE, "apply_one_qubit: unknown kind %d", kind);
  }
  {
    cplx<R> dense[4];
    dense_one<R>(kind, g.m, dense);
    int wrc = QJ_OK;
    if (try_wave<R>(static_cast<cplx<R>*>(state), nqubits, &m, 1, ins, dense, s, &wrc)) return wrc;
  }
  const u64 npairs = u64(1) << (nqubits - ins.npos);
  // complex64 lanes move two amplitudes (16 B) when bit 0 is free, except
  // in the kernels that read both halves of each pair (dense, X, Y) with the
  // target at bit >= 22 or <= 10: there one amplitude (8 B) per lane was
  // measured faster at 30 qubits (H on bit 29: 5.2 -> 5.9 TB/s; bits 1-10:
  // +6-8%), while bits 11-21 and the one-sided Z kernels are faster with 16 B
  const bool two_sided = kind != QJ_APPLY_Z && kind != QJ_APPLY_Z_POW;
  const int lanes_env = c64_lanes_env();
  const bool narrow = lanes_env ? lanes_env == 1 : (two_sided && (m >= 22 || m <= 10));
  const bool vec = sizeof(R) == 4 && ins.pos[0] >= 1 && !narrow;
  const u64 ngroups = vec ? npairs / 2 : npairs;
  const u64 tmask = u64(1) << m;
  cplx<R>* st = static_cast<cplx<R>*>(state);
  // workgroup size: 128 threads for complex64 one-amplitude lanes with the
  // target at bit >= 22 (H at bits 24 / 27 / 29 of 30 qubits: +7-14% over
  // 256), else 256 (128 or 512 lost up to 17% elsewhere); QJ_GATE_BLOCK
  // overrides
  static const bool nt_forced = std::getenv("QJ_GATE_BLOCK") != nullptr;
  const int nt = nt_forced ? gate_block_env() : ((sizeof(R) == 4 && !vec && m >= 22) ? 128 : kBlock);
#define QJ_ONE(K)                                                  \
  if (vec) launch_one<R, K, 2>(st, ngroups, tmask, ins, g, s, nt); \
  else launch_one<R, K, 1>(st, ngroups, tmask, ins, g, s, nt);
  switch (kind) {
    case QJ_APPLY_GATE: QJ_ONE(QJ_APPLY_GATE); break;
    case QJ_APPLY_X: QJ_ONE(QJ_APPLY_X); break;
    case QJ_APPLY_Y: QJ_ONE(QJ_APPLY_Y); break;
    case QJ_APPLY_Z: QJ_ONE(QJ_APPLY_Z); break;
    default: QJ_ONE(QJ_APPLY_Z_POW); break;
  }
#undef QJ_ONE
  return check_launch("apply_one_qubit launch");
}

template <typename R, int KIND, int V>
static void launch_two(cplx<R>* st, u64 ngroups, u64 q1, u64 q2, const Ins& ins, const M2<R>& g,
                       hipStream_t s) {
  const int nt = gate_block_env();
  const u64 grid = grid_for(ngroups, (u64)nt * kIlp);
  if (nt == 128)
    k_two_qubit<R, KIND, V, 128><<<dim3((unsigned)grid), dim3(128), 0, s>>>(st, ngroups, q1, q2, ins, g);
  else if (nt == 512)
    k_two_qubit<R, KIND, V, 512><<<dim3((unsigned)grid), dim3(512), 0, s>>>(st, ngroups, q1, q2, ins, g);
  else
    k_two_qubit<R, KIND, V><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(st, ngroups, q1, q2, ins, g);
}

template <typename R>
static int two_qubit_impl(void* state, int nqubits, int kind, int bit1, int bit2,
                          const int32_t* qubits, int nq, const void* gate, hipStream_t s) {
  int tb[2] = {bit1, bit2};
  Ins ins;
  int rc = build_spec(nqubits, tb, 2, qubits, nq, &ins, "apply_two_qubit");
  if (rc) return rc;
  M2<R> g;
  std::memset(&g, 0, sizeof(g));
  if (kind == QJ_APPLY_TWO_QUBIT_GATE || kind == QJ_APPLY_FSIM) {
    if (!gate) return set_error(QJ_EVALUE, "apply_two_qubit: gate matrix required");
    std::memcpy(g.m, gate, (kind == QJ_APPLY_FSIM ? 5 : 16) * sizeof(cplx<R>));
  } else if (kind != QJ_APPLY_SWAP) {
    return set_error(QJ_EVALUE, "apply_two_qubit: unknown kind %d", kind);
  }
  {
    cplx<R> dense[16];
    dense_two<R>(kind, g.m, dense);
    int wrc = QJ_OK;
    if (try_wave<R>(static_cast<cplx<R>*>(state), nqubits, tb, 2, ins, dense, s, &wrc)) return wrc;
  }
  const u64 nquads = u64(1) << (nqubits - ins.npos);
  const bool vec = sizeof(R) == 4 && ins.pos[0] >= 1 && c64_lanes_env() != 1;
  const u64 ngroups = vec ? nquads / 2 : nquads;
  const u64 q1 = u64(1) << bit1, q2 = u64(1) << bit2;
  cplx<R>* st = static_cast<cplx<R>*>(state);
#define QJ_TWO(K)                                                     \
  if (vec) launch_two<R, K, 2>(st, ngroups, q1, q2, ins, g, s);      \
  else launch_two<R, K, 1>(st, ngroups, q1, q2, ins, g, s);
  switch (kind) {
    case QJ_APPLY_TWO_QUBIT_GATE: QJ_TWO(QJ_APPLY_TWO_QUBIT_GATE); break;
    case QJ_APPLY_SWAP: QJ_TWO(QJ_APPLY_SWAP); break;
    default: QJ_TWO(QJ_APPLY_FSIM); break;
  }
#undef QJ_TWO
  return check_launch("apply_two_qubit launch");
}

template <typename R, int K>
static void launch_multi_reg(cplx<R>* st, u64 ngroups, const Ins& ins, const int* tb,
                             const cplx<R>* dmat, hipStream_t s) {
  Offs<K> off;
  for (int r = 0; r < (1 << K); ++r) {
    u64 o = 0;
    for (int u = 0; u < K; ++u)
      if ((r >> (K - 1 - u)) & 1) o |= u64(1) << tb[u];
    off.o[r] = o;
  }
  static const int nt_env = [] {
    const char* e = std::getenv("QJ_MREG_BLOCK");
    const int v = e ? std::atoi(e) : 0;
    return (v == 128 || v == 256 || v == 512) ? v : 0;
  }();
  // Workgroup size of the 3-target kernel by precision and lowest target
  // bit m, measured for targets m..m+2 at 30 qubits (m = 3, 6, ..., 27):
  // complex128 256 threads for 9 <= m <= 21 (128 threads lost 5-18% there)
  // and 128 elsewhere (+2-6%); complex64 128 threads for m >= 9 (+5-30%,
  // -1.5% at m = 24) and 256 below.  4 and 5 targets keep 256.
  int nt = kBlock;
  if (K == 3) {
    int m = tb[0];
    for (int u = 1; u < K; ++u) m = tb[u] < m ? tb[u] : m;
    nt = sizeof(R) == 8 ? ((m >= 9 && m <= 21) ? 256 : 128) : (m >= 9 ? 128 : 256);
  }
  if (nt_env) nt = nt_env;
  const u64 grid = grid_for(ngroups, (u64)nt);
  if (nt == 128)
    k_multi_reg<R, K, 128><<<dim3((unsigned)grid), dim3(128), 0, s>>>(st, ngroups, ins, off, dmat);
  else if (nt == 512)
    k_multi_reg<R, K, 512><<<dim3((unsigned)grid), dim3(512), 0, s>>>(st, ngroups, ins, off, dmat);
  else
    k_multi_reg<R, K><<<dim3((unsigned)grid), dim3(kBlock), 0, s>>>(st, ngroups, ins, off, dmat);
}

template <typename R>
static int multi_qubit_impl(void* state, int nqubits, const int32_t* target_bits, int ntargets,
                            const int32_t* qubits, int nq, const void* gate, hipStream_t s) {
  if (ntargets < 1 || ntargets > 12)
    return set_error(QJ_EVALUE, "Number of target qubits must be <= 12 but is %d.", ntargets);
  if (!gate) return set_error(QJ_EVALUE, "apply_multi_qubit: gate matrix required");
  std::vector<int> tb(target_bits, target_bits + ntargets);
  if (ntargets == 1)
    return one_qubit_impl<R>(state, nqubits, QJ_APPLY_GATE, tb[0], qubits, nq, gate, s);
  if (ntargets == 2)
    return two_qubit_impl<R>(state, nqubits, QJ_APPLY_TWO_QUBIT_GATE, tb[0], tb[1], qubits, nq,
                             gate, s);
  Ins ins;
  int rc = build_spec(nqubits, tb.data(), ntargets, qubits, nq, &ins, "apply_multi_qubit");
  if (rc) return rc;
  const int K = ntargets, N = 1 << K;
  const u64 ngroups = u64(1) << (nqubits - ins.npos);
  cplx<R>* st = static_cast<cplx<R>*>(state);
  if (K == 3 && ins.npos == 3 && nqubits >= 11) {
    bool low = true;
    for (int u = 0; u < 3; ++u) low &= tb[u] < 3;
    if (low) {
      Offs32 roff;
      std::memset(&roff, 0, sizeof(roff));
      for (int r = 0; r < 8; ++r) {
        u64 o = 0;
        for (int u = 0; u < 3; ++u)
          if ((r >> (2 - u)) & 1) o |= u64(1) << tb[u];
        roff.o[r] = o;
      }
      void* dmat = nullptr;
      int ticket = -1;
      rc = ring_upload(gate, mat_bytes_of<R>(8), s, &dmat, &ticket);
      if (rc) return rc;
      static const int chunk_env = [] {
        const char* e = std::getenv("QJ_LOW3_CHUNK");
        const int v = e ? std::atoi(e) : 0;
        return (v == 1024 || v == 2048 || v == 4096) ? v : 0;
      }();
      // 1024 amplitudes per workgroup for complex128 (5.2 -> 5.45 TB/s at
      // 30 qubits; 4096: 5.42), 2048 for complex64 (1024: -3%, 4096: -4%)
      const int dflt = sizeof(R) == 8 ? 1024 : kLow3Chunk;
      const int chunk = (chunk_env && (u64(1) << nqubits) >= (u64)chunk_env) ? chunk_env : dflt;
      const u64 nchunks = (u64(1) << nqubits) / (u64)chunk;
      const cplx<R>* dm = static_cast<const cplx<R>*>(dmat);
      if (chunk == 1024)
        k_multi_low3<R, 1024><<<dim3((unsigned)grid_for(nchunks, 1)), dim3(128), 0, s>>>(st, nchunks, roff, dm);
      else if (chunk == 4096)
        k_multi_low3<R, 4096><<<dim3((unsigned)grid_for(nchunks, 1)), dim3(512), 0, s>>>(st, nchunks, roff, dm);
      else
        k_multi_low3<R, 2048><<<dim3((unsigned)grid_for(nchunks, 1)), dim3(256), 0, s>>>(st, nchunks, roff, dm);
      rc = check_launch("apply_multi_qubit launch");
      ring_release(ticket, s);
      return rc;
    }
  }

  const size_t mat_bytes = (size_t)N * N * sizeof(cplx<R>);
  if (K <= 5) {
    void* dmat = nullptr;
    int ticket = -1;
    rc = ring_upload(gate, mat_bytes, s, &dmat, &ticket);
    if (rc) return rc;
    const cplx<R>* dm = static_cast<const cplx<R>*>(dmat);
    if (K == 3) launch_multi_reg<R, 3>(st, ngroups, ins, tb.data(), dm, s);
    else if (K == 4) launch_multi_reg<R, 4>(st, ngroups, ins, tb.data(), dm, s);
    else launch_multi_reg<R, 5>(st, ngroups, ins, tb.data(), dm, s);
    rc = check_launch("apply_multi_qubit launch");
    ring_release(ticket, s);
    return rc;
  }
  // k > 5: matrix followed by the 2^k offsets in one upload.
  std::vector<char> buf(mat_bytes + (size_t)N * sizeof(u64));
  std::memcpy(buf.data(), gate, mat_bytes);
  u64* offs = reinterpret_cast<u64*>(buf.data() + mat_bytes);
  for (int r = 0; r < N; ++r) {
    u64 o = 0;
    for (int u = 0; u < K; ++u)
      if ((r >> (K - 1 - u)) & 1) o |= u64(1) << tb[u];
    offs[r] = o;
  }
  void* dbuf = nullptr;
  int ticket = -1;
  rc = ring_upload(buf.data(), buf.size(), s, &dbuf, &ticket);
  if (rc) return rc;
  const cplx<R>* dm = static_cast<const cplx<R>*>(dbuf);
  const u64* doffs = reinterpret_cast<const u64*>(static_cast<char*>(dbuf) + mat_bytes);
  const int GP = N >= kBlock ? 1 : kBlock / N;
  const size_t lds = (size_t)GP * N * sizeof(cplx<R>);
  const u64 grid = grid_for(ngroups, GP);
  k_multi_lds<R><<<dim3((unsigned)grid), dim3(kBlock), lds, s>>>(st, ngroups, ins, K, doffs, dm);
  rc = check_launch("apply_multi_qubit launch");
  ring_release(ticket, s);
  return rc;
}

}  // namespace qj

using namespace qj;

extern "C" {

int qj_apply_one_qubit(void* state, int dtype, int nqubits, int kind, int m,
                       const int32_t* qubits, int nqubits_sorted, const void* gate, void* stream) {
  if (!state) return set_error(QJ_EVALUE, "apply_one_qubit: null state");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64)
    return one_qubit_impl<float>(state, nqubits, kind, m, qubits, nqubits_sorted, gate, s);
  if (dtype == QJ_COMPLEX128)
    return one_qubit_impl<double>(state, nqubits, kind, m, qubits, nqubits_sorted, gate, s);
  return set_error(QJ_EVALUE, "apply_one_qubit: state dtype must be complex64/complex128");
}

int qj_apply_two_qubit(void* state, int dtype, int nqubits, int kind, int bit1, int bit2,
                       const int32_t* qubits, int nqubits_sorted, const void* gate, void* stream) {
  if (!state) return set_error(QJ_EVALUE, "apply_two_qubit: null state");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64)
    return two_qubit_impl<float>(state, nqubits, kind, bit1, bit2, qubits, nqubits_sorted, gate, s);
  if (dtype == QJ_COMPLEX128)
    return two_qubit_impl<double>(state, nqubits, kind, bit1, bit2, qubits, nqubits_sorted, gate, s);
  return set_error(QJ_EVALUE, "apply_two_qubit: state dtype must be complex64/complex128");
}

int qj_apply_multi_qubit(void* state, int dtype, int nqubits, const int32_t* target_bits,
                         int ntargets, const int32_t* qubits, int nqubits_sorted, const void* gate,
                         void* stream) {
  if (!state) return set_error(QJ_EVALUE, "apply_multi_qubit: null state");
  if (!target_bits) return set_error(QJ_EVALUE, "apply_multi_qubit: null targets");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == QJ_COMPLEX64)
    return multi_qubit_impl<float>(state, nqubits, target_bits, ntargets, qubits, nqubits_sorted,
                                   gate, s);
  if (dtype == QJ_COMPLEX128)
    return multi_qubit_impl<double>(state, nqubits, target_bits, ntargets, qubits, nqubits_sorted,
                                    gate, s);
  return set_error(QJ_EVALUE, "apply_multi_qubit: state dtype must be complex64/complex128");
}

}  // extern "C"
