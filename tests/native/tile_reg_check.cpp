// Host check of the register-tile planner (qibojit_amd/csrc/qj_tile_reg.h).
// Random tile-op lists are planned with reg_plan and executed on the CPU with
// the kernel's own per-thread steps (reg_layout, reg_dense_any, reg_phase,
// lds_swz, reg_state_offsets), thread by thread, with the LDS re-layouts
// done for all threads before any thread reads; the result is compared with
// the op semantics of include/qjhip.h applied directly to the state.
// Exit status 0 when every case agrees; prints one line per case.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "../../qibojit_amd/csrc/qj_tile_reg.h"

using namespace qj;

static u64 insert_bits_host(u64 g, const std::vector<int>& pos) {
  for (int p : pos) {
    const u64 lo = g & ((u64(1) << p) - 1);
    g = ((g ^ lo) << 1) | lo;
  }
  return g;
}

template <typename R>
struct Case {
  int nq, L, TB, nrb;
  std::vector<int> gpos;  // tile bit -> state bit
  std::vector<qj_tile_op> ops;
  std::vector<char> data;
};

template <typename R>
static size_t put(std::vector<char>& d, const void* src, size_t bytes) {
  const size_t off = d.size();
  d.resize(off + ((bytes + 15) & ~size_t(15)), 0);
  std::memcpy(d.data() + off, src, bytes);
  return off;
}

template <typename R>
static Case<R> make_case(std::mt19937_64& rng, int nq, int L, int nhigh, int nrb, int nops, int kmax) {
  Case<R> c;
  c.nq = nq;
  c.L = L;
  c.TB = L + nhigh;
  c.nrb = nrb;
  for (int b = 0; b < L; ++b) c.gpos.push_back(b);
  std::vector<int> rest;
  for (int b = L; b < nq; ++b) rest.push_back(b);
  std::shuffle(rest.begin(), rest.end(), rng);
  std::vector<int> high(rest.begin(), rest.begin() + nhigh);
  std::sort(high.begin(), high.end());
  for (int b : high) c.gpos.push_back(b);
  std::vector<int> outside;
  for (int b = 0; b < nq; ++b)
    if (std::find(c.gpos.begin(), c.gpos.end(), b) == c.gpos.end()) outside.push_back(b);
  const int nout = (int)outside.size(), ga = (nout + 1) / 2, gb = nout - ga;
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  auto rc = [&]() { return mk((R)U(rng), (R)U(rng)); };
  auto unit = [&]() {
    const double a = 3.14159265358979 * U(rng);
    return mk((R)std::cos(a), (R)std::sin(a));
  };
  for (int o = 0; o < nops; ++o) {
    qj_tile_op op;
    std::memset(&op, 0, sizeof(op));
    const int kind = (int)(rng() % 3);
    std::vector<int> bits(c.TB);
    for (int b = 0; b < c.TB; ++b) bits[b] = b;
    std::shuffle(bits.begin(), bits.end(), rng);
    if (kind == 2) {
      op.kind = QJ_TILE_SWAP;
      op.k = 2;
      op.target[0] = bits[0];
      op.target[1] = bits[1];
    } else if (kind == 0) {
      op.kind = QJ_TILE_DENSE;
      op.k = 1 + (int)(rng() % kmax);
      for (int u = 0; u < op.k; ++u) op.target[u] = bits[u];
      if (rng() % 3 == 0) op.local_ctrl = u64(1) << bits[op.k];
      if (nout && rng() % 3 == 0) op.global_ctrl = u64(1) << outside[rng() % nout];
      std::vector<cplx<R>> m((size_t)1 << (2 * op.k));
      const int form = op.k == 1 ? (int)(rng() % 3) : 0;  // real and Hadamard 2x2 take their own paths
      for (auto& v : m) v = form == 1 ? mk((R)U(rng), (R)0) : rc();
      if (form == 2) {
        // Hadamard form; deferred when it has no controls (cleared for half of them)
        const R h = (R)(rng() % 2 ? 0.70710678118654752 : -0.5);
        m = {mk(h, (R)0), mk(h, (R)0), mk(h, (R)0), mk(-h, (R)0)};
        if (rng() % 2) op.local_ctrl = op.global_ctrl = 0;
      }
      op.offset = put<R>(c.data, m.data(), m.size() * sizeof(cplx<R>));
    } else {
      op.kind = QJ_TILE_PHASE;
      if (rng() % 2) op.local_ctrl = u64(1) << bits[0];
      if (nout && rng() % 3 == 0) op.global_ctrl = u64(1) << outside[rng() % nout];
      std::vector<cplx<R>> blk(((size_t)1 << ga) + ((size_t)1 << gb) + c.TB);
      for (auto& v : blk) v = unit();
      const int32_t head[4] = {ga, gb, c.TB, 0};
      std::vector<char> raw(16 + blk.size() * sizeof(cplx<R>));
      std::memcpy(raw.data(), head, 16);
      std::memcpy(raw.data() + 16, blk.data(), blk.size() * sizeof(cplx<R>));
      op.offset = put<R>(c.data, raw.data(), raw.size());
    }
    c.ops.push_back(op);
  }
  return c;
}

// The documented semantics, applied op by op to the whole state.
template <typename R>
static void apply_direct(const Case<R>& c, std::vector<std::complex<double>>& psi) {
  const u64 n = psi.size();
  std::vector<int> outside;
  for (int b = 0; b < c.nq; ++b)
    if (std::find(c.gpos.begin(), c.gpos.end(), b) == c.gpos.end()) outside.push_back(b);
  const int nout = (int)outside.size(), ga = (nout + 1) / 2, gb = nout - ga;
  auto local_of = [&](u64 i) {
    int e = 0;
    for (int b = 0; b < c.TB; ++b) e |= (int)((i >> c.gpos[b]) & 1) << b;
    return e;
  };
  auto tile_of = [&](u64 i) {
    u64 t = 0;
    for (int r = 0; r < nout; ++r) t |= ((i >> outside[r]) & 1) << r;
    return t;
  };
  for (const qj_tile_op& op : c.ops) {
    auto ctrl_ok = [&](u64 i) {
      return (i & op.global_ctrl) == op.global_ctrl && ((u64)local_of(i) & op.local_ctrl) == op.local_ctrl;
    };
    if (op.kind == QJ_TILE_SWAP) {
      const u64 ma = u64(1) << c.gpos[op.target[0]], mb = u64(1) << c.gpos[op.target[1]];
      for (u64 i = 0; i < n; ++i)
        if ((i & ma) && !(i & mb)) std::swap(psi[i], psi[i ^ ma ^ mb]);
    } else if (op.kind == QJ_TILE_DENSE) {
      const int k = op.k, N = 1 << k;
      const cplx<R>* m = reinterpret_cast<const cplx<R>*>(c.data.data() + op.offset);
      u64 tmask = 0;
      for (int u = 0; u < k; ++u) tmask |= u64(1) << c.gpos[op.target[u]];
      std::vector<std::complex<double>> in(N);
      for (u64 i = 0; i < n; ++i) {
        if ((i & tmask) || !ctrl_ok(i)) continue;
        std::vector<u64> rows(N);
        for (int r = 0; r < N; ++r) {
          u64 j = i;
          for (int u = 0; u < k; ++u)
            if ((r >> (k - 1 - u)) & 1) j |= u64(1) << c.gpos[op.target[u]];
          rows[r] = j;
          in[r] = psi[j];
        }
        for (int r = 0; r < N; ++r) {
          std::complex<double> acc = 0;
          for (int q = 0; q < N; ++q) acc += std::complex<double>(m[r * N + q].re, m[r * N + q].im) * in[q];
          psi[rows[r]] = acc;
        }
      }
    } else {
      const cplx<R>* oa = reinterpret_cast<const cplx<R>*>(c.data.data() + op.offset + 16);
      const cplx<R>* ob = oa + ((size_t)1 << ga);
      const cplx<R>* w = ob + ((size_t)1 << gb);
      auto cd = [](cplx<R> v) { return std::complex<double>(v.re, v.im); };
      for (u64 i = 0; i < n; ++i) {
        if (!ctrl_ok(i)) continue;
        const u64 t = tile_of(i);
        const int e = local_of(i);
        std::complex<double> f = cd(oa[t & ((u64(1) << ga) - 1)]) * cd(ob[t >> ga]);
        for (int b = 0; b < c.TB; ++b)
          if ((e >> b) & 1) f *= cd(w[b]);
        psi[i] *= f;
      }
    }
  }
}

// The register kernel, executed thread by thread with its own steps.
template <typename R, int NRB>
static bool run_kernel(const Case<R>& c, const std::vector<char>& blob, std::vector<cplx<R>>& st) {
  constexpr int NR = 1 << NRB;
  const RegHead& h = *reinterpret_cast<const RegHead*>(blob.data());
  const char* lp = blob.data() + h.lds_off;
  const RegSeg* segs = reinterpret_cast<const RegSeg*>(lp + h.segs_off);
  const RegOp* ops = reinterpret_cast<const RegOp*>(lp + h.ops_off);
  const u64* ftab = reinterpret_cast<const u64*>(lp + h.ftab_off);
  const char* ldata = lp + h.ldata_off;
  const int NT = h.NT, nthr = 1 << NT, nseg = h.nseg;
  std::vector<int> tpos(c.gpos.begin(), c.gpos.end());
  std::sort(tpos.begin(), tpos.end());
  const u64 ntiles = u64(1) << (c.nq - c.TB);
  std::vector<cplx<R>> lds((size_t)1 << c.TB);
  std::vector<int> hits((size_t)1 << c.TB);
  std::vector<int> written(st.size(), 0);
  std::vector<std::array<cplx<R>, NR>> x(nthr);
  std::vector<RegLay<NRB>> cur(nthr);
  for (u64 t = 0; t < ntiles; ++t) {
    const u64 base = insert_bits_host(t, tpos);
    std::vector<cplx<R>> f(h.nphase);
    for (int j = 0; j < h.nphase; ++j)
      f[j] = reg_phase_outside<R>(reinterpret_cast<const cplx<R>*>(blob.data() + ftab[j]), h.ga, t);
    for (int tid = 0; tid < nthr; ++tid) {
      u64 thr, rb[NRB];
      reg_state_offsets<NRB>(segs[0], h.gbit_ld, NT, tid, thr, rb);
      for (int r = 0; r < NR; ++r) x[tid][r] = st[base | thr | reg_state_offset<NRB>(rb, r)];
      cur[tid] = reg_layout<NRB>(segs[0], NT, tid);
    }
    for (int s = 0; s < nseg; ++s) {
      const RegSeg& sg = segs[s];
      if (s > 0) {
        std::fill(hits.begin(), hits.end(), 0);
        for (int tid = 0; tid < nthr; ++tid)
          for (int r = 0; r < NR; ++r) {
            const int a = reg_lds<R, NRB>(lds_swz<R>(cur[tid].ethr), cur[tid], r);
            if (a != lds_swz<R>(reg_elem<NRB>(cur[tid], r))) return false;  // linearity of the swizzle
            // the planner's per-register byte offsets give the same address
            if (h.ub_off) {
              const uint16_t* ub = reinterpret_cast<const uint16_t*>(blob.data() + h.lds_off + h.ub_off);
              if ((lds_swz<R>(cur[tid].ethr) * (int)sizeof(cplx<R>) ^ ub[(s - 1) * NR + r]) !=
                  a * (int)sizeof(cplx<R>))
                return false;
            }
            lds[a] = x[tid][r];
            ++hits[a];
          }
        for (int v : hits)
          if (v != 1) return false;  // a re-layout must write every LDS slot once
        for (int tid = 0; tid < nthr; ++tid) {
          cur[tid] = reg_layout<NRB>(sg, NT, tid);
          for (int r = 0; r < NR; ++r) x[tid][r] = lds[reg_lds<R, NRB>(lds_swz<R>(cur[tid].ethr), cur[tid], r)];
        }
      }
      for (int o = sg.o0; o < sg.o1; ++o) {
        const RegOp& op = ops[o];
        if ((base & op.gctrl) != op.gctrl) continue;
        const cplx<R>* d = reinterpret_cast<const cplx<R>*>(ldata + op.loff);
        for (int tid = 0; tid < nthr; ++tid) {
          if (((uint32_t)cur[tid].ethr & op.lct) != op.lct) continue;
          cplx<R>(&xr)[NR] = *reinterpret_cast<cplx<R>(*)[NR]>(x[tid].data());
          if (op.kind == kRegPhase)
            reg_phase<R, NRB>(xr, f[op.fidx], d, NT, tid, op.rcm);
          else
            reg_dense_any<R, NRB>(xr, op.kind, op.mask, d, op.rcm);
        }
      }
    }
    for (int tid = 0; tid < nthr; ++tid) {
      u64 thr, rb[NRB];
      reg_state_offsets<NRB>(segs[nseg - 1], h.gbit_st, NT, tid, thr, rb);
      for (int r = 0; r < NR; ++r) {
        const u64 i = base | thr | reg_state_offset<NRB>(rb, r);
        if (h.nhad) {
          x[tid][r].re *= (R)h.had_scale;
          x[tid][r].im *= (R)h.had_scale;
        }
        st[i] = x[tid][r];
        ++written[i];
      }
    }
  }
  for (int v : written)
    if (v != 1) return false;  // every amplitude stored exactly once
  return true;
}

template <typename R>
static int check(std::mt19937_64& rng, int nq, int L, int nhigh, int nrb, int nops, int kmax, double tol) {
  Case<R> c = make_case<R>(rng, nq, L, nhigh, nrb, nops, kmax);
  RegPlanIn in;
  std::memset(&in, 0, sizeof(in));
  in.TB = c.TB;
  in.NRB = nrb;
  in.S = sizeof(R) == 4 ? 5 : 4;
  in.esize = (int)sizeof(cplx<R>);
  const int nout = nq - c.TB;
  in.ga = (nout + 1) / 2;
  in.gb = nout - in.ga;
  for (int b = 0; b < c.TB; ++b) in.gpos[b] = c.gpos[b];
  in.ops = c.ops.data();
  in.nops = (int)c.ops.size();
  in.data = c.data.data();
  in.lds_cap = u64(1) << 30;
  std::vector<char> blob;
  if (!reg_plan(in, blob)) {
    std::printf("FAIL plan rejected nq=%d TB=%d nrb=%d\n", nq, c.TB, nrb);
    return 1;
  }
  const RegHead& h = *reinterpret_cast<const RegHead*>(blob.data());
  std::normal_distribution<double> N(0.0, 1.0);
  std::vector<std::complex<double>> psi((size_t)1 << nq);
  std::vector<cplx<R>> st(psi.size());
  for (size_t i = 0; i < psi.size(); ++i) {
    st[i] = mk((R)N(rng), (R)N(rng));
    psi[i] = std::complex<double>(st[i].re, st[i].im);
  }
  apply_direct<R>(c, psi);
  bool ok;
  if (nrb == 3)
    ok = run_kernel<R, 3>(c, blob, st);
  else if (nrb == 4)
    ok = run_kernel<R, 4>(c, blob, st);
  else
    ok = run_kernel<R, 5>(c, blob, st);
  double err = 0, scale = 0;
  for (size_t i = 0; i < psi.size(); ++i) {
    err = std::max(err, std::abs(std::complex<double>(st[i].re, st[i].im) - psi[i]));
    scale = std::max(scale, std::abs(psi[i]));
  }
  const bool pass = ok && err <= tol * scale;
  std::printf("%s %s nq=%d TB=%d nrb=%d ops=%d segs=%d deferred_h=%d rel_err=%.3e\n", pass ? "ok" : "FAIL",
              sizeof(R) == 4 ? "c64" : "c128", nq, c.TB, nrb, nops, h.nseg, h.nhad, err / scale);
  return pass ? 0 : 1;
}

int main() {
  std::mt19937_64 rng(20241202);
  int bad = 0;
  for (int rep = 0; rep < 6; ++rep) {
    bad += check<double>(rng, 12, 5, 7, 4, 24, 4, 1e-12);
    bad += check<double>(rng, 13, 5, 7, 3, 24, 3, 1e-12);
    bad += check<float>(rng, 14, 6, 7, 4, 24, 4, 1e-4);
    bad += check<float>(rng, 13, 6, 7, 5, 16, 5, 1e-4);
    bad += check<double>(rng, 10, 5, 5, 4, 20, 4, 1e-12);
    bad += check<float>(rng, 11, 6, 4, 4, 20, 4, 1e-4);
    bad += check<float>(rng, 14, 3, 10, 4, 24, 4, 1e-4);  // ten high bits, 64-byte runs
  }
  return bad ? 1 : 0;
}
