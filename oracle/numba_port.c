/*
 * numba_port.c -- CPU restatement of qibojit's numba kernels.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker that the HIP kernels are
 * compared against, and the CPU baseline that bench.py times.  It is never
 * linked into, loaded by, or called from the product path.
 *
 * Each function restates one numba kernel of the reference with the same
 * index arithmetic, the same one-pass-per-gate structure, the same summation
 * order and the same parallel loop (OpenMP `parallel for` for numba's
 * `prange`).  Complex products use the textbook formula numba emits, and the
 * file is built with -ffp-contract=off so no FMA is formed.
 *
 * Citations are to src/qibojit/custom_operators/{gates,ops}.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float re, im; } c64;
typedef struct { double re, im; } c128;

/* multicontrol_index: gates.py:5-12 */
static inline int64_t multicontrol_index(int64_t g, const int32_t* qubits, int nq) {
  int64_t i = g;
  for (int j = 0; j < nq; ++j) {
    const int n = qubits[j];
    const int64_t k = (int64_t)1 << n;
    i = ((i >> n) << (n + 1)) + (i & (k - 1)) + k;
  }
  return i;
}

/* multitarget_index: gates.py:257-262 */
static inline int64_t multitarget_index(int64_t i, const int64_t* targets, int nt) {
  int64_t t = 0;
  for (int u = 0; u < nt; ++u) t += ((i >> u) & 1) * targets[u];
  return t;
}

#define DEFINE_PORT(T, R, SUF)                                                              \
  static inline T mul_##SUF(T a, T b) {                                                     \
    T c;                                                                                    \
    c.re = a.re * b.re - a.im * b.im;                                                       \
    c.im = a.re * b.im + a.im * b.re;                                                       \
    return c;                                                                               \
  }                                                                                         \
  static inline T add_##SUF(T a, T b) {                                                     \
    T c;                                                                                    \
    c.re = a.re + b.re;                                                                     \
    c.im = a.im + b.im;                                                                     \
    return c;                                                                               \
  }                                                                                         \
  /* apply_gate_kernel: gates.py:15-25 ; multicontrol: gates.py:28-38 */                    \
  void port_apply_gate_##SUF(T* st, const T* g, int64_t nstates, int m, const int32_t* q,  \
                             int nq) {                                                      \
    const int64_t tk = (int64_t)1 << m;                                                     \
    _Pragma("omp parallel for schedule(static)") for (int64_t gi = 0; gi < nstates; ++gi) { \
      int64_t i1, i2;                                                                       \
      if (nq) {                                                                             \
        const int64_t i = multicontrol_index(gi, q, nq);                                   \
        i1 = i - tk;                                                                        \
        i2 = i;                                                                             \
      } else {                                                                              \
        i1 = ((gi >> m) << (m + 1)) + (gi & (tk - 1));                                      \
        i2 = i1 + tk;                                                                       \
      }                                                                                     \
      const T a = st[i1], b = st[i2];                                                       \
      st[i1] = add_##SUF(mul_##SUF(g[0], a), mul_##SUF(g[1], b));                           \
      st[i2] = add_##SUF(mul_##SUF(g[2], a), mul_##SUF(g[3], b));                           \
    }                                                                                       \
  }                                                                                         \
  /* apply_x/y/z/z_pow_kernel and multicontrol variants: gates.py:41-114.                 \
     kind: 1 = x, 2 = y, 3 = z, 4 = z_pow (g[0] is the phase). */                          \
  void port_apply_pauli_##SUF(T* st, const T* g, int64_t nstates, int m, const int32_t* q, \
                              int nq, int kind) {                                           \
    const int64_t tk = (int64_t)1 << m;                                                     \
    _Pragma("omp parallel for schedule(static)") for (int64_t gi = 0; gi < nstates; ++gi) { \
      int64_t i1, i2;                                                                       \
      if (nq) {                                                                             \
        const int64_t i = multicontrol_index(gi, q, nq);                                   \
        i1 = i - tk;                                                                        \
        i2 = i;                                                                             \
      } else {                                                                              \
        i1 = ((gi >> m) << (m + 1)) + (gi & (tk - 1));                                      \
        i2 = i1 + tk;                                                                       \
      }                                                                                     \
      const T a = st[i1], b = st[i2];                                                       \
      if (kind == 1) {                                                                      \
        st[i1] = b;                                                                         \
        st[i2] = a;                                                                         \
      } else if (kind == 2) {                                                               \
        T na, nb;                                                                           \
        na.re = b.im;                                                                       \
        na.im = -b.re; /* -1j * b */                                                        \
        nb.re = -a.im;                                                                      \
        nb.im = a.re; /* 1j * a */                                                          \
        st[i1] = na;                                                                        \
        st[i2] = nb;                                                                        \
      } else if (kind == 3) {                                                               \
        T nb;                                                                               \
        nb.re = -b.re;                                                                      \
        nb.im = -b.im;                                                                      \
        st[i2] = nb;                                                                        \
      } else {                                                                              \
        st[i2] = mul_##SUF(g[0], b);                                                        \
      }                                                                                     \
    }                                                                                       \
  }                                                                                         \
  /* apply_two_qubit_gate_kernel: gates.py:117-153 ; multicontrol: gates.py:156-193;      \
     apply_swap: gates.py:196-216 ; apply_fsim: gates.py:219-254.                          \
     kind: 0 = dense 4x4, 1 = swap, 2 = fsim (5 entries). */                                \
  void port_apply_two_##SUF(T* st, const T* g, int64_t nstates, int m1, int m2,            \
                            int swap_targets, const int32_t* q, int nq, int kind) {         \
    const int64_t tk1 = (int64_t)1 << m1, tk2 = (int64_t)1 << m2;                          \
    int64_t uk1 = tk1, uk2 = tk2;                                                           \
    if (swap_targets) {                                                                     \
      uk1 = tk2;                                                                            \
      uk2 = tk1;                                                                            \
    }                                                                                       \
    _Pragma("omp parallel for schedule(static)") for (int64_t gi = 0; gi < nstates; ++gi) { \
      int64_t i0, i1, i2, i3;                                                               \
      if (nq) {                                                                             \
        const int64_t i = multicontrol_index(gi, q, nq);                                   \
        if (kind == 1) {                                                                    \
          i1 = i - tk2;                                                                     \
          i2 = i - tk1;                                                                     \
        } else {                                                                            \
          i1 = i - uk2;                                                                     \
          i2 = i - uk1;                                                                     \
        }                                                                                   \
        i0 = i1 - uk1;                                                                      \
        i3 = i;                                                                             \
      } else {                                                                              \
        int64_t i = ((gi >> m1) << (m1 + 1)) + (gi & (tk1 - 1));                            \
        i = ((i >> m2) << (m2 + 1)) + (i & (tk2 - 1));                                      \
        if (kind == 1) {                                                                    \
          i1 = i + tk1;                                                                     \
          i2 = i + tk2;                                                                     \
        } else {                                                                            \
          i1 = i + uk1;                                                                     \
          i2 = i + uk2;                                                                     \
        }                                                                                   \
        i0 = i;                                                                             \
        i3 = i + tk1 + tk2;                                                                 \
      }                                                                                     \
      if (kind == 1) {                                                                      \
        const T a = st[i1];                                                                 \
        st[i1] = st[i2];                                                                    \
        st[i2] = a;                                                                         \
      } else if (kind == 2) {                                                               \
        const T a = st[i1], b = st[i2];                                                     \
        st[i1] = add_##SUF(mul_##SUF(g[0], a), mul_##SUF(g[1], b));                         \
        st[i2] = add_##SUF(mul_##SUF(g[2], a), mul_##SUF(g[3], b));                         \
        st[i3] = mul_##SUF(st[i3], g[4]);                                                   \
      } else {                                                                              \
        const T b0 = st[i0], b1 = st[i1], b2 = st[i2], b3 = st[i3];                         \
        T o[4];                                                                             \
        for (int r = 0; r < 4; ++r) {                                                       \
          T acc = mul_##SUF(g[4 * r + 0], b0);                                              \
          acc = add_##SUF(acc, mul_##SUF(g[4 * r + 1], b1));                                \
          acc = add_##SUF(acc, mul_##SUF(g[4 * r + 2], b2));                                \
          acc = add_##SUF(acc, mul_##SUF(g[4 * r + 3], b3));                                \
          o[r] = acc;                                                                       \
        }                                                                                   \
        st[i0] = o[0];                                                                      \
        st[i1] = o[1];                                                                      \
        st[i2] = o[2];                                                                      \
        st[i3] = o[3];                                                                      \
      }                                                                                     \
    }                                                                                       \
  }                                                                                         \
  /* apply_{three,four,five}_qubit_gate_kernel / apply_multi_qubit_gate_kernel:            \
     gates.py:265-424.  `targets` are the int64 strides in reversed target order            \
     (cpu.py:594-596); row i of the matrix acts on ig - multitarget_index(2^k-1-i).         \
     The sum runs over columns left to right, as the unrolled kernels write it. */          \
  void port_apply_multi_##SUF(T* st, const T* g, int64_t nstates, const int64_t* targets,  \
                              int nt, const int32_t* q, int nq) {                           \
    const int ns = 1 << nt;                                                                 \
    _Pragma("omp parallel") {                                                               \
      T* buf = (T*)malloc(sizeof(T) * ns);                                                  \
      _Pragma("omp for schedule(static)") for (int64_t gi = 0; gi < nstates; ++gi) {        \
        const int64_t ig = multicontrol_index(gi, q, nq);                                  \
        for (int i = 0; i < ns; ++i) buf[i] = st[ig - multitarget_index(ns - i - 1, targets, nt)]; \
        for (int i = 0; i < ns; ++i) {                                                      \
          T acc = mul_##SUF(g[(int64_t)i * ns], buf[0]);                                    \
          for (int j = 1; j < ns; ++j) acc = add_##SUF(acc, mul_##SUF(g[(int64_t)i * ns + j], buf[j])); \
          st[ig - multitarget_index(ns - i - 1, targets, nt)] = acc;                        \
        }                                                                                   \
      }                                                                                     \
      free(buf);                                                                            \
    }                                                                                       \
  }                                                                                         \
  /* initial_state_vector: ops.py:10-18 (a density matrix is the same on d*d) */           \
  void port_initial_state_##SUF(T* st, int64_t n) {                                         \
    st[0].re = 1;                                                                           \
    st[0].im = 0;                                                                           \
    _Pragma("omp parallel for schedule(static)") for (int64_t i = 1; i < n; ++i) {         \
      st[i].re = 0;                                                                         \
      st[i].im = 0;                                                                         \
    }                                                                                       \
  }                                                                                         \
  /* collapse_state / collapse_state_normalized: ops.py:43-79.  Returns the norm. */      \
  double port_collapse_##SUF(T* st, const int32_t* qk, int nm, int64_t result, int nqubits, \
                             int normalize) {                                               \
    const int64_t nstates = (int64_t)1 << (nqubits - nm);                                   \
    const int64_t nsub = (int64_t)1 << nm;                                                  \
    double norms = 0.0;                                                                     \
    _Pragma("omp parallel for schedule(static) reduction(+ : norms)") for (int64_t gi = 0;  \
                                                                          gi < nstates;     \
                                                                          ++gi) {           \
      for (int64_t h = 0; h < nsub; ++h) {                                                  \
        int64_t i = gi;                                                                     \
        for (int j = 0; j < nm; ++j) {                                                      \
          const int n = qk[j];                                                              \
          const int64_t k = (int64_t)1 << n;                                                \
          i = ((i >> n) << (n + 1)) + (i & (k - 1)) + ((h >> j) % 2) * k;                   \
        }                                                                                   \
        if (h == result) {                                                                  \
          if (normalize) {                                                                  \
            const double a = hypot((double)st[i].re, (double)st[i].im);                     \
            norms += a * a;                                                                 \
          }                                                                                 \
        } else {                                                                            \
          st[i].re = 0;                                                                     \
          st[i].im = 0;                                                                     \
        }                                                                                   \
      }                                                                                     \
    }                                                                                       \
    if (!normalize) return 0.0;                                                             \
    const double norm = sqrt(norms);                                                        \
    _Pragma("omp parallel for schedule(static)") for (int64_t gi = 0; gi < nstates; ++gi) { \
      int64_t i = gi;                                                                       \
      for (int j = 0; j < nm; ++j) {                                                        \
        const int n = qk[j];                                                                \
        const int64_t k = (int64_t)1 << n;                                                  \
        i = ((i >> n) << (n + 1)) + (i & (k - 1)) + ((result >> j) % 2) * k;                \
      }                                                                                     \
      st[i].re = (R)((double)st[i].re / norm);                                              \
      st[i].im = (R)((double)st[i].im / norm);                                              \
    }                                                                                       \
    return norm;                                                                            \
  }

DEFINE_PORT(c64, float, c64)
DEFINE_PORT(c128, double, c128)

/* ------------------------------------------------------------------------ */
/* MT19937 with NumPy legacy RandomState semantics (numpy/random/src/mt19937) */
typedef struct {
  uint32_t key[624];
  int pos;
} mt_state;

static void mt_seed(mt_state* s, uint32_t seed) {
  for (int i = 0; i < 624; ++i) {
    s->key[i] = seed;
    seed = 1812433253u * (seed ^ (seed >> 30)) + (uint32_t)(i + 1);
  }
  s->pos = 624;
}

static void mt_gen(mt_state* s) {
  uint32_t y;
  int i;
  for (i = 0; i < 624 - 397; ++i) {
    y = (s->key[i] & 0x80000000u) | (s->key[i + 1] & 0x7fffffffu);
    s->key[i] = s->key[i + 397] ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
  }
  for (; i < 623; ++i) {
    y = (s->key[i] & 0x80000000u) | (s->key[i + 1] & 0x7fffffffu);
    s->key[i] = s->key[i + (397 - 624)] ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
  }
  y = (s->key[623] & 0x80000000u) | (s->key[0] & 0x7fffffffu);
  s->key[623] = s->key[396] ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
  s->pos = 0;
}

static uint32_t mt_next32(mt_state* s) {
  if (s->pos == 624) mt_gen(s);
  uint32_t y = s->key[s->pos++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

/* RandomState.random_sample(): 53-bit double */
static double mt_double(mt_state* s) {
  const int32_t a = (int32_t)(mt_next32(s) >> 5), b = (int32_t)(mt_next32(s) >> 6);
  return (a * 67108864.0 + b) / 9007199254740992.0;
}

/* RandomState.randint(low, high) for int64: masked rejection on 32-bit draws
 * when high - low - 1 fits in 32 bits. */
static int64_t mt_randint(mt_state* s, int64_t low, int64_t high) {
  const uint64_t rng = (uint64_t)(high - low - 1);
  if (rng == 0) return low;
  if (rng <= 0xFFFFFFFFull) {
    if (rng == 0xFFFFFFFFull) return low + (int64_t)mt_next32(s);
    uint32_t mask = (uint32_t)rng;
    mask |= mask >> 1; mask |= mask >> 2; mask |= mask >> 4; mask |= mask >> 8; mask |= mask >> 16;
    uint32_t v;
    while ((v = (mt_next32(s) & mask)) > (uint32_t)rng) {
    }
    return low + (int64_t)v;
  }
  uint64_t mask = rng;
  mask |= mask >> 1; mask |= mask >> 2; mask |= mask >> 4; mask |= mask >> 8; mask |= mask >> 16;
  mask |= mask >> 32;
  uint64_t v;
  do {
    const uint64_t hi = mt_next32(s), lo = mt_next32(s);
    v = ((hi << 32) | lo) & mask;
  } while (v > rng);
  return low + (int64_t)v;
}

/* Test hooks: raw streams of the generator, to check it against NumPy. */
void port_mt_words(uint32_t seed, uint32_t* out, int64_t n) {
  mt_state s;
  mt_seed(&s, seed);
  for (int64_t i = 0; i < n; ++i) out[i] = mt_next32(&s);
}
void port_mt_randint(uint32_t seed, int64_t low, int64_t high, int64_t* out, int64_t n) {
  mt_state s;
  mt_seed(&s, seed);
  for (int64_t i = 0; i < n; ++i) out[i] = mt_randint(&s, low, high);
}
void port_mt_random(uint32_t seed, double* out, int64_t n) {
  mt_state s;
  mt_seed(&s, seed);
  for (int64_t i = 0; i < n; ++i) out[i] = mt_double(&s);
}

/* Per-chain seeds: seed(seed); [randint(0, int(1e8)) for _ in range(nthreads)]
 * (ops.py:93-94). */
void port_chain_seeds(uint32_t seed, int nthreads, uint32_t* out) {
  mt_state s;
  mt_seed(&s, seed);
  for (int t = 0; t < nthreads; ++t) out[t] = (uint32_t)mt_randint(&s, 0, 100000000);
}

/* measure_frequencies: ops.py:82-108.  `probs_f32` selects the probs type;
 * the ratio is computed in that type and compared in double with random(). */
#define DEFINE_MF(T, SUF)                                                                     \
  void port_measure_frequencies_##SUF(int64_t* freqs, const T* probs, int64_t nstates,        \
                                      int64_t nshots, uint32_t seed, int nthreads) {          \
    int64_t* thread_nshots = (int64_t*)malloc(sizeof(int64_t) * nthreads);                    \
    for (int t = 0; t < nthreads; ++t) thread_nshots[t] = nshots / nthreads;                   \
    thread_nshots[nthreads - 1] += nshots % nthreads;                                         \
    uint32_t* seeds = (uint32_t*)malloc(sizeof(uint32_t) * nthreads);                        \
    port_chain_seeds(seed, nthreads, seeds);                                                  \
    int64_t start = 0;                                                                        \
    for (int64_t i = 1; i < nstates; ++i)                                                     \
      if (probs[i] > probs[start]) start = i;                                                 \
    int64_t* tf = (int64_t*)calloc((size_t)nthreads * nstates, sizeof(int64_t));              \
    _Pragma("omp parallel for schedule(static, 1)") for (int n = 0; n < nthreads; ++n) {      \
      mt_state s;                                                                             \
      mt_seed(&s, seeds[n]);                                                                  \
      int64_t shot = start;                                                                   \
      int64_t* priv = tf + (size_t)n * nstates;                                               \
      for (int64_t i = 0; i < thread_nshots[n]; ++i) {                                        \
        const int64_t new_shot = (shot + mt_randint(&s, 0, nstates)) % nstates;              \
        const T ratio = probs[new_shot] / probs[shot];                                        \
        if ((double)ratio > mt_double(&s)) shot = new_shot;                                   \
        priv[shot] += 1;                                                                      \
      }                                                                                       \
    }                                                                                         \
    for (int n = 0; n < nthreads; ++n)                                                        \
      for (int64_t i = 0; i < nstates; ++i) freqs[i] += tf[(size_t)n * nstates + i];          \
    free(tf);                                                                                 \
    free(seeds);                                                                              \
    free(thread_nshots);                                                                      \
  }

DEFINE_MF(float, f32)
DEFINE_MF(double, f64)

int port_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void port_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
