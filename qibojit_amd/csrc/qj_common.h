// Shared device/host helpers of the gfx950 state-vector kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/qjhip.h"

namespace qj {

typedef uint64_t u64;
typedef int64_t i64;

constexpr int kBlock = 256;      // threads per workgroup (4 wave64)
constexpr int kMaxPos = 48;      // max bit positions in an insertion spec
constexpr u64 kMaxGrid = 1u << 23;  // grid-stride cap: keeps grid*block < 2^32

// Complex number stored as {re, im}; aligned to its size so that one
// amplitude is one 8 B (complex64) or 16 B (complex128) vector access.
template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R re, im;
};

template <typename R>
__host__ __device__ __forceinline__ cplx<R> mk(R a, R b) {
  cplx<R> c;
  c.re = a;
  c.im = b;
  return c;
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) {
  return mk(a.re + b.re, a.im + b.im);
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// a*x + b*y, summed left to right as in the reference kernels
// (custom_operators/gates.py:21-24).
template <typename R>
__device__ __forceinline__ cplx<R> mac2(cplx<R> a, cplx<R> x, cplx<R> b,
                                        cplx<R> y) {
  return cadd(cmul(a, x), cmul(b, y));
}

// V consecutive amplitudes moved with one vector memory instruction.
template <typename R, int V>
struct alignas(V * 2 * sizeof(R)) cvec {
  cplx<R> v[V];
};

template <typename R, int V>
__device__ __forceinline__ cvec<R, V> ldv(const cplx<R>* p) {
  return *reinterpret_cast<const cvec<R, V>*>(p);
}
template <typename R, int V>
__device__ __forceinline__ void stv(cplx<R>* p, const cvec<R, V>& x) {
  *reinterpret_cast<cvec<R, V>*>(p) = x;
}

// Insertion of fixed bits into a compact index: the positions are ascending;
// each inserts a 0 below which the lower bits of g stay in place.  `ones`
// then sets the bits that must be 1 (controls, or a collapse outcome).  This
// is the bijection of multicontrol_index (custom_operators/gates.py:5-12)
// written for an ascending position list.
struct Ins {
  int npos;
  int pos[kMaxPos];
  u64 ones;
};

__device__ __forceinline__ u64 insert_bits(u64 g, const Ins& s) {
  for (int j = 0; j < s.npos; ++j) {
    const u64 lo = g & ((u64(1) << s.pos[j]) - 1);
    g = ((g ^ lo) << 1) | lo;
  }
  return g | s.ones;
}

// 32 precomputed index offsets (outcome or row r -> amplitude offset).
struct Offs32 {
  u64 o[32];
};

inline u64 grid_for(u64 work_items, u64 per_block) {
  u64 g = (work_items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > kMaxGrid) g = kMaxGrid;
  return g;
}

// ---- host-side error plumbing (qj_runtime.hip) ----
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);
int hip_status(hipError_t e, const char* what);

// Per-device scratch (qj_runtime.hip).
struct DeviceScratch {
  double* partials;       // reduction partials
  size_t partials_cap;    // in doubles
  double* scalars;        // small device scalars (norms, factors)
  void* big;              // growable work buffer (sampler words)
  size_t big_cap;         // in bytes
};
int scratch_for_current_device(DeviceScratch** out);
// Ensures sc->big holds `bytes`; growing it synchronises `stream` first.
int scratch_reserve_big(DeviceScratch* sc, size_t bytes, hipStream_t stream);
// Matrix staging: copies `bytes` from host `src` into a device slot ordered on
// `stream`; returns the device pointer and a ticket to pass to ring_release
// after the consuming kernel has been enqueued.
int ring_upload(const void* src, size_t bytes, hipStream_t stream, void** dev,
                int* ticket);
void ring_release(int ticket, hipStream_t stream);

inline bool is_complex(int dtype) {
  return dtype == QJ_COMPLEX64 || dtype == QJ_COMPLEX128;
}
inline size_t elem_size(int dtype) {
  switch (dtype) {
    case QJ_COMPLEX64: return 8;
    case QJ_COMPLEX128: return 16;
    case QJ_FLOAT32: return 4;
    case QJ_FLOAT64: return 8;
  }
  return 0;
}

}  // namespace qj
