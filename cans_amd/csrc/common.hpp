// Shared definitions for the CDNA4 (gfx950) kernels of the pressure-correction step.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

#include "cans_hip.h"

namespace cans {

// Halo'd field layout: element (i,j,k), i in [0,n1+1], j in [0,n2+1], k in [0,n3+1],
// lives at origin[i + sy*j + sz*k].  sy >= n1+2, sz >= sy*(n2+2).
struct Layout {
  int n1, n2, n3;
  long sy, sz;
  __host__ __device__ long at(int i, int j, int k) const { return i + sy * j + sz * k; }
  __host__ __device__ long cells() const { return (long)n1 * n2 * n3; }
};

// aligned to its full size so that LDS and global accesses move it as one 16-byte
// (double) or 8-byte (float) operation
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
  T re, im;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {  // a * conj(b)
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.re - b.re, a.im - b.im}; }

// error reporting (thread-local message, returned by cans_last_error)
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define CANS_HIP_TRY(expr)                                                               \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::cans::fail(CANS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CANS_LAUNCH_CHECK(name)                                                                  \
  do {                                                                                           \
    hipError_t e_ = hipGetLastError();                                                           \
    if (e_ != hipSuccess) return ::cans::fail(CANS_EHIP, std::string(name) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Run CALL with T bound to the field storage type of context c: double (CANS_FP64)
// or float (CANS_MIXED).  Arithmetic stays double in both; only storage differs.
#define CANS_DISPATCH(c, ...)                                                  \
  ((c)->precision == CANS_MIXED ? [&]() { using T = float; return __VA_ARGS__; }() \
                                : [&]() { using T = double; return __VA_ARGS__; }())

}  // namespace cans

// opaque handles
constexpr int DSCAL_SCAL = 16;  // first device-scalar slot of the scalars' bulk forcing (CANS_MAX_SCALARS of them)

struct cans_ctx {
  cans::Layout L;
  int precision;          // CANS_FP64, or CANS_MIXED (pressure-like arrays stored as float)
  hipStream_t stream;
  int device;
  // scratch owned by the context (grown on demand, freed at destroy):
  // 0 RK intermediate fields, 1 transform work buffer, 2 Thomas checkpoints,
  // 3 small z tables, 4 divergence of the old velocity (mixed mode), 5 Helmholtz z tables,
  // 6 distributed z solve (factors and reduced systems), 7 refinement corrections (mixed mode),
  // 8 the second set of RK intermediate fields, 9 the second pressure field, 10 the folded
  // projection's ghost values (cans_rk3_step) and 11 the input copy of a dense y transform
  static constexpr int NWORK = 12;
  void* work[NWORK] = {};
  size_t work_bytes[NWORK] = {};
  double* dscal = nullptr;     // device scalars (64): [0..2] forcing f, [8..11] reductions,
                               // [DSCAL_SCAL + i] forcing of scalar i
  double* partials = nullptr;  // per-block reduction partials
  size_t partials_n = 0;
  bool rk_first = true;        // the next rk call is the first on this context
  int imp = 0;                 // &numerics: 0 explicit diffusion, 1 is_impdiff, 2 is_impdiff_1d
  int dtdma = 0;               // &numerics is_poisson_dtdma: z slabs of the distributed solve (<= 1: off)
  int refine_sweeps = 0;       // refinement sweeps of the fused mixed substep's pressure solve
  int mixed_incr = 0;          // mixed mode: the RK intermediate velocity as FP32 increments (else FP64)
  int refine_mask = 0;         // substeps (bit irk-1) that refine
};
