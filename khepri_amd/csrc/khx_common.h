// Shared device helpers of the khx RCWA kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

// ---------------------------------------------------------------- complex
// Interleaved complex double, layout-compatible with numpy complex128.
struct cd {
  double x, y;
};

__host__ __device__ __forceinline__ cd mkc(double a, double b) {
  cd r;
  r.x = a;
  r.y = b;
  return r;
}
__host__ __device__ __forceinline__ cd operator+(cd a, cd b) { return mkc(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ cd operator-(cd a, cd b) { return mkc(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ cd operator-(cd a) { return mkc(-a.x, -a.y); }
__host__ __device__ __forceinline__ cd operator*(cd a, cd b) {
  return mkc(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ cd operator*(double s, cd a) { return mkc(s * a.x, s * a.y); }
__host__ __device__ __forceinline__ cd conjc(cd a) { return mkc(a.x, -a.y); }
__host__ __device__ __forceinline__ double abs2c(cd a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ __forceinline__ double absc(cd a) { return hypot(a.x, a.y); }
__host__ __device__ __forceinline__ double cabs1(cd a) { return fabs(a.x) + fabs(a.y); }

// a*conj(b) and conj(a)*b accumulations written as fused multiply-adds
__host__ __device__ __forceinline__ cd fmac(cd acc, cd a, cd b) {  // acc + a*b
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}
__host__ __device__ __forceinline__ cd fmsc(cd acc, cd a, cd b) {  // acc - a*b
  acc.x = fma(-a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(-a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
  return acc;
}
__host__ __device__ __forceinline__ cd fmac_cj(cd acc, cd a, cd b) {  // acc + conj(a)*b
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
  return acc;
}

// Smith's algorithm (robust complex division).
__host__ __device__ __forceinline__ cd divc(cd a, cd b) {
  if (fabs(b.x) >= fabs(b.y)) {
    if (b.x == 0.0 && b.y == 0.0) return mkc(a.x / fabs(b.x), a.y / fabs(b.x));  // inf/nan like numpy
    double r = b.y / b.x, d = b.x + b.y * r;
    return mkc((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    double r = b.x / b.y, d = b.y + b.x * r;
    return mkc((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
__host__ __device__ __forceinline__ cd recipc(cd b) { return divc(mkc(1.0, 0.0), b); }

// Principal square root (C99 csqrt semantics: the sign of a zero imaginary
// part selects the side of the branch cut).
__host__ __device__ __forceinline__ cd sqrtc(cd z) {
  if (z.x == 0.0 && z.y == 0.0) return mkc(0.0, z.y);
  if (isinf(z.y)) return mkc(INFINITY, z.y);
  double r = hypot(z.x, z.y);
  if (z.x >= 0.0) {
    double t = sqrt(0.5 * (r + z.x));
    return mkc(t, z.y / (2.0 * t));
  }
  double t = sqrt(0.5 * (r - z.x));
  return mkc(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}
// sqrt(z + 0j): a negative zero imaginary part becomes +0 first.
__host__ __device__ __forceinline__ cd sqrtc_p0(cd z) {
  if (z.y == 0.0) z.y = 0.0;
  return sqrtc(z);
}
__host__ __device__ __forceinline__ cd expc(cd z) {
  double e = exp(z.x), s, c;
  sincos(z.y, &s, &c);
  return mkc(e * c, e * s);
}

// ---------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// ---------------------------------------------------------------- host side
namespace khx {
void set_error(const std::string& msg);
const char* last_error();
}  // namespace khx

#define KHX_CHECK_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      khx::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));      \
      return -3;                                                              \
    }                                                                         \
  } while (0)

#define KHX_LAUNCH_CHECK()                                                    \
  do {                                                                        \
    hipError_t e_ = hipGetLastError();                                        \
    if (e_ != hipSuccess) {                                                   \
      khx::set_error(std::string("kernel launch: ") + hipGetErrorString(e_)); \
      return -3;                                                              \
    }                                                                         \
  } while (0)

#define KHX_REQUIRE(cond, msg)        \
  do {                                \
    if (!(cond)) {                    \
      khx::set_error(msg);            \
      return -1;                      \
    }                                 \
  } while (0)

#define KHX_TRY(expr)      \
  do {                     \
    int rc_ = (expr);      \
    if (rc_ != 0) return rc_; \
  } while (0)
