// Shared device/host definitions for the MI355X AIFMM library.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace rcs {

// Interleaved complex number; R is double (complex128 path) or float (complex64 path).
template <typename R>
struct cpx {
  R re, im;
  cpx() = default;  // trivial, so that LDS arrays of cpx are allowed
  __host__ __device__ cpx(R a, R b = 0) : re(a), im(b) {}
};
template <typename R>
__host__ __device__ inline cpx<R> operator+(cpx<R> a, cpx<R> b) { return cpx<R>(a.re + b.re, a.im + b.im); }
template <typename R>
__host__ __device__ inline cpx<R> operator-(cpx<R> a, cpx<R> b) { return cpx<R>(a.re - b.re, a.im - b.im); }
template <typename R>
__host__ __device__ inline cpx<R> operator-(cpx<R> a) { return cpx<R>(-a.re, -a.im); }
template <typename R>
__host__ __device__ inline cpx<R> operator*(cpx<R> a, cpx<R> b) {
  return cpx<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename R>
__host__ __device__ inline cpx<R> operator*(cpx<R> a, R s) { return cpx<R>(a.re * s, a.im * s); }
template <typename R>
__host__ __device__ inline cpx<R> conj(cpx<R> a) { return cpx<R>(a.re, -a.im); }
template <typename R>
__host__ __device__ inline R norm2(cpx<R> a) { return a.re * a.re + a.im * a.im; }
template <typename R>
__host__ __device__ inline cpx<R> cdiv(cpx<R> a, cpx<R> b) {
  R d = b.re * b.re + b.im * b.im;
  return cpx<R>((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
// a += b*c
template <typename R>
__host__ __device__ inline void cfma(cpx<R>& a, cpx<R> b, cpx<R> c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(-b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(b.im, c.re, a.im);
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(x) ::rcs::hip_check((x), #x)

// Precision codes of the C ABI.
enum Precision { PREC_C128 = 0, PREC_C64 = 1 };
enum KernelKind { KIND_EFIE = 0, KIND_LAPLACE = 1 };

}  // namespace rcs
