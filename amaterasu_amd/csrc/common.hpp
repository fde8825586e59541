// Shared device/host helpers for the amaterasu_amd HIP library (gfx950 / CDNA4).
//
// Complex numbers are interleaved (re, im) pairs: float2 (complex64) or double2
// (complex128), the layout torch.complex64/complex128 tensors use. Every kernel is
// templated on the real type R; Cplx<R>::T is the matching vector type.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace amt {

template <typename R>
struct Cplx;
template <>
struct Cplx<float> {
  using T = float2;
};
template <>
struct Cplx<double> {
  using T = double2;
};

template <typename C>
__host__ __device__ __forceinline__ C cmake(decltype(C::x) re, decltype(C::x) im) {
  C c;
  c.x = re;
  c.y = im;
  return c;
}
template <typename C>
__host__ __device__ __forceinline__ C cadd(C a, C b) {
  return cmake<C>(a.x + b.x, a.y + b.y);
}
template <typename C>
__host__ __device__ __forceinline__ C csub(C a, C b) {
  return cmake<C>(a.x - b.x, a.y - b.y);
}
template <typename C>
__host__ __device__ __forceinline__ C cmul(C a, C b) {
  return cmake<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C>
__host__ __device__ __forceinline__ C cscale(C a, decltype(C::x) s) {
  return cmake<C>(a.x * s, a.y * s);
}
// multiply by SIGN * i  (SIGN = +1 or -1)
template <int SIGN, typename C>
__host__ __device__ __forceinline__ C cmul_i(C a) {
  return SIGN > 0 ? cmake<C>(-a.y, a.x) : cmake<C>(a.y, -a.x);
}

// Wavefront size on CDNA is 64.
constexpr int kWave = 64;

}  // namespace amt
