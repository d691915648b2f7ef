// Shared device/host helpers for libhfx (CDNA4 / gfx950).
//
// Complex values are stored interleaved (re, im) exactly like numpy/torch
// complex128 / complex64, so device pointers from torch tensors are used as-is.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace hfx {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T re, im;
};

template <typename T> __host__ __device__ __forceinline__ cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> __host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __host__ __device__ __forceinline__ cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T> __host__ __device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
// a * conj(b)
template <typename T> __host__ __device__ __forceinline__ cx<T> mulc(cx<T> a, cx<T> b) {
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
// conj(a) * b
template <typename T> __host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// multiply by -i (fwd=true) or +i
template <typename T> __host__ __device__ __forceinline__ cx<T> rot(cx<T> a, bool minus_i) {
  return minus_i ? cx<T>{a.im, -a.re} : cx<T>{-a.im, a.re};
}
template <typename T> __host__ __device__ __forceinline__ T re_mul(cx<T> a, cx<T> b) {  // Re(a*b)
  return a.re * b.re - a.im * b.im;
}
template <typename T, typename U> __host__ __device__ __forceinline__ cx<T> cvt(cx<U> a) {
  return {static_cast<T>(a.re), static_cast<T>(a.im)};
}

// ---- wavefront / block reductions (wave64) -------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// Deterministic block sum of NV doubles per thread; result valid in thread 0.
// `sh` needs NV * (blockDim.x / 64) doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[q * nw + wid] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += sh[q * nw + w];
      v[q] = s;
    }
  }
  __syncthreads();
}

}  // namespace hfx

// ---- host-side error reporting (definitions in hfx_api.hip) ---------------
namespace hfx {
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);
}  // namespace hfx

#define HFX_OK 0
#define HFX_EINVAL (-1)
#define HFX_EHIP (-2)
#define HFX_EUNSUPPORTED (-3)

#define HFX_HIP_TRY(expr)                                                           \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return hfx::set_error(HFX_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)
