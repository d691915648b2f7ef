// QR retraction (solver.py:108-119): U = Q of I - tau G, phases normalised
// so that R has a positive real diagonal.  That Q is unique for a nonsingular
// I - tau G (A = QR = (Q D)(D^* R) for any diagonal unitary D), so it does not
// depend on how the factorisation is computed: here modified Gram-Schmidt
// with one full re-orthogonalisation pass per column ("twice is enough"),
// whose R diagonal is the column norm and so already positive.  I - tau G has singular
// values in [1, sqrt(1 + tau^2 |G|^2)] for skew-Hermitian G, the solvers'
// case, so the orthogonality of Q is at rounding level.
//
// Layout: NB lanes per matrix (NB = 4, 8, 16 or 32; a smaller nb is embedded in
// the top-left block of an NB x NB identity, whose QR is block-diagonal with
// the identity, exactly), lane r holding row r of every column in registers;
// 64 / NB matrices per wavefront, four wavefronts per block.  The column dot
// products are butterfly sums over the NB lanes of a matrix.
#include "hfx_common.h"
#include "hfx_internal.h"

namespace hfx {

template <int NB>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = NB / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <int NB>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int o = NB / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int QR_BLOCK = 256;

template <typename T, int NB>
__global__ __launch_bounds__(QR_BLOCK) void qr_retract_kernel(int nb, size_t nk, const cx<T>* __restrict__ G,
                                                               const double* __restrict__ tau,
                                                               cx<T>* __restrict__ U) {
  const int r = threadIdx.x % NB;
  const size_t k = ((size_t)blockIdx.x * QR_BLOCK + threadIdx.x) / NB;
  const bool live = k < nk;   // whole groups are live or not: no early exit (the sums shuffle)
  const size_t nb2 = (size_t)nb * nb;
  cx<T> q[NB];
  const T t = live ? (T)tau[k] : (T)0;
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    cx<T> g = cx<T>{(T)0, (T)0};
    if (live && r < nb && c < nb) g = G[k * nb2 + (size_t)r * nb + c];
    q[c] = cx<T>{(r == c ? (T)1 : (T)0) - t * g.re, -t * g.im};   // I - tau G
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    cx<T> v = q[j];
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
      for (int i = 0; i < NB; ++i) {   // v -= q_i (q_i^H v), one column at a time
        if (i < j) {
          const cx<T> p = cmul(q[i], v);
          const cx<T> ri{group_sum<NB>(p.re), group_sum<NB>(p.im)};
          v = v - q[i] * ri;
        }
      }
    }
    const T nrm = sqrt(group_sum<NB>(v.re * v.re + v.im * v.im));
    const T inv = nrm > (T)0 ? (T)1 / nrm : (T)1;
    q[j] = inv * v;
  }
  if (live && r < nb) {
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c < nb) U[k * nb2 + (size_t)r * nb + c] = q[c];
  }
}

template <typename T, int NB>
static void qr_launch(int nb, size_t nk, const void* G, const double* tau, void* U, hipStream_t s) {
  const size_t per_block = QR_BLOCK / NB;
  const size_t blocks = (nk + per_block - 1) / per_block;
  hipLaunchKernelGGL((qr_retract_kernel<T, NB>), dim3((unsigned)blocks), dim3(QR_BLOCK), 0, s, nb, nk,
                     (const cx<T>*)G, tau, (cx<T>*)U);
}

}  // namespace hfx

using namespace hfx;

extern "C" {

int hfx_qr_retract(int dtype, int nb, size_t nk, const void* G, const double* tau, void* U, void* stream) {
  if (dtype != HFX_DTYPE_C128 && dtype != HFX_DTYPE_C64)
    return set_error(HFX_EINVAL, "hfx_qr_retract: dtype must be HFX_C128 or HFX_C64");
  if (nb < 1 || nb > 32) return set_error(HFX_EUNSUPPORTED, "hfx_qr_retract: 1 <= nb <= 32 (got %d)", nb);
  if (!G || !tau || !U) return set_error(HFX_EINVAL, "hfx_qr_retract: NULL argument");
  if (nk == 0) return HFX_OK;
  if (nk > ((size_t)1 << 34)) return set_error(HFX_EINVAL, "hfx_qr_retract: nk=%zu too large", nk);
  hipStream_t s = (hipStream_t)stream;
  const bool d = dtype == HFX_DTYPE_C128;
  if (nb <= 4) d ? qr_launch<double, 4>(nb, nk, G, tau, U, s) : qr_launch<float, 4>(nb, nk, G, tau, U, s);
  else if (nb <= 8) d ? qr_launch<double, 8>(nb, nk, G, tau, U, s) : qr_launch<float, 8>(nb, nk, G, tau, U, s);
  else if (nb <= 16) d ? qr_launch<double, 16>(nb, nk, G, tau, U, s) : qr_launch<float, 16>(nb, nk, G, tau, U, s);
  else d ? qr_launch<double, 32>(nb, nk, G, tau, U, s) : qr_launch<float, 32>(nb, nk, G, tau, U, s);
  return check_launch("qr_retract_kernel");
}

}  // extern "C"
