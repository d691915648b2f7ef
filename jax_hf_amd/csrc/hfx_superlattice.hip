// Superlattice (moire) Fock build on the extended momentum grid
// (superlattice.py:167-377).
//
// The exchange of a plane-wave basis indexed by reciprocal vectors G is a 2D
// convolution in absolute momentum, one independent convolution per
// (dG = G_a - G_b, alpha, beta) channel.  The GPU form runs every channel at
// once instead of streaming them:
//   1. sl_scatter_kernel writes the slot planes X[d][a,b](x, y) of the extended
//      grid: the (nkx, nky) block of (P - refP)[(i,a), (j,b)] sits at tile
//      g_off[i] of plane d = dG(i, j); everything else is zero (filled once by
//      the caller, never written by a build).
//   2. the planar exchange of the plan (hfx_selfenergy_planar, one shared
//      kernel plane V_lag_fft) turns each plane into -ifft2(fft2(X) V), written
//      to a second set of planes.
//   3. sl_gather_kernel reads every (k, p, q) of Sigma back from its plane and
//      tile, adds the k-independent Hartree shift and forms F = herm(h + Sigma)
//      in the same pass.
// The Hartree shift (superlattice.py:322-357) needs one small launch before
// the gather: the weighted k sums of the orbital-diagonal blocks summed over
// the pairs of each dG, R(d, c); the gather contracts them with the
// orbital-resolved Coulomb table per entry.
#include "hfx_common.h"
#include "hfx_internal.h"

namespace hfx {

constexpr int SL_BLOCK = 256;
constexpr int SL_MAX_BLOCKS = 4096;

static int sl_grid(size_t n) {
  size_t g = (n + SL_BLOCK - 1) / SL_BLOCK;
  if (g > SL_MAX_BLOCKS) g = SL_MAX_BLOCKS;
  return (int)(g < 1 ? 1 : g);
}

struct SlGeom {
  int nkx, nky, nG, dorb, ndelta, Nx, Ny;
};

// Writes only the pair blocks: one workgroup row per (G_i, G_j) pair
// (blockIdx.y = CSR position s), threads over (channel, k) with k fastest so
// the stores run along the extended-grid rows.  The rest of every plane is
// zero and is never written (the exchange reads planes and writes a second
// buffer), so the caller zero-fills planes once.
template <typename T>
__global__ void sl_scatter_kernel(SlGeom g, const cx<T>* __restrict__ P, const cx<T>* __restrict__ refP,
                                  const int* __restrict__ di, const int* __restrict__ dj,
                                  const int* __restrict__ p2d, const int* __restrict__ goff,
                                  cx<T>* __restrict__ planes) {
  const int D = g.nG * g.dorb, nch = g.dorb * g.dorb, nk = g.nkx * g.nky;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nch * nk) return;
  const int s = blockIdx.y, i = di[s], j = dj[s], d = p2d[i * g.nG + j];
  const int t = u / nk, k = u - t * nk, a = t / g.dorb, b = t - a * g.dorb;
  const int kx = k / g.nky, ky = k - kx * g.nky;
  const size_t src = ((size_t)k * D + i * g.dorb + a) * D + j * g.dorb + b;
  cx<T> v = P[src];
  if (refP) v = v - refP[src];
  const int x = g.nkx * goff[2 * i] + kx, y = g.nky * goff[2 * i + 1] + ky;
  planes[(((size_t)d * nch + t) * g.Nx + x) * g.Ny + y] = v;
}

// R[d dorb + c] = sum over the (A, B) pairs of dG_d and over k of
// w[k] (P - refP)[k, (A, c), (B, c)]: one workgroup per (d, c).
template <typename T>
__global__ void sl_hartree_moments_kernel(int nk, int nG, int dorb, const cx<T>* __restrict__ P,
                                          const cx<T>* __restrict__ refP, const double* __restrict__ w,
                                          const int* __restrict__ dstart, const int* __restrict__ di,
                                          const int* __restrict__ dj, cx<double>* __restrict__ R) {
  const int D = nG * dorb, d = blockIdx.x / dorb, c = blockIdx.x - d * dorb;
  const int s0 = dstart[d], npair = dstart[d + 1] - s0;
  double re = 0.0, im = 0.0;
  for (int u = threadIdx.x; u < npair * nk; u += blockDim.x) {
    const int s = s0 + u / nk, k = u % nk;
    const size_t off = (size_t)k * D * D + (size_t)(di[s] * dorb + c) * D + dj[s] * dorb + c;
    cx<T> v = P[off];
    if (refP) v = v - refP[off];
    re += w[k] * (double)v.re;
    im += w[k] * (double)v.im;
  }
  __shared__ double sh[2 * (SL_BLOCK / 64)];
  re = wave_sum(re);
  im = wave_sum(im);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) {
    sh[wid] = re;
    sh[nw + wid] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sr = 0.0, si = 0.0;
    for (int q = 0; q < nw; ++q) {
      sr += sh[q];
      si += sh[nw + q];
    }
    R[blockIdx.x] = cx<double>{sr, si};
  }
}

struct SlHartree {
  const cx<double>* R;   // (n_delta, dorb) or NULL
  const double* HH;      // (nG, nG, dorb, dorb)
  double deg;
};

// sigma_H[(A, a), (B, b)] = delta_ab deg sum_c HH[A, B, a, c] R[dG(A, B), c]
__device__ __forceinline__ cx<double> sl_hartree_at(const SlGeom& g, const SlHartree& hs, const int* p2d, int A,
                                                    int a, int B, int b) {
  double re = 0.0, im = 0.0;
  if (a == b) {
    const cx<double>* r = hs.R + p2d[A * g.nG + B] * g.dorb;
    const double* hh = hs.HH + ((size_t)(A * g.nG + B) * g.dorb + a) * g.dorb;
    for (int c = 0; c < g.dorb; ++c) {
      re += hh[c] * r[c].re;
      im += hh[c] * r[c].im;
    }
  }
  return cx<double>{hs.deg * re, hs.deg * im};
}

template <typename T>
__device__ __forceinline__ cx<T> sl_plane_at(const SlGeom& g, const cx<T>* planes, const int* p2d, const int* goff,
                                             int i, int a, int j, int b, int kx, int ky) {
  const int d = p2d[i * g.nG + j];
  const int x = g.nkx * goff[2 * i] + kx, y = g.nky * goff[2 * i + 1] + ky;
  return planes[(((size_t)d * g.dorb * g.dorb + a * g.dorb + b) * g.Nx + x) * g.Ny + y];
}

// blockIdx.y = k, threads over the D x D entries of that k-point.
template <typename T>
__global__ void sl_gather_kernel(SlGeom g, const cx<T>* __restrict__ planes, const int* __restrict__ p2d,
                                 const int* __restrict__ goff, SlHartree hs, const cx<T>* __restrict__ h,
                                 cx<T>* __restrict__ Sigma, cx<T>* __restrict__ F) {
  const int D = g.nG * g.dorb, DD = D * D;
  const int pq = blockIdx.x * blockDim.x + threadIdx.x;
  if (pq >= DD) return;
  const int k = blockIdx.y, kx = k / g.nky, ky = k - kx * g.nky;
  const int p = pq / D, q = pq - p * D;
  const int i = p / g.dorb, a = p - i * g.dorb, j = q / g.dorb, b = q - j * g.dorb;
  cx<T> s{0, 0}, st{0, 0};
  if (planes) {
    s = sl_plane_at(g, planes, p2d, goff, i, a, j, b, kx, ky);
    st = sl_plane_at(g, planes, p2d, goff, j, b, i, a, kx, ky);
  }
  if (hs.R) {
    s = s + cvt<T>(sl_hartree_at(g, hs, p2d, i, a, j, b));
    st = st + cvt<T>(sl_hartree_at(g, hs, p2d, j, b, i, a));
  }
  const size_t base = (size_t)k * DD;
  Sigma[base + pq] = s;
  const cx<T> f = h[base + pq] + s, ft = h[base + (size_t)q * D + p] + st;
  F[base + pq] = cx<T>{(T)0.5 * (f.re + ft.re), (T)0.5 * (f.im - ft.im)};
}

template <typename T>
static int sl_build(const SlGeom& g, const void* P, const void* refP, const int* p2d, const int* goff, const hfx_plan* plan, const void* Vlag, void* planes,
                    void* splanes, int hartree, const double* w, const int* dstart, const int* di, const int* dj,
                    const double* HH, double deg, void* R, const void* h, void* Sigma, void* F, hipStream_t s) {
  const int D = g.nG * g.dorb, nk = g.nkx * g.nky;
  if (planes) {
    const dim3 grid((g.dorb * g.dorb * nk + SL_BLOCK - 1) / SL_BLOCK, g.nG * g.nG);
    hipLaunchKernelGGL(sl_scatter_kernel<T>, grid, dim3(SL_BLOCK), 0, s, g, (const cx<T>*)P, (const cx<T>*)refP, di,
                       dj, p2d, goff, (cx<T>*)planes);
    int rc = check_launch("sl_scatter_kernel");
    if (rc) return rc;
    rc = hfx_selfenergy_planar(plan, g.ndelta, planes, Vlag, 0, nullptr, nullptr, g.dorb * g.dorb, 0, 0, nullptr,
                               splanes, nullptr, (void*)s);
    if (rc) return rc;
  }
  if (hartree) {
    hipLaunchKernelGGL(sl_hartree_moments_kernel<T>, dim3(g.ndelta * g.dorb), dim3(SL_BLOCK), 0, s, nk, g.nG,
                       g.dorb, (const cx<T>*)P, (const cx<T>*)refP, w, dstart, di, dj, (cx<double>*)R);
    int rc = check_launch("sl_hartree_moments_kernel");
    if (rc) return rc;
  }
  const SlHartree hs{hartree ? (const cx<double>*)R : nullptr, HH, deg};
  const dim3 grid((D * D + SL_BLOCK - 1) / SL_BLOCK, nk);
  hipLaunchKernelGGL(sl_gather_kernel<T>, grid, dim3(SL_BLOCK), 0, s, g, planes ? (const cx<T>*)splanes : nullptr,
                     p2d, goff, hs, (const cx<T>*)h, (cx<T>*)Sigma, (cx<T>*)F);
  return check_launch("sl_gather_kernel");
}

}  // namespace hfx

using namespace hfx;

extern "C" {

int hfx_superlattice_fock(int dtype, int nkx, int nky, int n_G, int dim_orb, int n_delta, int Nx, int Ny,
                          const void* P, const void* refP, const int* pair_to_delta, const int* g_off, const hfx_plan* plan, const void* V_lag_fft,
                          void* planes, void* sigma_planes, int hartree, const double* w2d,
                          const int* delta_start, const int* delta_i, const int* delta_j, const double* HH_orb,
                          double degeneracy, void* delta_scratch, const void* h, void* Sigma, void* F,
                          void* stream) {
  if (dtype != HFX_C128 && dtype != HFX_C64)
    return set_error(HFX_EINVAL, "hfx_superlattice_fock: dtype must be HFX_C128 or HFX_C64");
  if (nkx < 1 || nky < 1 || n_G < 1 || dim_orb < 1 || n_delta < 1 || Nx < nkx || Ny < nky || Nx % nkx ||
      Ny % nky)
    return set_error(HFX_EINVAL, "hfx_superlattice_fock: bad geometry (nk %dx%d, n_G %d, dim_orb %d, grid %dx%d)",
                     nkx, nky, n_G, dim_orb, Nx, Ny);
  if ((long long)n_delta * dim_orb * dim_orb > 65535 || (long long)nkx * nky > 65535 || n_G * n_G > 65535 ||
      (long long)n_G * dim_orb * n_G * dim_orb > (1LL << 30))
    return set_error(HFX_EUNSUPPORTED, "hfx_superlattice_fock: %d planes / %d k-points / D=%d out of range",
                     n_delta * dim_orb * dim_orb, nkx * nky, n_G * dim_orb);
  if (!P || !pair_to_delta || !g_off || !h || !Sigma || !F)
    return set_error(HFX_EINVAL, "hfx_superlattice_fock: NULL argument");
  if (planes && (!delta_i || !delta_j || !plan || !V_lag_fft || !sigma_planes || sigma_planes == planes))
    return set_error(HFX_EINVAL,
                     "hfx_superlattice_fock: the exchange needs delta_i, delta_j, plan, V_lag_fft and a separate "
                     "sigma_planes buffer");
  if (planes && (plan->n1 != Nx || plan->n2 != Ny || plan->nb != dim_orb || plan->dtype != dtype))
    return set_error(HFX_EINVAL,
                     "hfx_superlattice_fock: plan (%dx%d, nb %d, dtype %d) does not match the grid %dx%d, dim_orb %d, "
                     "dtype %d", plan->n1, plan->n2, plan->nb, plan->dtype, Nx, Ny, dim_orb, dtype);
  if (hartree && (!w2d || !delta_start || !delta_i || !delta_j || !HH_orb || !delta_scratch))
    return set_error(HFX_EINVAL, "hfx_superlattice_fock: the Hartree shift needs its tables and scratch");
  const SlGeom g{nkx, nky, n_G, dim_orb, n_delta, Nx, Ny};
  const hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_C128)
    return sl_build<double>(g, P, refP, pair_to_delta, g_off, plan, V_lag_fft, planes, sigma_planes,
                            hartree, w2d, delta_start, delta_i, delta_j, HH_orb, degeneracy, delta_scratch, h, Sigma,
                            F, s);
  return sl_build<float>(g, P, refP, pair_to_delta, g_off, plan, V_lag_fft, planes, sigma_planes,
                         hartree, w2d, delta_start, delta_i, delta_j, HH_orb, degeneracy, delta_scratch, h, Sigma, F,
                         s);
}

}  // extern "C"
