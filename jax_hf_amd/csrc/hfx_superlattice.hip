// Superlattice (moire) Fock build on the extended momentum grid
// (superlattice.py:167-377).
//
// The exchange of a plane-wave basis indexed by reciprocal vectors G is a 2D
// convolution in absolute momentum, one independent convolution per
// (dG = G_a - G_b, alpha, beta) channel.  The GPU form runs every channel at
// once instead of streaming them:
//   1. sl_scatter_kernel writes the slot planes X[d][a,b](x, y) of the extended
//      grid: the (nkx, nky) block of (P - refP)[(i,a), (j,b)] sits at tile
//      g_off[i] of plane d = dG(i, j); everything else is zero.
//   2. the planar exchange of the plan (hfx_selfenergy_planar, one shared
//      kernel plane V_lag_fft) turns each plane into -ifft2(fft2(X) V), written
//      to a second set of planes.
//   3. sl_gather_kernel reads every (k, p, q) of Sigma back from its plane and
//      tile, adds the k-independent Hartree shift and forms F = herm(h + Sigma)
//      in the same pass.
// The Hartree shift (superlattice.py:322-357) is two small launches: weighted
// k sums of the orbital-diagonal blocks, then the per-dG sums contracted with
// the orbital-resolved Coulomb table.
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

template <typename T>
__global__ void sl_scatter_kernel(SlGeom g, const cx<T>* __restrict__ P, const cx<T>* __restrict__ refP,
                                  const int* __restrict__ tile_G, const int* __restrict__ partner,
                                  cx<T>* __restrict__ planes) {
  const int D = g.nG * g.dorb, nch = g.dorb * g.dorb, tiles_y = g.Ny / g.nky;
  const size_t plane = (size_t)g.Nx * g.Ny, total = (size_t)g.ndelta * nch * plane;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const int xy = (int)(e % plane);
    const int t = (int)((e / plane) % nch), d = (int)(e / (plane * nch));
    const int x = xy / g.Ny, y = xy % g.Ny, tx = x / g.nkx, ty = y / g.nky;
    cx<T> v{0, 0};
    const int i = tile_G[tx * tiles_y + ty];
    if (i >= 0) {
      const int j = partner[d * g.nG + i];
      if (j >= 0) {
        const int a = t / g.dorb, b = t % g.dorb;
        const size_t k = (size_t)(x - tx * g.nkx) * g.nky + (y - ty * g.nky);
        const size_t src = (k * D + i * g.dorb + a) * D + j * g.dorb + b;
        v = P[src];
        if (refP) v = v - refP[src];
      }
    }
    planes[e] = v;
  }
}

// rho_po[(A nG + B) dorb + c] = sum_k w[k] (P - refP)[k, (A, c), (B, c)]
template <typename T>
__global__ void sl_hartree_moments_kernel(int nk, int nG, int dorb, const cx<T>* __restrict__ P,
                                          const cx<T>* __restrict__ refP, const double* __restrict__ w,
                                          cx<double>* __restrict__ rho_po) {
  const int D = nG * dorb, n = nG * nG * dorb;
  const int o = blockIdx.x;                      // one workgroup per output
  if (o >= n) return;
  const int c = o % dorb, AB = o / dorb, A = AB / nG, B = AB % nG;
  const size_t off = (size_t)(A * dorb + c) * D + B * dorb + c;
  double re = 0.0, im = 0.0;
  for (int k = threadIdx.x; k < nk; k += blockDim.x) {
    cx<T> v = P[(size_t)k * D * D + off];
    if (refP) v = v - refP[(size_t)k * D * D + off];
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
    rho_po[o] = cx<double>{sr, si};
  }
}

// sigma_H[(A, a), (B, b)] = delta_ab deg sum_c HH[A, B, a, c] R(dG(A, B), c),
// R(d, c) = sum over the (A', B') pairs of d of rho_po[A', B', c]
template <typename T>
__global__ void sl_hartree_shift_kernel(int nG, int dorb, const int* __restrict__ pair_to_delta,
                                        const int* __restrict__ dstart, const int* __restrict__ di,
                                        const int* __restrict__ dj, const double* __restrict__ HH, double deg,
                                        const cx<double>* __restrict__ rho_po, cx<T>* __restrict__ sigH) {
  const int D = nG * dorb;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= D * D) return;
  const int p = e / D, q = e % D, A = p / dorb, a = p % dorb, B = q / dorb, b = q % dorb;
  double re = 0.0, im = 0.0;
  if (a == b) {
    const int d = pair_to_delta[A * nG + B];
    for (int c = 0; c < dorb; ++c) {
      double rr = 0.0, ri = 0.0;
      for (int s = dstart[d]; s < dstart[d + 1]; ++s) {
        const cx<double> v = rho_po[(di[s] * nG + dj[s]) * dorb + c];
        rr += v.re;
        ri += v.im;
      }
      const double hv = HH[((size_t)(A * nG + B) * dorb + a) * dorb + c];
      re += hv * rr;
      im += hv * ri;
    }
    re *= deg;
    im *= deg;
  }
  sigH[e] = cx<T>{(T)re, (T)im};
}

template <typename T>
__device__ __forceinline__ cx<T> sl_plane_at(const SlGeom& g, const cx<T>* planes, const int* p2d, const int* goff,
                                             int i, int a, int j, int b, int kx, int ky) {
  const int d = p2d[i * g.nG + j];
  const int x = g.nkx * goff[2 * i] + kx, y = g.nky * goff[2 * i + 1] + ky;
  return planes[(((size_t)d * g.dorb * g.dorb + a * g.dorb + b) * g.Nx + x) * g.Ny + y];
}

template <typename T>
__global__ void sl_gather_kernel(SlGeom g, const cx<T>* __restrict__ planes, const int* __restrict__ p2d,
                                 const int* __restrict__ goff, const cx<T>* __restrict__ sigH,
                                 const cx<T>* __restrict__ h, cx<T>* __restrict__ Sigma, cx<T>* __restrict__ F) {
  const int D = g.nG * g.dorb;
  const size_t DD = (size_t)D * D, total = (size_t)g.nkx * g.nky * DD;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t k = e / DD;
    const int pq = (int)(e % DD), p = pq / D, q = pq % D;
    const int kx = (int)(k / g.nky), ky = (int)(k % g.nky);
    const int i = p / g.dorb, a = p % g.dorb, j = q / g.dorb, b = q % g.dorb;
    cx<T> s{0, 0}, st{0, 0};
    if (planes) {
      s = sl_plane_at(g, planes, p2d, goff, i, a, j, b, kx, ky);
      st = sl_plane_at(g, planes, p2d, goff, j, b, i, a, kx, ky);
    }
    if (sigH) {
      s = s + sigH[pq];
      st = st + sigH[(size_t)q * D + p];
    }
    Sigma[e] = s;
    const cx<T> f = h[e] + s, ft = h[k * DD + (size_t)q * D + p] + st;
    F[e] = cx<T>{(T)0.5 * (f.re + ft.re), (T)0.5 * (f.im - ft.im)};
  }
}

template <typename T>
static int sl_build(const SlGeom& g, const void* P, const void* refP, const int* tile_G, const int* partner,
                    const int* p2d, const int* goff, const hfx_plan* plan, const void* Vlag, void* planes,
                    void* splanes, int hartree, const double* w, const int* dstart, const int* di, const int* dj, const double* HH,
                    double deg, void* rho_po, void* sigH, const void* h, void* Sigma, void* F, hipStream_t s) {
  const int D = g.nG * g.dorb;
  const size_t nplanes = (size_t)g.ndelta * g.dorb * g.dorb * g.Nx * g.Ny;
  if (planes) {
    hipLaunchKernelGGL(sl_scatter_kernel<T>, dim3(sl_grid(nplanes)), dim3(SL_BLOCK), 0, s, g, (const cx<T>*)P,
                       (const cx<T>*)refP, tile_G, partner, (cx<T>*)planes);
    int rc = check_launch("sl_scatter_kernel");
    if (rc) return rc;
    rc = hfx_selfenergy_planar(plan, g.ndelta, planes, Vlag, 0, nullptr, g.dorb * g.dorb, 0, 0, nullptr, splanes,
                               nullptr, (void*)s);
    if (rc) return rc;
  }
  if (hartree) {
    const int nmom = g.nG * g.nG * g.dorb;
    hipLaunchKernelGGL(sl_hartree_moments_kernel<T>, dim3(nmom), dim3(SL_BLOCK), 0, s, g.nkx * g.nky, g.nG, g.dorb,
                       (const cx<T>*)P, (const cx<T>*)refP, w, (cx<double>*)rho_po);
    int rc = check_launch("sl_hartree_moments_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(sl_hartree_shift_kernel<T>, dim3((D * D + SL_BLOCK - 1) / SL_BLOCK), dim3(SL_BLOCK), 0, s,
                       g.nG, g.dorb, p2d, dstart, di, dj, HH, deg, (const cx<double>*)rho_po, (cx<T>*)sigH);
    rc = check_launch("sl_hartree_shift_kernel");
    if (rc) return rc;
  }
  const size_t nout = (size_t)g.nkx * g.nky * D * D;
  hipLaunchKernelGGL(sl_gather_kernel<T>, dim3(sl_grid(nout)), dim3(SL_BLOCK), 0, s, g,
                     planes ? (const cx<T>*)splanes : nullptr, p2d,
                     goff, hartree ? (const cx<T>*)sigH : nullptr, (const cx<T>*)h, (cx<T>*)Sigma, (cx<T>*)F);
  return check_launch("sl_gather_kernel");
}

}  // namespace hfx

using namespace hfx;

extern "C" {

int hfx_superlattice_fock(int dtype, int nkx, int nky, int n_G, int dim_orb, int n_delta, int Nx, int Ny,
                          const void* P, const void* refP, const int* tile_G, const int* partner,
                          const int* pair_to_delta, const int* g_off, const hfx_plan* plan, const void* V_lag_fft,
                          void* planes, void* sigma_planes, int hartree, const double* w2d, const int* delta_start,
                          const int* delta_i, const int* delta_j, const double* HH_orb, double degeneracy,
                          void* rho_scratch, void* sigma_H, const void* h, void* Sigma, void* F, void* stream) {
  if (dtype != HFX_C128 && dtype != HFX_C64)
    return set_error(HFX_EINVAL, "hfx_superlattice_fock: dtype must be HFX_C128 or HFX_C64");
  if (nkx < 1 || nky < 1 || n_G < 1 || dim_orb < 1 || n_delta < 1 || Nx < nkx || Ny < nky || Nx % nkx ||
      Ny % nky)
    return set_error(HFX_EINVAL, "hfx_superlattice_fock: bad geometry (nk %dx%d, n_G %d, dim_orb %d, grid %dx%d)",
                     nkx, nky, n_G, dim_orb, Nx, Ny);
  if (!P || !pair_to_delta || !g_off || !h || !Sigma || !F)
    return set_error(HFX_EINVAL, "hfx_superlattice_fock: NULL argument");
  if (planes && (!tile_G || !partner || !plan || !V_lag_fft || !sigma_planes || sigma_planes == planes))
    return set_error(HFX_EINVAL,
                     "hfx_superlattice_fock: the exchange needs tile_G, partner, plan, V_lag_fft and a separate "
                     "sigma_planes buffer");
  if (hartree && (!w2d || !delta_start || !delta_i || !delta_j || !HH_orb || !rho_scratch || !sigma_H))
    return set_error(HFX_EINVAL, "hfx_superlattice_fock: the Hartree shift needs its tables and scratch");
  const SlGeom g{nkx, nky, n_G, dim_orb, n_delta, Nx, Ny};
  const hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_C128)
    return sl_build<double>(g, P, refP, tile_G, partner, pair_to_delta, g_off, plan, V_lag_fft, planes, sigma_planes,
                            hartree,
                            w2d, delta_start, delta_i, delta_j, HH_orb, degeneracy, rho_scratch, sigma_H, h, Sigma,
                            F, s);
  return sl_build<float>(g, P, refP, tile_G, partner, pair_to_delta, g_off, plan, V_lag_fft, planes, sigma_planes,
                         hartree, w2d,
                         delta_start, delta_i, delta_j, HH_orb, degeneracy, rho_scratch, sigma_H, h, Sigma, F, s);
}

}  // extern "C"
