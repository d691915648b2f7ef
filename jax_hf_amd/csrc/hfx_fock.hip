// Fock-matrix epilogue, k-independent (Hartree / contact) terms and the HF
// energy reduction.
//
//   F     = herm(h + Sigma + sigma_contact + diag(H))            fock.py:94
//   Sigma_out = Sigma + sigma_contact                             fock.py:92
//   E     = sum_k w_k Re Tr[(h + (Sigma_out + H)/2) P]            fock.py:109-111
//   H_a   = sum_b HH_ab Re rho_bb,  rho = sum_k w_k (P - refP)(k) fock.py:66-74
//   sigma_contact = sum_t g_t [O_i tr(O_j rho) - O_i rho O_j]     fock.py:76-92
//
// Energies are accumulated in FP64 for both storage precisions; per-block
// partial sums are reduced by one fixed-order pass, so results are bitwise
// reproducible run to run.
#include "hfx_common.h"
#include "hfx_internal.h"

namespace hfx {

constexpr int EPI_BLOCK = 256;
constexpr int EPI_MAX_BLOCKS = 1024;

// ---- k-independent terms: one small block -------------------------------
__global__ void kterms_kernel(int nb, const cx<double>* __restrict__ rho, const double* __restrict__ HH,
                              int ntc, const double* __restrict__ g, const cx<double>* __restrict__ Oi,
                              const cx<double>* __restrict__ Oj, double* __restrict__ Hdiag,
                              cx<double>* __restrict__ sigc) {
  const int nb2 = nb * nb;
  for (int e = threadIdx.x; e < nb2; e += blockDim.x) {
    const int a = e / nb, b = e - a * nb;
    cx<double> acc{0.0, 0.0};
    for (int t = 0; t < ntc && sigc; ++t) {
      const cx<double>* oi = Oi + (size_t)t * nb2;
      const cx<double>* oj = Oj + (size_t)t * nb2;
      // tr(O_j rho)
      cx<double> tr{0.0, 0.0};
      for (int x = 0; x < nb; ++x)
        for (int y = 0; y < nb; ++y) tr = tr + oj[x * nb + y] * rho[y * nb + x];
      // (O_i rho O_j)_ab
      cx<double> m{0.0, 0.0};
      for (int x = 0; x < nb; ++x) {
        cx<double> r{0.0, 0.0};
        for (int y = 0; y < nb; ++y) r = r + rho[x * nb + y] * oj[y * nb + b];
        m = m + oi[a * nb + x] * r;
      }
      acc = acc + g[t] * (oi[a * nb + b] * tr - m);
    }
    if (sigc) sigc[e] = acc;
    if (Hdiag && b == 0) {
      double s = 0.0;
      for (int c = 0; c < nb; ++c) s += HH[a * nb + c] * rho[c * nb + c].re;
      Hdiag[a] = s;
    }
  }
}

// ---- weighted moments rho_c = sum_k w_k (P - refP)_c when exchange is off --
template <typename T>
__global__ void moments_kernel(int Nk, int nb, const cx<T>* __restrict__ P, const cx<T>* __restrict__ refP,
                               const double* __restrict__ w2d, cx<double>* __restrict__ rho) {
  __shared__ double red[2 * (EPI_BLOCK / 64)];
  const int nb2 = nb * nb;
  const int c = blockIdx.x;
  double re = 0.0, im = 0.0;
  for (int k = threadIdx.x; k < Nk; k += EPI_BLOCK) {
    cx<T> v = P[(size_t)k * nb2 + c];
    if (refP) v = v - refP[(size_t)k * nb2 + c];
    re += w2d[k] * (double)v.re;
    im += w2d[k] * (double)v.im;
  }
  double acc[2] = {re, im};
  block_sum<2>(acc, red);
  if (threadIdx.x == 0) rho[c] = cx<double>{acc[0], acc[1]};
}

// ---- Fock epilogue + energy partials ------------------------------------
template <typename T>
__global__ __launch_bounds__(EPI_BLOCK) void fock_epilogue_kernel(
    int Nk, int nb, const cx<T>* __restrict__ h, const cx<T>* Sig, const double* __restrict__ Hdiag,
    const cx<double>* __restrict__ sigc, const cx<T>* __restrict__ P, const double* __restrict__ w2d,
    cx<T>* __restrict__ F, cx<T>* Sig_out, double* __restrict__ partials) {
  __shared__ double red[EPI_BLOCK / 64];
  const int nb2 = nb * nb;
  const size_t total = (size_t)Nk * nb2;
  double eacc = 0.0;
  for (size_t e = (size_t)blockIdx.x * EPI_BLOCK + threadIdx.x; e < total; e += (size_t)gridDim.x * EPI_BLOCK) {
    const size_t k = e / nb2;
    const int c = (int)(e - k * nb2);
    const int i = c / nb, j = c - i * nb;
    const size_t et = k * nb2 + (size_t)j * nb + i;
    cx<T> s_ij = Sig ? Sig[e] : cx<T>{0, 0};
    cx<T> s_ji = Sig ? Sig[et] : cx<T>{0, 0};
    if (sigc) {
      s_ij = s_ij + cvt<T>(sigc[c]);
      s_ji = s_ji + cvt<T>(sigc[j * nb + i]);
    }
    const T hd = (Hdiag && i == j) ? (T)Hdiag[i] : (T)0;
    cx<T> a_ij = h[e] + s_ij;
    cx<T> a_ji = h[et] + s_ji;
    a_ij.re += hd;
    a_ji.re += hd;
    cx<T> f = (T)0.5 * (a_ij + conj(a_ji));
    F[e] = f;
    if (Sig_out) Sig_out[e] = s_ij;
    if (P) {
      // w_k Re[(h + (Sigma + H)/2)_ij P_ji]
      cx<T> m = h[e] + (T)0.5 * s_ij;
      m.re += (T)0.5 * hd;
      const cx<T> pji = P[et];
      eacc += w2d[k] * ((double)m.re * (double)pji.re - (double)m.im * (double)pji.im);
    }
  }
  if (partials) {
    double acc[1] = {eacc};
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
  }
}

// ---- API hf_energy: H either (1,1,nb,nb) broadcast or full ---------------
template <typename T>
__global__ __launch_bounds__(EPI_BLOCK) void hf_energy_kernel(
    int Nk, int nb, const cx<T>* __restrict__ P, const cx<T>* __restrict__ h, const cx<T>* __restrict__ Sig,
    const cx<T>* __restrict__ H, int H_bcast, const double* __restrict__ w2d, double* __restrict__ partials) {
  __shared__ double red[EPI_BLOCK / 64];
  const int nb2 = nb * nb;
  const size_t total = (size_t)Nk * nb2;
  double eacc = 0.0;
  for (size_t e = (size_t)blockIdx.x * EPI_BLOCK + threadIdx.x; e < total; e += (size_t)gridDim.x * EPI_BLOCK) {
    const size_t k = e / nb2;
    const int c = (int)(e - k * nb2);
    const int i = c / nb, j = c - i * nb;
    cx<T> s = Sig ? Sig[e] : cx<T>{0, 0};
    if (H) s = s + H[H_bcast ? (size_t)c : e];
    const cx<T> m = h[e] + (T)0.5 * s;
    const cx<T> pji = P[k * nb2 + (size_t)j * nb + i];
    eacc += w2d[k] * ((double)m.re * (double)pji.re - (double)m.im * (double)pji.im);
  }
  double acc[1] = {eacc};
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

__global__ void reduce_partials_kernel(int n, const double* __restrict__ partials, double offset,
                                       double* __restrict__ out) {
  __shared__ double red[1024 / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
  double acc[1] = {s};
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) *out = acc[0] + offset;
}

static int epi_grid(size_t total) {
  size_t g = (total + EPI_BLOCK - 1) / EPI_BLOCK;
  if (g > (size_t)EPI_MAX_BLOCKS) g = EPI_MAX_BLOCKS;
  if (g < 1) g = 1;
  return (int)g;
}

int energy_partials_max() { return EPI_MAX_BLOCKS; }

int kterms_launch(const hfx_plan* p, const void* rho, const double* HH, int ntc, const double* g,
                  const void* Oi, const void* Oj, double* Hdiag, void* sigc, hipStream_t s) {
  hipLaunchKernelGGL(kterms_kernel, dim3(1), dim3(256), 0, s, p->nb, (const cx<double>*)rho, HH, ntc, g,
                     (const cx<double>*)Oi, (const cx<double>*)Oj, Hdiag, (cx<double>*)sigc);
  return check_launch("kterms_kernel");
}

int moments_launch(const hfx_plan* p, const void* P, const void* refP, const double* w2d, void* rho,
                   hipStream_t s) {
  const int Nk = p->n1 * p->n2, nb2 = p->nb * p->nb;
  if (p->dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL(moments_kernel<double>, dim3(nb2), dim3(EPI_BLOCK), 0, s, Nk, p->nb,
                       (const cx<double>*)P, (const cx<double>*)refP, w2d, (cx<double>*)rho);
  else
    hipLaunchKernelGGL(moments_kernel<float>, dim3(nb2), dim3(EPI_BLOCK), 0, s, Nk, p->nb,
                       (const cx<float>*)P, (const cx<float>*)refP, w2d, (cx<double>*)rho);
  return check_launch("moments_kernel");
}

int epilogue_launch(const hfx_plan* p, const void* h, const void* Sig, const double* Hdiag, const void* sigc,
                    const void* P, const double* w2d, void* F, void* Sig_out, double* partials, double* E_out,
                    double e_offset, hipStream_t s) {
  const int Nk = p->n1 * p->n2;
  const int grid = epi_grid((size_t)Nk * p->nb * p->nb);
  double* part = (P && E_out) ? partials : nullptr;
  const void* Pe = (P && E_out) ? P : nullptr;
  if (p->dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL(fock_epilogue_kernel<double>, dim3(grid), dim3(EPI_BLOCK), 0, s, Nk, p->nb,
                       (const cx<double>*)h, (const cx<double>*)Sig, Hdiag, (const cx<double>*)sigc,
                       (const cx<double>*)Pe, w2d, (cx<double>*)F, (cx<double>*)Sig_out, part);
  else
    hipLaunchKernelGGL(fock_epilogue_kernel<float>, dim3(grid), dim3(EPI_BLOCK), 0, s, Nk, p->nb,
                       (const cx<float>*)h, (const cx<float>*)Sig, Hdiag, (const cx<double>*)sigc,
                       (const cx<float>*)Pe, w2d, (cx<float>*)F, (cx<float>*)Sig_out, part);
  int rc = check_launch("fock_epilogue_kernel");
  if (rc || !part) return rc;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(1024), 0, s, grid, partials, e_offset, E_out);
  return check_launch("reduce_partials_kernel");
}

// Batched epilogue: nsolve consecutive solves of Nk k-points; blockIdx.y is
// the solve, energy partials per (solve, block); frozen solves are skipped.
template <typename T>
__global__ __launch_bounds__(EPI_BLOCK) void fock_epilogue_seg_kernel(
    int Nk, int nb, const cx<T>* __restrict__ h, const cx<T>* __restrict__ Sig, const cx<T>* __restrict__ P,
    const double* __restrict__ w2d, cx<T>* __restrict__ F, const double* __restrict__ st, double* __restrict__ part) {
  __shared__ double red[EPI_BLOCK / 64];
  const int sidx = blockIdx.y;
  const int nb2 = nb * nb;
  const size_t total = (size_t)Nk * nb2, base = (size_t)sidx * total;
  double eacc = 0.0;
  if (st[(size_t)sidx * HFX_ST_LEN + HFX_ST_FROZEN] == 0.0) {
    for (size_t e0 = (size_t)blockIdx.x * EPI_BLOCK + threadIdx.x; e0 < total; e0 += (size_t)gridDim.x * EPI_BLOCK) {
      const size_t e = base + e0;
      const size_t k = e0 / nb2;
      const int c = (int)(e0 - k * nb2);
      const int i = c / nb, j = c - i * nb;
      const size_t et = base + k * nb2 + (size_t)j * nb + i;
      const cx<T> a_ij = h[e] + Sig[e], a_ji = h[et] + Sig[et];
      F[e] = (T)0.5 * (a_ij + conj(a_ji));
      const cx<T> m = h[e] + (T)0.5 * Sig[e];
      const cx<T> pji = P[et];
      eacc += w2d[(size_t)sidx * Nk + k] * ((double)m.re * (double)pji.re - (double)m.im * (double)pji.im);
    }
  }
  double acc[1] = {eacc};
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) part[(size_t)sidx * gridDim.x + blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(256) void energy_seg_kernel(const double* __restrict__ part, int nblk, double* st) {
  __shared__ double red[4];
  const int sidx = blockIdx.x;
  double* sts = st + (size_t)sidx * HFX_ST_LEN;
  if (sts[HFX_ST_FROZEN] != 0.0) return;
  double acc[1] = {0.0};
  for (int b = threadIdx.x; b < nblk; b += 256) acc[0] += part[(size_t)sidx * nblk + b];
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) sts[HFX_ST_E] = acc[0];
}

int epilogue_seg_launch(const hfx_plan* p, int nsolve, const void* h, const void* Sig, const void* P, const double* w2d,
                        void* F, double* st, double* part, hipStream_t s) {
  const int Nk = p->n1 * p->n2;
  int gx = epi_grid((size_t)Nk * p->nb * p->nb);
  if (gx > 64) gx = 64;
  const dim3 grid((unsigned)gx, (unsigned)nsolve);
  if (p->dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL(fock_epilogue_seg_kernel<double>, grid, dim3(EPI_BLOCK), 0, s, Nk, p->nb,
                       (const cx<double>*)h, (const cx<double>*)Sig, (const cx<double>*)P, w2d, (cx<double>*)F, st,
                       part);
  else
    hipLaunchKernelGGL(fock_epilogue_seg_kernel<float>, grid, dim3(EPI_BLOCK), 0, s, Nk, p->nb,
                       (const cx<float>*)h, (const cx<float>*)Sig, (const cx<float>*)P, w2d, (cx<float>*)F, st, part);
  int rc = check_launch("fock_epilogue_seg_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(energy_seg_kernel, dim3((unsigned)nsolve), dim3(256), 0, s, part, gx, st);
  return check_launch("energy_seg_kernel");
}

int hf_energy_launch(const hfx_plan* p, const void* P, const void* h, const void* Sig, const void* H,
                     int H_bcast, const double* w2d, double* partials, double* E_out, hipStream_t s) {
  const int Nk = p->n1 * p->n2;
  const int grid = epi_grid((size_t)Nk * p->nb * p->nb);
  if (p->dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL(hf_energy_kernel<double>, dim3(grid), dim3(EPI_BLOCK), 0, s, Nk, p->nb,
                       (const cx<double>*)P, (const cx<double>*)h, (const cx<double>*)Sig,
                       (const cx<double>*)H, H_bcast, w2d, partials);
  else
    hipLaunchKernelGGL(hf_energy_kernel<float>, dim3(grid), dim3(EPI_BLOCK), 0, s, Nk, p->nb,
                       (const cx<float>*)P, (const cx<float>*)h, (const cx<float>*)Sig,
                       (const cx<float>*)H, H_bcast, w2d, partials);
  int rc = check_launch("hf_energy_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(1024), 0, s, grid, partials, 0.0, E_out);
  return check_launch("reduce_partials_kernel");
}

}  // namespace hfx
