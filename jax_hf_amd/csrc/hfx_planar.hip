// Planar (channel-major) staging of the exchange operands.
//
// The exchange transforms one channel c = (i, j) of P over the whole k mesh
// per workgroup.  In the (k, i, j) storage of P / Sigma the 16 bytes of one
// channel at consecutive k lie nb^2 elements apart, so every lane of a load
// touches its own cache line.  The Fock build therefore moves the channels it
// needs into slot planes Xt[s][t][k] = X[s][k][chans[t]] first (one
// coalesced pass through an LDS tile of KT consecutive k-points), runs the
// exchange on the planes, and reads the Sigma planes back in the epilogue
// through the same kind of tile:
//
//   pack:      X (nsolve, Nk, nb, nb) [- refX]  ->  Xt (nsolve, nchan, Nk)
//   epilogue:  Sigma_t (nsolve, nchan, Nk)      ->  F = herm(h + Sigma + sigma_contact + diag(H)),
//              E = sum_k w_k Re Tr[(h + (Sigma + H)/2) P]           (fock.py:92-111)
//
// With the Hermitian channel set (i <= j) only the upper triangle is packed,
// transformed and read back; the epilogue mirrors Sigma_ji = conj(Sigma_ij)
// from the LDS tile.
#include "hfx_common.h"
#include "hfx_internal.h"

namespace hfx {

constexpr int PL_BLOCK = 256;
constexpr int PL_MAX_LDS = 40 * 1024;   // four epilogue workgroups per CU
constexpr int PL_EPI_MAX_BLOCKS = 4096;
constexpr int PL_TILE_LIMIT = 150 * 1024;   // largest epilogue tile (nb <= 96 in complex128)
#ifndef HFX_PL_U
#define HFX_PL_U 4
#endif
#ifndef HFX_PL_EPI_BLOCK
#define HFX_PL_EPI_BLOCK 512
#endif
constexpr int PL_U = HFX_PL_U;   // loads each thread issues before the first use (memory-level parallelism)
constexpr int EPI_BLK = HFX_PL_EPI_BLOCK;

// Bytes of one slot's k-segment in the epilogue tile (one cache line).
#ifndef HFX_PL_SEG
#define HFX_PL_SEG 128
#endif

// log2 of the epilogue k-tile: the largest power of two with KT * sizeof(cx<T>)
// <= HFX_PL_SEG whose padded tile fits PL_MAX_LDS.
int planar_tile_log2(int dtype, int nb) {
  const int esz = dtype == HFX_DTYPE_C128 ? 16 : 8;
  const int pitch = nb * nb + 1;
  int l = 0;
  while ((2 << l) <= HFX_PL_SEG / esz && (size_t)(2 << l) * pitch * esz <= (size_t)PL_MAX_LDS) ++l;
  return l;
}

static size_t planar_tile_bytes(int dtype, int nb) {
  const int esz = dtype == HFX_DTYPE_C128 ? 16 : 8;
  return ((size_t)1 << planar_tile_log2(dtype, nb)) * (size_t)(nb * nb + 1) * esz;
}

// ---- pack: X[s][k][c] - refX[k][c]  ->  Xt[s][t][k], c = chans[t] ----------
// One workgroup per (tile of KT k-points, chunk of CH slots, solve): lanes read
// along the channel index (consecutive slots are consecutive channels in the
// default lists) and write along k, KT * sizeof(cx<T>) = 512 B per slot.
template <typename T>
constexpr int pack_kt() { return 512 / (int)sizeof(cx<T>); }
constexpr int PACK_CH = 16;

template <typename T>
__global__ __launch_bounds__(PL_BLOCK) void pack_planar_kernel(int Nk, int nb2, const cx<T>* __restrict__ X,
                                                               const cx<T>* __restrict__ refX,
                                                               const int* __restrict__ chans, int nchan,
                                                               cx<T>* __restrict__ Xt) {
  constexpr int KT = pack_kt<T>(), CH = PACK_CH, ITEMS = KT * CH;
  static_assert(ITEMS % PL_BLOCK == 0, "pack tile must be a whole number of passes");
  __shared__ cx<T> tile[KT][CH + 1];
  const int e0 = blockIdx.x * KT, t0 = blockIdx.y * CH, s = blockIdx.z;
  const int kt = min(KT, Nk - e0), ch = min(CH, nchan - t0);
  const cx<T>* src = X + ((size_t)s * Nk + e0) * nb2;
  const cx<T>* rsrc = refX ? refX + (size_t)e0 * nb2 : nullptr;
  cx<T> v[ITEMS / PL_BLOCK];
#pragma unroll
  for (int u = 0; u < ITEMS / PL_BLOCK; ++u) {
    const int i = u * PL_BLOCK + threadIdx.x, kk = i / CH, q = i % CH;
    if (kk < kt && q < ch) {
      const size_t o = (size_t)kk * nb2 + chans[t0 + q];
      v[u] = src[o];
      if (rsrc) v[u] = v[u] - rsrc[o];
    }
  }
#pragma unroll
  for (int u = 0; u < ITEMS / PL_BLOCK; ++u) {
    const int i = u * PL_BLOCK + threadIdx.x, kk = i / CH, q = i % CH;
    if (kk < kt && q < ch) tile[kk][q] = v[u];
  }
  __syncthreads();
  cx<T>* dst = Xt + ((size_t)s * nchan + t0) * Nk + e0;
#pragma unroll
  for (int u = 0; u < ITEMS / PL_BLOCK; ++u) {
    const int i = u * PL_BLOCK + threadIdx.x, q = i / KT, kk = i % KT;
    if (kk < kt && q < ch) dst[(size_t)q * Nk + kk] = tile[kk][q];
  }
}

// ---- epilogue on Sigma planes -------------------------------------------
// blockIdx.y = solve.  Single-solve form: Hdiag / sigc / Sig_out optional,
// w2d shared.  Batched form (st != null): per-solve weights at s * w_stride,
// frozen solves skipped.  Energy partials at part[s * gridDim.x + blockIdx.x],
// summed in block order by planar_energy_kernel.  (A completion ticket in this
// kernel would need a device-scope release per block, which writes back the
// block's XCD L2 full of dirty F lines: 4x slower than the second launch.)
template <typename T>
__global__ __launch_bounds__(EPI_BLK) void epilogue_planar_kernel(
    int Nk, int nb, int ktl, const cx<T>* __restrict__ h, const cx<T>* __restrict__ St,
    const int* __restrict__ chans, int nchan, int herm, int cover, const double* __restrict__ Hdiag,
    const cx<double>* __restrict__ sigc, const cx<T>* __restrict__ P, const double* __restrict__ w2d,
    size_t w_stride, cx<T>* __restrict__ F, cx<T>* __restrict__ Sig_out, double* __restrict__ part,
    const double* __restrict__ st) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
  __shared__ double red[EPI_BLK / 64];
  const int nb2 = nb * nb, KT = 1 << ktl, pitch = nb2 + 1;
  const int s = blockIdx.y;
  double eacc = 0.0;
  if (!st || st[(size_t)s * HFX_ST_LEN + HFX_ST_FROZEN] == 0.0) {
    const size_t base = (size_t)s * Nk * nb2;
    const cx<T>* Ss = St + (size_t)s * nchan * Nk;
    const double* ws = w2d ? w2d + (size_t)s * w_stride : nullptr;
    const int ntiles = (Nk + KT - 1) >> ktl;
    for (int b = blockIdx.x; b < ntiles; b += gridDim.x) {
      const int e0 = b << ktl;
      const int kt = min(KT, Nk - e0);
      if (!cover) {
        for (int i = threadIdx.x; i < kt * pitch; i += EPI_BLK) tile[i] = cx<T>{0, 0};
        __syncthreads();
      }
      const int ns = nchan * KT;
      for (int i0 = 0; i0 < ns; i0 += PL_U * EPI_BLK) {
        cx<T> v[PL_U];
#pragma unroll
        for (int u = 0; u < PL_U; ++u) {
          const int i = i0 + u * EPI_BLK + threadIdx.x;
          const int t = i >> ktl, kk = i & (KT - 1);
          if (i < ns && kk < kt) v[u] = Ss[(size_t)t * Nk + e0 + kk];
        }
#pragma unroll
        for (int u = 0; u < PL_U; ++u) {
          const int i = i0 + u * EPI_BLK + threadIdx.x;
          const int t = i >> ktl, kk = i & (KT - 1);
          if (i < ns && kk < kt) {
            const int c = chans[t], ci = c / nb, cj = c - ci * nb;
            tile[kk * pitch + c] = v[u];
            if (herm && ci != cj) tile[kk * pitch + cj * nb + ci] = conj(v[u]);
          }
        }
      }
      __syncthreads();
      const int n = kt * nb2;
      const size_t eb = base + (size_t)e0 * nb2;
      for (int i0 = 0; i0 < n; i0 += PL_U * EPI_BLK) {
        cx<T> hij[PL_U], hji[PL_U], pji[PL_U];
#pragma unroll
        for (int u = 0; u < PL_U; ++u) {
          const int i = i0 + u * EPI_BLK + threadIdx.x;
          if (i < n) {
            const int kk = i / nb2, c = i - kk * nb2;
            const int ci = c / nb, cj = c - ci * nb;
            const size_t e = eb + i, et = e - c + cj * nb + ci;
            hij[u] = h[e];
            hji[u] = h[et];
            if (part) pji[u] = P[et];
          }
        }
#pragma unroll
        for (int u = 0; u < PL_U; ++u) {
          const int i = i0 + u * EPI_BLK + threadIdx.x;
          if (i < n) {
            const int kk = i / nb2, c = i - kk * nb2;
            const int ci = c / nb, cj = c - ci * nb, ct = cj * nb + ci;
            const size_t e = eb + i;
            cx<T> s_ij = tile[kk * pitch + c], s_ji = tile[kk * pitch + ct];
            if (sigc) {
              s_ij = s_ij + cvt<T>(sigc[c]);
              s_ji = s_ji + cvt<T>(sigc[ct]);
            }
            const T hd = (Hdiag && ci == cj) ? (T)Hdiag[ci] : (T)0;
            cx<T> a_ij = hij[u] + s_ij, a_ji = hji[u] + s_ji;
            a_ij.re += hd;
            a_ji.re += hd;
            F[e] = (T)0.5 * (a_ij + conj(a_ji));
            if (Sig_out) Sig_out[e] = s_ij;
            if (part) {
              // w_k Re[(h + (Sigma + H)/2)_ij P_ji]
              cx<T> m = hij[u] + (T)0.5 * s_ij;
              m.re += (T)0.5 * hd;
              eacc += ws[e0 + kk] * ((double)m.re * (double)pji[u].re - (double)m.im * (double)pji[u].im);
            }
          }
        }
      }
      __syncthreads();
    }
  }
  if (part) {
    double acc[1] = {eacc};
    block_sum<1>(acc, red);
    if (threadIdx.x == 0) part[(size_t)s * gridDim.x + blockIdx.x] = acc[0];
  }
}

__global__ __launch_bounds__(256) void planar_energy_kernel(const double* __restrict__ part, int nblk,
                                                            double offset, double* __restrict__ E_out,
                                                            double* __restrict__ st) {
  __shared__ double red[4];
  const int s = blockIdx.x;
  if (st && st[(size_t)s * HFX_ST_LEN + HFX_ST_FROZEN] != 0.0) return;
  double acc[1] = {0.0};
  for (int b = threadIdx.x; b < nblk; b += 256) acc[0] += part[(size_t)s * nblk + b];
  block_sum<1>(acc, red);
  if (threadIdx.x == 0) {
    if (st) st[(size_t)s * HFX_ST_LEN + HFX_ST_E] = acc[0] + offset;
    else *E_out = acc[0] + offset;
  }
}

// Raise the dynamic-LDS limit of a kernel to at least `bytes` (kept per kernel).
static int set_lds(const void* kern, int& limit, size_t bytes) {
  if ((size_t)limit < bytes) {
    HFX_HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    limit = (int)bytes;
  }
  return HFX_OK;
}

int pack_planar_launch(const hfx_plan* p, int nsolve, const void* X, const void* refX, const int* chans, int nchan,
                       void* Xt, hipStream_t s) {
  const int Nk = p->n1 * p->n2, nb2 = p->nb * p->nb;
  const int kt = p->dtype == HFX_DTYPE_C128 ? pack_kt<double>() : pack_kt<float>();
  const dim3 grid((unsigned)((Nk + kt - 1) / kt), (unsigned)((nchan + PACK_CH - 1) / PACK_CH), (unsigned)nsolve);
  if (p->dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL(pack_planar_kernel<double>, grid, dim3(PL_BLOCK), 0, s, Nk, nb2, (const cx<double>*)X,
                       (const cx<double>*)refX, chans, nchan, (cx<double>*)Xt);
  else
    hipLaunchKernelGGL(pack_planar_kernel<float>, grid, dim3(PL_BLOCK), 0, s, Nk, nb2, (const cx<float>*)X,
                       (const cx<float>*)refX, chans, nchan, (cx<float>*)Xt);
  return check_launch("pack_planar_kernel");
}

int epilogue_planar_launch(const hfx_plan* p, int nsolve, int max_blocks, const void* h, const void* St,
                           const int* chans, int nchan, int herm, int cover, const double* Hdiag, const void* sigc,
                           const void* P, const double* w2d, size_t w_stride, void* F, void* Sig_out,
                           double* part, double* E_out, double e_offset, double* st, hipStream_t s) {
  const int Nk = p->n1 * p->n2;
  const int ktl = planar_tile_log2(p->dtype, p->nb);
  const size_t lds = planar_tile_bytes(p->dtype, p->nb);
  if (lds > (size_t)PL_TILE_LIMIT)
    return set_error(HFX_EINVAL, "planar epilogue: nb=%d needs %zu B of LDS per k-point tile", p->nb, lds);
  const int ntiles = (Nk + (1 << ktl) - 1) >> ktl;
  const int gx = ntiles < max_blocks ? ntiles : max_blocks;
  const dim3 grid((unsigned)gx, (unsigned)nsolve);
  const bool energy = part && P && w2d && (E_out || st);
  double* pe = energy ? part : nullptr;
  if (p->dtype == HFX_DTYPE_C128) {
    auto kern = epilogue_planar_kernel<double>;
    static int limit = 64 * 1024;
    if (int rc = set_lds((const void*)kern, limit, lds)) return rc;
    hipLaunchKernelGGL(kern, grid, dim3(EPI_BLK), lds, s, Nk, p->nb, ktl, (const cx<double>*)h,
                       (const cx<double>*)St, chans, nchan, herm, cover, Hdiag, (const cx<double>*)sigc,
                       (const cx<double>*)P, w2d, w_stride, (cx<double>*)F, (cx<double>*)Sig_out, pe, st);
  } else {
    auto kern = epilogue_planar_kernel<float>;
    static int limit = 64 * 1024;
    if (int rc = set_lds((const void*)kern, limit, lds)) return rc;
    hipLaunchKernelGGL(kern, grid, dim3(EPI_BLK), lds, s, Nk, p->nb, ktl, (const cx<float>*)h,
                       (const cx<float>*)St, chans, nchan, herm, cover, Hdiag, (const cx<double>*)sigc,
                       (const cx<float>*)P, w2d, w_stride, (cx<float>*)F, (cx<float>*)Sig_out, pe, st);
  }
  int rc = check_launch("epilogue_planar_kernel");
  if (rc || !energy) return rc;
  hipLaunchKernelGGL(planar_energy_kernel, dim3((unsigned)nsolve), dim3(256), 0, s, part, gx, e_offset, E_out, st);
  return check_launch("planar_energy_kernel");
}

int planar_epi_max_blocks() { return PL_EPI_MAX_BLOCKS; }

}  // namespace hfx
