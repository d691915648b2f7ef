// Small elementwise / reduction kernels around the Fock build:
//   * precision conversion between the storage kinds (mixed-precision plumbing),
//   * hermitisation 0.5 (X + X^dagger) (fock.py:13-15, used after project_fn),
//   * max |P| and max |P| over an off-block channel mask (utils.py:263-266),
//   * occupation entropy / free energy (fock.py:114-129).
#include "hfx_common.h"
#include "hfx_internal.h"

namespace hfx {

constexpr int UTIL_BLOCK = 256;
constexpr int UTIL_MAX_BLOCKS = 1024;

enum Kind { K_F64 = 0, K_F32 = 1, K_C128 = 2, K_C64 = 3 };

__device__ __forceinline__ void load_any(const void* p, int kind, size_t i, double& re, double& im) {
  switch (kind) {
    case K_F64: re = ((const double*)p)[i]; im = 0.0; break;
    case K_F32: re = ((const float*)p)[i]; im = 0.0; break;
    case K_C128: { cx<double> v = ((const cx<double>*)p)[i]; re = v.re; im = v.im; } break;
    default: { cx<float> v = ((const cx<float>*)p)[i]; re = v.re; im = v.im; } break;
  }
}
__device__ __forceinline__ void store_any(void* p, int kind, size_t i, double re, double im) {
  switch (kind) {
    case K_F64: ((double*)p)[i] = re; break;
    case K_F32: ((float*)p)[i] = (float)re; break;
    case K_C128: ((cx<double>*)p)[i] = cx<double>{re, im}; break;
    default: ((cx<float>*)p)[i] = cx<float>{(float)re, (float)im}; break;
  }
}

__global__ void convert_kernel(const void* src, int sk, void* dst, int dk, size_t n) {
  for (size_t i = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * UTIL_BLOCK) {
    double re, im;
    load_any(src, sk, i, re, im);
    store_any(dst, dk, i, re, im);
  }
}

template <typename T, typename TY = T>
__global__ void hermitize_kernel(const cx<T>* X, cx<TY>* Y, size_t nmat, int nb) {
  const size_t nb2 = (size_t)nb * nb, total = nmat * nb2;
  for (size_t e = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; e < total; e += (size_t)gridDim.x * UTIL_BLOCK) {
    const size_t k = e / nb2;
    const int c = (int)(e - k * nb2), i = c / nb, j = c - i * nb;
    const cx<T> a = X[e], b = X[k * nb2 + (size_t)j * nb + i];
    const cx<T> h = (T)0.5 * (a + conj(b));
    Y[e] = cx<TY>{(TY)h.re, (TY)h.im};
  }
}

// out[0] = max |X|, out[1] = max |X| over channels with mask[c] != 0 (per block, then reduced)
template <typename T>
__global__ void absmax_kernel(const cx<T>* X, size_t nmat, int nb2, const unsigned char* mask, double* part) {
  __shared__ double sh[2 * (UTIL_BLOCK / 64)];
  double m0 = 0.0, m1 = 0.0;
  const size_t total = nmat * nb2;
  for (size_t e = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; e < total; e += (size_t)gridDim.x * UTIL_BLOCK) {
    const cx<T> v = X[e];
    const double a = sqrt((double)v.re * v.re + (double)v.im * v.im);
    m0 = fmax(m0, a);
    if (mask && mask[e % nb2]) m1 = fmax(m1, a);
  }
  m0 = wave_max(m0);
  m1 = wave_max(m1);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { sh[wid] = m0; sh[UTIL_BLOCK / 64 + wid] = m1; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < UTIL_BLOCK / 64; ++w) { m0 = fmax(m0, sh[w]); m1 = fmax(m1, sh[UTIL_BLOCK / 64 + w]); }
    m0 = fmax(m0, sh[0]);
    m1 = fmax(m1, sh[UTIL_BLOCK / 64]);
    part[2 * blockIdx.x] = m0;
    part[2 * blockIdx.x + 1] = m1;
  }
}

__global__ void absmax_final_kernel(const double* part, int n, double* out) {
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < n; ++i) { a = fmax(a, part[2 * i]); b = fmax(b, part[2 * i + 1]); }
    out[0] = a;
    out[1] = b;
  }
}

// S = -sum_k w_k sum_i [p ln p + (1-p) ln(1-p)], p clipped to [1e-14, 1-1e-14]
template <typename R>
__global__ void entropy_kernel(const R* p, const double* w, size_t nk, int nb, double* part) {
  __shared__ double sh[UTIL_BLOCK / 64];
  double s = 0.0;
  const size_t total = nk * nb;
  for (size_t e = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; e < total; e += (size_t)gridDim.x * UTIL_BLOCK) {
    double x = (double)p[e];
    x = fmin(fmax(x, 1e-14), 1.0 - 1e-14);
    s += w[e / nb] * (x * log(x) + (1.0 - x) * log1p(-x));
  }
  double acc[1] = {s};
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = -acc[0];
}

__global__ void entropy_final_kernel(const double* part, int n, const double* E, double T, double* out) {
  __shared__ double sh[1024 / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
  double acc[1] = {s};
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) *out = E ? (*E - fmax(T, 1e-14) * acc[0]) : acc[0];
}

// out = a X + b Y (complex a, b; Y may be NULL), elementwise over n values
__global__ void lincomb_kernel(size_t n, int kind, double ar, double ai, const void* X, double br, double bi,
                               const void* Y, void* out) {
  for (size_t i = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * UTIL_BLOCK) {
    double xr, xi, yr = 0.0, yi = 0.0;
    load_any(X, kind, i, xr, xi);
    if (Y) load_any(Y, kind, i, yr, yi);
    const double r = ar * xr - ai * xi + br * yr - bi * yi;
    const double m = ar * xi + ai * xr + br * yi + bi * yr;
    store_any(out, kind, i, r, m);
  }
}

// out_i = a' X_i + b' Y_i over two arrays (the matrix part and the occupation
// part of a Newton vector) in one launch; a' = a (* a_dev), b' = b (* b_dev)
// with a_dev / b_dev device scalars read by every thread, so the coefficients
// of the truncated-CG updates never leave the device.
__global__ void lincomb_dev_kernel(size_t n1, int kind1, const void* X1, const void* Y1, void* out1, size_t n2,
                                   int kind2, const void* X2, const void* Y2, void* out2, double a,
                                   const double* a_dev, double b, const double* b_dev) {
  const double ca = a_dev ? a * a_dev[0] : a, cb = b_dev ? b * b_dev[0] : b;
  const size_t n = n1 + n2;
  for (size_t i = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * UTIL_BLOCK) {
    const bool first = i < n1;
    const size_t j = first ? i : i - n1;
    const int kind = first ? kind1 : kind2;
    const void* X = first ? X1 : X2;
    const void* Y = first ? Y1 : Y2;
    double xr, xi, yr = 0.0, yi = 0.0;
    load_any(X, kind, j, xr, xi);
    if (Y) load_any(Y, kind, j, yr, yi);
    store_any(first ? out1 : out2, kind, j, ca * xr + cb * yr, ca * xi + cb * yi);
  }
}

// Up to four device-coefficient updates in one pass, applied per element in
// order (so an op may read what an earlier op of the list reads and a later
// op writes, at the same index): the truncated-CG updates of one inner step.
constexpr int LCD_MAX = 4;
struct LcdOps {
  const void* x1[LCD_MAX];
  const void* y1[LCD_MAX];
  void* o1[LCD_MAX];
  const void* x2[LCD_MAX];
  const void* y2[LCD_MAX];
  void* o2[LCD_MAX];
  double a[LCD_MAX], b[LCD_MAX];
  int ai[LCD_MAX], bi[LCD_MAX];
};

__global__ void lincomb_dev_batch_kernel(int nops, size_t n1, int kind1, size_t n2, int kind2, LcdOps op,
                                         const double* sc) {
  double ca[LCD_MAX], cb[LCD_MAX];
#pragma unroll
  for (int j = 0; j < LCD_MAX; ++j) {
    ca[j] = (j < nops && op.ai[j] >= 0) ? op.a[j] * sc[op.ai[j]] : op.a[j];
    cb[j] = (j < nops && op.bi[j] >= 0) ? op.b[j] * sc[op.bi[j]] : op.b[j];
  }
  const size_t n = n1 + n2;
  for (size_t i = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * UTIL_BLOCK) {
    const bool first = i < n1;
    const size_t e = first ? i : i - n1;
    const int kind = first ? kind1 : kind2;
#pragma unroll
    for (int j = 0; j < LCD_MAX; ++j)
      if (j < nops) {
        const void* X = first ? op.x1[j] : op.x2[j];
        const void* Y = first ? op.y1[j] : op.y2[j];
        double xr, xi, yr = 0.0, yi = 0.0;
        load_any(X, kind, e, xr, xi);
        if (Y) load_any(Y, kind, e, yr, yi);
        store_any(first ? op.o1[j] : op.o2[j], kind, e, ca[j] * xr + cb[j] * yr, ca[j] * xi + cb[j] * yi);
      }
  }
}

// out = sum_v c_v X_v over nmat nb x nb matrices, optionally Hermitised per
// matrix: out_ij = (Y_ij + conj(Y_ji)) / 2 with Y = sum_v c_v X_v.
constexpr int LCM_MAX = 8;
struct LcmInputs {
  const void* x[LCM_MAX];
  double c[LCM_MAX];
};

__global__ void lincomb_multi_kernel(size_t nmat, int nb, int kind, int nin, LcmInputs in, int herm, void* out) {
  const size_t nb2 = (size_t)nb * nb, total = nmat * nb2;
  for (size_t i = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; i < total; i += (size_t)gridDim.x * UTIL_BLOCK) {
    double r = 0.0, m = 0.0, rt = 0.0, mt = 0.0;
    size_t it = i;
    if (herm) {
      const size_t k = i / nb2, c = i - k * nb2, a = c / nb, b = c - a * nb;
      it = k * nb2 + b * nb + a;
    }
    for (int v = 0; v < nin; ++v) {
      double xr, xi;
      load_any(in.x[v], kind, i, xr, xi);
      r += in.c[v] * xr;
      m += in.c[v] * xi;
      if (herm) {
        load_any(in.x[v], kind, it, xr, xi);
        rt += in.c[v] * xr;
        mt += in.c[v] * xi;
      }
    }
    if (herm) {
      r = 0.5 * (r + rt);
      m = 0.5 * (m - mt);
    }
    store_any(out, kind, i, r, m);
  }
}

// out = expit(-(x - mu) / (T + 1e-12))  (utils.py:51-53)
__global__ void fermi_kernel(const double* x, const double* mu, double T, size_t n, double* out) {
  const double m = mu ? *mu : 0.0;
  for (size_t i = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * UTIL_BLOCK)
    out[i] = 1.0 / (1.0 + exp((x[i] - m) / (T + 1e-12)));
}

// x[i] -= scale * (*s) : the particle-conserving shift of solver.py:224-233
__global__ void vec_shift_kernel(double* x, size_t n, const double* s, double scale) {
  const double d = (*s) * scale;
  for (size_t i = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * UTIL_BLOCK) x[i] -= d;
}

// x[i] -= scale * sum_k w[k] part[k] with the sum formed by every block (the
// same strided partial sums and block tree in each, so the totals are
// bitwise identical): the projection of a Newton vector in one launch instead
// of a two-launch reduction and a shift.
constexpr int VSS_MAX_BLOCKS = 64;
__global__ void vec_shift_sum_kernel(double* x, size_t n, const double* __restrict__ part,
                                     const double* __restrict__ w, size_t nk, double scale) {
  __shared__ double sh[UTIL_BLOCK / 64];
  __shared__ double tot;
  double acc[1] = {0.0};
  for (size_t k = threadIdx.x; k < nk; k += UTIL_BLOCK) acc[0] += w[k] * part[k];
  block_sum<1>(acc, sh);
  if (threadIdx.x == 0) tot = acc[0];
  __syncthreads();
  const double d = tot * scale;
  for (size_t i = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * UTIL_BLOCK) x[i] -= d;
}

// part[v][k] = sum_b x_v[k, b] y_v[k, b]
struct VecPairs {
  const double* x[8];
  const double* y[8];
};
__global__ void vec_dots_kernel(VecPairs P, int nv, size_t nk, int nb, double* part) {
  for (size_t k = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; k < nk; k += (size_t)gridDim.x * UTIL_BLOCK)
    for (int v = 0; v < nv; ++v) {
      double acc = 0.0;
      for (int b = 0; b < nb; ++b) acc += P.x[v][k * nb + b] * P.y[v][k * nb + b];
      part[(size_t)v * nk + k] = acc;
    }
}

// out[k, c] = a conj(X[k, c]) Y[k, c], or summed over the nch channels per k
template <typename T>
__global__ void cmul_kernel(size_t nk, int nch, const cx<T>* X, const cx<T>* Y, double ar, double ai, int sum_ch,
                            cx<T>* out) {
  const size_t total = sum_ch ? nk : nk * nch;
  for (size_t i = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; i < total; i += (size_t)gridDim.x * UTIL_BLOCK) {
    double re = 0.0, im = 0.0;
    const size_t c0 = sum_ch ? i * nch : i, c1 = sum_ch ? c0 + nch : i + 1;
    for (size_t c = c0; c < c1; ++c) {
      const cx<T> x = X[c], y = Y[c];
      re += (double)x.re * y.re + (double)x.im * y.im;   // conj(x) y
      im += (double)x.re * y.im - (double)x.im * y.re;
    }
    out[i] = cx<T>{(T)(ar * re - ai * im), (T)(ar * im + ai * re)};
  }
}

// out[k, c] = alpha (*alpha_dev) w[k] X[k, c]  (per-k weight scaling, complex
// storage kinds; alpha_dev and w optional).  dag_nb > 0: X[k] is dag_nb x dag_nb
// and out[k] = alpha ... X[k]^dagger (out must not alias X then).
__global__ void scale_k_kernel(size_t nk, int nch, int kind, const void* X, const double* w, double alpha,
                               const double* alpha_dev, int dag_nb, void* out) {
  const size_t total = nk * nch;
  const double a = alpha * (alpha_dev ? *alpha_dev : 1.0);
  for (size_t i = (size_t)blockIdx.x * UTIL_BLOCK + threadIdx.x; i < total; i += (size_t)gridDim.x * UTIL_BLOCK) {
    double re, im;
    size_t src = i;
    if (dag_nb > 0) {
      const size_t k = i / nch;
      const int c = (int)(i - k * nch), r = c / dag_nb, q = c - r * dag_nb;
      src = k * nch + (size_t)q * dag_nb + r;
    }
    load_any(X, kind, src, re, im);
    if (dag_nb > 0) im = -im;
    const double s = a * (w ? w[i / nch] : 1.0);
    store_any(out, kind, i, s * re, s * im);
  }
}

static int grid_for(size_t n) {
  size_t g = (n + UTIL_BLOCK - 1) / UTIL_BLOCK;
  if (g > UTIL_MAX_BLOCKS) g = UTIL_MAX_BLOCKS;
  return (int)(g < 1 ? 1 : g);
}

// Periodic linear resampling of a centred k-grid (utils.py:477-535): axis 0 is
// interpolated first, then axis 1, each from the source coordinate
// u = (j - n_new/2) * (n_old/n_new) + n_old/2 in the array's real precision.
// Every product and sum is rounded on its own (no fused multiply-add), so the
// floor index and the weights are those of the same expression evaluated in
// IEEE arithmetic on the host.  One thread per output real scalar; a complex
// array interpolates its real and imaginary parts with the same weights.
// (__dmul_rn and friends are plain operators in clang's HIP headers, which
// HIP's default -ffp-contract=fast-honor-pragmas would fuse into an FMA with
// the following add; the pragma keeps each operation rounded on its own.)
__device__ __forceinline__ double rn_mul(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ double rn_add(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ float rn_mul(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ float rn_add(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}

template <typename R>
__device__ __forceinline__ void resample_src(int j, int n_old, int n_new, R ratio, int& i0, int& i1, R& f) {
  const R u = rn_add(rn_mul((R)j - (R)(n_new / 2), ratio), (R)(n_old / 2));
  const R fl = floor(u);
  f = rn_add(u, -fl);
  int i = (int)fl % n_old;
  if (i < 0) i += n_old;
  i0 = i;
  i1 = i + 1 == n_old ? 0 : i + 1;
}

template <typename R>
__device__ __forceinline__ R resample_lerp(R x0, R x1, R f) {
  return rn_add(rn_mul(rn_add((R)1, -f), x0), rn_mul(f, x1));
}

template <typename R>
__global__ void resample_kernel(int n0, int n1, int nk, size_t m, R r0, R r1, const R* __restrict__ X,
                                R* __restrict__ out) {
  const size_t total = (size_t)nk * nk * m;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const size_t c = t % m, ij = t / m;
    const int i = (int)(ij / nk), j = (int)(ij % nk);
    int a0 = i, a1 = i, b0 = j, b1 = j;
    R fa = 0, fb = 0;
    if (n0 != nk) resample_src(i, n0, nk, r0, a0, a1, fa);
    if (n1 != nk) resample_src(j, n1, nk, r1, b0, b1, fb);
    const R* row0 = X + (size_t)a0 * n1 * m + c;
    const R* row1 = X + (size_t)a1 * n1 * m + c;
    R y0 = row0[(size_t)b0 * m], y1 = 0;
    if (n0 != nk) y0 = resample_lerp(y0, row1[(size_t)b0 * m], fa);
    if (n1 != nk) {
      y1 = row0[(size_t)b1 * m];
      if (n0 != nk) y1 = resample_lerp(y1, row1[(size_t)b1 * m], fa);
      y0 = resample_lerp(y0, y1, fb);
    }
    out[t] = y0;
  }
}

}  // namespace hfx

using namespace hfx;

extern "C" {

int hfx_convert(const void* src, int src_kind, void* dst, int dst_kind, size_t n, void* stream) {
  if (!src || !dst || src_kind < 0 || src_kind > 3 || dst_kind < 0 || dst_kind > 3)
    return set_error(HFX_EINVAL, "hfx_convert: bad argument");
  if (n == 0) return HFX_OK;
  hipLaunchKernelGGL(convert_kernel, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, src, src_kind,
                     dst, dst_kind, n);
  return check_launch("convert_kernel");
}

int hfx_lincomb(size_t n, int kind, double ar, double ai, const void* X, double br, double bi, const void* Y,
                void* out, void* stream) {
  if (!X || !out || kind < 0 || kind > 3) return set_error(HFX_EINVAL, "hfx_lincomb: bad argument");
  if (n == 0) return HFX_OK;
  hipLaunchKernelGGL(lincomb_kernel, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, n, kind, ar, ai,
                     X, br, bi, Y, out);
  return check_launch("lincomb_kernel");
}

int hfx_lincomb_dev(size_t n1, int kind1, const void* X1, const void* Y1, void* out1, size_t n2, int kind2,
                    const void* X2, const void* Y2, void* out2, double a, const double* a_dev, double b,
                    const double* b_dev, void* stream) {
  if (!X1 || !out1 || kind1 < 0 || kind1 > 3 || (n2 && (!X2 || !out2 || kind2 < 0 || kind2 > 3)))
    return set_error(HFX_EINVAL, "hfx_lincomb_dev: bad argument");
  const size_t n = n1 + n2;
  if (n == 0) return HFX_OK;
  hipLaunchKernelGGL(lincomb_dev_kernel, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, n1, kind1, X1,
                     Y1, out1, n2, kind2, X2, Y2, out2, a, a_dev, b, b_dev);
  return check_launch("lincomb_dev_kernel");
}

int hfx_lincomb_dev_batch(int nops, size_t n1, int kind1, size_t n2, int kind2, const void* const* X1,
                          const void* const* Y1, void* const* out1, const void* const* X2, const void* const* Y2,
                          void* const* out2, const double* a, const int* a_idx, const double* b, const int* b_idx,
                          const double* dev_scalars, void* stream) {
  if (nops < 1 || nops > LCD_MAX || !X1 || !Y1 || !out1 || !a || !a_idx || !b || !b_idx || kind1 < 0 || kind1 > 3 ||
      (n2 && (!X2 || !Y2 || !out2 || kind2 < 0 || kind2 > 3)))
    return set_error(HFX_EINVAL, "hfx_lincomb_dev_batch: bad argument");
  LcdOps op{};
  for (int j = 0; j < nops; ++j) {
    if (!X1[j] || !out1[j] || (n2 && (!X2[j] || !out2[j])) || ((a_idx[j] >= 0 || b_idx[j] >= 0) && !dev_scalars))
      return set_error(HFX_EINVAL, "hfx_lincomb_dev_batch: op %d is incomplete", j);
    op.x1[j] = X1[j];
    op.y1[j] = Y1[j];
    op.o1[j] = out1[j];
    if (n2) {
      op.x2[j] = X2[j];
      op.y2[j] = Y2[j];
      op.o2[j] = out2[j];
    }
    op.a[j] = a[j];
    op.b[j] = b[j];
    op.ai[j] = a_idx[j];
    op.bi[j] = b_idx[j];
  }
  const size_t n = n1 + n2;
  if (n == 0) return HFX_OK;
  hipLaunchKernelGGL(lincomb_dev_batch_kernel, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, nops, n1,
                     kind1, n2, kind2, op, dev_scalars);
  return check_launch("lincomb_dev_batch_kernel");
}

int hfx_lincomb_multi(size_t nmat, int nb, int kind, int nin, const void* const* X, const double* coef, int herm,
                      void* out, void* stream) {
  if (!X || !coef || !out || nin < 1 || nin > LCM_MAX || nb < 1 || kind < 0 || kind > 3)
    return set_error(HFX_EINVAL, "hfx_lincomb_multi: bad argument (nin=%d)", nin);
  LcmInputs in{};
  for (int v = 0; v < nin; ++v) {
    if (!X[v]) return set_error(HFX_EINVAL, "hfx_lincomb_multi: input %d is NULL", v);
    if (herm && X[v] == out) return set_error(HFX_EINVAL, "hfx_lincomb_multi: out aliases an input with herm");
    in.x[v] = X[v];
    in.c[v] = coef[v];
  }
  const size_t n = nmat * (size_t)nb * nb;
  if (n == 0) return HFX_OK;
  hipLaunchKernelGGL(lincomb_multi_kernel, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, nmat, nb, kind,
                     nin, in, herm, out);
  return check_launch("lincomb_multi_kernel");
}

int hfx_fermi(const double* x, const double* mu, double T, size_t n, double* out, void* stream) {
  if (!x || !out) return set_error(HFX_EINVAL, "hfx_fermi: bad argument");
  if (n == 0) return HFX_OK;
  hipLaunchKernelGGL(fermi_kernel, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, x, mu, T, n, out);
  return check_launch("fermi_kernel");
}

int hfx_vec_shift_sum(double* x, size_t n, const double* part, const double* w, size_t nk, double scale,
                      void* stream) {
  if (!x || !part || !w || nk == 0) return set_error(HFX_EINVAL, "hfx_vec_shift_sum: bad argument");
  if (n == 0) return HFX_OK;
  size_t g = grid_for(n);
  if (g > (size_t)VSS_MAX_BLOCKS) g = VSS_MAX_BLOCKS;
  hipLaunchKernelGGL(vec_shift_sum_kernel, dim3((unsigned)g), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, x, n, part, w,
                     nk, scale);
  return check_launch("vec_shift_sum_kernel");
}

int hfx_vec_shift(double* x, size_t n, const double* s, double scale, void* stream) {
  if (!x || !s) return set_error(HFX_EINVAL, "hfx_vec_shift: bad argument");
  if (n == 0) return HFX_OK;
  hipLaunchKernelGGL(vec_shift_kernel, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, x, n, s, scale);
  return check_launch("vec_shift_kernel");
}

int hfx_vec_dots(const double* const* xs, const double* const* ys, int nv, size_t nk, int nb, double* part,
                 void* stream) {
  if (!xs || !ys || !part || nv < 1 || nv > 8 || nb < 1) return set_error(HFX_EINVAL, "hfx_vec_dots: bad argument");
  VecPairs P{};
  for (int v = 0; v < nv; ++v) {
    P.x[v] = xs[v];
    P.y[v] = ys[v];
  }
  hipLaunchKernelGGL(vec_dots_kernel, dim3(grid_for(nk)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, P, nv, nk, nb, part);
  return check_launch("vec_dots_kernel");
}

int hfx_cmul(int dtype, size_t nk, int nch, const void* X, const void* Y, double ar, double ai, int sum_ch,
             void* out, void* stream) {
  if (!X || !Y || !out || nch < 1) return set_error(HFX_EINVAL, "hfx_cmul: bad argument");
  const size_t n = sum_ch ? nk : nk * nch;
  if (dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL(cmul_kernel<double>, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, nk, nch,
                       (const cx<double>*)X, (const cx<double>*)Y, ar, ai, sum_ch, (cx<double>*)out);
  else
    hipLaunchKernelGGL(cmul_kernel<float>, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, nk, nch,
                       (const cx<float>*)X, (const cx<float>*)Y, ar, ai, sum_ch, (cx<float>*)out);
  return check_launch("cmul_kernel");
}

int hfx_scale_k(size_t nk, int nch, int kind, const void* X, const double* w, double alpha, void* out,
                void* stream) {
  if (!X || !out || nch < 1 || kind < 0 || kind > 3) return set_error(HFX_EINVAL, "hfx_scale_k: bad argument");
  hipLaunchKernelGGL(scale_k_kernel, dim3(grid_for(nk * nch)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, nk, nch, kind,
                     X, w, alpha, nullptr, 0, out);
  return check_launch("scale_k_kernel");
}

int hfx_scale_k_ex(size_t nk, int nb, int kind, const void* X, const double* w, double alpha, const double* alpha_dev,
                   int dagger, void* out, void* stream) {
  if (!X || !out || nb < 1 || kind < 2 || kind > 3 || (dagger && X == out))
    return set_error(HFX_EINVAL, "hfx_scale_k_ex: bad argument");
  hipLaunchKernelGGL(scale_k_kernel, dim3(grid_for(nk * (size_t)nb * nb)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream,
                     nk, nb * nb, kind, X, w, alpha, alpha_dev, dagger ? nb : 0, out);
  return check_launch("scale_k_kernel");
}

int hfx_hermitize(const void* X, void* Y, size_t nmat, int nb, int dtype, void* stream) {
  if (!X || !Y || nb <= 0) return set_error(HFX_EINVAL, "hfx_hermitize: bad argument");
  if (X == Y) return set_error(HFX_EINVAL, "hfx_hermitize: in-place not supported");
  const size_t n = nmat * nb * nb;
  if (dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL(hermitize_kernel<double>, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream,
                       (const cx<double>*)X, (cx<double>*)Y, nmat, nb);
  else
    hipLaunchKernelGGL(hermitize_kernel<float>, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream,
                       (const cx<float>*)X, (cx<float>*)Y, nmat, nb);
  return check_launch("hermitize_kernel");
}

int hfx_hermitize_convert(const void* X, int x_dtype, void* Y, int y_dtype, size_t nmat, int nb, void* stream) {
  if (!X || !Y || X == Y || nb <= 0 || (x_dtype != HFX_DTYPE_C128 && x_dtype != HFX_DTYPE_C64) ||
      (y_dtype != HFX_DTYPE_C128 && y_dtype != HFX_DTYPE_C64))
    return set_error(HFX_EINVAL, "hfx_hermitize_convert: bad argument");
  const size_t n = nmat * nb * nb;
  if (n == 0) return HFX_OK;
  hipStream_t s = (hipStream_t)stream;
  const dim3 g(grid_for(n)), b(UTIL_BLOCK);
  if (x_dtype == HFX_DTYPE_C128 && y_dtype == HFX_DTYPE_C64)
    hipLaunchKernelGGL((hermitize_kernel<double, float>), g, b, 0, s, (const cx<double>*)X, (cx<float>*)Y, nmat, nb);
  else if (x_dtype == HFX_DTYPE_C64 && y_dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL((hermitize_kernel<float, double>), g, b, 0, s, (const cx<float>*)X, (cx<double>*)Y, nmat, nb);
  else if (x_dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL((hermitize_kernel<double>), g, b, 0, s, (const cx<double>*)X, (cx<double>*)Y, nmat, nb);
  else
    hipLaunchKernelGGL((hermitize_kernel<float>), g, b, 0, s, (const cx<float>*)X, (cx<float>*)Y, nmat, nb);
  return check_launch("hermitize_kernel");
}

int hfx_absmax_masked(const void* X, size_t nmat, int nb, int dtype, const unsigned char* mask, double* scratch,
                      double* out, void* stream) {
  if (!X || !scratch || !out || nb <= 0) return set_error(HFX_EINVAL, "hfx_absmax_masked: bad argument");
  const size_t n = nmat * nb * nb;
  const int g = grid_for(n) > 512 ? 512 : grid_for(n);
  if (dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL(absmax_kernel<double>, dim3(g), dim3(UTIL_BLOCK), 0, (hipStream_t)stream,
                       (const cx<double>*)X, nmat, nb * nb, mask, scratch);
  else
    hipLaunchKernelGGL(absmax_kernel<float>, dim3(g), dim3(UTIL_BLOCK), 0, (hipStream_t)stream,
                       (const cx<float>*)X, nmat, nb * nb, mask, scratch);
  int rc = check_launch("absmax_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(absmax_final_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, scratch, g, out);
  return check_launch("absmax_final_kernel");
}

int hfx_entropy(const void* p, int real_kind, const double* w, size_t nk, int nb, const double* E, double T,
                double* scratch, double* out, void* stream) {
  if (!p || !w || !scratch || !out || nb <= 0 || (real_kind != 0 && real_kind != 1))
    return set_error(HFX_EINVAL, "hfx_entropy: bad argument");
  const size_t n = nk * nb;
  const int g = grid_for(n);
  if (real_kind == 0)
    hipLaunchKernelGGL(entropy_kernel<double>, dim3(g), dim3(UTIL_BLOCK), 0, (hipStream_t)stream,
                       (const double*)p, w, nk, nb, scratch);
  else
    hipLaunchKernelGGL(entropy_kernel<float>, dim3(g), dim3(UTIL_BLOCK), 0, (hipStream_t)stream,
                       (const float*)p, w, nk, nb, scratch);
  int rc = check_launch("entropy_kernel");
  if (rc) return rc;
  hipLaunchKernelGGL(entropy_final_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, scratch, g, E, T, out);
  return check_launch("entropy_final_kernel");
}

int hfx_resample_kgrid(int kind, int n0, int n1, size_t m, int nk, const void* X, void* out, void* stream) {
  if (!X || !out || kind < 0 || kind > 3 || n0 < 1 || n1 < 1 || nk < 1)
    return set_error(HFX_EINVAL, "hfx_resample_kgrid: bad argument (kind=%d, n0=%d, n1=%d, nk=%d)", kind, n0, n1, nk);
  if (X == out) return set_error(HFX_EINVAL, "hfx_resample_kgrid: out aliases the input");
  const size_t reals = m * (kind >= K_C128 ? 2 : 1);
  const size_t n = (size_t)nk * nk * reals;
  if (n == 0) return HFX_OK;
  if (kind == K_F64 || kind == K_C128) {
    hipLaunchKernelGGL(resample_kernel<double>, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, n0, n1,
                       nk, reals, (double)n0 / (double)nk, (double)n1 / (double)nk, (const double*)X, (double*)out);
  } else {
    hipLaunchKernelGGL(resample_kernel<float>, dim3(grid_for(n)), dim3(UTIL_BLOCK), 0, (hipStream_t)stream, n0, n1,
                       nk, reals, (float)n0 / (float)nk, (float)n1 / (float)nk, (const float*)X, (float*)out);
  }
  return check_launch("resample_kernel");
}

}  // extern "C"
