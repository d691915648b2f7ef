// Exchange self-energy as a fused 2D FFT convolution over the k-mesh.
//
//   Sigma_c(k) = -IFFT2[ FFT2(P_c - refP_c) * VR_c ](k)        (one channel c = (i,j))
//
// Semantics follow src/jax_hf/utils.py:140-173 with the ifftshift phase already
// absorbed into VR (src/jax_hf/problem.py:88-98).  One workgroup owns one
// channel's whole (n1 x n2) plane in LDS: load (strided gather from the
// (n1, n2, nb, nb) layout, fused refP subtraction and weighted channel moment),
// forward row + column passes, VR multiply with the -1/(n1 n2) inverse scale
// folded in, inverse column + row passes, store (fused conjugate mirror in the
// Hermitian-packed mode).  HBM traffic is P + VR in, Sigma out: nothing else.
//
// Each 1D transform is a mixed-radix Stockham autosort done in place in LDS:
// every thread first pulls all the butterfly inputs it owns into registers,
// the workgroup synchronises, then every thread writes its outputs.  Twiddles
// come from per-axis tables W_N^m = exp(-2 pi i m / N) computed on the host in
// extended precision (so the FP64 path is exact to rounding).
#include "hfx_common.h"
#include "hfx_internal.h"

namespace hfx {

// Math constants of the small DFT kernels (cos / sin of 2 pi q / R).
#define HFX_S3 0.86602540378443864676   // sin(2 pi / 3) = sqrt(3)/2
#define HFX_C5A 0.30901699437494742410  // cos(2 pi / 5)
#define HFX_C5B -0.80901699437494742410 // cos(4 pi / 5)
#define HFX_S5A 0.95105651629515357212  // sin(2 pi / 5)
#define HFX_S5B 0.58778525229247312917  // sin(4 pi / 5)
#define HFX_R2 0.70710678118654752440   // sqrt(1/2)

// In-place R-point DFT of v[0..R), forward sign exp(-2 pi i rm/R) unless inv.
template <typename T, int R>
__device__ __forceinline__ void small_dft(cx<T> (&v)[R], bool inv, const cx<T>* __restrict__ tw, int N) {
  const bool fwd = !inv;
  if constexpr (R == 2) {
    cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (R == 3) {
    const T s = (T)HFX_S3;
    cx<T> t1 = v[1] + v[2];
    cx<T> t2 = v[0] - (T)0.5 * t1;
    cx<T> t3 = s * rot(v[1] - v[2], fwd);  // -i s (v1 - v2) forward
    v[0] = v[0] + t1;
    v[1] = t2 + t3;
    v[2] = t2 - t3;
  } else if constexpr (R == 4) {
    cx<T> a = v[0] + v[2], b = v[0] - v[2];
    cx<T> c = v[1] + v[3], d = rot(v[1] - v[3], fwd);
    v[0] = a + c;
    v[2] = a - c;
    v[1] = b + d;
    v[3] = b - d;
  } else if constexpr (R == 5) {
    const T c1 = (T)HFX_C5A, c2 = (T)HFX_C5B, s1 = (T)HFX_S5A, s2 = (T)HFX_S5B;
    cx<T> a1 = v[1] + v[4], a2 = v[2] + v[3];
    cx<T> b1 = v[1] - v[4], b2 = v[2] - v[3];
    cx<T> x0 = v[0];
    cx<T> r1 = x0 + c1 * a1 + c2 * a2;
    cx<T> r2 = x0 + c2 * a1 + c1 * a2;
    cx<T> i1 = rot(s1 * b1 + s2 * b2, fwd);
    cx<T> i2 = rot(s2 * b1 - s1 * b2, fwd);
    v[0] = x0 + a1 + a2;
    v[1] = r1 + i1;
    v[4] = r1 - i1;
    v[2] = r2 + i2;
    v[3] = r2 - i2;
  } else if constexpr (R == 8) {
    // even/odd split into two 4-point DFTs and one layer of W8 twiddles
    cx<T> e0 = v[0] + v[4], e1 = v[0] - v[4], e2 = v[2] + v[6], e3 = rot(v[2] - v[6], fwd);
    cx<T> E0 = e0 + e2, E2 = e0 - e2, E1 = e1 + e3, E3 = e1 - e3;
    cx<T> o0 = v[1] + v[5], o1 = v[1] - v[5], o2 = v[3] + v[7], o3 = rot(v[3] - v[7], fwd);
    cx<T> O0 = o0 + o2, O2 = o0 - o2, O1 = o1 + o3, O3 = o1 - o3;
    const T r = (T)HFX_R2;
    // W8^1 = r(1 -+ i), W8^2 = -+ i, W8^3 = r(-1 -+ i)
    cx<T> w1 = fwd ? mk<T>(r * (O1.re + O1.im), r * (O1.im - O1.re)) : mk<T>(r * (O1.re - O1.im), r * (O1.im + O1.re));
    cx<T> w2 = rot(O2, fwd);
    cx<T> w3 = fwd ? mk<T>(r * (O3.im - O3.re), -r * (O3.re + O3.im)) : mk<T>(-r * (O3.re + O3.im), r * (O3.re - O3.im));
    v[0] = E0 + O0;
    v[4] = E0 - O0;
    v[1] = E1 + w1;
    v[5] = E1 - w1;
    v[2] = E2 + w2;
    v[6] = E2 - w2;
    v[3] = E3 + w3;
    v[7] = E3 - w3;
  } else {
    // generic odd radix: O(R^2) with W_R^q = W_N^{q N / R} from the table
    cx<T> out[R];
    const int step = N / R;
#pragma unroll
    for (int m = 0; m < R; ++m) {
      cx<T> acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cx<T> w = tw[((r * m) % R) * step];
        if (inv) w = conj(w);
        acc = acc + v[r] * w;
      }
      out[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = out[m];
  }
}

// One Stockham radix-R stage over `nlines` independent lines of length N.
// Element q of line l lives at buf[l * lstride + q * estride].
template <typename T, int R, int BLOCK, int MAXE>
__device__ __forceinline__ void fft_stage(cx<T>* buf, int nlines, int N, int lstride, int estride,
                                          bool lines_fast, int Ns, const cx<T>* __restrict__ tw,
                                          bool inv) {
  constexpr int MAXIT = (MAXE + R * BLOCK - 1) / (R * BLOCK);
  const int NR = N / R;
  const int total = nlines * NR;
  const int tstep = N / (Ns * R);
  cx<T> v[MAXIT][R];
#pragma unroll
  for (int m = 0; m < MAXIT; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (it < total) {
      int line, j;
      if (lines_fast) { line = it % nlines; j = it / nlines; }
      else { line = it / NR; j = it - line * NR; }
      const int k = j % Ns;
      const cx<T>* src = buf + line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cx<T> x = src[(j + r * NR) * estride];
        if (r > 0 && k > 0) {
          cx<T> w = tw[r * k * tstep];
          x = inv ? mulc(x, w) : x * w;
        }
        v[m][r] = x;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MAXIT; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (it < total) {
      int line, j;
      if (lines_fast) { line = it % nlines; j = it / nlines; }
      else { line = it / NR; j = it - line * NR; }
      const int k = j % Ns;
      small_dft<T, R>(v[m], inv, tw, N);
      const int base = (j / Ns) * Ns * R + k;
      cx<T>* dst = buf + line * lstride;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[(base + r * Ns) * estride] = v[m][r];
    }
  }
  __syncthreads();
}

// One Stockham stage of a runtime radix R (a prime factor above 7), computed
// output by output.  Output m of butterfly j (k = j mod Ns) is
//   y = sum_r x[j + r N/R] W_N^{r (k tstep + m N/R)},
// the stage twiddle and the R-point DFT folded into one table entry per term
// (W_R^{r m} = W_N^{r m N/R}); conjugated for the inverse.  O(N R) per line
// instead of O(N log N): meant for the odd prime factor of an unusual mesh.
template <typename T, int BLOCK, int MAXE>
__device__ void fft_stage_any(cx<T>* buf, int nlines, int N, int lstride, int estride, bool lines_fast, int Ns,
                              int R, const cx<T>* __restrict__ tw, bool inv) {
  constexpr int MAXO = (MAXE + BLOCK - 1) / BLOCK;
  const int NR = N / R, total = nlines * N, tstep = N / (Ns * R);
  cx<T> y[MAXO];
#pragma unroll
  for (int m = 0; m < MAXO; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (it < total) {
      int line, q;
      if (lines_fast) { line = it % nlines; q = it / nlines; }
      else { line = it / N; q = it - line * N; }
      // q = (j / Ns) Ns R + mm Ns + k with k < Ns, mm < R
      const int k = q % Ns, t = q / Ns, mm = t % R, j = (t / R) * Ns + k;
      const int step = (k * tstep + mm * NR) % N;
      const cx<T>* src = buf + line * lstride;
      cx<T> acc = src[j * estride];
      int ex = 0;
      for (int r = 1; r < R; ++r) {
        ex += step;
        if (ex >= N) ex -= N;
        const cx<T> w = tw[ex], x = src[(j + r * NR) * estride];
        acc = acc + (inv ? mulc(x, w) : x * w);
      }
      y[m] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MAXO; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (it < total) {
      int line, q;
      if (lines_fast) { line = it % nlines; q = it / nlines; }
      else { line = it / N; q = it - line * N; }
      buf[line * lstride + q * estride] = y[m];
    }
  }
  __syncthreads();
}

template <typename T, int BLOCK, int MAXE>
__device__ void fft_lines(cx<T>* buf, int nlines, int N, int lstride, int estride, bool lines_fast,
                          int nst, const int* rad, const cx<T>* __restrict__ tw, bool inv) {
  int Ns = 1;
  for (int s = 0; s < nst; ++s) {
    const int R = rad[s];
    switch (R) {
      case 2: fft_stage<T, 2, BLOCK, MAXE>(buf, nlines, N, lstride, estride, lines_fast, Ns, tw, inv); break;
      case 3: fft_stage<T, 3, BLOCK, MAXE>(buf, nlines, N, lstride, estride, lines_fast, Ns, tw, inv); break;
      case 4: fft_stage<T, 4, BLOCK, MAXE>(buf, nlines, N, lstride, estride, lines_fast, Ns, tw, inv); break;
      case 5: fft_stage<T, 5, BLOCK, MAXE>(buf, nlines, N, lstride, estride, lines_fast, Ns, tw, inv); break;
      case 7: fft_stage<T, 7, BLOCK, MAXE>(buf, nlines, N, lstride, estride, lines_fast, Ns, tw, inv); break;
      case 8: fft_stage<T, 8, BLOCK, MAXE>(buf, nlines, N, lstride, estride, lines_fast, Ns, tw, inv); break;
      default:
        fft_stage_any<T, BLOCK, MAXE>(buf, nlines, N, lstride, estride, lines_fast, Ns, R, tw, inv);
        break;
    }
    Ns *= R;
  }
}

template <typename T>
struct ExchArgs {
  const cx<T>* P;
  const cx<T>* refP;   // may be null
  const cx<T>* VR;     // (n1, n2, vb, vb)
  const int* chans;    // channel list (i * nb + j)
  const int* vmap;     // full VR: VR plane of channel c is vmap[c] (null: plane c); planes that
                       // repeat (a V depending only on the layer pair) are then read once from HBM
  const double* w2d;   // (n1, n2) weights for the channel moment, may be null
  cx<T>* Sigma;        // (n1, n2, nb, nb)
  cx<double>* rho;     // (nb, nb) weighted channel moments, may be null
  int nchan, herm, vfull, xcd_remap;
  int nsolve;          // planes = nsolve * nchan; solve s reads / writes at s * plane_stride
  size_t plane_stride; // elements between consecutive solves' P / refP / Sigma (Nk * nb^2)
  int planar;          // 1: P / refP / Sigma are slot planes (nsolve * nchan, n1, n2), plane s * nchan + t
                       //    holding channel chans[t]; VR (vfull) is channel-major (nb^2, n1, n2)
  int mode;   // 0: exchange convolution; 1: forward transform of w * V (VR preparation)
  int shift;  // mode 0: ifftshift the result on store; mode 1: apply the ifftshift phase
  int conjv;  // mode 0: multiply by conj(VR) (the adjoint / cross-correlation)
  FftGeom<T> g;
};
static_assert(offsetof(ExchArgs<double>, shift) < offsetof(ExchArgs<double>, conjv),
              "exchange_launch initialises ExchArgs positionally: mode, shift, conjv");

// LDS row pitch: odd, so lanes walking down a column (stride = pitch) and
// lanes walking along a row (stride 1) both spread over all 64 banks.
__host__ __device__ constexpr int lds_pitch(int n2) { return (n2 % 2 == 0) ? n2 + 1 : n2; }

// ---- shared pieces of both kernel flavours --------------------------------
// Channel of this workgroup; for a batch of solves also moves the P / refP /
// Sigma pointers to this workgroup's solve (VR is shared by the batch).  In
// the planar layout the pointers move to this workgroup's planes, after which
// k-point e of the channel is element e (exch_ix / exch_vix).
template <typename T>
__device__ __forceinline__ int vr_chan(const ExchArgs<T>& a, int c) {
  return a.vmap ? a.vmap[c] : c;
}

template <typename T>
__device__ __forceinline__ int exch_channel_at(ExchArgs<T>& a, int b) {
  const int s = b / a.nchan;
  const int c = a.chans[b - s * a.nchan];
  if (a.planar) {
    const size_t nk = (size_t)a.g.n1 * a.g.n2, off = (size_t)b * nk;
    a.P += off;
    if (a.refP) a.refP += off;
    a.Sigma += off;
    if (a.vfull) a.VR += (size_t)vr_chan(a, c) * nk;
  } else if (s > 0) {
    const size_t off = (size_t)s * a.plane_stride;
    a.P += off;
    if (a.refP) a.refP += off;
    a.Sigma += off;
  }
  return c;
}

template <typename T, int BLOCK>
__device__ __forceinline__ int exch_channel(ExchArgs<T>& a) {
  int b = blockIdx.x;
  if (a.xcd_remap) {  // consecutive planes on one XCD (blocks round-robin over 8 XCDs)
    const int per = (a.nchan * a.nsolve) >> 3;
    b = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  return exch_channel_at(a, b);
}

// Slot planes: the plane's P / refP / Sigma offsets need only the plane index.
// Returns the channel (VR plane, moment slot) without using it, so that a
// kernel can issue its first P loads before the channel lookup returns; the
// caller applies the VR offset (exch_plane_vr) once it needs the channel.
template <typename T>
__device__ __forceinline__ int exch_plane_begin(ExchArgs<T>& a) {
  int b = blockIdx.x;
  if (a.xcd_remap) {  // consecutive planes on one XCD (blocks round-robin over 8 XCDs)
    const int per = (a.nchan * a.nsolve) >> 3;
    b = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  const int s = b / a.nchan;
  const int c = a.chans[b - s * a.nchan];
  const size_t off = (size_t)b * a.g.n1 * a.g.n2;
  a.P += off;
  if (a.refP) a.refP += off;
  a.Sigma += off;
  return c;
}
template <typename T>
__device__ __forceinline__ void exch_plane_vr(ExchArgs<T>& a, int c) {
  if (a.vfull) a.VR += (size_t)vr_chan(a, c) * a.g.n1 * a.g.n2;
}

// Index of channel c at k-point e in P / refP / Sigma, and in VR.
template <typename T>
__device__ __forceinline__ size_t exch_ix(const ExchArgs<T>& a, size_t e, int c, int nb2) {
  return a.planar ? e : e * nb2 + c;
}
template <typename T>
__device__ __forceinline__ size_t exch_vix(const ExchArgs<T>& a, size_t e, int c, int nb2) {
  return (a.planar || !a.vfull) ? e : e * nb2 + vr_chan(a, c);
}

template <typename T, int BLOCK>
__device__ __forceinline__ void exch_load(const ExchArgs<T>& a, cx<T>* plane, int n1, int n2, int pitch,
                                          int c, double* red) {
  const int nb = a.g.nb, nb2 = nb * nb, N = n1 * n2;
  const int ci = c / nb, cj = c - ci * nb;
  double mre = 0.0, mim = 0.0;
  for (int e = threadIdx.x; e < N; e += BLOCK) {
    const size_t ix = exch_ix(a, (size_t)e, c, nb2);
    cx<T> v = a.P[ix];
    if (a.refP) v = v - a.refP[ix];
    const int k1 = e / n2, k2 = e - k1 * n2;
    if (a.mode == 1) {  // VR preparation: w[q] V[q] (weight at the q index, problem.py:88)
      plane[k1 * pitch + k2] = (T)a.w2d[e] * v;
      continue;
    }
    plane[k1 * pitch + k2] = v;
    if (a.w2d) {
      const double w = a.w2d[e];
      mre += w * (double)v.re;
      mim += w * (double)v.im;
    }
  }
  if (a.rho && a.mode == 0) {
    double acc[2] = {mre, mim};
    block_sum<2>(acc, red);
    if (threadIdx.x == 0) {
      a.rho[c] = cx<double>{acc[0], acc[1]};
      if (a.herm && ci != cj) a.rho[cj * nb + ci] = cx<double>{acc[0], -acc[1]};
    }
  } else {
    __syncthreads();
  }
}

template <typename T, int BLOCK>
__device__ __forceinline__ void exch_multiply(const ExchArgs<T>& a, cx<T>* plane, int n1, int n2, int pitch,
                                              int c) {
  const int nb2 = a.g.nb * a.g.nb, N = n1 * n2;
  const T scale = (T)(-1.0 / (double)N);
  for (int e = threadIdx.x; e < N; e += BLOCK) {
    const int k1 = e / n2, k2 = e - k1 * n2;
    const cx<T> vr = a.VR[exch_vix(a, (size_t)e, c, nb2)];
    cx<T>& x = plane[k1 * pitch + k2];
    x = scale * (a.conjv ? mulc(x, vr) : x * vr);
  }
  __syncthreads();
}

template <typename T, int BLOCK>
__device__ __forceinline__ void exch_store(const ExchArgs<T>& a, const cx<T>* plane, int n1, int n2, int pitch,
                                           int c) {
  const int nb = a.g.nb, nb2 = nb * nb, N = n1 * n2;
  const int ci = c / nb, cj = c - ci * nb;
  const bool mirror = a.herm && ci != cj && !a.planar;
  const int cm = cj * nb + ci;
  const int s1 = n1 / 2, s2 = n2 / 2;
  for (int e = threadIdx.x; e < N; e += BLOCK) {
    const int k1 = e / n2, k2 = e - k1 * n2;
    cx<T> v = plane[k1 * pitch + k2];
    size_t eo = (size_t)e;
    if (a.mode == 1) {
      if (a.shift) {  // VR * exp(2 pi i (k1 s1 / n1 + k2 s2 / n2))  (problem.py:90-98)
        const cx<T> ph = a.g.tw1[(k1 * s1) % n1] * a.g.tw2[(k2 * s2) % n2];
        v = mulc(v, ph);
      }
    } else if (a.shift) {  // ifftshift on store: out[i] = in[(i + n//2) mod n]
      const int o1 = (k1 - s1 + n1) % n1, o2 = (k2 - s2 + n2) % n2;
      eo = (size_t)o1 * n2 + o2;
    }
    a.Sigma[exch_ix(a, eo, c, nb2)] = v;
    if (mirror) a.Sigma[eo * nb2 + cm] = conj(v);
  }
}

// ---- generic (runtime mesh) kernel -----------------------------------------
template <typename T, int BLOCK, int MAXE>
__global__ __launch_bounds__(BLOCK) void exch_conv_kernel(ExchArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* plane = reinterpret_cast<cx<T>*>(smem_raw);
  __shared__ double red[2 * (BLOCK / 64)];
  const int n1 = a.g.n1, n2 = a.g.n2, pitch = lds_pitch(n2);
  const int c = exch_channel<T, BLOCK>(a);
  exch_load<T, BLOCK>(a, plane, n1, n2, pitch, c, red);
  fft_lines<T, BLOCK, MAXE>(plane, n1, n2, pitch, 1, true, a.g.nst2, a.g.rad2, a.g.tw2, false);
  fft_lines<T, BLOCK, MAXE>(plane, n2, n1, 1, pitch, true, a.g.nst1, a.g.rad1, a.g.tw1, false);
  if (a.mode == 0) {
    exch_multiply<T, BLOCK>(a, plane, n1, n2, pitch, c);
    fft_lines<T, BLOCK, MAXE>(plane, n2, n1, 1, pitch, true, a.g.nst1, a.g.rad1, a.g.tw1, true);
    fft_lines<T, BLOCK, MAXE>(plane, n1, n2, pitch, 1, true, a.g.nst2, a.g.rad2, a.g.tw2, true);
  }
  exch_store<T, BLOCK>(a, plane, n1, n2, pitch, c);
}

// ---- compile-time mesh kernels ----------------------------------------------
// Radix sequence of a length-N axis (8s first, then 4, 2, 3, 5, 7), matching
// the host plan's factorisation.
constexpr int ct_rad(int N, int s) {
  int k = 0;
  while (N % 8 == 0) { if (k == s) return 8; N /= 8; ++k; }
  const int rs[5] = {4, 2, 3, 5, 7};
  for (int q = 0; q < 5; ++q)
    while (N % rs[q] == 0) { if (k == s) return rs[q]; N /= rs[q]; ++k; }
  return 0;
}
constexpr int ct_nst(int N) {
  int k = 0;
  while (ct_rad(N, k) != 0) ++k;
  return k;
}

template <typename T, int R, int N, int NL, int LS, int ES, int BLOCK, int NS, bool INV>
__device__ __forceinline__ void ct_stage(cx<T>* buf, const cx<T>* __restrict__ tw) {
  constexpr int NR = N / R, ITEMS = NL * NR, MAXIT = (ITEMS + BLOCK - 1) / BLOCK, TSTEP = N / (NS * R);
  cx<T> v[MAXIT][R];
#pragma unroll
  for (int m = 0; m < MAXIT; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (ITEMS % BLOCK == 0 || it < ITEMS) {
      const int line = it % NL, j = it / NL, k = j % NS;
      const cx<T>* src = buf + line * LS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cx<T> x = src[(j + r * NR) * ES];
        if (NS > 1 && r > 0) {
          const cx<T> w = tw[r * k * TSTEP];
          x = INV ? mulc(x, w) : x * w;
        }
        v[m][r] = x;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MAXIT; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (ITEMS % BLOCK == 0 || it < ITEMS) {
      const int line = it % NL, j = it / NL, k = j % NS;
      small_dft<T, R>(v[m], INV, tw, N);
      const int base = (j / NS) * NS * R + k;
      cx<T>* dst = buf + line * LS;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[(base + r * NS) * ES] = v[m][r];
    }
  }
  __syncthreads();
}

template <typename T, int N, int NL, int LS, int ES, int BLOCK, bool INV, int S = 0, int NS = 1>
__device__ __forceinline__ void ct_lines(cx<T>* buf, const cx<T>* __restrict__ tw) {
  if constexpr (S < ct_nst(N)) {
    constexpr int R = ct_rad(N, S);
    ct_stage<T, R, N, NL, LS, ES, BLOCK, NS, INV>(buf, tw);
    ct_lines<T, N, NL, LS, ES, BLOCK, INV, S + 1, NS * R>(buf, tw);
  }
}

template <typename T, int N1, int N2, int BLOCK>
__global__ __launch_bounds__(BLOCK) void exch_ct_kernel(ExchArgs<T> a) {
  constexpr int PITCH = lds_pitch(N2);
  __shared__ cx<T> plane[N1 * PITCH];
  __shared__ double red[2 * (BLOCK / 64)];
  const int c = exch_channel<T, BLOCK>(a);
  exch_load<T, BLOCK>(a, plane, N1, N2, PITCH, c, red);
  ct_lines<T, N2, N1, PITCH, 1, BLOCK, false>(plane, a.g.tw2);   // rows (along k2)
  ct_lines<T, N1, N2, 1, PITCH, BLOCK, false>(plane, a.g.tw1);   // columns (along k1)
  if (a.mode == 0) {
    exch_multiply<T, BLOCK>(a, plane, N1, N2, PITCH, c);
    ct_lines<T, N1, N2, 1, PITCH, BLOCK, true>(plane, a.g.tw1);
    ct_lines<T, N2, N1, PITCH, 1, BLOCK, true>(plane, a.g.tw2);
  }
  exch_store<T, BLOCK>(a, plane, N1, N2, PITCH, c);
}

constexpr int EXCH_BLOCK = 1024;
// Largest plane (n1 * n2) held in LDS: 159 KiB, leaving room for the static
// reduction scratch inside the 160 KiB per-workgroup limit of gfx950.
constexpr int MAXE_F64 = 10176;
constexpr int MAXE_F32 = 20352;

// ---- v2: two composite-radix Stockham stages per axis ------------------------
// Row pass: stage 1 loads straight from HBM (fused refP subtraction and channel
// moment), stage 2 in LDS.  Column pass: stage 1 in LDS, stage 2 multiplies by
// VR (prefetched into registers before the stage barrier) on the way out.
// Inverse passes mirror this; the last inverse row stage stores straight to HBM
// (with the Hermitian mirror).  Twiddle tables live in LDS.  7 LDS round trips
// per channel instead of 15.

// Composite sizes R = A * B are done in registers by the four-step split
// X[ka + A kb] = DFT_B over b of { W_R^{b ka} DFT_A over a of x[B a + b] }.
constexpr int v2_split(int R) {
  return R == 6 ? 2 : R == 9 ? 3 : R == 10 ? 2 : R == 12 ? 3 : R == 15 ? 3 : R == 16 ? 4 : 0;
}

// W_R^m = exp(-+ 2 pi i m / R) for the radices that dft_any splits, as
// compile-time constants (their indices are constants once the split loops
// unroll): no table reads for the butterflies' internal twiddles.
template <typename T, int R>
__device__ __forceinline__ cx<T> croot(int m, bool inv) {
  if constexpr (R == 6) {
    constexpr double c[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
    constexpr double s[6] = {0.0, 0.8660254037844386467637, 0.8660254037844386467637, 0.0, -0.8660254037844386467637, -0.8660254037844386467637};
    return cx<T>{(T)c[m], (T)(inv ? s[m] : -s[m])};
  }
  if constexpr (R == 9) {
    constexpr double c[9] = {1.0, 0.7660444431189780352024, 0.1736481776669303488517, -0.5, -0.9396926207859083840541, -0.9396926207859083840541, -0.5, 0.1736481776669303488517, 0.7660444431189780352024};
    constexpr double s[9] = {0.0, 0.6427876096865393263226, 0.9848077530122080593667, 0.8660254037844386467637, 0.3420201433256687330441, -0.3420201433256687330441, -0.8660254037844386467637, -0.9848077530122080593667, -0.6427876096865393263226};
    return cx<T>{(T)c[m], (T)(inv ? s[m] : -s[m])};
  }
  if constexpr (R == 10) {
    constexpr double c[10] = {1.0, 0.8090169943749474241023, 0.3090169943749474241023, -0.3090169943749474241023, -0.8090169943749474241023, -1.0, -0.8090169943749474241023, -0.3090169943749474241023, 0.3090169943749474241023, 0.8090169943749474241023};
    constexpr double s[10] = {0.0, 0.5877852522924731291687, 0.9510565162951535721164, 0.9510565162951535721164, 0.5877852522924731291687, 0.0, -0.5877852522924731291687, -0.9510565162951535721164, -0.9510565162951535721164, -0.5877852522924731291687};
    return cx<T>{(T)c[m], (T)(inv ? s[m] : -s[m])};
  }
  if constexpr (R == 12) {
    constexpr double c[12] = {1.0, 0.8660254037844386467637, 0.5, 0.0, -0.5, -0.8660254037844386467637, -1.0, -0.8660254037844386467637, -0.5, 0.0, 0.5, 0.8660254037844386467637};
    constexpr double s[12] = {0.0, 0.5, 0.8660254037844386467637, 1.0, 0.8660254037844386467637, 0.5, 0.0, -0.5, -0.8660254037844386467637, -1.0, -0.8660254037844386467637, -0.5};
    return cx<T>{(T)c[m], (T)(inv ? s[m] : -s[m])};
  }
  if constexpr (R == 15) {
    constexpr double c[15] = {1.0, 0.9135454576426008955021, 0.6691306063588582138263, 0.3090169943749474241023, -0.1045284632676534713998, -0.5, -0.8090169943749474241023, -0.9781476007338056379286, -0.9781476007338056379286, -0.8090169943749474241023, -0.5, -0.1045284632676534713998, 0.3090169943749474241023, 0.6691306063588582138263, 0.9135454576426008955021};
    constexpr double s[15] = {0.0, 0.406736643075800207754, 0.7431448254773942350147, 0.9510565162951535721164, 0.9945218953682733369227, 0.8660254037844386467637, 0.5877852522924731291687, 0.2079116908177593371017, -0.2079116908177593371017, -0.5877852522924731291687, -0.8660254037844386467637, -0.9945218953682733369227, -0.9510565162951535721164, -0.7431448254773942350147, -0.406736643075800207754};
    return cx<T>{(T)c[m], (T)(inv ? s[m] : -s[m])};
  }
  if constexpr (R == 16) {
    constexpr double c[16] = {1.0, 0.9238795325112867561282, 0.7071067811865475244008, 0.3826834323650897717285, 0.0, -0.3826834323650897717285, -0.7071067811865475244008, -0.9238795325112867561282, -1.0, -0.9238795325112867561282, -0.7071067811865475244008, -0.3826834323650897717285, 0.0, 0.3826834323650897717285, 0.7071067811865475244008, 0.9238795325112867561282};
    constexpr double s[16] = {0.0, 0.3826834323650897717285, 0.7071067811865475244008, 0.9238795325112867561282, 1.0, 0.9238795325112867561282, 0.7071067811865475244008, 0.3826834323650897717285, 0.0, -0.3826834323650897717285, -0.7071067811865475244008, -0.9238795325112867561282, -1.0, -0.9238795325112867561282, -0.7071067811865475244008, -0.3826834323650897717285};
    return cx<T>{(T)c[m], (T)(inv ? s[m] : -s[m])};
  }
  return cx<T>{(T)1, (T)0};
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft_any(cx<T> (&v)[R], const cx<T>* tw, int NT) {
  constexpr int A = v2_split(R);
  if constexpr (A == 0) {
    small_dft<T, R>(v, INV, tw, NT);
  } else {
    constexpr int B = R / A;
    cx<T> y[B][A];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      cx<T> t[A];
#pragma unroll
      for (int q = 0; q < A; ++q) t[q] = v[B * q + b];
      dft_any<T, A, INV>(t, tw, NT);
#pragma unroll
      for (int q = 0; q < A; ++q) {
        cx<T> x = t[q];
        if (b > 0 && q > 0) x = x * croot<T, R>((b * q) % R, INV);
        y[b][q] = x;
      }
    }
#pragma unroll
    for (int q = 0; q < A; ++q) {
      cx<T> t[B];
#pragma unroll
      for (int b = 0; b < B; ++b) t[b] = y[b][q];
      dft_any<T, B, INV>(t, tw, NT);
#pragma unroll
      for (int b = 0; b < B; ++b) v[q + A * b] = t[b];
    }
  }
}

// radices whose butterflies use compile-time twiddles only (no table reads)
constexpr bool tw_free_radix(int R) {
  return R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || v2_split(R) != 0;
}

// first (row-pass) and column-pass radix pairs per mesh size
constexpr int v2_row_a(int N) {
  return N == 96 ? 12 : N == 90 ? 10 : N == 72 ? 8 : N == 64 ? 8 : N == 60 ? 6 : N == 48 ? 6 : N == 36 ? 6
       : N == 32 ? 8 : N == 24 ? 6 : N == 16 ? 4 : N == 12 ? 4 : 0;
}
constexpr int v2_col_a(int N) {
  return N == 96 ? 8 : N == 90 ? 9 : N == 72 ? 9 : N == 64 ? 8 : N == 60 ? 10 : N == 48 ? 8 : N == 36 ? 6
       : N == 32 ? 4 : N == 24 ? 4 : N == 16 ? 4 : N == 12 ? 3 : 0;
}

// One stage.  SRC 1: inputs from HBM (row pass, lines = rows k1, positions = k2).
// DST 1: outputs to HBM (same addressing).  VRM: multiply outputs by VR
// (column pass, lines = columns k2, positions = k1).
template <typename T, int R, int N, int NL, int LS, int ES, int BLOCK, int NS, bool INV, int SRC, int DST,
          bool VRM, int N2, bool PL>
__device__ __forceinline__ void v2_stage(cx<T>* plane, const cx<T>* tw, const ExchArgs<T>& a, int c,
                                         double& mre, double& mim) {
  constexpr int NR = N / R, ITEMS = NL * NR, MAXIT = (ITEMS + BLOCK - 1) / BLOCK, TSTEP = N / (NS * R);
  const int nb = a.g.nb, nb2 = nb * nb;
  cx<T> v[MAXIT][R];
  cx<T> vr[VRM ? MAXIT : 1][VRM ? R : 1];
#pragma unroll
  for (int m = 0; m < MAXIT; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (ITEMS % BLOCK == 0 || it < ITEMS) {
      // planar HBM stages: lanes walk along the row (coalesced); otherwise across lines
      constexpr bool JF = PL && (SRC == 1 || DST == 1);
      const int line = JF ? it / NR : it % NL, j = JF ? it % NR : it / NL, k = j % NS;
      if constexpr (SRC == 1) {
        // all R loads of the item first, then the optional terms: the optional
        // branches would otherwise split the loop and serialise the loads
        // (one HBM round trip per element)
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const size_t e = (size_t)line * N2 + j + r * NR;
          v[m][r] = a.P[PL ? e : e * nb2 + c];
        }
        if (a.refP) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const size_t e = (size_t)line * N2 + j + r * NR;
            v[m][r] = v[m][r] - a.refP[PL ? e : e * nb2 + c];
          }
        }
        if (a.w2d) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const double w = a.w2d[(size_t)line * N2 + j + r * NR];
            mre += w * (double)v[m][r].re;
            mim += w * (double)v[m][r].im;
          }
        }
        static_assert(SRC != 1 || NS == 1, "the HBM stage is the first stage of its pass");
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          cx<T> x = plane[line * LS + (j + r * NR) * ES];
          if (NS > 1 && r > 0) {
            const cx<T> w = tw[r * k * TSTEP];
            x = INV ? mulc(x, w) : x * w;
          }
          v[m][r] = x;
        }
      }
      if constexpr (VRM) {   // prefetch VR at the output positions (k1 = base + r NS, k2 = line)
        const int base = (j / NS) * NS * R + k;
        const int vc = a.vfull ? vr_chan(a, c) : 0, vstride = a.vfull ? nb2 : 1;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const size_t e = (size_t)(base + r * NS) * N2 + line;
          vr[m][r] = a.VR[PL ? e : e * vstride + vc];
        }
      }
    }
  }
  if constexpr (SRC == 0) __syncthreads();
#pragma unroll
  for (int m = 0; m < MAXIT; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (ITEMS % BLOCK == 0 || it < ITEMS) {
      constexpr bool JF = PL && (SRC == 1 || DST == 1);
      const int line = JF ? it / NR : it % NL, j = JF ? it % NR : it / NL, k = j % NS;
      dft_any<T, R, INV>(v[m], tw, N);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int pos = base + r * NS;
        cx<T> y = v[m][r];
        if constexpr (VRM) {
          const T scale = (T)(-1.0 / (double)(N * N2));
          y = scale * (a.conjv ? mulc(y, vr[m][r]) : y * vr[m][r]);
        }
        if constexpr (DST == 1) {
          const size_t e = (size_t)line * N2 + pos;
          if constexpr (PL) {
            a.Sigma[e] = y;
            continue;
          }
          a.Sigma[e * nb2 + c] = y;
          const int ci = c / nb, cj = c - ci * nb;
          if (a.herm && ci != cj) a.Sigma[e * nb2 + cj * nb + ci] = conj(y);
        } else {
          plane[line * LS + pos * ES] = y;
        }
      }
    }
  }
  if constexpr (DST == 0) __syncthreads();
}

template <typename T, int N1, int N2, int BLOCK, bool PL>
__global__ __launch_bounds__(BLOCK) void exch_v2_kernel(ExchArgs<T> a) {
  constexpr int PITCH = lds_pitch(N2);
  constexpr int RA2 = v2_row_a(N2), RB2 = N2 / RA2;   // row pass (along k2)
  constexpr int RA1 = v2_col_a(N1), RB1 = N1 / RA1;   // column pass (along k1)
  __shared__ cx<T> plane[N1 * PITCH];
  __shared__ cx<T> tw1[N1], tw2[N2];
  __shared__ double red[2 * (BLOCK / 64)];
  double mre = 0.0, mim = 0.0;
  int c;
  if constexpr (PL) {
    // slot planes: the first stage's P loads go out before the channel and the
    // twiddle tables arrive (as in exch_v3_kernel)
    static_assert(tw_free_radix(RA2), "the first row stage must not read the twiddle table");
    c = exch_plane_begin<T>(a);
    constexpr int T1 = (N1 + BLOCK - 1) / BLOCK, T2 = (N2 + BLOCK - 1) / BLOCK;
    cx<T> t1r[T1], t2r[T2];
#pragma unroll
    for (int m = 0; m < T1; ++m)
      if (threadIdx.x + m * BLOCK < N1) t1r[m] = a.g.tw1[threadIdx.x + m * BLOCK];
#pragma unroll
    for (int m = 0; m < T2; ++m)
      if (threadIdx.x + m * BLOCK < N2) t2r[m] = a.g.tw2[threadIdx.x + m * BLOCK];
    v2_stage<T, RA2, N2, N1, PITCH, 1, BLOCK, 1, false, 1, 0, false, N2, PL>(plane, tw2, a, c, mre, mim);
    asm volatile("" : "+v"(c));
#pragma unroll
    for (int m = 0; m < T1; ++m) {
      asm volatile("" : "+v"(t1r[m].re), "+v"(t1r[m].im));
      if (threadIdx.x + m * BLOCK < N1) tw1[threadIdx.x + m * BLOCK] = t1r[m];
    }
#pragma unroll
    for (int m = 0; m < T2; ++m) {
      asm volatile("" : "+v"(t2r[m].re), "+v"(t2r[m].im));
      if (threadIdx.x + m * BLOCK < N2) tw2[threadIdx.x + m * BLOCK] = t2r[m];
    }
    exch_plane_vr<T>(a, c);
    __syncthreads();
  } else {
    c = exch_channel<T, BLOCK>(a);
    for (int m = threadIdx.x; m < N1; m += BLOCK) tw1[m] = a.g.tw1[m];
    for (int m = threadIdx.x; m < N2; m += BLOCK) tw2[m] = a.g.tw2[m];
    __syncthreads();
    // forward rows: lines = k1 rows (N1 of them), stride PITCH; positions along k2
    v2_stage<T, RA2, N2, N1, PITCH, 1, BLOCK, 1, false, 1, 0, false, N2, PL>(plane, tw2, a, c, mre, mim);
  }
  v2_stage<T, RB2, N2, N1, PITCH, 1, BLOCK, RA2, false, 0, 0, false, N2, PL>(plane, tw2, a, c, mre, mim);
  // forward columns (+ VR multiply), inverse columns
  v2_stage<T, RA1, N1, N2, 1, PITCH, BLOCK, 1, false, 0, 0, false, N2, PL>(plane, tw1, a, c, mre, mim);
  v2_stage<T, RB1, N1, N2, 1, PITCH, BLOCK, RA1, false, 0, 0, true, N2, PL>(plane, tw1, a, c, mre, mim);
  v2_stage<T, RA1, N1, N2, 1, PITCH, BLOCK, 1, true, 0, 0, false, N2, PL>(plane, tw1, a, c, mre, mim);
  v2_stage<T, RB1, N1, N2, 1, PITCH, BLOCK, RA1, true, 0, 0, false, N2, PL>(plane, tw1, a, c, mre, mim);
  // inverse rows, last stage straight to HBM
  v2_stage<T, RA2, N2, N1, PITCH, 1, BLOCK, 1, true, 0, 0, false, N2, PL>(plane, tw2, a, c, mre, mim);
  v2_stage<T, RB2, N2, N1, PITCH, 1, BLOCK, RA2, true, 0, 1, false, N2, PL>(plane, tw2, a, c, mre, mim);
  if (a.rho) {
    double acc[2] = {mre, mim};
    block_sum<2>(acc, red);
    if (threadIdx.x == 0) {
      const int nb = a.g.nb, ci = c / nb, cj = c - ci * nb;
      a.rho[c] = cx<double>{acc[0], acc[1]};
      if (a.herm && ci != cj) a.rho[cj * nb + ci] = cx<double>{acc[0], -acc[1]};
    }
  }
}

// ---- v3 (slot planes): the VR multiply between a forward and an inverse column
// stage of the same radix ---------------------------------------------------
// The forward column pass runs its radices in the order (A, B) and the inverse
// pass in the order (B, A).  The last forward stage (radix B, stride A) leaves
// each thread holding the outputs j + A r (r < B) of its column, in natural
// order; the first inverse stage (radix B, stride 1) reads exactly the inputs
// j + A r.  So one stage does the forward radix-B butterflies, the VR multiply
// and the inverse radix-B butterflies in registers: one LDS round trip and one
// barrier fewer than v2 (six in all).  The fused stage's VR values are loaded
// right after the first row stage, as global loads: the barriers wait only for
// LDS (lgkmcnt), so that HBM read overlaps the second row and the first column
// stages.
template <typename T>
__device__ __forceinline__ cx<T> ld_global(const cx<T>* p, size_t i) {
  typedef T v2t __attribute__((ext_vector_type(2)));
  const v2t x = ((const __attribute__((address_space(1))) v2t*)(p + i))[0];
  return cx<T>{x.x, x.y};
}

template <typename T, int R, int N, int BLOCK, int A, int MAXIT>
__device__ __forceinline__ void v3_fused_cols(cx<T>* plane, const cx<T>* tw, const cx<T> (&vr)[MAXIT][R],
                                              T scale, bool conjv) {
  constexpr int PITCH = lds_pitch(N), ITEMS = N * A;
  static_assert(A * R == N, "column radices");
  cx<T> v[MAXIT][R];
#pragma unroll
  for (int m = 0; m < MAXIT; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (ITEMS % BLOCK == 0 || it < ITEMS) {
      const int line = it % N, j = it / N;   // column k2 = line, butterfly j < A
#pragma unroll
      for (int r = 0; r < R; ++r) {
        cx<T> x = plane[line + (j + r * A) * PITCH];
        if (r > 0) x = x * tw[r * j];          // forward stride-A stage: W_N^{r j}
        v[m][r] = x;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MAXIT; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (ITEMS % BLOCK == 0 || it < ITEMS) {
      const int line = it % N, j = it / N;
      dft_any<T, R, false>(v[m], tw, N);        // outputs at k1 = j + A r
#pragma unroll
      for (int r = 0; r < R; ++r) v[m][r] = scale * (conjv ? mulc(v[m][r], vr[m][r]) : v[m][r] * vr[m][r]);
      dft_any<T, R, true>(v[m], tw, N);         // inverse stride-1 stage: outputs at j R + r
#pragma unroll
      for (int r = 0; r < R; ++r) plane[line + (j * R + r) * PITCH] = v[m][r];
    }
  }
  __syncthreads();
}

template <typename T, int N, int BLOCK>
__global__ __launch_bounds__(BLOCK) void exch_v3_kernel(ExchArgs<T> a) {
  constexpr int PITCH = lds_pitch(N);
  constexpr int RA2 = v2_row_a(N), RB2 = N / RA2;   // row pass
  constexpr int RA1 = v2_col_a(N), RB1 = N / RA1;   // column pass: forward (RA1, RB1), inverse (RB1, RA1)
  constexpr int FITEMS = N * RA1, FMAX = (FITEMS + BLOCK - 1) / BLOCK;
  __shared__ cx<T> plane[N * PITCH];
  __shared__ cx<T> tw[N];
  __shared__ double red[2 * (BLOCK / 64)];
  // The slot-plane offsets need only the plane index, so the first stage's P
  // loads go out at once; the channel (for VR and the moment) and the twiddle
  // table arrive meanwhile (the first stage's radix reads no table twiddles)
  static_assert(tw_free_radix(RA2), "the first row stage must not read the twiddle table");
  int c = exch_plane_begin<T>(a);
  constexpr int TWM = (N + BLOCK - 1) / BLOCK;
  cx<T> twr[TWM];
#pragma unroll
  for (int m = 0; m < TWM; ++m)
    if (threadIdx.x + m * BLOCK < N) twr[m] = a.g.tw1[threadIdx.x + m * BLOCK];
  double mre = 0.0, mim = 0.0;
  // forward rows: lines = k1 rows, positions along k2; the first stage loads from HBM
  v2_stage<T, RA2, N, N, PITCH, 1, BLOCK, 1, false, 1, 0, false, N, true>(plane, tw, a, c, mre, mim);
  // opaque after the stage: the compiler would otherwise consume the channel
  // and the table (and so wait for them) before the stage's loads
  asm volatile("" : "+v"(c));
#pragma unroll
  for (int m = 0; m < TWM; ++m) {
    asm volatile("" : "+v"(twr[m].re), "+v"(twr[m].im));
    if (threadIdx.x + m * BLOCK < N) tw[threadIdx.x + m * BLOCK] = twr[m];
  }
  exch_plane_vr<T>(a, c);
  __syncthreads();
  // VR of the fused column stage, loaded after P (loads retire in order, so a
  // wait for P would also wait for anything issued before it)
  cx<T> vr[FMAX][RB1];
#pragma unroll
  for (int m = 0; m < FMAX; ++m) {
    const int it = threadIdx.x + m * BLOCK;
    if (FITEMS % BLOCK == 0 || it < FITEMS) {
      const int line = it % N, j = it / N;
#pragma unroll
      for (int r = 0; r < RB1; ++r) vr[m][r] = ld_global(a.VR, (size_t)(j + RA1 * r) * N + line);
    }
  }
  v2_stage<T, RB2, N, N, PITCH, 1, BLOCK, RA2, false, 0, 0, false, N, true>(plane, tw, a, c, mre, mim);
  // columns: forward radix RA1, then [forward RB1, x VR, inverse RB1], then inverse RA1
  v2_stage<T, RA1, N, N, 1, PITCH, BLOCK, 1, false, 0, 0, false, N, true>(plane, tw, a, c, mre, mim);
  v3_fused_cols<T, RB1, N, BLOCK, RA1, FMAX>(plane, tw, vr, (T)(-1.0 / ((double)N * N)), a.conjv != 0);
  v2_stage<T, RA1, N, N, 1, PITCH, BLOCK, RB1, true, 0, 0, false, N, true>(plane, tw, a, c, mre, mim);
  // inverse rows, last stage straight to HBM
  v2_stage<T, RA2, N, N, PITCH, 1, BLOCK, 1, true, 0, 0, false, N, true>(plane, tw, a, c, mre, mim);
  v2_stage<T, RB2, N, N, PITCH, 1, BLOCK, RA2, true, 0, 1, false, N, true>(plane, tw, a, c, mre, mim);
  if (a.rho) {
    double acc[2] = {mre, mim};
    block_sum<2>(acc, red);
    if (threadIdx.x == 0) {
      const int nb = a.g.nb, ci = c / nb, cj = c - ci * nb;
      a.rho[c] = cx<double>{acc[0], acc[1]};
      if (a.herm && ci != cj) a.rho[cj * nb + ci] = cx<double>{acc[0], -acc[1]};
    }
  }
}

template <int N>
constexpr int ct_block() { return N * N >= 4096 ? 1024 : (N * N >= 1024 ? 512 : 256); }

#define HFX_CT_SIZES(X) X(12) X(16) X(24) X(32) X(36) X(48) X(60) X(64) X(72) X(90) X(96)

// Workgroup sizes of the two-stage kernels, per variant (HFX_EXCH_V2_* / V3_*).
template <int N>
constexpr int v2_block(int variant) {
  constexpr int big[3] = {1024, 768, 512}, mid[3] = {512, 384, 256}, small[3] = {256, 192, 128};
  const int i = (variant == HFX_EXCH_V2_B || variant == HFX_EXCH_V3_B) ? 1
              : (variant == HFX_EXCH_V2_C || variant == HFX_EXCH_V3_C) ? 2 : 0;
  return N * N >= 4096 ? big[i] : (N * N >= 1024 ? mid[i] : small[i]);
}

// Default variant per mesh (used when the plan's variant is HFX_EXCH_AUTO): at
// 96 x 96 the fused-column form on slot planes (29.7 vs 31.1 us for 256 FP64
// channels); other meshes keep the two-stage kernel until timed
// (hfx_exchange_autotune picks per plan).
constexpr int v2_default(int N, bool planar) {
  return N == 96 ? (planar ? HFX_EXCH_V3_B : HFX_EXCH_V2_B) : HFX_EXCH_V2_A;
}

template <typename T, int N, int V>
static int launch_v2(const ExchArgs<T>& a, hipStream_t s) {
  constexpr int B = v2_block<N>(V);
  if (a.planar && V >= HFX_EXCH_V3_A) {
    hipLaunchKernelGGL((exch_v3_kernel<T, N, B>), dim3(a.nchan * a.nsolve), dim3(B), 0, s, a);
    return check_launch("exch_v3_kernel");
  }
  if (a.planar)
    hipLaunchKernelGGL((exch_v2_kernel<T, N, N, B, true>), dim3(a.nchan * a.nsolve), dim3(B), 0, s, a);
  else
    hipLaunchKernelGGL((exch_v2_kernel<T, N, N, B, false>), dim3(a.nchan * a.nsolve), dim3(B), 0, s, a);
  return check_launch("exch_v2_kernel");
}

// ---- HBM-staged exchange (planes larger than LDS) ---------------------------
// Three passes over tiles that fit in LDS; the channel's own positions of the
// output array hold the intermediate, so no workspace is needed:
//   rows          a tile of k1-rows, forward along k2 (loads P - refP, or w V
//                 in mode 1) -> X;
//   columns       a tile of k2-columns, forward along k1, then (mode 0) the VR
//                 multiply with -1/(n1 n2) and the inverse along k1, or (mode 1)
//                 the ifftshift phase -> X (the result in mode 1);
//   inverse rows  (mode 0) inverse along k2 -> Sigma, with the Hermitian mirror.
// An ifftshift of the result permutes k1 in the column pass and k2 in the
// inverse row pass, so each pass writes only its own tile's lines.  The
// weighted channel moment is a separate per-plane reduction run first.
// Traffic per channel: P in, VR in, 5 plane-sized transfers of the
// intermediate, Sigma out (vs P + VR + Sigma for the LDS-resident kernels).
struct XsTiles {
  int tr, ntr;   // rows per row tile, row tiles per plane
  int tc, ntc;   // columns per column tile, column tiles per plane
};

template <typename T, int BLOCK, int MAXE>
__global__ __launch_bounds__(BLOCK) void exst_rows_kernel(ExchArgs<T> a, XsTiles t, int inv) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
  const int n1 = a.g.n1, n2 = a.g.n2, pitch = lds_pitch(n2), nb = a.g.nb, nb2 = nb * nb;
  const int b = blockIdx.x / t.ntr, r0 = (blockIdx.x - b * t.ntr) * t.tr;
  const int nr = min(t.tr, n1 - r0), cnt = nr * n2;
  const int c = exch_channel_at(a, b);
  for (int it = threadIdx.x; it < cnt; it += BLOCK) {
    const int r = it / n2, k2 = it - r * n2;
    const size_t e = (size_t)(r0 + r) * n2 + k2, ix = exch_ix(a, e, c, nb2);
    cx<T> v;
    if (inv) {
      v = a.Sigma[ix];
    } else {
      v = a.P[ix];
      if (a.refP) v = v - a.refP[ix];
      if (a.mode == 1) v = (T)a.w2d[e] * v;   // w[q] V[q] (problem.py:88)
    }
    tile[r * pitch + k2] = v;
  }
  __syncthreads();
  fft_lines<T, BLOCK, MAXE>(tile, nr, n2, pitch, 1, true, a.g.nst2, a.g.rad2, a.g.tw2, inv != 0);
  const int ci = c / nb, cj = c - ci * nb, cm = cj * nb + ci, s2 = n2 / 2;
  const bool mirror = inv && a.herm && ci != cj && !a.planar;
  for (int it = threadIdx.x; it < cnt; it += BLOCK) {
    const int r = it / n2, k2 = it - r * n2;
    const int o2 = (inv && a.shift) ? (k2 - s2 + n2) % n2 : k2;   // ifftshift along k2
    const size_t e = (size_t)(r0 + r) * n2 + o2;
    const cx<T> v = tile[r * pitch + k2];
    a.Sigma[exch_ix(a, e, c, nb2)] = v;
    if (mirror) a.Sigma[e * nb2 + cm] = conj(v);
  }
}

template <typename T, int BLOCK, int MAXE>
__global__ __launch_bounds__(BLOCK) void exst_cols_kernel(ExchArgs<T> a, XsTiles t) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cx<T>* tile = reinterpret_cast<cx<T>*>(smem_raw);
  const int n1 = a.g.n1, n2 = a.g.n2, nb2 = a.g.nb * a.g.nb;
  const int b = blockIdx.x / t.ntc, c0 = (blockIdx.x - b * t.ntc) * t.tc;
  const int nc = min(t.tc, n2 - c0), cp = lds_pitch(nc), cnt = n1 * nc;
  const int c = exch_channel_at(a, b);
  for (int it = threadIdx.x; it < cnt; it += BLOCK) {
    const int k1 = it / nc, q = it - k1 * nc;
    tile[k1 * cp + q] = a.Sigma[exch_ix(a, (size_t)k1 * n2 + c0 + q, c, nb2)];
  }
  __syncthreads();
  fft_lines<T, BLOCK, MAXE>(tile, nc, n1, 1, cp, true, a.g.nst1, a.g.rad1, a.g.tw1, false);
  if (a.mode == 0) {
    const T scale = (T)(-1.0 / ((double)n1 * n2));
    for (int it = threadIdx.x; it < cnt; it += BLOCK) {
      const int k1 = it / nc, q = it - k1 * nc;
      const cx<T> vr = a.VR[exch_vix(a, (size_t)k1 * n2 + c0 + q, c, nb2)];
      cx<T>& x = tile[k1 * cp + q];
      x = scale * (a.conjv ? mulc(x, vr) : x * vr);
    }
    __syncthreads();
    fft_lines<T, BLOCK, MAXE>(tile, nc, n1, 1, cp, true, a.g.nst1, a.g.rad1, a.g.tw1, true);
  }
  const int s1 = n1 / 2, s2 = n2 / 2;
  for (int it = threadIdx.x; it < cnt; it += BLOCK) {
    const int k1 = it / nc, q = it - k1 * nc, k2 = c0 + q;
    cx<T> v = tile[k1 * cp + q];
    int o1 = k1;
    if (a.mode == 1) {
      if (a.shift) {  // VR * exp(2 pi i (k1 s1 / n1 + k2 s2 / n2))  (problem.py:90-98)
        const cx<T> ph = a.g.tw1[(int)(((long long)k1 * s1) % n1)] * a.g.tw2[(int)(((long long)k2 * s2) % n2)];
        v = mulc(v, ph);
      }
    } else if (a.shift) {
      o1 = (k1 - s1 + n1) % n1;   // ifftshift along k1
    }
    a.Sigma[exch_ix(a, (size_t)o1 * n2 + k2, c, nb2)] = v;
  }
}

// Weighted channel moment rho_c = sum_k w_k (P - refP)_c(k) of the staged path
// (the LDS kernels fuse it into their load), one workgroup per plane.
template <typename T, int BLOCK>
__global__ __launch_bounds__(BLOCK) void exst_moment_kernel(ExchArgs<T> a) {
  __shared__ double red[2 * (BLOCK / 64)];
  const int c = exch_channel_at(a, blockIdx.x);
  const int nb = a.g.nb, nb2 = nb * nb, N = a.g.n1 * a.g.n2;
  double acc[2] = {0.0, 0.0};
  for (int e = threadIdx.x; e < N; e += BLOCK) {
    const size_t ix = exch_ix(a, (size_t)e, c, nb2);
    cx<T> v = a.P[ix];
    if (a.refP) v = v - a.refP[ix];
    const double w = a.w2d[e];
    acc[0] += w * (double)v.re;
    acc[1] += w * (double)v.im;
  }
  block_sum<2>(acc, red);
  if (threadIdx.x == 0) {
    const int ci = c / nb, cj = c - ci * nb;
    a.rho[c] = cx<double>{acc[0], acc[1]};
    if (a.herm && ci != cj) a.rho[cj * nb + ci] = cx<double>{acc[0], -acc[1]};
  }
}

// Largest mesh axis the staged path takes: one row (odd pitch) and one column
// of the plane must fit in a tile.
template <typename T>
constexpr int staged_max_axis() { return (sizeof(T) == 8 ? MAXE_F64 : MAXE_F32) - 1; }

template <typename T, int BLOCK, int MAXE>
static int launch_staged_form(const ExchArgs<T>& a, hipStream_t s) {
  const int n1 = a.g.n1, n2 = a.g.n2, planes = a.nchan * a.nsolve, p2 = lds_pitch(n2);
  XsTiles t;
  t.tr = max(1, min(n1, MAXE / p2));
  t.ntr = (n1 + t.tr - 1) / t.tr;
  int tc = min(16, n2);
  while (tc > 1 && n1 * lds_pitch(tc) > MAXE) --tc;
  t.tc = tc;
  t.ntc = (n2 + tc - 1) / tc;
  const size_t rows_lds = (size_t)t.tr * p2 * sizeof(cx<T>);
  const size_t cols_lds = (size_t)n1 * lds_pitch(tc) * sizeof(cx<T>);
  auto krows = exst_rows_kernel<T, BLOCK, MAXE>;
  auto kcols = exst_cols_kernel<T, BLOCK, MAXE>;
  static bool attr_set = false;
  if (!attr_set) {
    HFX_HIP_TRY(hipFuncSetAttribute((const void*)krows, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(MAXE * sizeof(cx<T>))));
    HFX_HIP_TRY(hipFuncSetAttribute((const void*)kcols, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(MAXE * sizeof(cx<T>))));
    attr_set = true;
  }
  if (a.mode == 0 && a.rho) {
    hipLaunchKernelGGL((exst_moment_kernel<T, 256>), dim3(planes), dim3(256), 0, s, a);
    if (int rc = check_launch("exst_moment_kernel")) return rc;
  }
  hipLaunchKernelGGL(krows, dim3((unsigned)planes * t.ntr), dim3(BLOCK), rows_lds, s, a, t, 0);
  if (int rc = check_launch("exst_rows_kernel")) return rc;
  hipLaunchKernelGGL(kcols, dim3((unsigned)planes * t.ntc), dim3(BLOCK), cols_lds, s, a, t);
  if (int rc = check_launch("exst_cols_kernel")) return rc;
  if (a.mode == 0) {
    hipLaunchKernelGGL(krows, dim3((unsigned)planes * t.ntr), dim3(BLOCK), rows_lds, s, a, t, 1);
    return check_launch("exst_rows_kernel");
  }
  return HFX_OK;
}

template <typename T>
static int launch_staged(const ExchArgs<T>& a, hipStream_t s) {
  constexpr int MAXE = sizeof(T) == 8 ? MAXE_F64 : MAXE_F32;
  constexpr int SMALL = 2048;   // tile elements of the 256-thread form (32 KiB in FP64)
  const int n1 = a.g.n1, n2 = a.g.n2;
  if (n1 > staged_max_axis<T>() || n2 > staged_max_axis<T>())
    return set_error(HFX_EUNSUPPORTED, "exchange: mesh axis beyond %d", staged_max_axis<T>());
  if (lds_pitch(n2) <= SMALL && n1 <= SMALL) return launch_staged_form<T, 256, SMALL>(a, s);
  return launch_staged_form<T, EXCH_BLOCK, MAXE>(a, s);
}

template <typename T>
static int launch_exch(const ExchArgs<T>& a, int variant, hipStream_t s) {
  constexpr int MAXE = sizeof(T) == 8 ? MAXE_F64 : MAXE_F32;
  const int n1 = a.g.n1, n2 = a.g.n2;
  if (n1 == n2) {
    switch (n1) {
#define HFX_CASE(N)                                                                        \
  case N:                                                                                  \
    if ((size_t)N * lds_pitch(N) * sizeof(cx<T>) <= (size_t)MAXE * sizeof(cx<T>)) {        \
      const int v = variant == HFX_EXCH_AUTO ? v2_default(N, a.planar != 0) : variant;     \
      if (a.mode == 0 && a.shift == 0 && v != HFX_EXCH_V1) {                               \
        if (v == HFX_EXCH_V2_B) return launch_v2<T, N, HFX_EXCH_V2_B>(a, s);                \
        if (v == HFX_EXCH_V2_C) return launch_v2<T, N, HFX_EXCH_V2_C>(a, s);                \
        if (v == HFX_EXCH_V3_A) return launch_v2<T, N, HFX_EXCH_V3_A>(a, s);                \
        if (v == HFX_EXCH_V3_B) return launch_v2<T, N, HFX_EXCH_V3_B>(a, s);                \
        if (v == HFX_EXCH_V3_C) return launch_v2<T, N, HFX_EXCH_V3_C>(a, s);                \
        return launch_v2<T, N, HFX_EXCH_V2_A>(a, s);                                       \
      }                                                                                    \
      hipLaunchKernelGGL((exch_ct_kernel<T, N, N, ct_block<N>()>), dim3(a.nchan * a.nsolve),           \
                         dim3(ct_block<N>()), 0, s, a);                                    \
      return check_launch("exch_ct_kernel");                                               \
    }                                                                                      \
    break;
      HFX_CT_SIZES(HFX_CASE)
#undef HFX_CASE
      default: break;
    }
  }
  if (exchange_lds_elems(n1, n2) > MAXE) return launch_staged<T>(a, s);
  const size_t lds = (size_t)n1 * lds_pitch(n2) * sizeof(cx<T>);
  auto kern = exch_conv_kernel<T, EXCH_BLOCK, MAXE>;
  static bool attr_set = false;
  if (!attr_set) {
    HFX_HIP_TRY(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(MAXE * sizeof(cx<T>))));
    attr_set = true;
  }
  hipLaunchKernelGGL(kern, dim3(a.nchan * a.nsolve), dim3(EXCH_BLOCK), lds, s, a);
  return check_launch("exch_conv_kernel");
}

int exchange_launch(const hfx_plan* plan, const void* P, const void* refP, const void* VR, int vfull,
                    const int* chans, int nchan, int herm, const double* w2d, void* Sigma, void* rho,
                    int mode, int shift, int conjv, hipStream_t s, int nsolve, int planar, const int* vmap) {
  if (nchan <= 0 || nsolve <= 0) return HFX_OK;
  const int remap = ((nchan * nsolve) % 8 == 0) ? 1 : 0;
  const size_t stride = (size_t)plan->n1 * plan->n2 * plan->nb * plan->nb;
  if (plan->dtype == HFX_DTYPE_C128) {
    ExchArgs<double> a{(const cx<double>*)P, (const cx<double>*)refP, (const cx<double>*)VR, chans,
                       vfull ? vmap : nullptr, w2d,
                       (cx<double>*)Sigma, (cx<double>*)rho, nchan, herm, vfull, remap, nsolve, stride, planar,
                       /*mode=*/mode, /*shift=*/shift, /*conjv=*/conjv, plan->g64};
    return launch_exch<double>(a, plan->exch_variant, s);
  }
  ExchArgs<float> a{(const cx<float>*)P, (const cx<float>*)refP, (const cx<float>*)VR, chans,
                    vfull ? vmap : nullptr, w2d,
                    (cx<float>*)Sigma, (cx<double>*)rho, nchan, herm, vfull, remap, nsolve, stride, planar,
                    /*mode=*/mode, /*shift=*/shift, /*conjv=*/conjv, plan->g32};
  return launch_exch<float>(a, plan->exch_variant, s);
}

int exchange_max_plane(int dtype) { return dtype == HFX_DTYPE_C128 ? MAXE_F64 : MAXE_F32; }
int exchange_lds_elems(int n1, int n2) { return n1 * lds_pitch(n2); }

}  // namespace hfx
