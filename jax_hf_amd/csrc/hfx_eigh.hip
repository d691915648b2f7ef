// Batched Hermitian eigensolver for nb <= 16: two matrices per 64-lane
// wavefront, each row split over two lanes, everything in registers.
//
// Parallel cyclic Jacobi with the XOR pairing schedule: step M (1 <= M < P,
// P = the power of two >= nb) rotates the disjoint pairs (i, i ^ M); over
// the P - 1 steps of a sweep every pair occurs once.  Row i of A and of V is
// kept in XOR-permuted order, b[t] = A[i][i ^ t] and w[t] = V[i][i ^ t], the
// half h of the row (t = 8h + s, s < 8) on lane 32 m + 16 h + i of matrix m,
// so that for a compile-time M every register index is a constant:
//   * the diagonal is b[0] and the pair element A[i][i^M] is b[M];
//   * the column rotation A <- A J (columns x = i ^ t and x ^ M) mixes b[t]
//     with b[t ^ M]: in the same lane for M < 8, from the other half-row lane
//     (one ds_swizzle, xor 16) for M >= 8; the parameters of column x are read
//     from LDS, where the lane owning row x published them;
//   * the row rotation A <- J^H A needs the partner row's entry for our column
//     i ^ t, which sits at its position t ^ M: one ds_swizzle (xor M, or
//     xor 16 + M when that position is in the other half).
// Rotation of pair (p, q), c = A_pq = |c| e^{i phi}: zeta = (A_qq - A_pp)/(2|c|),
// t = sign(zeta)/(|zeta| + sqrt(1 + zeta^2)), cs = 1/sqrt(1 + t^2), sn = t cs,
// J_pp = J_qq = cs, J_pq = sn e^{i phi}, J_qp = -sn e^{-i phi}; afterwards
// A_pp = A_pp - t|c|, A_qq = A_qq + t|c|, A_pq = 0 exactly.  Threshold
// strategy: a pair rotates only while |A_pq| > tol ||A||_F, tol = eps (below
// eps ||A||_F the element is a perturbation under the rounding of the
// eigenvalues) or a caller's larger relative tolerance; the sweeps end after
// the first sweep without a rotation in either of the wave's matrices; a step
// in which no pair of the wave rotates is skipped.
// Eigenvalues ascending, eigenvectors in the columns of V (the
// reference's jnp.linalg.eigh convention, linalg.py:96-160, reference_scf.py:89).
// With a basis U0 (A = U0^H F U0 for some F) the rotations are accumulated
// onto the rows of U0 instead of the identity: V = U0 W, the eigenvectors of F,
// with no separate product (V may be U0: each wave reads its rows first).
#include "hfx_common.h"
#include "hfx_internal.h"

namespace hfx {

// lane' = lane ^ X within each 32-lane half of the wavefront (ds_swizzle bit mode)
template <int X>
__device__ __forceinline__ double swz(double v) {
  constexpr int pat = (X << 10) | 0x1F;
  const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), pat);
  const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), pat);
  return __hiloint2double(hi, lo);
}

template <int X>
__device__ __forceinline__ float swz(float v) {
  constexpr int pat = (X << 10) | 0x1F;
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), pat));
}

template <int X, typename T>
__device__ __forceinline__ cx<T> swzc(cx<T> v) {
  return cx<T>{swz<X>(v.re), swz<X>(v.im)};
}

template <typename T>
struct J16Lds {
  // per (matrix in the wave, row/column x): J_xx, J_{x^M, x}, the new
  // diagonal entry, and whether pair (x, x^M) rotated
  T jd[2][16];
  T jre[2][16];
  T jim[2][16];
  T dn[2][16];
  int rot[2][16];
};

template <typename T, int M>
__device__ __forceinline__ void j16_step(cx<T> (&b)[8], cx<T> (&w)[8], int m, int h, int i_in, int nb, bool act,
                                         T thr2, int& rot, J16Lds<T>& sh) {
  __builtin_amdgcn_sched_barrier(0);
  // an opaque copy of the row index, so that the compiler does not hoist every
  // step's index arithmetic out of the sweep loop
  int i = i_in;
  asm volatile("" : "+v"(i));
  constexpr int MH = M >> 3, ML = M & 7;   // half and in-half slot of position M
  const int j = i ^ M;
  // parameters, on the lanes holding the first half of each row (b[0] = A_ii)
  const T aqq = swz<M>(b[0].re);                               // partner row's diagonal
  cx<T> cij;                                                   // A[i][j]
  if constexpr (MH) cij = swzc<16>(b[ML]);
  else cij = b[ML];
  int r = 0;
  if (h == 0) {
    const T app = b[0].re;
    T jd = (T)1, dn = app;
    cx<T> jo{(T)0, (T)0};
    if (act && j < nb) {
      const T ac2 = cij.re * cij.re + cij.im * cij.im;
      if (ac2 > thr2) {
        r = 1;
        const bool lo = i < j;          // this row is p, the smaller index of the pair
        // 1/|c| once (zeta and the phase are products with it), cs = rsqrt(1 + t^2):
        // one division and one square root per rotation
        const T rac = rsqrt(ac2);
        const T ac = ac2 * rac;
        const cx<T> c = lo ? cij : conj(cij);
        const T App = lo ? app : aqq, Aqq = lo ? aqq : app;
        const T zeta = (Aqq - App) * ((T)0.5 * rac);
        const T t = (zeta >= (T)0 ? (T)1 : (T)-1) / (fabs(zeta) + sqrt((T)1 + zeta * zeta));
        const T cs = rsqrt((T)1 + t * t), sn = t * cs;
        const T pre = c.re * rac, pim = c.im * rac;
        jd = cs;
        jo = lo ? cx<T>{-sn * pre, sn * pim} : cx<T>{sn * pre, sn * pim};   // J_qp : J_pq
        dn = lo ? App - t * ac : Aqq + t * ac;
      }
    }
    rot |= r;
    sh.jd[m][i] = jd;
    sh.jre[m][i] = jo.re;
    sh.jim[m][i] = jo.im;
    sh.dn[m][i] = dn;
    sh.rot[m][i] = r;
  }
  __syncthreads();
  // no pair of either matrix rotates: the step is the identity (the closing
  // sweep, and most late steps of a warm start, end here)
  if (!__any(r)) return;
  const int ih = i ^ (h << 3);          // the column of local slot s is ih ^ s
  // A <- A J, V <- V J:  new X[t] = J_xx X[t] + J_{x^M, x} X[t ^ M],  x = i ^ t
  cx<T> pb[8], pw[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    if constexpr (MH) {
      pb[s] = swzc<16>(b[s ^ ML]);
      pw[s] = swzc<16>(w[s ^ ML]);
    } else {
      pb[s] = b[s ^ ML];
      pw[s] = w[s ^ ML];
    }
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int x = ih ^ s;
    const T dx = sh.jd[m][x];
    const cx<T> ox{sh.jre[m][x], sh.jim[m][x]};
    b[s] = dx * b[s] + ox * pb[s];
    w[s] = dx * w[s] + ox * pw[s];
    // materialise V here: otherwise the compiler sinks every step's V update
    // to the final store and keeps all steps' parameters live until then
    asm volatile("" : "+v"(w[s].re), "+v"(w[s].im));
  }
  // A <- J^H A: row i with the partner row i ^ M
  const T jdi = sh.jd[m][i];
  const cx<T> cjo{sh.jre[m][i], -sh.jim[m][i]};
  constexpr int XR = MH ? (16 | M) : M;
#pragma unroll
  for (int s = 0; s < 8; ++s) pb[s] = swzc<XR>(b[s ^ ML]);
#pragma unroll
  for (int s = 0; s < 8; ++s) b[s] = jdi * b[s] + cjo * pb[s];
  // exact values on a rotated pair: A_ii = the new diagonal, A_{i, i^M} = 0
  if (sh.rot[m][i]) {
    if (h == 0) b[0] = cx<T>{sh.dn[m][i], (T)0};
    if (h == MH) b[ML] = cx<T>{(T)0, (T)0};
  }
  __syncthreads();
}

template <typename T, int P, int M = 1>
__device__ __forceinline__ void j16_sweep(cx<T> (&b)[8], cx<T> (&w)[8], int m, int h, int i, int nb, bool act,
                                          T thr2, int& rot, J16Lds<T>& sh) {
  j16_step<T, M>(b, w, m, h, i, nb, act, thr2, rot, sh);
  if constexpr (M + 1 < P) j16_sweep<T, P, M + 1>(b, w, m, h, i, nb, act, thr2, rot, sh);
}

template <typename T, int P>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void eigh16_kernel(int nb, size_t nk, const cx<T>* __restrict__ A,
                                                    double* __restrict__ lam, cx<T>* V, const cx<T>* U0,
                                                    int herm_in, double rel_tol) {
  __shared__ J16Lds<T> sh;
  __shared__ double sd[2][16];
  __shared__ int srank[2][16];
  const int lane = threadIdx.x, m = lane >> 5, h = (lane >> 4) & 1, i = lane & 15;
  const size_t k = (size_t)blockIdx.x * 2 + m;
  const bool act = k < nk && i < nb;
  const size_t nb2 = (size_t)nb * nb, base = act ? k * nb2 : 0;
  const int ih = i ^ (h << 3);
  cx<T> b[8], w[8];
  T rs = (T)0;
  // every load issued unconditionally (clamped to an element of this matrix,
  // or of matrix 0 on idle lanes) and the padding selected away afterwards:
  // loads under per-lane branches are waited for one branch at a time
  const int ic = i < nb ? i : 0;
  const bool u0 = U0 != nullptr;
  cx<T> la[8], lt[8], lu[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int x = ih ^ s, xc = x < nb ? x : 0;
    la[s] = A[base + (size_t)ic * nb + xc];
    if (herm_in) lt[s] = A[base + (size_t)xc * nb + ic];
    if (u0) lu[s] = U0[base + (size_t)ic * nb + xc];
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int x = ih ^ s;
    const bool in = act && x < nb;
    cx<T> v = la[s];
    if (herm_in) v = (T)0.5 * (v + conj(lt[s]));
    if (x == i) v.im = (T)0;            // exact Hermitian diagonal
    if (!in) v = cx<T>{(T)0, (T)0};
    b[s] = v;
    if (u0) w[s] = in ? lu[s] : cx<T>{(T)0, (T)0};
    else w[s] = cx<T>{(x == i && act) ? (T)1 : (T)0, (T)0};
    rs += v.re * v.re + v.im * v.im;
  }
  rs += swz<1>(rs);
  rs += swz<2>(rs);
  rs += swz<4>(rs);
  rs += swz<8>(rs);
  rs += swz<16>(rs);                    // ||A||_F^2 of this lane's matrix
  const T eps = sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920929e-07;
  const T tol = rel_tol > (double)eps ? (T)rel_tol : eps;   // rotate while |A_pq| > tol ||A||_F
  const T thr2 = tol * tol * rs;
  for (int sweep = 0; sweep < 30; ++sweep) {
    if (sweep > 0) {
      // A sweep rotates exactly when some off-diagonal element is above the
      // threshold at its start (a pair's element changes only through a
      // rotation, and the first such pair in step order is reached unchanged):
      // test that directly instead of running a closing sweep that rotates
      // nothing.
      T big = (T)0;
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const T a2 = b[s].re * b[s].re + b[s].im * b[s].im;
        big = (s == 0 && h == 0) ? big : fmax(big, a2);   // slot 0 of the first half is the diagonal
      }
      if (!__any(act && big > thr2)) break;
    }
    int rot = 0;
    j16_sweep<T, P>(b, w, m, h, i, nb, act, thr2, rot, sh);
    if (!__any(rot)) break;
  }
  // ascending order: rank_i = #{j : d_j < d_i or (d_j == d_i and j < i)}
  if (h == 0) sd[m][i] = (double)b[0].re;
  __syncthreads();
  const double d = sd[m][i];
  int rk = 0;
  for (int q = 0; q < nb; ++q) {
    const double dq = sd[m][q];
    rk += (dq < d) || (dq == d && q < i);
  }
  if (h == 0) srank[m][i] = rk;
  __syncthreads();
  if (act) {
    if (h == 0) lam[k * nb + rk] = d;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int x = ih ^ s;
      if (x < nb) V[base + (size_t)i * nb + srank[m][x]] = w[s];
    }
  }
}

template <typename T>
static void launch_eigh16(int nb, size_t nk, const cx<T>* A, double* lam, cx<T>* V, const cx<T>* U0, int herm_in,
                          double rel_tol, hipStream_t s) {
  const dim3 grid((unsigned)((nk + 1) / 2));
  if (nb <= 2)
    hipLaunchKernelGGL((eigh16_kernel<T, 2>), grid, dim3(64), 0, s, nb, nk, A, lam, V, U0, herm_in, rel_tol);
  else if (nb <= 4)
    hipLaunchKernelGGL((eigh16_kernel<T, 4>), grid, dim3(64), 0, s, nb, nk, A, lam, V, U0, herm_in, rel_tol);
  else if (nb <= 8)
    hipLaunchKernelGGL((eigh16_kernel<T, 8>), grid, dim3(64), 0, s, nb, nk, A, lam, V, U0, herm_in, rel_tol);
  else
    hipLaunchKernelGGL((eigh16_kernel<T, 16>), grid, dim3(64), 0, s, nb, nk, A, lam, V, U0, herm_in, rel_tol);
}

int eigh16_launch(int dtype, int nb, size_t nk, const void* A, double* lam, void* V, int herm_in, double rel_tol,
                  hipStream_t s, const void* U0) {
  if (dtype == HFX_DTYPE_C128)
    launch_eigh16<double>(nb, nk, (const cx<double>*)A, lam, (cx<double>*)V, (const cx<double>*)U0, herm_in,
                          rel_tol, s);
  else
    launch_eigh16<float>(nb, nk, (const cx<float>*)A, lam, (cx<float>*)V, (const cx<float>*)U0, herm_in, rel_tol,
                         s);
  return check_launch("eigh16_kernel");
}

}  // namespace hfx
