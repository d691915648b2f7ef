// Per-k batched nb x nb complex kernels of the direct-minimisation and SCF
// loops (src/jax_hf/solver.py:34-233, 451-600; reference_scf.py:87-96).
//
// Execution model: one 64-lane wavefront (= one 64-thread workgroup) owns one
// k-point.  Its matrices are staged in LDS with an odd row pitch (nb + 1) so
// row- and column-walking lanes hit distinct banks; lane l owns the matrix
// elements e = l, l + 64, ... (<= 4 for nb <= 16).  Per-k real vectors
// (occupations, eigenvalues, gradients) are float64 in both storage
// precisions; complex matrices are in the storage precision T.  Products
// accumulate in T (FP32 in the mixed form).  Per-k scalar partials go to a
// (nval, Nk) float64 array that hfx_reduce sums in a fixed order.
#include "hfx_common.h"
#include "hfx_internal.h"

namespace hfx {

// Size class NB (16 or 32): the largest nb one kernel instance takes.  NB = 16
// runs one 64-lane wavefront per k-point (with the matrix-core products at
// nb = 16); NB = 32 a 256-thread workgroup per k-point.  Either way every
// thread owns PK_EPL <= 4 matrix elements and a staged matrix has pitch NB + 1.
// The macros below are written in terms of the template parameter NB of the
// function they appear in.
#define PK_NBMAX (NB)
#define PK_PITCH ((NB) + 1)
#define PK_MAT ((NB) * ((NB) + 1))
#define PKB ((NB) <= 16 ? 64 : 256)
#define PK_EPL (((NB) * (NB) + PKB - 1) / PKB)

// state vector of the solve that k-point k belongs to (seg_len = 0: one solve)
__device__ __forceinline__ const double* seg_state(const double* st, size_t k, int seg_len) {
  return seg_len > 0 ? st + (k / (size_t)seg_len) * HFX_ST_LEN : st;
}

// Sum over the workgroup of one value per thread (deterministic order; the
// result is valid in every thread).  Every thread of the workgroup must call it.
template <int BLOCK>
__device__ __forceinline__ double pk_sum(double v) {
  v = wave_sum(v);
  if constexpr (BLOCK == 64) {
    return v;
  } else {
    __shared__ double sh[BLOCK / 64];
    __syncthreads();   // the previous call's readers are done with sh
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) t += sh[w];
    return t;
  }
}
#define PK_SUM(x) pk_sum<PKB>(x)

// Element loop of one nb x nb matrix over the workgroup: element e = (i, j).
// nb equal to the size class: every slot is an element, so the body runs
// without a per-lane guard and the compiler issues its global loads together
// (a load under a guard is waited for before the next guard's load); smaller
// nb: the guarded form.
#define PK_FOR_ELEM(nb, ...)                                                   \
  if (PK_EPL * PKB == PK_NBMAX * PK_NBMAX && (nb) == PK_NBMAX) {               \
    _Pragma("unroll") for (int m_ = 0; m_ < PK_EPL; ++m_) {                    \
      const int e = threadIdx.x + PKB * m_;                                    \
      const int i = e / PK_NBMAX, j = e % PK_NBMAX;                            \
      __VA_ARGS__                                                              \
    }                                                                          \
  } else {                                                                     \
    _Pragma("unroll") for (int m_ = 0; m_ < PK_EPL; ++m_) {                    \
      const int e = threadIdx.x + PKB * m_;                                    \
      if (e < (nb) * (nb)) {                                                   \
        const int i = e / (nb), j = e - i * (nb);                              \
        __VA_ARGS__                                                            \
      }                                                                        \
    }                                                                          \
  }

template <int NB, typename T>
__device__ __forceinline__ void m_load(cx<T>* s, const cx<T>* __restrict__ g, int nb) {
  PK_FOR_ELEM(nb, { s[i * PK_PITCH + j] = g[e]; })
}
template <int NB, typename T>
__device__ __forceinline__ void m_store(cx<T>* __restrict__ g, const cx<T>* s, int nb) {
  PK_FOR_ELEM(nb, { g[e] = s[i * PK_PITCH + j]; })
}

// ---- 16 x 16 complex products on the matrix cores ---------------------------
// One wave computes C = op(A) op(B) with the 16x16x4 MFMA of its precision
// (v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32): K = 16 in four slices,
// four real products per slice (re*re - im*im, re*im + im*re).  Operand lanes:
// A[i = l % 16][k = 4 s + l / 16], B[k = 4 s + l / 16][j = l % 16].  The
// accumulator element r of lane l is C[row][l % 16] with row = l / 16 + 4 r
// (f64) or 4 (l / 16) + r (f32) -- the two instructions lay rows out
// differently.  Each operand element is read from LDS once (the scalar form
// reads it nb times).
typedef double pk_v4d __attribute__((ext_vector_type(4)));
typedef float pk_v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef pk_v4d v4;
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
  typedef pk_v4f v4;
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

template <typename T>
struct MAcc {
  typename Mfma<T>::v4 re, im;
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int r = 0; r < 4; ++r) { re[r] = (T)0; im[r] = (T)0; }
  }
};

// acc += sgn * op(A) op(B) for 16 x 16 LDS matrices (pitch PK_PITCH); all 64 lanes.
template <int NB, typename T, bool AH, bool BH>
__device__ __forceinline__ void mf_gemm16(MAcc<T>& acc, const cx<T>* A, const cx<T>* B, T sgn = (T)1) {
  const int l = threadIdx.x & 63, x = l & 15, kq = l >> 4;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int kk = 4 * s + kq;
    cx<T> a = AH ? conj(A[kk * PK_PITCH + x]) : A[x * PK_PITCH + kk];
    const cx<T> b = BH ? conj(B[x * PK_PITCH + kk]) : B[kk * PK_PITCH + x];
    a.re *= sgn;
    a.im *= sgn;
    acc.re = Mfma<T>::mma(a.re, b.re, acc.re);
    acc.re = Mfma<T>::mma(-a.im, b.im, acc.re);
    acc.im = Mfma<T>::mma(a.re, b.im, acc.im);
    acc.im = Mfma<T>::mma(a.im, b.re, acc.im);
  }
}

template <int NB, typename T>
__device__ __forceinline__ void mf_store16(cx<T>* C, const MAcc<T>& acc) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < 4; ++r) C[Mfma<T>::row(l, r) * PK_PITCH + (l & 15)] = cx<T>{acc.re[r], acc.im[r]};
}

// C = op(A) op(B); op = conjugate transpose when the flag is set.  C must not
// alias A or B.  Ends with a barrier.
template <int NB, typename T, bool AH, bool BH>
__device__ __forceinline__ void m_gemm(cx<T>* C, const cx<T>* A, const cx<T>* B, int nb) {
  if constexpr (NB == 16) {
    if (nb == 16) {
      MAcc<T> acc;
      acc.zero();
      mf_gemm16<NB, T, AH, BH>(acc, A, B);
      mf_store16<NB>(C, acc);
      __syncthreads();
      return;
    }
  }
  PK_FOR_ELEM(nb, {
    cx<T> acc{0, 0};
    for (int l = 0; l < nb; ++l) {
      const cx<T> a = AH ? conj(A[l * PK_PITCH + i]) : A[i * PK_PITCH + l];
      const cx<T> b = BH ? conj(B[j * PK_PITCH + l]) : B[l * PK_PITCH + j];
      acc = acc + a * b;
    }
    C[i * PK_PITCH + j] = acc;
  })
  __syncthreads();
}

// Cayley factor c = (1 + i a) / (1 - i a), a = tau lam / 2  (solver.py:169-175)
template <typename T>
__device__ __forceinline__ cx<T> cayley(double lam, double tau) {
  const T a = (T)(tau * lam * 0.5);
  // (1 + ia)/(1 - ia) = (1 + ia)^2 / (1 + a^2)
  const T d = (T)1 / ((T)1 + a * a);
  return cx<T>{((T)1 - a * a) * d, (T)2 * a * d};
}

// diag(U^dagger Ft U) from the spectral form (solver.py:185-197):
// eps_b = Re sum_j (V M)_bj conj(V_bj),  M = conj(c) (x) Ft_eig (x) c.
// V and Fe are LDS matrices, c per column in cs[], scratch W (LDS).
template <int NB, typename T>
__device__ __forceinline__ void spectral_diag(const cx<T>* V, const cx<T>* Fe, const cx<T>* cs, cx<T>* W,
                                              double* outd, int nb) {
  PK_FOR_ELEM(nb, { W[i * PK_PITCH + j] = conj(cs[i]) * Fe[i * PK_PITCH + j] * cs[j]; })
  __syncthreads();
  // lane b < nb: row b of V * W, dotted with conj(V_b.)
  if ((int)threadIdx.x < nb) {
    const int b = threadIdx.x;
    double acc = 0.0;
    for (int j = 0; j < nb; ++j) {
      cx<T> a{0, 0};
      for (int l = 0; l < nb; ++l) a = a + V[b * PK_PITCH + l] * W[l * PK_PITCH + j];
      const cx<T> v = V[b * PK_PITCH + j];
      acc += (double)a.re * (double)v.re + (double)a.im * (double)v.im;   // Re(a conj(v))
    }
    outd[b] = acc;
  }
  __syncthreads();
}

// ---- batched Hermitian eigensolver: cyclic parallel (round-robin) Jacobi ----
// A (LDS, Hermitian) is diagonalised in place; V (LDS) receives the
// eigenvectors (columns), lam[] (LDS, double) the eigenvalues, ascending.
// For pair (p, q) with c = A_pq = |c| e^{i phi}: zeta = (A_qq - A_pp)/(2|c|),
// t = sign(zeta)/(|zeta| + sqrt(1 + zeta^2)), cs = 1/sqrt(1+t^2), sn = t cs,
// J_pp = J_qq = cs, J_pq = sn e^{i phi}, J_qp = -sn e^{-i phi}; A <- J^H A J.
template <int NB, typename T>
__device__ void m_eigh(cx<T>* A, cx<T>* V, cx<T>* Vt, double* lam, int nb, double* red) {
  __shared__ int partner[PK_NBMAX];
  __shared__ T jd[PK_NBMAX];
  __shared__ cx<T> jo[PK_NBMAX];
  __shared__ T dnew[PK_NBMAX];
  const T eps = sizeof(T) == 8 ? (T)2.220446049250313e-16 : (T)1.1920929e-07;
  PK_FOR_ELEM(nb, { V[i * PK_PITCH + j] = cx<T>{i == j ? (T)1 : (T)0, (T)0}; })
  // exact hermitian diagonal
  if ((int)threadIdx.x < nb) A[threadIdx.x * PK_PITCH + threadIdx.x].im = (T)0;
  __syncthreads();
  const int m = nb + (nb & 1);    // even round-robin size (dummy index = nb when odd)
  cx<T> newA[PK_EPL], newV[PK_EPL];
  __shared__ int rf[PK_NBMAX];    // 1: the index is in a pair rotated this step
  // Threshold Jacobi: a pair (p, q) is rotated only while
  //   |a_pq| > eps * max(sqrt(|a_pp a_qq|), 0.01 * ||A||_F / sqrt(nb)),
  // otherwise it is left as is (below the rounding of the diagonal); the
  // sweeps stop after the first sweep with no rotation.  The relative part
  // keeps small eigenvalues accurate, the absolute floor stops rotations that
  // would only stir rounding noise.  A step touches only the rows / columns
  // of its rotated pairs (A) and their columns (V), and a step with no
  // rotation is skipped: a warm start (nearly diagonal A) and the closing
  // sweep cost little more than the pair tests.
  T floor_abs;
  {
    T tot = 0;
    PK_FOR_ELEM(nb, {
      const cx<T> a = A[i * PK_PITCH + j];
      tot += a.re * a.re + a.im * a.im;
    })
    const double tot_s = PK_SUM((double)tot);
    floor_abs = (T)(0.01 * sqrt(tot_s / (double)nb));
  }
  for (int sweep = 0; sweep < 30 && nb > 1; ++sweep) {
    bool any = false;
    for (int step = 0; step < m - 1; ++step) {
      // pair table for this step: index m-1 meets `step`; others (step+k, step-k) mod (m-1)
      if ((int)threadIdx.x < m / 2) {
        const int k = threadIdx.x;
        int p, q;
        if (k == 0) { p = step; q = m - 1; }
        else { p = (step + k) % (m - 1); q = (step - k + (m - 1)) % (m - 1); }
        if (p > q) { const int t = p; p = q; q = t; }
        if (q < nb) {
          partner[p] = q;
          partner[q] = p;
          const cx<T> c = A[p * PK_PITCH + q];
          const T ac = sqrt(c.re * c.re + c.im * c.im);
          const T app = A[p * PK_PITCH + p].re, aqq = A[q * PK_PITCH + q].re;
          if (ac > eps * fmax(sqrt(fabs(app * aqq)), floor_abs)) {
            rf[p] = rf[q] = 1;
            const T zeta = (aqq - app) / ((T)2 * ac);
            const T t = (zeta >= (T)0 ? (T)1 : (T)-1) / (fabs(zeta) + sqrt((T)1 + zeta * zeta));
            const T cs = (T)1 / sqrt((T)1 + t * t), sn = t * cs;
            const cx<T> ph{c.re / ac, c.im / ac};   // e^{i phi}
            jd[p] = cs;
            jd[q] = cs;
            jo[p] = cx<T>{-sn * ph.re, sn * ph.im};   // J_qp = -sn e^{-i phi}
            jo[q] = cx<T>{sn * ph.re, sn * ph.im};    // J_pq =  sn e^{i phi}
            dnew[p] = app - t * ac;
            dnew[q] = aqq + t * ac;
          } else {
            rf[p] = rf[q] = 0;
          }
        } else if (p < nb) {   // paired with the dummy: untouched this step
          rf[p] = 0;
        }
      }
      __syncthreads();
      // every wave forms the same mask from the table (lane l reads rf[l])
      const int ln = threadIdx.x & 63;
      const unsigned long long rot = __ballot(ln < nb && rf[ln] != 0);
      if (rot == 0ull) continue;   // uniform: nothing rotates in this step
      any = true;
      // A'_ij = sum_{a in {i,i'}, b in {j,j'}} conj(J_ai) A_ab J_bj ;  V' = V J
      // (only where i or j belongs to a rotated pair; V' only in rotated columns)
      PK_FOR_ELEM(nb, {
        const bool ri = (rot >> i) & 1ull, rj = (rot >> j) & 1ull;
        if (!(ri || rj)) continue;
        const int ip = ri ? partner[i] : i, jp = rj ? partner[j] : j;
        const T di = ri ? jd[i] : (T)1, dj = rj ? jd[j] : (T)1;
        const cx<T> oi = ri ? jo[i] : cx<T>{0, 0}, oj = rj ? jo[j] : cx<T>{0, 0};
        cx<T> r;
        if (i == j) {
          r = cx<T>{dnew[i], (T)0};
        } else if (ip == j) {
          r = cx<T>{0, 0};
        } else {
          const cx<T> aij = A[i * PK_PITCH + j];
          cx<T> x = dj * aij;
          if (rj) x = x + A[i * PK_PITCH + jp] * oj;
          cx<T> y{0, 0};
          if (ri) {
            y = dj * A[ip * PK_PITCH + j];
            if (rj) y = y + A[ip * PK_PITCH + jp] * oj;
          }
          r = di * x + cmul(oi, y);
        }
        newA[m_] = r;
        if (rj) newV[m_] = dj * V[i * PK_PITCH + j] + V[i * PK_PITCH + jp] * oj;
      })
      __syncthreads();
      PK_FOR_ELEM(nb, {
        const bool ri = (rot >> i) & 1ull, rj = (rot >> j) & 1ull;
        if (ri || rj) A[i * PK_PITCH + j] = newA[m_];
        if (rj) V[i * PK_PITCH + j] = newV[m_];
      })
      __syncthreads();
    }
    if (!any) break;
  }
  // ascending order: rank_i = #{j : l_j < l_i or (l_j == l_i and j < i)}
  __shared__ int rank[PK_NBMAX];
  if ((int)threadIdx.x < nb) {
    const int i = threadIdx.x;
    const T li = A[i * PK_PITCH + i].re;
    int r = 0;
    for (int j = 0; j < nb; ++j) {
      const T lj = A[j * PK_PITCH + j].re;
      r += (lj < li) || (lj == li && j < i);
    }
    rank[i] = r;
    lam[r] = (double)li;
  }
  __syncthreads();
  PK_FOR_ELEM(nb, { Vt[i * PK_PITCH + rank[j]] = V[i * PK_PITCH + j]; })
  __syncthreads();
  PK_FOR_ELEM(nb, { V[i * PK_PITCH + j] = Vt[i * PK_PITCH + j]; })
  __syncthreads();
  (void)red;
}

// ============================ kernels ========================================

// Batched eigh: A (nk, nb, nb) Hermitian -> lam (nk, nb) ascending, V (nk, nb, nb).
// herm_in: hermitise the input first (reference_scf.py:89 eigh(_herm(F))).
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_eigh_kernel(int nb, const cx<T>* __restrict__ A, double* __restrict__ lam,
                                                     cx<T>* V, const cx<T>* U0, int herm_in) {
  __shared__ cx<T> sA[PK_MAT], sV[PK_MAT], sT[PK_MAT];
  __shared__ double sl[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  m_load<NB>(sA, A + k * nb2, nb);
  __syncthreads();
  if (herm_in) {
    PK_FOR_ELEM(nb, { sT[i * PK_PITCH + j] = (T)0.5 * (sA[i * PK_PITCH + j] + conj(sA[j * PK_PITCH + i])); })
    __syncthreads();
    PK_FOR_ELEM(nb, { sA[i * PK_PITCH + j] = sT[i * PK_PITCH + j]; })
    __syncthreads();
  }
  m_eigh<NB>(sA, sV, sT, sl, nb, nullptr);
  if ((int)threadIdx.x < nb) lam[k * nb + threadIdx.x] = sl[threadIdx.x];
  if (U0) {   // eigenvectors in the basis U0: V = U0 W (V may be U0)
    m_load<NB>(sT, U0 + k * nb2, nb);
    __syncthreads();
    m_gemm<NB, T, false, false>(sA, sT, sV, nb);
    m_store<NB>(V + k * nb2, sA, nb);
  } else {
    m_store<NB>(V + k * nb2, sV, nb);
  }
}

// P = herm(Q diag(p) Q^dagger)  (solver.py:42-44, 462; utils.py:473-474)
// p from pvec, or (when pvec is NULL) p = sigmoid((mu - eps)/T) and written to pout.
// With Pref: part[k] = sum |P - Pref|^2 as well (the density residual of a
// Roothaan step from Pref, reference_scf.py:291-294).
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_density_kernel(int nb, const cx<T>* __restrict__ Q,
                                                        const double* __restrict__ pvec,
                                                        const double* __restrict__ eps, const double* __restrict__ mu,
                                                        double T_, double* __restrict__ pout,
                                                        cx<T>* __restrict__ P, const cx<T>* __restrict__ Pref,
                                                        double* __restrict__ part) {
  __shared__ cx<T> sQ[PK_MAT], sX[PK_MAT], sP[PK_MAT];
  __shared__ double sp[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  if ((int)threadIdx.x < nb) {
    double pv;
    if (pvec) pv = pvec[k * nb + threadIdx.x];
    else {
      pv = 1.0 / (1.0 + exp(-((*mu) - eps[k * nb + threadIdx.x]) / T_));
      if (pout) pout[k * nb + threadIdx.x] = pv;
    }
    sp[threadIdx.x] = pv;
  }
  m_load<NB>(sQ, Q + k * nb2, nb);
  __syncthreads();
  PK_FOR_ELEM(nb, { sX[i * PK_PITCH + j] = (T)sp[j] * sQ[i * PK_PITCH + j]; })
  __syncthreads();
  m_gemm<NB, T, false, true>(sP, sX, sQ, nb);
  double d2 = 0.0;
  PK_FOR_ELEM(nb, {
    const cx<T> v = (T)0.5 * (sP[i * PK_PITCH + j] + conj(sP[j * PK_PITCH + i]));
    P[k * nb2 + e] = v;
    if (Pref) {
      const cx<T> d = v - Pref[k * nb2 + e];
      d2 += (double)d.re * d.re + (double)d.im * d.im;
    }
  })
  if (Pref) {
    d2 = PK_SUM(d2);
    if (threadIdx.x == 0) part[k] = d2;
  }
}

// CG step 1 (solver.py:472-491): Ft = Q^dagger F Q, eps = Re diag Ft,
// G = skew((p_j - p_i) Ft) offdiag.  part[0][k] = sum |G|^2, part[1][k] = sum eps^2.
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_cg_grad_kernel(int nb, size_t Nk, const cx<T>* __restrict__ Q,
                                                        const cx<T>* __restrict__ F, const double* __restrict__ p,
                                                        cx<T>* __restrict__ Ft, double* __restrict__ eps,
                                                        cx<T>* __restrict__ G, double* __restrict__ part) {
  __shared__ cx<T> sQ[PK_MAT], sF[PK_MAT], sX[PK_MAT], sT[PK_MAT];
  __shared__ double sp[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  m_load<NB>(sQ, Q + k * nb2, nb);
  m_load<NB>(sF, F + k * nb2, nb);
  if ((int)threadIdx.x < nb) sp[threadIdx.x] = p[k * nb + threadIdx.x];
  __syncthreads();
  m_gemm<NB, T, false, false>(sX, sF, sQ, nb);
  m_gemm<NB, T, true, false>(sT, sQ, sX, nb);
  double g2 = 0.0, e2 = 0.0;
  PK_FOR_ELEM(nb, {
    const cx<T> fij = sT[i * PK_PITCH + j];
    Ft[k * nb2 + e] = fij;
    cx<T> g{0, 0};
    if (i != j) {
      const T dij = (T)(sp[j] - sp[i]), dji = (T)(sp[i] - sp[j]);
      const cx<T> a = dij * fij, b = dji * sT[j * PK_PITCH + i];
      g = (T)0.5 * (a - conj(b));
    }
    G[k * nb2 + e] = g;
    g2 += (double)g.re * g.re + (double)g.im * g.im;
    if (i == j) {
      eps[k * nb + i] = (double)fij.re;
      e2 += (double)fij.re * (double)fij.re;
    }
  })
  g2 = PK_SUM(g2);
  e2 = PK_SUM(e2);
  if (threadIdx.x == 0) {
    part[k] = g2;
    part[Nk + k] = e2;
  }
}

// CG step 2 (solver.py:486, 493-519): preconditioned gradients and the six
// Polak-Ribiere inner products per k (weights applied in the reduction).
// st: device scalar state (mu, lam).  Outputs H, gp, hp; part[0..5][k].
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_cg_precond_kernel(int nb, size_t Nk, const cx<T>* __restrict__ G,
                                                           const cx<T>* __restrict__ Gprev,
                                                           const double* __restrict__ eps,
                                                           const double* __restrict__ p,
                                                           const double* __restrict__ gp_prev,
                                                           const double* __restrict__ st_all, double T_,
                                                           cx<T>* __restrict__ H, double* __restrict__ gp,
                                                           double* __restrict__ hp, double* __restrict__ part,
                                                           int seg_len) {
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  const double* st = seg_state(st_all, k, seg_len);
  const double lam = st[HFX_ST_LAM], mu = st[HFX_ST_MU];
  double s_gh = 0, s_ph = 0, s_pp = 0;
  PK_FOR_ELEM(nb, {
    const double gap = eps[k * nb + i] - eps[k * nb + j];
    const T den = (T)sqrt(gap * gap + lam * lam);
    const cx<T> g = G[k * nb2 + e], gpv = Gprev[k * nb2 + e];
    const cx<T> hq{g.re / den, g.im / den};
    H[k * nb2 + e] = hq;
    s_gh += (double)g.re * hq.re + (double)g.im * hq.im;
    s_ph += (double)gpv.re * hq.re + (double)gpv.im * hq.im;
    s_pp += (double)gpv.re * gpv.re + (double)gpv.im * gpv.im;
  })
  double v_gh = 0, v_ph = 0, v_pp = 0;
  if ((int)threadIdx.x < nb) {
    const size_t q = k * nb + threadIdx.x;
    const double pv = p[q];
    const double pl = fmin(fmax(pv, 1e-6), 1.0 - 1e-6);
    const double g = eps[q] + T_ * (log(pl) - log1p(-pl)) - mu;   // solver.py:52-55, 486
    const double ps = fmin(fmax(pv, 1e-8), 1.0 - 1e-8);
    const double h = g * (ps * (1.0 - ps) / T_);
    gp[q] = g;
    hp[q] = h;
    const double gq = gp_prev[q];
    v_gh = g * h;
    v_ph = gq * h;
    v_pp = gq * gq;
  }
  s_gh = PK_SUM(s_gh);
  s_ph = PK_SUM(s_ph);
  s_pp = PK_SUM(s_pp);
  v_gh = PK_SUM(v_gh);
  v_ph = PK_SUM(v_ph);
  v_pp = PK_SUM(v_pp);
  if (threadIdx.x == 0) {
    part[0 * Nk + k] = s_gh;
    part[1 * Nk + k] = s_ph;
    part[2 * Nk + k] = v_gh;
    part[3 * Nk + k] = v_ph;
    part[4 * Nk + k] = s_pp;
    part[5 * Nk + k] = v_pp;
  }
}

// CG step 3 (solver.py:531-546): d = H + beta d_prev (in place in D),
// dp likewise, eigh(herm(i d)) -> V, lam_d, Ft_eig = V^dagger Ft V,
// part[k] = sum |d|^2.
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_cg_direction_kernel(int nb, const cx<T>* __restrict__ H,
                                                             cx<T>* __restrict__ D, const double* __restrict__ hp,
                                                             double* __restrict__ dp, const cx<T>* __restrict__ Ft,
                                                             const double* __restrict__ st, cx<T>* __restrict__ Vout,
                                                             double* __restrict__ lam_d, cx<T>* __restrict__ Fe,
                                                             double* __restrict__ part) {
  __shared__ cx<T> sA[PK_MAT], sV[PK_MAT], sT[PK_MAT], sF[PK_MAT];
  __shared__ double sl[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  const T beta = (T)st[HFX_ST_BETA];
  double dn = 0.0;
  PK_FOR_ELEM(nb, {
    const cx<T> hq = H[k * nb2 + e];
    const cx<T> dv = D[k * nb2 + e];
    const cx<T> d = hq + beta * dv;
    D[k * nb2 + e] = d;
    sT[i * PK_PITCH + j] = d;
    dn += (double)d.re * d.re + (double)d.im * d.im;
  })
  if ((int)threadIdx.x < nb) {
    const size_t q = k * nb + threadIdx.x;
    dp[q] = hp[q] + (double)beta * dp[q];
  }
  m_load<NB>(sF, Ft + k * nb2, nb);
  __syncthreads();
  // herm(i d): x = i d_ij, y = i d_ji, 0.5 (x + conj(y))  (solver.py:163-164)
  PK_FOR_ELEM(nb, {
    const cx<T> a = sT[i * PK_PITCH + j], b = sT[j * PK_PITCH + i];
    const cx<T> x{-a.im, a.re}, y{-b.im, b.re};
    sA[i * PK_PITCH + j] = (T)0.5 * (x + conj(y));
  })
  __syncthreads();
  m_eigh<NB>(sA, sV, sT, sl, nb, nullptr);
  if ((int)threadIdx.x < nb) lam_d[k * nb + threadIdx.x] = sl[threadIdx.x];
  m_store<NB>(Vout + k * nb2, sV, nb);
  m_gemm<NB, T, false, false>(sT, sF, sV, nb);
  m_gemm<NB, T, true, false>(sA, sV, sT, nb);
  m_store<NB>(Fe + k * nb2, sA, nb);
  dn = PK_SUM(dn);
  if (threadIdx.x == 0) part[k] = dn;
}

// CG line-search trial (solver.py:549-553; tau from st[HFX_ST_TAU]):
// part[0][k] = sum_b p_t eps_t, part[1][k] = sum_b [p_t ln p_t + (1-p_t) ln(1-p_t)].
// With omega0 != 0 also the tau = 0 point (solver.py:559-561) into part[2], part[3].
// Skipped (no writes) once st[HFX_ST_ACC] is set.
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_cg_trial_kernel(int nb, size_t Nk, const cx<T>* __restrict__ V,
                                                         const double* __restrict__ lam_d,
                                                         const cx<T>* __restrict__ Fe, const double* __restrict__ p,
                                                         const double* __restrict__ dp,
                                                         const double* __restrict__ eps, const double* __restrict__ st,
                                                         int omega0, double* __restrict__ part) {
  if (st[HFX_ST_ACC] != 0.0) return;
  __shared__ cx<T> sV[PK_MAT], sF[PK_MAT], sW[PK_MAT];
  __shared__ cx<T> cs[PK_NBMAX];
  __shared__ double et[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  const double tau = st[HFX_ST_TAU];
  m_load<NB>(sV, V + k * nb2, nb);
  m_load<NB>(sF, Fe + k * nb2, nb);
  if ((int)threadIdx.x < nb) cs[threadIdx.x] = cayley<T>(lam_d[k * nb + threadIdx.x], tau);
  __syncthreads();
  spectral_diag<NB>(sV, sF, cs, sW, et, nb);
  double ef = 0.0, sf = 0.0, e0 = 0.0, s0 = 0.0;
  if ((int)threadIdx.x < nb) {
    const size_t q = k * nb + threadIdx.x;
    double pt = fmin(fmax(p[q] - tau * dp[q], 1e-8), 1.0 - 1e-8);
    ef = pt * et[threadIdx.x];
    double x = fmin(fmax(pt, 1e-14), 1.0 - 1e-14);
    sf = x * log(x) + (1.0 - x) * log1p(-x);
    if (omega0) {
      const double p0 = fmin(fmax(p[q], 1e-8), 1.0 - 1e-8);
      e0 = p0 * eps[q];
      x = fmin(fmax(p0, 1e-14), 1.0 - 1e-14);
      s0 = x * log(x) + (1.0 - x) * log1p(-x);
    }
  }
  ef = PK_SUM(ef);
  sf = PK_SUM(sf);
  e0 = PK_SUM(e0);
  s0 = PK_SUM(s0);
  if (threadIdx.x == 0) {
    part[k] = ef;
    part[Nk + k] = sf;
    if (omega0) {
      part[2 * Nk + k] = e0;
      part[3 * Nk + k] = s0;
    }
  }
}

// CG retraction (solver.py:586-590): U = V diag(c) V^dagger, Q <- Q U,
// eps_new = diag(U^dagger Ft U) (spectral), tau from st[HFX_ST_TAUF].
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_cg_retract_kernel(int nb, cx<T>* __restrict__ Q, const cx<T>* __restrict__ V,
                                                           const double* __restrict__ lam_d,
                                                           const cx<T>* __restrict__ Fe, const double* __restrict__ st,
                                                           double tau_override, int use_override,
                                                           double* __restrict__ eps_new, cx<T>* __restrict__ Ft_out) {
  __shared__ cx<T> sQ[PK_MAT], sV[PK_MAT], sX[PK_MAT], sU[PK_MAT];
  __shared__ cx<T> cs[PK_NBMAX];
  __shared__ double et[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  const double tau = use_override ? tau_override : st[HFX_ST_TAUF];
  m_load<NB>(sQ, Q + k * nb2, nb);
  m_load<NB>(sV, V + k * nb2, nb);
  if ((int)threadIdx.x < nb) cs[threadIdx.x] = cayley<T>(lam_d[k * nb + threadIdx.x], tau);
  __syncthreads();
  PK_FOR_ELEM(nb, { sX[i * PK_PITCH + j] = sV[i * PK_PITCH + j] * cs[j]; })
  __syncthreads();
  m_gemm<NB, T, false, true>(sU, sX, sV, nb);     // U
  m_gemm<NB, T, false, false>(sX, sQ, sU, nb);    // Q U
  if constexpr (sizeof(T) == 4) {
    // FP32: V (and so U) is unitary only to ~1e-7, and the error would build up
    // over the iterations (the density then holds more or fewer than n electrons
    // and the energy drifts).  One Newton-Schulz step Q <- Q (3 I - Q^dagger Q) / 2
    // squares the deviation from unitarity, holding Q at FP32 rounding.
    m_gemm<NB, T, true, false>(sQ, sX, sX, nb);   // S = X^dagger X  (sQ is free here)
    PK_FOR_ELEM(nb, {
      const cx<T> v = sQ[i * PK_PITCH + j];
      sQ[i * PK_PITCH + j] = cx<T>{(i == j ? (T)1.5 : (T)0) - (T)0.5 * v.re, (T)-0.5 * v.im};
    })
    __syncthreads();
    PK_FOR_ELEM(nb, {   // Q <- X M, straight to global (sU, sV keep U and V)
      cx<T> acc{0, 0};
      for (int l = 0; l < nb; ++l) acc = acc + sX[i * PK_PITCH + l] * sQ[l * PK_PITCH + j];
      Q[k * nb2 + e] = acc;
    })
    __syncthreads();
  } else {
    m_store<NB>(Q + k * nb2, sX, nb);
  }
  if (eps_new) {
    m_load<NB>(sQ, Fe + k * nb2, nb);             // reuse sQ for Ft_eig
    __syncthreads();
    spectral_diag<NB>(sV, sQ, cs, sX, et, nb);
    if ((int)threadIdx.x < nb) eps_new[k * nb + threadIdx.x] = et[threadIdx.x];
  }
  if (Ft_out) {   // Ft_out = U^dagger Ft U with Ft given in Fe (Newton trial, solver.py:1008)
    m_load<NB>(sQ, Fe + k * nb2, nb);
    __syncthreads();
    m_gemm<NB, T, false, false>(sX, sQ, sU, nb);
    m_gemm<NB, T, true, false>(sV, sU, sX, nb);
    m_store<NB>(Ft_out + k * nb2, sV, nb);
  }
}

// ---- Cayley retraction by Gauss-Jordan (no eigh) -----------------------------
// U(tau) = (I + (tau/2) d)^{-1} (I - (tau/2) d), d = (D - D^dagger)/2 the
// skew-Hermitian part of the direction -- the same unitary as the spectral
// form V diag((1 + i tau lam/2)/(1 - i tau lam/2)) V^dagger of solver.py:140-197
// (i d = V diag(lam) V^dagger), obtained without diagonalising i d.  The
// elimination needs no pivoting: I + (tau/2) d has Hermitian part I, so every
// Schur complement keeps a positive definite Hermitian part and every pivot a
// real part >= 1.  On entry sX holds D; on exit sU holds U (sM, sX clobbered).
template <int NB, typename T>
__device__ void m_cayley_gj(cx<T>* sM, cx<T>* sU, const cx<T>* sX, double tau, int nb) {
  const T h = (T)(0.5 * tau);
  PK_FOR_ELEM(nb, {
    const cx<T> a = sX[i * PK_PITCH + j], b = sX[j * PK_PITCH + i];
    const cx<T> d{(T)0.5 * (a.re - b.re), (T)0.5 * (a.im + b.im)};   // (D - D^dagger)/2
    const T one = i == j ? (T)1 : (T)0;
    sM[i * PK_PITCH + j] = cx<T>{one + h * d.re, h * d.im};
    sU[i * PK_PITCH + j] = cx<T>{one - h * d.re, -h * d.im};
  })
  __syncthreads();
  for (int pv = 0; pv < nb; ++pv) {
    const cx<T> mp = sM[pv * PK_PITCH + pv];
    const T n2 = mp.re * mp.re + mp.im * mp.im;
    const cx<T> ip{mp.re / n2, -mp.im / n2};   // 1 / M_pp
    cx<T> nm[PK_EPL], nu[PK_EPL];
    PK_FOR_ELEM(nb, {
      const cx<T> mpj = sM[pv * PK_PITCH + j] * ip, upj = sU[pv * PK_PITCH + j] * ip;
      if (i == pv) {
        nm[m_] = mpj;
        nu[m_] = upj;
      } else {
        const cx<T> f = sM[i * PK_PITCH + pv];
        nm[m_] = sM[i * PK_PITCH + j] - f * mpj;
        nu[m_] = sU[i * PK_PITCH + j] - f * upj;
      }
    })
    __syncthreads();
    PK_FOR_ELEM(nb, {
      sM[i * PK_PITCH + j] = nm[m_];
      sU[i * PK_PITCH + j] = nu[m_];
    })
    __syncthreads();
  }
}

// nb = 16: the same elimination with the matrices in registers.  Lane l holds
// column j = l % 16 of rows i = l / 16 + 4 r (r = 0..3); the pivot row comes
// from the lanes of row p by a wave shuffle, the pivot column entry of each
// row from the lane of column p in the same 16-lane group; no LDS round trip
// or barrier inside the elimination.  Reads D from sX, leaves U in sU (sU may
// be sX: every lane has read D before the first store).  Same pivots, same
// arithmetic per element as m_cayley_gj.
template <typename T>
__device__ __forceinline__ cx<T> shfl_cx(cx<T> v, int src) {
  return cx<T>{__shfl(v.re, src), __shfl(v.im, src)};
}

template <int NB, typename T>
__device__ void m_cayley_gj16(cx<T>* sU, const cx<T>* sX, double tau) {
  const T h = (T)(0.5 * tau);
  const int l = threadIdx.x & 63, j = l & 15, g = l >> 4;
  cx<T> M[4], U[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = g + 4 * r;
    const cx<T> a = sX[i * PK_PITCH + j], b = sX[j * PK_PITCH + i];
    const cx<T> d{(T)0.5 * (a.re - b.re), (T)0.5 * (a.im + b.im)};   // (D - D^dagger)/2
    const T one = i == j ? (T)1 : (T)0;
    M[r] = cx<T>{one + h * d.re, h * d.im};
    U[r] = cx<T>{one - h * d.re, -h * d.im};
  }
  __syncthreads();   // sU may alias sX
  // pivot p = 4 rp + q: rp selects the register (compile time), q stays a loop
  // variable so that the 16 steps are not unrolled into one long schedule
#pragma unroll
  for (int rp = 0; rp < 4; ++rp) {
#pragma unroll 1
    for (int q = 0; q < 4; ++q) {
      const int p = 4 * rp + q, src_row = (q << 4) | j;
      const cx<T> mpj0 = shfl_cx(M[rp], src_row), upj0 = shfl_cx(U[rp], src_row);
      const cx<T> mp = shfl_cx(M[rp], (q << 4) | p);              // M_pp
      const T n2 = mp.re * mp.re + mp.im * mp.im;
      const cx<T> ip{mp.re / n2, -mp.im / n2};                     // 1 / M_pp
      const cx<T> mpj = mpj0 * ip, upj = upj0 * ip;
      cx<T> f[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) f[r] = shfl_cx(M[r], (g << 4) | p);   // M[i][p]
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (g + 4 * r == p) {
          M[r] = mpj;
          U[r] = upj;
        } else {
          M[r] = M[r] - f[r] * mpj;
          U[r] = U[r] - f[r] * upj;
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) sU[(g + 4 * r) * PK_PITCH + j] = U[r];
  __syncthreads();
}

// eps_b = Re sum_i conj(U_ib) (Ft U)_ib into et[b] (Y: LDS scratch)
template <int NB, typename T>
__device__ __forceinline__ void m_diag_UFU(const cx<T>* sF, const cx<T>* sU, cx<T>* sY, double* et, int nb) {
  m_gemm<NB, T, false, false>(sY, sF, sU, nb);
  if ((int)threadIdx.x < nb) {
    const int b = threadIdx.x;
    double acc = 0.0;
    for (int i = 0; i < nb; ++i) {
      const cx<T> u = sU[i * PK_PITCH + b], y = sY[i * PK_PITCH + b];
      acc += (double)u.re * (double)y.re + (double)u.im * (double)y.im;
    }
    et[b] = acc;
  }
  __syncthreads();
}

// Direction without the spectral setup: D <- H + beta D, dp <- hp + beta dp,
// part[k] = |D_k|^2 (solver.py:530-531, 545).
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_cg_dir_kernel(int nb, const cx<T>* __restrict__ H, cx<T>* __restrict__ D,
                                                       const double* __restrict__ hp, double* __restrict__ dp,
                                                       const double* __restrict__ st_all, double* __restrict__ part,
                                                       int seg_len) {
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  const double* st = seg_state(st_all, k, seg_len);
  const T beta = (T)st[HFX_ST_BETA];
  double dn = 0.0;
  PK_FOR_ELEM(nb, {
    const cx<T> d = H[k * nb2 + e] + beta * D[k * nb2 + e];
    D[k * nb2 + e] = d;
    dn += (double)d.re * d.re + (double)d.im * d.im;
  })
  if ((int)threadIdx.x < nb) {
    const size_t q = k * nb + threadIdx.x;
    dp[q] = hp[q] + (double)beta * dp[q];
  }
  dn = PK_SUM(dn);
  if (threadIdx.x == 0) part[k] = dn;
}

// Line-search trial with the Gauss-Jordan Cayley factor; same outputs as
// pk_cg_trial_kernel.  Skipped once st[HFX_ST_ACC] is set.
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_cg_trial_gj_kernel(int nb, size_t Nk, const cx<T>* __restrict__ D,
                                                            const cx<T>* __restrict__ Ft,
                                                            const double* __restrict__ p,
                                                            const double* __restrict__ dp,
                                                            const double* __restrict__ eps,
                                                            const double* __restrict__ st_all, int omega0,
                                                            double* __restrict__ part, cx<T>* __restrict__ U_out,
                                                            double* __restrict__ eps_out, int seg_len) {
  const double* st = seg_state(st_all, blockIdx.x, seg_len);
  if (st[HFX_ST_ACC] != 0.0 || st[HFX_ST_FROZEN] != 0.0) return;
  __shared__ cx<T> sX[PK_MAT], sM[PK_MAT], sU[PK_MAT], sF[PK_MAT];
  __shared__ double et[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  const double tau = st[HFX_ST_TAU];
  m_load<NB>(sX, D + k * nb2, nb);
  m_load<NB>(sF, Ft + k * nb2, nb);
  __syncthreads();
  if constexpr (NB == 16) {
    if (nb == 16) m_cayley_gj16<NB>(sU, sX, tau);
    else m_cayley_gj<NB>(sM, sU, sX, tau, nb);
  } else {
    m_cayley_gj<NB>(sM, sU, sX, tau, nb);
  }
  m_diag_UFU<NB>(sF, sU, sX, et, nb);
  if (U_out) {   // kept for the retraction when this trial is the accepted one
    m_store<NB>(U_out + k * nb2, sU, nb);
    if ((int)threadIdx.x < nb) eps_out[k * nb + threadIdx.x] = et[threadIdx.x];
  }
  double ef = 0.0, sf = 0.0, e0 = 0.0, s0 = 0.0;
  if ((int)threadIdx.x < nb) {
    const size_t q = k * nb + threadIdx.x;
    double pt = fmin(fmax(p[q] - tau * dp[q], 1e-8), 1.0 - 1e-8);
    ef = pt * et[threadIdx.x];
    double x = fmin(fmax(pt, 1e-14), 1.0 - 1e-14);
    sf = x * log(x) + (1.0 - x) * log1p(-x);
    if (omega0) {
      const double p0 = fmin(fmax(p[q], 1e-8), 1.0 - 1e-8);
      e0 = p0 * eps[q];
      x = fmin(fmax(p0, 1e-14), 1.0 - 1e-14);
      s0 = x * log(x) + (1.0 - x) * log1p(-x);
    }
  }
  ef = PK_SUM(ef);
  sf = PK_SUM(sf);
  e0 = PK_SUM(e0);
  s0 = PK_SUM(s0);
  if (threadIdx.x == 0) {
    part[k] = ef;
    part[Nk + k] = sf;
    if (omega0) {
      part[2 * Nk + k] = e0;
      part[3 * Nk + k] = s0;
    }
  }
}

// Retraction with the Gauss-Jordan Cayley factor (tau = st[HFX_ST_TAUF]):
// Q <- Q U (FP32: plus one Newton-Schulz re-orthonormalisation step),
// eps_new = diag(U^dagger Ft U)  (solver.py:586-592).
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_cg_retract_gj_kernel(int nb, cx<T>* __restrict__ Q,
                                                              const cx<T>* __restrict__ D,
                                                              const cx<T>* __restrict__ Ft,
                                                              const double* __restrict__ st_all,
                                                              double* __restrict__ eps_new,
                                                              const cx<T>* __restrict__ U_acc,
                                                              const double* __restrict__ eps_acc, int seg_len) {
  __shared__ cx<T> sX[PK_MAT], sM[PK_MAT], sU[PK_MAT], sF[PK_MAT];
  __shared__ double et[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  const double* st = seg_state(st_all, k, seg_len);
  if (st[HFX_ST_FROZEN] != 0.0) return;
  if (U_acc && st[HFX_ST_ACC] != 0.0) {
    // a trial was accepted: tau_final is its tau, and its U / eps were kept
    m_load<NB>(sU, U_acc + k * nb2, nb);
    if ((int)threadIdx.x < nb) eps_new[k * nb + threadIdx.x] = eps_acc[k * nb + threadIdx.x];
  } else {
    const double tau = st[HFX_ST_TAUF];
    m_load<NB>(sX, D + k * nb2, nb);
    m_load<NB>(sF, Ft + k * nb2, nb);
    __syncthreads();
    if constexpr (NB == 16) {
      if (nb == 16) m_cayley_gj16<NB>(sU, sX, tau);
      else m_cayley_gj<NB>(sM, sU, sX, tau, nb);
    } else {
      m_cayley_gj<NB>(sM, sU, sX, tau, nb);
    }
    m_diag_UFU<NB>(sF, sU, sX, et, nb);
    if ((int)threadIdx.x < nb) eps_new[k * nb + threadIdx.x] = et[threadIdx.x];
  }
  m_load<NB>(sF, Q + k * nb2, nb);
  __syncthreads();
  m_gemm<NB, T, false, false>(sX, sF, sU, nb);    // X = Q U
  if constexpr (sizeof(T) == 4) {
    m_gemm<NB, T, true, false>(sM, sX, sX, nb);   // S = X^dagger X
    PK_FOR_ELEM(nb, {
      const cx<T> v = sM[i * PK_PITCH + j];
      sM[i * PK_PITCH + j] = cx<T>{(i == j ? (T)1.5 : (T)0) - (T)0.5 * v.re, (T)-0.5 * v.im};
    })
    __syncthreads();
    m_gemm<NB, T, false, false>(sU, sX, sM, nb);  // Q <- X (3 I - S) / 2
    m_store<NB>(Q + k * nb2, sU, nb);
  } else {
    m_store<NB>(Q + k * nb2, sX, nb);
  }
}

// Ft = Q^dagger F Q, eigh(Ft) -> (eps, W), Q <- Q W  (solver.py:622-625)
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_rotate_eigh_kernel(int nb, cx<T>* __restrict__ Q, const cx<T>* __restrict__ F,
                                                            double* __restrict__ eps) {
  __shared__ cx<T> sQ[PK_MAT], sF[PK_MAT], sX[PK_MAT], sT[PK_MAT];
  __shared__ double sl[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  m_load<NB>(sQ, Q + k * nb2, nb);
  m_load<NB>(sF, F + k * nb2, nb);
  __syncthreads();
  m_gemm<NB, T, false, false>(sX, sF, sQ, nb);
  m_gemm<NB, T, true, false>(sT, sQ, sX, nb);     // Ft
  m_eigh<NB>(sT, sX, sF, sl, nb, nullptr);        // W in sX
  m_gemm<NB, T, false, false>(sF, sQ, sX, nb);    // Q W
  m_store<NB>(Q + k * nb2, sF, nb);
  if ((int)threadIdx.x < nb) eps[k * nb + threadIdx.x] = sl[threadIdx.x];
}

// C = op(A) op(B) op(C0)?  Generic batched products: out = opA(A) opB(B) [opC(Cm)]
template <int NB, typename T, bool AH, bool BH, bool CH>
__global__ __launch_bounds__(PKB) void pk_mm_kernel(int nb, const cx<T>* __restrict__ A, const cx<T>* __restrict__ B,
                                                   const cx<T>* __restrict__ Cm, cx<T>* __restrict__ out,
                                                   const cx<T>* __restrict__ Bsub, int herm_out) {
  __shared__ cx<T> sA[PK_MAT], sB[PK_MAT], sC[PK_MAT], sX[PK_MAT];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  m_load<NB>(sA, A + k * nb2, nb);
  m_load<NB>(sB, B + k * nb2, nb);
  if (Cm) m_load<NB>(sC, Cm + k * nb2, nb);
  __syncthreads();
  if (Bsub) {   // B <- B - Bsub   (e.g. F[dP] - h)
    PK_FOR_ELEM(nb, { sB[i * PK_PITCH + j] = sB[i * PK_PITCH + j] - Bsub[k * nb2 + e]; })
    __syncthreads();
  }
  m_gemm<NB, T, AH, BH>(sX, sA, sB, nb);
  const cx<T>* R = sX;
  if (Cm) {
    m_gemm<NB, T, false, CH>(sA, sX, sC, nb);
    R = sA;
  }
  if (herm_out) {
    PK_FOR_ELEM(nb, { out[k * nb2 + e] = (T)0.5 * (R[i * PK_PITCH + j] + conj(R[j * PK_PITCH + i])); })
  } else {
    m_store<NB>(out + k * nb2, R, nb);
  }
}

// DIIS extrapolation in the current basis (reference_scf.py:298-314, then the
// eigh of reference_scf.py:87-96 started from the basis U):
//   Ft = U^dagger herm(sum_v c_v X_v) U   (Hermitised), one wave per k.
constexpr int PK_XMAX = 8;
struct PkInputs {
  const void* x[PK_XMAX];
  double c[PK_XMAX];
};

// coef_dev (optional): per-solve coefficients, coef_dev[s * PK_XMAX + v] for the
// solve s = k / seg_len of a stack of solves (the batched polish).
// sX = sum_v c_v X_v[k] with the input count a compile-time constant, so that
// the loads of every input go out together
template <int NB, typename T, int NIN>
__device__ __forceinline__ void extrap_sum(const PkInputs& in, const double* sc, size_t k, int nb, cx<T>* sX) {
  const size_t nb2 = (size_t)nb * nb;
  cx<T> acc[PK_EPL];
#pragma unroll
  for (int m_ = 0; m_ < PK_EPL; ++m_) acc[m_] = cx<T>{(T)0, (T)0};
#pragma unroll
  for (int v = 0; v < NIN; ++v) {
    const cx<T>* Xv = reinterpret_cast<const cx<T>*>(in.x[v]);
    const T cv = (T)sc[v];
    PK_FOR_ELEM(nb, { acc[m_] = acc[m_] + cv * Xv[k * nb2 + e]; })
  }
  PK_FOR_ELEM(nb, { sX[i * PK_PITCH + j] = acc[m_]; })
}

template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_extrap_rotate_kernel(int nb, int nin, PkInputs in,
                                                              const double* __restrict__ coef_dev, int seg_len,
                                                              const cx<T>* __restrict__ U, cx<T>* __restrict__ Ft) {
  __shared__ cx<T> sF[PK_MAT], sU[PK_MAT], sX[PK_MAT];
  __shared__ double sc[PK_XMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  if ((int)threadIdx.x < nin)
    sc[threadIdx.x] = coef_dev ? coef_dev[(k / (size_t)seg_len) * PK_XMAX + threadIdx.x] : in.c[threadIdx.x];
  __syncthreads();
  switch (nin) {
    case 1: extrap_sum<NB, T, 1>(in, sc, k, nb, sX); break;
    case 2: extrap_sum<NB, T, 2>(in, sc, k, nb, sX); break;
    case 3: extrap_sum<NB, T, 3>(in, sc, k, nb, sX); break;
    case 4: extrap_sum<NB, T, 4>(in, sc, k, nb, sX); break;
    case 5: extrap_sum<NB, T, 5>(in, sc, k, nb, sX); break;
    case 6: extrap_sum<NB, T, 6>(in, sc, k, nb, sX); break;
    case 7: extrap_sum<NB, T, 7>(in, sc, k, nb, sX); break;
    default: extrap_sum<NB, T, 8>(in, sc, k, nb, sX); break;
  }
  m_load<NB>(sU, U + k * nb2, nb);
  __syncthreads();
  PK_FOR_ELEM(nb, { sF[i * PK_PITCH + j] = (T)0.5 * (sX[i * PK_PITCH + j] + conj(sX[j * PK_PITCH + i])); })
  __syncthreads();
  m_gemm<NB, T, true, false>(sX, sU, sF, nb);     // U^H F
  m_gemm<NB, T, false, false>(sF, sX, sU, nb);    // (U^H F) U
  PK_FOR_ELEM(nb, { Ft[k * nb2 + e] = (T)0.5 * (sF[i * PK_PITCH + j] + conj(sF[j * PK_PITCH + i])); })
}

// Back from the rotated basis: U <- U W (in place, per k), and the density
// P = herm(U diag(occ) U^dagger) of the new orbitals (solver.py:42-44).
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_rotate_density_kernel(int nb, cx<T>* __restrict__ U,
                                                               const cx<T>* __restrict__ W,
                                                               const double* __restrict__ occ,
                                                               cx<T>* __restrict__ P) {
  __shared__ cx<T> sU[PK_MAT], sW[PK_MAT], sX[PK_MAT];
  __shared__ double sp[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  m_load<NB>(sU, U + k * nb2, nb);
  m_load<NB>(sW, W + k * nb2, nb);
  if ((int)threadIdx.x < nb) sp[threadIdx.x] = occ[k * nb + threadIdx.x];
  __syncthreads();
  m_gemm<NB, T, false, false>(sX, sU, sW, nb);    // new orbitals U W
  PK_FOR_ELEM(nb, {
    U[k * nb2 + e] = sX[i * PK_PITCH + j];
    sW[i * PK_PITCH + j] = (T)sp[j] * sX[i * PK_PITCH + j];
  })
  __syncthreads();
  m_gemm<NB, T, false, true>(sU, sW, sX, nb);     // (U W) diag(occ) (U W)^H
  PK_FOR_ELEM(nb, { P[k * nb2 + e] = (T)0.5 * (sU[i * PK_PITCH + j] + conj(sU[j * PK_PITCH + i])); })
}

// Newton-Schulz re-orthonormalisation in place, `steps` times:
// U <- U (3 I - U^dagger U) / 2.  Each step squares the deviation of U^dagger U
// from I: an orbital basis carried in FP32 (~1e-7) reaches FP64 rounding in two.
// src (optional): start from these orbitals in precision TS instead of U's own
// (the FP32 CG orbitals converted on the way in).
template <int NB, typename T, typename TS = T>
__global__ __launch_bounds__(PKB) void pk_orthonormalize_kernel(int nb, cx<T>* __restrict__ U, int steps,
                                                               const cx<TS>* __restrict__ src) {
  __shared__ cx<T> sU[PK_MAT], sS[PK_MAT], sX[PK_MAT];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  if (src) {
    PK_FOR_ELEM(nb, {
      const cx<TS> v = src[k * nb2 + e];
      sU[i * PK_PITCH + j] = cx<T>{(T)v.re, (T)v.im};
    })
  } else {
    m_load<NB>(sU, U + k * nb2, nb);
  }
  __syncthreads();
  for (int it = 0; it < steps; ++it) {
    m_gemm<NB, T, true, false>(sS, sU, sU, nb);     // S = U^H U
    PK_FOR_ELEM(nb, {
      const cx<T> v = sS[i * PK_PITCH + j];
      sS[i * PK_PITCH + j] = cx<T>{(i == j ? (T)1.5 : (T)0) - (T)0.5 * v.re, (T)-0.5 * v.im};
    })
    __syncthreads();
    m_gemm<NB, T, false, false>(sX, sU, sS, nb);    // U (3 I - S) / 2
    PK_FOR_ELEM(nb, { sU[i * PK_PITCH + j] = sX[i * PK_PITCH + j]; })
    __syncthreads();
  }
  m_store<NB>(U + k * nb2, sU, nb);
}

// SCF residuals (reference_scf.py:291-294): comm = F D - D F (stored when
// comm != NULL), part[0][k] = sum |comm|^2, part[1][k] = sum |Pn - D|^2, and
// with nr > 0 the new row of the DIIS B matrix in the same pass:
// part[2 + v][k] = Re sum conj(comm) R_v over the stored residuals R_v
// (reference_scf.py:110-111); for v == self (the slot comm is written to) the
// value is part[0][k] itself.
struct PkRes {
  const void* r[PK_XMAX];
  int nr, self;
  int r32;   // FP64 kernels: the residual history (comm included) stored as complex64
};
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_scf_resid_kernel(int nb, size_t Nk, const cx<T>* __restrict__ F,
                                                          const cx<T>* __restrict__ D, const cx<T>* __restrict__ Pn,
                                                          cx<T>* __restrict__ comm, double* __restrict__ part,
                                                          PkRes rs) {
  __shared__ cx<T> sF[PK_MAT], sD[PK_MAT], sX[PK_MAT], sY[PK_MAT];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  m_load<NB>(sF, F + k * nb2, nb);
  m_load<NB>(sD, D + k * nb2, nb);
  __syncthreads();
  m_gemm<NB, T, false, false>(sX, sF, sD, nb);
  m_gemm<NB, T, false, false>(sY, sD, sF, nb);
  double c2 = 0.0, d2 = 0.0, bd[PK_XMAX];
#pragma unroll
  for (int v = 0; v < PK_XMAX; ++v) bd[v] = 0.0;
  PK_FOR_ELEM(nb, {
    const cx<T> c = sX[i * PK_PITCH + j] - sY[i * PK_PITCH + j];
    if (comm) comm[k * nb2 + e] = c;
    c2 += (double)c.re * c.re + (double)c.im * c.im;
    if (Pn) {
      const cx<T> d = Pn[k * nb2 + e] - sD[i * PK_PITCH + j];
      d2 += (double)d.re * d.re + (double)d.im * d.im;
    }
    _Pragma("unroll") for (int v = 0; v < PK_XMAX; ++v)
      if (v < rs.nr && v != rs.self) {
        const cx<T> r = reinterpret_cast<const cx<T>*>(rs.r[v])[k * nb2 + e];
        bd[v] += (double)c.re * r.re + (double)c.im * r.im;
      }
  })
  c2 = PK_SUM(c2);
  d2 = PK_SUM(d2);
  if (threadIdx.x == 0) {
    part[k] = c2;
    part[Nk + k] = d2;
  }
#pragma unroll
  for (int v = 0; v < PK_XMAX; ++v)
    if (v < rs.nr) {
      const double b = v == rs.self ? c2 : PK_SUM(bd[v]);
      if (threadIdx.x == 0) part[(size_t)(2 + v) * Nk + k] = b;
    }
}

// nb = 16: the same outputs with F D - D F accumulated on the matrix cores and
// kept in registers (lane l holds the elements (row(l, r), l % 16)); two staged
// matrices instead of four, so twice the waves per CU for this streaming pass.
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_scf_resid16_kernel(size_t Nk, const cx<T>* __restrict__ F,
                                                            const cx<T>* __restrict__ D,
                                                            const cx<T>* __restrict__ Pn, cx<T>* __restrict__ comm,
                                                            double* __restrict__ part, PkRes rs) {
  static_assert(NB == 16, "the matrix-core residual pass is the nb = 16 form");
  constexpr int nb = 16, nb2 = 256;
  __shared__ cx<T> sF[PK_MAT], sD[PK_MAT];
  const size_t k = blockIdx.x;
  m_load<NB>(sF, F + k * nb2, nb);
  m_load<NB>(sD, D + k * nb2, nb);
  __syncthreads();
  MAcc<T> acc;
  acc.zero();
  mf_gemm16<NB, T, false, false>(acc, sF, sD);
  mf_gemm16<NB, T, false, false>(acc, sD, sF, (T)-1);
  const int l = threadIdx.x & 63, col = l & 15;
  double c2 = 0.0, d2 = 0.0, bd[PK_XMAX];
#pragma unroll
  for (int v = 0; v < PK_XMAX; ++v) bd[v] = 0.0;
  // this lane's four elements; each stream below issues its four loads
  // together (one memory round trip per stored residual, not four)
  size_t el[4];
  cx<T> cv[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    el[r] = k * nb2 + (size_t)Mfma<T>::row(l, r) * nb + col;
    cv[r] = cx<T>{acc.re[r], acc.im[r]};
    c2 += (double)cv[r].re * cv[r].re + (double)cv[r].im * cv[r].im;
  }
  if (Pn) {
    cx<T> pn[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) pn[r] = Pn[el[r]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const cx<T> d = pn[r] - sD[Mfma<T>::row(l, r) * PK_PITCH + col];
      d2 += (double)d.re * d.re + (double)d.im * d.im;
    }
  }
#pragma unroll
  for (int v = 0; v < PK_XMAX; ++v)
    if (v < rs.nr && v != rs.self) {
      double qre[4], qim[4];
      if (rs.r32) {
        const cx<float>* R = reinterpret_cast<const cx<float>*>(rs.r[v]);
        cx<float> q[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) q[r] = R[el[r]];
#pragma unroll
        for (int r = 0; r < 4; ++r) { qre[r] = q[r].re; qim[r] = q[r].im; }
      } else {
        const cx<T>* R = reinterpret_cast<const cx<T>*>(rs.r[v]);
        cx<T> q[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) q[r] = R[el[r]];
#pragma unroll
        for (int r = 0; r < 4; ++r) { qre[r] = (double)q[r].re; qim[r] = (double)q[r].im; }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) bd[v] += (double)cv[r].re * qre[r] + (double)cv[r].im * qim[r];
    }
  // the new residual last: the stored history above is read before this slot is rewritten
  if (comm) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (rs.r32) reinterpret_cast<cx<float>*>(comm)[el[r]] = cx<float>{(float)cv[r].re, (float)cv[r].im};
      else comm[el[r]] = cv[r];
    }
  }
  c2 = PK_SUM(c2);
  d2 = PK_SUM(d2);
  if (threadIdx.x == 0) {
    part[k] = c2;
    part[Nk + k] = d2;
  }
#pragma unroll
  for (int v = 0; v < PK_XMAX; ++v)
    if (v < rs.nr) {
      const double b = v == rs.self ? c2 : PK_SUM(bd[v]);
      if (threadIdx.x == 0) part[(size_t)(2 + v) * Nk + k] = b;
    }
}

// ---- symmetry averaging (src/jax_hf/symmetry.py:46-150) ----------------------
//   out[k] = mask[k] ? weight sum_g U_g op_g(A[src_g(k)]) U_g^dagger : A[k]
// op_g = complex conjugation for the elements flagged in conj_bits (time
// reversal).  Every k-map of the reference (identity, the mod / flip
// negations, per-element index maps) is a gather table src (ng, Nk) made on
// the host.  One wave per k; the two products per element run on the matrix
// cores at nb = 16.  Out-of-range sources (rejected by the host) read A[k].
constexpr int SYM_MAX = 64;
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_sym_avg_kernel(int nb, size_t Nk, int ng, const cx<T>* __restrict__ A,
                                                        const cx<T>* __restrict__ Ug, const int* __restrict__ src,
                                                        unsigned long long conj_bits,
                                                        const unsigned char* __restrict__ mask, double weight,
                                                        cx<T>* __restrict__ out) {
  __shared__ cx<T> sA[PK_MAT], sU[PK_MAT], sX[PK_MAT];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  if (mask && !mask[k]) {
    PK_FOR_ELEM(nb, { out[k * nb2 + e] = A[k * nb2 + e]; })
    return;
  }
  cx<T> acc[PK_EPL];
#pragma unroll
  for (int m = 0; m < PK_EPL; ++m) acc[m] = cx<T>{0, 0};
  for (int g = 0; g < ng; ++g) {
    long long ks = src[(size_t)g * Nk + k];
    if (ks < 0 || ks >= (long long)Nk) ks = (long long)k;
    const bool cj = (conj_bits >> g) & 1ull;
    PK_FOR_ELEM(nb, {
      cx<T> a = A[(size_t)ks * nb2 + e];
      if (cj) a.im = -a.im;
      sA[i * PK_PITCH + j] = a;
      sU[i * PK_PITCH + j] = Ug[(size_t)g * nb2 + e];
    })
    __syncthreads();
    m_gemm<NB, T, false, false>(sX, sU, sA, nb);   // U op(A)
    m_gemm<NB, T, false, true>(sA, sX, sU, nb);    // (U op(A)) U^dagger
    PK_FOR_ELEM(nb, { acc[m_] = acc[m_] + sA[i * PK_PITCH + j]; })
    __syncthreads();
  }
  PK_FOR_ELEM(nb, { out[k * nb2 + e] = (T)weight * acc[m_]; })
}

// Per-k inner products of X with up to 8 matrices Y_v:
//   mode 0: part[v][k] = Re sum_ij conj(X_ij) Y_v,ij   (Frobenius, solver.py:205-208, reference_scf.py:110-111)
//   mode 1: part[v][k] = Re sum_ij X_ij Y_v,ji          (trace of the product, reference_scf.py:319)
struct DotList {
  const void* x[8];
  const void* y[8];
};
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_dots_kernel(int nb, size_t Nk, DotList L, int nv, int mode,
                                                     double* __restrict__ part) {
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  for (int v = 0; v < nv; ++v) {
    const cx<T>* X = (const cx<T>*)L.x[v];
    const cx<T>* Y = (const cx<T>*)L.y[v];
    double acc = 0.0;
    PK_FOR_ELEM(nb, {
      const cx<T> x = X[k * nb2 + e];
      const cx<T> y = mode == 0 ? Y[k * nb2 + e] : Y[k * nb2 + (size_t)j * nb + i];
      acc += mode == 0 ? ((double)x.re * y.re + (double)x.im * y.im) : ((double)x.re * y.re - (double)x.im * y.im);
    })
    acc = PK_SUM(acc);
    if (threadIdx.x == 0) part[(size_t)v * Nk + k] = acc;
  }
}

// Newton HVP, density part (solver.py:917-919):
//   M = X (x) diff_p + diag(dp), diff_p[i][j] = p_j - p_i;  dP = herm(Q M Q^dagger).
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_newton_dP_kernel(int nb, const cx<T>* __restrict__ Q, const cx<T>* __restrict__ X,
                                                          const double* __restrict__ p, const double* __restrict__ dp,
                                                          cx<T>* __restrict__ dP) {
  __shared__ cx<T> sQ[PK_MAT], sM[PK_MAT], sX[PK_MAT], sP[PK_MAT];
  __shared__ double sp[PK_NBMAX], sd[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  m_load<NB>(sQ, Q + k * nb2, nb);
  if ((int)threadIdx.x < nb) {
    sp[threadIdx.x] = p[k * nb + threadIdx.x];
    sd[threadIdx.x] = dp[k * nb + threadIdx.x];
  }
  __syncthreads();
  PK_FOR_ELEM(nb, {
    cx<T> m = (T)(sp[j] - sp[i]) * X[k * nb2 + e];
    if (i == j) m.re += (T)sd[i];
    sM[i * PK_PITCH + j] = m;
  })
  __syncthreads();
  m_gemm<NB, T, false, false>(sX, sQ, sM, nb);
  m_gemm<NB, T, false, true>(sP, sX, sQ, nb);
  PK_FOR_ELEM(nb, { dP[k * nb2 + e] = (T)0.5 * (sP[i * PK_PITCH + j] + conj(sP[j * PK_PITCH + i])); })
}

// Newton HVP, response part (solver.py:923-929) with R = F[dP] - h (+ penalty):
//   St = Q^dagger R Q, A = Ft X - X Ft + St,
//   HX = skew(diff_dp (x) Ft + diff_p (x) A) offdiag,  Hp = Re diag(A) + T/(pc(1-pc)) dp
// (Hp before the particle-conserving projection); part[k] = sum_b Hp.
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_newton_hvp_kernel(int nb, const cx<T>* __restrict__ Q,
                                                           const cx<T>* __restrict__ R, const cx<T>* __restrict__ Ft,
                                                           const cx<T>* __restrict__ X, const double* __restrict__ p,
                                                           const double* __restrict__ dp, double T_,
                                                           cx<T>* __restrict__ HX, double* __restrict__ Hp,
                                                           double* __restrict__ part) {
  __shared__ cx<T> sQ[PK_MAT], sR[PK_MAT], sF[PK_MAT], sX[PK_MAT], sY[PK_MAT], sA[PK_MAT];
  __shared__ double sp[PK_NBMAX], sd[PK_NBMAX];
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  m_load<NB>(sQ, Q + k * nb2, nb);
  m_load<NB>(sR, R + k * nb2, nb);
  m_load<NB>(sF, Ft + k * nb2, nb);
  m_load<NB>(sX, X + k * nb2, nb);
  if ((int)threadIdx.x < nb) {
    sp[threadIdx.x] = p[k * nb + threadIdx.x];
    sd[threadIdx.x] = dp[k * nb + threadIdx.x];
  }
  __syncthreads();
  m_gemm<NB, T, false, false>(sY, sR, sQ, nb);    // R Q
  m_gemm<NB, T, true, false>(sA, sQ, sY, nb);     // St
  m_gemm<NB, T, false, false>(sY, sF, sX, nb);    // Ft X
  PK_FOR_ELEM(nb, { sA[i * PK_PITCH + j] = sA[i * PK_PITCH + j] + sY[i * PK_PITCH + j]; })
  __syncthreads();
  m_gemm<NB, T, false, false>(sY, sX, sF, nb);    // X Ft
  PK_FOR_ELEM(nb, { sR[i * PK_PITCH + j] = sA[i * PK_PITCH + j] - sY[i * PK_PITCH + j]; })   // A
  __syncthreads();
  PK_FOR_ELEM(nb, {
    // C = diff_dp (x) Ft + diff_p (x) A ; HX = 0.5 (C - C^H) offdiag
    cx<T> h{0, 0};
    if (i != j) {
      const cx<T> cij = (T)(sd[j] - sd[i]) * sF[i * PK_PITCH + j] + (T)(sp[j] - sp[i]) * sR[i * PK_PITCH + j];
      const cx<T> cji = (T)(sd[i] - sd[j]) * sF[j * PK_PITCH + i] + (T)(sp[i] - sp[j]) * sR[j * PK_PITCH + i];
      h = (T)0.5 * (cij - conj(cji));
    }
    HX[k * nb2 + e] = h;
  })
  double acc = 0.0;
  if ((int)threadIdx.x < nb) {
    const int b = threadIdx.x;
    const double pc = fmin(fmax(sp[b], 1e-8), 1.0 - 1e-8);
    const double v = (double)sR[b * PK_PITCH + b].re + (T_ / (pc * (1.0 - pc))) * sd[b];
    Hp[k * nb + b] = v;
    acc = v;
  }
  acc = PK_SUM(acc);
  if (threadIdx.x == 0) part[k] = acc;
}

// Newton preconditioner (solver.py:905-913): ZX = RX / sqrt(gap^2 + lam^2)
// (lam = st[HFX_ST_LAM]), zp = rp pc(1-pc)/T before projection; part[k] = sum_b zp.
template <int NB, typename T>
__global__ __launch_bounds__(PKB) void pk_newton_prec_kernel(int nb, const cx<T>* __restrict__ RX,
                                                            const double* __restrict__ rp, const double* __restrict__ eps,
                                                            const double* __restrict__ p, const double* __restrict__ st,
                                                            double T_, cx<T>* __restrict__ ZX, double* __restrict__ zp,
                                                            double* __restrict__ part) {
  const size_t k = blockIdx.x, nb2 = (size_t)nb * nb;
  const double lam = st[HFX_ST_LAM];
  PK_FOR_ELEM(nb, {
    const double gap = eps[k * nb + i] - eps[k * nb + j];
    const T den = (T)sqrt(gap * gap + lam * lam);
    const cx<T> r = RX[k * nb2 + e];
    ZX[k * nb2 + e] = cx<T>{r.re / den, r.im / den};
  })
  double acc = 0.0;
  if ((int)threadIdx.x < nb) {
    const size_t q = k * nb + threadIdx.x;
    const double pc = fmin(fmax(p[q], 1e-8), 1.0 - 1e-8);
    const double v = rp[q] * (pc * (1.0 - pc) / T_);
    zp[q] = v;
    acc = v;
  }
  acc = PK_SUM(acc);
  if (threadIdx.x == 0) part[k] = acc;
}

// ============================ launchers ======================================

// Size-class dispatch: nb <= 16 -> one wave per k (NB = 16), nb <= 32 -> a
// 256-thread workgroup per k (NB = 32).  `nb`, `nk` and `s` in scope.
constexpr int PK_NB_LIMIT = 32;
#define PK_LAUNCH(K, T, ...)                                                                 \
  do {                                                                                       \
    if (nb <= 16)                                                                            \
      hipLaunchKernelGGL((K<16, T>), dim3(nk), dim3(64), 0, s, __VA_ARGS__);                  \
    else                                                                                     \
      hipLaunchKernelGGL((K<32, T>), dim3(nk), dim3(256), 0, s, __VA_ARGS__);                 \
  } while (0)
#define PK_LAUNCH_MM(T, a, b, c, ...)                                                        \
  do {                                                                                       \
    if (nb <= 16)                                                                            \
      hipLaunchKernelGGL((pk_mm_kernel<16, T, a, b, c>), dim3(nk), dim3(64), 0, s, __VA_ARGS__); \
    else                                                                                     \
      hipLaunchKernelGGL((pk_mm_kernel<32, T, a, b, c>), dim3(nk), dim3(256), 0, s, __VA_ARGS__); \
  } while (0)

template <typename T> using C_ = cx<T>;

int pk_check_nb(int nb) {
  if (nb < 1 || nb > PK_NB_LIMIT)
    return set_error(HFX_EUNSUPPORTED, "per-k kernels support 1 <= nb <= %d (got %d)", PK_NB_LIMIT, nb);
  return HFX_OK;
}

}  // namespace hfx

using namespace hfx;

#define PKC(T, p) ((const cx<T>*)(p))
#define PKM(T, p) ((cx<T>*)(p))

extern "C" {

int hfx_eigh_batched(int dtype, int nb, size_t nk, const void* A, double* lam, void* V, int herm_in, void* stream) {
  return hfx_eigh_batched_tol(dtype, nb, nk, A, lam, V, herm_in, 0.0, stream);
}

static int eigh_any(int dtype, int nb, size_t nk, const void* A, double* lam, void* V, const void* U0, int herm_in,
                    double rel_tol, void* stream) {
  if (nk == 0) return HFX_OK;
  if (nb <= 16) return eigh16_launch(dtype, nb, nk, A, lam, V, herm_in, rel_tol, (hipStream_t)stream, U0);
  // 16 < nb <= 32: LDS-staged cyclic Jacobi, one 256-thread workgroup per
  // matrix, rotating to its own threshold (rel_tol only ever stops earlier)
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL((pk_eigh_kernel<32, double>), dim3(nk), dim3(256), 0, s, nb, PKC(double, A), lam,
                       PKM(double, V), PKC(double, U0), herm_in);
  else
    hipLaunchKernelGGL((pk_eigh_kernel<32, float>), dim3(nk), dim3(256), 0, s, nb, PKC(float, A), lam,
                       PKM(float, V), PKC(float, U0), herm_in);
  return check_launch("pk_eigh_kernel");
}

int hfx_eigh_batched_tol(int dtype, int nb, size_t nk, const void* A, double* lam, void* V, int herm_in,
                         double rel_tol, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!A || !lam || !V || !(rel_tol >= 0.0) || rel_tol > 1e-3)
    return set_error(HFX_EINVAL, "hfx_eigh_batched: bad argument (rel_tol=%g)", rel_tol);
  return eigh_any(dtype, nb, nk, A, lam, V, nullptr, herm_in, rel_tol, stream);
}

int hfx_eigh_batched_basis(int dtype, int nb, size_t nk, const void* A, double* lam, const void* U, void* V,
                           int herm_in, double rel_tol, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!A || !lam || !U || !V || A == U || A == V || !(rel_tol >= 0.0) || rel_tol > 1e-3)
    return set_error(HFX_EINVAL, "hfx_eigh_batched_basis: bad argument (rel_tol=%g)", rel_tol);
  return eigh_any(dtype, nb, nk, A, lam, V, U, herm_in, rel_tol, stream);
}

int hfx_density_from_Qp(int dtype, int nb, size_t nk, const void* Q, const double* p, const double* eps,
                        const double* mu, double T, double* pout, void* P, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!Q || !P || (!p && (!eps || !mu))) return set_error(HFX_EINVAL, "hfx_density_from_Qp: NULL argument");
  if (nk == 0) return HFX_OK;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_density_kernel, double, nb, PKC(double, Q), p, eps, mu, T, pout,
                       PKM(double, P), (const cx<double>*)nullptr, (double*)nullptr);
  else
    PK_LAUNCH(pk_density_kernel, float, nb, PKC(float, Q), p, eps, mu, T, pout,
                       PKM(float, P), (const cx<float>*)nullptr, (double*)nullptr);
  return check_launch("pk_density_kernel");
}

int hfx_density_diff(int dtype, int nb, size_t nk, const void* Q, const double* p, const void* Pref, void* P,
                     double* part, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!Q || !p || !Pref || !P || !part || P == Pref) return set_error(HFX_EINVAL, "hfx_density_diff: bad argument");
  if (nk == 0) return HFX_OK;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_density_kernel, double, nb, PKC(double, Q), p, (const double*)nullptr, (const double*)nullptr, 1.0,
              (double*)nullptr, PKM(double, P),
              PKC(double, Pref), part);
  else
    PK_LAUNCH(pk_density_kernel, float, nb, PKC(float, Q), p, (const double*)nullptr, (const double*)nullptr, 1.0,
              (double*)nullptr, PKM(float, P),
              PKC(float, Pref), part);
  return check_launch("pk_density_kernel");
}

int hfx_cg_grad(int dtype, int nb, size_t nk, const void* Q, const void* F, const double* p, void* Ft, double* eps,
                void* G, double* part, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_cg_grad_kernel, double, nb, nk, PKC(double, Q), PKC(double, F), p,
                       PKM(double, Ft), eps, PKM(double, G), part);
  else
    PK_LAUNCH(pk_cg_grad_kernel, float, nb, nk, PKC(float, Q), PKC(float, F), p,
                       PKM(float, Ft), eps, PKM(float, G), part);
  return check_launch("pk_cg_grad_kernel");
}

int hfx_cg_precond(int dtype, int nb, size_t nk, const void* G, const void* Gprev, const double* eps, const double* p,
                   const double* gp_prev, const double* st, double T, void* H, double* gp, double* hp, double* part,
                   int seg_len, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_cg_precond_kernel, double, nb, nk, PKC(double, G),
                       PKC(double, Gprev), eps, p, gp_prev, st, T, PKM(double, H), gp, hp, part, seg_len);
  else
    PK_LAUNCH(pk_cg_precond_kernel, float, nb, nk, PKC(float, G), PKC(float, Gprev),
                       eps, p, gp_prev, st, T, PKM(float, H), gp, hp, part, seg_len);
  return check_launch("pk_cg_precond_kernel");
}

int hfx_cg_direction(int dtype, int nb, size_t nk, const void* H, void* D, const double* hp, double* dp, const void* Ft,
                     const double* st, void* V, double* lam_d, void* Fe, double* part, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_cg_direction_kernel, double, nb, PKC(double, H), PKM(double, D),
                       hp, dp, PKC(double, Ft), st, PKM(double, V), lam_d, PKM(double, Fe), part);
  else
    PK_LAUNCH(pk_cg_direction_kernel, float, nb, PKC(float, H), PKM(float, D), hp,
                       dp, PKC(float, Ft), st, PKM(float, V), lam_d, PKM(float, Fe), part);
  return check_launch("pk_cg_direction_kernel");
}

int hfx_cg_trial(int dtype, int nb, size_t nk, const void* V, const double* lam_d, const void* Fe, const double* p,
                 const double* dp, const double* eps, const double* st, int omega0, double* part, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_cg_trial_kernel, double, nb, nk, PKC(double, V), lam_d,
                       PKC(double, Fe), p, dp, eps, st, omega0, part);
  else
    PK_LAUNCH(pk_cg_trial_kernel, float, nb, nk, PKC(float, V), lam_d,
                       PKC(float, Fe), p, dp, eps, st, omega0, part);
  return check_launch("pk_cg_trial_kernel");
}

int hfx_cg_retract(int dtype, int nb, size_t nk, void* Q, const void* V, const double* lam_d, const void* Fe,
                   const double* st, double tau, int use_tau, double* eps_new, void* Ft_out, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_cg_retract_kernel, double, nb, PKM(double, Q), PKC(double, V),
                       lam_d, PKC(double, Fe), st, tau, use_tau, eps_new, PKM(double, Ft_out));
  else
    PK_LAUNCH(pk_cg_retract_kernel, float, nb, PKM(float, Q), PKC(float, V), lam_d,
                       PKC(float, Fe), st, tau, use_tau, eps_new, PKM(float, Ft_out));
  return check_launch("pk_cg_retract_kernel");
}

int hfx_cg_direction_gj(int dtype, int nb, size_t nk, const void* H, void* D, const double* hp, double* dp,
                        const double* st, double* part, int seg_len, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_cg_dir_kernel, double, nb, PKC(double, H), PKM(double, D), hp,
                       dp, st, part, seg_len);
  else
    PK_LAUNCH(pk_cg_dir_kernel, float, nb, PKC(float, H), PKM(float, D), hp, dp,
                       st, part, seg_len);
  return check_launch("pk_cg_dir_kernel");
}

int hfx_cg_trial_gj(int dtype, int nb, size_t nk, const void* D, const void* Ft, const double* p, const double* dp,
                    const double* eps, const double* st, int omega0, double* part, void* U_keep,
                    double* eps_keep, int seg_len, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!U_keep != !eps_keep) return set_error(HFX_EINVAL, "hfx_cg_trial_gj: U_keep and eps_keep go together");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_cg_trial_gj_kernel, double, nb, nk, PKC(double, D),
                       PKC(double, Ft), p, dp, eps, st, omega0, part, PKM(double, U_keep), eps_keep, seg_len);
  else
    PK_LAUNCH(pk_cg_trial_gj_kernel, float, nb, nk, PKC(float, D),
                       PKC(float, Ft), p, dp, eps, st, omega0, part, PKM(float, U_keep), eps_keep, seg_len);
  return check_launch("pk_cg_trial_gj_kernel");
}

int hfx_cg_retract_gj(int dtype, int nb, size_t nk, void* Q, const void* D, const void* Ft, const double* st,
                      double* eps_new, const void* U_keep, const double* eps_keep, int seg_len, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!eps_new) return set_error(HFX_EINVAL, "hfx_cg_retract_gj: eps_new is required");
  if (!U_keep != !eps_keep) return set_error(HFX_EINVAL, "hfx_cg_retract_gj: U_keep and eps_keep go together");
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_cg_retract_gj_kernel, double, nb, PKM(double, Q),
                       PKC(double, D), PKC(double, Ft), st, eps_new, PKC(double, U_keep), eps_keep, seg_len);
  else
    PK_LAUNCH(pk_cg_retract_gj_kernel, float, nb, PKM(float, Q),
                       PKC(float, D), PKC(float, Ft), st, eps_new, PKC(float, U_keep), eps_keep, seg_len);
  return check_launch("pk_cg_retract_gj_kernel");
}

int hfx_rotate_eigh(int dtype, int nb, size_t nk, void* Q, const void* F, double* eps, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_rotate_eigh_kernel, double, nb, PKM(double, Q), PKC(double, F), eps);
  else
    PK_LAUNCH(pk_rotate_eigh_kernel, float, nb, PKM(float, Q), PKC(float, F), eps);
  return check_launch("pk_rotate_eigh_kernel");
}

// out = opA(A) opB(B - Bsub) [opC(Cm)], optionally hermitised.  ops: bit0 A^H, bit1 B^H, bit2 C^H.
int hfx_matmul_batched(int dtype, int nb, size_t nk, const void* A, const void* B, const void* Cm, int ops,
                       const void* Bsub, int herm_out, void* out, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!A || !B || !out) return set_error(HFX_EINVAL, "hfx_matmul_batched: NULL argument");
  hipStream_t s = (hipStream_t)stream;
#define MM_CASE(o, a, b, c)                                                                                  \
  case o:                                                                                                   \
    if (dtype == HFX_DTYPE_C128)                                                                            \
      PK_LAUNCH_MM(double, a, b, c, nb, PKC(double, A),        \
                         PKC(double, B), PKC(double, Cm), PKM(double, out), PKC(double, Bsub), herm_out);        \
    else                                                                                                    \
      PK_LAUNCH_MM(float, a, b, c, nb, PKC(float, A),          \
                         PKC(float, B), PKC(float, Cm), PKM(float, out), PKC(float, Bsub), herm_out);            \
    break;
  switch (ops & 7) {
    MM_CASE(0, false, false, false)
    MM_CASE(1, true, false, false)
    MM_CASE(2, false, true, false)
    MM_CASE(3, true, true, false)
    MM_CASE(4, false, false, true)
    MM_CASE(5, true, false, true)
    MM_CASE(6, false, true, true)
    MM_CASE(7, true, true, true)
  }
#undef MM_CASE
  return check_launch("pk_mm_kernel");
}

static int extrap_rotate_launch(int dtype, int nb, size_t nk, int nin, const void* const* X, const double* coef,
                                const double* coef_dev, int seg_len, const void* U, void* Ft, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!X || (!coef && !coef_dev) || !U || !Ft || nin < 1 || nin > PK_XMAX || (coef_dev && seg_len < 1))
    return set_error(HFX_EINVAL, "hfx_extrap_rotate: bad argument (nin=%d)", nin);
  PkInputs in{};
  for (int v = 0; v < nin; ++v) {
    if (!X[v]) return set_error(HFX_EINVAL, "hfx_extrap_rotate: input %d is NULL", v);
    in.x[v] = X[v];
    in.c[v] = coef ? coef[v] : 0.0;
  }
  if (nk == 0) return HFX_OK;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_extrap_rotate_kernel, double, nb, nin, in, coef_dev, seg_len,
                       PKC(double, U), PKM(double, Ft));
  else
    PK_LAUNCH(pk_extrap_rotate_kernel, float, nb, nin, in, coef_dev, seg_len,
                       PKC(float, U), PKM(float, Ft));
  return check_launch("pk_extrap_rotate_kernel");
}

int hfx_extrap_rotate(int dtype, int nb, size_t nk, int nin, const void* const* X, const double* coef, const void* U,
                      void* Ft, void* stream) {
  return extrap_rotate_launch(dtype, nb, nk, nin, X, coef, nullptr, 1, U, Ft, stream);
}

int hfx_extrap_rotate_batched(int dtype, int nb, size_t nk, int seg_len, int nin, const void* const* X,
                              const double* coef_dev, const void* U, void* Ft, void* stream) {
  if (!coef_dev || seg_len < 1 || nk % (size_t)seg_len)
    return set_error(HFX_EINVAL, "hfx_extrap_rotate_batched: bad argument");
  return extrap_rotate_launch(dtype, nb, nk, nin, X, nullptr, coef_dev, seg_len, U, Ft, stream);
}

int hfx_rotate_density(int dtype, int nb, size_t nk, void* U, const void* W, const double* occ, void* P,
                       void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!U || !W || !occ || !P || U == W || U == P) return set_error(HFX_EINVAL, "hfx_rotate_density: bad argument");
  if (nk == 0) return HFX_OK;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_rotate_density_kernel, double, nb, PKM(double, U),
                       PKC(double, W), occ, PKM(double, P));
  else
    PK_LAUNCH(pk_rotate_density_kernel, float, nb, PKM(float, U),
                       PKC(float, W), occ, PKM(float, P));
  return check_launch("pk_rotate_density_kernel");
}

static int scf_resid_launch(int dtype, int nb, size_t nk, const void* F, const void* D, const void* Pn, void* comm,
                            PkRes rs, double* part, hipStream_t s) {
  if (nb == 16) {
    if (dtype == HFX_DTYPE_C128)
      hipLaunchKernelGGL((pk_scf_resid16_kernel<16, double>), dim3(nk), dim3(64), 0, s, nk, PKC(double, F),
                         PKC(double, D), PKC(double, Pn), PKM(double, comm), part, rs);
    else
      hipLaunchKernelGGL((pk_scf_resid16_kernel<16, float>), dim3(nk), dim3(64), 0, s, nk, PKC(float, F), PKC(float, D),
                         PKC(float, Pn), PKM(float, comm), part, rs);
    return check_launch("pk_scf_resid16_kernel");
  }
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_scf_resid_kernel, double, nb, nk, PKC(double, F), PKC(double, D),
                       PKC(double, Pn), PKM(double, comm), part, rs);
  else
    PK_LAUNCH(pk_scf_resid_kernel, float, nb, nk, PKC(float, F), PKC(float, D),
                       PKC(float, Pn), PKM(float, comm), part, rs);
  return check_launch("pk_scf_resid_kernel");
}

int hfx_scf_resid(int dtype, int nb, size_t nk, const void* F, const void* D, const void* Pn, void* comm,
                  double* part, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!F || !D || !part) return set_error(HFX_EINVAL, "hfx_scf_resid: NULL argument");
  return scf_resid_launch(dtype, nb, nk, F, D, Pn, comm, PkRes{}, part, (hipStream_t)stream);
}

int hfx_symmetry_average(int dtype, int nb, size_t nk, int ng, const void* A, const void* Ug, const int* src,
                         unsigned long long conj_bits, const unsigned char* mask, double weight, void* out,
                         void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!A || !Ug || !src || !out || A == out || ng < 1 || ng > SYM_MAX)
    return set_error(HFX_EINVAL, "hfx_symmetry_average: bad argument (ng=%d, out must not alias A)", ng);
  if (nk == 0) return HFX_OK;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_sym_avg_kernel, double, nb, nk, ng, PKC(double, A),
                       PKC(double, Ug), src, conj_bits, mask, weight, PKM(double, out));
  else
    PK_LAUNCH(pk_sym_avg_kernel, float, nb, nk, ng, PKC(float, A),
                       PKC(float, Ug), src, conj_bits, mask, weight, PKM(float, out));
  return check_launch("pk_sym_avg_kernel");
}

int hfx_orthonormalize(int dtype, int nb, size_t nk, void* U, int steps, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!U || steps < 1 || steps > 4) return set_error(HFX_EINVAL, "hfx_orthonormalize: bad argument (steps=%d)", steps);
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_orthonormalize_kernel, double, nb, PKM(double, U), steps, (const cx<double>*)nullptr);
  else
    PK_LAUNCH(pk_orthonormalize_kernel, float, nb, PKM(float, U), steps, (const cx<float>*)nullptr);
  return check_launch("pk_orthonormalize_kernel");
}

int hfx_orthonormalize_from(int nb, size_t nk, const void* Q32, void* U, int steps, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!Q32 || !U || steps < 1 || steps > 4)
    return set_error(HFX_EINVAL, "hfx_orthonormalize_from: bad argument (steps=%d)", steps);
  if (nk == 0) return HFX_OK;
  hipStream_t s = (hipStream_t)stream;
  if (nb <= 16)
    hipLaunchKernelGGL((pk_orthonormalize_kernel<16, double, float>), dim3(nk), dim3(64), 0, s, nb, PKM(double, U),
                       steps, PKC(float, Q32));
  else
    hipLaunchKernelGGL((pk_orthonormalize_kernel<32, double, float>), dim3(nk), dim3(256), 0, s, nb, PKM(double, U),
                       steps, PKC(float, Q32));
  return check_launch("pk_orthonormalize_kernel");
}

int hfx_scf_resid_dots(int dtype, int nb, size_t nk, const void* F, const void* D, const void* Pn, void* comm,
                       const void* const* Rs, int nr, int self, int hist_dtype, double* part, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  if (!F || !D || !comm || !part || !Rs || nr < 1 || nr > PK_XMAX || self < -1 || self >= nr ||
      (hist_dtype != dtype && !(dtype == HFX_DTYPE_C128 && hist_dtype == HFX_DTYPE_C64 && nb == 16)))
    return set_error(HFX_EINVAL, "hfx_scf_resid_dots: bad argument (nr=%d, self=%d, hist_dtype=%d)", nr, self,
                     hist_dtype);
  PkRes rs{};
  rs.nr = nr;
  rs.self = self;
  rs.r32 = hist_dtype != dtype;
  for (int v = 0; v < nr; ++v) {
    if (!Rs[v] && v != self) return set_error(HFX_EINVAL, "hfx_scf_resid_dots: NULL residual %d", v);
    rs.r[v] = Rs[v];
  }
  return scf_resid_launch(dtype, nb, nk, F, D, Pn, comm, rs, part, (hipStream_t)stream);
}

int hfx_batched_dots(int dtype, int nb, size_t nk, const void* const* Xs, const void* const* Ys, int nv, int mode,
                     double* part, void* stream) {
  if (nb < 1 || !Xs || !Ys || !part || nv < 1 || nv > 8 || (mode != 0 && mode != 1))
    return set_error(HFX_EINVAL, "hfx_batched_dots: bad argument");
  DotList L{};
  for (int v = 0; v < nv; ++v) {
    L.x[v] = Xs[v];
    L.y[v] = Ys[v];
  }
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_dots_kernel, double, nb, nk, L, nv, mode, part);
  else
    PK_LAUNCH(pk_dots_kernel, float, nb, nk, L, nv, mode, part);
  return check_launch("pk_dots_kernel");
}

int hfx_newton_dP(int dtype, int nb, size_t nk, const void* Q, const void* X, const double* p, const double* dp,
                  void* dP, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_newton_dP_kernel, double, nb, PKC(double, Q), PKC(double, X), p,
                       dp, PKM(double, dP));
  else
    PK_LAUNCH(pk_newton_dP_kernel, float, nb, PKC(float, Q), PKC(float, X), p, dp,
                       PKM(float, dP));
  return check_launch("pk_newton_dP_kernel");
}

int hfx_newton_hvp(int dtype, int nb, size_t nk, const void* Q, const void* R, const void* Ft, const void* X,
                   const double* p, const double* dp, double T, void* HX, double* Hp, double* part, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_newton_hvp_kernel, double, nb, PKC(double, Q), PKC(double, R),
                       PKC(double, Ft), PKC(double, X), p, dp, T, PKM(double, HX), Hp, part);
  else
    PK_LAUNCH(pk_newton_hvp_kernel, float, nb, PKC(float, Q), PKC(float, R),
                       PKC(float, Ft), PKC(float, X), p, dp, T, PKM(float, HX), Hp, part);
  return check_launch("pk_newton_hvp_kernel");
}

int hfx_newton_prec(int dtype, int nb, size_t nk, const void* RX, const double* rp, const double* eps,
                    const double* p, const double* st, double T, void* ZX, double* zp, double* part, void* stream) {
  if (int rc = pk_check_nb(nb)) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    PK_LAUNCH(pk_newton_prec_kernel, double, nb, PKC(double, RX), rp, eps, p, st, T,
                       PKM(double, ZX), zp, part);
  else
    PK_LAUNCH(pk_newton_prec_kernel, float, nb, PKC(float, RX), rp, eps, p, st, T,
                       PKM(float, ZX), zp, part);
  return check_launch("pk_newton_prec_kernel");
}

}  // extern "C"
