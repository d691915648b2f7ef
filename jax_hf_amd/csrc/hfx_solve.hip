// Deterministic reductions with fused scalar post-ops, and the chemical-
// potential solvers, for the GPU-resident solver loops.
//
// Reductions: per-k partial values (nval, Nk) float64 -> nval sums in a fixed
// two-level order (bitwise reproducible), optionally weighted by w[k]; thread
// 0 of the final block then applies the scalar step of the algorithm
// (gradient norm / preconditioner scale, Polak-Ribiere beta, tau0, line-search
// acceptance) directly into the device scalar state, so a CG iteration needs
// no host round trip before its convergence test.
//
// Chemical potential:
//   * one cooperative launch runs all iterations of the Newton-with-bracket
//     solver of solver.py:62-101 or the Newton of utils.py:385-442, each
//     iteration a grid-wide (count, slope) reduction followed by a grid
//     barrier; every block then evaluates the same fixed-order sum, so all
//     blocks step mu identically;
//   * the pure bisection of utils.py:353-382 (optionally twice, with the level
//     shift of reference_scf.py:87-96) runs as a multisection: ordinary
//     launches, three levels of the bisection tree per pass.
#include <hip/hip_cooperative_groups.h>

#include "hfx_common.h"
#include "hfx_internal.h"

namespace cg = cooperative_groups;

namespace hfx {

constexpr int RED_BLOCK = 256;
constexpr int RED_MAX_BLOCKS = 256;
constexpr int RED_MAX_VALS = 16;

__device__ __forceinline__ void stage1_body(const double* __restrict__ part, int nval, size_t nk,
                                            const double* __restrict__ w, unsigned wmask, double* __restrict__ s1,
                                            double* sh) {
  const size_t chunk = (nk + gridDim.x - 1) / gridDim.x;
  const size_t k0 = (size_t)blockIdx.x * chunk, k1 = k0 + chunk < nk ? k0 + chunk : nk;
  double acc[RED_MAX_VALS];
#pragma unroll
  for (int v = 0; v < RED_MAX_VALS; ++v) acc[v] = 0.0;
  for (size_t k = k0 + threadIdx.x; k < k1; k += RED_BLOCK) {
    const double wk = w ? w[k] : 1.0;
#pragma unroll
    for (int v = 0; v < RED_MAX_VALS; ++v)
      if (v < nval) acc[v] += ((wmask >> v) & 1u ? wk : 1.0) * part[(size_t)v * nk + k];
  }
  block_sum<RED_MAX_VALS>(acc, sh);
  if (threadIdx.x == 0)
    for (int v = 0; v < nval; ++v) s1[(size_t)v * gridDim.x + blockIdx.x] = acc[v];
}

struct PostArgs {
  int op;
  double T, denom_scale, count, max_step, shrink, tau_fallback;
  int restart, last, omega0, restart_every;
};

__device__ __forceinline__ void post_op(const double (&acc)[RED_MAX_VALS], int nval, double* __restrict__ out,
                                        double* __restrict__ st, const PostArgs& pa);

__device__ __forceinline__ void stage2_body(const double* __restrict__ s1, int nval, int nblk,
                                            double* __restrict__ out, double* __restrict__ st, const PostArgs& pa,
                                            double* sh) {
  double acc[RED_MAX_VALS];
#pragma unroll
  for (int v = 0; v < RED_MAX_VALS; ++v) acc[v] = 0.0;
  for (int b = threadIdx.x; b < nblk; b += RED_BLOCK)
#pragma unroll
    for (int v = 0; v < RED_MAX_VALS; ++v)
      if (v < nval) acc[v] += s1[(size_t)v * nblk + b];
  block_sum<RED_MAX_VALS>(acc, sh);
  if (threadIdx.x == 0) post_op(acc, nval, out, st, pa);
}

// The scalar step applied to the weighted sums (thread 0).
__device__ __forceinline__ void post_op(const double (&acc)[RED_MAX_VALS], int nval, double* __restrict__ out,
                                        double* __restrict__ st, const PostArgs& pa) {
  if (out)
    for (int v = 0; v < nval; ++v) out[v] = acc[v];
  const double tiny = 1e-30;
  switch (pa.op) {
    case HFX_OP_GRAD: {   // solver.py:491, 496-497
      st[HFX_ST_GRAD] = sqrt(fmax(0.0, acc[0]));
      const double eps_scale = sqrt(acc[1] / pa.count + tiny);
      st[HFX_ST_LAM] = fmax(pa.T, pa.denom_scale * eps_scale);
    } break;
    case HFX_OP_BETA: {   // solver.py:517-526
      const double num = acc[0] - acc[1] + acc[2] - acc[3];
      const double den = acc[4] + acc[5];
      double beta = den > tiny ? fmax(0.0, num / den) : 0.0;
      beta = fmin(beta, 5.0);
      const int it = (int)st[HFX_ST_ITER];
      if (pa.restart || (pa.restart_every > 0 && it % pa.restart_every == 0)) beta = 0.0;
      st[HFX_ST_BETA] = beta;
      st[HFX_ST_ITER] = (double)(it + 1);
    } break;
    case HFX_OP_TAU0: {   // solver.py:546-547
      const double dn = sqrt(fmax(0.0, acc[0]));
      const double tau0 = fmin(1.0, pa.max_step / fmax(dn, tiny));
      st[HFX_ST_DNORM] = dn;
      st[HFX_ST_TAU0] = tau0;
      st[HFX_ST_TAU] = tau0;
      st[HFX_ST_ACC] = 0.0;
      st[HFX_ST_TAUF] = tau0 * pa.tau_fallback;   // used when no trial runs (bt_max = 0)
    } break;
    case HFX_OP_TRIAL: {  // solver.py:549-579 (free_energy = E - T S, S = -sum w s)
      const double Tv = fmax(pa.T, 1e-14);
      if (pa.omega0) st[HFX_ST_OMEGA0] = acc[2] + Tv * acc[3];
      const double om = acc[0] + Tv * acc[1];
      st[HFX_ST_OMEGA] = om;
      if (om < st[HFX_ST_OMEGA0]) {
        st[HFX_ST_ACC] = 1.0;
        st[HFX_ST_TAUF] = st[HFX_ST_TAU];
      } else {
        st[HFX_ST_TAU] = st[HFX_ST_TAU] * pa.shrink;
        if (pa.last) st[HFX_ST_TAUF] = st[HFX_ST_TAU0] * pa.tau_fallback;
      }
    } break;
    case HFX_OP_NT0:
    case HFX_OP_NT1:
    case HFX_OP_NT2:      // Steihaug scalars from the reduced pairs (st: the loop's HFX_NT_* block)
      newton_tr_phase(pa.op == HFX_OP_NT0 ? 0 : pa.op == HFX_OP_NT1 ? 1 : 2, acc, nval / 2, st);
      break;
    default:
      break;
  }
}

// Two-launch form: block partials, then one workgroup sums them in block order
// and applies the scalar step.  Inside a graph this is ~2.5 us cheaper than the
// completion-ticket form, whose device-scope fence writes back the XCD's L2.
__global__ __launch_bounds__(RED_BLOCK) void reduce_stage1(const double* __restrict__ part, int nval, size_t nk,
                                                           const double* __restrict__ w, unsigned wmask,
                                                           double* __restrict__ s1, const double* __restrict__ st,
                                                           int op) {
  __shared__ double sh[RED_MAX_VALS * (RED_BLOCK / 64)];
  if (op == HFX_OP_TRIAL && st[HFX_ST_ACC] != 0.0) return;   // line search already done
  stage1_body(part, nval, nk, w, wmask, s1, sh);
}

__global__ __launch_bounds__(RED_BLOCK) void reduce_stage2(const double* __restrict__ s1, int nval, int nblk,
                                                           double* __restrict__ out, double* __restrict__ st,
                                                           PostArgs pa) {
  __shared__ double sh[RED_MAX_VALS * (RED_BLOCK / 64)];
  if (pa.op == HFX_OP_TRIAL && st[HFX_ST_ACC] != 0.0) return;
  stage2_body(s1, nval, nblk, out, st, pa, sh);
}

// Small reductions in one workgroup: no block partials, no completion ticket and
// so no device-scope fence (which writes back the XCD's L2 and costs more than
// the sum itself at these sizes).
constexpr int RED1_BLOCK = 1024;
constexpr size_t RED1_MAX_NK = 2048;   // above this one CU reads the partials too slowly

__global__ __launch_bounds__(RED1_BLOCK) void reduce_single(const double* __restrict__ part, int nval, size_t nk,
                                                            const double* __restrict__ w, unsigned wmask,
                                                            double* __restrict__ out, double* __restrict__ st,
                                                            PostArgs pa) {
  __shared__ double sh[RED_MAX_VALS * (RED1_BLOCK / 64)];
  if (pa.op == HFX_OP_TRIAL && st[HFX_ST_ACC] != 0.0) return;   // line search already done
  double acc[RED_MAX_VALS];
#pragma unroll
  for (int v = 0; v < RED_MAX_VALS; ++v) acc[v] = 0.0;
  for (size_t k = threadIdx.x; k < nk; k += RED1_BLOCK) {
    const double wk = w ? w[k] : 1.0;
#pragma unroll
    for (int v = 0; v < RED_MAX_VALS; ++v)
      if (v < nval) acc[v] += ((wmask >> v) & 1u ? wk : 1.0) * part[(size_t)v * nk + k];
  }
  block_sum<RED_MAX_VALS>(acc, sh);
  if (threadIdx.x == 0) post_op(acc, nval, out, st, pa);
}

int reduce_blocks(size_t nk) {
  size_t g = (nk + RED_BLOCK - 1) / RED_BLOCK;
  if (g > RED_MAX_BLOCKS) g = RED_MAX_BLOCKS;
  return (int)(g < 1 ? 1 : g);
}

// Inner products of (matrix, vector) pairs straight into block partials:
//   value v      = sum_k (w_k if bit v of wmask) Re <X_v[k], Y_v[k]>_F
//   value np + v = sum_k (w_k if bit np + v)     <vx_v[k], vy_v[k]>
// (the weighted inner products of the Newton subproblem, solver.py:947-991);
// a grid-stride pass over the elements, one block sum per value, then
// reduce_stage2's fixed-order sum and scalar step.  Two launches in place of a
// per-k dots pass, a vector dots pass and a two-launch reduction.
constexpr int PD_MAX = 4;
struct PairDots {
  const void* x[PD_MAX];
  const void* y[PD_MAX];
  const double* vx[PD_MAX];
  const double* vy[PD_MAX];
};

template <typename T>
__global__ __launch_bounds__(RED_BLOCK) void pair_dots_kernel(PairDots d, int np, size_t nk, int nb,
                                                              const double* __restrict__ w, unsigned wmask,
                                                              double* __restrict__ s1) {
  __shared__ double sh[RED_MAX_VALS * (RED_BLOCK / 64)];
  double acc[2 * PD_MAX];
#pragma unroll
  for (int v = 0; v < 2 * PD_MAX; ++v) acc[v] = 0.0;
  const unsigned nb2 = (unsigned)(nb * nb), unb = (unsigned)nb;
  const size_t nm = nk * nb2, nvec = nk * unb, stride = (size_t)gridDim.x * RED_BLOCK;
  for (size_t e = (size_t)blockIdx.x * RED_BLOCK + threadIdx.x; e < nm; e += stride) {
    const double wk = w ? w[(unsigned)e / nb2] : 1.0;   // nk nb^2 < 2^32
#pragma unroll
    for (int v = 0; v < PD_MAX; ++v)
      if (v < np) {
        const cx<T> x = reinterpret_cast<const cx<T>*>(d.x[v])[e];
        const cx<T> y = reinterpret_cast<const cx<T>*>(d.y[v])[e];
        const double t = (double)x.re * y.re + (double)x.im * y.im;
        acc[v] += ((wmask >> v) & 1u) ? wk * t : t;
      }
  }
  for (size_t e = (size_t)blockIdx.x * RED_BLOCK + threadIdx.x; e < nvec; e += stride) {
    const double wk = w ? w[(unsigned)e / unb] : 1.0;
#pragma unroll
    for (int v = 0; v < PD_MAX; ++v)
      if (v < np) {
        const double t = d.vx[v][e] * d.vy[v][e];
        acc[PD_MAX + v] += ((wmask >> (np + v)) & 1u) ? wk * t : t;
      }
  }
  block_sum<2 * PD_MAX>(acc, sh);
  if (threadIdx.x == 0)
    for (int v = 0; v < np; ++v) {
      s1[(size_t)v * gridDim.x + blockIdx.x] = acc[v];
      s1[(size_t)(np + v) * gridDim.x + blockIdx.x] = acc[PD_MAX + v];
    }
}

// ---------------------------------------------------------------------------
// chemical potential
// ---------------------------------------------------------------------------
constexpr int MU_BLOCK = 256;

struct MuArgs {
  const double* eps;   // (nk, nb)
  const double* w;     // (nk)
  size_t nk;
  int nb;
  double target, T, mu0, level_shift;
  int maxiter, mode, mu0_from_state;
  double* st;          // reads st[HFX_ST_MU] as mu0 when mu0_from_state; writes st[HFX_ST_MU]
  double* mu_out;      // optional extra output
  double* p_out;       // optional occupations (nk, nb)
  double* scratch;     // >= 4 * gridDim.x doubles
};

__device__ __forceinline__ double sigm(double x) { return 1.0 / (1.0 + exp(-x)); }

// Newton mu is converged when the step is below 4 ulp of max(|mu|, T), when
// N(mu) matches the target to the rounding of its own sum (64 eps N: past
// that the sign of g is noise and further steps only shrink the bracket), or
// when the bracket has closed to 4 ulp.
__device__ __forceinline__ bool mu_converged(double g, double Z, double N, double mu, double Tj, double lo,
                                             double hi) {
  const double tol_mu = 8.9e-16 * fmax(fabs(mu), Tj);
  return fabs(g / fmax(Z, 1e-18)) <= tol_mu || fabs(g) <= 64.0 * 2.220446049250313e-16 * fabs(N) ||
         hi - lo <= tol_mu;
}

// The four sums of one mu evaluation, accumulated per eigenvalue: N and its
// first three derivatives in mu, N^(j) = sum_e w_e f^(j)(x_e) / T^j with
// x_e = (mu - eps_e) / T, f' = f (1 - f), f'' = f' (1 - 2 f) and
// f^(3) = f' (1 - 6 f + 6 f^2): one exponential for all four.
constexpr int MU_NS = 4;
__device__ __forceinline__ void mu_accum(double (&acc)[MU_NS], double mu, double ev, double wk, double Tj) {
  const double f = sigm((mu - ev) / Tj);
  const double q = wk * f * (1.0 - f) / Tj;
  acc[0] += wk * f;
  acc[1] += q;
  acc[2] += q * (1.0 - 2.0 * f) / Tj;
  acc[3] += q * ((1.0 - 6.0 * f) + 6.0 * f * f) / (Tj * Tj);
}

// One safeguarded update of mu from g = N - target and N^(1..3) at mu: the
// step is the root of the cubic Taylor model
//   g + N^(1) d + N^(2) d^2 / 2 + N^(3) d^3 / 6 = 0
// reached by Newton iterations on the model from the Newton step -g / N^(1)
// (solver.py:80-99 takes that Newton step itself; both iterations converge to
// the same root, this one with a fourth-order instead of a second-order
// error, so a warm-started solve needs one or two evaluations instead of
// three or four).  A model step that fails (no positive model slope, a
// non-finite or a much longer step than Newton's) falls back to the Newton
// step; a step leaving the bracket (updated from the sign of g) becomes the
// bisection midpoint, as in the reference loop.  Returns true
// when the step leaves an error below the convergence tolerance, so the solve
// ends without a confirming evaluation: |f^(4)| <= f' bounds the model's
// remainder by d^4 / (24 T^3) (|d| <= T / 10 keeps N^(1) at the intermediate
// point within 11 % of its value at mu); a Newton step's by d^2 / (2 T).
__device__ __forceinline__ bool mu_cubic_update(double g, double N1, double N2, double N3, double& mu, double Tj,
                                                double& lo, double& hi) {
  if (g < 0) lo = mu;
  if (g > 0) hi = mu;
  const double Z = fmax(N1, 1e-18);
  const double dn = -g / Z;   // Newton step
  double d = dn;
  bool model = true;
#pragma unroll 1
  for (int i = 0; i < 6; ++i) {
    const double pm = g + d * (Z + d * (0.5 * N2 + d * (N3 / 6.0)));
    const double dp = Z + d * (N2 + 0.5 * d * N3);
    if (!(dp > 0.0)) { model = false; break; }
    d -= pm / dp;
  }
  if (!model || !isfinite(d) || fabs(d) > 2.0 * fabs(dn)) {
    d = dn;
    model = false;
  }
  double mu_new = mu + d;
  const double mu_bis = 0.5 * (lo + hi);
  bool step = true;
  if (mu_new <= lo || mu_new >= hi) {
    mu_new = mu_bis;
    step = false;
  }
  mu_new = fmin(fmax(mu_new, lo), hi);
  if (!isfinite(mu_new)) {
    mu_new = mu_bis;
    step = false;
  }
  mu = mu_new;
  if (!step) return false;
  const double tol = 8.9e-16 * fmax(fabs(mu_new), Tj);
  if (model) {
    const double d2 = d * d;
    return fabs(d) <= 0.1 * Tj && d2 * d2 <= 3.0 * Tj * Tj * Tj * tol;   // d^4 / (24 T^3) <= tol / 8
  }
  return d * d <= 0.5 * Tj * tol;
}

// Block-wide sums of NV interleaved values per entry, src[stride q + v] for
// q < n (strided per thread, then a block tree; every block evaluates the
// same expression tree, so all blocks get bitwise-identical totals).  All
// threads receive the totals.
template <int NV>
__device__ __forceinline__ void block_gather(const double* src, int n, int stride, double* sh, double (&tot)[NV]) {
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += src[(size_t)stride * q + v];
  }
  block_sum<NV>(acc, sh);
  __shared__ double t[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) t[v] = acc[v];
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < NV; ++v) tot[v] = t[v];
  __syncthreads();
}

// Block-wide sum of the first n entries of a global array (strided per
// thread, then a block tree): every block of the grid evaluates the same
// expression tree over the same data, so all blocks get bitwise-identical totals.
__device__ __forceinline__ void block_gather2(const double* src, int n, double* sh, double& ta, double& tb) {
  double acc[2] = {0.0, 0.0};
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    acc[0] += src[2 * q];
    acc[1] += src[2 * q + 1];
  }
  block_sum<2>(acc, sh);
  __shared__ double tot[2];
  if (threadIdx.x == 0) {
    tot[0] = acc[0];
    tot[1] = acc[1];
  }
  __syncthreads();
  ta = tot[0];
  tb = tot[1];
  __syncthreads();
}

// grid-wide deterministic sum of two values; all threads receive the totals.
// `slot` alternates between calls so a block that runs ahead never overwrites
// partials another block is still reading.
__device__ void grid_sum2(cg::grid_group& grid, double a, double b, double* scratch, int slot, double* sh,
                          double& ta, double& tb) {
  double acc[2] = {a, b};
  block_sum<2>(acc, sh);
  double* part = scratch + (size_t)slot * 2 * gridDim.x;
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
  grid.sync();
  block_gather2(part, (int)gridDim.x, sh, ta, tb);
}

__device__ void grid_minmax(cg::grid_group& grid, const double* v, size_t n, double* scratch, double* sh,
                            double& mn, double& mx) {
  double lo = INFINITY, hi = -INFINITY;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    lo = fmin(lo, v[e]);
    hi = fmax(hi, v[e]);
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0) { sh[wid] = lo; sh[nw + wid] = hi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 0; w < nw; ++w) { lo = fmin(lo, sh[w]); hi = fmax(hi, sh[nw + w]); }
    scratch[2 * blockIdx.x] = lo;
    scratch[2 * blockIdx.x + 1] = hi;
  }
  grid.sync();
  double a = INFINITY, b = -INFINITY;
  for (unsigned q = threadIdx.x; q < gridDim.x; q += blockDim.x) {
    a = fmin(a, scratch[2 * q]);
    b = fmax(b, scratch[2 * q + 1]);
  }
  a = wave_min(a);
  b = wave_max(b);
  __syncthreads();
  if (lane == 0) { sh[wid] = a; sh[nw + wid] = b; }
  __syncthreads();
  a = INFINITY;
  b = -INFINITY;
  for (int w = 0; w < nw; ++w) { a = fmin(a, sh[w]); b = fmax(b, sh[nw + w]); }
  mn = a;
  mx = b;
  __syncthreads();
}

__global__ __launch_bounds__(MU_BLOCK) void mu_kernel(MuArgs a) {
  cg::grid_group grid = cg::this_grid();
  __shared__ double sh[2 * (MU_BLOCK / 64)];
  const size_t n = a.nk * a.nb, stride = (size_t)gridDim.x * blockDim.x;
  double mu;
  if (a.mode == HFX_MU_NEWTON_BRACKET) {   // solver.py:62-101
    const double Tj = fmax(a.T, 1e-12);
    double mn, mx;
    grid_minmax(grid, a.eps, n, a.scratch, sh, mn, mx);
    double lo = mn - 50.0 * Tj, hi = mx + 50.0 * Tj;
    mu = a.mu0_from_state ? a.st[HFX_ST_MU] : a.mu0;
    mu = fmin(fmax(mu, lo), hi);
    for (int it = 0; it < a.maxiter; ++it) {
      double Np = 0.0, Zp = 0.0;
      for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const double pv = sigm((mu - a.eps[e]) / Tj);
        const double wk = a.w[(unsigned)e / (unsigned)a.nb];   // n < 2^32
        Np += wk * pv;
        Zp += wk * pv * (1.0 - pv) / Tj;
      }
      double N, Z;
      grid_sum2(grid, Np, Zp, a.scratch, 1 + (it & 1), sh, N, Z);
      const double g = N - a.target;
      // every block takes this branch together, the totals being bitwise identical
      if (mu_converged(g, Z, N, mu, Tj, lo, hi)) break;
      if (g < 0) lo = mu;
      if (g > 0) hi = mu;
      double mu_new = mu - g / fmax(Z, 1e-18);
      const double mu_bis = 0.5 * (lo + hi);
      if (mu_new <= lo || mu_new >= hi) mu_new = mu_bis;
      mu_new = fmin(fmax(mu_new, lo), hi);
      if (!isfinite(mu_new)) mu_new = mu_bis;
      mu = mu_new;
    }
  } else {   // HFX_MU_NEWTON, utils.py:385-442 (bisection: ms_bisection_launch)
    const double Tj = fmax(a.T, 1e-12);
    double mn, mx;
    grid_minmax(grid, a.eps, n, a.scratch, sh, mn, mx);
    const double span = mx - mn + 10.0 * fmax(Tj, 1e-6);
    double lo = mn - span, hi = mx + span;
    mu = 0.5 * (lo + hi);
    double N = 0.0, Z = 0.0;
    for (int it = 0; it <= a.maxiter; ++it) {
      double Np = 0.0, Zp = 0.0;
      for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const double pv = sigm((mu - a.eps[e]) / Tj);
        const double wk = a.w[(unsigned)e / (unsigned)a.nb];   // n < 2^32
        Np += wk * pv;
        Zp += wk * pv * (1.0 - pv) / Tj;
      }
      grid_sum2(grid, Np, Zp, a.scratch, 1 + (it & 1), sh, N, Z);
      if (it == a.maxiter) break;   // final residual check below
      const double g = N - a.target;
      // converged: the Newton step is below 4 ulp of max(|mu|, T) (every block
      // takes this branch together, the totals being bitwise identical); no
      // noise-floor stop here: the final check below wants |g| <= 1e-12
      if (fabs(g / fmax(Z, 1e-18)) <= 8.9e-16 * fmax(fabs(mu), Tj)) break;
      if (g < 0) lo = mu;
      if (g > 0) hi = mu;
      double mu_new = mu - g / fmax(Z, 1e-18);
      const double mu_bis = 0.5 * (lo + hi);
      if (mu_new <= lo || mu_new >= hi) mu_new = mu_bis;
      mu_new = fmin(fmax(mu_new, lo), hi);
      if (!isfinite(mu_new)) mu_new = mu_bis;
      mu = mu_new;
    }
    if (fabs(N - a.target) > 1e-12) mu = 0.5 * (lo + hi);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (a.st) a.st[HFX_ST_MU] = mu;
    if (a.mu_out) *a.mu_out = mu;
  }
  if (a.p_out) {
    const double Tj = fmax(a.T, 1e-12);
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
      a.p_out[e] = sigm((mu - a.eps[e]) / Tj);   // solver.py:394, 592
    }
  }
}

// ---- graph-capturable mu (Newton with bracket, two launches per step) -------
constexpr int MUS_BLOCK = 256;
constexpr size_t MU_PART_OFF = (size_t)RED_MAX_VALS * RED_MAX_BLOCKS;   // mu partials
constexpr size_t MU_TICKET_OFF = MU_PART_OFF + 8 * 512;                 // end of the mu partials

struct MuStepArgs {
  const double* eps;
  const double* w;
  size_t nk;
  int nb;
  double target, T, mu0;
  int mu0_from_state;
  double* st;
  double* part;        // 6 * gridDim.x doubles
  double* p_out;
};

// Block min / max of eps and the four sums at mu into out[0..5]: lo, hi, N,
// N^(1), N^(2), N^(3) (min / max only with `minmax`).
__device__ __forceinline__ void mu_eval_block(const double* eps, const double* w, size_t n, int nb, size_t stride,
                                              double mu, double Tj, bool minmax, double* sh, double* out) {
  double lo = INFINITY, hi = -INFINITY, acc[MU_NS] = {0.0, 0.0, 0.0, 0.0};
  for (size_t e = (size_t)blockIdx.x * MUS_BLOCK + threadIdx.x; e < n; e += stride) {
    const double ev = eps[e];
    if (minmax) {
      lo = fmin(lo, ev);
      hi = fmax(hi, ev);
    }
    mu_accum(acc, mu, ev, w[(unsigned)e / (unsigned)nb], Tj);   // n < 2^32
  }
  block_sum<MU_NS>(acc, sh);
  if (minmax) {
    lo = wave_min(lo);
    hi = wave_max(hi);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = MUS_BLOCK / 64;
    if (lane == 0) { sh[wid] = lo; sh[nw + wid] = hi; }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int q = 0; q < nw; ++q) { lo = fmin(lo, sh[q]); hi = fmax(hi, sh[nw + q]); }
  }
  if (threadIdx.x == 0) {
    out[0] = lo;
    out[1] = hi;
#pragma unroll
    for (int v = 0; v < MU_NS; ++v) out[2 + v] = acc[v];
  }
}

// Convergence test or one update from the sums at st[MU]; bracket in st.
__device__ __forceinline__ void mu_finish_step(double* st, const double (&t)[MU_NS], double target, double Tj) {
  const double mu = st[HFX_ST_MU];
  const double g = t[0] - target;   // solver.py:80-99
  double lo = st[HFX_ST_MULO], hi = st[HFX_ST_MUHI];
  if (mu_converged(g, t[1], t[0], mu, Tj, lo, hi)) {
    st[HFX_ST_MUDONE] = 1.0;
    return;
  }
  double mu_new = mu;
  const bool fin = mu_cubic_update(g, t[1], t[2], t[3], mu_new, Tj, lo, hi);
  st[HFX_ST_MU] = mu_new;
  st[HFX_ST_MULO] = lo;
  st[HFX_ST_MUHI] = hi;
  if (fin) st[HFX_ST_MUDONE] = 1.0;
}

// Bracket lo / hi = min / max(eps) -/+ 50 T (solver.py:70-72), mu0 clamped into
// it; when mu0 was inside (so the sums t were taken at the clamped mu) the
// first step as well.
__device__ __forceinline__ void mu_finish_bracket(double* st, double mu0, double mn, double mx,
                                                  const double (&t)[MU_NS], double target, double Tj) {
  const double blo = mn - 50.0 * Tj, bhi = mx + 50.0 * Tj;
  const double mu = fmin(fmax(mu0, blo), bhi);
  st[HFX_ST_MU] = mu;
  st[HFX_ST_MULO] = blo;
  st[HFX_ST_MUHI] = bhi;
  st[HFX_ST_MUDONE] = 0.0;
  if (mu == mu0) mu_finish_step(st, t, target, Tj);
}

// Block min / max over the first n (lo, hi) pairs at stride 6, order-free; the
// result in thread 0.
__device__ __forceinline__ void block_minmax6(const double* src, int n, double* sh, double& mn, double& mx) {
  mn = INFINITY;
  mx = -INFINITY;
  for (int b = threadIdx.x; b < n; b += blockDim.x) {
    mn = fmin(mn, src[6 * b]);
    mx = fmax(mx, src[6 * b + 1]);
  }
  mn = wave_min(mn);
  mx = wave_max(mx);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) { sh[wid] = mn; sh[nw + wid] = mx; }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int q = 0; q < nw; ++q) { mn = fmin(mn, sh[q]); mx = fmax(mx, sh[nw + q]); }
  __syncthreads();
}

// The bracket and the first evaluation, at the warm-start mu0, share one pass
// over eps (block partials, 6 doubles per block).  Each later launch of the
// chain finishes the previous evaluation itself -- every block sums the
// previous launch's partials in the same order, so all blocks reach the same
// state bit for bit -- and then evaluates at the new mu; block 0 records the
// state in the slot the next launch reads.  Two partial buffers and two state
// slots alternate, so a launch never reads what it writes.  One launch per
// step: no one-block finish launch in between, and no completion ticket (whose
// device-scope fence writes back the XCD's L2).  A launch after convergence
// only carries the state forward.
constexpr int MU_CHAIN_MAX_BLK = 320;
constexpr size_t MU_CHAIN_BUF = 6 * (size_t)MU_CHAIN_MAX_BLK;   // doubles per partial buffer
static_assert(2 * MU_CHAIN_BUF + 8 <= MU_TICKET_OFF - MU_PART_OFF, "mu chain scratch");

__device__ __forceinline__ double* mu_chain_buf(const MuStepArgs& a, int j) {
  return a.part + (size_t)(j & 1) * MU_CHAIN_BUF;
}
__device__ __forceinline__ double* mu_chain_slot(const MuStepArgs& a, int j) {   // mu, lo, hi, done
  return a.part + 2 * MU_CHAIN_BUF + 4 * (j & 1);
}

__global__ __launch_bounds__(MUS_BLOCK) void mu_bracket_part(MuStepArgs a) {
  __shared__ double sh[MU_NS * (MUS_BLOCK / 64)];
  const size_t n = a.nk * a.nb, stride = (size_t)gridDim.x * MUS_BLOCK;
  const double mu0 = a.mu0_from_state ? a.st[HFX_ST_MU] : a.mu0;
  mu_eval_block(a.eps, a.w, n, a.nb, stride, mu0, fmax(a.T, 1e-12), true, sh, mu_chain_buf(a, 0) + 6 * (size_t)blockIdx.x);
}

// The state after evaluation j - 1 of the chain (evaluation 0 is the bracket
// pass) into sst[HFX_ST_MU / MULO / MUHI / MUDONE], the same in every block.
__device__ __forceinline__ void mu_chain_finish(const MuStepArgs& a, int nblk, int j, double* sst, double* sh) {
  const double Tj = fmax(a.T, 1e-12);
  const double* src = mu_chain_buf(a, j - 1);
  if (j == 1) {
    double mn, mx, t[MU_NS];
    block_minmax6(src, nblk, sh, mn, mx);
    block_gather<MU_NS>(src + 2, nblk, 6, sh, t);
    if (threadIdx.x == 0) {
      const double mu0 = a.mu0_from_state ? a.st[HFX_ST_MU] : a.mu0;
      mu_finish_bracket(sst, mu0, mn, mx, t, a.target, Tj);
    }
  } else {
    const double* slot = mu_chain_slot(a, j - 1);
    const bool done = slot[3] != 0.0;
    if (threadIdx.x == 0) {
      sst[HFX_ST_MU] = slot[0];
      sst[HFX_ST_MULO] = slot[1];
      sst[HFX_ST_MUHI] = slot[2];
      sst[HFX_ST_MUDONE] = slot[3];
    }
    if (!done) {
      double t[MU_NS];
      block_gather<MU_NS>(src + 2, nblk, 6, sh, t);
      if (threadIdx.x == 0) mu_finish_step(sst, t, a.target, Tj);
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(MUS_BLOCK) void mu_chain_step(MuStepArgs a, int nblk, int j) {
  __shared__ double sh[MU_NS * (MUS_BLOCK / 64)];
  __shared__ double sst[HFX_ST_MUDONE + 1];
  mu_chain_finish(a, nblk, j, sst, sh);
  const double mu = sst[HFX_ST_MU];
  const bool done = sst[HFX_ST_MUDONE] != 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double* o = mu_chain_slot(a, j);
    o[0] = mu;
    o[1] = sst[HFX_ST_MULO];
    o[2] = sst[HFX_ST_MUHI];
    o[3] = sst[HFX_ST_MUDONE];
  }
  if (done) return;
  const size_t n = a.nk * a.nb, stride = (size_t)gridDim.x * MUS_BLOCK;
  mu_eval_block(a.eps, a.w, n, a.nb, stride, mu, fmax(a.T, 1e-12), false, sh, mu_chain_buf(a, j) + 6 * (size_t)blockIdx.x);
}

// Last launch: finish evaluation j - 1 into the solver state, and the
// occupations at the final mu (all blocks) when p_out is given.
__global__ __launch_bounds__(MUS_BLOCK) void mu_chain_fin(MuStepArgs a, int nblk, int j) {
  __shared__ double sh[MU_NS * (MUS_BLOCK / 64)];
  __shared__ double sst[HFX_ST_MUDONE + 1];
  mu_chain_finish(a, nblk, j, sst, sh);
  const double mu = sst[HFX_ST_MU];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.st[HFX_ST_MU] = mu;
    a.st[HFX_ST_MULO] = sst[HFX_ST_MULO];
    a.st[HFX_ST_MUHI] = sst[HFX_ST_MUHI];
    a.st[HFX_ST_MUDONE] = sst[HFX_ST_MUDONE];
  }
  if (!a.p_out) return;
  const size_t n = a.nk * a.nb, stride = (size_t)gridDim.x * MUS_BLOCK;
  const double Tj = fmax(a.T, 1e-12);
  for (size_t e = (size_t)blockIdx.x * MUS_BLOCK + threadIdx.x; e < n; e += stride)
    a.p_out[e] = sigm((mu - a.eps[e]) / Tj);   // solver.py:394, 592
}

// ---- batched solves: segmented reductions and mu -------------------------------
// nseg consecutive solves of seg_len k-points; blockIdx.y is the solve, each
// solve has its own 32-double state vector (st + HFX_ST_LEN * s), block
// partials and completion ticket; frozen solves skip their launches.
constexpr int SEG_BLOCKS = 32;   // blocks per solve (k-points per solve <= 8192 at 256 threads each)
constexpr int MU_SEG_VALS = 6;   // mu partials per block: lo, hi, N, N^(1), N^(2), N^(3)

struct SegScratch {
  double* s1;          // [nseg][RED_MAX_VALS][SEG_BLOCKS]
  double* mpart;       // [nseg][MU_SEG_VALS * SEG_BLOCKS]: per block lo, hi and the four mu sums
  unsigned* tickets;   // [nseg][2]
};

__host__ __device__ inline SegScratch seg_scratch(double* base, int nseg) {
  SegScratch t;
  t.s1 = base;
  t.mpart = base + (size_t)nseg * RED_MAX_VALS * SEG_BLOCKS;
  t.tickets = reinterpret_cast<unsigned*>(t.mpart + (size_t)nseg * MU_SEG_VALS * SEG_BLOCKS);
  return t;
}

__device__ __forceinline__ bool seg_last_block(unsigned* ticket, unsigned nblk) {
  __shared__ bool last;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(ticket, 1u) == nblk - 1;
  __syncthreads();
  if (last) __threadfence();
  return last;
}

__global__ __launch_bounds__(RED_BLOCK) void reduce_seg_kernel(const double* __restrict__ part, int nval,
                                                               size_t nk_total, int seg_len,
                                                               const double* __restrict__ w, unsigned wmask,
                                                               double* __restrict__ st_all, PostArgs pa,
                                                               int out_slot, SegScratch scr) {
  __shared__ double sh[RED_MAX_VALS * (RED_BLOCK / 64)];
  const int sidx = blockIdx.y;
  double* st = st_all + (size_t)sidx * HFX_ST_LEN;
  if (st[HFX_ST_FROZEN] != 0.0) return;
  if (pa.op == HFX_OP_TRIAL && st[HFX_ST_ACC] != 0.0) return;
  const size_t chunk = ((size_t)seg_len + gridDim.x - 1) / gridDim.x;
  const size_t k0 = (size_t)sidx * seg_len + (size_t)blockIdx.x * chunk;
  size_t k1 = k0 + chunk;
  const size_t kend = (size_t)(sidx + 1) * seg_len;
  if (k1 > kend) k1 = kend;
  double acc[RED_MAX_VALS];
#pragma unroll
  for (int v = 0; v < RED_MAX_VALS; ++v) acc[v] = 0.0;
  for (size_t k = k0 + threadIdx.x; k < k1; k += RED_BLOCK) {
    const double wk = w ? w[k] : 1.0;
#pragma unroll
    for (int v = 0; v < RED_MAX_VALS; ++v)
      if (v < nval) acc[v] += ((wmask >> v) & 1u ? wk : 1.0) * part[(size_t)v * nk_total + k];
  }
  block_sum<RED_MAX_VALS>(acc, sh);
  double* s1 = scr.s1 + (size_t)sidx * RED_MAX_VALS * SEG_BLOCKS;
  if (threadIdx.x == 0)
    for (int v = 0; v < nval; ++v) s1[(size_t)v * gridDim.x + blockIdx.x] = acc[v];
  unsigned* ticket = scr.tickets + 2 * sidx;
  if (!seg_last_block(ticket, gridDim.x)) return;
  if (threadIdx.x == 0) *ticket = 0u;
  if (pa.op == HFX_OP_SUM && out_slot >= 0) {
    double tot[RED_MAX_VALS];
#pragma unroll
    for (int v = 0; v < RED_MAX_VALS; ++v) tot[v] = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += RED_BLOCK)
#pragma unroll
      for (int v = 0; v < RED_MAX_VALS; ++v)
        if (v < nval) tot[v] += s1[(size_t)v * gridDim.x + b];
    block_sum<RED_MAX_VALS>(tot, sh);
    if (threadIdx.x == 0)
      for (int v = 0; v < nval; ++v) st[out_slot + v] = tot[v];
    return;
  }
  stage2_body(s1, nval, (int)gridDim.x, nullptr, st, pa, sh);
}

struct MuSegArgs {
  const double* eps;
  const double* w;
  const double* targets;   // electrons per solve (normalised weights)
  int seg_len, nb;
  double T;
  double* st_all;
  double* p_out;
  SegScratch scr;
};

// Batched form of the bracket pass with the first evaluation (the last block
// of each solve finishes it through the solve's completion ticket).
__global__ __launch_bounds__(MUS_BLOCK) void mu_bracket_seg_kernel(MuSegArgs a) {
  __shared__ double sh[MU_NS * (MUS_BLOCK / 64)];
  const int sidx = blockIdx.y;
  double* st = a.st_all + (size_t)sidx * HFX_ST_LEN;
  if (st[HFX_ST_FROZEN] != 0.0) return;
  const size_t n = (size_t)a.seg_len * a.nb, base = (size_t)sidx * n, stride = (size_t)gridDim.x * MUS_BLOCK;
  const double Tj = fmax(a.T, 1e-12);
  const double mu0 = st[HFX_ST_MU];   // warm start from the state
  double* mp = a.scr.mpart + (size_t)sidx * MU_SEG_VALS * SEG_BLOCKS;
  mu_eval_block(a.eps + base, a.w + (size_t)sidx * a.seg_len, n, a.nb, stride, mu0, Tj, true, sh,
                mp + 6 * (size_t)blockIdx.x);
  unsigned* ticket = a.scr.tickets + 2 * sidx + 1;
  if (!seg_last_block(ticket, gridDim.x)) return;
  double mn, mx, t[MU_NS];
  block_minmax6(mp, (int)gridDim.x, sh, mn, mx);
  block_gather<MU_NS>(mp + 2, (int)gridDim.x, 6, sh, t);
  if (threadIdx.x == 0) {
    mu_finish_bracket(st, mu0, mn, mx, t, a.targets[sidx], Tj);
    *ticket = 0u;
  }
}

__global__ __launch_bounds__(MUS_BLOCK) void mu_step_seg_kernel(MuSegArgs a) {
  const int sidx = blockIdx.y;
  double* st = a.st_all + (size_t)sidx * HFX_ST_LEN;
  if (st[HFX_ST_FROZEN] != 0.0 || st[HFX_ST_MUDONE] != 0.0) return;
  __shared__ double sh[MU_NS * (MUS_BLOCK / 64)];
  const size_t n = (size_t)a.seg_len * a.nb, base = (size_t)sidx * n, stride = (size_t)gridDim.x * MUS_BLOCK;
  const double Tj = fmax(a.T, 1e-12);
  double* mp = a.scr.mpart + (size_t)sidx * MU_SEG_VALS * SEG_BLOCKS;
  mu_eval_block(a.eps + base, a.w + (size_t)sidx * a.seg_len, n, a.nb, stride, st[HFX_ST_MU], Tj, false, sh,
                mp + 6 * (size_t)blockIdx.x);
  unsigned* ticket = a.scr.tickets + 2 * sidx + 1;
  if (!seg_last_block(ticket, gridDim.x)) return;
  double t[MU_NS];
  block_gather<MU_NS>(mp + 2, (int)gridDim.x, 6, sh, t);
  if (threadIdx.x == 0) {
    mu_finish_step(st, t, a.targets[sidx], Tj);
    *ticket = 0u;
  }
}

__global__ __launch_bounds__(MUS_BLOCK) void mu_occ_seg_kernel(MuSegArgs a) {
  const int sidx = blockIdx.y;
  const double* st = a.st_all + (size_t)sidx * HFX_ST_LEN;
  if (st[HFX_ST_FROZEN] != 0.0) return;
  const size_t n = (size_t)a.seg_len * a.nb, base = (size_t)sidx * n, stride = (size_t)gridDim.x * MUS_BLOCK;
  const double Tj = fmax(a.T, 1e-12), mu = st[HFX_ST_MU];
  for (size_t e = (size_t)blockIdx.x * MUS_BLOCK + threadIdx.x; e < n; e += stride)
    a.p_out[base + e] = sigm((mu - a.eps[base + e]) / Tj);
}

// ---- Pulay C-DIIS coefficients on the device (reference_scf.py:99-121) ----------
// One wavefront per solve: B[idx][:] = B[:][idx] = brow (the new row, already
// reduced), then the masked, max-normalised augmented system
//   [ Bn + diag(1 - a)  -a ] [c]   [ 0]
//   [ -a^T               0 ] [l] = [-1],   a_i = (i < count),
// solved by LU with partial pivoting (first maximal pivot, as LAPACK's getrf).
// Lane r holds row r of the (padded) system in registers; the pivot search,
// row swap and pivot-row broadcast are wave shuffles, so the dependent chain
// per column is a few shuffles and one FMA per element.  coef = c when
// count >= 2 and the solve is regular and finite, else the plain step e_idx;
// coef[nh..7] = 0.
constexpr int DIIS_MAX = 8;
constexpr int DIIS_M = DIIS_MAX + 1;   // augmented size; rows past nh + 1 are identity padding
// s1 (optional): first sum hfx_reduce's block partials of nval values over
// nblk blocks in reduce_stage2's order (four 64-lane wave sums, then those in
// order), write the totals to out, and take the B row from out + 2.
__global__ __launch_bounds__(64) void diis_coef_kernel(const double* __restrict__ brow_all, size_t brow_stride,
                                                       double* __restrict__ B_all, int nsolve, int nh, int idx,
                                                       int count, double* __restrict__ coef_all,
                                                       const double* __restrict__ s1, int nblk, int nval,
                                                       double* __restrict__ out) {
  const int sidx = blockIdx.x;
  if (sidx >= nsolve) return;
  const double* brow = brow_all + (size_t)sidx * brow_stride;
  double* B = B_all + (size_t)sidx * nh * nh;
  double* coef = coef_all + (size_t)sidx * DIIS_MAX;
  const int r = threadIdx.x;
  const int N = nh;
  double br;
  if (s1) {
    double mine = 0.0;   // lane v keeps total v
    for (int v = 0; v < nval; ++v) {
      double tot = 0.0;
#pragma unroll
      for (int wv = 0; wv < RED_BLOCK / 64; ++wv) {
        const int b = wv * 64 + r;
        double x = 0.0;
        if (b < nblk) x += s1[(size_t)v * nblk + b];
        tot += wave_sum(x);
      }
      if (r == v) mine = tot;
      if (r == 0) out[v] = tot;
    }
    br = __shfl(mine, r + 2 < 64 ? r + 2 : 63);
    if (r >= nh) br = 0.0;
  } else {
    br = r < nh ? brow[r] : 0.0;
  }
  double a[DIIS_M];
#pragma unroll
  for (int j = 0; j < DIIS_M; ++j) {
    const double bj = __shfl(br, j);
    double v = 0.0;
    if (r < count && j < count) v = r == idx ? bj : (j == idx ? br : B[r * nh + j]);
    a[j] = v;
  }
  if (r < nh) {
    B[idx * nh + r] = br;
    B[r * nh + idx] = br;
  }
  double mx = 0.0;
#pragma unroll
  for (int j = 0; j < DIIS_M; ++j) mx = fmax(mx, fabs(a[j]));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  mx = fmax(mx, 1e-30);
  const bool ar = r < count;
#pragma unroll
  for (int j = 0; j < DIIS_M; ++j) {
    double v;
    if (r < N) v = j < N ? a[j] / mx + ((j == r && !ar) ? 1.0 : 0.0) : (j == N ? (ar ? -1.0 : 0.0) : 0.0);
    else if (r == N) v = j < N ? (j < count ? -1.0 : 0.0) : 0.0;
    else v = (j == r) ? 1.0 : 0.0;   // padding rows (and lanes past the system)
    a[j] = v;
  }
  double x = r == N ? -1.0 : 0.0;
  bool ok = true;
#pragma unroll
  for (int c = 0; c < DIIS_M; ++c) {
    // pivot: the first row r >= c with the largest |a[r][c]|
    double bv = (r >= c && r < DIIS_M) ? fabs(a[c]) : -1.0;
    int bi = r;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double ov = __shfl_xor(bv, o);
      const int oi = __shfl_xor(bi, o);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (bv == 0.0) ok = false;
    const int src = r == c ? bi : (r == bi ? c : r);
#pragma unroll
    for (int j = c; j < DIIS_M; ++j) a[j] = __shfl(a[j], src);
    x = __shfl(x, src);
    double pj[DIIS_M];
#pragma unroll
    for (int j = c; j < DIIS_M; ++j) pj[j] = __shfl(a[j], c);
    const double px = __shfl(x, c);
    const double inv = bv == 0.0 ? 0.0 : 1.0 / pj[c];
    if (r > c && r < DIIS_M) {
      const double f = a[c] * inv;
#pragma unroll
      for (int j = c; j < DIIS_M; ++j) a[j] -= f * pj[j];
      x -= f * px;
    }
  }
  double xs[DIIS_M];
#pragma unroll
  for (int c = DIIS_M - 1; c >= 0; --c) {
    double t = x;
#pragma unroll
    for (int j = c + 1; j < DIIS_M; ++j) t -= a[j] * xs[j];
    xs[c] = __shfl(t / (a[c] == 0.0 ? 1.0 : a[c]), c);
  }
#pragma unroll
  for (int i = 0; i < DIIS_MAX; ++i)
    if (i < N) ok = ok && isfinite(xs[i]);
  const bool use = ok && count >= 2;
  if (r < DIIS_MAX) {
    double mine = 0.0;
#pragma unroll
    for (int j = 0; j < DIIS_MAX; ++j)
      if (j == r) mine = xs[j];
    coef[r] = r < N ? (use ? (r < count ? mine : 0.0) : (r == idx ? 1.0 : 0.0)) : 0.0;
  }
}

// ---- bisection by multisection (utils.py:353-382, reference_scf.py:87-96) ---
// The bisection walks a binary tree: each level takes mid = 0.5 * (lo + hi)
// and keeps [lo, mid] when the count at mid exceeds the target, else
// [mid, hi].  One pass evaluates the counts at all 2^L - 1 midpoints of the
// next L <= 3 levels of that tree (every block builds the same tree from the
// same bracket with the same expression) and writes per-block node counts;
// the next pass starts with every block summing those counts in one fixed
// order (so all blocks get bitwise-identical totals) and walking the previous
// pass's L levels: the same midpoints and the same tests as the level-by-level
// loop, in ceil(maxiter / 3) ordinary launches (no grid barrier, so a graph may
// capture them), and a one-workgroup launch walks the last pass.  Three
// levels per pass balance the launch count against the 2^L - 1 Fermi
// functions per eigenvalue (six levels spent 64 us per pass in them at
// 96 x 96 x 16).  An eigenvalue 37.5 T or more below the bracket contributes
// exactly its weight at every node (1 + e^x rounds to 1); one 37.5 T above it
// contributes less than 6e-17 of its weight and is skipped.
constexpr int MS_BLOCK = 256;
constexpr int MS_MAX_BLOCKS = 512;
constexpr int MS_LEVELS = 3;
constexpr int MS_NODES = 8;    // 2^3 - 1 tree nodes, padded
constexpr size_t MS_OFF = MU_TICKET_OFF + 8;                         // state
constexpr size_t MS_MM_OFF = MS_OFF + 16;                            // [block] min, max
constexpr size_t MS_PART_OFF = MS_MM_OFF + 2 * (size_t)MS_MAX_BLOCKS;  // [2][block][MS_NODES]
constexpr size_t MS_END = MS_PART_OFF + 2 * (size_t)MS_MAX_BLOCKS * MS_NODES;
// state: bracket slot s (pass parity) at 4 s: lo, hi, collapsed; then mu, mu_raw
enum { MS_LO = 0, MS_HI = 1, MS_DONE = 2, MS_MU = 8, MS_MURAW = 9 };
constexpr double MS_SAT = 37.5;   // e^-37.5 < 2^-54

struct MsArgs {
  const double* eps;   // (n) eigenvalues, nb per k-point
  const double* w;     // (n / nb) weights
  size_t n;
  int nb, nblk;
  double T, target, level_shift;
  int shift;           // count the level-shifted values (reads state[MS_MURAW])
  double* state;       // scratch + MS_OFF
  double* mm;          // scratch + MS_MM_OFF
  double* part;        // scratch + MS_PART_OFF
};

// eigenvalue e, level-shifted when a.shift (reference_scf.py:91-92)
__device__ __forceinline__ double ms_value(const MsArgs& a, size_t e, double mu_raw) {
  const double x = a.eps[e];
  if (!a.shift) return x;
  const double f = sigm(-(x - mu_raw) / (a.T + 1e-12));
  return x + a.level_shift * (1.0 - f);
}

__global__ __launch_bounds__(MS_BLOCK) void ms_minmax_part(MsArgs a) {
  __shared__ double sh[2 * (MS_BLOCK / 64)];
  const double mu_raw = a.shift ? a.state[MS_MURAW] : 0.0;
  double lo = INFINITY, hi = -INFINITY;
  for (size_t e = (size_t)blockIdx.x * MS_BLOCK + threadIdx.x; e < a.n; e += (size_t)gridDim.x * MS_BLOCK) {
    const double v = ms_value(a, e, mu_raw);
    lo = fmin(lo, v);
    hi = fmax(hi, v);
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { sh[wid] = lo; sh[MS_BLOCK / 64 + wid] = hi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < MS_BLOCK / 64; ++q) { lo = fmin(lo, sh[q]); hi = fmax(hi, sh[MS_BLOCK / 64 + q]); }
    a.mm[2 * blockIdx.x] = lo;
    a.mm[2 * blockIdx.x + 1] = hi;
  }
}

// The bracket at the start of pass `pass` into br[0..2] (lo, hi, collapsed),
// identical in every block.  Pass 0: the initial bracket from the min / max
// partials, span = max - min + 10 max(T, 1e-6) on both sides (utils.py:366-370).
// Later passes: the previous pass's bracket walked down its `prev_levels`
// levels with the summed node counts.
__device__ void ms_bracket(const MsArgs& a, int pass, int prev_levels, double* sh, double* br) {
  if (pass == 0) {
    double lo = INFINITY, hi = -INFINITY;
    for (int b = threadIdx.x; b < a.nblk; b += MS_BLOCK) {
      lo = fmin(lo, a.mm[2 * b]);
      hi = fmax(hi, a.mm[2 * b + 1]);
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) { sh[wid] = lo; sh[MS_BLOCK / 64 + wid] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int q = 0; q < MS_BLOCK / 64; ++q) { lo = fmin(lo, sh[q]); hi = fmax(hi, sh[MS_BLOCK / 64 + q]); }
      const double span = hi - lo + 10.0 * fmax(a.T, 1e-6);
      br[0] = lo - span;
      br[1] = hi + span;
      br[2] = 0.0;
    }
    __syncthreads();
    return;
  }
  const double* st = a.state + 4 * ((pass - 1) & 1);
  const bool done = st[MS_DONE] != 0.0;   // uniform
  double tot[MS_NODES];
#pragma unroll
  for (int j = 0; j < MS_NODES; ++j) tot[j] = 0.0;
  if (!done) {
    const double* part = a.part + (size_t)((pass - 1) & 1) * MS_MAX_BLOCKS * MS_NODES;
    for (int b = threadIdx.x; b < a.nblk; b += MS_BLOCK)
#pragma unroll
      for (int j = 0; j < MS_NODES; ++j) tot[j] += part[(size_t)b * MS_NODES + j];
    block_sum<MS_NODES>(tot, sh);
  }
  if (threadIdx.x == 0) {
    double lo = st[MS_LO], hi = st[MS_HI], col = done ? 1.0 : 0.0;
    int node = 0;
    for (int lev = 0; lev < prev_levels && col == 0.0; ++lev) {
      const double mid = 0.5 * (lo + hi);
      if (mid == lo || mid == hi) {   // collapsed: later levels change nothing
        col = 1.0;
        break;
      }
      if (tot[node] > a.target) {
        hi = mid;
        node = 2 * node + 1;
      } else {
        lo = mid;
        node = 2 * node + 2;
      }
    }
    br[0] = lo;
    br[1] = hi;
    br[2] = col;
  }
  __syncthreads();
}

// One pass: the bracket, then the counts at the NODES = 2^L - 1 nodes of the
// next L levels (breadth-first: node i has children 2i+1 = [lo_i, mid_i] and
// 2i+2 = [mid_i, hi_i]) into part[pass & 1][block][MS_NODES].
template <int NODES>
__global__ __launch_bounds__(MS_BLOCK) void ms_count_pass(MsArgs a, int pass, int prev_levels) {
  __shared__ double sh[MS_NODES * (MS_BLOCK / 64)];
  __shared__ double br[3];
  ms_bracket(a, pass, prev_levels, sh, br);
  const double lo = br[0], hi = br[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double* st = a.state + 4 * (pass & 1);
    st[MS_LO] = lo;
    st[MS_HI] = hi;
    st[MS_DONE] = br[2];
  }
  if (br[2] != 0.0) return;   // collapsed in an earlier pass (uniform)
  double plo[NODES], phi[NODES], mid[NODES];
  plo[0] = lo;
  phi[0] = hi;
#pragma unroll
  for (int i = 0; i < NODES; ++i) {
    mid[i] = 0.5 * (plo[i] + phi[i]);
    if (2 * i + 2 < NODES) {
      plo[2 * i + 1] = plo[i];
      phi[2 * i + 1] = mid[i];
      plo[2 * i + 2] = mid[i];
      phi[2 * i + 2] = phi[i];
    }
  }
  const double mu_raw = a.shift ? a.state[MS_MURAW] : 0.0;
  const double Td = a.T + 1e-12;
  const double win = MS_SAT * Td;
  double acc[MS_NODES], below = 0.0;
#pragma unroll
  for (int j = 0; j < MS_NODES; ++j) acc[j] = 0.0;
  for (size_t e = (size_t)blockIdx.x * MS_BLOCK + threadIdx.x; e < a.n; e += (size_t)gridDim.x * MS_BLOCK) {
    const double v = ms_value(a, e, mu_raw);
    const double wk = a.w[e / a.nb];
    if (v - lo <= -win) {
      below += wk;                // occupied at every node
    } else if (v - hi < win) {
#pragma unroll
      for (int j = 0; j < NODES; ++j) acc[j] += wk * sigm(-(v - mid[j]) / Td);   // fermidirac(eps - mid, T)
    }
  }
#pragma unroll
  for (int j = 0; j < MS_NODES; ++j) acc[j] += below;
  __syncthreads();   // sh is reused
  block_sum<MS_NODES>(acc, sh);
  if (threadIdx.x == 0) {
    double* part = a.part + (size_t)(pass & 1) * MS_MAX_BLOCKS * MS_NODES + (size_t)blockIdx.x * MS_NODES;
#pragma unroll
    for (int j = 0; j < MS_NODES; ++j) part[j] = acc[j];
  }
}

// walk the last pass (`passes` passes done, the last of `levels` levels):
// mu = the midpoint of the final bracket; store 2 also keeps it as mu_raw
__global__ __launch_bounds__(MS_BLOCK) void ms_finish(MsArgs a, int passes, int levels, int store) {
  __shared__ double sh[MS_NODES * (MS_BLOCK / 64)];
  __shared__ double br[3];
  ms_bracket(a, passes, levels, sh, br);
  if (threadIdx.x == 0) {
    const double mu = 0.5 * (br[0] + br[1]);
    a.state[MS_MU] = mu;
    if (store == 2) a.state[MS_MURAW] = mu;
  }
}

__global__ __launch_bounds__(MS_BLOCK) void ms_out(MsArgs a, double* st, double* mu_out, double* p_out) {
  const double mu = a.state[MS_MU];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (st) st[HFX_ST_MU] = mu;
    if (mu_out) *mu_out = mu;
  }
  if (!p_out) return;
  const double mu_raw = a.shift ? a.state[MS_MURAW] : 0.0;
  for (size_t e = (size_t)blockIdx.x * MS_BLOCK + threadIdx.x; e < a.n; e += (size_t)gridDim.x * MS_BLOCK)
    p_out[e] = sigm(-(ms_value(a, e, mu_raw) - mu) / (a.T + 1e-12));   // fermidirac
}

// Bisection mu (and occupations) for nk x nb eigenvalues; scratch >= MS_END doubles.
int ms_bisection_launch(const double* eps, const double* w, size_t nk, int nb, double target, double T,
                        int maxiter, double level_shift, double* st, double* mu_out, double* p_out, double* scratch,
                        hipStream_t s) {
  const size_t n = nk * (size_t)nb;
  size_t g = (n + MS_BLOCK - 1) / MS_BLOCK;
  if (g > (size_t)MS_MAX_BLOCKS) g = MS_MAX_BLOCKS;
  if (g < 1) g = 1;
  MsArgs a{eps, w, n, nb, (int)g, T, target, level_shift, 0,
           scratch + MS_OFF, scratch + MS_MM_OFF, scratch + MS_PART_OFF};
  const int phases = level_shift != 0.0 ? 2 : 1;
  for (int ph = 0; ph < phases; ++ph) {
    a.shift = ph;
    hipLaunchKernelGGL(ms_minmax_part, dim3(g), dim3(MS_BLOCK), 0, s, a);
    if (int rc = check_launch("ms_minmax_part")) return rc;
    int left = maxiter > 0 ? maxiter : 0, pass = 0, prev = 0;
    while (left > 0) {
      const int L = left < MS_LEVELS ? left : MS_LEVELS;
      if (L == 3) hipLaunchKernelGGL(ms_count_pass<7>, dim3(g), dim3(MS_BLOCK), 0, s, a, pass, prev);
      if (L == 2) hipLaunchKernelGGL(ms_count_pass<3>, dim3(g), dim3(MS_BLOCK), 0, s, a, pass, prev);
      if (L == 1) hipLaunchKernelGGL(ms_count_pass<1>, dim3(g), dim3(MS_BLOCK), 0, s, a, pass, prev);
      if (int rc = check_launch("ms_count_pass")) return rc;
      left -= L;
      prev = L;
      ++pass;
    }
    hipLaunchKernelGGL(ms_finish, dim3(1), dim3(MS_BLOCK), 0, s, a, pass, prev, ph == 0 && phases == 2 ? 2 : 1);
    if (int rc = check_launch("ms_finish")) return rc;
  }
  if (!st && !mu_out && !p_out) return HFX_OK;
  hipLaunchKernelGGL(ms_out, dim3(g), dim3(MS_BLOCK), 0, s, a, st, mu_out, p_out);
  return check_launch("ms_out");
}

static int g_mu_blocks = 0;

// Co-resident grid for the cooperative launch, sized to ~8 eigenvalues per
// thread: fewer blocks make each grid barrier cheaper, and the per-iteration
// work (one exp per eigenvalue) is small.
int mu_blocks(size_t n) {
  if (!g_mu_blocks) {
    int dev = 0, ncu = 0, occ = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, mu_kernel, MU_BLOCK, 0);
    int b = ncu * (occ > 0 ? (occ < 2 ? occ : 2) : 1);
    if (b > 512) b = 512;
    g_mu_blocks = b < 1 ? 1 : b;
  }
  size_t want = (n + (size_t)MU_BLOCK * 8 - 1) / ((size_t)MU_BLOCK * 8);
  if (want < 1) want = 1;
  return want < (size_t)g_mu_blocks ? (int)want : g_mu_blocks;
}

}  // namespace hfx

using namespace hfx;

extern "C" {

size_t hfx_solver_scratch_doubles(void) {
  return MS_END;
}

int hfx_reduce(const double* part, int nval, size_t nk, const double* w, unsigned wmask, double* out, double* st,
               int op, const double* params, double* scratch, void* stream) {
  if (!part || !scratch || nval < 1 || nval > RED_MAX_VALS) return set_error(HFX_EINVAL, "hfx_reduce: bad argument");
  if (op != HFX_OP_SUM && op != HFX_OP_PARTIALS && !st)
    return set_error(HFX_EINVAL, "hfx_reduce: op needs the state vector");
  hipStream_t s = (hipStream_t)stream;
  const int g = reduce_blocks(nk);
  PostArgs pa{};
  pa.op = op;
  if (params) {
    pa.T = params[0];
    pa.denom_scale = params[1];
    pa.count = params[2];
    pa.max_step = params[3];
    pa.shrink = params[4];
    pa.tau_fallback = params[5];
    pa.restart = (int)params[6];
    pa.last = (int)params[7];
    pa.omega0 = (int)params[8];
    pa.restart_every = (int)params[9];
  }
  if (nk <= RED1_MAX_NK && op != HFX_OP_PARTIALS) {
    hipLaunchKernelGGL(reduce_single, dim3(1), dim3(RED1_BLOCK), 0, s, part, nval, nk, w, wmask, out, st, pa);
    return check_launch("reduce_single");
  }
  hipLaunchKernelGGL(reduce_stage1, dim3(g), dim3(RED_BLOCK), 0, s, part, nval, nk, w, wmask, scratch, st, op);
  if (int rc = check_launch("reduce_stage1")) return rc;
  if (op == HFX_OP_PARTIALS) return HFX_OK;
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(RED_BLOCK), 0, s, scratch, nval, g, out, st, pa);
  return check_launch("reduce_stage2");
}

int hfx_pair_dots(int dtype, int nb, size_t nk, int np, const void* const* X, const void* const* Y,
                  const double* const* vx, const double* const* vy, const double* w, unsigned wmask, double* out,
                  double* st, int op, double* scratch, void* stream) {
  if (nb < 1 || np < 1 || np > PD_MAX || !X || !Y || !vx || !vy || !scratch || (!out && op == HFX_OP_SUM) ||
      (op != HFX_OP_SUM && op != HFX_OP_NT0 && op != HFX_OP_NT1 && op != HFX_OP_NT2) || (op != HFX_OP_SUM && !st) ||
      nk * (size_t)nb * nb >= 0xFFFFFFFFull || (dtype != HFX_DTYPE_C128 && dtype != HFX_DTYPE_C64))
    return set_error(HFX_EINVAL, "hfx_pair_dots: bad argument");
  PairDots d{};
  for (int v = 0; v < np; ++v) {
    if (!X[v] || !Y[v] || !vx[v] || !vy[v]) return set_error(HFX_EINVAL, "hfx_pair_dots: pair %d is NULL", v);
    d.x[v] = X[v];
    d.y[v] = Y[v];
    d.vx[v] = vx[v];
    d.vy[v] = vy[v];
  }
  if (nk == 0) return HFX_OK;
  const size_t nm = nk * (size_t)nb * nb;
  size_t g = (nm + (size_t)RED_BLOCK * 4 - 1) / ((size_t)RED_BLOCK * 4);
  if (g > (size_t)RED_MAX_BLOCKS) g = RED_MAX_BLOCKS;
  if (g < 1) g = 1;
  hipStream_t s = (hipStream_t)stream;
  if (dtype == HFX_DTYPE_C128)
    hipLaunchKernelGGL(pair_dots_kernel<double>, dim3((unsigned)g), dim3(RED_BLOCK), 0, s, d, np, nk, nb, w, wmask,
                       scratch);
  else
    hipLaunchKernelGGL(pair_dots_kernel<float>, dim3((unsigned)g), dim3(RED_BLOCK), 0, s, d, np, nk, nb, w, wmask,
                       scratch);
  if (int rc = check_launch("pair_dots_kernel")) return rc;
  PostArgs pa{};
  pa.op = op;
  hipLaunchKernelGGL(reduce_stage2, dim3(1), dim3(RED_BLOCK), 0, s, scratch, 2 * np, (int)g, out, st, pa);
  return check_launch("reduce_stage2");
}

int hfx_solve_mu_steps(const double* eps, const double* w, size_t nk, int nb, double target, double T,
                       double mu0, int mu0_from_state, int maxiter, double* st, double* p_out, double* scratch,
                       void* stream) {
  if (!eps || !w || !scratch || !st || nb < 1 || nk == 0 || maxiter < 0 || nk * (size_t)nb >= 0xFFFFFFFFull)
    return set_error(HFX_EINVAL, "hfx_solve_mu_steps: bad argument");
  const size_t n = nk * (size_t)nb;
  size_t nbk = (n + (size_t)MUS_BLOCK * 4 - 1) / ((size_t)MUS_BLOCK * 4);
  if (nbk > (size_t)MU_CHAIN_MAX_BLK) nbk = MU_CHAIN_MAX_BLK;
  if (nbk < 1) nbk = 1;
  MuStepArgs a{eps, w, nk, nb, target, T, mu0, mu0_from_state, st, scratch + MU_PART_OFF, p_out};
  hipStream_t s = (hipStream_t)stream;
  // bracket pass, maxiter chained steps, then the finish (with the occupations)
  hipLaunchKernelGGL(mu_bracket_part, dim3((unsigned)nbk), dim3(MUS_BLOCK), 0, s, a);
  if (int rc = check_launch("mu_bracket_part")) return rc;
  for (int it = 1; it <= maxiter; ++it) {
    hipLaunchKernelGGL(mu_chain_step, dim3((unsigned)nbk), dim3(MUS_BLOCK), 0, s, a, (int)nbk, it);
    if (int rc = check_launch("mu_chain_step")) return rc;
  }
  hipLaunchKernelGGL(mu_chain_fin, dim3(p_out ? (unsigned)nbk : 1u), dim3(MUS_BLOCK), 0, s, a, (int)nbk,
                     maxiter + 1);
  return check_launch("mu_chain_fin");
}

int hfx_diis_coefficients(const double* brow, double* B, int nh, int idx, int count, double* coef, void* stream) {
  return hfx_diis_coefficients_batched(brow, 0, B, 1, nh, idx, count, coef, stream);
}

int hfx_diis_coefficients_batched(const double* brow, size_t brow_stride, double* B, int nsolve, int nh, int idx,
                                  int count, double* coef, void* stream) {
  if (!brow || !B || !coef || nsolve < 1 || nh < 1 || nh > DIIS_MAX || idx < 0 || idx >= nh || count < 1 ||
      count > nh)
    return set_error(HFX_EINVAL, "hfx_diis_coefficients: bad argument (nh=%d idx=%d count=%d)", nh, idx, count);
  hipLaunchKernelGGL(diis_coef_kernel, dim3((unsigned)nsolve), dim3(64), 0, (hipStream_t)stream, brow, brow_stride, B,
                     nsolve, nh, idx, count, coef, (const double*)nullptr, 0, 0, (double*)nullptr);
  return check_launch("diis_coef_kernel");
}

int hfx_diis_coefficients_reduce(const double* scratch, size_t nk, int nval, double* out, double* B, int nh, int idx,
                                 int count, double* coef, void* stream) {
  if (!scratch || !out || !B || !coef || nk == 0 || nh < 1 || nh > DIIS_MAX || nval != nh + 2 || nval > 64 ||
      idx < 0 || idx >= nh || count < 1 || count > nh)
    return set_error(HFX_EINVAL, "hfx_diis_coefficients_reduce: bad argument (nval=%d nh=%d idx=%d count=%d)", nval,
                     nh, idx, count);
  hipLaunchKernelGGL(diis_coef_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, (const double*)nullptr, (size_t)0, B,
                     1, nh, idx, count, coef, scratch, reduce_blocks(nk), nval, out);
  return check_launch("diis_coef_kernel");
}

size_t hfx_batch_scratch_doubles(int nseg) {
  if (nseg < 1) nseg = 1;
  return (size_t)nseg * (RED_MAX_VALS * SEG_BLOCKS + MU_SEG_VALS * SEG_BLOCKS + 1);
}

int hfx_reduce_batched(const double* part, int nval, size_t nk_total, int seg_len, const double* w, unsigned wmask,
                       double* st, int op, const double* params, int out_slot, double* scratch,
                       size_t scratch_doubles, void* stream) {
  if (!part || !st || !scratch || nval < 1 || nval > RED_MAX_VALS || seg_len < 1 || nk_total % (size_t)seg_len)
    return set_error(HFX_EINVAL, "hfx_reduce_batched: bad argument");
  if (nk_total / seg_len > 0x7FFFFFFFull)
    return set_error(HFX_EINVAL, "hfx_reduce_batched: too many solves");
  if (scratch_doubles < hfx_batch_scratch_doubles((int)(nk_total / seg_len)))
    return set_error(HFX_EINVAL, "hfx_reduce_batched: scratch of %zu doubles, need %zu", scratch_doubles,
                     hfx_batch_scratch_doubles((int)(nk_total / seg_len)));
  if (out_slot >= 0 && (op != HFX_OP_SUM || out_slot + nval > HFX_ST_LEN))
    return set_error(HFX_EINVAL, "hfx_reduce_batched: out_slot only with HFX_OP_SUM, inside the state vector");
  const int nseg = (int)(nk_total / seg_len);
  PostArgs pa{};
  pa.op = op;
  if (params) {
    pa.T = params[0];
    pa.denom_scale = params[1];
    pa.count = params[2];
    pa.max_step = params[3];
    pa.shrink = params[4];
    pa.tau_fallback = params[5];
    pa.restart = (int)params[6];
    pa.last = (int)params[7];
    pa.omega0 = (int)params[8];
    pa.restart_every = (int)params[9];
  }
  int g = (seg_len + RED_BLOCK - 1) / RED_BLOCK;
  if (g > SEG_BLOCKS) g = SEG_BLOCKS;
  hipLaunchKernelGGL(reduce_seg_kernel, dim3((unsigned)g, (unsigned)nseg), dim3(RED_BLOCK), 0, (hipStream_t)stream,
                     part, nval, nk_total, seg_len, w, wmask, st, pa, out_slot, seg_scratch(scratch, nseg));
  return check_launch("reduce_seg_kernel");
}

int hfx_solve_mu_batched(const double* eps, const double* w, const double* targets, size_t nk_total, int seg_len,
                         int nb, double T, int maxiter, double* st, double* p_out, double* scratch,
                         size_t scratch_doubles, void* stream) {
  if (!eps || !w || !targets || !st || !scratch || nb < 1 || seg_len < 1 || nk_total % (size_t)seg_len ||
      maxiter < 0 || nk_total / seg_len > 0x7FFFFFFFull)
    return set_error(HFX_EINVAL, "hfx_solve_mu_batched: bad argument");
  if (scratch_doubles < hfx_batch_scratch_doubles((int)(nk_total / seg_len)))
    return set_error(HFX_EINVAL, "hfx_solve_mu_batched: scratch of %zu doubles, need %zu", scratch_doubles,
                     hfx_batch_scratch_doubles((int)(nk_total / seg_len)));
  const int nseg = (int)(nk_total / seg_len);
  const size_t n = (size_t)seg_len * nb;
  size_t g = (n + (size_t)MUS_BLOCK * 4 - 1) / ((size_t)MUS_BLOCK * 4);
  if (g > SEG_BLOCKS) g = SEG_BLOCKS;
  if (g < 1) g = 1;
  MuSegArgs a{eps, w, targets, seg_len, nb, T, st, p_out, seg_scratch(scratch, nseg)};
  const dim3 grid((unsigned)g, (unsigned)nseg);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(mu_bracket_seg_kernel, grid, dim3(MUS_BLOCK), 0, s, a);
  if (int rc = check_launch("mu_bracket_seg_kernel")) return rc;
  for (int it = 0; it < maxiter; ++it) {
    hipLaunchKernelGGL(mu_step_seg_kernel, grid, dim3(MUS_BLOCK), 0, s, a);
    if (int rc = check_launch("mu_step_seg_kernel")) return rc;
  }
  if (!p_out) return HFX_OK;
  hipLaunchKernelGGL(mu_occ_seg_kernel, grid, dim3(MUS_BLOCK), 0, s, a);
  return check_launch("mu_occ_seg_kernel");
}

int hfx_solve_mu(const double* eps, const double* w, size_t nk, int nb, double target, double T, double mu0,
                 int mu0_from_state, int maxiter, int mode, double level_shift, double* st, double* mu_out,
                 double* p_out, double* scratch, void* stream) {
  if (!eps || !w || !scratch || nb < 1 || nk == 0 || nk * (size_t)nb >= 0xFFFFFFFFull)
    return set_error(HFX_EINVAL, "hfx_solve_mu: bad argument");
  if (mode != HFX_MU_NEWTON_BRACKET && mode != HFX_MU_BISECTION && mode != HFX_MU_NEWTON)
    return set_error(HFX_EINVAL, "hfx_solve_mu: unknown mode %d", mode);
  if (mu0_from_state && !st) return set_error(HFX_EINVAL, "hfx_solve_mu: mu0_from_state needs st");
  if (maxiter < 0) return set_error(HFX_EINVAL, "hfx_solve_mu: maxiter=%d", maxiter);
  if (mode == HFX_MU_BISECTION)
    return ms_bisection_launch(eps, w, nk, nb, target, T, maxiter, level_shift, st, mu_out, p_out, scratch,
                               (hipStream_t)stream);
  MuArgs a{eps, w, nk, nb, target, T, mu0, level_shift, maxiter, mode, mu0_from_state, st, mu_out, p_out, scratch};
  const int blocks = mu_blocks(nk * (size_t)nb);
  void* args[] = {&a};
  hipError_t e = hipLaunchCooperativeKernel((const void*)mu_kernel, dim3(blocks), dim3(MU_BLOCK), args, 0,
                                            (hipStream_t)stream);
  if (e != hipSuccess) return set_error(HFX_EHIP, "mu_kernel cooperative launch: %s", hipGetErrorString(e));
  return check_launch("mu_kernel");
}

}  // extern "C"
