// Internal (non-exported) declarations shared by the libhfx translation units.
#pragma once

#include "hfx_common.h"
#include "../../include/hfx.h"

#define HFX_DTYPE_C128 0
#define HFX_DTYPE_C64 1
#define HFX_MAX_STAGES 16

namespace hfx {

// Per-precision FFT geometry passed by value to the exchange kernel.
template <typename T>
struct FftGeom {
  int n1, n2, nb;
  int nst1, nst2;
  int rad1[HFX_MAX_STAGES];
  int rad2[HFX_MAX_STAGES];
  const cx<T>* tw1;  // W_{n1}^m, m < n1 (device)
  const cx<T>* tw2;  // W_{n2}^m, m < n2 (device)
};

// Scalar steps of the Steihaug truncated CG of the trust-region Newton
// subproblem (solver.py:947-991), one thread (a reduction's post-op).  sums:
// the reduced inner products of np pairs (matrix parts at [0, np), w-weighted
// occupation parts at [np, 2 np)); sc: the loop's device scalars (HFX_NT_*).
// phase 0 (after d^T H d, d^T d, v^T d, v^T v): alpha and the boundary step t
// of |v + t d| = Delta; phase 1 (after |v + alpha d|^2, |r - alpha H d|^2,
// (r - alpha H d)^T z): the stop test and the update coefficients, identity
// once the loop has stopped; phase 2: start (rz = r^T z).
__device__ __forceinline__ void newton_tr_phase(int phase, const double* sums, int np, double* sc) {
  const double tiny = 1e-30;
  auto jip = [&](int i) { return sums[i] + sums[np + i]; };
  if (phase == 2) {
    sc[HFX_NT_RZ] = jip(0);
    sc[HFX_NT_DONE] = 0.0;
    sc[HFX_NT_STEPS] = 0.0;
    return;
  }
  const bool done = sc[HFX_NT_DONE] != 0.0;
  if (phase == 0) {
    if (done) {
      sc[HFX_NT_ALPHA] = 0.0;
      sc[HFX_NT_NALPHA] = 0.0;
      return;
    }
    const double dHd = jip(0), dd = jip(1), vd = jip(2), vv = jip(3), Delta = sc[HFX_NT_DELTA];
    const bool neg = dHd <= 1e-14 * dd;
    const double b = 2.0 * vd, cc = vv - Delta * Delta;
    const double disc = fmax(b * b - 4.0 * dd * cc, 0.0);
    const double alpha = sc[HFX_NT_RZ] / (neg ? 1.0 : dHd);
    sc[HFX_NT_TBD] = (-b + sqrt(disc)) / (2.0 * dd + 1e-30);
    sc[HFX_NT_NEG] = neg ? 1.0 : 0.0;
    sc[HFX_NT_ALPHA] = alpha;
    sc[HFX_NT_NALPHA] = -alpha;
    return;
  }
  if (done) {
    sc[HFX_NT_CV] = 0.0;
    sc[HFX_NT_CR] = 0.0;
    sc[HFX_NT_CDA] = 0.0;
    sc[HFX_NT_CDB] = 1.0;
    return;
  }
  const double na2 = jip(0), rn2 = jip(1), rz_new = jip(2), rz = sc[HFX_NT_RZ], alpha = sc[HFX_NT_ALPHA];
  const bool take_bd = sc[HFX_NT_NEG] != 0.0 || sqrt(fmax(0.0, na2)) >= sc[HFX_NT_DELTA];
  const bool inner = sqrt(fmax(0.0, rn2)) < 0.05 * sc[HFX_NT_GRAD];
  const double beta = rz_new / (fabs(rz) > tiny ? rz : 1.0);
  const bool stop = take_bd || inner;
  sc[HFX_NT_CV] = take_bd ? sc[HFX_NT_TBD] : alpha;   // v += c d, H v += c H d
  sc[HFX_NT_CR] = take_bd ? 0.0 : -alpha;             // r += c H d
  if (stop) {
    sc[HFX_NT_CDA] = 0.0;
    sc[HFX_NT_CDB] = 1.0;
    sc[HFX_NT_DONE] = 1.0;
  } else {
    sc[HFX_NT_CDA] = 1.0;                              // d = z + beta d
    sc[HFX_NT_CDB] = beta;
    sc[HFX_NT_RZ] = rz_new;
  }
  sc[HFX_NT_STEPS] += 1.0;
}

}  // namespace hfx

// Device twiddle tables owned by the plan; only exch_variant changes after
// creation (hfx_plan_set_exchange_variant / hfx_exchange_autotune).
struct hfx_plan {
  int n1, n2, nb, dtype;
  int exch_variant;  // HFX_EXCH_*
  void* tw_dev;  // one allocation: [tw1 | tw2 | channel lists]
  const int* chans_full;  // c = 0 .. nb*nb-1
  const int* chans_triu;  // i * nb + j for i <= j
  hfx::FftGeom<double> g64;
  hfx::FftGeom<float> g32;
};

namespace hfx {
int exchange_launch(const hfx_plan* plan, const void* P, const void* refP, const void* VR, int vfull,
                    const int* chans, int nchan, int herm, const double* w2d, void* Sigma, void* rho,
                    int mode, int shift, int conjv, hipStream_t s, int nsolve = 1, int planar = 0,
                    const int* vmap = nullptr);
int exchange_max_plane(int dtype);
int epilogue_seg_launch(const hfx_plan* p, int nsolve, const void* h, const void* Sig, const void* P, const double* w2d,
                        void* F, double* st, double* part, hipStream_t s);
int eigh16_launch(int dtype, int nb, size_t nk, const void* A, double* lam, void* V, int herm_in, double rel_tol,
                  hipStream_t s, const void* U0 = nullptr);
int exchange_lds_elems(int n1, int n2);
int pack_planar_launch(const hfx_plan* p, int nsolve, const void* X, const void* refX, const int* chans, int nchan,
                       void* Xt, hipStream_t s);
int epilogue_planar_launch(const hfx_plan* p, int nsolve, int max_blocks, const void* h, const void* St,
                           const int* chans, int nchan, int herm, int cover, const double* Hdiag, const void* sigc,
                           const void* P, const double* w2d, size_t w_stride, void* F, void* Sig_out,
                           double* part, double* E_out, double e_offset, double* st, hipStream_t s);
int planar_epi_max_blocks();
}  // namespace hfx
