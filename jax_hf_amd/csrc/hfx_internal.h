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
                  hipStream_t s);
int exchange_lds_elems(int n1, int n2);
int pack_planar_launch(const hfx_plan* p, int nsolve, const void* X, const void* refX, const int* chans, int nchan,
                       void* Xt, hipStream_t s);
int epilogue_planar_launch(const hfx_plan* p, int nsolve, int max_blocks, const void* h, const void* St,
                           const int* chans, int nchan, int herm, int cover, const double* Hdiag, const void* sigc,
                           const void* P, const double* w2d, size_t w_stride, void* F, void* Sig_out,
                           double* part, double* E_out, double e_offset, double* st, hipStream_t s);
int planar_epi_max_blocks();
}  // namespace hfx
