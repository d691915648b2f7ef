// C-ABI entry points of libhfx: plans, argument validation, error reporting.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../include/hfx.h"
#include "hfx_common.h"
#include "hfx_internal.h"

namespace hfx {

static thread_local char g_err[1024] = "";

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(HFX_EHIP, "%s launch failed: %s", what, hipGetErrorString(e));
  return HFX_OK;
}

int kterms_launch(const hfx_plan* p, const void* rho, const double* HH, int ntc, const double* g,
                  const void* Oi, const void* Oj, double* Hdiag, void* sigc, hipStream_t s);
int moments_launch(const hfx_plan* p, const void* P, const void* refP, const double* w2d, void* rho,
                   hipStream_t s);
int epilogue_launch(const hfx_plan* p, const void* h, const void* Sig, const double* Hdiag, const void* sigc,
                    const void* P, const double* w2d, void* F, void* Sig_out, double* partials, double* E_out,
                    double e_offset, hipStream_t s);
int hf_energy_launch(const hfx_plan* p, const void* P, const void* h, const void* Sig, const void* H,
                     int H_bcast, const double* w2d, double* partials, double* E_out, hipStream_t s);
int energy_partials_max();

// Radix plan for a length-N axis: 8s first, then 4, 2, 3, 5, 7, then any
// larger prime factor as a stage of its own (Stockham stages may come in any
// order; the exchange kernels run a prime above 7 as an O(N p) generic stage).
static int factorize(int N, int* rad, int* nst) {
  int n = N, k = 0;
  while (n % 8 == 0 && k < HFX_MAX_STAGES) { rad[k++] = 8; n /= 8; }
  const int rs[5] = {4, 2, 3, 5, 7};
  for (int q = 0; q < 5; ++q)
    while (n % rs[q] == 0 && k < HFX_MAX_STAGES) { rad[k++] = rs[q]; n /= rs[q]; }
  for (int p = 11; n > 1 && k < HFX_MAX_STAGES; p += 2) {
    if ((long long)p * p > n) p = n;   // what is left is prime
    while (n % p == 0 && k < HFX_MAX_STAGES) { rad[k++] = p; n /= p; }
  }
  *nst = k;
  return n == 1 ? 0 : -1;
}

// W_N^m = exp(-2 pi i m / N) in extended precision, rounded once.
static void twiddles(int N, std::vector<long double>& re, std::vector<long double>& im) {
  re.resize(N);
  im.resize(N);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < N; ++m) {
    // reduce the angle to the first octant-friendly range m/N in [0, 1)
    const long double t = two_pi * (long double)m / (long double)N;
    re[m] = cosl(t);
    im[m] = -sinl(t);
  }
}

}  // namespace hfx

using namespace hfx;

extern "C" {

int hfx_abi_version(void) { return 24; }

int hfx_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return HFX_EINVAL;
  strncpy(buf, g_err, n - 1);
  buf[n - 1] = '\0';
  return HFX_OK;
}

int hfx_plan_create(int nk1, int nk2, int nb, int dtype, hfx_plan** out) {
  if (!out) return set_error(HFX_EINVAL, "hfx_plan_create: out is NULL");
  *out = nullptr;
  if (nk1 <= 0 || nk2 <= 0 || nb <= 0)
    return set_error(HFX_EINVAL, "hfx_plan_create: bad shape (%d, %d, %d)", nk1, nk2, nb);
  if (dtype != HFX_C128 && dtype != HFX_C64)
    return set_error(HFX_EINVAL, "hfx_plan_create: dtype must be HFX_C128 or HFX_C64");
  if (nb > 64) return set_error(HFX_EUNSUPPORTED, "hfx_plan_create: nb=%d > 64 unsupported", nb);
  hfx_plan* p = new hfx_plan();
  p->n1 = nk1;
  p->n2 = nk2;
  p->nb = nb;
  p->dtype = dtype;
  p->exch_variant = HFX_EXCH_AUTO;
  int rad1[HFX_MAX_STAGES], rad2[HFX_MAX_STAGES], ns1 = 0, ns2 = 0;
  if (factorize(nk1, rad1, &ns1) || factorize(nk2, rad2, &ns2)) {
    delete p;
    return set_error(HFX_EUNSUPPORTED, "hfx_plan_create: mesh (%d, %d) needs more than %d FFT stages", nk1, nk2,
                     HFX_MAX_STAGES);
  }
  // planes beyond LDS run the HBM-staged exchange, whose tiles hold one whole
  // row (odd pitch) and one whole column of the plane
  if (nk1 > exchange_max_plane(dtype) - 1 || nk2 > exchange_max_plane(dtype) - 1) {
    delete p;
    return set_error(HFX_EUNSUPPORTED, "hfx_plan_create: mesh axis %d exceeds %d", nk1 > nk2 ? nk1 : nk2,
                     exchange_max_plane(dtype) - 1);
  }
  std::vector<long double> r1, i1, r2, i2;
  twiddles(nk1, r1, i1);
  twiddles(nk2, r2, i2);
  const int nb2 = nb * nb, ntri = nb * (nb + 1) / 2;
  const size_t esz = dtype == HFX_C128 ? 16 : 8;
  const size_t tw_bytes = (size_t)(nk1 + nk2) * esz;
  const size_t ch_off = (tw_bytes + 255) / 256 * 256;
  const size_t bytes = ch_off + (size_t)(nb2 + ntri) * sizeof(int);
  std::vector<unsigned char> host(bytes, 0);
  if (dtype == HFX_C128) {
    cx<double>* t = (cx<double>*)host.data();
    for (int m = 0; m < nk1; ++m) t[m] = {(double)r1[m], (double)i1[m]};
    for (int m = 0; m < nk2; ++m) t[nk1 + m] = {(double)r2[m], (double)i2[m]};
  } else {
    cx<float>* t = (cx<float>*)host.data();
    for (int m = 0; m < nk1; ++m) t[m] = {(float)r1[m], (float)i1[m]};
    for (int m = 0; m < nk2; ++m) t[nk1 + m] = {(float)r2[m], (float)i2[m]};
  }
  int* ch = (int*)(host.data() + ch_off);
  for (int c = 0; c < nb2; ++c) ch[c] = c;
  int q = nb2;
  for (int i = 0; i < nb; ++i)
    for (int j = i; j < nb; ++j) ch[q++] = i * nb + j;
  hipError_t e = hipMalloc(&p->tw_dev, bytes);
  if (e != hipSuccess) {
    delete p;
    return set_error(HFX_EHIP, "hfx_plan_create: hipMalloc: %s", hipGetErrorString(e));
  }
  e = hipMemcpy(p->tw_dev, host.data(), bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(p->tw_dev);
    delete p;
    return set_error(HFX_EHIP, "hfx_plan_create: hipMemcpy: %s", hipGetErrorString(e));
  }
  unsigned char* base = (unsigned char*)p->tw_dev;
  p->chans_full = (const int*)(base + ch_off);
  p->chans_triu = p->chans_full + nb2;
  for (auto* g : {&p->g64}) {
    g->n1 = nk1; g->n2 = nk2; g->nb = nb; g->nst1 = ns1; g->nst2 = ns2;
    for (int s = 0; s < HFX_MAX_STAGES; ++s) { g->rad1[s] = rad1[s]; g->rad2[s] = rad2[s]; }
    g->tw1 = (const cx<double>*)base;
    g->tw2 = (const cx<double>*)base + nk1;
  }
  {
    auto* g = &p->g32;
    g->n1 = nk1; g->n2 = nk2; g->nb = nb; g->nst1 = ns1; g->nst2 = ns2;
    for (int s = 0; s < HFX_MAX_STAGES; ++s) { g->rad1[s] = rad1[s]; g->rad2[s] = rad2[s]; }
    g->tw1 = (const cx<float>*)base;
    g->tw2 = (const cx<float>*)base + nk1;
  }
  *out = p;
  return HFX_OK;
}

int hfx_plan_destroy(hfx_plan* plan) {
  if (!plan) return HFX_OK;
  if (plan->tw_dev) (void)hipFree(plan->tw_dev);
  delete plan;
  return HFX_OK;
}

size_t hfx_energy_scratch_bytes(const hfx_plan* plan) {
  (void)plan;
  const int n = energy_partials_max() > planar_epi_max_blocks() ? energy_partials_max() : planar_epi_max_blocks();
  return (size_t)n * sizeof(double);
}

#define HFX_REQUIRE(cond, ...) \
  do {                         \
    if (!(cond)) return set_error(HFX_EINVAL, __VA_ARGS__); \
  } while (0)

int hfx_selfenergy(const hfx_plan* plan, const void* P, const void* refP, const void* VR, int vr_full,
                   const int* chans, int nchan, int herm, int ifftshift, const double* w2d, void* Sigma,
                   void* rho, void* stream) {
  HFX_REQUIRE(plan, "hfx_selfenergy: plan is NULL");
  HFX_REQUIRE(P && VR && Sigma, "hfx_selfenergy: P, VR and Sigma are required");
  HFX_REQUIRE(!rho || w2d, "hfx_selfenergy: rho requires w2d");
  const int nb2 = plan->nb * plan->nb;
  if (!chans) {
    chans = herm ? plan->chans_triu : plan->chans_full;
    nchan = herm ? plan->nb * (plan->nb + 1) / 2 : nb2;
  }
  HFX_REQUIRE(nchan >= 0 && nchan <= nb2, "hfx_selfenergy: nchan=%d out of range", nchan);
  return exchange_launch(plan, P, refP, VR, vr_full ? 1 : 0, chans, nchan, herm ? 1 : 0, w2d, Sigma, rho,
                         0, ifftshift ? 1 : 0, 0, (hipStream_t)stream);
}

int hfx_selfenergy_batched(const hfx_plan* plan, int nsolve, const void* P, const void* VR, int vr_full, int herm,
                           void* Sigma, void* stream) {
  HFX_REQUIRE(plan && P && VR && Sigma, "hfx_selfenergy_batched: NULL argument");
  HFX_REQUIRE(nsolve >= 1 && (long long)nsolve * plan->nb * plan->nb <= (1LL << 30),
              "hfx_selfenergy_batched: nsolve=%d out of range", nsolve);
  const int* chans = herm ? plan->chans_triu : plan->chans_full;
  const int nchan = herm ? plan->nb * (plan->nb + 1) / 2 : plan->nb * plan->nb;
  return exchange_launch(plan, P, nullptr, VR, vr_full ? 1 : 0, chans, nchan, herm ? 1 : 0, nullptr, Sigma, nullptr,
                         0, 0, 0, (hipStream_t)stream, nsolve);
}

// Default channel set when chans is NULL: upper triangle (herm) or all nb^2.
// *cover: the set (with the Hermitian mirror) reaches every channel.
static void resolve_chans(const hfx_plan* plan, const int*& chans, int& nchan, int herm, int* cover) {
  if (!chans) {
    chans = herm ? plan->chans_triu : plan->chans_full;
    nchan = herm ? plan->nb * (plan->nb + 1) / 2 : plan->nb * plan->nb;
    if (cover) *cover = 1;
  } else if (cover) {
    *cover = 0;
  }
}

int hfx_planar_pack(const hfx_plan* plan, int nsolve, const void* X, const void* refX, const int* chans, int nchan,
                    int herm, void* Xt, void* stream) {
  HFX_REQUIRE(plan && X && Xt, "hfx_planar_pack: NULL argument");
  HFX_REQUIRE(nsolve >= 1, "hfx_planar_pack: nsolve=%d", nsolve);
  resolve_chans(plan, chans, nchan, herm, nullptr);
  HFX_REQUIRE(nchan >= 1 && nchan <= plan->nb * plan->nb, "hfx_planar_pack: nchan=%d out of range", nchan);
  return pack_planar_launch(plan, nsolve, X, refX, chans, nchan, Xt, (hipStream_t)stream);
}

int hfx_selfenergy_planar(const hfx_plan* plan, int nsolve, const void* Pt, const void* VRc, int vr_full,
                          const int* vr_map, const int* chans, int nchan, int herm, int ifftshift,
                          const double* w2d, void* Sigt, void* rho, void* stream) {
  HFX_REQUIRE(plan && Pt && VRc && Sigt, "hfx_selfenergy_planar: NULL argument");
  HFX_REQUIRE(!rho || (w2d && nsolve == 1), "hfx_selfenergy_planar: rho needs w2d and one solve");
  HFX_REQUIRE(nsolve >= 1 && (long long)nsolve * plan->nb * plan->nb <= (1LL << 30),
              "hfx_selfenergy_planar: nsolve=%d out of range", nsolve);
  resolve_chans(plan, chans, nchan, herm, nullptr);
  HFX_REQUIRE(nchan >= 0 && nchan <= plan->nb * plan->nb, "hfx_selfenergy_planar: nchan=%d out of range", nchan);
  return exchange_launch(plan, Pt, nullptr, VRc, vr_full ? 1 : 0, chans, nchan, herm ? 1 : 0, w2d, Sigt, rho, 0,
                         ifftshift ? 1 : 0, 0, (hipStream_t)stream, nsolve, /*planar=*/1, vr_map);
}

int hfx_selfenergy_adjoint(const hfx_plan* plan, const void* G, const void* VR, int vr_full, void* out,
                           void* stream) {
  HFX_REQUIRE(plan && G && VR && out, "hfx_selfenergy_adjoint: NULL argument");
  return exchange_launch(plan, G, nullptr, VR, vr_full ? 1 : 0, plan->chans_full, plan->nb * plan->nb, 0,
                         nullptr, out, nullptr, 0, 0, 1, (hipStream_t)stream);
}

int hfx_fft2_weighted(const hfx_plan* plan, const void* X, const double* w, void* out, void* stream) {
  HFX_REQUIRE(plan && X && w && out, "hfx_fft2_weighted: NULL argument");
  return exchange_launch(plan, X, nullptr, X, 1, plan->chans_full, plan->nb * plan->nb, 0, w, out, nullptr, 1, 0,
                         0, (hipStream_t)stream);
}

int hfx_plan_set_exchange_variant(hfx_plan* plan, int variant) {
  HFX_REQUIRE(plan, "hfx_plan_set_exchange_variant: NULL plan");
  HFX_REQUIRE(variant >= HFX_EXCH_AUTO && variant <= HFX_EXCH_V3_C,
              "hfx_plan_set_exchange_variant: variant %d out of range", variant);
  plan->exch_variant = variant;
  return HFX_OK;
}

int hfx_plan_get_exchange_variant(const hfx_plan* plan) { return plan ? plan->exch_variant : HFX_EINVAL; }

int hfx_exchange_autotune(hfx_plan* plan, const void* P, const void* VR, int vr_full, const int* vr_map, int planar,
                          void* Sigma, int reps, double* times_ms, int* best, void* stream) {
  HFX_REQUIRE(plan && P && VR && Sigma, "hfx_exchange_autotune: NULL argument");
  HFX_REQUIRE(reps > 0 && reps <= 1000, "hfx_exchange_autotune: reps must be in 1..1000");
  const hipStream_t s = (hipStream_t)stream;
  const int keep = plan->exch_variant;
  hipEvent_t e0, e1;
  HFX_HIP_TRY(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return set_error(HFX_EHIP, "hfx_exchange_autotune: hipEventCreate failed");
  }
  int rc = HFX_OK, arg = HFX_EXCH_V1;
  double tbest = 1e300;
  if (times_ms) times_ms[0] = 0.0;
  for (int v = HFX_EXCH_V1; v <= HFX_EXCH_V3_C && rc == HFX_OK; ++v) {
    plan->exch_variant = v;
    const int nch = plan->nb * plan->nb;
    for (int r = -1; r < reps && rc == HFX_OK; ++r) {  // r = -1: warm-up launch
      if (r == 0 && hipEventRecord(e0, s) != hipSuccess) rc = set_error(HFX_EHIP, "hipEventRecord failed");
      if (rc == HFX_OK)
        rc = exchange_launch(plan, P, nullptr, VR, vr_full, plan->chans_full, nch, 0, nullptr, Sigma, nullptr,
                             0, 0, 0, s, 1, planar ? 1 : 0, planar ? vr_map : nullptr);
    }
    float ms = 0.f;
    if (rc == HFX_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = set_error(HFX_EHIP, "hfx_exchange_autotune: event timing failed");
    const double t = (double)ms / reps;
    if (times_ms) times_ms[v] = t;
    if (rc == HFX_OK && t < tbest) { tbest = t; arg = v; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  plan->exch_variant = rc == HFX_OK ? arg : keep;
  if (best) *best = plan->exch_variant;
  return rc;
}

int hfx_prepare_vr(const hfx_plan* plan, const void* Vq, int vr_full, const double* w2d, int phase,
                   void* VR, void* stream) {
  HFX_REQUIRE(plan && Vq && w2d && VR, "hfx_prepare_vr: NULL argument");
  const int nchan = vr_full ? plan->nb * plan->nb : 1;
  const int* chans = plan->chans_full;  // channel c of a (.., 1, 1) V is channel 0
  if (!vr_full) {
    // a scalar V is a single channel with nb = 1 strides: use a 1-band view
    return set_error(HFX_EUNSUPPORTED, "hfx_prepare_vr: use a plan with nb = 1 for scalar V");
  }
  return exchange_launch(plan, Vq, nullptr, Vq, 1, chans, nchan, 0, w2d, VR, nullptr, 1, phase ? 1 : 0, 0,
                         (hipStream_t)stream);
}

int hfx_density_moments(const hfx_plan* plan, const void* P, const void* refP, const double* w2d,
                        void* rho, void* stream) {
  HFX_REQUIRE(plan && P && w2d && rho, "hfx_density_moments: NULL argument");
  return moments_launch(plan, P, refP, w2d, rho, (hipStream_t)stream);
}

int hfx_kterms(const hfx_plan* plan, const void* rho, const double* HH, int ntc, const double* g,
               const void* Oi, const void* Oj, double* Hdiag, void* sigc, void* stream) {
  HFX_REQUIRE(plan && rho, "hfx_kterms: NULL argument");
  HFX_REQUIRE(!Hdiag || HH, "hfx_kterms: Hdiag requires HH");
  HFX_REQUIRE(!sigc || (ntc >= 0 && (ntc == 0 || (g && Oi && Oj))), "hfx_kterms: contact arrays missing");
  return kterms_launch(plan, rho, HH, sigc ? ntc : 0, g, Oi, Oj, Hdiag, sigc, (hipStream_t)stream);
}

int hfx_fock_epilogue_batched(const hfx_plan* plan, int nsolve, const void* h, const void* Sigma, const void* P,
                              const double* w2d, void* F, double* st, double* scratch, size_t scratch_doubles,
                              void* stream) {
  HFX_REQUIRE(plan && h && Sigma && P && w2d && F && st && scratch, "hfx_fock_epilogue_batched: NULL argument");
  HFX_REQUIRE(nsolve >= 1, "hfx_fock_epilogue_batched: nsolve=%d", nsolve);
  HFX_REQUIRE(scratch_doubles >= (size_t)64 * nsolve,
              "hfx_fock_epilogue_batched: scratch of %zu doubles, need %zu", scratch_doubles, (size_t)64 * nsolve);
  return epilogue_seg_launch(plan, nsolve, h, Sigma, P, w2d, F, st, scratch, (hipStream_t)stream);
}

int hfx_fock_epilogue(const hfx_plan* plan, const void* h, const void* Sigma, const double* Hdiag,
                      const void* sigc, const void* P, const double* w2d, void* F, void* Sigma_out,
                      double* scratch, double* E_out, double e_offset, void* stream) {
  HFX_REQUIRE(plan && h && F, "hfx_fock_epilogue: plan, h and F are required");
  HFX_REQUIRE(!(Sigma_out && sigc && Sigma_out == Sigma), "hfx_fock_epilogue: Sigma_out aliases Sigma");
  HFX_REQUIRE(!E_out || (P && w2d && scratch), "hfx_fock_epilogue: energy needs P, w2d and scratch");
  return epilogue_launch(plan, h, Sigma, Hdiag, sigc, P, w2d, F, Sigma_out, scratch, E_out, e_offset,
                         (hipStream_t)stream);
}

int hfx_fock_epilogue_planar(const hfx_plan* plan, const void* h, const void* Sigt, const int* chans, int nchan,
                             int herm, const double* Hdiag, const void* sigc, const void* P, const double* w2d,
                             void* F, void* Sigma_out, double* scratch, double* E_out, double e_offset,
                             void* stream) {
  HFX_REQUIRE(plan && h && Sigt && F, "hfx_fock_epilogue_planar: h, Sigt and F are required");
  HFX_REQUIRE(!E_out || (P && w2d && scratch), "hfx_fock_epilogue_planar: energy needs P, w2d and scratch");
  int cover = 0;
  resolve_chans(plan, chans, nchan, herm, &cover);
  HFX_REQUIRE(nchan >= 1 && nchan <= plan->nb * plan->nb, "hfx_fock_epilogue_planar: nchan=%d out of range", nchan);
  return epilogue_planar_launch(plan, 1, planar_epi_max_blocks(), h, Sigt, chans, nchan, herm ? 1 : 0, cover, Hdiag,
                                sigc, E_out ? P : nullptr, w2d, 0, F, Sigma_out, E_out ? scratch : nullptr, E_out,
                                e_offset, nullptr, (hipStream_t)stream);
}

int hfx_fock_epilogue_planar_batched(const hfx_plan* plan, int nsolve, const void* h, const void* Sigt, int herm,
                                     const void* P, const double* w2d, void* F, double* st, double* scratch,
                                     size_t scratch_doubles, void* stream) {
  HFX_REQUIRE(plan && h && Sigt && P && w2d && F && st && scratch,
              "hfx_fock_epilogue_planar_batched: NULL argument");
  HFX_REQUIRE(nsolve >= 1, "hfx_fock_epilogue_planar_batched: nsolve=%d", nsolve);
  HFX_REQUIRE(scratch_doubles >= (size_t)256 * nsolve,
              "hfx_fock_epilogue_planar_batched: scratch of %zu doubles, need %zu", scratch_doubles,
              (size_t)256 * nsolve);
  const int* chans = nullptr;
  int nchan = 0, cover = 0;
  resolve_chans(plan, chans, nchan, herm, &cover);
  return epilogue_planar_launch(plan, nsolve, 256, h, Sigt, chans, nchan, herm ? 1 : 0, cover, nullptr, nullptr, P,
                                w2d, (size_t)plan->n1 * plan->n2, F, nullptr, scratch, nullptr, 0.0, st,
                                (hipStream_t)stream);
}

int hfx_hf_energy(const hfx_plan* plan, const void* P, const void* h, const void* Sigma, const void* H,
                  int H_bcast, const double* w2d, double* scratch, double* E_out, void* stream) {
  HFX_REQUIRE(plan && P && h && w2d && scratch && E_out, "hfx_hf_energy: NULL argument");
  return hf_energy_launch(plan, P, h, Sigma, H, H_bcast, w2d, scratch, E_out, (hipStream_t)stream);
}

}  // extern "C"
