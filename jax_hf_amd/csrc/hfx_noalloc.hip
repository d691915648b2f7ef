// hfx_solve_mu without a scratch argument, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  The scratch is a library-owned file-scope device
// array; each GPU the library runs on holds its own copy until the process
// ends.  Its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and only that address is
// passed to kernels: the array's name is a host-side stand-in in host code.
//
// The mu kernels write every scratch word before reading it (block partials,
// the multisection bracket, the node counts), so nothing needs clearing
// between calls.  All calls on one GPU share the array: calls that overlap on
// that GPU (several streams, or several host threads without ordering) race.
#include <mutex>

#include "hfx_common.h"
#include "hfx_internal.h"

namespace {

constexpr size_t kScratchDoubles = 32768;   // 256 KiB per GPU
constexpr int kMaxDevices = 64;

std::mutex g_lock;
double* g_scratch[kMaxDevices] = {};

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

// Device address of this GPU's copy of the scratch array (looked up on first use).
static int scratch_ptr(double** out) {
  int dev = 0;
  HFX_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return hfx::set_error(HFX_EUNSUPPORTED, "device index %d out of range", dev);
  std::lock_guard<std::mutex> g(g_lock);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return hfx::set_error(HFX_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_scratch[dev];
  return HFX_OK;
}

extern "C" {

size_t hfx_noalloc_scratch_doubles(void) { return kScratchDoubles; }

int hfx_solve_mu_noalloc(const double* eps, const double* w, size_t nk, int nb, double target, double T,
                         double mu0, int mu0_from_state, int maxiter, int mode, double level_shift, double* st,
                         double* mu_out, double* p_out, void* stream) {
  const size_t need = hfx_solver_scratch_doubles();
  if (need > kScratchDoubles)
    return hfx::set_error(HFX_EUNSUPPORTED, "hfx_solve_mu_noalloc: needs %zu scratch doubles, the library holds %zu",
                          need, kScratchDoubles);
  if (!eps || !w || nb < 1 || nk == 0 || nk * (size_t)nb >= 0xFFFFFFFFull)
    return hfx::set_error(HFX_EINVAL, "hfx_solve_mu_noalloc: bad argument (nk=%zu, nb=%d; nk * nb must be below 2^32)",
                          nk, nb);
  double* s = nullptr;
  if (int rc = scratch_ptr(&s)) return rc;
  return hfx_solve_mu(eps, w, nk, nb, target, T, mu0, mu0_from_state, maxiter, mode, level_shift, st, mu_out, p_out,
                      s, stream);
}

}  // extern "C"
