// Local H.psi without any device allocation: the stick workspace that amt_hpsi_local takes
// from its caller lives here in a fixed, file-scope device array instead, so callers that
// cannot allocate (code under stream or graph capture, tight serving loops, harnesses that
// forbid hipMalloc) still get the same kernels and bit-identical results.
//
// The array is per GPU: each device that loads the library holds its own copy until the
// process ends. Host code must never pass the array's name to a kernel (on the host it is
// only a stand-in symbol); its device address is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch_ptr. Every call on one GPU shares that GPU's
// array: calls that overlap on one device (different streams, or different host threads)
// race on it. Calls on one stream are ordered and safe.
#include <mutex>

#include "../../include/amaterasu_amd.h"
#include "kernels.hpp"

namespace amt {

static_assert(AMT_NOALLOC_SCRATCH_BYTES % 256 == 0, "the scratch size is a multiple of its alignment");

// Stick workspace of the allocation-free entry point: [rows][n1][smax] complex per launch.
static __device__ __align__(256) unsigned char scratch[AMT_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The device address of this device's scratch array (looked up on first use per device,
// with that device current). Null, with the error set, on failure.
static void* scratch_on_current_device() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error(std::string("hpsi_local_noalloc: hipGetDevice: ") + hipGetErrorString(e));
    return nullptr;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("hpsi_local_noalloc: device ordinal " + std::to_string(dev) + " beyond the supported 64");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(std::string("hpsi_local_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

// Band rows whose stick workspace fits the scratch array in one launch (0: not even one).
long long noalloc_rows_per_launch(const Plan& p, int prec) {
  const long long row_bytes = (long long)p.n1 * p.smax * (prec == kFP64 ? 16 : 8);
  if (row_bytes <= 0) return 0;
  const long long rows = (long long)AMT_NOALLOC_SCRATCH_BYTES / row_bytes;
  return rows < 65535 ? rows : 65535;
}

// amt_hpsi_local over the scratch array: the [nkk][nb] rows go in launches of whole k-points
// when a k-point's bands fit, else in band slices of one k-point at a time (rows are
// independent, so the split does not change a single bit of the result).
int hpsi_local_noalloc(const Plan& p, int prec, const void* psi, void* hpsi, int nb, int k0, int nkk,
                       const void* veff, hipStream_t s) {
  const long long rows_max = noalloc_rows_per_launch(p, prec);
  if (rows_max < 1) {
    const long long row_bytes = (long long)p.n1 * p.smax * (prec == kFP64 ? 16 : 8);
    set_error("hpsi_local_noalloc: one band row needs " + std::to_string(row_bytes) +
              " bytes of stick workspace, more than the fixed scratch of " +
              std::to_string((long long)AMT_NOALLOC_SCRATCH_BYTES) + " bytes; use amt_hpsi_local with a workspace");
    return 1;
  }
  unsigned char* ws = static_cast<unsigned char*>(scratch_on_current_device());
  if (!ws) return 1;
  const size_t elem = prec == kFP64 ? 16 : 8;
  const size_t row_stride = (size_t)p.npw_pad * elem;
  const unsigned char* in = static_cast<const unsigned char*>(psi);
  unsigned char* out = static_cast<unsigned char*>(hpsi);
  if (nb <= rows_max) {
    const int kper = (int)(rows_max / nb);
    for (int kk = 0; kk < nkk; kk += kper) {
      const int nk_here = nkk - kk < kper ? nkk - kk : kper;
      const size_t off = (size_t)kk * nb * row_stride;
      if (launch_hpsi_local(p, prec, in + off, out + off, nb, k0 + kk, nk_here, veff, ws, s)) return 1;
    }
    return 0;
  }
  for (int kk = 0; kk < nkk; ++kk) {
    for (int b0 = 0; b0 < nb; b0 += (int)rows_max) {
      const int nbc = nb - b0 < rows_max ? nb - b0 : (int)rows_max;
      const size_t off = ((size_t)kk * nb + b0) * row_stride;
      if (launch_hpsi_local(p, prec, in + off, out + off, nbc, k0 + kk, 1, veff, ws, s)) return 1;
    }
  }
  return 0;
}

}  // namespace amt
