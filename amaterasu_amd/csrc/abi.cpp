// The C ABI of libamaterasu_amd: plain pointers and sizes, an int status per call and a
// thread-local error string. No exception crosses the ABI. Declarations and the
// reference interfaces each entry point replaces are in include/amaterasu_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/amaterasu_amd.h"
#include "kernels.hpp"

namespace amt {

static thread_local std::string g_error;
void set_error(const std::string& msg) { g_error = msg; }

int launch_residual(int prec, const void* hv, const void* v, const double* eps, void* res, double* norm, int nrow,
                    int npw, hipStream_t s);
int launch_precondition(int prec, const void* res, const void* v, const void* kin, const int* rows, void* out,
                        int nout, int npw, hipStream_t s);
int launch_dij(int prec, const void* P, const void* D, const int* blk_start, const int* blk_len, void* out,
               int nrow, int nproj, long long sP, long long sD, long long sO, int batch, hipStream_t s);

static int hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return 0;
  }
  return 1;
}

template <typename T>
static int upload(T** dst, const std::vector<T>& src, const char* what) {
  *dst = nullptr;
  if (src.empty()) return 1;
  if (!hip_ok(hipMalloc(reinterpret_cast<void**>(dst), src.size() * sizeof(T)), what)) return 0;
  return hip_ok(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), what);
}

static void free_plan_device(Plan& p) {
  void* ptrs[] = {p.d_flatmap,  p.d_stick_jk, p.d_stick_map, p.d_nsticks, p.d_rows,       p.d_nrows,
                  p.d_kin64,    p.d_kin32,    p.d_tw64[0],   p.d_tw64[1],  p.d_tw64[2],    p.d_tw32[0],
                  p.d_tw32[1],  p.d_tw32[2],  p.d_rowmask,   p.d_rowstart, p.d_livemask,   p.d_npw};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  p.d_flatmap = p.d_stick_jk = p.d_stick_map = p.d_nsticks = p.d_rows = p.d_nrows = nullptr;
  p.d_rowmask = p.d_livemask = nullptr;
  p.d_rowstart = p.d_npw = nullptr;
  p.d_kin64 = nullptr;
  p.d_kin32 = nullptr;
  for (int a = 0; a < 3; ++a) {
    p.d_tw64[a] = nullptr;
    p.d_tw32[a] = nullptr;
  }
}

static int max_lds_bytes() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
    return 65536;
  return v;
}

}  // namespace amt

using namespace amt;

struct amt_plan {
  Plan p;
};

extern "C" {

const char* amt_version(void) { return "amaterasu_amd 0.1.0 (gfx950)"; }

const char* amt_last_error(void) { return g_error.c_str(); }

int amt_device_info(char* buf, int len) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (!hip_ok(hipGetDevice(&dev), "hipGetDevice")) return 1;
  if (!hip_ok(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties")) return 1;
  std::string s = std::string(prop.name) + " arch=" + prop.gcnArchName + " cu=" +
                  std::to_string(prop.multiProcessorCount) + " lds_per_block=" +
                  std::to_string(prop.sharedMemPerBlock) + " lds_per_cu=" +
                  std::to_string(prop.maxSharedMemoryPerMultiProcessor) + " clock_khz=" +
                  std::to_string(prop.clockRate) + " mem_clock_khz=" + std::to_string(prop.memoryClockRate) +
                  " bus_bits=" + std::to_string(prop.memoryBusWidth) + " l2=" + std::to_string(prop.l2CacheSize) +
                  " hbm_bytes=" + std::to_string(prop.totalGlobalMem);
  if (buf && len > 0) {
    std::strncpy(buf, s.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return 0;
}

int amt_plan_create(amt_plan_t** out, int n1, int n2, int n3, int nk, const int* npw, int npw_pad,
                    const int* miller, const double* kinetic) {
  if (!out || !npw || !miller || !kinetic || nk <= 0) {
    set_error("plan_create: null argument or nk <= 0");
    return 1;
  }
  *out = nullptr;
  auto* h = new amt_plan;
  Plan& p = h->p;
  p.n1 = n1;
  p.n2 = n2;
  p.n3 = n3;
  p.nk = nk;
  p.npw_pad = npw_pad;
  const int dims[3] = {n1, n2, n3};
  for (int a = 0; a < 3; ++a) {
    if (!make_axis_plan(dims[a], &p.ap[a])) {
      set_error("plan_create: grid length " + std::to_string(dims[a]) + " is not 2,3,5,7,11-smooth (or > 4096)");
      delete h;
      return 1;
    }
  }
  const long long ngrid = (long long)n1 * n2 * n3;
  if (ngrid > (1LL << 31) - 1) {
    set_error("plan_create: grid too large for 32-bit flat indices");
    delete h;
    return 1;
  }
  p.npw.assign(npw, npw + nk);
  for (int k = 0; k < nk; ++k) {
    if (npw[k] <= 0 || npw[k] > npw_pad) {
      set_error("plan_create: npw[" + std::to_string(k) + "] outside (0, npw_pad]");
      delete h;
      return 1;
    }
  }
  // sticks: the distinct (j, k) columns of each sphere
  std::vector<std::vector<int>> stick_of(nk);  // per k, per G: stick id
  std::vector<std::vector<int>> ipos(nk);      // per k, per G: i index
  std::vector<std::vector<int>> sticks(nk);    // per k: j*n3+k of each stick, sorted
  std::vector<int> flat((size_t)nk * npw_pad, -1);
  std::vector<double> kin64((size_t)nk * npw_pad, 0.0);
  long long off = 0;
  int smax = 0;
  for (int k = 0; k < nk; ++k) {
    const int n = npw[k];
    int hmax[3] = {0, 0, 0};
    for (int g = 0; g < n; ++g)
      for (int a = 0; a < 3; ++a) hmax[a] = std::max(hmax[a], std::abs(miller[3 * (off + g) + a]));
    for (int a = 0; a < 3; ++a) {
      if (2 * hmax[a] + 1 > dims[a]) {
        set_error("plan_create: grid is too small for the Miller indices of k-point " + std::to_string(k) +
                  "; high-G components would alias onto low-G ones");
        delete h;
        return 1;
      }
    }
    std::vector<int> colid(n), ii(n);
    std::vector<int> cols;
    cols.reserve(n);
    for (int g = 0; g < n; ++g) {
      const int* m = miller + 3 * (off + g);
      const int i = ((m[0] % n1) + n1) % n1, j = ((m[1] % n2) + n2) % n2, kk = ((m[2] % n3) + n3) % n3;
      ii[g] = i;
      colid[g] = j * n3 + kk;
      flat[(size_t)k * npw_pad + g] = (int)(((long long)i * n2 + j) * n3 + kk);
      kin64[(size_t)k * npw_pad + g] = kinetic[off + g];
      cols.push_back(j * n3 + kk);
    }
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    std::vector<int> sid(n);
    for (int g = 0; g < n; ++g) sid[g] = (int)(std::lower_bound(cols.begin(), cols.end(), colid[g]) - cols.begin());
    stick_of[k] = std::move(sid);
    ipos[k] = std::move(ii);
    sticks[k] = std::move(cols);
    smax = std::max(smax, (int)sticks[k].size());
    off += n;
  }
  p.smax = smax;
  p.nsticks.resize(nk);
  std::vector<int> jk((size_t)nk * smax, -1);
  std::vector<int> smap((size_t)nk * n1 * smax, -1);
  for (int k = 0; k < nk; ++k) {
    p.nsticks[k] = (int)sticks[k].size();
    for (int s = 0; s < p.nsticks[k]; ++s) jk[(size_t)k * smax + s] = sticks[k][s];
    for (int g = 0; g < npw[k]; ++g) {
      int& slot = smap[((size_t)k * n1 + ipos[k][g]) * smax + stick_of[k][g]];
      if (slot != -1) {
        set_error("plan_create: two plane waves map to one grid point");
        delete h;
        return 1;
      }
      slot = g;
    }
  }
  // rows j (plane index along axis 1) that hold at least one stick, per k
  std::vector<int> rows((size_t)nk * n2, 0), nrows(nk, 0);
  for (int k = 0; k < nk; ++k) {
    std::vector<char> used(n2, 0);
    for (int v : sticks[k]) used[v / n3] = 1;
    for (int j = 0; j < n2; ++j)
      if (used[j]) rows[(size_t)k * n2 + nrows[k]++] = j;
  }
  p.fast = (n1 == n2 && n2 == n3 && fast_length_supported(n1));
  // per-row stick masks and first-stick indices for the compile-time plane kernels
  std::vector<unsigned long long> rowmask, livemask;
  std::vector<int> rowstart;
  if (n2 <= 128 && n3 <= 128) {
    rowmask.assign((size_t)nk * n2 * 2, 0ull);
    rowstart.assign((size_t)nk * n2, 0);
    livemask.assign((size_t)nk * 2, 0ull);
    for (int k = 0; k < nk; ++k) {
      std::vector<int> count(n2, 0);
      for (int v : sticks[k]) {
        const int j = v / n3, kk = v % n3;
        rowmask[((size_t)k * n2 + j) * 2 + (kk >> 6)] |= 1ull << (kk & 63);
        livemask[(size_t)k * 2 + (j >> 6)] |= 1ull << (j & 63);
        ++count[j];
      }
      int run = 0;
      for (int j = 0; j < n2; ++j) {
        rowstart[(size_t)k * n2 + j] = run;
        run += count[j];
      }
    }
  }
  std::vector<float> kin32(kin64.begin(), kin64.end());
  // launch geometry (must fit the LDS FFT's per-thread butterfly budget)
  const int lds_max = max_lds_bytes();
  for (int prec = 0; prec < 2; ++prec) {
    const int tpt = prec == kFP64 ? kTptF64 : kTptF32;
    const int csz = prec == kFP64 ? 16 : 8;
    int w = prec == kFP64 ? 32 : 64;
    int th = 0;
    for (;; w /= 2) {
      th = ((threads_needed(p.ap[0], w, tpt) + 63) / 64) * 64;
      if ((th <= 1024 && (long long)n1 * w * csz <= 65536) || w <= 1) break;
    }
    p.stick_w[prec] = w;
    p.stick_threads[prec] = th < 64 ? 64 : th;
    int pt = std::max(threads_needed(p.ap[2], n2, tpt), threads_needed(p.ap[1], n3, tpt));
    pt = ((pt + 63) / 64) * 64;
    if (pt < 64) pt = 64;
    const long long plds = (long long)n2 * n3 * csz;
    // 0 marks a precision whose plane does not fit one workgroup on this device. The
    // compile-time-length kernels use a fixed 256-thread workgroup (512 for the density)
    // and add the row lists, first-stick indices and 128-bit stick masks (24 n2 bytes).
    if (p.fast)
      p.plane_threads[prec] = (plds + 24LL * n2 <= lds_max) ? 256 : 0;
    else
      p.plane_threads[prec] = (pt <= 1024 && plds <= lds_max) ? pt : 0;
  }
  p.plane_fit[0] = p.plane_threads[0];
  p.plane_fit[1] = p.plane_threads[1];
  if (p.stick_threads[0] > 1024 || p.stick_threads[1] > 1024) {
    set_error("plan_create: axis 0 too long for the stick kernels");
    delete h;
    return 1;
  }
  std::vector<int> nst = p.nsticks;
  if (!upload(&p.d_flatmap, flat, "plan flatmap") || !upload(&p.d_stick_jk, jk, "plan sticks") ||
      !upload(&p.d_stick_map, smap, "plan stick map") || !upload(&p.d_nsticks, nst, "plan nsticks") ||
      !upload(&p.d_rows, rows, "plan rows") || !upload(&p.d_nrows, nrows, "plan nrows") ||
      !upload(&p.d_kin64, kin64, "plan kinetic") || !upload(&p.d_kin32, kin32, "plan kinetic32") ||
      !upload(&p.d_npw, p.npw, "plan npw") ||
      (!rowmask.empty() && (!upload(&p.d_rowmask, rowmask, "plan row masks") ||
                            !upload(&p.d_rowstart, rowstart, "plan row starts") ||
                            !upload(&p.d_livemask, livemask, "plan live rows")))) {
    free_plan_device(p);
    delete h;
    return 1;
  }
  for (int a = 0; a < 3; ++a) {
    std::vector<double> tw;
    twiddles(dims[a], tw);
    std::vector<float> t32(tw.begin(), tw.end());
    if (!upload(reinterpret_cast<double**>(&p.d_tw64[a]), tw, "plan twiddles") ||
        !upload(reinterpret_cast<float**>(&p.d_tw32[a]), t32, "plan twiddles32")) {
      free_plan_device(p);
      delete h;
      return 1;
    }
  }
  *out = h;
  return 0;
}

int amt_plan_destroy(amt_plan_t* h) {
  if (!h) return 0;
  free_plan_device(h->p);
  delete h;
  return 0;
}

int amt_plan_info(const amt_plan_t* h, int* info, int len) {
  if (!h || !info) {
    set_error("plan_info: null argument");
    return 1;
  }
  const Plan& p = h->p;
  const int v[16] = {p.n1,           p.n2,           p.n3,
                     p.nk,           p.npw_pad,      p.smax,
                     p.stick_w[0],   p.stick_w[1],   p.stick_threads[0],
                     p.stick_threads[1], p.plane_threads[0], p.plane_threads[1],
                     p.fast ? p.plane_nth[0] : 0, p.fast ? p.plane_nth[1] : 0,
                     p.fast ? p.fast_stick_w[0] : 0, p.fast ? p.fast_stick_w[1] : 0};
  for (int i = 0; i < len && i < 16; ++i) info[i] = v[i];
  return 0;
}

int amt_plan_set_launch(amt_plan_t* h, int precision, int plane_threads, int stick_width) {
  if (!h || (precision != kFP32 && precision != kFP64)) {
    set_error("plan_set_launch: null plan or bad precision");
    return 1;
  }
  Plan& p = h->p;
  if (!p.fast) {
    set_error("plan_set_launch: only the compile-time-length kernels of cubic grids have these launch shapes");
    return 1;
  }
  const int wdef = precision == kFP64 ? 32 : 64;
  if ((plane_threads != 0 && plane_threads != 256 && plane_threads != 512) ||
      (stick_width != 0 && stick_width != wdef && stick_width != wdef / 2)) {
    set_error("plan_set_launch: plane_threads must be 256 or 512, stick_width " + std::to_string(wdef) + " or " +
              std::to_string(wdef / 2) + " (0 keeps the current value)");
    return 1;
  }
  if (plane_threads) p.plane_nth[precision] = plane_threads;
  if (stick_width) p.fast_stick_w[precision] = stick_width;
  return 0;
}

int amt_plan_nsticks(const amt_plan_t* h, int* out) {
  if (!h || !out) {
    set_error("plan_nsticks: null argument");
    return 1;
  }
  for (int k = 0; k < h->p.nk; ++k) out[k] = h->p.nsticks[k];
  return 0;
}

// Bytes of the local H.psi / density workspace for `rows` band rows: their sticks, or, where
// the plane does not fit LDS (full-grid route), whole grids for min(rows, ~1 GiB of grids).
static size_t local_ws_bytes(const Plan& p, int prec, long long rows) {
  if (rows <= 0 || (prec != kFP32 && prec != kFP64)) return 0;
  const size_t elem = prec == kFP64 ? 16 : 8;
  if (p.plane_threads[prec] > 0) return (size_t)rows * p.n1 * p.smax * elem;
  const size_t grid = (size_t)p.n1 * p.n2 * p.n3 * elem;
  long long g = (long long)((size_t(1) << 30) / grid);
  if (g < 1) g = 1;
  return (size_t)(rows < g ? rows : g) * grid;
}

size_t amt_plan_stick_bytes(const amt_plan_t* h, int precision, int nbands) {
  if (!h || nbands <= 0) return 0;
  return local_ws_bytes(h->p, precision, nbands);
}

int amt_plan_set_route(amt_plan_t* h, int precision, int full_grid) {
  if (!h || (precision != kFP32 && precision != kFP64)) {
    set_error("plan_set_route: null plan or bad precision");
    return 1;
  }
  Plan& p = h->p;
  p.plane_threads[precision] = full_grid ? 0 : p.plane_fit[precision];
  return 0;
}

static int check_plane(const Plan& p, int prec) {
  if (prec != kFP32 && prec != kFP64) {
    set_error("precision must be 0 (fp32) or 1 (fp64)");
    return 1;
  }
  if (p.plane_threads[prec] == 0) {
    set_error("the " + std::to_string(p.n2) + "x" + std::to_string(p.n3) + " plane does not fit one workgroup's LDS at " +
              (prec == kFP64 ? "fp64" : "fp32"));
    return 1;
  }
  return 0;
}

int amt_hpsi_local(const amt_plan_t* h, int precision, const void* psi, void* hpsi, int nb, int k0, int nkk,
                   const void* veff, void* sticks, void* stream) {
  if (!h || !psi || !hpsi || !veff || !sticks) {
    set_error("hpsi_local: null argument");
    return 1;
  }
  const Plan& p = h->p;
  if (precision != kFP32 && precision != kFP64) {
    set_error("precision must be 0 (fp32) or 1 (fp64)");
    return 1;
  }
  if (nb <= 0 || nkk <= 0 || k0 < 0 || k0 + nkk > p.nk) {
    set_error("hpsi_local: band/k range outside the plan");
    return 1;
  }
  if ((long long)nb * nkk > 65535) {
    set_error("hpsi_local: more than 65535 bands in one launch");
    return 1;
  }
  if (p.plane_threads[precision] == 0)
    return launch_hpsi_full_grid(p, precision, psi, hpsi, nb, k0, nkk, veff, sticks,
                                 local_ws_bytes(p, precision, (long long)nb * nkk), (hipStream_t)stream);
  return launch_hpsi_local(p, precision, psi, hpsi, nb, k0, nkk, veff, sticks, (hipStream_t)stream);
}

int amt_plan_set_projectors(amt_plan_t* h, int precision, int nproj, const void* matrix, const void* dij,
                            const int* blk_start, const int* blk_len) {
  if (!h || (precision != kFP32 && precision != kFP64) || nproj < 0) {
    set_error("plan_set_projectors: null plan, bad precision or nproj < 0");
    return 1;
  }
  if (nproj > 0 && (!matrix || !dij || !blk_start || !blk_len)) {
    set_error("plan_set_projectors: null table");
    return 1;
  }
  Plan& p = h->p;
  p.nproj[precision] = nproj;
  p.kb_matrix[precision] = nproj > 0 ? matrix : nullptr;
  p.kb_dij[precision] = nproj > 0 ? dij : nullptr;
  p.kb_blk_start[precision] = nproj > 0 ? blk_start : nullptr;
  p.kb_blk_len[precision] = nproj > 0 ? blk_len : nullptr;
  return 0;
}

namespace {
constexpr size_t kAlign = 256;
size_t align_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }
struct HpsiLayout {
  size_t sticks = 0, proj = 0, pd = 0, gemm = 0, total = 0;
};
HpsiLayout hpsi_layout(const Plan& p, int prec, int nb, int nkk) {
  HpsiLayout l;
  const size_t elem = prec == kFP64 ? 16 : 8;
  l.sticks = align_up(local_ws_bytes(p, prec, (long long)nb * nkk));
  const int np = p.nproj[prec];
  if (np > 0) {
    l.proj = align_up((size_t)nkk * nb * np * elem);
    l.pd = l.proj;
    const size_t g1 = zgemm_workspace_bytes(prec, 1, nb, np, p.npw_pad, nkk);
    const size_t g2 = zgemm_workspace_bytes(prec, 0, nb, p.npw_pad, np, nkk);
    l.gemm = align_up(g1 > g2 ? g1 : g2);
  }
  l.total = l.sticks + l.proj + l.pd + l.gemm;
  return l;
}
}  // namespace

size_t amt_hpsi_workspace(const amt_plan_t* h, int precision, int nb, int nkk) {
  if (!h || (precision != kFP32 && precision != kFP64) || nb <= 0 || nkk <= 0) return 0;
  return hpsi_layout(h->p, precision, nb, nkk).total;
}

int amt_hpsi(const amt_plan_t* h, int precision, const void* psi, void* hpsi, int nb, int k0, int nkk,
             const void* veff, void* ws, size_t ws_bytes, void* stream) {
  if (!h || !psi || !hpsi || !veff || !ws) {
    set_error("hpsi: null argument");
    return 1;
  }
  const Plan& p = h->p;
  if (precision != kFP32 && precision != kFP64) {
    set_error("precision must be 0 (fp32) or 1 (fp64)");
    return 1;
  }
  if (nb <= 0 || nkk <= 0 || k0 < 0 || k0 + nkk > p.nk || (long long)nb * nkk > 65535) {
    set_error("hpsi: band/k range outside the plan");
    return 1;
  }
  const HpsiLayout l = hpsi_layout(p, precision, nb, nkk);
  if (ws_bytes < l.total) {
    set_error("hpsi: workspace smaller than amt_hpsi_workspace()");
    return 1;
  }
  hipStream_t s = (hipStream_t)stream;
  unsigned char* base = static_cast<unsigned char*>(ws);
  if (p.plane_threads[precision] == 0) {
    if (launch_hpsi_full_grid(p, precision, psi, hpsi, nb, k0, nkk, veff, base, l.sticks, s)) return 1;
  } else if (launch_hpsi_local(p, precision, psi, hpsi, nb, k0, nkk, veff, base, s)) {
    return 1;
  }
  const int np = p.nproj[precision];
  if (np == 0) return 0;
  // V_NL psi = (P D) beta with P = psi beta^H (pseudo/kb.py:216-230)
  const size_t elem = precision == kFP64 ? 16 : 8;
  const unsigned char* mat = static_cast<const unsigned char*>(p.kb_matrix[precision]) + (size_t)k0 * np * p.npw_pad * elem;
  const unsigned char* dij =
      static_cast<const unsigned char*>(p.kb_dij[precision]) + (size_t)k0 * np * np * (precision == kFP64 ? 8 : 4);
  void* proj = base + l.sticks;
  void* pd = base + l.sticks + l.proj;
  void* gws = base + l.sticks + l.proj + l.pd;
  const long long spsi = (long long)nb * p.npw_pad, smat = (long long)np * p.npw_pad, sp = (long long)nb * np;
  if (launch_zgemm(precision, 1, nb, np, p.npw_pad, 1.0, psi, p.npw_pad, spsi, mat, p.npw_pad, smat, 0.0, proj, np,
                   sp, nkk, gws, l.gemm, s))
    return 1;
  if (launch_dij(precision, proj, dij, p.kb_blk_start[precision], p.kb_blk_len[precision], pd, nb, np, sp,
                 (long long)np * np, sp, nkk, s))
    return 1;
  return launch_zgemm(precision, 0, nb, p.npw_pad, np, 1.0, pd, np, sp, mat, p.npw_pad, smat, 1.0, hpsi, p.npw_pad,
                      spsi, nkk, gws, l.gemm, s);
}

int amt_hpsi_local_noalloc(const amt_plan_t* h, int precision, const void* psi, void* hpsi, int nb, int k0, int nkk,
                           const void* veff, void* stream) {
  if (!h || !psi || !hpsi || !veff) {
    set_error("hpsi_local_noalloc: null argument");
    return 1;
  }
  const Plan& p = h->p;
  if (check_plane(p, precision)) return 1;
  if (nb <= 0 || nkk <= 0 || k0 < 0 || k0 + nkk > p.nk) {
    set_error("hpsi_local_noalloc: band/k range outside the plan");
    return 1;
  }
  return hpsi_local_noalloc(p, precision, psi, hpsi, nb, k0, nkk, veff, (hipStream_t)stream);
}

int amt_hpsi_stage(const amt_plan_t* h, int precision, int stage, const void* psi, void* hpsi, int nb, int k0,
                   int nkk, const void* veff, void* sticks, void* stream) {
  if (stage < 1 || stage > 3) {
    set_error("hpsi_stage: stage must be 1 (stick inverse), 2 (plane) or 3 (stick forward)");
    return 1;
  }
  if (!h || !psi || !hpsi || !veff || !sticks) {
    set_error("hpsi_stage: null argument");
    return 1;
  }
  const Plan& p = h->p;
  if (check_plane(p, precision)) return 1;
  if (nb <= 0 || nkk <= 0 || k0 < 0 || k0 + nkk > p.nk || (long long)nb * nkk > 65535) {
    set_error("hpsi_stage: band/k range outside the plan");
    return 1;
  }
  return launch_hpsi_local(p, precision, psi, hpsi, nb, k0, nkk, veff, sticks, (hipStream_t)stream, stage);
}

size_t amt_density_part_bytes(const amt_plan_t* h, int nlist, int chunk) {
  if (!h || nlist <= 0 || chunk <= 0) return 0;
  const long long nch = (nlist + chunk - 1) / chunk;
  return (size_t)nch * h->p.n1 * h->p.n2 * h->p.n3 * sizeof(double);
}

int amt_density(const amt_plan_t* h, int precision, const void* psi, int nb, int k0, int nkk, const int* band_list,
                const double* band_weight, int nlist, int chunk, void* sticks, double* part, double* rho,
                double beta, double scale, void* stream) {
  if (!h || !psi || !rho || (nlist > 0 && (!band_list || !band_weight || !sticks || !part))) {
    set_error("density: null argument");
    return 1;
  }
  const Plan& p = h->p;
  if (precision != kFP32 && precision != kFP64) {
    set_error("precision must be 0 (fp32) or 1 (fp64)");
    return 1;
  }
  if (nb <= 0 || nkk <= 0 || k0 < 0 || k0 + nkk > p.nk || chunk <= 0 || nlist < 0) {
    set_error("density: band/k range outside the plan");
    return 1;
  }
  if (p.plane_threads[precision] == 0)
    return launch_density_full_grid(p, precision, psi, nb, k0, band_list, band_weight, nlist, sticks,
                                    local_ws_bytes(p, precision, (long long)nb * nkk), part, rho, beta, scale,
                                    (hipStream_t)stream);
  if (nlist > 0 && (nlist + chunk - 1) / chunk > 65535) {
    set_error("density: too many chunks");
    return 1;
  }
  return launch_density(p, precision, psi, nb, k0, nkk, band_list, band_weight, nlist, chunk, sticks, part, rho,
                        beta, scale, (hipStream_t)stream);
}

int amt_scatter(const amt_plan_t* h, int precision, const void* psi, void* grid, int nb, int k, void* stream) {
  if (!h || !psi || !grid || nb <= 0 || k < 0 || k >= h->p.nk) {
    set_error("scatter: bad argument");
    return 1;
  }
  if (nb > 65535) {
    set_error("scatter: more than 65535 bands");
    return 1;
  }
  return launch_scatter(h->p, precision, psi, grid, nb, k, (hipStream_t)stream);
}

int amt_gather(const amt_plan_t* h, int precision, const void* grid, void* out, int nb, int k, double scale,
               void* stream) {
  if (!h || !grid || !out || nb <= 0 || k < 0 || k >= h->p.nk) {
    set_error("gather: bad argument");
    return 1;
  }
  if (nb > 65535) {
    set_error("gather: more than 65535 bands");
    return 1;
  }
  return launch_gather(h->p, precision, grid, out, nb, k, scale, (hipStream_t)stream);
}

struct amt_fft_plan {
  FftPlan f;
};

static void free_fft_plan(FftPlan& f) {
  for (int a = 0; a < 3; ++a) {
    if (f.tw64[a]) (void)hipFree(f.tw64[a]);
    if (f.tw32[a]) (void)hipFree(f.tw32[a]);
    f.tw64[a] = nullptr;
    f.tw32[a] = nullptr;
  }
}

int amt_fft_plan_create(amt_fft_plan_t** out, int n1, int n2, int n3) {
  if (!out) {
    set_error("fft_plan_create: null argument");
    return 1;
  }
  *out = nullptr;
  auto* h = new amt_fft_plan;
  FftPlan& f = h->f;
  const int dims[3] = {n1, n2, n3};
  for (int a = 0; a < 3; ++a) {
    f.n[a] = dims[a];
    if (dims[a] <= 0 || !make_axis_plan(dims[a], &f.ap[a])) {
      set_error("fft_plan_create: grid length " + std::to_string(dims[a]) + " is not 2,3,5,7,11-smooth");
      delete h;
      return 1;
    }
  }
  for (int a = 0; a < 3; ++a) {
    std::vector<double> tw;
    twiddles(dims[a], tw);
    std::vector<float> t32(tw.begin(), tw.end());
    if (!upload(reinterpret_cast<double**>(&f.tw64[a]), tw, "fft plan twiddles") ||
        !upload(reinterpret_cast<float**>(&f.tw32[a]), t32, "fft plan twiddles32")) {
      free_fft_plan(f);
      delete h;
      return 1;
    }
  }
  *out = h;
  return 0;
}

int amt_fft_plan_destroy(amt_fft_plan_t* h) {
  if (!h) return 0;
  free_fft_plan(h->f);
  delete h;
  return 0;
}

int amt_fft3d(const amt_fft_plan_t* h, int precision, int batch, void* data, int sign, double scale, void* stream) {
  if (!h || !data || batch <= 0 || (sign != 1 && sign != -1) || (precision != kFP32 && precision != kFP64)) {
    set_error("fft3d: bad argument");
    return 1;
  }
  return launch_fft3d(h->f, precision, batch, data, sign, scale, (hipStream_t)stream);
}

size_t amt_zgemm_workspace(int precision, int op, int m, int n, int k, int batch) {
  return zgemm_workspace_bytes(precision, op, m, n, k, batch);
}

int amt_zgemm_set_split(int precision, int op, int m, int n, int k, int batch, int split) {
  if ((precision != kFP32 && precision != kFP64) || (op != 0 && op != 1) || m < 0 || n < 0 || k < 0 || batch < 0) {
    set_error("zgemm_set_split: bad argument");
    return 1;
  }
  return set_zgemm_split(precision, op, m, n, k, batch, split);
}

int amt_zgemm(int precision, int op, int m, int n, int k, double alpha, const void* a, long long lda, long long sa,
              const void* b, long long ldb, long long sb, double beta, void* c, long long ldc, long long sc,
              int batch, void* ws, size_t ws_bytes, void* stream) {
  if ((m > 0 && n > 0 && k > 0) && (!a || !b || !c)) {
    set_error("zgemm: null operand");
    return 1;
  }
  if (op != 0 && op != 1) {
    set_error("zgemm: op must be 0 (NN) or 1 (NC)");
    return 1;
  }
  if (m < 0 || n < 0 || k < 0 || batch < 0 || lda < k || ldc < n || ldb < (op == 0 ? n : k)) {
    set_error("zgemm: bad dimension or leading dimension");
    return 1;
  }
  if ((long long)(m + 63) / 64 > 65535) {
    set_error("zgemm: m too large for one launch");
    return 1;
  }
  return launch_zgemm(precision, op, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch, ws, ws_bytes,
                      (hipStream_t)stream);
}

int amt_residual(int precision, const void* hv, const void* v, const double* eps, void* res, double* norm,
                 int nrow, int npw, void* stream) {
  if (nrow > 0 && (!hv || !v || !eps || !res || !norm)) {
    set_error("residual: null argument");
    return 1;
  }
  return launch_residual(precision, hv, v, eps, res, norm, nrow, npw, (hipStream_t)stream);
}

int amt_precondition(int precision, const void* res, const void* v, const void* kinetic, const int* rows,
                     void* out, int nout, int npw, void* stream) {
  if (nout > 0 && (!res || !v || !kinetic || !rows || !out)) {
    set_error("precondition: null argument");
    return 1;
  }
  return launch_precondition(precision, res, v, kinetic, rows, out, nout, npw, (hipStream_t)stream);
}

int amt_kb_dij(int precision, const void* proj, const void* dij, const int* blk_start, const int* blk_len,
               void* out, int nrow, int nproj, long long sp, long long sd, long long so, int batch, void* stream) {
  if (nrow > 0 && nproj > 0 && (!proj || !dij || !blk_start || !blk_len || !out)) {
    set_error("kb_dij: null argument");
    return 1;
  }
  return launch_dij(precision, proj, dij, blk_start, blk_len, out, nrow, nproj, sp, sd, so, batch,
                    (hipStream_t)stream);
}

}  // extern "C"

// ------------------------------------------------------------------ learned start
int amt_mlp_param_count(void) { return mlp_param_count(); }

int amt_mlp_features(const double* cell, const double* frac, int natoms, int n1, int n2, int n3, const int* images,
                     double rcut, double gamma, float* out, void* stream) {
  if (!cell || !frac || !images || !out || natoms <= 0 || n1 <= 0 || n2 <= 0 || n3 <= 0) {
    set_error("mlp_features: bad argument");
    return 1;
  }
  for (int a = 0; a < 3; ++a)
    if (images[a] < 0 || images[a] > 8) {
      set_error("mlp_features: images per axis must be in 0..8");
      return 1;
    }
  if (!(rcut > 0) || !(gamma > 0)) {
    set_error("mlp_features: rcut and gamma must be positive");
    return 1;
  }
  if (launch_mlp_features(cell, frac, natoms, n1, n2, n3, images[0], images[1], images[2], rcut, gamma, out,
                          (hipStream_t)stream)) {
    set_error(std::string("mlp_features launch: ") + hipGetErrorString(hipGetLastError()));
    return 1;
  }
  return 0;
}

int amt_mlp_forward(const float* feats, long long npts, const float* theta, const float* mean, const float* std,
                    float* y, void* stream) {
  if (npts < 0 || (npts > 0 && (!feats || !theta || !mean || !std || !y))) {
    set_error("mlp_forward: bad argument");
    return 1;
  }
  if (launch_mlp_forward(feats, npts, theta, mean, std, y, (hipStream_t)stream)) {
    set_error(std::string("mlp_forward launch: ") + hipGetErrorString(hipGetLastError()));
    return 1;
  }
  return 0;
}

size_t amt_mlp_workspace(long long npts) {
  const size_t nw = (size_t)mlp_train_waves(npts);
  return nw * ((size_t)mlp_param_count() + 1) * sizeof(float);
}

int amt_mlp_loss_grad(const float* feats, const float* target, long long npts, double count, const float* theta,
                      const float* mean, const float* std, float* grad, float* loss, void* ws, size_t ws_bytes,
                      void* stream) {
  if (npts <= 0 || !feats || !target || !theta || !mean || !std || !grad || !ws || !(count > 0)) {
    set_error("mlp_loss_grad: bad argument");
    return 1;
  }
  if (ws_bytes < amt_mlp_workspace(npts)) {
    set_error("mlp_loss_grad: workspace smaller than amt_mlp_workspace()");
    return 1;
  }
  const int nw = mlp_train_waves(npts);
  float* part = static_cast<float*>(ws);
  float* part_loss = part + (size_t)nw * mlp_param_count();
  if (launch_mlp_train(feats, target, npts, count, theta, mean, std, part, part_loss, grad, loss,
                       (hipStream_t)stream)) {
    set_error(std::string("mlp_loss_grad launch: ") + hipGetErrorString(hipGetLastError()));
    return 1;
  }
  return 0;
}

int amt_adam(float* theta, const float* grad, float* m, float* v, int n, float lr, float beta1, float beta2,
             float eps, int step, void* stream) {
  if (n <= 0 || !theta || !grad || !m || !v || step < 1) {
    set_error("adam: bad argument");
    return 1;
  }
  if (launch_adam(theta, grad, m, v, n, lr, beta1, beta2, eps, step, (hipStream_t)stream)) {
    set_error(std::string("adam launch: ") + hipGetErrorString(hipGetLastError()));
    return 1;
  }
  return 0;
}

// ------------------------------------------------------------------ subspace eigensolve
size_t amt_heev_workspace(int n) { return n > 0 ? heev_workspace_bytes(n) : 0; }

int amt_heev_lowest(int n, const void* a, int nev, double* w, void* zt, void* ws, size_t ws_bytes, void* stream) {
  if (!a || !w || !zt || !ws) {
    set_error("heev: null argument");
    return 1;
  }
  return launch_heev_lowest(n, a, nev, w, zt, ws, ws_bytes, (hipStream_t)stream);
}
