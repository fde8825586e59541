// Lowest eigenpairs of a dense Hermitian matrix (FP64 complex, n <= 1024): the subspace
// eigensolve of Rayleigh-Ritz (scf/eigen/subspace.py:74).
//
// 1. Householder tridiagonalisation A = Q T Q^H (the conventions of LAPACK's zhetd2,
//    lower): for column j, beta/tau/v from the column below the diagonal
//    (zlarfg), p = tau A22 v, w = p - (tau/2)(p^H v) v, A22 -= v w^H + w v^H. The whole
//    trailing matrix is updated (both triangles), so row j is the conjugate of column j.
//    One launch per column: launch j finishes step j (reads the complete p_j, forms w_j,
//    updates the rows it owns) and starts step j+1 (every workgroup forms the updated pivot
//    row j+1 in LDS from the untouched global copy, derives v_{j+1}, and multiplies its own
//    rows by it). The kernel boundary is the only global synchronisation; rows are owned
//    cyclically (row r by workgroup r % G).
// 2. The real symmetric tridiagonal eigenproblem by rocSOLVER's divide and conquer (dstedc).
// 3. Back-transformation Z = Q S for the lowest nev eigenvectors: one workgroup per
//    eigenvector applies the reflectors in reverse order (no global synchronisation).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstdlib>
#include <mutex>
#include <vector>

#include "kernels.hpp"

namespace amt {

namespace eig {
constexpr int NT = 256;
constexpr int MAXN = 1024;
}  // namespace eig

__device__ __forceinline__ double2 zmul(double2 a, double2 b) { return cmul(a, b); }
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }

// Block-wide sum of a double2 (all threads get the result). red must hold 2 * NT/64 + 2 doubles.
__device__ __forceinline__ double2 block_sum2(double2 v, double* red) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[2 * wave] = v.x;
    red[2 * wave + 1] = v.y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sx = 0.0, sy = 0.0;
    for (int w = 0; w < eig::NT / 64; ++w) {
      sx += red[2 * w];
      sy += red[2 * w + 1];
    }
    red[2 * (eig::NT / 64)] = sx;
    red[2 * (eig::NT / 64) + 1] = sy;
  }
  __syncthreads();
  const double2 r = make_double2(red[2 * (eig::NT / 64)], red[2 * (eig::NT / 64) + 1]);
  __syncthreads();
  return r;
}

// zlarfg on the column x[0..len) (x[0] = alpha) held in LDS: on return x[1..len) = v[1..],
// v[0] = 1 implicitly; returns beta (real) and tau. Every thread of the workgroup calls it.
__device__ void larfg(double2* x, int len, double* red, double* beta_out, double2* tau_out) {
  double2 part = make_double2(0.0, 0.0);
  for (int i = 1 + threadIdx.x; i < len; i += eig::NT) part.x += x[i].x * x[i].x + x[i].y * x[i].y;
  const double xnorm2 = block_sum2(part, red).x;
  const double2 alpha = x[0];
  if (xnorm2 == 0.0 && alpha.y == 0.0) {
    *beta_out = alpha.x;
    *tau_out = make_double2(0.0, 0.0);
    return;  // H = I; v[1..] = x[1..] = 0 already
  }
  const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
  const double beta = alpha.x >= 0.0 ? -nrm : nrm;
  *beta_out = beta;
  *tau_out = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
  // scale = 1 / (alpha - beta)
  const double2 d = make_double2(alpha.x - beta, alpha.y);
  const double den = d.x * d.x + d.y * d.y;
  const double2 sc = make_double2(d.x / den, -d.y / den);
  for (int i = 1 + threadIdx.x; i < len; i += eig::NT) x[i] = zmul(x[i], sc);
  __syncthreads();
}

// Shared state of one launch.
struct HetrdShared {
  double2 v[eig::MAXN];   // reflector of the current step (v[0] = 1), trailing numbering
  double2 w[eig::MAXN];   // w of the current step
  double2 nr[eig::MAXN];  // the updated pivot row of the next step
  double red[2 * (eig::NT / 64) + 2];
  double2 tau;
  double beta;
};

// Prologue: v_0 from row 0, d_0 / e_0 / tau_0, V[0], and p_0 = tau_0 A[1:,1:] v_0 for own rows.
__global__ __launch_bounds__(eig::NT) void k_hetrd_start(const double2* __restrict__ A, int n, double2* __restrict__ V,
                                                         double2* __restrict__ taus, double* __restrict__ d,
                                                         double* __restrict__ e, double2* __restrict__ p) {
  __shared__ HetrdShared sh;
  const int len = n - 1;  // length of column 0 below the diagonal
  for (int t = threadIdx.x; t < len; t += eig::NT) sh.v[t] = zconj(A[1 + t]);  // column 0 = conj(row 0)
  __syncthreads();
  double beta;
  double2 tau;
  larfg(sh.v, len, sh.red, &beta, &tau);
  if (blockIdx.x == 0) {
    for (int t = threadIdx.x; t < len; t += eig::NT) V[t] = t == 0 ? make_double2(1.0, 0.0) : sh.v[t];
    if (threadIdx.x == 0) {
      taus[0] = tau;
      d[0] = A[0].x;
      e[0] = beta;
    }
  }
  // own rows r >= 1 (trailing index t = r - 1): p[r-1] = tau * sum_c A[r][c] v[c-1], c >= 1
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int k = 0;
  for (int r = 1 + blockIdx.x; r < n; r += gridDim.x, ++k) {
    if ((k & 3) != wave) continue;  // rows of this workgroup spread over its 4 waves
    double2 acc = make_double2(0.0, 0.0);
    for (int c = 1 + lane; c < n; c += 64) {
      const double2 a = A[(long long)r * n + c];
      const double2 vv = c == 1 ? make_double2(1.0, 0.0) : sh.v[c - 1];
      acc.x += a.x * vv.x - a.y * vv.y;
      acc.y += a.x * vv.y + a.y * vv.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
      acc.x += __shfl_down(acc.x, off, 64);
      acc.y += __shfl_down(acc.y, off, 64);
    }
    if (lane == 0) p[r - 1] = zmul(tau, acc);
  }
}

// Launch j (0 <= j <= n-2): finish step j, start step j+1. Trailing numbering of step j:
// index t <-> row/column j+1+t. p_in: p_j (length n-j-1); p_out: p_{j+1}.
// Each workgroup owns at most one trailing row per wave (n - 2 <= 4 G, checked by the host):
// the wave keeps that row in registers from its update to its product with v_{j+1}, and the
// old pivot row j+1 is fetched into registers at the start, so the only global round trips
// on the critical path are the loads of v_j, p_j and the owned row.
constexpr int kRowRegs = eig::MAXN / 64;  // elements of one row per lane

__global__ __launch_bounds__(eig::NT) void k_hetrd_step(double2* __restrict__ A, int n, int j,
                                                        double2* __restrict__ V, double2* __restrict__ taus,
                                                        double* __restrict__ d, double* __restrict__ e,
                                                        const double2* __restrict__ p_in, double2* __restrict__ p_out) {
  __shared__ HetrdShared sh;
  const int len = n - j - 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // the owned row of this wave (if any) and the old pivot row, straight into registers
  const int myrow = j + 2 + blockIdx.x + wave * gridDim.x;
  const bool own = myrow < n;
  double2 rowv[kRowRegs];
#pragma unroll
  for (int q = 0; q < kRowRegs; ++q) {
    const int c = j + 1 + lane + 64 * q;
    rowv[q] = (own && c < n) ? A[(long long)myrow * n + c] : make_double2(0.0, 0.0);
  }
  double2 piv[kRowRegs / 4];  // pivot row j+1, columns j+1 + threadIdx.x + 256 q
#pragma unroll
  for (int q = 0; q < kRowRegs / 4; ++q) {
    const int c = j + 1 + threadIdx.x + eig::NT * q;
    piv[q] = c < n ? A[(long long)(j + 1) * n + c] : make_double2(0.0, 0.0);
  }
  const double2 tau = taus[j];
  for (int t = threadIdx.x; t < len; t += eig::NT) {
    sh.v[t] = V[(long long)j * n + t];
    sh.w[t] = p_in[t];
  }
  __syncthreads();
  // K = -(tau/2) p^H v; w = p + K v
  double2 part = make_double2(0.0, 0.0);
  for (int t = threadIdx.x; t < len; t += eig::NT) {
    const double2 pv = zmul(zconj(sh.w[t]), sh.v[t]);
    part.x += pv.x;
    part.y += pv.y;
  }
  const double2 phv = block_sum2(part, sh.red);
  const double2 K = zmul(make_double2(-0.5 * tau.x, -0.5 * tau.y), phv);
  for (int t = threadIdx.x; t < len; t += eig::NT) {
    const double2 kv = zmul(K, sh.v[t]);
    sh.w[t] = make_double2(sh.w[t].x + kv.x, sh.w[t].y + kv.y);
  }
  __syncthreads();
  // update the owned row in registers and store it:  A -= v w^H + w v^H
  if (own) {
    const int tr = myrow - j - 1;
    const double2 vr = sh.v[tr], wr = sh.w[tr];
    double2* row = A + (long long)myrow * n;
#pragma unroll
    for (int q = 0; q < kRowRegs; ++q) {
      const int c = j + 1 + lane + 64 * q;
      if (c < n) {
        const int tc = c - j - 1;
        const double2 a = zmul(vr, zconj(sh.w[tc]));
        const double2 b = zmul(wr, zconj(sh.v[tc]));
        rowv[q].x -= a.x + b.x;
        rowv[q].y -= a.y + b.y;
        row[c] = rowv[q];
      }
    }
  }
  const int len2 = n - j - 2;  // length of column j+1 below the diagonal
  if (len2 <= 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const double2 a = zmul(sh.v[0], zconj(sh.w[0]));
      d[j + 1] = piv[0].x - 2.0 * a.x;
    }
    return;
  }
  // the updated pivot row j+1: old - (v_0 conj(w_c) + w_0 conj(v_c)), v_0 = 1
#pragma unroll
  for (int q = 0; q < kRowRegs / 4; ++q) {
    const int c = j + 1 + threadIdx.x + eig::NT * q;
    if (c < n) {
      const int tc = c - j - 1;
      const double2 a = zmul(sh.v[0], zconj(sh.w[tc]));
      const double2 b = zmul(sh.w[0], zconj(sh.v[tc]));
      sh.nr[tc] = make_double2(piv[q].x - a.x - b.x, piv[q].y - a.y - b.y);
    }
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) d[j + 1] = sh.nr[0].x;
  // column j+1 below the diagonal = conj(new row j+1, c >= j+2) -> reflector v_{j+1} (in sh.v)
  for (int t = threadIdx.x; t < len2; t += eig::NT) sh.v[t] = zconj(sh.nr[1 + t]);
  __syncthreads();
  double beta;
  double2 tau2;
  larfg(sh.v, len2, sh.red, &beta, &tau2);
  if (blockIdx.x == 0) {
    for (int t = threadIdx.x; t < len2; t += eig::NT)
      V[(long long)(j + 1) * n + t] = t == 0 ? make_double2(1.0, 0.0) : sh.v[t];
    if (threadIdx.x == 0) {
      taus[j + 1] = tau2;
      e[j + 1] = beta;
    }
  }
  // p_{j+1} for the owned row: tau2 * sum_{c >= j+2} A[r][c] v_{j+1}[c - j - 2], from registers
  if (own) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < kRowRegs; ++q) {
      const int c = j + 1 + lane + 64 * q;
      if (c >= j + 2 && c < n) {
        const double2 vv = c == j + 2 ? make_double2(1.0, 0.0) : sh.v[c - j - 2];
        acc.x += rowv[q].x * vv.x - rowv[q].y * vv.y;
        acc.y += rowv[q].x * vv.y + rowv[q].y * vv.x;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      acc.x += __shfl_down(acc.x, off, 64);
      acc.y += __shfl_down(acc.y, off, 64);
    }
    if (lane == 0) p_out[myrow - j - 2] = zmul(tau2, acc);
  }
}

// Z^T[k][:] = Q S[:, k] for k < nev; S column-major (ldc = n), real.
__global__ __launch_bounds__(eig::NT) void k_backtransform(const double2* __restrict__ V,
                                                           const double2* __restrict__ taus,
                                                           const double* __restrict__ S, int n,
                                                           double2* __restrict__ Zt) {
  __shared__ double2 z[eig::MAXN];
  __shared__ double red[2 * (eig::NT / 64) + 2];
  const int k = blockIdx.x;
  for (int i = threadIdx.x; i < n; i += eig::NT) z[i] = make_double2(S[(long long)k * n + i], 0.0);
  __syncthreads();
  for (int j = n - 2; j >= 0; --j) {
    const int len = n - j - 1;
    const double2* v = V + (long long)j * n;
    double2 part = make_double2(0.0, 0.0);
    for (int t = threadIdx.x; t < len; t += eig::NT) {
      const double2 c = zmul(zconj(v[t]), z[j + 1 + t]);
      part.x += c.x;
      part.y += c.y;
    }
    const double2 y = block_sum2(part, red);
    const double2 ty = zmul(taus[j], y);
    for (int t = threadIdx.x; t < len; t += eig::NT) {
      const double2 u = zmul(v[t], ty);
      z[j + 1 + t].x -= u.x;
      z[j + 1 + t].y -= u.y;
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < n; i += eig::NT) Zt[(long long)k * n + i] = z[i];
}

// ------------------------------------------------------------------ host
static std::mutex g_handle_mutex;
static rocblas_handle g_handles[64] = {};

static rocblas_handle blas_handle() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(g_handle_mutex);
  if (!g_handles[dev] && rocblas_create_handle(&g_handles[dev]) != rocblas_status_success) g_handles[dev] = nullptr;
  return g_handles[dev];
}

// The n launches of the tridiagonalisation, captured once per (device, workspace, n) into
// a HIP graph and replayed: the launch sequence is fixed by n and every pointer lives in the
// caller's workspace, so only the input copy (outside the graph) differs between calls.
struct TridiagGraph {
  int dev;
  const void* ws;
  int n;
  hipGraphExec_t exec;
};
static std::mutex g_graph_mutex;
static std::vector<TridiagGraph> g_graphs;

// Workgroups of the tridiagonalisation: one trailing row per wave at most (n - 2 <= 4 G).
static int hetrd_groups(int n) {
  int g = (n - 2 + 3) / 4;
  if (g < 1) g = 1;
  if (g > 256) g = 256;
  return g;
}

static void enqueue_tridiag(hipStream_t s, int n, double2* A, double2* V, double2* taus, double* d, double* e,
                            double2* p0, double2* p1) {
  const int G = hetrd_groups(n);
  hipLaunchKernelGGL(k_hetrd_start, dim3(G), dim3(eig::NT), 0, s, A, n, V, taus, d, e, p0);
  for (int j = 0; j <= n - 2; ++j) {
    const double2* pin = (j & 1) ? p1 : p0;
    double2* pout = (j & 1) ? p0 : p1;
    hipLaunchKernelGGL(k_hetrd_step, dim3(G), dim3(eig::NT), 0, s, A, n, j, V, taus, d, e, pin, pout);
  }
}

static hipGraphExec_t tridiag_graph(int n, const void* ws, double2* A, double2* V, double2* taus, double* d,
                                    double* e, double2* p0, double2* p1) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lock(g_graph_mutex);
  for (const TridiagGraph& g : g_graphs)
    if (g.dev == dev && g.ws == ws && g.n == n) return g.exec;
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ok = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess;
  if (ok) {
    enqueue_tridiag(cs, n, A, V, taus, d, e, p0, p1);
    ok = hipStreamEndCapture(cs, &graph) == hipSuccess && graph != nullptr;
  }
  if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipStreamDestroy(cs);
  if (!ok) {
    (void)hipGetLastError();
    return nullptr;
  }
  if (g_graphs.size() >= 32) {  // bounded cache: drop the oldest
    (void)hipGraphExecDestroy(g_graphs.front().exec);
    g_graphs.erase(g_graphs.begin());
  }
  g_graphs.push_back({dev, ws, n, exec});
  return exec;
}

size_t heev_workspace_bytes(int n) {
  // A copy, V, p (2), taus, S (n x n real), d, e, info
  return (size_t)n * n * 16 * 2 + (size_t)2 * n * 16 + (size_t)n * 16 + (size_t)n * n * 8 + (size_t)2 * n * 8 + 256;
}

int launch_heev_lowest(int n, const void* a, int nev, double* w, void* zt, void* ws, size_t ws_bytes,
                       hipStream_t s) {
  if (n < 2 || n > eig::MAXN || nev < 1 || nev > n) {
    set_error("heev: need 2 <= n <= 1024 and 1 <= nev <= n");
    return 1;
  }
  if (ws_bytes < heev_workspace_bytes(n)) {
    set_error("heev: workspace smaller than amt_heev_workspace(n)");
    return 1;
  }
  char* base = static_cast<char*>(ws);
  double2* A = reinterpret_cast<double2*>(base);
  double2* V = A + (size_t)n * n;
  double2* p0 = V + (size_t)n * n;
  double2* p1 = p0 + n;
  double2* taus = p1 + n;
  double* S = reinterpret_cast<double*>(taus + n);
  double* d = S + (size_t)n * n;
  double* e = d + n;
  rocblas_int* info = reinterpret_cast<rocblas_int*>(e + n);
  if (hipMemcpyAsync(A, a, (size_t)n * n * 16, hipMemcpyDeviceToDevice, s) != hipSuccess) {
    set_error("heev: copy of the input failed");
    return 1;
  }
  const char* nog = std::getenv("AMT_HEEV_NO_GRAPH");
  hipGraphExec_t exec = (nog && nog[0] == '1') ? nullptr : tridiag_graph(n, ws, A, V, taus, d, e, p0, p1);
  if (exec) {
    if (hipGraphLaunch(exec, s) != hipSuccess) {
      set_error("heev: tridiagonalisation graph launch failed");
      return 1;
    }
  } else {
    enqueue_tridiag(s, n, A, V, taus, d, e, p0, p1);
  }
  if (hipGetLastError() != hipSuccess) {
    set_error("heev: tridiagonalisation launch failed");
    return 1;
  }
  rocblas_handle h = blas_handle();
  if (!h || rocblas_set_stream(h, s) != rocblas_status_success) {
    set_error("heev: rocBLAS handle unavailable");
    return 1;
  }
  if (rocsolver_dstedc(h, rocblas_evect_tridiagonal, n, d, e, S, n, info) != rocblas_status_success) {
    set_error("heev: rocsolver_dstedc failed");
    return 1;
  }
  hipLaunchKernelGGL(k_backtransform, dim3(nev), dim3(eig::NT), 0, s, V, taus, S, n,
                     static_cast<double2*>(zt));
  if (hipGetLastError() != hipSuccess) {
    set_error("heev: back-transformation launch failed");
    return 1;
  }
  if (hipMemcpyAsync(w, d, (size_t)nev * 8, hipMemcpyDeviceToDevice, s) != hipSuccess) {
    set_error("heev: copy of the eigenvalues failed");
    return 1;
  }
  return 0;
}

}  // namespace amt
