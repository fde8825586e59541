// Host-side launcher declarations shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "fft_lds.hpp"

namespace amt {

enum Precision { kFP32 = 0, kFP64 = 1 };

// Butterflies per thread held in registers during one Stockham pass (radix 4..11; radix 2
// and 3 passes hold twice as many). FP64 holds half as many to bound register use.
constexpr int kTptF32 = 4;
constexpr int kTptF64 = 2;

// Factorise n into the supported radices; false if n is not 2,3,5,7,11-smooth or too long.
bool make_axis_plan(int n, AxisPlan* out);
// Threads a workgroup needs so every pass of `ap` over `nseq` sequences fits TPT per thread.
int threads_needed(const AxisPlan& ap, int nseq, int tpt);
// Twiddle table exp(-2 pi i m / n), m = 0..n-1, in double and float.
void twiddles(int n, std::vector<double>& tw);

// ------------------------------------------------------------------ plan (device state)
struct Plan {
  int n1 = 0, n2 = 0, n3 = 0;
  int nk = 0;
  int npw_pad = 0;
  int smax = 0;  // sticks of the widest k-point
  std::vector<int> npw;
  std::vector<int> nsticks;
  AxisPlan ap[3];
  // device tables
  int* d_flatmap = nullptr;    // [nk][npw_pad] flat C-order grid index, -1 on padding
  int* d_stick_jk = nullptr;   // [nk][smax] j*n3 + k of each stick, -1 on padding
  int* d_stick_map = nullptr;  // [nk][n1][smax] sphere index at (i, stick) or -1
  int* d_nsticks = nullptr;    // [nk]
  int* d_rows = nullptr;       // [nk][n2] the rows j that hold sticks (first nrows[k] entries)
  int* d_nrows = nullptr;      // [nk]
  int* d_npw = nullptr;        // [nk]
  // Per k-point and row j (n3 <= 128 only): which columns k of row j hold sticks, as a
  // 128-bit mask (two words), and the index of the row's first stick (sticks are sorted by
  // j * n3 + k, so stick (j, k) = rowstart + number of set bits of the mask below k).
  unsigned long long* d_rowmask = nullptr;  // [nk][n2][2]
  int* d_rowstart = nullptr;                // [nk][n2]
  unsigned long long* d_livemask = nullptr; // [nk][2]: rows j that hold sticks (n2 <= 128)
  bool fast = false;           // cubic grid of a length with compile-time kernels
  double* d_kin64 = nullptr;   // [nk][npw_pad] |k+G|^2/2 (0 on padding)
  float* d_kin32 = nullptr;
  double2* d_tw64[3] = {nullptr, nullptr, nullptr};
  float2* d_tw32[3] = {nullptr, nullptr, nullptr};
  // Kleinman-Bylander tables the caller attached (amt_plan_set_projectors), per precision:
  // matrix [nk][nproj][npw_pad] complex, D [nk][nproj][nproj] real, atom blocks [nproj].
  // Caller-owned device memory: the plan keeps the pointers, not copies.
  int nproj[2] = {0, 0};
  const void* kb_matrix[2] = {nullptr, nullptr};
  const void* kb_dij[2] = {nullptr, nullptr};
  const int* kb_blk_start[2] = {nullptr, nullptr};
  const int* kb_blk_len[2] = {nullptr, nullptr};
  // launch geometry
  int stick_w[2] = {0, 0};         // sticks per workgroup, [prec]
  int stick_threads[2] = {0, 0};   // workgroup size of the stick kernels
  int plane_threads[2] = {0, 0};   // workgroup size of the plane kernel (0: full-grid route)
  int plane_fit[2] = {0, 0};       // the plane kernel's size where it fits (amt_plan_set_route restores it)
  // launch shapes of the compile-time-length kernels (amt_plan_set_launch; an autotuner picks them)
  int plane_nth[2] = {256, 256};   // plane kernel workgroup: 256 or 512 threads
  int fast_stick_w[2] = {64, 32};  // sticks per workgroup of the stick kernels: the default or half
};

void set_error(const std::string& msg);

// Compile-time-length kernels (fft_fast.hip) for cubic grids of the listed lengths.
bool fast_length_supported(int n);
int launch_hpsi_fast(const Plan& p, int prec, const void* psi, void* hpsi, int nb, int k0, int nkk,
                     const void* veff, void* sticks, hipStream_t s, int stage);
int launch_density_fast(const Plan& p, int prec, const void* psi, int nb, int k0, int nkk, const int* list,
                        const double* weight, int nlist, int chunk, void* sticks, double* part, hipStream_t s);

// Local part of H.psi (kinetic + V_eff) for nbk = nkb bands laid out [kb][npw_pad],
// kb = k * nb + b for k in [k0, k0 + nkb / nb). `sticks` is nkb * n1 * smax complex.
int launch_hpsi_local(const Plan& p, int prec, const void* psi, void* hpsi, int nb, int k0, int nkk,
                      const void* veff, void* sticks, hipStream_t s, int stage = 0);
// launch_hpsi_local over the library's fixed scratch array (noalloc.hip).
long long noalloc_rows_per_launch(const Plan& p, int prec);
int hpsi_local_noalloc(const Plan& p, int prec, const void* psi, void* hpsi, int nb, int k0, int nkk,
                       const void* veff, hipStream_t s);
// sum over listed bands of w |u(r)|^2 into part[chunk][N] (double), then
// rho = beta * rho + scale * sum_chunks part.
int launch_density(const Plan& p, int prec, const void* psi, int nb, int k0, int nkk, const int* d_band_list,
                   const double* d_band_weight, int nlist, int chunk, void* sticks, double* part, double* rho,
                   double beta, double scale, hipStream_t s);
// The full-grid route (whole grids, a chunk of rows at a time in `ws`) for planes that do not
// fit one workgroup's LDS; full_grid_rows = grids the workspace holds.
long long full_grid_rows(const Plan& p, int prec, size_t ws_bytes);
int launch_hpsi_full_grid(const Plan& p, int prec, const void* psi, void* hpsi, int nb, int k0, int nkk,
                          const void* veff, void* ws, size_t ws_bytes, hipStream_t s);
int launch_density_full_grid(const Plan& p, int prec, const void* psi, int nb, int k0, const int* list,
                             const double* weight, int nlist, void* ws, size_t ws_bytes, double* acc, double* rho,
                             double beta, double scale, hipStream_t s);
int launch_scatter(const Plan& p, int prec, const void* psi, void* grid, int nb, int k, hipStream_t s);
int launch_gather(const Plan& p, int prec, const void* grid, void* out, int nb, int k, double scale,
                  hipStream_t s);
// Twiddle tables and axis factorisations of one 3-D grid (amt_fft_plan_create).
struct FftPlan {
  int n[3] = {0, 0, 0};
  AxisPlan ap[3];
  double2* tw64[3] = {nullptr, nullptr, nullptr};
  float2* tw32[3] = {nullptr, nullptr, nullptr};
};
int launch_fft3d(const FftPlan& fp, int prec, int batch, void* data, int sign, double scale, hipStream_t s);

// C = alpha * A op(B) + beta * C, complex, row-major; op = 0: B is K x N, op = 1: B is N x K
// and conj(B)^T is used. Batched with element strides. `ws` holds split-K partials.
int launch_zgemm(int prec, int op, int m, int n, int k, double alpha, const void* a, long long lda,
                 long long sa, const void* b, long long ldb, long long sb, double beta, void* c, long long ldc,
                 long long sc, int batch, void* ws, size_t ws_bytes, hipStream_t s);
size_t zgemm_workspace_bytes(int prec, int op, int m, int n, int k, int batch);
int set_zgemm_split(int prec, int op, int m, int n, int k, int batch, int split);

// Lowest eigenpairs of a dense Hermitian FP64 matrix (eig.hip).
size_t heev_workspace_bytes(int n);
int launch_heev_lowest(int n, const void* a, int nev, double* w, void* zt, void* ws, size_t ws_bytes, hipStream_t s);

// Learned starting density (mlp.hip).
int mlp_param_count();
int mlp_train_waves(long long npts);
int launch_mlp_features(const double* cell, const double* frac, int natoms, int n1, int n2, int n3, int r1, int r2,
                        int r3, double rcut, double gamma, float* out, hipStream_t s);
int launch_mlp_forward(const float* feats, long long npts, const float* theta, const float* mean, const float* stdv,
                       float* y, hipStream_t s);
int launch_mlp_train(const float* feats, const float* target, long long npts, double count, const float* theta,
                     const float* mean, const float* stdv, float* part, float* part_loss, float* grad, float* loss,
                     hipStream_t s);
int launch_adam(float* theta, const float* grad, float* m, float* v, int n, float lr, float beta1, float beta2,
                float eps, int step, hipStream_t s);

}  // namespace amt
