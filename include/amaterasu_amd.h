/*
 * amaterasu_amd -- C ABI of the MI355X (gfx950) H.psi / Davidson / density engine.
 *
 * Conventions
 *   - Every entry point returns 0 on success and nonzero on failure; amt_last_error()
 *     then holds a message (thread-local). Nothing throws across the ABI.
 *   - Device pointers are plain HIP device pointers (e.g. torch tensor data_ptr()).
 *     Complex data is interleaved (re, im): complex64 = float2, complex128 = double2.
 *     `precision` is 0 for FP32/complex64 and 1 for FP64/complex128.
 *   - Band blocks are row-major [band][npw_pad]; a k-batched block is
 *     [k][band][npw_pad]. Grids are C-ordered, flat = (i*n2 + j)*n3 + k.
 *   - `stream` is a hipStream_t (0 = the null stream); every call is asynchronous on it.
 *   - Workspaces (sticks, density partials, GEMM split-K partials) are passed in by the
 *     caller; compute entry points allocate nothing. Only amt_plan_create and
 *     amt_fft_plan_create allocate (amt_plan_set_projectors keeps caller pointers).
 *
 * Reference interfaces (amaterasu 0.1.0a2, paths relative to its source tree) that the
 * entry points replace are cited per function. The Python binding a maintainer adds to
 * the reference is in INTEGRATION.md.
 */
#ifndef AMATERASU_AMD_H
#define AMATERASU_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct amt_plan amt_plan_t;

/* Library identification and the last error of the calling thread. */
const char* amt_version(void);
const char* amt_last_error(void);
/* One line describing the current HIP device (name, arch, CUs, LDS, clocks). */
int amt_device_info(char* buf, int len);

/*
 * Device-resident tables for one FFT grid and a set of k-point spheres: the
 * sphere->grid maps of FFTGrid.indices (basis/fft.py:174-190), the stick
 * decomposition used by the pruned transforms, |k+G|^2/2 (basis/gvectors.py:100-103,
 * GSphere.kinetic_energy) and the twiddle tables.
 *   miller  : sum(npw) x 3 int32 Miller indices, k-point blocks concatenated in
 *             GSphere order (basis/gvectors.py:153-155);
 *   kinetic : sum(npw) float64 |k+G|^2/2 in the same order.
 * Raises the reference's aliasing error (basis/fft.py:183-189) if 2 h_max + 1 > n.
 */
int amt_plan_create(amt_plan_t** out, int n1, int n2, int n3, int nk, const int* npw, int npw_pad,
                    const int* miller, const double* kinetic);
int amt_plan_destroy(amt_plan_t* plan);
/* info[0..15] = n1, n2, n3, nk, npw_pad, smax, stick_w{fp32,fp64}, stick_threads{fp32,fp64},
 * plane_threads{fp32,fp64} (0 = plane does not fit LDS at that precision: full-grid route),
 * and for cubic grids with compile-time-length kernels (else 0) the launch shapes of
 * amt_plan_set_launch: plane workgroup{fp32,fp64}, sticks per workgroup{fp32,fp64}. */
int amt_plan_info(const amt_plan_t* plan, int* info, int len);
/* Tuning hook, no counterpart in the reference: the launch shape of the compile-time-length
 * H.psi kernels at `precision` -- plane kernel workgroup of 256 or 512 threads, stick kernels
 * over 64 or 32 sticks per workgroup (fp32; 32 or 16 in fp64); 0 keeps a value. Results do
 * not depend on it; amaterasu_amd.autotune times the candidates per grid length and keeps
 * the fastest. */
int amt_plan_set_launch(amt_plan_t* plan, int precision, int plane_threads, int stick_width);
int amt_plan_nsticks(const amt_plan_t* plan, int* out);
/* Bytes of the local H.psi / density workspace for `nbands` band rows (all k of one launch):
 * their sticks, or, where the (n2 x n3) plane does not fit one workgroup's LDS at this
 * precision (info[10 + precision] = 0, e.g. FP64 planes above ~100^2), whole grids for
 * min(nbands, ~1 GiB of grids) rows -- the full-grid route (the plan's scatter, 3-D
 * transforms and gather, a chunk of rows at a time). */
size_t amt_plan_stick_bytes(const amt_plan_t* plan, int precision, int nbands);
/* Tuning hook, no counterpart in the reference: full_grid = 1 sends this precision's local
 * H.psi and density through the full-grid route even where the plane fits (for comparison);
 * 0 restores the plane kernels where they fit. Query amt_plan_stick_bytes again afterwards. */
int amt_plan_set_route(amt_plan_t* plan, int precision, int full_grid);

/*
 * Local part of H.psi for nb bands at each of the k-points k0..k0+nk-1:
 *   hpsi = T_G psi + gather(FFT(V_eff . IFFT(scatter psi)))
 * Replaces the first three lines of Hamiltonian.apply (hamiltonian/operator.py:60-66),
 * i.e. apply_kinetic (hamiltonian/kinetic.py:22-24), FFTGrid.wave_to_real and
 * FFTGrid.real_to_wave (basis/fft.py:237-243). V_eff is one spin channel, N reals of
 * `precision`. psi/hpsi: [nk][nb][npw_pad]; padded columns of hpsi are not written.
 * `sticks` holds amt_plan_stick_bytes(plan, precision, nb * nk) bytes. Pruned stick/plane
 * transforms where the plane fits LDS, the full-grid route elsewhere.
 * The non-local term is added afterwards with amt_zgemm / amt_kb_dij (or use amt_hpsi).
 */
int amt_hpsi_local(const amt_plan_t* plan, int precision, const void* psi, void* hpsi, int nb, int k0, int nk,
                   const void* veff, void* sticks, void* stream);

/*
 * Attach the Kleinman-Bylander tables of KBProjectors (pseudo/kb.py:199-260, built by
 * build_projectors kb.py:260-343) to the plan, for amt_hpsi at `precision`:
 *   matrix   : [nk][nproj][npw_pad] complex (beta, zero on padded columns), device;
 *   dij      : [nk][nproj][nproj] real D, block-diagonal over atoms, device;
 *   blk_start, blk_len : [nproj] int, the atom block (first row, length) of each row, device.
 * The plan keeps the pointers (the memory stays the caller's and must outlive its use by
 * amt_hpsi); nothing is copied or allocated. nproj = 0 detaches the tables.
 */
int amt_plan_set_projectors(amt_plan_t* plan, int precision, int nproj, const void* matrix, const void* dij,
                            const int* blk_start, const int* blk_len);
/* Bytes of the workspace amt_hpsi needs for nb bands at each of nk k-points. */
size_t amt_hpsi_workspace(const amt_plan_t* plan, int precision, int nb, int nk);
/*
 * The whole operator of Hamiltonian.apply (hamiltonian/operator.py:60-69) in one call:
 *   hpsi = T_G psi + gather(FFT(V_eff . IFFT(scatter psi))) + (psi beta^H) D beta
 * for nb bands at each of the k-points k0..k0+nk-1 ([nk][nb][npw_pad], the projector term of
 * KBProjectors.apply, pseudo/kb.py:216-230, with the tables of amt_plan_set_projectors; no
 * projector term when none are attached). `ws` holds amt_hpsi_workspace(...) bytes.
 */
int amt_hpsi(const amt_plan_t* plan, int precision, const void* psi, void* hpsi, int nb, int k0, int nk,
             const void* veff, void* ws, size_t ws_bytes, void* stream);

/*
 * amt_hpsi_local without a workspace argument and without any device allocation. No
 * counterpart in the reference: for callers that cannot allocate (code under stream or
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The stick
 * workspace is a fixed per-device array of AMT_NOALLOC_SCRATCH_BYTES inside the library,
 * held by each GPU that runs it until the process ends; rows go through it in launches of
 * at most that many bytes of sticks (whole k-points, or band slices of one k-point), with
 * results bit-identical to amt_hpsi_local. Fails with an error if a single band row needs
 * more sticks than the array holds (n1 * smax * 8 bytes in fp32, 16 in fp64). All calls on
 * one GPU share its array: calls that overlap on one device (other streams, other host
 * threads) race; calls on one stream are ordered.
 */
#define AMT_NOALLOC_SCRATCH_BYTES (64u << 20)
int amt_hpsi_local_noalloc(const amt_plan_t* plan, int precision, const void* psi, void* hpsi, int nb, int k0,
                           int nk, const void* veff, void* stream);

/* One kernel of amt_hpsi_local (stage 1: stick inverse, 2: plane V_eff, 3: stick forward),
 * for timing a single kernel; running 1, 2, 3 in order equals amt_hpsi_local. */
int amt_hpsi_stage(const amt_plan_t* plan, int precision, int stage, const void* psi, void* hpsi, int nb, int k0,
                   int nk, const void* veff, void* sticks, void* stream);

/*
 * Density accumulation (scf/density.py:60-71): for each listed band row l
 * (band_list[l] indexes rows of the [nk][nb][npw_pad] block),
 *   rho = beta*rho + scale * sum_l band_weight[l] |u_l(r)|^2,  u = FFTGrid.wave_to_real.
 * The caller passes scale = 1/Omega and weights (2/nspin) w_k f_nk, as density.py does.
 * rho is float64 N; partials are float64 [ceil(nlist/chunk)][N].
 */
size_t amt_density_part_bytes(const amt_plan_t* plan, int nlist, int chunk);
int amt_density(const amt_plan_t* plan, int precision, const void* psi, int nb, int k0, int nk,
                const int* band_list, const double* band_weight, int nlist, int chunk, void* sticks,
                double* part, double* rho, double beta, double scale, void* stream);

/* FFTGrid.scatter (basis/fft.py:192-206): zero grid, place nb rows of sphere k. */
int amt_scatter(const amt_plan_t* plan, int precision, const void* psi, void* grid, int nb, int k, void* stream);
/* FFTGrid.gather (basis/fft.py:208-212) times `scale`; padded columns are zeroed. */
int amt_gather(const amt_plan_t* plan, int precision, const void* grid, void* out, int nb, int k, double scale,
               void* stream);
/*
 * In-place batched 3D transform over the last three axes of an n1 x n2 x n3 grid, times `scale`:
 *   sign = +1: u(r) = sum_G c_G e^{+iG.r}      (FFTGrid.to_real,      basis/fft.py:219-226, scale 1)
 *   sign = -1: c_G = sum_r u(r) e^{-iG.r}      (FFTGrid.to_reciprocal, basis/fft.py:228-235, scale 1/N)
 * Lengths must be 2,3,5,7,11-smooth (the grid rule of basis/fft.py:72-81). amt_fft_plan_create
 * factorises the lengths and allocates the twiddle tables of both precisions (the only call
 * here that allocates); amt_fft3d then allocates nothing and is asynchronous on `stream`.
 * A plan may be used from several streams at once (it is read-only after creation).
 */
typedef struct amt_fft_plan amt_fft_plan_t;
int amt_fft_plan_create(amt_fft_plan_t** out, int n1, int n2, int n3);
int amt_fft_plan_destroy(amt_fft_plan_t* plan);
int amt_fft3d(const amt_fft_plan_t* plan, int precision, int batch, void* data, int sign, double scale, void* stream);

/*
 * C = alpha A op(B) + beta C on the matrix cores (complex, row-major, real alpha/beta):
 *   op = 0: B is k x n;  op = 1: B is n x k and conj(B)^T is used.
 * Replaces the projector products of KBProjectors.project/apply (pseudo/kb.py:216-230),
 * the Gram and projected matrices of orthonormalize / orthogonalize_against /
 * rayleigh_ritz (scf/eigen/subspace.py:24, 53-54, 72) and their rotations (:75-76).
 * Batched over `batch` with element strides sa/sb/sc. `ws` must hold
 * amt_zgemm_workspace(...) bytes (0 when no split along k is used).
 */
size_t amt_zgemm_workspace(int precision, int op, int m, int n, int k, int batch);
/*
 * Tuning hook, no counterpart in the reference: records the split along k that amt_zgemm
 * (and amt_zgemm_workspace) use for every product in the shape's bucket (precision, op,
 * ceil(m/64), ceil(n/64), ceil(k/256), batch); split 0 restores the built-in heuristic.
 * For callers that time the candidates themselves (amaterasu_amd.autotune does).
 * Returns 2 and records nothing for a product on the 128 x 128-tile path (FP32, op = 1,
 * m and n >= 256), whose split is fixed by its K-chunk bound.
 */
int amt_zgemm_set_split(int precision, int op, int m, int n, int k, int batch, int split);
int amt_zgemm(int precision, int op, int m, int n, int k, double alpha, const void* a, long long lda, long long sa,
              const void* b, long long ldb, long long sb, double beta, void* c, long long ldc, long long sc,
              int batch, void* ws, size_t ws_bytes, void* stream);

/* Davidson residual (scf/eigen/davidson.py:81-82): res = hv - eps v, norm = ||res|| per row (float64). */
int amt_residual(int precision, const void* hv, const void* v, const double* eps, void* res, double* norm,
                 int nrow, int npw, void* stream);
/*
 * Teter-Payne-Allan preconditioning and normalisation of the active residuals
 * (scf/eigen/davidson.py:86-92, scf/eigen/precond.py:27-46): out row a is built from
 * res/v row rows[a]. kinetic has the precision of the data.
 */
int amt_precondition(int precision, const void* res, const void* v, const void* kinetic, const int* rows,
                     void* out, int nout, int npw, void* stream);
/* (P D) of KBProjectors.apply (pseudo/kb.py:230) for block-diagonal D (real, nproj x nproj). */
int amt_kb_dij(int precision, const void* proj, const void* dij, const int* blk_start, const int* blk_len,
               void* out, int nrow, int nproj, long long sp, long long sd, long long so, int batch, void* stream);

/*
 * Lowest nev eigenpairs of a dense Hermitian n x n complex128 matrix (row-major, 2 <= n <= 2048):
 * the subspace eigensolve of rayleigh_ritz (scf/eigen/subspace.py:74, scipy.linalg.eigh there).
 * w: nev eigenvalues ascending (float64, device); zt: nev x n, row k = eigenvector k (device).
 * Householder tridiagonalisation on the device (one launch per column); the lowest
 * eigenpairs of the tridiagonal by multisection bisection and inverse iteration, with a
 * Rayleigh-Ritz step inside clusters of close eigenvalues; back-transformation on the
 * device. a is not modified; ws must hold amt_heev_workspace(n) bytes.
 */
size_t amt_heev_workspace(int n);
int amt_heev_lowest(int n, const void* a, int nev, double* w, void* zt, void* ws, size_t ws_bytes, void* stream);

/*
 * Learned starting density. No counterpart in the reference: it produces the density=
 * argument of run_scf (scf/driver.py:267-299) as a correction over the atomic superposition
 * of PlaneWaveSystem.initial_density (hamiltonian/system.py:147-190), which stays the
 * fallback. Model: 16 radial features of the atoms within r_c of a grid point, three
 * hidden layers of 64 SiLU units, one linear output; bf16 matrix-core operands, FP32 sums.
 * Parameters are one FP32 vector of amt_mlp_param_count() entries:
 *   W1[64][16] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4.
 */
int amt_mlp_param_count(void);
/* Features of every point of an n1 x n2 x n3 grid (C order) into out[N][16] (float32).
 * cell: 3x3 row vectors (Bohr); frac: natoms x 3 fractional coordinates (host pointers are
 * not accepted: both live in device memory). images[3]: lattice images searched per axis. */
int amt_mlp_features(const double* cell, const double* frac, int natoms, int n1, int n2, int n3,
                     const int* images, double rcut, double gamma, float* out, void* stream);
/* y[p] = model output at each of npts feature rows (mean/std: 16 floats each). */
int amt_mlp_forward(const float* feats, long long npts, const float* theta, const float* mean,
                    const float* std, float* y, void* stream);
/* Bytes of workspace amt_mlp_loss_grad needs for npts points. */
size_t amt_mlp_workspace(long long npts);
/* grad = d/dtheta of (1/count) sum_p (y_p - target_p)^2 over the npts rows, loss = that mean
 * (count = npts for a plain mean). Deterministic (fixed reduction order). */
int amt_mlp_loss_grad(const float* feats, const float* target, long long npts, double count, const float* theta,
                      const float* mean, const float* std, float* grad, float* loss, void* ws, size_t ws_bytes,
                      void* stream);
/* One Adam step (Kingma & Ba, bias-corrected; step counts from 1) on n parameters. */
int amt_adam(float* theta, const float* grad, float* m, float* v, int n, float lr, float beta1, float beta2,
             float eps, int step, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AMATERASU_AMD_H */
