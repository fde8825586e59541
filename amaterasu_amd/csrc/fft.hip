// Sphere<->grid maps, the 3D FFT pair and the fused local part of H.psi.
//
// Layout. A band's coefficients are a row of npw_pad complex numbers in the sphere's
// own order (|k+G|^2, then Miller index). The real-space grid is C-ordered,
// flat = (i n2 + j) n3 + k with (i, j, k) = Miller index mod (n1, n2, n3).
//
// Pruned transform. The wavefunction sphere occupies ~7% of the grid and its
// projection on the (j, k) plane ~20% of the columns along i ("sticks"). The local
// part of H.psi therefore runs as three kernels that never write a full grid:
//   stick_inverse : gather the sphere into sticks, FFT along i        -> S x n1 values
//   plane_vloc    : per i-plane, place sticks, 2D inverse FFT, x V_eff(r),
//                   2D forward FFT, pick the sticks back out           (in place)
//   stick_forward : FFT along i, scale 1/N, pick sphere points, add T_G c_G
// HBM traffic per band is ~4 S n1 complex plus the sphere in and out, against
// ~8 N for three full-grid transform passes each way.
#include <cmath>
#include <vector>

#include "kernels.hpp"

namespace amt {

// ----------------------------------------------------------------------------- host
bool make_axis_plan(int n, AxisPlan* out) {
  if (n < 1 || n > 4096) return false;
  AxisPlan ap{};
  ap.n = n;
  int rem = n, np = 0;
  int radix[kMaxPasses];
  // powers of two: radix 8 first, then one 4 and/or one 2 for the remainder
  while (rem % 8 == 0 && np < kMaxPasses) {
    radix[np++] = 8;
    rem /= 8;
  }
  if (rem % 4 == 0 && np < kMaxPasses) {
    radix[np++] = 4;
    rem /= 4;
  }
  if (rem % 2 == 0 && np < kMaxPasses) {
    radix[np++] = 2;
    rem /= 2;
  }
  const int odd[4] = {3, 5, 7, 11};
  for (int f : odd) {
    while (rem % f == 0 && np < kMaxPasses) {
      radix[np++] = f;
      rem /= f;
    }
  }
  if (rem != 1) return false;
  ap.npass = np;
  ap.code = 0;
  for (int p = 0; p < np; ++p) ap.code |= (unsigned long long)radix[p] << (4 * p);
  *out = ap;
  return true;
}

int threads_needed(const AxisPlan& ap, int nseq, int tpt) {
  int need = 1;
  for (int p = 0; p < ap.npass; ++p) {
    const int r = ap.radix(p);
    const int t = tasks_per_thread(r, tpt);
    const long long tasks = (long long)nseq * (ap.n / r);
    const int th = (int)((tasks + t - 1) / t);
    if (th > need) need = th;
  }
  return need;
}

void twiddles(int n, std::vector<double>& tw) {
  tw.resize(2 * (size_t)n);
  for (int m = 0; m < n; ++m) {
    // exp(-2 pi i m/n); reduce the angle to the first octant-free form m/n in [0,1)
    const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
    tw[2 * m] = (double)cosl(a);
    tw[2 * m + 1] = (double)sinl(a);
  }
}

// --------------------------------------------------------------------------- kernels
template <typename C>
__global__ void k_fill_zero(C* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = cmake<C>(0, 0);
}

// grid[b][flatmap[g]] = psi[b][g] for the npw real entries (grid zeroed beforehand)
template <typename C>
__global__ void k_scatter(const C* __restrict__ psi, C* __restrict__ grid, const int* __restrict__ flatmap,
                          int npw, int npw_pad, long long ngrid) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (g < npw) {
    const int f = flatmap[g];
    if (f >= 0) grid[(long long)b * ngrid + f] = psi[(long long)b * npw_pad + g];
  }
}

template <typename C>
__global__ void k_gather(const C* __restrict__ grid, C* __restrict__ out, const int* __restrict__ flatmap,
                         int npw, int npw_pad, long long ngrid, decltype(C::x) scale) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (g < npw_pad) {
    C v = cmake<C>(0, 0);
    if (g < npw) {
      const int f = flatmap[g];
      if (f >= 0) v = cscale(grid[(long long)b * ngrid + f], scale);
    }
    out[(long long)b * npw_pad + g] = v;
  }
}

// Batched 1D FFT along one axis of a (nouter, n, inner) array (element stride `inner`
// along the axis). inner > 1: a workgroup owns W adjacent columns of one outer slab.
// inner == 1: a workgroup owns W adjacent contiguous rows.
template <int SIGN, int TPT, typename C>
__global__ void k_axis_fft(C* __restrict__ data, AxisPlan ap, long long inner, long long nouter, int W,
                           const C* __restrict__ tw, decltype(C::x) scale) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int n = ap.n;
  const int tid = threadIdx.x, nth = blockDim.x;
  if (inner > 1) {
    const long long ncolblk = (inner + W - 1) / W;
    const long long o = blockIdx.x / ncolblk;
    const long long c0 = (blockIdx.x % ncolblk) * W;
    const int wcnt = (int)((inner - c0) < W ? (inner - c0) : W);
    C* base = data + o * n * inner + c0;
    for (int e = tid; e < n * W; e += nth) {
      const int i = e / W, w = e % W;
      lds[e] = (w < wcnt) ? base[(long long)i * inner + w] : cmake<C>(0, 0);
    }
    __syncthreads();
    lds_fft<SIGN, TPT>(lds, ap, W, W, 1, true, tw, tid, nth);
    for (int e = tid; e < n * W; e += nth) {
      const int i = e / W, w = e % W;
      if (w < wcnt) base[(long long)i * inner + w] = cscale(lds[e], scale);
    }
  } else {
    const long long r0 = (long long)blockIdx.x * W;
    const int rcnt = (int)((nouter - r0) < W ? (nouter - r0) : W);
    C* base = data + r0 * n;
    for (int e = tid; e < n * W; e += nth) lds[e] = (e < n * rcnt) ? base[e] : cmake<C>(0, 0);
    __syncthreads();
    lds_fft<SIGN, TPT>(lds, ap, W, 1, n, false, tw, tid, nth);
    for (int e = tid; e < n * rcnt; e += nth) base[e] = cscale(lds[e], scale);
  }
}

// Gather the sphere of band kb into W sticks and transform along i (inverse).
// sticks layout [kb][i][smax].
template <int TPT, typename C>
__global__ void k_stick_inverse(const C* __restrict__ psi, C* __restrict__ sticks,
                                const int* __restrict__ stick_map, int n1, int smax, int npw_pad, int nb, int k0,
                                AxisPlan ap1, const C* __restrict__ tw1, int W) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int kb = blockIdx.y;
  const int k = k0 + kb / nb;
  const int s0 = blockIdx.x * W;
  const C* c = psi + (long long)kb * npw_pad;
  const int* mp = stick_map + (long long)k * n1 * smax;
  for (int e = tid; e < n1 * W; e += nth) {
    const int i = e / W, w = e % W, s = s0 + w;
    C v = cmake<C>(0, 0);
    if (s < smax) {
      const int g = mp[i * smax + s];
      if (g >= 0) v = c[g];
    }
    lds[e] = v;
  }
  __syncthreads();
  lds_fft<1, TPT>(lds, ap1, W, W, 1, true, tw1, tid, nth);
  C* out = sticks + (long long)kb * n1 * smax;
  for (int e = tid; e < n1 * W; e += nth) {
    const int i = e / W, w = e % W, s = s0 + w;
    if (s < smax) out[i * smax + s] = lds[e];
  }
}

// Per (band, i-plane): sticks -> plane, 2D inverse FFT, x V_eff, 2D forward FFT, plane -> sticks.
template <int TPT, typename C, typename R>
__global__ void k_plane_vloc(C* __restrict__ sticks, const int* __restrict__ stick_jk,
                             const int* __restrict__ nsticks, const R* __restrict__ veff, int n1, int n2, int n3,
                             int smax, int nb, int k0, AxisPlan ap2, AxisPlan ap3, const C* __restrict__ tw2,
                             const C* __restrict__ tw3) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  C* plane = reinterpret_cast<C*>(smem_raw);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int i = blockIdx.x;
  const int kb = blockIdx.y;
  const int k = k0 + kb / nb;
  const int S = nsticks[k];
  const int np = n2 * n3;
  const int* jk = stick_jk + (long long)k * smax;
  C* st = sticks + ((long long)kb * n1 + i) * smax;
  for (int e = tid; e < np; e += nth) plane[e] = cmake<C>(0, 0);
  __syncthreads();
  for (int s = tid; s < S; s += nth) plane[jk[s]] = st[s];
  __syncthreads();
  lds_fft<1, TPT>(plane, ap3, n2, 1, n3, false, tw3, tid, nth);  // along k (rows)
  lds_fft<1, TPT>(plane, ap2, n3, n3, 1, true, tw2, tid, nth);   // along j (columns)
  const R* v = veff + (long long)i * np;
  for (int e = tid; e < np; e += nth) plane[e] = cscale(plane[e], v[e]);
  __syncthreads();
  lds_fft<-1, TPT>(plane, ap2, n3, n3, 1, true, tw2, tid, nth);
  lds_fft<-1, TPT>(plane, ap3, n2, 1, n3, false, tw3, tid, nth);
  for (int s = tid; s < S; s += nth) st[s] = plane[jk[s]];
}

// Forward transform along i, then out[g] = T_g psi[g] + (1/N) sticks at g.
template <int TPT, typename C, typename R>
__global__ void k_stick_forward(const C* __restrict__ sticks, const C* __restrict__ psi, C* __restrict__ out,
                                const int* __restrict__ stick_map, const R* __restrict__ kin, int n1, int smax,
                                int npw_pad, int nb, int k0, AxisPlan ap1, const C* __restrict__ tw1, int W,
                                R scale) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int kb = blockIdx.y;
  const int k = k0 + kb / nb;
  const int s0 = blockIdx.x * W;
  const C* in = sticks + (long long)kb * n1 * smax;
  for (int e = tid; e < n1 * W; e += nth) {
    const int i = e / W, w = e % W, s = s0 + w;
    lds[e] = (s < smax) ? in[i * smax + s] : cmake<C>(0, 0);
  }
  __syncthreads();
  lds_fft<-1, TPT>(lds, ap1, W, W, 1, true, tw1, tid, nth);
  const int* mp = stick_map + (long long)k * n1 * smax;
  const C* c = psi + (long long)kb * npw_pad;
  C* o = out + (long long)kb * npw_pad;
  const R* t = kin + (long long)k * npw_pad;
  for (int e = tid; e < n1 * W; e += nth) {
    const int i = e / W, w = e % W, s = s0 + w;
    if (s < smax) {
      const int g = mp[i * smax + s];
      if (g >= 0) {
        const C x = c[g];
        const C y = lds[e];
        o[g] = cmake<C>(t[g] * x.x + scale * y.x, t[g] * x.y + scale * y.y);
      }
    }
  }
}

// Density: per (i-plane, chunk of listed bands): sum w |u|^2 of the 2D inverse transform
// of the band's sticks (already transformed along i) into part[chunk][i][plane].
template <int TPT, int MAXPTS, typename C>
__global__ void k_density_plane(const C* __restrict__ sticks, const int* __restrict__ stick_jk,
                                const int* __restrict__ nsticks, const int* __restrict__ band_list,
                                const double* __restrict__ band_weight, int nlist, int chunk, int n1, int n2,
                                int n3, int smax, int nb, int k0, AxisPlan ap2, AxisPlan ap3,
                                const C* __restrict__ tw2, const C* __restrict__ tw3, double* __restrict__ part) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  C* plane = reinterpret_cast<C*>(smem_raw);
  const int tid = threadIdx.x, nth = blockDim.x;
  const int i = blockIdx.x;
  const int ch = blockIdx.y;
  const int np = n2 * n3;
  double acc[MAXPTS];
#pragma unroll
  for (int q = 0; q < MAXPTS; ++q) acc[q] = 0.0;
  const int l0 = ch * chunk;
  const int l1 = (l0 + chunk < nlist) ? l0 + chunk : nlist;
  for (int l = l0; l < l1; ++l) {
    const int kb = band_list[l];
    const double wgt = band_weight[l];
    const int k = k0 + kb / nb;
    const int S = nsticks[k];
    const int* jk = stick_jk + (long long)k * smax;
    const C* st = sticks + ((long long)kb * n1 + i) * smax;
    for (int e = tid; e < np; e += nth) plane[e] = cmake<C>(0, 0);
    __syncthreads();
    for (int s = tid; s < S; s += nth) plane[jk[s]] = st[s];
    __syncthreads();
    lds_fft<1, TPT>(plane, ap3, n2, 1, n3, false, tw3, tid, nth);
    lds_fft<1, TPT>(plane, ap2, n3, n3, 1, true, tw2, tid, nth);
#pragma unroll
    for (int q = 0; q < MAXPTS; ++q) {
      const int e = tid + q * nth;
      if (e < np) {
        const C u = plane[e];
        acc[q] += wgt * ((double)u.x * (double)u.x + (double)u.y * (double)u.y);
      }
    }
    __syncthreads();
  }
  double* out = part + ((long long)ch * n1 + i) * np;
#pragma unroll
  for (int q = 0; q < MAXPTS; ++q) {
    const int e = tid + q * nth;
    if (e < np) out[e] = acc[q];
  }
}

__global__ void k_density_reduce(const double* __restrict__ part, int nchunk, long long ngrid,
                                 double* __restrict__ rho, double beta, double scale) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < ngrid) {
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(long long)c * ngrid + p];
    rho[p] = (beta == 0.0 ? 0.0 : beta * rho[p]) + scale * s;
  }
}

// --------------------------------------------------------------------------- full-grid route
// For grids whose (n2 x n3) plane does not fit one workgroup's LDS at a precision (FP64 planes
// above ~100^2): whole grids through the library's scatter, 3-D transforms and gather, a chunk
// of band rows at a time in the caller's workspace.

// grids[b][e] *= v[e]
template <typename C, typename R>
__global__ void k_grid_vmul(C* __restrict__ grids, const R* __restrict__ v, long long ngrid, int nbc) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ngrid) return;
  const R ve = v[e];
  for (int b = 0; b < nbc; ++b) grids[(long long)b * ngrid + e] = cscale(grids[(long long)b * ngrid + e], ve);
}

// out[b][g] = kin[g] psi[b][g] + scale grid[b][flatmap[g]] for g < npw (padding not written)
template <typename C, typename R>
__global__ void k_gather_kinetic(const C* __restrict__ grid, const C* __restrict__ psi, C* __restrict__ out,
                                 const int* __restrict__ flatmap, const R* __restrict__ kin, int npw, int npw_pad,
                                 long long ngrid, R scale) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (g >= npw) return;
  const C x = psi[(long long)b * npw_pad + g];
  const C y = grid[(long long)b * ngrid + flatmap[g]];
  out[(long long)b * npw_pad + g] = cmake<C>(kin[g] * x.x + scale * y.x, kin[g] * x.y + scale * y.y);
}

// grids[c][flatmap_k[g]] = psi[row][g] for the listed rows list[c0 + c] (k = k0 + row / nb);
// grids zeroed beforehand
template <typename C>
__global__ void k_scatter_list(const C* __restrict__ psi, C* __restrict__ grids, const int* __restrict__ flatmap,
                               const int* __restrict__ npw, const int* __restrict__ list, int c0, int nb, int k0,
                               int npw_pad, long long ngrid) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  const int row = list[c0 + c];
  const int k = k0 + row / nb;
  if (g >= npw[k]) return;
  grids[(long long)c * ngrid + flatmap[(long long)k * npw_pad + g]] = psi[(long long)row * npw_pad + g];
}

// acc[e] += sum_c w[c0 + c] |grids[c][e]|^2 (FP64)
template <typename C>
__global__ void k_grid_density(const C* __restrict__ grids, const double* __restrict__ weight, int c0, int nbc,
                               long long ngrid, double* __restrict__ acc) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ngrid) return;
  double s = acc[e];
  for (int c = 0; c < nbc; ++c) {
    const C u = grids[(long long)c * ngrid + e];
    s += weight[c0 + c] * ((double)u.x * (double)u.x + (double)u.y * (double)u.y);
  }
  acc[e] = s;
}

// --------------------------------------------------------------------------- launchers
static int check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return 1;
  }
  return 0;
}

template <typename R>
static int hpsi_local_t(const Plan& p, const void* psi_, void* hpsi_, int nb, int k0, int nkk, const void* veff_,
                        void* sticks_, hipStream_t s, int stage) {
  using C = typename Cplx<R>::T;
  constexpr int prec = sizeof(R) == 8 ? kFP64 : kFP32;
  constexpr int TPT = sizeof(R) == 8 ? kTptF64 : kTptF32;
  const C* psi = static_cast<const C*>(psi_);
  C* hpsi = static_cast<C*>(hpsi_);
  C* sticks = static_cast<C*>(sticks_);
  const R* veff = static_cast<const R*>(veff_);
  const C* tw[3];
  for (int a = 0; a < 3; ++a) tw[a] = prec == kFP64 ? (const C*)p.d_tw64[a] : (const C*)p.d_tw32[a];
  const R* kin = prec == kFP64 ? (const R*)p.d_kin64 : (const R*)p.d_kin32;
  const int nkb = nb * nkk;
  const int W = p.stick_w[prec];
  const dim3 sgrid((p.smax + W - 1) / W, nkb);
  const size_t slds = (size_t)p.n1 * W * sizeof(C);
  const size_t plds = (size_t)p.n2 * p.n3 * sizeof(C);
  const R scale = (R)(1.0 / ((double)p.n1 * p.n2 * p.n3));
  // stage 0 runs all three kernels; 1, 2, 3 run one of them (for timing a single kernel)
  if (stage == 0 || stage == 1) {
    hipLaunchKernelGGL((k_stick_inverse<TPT, C>), sgrid, dim3(p.stick_threads[prec]), slds, s, psi, sticks,
                       p.d_stick_map, p.n1, p.smax, p.npw_pad, nb, k0, p.ap[0], tw[0], W);
    if (check(hipGetLastError(), "stick_inverse launch")) return 1;
  }
  if (stage == 0 || stage == 2) {
    hipLaunchKernelGGL((k_plane_vloc<TPT, C, R>), dim3(p.n1, nkb), dim3(p.plane_threads[prec]), plds, s, sticks,
                       p.d_stick_jk, p.d_nsticks, veff, p.n1, p.n2, p.n3, p.smax, nb, k0, p.ap[1], p.ap[2], tw[1],
                       tw[2]);
    if (check(hipGetLastError(), "plane_vloc launch")) return 1;
  }
  if (stage == 0 || stage == 3) {
    hipLaunchKernelGGL((k_stick_forward<TPT, C, R>), sgrid, dim3(p.stick_threads[prec]), slds, s, sticks, psi,
                       hpsi, p.d_stick_map, kin, p.n1, p.smax, p.npw_pad, nb, k0, p.ap[0], tw[0], W, scale);
    if (check(hipGetLastError(), "stick_forward launch")) return 1;
  }
  return 0;
}

int launch_hpsi_local(const Plan& p, int prec, const void* psi, void* hpsi, int nb, int k0, int nkk,
                      const void* veff, void* sticks, hipStream_t s, int stage) {
  if (p.fast) {
    if (launch_hpsi_fast(p, prec, psi, hpsi, nb, k0, nkk, veff, sticks, s, stage)) {
      set_error(std::string("hpsi (compile-time length) launch: ") + hipGetErrorString(hipGetLastError()));
      return 1;
    }
    return 0;
  }
  return prec == kFP64 ? hpsi_local_t<double>(p, psi, hpsi, nb, k0, nkk, veff, sticks, s, stage)
                       : hpsi_local_t<float>(p, psi, hpsi, nb, k0, nkk, veff, sticks, s, stage);
}

template <typename R>
constexpr int density_max_pts() {
  return sizeof(R) == 8 ? 8 : 16;
}

template <typename R>
static int density_t(const Plan& p, const void* psi_, int nb, int k0, int nkk, const int* list,
                     const double* weight, int nlist, int chunk, void* sticks_, double* part, double* rho,
                     double beta, double scale, hipStream_t s) {
  using C = typename Cplx<R>::T;
  constexpr int prec = sizeof(R) == 8 ? kFP64 : kFP32;
  constexpr int TPT = sizeof(R) == 8 ? kTptF64 : kTptF32;
  const C* psi = static_cast<const C*>(psi_);
  C* sticks = static_cast<C*>(sticks_);
  const C* tw[3];
  for (int a = 0; a < 3; ++a) tw[a] = prec == kFP64 ? (const C*)p.d_tw64[a] : (const C*)p.d_tw32[a];
  const int nkb = nb * nkk;
  const int W = p.stick_w[prec];
  const long long ngrid = (long long)p.n1 * p.n2 * p.n3;
  const int nchunk = nlist > 0 ? (nlist + chunk - 1) / chunk : 0;
  if (nlist > 0 && p.fast) {
    if (launch_density_fast(p, prec, psi_, nb, k0, nkk, list, weight, nlist, chunk, sticks_, part, s)) {
      set_error(std::string("density (compile-time length) launch: ") + hipGetErrorString(hipGetLastError()));
      return 1;
    }
  } else if (nlist > 0) {
    hipLaunchKernelGGL((k_stick_inverse<TPT, C>), dim3((p.smax + W - 1) / W, nkb), dim3(p.stick_threads[prec]),
                       (size_t)p.n1 * W * sizeof(C), s, psi, sticks, p.d_stick_map, p.n1, p.smax, p.npw_pad, nb,
                       k0, p.ap[0], tw[0], W);
    if (check(hipGetLastError(), "stick_inverse launch")) return 1;
    constexpr int MAXPTS = density_max_pts<R>();
    const int np = p.n2 * p.n3;
    int dth = p.plane_threads[prec];
    const int need = ((np + MAXPTS - 1) / MAXPTS + 63) / 64 * 64;
    if (need > dth) dth = need;
    hipLaunchKernelGGL((k_density_plane<TPT, MAXPTS, C>), dim3(p.n1, nchunk), dim3(dth),
                       (size_t)p.n2 * p.n3 * sizeof(C), s, sticks, p.d_stick_jk, p.d_nsticks, list, weight, nlist,
                       chunk, p.n1, p.n2, p.n3, p.smax, nb, k0, p.ap[1], p.ap[2], tw[1], tw[2], part);
    if (check(hipGetLastError(), "density_plane launch")) return 1;
  }
  hipLaunchKernelGGL(k_density_reduce, dim3((unsigned)((ngrid + 255) / 256)), dim3(256), 0, s, part, nchunk,
                     ngrid, rho, beta, scale);
  return check(hipGetLastError(), "density_reduce launch");
}

int launch_density(const Plan& p, int prec, const void* psi, int nb, int k0, int nkk, const int* list,
                   const double* weight, int nlist, int chunk, void* sticks, double* part, double* rho, double beta,
                   double scale, hipStream_t s) {
  const int maxpts = prec == kFP64 ? density_max_pts<double>() : density_max_pts<float>();
  if ((p.n2 * p.n3 + maxpts - 1) / maxpts > 1024) {
    set_error("density: plane too large for the register accumulator");
    return 1;
  }
  return prec == kFP64 ? density_t<double>(p, psi, nb, k0, nkk, list, weight, nlist, chunk, sticks, part, rho,
                                           beta, scale, s)
                       : density_t<float>(p, psi, nb, k0, nkk, list, weight, nlist, chunk, sticks, part, rho,
                                          beta, scale, s);
}

template <typename R>
static int scatter_t(const Plan& p, const void* psi, void* grid, int nb, int k, hipStream_t s) {
  using C = typename Cplx<R>::T;
  const long long ngrid = (long long)p.n1 * p.n2 * p.n3;
  const long long tot = ngrid * nb;
  hipLaunchKernelGGL((k_fill_zero<C>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, (C*)grid, tot);
  if (check(hipGetLastError(), "fill launch")) return 1;
  hipLaunchKernelGGL((k_scatter<C>), dim3((p.npw[k] + 255) / 256, nb), dim3(256), 0, s, (const C*)psi, (C*)grid,
                     p.d_flatmap + (long long)k * p.npw_pad, p.npw[k], p.npw_pad, ngrid);
  return check(hipGetLastError(), "scatter launch");
}

int launch_scatter(const Plan& p, int prec, const void* psi, void* grid, int nb, int k, hipStream_t s) {
  return prec == kFP64 ? scatter_t<double>(p, psi, grid, nb, k, s) : scatter_t<float>(p, psi, grid, nb, k, s);
}

template <typename R>
static int gather_t(const Plan& p, const void* grid, void* out, int nb, int k, double scale, hipStream_t s) {
  using C = typename Cplx<R>::T;
  const long long ngrid = (long long)p.n1 * p.n2 * p.n3;
  hipLaunchKernelGGL((k_gather<C>), dim3((p.npw_pad + 255) / 256, nb), dim3(256), 0, s, (const C*)grid, (C*)out,
                     p.d_flatmap + (long long)k * p.npw_pad, p.npw[k], p.npw_pad, ngrid, (R)scale);
  return check(hipGetLastError(), "gather launch");
}

int launch_gather(const Plan& p, int prec, const void* grid, void* out, int nb, int k, double scale,
                  hipStream_t s) {
  return prec == kFP64 ? gather_t<double>(p, grid, out, nb, k, scale, s)
                       : gather_t<float>(p, grid, out, nb, k, scale, s);
}

template <typename R>
static int fft3d_t(const FftPlan& fp, int batch, void* data_, int sign, double scale, hipStream_t s) {
  using C = typename Cplx<R>::T;
  constexpr int prec = sizeof(R) == 8 ? kFP64 : kFP32;
  constexpr int TPT = sizeof(R) == 8 ? kTptF64 : kTptF32;
  C* data = static_cast<C*>(data_);
  const int n1 = fp.n[0], n2 = fp.n[1], n3 = fp.n[2];
  const int* dims = fp.n;
  const AxisPlan* ap = fp.ap;
  // axis 2 (contiguous), then 1, then 0; scale folded into the last pass
  for (int step = 0; step < 3; ++step) {
    const int a = 2 - step;
    const long long inner = a == 2 ? 1 : (a == 1 ? (long long)n3 : (long long)n2 * n3);
    const long long nouter = (long long)batch * (a == 2 ? (long long)n1 * n2 : (a == 1 ? (long long)n1 : 1));
    const int W = inner == 1 ? 8 : 16;
    const int nth0 = threads_needed(ap[a], W, TPT);
    const int nth = ((nth0 + 63) / 64) * 64;
    if (nth > 1024) {
      set_error("fft3d: axis too long for one workgroup");
      return 1;
    }
    const C* tw = static_cast<const C*>(prec == kFP64 ? (const void*)fp.tw64[a] : (const void*)fp.tw32[a]);
    const long long nblk = inner == 1 ? (nouter + W - 1) / W : nouter * ((inner + W - 1) / W);
    const size_t lds = (size_t)dims[a] * W * sizeof(C);
    const R sc = (R)(step == 2 ? scale : 1.0);
    if (sign > 0)
      hipLaunchKernelGGL((k_axis_fft<1, TPT, C>), dim3((unsigned)nblk), dim3(nth), lds, s, data, ap[a], inner,
                         nouter, W, tw, sc);
    else
      hipLaunchKernelGGL((k_axis_fft<-1, TPT, C>), dim3((unsigned)nblk), dim3(nth), lds, s, data, ap[a], inner,
                         nouter, W, tw, sc);
    if (check(hipGetLastError(), "axis fft launch")) return 1;
  }
  return 0;
}

// The plan's own grid tables as a transform plan (no allocation).
static FftPlan grid_fft(const Plan& p) {
  FftPlan f;
  f.n[0] = p.n1;
  f.n[1] = p.n2;
  f.n[2] = p.n3;
  for (int a = 0; a < 3; ++a) {
    f.ap[a] = p.ap[a];
    f.tw64[a] = p.d_tw64[a];
    f.tw32[a] = p.d_tw32[a];
  }
  return f;
}

long long full_grid_rows(const Plan& p, int prec, size_t ws_bytes) {
  const long long gb = (long long)p.n1 * p.n2 * p.n3 * (prec == kFP64 ? 16 : 8);
  return gb > 0 ? (long long)(ws_bytes / (size_t)gb) : 0;
}

template <typename R>
static int hpsi_full_grid_t(const Plan& p, const void* psi_, void* hpsi_, int nb, int k0, int nkk, const void* veff_,
                            void* ws, size_t ws_bytes, hipStream_t s) {
  using C = typename Cplx<R>::T;
  constexpr int prec = sizeof(R) == 8 ? kFP64 : kFP32;
  const long long ngrid = (long long)p.n1 * p.n2 * p.n3;
  const long long cap = full_grid_rows(p, prec, ws_bytes);
  if (cap < 1) {
    set_error("hpsi (full-grid route): the workspace holds no whole grid");
    return 1;
  }
  const FftPlan f = grid_fft(p);
  const C* psi = static_cast<const C*>(psi_);
  C* hpsi = static_cast<C*>(hpsi_);
  C* grids = static_cast<C*>(ws);
  const R* veff = static_cast<const R*>(veff_);
  const R* kin = prec == kFP64 ? (const R*)p.d_kin64 : (const R*)p.d_kin32;
  for (int kk = 0; kk < nkk; ++kk) {
    const int k = k0 + kk;
    for (int b0 = 0; b0 < nb; b0 += (int)cap) {
      const int nbc = (int)(nb - b0 < cap ? nb - b0 : cap);
      const C* in = psi + ((long long)kk * nb + b0) * p.npw_pad;
      C* out = hpsi + ((long long)kk * nb + b0) * p.npw_pad;
      const long long tot = ngrid * nbc;
      hipLaunchKernelGGL((k_fill_zero<C>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, grids, tot);
      hipLaunchKernelGGL((k_scatter<C>), dim3((p.npw[k] + 255) / 256, nbc), dim3(256), 0, s, in, grids,
                         p.d_flatmap + (long long)k * p.npw_pad, p.npw[k], p.npw_pad, ngrid);
      if (check(hipGetLastError(), "full-grid scatter launch")) return 1;
      if (launch_fft3d(f, prec, nbc, grids, +1, 1.0, s)) return 1;
      hipLaunchKernelGGL((k_grid_vmul<C, R>), dim3((unsigned)((ngrid + 255) / 256)), dim3(256), 0, s, grids, veff,
                         ngrid, nbc);
      if (check(hipGetLastError(), "full-grid V_eff launch")) return 1;
      if (launch_fft3d(f, prec, nbc, grids, -1, 1.0, s)) return 1;
      hipLaunchKernelGGL((k_gather_kinetic<C, R>), dim3((p.npw[k] + 255) / 256, nbc), dim3(256), 0, s, grids, in,
                         out, p.d_flatmap + (long long)k * p.npw_pad, kin + (long long)k * p.npw_pad, p.npw[k],
                         p.npw_pad, ngrid, (R)(1.0 / (double)ngrid));
      if (check(hipGetLastError(), "full-grid gather launch")) return 1;
    }
  }
  return 0;
}

int launch_hpsi_full_grid(const Plan& p, int prec, const void* psi, void* hpsi, int nb, int k0, int nkk,
                          const void* veff, void* ws, size_t ws_bytes, hipStream_t s) {
  return prec == kFP64 ? hpsi_full_grid_t<double>(p, psi, hpsi, nb, k0, nkk, veff, ws, ws_bytes, s)
                       : hpsi_full_grid_t<float>(p, psi, hpsi, nb, k0, nkk, veff, ws, ws_bytes, s);
}

// rho = beta rho + scale sum_l w_l |u_l|^2 through whole grids; `acc` is one FP64 grid.
template <typename R>
static int density_full_grid_t(const Plan& p, const void* psi_, int nb, int k0, const int* list,
                               const double* weight, int nlist, void* ws, size_t ws_bytes, double* acc, double* rho,
                               double beta, double scale, hipStream_t s) {
  using C = typename Cplx<R>::T;
  constexpr int prec = sizeof(R) == 8 ? kFP64 : kFP32;
  const long long ngrid = (long long)p.n1 * p.n2 * p.n3;
  const unsigned gb = (unsigned)((ngrid + 255) / 256);
  if (nlist == 0) {  // rho = beta rho
    hipLaunchKernelGGL(k_density_reduce, dim3(gb), dim3(256), 0, s, (const double*)nullptr, 0, ngrid, rho, beta,
                       scale);
    return check(hipGetLastError(), "density reduce launch");
  }
  const long long cap = full_grid_rows(p, prec, ws_bytes);
  if (cap < 1) {
    set_error("density (full-grid route): the workspace holds no whole grid");
    return 1;
  }
  const FftPlan f = grid_fft(p);
  C* grids = static_cast<C*>(ws);
  hipLaunchKernelGGL(k_density_reduce, dim3(gb), dim3(256), 0, s, (const double*)nullptr, 0, ngrid, acc, 0.0, 0.0);
  if (check(hipGetLastError(), "full-grid density clear launch")) return 1;
  for (int c0 = 0; c0 < nlist; c0 += (int)cap) {
    const int nbc = (int)(nlist - c0 < cap ? nlist - c0 : cap);
    const long long tot = ngrid * nbc;
    hipLaunchKernelGGL((k_fill_zero<C>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, grids, tot);
    hipLaunchKernelGGL((k_scatter_list<C>), dim3((p.npw_pad + 255) / 256, nbc), dim3(256), 0, s,
                       static_cast<const C*>(psi_), grids, p.d_flatmap, p.d_npw, list, c0, nb, k0, p.npw_pad, ngrid);
    if (check(hipGetLastError(), "full-grid scatter launch")) return 1;
    if (launch_fft3d(f, prec, nbc, grids, +1, 1.0, s)) return 1;
    hipLaunchKernelGGL((k_grid_density<C>), dim3(gb), dim3(256), 0, s, grids, weight, c0, nbc, ngrid, acc);
    if (check(hipGetLastError(), "full-grid density launch")) return 1;
  }
  hipLaunchKernelGGL(k_density_reduce, dim3(gb), dim3(256), 0, s, acc, 1, ngrid, rho, beta, scale);
  return check(hipGetLastError(), "full-grid density reduce launch");
}

int launch_density_full_grid(const Plan& p, int prec, const void* psi, int nb, int k0, const int* list,
                             const double* weight, int nlist, void* ws, size_t ws_bytes, double* acc, double* rho,
                             double beta, double scale, hipStream_t s) {
  return prec == kFP64
             ? density_full_grid_t<double>(p, psi, nb, k0, list, weight, nlist, ws, ws_bytes, acc, rho, beta, scale, s)
             : density_full_grid_t<float>(p, psi, nb, k0, list, weight, nlist, ws, ws_bytes, acc, rho, beta, scale, s);
}

int launch_fft3d(const FftPlan& fp, int prec, int batch, void* data, int sign, double scale, hipStream_t s) {
  return prec == kFP64 ? fft3d_t<double>(fp, batch, data, sign, scale, s)
                       : fft3d_t<float>(fp, batch, data, sign, scale, s);
}

}  // namespace amt
