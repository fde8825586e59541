// Davidson's elementwise/reduction steps and the Kleinman-Bylander D_ij contraction.
//
// residual:      R = HV - eps V, ||R|| per band (FP64 accumulation)
// precondition:  T_band = sum |V|^2 T_G; x = T_G / max(T_band, 1e-8);
//                K = P / (P + 16 x^4), P = 27 + x (18 + x (12 + 8 x));
//                c = R K / max(||R K||, 1e-300)      (Teter-Payne-Allan, one pass per band)
// dij:           (P D)[b][i] = sum_j P[b][j] D[j][i] over the atom block holding i.
#include "kernels.hpp"

namespace amt {

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < NT / 64; ++w) t += red[w];
    red[NT / 64] = t;
  }
  __syncthreads();
  t = red[NT / 64];
  __syncthreads();
  return t;
}

// One workgroup of kRowThreads per band row (rows laid out [row][npw_pad]): with a few hundred
// rows of 10^4-10^5 plane waves each, 16 waves per row keep enough loads in flight per CU.
constexpr int kRowThreads = 1024;

template <typename R>
__global__ __launch_bounds__(kRowThreads) void k_residual(const typename Cplx<R>::T* __restrict__ hv,
                                                  const typename Cplx<R>::T* __restrict__ v,
                                                  const double* __restrict__ eps, typename Cplx<R>::T* __restrict__ res,
                                                  double* __restrict__ norm, int npw) {
  using C = typename Cplx<R>::T;
  __shared__ double red[kRowThreads / 64 + 1];
  const long long off = (long long)blockIdx.x * npw;
  const R e = (R)eps[blockIdx.x];
  double acc = 0.0;
  for (int g = threadIdx.x; g < npw; g += kRowThreads) {
    const C a = hv[off + g], b = v[off + g];
    const C r = cmake<C>(a.x - e * b.x, a.y - e * b.y);
    res[off + g] = r;
    acc += (double)r.x * r.x + (double)r.y * r.y;
  }
  acc = block_sum<kRowThreads>(acc, red);
  if (threadIdx.x == 0) norm[blockIdx.x] = sqrt(acc);
}

// rows[a] selects the source band of output row a.
template <typename R>
__global__ __launch_bounds__(kRowThreads) void k_precondition(const typename Cplx<R>::T* __restrict__ res,
                                                      const typename Cplx<R>::T* __restrict__ v,
                                                      const R* __restrict__ kin, const int* __restrict__ rows,
                                                      typename Cplx<R>::T* __restrict__ out, int npw) {
  using C = typename Cplx<R>::T;
  __shared__ double red[kRowThreads / 64 + 1];
  const int src = rows[blockIdx.x];
  const long long si = (long long)src * npw, oi = (long long)blockIdx.x * npw;
  double tb = 0.0;
  for (int g = threadIdx.x; g < npw; g += kRowThreads) {
    const C b = v[si + g];
    tb += ((double)b.x * b.x + (double)b.y * b.y) * (double)kin[g];
  }
  tb = block_sum<kRowThreads>(tb, red);
  const double ref = tb > 1e-8 ? tb : 1e-8;
  const double rinv = 1.0 / ref;
  double nn = 0.0;
  for (int g = threadIdx.x; g < npw; g += kRowThreads) {
    const double x = (double)kin[g] * rinv;
    const double p = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x));
    const double f = p / (p + 16.0 * x * x * x * x);
    const C r = res[si + g];
    const C c = cmake<C>((R)(r.x * f), (R)(r.y * f));
    out[oi + g] = c;
    nn += (double)c.x * c.x + (double)c.y * c.y;
  }
  nn = block_sum<kRowThreads>(nn, red);
  const double inv = 1.0 / (sqrt(nn) > 1e-300 ? sqrt(nn) : 1e-300);
  for (int g = threadIdx.x; g < npw; g += kRowThreads) {
    const C c = out[oi + g];
    out[oi + g] = cmake<C>((R)(c.x * inv), (R)(c.y * inv));
  }
}

// (P D): P [nrow][nproj] complex, D [nproj][nproj] real (block-diagonal over atoms);
// blk_start/blk_len give the atom block of each projector.
template <typename R>
__global__ void k_dij(const typename Cplx<R>::T* __restrict__ P, const R* __restrict__ D,
                      const int* __restrict__ blk_start, const int* __restrict__ blk_len,
                      typename Cplx<R>::T* __restrict__ out, int nrow, int nproj, long long sP, long long sD,
                      long long sO) {
  using C = typename Cplx<R>::T;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int batch = blockIdx.y;
  if (e >= (long long)nrow * nproj) return;
  const int b = (int)(e / nproj), i = (int)(e % nproj);
  const C* p = P + batch * sP + (long long)b * nproj;
  const R* d = D + batch * sD;
  const int j0 = blk_start[i], j1 = j0 + blk_len[i];
  R sr = 0, si = 0;
  for (int j = j0; j < j1; ++j) {
    const R dj = d[(long long)j * nproj + i];
    sr += p[j].x * dj;
    si += p[j].y * dj;
  }
  out[batch * sO + (long long)b * nproj + i] = cmake<C>(sr, si);
}

int launch_residual(int prec, const void* hv, const void* v, const double* eps, void* res, double* norm, int nrow,
                    int npw, hipStream_t s) {
  if (nrow <= 0) return 0;
  if (prec == kFP64)
    hipLaunchKernelGGL(k_residual<double>, dim3(nrow), dim3(kRowThreads), 0, s, (const double2*)hv, (const double2*)v, eps,
                       (double2*)res, norm, npw);
  else
    hipLaunchKernelGGL(k_residual<float>, dim3(nrow), dim3(kRowThreads), 0, s, (const float2*)hv, (const float2*)v, eps,
                       (float2*)res, norm, npw);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("residual launch: ") + hipGetErrorString(e));
    return 1;
  }
  return 0;
}

int launch_precondition(int prec, const void* res, const void* v, const void* kin, const int* rows, void* out,
                        int nout, int npw, hipStream_t s) {
  if (nout <= 0) return 0;
  if (prec == kFP64)
    hipLaunchKernelGGL(k_precondition<double>, dim3(nout), dim3(kRowThreads), 0, s, (const double2*)res, (const double2*)v,
                       (const double*)kin, rows, (double2*)out, npw);
  else
    hipLaunchKernelGGL(k_precondition<float>, dim3(nout), dim3(kRowThreads), 0, s, (const float2*)res, (const float2*)v,
                       (const float*)kin, rows, (float2*)out, npw);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("precondition launch: ") + hipGetErrorString(e));
    return 1;
  }
  return 0;
}

int launch_dij(int prec, const void* P, const void* D, const int* blk_start, const int* blk_len, void* out,
               int nrow, int nproj, long long sP, long long sD, long long sO, int batch, hipStream_t s) {
  if (nrow <= 0 || nproj <= 0 || batch <= 0) return 0;
  const long long tot = (long long)nrow * nproj;
  const dim3 grid((unsigned)((tot + 255) / 256), batch);
  if (prec == kFP64)
    hipLaunchKernelGGL(k_dij<double>, grid, dim3(256), 0, s, (const double2*)P, (const double*)D, blk_start,
                       blk_len, (double2*)out, nrow, nproj, sP, sD, sO);
  else
    hipLaunchKernelGGL(k_dij<float>, grid, dim3(256), 0, s, (const float2*)P, (const float*)D, blk_start, blk_len,
                       (float2*)out, nrow, nproj, sP, sD, sO);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("dij launch: ") + hipGetErrorString(e));
    return 1;
  }
  return 0;
}

}  // namespace amt
