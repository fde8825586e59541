// Local H.psi and density kernels specialised to a compile-time grid length N
// (cubic grids n1 = n2 = n3 = N, N in AMT_FAST_LENGTHS), built on the two-level
// register FFT of fft_reg.hpp. Same data layout and results as the generic kernels of
// fft.hip, which remain the path for every other grid. Two further savings over the
// generic plane kernel: the transforms along k touch only the rows j that hold sticks
// (about half of them: the others are zero on the way in and discarded on the way out),
// and every index division is by a compile-time constant.
#include "fft_reg.hpp"
#include "kernels.hpp"

namespace amt {

#define AMT_FAST_LENGTHS(X) \
  X(20) X(24) X(28) X(30) X(32) X(35) X(36) X(40) X(42) X(44) X(45) X(48) X(54) X(56) X(60) X(64) X(70) X(72) \
  X(80) X(84) X(90) X(96) X(108) X(112) X(120) X(128)

bool fast_length_supported(int n) {
  switch (n) {
#define AMT_CASE(N) case N:
    AMT_FAST_LENGTHS(AMT_CASE)
#undef AMT_CASE
    return true;
    default:
      return false;
  }
}

constexpr int ceil_div(int a, int b) { return (a + b - 1) / b; }

template <typename C>
constexpr int stick_width() {
  return sizeof(C) == 16 ? 32 : 64;
}

// ----------------------------------------------------------------------- kernels
template <int N, typename C>
__global__ __launch_bounds__(256) void k_stick_inverse_n(const C* __restrict__ psi, C* __restrict__ sticks,
                                                         const int* __restrict__ stick_map, int smax, int npw_pad,
                                                         int nb, int k0, const C* __restrict__ tw) {
  constexpr int W = stick_width<C>();
  constexpr int T2 = ceil_div(W * Split<N>::N1, 256);
  __shared__ C tile[N * W];
  const int tid = threadIdx.x;
  const int kb = blockIdx.y;
  const int k = k0 + kb / nb;
  const int s0 = blockIdx.x * W;
  const C* c = psi + (long long)kb * npw_pad;
  const int* mp = stick_map + (long long)k * N * smax;
  for (int e = tid; e < N * W; e += 256) {
    const int i = e / W, w = e % W, s = s0 + w;
    C v = cmake<C>(0, 0);
    if (s < smax) {
      const int g = mp[i * smax + s];
      if (g >= 0) v = c[g];
    }
    tile[e] = v;
  }
  __syncthreads();
  fft2<N, 1, T2, false>(tile, W, 1, W, nullptr, tw, tid, 256);
  C* out = sticks + (long long)kb * N * smax;
  for (int e = tid; e < N * W; e += 256) {
    const int i = e / W, w = e % W, s = s0 + w;
    if (s < smax) out[i * smax + s] = tile[e];
  }
}

template <int N, typename C, typename R>
__global__ __launch_bounds__(256) void k_stick_forward_n(const C* __restrict__ sticks, const C* __restrict__ psi,
                                                         C* __restrict__ out, const int* __restrict__ stick_map,
                                                         const R* __restrict__ kin, int smax, int npw_pad, int nb,
                                                         int k0, const C* __restrict__ tw, R scale) {
  constexpr int W = stick_width<C>();
  constexpr int T2 = ceil_div(W * Split<N>::N1, 256);
  __shared__ C tile[N * W];
  const int tid = threadIdx.x;
  const int kb = blockIdx.y;
  const int k = k0 + kb / nb;
  const int s0 = blockIdx.x * W;
  const C* in = sticks + (long long)kb * N * smax;
  for (int e = tid; e < N * W; e += 256) {
    const int i = e / W, w = e % W, s = s0 + w;
    tile[e] = (s < smax) ? in[i * smax + s] : cmake<C>(0, 0);
  }
  __syncthreads();
  fft2<N, -1, T2, false>(tile, W, 1, W, nullptr, tw, tid, 256);
  const int* mp = stick_map + (long long)k * N * smax;
  const C* cc = psi + (long long)kb * npw_pad;
  C* o = out + (long long)kb * npw_pad;
  const R* t = kin + (long long)k * npw_pad;
  for (int e = tid; e < N * W; e += 256) {
    const int i = e / W, w = e % W, s = s0 + w;
    if (s < smax) {
      const int g = mp[i * smax + s];
      if (g >= 0) {
        const C x = cc[g];
        const C y = tile[e];
        o[g] = cmake<C>(t[g] * x.x + scale * y.x, t[g] * x.y + scale * y.y);
      }
    }
  }
}

template <int N, typename C, typename R>
__global__ __launch_bounds__(256) void k_plane_vloc_n(C* __restrict__ sticks, const int* __restrict__ stick_jk,
                                                      const int* __restrict__ nsticks, const int* __restrict__ rows,
                                                      const int* __restrict__ nrows, const R* __restrict__ veff,
                                                      int smax, int nb, int k0, const C* __restrict__ tw) {
  constexpr int NP = N * N;
  constexpr int T2 = ceil_div(N * Split<N>::N1, 256);
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  C* plane = reinterpret_cast<C*>(smem_raw);
  __shared__ int rl[(N + 3) / 4 * 4];  // a multiple of 16 bytes: the dynamic plane after it stays 16-byte aligned
  const int tid = threadIdx.x;
  const int i = blockIdx.x;
  const int kb = blockIdx.y;
  const int k = k0 + kb / nb;
  const int S = nsticks[k];
  const int nr = nrows[k];
  const int* jk = stick_jk + (long long)k * smax;
  C* st = sticks + ((long long)kb * N + i) * smax;
  for (int e = tid; e < NP; e += 256) plane[e] = cmake<C>(0, 0);
  for (int r = tid; r < nr; r += 256) rl[r] = rows[(long long)k * N + r];
  __syncthreads();
  for (int s = tid; s < S; s += 256) plane[jk[s]] = st[s];
  __syncthreads();
  fft2<N, 1, T2, true>(plane, 1, N, nr, rl, tw, tid, 256);       // along k, rows with sticks
  fft2<N, 1, T2, false>(plane, N, 1, N, nullptr, tw, tid, 256);  // along j, every column
  const R* v = veff + (long long)i * NP;
  for (int e = tid; e < NP; e += 256) plane[e] = cscale(plane[e], v[e]);
  __syncthreads();
  fft2<N, -1, T2, false>(plane, N, 1, N, nullptr, tw, tid, 256);
  fft2<N, -1, T2, true>(plane, 1, N, nr, rl, tw, tid, 256);
  for (int s = tid; s < S; s += 256) st[s] = plane[jk[s]];
}

template <int N, int MAXPTS, typename C>
__global__ __launch_bounds__(512) void k_density_plane_n(const C* __restrict__ sticks,
                                                         const int* __restrict__ stick_jk,
                                                         const int* __restrict__ nsticks,
                                                         const int* __restrict__ rows, const int* __restrict__ nrows,
                                                         const int* __restrict__ band_list,
                                                         const double* __restrict__ band_weight, int nlist,
                                                         int chunk, int smax, int nb, int k0,
                                                         const C* __restrict__ tw, double* __restrict__ part) {
  constexpr int NP = N * N;
  constexpr int T2 = ceil_div(N * Split<N>::N1, 512);
  static_assert(MAXPTS * 512 >= NP, "density accumulator too small");
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  C* plane = reinterpret_cast<C*>(smem_raw);
  __shared__ int rl[(N + 3) / 4 * 4];  // a multiple of 16 bytes: the dynamic plane after it stays 16-byte aligned
  const int tid = threadIdx.x;
  const int i = blockIdx.x;
  const int ch = blockIdx.y;
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
    const int nr = nrows[k];
    const int* jk = stick_jk + (long long)k * smax;
    const C* st = sticks + ((long long)kb * N + i) * smax;
    for (int e = tid; e < NP; e += 512) plane[e] = cmake<C>(0, 0);
    for (int r = tid; r < nr; r += 512) rl[r] = rows[(long long)k * N + r];
    __syncthreads();
    for (int s = tid; s < S; s += 512) plane[jk[s]] = st[s];
    __syncthreads();
    fft2<N, 1, T2, true>(plane, 1, N, nr, rl, tw, tid, 512);
    fft2<N, 1, T2, false>(plane, N, 1, N, nullptr, tw, tid, 512);
#pragma unroll
    for (int q = 0; q < MAXPTS; ++q) {
      const int e = tid + q * 512;
      if (e < NP) {
        const C u = plane[e];
        acc[q] += wgt * ((double)u.x * (double)u.x + (double)u.y * (double)u.y);
      }
    }
    __syncthreads();
  }
  double* out = part + ((long long)ch * N + i) * NP;
#pragma unroll
  for (int q = 0; q < MAXPTS; ++q) {
    const int e = tid + q * 512;
    if (e < NP) out[e] = acc[q];
  }
}

// ----------------------------------------------------------------------- launchers
template <int N, typename R>
static int hpsi_fast_t(const Plan& p, const void* psi_, void* hpsi_, int nb, int k0, int nkk, const void* veff_,
                       void* sticks_, hipStream_t s, int stage) {
  using C = typename Cplx<R>::T;
  constexpr int prec = sizeof(R) == 8 ? kFP64 : kFP32;
  constexpr int W = stick_width<C>();
  const C* psi = static_cast<const C*>(psi_);
  C* hpsi = static_cast<C*>(hpsi_);
  C* sticks = static_cast<C*>(sticks_);
  const R* veff = static_cast<const R*>(veff_);
  const C* tw = prec == kFP64 ? (const C*)p.d_tw64[0] : (const C*)p.d_tw32[0];
  const R* kin = prec == kFP64 ? (const R*)p.d_kin64 : (const R*)p.d_kin32;
  const int nkb = nb * nkk;
  const dim3 sgrid((p.smax + W - 1) / W, nkb);
  const size_t plds = (size_t)N * N * sizeof(C);
  if (stage == 0 || stage == 1) {
    hipLaunchKernelGGL((k_stick_inverse_n<N, C>), sgrid, dim3(256), 0, s, psi, sticks, p.d_stick_map, p.smax,
                       p.npw_pad, nb, k0, tw);
    if (hipGetLastError() != hipSuccess) return 1;
  }
  if (stage == 0 || stage == 2) {
    hipLaunchKernelGGL((k_plane_vloc_n<N, C, R>), dim3(N, nkb), dim3(256), plds, s, sticks, p.d_stick_jk,
                       p.d_nsticks, p.d_rows, p.d_nrows, veff, p.smax, nb, k0, tw);
    if (hipGetLastError() != hipSuccess) return 1;
  }
  if (stage == 0 || stage == 3) {
    hipLaunchKernelGGL((k_stick_forward_n<N, C, R>), sgrid, dim3(256), 0, s, sticks, psi, hpsi, p.d_stick_map, kin,
                       p.smax, p.npw_pad, nb, k0, tw, (R)(1.0 / ((double)N * N * N)));
    if (hipGetLastError() != hipSuccess) return 1;
  }
  return 0;
}

template <int N, typename R>
static int density_fast_t(const Plan& p, const void* psi_, int nb, int k0, int nkk, const int* list,
                          const double* weight, int nlist, int chunk, void* sticks_, double* part, hipStream_t s) {
  using C = typename Cplx<R>::T;
  constexpr int prec = sizeof(R) == 8 ? kFP64 : kFP32;
  constexpr int W = stick_width<C>();
  constexpr int MAXPTS = ceil_div(N * N, 512);
  const C* tw = prec == kFP64 ? (const C*)p.d_tw64[0] : (const C*)p.d_tw32[0];
  const int nkb = nb * nkk;
  const int nchunk = (nlist + chunk - 1) / chunk;
  hipLaunchKernelGGL((k_stick_inverse_n<N, C>), dim3((p.smax + W - 1) / W, nkb), dim3(256), 0, s,
                     static_cast<const C*>(psi_), static_cast<C*>(sticks_), p.d_stick_map, p.smax, p.npw_pad, nb,
                     k0, tw);
  if (hipGetLastError() != hipSuccess) return 1;
  hipLaunchKernelGGL((k_density_plane_n<N, MAXPTS, C>), dim3(N, nchunk), dim3(512), (size_t)N * N * sizeof(C), s,
                     static_cast<const C*>(sticks_), p.d_stick_jk, p.d_nsticks, p.d_rows, p.d_nrows, list, weight,
                     nlist, chunk, p.smax, nb, k0, tw, part);
  return hipGetLastError() != hipSuccess ? 1 : 0;
}

int launch_hpsi_fast(const Plan& p, int prec, const void* psi, void* hpsi, int nb, int k0, int nkk,
                     const void* veff, void* sticks, hipStream_t s, int stage) {
  switch (p.n1) {
#define AMT_CASE(N)                                                                                   \
  case N:                                                                                             \
    return prec == kFP64 ? hpsi_fast_t<N, double>(p, psi, hpsi, nb, k0, nkk, veff, sticks, s, stage) \
                         : hpsi_fast_t<N, float>(p, psi, hpsi, nb, k0, nkk, veff, sticks, s, stage);
    AMT_FAST_LENGTHS(AMT_CASE)
#undef AMT_CASE
    default:
      return 1;
  }
}

int launch_density_fast(const Plan& p, int prec, const void* psi, int nb, int k0, int nkk, const int* list,
                        const double* weight, int nlist, int chunk, void* sticks, double* part, hipStream_t s) {
  switch (p.n1) {
#define AMT_CASE(N)                                                                                             \
  case N:                                                                                                       \
    return prec == kFP64                                                                                        \
               ? density_fast_t<N, double>(p, psi, nb, k0, nkk, list, weight, nlist, chunk, sticks, part, s)    \
               : density_fast_t<N, float>(p, psi, nb, k0, nkk, list, weight, nlist, chunk, sticks, part, s);
    AMT_FAST_LENGTHS(AMT_CASE)
#undef AMT_CASE
    default:
      return 1;
  }
}

}  // namespace amt
