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
// Stick kernels run on a 1-D grid of 8 * nch * ceil(nkb / 8) workgroups. Workgroup w goes to
// XCD w % 8, and all nch stick chunks of band row kb get w % 8 = kb % 8, so every chunk of a
// band row runs on one XCD and its scattered sphere-order reads and writes of that row meet in
// one L2 instead of being fetched once per XCD.
__device__ __forceinline__ bool stick_block(int nch, int nkb, int& chunk, int& kb) {
  const int w = blockIdx.x, slot = w >> 3;
  chunk = slot % nch;
  kb = (slot / nch) * 8 + (w & 7);
  return kb < nkb;
}
inline unsigned stick_grid(int nch, int nkb) { return 8u * (unsigned)nch * (unsigned)((nkb + 7) / 8); }

template <int N, typename C, int W>
__global__ __launch_bounds__(256) void k_stick_inverse_n(const C* __restrict__ psi, C* __restrict__ sticks,
                                                         const int* __restrict__ stick_map, int smax, int npw_pad,
                                                         int nb, int k0, const C* __restrict__ tw, int nch,
                                                         int nkb) {
  constexpr int T2 = ceil_div(W * Split<N>::N1, 256);
  __shared__ C tile[N * W];
  const int tid = threadIdx.x;
  int chunk, kb;
  if (!stick_block(nch, nkb, chunk, kb)) return;
  const int k = k0 + kb / nb;
  const int s0 = chunk * W;
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

template <int N, typename C, typename R, int W>
__global__ __launch_bounds__(256) void k_stick_forward_n(const C* __restrict__ sticks, const C* __restrict__ psi,
                                                         C* __restrict__ out, const int* __restrict__ stick_map,
                                                         const R* __restrict__ kin, int smax, int npw_pad, int nb,
                                                         int k0, const C* __restrict__ tw, R scale, int nch, int nkb) {
  constexpr int T2 = ceil_div(W * Split<N>::N1, 256);
  __shared__ C tile[N * W];
  const int tid = threadIdx.x;
  int chunk, kb;
  if (!stick_block(nch, nkb, chunk, kb)) return;
  const int k = k0 + kb / nb;
  const int s0 = chunk * W;
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

// ----------------------------------------------------------------------- plane pieces
// Bit kk of a row's 128-bit stick mask, and the number of set bits below it.
__device__ __forceinline__ bool mask_bit(unsigned long long lo, unsigned long long hi, int kk) {
  return kk < 64 ? ((lo >> kk) & 1ull) : ((hi >> (kk - 64)) & 1ull);
}
__device__ __forceinline__ int mask_below(unsigned long long lo, unsigned long long hi, int kk) {
  return kk < 64 ? __popcll(lo & ((1ull << kk) - 1ull)) : __popcll(lo) + __popcll(hi & ((1ull << (kk - 64)) - 1ull));
}

// u[k1] *= w_N^{SIGN j2 k1} (tw = exp(-2 pi i m / N); j2 k1 < N for the four-step indices).
template <int N, int NA, int SIGN, typename C>
__device__ __forceinline__ void twiddle_row(C (&u)[NA], int j2, const C* __restrict__ tw) {
#pragma unroll
  for (int k1 = 1; k1 < NA; ++k1) {
    C w = tw[j2 * k1];
    if (SIGN > 0) w.y = -w.y;
    u[k1] = cmul(u[k1], w);
  }
}

// Local part of H.psi for one i-plane of one band row: the inverse 2D transform of the
// plane's sticks, times V_eff, and the forward 2D transform back onto the sticks. The length
// N = N1 N2 transforms are four-step (fft_reg.hpp) with these savings over a plain pair of
// 2D transforms:
//  * the sticks are read straight from the stick workspace by the first phase along k (the
//    row masks give each (j, k) its stick index), so the plane is never zeroed or scattered;
//  * along k only the rows j that hold sticks are transformed, and the first phase along j
//    reads the other rows as zeros;
//  * the second inverse phase along j, the V_eff multiply and the first forward phase along j
//    are one register pass: the forward transform uses the transposed split (N2 x N1), whose
//    first phase consumes exactly the N2 outputs one inverse task holds;
//  * the forward phase along j writes only rows that hold sticks, and the last phase along k
//    writes its outputs straight back to the sticks.
template <int N, typename C, typename R, int NTH>
__global__ __launch_bounds__(NTH) void k_plane_vloc_n(C* __restrict__ sticks, const int* __restrict__ rows,
                                                      const int* __restrict__ nrows,
                                                      const unsigned long long* __restrict__ rowmask,
                                                      const int* __restrict__ rowstart,
                                                      const unsigned long long* __restrict__ livemask,
                                                      const R* __restrict__ veff, int smax, int nb, int k0,
                                                      const C* __restrict__ tw) {
  constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  constexpr int TR = ceil_div(N * N1, NTH);  // rounds of the (row or column, k1) tasks
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  C* plane = reinterpret_cast<C*>(smem_raw);
  __shared__ int rl[(N + 3) / 4 * 4];  // sizes in multiples of 16 bytes keep the dynamic plane 16-byte aligned
  __shared__ int rs[(N + 3) / 4 * 4];
  __shared__ unsigned long long rm[2 * N];
  const int tid = threadIdx.x;
  const int i = blockIdx.x;
  const int kb = blockIdx.y;
  const int k = k0 + kb / nb;
  const int nr = nrows[k];
  const unsigned long long live_lo = livemask[2 * k], live_hi = livemask[2 * k + 1];
  C* st = sticks + ((long long)kb * N + i) * smax;
  for (int r = tid; r < nr; r += NTH) {
    const int j = rows[(long long)k * N + r];
    rl[r] = j;
    rs[r] = rowstart[(long long)k * N + j];
    rm[2 * r] = rowmask[((long long)k * N + j) * 2];
    rm[2 * r + 1] = rowmask[((long long)k * N + j) * 2 + 1];
  }
  __syncthreads();

  // 1. inverse along k, phase 1, rows with sticks: read the sticks, DFT_N1, twiddle
  for (int t = tid; t < nr * N2; t += NTH) {
    const int s = t / N2, j2 = t % N2;
    const unsigned long long lo = rm[2 * s], hi = rm[2 * s + 1];
    const C* src = st + rs[s];
    C u[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      const int kk = j2 + N2 * n1;
      u[n1] = mask_bit(lo, hi, kk) ? src[mask_below(lo, hi, kk)] : cmake<C>(0, 0);
    }
    small_dft<N1, 1>(u);
    twiddle_row<N, N1, 1>(u, j2, tw);
    C* row = plane + rl[s] * N;
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) row[j2 + N2 * k1] = u[k1];
  }
  __syncthreads();
  // 2. inverse along k, phase 2 (in place: all reads, barrier, all writes)
  {
    C v[TR][N2];
#pragma unroll
    for (int q = 0; q < TR; ++q) {
      const int t = tid + q * NTH;
      if (t < nr * N1) {
        const C* row = plane + rl[t / N1] * N + N2 * (t % N1);
#pragma unroll
        for (int j2 = 0; j2 < N2; ++j2) v[q][j2] = row[j2];
        small_dft<N2, 1>(v[q]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TR; ++q) {
      const int t = tid + q * NTH;
      if (t < nr * N1) {
        C* row = plane + rl[t / N1] * N + t % N1;
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) row[N1 * k2] = v[q][k2];
      }
    }
  }
  __syncthreads();
  // 3. inverse along j, phase 1, every column; rows without sticks read as zeros
  for (int t = tid; t < N * N2; t += NTH) {
    const int c = t % N, j2 = t / N;
    C u[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
      const int j = j2 + N2 * n1;
      u[n1] = mask_bit(live_lo, live_hi, j) ? plane[j * N + c] : cmake<C>(0, 0);
    }
    small_dft<N1, 1>(u);
    twiddle_row<N, N1, 1>(u, j2, tw);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) plane[(j2 + N2 * k1) * N + c] = u[k1];
  }
  __syncthreads();
  // 4. inverse along j phase 2, times V_eff, forward along j phase 1 of the transposed split
  //    (N = N2 x N1): task (c, k1) holds rows k1 + N1 m, m < N2, from the DFT to the twiddle.
  {
    const R* v = veff + (long long)i * N * N;
    C x[TR][N2];
#pragma unroll
    for (int q = 0; q < TR; ++q) {
      const int t = tid + q * NTH;
      if (t < N * N1) {
        const int c = t % N, k1 = t / N;
#pragma unroll
        for (int j2 = 0; j2 < N2; ++j2) x[q][j2] = plane[(N2 * k1 + j2) * N + c];
        small_dft<N2, 1>(x[q]);
#pragma unroll
        for (int m = 0; m < N2; ++m) x[q][m] = cscale(x[q][m], v[(k1 + N1 * m) * N + c]);
        small_dft<N2, -1>(x[q]);
        twiddle_row<N, N2, -1>(x[q], k1, tw);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TR; ++q) {
      const int t = tid + q * NTH;
      if (t < N * N1) {
        const int c = t % N, k1 = t / N;
#pragma unroll
        for (int m = 0; m < N2; ++m) plane[(k1 + N1 * m) * N + c] = x[q][m];
      }
    }
  }
  __syncthreads();
  // 5. forward along j, phase 2 of the transposed split: task (c, m) reads rows N1 m + k1,
  //    k1 < N1, and produces rows m + N2 p, p < N1; only rows with sticks are kept
  {
    constexpr int TC = ceil_div(N * N2, NTH);
    C y[TC][N1];
#pragma unroll
    for (int q = 0; q < TC; ++q) {
      const int t = tid + q * NTH;
      if (t < N * N2) {
        const int c = t % N, m = t / N;
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) y[q][k1] = plane[(N1 * m + k1) * N + c];
        small_dft<N1, -1>(y[q]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < TC; ++q) {
      const int t = tid + q * NTH;
      if (t < N * N2) {
        const int c = t % N, m = t / N;
#pragma unroll
        for (int p = 0; p < N1; ++p) {
          const int j = m + N2 * p;
          if (mask_bit(live_lo, live_hi, j)) plane[j * N + c] = y[q][p];
        }
      }
    }
  }
  __syncthreads();
  // 6. forward along k, phase 1, rows with sticks
  for (int t = tid; t < nr * N2; t += NTH) {
    const int s = t / N2, j2 = t % N2;
    C* row = plane + rl[s] * N;
    C u[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) u[n1] = row[j2 + N2 * n1];
    small_dft<N1, -1>(u);
    twiddle_row<N, N1, -1>(u, j2, tw);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) row[j2 + N2 * k1] = u[k1];
  }
  __syncthreads();
  // 7. forward along k, phase 2: outputs at k = k1 + N1 k2 go straight to their sticks
  for (int t = tid; t < nr * N1; t += NTH) {
    const int s = t / N1, k1 = t % N1;
    const C* row = plane + rl[s] * N + N2 * k1;
    C v[N2];
#pragma unroll
    for (int j2 = 0; j2 < N2; ++j2) v[j2] = row[j2];
    small_dft<N2, -1>(v);
    const unsigned long long lo = rm[2 * s], hi = rm[2 * s + 1];
    C* dst = st + rs[s];
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
      const int kk = k1 + N1 * k2;
      if (mask_bit(lo, hi, kk)) dst[mask_below(lo, hi, kk)] = v[k2];
    }
  }
}

// sum_l w_l |u_l(r)|^2 over the bands of one chunk of the list, on one i-plane: the inverse 2D
// transform of each band's plane as in k_plane_vloc_n (sticks read through the row masks, rows
// without sticks read as zeros along j), and the second phase along j squares its outputs into
// FP64 accumulators held by the task that computes them (task (c, k1) always owns rows
// k1 + N1 m of column c), so the transformed plane is never written back.
template <int N, typename C>
__global__ __launch_bounds__(512) void k_density_plane_n(const C* __restrict__ sticks, const int* __restrict__ rows,
                                                         const int* __restrict__ nrows,
                                                         const unsigned long long* __restrict__ rowmask,
                                                         const int* __restrict__ rowstart,
                                                         const unsigned long long* __restrict__ livemask,
                                                         const int* __restrict__ band_list,
                                                         const double* __restrict__ band_weight, int nlist,
                                                         int chunk, int smax, int nb, int k0,
                                                         const C* __restrict__ tw, double* __restrict__ part) {
  constexpr int NTH = 512;
  constexpr int NP = N * N;
  constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  constexpr int TR = ceil_div(N * N1, NTH);
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  C* plane = reinterpret_cast<C*>(smem_raw);
  __shared__ int rl[(N + 3) / 4 * 4];  // sizes in multiples of 16 bytes keep the dynamic plane 16-byte aligned
  __shared__ int rs[(N + 3) / 4 * 4];
  __shared__ unsigned long long rm[2 * N];
  const int tid = threadIdx.x;
  const int i = blockIdx.x;
  const int ch = blockIdx.y;
  double acc[TR][N2];
#pragma unroll
  for (int q = 0; q < TR; ++q)
#pragma unroll
    for (int m = 0; m < N2; ++m) acc[q][m] = 0.0;
  const int l0 = ch * chunk;
  const int l1 = (l0 + chunk < nlist) ? l0 + chunk : nlist;
  for (int l = l0; l < l1; ++l) {
    const int kb = band_list[l];
    const double wgt = band_weight[l];
    const int k = k0 + kb / nb;
    const int nr = nrows[k];
    const unsigned long long live_lo = livemask[2 * k], live_hi = livemask[2 * k + 1];
    const C* st = sticks + ((long long)kb * N + i) * smax;
    for (int r = tid; r < nr; r += NTH) {
      const int j = rows[(long long)k * N + r];
      rl[r] = j;
      rs[r] = rowstart[(long long)k * N + j];
      rm[2 * r] = rowmask[((long long)k * N + j) * 2];
      rm[2 * r + 1] = rowmask[((long long)k * N + j) * 2 + 1];
    }
    __syncthreads();
    // inverse along k, phase 1 (sticks read through the row masks)
    for (int t = tid; t < nr * N2; t += NTH) {
      const int s = t / N2, j2 = t % N2;
      const unsigned long long lo = rm[2 * s], hi = rm[2 * s + 1];
      const C* src = st + rs[s];
      C u[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) {
        const int kk = j2 + N2 * n1;
        u[n1] = mask_bit(lo, hi, kk) ? src[mask_below(lo, hi, kk)] : cmake<C>(0, 0);
      }
      small_dft<N1, 1>(u);
      twiddle_row<N, N1, 1>(u, j2, tw);
      C* row = plane + rl[s] * N;
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) row[j2 + N2 * k1] = u[k1];
    }
    __syncthreads();
    // inverse along k, phase 2 (in place)
    {
      C v[TR][N2];
#pragma unroll
      for (int q = 0; q < TR; ++q) {
        const int t = tid + q * NTH;
        if (t < nr * N1) {
          const C* row = plane + rl[t / N1] * N + N2 * (t % N1);
#pragma unroll
          for (int j2 = 0; j2 < N2; ++j2) v[q][j2] = row[j2];
          small_dft<N2, 1>(v[q]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < TR; ++q) {
        const int t = tid + q * NTH;
        if (t < nr * N1) {
          C* row = plane + rl[t / N1] * N + t % N1;
#pragma unroll
          for (int k2 = 0; k2 < N2; ++k2) row[N1 * k2] = v[q][k2];
        }
      }
    }
    __syncthreads();
    // inverse along j, phase 1 (rows without sticks read as zeros)
    for (int t = tid; t < N * N2; t += NTH) {
      const int c = t % N, j2 = t / N;
      C u[N1];
#pragma unroll
      for (int n1 = 0; n1 < N1; ++n1) {
        const int j = j2 + N2 * n1;
        u[n1] = mask_bit(live_lo, live_hi, j) ? plane[j * N + c] : cmake<C>(0, 0);
      }
      small_dft<N1, 1>(u);
      twiddle_row<N, N1, 1>(u, j2, tw);
#pragma unroll
      for (int k1 = 0; k1 < N1; ++k1) plane[(j2 + N2 * k1) * N + c] = u[k1];
    }
    __syncthreads();
    // inverse along j, phase 2, squared into the accumulators
#pragma unroll
    for (int q = 0; q < TR; ++q) {
      const int t = tid + q * NTH;
      if (t < N * N1) {
        const int c = t % N, k1 = t / N;
        C x[N2];
#pragma unroll
        for (int j2 = 0; j2 < N2; ++j2) x[j2] = plane[(N2 * k1 + j2) * N + c];
        small_dft<N2, 1>(x);
#pragma unroll
        for (int m = 0; m < N2; ++m) acc[q][m] += wgt * ((double)x[m].x * (double)x[m].x + (double)x[m].y * (double)x[m].y);
      }
    }
    __syncthreads();  // the next band overwrites the plane and the row lists
  }
  double* out = part + ((long long)ch * N + i) * NP;
#pragma unroll
  for (int q = 0; q < TR; ++q) {
    const int t = tid + q * NTH;
    if (t < N * N1) {
      const int c = t % N, k1 = t / N;
#pragma unroll
      for (int m = 0; m < N2; ++m) out[(k1 + N1 * m) * N + c] = acc[q][m];
    }
  }
}

// ----------------------------------------------------------------------- launchers
template <int N, typename R, int W>
static int stick_pass(const Plan& p, int inverse, const void* psi_, void* hpsi_, void* sticks_, int nb, int k0,
                      int nkb, hipStream_t s) {
  using C = typename Cplx<R>::T;
  constexpr int prec = sizeof(R) == 8 ? kFP64 : kFP32;
  const C* tw = prec == kFP64 ? (const C*)p.d_tw64[0] : (const C*)p.d_tw32[0];
  const R* kin = prec == kFP64 ? (const R*)p.d_kin64 : (const R*)p.d_kin32;
  const int nch = (p.smax + W - 1) / W;
  const dim3 grid(stick_grid(nch, nkb));
  if (inverse)
    hipLaunchKernelGGL((k_stick_inverse_n<N, C, W>), grid, dim3(256), 0, s, static_cast<const C*>(psi_),
                       static_cast<C*>(sticks_), p.d_stick_map, p.smax, p.npw_pad, nb, k0, tw, nch, nkb);
  else
    hipLaunchKernelGGL((k_stick_forward_n<N, C, R, W>), grid, dim3(256), 0, s, static_cast<const C*>(sticks_),
                       static_cast<const C*>(psi_), static_cast<C*>(hpsi_), p.d_stick_map, kin, p.smax, p.npw_pad,
                       nb, k0, tw, (R)(1.0 / ((double)N * N * N)), nch, nkb);
  return hipGetLastError() != hipSuccess ? 1 : 0;
}

template <int N, typename R>
static int hpsi_fast_t(const Plan& p, const void* psi_, void* hpsi_, int nb, int k0, int nkk, const void* veff_,
                       void* sticks_, hipStream_t s, int stage) {
  using C = typename Cplx<R>::T;
  constexpr int prec = sizeof(R) == 8 ? kFP64 : kFP32;
  constexpr int W = stick_width<C>();
  C* sticks = static_cast<C*>(sticks_);
  const R* veff = static_cast<const R*>(veff_);
  const C* tw = prec == kFP64 ? (const C*)p.d_tw64[0] : (const C*)p.d_tw32[0];
  const int nkb = nb * nkk;
  const bool half = p.fast_stick_w[prec] == W / 2;  // the launch shape the plan holds (autotuner)
  const size_t plds = (size_t)N * N * sizeof(C);
  if (stage == 0 || stage == 1) {
    if (half ? stick_pass<N, R, W / 2>(p, 1, psi_, hpsi_, sticks_, nb, k0, nkb, s)
             : stick_pass<N, R, W>(p, 1, psi_, hpsi_, sticks_, nb, k0, nkb, s))
      return 1;
  }
  if (stage == 0 || stage == 2) {
    if (p.plane_nth[prec] == 512)
      hipLaunchKernelGGL((k_plane_vloc_n<N, C, R, 512>), dim3(N, nkb), dim3(512), plds, s, sticks, p.d_rows,
                         p.d_nrows, p.d_rowmask, p.d_rowstart, p.d_livemask, veff, p.smax, nb, k0, tw);
    else
      hipLaunchKernelGGL((k_plane_vloc_n<N, C, R, 256>), dim3(N, nkb), dim3(256), plds, s, sticks, p.d_rows,
                         p.d_nrows, p.d_rowmask, p.d_rowstart, p.d_livemask, veff, p.smax, nb, k0, tw);
    if (hipGetLastError() != hipSuccess) return 1;
  }
  if (stage == 0 || stage == 3) {
    if (half ? stick_pass<N, R, W / 2>(p, 0, psi_, hpsi_, sticks_, nb, k0, nkb, s)
             : stick_pass<N, R, W>(p, 0, psi_, hpsi_, sticks_, nb, k0, nkb, s))
      return 1;
  }
  return 0;
}

template <int N, typename R>
static int density_fast_t(const Plan& p, const void* psi_, int nb, int k0, int nkk, const int* list,
                          const double* weight, int nlist, int chunk, void* sticks_, double* part, hipStream_t s) {
  using C = typename Cplx<R>::T;
  constexpr int prec = sizeof(R) == 8 ? kFP64 : kFP32;
  constexpr int W = stick_width<C>();
  const C* tw = prec == kFP64 ? (const C*)p.d_tw64[0] : (const C*)p.d_tw32[0];
  const int nkb = nb * nkk;
  const int nchunk = (nlist + chunk - 1) / chunk;
  const bool half = p.fast_stick_w[prec] == W / 2;  // the launch shape the plan holds (autotuner)
  if (half ? stick_pass<N, R, W / 2>(p, 1, psi_, nullptr, sticks_, nb, k0, nkb, s)
           : stick_pass<N, R, W>(p, 1, psi_, nullptr, sticks_, nb, k0, nkb, s))
    return 1;
  hipLaunchKernelGGL((k_density_plane_n<N, C>), dim3(N, nchunk), dim3(512), (size_t)N * N * sizeof(C), s,
                     static_cast<const C*>(sticks_), p.d_rows, p.d_nrows, p.d_rowmask, p.d_rowstart, p.d_livemask,
                     list, weight, nlist, chunk, p.smax, nb, k0, tw, part);
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
