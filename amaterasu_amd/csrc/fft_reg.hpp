// Compile-time-length FFTs for the grid lengths the workloads use.
//
// A length-N transform (N = N1 * N2, both <= 16) is done as the four-step
// decomposition with every small DFT held in registers:
//   input index  n = j2 + N2 n1,   output index k = k1 + N1 k2
//   X[k1 + N1 k2] = sum_j2 w_N^{j2 k1} w_N2^{j2 k2} sum_n1 x[j2 + N2 n1] w_N1^{n1 k1}
// Phase 1: one task per (sequence, j2) transforms the N1 elements j2 + N2 n1 in place
// (each task owns its positions, so no barrier is needed inside the phase) and applies
// the twiddle w_N^{j2 k1}. Phase 2: one task per (sequence, k1) reads the N2 elements
// N2 k1 + j2, transforms them and -- after a barrier, since the outputs land on other
// tasks' inputs -- writes X[k1 + N1 k2]. Two LDS round trips per transform, against one
// per radix pass for the generic Stockham path, with all inner twiddles folded to
// constants (constexpr trigonometry below) and no runtime radix dispatch.
#pragma once

#include "fft_lds.hpp"

namespace amt {

// ------------------------------------------------------------ constexpr trigonometry
constexpr double kTwoPi = 6.283185307179586476925286766559005768;

constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 14; ++i) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 14; ++i) {
    term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi num / den: reduce to the nearest quarter turn, then |x| <= pi/4.
constexpr double cos2pi(long num, long den) {
  long n = num % den;
  if (n < 0) n += den;
  const long m = (8 * n + den) / (2 * den);  // nearest quarter: round(4 n / den)
  const double x = kTwoPi * (double)(4 * n - m * den) / (4.0 * (double)den);
  switch (m % 4) {
    case 0: return taylor_cos(x);
    case 1: return -taylor_sin(x);
    case 2: return -taylor_cos(x);
    default: return taylor_sin(x);
  }
}
constexpr double sin2pi(long num, long den) { return cos2pi(num * 4 - den, 4 * den); }

// ------------------------------------------------------------ small in-register DFTs
template <int L>
struct SmallSplit {  // L = A * B for composite L not handled directly
  static constexpr int A = (L % 4 == 0 && L != 4 && L != 8) ? 4 : (L % 2 == 0 ? 2 : (L % 3 == 0 ? 3 : (L % 5 == 0 ? 5 : 7)));
  static constexpr int B = L / A;
};

template <int L, int SIGN, typename C>
__host__ __device__ __forceinline__ void small_dft(C* v) {
  using Rt = decltype(C::x);
  if constexpr (L == 1) {
    return;
  } else if constexpr (L == 2 || L == 4 || L == 8 || L == 3 || L == 5 || L == 7 || L == 11) {
    dft_any<L, SIGN>(v);
  } else {
    constexpr int A = SmallSplit<L>::A, B = SmallSplit<L>::B;
    static_assert(A * B == L, "small DFT split");
    // input index a*? : n = b + B a, output k = ka + A kb
    C t[L];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      C u[A];
#pragma unroll
      for (int a = 0; a < A; ++a) u[a] = v[b + B * a];
      small_dft<A, SIGN>(u);
#pragma unroll
      for (int ka = 0; ka < A; ++ka) {
        const Rt c = (Rt)cos2pi((long)b * ka, L);
        const Rt s = (Rt)(SIGN * sin2pi((long)b * ka, L));
        t[b + B * ka] = (b * ka == 0) ? u[ka] : cmake<C>(u[ka].x * c - u[ka].y * s, u[ka].x * s + u[ka].y * c);
      }
    }
#pragma unroll
    for (int ka = 0; ka < A; ++ka) {
      C w[B];
#pragma unroll
      for (int b = 0; b < B; ++b) w[b] = t[b + B * ka];
      small_dft<B, SIGN>(w);
#pragma unroll
      for (int kb = 0; kb < B; ++kb) v[ka + A * kb] = w[kb];
    }
  }
}

// ------------------------------------------------------------ length splits
// N = N1 * N2 with both factors <= 16 and close to sqrt(N); N1 >= N2.
constexpr int split_n1(int n) {
  int best = 0;
  for (int a = 1; a <= 16; ++a) {
    if (n % a == 0 && n / a <= 16 && a >= n / a) {
      if (best == 0 || a < best) best = a;
    }
  }
  return best;
}
template <int N>
struct Split {
  static constexpr int N1 = split_n1(N);
  static constexpr int N2 = N1 > 0 ? N / N1 : 0;
  static constexpr bool ok = N1 > 0;
};

// Batched length-N transform of `nseq` sequences in LDS: element e of sequence s at
// buf[seqbase(s) + e * es], seqbase(s) = (rows ? rows[s] : s) * ss. Sequences listed
// by `rows` (when not null) let a caller transform only the rows that hold data.
// `tw` is exp(-2 pi i m / N), m = 0..N-1. Every thread of the workgroup must call it.
// T2 bounds the phase-2 tasks per thread: nseq * N1 <= T2 * nth (checked by the caller).
// ROWS selects the task order: true when es == 1 (consecutive threads walk one
// sequence, so their LDS addresses are consecutive), false when ss == 1.
template <int N, int SIGN, int T2, bool ROWS, typename C>
__device__ __forceinline__ void fft2(C* buf, int es, int ss, int nseq, const int* rows, const C* __restrict__ tw,
                                     int tid, int nth) {
  constexpr int N1 = Split<N>::N1, N2 = Split<N>::N2;
  static_assert(Split<N>::ok, "length has no factorisation into two factors <= 16");
  // phase 1: (s, j2)
  const int t1 = nseq * N2;
  for (int t = tid; t < t1; t += nth) {
    const int s = ROWS ? t / N2 : t % nseq, j2 = ROWS ? t % N2 : t / nseq;
    C* base = buf + (rows ? rows[s] : s) * ss;
    C u[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) u[n1] = base[(j2 + N2 * n1) * es];
    small_dft<N1, SIGN>(u);
#pragma unroll
    for (int k1 = 1; k1 < N1; ++k1) {
      C w = tw[j2 * k1];
      if (SIGN > 0) w.y = -w.y;
      u[k1] = cmul(u[k1], w);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) base[(j2 + N2 * k1) * es] = u[k1];
  }
  __syncthreads();
  // phase 2: (s, k1)
  const int t2 = nseq * N1;
  C v[T2][N2];
#pragma unroll
  for (int q = 0; q < T2; ++q) {
    const int t = tid + q * nth;
    if (t < t2) {
      const int s = ROWS ? t / N1 : t % nseq, k1 = ROWS ? t % N1 : t / nseq;
      const C* base = buf + (rows ? rows[s] : s) * ss;
#pragma unroll
      for (int j2 = 0; j2 < N2; ++j2) v[q][j2] = base[(N2 * k1 + j2) * es];
      small_dft<N2, SIGN>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < T2; ++q) {
    const int t = tid + q * nth;
    if (t < t2) {
      const int s = ROWS ? t / N1 : t % nseq, k1 = ROWS ? t % N1 : t / nseq;
      C* base = buf + (rows ? rows[s] : s) * ss;
#pragma unroll
      for (int k2 = 0; k2 < N2; ++k2) base[(k1 + N1 * k2) * es] = v[q][k2];
    }
  }
  __syncthreads();
}

}  // namespace amt
