// Batched mixed-radix (2, 3, 4, 5, 7, 8, 11) complex FFTs on data held in LDS.
//
// Every grid length the plane-wave grid rule can produce is 2,3,5,7,11-smooth, so a
// length-n transform is a sequence of Stockham passes, one per radix factor. A pass
// of radix R over a sequence of length n is n/R independent butterflies; a
// workgroup owns a batch of sequences in LDS and its threads share all butterflies
// of the batch. Each thread stages its butterflies' inputs in registers, the
// workgroup synchronises, then every thread writes its outputs back to the same LDS
// buffer (in-place Stockham: no ping-pong buffer, which halves the LDS footprint
// and so doubles the number of resident workgroups per CU).
//
// Stockham pass (radix R, Ns = product of the radices already applied):
//   butterfly j in [0, n/R):  k = j mod Ns
//     v[r] = x[j + r n/R] * w^(r k n/(Ns R)),   r = 0..R-1
//     V    = DFT_R(v)
//     x'[(j/Ns) Ns R + k + r Ns] = V[r]
// w = exp(SIGN 2 pi i / n): SIGN = +1 is the inverse transform (u = sum c e^{+iGr}),
// SIGN = -1 the forward one. No 1/n is applied here.
#pragma once

#include "common.hpp"

namespace amt {

constexpr int kMaxPasses = 12;

// Radices packed 4 bits per pass (pass p in bits 4p..4p+3), so kernels read them from
// a scalar argument instead of indexing a struct array (which would live in scratch).
struct AxisPlan {
  int n;
  int npass;
  unsigned long long code;
  __host__ __device__ int radix(int p) const { return (int)((code >> (4 * p)) & 15ull); }
};

// cos(2 pi m / R) and sin(2 pi m / R), m = 0..R-1, for the odd prime radices.
template <int R>
struct Trig;
template <>
struct Trig<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.8660254037844386, -0.8660254037844386};
};
template <>
struct Trig<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494745, -0.8090169943749475, -0.8090169943749475,
                                  0.30901699437494745};
  static constexpr double s[5] = {0.0, 0.9510565162951535, 0.5877852522924731, -0.5877852522924731,
                                  -0.9510565162951535};
};
template <>
struct Trig<7> {
  static constexpr double c[7] = {1.0, 0.6234898018587335, -0.2225209339563144, -0.9009688679024191,
                                  -0.9009688679024191, -0.2225209339563144, 0.6234898018587335};
  static constexpr double s[7] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.4338837391175581,
                                  -0.4338837391175581, -0.9749279121818236, -0.7818314824680298};
};
template <>
struct Trig<11> {
  static constexpr double c[11] = {1.0, 0.8412535328311812, 0.41541501300188644, -0.14231483827328514,
                                   -0.6548607339452851, -0.9594929736144974, -0.9594929736144974,
                                   -0.6548607339452851, -0.14231483827328514, 0.41541501300188644,
                                   0.8412535328311812};
  static constexpr double s[11] = {0.0, 0.5406408174555976, 0.9096319953545183, 0.9898214418809327,
                                   0.7557495743542583, 0.28173255684142967, -0.28173255684142967,
                                   -0.7557495743542583, -0.9898214418809327, -0.9096319953545183,
                                   -0.5406408174555976};
};

// In-register DFT of length R: v[q] <- sum_r v[r] exp(SIGN 2 pi i r q / R).
template <int R, int SIGN, typename C>
__device__ __forceinline__ void dft(C* v);

template <int R, int SIGN, typename C>
__host__ __device__ __forceinline__ void dft_odd(C* v) {
  using Rt = decltype(C::x);
  constexpr int H = (R - 1) / 2;
  C a[H], b[H];
#pragma unroll
  for (int r = 1; r <= H; ++r) {
    a[r - 1] = cadd(v[r], v[R - r]);
    b[r - 1] = csub(v[r], v[R - r]);
  }
  C out[R];
  C x0 = v[0];
#pragma unroll
  for (int r = 0; r < H; ++r) x0 = cadd(x0, a[r]);
  out[0] = x0;
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    Rt ar = v[0].x, ai = v[0].y, br = 0, bi = 0;
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      const int m = (r * q) % R;
      const Rt cs = static_cast<Rt>(Trig<R>::c[m]);
      const Rt sn = static_cast<Rt>(Trig<R>::s[m]);
      ar += a[r - 1].x * cs;
      ai += a[r - 1].y * cs;
      br += b[r - 1].x * sn;
      bi += b[r - 1].y * sn;
    }
    out[q] = cmake<C>(ar - SIGN * bi, ai + SIGN * br);
    out[R - q] = cmake<C>(ar + SIGN * bi, ai - SIGN * br);
  }
#pragma unroll
  for (int q = 0; q < R; ++q) v[q] = out[q];
}

template <int SIGN, typename C>
__host__ __device__ __forceinline__ void dft2(C& a, C& b) {
  C t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int SIGN, typename C>
__host__ __device__ __forceinline__ void dft4v(C& v0, C& v1, C& v2, C& v3) {
  C t0 = cadd(v0, v2), t1 = csub(v0, v2);
  C t2 = cadd(v1, v3), t3 = cmul_i<SIGN>(csub(v1, v3));
  v0 = cadd(t0, t2);
  v2 = csub(t0, t2);
  v1 = cadd(t1, t3);
  v3 = csub(t1, t3);
}

template <>
__device__ __forceinline__ void dft<2, 1, float2>(float2* v) { dft2<1>(v[0], v[1]); }
template <>
__device__ __forceinline__ void dft<2, -1, float2>(float2* v) { dft2<-1>(v[0], v[1]); }
template <>
__device__ __forceinline__ void dft<2, 1, double2>(double2* v) { dft2<1>(v[0], v[1]); }
template <>
__device__ __forceinline__ void dft<2, -1, double2>(double2* v) { dft2<-1>(v[0], v[1]); }

template <int SIGN, typename C>
__host__ __device__ __forceinline__ void dft4(C* v) {
  dft4v<SIGN>(v[0], v[1], v[2], v[3]);
}

template <int SIGN, typename C>
__host__ __device__ __forceinline__ void dft8(C* v) {
  using Rt = decltype(C::x);
  // even/odd split: X[q] = E[q] + w8^q O[q], X[q+4] = E[q] - w8^q O[q]
  C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4v<SIGN>(e0, e1, e2, e3);
  dft4v<SIGN>(o0, o1, o2, o3);
  const Rt h = static_cast<Rt>(0.7071067811865476);
  // w8 = (1 + SIGN i)/sqrt2, w8^2 = SIGN i, w8^3 = (-1 + SIGN i)/sqrt2
  C w1o = cmake<C>(h * (o1.x - SIGN * o1.y), h * (o1.y + SIGN * o1.x));
  C w2o = cmul_i<SIGN>(o2);
  C w3o = cmake<C>(h * (-o3.x - SIGN * o3.y), h * (-o3.y + SIGN * o3.x));
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, w1o);
  v[5] = csub(e1, w1o);
  v[2] = cadd(e2, w2o);
  v[6] = csub(e2, w2o);
  v[3] = cadd(e3, w3o);
  v[7] = csub(e3, w3o);
}

template <int R, int SIGN, typename C>
__host__ __device__ __forceinline__ void dft_any(C* v) {
  if constexpr (R == 2) {
    dft2<SIGN>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<SIGN>(v);
  } else if constexpr (R == 8) {
    dft8<SIGN>(v);
  } else {
    dft_odd<R, SIGN>(v);
  }
}

// One Stockham pass of radix RAD over `nseq` sequences in LDS. Element e of sequence
// s lives at buf[s * ss + e * es]. `seq_fastest` orders butterflies so consecutive
// threads walk the unit-stride index (sequence index when ss == 1, butterfly index
// when es == 1), which keeps LDS accesses bank-friendly. `tw` is the table
// exp(-2 pi i m / n), m = 0..n-1. A thread stages up to tasks_per_thread(RAD, TPT)
// butterflies (about 4*TPT complex values whatever the radix, which bounds register
// use); the caller guarantees nseq * n / RAD <= tasks_per_thread * nth.
__host__ __device__ constexpr int tasks_per_thread(int rad, int tpt) {
  return (4 * tpt) / rad > 1 ? (4 * tpt) / rad : 1;
}

template <int RAD, int SIGN, int TPT_BASE, typename C>
__device__ __forceinline__ void stockham_pass(C* buf, int n, int ns, int nseq, int es, int ss, bool seq_fastest,
                                              const C* __restrict__ tw, int tid, int nth) {
  constexpr int TPT = tasks_per_thread(RAD, TPT_BASE);
  const int m = n / RAD;
  const int ntask = m * nseq;
  const int twstep = n / (ns * RAD);
  C v[TPT][RAD];
#pragma unroll
  for (int u = 0; u < TPT; ++u) {
    const int t = tid + u * nth;
    if (t < ntask) {
      const int s = seq_fastest ? t % nseq : t / m;
      const int j = seq_fastest ? t / nseq : t % m;
      const C* base = buf + s * ss;
      const int k = j % ns;
#pragma unroll
      for (int r = 0; r < RAD; ++r) v[u][r] = base[(j + r * m) * es];
      if (ns > 1) {
#pragma unroll
        for (int r = 1; r < RAD; ++r) {
          C w = tw[r * k * twstep];
          if (SIGN > 0) w.y = -w.y;
          v[u][r] = cmul(v[u][r], w);
        }
      }
      dft_any<RAD, SIGN>(v[u]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < TPT; ++u) {
    const int t = tid + u * nth;
    if (t < ntask) {
      const int s = seq_fastest ? t % nseq : t / m;
      const int j = seq_fastest ? t / nseq : t % m;
      C* base = buf + s * ss;
      const int k = j % ns;
      const int d = (j / ns) * ns * RAD + k;
#pragma unroll
      for (int r = 0; r < RAD; ++r) base[(d + r * ns) * es] = v[u][r];
    }
  }
  __syncthreads();
}

// Full length-n transform of a batch in LDS (all passes). Must be called by every
// thread of the workgroup (it synchronises).
template <int SIGN, int TPT, typename C>
__device__ __forceinline__ void lds_fft(C* buf, const AxisPlan& ap, int nseq, int es, int ss, bool seq_fastest,
                        const C* __restrict__ tw, int tid, int nth) {
  int ns = 1;
  for (int p = 0; p < ap.npass; ++p) {
    const int rad = ap.radix(p);
    // Opaque per-pass copies: without them the compiler hoists every case's address
    // arithmetic out of the pass loop and the register demand becomes the sum, not the
    // maximum, over the radix cases (then spills to scratch).
    int tid_l = tid, es_l = es, ss_l = ss, nseq_l = nseq;
    __asm__ volatile("" : "+v"(tid_l));
    __asm__ volatile("" : "+s"(es_l), "+s"(ss_l), "+s"(nseq_l));
    switch (rad) {
      case 2:
        stockham_pass<2, SIGN, TPT>(buf, ap.n, ns, nseq_l, es_l, ss_l, seq_fastest, tw, tid_l, nth);
        break;
      case 3:
        stockham_pass<3, SIGN, TPT>(buf, ap.n, ns, nseq_l, es_l, ss_l, seq_fastest, tw, tid_l, nth);
        break;
      case 4:
        stockham_pass<4, SIGN, TPT>(buf, ap.n, ns, nseq_l, es_l, ss_l, seq_fastest, tw, tid_l, nth);
        break;
      case 5:
        stockham_pass<5, SIGN, TPT>(buf, ap.n, ns, nseq_l, es_l, ss_l, seq_fastest, tw, tid_l, nth);
        break;
      case 7:
        stockham_pass<7, SIGN, TPT>(buf, ap.n, ns, nseq_l, es_l, ss_l, seq_fastest, tw, tid_l, nth);
        break;
      case 8:
        stockham_pass<8, SIGN, TPT>(buf, ap.n, ns, nseq_l, es_l, ss_l, seq_fastest, tw, tid_l, nth);
        break;
      case 11:
        stockham_pass<11, SIGN, TPT>(buf, ap.n, ns, nseq_l, es_l, ss_l, seq_fastest, tw, tid_l, nth);
        break;
      default:
        break;  // rejected on the host: every plan is factorised into these radices
    }
    ns *= rad;
  }
}

}  // namespace amt
