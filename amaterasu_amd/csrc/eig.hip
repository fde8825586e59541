// Lowest eigenpairs of a dense Hermitian matrix (FP64 complex, n <= 2048): the subspace
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
// 2. The lowest eigenpairs of the real symmetric tridiagonal T (d, e) only:
//    a. eigenvalues 0..m-1 (m = min(n, nev + kClusterMax)) by multisection: one workgroup
//       per eigenvalue evaluates 512 Sturm counts per step (the pivmin-guarded LDL^T
//       recurrence of LAPACK's dstebz), so each step narrows the bracket 513-fold;
//    b. eigenvectors by inverse iteration with each eigenvalue as its own shift (tridiagonal
//       LU with partial pivoting, pivots floored at eps ||T||, as LAPACK's dlagtf/dlagts),
//       two rounds from pseudo-random starts (with a shift within 2 eps ||T|| of the
//       eigenvalue, one round shrinks every other component by eps ||T|| / gap, below 1e-12
//       for gaps above the cluster threshold). Eigenvalues closer than kClusterGap ||T||
//       to a neighbour form a cluster, solved by one workgroup (a chain longer than
//       kClusterMax is cut at its largest gaps into pieces that the workgroup solves in
//       order, each orthogonalised against the piece before): after every round the
//       cluster's vectors are orthonormalised together (classical Gram-Schmidt, twice), and
//       a Rayleigh-Ritz step inside the cluster (cyclic Jacobi on the c x c projection)
//       makes them eigenvectors, so near-degenerate and degenerate eigenvalues come out
//       orthogonal with small residuals. Vectors of well-separated eigenvalues are accurate
//       to about eps ||T|| / gap, below 1e-12 for gaps above the cluster threshold.
// 3. Back-transformation Z = Q S for the lowest nev eigenvectors in compact-WY blocks of 32
//    reflectors: the blocks' T factors from their Gram matrices, then one workgroup per two
//    eigenvectors applies the blocks in reverse order (no global synchronisation).
#include <cfloat>

#include "kernels.hpp"

namespace amt {

namespace eig {
constexpr int NT = 256;
constexpr int MAXN = 2048;
constexpr int kClusterMax = 64;      // largest cluster solved together
constexpr double kClusterGap = 2e-4;  // cluster threshold, relative to the Gershgorin bound of ||T||
constexpr int kInverseRounds = 2;
}  // namespace eig

// Columns per LDS slice of the Rayleigh-Ritz products inside a cluster.
constexpr int kRRCols = 64;

// Steps of a serial tridiagonal chain whose independent LDS loads are issued together (the
// chain itself stays in registers, so LDS latency is paid once per chunk).
constexpr int kTriChunk = 8;

__device__ __forceinline__ double2 zmul(double2 a, double2 b) { return cmul(a, b); }

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double lane_f64(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Sum over the wavefront, the same bits in every lane: rotations within each row of 16 lanes
// (commutative pairs, so every lane of a row ends with the same sum), then the four row sums
// added in a fixed order.
__device__ __forceinline__ double wave_allsum(double x) {
  x += dpp_f64<0x128>(x);  // row_ror:8
  x += dpp_f64<0x124>(x);  // row_ror:4
  x += dpp_f64<0x122>(x);  // row_ror:2
  x += dpp_f64<0x121>(x);  // row_ror:1
  return (lane_f64(x, 0) + lane_f64(x, 16)) + (lane_f64(x, 32) + lane_f64(x, 48));
}

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

// LDS of one launch: three vectors of n in dynamic LDS (sized to the matrix, so that the
// solves of several k-points can share the CUs) and the reduction scratch.
struct HetrdShared {
  double2* v;   // reflector of the current step (v[0] = 1), trailing numbering
  double2* w;   // w of the current step
  double2* nr;  // the updated pivot row of the next step
  double* red;
};
extern __shared__ __attribute__((aligned(16))) double2 hetrd_dyn[];
extern __shared__ __attribute__((aligned(16))) double tri_dyn[];
#define HETRD_SHARED(n)                              \
  __shared__ double hetrd_red[2 * (eig::NT / 64) + 2]; \
  HetrdShared sh{hetrd_dyn, hetrd_dyn + (n), hetrd_dyn + 2 * (n), hetrd_red}

// Prologue: v_0 from row 0, d_0 / e_0 / tau_0, V[0], and p_0 = tau_0 A[1:,1:] v_0 for own rows.
__global__ __launch_bounds__(eig::NT) void k_hetrd_start(const double2* __restrict__ A, int n, double2* __restrict__ V,
                                                         double2* __restrict__ taus, double* __restrict__ d,
                                                         double* __restrict__ e, double2* __restrict__ p) {
  HETRD_SHARED(n);
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

// Block-wide sum of a double2 with one barrier: every thread adds the per-wave partials
// itself, in the same order. red (2 * NT/64 doubles) must not be reused in the same launch.
__device__ __forceinline__ double2 block_sum2_once(double2 v, double* red) {
  v = make_double2(wave_allsum(v.x), wave_allsum(v.y));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[2 * wave] = v.x;
    red[2 * wave + 1] = v.y;
  }
  __syncthreads();
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int w = 0; w < eig::NT / 64; ++w) {
    sx += red[2 * w];
    sy += red[2 * w + 1];
  }
  return make_double2(sx, sy);
}

// Launch j (0 <= j <= n-2): finish step j, start step j+1. Trailing numbering of step j:
// index t <-> row/column j+1+t. p_in: p_j (length n-j-1); p_out: p_{j+1}.
// Each workgroup owns at most one trailing row per wave (n - 2 <= 4 G, checked by the host):
// the wave keeps that row in registers from its update to its product with v_{j+1}. Thread
// x holds elements t = x + NT q of v_j, p_j, w_j and of the pivot row j+1 in registers, so
// all global loads are issued at once at the start, and the launch has three barriers: the
// sum p^H v, the publication of w, and the norm of the new column. The reflector of column
// j+1 (zlarfg) is kept unscaled in LDS; its scale enters the product with the owned row
// once, after the row's dot product, and the stored V row.
// kRowRegs = elements of one row per lane, 64 kRowRegs >= the trailing length n - j - 1 (8, 12
// or 16; longer rows: k_hetrd_step_wide, shorter ones: k_hetrd_step_w64).

template <int kRowRegs>
__global__ __launch_bounds__(eig::NT) void k_hetrd_step(double2* __restrict__ A, int n, int j,
                                                        double2* __restrict__ V, double2* __restrict__ taus,
                                                        double* __restrict__ d, double* __restrict__ e,
                                                        const double2* __restrict__ p_in, double2* __restrict__ p_out) {
  constexpr int QV = kRowRegs / 4;  // vector elements per thread (n - j - 1 <= NT * QV)
  double2* sv = hetrd_dyn;          // v_j
  double2* sw = hetrd_dyn + n;      // w_j
  double2* sx = hetrd_dyn + 2 * n;  // column j+1 below the diagonal, unscaled
  __shared__ double red[2][2 * (eig::NT / 64)];
  const int len = n - j - 1;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const double2 zero = make_double2(0.0, 0.0);
  const int myrow = j + 2 + blockIdx.x + wave * gridDim.x;
  const bool own = myrow < n;
  double2 rowv[kRowRegs];
#pragma unroll
  for (int q = 0; q < kRowRegs; ++q) {
    const int c = j + 1 + lane + 64 * q;
    rowv[q] = (own && c < n) ? A[(long long)myrow * n + c] : make_double2(0.0, 0.0);
  }
  double2 piv[QV], vq[QV], pq[QV];
#pragma unroll
  for (int q = 0; q < QV; ++q) {
    const int t = tid + eig::NT * q;
    piv[q] = t < len ? A[(long long)(j + 1) * n + j + 1 + t] : make_double2(0.0, 0.0);
    vq[q] = t < len ? V[(long long)j * n + t] : make_double2(0.0, 0.0);
    pq[q] = t < len ? p_in[t] : make_double2(0.0, 0.0);
  }
  const double2 tau = taus[j];
  // K = -(tau/2) p^H v; w = p + K v
  double2 part = zero;
#pragma unroll
  for (int q = 0; q < QV; ++q) {
    const int t = tid + eig::NT * q;
    if (t < len) sv[t] = vq[q];
    const double2 pv = zmul(zconj(pq[q]), vq[q]);
    part.x += pv.x;
    part.y += pv.y;
  }
  const double2 phv = block_sum2_once(part, red[0]);
  const double2 K = zmul(make_double2(-0.5 * tau.x, -0.5 * tau.y), phv);
  double2 wq[QV];
#pragma unroll
  for (int q = 0; q < QV; ++q) {
    const int t = tid + eig::NT * q;
    const double2 kv = zmul(K, vq[q]);
    wq[q] = make_double2(pq[q].x + kv.x, pq[q].y + kv.y);
    if (t < len) sw[t] = wq[q];
  }
  __syncthreads();
  // update the owned row in registers and store it:  A -= v w^H + w v^H
  if (own) {
    const int tr = myrow - j - 1;
    const double2 vr = sv[tr], wr = sw[tr];
    double2* row = A + (long long)myrow * n;
#pragma unroll
    for (int q = 0; q < kRowRegs; ++q) {
      const int c = j + 1 + lane + 64 * q;
      if (c < n) {
        const int tc = c - j - 1;
        const double2 a = zmul(vr, zconj(sw[tc]));
        const double2 b = zmul(wr, zconj(sv[tc]));
        rowv[q].x -= a.x + b.x;
        rowv[q].y -= a.y + b.y;
        row[c] = rowv[q];
      }
    }
  }
  const double2 v0 = sv[0], w0 = sw[0];
  const int len2 = n - j - 2;  // length of column j+1 below the diagonal
  if (len2 <= 0) {
    if (blockIdx.x == 0 && tid == 0) {
      const double2 a = zmul(v0, zconj(w0));
      d[j + 1] = piv[0].x - 2.0 * a.x;
    }
    return;
  }
  // the updated pivot row j+1 (element t = column j+1+t): old - (v_0 conj(w_t) + w_0 conj(v_t));
  // the column below the diagonal is its conjugate from t = 1 on: x[t-1]
  double2 xq[QV];
  double nrm2 = 0.0;
#pragma unroll
  for (int q = 0; q < QV; ++q) {
    const int t = tid + eig::NT * q;
    const double2 a = zmul(v0, zconj(wq[q]));
    const double2 b = zmul(w0, zconj(vq[q]));
    const double2 nr = make_double2(piv[q].x - a.x - b.x, piv[q].y - a.y - b.y);
    xq[q] = zconj(nr);
    if (t == 0) {
      if (blockIdx.x == 0) d[j + 1] = nr.x;
    } else if (t < len) {
      sx[t - 1] = xq[q];
      if (t >= 2) nrm2 += nr.x * nr.x + nr.y * nr.y;
    }
  }
  const double xnorm2 = block_sum2_once(make_double2(nrm2, 0.0), red[1]).x;
  // zlarfg on x: beta, tau2 and the scale 1 / (alpha - beta) of x[1..]
  const double2 alpha = sx[0];
  double beta;
  double2 tau2, sc;
  if (xnorm2 == 0.0 && alpha.y == 0.0) {
    beta = alpha.x;
    tau2 = zero;
    sc = zero;  // x[1..] = 0: H = I
  } else {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tau2 = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    const double2 dd = make_double2(alpha.x - beta, alpha.y);
    const double den = dd.x * dd.x + dd.y * dd.y;
    sc = make_double2(dd.x / den, -dd.y / den);
  }
  if (blockIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < QV; ++q) {
      const int t = tid + eig::NT * q;  // x index t - 1
      if (t >= 1 && t < len) V[(long long)(j + 1) * n + t - 1] = t == 1 ? make_double2(1.0, 0.0) : zmul(xq[q], sc);
    }
    if (tid == 0) {
      taus[j + 1] = tau2;
      e[j + 1] = beta;
    }
  }
  // p_{j+1} for the owned row: tau2 (A[r][j+2] + sc sum_{c >= j+3} A[r][c] x[c - j - 2])
  if (own) {
    double2 a0 = zero, a1 = zero;
#pragma unroll
    for (int q = 0; q < kRowRegs; ++q) {
      const int c = j + 1 + lane + 64 * q;
      if (c == j + 2) {
        a0 = rowv[q];
      } else if (c > j + 2 && c < n) {
        const double2 xv = sx[c - j - 2];
        a1.x += rowv[q].x * xv.x - rowv[q].y * xv.y;
        a1.y += rowv[q].x * xv.y + rowv[q].y * xv.x;
      }
    }
    const double2 s1 = zmul(sc, a1);
    double2 acc = make_double2(a0.x + s1.x, a0.y + s1.y);
    acc = make_double2(wave_allsum(acc.x), wave_allsum(acc.y));
    if (lane == 0) p_out[myrow - j - 2] = zmul(tau2, acc);
  }
}

// Steps whose trailing length is at most 64 Q, one wavefront (one trailing row) per
// workgroup: lane l holds the elements t = l + 64 q (q < Q) of v_j, p_j, w_j and of the pivot
// row, and the same columns of its row, so the row update and the product with v_{j+1} use
// the lane's own elements only and the three sums are wavefront reductions: no LDS and no
// barrier. Same formulas as k_hetrd_step.
constexpr int kW64Max = 8;  // Q of the longest wavefront step (trailing length 512)

__device__ __forceinline__ double2 wave_sum2(double2 v) { return make_double2(wave_allsum(v.x), wave_allsum(v.y)); }

// Element idx (uniform across the wavefront) of a vector distributed as t = lane + 64 q; the
// element is selected from the registers by unrolled compares (no dynamic register index).
template <int Q>
__device__ __forceinline__ double2 w64_element(const double2 (&x)[Q], int idx) {
  const int q = idx >> 6, src = idx & 63;
  double vx = x[0].x, vy = x[0].y;
#pragma unroll
  for (int u = 1; u < Q; ++u) {
    vx = q == u ? x[u].x : vx;
    vy = q == u ? x[u].y : vy;
  }
  return make_double2(lane_f64(vx, src), lane_f64(vy, src));
}
#define AMT_W64(x, idx) w64_element<Q>(x, idx)

template <int Q>
__global__ __launch_bounds__(64) void k_hetrd_step_w64(double2* __restrict__ A, int n, int j,
                                                       double2* __restrict__ V, double2* __restrict__ taus,
                                                       double* __restrict__ d, double* __restrict__ e,
                                                       const double2* __restrict__ p_in, double2* __restrict__ p_out) {
  const int len = n - j - 1;
  const int lane = threadIdx.x;
  const int myrow = j + 2 + blockIdx.x;
  const bool own = myrow < n;
  const double2 zero = make_double2(0.0, 0.0);
  double2 rowv[Q], piv[Q], vq[Q], pq[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int t = lane + 64 * q, c = j + 1 + t;
    rowv[q] = (own && c < n) ? A[(long long)myrow * n + c] : make_double2(0.0, 0.0);
    piv[q] = t < len ? A[(long long)(j + 1) * n + c] : make_double2(0.0, 0.0);
    vq[q] = t < len ? V[(long long)j * n + t] : make_double2(0.0, 0.0);
    pq[q] = t < len ? p_in[t] : make_double2(0.0, 0.0);
  }
  const double2 tau = taus[j];
  // K = -(tau/2) p^H v; w = p + K v
  double2 part = zero;
#pragma unroll
  for (int q = 0; q < Q; ++q) part = cadd(part, zmul(zconj(pq[q]), vq[q]));
  const double2 K = zmul(make_double2(-0.5 * tau.x, -0.5 * tau.y), wave_sum2(part));
  double2 wq[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) wq[q] = cadd(pq[q], zmul(K, vq[q]));
  // the owned row: A -= v w^H + w v^H
  if (own) {
    const int tr = myrow - j - 1;
    const double2 vr = AMT_W64(vq, tr), wr = AMT_W64(wq, tr);
    double2* row = A + (long long)myrow * n;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int c = j + 1 + lane + 64 * q;
      if (c < n) {
        rowv[q] = csub(rowv[q], cadd(zmul(vr, zconj(wq[q])), zmul(wr, zconj(vq[q]))));
        row[c] = rowv[q];
      }
    }
  }
  const double2 v0 = AMT_W64(vq, 0), w0 = AMT_W64(wq, 0);
  const int len2 = n - j - 2;
  if (len2 <= 0) {
    if (blockIdx.x == 0 && lane == 0) d[j + 1] = piv[0].x - 2.0 * zmul(v0, zconj(w0)).x;
    return;
  }
  // the updated pivot row j+1 and the column below the diagonal (its conjugate, t >= 1)
  double2 xq[Q];
  double nrm2 = 0.0;
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int t = lane + 64 * q;
    const double2 nr = csub(piv[q], cadd(zmul(v0, zconj(wq[q])), zmul(w0, zconj(vq[q]))));
    xq[q] = zconj(nr);
    if (t == 0 && blockIdx.x == 0) d[j + 1] = nr.x;
    if (t >= 2 && t < len) nrm2 += nr.x * nr.x + nr.y * nr.y;
  }
  const double xnorm2 = wave_sum2(make_double2(nrm2, 0.0)).x;
  const double2 alpha = AMT_W64(xq, 1);
  double beta;
  double2 tau2, sc;
  if (xnorm2 == 0.0 && alpha.y == 0.0) {
    beta = alpha.x;
    tau2 = zero;
    sc = zero;
  } else {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tau2 = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    const double2 dd = make_double2(alpha.x - beta, alpha.y);
    const double den = dd.x * dd.x + dd.y * dd.y;
    sc = make_double2(dd.x / den, -dd.y / den);
  }
  if (blockIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int t = lane + 64 * q;
      if (t >= 1 && t < len) V[(long long)(j + 1) * n + t - 1] = t == 1 ? make_double2(1.0, 0.0) : zmul(xq[q], sc);
    }
    if (lane == 0) {
      taus[j + 1] = tau2;
      e[j + 1] = beta;
    }
  }
  // p_{j+1} for the owned row: tau2 (A[r][j+2] + sc sum_{c >= j+3} A[r][c] x[c - j - 2]); column
  // c = j + 1 + t pairs with x at t - 1, which is this lane's xq
  if (own) {
    double2 a0 = zero, a1 = zero;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int t = lane + 64 * q, c = j + 1 + t;
      if (t == 1) {
        a0 = rowv[q];
      } else if (t >= 2 && c < n) {
        a1 = cadd(a1, zmul(rowv[q], xq[q]));
      }
    }
    const double2 acc = wave_sum2(cadd(a0, zmul(sc, a1)));
    if (lane == 0) p_out[myrow - j - 2] = zmul(tau2, acc);
  }
}

// The last kTailLen steps of the reduction in one workgroup of NW wavefronts: the trailing
// matrix (at most 64 x 64) stays in registers, lane r of every wavefront holding row r and
// wavefront q columns 64/NW q .. 64/NW (q + 1) - 1, and each step takes one workgroup barrier instead of
// a launch. Every wavefront keeps the whole of v, w and p in its lanes (lane r holds element
// r) and reduces them itself, in the same order, so all of them hold identical values; a
// wavefront broadcasts v and w to its own columns through a private LDS copy. The pivot
// column of the next step is published by the wavefront that holds it before the barrier
// that closes a step, and every wavefront applies the next step's update to it itself (the
// same arithmetic as the owner's), so the reflector needs no barrier of its own. Takes over
// at step j0 = n - 1 - L0 with the state a step launch would find (rows j0+1.. updated
// through step j0-1, v_j0 in V, tau_j0, p_j0) and writes d, e, V and taus of steps j0 .. n-2
// as the step launches would.
constexpr int kTailLen = 64;
constexpr int kTailWaves = 8;

__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_wave_barrier();
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

template <int K>
__device__ __forceinline__ double2 reg_element(const double2 (&x)[K], int k) {
  double vx = x[0].x, vy = x[0].y;
#pragma unroll
  for (int u = 1; u < K; ++u) {
    vx = k == u ? x[u].x : vx;
    vy = k == u ? x[u].y : vy;
  }
  return make_double2(vx, vy);
}

template <int NW, int RPL>
__global__ __launch_bounds__(64 * NW) void k_hetrd_tail(double2* __restrict__ A, int n, int j0,
                                                        double2* __restrict__ V, double2* __restrict__ taus,
                                                        double* __restrict__ d, double* __restrict__ e,
                                                        const double2* __restrict__ p_in) {
  constexpr int TL = 64 * RPL;  // rows (and columns) held
  constexpr int CW = TL / NW;   // columns per wavefront
  __shared__ double2 vw[NW][2][TL];    // per-wavefront copies of v and w
  __shared__ double2 col[2][TL];       // the next pivot column before its update, by step parity
  __shared__ double2 part[2][NW][TL];  // per-wavefront partial products M v, by step parity
  const int lane = threadIdx.x & 63, wq = threadIdx.x >> 6;
  const int L0 = n - j0 - 1;  // trailing length at step j0, <= TL (checked by the host)
  const int c0 = CW * wq;
  const double2 zero = make_double2(0.0, 0.0);
  double2 m[RPL][CW];  // rows lane + 64 q, columns c0 ..
  double2 v[RPL], p[RPL];
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = lane + 64 * q;
#pragma unroll
    for (int k = 0; k < CW; ++k) {
      const int c = c0 + k;
      double2 a = zero;
      if (r < L0 && c < L0) a = A[(long long)(j0 + 1 + r) * n + j0 + 1 + c];
      m[q][k] = make_double2(a.x, a.y);
    }
    v[q] = zero;
    p[q] = zero;
    if (r < L0) {
      v[q] = V[(long long)j0 * n + r];
      p[q] = p_in[r];
    }
  }
  double2 tau = taus[j0];
  if (wq == 0) {
#pragma unroll
    for (int q = 0; q < RPL; ++q) col[0][lane + 64 * q] = m[q][0];
  }
  __syncthreads();
  for (int s = 0; s < L0; ++s) {
    const int j = j0 + s;
    // step j: v, p hold elements r >= s (trailing index r - s), zero below
    double2 pv = zero;
#pragma unroll
    for (int q = 0; q < RPL; ++q) pv = cadd(pv, zmul(zconj(p[q]), v[q]));
    const double2 K =
        zmul(make_double2(-0.5 * tau.x, -0.5 * tau.y), make_double2(wave_allsum(pv.x), wave_allsum(pv.y)));
    double2 w[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      w[q] = cadd(p[q], zmul(K, v[q]));
      vw[wq][0][lane + 64 * q] = v[q];
      vw[wq][1][lane + 64 * q] = w[q];
    }
    wave_lds_order();
    if (s < c0 + CW) {  // columns below s see v = w = 0 and stay as they are
#pragma unroll
      for (int k = 0; k < CW; ++k) {
        const double2 vc = vw[wq][0][c0 + k], wc = vw[wq][1][c0 + k];
#pragma unroll
        for (int q = 0; q < RPL; ++q)
          m[q][k] = csub(m[q][k], cadd(zmul(v[q], zconj(wc)), zmul(w[q], zconj(vc))));
      }
    }
    // the pivot column s after this step's update, as its owner computes it
    const double2 vs = vw[wq][0][s], ws = vw[wq][1][s];
    double2 xc[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q)
      xc[q] = csub(col[s & 1][lane + 64 * q], cadd(zmul(v[q], zconj(ws)), zmul(w[q], zconj(vs))));
    if (wq == 0 && lane == (s & 63)) d[j + 1] = reg_element(xc, s >> 6).x;
    if (s == L0 - 1) break;
    // zlarfg on the column below the diagonal (rows s+1 ..)
    const double2 xa = reg_element(xc, (s + 1) >> 6);
    const double2 alpha = make_double2(lane_f64(xa.x, (s + 1) & 63), lane_f64(xa.y, (s + 1) & 63));
    double nx = 0.0;
#pragma unroll
    for (int q = 0; q < RPL; ++q)
      if (lane + 64 * q > s + 1) nx += xc[q].x * xc[q].x + xc[q].y * xc[q].y;
    const double xnorm2 = wave_allsum(nx);
    double beta;
    double2 tau2, sc;
    if (xnorm2 == 0.0 && alpha.y == 0.0) {
      beta = alpha.x;
      tau2 = zero;
      sc = zero;
    } else {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
      beta = alpha.x >= 0.0 ? -nrm : nrm;
      const double rb = 1.0 / beta;
      tau2 = make_double2((beta - alpha.x) * rb, -alpha.y * rb);
      const double2 dd = make_double2(alpha.x - beta, alpha.y);
      const double rden = 1.0 / (dd.x * dd.x + dd.y * dd.y);
      sc = make_double2(dd.x * rden, -dd.y * rden);
    }
    double2 vn[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = lane + 64 * q;
      vn[q] = zero;
      if (r > s + 1) vn[q] = zmul(xc[q], sc);
      if (r == s + 1) vn[q] = make_double2(1.0, 0.0);
      if (wq == 0 && r > s && r < L0) V[(long long)(j + 1) * n + r - s - 1] = vn[q];
    }
    if (threadIdx.x == 0) {
      taus[j + 1] = tau2;
      e[j + 1] = beta;
    }
    // publish the next pivot column (after this step's update) for the next step
    if (wq == (s + 1) / CW) {
#pragma unroll
      for (int q = 0; q < RPL; ++q) col[(s + 1) & 1][lane + 64 * q] = reg_element(m[q], (s + 1) % CW);
    }
    // p_{j+1} = tau2 M[s+1:, s+1:] v_{j+1}: this wavefront's columns, then the partials
#pragma unroll
    for (int q = 0; q < RPL; ++q) vw[wq][0][lane + 64 * q] = vn[q];
    wave_lds_order();
    double2 acc[RPL];
#pragma unroll
    for (int q = 0; q < RPL; ++q) acc[q] = zero;
    if (s + 1 < c0 + CW) {  // v_{j+1} is zero at columns up to s
#pragma unroll
      for (int k = 0; k < CW; ++k) {
        const double2 vc = vw[wq][0][c0 + k];
#pragma unroll
        for (int q = 0; q < RPL; ++q) acc[q] = cadd(acc[q], zmul(m[q][k], vc));
      }
    }
#pragma unroll
    for (int q = 0; q < RPL; ++q) part[s & 1][wq][lane + 64 * q] = acc[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = lane + 64 * q;
      double2 tot = part[s & 1][0][r];
#pragma unroll
      for (int u = 1; u < NW; ++u) tot = cadd(tot, part[s & 1][u][r]);
      p[q] = zmul(tau2, tot);
      if (r <= s) p[q] = zero;
      v[q] = vn[q];
    }
    tau = tau2;
  }
}

// The step for 1024 < n <= 2048: 32 row elements per lane take 128 VGPRs, so v and w live
// in LDS only (not also in registers) and the new column reuses v's LDS after one more
// barrier: at most 256 registers and 2n elements of LDS per workgroup (56 KB at n = 1760),
// which lets two workgroups share a CU. With three vectors in LDS and in registers (the
// n <= 1024 form) one workgroup per CU fit, and a 440-workgroup launch ran in two rounds.
__global__ __launch_bounds__(eig::NT, 2) void k_hetrd_step_wide(double2* __restrict__ A, int n, int j,
                                                               double2* __restrict__ V, double2* __restrict__ taus,
                                                               double* __restrict__ d, double* __restrict__ e,
                                                               const double2* __restrict__ p_in,
                                                               double2* __restrict__ p_out) {
  constexpr int RR = eig::MAXN / 64;      // row elements per lane
  constexpr int QV = eig::MAXN / eig::NT;  // vector elements per thread
  double2* sv = hetrd_dyn;      // v_j, then the column j+1 below the diagonal (unscaled)
  double2* sw = hetrd_dyn + n;  // p_j, then w_j
  __shared__ double red[2][2 * (eig::NT / 64)];
  const int len = n - j - 1;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const double2 zero = make_double2(0.0, 0.0);
  const int myrow = j + 2 + blockIdx.x + wave * gridDim.x;
  const bool own = myrow < n;
  double2 rowv[RR];
#pragma unroll
  for (int q = 0; q < RR; ++q) {
    const int c = j + 1 + lane + 64 * q;
    rowv[q] = (own && c < n) ? A[(long long)myrow * n + c] : make_double2(0.0, 0.0);
  }
  double2 piv[QV];
#pragma unroll
  for (int q = 0; q < QV; ++q) {
    const int t = tid + eig::NT * q;
    piv[q] = t < len ? A[(long long)(j + 1) * n + j + 1 + t] : make_double2(0.0, 0.0);
    if (t < len) {
      sv[t] = V[(long long)j * n + t];
      sw[t] = p_in[t];
    }
  }
  const double2 tau = taus[j];
  // K = -(tau/2) p^H v; w = p + K v (each thread on its own elements)
  double2 part = zero;
#pragma unroll
  for (int q = 0; q < QV; ++q) {
    const int t = tid + eig::NT * q;
    if (t < len) {
      const double2 pv = zmul(zconj(sw[t]), sv[t]);
      part.x += pv.x;
      part.y += pv.y;
    }
  }
  const double2 phv = block_sum2_once(part, red[0]);
  const double2 K = zmul(make_double2(-0.5 * tau.x, -0.5 * tau.y), phv);
#pragma unroll
  for (int q = 0; q < QV; ++q) {
    const int t = tid + eig::NT * q;
    if (t < len) {
      const double2 kv = zmul(K, sv[t]);
      sw[t] = make_double2(sw[t].x + kv.x, sw[t].y + kv.y);
    }
  }
  __syncthreads();
  if (own) {
    const int tr = myrow - j - 1;
    const double2 vr = sv[tr], wr = sw[tr];
    double2* row = A + (long long)myrow * n;
#pragma unroll
    for (int q = 0; q < RR; ++q) {
      const int c = j + 1 + lane + 64 * q;
      if (c < n) {
        const int tc = c - j - 1;
        const double2 a = zmul(vr, zconj(sw[tc]));
        const double2 b = zmul(wr, zconj(sv[tc]));
        rowv[q].x -= a.x + b.x;
        rowv[q].y -= a.y + b.y;
        row[c] = rowv[q];
      }
    }
  }
  const double2 v0 = sv[0], w0 = sw[0];
  const int len2 = n - j - 2;
  if (len2 <= 0) {
    if (blockIdx.x == 0 && tid == 0) {
      const double2 a = zmul(v0, zconj(w0));
      d[j + 1] = piv[0].x - 2.0 * a.x;
    }
    return;
  }
  double2 xq[QV];
  double nrm2 = 0.0;
#pragma unroll
  for (int q = 0; q < QV; ++q) {
    const int t = tid + eig::NT * q;
    xq[q] = zero;
    if (t < len) {
      const double2 a = zmul(v0, zconj(sw[t]));
      const double2 b = zmul(w0, zconj(sv[t]));
      const double2 nr = make_double2(piv[q].x - a.x - b.x, piv[q].y - a.y - b.y);
      xq[q] = zconj(nr);
      if (t == 0) {
        if (blockIdx.x == 0) d[j + 1] = nr.x;
      } else if (t >= 2) {
        nrm2 += nr.x * nr.x + nr.y * nr.y;
      }
    }
  }
  __syncthreads();  // every wave has finished reading v: its LDS becomes the new column
#pragma unroll
  for (int q = 0; q < QV; ++q) {
    const int t = tid + eig::NT * q;
    if (t >= 1 && t < len) sv[t - 1] = xq[q];
  }
  const double xnorm2 = block_sum2_once(make_double2(nrm2, 0.0), red[1]).x;
  const double2 alpha = sv[0];
  double beta;
  double2 tau2, sc;
  if (xnorm2 == 0.0 && alpha.y == 0.0) {
    beta = alpha.x;
    tau2 = zero;
    sc = zero;
  } else {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    tau2 = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
    const double2 dd = make_double2(alpha.x - beta, alpha.y);
    const double den = dd.x * dd.x + dd.y * dd.y;
    sc = make_double2(dd.x / den, -dd.y / den);
  }
  if (blockIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < QV; ++q) {
      const int t = tid + eig::NT * q;
      if (t >= 1 && t < len) V[(long long)(j + 1) * n + t - 1] = t == 1 ? make_double2(1.0, 0.0) : zmul(xq[q], sc);
    }
    if (tid == 0) {
      taus[j + 1] = tau2;
      e[j + 1] = beta;
    }
  }
  if (own) {
    double2 a0 = zero, a1 = zero;
#pragma unroll
    for (int q = 0; q < RR; ++q) {
      const int c = j + 1 + lane + 64 * q;
      if (c == j + 2) {
        a0 = rowv[q];
      } else if (c > j + 2 && c < n) {
        const double2 xv = sv[c - j - 2];
        a1.x += rowv[q].x * xv.x - rowv[q].y * xv.y;
        a1.y += rowv[q].x * xv.y + rowv[q].y * xv.x;
      }
    }
    const double2 s1 = zmul(sc, a1);
    double2 acc = make_double2(a0.x + s1.x, a0.y + s1.y);
    acc = make_double2(wave_allsum(acc.x), wave_allsum(acc.y));
    if (lane == 0) p_out[myrow - j - 2] = zmul(tau2, acc);
  }
}

// Z^T[k][:] = Q S[:, k] for k < nev; S column-major (ldc = n), real.
// Back-transformation in compact-WY blocks (LAPACK's zlarft/zlarfb, forward, columnwise):
// the reflectors H_j0 .. H_{j0+kb-1} of one block multiply to I - V T V^H with T upper
// triangular, T[i][i] = tau_i and T[0:i][i] = -tau_i T[0:i][0:i] (V[:, 0:i]^H v_i). Reflector
// j is stored from V[j n] with element t at row j + 1 + t (v_j[0] = 1). Applying a block to
// an eigenvector is two sums over rows and a kBtBlock x kBtBlock triangular product, so the
// serial chain of one eigenvector is n / kBtBlock blocks instead of n - 1 reflectors.
constexpr int kBtBlock = 32;
constexpr int kBtRows = 64;  // rows per partial Gram workgroup
constexpr int kBtVecs = 2;   // eigenvectors per back-transformation workgroup

__host__ __device__ inline int bt_blocks(int n) { return (n - 1 + kBtBlock - 1) / kBtBlock; }
__host__ __device__ inline int bt_chunks(int n) { return (n - 1 + kBtRows - 1) / kBtRows; }

// Partial Gram matrices of the blocks: Gp[b][ch][a][c] = sum over rows r in chunk ch of
// block b (r = j0 + 1 + ch kBtRows + rr) of conj(v_a[r]) v_c[r], grid (chunks, blocks).
__global__ __launch_bounds__(eig::NT) void k_bt_gram(const double2* __restrict__ V, int n,
                                                     double2* __restrict__ Gp) {
  constexpr int B = kBtBlock, RC = kBtRows;
  __shared__ double2 tile[RC][B + 1];
  const int ch = blockIdx.x, b = blockIdx.y, j0 = b * B;
  const int kb = min(B, n - 1 - j0);
  const int r0 = j0 + 1 + ch * RC;
  if (r0 >= n) return;  // past this block's rows (uniform over the workgroup)
  for (int idx = threadIdx.x; idx < RC * B; idx += eig::NT) {
    const int a = idx / RC, rr = idx % RC;  // consecutive threads: consecutive rows of one reflector
    const int r = r0 + rr, t = r - j0 - a - 1;
    tile[rr][a] = (a < kb && t >= 0 && r < n) ? V[(long long)(j0 + a) * n + t] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int a0 = 2 * (threadIdx.x >> 4), c0 = 2 * (threadIdx.x & 15);
  double2 acc[2][2] = {};
  for (int rr = 0; rr < RC; ++rr) {
    const double2 x0 = zconj(tile[rr][a0]), x1 = zconj(tile[rr][a0 + 1]);
    const double2 y0 = tile[rr][c0], y1 = tile[rr][c0 + 1];
    acc[0][0] = cadd(acc[0][0], zmul(x0, y0));
    acc[0][1] = cadd(acc[0][1], zmul(x0, y1));
    acc[1][0] = cadd(acc[1][0], zmul(x1, y0));
    acc[1][1] = cadd(acc[1][1], zmul(x1, y1));
  }
  double2* g = Gp + ((long long)b * bt_chunks(n) + ch) * B * B;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) g[(a0 + i) * B + c0 + j] = acc[i][j];
}

// T of each block from the summed Gram entries and tau (one workgroup per block). The
// recursion runs on one wave with lane a holding row a of T in registers (both loops
// unrolled), reading column i of the Gram matrix from LDS.
__global__ __launch_bounds__(eig::NT) void k_bt_tfactor(const double2* __restrict__ Gp,
                                                        const double2* __restrict__ taus, int n,
                                                        double2* __restrict__ Tf) {
  constexpr int B = kBtBlock;
  __shared__ double2 G[B][B + 1];
  const int b = blockIdx.x, j0 = b * B;
  const int kb = min(B, n - 1 - j0);
  const int nch = (n - 1 - j0 + kBtRows - 1) / kBtRows;
  const double2* gb = Gp + (long long)b * bt_chunks(n) * B * B;
  for (int idx = threadIdx.x; idx < B * B; idx += eig::NT) {
    const int a = idx / B, c = idx % B;
    double2 s = make_double2(0.0, 0.0);
    if (a < c && c < kb)
      for (int ch = 0; ch < nch; ++ch) s = cadd(s, gb[(long long)ch * B * B + idx]);
    G[a][c] = s;
  }
  __syncthreads();
  if (threadIdx.x < B) {
    const int a = threadIdx.x;
    double2 t[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      double2 s = make_double2(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < i; ++c)
        if (c >= a) s = cadd(s, zmul(t[c], G[c][i]));
      const double2 ti = i < kb ? taus[j0 + i] : make_double2(0.0, 0.0);
      const double2 q = zmul(ti, s);
      t[i] = a < i ? make_double2(-q.x, -q.y) : (a == i ? ti : make_double2(0.0, 0.0));
    }
    double2* out = Tf + ((long long)b * B + a) * B;
#pragma unroll
    for (int i = 0; i < B; ++i) out[i] = t[i];
  }
}

// Z = Q S for kBtVecs eigenvectors per workgroup (kept in LDS): for the blocks in reverse
// order, w = V^H z, u = T w, z -= V u. Both row sums issue their global loads in batches
// (a batch of reflector elements per thread for w, a whole row of the block for the
// update), so a block costs a few memory round trips rather than one per row.
constexpr int kBtThreads = 512;
constexpr int kBtSlices = kBtThreads / kBtBlock;  // threads per reflector in the w sums
constexpr int kBtBatch = 8;    // loads per thread in flight in the update's row sums
constexpr int kBtBatchW = 16;  // and in the sums w = V^H z

__global__ __launch_bounds__(kBtThreads) void k_bt_apply(const double2* __restrict__ V,
                                                         const double2* __restrict__ Tf,
                                                         const double* __restrict__ S, int n, int nev,
                                                         double2* __restrict__ Zt) {
  constexpr int B = kBtBlock, E = kBtVecs, NT = kBtThreads, SL = kBtSlices;
  double2* z = hetrd_dyn;  // E x n
  __shared__ double2 w[E][B], u[E][B];
  const int tid = threadIdx.x, e0 = blockIdx.x * E;
  for (int i = tid; i < E * n; i += NT) {
    const int e = i / n, r = i % n;
    z[i] = make_double2(e0 + e < nev ? S[(long long)(e0 + e) * n + r] : 0.0, 0.0);
  }
  __syncthreads();
  for (int b = bt_blocks(n) - 1; b >= 0; --b) {
    const int j0 = b * B;
    const int kb = min(B, n - 1 - j0);
    {
      const int a = tid / SL, sl = tid % SL;  // SL consecutive lanes per reflector
      double2 acc[E];
#pragma unroll
      for (int e = 0; e < E; ++e) acc[e] = make_double2(0.0, 0.0);
      if (a < kb) {
        const double2* v = V + (long long)(j0 + a) * n;
        const double2* zz = z + j0 + a + 1;
        const int len = n - j0 - a - 1;
        for (int t0 = sl; t0 < len; t0 += SL * kBtBatchW) {
          double2 vv[kBtBatchW];
#pragma unroll
          for (int q = 0; q < kBtBatchW; ++q) {
            const int t = t0 + SL * q;
            vv[q] = t < len ? v[t] : make_double2(0.0, 0.0);
          }
#pragma unroll
          for (int q = 0; q < kBtBatchW; ++q) {
            const int t = t0 + SL * q;
            if (t < len) {
              const double2 cv = zconj(vv[q]);
#pragma unroll
              for (int e = 0; e < E; ++e) acc[e] = cadd(acc[e], zmul(cv, zz[e * n + t]));
            }
          }
        }
      }
#pragma unroll
      for (int e = 0; e < E; ++e)
        for (int o = 1; o < SL; o <<= 1) {
          acc[e].x += __shfl_xor(acc[e].x, o, 64);
          acc[e].y += __shfl_xor(acc[e].y, o, 64);
        }
      if (sl == 0)
#pragma unroll
        for (int e = 0; e < E; ++e) w[e][a] = acc[e];
    }
    __syncthreads();
    if (tid < E * B) {
      const int e = tid / B, a = tid % B;
      double2 s = make_double2(0.0, 0.0);
      if (a < kb) {
        const double2* t = Tf + ((long long)b * B + a) * B;
        for (int c = a; c < kb; ++c) s = cadd(s, zmul(t[c], w[e][c]));
      }
      u[e][a] = s;
    }
    __syncthreads();
#pragma unroll 1
    for (int r = j0 + 1 + tid; r < n; r += NT) {
      const int amax = min(kb, r - j0);
      double2 d[E];
#pragma unroll
      for (int e = 0; e < E; ++e) d[e] = make_double2(0.0, 0.0);
#pragma unroll 1
      for (int h = 0; h < B; h += kBtBatch) {
        double2 vv[kBtBatch];
#pragma unroll
        for (int q = 0; q < kBtBatch; ++q)
          vv[q] = h + q < amax ? V[(long long)(j0 + h + q) * n + (r - j0 - h - q - 1)] : make_double2(0.0, 0.0);
#pragma unroll
        for (int e = 0; e < E; ++e)
#pragma unroll
          for (int q = 0; q < kBtBatch; ++q) d[e] = cadd(d[e], zmul(vv[q], u[e][h + q]));
      }
#pragma unroll
      for (int e = 0; e < E; ++e) z[e * n + r] = csub(z[e * n + r], d[e]);
    }
    __syncthreads();
  }
  for (int i = tid; i < E * n; i += NT) {
    const int e = i / n;
    if (e0 + e < nev) Zt[(long long)(e0 + e) * n + i % n] = z[i];
  }
}

// ------------------------------------------------------------ tridiagonal eigenpairs
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Eigenvalue k (0-based, ascending) of T for k < m, one workgroup per eigenvalue: each
// step evaluates the Sturm count at 512 points of the bracket (two independent recurrences
// per thread), so it narrows the bracket 513-fold. scal[0] = the Gershgorin bound of ||T||,
// scal[1] = pivmin (written by eigenvalue 0's workgroup).
__device__ __forceinline__ double recip(double q) {
  // v_rcp_f64 refined by two Newton steps: a correctly signed reciprocal to about 1 ulp,
  // a third of the latency of the IEEE division sequence on the recurrence's critical path
  double r = __builtin_amdgcn_rcp(q);
  r = fma(fma(-q, r, 1.0), r, r);
  return fma(fma(-q, r, 1.0), r, r);
}

__global__ __launch_bounds__(eig::NT) void k_tri_bisect(const double* __restrict__ d, const double* __restrict__ e,
                                                        int n, int m, double* __restrict__ wb,
                                                        double* __restrict__ scal) {
  double* dl = tri_dyn;  // n entries each
  double* e2 = tri_dyn + n;
  __shared__ double red[3 * (eig::NT / 64)];
  __shared__ int first[eig::NT / 64];
  const int k = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (k >= m) return;
  double lo = DBL_MAX, hi = -DBL_MAX, emax2 = 0.0;
  for (int i = tid; i < n; i += eig::NT) {
    const double di = d[i];
    const double ea = i > 0 ? fabs(e[i - 1]) : 0.0, eb = i + 1 < n ? fabs(e[i]) : 0.0;
    dl[i] = di;
    e2[i] = eb * eb;
    emax2 = fmax(emax2, eb * eb);
    lo = fmin(lo, di - ea - eb);
    hi = fmax(hi, di + ea + eb);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, off, 64));
    hi = fmax(hi, __shfl_xor(hi, off, 64));
    emax2 = fmax(emax2, __shfl_xor(emax2, off, 64));
  }
  if (lane == 0) {
    red[3 * wave] = lo;
    red[3 * wave + 1] = hi;
    red[3 * wave + 2] = emax2;
  }
  __syncthreads();
  for (int w = 0; w < eig::NT / 64; ++w) {
    lo = fmin(lo, red[3 * w]);
    hi = fmax(hi, red[3 * w + 1]);
    emax2 = fmax(emax2, red[3 * w + 2]);
  }
  const double tnorm = fmax(fabs(lo), fabs(hi));
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  const double pad = 2.0 * n * DBL_EPSILON * tnorm + 4.0 * pivmin;
  lo -= pad;
  hi += pad;
  if (k == 0 && tid == 0) {
    scal[0] = tnorm;
    scal[1] = pivmin;
  }
  constexpr int P = 2 * eig::NT;  // points per step
  for (int it = 0; it < 32; ++it) {
    if (hi - lo <= 2.0 * DBL_EPSILON * tnorm + 4.0 * pivmin) break;  // dstebz's default accuracy
    const double width = hi - lo;
    const double x0 = lo + width * (double)(2 * tid + 1) / (double)(P + 1);
    const double x1 = lo + width * (double)(2 * tid + 2) / (double)(P + 1);
    double q0 = dl[0] - x0, q1 = dl[0] - x1;
    if (fabs(q0) < pivmin) q0 = -pivmin;
    if (fabs(q1) < pivmin) q1 = -pivmin;
    int c0 = q0 < 0.0, c1 = q1 < 0.0;
    auto sturm = [&](double di, double ei) {
      q0 = (di - x0) - ei * recip(q0);
      q1 = (di - x1) - ei * recip(q1);
      if (fabs(q0) < pivmin) q0 = -pivmin;
      if (fabs(q1) < pivmin) q1 = -pivmin;
      c0 += q0 < 0.0;
      c1 += q1 < 0.0;
    };
    int i = 1;
    for (; i + kTriChunk <= n; i += kTriChunk) {  // the (broadcast) LDS loads of a chunk issued at once
      double dd[kTriChunk], ee[kTriChunk];
#pragma unroll
      for (int u = 0; u < kTriChunk; ++u) {
        dd[u] = dl[i + u];
        ee[u] = e2[i - 1 + u];
      }
#pragma unroll
      for (int u = 0; u < kTriChunk; ++u) sturm(dd[u], ee[u]);
    }
    for (; i < n; ++i) sturm(dl[i], e2[i - 1]);
    // the first point (0-based j, x_j = lo + width (j+1)/(P+1)) whose count exceeds k
    int f = c0 >= k + 1 ? 2 * tid : (c1 >= k + 1 ? 2 * tid + 1 : P);
    for (int off = 32; off > 0; off >>= 1) f = min(f, __shfl_xor(f, off, 64));
    __syncthreads();
    if (lane == 0) first[wave] = f;
    __syncthreads();
    for (int w = 0; w < eig::NT / 64; ++w) f = min(f, first[w]);
    const double nlo = f > 0 ? lo + width * (double)f / (double)(P + 1) : lo;
    const double nhi = f < P ? lo + width * (double)(f + 1) / (double)(P + 1) : hi;
    lo = nlo;
    hi = nhi;
  }
  if (tid == 0) wb[k] = 0.5 * (lo + hi);
}

__device__ __forceinline__ double start_value(unsigned long long key) {
  key += 0x9E3779B97F4A7C15ull;  // splitmix64
  key = (key ^ (key >> 30)) * 0xBF58476D1CE4E5B9ull;
  key = (key ^ (key >> 27)) * 0x94D049BB133111EBull;
  key ^= key >> 31;
  return (double)(key >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Factors of T - lam I in LDS, 5 arrays of n: U row i = (1/u0, u1, u2) on columns i, i+1,
// i+2 (the reciprocal of the pivot, floored at +-tiny), multiplier l[i], and sw[i] != 0
// when rows i and i+1 were swapped (partial pivoting, as LAPACK's dlagtf).
struct TriLU {
  double *rinv, *u1, *u2, *l, *sw;
};

// Both routines below are serial chains run by one thread over LDS-resident arrays. The chain
// itself lives in registers (a, b in the factorisation; the running element in the solves);
// every load the chain does not depend on is issued kTriChunk steps ahead.
__device__ void tri_factor(const double* __restrict__ dl, const double* __restrict__ el, int n, double lam,
                           double tiny, const TriLU& f) {
  double a = dl[0] - lam, b = n > 1 ? el[0] : 0.0;
  int i = 0;
  auto step = [&](int ii, double s, double dd, double g) {
    double p;
    if (fabs(s) > fabs(a)) {
      const double mlt = a / s;
      p = s;
      f.u1[ii] = dd;
      f.u2[ii] = g;
      f.l[ii] = mlt;
      f.sw[ii] = 1.0;
      a = b - mlt * dd;
      b = -mlt * g;
    } else {
      const double mlt = a != 0.0 ? s / a : 0.0;
      p = a;
      f.u1[ii] = b;
      f.u2[ii] = 0.0;
      f.l[ii] = mlt;
      f.sw[ii] = 0.0;
      a = dd - mlt * b;
      b = g;
    }
    if (fabs(p) < tiny) p = p < 0.0 ? -tiny : tiny;
    f.rinv[ii] = 1.0 / p;
  };
  for (; i + kTriChunk + 1 < n; i += kTriChunk) {  // steps i..i+7 read el[i..i+8], dl[i+1..i+8]
    double s[kTriChunk], dd[kTriChunk], g[kTriChunk];
#pragma unroll
    for (int u = 0; u < kTriChunk; ++u) {
      s[u] = el[i + u];
      dd[u] = dl[i + 1 + u] - lam;
      g[u] = el[i + 1 + u];
    }
#pragma unroll
    for (int u = 0; u < kTriChunk; ++u) step(i + u, s[u], dd[u], g[u]);
  }
  for (; i + 1 < n; ++i) step(i, el[i], dl[i + 1] - lam, i + 2 < n ? el[i + 1] : 0.0);
  if (fabs(a) < tiny) a = a < 0.0 ? -tiny : tiny;
  f.rinv[n - 1] = 1.0 / a;
  f.u1[n - 1] = 0.0;
  f.u2[n - 1] = 0.0;
  f.l[n - 1] = 0.0;
  f.sw[n - 1] = 0.0;
}

// x <- (T - lam I)^{-1} x with the factors.
__device__ void tri_solve(const TriLU& f, int n, double* __restrict__ x) {
  // forward elimination: step i reads the original x[i+1] and the running x[i] (cur) and
  // writes the final x[i]; cur carries the new x[i+1]
  double cur = x[0];
  int i = 0;
  for (; i + kTriChunk < n; i += kTriChunk) {
    double o[kTriChunk], l[kTriChunk], sw[kTriChunk], out[kTriChunk];
#pragma unroll
    for (int u = 0; u < kTriChunk; ++u) {
      o[u] = x[i + 1 + u];
      l[u] = f.l[i + u];
      sw[u] = f.sw[i + u];
    }
#pragma unroll
    for (int u = 0; u < kTriChunk; ++u) {
      if (sw[u] != 0.0) {
        out[u] = o[u];
        cur = cur - l[u] * o[u];
      } else {
        out[u] = cur;
        cur = o[u] - l[u] * cur;
      }
    }
#pragma unroll
    for (int u = 0; u < kTriChunk; ++u) x[i + u] = out[u];
  }
  for (; i + 1 < n; ++i) {
    const double o = x[i + 1];
    if (f.sw[i] != 0.0) {
      x[i] = o;
      cur = cur - f.l[i] * o;
    } else {
      x[i] = cur;
      cur = o - f.l[i] * cur;
    }
  }
  x[n - 1] = cur;
  // back substitution from the end, x1 = x[i+1], x2 = x[i+2] in registers
  double x1 = 0.0, x2 = 0.0;
  int j = n - 1;
  for (; j - kTriChunk + 1 >= 0; j -= kTriChunk) {
    double xv[kTriChunk], u1[kTriChunk], u2[kTriChunk], ri[kTriChunk];
#pragma unroll
    for (int u = 0; u < kTriChunk; ++u) {
      xv[u] = x[j - u];
      u1[u] = f.u1[j - u];
      u2[u] = f.u2[j - u];
      ri[u] = f.rinv[j - u];
    }
#pragma unroll
    for (int u = 0; u < kTriChunk; ++u) {
      const double v = (xv[u] - u1[u] * x1 - u2[u] * x2) * ri[u];
      xv[u] = v;
      x2 = x1;
      x1 = v;
    }
#pragma unroll
    for (int u = 0; u < kTriChunk; ++u) x[j - u] = xv[u];
  }
  for (; j >= 0; --j) {
    const double v = (x[j] - f.u1[j] * x1 - f.u2[j] * x2) * f.rinv[j];
    x[j] = v;
    x2 = x1;
    x1 = v;
  }
}

// One inverse-iteration round for row r < m of S: round 0 draws the pseudo-random start
// and factors T - w_r I (factors kept in work + 5 n r for the later rounds); every round
// scales the row by tiny (it is unit-norm from the previous round) and solves once.
__global__ __launch_bounds__(64) void k_tri_inverse(const double* __restrict__ d, const double* __restrict__ e,
                                                    int n, int m, int round, const double* __restrict__ wb,
                                                    const double* __restrict__ scal, double* __restrict__ S,
                                                    double* __restrict__ work) {
  // dynamic LDS of 8 n doubles: the factors, the vector and T
  const TriLU f{tri_dyn, tri_dyn + n, tri_dyn + 2 * n, tri_dyn + 3 * n, tri_dyn + 4 * n};
  double* x = tri_dyn + 5 * n;
  double* dl = tri_dyn + 6 * n;
  double* el = tri_dyn + 7 * n;
  const int r = blockIdx.x, tid = threadIdx.x;
  if (r >= m) return;
  const double tnorm = scal[0];
  const double tiny = fmax(DBL_EPSILON * tnorm, DBL_MIN);
  double* row = S + (long long)r * n;
  double* g = work + 5LL * n * r;
  if (round == 0) {
    for (int i = tid; i < n; i += 64) {
      dl[i] = d[i];
      el[i] = i + 1 < n ? e[i] : 0.0;
      x[i] = tiny * start_value(((unsigned long long)r << 32) ^ (unsigned long long)i);
    }
    __syncthreads();
    if (tid == 0) tri_factor(dl, el, n, wb[r], tiny, f);
    __syncthreads();
    for (int i = tid; i < n; i += 64) {
      g[i] = f.rinv[i];
      g[n + i] = f.u1[i];
      g[2 * n + i] = f.u2[i];
      g[3 * n + i] = f.l[i];
      g[4 * n + i] = f.sw[i];
    }
  } else {
    for (int i = tid; i < n; i += 64) {
      f.rinv[i] = g[i];
      f.u1[i] = g[n + i];
      f.u2[i] = g[2 * n + i];
      f.l[i] = g[3 * n + i];
      f.sw[i] = g[4 * n + i];
      x[i] = tiny * row[i];
    }
    __syncthreads();
  }
  if (tid == 0) tri_solve(f, n, x);
  __syncthreads();
  for (int i = tid; i < n; i += 64) row[i] = x[i];
}

struct TriShared {
  double K[eig::kClusterMax][eig::kClusterMax + 1];
  double Q[eig::kClusterMax][eig::kClusterMax + 1];
  double dots[2 * eig::kClusterMax];  // a piece's rows and the previous piece's
  double red[eig::NT / 64 + 1];
  int order[eig::kClusterMax];
  int flag;
};

// Workgroup-wide sum of a double (all threads get it).
__device__ __forceinline__ double group_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < eig::NT / 64; ++w) s += red[w];
  return s;
}

// Orthonormalise row t of X against rows 0..t-1 (classical Gram-Schmidt, twice) and
// normalise it. Returns false when the row is numerically dependent on the earlier ones.
__device__ bool orthonormalize_row(double* X, int n, int t, TriShared& sh) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* xt = X + (long long)t * n;
  double before = 0.0;
  for (int i = threadIdx.x; i < n; i += eig::NT) before += xt[i] * xt[i];
  before = group_sum(before, sh.red);
  for (int rep = 0; rep < 2 && t > 0; ++rep) {
    for (int j = wave; j < t; j += eig::NT / 64) {
      const double* xj = X + (long long)j * n;
      double s = 0.0;
      for (int i = lane; i < n; i += 64) s += xj[i] * xt[i];
      s = wave_sum(s);
      if (lane == 0) sh.dots[j] = s;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += eig::NT) {
      double v = xt[i];
      for (int j = 0; j < t; ++j) v -= sh.dots[j] * X[(long long)j * n + i];
      xt[i] = v;
    }
    __syncthreads();
  }
  double nrm2 = 0.0;
  for (int i = threadIdx.x; i < n; i += eig::NT) nrm2 += xt[i] * xt[i];
  nrm2 = group_sum(nrm2, sh.red);
  if (!(nrm2 > 1e-24 * before) || !(nrm2 > 0.0)) return false;
  const double sc = 1.0 / sqrt(nrm2);
  for (int i = threadIdx.x; i < n; i += eig::NT) xt[i] *= sc;
  __syncthreads();
  return true;
}

// out[a][b] = sum_i A[a n + i] B[b n + i] for a < na, b < nb (both <= kClusterMax): slices of
// kRRCols columns of both staged in LDS, every thread accumulating its own entries.
__device__ void slice_dots(const double* A, int na, const double* B, int nb, int n,
                           double (*out)[eig::kClusterMax + 1]) {
  constexpr int PPT = (eig::kClusterMax * eig::kClusterMax + eig::NT - 1) / eig::NT;
  const int tid = threadIdx.x, npair = na * nb;
  int pa[PPT], pb[PPT];
  double acc[PPT];
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    const int pr = tid + eig::NT * q;
    pa[q] = pr / nb;
    pb[q] = pr - pa[q] * nb;
    acc[q] = 0.0;
  }
  double* xs = tri_dyn;  // [kClusterMax][kRRCols + 1]
  double* ys = tri_dyn + eig::kClusterMax * (kRRCols + 1);
  for (int i0 = 0; i0 < n; i0 += kRRCols) {
    __syncthreads();
    for (int idx = tid; idx < na * kRRCols; idx += eig::NT) {
      const int t = idx / kRRCols, u = idx - t * kRRCols, i = i0 + u;
      xs[t * (kRRCols + 1) + u] = i < n ? A[(long long)t * n + i] : 0.0;
    }
    for (int idx = tid; idx < nb * kRRCols; idx += eig::NT) {
      const int t = idx / kRRCols, u = idx - t * kRRCols, i = i0 + u;
      ys[t * (kRRCols + 1) + u] = i < n ? B[(long long)t * n + i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      if (tid + eig::NT * q < npair) {
        const double* xa = xs + pa[q] * (kRRCols + 1);
        const double* yb = ys + pb[q] * (kRRCols + 1);
        double sum = 0.0;
#pragma unroll 8
        for (int u = 0; u < kRRCols; ++u) sum += xa[u] * yb[u];
        acc[q] += sum;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PPT; ++q)
    if (tid + eig::NT * q < npair) out[pa[q]][pb[q]] = acc[q];
  __syncthreads();
}

// The rows P of one piece (c rows at X + off n) made orthonormal and orthogonal to the rows
// X[0, off) of the previous piece, as a block: the projection off the previous piece twice,
// then CholeskyQR twice (Gram matrix, Cholesky factor L and L^{-1} in LDS on one wavefront,
// P <- L^{-1} P in place). The same rows Gram-Schmidt would give, in exact arithmetic. Uses
// no global scratch: the workspace rows beside the piece hold the LU factors that the next
// inverse-iteration round reads.
// Returns false when a Cholesky pivot shows a row within 1e-7 (relative) of the span of the
// rows before it; the caller then takes the row-by-row route, which restarts such rows.
__device__ bool orthonormalize_block(double* X, int n, int off, int c, TriShared& sh) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  double* P = X + (long long)off * n;
  for (int rep = 0; rep < 2 && off > 0; ++rep) {
    slice_dots(P, c, X, off, n, sh.K);  // D = P Q^T
    for (int idx = tid; idx < c * n; idx += eig::NT) {
      const int a = idx / n, i = idx - a * n;
      double v = P[idx];
      for (int b = 0; b < off; ++b) v -= sh.K[a][b] * X[(long long)b * n + i];
      P[idx] = v;
    }
    __syncthreads();
  }
  for (int rep = 0; rep < 2; ++rep) {
    slice_dots(P, c, P, c, n, sh.K);  // G = P P^T
    if (wave == 0) {  // lane r owns row r: Cholesky in the lower triangle of K, L^{-1} into Q
      if (lane < c) sh.dots[lane] = sh.K[lane][lane];
      bool good = true;
      wave_lds_order();
      for (int k = 0; k < c && good; ++k) {
        const double gkk = sh.K[k][k];
        if (!(gkk > 1e-14 * sh.dots[k])) {
          good = false;
          break;
        }
        const double lkk = sqrt(gkk);
        wave_lds_order();
        if (lane > k && lane < c) sh.K[lane][k] /= lkk;
        if (lane == k) sh.K[k][k] = lkk;
        wave_lds_order();
        if (lane > k && lane < c) {
          const double lrk = sh.K[lane][k];
          for (int m = k + 1; m <= lane; ++m) sh.K[lane][m] -= lrk * sh.K[m][k];
        }
        wave_lds_order();
      }
      if (good) {
        for (int r = 0; r < c; ++r) {  // row r of L^{-1}: lane k < r from rows m < r
          const double lrr = sh.K[r][r];
          if (lane < r) {
            double sum = 0.0;
            for (int m = lane; m < r; ++m) sum += sh.K[r][m] * sh.Q[m][lane];
            sh.Q[r][lane] = -sum / lrr;
          } else if (lane == r) {
            sh.Q[r][r] = 1.0 / lrr;
          }
          wave_lds_order();
        }
      }
      if (lane == 0) sh.flag = good ? 1 : 0;
    }
    __syncthreads();
    if (!sh.flag) return false;
    // P <- L^{-1} P in place: each thread owns whole columns and, L^{-1} being lower
    // triangular, overwrites row a only after the rows above it no longer need it (a
    // descending)
    for (int i = tid; i < n; i += eig::NT) {
      for (int a = c - 1; a >= 0; --a) {
        double v = 0.0;
        for (int b = 0; b <= a; ++b) v += sh.Q[a][b] * P[(long long)b * n + i];
        P[(long long)a * n + i] = v;
      }
    }
    __syncthreads();
  }
  return true;
}

// One piece of a chain of close eigenvalues, rows [ps, ps + c) of S, in one workgroup. X
// points at the row `off` before the piece: rows 0..off-1 of X are the previous piece of the
// same chain (already orthonormal), so the Gram-Schmidt pass makes the piece orthogonal to
// them as well as orthonormal in itself. After the last round a Rayleigh-Ritz step inside
// the piece makes its rows eigenvectors and writes their eigenvalues w[ps + t] (< nev).
__device__ void solve_piece(const double* __restrict__ d, const double* __restrict__ e, int n, int nev, int round,
                            int last, const double* __restrict__ wb, double tnorm, double* __restrict__ w,
                            double* X, int off, int ps, int c, double* __restrict__ tmp, TriShared& sh) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const bool block_done = c > 1 && orthonormalize_block(X, n, off, c, sh);
  for (int t = off; t < off + c && !block_done; ++t) {
    for (int attempt = 0; attempt < 4; ++attempt) {
      if (orthonormalize_row(X, n, t, sh)) break;
      // dependent on the earlier rows: restart this row from fresh pseudo-random values
      double* xt = X + (long long)t * n;
      for (int i = tid; i < n; i += eig::NT)
        xt[i] = start_value(((unsigned long long)(ps + t - off) << 32) ^ (unsigned long long)i ^
                            ((unsigned long long)(attempt + 1 + 4 * round) << 56));
      __syncthreads();
    }
  }
  if (!last) return;
  if (c == 1) {
    if (tid == 0 && ps < nev) w[ps] = wb[ps];
    __syncthreads();
    return;
  }
  double* P = X + (long long)off * n;  // the piece's own rows
  // Rayleigh-Ritz inside the piece: K = P (T P^T) (c x c). The rows of T P^T go to tmp (free
  // until the rotation below); then slices of kRRCols columns of P and T P^T are staged in
  // LDS and every thread accumulates its own entries (a <= b) over each slice.
  for (int idx = tid; idx < c * n; idx += eig::NT) {
    const int t = idx / n, i = idx - t * n;
    const double* xb = P + (long long)t * n;
    double y = d[i] * xb[i];
    if (i > 0) y += e[i - 1] * xb[i - 1];
    if (i + 1 < n) y += e[i] * xb[i + 1];
    tmp[idx] = y;
  }
  constexpr int PPT = (eig::kClusterMax * (eig::kClusterMax + 1) / 2 + eig::NT - 1) / eig::NT;
  const int npair = c * (c + 1) / 2;
  int pa[PPT], pb[PPT];
  double kacc[PPT];
#pragma unroll
  for (int q = 0; q < PPT; ++q) {
    int a = 0, rem = tid + eig::NT * q;
    while (a < c && rem >= c - a) {
      rem -= c - a;
      ++a;
    }
    pa[q] = a;
    pb[q] = a + rem;
    kacc[q] = 0.0;
  }
  double* xs = tri_dyn;                               // [kClusterMax][kRRCols + 1]
  double* ys = tri_dyn + eig::kClusterMax * (kRRCols + 1);
  for (int i0 = 0; i0 < n; i0 += kRRCols) {
    __syncthreads();
    for (int idx = tid; idx < c * kRRCols; idx += eig::NT) {
      const int t = idx / kRRCols, u = idx - t * kRRCols, i = i0 + u;
      xs[t * (kRRCols + 1) + u] = i < n ? P[(long long)t * n + i] : 0.0;
      ys[t * (kRRCols + 1) + u] = i < n ? tmp[(long long)t * n + i] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
      if (tid + eig::NT * q < npair) {
        const double* xa = xs + pa[q] * (kRRCols + 1);
        const double* yb = ys + pb[q] * (kRRCols + 1);
        double acc = 0.0;
#pragma unroll 8
        for (int u = 0; u < kRRCols; ++u) acc += xa[u] * yb[u];
        kacc[q] += acc;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < PPT; ++q)
    if (tid + eig::NT * q < npair) {
      sh.K[pa[q]][pb[q]] = kacc[q];
      sh.K[pb[q]][pa[q]] = kacc[q];
    }
  for (int idx = tid; idx < c * c; idx += eig::NT) sh.Q[idx / c][idx % c] = (idx / c == idx % c) ? 1.0 : 0.0;
  __syncthreads();
  if (wave == 0) {  // cyclic Jacobi, lane r owns row r of K and Q
    const double skip = 1e-18 * tnorm;
    for (int sweep = 0; sweep < 30; ++sweep) {
      bool rotated = false;
      for (int p = 0; p < c - 1; ++p) {
        for (int q = p + 1; q < c; ++q) {
          const double apq = sh.K[p][q];
          if (fabs(apq) <= skip) continue;
          const double app = sh.K[p][p], aqq = sh.K[q][q];
          const double theta = (aqq - app) / (2.0 * apq);
          const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
          const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
          const int r = lane;
          double krp = 0.0, krq = 0.0, qrp = 0.0, qrq = 0.0;
          if (r < c) {
            krp = sh.K[r][p];
            krq = sh.K[r][q];
            qrp = sh.Q[r][p];
            qrq = sh.Q[r][q];
          }
          __builtin_amdgcn_wave_barrier();
          if (r < c) {
            if (r != p && r != q) {
              const double np_ = cs * krp - sn * krq, nq = sn * krp + cs * krq;
              sh.K[r][p] = np_;
              sh.K[p][r] = np_;
              sh.K[r][q] = nq;
              sh.K[q][r] = nq;
            }
            sh.Q[r][p] = cs * qrp - sn * qrq;
            sh.Q[r][q] = sn * qrp + cs * qrq;
          }
          if (r == 0) {
            sh.K[p][p] = app - t * apq;
            sh.K[q][q] = aqq + t * apq;
            sh.K[p][q] = 0.0;
            sh.K[q][p] = 0.0;
          }
          __builtin_amdgcn_wave_barrier();
          rotated = true;
        }
      }
      if (!rotated) break;
    }
    if (lane == 0) {  // ascending order of the Ritz values
      for (int t = 0; t < c; ++t) sh.order[t] = t;
      for (int t = 0; t < c; ++t) {
        int best = t;
        for (int u = t + 1; u < c; ++u)
          if (sh.K[sh.order[u]][sh.order[u]] < sh.K[sh.order[best]][sh.order[best]]) best = u;
        const int tmpo = sh.order[t];
        sh.order[t] = sh.order[best];
        sh.order[best] = tmpo;
      }
    }
  }
  __syncthreads();
  // rotated vectors into the (no longer needed) factor storage of the piece's rows, then back
  for (int idx = tid; idx < c * n; idx += eig::NT) {
    const int t = idx / n, i = idx - t * n;
    const int o = sh.order[t];
    double v = 0.0;
    for (int a = 0; a < c; ++a) v += sh.Q[a][o] * P[(long long)a * n + i];
    tmp[idx] = v;
  }
  __syncthreads();
  for (int idx = tid; idx < c * n; idx += eig::NT) P[idx] = tmp[idx];
  for (int t = tid; t < c; t += eig::NT)
    if (ps + t < nev) w[ps + t] = sh.K[sh.order[t]][sh.order[t]];
  __syncthreads();
}

// After each inverse-iteration round: workgroup k leads the chain of close eigenvalues that
// starts at k (if one does; an isolated eigenvalue is a chain of one). A chain of at most
// kClusterMax members is one piece. A longer chain is cut into pieces of at most kClusterMax,
// each cut placed at the largest gap inside the next kClusterMax members, and the pieces are
// solved in order by this workgroup, each orthogonalised against the piece before it, so
// (near-)degenerate eigenvalues never straddle a cut and neighbouring pieces stay orthogonal.
__global__ __launch_bounds__(eig::NT) void k_tri_cluster(const double* __restrict__ d, const double* __restrict__ e,
                                                         int n, int m, int nev, int round, int last,
                                                         const double* __restrict__ wb,
                                                         const double* __restrict__ scal, double* __restrict__ w,
                                                         double* __restrict__ S, double* __restrict__ work) {
  __shared__ TriShared sh;
  const int k = blockIdx.x;
  if (k >= nev) return;
  const double tnorm = scal[0];
  const double delta = eig::kClusterGap * tnorm;
  if (k > 0 && wb[k] - wb[k - 1] < delta) return;  // not the first member of a chain
  int ce = k + 1;
  while (ce < m && wb[ce] - wb[ce - 1] < delta) ++ce;
  int prev = k;  // first row of the previous piece (= ps for the first piece)
  for (int ps = k; ps < ce;) {
    int pe = ce;
    if (ce - ps > eig::kClusterMax) {  // cut before the largest gap in (ps, ps + kClusterMax]
      pe = ps + 1;
      double gbest = -1.0;
      for (int q = ps + 1; q <= ps + eig::kClusterMax; ++q) {
        const double g = wb[q] - wb[q - 1];
        if (g >= gbest) {  // ties go to the later gap: the largest piece
          gbest = g;
          pe = q;
        }
      }
    }
    solve_piece(d, e, n, nev, round, last, wb, tnorm, w, S + (long long)prev * n, ps - prev, ps, pe - ps,
                work + 5LL * n * ps, sh);
    prev = ps;
    ps = pe;
  }
}

// ------------------------------------------------------------------ host
// Workgroups of the tridiagonalisation: one trailing row per wave at most (n - 2 <= 4 G).
static int hetrd_groups(int n) {
  int g = (n - 2 + 3) / 4;
  if (g < 1) g = 1;
  if (g > 512) g = 512;
  return g;
}

static void enqueue_tridiag(hipStream_t s, int n, double2* A, double2* V, double2* taus, double* d, double* e,
                            double2* p0, double2* p1) {
  const int G = hetrd_groups(n);
  hipLaunchKernelGGL(k_hetrd_start, dim3(G), dim3(eig::NT), 3 * (size_t)n * sizeof(double2), s, A, n, V, taus,
                     d, e, p0);
  // Step j touches the trailing rows j+2 .. n-1 and columns j+1 .. n-1 only: its grid covers
  // those rows (one per wave) and its row template the trailing length, so the late, short
  // steps launch few workgroups with short loops.
  const size_t lds3 = 3 * (size_t)n * sizeof(double2), lds2 = 2 * (size_t)n * sizeof(double2);
  for (int j = 0; j <= n - 2; ++j) {
    const double2* pin = (j & 1) ? p1 : p0;
    double2* pout = (j & 1) ? p0 : p1;
    const int len = n - j - 1, rows = n - j - 2;
    if (len <= kTailLen) {  // steps j .. n-2 in one launch
      hipLaunchKernelGGL((k_hetrd_tail<kTailWaves, kTailLen / 64>), dim3(1), dim3(64 * kTailWaves), 0, s, A, n, j, V,
                         taus, d, e, pin);
      break;
    }
    const dim3 grid(rows > 0 ? (rows + 3) / 4 : 1);
    if (len > eig::MAXN / 2)
      hipLaunchKernelGGL(k_hetrd_step_wide, grid, dim3(eig::NT), lds2, s, A, n, j, V, taus, d, e, pin, pout);
    else if (len > 768)
      hipLaunchKernelGGL(k_hetrd_step<16>, grid, dim3(eig::NT), lds3, s, A, n, j, V, taus, d, e, pin, pout);
    else if (len > 512)
      hipLaunchKernelGGL(k_hetrd_step<12>, grid, dim3(eig::NT), lds3, s, A, n, j, V, taus, d, e, pin, pout);
    else if (len > 64 * kW64Max)
      hipLaunchKernelGGL(k_hetrd_step<8>, grid, dim3(eig::NT), lds3, s, A, n, j, V, taus, d, e, pin, pout);
    else if (len > 256)
      hipLaunchKernelGGL(k_hetrd_step_w64<kW64Max>, dim3(rows > 0 ? rows : 1), dim3(64), 0, s, A, n, j, V, taus, d,
                         e, pin, pout);
    else
      hipLaunchKernelGGL(k_hetrd_step_w64<4>, dim3(rows > 0 ? rows : 1), dim3(64), 0, s, A, n, j, V, taus, d, e,
                         pin, pout);
  }
}

size_t heev_workspace_bytes(int n) {
  // A copy, V, p (2), taus, S (n x n real), d, e, bisection values, 4 scalars, LU factors (5 n x n),
  // partial Gram matrices and T factors of the back-transformation blocks
  const size_t bb = n >= 2 ? (size_t)bt_blocks(n) : 0, bc = n >= 2 ? (size_t)bt_chunks(n) : 0;
  return (size_t)n * n * 16 * 2 + (size_t)2 * n * 16 + (size_t)n * 16 + (size_t)n * n * 8 + (size_t)3 * n * 8 + 32 +
         (size_t)5 * n * n * 8 + 16 + bb * (bc + 1) * kBtBlock * kBtBlock * 16;
}

int launch_heev_lowest(int n, const void* a, int nev, double* w, void* zt, void* ws, size_t ws_bytes,
                       hipStream_t s) {
  if (n < 2 || n > eig::MAXN || nev < 1 || nev > n) {
    set_error("heev: need 2 <= n <= 2048 and 1 <= nev <= n");
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
  double* wb = e + n;
  double* scal = wb + n;
  double* lu = scal + 4;
  if (hipMemcpyAsync(A, a, (size_t)n * n * 16, hipMemcpyDeviceToDevice, s) != hipSuccess) {
    set_error("heev: copy of the input failed");
    return 1;
  }
  enqueue_tridiag(s, n, A, V, taus, d, e, p0, p1);
  if (hipGetLastError() != hipSuccess) {
    set_error("heev: tridiagonalisation launch failed");
    return 1;
  }
  const int m = nev + eig::kClusterMax < n ? nev + eig::kClusterMax : n;
  hipLaunchKernelGGL(k_tri_bisect, dim3(m), dim3(eig::NT), 2 * (size_t)n * sizeof(double), s, d, e, n, m, wb,
                     scal);
  for (int round = 0; round < eig::kInverseRounds; ++round) {
    hipLaunchKernelGGL(k_tri_inverse, dim3(m), dim3(64), 8 * (size_t)n * sizeof(double), s, d, e, n, m, round,
                       wb, scal, S, lu);
    hipLaunchKernelGGL(k_tri_cluster, dim3(nev), dim3(eig::NT), 2 * eig::kClusterMax * (kRRCols + 1) * sizeof(double),
                       s, d, e, n, m, nev, round,
                       round + 1 == eig::kInverseRounds ? 1 : 0, wb, scal, w, S, lu);
  }
  if (hipGetLastError() != hipSuccess) {
    set_error("heev: tridiagonal eigenpair launch failed");
    return 1;
  }
  double2* Gp = reinterpret_cast<double2*>((reinterpret_cast<uintptr_t>(lu + (size_t)5 * n * n) + 15) & ~uintptr_t(15));
  double2* Tf = Gp + (size_t)bt_blocks(n) * bt_chunks(n) * kBtBlock * kBtBlock;
  hipLaunchKernelGGL(k_bt_gram, dim3(bt_chunks(n), bt_blocks(n)), dim3(eig::NT), 0, s, V, n, Gp);
  hipLaunchKernelGGL(k_bt_tfactor, dim3(bt_blocks(n)), dim3(eig::NT), 0, s, Gp, taus, n, Tf);
  hipLaunchKernelGGL(k_bt_apply, dim3((nev + kBtVecs - 1) / kBtVecs), dim3(kBtThreads),
                     (size_t)kBtVecs * n * sizeof(double2), s, V, Tf, S, n, nev, static_cast<double2*>(zt));
  if (hipGetLastError() != hipSuccess) {
    set_error("heev: back-transformation launch failed");
    return 1;
  }
  return 0;
}

}  // namespace amt
