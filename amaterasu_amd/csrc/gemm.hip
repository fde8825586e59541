// Complex GEMM on the CDNA matrix cores (MFMA), FP32 and FP64.
//
// C = alpha * A op(B) + beta * C with row-major interleaved-complex operands and real
// alpha, beta. op = 0 (NN): B is K x N. op = 1 (NC): B is N x K and conj(B)^T is
// used -- the projector contraction <beta|psi> and every subspace overlap is of this
// form, with the reduction along the contiguous plane-wave axis of both operands.
//
// Complex arithmetic runs on the real matrix cores with Gauss's three-product form:
//   P1 += A_re B_re,  P2 += A_im B_im,  P3 += (A_re + A_im)(B_re + B_im)
//   C_re = P1 - P2,   C_im = P3 - P1 - P2
// i.e. three v_mfma_*_16x16x4 per 16x16x4 complex step instead of four (measured 12-27 %
// faster on every projector and subspace shape, with the same error against an FP64
// product: 2-6e-7 relative in FP32).
// The FP32 and FP64 instructions share one operand layout, so one template serves
// both on the input side: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]. The
// accumulator differs (measured on gfx950): accumulator element q of lane l is
// D[4 (l / 16) + q][l % 16] for FP32 and D[l / 16 + 4 q][l % 16] for FP64.
//
// Tiling: a 256-thread workgroup (4 wavefronts) computes a 64 x 64 complex tile; each
// wavefront a 32 x 32 quarter as 2 x 2 MFMA blocks. K advances 16 at a time through two
// LDS buffers, stored k-major so fragment reads are unit-stride across lanes: the next K
// tile's global loads are issued before this tile's MFMAs and stored into the other buffer
// after them, so one barrier per tile suffices and the load latency hides under the MFMAs. Long-K /
// small-output products (projections onto a few hundred bands) are split along K
// into enough workgroups to fill 256 CUs, with a deterministic second-pass reduction.
//
// FP32 accumulation over a long K chunk is two-level: the MFMA's FP32 accumulators cover
// kFlushTiles * BK of K and are then added into FP64 registers, so the rounding error of a
// long contraction (subspace overlaps over ~10^5 plane waves) grows with that length, not
// with K. One FP32 accumulator over all of K = 81k left relative errors of ~6e-5 in the
// Gram matrices of a 440-band Davidson, enough to stall an SCF. Chunks of at most
// kLongChunk (every split-K launch of the projector products) keep the single level and
// its higher occupancy; split partials are summed in FP64 either way.
#include <array>
#include <cmath>
#include <map>
#include <mutex>
#include <type_traits>

#include "kernels.hpp"

namespace amt {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename R>
struct Mfma;
template <>
struct Mfma<float> {
  using acc_t = floatx4;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int q) { return (lane >> 4) * 4 + q; }
};
template <>
struct Mfma<double> {
  using acc_t = doublex4;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int q) { return (lane >> 4) + 4 * q; }
};

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int kFlushTiles = 4;

// Tile shapes TM: 1 = 64 x 64 tiles (each wavefront 32 x 32 as 2 x 2 MFMA blocks); 2 (FP32,
// single level) = 128 x 128 tiles, each wavefront 64 x 64 as 4 x 4 blocks, half the operand
// traffic per MFMA for large products; kTall = 80 x 64 tiles, the four wavefronts side by
// side, each 80 x 16 as 5 x 1 blocks: 160-, 480- or 800-row operands (the Si64 band blocks
// and their multiples) fill them exactly, where 64-row tiles compute 20 % of padding.
constexpr int kTall = 5;

template <int TM>
struct TileShape {
  static constexpr int rows = TM == kTall ? 80 : BM * TM, cols = TM == kTall ? 64 : BN * TM;
  static constexpr int wave_rows = TM == kTall ? 5 : 2 * TM, wave_cols = TM == kTall ? 1 : 2 * TM;  // MFMA blocks
};

template <typename R, int OP, bool kTwoLevel, int TM>
__global__ __launch_bounds__(256, TM == 2 ? 2 : 1) void k_zgemm(int M, int N, int K, int kchunk, int nsplit,
                                               const typename Cplx<R>::T* __restrict__ A, long long lda,
                                               long long sA, const typename Cplx<R>::T* __restrict__ B,
                                               long long ldb, long long sB, typename Cplx<R>::T* __restrict__ Cm,
                                               long long ldc, long long sC, R alpha, R beta,
                                               typename Cplx<R>::T* __restrict__ part, int mfast) {
  static_assert(TM != 2 || !kTwoLevel, "the two-level accumulation uses the 64- or 80-row tiles");
  using C = typename Cplx<R>::T;
  using acc_t = typename Mfma<R>::acc_t;
  constexpr int TBM = TileShape<TM>::rows, TBN = TileShape<TM>::cols;
  constexpr int WBR = TileShape<TM>::wave_rows, WBC = TileShape<TM>::wave_cols;  // MFMA blocks per wave
  constexpr int KQ = TM == kTall ? 4 : 4 * TM;  // consecutive k (or columns) per thread in a 64-wide load
  constexpr int AE = TBM * BK / 256;            // A elements per thread and K tile
  __shared__ C As[2][BK][TBM + 1];  // double-buffered K tiles
  __shared__ C Bs[2][BK][TBN + 1];
  // XCD-aware order: workgroups are handed to the 8 XCDs round-robin by linear id, so
  // logical tile t = (id % 8) * (T8 / 8) + id / 8 gives each XCD a contiguous run of tiles
  // that share an operand panel in its L2: the N-tiles of one M-tile and K-chunk (A panel),
  // or with mfast the M-tiles of one N-tile (B panel; NN products with few M-tiles).
  const unsigned gx = gridDim.x, gy = gridDim.y, gxy = gridDim.x * gridDim.y;
  const unsigned T = gxy * gridDim.z, T8 = T & ~7u;
  const unsigned id = blockIdx.x + gx * blockIdx.y + gxy * blockIdx.z;
  const unsigned tl = id < T8 ? (id & 7u) * (T8 >> 3) + (id >> 3) : id;
  const unsigned txy = tl % gxy;
  const int bz = (int)(tl / gxy);
  const int batch = bz / nsplit, split = bz % nsplit;
  const int m0 = (int)(mfast ? txy % gy : txy / gx) * TBM, n0 = (int)(mfast ? txy / gy : txy % gx) * TBN;
  const int kbeg = split * kchunk;
  const int kend = (kbeg + kchunk < K) ? kbeg + kchunk : K;
  A += (long long)batch * sA;
  B += (long long)batch * sB;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = TM == kTall ? 0 : (wave >> 1) * 32 * TM, wc = TM == kTall ? wave * 16 : (wave & 1) * 32 * TM;
  acc_t p1[WBR][WBC], p2[WBR][WBC], p3[WBR][WBC];  // Gauss's three products
#pragma unroll
  for (int a = 0; a < WBR; ++a)
#pragma unroll
    for (int b = 0; b < WBC; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        p1[a][b][q] = 0;
        p2[a][b][q] = 0;
        p3[a][b][q] = 0;
      }
  // second-level FP64 sums of the FP32 accumulators (two-level form only)
  double sumr[kTwoLevel ? WBR : 1][kTwoLevel ? WBC : 1][4], sumi[kTwoLevel ? WBR : 1][kTwoLevel ? WBC : 1][4];
  auto flush = [&]() {
    if constexpr (kTwoLevel) {
#pragma unroll
      for (int a = 0; a < WBR; ++a)
#pragma unroll
        for (int b = 0; b < WBC; ++b)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double t1 = p1[a][b][q], t2 = p2[a][b][q], t3 = p3[a][b][q];
            sumr[a][b][q] += t1 - t2;
            sumi[a][b][q] += t3 - t1 - t2;
            p1[a][b][q] = 0;
            p2[a][b][q] = 0;
            p3[a][b][q] = 0;
          }
    }
  };
  if constexpr (kTwoLevel) {
#pragma unroll
    for (int a = 0; a < WBR; ++a)
#pragma unroll
      for (int b = 0; b < WBC; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          sumr[a][b][q] = 0.0;
          sumi[a][b][q] = 0.0;
        }
  }
  // Global -> register loads of one K tile (bounds-checked, conj(B) for NC), and the store of
  // those registers into an LDS buffer: A k-major As[k][row]; B k-major Bs[k][col].
  // The 80-row tile loads A as AE single elements per thread (16 threads per row of 16 k).
  C ra[TM == kTall ? AE : KQ], rb[KQ];
  auto gload = [&](int k0) {
    if constexpr (TM == kTall) {
#pragma unroll
      for (int u = 0; u < AE; ++u) {
        const int idx = tid + 256 * u, row = idx / BK, gm = m0 + row, gk = k0 + idx % BK;
        ra[u] = (gm < M && gk < kend) ? A[(long long)gm * lda + gk] : cmake<C>(0, 0);
      }
    } else {
      const int row = tid / (BK / KQ), kq = (tid % (BK / KQ)) * KQ;
      const int gm = m0 + row;
#pragma unroll
      for (int u = 0; u < KQ; ++u) {
        const int gk = k0 + kq + u;
        ra[u] = (gm < M && gk < kend) ? A[(long long)gm * lda + gk] : cmake<C>(0, 0);
      }
    }
    if (OP == 0) {
      const int kk = tid >> 4, c0 = (tid & 15) * KQ;
      const int gk = k0 + kk;
#pragma unroll
      for (int u = 0; u < KQ; ++u) {
        const int gn = n0 + c0 + u;
        rb[u] = (gk < kend && gn < N) ? B[(long long)gk * ldb + gn] : cmake<C>(0, 0);
      }
    } else {
      const int col = tid / (BK / KQ), kq = (tid % (BK / KQ)) * KQ;
      const int gn = n0 + col;
#pragma unroll
      for (int u = 0; u < KQ; ++u) {
        const int gk = k0 + kq + u;
        C v = (gn < N && gk < kend) ? B[(long long)gn * ldb + gk] : cmake<C>(0, 0);
        v.y = -v.y;
        rb[u] = v;
      }
    }
  };
  auto sstore = [&](int buf) {
    if constexpr (TM == kTall) {
#pragma unroll
      for (int u = 0; u < AE; ++u) {
        const int idx = tid + 256 * u;
        As[buf][idx % BK][idx / BK] = ra[u];
      }
    } else {
      const int row = tid / (BK / KQ), kq = (tid % (BK / KQ)) * KQ;
#pragma unroll
      for (int u = 0; u < KQ; ++u) As[buf][kq + u][row] = ra[u];
    }
    if (OP == 0) {
      const int kk = tid >> 4, c0 = (tid & 15) * KQ;
#pragma unroll
      for (int u = 0; u < KQ; ++u) Bs[buf][kk][c0 + u] = rb[u];
    } else {
      const int col = tid / (BK / KQ), kq = (tid % (BK / KQ)) * KQ;
#pragma unroll
      for (int u = 0; u < KQ; ++u) Bs[buf][kq + u][col] = rb[u];
    }
  };
  auto compute = [&](int buf) {
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      C a[WBR], b[WBC];
#pragma unroll
      for (int rb2 = 0; rb2 < WBR; ++rb2) a[rb2] = As[buf][kr][wr + rb2 * 16 + (lane & 15)];
#pragma unroll
      for (int cb = 0; cb < WBC; ++cb) b[cb] = Bs[buf][kr][wc + cb * 16 + (lane & 15)];
#pragma unroll
      for (int rb2 = 0; rb2 < WBR; ++rb2)
#pragma unroll
        for (int cb = 0; cb < WBC; ++cb) {
          p1[rb2][cb] = Mfma<R>::mma(a[rb2].x, b[cb].x, p1[rb2][cb]);
          p2[rb2][cb] = Mfma<R>::mma(a[rb2].y, b[cb].y, p2[rb2][cb]);
          p3[rb2][cb] = Mfma<R>::mma(a[rb2].x + a[rb2].y, b[cb].x + b[cb].y, p3[rb2][cb]);
        }
    }
  };
  int tiles = 0;
  // double-buffered: the next tile's global loads are in flight while this tile's MFMAs run,
  // and one barrier per tile separates the store of the next buffer from the reads of this one
  if (kbeg < kend) {
    gload(kbeg);
    sstore(0);
  }
  __syncthreads();
  int buf = 0;
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    const bool more = k0 + BK < kend;
    if (more) gload(k0 + BK);
    compute(buf);
    if (more) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
    if constexpr (kTwoLevel) {
      if (++tiles == kFlushTiles) {
        flush();
        tiles = 0;
      }
    }
  }
  flush();
#pragma unroll
  for (int rb = 0; rb < WBR; ++rb)
#pragma unroll
    for (int cb = 0; cb < WBC; ++cb)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wr + rb * 16 + Mfma<R>::row(lane, q);
        const int col = n0 + wc + cb * 16 + (lane & 15);
        if (row < M && col < N) {
          C v;
          if constexpr (kTwoLevel)
            v = cmake<C>((R)sumr[rb][cb][q], (R)sumi[rb][cb][q]);
          else
            v = cmake<C>(p1[rb][cb][q] - p2[rb][cb][q], p3[rb][cb][q] - p1[rb][cb][q] - p2[rb][cb][q]);
          if (part) {
            part[(((long long)batch * nsplit + split) * M + row) * N + col] = v;
          } else {
            C* dst = Cm + (long long)batch * sC + (long long)row * ldc + col;
            C o = cmake<C>(alpha * v.x, alpha * v.y);
            if (beta != (R)0) {
              const C old = *dst;
              o.x += beta * old.x;
              o.y += beta * old.y;
            }
            *dst = o;
          }
        }
      }
}

// Sum of the split partials in FP64: each output element is summed by kReduceGroups threads
// (splits g, g + kReduceGroups, ...) whose partial sums are added in a fixed order, so a
// product's result does not depend on scheduling; 32 consecutive elements per group keep the
// loads of one split contiguous.
constexpr int kReduceGroups = 8, kReduceCols = 32;

template <typename R>
__global__ __launch_bounds__(kReduceGroups * kReduceCols) void k_splitk_reduce(
    int M, int N, int nsplit, const typename Cplx<R>::T* __restrict__ part, typename Cplx<R>::T* __restrict__ Cm,
    long long ldc, long long sC, R alpha, R beta) {
  using C = typename Cplx<R>::T;
  __shared__ double red[kReduceGroups][kReduceCols][2];
  const int g = threadIdx.x / kReduceCols, el = threadIdx.x % kReduceCols;
  const long long mn = (long long)M * N;
  const long long e = (long long)blockIdx.x * kReduceCols + el;
  const int batch = blockIdx.y;
  double sr = 0, si = 0;
  if (e < mn) {
    const C* p = part + (long long)batch * nsplit * mn + e;
#pragma unroll 4
    for (int s = g; s < nsplit; s += kReduceGroups) {
      const C v = p[(long long)s * mn];
      sr += v.x;
      si += v.y;
    }
  }
  red[g][el][0] = sr;
  red[g][el][1] = si;
  __syncthreads();
  if (g != 0 || e >= mn) return;
  for (int q = 1; q < kReduceGroups; ++q) {
    sr += red[q][el][0];
    si += red[q][el][1];
  }
  const int row = (int)(e / N), col = (int)(e % N);
  C* dst = Cm + (long long)batch * sC + (long long)row * ldc + col;
  C o = cmake<C>((R)(alpha * sr), (R)(alpha * si));
  if (beta != (R)0) {
    const C old = *dst;
    o.x += beta * old.x;
    o.y += beta * old.y;
  }
  *dst = o;
}

// Split-K factor: a tuned value for the shape's bucket (tiles of M and N, K in 256s,
// batch) when one was recorded with set_zgemm_split, else: the smallest split that gives
// about 1024 workgroups (four 64 x 64 tiles per CU) and, in FP32, K chunks of at most
// kLongChunk (the single-level kernel, at twice the two-level one's occupancy); then as
// many splits as still fit in the same number of 1024-workgroup rounds (a few per cent over
// a round does not count as another), so the chunks get shorter at no extra round. At least
// 256 of K per split, at most 256 splits. Measured on the Si64 projections and overlaps
// (K = 37 073): 1.5-1.7x faster than about 1024 workgroups alone.
static std::mutex g_split_mutex;
static std::map<std::array<int, 6>, int> g_split_table;

static std::array<int, 6> split_key(int prec, int op, int m, int n, int k, int batch) {
  return {prec, op, (m + BM - 1) / BM, (n + BN - 1) / BN, (k + 255) / 256, batch};
}

constexpr int kLongChunk = 1024;

static int choose_split(int prec, int op, int m, int n, int k, int batch, int tile_rows) {
  const long long kmax = (k + 255) / 256;
  {
    std::lock_guard<std::mutex> lock(g_split_mutex);
    auto it = g_split_table.find(split_key(prec, op, m, n, k, batch));
    if (it != g_split_table.end()) {
      long long s = it->second;
      if (s > kmax) s = kmax;
      return s < 1 ? 1 : (int)s;
    }
  }
  const long long tiles = (long long)((m + tile_rows - 1) / tile_rows) * ((n + BN - 1) / BN) * batch;
  if (tiles >= 512) return 1;
  long long want = (1024 + tiles - 1) / tiles;
  const long long single = prec == kFP32 ? (k + kLongChunk - 1) / kLongChunk : 1;
  if (single > want) want = single;
  long long rounds = (long long)std::ceil((double)(tiles * want) / 1024.0 - 0.05);
  if (rounds < 1) rounds = 1;
  const long long fill = rounds * 1024 / tiles;
  want = fill > single ? fill : single;
  if (want > kmax) want = kmax;
  if (want > 256) want = 256;
  if (want < 1) want = 1;
  return (int)want;
}

static bool takes_large_tiles(int prec, int op, int m, int n, int k, int batch);

int set_zgemm_split(int prec, int op, int m, int n, int k, int batch, int split) {
  if (split < 0 || split > 256) {
    set_error("zgemm split must be 0 (heuristic) .. 256");
    return 1;
  }
  if (takes_large_tiles(prec, op, m, n, k, batch)) return 2;  // K chunks fixed by accuracy
  std::lock_guard<std::mutex> lock(g_split_mutex);
  const auto key = split_key(prec, op, m, n, k, batch);
  if (split == 0)
    g_split_table.erase(key);
  else
    g_split_table[key] = split;
  return 0;
}

// Large FP32 NC products (overlaps and projections: both output sides >= kBigSide, K the
// plane waves) use 128 x 128 tiles with K chunks of at most kBigChunk: one FP32
// accumulator per chunk (as accurate as the two-level form for these lengths) and the
// FP64 split reduction, at the cost of writing and reading the partial products once per
// chunk (about 5 % of the product's time). Measured on the Si214 overlaps 1.4-2x faster
// than 64 x 64 tiles; NN products (rotations, projector application) were not faster and
// keep 64 x 64. Beyond kBigWorkspace of partials the two-level form is used instead.
constexpr int kBigSide = 256;
constexpr int kBigChunk = 512;
constexpr size_t kBigWorkspace = size_t(24) << 30;

struct GemmPlan {
  int tm;  // tile shape (1, 2 or kTall)
  int ns;  // K splits (before rounding the chunk to BK)
};

static GemmPlan plan_gemm(int prec, int op, int m, int n, int k, int batch) {
  if (prec == kFP32 && op == 1 && m >= kBigSide && n >= kBigSide && k > 0) {
    const long long tiles = (long long)((m + 2 * BM - 1) / (2 * BM)) * ((n + 2 * BN - 1) / (2 * BN)) * batch;
    long long ns = (k + kBigChunk - 1) / kBigChunk;
    if (tiles < 512) {
      long long want = (1024 + tiles - 1) / tiles;
      const long long kmax = (k + 255) / 256;
      if (want > kmax) want = kmax;
      if (want > ns) ns = want;
    }
    const size_t bytes = ns > 1 ? (size_t)batch * ns * m * n * 8 : 0;
    if (batch * ns <= 65535 && bytes <= kBigWorkspace) return {2, (int)ns};
  }
  // 80-row tiles where they leave fewer padded rows than 64-row ones (m = 160, 480, ...), for
  // the products where they measured faster (Si64 shapes, 160-row blocks): FP32 NN and NC with
  // m < 256 on the single-level kernel (10-21 %), FP64 NN (3-5 %). Their two-level FP32 form
  // and FP64 NC products run at one wave per SIMD and were slower.
  bool tall = (m + 79) / 80 * 80 < (m + 63) / 64 * 64 && (prec == kFP32 ? (op == 0 || m < 256) : op == 0);
  int ns = choose_split(prec, op, m, n, k, batch, tall ? TileShape<kTall>::rows : TileShape<1>::rows);
  if (tall && prec == kFP32 && (k + ns - 1) / ns > kLongChunk) {
    tall = false;
    ns = choose_split(prec, op, m, n, k, batch, TileShape<1>::rows);
  }
  return {tall ? kTall : 1, ns};
}

static bool takes_large_tiles(int prec, int op, int m, int n, int k, int batch) {
  return plan_gemm(prec, op, m, n, k, batch).tm == 2;
}

size_t zgemm_workspace_bytes(int prec, int op, int m, int n, int k, int batch) {
  const GemmPlan g = plan_gemm(prec, op, m, n, k, batch);
  if (g.ns <= 1) return 0;
  const size_t c = prec == kFP64 ? 16 : 8;
  return (size_t)batch * g.ns * m * n * c;
}

template <typename R, bool kTwoLevel, int TM>
static void launch_tile(int op, dim3 grid, hipStream_t s, int m, int n, int k, int kchunk, int ns, const void* a,
                        long long lda, long long sa, const void* b, long long ldb, long long sb, void* c,
                        long long ldc, long long sc, double alpha, double beta, typename Cplx<R>::T* part) {
  using C = typename Cplx<R>::T;
  const int mfast = op == 0 && grid.y > 1 && grid.x >= grid.y;
  if constexpr (TM == 1 || TM == kTall) {
    if (op == 0) {
      hipLaunchKernelGGL((k_zgemm<R, 0, kTwoLevel, TM>), grid, dim3(256), 0, s, m, n, k, kchunk, ns, (const C*)a,
                         lda, sa, (const C*)b, ldb, sb, (C*)c, ldc, sc, (R)alpha, (R)beta, part, mfast);
      return;
    }
  }
  // 128 x 128 tiles are planned for NC products only
  hipLaunchKernelGGL((k_zgemm<R, 1, kTwoLevel, TM>), grid, dim3(256), 0, s, m, n, k, kchunk, ns, (const C*)a,
                     lda, sa, (const C*)b, ldb, sb, (C*)c, ldc, sc, (R)alpha, (R)beta, part, 0);
}


template <typename R>
static int zgemm_t(int op, int m, int n, int k, double alpha, const void* a, long long lda, long long sa,
                   const void* b, long long ldb, long long sb, double beta, void* c, long long ldc, long long sc,
                   int batch, void* ws, size_t ws_bytes, hipStream_t s) {
  using C = typename Cplx<R>::T;
  if (m <= 0 || n <= 0 || batch <= 0) return 0;
  const GemmPlan g = plan_gemm(std::is_same<R, double>::value ? kFP64 : kFP32, op, m, n, k, batch);
  int ns = k <= 0 ? 1 : g.ns;
  const size_t need = ns > 1 ? (size_t)batch * ns * m * n * sizeof(C) : 0;
  C* part = nullptr;
  if (ns > 1) {
    if (ws == nullptr || ws_bytes < need) {
      set_error("zgemm: split-K workspace too small (" + std::to_string(ws_bytes) + " < " + std::to_string(need) +
                ")");
      return 1;
    }
    part = static_cast<C*>(ws);
  }
  int kchunk = (k + ns - 1) / ns;
  kchunk = ((kchunk + BK - 1) / BK) * BK;
  if (kchunk < BK) kchunk = BK;
  ns = (k + kchunk - 1) / kchunk;
  if (ns < 1) ns = 1;
  if ((long long)batch * ns > 65535) {
    set_error("zgemm: batch x split exceeds the grid limit");
    return 1;
  }
  const int trows = g.tm == 2 ? TileShape<2>::rows : g.tm == kTall ? TileShape<kTall>::rows : TileShape<1>::rows;
  const int tcols = g.tm == 2 ? TileShape<2>::cols : g.tm == kTall ? TileShape<kTall>::cols : TileShape<1>::cols;
  const dim3 grid((n + tcols - 1) / tcols, (m + trows - 1) / trows, batch * ns);
  C* const p = ns > 1 ? part : nullptr;
#define AMT_TILE(TWO, SHAPE) \
  launch_tile<R, TWO, SHAPE>(op, grid, s, m, n, k, kchunk, ns, a, lda, sa, b, ldb, sb, c, ldc, sc, alpha, beta, p)
  if constexpr (std::is_same<R, float>::value) {
    if (g.tm == 2)
      AMT_TILE(false, 2);
    else if (kchunk > kLongChunk)
      g.tm == kTall ? AMT_TILE(true, kTall) : AMT_TILE(true, 1);
    else
      g.tm == kTall ? AMT_TILE(false, kTall) : AMT_TILE(false, 1);
  } else {
    g.tm == kTall ? AMT_TILE(false, kTall) : AMT_TILE(false, 1);
  }
#undef AMT_TILE
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("zgemm launch: ") + hipGetErrorString(e));
    return 1;
  }
  if (ns > 1) {
    const long long mn = (long long)m * n;
    hipLaunchKernelGGL((k_splitk_reduce<R>), dim3((unsigned)((mn + kReduceCols - 1) / kReduceCols), batch),
                       dim3(kReduceGroups * kReduceCols), 0, s, m, n, ns, (const C*)part, (C*)c, ldc, sc, (R)alpha,
                       (R)beta);
    e = hipGetLastError();
    if (e != hipSuccess) {
      set_error(std::string("zgemm reduce launch: ") + hipGetErrorString(e));
      return 1;
    }
  }
  return 0;
}

int launch_zgemm(int prec, int op, int m, int n, int k, double alpha, const void* a, long long lda, long long sa,
                 const void* b, long long ldb, long long sb, double beta, void* c, long long ldc, long long sc,
                 int batch, void* ws, size_t ws_bytes, hipStream_t s) {
  return prec == kFP64 ? zgemm_t<double>(op, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch, ws,
                                         ws_bytes, s)
                       : zgemm_t<float>(op, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch, ws,
                                        ws_bytes, s);
}

}  // namespace amt
