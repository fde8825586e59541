// Learned starting density: features, MLP inference and training kernels.
//
// Model (fixed shape): 16 Gaussian radial features of the atoms within r_c of a grid point
// -> three hidden layers of 64 SiLU units -> one linear output, the density correction
// over the atomic superposition in units of `out_scale` (applied by the caller).
//
// The matrix products run on the bf16 matrix cores (v_mfma_f32_16x16x16bf16_1k):
// weights, layer inputs and back-propagated pre-activation gradients are rounded to
// bfloat16 (nearest, ties to even), every sum is FP32. Layers are computed transposed,
// Z^T = W H^T, so the accumulator layout of one layer (lane l: units 4(l/16)+v of point
// l%16) is exactly the B-operand layout of the next and no shuffle is needed between
// layers. Weight gradients (sum over points) need the other orientation and go through
// a per-wave LDS transpose. A wave owns 16 points per tile; all four waves of a
// workgroup walk the same number of tiles so workgroup barriers stay uniform.
//
// Parameters live in one FP32 vector (the optimiser's master copy):
//   W1[64][16] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4
// and are rounded to bf16 into LDS when a kernel starts.
#include "kernels.hpp"

namespace amt {

namespace mlp {
constexpr int NF = 16;
constexpr int H = 64;
constexpr int OW1 = 0, OB1 = OW1 + H * NF, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * H,
              OW4 = OB3 + H, OB4 = OW4 + H, NPARAM = OB4 + 1;
constexpr int NT = 256;             // threads per workgroup (4 waves)
constexpr int S1 = NF + 4;          // LDS row stride (bf16) of W1
constexpr int SH = H + 4;           // LDS row stride (bf16) of W2, W3 and their transposes
constexpr int ST = 16 + 4;          // LDS row stride (bf16) of the per-wave transpose tiles
constexpr int MAX_GRID = 512;
}  // namespace mlp

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_bits(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_value(unsigned short b) { return __uint_as_float((unsigned int)b << 16); }
__device__ __forceinline__ float bf16_round(float f) { return bf16_value(bf16_bits(f)); }

__device__ __forceinline__ s16x4 pack4(float a, float b, float c, float d) {
  s16x4 r;
  r[0] = (short)bf16_bits(a);
  r[1] = (short)bf16_bits(b);
  r[2] = (short)bf16_bits(c);
  r[3] = (short)bf16_bits(d);
  return r;
}
__device__ __forceinline__ s16x4 load4(const unsigned short* p) {
  // 8-byte aligned by construction of the strides
  return *reinterpret_cast<const s16x4*>(p);
}
__device__ __forceinline__ f32x4 mfma(s16x4 a, s16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

struct MlpShared {
  unsigned short w1[mlp::H * mlp::S1];
  unsigned short w2[mlp::H * mlp::SH];
  unsigned short w3[mlp::H * mlp::SH];
  unsigned short w2t[mlp::H * mlp::SH];  // w2t[i][o] = W2[o][i]
  unsigned short w3t[mlp::H * mlp::SH];
  float b1[mlp::H], b2[mlp::H], b3[mlp::H], w4[mlp::H];  // w4 already rounded to bf16
  float mean[mlp::NF], rstd[mlp::NF];
  float b4;
};

__device__ void load_params(MlpShared& sh, const float* __restrict__ theta, const float* __restrict__ mean,
                            const float* __restrict__ stdv, bool transposes) {
  using namespace mlp;
  for (int e = threadIdx.x; e < H * NF; e += NT) sh.w1[(e / NF) * S1 + e % NF] = bf16_bits(theta[OW1 + e]);
  for (int e = threadIdx.x; e < H * H; e += NT) {
    const int o = e / H, i = e % H;
    const unsigned short a = bf16_bits(theta[OW2 + e]), b = bf16_bits(theta[OW3 + e]);
    sh.w2[o * SH + i] = a;
    sh.w3[o * SH + i] = b;
    if (transposes) {
      sh.w2t[i * SH + o] = a;
      sh.w3t[i * SH + o] = b;
    }
  }
  for (int e = threadIdx.x; e < H; e += NT) {
    sh.b1[e] = theta[OB1 + e];
    sh.b2[e] = theta[OB2 + e];
    sh.b3[e] = theta[OB3 + e];
    sh.w4[e] = bf16_round(theta[OW4 + e]);
  }
  for (int e = threadIdx.x; e < NF; e += NT) {
    sh.mean[e] = mean[e];
    sh.rstd[e] = 1.0f / stdv[e];
  }
  if (threadIdx.x == 0) sh.b4 = theta[OB4];
}

// Forward pass of one 16-point tile. Lane l holds point p = l % 16 and, per 16-unit block
// ob, units ob*16 + 4 g + v (g = l / 16, v = 0..3).
struct Tile {
  s16x4 x;            // layer-1 input, features 4g..4g+3 of point p (bf16)
  f32x4 z1[4], z2[4], z3[4];
  s16x4 h1[4], h2[4], h3[4];
  float y;            // output of point p (every lane of the point holds it)
};

__device__ __forceinline__ void tile_forward(const MlpShared& sh, const float* __restrict__ feats, long long npts,
                                             long long tile, Tile& t) {
  using namespace mlp;
  const int lane = threadIdx.x & 63, p = lane & 15, g = lane >> 4;
  const long long q = tile * 16 + p;
  float f[4] = {0.f, 0.f, 0.f, 0.f};
  if (q < npts) {
    const float4 r = *reinterpret_cast<const float4*>(feats + q * NF + 4 * g);
    f[0] = (r.x - sh.mean[4 * g + 0]) * sh.rstd[4 * g + 0];
    f[1] = (r.y - sh.mean[4 * g + 1]) * sh.rstd[4 * g + 1];
    f[2] = (r.z - sh.mean[4 * g + 2]) * sh.rstd[4 * g + 2];
    f[3] = (r.w - sh.mean[4 * g + 3]) * sh.rstd[4 * g + 3];
  }
  t.x = pack4(f[0], f[1], f[2], f[3]);
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(load4(&sh.w1[(ob * 16 + p) * S1 + 4 * g]), t.x, acc);
    float h[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      acc[v] += sh.b1[ob * 16 + 4 * g + v];
      h[v] = silu(acc[v]);
    }
    t.z1[ob] = acc;
    t.h1[ob] = pack4(h[0], h[1], h[2], h[3]);
  }
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(load4(&sh.w2[(ob * 16 + p) * SH + kb * 16 + 4 * g]), t.h1[kb], acc);
    float h[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      acc[v] += sh.b2[ob * 16 + 4 * g + v];
      h[v] = silu(acc[v]);
    }
    t.z2[ob] = acc;
    t.h2[ob] = pack4(h[0], h[1], h[2], h[3]);
  }
  float part = 0.f;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) acc = mfma(load4(&sh.w3[(ob * 16 + p) * SH + kb * 16 + 4 * g]), t.h2[kb], acc);
    float h[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      acc[v] += sh.b3[ob * 16 + 4 * g + v];
      h[v] = silu(acc[v]);
    }
    t.z3[ob] = acc;
    t.h3[ob] = pack4(h[0], h[1], h[2], h[3]);
#pragma unroll
    for (int v = 0; v < 4; ++v) part += sh.w4[ob * 16 + 4 * g + v] * bf16_value((unsigned short)t.h3[ob][v]);
  }
  // sum over the four lanes of the point (g = 0..3), fixed order
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  t.y = part + sh.b4;
}

__global__ __launch_bounds__(mlp::NT) void k_mlp_forward(const float* __restrict__ feats, long long npts,
                                                         const float* __restrict__ theta,
                                                         const float* __restrict__ mean,
                                                         const float* __restrict__ stdv, float* __restrict__ y) {
  __shared__ MlpShared sh;
  load_params(sh, theta, mean, stdv, false);
  __syncthreads();
  const long long ntiles = (npts + 15) / 16;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (long long tile = (long long)blockIdx.x * 4 + wave; tile < ntiles; tile += (long long)gridDim.x * 4) {
    Tile t;
    tile_forward(sh, feats, npts, tile, t);
    const long long q = tile * 16 + lane;
    if (lane < 16 && q < npts) y[q] = t.y;
  }
}

// Per-wave transpose tile: rows = units (up to 64), columns = the 16 points of the tile.
__device__ __forceinline__ void put_tile(unsigned short* tt, int ob, const float* vals /*4*/) {
  const int lane = threadIdx.x & 63, p = lane & 15, g = lane >> 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) tt[(ob * 16 + 4 * g + v) * mlp::ST + p] = bf16_bits(vals[v]);
}
__device__ __forceinline__ void put_tile_bits(unsigned short* tt, int ob, s16x4 bits) {
  const int lane = threadIdx.x & 63, p = lane & 15, g = lane >> 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) tt[(ob * 16 + 4 * g + v) * mlp::ST + p] = (unsigned short)bits[v];
}
// A/B operand of a sum over the tile's points: unit ub*16 + l%16, points 4g..4g+3.
__device__ __forceinline__ s16x4 get_tile(const unsigned short* tt, int ub) {
  const int lane = threadIdx.x & 63, p = lane & 15, g = lane >> 4;
  return load4(&tt[(ub * 16 + p) * mlp::ST + 4 * g]);
}

// Fused forward + backward of L = (1/count) sum_p (y_p - t_p)^2 over the points of this call
// (count may exceed npts when one batch is split over several calls). Writes one partial
// gradient vector per wave: part[wave][NPARAM], and part_loss[wave] = sum (y - t)^2.
__global__ __launch_bounds__(mlp::NT) void k_mlp_train(const float* __restrict__ feats,
                                                       const float* __restrict__ target, long long npts,
                                                       double count, const float* __restrict__ theta,
                                                       const float* __restrict__ mean,
                                                       const float* __restrict__ stdv, float* __restrict__ part,
                                                       float* __restrict__ part_loss) {
  using namespace mlp;
  __shared__ MlpShared sh;
  __shared__ unsigned short tz[4][H * ST];  // per wave: bf16(dz) of the current layer
  __shared__ unsigned short th[4][H * ST];  // per wave: the layer's input
  load_params(sh, theta, mean, stdv, true);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, p = lane & 15, g = lane >> 4;
  unsigned short* TZ = tz[wave];
  unsigned short* TH = th[wave];
  const float scale = (float)(2.0 / count);
  f32x4 gw1[4], gw2[4][4], gw3[4][4];
  float gb1[4][4], gb2[4][4], gb3[4][4], gw4[4][4];
  float gb4 = 0.f, loss = 0.f;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    gw1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      gw2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
      gw3[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
      gb1[a][b] = gb2[a][b] = gb3[a][b] = gw4[a][b] = 0.f;
    }
  }
  const long long ntiles = (npts + 15) / 16;
  const long long stride = (long long)gridDim.x * 4;
  for (long long base = (long long)blockIdx.x * 4; base < ntiles; base += stride) {
    const long long tile = base + wave;  // may be past the end: then every point is masked
    Tile t;
    tile_forward(sh, feats, npts, tile, t);
    const long long q = tile * 16 + p;
    const bool valid = q < npts;
    const float err = valid ? t.y - target[q] : 0.f;
    const float dz4 = scale * err;
    const float dz4r = bf16_round(dz4);
    if (g == 0) {
      loss += err * err;
      gb4 += dz4;
    }
    // layer 4 -> layer 3
    float dz3[4][4];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        gw4[ob][v] += dz4r * bf16_value((unsigned short)t.h3[ob][v]);
        const float dh = dz4r * sh.w4[ob * 16 + 4 * g + v];
        dz3[ob][v] = dh * dsilu(t.z3[ob][v]);
        gb3[ob][v] += dz3[ob][v];
      }
      put_tile(TZ, ob, dz3[ob]);
      put_tile_bits(TH, ob, t.h2[ob]);
    }
    __syncthreads();
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      const s16x4 a = get_tile(TZ, ob);
#pragma unroll
      for (int ib = 0; ib < 4; ++ib) gw3[ob][ib] = mfma(a, get_tile(TH, ib), gw3[ob][ib]);
    }
    // dh2 = W3^T bf16(dz3); B operand = this lane's own bf16(dz3)
    s16x4 dz3r[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) dz3r[kb] = pack4(dz3[kb][0], dz3[kb][1], dz3[kb][2], dz3[kb][3]);
    float dz2[4][4];
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) acc = mfma(load4(&sh.w3t[(ib * 16 + p) * SH + kb * 16 + 4 * g]), dz3r[kb], acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        dz2[ib][v] = acc[v] * dsilu(t.z2[ib][v]);
        gb2[ib][v] += dz2[ib][v];
      }
    }
    __syncthreads();  // every wave is done reading its tiles of layer 3
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      put_tile(TZ, ob, dz2[ob]);
      put_tile_bits(TH, ob, t.h1[ob]);
    }
    __syncthreads();
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) {
      const s16x4 a = get_tile(TZ, ob);
#pragma unroll
      for (int ib = 0; ib < 4; ++ib) gw2[ob][ib] = mfma(a, get_tile(TH, ib), gw2[ob][ib]);
    }
    s16x4 dz2r[4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) dz2r[kb] = pack4(dz2[kb][0], dz2[kb][1], dz2[kb][2], dz2[kb][3]);
    float dz1[4][4];
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) acc = mfma(load4(&sh.w2t[(ib * 16 + p) * SH + kb * 16 + 4 * g]), dz2r[kb], acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        dz1[ib][v] = acc[v] * dsilu(t.z1[ib][v]);
        gb1[ib][v] += dz1[ib][v];
      }
    }
    __syncthreads();
#pragma unroll
    for (int ob = 0; ob < 4; ++ob) put_tile(TZ, ob, dz1[ob]);
    put_tile_bits(TH, 0, t.x);  // 16 feature rows
    __syncthreads();
    {
      const s16x4 b = get_tile(TH, 0);
#pragma unroll
      for (int ob = 0; ob < 4; ++ob) gw1[ob] = mfma(get_tile(TZ, ob), b, gw1[ob]);
    }
    __syncthreads();  // tiles are rewritten by the next iteration
  }
  // ---- write this wave's partial gradient
  float* out = part + (long long)(blockIdx.x * 4 + wave) * NPARAM;
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int o = ob * 16 + 4 * g + v;
      out[OW1 + o * NF + p] = gw1[ob][v];
#pragma unroll
      for (int ib = 0; ib < 4; ++ib) {
        out[OW2 + o * H + ib * 16 + p] = gw2[ob][ib][v];
        out[OW3 + o * H + ib * 16 + p] = gw3[ob][ib][v];
      }
    }
  }
  // biases and w4: sum over the 16 points (lanes with the same g), fixed order
#pragma unroll
  for (int ob = 0; ob < 4; ++ob) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float a = gb1[ob][v], b = gb2[ob][v], c = gb3[ob][v], d = gw4[ob][v];
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
        d += __shfl_xor(d, off, 64);
      }
      if (p == 0) {
        const int o = ob * 16 + 4 * g + v;
        out[OB1 + o] = a;
        out[OB2 + o] = b;
        out[OB3 + o] = c;
        out[OW4 + o] = d;
      }
    }
  }
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    gb4 += __shfl_xor(gb4, off, 64);
    loss += __shfl_xor(loss, off, 64);
  }
  if (lane == 0) {
    out[OB4] = gb4;
    part_loss[blockIdx.x * 4 + wave] = loss;
  }
}

// grad[j] = sum_w part[w][j] (w ascending); loss = sum_w part_loss[w] / count.
__global__ void k_mlp_reduce(const float* __restrict__ part, const float* __restrict__ part_loss, int nw,
                             double count, float* __restrict__ grad, float* __restrict__ loss) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < mlp::NPARAM) {
    float s = 0.f;
    for (int w = 0; w < nw; ++w) s += part[(long long)w * mlp::NPARAM + j];
    grad[j] = s;
  }
  if (j == 0 && loss) {
    double s = 0.0;
    for (int w = 0; w < nw; ++w) s += part_loss[w];
    *loss = (float)(s / count);
  }
}

// Adam (Kingma & Ba 2015, algorithm 1) with bias correction; `step` counts from 1.
__global__ void k_adam(float* __restrict__ theta, const float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float beta1, float beta2, float eps, float bc1,
                       float bc2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const float gj = grad[j];
  const float mj = beta1 * m[j] + (1.0f - beta1) * gj;
  const float vj = beta2 * v[j] + (1.0f - beta2) * gj * gj;
  m[j] = mj;
  v[j] = vj;
  theta[j] -= lr * (mj / bc1) / (sqrtf(vj / bc2) + eps);
}

// Features of every point of an n1 x n2 x n3 grid (C order), FP64 arithmetic.
// frac: atoms' fractional coordinates; cell rows a1, a2, a3 (Bohr).
__global__ __launch_bounds__(128) void k_mlp_features(const double* __restrict__ cell, const double* __restrict__ frac, int natoms,
                               int n1, int n2, int n3, int r1, int r2, int r3, double rcut, double gamma,
                               float* __restrict__ out) {
  const long long npts = (long long)n1 * n2 * n3;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  const int k = (int)(q % n3);
  const int j = (int)((q / n3) % n2);
  const int i = (int)(q / ((long long)n2 * n3));
  const double fp[3] = {(double)i / n1, (double)j / n2, (double)k / n3};
  double c[9];
#pragma unroll
  for (int e = 0; e < 9; ++e) c[e] = cell[e];
  double acc[mlp::NF];
#pragma unroll
  for (int e = 0; e < mlp::NF; ++e) acc[e] = 0.0;
  const double dmu = rcut / (mlp::NF - 1);
  for (int a = 0; a < natoms; ++a) {
    double df[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
      df[x] = fp[x] - frac[3 * a + x];
      df[x] -= rint(df[x]);
    }
    for (int s1 = -r1; s1 <= r1; ++s1)
      for (int s2 = -r2; s2 <= r2; ++s2)
        for (int s3 = -r3; s3 <= r3; ++s3) {
          const double u0 = df[0] + s1, u1 = df[1] + s2, u2 = df[2] + s3;
          const double dx = u0 * c[0] + u1 * c[3] + u2 * c[6];
          const double dy = u0 * c[1] + u1 * c[4] + u2 * c[7];
          const double dz = u0 * c[2] + u1 * c[5] + u2 * c[8];
          const double d = sqrt(dx * dx + dy * dy + dz * dz);
          if (d >= rcut) continue;
          const double fc = 0.5 * (cos(M_PI * d / rcut) + 1.0);
#pragma unroll
          for (int e = 0; e < mlp::NF; ++e) {
            const double t = d - dmu * e;
            acc[e] += fc * exp(-gamma * t * t);
          }
        }
  }
  float* o = out + q * mlp::NF;
#pragma unroll
  for (int e = 0; e < mlp::NF; ++e) o[e] = (float)acc[e];
}

// ------------------------------------------------------------------ launchers
int mlp_param_count() { return mlp::NPARAM; }

int mlp_train_waves(long long npts) {
  const long long ntiles = (npts + 15) / 16;
  long long grid = (ntiles + 3) / 4;
  if (grid > mlp::MAX_GRID) grid = mlp::MAX_GRID;
  if (grid < 1) grid = 1;
  return (int)grid * 4;
}

static int grid_for(long long npts) { return mlp_train_waves(npts) / 4; }

int launch_mlp_forward(const float* feats, long long npts, const float* theta, const float* mean, const float* stdv,
                       float* y, hipStream_t s) {
  if (npts == 0) return 0;
  hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(npts)), dim3(mlp::NT), 0, s, feats, npts, theta, mean, stdv, y);
  return hipGetLastError() != hipSuccess;
}

int launch_mlp_train(const float* feats, const float* target, long long npts, double count, const float* theta,
                     const float* mean, const float* stdv, float* part, float* part_loss, float* grad, float* loss,
                     hipStream_t s) {
  const int grid = grid_for(npts);
  hipLaunchKernelGGL(k_mlp_train, dim3(grid), dim3(mlp::NT), 0, s, feats, target, npts, count, theta, mean, stdv,
                     part, part_loss);
  if (hipGetLastError() != hipSuccess) return 1;
  hipLaunchKernelGGL(k_mlp_reduce, dim3((mlp::NPARAM + 255) / 256), dim3(256), 0, s, part, part_loss, grid * 4,
                     count, grad, loss);
  return hipGetLastError() != hipSuccess;
}

int launch_adam(float* theta, const float* grad, float* m, float* v, int n, float lr, float beta1, float beta2,
                float eps, int step, hipStream_t s) {
  const float bc1 = 1.0f - powf(beta1, (float)step), bc2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, theta, grad, m, v, n, lr, beta1, beta2, eps,
                     bc1, bc2);
  return hipGetLastError() != hipSuccess;
}

int launch_mlp_features(const double* cell, const double* frac, int natoms, int n1, int n2, int n3, int r1, int r2,
                        int r3, double rcut, double gamma, float* out, hipStream_t s) {
  const long long npts = (long long)n1 * n2 * n3;
  if (npts == 0) return 0;
  hipLaunchKernelGGL(k_mlp_features, dim3((unsigned)((npts + 127) / 128)), dim3(128), 0, s, cell, frac, natoms, n1,
                     n2, n3, r1, r2, r3, rcut, gamma, out);
  return hipGetLastError() != hipSuccess;
}

}  // namespace amt
