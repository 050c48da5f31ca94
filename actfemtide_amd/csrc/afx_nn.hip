// Surrogate (learned warm start) kernels: BF16 MFMA GEMM and the data-movement, activation,
// loss and optimizer kernels a 3-D convolutional encoder-decoder needs for training and
// inference on gfx950. There is no counterpart in the reference.
//
// Layouts. Activations are channels-last NDHWC bf16 ([batch][z][y][x][c], c a multiple of 8).
// Every convolution is lowered to one "NT" GEMM, C[m][n] = sum_k A[m][k] * B[n][k], with both
// operands K-contiguous and K padded to a multiple of 32 with zeros:
//   forward 3x3x3:      A = im2col(X) [voxels][27*Ci], B = W [Co][27*Ci]
//   backward data:      A = im2col(dY, flipped taps) [voxels][27*Co], B = W^T [Ci][27*Co]
//   backward weights:   A = dY^T [Co][voxels], B = im2col(X)^T [27*Ci][voxels]
//   2x2x2 stride 2:     im2col is a space-to-depth permutation; the transposed (up) convolution
//                       is a GEMM followed by depth-to-space.
// The GEMM runs on v_mfma_f32_32x32x16_bf16: each wave owns a 32-row strip of C and NW
// 32-column tiles; operands are loaded straight into registers (16 bytes per lane per tile),
// accumulation is FP32, the epilogue adds a bias, applies ReLU and writes FP32 and/or bf16.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/afx_nn.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t *p) {
  const uint4 v = *(const uint4 *)p;
  return __builtin_bit_cast(bf16x8, v);
}

// ---------------------------------------------------------------------------------------------
// GEMM: C[m][n] = sum_k A[m][k] B[n][k] (+ bias[n % bias_period]) (ReLU), K % 16 == 0,
// lda/ldb % 8 == 0.
// grid = (ceil(M/128), ceil(N/(32*NW)), splits); block = 256 (4 waves stacked along M).
// With splits > 1, split s covers k in [s*kchunk, min(K,(s+1)*kchunk)) and writes its partial
// sums to P[s][M][N] (no epilogue); afx_nn_splitk_reduce finishes.
template <int NW>
__global__ void __launch_bounds__(256) k_gemm_nt(const uint16_t *__restrict__ A, int lda,
                                                 const uint16_t *__restrict__ B, int ldb, int M, int N,
                                                 int K, int kchunk, float *__restrict__ C, int ldc,
                                                 uint16_t *__restrict__ Cb, int ldcb,
                                                 const float *__restrict__ bias, int bias_period,
                                                 int relu, float *__restrict__ P) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int m0 = blockIdx.x * 128 + wave * 32;
  const int n0 = blockIdx.y * (32 * NW);
  const int k_begin = blockIdx.z * kchunk;
  const int k_end = min(K, k_begin + kchunk);
  const int r = lane & 31;       // operand row within the 32-row tile
  const int kh = (lane >> 5) * 8;  // which 8-wide half of the 16-wide k step
  const bool arow = (m0 + r) < M;
  const uint16_t *pa = A + (size_t)(arow ? m0 + r : 0) * lda + kh;
  const uint16_t *pb[NW];
  bool brow[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    const int n = n0 + j * 32 + r;
    brow[j] = n < N;
    pb[j] = B + (size_t)(brow[j] ? n : 0) * ldb + kh;
  }
  f32x16 acc[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) acc[j] = f32x16{};
  const bf16x8 zero8 = __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
  if (m0 < M) {
    for (int k = k_begin; k < k_end; k += 32) {
      const bool two = k + 16 < k_end;  // K is a multiple of 16: the last step may be half
      bf16x8 a0 = arow ? load8(pa + k) : zero8;
      bf16x8 a1 = (arow && two) ? load8(pa + k + 16) : zero8;
      bf16x8 b0[NW], b1[NW];
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        b0[j] = brow[j] ? load8(pb[j] + k) : zero8;
        b1[j] = (brow[j] && two) ? load8(pb[j] + k + 16) : zero8;
      }
#pragma unroll
      for (int j = 0; j < NW; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0[j], acc[j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < NW; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1[j], acc[j], 0, 0, 0);
    }
  }
  // D layout: lane holds column (lane & 31), rows 8*(i/4) + 4*(lane/32) + (i%4)
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    const int n = n0 + j * 32 + (lane & 31);
    if (n >= N) continue;
    const float bv = (bias != nullptr && P == nullptr) ? bias[n % bias_period] : 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int m = m0 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
      if (m >= M) continue;
      float v = acc[j][i];
      if (P != nullptr) {
        P[((size_t)blockIdx.z * M + m) * N + n] = v;
      } else {
        v += bv;
        if (relu) v = fmaxf(v, 0.f);
        if (C != nullptr) C[(size_t)m * ldc + n] = v;
        if (Cb != nullptr) Cb[(size_t)m * ldcb + n] = f2bf(v);
      }
    }
  }
}

// the split-K partials added in split order, then the GEMM's epilogue (bias, ReLU, FP32 and/or bf16)
__global__ void k_splitk_epilogue(const float *__restrict__ P, int S, int M, int N, float *__restrict__ C, int ldc,
                                  uint16_t *__restrict__ Cb, int ldcb, const float *__restrict__ bias,
                                  int bias_period, int relu) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)M * N) return;
  const int m = (int)(i / N), n = (int)(i % N);
  float v = 0.f;
  for (int q = 0; q < S; ++q) v += P[(size_t)q * M * N + i];
  if (bias != nullptr) v += bias[n % bias_period];
  if (relu) v = fmaxf(v, 0.f);
  if (C != nullptr) C[(size_t)m * ldc + n] = v;
  if (Cb != nullptr) Cb[(size_t)m * ldcb + n] = f2bf(v);
}

// ---------------------------------------------------------------------------------------------
// im2col for 3x3x3, padding 1, stride 1. cols[v][tap*C + c], tap = (dz*3 + dy)*3 + dx with the
// neighbour at offset (dz-1, dy-1, dx-1); flip uses offset (1-dz, 1-dy, 1-dx) (backward data).
// One thread per (voxel, 8-channel group of the padded row).
__global__ void k_im2col3(const uint16_t *__restrict__ X, uint16_t *__restrict__ cols, int nb, int D, int H, int W,
                          int C, int Kp, int flip) {
  const int g8 = Kp / 8;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nv = (long long)nb * D * H * W;
  if (i >= nv * g8) return;
  const long long v = i / g8;
  const int k = (int)(i % g8) * 8;
  uint4 out = make_uint4(0, 0, 0, 0);
  if (k < 27 * C) {
    const int tap = k / C, c = k % C;
    int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    if (flip) {
      dz = -dz;
      dy = -dy;
      dx = -dx;
    }
    const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)((v / ((long long)W * H)) % D);
    const long long b = v / ((long long)W * H * D);
    const int zz = z + dz, yy = y + dy, xx = x + dx;
    if (zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W)
      out = *(const uint4 *)(X + ((((b * D + zz) * H + yy) * W + xx) * C + c));
  }
  *(uint4 *)(cols + v * Kp + k) = out;
}

// colsT[tap*C + c][v] (row length Vp >= voxels, padded rows/columns zero). One thread per
// (8-channel group, voxel); consecutive threads take consecutive voxels.
__global__ void k_im2col3_t(const uint16_t *__restrict__ X, uint16_t *__restrict__ colsT, int nb, int D, int H,
                            int W, int C, int Kp, int Vp) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int g8 = Kp / 8;
  if (i >= (long long)g8 * Vp) return;
  const int k = (int)(i / Vp) * 8;
  const long long v = i % Vp;
  const long long nv = (long long)nb * D * H * W;
  uint4 val = make_uint4(0, 0, 0, 0);
  if (v < nv && k < 27 * C) {
    const int tap = k / C, c = k % C;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)((v / ((long long)W * H)) % D);
    const long long b = v / ((long long)W * H * D);
    const int zz = z + dz, yy = y + dy, xx = x + dx;
    if (zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W)
      val = *(const uint4 *)(X + ((((b * D + zz) * H + yy) * W + xx) * C + c));
  }
  const uint16_t *h = (const uint16_t *)&val;
#pragma unroll
  for (int j = 0; j < 8; ++j) colsT[(size_t)(k + j) * Vp + v] = h[j];
}

// W [Co][27][Ci] -> Wt [Ci][27*Co (padded to Kp)] with Wt[ci][tap*Co + co] = W[co][tap][ci]
__global__ void k_wt3(const uint16_t *__restrict__ Wsrc, int ldw, uint16_t *__restrict__ Wt, int Co, int Ci,
                      int Kp) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)Ci * Kp) return;
  const int ci = (int)(i / Kp), k = (int)(i % Kp);
  uint16_t v = 0;
  if (k < 27 * Co) {
    const int tap = k / Co, co = k % Co;
    v = Wsrc[(size_t)co * ldw + tap * Ci + ci];
  }
  Wt[i] = v;
}

// space-to-depth by 2: X [nb][D][H][W][C] <-> Y [nb][D/2][H/2][W/2][8*C], sub-voxel
// s = (dz*2 + dy)*2 + dx at channel s*C + c. dir 0 gathers X into Y, dir 1 scatters Y into X.
// Moves 16-byte chunks: q16 = C * element size / 16 chunks per voxel.
__global__ void k_s2d(uint4 *__restrict__ X, uint4 *__restrict__ Y, int nb, int D, int H, int W, int q16, int dir) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nv = (long long)nb * D * H * W;
  if (i >= nv * q16) return;
  const long long v = i / q16;
  const int c = (int)(i % q16);
  const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)((v / ((long long)W * H)) % D);
  const long long b = v / ((long long)W * H * D);
  const int D2 = D / 2, H2 = H / 2, W2 = W / 2;
  const int s = ((z & 1) * 2 + (y & 1)) * 2 + (x & 1);
  const long long vo = ((b * D2 + z / 2) * H2 + y / 2) * W2 + x / 2;
  uint4 *px = X + v * q16 + c;
  uint4 *py = Y + vo * 8 * q16 + s * q16 + c;
  if (dir == 0)
    *py = *px;
  else
    *px = *py;
}

// transpose to bf16 with zero padding: X [R][ldx] (first Cc columns, f32 or bf16) -> Y [Cc][Rp]
template <typename TI>
__device__ __forceinline__ float tofloat(TI v);
template <>
__device__ __forceinline__ float tofloat<float>(float v) { return v; }
template <>
__device__ __forceinline__ float tofloat<uint16_t>(uint16_t v) { return bf2f(v); }

template <typename TI>
__global__ void k_transpose(const TI *__restrict__ X, int ldx, int R, int Cc, uint16_t *__restrict__ Y, int Rp) {
  __shared__ float tile[32][33];
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: ty 0..7
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, c = c0 + tx;
    tile[j][tx] = (r < R && c < Cc) ? tofloat<TI>(X[(size_t)r * ldx + c]) : 0.f;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j, r = r0 + tx;
    if (c < Cc && r < Rp) Y[(size_t)c * Rp + r] = f2bf(tile[tx][j]);
  }
}

__global__ void k_cast_f2b(const float *__restrict__ x, uint16_t *__restrict__ y, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// dPre = dY * (Y > 0) (Y the bf16 ReLU output); writes f32 and bf16
__global__ void k_relu_bwd(const float *__restrict__ dY, const uint16_t *__restrict__ Y, float *__restrict__ dP,
                           uint16_t *__restrict__ dPb, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = (bf2f(Y[i]) > 0.f) ? dY[i] : 0.f;
  dP[i] = g;
  if (dPb != nullptr) dPb[i] = f2bf(g);
}

// column sums of X [R][C] (f32): out[c] = sum_r X[r][c], in two deterministic passes. Pass 1: block
// (chunk, tile) sums the COLSUM_ROWS rows of its chunk for 64 columns (4 row lanes x 64 columns,
// then the 4 lanes in order) into part[chunk][c]; pass 2: one thread per column adds the chunks in
// order. (One block per 64 columns summing every row took 6.6 ms for the 98,304 x 32 bias gradient
// of the first level; the chunked passes spread those rows over the whole GPU.)
constexpr int COLSUM_ROWS = 256;
__global__ void k_colsum_part(const float *__restrict__ X, int R, int Cc, float *__restrict__ part) {
  __shared__ float lane_sum[4][64];
  const int c = blockIdx.y * 64 + (threadIdx.x & 63);
  const int ry = threadIdx.x >> 6;
  const int r0 = blockIdx.x * COLSUM_ROWS, r1 = min(R, r0 + COLSUM_ROWS);
  float s = 0.f;
  if (c < Cc)
    for (int r = r0 + ry; r < r1; r += 4) s += X[(size_t)r * Cc + c];
  lane_sum[ry][threadIdx.x & 63] = s;
  __syncthreads();
  if (ry == 0 && c < Cc)
    part[(size_t)blockIdx.x * Cc + c] =
        lane_sum[0][threadIdx.x] + lane_sum[1][threadIdx.x] + lane_sum[2][threadIdx.x] + lane_sum[3][threadIdx.x];
}
__global__ void k_colsum_fin(const float *__restrict__ part, int nchunk, int Cc, float *__restrict__ out,
                             int accumulate) {
  // one wave per column: lane l adds chunks l, l + 64, ... in order, then a fixed XOR tree
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
  if (c >= Cc) return;  // wave-uniform
  float t = 0.f;
  for (int q = l; q < nchunk; q += 64) t += part[(size_t)q * Cc + c];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) t += __shfl_xor(t, m, 64);
  if (l == 0) out[c] = accumulate ? out[c] + t : t;
}

__global__ void k_add_b(const uint16_t *__restrict__ a, const uint16_t *__restrict__ b, uint16_t *__restrict__ y,
                        long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(bf2f(a[i]) + bf2f(b[i]));
}

__global__ void k_add_f(const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ y, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a[i] + b[i];
}

// masked mean squared error over the first Cc columns of out [rows][ld] against tgt [rows][Cc];
// rows with mask == 0 excluded. Loss partial per block; grad [rows][ld] (zero beyond Cc).
__global__ void k_mse(const float *__restrict__ out, int ld, const float *__restrict__ tgt,
                      const uint8_t *__restrict__ mask, long long rows, int Cc, float inv_count,
                      float *__restrict__ grad, float *__restrict__ part) {
  __shared__ float red[256];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < rows * ld) {
    const long long r = i / ld;
    const int c = (int)(i % ld);
    const bool on = c < Cc && (mask == nullptr || mask[r] != 0);
    const float d = on ? out[i] - tgt[r * Cc + c] : 0.f;
    l = d * d * inv_count;
    grad[i] = 2.f * d * inv_count;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_sum(const float *__restrict__ x, int n, float *__restrict__ y) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[0] = red[0];
}

// Adam (decoupled weight decay 0) on FP32 master weights, bf16 copy refreshed
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       uint16_t *__restrict__ wb, long long n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb != nullptr) wb[i] = f2bf(wi);
}

// cell field -> edge values: x[e] = mean over incident cells c of sum_d F[c][d] * t[e][d],
// F complex as 6 floats (Re x, Re y, Re z, Im x, Im y, Im z) scaled by `scale`
__global__ void k_cells_to_edges(const int *__restrict__ eptr, const int *__restrict__ ecell,
                                 const double *__restrict__ evec, const float *__restrict__ F, int ldf, double scale,
                                 const uint8_t *__restrict__ zero, int ne, double *__restrict__ x) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  if (zero != nullptr && zero[e]) {
    x[2 * e] = 0.0;
    x[2 * e + 1] = 0.0;
    return;
  }
  const int b = eptr[e], en = eptr[e + 1];
  const double tx = evec[3 * e], ty = evec[3 * e + 1], tz = evec[3 * e + 2];
  double re = 0.0, im = 0.0;
  for (int q = b; q < en; ++q) {
    const float *f = F + (size_t)ecell[q] * ldf;
    re += f[0] * tx + f[1] * ty + f[2] * tz;
    im += f[3] * tx + f[4] * ty + f[5] * tz;
  }
  const double s = (en > b) ? scale / (double)(en - b) : 0.0;
  x[2 * e] = re * s;
  x[2 * e + 1] = im * s;
}

// edge values -> cell field by least squares per cell: F = (T^T T)^-1 T^T x over the cell's
// edges (rows t_e), written as 6 floats divided by `scale`
__global__ void k_edges_to_cells(const int *__restrict__ cptr, const int *__restrict__ cedge,
                                 const double *__restrict__ evec, const double *__restrict__ x, double scale,
                                 int nc, float *__restrict__ F, int ldf) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double g[6] = {0, 0, 0, 0, 0, 0}, br[3] = {0, 0, 0}, bi[3] = {0, 0, 0};
  for (int q = cptr[c]; q < cptr[c + 1]; ++q) {
    const int e = cedge[q];
    const double t[3] = {evec[3 * e], evec[3 * e + 1], evec[3 * e + 2]};
    g[0] += t[0] * t[0];
    g[1] += t[0] * t[1];
    g[2] += t[0] * t[2];
    g[3] += t[1] * t[1];
    g[4] += t[1] * t[2];
    g[5] += t[2] * t[2];
    for (int d = 0; d < 3; ++d) {
      br[d] += t[d] * x[2 * e];
      bi[d] += t[d] * x[2 * e + 1];
    }
  }
  // symmetric 3x3 inverse by cofactors
  const double a = g[0], b = g[1], cc = g[2], d = g[3], e2 = g[4], f = g[5];
  const double i00 = d * f - e2 * e2, i01 = cc * e2 - b * f, i02 = b * e2 - cc * d;
  const double i11 = a * f - cc * cc, i12 = b * cc - a * e2, i22 = a * d - b * b;
  const double det = a * i00 + b * i01 + cc * i02;
  const double s = (det != 0.0) ? 1.0 / (det * scale) : 0.0;
  float *o = F + (size_t)c * ldf;
  o[0] = (float)((i00 * br[0] + i01 * br[1] + i02 * br[2]) * s);
  o[1] = (float)((i01 * br[0] + i11 * br[1] + i12 * br[2]) * s);
  o[2] = (float)((i02 * br[0] + i12 * br[1] + i22 * br[2]) * s);
  o[3] = (float)((i00 * bi[0] + i01 * bi[1] + i02 * bi[2]) * s);
  o[4] = (float)((i01 * bi[0] + i11 * bi[1] + i12 * bi[2]) * s);
  o[5] = (float)((i02 * bi[0] + i12 * bi[1] + i22 * bi[2]) * s);
}

inline unsigned nblk(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

thread_local const char *g_err = "";

int fail(const char *m) {
  g_err = m;
  return AFX_NN_EINVAL;
}

int check(hipError_t e) {
  if (e != hipSuccess) {
    g_err = hipGetErrorString(e);
    return AFX_NN_EHIP;
  }
  return 0;
}

}  // namespace

extern "C" {

const char *afx_nn_last_error(void) { return g_err; }

int afx_nn_gemm_nt(const uint16_t *A, int lda, const uint16_t *B, int ldb, int M, int N, int K, float *C, int ldc,
                   uint16_t *Cb, int ldcb, const float *bias, int bias_period, int relu, int splits, float *partial,
                   void *stream) {
  if (M <= 0 || N <= 0 || K <= 0) return fail("gemm: empty shape");
  if (K % 16 != 0) return fail("gemm: K must be a multiple of 16 (pad the operands)");
  if (lda % 8 || ldb % 8 || lda < K || ldb < K) return fail("gemm: lda/ldb must be >= K and multiples of 8");
  if (((uintptr_t)A | (uintptr_t)B) & 15) return fail("gemm: operands must be 16-byte aligned");
  if (splits < 1) splits = 1;
  if (splits > 1 && partial == nullptr) return fail("gemm: split-K needs a partial buffer");
  if (C == nullptr && Cb == nullptr) return fail("gemm: no output");
  if (bias_period <= 0) bias_period = N;
  hipStream_t s = (hipStream_t)stream;
  const int nw = N > 64 ? 4 : (N > 32 ? 2 : 1);
  const int kchunk = ((K / 16 + splits - 1) / splits) * 16;  // split boundaries on 16
  splits = (K + kchunk - 1) / kchunk;
  if (splits > 1 && partial == nullptr) return fail("gemm: split-K needs a partial buffer");
  dim3 grid((M + 127) / 128, (N + 32 * nw - 1) / (32 * nw), splits);
  float *P = splits > 1 ? partial : nullptr;
  if (nw == 4)
    hipLaunchKernelGGL(k_gemm_nt<4>, grid, dim3(256), 0, s, A, lda, B, ldb, M, N, K, kchunk, C, ldc, Cb, ldcb, bias,
                       bias_period, relu, P);
  else if (nw == 2)
    hipLaunchKernelGGL(k_gemm_nt<2>, grid, dim3(256), 0, s, A, lda, B, ldb, M, N, K, kchunk, C, ldc, Cb, ldcb, bias,
                       bias_period, relu, P);
  else
    hipLaunchKernelGGL(k_gemm_nt<1>, grid, dim3(256), 0, s, A, lda, B, ldb, M, N, K, kchunk, C, ldc, Cb, ldcb, bias,
                       bias_period, relu, P);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && P != nullptr) {
    hipLaunchKernelGGL(k_splitk_epilogue, dim3(nblk((long long)M * N)), dim3(256), 0, s, P, splits, M, N, C, ldc, Cb,
                       ldcb, bias, bias_period, relu);
    e = hipGetLastError();
  }
  return check(e);
}

int afx_nn_im2col3(const uint16_t *X, uint16_t *cols, int nb, int D, int H, int W, int C, int Kp, int flip,
                   void *stream) {
  if (C % 8 || Kp % 8 || Kp < 27 * C) return fail("im2col3: C and Kp must be multiples of 8, Kp >= 27*C");
  const long long n = (long long)nb * D * H * W * (Kp / 8);
  hipLaunchKernelGGL(k_im2col3, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, X, cols, nb, D, H, W, C, Kp, flip);
  return check(hipGetLastError());
}

int afx_nn_im2col3_t(const uint16_t *X, uint16_t *colsT, int nb, int D, int H, int W, int C, int Kp, int Vp,
                     void *stream) {
  if (C % 8 || Kp % 8 || Kp < 27 * C) return fail("im2col3_t: C and Kp must be multiples of 8, Kp >= 27*C");
  if ((long long)Vp < (long long)nb * D * H * W) return fail("im2col3_t: Vp smaller than the voxel count");
  const long long n = (long long)(Kp / 8) * Vp;
  hipLaunchKernelGGL(k_im2col3_t, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, X, colsT, nb, D, H, W, C, Kp, Vp);
  return check(hipGetLastError());
}

int afx_nn_weight_t3(const uint16_t *W, int ldw, uint16_t *Wt, int Co, int Ci, int Kp, void *stream) {
  if (Kp < 27 * Co || ldw < 27 * Ci) return fail("weight_t3: bad leading dimensions");
  hipLaunchKernelGGL(k_wt3, dim3(nblk((long long)Ci * Kp)), dim3(256), 0, (hipStream_t)stream, W, ldw, Wt, Co, Ci,
                     Kp);
  return check(hipGetLastError());
}

int afx_nn_space_to_depth(void *X, void *Y, int nb, int D, int H, int W, int C, int esize, int inverse,
                          void *stream) {
  if ((esize != 2 && esize != 4) || (C * esize) % 16 || D % 2 || H % 2 || W % 2)
    return fail("space_to_depth: element size 2 or 4, C*esize % 16 == 0 and even D, H, W required");
  const int q16 = C * esize / 16;
  const long long n = (long long)nb * D * H * W * q16;
  hipLaunchKernelGGL(k_s2d, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, (uint4 *)X, (uint4 *)Y, nb, D, H, W,
                     q16, inverse ? 1 : 0);
  return check(hipGetLastError());
}

int afx_nn_transpose(const void *X, int esize, int ldx, int R, int Cc, uint16_t *Y, int Rp, void *stream) {
  if (Rp < R || ldx < Cc || (esize != 2 && esize != 4)) return fail("transpose: bad sizes");
  dim3 grid((Rp + 31) / 32, (Cc + 31) / 32);
  if (esize == 4)
    hipLaunchKernelGGL(k_transpose<float>, grid, dim3(256), 0, (hipStream_t)stream, (const float *)X, ldx, R, Cc, Y,
                       Rp);
  else
    hipLaunchKernelGGL(k_transpose<uint16_t>, grid, dim3(256), 0, (hipStream_t)stream, (const uint16_t *)X, ldx, R,
                       Cc, Y, Rp);
  return check(hipGetLastError());
}

int afx_nn_cast_f2b(const float *x, uint16_t *y, long long n, void *stream) {
  hipLaunchKernelGGL(k_cast_f2b, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, x, y, n);
  return check(hipGetLastError());
}

int afx_nn_relu_bwd(const float *dY, const uint16_t *Y, float *dP, uint16_t *dPb, long long n, void *stream) {
  hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, dY, Y, dP, dPb, n);
  return check(hipGetLastError());
}

int afx_nn_colsum(const float *X, int R, int Cc, float *out, int accumulate, float *part, void *stream) {
  if (R <= 0 || Cc <= 0) return 0;
  if (part == nullptr) return fail("colsum: no workspace");
  const int nchunk = (R + COLSUM_ROWS - 1) / COLSUM_ROWS;
  hipLaunchKernelGGL(k_colsum_part, dim3(nchunk, (Cc + 63) / 64), dim3(256), 0, (hipStream_t)stream, X, R, Cc, part);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_colsum_fin, dim3((Cc + 3) / 4), dim3(256), 0, (hipStream_t)stream, part, nchunk, Cc, out,
                       accumulate);
    e = hipGetLastError();
  }
  return check(e);
}

int afx_nn_add_bf16(const uint16_t *a, const uint16_t *b, uint16_t *y, long long n, void *stream) {
  hipLaunchKernelGGL(k_add_b, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, a, b, y, n);
  return check(hipGetLastError());
}

int afx_nn_add_f32(const float *a, const float *b, float *y, long long n, void *stream) {
  hipLaunchKernelGGL(k_add_f, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, a, b, y, n);
  return check(hipGetLastError());
}

int afx_nn_mse(const float *out, int ld, const float *tgt, const uint8_t *mask, long long rows, int Cc,
               float inv_count, float *grad, float *partial, float *loss, void *stream) {
  if (ld < Cc) return fail("mse: ld < Cc");
  const long long n = rows * ld;
  const unsigned nb = nblk(n);
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, (hipStream_t)stream, out, ld, tgt, mask, rows, Cc, inv_count,
                     grad, partial);
  if (int rc = check(hipGetLastError())) return rc;
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, (hipStream_t)stream, partial, (int)nb, loss);
  return check(hipGetLastError());
}

int afx_nn_adam(float *w, const float *g, float *m, float *v, uint16_t *wb, long long n, float lr, float b1, float b2,
                float eps, int step, void *stream) {
  if (step < 1) return fail("adam: step counts from 1");
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, wb, n, lr, b1, b2, eps,
                     c1, c2);
  return check(hipGetLastError());
}

int afx_nn_cells_to_edges(const int *eptr, const int *ecell, const double *evec, const float *F, int ldf,
                          double scale, const uint8_t *zero, int ne, double *x, void *stream) {
  if (ldf < 6) return fail("cells_to_edges: ldf < 6");
  hipLaunchKernelGGL(k_cells_to_edges, dim3(nblk(ne)), dim3(256), 0, (hipStream_t)stream, eptr, ecell, evec, F, ldf,
                     scale, zero, ne, x);
  return check(hipGetLastError());
}

int afx_nn_edges_to_cells(const int *cptr, const int *cedge, const double *evec, const double *x, double scale,
                          int nc, float *F, int ldf, void *stream) {
  if (ldf < 6 || scale == 0.0) return fail("edges_to_cells: ldf < 6 or zero scale");
  hipLaunchKernelGGL(k_edges_to_cells, dim3(nblk(nc)), dim3(256), 0, (hipStream_t)stream, cptr, cedge, evec, x, scale,
                     nc, F, ldf);
  return check(hipGetLastError());
}

}  // extern "C"
