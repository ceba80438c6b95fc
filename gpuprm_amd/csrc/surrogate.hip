// surrogate.hip -- collision-probability MLP on MFMA.
//
// 7 -> 256 -> 256 -> 256 -> 1, SiLU hidden activations, sigmoid output,
// inputs the joint angles divided by pi.  Weights are kept in FP32 (Adam
// state too) with BF16 copies; activations are BF16; every product runs on
// v_mfma_f32_16x16x16bf16_1k with FP32 accumulation.  Training minimises the
// mean binary cross-entropy against labels of the exact checker (code != 0).
//
// One tiled GEMM serves all six products of a step, through strides:
//   forward      H_l   = silu(H_{l-1} W_l^T + b_l)     (epilogue: bias, SiLU)
//   input grad   dA_l  = (dA_{l+1} W_{l+1}) * silu'(A_l) (epilogue: SiLU')
//   weight grad  dW_l  = dA_l^T H_{l-1}                 (reduction over the batch)
// Tiles of 64x64 outputs per 256-thread block (four waves, 2x2 MFMA blocks of
// 16x16 each), K in steps of 32 staged through LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/gprm.h"
#include "scene.h"
#include "surrogate.h"

namespace gprm {

namespace {

constexpr int kIn = 7;      // joints of the arm the model is sized for
constexpr int kInPad = 16;  // input width padded to one MFMA k-step
constexpr int kH = 256;
constexpr int kTile = 64;
constexpr int kTK = 32;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ float silu(float a) { return a / (1.0f + __expf(-a)); }
__device__ __forceinline__ float dsilu(float a) {
  float s = 1.0f / (1.0f + __expf(-a));
  return s * (1.0f + a * (1.0f - s));
}

enum Epi { kEpiFwd = 0, kEpiBwd = 1, kEpiGrad = 2 };

struct GemmArgs {
  int M, N, K;
  const unsigned short* A; long long lam, lak;  // A(m,k) = A[m*lam + k*lak]
  const unsigned short* B; long long lbk, lbn;  // B(k,n) = B[k*lbk + n*lbn]
  float* Cf; int ldc;                           // FP32 output (pre-activation or gradient)
  unsigned short* Ch;                           // BF16 output (activation / activation grad)
  const float* bias;                            // forward: bias[n]
  const float* aux;                             // backward: pre-activation A_l[m*ldc + n]
  float* ws;                                    // weight gradients: per-split partial sums
  long long zstride;                            //   (ws + z*zstride), summed into Cf afterwards
};

template <int EPI>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
  __shared__ unsigned short As[kTile][kTK + 4];
  __shared__ unsigned short Bs[kTile][kTK + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // split K: block z of gridDim.z covers K range [kb, ke)
  const int kc = (g.K + (int)gridDim.z - 1) / (int)gridDim.z;
  const int kb = (int)blockIdx.z * kc, ke = min(g.K, kb + kc);
  for (int k0 = kb; k0 < ke; k0 += kTK) {
    // stage A (64 x 32) and B^T (64 x 32) tiles, zero outside the matrices
    // and the block's K range; consecutive threads walk each operand's
    // contiguous dimension
    for (int e = tid; e < kTile * kTK; e += 256) {
      int r, c;
      if (g.lak == 1) { r = e / kTK; c = e % kTK; } else { r = e % kTile; c = e / kTile; }
      const int m = m0 + r, k = k0 + c;
      As[r][c] = (m < g.M && k < ke) ? g.A[m * g.lam + k * g.lak] : (unsigned short)0;
      if (g.lbk == 1) { r = e / kTK; c = e % kTK; } else { r = e % kTile; c = e / kTile; }
      const int n = n0 + r, k2 = k0 + c;
      Bs[r][c] = (n < g.N && k2 < ke) ? g.B[k2 * g.lbk + n * g.lbn] : (unsigned short)0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 16) {
      const int kq = kk + 4 * (lane >> 4);
      bf16x4 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const unsigned short* pa = &As[wm + 16 * i + (lane & 15)][kq];
        const unsigned short* pb = &Bs[wn + 16 * i + (lane & 15)][kq];
        a[i] = bf16x4{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
        b[i] = bf16x4{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // D(i, n) of a 16x16 block: lane holds column n = lane%16, rows 4*(lane/16)+r
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int n = n0 + wn + 16 * j + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        const float v = acc[i][j][r];
        const long long o = (long long)m * g.ldc + n;
        if (EPI == kEpiFwd) {
          const float a = v + g.bias[n];
          g.Cf[o] = a;
          g.Ch[o] = f2bf(silu(a));
        } else if (EPI == kEpiBwd) {
          g.Ch[o] = f2bf(v * dsilu(g.aux[o]));
        } else if (gridDim.z > 1) {
          g.ws[blockIdx.z * g.zstride + o] = v;  // split K: this block's partial sum
        } else {
          g.Cf[o] = v;
        }
      }
}

// X[b][0..15] = bf16(q_i / pi), zero padding; labels y = (code != 0).
__global__ void prep_kernel(const float* __restrict__ q, long long off, int n, int nd, unsigned short* __restrict__ X,
                            const int* __restrict__ codes, float* __restrict__ y) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n) return;
  for (int i = 0; i < kInPad; ++i) X[(long long)b * kInPad + i] = i < nd ? f2bf(q[b + off * i] * 0.3183098861837907f) : 0;
  if (codes) y[b] = codes[b] != 0 ? 1.0f : 0.0f;
}

// z = H3 w4 + b4, p = sigmoid(z); one wave per row, lanes over the 256 inputs.
__global__ void head_fwd_kernel(const unsigned short* __restrict__ H3, const float* __restrict__ w4,
                                const float* __restrict__ b4, int n, float* __restrict__ z, float* __restrict__ p) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  float s = 0.0f;
  for (int j = lane; j < kH; j += 64) s += bf2f(H3[(long long)row * kH + j]) * w4[j];
  for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
  if (lane == 0) {
    float zz = s + b4[0];
    z[row] = zz;
    p[row] = 1.0f / (1.0f + __expf(-zz));
  }
}

// dz = (p - y)/n; dA3 = dz w4 * silu'(A3) (bf16); loss terms.
__global__ void head_bwd_kernel(const float* __restrict__ p, const float* __restrict__ z, const float* __restrict__ y,
                                const float* __restrict__ w4, const float* __restrict__ A3, int n,
                                unsigned short* __restrict__ dA3, float* __restrict__ dz, float* __restrict__ lrow) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  const float d = (p[row] - y[row]) / (float)n;
  for (int j = lane; j < kH; j += 64) dA3[(long long)row * kH + j] = f2bf(d * w4[j] * dsilu(A3[(long long)row * kH + j]));
  if (lane == 0) {
    dz[row] = d;
    // BCE from the logit, stable form: max(z,0) - z*y + log(1 + exp(-|z|))
    float zz = z[row];
    float l = fmaxf(zz, 0.0f) - zz * y[row] + log1pf(__expf(-fabsf(zz)));
    lrow[row] = l / (float)n;
  }
}

// dW4[j] = sum_b dz[b] H3[b][j]; db4 = sum dz; and column sums of bf16 dA
// (bias gradients): one thread per column, block y sums rows
// [y * rows, (y + 1) * rows) into out[y * cols + j].
__global__ void colsum_kernel(const unsigned short* __restrict__ Hh, const float* __restrict__ w, int n, int cols,
                              int rows, float* __restrict__ out) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  const int b0 = blockIdx.y * rows, b1 = min(n, b0 + rows);
  float s = 0.0f;
  for (int b = b0; b < b1; ++b) s += (w ? w[b] : 1.0f) * bf2f(Hh[(long long)b * cols + j]);
  out[(long long)blockIdx.y * cols + j] = s;
}
__global__ void sum_kernel(const float* __restrict__ v, int n, float* __restrict__ out) {
  __shared__ float sh[256];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) s += v[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) sh[threadIdx.x] += sh[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// Adam on one FP32 tensor, refreshing its BF16 copy (and transposed copy).
__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, int n, float lr, float b1, float b2, float eps, float c1,
                            float c2, unsigned short* __restrict__ wh, unsigned short* __restrict__ wt, int rows,
                            int cols) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wh) wh[i] = f2bf(wi);
  if (wt) wt[(long long)(i % cols) * rows + i / cols] = f2bf(wi);
}

__global__ void cast_kernel(const float* __restrict__ w, int n, unsigned short* __restrict__ wh,
                            unsigned short* __restrict__ wt, int rows, int cols) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  wh[i] = f2bf(w[i]);
  if (wt) wt[(long long)(i % cols) * rows + i / cols] = f2bf(w[i]);
}

}  // namespace

// Parameter tensors: W1 [256 x 16] (columns 7..15 zero and kept zero), b1,
// W2 [256 x 256], b2, W3, b3, w4 [256], b4 [1].
struct Surrogate {
  int cap = 0, ndof = 0, step = 0;
  float* P[8] = {};        // FP32 parameters
  float* G[8] = {};        // gradients
  float* Mm[8] = {};
  float* Vv[8] = {};
  int size[8] = {};
  unsigned short *W1h = nullptr, *W2h = nullptr, *W3h = nullptr, *W2t = nullptr, *W3t = nullptr;
  unsigned short *X = nullptr, *H1 = nullptr, *H2 = nullptr, *H3 = nullptr, *dA1 = nullptr, *dA2 = nullptr,
                 *dA3 = nullptr;
  float *A1 = nullptr, *A2 = nullptr, *A3 = nullptr, *z = nullptr, *p = nullptr, *y = nullptr, *dz = nullptr,
        *loss = nullptr;
  float* ws = nullptr;     // split-K partial sums of the weight gradients and column sums
  float* lrow = nullptr;   // per-row loss terms
  std::vector<void*> allocs;
};

namespace {

template <class T>
bool dalloc(Surrogate* s, T** p, size_t n) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(T) * (n ? n : 1)) != hipSuccess) return false;
  s->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return true;
}

// ---- fused inference ------------------------------------------------------
// The whole forward of 64 rows per 256-thread block, activations kept in LDS
// (two 64 x 264 bf16 buffers; rows padded to spread the banks), so no
// activation goes to HBM.  Wave w computes output columns [64w, 64w + 64) of
// every layer for all 64 rows: 4 x 4 tiles of 16 x 16 on
// v_mfma_f32_16x16x16bf16_1k, K in the same order as the layered GEMM, so
// the hidden activations equal the training path's bit for bit.  Weight
// fragments (8 bytes per lane) come straight from global memory, where the
// 2 x 128 KB of W2, W3 stay resident in L2.
constexpr int kFR = 64;          // rows per block
constexpr int kHP = kH + 8;      // LDS row pitch (bf16 elements)
constexpr size_t kPredictLds = 2 * (size_t)kFR * kHP * sizeof(unsigned short);

__device__ __forceinline__ bf16x4 ld_bf16x4(const unsigned short* p) {
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  bf16x4 r;
  r[0] = (short)(v.x & 0xffff); r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffff); r[3] = (short)(v.y >> 16);
  return r;
}

// acc (+ bias, SiLU, bf16) into the LDS activation buffer: D(row, col) of a
// 16x16 tile sits in lane (col % 16), rows 4 (lane / 16) + r.
__device__ __forceinline__ void fused_epilogue(const f32x4 (&acc)[4][4], const float* __restrict__ bias,
                                               unsigned short* __restrict__ out, int wave, int lane) {
#pragma unroll
  for (int cj = 0; cj < 4; ++cj) {
    const int col = 64 * wave + 16 * cj + (lane & 15);
    const float b = bias[col];
#pragma unroll
    for (int ri = 0; ri < 4; ++ri)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * ri + 4 * (lane >> 4) + r;
        out[row * kHP + col] = f2bf(silu(acc[ri][cj][r] + b));
      }
  }
}

// One hidden layer 256 -> 256 from LDS buffer `in` into `out`.
__device__ __forceinline__ void fused_layer(const unsigned short* __restrict__ in, const unsigned short* __restrict__ W,
                                            const float* __restrict__ bias, unsigned short* __restrict__ out,
                                            int wave, int lane) {
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int li = lane & 15, lk = 4 * (lane >> 4);
  for (int k0 = 0; k0 < kH; k0 += 16) {
    bf16x4 a[4], b[4];
#pragma unroll
    for (int ri = 0; ri < 4; ++ri) a[ri] = ld_bf16x4(in + (16 * ri + li) * kHP + k0 + lk);
#pragma unroll
    for (int cj = 0; cj < 4; ++cj) b[cj] = ld_bf16x4(W + (size_t)(64 * wave + 16 * cj + li) * kH + k0 + lk);
#pragma unroll
    for (int ri = 0; ri < 4; ++ri)
#pragma unroll
      for (int cj = 0; cj < 4; ++cj)
        acc[ri][cj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ri], b[cj], acc[ri][cj], 0, 0, 0);
  }
  fused_epilogue(acc, bias, out, wave, lane);
}

__global__ void __launch_bounds__(256)
    mlp_predict_kernel(const float* __restrict__ q, long long off, int n, int nd, const unsigned short* __restrict__ W1h,
                       const float* __restrict__ b1, const unsigned short* __restrict__ W2h,
                       const float* __restrict__ b2, const unsigned short* __restrict__ W3h,
                       const float* __restrict__ b3, const float* __restrict__ w4, const float* __restrict__ b4,
                       float* __restrict__ prob) {
  extern __shared__ unsigned short hs[];
  unsigned short* const h0 = hs;
  unsigned short* const h1 = hs + kFR * kHP;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r0 = blockIdx.x * kFR;
  const int li = lane & 15, lk = 4 * (lane >> 4);
  // layer 1: the inputs bf16(q / pi) (zero beyond ndof and n) as A
  // fragments, K = 16 in one step
  {
    bf16x4 xa[4];
#pragma unroll
    for (int ri = 0; ri < 4; ++ri) {
      const int row = r0 + 16 * ri + li;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = lk + j;
        xa[ri][j] = (short)((row < n && k < nd) ? f2bf(q[row + off * k] * 0.3183098861837907f) : 0);
      }
    }
    f32x4 acc[4][4];
#pragma unroll
    for (int cj = 0; cj < 4; ++cj) {
      const bf16x4 wb = ld_bf16x4(W1h + (size_t)(64 * wave + 16 * cj + li) * kInPad + lk);
#pragma unroll
      for (int ri = 0; ri < 4; ++ri)
        acc[ri][cj] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xa[ri], wb, f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
    }
    fused_epilogue(acc, b1, h0, wave, lane);
  }
  __syncthreads();
  fused_layer(h0, W2h, b2, h1, wave, lane);
  __syncthreads();
  fused_layer(h1, W3h, b3, h0, wave, lane);
  __syncthreads();
  // head: z = h3 . w4 + b4 (four threads per row, a quarter of the inputs
  // each), p = sigmoid(z)
  {
    const int row = tid >> 2, part = tid & 3;
    float sacc = 0.0f;
    for (int k = 64 * part; k < 64 * part + 64; ++k) sacc += bf2f(h0[row * kHP + k]) * w4[k];
    sacc += __shfl_xor(sacc, 1, 64);
    sacc += __shfl_xor(sacc, 2, 64);
    if (part == 0 && r0 + row < n) prob[r0 + row] = 1.0f / (1.0f + __expf(-(sacc + b4[0])));
  }
}

// out[i] = sum over z < nz of ws[z * stride + i]
__global__ void split_sum_kernel(const float* __restrict__ ws, long long stride, int nz, int n, float* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc = 0.0f;
  for (int z = 0; z < nz; ++z) acc += ws[z * stride + i];
  out[i] = acc;
}

constexpr int kMaxSplit = 64;

// Column sums of n rows (bias gradients) in up to kMaxSplit row chunks,
// partials through the workspace.
hipError_t colsum(const unsigned short* Hh, const float* w, int n, float* ws, float* out, hipStream_t st) {
  const int chunks = std::max(1, std::min(kMaxSplit, (n + 255) / 256));
  const int rows = (n + chunks - 1) / chunks;
  hipLaunchKernelGGL(colsum_kernel, dim3(1, chunks), dim3(kH), 0, st, Hh, w, n, kH, rows, ws);
  hipLaunchKernelGGL(split_sum_kernel, dim3(1), dim3(kH), 0, st, ws, (long long)kH, chunks, kH, out);
  return hipGetLastError();
}

// Weight gradients reduce over the batch (K = rows): K is split over
// gridDim.z blocks of about 256 rows each, whose partial sums go to the
// workspace and are added up by split_sum_kernel.
hipError_t gemm(int epi, const GemmArgs& g0, hipStream_t st) {
  GemmArgs g = g0;
  int split = 1;
  if (epi == kEpiGrad && g.K > 512 && g.ws) {
    split = (g.K + 255) / 256;
    if (split > kMaxSplit) split = kMaxSplit;
    g.zstride = (long long)g.M * g.ldc;
  }
  dim3 grid((g.N + kTile - 1) / kTile, (g.M + kTile - 1) / kTile, split), block(256);
  if (epi == kEpiFwd) hipLaunchKernelGGL(gemm_bf16_kernel<kEpiFwd>, grid, block, 0, st, g);
  else if (epi == kEpiBwd) hipLaunchKernelGGL(gemm_bf16_kernel<kEpiBwd>, grid, block, 0, st, g);
  else hipLaunchKernelGGL(gemm_bf16_kernel<kEpiGrad>, grid, block, 0, st, g);
  if (split > 1) {
    const int nout = g.M * g.ldc;
    hipLaunchKernelGGL(split_sum_kernel, dim3((nout + 255) / 256), dim3(256), 0, st, g.ws, g.zstride, split, nout,
                       g.Cf);
  }
  return hipGetLastError();
}

hipError_t refresh_bf16(Surrogate* s, hipStream_t st) {
  hipLaunchKernelGGL(cast_kernel, dim3((kH * kInPad + 255) / 256), dim3(256), 0, st, s->P[0], kH * kInPad, s->W1h,
                     (unsigned short*)nullptr, kH, kInPad);
  hipLaunchKernelGGL(cast_kernel, dim3((kH * kH + 255) / 256), dim3(256), 0, st, s->P[2], kH * kH, s->W2h, s->W2t, kH, kH);
  hipLaunchKernelGGL(cast_kernel, dim3((kH * kH + 255) / 256), dim3(256), 0, st, s->P[4], kH * kH, s->W3h, s->W3t, kH, kH);
  return hipGetLastError();
}

// Forward of n <= cap rows already prepared in X.
hipError_t forward(Surrogate* s, int n, hipStream_t st) {
  GemmArgs g{};
  g.M = n; g.N = kH; g.ldc = kH;
  g.K = kInPad; g.A = s->X; g.lam = kInPad; g.lak = 1; g.B = s->W1h; g.lbk = 1; g.lbn = kInPad;
  g.Cf = s->A1; g.Ch = s->H1; g.bias = s->P[1];
  hipError_t e = gemm(kEpiFwd, g, st);
  if (e != hipSuccess) return e;
  g.K = kH; g.A = s->H1; g.lam = kH; g.lak = 1; g.B = s->W2h; g.lbk = 1; g.lbn = kH;
  g.Cf = s->A2; g.Ch = s->H2; g.bias = s->P[3];
  if ((e = gemm(kEpiFwd, g, st)) != hipSuccess) return e;
  g.A = s->H2; g.B = s->W3h; g.Cf = s->A3; g.Ch = s->H3; g.bias = s->P[5];
  if ((e = gemm(kEpiFwd, g, st)) != hipSuccess) return e;
  hipLaunchKernelGGL(head_fwd_kernel, dim3((n + 3) / 4), dim3(256), 0, st, s->H3, s->P[6], s->P[7], n, s->z, s->p);
  return hipGetLastError();
}

// Backward of the last forward against labels in y; gradients into G.
hipError_t backward(Surrogate* s, int n, hipStream_t st) {
  hipLaunchKernelGGL(head_bwd_kernel, dim3((n + 3) / 4), dim3(256), 0, st, s->p, s->z, s->y, s->P[6], s->A3, n,
                     s->dA3, s->dz, s->lrow);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, st, s->lrow, n, s->loss);
  if ((e = colsum(s->H3, s->dz, n, s->ws, s->G[6], st)) != hipSuccess) return e;
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, st, s->dz, n, s->G[7]);
  // layer 3: dW3 = dA3^T H2, db3, dA2 = (dA3 W3) * silu'(A2)
  GemmArgs g{};
  g.M = kH; g.N = kH; g.K = n; g.ldc = kH;
  g.A = s->dA3; g.lam = 1; g.lak = kH; g.B = s->H2; g.lbk = kH; g.lbn = 1; g.Cf = s->G[4];
  g.ws = s->ws;
  if ((e = gemm(kEpiGrad, g, st)) != hipSuccess) return e;
  if ((e = colsum(s->dA3, nullptr, n, s->ws, s->G[5], st)) != hipSuccess) return e;
  GemmArgs d{};
  d.M = n; d.N = kH; d.K = kH; d.ldc = kH;
  d.A = s->dA3; d.lam = kH; d.lak = 1; d.B = s->W3h; d.lbk = kH; d.lbn = 1; d.Ch = s->dA2; d.aux = s->A2;
  if ((e = gemm(kEpiBwd, d, st)) != hipSuccess) return e;
  // layer 2
  g.A = s->dA2; g.B = s->H1; g.Cf = s->G[2];
  if ((e = gemm(kEpiGrad, g, st)) != hipSuccess) return e;
  if ((e = colsum(s->dA2, nullptr, n, s->ws, s->G[3], st)) != hipSuccess) return e;
  d.A = s->dA2; d.B = s->W2h; d.Ch = s->dA1; d.aux = s->A1;
  if ((e = gemm(kEpiBwd, d, st)) != hipSuccess) return e;
  // layer 1: dW1 = dA1^T X (256 x 16), db1
  g.N = kInPad; g.ldc = kInPad; g.A = s->dA1; g.B = s->X; g.lbk = kInPad; g.Cf = s->G[0];
  if ((e = gemm(kEpiGrad, g, st)) != hipSuccess) return e;
  if ((e = colsum(s->dA1, nullptr, n, s->ws, s->G[1], st)) != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace

int surrogate_predict(Surrogate* s, const float* q, int n, long long offset, float* prob, hipStream_t st) {
  if (!s || n < 0) return -1;
  if (n == 0) return 0;
  // the kernel's dynamic LDS exceeds the 64 KB default: raise its limit once
  static std::once_flag once;
  static hipError_t attr = hipSuccess;
  std::call_once(once, [] {
    attr = hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_predict_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPredictLds);
  });
  if (attr != hipSuccess) return -1;
  // grid.x is limited; 2^30 rows per launch
  for (long long c0 = 0; c0 < n; c0 += (1LL << 30)) {
    const int cn = (int)std::min<long long>(1LL << 30, n - c0);
    hipLaunchKernelGGL(mlp_predict_kernel, dim3((cn + kFR - 1) / kFR), dim3(256), kPredictLds, st, q + c0, offset,
                       cn, s->ndof, s->W1h, s->P[1], s->W2h, s->P[3], s->W3h, s->P[5], s->P[6], s->P[7], prob + c0);
    if (hipGetLastError() != hipSuccess) return -1;
  }
  return 0;
}

}  // namespace gprm

using namespace gprm;

extern "C" {

int gprm_surrogate_create(int ndof, int batch, unsigned long long seed, void** out) {
  if (!out || ndof < 1 || ndof > kInPad || batch < 1) return GPRM_E_INVALID;
  *out = nullptr;
  Surrogate* s = new Surrogate();
  s->cap = batch;
  s->ndof = ndof;
  const int sizes[8] = {kH * kInPad, kH, kH * kH, kH, kH * kH, kH, kH, 1};
  bool ok = true;
  for (int t = 0; t < 8 && ok; ++t) {
    s->size[t] = sizes[t];
    ok = dalloc(s, &s->P[t], sizes[t]) && dalloc(s, &s->G[t], sizes[t]) && dalloc(s, &s->Mm[t], sizes[t]) &&
         dalloc(s, &s->Vv[t], sizes[t]);
  }
  const size_t B = batch;
  ok = ok && dalloc(s, &s->W1h, kH * kInPad) && dalloc(s, &s->W2h, kH * kH) && dalloc(s, &s->W3h, kH * kH) &&
       dalloc(s, &s->W2t, kH * kH) && dalloc(s, &s->W3t, kH * kH) && dalloc(s, &s->X, B * kInPad) &&
       dalloc(s, &s->H1, B * kH) && dalloc(s, &s->H2, B * kH) && dalloc(s, &s->H3, B * kH) &&
       dalloc(s, &s->dA1, B * kH) && dalloc(s, &s->dA2, B * kH) && dalloc(s, &s->dA3, B * kH) &&
       dalloc(s, &s->A1, B * kH) && dalloc(s, &s->A2, B * kH) && dalloc(s, &s->A3, B * kH) && dalloc(s, &s->z, B) &&
       dalloc(s, &s->p, B) && dalloc(s, &s->y, B) && dalloc(s, &s->dz, B) && dalloc(s, &s->loss, 1) &&
       dalloc(s, &s->ws, (size_t)kMaxSplit * kH * kH) && dalloc(s, &s->lrow, B);
  if (!ok) {
    for (void* p : s->allocs) (void)hipFree(p);
    delete s;
    return GPRM_E_HIP;
  }
  // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases, seeded
  Rng rng(seed ^ 0x5DEECE66DULL);
  const int fan_in[8] = {kIn, kIn, kH, kH, kH, kH, kH, kH};
  for (int t = 0; t < 8; ++t) {
    std::vector<float> h(sizes[t]), zeros(sizes[t], 0.0f);
    const double b = 1.0 / std::sqrt((double)fan_in[t]);
    for (int i = 0; i < sizes[t]; ++i) h[i] = (float)rng.uniform(-b, b);
    if (t == 0)
      for (int i = 0; i < sizes[t]; ++i)
        if (i % kInPad >= ndof) h[i] = 0.0f;
    if (hipMemcpy(s->P[t], h.data(), 4 * h.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->Mm[t], zeros.data(), 4 * h.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->Vv[t], zeros.data(), 4 * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
      for (void* p : s->allocs) (void)hipFree(p);
      delete s;
      return GPRM_E_HIP;
    }
  }
  if (refresh_bf16(s, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    for (void* p : s->allocs) (void)hipFree(p);
    delete s;
    return GPRM_E_HIP;
  }
  *out = s;
  return GPRM_OK;
}

void gprm_surrogate_destroy(void* h) {
  Surrogate* s = static_cast<Surrogate*>(h);
  if (!s) return;
  (void)hipDeviceSynchronize();
  for (void* p : s->allocs) (void)hipFree(p);
  delete s;
}

// Parameters in the order W1 (256x16), b1, W2 (256x256), b2, W3, b3, w4, b4
// as one FP32 host array; get and set (set refreshes the BF16 copies).
int gprm_surrogate_param_count(void) { return kH * kInPad + kH + 2 * (kH * kH + kH) + kH + 1; }

int gprm_surrogate_get_params(void* h, float* out) {
  Surrogate* s = static_cast<Surrogate*>(h);
  if (!s || !out) return GPRM_E_INVALID;
  size_t o = 0;
  for (int t = 0; t < 8; ++t) {
    if (hipMemcpy(out + o, s->P[t], 4 * (size_t)s->size[t], hipMemcpyDeviceToHost) != hipSuccess) return GPRM_E_HIP;
    o += s->size[t];
  }
  return GPRM_OK;
}

int gprm_surrogate_set_params(void* h, const float* in) {
  Surrogate* s = static_cast<Surrogate*>(h);
  if (!s || !in) return GPRM_E_INVALID;
  size_t o = 0;
  for (int t = 0; t < 8; ++t) {
    if (hipMemcpy(s->P[t], in + o, 4 * (size_t)s->size[t], hipMemcpyHostToDevice) != hipSuccess) return GPRM_E_HIP;
    o += s->size[t];
  }
  if (refresh_bf16(s, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return GPRM_E_HIP;
  return GPRM_OK;
}

// Forward + backward on n <= batch device configurations (SoA, offset) with
// device codes as labels; gradients (same order as the parameters) to
// grads_out (host, may be NULL) and the mean loss to *loss_out.
int gprm_surrogate_gradients(void* h, const float* d_q, long long offset, const int* d_codes, int n, float* grads_out,
                             float* loss_out) {
  Surrogate* s = static_cast<Surrogate*>(h);
  if (!s || !d_q || !d_codes || n < 1 || n > s->cap) return GPRM_E_INVALID;
  hipLaunchKernelGGL(prep_kernel, dim3((n + 255) / 256), dim3(256), 0, nullptr, d_q, offset, n, s->ndof, s->X,
                     d_codes, s->y);
  if (forward(s, n, nullptr) != hipSuccess || backward(s, n, nullptr) != hipSuccess) return GPRM_E_HIP;
  if (hipDeviceSynchronize() != hipSuccess) return GPRM_E_HIP;
  if (loss_out && hipMemcpy(loss_out, s->loss, 4, hipMemcpyDeviceToHost) != hipSuccess) return GPRM_E_HIP;
  if (grads_out) {
    size_t o = 0;
    for (int t = 0; t < 8; ++t) {
      if (hipMemcpy(grads_out + o, s->G[t], 4 * (size_t)s->size[t], hipMemcpyDeviceToHost) != hipSuccess)
        return GPRM_E_HIP;
      o += s->size[t];
    }
  }
  return GPRM_OK;
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on n <= batch device
// configurations; *loss_out (may be NULL) gets the batch loss before the step.
int gprm_surrogate_train_step(void* h, const float* d_q, long long offset, const int* d_codes, int n, float lr,
                              float* loss_out) {
  Surrogate* s = static_cast<Surrogate*>(h);
  if (!s || !d_q || !d_codes || n < 1 || n > s->cap) return GPRM_E_INVALID;
  hipStream_t st = nullptr;
  hipLaunchKernelGGL(prep_kernel, dim3((n + 255) / 256), dim3(256), 0, st, d_q, offset, n, s->ndof, s->X, d_codes,
                     s->y);
  if (forward(s, n, st) != hipSuccess || backward(s, n, st) != hipSuccess) return GPRM_E_HIP;
  // W1's padding columns keep a zero gradient: dA1^T X has zero X columns there
  ++s->step;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)s->step), c2 = 1.0f - std::pow(b2, (float)s->step);
  for (int t = 0; t < 8; ++t) {
    unsigned short* wh = t == 0 ? s->W1h : t == 2 ? s->W2h : t == 4 ? s->W3h : nullptr;
    unsigned short* wt = t == 2 ? s->W2t : t == 4 ? s->W3t : nullptr;
    const int rows = kH, cols = t == 0 ? kInPad : kH;
    hipLaunchKernelGGL(adam_kernel, dim3((s->size[t] + 255) / 256), dim3(256), 0, st, s->P[t], s->Mm[t], s->Vv[t],
                       s->G[t], s->size[t], lr, b1, b2, eps, c1, c2, wh, wt, rows, cols);
  }
  if (hipGetLastError() != hipSuccess) return GPRM_E_HIP;
  if (loss_out && hipMemcpy(loss_out, s->loss, 4, hipMemcpyDeviceToHost) != hipSuccess) return GPRM_E_HIP;
  return GPRM_OK;
}

int gprm_surrogate_predict(void* h, const float* d_q, int n, long long offset, float* d_prob, void* stream) {
  Surrogate* s = static_cast<Surrogate*>(h);
  if (!s || n < 0 || (n > 0 && (!d_q || !d_prob))) return GPRM_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  if (surrogate_predict(s, d_q, n, offset, d_prob, st) != 0) return GPRM_E_HIP;
  if (hipStreamSynchronize(st) != hipSuccess) return GPRM_E_HIP;
  return GPRM_OK;
}

}  // extern "C"
