// surrogate.hip -- collision-probability MLP on MFMA.
//
// 7 -> 256 -> 256 -> 256 -> 1, SiLU hidden activations, sigmoid output,
// inputs the joint angles divided by pi.  Weights are kept in FP32 (Adam
// state too) with BF16 copies; activations and activation gradients are
// BF16; every product runs on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation.  Training minimises the mean binary cross-entropy against
// labels of the exact checker (code != 0).
//
// Every kernel works on tiles of 64 rows (configurations) with 512 threads:
// wave w owns 32 output columns of each 256-wide layer and keeps its slice of
// that layer's weights in registers for the whole (persistent) launch, so a
// layer's weights are read from memory once per block, not once per tile.
// Layer inputs come from LDS; the weights are the MFMA's A operand and the
// activations its B operand, so a lane's four accumulator values are four
// consecutive output columns of one row (one 8-byte store).
//
// A training step is four launches:
//   F  forward of all three layers and the head, the loss, dz = (p - y)/n,
//      dA3 = dz w4 silu'(A3); writes H1, H2 and the inputs transposed (for the
//      weight gradients), silu'(A1), silu'(A2) (for the backward), dA3, and
//      per-block partial sums of dw4, db4 and the loss;
//   B  dA2 = (dA3 W3) silu'(A2), dA1 = (dA2 W2) silu'(A1) with W3^T and W2^T
//      in registers; writes dA3, dA2, dA1 transposed;
//   W  dW_l = dA_l^T H_(l-1) and db_l = column sums of dA_l, K = the batch
//      split over blocks, partial sums per split (no atomics: the result
//      does not depend on scheduling);
//   R  sums of the partials, Adam, and the BF16 copies (W, then W^T by a
//      tiled transpose).
// BF16 roundings: inputs, weight copies, hidden activations, silu' and the
// activation gradients; everything else is FP32.
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
constexpr int kInPad = 16;  // input width, padded with zero columns
constexpr int kH = 256;

// Parameters as one flat FP32 array: W1 [256 x 16], b1, W2 [256 x 256], b2,
// W3, b3, w4 [256], b4.
constexpr int kOffW1 = 0, kOffB1 = kH * kInPad, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kH, kOffW4 = kOffB3 + kH, kOffB4 = kOffW4 + kH,
              kPtot = kOffB4 + 1;
constexpr int kTensorOff[9] = {kOffW1, kOffB1, kOffW2, kOffB2, kOffW3, kOffB3, kOffW4, kOffB4, kPtot};
// Partial-sum rows: the weight-gradient kernel's cover W1 .. b3 (flat
// offsets), the forward kernel's dw4 [256], db4, loss.
constexpr int kRowW = kOffW4;
constexpr int kRowF = 260;
constexpr int kMaxSplit = 32;    // K splits of the layer-2/3 weight gradients
constexpr int kMaxSplit1 = 128;  // and of layer 1's

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4h __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kFR = 64;          // rows per tile
constexpr int kPW = 8;           // waves per block
constexpr int kPC = kH / kPW;    // output columns per wave
constexpr int kPJ = kPC / 16;    // 16-column blocks per wave
constexpr int kHP = kH + 8;      // LDS row pitch of a 64 x 256 tile (bf16; 528 bytes, 16-byte aligned)
constexpr int kXP = 16;          // LDS row pitch of the staged inputs
constexpr int kTileElems = kFR * kHP;

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// The thread index as a value the compiler cannot see through: addresses
// derived from it inside a tile loop are recomputed per tile (a few integer
// operations) instead of being hoisted out of the loop and kept live, which
// with the weights resident would not fit in the registers.
__device__ __forceinline__ int opaque_tid() {
  int t = (int)threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// ---- shared pieces -----------------------------------------------------------

// This wave's fragments of a 256 x 256 bf16 matrix W (row-major, W[n][k]):
// w[cj][s] = W[kPC w + 16 cj + li][32 s + 8 g .. + 7] (lane = 16 g + li).
__device__ __forceinline__ void load_frags(const unsigned short* __restrict__ W, bf16x8 (&w)[kPJ][8], int wave,
                                           int lane) {
  const int li = lane & 15, g = lane >> 4;
#pragma unroll
  for (int cj = 0; cj < kPJ; ++cj) {
    const int row = kPC * wave + 16 * cj + li;
#pragma unroll
    for (int s = 0; s < 8; ++s) w[cj][s] = *reinterpret_cast<const bf16x8*>(W + (size_t)row * kH + 32 * s + 8 * g);
  }
}

// W1 is 256 x 16: k = 8 g + j exists for g < 2; K is padded to 32 with zeros.
__device__ __forceinline__ void load_w1_frags(const unsigned short* __restrict__ W1, bf16x8 (&w)[kPJ], int wave,
                                              int lane) {
  const int li = lane & 15, g = lane >> 4;
#pragma unroll
  for (int cj = 0; cj < kPJ; ++cj) {
    const int row = kPC * wave + 16 * cj + li;
    w[cj] = g < 2 ? *reinterpret_cast<const bf16x8*>(W1 + (size_t)row * kInPad + 8 * g) : bf16x8{};
  }
}

// acc[cj][ri] = sum_k W[kPC w + 16 cj + 4 g + r][k] in[16 ri + li][k] over
// the 256 columns of the LDS tile `in`.
__device__ __forceinline__ void layer_mfma(const __bf16* __restrict__ in, const bf16x8 (&w)[kPJ][8],
                                           f32x4 (&acc)[kPJ][4], int lane) {
  const int li = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < kPJ; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  // fragments one k-step ahead; the scheduling barrier keeps the compiler
  // from hoisting all eight steps' loads (128 more registers than there are)
  const __bf16* src = in + li * kHP + 8 * g;
  bf16x8 a[4], an[4];
#pragma unroll
  for (int ri = 0; ri < 4; ++ri) a[ri] = *reinterpret_cast<const bf16x8*>(src + 16 * ri * kHP);
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    if (s < 7) {
#pragma unroll
      for (int ri = 0; ri < 4; ++ri) an[ri] = *reinterpret_cast<const bf16x8*>(src + 16 * ri * kHP + 32 * (s + 1));
    }
#pragma unroll
    for (int cj = 0; cj < kPJ; ++cj)
#pragma unroll
      for (int ri = 0; ri < 4; ++ri)
        acc[cj][ri] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[cj][s], a[ri], acc[cj][ri], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ri = 0; ri < 4; ++ri) a[ri] = an[ri];
  }
}

// Layer 1 from the staged inputs xs (64 x kXP).
__device__ __forceinline__ void layer1_mfma(const __bf16* __restrict__ xs, const bf16x8 (&w1)[kPJ],
                                            f32x4 (&acc)[kPJ][4], int lane) {
  const int li = lane & 15, g = lane >> 4;
  bf16x8 xa[4];
#pragma unroll
  for (int ri = 0; ri < 4; ++ri)
    xa[ri] = g < 2 ? *reinterpret_cast<const bf16x8*>(xs + (16 * ri + li) * kXP + 8 * g) : bf16x8{};
#pragma unroll
  for (int cj = 0; cj < kPJ; ++cj)
#pragma unroll
    for (int ri = 0; ri < 4; ++ri)
      acc[cj][ri] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1[cj], xa[ri], f32x4{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
}

// SiLU and its derivative from one exponential and one reciprocal:
// sigma = 1 / (1 + e^-a), silu = a sigma, silu' = sigma (1 + a (1 - sigma)).
__device__ __forceinline__ float silu_fast(float a) { return a * rcp(1.0f + __expf(-a)); }
__device__ __forceinline__ void silu_both(float a, float& h, float& d) {
  const float sg = rcp(1.0f + __expf(-a));
  h = a * sg;
  d = sg * (1.0f + a * (1.0f - sg));
}

// Inference epilogue: out[row][col..col+3] = bf16(silu(acc + bias)), row =
// 16 ri + li, col = kPC w + 16 cj + 4 g.
__device__ __forceinline__ void epi_predict(const f32x4 (&acc)[kPJ][4], const float* __restrict__ bias,
                                            __bf16* __restrict__ out, int wave, int lane) {
  const int li = lane & 15, g = lane >> 4;
#pragma unroll
  for (int cj = 0; cj < kPJ; ++cj) {
    const int col = kPC * wave + 16 * cj + 4 * g;
    const float4 bb = *reinterpret_cast<const float4*>(bias + col);
#pragma unroll
    for (int ri = 0; ri < 4; ++ri) {
      bf16x4h v;
      v[0] = (__bf16)silu_fast(acc[cj][ri][0] + bb.x);
      v[1] = (__bf16)silu_fast(acc[cj][ri][1] + bb.y);
      v[2] = (__bf16)silu_fast(acc[cj][ri][2] + bb.z);
      v[3] = (__bf16)silu_fast(acc[cj][ri][3] + bb.w);
      *reinterpret_cast<bf16x4h*>(out + (16 * ri + li) * kHP + col) = v;
    }
  }
}

// Training epilogue: a = acc + bias; out (LDS tile) = bf16(silu(a)) and
// sd[row * sld + col] = bf16(silu'(a)) (sd in LDS or global memory).
__device__ __forceinline__ void epi_train(const f32x4 (&acc)[kPJ][4], const float* __restrict__ bias,
                                          __bf16* __restrict__ out, __bf16* __restrict__ sd, int sld, int wave,
                                          int lane) {
  const int li = lane & 15, g = lane >> 4;
#pragma unroll
  for (int cj = 0; cj < kPJ; ++cj) {
    const int col = kPC * wave + 16 * cj + 4 * g;
    const float4 bb = *reinterpret_cast<const float4*>(bias + col);
    const float b4[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
    for (int ri = 0; ri < 4; ++ri) {
      bf16x4h hv, dv;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float h, d;
        silu_both(acc[cj][ri][r] + b4[r], h, d);
        hv[r] = (__bf16)h;
        dv[r] = (__bf16)d;
      }
      const int row = 16 * ri + li;
      *reinterpret_cast<bf16x4h*>(out + row * kHP + col) = hv;
      *reinterpret_cast<bf16x4h*>(sd + (size_t)row * sld + col) = dv;
    }
  }
}

// Backward epilogue: out[row][col..] = bf16(acc * sv) with sv = silu' of the
// layer's pre-activations at the same positions.
__device__ __forceinline__ void epi_bwd(const f32x4 (&acc)[kPJ][4], const bf16x4h (&sv)[kPJ][4],
                                        __bf16* __restrict__ out, int wave, int lane) {
  const int li = lane & 15, g = lane >> 4;
#pragma unroll
  for (int cj = 0; cj < kPJ; ++cj) {
    const int col = kPC * wave + 16 * cj + 4 * g;
#pragma unroll
    for (int ri = 0; ri < 4; ++ri) {
      bf16x4h v;
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = (__bf16)(acc[cj][ri][r] * (float)sv[cj][ri][r]);
      *reinterpret_cast<bf16x4h*>(out + (16 * ri + li) * kHP + col) = v;
    }
  }
}

// This lane's silu' values for epi_bwd: rows r0 + 16 ri + li, columns
// kPC w + 16 cj + 4 g .. + 3 of S (row-major, 256 columns).
__device__ __forceinline__ void load_sv(const unsigned short* __restrict__ S, long long r0, bf16x4h (&sv)[kPJ][4],
                                        int wave, int lane) {
  const int li = lane & 15, g = lane >> 4;
#pragma unroll
  for (int cj = 0; cj < kPJ; ++cj)
#pragma unroll
    for (int ri = 0; ri < 4; ++ri)
      sv[cj][ri] = *reinterpret_cast<const bf16x4h*>(S + (size_t)(r0 + 16 * ri + li) * kH + kPC * wave + 16 * cj +
                                                     4 * g);
}

// The 64 x 256 LDS tile src to dst[c * ldT + r0 + row] (the transposed
// layout, 256 x ldT); rows at or beyond `valid` become zeros.  Lane l reads
// column 64 i + l (consecutive 2-byte words: no bank conflicts), wave w rows
// [8 w, 8 w + 8), one 16-byte store each.
__device__ __forceinline__ void store_transposed(const __bf16* __restrict__ src, unsigned short* __restrict__ dst,
                                                 long long ldT, long long r0, int valid, int tid) {
  const int lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int i = 0; i < kH / 64; ++i) {
    const int c = 64 * i + lane;
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int row = 8 * w + e;
      v[e] = row < valid ? src[row * kHP + c] : (__bf16)0.0f;
    }
    *reinterpret_cast<bf16x8*>(dst + (size_t)c * ldT + r0 + 8 * w) = v;
  }
}

// bf16(q / pi) of rows [r0, r0 + 64) into xs (64 x kXP, zero beyond ndof and
// n); with xt also to xt[k * ldT + r0 + row].
__device__ __forceinline__ void stage_inputs(const float* __restrict__ q, long long off, long long r0, int n, int nd,
                                             __bf16* __restrict__ xs, unsigned short* __restrict__ xt, long long ldT,
                                             int tid) {
  for (int e = tid; e < kFR * kXP; e += kPW * 64) {
    const int row = e & (kFR - 1), k = e / kFR;
    const long long r = r0 + row;
    const __bf16 v = (__bf16)((r < n && k < nd) ? q[r + off * k] * 0.3183098861837907f : 0.0f);
    xs[row * kXP + k] = v;
    if (xt) reinterpret_cast<__bf16*>(xt)[(size_t)k * ldT + r] = v;
  }
}

// ---- inference ---------------------------------------------------------------
// Persistent, one block per CU; W1, W2, W3 in registers (136 VGPRs per
// wave), two waves per SIMD so one wave's epilogue (SiLU costs two
// quarter-rate transcendentals per value) overlaps the other's products.
// LDS: two activation tiles, the staged inputs, b1 | b2 | b3 | w4.
constexpr size_t kPredictLds = 2 * (size_t)kTileElems * 2 + (size_t)kFR * kXP * 2 + 4 * kH * sizeof(float);

__global__ void __launch_bounds__(kPW * 64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    mlp_predict_kernel(const float* __restrict__ q, long long off, int n, int nd, const unsigned short* __restrict__ W1h,
                       const unsigned short* __restrict__ W2h, const unsigned short* __restrict__ W3h,
                       const float* __restrict__ P, float* __restrict__ prob) {
  extern __shared__ __bf16 hs[];
  __bf16* const h0 = hs;
  __bf16* const h1 = hs + kTileElems;
  __bf16* const xs = hs + 2 * kTileElems;
  float* const bs = reinterpret_cast<float*>(xs + kFR * kXP);  // b1 | b2 | b3 | w4
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < kH; e += kPW * 64) {
    bs[e] = P[kOffB1 + e];
    bs[kH + e] = P[kOffB2 + e];
    bs[2 * kH + e] = P[kOffB3 + e];
    bs[3 * kH + e] = P[kOffW4 + e];
  }
  bf16x8 w1[kPJ], w2[kPJ][8], w3[kPJ][8];
  load_w1_frags(W1h, w1, wave, lane);
  load_frags(W2h, w2, wave, lane);
  load_frags(W3h, w3, wave, lane);
  const float bias4 = P[kOffB4];
  const long long stride = (long long)gridDim.x * kFR;
  stage_inputs(q, off, (long long)blockIdx.x * kFR, n, nd, xs, nullptr, 0, tid);
  __syncthreads();
  for (long long r0 = (long long)blockIdx.x * kFR; r0 < n; r0 += stride) {
    f32x4 acc[kPJ][4];
    layer1_mfma(xs, w1, acc, lane);
    epi_predict(acc, bs, h0, wave, lane);
    __syncthreads();
    // xs is free: the next tile's inputs load while the layers run
    stage_inputs(q, off, r0 + stride, n, nd, xs, nullptr, 0, tid);
    layer_mfma(h0, w2, acc, lane);
    epi_predict(acc, bs + kH, h1, wave, lane);
    __syncthreads();
    layer_mfma(h1, w3, acc, lane);
    epi_predict(acc, bs + 2 * kH, h0, wave, lane);
    __syncthreads();
    // head: z = h3 . w4 + b4 (eight threads per row, 32 inputs each),
    // p = sigmoid(z)
    {
      const int row = tid >> 3, part = tid & 7;
      const float* const w4s = bs + 3 * kH;
      float sacc = 0.0f;
#pragma unroll
      for (int k = 32 * part; k < 32 * part + 32; k += 8) {
        const bf16x8 hv = *reinterpret_cast<const bf16x8*>(h0 + row * kHP + k);
#pragma unroll
        for (int j = 0; j < 8; ++j) sacc += (float)hv[j] * w4s[k + j];
      }
      sacc += __shfl_xor(sacc, 1, 64);
      sacc += __shfl_xor(sacc, 2, 64);
      sacc += __shfl_xor(sacc, 4, 64);
      if (part == 0 && r0 + row < n) prob[r0 + row] = rcp(1.0f + __expf(-(sacc + bias4)));
    }
    __syncthreads();  // h0 is the next tile's layer-1 output
  }
}

// ---- training ------------------------------------------------------------------

struct TrainArgs {
  const float* q;
  long long off;
  int n, nd;
  long long ldT;  // row pitch of the transposed arrays (rows >= n padded to a multiple of 64)
  const int* codes;
  const unsigned short *W1h, *W2h, *W3h, *W2t, *W3t;
  const float* P;
  unsigned short *XT, *H1T, *H2T;  // transposed: 16 x ldT, 256 x ldT, 256 x ldT
  unsigned short *S1, *S2, *dA3;   // row-major, 256 columns
  unsigned short *dA3T, *dA2T, *dA1T;
  float* wsF;  // [grid of F][kRowF]
};

// F: LDS h0, h1, s3 (64 x kHP each), xs, b1 | b2 | b3 | w4, dz[64], and
// 16 floats for the block's final sums.
constexpr size_t kTrainFwdLds =
    3 * (size_t)kTileElems * 2 + (size_t)kFR * kXP * 2 + 4 * kH * sizeof(float) + kFR * sizeof(float) + 16 * 4;

__global__ void __launch_bounds__(kPW * 64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    mlp_train_fwd_kernel(TrainArgs a) {
  extern __shared__ __bf16 hs[];
  __bf16* const h0 = hs;
  __bf16* const h1 = hs + kTileElems;
  __bf16* const s3 = hs + 2 * kTileElems;
  __bf16* const xs = hs + 3 * kTileElems;
  float* const bs = reinterpret_cast<float*>(xs + kFR * kXP);
  float* const dzs = bs + 4 * kH;
  float* const fin = dzs + kFR;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int e = tid; e < kH; e += kPW * 64) {
    bs[e] = a.P[kOffB1 + e];
    bs[kH + e] = a.P[kOffB2 + e];
    bs[2 * kH + e] = a.P[kOffB3 + e];
    bs[3 * kH + e] = a.P[kOffW4 + e];
  }
  bf16x8 w1[kPJ], w2[kPJ][8], w3[kPJ][8];
  load_w1_frags(a.W1h, w1, wave, lane);
  load_frags(a.W2h, w2, wave, lane);
  load_frags(a.W3h, w3, wave, lane);
  const float bias4 = a.P[kOffB4];
  const float nf = (float)a.n;
  const float* const w4s = bs + 3 * kH;
  float dw4 = 0.0f, db4 = 0.0f, lsum = 0.0f;
  const long long stride = (long long)gridDim.x * kFR;
  stage_inputs(a.q, a.off, (long long)blockIdx.x * kFR, a.n, a.nd, xs, a.XT, a.ldT, tid);
  __syncthreads();
  for (long long r0 = (long long)blockIdx.x * kFR; r0 < a.n; r0 += stride) {
    const int tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
    const int valid = (int)min<long long>(kFR, a.n - r0);
    f32x4 acc[kPJ][4];
    layer1_mfma(xs, w1, acc, lane);
    epi_train(acc, bs, h0, reinterpret_cast<__bf16*>(a.S1) + (size_t)r0 * kH, kH, wave, lane);
    __syncthreads();
    stage_inputs(a.q, a.off, r0 + stride, a.n, a.nd, xs, r0 + stride < a.n ? a.XT : nullptr, a.ldT, tid);
    store_transposed(h0, a.H1T, a.ldT, r0, valid, tid);
    layer_mfma(h0, w2, acc, lane);
    epi_train(acc, bs + kH, h1, reinterpret_cast<__bf16*>(a.S2) + (size_t)r0 * kH, kH, wave, lane);
    __syncthreads();
    store_transposed(h1, a.H2T, a.ldT, r0, valid, tid);
    layer_mfma(h1, w3, acc, lane);
    epi_train(acc, bs + 2 * kH, h0, s3, kHP, wave, lane);
    __syncthreads();
    // head: z = h3 . w4 + b4, p = sigmoid(z), dz = (p - y) / n (zero for
    // rows beyond n), loss terms
    {
      const int row = tid >> 3, part = tid & 7;
      float sacc = 0.0f;
#pragma unroll
      for (int k = 32 * part; k < 32 * part + 32; k += 8) {
        const bf16x8 hv = *reinterpret_cast<const bf16x8*>(h0 + row * kHP + k);
#pragma unroll
        for (int j = 0; j < 8; ++j) sacc += (float)hv[j] * w4s[k + j];
      }
      sacc += __shfl_xor(sacc, 1, 64);
      sacc += __shfl_xor(sacc, 2, 64);
      sacc += __shfl_xor(sacc, 4, 64);
      if (part == 0) {
        float d = 0.0f;
        if (row < valid) {
          const float z = sacc + bias4;
          const float p = rcp(1.0f + __expf(-z));
          const float y = a.codes[r0 + row] != 0 ? 1.0f : 0.0f;
          d = (p - y) / nf;
          // binary cross-entropy from the logit: max(z,0) - z y + log(1 + e^-|z|)
          lsum += (fmaxf(z, 0.0f) - z * y + log1pf(__expf(-fabsf(z)))) / nf;
          db4 += d;
        }
        dzs[row] = d;
      }
    }
    __syncthreads();
    // dA3 = bf16(dz w4 silu'(A3)), rows of 8 columns per thread; dw4 sums
    // dz H3 over this tile's rows (thread: column tid % 256, half the rows)
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int item = tid + kPW * 64 * it, row = item >> 5, c = (item & 31) * 8;
      const bf16x8 sv = *reinterpret_cast<const bf16x8*>(s3 + row * kHP + c);
      const float d = dzs[row];
      bf16x8 o;
#pragma unroll
      for (int e = 0; e < 8; ++e) o[e] = (__bf16)((d * w4s[c + e]) * (float)sv[e]);
      *reinterpret_cast<bf16x8*>(a.dA3 + (size_t)(r0 + row) * kH + c) = o;
    }
    {
      const int j = tid & (kH - 1), half = tid >> 8;
      for (int rr = 0; rr < 32; ++rr) {
        const int row = 32 * half + rr;
        dw4 += dzs[row] * (float)h0[row * kHP + j];
      }
    }
    __syncthreads();  // h0, s3 and dz are the next tile's
  }
  // this block's partial sums: dw4 (two halves per column), db4 and the loss
  float* const red = reinterpret_cast<float*>(h1);
  red[tid] = dw4;
  float x = db4, l = lsum;
  for (int d = 32; d > 0; d >>= 1) {
    x += __shfl_xor(x, d, 64);
    l += __shfl_xor(l, d, 64);
  }
  if (lane == 0) {
    fin[wave] = x;
    fin[8 + wave] = l;
  }
  __syncthreads();
  float* const out = a.wsF + (size_t)blockIdx.x * kRowF;
  if (tid < kH) out[tid] = red[tid] + red[tid + kH];
  if (tid == 0) {
    float sx = 0.0f, sl = 0.0f;
    for (int w = 0; w < kPW; ++w) {
      sx += fin[w];
      sl += fin[8 + w];
    }
    out[kH] = sx;
    out[kH + 1] = sl;
  }
}

// B: LDS g0, g1 (64 x kHP each).
constexpr size_t kTrainBwdLds = 2 * (size_t)kTileElems * 2;

__global__ void __launch_bounds__(kPW * 64) __attribute__((amdgpu_waves_per_eu(2, 2)))
    mlp_train_bwd_kernel(TrainArgs a) {
  extern __shared__ __bf16 hs[];
  __bf16* const g0 = hs;
  __bf16* const g1 = hs + kTileElems;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // dA2 = dA3 W3: out[row][n] = sum_k dA3[row][k] W3[k][n] = sum_k dA3[row][k] W3t[n][k]
  bf16x8 w3t[kPJ][8], w2t[kPJ][8];
  load_frags(a.W3t, w3t, wave, lane);
  load_frags(a.W2t, w2t, wave, lane);
  for (long long r0 = (long long)blockIdx.x * kFR; r0 < a.n; r0 += (long long)gridDim.x * kFR) {
    const int tid = opaque_tid(), lane = tid & 63, wave = tid >> 6;
    const int valid = (int)min<long long>(kFR, a.n - r0);
    // the dA3 tile (rows beyond n are zeros: dz is zero there)
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const int item = tid + kPW * 64 * it, row = item >> 5, c = (item & 31) * 8;
      *reinterpret_cast<bf16x8*>(g0 + row * kHP + c) =
          *reinterpret_cast<const bf16x8*>(a.dA3 + (size_t)(r0 + row) * kH + c);
    }
    bf16x4h sv[kPJ][4];
    load_sv(a.S2, r0, sv, wave, lane);
    __syncthreads();
    store_transposed(g0, a.dA3T, a.ldT, r0, valid, tid);
    f32x4 acc[kPJ][4];
    layer_mfma(g0, w3t, acc, lane);
    epi_bwd(acc, sv, g1, wave, lane);
    load_sv(a.S1, r0, sv, wave, lane);
    __syncthreads();
    store_transposed(g1, a.dA2T, a.ldT, r0, valid, tid);
    layer_mfma(g1, w2t, acc, lane);
    epi_bwd(acc, sv, g0, wave, lane);
    __syncthreads();
    store_transposed(g0, a.dA1T, a.ldT, r0, valid, tid);
    __syncthreads();  // g0 is the next tile's
  }
}

// W: dW_l[i][j] = sum_k dA_l^T[i][k] H^T[j][k] (both K-major: 16-byte
// fragment loads), db_l[i] = sum_k dA_l^T[i][k].  Block tile TI x TJ, waves
// BI x BJ, each 64 x 16 WN; K in chunks of 32 through double-buffered LDS.
// blockIdx.x: layer (of the two in the launch) and output tile; blockIdx.y:
// the split of K, whose partial sums go to ws[split][...].
struct WgradLayer {
  const unsigned short* A;  // dA_l^T [256 x ldT]
  const unsigned short* B;  // H_(l-1)^T [J x ldT]
  int woff, boff, J;        // W_l and b_l offsets in a partial row; columns of W_l
};
struct WgradArgs {
  WgradLayer L0, L1;
  int tiles;  // output tiles per layer
  long long ldT;
  long long K, kc;  // rows (multiple of 64), rows per split (multiple of 32)
  float* ws;        // [split][row]
  int row;          // floats per partial row
};

template <int BI, int BJ, int WN>
__global__ void __launch_bounds__(256) mlp_wgrad_kernel(WgradArgs a) {
  constexpr int TI = 64 * BI, TJ = 16 * WN * BJ, KC = 32, KP = KC + 8;
  constexpr int NA = TI * (KC / 8) / 256, NB = (TJ * (KC / 8) + 255) / 256;
  __shared__ __bf16 As[2][TI * KP];
  __shared__ __bf16 Bs[2][TJ * KP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, li = lane & 15, g = lane >> 4;
  const bool second = (int)blockIdx.x >= a.tiles;
  const WgradLayer L = second ? a.L1 : a.L0;
  const int tile = (int)blockIdx.x - (second ? a.tiles : 0);
  const int tilesJ = L.J / TJ;
  const int i0 = (tile / tilesJ) * TI, j0 = (tile % tilesJ) * TJ;
  const int wi = wave / BJ, wj = wave % BJ;
  const long long kb = (long long)blockIdx.y * a.kc, ke = min(a.K, kb + a.kc);
  const bool do_bias = j0 == 0 && wj == 0;
  bf16x8 ra[NA], rb[NB];
  auto gload = [&](long long k0) {
#pragma unroll
    for (int m = 0; m < NA; ++m) {
      const int c = tid + 256 * m, row = c / (KC / 8), part = c % (KC / 8);
      ra[m] = *reinterpret_cast<const bf16x8*>(L.A + (size_t)(i0 + row) * a.ldT + k0 + 8 * part);
    }
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int c = tid + 256 * m, row = c / (KC / 8), part = c % (KC / 8);
      if (c < TJ * (KC / 8))
        rb[m] = *reinterpret_cast<const bf16x8*>(L.B + (size_t)(j0 + row) * a.ldT + k0 + 8 * part);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int m = 0; m < NA; ++m) {
      const int c = tid + 256 * m, row = c / (KC / 8), part = c % (KC / 8);
      *reinterpret_cast<bf16x8*>(&As[buf][row * KP + 8 * part]) = ra[m];
    }
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      const int c = tid + 256 * m, row = c / (KC / 8), part = c % (KC / 8);
      if (c < TJ * (KC / 8)) *reinterpret_cast<bf16x8*>(&Bs[buf][row * KP + 8 * part]) = rb[m];
    }
  };
  f32x4 acc[4][WN];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < WN; ++bj) acc[bi][bj] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  float bsum[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (kb < ke) {
    gload(kb);
    sstore(0);
  }
  __syncthreads();
  int buf = 0;
  for (long long k0 = kb; k0 < ke; k0 += KC) {
    const bool more = k0 + KC < ke;
    if (more) gload(k0 + KC);
    bf16x8 af[4], bfr[WN];
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
      af[bi] = *reinterpret_cast<const bf16x8*>(&As[buf][(wi * 64 + 16 * bi + li) * KP + 8 * g]);
#pragma unroll
    for (int bj = 0; bj < WN; ++bj)
      bfr[bj] = *reinterpret_cast<const bf16x8*>(&Bs[buf][(wj * 16 * WN + 16 * bj + li) * KP + 8 * g]);
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < WN; ++bj)
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[bi], bfr[bj], acc[bi][bj], 0, 0, 0);
    if (do_bias) {
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int e = 0; e < 8; ++e) bsum[bi] += (float)af[bi][e];
    }
    if (more) sstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // D(i, j) of a 16 x 16 block: lane holds column j = li, rows 4 g + r
  float* const out = a.ws + (size_t)blockIdx.y * a.row;
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < WN; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wi * 64 + 16 * bi + 4 * g + r, j = j0 + wj * 16 * WN + 16 * bj + li;
        out[L.woff + i * L.J + j] = acc[bi][bj][r];
      }
  if (do_bias) {
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
      float v = bsum[bi];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (g == 0) out[L.boff + i0 + wi * 64 + 16 * bi + li] = v;
    }
  }
}

// R: one thread per parameter (and one for the loss).  mode 0: gradients
// (sums of the partials) into G; 1: also one Adam step on P, M, V; 2: only
// refresh the BF16 copies of W1, W2, W3 from P.  The transposed copies are
// mlp_transpose_kernel's.
struct UpdateArgs {
  float *P, *G, *M, *V, *loss;
  unsigned short *W1h, *W2h, *W3h;
  const float* wsW1;  // [SW1][kOffW2]: W1, b1
  int SW1;
  const float* wsW;   // [SW][kRowW]: W2 .. b3 (at their flat offsets)
  int SW;
  const float* wsF;   // [SF][kRowF]: w4, b4, loss
  int SF;
  int mode;
  float lr, b1, b2, eps, c1, c2;
};

__global__ void mlp_update_kernel(UpdateArgs a) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx > kPtot) return;
  if (idx == kPtot) {
    if (a.mode != 2) {
      float l = 0.0f;
      for (int s0 = 0; s0 < a.SF; s0 += 8) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = s0 + j < a.SF ? a.wsF[(size_t)(s0 + j) * kRowF + kH + 1] : 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j) l += v[j];
      }
      a.loss[0] = l;
    }
    return;
  }
  float w = a.P[idx];
  if (a.mode != 2) {
    const float* src;
    int rows, pitch;
    if (idx < kOffW2) {
      src = a.wsW1 + idx; rows = a.SW1; pitch = kOffW2;
    } else if (idx < kRowW) {
      src = a.wsW + idx; rows = a.SW; pitch = kRowW;
    } else {
      src = a.wsF + (idx - kOffW4); rows = a.SF; pitch = kRowF;
    }
    // in split order, 32 loads in flight at a time
    float gsum = 0.0f;
    for (int s0 = 0; s0 < rows; s0 += 32) {
      float v[32];
#pragma unroll
      for (int j = 0; j < 32; ++j) v[j] = s0 + j < rows ? src[(size_t)(s0 + j) * pitch] : 0.0f;
#pragma unroll
      for (int j = 0; j < 32; ++j) gsum += v[j];
    }
    a.G[idx] = gsum;
    if (a.mode == 1) {
      const float mi = a.b1 * a.M[idx] + (1.0f - a.b1) * gsum;
      const float vi = a.b2 * a.V[idx] + (1.0f - a.b2) * gsum * gsum;
      a.M[idx] = mi;
      a.V[idx] = vi;
      w -= a.lr * (mi / a.c1) / (sqrtf(vi / a.c2) + a.eps);
      a.P[idx] = w;
    }
  }
  const __bf16 h = (__bf16)w;
  if (idx < kOffB1) reinterpret_cast<__bf16*>(a.W1h)[idx] = h;
  else if (idx >= kOffW2 && idx < kOffB2) reinterpret_cast<__bf16*>(a.W2h)[idx - kOffW2] = h;
  else if (idx >= kOffW3 && idx < kOffB3) reinterpret_cast<__bf16*>(a.W3h)[idx - kOffW3] = h;
}

// W2^T and W3^T from the row-major copies, 64 x 64 tiles through LDS (both
// sides coalesced); blockIdx.z picks the matrix.
__global__ void __launch_bounds__(256) mlp_transpose_kernel(const unsigned short* __restrict__ W2h,
                                                            const unsigned short* __restrict__ W3h,
                                                            unsigned short* __restrict__ W2t,
                                                            unsigned short* __restrict__ W3t) {
  __shared__ unsigned short t[64][65];
  const unsigned short* src = blockIdx.z ? W3h : W2h;
  unsigned short* dst = blockIdx.z ? W3t : W2t;
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    t[r][c] = src[(r0 + r) * kH + c0 + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    dst[(c0 + r) * kH + r0 + c] = t[c][r];
  }
}

int device_cus() {
  static const int cus = [] {
    int dev = 0, c = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c < 1)
      c = 256;
    return c;
  }();
  return cus;
}

// The kernels' dynamic LDS exceeds the 64 KB default: raise their limits once.
hipError_t lds_limits() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    err = hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_predict_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPredictLds);
    if (err == hipSuccess)
      err = hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_train_fwd_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTrainFwdLds);
    if (err == hipSuccess)
      err = hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_train_bwd_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kTrainBwdLds);
  });
  return err;
}

}  // namespace

struct Surrogate {
  int cap = 0, ndof = 0, step = 0;
  long long ldT = 0;       // cap rounded up to a multiple of 64
  float *P = nullptr, *G = nullptr, *Mm = nullptr, *Vv = nullptr, *loss = nullptr;
  unsigned short *W1h = nullptr, *W2h = nullptr, *W3h = nullptr, *W2t = nullptr, *W3t = nullptr;
  unsigned short *XT = nullptr, *H1T = nullptr, *H2T = nullptr, *S1 = nullptr, *S2 = nullptr, *dA3 = nullptr,
                 *dA3T = nullptr, *dA2T = nullptr, *dA1T = nullptr;
  float *wsW1 = nullptr, *wsW = nullptr, *wsF = nullptr;  // partial sums (layer-1 and layer-2/3 gradients, F)
  int SW1 = 0, SW = 0, SF = 0;                            // their rows from the last fwd_bwd
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

// mode as mlp_update_kernel's; the partial sums are those of the last
// fwd_bwd.
hipError_t update(Surrogate* s, int mode, float lr, hipStream_t st) {
  UpdateArgs u{};
  u.P = s->P; u.G = s->G; u.M = s->Mm; u.V = s->Vv; u.loss = s->loss;
  u.W1h = s->W1h; u.W2h = s->W2h; u.W3h = s->W3h;
  u.wsW1 = s->wsW1; u.SW1 = s->SW1; u.wsW = s->wsW; u.SW = s->SW; u.wsF = s->wsF; u.SF = s->SF; u.mode = mode;
  u.lr = lr; u.b1 = 0.9f; u.b2 = 0.999f; u.eps = 1e-8f;
  u.c1 = 1.0f - std::pow(u.b1, (float)s->step);
  u.c2 = 1.0f - std::pow(u.b2, (float)s->step);
  hipLaunchKernelGGL(mlp_update_kernel, dim3((kPtot + 1 + 255) / 256), dim3(256), 0, st, u);
  if (mode != 0)
    hipLaunchKernelGGL(mlp_transpose_kernel, dim3(kH / 64, kH / 64, 2), dim3(256), 0, st, s->W2h, s->W3h, s->W2t,
                       s->W3t);
  return hipGetLastError();
}

// F, B, W on n rows (partial sums and their counts in s).
hipError_t fwd_bwd(Surrogate* s, const float* q, long long offset, const int* codes, int n, hipStream_t st) {
  hipError_t e = lds_limits();
  if (e != hipSuccess) return e;
  TrainArgs a{};
  a.q = q; a.off = offset; a.n = n; a.nd = s->ndof; a.ldT = s->ldT; a.codes = codes;
  a.W1h = s->W1h; a.W2h = s->W2h; a.W3h = s->W3h; a.W2t = s->W2t; a.W3t = s->W3t; a.P = s->P;
  a.XT = s->XT; a.H1T = s->H1T; a.H2T = s->H2T; a.S1 = s->S1; a.S2 = s->S2; a.dA3 = s->dA3;
  a.dA3T = s->dA3T; a.dA2T = s->dA2T; a.dA1T = s->dA1T; a.wsF = s->wsF;
  const long long tiles = (n + kFR - 1) / kFR;
  const int grid = (int)std::min<long long>(tiles, device_cus());
  hipLaunchKernelGGL(mlp_train_fwd_kernel, dim3(grid), dim3(kPW * 64), kTrainFwdLds, st, a);
  hipLaunchKernelGGL(mlp_train_bwd_kernel, dim3(grid), dim3(kPW * 64), kTrainBwdLds, st, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // weight gradients: K in splits of at least kMin rows, at most maxS
  // splits (a multiple of 32 rows each)
  const long long K = tiles * kFR;
  auto splits = [K](long long kmin, int maxS, long long* kc) {
    int S = (int)std::max<long long>(1, std::min<long long>(maxS, K / kmin));
    *kc = ((K + S - 1) / S + 31) / 32 * 32;
    return (int)((K + *kc - 1) / *kc);
  };
  WgradArgs w{};
  w.ldT = s->ldT; w.K = K; w.ws = s->wsW; w.row = kRowW;
  s->SW = splits(256, kMaxSplit, &w.kc);
  w.L0 = WgradLayer{s->dA3T, s->H2T, kOffW3, kOffB3, kH};
  w.L1 = WgradLayer{s->dA2T, s->H1T, kOffW2, kOffB2, kH};
  w.tiles = 4;  // 128 x 128 tiles of a 256 x 256 gradient
  hipLaunchKernelGGL((mlp_wgrad_kernel<2, 2, 4>), dim3(2 * w.tiles, s->SW), dim3(256), 0, st, w);
  // layer 1 (256 x 16, one tile per split): more, shorter splits
  w.ws = s->wsW1; w.row = kOffW2;
  s->SW1 = splits(128, kMaxSplit1, &w.kc);
  w.L0 = WgradLayer{s->dA1T, s->XT, kOffW1, kOffB1, kInPad};
  w.L1 = w.L0;
  w.tiles = 1;
  hipLaunchKernelGGL((mlp_wgrad_kernel<4, 1, 1>), dim3(1, s->SW1), dim3(256), 0, st, w);
  s->SF = grid;
  return hipGetLastError();
}

}  // namespace

int surrogate_predict(Surrogate* s, const float* q, int n, long long offset, float* prob, hipStream_t st) {
  if (!s || n < 0) return -1;
  if (n == 0) return 0;
  if (lds_limits() != hipSuccess) return -1;
  // persistent: one block per CU (its waves keep W1, W2 and W3, 264 KB of
  // bf16, in registers), fewer when there are fewer row tiles
  const long long tiles = (n + kFR - 1) / kFR;
  const int grid = (int)std::min<long long>(tiles, device_cus());
  hipLaunchKernelGGL(mlp_predict_kernel, dim3(grid), dim3(kPW * 64), kPredictLds, st, q, offset, n, s->ndof, s->W1h,
                     s->W2h, s->W3h, s->P, prob);
  if (hipGetLastError() != hipSuccess) return -1;
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
  s->ldT = ((long long)batch + kFR - 1) / kFR * kFR;
  const size_t T = (size_t)s->ldT;
  bool ok = dalloc(s, &s->P, kPtot) && dalloc(s, &s->G, kPtot) && dalloc(s, &s->Mm, kPtot) &&
            dalloc(s, &s->Vv, kPtot) && dalloc(s, &s->loss, 1) && dalloc(s, &s->W1h, kH * kInPad) &&
            dalloc(s, &s->W2h, kH * kH) && dalloc(s, &s->W3h, kH * kH) && dalloc(s, &s->W2t, kH * kH) &&
            dalloc(s, &s->W3t, kH * kH) && dalloc(s, &s->XT, kInPad * T) && dalloc(s, &s->H1T, kH * T) &&
            dalloc(s, &s->H2T, kH * T) && dalloc(s, &s->S1, kH * T) && dalloc(s, &s->S2, kH * T) &&
            dalloc(s, &s->dA3, kH * T) && dalloc(s, &s->dA3T, kH * T) && dalloc(s, &s->dA2T, kH * T) &&
            dalloc(s, &s->dA1T, kH * T) && dalloc(s, &s->wsW, (size_t)kMaxSplit * kRowW) &&
            dalloc(s, &s->wsW1, (size_t)kMaxSplit1 * kOffW2) &&
            dalloc(s, &s->wsF, (size_t)device_cus() * kRowF);
  // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases, seeded
  if (ok) {
    Rng rng(seed ^ 0x5DEECE66DULL);
    const int fan_in[8] = {kIn, kIn, kH, kH, kH, kH, kH, kH};
    std::vector<float> h(kPtot), zeros(kPtot, 0.0f);
    for (int t = 0; t < 8; ++t) {
      const double b = 1.0 / std::sqrt((double)fan_in[t]);
      for (int i = kTensorOff[t]; i < kTensorOff[t + 1]; ++i) h[i] = (float)rng.uniform(-b, b);
    }
    for (int i = 0; i < kH * kInPad; ++i)
      if (i % kInPad >= ndof) h[i] = 0.0f;
    ok = hipMemcpy(s->P, h.data(), 4 * (size_t)kPtot, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(s->Mm, zeros.data(), 4 * (size_t)kPtot, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(s->Vv, zeros.data(), 4 * (size_t)kPtot, hipMemcpyHostToDevice) == hipSuccess &&
         update(s, 2, 0.0f, nullptr) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
  }
  if (!ok) {
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
int gprm_surrogate_param_count(void) { return kPtot; }

int gprm_surrogate_get_params(void* h, float* out) {
  Surrogate* s = static_cast<Surrogate*>(h);
  if (!s || !out) return GPRM_E_INVALID;
  if (hipMemcpy(out, s->P, 4 * (size_t)kPtot, hipMemcpyDeviceToHost) != hipSuccess) return GPRM_E_HIP;
  return GPRM_OK;
}

int gprm_surrogate_set_params(void* h, const float* in) {
  Surrogate* s = static_cast<Surrogate*>(h);
  if (!s || !in) return GPRM_E_INVALID;
  if (hipMemcpy(s->P, in, 4 * (size_t)kPtot, hipMemcpyHostToDevice) != hipSuccess) return GPRM_E_HIP;
  if (update(s, 2, 0.0f, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return GPRM_E_HIP;
  return GPRM_OK;
}

// Forward + backward on n <= batch device configurations (SoA, offset) with
// device codes as labels; gradients (same order as the parameters) to
// grads_out (host, may be NULL) and the mean loss to *loss_out.
int gprm_surrogate_gradients(void* h, const float* d_q, long long offset, const int* d_codes, int n, float* grads_out,
                             float* loss_out) {
  Surrogate* s = static_cast<Surrogate*>(h);
  if (!s || !d_q || !d_codes || n < 1 || n > s->cap) return GPRM_E_INVALID;
  if (fwd_bwd(s, d_q, offset, d_codes, n, nullptr) != hipSuccess || update(s, 0, 0.0f, nullptr) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess)
    return GPRM_E_HIP;
  if (loss_out && hipMemcpy(loss_out, s->loss, 4, hipMemcpyDeviceToHost) != hipSuccess) return GPRM_E_HIP;
  if (grads_out && hipMemcpy(grads_out, s->G, 4 * (size_t)kPtot, hipMemcpyDeviceToHost) != hipSuccess)
    return GPRM_E_HIP;
  return GPRM_OK;
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on n <= batch device
// configurations; *loss_out (may be NULL) gets the batch loss before the
// step.  W1's padding columns keep a zero gradient (their inputs are zero),
// so they stay zero.
int gprm_surrogate_train_step(void* h, const float* d_q, long long offset, const int* d_codes, int n, float lr,
                              float* loss_out) {
  Surrogate* s = static_cast<Surrogate*>(h);
  if (!s || !d_q || !d_codes || n < 1 || n > s->cap) return GPRM_E_INVALID;
  if (fwd_bwd(s, d_q, offset, d_codes, n, nullptr) != hipSuccess) return GPRM_E_HIP;
  ++s->step;
  if (update(s, 1, lr, nullptr) != hipSuccess) return GPRM_E_HIP;
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
