// mlp.hip -- kernels of the pose-energy surrogate (see mlp.h).
//
// GEMMs run on gfx950 BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulate): one wavefront owns a 32x32 tile of C = A * B^T as 2x2 MFMA
// blocks, both operands k-contiguous, so every fragment is one 16-byte load
// per lane, and the next K step's four fragments are in flight while the
// current four multiply.  MFMA 16x16x32 lane layout (lane l, r = l % 16, g = l / 16):
//   A fragment  A[r][8g .. 8g+7]       B fragment  B^T[8g .. 8g+7][r]
//   accumulator C[4g .. 4g+3][r]
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mlp.h"

namespace bude {
namespace mlp {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even: v_cvt_pk_bf16_f32, two values per
// instruction (lo in bits 0-15).
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float floatx2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(floatx2_t{lo, hi}, bf16x2_t));
}
__device__ __forceinline__ bf16 to_bf16(float f) { return bf16(pack_bf16(f, 0.f) & 0xffffu); }

__device__ __forceinline__ float from_bf16(bf16 h) { return __uint_as_float(uint32_t(h) << 16); }

// GELU in its tanh form, 0.5 z (1 + tanh(u)) with u = sqrt(2/pi) (z +
// 0.044715 z^3), written branch-free as z * sigmoid(2u) = z / (1 + 2^q) with
// q = -2u * log2(e): 4 VALU + v_exp_f32 + v_rcp_f32.  Saturates correctly
// (2^q -> inf gives 0, 2^q -> 0 gives z).
constexpr float kG1 = -2.3022082f;   // -2 sqrt(2/pi) log2(e)
constexpr float kG3 = -0.10294324f;  // -2 sqrt(2/pi) 0.044715 log2(e)

__device__ __forceinline__ float gelu_sigma(float z) {
  const float q = z * __builtin_fmaf(z * z, kG3, kG1);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(q));
}

__device__ __forceinline__ float gelu(float z) { return z * gelu_sigma(z); }

// d/dz [z s(z)], s = sigmoid(2u): s + z s (1 - s) 2u'(z),
// 2u' = 2 sqrt(2/pi) (1 + 3 * 0.044715 z^2).
__device__ __forceinline__ float gelu_grad(float z) {
  const float s = gelu_sigma(z);
  return s + z * s * (1.0f - s) * __builtin_fmaf(z * z, 0.21406444f, 1.5957691f);
}

__device__ __forceinline__ bf16x8 load8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// Surrogate inputs of pose q: angles / pi, translations / 5, sin and cos of
// the angles, and the 9 entries of the rotation R = Rz(pan) Ry(roll) Rx(tilt)
// they determine (src/serial/fasten.hpp:39-50); the rest of the kIn-wide row
// is zero.
__device__ __forceinline__ void pose_features(const float *__restrict__ poses, size_t stride, size_t q,
                                              float (&f)[kIn]) {
#pragma unroll
  for (int c = 0; c < kIn; c++) f[c] = 0.f;
  float sn[3], cs[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float ang = poses[size_t(a) * stride + q];
    sincosf(ang, &sn[a], &cs[a]);
    f[a] = ang * 0.31830988618379067f;
    f[3 + a] = poses[size_t(3 + a) * stride + q] * 0.2f;
    f[6 + a] = sn[a];
    f[9 + a] = cs[a];
  }
  const float sx = sn[0], cx = cs[0], sy = sn[1], cy = cs[1], sz = sn[2], cz = cs[2];
  f[12] = cy * cz;
  f[13] = sx * sy * cz - cx * sz;
  f[14] = cx * sy * cz + sx * sz;
  f[15] = cy * sz;
  f[16] = sx * sy * sz + cx * cz;
  f[17] = cx * sy * sz - sx * cz;
  f[18] = -sy;
  f[19] = sx * cy;
  f[20] = cx * cy;
}

// One wave per (16 BM) x (16 BN) tile of C: BM x BN MFMA blocks, the next K
// step's BM + BN fragments loaded while the current ones multiply.  32 x 32
// (BM = BN = 2) for the weight gradients and small products; 64 x 64 for
// the large activation products, which then read their A rows half as
// often and issue 16 MFMAs per K step.
template <int BM, int BN>
__global__ __launch_bounds__(64) void gemm_nt_kernel(const GemmArgs g) {
  const int lane = threadIdx.x;
  const int n0 = blockIdx.x * 16 * BN, m0 = blockIdx.y * 16 * BM;
  const int kc = g.K / g.ksplit, kbeg = blockIdx.z * kc, kend = kbeg + kc;
  const int r = lane & 15, grp = lane >> 4;
  const bf16 *ap = g.a + size_t(m0 + r) * g.lda + 8 * grp;
  const bf16 *bp = g.b + size_t(n0 + r) * g.ldb + 8 * grp;
  floatx4 acc[BM][BN];
#pragma unroll
  for (int i = 0; i < BM; i++)
#pragma unroll
    for (int j = 0; j < BN; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // row sums of B over this K slice (the bias gradient when B is dZ^T),
  // by the waves of the first row of tiles: lane (r, grp) sums its own
  // fragment columns, the four lane groups are added at the end
  const bool bsum = g.bias_part && blockIdx.y == 0;
  float bs[BN];
#pragma unroll
  for (int j = 0; j < BN; j++) bs[j] = 0.f;

  bf16x8 fa[BM], fb[BN];
#pragma unroll
  for (int i = 0; i < BM; i++) fa[i] = load8(ap + size_t(16 * i) * g.lda + kbeg);
#pragma unroll
  for (int j = 0; j < BN; j++) fb[j] = load8(bp + size_t(16 * j) * g.ldb + kbeg);
  for (int k = kbeg; k < kend; k += 32) {
    const int kn = k + 32 < kend ? k + 32 : k;  // (the last step reloads its own fragments)
    bf16x8 na[BM], nb[BN];
#pragma unroll
    for (int i = 0; i < BM; i++) na[i] = load8(ap + size_t(16 * i) * g.lda + kn);
#pragma unroll
    for (int j = 0; j < BN; j++) nb[j] = load8(bp + size_t(16 * j) * g.ldb + kn);
#pragma unroll
    for (int i = 0; i < BM; i++)
#pragma unroll
      for (int j = 0; j < BN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    if (bsum) {
#pragma unroll
      for (int j = 0; j < BN; j++) {
        const uint4 u = __builtin_bit_cast(uint4, fb[j]);
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int e = 0; e < 4; e++) bs[j] += __uint_as_float(w[e] << 16) + __uint_as_float(w[e] & 0xffff0000u);
      }
    }
#pragma unroll
    for (int i = 0; i < BM; i++) fa[i] = na[i];
#pragma unroll
    for (int j = 0; j < BN; j++) fb[j] = nb[j];
  }
  if (bsum) {
#pragma unroll
    for (int j = 0; j < BN; j++) {
      bs[j] += __shfl_xor(bs[j], 16);
      bs[j] += __shfl_xor(bs[j], 32);
      if (grp == 0) g.bias_part[size_t(blockIdx.z) * g.N + n0 + 16 * j + r] = bs[j];
    }
  }

#pragma unroll
  for (int i = 0; i < BM; i++) {
#pragma unroll
    for (int j = 0; j < BN; j++) {
      const int n = n0 + 16 * j + r;
      const int mb = m0 + 16 * i + 4 * grp;
      if (g.epi == kEpiOut) {
        if (n == 0)
#pragma unroll
          for (int q = 0; q < 4; q++) g.out_f32[mb + q] = acc[i][j][q] + g.bias[0];
        continue;
      }
      bf16 t[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const size_t o = size_t(mb + q) * g.ldo + n;
        const float c = acc[i][j][q];
        if (g.epi == kEpiHidden) {
          const float z = c + g.bias[n];
          if (g.out_f32) g.out_f32[o] = z;
          t[q] = to_bf16(gelu(z));
          g.out_bf16[o] = t[q];
        } else if (g.epi == kEpiBackGelu) {
          t[q] = to_bf16(c * gelu_grad(g.zprev[o]));
          g.out_bf16[o] = t[q];
        } else if (g.ksplit > 1 || g.defer) {
          g.work[size_t(blockIdx.z) * size_t(g.M) * g.ldo + o] = c;
        } else {
          g.out_f32[o] = c;
        }
      }
      if (g.out_t && g.epi != kEpiF32) {
        bf16 *dst = g.out_t + size_t(n) * g.ldt + mb;  // 4 consecutive rows -> one 8-byte store
        *reinterpret_cast<uint2 *>(dst) =
            make_uint2(uint32_t(t[0]) | (uint32_t(t[1]) << 16), uint32_t(t[2]) | (uint32_t(t[3]) << 16));
      }
    }
  }
}

// C = A B^T for K = 32 KS known at compile time (no split): every K step's
// fragments are requested at once, so the wave waits for one memory
// latency instead of KS in a row, then the epilogue of gemm_nt_kernel.
template <int KS>
__global__ __launch_bounds__(64) void gemm_nt_preload_kernel(const GemmArgs g) {
  const int lane = threadIdx.x;
  const int n0 = blockIdx.x * 32, m0 = blockIdx.y * 32;
  const int r = lane & 15, grp = lane >> 4;
  const bf16 *ap = g.a + size_t(m0 + r) * g.lda + 8 * grp;
  const bf16 *bp = g.b + size_t(n0 + r) * g.ldb + 8 * grp;
  bf16x8 fa[KS][2], fb[KS][2];
#pragma unroll
  for (int k = 0; k < KS; k++)
#pragma unroll
    for (int i = 0; i < 2; i++) {
      fa[k][i] = load8(ap + size_t(16 * i) * g.lda + 32 * k);
      fb[k][i] = load8(bp + size_t(16 * i) * g.ldb + 32 * k);
    }
  __builtin_amdgcn_sched_barrier(0);  // (all loads issued before the first multiply)
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < KS; k++)
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[k][i], fb[k][j], acc[i][j], 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 2; i++) {
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const int n = n0 + 16 * j + r;
      const int mb = m0 + 16 * i + 4 * grp;
      if (g.epi == kEpiOut) {
        if (n == 0)
#pragma unroll
          for (int q = 0; q < 4; q++) g.out_f32[mb + q] = acc[i][j][q] + g.bias[0];
        continue;
      }
      bf16 t[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const size_t o = size_t(mb + q) * g.ldo + n;
        const float c = acc[i][j][q];
        if (g.epi == kEpiHidden) {
          const float z = c + g.bias[n];
          if (g.out_f32) g.out_f32[o] = z;
          t[q] = to_bf16(gelu(z));
          g.out_bf16[o] = t[q];
        } else if (g.epi == kEpiBackGelu) {
          t[q] = to_bf16(c * gelu_grad(g.zprev[o]));
          g.out_bf16[o] = t[q];
        } else {
          g.out_f32[o] = c;
        }
      }
      if (g.out_t && g.epi != kEpiF32) {
        bf16 *dst = g.out_t + size_t(n) * g.ldt + mb;
        *reinterpret_cast<uint2 *>(dst) =
            make_uint2(uint32_t(t[0]) | (uint32_t(t[1]) << 16), uint32_t(t[2]) | (uint32_t(t[3]) << 16));
      }
    }
  }
}

// out[i] = sum over z of work[z][i], z ascending (fixed order).
__global__ void ksplit_reduce_kernel(const float *__restrict__ work, size_t n, int splits, float *__restrict__ out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc = 0.f;
  for (int z = 0; z < splits; z++) acc += work[size_t(z) * n + i];
  out[i] = acc;
}

__global__ void features_kernel(const float *__restrict__ poses, size_t stride, const uint32_t *__restrict__ idx,
                                size_t n, bf16 *__restrict__ out, bf16 *__restrict__ out_t, int ldt,
                                const float *__restrict__ desc, const uint32_t *__restrict__ lig, int nfeat) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t q = idx ? idx[i] : i;
  float f[kIn];
  pose_features(poses, stride, q, f);
#pragma unroll
  for (int c = 0; c < kFeat; c++)
    if (c >= nfeat) f[c] = 0.f;
  if (desc) {
    const float *d = desc + size_t(lig[q]) * kDesc;
#pragma unroll
    for (int c = 0; c < kDesc; c++) f[kFeat + c] = d[c];
  }
  bf16 h[kIn];
#pragma unroll
  for (int c = 0; c < kIn; c++) h[c] = to_bf16(f[c]);
  uint4 *dst = reinterpret_cast<uint4 *>(out + i * kIn);
#pragma unroll
  for (int v = 0; v < kIn / 8; v++)
    dst[v] = make_uint4(uint32_t(h[8 * v + 0]) | (uint32_t(h[8 * v + 1]) << 16),
                        uint32_t(h[8 * v + 2]) | (uint32_t(h[8 * v + 3]) << 16),
                        uint32_t(h[8 * v + 4]) | (uint32_t(h[8 * v + 5]) << 16),
                        uint32_t(h[8 * v + 6]) | (uint32_t(h[8 * v + 7]) << 16));
  if (out_t)
#pragma unroll
    for (int c = 0; c < kIn; c++) out_t[size_t(c) * ldt + i] = h[c];
}

__global__ void loss_grad_kernel(const float *__restrict__ pred, const float *__restrict__ energy,
                                 const uint32_t *__restrict__ idx, size_t n, TargetNorm norm,
                                 const TargetNorm *__restrict__ table, const uint32_t *__restrict__ lig,
                                 float tail_weight, bf16 *__restrict__ dz, bf16 *__restrict__ dz_t, int ldt,
                                 unsigned long long *__restrict__ loss) {
  const size_t b = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float sq = 0.f;
  if (b < n) {
    const size_t q = idx ? idx[b] : b;
    const float e = energy[q];
    const TargetNorm t = table ? table[lig[q]] : norm;
    const float y = (fminf(e, t.clip) - t.mu) * t.inv_sigma;
  const float d = pred[b] - y;
    sq = d * d;
    const float wgt = e < t.tail ? tail_weight : 1.0f;
    const bf16 g = to_bf16(2.0f * wgt * d / float(n));
    dz[b * kOut] = g;
#pragma unroll
    for (int c = 1; c < kOut; c++) dz[b * kOut + c] = 0;
    dz_t[b] = g;
#pragma unroll
    for (int c = 1; c < kOut; c++) dz_t[size_t(c) * ldt + b] = 0;
  }
  // wave reduction of the squared error, one fixed-point atomic per wave
  // (clamped at 2^40 units so that no run of waves wraps the sum)
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off);
  if ((threadIdx.x & 63) == 0 && loss)
    atomicAdd(loss, static_cast<unsigned long long>(fmin(double(sq) * kLossScale + 0.5, 1099511627776.0)));
}

// One work-group of W waves per 64 consecutive entries of the flat index
// over the batch's tensors: wave w sums the lane's entry over slices w,
// w + W, ... (each load a coalesced row of 64 entries), the work-group adds
// the waves' partial sums in wave order, and the first wave applies Adam.
// The slice sums are spread over the waves so that their loads are in
// flight together (W = 16 for the hundreds of slices of the fused step, 4
// for the K splits of the GEMMs); the order of every sum is fixed.
template <int W>
__global__ __launch_bounds__(64 * W) void adam_multi_kernel(const AdamBatch b, float lr, float b1, float b2, float eps,
                                                            float bc1, float bc2, const float *__restrict__ hyper,
                                                            StepState *__restrict__ advance) {
  constexpr int kAdamWaves = W;
  __shared__ float part[kAdamWaves][64];
  if (hyper) {  // learning rate and bias corrections of this step from device memory
    lr = hyper[0];
    bc1 = hyper[2];
    bc2 = hyper[3];
  }
  // (the step count is read only by the next step's first kernel)
  if (advance && blockIdx.x == 0 && threadIdx.x == 0) advance->step += 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  size_t i = size_t(blockIdx.x) * 64 + lane;
  int t = 0;
  for (; t < b.count; t++) {
    const size_t n = size_t(b.t[t].rows) * b.t[t].cols;
    if (i < n) break;
    i -= n;
  }
  const bool valid = t < b.count;
  float gs = 0.f;
  if (valid) {
    const float *g = b.t[t].g + i;
    const size_t stride = b.t[t].stride;
    const int splits = b.t[t].splits;
    int z = wave;
    for (; z + 3 * kAdamWaves < splits; z += 4 * kAdamWaves) {
      const float s0 = g[size_t(z) * stride], s1 = g[size_t(z + kAdamWaves) * stride];
      const float s2 = g[size_t(z + 2 * kAdamWaves) * stride], s3 = g[size_t(z + 3 * kAdamWaves) * stride];
      gs = ((gs + s0) + s1) + s2 + s3;
    }
    for (; z < splits; z += kAdamWaves) gs += g[size_t(z) * stride];
  }
  part[wave][lane] = gs;
  __syncthreads();
  if (wave != 0 || !valid) return;
  float gi = part[0][lane];
#pragma unroll
  for (int w = 1; w < kAdamWaves; w++) gi += part[w][lane];
  const AdamTensor &T = b.t[t];
  const float mi = b1 * T.m[i] + (1.f - b1) * gi;
  const float vi = b2 * T.v[i] + (1.f - b2) * gi * gi;
  T.m[i] = mi;
  T.v[i] = vi;
  const float wi = T.w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  T.w[i] = wi;
  const bf16 h = to_bf16(wi);
  if (T.w16) T.w16[i] = h;
  if (T.w16t) T.w16t[size_t(i % T.cols) * T.rows + i / T.cols] = h;
}

__global__ void begin_step_kernel(StepState *__restrict__ st, const uint32_t *__restrict__ perm, uint32_t per_epoch,
                                  uint32_t B, uint32_t *__restrict__ idx, uint32_t steps, uint32_t warm, float lr0,
                                  float lr1, float b1, float b2, uint32_t step0) {
  const uint32_t s = st->step;
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < B) idx[i] = perm[size_t(s % per_epoch) * B + i];
  if (i == 0) schedule_hyper(Schedule{steps, warm, step0, lr0, lr1, b1, b2}, s, st->hyper);
}

__global__ void to_bf16_kernel(const float *__restrict__ w, int rows, int cols, bf16 *__restrict__ w16,
                               bf16 *__restrict__ w16t) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(rows) * cols) return;
  const bf16 h = to_bf16(w[i]);
  if (w16) w16[i] = h;
  if (w16t) w16t[size_t(i % cols) * rows + i / cols] = h;
}

// ---- fused inference ---------------------------------------------------------
// One work-group (8 waves) scores tiles of kTile poses through the whole
// network with the activations in one LDS buffer (pose-major rows):
// features -> 3 x (MFMA GEMM + bias + GELU) -> output layer, also on MFMA.
// A hidden layer computes its transpose, D[neuron][pose] = W . act^T, so the
// weights are the A operand (streamed from L2 as k-contiguous rows of W^T,
// the next k-step's fragments loaded while the current one multiplies) and
// the activations the B operand (16-byte LDS reads along k); a lane's four
// results are then four consecutive neurons of one pose (accumulated onto
// their biases), which GELU, two packed BF16 conversions and one 8-byte LDS
// store finish.  Wave w owns
// neurons [32w, 32w + 32) and holds them in registers until every wave has
// read the layer's input, then overwrites the buffer.  128-pose tiles halve
// the weight traffic per pose against 64-pose tiles, and one 66 KiB buffer
// leaves room for two work-groups per CU.
constexpr int kTile = 128;
constexpr int kInferThreads = 512;
constexpr int kLdsRow = kHidden + 8;  // bf16 per LDS row (padding staggers banks)

__device__ __forceinline__ bf16x8 lds_load8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

template <int K>
__device__ __forceinline__ void fused_layer(bf16 *__restrict__ buf, const bf16 *__restrict__ wt,
                                            const float *__restrict__ bias, int lane, int wave) {
  const int r = lane & 15, grp = lane >> 4;
  floatx4 acc[2][kTile / 16];  // [neuron block of 16][pose block of 16], from the biases
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const float4 b4 = *reinterpret_cast<const float4 *>(bias + 32 * wave + 16 * j + 4 * grp);
#pragma unroll
    for (int i = 0; i < kTile / 16; i++) acc[j][i] = floatx4{b4.x, b4.y, b4.z, b4.w};
  }
  const bf16 *wrow = wt + size_t(32 * wave + r) * K + 8 * grp;
  bf16x8 fw[2];
#pragma unroll
  for (int j = 0; j < 2; j++) fw[j] = load8(wrow + size_t(16 * j) * K);
#pragma unroll 1
  for (int k = 0; k < K; k += 32) {
    bf16x8 fn[2];
    const int kn = k + 32 < K ? k + 32 : k;  // (the last step reloads its own fragments)
#pragma unroll
    for (int j = 0; j < 2; j++) fn[j] = load8(wrow + size_t(16 * j) * K + kn);
#pragma unroll
    for (int i = 0; i < kTile / 16; i++) {
      // (two halves: the compiler would otherwise hold all eight B fragments)
      if (i == kTile / 32) __builtin_amdgcn_sched_barrier(0);
      const bf16x8 fa = lds_load8(buf + (16 * i + r) * kLdsRow + k + 8 * grp);
#pragma unroll
      for (int j = 0; j < 2; j++) acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw[j], fa, acc[j][i], 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 2; j++) fw[j] = fn[j];
  }
  // The epilogue (GELU, conversion) is the VALU-heavy phase; raising
  // the wave's issue priority through it lets the other work-group's MFMA
  // waves fill the gaps instead of delaying it (-4 to -5 % per launch).
  __builtin_amdgcn_s_setprio(2);
  __syncthreads();  // every wave has read this layer's input
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const int n0 = 32 * wave + 16 * j + 4 * grp;  // this lane's four neurons
#pragma unroll
    for (int i = 0; i < kTile / 16; i++) {
      const floatx4 z = acc[j][i];
      uint2 v;
      v.x = pack_bf16(gelu(z[0]), gelu(z[1]));
      v.y = pack_bf16(gelu(z[2]), gelu(z[3]));
      *reinterpret_cast<uint2 *>(buf + (16 * i + r) * kLdsRow + n0) = v;
    }
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();
}

// (4 waves per SIMD, two work-groups per CU: within the 128 registers this
// allows, the compiler spills a few kernel-lifetime values once and reloads
// them a few times per layer, which measured faster than the 146 registers
// and one work-group per CU it picks otherwise)
__global__ __launch_bounds__(kInferThreads, 4) void mlp_infer_kernel(
    const float *__restrict__ poses, size_t stride, size_t n, const bf16 *__restrict__ w0t,
    const float *__restrict__ b0, const bf16 *__restrict__ w1t, const float *__restrict__ b1,
    const bf16 *__restrict__ w2t, const float *__restrict__ b2, const bf16 *__restrict__ w3,
    const float *__restrict__ b3, float *__restrict__ out) {
  __shared__ bf16 act[kTile * kLdsRow];
  __shared__ bf16 w3s[kHidden];  // column 0 of W_3: the output neuron's weights
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid < kHidden) w3s[tid] = w3[size_t(tid) * kOut];
  const float bias3 = b3[0];
  const size_t ntiles = (n + kTile - 1) / kTile;
  for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const size_t base = tile * kTile;
    // features of the tile's poses (rows past n are zero), as packed pairs
    if (tid < kTile) {
      const size_t i = base + tid;
      uint32_t *row = reinterpret_cast<uint32_t *>(act + tid * kLdsRow);
      float f[kIn];
      if (i < n) {
        pose_features(poses, stride, i, f);
      } else {
#pragma unroll
        for (int c = 0; c < kIn; c++) f[c] = 0.f;
      }
#pragma unroll
      for (int c = 0; c < kIn; c += 2) row[c / 2] = pack_bf16(f[c], f[c + 1]);
    }
    __syncthreads();
    fused_layer<kIn>(act, w0t, b0, lane, wave);
    fused_layer<kHidden>(act, w1t, b1, lane, wave);
    fused_layer<kHidden>(act, w2t, b2, lane, wave);
    // output layer on MFMA: wave w scores poses [16w, 16w + 16) as row 0 of
    // D = W_3^T . act^T (A rows 1-15 are zero); lanes 0-15 hold the results
    {
      const int r = lane & 15, grp = lane >> 4;
      floatx4 acc = floatx4{0.f, 0.f, 0.f, 0.f};
      const bf16x8 zero = bf16x8{};
#pragma unroll
      for (int k = 0; k < kHidden; k += 32) {
        const bf16x8 fb = lds_load8(act + (16 * wave + r) * kLdsRow + k + 8 * grp);
        const bf16x8 fa = r == 0 ? lds_load8(w3s + k + 8 * grp) : zero;
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, acc, 0, 0, 0);
      }
      if (lane < 16 && base + 16 * wave + lane < n) out[base + 16 * wave + lane] = acc[0] + bias3;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void fold_bias_kernel(const float *__restrict__ b, const bf16 *__restrict__ w16,
                                                        const float *__restrict__ desc, float *__restrict__ out) {
  const int l = blockIdx.x, j = threadIdx.x;
  float acc = b[j];
#pragma unroll
  for (int c = 0; c < kDesc; c++)
    acc = __builtin_fmaf(from_bf16(to_bf16(desc[size_t(l) * kDesc + c])), from_bf16(w16[size_t(kFeat + c) * kHidden + j]),
                         acc);
  out[size_t(l) * kHidden + j] = acc;
}

unsigned blocks_for(size_t n, unsigned t) { return unsigned((n + t - 1) / t); }

}  // namespace

hipError_t gemm(const GemmArgs &g0, hipStream_t s) {
  GemmArgs g = g0;
  if (g.ksplit < 1) g.ksplit = 1;
  if (g.M % 32 || g.N % 32 || g.K % 32 || g.lda % 8 || g.ldb % 8 || (g.K / 32) % g.ksplit) return hipErrorInvalidValue;
  if ((g.ksplit > 1 || g.defer || g.bias_part) && (g.epi != kEpiF32 || !g.work)) return hipErrorInvalidValue;
  // K = 256 without a split (the hidden layers' products): every fragment requested at once
  if (g.ksplit == 1 && g.K == 256 && !g.defer && !g.bias_part) {
    hipLaunchKernelGGL((gemm_nt_preload_kernel<8>), dim3(unsigned(g.N / 32), unsigned(g.M / 32), 1u), dim3(64), 0, s,
                       g);
  } else {
    hipLaunchKernelGGL((gemm_nt_kernel<2, 2>), dim3(unsigned(g.N / 32), unsigned(g.M / 32), unsigned(g.ksplit)),
                       dim3(64), 0, s, g);
  }
  if (g.ksplit > 1 && !g.defer) {
    if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
    const size_t n = size_t(g.M) * g.ldo;
    hipLaunchKernelGGL(ksplit_reduce_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, g.work, n, g.ksplit,
                       g.out_f32);
  }
  return hipGetLastError();
}

int auto_ksplit(int M, int N, int K) {
  const int tiles = (M / 32) * (N / 32);
  int split = 1;
  while (split * 2 * tiles <= 1024 && (K / 32) % (split * 2) == 0 && K / (split * 2) >= 256) split *= 2;
  return split;
}

int max_ksplit(int M, int N) { return std::max(1, 1024 / std::max(1, (M / 32) * (N / 32))); }

hipError_t sum_slices(const float *work, int splits, size_t n, float *out, hipStream_t s) {
  if (splits < 1) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(ksplit_reduce_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, work, n, splits, out);
  return hipGetLastError();
}

hipError_t features(const float *poses, size_t stride, const uint32_t *idx, size_t n, bf16 *out, bf16 *out_t, int ldt,
                    hipStream_t s, const float *desc, const uint32_t *lig, int nfeat) {
  if (n == 0) return hipSuccess;
  if ((desc && !lig) || nfeat < 0 || nfeat > kFeat) return hipErrorInvalidValue;
  hipLaunchKernelGGL(features_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, poses, stride, idx, n, out, out_t,
                     ldt, desc, lig, nfeat);
  return hipGetLastError();
}

hipError_t loss_grad(const float *pred, const float *energy, const uint32_t *idx, size_t n, TargetNorm norm,
                     const TargetNorm *table, const uint32_t *lig, float tail_weight, bf16 *dz, bf16 *dz_t, int ldt,
                     unsigned long long *loss_acc, hipStream_t s) {
  if (table && !lig) return hipErrorInvalidValue;
  hipLaunchKernelGGL(loss_grad_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, pred, energy, idx, n, norm, table,
                     lig, tail_weight, dz, dz_t, ldt, loss_acc);
  return hipGetLastError();
}

hipError_t fold_bias(const float *b, const bf16 *w16, const float *desc, int nlig, float *out, hipStream_t s) {
  if (nlig <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fold_bias_kernel, dim3(unsigned(nlig)), dim3(kHidden), 0, s, b, w16, desc, out);
  return hipGetLastError();
}

hipError_t begin_step(StepState *st, const uint32_t *perm, uint32_t per_epoch, uint32_t B, uint32_t *idx,
                      uint32_t steps, uint32_t warm, float lr0, float lr1, float b1, float b2, uint32_t step0,
                      hipStream_t s) {
  if (B == 0 || per_epoch == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(begin_step_kernel, dim3(blocks_for(B, 256)), dim3(256), 0, s, st, perm, per_epoch, B, idx, steps,
                     warm, lr0, lr1, b1, b2, step0);
  return hipGetLastError();
}

hipError_t adam_multi_dev(const AdamBatch &b, const float *hyper, float b1, float b2, float eps, hipStream_t s,
                          StepState *advance) {
  size_t total = 0;
  int splits = 1;
  for (int t = 0; t < b.count; t++) {
    total += size_t(b.t[t].rows) * b.t[t].cols;
    splits = std::max(splits, b.t[t].splits);
  }
  if (splits > 32)
    hipLaunchKernelGGL(adam_multi_kernel<16>, dim3(blocks_for(total, 64)), dim3(64 * 16), 0, s, b, 0.f, b1, b2, eps,
                       1.f, 1.f, hyper, advance);
  else
    hipLaunchKernelGGL(adam_multi_kernel<4>, dim3(blocks_for(total, 64)), dim3(64 * 4), 0, s, b, 0.f, b1, b2, eps, 1.f,
                       1.f, hyper, advance);
  return hipGetLastError();
}

hipError_t to_bf16(const float *w, int rows, int cols, bf16 *w16, bf16 *w16t, hipStream_t s) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(size_t(rows) * cols, 256)), dim3(256), 0, s, w, rows, cols, w16,
                     w16t);
  return hipGetLastError();
}

hipError_t infer(const float *poses, size_t stride, size_t n, const bf16 *const wt[4], const bf16 *w3,
                 const float *const bias[4], float *out, int blocks, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t tiles = (n + kTile - 1) / kTile;
  const unsigned grid = unsigned(std::min<size_t>(tiles, size_t(blocks)));
  hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(kInferThreads), 0, s, poses, stride, n, wt[0], bias[0], wt[1], bias[1],
                     wt[2], bias[2], w3, bias[3], out);
  return hipGetLastError();
}

}  // namespace mlp
}  // namespace bude
