// mlp.hip -- kernels of the pose-energy surrogate (see mlp.h).
//
// Products run on gfx950 BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulate); operands read from global memory are fragment-tiled
// (tiled_index), so every fragment is one contiguous 1 KB load (16 bytes per
// lane).  MFMA 16x16x32 lane layout (lane l, r = l % 16, g = l / 16):
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

// GELU(z) and its derivative d/dz [z s(z)] = s + z s (1 - s) 2u'(z),
// s = sigmoid(2u), 2u' = 2 sqrt(2/pi) (1 + 3 * 0.044715 z^2): one s for both.
__device__ __forceinline__ void gelu_both(float z, float &a, float &g) {
  const float s = gelu_sigma(z);
  a = z * s;
  g = s + z * s * (1.0f - s) * __builtin_fmaf(z * z, 0.21406444f, 1.5957691f);
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

// ---- the rows of a training step, fused ---------------------------------------
// One work-group of kRBW = 8 waves per kRB = 32 rows of the batch does what a
// training step computes row by row: the input features, the forward pass of
// the three hidden layers and of the output neuron, the loss and its
// gradient, and the backward pass down to dZ0.  It writes the transposed
// BF16 rows the weight-gradient products read (X^T in pose mode, A_l^T,
// dZ_l^T, dz3^T) and the loss; the weight gradients are the products after
// it.  The layers run transposed, D[neuron][row] = W^T X^T: wave w owns
// neurons 32 w .. 32 w + 31 of every hidden layer (kNB = 2 MFMA blocks) and
// all 32 rows (2 blocks), two waves per SIMD, so that one's MFMAs run beside
// the other's epilogue; lane (r, grp) of row block i holds row 2 r + i, so a
// lane's two rows are adjacent in the transposed outputs (one 4-byte store
// per neuron).  The weights are the A operands, read from global memory (all
// work-groups read the same 0.3 MB: L2); the layer below is the B operand,
// from LDS, row 2 r + i kept at LDS row r + 16 i (the 16 rows one operand
// read touches then fall on distinct bank groups).  GELU'(Z_l), rounded to
// BF16, waits in LDS for the backward pass, where the lane that wrote it
// reads it back: wave w holds the same neurons of dA_l as of Z_l.  Each
// product's weight fragments are all loaded two products ahead (two register
// buffers; the first two while the input rows are formed), so their latency
// passes behind the work in between.
constexpr int kRB = kRowBlock;
constexpr int kRBW = 8;                    // waves per work-group
constexpr int kNB = kHidden / (16 * kRBW);  // 16-neuron MFMA blocks per wave
constexpr int kRBPitch = kHidden + 8;  // bf16 per LDS row (528 bytes: 4-bank stagger between rows)
constexpr int kRBIn = kIn + 8;

// A fragments of the wave's 16 kNB rows of the fragment-tiled [256][K] A
// for all K / 32 k steps at once (each one contiguous 1 KB load): issued a
// layer ahead, so their latency passes behind the current layer's epilogue
// and barrier.
template <int K>
__device__ __forceinline__ void rows_load(const bf16 *__restrict__ A, int wave, int r, int grp,
                                          bf16x8 (&fa)[8][kNB]) {
  const bf16 *ap = A + 8 * (r + 16 * grp);
#pragma unroll
  for (int k = 0; k < K / 32; k++)
#pragma unroll
    for (int jn = 0; jn < kNB; jn++) fa[k][jn] = load8(ap + size_t((kNB * wave + jn) * (K / 32) + k) * 512);
}

// acc[jn][i] = sum_k A[16 kNB w + 16 jn + r][k] B[k][row 2 r + i] from the
// loaded fragments, B rows in LDS (row 2 r + i at r + 16 i).
template <int K>
__device__ __forceinline__ void rows_mma(const bf16x8 (&fa)[8][kNB], const bf16 *B, int pitch, int r, int grp,
                                         floatx4 (&acc)[kNB][2]) {
#pragma unroll
  for (int jn = 0; jn < kNB; jn++) acc[jn][0] = acc[jn][1] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < K / 32; k++) {
    bf16x8 fb[2];
#pragma unroll
    for (int i = 0; i < 2; i++) fb[i] = *reinterpret_cast<const bf16x8 *>(B + (r + 16 * i) * pitch + 32 * k + 8 * grp);
#pragma unroll
    for (int jn = 0; jn < kNB; jn++)
#pragma unroll
      for (int i = 0; i < 2; i++) acc[jn][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[k][jn], fb[i], acc[jn][i], 0, 0, 0);
  }
}

__device__ __forceinline__ float bf16_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t v) { return __uint_as_float(v & 0xffff0000u); }

// The 4 x 4 transpose of w[c] across the lanes q = lane mod 4 of each quad:
// afterwards lane q holds w[c] of lane c, word q (two exchanges of two words
// each, by DPP quad permutations).
__device__ __forceinline__ void quad_transpose(uint32_t (&w)[4], int q) {
  {  // lanes q, q ^ 1 swap the words whose bit 0 differs from q's
    const bool odd = q & 1;
    const uint32_t s0 = odd ? w[0] : w[1], s1 = odd ? w[2] : w[3];
    const uint32_t r0 = uint32_t(__builtin_amdgcn_mov_dpp(int(s0), 0xb1, 0xf, 0xf, false));  // quad_perm 1 0 3 2
    const uint32_t r1 = uint32_t(__builtin_amdgcn_mov_dpp(int(s1), 0xb1, 0xf, 0xf, false));
    if (odd) {
      w[0] = r0;
      w[2] = r1;
    } else {
      w[1] = r0;
      w[3] = r1;
    }
  }
  {  // lanes q, q ^ 2 swap the words whose bit 1 differs from q's
    const bool hi = q & 2;
    const uint32_t s0 = hi ? w[0] : w[2], s1 = hi ? w[1] : w[3];
    const uint32_t r0 = uint32_t(__builtin_amdgcn_mov_dpp(int(s0), 0x4e, 0xf, 0xf, false));  // quad_perm 2 3 0 1
    const uint32_t r1 = uint32_t(__builtin_amdgcn_mov_dpp(int(s1), 0x4e, 0xf, 0xf, false));
    if (hi) {
      w[0] = r0;
      w[1] = r1;
    } else {
      w[2] = r0;
      w[3] = r1;
    }
  }
}

// v[jn][i][c] (neuron 16 kNB w + 16 jn + 4 grp + c, row 2 r + i) as BF16: to the
// LDS rows (if set) and to the tiled transposed rows out_t [neuron][ld].  A
// lane's word c holds rows 2 r, 2 r + 1 of neuron 4 grp + c; after the quad
// transpose lane (r, grp) holds rows 8 (r / 4) .. + 7 of neuron 4 grp + r % 4,
// 16 contiguous bytes of the tile, so one store per lane writes the whole
// 16-neuron x 32-row tile (1 KB).
__device__ __forceinline__ void rows_store(const uint32_t (&v)[kNB][2][2], bf16 *lds, bf16 *__restrict__ out_t,
                                           size_t ld, size_t row0, int wave, int r, int grp) {
#pragma unroll
  for (int jn = 0; jn < kNB; jn++) {
    const int n0 = 16 * kNB * wave + 16 * jn + 4 * grp;
    if (lds) {
#pragma unroll
      for (int i = 0; i < 2; i++)
        *reinterpret_cast<uint2 *>(lds + (r + 16 * i) * kRBPitch + n0) = make_uint2(v[jn][i][0], v[jn][i][1]);
    }
    uint32_t w[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const uint32_t e0 = c & 1 ? v[jn][0][c >> 1] >> 16 : v[jn][0][c >> 1] & 0xffffu;
      const uint32_t e1 = c & 1 ? v[jn][1][c >> 1] & 0xffff0000u : v[jn][1][c >> 1] << 16;
      w[c] = e0 | e1;
    }
    quad_transpose(w, r & 3);
    const size_t tile = tiled_index(size_t(16 * kNB * wave + 16 * jn), row0, ld);
    const int slot = 4 * grp + (r & 3) + 16 * (r >> 2);
    *reinterpret_cast<uint4 *>(out_t + tile + size_t(8 * slot)) = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

__global__ __launch_bounds__(64 * kRBW) void rows_kernel(const RowsArgs a) {
  __shared__ __align__(16) bf16 xs[kRB][kRBIn];
  __shared__ __align__(16) bf16 hs[2][kRB][kRBPitch];
  __shared__ __align__(16) bf16 gs[3][kRB][kRBPitch];  // GELU'(Z_l), read back by the lane that wrote it
  __shared__ float ypart[kRBW][kRB];
  __shared__ float dz3s[kRB];
  __shared__ uint32_t qs[kRB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, r = lane & 15, grp = lane >> 4;
  const size_t row0 = size_t(blockIdx.x) * kRB;
  if (a.perm && blockIdx.x == 0 && tid == 0) schedule_hyper(a.sched, a.state->step, a.state->hyper);
  // the weight fragments of the products, two products ahead in two buffers:
  // layers 0 and 1 in flight while the input rows are formed
  bf16x8 fx[8][kNB], fy[8][kNB];
  rows_load<kIn>(a.wt[0], wave, r, grp, fx);
  rows_load<kHidden>(a.wt[1], wave, r, grp, fy);
  // ---- the input row p = tid < kRB (LDS row p / 2 + 16 (p % 2))
  if (tid < kRB) {
    const int p = tid;
    const size_t row = row0 + p;
    const uint32_t q = a.perm ? a.perm[size_t(a.state->step % a.per_epoch) * a.n + row]
                              : (a.idx ? a.idx[row] : uint32_t(row));
    qs[p] = q;
    uint4 h[kIn / 8];
    if (a.feat) {
#pragma unroll
      for (int v = 0; v < kIn / 8; v++) h[v] = reinterpret_cast<const uint4 *>(a.feat + row * kIn)[v];
    } else {
      float f[kIn];
      pose_features(a.poses, a.stride, q, f);
#pragma unroll
      for (int c = 0; c < kFeat; c++)
        if (c >= a.nfeat) f[c] = 0.f;
      if (a.desc) {
        const float *d = a.desc + size_t(a.lig[q]) * kDesc;
#pragma unroll
        for (int c = 0; c < kDesc; c++) f[kFeat + c] = d[c];
      }
#pragma unroll
      for (int v = 0; v < kIn / 8; v++)
        h[v] = make_uint4(pack_bf16(f[8 * v], f[8 * v + 1]), pack_bf16(f[8 * v + 2], f[8 * v + 3]),
                          pack_bf16(f[8 * v + 4], f[8 * v + 5]), pack_bf16(f[8 * v + 6], f[8 * v + 7]));
      const uint32_t *hw = reinterpret_cast<const uint32_t *>(h);
#pragma unroll
      for (int c = 0; c < kIn; c++) a.feat_t[tiled_index(c, row, a.ld_t)] = bf16(hw[c >> 1] >> (16 * (c & 1)));
    }
#pragma unroll
    for (int v = 0; v < kIn / 8; v++) reinterpret_cast<uint4 *>(&xs[(p >> 1) + 16 * (p & 1)][0])[v] = h[v];
  }

  __syncthreads();
  // ---- forward: three hidden layers, GELU'(Z) kept in LDS
  float w3[kNB][4];  // W3[n][0] of the wave's neurons n = 16 kNB w + 16 jn + 4 grp + c
#pragma unroll
  for (int jn = 0; jn < kNB; jn++)
#pragma unroll
    for (int c = 0; c < 4; c++) w3[jn][c] = from_bf16(a.w[3][size_t(16 * kNB * wave + 16 * jn + 4 * grp + c) * kOut]);
  float yd[2] = {0.f, 0.f};
#pragma unroll
  for (int l = 0; l < 3; l++) {
    floatx4 acc[kNB][2];
    // products 0 .. 4 (layers 0, 1, 2, backward W2, W1) alternate between
    // the buffers; product p + 2's fragments load as soon as p's are used
    bf16x8 (&fa)[8][kNB] = l & 1 ? fy : fx;
    if (l == 0)
      rows_mma<kIn>(fa, &xs[0][0], kRBIn, r, grp, acc);
    else
      rows_mma<kHidden>(fa, &hs[(l - 1) & 1][0][0], kRBPitch, r, grp, acc);
    if (l == 0) rows_load<kHidden>(a.wt[2], wave, r, grp, fx);
    if (l == 1) rows_load<kHidden>(a.w[2], wave, r, grp, fy);
    if (l == 2) rows_load<kHidden>(a.w[1], wave, r, grp, fx);
    uint32_t v[kNB][2][2];
#pragma unroll
    for (int jn = 0; jn < kNB; jn++) {
      const float4 bb = *reinterpret_cast<const float4 *>(a.bias[l] + 16 * kNB * wave + 16 * jn + 4 * grp);
      const float bv[4] = {bb.x, bb.y, bb.z, bb.w};
#pragma unroll
      for (int i = 0; i < 2; i++) {
        float av[4], gv[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          gelu_both(acc[jn][i][c] + bv[c], av[c], gv[c]);
        }
        v[jn][i][0] = pack_bf16(av[0], av[1]);
        v[jn][i][1] = pack_bf16(av[2], av[3]);
        *reinterpret_cast<uint2 *>(&gs[l][r + 16 * i][16 * kNB * wave + 16 * jn + 4 * grp]) =
            make_uint2(pack_bf16(gv[0], gv[1]), pack_bf16(gv[2], gv[3]));
        if (l == 2)
#pragma unroll
          for (int c = 0; c < 4; c++) {
            const uint32_t u = v[jn][i][c >> 1];
            yd[i] = __builtin_fmaf(c & 1 ? bf16_hi(u) : bf16_lo(u), w3[jn][c], yd[i]);
          }
      }
    }
    rows_store(v, &hs[l & 1][0][0], a.a_t[l], a.ld_t, row0, wave, r, grp);
    __syncthreads();
  }
  // ---- output neuron, loss, dz3 (the four lane groups, then the four waves, in order)
#pragma unroll
  for (int i = 0; i < 2; i++) {
    yd[i] += __shfl_xor(yd[i], 16);
    yd[i] += __shfl_xor(yd[i], 32);
  }
  if (grp == 0) {
    ypart[wave][2 * r] = yd[0];
    ypart[wave][2 * r + 1] = yd[1];
  }
  __syncthreads();
  if (tid < kRB) {
    float y = ypart[0][tid];
#pragma unroll
    for (int w = 1; w < kRBW; w++) y += ypart[w][tid];
    y += a.bias[3][0];
    const uint32_t q = qs[tid];
    const float e = a.energy[q];
    const TargetNorm t = a.table ? a.table[a.lig[q]] : a.norm;
    const float d = y - (fminf(e, t.clip) - t.mu) * t.inv_sigma;
    const float wgt = e < t.tail ? a.tail_weight : 1.0f;
    const uint32_t gq = pack_bf16(2.0f * wgt * d / float(a.n), 0.f) & 0xffffu;
    dz3s[tid] = bf16_lo(gq);
    a.dz3_t[tiled_index(0, row0 + tid, a.ld_t)] = bf16(gq);
    if (a.pred) a.pred[row0 + tid] = y;
    // the squared errors of the 32 rows: one fixed-point atomic per work-group
    float sq = d * d;
    for (int off = 16; off > 0; off >>= 1) sq += __shfl_down(sq, off, 32);
    if (tid == 0 && a.loss)
      atomicAdd(a.loss, static_cast<unsigned long long>(fmin(double(sq) * kLossScale + 0.5, 1099511627776.0)));
  }
  __syncthreads();
  // ---- backward: dZ2 = dz3 W3^T * GELU'(Z2) (rank one), then two products
  {
    const float dq[2] = {dz3s[2 * r], dz3s[2 * r + 1]};
    uint32_t v[kNB][2][2];
#pragma unroll
    for (int jn = 0; jn < kNB; jn++)
#pragma unroll
      for (int i = 0; i < 2; i++) {
        const uint2 gp = *reinterpret_cast<const uint2 *>(&gs[2][r + 16 * i][16 * kNB * wave + 16 * jn + 4 * grp]);
        float dv[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const uint32_t gu = c < 2 ? gp.x : gp.y;
          dv[c] = (w3[jn][c] * dq[i]) * (c & 1 ? bf16_hi(gu) : bf16_lo(gu));
        }
        v[jn][i][0] = pack_bf16(dv[0], dv[1]);
        v[jn][i][1] = pack_bf16(dv[2], dv[3]);
      }
    rows_store(v, &hs[1][0][0], a.dz_t[2], a.ld_t, row0, wave, r, grp);
  }
  __syncthreads();
#pragma unroll
  for (int l = 2; l >= 1; l--) {
    // dZ_{l-1} = (dZ_l W_l^T) * GELU'(Z_{l-1}): A = W_l rows (input neurons), B = dZ_l from LDS
    floatx4 acc[kNB][2];
    rows_mma<kHidden>(l == 2 ? fy : fx, &hs[(l + 1) & 1][0][0], kRBPitch, r, grp, acc);  // dZ2: hs[1], dZ1: hs[0]
    uint32_t v[kNB][2][2];
#pragma unroll
    for (int jn = 0; jn < kNB; jn++)
#pragma unroll
      for (int i = 0; i < 2; i++) {
        const uint2 gp = *reinterpret_cast<const uint2 *>(&gs[l - 1][r + 16 * i][16 * kNB * wave + 16 * jn + 4 * grp]);
        float dv[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const uint32_t gu = c < 2 ? gp.x : gp.y;
          dv[c] = acc[jn][i][c] * (c & 1 ? bf16_hi(gu) : bf16_lo(gu));
        }
        v[jn][i][0] = pack_bf16(dv[0], dv[1]);
        v[jn][i][1] = pack_bf16(dv[2], dv[3]);
      }
    rows_store(v, l == 2 ? &hs[0][0][0] : nullptr, a.dz_t[l - 1], a.ld_t, row0, wave, r, grp);
    if (l == 2) __syncthreads();
  }
}

// ---- the weight gradients of a step, one launch ------------------------------
// One wave per (product, 64 x 64 tile, K slice): BM x BN 16 x 16 MFMA blocks
// (2 at the edge of a 32-row or 32-column product) over a slice of k / splits
// batch rows.  The launch has fewer waves than the device has SIMDs, so a
// wave's own instruction stream is the kernel's time: the tile shape is a
// template parameter (no branch around an MFMA, no accumulator copies), the
// fragments come by buffer loads at per-lane row offsets computed once plus
// a scalar step offset (no 64-bit address arithmetic per load), four k steps
// of fragments load while the four before them multiply (the rows were just
// written by the rows kernel: their latency is long; the re-reads of a slice
// by the tiles beside it hit L2, since a slice's waves all run on one XCD:
// item mod 8 = split mod 8), and the waves of the first tile row also sum
// Q's rows, the bias gradient, as an MFMA against rows of ones.
struct DwItem {
  int d, m0, n0, split;
};

__device__ __forceinline__ DwItem dw_item(const DwArgs &a, int item) {
  DwItem it{0, 0, 0, 0};
  for (int d = 0; d < a.count; d++) {
    const int tn = (a.d[d].N + 63) / 64, tm = (a.d[d].M + 63) / 64, n = tn * tm * a.splits;
    if (item < n) {
      it.d = d;
      it.split = item % a.splits;
      const int tile = item / a.splits;
      it.n0 = 64 * (tile % tn);
      it.m0 = 64 * (tile / tn);
      return it;
    }
    item -= n;
  }
  it.d = -1;
  return it;
}

typedef int intx4 __attribute__((ext_vector_type(4)));
constexpr int kRawBufferWord3 = 0x00020000;  // raw (unformatted) buffer access, 32-bit elements

__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_buffer(const void *base, size_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, int(bytes), kRawBufferWord3);
}

template <int BM, int BN, bool BSUM>
__device__ __forceinline__ void dw_tile(const DwArgs &a, const DwProduct &d, const DwItem &it) {
  constexpr int kG = 4;  // k steps per group of fragments
  const int lane = threadIdx.x, r = lane & 15, grp = lane >> 4;
  const int kc = int(a.k / size_t(a.splits)), nsteps = kc / 32;
  const uint32_t k0 = uint32_t(it.split) * uint32_t(kc);
  const __amdgpu_buffer_rsrc_t rp = raw_buffer(d.p, size_t(d.M) * a.ld * 2);
  const __amdgpu_buffer_rsrc_t rq = raw_buffer(d.q, size_t(d.N) * a.ld * 2);
  // byte offsets of the lane's 16 bytes of the fragments at the slice's first
  // k step (tiled operands: a fragment is 1 KB, the next k step the next KB)
  const uint32_t kb = uint32_t(a.ld / 32), kb0 = k0 / 32u;
  uint32_t vp[BM], vq[BN];
#pragma unroll
  for (int i = 0; i < BM; i++) vp[i] = 1024u * (uint32_t(it.m0 / 16 + i) * kb + kb0) + 16u * uint32_t(lane);
#pragma unroll
  for (int j = 0; j < BN; j++) vq[j] = 1024u * (uint32_t(it.n0 / 16 + j) * kb + kb0) + 16u * uint32_t(lane);
  floatx4 acc[BM][BN], accb[BN];
#pragma unroll
  for (int i = 0; i < BM; i++)
#pragma unroll
    for (int j = 0; j < BN; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j = 0; j < BN; j++) accb[j] = floatx4{0.f, 0.f, 0.f, 0.f};
  bf16x8 ones;
#pragma unroll
  for (int e = 0; e < 8; e++) ones[e] = __builtin_bit_cast(__bf16, uint16_t(0x3f80));
  // step s of the slice: byte offset 1024 s (uniform)
  auto load_step = [&](int s, bf16x8 (&ap)[BM], bf16x8 (&aq)[BN]) {
    const int so = 1024 * s;
#pragma unroll
    for (int i = 0; i < BM; i++) ap[i] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rp, vp[i], so, 0));
#pragma unroll
    for (int j = 0; j < BN; j++) aq[j] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rq, vq[j], so, 0));
  };
  auto mul_step = [&](const bf16x8 (&ap)[BM], const bf16x8 (&aq)[BN]) {
#pragma unroll
    for (int i = 0; i < BM; i++)
#pragma unroll
      for (int j = 0; j < BN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ap[i], aq[j], acc[i][j], 0, 0, 0);
    if constexpr (BSUM) {
#pragma unroll
      for (int j = 0; j < BN; j++) accb[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ones, aq[j], accb[j], 0, 0, 0);
    }
  };
  bf16x8 fp[2][kG][BM], fq[2][kG][BN];
  const int groups = nsteps / kG;  // whole groups; the remaining steps one at a time after them
  if (groups > 0) {
#pragma unroll
    for (int t = 0; t < kG; t++) load_step(t, fp[0][t], fq[0][t]);
    // (the scheduling barriers keep each group's loads ahead of the
    // multiplies before them: left alone, the compiler moves every load next
    // to its first use)
    for (int g = 0; g < groups; g += 2) {
      const int g1 = min(g + 1, groups - 1), g2 = min(g + 2, groups - 1);  // (reloads of the last group: unused)
#pragma unroll
      for (int t = 0; t < kG; t++) load_step(kG * g1 + t, fp[1][t], fq[1][t]);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < kG; t++) mul_step(fp[0][t], fq[0][t]);
      __builtin_amdgcn_sched_barrier(0);
      if (g + 1 < groups) {
#pragma unroll
        for (int t = 0; t < kG; t++) load_step(kG * g2 + t, fp[0][t], fq[0][t]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < kG; t++) mul_step(fp[1][t], fq[1][t]);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  for (int s = kG * groups; s < nsteps; s++) {
    load_step(s, fp[0][0], fq[0][0]);
    mul_step(fp[0][0], fq[0][0]);
  }
  // D to the slice: element (row m0 + 16 i + 4 grp + c, column n0 + 16 j + r)
  float *out = d.work + size_t(it.split) * size_t(d.M) * size_t(d.N) + size_t(it.m0 + 4 * grp) * size_t(d.N) +
               size_t(it.n0 + r);
#pragma unroll
  for (int i = 0; i < BM; i++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      float *o = out + size_t(16 * i + c) * size_t(d.N);
#pragma unroll
      for (int j = 0; j < BN; j++) o[16 * j] = acc[i][j][c];
    }
  if constexpr (BSUM) {
    if (grp == 0)
#pragma unroll
      for (int j = 0; j < BN; j++) d.bias_part[size_t(it.split) * d.N + it.n0 + 16 * j + r] = accb[j][0];
  }
}

__global__ __launch_bounds__(64) void dw_kernel(const DwArgs a) {
  const DwItem it = dw_item(a, blockIdx.x);
  if (it.d < 0) return;
  const DwProduct &d = a.d[it.d];
  const bool wide_m = d.M - it.m0 >= 64, wide_n = d.N - it.n0 >= 64;
  const bool bsum = d.bias_part && it.m0 == 0;
  if (wide_m && wide_n) {
    if (bsum) dw_tile<4, 4, true>(a, d, it);
    else dw_tile<4, 4, false>(a, d, it);
  } else if (wide_m) {
    if (bsum) dw_tile<4, 2, true>(a, d, it);
    else dw_tile<4, 2, false>(a, d, it);
  } else if (wide_n) {
    if (bsum) dw_tile<2, 4, true>(a, d, it);
    else dw_tile<2, 4, false>(a, d, it);
  } else {
    if (bsum) dw_tile<2, 2, true>(a, d, it);
    else dw_tile<2, 2, false>(a, d, it);
  }
}

__global__ void ksplit_reduce_kernel(const float *__restrict__ work, size_t n, int splits, float *__restrict__ out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float acc = 0.f;
  for (int z = 0; z < splits; z++) acc += work[size_t(z) * n + i];
  out[i] = acc;
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
  if (int(blockIdx.x) >= b.tensor_blocks) {  // the fixed-point tensor: 64 W entries per work-group
    const AdamFixed &F = b.fixed;
    const size_t k = size_t(blockIdx.x - b.tensor_blocks) * (64 * W) + threadIdx.x;
    if (k >= F.n) return;
    const float gk = float(double(static_cast<long long>(F.acc[k])) / F.scale);
    F.acc[k] = 0ull;
    const float mk = b1 * F.m[k] + (1.f - b1) * gk;
    const float vk = b2 * F.v[k] + (1.f - b2) * gk * gk;
    F.m[k] = mk;
    F.v[k] = vk;
    F.w[k] = F.w[k] - hyper[1] * (mk / bc1) / (sqrtf(vk / bc2) + eps);
    return;
  }
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
    // slices wave, wave + W, ... added in that order, eight (or four) loads in flight
    int z = wave;
    for (; z + 7 * kAdamWaves < splits; z += 8 * kAdamWaves) {
      float sl[8];
#pragma unroll
      for (int u = 0; u < 8; u++) sl[u] = g[size_t(z + u * kAdamWaves) * stride];
#pragma unroll
      for (int u = 0; u < 8; u++) gs += sl[u];
    }
    for (; z + 3 * kAdamWaves < splits; z += 4 * kAdamWaves) {
      float sl[4];
#pragma unroll
      for (int u = 0; u < 4; u++) sl[u] = g[size_t(z + u * kAdamWaves) * stride];
#pragma unroll
      for (int u = 0; u < 4; u++) gs += sl[u];
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
  if (T.w16x) T.w16x[tiled_index(i / T.cols, i % T.cols, size_t(T.cols))] = h;
  if (T.w16tx) T.w16tx[tiled_index(i % T.cols, i / T.cols, size_t(T.rows))] = h;
}

__global__ void tile_bf16_kernel(const float *__restrict__ w, int rows, int cols, bf16 *__restrict__ wx,
                                 bf16 *__restrict__ wtx) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(rows) * cols) return;
  const bf16 h = to_bf16(w[i]);
  if (wx) wx[tiled_index(i / cols, i % cols, size_t(cols))] = h;
  if (wtx) wtx[tiled_index(i % cols, i / cols, size_t(rows))] = h;
}

__global__ void to_bf16_kernel(const float *__restrict__ w, int rows, int cols, bf16 *__restrict__ w16,
                               bf16 *__restrict__ w16t, size_t ldt, int tiled) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(rows) * cols) return;
  const bf16 h = to_bf16(w[i]);
  if (w16) w16[i] = h;
  if (w16t) w16t[tiled ? tiled_index(i % cols, i / cols, ldt) : size_t(i % cols) * ldt + i / cols] = h;
}

// ---- fused inference ---------------------------------------------------------
// One work-group (8 waves) scores tiles of kTile poses through the whole
// network with the activations in one LDS buffer (pose-major rows):
// features -> 3 x (MFMA GEMM + bias + GELU) -> output layer, also on MFMA.
// A hidden layer computes its transpose, D[neuron][pose] = W . act^T, so the
// weights are the A operand (streamed from L2 as fragment tiles of W^T,
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
  // W^T fragment-tiled: neuron block 2 w + j, k step k / 32 is 1 KB at
  // ((2 w + j) K / 32 + k / 32) KB, this lane's 16 bytes at 16 lane
  const bf16 *wl = wt + 8 * lane;
  bf16x8 fw[2];
#pragma unroll
  for (int j = 0; j < 2; j++) fw[j] = load8(wl + size_t((2 * wave + j) * (K / 32)) * 512);
#pragma unroll 1
  for (int k = 0; k < K; k += 32) {
    bf16x8 fn[2];
    const int kn = k + 32 < K ? k + 32 : k;  // (the last step reloads its own fragments)
#pragma unroll
    for (int j = 0; j < 2; j++) fn[j] = load8(wl + size_t((2 * wave + j) * (K / 32) + kn / 32) * 512);
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

__global__ __launch_bounds__(256) void fold_bias_kernel(const float *__restrict__ b, const bf16 *__restrict__ w0tx,
                                                        const float *__restrict__ desc, float *__restrict__ out) {
  const int l = blockIdx.x, j = threadIdx.x;
  float acc = b[j];
#pragma unroll
  for (int c = 0; c < kDesc; c++)
    acc = __builtin_fmaf(from_bf16(to_bf16(desc[size_t(l) * kDesc + c])),
                         from_bf16(w0tx[tiled_index(size_t(j), size_t(kFeat + c), kIn)]), acc);
  out[size_t(l) * kHidden + j] = acc;
}

unsigned blocks_for(size_t n, unsigned t) { return unsigned((n + t - 1) / t); }

}  // namespace

hipError_t rows(const RowsArgs &a, hipStream_t s) {
  if (a.n == 0 || a.n % kRB || a.ld_t < a.n || a.ld_t % 32 || (!a.feat && (!a.poses || !a.feat_t)) ||
      (a.perm && (!a.state || a.per_epoch == 0)) || !a.energy || !a.dz3_t)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(rows_kernel, dim3(unsigned(a.n / kRB)), dim3(64 * kRBW), 0, s, a);
  return hipGetLastError();
}

hipError_t weight_grads(const DwArgs &a, hipStream_t s) {
  if (a.count < 1 || a.count > 4 || a.splits < 1 || a.k == 0 || a.k % (32 * size_t(a.splits)) || a.ld < a.k ||
      a.ld % 32)
    return hipErrorInvalidValue;
  int items = 0;
  for (int d = 0; d < a.count; d++) {
    // (buffer offsets and sizes are 32-bit: operands and slices under 2 GiB)
    if (a.d[d].M % 32 || a.d[d].N % 32 || !a.d[d].p || !a.d[d].q || !a.d[d].work ||
        size_t(std::max(a.d[d].M, a.d[d].N)) * a.ld * 2 >= (size_t(1) << 31) ||
        size_t(a.d[d].M) * size_t(a.d[d].N) * 4 >= (size_t(1) << 31))
      return hipErrorInvalidValue;
    items += ((a.d[d].M + 63) / 64) * ((a.d[d].N + 63) / 64) * a.splits;
  }
  hipLaunchKernelGGL(dw_kernel, dim3(unsigned(items)), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t sum_slices(const float *work, int splits, size_t n, float *out, hipStream_t s) {
  if (splits < 1) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(ksplit_reduce_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, work, n, splits, out);
  return hipGetLastError();
}

hipError_t fold_bias(const float *b, const bf16 *w0tx, const float *desc, int nlig, float *out, hipStream_t s) {
  if (nlig <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fold_bias_kernel, dim3(unsigned(nlig)), dim3(kHidden), 0, s, b, w0tx, desc, out);
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
  if (total == 0 || !hyper || (b.fixed.n && (!b.fixed.w || !b.fixed.acc || !b.fixed.m || !b.fixed.v)))
    return hipErrorInvalidValue;
  const int W = splits > 32 ? 16 : 4;
  AdamBatch a = b;
  a.tensor_blocks = int(blocks_for(total, 64));
  const unsigned blocks = unsigned(a.tensor_blocks) + blocks_for(b.fixed.n, 64u * W);
  if (W == 16)
    hipLaunchKernelGGL(adam_multi_kernel<16>, dim3(blocks), dim3(64 * 16), 0, s, a, 0.f, b1, b2, eps, 1.f, 1.f, hyper,
                       advance);
  else
    hipLaunchKernelGGL(adam_multi_kernel<4>, dim3(blocks), dim3(64 * 4), 0, s, a, 0.f, b1, b2, eps, 1.f, 1.f, hyper,
                       advance);
  return hipGetLastError();
}

hipError_t tile_bf16(const float *w, int rows, int cols, bf16 *wx, bf16 *wtx, hipStream_t s) {
  if (rows <= 0 || cols <= 0 || (wx && (rows % 16 || cols % 32)) || (wtx && (cols % 16 || rows % 32)))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(tile_bf16_kernel, dim3(blocks_for(size_t(rows) * cols, 256)), dim3(256), 0, s, w, rows, cols, wx,
                     wtx);
  return hipGetLastError();
}

hipError_t to_bf16(const float *w, int rows, int cols, bf16 *w16, bf16 *w16t, hipStream_t s, size_t ldt,
                   bool tiled) {
  if (ldt == 0) ldt = size_t(rows);
  if (ldt < size_t(rows) || (tiled && (cols % 16 || ldt % 32))) return hipErrorInvalidValue;
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(size_t(rows) * cols, 256)), dim3(256), 0, s, w, rows, cols, w16,
                     w16t, ldt, int(tiled));
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
