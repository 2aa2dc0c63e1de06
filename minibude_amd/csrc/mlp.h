// mlp.h -- the pose-energy surrogate: a 4-layer MLP (21 pose features ->
// 256 -> 256 -> 256 -> 1, GELU) on BF16 MFMA with FP32 accumulation, its
// training kernels (backward GEMMs, loss, Adam), and the gather that passes
// the surrogate's best poses to exact rescoring.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bude {
namespace mlp {

using bf16 = uint16_t;  // storage type of BF16 tensors (round-to-nearest-even from FP32)

constexpr int kFeat = 21;    // tilt, roll, pan, tx, ty, tz, sin x3, cos x3, rotation matrix x9
constexpr int kIn = 32;      // feature width padded to the MFMA K step
constexpr int kHidden = 256;
constexpr int kOut = 32;     // output layer padded to one MFMA tile (column 0 is the energy)
constexpr int kLayers = 4;
constexpr int kDesc = kIn - kFeat;  // ligand descriptor inputs (zero when one ligand is screened)

// Target transform of one ligand's energies: y = (min(E, clip) - mu) * inv_sigma;
// energies below `tail` get the tail loss weight.
struct TargetNorm {
  float clip, mu, inv_sigma, tail;
};

// GEMM epilogues.
enum Epilogue : int {
  kEpiHidden = 0,  // z = acc + b[n]; a = gelu(z); store z (f32, optional), a (bf16), a^T (bf16, optional)
  kEpiOut = 1,     // pred[m] = acc[m][0] + b[0]
  kEpiBackGelu = 2,  // dz = acc * gelu'(zprev[m][n]); store dz (bf16) and dz^T (bf16)
  kEpiF32 = 3,     // store acc (f32), e.g. a weight gradient
  kEpiHiddenRelu = 4,  // as kEpiHidden with a = max(z, 0)
  kEpiBackRelu = 5,    // as kEpiBackGelu with the ReLU derivative (zprev > 0)
};

struct GemmArgs {
  const bf16 *a;  // [M][lda], rows k-contiguous
  const bf16 *b;  // [N][ldb], rows k-contiguous: C = A * B^T
  int M, N, K, lda, ldb;
  int epi;
  const float *bias;   // [N]            (kEpiHidden, kEpiOut)
  float *out_f32;      // [M][ldo]       (kEpiHidden z, kEpiF32 acc, kEpiOut pred[M])
  bf16 *out_bf16;      // [M][ldo]       (kEpiHidden a, kEpiBackGelu dz)
  bf16 *out_t;         // [N][ldt]       (transposed copy, optional)
  const float *zprev;  // [M][ldo]       (kEpiBackGelu)
  int ldo, ldt;
  int ksplit;          // kEpiF32 only: > 1 splits K over ksplit work-groups per tile
  float *work;         // [ksplit][M][ldo] partial sums when ksplit > 1
};

// C = A * B^T on 32x32 tiles (M, N, K multiples of 32).  With ksplit > 1
// (kEpiF32) every tile's K range is split over ksplit work-groups writing
// partial sums to `work`, then summed in a fixed order (deterministic).
hipError_t gemm(const GemmArgs &g, hipStream_t s);

// Split count that gives a kEpiF32 GEMM about 1,024 waves (K split in
// multiples of 256).
int auto_ksplit(int M, int N, int K);

// Features of poses idx[i] (or i when idx is null) from 6 pose rows of
// `stride` floats: row i of out [n][kIn] (bf16), and out_t [kIn][ldt] if set.
// With desc [nlig][kDesc] and lig (ligand of each pose row), columns kFeat..
// hold desc[lig[q]]; otherwise they are zero.  Feature columns from nfeat
// up are zero (nfeat = 12: the six pose parameters and the sines and
// cosines of the angles only; a network whose first-layer rows for the
// other inputs start at zero then keeps them at zero).
hipError_t features(const float *poses, size_t stride, const uint32_t *idx, size_t n, bf16 *out, bf16 *out_t,
                    int ldt, hipStream_t s, const float *desc = nullptr, const uint32_t *lig = nullptr,
                    int nfeat = kFeat);

// Weighted MSE on the transformed target of pose q = idx[b] (or b):
// y = (min(E, clip) - mu) * inv_sigma with the TargetNorm table[lig[q]]
// (device memory) when table is set, else `norm`; weight tail_weight when
// E < tail else 1.
// dz[b][0] = 2 w (pred - y) / n, other columns 0 (also transposed);
// loss_acc += sum (pred - y)^2 (f32 atomics, for monitoring).
hipError_t loss_grad(const float *pred, const float *energy, const uint32_t *idx, size_t n, TargetNorm norm,
                     const TargetNorm *table, const uint32_t *lig, float tail_weight, bf16 *dz, bf16 *dz_t, int ldt,
                     float *loss_acc, hipStream_t s);

// Per-ligand first-layer bias with the descriptor inputs folded in, for the
// fused inference kernel (whose descriptor columns are zero):
// out[l][j] = b[j] + sum_c bf16(desc[l][c]) * w16[kFeat + c][j], w16 = W_0 [kIn][kHidden].
hipError_t fold_bias(const float *b, const bf16 *w16, const float *desc, int nlig, float *out, hipStream_t s);

// db[n] = sum_b dz_t[n][b] for n < N rows of length B.
hipError_t bias_grad(const bf16 *dz_t, int N, int B, int ldt, float *db, hipStream_t s);

// Adam step on n parameters; refreshes the BF16 copy w16 [rows][cols] and
// its transpose w16t [cols][rows] (either may be null).
hipError_t adam(float *w, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                float bc1, float bc2, bf16 *w16, bf16 *w16t, int rows, int cols, hipStream_t s);

// One Adam step over up to 8 parameter tensors in a single launch.
struct AdamTensor {
  float *w;
  const float *g;
  float *m, *v;
  bf16 *w16, *w16t;  // optional BF16 copies (row-major and transposed)
  int rows, cols;    // w is rows x cols row-major
};
struct AdamBatch {
  AdamTensor t[8];
  int count;
};
hipError_t adam_multi(const AdamBatch &b, float lr, float b1, float b2, float eps, float bc1, float bc2,
                      hipStream_t s);

// Bias gradients of up to 4 layers in one launch: db_l[n] = sum_b dz_t_l[n][b].
struct BiasGradBatch {
  const bf16 *dz_t[4];
  float *db[4];
  int N[4];
  int count, B, ldt;
};
hipError_t bias_grad_multi(const BiasGradBatch &b, hipStream_t s);

// FP32 -> BF16 of a rows x cols matrix and its transpose.
hipError_t to_bf16(const float *w, int rows, int cols, bf16 *w16, bf16 *w16t, hipStream_t s);

// Fused inference: predictions (standardised units) of n poses through the
// whole network, activations kept in LDS; wt[l] = W_l^T [N_l][K_l] (BF16),
// w3 = W_3 [256][kOut] (column 0 used); `blocks` persistent work-groups.
hipError_t infer(const float *poses, size_t stride, size_t n, const bf16 *const wt[4], const bf16 *w3,
                 const float *const bias[4], float *out, int blocks, hipStream_t s);


}  // namespace mlp
}  // namespace bude
