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
  int defer;           // kEpiF32: leave the K slices in `work` (slice 0 of ksplit = 1 too) for
                       // the consumer to sum (Adam over AdamTensor slices), no reduction pass
  float *bias_part;    // kEpiF32, optional: [ksplit][N] row sums of B over each K slice (the
                       // bias gradient when B is dZ^T)
};

// C = A * B^T on 32x32 tiles (M, N, K multiples of 32), each wave loading
// the next K step's fragments while it multiplies the current ones.  With
// ksplit > 1 (kEpiF32) every tile's K range is split over ksplit
// work-groups writing partial sums to `work`, then summed in a fixed order
// (deterministic) by a second pass, or by the consumer (`defer`).
hipError_t gemm(const GemmArgs &g, hipStream_t s);

// out[i] = sum over z < splits of work[z * n + i], z ascending.
hipError_t sum_slices(const float *work, int splits, size_t n, float *out, hipStream_t s);

// Split count that gives a kEpiF32 GEMM about 1,024 waves (K split in
// multiples of 256).
int auto_ksplit(int M, int N, int K);
// Largest split auto_ksplit returns for an M x N product (any K).
int max_ksplit(int M, int N);

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
// loss_acc += sum (pred - y)^2 in fixed point (kLossScale units per 1,
// integer atomics: the same sum whatever order the waves land in; for
// monitoring).
constexpr double kLossScale = 16777216.0;
hipError_t loss_grad(const float *pred, const float *energy, const uint32_t *idx, size_t n, TargetNorm norm,
                     const TargetNorm *table, const uint32_t *lig, float tail_weight, bf16 *dz, bf16 *dz_t, int ldt,
                     unsigned long long *loss_acc, hipStream_t s);

// Per-ligand first-layer bias with the descriptor inputs folded in, for the
// fused inference kernel (whose descriptor columns are zero):
// out[l][j] = b[j] + sum_c bf16(desc[l][c]) * w16[kFeat + c][j], w16 = W_0 [kIn][kHidden].
hipError_t fold_bias(const float *b, const bf16 *w16, const float *desc, int nlig, float *out, hipStream_t s);

// Device-side state of a training loop, so that one captured hipGraph of a
// step replays as every step: the step count, and this step's two learning
// rates (weights, and a second parameter group) and Adam bias corrections.
struct StepState {
  uint32_t step, pad;
  float hyper[4];  // lr0, lr1, 1 - b1^t, 1 - b2^t
};
inline const float *step_hyper(const StepState *st) {
  return reinterpret_cast<const float *>(reinterpret_cast<const char *>(st) + 8);
}

// The learning-rate schedule of a training run and Adam's bias corrections:
// linear warm-up over `warm` steps, then cosine decay to 5 % at the last of
// `steps`; step0 counts the steps of earlier runs on the same network.
struct Schedule {
  uint32_t steps, warm, step0;
  float lr0, lr1, b1, b2;
};
// hyper (lr0, lr1, 1 - b1^t, 1 - b2^t) of step s of the run.
__device__ inline void schedule_hyper(const Schedule &c, uint32_t s, float *hyper) {
  const float warmf = s < c.warm ? float(s + 1) / float(c.warm) : 1.f;
  const float prog = float(s) / float(c.steps > 1 ? c.steps - 1 : 1);
  const float f = warmf * (0.05f + 0.95f * 0.5f * (1.f + cosf(3.14159265f * prog)));
  const float t = float(c.step0 + s + 1);
  hyper[0] = c.lr0 * f;
  hyper[1] = c.lr1 * f;
  hyper[2] = 1.f - powf(c.b1, t);
  hyper[3] = 1.f - powf(c.b2, t);
}

// One Adam step over up to 8 parameter tensors in a single launch.
struct AdamTensor {
  float *w;
  const float *g;    // gradient: `splits` slices `stride` floats apart, summed in a fixed order
  float *m, *v;
  bf16 *w16, *w16t;  // optional BF16 copies (row-major and transposed)
  int rows, cols;    // w is rows x cols row-major
  int splits;        // 0 or 1: g is the gradient itself
  size_t stride;
};
struct AdamBatch {
  AdamTensor t[8];
  int count;
};
// The learning rate and the bias corrections are read from device memory
// when the kernel runs: hyper[0] = lr, hyper[2] = bc1, hyper[3] = bc2 (so a
// captured graph of a training step can be replayed step after step).  With
// `advance` set the launch also ends the step: advance->step += 1.
hipError_t adam_multi_dev(const AdamBatch &b, const float *hyper, float b1, float b2, float eps, hipStream_t s,
                          StepState *advance = nullptr);

// Start of step s = st->step: idx[i] = perm[(s % per_epoch) * B + i] for
// i < B (the batch), and the step's rates: linear warm-up over `warm` steps,
// then cosine decay to 5 % of lr0 / lr1 at the last of `steps`; t = step0 +
// s + 1 for the bias corrections.
hipError_t begin_step(StepState *st, const uint32_t *perm, uint32_t per_epoch, uint32_t B, uint32_t *idx,
                      uint32_t steps, uint32_t warm, float lr0, float lr1, float b1, float b2, uint32_t step0,
                      hipStream_t s);

// FP32 -> BF16 of a rows x cols matrix and its transpose.
hipError_t to_bf16(const float *w, int rows, int cols, bf16 *w16, bf16 *w16t, hipStream_t s);

// Fused inference: predictions (standardised units) of n poses through the
// whole network, activations kept in LDS; wt[l] = W_l^T [N_l][K_l] (BF16),
// w3 = W_3 [256][kOut] (column 0 used); `blocks` persistent work-groups.
hipError_t infer(const float *poses, size_t stride, size_t n, const bf16 *const wt[4], const bf16 *w3,
                 const float *const bias[4], float *out, int blocks, hipStream_t s);


}  // namespace mlp
}  // namespace bude
