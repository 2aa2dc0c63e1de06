// mlp.h -- the pose-energy surrogate: a 4-layer MLP (21 pose features ->
// 256 -> 256 -> 256 -> 1, GELU) on BF16 MFMA with FP32 accumulation, its
// training kernels (the fused rows kernel, the weight-gradient products,
// Adam) and its fused inference kernel.
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

// out[i] = sum over z < splits of work[z * n + i], z ascending.
hipError_t sum_slices(const float *work, int splits, size_t n, float *out, hipStream_t s);

// Squared errors are summed in fixed point, kLossScale units per 1 (integer
// atomics: the same sum whatever order the waves land in; for monitoring).
constexpr double kLossScale = 16777216.0;

// Per-ligand first-layer bias with the descriptor inputs folded in, for the
// fused inference kernel (whose descriptor columns are zero):
// out[l][j] = b[j] + sum_c bf16(desc[l][c]) * W_0[kFeat + c][j], read from
// w0tx = W_0^T [kHidden][kIn] fragment-tiled (tiled_index).
hipError_t fold_bias(const float *b, const bf16 *w0tx, const float *desc, int nlig, float *out, hipStream_t s);

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

// The fused rows kernel of a training step (forward, loss, backward to dZ0
// for kRowBlock rows per work-group; see mlp.hip).  Input rows: features
// computed from poses (pose_features, the first nfeat of them, desc[lig[q]]
// in the last kDesc columns if desc is set; X^T written to feat_t), or the
// BF16 feature rows feat [n][kIn].  Row r of the launch is sample q =
// perm[(state->step mod per_epoch) n + r] with perm set (the launch then
// also writes the step's schedule into state->hyper), else idx[r], else r.
// Target of sample q: y = (min(E, clip) - mu) * inv_sigma with table[lig[q]]
// (device memory) when table is set, else norm; dz3 = bf16(2 w (pred - y) /
// n), w = tail_weight when E < tail, else 1.  Outputs: pred (optional), the
// loss sum (kLossScale fixed point),
// and the transposed BF16 rows A_l^T, dZ_l^T [kHidden][ld_t] and dz3^T (row
// 0 of [kOut][ld_t]), fragment-tiled (tiled_index), for the weight-gradient
// products.
constexpr int kRowBlock = 32;

// The transposed rows of a batch (X^T, A_l^T, dZ_l^T, dz3^T: M x K, K the
// batch) are stored as MFMA fragments: the 16 x 32 block of rows 16 mb ..
// 16 mb + 15 and columns 32 kb .. 32 kb + 31 is 1 KB contiguous at block
// mb * (K / 32) + kb, in the order a 16x16x32 operand takes it (lane
// r + 16 g: row r, columns 8 g .. 8 g + 7).  A fragment is then one
// contiguous 1 KB load (8 whole lines) instead of 64 bytes of 16 rows (16
// lines, half used each), which the weight-gradient product's loads wait
// on far less.
__host__ __device__ inline size_t tiled_index(size_t m, size_t k, size_t K) {
  return ((m >> 4) * (K >> 5) + (k >> 5)) * 512 + ((m & 15) + 16 * ((k & 31) >> 3)) * 8 + (k & 7);
}

// Columns of the tiled transposed rows of a batch of n: n rounded up to the
// 32 of a block.
inline size_t t_ld(size_t n) { return (n + 31) / 32 * 32; }
struct RowsArgs {
  const float *poses;
  size_t stride;
  const float *desc;
  const uint32_t *lig;
  int nfeat;
  const bf16 *feat;
  const uint32_t *perm;
  uint32_t per_epoch;
  StepState *state;
  Schedule sched;
  const uint32_t *idx;
  size_t n;     // rows, a multiple of kRowBlock
  size_t ld_t;  // columns of the tiled transposed rows (t_ld(n): a multiple of 32, >= n)
  const float *energy;
  TargetNorm norm;
  const TargetNorm *table;
  float tail_weight;
  const bf16 *wt[3];   // W_l^T [N_l][K_l], l < 3, fragment-tiled (tiled_index(n, k, K_l))
  const bf16 *w[4];    // W_l [K_l][N_l]: w[1], w[2] fragment-tiled (tiled_index(k, n, N_l)); w[3] row-major,
                       // column 0 read; w[0] unused
  const float *bias[4];
  float *pred;
  unsigned long long *loss;
  bf16 *feat_t;        // [kIn][ld_t], tiled
  bf16 *a_t[3], *dz_t[3];
  bf16 *dz3_t;
};
hipError_t rows(const RowsArgs &a, hipStream_t s);

// The weight gradients of a training step in one launch: for each product
// d, D[M][N] = P[M][0:k] Q[N][0:k]^T (P = A_{l-1}^T or X^T, Q = dZ_l^T, the
// transposed rows of the batch, tiled with ld columns: tiled_index) as
// `splits` K slices in work [splits][M][N], and the row sums of Q per slice
// in bias_part [splits][N] (the bias gradient).  M, N and k multiples of 32,
// ld a multiple of 32 and >= k; `splits` = dw_splits(k).
constexpr int kDwSplits = 16;
// The K slices of a batch of k rows: the largest power of two up to
// kDwSplits that divides k into multiples of 32 rows.
inline int dw_splits(size_t k) {
  int s = kDwSplits;
  while (s > 1 && k % (32 * size_t(s))) s /= 2;
  return s;
}
struct DwProduct {
  const bf16 *p, *q;
  int M, N;
  float *work, *bias_part;
};
struct DwArgs {
  DwProduct d[4];
  int count;
  size_t k;   // batch rows summed over
  size_t ld;  // columns of the tiled P and Q
  int splits;
};
hipError_t weight_grads(const DwArgs &a, hipStream_t s);

// One Adam step over up to 8 parameter tensors in a single launch.
struct AdamTensor {
  float *w;
  const float *g;    // gradient: `splits` slices `stride` floats apart, summed in a fixed order
  float *m, *v;
  bf16 *w16, *w16t;  // optional BF16 copies (row-major and transposed)
  int rows, cols;    // w is rows x cols row-major
  int splits;        // 0 or 1: g is the gradient itself
  size_t stride;
  bf16 *w16x = nullptr, *w16tx = nullptr;  // optional fragment-tiled BF16 copies of w and w^T (tiled_index)
};
// A tensor whose gradient arrives as a fixed-point sum of integer atomics:
// g = acc / scale, and the Adam step clears acc.  Its learning rate is
// hyper[1] (adam_multi_dev with hyper set).
struct AdamFixed {
  float *w;
  unsigned long long *acc;
  float *m, *v;
  size_t n;  // 0: none
  double scale;
};
struct AdamBatch {
  AdamTensor t[8];
  int count;
  AdamFixed fixed{};  // optional: its work-groups follow the tensors' in the same launch
  int tensor_blocks = 0;  // (set by adam_multi_dev)
};
// The learning rate and the bias corrections are read from device memory
// when the kernel runs: hyper[0] = lr, hyper[2] = bc1, hyper[3] = bc2 (so a
// captured graph of a training step can be replayed step after step).  With
// `advance` set the launch also ends the step: advance->step += 1.
hipError_t adam_multi_dev(const AdamBatch &b, const float *hyper, float b1, float b2, float eps, hipStream_t s,
                          StepState *advance = nullptr);

// FP32 -> BF16 fragment-tiled copies (tiled_index) of a rows x cols matrix
// (wx, rows a multiple of 16, cols of 32) and of its transpose (wtx, cols a
// multiple of 16, rows of 32); either may be null.
hipError_t tile_bf16(const float *w, int rows, int cols, bf16 *wx, bf16 *wtx, hipStream_t s);

// FP32 -> BF16 of a rows x cols matrix and its transpose (cols x ldt, ldt 0:
// rows; tiled: fragment-tiled, tiled_index, cols a multiple of 16 and ldt of
// 32).
hipError_t to_bf16(const float *w, int rows, int cols, bf16 *w16, bf16 *w16t, hipStream_t s, size_t ldt = 0,
                   bool tiled = false);

// Fused inference: predictions (standardised units) of n poses through the
// whole network, activations kept in LDS; wt[l] = W_l^T [N_l][K_l] (BF16,
// fragment-tiled: tiled_index(n, k, K_l); l < 3), w3 = W_3 [256][kOut]
// row-major (column 0 used); `blocks` persistent work-groups.
hipError_t infer(const float *poses, size_t stride, size_t n, const bf16 *const wt[4], const bf16 *w3,
                 const float *const bias[4], float *out, int blocks, hipStream_t s);


}  // namespace mlp
}  // namespace bude
