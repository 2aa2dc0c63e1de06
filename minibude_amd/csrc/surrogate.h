// surrogate.h -- training state of the MLP pose-energy surrogate (weights,
// Adam moments, activations) and the training loop shared by the learned
// prescreen of one ligand and the per-protein surrogate of a combined screen.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "hip_util.h"
#include "mlp.h"

namespace bude {
namespace surrogate {

using mlp::bf16;
constexpr int kK[4] = {mlp::kIn, mlp::kHidden, mlp::kHidden, mlp::kHidden};
constexpr int kN[4] = {mlp::kHidden, mlp::kHidden, mlp::kHidden, mlp::kOut};

struct Net {
  DeviceBuffer w[4], b[4], mw[4], vw[4], mb[4], vb[4], gw[4], gb[4];
  // BF16 copies the kernels read, refreshed by every Adam step:
  // fragment-tiled W_l^T (l < 3: the forward products of training and
  // inference, the descriptor fold) and W_l (l = 1, 2: the backward
  // products), and W_3 row-major in w16[3] (the output neuron's column)
  DeviceBuffer w16[4], w16x[4], w16tx[4];
  int step = 0;
};

// The rows of one batch of `rows` poses: BF16 features, and in training the
// transposed rows the weight-gradient products read (X^T, A_l^T, dZ_l^T,
// dz3^T: fragment-tiled with ld_t = mlp::t_ld(rows) columns), predictions and
// the loss sum.
struct Acts {
  size_t rows = 0, ld_t = 0;
  DeviceBuffer feat, feat_t, a_t[3], pred, dz3_t, dz_t[3], loss;
  // weight and bias gradients of layer l as splits[l] K slices (summed in a
  // fixed order by the Adam step, in slice order by sum_grads)
  DeviceBuffer gpart[4], bpart[4];
  int splits[4] = {1, 1, 1, 1};
};

int alloc_acts(Acts &A, size_t rows, bool training);
// The fused rows kernel over `rows` rows (mlp::rows) with the network's
// weights and A's buffers filled in from in's inputs and targets.
int rows_step(const Net &net, Acts &A, size_t rows, const mlp::RowsArgs &in, hipStream_t s);
// From the transposed rows: the weight and bias gradients as K slices in
// A.gpart / bpart.
int weight_grads(Net &net, Acts &A, size_t rows, hipStream_t s);
// The slices summed in order into net.gw / gb.
int sum_grads(Net &net, const Acts &A, hipStream_t s);
// Uploads FP32 weights W_l [kK[l]][kN[l]] (zero padded) and biases, clears
// the Adam state and refreshes the BF16 copies.
int set_weights(Net &net, const std::vector<std::vector<float>> &w, const std::vector<std::vector<float>> &b,
                hipStream_t s);
// He-normal hidden weights, small output weights, zero biases; `inputs`
// real input rows (the rest of the padded input stays zero) and one output.
void init_weights(uint64_t seed, int inputs, std::vector<std::vector<float>> &w, std::vector<std::vector<float>> &b);

// What the training loop reads, all device memory: label poses (6 rows of
// `stride` floats), their exact energies, and either one TargetNorm for all
// labels (table null) or a per-ligand table with the ligand of each label
// and the per-ligand descriptor inputs.
struct TrainData {
  const float *poses = nullptr;
  size_t stride = 0;
  const float *energy = nullptr;
  mlp::TargetNorm norm{};
  const mlp::TargetNorm *table = nullptr;
  const uint32_t *lig = nullptr;
  const float *desc = nullptr;
  int nfeat = mlp::kFeat;  // pose inputs used (the rest zero): 12 or 21
};

struct TrainParams {
  uint32_t batch = 8192, epochs = 1;
  float lr = 1e-3f, tail_weight = 1.f;
  uint64_t seed = 0;
};

struct TrainReport {
  float final_loss = 0.f;
  uint32_t steps = 0;
};

// Adam over mini-batches of the label indices `train` (reshuffled every
// epoch with a seeded mt19937_64); learning rate: linear warm-up, then
// cosine decay to 5 %.  loss_curve[e] (if set, e < cap) = epoch e's mean
// squared error in standardised units.  A step runs as one captured
// hipGraph replayed step after step (every per-step value is computed on
// the device from the step count), or, with BUDE_MI355X_TRAIN_GRAPH=0 in the
// environment, as the same kernels launched step by step: the same network
// bit for bit.
int train(Net &net, const TrainData &d, std::vector<uint32_t> train, const TrainParams &p, hipStream_t s,
          float *loss_curve, size_t cap, TrainReport *rep);

// The fused inference kernel over n poses (6 rows of `stride` floats) on a
// persistent grid of `blocks` work-groups; bias0 replaces the first-layer
// bias when set (a per-ligand bias with descriptor inputs folded in).
int predict(const Net &net, const float *poses, size_t stride, size_t n, float *out, int blocks, hipStream_t s,
            const float *bias0 = nullptr);

}  // namespace surrogate
}  // namespace bude
