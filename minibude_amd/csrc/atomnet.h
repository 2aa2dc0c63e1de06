// atomnet.h -- the learned per-atom field: a per-protein MLP surrogate of
// the energy one ligand atom of a given forcefield type has at a given
// position, summed over a ligand's atoms to score a pose.
//
// fasten_main's pose energy is a sum over ligand atoms of terms that depend
// only on the atom's type and its transformed position
// (src/serial/fasten.hpp:56-131), so a surrogate of that one-atom energy
// scores every ligand of a screen with one network per protein.  The
// network:
//
//   phi(p)   = trilinear interpolation of a learned feature grid
//              G [n0][n1][n2][kF] (spacing h, cell and fraction clamped)
//   x0       = [phi(p), onehot(slot), 0 ...]   (k0 = 32 * ceil((kF + nslots) / 32))
//   a1       = relu(x0 W0 + b0)                (kH)
//   a2       = relu(a1 W1 + b1)                (kH)
//   y        = a2 W2[:, 0] + b2[0]
//   E(p, s)  = sigma * y + mu                  (energy units)
//
// BF16 weights, activations and inputs on MFMA with FP32 accumulation,
// trained by Adam (grid and weights) on exact one-atom energies from the
// plain kernel's pair loop at positions the screened poses put ligand atoms
// at.  Inference folds the one-hot row of W0 into a per-slot first-layer
// bias and reads the grid as FP16 pairs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fasten_batch.h"
#include "field.h"
#include "hip_util.h"
#include "mlp.h"

namespace bude {
namespace atomnet {

using mlp::bf16;

constexpr int kF = 8;               // grid features per node
constexpr int kH = 32;              // hidden width
constexpr int kOut = mlp::kOut;     // output layer padded to one MFMA tile (column 0 used)
constexpr int kMaxSlots = 64 - kF;  // forcefield slots: k0 <= 64
constexpr double kGradScale = 1099511627776.0;  // 2^40: fixed-point grid gradients (int64 atomics)

inline int input_width(int nslots) { return 32 * ((kF + nslots + 31) / 32); }

struct Net {
  field::Grid grid{};
  int nslots = 0, k0 = 0;
  size_t nodes = 0;
  // grid features: FP32 master [nodes][kF], Adam moments, int64 fixed-point
  // gradient accumulators (cleared by the Adam step); for inference the
  // cell table gzp [n0 - 1][n1 - 1][n2 - 1][4][kF] of FP16 pairs: per cell
  // (i, j, l) and corner column q = (dx, dy), feature k of node (i + dx,
  // j + dy, l) and of node (i + dx, j + dy, l + 1) -- a cell's 8 corners
  // in one 128-byte record, one packed dot product per pair
  DeviceBuffer g, gm, gv, gacc, gzp;
  // layers l = 0..2: W_l [K_l][N_l] (FP32 master, Adam moments, gradient,
  // BF16 row-major and transposed copies), b_l [N_l]
  DeviceBuffer w[3], b[3], mw[3], vw[3], mb[3], vb[3], gw[3], gb[3], w16[3], w16t[3];
  // [nslots][kH] first-layer bias of each slot, b0 + W0[kF + s] (FP32):
  // inference starts the layer's MFMA accumulators from it
  DeviceBuffer slot_bias;
  float mu = 0.f, sigma = 1.f, clip = 0.f;
  float b2 = 0.f;  // b_2[0], read back by finalize
  int step = 0;
};

// K and N of layer l.
inline int layer_k(const Net &n, int l) { return l == 0 ? n.k0 : kH; }
inline int layer_n(int l) { return l == 2 ? kOut : kH; }

// Sizes the network for `grid` and `nslots` (keeping allocations large
// enough from an earlier call) and initialises it on stream s: grid
// features N(0, 0.1^2), He-normal hidden weights, 1/sqrt(fan_in) output
// weights, zero biases (from a hash of seed), Adam state cleared.
int init(Net &net, const field::Grid &grid, int nslots, uint64_t seed, hipStream_t s);
// Sets every parameter from host arrays (grid [nodes][kF], W_l [K_l][N_l]
// with W2 [kH][1] placed in column 0, b_l) and clears the Adam state.
int set_params(Net &net, const float *grid, const float *const w[3], const float *const b[3], hipStream_t s);

// n points at positions ligand atoms take under the screened poses: point
// i is atom a of ligand l under pose j (l, j, a drawn from a hash of seed and
// i), as (x, y, z, slot bits) with slot = (i / 64) % nslots (one slot per
// 64 consecutive points).  ligands: (x, y, z, slot bits) per atom.
hipError_t sample_points(const float4 *ligands, const LigandMeta *meta, int nligands, const float *poses,
                         size_t stride, int nslots, uint64_t seed, size_t n, float4 *out, hipStream_t s);

// Exact energy of one ligand atom at each point (n a multiple of 64, one
// slot per 64 points): the plain kernel's pair loop over the grouped protein
// with the pair coefficients slot_coef[slot * ngroups + g], x 0.5 -- the
// plain kernel's energy of a one-atom ligand there, bit for bit.
hipError_t label_points(const float4 *protein, const int32_t *group_begin, const PairCoef *slot_coef, int ngroups,
                        const float4 *points, size_t n, int form, float *out, hipStream_t s);

struct TrainParams {
  uint32_t batch = 16384;  // multiple of 64
  uint32_t steps = 2000;
  float lr = 2e-3f;        // weights: peak Adam learning rate (warm-up, cosine decay)
  float grid_lr = 1e-2f;   // grid features
  uint64_t seed = 0;
};

struct TrainReport {
  float val_rmse = 0.f;    // one-atom energy RMSE on the held-out points (energy units)
  float final_loss = 0.f;  // mean squared error over all training steps (standardised units)
  uint32_t steps = 0;
};

// Training buffers of one batch (reused from call to call: an allocation
// can wait for the kernels in flight).  A step is one fused kernel
// (forward, loss, backward, grid gradient) between the step's batch
// indices and its Adam updates; its activations never leave LDS.
struct Workspace {
  size_t rows = 0;
  int k0 = 0;
  DeviceBuffer pred, loss, part, vpred;
  // weight and bias gradients of layer l as splits[l] slices, one per
  // work-group of the step (summed in a fixed order by the Adam step); room for
  // `slices` of them
  DeviceBuffer gpart[3], bpart[3];
  int splits[3] = {1, 1, 1};
  unsigned slices = 0;
  // per-step state in device memory (step count, learning rates, Adam bias
  // corrections) and the step's batch indices, so that one captured step
  // replays as every step
  DeviceBuffer state, idx;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  Workspace() = default;
  Workspace(const Workspace &) = delete;
  Workspace &operator=(const Workspace &) = delete;
  ~Workspace() { drop_graph(); }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};
int reserve(const Net &net, Workspace &ws, size_t batch, size_t nval);

// Training runs as one captured hipGraph of a whole step (the fused step
// kernel, which reads its batch through the epoch's permutation, then one
// Adam launch over the weights and the grid) replayed `steps` times (every per-step value is computed on the device
// from the step count), or, with BUDE_MI355X_TRAIN_GRAPH=0 in the
// environment, as the same kernels launched step by step; both give the
// same network bit for bit.
// Trains on points[0, ntrain) with labels (energy units), validates on
// points[ntrain, ntrain + nval); sets mu / sigma / clip from the training
// labels (target y = (min(E, clip) - mu) / sigma, clip = mu + 8 sigma).
// perm: device array of ntrain point indices, read in batches (wrapping).
int train(Net &net, Workspace &ws, const float4 *points, const float *labels, const uint32_t *perm, size_t ntrain,
          size_t nval, const TrainParams &p, hipStream_t s, TrainReport *rep);

// One forward + backward pass of the training kernels over `batch` points
// (multiple of 64) against standardised targets: predictions, loss and the
// FP32 gradients (grid gradient in grad_grid [nodes][kF]).  For gradient
// checks.
int forward_backward(Net &net, const float4 *points, const float *target, size_t batch, float *pred, float *loss,
                     float *grad_grid, hipStream_t s);

// Per-slot first-layer biases for inference and b_2[0] on the host (after
// training or set_params; synchronises the stream).
int finalize(Net &net, hipStream_t s);

struct ScoreArgs {
  const uint32_t *gzp;        // cell table (Net::gzp)
  field::Grid grid;
  const bf16 *w0t;            // W0^T [kH][k0]: columns 0..kF-1 read
  int k0;
  const float *slot_bias;     // [nslots][kH] b0 + W0[kF + s]
  int nslots;
  const bf16 *w1t;            // W1^T [kH][kH]
  const float *b1;            // [kH]
  const bf16 *w2;             // W2 [kH][kOut], column 0
  float b2, sigma, mu;
  const float4 *ligands;      // x, y, z, slot (int bits)
  const LigandMeta *meta;
  const int64_t *tile_begin;  // nligands + 1: first 64-pose tile of each ligand
  int nligands;
  const float *poses;         // 6 rows of `stride` floats
  size_t stride;
  float *out;                 // out[meta.pose_off + j]: sum over atoms of E
  int64_t tile0;
  bool raise_priority;
};

// Surrogate energy of every pose of tiles [tile0, tile0 + ntiles): one wave
// per tile of 64 poses of one ligand (the tiling of field::score).
hipError_t score(const ScoreArgs &a, int64_t ntiles, hipStream_t s);
ScoreArgs score_args(const Net &net);

}  // namespace atomnet
}  // namespace bude
