// surrogate.hip -- training state and training loop of the MLP surrogate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <random>
#include <vector>

#include "surrogate.h"

namespace bude {
namespace surrogate {

int alloc_acts(Acts &A, size_t rows, bool training) {
  A.rows = rows;
  A.ld_t = mlp::t_ld(rows);
  const size_t H = mlp::kHidden, O = mlp::kOut, I = mlp::kIn, ld = A.ld_t;
  int st = A.feat.alloc(rows * I * 2);
  if (!st) st = A.pred.alloc(rows * 4);
  if (training) {
    if (!st) st = A.feat_t.alloc(ld * I * 2);
    for (int l = 0; l < 3 && !st; l++) {
      st = A.a_t[l].alloc(ld * H * 2);
      if (!st) st = A.dz_t[l].alloc(ld * H * 2);
    }
    // dz3^T: row 0 is written every step, rows 1.. stay zero (the padded
    // output columns of the last weight gradient)
    if (!st) st = A.dz3_t.alloc(ld * O * 2);
    if (!st) st = hip_check(hipMemset(A.dz3_t.ptr, 0, ld * O * 2), "hipMemset");
    if (!st) st = A.loss.alloc(256);
    for (int l = 0; l < 4 && !st; l++) {
      const size_t cap = size_t(mlp::kDwSplits);
      st = A.gpart[l].alloc(cap * kK[l] * kN[l] * 4);
      if (!st) st = A.bpart[l].alloc(cap * kN[l] * 4);
    }
  }
  return st;
}

int rows_step(const Net &net, Acts &A, size_t rows, const mlp::RowsArgs &in, hipStream_t s) {
  mlp::RowsArgs a = in;
  a.n = rows;
  a.ld_t = A.ld_t;
  for (int l = 0; l < 3; l++) {
    a.wt[l] = static_cast<const bf16 *>(net.w16tx[l].ptr);
    a.a_t[l] = static_cast<bf16 *>(A.a_t[l].ptr);
    a.dz_t[l] = static_cast<bf16 *>(A.dz_t[l].ptr);
  }
  for (int l = 0; l < 4; l++) {
    a.w[l] = static_cast<const bf16 *>(l == 1 || l == 2 ? net.w16x[l].ptr : net.w16[l].ptr);  // (w[0] unused)
    a.bias[l] = static_cast<const float *>(net.b[l].ptr);
  }
  a.loss = static_cast<unsigned long long *>(A.loss.ptr);
  a.feat_t = static_cast<bf16 *>(A.feat_t.ptr);
  a.dz3_t = static_cast<bf16 *>(A.dz3_t.ptr);
  return hip_check(mlp::rows(a, s), "mlp rows");
}

// dW_l = A_{l-1}^T dZ_l -> [K_l][N_l] and db_l = the column sums of dZ_l,
// as mlp::dw_splits(rows) K slices (split over the batch rows) from the transposed
// rows, all four layers in one launch.
int weight_grads(Net &net, Acts &A, size_t rows, hipStream_t s) {
  (void)net;
  mlp::DwArgs a{};
  a.count = 4;
  a.k = rows;
  a.ld = A.ld_t;
  a.splits = mlp::dw_splits(rows);
  for (int l = 0; l < 4; l++) {
    a.d[l].p = static_cast<const bf16 *>(l == 0 ? A.feat_t.ptr : A.a_t[l - 1].ptr);
    a.d[l].q = static_cast<const bf16 *>(l == 3 ? A.dz3_t.ptr : A.dz_t[l].ptr);
    a.d[l].M = kK[l];
    a.d[l].N = kN[l];
    a.d[l].work = static_cast<float *>(A.gpart[l].ptr);
    a.d[l].bias_part = static_cast<float *>(A.bpart[l].ptr);
    A.splits[l] = a.splits;
  }
  return hip_check(mlp::weight_grads(a, s), "mlp weight gradients");
}

int sum_grads(Net &net, const Acts &A, hipStream_t s) {
  for (int l = 0; l < 4; l++) {
    if (int st = hip_check(mlp::sum_slices(static_cast<const float *>(A.gpart[l].ptr), A.splits[l],
                                           size_t(kK[l]) * kN[l], static_cast<float *>(net.gw[l].ptr), s),
                           "mlp dW"))
      return st;
    if (int st = hip_check(mlp::sum_slices(static_cast<const float *>(A.bpart[l].ptr), A.splits[l], size_t(kN[l]),
                                           static_cast<float *>(net.gb[l].ptr), s),
                           "mlp db"))
      return st;
  }
  return BUDE_OK;
}

// Upload FP32 weights W_l [kK[l]][kN[l]] (zero padded) and biases; refresh BF16 copies.
int set_weights(Net &net, const std::vector<std::vector<float>> &w, const std::vector<std::vector<float>> &b,
                hipStream_t s) {
  for (int l = 0; l < 4; l++) {
    const size_t nw = size_t(kK[l]) * kN[l];
    int st = net.w[l].upload(w[l].data(), nw * 4);
    if (!st) st = net.b[l].upload(b[l].data(), size_t(kN[l]) * 4);
    for (DeviceBuffer *d : {&net.mw[l], &net.vw[l], &net.gw[l]})
      if (!st) {
        st = d->alloc(nw * 4);
        if (!st) st = hip_check(hipMemset(d->ptr, 0, nw * 4), "hipMemset");
      }
    for (DeviceBuffer *d : {&net.mb[l], &net.vb[l], &net.gb[l]})
      if (!st) {
        st = d->alloc(size_t(kN[l]) * 4);
        if (!st) st = hip_check(hipMemset(d->ptr, 0, size_t(kN[l]) * 4), "hipMemset");
      }
    if (!st && l == 3) st = net.w16[l].alloc(nw * 2);
    if (!st && l == 3)
      st = hip_check(mlp::to_bf16(static_cast<const float *>(net.w[l].ptr), kK[l], kN[l],
                                  static_cast<bf16 *>(net.w16[l].ptr), nullptr, s),
                     "to_bf16");
    if (!st && l < 3) st = net.w16tx[l].alloc(nw * 2);
    if (!st && (l == 1 || l == 2)) st = net.w16x[l].alloc(nw * 2);
    if (!st && l < 3)
      st = hip_check(mlp::tile_bf16(static_cast<const float *>(net.w[l].ptr), kK[l], kN[l],
                                    static_cast<bf16 *>(net.w16x[l].ptr), static_cast<bf16 *>(net.w16tx[l].ptr), s),
                     "tile_bf16");
    if (st) return st;
  }
  net.step = 0;
  return hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

void init_weights(uint64_t seed, int inputs, std::vector<std::vector<float>> &w, std::vector<std::vector<float>> &b) {
  std::mt19937_64 g(seed ^ 0x5eedf00dull);
  std::normal_distribution<float> nd(0.f, 1.f);
  w.assign(4, {});
  b.assign(4, {});
  for (int l = 0; l < 4; l++) {
    w[l].assign(size_t(kK[l]) * kN[l], 0.f);
    b[l].assign(size_t(kN[l]), 0.f);
    const int fan_in = l == 0 ? inputs : kK[l];
    const int rows = l == 0 ? inputs : kK[l];
    const int cols = l == 3 ? 1 : kN[l];
    const float sd = l == 3 ? std::sqrt(1.f / float(fan_in)) : std::sqrt(2.f / float(fan_in));
    for (int r = 0; r < rows; r++)
      for (int c = 0; c < cols; c++) w[l][size_t(r) * kN[l] + c] = sd * nd(g);
  }
}

int train(Net &net, const TrainData &d, std::vector<uint32_t> perm, const TrainParams &p, hipStream_t s,
          float *loss_curve, size_t cap, TrainReport *rep) {
  const size_t B = p.batch, ntrain = perm.size();
  if (B == 0 || B % 32 || ntrain < B || !d.poses || !d.energy || (d.table && !d.lig) || (d.desc && !d.lig))
    return fail(BUDE_ERR_INVALID_ARGUMENT, "surrogate train: need at least one batch of labels");
  Acts A;
  if (int st = alloc_acts(A, B, true)) return st;
  std::mt19937_64 rng(p.seed);
  DeviceBuffer dperm, dstate;
  if (int st = dperm.alloc(ntrain * 4)) return st;
  if (int st = dstate.alloc(sizeof(mlp::StepState))) return st;
  const uint32_t epochs = std::max<uint32_t>(1, p.epochs);
  const uint32_t per_epoch = uint32_t(ntrain / B);
  const uint32_t total = per_epoch * epochs;
  const uint32_t warm = std::min<uint32_t>(200, total / 10 + 1);
  const uint32_t step0 = uint32_t(net.step);
  auto *state = static_cast<mlp::StepState *>(dstate.ptr);
  // one step with every per-step value (batch indices, learning rate, Adam
  // bias corrections) computed on the device from the step count, captured
  // once as a hipGraph and replayed step after step
  mlp::RowsArgs in{};
  in.poses = d.poses;
  in.stride = d.stride;
  in.desc = d.desc;
  in.lig = d.lig;
  in.nfeat = d.nfeat;
  in.perm = static_cast<const uint32_t *>(dperm.ptr);
  in.per_epoch = per_epoch;
  in.state = state;
  in.sched = mlp::Schedule{total, warm, step0, p.lr, 0.f, 0.9f, 0.999f};
  in.energy = d.energy;
  in.norm = d.norm;
  in.table = d.table;
  in.tail_weight = p.tail_weight;
  auto one_step = [&]() -> int {
    if (int st = rows_step(net, A, B, in, s)) return st;
    if (int st = weight_grads(net, A, B, s)) return st;
    // Adam over the gradient slices; the same launch ends the step
    mlp::AdamBatch ab{};
    for (int l = 0; l < 4; l++) {
      ab.t[2 * l] = {static_cast<float *>(net.w[l].ptr), static_cast<const float *>(A.gpart[l].ptr),
                     static_cast<float *>(net.mw[l].ptr), static_cast<float *>(net.vw[l].ptr),
                     l == 3 ? static_cast<bf16 *>(net.w16[l].ptr) : nullptr, nullptr, kK[l], kN[l],
                     A.splits[l], size_t(kK[l]) * kN[l], static_cast<bf16 *>(net.w16x[l].ptr),
                     static_cast<bf16 *>(net.w16tx[l].ptr)};
      ab.t[2 * l + 1] = {static_cast<float *>(net.b[l].ptr), static_cast<const float *>(A.bpart[l].ptr),
                         static_cast<float *>(net.mb[l].ptr), static_cast<float *>(net.vb[l].ptr), nullptr, nullptr, 1,
                         kN[l], A.splits[l], size_t(kN[l])};
    }
    ab.count = 8;
    return hip_check(mlp::adam_multi_dev(ab, mlp::step_hyper(state), 0.9f, 0.999f, 1e-8f, s, state), "adam");
  };
  if (int st = hip_check(hipMemsetAsync(dstate.ptr, 0, sizeof(mlp::StepState), s), "hipMemsetAsync")) return st;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  struct GraphGuard {
    hipGraph_t &g;
    hipGraphExec_t &e;
    ~GraphGuard() {
      if (e) (void)hipGraphExecDestroy(e);
      if (g) (void)hipGraphDestroy(g);
    }
  } guard{graph, exec};
  const char *env = std::getenv("BUDE_MI355X_TRAIN_GRAPH");
  const bool use_graph = !(env && env[0] == '0');
  if (use_graph) {
    if (int st = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
      return st;
    const int cst = one_step();
    const hipError_t ce = hipStreamEndCapture(s, &graph);
    if (cst) return cst;
    if (int st = hip_check(ce, "hipStreamEndCapture")) return st;
    if (int st = hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate")) return st;
  }
  // epoch e's order is shuffled on the host while epoch e - 1 runs on the
  // device: two pinned staging buffers (one is refilled only after its last
  // copy has ended), the epoch's loss sum set aside on the device, and one
  // synchronization at the end
  PinnedBuffer host_perm[2];
  DeviceBuffer dlosses;
  hipEvent_t copied[2] = {nullptr, nullptr};
  struct EventGuard {
    hipEvent_t (&e)[2];
    ~EventGuard() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } event_guard{copied};
  for (int k = 0; k < 2; k++) {
    if (int st = host_perm[k].alloc(ntrain * 4)) return st;
    if (int st = hip_check(hipEventCreateWithFlags(&copied[k], hipEventDisableTiming), "hipEventCreate")) return st;
  }
  if (int st = dlosses.alloc(size_t(epochs) * 8)) return st;
  for (uint32_t ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    const int k = int(ep & 1);
    if (ep >= 2)
      if (int st = hip_check(hipEventSynchronize(copied[k]), "hipEventSynchronize")) return st;
    std::copy(perm.begin(), perm.end(), static_cast<uint32_t *>(host_perm[k].ptr));
    if (int st = hip_check(hipMemcpyAsync(dperm.ptr, host_perm[k].ptr, ntrain * 4, hipMemcpyHostToDevice, s),
                           "hipMemcpyAsync"))
      return st;
    if (int st = hip_check(hipEventRecord(copied[k], s), "hipEventRecord")) return st;
    if (int st = hip_check(hipMemsetAsync(A.loss.ptr, 0, 8, s), "hipMemsetAsync")) return st;
    for (uint32_t it = 0; it < per_epoch; it++) {
      const int st = use_graph ? hip_check(hipGraphLaunch(exec, s), "hipGraphLaunch") : one_step();
      if (st) return st;
    }
    if (int st = hip_check(hipMemcpyAsync(static_cast<unsigned long long *>(dlosses.ptr) + ep, A.loss.ptr, 8,
                                          hipMemcpyDeviceToDevice, s),
                           "hipMemcpyAsync"))
      return st;
  }
  std::vector<unsigned long long> lsum(epochs);
  if (int st = hip_check(hipMemcpyAsync(lsum.data(), dlosses.ptr, size_t(epochs) * 8, hipMemcpyDeviceToHost, s),
                         "hipMemcpyAsync"))
    return st;
  if (int st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")) return st;
  TrainReport r{};
  for (uint32_t ep = 0; ep < epochs; ep++) {
    const float mse = float(double(lsum[ep]) / mlp::kLossScale / double(size_t(per_epoch) * B));
    if (loss_curve && ep < cap) loss_curve[ep] = mse;
    r.final_loss = mse;
  }
  net.step += int(total);
  r.steps = total;
  if (rep) *rep = r;
  return BUDE_OK;
}

int predict(const Net &net, const float *poses, size_t stride, size_t n, float *out, int blocks, hipStream_t s,
            const float *bias0) {
  const bf16 *wt[4] = {static_cast<const bf16 *>(net.w16tx[0].ptr), static_cast<const bf16 *>(net.w16tx[1].ptr),
                       static_cast<const bf16 *>(net.w16tx[2].ptr), nullptr};
  const float *bias[4] = {bias0 ? bias0 : static_cast<const float *>(net.b[0].ptr),
                          static_cast<const float *>(net.b[1].ptr), static_cast<const float *>(net.b[2].ptr),
                          static_cast<const float *>(net.b[3].ptr)};
  return hip_check(mlp::infer(poses, stride, n, wt, static_cast<const bf16 *>(net.w16[3].ptr), bias, out, blocks, s),
                   "mlp infer");
}

}  // namespace surrogate
}  // namespace bude
