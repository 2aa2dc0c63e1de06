// surrogate.hip -- training state and training loop of the MLP surrogate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>

#include "surrogate.h"

namespace bude {
namespace surrogate {

namespace {

mlp::GemmArgs gemm_args(const void *a, const void *b, int M, int N, int K, int lda, int ldb, int epi) {
  mlp::GemmArgs g{};
  g.a = static_cast<const bf16 *>(a);
  g.b = static_cast<const bf16 *>(b);
  g.M = M;
  g.N = N;
  g.K = K;
  g.lda = lda;
  g.ldb = ldb;
  g.epi = epi;
  return g;
}

}  // namespace

int alloc_acts(Acts &A, size_t rows, bool training) {
  A.rows = rows;
  const size_t H = mlp::kHidden, O = mlp::kOut, I = mlp::kIn;
  int st = A.feat.alloc(rows * I * 2);
  if (!st) st = A.pred.alloc(rows * 4);
  for (int l = 0; l < 3 && !st; l++) st = A.a[l].alloc(rows * H * 2);
  if (training) {
    if (!st) st = A.feat_t.alloc(rows * I * 2);
    for (int l = 0; l < 3 && !st; l++) {
      st = A.z[l].alloc(rows * H * 4);
      if (!st) st = A.a_t[l].alloc(rows * H * 2);
      if (!st) st = A.dz[l].alloc(rows * H * 2);
      if (!st) st = A.dz_t[l].alloc(rows * H * 2);
    }
    if (!st) st = A.dz3.alloc(rows * O * 2);
    if (!st) st = A.dz3_t.alloc(rows * O * 2);
    if (!st) st = A.loss.alloc(256);
    // split-K partial sums of the largest weight gradient
    size_t wmax = 0;
    for (int l = 0; l < 4; l++)
      wmax = std::max(wmax, size_t(mlp::auto_ksplit(kK[l], kN[l], int(rows))) * kK[l] * kN[l]);
    if (!st) st = A.work.alloc(wmax * 4);
  }
  return st;
}

// Forward over `rows` rows already featurised in A.feat; training stores
// z, a, a^T for the backward pass.
int forward(const Net &net, Acts &A, size_t rows, bool training, hipStream_t s) {
  const int R = int(rows);
  const void *in = A.feat.ptr;
  int in_ld = mlp::kIn;
  for (int l = 0; l < 3; l++) {
    auto g = gemm_args(in, net.w16t[l].ptr, R, kN[l], kK[l], in_ld, kK[l], mlp::kEpiHidden);
    g.bias = static_cast<const float *>(net.b[l].ptr);
    g.out_f32 = training ? static_cast<float *>(A.z[l].ptr) : nullptr;
    g.out_bf16 = static_cast<bf16 *>(A.a[training ? l : (l & 1)].ptr);
    g.out_t = training ? static_cast<bf16 *>(A.a_t[l].ptr) : nullptr;
    g.ldo = kN[l];
    g.ldt = R;
    if (int st = hip_check(mlp::gemm(g, s), "mlp gemm forward")) return st;
    in = g.out_bf16;
    in_ld = kN[l];
  }
  auto g = gemm_args(in, net.w16t[3].ptr, R, kN[3], kK[3], in_ld, kK[3], mlp::kEpiOut);
  g.bias = static_cast<const float *>(net.b[3].ptr);
  g.out_f32 = static_cast<float *>(A.pred.ptr);
  return hip_check(mlp::gemm(g, s), "mlp gemm output");
}

// Backward from dz3 (already written by the loss kernel) into net.gw / gb.
int backward(Net &net, Acts &A, size_t rows, hipStream_t s) {
  const int R = int(rows);
  mlp::BiasGradBatch bg{};
  bg.count = 4;
  bg.B = R;
  bg.ldt = R;
  const void *dz_t_next = A.dz3_t.ptr;  // [N_l][R]
  const void *dz_next = A.dz3.ptr;      // [R][N_l]
  for (int l = 3; l >= 0; l--) {
    // dW_l = A_{l-1}^T dZ_l  -> [K_l][N_l]
    const void *prev_t = l == 0 ? A.feat_t.ptr : A.a_t[l - 1].ptr;
    auto gw = gemm_args(prev_t, dz_t_next, kK[l], kN[l], R, R, R, mlp::kEpiF32);
    gw.out_f32 = static_cast<float *>(net.gw[l].ptr);
    gw.ldo = kN[l];
    gw.ksplit = mlp::auto_ksplit(kK[l], kN[l], R);
    gw.work = static_cast<float *>(A.work.ptr);
    if (int st = hip_check(mlp::gemm(gw, s), "mlp gemm dW")) return st;
    bg.dz_t[l] = static_cast<const bf16 *>(dz_t_next);
    bg.db[l] = static_cast<float *>(net.gb[l].ptr);
    bg.N[l] = kN[l];
    if (l == 0) break;
    // dZ_{l-1} = (dZ_l W_l^T) * gelu'(Z_{l-1})  -> [R][K_l]
    auto ga = gemm_args(dz_next, net.w16[l].ptr, R, kK[l], kN[l], kN[l], kN[l], mlp::kEpiBackGelu);
    ga.zprev = static_cast<const float *>(A.z[l - 1].ptr);
    ga.out_bf16 = static_cast<bf16 *>(A.dz[l - 1].ptr);
    ga.out_t = static_cast<bf16 *>(A.dz_t[l - 1].ptr);
    ga.ldo = kK[l];
    ga.ldt = R;
    if (int st = hip_check(mlp::gemm(ga, s), "mlp gemm dA")) return st;
    dz_next = A.dz[l - 1].ptr;
    dz_t_next = A.dz_t[l - 1].ptr;
  }
  return hip_check(mlp::bias_grad_multi(bg, s), "mlp bias grads");
}

int adam_step(Net &net, float lr, hipStream_t s) {
  net.step++;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float bc1 = 1.f - std::pow(b1, float(net.step)), bc2 = 1.f - std::pow(b2, float(net.step));
  mlp::AdamBatch ab{};
  for (int l = 0; l < 4; l++) {
    ab.t[2 * l] = {static_cast<float *>(net.w[l].ptr), static_cast<const float *>(net.gw[l].ptr),
                   static_cast<float *>(net.mw[l].ptr), static_cast<float *>(net.vw[l].ptr),
                   static_cast<bf16 *>(net.w16[l].ptr), static_cast<bf16 *>(net.w16t[l].ptr), kK[l], kN[l]};
    ab.t[2 * l + 1] = {static_cast<float *>(net.b[l].ptr), static_cast<const float *>(net.gb[l].ptr),
                       static_cast<float *>(net.mb[l].ptr), static_cast<float *>(net.vb[l].ptr), nullptr, nullptr, 1,
                       kN[l]};
  }
  ab.count = 8;
  return hip_check(mlp::adam_multi(ab, lr, b1, b2, eps, bc1, bc2, s), "adam");
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
    if (!st) st = net.w16[l].alloc(nw * 2);
    if (!st) st = net.w16t[l].alloc(nw * 2);
    if (!st)
      st = hip_check(mlp::to_bf16(static_cast<const float *>(net.w[l].ptr), kK[l], kN[l],
                                              static_cast<bf16 *>(net.w16[l].ptr),
                                              static_cast<bf16 *>(net.w16t[l].ptr), s),
                           "to_bf16");
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
  DeviceBuffer dperm;
  if (int st = dperm.alloc(ntrain * 4)) return st;
  const uint32_t epochs = std::max<uint32_t>(1, p.epochs);
  const size_t steps_per_epoch = ntrain / B;
  const size_t total = steps_per_epoch * epochs;
  const size_t warm = std::min<size_t>(200, total / 10 + 1);
  size_t step = 0;
  TrainReport r{};
  for (uint32_t ep = 0; ep < epochs; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    if (int st = hip_check(hipMemcpyAsync(dperm.ptr, perm.data(), ntrain * 4, hipMemcpyHostToDevice, s),
                           "hipMemcpyAsync"))
      return st;
    if (int st = hip_check(hipMemsetAsync(A.loss.ptr, 0, 4, s), "hipMemsetAsync")) return st;
    for (size_t it = 0; it < steps_per_epoch; it++, step++) {
      const uint32_t *idx = static_cast<const uint32_t *>(dperm.ptr) + it * B;
      if (int st = hip_check(mlp::features(d.poses, d.stride, idx, B, static_cast<bf16 *>(A.feat.ptr),
                                           static_cast<bf16 *>(A.feat_t.ptr), int(B), s, d.desc, d.lig, d.nfeat),
                             "features"))
        return st;
      if (int st = forward(net, A, B, true, s)) return st;
      if (int st = hip_check(mlp::loss_grad(static_cast<const float *>(A.pred.ptr), d.energy, idx, B, d.norm, d.table,
                                            d.lig, p.tail_weight, static_cast<bf16 *>(A.dz3.ptr),
                                            static_cast<bf16 *>(A.dz3_t.ptr), int(B),
                                            static_cast<float *>(A.loss.ptr), s),
                             "loss"))
        return st;
      if (int st = backward(net, A, B, s)) return st;
      const float warmf = step < warm ? float(step + 1) / float(warm) : 1.f;
      const float prog = float(step) / float(std::max<size_t>(1, total - 1));
      const float lr = p.lr * warmf * (0.05f + 0.95f * 0.5f * (1.f + std::cos(3.14159265f * prog)));
      if (int st = adam_step(net, lr, s)) return st;
    }
    float lsum = 0.f;
    if (int st = hip_check(hipMemcpyAsync(&lsum, A.loss.ptr, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync"))
      return st;
    if (int st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")) return st;
    const float mse = lsum / float(steps_per_epoch * B);
    if (loss_curve && ep < cap) loss_curve[ep] = mse;
    r.final_loss = mse;
  }
  r.steps = uint32_t(step);
  if (rep) *rep = r;
  return BUDE_OK;
}

int predict(const Net &net, const float *poses, size_t stride, size_t n, float *out, int blocks, hipStream_t s,
            const float *bias0) {
  const bf16 *wt[4] = {static_cast<const bf16 *>(net.w16t[0].ptr), static_cast<const bf16 *>(net.w16t[1].ptr),
                       static_cast<const bf16 *>(net.w16t[2].ptr), static_cast<const bf16 *>(net.w16t[3].ptr)};
  const float *bias[4] = {bias0 ? bias0 : static_cast<const float *>(net.b[0].ptr),
                          static_cast<const float *>(net.b[1].ptr), static_cast<const float *>(net.b[2].ptr),
                          static_cast<const float *>(net.b[3].ptr)};
  return hip_check(mlp::infer(poses, stride, n, wt, static_cast<const bf16 *>(net.w16[3].ptr), bias, out, blocks, s),
                   "mlp infer");
}

}  // namespace surrogate
}  // namespace bude
