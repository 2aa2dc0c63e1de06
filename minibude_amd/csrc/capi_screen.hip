// capi_screen.hip -- C ABI of the learned prescreen: an MLP pose-energy
// surrogate trained on the GPU (own BF16-MFMA forward/backward kernels,
// Adam) on energies the plain kernel computes, then used to screen a pose
// set: surrogate on every pose, exact plain-kernel rescoring of the lowest
// keep_fraction, plain kernel on everything when the surrogate's validation
// error is above the configured bound (fallback).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "bude_internal.h"
#include "fasten_batch.h"
#include "fasten_plain.h"
#include "hip_util.h"
#include "mlp.h"
#include "select.h"
#include "surrogate.h"

namespace {

using bude::DeviceBuffer;
using bude::mlp::bf16;
using bude::surrogate::Acts;
using bude::surrogate::Net;
using bude::surrogate::kK;
using bude::surrogate::kN;

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct bude_screen {
  int device = 0;
  int cus = 0;
  size_t lds_per_cu = 0;
  BudeScreenConfig cfg{};
  // plain-kernel deck (labels, rescoring, fallback)
  size_t natlig = 0, ngroups = 0;
  DeviceBuffer protein, group_begin, coef, ligand;
  // surrogate
  Net net;
  bool trained = false, fallback = false;
  float clip = 0.f, mu = 0.f, sigma = 1.f, tail = -3.0e38f, val_rmse = 0.f;
  // candidate poses
  size_t ncand = 0;
  DeviceBuffer cand, pred, sel_idx, sel_n, exact;
  // rescoring: the persistent batched kernel over the selection (its rows
  // through sel_idx), one ligand whose pose count and tiles (0, 64 j) the
  // device writes from the selection's size
  DeviceBuffer rs_tiles, rs_meta, rs_counter, rs_order, rs_ntiles;
  bude::sel::Plan plan;  // one segment: the loaded poses, k = ceil(keep_fraction * n)
  size_t plan_n = 0, plan_k = 0;
  size_t nsel = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[5] = {};

  ~bude_screen() {
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  bude::PlainArgs plain(const float *poses, size_t stride, size_t n, float *energies) const {
    bude::PlainArgs a{};
    a.protein = static_cast<const float4 *>(protein.ptr);
    a.group_begin = static_cast<const int32_t *>(group_begin.ptr);
    a.coef = static_cast<const bude::PairCoef *>(coef.ptr);
    a.ligand = static_cast<const float4 *>(ligand.ptr);
    a.poses = poses;
    a.energies = energies;
    a.stride = stride;
    a.nposes = n;
    a.natlig = int(natlig);
    a.ngroups = int(ngroups);
    a.ppwi = 1;
    a.form = bude::plain_form();
    return a;
  }
};

namespace {

int upload_poses(DeviceBuffer &d, const float *const poses[6], size_t n) {
  if (int st = d.alloc(6 * n * sizeof(float))) return st;
  for (int k = 0; k < 6; k++)
    if (int st = bude::hip_check(hipMemcpy(static_cast<float *>(d.ptr) + size_t(k) * n, poses[k], n * sizeof(float),
                                           hipMemcpyHostToDevice),
                                 "hipMemcpy poses"))
      return st;
  return BUDE_OK;
}

// Exact energies of the selected poses (rows sel_idx[0, nsel) of the
// candidates) into exact[0, nsel): the persistent batched kernel
// (bit-identical to the plain kernel) on tiles of 64 poses pulled from a
// queue by a grid pinned to 3 waves on every SIMD, so the SIMDs finish
// together on a pass of a few thousand tiles.  The tables come from the
// selection's count on the device; the host does not wait for it.
int rescore(bude_screen *s) {
  if (int st = bude::hip_check(bude::build_rescore_tables(static_cast<const int32_t *>(s->rs_order.ptr), 1,
                                                          static_cast<const uint32_t *>(s->sel_n.ptr),
                                                          static_cast<bude::LigandMeta *>(s->rs_meta.ptr),
                                                          static_cast<int2 *>(s->rs_tiles.ptr),
                                                          static_cast<int32_t *>(s->rs_ntiles.ptr), s->stream),
                               "rescoring tables"))
    return st;
  bude::BatchArgs a{};
  a.protein = static_cast<const float4 *>(s->protein.ptr);
  a.group_begin = static_cast<const int32_t *>(s->group_begin.ptr);
  a.coef = static_cast<const bude::PairCoef *>(s->coef.ptr);
  a.ligands = static_cast<const float4 *>(s->ligand.ptr);
  a.meta = static_cast<const bude::LigandMeta *>(s->rs_meta.ptr);
  a.tiles = static_cast<const int2 *>(s->rs_tiles.ptr);
  a.poses = static_cast<const float *>(s->cand.ptr);
  a.energies = static_cast<float *>(s->exact.ptr);
  a.counter = static_cast<unsigned int *>(s->rs_counter.ptr);
  a.stride = s->ncand;
  a.ngroups = int(s->ngroups);
  a.form = bude::plain_form();
  a.pose_index = static_cast<const uint32_t *>(s->sel_idx.ptr);
  a.ntiles_dev = static_cast<const int32_t *>(s->rs_ntiles.ptr);
  const bude::PinnedGrid g = bude::pinned_grid(s->cus, s->lds_per_cu, 3);
  a.lds_bytes = g.lds_bytes;
  return bude::hip_check(bude::launch_batch_indexed(a, g.blocks, s->stream), "rescore launch");
}

// Surrogate predictions (normalised target units) for n poses whose 6 rows
// start at `poses` with row stride `stride`: the fused inference kernel on a
// persistent grid of 2 work-groups per CU.
int predict_all(bude_screen *s, const float *poses, size_t stride, size_t n, float *out) {
  if (int st = bude::surrogate::predict(s->net, poses, stride, n, out, 2 * s->cus, s->stream)) return st;
  return bude::hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
}

}  // namespace

extern "C" int bude_mi355x_screen_create(int device, const BudeAtom *protein, size_t natpro, const BudeAtom *ligand,
                                         size_t natlig, const BudeFFParams *forcefield, size_t ntypes,
                                         const BudeScreenConfig *cfg, bude_screen **out) {
  if (!out || !cfg) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_create: NULL argument");
  *out = nullptr;
  if (int st = bude::check_deck(protein, natpro, ligand, natlig, forcefield, ntypes, "bude_mi355x_screen_create"))
    return st;
  if (cfg->batch < 32 || cfg->batch % 32 || cfg->batch > (1u << 20) || !(cfg->keep_fraction > 0.f) ||
      cfg->keep_fraction > 1.f || !(cfg->lr > 0.f) || !(cfg->clip_quantile > 0.f) || cfg->clip_quantile > 1.f ||
      cfg->tail_quantile < 0.f || cfg->tail_quantile > 1.f || !(cfg->tail_weight > 0.f) ||
      (cfg->features != 12 && cfg->features != uint32_t(bude::mlp::kFeat)))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      "bude_mi355x_screen_create: batch must be a multiple of 32 in [32, 2^20], keep_fraction and "
                      "clip_quantile in (0, 1], tail_quantile in [0, 1], lr > 0, tail_weight > 0, features 12 or 21");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || device < 0 || device >= ndev)
    return bude::fail(BUDE_ERR_NO_DEVICE, "bude_mi355x_screen_create: no such HIP device");
  if (int st = bude::hip_check(hipSetDevice(device), "hipSetDevice")) return st;

  auto s = std::make_unique<bude_screen>();
  s->device = device;
  s->cfg = *cfg;
  hipDeviceProp_t props{};
  if (int st = bude::hip_check(hipGetDeviceProperties(&props, device), "hipGetDeviceProperties")) return st;
  s->cus = props.multiProcessorCount;
  s->lds_per_cu = props.maxSharedMemoryPerMultiProcessor;
  s->natlig = natlig;
  const bude::ProteinGroups groups = bude::group_protein(protein, natpro);
  s->ngroups = groups.ngroups();
  const std::vector<bude::PairCoef> coef = bude::coef_table(ligand, natlig, groups, forcefield);
  std::vector<float> lig4(natlig * 4);
  for (size_t i = 0; i < natlig; i++) {
    lig4[4 * i + 0] = ligand[i].x;
    lig4[4 * i + 1] = ligand[i].y;
    lig4[4 * i + 2] = ligand[i].z;
  }
  int st = s->protein.upload(groups.xyzw.data(), groups.xyzw.size() * 4);
  if (!st) st = s->group_begin.upload(groups.begin.data(), groups.begin.size() * 4);
  if (!st) st = s->coef.upload(coef.data(), coef.size() * sizeof(bude::PairCoef));
  if (!st) st = s->ligand.upload(lig4.data(), lig4.size() * 4);
  if (!st) st = s->sel_n.alloc(256);
  if (!st) st = bude::hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
  for (auto &e : s->ev)
    if (!st) st = bude::hip_check(hipEventCreate(&e), "hipEventCreate");
  std::vector<std::vector<float>> w, b;
  // (with 12 inputs the first-layer rows of the rotation entries start at
  // zero and, their inputs being zero in training, stay there)
  bude::surrogate::init_weights(cfg->seed, int(cfg->features), w, b);
  if (!st) st = bude::surrogate::set_weights(s->net, w, b, s->stream);
  if (st) return st;
  *out = s.release();
  return bude::ok();
}

extern "C" int bude_mi355x_screen_train(bude_screen *s, const float *const poses[6], size_t n,
                                        BudeScreenTrainStats *stats, float *loss_curve, size_t loss_cap) {
  if (!s || !poses || !stats) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_train: NULL argument");
  const size_t B = s->cfg.batch;
  const size_t nval = std::max<size_t>(32, n / 16);
  if (n < 2 * B + nval || n > 0xffffffffu)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_train: need at least 2 batches + a validation "
                                                 "split (n/16) of poses");
  if (int st = bude::hip_check(hipSetDevice(s->device), "hipSetDevice")) return st;
  *stats = BudeScreenTrainStats{};
  hipStream_t st_ = s->stream;

  // 1. exact labels from the plain kernel
  DeviceBuffer tp, te;
  if (int st = upload_poses(tp, poses, n)) return st;
  if (int st = te.alloc(n * 4)) return st;
  auto t0 = std::chrono::steady_clock::now();
  const bude::PlainArgs pa = s->plain(static_cast<const float *>(tp.ptr), n, n, static_cast<float *>(te.ptr));
  if (int st = bude::hip_check(bude::launch_plain(pa, 256, st_), "label launch")) return st;
  if (int st = bude::hip_check(hipStreamSynchronize(st_), "hipStreamSynchronize")) return st;
  stats->label_ms = ms_since(t0);

  // 2. target transform from the training split: clip at a quantile (the
  //    screen ranks the low tail), standardise
  const size_t ntrain = n - nval;
  std::vector<float> e(n);
  if (int st = bude::hip_check(hipMemcpy(e.data(), te.ptr, n * 4, hipMemcpyDeviceToHost), "hipMemcpy")) return st;
  {
    std::vector<float> sorted(e.begin(), e.begin() + long(ntrain));
    const size_t q = std::min(ntrain - 1, size_t(double(s->cfg.clip_quantile) * double(ntrain)));
    std::nth_element(sorted.begin(), sorted.begin() + long(q), sorted.end());
    s->clip = sorted[q];
    if (s->cfg.tail_quantile > 0.f) {
      const size_t qt = std::min(ntrain - 1, size_t(double(s->cfg.tail_quantile) * double(ntrain)));
      std::nth_element(sorted.begin(), sorted.begin() + long(qt), sorted.end());
      s->tail = sorted[qt];
    }
    double sum = 0, sq = 0;
    for (size_t i = 0; i < ntrain; i++) {
      const double y = std::min(e[i], s->clip);
      sum += y;
      sq += y * y;
    }
    s->mu = float(sum / double(ntrain));
    s->sigma = float(std::sqrt(std::max(1e-12, sq / double(ntrain) - double(s->mu) * s->mu)));
  }

  // 3. Adam over shuffled mini-batches of the training split
  std::vector<uint32_t> perm(ntrain);
  std::iota(perm.begin(), perm.end(), 0u);
  bude::surrogate::TrainData td;
  td.poses = static_cast<const float *>(tp.ptr);
  td.stride = n;
  td.energy = static_cast<const float *>(te.ptr);
  td.norm = bude::mlp::TargetNorm{s->clip, s->mu, 1.0f / s->sigma, s->tail};
  td.nfeat = int(s->cfg.features);
  bude::surrogate::TrainParams tpar;
  tpar.batch = uint32_t(B);
  tpar.epochs = s->cfg.epochs;
  tpar.lr = s->cfg.lr;
  tpar.tail_weight = s->cfg.tail_weight;
  tpar.seed = s->cfg.seed;
  bude::surrogate::TrainReport rep;
  t0 = std::chrono::steady_clock::now();
  if (int st = bude::surrogate::train(s->net, td, std::move(perm), tpar, st_, loss_curve, loss_cap, &rep)) return st;
  stats->final_loss = rep.final_loss;
  stats->train_ms = ms_since(t0);
  stats->steps = rep.steps;

  // 4. validation RMSE (normalised units) on the held-out split
  DeviceBuffer vpred;
  if (int st = vpred.alloc(nval * 4)) return st;
  if (int st = predict_all(s, static_cast<const float *>(tp.ptr) + ntrain, n, nval, static_cast<float *>(vpred.ptr)))
    return st;
  {
    std::vector<float> vp(nval);
    if (int st = bude::hip_check(hipMemcpy(vp.data(), vpred.ptr, nval * 4, hipMemcpyDeviceToHost), "hipMemcpy"))
      return st;
    double se = 0;
    for (size_t i = 0; i < nval; i++) {
      const double y = (double(std::min(e[ntrain + i], s->clip)) - s->mu) / s->sigma;
      se += (vp[i] - y) * (vp[i] - y);
    }
    s->val_rmse = float(std::sqrt(se / double(nval)));
  }
  s->trained = true;
  s->fallback = !(s->val_rmse <= s->cfg.max_val_rmse);
  stats->val_rmse = s->val_rmse;
  stats->clip = s->clip;
  stats->mu = s->mu;
  stats->sigma = s->sigma;
  stats->fallback = s->fallback ? 1 : 0;
  return bude::ok();
}

extern "C" int bude_mi355x_screen_load(bude_screen *s, const float *const poses[6], size_t n) {
  if (!s || !poses || n == 0 || n > 0xffffffffu)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_load: bad arguments");
  if (int st = bude::hip_check(hipSetDevice(s->device), "hipSetDevice")) return st;
  s->ncand = n;
  s->nsel = 0;
  int st = upload_poses(s->cand, poses, n);
  if (!st) st = s->pred.alloc(n * 4);
  if (!st) st = s->sel_idx.alloc(n * 4);
  if (!st) st = s->exact.alloc(n * 4);
  if (!st) st = s->rs_tiles.alloc((n + 63) / 64 * sizeof(int2));
  if (!st) st = s->rs_counter.alloc(sizeof(unsigned int));
  if (!st) st = s->rs_ntiles.alloc(sizeof(int32_t));
  if (!st) {  // the one ligand's row of the rescoring meta (its pose count comes from the selection)
    const bude::LigandMeta m{0, int32_t(s->natlig), 0, 0, 0, 0};
    st = s->rs_meta.assign(&m, sizeof(m));
  }
  if (!st) {
    const int32_t zero = 0;
    st = s->rs_order.assign(&zero, sizeof(zero));
  }
  return st;
}

extern "C" int bude_mi355x_screen_run(bude_screen *s, BudeScreenResult *res) {
  if (!s || !res) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_run: NULL argument");
  if (!s->trained || s->ncand == 0)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_run: train and load poses first");
  if (int st = bude::hip_check(hipSetDevice(s->device), "hipSetDevice")) return st;
  *res = BudeScreenResult{};
  const size_t n = s->ncand;
  const size_t k = std::max<size_t>(1, size_t(std::ceil(double(s->cfg.keep_fraction) * double(n))));
  hipStream_t st_ = s->stream;
  const auto t0 = std::chrono::steady_clock::now();
  if (s->plan_n != n || s->plan_k != k) {  // (uploads before any launch of this run)
    if (int st = s->plan.prepare({bude::sel::Segment{0, uint32_t(n), uint32_t(k)}})) return st;
    s->plan_n = n;
    s->plan_k = k;
  }
  float ms01 = 0, ms12 = 0, ms23 = 0;
  if (s->fallback) {
    // plain kernel on every pose, then the same selection on exact energies
    if (int st = bude::hip_check(hipEventRecord(s->ev[0], st_), "hipEventRecord")) return st;
    const bude::PlainArgs pa = s->plain(static_cast<const float *>(s->cand.ptr), n, n, static_cast<float *>(s->pred.ptr));
    if (int st = bude::hip_check(bude::launch_plain(pa, 256, st_), "fallback launch")) return st;
    if (int st = bude::hip_check(hipEventRecord(s->ev[1], st_), "hipEventRecord")) return st;
  } else {
    if (int st = bude::hip_check(hipEventRecord(s->ev[0], st_), "hipEventRecord")) return st;
    if (int st = bude::surrogate::predict(s->net, static_cast<const float *>(s->cand.ptr), n, n,
                                          static_cast<float *>(s->pred.ptr), 2 * s->cus, st_))
      return st;
    if (int st = bude::hip_check(hipEventRecord(s->ev[1], st_), "hipEventRecord")) return st;
  }
  if (int st = bude::hip_check(bude::sel::select_lowest_segments(s->plan, static_cast<const float *>(s->pred.ptr),
                                                                 static_cast<uint32_t *>(s->sel_idx.ptr),
                                                                 static_cast<uint32_t *>(s->sel_n.ptr), st_),
                               "select"))
    return st;
  if (int st = bude::hip_check(hipEventRecord(s->ev[2], st_), "hipEventRecord")) return st;
  if (!s->fallback)  // (in fallback the exact energies are already in pred)
    if (int st = rescore(s)) return st;
  if (int st = bude::hip_check(hipEventRecord(s->ev[3], st_), "hipEventRecord")) return st;
  uint32_t nsel = 0;
  if (int st = bude::hip_check(hipMemcpyAsync(&nsel, s->sel_n.ptr, 4, hipMemcpyDeviceToHost, st_), "hipMemcpyAsync"))
    return st;
  if (int st = bude::hip_check(hipStreamSynchronize(st_), "hipStreamSynchronize")) return st;
  s->nsel = nsel;
  (void)hipEventElapsedTime(&ms01, s->ev[0], s->ev[1]);
  (void)hipEventElapsedTime(&ms12, s->ev[1], s->ev[2]);
  (void)hipEventElapsedTime(&ms23, s->ev[2], s->ev[3]);
  res->total_ms = ms_since(t0);
  res->score_ms = ms01;
  res->select_ms = ms12;
  res->rescore_ms = ms23;
  res->nposes = n;
  res->nselected = nsel;
  res->fallback = s->fallback ? 1 : 0;
  return bude::ok();
}

extern "C" int bude_mi355x_screen_set_keep(bude_screen *s, float keep_fraction) {
  if (!s || !(keep_fraction > 0.f) || keep_fraction > 1.f)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_set_keep: keep_fraction must be in (0, 1]");
  s->cfg.keep_fraction = keep_fraction;
  return bude::ok();
}

extern "C" int bude_mi355x_screen_selected(bude_screen *s, uint32_t *idx, float *energies, size_t cap, size_t *n) {
  if (!s || !n) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_selected: NULL argument");
  *n = s->nsel;
  if (!idx && !energies) return bude::ok();
  if (cap < s->nsel) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_selected: buffer too small");
  if (int st = bude::hip_check(hipSetDevice(s->device), "hipSetDevice")) return st;
  std::vector<uint32_t> id(s->nsel);
  if (int st = bude::hip_check(hipMemcpy(id.data(), s->sel_idx.ptr, s->nsel * 4, hipMemcpyDeviceToHost), "hipMemcpy"))
    return st;
  std::vector<float> ex(s->nsel);
  if (s->fallback) {
    std::vector<float> all(s->ncand);
    if (int st = bude::hip_check(hipMemcpy(all.data(), s->pred.ptr, s->ncand * 4, hipMemcpyDeviceToHost), "hipMemcpy"))
      return st;
    for (size_t j = 0; j < s->nsel; j++) ex[j] = all[id[j]];
  } else if (s->nsel) {
    if (int st = bude::hip_check(hipMemcpy(ex.data(), s->exact.ptr, s->nsel * 4, hipMemcpyDeviceToHost), "hipMemcpy"))
      return st;
  }
  // report in pose order
  std::vector<size_t> ord(s->nsel);
  std::iota(ord.begin(), ord.end(), size_t(0));
  std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return id[a] < id[b]; });
  for (size_t j = 0; j < s->nsel; j++) {
    if (idx) idx[j] = id[ord[j]];
    if (energies) energies[j] = ex[ord[j]];
  }
  return bude::ok();
}

extern "C" int bude_mi355x_screen_predict(bude_screen *s, float *pred) {
  if (!s || !pred || s->ncand == 0)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_predict: load poses first");
  if (int st = bude::hip_check(hipSetDevice(s->device), "hipSetDevice")) return st;
  DeviceBuffer out;
  if (int st = out.alloc(s->ncand * 4)) return st;
  if (int st = predict_all(s, static_cast<const float *>(s->cand.ptr), s->ncand, s->ncand,
                           static_cast<float *>(out.ptr)))
    return st;
  std::vector<float> p(s->ncand);
  if (int st = bude::hip_check(hipMemcpy(p.data(), out.ptr, s->ncand * 4, hipMemcpyDeviceToHost), "hipMemcpy"))
    return st;
  for (size_t i = 0; i < s->ncand; i++) pred[i] = p[i] * s->sigma + s->mu;  // back to energy units (clipped scale)
  return bude::ok();
}

extern "C" int bude_mi355x_screen_exhaustive(bude_screen *s, float *energies, double *ms) {
  if (!s || s->ncand == 0) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_screen_exhaustive: no poses");
  if (int st = bude::hip_check(hipSetDevice(s->device), "hipSetDevice")) return st;
  DeviceBuffer out;
  if (int st = out.alloc(s->ncand * 4)) return st;
  const bude::PlainArgs pa = s->plain(static_cast<const float *>(s->cand.ptr), s->ncand, s->ncand,
                                      static_cast<float *>(out.ptr));
  if (int st = bude::hip_check(hipEventRecord(s->ev[0], s->stream), "hipEventRecord")) return st;
  if (int st = bude::hip_check(bude::launch_plain(pa, 256, s->stream), "exhaustive launch")) return st;
  if (int st = bude::hip_check(hipEventRecord(s->ev[1], s->stream), "hipEventRecord")) return st;
  if (int st = bude::hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize")) return st;
  float t = 0.f;
  (void)hipEventElapsedTime(&t, s->ev[0], s->ev[1]);
  if (ms) *ms = t;
  if (energies)
    return bude::hip_check(hipMemcpy(energies, out.ptr, s->ncand * 4, hipMemcpyDeviceToHost), "hipMemcpy");
  return bude::ok();
}

extern "C" int bude_mi355x_screen_destroy(bude_screen *s) {
  if (!s) return bude::ok();
  (void)hipSetDevice(s->device);
  delete s;
  return bude::ok();
}

extern "C" int bude_mi355x_mlp_forward_backward(int device, size_t batch, const float *features, const float *target,
                                                const float *const weights[4], const float *const biases[4],
                                                float *pred, float *loss, float *const grad_w[4],
                                                float *const grad_b[4]) {
  if (!features || !target || !weights || !biases || !pred || !loss || !grad_w || !grad_b || batch == 0 ||
      batch % 32 || batch > (1u << 16))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      "bude_mi355x_mlp_forward_backward: bad arguments (batch a multiple of 32, <= 65536)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || device < 0 || device >= ndev)
    return bude::fail(BUDE_ERR_NO_DEVICE, "bude_mi355x_mlp_forward_backward: no such HIP device");
  if (int st = bude::hip_check(hipSetDevice(device), "hipSetDevice")) return st;
  hipStream_t s = nullptr;
  if (int st = bude::hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate")) return st;
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  // real shapes: 21x256, 256x256, 256x256, 256x1 -> padded device layout
  const int realK[4] = {bude::mlp::kFeat, 256, 256, 256}, realN[4] = {256, 256, 256, 1};
  std::vector<std::vector<float>> w(4), b(4);
  for (int l = 0; l < 4; l++) {
    w[l].assign(size_t(kK[l]) * kN[l], 0.f);
    b[l].assign(size_t(kN[l]), 0.f);
    for (int r = 0; r < realK[l]; r++)
      for (int c = 0; c < realN[l]; c++) w[l][size_t(r) * kN[l] + c] = weights[l][size_t(r) * realN[l] + c];
    for (int c = 0; c < realN[l]; c++) b[l][c] = biases[l][c];
  }
  Net net;
  if (int st = bude::surrogate::set_weights(net, w, b, s)) return st;
  Acts A;
  if (int st = bude::surrogate::alloc_acts(A, batch, true)) return st;
  std::vector<float> fpad(batch * bude::mlp::kIn, 0.f);
  for (size_t i = 0; i < batch; i++)
    for (int c = 0; c < bude::mlp::kFeat; c++) fpad[i * bude::mlp::kIn + c] = features[i * bude::mlp::kFeat + c];
  DeviceBuffer f32, tgt;
  if (int st = f32.upload(fpad.data(), fpad.size() * 4)) return st;
  if (int st = tgt.upload(target, batch * 4)) return st;
  if (int st = bude::hip_check(bude::mlp::to_bf16(static_cast<const float *>(f32.ptr), int(batch), bude::mlp::kIn,
                                                  static_cast<bf16 *>(A.feat.ptr), static_cast<bf16 *>(A.feat_t.ptr),
                                                  s, A.ld_t, true),
                               "to_bf16"))
    return st;
  if (int st = bude::hip_check(hipMemsetAsync(A.loss.ptr, 0, 8, s), "hipMemsetAsync")) return st;
  bude::mlp::RowsArgs in{};
  in.feat = static_cast<const bf16 *>(A.feat.ptr);
  in.energy = static_cast<const float *>(tgt.ptr);
  in.norm = bude::mlp::TargetNorm{3.0e38f, 0.f, 1.f, -3.0e38f};
  in.tail_weight = 1.f;
  in.pred = static_cast<float *>(A.pred.ptr);
  if (int st = bude::surrogate::rows_step(net, A, batch, in, s)) return st;
  if (int st = bude::surrogate::weight_grads(net, A, batch, s)) return st;
  if (int st = bude::surrogate::sum_grads(net, A, s)) return st;
  if (int st = bude::hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")) return st;
  if (int st = bude::hip_check(hipMemcpy(pred, A.pred.ptr, batch * 4, hipMemcpyDeviceToHost), "hipMemcpy")) return st;
  unsigned long long lsum = 0;
  if (int st = bude::hip_check(hipMemcpy(&lsum, A.loss.ptr, 8, hipMemcpyDeviceToHost), "hipMemcpy")) return st;
  *loss = float(double(lsum) / bude::mlp::kLossScale / double(batch));
  for (int l = 0; l < 4; l++) {
    std::vector<float> gw(size_t(kK[l]) * kN[l]);
    std::vector<float> gb(static_cast<size_t>(kN[l]));
    if (int st = bude::hip_check(hipMemcpy(gw.data(), net.gw[l].ptr, gw.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy"))
      return st;
    if (int st = bude::hip_check(hipMemcpy(gb.data(), net.gb[l].ptr, gb.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy"))
      return st;
    for (int r = 0; r < realK[l]; r++)
      for (int c = 0; c < realN[l]; c++) grad_w[l][size_t(r) * realN[l] + c] = gw[size_t(r) * kN[l] + c];
    for (int c = 0; c < realN[l]; c++) grad_b[l][c] = gb[c];
  }
  return bude::ok();
}

extern "C" int bude_mi355x_mlp_predict(int device, const float *const poses[6], size_t n,
                                       const float *const weights[4], const float *const biases[4], float *pred) {
  if (!poses || !weights || !biases || !pred || n == 0 || n > 0xffffffffu)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_mlp_predict: bad arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || device < 0 || device >= ndev)
    return bude::fail(BUDE_ERR_NO_DEVICE, "bude_mi355x_mlp_predict: no such HIP device");
  if (int st = bude::hip_check(hipSetDevice(device), "hipSetDevice")) return st;
  hipStream_t s = nullptr;
  if (int st = bude::hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate")) return st;
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{s};
  const int realK[4] = {bude::mlp::kFeat, 256, 256, 256}, realN[4] = {256, 256, 256, 1};
  std::vector<std::vector<float>> w(4), b(4);
  for (int l = 0; l < 4; l++) {
    w[l].assign(size_t(kK[l]) * kN[l], 0.f);
    b[l].assign(size_t(kN[l]), 0.f);
    for (int r = 0; r < realK[l]; r++)
      for (int c = 0; c < realN[l]; c++) w[l][size_t(r) * kN[l] + c] = weights[l][size_t(r) * realN[l] + c];
    for (int c = 0; c < realN[l]; c++) b[l][c] = biases[l][c];
  }
  Net net;
  if (int st = bude::surrogate::set_weights(net, w, b, s)) return st;
  DeviceBuffer dp, out;
  if (int st = upload_poses(dp, poses, n)) return st;
  if (int st = out.alloc(n * 4)) return st;
  hipDeviceProp_t props{};
  if (int st = bude::hip_check(hipGetDeviceProperties(&props, device), "hipGetDeviceProperties")) return st;
  const bf16 *wt[4] = {static_cast<const bf16 *>(net.w16tx[0].ptr), static_cast<const bf16 *>(net.w16tx[1].ptr),
                       static_cast<const bf16 *>(net.w16tx[2].ptr), nullptr};
  const float *bias[4] = {static_cast<const float *>(net.b[0].ptr), static_cast<const float *>(net.b[1].ptr),
                          static_cast<const float *>(net.b[2].ptr), static_cast<const float *>(net.b[3].ptr)};
  if (int st = bude::hip_check(bude::mlp::infer(static_cast<const float *>(dp.ptr), n, n, wt,
                                                static_cast<const bf16 *>(net.w16[3].ptr), bias,
                                                static_cast<float *>(out.ptr), 2 * props.multiProcessorCount, s),
                               "mlp infer"))
    return st;
  if (int st = bude::hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")) return st;
  return bude::hip_check(hipMemcpy(pred, out.ptr, n * 4, hipMemcpyDeviceToHost), "hipMemcpy");
}
