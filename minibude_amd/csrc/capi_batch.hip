// capi_batch.hip -- C ABI of the batched ragged fasten_main: one protein,
// many ligands with their own atom and pose counts, scored in one persistent
// launch replayed from a captured hipGraph; plus the one-launch-per-ligand
// baseline and the batch autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "bude_internal.h"
#include "fasten_batch.h"
#include "fasten_plain.h"
#include "hip_util.h"

namespace {

struct GraphExec {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~GraphExec() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

}  // namespace

struct bude_batch {
  int device = 0;
  int cus = 0;
  size_t natpro = 0, ngroups = 0, nligands = 0, total_poses = 0;
  std::vector<int32_t> natlig;
  std::vector<size_t> nposes, pose_off;
  std::vector<bude::LigandMeta> meta_host;
  bude::DeviceBuffer protein, group_begin, coef, ligands, meta, poses, energies, counter;
  // tile lists per (ppwi, order)
  std::map<std::pair<int, int>, std::unique_ptr<bude::DeviceBuffer>> tiles;
  std::map<std::pair<int, int>, int> ntiles;
  std::map<std::tuple<int, int, int, int>, std::unique_ptr<GraphExec>> graphs;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  ~bude_batch() {
    graphs.clear();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int check_config(const BudeBatchConfig *cfg) {
  if (!cfg) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "batch: config is NULL");
  if (!bude::batch_shape_supported(cfg->ppwi, cfg->wgsize))
    return bude::fail(BUDE_ERR_UNSUPPORTED, "batch: ppwi must be 1, 2, 4 or 8 and wgsize 64, 128, 256 or 512 (got " +
                                                std::to_string(cfg->ppwi) + ", " + std::to_string(cfg->wgsize) + ")");
  if (cfg->waves_per_simd < 1 || cfg->waves_per_simd > 16)
    return bude::fail(BUDE_ERR_UNSUPPORTED, "batch: waves_per_simd must be in [1, 16]");
  if (cfg->order != 0 && cfg->order != 1) return bude::fail(BUDE_ERR_UNSUPPORTED, "batch: order must be 0 or 1");
  return BUDE_OK;
}

// Tiles of 64*ppwi poses; order 0 = heaviest ligand first (cost ~ natlig),
// order 1 = ligand order.
int tiles_for(bude_batch *b, int ppwi, int order, const bude::DeviceBuffer **buf, int *count) {
  const auto key = std::make_pair(ppwi, order);
  auto it = b->tiles.find(key);
  if (it == b->tiles.end()) {
    const size_t span = size_t(64) * size_t(ppwi);
    std::vector<size_t> lig(b->nligands);
    std::iota(lig.begin(), lig.end(), size_t(0));
    if (order == 0)
      std::stable_sort(lig.begin(), lig.end(), [&](size_t x, size_t y) { return b->natlig[x] > b->natlig[y]; });
    std::vector<int2> t;
    for (size_t i : lig)
      for (size_t j = 0; j < b->nposes[i]; j += span) t.push_back(make_int2(int(i), int(j)));
    if (t.size() > 0x7fffffffu) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "batch: too many tiles");
    auto d = std::make_unique<bude::DeviceBuffer>();
    if (int st = d->upload(t.data(), t.size() * sizeof(int2))) return st;
    b->ntiles[key] = int(t.size());
    it = b->tiles.emplace(key, std::move(d)).first;
  }
  *buf = it->second.get();
  *count = b->ntiles[key];
  return BUDE_OK;
}

int launch_once(bude_batch *b, const BudeBatchConfig *cfg) {
  const bude::DeviceBuffer *tb = nullptr;
  int nt = 0;
  if (int st = tiles_for(b, cfg->ppwi, cfg->order, &tb, &nt)) return st;
  bude::BatchArgs a{};
  a.protein = static_cast<const float4 *>(b->protein.ptr);
  a.group_begin = static_cast<const int32_t *>(b->group_begin.ptr);
  a.coef = static_cast<const bude::PairCoef *>(b->coef.ptr);
  a.ligands = static_cast<const float4 *>(b->ligands.ptr);
  a.meta = static_cast<const bude::LigandMeta *>(b->meta.ptr);
  a.tiles = static_cast<const int2 *>(tb->ptr);
  a.poses = static_cast<const float *>(b->poses.ptr);
  a.energies = static_cast<float *>(b->energies.ptr);
  a.counter = static_cast<unsigned int *>(b->counter.ptr);
  a.stride = b->total_poses;
  a.ngroups = int(b->ngroups);
  a.ntiles = nt;
  a.form = bude::plain_form();
  const int waves = b->cus * 4 * cfg->waves_per_simd;
  const int blocks = std::max(1, waves / (cfg->wgsize / 64));
  return bude::hip_check(bude::launch_batch(a, cfg->ppwi, cfg->wgsize, blocks, b->stream), "fasten_batch launch");
}

int graph_for(bude_batch *b, const BudeBatchConfig *cfg, hipGraphExec_t *exec) {
  const auto key = std::make_tuple(cfg->ppwi, cfg->wgsize, cfg->waves_per_simd, cfg->order);
  auto it = b->graphs.find(key);
  if (it == b->graphs.end()) {
    // build tiles before capture (the upload is not capturable)
    const bude::DeviceBuffer *tb = nullptr;
    int nt = 0;
    if (int st = tiles_for(b, cfg->ppwi, cfg->order, &tb, &nt)) return st;
    auto g = std::make_unique<GraphExec>();
    if (int st = bude::hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal),
                                 "hipStreamBeginCapture"))
      return st;
    const int lst = launch_once(b, cfg);
    const hipError_t ce = hipStreamEndCapture(b->stream, &g->graph);  // g owns (and frees) it on every path
    if (lst) return lst;
    if (int st = bude::hip_check(ce, "hipStreamEndCapture")) return st;
    if (int st = bude::hip_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate"))
      return st;
    it = b->graphs.emplace(key, std::move(g)).first;
  }
  *exec = it->second->exec;
  return BUDE_OK;
}

template <typename Fn> int timed(bude_batch *b, size_t warmup, size_t iterations, double *device_ms, Fn &&fn) {
  for (size_t i = 0; i < warmup + iterations; i++) {
    if (int st = bude::hip_check(hipEventRecord(b->ev0, b->stream), "hipEventRecord")) return st;
    if (int st = fn()) return st;
    if (int st = bude::hip_check(hipEventRecord(b->ev1, b->stream), "hipEventRecord")) return st;
    if (int st = bude::hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize")) return st;
    if (device_ms) {
      float ms = 0.f;
      if (int st = bude::hip_check(hipEventElapsedTime(&ms, b->ev0, b->ev1), "hipEventElapsedTime")) return st;
      device_ms[i] = ms;
    }
  }
  return BUDE_OK;
}

}  // namespace

extern "C" int bude_mi355x_batch_create(int device, const BudeAtom *protein, size_t natpro,
                                        const BudeFFParams *forcefield, size_t ntypes, size_t nligands,
                                        const BudeAtom *const *ligands, const size_t *natlig,
                                        const float *const *poses, const size_t *nposes, bude_batch **out) {
  if (!out) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_batch_create: out is NULL");
  *out = nullptr;
  if (nligands == 0 || !ligands || !natlig || !poses || !nposes)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_batch_create: empty batch");
  for (size_t i = 0; i < nligands; i++) {
    const std::string who = "bude_mi355x_batch_create: ligand " + std::to_string(i);
    if (int st = bude::check_deck(protein, natpro, ligands[i], natlig[i], forcefield, ntypes, who.c_str())) return st;
    if (!poses[i] || nposes[i] == 0 || nposes[i] > 0x7fffffffu)
      return bude::fail(BUDE_ERR_INVALID_ARGUMENT, who + ": no poses (or more than 2^31)");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return bude::fail(BUDE_ERR_NO_DEVICE, "bude_mi355x_batch_create: no HIP device visible");
  if (device < 0 || device >= ndev)
    return bude::fail(BUDE_ERR_NO_DEVICE, "bude_mi355x_batch_create: device out of range");
  if (int st = bude::hip_check(hipSetDevice(device), "hipSetDevice")) return st;

  auto b = std::make_unique<bude_batch>();
  b->device = device;
  hipDeviceProp_t props{};
  if (int st = bude::hip_check(hipGetDeviceProperties(&props, device), "hipGetDeviceProperties")) return st;
  b->cus = props.multiProcessorCount;
  b->natpro = natpro;
  b->nligands = nligands;
  const bude::ProteinGroups groups = bude::group_protein(protein, natpro);
  b->ngroups = groups.ngroups();

  std::vector<float> lig4;
  std::vector<bude::PairCoef> coef;
  size_t total = 0;
  for (size_t i = 0; i < nligands; i++) {
    bude::LigandMeta m{};
    m.lig_off = int32_t(lig4.size() / 4);
    m.natlig = int32_t(natlig[i]);
    m.coef_off = int32_t(coef.size());
    m.nposes = int32_t(nposes[i]);
    m.pose_off = int64_t(total);
    for (size_t a = 0; a < natlig[i]; a++) {
      const BudeAtom &at = ligands[i][a];
      lig4.insert(lig4.end(), {at.x, at.y, at.z, 0.f});
    }
    const std::vector<bude::PairCoef> c = bude::coef_table(ligands[i], natlig[i], groups, forcefield);
    coef.insert(coef.end(), c.begin(), c.end());
    if (coef.size() > 0x7fffffffu || lig4.size() / 4 > 0x7fffffffu)
      return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_batch_create: batch too large");
    b->meta_host.push_back(m);
    b->natlig.push_back(int32_t(natlig[i]));
    b->nposes.push_back(nposes[i]);
    b->pose_off.push_back(total);
    total += nposes[i];
  }
  b->total_poses = total;

  int st = BUDE_OK;
  if (!st) st = b->protein.upload(groups.xyzw.data(), groups.xyzw.size() * sizeof(float));
  if (!st) st = b->group_begin.upload(groups.begin.data(), groups.begin.size() * sizeof(int32_t));
  if (!st) st = b->coef.upload(coef.data(), coef.size() * sizeof(bude::PairCoef));
  if (!st) st = b->ligands.upload(lig4.data(), lig4.size() * sizeof(float));
  if (!st) st = b->meta.upload(b->meta_host.data(), b->meta_host.size() * sizeof(bude::LigandMeta));
  if (!st) st = b->poses.alloc(6 * total * sizeof(float));
  for (size_t i = 0; i < nligands && !st; i++)
    for (int k = 0; k < 6 && !st; k++)
      st = bude::hip_check(hipMemcpy(static_cast<float *>(b->poses.ptr) + size_t(k) * total + b->pose_off[i],
                                     poses[i] + size_t(k) * nposes[i], nposes[i] * sizeof(float),
                                     hipMemcpyHostToDevice),
                           "hipMemcpy poses");
  if (!st) st = b->energies.alloc(total * sizeof(float));
  if (!st) st = b->counter.alloc(256);
  if (!st) st = bude::hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
  if (!st) st = bude::hip_check(hipEventCreate(&b->ev0), "hipEventCreate");
  if (!st) st = bude::hip_check(hipEventCreate(&b->ev1), "hipEventCreate");
  if (!st) st = bude::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  if (st) return st;
  *out = b.release();
  return bude::ok();
}

extern "C" int bude_mi355x_batch_run(bude_batch *b, const BudeBatchConfig *cfg, size_t warmup, size_t iterations,
                                     double *device_ms) {
  if (!b) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_batch_run: batch is NULL");
  if (int st = check_config(cfg)) return st;
  if (int st = bude::hip_check(hipSetDevice(b->device), "hipSetDevice")) return st;
  if (cfg->use_graph) {
    hipGraphExec_t exec = nullptr;
    if (int st = graph_for(b, cfg, &exec)) return st;
    return timed(b, warmup, iterations, device_ms,
                 [&] { return bude::hip_check(hipGraphLaunch(exec, b->stream), "hipGraphLaunch"); });
  }
  return timed(b, warmup, iterations, device_ms, [&] { return launch_once(b, cfg); });
}

extern "C" int bude_mi355x_batch_run_separate(bude_batch *b, size_t ppwi, size_t wgsize, size_t warmup,
                                              size_t iterations, double *device_ms) {
  if (!b) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_batch_run_separate: batch is NULL");
  if (!(ppwi == 1 || ppwi == 2 || ppwi == 4 || ppwi == 8 || ppwi == 16 || ppwi == 32 || ppwi == 64 || ppwi == 128) ||
      !(wgsize == 64 || wgsize == 128 || wgsize == 256 || wgsize == 512 || wgsize == 1024))
    return bude::fail(BUDE_ERR_UNSUPPORTED, "bude_mi355x_batch_run_separate: unsupported ppwi/wgsize");
  if (int st = bude::hip_check(hipSetDevice(b->device), "hipSetDevice")) return st;
  return timed(b, warmup, iterations, device_ms, [&]() -> int {
    for (size_t i = 0; i < b->nligands; i++) {
      const bude::LigandMeta &m = b->meta_host[i];
      bude::PlainArgs a{};
      a.protein = static_cast<const float4 *>(b->protein.ptr);
      a.group_begin = static_cast<const int32_t *>(b->group_begin.ptr);
      a.coef = static_cast<const bude::PairCoef *>(b->coef.ptr) + m.coef_off;
      a.ligand = static_cast<const float4 *>(b->ligands.ptr) + m.lig_off;
      a.poses = static_cast<const float *>(b->poses.ptr) + m.pose_off;
      a.energies = static_cast<float *>(b->energies.ptr) + m.pose_off;
      a.stride = b->total_poses;
      a.nposes = size_t(m.nposes);
      a.natlig = m.natlig;
      a.ngroups = int(b->ngroups);
      a.ppwi = int(ppwi);
      a.form = bude::plain_form();
      if (int st = bude::hip_check(bude::launch_plain(a, int(wgsize), b->stream), "fasten_plain launch")) return st;
    }
    return BUDE_OK;
  });
}

extern "C" int bude_mi355x_batch_autotune(bude_batch *b, BudeBatchConfig *best, double *best_ms) {
  if (!b || !best) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_batch_autotune: NULL argument");
  // Candidate tile widths, work-group sizes, persistent-grid depths and tile
  // orders, by successive halving: every candidate's graph is built and
  // replayed twice (the second replay timed), then the four fastest are
  // timed as the median of three more replays each.
  std::vector<std::pair<double, BudeBatchConfig>> round1;
  for (int ppwi : {1, 2, 4})
    for (int wg : {64, 256})
      for (int wps : {4, 8})
        for (int order : {0, 1}) {
          BudeBatchConfig c{ppwi, wg, wps, order, 1};
          double ms[2] = {0, 0};
          if (int st = bude_mi355x_batch_run(b, &c, 1, 1, ms)) return st;
          round1.emplace_back(ms[1], c);
        }
  std::stable_sort(round1.begin(), round1.end(),
                   [](const auto &x, const auto &y) { return x.first < y.first; });
  double best_t = 1e300;
  BudeBatchConfig winner{};
  for (size_t i = 0; i < round1.size() && i < 4; i++) {
    double ms[3] = {0, 0, 0};
    if (int st = bude_mi355x_batch_run(b, &round1[i].second, 0, 3, ms)) return st;
    std::sort(ms, ms + 3);
    if (ms[1] < best_t) {
      best_t = ms[1];
      winner = round1[i].second;
    }
  }
  *best = winner;
  if (best_ms) *best_ms = best_t;
  return bude::ok();
}

extern "C" int bude_mi355x_batch_energies(bude_batch *b, size_t ligand, float *out) {
  if (!b || !out || ligand >= b->nligands)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_batch_energies: bad arguments");
  if (int st = bude::hip_check(hipSetDevice(b->device), "hipSetDevice")) return st;
  return bude::hip_check(hipMemcpy(out, static_cast<float *>(b->energies.ptr) + b->pose_off[ligand],
                                   b->nposes[ligand] * sizeof(float), hipMemcpyDeviceToHost),
                         "hipMemcpy energies");
}

extern "C" int bude_mi355x_batch_clear(bude_batch *b) {
  if (!b) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_batch_clear: batch is NULL");
  if (int st = bude::hip_check(hipSetDevice(b->device), "hipSetDevice")) return st;
  return bude::hip_check(hipMemset(b->energies.ptr, 0xff, b->total_poses * sizeof(float)), "hipMemset");
}

extern "C" int bude_mi355x_batch_destroy(bude_batch *b) {
  if (!b) return bude::ok();
  (void)hipSetDevice(b->device);
  delete b;
  return bude::ok();
}
