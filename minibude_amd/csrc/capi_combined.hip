// capi_combined.hip -- C ABI of the combined screen: a fixed set of ligands,
// each with its own pose set, screened against one protein after another.
// Per protein: exact energies of a few sampled poses of every (protein,
// ligand) pair (one ragged launch), a surrogate (the per-protein MLP or the
// protein's single-atom energy maps) validated per pair on them, the
// surrogate over every pose of every pair, per-pair top-k of its scores, and
// exact rescoring of every pair's selection (with the maps, chunk by chunk of
// ligands, each chunk's rescoring beside the next chunk's scoring); pairs
// whose validation error is too large are scored exhaustively instead.  A
// batch of proteins runs each protein's scoring beside the previous one's
// rescoring, in two protein slots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "atomnet.h"
#include "bude_internal.h"
#include "fasten_batch.h"
#include "field.h"
#include "hip_util.h"
#include "mlp.h"
#include "select.h"
#include "surrogate.h"

namespace {

using bude::DeviceBuffer;

// Persistent batched launch shape used for every exact pass: tiles of 64
// poses on a grid pinned to the same number of waves on every SIMD.  The
// depth follows the tile count: a pass of fewer than 16 tiles per SIMD (the
// rescoring of a few per cent of the poses) runs 3 waves per SIMD, so every
// wave pulls about two tiles from the heaviest-first queue and the SIMDs
// finish together; larger passes run 4, enough to keep the VALU busy
// (measured on ragged 64-ligand sets: 12.6 vs 15.3 ms at 8 waves per SIMD
// for 5,243 tiles, 136.2 vs 137.3 ms for 65,536).
constexpr int kBatchPpwi = 1;

int batch_waves_per_simd(int64_t ntiles, int cus) { return ntiles < 16 * 4 * int64_t(cus) ? 3 : 4; }

// Field pipeline depth: at most 16 chunks, and (chosen automatically) each
// expected to give at least 64 rescoring tiles per CU, so that every chunk's
// rescoring launch still fills the machine; smaller screens run as one chunk.
constexpr size_t kMaxChunks = 16;
constexpr int64_t kChunkTilesPerCu = 64;

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

__global__ void gather_index_kernel(const float *__restrict__ src, size_t src_stride, const uint32_t *__restrict__ idx,
                                    size_t n, float *__restrict__ dst, size_t dst_stride) {
  const size_t j = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const size_t q = idx[j];
#pragma unroll
  for (int k = 0; k < 6; k++) dst[size_t(k) * dst_stride + j] = src[size_t(k) * src_stride + q];
}

unsigned blocks_for(size_t n, unsigned t) { return unsigned((n + t - 1) / t); }

// Pair coefficients of every ligand atom against every protein group:
// coef[a * ngroups + g] = slot_coef[slot(a) * ngroups + g], the atom's map
// slot in the w bits of its float4 (a coefficient depends on the ligand
// atom's type and the group's type only).
__global__ __launch_bounds__(256) void expand_coef_kernel(const float4 *__restrict__ ligands, size_t n, int ngroups,
                                                          const bude::PairCoef *__restrict__ slot_coef,
                                                          bude::PairCoef *__restrict__ coef) {
  const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const size_t a = i / size_t(ngroups), g = i - a * size_t(ngroups);
  const int slot = __float_as_int(ligands[a].w);
  coef[i] = slot_coef[size_t(slot) * size_t(ngroups) + g];
}

// Eleven numbers describing a ligand to the per-protein MLP: atom count,
// radius of gyration, the normalised second-moment tensor (6) in the
// ligand's own frame (with the rotation-matrix inputs this orients the
// ligand's shape), and the centroid.
std::vector<float> descriptor(const BudeAtom *a, size_t n) {
  double c[3] = {0, 0, 0};
  for (size_t i = 0; i < n; i++) {
    c[0] += a[i].x;
    c[1] += a[i].y;
    c[2] += a[i].z;
  }
  for (double &v : c) v /= double(n);
  double m[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy xz yz
  for (size_t i = 0; i < n; i++) {
    const double x = a[i].x - c[0], y = a[i].y - c[1], z = a[i].z - c[2];
    m[0] += x * x;
    m[1] += y * y;
    m[2] += z * z;
    m[3] += x * y;
    m[4] += x * z;
    m[5] += y * z;
  }
  for (double &v : m) v /= double(n);
  const double rg2 = std::max(1e-12, m[0] + m[1] + m[2]);
  std::vector<float> d(bude::mlp::kDesc, 0.f);
  d[0] = float(double(n) / 96.0);
  d[1] = float(std::sqrt(rg2) / 5.0);
  for (int k = 0; k < 6; k++) d[2 + k] = float(m[k] / rg2);
  for (int k = 0; k < 3; k++) d[8 + k] = float(c[k] / 5.0);
  return d;
}

struct Norm {
  float clip = 0.f, mu = 0.f, sigma = 1.f, tail = -3.0e38f;
};

// Target transform of one pair from its training energies (as the learned
// prescreen: clip at a quantile, standardise, tail threshold).
Norm norm_of(std::vector<float> e, float clip_q, float tail_q) {
  Norm r;
  if (e.empty()) return r;
  const size_t n = e.size();
  const size_t q = std::min(n - 1, size_t(double(clip_q) * double(n)));
  std::nth_element(e.begin(), e.begin() + long(q), e.end());
  r.clip = e[q];
  if (tail_q > 0.f) {
    const size_t qt = std::min(n - 1, size_t(double(tail_q) * double(n)));
    std::nth_element(e.begin(), e.begin() + long(qt), e.end());
    r.tail = e[qt];
  }
  double sum = 0, sq = 0;
  for (float v : e) {
    const double y = std::min(v, r.clip);
    sum += y;
    sq += y * y;
  }
  r.mu = float(sum / double(n));
  r.sigma = float(std::sqrt(std::max(1e-12, sq / double(n) - double(r.mu) * r.mu)));
  return r;
}

}  // namespace

// One protein's screen: its tables, selection, exact energies of the
// selection, per-pair results, sampled poses and events.
struct ProteinSlot {
  size_t ngroups = 0;
  std::vector<int32_t> coef_off;
  DeviceBuffer protein, group_begin, coef, slot_coef;
  // exact_sel[pose_off[l] + j] is the exact energy of pose
  // sel_idx[pose_off[l] + j] of pair l (j < nsel[l]); meta_rs / tiles_rs /
  // rs_ntiles the rescoring tables (tiles of each chunk, counts per chunk
  // then the whole set's); rs_sel the non-fallback ligands when the whole
  // set is rescored at once
  DeviceBuffer sel_idx, sel_n, exact_sel, meta_rs, tiles_rs, rs_ntiles, rs_sel;
  std::vector<hipEvent_t> chunk_ev;  // 2 per chunk: scores done, tables done
  hipEvent_t ev[12] = {};
  std::vector<uint32_t> nsel;
  std::vector<int32_t> fallback;
  std::vector<float> val_rmse, val_rmse_cal, gate_z, sample_recall, keep;
  // the selection of a protein with widened pairs: the object's segments
  // with this protein's per-pair k
  bude::sel::Plan wplan;
  float atom_val_rmse = 0.f;  // learned per-atom field: one-atom RMSE on held-out points (energy units)
  // the sampled poses: pose index within the pair, exact energy, per-pair offsets
  std::vector<uint32_t> lab_j;
  std::vector<float> lab_energy;
  std::vector<size_t> lab_offs;
  uint64_t protein_sum = 0;
  // begun and not yet ended; how its selection and rescoring run
  // overlap: the next protein of a batch may begin before this one ends (its
  // selection ran chunk by chunk, or as one pass over the scores that the
  // next protein's scoring waits for: ev[10]; not with fallback pairs, whose
  // exhaustive pass writes shared tables, or the pose MLP)
  bool busy = false, valid = false, pipelined = false, mlp = false, overlap = false;
  uint32_t *nsel_host = nullptr;  // pinned: the counts' copy lands here without blocking the host
  // pinned staging of the protein's uploads (a synchronous copy would wait
  // for the previous protein's kernels on the other streams)
  char *stage = nullptr;
  size_t stage_cap = 0, stage_off = 0;
  size_t K = 0;
  BudeCombinedResult res{};
  std::chrono::steady_clock::time_point t0;

  ProteinSlot() = default;
  ProteinSlot(const ProteinSlot &) = delete;
  ProteinSlot &operator=(const ProteinSlot &) = delete;
  ~ProteinSlot() {
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto &e : chunk_ev)
      if (e) (void)hipEventDestroy(e);
    if (nsel_host) (void)hipHostFree(nsel_host);
    if (stage) (void)hipHostFree(stage);
  }
};

struct bude_combined {
  int device = 0, cus = 0;
  size_t lds_per_cu = 0;
  BudeCombinedConfig cfg{};
  std::vector<BudeFFParams> ff;
  // ligands (fixed for the object's life)
  size_t nlig = 0, total = 0;
  std::vector<std::vector<BudeAtom>> lig;
  std::vector<int32_t> lig_off;
  std::vector<size_t> nposes, pose_off;
  std::vector<int32_t> slot_of_type;  // forcefield type -> map slot (-1: unused)
  std::vector<int32_t> slot_type;     // map slot -> forcefield type
  DeviceBuffer ligands;               // float4: x, y, z, slot bits
  // the field scorer's atoms, grouped by map slot for the XCDs
  // (field::split_atoms), the groups' bounds, and partial-sum scratch of the
  // chunk scores (second stream) and the sampled poses' scores (third)
  int field_groups = 1;
  int64_t field_launch_tiles = 1;  // tiles per field launch when field_groups > 1
  DeviceBuffer fgrouped, fgroup_begin, fpart, fpart_lab;
  DeviceBuffer poses;                 // 6 rows of `total`
  DeviceBuffer pred, exact_all;
  DeviceBuffer desc;                  // [nlig][kDesc]
  bude::sel::Plan plan;               // every pair
  // The field screen runs as a pipeline over ligand ranges
  // [chunk_lig[k], chunk_lig[k + 1]): the second stream scores, selects and
  // builds the rescoring tables of chunk k while the first rescores chunk
  // k - 1.  Per chunk: its selection plan, the first slot of its rescoring
  // tiles (of Σ ceil(n / 64) over its pairs), the tiles its selection is
  // expected to give, and two events (scores done, tables done).
  std::vector<size_t> chunk_lig;
  std::vector<std::unique_ptr<bude::sel::Plan>> chunk_plan;
  std::vector<int64_t> chunk_tile;
  std::vector<int64_t> chunk_expect;
  std::vector<int64_t> lig_expect;    // expected rescoring tiles of each pair: ceil(k / 64)
  bool plan_ok = false;               // the selection tables above are complete
  bude::field::Grid grid{};
  // the screen's proteins: the one screened last is slots[cur]; a batch
  // begins the next protein in the other slot while this one is rescored
  ProteinSlot slots[2];
  int cur = 0;
  ProteinSlot &s() { return slots[cur]; }
  const ProteinSlot &s() const { return slots[cur]; }
  DeviceBuffer maps, cells, counter;
  // per-call work tables (each its own buffer; uploaded only while no queued
  // kernel reads it)
  DeviceBuffer meta_label, tiles_label, meta_lf, meta_fb, tiles_fb, meta_all, meta_ex, tiles_all;
  DeviceBuffer tb_label, tb_all, lab_src, lab_poses, lab_e, lab_lig, lab_pred, norms, bias0;
  std::vector<int64_t> tb_all_host;  // field tile prefix of every pair (fixed: the poses are)
  bool field_tables = false;
  // rs_order lists each pipeline chunk's ligands heaviest first
  DeviceBuffer rs_order;
  // MLP
  bude::surrogate::Net net;
  // learned per-atom field: the network and its training workspace, the
  // training and held-out points (positions of ligand atoms under the
  // resident poses, fixed for the object's life), their exact labels
  // (per protein) and the order the batches read the training points in
  bude::atomnet::Net anet;
  bude::atomnet::Workspace aws;
  DeviceBuffer apoints, alabels, aperm, ameta;
  size_t antrain = 0, anval = 0;
  // exact energies of every pair (exhaustive) and the protein they belong to
  uint64_t exact_sum = 0;
  bool have_exact = false;
  // stream: exact passes (work-queue counter), selection of the whole set,
  // MLP; stream2: maps, field scores, chunk selections and tables; stream3:
  // the sampled-pose pass and its field scores
  hipStream_t stream = nullptr, stream2 = nullptr, stream3 = nullptr;

  ~bude_combined() {
    if (stream) (void)hipStreamDestroy(stream);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream3) (void)hipStreamDestroy(stream3);
  }
};

namespace {

struct Span {
  size_t lig, off, n;  // ligand, first pose in the pose array, poses
};

// Bytes one protein of natpro atoms uploads (its tables, the rescoring meta,
// the sampled poses' tables), with room for the 256-byte alignment.
size_t stage_bound(const bude_combined *c, size_t natpro) {
  const size_t L = c->nlig, lpp = std::min<size_t>(c->cfg.labels_per_pair, c->total);
  return natpro * 16 + (natpro + 2) * 4 + c->slot_type.size() * c->ff.size() * sizeof(bude::PairCoef) +
         3 * L * sizeof(bude::LigandMeta) + 2 * L * lpp * 4 + L * (lpp / bude::kSplitTile + 1) * 8 + (L + 1) * 8 +
         L * sizeof(bude::sel::Segment) + 16 * 256;
}

// Room for `bytes` of this protein's uploads in the slot's staging buffer
// (grown with headroom, before any upload of the protein is queued: the
// slot's earlier uploads finished with its previous protein).
int stage_reserve(ProteinSlot &ps, size_t bytes) {
  ps.stage_off = 0;
  if (bytes <= ps.stage_cap) return BUDE_OK;
  if (ps.stage) (void)hipHostFree(ps.stage);
  ps.stage = nullptr;
  ps.stage_cap = 0;
  const size_t cap = bytes + bytes / 2;
  if (int st = bude::hip_check(hipHostMalloc(reinterpret_cast<void **>(&ps.stage), cap), "hipHostMalloc")) return st;
  ps.stage_cap = cap;
  return BUDE_OK;
}

// Queues the upload of n bytes into dst on stream s through the staging
// buffer; the host does not wait for it.
int stage_upload(ProteinSlot &ps, DeviceBuffer &dst, const void *src, size_t n, hipStream_t s) {
  if (int st = dst.reserve(n)) return st;
  if (n == 0) return BUDE_OK;
  const size_t end = ps.stage_off + ((n + 255) & ~size_t(255));
  if (end > ps.stage_cap) return bude::fail(BUDE_ERR_RUNTIME, "combined: staging buffer too small");
  std::memcpy(ps.stage + ps.stage_off, src, n);
  if (int st = bude::hip_check(hipMemcpyAsync(dst.ptr, ps.stage + ps.stage_off, n, hipMemcpyHostToDevice, s),
                               "hipMemcpyAsync H2D"))
    return st;
  ps.stage_off = end;
  return BUDE_OK;
}

// Synchronous upload, or staged on stream s when s is not null.
int put(bude_combined *c, DeviceBuffer &dst, const void *src, size_t n, hipStream_t s) {
  return s ? stage_upload(c->s(), dst, src, n, s) : dst.assign(src, n);
}

// Work tables of one persistent batched launch scoring spans[i].n poses of
// ligand spans[i].lig starting at row spans[i].off of the pose array:
// per-ligand meta and the tile list (heaviest ligands first), uploaded into
// this launch's own buffers (synchronously: nothing may be reading them).
// Returns the tile count in *ntiles (0: nothing to score).
int prepare_exact(bude_combined *c, const std::vector<Span> &spans, DeviceBuffer &meta, DeviceBuffer &tiles,
                  int *ntiles, size_t tile_poses = 64 * kBatchPpwi, hipStream_t stage_on = nullptr) {
  *ntiles = 0;
  std::vector<bude::LigandMeta> m(c->nlig);
  for (size_t l = 0; l < c->nlig; l++) {
    m[l].lig_off = c->lig_off[l];
    m[l].natlig = int32_t(c->lig[l].size());
    m[l].coef_off = c->s().coef_off[l];
  }
  std::vector<size_t> order;
  for (size_t i = 0; i < spans.size(); i++)
    if (spans[i].n) {
      m[spans[i].lig].nposes = int32_t(spans[i].n);
      m[spans[i].lig].pose_off = int64_t(spans[i].off);
      order.push_back(i);
    }
  if (order.empty()) return BUDE_OK;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return c->lig[spans[a].lig].size() > c->lig[spans[b].lig].size(); });
  std::vector<int2> t;
  for (size_t i : order)
    for (size_t j = 0; j < spans[i].n; j += tile_poses) t.push_back(make_int2(int(spans[i].lig), int(j)));
  if (t.size() > 0x7fffffffu) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: too many tiles");
  if (int st = put(c, meta, m.data(), m.size() * sizeof(bude::LigandMeta), stage_on)) return st;
  if (int st = put(c, tiles, t.data(), t.size() * sizeof(int2), stage_on)) return st;
  *ntiles = int(t.size());
  return BUDE_OK;
}

// The launch prepared by prepare_exact: energies[off + j] of every span, on
// the first stream (which owns the work-queue counter).
int launch_exact(bude_combined *c, const float *poses, size_t stride, float *energies, const DeviceBuffer &meta,
                 const DeviceBuffer &tiles, int ntiles) {
  if (ntiles == 0) return BUDE_OK;
  bude::BatchArgs a{};
  a.protein = static_cast<const float4 *>(c->s().protein.ptr);
  a.group_begin = static_cast<const int32_t *>(c->s().group_begin.ptr);
  a.coef = static_cast<const bude::PairCoef *>(c->s().coef.ptr);
  a.ligands = static_cast<const float4 *>(c->ligands.ptr);
  a.meta = static_cast<const bude::LigandMeta *>(meta.ptr);
  a.tiles = static_cast<const int2 *>(tiles.ptr);
  a.poses = poses;
  a.energies = energies;
  a.counter = static_cast<unsigned int *>(c->counter.ptr);
  a.stride = stride;
  a.ngroups = int(c->s().ngroups);
  a.ntiles = ntiles;
  a.form = bude::plain_form();
  const bude::PinnedGrid g = bude::pinned_grid(c->cus, c->lds_per_cu, batch_waves_per_simd(ntiles, c->cus));
  a.lds_bytes = g.lds_bytes;
  return bude::hip_check(bude::launch_batch(a, kBatchPpwi, g.wgsize, g.blocks, c->stream), "combined batch launch");
}

// The sampled-pose pass (about a thousand poses per pair): the split form,
// 16-pose tiles, four lanes per pose, so the short pass fills the SIMDs
// instead of leaving one latency-bound wave on each.  Its energies only
// train or validate the surrogate; every energy the screen returns comes
// from the tiled kernel.
int launch_sampled(bude_combined *c, const float *poses, size_t stride, float *energies, const DeviceBuffer &meta,
                   const DeviceBuffer &tiles, int ntiles, hipStream_t stream) {
  if (ntiles == 0) return BUDE_OK;
  bude::BatchArgs a{};
  a.protein = static_cast<const float4 *>(c->s().protein.ptr);
  a.group_begin = static_cast<const int32_t *>(c->s().group_begin.ptr);
  a.coef = static_cast<const bude::PairCoef *>(c->s().coef.ptr);
  a.ligands = static_cast<const float4 *>(c->ligands.ptr);
  a.meta = static_cast<const bude::LigandMeta *>(meta.ptr);
  a.tiles = static_cast<const int2 *>(tiles.ptr);
  a.poses = poses;
  a.energies = energies;
  a.stride = stride;
  a.ngroups = int(c->s().ngroups);
  a.ntiles = ntiles;
  a.form = bude::plain_form();
  return bude::hip_check(bude::launch_split(a, stream), "combined sampled-pose launch");
}

int exact_ragged(bude_combined *c, const std::vector<Span> &spans, const float *poses, size_t stride,
                 float *energies, DeviceBuffer &meta, DeviceBuffer &tiles) {
  int nt = 0;
  if (int st = prepare_exact(c, spans, meta, tiles, &nt)) return st;
  return launch_exact(c, poses, stride, energies, meta, tiles, nt);
}

// Work tables of a field scoring of spans (as prepare_exact): meta and the
// per-ligand tile prefix.  Returns the tile count in *ntiles.
int prepare_field(bude_combined *c, const std::vector<Span> &spans, DeviceBuffer &meta, DeviceBuffer &tile_begin,
                  int64_t *ntiles, std::vector<int64_t> *tb_host = nullptr, hipStream_t stage_on = nullptr) {
  *ntiles = 0;
  std::vector<bude::LigandMeta> m(c->nlig);
  std::vector<int64_t> tb(c->nlig + 1, 0);
  for (size_t l = 0; l < c->nlig; l++) {
    m[l].lig_off = c->lig_off[l];
    m[l].natlig = int32_t(c->lig[l].size());
  }
  for (const Span &s : spans) {
    m[s.lig].nposes = int32_t(s.n);
    m[s.lig].pose_off = int64_t(s.off);
  }
  for (size_t l = 0; l < c->nlig; l++) tb[l + 1] = tb[l] + int64_t((size_t(m[l].nposes) + 63) / 64);
  if (tb_host) *tb_host = tb;
  if (tb[c->nlig] == 0) return BUDE_OK;
  if (int st = put(c, meta, m.data(), m.size() * sizeof(bude::LigandMeta), stage_on)) return st;
  if (int st = put(c, tile_begin, tb.data(), tb.size() * sizeof(int64_t), stage_on)) return st;
  *ntiles = tb[c->nlig];
  return BUDE_OK;
}

// Interpolated field scores (out[off + j]) of tiles [tile0, tile0 + ntiles)
// of the spans prepared by prepare_field, from the current cells.
// Tiles of one field launch when the atoms are split over the XCDs: 16.8M
// poses, 512 MiB of partial sums at 8 groups (BUDE_MI355X_FIELD_LAUNCH_TILES
// sets a smaller count, so tests can run many launches on small sets).
int64_t field_launch_tiles() {
  const char *env = std::getenv("BUDE_MI355X_FIELD_LAUNCH_TILES");
  const long long v = env ? std::atoll(env) : 0;
  return v > 0 && v < (1ll << 18) ? int64_t(v) : int64_t(1) << 18;
}

// The field scorer's atom groups (field::split_atoms): one per XCD, or one
// in all when BUDE_MI355X_FIELD_GROUPS=1 (the form without the split, kept
// to measure the split against).
int group_field_atoms(bude_combined *c, const std::vector<float> &lig4) {
  const char *env = std::getenv("BUDE_MI355X_FIELD_GROUPS");
  c->field_groups = env && std::strcmp(env, "1") == 0 ? 1 : bude::field::kXcdGroups;
  c->field_launch_tiles = field_launch_tiles();
  std::vector<float4> atoms(lig4.size() / 4), grouped;
  std::memcpy(atoms.data(), lig4.data(), atoms.size() * sizeof(float4));
  std::vector<int32_t> off(c->lig_off), begin;
  off.push_back(int32_t(atoms.size()));
  bude::field::split_atoms(c->field_groups, atoms, off, c->nposes, int(c->slot_type.size()), grouped, begin);
  if (int st = c->fgrouped.upload(grouped.data(), grouped.size() * sizeof(float4))) return st;
  return c->fgroup_begin.upload(begin.data(), begin.size() * sizeof(int32_t));
}

// Bytes of partial-sum scratch a field launch over `ntiles` tiles needs.
size_t field_partial_bytes(const bude_combined *c, int64_t ntiles) {
  return c->field_groups > 1 ? size_t(std::min(ntiles, c->field_launch_tiles)) * size_t(c->field_groups) * 64 * 4
                             : 0;
}

// partial: scratch of field_partial_bytes(ntiles), used by nothing else on
// another stream meanwhile.
int launch_field(bude_combined *c, const float *poses, size_t stride, float *out, const DeviceBuffer &meta,
                 const DeviceBuffer &tile_begin, int64_t tile0, int64_t ntiles, hipStream_t stream,
                 DeviceBuffer &partial, bool raise_priority = false) {
  if (ntiles == 0) return BUDE_OK;
  if (partial.bytes < field_partial_bytes(c, ntiles))
    return bude::fail(BUDE_ERR_RUNTIME, "field score: partial-sum scratch too small");
  bude::field::ScoreArgs a{};
  a.cells = static_cast<const uint4 *>(c->cells.ptr);
  a.grid = c->grid;
  a.grouped = static_cast<const float4 *>(c->fgrouped.ptr);
  a.group_begin = static_cast<const int32_t *>(c->fgroup_begin.ptr);
  a.ngroups = c->field_groups;
  a.partial = static_cast<float *>(partial.ptr);
  a.meta = static_cast<const bude::LigandMeta *>(meta.ptr);
  a.tile_begin = static_cast<const int64_t *>(tile_begin.ptr);
  a.nligands = int(c->nlig);
  a.poses = poses;
  a.stride = stride;
  a.out = out;
  a.raise_priority = raise_priority;
  const int64_t step = c->field_groups > 1 ? c->field_launch_tiles : ntiles;
  for (int64_t t = 0; t < ntiles; t += step) {
    a.tile0 = tile0 + t;
    if (int st = bude::hip_check(bude::field::score(a, std::min(step, ntiles - t), stream), "field score")) return st;
  }
  return BUDE_OK;
}

// Surrogate scores of the per-atom kinds (the field's maps or the learned
// per-atom field) of tiles [tile0, tile0 + ntiles) of the spans prepared by
// prepare_field.
int launch_surrogate(bude_combined *c, const float *poses, size_t stride, float *out, const DeviceBuffer &meta,
                     const DeviceBuffer &tile_begin, int64_t tile0, int64_t ntiles, hipStream_t stream,
                     DeviceBuffer &partial, bool raise_priority = false) {
  if (c->cfg.surrogate != BUDE_SURROGATE_ATOM_MLP)
    return launch_field(c, poses, stride, out, meta, tile_begin, tile0, ntiles, stream, partial, raise_priority);
  if (ntiles == 0) return BUDE_OK;
  bude::atomnet::ScoreArgs a = bude::atomnet::score_args(c->anet);
  a.ligands = static_cast<const float4 *>(c->ligands.ptr);
  a.meta = static_cast<const bude::LigandMeta *>(meta.ptr);
  a.tile_begin = static_cast<const int64_t *>(tile_begin.ptr);
  a.nligands = int(c->nlig);
  a.poses = poses;
  a.stride = stride;
  a.out = out;
  a.tile0 = tile0;
  a.raise_priority = raise_priority;
  return bude::hip_check(bude::atomnet::score(a, ntiles, stream), "atomnet score");
}

// Fallback statistics of one pair from its validation samples: exact
// energies e and surrogate scores p (energy units).  z = (Q(keep) -
// Q(top)) / sd, with Q the samples' exact energy quantiles and sd the
// surrogate's error in energy units over the lowest tail_q of the samples
// (by exact energy): the residual deviation of the least-squares line
// p = a e + b there, over a (a ranking is blind to a and b).  A pose of
// exact energy e is selected unless its error exceeds the selection
// threshold's by Q(keep) - e; with Gaussian errors of deviation sd and the
// exact energies spread evenly over the gap, the exact top fraction loses
// G(z) / z of its poses, G(z) = phi(z) - z Phi(-z): 0.0093 at z = 1.75,
// 0.0043 at z = 2.  need: the fraction of the samples whose exact energy is
// at most Q(top) + z_req sd, the smallest rescored fraction whose gap
// reaches z_req (infinite without a usable slope).  recall: of the
// samples' exact lowest 5 %, the fraction among the surrogate's lowest 10 %.
struct Gate {
  float z, recall;
  double need;
};

Gate gate_stats(const std::vector<double> &e, const std::vector<double> &p, double keep, double top, double tail_q,
                double z_req) {
  const size_t n = e.size();
  if (n < 2) return Gate{0.f, 0.f, INFINITY};
  std::vector<size_t> oe(n), op(n);
  std::iota(oe.begin(), oe.end(), size_t(0));
  std::iota(op.begin(), op.end(), size_t(0));
  std::stable_sort(oe.begin(), oe.end(), [&](size_t a, size_t b) { return e[a] < e[b]; });
  std::stable_sort(op.begin(), op.end(), [&](size_t a, size_t b) { return p[a] < p[b]; });
  auto q = [&](double f) {
    const double x = std::min(1.0, std::max(0.0, f)) * double(n - 1);
    const size_t i = size_t(x);
    const double t = x - double(i);
    return i + 1 < n ? e[oe[i]] * (1 - t) + e[oe[i + 1]] * t : e[oe[n - 1]];
  };
  const double gap = q(keep) - q(top);
  const size_t nt = std::min(n, std::max<size_t>(8, size_t(std::ceil(tail_q * double(n)))));
  double se = 0, sp = 0, see = 0, sep = 0, spp = 0;
  for (size_t k = 0; k < nt; k++) {
    const double x = e[oe[k]], y = p[oe[k]];
    se += x;
    sp += y;
    see += x * x;
    sep += x * y;
    spp += y * y;
  }
  const double nn = double(nt), ve = see / nn - (se / nn) * (se / nn), vp = spp / nn - (sp / nn) * (sp / nn),
               cv = sep / nn - (se / nn) * (sp / nn);
  // slope of p on e and the residual deviation, in energy units; a flat or
  // inverted relation gets no room
  const double a = ve > 0 ? cv / ve : 0.0;
  const double sd = a > 0 ? std::sqrt(std::max(0.0, vp - a * cv)) / a : INFINITY;
  Gate g{};
  g.z = sd > 0 ? float(gap / sd) : (gap > 0 ? INFINITY : 0.f);
  g.need = INFINITY;
  if (std::isfinite(sd)) {
    const double thr = q(top) + z_req * sd;
    size_t below = 0;
    while (below < n && e[oe[below]] <= thr) below++;
    g.need = double(below) / double(n);
  }
  const size_t kt = std::max<size_t>(1, size_t(std::lround(0.05 * double(n))));
  const size_t ks = std::min(n, std::max(kt, size_t(std::lround(0.10 * double(n)))));
  std::vector<char> sel(n, 0);
  for (size_t k = 0; k < ks; k++) sel[op[k]] = 1;
  size_t hits = 0;
  for (size_t k = 0; k < kt; k++) hits += size_t(sel[oe[k]]);
  g.recall = float(double(hits) / double(kt));
  return g;
}

// Learned per-atom field for the protein in slot c->s() on the second
// stream, after its tables: exact labels of the training and held-out
// points, a fresh network (seeded by the protein), training.  The host waits
// for it.
int train_atom(bude_combined *c, uint64_t pseed) {
  ProteinSlot &ps = c->s();
  const size_t n = c->antrain + c->anval;
  if (int st = bude::hip_check(bude::atomnet::label_points(static_cast<const float4 *>(ps.protein.ptr),
                                                           static_cast<const int32_t *>(ps.group_begin.ptr),
                                                           static_cast<const bude::PairCoef *>(ps.slot_coef.ptr),
                                                           int(ps.ngroups), static_cast<const float4 *>(c->apoints.ptr),
                                                           n, bude::plain_form(), static_cast<float *>(c->alabels.ptr),
                                                           c->stream2),
                               "label points"))
    return st;
  if (int st = bude::atomnet::init(c->anet, c->grid, int(c->slot_type.size()), pseed, c->stream2)) return st;
  bude::atomnet::TrainParams tp;
  tp.batch = c->cfg.batch;
  tp.steps = c->cfg.steps;
  tp.lr = c->cfg.lr;
  tp.grid_lr = c->cfg.grid_lr;
  tp.seed = pseed;
  bude::atomnet::TrainReport rep{};
  if (int st = bude::atomnet::train(c->anet, c->aws, static_cast<const float4 *>(c->apoints.ptr),
                                    static_cast<const float *>(c->alabels.ptr),
                                    static_cast<const uint32_t *>(c->aperm.ptr), c->antrain, c->anval, tp, c->stream2,
                                    &rep))
    return st;
  ps.atom_val_rmse = rep.val_rmse;
  ps.res.surrogate_rmse = rep.val_rmse;
  return BUDE_OK;
}

// Per-pair top-k of pred over the pairs of `plan` (pairs first .. first +
// n - 1) and their rescoring tables from the ligand list `order` (n
// entries) into tiles / *ntiles, on stream s.
int select_and_tables(bude_combined *c, const bude::sel::Plan &plan, size_t first, const int32_t *order, int n,
                      int2 *tiles, int32_t *ntiles, hipStream_t s) {
  uint32_t *sel_n = static_cast<uint32_t *>(c->s().sel_n.ptr);
  if (int st = bude::hip_check(bude::sel::select_lowest_segments(plan, static_cast<const float *>(c->pred.ptr),
                                                                 static_cast<uint32_t *>(c->s().sel_idx.ptr),
                                                                 sel_n + first, s),
                               "select"))
    return st;
  return bude::hip_check(
      bude::build_rescore_tables(order, n, sel_n, static_cast<bude::LigandMeta *>(c->s().meta_rs.ptr), tiles, ntiles, s),
      "rescoring tables");
}

// Exact energies of the selected poses listed by tiles / *ntiles (written
// on the device by select_and_tables) into exact_sel, on the first stream;
// `expect` tiles set the grid depth.
int launch_rescore(bude_combined *c, const int2 *tiles, const int32_t *ntiles, int64_t expect) {
  bude::BatchArgs a{};
  a.protein = static_cast<const float4 *>(c->s().protein.ptr);
  a.group_begin = static_cast<const int32_t *>(c->s().group_begin.ptr);
  a.coef = static_cast<const bude::PairCoef *>(c->s().coef.ptr);
  a.ligands = static_cast<const float4 *>(c->ligands.ptr);
  a.meta = static_cast<const bude::LigandMeta *>(c->s().meta_rs.ptr);
  a.tiles = tiles;
  a.poses = static_cast<const float *>(c->poses.ptr);
  a.energies = static_cast<float *>(c->s().exact_sel.ptr);
  a.counter = static_cast<unsigned int *>(c->counter.ptr);
  a.stride = c->total;
  a.ngroups = int(c->s().ngroups);
  a.form = bude::plain_form();
  a.pose_index = static_cast<const uint32_t *>(c->s().sel_idx.ptr);
  a.ntiles_dev = ntiles;
  const bude::PinnedGrid g = bude::pinned_grid(c->cus, c->lds_per_cu, batch_waves_per_simd(expect, c->cus));
  a.lds_bytes = g.lds_bytes;
  return bude::hip_check(bude::launch_batch_indexed(a, g.blocks, c->stream), "combined rescoring launch");
}

// The ligands [a, b) heaviest first (ties in index order).
std::vector<int32_t> heaviest_first(const bude_combined *c, size_t a, size_t b, const std::vector<int32_t> *skip) {
  std::vector<int32_t> o;
  for (size_t l = a; l < b; l++)
    if (!skip || !(*skip)[l]) o.push_back(int32_t(l));
  std::stable_sort(o.begin(), o.end(), [&](int32_t x, int32_t y) { return c->lig[x].size() > c->lig[y].size(); });
  return o;
}

// Selection segments (the lowest ceil(keep n) of each pair) of the whole set
// and of each pipeline chunk, the chunks' rescoring tile slots and ligand
// lists, and their events.  Both streams must be idle.  Until it succeeds
// the object screens nothing (a failure part way leaves mixed tables).
int plan_selection(bude_combined *c, float keep) {
  c->plan_ok = false;
  const size_t L = c->nlig;
  std::vector<bude::sel::Segment> segs(L);
  std::vector<int64_t> &kt = c->lig_expect, cap(L);
  kt.assign(L, 0);
  int64_t expect_all = 0;
  double cost = 0;
  for (size_t i = 0; i < L; i++) {
    const double kf = std::ceil(double(keep) * double(c->nposes[i]));
    const uint32_t k = uint32_t(std::min<double>(double(c->nposes[i]), std::max(1.0, kf)));
    segs[i] = bude::sel::Segment{int64_t(c->pose_off[i]), uint32_t(c->nposes[i]), k};
    kt[i] = int64_t((k + 63) / 64);
    cap[i] = int64_t((c->nposes[i] + 63) / 64);
    expect_all += kt[i];
    cost += double(c->nposes[i]) * double(c->lig[i].size());
  }
  if (int st = c->plan.prepare(segs)) return st;
  for (ProteinSlot &ps : c->slots)
    if (int st = ps.wplan.prepare(segs)) return st;
  size_t K = 1;
  if (c->cfg.surrogate != BUDE_SURROGATE_MLP)
    K = c->cfg.pipeline_chunks ? size_t(c->cfg.pipeline_chunks)
                               : size_t(std::max<int64_t>(1, expect_all / (kChunkTilesPerCu * c->cus)));
  K = std::min({K, L, kMaxChunks});
  // boundaries at equal shares of the field work (poses x atoms)
  c->chunk_lig.assign(1, 0);
  double acc = 0;
  for (size_t l = 0; l + 1 < L && c->chunk_lig.size() < K; l++) {
    acc += double(c->nposes[l]) * double(c->lig[l].size());
    if (acc >= cost * double(c->chunk_lig.size()) / double(K)) c->chunk_lig.push_back(l + 1);
  }
  c->chunk_lig.push_back(L);
  K = c->chunk_lig.size() - 1;
  c->chunk_plan.clear();
  c->chunk_tile.assign(K + 1, 0);
  c->chunk_expect.assign(K, 0);
  std::vector<int32_t> order;
  for (size_t k = 0; k < K; k++) {
    const size_t a = c->chunk_lig[k], b = c->chunk_lig[k + 1];
    c->chunk_plan.push_back(std::make_unique<bude::sel::Plan>());
    const std::vector<bude::sel::Segment> cs(segs.begin() + long(a), segs.begin() + long(b));
    if (int st = c->chunk_plan[k]->prepare(cs)) return st;
    int64_t t = 0;
    for (size_t l = a; l < b; l++) {
      t += cap[l];
      c->chunk_expect[k] += kt[l];
    }
    c->chunk_tile[k + 1] = c->chunk_tile[k] + t;
    const std::vector<int32_t> o = heaviest_first(c, a, b, nullptr);
    order.insert(order.end(), o.begin(), o.end());
  }
  if (c->chunk_tile[K] > 0x7fffffff) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: too many rescoring tiles");
  if (int st = c->rs_order.assign(order.data(), order.size() * 4)) return st;
  for (ProteinSlot &ps : c->slots) {
    if (int st = ps.tiles_rs.reserve(size_t(c->chunk_tile[K]) * sizeof(int2))) return st;
    if (int st = ps.rs_ntiles.reserve((K + 1) * 4)) return st;
    while (ps.chunk_ev.size() < 2 * K) {
      hipEvent_t e = nullptr;
      if (int st = bude::hip_check(hipEventCreate(&e), "hipEventCreate")) return st;
      ps.chunk_ev.push_back(e);
    }
  }
  c->plan_ok = true;
  return BUDE_OK;
}

int sync(bude_combined *c) { return bude::hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize"); }

int record(bude_combined *c, int i, hipStream_t s = nullptr) {
  return bude::hip_check(hipEventRecord(c->s().ev[i], s ? s : c->stream), "hipEventRecord");
}

float elapsed_events(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

float elapsed(bude_combined *c, int a, int b) { return elapsed_events(c->s().ev[a], c->s().ev[b]); }

int check_config(const BudeCombinedConfig *cfg) {
  if (!cfg) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: config is NULL");
  if (cfg->surrogate != BUDE_SURROGATE_MLP && cfg->surrogate != BUDE_SURROGATE_FIELD &&
      cfg->surrogate != BUDE_SURROGATE_ATOM_MLP)
    return bude::fail(BUDE_ERR_UNSUPPORTED, "combined: surrogate must be BUDE_SURROGATE_MLP, BUDE_SURROGATE_FIELD or "
                                            "BUDE_SURROGATE_ATOM_MLP");
  if (!(cfg->gate_z >= 0.f) || !(cfg->gate_top_fraction > 0.f) || cfg->gate_top_fraction > cfg->keep_fraction ||
      !(cfg->widen_max >= 0.f) || cfg->widen_max > 1.f)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      "combined: gate_z >= 0, gate_top_fraction in (0, keep_fraction], widen_max in [0, 1]");
  if (!(cfg->keep_fraction > 0.f) || cfg->keep_fraction > 1.f || !(cfg->clip_quantile > 0.f) ||
      cfg->clip_quantile > 1.f || cfg->tail_quantile < 0.f || cfg->tail_quantile > 1.f || !(cfg->tail_weight > 0.f) ||
      cfg->max_val_rmse < 0.f)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      "combined: keep_fraction and clip_quantile in (0, 1], tail_quantile in [0, 1], tail_weight > 0, "
                      "max_val_rmse >= 0");
  if (cfg->labels_per_pair < 2 || cfg->labels_per_pair > (1u << 24))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: labels_per_pair must be in [2, 2^24]");
  if (cfg->surrogate == BUDE_SURROGATE_MLP &&
      (cfg->batch < 32 || cfg->batch % 32 || cfg->batch > (1u << 20) || !(cfg->lr > 0.f)))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: MLP batch must be a multiple of 32 in [32, 2^20], lr > 0");
  if (cfg->surrogate != BUDE_SURROGATE_MLP && !(cfg->grid_spacing >= 0.05f && cfg->grid_spacing <= 4.f))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: grid_spacing must be in [0.05, 4]");
  if (cfg->surrogate == BUDE_SURROGATE_ATOM_MLP &&
      (cfg->batch < 64 || cfg->batch % 64 || cfg->batch > (1u << 20) || cfg->train_points < cfg->batch ||
       cfg->train_points % cfg->batch || cfg->train_points > (1u << 28) || cfg->steps == 0 || !(cfg->lr > 0.f) ||
       !(cfg->grid_lr > 0.f)))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      "combined: ATOM_MLP batch must be a multiple of 64 in [64, 2^20], train_points a multiple of "
                      "batch up to 2^28, steps > 0, lr > 0, grid_lr > 0");
  if (cfg->pipeline_chunks > kMaxChunks)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: pipeline_chunks must be in [0, 16]");
  return BUDE_OK;
}

// Map grid covering every position a ligand atom can take: the poses rotate
// about the ligand frame's origin, so |R x| = |x| <= r, and translate by t.
int make_grid(bude_combined *c, const float *const *poses) {
  double r = 0;
  for (const auto &l : c->lig)
    for (const BudeAtom &a : l) r = std::max(r, std::sqrt(double(a.x) * a.x + double(a.y) * a.y + double(a.z) * a.z));
  double tmin[3] = {1e300, 1e300, 1e300}, tmax[3] = {-1e300, -1e300, -1e300};
  for (size_t l = 0; l < c->nlig; l++)
    for (int k = 0; k < 3; k++) {
      const float *row = poses[l] + size_t(3 + k) * c->nposes[l];
      for (size_t j = 0; j < c->nposes[l]; j++)
        if (std::isfinite(row[j])) {
          tmin[k] = std::min(tmin[k], double(row[j]));
          tmax[k] = std::max(tmax[k], double(row[j]));
        }
    }
  const double h = c->cfg.grid_spacing;
  size_t points = 1;
  for (int k = 0; k < 3; k++) {
    if (tmin[k] > tmax[k]) tmin[k] = tmax[k] = 0;
    // lo on a multiple of h: grid coordinates lo + i*h are then exact in
    // FP32 for spacings that are powers of two
    const double lo = std::floor((tmin[k] - r - h) / h) * h, hi = tmax[k] + r + h;
    c->grid.lo[k] = float(lo);
    const double n = std::ceil((hi - lo) / h) + 1;
    if (n > 4096) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: grid too large (raise grid_spacing)");
    c->grid.n[k] = std::max(2, int(n));
    points *= size_t(c->grid.n[k]);
  }
  c->grid.h = float(h);
  if (points > 0x7fffffffu || points * c->slot_type.size() * 4 > (size_t(32) << 30))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: maps would exceed 32 GiB (raise grid_spacing)");
  return BUDE_OK;
}

// Learned per-atom field: its training and held-out points (positions of
// random atoms of random ligands under random resident poses, one forcefield
// slot per 64 consecutive points) and the order training batches read them
// in.  They depend on the ligand set only, so every protein trains on the
// same positions (with its own exact labels).
int prepare_atom_points(bude_combined *c) {
  const size_t L = c->nlig, nslots = c->slot_type.size();
  if (nslots > size_t(bude::atomnet::kMaxSlots))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: the ligands use " + std::to_string(nslots) +
                                                     " forcefield types, the learned field at most " +
                                                     std::to_string(bude::atomnet::kMaxSlots));
  std::vector<bude::LigandMeta> m(L);
  for (size_t l = 0; l < L; l++)
    m[l] = bude::LigandMeta{c->lig_off[l], int32_t(c->lig[l].size()), 0, int32_t(c->nposes[l]),
                            int64_t(c->pose_off[l]), 0};
  if (int st = c->ameta.upload(m.data(), L * sizeof(bude::LigandMeta))) return st;
  c->antrain = c->cfg.train_points;
  c->anval = 65536;
  const size_t n = c->antrain + c->anval;
  if (int st = c->apoints.alloc(n * sizeof(float4))) return st;
  if (int st = c->alabels.alloc(n * 4)) return st;
  if (int st = bude::hip_check(bude::atomnet::sample_points(static_cast<const float4 *>(c->ligands.ptr),
                                                            static_cast<const bude::LigandMeta *>(c->ameta.ptr), int(L),
                                                            static_cast<const float *>(c->poses.ptr), c->total,
                                                            int(nslots), c->cfg.seed ^ 0x70017ull, n,
                                                            static_cast<float4 *>(c->apoints.ptr), nullptr),
                               "sample points"))
    return st;
  std::vector<uint32_t> perm(c->antrain);
  std::iota(perm.begin(), perm.end(), 0u);
  std::mt19937_64 rng(c->cfg.seed ^ 0x9e37ull);
  std::shuffle(perm.begin(), perm.end(), rng);
  return c->aperm.upload(perm.data(), perm.size() * 4);
}

}  // namespace

extern "C" int bude_mi355x_combined_create(int device, const BudeFFParams *forcefield, size_t ntypes, size_t nligands,
                                           const BudeAtom *const *ligands, const size_t *natlig,
                                           const float *const *poses, const size_t *nposes,
                                           const BudeCombinedConfig *cfg, bude_combined **out) {
  if (!out) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_create: out is NULL");
  *out = nullptr;
  if (int st = check_config(cfg)) return st;
  if (!forcefield || ntypes == 0 || nligands == 0 || nligands > (1u << 20) || !ligands || !natlig || !poses ||
      !nposes)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_create: empty forcefield or ligand set");
  size_t total = 0;
  for (size_t i = 0; i < nligands; i++) {
    const std::string who = "bude_mi355x_combined_create: ligand " + std::to_string(i);
    // the protein comes later; check the ligand against the forcefield with itself as a stand-in protein
    if (int st = bude::check_deck(ligands[i], natlig[i], ligands[i], natlig[i], forcefield, ntypes, who.c_str()))
      return st;
    if (!poses[i] || nposes[i] == 0 || nposes[i] > 0x7fffffffu)
      return bude::fail(BUDE_ERR_INVALID_ARGUMENT, who + ": no poses (or more than 2^31)");
    total += nposes[i];
  }
  if (total > 0xffffffffu) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_create: over 2^32 poses");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || device < 0 || device >= ndev)
    return bude::fail(BUDE_ERR_NO_DEVICE, "bude_mi355x_combined_create: no such HIP device");
  if (int st = bude::hip_check(hipSetDevice(device), "hipSetDevice")) return st;

  auto c = std::make_unique<bude_combined>();
  c->device = device;
  c->cfg = *cfg;
  hipDeviceProp_t props{};
  if (int st = bude::hip_check(hipGetDeviceProperties(&props, device), "hipGetDeviceProperties")) return st;
  c->cus = props.multiProcessorCount;
  c->lds_per_cu = props.maxSharedMemoryPerMultiProcessor;
  c->ff.assign(forcefield, forcefield + ntypes);
  c->nlig = nligands;
  c->total = total;
  c->slot_of_type.assign(ntypes, -1);
  std::vector<float> lig4, desc;
  size_t off = 0;
  for (size_t i = 0; i < nligands; i++) {
    c->lig.emplace_back(ligands[i], ligands[i] + natlig[i]);
    c->lig_off.push_back(int32_t(lig4.size() / 4));
    for (size_t a = 0; a < natlig[i]; a++) {
      const BudeAtom &at = ligands[i][a];
      int32_t &slot = c->slot_of_type[size_t(at.type)];
      if (slot < 0) {
        slot = int32_t(c->slot_type.size());
        c->slot_type.push_back(at.type);
      }
      float w;
      std::memcpy(&w, &slot, 4);
      lig4.insert(lig4.end(), {at.x, at.y, at.z, w});
    }
    if (lig4.size() / 4 > 0x7fffffffu) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: too many ligand atoms");
    const std::vector<float> d = descriptor(ligands[i], natlig[i]);
    desc.insert(desc.end(), d.begin(), d.end());
    c->nposes.push_back(nposes[i]);
    c->pose_off.push_back(off);
    off += nposes[i];
  }
  if (cfg->surrogate != BUDE_SURROGATE_MLP)
    if (int st = make_grid(c.get(), poses)) return st;

  int st = plan_selection(c.get(), cfg->keep_fraction);
  if (!st) st = c->ligands.upload(lig4.data(), lig4.size() * 4);
  if (!st && cfg->surrogate == BUDE_SURROGATE_FIELD) st = group_field_atoms(c.get(), lig4);
  if (!st) st = c->desc.upload(desc.data(), desc.size() * 4);
  if (!st) st = c->poses.alloc(6 * total * sizeof(float));
  for (size_t i = 0; i < nligands && !st; i++)
    for (int k = 0; k < 6 && !st; k++)
      st = bude::hip_check(hipMemcpy(static_cast<float *>(c->poses.ptr) + size_t(k) * total + c->pose_off[i],
                                     poses[i] + size_t(k) * nposes[i], nposes[i] * sizeof(float),
                                     hipMemcpyHostToDevice),
                           "hipMemcpy poses");
  if (!st) st = c->pred.alloc(total * sizeof(float));
  if (!st && cfg->surrogate == BUDE_SURROGATE_ATOM_MLP) st = prepare_atom_points(c.get());
  for (ProteinSlot &ps : c->slots) {
    if (!st) st = ps.sel_idx.alloc(total * sizeof(uint32_t));
    if (!st) st = ps.sel_n.alloc(nligands * sizeof(uint32_t));
    if (!st) st = ps.exact_sel.alloc(total * sizeof(float));
    for (auto &e : ps.ev)
      if (!st) st = bude::hip_check(hipEventCreate(&e), "hipEventCreate");
    if (!st)
      st = bude::hip_check(hipHostMalloc(reinterpret_cast<void **>(&ps.nsel_host), nligands * sizeof(uint32_t)),
                           "hipHostMalloc");
    // each slot's per-protein buffers sized now for proteins of up to 2^16
    // atoms (a batch then allocates nothing: an allocation can wait for the
    // kernels in flight)
    const size_t natoms = lig4.size() / 4, ntypes_ = c->ff.size(), pro_max = size_t(1) << 16;
    if (!st) st = ps.protein.reserve(pro_max * 16);
    if (!st) st = ps.group_begin.reserve((ntypes_ + 1) * 4);
    if (!st) st = ps.slot_coef.reserve(c->slot_type.size() * ntypes_ * sizeof(bude::PairCoef));
    if (!st) st = ps.meta_rs.reserve(nligands * sizeof(bude::LigandMeta));
    if (!st) st = ps.coef.reserve(natoms * ntypes_ * sizeof(bude::PairCoef));
    if (!st) st = stage_reserve(ps, stage_bound(c.get(), pro_max));
  }
  if (!st) st = c->counter.alloc(256);
  if (!st) st = bude::hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  if (!st) st = bude::hip_check(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking), "hipStreamCreate");
  if (!st) st = bude::hip_check(hipStreamCreateWithFlags(&c->stream3, hipStreamNonBlocking), "hipStreamCreate");
  if (!st) st = bude::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  if (st) return st;
  *out = c.release();
  return bude::ok();
}

namespace {

// First half of one protein's screen, into slot c->s(): everything up to
// the exact rescoring of the selection, which is left queued on the first
// stream (with the copy of the selection counts) for screen_end.  The host
// waits only for the sampled-pose pass.
int screen_begin(bude_combined *c, const BudeAtom *protein, size_t natpro) {
  if (int st = bude::check_deck(protein, natpro, c->lig[0].data(), c->lig[0].size(), c->ff.data(), c->ff.size(),
                                "bude_mi355x_combined_screen"))
    return st;
  if (!c->plan_ok)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      "bude_mi355x_combined_screen: the selection tables are incomplete (a failed set_keep); "
                      "call bude_mi355x_combined_set_keep again");
  BudeCombinedResult *res = &c->s().res;
  *res = BudeCombinedResult{};
  c->s().valid = false;
  c->s().busy = true;
  {  // FNV-1a of the protein: exact energies stay valid while it is unchanged
    uint64_t hsum = 1469598103934665603ull ^ uint64_t(natpro);
    const auto *bytes = reinterpret_cast<const unsigned char *>(protein);
    for (size_t i = 0; i < natpro * sizeof(BudeAtom); i++) hsum = (hsum ^ bytes[i]) * 1099511628211ull;
    c->s().protein_sum = hsum;
  }
  const auto t0 = std::chrono::steady_clock::now();
  c->s().t0 = t0;
  const size_t L = c->nlig;
  const BudeCombinedConfig &cfg = c->cfg;
  const bool mlp = cfg.surrogate == BUDE_SURROGATE_MLP;
  const bool atom = cfg.surrogate == BUDE_SURROGATE_ATOM_MLP;

  // 1. protein tables: grouped atoms and the pair coefficients of every map
  //    slot's type against every protein group (a coefficient depends on
  //    the two types only); the table of every ligand atom ([atom][group],
  //    pair l's at coef_off[l] = lig_off[l] * ngroups) is expanded from it
  //    on the device
  const bude::ProteinGroups groups = bude::group_protein(protein, natpro);
  c->s().ngroups = groups.ngroups();
  const size_t natoms = size_t(c->lig_off.back()) + c->lig.back().size(), ncoef = natoms * c->s().ngroups;
  if (ncoef > 0x7fffffffu) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: pair tables too large");
  c->s().coef_off.assign(L, 0);
  for (size_t l = 0; l < L; l++) c->s().coef_off[l] = int32_t(size_t(c->lig_off[l]) * c->s().ngroups);
  // every upload of this protein goes through the slot's pinned staging
  if (int st = stage_reserve(c->s(), stage_bound(c, natpro))) return st;
  {  // protein tables and the rescoring meta on the third stream (idle here:
     // the previous protein's sampled poses are done; the second stream may
     // still be scoring it); the second stream and, through the third, the
     // first wait for them
    std::vector<BudeAtom> probe(c->slot_type.size());
    for (size_t s = 0; s < probe.size(); s++) probe[s] = BudeAtom{0.f, 0.f, 0.f, c->slot_type[s]};
    const std::vector<bude::PairCoef> sc = bude::coef_table(probe.data(), probe.size(), groups, c->ff.data());
    if (int st = stage_upload(c->s(), c->s().slot_coef, sc.data(), sc.size() * sizeof(bude::PairCoef), c->stream3))
      return st;
    std::vector<bude::LigandMeta> m(L);  // everything but the pose counts, which the selection writes
    for (size_t l = 0; l < L; l++)
      m[l] = bude::LigandMeta{c->lig_off[l], int32_t(c->lig[l].size()), c->s().coef_off[l], 0, int64_t(c->pose_off[l]), 0};
    if (int st = stage_upload(c->s(), c->s().meta_rs, m.data(), m.size() * sizeof(bude::LigandMeta), c->stream3))
      return st;
  }
  if (int st = stage_upload(c->s(), c->s().protein, groups.xyzw.data(), groups.xyzw.size() * 4, c->stream3)) return st;
  if (int st = stage_upload(c->s(), c->s().group_begin, groups.begin.data(), groups.begin.size() * 4, c->stream3))
    return st;
  if (int st = record(c, 9, c->stream3)) return st;
  if (int st = bude::hip_check(hipStreamWaitEvent(c->stream2, c->s().ev[9], 0), "hipStreamWaitEvent")) return st;
  if (int st = c->s().coef.reserve(ncoef * sizeof(bude::PairCoef))) return st;
  hipLaunchKernelGGL(expand_coef_kernel, dim3(blocks_for(ncoef, 256)), dim3(256), 0, c->stream3,
                     static_cast<const float4 *>(c->ligands.ptr), ncoef, int(c->s().ngroups),
                     static_cast<const bude::PairCoef *>(c->s().slot_coef.ptr), static_cast<bude::PairCoef *>(c->s().coef.ptr));
  if (int st = bude::hip_check(hipGetLastError(), "expand pair tables")) return st;
  res->prepare_ms = float(ms_since(t0));

  // 2. sampled poses of every pair: first 15/16 train the MLP, the rest
  //    validate (the field validates on all of them)
  // (seeded by the protein itself: a protein draws the same samples whatever
  // was screened before it, alone or in a batch)
  const uint64_t pseed = cfg.seed + 0x9e3779b97f4a7c15ull * c->s().protein_sum;
  std::mt19937_64 rng(pseed);
  std::vector<uint32_t> src, lab_lig;
  std::vector<size_t> lab_off(L + 1, 0), lab_train(L, 0);
  for (size_t l = 0; l < L; l++) {
    const size_t n = c->nposes[l], m = std::min<size_t>(cfg.labels_per_pair, n);
    std::vector<uint32_t> pick;
    if (m == n) {
      pick.resize(n);
      std::iota(pick.begin(), pick.end(), 0u);
    } else {  // Floyd's sampling without replacement
      std::vector<bool> seen;
      std::vector<uint32_t> sorted;
      if (n <= (size_t(1) << 26)) seen.assign(n, false);
      for (size_t j = n - m; j < n; j++) {
        const uint32_t t = uint32_t(std::uniform_int_distribution<size_t>(0, j)(rng));
        const bool dup = seen.empty() ? std::binary_search(sorted.begin(), sorted.end(), t) : bool(seen[t]);
        const uint32_t v = dup ? uint32_t(j) : t;
        pick.push_back(v);
        if (!seen.empty()) seen[v] = true;
        else sorted.insert(std::upper_bound(sorted.begin(), sorted.end(), v), v);
      }
    }
    std::shuffle(pick.begin(), pick.end(), rng);
    for (uint32_t j : pick) {
      src.push_back(uint32_t(c->pose_off[l] + j));
      lab_lig.push_back(uint32_t(l));
    }
    lab_off[l + 1] = src.size();
    lab_train[l] = mlp ? m - std::max<size_t>(1, m / 16) : 0;
  }
  const size_t nlab = src.size();
  if (int st = stage_upload(c->s(), c->lab_src, src.data(), nlab * 4, c->stream3)) return st;
  if (int st = stage_upload(c->s(), c->lab_lig, lab_lig.data(), nlab * 4, c->stream3)) return st;
  if (int st = c->lab_poses.reserve(6 * nlab * 4)) return st;
  if (int st = c->lab_e.reserve(nlab * 4)) return st;
  if (int st = c->lab_pred.reserve(nlab * 4)) return st;
  std::vector<Span> lab_spans(L), all(L);
  for (size_t l = 0; l < L; l++) {
    lab_spans[l] = Span{l, lab_off[l], lab_off[l + 1] - lab_off[l]};
    all[l] = Span{l, c->pose_off[l], c->nposes[l]};
  }
  float *pred = static_cast<float *>(c->pred.ptr);
  const float *poses = static_cast<const float *>(c->poses.ptr);
  float *lp = static_cast<float *>(c->lab_pred.ptr);
  const float *lpos = static_cast<const float *>(c->lab_poses.ptr);

  // 3. every work table of the exact and field passes first: an upload
  //    issued while a kernel fills the device waits for it
  int nt_label = 0;
  int64_t ft_all = 0, ft_label = 0;
  const std::vector<int64_t> &tb_all = c->tb_all_host;
  if (int st = prepare_exact(c, lab_spans, c->meta_label, c->tiles_label, &nt_label, bude::kSplitTile, c->stream3))
    return st;
  if (!mlp) {
    const size_t np = c->grid.points(), nslots = c->slot_type.size();
    if (!atom) {
      if (int st = c->maps.reserve(np * nslots * 4)) return st;
      if (int st = c->cells.reserve(c->grid.cells() * nslots * 16)) return st;
    }
    if (!c->field_tables) {  // (the pose counts never change: built once, never rewritten while read)
      if (int st = prepare_field(c, all, c->meta_all, c->tb_all, &ft_all, &c->tb_all_host)) return st;
      c->field_tables = true;
    }
    ft_all = c->tb_all_host[L];
    if (int st = prepare_field(c, lab_spans, c->meta_lf, c->tb_label, &ft_label, nullptr, c->stream3)) return st;
    if (!atom && c->field_groups > 1) {  // partial sums of the largest chunk, and of the sampled poses
      int64_t most = 0;
      for (size_t k = 0; k + 1 < c->chunk_lig.size(); k++)
        most = std::max(most, c->tb_all_host[c->chunk_lig[k + 1]] - c->tb_all_host[c->chunk_lig[k]]);
      if (int st = c->fpart.reserve(field_partial_bytes(c, most))) return st;
      if (int st = c->fpart_lab.reserve(field_partial_bytes(c, ft_label))) return st;
    }
  }
  const size_t K = c->chunk_lig.size() - 1;
  int2 *rs_tiles = static_cast<int2 *>(c->s().tiles_rs.ptr);
  int32_t *rs_ntiles = static_cast<int32_t *>(c->s().rs_ntiles.ptr);

  // 4. exact energies of the sampled poses (one ragged split launch) on the
  //    third stream, and, for the field, maps -> cells -> scores of every
  //    pose of every pair on the second stream, concurrently (and beside the
  //    previous protein's rescoring on the first stream, in a batch)
  if (int st = record(c, 0, c->stream3)) return st;
  hipLaunchKernelGGL(gather_index_kernel, dim3(blocks_for(nlab, 256)), dim3(256), 0, c->stream3, poses, c->total,
                     static_cast<const uint32_t *>(c->lab_src.ptr), nlab, static_cast<float *>(c->lab_poses.ptr),
                     nlab);
  if (int st = bude::hip_check(hipGetLastError(), "gather labels")) return st;
  if (int st = launch_sampled(c, lpos, nlab, static_cast<float *>(c->lab_e.ptr), c->meta_label, c->tiles_label,
                              nt_label, c->stream3))
    return st;
  if (int st = record(c, 1, c->stream3)) return st;
  if (!mlp) {
    const size_t nslots = c->slot_type.size();
    if (int st = record(c, 2, c->stream2)) return st;
    if (atom) {
      if (int st = train_atom(c, pseed)) return st;
    } else {
      if (int st = bude::hip_check(
              bude::field::build_maps(static_cast<const float4 *>(c->s().protein.ptr),
                                      static_cast<const int32_t *>(c->s().group_begin.ptr),
                                      static_cast<const bude::PairCoef *>(c->s().slot_coef.ptr), int(c->s().ngroups),
                                      int(nslots), c->grid, bude::plain_form(), static_cast<float *>(c->maps.ptr),
                                      c->stream2),
              "build maps"))
        return st;
      if (int st = bude::hip_check(bude::field::pack_cells(static_cast<const float *>(c->maps.ptr), int(nslots),
                                                           c->grid, static_cast<uint4 *>(c->cells.ptr), c->stream2),
                                   "pack cells"))
        return st;
    }
    if (int st = record(c, 3, c->stream2)) return st;
    // chunk by chunk: scores, top-k and rescoring tables, all on the device;
    // the scoring of chunk k > 0 runs beside the rescoring of chunk k - 1
    // and issues ahead of it (chunk 0 runs beside the sampled-pose pass,
    // which the host waits for, and, in a batch, beside the previous
    // protein's rescoring, which it issues ahead of too)
    const bool overlap = c->slots[c->cur ^ 1].busy;
    {  // a previous protein still in flight may select from the scores this one overwrites
      const ProteinSlot &other = c->slots[c->cur ^ 1];
      if (other.busy && !other.pipelined)
        if (int st = bude::hip_check(hipStreamWaitEvent(c->stream2, other.ev[10], 0), "hipStreamWaitEvent")) return st;
    }
    for (size_t k = 0; k < K; k++) {
      const size_t a = c->chunk_lig[k], b = c->chunk_lig[k + 1];
      if (int st = launch_surrogate(c, poses, c->total, pred, c->meta_all, c->tb_all, tb_all[a], tb_all[b] - tb_all[a],
                                    c->stream2, c->fpart, k > 0 || overlap))
        return st;
      if (int st = bude::hip_check(hipEventRecord(c->s().chunk_ev[2 * k], c->stream2), "hipEventRecord")) return st;
      if (int st = select_and_tables(c, *c->chunk_plan[k], a, static_cast<const int32_t *>(c->rs_order.ptr) + a,
                                     int(b - a), rs_tiles + c->chunk_tile[k], rs_ntiles + k, c->stream2))
        return st;
      if (int st = bude::hip_check(hipEventRecord(c->s().chunk_ev[2 * k + 1], c->stream2), "hipEventRecord")) return st;
    }
    if (int st = record(c, 4, c->stream2)) return st;
  }
  std::vector<float> le(nlab), lpred(nlab);
  if (int st = bude::hip_check(hipMemcpyAsync(le.data(), c->lab_e.ptr, nlab * 4, hipMemcpyDeviceToHost, c->stream3),
                               "hipMemcpyAsync"))
    return st;
  if (!mlp) {  // field scores of the sampled poses, once the cells exist
    if (int st = bude::hip_check(hipStreamWaitEvent(c->stream3, c->s().ev[3], 0), "hipStreamWaitEvent")) return st;
    if (int st = launch_surrogate(c, lpos, nlab, lp, c->meta_lf, c->tb_label, 0, ft_label, c->stream3, c->fpart_lab))
      return st;
    if (int st = bude::hip_check(hipMemcpyAsync(lpred.data(), lp, nlab * 4, hipMemcpyDeviceToHost, c->stream3),
                                 "hipMemcpyAsync"))
      return st;
  }
  if (int st = bude::hip_check(hipStreamSynchronize(c->stream3), "hipStreamSynchronize")) return st;
  res->label_ms = elapsed(c, 0, 1);
  c->s().lab_energy = le;
  c->s().lab_offs = lab_off;
  c->s().lab_j.resize(nlab);
  for (size_t l = 0; l < L; l++)
    for (size_t i = lab_off[l]; i < lab_off[l + 1]; i++) c->s().lab_j[i] = uint32_t(src[i] - c->pose_off[l]);

  // per-pair target transform from the training part (all samples for the field)
  std::vector<Norm> norms(L);
  std::vector<bude::mlp::TargetNorm> tn(L);
  for (size_t l = 0; l < L; l++) {
    const size_t b = lab_off[l], e = (mlp && lab_train[l] > 0) ? b + lab_train[l] : lab_off[l + 1];
    norms[l] = norm_of(std::vector<float>(le.begin() + long(b), le.begin() + long(e)), cfg.clip_quantile,
                       cfg.tail_quantile);
    tn[l] = bude::mlp::TargetNorm{norms[l].clip, norms[l].mu, 1.0f / norms[l].sigma, norms[l].tail};
  }

  // 5. MLP: train on the training part, predict the validation part
  if (mlp) {
    if (int st = record(c, 2)) return st;
    if (int st = c->norms.assign(tn.data(), tn.size() * sizeof(tn[0]))) return st;
    std::vector<std::vector<float>> w, b;
    bude::surrogate::init_weights(pseed, bude::mlp::kIn, w, b);
    if (int st = bude::surrogate::set_weights(c->net, w, b, c->stream)) return st;
    std::vector<uint32_t> train;
    for (size_t l = 0; l < L; l++)
      for (size_t j = 0; j < lab_train[l]; j++) train.push_back(uint32_t(lab_off[l] + j));
    if (train.size() < cfg.batch)
      return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "combined: fewer training labels than one MLP batch");
    bude::surrogate::TrainData td;
    td.poses = lpos;
    td.stride = nlab;
    td.energy = static_cast<const float *>(c->lab_e.ptr);
    td.table = static_cast<const bude::mlp::TargetNorm *>(c->norms.ptr);
    td.lig = static_cast<const uint32_t *>(c->lab_lig.ptr);
    td.desc = static_cast<const float *>(c->desc.ptr);
    bude::surrogate::TrainParams tp;
    tp.batch = cfg.batch;
    tp.epochs = cfg.epochs;
    tp.lr = cfg.lr;
    tp.tail_weight = cfg.tail_weight;
    tp.seed = pseed;
    if (int st = bude::surrogate::train(c->net, td, std::move(train), tp, c->stream, nullptr, 0, nullptr)) return st;
    if (int st = c->bias0.reserve(L * bude::mlp::kHidden * 4)) return st;
    if (int st = bude::hip_check(bude::mlp::fold_bias(static_cast<const float *>(c->net.b[0].ptr),
                                                      static_cast<const bude::mlp::bf16 *>(c->net.w16tx[0].ptr),
                                                      static_cast<const float *>(c->desc.ptr), int(L),
                                                      static_cast<float *>(c->bias0.ptr), c->stream),
                                 "fold_bias"))
      return st;
    if (int st = record(c, 3)) return st;
    for (size_t l = 0; l < L; l++) {
      const size_t b0 = lab_off[l] + lab_train[l], n = lab_off[l + 1] - b0;
      if (n == 0) continue;
      if (int st = bude::surrogate::predict(c->net, lpos + b0, nlab, n, lp + b0, 2 * c->cus, c->stream,
                                            static_cast<const float *>(c->bias0.ptr) + l * bude::mlp::kHidden))
        return st;
    }
  }
  if (mlp) {
    if (int st = bude::hip_check(hipMemcpyAsync(lpred.data(), lp, nlab * 4, hipMemcpyDeviceToHost, c->stream),
                                 "hipMemcpyAsync"))
      return st;
    if (int st = sync(c)) return st;
  }

  // 6. validation per pair: RMSE on the clipped, standardised scale of that
  //    pair, and the gate statistics in energy units (the MLP's scores are
  //    standardised: mapped back, they are energies below the clip)
  c->s().fallback.assign(L, 0);
  c->s().val_rmse.assign(L, 0.f);
  c->s().val_rmse_cal.assign(L, 0.f);
  c->s().gate_z.assign(L, 0.f);
  c->s().sample_recall.assign(L, 0.f);
  c->s().keep.assign(L, cfg.keep_fraction);
  std::vector<double> need(L, 0.0);
  const double tail_q = mlp ? std::min(0.25, double(cfg.clip_quantile)) : 0.25;
  for (size_t l = 0; l < L; l++) {
    const size_t b = lab_off[l] + lab_train[l], e = lab_off[l + 1];
    if (e <= b) continue;
    {
      std::vector<double> ev, pv;
      for (size_t i = b; i < e; i++) {
        ev.push_back(double(le[i]));
        pv.push_back(mlp ? double(lpred[i]) * norms[l].sigma + norms[l].mu : double(lpred[i]));
      }
      const Gate g = gate_stats(ev, pv, double(cfg.keep_fraction), double(cfg.gate_top_fraction), tail_q,
                                double(cfg.gate_z));
      c->s().gate_z[l] = g.z;
      c->s().sample_recall[l] = g.recall;
      need[l] = g.need;
    }
    double se = 0, sy = 0, sp = 0, syy = 0, spp = 0, syp = 0;
    for (size_t i = b; i < e; i++) {
      const double y = (double(std::min(le[i], norms[l].clip)) - norms[l].mu) / norms[l].sigma;
      const double p = mlp ? double(lpred[i])
                           : (double(std::min(lpred[i], norms[l].clip)) - norms[l].mu) / norms[l].sigma;
      se += (p - y) * (p - y);
      sy += y;
      sp += p;
      syy += y * y;
      spp += p * p;
      syp += y * p;
    }
    const double n = double(e - b);
    const double vy = std::max(0.0, syy / n - (sy / n) * (sy / n));
    const double vp = std::max(0.0, spp / n - (sp / n) * (sp / n));
    const double cv = syp / n - (sy / n) * (sp / n);
    c->s().val_rmse[l] = float(std::sqrt(se / n));
    // residual of the best increasing affine map of the scores onto the
    // targets: the error that matters for a ranking
    c->s().val_rmse_cal[l] = float(std::sqrt(vp > 0 && cv > 0 ? std::max(0.0, vy - cv * cv / vp) : vy));
    // short of the gate at keep_fraction: widened to the fraction that
    // reaches it (with a 5 % margin) if that is within widen_max
    bool fb = !(c->s().val_rmse_cal[l] <= cfg.max_val_rmse);
    if (!fb && !(c->s().gate_z[l] >= cfg.gate_z)) {
      const double f = need[l] * 1.05;
      if (cfg.widen_max > 0.f && f <= double(cfg.widen_max)) {
        c->s().keep[l] = float(std::max(double(cfg.keep_fraction), f));
        res->widened_pairs++;
      } else {
        fb = true;
      }
    }
    c->s().fallback[l] = fb;
    if (fb) c->s().keep[l] = 1.f;
    res->fallback_pairs += uint32_t(fb);
  }

  // 7. Without fallback pairs the field screen is already selecting chunk
  //    by chunk on the second stream.  Otherwise: MLP scores of every pose
  //    of every non-fallback pair (the field's are on the second stream),
  //    exhaustive exact scores of the fallback pairs overwriting theirs,
  //    and the rescoring list of the non-fallback pairs
  std::vector<Span> fb;
  for (size_t l = 0; l < L; l++)
    if (c->s().fallback[l]) fb.push_back(all[l]);
  const bool widened = res->widened_pairs > 0;
  const bool pipelined = !mlp && fb.empty() && !widened;
  int64_t expect = 0;
  int nrs = 0;
  if (!pipelined) {
    std::vector<int32_t> o = heaviest_first(c, 0, L, &c->s().fallback);
    for (int32_t l : o) expect += c->lig_expect[size_t(l)];
    nrs = int(o.size());
    if (o.empty()) o.push_back(0);  // every pair fell back: nothing to rescore
    if (int st = c->s().rs_sel.assign(o.data(), o.size() * 4)) return st;
  }
  if (widened) {  // this protein's per-pair k, into the slot's own segment table
    std::vector<bude::sel::Segment> segs(L);
    expect = 0;
    for (size_t l = 0; l < L; l++) {  // (a fallback pair selects its lowest keep_fraction of exact energies)
      const double f = c->s().fallback[l] ? double(cfg.keep_fraction) : double(c->s().keep[l]);
      const double kf = std::ceil(f * double(c->nposes[l]));
      const uint32_t k = uint32_t(std::min<double>(double(c->nposes[l]), std::max(1.0, kf)));
      segs[l] = bude::sel::Segment{int64_t(c->pose_off[l]), uint32_t(c->nposes[l]), k};
      if (!c->s().fallback[l]) expect += int64_t((k + 63) / 64);
    }
    if (int st = stage_upload(c->s(), c->s().wplan.segs, segs.data(), L * sizeof(segs[0]), c->stream)) return st;
  }
  if (mlp) {
    if (int st = record(c, 4)) return st;
    for (size_t l = 0; l < L; l++)
      if (!c->s().fallback[l])
        if (int st = bude::surrogate::predict(c->net, poses + c->pose_off[l], c->total, c->nposes[l],
                                              pred + c->pose_off[l], 2 * c->cus, c->stream,
                                              static_cast<const float *>(c->bias0.ptr) + l * bude::mlp::kHidden))
          return st;
  } else if (!pipelined) {
    if (int st = bude::hip_check(hipStreamWaitEvent(c->stream, c->s().ev[4], 0), "hipStreamWaitEvent")) return st;
  }
  if (int st = record(c, 5)) return st;
  if (int st = exact_ragged(c, fb, poses, c->total, pred, c->meta_fb, c->tiles_fb)) return st;
  if (int st = record(c, 6)) return st;

  // 8-9. per-pair top-k of the scores and exact rescoring of every
  //      non-fallback pair's selection: chunk by chunk behind the second
  //      stream, or the whole set at once
  if (pipelined) {
    for (size_t k = 0; k < K; k++) {
      if (int st = bude::hip_check(hipStreamWaitEvent(c->stream, c->s().chunk_ev[2 * k + 1], 0), "hipStreamWaitEvent"))
        return st;
      if (k == 0)
        if (int st = record(c, 7)) return st;
      if (int st = launch_rescore(c, rs_tiles + c->chunk_tile[k], rs_ntiles + k, c->chunk_expect[k])) return st;
    }
  } else {
    if (int st = select_and_tables(c, widened ? c->s().wplan : c->plan, 0,
                                   static_cast<const int32_t *>(c->s().rs_sel.ptr), nrs, rs_tiles, rs_ntiles + K,
                                   c->stream))
      return st;
    if (int st = record(c, 10)) return st;  // the scores are read
    if (int st = record(c, 7)) return st;
    if (nrs)
      if (int st = launch_rescore(c, rs_tiles, rs_ntiles + K, expect)) return st;
  }
  // the counts land in pinned memory: the host does not wait here
  if (int st = bude::hip_check(hipMemcpyAsync(c->s().nsel_host, c->s().sel_n.ptr, L * 4, hipMemcpyDeviceToHost,
                                              c->stream),
                               "hipMemcpyAsync"))
    return st;
  if (int st = record(c, 8)) return st;
  c->s().pipelined = pipelined;
  c->s().overlap = pipelined || (!mlp && fb.empty());
  c->s().mlp = mlp;
  c->s().K = K;
  res->nposes = c->total;
  res->labels = nlab;
  return BUDE_OK;
}

// Second half: waits for slot c->s()'s rescoring and fills its result.
int screen_end(bude_combined *c) {
  ProteinSlot &ps = c->s();
  BudeCombinedResult *res = &ps.res;
  ps.busy = false;
  if (int st = bude::hip_check(hipEventSynchronize(ps.ev[8]), "hipEventSynchronize")) return st;
  const size_t L = c->nlig;
  ps.nsel.assign(ps.nsel_host, ps.nsel_host + L);
  res->train_ms = elapsed(c, 2, 3);
  if (ps.pipelined) {
    res->score_ms = elapsed_events(ps.ev[3], ps.chunk_ev[0]);
    for (size_t k = 0; k < ps.K; k++) {
      if (k) res->score_ms += elapsed_events(ps.chunk_ev[2 * k - 1], ps.chunk_ev[2 * k]);
      res->select_ms += elapsed_events(ps.chunk_ev[2 * k], ps.chunk_ev[2 * k + 1]);
    }
  } else {
    res->score_ms = elapsed(c, ps.mlp ? 4 : 3, ps.mlp ? 5 : 4) + elapsed(c, 5, 6);
    res->select_ms = elapsed(c, 6, 7);
  }
  res->rescore_ms = elapsed(c, 7, 8);
  res->total_ms = ms_since(ps.t0);
  res->nselected = 0;
  for (size_t l = 0; l < L; l++) res->nselected += ps.nsel[l];
  ps.valid = true;
  return BUDE_OK;
}

}  // namespace

extern "C" int bude_mi355x_combined_screen(bude_combined *c, const BudeAtom *protein, size_t natpro,
                                           BudeCombinedResult *res) {
  if (!c || !res) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_screen: NULL argument");
  if (c->slots[c->cur ^ 1].busy)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_screen: a batch is running");
  if (int st = bude::hip_check(hipSetDevice(c->device), "hipSetDevice")) return st;
  *res = BudeCombinedResult{};
  int st = screen_begin(c, protein, natpro);
  if (!st) st = screen_end(c);
  if (st) {  // leave nothing in flight behind an error
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamSynchronize(c->stream3);
    c->s().busy = false;
    c->s().valid = false;
    return st;
  }
  *res = c->s().res;
  return bude::ok();
}

extern "C" int bude_mi355x_combined_screen_batch(bude_combined *c, size_t nproteins, const BudeAtom *const *proteins,
                                                 const size_t *natpro, BudeCombinedResult *res,
                                                 bude_combined_done_fn done, void *user, double *wall_ms) {
  if (!c || !proteins || !natpro || (nproteins && !res))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_screen_batch: NULL argument");
  if (c->slots[0].busy || c->slots[1].busy)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_screen_batch: a batch is running");
  if (int st = bude::hip_check(hipSetDevice(c->device), "hipSetDevice")) return st;
  const auto t0 = std::chrono::steady_clock::now();
  // protein i begins in one slot while protein i - 1, in the other, is
  // rescored; i - 1 then ends (and its callback runs with it as the protein
  // screened last).  A protein with fallback pairs, or screened by the pose
  // MLP, shares buffers with the next one's scoring and ends first.
  int prev = -1, st = BUDE_OK;
  auto finish = [&](int slot, size_t i) -> int {
    c->cur = slot;
    if (int e = screen_end(c)) return e;
    res[i] = c->s().res;
    if (done) done(user, i);
    return BUDE_OK;
  };
  for (size_t i = 0; i < nproteins && !st; i++) {
    if (prev >= 0 && !c->slots[prev].overlap) {
      st = finish(prev, i - 1);
      if (st) break;
      prev = -1;
    }
    const int slot = prev >= 0 ? prev ^ 1 : c->cur;
    c->cur = slot;
    st = screen_begin(c, proteins[i], natpro[i]);
    if (st) break;
    if (prev >= 0) st = finish(prev, i - 1);
    c->cur = slot;
    prev = slot;
  }
  if (!st && prev >= 0) st = finish(prev, nproteins - 1);
  if (st) {  // leave nothing in flight behind an error
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->stream2);
    (void)hipStreamSynchronize(c->stream3);
    for (ProteinSlot &ps : c->slots) {
      ps.busy = false;
      ps.valid = false;
    }
    return st;
  }
  if (wall_ms) *wall_ms = ms_since(t0);
  return bude::ok();
}

extern "C" int bude_mi355x_combined_set_keep(bude_combined *c, float keep_fraction) {
  if (!c || !(keep_fraction > 0.f) || keep_fraction > 1.f)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_set_keep: keep_fraction must be in (0, 1]");
  if (keep_fraction < c->cfg.gate_top_fraction)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      "bude_mi355x_combined_set_keep: keep_fraction below the configured gate_top_fraction");
  if (c->slots[0].busy || c->slots[1].busy)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_set_keep: not during a batch");
  if (int st = bude::hip_check(hipSetDevice(c->device), "hipSetDevice")) return st;
  if (int st = bude::hip_check(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return st;
  if (int st = bude::hip_check(hipStreamSynchronize(c->stream2), "hipStreamSynchronize")) return st;
  if (int st = plan_selection(c, keep_fraction)) {
    c->s().valid = false;
    return st;
  }
  c->cfg.keep_fraction = keep_fraction;
  return bude::ok();
}

extern "C" int bude_mi355x_combined_pair(bude_combined *c, size_t ligand, BudePairResult *r) {
  if (!c || !r || ligand >= c->nlig) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_pair: bad arguments");
  if (!c->s().valid) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_pair: no protein screened yet");
  r->nposes = uint32_t(c->nposes[ligand]);
  r->nselected = c->s().nsel[ligand];
  r->fallback = c->s().fallback[ligand];
  r->val_rmse = c->s().val_rmse[ligand];
  r->val_rmse_calibrated = c->s().val_rmse_cal[ligand];
  r->gate_z = c->s().gate_z[ligand];
  r->sample_recall = c->s().sample_recall[ligand];
  r->keep_fraction = c->s().keep[ligand];
  return bude::ok();
}

extern "C" int bude_mi355x_combined_selected(bude_combined *c, size_t ligand, uint32_t *idx, float *energies, size_t cap,
                                             size_t *n) {
  if (!c || !n || ligand >= c->nlig)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_selected: bad arguments");
  if (!c->s().valid) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_selected: no protein screened yet");
  const size_t k = c->s().nsel[ligand];
  *n = k;
  if (!idx && !energies) return bude::ok();
  if (cap < k) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_selected: buffer too small");
  if (int st = bude::hip_check(hipSetDevice(c->device), "hipSetDevice")) return st;
  std::vector<uint32_t> id(k);
  std::vector<float> ex(k);
  const size_t po = c->pose_off[ligand];
  if (k) {
    if (int st = bude::hip_check(hipMemcpy(id.data(), static_cast<uint32_t *>(c->s().sel_idx.ptr) + po, k * 4,
                                           hipMemcpyDeviceToHost),
                                 "hipMemcpy"))
      return st;
    if (c->s().fallback[ligand]) {  // the scores of a fallback pair are its exact energies
      std::vector<float> all(c->nposes[ligand]);
      if (int st = bude::hip_check(hipMemcpy(all.data(), static_cast<float *>(c->pred.ptr) + po, all.size() * 4,
                                             hipMemcpyDeviceToHost),
                                   "hipMemcpy"))
        return st;
      for (size_t j = 0; j < k; j++) ex[j] = all[id[j]];
    } else if (int st = bude::hip_check(hipMemcpy(ex.data(), static_cast<float *>(c->s().exact_sel.ptr) + po,
                                                  k * 4, hipMemcpyDeviceToHost),
                                        "hipMemcpy")) {
      return st;
    }
  }
  std::vector<size_t> ord(k);
  std::iota(ord.begin(), ord.end(), size_t(0));
  std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return id[a] < id[b]; });
  for (size_t j = 0; j < k; j++) {
    if (idx) idx[j] = id[ord[j]];
    if (energies) energies[j] = ex[ord[j]];
  }
  return bude::ok();
}

extern "C" int bude_mi355x_combined_labels(const bude_combined *c, size_t ligand, uint32_t *idx, float *energies,
                                           size_t cap, size_t *n) {
  if (!c || !n) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_labels: NULL argument");
  if (!c->s().valid) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_labels: screen a protein first");
  if (ligand >= c->nlig) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_labels: no such ligand");
  const size_t b = c->s().lab_offs[ligand], e = c->s().lab_offs[ligand + 1];
  *n = e - b;
  if (!idx && !energies) return bude::ok();
  if (cap < e - b) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_labels: cap too small");
  for (size_t i = b; i < e; i++) {
    if (idx) idx[i - b] = c->s().lab_j[i];
    if (energies) energies[i - b] = c->s().lab_energy[i];
  }
  return bude::ok();
}

extern "C" int bude_mi355x_combined_scores(bude_combined *c, size_t ligand, float *out) {
  if (!c || !out || ligand >= c->nlig)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_scores: bad arguments");
  if (!c->s().valid) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_scores: no protein screened yet");
  if (c->slots[c->cur ^ 1].busy)  // (the scores and maps are the next protein's by now)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_scores: not for an earlier protein of a batch");
  if (int st = bude::hip_check(hipSetDevice(c->device), "hipSetDevice")) return st;
  return bude::hip_check(hipMemcpy(out, static_cast<float *>(c->pred.ptr) + c->pose_off[ligand],
                                   c->nposes[ligand] * 4, hipMemcpyDeviceToHost),
                         "hipMemcpy");
}

extern "C" int bude_mi355x_combined_exhaustive(bude_combined *c, double *device_ms) {
  if (!c) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_exhaustive: NULL argument");
  if (!c->s().valid)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_exhaustive: no protein screened yet");
  if (int st = bude::hip_check(hipSetDevice(c->device), "hipSetDevice")) return st;
  if (int st = c->exact_all.reserve(c->total * 4)) return st;
  std::vector<Span> all(c->nlig);
  for (size_t l = 0; l < c->nlig; l++) all[l] = Span{l, c->pose_off[l], c->nposes[l]};
  if (int st = record(c, 0)) return st;
  if (int st = exact_ragged(c, all, static_cast<const float *>(c->poses.ptr), c->total,
                            static_cast<float *>(c->exact_all.ptr), c->meta_ex, c->tiles_all))
    return st;
  if (int st = record(c, 1)) return st;
  if (int st = sync(c)) return st;
  if (device_ms) *device_ms = elapsed(c, 0, 1);
  c->have_exact = true;
  c->exact_sum = c->s().protein_sum;
  return bude::ok();
}

extern "C" int bude_mi355x_combined_exact(bude_combined *c, size_t ligand, float *out) {
  if (!c || !out || ligand >= c->nlig)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_exact: bad arguments");
  if (!c->have_exact || c->exact_sum != c->s().protein_sum)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      "bude_mi355x_combined_exact: run bude_mi355x_combined_exhaustive for this protein first");
  if (int st = bude::hip_check(hipSetDevice(c->device), "hipSetDevice")) return st;
  return bude::hip_check(hipMemcpy(out, static_cast<float *>(c->exact_all.ptr) + c->pose_off[ligand],
                                   c->nposes[ligand] * 4, hipMemcpyDeviceToHost),
                         "hipMemcpy");
}

extern "C" int bude_mi355x_combined_recall(bude_combined *c, float top_fraction, float *recall) {
  if (!c || !recall || !(top_fraction > 0.f) || top_fraction > 1.f)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_recall: bad arguments");
  if (!c->have_exact || c->exact_sum != c->s().protein_sum)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      "bude_mi355x_combined_recall: run bude_mi355x_combined_exhaustive for this protein first");
  if (int st = bude::hip_check(hipSetDevice(c->device), "hipSetDevice")) return st;
  // the exact top set of every pair by the same per-segment selection
  std::vector<bude::sel::Segment> segs(c->nlig);
  std::vector<uint32_t> kt(c->nlig);
  for (size_t l = 0; l < c->nlig; l++) {
    kt[l] = uint32_t(std::max(1.0, std::nearbyint(double(top_fraction) * double(c->nposes[l]))));
    segs[l] = bude::sel::Segment{int64_t(c->pose_off[l]), uint32_t(c->nposes[l]), kt[l]};
  }
  bude::sel::Plan plan;
  DeviceBuffer tidx, tn;
  int st = plan.prepare(segs);
  if (!st) st = tidx.alloc(c->total * 4);
  if (!st) st = tn.alloc(c->nlig * 4);
  if (!st)
    st = bude::hip_check(bude::sel::select_lowest_segments(plan, static_cast<const float *>(c->exact_all.ptr),
                                                           static_cast<uint32_t *>(tidx.ptr),
                                                           static_cast<uint32_t *>(tn.ptr), c->stream),
                         "select");
  if (!st) st = sync(c);
  if (st) return st;
  std::vector<uint32_t> tcount(c->nlig);
  if (int e = bude::hip_check(hipMemcpy(tcount.data(), tn.ptr, c->nlig * 4, hipMemcpyDeviceToHost), "hipMemcpy"))
    return e;
  for (size_t l = 0; l < c->nlig; l++) {
    const size_t po = c->pose_off[l];
    std::vector<uint32_t> t(tcount[l]), s(c->s().nsel[l]);
    if (int e = bude::hip_check(hipMemcpy(t.data(), static_cast<uint32_t *>(tidx.ptr) + po, t.size() * 4,
                                          hipMemcpyDeviceToHost),
                                "hipMemcpy"))
      return e;
    if (int e = bude::hip_check(hipMemcpy(s.data(), static_cast<uint32_t *>(c->s().sel_idx.ptr) + po, s.size() * 4,
                                          hipMemcpyDeviceToHost),
                                "hipMemcpy"))
      return e;
    std::vector<char> mark(c->nposes[l], 0);
    for (uint32_t j : s) mark[j] = 1;
    size_t hits = 0;
    for (uint32_t j : t) hits += mark[j];
    recall[l] = float(double(std::min<size_t>(hits, kt[l])) / double(kt[l]));
  }
  return bude::ok();
}

extern "C" int bude_mi355x_combined_grid(bude_combined *c, float lo[3], float *h, int32_t n[3]) {
  if (!c || !lo || !h || !n) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_grid: NULL argument");
  if (c->cfg.surrogate == BUDE_SURROGATE_MLP)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_grid: not a field or learned-field screen");
  for (int k = 0; k < 3; k++) {
    lo[k] = c->grid.lo[k];
    n[k] = c->grid.n[k];
  }
  *h = c->grid.h;
  return bude::ok();
}

extern "C" int bude_mi355x_combined_map(bude_combined *c, int32_t type, float *out) {
  if (!c || !out || type < 0 || size_t(type) >= c->slot_of_type.size())
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_map: bad arguments");
  if (c->cfg.surrogate != BUDE_SURROGATE_FIELD || !c->s().valid)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_map: no field screen run yet");
  if (c->slots[c->cur ^ 1].busy)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_map: not for an earlier protein of a batch");
  const int32_t slot = c->slot_of_type[size_t(type)];
  if (slot < 0) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_combined_map: no ligand atom has this type");
  if (int st = bude::hip_check(hipSetDevice(c->device), "hipSetDevice")) return st;
  const size_t np = c->grid.points();
  return bude::hip_check(hipMemcpy(out, static_cast<float *>(c->maps.ptr) + size_t(slot) * np, np * 4,
                                   hipMemcpyDeviceToHost),
                         "hipMemcpy");
}

extern "C" int bude_mi355x_combined_destroy(bude_combined *c) {
  if (!c) return bude::ok();
  (void)hipSetDevice(c->device);
  delete c;
  return bude::ok();
}
