// capi_atomnet.hip -- C ABI of the learned per-atom field's parts: exact
// one-atom energies at arbitrary points (the labels it trains on), one
// forward + backward pass of its training kernels, and its fused scorer
// over the poses of one ligand.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "atomnet.h"
#include "bude_internal.h"
#include "hip_util.h"

namespace {

int set_device(int device, const char *who) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || device < 0 || device >= ndev)
    return bude::fail(BUDE_ERR_NO_DEVICE, std::string(who) + ": no such HIP device");
  return bude::hip_check(hipSetDevice(device), "hipSetDevice");
}

int make_net(bude::atomnet::Net &net, const float lo[3], float h, const int32_t n[3], const float *grid, size_t nslots,
             const float *const weights[3], const float *const biases[3], hipStream_t s, const char *who) {
  using namespace bude::atomnet;
  if (!lo || !n || !grid || !weights || !biases || !(h > 0.f))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": NULL or empty network");
  for (int k = 0; k < 3; k++)
    if (!weights[k] || !biases[k] || n[k] < 2 || n[k] > 4096)
      return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": NULL weights or a grid axis outside [2, 4096]");
  if (nslots == 0 || nslots > size_t(kMaxSlots))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      std::string(who) + ": nslots must be in [1, " + std::to_string(kMaxSlots) + "]");
  bude::field::Grid g{};
  for (int k = 0; k < 3; k++) {
    g.lo[k] = lo[k];
    g.n[k] = n[k];
  }
  g.h = h;
  net.grid = g;
  net.nslots = int(nslots);
  net.k0 = input_width(int(nslots));
  net.nodes = g.points();
  // W2 [kH][1] and b2 [1] from the caller into column 0 of [kH][kOut] / [kOut]
  std::vector<float> w2(size_t(kH) * kOut, 0.f), b2(kOut, 0.f);
  for (int r = 0; r < kH; r++) w2[size_t(r) * kOut] = weights[2][r];
  b2[0] = biases[2][0];
  const float *w[3] = {weights[0], weights[1], w2.data()}, *b[3] = {biases[0], biases[1], b2.data()};
  if (int st = set_params(net, grid, w, b, s)) return st;
  return finalize(net, s);
}

}  // namespace

extern "C" int bude_mi355x_atom_energies(int device, const BudeAtom *protein, size_t natpro,
                                         const BudeFFParams *forcefield, size_t ntypes, const float *points,
                                         const int32_t *types, size_t n, float *out) {
  const char *who = "bude_mi355x_atom_energies";
  if (!points || !types || !out || n == 0)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": NULL or empty points");
  if (n > (size_t(1) << 28)) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": more than 2^28 points");
  // (the protein stands in for the ligand in the deck check)
  if (int st = bude::check_deck(protein, natpro, protein, natpro, forcefield, ntypes, who)) return st;
  for (size_t i = 0; i < n; i++)
    if (types[i] < 0 || size_t(types[i]) >= ntypes)
      return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": point " + std::to_string(i) + " has type " +
                                                       std::to_string(types[i]) + " outside the forcefield");
  if (int st = set_device(device, who)) return st;
  const bude::ProteinGroups groups = bude::group_protein(protein, natpro);
  const size_t ng = groups.ngroups();
  std::vector<BudeAtom> probe(ntypes);
  for (size_t t = 0; t < ntypes; t++) probe[t] = BudeAtom{0.f, 0.f, 0.f, int32_t(t)};
  const std::vector<bude::PairCoef> coef = bude::coef_table(probe.data(), ntypes, groups, forcefield);
  // points grouped by type into whole 64-point blocks (a block's padding
  // repeats its last point)
  std::vector<std::vector<size_t>> by(ntypes);
  for (size_t i = 0; i < n; i++) by[size_t(types[i])].push_back(i);
  std::vector<float> pts;
  std::vector<size_t> where;
  for (size_t t = 0; t < ntypes; t++) {
    if (by[t].empty()) continue;
    float slot;
    const int32_t ti = int32_t(t);
    std::memcpy(&slot, &ti, 4);
    const size_t m = by[t].size(), padded = (m + 63) / 64 * 64;
    for (size_t k = 0; k < padded; k++) {
      const size_t i = by[t][std::min(k, m - 1)];
      pts.insert(pts.end(), {points[3 * i], points[3 * i + 1], points[3 * i + 2], slot});
      where.push_back(k < m ? i : size_t(-1));
    }
  }
  const size_t np = where.size();
  bude::DeviceBuffer dprot, dbegin, dcoef, dpts, dout;
  int st = dprot.upload(groups.xyzw.data(), groups.xyzw.size() * 4);
  if (!st) st = dbegin.upload(groups.begin.data(), groups.begin.size() * 4);
  if (!st) st = dcoef.upload(coef.data(), coef.size() * sizeof(bude::PairCoef));
  if (!st) st = dpts.upload(pts.data(), pts.size() * 4);
  if (!st) st = dout.alloc(np * 4);
  if (!st)
    st = bude::hip_check(bude::atomnet::label_points(static_cast<const float4 *>(dprot.ptr),
                                                     static_cast<const int32_t *>(dbegin.ptr),
                                                     static_cast<const bude::PairCoef *>(dcoef.ptr), int(ng),
                                                     static_cast<const float4 *>(dpts.ptr), np, bude::plain_form(),
                                                     static_cast<float *>(dout.ptr), nullptr),
                         "atom energies");
  std::vector<float> e(np);
  if (!st) st = bude::hip_check(hipMemcpy(e.data(), dout.ptr, np * 4, hipMemcpyDeviceToHost), "hipMemcpy");
  if (st) return st;
  for (size_t k = 0; k < np; k++)
    if (where[k] != size_t(-1)) out[where[k]] = e[k];
  return bude::ok();
}

extern "C" int bude_mi355x_atomnet_forward_backward(int device, size_t batch, const float *points,
                                                    const int32_t *slots, const float *target, const float grid_lo[3],
                                                    float grid_h, const int32_t grid_n[3], const float *grid,
                                                    size_t nslots, const float *const weights[3],
                                                    const float *const biases[3], float *pred, float *loss,
                                                    float *grad_grid, float *const grad_w[3], float *const grad_b[3]) {
  using namespace bude::atomnet;
  const char *who = "bude_mi355x_atomnet_forward_backward";
  if (batch == 0 || batch % 64 || !points || !slots || !target || !pred || !loss || !grad_grid || !grad_w || !grad_b)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": batch must be a multiple of 64, no NULL arrays");
  for (size_t i = 0; i < batch; i++)
    if (slots[i] < 0 || size_t(slots[i]) >= nslots)
      return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": slot out of range");
  if (int st = set_device(device, who)) return st;
  Net net;
  if (int st = make_net(net, grid_lo, grid_h, grid_n, grid, nslots, weights, biases, nullptr, who)) return st;
  std::vector<float> pts(batch * 4);
  for (size_t i = 0; i < batch; i++) {
    pts[4 * i] = points[3 * i];
    pts[4 * i + 1] = points[3 * i + 1];
    pts[4 * i + 2] = points[3 * i + 2];
    std::memcpy(&pts[4 * i + 3], &slots[i], 4);
  }
  bude::DeviceBuffer dpts, dtarget, dpred, dgrid;
  int st = dpts.upload(pts.data(), pts.size() * 4);
  if (!st) st = dtarget.upload(target, batch * 4);
  if (!st) st = dpred.alloc(batch * 4);
  if (!st) st = dgrid.alloc(net.nodes * kF * 4);
  if (!st)
    st = forward_backward(net, static_cast<const float4 *>(dpts.ptr), static_cast<const float *>(dtarget.ptr), batch,
                          static_cast<float *>(dpred.ptr), loss, static_cast<float *>(dgrid.ptr), nullptr);
  if (!st) st = bude::hip_check(hipMemcpy(pred, dpred.ptr, batch * 4, hipMemcpyDeviceToHost), "hipMemcpy");
  if (!st)
    st = bude::hip_check(hipMemcpy(grad_grid, dgrid.ptr, net.nodes * kF * 4, hipMemcpyDeviceToHost), "hipMemcpy");
  for (int l = 0; l < 3 && !st; l++) {
    const int K = layer_k(net, l), N = layer_n(l);
    std::vector<float> gw(size_t(K) * N), gb(N);
    st = bude::hip_check(hipMemcpy(gw.data(), net.gw[l].ptr, gw.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if (!st) st = bude::hip_check(hipMemcpy(gb.data(), net.gb[l].ptr, gb.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if (st) break;
    if (l < 2) {
      std::copy(gw.begin(), gw.end(), grad_w[l]);
      std::copy(gb.begin(), gb.end(), grad_b[l]);
    } else {
      for (int r = 0; r < kH; r++) grad_w[2][r] = gw[size_t(r) * kOut];
      grad_b[2][0] = gb[0];
    }
  }
  return st;
}

extern "C" int bude_mi355x_atomnet_score(int device, const BudeAtom *ligand, size_t natlig,
                                         const float *const poses[6], size_t nposes, const float grid_lo[3],
                                         float grid_h, const int32_t grid_n[3], const float *grid, size_t nslots,
                                         const float *const weights[3], const float *const biases[3], float sigma,
                                         float mu, float *scores) {
  using namespace bude::atomnet;
  const char *who = "bude_mi355x_atomnet_score";
  if (!ligand || natlig == 0 || !poses || nposes == 0 || nposes > 0x7fffffffu || !scores)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": empty ligand or poses");
  for (int k = 0; k < 6; k++)
    if (!poses[k]) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": pose row is NULL");
  for (size_t a = 0; a < natlig; a++)
    if (ligand[a].type < 0 || size_t(ligand[a].type) >= nslots)
      return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": ligand atom slot out of range");
  if (int st = set_device(device, who)) return st;
  Net net;
  if (int st = make_net(net, grid_lo, grid_h, grid_n, grid, nslots, weights, biases, nullptr, who)) return st;
  std::vector<float> lig4(natlig * 4);
  for (size_t a = 0; a < natlig; a++) {
    lig4[4 * a] = ligand[a].x;
    lig4[4 * a + 1] = ligand[a].y;
    lig4[4 * a + 2] = ligand[a].z;
    std::memcpy(&lig4[4 * a + 3], &ligand[a].type, 4);
  }
  const bude::LigandMeta meta{0, int32_t(natlig), 0, int32_t(nposes), 0, 0};
  const int64_t tb[2] = {0, int64_t((nposes + 63) / 64)};
  bude::DeviceBuffer dlig, dmeta, dtb, dposes, dout;
  int st = dlig.upload(lig4.data(), lig4.size() * 4);
  if (!st) st = dmeta.upload(&meta, sizeof(meta));
  if (!st) st = dtb.upload(tb, sizeof(tb));
  if (!st) st = dposes.alloc(6 * nposes * 4);
  for (int k = 0; k < 6 && !st; k++)
    st = bude::hip_check(hipMemcpy(static_cast<float *>(dposes.ptr) + size_t(k) * nposes, poses[k], nposes * 4,
                                   hipMemcpyHostToDevice),
                         "hipMemcpy");
  if (!st) st = dout.alloc(nposes * 4);
  if (st) return st;
  ScoreArgs a = score_args(net);
  a.sigma = sigma;
  a.mu = mu;
  a.ligands = static_cast<const float4 *>(dlig.ptr);
  a.meta = static_cast<const bude::LigandMeta *>(dmeta.ptr);
  a.tile_begin = static_cast<const int64_t *>(dtb.ptr);
  a.nligands = 1;
  a.poses = static_cast<const float *>(dposes.ptr);
  a.stride = nposes;
  a.out = static_cast<float *>(dout.ptr);
  a.tile0 = 0;
  if (int st2 = bude::hip_check(score(a, tb[1], nullptr), "atomnet score")) return st2;
  return bude::hip_check(hipMemcpy(scores, dout.ptr, nposes * 4, hipMemcpyDeviceToHost), "hipMemcpy");
}
