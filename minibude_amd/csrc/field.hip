// field.hip -- single-atom energy maps and the interpolating pose scorer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "fasten_core.h"
#include "field.h"

namespace bude {
namespace field {

namespace {

using namespace core;

// Distance from a protein atom beyond which all three hinges of a pair with
// coefficients c are exactly zero in either hinge form (the largest zero
// crossing of the terms present, widened by a margin far above the
// rounding of the distances compared with it).  Past it a pair adds
// fma(w, 0, e) == e, so leaving it out changes no bit of the sum.
__device__ __forceinline__ float pair_reach(const PairCoef &c) {
  float r = fmaxf(c.radij, -1.0f / c.neg_rr);
  if (c.wq != 0.f) r = fmaxf(r, fmaxf(c.c0 / -c.neg_e1, c.radij - 1.0f / c.neg_e1));
  if (c.wd != 0.f && c.neg_rd != 0.f) r = fmaxf(r, fmaxf(c.e0 / -c.neg_rd, c.radij - 1.0f / c.neg_rd));
  return r * 1.0001f + 1e-3f;
}

// culled atoms per wave held in LDS before they are scored: 8 KiB per
// work-group, so the maps still fit beside a pinned rescoring grid
constexpr int kMapList = 128;

// The pairs of one group's culled atoms (LDS, in atom order) with this
// work-item's point, into part.
template <int FORM, bool CHG, bool DSLV>
__device__ __forceinline__ void list_pairs(const float4 *list, int n, const PairCoef c, float lx, float ly, float lz,
                                           float &part) {
  float c0 = 0.f, e0 = 0.f;
  if constexpr (FORM == kFormD) {
    if constexpr (CHG) c0 = vgpr_copy(c.c0);
    if constexpr (DSLV) e0 = vgpr_copy(c.e0);
  }
  const float px[1] = {lx}, py[1] = {ly}, pz[1] = {lz};
  float e[1] = {part};
  int t = 0;
  for (; t + 4 <= n; t += 4) {
    const float4 a0 = list[t], a1 = list[t + 1], a2 = list[t + 2], a3 = list[t + 3];
    atom_pairs<1, FORM, CHG, DSLV>(a0, c, c0, e0, px, py, pz, e);
    atom_pairs<1, FORM, CHG, DSLV>(a1, c, c0, e0, px, py, pz, e);
    atom_pairs<1, FORM, CHG, DSLV>(a2, c, c0, e0, px, py, pz, e);
    atom_pairs<1, FORM, CHG, DSLV>(a3, c, c0, e0, px, py, pz, e);
  }
  for (; t < n; ++t) atom_pairs<1, FORM, CHG, DSLV>(list[t], c, c0, e0, px, py, pz, e);
  part = e[0];
}

template <int FORM>
__device__ __forceinline__ void list_dispatch(const float4 *list, int n, const PairCoef c, float lx, float ly,
                                              float lz, float &part) {
  const bool chg = c.wq != 0.f, dslv = c.wd != 0.f;  // wave-uniform
  if (chg && dslv) list_pairs<FORM, true, true>(list, n, c, lx, ly, lz, part);
  else if (chg) list_pairs<FORM, true, false>(list, n, c, lx, ly, lz, part);
  else if (dslv) list_pairs<FORM, false, true>(list, n, c, lx, ly, lz, part);
  else list_pairs<FORM, false, false>(list, n, c, lx, ly, lz, part);
}

// Makes a wave's LDS writes visible to its own later reads.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wave per 4 x 4 x 4 brick of grid points (lane = (x, y, z) within the
// brick) and per map slot.  For each protein group the wave first keeps the
// atoms within reach of its brick (64 atoms per step, compacted in order
// into its LDS list), then scores only those: at 0.5 A spacing a brick
// reaches about a tenth of a protein's atoms.  The pairs left out are
// exactly zero and the kept ones are summed in the same order by the same
// code as the plain kernel's, so every map value is the plain kernel's
// one-atom energy bit for bit.
template <int FORM>
__global__ __launch_bounds__(256) void map_kernel(const float4 *__restrict__ protein,
                                                  const int32_t *__restrict__ group_begin,
                                                  const PairCoef *__restrict__ coef, int ngroups, const Grid g,
                                                  float *__restrict__ maps) {
  // The maps run beside the sampled-pose pass and are on the screen's
  // critical path (maps -> cells -> field scores), the sampled poses are
  // not: raised issue priority lets the maps go first (2-4 % per screen).
  __builtin_amdgcn_s_setprio(2);
  __shared__ float4 lists[4][kMapList];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nb1 = (g.n[1] + 3) / 4, nb2 = (g.n[2] + 3) / 4;
  const int64_t nbricks = int64_t((g.n[0] + 3) / 4) * nb1 * nb2;
  const int64_t brick = int64_t(blockIdx.x) * 4 + wave;
  if (brick >= nbricks) return;  // whole wave
  const int bx = int(brick / (int64_t(nb1) * nb2)), by = int((brick / nb2) % nb1), bz = int(brick % nb2);
  const int ix = 4 * bx + (lane >> 4), iy = 4 * by + ((lane >> 2) & 3), iz = 4 * bz + (lane & 3);
  const bool valid = ix < g.n[0] && iy < g.n[1] && iz < g.n[2];
  const float lx = g.lo[0] + float(min(ix, g.n[0] - 1)) * g.h;
  const float ly = g.lo[1] + float(min(iy, g.n[1] - 1)) * g.h;
  const float lz = g.lo[2] + float(min(iz, g.n[2] - 1)) * g.h;
  // the brick's box (its points, clamped or not, lie inside it)
  const float cx = g.lo[0] + (float(4 * bx) + 1.5f) * g.h, cy = g.lo[1] + (float(4 * by) + 1.5f) * g.h,
              cz = g.lo[2] + (float(4 * bz) + 1.5f) * g.h;
  const float half_diag = 1.5f * 1.7320508f * g.h * 1.0001f;
  float4 *list = lists[wave];
  const PairCoef *c = coef + size_t(blockIdx.y) * size_t(ngroups);
  float e = 0.f;
  for (int k = 0; k < ngroups; ++k) {
    const PairCoef ck = c[k];
    const float reach = pair_reach(ck) + half_diag, reach2 = reach * reach;
    const int i0 = group_begin[k], i1 = group_begin[k + 1];
    float part = 0.f;
    int n = 0;
    for (int base = i0; base < i1; base += 64) {
      const int j = base + lane;
      float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
      bool keep = false;
      if (j < i1) {
        p = protein[j];
        const float dx = p.x - cx, dy = p.y - cy, dz = p.z - cz;
        keep = dx * dx + dy * dy + dz * dz < reach2;
      }
      const uint64_t mask = __ballot(keep);
      if (keep) list[n + __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u))] = p;
      n += __popcll(mask);
      if (n > kMapList - 64 || base + 64 >= i1) {  // (wave-uniform)
        wave_sync();
        list_dispatch<FORM>(list, n, ck, lx, ly, lz, part);
        wave_sync();  // the list is rewritten next
        n = 0;
      }
    }
    e += part;
  }
  if (valid) maps[size_t(blockIdx.y) * g.points() + (size_t(ix) * size_t(g.n[1]) + size_t(iy)) * size_t(g.n[2]) + size_t(iz)] = e * 0.5f;
}

// Cell index and fraction along one axis, clamped to the grid.
__device__ __forceinline__ void cell(float w, float lo, float inv_h, int n, int &i0, float &f) {
  const float u = (w - lo) * inv_h;
  i0 = min(max(int(floorf(u)), 0), n - 2);
  f = __builtin_amdgcn_fmed3f(u - float(i0), 0.0f, 1.0f);
}

union Cell {
  uint4 u;
  _Float16 h[8];
};

__device__ __forceinline__ _Float16 to_half(float v) { return _Float16(fminf(fmaxf(v, -65504.0f), 65504.0f)); }

__global__ __launch_bounds__(256) void pack_kernel(const float *__restrict__ maps, const Grid g,
                                                   uint4 *__restrict__ cells) {
  // on the critical path like the maps: issue ahead of the concurrent
  // sampled-pose pass (0.26 ms beside it, 0.03 ms alone)
  __builtin_amdgcn_s_setprio(2);
  const size_t nc = g.cells(), np = g.points();
  const size_t c = size_t(blockIdx.x) * 256 + threadIdx.x;
  if (c >= nc) return;
  const size_t c2 = size_t(g.n[2] - 1), c12 = size_t(g.n[1] - 1) * c2;
  const size_t ix = c / c12, iy = (c / c2) % size_t(g.n[1] - 1), iz = c % c2;
  const size_t s1 = size_t(g.n[2]), s0 = size_t(g.n[1]) * s1;
  const float *m = maps + size_t(blockIdx.y) * np + ix * s0 + iy * s1 + iz;
  const size_t off[8] = {0, 1, s1, s1 + 1, s0, s0 + 1, s0 + s1, s0 + s1 + 1};
  Cell v;
#pragma unroll
  for (int k = 0; k < 8; k++) v.h[k] = to_half(m[off[k]]);
  cells[size_t(blockIdx.y) * nc + c] = v.u;
}

// The ligand of tile t: the last l with tile_begin[l] <= t (wave-uniform).
__device__ __forceinline__ int tile_ligand(const int64_t *__restrict__ tile_begin, int nligands, int64_t t) {
  int lo = 0, hi = nligands - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tile_begin[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Work-group b scores atom group b mod ngroups of the 4 tiles
// tile0 + 4 (b / ngroups) + wave: with ngroups = 8 every work-group of one
// XCD reads the cells of one group's map slots.
__global__ __launch_bounds__(256) void score_kernel(const uint4 *__restrict__ cells, const Grid g,
                                                    const float4 *__restrict__ grouped,
                                                    const int32_t *__restrict__ group_begin, int ngroups,
                                                    const LigandMeta *__restrict__ meta,
                                                    const int64_t *__restrict__ tile_begin, int nligands,
                                                    const float *__restrict__ poses, size_t stride,
                                                    float *__restrict__ out, float *__restrict__ partial,
                                                    int64_t tile0, int64_t tend, int raise_priority) {
  // Optional issue priority over a concurrent exact pass (the pipelined
  // combined screen rescoring the previous chunk): these waves wait on cell
  // reads most of the time and get the VALU the moment their data arrives.
  if (raise_priority) __builtin_amdgcn_s_setprio(2);
  const int lane = threadIdx.x & 63;
  const int grp = int(blockIdx.x % unsigned(ngroups));
  const int64_t t = tile0 + int64_t(blockIdx.x / unsigned(ngroups)) * 4 + (threadIdx.x >> 6);
  if (t >= tend) return;
  const int l = tile_ligand(tile_begin, nligands, t);
  const LigandMeta m = meta[l];
  const int j = int(t - tile_begin[l]) * 64 + lane;
  const Xform x = make_xform(poses, stride, size_t(m.pose_off) + size_t(j < m.nposes ? j : m.nposes - 1));
  const float inv_h = 1.0f / g.h;
  const size_t nc = g.cells();
  const int c2 = g.n[2] - 1, c12 = (g.n[1] - 1) * c2;
  const int a0 = group_begin[size_t(l) * size_t(ngroups + 1) + size_t(grp)];
  const int a1 = group_begin[size_t(l) * size_t(ngroups + 1) + size_t(grp) + 1];
  float e = 0.f;
  for (int il = a0; il < a1; ++il) {
    const float4 la = grouped[il];
    const float wx = x.m[3] + la.x * x.m[0] + la.y * x.m[1] + la.z * x.m[2];
    const float wy = x.m[7] + la.x * x.m[4] + la.y * x.m[5] + la.z * x.m[6];
    const float wz = x.m[11] + la.x * x.m[8] + la.y * x.m[9] + la.z * x.m[10];
    int ix, iy, iz;
    float fx, fy, fz;
    cell(wx, g.lo[0], inv_h, g.n[0], ix, fx);
    cell(wy, g.lo[1], inv_h, g.n[1], iy, fy);
    cell(wz, g.lo[2], inv_h, g.n[2], iz, fz);
    Cell v;
    v.u = cells[size_t(__float_as_int(la.w)) * nc + size_t(ix * c12 + iy * c2 + iz)];
    float q[8];
#pragma unroll
    for (int k = 0; k < 8; k++) q[k] = float(v.h[k]);  // v000 v001 v010 v011 v100 v101 v110 v111
    const float c00 = __builtin_fmaf(fz, q[1] - q[0], q[0]), c01 = __builtin_fmaf(fz, q[3] - q[2], q[2]);
    const float c10 = __builtin_fmaf(fz, q[5] - q[4], q[4]), c11 = __builtin_fmaf(fz, q[7] - q[6], q[6]);
    const float c0 = __builtin_fmaf(fy, c01 - c00, c00), c1 = __builtin_fmaf(fy, c11 - c10, c10);
    e += __builtin_fmaf(fx, c1 - c0, c0);
  }
  if (ngroups == 1) {
    if (j < m.nposes) out[size_t(m.pose_off) + size_t(j)] = e;
  } else {
    // every lane of the tile, so the sum kernel reads whole rows
    partial[size_t(grp) * size_t(tend - tile0) * 64 + size_t(t - tile0) * 64 + size_t(lane)] = e;
  }
}

// out[pose] = the groups' partial sums of the pose added in group order
// (one wave per tile of tiles [tile0, tend)).
__global__ __launch_bounds__(256) void sum_kernel(const float *__restrict__ partial, int ngroups,
                                                  const LigandMeta *__restrict__ meta,
                                                  const int64_t *__restrict__ tile_begin, int nligands,
                                                  float *__restrict__ out, int64_t tile0, int64_t tend) {
  const int lane = threadIdx.x & 63;
  const int64_t t = tile0 + int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (t >= tend) return;
  const int l = tile_ligand(tile_begin, nligands, t);
  const LigandMeta m = meta[l];
  const int j = int(t - tile_begin[l]) * 64 + lane;
  if (j >= m.nposes) return;
  const size_t n = size_t(tend - tile0) * 64, i = size_t(t - tile0) * 64 + size_t(lane);
  float e = partial[i];
  for (int grp = 1; grp < ngroups; grp++) e += partial[size_t(grp) * n + i];
  out[size_t(m.pose_off) + size_t(j)] = e;
}

}  // namespace

hipError_t build_maps(const float4 *protein, const int32_t *group_begin, const PairCoef *coef, int ngroups,
                      int nslots, const Grid &g, int form, float *maps, hipStream_t s) {
  const size_t np = g.points();
  if (np == 0 || nslots <= 0 || nslots > 65535 || g.n[0] < 2 || g.n[1] < 2 || g.n[2] < 2 || np > 0x7fffffffu ||
      (np + 255) / 256 > 0x7fffffffu)
    return hipErrorInvalidValue;
  const size_t nbricks = size_t((g.n[0] + 3) / 4) * size_t((g.n[1] + 3) / 4) * size_t((g.n[2] + 3) / 4);
  const dim3 grid(unsigned((nbricks + 3) / 4), unsigned(nslots));
  if (form == kFormD)
    hipLaunchKernelGGL(map_kernel<kFormD>, grid, dim3(256), 0, s, protein, group_begin, coef, ngroups, g, maps);
  else
    hipLaunchKernelGGL(map_kernel<kFormB>, grid, dim3(256), 0, s, protein, group_begin, coef, ngroups, g, maps);
  return hipGetLastError();
}

hipError_t pack_cells(const float *maps, int nslots, const Grid &g, uint4 *cells, hipStream_t s) {
  const size_t nc = g.cells();
  if (nc == 0 || nslots <= 0 || nslots > 65535 || (nc + 255) / 256 > 0x7fffffffu || nc > 0x7fffffffu)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(pack_kernel, dim3(unsigned((nc + 255) / 256), unsigned(nslots)), dim3(256), 0, s, maps, g, cells);
  return hipGetLastError();
}

hipError_t score(const ScoreArgs &a, int64_t ntiles, hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  if (a.nligands <= 0 || a.tile0 < 0 || a.grid.n[0] < 2 || a.grid.n[1] < 2 || a.grid.n[2] < 2 ||
      !(a.ngroups == 1 || a.ngroups == kXcdGroups) || (a.ngroups > 1 && !a.partial) ||
      (ntiles + 3) / 4 * a.ngroups > 0x7fffffff)
    return hipErrorInvalidValue;
  const unsigned blocks = unsigned((ntiles + 3) / 4);
  hipLaunchKernelGGL(score_kernel, dim3(blocks * unsigned(a.ngroups)), dim3(256), 0, s, a.cells, a.grid, a.grouped,
                     a.group_begin, a.ngroups, a.meta, a.tile_begin, a.nligands, a.poses, a.stride, a.out, a.partial,
                     a.tile0, a.tile0 + ntiles, int(a.raise_priority));
  if (a.ngroups > 1)
    hipLaunchKernelGGL(sum_kernel, dim3(blocks), dim3(256), 0, s, a.partial, a.ngroups, a.meta, a.tile_begin,
                       a.nligands, a.out, a.tile0, a.tile0 + ntiles);
  return hipGetLastError();
}

namespace {
int slot_bits(float w) {
  int s;
  std::memcpy(&s, &w, sizeof s);
  return s;
}
}  // namespace

void split_atoms(int ngroups, const std::vector<float4> &ligands, const std::vector<int32_t> &lig_off,
                 const std::vector<size_t> &nposes, int nslots, std::vector<float4> &grouped,
                 std::vector<int32_t> &group_begin) {
  const size_t L = nposes.size();
  // weight of every (ligand, slot) item and of every slot
  std::vector<double> item(L * size_t(nslots), 0.0), total(size_t(nslots), 0.0);
  for (size_t l = 0; l < L; l++)
    for (int32_t a = lig_off[l]; a < lig_off[l + 1]; a++) {
      const int s = slot_bits(ligands[size_t(a)].w);
      item[l * size_t(nslots) + size_t(s)] += double(nposes[l]);
      total[size_t(s)] += double(nposes[l]);
    }
  std::vector<int> order(static_cast<size_t>(nslots), 0);
  for (int s = 0; s < nslots; s++) order[size_t(s)] = s;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return total[size_t(x)] > total[size_t(y)]; });
  double sum = 0.0;
  for (double w : total) sum += w;
  // group of each item: walk the items slot by slot, ligand by ligand, and
  // move to the next group once the running weight passes its share
  std::vector<int> group_of(L * size_t(nslots), 0);
  double run = 0.0;
  int grp = 0;
  for (int s : order)
    for (size_t l = 0; l < L; l++) {
      const double w = item[l * size_t(nslots) + size_t(s)];
      if (w == 0.0) continue;
      group_of[l * size_t(nslots) + size_t(s)] = grp;
      run += w;
      while (grp + 1 < ngroups && run >= sum * double(grp + 1) / double(ngroups)) grp++;
    }
  grouped.clear();
  group_begin.assign(L * size_t(ngroups + 1), 0);
  for (size_t l = 0; l < L; l++) {
    for (int g = 0; g < ngroups; g++) {
      group_begin[l * size_t(ngroups + 1) + size_t(g)] = int32_t(grouped.size());
      for (int32_t a = lig_off[l]; a < lig_off[l + 1]; a++)
        if (group_of[l * size_t(nslots) + size_t(slot_bits(ligands[size_t(a)].w))] == g)
          grouped.push_back(ligands[size_t(a)]);
    }
    group_begin[l * size_t(ngroups + 1) + size_t(ngroups)] = int32_t(grouped.size());
  }
}

}  // namespace field
}  // namespace bude
