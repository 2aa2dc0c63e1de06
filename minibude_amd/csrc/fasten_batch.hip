// fasten_batch.hip -- batched ragged fasten_main: many ligands, each with its
// own atom count and pose count, scored against one protein in ONE
// persistent launch.
//
// The reference scores one (ligand, pose set) per fasten() call and its
// driver never runs a partial work-group (src/main.cpp:525-530).  Here the
// unit of work is a tile: 64 x RB consecutive poses of one ligand (one
// wavefront's register-blocked share, exactly the plain kernel's per-wave
// work, src/hip/fasten.hpp:21,48).  Tiles are listed on the host in
// decreasing cost order (cost ~ natlig of the tile's ligand) and every wave
// of a fixed-size grid pulls the next tile from a device counter until the
// list is exhausted, so ligands of 8 and 96 atoms share the machine without
// a tail of idle CUs and pose counts need not divide anything (the last tile
// of a ligand is masked).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bude_internal.h"
#include "fasten_batch.h"
#include "fasten_core.h"

namespace bude {

namespace {

using namespace core;

__global__ void reset_counter_kernel(unsigned int *__restrict__ counter) {
  if (threadIdx.x == 0) counter[0] = 0u;
}

// INDEXED: pose rows through pose_index, tile count from ntiles_dev (both
// written by earlier kernels on the device).
template <int RB, int WG, int FORM, bool INDEXED = false>
__global__ __launch_bounds__(WG) void fasten_batch_kernel(
    const float4 *__restrict__ protein, const int32_t *__restrict__ group_begin, const PairCoef *__restrict__ coef,
    const float4 *__restrict__ ligands, const LigandMeta *__restrict__ meta, const int2 *__restrict__ tiles,
    const float *__restrict__ poses, float *__restrict__ energies, unsigned int *__restrict__ counter, size_t stride,
    int ngroups, int ntiles, const uint32_t *__restrict__ pose_index, const int32_t *__restrict__ ntiles_dev) {
  const int lane = threadIdx.x & 63;
  if constexpr (INDEXED) ntiles = *ntiles_dev;
  for (;;) {
    // One atomic per wave; every wave leaves once the list is exhausted.
    unsigned int t = 0;
    if (lane == 0) t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    t = __builtin_amdgcn_readfirstlane(__shfl(t, 0));
    if (t >= unsigned(ntiles)) break;

    const int2 tile = tiles[t];
    const LigandMeta m = meta[tile.x];
    Xform tr[RB];
    float e[RB];
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      const int j = tile.y + lane + k * 64;
      size_t q = size_t(m.pose_off) + size_t(j < m.nposes ? j : m.nposes - 1);
      if constexpr (INDEXED) q = size_t(m.pose_off) + pose_index[q];
      tr[k] = make_xform(poses, stride, q);
      e[k] = 0.f;
    }
    for (int il = 0; il < m.natlig; ++il) {
      const float4 la = ligands[m.lig_off + il];
      float lx[RB], ly[RB], lz[RB];
#pragma unroll
      for (int k = 0; k < RB; ++k) {
        const float *mm = tr[k].m;
        lx[k] = mm[3] + la.x * mm[0] + la.y * mm[1] + la.z * mm[2];
        ly[k] = mm[7] + la.x * mm[4] + la.y * mm[5] + la.z * mm[6];
        lz[k] = mm[11] + la.x * mm[8] + la.y * mm[9] + la.z * mm[10];
      }
      const PairCoef *cl = coef + size_t(m.coef_off) + size_t(il) * ngroups;
      for (int g = 0; g < ngroups; ++g)
        group_dispatch<RB, FORM>(protein, group_begin[g], group_begin[g + 1], cl[g], lx, ly, lz, e);
    }
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      const int j = tile.y + lane + k * 64;
      if (j < m.nposes) energies[size_t(m.pose_off) + size_t(j)] = e[k] * 0.5f;
    }
  }
}

// Split form for short passes (a few tiles per SIMD, where a wave scoring
// 64 poses runs nearly alone on its SIMD and its scalar-load latency is not
// hidden): a wave scores 16 poses, four lanes per pose, lane s of a pose
// taking the ligand atoms il = s, s + 4, ...; the four partial energies are
// then added across the lanes.  Four times the waves for the same work, so
// the pass fills the SIMDs.  The pair coefficients differ per lane and sit
// in VGPRs (the next group's are loaded while the current group runs);
// protein atoms stay wave-uniform.  The sums are grouped differently from
// the tiled kernels', so energies agree with theirs to rounding, not bit
// for bit.  One wave per tile (no work queue); tiles are (ligand, first
// pose) pairs of 16 poses.
template <int FORM>
__global__ __launch_bounds__(256) void fasten_split_kernel(
    const float4 *__restrict__ protein, const int32_t *__restrict__ group_begin, const PairCoef *__restrict__ coef,
    const float4 *__restrict__ ligands, const LigandMeta *__restrict__ meta, const int2 *__restrict__ tiles,
    const float *__restrict__ poses, float *__restrict__ energies, size_t stride, int ngroups, int ntiles) {
  const int t = int(blockIdx.x) * 4 + int(threadIdx.x >> 6);  // wave-uniform
  if (t >= ntiles) return;
  const int lane = threadIdx.x & 63, p = lane >> 2, sl = lane & 3;
  const int2 tile = tiles[t];
  const LigandMeta m = meta[tile.x];
  const int j = tile.y + p;
  const Xform tr = make_xform(poses, stride, size_t(m.pose_off) + size_t(j < m.nposes ? j : m.nposes - 1));
  float e = 0.f;
  for (int base = 0; base < m.natlig; base += 4) {  // uniform trip count
    const int il = base + sl;
    const bool on = il < m.natlig;
    const int ia = on ? il : m.natlig - 1;  // (inactive lanes score a real atom and drop it)
    const float4 la = ligands[m.lig_off + ia];
    const float *mm = tr.m;
    const float lx = mm[3] + la.x * mm[0] + la.y * mm[1] + la.z * mm[2];
    const float ly = mm[7] + la.x * mm[4] + la.y * mm[5] + la.z * mm[6];
    const float lz = mm[11] + la.x * mm[8] + la.y * mm[9] + la.z * mm[10];
    const PairCoef *cl = coef + size_t(m.coef_off) + size_t(ia) * ngroups;
    PairCoef c = cl[0];
    float sum = 0.f;
    for (int g = 0; g < ngroups; ++g) {
      const PairCoef cn = cl[g + 1 < ngroups ? g + 1 : g];
      float part = 0.f;
      const int i1 = group_begin[g + 1];
#pragma unroll 4
      for (int i = group_begin[g]; i < i1; ++i) {
        const float4 pa = protein[i];
        const float dx = lx - pa.x, dy = ly - pa.y, dz = lz - pa.z;
        const float d = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
        if constexpr (FORM == kFormB) {
          const float b = d - c.radij;
          part = __builtin_fmaf(2.0f * kHardness, sat(b * c.neg_rr), part);
          part = __builtin_fmaf(c.wq, sat(__builtin_fmaf(b, c.neg_e1, 1.0f)), part);
          part = __builtin_fmaf(c.wd, sat(__builtin_fmaf(b, c.neg_rd, 1.0f)), part);
        } else {
          part = __builtin_fmaf(2.0f * kHardness, sat(__builtin_fmaf(d, c.neg_rr, 1.0f)), part);
          part = __builtin_fmaf(c.wq, sat(__builtin_fmaf(d, c.neg_e1, c.c0)), part);
          part = __builtin_fmaf(c.wd, sat(__builtin_fmaf(d, c.neg_rd, c.e0)), part);
        }
      }
      sum += part;
      c = cn;
    }
    if (on) e += sum;
  }
  e += __shfl_xor(e, 1);
  e += __shfl_xor(e, 2);
  if (sl == 0 && j < m.nposes) energies[size_t(m.pose_off) + size_t(j)] = e * 0.5f;
}

// Rescoring work tables of the ligands order[0..n) from their selection
// counts, on the device right after the selection:
// meta[l].nposes = sel_n[l], tiles (l, 64 i) for i < ceil(sel_n[l] / 64) in
// list order from tiles[0], and the tile count in *ntiles.  One work-group
// of 256; the host never waits for the counts.
__global__ __launch_bounds__(256) void rescore_tables_kernel(const int32_t *__restrict__ order, int n,
                                                             const uint32_t *__restrict__ sel_n,
                                                             LigandMeta *__restrict__ meta,
                                                             int2 *__restrict__ tiles, int32_t *__restrict__ ntiles) {
  __shared__ int32_t lig[256], cnt[256], off[256], wsum[4];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t base = 0;
  for (int b = 0; b < n; b += 256) {
    const int i = b + int(threadIdx.x);
    int32_t c = 0, l = 0;
    if (i < n) {
      l = order[i];
      const uint32_t k = sel_n[l];
      meta[l].nposes = int32_t(k);
      c = int32_t((k + 63) / 64);
    }
    int32_t x = c;  // inclusive scan over the wave, then over the four waves
    for (int d = 1; d < 64; d <<= 1) {
      const int32_t y = __shfl_up(x, d);
      if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int32_t wb = 0;
    for (int q = 0; q < w; q++) wb += wsum[q];
    lig[threadIdx.x] = l;
    cnt[threadIdx.x] = c;
    off[threadIdx.x] = base + wb + x - c;
    const int32_t sum = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    const int m = min(256, n - b);
    for (int q = 0; q < m; q++)
      for (int t = int(threadIdx.x); t < cnt[q]; t += 256) tiles[off[q] + t] = make_int2(lig[q], 64 * t);
    base += sum;
    __syncthreads();  // lig/cnt/off/wsum are rewritten by the next batch
  }
  if (threadIdx.x == 0) *ntiles = base;
}

template <int RB, int WG>
hipError_t launch_rb_wg(const BatchArgs &a, int blocks, hipStream_t s) {
  if (a.form == kFormD)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(fasten_batch_kernel<RB, WG, kFormD>), dim3(unsigned(blocks)), dim3(WG),
                       a.lds_bytes, s, a.protein, a.group_begin, a.coef, a.ligands, a.meta, a.tiles, a.poses,
                       a.energies, a.counter, a.stride, a.ngroups, a.ntiles, nullptr, nullptr);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(fasten_batch_kernel<RB, WG, kFormB>), dim3(unsigned(blocks)), dim3(WG),
                       a.lds_bytes, s, a.protein, a.group_begin, a.coef, a.ligands, a.meta, a.tiles, a.poses,
                       a.energies, a.counter, a.stride, a.ngroups, a.ntiles, nullptr, nullptr);
  return hipGetLastError();
}

template <int RB>
hipError_t launch_rb(const BatchArgs &a, int wg, int blocks, hipStream_t s) {
  switch (wg) {
    case 64: return launch_rb_wg<RB, 64>(a, blocks, s);
    case 128: return launch_rb_wg<RB, 128>(a, blocks, s);
    case 256: return launch_rb_wg<RB, 256>(a, blocks, s);
    case 512: return launch_rb_wg<RB, 512>(a, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

bool batch_shape_supported(int rb, int wgsize) {
  return (rb == 1 || rb == 2 || rb == 4 || rb == 8) && (wgsize == 64 || wgsize == 128 || wgsize == 256 ||
                                                         wgsize == 512);
}

PinnedGrid pinned_grid(int cus, size_t lds_per_cu, int wps) {
  PinnedGrid g{256, std::max(1, cus * wps), 0u};
  if (wps < 1) return g;
  // the least whole KiB (at most the 64 KiB a work-group may allocate) with
  // which wps + 1 work-groups do not fit, while wps fit with 2 KiB of slack
  // each for the allocation granularity; the rest of the CU's LDS stays free
  // for kernels running beside the grid (the selection's histograms)
  const size_t b = (lds_per_cu / size_t(wps + 1) + 1024) & ~size_t(1023);
  if (b <= 65536 && size_t(wps) * (b + 2048) <= lds_per_cu && size_t(wps + 1) * b > lds_per_cu)
    g.lds_bytes = unsigned(b);
  return g;
}

hipError_t launch_split(const BatchArgs &a, hipStream_t s) {
  if (a.ntiles <= 0) return hipSuccess;
  const unsigned blocks = unsigned((a.ntiles + 3) / 4);
  if (a.form == kFormD)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(fasten_split_kernel<kFormD>), dim3(blocks), dim3(256), 0, s, a.protein,
                       a.group_begin, a.coef, a.ligands, a.meta, a.tiles, a.poses, a.energies, a.stride, a.ngroups,
                       a.ntiles);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(fasten_split_kernel<kFormB>), dim3(blocks), dim3(256), 0, s, a.protein,
                       a.group_begin, a.coef, a.ligands, a.meta, a.tiles, a.poses, a.energies, a.stride, a.ngroups,
                       a.ntiles);
  return hipGetLastError();
}

hipError_t build_rescore_tables(const int32_t *order, int n, const uint32_t *sel_n, LigandMeta *meta, int2 *tiles,
                                int32_t *ntiles, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(rescore_tables_kernel, dim3(1), dim3(256), 0, s, order, n, sel_n, meta, tiles, ntiles);
  return hipGetLastError();
}

hipError_t launch_batch_indexed(const BatchArgs &a, int blocks, hipStream_t s) {
  if (!a.pose_index || !a.ntiles_dev || blocks < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, a.counter);
  if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
  if (a.form == kFormD)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(fasten_batch_kernel<1, 256, kFormD, true>), dim3(unsigned(blocks)), dim3(256),
                       a.lds_bytes, s, a.protein, a.group_begin, a.coef, a.ligands, a.meta, a.tiles, a.poses,
                       a.energies, a.counter, a.stride, a.ngroups, 0, a.pose_index, a.ntiles_dev);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(fasten_batch_kernel<1, 256, kFormB, true>), dim3(unsigned(blocks)), dim3(256),
                       a.lds_bytes, s, a.protein, a.group_begin, a.coef, a.ligands, a.meta, a.tiles, a.poses,
                       a.energies, a.counter, a.stride, a.ngroups, 0, a.pose_index, a.ntiles_dev);
  return hipGetLastError();
}

hipError_t launch_batch(const BatchArgs &a, int rb, int wgsize, int blocks, hipStream_t s) {
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, s, a.counter);
  if (hipError_t e = hipGetLastError(); e != hipSuccess) return e;
  switch (rb) {
    case 1: return launch_rb<1>(a, wgsize, blocks, s);
    case 2: return launch_rb<2>(a, wgsize, blocks, s);
    case 4: return launch_rb<4>(a, wgsize, blocks, s);
    case 8: return launch_rb<8>(a, wgsize, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace bude
