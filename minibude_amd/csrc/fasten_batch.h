// fasten_batch.h -- launch interface of the batched ragged fasten_main.
#pragma once

#include <hip/hip_runtime.h>

#include "bude_internal.h"

namespace bude {

// Per-ligand offsets into the concatenated batch arrays.
struct alignas(32) LigandMeta {
  int32_t lig_off;   // first atom in the concatenated ligand atoms
  int32_t natlig;
  int32_t coef_off;  // first PairCoef of this ligand ([natlig][ngroups])
  int32_t nposes;
  int64_t pose_off;  // first pose in the concatenated pose rows / energies
  int64_t pad;
};
static_assert(sizeof(LigandMeta) == 32, "LigandMeta is 32 bytes");

struct BatchArgs {
  const float4 *protein;        // grouped protein atoms
  const int32_t *group_begin;   // ngroups + 1
  const PairCoef *coef;         // concatenated per-ligand tables
  const float4 *ligands;        // concatenated ligand atoms (x, y, z, 0)
  const LigandMeta *meta;       // nligands
  const int2 *tiles;            // ntiles (ligand, first pose), heaviest first
  const float *poses;           // 6 rows of `stride` floats (all ligands' poses)
  float *energies;              // stride
  unsigned int *counter;        // work-queue head, reset by every launch
  size_t stride;
  int ngroups;
  int ntiles;
  int form;                     // hinge form (plain_form())
  unsigned lds_bytes;           // dynamic LDS per work-group, unused by the kernel: sized so that exactly
                                // n work-groups fit on a CU, it pins the waves per SIMD (0: no pinning)
  // launch_batch_indexed only: row j of ligand l is pose meta[l].pose_off + pose_index[meta[l].pose_off + j]
  // (energies still at meta[l].pose_off + j), and the tile count is read from device memory
  const uint32_t *pose_index;
  const int32_t *ntiles_dev;
};

bool batch_shape_supported(int rb, int wgsize);

// A persistent grid pinned to `wps` waves on every SIMD: work-groups of 256
// work-items (one wave per SIMD), `wps` of them per CU, each carrying unused
// dynamic LDS sized so that no more than `wps` fit on a CU.  Left to itself
// the dispatcher can stack waves unevenly over the SIMDs of a launch that
// does not fill the machine, and a SIMD holding one or two waves runs its
// tiles at a fraction of the rate of one holding four.  lds_bytes is 0 (no
// pinning) when the device's LDS does not allow it.
struct PinnedGrid {
  int wgsize, blocks;
  unsigned lds_bytes;
};
PinnedGrid pinned_grid(int cus, size_t lds_per_cu, int wps);

// Counter reset + one persistent launch of `blocks` work-groups of `wgsize`
// work-items, each wave scoring tiles of 64 * rb poses.
hipError_t launch_batch(const BatchArgs &a, int rb, int wgsize, int blocks, hipStream_t s);

// Work tables of a rescoring pass over per-segment selections, built on the
// device from the selection counts: for the ligands order[0..n) (list
// order, e.g. heaviest first) meta[l].nposes = sel_n[l] (every other field
// of meta[l] set beforehand), tiles (l, 64 i) for i < ceil(sel_n[l] / 64)
// from tiles[0] (room for the worst case is the caller's), and the tile
// count in *ntiles; one work-group, on stream s after the selection.
hipError_t build_rescore_tables(const int32_t *order, int n, const uint32_t *sel_n, LigandMeta *meta, int2 *tiles,
                                int32_t *ntiles, hipStream_t s);

// launch_batch over rows that are a per-segment selection (pose_index) with
// tiles and tile count written by an earlier kernel (ntiles_dev): one pose
// per work-item (rb 1), work-groups of 256.  Energies are bit-equal to
// launch_batch's for the same poses.
hipError_t launch_batch_indexed(const BatchArgs &a, int blocks, hipStream_t s);

// The split form for short passes: a.tiles are (ligand, first pose) pairs of
// 16 poses, one wave each, four lanes per pose; a.counter and a.lds_bytes
// are not used.  Energies agree with launch_batch's to rounding (the sums
// are grouped by ligand-atom slice), not bit for bit.
constexpr int kSplitTile = 16;
hipError_t launch_split(const BatchArgs &a, hipStream_t s);

}  // namespace bude
