// field.h -- per-protein energy maps of single ligand atoms, and the pose
// scorer that sums trilinear interpolations of them over a ligand's atoms.
//
// fasten_main's energy is a sum over ligand atoms of terms that depend only
// on that atom's forcefield type and its position after the pose transform
// (src/serial/fasten.hpp:56-131: the pair parameters of a3 depend on the
// (ligand type, protein type) pair, the pair energy on the distance).  So
// for one protein, E(pose) = sum_il F[type(il)](R x_il + t) with
// F[t](p) = the energy of one ligand atom of type t at p.  The maps hold F
// exactly (computed by the same pair loop as the plain kernel) at the points
// of a regular grid; the scorer interpolates between them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "bude_internal.h"
#include "fasten_batch.h"

namespace bude {
namespace field {

struct Grid {
  float lo[3];  // coordinates of grid point (0, 0, 0)
  float h;      // spacing
  int n[3];     // points per axis, each >= 2
  __host__ __device__ size_t points() const { return size_t(n[0]) * size_t(n[1]) * size_t(n[2]); }
  __host__ __device__ size_t cells() const { return size_t(n[0] - 1) * size_t(n[1] - 1) * size_t(n[2] - 1); }
};

// maps[(s * n0 + ix) * n1 + iy) * n2 + iz] = energy of one ligand atom whose
// pair coefficients are coef[s * ngroups + g] (g over the protein groups) at
// point lo + h * (ix, iy, iz): 0.5 * the sum of the pair energies, groups in
// order, exactly as the plain kernel sums one ligand atom.
hipError_t build_maps(const float4 *protein, const int32_t *group_begin, const PairCoef *coef, int ngroups,
                      int nslots, const Grid &g, int form, float *maps, hipStream_t s);

// Cells of the maps: cell (ix, iy, iz) of slot s holds its 8 corner values
// [v000 v001 v010 v011 v100 v101 v110 v111] (index bits x y z) as FP16
// (round to nearest even, clamped to +-65504) in 16 contiguous bytes, so an
// interpolation is one 16-byte read instead of 8 scattered ones and the
// cells of every slot stay small enough for the device's last-level cache.
// cells has nslots * g.cells() entries.
hipError_t pack_cells(const float *maps, int nslots, const Grid &g, uint4 *cells, hipStream_t s);

// The device's work-groups go round-robin to its 8 XCDs (work-group i to
// XCD i mod 8), each with its own 4 MB L2.  The cells of every map slot
// (143 MB for 34 slots at 0.5 A) do not fit one L2, and a wave's 64 poses
// put its atoms anywhere in the box, so every cell read misses L2.  The
// scorer therefore splits each ligand's atoms into kXcdGroups groups by map
// slot (split_atoms) and gives group g to work-groups i with i mod 8 == g:
// each L2 then caches only its group's slots.  The groups' partial sums of
// a pose are added in group order by a second kernel.
constexpr int kXcdGroups = 8;

// Assignment of the ligands' atoms to `ngroups` groups (1 or kXcdGroups):
// items (ligand, slot) weighted by atoms x poses, slots in decreasing order
// of their total weight (ties by slot), each slot's ligands in order, filled
// into groups 0, 1, ... up to an equal share of the total each, so a group
// holds a few whole slots and at most two partial ones.  grouped holds every
// ligand's atoms, group by group, each group's in their ligand order;
// group_begin[l * (ngroups + 1) + g] is the index in grouped of ligand l's
// first atom of group g (group_begin[l * (ngroups + 1) + ngroups] its end).
// ligands: x, y, z, map slot (int bits); ligand l's atoms are
// [lig_off[l], lig_off[l + 1]); nposes[l] its pose count.
void split_atoms(int ngroups, const std::vector<float4> &ligands, const std::vector<int32_t> &lig_off,
                 const std::vector<size_t> &nposes, int nslots, std::vector<float4> &grouped,
                 std::vector<int32_t> &group_begin);

struct ScoreArgs {
  const uint4 *cells;          // nslots * grid.cells() packed cells (pack_cells)
  Grid grid;
  const float4 *grouped;       // x, y, z, map slot (int bits), grouped by split_atoms
  const int32_t *group_begin;  // nligands * (ngroups + 1) (split_atoms)
  int ngroups;                 // 1 or kXcdGroups
  const LigandMeta *meta;      // nposes, pose_off of every ligand
  const int64_t *tile_begin;   // nligands + 1: first 64-pose tile of each ligand
  int nligands;
  const float *poses;          // 6 rows of `stride` floats
  size_t stride;
  float *out;                  // out[meta.pose_off + j]
  float *partial;              // ngroups > 1: ngroups * ntiles * 64 floats of scratch
  int64_t tile0;               // first tile scored
  bool raise_priority;         // waves issue ahead of a concurrent kernel's
};

// Interpolated energy of every pose of tiles [tile0, tile0 + ntiles): one
// wave per tile of 64 poses of one ligand and group of its atoms; every tile
// is tile0 = 0, ntiles = tile_begin[nligands].
hipError_t score(const ScoreArgs &a, int64_t ntiles, hipStream_t s);

}  // namespace field
}  // namespace bude
