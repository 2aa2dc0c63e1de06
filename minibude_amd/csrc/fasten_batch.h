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
};

bool batch_shape_supported(int rb, int wgsize);

// Counter reset + one persistent launch of `blocks` work-groups of `wgsize`
// work-items, each wave scoring tiles of 64 * rb poses.
hipError_t launch_batch(const BatchArgs &a, int rb, int wgsize, int blocks, hipStream_t s);

}  // namespace bude
