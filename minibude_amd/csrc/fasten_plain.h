// fasten_plain.h -- launch interface of the plain fasten_main kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "bude_internal.h"

namespace bude {

// Device pointers and sizes of one plain launch.
struct PlainArgs {
  const float4 *protein;     // grouped protein atoms (x, y, z, 0)
  const int32_t *group_begin;  // ngroups + 1 offsets
  const PairCoef *coef;      // [natlig][ngroups]
  const float4 *ligand;      // ligand atoms (x, y, z, 0), deck order
  const float *poses;        // 6 rows of `stride` floats
  float *energies;           // nposes
  size_t stride;
  size_t nposes;
  int natlig;
  int ngroups;
  int ppwi;
  int form;                  // 0: hinges in b = d - radij, 1: hinges in d
};

// Poses per register block actually held in VGPRs for (ppwi, wgsize).
int plain_register_block(int ppwi, int wgsize);

// One launch over all poses; wgsize in {64, 128, 256, 512, 1024}; ppwi a
// power of two in [1, 128].
hipError_t launch_plain(const PlainArgs &a, int wgsize, hipStream_t s);

}  // namespace bude
