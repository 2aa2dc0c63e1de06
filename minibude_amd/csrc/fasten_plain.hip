// fasten_plain.hip -- the plain fasten_main on gfx950.
//
// Same function as the reference kernel (src/hip/fasten.hpp:13-154,
// semantics src/serial/fasten.hpp:18-132): work-item w of work-group b owns
// the PPWI poses b*wgsize*PPWI + w + i*wgsize (i < PPWI), as in
// src/hip/fasten.hpp:21,48,148-151, and writes 0.5 * sum over (ligand atom,
// protein atom) pairs of steric + charge + desolvation energy.
//
// Built for the FP32 VALU roofline (the loop is not a contraction, so no
// MFMA):
//  * Protein atoms are grouped by forcefield type on the host, so the pair
//    coefficients are constant over a group; the ligand atom, the protein
//    coordinates and the coefficients are wave-uniform and travel through
//    scalar loads into SGPRs, which VALU instructions read for free.
//  * Each work-item keeps RB poses' transforms, transformed ligand atom
//    and energy in VGPRs (register blocking); PPWI > RB is done in RB-sized
//    chunks so no PPWI spills.
//  * The pair energy is three clamped hinges of b = d - radij (see
//    PairCoef): per interaction 3 sub + 3 mul/fma + 1 sqrt + 1 sub +
//    3 clamped mul/fma + 3 fma accumulate, branch-free.
#include <hip/hip_runtime.h>

#include "bude_internal.h"
#include "fasten_core.h"
#include "fasten_plain.h"

namespace bude {

namespace {

using namespace core;

// Separate __restrict__ kernel arguments (not a struct) so the compiler can
// prove the uniform loads are never clobbered and issue them as scalar loads.
template <int RB, int WG, int FORM>
__global__ __launch_bounds__(WG) void fasten_plain_kernel(
    const float4 *__restrict__ protein, const int32_t *__restrict__ group_begin, const PairCoef *__restrict__ coef,
    const float4 *__restrict__ ligand, const float *__restrict__ poses, float *__restrict__ energies, size_t stride,
    size_t n, int natlig, int ngroups, int ppwi) {
  const int tid = threadIdx.x;
  const size_t block_base = size_t(blockIdx.x) * size_t(WG) * size_t(ppwi);

  for (int c0 = 0; c0 < ppwi; c0 += RB) {
    Xform t[RB];
    float e[RB];
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      const size_t pose = block_base + size_t(c0 + k) * WG + tid;
      t[k] = make_xform(poses, stride, pose < n ? pose : n - 1);
      e[k] = 0.f;
    }

    for (int il = 0; il < natlig; ++il) {
      const float4 la = ligand[il];
      float lx[RB], ly[RB], lz[RB];
#pragma unroll
      for (int k = 0; k < RB; ++k) {
        const float *m = t[k].m;
        lx[k] = m[3] + la.x * m[0] + la.y * m[1] + la.z * m[2];
        ly[k] = m[7] + la.x * m[4] + la.y * m[5] + la.z * m[6];
        lz[k] = m[11] + la.x * m[8] + la.y * m[9] + la.z * m[10];
      }
      const PairCoef *cl = coef + size_t(il) * ngroups;
      for (int g = 0; g < ngroups; ++g)
        group_dispatch<RB, FORM>(protein, group_begin[g], group_begin[g + 1], cl[g], lx, ly, lz, e);
    }

#pragma unroll
    for (int k = 0; k < RB; ++k) {
      const size_t pose = block_base + size_t(c0 + k) * WG + tid;
      if (pose < n) energies[pose] = e[k] * 0.5f;
    }
  }
}

template <int RB, int WG>
hipError_t launch_rb_wg(const PlainArgs &a, size_t blocks, hipStream_t s) {
  if (a.form == kFormD)
    hipLaunchKernelGGL(HIP_KERNEL_NAME(fasten_plain_kernel<RB, WG, kFormD>), dim3(unsigned(blocks)), dim3(WG), 0,
                       s, a.protein, a.group_begin, a.coef, a.ligand, a.poses, a.energies, a.stride, a.nposes,
                       a.natlig, a.ngroups, a.ppwi);
  else
    hipLaunchKernelGGL(HIP_KERNEL_NAME(fasten_plain_kernel<RB, WG, kFormB>), dim3(unsigned(blocks)), dim3(WG), 0,
                       s, a.protein, a.group_begin, a.coef, a.ligand, a.poses, a.energies, a.stride, a.nposes,
                       a.natlig, a.ngroups, a.ppwi);
  return hipGetLastError();
}

template <int RB>
hipError_t launch_rb(const PlainArgs &a, int wg, size_t blocks, hipStream_t s) {
  switch (wg) {
    case 64: return launch_rb_wg<RB, 64>(a, blocks, s);
    case 128: return launch_rb_wg<RB, 128>(a, blocks, s);
    case 256: return launch_rb_wg<RB, 256>(a, blocks, s);
    case 512: return launch_rb_wg<RB, 512>(a, blocks, s);
    case 1024: return launch_rb_wg<RB, 1024>(a, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

int plain_register_block(int ppwi, int wgsize) {
  // VGPR budget per work-item shrinks with the work-group size (1024
  // work-items = 4 waves per SIMD = 128 VGPRs each).
  const int cap = wgsize >= 1024 ? 4 : 8;
  return ppwi < cap ? ppwi : cap;
}

hipError_t launch_plain(const PlainArgs &a, int wgsize, hipStream_t s) {
  const size_t items = (a.nposes + size_t(a.ppwi) - 1) / size_t(a.ppwi);
  const size_t blocks = (items + size_t(wgsize) - 1) / size_t(wgsize);
  switch (plain_register_block(a.ppwi, wgsize)) {
    case 1: return launch_rb<1>(a, wgsize, blocks, s);
    case 2: return launch_rb<2>(a, wgsize, blocks, s);
    case 4: return launch_rb<4>(a, wgsize, blocks, s);
    case 8: return launch_rb<8>(a, wgsize, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace bude
