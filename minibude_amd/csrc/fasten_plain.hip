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
#include "fasten_plain.h"

namespace bude {

namespace {

__device__ __forceinline__ float sat(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }

// 3x4 rigid transform of one pose (src/serial/fasten.hpp:32-50):
// R = Rz(pan) * Ry(roll) * Rx(tilt), translation in the last column.
struct Xform {
  float m[12];
};

__device__ __forceinline__ Xform make_xform(const float *__restrict__ poses, size_t stride, size_t q) {
  const float tilt = poses[q], roll = poses[stride + q], pan = poses[2 * stride + q];
  float sx, cx, sy, cy, sz, cz;
  sincosf(tilt, &sx, &cx);
  sincosf(roll, &sy, &cy);
  sincosf(pan, &sz, &cz);
  Xform t;
  t.m[0] = cy * cz;
  t.m[1] = sx * sy * cz - cx * sz;
  t.m[2] = cx * sy * cz + sx * sz;
  t.m[3] = poses[3 * stride + q];
  t.m[4] = cy * sz;
  t.m[5] = sx * sy * sz + cx * cz;
  t.m[6] = cx * sy * sz - sx * cz;
  t.m[7] = poses[4 * stride + q];
  t.m[8] = -sy;
  t.m[9] = sx * cy;
  t.m[10] = cx * cy;
  t.m[11] = poses[5 * stride + q];
  return t;
}

template <int RB>
__device__ __forceinline__ void group_pairs(const float4 *__restrict__ prot, int i0, int i1, const PairCoef c,
                                            const float (&lx)[RB], const float (&ly)[RB], const float (&lz)[RB],
                                            float (&e)[RB]) {
#pragma unroll 2
  for (int i = i0; i < i1; ++i) {
    const float4 p = prot[i];
#pragma unroll
    for (int k = 0; k < RB; ++k) {
      const float dx = lx[k] - p.x;
      const float dy = ly[k] - p.y;
      const float dz = lz[k] - p.z;
      const float d = __builtin_amdgcn_sqrtf(dx * dx + dy * dy + dz * dz);
      const float b = d - c.radij;
      const float st = sat(b * c.neg_rr);
      const float ch = sat(__builtin_fmaf(b, c.neg_e1, 1.0f));
      const float ds = sat(__builtin_fmaf(b, c.neg_rd, 1.0f));
      e[k] = __builtin_fmaf(2.0f * kHardness, st, e[k]);
      e[k] = __builtin_fmaf(c.wq, ch, e[k]);
      e[k] = __builtin_fmaf(c.wd, ds, e[k]);
    }
  }
}

// Separate __restrict__ kernel arguments (not a struct) so the compiler can
// prove the uniform loads are never clobbered and issue them as scalar loads.
template <int RB, int WG>
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
        group_pairs<RB>(protein, group_begin[g], group_begin[g + 1], cl[g], lx, ly, lz, e);
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
  hipLaunchKernelGGL(HIP_KERNEL_NAME(fasten_plain_kernel<RB, WG>), dim3(unsigned(blocks)), dim3(WG), 0, s,
                     a.protein, a.group_begin, a.coef, a.ligand, a.poses, a.energies, a.stride, a.nposes, a.natlig,
                     a.ngroups, a.ppwi);
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
