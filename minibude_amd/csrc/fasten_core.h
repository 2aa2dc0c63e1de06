// fasten_core.h -- device building blocks shared by the plain and batched
// fasten_main kernels: pose transform, and the pair loop over one protein
// group with its hinge-form and zero-weight specialisations.
#pragma once

#include <hip/hip_runtime.h>

#include "bude_internal.h"

namespace bude {
namespace core {

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

// Hinge forms of the charge / desolvation terms (see PairCoef):
//   kFormB  sat(1 + b * neg_e1) with b = d - radij (one extra subtraction,
//           matches the reference's rounding of 1 - b*elcdst1 exactly)
//   kFormD  sat(c0 + d * neg_e1) (one rounding of c0 per type pair)
enum HingeForm { kFormB = 0, kFormD = 1 };

// A VGPR copy of a wave-uniform value, opaque to the compiler.  A VALU
// instruction reads at most one SGPR, so fma(d, neg_e1, c0) needs one of the
// two group coefficients in a VGPR; left to itself the compiler re-copies it
// inside the pair loop (it reuses the copy's register for the result), one
// extra VALU per pair.  Copied once per group here, the value stays live.
__device__ __forceinline__ float vgpr_copy(float s) {
  float v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
  return v;
}

// The pairs of one protein atom p with the RB poses of this work-item.
template <int RB, int FORM, bool CHG, bool DSLV>
__device__ __forceinline__ void atom_pairs(const float4 p, const PairCoef &c, float c0, float e0,
                                           const float (&lx)[RB], const float (&ly)[RB], const float (&lz)[RB],
                                           float (&e)[RB]) {
#pragma unroll
  for (int k = 0; k < RB; ++k) {
    const float dx = lx[k] - p.x;
    const float dy = ly[k] - p.y;
    const float dz = lz[k] - p.z;
    // the contraction written out (dx^2, then fma dy, then fma dz): left to
    // the compiler it can pick another order in another kernel, and every
    // kernel sharing this code must round the same way
    const float d = __builtin_amdgcn_sqrtf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)));
    if constexpr (FORM == kFormB) {
      const float b = d - c.radij;
      e[k] = __builtin_fmaf(2.0f * kHardness, sat(b * c.neg_rr), e[k]);
      if constexpr (CHG) e[k] = __builtin_fmaf(c.wq, sat(__builtin_fmaf(b, c.neg_e1, 1.0f)), e[k]);
      if constexpr (DSLV) e[k] = __builtin_fmaf(c.wd, sat(__builtin_fmaf(b, c.neg_rd, 1.0f)), e[k]);
    } else {
      e[k] = __builtin_fmaf(2.0f * kHardness, sat(__builtin_fmaf(d, c.neg_rr, 1.0f)), e[k]);
      if constexpr (CHG) e[k] = __builtin_fmaf(c.wq, sat(__builtin_fmaf(d, c.neg_e1, c0)), e[k]);
      if constexpr (DSLV) e[k] = __builtin_fmaf(c.wd, sat(__builtin_fmaf(d, c.neg_rd, e0)), e[k]);
    }
  }
}

// All pairs of one protein group with the RB poses of this work-item, atoms
// in order.  The protein coordinates are wave-uniform scalar loads, and a
// scalar load can only be waited for together with every other one in
// flight, so the loop runs on two 4-atom buffers: the first use of the
// current buffer waits for it (the only load outstanding), the next buffer is
// requested right after, and the current buffer's other three atoms cover its
// latency.  The scheduling barriers keep the compiler from hoisting the next
// request above that first use (which would make the wait cover both).  A
// wave then keeps its VALU busy even when it is alone on its SIMD.
// CHG / DSLV drop the charge / desolvation hinge when the group's weight is
// exactly zero for this ligand atom (adding fma(0, h, e) == e is a no-op, so
// results are bit-identical).
template <int RB, int FORM, bool CHG, bool DSLV>
__device__ __forceinline__ void group_pairs(const float4 *__restrict__ prot, int i0, int i1, const PairCoef c,
                                            const float (&lx)[RB], const float (&ly)[RB], const float (&lz)[RB],
                                            float (&e)[RB]) {
  float c0 = 0.f, e0 = 0.f;
  if constexpr (FORM == kFormD) {
    if constexpr (CHG) c0 = vgpr_copy(c.c0);
    if constexpr (DSLV) e0 = vgpr_copy(c.e0);
  }
  int i = i0;
  if (i1 - i >= 8) {
    float4 x0 = prot[i], x1 = prot[i + 1], x2 = prot[i + 2], x3 = prot[i + 3];
    // invariant: x0..x3 hold atoms i..i+3 and i + 4 <= i1
    for (; i + 12 <= i1; i += 8) {
      atom_pairs<RB, FORM, CHG, DSLV>(x0, c, c0, e0, lx, ly, lz, e);
      __builtin_amdgcn_sched_barrier(0);
      const float4 y0 = prot[i + 4], y1 = prot[i + 5], y2 = prot[i + 6], y3 = prot[i + 7];
      __builtin_amdgcn_sched_barrier(0);
      atom_pairs<RB, FORM, CHG, DSLV>(x1, c, c0, e0, lx, ly, lz, e);
      atom_pairs<RB, FORM, CHG, DSLV>(x2, c, c0, e0, lx, ly, lz, e);
      atom_pairs<RB, FORM, CHG, DSLV>(x3, c, c0, e0, lx, ly, lz, e);
      atom_pairs<RB, FORM, CHG, DSLV>(y0, c, c0, e0, lx, ly, lz, e);
      __builtin_amdgcn_sched_barrier(0);
      x0 = prot[i + 8];
      x1 = prot[i + 9];
      x2 = prot[i + 10];
      x3 = prot[i + 11];
      __builtin_amdgcn_sched_barrier(0);
      atom_pairs<RB, FORM, CHG, DSLV>(y1, c, c0, e0, lx, ly, lz, e);
      atom_pairs<RB, FORM, CHG, DSLV>(y2, c, c0, e0, lx, ly, lz, e);
      atom_pairs<RB, FORM, CHG, DSLV>(y3, c, c0, e0, lx, ly, lz, e);
    }
    atom_pairs<RB, FORM, CHG, DSLV>(x0, c, c0, e0, lx, ly, lz, e);
    atom_pairs<RB, FORM, CHG, DSLV>(x1, c, c0, e0, lx, ly, lz, e);
    atom_pairs<RB, FORM, CHG, DSLV>(x2, c, c0, e0, lx, ly, lz, e);
    atom_pairs<RB, FORM, CHG, DSLV>(x3, c, c0, e0, lx, ly, lz, e);
    i += 4;
  }
  for (; i + 2 <= i1; i += 2) {
    const float4 p0 = prot[i], p1 = prot[i + 1];
    atom_pairs<RB, FORM, CHG, DSLV>(p0, c, c0, e0, lx, ly, lz, e);
    atom_pairs<RB, FORM, CHG, DSLV>(p1, c, c0, e0, lx, ly, lz, e);
  }
  if (i < i1) atom_pairs<RB, FORM, CHG, DSLV>(prot[i], c, c0, e0, lx, ly, lz, e);
}

// One protein group: its pairs are summed into a fresh partial sum that is
// then added to the pose total, so the running FP32 total takes one rounding
// per (ligand atom, group) instead of three per pair (a shorter, smaller-
// magnitude accumulation than the reference's single running sum).
template <int RB, int FORM>
__device__ __forceinline__ void group_dispatch(const float4 *__restrict__ prot, int i0, int i1, const PairCoef c,
                                               const float (&lx)[RB], const float (&ly)[RB],
                                               const float (&lz)[RB], float (&e)[RB]) {
  float part[RB];
#pragma unroll
  for (int k = 0; k < RB; ++k) part[k] = 0.f;
  const bool chg = c.wq != 0.f, dslv = c.wd != 0.f;  // wave-uniform
  if (chg && dslv) group_pairs<RB, FORM, true, true>(prot, i0, i1, c, lx, ly, lz, part);
  else if (chg) group_pairs<RB, FORM, true, false>(prot, i0, i1, c, lx, ly, lz, part);
  else if (dslv) group_pairs<RB, FORM, false, true>(prot, i0, i1, c, lx, ly, lz, part);
  else group_pairs<RB, FORM, false, false>(prot, i0, i1, c, lx, ly, lz, part);
#pragma unroll
  for (int k = 0; k < RB; ++k) e[k] += part[k];
}

}  // namespace core
}  // namespace bude
