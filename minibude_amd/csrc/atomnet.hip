// atomnet.hip -- the learned per-atom field (see atomnet.h): training
// points and their exact labels, the fused training step (forward and
// backward on BF16 MFMA, deterministic fixed-point grid gradients, Adam), and
// the fused scorer that sums the network over a ligand's atoms for every pose.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdlib>
#include <random>
#include <vector>

#include "atomnet.h"
#include "fasten_core.h"

namespace bude {
namespace atomnet {

namespace {

using namespace core;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float floatx2_t __attribute__((ext_vector_type(2)));

// FP32 -> BF16 (round to nearest even), two values per instruction.
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(floatx2_t{lo, hi}, bf16x2_t));
}
__device__ __forceinline__ float lo_bf16(uint32_t v) { return __uint_as_float(v << 16); }

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Cell index and fraction along one axis, clamped to the grid (as the
// field scorer's).
__device__ __forceinline__ void cell(float w, float lo, float inv_h, int n, int &i0, float &f) {
  const float u = (w - lo) * inv_h;
  i0 = min(max(int(floorf(u)), 0), n - 2);
  f = __builtin_amdgcn_fmed3f(u - float(i0), 0.0f, 1.0f);
}

// ---- training points and labels ---------------------------------------------

__global__ void sample_kernel(const float4 *__restrict__ ligands, const LigandMeta *__restrict__ meta, int nligands,
                              const float *__restrict__ poses, size_t stride, int nslots, uint64_t seed, size_t n,
                              float4 *__restrict__ out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t h1 = mix64(seed ^ (uint64_t(i) * 0xd1342543de82ef95ull));
  const uint64_t h2 = mix64(h1 ^ 0x632be59bd9b4e019ull);
  const LigandMeta m = meta[h1 % uint64_t(nligands)];
  const size_t j = size_t((h2 >> 32) % uint64_t(m.nposes));
  const int a = int((h2 & 0xffffffffull) % uint64_t(m.natlig));
  const Xform x = make_xform(poses, stride, size_t(m.pose_off) + j);
  const float4 la = ligands[m.lig_off + a];
  const float px = x.m[3] + la.x * x.m[0] + la.y * x.m[1] + la.z * x.m[2];
  const float py = x.m[7] + la.x * x.m[4] + la.y * x.m[5] + la.z * x.m[6];
  const float pz = x.m[11] + la.x * x.m[8] + la.y * x.m[9] + la.z * x.m[10];
  out[i] = make_float4(px, py, pz, __int_as_float(int((i / 64) % size_t(nslots))));
}

// One wave per 64 points of one slot: the plain kernel's pair loop for a
// single ligand atom at each point (lx = the point, exactly what a pose of
// zero angles and translation p gives an atom at the origin).
template <int FORM>
__global__ __launch_bounds__(256) void label_kernel(const float4 *__restrict__ protein,
                                                    const int32_t *__restrict__ group_begin,
                                                    const PairCoef *__restrict__ slot_coef, int ngroups,
                                                    const float4 *__restrict__ points, size_t n,
                                                    float *__restrict__ out) {
  const int lane = threadIdx.x & 63;
  const size_t base = (size_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * 64;
  if (base >= n) return;
  const float4 p = points[base + lane];
  const int slot = __builtin_amdgcn_readfirstlane(__float_as_int(p.w));
  const PairCoef *c = slot_coef + size_t(slot) * size_t(ngroups);
  const float lx[1] = {p.x}, ly[1] = {p.y}, lz[1] = {p.z};
  float e[1] = {0.f};
  for (int g = 0; g < ngroups; ++g) group_dispatch<1, FORM>(protein, group_begin[g], group_begin[g + 1], c[g], lx, ly, lz, e);
  out[base + lane] = e[0] * 0.5f;
}

// Partial sums (sum, sum of squares, in FP64) of labels[0, n) per block.
__global__ __launch_bounds__(256) void stats_kernel(const float *__restrict__ labels, size_t n,
                                                    double *__restrict__ part) {
  __shared__ double s1[256], s2[256];
  double a = 0, b = 0;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const double v = labels[i];
    a += v;
    b += v * v;
  }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

// ---- training kernels -------------------------------------------------------

// Fixed-point accumulators -> FP32 gradient (forward_backward for tests).
__global__ void grid_grad_out_kernel(unsigned long long *__restrict__ acc, float *__restrict__ out, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = float(double(static_cast<long long>(acc[i])) / kGradScale);
  acc[i] = 0ull;
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pack_half2(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, half2_t{_Float16(lo), _Float16(hi)});
}

// FP32 grid -> cell table: out[(cell 4 + q) kF + k] = FP16 (g[node][k],
// g[node + 1][k]), node = the cell's corner (i + dx, j + dy, l) of column
// q = 2 dx + dy: the 8 corners a trilinear interpolation reads, in one
// 128-byte record per cell.
__global__ void cell_table_kernel(const float *__restrict__ g, const field::Grid grid, uint32_t *__restrict__ out) {
  const size_t c1 = size_t(grid.n[1] - 1), c2 = size_t(grid.n[2] - 1);
  const size_t n = size_t(grid.n[0] - 1) * c1 * c2 * 4 * kF;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int k = int(t % kF), q = int((t / kF) % 4);
  const size_t cell = t / (4 * kF), l = cell % c2, j = (cell / c2) % c1, i = cell / (c2 * c1);
  const size_t node = ((i + size_t(q >> 1)) * size_t(grid.n[1]) + j + size_t(q & 1)) * size_t(grid.n[2]) + l;
  out[t] = pack_half2(g[node * kF + k], g[(node + 1) * kF + k]);
}

// Squared error of sigma * pred + mu against min(label, clip), summed in FP64 per block.
__global__ __launch_bounds__(256) void sq_err_kernel(const float *__restrict__ pred, const float *__restrict__ labels,
                                                     size_t n, float sigma, float mu, float clip,
                                                     double *__restrict__ part) {
  __shared__ double s[256];
  double a = 0;
  for (size_t i = size_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += size_t(gridDim.x) * 256) {
    const double d = double(sigma) * pred[i] + mu - double(fminf(labels[i], clip));
    a += d * d;
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// A standard normal deviate from a hash of (seed, i) (Box-Muller).
__device__ __forceinline__ float normal(uint64_t seed, uint64_t i) {
  const uint64_t h = mix64(seed ^ (i * 0x9e3779b97f4a7c15ull));
  const float u1 = (float(uint32_t(h >> 40)) + 1.0f) * (1.0f / 16777217.0f);  // (0, 1]
  const float u2 = float(uint32_t(h) >> 8) * (1.0f / 16777216.0f);
  return sqrtf(-2.0f * logf(u1)) * cosf(6.2831853f * u2);
}

// w[r][c] = sd * N(0, 1) for r < rows, c < cols of a [.][ld] matrix, 0
// elsewhere (n entries in all); m, v (Adam moments) cleared.
__global__ void init_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v, size_t n, int ld,
                            int rows, int cols, float sd, uint64_t seed) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = int(i / size_t(ld)), c = int(i % size_t(ld));
  w[i] = (r < rows && c < cols) ? sd * normal(seed, i) : 0.f;
  m[i] = 0.f;
  v[i] = 0.f;
}

// slot_bias[s][j] = b0[j] + W0[kF + s][j] (W0's BF16 copy, row-major [k0][kH]).
__global__ void slot_bias_kernel(const float *__restrict__ b0, const bf16 *__restrict__ w16, int nslots,
                                 float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots * kH) return;
  const int s = i / kH, j = i % kH;
  out[i] = b0[j] + __uint_as_float(uint32_t(w16[size_t(kF + s) * kH + j]) << 16);
}

// ---- fused scorer -------------------------------------------------------------
// One wave per 64 poses of one ligand (lane = pose).  Per ligand atom:
//   gather    each pose's transformed atom, its cell and the four FP16
//             weight pairs (wx wy (1 - fz), wx wy fz) of the cell's z-pairs
//             go through the wave's LDS to the eight lanes that then load the
//             cell's 128-byte record together (16 bytes each: one z-pair's
//             four features at both z): eight loads bring the 64 records,
//             each touching 8 cache lines, where one record per lane would
//             touch 64 -- the vector L1's line lookups bound the
//             one-record-per-lane form;
//   features  each lane interpolates its four features of its z-pair (one
//             packed FP16 dot product each, FP32), rounds them to BF16 and
//             writes them to the pose's LDS row, K slot 8 q + f (pair q,
//             feature f);
//   decoder   on MFMA, as the transpose D[neuron][pose] = W^T . x^T so the
//             weights are the (register-held) A operand and the wave's 64
//             poses four 16-pose B operands: layer 1 over K = 32 with W0's
//             feature rows repeated for the four pairs (the sum over pairs
//             is the MFMA's own), onto the slot's FP32 bias (b0 + W0[kF +
//             slot]); ReLU; layer 2 onto b1 with layer 1's accumulators as
//             its B operand in registers; ReLU; the output neuron as a third
//             MFMA with W2 in every row of A, accumulated over the atoms in
//             its accumulators.  The ReLUs act on BF16 pairs (one packed
//             16-bit integer max each), after the rounding the next layer's
//             operand needs anyway.

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x8 lds8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

// The four features of one z-pair chunk, interpolated and rounded to BF16
// pairs: four v_dot2_f32_f16 with a zero addend (the fdot2 builtin becomes
// the accumulating VOP2 form after a move of zero into each accumulator),
// then the packs, after the 3 wait states a DOT result needs before another
// VALU instruction reads it (s_nop 1 plus the two DOTs after the second).
__device__ __forceinline__ uint2 pair_features(uint32_t w, uint4 v) {
  uint32_t o0, o1;
  float t0, t1, t2, t3;
  asm("v_dot2_f32_f16 %2, %6, %7, 0\n\t"
      "v_dot2_f32_f16 %3, %6, %8, 0\n\t"
      "v_dot2_f32_f16 %4, %6, %9, 0\n\t"
      "v_dot2_f32_f16 %5, %6, %10, 0\n\t"
      "s_nop 2\n\t"
      "v_cvt_pk_bf16_f32 %0, %2, %3\n\t"
      "v_cvt_pk_bf16_f32 %1, %4, %5"
      : "=v"(o0), "=v"(o1), "=&v"(t0), "=&v"(t1), "=&v"(t2), "=&v"(t3)
      : "v"(w), "v"(v.x), "v"(v.y), "v"(v.z), "v"(v.w));
  return make_uint2(o0, o1);
}

constexpr int kStage = 128;  // ligand atoms staged in LDS per wave at a time
// Layer 1's B rows in LDS: pose p's four pair slots (uint4 each) start at
// uint4 4 p + 4 (p / 16) + (p / 4) % 4.  The stagger makes both accesses
// conflict-free: the 8-byte feature writes (32 lanes = 4 poses x 64 bytes
// fill the 64 banks) and the 16-byte operand reads (16 lanes = 16 poses of
// one pair slot, 16 distinct 4-bank groups).
constexpr int kRowsPerWave = 272;

// One atom's share of a lane: the 16 bytes of each of its eight records
// (record of pose 8 m + lane / 8, bytes 16 (lane % 8) ..).
struct AtomFetch {
  uint4 v[8];
  int slot;
};

// ReLU of two BF16 values packed in one word, as one packed 16-bit integer
// max with zero: a BF16 value with its sign bit set (negative, or -0) is a
// negative int16.  Equal, bit for bit, to packing relu() of the FP32 values.
typedef short short2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t relu_bf16x2(uint32_t v) {
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(short2_t, v), short2_t{0, 0}));
}

// ReLU of the neurons of one 16-point block held as the accumulators of
// the two 16-neuron blocks, as the next layer's BF16 B operand (K slot
// 8 grp + e = neuron 16 (e / 4) + 4 grp + e % 4).
__device__ __forceinline__ bf16x8 relu_operand(const floatx4 &v0, const floatx4 &v1) {
  return __builtin_bit_cast(bf16x8, make_uint4(relu_bf16x2(pack_bf16(v0[0], v0[1])), relu_bf16x2(pack_bf16(v0[2], v0[3])),
                                               relu_bf16x2(pack_bf16(v1[0], v1[1])), relu_bf16x2(pack_bf16(v1[2], v1[3]))));
}

// K slot 8 grp + e of a permuted-K operand holds hidden neuron
// 16 (e / 4) + 4 grp + e % 4: a row's two 4-element halves.
__device__ __forceinline__ bf16x8 perm_row(const bf16 *row, int grp) {
  const uint2 lo = *reinterpret_cast<const uint2 *>(row + 4 * grp), hi = *reinterpret_cast<const uint2 *>(row + 16 + 4 * grp);
  return __builtin_bit_cast(bf16x8, make_uint4(lo.x, lo.y, hi.x, hi.y));
}

__global__ __launch_bounds__(256, 3) void score_kernel(const ScoreArgs a, int64_t tend) {
  __shared__ uint4 xrows[4][kRowsPerWave];  // per wave: layer 1's B rows, 4 pairs x 8 BF16 features per pose
  // per wave and atom parity: each pose's record byte offset, and its weight
  // pair of z-pair q at [q][pose] (rows padded to 80 words: the readers of
  // one instruction, 8 poses x 4 pairs, hit 32 distinct banks)
  __shared__ uint32_t xoff[4][2][64];
  __shared__ uint32_t xw[4][2][4][80];
  __shared__ float w2s[kH], b1s[kH];
  __shared__ float sbias[kMaxSlots * kH];
  __shared__ float4 latoms[4][kStage];
  const int tid = threadIdx.x;
  if (tid < kH) {
    w2s[tid] = __uint_as_float(uint32_t(a.w2[size_t(tid) * kOut]) << 16);
    b1s[tid] = a.b1[tid];
  }
  for (int i = tid; i < a.nslots * kH; i += 256) sbias[i] = a.slot_bias[i];
  __syncthreads();
  if (a.raise_priority) __builtin_amdgcn_s_setprio(2);
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t t = a.tile0 + int64_t(blockIdx.x) * 4 + wave;
  if (t >= tend) return;
  int lo = 0, hi = a.nligands - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.tile_begin[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  const LigandMeta m = a.meta[lo];
  const int j = int(t - a.tile_begin[lo]) * 64 + lane;
  const Xform x = make_xform(a.poses, a.stride, size_t(m.pose_off) + size_t(j < m.nposes ? j : m.nposes - 1));
  const int r = lane & 15, grp = lane >> 4;
  // register-held A fragments: layer 1's, the feature row 16 jn + r of W0^T
  // in every lane group (K slot 8 q + f: pair q, feature f); layer 2's,
  // W1^T with its K (input neuron) order permuted so that layer 1's
  // accumulators are layer 2's B operand as they stand: lane (r, grp) holds
  // hidden neurons 16 jn + 4 grp + c of point r (jn = 0, 1; c = 0..3) after
  // layer 1, so layer 2's K slot 8 grp + e is neuron 16 (e / 4) + 4 grp +
  // e % 4; the output layer's, W2 in that order in every row
  bf16x8 fw0[kH / 16], fw1[kH / 16];
#pragma unroll
  for (int jn = 0; jn < kH / 16; jn++) {
    fw0[jn] = *reinterpret_cast<const bf16x8 *>(a.w0t + size_t(16 * jn + r) * a.k0);
    fw1[jn] = perm_row(a.w1t + size_t(16 * jn + r) * kH, grp);
  }
  const bf16x8 fw2 = __builtin_bit_cast(
      bf16x8, make_uint4(pack_bf16(w2s[4 * grp], w2s[4 * grp + 1]), pack_bf16(w2s[4 * grp + 2], w2s[4 * grp + 3]),
                         pack_bf16(w2s[16 + 4 * grp], w2s[17 + 4 * grp]),
                         pack_bf16(w2s[18 + 4 * grp], w2s[19 + 4 * grp])));
  // this lane's place in the cooperative gathers: record of pose 8 m + pl,
  // 16-byte chunk ch = (pair ch / 2, features 4 (ch % 2) ..)
  const int pl = lane >> 3, ch = lane & 7;
  uint4 *rows = xrows[wave];
  // pose 8 m + pl: uint4 32 m + 4 (m / 2) + 2 (m % 2) + 4 pl + pl / 4
  uint2 *wrow = reinterpret_cast<uint2 *>(rows + 4 * pl + (pl >> 2)) + ch;
  // pose 16 i + r: uint4 68 i + 4 r + (r / 4) % 4
  const uint4 *brow = rows + 4 * r + (r >> 2) + grp;
  const char *table = reinterpret_cast<const char *>(a.gzp);
  // the pose transform in grid coordinates, u = (R x + t - lo) / h
  const float inv_h = 1.0f / a.grid.h;
  float xg[12];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    xg[4 * d] = x.m[4 * d] * inv_h;
    xg[4 * d + 1] = x.m[4 * d + 1] * inv_h;
    xg[4 * d + 2] = x.m[4 * d + 2] * inv_h;
    xg[4 * d + 3] = (x.m[4 * d + 3] - a.grid.lo[d]) * inv_h;
  }
  // cell index and fraction along one axis from u, clamped to the grid: for
  // finite u the values cell() gives
  auto cell_u = [](float u, int n, int &i0, float &f) {
    const float uc = __builtin_amdgcn_fmed3f(u, 0.0f, float(n - 1));
    i0 = min(int(uc), n - 2);
    f = uc - float(i0);
  };
  // the output neuron of point 16 i + r (every accumulator row holds it),
  // summed over atoms
  floatx4 out[4];
#pragma unroll
  for (int i = 0; i < 4; i++) out[i] = floatx4{0.f, 0.f, 0.f, 0.f};
  // the ligand's atoms, staged kStage at a time in the wave's LDS rows (read
  // as broadcasts: no global load, and so no wait, inside the atom loop)
  const int natlig = __builtin_amdgcn_readfirstlane(m.natlig);
  const int lig_off = __builtin_amdgcn_readfirstlane(m.lig_off);
  float4 *ls = latoms[wave];
  auto stage = [&](int base) {
    wave_sync();  // (earlier reads of the rows are done)
    for (int i = lane; i < kStage && base + i < natlig; i += 64) ls[i] = a.ligands[lig_off + base + i];
    wave_sync();
  };
  // atom i's gathers through the transfer rows of its parity; the first
  // atom of a new stage reloads the rows
  auto fetch = [&](int i, AtomFetch &f, int par) {
    if (i > 0 && i % kStage == 0) stage(i);
    const float4 la = ls[i % kStage];
    f.slot = __float_as_int(la.w);  // (the same in every lane)
    int ix, iy, iz;
    float fx, fy, fz;
    cell_u(xg[3] + la.x * xg[0] + la.y * xg[1] + la.z * xg[2], a.grid.n[0], ix, fx);
    cell_u(xg[7] + la.x * xg[4] + la.y * xg[5] + la.z * xg[6], a.grid.n[1], iy, fy);
    cell_u(xg[11] + la.x * xg[8] + la.y * xg[9] + la.z * xg[10], a.grid.n[2], iz, fz);
    // the cell's record (4 z-pairs (dx, dy) x 8 features x 2 z, FP16: 128
    // bytes; 32-bit byte offsets: the host keeps the table under 2^31 bytes)
    const uint32_t c1 = uint32_t(a.grid.n[1] - 1), c2 = uint32_t(a.grid.n[2] - 1);
    const uint32_t off = ((uint32_t(ix) * c1 + uint32_t(iy)) * c2 + uint32_t(iz)) * (16 * kF);
    const float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy};
    uint32_t w2[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float wxy = wx[q >> 1] * wy[q & 1];
      w2[q] = pack_half2(wxy * (1.f - fz), wxy * fz);
    }
    wave_sync();  // (the reads of these rows by the atom before last are done)
    xoff[wave][par][lane] = off;
#pragma unroll
    for (int q = 0; q < 4; q++) xw[wave][par][q][lane] = w2[q];
    wave_sync();
#pragma unroll
    for (int mm = 0; mm < 8; mm++)
      f.v[mm] = *reinterpret_cast<const uint4 *>(table + (xoff[wave][par][8 * mm + pl] + uint32_t(16 * ch)));
  };
  // one atom through the network, its output added to out[]
  auto eval = [&](const AtomFetch &cur, int par) {
    const int slot = __builtin_amdgcn_readfirstlane(cur.slot);
    // this lane's four features of its z-pair for its eight poses
#pragma unroll
    for (int mm = 0; mm < 8; mm++) {
      const uint32_t w = xw[wave][par][ch >> 1][8 * mm + pl];
      const uint4 v = cur.v[mm];
      wrow[2 * (32 * mm + 4 * (mm >> 1) + 2 * (mm & 1))] = pair_features(w, v);
    }
    wave_sync();
    // layer 1 onto the slot's bias
    floatx4 sb[kH / 16];
#pragma unroll
    for (int jn = 0; jn < kH / 16; jn++) {
      const float4 v = *reinterpret_cast<const float4 *>(sbias + slot * kH + 16 * jn + 4 * grp);
      sb[jn] = floatx4{v.x, v.y, v.z, v.w};
    }
    floatx4 acc[kH / 16][4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const bf16x8 fb = __builtin_bit_cast(bf16x8, brow[68 * i]);
#pragma unroll
      for (int jn = 0; jn < kH / 16; jn++)
        acc[jn][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw0[jn], fb, sb[jn], 0, 0, 0);
    }
    wave_sync();  // (the rows are rewritten by the next atom)
    // layer 2 onto b1 with layer 1's ReLU outputs as its B operand straight
    // from the accumulators (K permuted, above); the output neuron on its
    // ReLU outputs likewise
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const bf16x8 fb = relu_operand(acc[0][i], acc[1][i]);
      floatx4 z[kH / 16];
#pragma unroll
      for (int jn = 0; jn < kH / 16; jn++) {
        const float4 bb = *reinterpret_cast<const float4 *>(b1s + 16 * jn + 4 * grp);
        z[jn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw1[jn], fb, floatx4{bb.x, bb.y, bb.z, bb.w}, 0, 0, 0);
      }
      out[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw2, relu_operand(z[0], z[1]), out[i], 0, 0, 0);
    }
  };
  // two fetch buffers in turn: atom i + 1's gathers are in flight while
  // atom i is evaluated.  The fetches are unconditional (past the last atom
  // they fetch it again: in bounds, unused), so the loads' waits are placed
  // for that order on every path.
  if (natlig == 0) {
    if (j < m.nposes) a.out[size_t(m.pose_off) + size_t(j)] = 0.f;
    return;
  }
  AtomFetch f0{}, f1{};
  stage(0);
  fetch(0, f0, 0);
  for (int il = 0; il < natlig; il += 2) {
    fetch(min(il + 1, natlig - 1), f1, 1);
    eval(f0, 0);
    fetch(min(il + 2, natlig - 1), f0, 0);
    if (il + 1 < natlig) eval(f1, 1);
  }
  // pose j = 16 grp + r of the tile is point r of block grp
  const float sy = grp == 0 ? out[0][0] : (grp == 1 ? out[1][0] : (grp == 2 ? out[2][0] : out[3][0]));
  if (j < m.nposes)
    a.out[size_t(m.pose_off) + size_t(j)] = a.sigma * (sy + float(m.natlig) * a.b2) + float(m.natlig) * a.mu;
}

// ---- fused training step ------------------------------------------------------
// One launch does a training step's forward, loss and backward: one wave per
// 64 points of the batch (lane = point while the point's own data are
// formed, then the MFMA lane layout of the scorer), kStepWaves waves per
// work-group.
//   forward   phi from the FP32 grid (trilinear), layer 0 on MFMA over
//             K = k0 ([phi, one-hot slot]), ReLU, layer 1 on MFMA with layer
//             0's accumulators as its B operand (W1^T's K order permuted as
//             in the scorer), ReLU, y = a1 . W2[:, 0] + b2 (VALU dot over the
//             lane's neurons, summed over the four lane groups), loss
//             gradient dz3 = bf16(2 (y - t) / batch).
//   backward  dz1 = bf16(dz3 W2 * relu'(z1)) elementwise; dz0 =
//             bf16((dz1 W1^T) * relu'(z0)) on MFMA with dz1 from registers
//             (W1's K order permuted likewise); dL/dphi = dz0 W0[0..kF)^T on
//             MFMA; weight and bias gradients X^T dZ as MFMAs over the
//             wave's 64 points from LDS tiles written transposed (bias
//             gradients against an all-ones A row); the grid gradient
//             scattered to the 8 corners in 2^-40 fixed point by 64-bit
//             integer atomics, one point per instruction (lane = (corner,
//             feature)).
// The work-group then sums its waves' gradients in wave order into its own
// slice of gw / gb, which the Adam step sums in a fixed order: the step is
// deterministic.  BF16 rounding sits where the layer-by-layer
// kernels put it (inputs, weights, activations, dz), so the oracle of those
// kernels is the oracle of this one.
constexpr int kTP = 64 + 8;  // bf16 pitch of a transposed [row][64 points] tile (padding staggers banks)
// Waves (of 64 points) per work-group of the fused step: a batch of 32,768
// points is then 256 work-groups, one per CU.
constexpr int kStepWaves = 2;

struct alignas(16) StepRow {
  uint4 phi;  // BF16 features of the point
  int node;   // grid node of the cell's lower corner
  float fx, fy, fz;
  int slot;
  float yt;   // standardised target
  float dz3;  // loss gradient (BF16 value)
  int pad;
};

// The B operand of point block i from per-neuron values held as the
// accumulators of the two 16-neuron blocks (the permuted K order).
__device__ __forceinline__ bf16x8 acc_operand(const floatx4 &v0, const floatx4 &v1) {
  return __builtin_bit_cast(bf16x8, make_uint4(pack_bf16(v0[0], v0[1]), pack_bf16(v0[2], v0[3]),
                                               pack_bf16(v1[0], v1[1]), pack_bf16(v1[2], v1[3])));
}

__device__ __forceinline__ float bf16_round(float x) { return lo_bf16(pack_bf16(x, 0.f)); }

struct StepArgs {
  const float4 *points;
  // batch rows -> points: with perm set, row r of step s (the count in
  // state) is point perm[(s mod per_epoch) n + r], and the launch also
  // writes the step's learning rates and bias corrections (sched) into
  // state for the Adam kernels after it; else through idx, if set; else
  // rows are points
  const uint32_t *perm;
  uint32_t per_epoch;
  mlp::StepState *state;
  mlp::Schedule sched;
  const uint32_t *idx;
  const float *labels;    // exact one-atom energy of every point
  size_t n;               // rows in this launch (a multiple of 64)
  const float *grid;      // FP32 grid features [nodes][kF]
  field::Grid g;
  const bf16 *w0t, *w0;   // W0^T [kH][k0], W0 [k0][kH]
  const bf16 *w1t, *w1;   // W1^T, W1 [kH][kH]
  const bf16 *w2;         // W2 [kH][kOut] (column 0)
  const float *b0, *b1, *b2;
  mlp::TargetNorm norm;
  float grad_scale;       // 2 / batch
  float *pred;            // optional: y of every row
  unsigned long long *loss;  // optional: fixed-point sum of squared errors (mlp::kLossScale)
  int backward;
  unsigned long long *gacc;  // grid gradient accumulators [nodes][kF]
  float *gw[3], *gb[3];   // this launch's work-group slices: gw[l] [nwg][K_l][N_l], gb[l] [nwg][N_l]
};

template <int K0S>
__global__ __launch_bounds__(64 * kStepWaves, 2) void step_kernel(const StepArgs a) {
  constexpr int K0 = 32 * K0S;
  constexpr int NP = K0 * kH + kH * kH + 3 * kH + 1;  // a wave's partial gradients (dW2, db2: column 0)
  constexpr int NT = 64 * kStepWaves;
  __shared__ StepRow rows[kStepWaves][64];
  __shared__ __align__(16) float dphi[kStepWaves][64 * kF];
  __shared__ __align__(16) bf16 tt[kStepWaves][64 * kTP + kH * kTP];  // transposed tiles; the wave partials at the end
  static_assert(sizeof(tt) >= kStepWaves * NP * sizeof(float), "reduction buffer");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, r = lane & 15, grp = lane >> 4;
  const size_t tile = size_t(blockIdx.x) * kStepWaves + wave;
  const bool active = tile * 64 < a.n;
  if (a.perm && blockIdx.x == 0 && tid == 0) mlp::schedule_hyper(a.sched, a.state->step, a.state->hyper);
  StepRow *R = rows[wave];
  bf16 *U = tt[wave];
  const bf16x8 zero8 = bf16x8{};
  float sq = 0.f;
  // this wave's partial gradients, in MFMA output layouts
  floatx4 gW0[K0 / 16][2], gW1[2][2], gW2[2], gB0[2], gB1[2];
  float gB2 = 0.f;
#pragma unroll
  for (int b = 0; b < 2; b++) {
#pragma unroll
    for (int kb = 0; kb < K0 / 16; kb++) gW0[kb][b] = floatx4{0.f, 0.f, 0.f, 0.f};
    gW1[0][b] = gW1[1][b] = gW2[b] = gB0[b] = gB1[b] = floatx4{0.f, 0.f, 0.f, 0.f};
  }
  if (active) {
    const float inv_h = 1.0f / a.g.h;
    const int s1 = a.g.n[2], s0 = a.g.n[1] * a.g.n[2];
    // ---- the point of this lane: features, target
    {
      const size_t row = tile * 64 + lane;
      const uint32_t q = a.perm ? a.perm[size_t(a.state->step % a.per_epoch) * a.n + row]
                                : (a.idx ? a.idx[row] : uint32_t(row));
      const float4 p = a.points[q];
      const float e = a.labels[q];
      int ix, iy, iz;
      float fx, fy, fz;
      cell(p.x, a.g.lo[0], inv_h, a.g.n[0], ix, fx);
      cell(p.y, a.g.lo[1], inv_h, a.g.n[1], iy, fy);
      cell(p.z, a.g.lo[2], inv_h, a.g.n[2], iz, fz);
      const int node = ix * s0 + iy * s1 + iz;
      const float gx[2] = {1.f - fx, fx}, gy[2] = {1.f - fy, fy}, gz[2] = {1.f - fz, fz};
      float phi[kF];
#pragma unroll
      for (int k = 0; k < kF; k++) phi[k] = 0.f;
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int dx = (c >> 2) & 1, dy = (c >> 1) & 1, dz = c & 1;
        const float w = gx[dx] * gy[dy] * gz[dz];
        const float4 *v = reinterpret_cast<const float4 *>(a.grid + size_t(node + dx * s0 + dy * s1 + dz) * kF);
        const float4 u0 = v[0], u1 = v[1];
        phi[0] = __builtin_fmaf(w, u0.x, phi[0]);
        phi[1] = __builtin_fmaf(w, u0.y, phi[1]);
        phi[2] = __builtin_fmaf(w, u0.z, phi[2]);
        phi[3] = __builtin_fmaf(w, u0.w, phi[3]);
        phi[4] = __builtin_fmaf(w, u1.x, phi[4]);
        phi[5] = __builtin_fmaf(w, u1.y, phi[5]);
        phi[6] = __builtin_fmaf(w, u1.z, phi[6]);
        phi[7] = __builtin_fmaf(w, u1.w, phi[7]);
      }
      StepRow row_{};
      row_.phi = make_uint4(pack_bf16(phi[0], phi[1]), pack_bf16(phi[2], phi[3]), pack_bf16(phi[4], phi[5]),
                            pack_bf16(phi[6], phi[7]));
      row_.node = node;
      row_.fx = fx;
      row_.fy = fy;
      row_.fz = fz;
      row_.slot = __float_as_int(p.w);
      row_.yt = (fminf(e, a.norm.clip) - a.norm.mu) * a.norm.inv_sigma;
      R[lane] = row_;
    }
    wave_sync();
    // ---- layer 0: [phi, one-hot] (B, from the rows) x W0^T (A, from L2)
    floatx4 z0[2][4];
    {
      bf16x8 fw0[2][K0S];
#pragma unroll
      for (int jn = 0; jn < 2; jn++)
#pragma unroll
        for (int s = 0; s < K0S; s++)
          fw0[jn][s] = *reinterpret_cast<const bf16x8 *>(a.w0t + size_t(16 * jn + r) * K0 + 32 * s + 8 * grp);
#pragma unroll
      for (int jn = 0; jn < 2; jn++) {
        const float4 b = *reinterpret_cast<const float4 *>(a.b0 + 16 * jn + 4 * grp);
#pragma unroll
        for (int i = 0; i < 4; i++) z0[jn][i] = floatx4{b.x, b.y, b.z, b.w};
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const StepRow &pr = R[16 * i + r];
        const uint4 ph = pr.phi;
        const int slot = pr.slot;
#pragma unroll
        for (int s = 0; s < K0S; s++) {
          // the one-hot column kF + slot, if it falls in this lane's 8 columns
          const int e = kF + slot - (32 * s + 8 * grp);
          const uint32_t one = (e >= 0 && e < 8) ? (0x3f80u << (16 * (e & 1))) : 0u;
          const int wsel = e >> 1;
          uint4 v = make_uint4(wsel == 0 ? one : 0u, wsel == 1 ? one : 0u, wsel == 2 ? one : 0u, wsel == 3 ? one : 0u);
          if (s == 0 && grp == 0) v = ph;  // (columns 0..7: the features; the slot column is >= kF)
          const bf16x8 fb = __builtin_bit_cast(bf16x8, v);
#pragma unroll
          for (int jn = 0; jn < 2; jn++) z0[jn][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw0[jn][s], fb, z0[jn][i], 0, 0, 0);
        }
      }
    }
    // ReLU masks (bit (jn * 4 + i) * 4 + c) and the BF16 activations
    uint32_t m0 = 0, m1 = 0;
    floatx4 a0[2][4];
#pragma unroll
    for (int jn = 0; jn < 2; jn++)
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const float z = z0[jn][i][c];
          m0 |= uint32_t(z > 0.f) << ((jn * 4 + i) * 4 + c);
          a0[jn][i][c] = bf16_round(z > 0.f ? z : 0.f);
        }
    // ---- layer 1
    floatx4 a1[2][4];
    {
      bf16x8 fw1[2];
#pragma unroll
      for (int jn = 0; jn < 2; jn++) fw1[jn] = perm_row(a.w1t + size_t(16 * jn + r) * kH, grp);
      floatx4 z1[2][4];
#pragma unroll
      for (int jn = 0; jn < 2; jn++) {
        const float4 b = *reinterpret_cast<const float4 *>(a.b1 + 16 * jn + 4 * grp);
#pragma unroll
        for (int i = 0; i < 4; i++) z1[jn][i] = floatx4{b.x, b.y, b.z, b.w};
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const bf16x8 fb = acc_operand(a0[0][i], a0[1][i]);
#pragma unroll
        for (int jn = 0; jn < 2; jn++) z1[jn][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw1[jn], fb, z1[jn][i], 0, 0, 0);
      }
#pragma unroll
      for (int jn = 0; jn < 2; jn++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int c = 0; c < 4; c++) {
            const float z = z1[jn][i][c];
            m1 |= uint32_t(z > 0.f) << ((jn * 4 + i) * 4 + c);
            a1[jn][i][c] = bf16_round(z > 0.f ? z : 0.f);
          }
    }
    // ---- output, loss gradient
    float w2[2][4];
#pragma unroll
    for (int jn = 0; jn < 2; jn++)
#pragma unroll
      for (int c = 0; c < 4; c++) w2[jn][c] = lo_bf16(uint32_t(a.w2[size_t(16 * jn + 4 * grp + c) * kOut]));
    const float b2 = a.b2[0];
    float dz3[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      float part = 0.f;
#pragma unroll
      for (int jn = 0; jn < 2; jn++)
#pragma unroll
        for (int c = 0; c < 4; c++) part = __builtin_fmaf(a1[jn][i][c], w2[jn][c], part);
      part += __shfl_xor(part, 16);
      part += __shfl_xor(part, 32);
      const float y = part + b2;
      const float d = y - R[16 * i + r].yt;
      if (grp == i) {
        sq += d * d;
        if (a.pred) a.pred[tile * 64 + 16 * i + r] = y;
      }
      dz3[i] = bf16_round(a.grad_scale * d);
    }
    if (a.backward) {
#pragma unroll
      for (int i = 0; i < 4; i++)
        if (grp == i) R[16 * i + r].dz3 = dz3[i];
      // transposed activations: A0T [n][p] at U, A1T at U + kH kTP
#pragma unroll
      for (int jn = 0; jn < 2; jn++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int c = 0; c < 4; c++) {
            const int n = 16 * jn + 4 * grp + c, p = 16 * i + r;
            U[n * kTP + p] = bf16(pack_bf16(a0[jn][i][c], 0.f));
            U[(kH + n) * kTP + p] = bf16(pack_bf16(a1[jn][i][c], 0.f));
          }
      // dz1 = dz3 W2 * relu'(z1), exact product then BF16
      floatx4 dz1[2][4];
#pragma unroll
      for (int jn = 0; jn < 2; jn++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int c = 0; c < 4; c++)
            dz1[jn][i][c] = ((m1 >> ((jn * 4 + i) * 4 + c)) & 1u) ? bf16_round(dz3[i] * w2[jn][c]) : 0.f;
      wave_sync();
      // dW2 = A1T . dz3 (column 0), db2
      const bf16x8 ones = __builtin_bit_cast(bf16x8, make_uint4(0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u));
      const bf16x8 onesA = r == 0 ? ones : zero8;
#pragma unroll
      for (int s = 0; s < 2; s++) {
        uint32_t dv[4] = {0u, 0u, 0u, 0u};
        if (r == 0)
#pragma unroll
          for (int h = 0; h < 4; h++)
            dv[h] = pack_bf16(R[32 * s + 8 * grp + 2 * h].dz3, R[32 * s + 8 * grp + 2 * h + 1].dz3);
        const bf16x8 fb = __builtin_bit_cast(bf16x8, make_uint4(dv[0], dv[1], dv[2], dv[3]));
#pragma unroll
        for (int jn = 0; jn < 2; jn++) {
          const bf16x8 fa = lds8(U + (kH + 16 * jn + r) * kTP + 32 * s + 8 * grp);
          gW2[jn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, gW2[jn], 0, 0, 0);
        }
      }
      {
        float v = 0.f;
#pragma unroll
        for (int i = 0; i < 4; i++) v += grp == i ? dz3[i] : 0.f;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off);
        gB2 = v;
      }
      wave_sync();
      // DZ1T [n][p] over A1T
#pragma unroll
      for (int jn = 0; jn < 2; jn++)
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
          for (int c = 0; c < 4; c++)
            U[(kH + 16 * jn + 4 * grp + c) * kTP + 16 * i + r] = bf16(pack_bf16(dz1[jn][i][c], 0.f));
      wave_sync();
      // dW1 = A0T . DZ1T^T, db1 = ones . DZ1T^T
#pragma unroll
      for (int s = 0; s < 2; s++) {
        bf16x8 fa[2], fb[2];
#pragma unroll
        for (int j = 0; j < 2; j++) {
          fa[j] = lds8(U + (16 * j + r) * kTP + 32 * s + 8 * grp);
          fb[j] = lds8(U + (kH + 16 * j + r) * kTP + 32 * s + 8 * grp);
        }
#pragma unroll
        for (int jo = 0; jo < 2; jo++) {
#pragma unroll
          for (int ji = 0; ji < 2; ji++) gW1[ji][jo] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[ji], fb[jo], gW1[ji][jo], 0, 0, 0);
          gB1[jo] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(onesA, fb[jo], gB1[jo], 0, 0, 0);
        }
      }
      // dz0 = (dz1 W1^T) * relu'(z0): A = W1 rows (K permuted), B = dz1 from registers
      floatx4 dz0[2][4];
      {
        bf16x8 fw[2];
#pragma unroll
        for (int ji = 0; ji < 2; ji++) fw[ji] = perm_row(a.w1 + size_t(16 * ji + r) * kH, grp);
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const bf16x8 fb = acc_operand(dz1[0][i], dz1[1][i]);
#pragma unroll
          for (int ji = 0; ji < 2; ji++) {
            const floatx4 v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw[ji], fb, floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
            for (int c = 0; c < 4; c++) dz0[ji][i][c] = ((m0 >> ((ji * 4 + i) * 4 + c)) & 1u) ? bf16_round(v[c]) : 0.f;
          }
        }
      }
      // dL/dphi = dz0 W0[0..kF)^T: A = W0 rows 0..kF-1 (K permuted), rows kF..15 zero
      {
        const bf16x8 fw = r < kF ? perm_row(a.w0 + size_t(r) * kH, grp) : zero8;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const floatx4 v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw, acc_operand(dz0[0][i], dz0[1][i]),
                                                                    floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
          if (grp < 2) *reinterpret_cast<float4 *>(&dphi[wave][(16 * i + r) * kF + 4 * grp]) = make_float4(v[0], v[1], v[2], v[3]);
        }
      }
      wave_sync();
      // X0T [k][p] (k < k0) at U, DZ0T [n][p] at U + 64 kTP
      {
        uint4 *z = reinterpret_cast<uint4 *>(U);
        for (int t = lane; t < K0 * kTP / 8; t += 64) z[t] = make_uint4(0u, 0u, 0u, 0u);
        const uint4 ph = R[lane].phi;
        const uint32_t pw[4] = {ph.x, ph.y, ph.z, ph.w};
#pragma unroll
        for (int k = 0; k < kF; k++) U[k * kTP + lane] = bf16((pw[k >> 1] >> (16 * (k & 1))) & 0xffffu);
        U[(kF + R[lane].slot) * kTP + lane] = bf16(0x3f80);
#pragma unroll
        for (int jn = 0; jn < 2; jn++)
#pragma unroll
          for (int i = 0; i < 4; i++)
#pragma unroll
            for (int c = 0; c < 4; c++)
              U[(64 + 16 * jn + 4 * grp + c) * kTP + 16 * i + r] = bf16(pack_bf16(dz0[jn][i][c], 0.f));
      }
      wave_sync();
      // dW0 = X0T . DZ0T^T, db0 = ones . DZ0T^T
#pragma unroll
      for (int s = 0; s < 2; s++) {
        bf16x8 fb[2];
#pragma unroll
        for (int jn = 0; jn < 2; jn++) fb[jn] = lds8(U + (64 + 16 * jn + r) * kTP + 32 * s + 8 * grp);
#pragma unroll
        for (int kb = 0; kb < K0 / 16; kb++) {
          const bf16x8 fa = lds8(U + (16 * kb + r) * kTP + 32 * s + 8 * grp);
#pragma unroll
          for (int jn = 0; jn < 2; jn++) gW0[kb][jn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb[jn], gW0[kb][jn], 0, 0, 0);
        }
#pragma unroll
        for (int jn = 0; jn < 2; jn++) gB0[jn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(onesA, fb[jn], gB0[jn], 0, 0, 0);
      }
      // grid gradient: one point per instruction, lane = (corner, feature)
      {
        const int cq = lane >> 3, k = lane & 7;
        const int dx = cq >> 2, dy = (cq >> 1) & 1, dz = cq & 1;
        const int off = dx * s0 + dy * s1 + dz;
        for (int pt = 0; pt < 64; pt++) {
          const StepRow &pr = R[pt];
          const float w = (dx ? pr.fx : 1.f - pr.fx) * (dy ? pr.fy : 1.f - pr.fy) * (dz ? pr.fz : 1.f - pr.fz);
          const long long v = llrint(double(w * dphi[wave][pt * kF + k]) * kGradScale);
          atomicAdd(a.gacc + size_t(pr.node + off) * kF + k, static_cast<unsigned long long>(v));
        }
      }
    }
  }
  // squared error of the wave's points (fixed point, as the layer kernels' loss)
  for (int off = 32; off > 0; off >>= 1) sq += __shfl_down(sq, off);
  if (lane == 0 && active && a.loss)
    atomicAdd(a.loss, static_cast<unsigned long long>(fmin(double(sq) * mlp::kLossScale + 0.5, 1099511627776.0)));
  if (!a.backward) return;  // (uniform over the launch: no barrier below is reached)
  // ---- the work-group's gradients: each wave's partials to LDS, summed in wave order
  __syncthreads();
  float *red = reinterpret_cast<float *>(&tt[0][0]);
  float *P = red + wave * NP;
  // layout: dW0 [K0][kH], dW1 [kH][kH], dW2 [kH], db0 [kH], db1 [kH], db2
#pragma unroll
  for (int kb = 0; kb < K0 / 16; kb++)
#pragma unroll
    for (int jn = 0; jn < 2; jn++)
#pragma unroll
      for (int c = 0; c < 4; c++) P[(16 * kb + 4 * grp + c) * kH + 16 * jn + r] = gW0[kb][jn][c];
#pragma unroll
  for (int ji = 0; ji < 2; ji++)
#pragma unroll
    for (int jo = 0; jo < 2; jo++)
#pragma unroll
      for (int c = 0; c < 4; c++) P[K0 * kH + (16 * ji + 4 * grp + c) * kH + 16 * jo + r] = gW1[ji][jo][c];
#pragma unroll
  for (int jn = 0; jn < 2; jn++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if (r == 0) P[K0 * kH + kH * kH + 16 * jn + 4 * grp + c] = gW2[jn][c];
      if (grp == 0 && c == 0) {
        P[K0 * kH + kH * kH + kH + 16 * jn + r] = gB0[jn][0];
        P[K0 * kH + kH * kH + 2 * kH + 16 * jn + r] = gB1[jn][0];
      }
    }
  if (lane == 0) P[NP - 1] = gB2;
  __syncthreads();
  const size_t wg = blockIdx.x;
  auto wsum = [&](int o) {  // the waves' partials of entry o, in wave order
    float v = red[o];
#pragma unroll
    for (int w = 1; w < kStepWaves; w++) v += red[w * NP + o];
    return v;
  };
  for (int t = tid; t < K0 * kH; t += NT) a.gw[0][wg * K0 * kH + t] = wsum(t);
  for (int t = tid; t < kH * kH; t += NT) a.gw[1][wg * kH * kH + t] = wsum(K0 * kH + t);
  for (int t = tid; t < kH * kOut; t += NT)  // W2 [kH][kOut]: column 0
    a.gw[2][wg * kH * kOut + t] = t % kOut ? 0.f : wsum(K0 * kH + kH * kH + t / kOut);
  if (tid < kH) {
    const int o0 = K0 * kH + kH * kH + kH + tid, o1 = o0 + kH;
    a.gb[0][wg * kH + tid] = wsum(o0);
    a.gb[1][wg * kH + tid] = wsum(o1);
  }
  if (tid < kOut) a.gb[2][wg * kOut + tid] = tid ? 0.f : wsum(NP - 1);
}

unsigned blocks_for(size_t n, unsigned t) { return unsigned((n + t - 1) / t); }

uint64_t mix_seed(uint64_t seed, uint64_t k) { return (seed ^ (k * 0xd6e8feb86659fd93ull)) * 0x9e3779b97f4a7c15ull; }

using Acts = Workspace;

// A training run's batch rows (see StepArgs::perm).
struct StepRows {
  const uint32_t *perm;
  uint32_t per_epoch;
  mlp::StepState *state;
  mlp::Schedule sched;
};

// The fused step over rows [0, n) of `points` (through idx, if set):
// forward (y into pred, if set; squared error into A.loss), and with
// `backward` the gradients -- grid into net.gacc, weights and biases as
// one slice per work-group into A.gpart / bpart (A.splits slices).
int launch_step(Net &net, Acts &A, const float4 *points, const uint32_t *idx, const float *labels, size_t n,
                size_t batch, mlp::TargetNorm norm, float *pred, bool backward, hipStream_t s,
                const StepRows *rows = nullptr) {
  if (n == 0 || n % 64) return fail(BUDE_ERR_INVALID_ARGUMENT, "atomnet: rows must be a multiple of 64");
  StepArgs a{};
  a.points = points;
  a.idx = idx;
  if (rows) {
    if (rows->per_epoch == 0) return fail(BUDE_ERR_INVALID_ARGUMENT, "atomnet: empty epoch");
    a.perm = rows->perm;
    a.per_epoch = rows->per_epoch;
    a.state = rows->state;
    a.sched = rows->sched;
  }
  a.labels = labels;
  a.n = n;
  a.grid = static_cast<const float *>(net.g.ptr);
  a.g = net.grid;
  a.w0t = static_cast<const bf16 *>(net.w16t[0].ptr);
  a.w0 = static_cast<const bf16 *>(net.w16[0].ptr);
  a.w1t = static_cast<const bf16 *>(net.w16t[1].ptr);
  a.w1 = static_cast<const bf16 *>(net.w16[1].ptr);
  a.w2 = static_cast<const bf16 *>(net.w16[2].ptr);
  a.b0 = static_cast<const float *>(net.b[0].ptr);
  a.b1 = static_cast<const float *>(net.b[1].ptr);
  a.b2 = static_cast<const float *>(net.b[2].ptr);
  a.norm = norm;
  a.grad_scale = 2.0f / float(batch);
  a.pred = pred;
  a.loss = backward ? static_cast<unsigned long long *>(A.loss.ptr) : nullptr;
  a.backward = backward ? 1 : 0;
  a.gacc = static_cast<unsigned long long *>(net.gacc.ptr);
  const unsigned blocks = blocks_for(n, 64 * kStepWaves);
  if (backward) {
    if (A.slices < blocks) return fail(BUDE_ERR_INVALID_ARGUMENT, "atomnet: workspace sized for fewer rows");
    for (int l = 0; l < 3; l++) {
      a.gw[l] = static_cast<float *>(A.gpart[l].ptr);
      a.gb[l] = static_cast<float *>(A.bpart[l].ptr);
      A.splits[l] = int(blocks);
    }
  }
  if (net.k0 == 32) {
    hipLaunchKernelGGL(step_kernel<1>, dim3(blocks), dim3(64 * kStepWaves), 0, s, a);
  } else if (net.k0 == 64) {
    hipLaunchKernelGGL(step_kernel<2>, dim3(blocks), dim3(64 * kStepWaves), 0, s, a);
  } else {
    return fail(BUDE_ERR_INVALID_ARGUMENT, "atomnet: input width must be 32 or 64");
  }
  return hip_check(hipGetLastError(), "atomnet step");
}

// Adam on every parameter with this step's values from the device state,
// which it then advances to the next step.
int adam_step_dev(Net &net, const Acts &A, mlp::StepState *state, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  mlp::AdamBatch ab{};
  for (int l = 0; l < 3; l++) {
    const int K = layer_k(net, l), N = layer_n(l);
    ab.t[2 * l] = {static_cast<float *>(net.w[l].ptr), static_cast<const float *>(A.gpart[l].ptr),
                   static_cast<float *>(net.mw[l].ptr), static_cast<float *>(net.vw[l].ptr),
                   static_cast<bf16 *>(net.w16[l].ptr), static_cast<bf16 *>(net.w16t[l].ptr), K, N,
                   A.splits[l], size_t(K) * N};
    ab.t[2 * l + 1] = {static_cast<float *>(net.b[l].ptr), static_cast<const float *>(A.bpart[l].ptr),
                       static_cast<float *>(net.mb[l].ptr), static_cast<float *>(net.vb[l].ptr), nullptr, nullptr, 1,
                       N, A.splits[l], size_t(N)};
  }
  ab.count = 6;
  // the grid, from its fixed-point accumulators (cleared), in the same launch
  ab.fixed = {static_cast<float *>(net.g.ptr), static_cast<unsigned long long *>(net.gacc.ptr),
              static_cast<float *>(net.gm.ptr), static_cast<float *>(net.gv.ptr), net.nodes * kF, kGradScale};
  return hip_check(mlp::adam_multi_dev(ab, mlp::step_hyper(state), b1, b2, eps, s, state), "atomnet adam");
}

int alloc_zero(DeviceBuffer &d, size_t bytes) {
  if (int st = d.alloc(bytes)) return st;
  return hip_check(hipMemset(d.ptr, 0, bytes), "hipMemset");
}

}  // namespace

int reserve(const Net &net, Workspace &A, size_t B, size_t nval) {
  const size_t slices = (B + 64 * kStepWaves - 1) / (64 * kStepWaves);
  if (A.rows >= B && A.k0 == net.k0 && A.vpred.bytes >= std::max<size_t>(nval, 1) * 4) return BUDE_OK;
  A.rows = B;
  A.k0 = net.k0;
  int st = A.pred.reserve(B * 4);
  if (!st) st = A.loss.reserve(256);
  if (!st) st = A.part.reserve(1024 * 2 * 8);
  if (!st) st = A.vpred.reserve(std::max<size_t>(nval, 1) * 4);
  if (!st) st = A.state.reserve(sizeof(mlp::StepState));
  if (!st) st = A.idx.reserve(B * 4);
  for (int l = 0; l < 3 && !st; l++) {
    const int K = layer_k(net, l), N = layer_n(l);
    st = A.gpart[l].reserve(slices * K * N * 4);
    if (!st) st = A.bpart[l].reserve(slices * N * 4);
  }
  if (!st) A.slices = unsigned(slices);
  return st;
}

int set_params(Net &net, const float *grid, const float *const w[3], const float *const b[3], hipStream_t s) {
  const size_t ng = net.nodes * kF;
  int st = net.g.upload(grid, ng * 4);
  if (!st) st = alloc_zero(net.gm, ng * 4);
  if (!st) st = alloc_zero(net.gv, ng * 4);
  if (!st) st = alloc_zero(net.gacc, ng * 8);
  if (!st) st = net.gzp.alloc(net.nodes * 4 * kF * 4);  // (cells < nodes)
  for (int l = 0; l < 3 && !st; l++) {
    const int K = layer_k(net, l), N = layer_n(l);
    const size_t nw = size_t(K) * N;
    st = net.w[l].upload(w[l], nw * 4);
    if (!st) st = net.b[l].upload(b[l], size_t(N) * 4);
    for (DeviceBuffer *d : {&net.mw[l], &net.vw[l], &net.gw[l]})
      if (!st) st = alloc_zero(*d, nw * 4);
    for (DeviceBuffer *d : {&net.mb[l], &net.vb[l], &net.gb[l]})
      if (!st) st = alloc_zero(*d, size_t(N) * 4);
    if (!st) st = net.w16[l].alloc(nw * 2);
    if (!st) st = net.w16t[l].alloc(nw * 2);
    if (!st)
      st = hip_check(mlp::to_bf16(static_cast<const float *>(net.w[l].ptr), K, N, static_cast<bf16 *>(net.w16[l].ptr),
                                  static_cast<bf16 *>(net.w16t[l].ptr), s),
                     "to_bf16");
  }
  if (!st) st = net.slot_bias.alloc(size_t(net.nslots) * kH * 4);
  net.step = 0;
  if (!st) st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  return st;
}

int init(Net &net, const field::Grid &grid, int nslots, uint64_t seed, hipStream_t s) {
  if (nslots <= 0 || nslots > kMaxSlots)
    return fail(BUDE_ERR_INVALID_ARGUMENT, "atomnet: " + std::to_string(nslots) + " forcefield slots, at most " +
                                               std::to_string(kMaxSlots));
  net.grid = grid;
  net.nslots = nslots;
  net.k0 = input_width(nslots);
  net.nodes = grid.points();
  net.step = 0;
  // buffers kept from call to call (an allocation can wait for the kernels
  // in flight); every value written on the stream below
  const size_t ng = net.nodes * kF;
  int st = net.g.reserve(ng * 4);
  if (!st) st = net.gm.reserve(ng * 4);
  if (!st) st = net.gv.reserve(ng * 4);
  if (!st) st = net.gacc.reserve(ng * 8);
  if (!st) st = net.gzp.reserve(net.nodes * 4 * kF * 4);
  for (int l = 0; l < 3 && !st; l++) {
    const size_t nw = size_t(layer_k(net, l)) * layer_n(l), nb = size_t(layer_n(l));
    for (DeviceBuffer *d : {&net.w[l], &net.mw[l], &net.vw[l], &net.gw[l]})
      if (!st) st = d->reserve(nw * 4);
    for (DeviceBuffer *d : {&net.b[l], &net.mb[l], &net.vb[l], &net.gb[l]})
      if (!st) st = d->reserve(nb * 4);
    if (!st) st = net.w16[l].reserve(nw * 2);
    if (!st) st = net.w16t[l].reserve(nw * 2);
  }
  if (!st) st = net.slot_bias.reserve(size_t(nslots) * kH * 4);
  if (st) return st;
  // grid features N(0, 0.1^2); He-normal hidden weights (first layer: fan-in
  // of the features and one hot input), 1 / sqrt(fan_in) output column;
  // zero biases and Adam state
  hipLaunchKernelGGL(init_kernel, dim3(blocks_for(ng, 256)), dim3(256), 0, s, static_cast<float *>(net.g.ptr),
                     static_cast<float *>(net.gm.ptr), static_cast<float *>(net.gv.ptr), ng, kF, int(net.nodes), kF,
                     0.1f, seed);
  if (int e = hip_check(hipGetLastError(), "atomnet init")) return e;
  if (int e = hip_check(hipMemsetAsync(net.gacc.ptr, 0, ng * 8, s), "hipMemsetAsync")) return e;
  for (int l = 0; l < 3; l++) {
    const int K = layer_k(net, l), N = layer_n(l);
    const size_t nw = size_t(K) * N;
    const int rows = l == 0 ? kF + nslots : K, cols = l == 2 ? 1 : N;
    const float fan_in = l == 0 ? float(kF + 1) : float(K);
    const float sd = l == 2 ? std::sqrt(1.f / fan_in) : std::sqrt(2.f / fan_in);
    hipLaunchKernelGGL(init_kernel, dim3(blocks_for(nw, 256)), dim3(256), 0, s, static_cast<float *>(net.w[l].ptr),
                       static_cast<float *>(net.mw[l].ptr), static_cast<float *>(net.vw[l].ptr), nw, N, rows, cols, sd,
                       mix_seed(seed, uint64_t(l) + 1));
    if (int e = hip_check(hipGetLastError(), "atomnet init")) return e;
    hipLaunchKernelGGL(init_kernel, dim3(1), dim3(256), 0, s, static_cast<float *>(net.b[l].ptr),
                       static_cast<float *>(net.mb[l].ptr), static_cast<float *>(net.vb[l].ptr), size_t(N), N, 0, 0,
                       0.f, seed);
    if (int e = hip_check(hipGetLastError(), "atomnet init")) return e;
    if (int e = hip_check(mlp::to_bf16(static_cast<const float *>(net.w[l].ptr), K, N,
                                       static_cast<bf16 *>(net.w16[l].ptr), static_cast<bf16 *>(net.w16t[l].ptr), s),
                          "to_bf16"))
      return e;
  }
  return BUDE_OK;
}

hipError_t sample_points(const float4 *ligands, const LigandMeta *meta, int nligands, const float *poses,
                         size_t stride, int nslots, uint64_t seed, size_t n, float4 *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (nligands <= 0 || nslots <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(sample_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, ligands, meta, nligands, poses, stride,
                     nslots, seed, n, out);
  return hipGetLastError();
}

hipError_t label_points(const float4 *protein, const int32_t *group_begin, const PairCoef *slot_coef, int ngroups,
                        const float4 *points, size_t n, int form, float *out, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (n % 64) return hipErrorInvalidValue;
  const unsigned blocks = blocks_for(n / 64, 4);
  if (form == kFormD)
    hipLaunchKernelGGL(label_kernel<kFormD>, dim3(blocks), dim3(256), 0, s, protein, group_begin, slot_coef, ngroups,
                       points, n, out);
  else
    hipLaunchKernelGGL(label_kernel<kFormB>, dim3(blocks), dim3(256), 0, s, protein, group_begin, slot_coef, ngroups,
                       points, n, out);
  return hipGetLastError();
}

int train(Net &net, Workspace &A, const float4 *points, const float *labels, const uint32_t *perm, size_t ntrain,
          size_t nval, const TrainParams &p, hipStream_t s, TrainReport *rep) {
  const size_t B = p.batch;
  if (B == 0 || B % 64 || ntrain < B || ntrain % B || nval % 64)
    return fail(BUDE_ERR_INVALID_ARGUMENT, "atomnet train: the training points must be whole batches "
                                           "(a multiple of 64) and the held-out points a multiple of 64");
  if (int st = reserve(net, A, B, nval)) return st;
  // target transform from the training labels
  {
    hipLaunchKernelGGL(stats_kernel, dim3(1024), dim3(256), 0, s, labels, ntrain, static_cast<double *>(A.part.ptr));
    if (int st = hip_check(hipGetLastError(), "atomnet stats")) return st;
    std::vector<double> part(2048);
    if (int st = hip_check(hipMemcpyAsync(part.data(), A.part.ptr, part.size() * 8, hipMemcpyDeviceToHost, s),
                           "hipMemcpyAsync"))
      return st;
    if (int st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")) return st;
    double s1 = 0, s2 = 0;
    for (size_t i = 0; i < 1024; i++) {
      s1 += part[2 * i];
      s2 += part[2 * i + 1];
    }
    const double mean = s1 / double(ntrain), var = std::max(1e-12, s2 / double(ntrain) - mean * mean);
    net.mu = float(mean);
    net.sigma = float(std::sqrt(var));
    net.clip = float(mean + 8.0 * std::sqrt(var));
  }
  const mlp::TargetNorm norm{net.clip, net.mu, 1.0f / net.sigma, -INFINITY};
  const uint32_t steps = std::max<uint32_t>(1, p.steps);
  const uint32_t warm = std::min<uint32_t>(100, steps / 10 + 1);
  const uint32_t step0 = uint32_t(net.step);
  auto *state = static_cast<mlp::StepState *>(A.state.ptr);
  // one training step, every per-step value taken from the device state:
  // the fused step reads its rows through the permutation at the step count
  // and writes the step's learning rates for the two Adam launches, the
  // grid's Adam advances the count
  const StepRows rows{perm, uint32_t(ntrain / B), state,
                      mlp::Schedule{steps, warm, step0, p.lr, p.grid_lr, 0.9f, 0.999f}};
  auto one_step = [&]() -> int {
    if (int st = launch_step(net, A, points, nullptr, labels, B, B, norm, nullptr, true, s, &rows)) return st;
    return adam_step_dev(net, A, state, s);
  };
  if (int st = hip_check(hipMemsetAsync(A.state.ptr, 0, sizeof(mlp::StepState), s), "hipMemsetAsync")) return st;
  if (int st = hip_check(hipMemsetAsync(A.loss.ptr, 0, 8, s), "hipMemsetAsync")) return st;
  const char *env = std::getenv("BUDE_MI355X_TRAIN_GRAPH");
  if (env && env[0] == '0') {
    for (uint32_t step = 0; step < steps; step++)
      if (int st = one_step()) return st;
  } else {
    A.drop_graph();
    if (int st = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"))
      return st;
    const int cst = one_step();
    const hipError_t ce = hipStreamEndCapture(s, &A.graph);
    if (cst) return cst;
    if (int st = hip_check(ce, "hipStreamEndCapture")) return st;
    if (int st = hip_check(hipGraphInstantiate(&A.exec, A.graph, nullptr, nullptr, 0), "hipGraphInstantiate"))
      return st;
    for (uint32_t step = 0; step < steps; step++)
      if (int st = hip_check(hipGraphLaunch(A.exec, s), "hipGraphLaunch")) return st;
  }
  net.step += int(steps);
  TrainReport r{};
  r.steps = steps;
  unsigned long long lsum = 0;
  if (int st = hip_check(hipMemcpyAsync(&lsum, A.loss.ptr, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync")) return st;
  // held-out RMSE in energy units (the training step's forward)
  double se = 0;
  if (nval) {
    DeviceBuffer &vpred = A.vpred;
    if (int st = launch_step(net, A, points + ntrain, nullptr, labels + ntrain, nval, B, norm,
                             static_cast<float *>(vpred.ptr), false, s))
      return st;
    hipLaunchKernelGGL(sq_err_kernel, dim3(256), dim3(256), 0, s, static_cast<const float *>(vpred.ptr),
                       labels + ntrain, nval, net.sigma, net.mu, net.clip, static_cast<double *>(A.part.ptr));
    if (int st = hip_check(hipGetLastError(), "atomnet val")) return st;
    std::vector<double> part(256);
    if (int st = hip_check(hipMemcpyAsync(part.data(), A.part.ptr, 256 * 8, hipMemcpyDeviceToHost, s),
                           "hipMemcpyAsync"))
      return st;
    if (int st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")) return st;
    for (double v : part) se += v;
  } else if (int st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")) {
    return st;
  }
  r.val_rmse = nval ? float(std::sqrt(se / double(nval))) : 0.f;
  r.final_loss = float(double(lsum) / mlp::kLossScale / (double(steps) * double(B)));
  if (rep) *rep = r;
  return finalize(net, s);
}

int forward_backward(Net &net, const float4 *points, const float *target, size_t batch, float *pred, float *loss,
                     float *grad_grid, hipStream_t s) {
  if (batch == 0 || batch % 64) return fail(BUDE_ERR_INVALID_ARGUMENT, "atomnet: batch must be a multiple of 64");
  Acts A;
  if (int st = reserve(net, A, batch, 0)) return st;
  if (int st = hip_check(hipMemsetAsync(A.loss.ptr, 0, 8, s), "hipMemsetAsync")) return st;
  const mlp::TargetNorm norm{INFINITY, 0.f, 1.f, -INFINITY};
  if (int st = launch_step(net, A, points, nullptr, target, batch, batch, norm, static_cast<float *>(A.pred.ptr), true,
                           s))
    return st;
  // the K slices of the weight and bias gradients, summed in order into net.gw / gb
  for (int l = 0; l < 3; l++) {
    const size_t K = size_t(layer_k(net, l)), N = size_t(layer_n(l));
    if (int st = hip_check(mlp::sum_slices(static_cast<const float *>(A.gpart[l].ptr), A.splits[l], K * N,
                                           static_cast<float *>(net.gw[l].ptr), s),
                           "atomnet dW"))
      return st;
    if (int st = hip_check(mlp::sum_slices(static_cast<const float *>(A.bpart[l].ptr), A.splits[l], N,
                                           static_cast<float *>(net.gb[l].ptr), s),
                           "atomnet db"))
      return st;
  }
  const size_t ng = net.nodes * kF;
  hipLaunchKernelGGL(grid_grad_out_kernel, dim3(blocks_for(ng, 256)), dim3(256), 0, s,
                     static_cast<unsigned long long *>(net.gacc.ptr), grad_grid, ng);
  if (int st = hip_check(hipGetLastError(), "atomnet grid grad out")) return st;
  if (int st = hip_check(hipMemcpyAsync(pred, A.pred.ptr, batch * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync"))
    return st;
  unsigned long long l = 0;
  if (int st = hip_check(hipMemcpyAsync(&l, A.loss.ptr, 8, hipMemcpyDeviceToHost, s), "hipMemcpyAsync")) return st;
  if (int st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")) return st;
  *loss = float(double(l) / mlp::kLossScale / double(batch));
  return BUDE_OK;
}

int finalize(Net &net, hipStream_t s) {
  const size_t nz =
      size_t(net.grid.n[0] - 1) * size_t(net.grid.n[1] - 1) * size_t(net.grid.n[2] - 1) * 4 * kF;
  hipLaunchKernelGGL(cell_table_kernel, dim3(blocks_for(nz, 256)), dim3(256), 0, s,
                     static_cast<const float *>(net.g.ptr), net.grid, static_cast<uint32_t *>(net.gzp.ptr));
  if (int st = hip_check(hipGetLastError(), "atomnet cell table")) return st;
  hipLaunchKernelGGL(slot_bias_kernel, dim3(blocks_for(size_t(net.nslots) * kH, 256)), dim3(256), 0, s,
                     static_cast<const float *>(net.b[0].ptr), static_cast<const bf16 *>(net.w16[0].ptr), net.nslots,
                     static_cast<float *>(net.slot_bias.ptr));
  if (int st = hip_check(hipGetLastError(), "atomnet slot bias")) return st;
  if (int st = hip_check(hipMemcpyAsync(&net.b2, net.b[2].ptr, 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync"))
    return st;
  return hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

ScoreArgs score_args(const Net &net) {
  ScoreArgs a{};
  a.gzp = static_cast<const uint32_t *>(net.gzp.ptr);
  a.grid = net.grid;
  a.w0t = static_cast<const bf16 *>(net.w16t[0].ptr);
  a.k0 = net.k0;
  a.slot_bias = static_cast<const float *>(net.slot_bias.ptr);
  a.nslots = net.nslots;
  a.w1t = static_cast<const bf16 *>(net.w16t[1].ptr);
  a.b1 = static_cast<const float *>(net.b[1].ptr);
  a.w2 = static_cast<const bf16 *>(net.w16[2].ptr);
  a.b2 = net.b2;
  a.sigma = net.sigma;
  a.mu = net.mu;
  return a;
}

hipError_t score(const ScoreArgs &a, int64_t ntiles, hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  if (a.nligands <= 0 || a.tile0 < 0 || a.grid.n[0] < 2 || a.grid.n[1] < 2 || a.grid.n[2] < 2 ||
      a.nslots < 1 || a.nslots > kMaxSlots || (ntiles + 3) / 4 > 0x7fffffff ||
      uint64_t(a.grid.n[0] - 1) * uint64_t(a.grid.n[1] - 1) * uint64_t(a.grid.n[2] - 1) * 16 * kF >=
          (uint64_t(1) << 31))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(score_kernel, dim3(unsigned((ntiles + 3) / 4)), dim3(256), 0, s, a, a.tile0 + ntiles);
  return hipGetLastError();
}

}  // namespace atomnet
}  // namespace bude
