// atomnet.hip -- the learned per-atom field (see atomnet.h): training
// points and their exact labels, the training step (BF16 MFMA GEMMs of
// mlp::gemm, deterministic fixed-point grid gradients, Adam), and the fused
// scorer that sums the network over a ligand's atoms for every pose.
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
__device__ __forceinline__ float hi_bf16(uint32_t v) { return __uint_as_float(v & 0xffff0000u); }

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

// Trilinear corners of position (x, y, z): node index of corner c = (dx, dy,
// dz) bits (x high, z low) and its weight.
struct Corners {
  int node[8];
  float w[8];
};

__device__ __forceinline__ Corners corners(float x, float y, float z, const field::Grid &g, float inv_h) {
  int ix, iy, iz;
  float fx, fy, fz;
  cell(x, g.lo[0], inv_h, g.n[0], ix, fx);
  cell(y, g.lo[1], inv_h, g.n[1], iy, fy);
  cell(z, g.lo[2], inv_h, g.n[2], iz, fz);
  const int s1 = g.n[2], s0 = g.n[1] * g.n[2];
  const int base = ix * s0 + iy * s1 + iz;
  Corners c;
  const float gx[2] = {1.f - fx, fx}, gy[2] = {1.f - fy, fy}, gz[2] = {1.f - fz, fz};
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int dx = (k >> 2) & 1, dy = (k >> 1) & 1, dz = k & 1;
    c.node[k] = base + dx * s0 + dy * s1 + dz;
    c.w[k] = gx[dx] * gy[dy] * gz[dz];
  }
  return c;
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

// Input rows of points idx[b]: [bf16(phi), onehot(slot), 0 ...] (k0 columns)
// and the transposed copy x0t [k0][B], from the FP32 grid.
__global__ void features_kernel(const float4 *__restrict__ points, const uint32_t *__restrict__ idx, size_t B,
                                const float *__restrict__ grid, const field::Grid g, int k0, bf16 *__restrict__ x0,
                                bf16 *__restrict__ x0t) {
  const size_t b = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float4 p = points[idx ? idx[b] : b];
  const int slot = __float_as_int(p.w);
  const Corners c = corners(p.x, p.y, p.z, g, 1.0f / g.h);
  float phi[kF];
#pragma unroll
  for (int k = 0; k < kF; k++) phi[k] = 0.f;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const float4 *v = reinterpret_cast<const float4 *>(grid + size_t(c.node[q]) * kF);
#pragma unroll
    for (int h = 0; h < kF / 4; h++) {
      const float4 u = v[h];
      phi[4 * h + 0] = __builtin_fmaf(c.w[q], u.x, phi[4 * h + 0]);
      phi[4 * h + 1] = __builtin_fmaf(c.w[q], u.y, phi[4 * h + 1]);
      phi[4 * h + 2] = __builtin_fmaf(c.w[q], u.z, phi[4 * h + 2]);
      phi[4 * h + 3] = __builtin_fmaf(c.w[q], u.w, phi[4 * h + 3]);
    }
  }
  uint32_t *row = reinterpret_cast<uint32_t *>(x0 + b * size_t(k0));
#pragma unroll
  for (int k = 0; k < kF; k += 2) {
    const uint32_t u = pack_bf16(phi[k], phi[k + 1]);
    row[k / 2] = u;
    x0t[size_t(k) * B + b] = bf16(u & 0xffffu);
    x0t[size_t(k + 1) * B + b] = bf16(u >> 16);
  }
  for (int k = kF; k < k0; k += 2) {  // one-hot slot columns (BF16 1.0 = 0x3f80)
    const bf16 h0 = k == kF + slot ? bf16(0x3f80) : bf16(0), h1 = k + 1 == kF + slot ? bf16(0x3f80) : bf16(0);
    row[k / 2] = uint32_t(h0) | (uint32_t(h1) << 16);
    x0t[size_t(k) * B + b] = h0;
    x0t[size_t(k + 1) * B + b] = h1;
  }
}

// Grid gradient: every point's dL/dphi (the first kF columns of dx0
// [B][32]) spread to its 8 corners with the trilinear weights, accumulated
// as 2^40 fixed point in 64-bit integer atomics (order-independent, so the
// training is deterministic).
// One wave per point, lane = (corner, feature): each atomic instruction
// then covers the point's 8 corner nodes, 4 cache lines (the two z-corners
// of a pair share one), where one work-item per (point, corner) spread
// every instruction over 8 points' 32 lines.
__global__ __launch_bounds__(256) void grid_grad_kernel(const float4 *__restrict__ points,
                                                        const uint32_t *__restrict__ idx, size_t B,
                                                        const field::Grid g, const float *__restrict__ dx0,
                                                        unsigned long long *__restrict__ acc) {
  const size_t b = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  if (b >= B) return;
  const int lane = threadIdx.x & 63, q = lane >> 3, k = lane & 7;
  const int dx = q >> 2, dy = (q >> 1) & 1, dz = q & 1;
  const float4 p = points[idx ? idx[b] : b];
  const float inv_h = 1.0f / g.h;
  int ix, iy, iz;
  float fx, fy, fz;
  cell(p.x, g.lo[0], inv_h, g.n[0], ix, fx);
  cell(p.y, g.lo[1], inv_h, g.n[1], iy, fy);
  cell(p.z, g.lo[2], inv_h, g.n[2], iz, fz);
  const int node = ((ix + dx) * g.n[1] + iy + dy) * g.n[2] + iz + dz;
  const float w = (dx ? fx : 1.f - fx) * (dy ? fy : 1.f - fy) * (dz ? fz : 1.f - fz);
  const long long v = llrint(double(w * dx0[b * 32 + k]) * kGradScale);
  atomicAdd(acc + size_t(node) * kF + k, static_cast<unsigned long long>(v));
}

// Adam on the grid features from the fixed-point accumulators (which it
// clears), refreshing the BF16 copy.
// With `state` set, this step's rates come from it, and the last kernel of
// the step advances its step count.
__global__ void grid_adam_kernel(float *__restrict__ w, unsigned long long *__restrict__ acc, float *__restrict__ m,
                                 float *__restrict__ v, bf16 *__restrict__ w16, size_t n, float lr, float b1, float b2,
                                 float eps, float bc1, float bc2, mlp::StepState *__restrict__ state) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (state) {
    lr = state->hyper[1];
    bc1 = state->hyper[2];
    bc2 = state->hyper[3];
    if (i == 0) state->step += 1;  // (no kernel of this step reads it after begin_step)
  }
  if (i >= n) return;
  const float gi = float(double(static_cast<long long>(acc[i])) / kGradScale);
  acc[i] = 0ull;
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wi;
  if (w16) w16[i] = bf16(pack_bf16(wi, 0.f) & 0xffffu);
}

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

// FP32 grid -> z-pair table: out[((i n1 + j)(n2 - 1) + l) kF + k] = FP16
// (g[i][j][l][k], g[i][j][l + 1][k]).
__global__ void zpair_kernel(const float *__restrict__ g, const field::Grid grid, uint32_t *__restrict__ out) {
  const size_t n2m = size_t(grid.n[2] - 1);
  const size_t n = size_t(grid.n[0]) * size_t(grid.n[1]) * n2m * kF;
  const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int k = int(t % kF);
  const size_t e = t / kF, l = e % n2m, ij = e / n2m;
  const size_t node = ij * size_t(grid.n[2]) + l;
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

// slot_bias[s][j] = b0[j] + W0[kF + s][j] (W0's BF16 copy, row-major [k0][kH])
// as the BF16 pair hi = bf16(b), lo = bf16(b - hi) packed (hi low).
__global__ void slot_bias_kernel(const float *__restrict__ b0, const bf16 *__restrict__ w16, int nslots,
                                 uint32_t *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots * kH) return;
  const int s = i / kH, j = i % kH;
  const float b = b0[j] + __uint_as_float(uint32_t(w16[size_t(kF + s) * kH + j]) << 16);
  const float hi = __uint_as_float(pack_bf16(b, 0.f) << 16);
  out[i] = (pack_bf16(b, 0.f) & 0xffffu) | (pack_bf16(b - hi, 0.f) << 16);
}

// ---- fused scorer -------------------------------------------------------------
// One wave per 64 poses of one ligand (lane = pose).  Per ligand atom: the
// transformed position, the interpolated features -- four z-pairs of FP16
// features (32 bytes each), one packed FP16 dot product per pair and feature
// against the pair's two trilinear weights, FP32 accumulation -- then the
// decoder on MFMA, as the transpose D[neuron][pose] = W^T . x^T so the
// weights are the (register-held) A operand and the wave's 64 poses four
// 16-pose B operands: layer 1 over K = 32 (the kF features from the wave's
// own LDS rows, then the slot's bias as a BF16 (hi, lo) pair against two
// inputs of value 1, then zeros), ReLU, layer 2 onto b1 with layer 1's
// accumulators as its B operand in registers, ReLU, and the output neuron
// as a dot product with W2 over the lane's neurons, summed over atoms and
// then over the four lane groups that share a pose.
constexpr int kRow = kF;  // bf16 per LDS row: the point's features

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x8 lds8(const bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

__device__ __forceinline__ float fdot2(uint32_t a, uint32_t b, float c) {
  return __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, a), __builtin_bit_cast(half2_t, b), c, false);
}

constexpr int kStage = 128;  // ligand atoms staged in LDS per wave at a time

// The four z-pairs around one transformed ligand atom and its cell
// fractions: the gathers of the next atom are issued before the current
// one is computed, so their latency hides behind its network.
struct AtomFetch {
  uint4 v[4][2];
  float fx, fy, fz;
  int slot;
};

__device__ __forceinline__ void fetch_atom(const ScoreArgs &a, const Xform &x, const float4 la, float inv_h,
                                           AtomFetch &f) {
  f.slot = __float_as_int(la.w);  // (the same in every lane)
  const float px = x.m[3] + la.x * x.m[0] + la.y * x.m[1] + la.z * x.m[2];
  const float py = x.m[7] + la.x * x.m[4] + la.y * x.m[5] + la.z * x.m[6];
  const float pz = x.m[11] + la.x * x.m[8] + la.y * x.m[9] + la.z * x.m[10];
  int ix, iy, iz;
  cell(px, a.grid.lo[0], inv_h, a.grid.n[0], ix, f.fx);
  cell(py, a.grid.lo[1], inv_h, a.grid.n[1], iy, f.fy);
  cell(pz, a.grid.lo[2], inv_h, a.grid.n[2], iz, f.fz);
  // the four z-pairs (dx, dy), 8 FP16 pairs = 32 bytes each (32-bit
  // offsets: the host keeps the table under 2^31 words)
  const uint32_t n1 = uint32_t(a.grid.n[1]), n2m = uint32_t(a.grid.n[2] - 1);
  const uint32_t base = ((uint32_t(ix) * n1 + uint32_t(iy)) * n2m + uint32_t(iz)) * kF;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const uint32_t dx = q >> 1, dy = q & 1;
    const uint4 *e = reinterpret_cast<const uint4 *>(a.gzp + base + (dx * n1 + dy) * n2m * kF);
    f.v[q][0] = e[0];
    f.v[q][1] = e[1];
  }
}

// max(z, 0) of a finite float as one integer max on its bits (negative
// floats are negative integers); fmaxf would add a canonicalisation.
__device__ __forceinline__ float relu(float z) { return __int_as_float(max(__float_as_int(z), 0)); }

__global__ __launch_bounds__(256, 3) void score_kernel(const ScoreArgs a, int64_t tend) {
  __shared__ bf16 act[4][64 * kRow];
  __shared__ float w2s[kH], b1s[kH];
  __shared__ uint32_t sbias[kMaxSlots * kH];
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
  // register-held A fragments of the feature rows of W0^T (k 0 .. kF - 1:
  // lane group 0 only) and of W1^T with its K (input neuron) order permuted
  // so that layer 1's accumulators are layer 2's B operand as they stand:
  // lane (r, grp) holds hidden neurons 16 jn + 4 grp + c of point r (jn =
  // 0, 1; c = 0..3) after layer 1, so layer 2's K slot 8 grp + e is neuron
  // 16 (e / 4) + 4 grp + e % 4
  const bf16x8 zero8 = bf16x8{};
  bf16x8 fw0[kH / 16], fw1[kH / 16];
#pragma unroll
  for (int jn = 0; jn < kH / 16; jn++) {
    fw0[jn] = grp == 0 ? *reinterpret_cast<const bf16x8 *>(a.w0t + size_t(16 * jn + r) * a.k0) : zero8;
    const bf16 *row = a.w1t + size_t(16 * jn + r) * kH + 4 * grp;
    const uint2 lo = *reinterpret_cast<const uint2 *>(row), hi = *reinterpret_cast<const uint2 *>(row + 16);
    fw1[jn] = __builtin_bit_cast(bf16x8, make_uint4(lo.x, lo.y, hi.x, hi.y));
  }
  // B operand of lane group 1 in layer 1: the two unit inputs the slot's
  // bias pair multiplies (k = kF, kF + 1)
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const bf16x8 ones2 = __builtin_bit_cast(bf16x8, u32x4{0x3f803f80u, 0u, 0u, 0u});
  bf16 *buf = act[wave];
  const float inv_h = 1.0f / a.grid.h;
  float part[4] = {0.f, 0.f, 0.f, 0.f};  // this lane's share of pose block i, summed over atoms
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
  // atom i's gathers; the first atom of a new stage reloads the rows
  auto fetch = [&](int i, AtomFetch &f) {
    if (i > 0 && i % kStage == 0) stage(i);
    fetch_atom(a, x, ls[i % kStage], inv_h, f);
  };
  // one atom through the network, its output added to part[]
  auto eval = [&](const AtomFetch &cur) {
    const int slot = __builtin_amdgcn_readfirstlane(cur.slot);
    const float fz = cur.fz;
    const float wx[2] = {1.f - cur.fx, cur.fx}, wy[2] = {1.f - cur.fy, cur.fy};
    float phi[kF];
#pragma unroll
    for (int k = 0; k < kF; k++) phi[k] = 0.f;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float wxy = wx[q >> 1] * wy[q & 1];
      const uint32_t w2 = pack_half2(wxy * (1.f - fz), wxy * fz);
      const uint32_t u[8] = {cur.v[q][0].x, cur.v[q][0].y, cur.v[q][0].z, cur.v[q][0].w,
                             cur.v[q][1].x, cur.v[q][1].y, cur.v[q][1].z, cur.v[q][1].w};
#pragma unroll
      for (int k = 0; k < kF; k++) phi[k] = fdot2(w2, u[k], phi[k]);
    }
    *reinterpret_cast<uint4 *>(buf + lane * kRow) =
        make_uint4(pack_bf16(phi[0], phi[1]), pack_bf16(phi[2], phi[3]), pack_bf16(phi[4], phi[5]),
                   pack_bf16(phi[6], phi[7]));
    wave_sync();
    // layer 1: features and the slot's bias on MFMA
    bf16x8 fa[kH / 16];
#pragma unroll
    for (int jn = 0; jn < kH / 16; jn++) {
      const uint32_t b = sbias[slot * kH + 16 * jn + r];
      fa[jn] = grp == 0 ? fw0[jn] : (grp == 1 ? __builtin_bit_cast(bf16x8, u32x4{b, 0u, 0u, 0u}) : zero8);
    }
    floatx4 acc[kH / 16][4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const bf16x8 fb = grp == 0 ? lds8(buf + (16 * i + r) * kRow) : (grp == 1 ? ones2 : zero8);
#pragma unroll
      for (int jn = 0; jn < kH / 16; jn++)
        acc[jn][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[jn], fb, floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
    }
    wave_sync();  // (the feature rows are rewritten by the next atom)
    // layer 2 onto b1 with layer 1's ReLU outputs, in BF16, as its B
    // operand straight from the accumulators (K permuted, above); ReLU, and
    // the dot with W2 over this lane's neurons
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const floatx4 z0 = acc[0][i], z1 = acc[1][i];
      const bf16x8 fb = __builtin_bit_cast(
          bf16x8, make_uint4(pack_bf16(relu(z0[0]), relu(z0[1])), pack_bf16(relu(z0[2]), relu(z0[3])),
                             pack_bf16(relu(z1[0]), relu(z1[1])), pack_bf16(relu(z1[2]), relu(z1[3]))));
      float sum = part[i];
#pragma unroll
      for (int jn = 0; jn < kH / 16; jn++) {
        const float4 bb = *reinterpret_cast<const float4 *>(b1s + 16 * jn + 4 * grp);
        const float4 ww = *reinterpret_cast<const float4 *>(w2s + 16 * jn + 4 * grp);
        floatx4 z = floatx4{bb.x, bb.y, bb.z, bb.w};
        z = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fw1[jn], fb, z, 0, 0, 0);
        sum = __builtin_fmaf(ww.x, relu(z[0]), sum);
        sum = __builtin_fmaf(ww.y, relu(z[1]), sum);
        sum = __builtin_fmaf(ww.z, relu(z[2]), sum);
        sum = __builtin_fmaf(ww.w, relu(z[3]), sum);
      }
      part[i] = sum;
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
  fetch(0, f0);
  for (int il = 0; il < natlig; il += 2) {
    fetch(min(il + 1, natlig - 1), f1);
    eval(f0);
    fetch(min(il + 2, natlig - 1), f0);
    if (il + 1 < natlig) eval(f1);
  }
  // sum over the four lane groups of pose block i, kept by group i:
  // groups {g, g ^ 2} first (each keeps its half of the blocks), then {g, g ^ 1}
  const bool upper = grp >= 2;
  const float s0 = upper ? part[0] : part[2], s1 = upper ? part[1] : part[3];
  const float m0 = (upper ? part[2] : part[0]) + __shfl_xor(s0, 32);
  const float m1 = (upper ? part[3] : part[1]) + __shfl_xor(s1, 32);
  const bool odd = grp & 1;
  const float sy = (odd ? m1 : m0) + __shfl_xor(odd ? m0 : m1, 16);
  if (j < m.nposes)
    a.out[size_t(m.pose_off) + size_t(j)] = a.sigma * (sy + float(m.natlig) * a.b2) + float(m.natlig) * a.mu;
}

unsigned blocks_for(size_t n, unsigned t) { return unsigned((n + t - 1) / t); }

uint64_t mix_seed(uint64_t seed, uint64_t k) { return (seed ^ (k * 0xd6e8feb86659fd93ull)) * 0x9e3779b97f4a7c15ull; }

mlp::GemmArgs gemm_args(const void *a, const void *b, int M, int N, int K, int lda, int ldb, int epi) {
  mlp::GemmArgs g{};
  g.a = static_cast<const bf16 *>(a);
  g.b = static_cast<const bf16 *>(b);
  g.M = M;
  g.N = N;
  g.K = K;
  g.lda = lda;
  g.ldb = ldb;
  g.epi = epi;
  return g;
}

using Acts = Workspace;

// Forward of B rows whose inputs are in A.x0; training keeps z, a, a^T.
int forward(const Net &net, Acts &A, size_t B, bool training, hipStream_t s) {
  const int R = int(B);
  const void *in = A.x0.ptr;
  int in_ld = net.k0;
  for (int l = 0; l < 2; l++) {
    auto g = gemm_args(in, net.w16t[l].ptr, R, kH, layer_k(net, l), in_ld, layer_k(net, l), mlp::kEpiHiddenRelu);
    g.bias = static_cast<const float *>(net.b[l].ptr);
    g.out_f32 = training ? static_cast<float *>(A.z[l].ptr) : nullptr;
    g.out_bf16 = static_cast<bf16 *>(A.a[l].ptr);
    g.out_t = training ? static_cast<bf16 *>(A.at[l].ptr) : nullptr;
    g.ldo = kH;
    g.ldt = R;
    if (int st = hip_check(mlp::gemm(g, s), "atomnet gemm forward")) return st;
    in = g.out_bf16;
    in_ld = kH;
  }
  auto g = gemm_args(in, net.w16t[2].ptr, R, kOut, kH, kH, kH, mlp::kEpiOut);
  g.bias = static_cast<const float *>(net.b[2].ptr);
  g.out_f32 = static_cast<float *>(A.pred.ptr);
  return hip_check(mlp::gemm(g, s), "atomnet gemm output");
}

// Backward from dz3 (written by the loss kernel): weight and bias gradients
// as K slices into A.gpart / bpart, dL/dx0 (first 32 columns) into A.dx0.
int backward(Net &net, Acts &A, size_t B, hipStream_t s) {
  const int R = int(B);
  const void *dzt_next = A.dz3t.ptr, *dz_next = A.dz3.ptr;
  for (int l = 2; l >= 0; l--) {
    const int K = layer_k(net, l), N = layer_n(l);
    const void *prev_t = l == 0 ? A.x0t.ptr : A.at[l - 1].ptr;
    // dW_l = A_{l-1}^T dZ_l and db_l = column sums of dZ_l, per K slice
    auto gw = gemm_args(prev_t, dzt_next, K, N, R, R, R, mlp::kEpiF32);
    gw.ldo = N;
    gw.ksplit = mlp::auto_ksplit(K, N, R);
    gw.work = gw.out_f32 = static_cast<float *>(A.gpart[l].ptr);
    gw.bias_part = static_cast<float *>(A.bpart[l].ptr);
    gw.defer = 1;
    A.splits[l] = gw.ksplit;
    if (int st = hip_check(mlp::gemm(gw, s), "atomnet gemm dW")) return st;
    if (l == 0) {
      // dL/dx0 for the grid: dz1 W0^T over the first 32 rows of W0
      auto gx = gemm_args(dz_next, net.w16[0].ptr, R, 32, kH, kH, kH, mlp::kEpiF32);
      gx.out_f32 = static_cast<float *>(A.dx0.ptr);
      gx.ldo = 32;
      if (int st = hip_check(mlp::gemm(gx, s), "atomnet gemm dx0")) return st;
      break;
    }
    auto ga = gemm_args(dz_next, net.w16[l].ptr, R, kH, N, N, N, mlp::kEpiBackRelu);
    ga.zprev = static_cast<const float *>(A.z[l - 1].ptr);
    ga.out_bf16 = static_cast<bf16 *>(A.dz[l - 1].ptr);
    ga.out_t = static_cast<bf16 *>(A.dzt[l - 1].ptr);
    ga.ldo = kH;
    ga.ldt = R;
    if (int st = hip_check(mlp::gemm(ga, s), "atomnet gemm dA")) return st;
    dz_next = A.dz[l - 1].ptr;
    dzt_next = A.dzt[l - 1].ptr;
  }
  return BUDE_OK;
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
  if (int st = hip_check(mlp::adam_multi_dev(ab, mlp::step_hyper(state), b1, b2, eps, s), "atomnet adam")) return st;
  const size_t n = net.nodes * kF;
  hipLaunchKernelGGL(grid_adam_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, static_cast<float *>(net.g.ptr),
                     static_cast<unsigned long long *>(net.gacc.ptr), static_cast<float *>(net.gm.ptr),
                     static_cast<float *>(net.gv.ptr), static_cast<bf16 *>(nullptr), n, 0.f, b1, b2, eps, 1.f, 1.f,
                     state);
  return hip_check(hipGetLastError(), "atomnet grid adam");
}

// Forward + loss + backward of one batch (grid gradients accumulated in net.gacc).
int step_grads(Net &net, Acts &A, const float4 *points, const uint32_t *idx, const float *labels, size_t B,
               mlp::TargetNorm norm, hipStream_t s) {
  hipLaunchKernelGGL(features_kernel, dim3(blocks_for(B, 256)), dim3(256), 0, s, points, idx, B,
                     static_cast<const float *>(net.g.ptr), net.grid, net.k0, static_cast<bf16 *>(A.x0.ptr),
                     static_cast<bf16 *>(A.x0t.ptr));
  if (int st = hip_check(hipGetLastError(), "atomnet features")) return st;
  if (int st = forward(net, A, B, true, s)) return st;
  if (int st = hip_check(mlp::loss_grad(static_cast<const float *>(A.pred.ptr), labels, idx, B, norm, nullptr, nullptr,
                                        1.f, static_cast<bf16 *>(A.dz3.ptr), static_cast<bf16 *>(A.dz3t.ptr), int(B),
                                        static_cast<unsigned long long *>(A.loss.ptr), s),
                         "atomnet loss"))
    return st;
  if (int st = backward(net, A, B, s)) return st;
  hipLaunchKernelGGL(grid_grad_kernel, dim3(blocks_for(B * 64, 256)), dim3(256), 0, s, points, idx, B, net.grid,
                     static_cast<const float *>(A.dx0.ptr), static_cast<unsigned long long *>(net.gacc.ptr));
  return hip_check(hipGetLastError(), "atomnet grid grad");
}

int alloc_zero(DeviceBuffer &d, size_t bytes) {
  if (int st = d.alloc(bytes)) return st;
  return hip_check(hipMemset(d.ptr, 0, bytes), "hipMemset");
}

}  // namespace

int reserve(const Net &net, Workspace &A, size_t B, size_t nval) {
  if (A.rows >= B && A.k0 == net.k0 && A.vpred.bytes >= std::max<size_t>(nval, 1) * 4) return BUDE_OK;
  A.rows = B;
  A.k0 = net.k0;
  int st = A.x0.reserve(B * net.k0 * 2);
  if (!st) st = A.x0t.reserve(B * net.k0 * 2);
  for (int l = 0; l < 2 && !st; l++) {
    st = A.z[l].reserve(B * kH * 4);
    if (!st) st = A.a[l].reserve(B * kH * 2);
    if (!st) st = A.at[l].reserve(B * kH * 2);
    if (!st) st = A.dz[l].reserve(B * kH * 2);
    if (!st) st = A.dzt[l].reserve(B * kH * 2);
  }
  if (!st) st = A.pred.reserve(B * 4);
  if (!st) st = A.dz3.reserve(B * kOut * 2);
  if (!st) st = A.dz3t.reserve(B * kOut * 2);
  if (!st) st = A.dx0.reserve(B * 32 * 4);
  if (!st) st = A.loss.reserve(256);
  if (!st) st = A.part.reserve(1024 * 2 * 8);
  if (!st) st = A.vpred.reserve(std::max<size_t>(nval, 1) * 4);
  if (!st) st = A.state.reserve(sizeof(mlp::StepState));
  if (!st) st = A.idx.reserve(B * 4);
  for (int l = 0; l < 3 && !st; l++) {
    const int K = layer_k(net, l), N = layer_n(l);
    const size_t cap = mlp::max_ksplit(K, N);
    st = A.gpart[l].reserve(cap * K * N * 4);
    if (!st) st = A.bpart[l].reserve(cap * N * 4);
  }
  return st;
}

int set_params(Net &net, const float *grid, const float *const w[3], const float *const b[3], hipStream_t s) {
  const size_t ng = net.nodes * kF;
  int st = net.g.upload(grid, ng * 4);
  if (!st) st = alloc_zero(net.gm, ng * 4);
  if (!st) st = alloc_zero(net.gv, ng * 4);
  if (!st) st = alloc_zero(net.gacc, ng * 8);
  if (!st) st = net.gzp.alloc(ng * 4);
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
  if (!st) st = net.slot_bias.alloc(size_t(net.nslots) * kH * 4);  // BF16 pairs
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
  if (!st) st = net.gzp.reserve(ng * 4);
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
  uint32_t *idx = static_cast<uint32_t *>(A.idx.ptr);
  // one training step, every per-step value taken from the device state
  auto one_step = [&]() -> int {
    if (int st = hip_check(mlp::begin_step(state, perm, uint32_t(ntrain / B), uint32_t(B), idx, steps, warm, p.lr,
                                           p.grid_lr, 0.9f, 0.999f, step0, s),
                           "atomnet begin step"))
      return st;
    if (int st = step_grads(net, A, points, idx, labels, B, norm, s)) return st;
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
  // held-out RMSE in energy units (forward of the training kernels)
  double se = 0;
  if (nval) {
    DeviceBuffer &vpred = A.vpred;
    for (size_t v0 = 0; v0 < nval; v0 += B) {
      const size_t nb = std::min(B, nval - v0);
      hipLaunchKernelGGL(features_kernel, dim3(blocks_for(nb, 256)), dim3(256), 0, s, points + ntrain + v0, nullptr, nb,
                         static_cast<const float *>(net.g.ptr), net.grid, net.k0, static_cast<bf16 *>(A.x0.ptr),
                         static_cast<bf16 *>(A.x0t.ptr));
      if (int st = hip_check(hipGetLastError(), "atomnet features")) return st;
      if (int st = forward(net, A, nb, false, s)) return st;
      if (int st = hip_check(hipMemcpyAsync(static_cast<float *>(vpred.ptr) + v0, A.pred.ptr, nb * 4,
                                            hipMemcpyDeviceToDevice, s),
                             "hipMemcpyAsync"))
        return st;
    }
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
  if (int st = step_grads(net, A, points, nullptr, target, batch, norm, s)) return st;
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
  const size_t nz = size_t(net.grid.n[0]) * size_t(net.grid.n[1]) * size_t(net.grid.n[2] - 1) * kF;
  hipLaunchKernelGGL(zpair_kernel, dim3(blocks_for(nz, 256)), dim3(256), 0, s, static_cast<const float *>(net.g.ptr),
                     net.grid, static_cast<uint32_t *>(net.gzp.ptr));
  if (int st = hip_check(hipGetLastError(), "atomnet z-pairs")) return st;
  hipLaunchKernelGGL(slot_bias_kernel, dim3(blocks_for(size_t(net.nslots) * kH, 256)), dim3(256), 0, s,
                     static_cast<const float *>(net.b[0].ptr), static_cast<const bf16 *>(net.w16[0].ptr), net.nslots,
                     static_cast<uint32_t *>(net.slot_bias.ptr));
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
  a.slot_bias = static_cast<const uint32_t *>(net.slot_bias.ptr);
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
      uint64_t(a.grid.n[0]) * uint64_t(a.grid.n[1]) * uint64_t(a.grid.n[2] - 1) * kF >= (uint64_t(1) << 31))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(score_kernel, dim3(unsigned((ntiles + 3) / 4)), dim3(256), 0, s, a, a.tile0 + ntiles);
  return hipGetLastError();
}

}  // namespace atomnet
}  // namespace bude
