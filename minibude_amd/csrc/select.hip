// select.hip -- per-segment top-k by three radix passes (11 + 11 + 10 bits)
// over order-preserving keys.  Every pass streams each segment's values once
// in chunks of 16,384 (one work-group per chunk, LDS histogram merged into
// the segment's global histogram), one wave per segment then narrows the
// key prefix; a last streaming pass writes the selected positions with one
// atomic per wave.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "select.h"

namespace bude {
namespace sel {

namespace {

constexpr int kBins = 2048;
constexpr int kHistStride = kBins + 4;  // bins, then prefix, mask, k remaining
constexpr uint32_t kChunk = 16384;

struct Chunk {
  uint32_t seg;
  uint32_t begin;  // relative to the segment start
};

// IEEE-754 float -> uint32 with the same order (negative values reversed).
__device__ __forceinline__ uint32_t order_key(float x) {
  const uint32_t u = __float_as_uint(x);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(256) void seg_init_kernel(const Segment *__restrict__ segs, uint32_t *__restrict__ hist,
                                                       uint32_t *__restrict__ out_n) {
  uint32_t *h = hist + size_t(blockIdx.x) * kHistStride;
  for (int i = threadIdx.x; i < kBins; i += 256) h[i] = 0;
  if (threadIdx.x == 0) {
    h[kBins] = 0;
    h[kBins + 1] = 0;
    h[kBins + 2] = segs[blockIdx.x].k;
    out_n[blockIdx.x] = 0;
  }
}

__global__ __launch_bounds__(256) void seg_hist_kernel(const Segment *__restrict__ segs,
                                                       const Chunk *__restrict__ chunks, const float *__restrict__ x,
                                                       int shift, uint32_t digit_mask, uint32_t *__restrict__ hist) {
  __shared__ uint32_t local[kBins];
  for (int i = threadIdx.x; i < kBins; i += 256) local[i] = 0;
  __syncthreads();
  const Chunk c = chunks[blockIdx.x];
  const Segment sg = segs[c.seg];
  uint32_t *h = hist + size_t(c.seg) * kHistStride;
  const uint32_t prefix = h[kBins], mask = h[kBins + 1];
  const uint32_t end = min(c.begin + kChunk, sg.n);
  const float *xs = x + sg.off;
  for (uint32_t j = c.begin + threadIdx.x; j < end; j += 256) {
    const uint32_t key = order_key(xs[j]);
    if ((key & mask) == prefix) atomicAdd(&local[(key >> shift) & digit_mask], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kBins; i += 256)
    if (local[i]) atomicAdd(&h[i], local[i]);
}

// One wave per segment: the bin holding the k-th smallest key extends the
// prefix; the bins are cleared for the next pass.
__global__ __launch_bounds__(64) void seg_pick_kernel(uint32_t *__restrict__ hist, int shift, int bits) {
  uint32_t *h = hist + size_t(blockIdx.x) * kHistStride;
  const int lane = threadIdx.x;
  const uint32_t nb = 1u << bits, per = nb / 64;
  uint32_t sum = 0;
  for (uint32_t i = 0; i < per; i++) sum += h[lane * per + i];
  uint32_t inc = sum;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t v = __shfl_up(inc, off);
    if (lane >= off) inc += v;
  }
  const uint32_t k = h[kBins + 2];
  const unsigned long long hit = __ballot(inc >= k);
  const int owner = hit ? __ffsll(static_cast<long long>(hit)) - 1 : 63;
  if (lane == owner) {
    uint32_t cum = inc - sum, bin = lane * per;
    const uint32_t last = (lane + 1) * per - 1;
    while (bin < last && cum + h[bin] < k) cum += h[bin++];
    h[kBins] |= bin << shift;
    h[kBins + 1] |= (nb - 1) << shift;
    h[kBins + 2] = k - cum;
  }
  __syncthreads();
  for (int i = lane; i < kBins; i += 64) h[i] = 0;
}

// Two passes over the chunk: count (per wave, over its quarter of the
// chunk), one atomic per chunk for the chunk's output base, then write with
// per-wave running offsets.  (An atomic per wave and step contends on the
// few per-segment counters.)
__global__ __launch_bounds__(256) void seg_compact_kernel(const Segment *__restrict__ segs,
                                                          const Chunk *__restrict__ chunks,
                                                          const float *__restrict__ x, const uint32_t *__restrict__ hist,
                                                          uint32_t *__restrict__ out_idx, uint32_t *__restrict__ out_n) {
  __shared__ uint32_t wave_base[4];
  const Chunk c = chunks[blockIdx.x];
  const Segment sg = segs[c.seg];
  const uint32_t thr = hist[size_t(c.seg) * kHistStride + kBins];
  const uint32_t end = min(c.begin + kChunk, sg.n);
  const float *xs = x + sg.off;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr uint32_t kQuarter = kChunk / 4;
  const uint32_t wb = c.begin + uint32_t(wave) * kQuarter, we = min(wb + kQuarter, end);
  uint32_t count = 0;
  for (uint32_t j0 = wb; j0 < we; j0 += 64) {
    const uint32_t j = j0 + lane;
    count += uint32_t(__popcll(__ballot(j < we && order_key(xs[j]) <= thr)));
  }
  if (lane == 0) wave_base[wave] = count;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t c0 = wave_base[0], c1 = wave_base[1], c2 = wave_base[2], c3 = wave_base[3];
    const uint32_t base = (c0 + c1 + c2 + c3) ? atomicAdd(&out_n[c.seg], c0 + c1 + c2 + c3) : 0u;
    wave_base[0] = base;
    wave_base[1] = base + c0;
    wave_base[2] = base + c0 + c1;
    wave_base[3] = base + c0 + c1 + c2;
  }
  __syncthreads();
  if (count == 0) return;
  uint32_t pos = wave_base[wave];
  for (uint32_t j0 = wb; j0 < we; j0 += 64) {
    const uint32_t j = j0 + lane;
    const bool take = j < we && order_key(xs[j]) <= thr;
    const unsigned long long m = __ballot(take);
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
    if (take) out_idx[sg.off + pos + below] = j;
    pos += uint32_t(__popcll(m));
  }
}

}  // namespace

int Plan::prepare(const std::vector<Segment> &s) {
  std::vector<Chunk> c;
  for (size_t i = 0; i < s.size(); i++) {
    if (s[i].n == 0 || s[i].k == 0 || s[i].k > s[i].n || s[i].off < 0)
      return fail(BUDE_ERR_INVALID_ARGUMENT, "select: every segment needs 1 <= k <= n");
    for (uint32_t b = 0; b < s[i].n; b += kChunk) {
      c.push_back(Chunk{uint32_t(i), b});
      if (b > 0xffffffffu - kChunk) break;
    }
  }
  if (s.empty() || c.size() > 0x7fffffffu) return fail(BUDE_ERR_INVALID_ARGUMENT, "select: no segments (or too many)");
  nseg = s.size();
  nchunks = c.size();
  int st = segs.upload(s.data(), s.size() * sizeof(Segment));
  if (!st) st = chunks.upload(c.data(), c.size() * sizeof(Chunk));
  if (!st) st = hist.reserve(nseg * kHistStride * sizeof(uint32_t));
  return st;
}

hipError_t select_lowest_segments(const Plan &p, const float *x, uint32_t *out_idx, uint32_t *out_n, hipStream_t s) {
  if (p.nseg == 0) return hipErrorInvalidValue;
  const auto *segs = static_cast<const Segment *>(p.segs.ptr);
  const auto *chunks = static_cast<const Chunk *>(p.chunks.ptr);
  auto *hist = static_cast<uint32_t *>(p.hist.ptr);
  hipLaunchKernelGGL(seg_init_kernel, dim3(unsigned(p.nseg)), dim3(256), 0, s, segs, hist, out_n);
  const int shifts[3] = {21, 10, 0}, bits[3] = {11, 11, 10};
  for (int q = 0; q < 3; q++) {
    hipLaunchKernelGGL(seg_hist_kernel, dim3(unsigned(p.nchunks)), dim3(256), 0, s, segs, chunks, x, shifts[q],
                       (1u << bits[q]) - 1u, hist);
    hipLaunchKernelGGL(seg_pick_kernel, dim3(unsigned(p.nseg)), dim3(64), 0, s, hist, shifts[q], bits[q]);
  }
  hipLaunchKernelGGL(seg_compact_kernel, dim3(unsigned(p.nchunks)), dim3(256), 0, s, segs, chunks, x, hist, out_idx,
                     out_n);
  return hipGetLastError();
}

}  // namespace sel
}  // namespace bude
