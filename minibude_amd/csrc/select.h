// select.h -- per-segment top-k: the k lowest values of every segment of a
// concatenated score array (one segment per (protein, ligand) pair of a
// screen), by radix selection on order-preserving integer keys.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "hip_util.h"

namespace bude {
namespace sel {

struct Segment {
  int64_t off;  // first value in x (also the first output slot)
  uint32_t n;   // values
  uint32_t k;   // 1 <= k <= n
};
static_assert(sizeof(Segment) == 16, "Segment is 16 bytes");

// Device state of one selection over a fixed segment list: the segments, the
// chunk table the passes iterate over and the per-segment histograms.
struct Plan {
  size_t nseg = 0, nchunks = 0;
  DeviceBuffer segs, chunks, hist;
  // Uploads the segment list and sizes the work tables.
  int prepare(const std::vector<Segment> &s);
};

// For every segment s: out_idx[s.off + j], j < out_n[s], are the positions
// (relative to s.off) of every value <= the k-th smallest value of the
// segment (ties at the threshold are all kept, so out_n[s] >= k), in no
// particular order.  x, out_idx (as long as x) and out_n (nseg) are device
// memory.
hipError_t select_lowest_segments(const Plan &p, const float *x, uint32_t *out_idx, uint32_t *out_n, hipStream_t s);

}  // namespace sel
}  // namespace bude
