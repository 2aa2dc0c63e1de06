// deck_synth.cpp -- seeded synthetic decks in the reference deck format.
//
// The reference builds its decks from real molecules with makedeck
// (makedeck/main.cpp:49-234); the benchmark decks here are synthetic with
// the distributions the benchmark configurations name.  Every value comes
// from std::mt19937 32-bit words (a fully specified generator) through
// fixed integer/IEEE-double mappings, so a seed yields byte-identical decks
// on every host and compiler (std::uniform_*_distribution is
// implementation-defined and is not used).
#include "bude_mi355x.h"
#include "bude_internal.h"

#include <cmath>
#include <random>

namespace {

// One independent mt19937 stream per (seed, component tag).
uint32_t stream_seed(uint64_t seed, uint64_t tag) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + tag * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return uint32_t(z);
}

enum : uint64_t { TAG_FF = 1, TAG_PROTEIN = 2, TAG_LIGAND = 3, TAG_POSES = 4 };

// Top 24 bits of a word scaled into [lo, hi), evaluated in IEEE double and
// rounded once to float.
float uniform(std::mt19937 &g, double lo, double hi) {
  const double u = double(g() >> 8) * (1.0 / 16777216.0);
  return float(lo + (hi - lo) * u);
}

// Integer in [0, n) by 32x32->64 multiply-shift.
int32_t below(std::mt19937 &g, uint32_t n) { return int32_t((uint64_t(g()) * n) >> 32); }

}  // namespace

extern "C" int bude_synth_forcefield(uint64_t seed, size_t ntypes, BudeFFParams *out) {
  if (!out && ntypes) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_synth_forcefield: out is NULL");
  static const int32_t kHbtypes[3] = {45, 69, 70};  // '-' normal, 'E', 'F' (src/bude.h:31-32)
  std::mt19937 g(stream_seed(seed, TAG_FF));
  for (size_t i = 0; i < ntypes; i++) {
    out[i].hbtype = kHbtypes[below(g, 3)];
    out[i].radius = uniform(g, 1.0, 2.2);
    out[i].hphb = uniform(g, -2.0, 2.0);
    out[i].elsc = uniform(g, -1.0, 1.0);
  }
  return bude::ok();
}

extern "C" int bude_synth_protein(uint64_t seed, size_t natpro, size_t ntypes, float half_width, BudeAtom *out) {
  if ((!out && natpro) || ntypes == 0 || ntypes > 0x7fffffff)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_synth_protein: bad arguments");
  std::mt19937 g(stream_seed(seed, TAG_PROTEIN));
  for (size_t i = 0; i < natpro; i++) {
    out[i].x = uniform(g, -half_width, half_width);
    out[i].y = uniform(g, -half_width, half_width);
    out[i].z = uniform(g, -half_width, half_width);
    out[i].type = below(g, uint32_t(ntypes));
  }
  return bude::ok();
}

extern "C" int bude_synth_ligand(uint64_t seed, size_t natlig, size_t ntypes, float radius, BudeAtom *out) {
  if ((!out && natlig) || ntypes == 0 || ntypes > 0x7fffffff || !(radius > 0.f))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_synth_ligand: bad arguments");
  std::mt19937 g(stream_seed(seed, TAG_LIGAND));
  const double r2 = double(radius) * double(radius);
  for (size_t i = 0; i < natlig; i++) {
    float x, y, z;
    do {  // rejection sampling from the enclosing cube
      x = uniform(g, -radius, radius);
      y = uniform(g, -radius, radius);
      z = uniform(g, -radius, radius);
    } while (double(x) * x + double(y) * y + double(z) * z > r2);
    out[i].x = x;
    out[i].y = y;
    out[i].z = z;
    out[i].type = below(g, uint32_t(ntypes));
  }
  return bude::ok();
}

extern "C" int bude_synth_poses(uint64_t seed, size_t first, size_t count, float max_trans, float *rows6,
                                size_t stride) {
  if ((!rows6 && count) || stride < count)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_synth_poses: bad arguments");
  std::mt19937 g(stream_seed(seed, TAG_POSES));
  g.discard(6ull * first);  // pose i always consumes words [6i, 6i+6)
  const double pi = 3.14159265358979323846;
  for (size_t i = 0; i < count; i++) {
    for (int k = 0; k < 3; k++) rows6[k * stride + i] = uniform(g, -pi, pi);
    for (int k = 3; k < 6; k++) rows6[k * stride + i] = uniform(g, -max_trans, max_trans);
  }
  return bude::ok();
}
