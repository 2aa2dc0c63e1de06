// bude.h -- the plugin interface of the miniBUDE driver, restated for this
// project's own driver.  The names, members and meaning are those of the
// reference's src/bude.h:11-101 (a model written against one compiles
// against the other): the deck records are the packed 16-byte Atom /
// FFParams of the deck format, Params is the driver's parsed deck + run
// configuration, Sample is what a model returns for one (PPWI, wgsize), and
// Bude<PPWI> is the model interface the driver instantiates per PPWI.
#pragma once

#include <array>
#include <chrono>
#include <cstdint>
#include <limits>
#include <optional>
#include <string>
#include <utility>
#include <vector>

#define DIFF_TOLERANCE_PCT 0.025f
#define DEFAULT_ITERS 8
#define DEFAULT_ENERGY_ENTRIES 8

#define DATA_DIR "../data/bm1"
#define FILE_LIGAND "/ligand.in"
#define FILE_PROTEIN "/protein.in"
#define FILE_FORCEFIELD "/forcefield.in"
#define FILE_POSES "/poses.in"
#define FILE_REF_ENERGIES "/ref_energies.out"

struct __attribute__((__packed__)) Atom {
  float x, y, z;
  int32_t type;
};

struct __attribute__((__packed__)) FFParams {
  int32_t hbtype;
  float radius;
  float hphb;
  float elsc;
};

static_assert(sizeof(Atom) == 16 && sizeof(FFParams) == 16, "deck records are 16 bytes");

struct Params {
  std::vector<Atom> protein;
  std::vector<Atom> ligand;
  std::vector<FFParams> forcefield;
  std::array<std::vector<float>, 6> poses;  // tilt, roll, pan, x, y, z rows

  std::vector<float> refEnergies;
  size_t maxPoses = 0, iterations = DEFAULT_ITERS, warmupIterations = 2, outRows = DEFAULT_ENERGY_ENTRIES;
  std::string deckDir, output, deviceSelector;
  bool csv = false;
  bool list = false;

  [[nodiscard]] size_t totalIterations() const { return iterations + warmupIterations; }
  [[nodiscard]] size_t natpro() const { return protein.size(); }
  [[nodiscard]] size_t natlig() const { return ligand.size(); }
  [[nodiscard]] size_t ntypes() const { return forcefield.size(); }
  [[nodiscard]] size_t nposes() const { return poses[0].size(); }
};

using TimePoint = std::chrono::high_resolution_clock::time_point;

[[nodiscard]] static inline TimePoint now() { return std::chrono::high_resolution_clock::now(); }

[[nodiscard]] static inline double elapsedMillis(const TimePoint &a, const TimePoint &b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Sample {
  size_t ppwi, wgsize;
  std::vector<float> energies;
  std::vector<std::pair<TimePoint, TimePoint>> kernelTimes;
  std::optional<std::pair<TimePoint, TimePoint>> contextTime;
  Sample(size_t ppwi, size_t wgsize, size_t nposes) : ppwi(ppwi), wgsize(wgsize), energies(nposes) {}
};

using Device = std::pair<size_t, std::string>;

template <size_t PPWI> class Bude {
public:
  virtual ~Bude() = default;
  [[nodiscard]] virtual std::string name() = 0;
  [[nodiscard]] virtual std::vector<Device> enumerateDevices() = 0;
  [[nodiscard]] virtual bool compatible(const Params &, size_t, size_t) const { return true; }
  [[nodiscard]] virtual Sample fasten(const Params &p, size_t wgsize, size_t device) const = 0;
};
