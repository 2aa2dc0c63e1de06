// mi355x/fasten.hpp -- the "mi355x" model: Bude<PPWI> implemented by
// libbude_mi355x through its C ABI (include/bude_mi355x.h).
//
// Placed at src/mi355x/fasten.hpp of a miniBUDE tree it is a model like
// src/hip/fasten.hpp (selected with -DMI355X: an `#elif defined(MI355X)`
// include in src/main.cpp and a register_model line in CMakeLists.txt); this
// project's own driver includes it the same way, next to its own bude.h.
//
// fasten() keeps HipBude::fasten's contract (src/hip/fasten.hpp:212-267):
// upload the deck (contextTime), run warmup + iterations launches each timed
// on the host around launch + synchronize (kernelTimes), copy the energies
// back, free the device buffers.  Library failures become
// std::runtime_error / std::invalid_argument with the library's message, as
// HipBude::checkError does (src/hip/fasten.hpp:161-165).
#pragma once

#include "../bude.h"
#include "bude_mi355x.h"

#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef IMPL_CLS
  #error IMPL_CLS was already defined
#endif
#define IMPL_CLS Mi355xBude

static_assert(sizeof(Atom) == sizeof(BudeAtom) && sizeof(FFParams) == sizeof(BudeFFParams),
              "deck records must match the C ABI records");

template <size_t PPWI> class IMPL_CLS final : public Bude<PPWI> {

  static void check(int status) {
    if (status == BUDE_OK) return;
    const std::string msg = std::string("mi355x: ") + bude_mi355x_last_error();
    if (status == BUDE_ERR_INVALID_ARGUMENT || status == BUDE_ERR_UNSUPPORTED) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
  }

  // Owns a library context for the duration of one fasten() call.
  struct Ctx {
    bude_ctx *h = nullptr;
    ~Ctx() {
      if (h) (void)bude_mi355x_destroy(h);
    }
  };

public:
  IMPL_CLS() = default;

  [[nodiscard]] std::string name() override { return "mi355x"; }

  [[nodiscard]] std::vector<Device> enumerateDevices() override {
    int count = 0;
    const int st = bude_mi355x_device_count(&count);
    if (st == BUDE_ERR_NO_DEVICE) return {};
    check(st);
    std::vector<Device> devices;
    for (int i = 0; i < count; i++) {
      char buf[256];
      check(bude_mi355x_device_name(i, buf, sizeof(buf)));
      devices.emplace_back(size_t(i), std::string(buf));
    }
    return devices;
  }

  [[nodiscard]] bool compatible(const Params &, size_t wgsize, size_t device) const override {
    int ok = 0;
    return bude_mi355x_compatible(int(device), PPWI, wgsize, &ok) == BUDE_OK && ok == 1;
  }

  [[nodiscard]] Sample fasten(const Params &p, size_t wgsize, size_t device) const override {
    Sample sample(PPWI, wgsize, p.nposes());
    const float *rows[6];
    for (int k = 0; k < 6; k++) rows[k] = p.poses[k].data();

    Ctx ctx;
    const auto contextStart = now();
    check(bude_mi355x_create(int(device), reinterpret_cast<const BudeAtom *>(p.protein.data()), p.natpro(),
                             reinterpret_cast<const BudeAtom *>(p.ligand.data()), p.natlig(),
                             reinterpret_cast<const BudeFFParams *>(p.forcefield.data()), p.ntypes(), rows,
                             p.nposes(), &ctx.h));
    const auto contextEnd = now();
    sample.contextTime = {contextStart, contextEnd};

    // All launches in one call; the library times each one on the host
    // around launch + stream synchronize and copies the energies out after
    // the last one, outside every timed bracket.
    std::vector<double> hostMs(p.totalIterations());
    check(bude_mi355x_fasten(ctx.h, PPWI, wgsize, p.warmupIterations, p.iterations, sample.energies.data(),
                             hostMs.data(), nullptr));
    auto t = contextEnd;
    for (double ms : hostMs) {
      const auto t1 = t + std::chrono::duration_cast<TimePoint::duration>(std::chrono::duration<double, std::milli>(ms));
      sample.kernelTimes.emplace_back(t, t1);
      t = t1;
    }
    return sample;
  }
};
