// ref_harness.cpp -- test infrastructure only.  A C-ABI wrapper, written for
// this project, around the reference's own per-pose kernel
// bude::kernel::fasten_main (makedeck/ref-kernel.cpp:15-139), the function
// that produced data/bm1/ref_energies.out and data/bm2/ref_energies.out.
// It is compiled from the reference sources where they lie (see
// oracle/Makefile) into oracle/_ref/libref_kernel.so and used only to pin the
// oracle and to write reference energies for the synthetic decks.
#include "ref-kernel.h"

#include <cstddef>
#include <cstring>
#include <vector>

extern "C" void ref_fasten(size_t natlig, size_t natpro, const void *protein, const void *ligand, const void *ff,
                           size_t ntypes, const float *p0, const float *p1, const float *p2, const float *p3,
                           const float *p4, const float *p5, size_t first, size_t count, float *energies) {
  std::vector<bude::Atom> pro(natpro), lig(natlig);
  std::vector<bude::FFParams> f(ntypes);
  std::memcpy(pro.data(), protein, natpro * sizeof(bude::Atom));
  std::memcpy(lig.data(), ligand, natlig * sizeof(bude::Atom));
  std::memcpy(f.data(), ff, ntypes * sizeof(bude::FFParams));
  std::vector<float> t0(p0 + first, p0 + first + count), t1(p1 + first, p1 + first + count),
      t2(p2 + first, p2 + first + count), t3(p3 + first, p3 + first + count), t4(p4 + first, p4 + first + count),
      t5(p5 + first, p5 + first + count);
  std::vector<float> out(count);
#pragma omp parallel for schedule(dynamic, 16)
  for (long long pose = 0; pose < (long long)count; pose++)
    bude::kernel::fasten_main(int(natlig), int(natpro), pro, lig, t0, t1, t2, t3, t4, t5, out, f, int(pose));
  std::memcpy(energies, out.data(), count * sizeof(float));
}
