// bude_internal.h -- shared internals of libbude_mi355x (not part of the ABI).
#pragma once

#include "bude_mi355x.h"

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace bude {

// ---- error plumbing: status code + thread-local message -------------------
int fail(int status, const std::string &msg);
inline int ok() { return BUDE_OK; }

// ---- forcefield constants (src/bude.h:22-35) ------------------------------
constexpr int kHbtypeF = 70;
constexpr int kHbtypeE = 69;
constexpr float kHardness = 38.0f;
constexpr float kCnstnt = 45.0f;
constexpr float kNpNpDist = 5.5f;
constexpr float kNpPDist = 1.0f;

// Per (ligand atom, protein type) coefficients of the pair energy written as
// three clamped hinges of b = d - radij (d = pair distance):
//
//   E(d) = 76 * sat(b * s_neg_rr)            steric        (b < 0 only)
//        + wq * sat(1 + b * neg_e1)          charge
//        + wd * sat(1 + b * neg_rd)          desolvation
//
// with sat(x) = min(max(x, 0), 1).  This is the reference pair energy
// (src/serial/fasten.hpp:80-123) rearranged, not approximated:
//  * steric: zone1 ? 76 (1 - d/radij) : 0 = 76 relu(-b/radij), and
//    -b/radij < 1 for d > 0, so the upper clamp never binds;
//  * charge: chrg_init * (zone1 ? 1 : 1 - b*elcdst1) * [b < elcdst];
//    elcdst1 = 1/elcdst exactly (0.5 or 0.25), so the factor is sat(1 -
//    b*elcdst1), and 1 - b*elcdst1 > 0 inside the zone, hence
//    type_E ? -|q*f| : q*f == (type_E ? -|q| : q) * f and wq folds the sign
//    and the 45 (CNSTNT);
//  * desolvation: dslv_init * [b < distdslv && hphb_p != 0] *
//    (zone1 ? 1 : 1 - b/distdslv) = wd * sat(1 - b/distdslv), wd = 0 when
//    hphb_p == 0 or distdslv = -FLT_MAX (b < -FLT_MAX never holds).
// Every hinge is continuous at its kinks, so evaluating the zone tests as
// clamps changes results by O(ulp) only.  The charge and desolvation hinges
// can equally be written in d with the offsets c0 / e0 (saves the b = d -
// radij subtraction, at the cost of one rounding of the offset per type
// pair).
struct alignas(32) PairCoef {
  float radij;     // rp + rl
  float neg_rr;    // -1/radij        (FP32 reciprocal, as the reference)
  float neg_e1;    // -elcdst1        (-0.5 or -0.25)
  float wq;        // 45 * (type_E ? -|chrg_init| : chrg_init)
  float neg_rd;    // -1/distdslv     (0 when the desolvation term is off)
  float wd;        // dslv_init       (0 when the desolvation term is off)
  float c0;        // 1 + radij*elcdst1: charge hinge as sat(c0 + neg_e1 * d)
  float e0;        // 1 + radij/distdslv (one rounding): desolvation hinge as sat(e0 + neg_rd * d)
};
static_assert(sizeof(PairCoef) == 32, "PairCoef is 32 bytes");

PairCoef pair_coef(const BudeFFParams &p, const BudeFFParams &l);

// Protein atoms regrouped by forcefield type.  The pair coefficients are
// constant over a group, so the kernels load them once per (ligand atom,
// group) and stream only coordinates inside a group.
struct ProteinGroups {
  std::vector<float> xyzw;        // 4 floats per atom (x, y, z, 0), grouped
  std::vector<int32_t> begin;     // ngroups + 1 offsets into the atoms
  std::vector<int32_t> type;      // forcefield type of each group
  size_t ngroups() const { return type.size(); }
};
ProteinGroups group_protein(const BudeAtom *protein, size_t natpro);

// coef[il * ngroups + g] for every ligand atom il and protein group g.
std::vector<PairCoef> coef_table(const BudeAtom *ligand, size_t natlig, const ProteinGroups &groups,
                                 const BudeFFParams *ff);

// Hinge form used by every kernel: 1 (hinges in d with per-type-pair
// offsets, the default) or 0 (hinges in b = d - radij); the environment
// variable BUDE_MI355X_PLAIN_FORM=b selects 0.  Both evaluate the same energy
// to within FP32 rounding.
int plain_form();

// Argument checks shared by the entry points.
int check_deck(const BudeAtom *protein, size_t natpro, const BudeAtom *ligand, size_t natlig,
               const BudeFFParams *ff, size_t ntypes, const char *who);

}  // namespace bude
