// prep.cpp -- host-side preparation shared by every kernel: error plumbing,
// pair-coefficient derivation, protein grouping by forcefield type.
#include "bude_internal.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <numeric>

namespace bude {

namespace {
thread_local std::string g_last_error;
}

int fail(int status, const std::string &msg) {
  g_last_error = msg;
  return status;
}

}  // namespace bude

extern "C" const char *bude_mi355x_last_error(void) { return bude::g_last_error.c_str(); }

extern "C" int bude_mi355x_abi_version(void) { return BUDE_MI355X_ABI_VERSION; }

namespace bude {

// Pair parameters exactly as the reference derives them per (protein,
// ligand) forcefield row (src/serial/fasten.hpp:59-96, FP32), folded into
// the hinge coefficients documented on PairCoef.
PairCoef pair_coef(const BudeFFParams &p, const BudeFFParams &l) {
  const bool lhphb_ltz = l.hphb < 0.f;
  const bool lhphb_gtz = l.hphb > 0.f;
  const float radij = p.radius + l.radius;
  const float r_radij = 1.0f / radij;
  const bool ff = p.hbtype == kHbtypeF && l.hbtype == kHbtypeF;
  const float elcdst1 = ff ? 0.25f : 0.5f;
  const bool type_E = p.hbtype == kHbtypeE || l.hbtype == kHbtypeE;
  const bool phphb_ltz = p.hphb < 0.f;
  const bool phphb_gtz = p.hphb > 0.f;
  const bool phphb_nz = p.hphb != 0.f;
  const float p_hphb = p.hphb * (phphb_ltz && lhphb_gtz ? -1.0f : 1.0f);
  const float l_hphb = l.hphb * (phphb_gtz && lhphb_ltz ? -1.0f : 1.0f);
  // distdslv = -FLT_MAX (both hphb >= 0) switches the term off entirely.
  const bool dslv_on = phphb_nz && (phphb_ltz || lhphb_ltz);
  const float distdslv = phphb_ltz ? (lhphb_ltz ? kNpNpDist : kNpPDist) : kNpPDist;
  const float chrg_init = l.elsc * p.elsc;
  const float dslv_init = p_hphb + l_hphb;

  PairCoef c{};
  c.radij = radij;
  c.neg_rr = -r_radij;
  c.neg_e1 = -elcdst1;
  c.wq = kCnstnt * (type_E ? -std::fabs(chrg_init) : chrg_init);
  c.neg_rd = dslv_on ? -(1.0f / distdslv) : 0.f;
  c.wd = dslv_on ? dslv_init : 0.f;
  c.c0 = 1.0f + radij * elcdst1;  // radij * elcdst1 is exact (power of two)
  c.e0 = dslv_on ? float(1.0 + double(radij) * double(1.0f / distdslv)) : 0.f;
  return c;
}

ProteinGroups group_protein(const BudeAtom *protein, size_t natpro) {
  std::vector<size_t> order(natpro);
  std::iota(order.begin(), order.end(), size_t(0));
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return protein[a].type < protein[b].type; });
  ProteinGroups g;
  g.xyzw.resize(natpro * 4);
  for (size_t i = 0; i < natpro; i++) {
    const BudeAtom &a = protein[order[i]];
    g.xyzw[4 * i + 0] = a.x;
    g.xyzw[4 * i + 1] = a.y;
    g.xyzw[4 * i + 2] = a.z;
    g.xyzw[4 * i + 3] = 0.f;
    if (i == 0 || a.type != protein[order[i - 1]].type) {
      g.begin.push_back(int32_t(i));
      g.type.push_back(a.type);
    }
  }
  g.begin.push_back(int32_t(natpro));
  return g;
}

std::vector<PairCoef> coef_table(const BudeAtom *ligand, size_t natlig, const ProteinGroups &groups,
                                 const BudeFFParams *ff) {
  const size_t ng = groups.ngroups();
  std::vector<PairCoef> t(natlig * ng);
  for (size_t il = 0; il < natlig; il++)
    for (size_t g = 0; g < ng; g++) t[il * ng + g] = pair_coef(ff[groups.type[g]], ff[ligand[il].type]);
  return t;
}

int plain_form() {
  static const int form = [] {
    const char *v = std::getenv("BUDE_MI355X_PLAIN_FORM");
    return (v && (v[0] == 'b' || v[0] == 'B')) ? 0 : 1;
  }();
  return form;
}

int check_deck(const BudeAtom *protein, size_t natpro, const BudeAtom *ligand, size_t natlig,
               const BudeFFParams *ff, size_t ntypes, const char *who) {
  const std::string w(who);
  if (!protein || !ligand || !ff) return fail(BUDE_ERR_INVALID_ARGUMENT, w + ": NULL deck array");
  if (natpro == 0 || natlig == 0 || ntypes == 0)
    return fail(BUDE_ERR_INVALID_ARGUMENT, w + ": empty protein, ligand or forcefield");
  if (natpro > (1u << 24) || natlig > (1u << 24))
    return fail(BUDE_ERR_INVALID_ARGUMENT, w + ": more than 2^24 atoms");
  for (size_t i = 0; i < natpro; i++)
    if (protein[i].type < 0 || size_t(protein[i].type) >= ntypes)
      return fail(BUDE_ERR_INVALID_ARGUMENT, w + ": protein atom " + std::to_string(i) + " has type " +
                                                 std::to_string(protein[i].type) + " outside the forcefield (" +
                                                 std::to_string(ntypes) + " rows)");
  for (size_t i = 0; i < natlig; i++)
    if (ligand[i].type < 0 || size_t(ligand[i].type) >= ntypes)
      return fail(BUDE_ERR_INVALID_ARGUMENT, w + ": ligand atom " + std::to_string(i) + " has type " +
                                                 std::to_string(ligand[i].type) + " outside the forcefield (" +
                                                 std::to_string(ntypes) + " rows)");
  return ok();
}

}  // namespace bude
