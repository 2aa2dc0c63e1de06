/*
 * bude_oracle.c -- CPU restatement of miniBUDE's fasten_main, used ONLY as
 * test infrastructure.
 *
 * This file is the checker, never the product: only tests/, the smoke check
 * in __graft_entry__.py and the cpu_baseline leg of bench.py may load it.
 * The product path (libbude_mi355x, the driver) never links or calls it.
 *
 * Two restatements of the same per-pose energy:
 *   bude_oracle_fasten_f32  FP32, in the reference's evaluation order
 *                           (ligand atom outer, protein atom inner; per pair
 *                           steric, then charge, then desolvation; x0.5).
 *                           Follows src/serial/fasten.hpp:18-132 and
 *                           makedeck/ref-kernel.cpp:15-139 (PPWI = 1).
 *   bude_oracle_fasten_f64  the same formula evaluated in FP64 (the "FP64
 *                           host oracle" every configuration reports its max
 *                           relative error against).
 * plus the pair-energy and transform primitives, exposed so closed-form
 * tests can pin each zone of the piecewise-linear pair energy.
 *
 * Built strict (no -ffast-math, -ffp-contract=off) so the FP32 path is
 * deterministic and matches a strict build of the reference bit for bit.
 *
 * Pinning (tests/test_oracle_*.py): golden energies data/bm1 and data/bm2
 * (ref_energies.out), the reference's own ref-kernel/serial build under
 * oracle/_ref, closed forms for every zone and hbtype/hphb branch, and
 * invariants (additivity, permutation, rigid motion, far field = 0).
 */
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>

/* Packed 16-byte records, identical to src/bude.h:39-49. */
typedef struct __attribute__((packed)) {
  float x, y, z;
  int32_t type;
} OAtom;

typedef struct __attribute__((packed)) {
  int32_t hbtype;
  float radius;
  float hphb;
  float elsc;
} OFF;

/* Constants, src/bude.h:22-35. */
#define O_CNSTNT 45.0
#define O_HBTYPE_F 70
#define O_HBTYPE_E 69
#define O_HARDNESS 38.0
#define O_NPNPDIST 5.5
#define O_NPPDIST 1.0

/* ---------------------------------------------------------------------- */
/* FP32 restatement                                                        */
/* ---------------------------------------------------------------------- */

/* Pose -> 3x4 transform, src/serial/fasten.hpp:32-50 (rows R = Rz*Ry*Rx). */
void bude_oracle_transform_f32(float t0, float t1, float t2, float t3, float t4, float t5, float out[12]) {
  const float sx = sinf(t0), cx = cosf(t0);
  const float sy = sinf(t1), cy = cosf(t1);
  const float sz = sinf(t2), cz = cosf(t2);
  out[0] = cy * cz;
  out[1] = sx * sy * cz - cx * sz;
  out[2] = cx * sy * cz + sx * sz;
  out[3] = t3;
  out[4] = cy * sz;
  out[5] = sx * sy * sz + cx * cz;
  out[6] = cx * sy * sz - sx * cz;
  out[7] = t4;
  out[8] = -sy;
  out[9] = sx * cy;
  out[10] = cx * cy;
  out[11] = t5;
}

/* One (protein atom, ligand atom) pair at distance-vector (x,y,z), FP32,
 * returning steric + charge + desolvation added to *etot in that order
 * (src/serial/fasten.hpp:80-123). */
static inline void pair_f32(const OFF *pp, const OFF *lp, float x, float y, float z, float *etot) {
  const int lhphb_ltz = lp->hphb < 0.f;
  const int lhphb_gtz = lp->hphb > 0.f;
  const float radij = pp->radius + lp->radius;
  const float r_radij = 1.0f / radij;
  const int ff = (pp->hbtype == O_HBTYPE_F && lp->hbtype == O_HBTYPE_F);
  const float elcdst = ff ? 4.0f : 2.0f;
  const float elcdst1 = ff ? 0.25f : 0.5f;
  const int type_E = (pp->hbtype == O_HBTYPE_E || lp->hbtype == O_HBTYPE_E);
  const int phphb_ltz = pp->hphb < 0.f;
  const int phphb_gtz = pp->hphb > 0.f;
  const int phphb_nz = pp->hphb != 0.f;
  const float p_hphb = pp->hphb * (phphb_ltz && lhphb_gtz ? -1.0f : 1.0f);
  const float l_hphb = lp->hphb * (phphb_gtz && lhphb_ltz ? -1.0f : 1.0f);
  const float distdslv = phphb_ltz ? (lhphb_ltz ? 5.5f : 1.0f) : (lhphb_ltz ? 1.0f : -FLT_MAX);
  const float r_distdslv = 1.0f / distdslv;
  const float chrg_init = lp->elsc * pp->elsc;
  const float dslv_init = p_hphb + l_hphb;

  const float distij = sqrtf(x * x + y * y + z * z);
  const float distbb = distij - radij;
  const int zone1 = distbb < 0.f;

  *etot += (1.0f - distij * r_radij) * (zone1 ? 2.0f * 38.0f : 0.f);

  float chrg_e = chrg_init * ((zone1 ? 1.0f : (1.0f - distbb * elcdst1)) * (distbb < elcdst ? 1.0f : 0.f));
  const float neg_chrg_e = -fabsf(chrg_e);
  chrg_e = type_E ? neg_chrg_e : chrg_e;
  *etot += chrg_e * 45.0f;

  const float coeff = 1.0f - distbb * r_distdslv;
  float dslv_e = dslv_init * ((distbb < distdslv && phphb_nz) ? 1.0f : 0.f);
  dslv_e *= zone1 ? 1.0f : coeff;
  *etot += dslv_e;
}

/* Energy of one pose, FP32, reference order (makedeck/ref-kernel.cpp:29-138). */
static float pose_f32(size_t natlig, size_t natpro, const OAtom *protein, const OAtom *ligand, const OFF *ff,
                      const float t[12]) {
  float etot = 0.f;
  for (size_t il = 0; il < natlig; il++) {
    const OAtom la = ligand[il];
    const OFF *lp = &ff[la.type];
    const float lx = t[3] + la.x * t[0] + la.y * t[1] + la.z * t[2];
    const float ly = t[7] + la.x * t[4] + la.y * t[5] + la.z * t[6];
    const float lz = t[11] + la.x * t[8] + la.y * t[9] + la.z * t[10];
    for (size_t ip = 0; ip < natpro; ip++) {
      const OAtom pa = protein[ip];
      pair_f32(&ff[pa.type], lp, lx - pa.x, ly - pa.y, lz - pa.z, &etot);
    }
  }
  return etot * 0.5f;
}

/* energies[i] for poses i in [first, first+count), SoA pose rows as in the
 * deck (src/main.cpp:294-311). */
void bude_oracle_fasten_f32(size_t natlig, size_t natpro, const void *protein, const void *ligand, const void *ff,
                            const float *p0, const float *p1, const float *p2, const float *p3, const float *p4,
                            const float *p5, size_t first, size_t count, float *energies) {
  const OAtom *pro = (const OAtom *)protein;
  const OAtom *lig = (const OAtom *)ligand;
  const OFF *f = (const OFF *)ff;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long k = 0; k < (long long)count; k++) {
    const size_t i = first + (size_t)k;
    float t[12];
    bude_oracle_transform_f32(p0[i], p1[i], p2[i], p3[i], p4[i], p5[i], t);
    energies[k] = pose_f32(natlig, natpro, pro, lig, f, t);
  }
}

/* ---------------------------------------------------------------------- */
/* FP64 restatement (same formula, every operation in double)              */
/* ---------------------------------------------------------------------- */

/* Pair energy as a function of the distance only (double).  Separate terms
 * are returned for the closed-form tests; the sum is the pair energy before
 * the final x0.5 of the pose. */
void bude_oracle_pair_terms_f64(const void *pparams, const void *lparams, double distij, double out[3]) {
  const OFF *pp = (const OFF *)pparams;
  const OFF *lp = (const OFF *)lparams;
  const int lhphb_ltz = lp->hphb < 0.f;
  const int lhphb_gtz = lp->hphb > 0.f;
  const double radij = (double)pp->radius + (double)lp->radius;
  const int ff = (pp->hbtype == O_HBTYPE_F && lp->hbtype == O_HBTYPE_F);
  const double elcdst = ff ? 4.0 : 2.0;
  const double elcdst1 = ff ? 0.25 : 0.5;
  const int type_E = (pp->hbtype == O_HBTYPE_E || lp->hbtype == O_HBTYPE_E);
  const int phphb_ltz = pp->hphb < 0.f;
  const int phphb_gtz = pp->hphb > 0.f;
  const int phphb_nz = pp->hphb != 0.f;
  const double p_hphb = (double)pp->hphb * (phphb_ltz && lhphb_gtz ? -1.0 : 1.0);
  const double l_hphb = (double)lp->hphb * (phphb_gtz && lhphb_ltz ? -1.0 : 1.0);
  /* -FLT_MAX branch: distbb < -FLT_MAX never holds, so the term is off. */
  const int dslv_on = phphb_nz && (phphb_ltz || lhphb_ltz);
  const double distdslv = phphb_ltz ? (lhphb_ltz ? O_NPNPDIST : O_NPPDIST) : (lhphb_ltz ? O_NPPDIST : 0.0);
  const double chrg_init = (double)lp->elsc * (double)pp->elsc;
  const double dslv_init = p_hphb + l_hphb;

  const double distbb = distij - radij;
  const int zone1 = distbb < 0.0;
  out[0] = zone1 ? (1.0 - distij / radij) * (2.0 * O_HARDNESS) : 0.0;
  double chrg_e = chrg_init * ((zone1 ? 1.0 : (1.0 - distbb * elcdst1)) * (distbb < elcdst ? 1.0 : 0.0));
  if (type_E) chrg_e = -fabs(chrg_e);
  out[1] = chrg_e * O_CNSTNT;
  double dslv_e = 0.0;
  if (dslv_on && distbb < distdslv) dslv_e = dslv_init * (zone1 ? 1.0 : (1.0 - distbb / distdslv));
  out[2] = dslv_e;
}

void bude_oracle_transform_f64(double t0, double t1, double t2, double t3, double t4, double t5, double out[12]) {
  const double sx = sin(t0), cx = cos(t0);
  const double sy = sin(t1), cy = cos(t1);
  const double sz = sin(t2), cz = cos(t2);
  out[0] = cy * cz;
  out[1] = sx * sy * cz - cx * sz;
  out[2] = cx * sy * cz + sx * sz;
  out[3] = t3;
  out[4] = cy * sz;
  out[5] = sx * sy * sz + cx * cz;
  out[6] = cx * sy * sz - sx * cz;
  out[7] = t4;
  out[8] = -sy;
  out[9] = sx * cy;
  out[10] = cx * cy;
  out[11] = t5;
}

static double pose_f64(size_t natlig, size_t natpro, const OAtom *protein, const OAtom *ligand, const OFF *ff,
                       const double t[12], double *magnitude) {
  double etot = 0.0, mag = 0.0;
  for (size_t il = 0; il < natlig; il++) {
    const OAtom la = ligand[il];
    const double lx = t[3] + la.x * t[0] + la.y * t[1] + la.z * t[2];
    const double ly = t[7] + la.x * t[4] + la.y * t[5] + la.z * t[6];
    const double lz = t[11] + la.x * t[8] + la.y * t[9] + la.z * t[10];
    for (size_t ip = 0; ip < natpro; ip++) {
      const OAtom pa = protein[ip];
      const double x = lx - pa.x, y = ly - pa.y, z = lz - pa.z;
      double terms[3];
      bude_oracle_pair_terms_f64(&ff[pa.type], &ff[la.type], sqrt(x * x + y * y + z * z), terms);
      etot += terms[0] + terms[1] + terms[2];
      mag += fabs(terms[0]) + fabs(terms[1]) + fabs(terms[2]);
    }
  }
  if (magnitude) *magnitude = mag * 0.5;
  return etot * 0.5;
}

void bude_oracle_fasten_f64(size_t natlig, size_t natpro, const void *protein, const void *ligand, const void *ff,
                            const float *p0, const float *p1, const float *p2, const float *p3, const float *p4,
                            const float *p5, size_t first, size_t count, double *energies) {
  const OAtom *pro = (const OAtom *)protein;
  const OAtom *lig = (const OAtom *)ligand;
  const OFF *f = (const OFF *)ff;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long k = 0; k < (long long)count; k++) {
    const size_t i = first + (size_t)k;
    double t[12];
    bude_oracle_transform_f64(p0[i], p1[i], p2[i], p3[i], p4[i], p5[i], t);
    energies[k] = pose_f64(natlig, natpro, pro, lig, f, t, NULL);
  }
}

/* FP64 energies plus the pose's magnitude 0.5 * sum |steric| + |charge| +
 * |dslv| over all pairs: the scale of the cancellation in the sum, which
 * bounds what any FP32 evaluation can resolve (|error| ~ eps * magnitude). */
void bude_oracle_fasten_f64_mag(size_t natlig, size_t natpro, const void *protein, const void *ligand, const void *ff,
                                const float *p0, const float *p1, const float *p2, const float *p3, const float *p4,
                                const float *p5, size_t first, size_t count, double *energies, double *magnitude) {
  const OAtom *pro = (const OAtom *)protein;
  const OAtom *lig = (const OAtom *)ligand;
  const OFF *f = (const OFF *)ff;
#pragma omp parallel for schedule(dynamic, 16)
  for (long long k = 0; k < (long long)count; k++) {
    const size_t i = first + (size_t)k;
    double t[12];
    bude_oracle_transform_f64(p0[i], p1[i], p2[i], p3[i], p4[i], p5[i], t);
    energies[k] = pose_f64(natlig, natpro, pro, lig, f, t, &magnitude[k]);
  }
}

/* Validation metric of the reference driver, src/main.cpp:329-342:
 * difference(a, b) = |a - b| / a with a = reference (signed), entries where
 * both |ref| < 1 and |sample| < 1 are skipped, an entry fails when
 * 100*diff > tol_pct.  Returns max_diff_% and writes the failure count. */
double bude_oracle_validate(const float *ref, const float *sample, size_t n, double tol_pct, size_t *failed) {
  double maxdiff = 0.0;
  size_t nf = 0;
  for (size_t i = 0; i < n; i++) {
    if (fabsf(ref[i]) < 1.f && fabsf(sample[i]) < 1.f) continue;
    const double diff = fabs((double)ref[i] - sample[i]) / ref[i];
    if (diff > maxdiff) maxdiff = diff;
    if (100.0 * diff > tol_pct) nf++;
  }
  if (failed) *failed = nf;
  return 100.0 * maxdiff;
}
