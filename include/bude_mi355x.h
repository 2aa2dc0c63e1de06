/*
 * bude_mi355x.h -- C ABI of libbude_mi355x, an MI355X (gfx950) native
 * implementation of miniBUDE's fasten_main pose-energy evaluation.
 *
 * Plain pointers and sizes only.  The caller owns every host array; the
 * library owns every device buffer.  Every entry point returns a status
 * (BUDE_OK == 0) and leaves a message for bude_mi355x_last_error() on the
 * calling thread when it fails.  Calls on one context must be serialised by
 * the caller; distinct contexts may be used from distinct threads.
 *
 * Each declaration cites the reference interface it replaces
 * (paths relative to the miniBUDE v2.0 source tree).
 */
#ifndef BUDE_MI355X_H
#define BUDE_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BUDE_MI355X_ABI_VERSION 1

/* Deck records: byte-identical to the reference's packed 16-byte structs
 * (src/bude.h:39-42 Atom, src/bude.h:44-49 FFParams); the deck files
 * protein.in / ligand.in / forcefield.in are raw arrays of them. */
typedef struct __attribute__((__packed__)) BudeAtom {
  float x, y, z;
  int32_t type;
} BudeAtom;

typedef struct __attribute__((__packed__)) BudeFFParams {
  int32_t hbtype;
  float radius;
  float hphb;
  float elsc;
} BudeFFParams;

enum {
  BUDE_OK = 0,
  BUDE_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference, e.g. src/serial/fasten.hpp:147-150 */
  BUDE_ERR_RUNTIME = 2,          /* std::runtime_error from HipBude::checkError, src/hip/fasten.hpp:161-165 */
  BUDE_ERR_NO_DEVICE = 3,
  BUDE_ERR_UNSUPPORTED = 4,
  BUDE_ERR_OUT_OF_MEMORY = 5
};

/* Message of the last failed call on this thread ("" if none). */
const char *bude_mi355x_last_error(void);

/* BUDE_MI355X_ABI_VERSION of the loaded library. */
int bude_mi355x_abi_version(void);

/* ---- devices: Bude<PPWI>::enumerateDevices, src/bude.h:98 and
 *      HipBude::enumerateDevices, src/hip/fasten.hpp:170-183 ---- */
int bude_mi355x_device_count(int *count);
/* "<name> (<MB>MB;<gcnArchName>;<CUs>CU)" into buf (NUL-terminated, truncated to len). */
int bude_mi355x_device_name(int device, char *buf, size_t len);

/* ---- tunables: the driver's PPWI list (src/main.cpp:48-53, :221-231) ---- */
/* Writes up to cap supported PPWI values (ascending) and their count. */
int bude_mi355x_ppwi_list(size_t *out, size_t cap, size_t *count);
/* Bude<PPWI>::compatible, src/bude.h:99: 1 when (ppwi, wgsize) can run on
 * device, 0 otherwise (the reason goes to last_error). */
int bude_mi355x_compatible(int device, size_t ppwi, size_t wgsize, int *ok);

/* ---- plain fasten_main: HipBude::fasten, src/hip/fasten.hpp:212-267 ---- */
typedef struct bude_ctx bude_ctx;

/* Allocates device buffers on `device` and copies the deck in
 * (src/hip/fasten.hpp:214-232, the reference's contextTime).  poses[k]
 * points at row k of the deck's SoA pose table, nposes floats each
 * (src/main.cpp:294-311): rows 0-2 are the tilt/roll/pan angles in radians,
 * rows 3-5 the x/y/z translations.  Atom types must index forcefield. */
int bude_mi355x_create(int device, const BudeAtom *protein, size_t natpro, const BudeAtom *ligand, size_t natlig,
                       const BudeFFParams *forcefield, size_t ntypes, const float *const poses[6], size_t nposes,
                       bude_ctx **out);

/* Runs warmup + iterations launches of the kernel with `ppwi` poses per
 * work-item and `wgsize` work-items per work-group over all poses
 * (src/hip/fasten.hpp:234-248), then copies the energies out
 * (src/hip/fasten.hpp:250-251; energies_out may be NULL to skip the copy).
 * Per launch, in order (warmup first), writes:
 *   host_ms[i]   host wall time of launch + hipStreamSynchronize, the
 *                reference's Sample::kernelTimes (may be NULL)
 *   device_ms[i] hipEvent time of the launch on the library's stream
 *                (may be NULL)
 * Both arrays hold warmup + iterations entries. */
int bude_mi355x_fasten(bude_ctx *ctx, size_t ppwi, size_t wgsize, size_t warmup, size_t iterations,
                       float *energies_out, double *host_ms, double *device_ms);

/* Milliseconds the create call spent on allocation + host-to-device copies. */
int bude_mi355x_context_ms(const bude_ctx *ctx, double *ms);

/* Frees every device buffer of the context (src/hip/fasten.hpp:253-264). */
int bude_mi355x_destroy(bude_ctx *ctx);

/* ---- batched ragged fasten_main: one protein, many ligands with their own
 *      atom and pose counts, in ONE persistent launch.  No counterpart in the
 *      reference, whose driver runs one (ligand, pose set) per
 *      Bude<PPWI>::fasten call (src/main.cpp:517-541) and skips pose counts
 *      that do not divide ppwi*wgsize (src/main.cpp:525-530); every pose of
 *      every ligand gets the energy fasten_main defines
 *      (src/serial/fasten.hpp:18-132). ---- */
typedef struct bude_batch bude_batch;

typedef struct BudeBatchConfig {
  int32_t ppwi;           /* poses per work-item; a tile is 64*ppwi poses of one ligand: 1, 2, 4, 8 */
  int32_t wgsize;         /* work-items per work-group: 64, 128, 256, 512 */
  int32_t waves_per_simd; /* persistent grid depth: CUs * 4 * waves_per_simd waves, 1..16 */
  int32_t order;          /* 0: tiles of the heaviest ligands first, 1: ligand order */
  int32_t use_graph;      /* 1: replay a captured hipGraph (counter reset + kernel) */
} BudeBatchConfig;

/* ligands[i] has natlig[i] atoms; poses[i] holds 6 rows of nposes[i] floats
 * (row k at poses[i] + k*nposes[i], rows as in the deck's poses.in). */
int bude_mi355x_batch_create(int device, const BudeAtom *protein, size_t natpro, const BudeFFParams *forcefield,
                             size_t ntypes, size_t nligands, const BudeAtom *const *ligands, const size_t *natlig,
                             const float *const *poses, const size_t *nposes, bude_batch **out);
/* warmup + iterations persistent launches; device_ms[i] = hipEvent time of
 * launch i (counter reset included), may be NULL. */
int bude_mi355x_batch_run(bude_batch *b, const BudeBatchConfig *cfg, size_t warmup, size_t iterations,
                          double *device_ms);
/* The baseline the batch replaces: one plain launch per ligand, in ligand
 * order on one stream; device_ms[i] covers all nligands launches of pass i. */
int bude_mi355x_batch_run_separate(bude_batch *b, size_t ppwi, size_t wgsize, size_t warmup, size_t iterations,
                                   double *device_ms);
/* Times candidate configurations (tile width, work-group size, grid depth,
 * tile order; graph replay) on this batch and returns the fastest. */
int bude_mi355x_batch_autotune(bude_batch *b, BudeBatchConfig *best, double *best_ms);
/* Energies of ligand `ligand` (nposes[ligand] floats) from the last run. */
int bude_mi355x_batch_energies(bude_batch *b, size_t ligand, float *out);
/* Sets every energy to NaN (so a later run can be checked to write all). */
int bude_mi355x_batch_clear(bude_batch *b);
int bude_mi355x_batch_destroy(bude_batch *b);

/* ---- deck tooling: the synthetic generator behind the benchmark decks
 *      (the role makedeck/main.cpp:49-234 plays for bm1/bm2).  Every value
 *      comes from std::mt19937 words through fixed integer/IEEE mappings, so
 *      a seed gives byte-identical decks on every host. ---- */
/* ntypes forcefield rows: hbtype uniform over {45 '-', 69 'E', 70 'F'},
 * radius U[1.0,2.2), hphb U[-2,2), elsc U[-1,1). */
int bude_synth_forcefield(uint64_t seed, size_t ntypes, BudeFFParams *out);
/* natpro atoms uniform in the cube [-half_width, half_width)^3, type U{0..ntypes-1}. */
int bude_synth_protein(uint64_t seed, size_t natpro, size_t ntypes, float half_width, BudeAtom *out);
/* natlig atoms uniform in the ball of `radius` at the origin, type U{0..ntypes-1}. */
int bude_synth_ligand(uint64_t seed, size_t natlig, size_t ntypes, float radius, BudeAtom *out);
/* Poses [first, first+count) of the seed's pose stream into six rows of
 * `stride` floats (rows6 + k*stride): angles U[-pi,pi), translations
 * U[-max_trans, max_trans).  Pose i is the same for every first/count. */
int bude_synth_poses(uint64_t seed, size_t first, size_t count, float max_trans, float *rows6, size_t stride);

#ifdef __cplusplus
}
#endif

#endif /* BUDE_MI355X_H */
