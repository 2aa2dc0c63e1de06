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

/* ---- plain fasten_main without allocation.  No counterpart in the
 *      reference (whose model allocates and frees in every fasten call); it
 *      is for callers that cannot allocate device memory: code under stream
 *      capture, tight serving loops, harnesses that forbid hipMalloc. ----
 * One pass of the plain kernel over caller device memory, with the
 * energies bude_mi355x_create + bude_mi355x_fasten give for the same deck,
 * bit for bit.  Every pointer is a device pointer: protein[natpro],
 * ligand[natlig] and forcefield[ntypes] deck records, `poses` six rows of
 * pose_stride floats (row k at poses + k * pose_stride, the first nposes of
 * each used), energies[nposes] written.  The deck's derived tables (atoms
 * grouped by type, pair coefficients) go to a fixed array of
 * BUDE_MI355X_NOALLOC_SCRATCH_BYTES per device, held until the process ends;
 * a deck that needs more (about 16 B per protein atom + 32 B per ligand atom
 * and forcefield type) or more than BUDE_MI355X_NOALLOC_MAX_TYPES types is
 * rejected with BUDE_ERR_INVALID_ARGUMENT before any device work.  Nothing is
 * allocated, on the device or by the library on the host.  Work runs on
 * `stream` (a hipStream_t; NULL: the null stream) and the call returns when
 * the energies are written.  An atom type outside the forcefield is reported
 * as BUDE_ERR_INVALID_ARGUMENT and leaves the energies untouched.  Calls on
 * one device share its array: calls that overlap in time on one device race
 * and must be serialised by the caller. */
#define BUDE_MI355X_NOALLOC_SCRATCH_BYTES (4u << 20)
#define BUDE_MI355X_NOALLOC_MAX_TYPES 256
int bude_mi355x_fasten_noalloc(int device, const BudeAtom *protein, size_t natpro, const BudeAtom *ligand,
                               size_t natlig, const BudeFFParams *forcefield, size_t ntypes, const float *poses,
                               size_t pose_stride, size_t nposes, size_t ppwi, size_t wgsize, float *energies,
                               void *stream);

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

/* ---- learned prescreen: an MLP pose-energy surrogate (12 or 21 pose
 *      features: the 6 pose parameters, sin and cos of the 3 angles, and
 *      optionally the 9 entries of the rotation matrix they determine ->
 *      256 -> 256 -> 256 -> 1, GELU,
 *      BF16 weights, FP32 accumulation on MFMA),
 *      trained on the GPU by this library's own forward/backward kernels and
 *      Adam on energies the plain kernel computes; screening runs the
 *      surrogate on every pose and the plain kernel on the lowest
 *      keep_fraction (exhaustive plain scoring when the validation RMSE is
 *      above max_val_rmse).  No counterpart in the reference; every energy it
 *      returns is fasten_main's (src/serial/fasten.hpp:18-132). ---- */
typedef struct bude_screen bude_screen;

typedef struct BudeScreenConfig {
  uint32_t batch;       /* training mini-batch, a multiple of 32 */
  uint32_t epochs;      /* passes over the training split */
  float lr;             /* peak Adam learning rate (warm-up, cosine decay) */
  float clip_quantile;  /* training energies above this quantile are clipped to it */
  float keep_fraction;  /* fraction of screened poses rescored exactly (lowest predictions) */
  float max_val_rmse;   /* validation RMSE bound (standardised units) above which screening falls back */
  float tail_quantile;  /* training energies below this quantile get loss weight tail_weight (0: none) */
  float tail_weight;
  uint64_t seed;        /* weight initialisation and shuffling */
  uint32_t features;    /* pose inputs: 12 (the 6 pose parameters and the sines and cosines of the angles) or 21
                           (those and the 9 entries of the rotation matrix) */
} BudeScreenConfig;

typedef struct BudeScreenTrainStats {
  double label_ms;   /* plain-kernel labelling of the training poses */
  double train_ms;   /* all training steps */
  float final_loss;  /* mean squared error of the last epoch (standardised units) */
  float val_rmse;    /* RMSE on the held-out 1/16 of the poses (standardised units) */
  float clip, mu, sigma;  /* target transform: y = (min(E, clip) - mu) / sigma */
  uint32_t steps;
  int32_t fallback;  /* 1 when val_rmse > max_val_rmse */
} BudeScreenTrainStats;

typedef struct BudeScreenResult {
  double total_ms;    /* host wall time of the whole screen */
  float score_ms;     /* surrogate (or exhaustive plain kernel in fallback) */
  float select_ms;    /* top-k selection */
  float rescore_ms;   /* gather + exact rescoring (the plain kernel's energies, bit for bit) */
  uint64_t nposes;
  uint64_t nselected;
  int32_t fallback;
} BudeScreenResult;

int bude_mi355x_screen_create(int device, const BudeAtom *protein, size_t natpro, const BudeAtom *ligand,
                              size_t natlig, const BudeFFParams *forcefield, size_t ntypes,
                              const BudeScreenConfig *cfg, bude_screen **out);
/* Labels n poses (6 rows of n floats) with the plain kernel and trains on
 * the first n - max(32, n/16), validating on the rest; loss_curve[e] = mean
 * squared error of epoch e (up to loss_cap entries, may be NULL). */
int bude_mi355x_screen_train(bude_screen *s, const float *const poses[6], size_t n, BudeScreenTrainStats *stats,
                             float *loss_curve, size_t loss_cap);
/* Makes n poses resident for screening. */
int bude_mi355x_screen_load(bude_screen *s, const float *const poses[6], size_t n);
/* Screens the loaded poses (see above). */
int bude_mi355x_screen_run(bude_screen *s, BudeScreenResult *res);
/* Changes the fraction the next screen_run rescores exactly (no retraining). */
int bude_mi355x_screen_set_keep(bude_screen *s, float keep_fraction);
/* The last screen's selected pose indices (ascending) and their exact energies. */
int bude_mi355x_screen_selected(bude_screen *s, uint32_t *idx, float *energies, size_t cap, size_t *n);
/* Surrogate prediction for every loaded pose, in energy units of the clipped scale. */
int bude_mi355x_screen_predict(bude_screen *s, float *pred);
/* Plain kernel on every loaded pose: the exhaustive baseline, and the exact energies recall is measured on. */
int bude_mi355x_screen_exhaustive(bude_screen *s, float *energies, double *ms);
int bude_mi355x_screen_destroy(bude_screen *s);

/* ---- combined screen: a fixed set of ligands, each with its own pose set,
 *      screened against one protein after another.  Per protein: exact
 *      energies of labels_per_pair sampled poses of every (protein, ligand)
 *      pair, a surrogate validated per pair on them, the surrogate over
 *      every pose of every pair, and exact rescoring of each pair's lowest
 *      keep_fraction in one ragged batched launch; a pair that fails
 *      validation is scored exhaustively instead.  Validation (every
 *      surrogate): on the pair's samples, the gap between the exact energies
 *      at the keep_fraction and gate_top_fraction quantiles, over the
 *      surrogate's error deviation on the lowest quarter of the samples (the
 *      residual of a least-squares line of score on exact energy there, in
 *      energy units), must reach gate_z -- with Gaussian errors and exact
 *      energies spread evenly over the gap the selection then loses a
 *      fraction G(z) / z, G(z) = phi(z) - z Phi(-z), of the exact top
 *      gate_top_fraction (0.0043 at z = 2) -- and the calibrated validation
 *      RMSE (below) must stay within max_val_rmse.  A pair short of gate_z
 *      at keep_fraction is widened instead of scored exhaustively when a
 *      larger fraction, at most widen_max, reaches it: the fraction of its
 *      samples whose exact energy is within gate_z deviations above the
 *      gate_top_fraction quantile (so its gap does), which its selection
 *      then keeps.
 *      Surrogates: BUDE_SURROGATE_MLP, the learned prescreen's network with
 *      11 ligand descriptor inputs, one per protein, trained on 15/16 of the
 *      samples; BUDE_SURROGATE_FIELD, single-ligand-atom energy maps of the
 *      protein on a grid (one per ligand atom type, computed by the pair loop
 *      of fasten_main) summed over the ligand's atoms by trilinear
 *      interpolation; BUDE_SURROGATE_ATOM_MLP, the learned per-atom field:
 *      one network per protein (a learned grid of BUDE_ATOMNET_FEATURES
 *      features at grid_spacing, trilinear, then a ReLU MLP of two
 *      BUDE_ATOMNET_HIDDEN-wide layers with the atom's forcefield type as a
 *      one-hot input) giving one ligand atom's energy at a position, summed
 *      over the ligand's atoms; trained by BF16 MFMA kernels and Adam for
 *      `steps` steps of `batch` points on train_points exact one-atom
 *      energies (the plain kernel's pair loop) at positions the screened
 *      poses put ligand atoms at.  No counterpart in the reference; every
 *      energy returned is fasten_main's (src/serial/fasten.hpp:18-132). ---- */
typedef struct bude_combined bude_combined;

enum { BUDE_SURROGATE_MLP = 0, BUDE_SURROGATE_FIELD = 1, BUDE_SURROGATE_ATOM_MLP = 2 };
#define BUDE_ATOMNET_FEATURES 8
#define BUDE_ATOMNET_HIDDEN 32

typedef struct BudeCombinedConfig {
  uint32_t surrogate;        /* BUDE_SURROGATE_MLP, BUDE_SURROGATE_FIELD or BUDE_SURROGATE_ATOM_MLP */
  uint32_t labels_per_pair;  /* sampled poses per pair scored exactly (2 .. 2^24) */
  float keep_fraction;       /* per pair: the lowest fraction of surrogate scores is rescored exactly */
  float max_val_rmse;        /* bound on a pair's calibrated validation RMSE (standardised units; +inf: none) */
  float clip_quantile;       /* per-pair target transform, as the learned prescreen */
  float tail_quantile;
  float tail_weight;
  uint32_t batch;            /* MLP: training mini-batch, a multiple of 32; ATOM_MLP: points, a multiple of 64 */
  uint32_t epochs;           /* MLP: passes over the training samples */
  float lr;                  /* MLP, ATOM_MLP: peak Adam learning rate of the weights */
  float grid_spacing;        /* FIELD, ATOM_MLP: grid spacing in Angstrom (0.05 .. 4) */
  uint32_t pipeline_chunks;  /* FIELD, ATOM_MLP: ligand chunks scored, selected and rescored as a pipeline
                                (rescoring of chunk k beside the scoring of chunk k + 1); 0 = automatic, <= 16 */
  uint64_t seed;             /* sampling, initialisation, shuffling */
  uint32_t train_points;     /* ATOM_MLP: training points per protein (a multiple of batch) */
  uint32_t steps;            /* ATOM_MLP: training steps */
  float grid_lr;             /* ATOM_MLP: peak Adam learning rate of the grid features */
  float gate_z;              /* validation: required (quantile gap) / (tail error deviation), see above */
  float gate_top_fraction;   /* validation: the exact top fraction the selection must keep (e.g. 0.01) */
  float widen_max;           /* validation: largest fraction a pair short of gate_z is widened to instead of
                                falling back (0: never widen; else keep_fraction .. 1) */
} BudeCombinedConfig;

typedef struct BudeCombinedResult {
  double total_ms;           /* host wall time of the whole protein */
  float prepare_ms;          /* protein grouping, pair tables, uploads (host) */
  float label_ms;            /* exact scoring of the sampled poses */
  float train_ms;            /* MLP training, or building the maps */
  float score_ms;            /* surrogate over every pose + exhaustive scoring of fallback pairs */
  float select_ms;           /* per-pair top-k and the rescoring tables */
  float rescore_ms;          /* exact rescoring of the selections, first launch to last; in a field pipeline of
                                several chunks it overlaps the scoring of the later chunks */
  uint64_t nposes;           /* poses screened (all pairs) */
  uint64_t nselected;        /* poses with exact energies returned */
  uint64_t labels;           /* sampled poses scored exactly */
  uint32_t fallback_pairs;
  float surrogate_rmse;      /* ATOM_MLP: one-atom energy RMSE on held-out points (energy units) */
  uint32_t widened_pairs;    /* pairs rescored at a fraction above keep_fraction (see widen_max) */
} BudeCombinedResult;

typedef struct BudePairResult {
  uint32_t nposes, nselected;
  int32_t fallback;          /* 1: scored exhaustively */
  float val_rmse;            /* surrogate vs exact on the pair's validation samples (clipped, standardised) */
  float val_rmse_calibrated; /* the same after the least-squares increasing affine map of the scores onto the
                                targets (the part of the error that changes the ranking); a fallback criterion */
  float gate_z;              /* validation gap / tail error deviation (energy units); fallback below cfg gate_z */
  float sample_recall;       /* of the samples' exact lowest 5 %, the fraction among the surrogate's lowest 10 % */
  float keep_fraction;       /* the fraction of the pair's surrogate scores rescored (keep_fraction, or more
                                for a widened pair; 1 for a fallback pair) */
} BudePairResult;

/* ligands[i] has natlig[i] atoms and poses[i] 6 rows of nposes[i] floats (as
 * bude_mi355x_batch_create); at most 2^32 - 1 poses in all.  The poses stay
 * resident in device memory for every protein screened. */
int bude_mi355x_combined_create(int device, const BudeFFParams *forcefield, size_t ntypes, size_t nligands,
                                const BudeAtom *const *ligands, const size_t *natlig, const float *const *poses,
                                const size_t *nposes, const BudeCombinedConfig *cfg, bude_combined **out);
/* Screens every ligand's poses against this protein. */
int bude_mi355x_combined_screen(bude_combined *c, const BudeAtom *protein, size_t natpro, BudeCombinedResult *res);
/* Screens proteins[0..nproteins) in order, protein i's sampled poses, maps
 * and field scores running beside protein i - 1's exact rescoring (field
 * screens without fallback pairs; otherwise one protein after another).
 * res[i] is protein i's result; done(user, i), when not NULL, runs as soon
 * as protein i is complete, with protein i as the protein screened last for
 * the accessors below (_pair, _selected, _labels, _exhaustive, _exact,
 * _recall), except _scores and _map, which then fail for every protein but
 * the batch's last (its successor's scoring has begun).  The selections of
 * every protein equal those of screening them one by one.  *wall_ms (may be
 * NULL) is the host time of the whole call.  No other call on the object
 * may run meanwhile, except the accessors from done. */
typedef void (*bude_combined_done_fn)(void *user, size_t index);
int bude_mi355x_combined_screen_batch(bude_combined *c, size_t nproteins, const BudeAtom *const *proteins,
                                      const size_t *natpro, BudeCombinedResult *res, bude_combined_done_fn done,
                                      void *user, double *wall_ms);
/* Changes the per-pair fraction later screens rescore exactly (not below
 * the configured gate_top_fraction). */
int bude_mi355x_combined_set_keep(bude_combined *c, float keep_fraction);
int bude_mi355x_combined_pair(bude_combined *c, size_t ligand, BudePairResult *r);
/* The last screen's selection of one pair: pose indices (ascending, within
 * the ligand's pose set) and their exact energies. */
int bude_mi355x_combined_selected(bude_combined *c, size_t ligand, uint32_t *idx, float *energies, size_t cap,
                                  size_t *n);
/* The last screen's sampled poses of one pair (what the MLP trained on, or
 * the field was validated on): pose indices within the ligand's pose set,
 * in sampling order, and their exact energies, which come from a split
 * kernel (four lanes per pose) and match the tiled kernel's to rounding.
 * idx and energies may both be NULL to query *n. */
int bude_mi355x_combined_labels(const bude_combined *c, size_t ligand, uint32_t *idx, float *energies, size_t cap,
                                size_t *n);
/* Surrogate score of every pose of one pair (exact energies for a fallback
 * pair): standardised units for the MLP, energy units for the field. */
int bude_mi355x_combined_scores(bude_combined *c, size_t ligand, float *out);
/* Exact energies of every pose of every pair against the last protein (one
 * batched launch, for measuring recall); _exact / _recall read them while
 * the protein screened last is the same one (e.g. after set_keep). */
int bude_mi355x_combined_exhaustive(bude_combined *c, double *device_ms);
int bude_mi355x_combined_exact(bude_combined *c, size_t ligand, float *out);
/* recall[l] = fraction of pair l's exact lowest round(top_fraction * n)
 * poses that the last screen selected. */
int bude_mi355x_combined_recall(bude_combined *c, float top_fraction, float *recall);
/* FIELD: the map grid (point (i,j,k) at lo + h*(i,j,k)), and the map of one
 * forcefield type, points in (i, j, k) row-major order. */
int bude_mi355x_combined_grid(bude_combined *c, float lo[3], float *h, int32_t n[3]);
int bude_mi355x_combined_map(bude_combined *c, int32_t type, float *out);
int bude_mi355x_combined_destroy(bude_combined *c);
/* One forward + backward pass of the surrogate on `batch` rows (a multiple
 * of 32) of 21 features with caller weights W_l [in][out] (21x256, 256x256,
 * 256x256, 256x1) and biases: predictions, mean squared error against
 * target, and the FP32 gradients of that loss.  Exposes the training
 * kernels to gradient checks. */
int bude_mi355x_mlp_forward_backward(int device, size_t batch, const float *features, const float *target,
                                     const float *const weights[4], const float *const biases[4], float *pred,
                                     float *loss, float *const grad_w[4], float *const grad_b[4]);
/* The fused inference kernel the screen uses, on n poses (6 rows of n
 * floats) with caller weights as above: predictions per pose. */
int bude_mi355x_mlp_predict(int device, const float *const poses[6], size_t n, const float *const weights[4],
                            const float *const biases[4], float *pred);

/* Exact energy of one ligand atom of forcefield type types[i] at
 * points[3i .. 3i+2] against the protein: fasten_main's sum for a one-atom
 * ligand there (src/serial/fasten.hpp:56-131), computed by the plain
 * kernel's pair loop (what the learned per-atom field trains on). */
int bude_mi355x_atom_energies(int device, const BudeAtom *protein, size_t natpro, const BudeFFParams *forcefield,
                              size_t ntypes, const float *points, const int32_t *types, size_t n, float *out);
/* One forward + backward pass of the learned per-atom field's training
 * kernels on `batch` points (a multiple of 64; slots[i] < nslots) against
 * standardised targets, with caller parameters: the feature grid
 * [n0 * n1 * n2][BUDE_ATOMNET_FEATURES] (node (i, j, k) at grid_lo + h (i, j,
 * k), row-major), W0 [32 * ceil((BUDE_ATOMNET_FEATURES + nslots) / 32)]
 * [BUDE_ATOMNET_HIDDEN] (feature rows, then one row per slot, then zeros),
 * W1 [BUDE_ATOMNET_HIDDEN][BUDE_ATOMNET_HIDDEN], W2 [BUDE_ATOMNET_HIDDEN][1]
 * and the biases; predictions, mean squared error and every FP32 gradient.
 * Exposes the training kernels to gradient checks. */
int bude_mi355x_atomnet_forward_backward(int device, size_t batch, const float *points, const int32_t *slots,
                                         const float *target, const float grid_lo[3], float grid_h,
                                         const int32_t grid_n[3], const float *grid, size_t nslots,
                                         const float *const weights[3], const float *const biases[3], float *pred,
                                         float *loss, float *grad_grid, float *const grad_w[3],
                                         float *const grad_b[3]);
/* The fused scorer the combined screen uses, on the poses of one ligand
 * (atom type fields hold slots < nslots) with caller parameters as above:
 * scores[j] = sum over the ligand's atoms of sigma * y + mu. */
int bude_mi355x_atomnet_score(int device, const BudeAtom *ligand, size_t natlig, const float *const poses[6],
                              size_t nposes, const float grid_lo[3], float grid_h, const int32_t grid_n[3],
                              const float *grid, size_t nslots, const float *const weights[3],
                              const float *const biases[3], float sigma, float mu, float *scores);

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

/* ---- deck tooling: decks from real molecules (the makedeck program,
 *      makedeck/main.cpp:49-234).  Host only: no device is touched.  A deck
 *      built from the same bhff/mol2 files, pose seed and pose count is
 *      byte-identical to makedeck's (the bm2 deck is reproduced exactly). ---- */
typedef struct bude_forcefield bude_forcefield;
/* Reads a BUDE bhff forcefield (bude::readForceField,
 * makedeck/input-utils.cpp:72-119): every 2-column "RESIDUE COUNT" line is
 * followed by COUNT 9-column atom rows "type hbchar radius hphb _ _ _ scaling
 * elsc"; an entry's parameters are {hbchar code, radius*scaling, hphb,
 * elsc}.  Lines whose first non-blank character is '#' or '%' are comments.
 * Malformed rows fail with BUDE_ERR_INVALID_ARGUMENT and "path:line". */
int bude_forcefield_read(const char *bhff_path, bude_forcefield **out);
/* Residue groups and total entries. */
int bude_forcefield_info(const bude_forcefield *ff, size_t *nresidues, size_t *nentries);
/* The deck's forcefield rows: all entries, residue groups in residue-name
 * order, entries in file order (makedeck/main.cpp:146-149).  out may be NULL
 * to query *n; otherwise cap must be >= the entry count. */
int bude_forcefield_params(const bude_forcefield *ff, BudeFFParams *out, size_t cap, size_t *n);
int bude_forcefield_destroy(bude_forcefield *ff);
/* Heavy atoms of a Tripos mol2 file against a forcefield (bude::readMol2,
 * makedeck/input-utils.cpp:122-202): the rows between "@<TRIPOS>ATOM" and
 * the next "@<TRIPOS>" line, hydrogens (type "H"/"h") skipped.  With a
 * one-residue forcefield an atom is looked up by its Sybyl type (column 6;
 * rows must have 9 columns), otherwise by residue (first 3 characters of
 * column 8) and atom name (column 2; rows must have 10 columns).  An atom's
 * type is the entry's index within its residue group, as makedeck writes
 * it.  "@<BUDE>CONF n" blocks after the atoms are counted in
 * *nconformations (may be NULL).  atoms may be NULL to query *natoms;
 * otherwise cap must be >= the heavy-atom count. */
int bude_mol2_read(const bude_forcefield *ff, const char *mol2_path, BudeAtom *atoms, size_t cap, size_t *natoms,
                   size_t *nconformations);
/* nposes distinct poses drawn from the grid values[k][0..counts[k]) of the
 * six pose fields (bude::generatePoses, makedeck/input-utils.cpp:205-262):
 * indices drawn per field in the order tilt, roll, pan, x, y, z from
 * std::mt19937(seed), duplicates dropped, poses emitted in the order of
 * makedeck's hash set.  Angle values are degrees and come out in radians.
 * rows[k] receives nposes floats.  Fails with BUDE_ERR_INVALID_ARGUMENT when
 * nposes exceeds the number of grid combinations. */
int bude_poses_generate(size_t nposes, uint64_t seed, const float *const values[6], const size_t counts[6],
                        float *const rows[6]);

#ifdef __cplusplus
}
#endif

#endif /* BUDE_MI355X_H */
