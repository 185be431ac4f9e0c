/*
 * o3d_biot.h -- C ABI of libo3dbiot.so: direct regularised Biot-Savart velocity
 * and velocity-gradient sums on one AMD Instinct MI355X (gfx950).
 *
 * Drop-in for the hot path of Omega3D (Applied-Scientific-Research/Omega3D):
 *   points_affect_points  src/Influence.h:67-551
 *   panels_affect_points  src/Influence.h:557-1099
 * and its external-solver plugin point
 *   external_vel_solver_f_ / external_vel_solver_d_   src/Influence.h:33-48
 *
 * Conventions shared by every entry point (the reference's, Influence.h:296-307,
 * ElementBase.h:170-208, Points.h:252-277):
 *   - arrays are structure-of-arrays, float32, caller-owned;
 *   - results are ACCUMULATED (+=) into the caller's velocity and gradient
 *     arrays, unscaled: the caller applies 1/(4 pi) and the freestream later;
 *   - gradient component k = 3*j + i holds du_i/dx_j (ux vx wx uy vy wy uz vz wz);
 *   - sources and targets may be the same arrays (a collection acting on itself);
 *   - when they are, the self pair j == i is INCLUDED, as in the reference,
 *     unless o3d_opts.self_term = O3D_SELF_EXCLUDE.
 * Every entry point returns an o3d_status; nothing asserts or aborts across the
 * ABI.  o3d_last_error() describes the last failure on the calling thread.
 */
#ifndef O3D_BIOT_H
#define O3D_BIOT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define O3D_ABI_VERSION 1

/* regularised cores, runtime-selected (the reference selects one at compile
 * time with a #define, CoreFunc.h:35-39) */
enum o3d_core {
  O3D_CORE_RM = 1,   /* Rosenhead-Moore           CoreFunc.h:43-83   */
  O3D_CORE_WL = 2,   /* Winckelmans-Leonard       CoreFunc.h:241-289 */
  O3D_CORE_EXP = 3,  /* compact exponential       CoreFunc.h:86-238  */
  O3D_CORE_V2 = 4    /* Vatistas n=2              CoreFunc.h:292-341 */
};

/* results_t values of ResultsType.h:27-33 that this path serves */
enum o3d_results { O3D_VELONLY = 1, O3D_VELANDGRAD = 2 };

enum o3d_precision {
  O3D_PREC_FP32 = 0,         /* FP32 pairs, plain FP32 accumulation per target            */
  O3D_PREC_COMPENSATED = 1,  /* FP32 pairs; per-tile FP32 partials folded with two-sum for
                                the velocity and in FP64 for the gradient tensor           */
  O3D_PREC_FP64 = 2          /* the FP64 oracle kernel: every operation in double          */
};

enum o3d_self { O3D_SELF_INCLUDE = 0, O3D_SELF_EXCLUDE = 1 };

/* How the point->point sums are organised.  Both give the same direct sums over
 * every pair; they differ in rounding only.
 *   DIRECT:   one pair at a time on the vector ALUs, sources in index order (an
 *             EXP call of >= 32,768 targets on one collection, tile and tpt left
 *             at 0, reads them in the targets' Morton order instead).
 *   FACTORED: particles in Morton order; a source tile well separated from a
 *             target block is summed through an exact factorisation
 *             K = Bp (|t'|^2 - 2 t'.s' + |s'|^2 + cores) as a product of the pair
 *             weights with a 29-column per-source moment table on the matrix
 *             cores; other pairs as in DIRECT.  RM and EXP cores, velocity and
 *             gradient, FP32 or compensated precision; tile 64, 128 or 256.
 *   AUTO:     FACTORED where it applies and both counts are >= 65536 with tile
 *             and tpt left at 0, DIRECT otherwise.
 * For panel sources (o3d_panels_affect_points[_dev]) the field selects the rule:
 *   AUTO / DIRECT:      the reference's recursive centroid rule (Kernels.h:1028-1211);
 *   O3D_PANEL_GAUSS7:   no counterpart in the reference -- per (target, panel) pair a
 *             7-point degree-5 rule (Strang & Fix; Dunavant) when the target is at
 *             least twice the panel's longest edge from its centroid, otherwise the
 *             exact integral of the panel's constant vortex and source sheet
 *             densities (ts/area, ss) over the flat triangle, in FP64 (edge
 *             logarithms and the signed solid angle).  The near/far decision is
 *             taken in float with a fixed operation order. */
enum o3d_method { O3D_METHOD_AUTO = 0, O3D_METHOD_DIRECT = 1, O3D_METHOD_FACTORED = 2, O3D_PANEL_GAUSS7 = 3 };

enum o3d_status {
  O3D_OK = 0,
  O3D_EINVAL = 1,          /* bad argument (null pointer, bad enum, size mismatch)  */
  O3D_ENOMEM = 2,          /* device allocation failed                              */
  O3D_EDEVICE = 3,         /* HIP runtime / kernel error                            */
  O3D_EINJECTED = 4,       /* failure forced by o3d_inject_fault (tests)            */
  O3D_EADMIT = 5,          /* request does not fit the serving memory budget        */
  O3D_EBUSY = 6,           /* try again: no free slot                               */
  O3D_ECANCELLED = 7,      /* request cancelled or timed out before it ran          */
  O3D_ENODEV = 8,          /* no usable gfx950 device                               */
  O3D_ENOTFOUND = 9        /* unknown handle / request id                           */
};

/* SoA view of a Points<float> collection (ElementBase.h:43-462, Points.h:50-1054).
 * As a SOURCE:  x, s and r are read (r may not be NULL).
 * As a TARGET:  x is read, r is read when non-NULL (thick-cored particles; NULL =
 *               inert field points, the reference's is_inert() branch), u is
 *               accumulated, ug is accumulated when results == O3D_VELANDGRAD. */
typedef struct o3d_points {
  int64_t n;
  const float* x[3];
  const float* s[3];
  const float* r;
  float* u[3];
  float* ug[9];
} o3d_points;

/* SoA view of a Surfaces<float> of flat triangles (Surfaces.h:50-251).
 * ts is the ABSOLUTE vortex strength (sheet strength x area, Surfaces.h:309-335),
 * ss the sheet SOURCE strength (NULL = none); idx holds 3 node indices per panel. */
typedef struct o3d_panels {
  int64_t np;          /* panels */
  int64_t nn;          /* nodes  */
  const float* x[3];
  const uint32_t* idx;
  const float* ts[3];
  const float* area;
  const float* ss;
} o3d_panels;

typedef struct o3d_opts {
  int32_t core;        /* enum o3d_core                                              */
  int32_t results;     /* enum o3d_results                                           */
  int32_t precision;   /* enum o3d_precision                                         */
  int32_t self_term;   /* enum o3d_self (exclusion needs src and targ to be the same
                          collection: same n, source i == target i)                  */
  int32_t tile;        /* sources per LDS tile (64..1024, multiple of 64); 0 = tuned */
  int32_t tpt;         /* targets per thread (1, 2, 4); 0 = tuned                    */
  int32_t maxlev;      /* panels: recursion depth (reference RECURSIVE_LEVELS = 3);
                          negative selects the default                               */
  int32_t method;      /* enum o3d_method (point->point only)                        */
  void* stream;        /* hipStream_t for the *_dev calls; NULL = null stream        */
} o3d_opts;

typedef struct o3d_stats {
  double seconds;      /* wall time of the call (host calls) or of the kernels      */
  double flops;        /* the reference's own count: N_t*(12 + F*N_s) etc.           */
  double interactions; /* target-source pairs evaluated                              */
} o3d_stats;

/* Fill *opts with the defaults: WL core (the reference's compiled default),
 * vel+grad, compensated, self included, tuned tile/tpt, maxlev 3. */
void o3d_default_opts(o3d_opts* opts);

/* Pick the device (ordinal) for this thread and create the library's context.
 * Optional: every call initialises lazily on the current device. */
int o3d_init(int device);
const char* o3d_last_error(void);
int o3d_abi_version(void);

/* ---- the reference's call shapes ------------------------------------------ */

/* points_affect_points (Influence.h:67-551) on HOST arrays: copies in, runs on the
 * GPU, synchronises, accumulates into targ->u / targ->ug.  stats may be NULL. */
int o3d_points_affect_points(const o3d_points* src, o3d_points* targ,
                             const o3d_opts* opts, o3d_stats* stats);

/* panels_affect_points (Influence.h:557-1099), recursive centroid rule of
 * rkernel_2vs_0p(g) (Kernels.h:1028-1211), on HOST arrays.  Target radii are
 * ignored, as in the reference (point kernels for every target, Influence.h:873). */
int o3d_panels_affect_points(const o3d_panels* src, o3d_points* targ,
                             const o3d_opts* opts, o3d_stats* stats);

typedef struct o3d_panel_profile {
  double total_ms;     /* the full evaluation (far points and near closed forms)       */
  double far_ms;       /* the same launch with the near pairs left out                 */
  double pairs;        /* target-panel pairs                                            */
  double near_pairs;   /* pairs evaluated in closed form                                */
  double far_pairs;    /* pairs evaluated by the 7-point rule                           */
} o3d_panel_profile;

/* The 7-point rule (opts->method = O3D_PANEL_GAUSS7) on DEVICE arrays, timed on
 * opts->stream: the full evaluation accumulates into targ as
 * o3d_panels_affect_points_dev does; a far-only pass (into library scratch) and a
 * count of the near pairs follow.  Velocity and gradient, FP32 or compensated.
 * No counterpart in the reference. */
int o3d_panels_g7_profile_dev(const o3d_panels* src, o3d_points* targ, const o3d_opts* opts,
                              o3d_panel_profile* prof);

/* points_affect_panels (Influence.h:1107-1221): the velocity of each target panel
 * induced by vortex particles, the reference's boundary-element right-hand side
 * input.  The panel kernel is run backwards (rkernel_2vs_0p with the panel carrying
 * alpha_j / area_i and the particle as the evaluation point) and the sum is
 * SUBTRACTED from vel[0..2][panel] (Influence.h:1208-1210).  Velocity only
 * (opts->results must be O3D_VELONLY; the reference asserts otherwise, :1111);
 * particle radii are not used; targ->ts and targ->ss are not read.  HOST arrays. */
int o3d_points_affect_panels(const o3d_points* src, const o3d_panels* targ, float* const vel[3],
                             const o3d_opts* opts, o3d_stats* stats);

/* Same two calls on DEVICE pointers, enqueued on opts->stream, no host sync.
 * stats->seconds is not filled (the call returns before the kernels end). */
int o3d_points_affect_points_dev(const o3d_points* src, o3d_points* targ,
                                 const o3d_opts* opts, o3d_stats* stats);
int o3d_panels_affect_points_dev(const o3d_panels* src, o3d_points* targ,
                                 const o3d_opts* opts, o3d_stats* stats);
int o3d_points_affect_panels_dev(const o3d_points* src, const o3d_panels* targ, float* const vel[3],
                                 const o3d_opts* opts, o3d_stats* stats);

/* No counterpart in the reference: o3d_points_affect_points_dev for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The same direct kernels (method AUTO
 * and DIRECT both run them; FACTORED and FP64 are refused with O3D_EINVAL), with
 * their scratch (the source split's partial sums, the Morton order of EXP
 * targets) in one static 66 MiB device array per GPU, held until the process
 * ends; nothing is allocated and no memory fill is issued, so the call can be
 * captured.  At most O3D_NOALLOC_MAX_TARGETS targets (O3D_EINVAL beyond).  Where
 * the _dev call's source split would need more partial sums than the array holds
 * this call takes fewer slices, so its rounding can differ; otherwise the results
 * are the _dev call's bits.  Calls that overlap on one GPU (other streams or
 * threads) share the array and race: keep them ordered. */
#define O3D_NOALLOC_MAX_TARGETS 4194304LL
int o3d_points_affect_points_noalloc_dev(const o3d_points* src, o3d_points* targ,
                                         const o3d_opts* opts, o3d_stats* stats);

/* panels_affect_panels (Influence.h:1225-1245): the velocity that the panels of src
 * induce at colocation points of the panels of targ -- each panel's centroid
 * pushed 1e-4 along its normal, as Surfaces::represent_as_particles(0.0001, -1.0)
 * places them (Surfaces.h:959-1010) -- by panels_affect_points' recursive rule,
 * velocity only (opts->results must be O3D_VELONLY; the reference asserts
 * otherwise, :1227), FP32 or compensated.  As in the reference, which copies the
 * temporary points' velocities over the target's (:1240-1244), vel[0..2][panel]
 * is OVERWRITTEN, not accumulated.  targ->ts, area and ss are not read.  HOST
 * arrays; the _dev form takes device pointers and runs on opts->stream. */
int o3d_panels_affect_panels(const o3d_panels* src, const o3d_panels* targ, float* const vel[3],
                             const o3d_opts* opts, o3d_stats* stats);
int o3d_panels_affect_panels_dev(const o3d_panels* src, const o3d_panels* targ, float* const vel[3],
                                 const o3d_opts* opts, o3d_stats* stats);

/* ---- BEM right-hand side of one boundary surface (BEMHelper.h:83-117) ---- */
typedef struct o3d_bem_rhs_opts {
  double fs[3];        /* freestream (Surfaces::finalize_vels, Surfaces.h:877-887)          */
  double vel[3];       /* the body's velocity at this time (Body::get_vel)                  */
  double rotvel[3];    /* the body's angular velocity vector (Body::get_rotvel_vec)         */
  float tc[3];         /* the surface's transformed geometric centre (Surfaces tc)          */
  int32_t body;        /* 0: no body, or attached to the ground: no motion term
                          (Surfaces.h:561-562)                                             */
  int32_t nunk;        /* unknowns per panel (RHS.h:71-95): 1 normal, 2 tangential (x1,
                          x2), 3 tangential then normal                                    */
} o3d_bem_rhs_opts;

/* DEVICE arrays, on opts->stream.  The panel-centre velocities pu[3] (np floats
 * each; prior contents discarded) are zeroed, receive points_affect_panels from
 * every particle collection vort[0..nvort) (velocity only, opts->precision FP32 or
 * compensated), are finalized (fs + pu/(4 pi)) and lose the body's motion at the
 * panel centroids (factor -1, BEMHelper.h:104); rhs (nunk * np floats) receives
 * minus their components on each panel's basis (x1, x2, normal; Surfaces.h:766-816,
 * RHS.h:47-98).  Float/double mixing as the reference's.  The augmented rows
 * (total circulation, BEMHelper.h:120-143) are the caller's. */
int o3d_bem_rhs_dev(const o3d_points* vort, int32_t nvort, const o3d_panels* bdry, const o3d_bem_rhs_opts* b,
                    const o3d_opts* opts, float* const pu[3], float* rhs);
/* The last two steps alone, on velocities pu[3] as points_affect_panels left them. */
int o3d_bem_rhs_from_vels_dev(const o3d_panels* bdry, const o3d_bem_rhs_opts* b, float* const pu[3], float* rhs,
                              void* stream);

/* BEM influence matrix block of boundary surface src on boundary surface targ
 * (panels_on_panels_coeff, Coefficients.h:169-444, over rkernel_2vs_2p,
 * Kernels.h:1213-1311; the block Surfaces<S>'s CoefficientVisitor returns to
 * BEMHelper.h:219-226).  Surfaces: nodes x[3], panel indices idx and areas (as
 * Surfaces::compute_bases leaves them); ts/ss are not read.  nunk: unknowns per
 * panel, 2 (vortex sheet along x1 and x2) or 3 (plus a source sheet), the same
 * for both (Surfaces.h:261).  same: the two are one surface (the reference's
 * &src == &targ), whose diagonal 2x2 or 3x3 blocks get the reference's constants
 * (Coefficients.h:399-421); targ must then be src.  maxlev: recursion depth, -1 =
 * the reference's RECURSIVE_LEVELS (3), at most 4.  coeffs: (nunk * targ->np)
 * rows x (nunk * src->np) columns, column-major (column j*nunk + k: source panel
 * j, unknown k; row i*nunk + m: target panel i, component on its x1, x2, normal),
 * scaled by 1/(4 pi).  Every tree decision is the reference's float arithmetic;
 * a recursed pair sums its leaves in FP64.  The device form synchronises the
 * stream once (the count of recursed pairs sizes a list). */
int o3d_bem_matrix(const o3d_panels* src, const o3d_panels* targ, int32_t nunk, int32_t same, int32_t maxlev,
                   float* coeffs);
int o3d_bem_matrix_dev(const o3d_panels* src, const o3d_panels* targ, int32_t nunk, int32_t same, int32_t maxlev,
                       float* coeffs, void* stream);
/* Diagnostics (no counterpart in the reference): the device form with its stage
 * times (root decisions and pair count; one-leaf and self coefficients written
 * and recursed pairs listed; recursed pairs) and the number of recursed pairs. */
typedef struct o3d_bem_matrix_profile {
  double count_ms;
  double far_ms;
  double near_ms;
  double total_ms;
  double near_pairs;
} o3d_bem_matrix_profile;
int o3d_bem_matrix_profile_dev(const o3d_panels* src, const o3d_panels* targ, int32_t nunk, int32_t same,
                               int32_t maxlev, float* coeffs, void* stream, o3d_bem_matrix_profile* prof);

/* Diagnostics of the factored method (no counterpart in the reference): runs one
 * o3d_points_affect_points_dev evaluation with method FACTORED on device arrays,
 * synchronises, and reports the time of its stages (preparation: box, Morton keys,
 * sort, gathers, bounds, classification; far field on the matrix cores; near field
 * and output) and the number of (target, source) pairs taken by the far field.
 * Results are accumulated as by the plain call. */
typedef struct o3d_factored_profile {
  double prep_ms;
  double far_ms;
  double near_ms;
  double total_ms;
  double far_pairs;
  double pairs;
  double far_tile_fraction;  /* far_pairs / pairs */
} o3d_factored_profile;
int o3d_factored_profile_dev(const o3d_points* src, o3d_points* targ, const o3d_opts* opts,
                             o3d_factored_profile* out);

/* Morton order of n points on DEVICE arrays (no counterpart in the reference: the
 * order in which the factored method, and direct EXP calls of >= 32,768 targets,
 * visit particles).  keys[i] receives point i's 30-bit Morton key (10 bits per
 * axis on the points' cubic bounding box), perm[k] the index of the k-th point in
 * key order; the sort is stable, so equal keys keep index order.  Enqueued on
 * `stream` (kernels only); scratch is the stream's library workspace. */
int o3d_morton_order_dev(const float* x, const float* y, const float* z, int64_t n, uint32_t* keys,
                         int32_t* perm, void* stream);

/* Double-precision collections (the reference's Points<double>): always run the
 * FP64 kernel, whatever opts->precision says.  Host and device variants. */
typedef struct o3d_points64 {
  int64_t n;
  const double* x[3];
  const double* s[3];
  const double* r;
  double* u[3];
  double* ug[9];
} o3d_points64;
int o3d_points_affect_points64(const o3d_points64* src, o3d_points64* targ,
                               const o3d_opts* opts, o3d_stats* stats);
int o3d_points_affect_points64_dev(const o3d_points64* src, o3d_points64* targ,
                                   const o3d_opts* opts, o3d_stats* stats);

/* The reference's external-solver plugin ABI (Influence.h:33-48, called at
 * :96-104 with gradients): host arrays, counts by pointer, point targets (no
 * target radius), results accumulated; returns the flop count.  The core is the
 * one set by o3d_set_external_core (default WL, the reference's compiled core).
 * _f_ runs the compensated FP32 kernel, _d_ the FP64 kernel.  On error they
 * return a negative value and o3d_last_error() says why. */
float external_vel_solver_f_(int* ns, const float* sx, const float* sy, const float* sz,
                             const float* ssx, const float* ssy, const float* ssz,
                             const float* sr,
                             int* nt, const float* tx, const float* ty, const float* tz,
                             float* tu, float* tv, float* tw, float* tux,
                             float* tvx, float* twx, float* tuy, float* tvy,
                             float* twy, float* tuz, float* tvz, float* twz);
float external_vel_solver_d_(int* ns, const double* sx, const double* sy, const double* sz,
                             const double* ssx, const double* ssy, const double* ssz,
                             const double* sr,
                             int* nt, const double* tx, const double* ty, const double* tz,
                             double* tu, double* tv, double* tw, double* tux,
                             double* tvx, double* twx, double* tuy, double* tvy,
                             double* twy, double* tuz, double* tvz, double* twz);
int o3d_set_external_core(int core);

/* Per-interaction flop count of the reference for (core, results, target kind,
 * vortex-only or vortex+source): Kernels.h flops_0v_0bg etc. -1 if invalid. */
int o3d_flops_per_interaction(int core, int results, int thick_target, int vortex_source);

/* ---- many independent collections in one persistent launch ------------------
 * (no counterpart in the reference, which evaluates one collection per call).
 * Instance i owns particles [off[i], off[i] + n[i]) of shared SoA DEVICE arrays
 * and acts on itself (sources = targets = its particles), velocity + gradient,
 * thick-cored targets.  A device-side work queue of (instance, 256*tpt-target
 * chunk) items, longest instances first, is drained by a persistent grid through
 * an atomic counter.  Each target's result is the same number whether its
 * instance runs alone or in a batch (same tile, fixed per-target order). */
typedef struct o3d_batch o3d_batch;

/* cores / precs / self_terms: per-instance arrays (NULL = the opts value for all).
 * opts->tile / opts->tpt: 0 = defaults (see o3d_batch_tune).
 * opts->method: AUTO sends every instance of >= 32,768 particles with an RM or EXP
 * core in FP32 or compensated precision to the factored far-field method; FACTORED
 * sends all instances there (an instance it cannot take is an error); DIRECT keeps
 * all in the persistent queue.  Factored instances run on up to 8 streams the
 * batch creates here, each with scratch allocated here; they fork from the
 * caller's stream after the zero fill and the caller's stream waits for them. */
int o3d_batch_create(o3d_batch** b, int32_t ninst, const int64_t* off, const int64_t* n,
                     const int32_t* cores, const int32_t* precs, const int32_t* self_terms,
                     const o3d_opts* opts);
int o3d_batch_destroy(o3d_batch* b);
/* Zero the instances' outputs and evaluate every instance, on `stream`.
 * all: the shared device arrays (x, s, r, u, ug with all->n = total particles). */
int o3d_batch_run(o3d_batch* b, const o3d_points* all, void* stream);
/* The same work as one ordinary call per instance, one after another on `stream`
 * (the comparison baseline). */
int o3d_batch_run_separate(o3d_batch* b, const o3d_points* all, void* stream);
/* Capture all of o3d_batch_run into one HIP graph (zero-fill kernel, queue reset
 * kernel, persistent kernel, and the factored instances' pipelines on their lanes,
 * forked and joined by events) and replay it; the graph reads the queue from
 * device memory and is bound to the arrays given at capture. */
int o3d_batch_capture(o3d_batch* b, const o3d_points* all, void* stream);
int o3d_batch_replay(o3d_batch* b, void* stream);
/* Autotuner: time every (tile, tpt) candidate `reps` times on the real batch's
 * queue part and keep the fastest in b; reports it and its mean time.  NULL lists =
 * the default candidates tile {128, 256, 512} x tpt {1, 2, 4}.  The outputs hold
 * the full evaluation afterwards. */
int o3d_batch_tune(o3d_batch* b, const o3d_points* all, void* stream,
                   const int32_t* tiles, int32_t ntiles, const int32_t* tpts, int32_t ntpts,
                   int32_t reps, int32_t* best_tile, int32_t* best_tpt, float* best_ms);
int o3d_batch_config(const o3d_batch* b, int32_t* tile, int32_t* tpt, int32_t* grid, int64_t* items);
/* How many instances the factored method evaluates, and the scratch they share. */
int o3d_batch_factored(const o3d_batch* b, int32_t* ninst, int64_t* scratch_bytes);

/* Autotuner for single-collection calls: time the candidates on the given device
 * collection and record the winner for (core, results, precision, target kind,
 * size bucket); later calls with tile = tpt = 0 use it.  The table can be saved
 * and loaded as text. */
int o3d_tune_p2p(const o3d_points* src, o3d_points* targ, const o3d_opts* opts, int32_t reps,
                 int32_t* best_tile, int32_t* best_tpt, float* best_ms);
int o3d_tune_save(const char* path);
int o3d_tune_load(const char* path);

/* ---- serving: device-resident simulations stepped by continuous batching ------
 * (no counterpart in the reference; each request is one inviscid Ralston RK2
 * convection step with vortex stretching, Convection.h:349-425, of a session whose
 * particle state stays in HBM between steps).
 *
 * A server owns one HBM arena of opts.hbm_budget bytes, one host thread and one
 * stream.  Session state costs 32 B/particle (position, strength, radius,
 * elongation); a step request needs 120 B/particle more while it runs (interim
 * state, both stages' velocity and gradient).  Requests wait in arrival order
 * and are admitted as soon as a slot and arena space are free, in any tick; each
 * tick evaluates a budget of (request, target-chunk) items from every admitted
 * request in one persistent launch, then applies the RK2 stage updates of the
 * requests whose stage completed.  A step's result does not depend on what else
 * was in the batch.  Steps of one session run in submission order. */
typedef struct o3d_server o3d_server;

typedef struct o3d_server_opts {
  size_t hbm_budget;          /* arena bytes (0 = 8 GiB)                               */
  int32_t max_active;         /* admitted requests at once (0 = 64)                    */
  int32_t tile;               /* kernel tile / targets per thread (0 = defaults)       */
  int32_t tpt;
  int32_t use_graph;          /* replay each tick from a captured HIP graph            */
  int32_t adaptive_tick;      /* raise the budget to (workgroup slots x the largest
                                 active chunk) so that large sessions fill the GPU      */
  int32_t method;             /* enum o3d_method.  AUTO (0): a step of a session of
                                 >= 32,768 particles with the RM or EXP core evaluates
                                 both stages by the factored far-field method on a
                                 second stream, overlapping the ticks; FACTORED: every
                                 RM/EXP session; DIRECT: ticks only.  The factored
                                 scratch is sized for the largest such session and is
                                 held outside the arena                                */
  double tick_interactions;   /* pair-evaluation budget of one tick (0 = 4e10)         */
} o3d_server_opts;

enum o3d_request_state { O3D_RQ_QUEUED = 0, O3D_RQ_RUNNING = 1, O3D_RQ_DONE = 2, O3D_RQ_FAILED = 3,
                         O3D_RQ_CANCELLED = 4 };

typedef struct o3d_request_info {
  int32_t state;              /* enum o3d_request_state                                */
  int32_t status;             /* o3d_status of the step (O3D_OK when done)             */
  double submit_s;            /* server clock, seconds                                 */
  double admit_s;
  double done_s;
  double latency_s;           /* done - submit                                         */
  double interactions;        /* 2 N^2 for a step                                      */
} o3d_request_info;

typedef struct o3d_server_stats {
  int64_t submitted, completed, failed, cancelled, queued, running;
  double p50_s, p99_s, mean_s;      /* latency of completed steps                       */
  double interactions;              /* pair evaluations of completed steps              */
  double wall_s;                    /* first submit -> last completion                  */
  double ginter_per_s;              /* interactions / wall_s / 1e9                      */
  size_t hbm_used, hbm_budget, hbm_peak;
  int64_t ticks;
  int64_t factored_steps;           /* steps evaluated by the factored method           */
  size_t factored_scratch;          /* bytes of its scratch (outside the arena)         */
} o3d_server_stats;

int o3d_server_create(o3d_server** s, const o3d_server_opts* opts);
int o3d_server_destroy(o3d_server* s);   /* waits for running steps, cancels queued ones */
/* host arrays in (x, s, r of *p); core fixed per session; O3D_EADMIT if the state
 * does not fit the arena, O3D_EINVAL on non-finite input */
int o3d_server_open_session(o3d_server* s, const o3d_points* p, int32_t core, int64_t* session);
int o3d_server_close_session(o3d_server* s, int64_t session);
/* copy a session's current state to host arrays (waits for its queued steps);
 * elong may be NULL */
int o3d_server_read_session(o3d_server* s, int64_t session, float* const x[3], float* const str[3],
                            float* elong);
/* queue one step: dt, precision (FP32 or compensated), freestream fs[3] (NULL = 0),
 * timeout_s (<= 0: none; a request still queued at its deadline is cancelled) */
int o3d_server_submit(o3d_server* s, int64_t session, double dt, int32_t precision, const double* fs,
                      double timeout_s, int64_t* request);
/* wait up to timeout_s (< 0: forever) for a request to end; O3D_EBUSY if it has not */
int o3d_server_wait(o3d_server* s, int64_t request, double timeout_s, o3d_request_info* info);
int o3d_server_cancel(o3d_server* s, int64_t request);
int o3d_server_get_stats(o3d_server* s, o3d_server_stats* st);
int o3d_server_reset_stats(o3d_server* s);
/* test hook: the next `count` admitted requests fail with O3D_EINJECTED instead of
 * running, as if their kernels had failed (kind 1); kind 0 clears */
int o3d_inject_fault(int32_t kind, int32_t count);

/* ---- device plumbing (no counterpart in the reference: device-resident use) -- */
int o3d_dev_malloc(void** ptr, size_t bytes);
int o3d_dev_free(void* ptr);
int o3d_memcpy_htod(void* dst, const void* src, size_t bytes);
int o3d_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int o3d_memset_dev(void* dst, int value, size_t bytes);
int o3d_device_synchronize(void);
int o3d_stream_create(void** stream);
/* Waits for the stream, frees the device workspace the library keeps for it (the
 * scratch of factored, split and panel calls on that stream), then destroys it. */
int o3d_stream_destroy(void* stream);
int o3d_stream_synchronize(void* stream);
int o3d_event_create(void** ev);
int o3d_event_destroy(void* ev);
int o3d_event_record(void* ev, void* stream);
int o3d_event_elapsed_ms(float* ms, void* start, void* stop);
/* A streaming float4 copy kernel of exactly `bytes` read and `bytes` written
 * (bytes a multiple of 16): the known-traffic reference that HBM byte counters
 * (FETCH_SIZE, WRITE_SIZE) are calibrated against. */
int o3d_copy_probe(void* dst, const void* src, size_t bytes, void* stream);
/* VALU issue-rate probe (no counterpart in the reference; calibrates the FP32
 * rooflines): a grid of 8 workgroups per CU runs 16 independent chains of one
 * instruction per lane, 64 instructions per loop iteration (op 0 v_fmac_f32 with a scalar-register operand, 1 v_add_f32,
 * 2 v_mul_f32, 3 v_rsq_f32, 4 v_fmac_f32 with vector-register operands only, all at
 * 32 waves per CU); ops 5-8 at one wave per SIMD: 5 v_fmac_f32 (vector-register
 * operands), 6 v_mfma_f32_32x32x16_f16, 7 both in the same waves (16 FMAs per MFMA),
 * 8 one MFMA wave and one FMA wave per SIMD.  Reports the second launch's time and
 * its wave64 count of the probed instruction (ops 6-8: the MFMAs).  Synchronous. */
int o3d_valu_probe(int32_t op, float* ms, double* wave_instructions);
/* device name, CU count, clock (kHz), HBM bytes */
int o3d_device_info(char* name, int name_len, int* cus, int* clock_khz, size_t* hbm_bytes);

#ifdef __cplusplus
}
#endif
#endif /* O3D_BIOT_H */
