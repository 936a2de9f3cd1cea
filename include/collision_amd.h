/*
 * collision_amd.h -- C ABI of the MI355X (gfx950) collision step.
 *
 * Drop-in boundary for the reference's CollisionSolver / CollisionSolver3d
 * (collid.h:128-244, dcollid.cpp:56-107, 317-362).  The reference mutates
 * FronTier POINT/STATE records in place; here a handle owns device-resident
 * copies of the same per-vertex state (Coords, x_old, vel, avgVel,
 * has_collsn) and element topology, and the caller moves state in and out
 * with plain host arrays.  All arrays are row-major [n][3] doubles, int32
 * indices, uint8 flags.  No torch or HIP types appear in the signatures.
 *
 * Every call returns a cs_status; cs_last_error() gives text.  Handles are
 * independent (no global state), each owns one HIP stream; calls on one
 * handle must not overlap.
 */
#ifndef COLLISION_AMD_H
#define COLLISION_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cs_status {
    CS_OK = 0,
    CS_ERR_ARG = 1,          /* bad argument or call order */
    CS_ERR_NONFINITE = 2,    /* NaN/inf where the reference calls clean_up(ERROR)
                                (dcollid.cpp:178-184, 713-718; dcollid3d.cpp:750-768,
                                 898-902, 1093-1105, 1291-1297) */
    CS_ERR_UNSUPPORTED = 3,  /* input outside the implemented model */
    CS_ERR_HIP = 4,          /* HIP runtime failure */
    CS_ERR_NOMEM = 5,
    CS_ERR_NO_DEVICE = 6,
    CS_ERR_INTERNAL = 7,
    CS_WARN_ZONE_CAP = 16    /* impact-zone loop stopped by max_zone_passes; the
                                reference loop is unbounded (dcollid.cpp:236) */
} cs_status;

typedef enum cs_pass_kind { CS_PASS_PROXIMITY = 0, CS_PASS_CCD = 1, CS_PASS_ZONE = 2 } cs_pass_kind;

/* Solver parameters: the reference's static setters (dcollid.cpp:56-89)
 * plus iteration caps. */
typedef struct cs_params {
    double dt;           /* setTimeStepSize       (s_dt,        default 1e-3)  */
    double thickness;    /* setFabricThickness    (s_thickness, default 1e-4)  */
    double round_eps;    /* setRoundingTolerance  (s_eps,       default 1e-6)  */
    double spring_k;     /* setSpringConstant     (s_k,         default 1000)  */
    double mass;         /* setPointMass          (s_m,         default 0.01)  */
    double friction;     /* setFrictionConstant   (s_lambda,    default 0.02)  */
    double restitution;  /* setRestitutionCoef    (s_cr,        default 0)     */
    double dom_lo[3];    /* setDomainBoundary L */
    double dom_hi[3];    /* setDomainBoundary U */
    int32_t max_ccd_passes;     /* 5  (dcollid.cpp:433) */
    int32_t max_zone_passes;    /* cap for the unbounded impact-zone loop */
    int32_t max_strain_sweeps;  /* 10 (dcollid.cpp:590) */
    int32_t log_pairs;          /* record hit pairs of every pass (parity taps) */
    int32_t device;             /* HIP device ordinal */
    int32_t force_full_passes;  /* 1: every pass re-runs the whole broad and narrow
                                   phase (0: the first CCD pass reuses the proximity
                                   pass's tree walk, later passes only redo pairs
                                   whose vertices changed; results are identical
                                   either way) */
    int32_t pass_timing;        /* per-pass times in cs_pass_info: 2 phase and kernel
                                   times (default), 1 broad- and narrow-phase kernel
                                   times only, 0 none; each timing point is a marker
                                   on the stream, which costs GPU time */
    int32_t record_impulses;    /* 1: keep each vertex's summed collsnImpulse and
                                   collsn_num of the last pass with a hit (cs_last_impulses,
                                   cs_write_vtk); the reference's CollisionImpulse
                                   debugging dump (dcollid.cpp:684-694) */
    int32_t precision;          /* 0: FP64 throughout (the plain form); 1: mixed -- the
                                   proximity pass screens every sub-test in FP32 from a
                                   local origin with certified forward-error bounds and
                                   runs the FP64 test only where the screen cannot
                                   prove a reject (results identical to 0) */
    int32_t learned;            /* 1: the learned CCD form (cs_learned_*): an MLP scores
                                   every CCD sub-test and seeds a bracketed Newton
                                   solve; the exact cubic still decides */
    int32_t precision_stats;    /* 1: with precision = 1, also evaluate the FP64
                                   quantities of every screened sub-test and report the
                                   largest FP32 error (cs_step_report.mx_*) */
    int32_t graph_step;         /* 1: cs_step / cs_step_batch replay one captured HIP graph
                                   of the predictor and the resolve prologue (every launch
                                   before the step's first host synchronisation) instead
                                   of launching them one by one; results are identical */
    int32_t autotune;           /* 1: the first 8 steps after cs_assemble time launch-shape
                                   candidates (cs_tuning) and keep the fastest; with the
                                   environment variable CS_TUNE_CACHE naming a file, the
                                   choice is stored there per device, size and batch, and
                                   a later assembly of the same key takes it without tuning.
                                   Results do not depend on the launch shapes. */
    int32_t pad2;
} cs_params;

/* Flat mesh: replaces assembleFromInterface's walk over FronTier surfaces and
 * string curves (dcollid3d.cpp:12-52).  Element order = triangles then
 * bonds, exactly the reference's hseList order. */
typedef struct cs_mesh_view {
    int32_t nv, nt, nb;
    const int32_t *tris;      /* [nt][3] vertex ids                          */
    const int32_t *tri_surf;  /* [nt] surface id (same-surface rigid skip)   */
    const double *tri_len0;   /* [nt][3] rest length of edge (j, j+1)        */
    const int32_t *bonds;     /* [nb][2] (start, end)                        */
    const double *bond_len0;  /* [nb] rest length                            */
    const uint8_t *is_fixed;  /* [nv] static rigid-body point (STATE.is_fixed) */
    /* movable rigid bodies (STATE.is_movableRG, cdinit.cpp:201-202); leave
     * nbody = 0 and the pointers NULL for none.  A body is one surface whose
     * triangles all have vertices of that body; bonds hold no movable vertex
     * (the reference's curve points are never movable). */
    int32_t nbody;
    const int32_t *vert_body; /* [nv] body of a movable vertex, -1 otherwise      */
    const double *body_mass;  /* [nbody] total_mass (dcollid3d.cpp:968-969)     */
    const double *body_com;   /* [nbody][3] center_of_mass (dcollid.cpp:641-650) */
    /* batched scenes (no counterpart in the reference, which runs one scene
     * per solver): nscene > 1 independent scenes in one mesh, elem_scene[e]
     * the scene of element e (triangles then bonds).  Elements of different
     * scenes share no vertex and are never paired, so each scene runs as it
     * would alone; cs_resolve_batch reports every scene's own pass loop.
     * Scenes laid out apart in space keep the broad phase from meeting
     * other scenes' boxes at all.  Leave nscene = 0 for one scene. */
    int32_t nscene;
    const int32_t *elem_scene;
} cs_mesh_view;

typedef struct cs_pass_info {
    int32_t kind;            /* cs_pass_kind */
    int32_t index;
    int64_t candidates;      /* box-overlapping element pairs; -1 in an incremental
                                pass unless log_pairs is set (the exact count needs
                                an extra query of the changed elements) */
    int64_t narrow_pairs;    /* candidate pairs after the early rejects; -1 in an
                                incremental pass unless log_pairs is set (exact in
                                the first CCD pass, which tests all its pairs) */
    int64_t hit_pairs;       /* pairs whose narrow phase reported contact */
    int64_t hit_subtests;
    int64_t tested_pairs;    /* pairs the narrow-phase kernel evaluated in this pass
                                (all narrow_pairs in a full pass and in the first CCD
                                pass; only pairs touching changed vertices in later
                                incremental passes) */
    int32_t incremental;     /* 1 if the pass reused an earlier pass's results (the
                                first CCD pass reuses the proximity pass's tree
                                walk; later passes also its narrow-phase results) */
    int32_t pad;
    double ms_broad, ms_narrow, ms_response;  /* phase wall time on the handle's stream */
    double ms_traverse_kernel;                /* broad-phase kernels (traversal / re-query) */
    double ms_narrow_kernel;                  /* narrow-phase kernel alone */
} cs_pass_info;

typedef struct cs_step_report {
    int32_t status;
    int32_t n_passes;
    int32_t n_ccd_passes;
    int32_t n_zone_passes;
    int32_t has_collision;   /* hasCollision() */
    int32_t n_zones;
    int32_t strain_sweeps;
    int32_t strain_edges;    /* edges over the limit in the last sweep */
    int32_t degenerate_edges;
    int32_t strain_levels;   /* level steps the strain sweeps ran (sweeps in flight
                                together advance one level per step; sum) */
    int64_t candidates;      /* sum over passes; -1 when a pass reports -1 */
    int64_t narrow_pairs;    /* sum over passes; -1 when a pass reports -1 */
    int64_t hit_pairs, hit_subtests;
    double ms_total, ms_broad, ms_narrow, ms_response, ms_zone, ms_strain;
    int64_t strain_slots;    /* element visits the strain sweeps evaluated (sum) */
    double ms_strain_kernel; /* strain-sweep kernel time (pass_timing >= 1) */
    int64_t host_syncs;      /* host synchronisations of the resolve */
    /* mixed precision (precision = 1): proximity sub-tests the FP32 screen
     * rejected, and those it passed on to FP64; with precision_stats the
     * largest FP32 error over the screened sub-tests of the distance (relative
     * to the thickness) and of the barycentric / segment parameters, and the
     * largest ratio of a screened quantity's FP32 error to the forward-error
     * bound the screen certified it with (at most 1 when the bounds hold) */
    int64_t mx_screened, mx_rechecked;
    double mx_err_dist, mx_err_param;
    double mx_bound_ratio;
} cs_step_report;

typedef struct cs_solver cs_solver;

/* ---- lifecycle (CollisionSolver3d(), ~CollisionSolver(), setters) ---- */
void cs_default_params(cs_params *p);
int cs_create(cs_solver **out, const cs_params *p);
int cs_destroy(cs_solver *h);
int cs_set_params(cs_solver *h, const cs_params *p);
int cs_get_params(cs_solver *h, cs_params *p);
/* CollisionSolver::setDomainBoundary (dcollid.cpp:109-114) */
int cs_set_domain_boundary(cs_solver *h, const double *L, const double *U);

/* ---- topology (CollisionSolver3d::assembleFromInterface) ---- */
int cs_assemble(cs_solver *h, const cs_mesh_view *mesh);

/* ---- state in / out (the FronTier POINT/STATE fields the solver reads and
 *      writes: Coords, vel, has_collsn; x_old and avgVel readable) ---- */
int cs_set_state(cs_solver *h, const double *x, const double *vel, const uint8_t *has_collsn);
int cs_get_state(cs_solver *h, double *x, double *vel, double *avg_vel, double *x_old,
                 uint8_t *has_collsn);
/* CollisionSolver::recordOriginPosition (dcollid.cpp:91-107): x_old <- Coords
 * (movable rigid-body points keep x_old), has_collsn <- false */
int cs_record_origin_position(cs_solver *h);
/* center_of_mass and center_of_mass_velo of every movable body as the last
 * resolve left them (updateFinalForRG, dcollid.cpp:626-675); either pointer
 * may be NULL.  [nbody][3] each. */
int cs_get_bodies(cs_solver *h, double *com, double *com_vel);

/* ---- step ---- */
/* the driver's per-step predictor (test.cpp:192-258): x_old <- x,
 * x <- x_old + dt vel */
int cs_predict(cs_solver *h);
/* CollisionSolver::resolveCollision (dcollid.cpp:317-362); includes the
 * per-step union-find reset done by assembleFromInterface (makeSet).
 * report may be NULL. */
int cs_resolve(cs_solver *h, cs_step_report *report);
/* cs_predict + cs_resolve */
int cs_step(cs_solver *h, cs_step_report *report);

/* One scene's share of a batched resolve: the pass loop resolveCollision
 * runs for that scene alone (dcollid.cpp:317-362) */
typedef struct cs_scene_report {
    int32_t n_passes;        /* 1 + n_ccd_passes + n_zone_passes */
    int32_t n_ccd_passes, n_zone_passes;
    int32_t has_collision;   /* hit in the scene's first CCD pass */
    int32_t zone_cap;        /* the impact-zone loop stopped at max_zone_passes */
    int32_t pad;
    int64_t hit_pairs;       /* hit pairs over the scene's passes */
} cs_scene_report;
/* cs_resolve over all scenes of a batched mesh (one launch per kernel for
 * every scene), then every scene's report: per_scene[nscene], nscene as
 * assembled.  report (may be NULL) covers the whole batch. */
int cs_resolve_batch(cs_solver *h, cs_step_report *report, cs_scene_report *per_scene,
                     int32_t nscene);
/* cs_predict + cs_resolve_batch */
int cs_step_batch(cs_solver *h, cs_step_report *report, cs_scene_report *per_scene,
                  int32_t nscene);
/* CollisionSolver::hasCollision(): 1 if the first CCD pass of the last
 * resolve found a collision */
int cs_has_collision(cs_solver *h);
int cs_synchronize(cs_solver *h);

/* ---- parity taps of the last resolve ---- */
int32_t cs_pass_count(cs_solver *h);
int cs_pass_infos(cs_solver *h, cs_pass_info *out, int32_t cap);
int64_t cs_pair_log_count(cs_solver *h);
/* rows (pass, lo, hi, hit mask); order unspecified, sort before comparing */
int cs_pair_log(cs_solver *h, int32_t *out4, int64_t cap);
/* zone label per vertex = smallest vertex id of its impact zone */
int cs_zone_labels(cs_solver *h, int32_t *out);
/* cs_zone_labels without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The labels are staged
 * in a fixed 16 MiB device array per GPU, so at most 4194304 vertices
 * (CS_ERR_UNSUPPORTED beyond).  All handles on one GPU share that array:
 * calls that overlap on one GPU race. */
int cs_zone_labels_noalloc(cs_solver *h, int32_t *out);

/* y[i] = f(x[i]) on the device with the collision step's own elementary
 * functions (fn 0 acos, 1 cos, 2 sin, 3 pow(x, 1.0/3.0); correctly rounded),
 * the ones isCoplanar and the impact-zone rotation use (dcollid3d.cpp:163-169,
 * 436-444); host arrays.  A parity tap with no counterpart in the reference. */
int cs_math_eval(int32_t device, int32_t fn, const double *x, double *y, int64_t n);

/* ---- learned time-of-impact proposer (no counterpart in the reference) ----
 * An MLP 24 -> 128 -> 128 -> 2 (ReLU; BF16 weights and activations, FP32
 * accumulation on the matrix cores) maps a CCD sub-test to a root-exists
 * logit and a time-of-impact seed in [0, 1] (fraction of dt).  The 24
 * features are the four points' x_old relative to their centroid and their
 * avgVel * dt, both divided by the thickness.  Labels come from the plain
 * CCD itself: root-exists = isCoplanar true (dcollid3d.cpp:371-482), seed =
 * first valid root / dt.  With cs_params.learned = 1 every CCD pass scores
 * its sub-tests, tests its pairs in order of decreasing score and runs a
 * safeguarded bracketed Newton solve of the coplanarity cubic from the seed
 * next to the exact closed form (which still decides; results are those of
 * the plain form bit for bit); cs_learned_stats reports the outcome. */
#define CS_MLP_PARAMS 19970   /* 24*128 + 128 + 128*128 + 128 + 128*2 + 2 */
typedef struct cs_learned_stats {
    int64_t scored;            /* sub-tests scored by the network */
    int64_t exists;            /* of those, with a valid root (exact cubic) */
    int64_t predicted_exists;  /* score > 0 */
    int64_t agree;             /* prediction == exact outcome */
    int64_t newton_roots;      /* roots the seeded Newton solve located */
    int64_t newton_iters;      /* its iterations (seeded at the network's TOI) */
    int64_t newton_iters_mid;  /* iterations of the same solve seeded at dt / 2 */
    double max_root_err;       /* largest |Newton root - closed-form root| / dt */
} cs_learned_stats;
/* parameters, FP32 master copy, in the order W1[128][24], b1[128],
 * W2[128][128], b2[128], W3[2][128], b3[2] (row = output unit) */
int cs_learned_set_weights(cs_solver *h, const float *w);
int cs_learned_get_weights(cs_solver *h, float *w);
/* random initialisation (He uniform, seeded) */
int cs_learned_init(cs_solver *h, uint64_t seed);
/* forward pass on host features x[n][24]: out[n][2] = (logit, seed) */
int cs_learned_infer(cs_solver *h, const float *x, float *out, int64_t n);
/* loss and gradient of the training objective on one batch (mean over the
 * batch of BCE(logit, exists) + exists * (sigmoid(s) - toi)^2), for checks */
int cs_learned_grad(cs_solver *h, const float *x, const float *exists, const float *toi,
                    int64_t n, float *grad, float *loss);
/* from now on, every CCD pass of cs_resolve appends (up to `capacity` in
 * total) sub-tests with their labels to the handle's dataset, keeping those
 * whose hash of (pair, sub-test, pass) is divisible by `stride`; a call
 * empties the dataset, capacity 0 stops collecting and frees it.  The
 * samples' order in the dataset is that of the device's atomics; training
 * puts them in a canonical order first. */
int cs_learned_collect(cs_solver *h, int64_t capacity, int32_t stride);
int64_t cs_learned_dataset_size(cs_solver *h);
/* copy the dataset out: x[n][24] (float), exists[n], toi[n] */
int cs_learned_dataset(cs_solver *h, float *x, float *exists, float *toi, int64_t cap);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) over shuffled batches of the
 * collected dataset, all on the device; loss[i] = mean loss of epoch i.
 * The classes are balanced by repetition: each sample with a root appears
 * round(#without / #with) times per epoch (1 to 64).  The same dataset,
 * weights and seed give the same weights bit for bit. */
int cs_learned_train(cs_solver *h, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                     float *loss);
/* statistics of the learned form since the previous call (which resets them) */
int cs_learned_get_stats(cs_solver *h, cs_learned_stats *out);

/* ---- launch shapes (no counterpart in the reference) ----
 * persistent traversal work-groups per CU (4, 8, 16, 32), narrow-phase
 * work-group size (64, 128, 256) and the dirty-element query's grid cap
 * (1024 .. 8192); the results of a step do not depend on them */
typedef struct cs_tuning {
    int32_t traverse_blocks_per_cu;
    int32_t narrow_threads;
    int32_t query_grid;
    int32_t tuned;             /* 0 defaults or set, 1 chosen by the autotuner, 2 from the cache */
} cs_tuning;
int cs_get_tuning(cs_solver *h, cs_tuning *out);
int cs_set_tuning(cs_solver *h, const cs_tuning *in);

/* ---- diagnostics and restart ---- */
/* the summed collsnImpulse [nv][3] and collsn_num [nv] every vertex received
 * in the last pass with a hit of the last resolve (needs record_impulses;
 * static points, points without a hit and resolves without one hold zero) */
int cs_last_impulses(cs_solver *h, double *imp, int32_t *num);
/* vtkplotVectorSurface (vtk.cpp:14-93): a VTK XML PolyData file (.vtp, ASCII)
 * of the current Coords, one polygon per element in element order, point
 * vectors "CollsnImpulse" and cell colours "CollsnRegion" (red: a point of
 * the element had a hit in the pass cs_last_impulses reports, green
 * otherwise); needs record_impulses */
int cs_write_vtk(cs_solver *h, const char *path);
/* checkpoint / restart (the FronTier driver's FT_Save / RestartRun,
 * test.cpp:72-91, 122-123): the per-vertex state (Coords, x_old, vel,
 * avgVel, has_collsn), the movable-body state (center_of_mass,
 * center_of_mass_velo, mrg_com, collsnImpulse_RG) and the parameters.
 * cs_load needs a handle assembled with the same topology (checked) and
 * restores everything but the device ordinal; stepping on gives bitwise
 * the run that was never interrupted. */
int cs_save(cs_solver *h, const char *path);
int cs_load(cs_solver *h, const char *path);

/* text of the last error on this handle (or of a failed cs_create if h is NULL) */
const char *cs_last_error(cs_solver *h);
/* library version string */
const char *cs_version(void);

#ifdef __cplusplus
}
#endif
#endif
