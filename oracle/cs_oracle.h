/*
 * cs_oracle: CPU FP64 restatement of the reference collision step.
 * TEST INFRASTRUCTURE ONLY.  It is the checker for the HIP library
 * (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
 * it); nothing in the product links or calls it.
 *
 * Reference files followed: dcollid.cpp (pass control, Jacobi apply,
 * union-find, strain limiting, domain box, finalize, element boxes),
 * dcollid3d.cpp (proximity, CCD cubic, impulses, impact-zone rigid motion,
 * sub-test enumeration), AABB.cpp (box margin, closed overlap, incremental
 * tree and its query order), test.cpp:134-258 (per-step predictor).
 */
#ifndef CS_ORACLE_H
#define CS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_ERR_ARG = 1,
    ORC_ERR_NONFINITE = 2,   /* where the reference calls clean_up(ERROR) */
    ORC_ERR_UNSUPPORTED = 3,
    ORC_WARN_ZONE_CAP = 16   /* impact-zone loop stopped by the cap */
};

enum { ORC_ORDER_CANONICAL = 0, ORC_ORDER_REFERENCE_TREE = 1 };
enum { ORC_MATH_DET = 0, ORC_MATH_LIBM = 1 };
enum { ORC_PASS_PROXIMITY = 0, ORC_PASS_CCD = 1, ORC_PASS_ZONE = 2 };

typedef struct orc_params {
    double dt;          /* s_dt                       (dcollid.cpp:31)  */
    double thickness;   /* s_thickness                (dcollid.cpp:30)  */
    double round_eps;   /* s_eps                      (dcollid.cpp:29)  */
    double spring_k;    /* s_k                        (dcollid.cpp:32)  */
    double mass;        /* s_m                        (dcollid.cpp:33)  */
    double friction;    /* s_lambda                   (dcollid.cpp:34)  */
    double restitution; /* s_cr                       (dcollid.cpp:35)  */
    double dom_lo[3], dom_hi[3]; /* Boundary[dir][side] (collid.h:142) */
    int32_t max_ccd_passes;      /* 5  (dcollid.cpp:433) */
    int32_t max_zone_passes;     /* reference: unbounded (dcollid.cpp:236) */
    int32_t max_strain_sweeps;   /* 10 (dcollid.cpp:590) */
    int32_t order_mode;          /* ORC_ORDER_* */
    int32_t math_mode;           /* ORC_MATH_* */
    int32_t log_pairs;           /* record every hit pair of every pass */
    int32_t record_impulses;     /* keep collsnImpulse / collsn_num of the last pass with a hit */
    double mach_eps;             /* FronTier MACH_EPS; DBL_EPSILON by default (the
                                    value is not in the reference tree) */
} orc_params;

typedef struct orc_pass_info {
    int32_t kind;            /* ORC_PASS_* */
    int32_t index;           /* pass number within its kind */
    int64_t candidates;      /* box-overlapping element pairs */
    int64_t narrow_pairs;    /* candidates surviving the shared-vertex / rigid checks */
    int64_t hit_pairs;       /* pairs whose narrow phase returned true */
    int64_t hit_subtests;    /* sub-tests that hit */
} orc_pass_info;

typedef struct orc_report {
    int32_t status;
    int32_t n_passes;
    int32_t n_ccd_passes;
    int32_t n_zone_passes;
    int32_t has_collision;   /* hasCollision() (dcollid.cpp:456-457) */
    int32_t n_zones;         /* zones of impact after the last zone pass */
    int32_t strain_sweeps;
    int32_t strain_edges;    /* edges over the limit in the last sweep */
    int32_t degenerate_edges;
    int32_t pad;
    int64_t candidates;      /* summed over passes */
    int64_t narrow_pairs;
    int64_t hit_pairs;
    int64_t hit_subtests;
} orc_report;

typedef struct orc_solver orc_solver;

void orc_default_params(orc_params *p);
orc_solver *orc_create(const orc_params *p);
void orc_destroy(orc_solver *s);
int orc_set_params(orc_solver *s, const orc_params *p);

/* topology: triangles (tri_surf = surface id, tri_len0 = rest lengths of
 * edges (j, j+1)), bonds, static flags.  Element order = tris then bonds. */
int orc_assemble(orc_solver *s, int32_t nv, int32_t nt, int32_t nb,
                 const int32_t *tris, const int32_t *tri_surf,
                 const double *tri_len0, const int32_t *bonds,
                 const double *bond_len0, const uint8_t *is_fixed);

/* movable rigid bodies (STATE.is_movableRG, cdinit.cpp:201-202): vert_body[v]
 * is the body of vertex v or -1; a body is one closed surface whose
 * triangles all have movable vertices of that body, and bonds hold no
 * movable vertex.  body_mass = total_mass, body_com = center_of_mass of
 * each body (dcollid.cpp:641-650).  Call after orc_assemble. */
int orc_set_bodies(orc_solver *s, int32_t nbody, const int32_t *vert_body,
                   const double *body_mass, const double *body_com);
/* center_of_mass and center_of_mass_velo of every body after the last
 * resolve (updateFinalForRG, dcollid.cpp:626-675) */
int orc_get_bodies(orc_solver *s, double *com, double *com_vel);

/* per-vertex state in [nv][3] row-major */
int orc_set_state(orc_solver *s, const double *x, const double *vel,
                  const uint8_t *has_collsn);
int orc_get_state(orc_solver *s, double *x, double *vel, double *avg_vel,
                  double *x_old, uint8_t *has_collsn);

/* driver step before resolve (test.cpp:192-258): x_old <- x, x <- x_old + dt vel */
int orc_predict(orc_solver *s);
/* CollisionSolver::resolveCollision (dcollid.cpp:317-362), including the
 * per-step makeSet done by assembleFromInterface (dcollid3d.cpp:44) */
int orc_resolve(orc_solver *s, orc_report *rep);

/* diagnostics of the last resolve */
int32_t orc_pass_count(orc_solver *s);
int orc_pass_infos(orc_solver *s, orc_pass_info *out, int32_t cap);
int64_t orc_pair_log_count(orc_solver *s);
/* rows of (pass, lo, hi, hitmask) */
int orc_pair_log(orc_solver *s, int32_t *out4, int64_t cap);
/* zone label per vertex = smallest vertex id in its impact zone */
int orc_zone_labels(orc_solver *s, int32_t *out);

/* summed collsnImpulse [nv][3] and collsn_num [nv] of the last pass with a
 * hit in the last resolve, zero if none (record_impulses) */
int orc_last_impulses(orc_solver *s, double *imp, int32_t *num);

/* ---- test hooks ---- */
/* overwrite avgVel (after orc_set_state) */
int orc_set_avg_vel(orc_solver *s, const double *avg_vel);
/* one pair through the narrow phase (isProximity / isCollision with the
 * Moving* dispatch), with the impact-zone merge rule active when zone_mode
 * is set; returns the hit mask of the sub-tests, impulses stay in the
 * accumulators (not applied).  Zone labels come from orc_zone_labels. */
int32_t orc_debug_pair(orc_solver *s, int32_t A, int32_t B, int32_t ccd, int32_t zone_mode);
/* reset the union-find to the per-step state (makeSet + body zones) */
int orc_debug_make_set(orc_solver *s);

/* ---- unit entry points (single sub-tests on explicit data) ---- */
/* isCoplanar: x_old[4][3], avg_vel[4][3] -> roots[4]; returns 0/1 */
int orc_is_coplanar(const double *x_old, const double *avg_vel, double dt,
                    int32_t math_mode, double *roots);
/* PointToTri / EdgeToEdge on coords c[4][3] with x_old[4][3], avg_vel[4][3];
 * returns 1 on a hit and writes the per-point impulse and friction deltas
 * (imp[4][3], fric[4][3]) and counts[4]; -1 on a nonfinite error */
int orc_point_to_tri(const orc_params *p, const double *c, const double *x_old,
                     const double *avg_vel, const uint8_t *is_fixed, double h,
                     double *imp, double *fric, int32_t *counts);
int orc_edge_to_edge(const orc_params *p, const double *c, const double *x_old,
                     const double *avg_vel, const uint8_t *is_fixed, double h,
                     double *imp, double *fric, int32_t *counts);
/* updateImpactListVelocity on one zone given in member order; lanes = 0
 * sums members sequentially (reference), 64 in the canonical lane order */
int orc_zone_rigid_update(const orc_params *p, int32_t n, const double *x_old,
                          double *avg_vel, const uint8_t *is_fixed, int32_t lanes);
/* elementary functions of the deterministic math mode (correctly rounded) */
double orc_math_acos(double x);
double orc_math_cos(double x);
double orc_math_sin(double x);
double orc_math_pow_third(double x);   /* pow(x, 1.0/3.0) */
/* y[i] = f(x[i]); fn 0..3: acos, cos, sin, pow(x, 1.0/3.0) of the
 * deterministic mode, 10..13: the same through the host libm */
int orc_math_eval(int32_t fn, const double *x, double *y, int64_t n);
/* calls in libm mode whose libm result was not the correctly rounded one,
 * since the last reset (process-wide) */
int64_t orc_libm_misrounded(int32_t reset);

#ifdef __cplusplus
}
#endif
#endif
