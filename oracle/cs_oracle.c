/*
 * cs_oracle.c -- CPU FP64 restatement of the reference collision step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  Every function cites the reference file:line it restates.
 *
 * Numerics follow the reference operation by operation (left-to-right sums,
 * no fused multiply-add; build with -ffp-contract=off).  FronTier macros the
 * reference relies on but that are absent from its tree are reconstructed as
 *   Dot3d(a,b)   = a0*b0 + a1*b1 + a2*b2        (left to right)
 *   Cross3d      = (a1b2-a2b1, a2b0-a0b2, a0b1-a1b0)
 *   Mag3d(a)     = sqrt(Dot3d(a,a))
 *   distance_between_positions(p,q,3) = sqrt(sum_i (p_i-q_i)^2), i = 0,1,2
 *   MACH_EPS     = DBL_EPSILON (a parameter, orc_params.mach_eps)
 * (parity unpinned at that boundary: no reference test fixes them).
 *
 * Two processing orders:
 *   ORC_ORDER_REFERENCE_TREE: the reference's incremental AABB tree
 *     (AABB.cpp:129-197), refit/rebuild rule (dcollid.cpp:366-428) and
 *     in-order leaf query (AABB.cpp:254-343); pairs reach the narrow phase in
 *     exactly the reference's order.
 *   ORC_ORDER_CANONICAL: the same pair set processed sorted by
 *     (lo, hi) element id, dispatched as (a = hi, b = lo), i.e. the reference's
 *     order for a tree whose leaves are in element order; impact-zone sums
 *     run over members in ascending vertex id.  The HIP library reproduces
 *     this order, so the two agree bitwise.
 */
#include "cs_oracle.h"
#include "orc_detmath.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ROUND_EPS 1e-10              /* collid.h:17 */
#define BOX_EPS 1e-6                 /* AABB.cpp:10-33 */
#define SUPERELAS_TOL 0.10           /* dcollid.cpp:489 */
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* std::min / std::max semantics (first argument wins ties and NaN compares) */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }

static inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double mag3(const double *a) { return sqrt(dot3(a, a)); }
static inline void sub3(const double *a, const double *b, double *c) {
    c[0] = a[0] - b[0]; c[1] = a[1] - b[1]; c[2] = a[2] - b[2];
}
static inline double dist3(const double *a, const double *b) {
    double s = 0.0;
    for (int i = 0; i < 3; ++i) { double d = a[i] - b[i]; s += d * d; }
    return sqrt(s);
}
/* myDet3d, dcollid.cpp:977-981 */
static double det3(const double a[3][3]) {
    return a[0][0] * (a[1][1] * a[2][2] - a[2][1] * a[1][2])
         - a[0][1] * (a[1][0] * a[2][2] - a[2][0] * a[1][2])
         + a[0][2] * (a[1][0] * a[2][1] - a[2][0] * a[1][1]);
}

/* ------------------------------------------------------------------------ */
typedef struct { double lo[3], hi[3]; } box_t;

typedef struct {
    int32_t *left, *right, *parent, *leaf; /* leaf = element id or -1 */
    uint8_t *has_box;
    box_t *box;
    int32_t n, cap, root;
    int32_t *leaf_node;                    /* element -> leaf node */
    int built;
} tree_t;

typedef struct { void *p; int64_t n, cap; size_t sz; } vec_t;

static int vec_push(vec_t *v, const void *item) {
    if (v->n == v->cap) {
        int64_t nc = v->cap ? v->cap * 2 : 1024;
        void *np = realloc(v->p, (size_t)nc * v->sz);
        if (!np) return -1;
        v->p = np; v->cap = nc;
    }
    memcpy((char *)v->p + (size_t)v->n * v->sz, item, v->sz);
    v->n++;
    return 0;
}

struct orc_solver {
    orc_params P;
    int32_t nv, nt, nb, ne;
    int32_t *ev;       /* [ne][3]; bonds have ev[3e+2] = -1 */
    int32_t *surf;
    double *len0;      /* [ne][3] */
    uint8_t *fixed;
    /* movable rigid bodies: per vertex body id (-1: none); per body mass,
     * center_of_mass, center_of_mass_velo, the solver's mrg_com copy
     * (collid.h:174) and the collsnImpulse_RG / collsn_num_RG accumulators,
     * which every point of a body holds identically (SpreadImpactZoneImpulse
     * adds the same amount to all of them, dcollid.cpp:1101-1115) */
    int32_t *vbody;
    int32_t nbody;
    double *bmass, *bcom, *bvel, *mrg_com, *rg_imp;
    int32_t *rg_num;
    double *x, *x_old, *vel, *av, *imp, *fric;
    double *last_imp;  /* record_impulses: the last pass's collsnImpulse sums */
    int32_t *last_num;
    int32_t *num;
    uint8_t *has;
    int32_t *uf_root, *uf_next, *uf_tail, *uf_w;
    uint8_t *sorted;
    box_t *ebox;       /* per element box of the current pass */
    tree_t prox, coll;
    double volume;     /* shared between both trees (collid.h:133) */
    int zone_mode;     /* s_detImpZone */
    int err;
    vec_t pairs;       /* int64 packed (a << 32 | b) in processing order */
    vec_t passes;      /* orc_pass_info */
    vec_t plog;        /* int32[4] */
    int32_t cur_pass;
    orc_pass_info cur;
    uint8_t *visited;  /* leaf already queried (nodeSet) */
    int32_t *stack;
};

static int elem_is_tri(const orc_solver *s, int32_t e) { return s->ev[3 * e + 2] >= 0; }

void orc_default_params(orc_params *p) {
    memset(p, 0, sizeof(*p));
    p->dt = 1e-3;            /* DT, collid.h:19 */
    p->thickness = 1e-4;     /* dcollid.cpp:30 */
    p->round_eps = 1e-6;     /* EPS, collid.h:18 */
    p->spring_k = 1000.0;    /* dcollid.cpp:32 */
    p->mass = 0.01;          /* dcollid.cpp:33 */
    p->friction = 0.02;      /* dcollid.cpp:34 */
    p->restitution = 0.0;    /* dcollid.cpp:35 */
    for (int i = 0; i < 3; ++i) { p->dom_lo[i] = -1e30; p->dom_hi[i] = 1e30; }
    p->max_ccd_passes = 5;
    p->max_zone_passes = 1000;
    p->max_strain_sweeps = 10;
    p->order_mode = ORC_ORDER_CANONICAL;
    p->math_mode = ORC_MATH_DET;
    p->log_pairs = 0;
    p->mach_eps = DBL_EPSILON;
}

orc_solver *orc_create(const orc_params *p) {
    orc_solver *s = (orc_solver *)calloc(1, sizeof(orc_solver));
    if (!s) return NULL;
    if (p) s->P = *p; else orc_default_params(&s->P);
    s->pairs.sz = sizeof(int64_t);
    s->passes.sz = sizeof(orc_pass_info);
    s->plog.sz = 4 * sizeof(int32_t);
    return s;
}

static void tree_free(tree_t *t) {
    free(t->left); free(t->right); free(t->parent); free(t->leaf);
    free(t->has_box); free(t->box); free(t->leaf_node);
    memset(t, 0, sizeof(*t));
}

static void free_topology(orc_solver *s) {
    free(s->ev); free(s->surf); free(s->len0); free(s->fixed);
    free(s->x); free(s->x_old); free(s->vel); free(s->av); free(s->imp); free(s->fric);
    free(s->num); free(s->has);
    free(s->uf_root); free(s->uf_next); free(s->uf_tail); free(s->uf_w); free(s->sorted);
    free(s->ebox); free(s->visited); free(s->stack);
    free(s->vbody); free(s->bmass); free(s->bcom); free(s->bvel); free(s->mrg_com);
    free(s->rg_imp); free(s->rg_num); free(s->last_imp); free(s->last_num);
    s->last_imp = NULL; s->last_num = NULL;
    s->vbody = NULL; s->bmass = s->bcom = s->bvel = s->mrg_com = s->rg_imp = NULL;
    s->rg_num = NULL; s->nbody = 0;
    tree_free(&s->prox); tree_free(&s->coll);
}

void orc_destroy(orc_solver *s) {
    if (!s) return;
    free_topology(s);
    free(s->pairs.p); free(s->passes.p); free(s->plog.p);
    free(s);
}

int orc_set_params(orc_solver *s, const orc_params *p) {
    if (!s || !p) return ORC_ERR_ARG;
    s->P = *p;
    return ORC_OK;
}

int orc_assemble(orc_solver *s, int32_t nv, int32_t nt, int32_t nb,
                 const int32_t *tris, const int32_t *tri_surf, const double *tri_len0,
                 const int32_t *bonds, const double *bond_len0, const uint8_t *is_fixed) {
    if (!s || nv <= 0 || nt < 0 || nb < 0) return ORC_ERR_ARG;
    free_topology(s);
    s->nv = nv; s->nt = nt; s->nb = nb; s->ne = nt + nb;
    int32_t ne = s->ne;
    s->ev = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(ne ? ne : 1));
    s->surf = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ne ? ne : 1));
    s->len0 = (double *)calloc(3 * (size_t)(ne ? ne : 1), sizeof(double));
    s->fixed = (uint8_t *)calloc((size_t)nv, 1);
    size_t n3 = 3 * (size_t)nv;
    s->x = (double *)calloc(n3, sizeof(double));
    s->x_old = (double *)calloc(n3, sizeof(double));
    s->vel = (double *)calloc(n3, sizeof(double));
    s->av = (double *)calloc(n3, sizeof(double));
    s->imp = (double *)calloc(n3, sizeof(double));
    s->fric = (double *)calloc(n3, sizeof(double));
    s->num = (int32_t *)calloc((size_t)nv, sizeof(int32_t));
    s->has = (uint8_t *)calloc((size_t)nv, 1);
    s->uf_root = (int32_t *)malloc(sizeof(int32_t) * (size_t)nv);
    s->uf_next = (int32_t *)malloc(sizeof(int32_t) * (size_t)nv);
    s->uf_tail = (int32_t *)malloc(sizeof(int32_t) * (size_t)nv);
    s->uf_w = (int32_t *)malloc(sizeof(int32_t) * (size_t)nv);
    s->sorted = (uint8_t *)calloc((size_t)nv, 1);
    s->ebox = (box_t *)malloc(sizeof(box_t) * (size_t)(ne ? ne : 1));
    s->visited = (uint8_t *)calloc((size_t)(ne ? ne : 1), 1);
    s->vbody = (int32_t *)malloc(sizeof(int32_t) * (size_t)nv);
    for (int32_t v = 0; v < nv; ++v) s->vbody[v] = -1;
    s->stack = (int32_t *)malloc(sizeof(int32_t) * (size_t)(2 * ne + 64));
    for (int32_t e = 0; e < nt; ++e) {
        for (int j = 0; j < 3; ++j) {
            int32_t v = tris[3 * e + j];
            if (v < 0 || v >= nv) return ORC_ERR_ARG;
            s->ev[3 * e + j] = v;
            s->len0[3 * e + j] = tri_len0 ? tri_len0[3 * e + j] : 0.0;
        }
        s->surf[e] = tri_surf ? tri_surf[e] : 0;
    }
    for (int32_t b = 0; b < nb; ++b) {
        int32_t e = nt + b;
        for (int j = 0; j < 2; ++j) {
            int32_t v = bonds[2 * b + j];
            if (v < 0 || v >= nv) return ORC_ERR_ARG;
            s->ev[3 * e + j] = v;
        }
        s->ev[3 * e + 2] = -1;
        s->surf[e] = -1;
        s->len0[3 * e] = bond_len0 ? bond_len0[b] : 0.0;
    }
    if (is_fixed) memcpy(s->fixed, is_fixed, (size_t)nv);
    return ORC_OK;
}

int orc_set_state(orc_solver *s, const double *x, const double *vel, const uint8_t *has) {
    if (!s || !s->x) return ORC_ERR_ARG;
    size_t n3 = 3 * (size_t)s->nv;
    if (x) { memcpy(s->x, x, n3 * sizeof(double)); memcpy(s->x_old, x, n3 * sizeof(double)); }
    if (vel) memcpy(s->vel, vel, n3 * sizeof(double));
    if (has) memcpy(s->has, has, (size_t)s->nv); else memset(s->has, 0, (size_t)s->nv);
    memset(s->imp, 0, n3 * sizeof(double));
    memset(s->fric, 0, n3 * sizeof(double));
    memset(s->num, 0, (size_t)s->nv * sizeof(int32_t));
    memset(s->av, 0, n3 * sizeof(double));
    /* initSurfaceState clears the RG accumulators (cdinit.cpp:192-197) */
    if (s->nbody) {
        memset(s->rg_imp, 0, 3 * (size_t)s->nbody * sizeof(double));
        memset(s->rg_num, 0, (size_t)s->nbody * sizeof(int32_t));
    }
    return ORC_OK;
}

int orc_get_state(orc_solver *s, double *x, double *vel, double *av, double *x_old, uint8_t *has) {
    if (!s || !s->x) return ORC_ERR_ARG;
    size_t n3 = 3 * (size_t)s->nv;
    if (x) memcpy(x, s->x, n3 * sizeof(double));
    if (vel) memcpy(vel, s->vel, n3 * sizeof(double));
    if (av) memcpy(av, s->av, n3 * sizeof(double));
    if (x_old) memcpy(x_old, s->x_old, n3 * sizeof(double));
    if (has) memcpy(has, s->has, (size_t)s->nv);
    return ORC_OK;
}

int orc_set_bodies(orc_solver *s, int32_t nbody, const int32_t *vert_body,
                   const double *body_mass, const double *body_com) {
    if (!s || !s->x || nbody < 0 || (nbody > 0 && (!vert_body || !body_mass || !body_com)))
        return ORC_ERR_ARG;
    free(s->bmass); free(s->bcom); free(s->bvel); free(s->mrg_com); free(s->rg_imp); free(s->rg_num);
    s->bmass = s->bcom = s->bvel = s->mrg_com = s->rg_imp = NULL;
    s->rg_num = NULL;
    s->nbody = 0;
    for (int32_t v = 0; v < s->nv; ++v) {
        int32_t b = vert_body ? vert_body[v] : -1;
        if (b < -1 || b >= nbody) return ORC_ERR_ARG;
        if (b >= 0 && s->fixed[v]) return ORC_ERR_ARG;
        s->vbody[v] = b;
    }
    /* a body is a set of whole triangles; bonds hold no movable point */
    for (int32_t e = 0; e < s->ne; ++e) {
        int np = elem_is_tri(s, e) ? 3 : 2;
        int32_t b0 = s->vbody[s->ev[3 * e]];
        for (int i = 0; i < np; ++i) {
            int32_t b = s->vbody[s->ev[3 * e + i]];
            if (b != b0 || (np == 2 && b >= 0)) return ORC_ERR_ARG;
        }
    }
    size_t nb3 = 3 * (size_t)(nbody ? nbody : 1);
    s->bmass = (double *)calloc((size_t)(nbody ? nbody : 1), sizeof(double));
    s->bcom = (double *)calloc(nb3, sizeof(double));
    s->bvel = (double *)calloc(nb3, sizeof(double));
    s->mrg_com = (double *)calloc(nb3, sizeof(double));
    s->rg_imp = (double *)calloc(nb3, sizeof(double));
    s->rg_num = (int32_t *)calloc((size_t)(nbody ? nbody : 1), sizeof(int32_t));
    if (nbody) {
        memcpy(s->bmass, body_mass, sizeof(double) * (size_t)nbody);
        memcpy(s->bcom, body_com, sizeof(double) * 3 * (size_t)nbody);
        memcpy(s->mrg_com, body_com, sizeof(double) * 3 * (size_t)nbody);
    }
    s->nbody = nbody;
    return ORC_OK;
}

int orc_get_bodies(orc_solver *s, double *com, double *com_vel) {
    if (!s) return ORC_ERR_ARG;
    if (com && s->nbody) memcpy(com, s->bcom, sizeof(double) * 3 * (size_t)s->nbody);
    if (com_vel && s->nbody) memcpy(com_vel, s->bvel, sizeof(double) * 3 * (size_t)s->nbody);
    return ORC_OK;
}

int orc_set_avg_vel(orc_solver *s, const double *avg_vel) {
    if (!s || !s->x || !avg_vel) return ORC_ERR_ARG;
    memcpy(s->av, avg_vel, sizeof(double) * 3 * (size_t)s->nv);
    return ORC_OK;
}

/* test.cpp:192-258: x_old <- Coords; Coords <- x_old + dt*vel.  The
 * accumulator resets there are no-ops after a resolve (every pass clears
 * them in updateAverageVelocity). */
int orc_predict(orc_solver *s) {
    if (!s || !s->x) return ORC_ERR_ARG;
    double dt = s->P.dt;
    for (int64_t i = 0; i < 3 * (int64_t)s->nv; ++i) {
        s->x_old[i] = s->x[i];
        s->x[i] = s->x_old[i] + dt * s->vel[i];
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* elementary math selected by mode.  In libm mode the calls whose libm
 * result differs from the correctly rounded one are counted (a diagnostic:
 * results may only diverge between the two modes after such a call). */
static int64_t g_libm_differs = 0;
static inline double libm_checked(double v, double cr) {
    if (v != cr && !(v != v && cr != cr)) g_libm_differs++;
    return v;
}
static inline double m_acos(const orc_solver *s, double x) {
    return s->P.math_mode == ORC_MATH_LIBM ? libm_checked(acos(x), orc_det_acos(x)) : orc_det_acos(x);
}
static inline double m_cos(int mode, double x) {
    return mode == ORC_MATH_LIBM ? libm_checked(cos(x), orc_det_cos(x)) : orc_det_cos(x);
}
static inline double m_sin(int mode, double x) {
    return mode == ORC_MATH_LIBM ? libm_checked(sin(x), orc_det_sin(x)) : orc_det_sin(x);
}
static inline double m_cbrt(int mode, double x) {
    return mode == ORC_MATH_LIBM ? libm_checked(pow(x, 1.0 / 3.0), orc_det_pow_third(x))
                                 : orc_det_pow_third(x);
}
int64_t orc_libm_misrounded(int32_t reset) {
    int64_t n = g_libm_differs;
    if (reset) g_libm_differs = 0;
    return n;
}

/* ------------------------------------------------------------------------ */
/* per-vertex flags (dcollid.cpp:1069-1099): static (is_fixed), movable
 * (is_movableRG), rigid = either */
static inline int is_static(const uint8_t *fixed, int32_t v) { return fixed[v] != 0; }

/* impulse context: the accumulator arrays and parameters */
typedef struct {
    const orc_params *P;
    const uint8_t *fixed;
    const int32_t *vbody;  /* body of a movable vertex, -1 otherwise */
    const double *bmass;   /* total_mass per body */
    double *rg_imp;        /* collsnImpulse_RG per body */
    int32_t *rg_num;       /* collsn_num_RG per body */
    const double *av;      /* avgVel, indexed by vertex id */
    const double *x_old;
    double *imp, *fric;
    int32_t *num;
    int err;
} imp_ctx;

static inline int c_static(const imp_ctx *C, int32_t v) { return C->fixed[v] != 0; }
static inline int c_movable(const imp_ctx *C, int32_t v) { return C->vbody && C->vbody[v] >= 0; }
static inline int c_rigid(const imp_ctx *C, int32_t v) { return c_static(C, v) || c_movable(C, v); }

/* SpreadImpactZoneImpulse (dcollid.cpp:1101-1115): the zone of a movable
 * point is its body (createImpZone never merges movable points outside
 * createImpZoneForRG, dcollid.cpp:473-484), so every point of the body
 * receives impulse * nor and one more count */
static void spread_rg(imp_ctx *C, int32_t v, double impulse, const double nor[3]) {
    int32_t b = C->vbody[v];
    for (int i = 0; i < 3; ++i) C->rg_imp[3 * b + i] += impulse * nor[i];
    C->rg_num[b] += 1;
}

/* PointToTriImpulse, dcollid3d.cpp:925-1107 */
static void pt_impulse(imp_ctx *C, const int32_t p[4], double nor[3], double w[3],
                       double dist, double root) {
    (void)root;
    const orc_params *P = C->P;
    const double eps_m = P->mach_eps;
    double v_rel[3] = {0.0, 0.0, 0.0}, vn, vt;
    double impulse = 0.0, m_impulse = 0.0, sum_w = 0.0;
    double rigid_impulse[2] = {0.0, 0.0};
    double k = P->spring_k, m = P->mass, dt = P->dt, lambda = P->friction, h = P->thickness;
    double cr = P->restitution;
    dist = h - dist;
    for (int i = 0; i < 3; ++i) {
        v_rel[i] += C->av[3 * p[3] + i];
        for (int j = 0; j < 3; ++j) v_rel[i] -= w[j] * C->av[3 * p[j] + i];
    }
    vn = dot3(v_rel, nor);
    if (dot3(v_rel, v_rel) > vn * vn) vt = sqrt(dot3(v_rel, v_rel) - vn * vn);
    else vt = 0.0;
    int s0 = c_static(C, p[0]), s1 = c_static(C, p[1]), s2 = c_static(C, p[2]), s3 = c_static(C, p[3]);
    int m0 = c_movable(C, p[0]), m1 = c_movable(C, p[1]), m2 = c_movable(C, p[2]), m3 = c_movable(C, p[3]);
    if (vn < 0) {
        if (s3 || (s0 && s1 && s2)) {
            impulse = vn;
            rigid_impulse[0] = vn;
            rigid_impulse[1] = vn;
        } else if (m0 && m1 && m2 && m3) {
            double ma = C->bmass[C->vbody[p[0]]], mb = C->bmass[C->vbody[p[3]]];
            rigid_impulse[0] = vn * mb / (ma + mb);
            rigid_impulse[1] = vn * ma / (ma + mb);
        } else if (m0 && m1 && m2) {
            rigid_impulse[0] = 0.5 * vn;
            impulse = 0.5 * vn;
        } else if (m3) {
            impulse = 0.5 * vn;
            rigid_impulse[1] = 0.5 * vn;
        } else {
            impulse = vn * 0.5;
        }
        for (int i = 0; i < 3; ++i) {
            if (c_static(C, p[i])) w[i] = 0.0;
            sum_w += w[i];
        }
        if (fabs(sum_w) > eps_m) {
            double inv = 1.0 / sum_w;
            for (int i = 0; i < 3; ++i) w[i] = inv * w[i];
        }
    }
    int all_rigid = c_rigid(C, p[0]) && c_rigid(C, p[1]) && c_rigid(C, p[2]) && c_rigid(C, p[3]);
    if (vn * dt < 0.1 * dist) {
        if (all_rigid) {
            rigid_impulse[0] *= 1.0 + cr;
            rigid_impulse[1] *= 1.0 + cr;
        } else {
            double tmp = -smin(dt * k * dist / m, (0.1 * dist / dt - vn));
            impulse += tmp;
            rigid_impulse[0] += tmp;
            rigid_impulse[1] += tmp;
        }
    }
    if (fabs(sum_w) < eps_m) m_impulse = impulse;
    else m_impulse = 2.0 * impulse / (1.0 + dot3(w, w));
    if (all_rigid) {
        if (m0) spread_rg(C, p[0], rigid_impulse[0], nor);
        if (m3) spread_rg(C, p[3], -1.0 * rigid_impulse[1], nor);
        return;
    }
    for (int i = 0; i < 3; ++i) {
        if (c_static(C, p[i])) continue;
        double t_impulse = m_impulse;
        if (c_movable(C, p[i])) t_impulse = rigid_impulse[0];
        for (int j = 0; j < 3; ++j) {
            C->imp[3 * p[i] + j] += w[i] * t_impulse * nor[j];
            if (fabs(vt) > ROUND_EPS)
                C->fric[3 * p[i] + j] += smax(-fabs(lambda * w[i] * t_impulse / vt), -1.0)
                                         * (v_rel[j] - vn * nor[j]);
        }
        C->num[p[i]] += 1;
    }
    if (!s3) {
        double t_impulse = m_impulse;
        if (m3) t_impulse = rigid_impulse[1];
        for (int j = 0; j < 3; ++j) {
            C->imp[3 * p[3] + j] -= t_impulse * nor[j];
            if (fabs(vt) > ROUND_EPS)
                C->fric[3 * p[3] + j] += smax(-fabs(lambda * t_impulse / vt), -1.0)
                                         * (v_rel[j] - vn * nor[j]);
        }
        C->num[p[3]] += 1;
    }
    /* static points keep a zero impulse (dcollid3d.cpp:1080-1082) */
    for (int q = 0; q < 4; ++q)
        for (int j = 0; j < 3; ++j) {
            double v = C->imp[3 * p[q] + j];
            if (isnan(v) || isinf(v)) C->err = ORC_ERR_NONFINITE;
        }
}

/* EdgeToEdgeImpulse, dcollid3d.cpp:1109-1300 */
static void ee_impulse(imp_ctx *C, const int32_t p[4], double nor[3], double a, double b,
                       double dist, double root) {
    (void)root;
    const orc_params *P = C->P;
    const double eps_m = P->mach_eps;
    double v_rel[3], vn, vt, impulse = 0.0, m_impulse = 0.0;
    double rigid_impulse[2] = {0.0, 0.0};
    double k = P->spring_k, m = P->mass, dt = P->dt, lambda = P->friction, h = P->thickness;
    double cr = P->restitution;
    dist = h - dist;
    double wa[2] = {1.0 - a, a}, wb[2] = {1.0 - b, b};
    for (int j = 0; j < 3; ++j) {
        v_rel[j] = (1.0 - b) * C->av[3 * p[2] + j] + b * C->av[3 * p[3] + j];
        v_rel[j] -= (1.0 - a) * C->av[3 * p[0] + j] + a * C->av[3 * p[1] + j];
    }
    vn = dot3(v_rel, nor);
    if (dot3(v_rel, v_rel) > vn * vn) vt = sqrt(dot3(v_rel, v_rel) - vn * vn);
    else vt = 0.0;
    int s0 = c_static(C, p[0]), s1 = c_static(C, p[1]), s2 = c_static(C, p[2]), s3 = c_static(C, p[3]);
    int m0 = c_movable(C, p[0]), m1 = c_movable(C, p[1]), m2 = c_movable(C, p[2]), m3 = c_movable(C, p[3]);
    if (vn < 0.0) {
        if ((s0 && s1) || (s2 && s3)) {
            impulse = vn;
            rigid_impulse[0] = vn;
            rigid_impulse[1] = vn;
        } else if (m0 && m1 && m2 && m3) {
            double ma = C->bmass[C->vbody[p[0]]], mb = C->bmass[C->vbody[p[2]]];
            rigid_impulse[0] = vn * mb / (ma + mb);
            rigid_impulse[1] = vn * ma / (ma + mb);
        } else if (m0 && m1) {
            rigid_impulse[0] = 0.5 * vn;
            impulse = 0.5 * vn;
        } else if (m2 && m3) {
            impulse = 0.5 * vn;
            rigid_impulse[1] = 0.5 * vn;
        } else {
            impulse = vn * 0.5;
        }
        if (s0) wa[0] = 0.0;
        if (s1) wa[1] = 0.0;
        if (s2) wb[0] = 0.0;
        if (s3) wb[1] = 0.0;
    }
    int all_rigid = c_rigid(C, p[0]) && c_rigid(C, p[1]) && c_rigid(C, p[2]) && c_rigid(C, p[3]);
    if (vn * dt < 0.1 * dist) {
        if (all_rigid) {
            rigid_impulse[0] *= 1.0 + cr;
            rigid_impulse[1] *= 1.0 + cr;
        } else {
            double tmp = -smin(dt * k * dist / m, (0.1 * dist / dt - vn));
            impulse += tmp;
            rigid_impulse[0] += tmp;
            rigid_impulse[1] += tmp;
        }
    }
    if (wa[0] + wa[1] < eps_m || wb[0] + wb[1] < eps_m) m_impulse = impulse;
    else m_impulse = 2.0 * impulse / (wa[0] * wa[0] + wa[1] * wa[1] + wb[0] * wb[0] + wb[1] * wb[1]);
    if (all_rigid) {
        if (m0) spread_rg(C, p[0], rigid_impulse[0], nor);
        if (m2) spread_rg(C, p[2], -1.0 * rigid_impulse[1], nor);
        return;
    }
    const double wgt[4] = {wa[0], wa[1], wb[0], wb[1]};
    for (int j = 0; j < 3; ++j) {
        for (int q = 0; q < 4; ++q) {
            if (c_static(C, p[q])) continue;
            double t_impulse = m_impulse;
            if (c_movable(C, p[q])) t_impulse = rigid_impulse[q < 2 ? 0 : 1];
            if (q < 2) C->imp[3 * p[q] + j] += wgt[q] * t_impulse * nor[j];
            else C->imp[3 * p[q] + j] -= wgt[q] * t_impulse * nor[j];
            if (fabs(vt) > ROUND_EPS)
                C->fric[3 * p[q] + j] += smax(-fabs(lambda * wgt[q] * t_impulse / vt), -1.0)
                                         * (v_rel[j] - vn * nor[j]);
            if (j == 0) C->num[p[q]] += 1;
        }
    }
    for (int q = 0; q < 4; ++q)
        for (int j = 0; j < 3; ++j) {
            double v = C->imp[3 * p[q] + j];
            if (isnan(v) || isinf(v)) C->err = ORC_ERR_NONFINITE;
        }
}

/* PointToTri, dcollid3d.cpp:778-922.  c = Coords of pts (triangle 0..2,
 * point 3); returns 1 on a hit (impulse applied). */
static int point_to_tri(imp_ctx *C, const int32_t p[4], const double c[4][3], double h,
                        double root) {
    double w[3] = {0.0, 0.0, 0.0};
    double x13[3], x23[3], x43[3], nor[3] = {0.0, 0.0, 0.0}, nor_mag, dist, det;
    sub3(c[0], c[2], x13);
    sub3(c[1], c[2], x23);
    sub3(c[3], c[2], x43);
    det = dot3(x13, x13) * dot3(x23, x23) - dot3(x13, x23) * dot3(x13, x23);
    if (fabs(det) < 1000 * C->P->mach_eps) return 0;   /* degenerate triangle */
    cross3(x13, x23, nor);
    nor_mag = mag3(nor);
    double x43_old[3];
    sub3(&C->x_old[3 * p[3]], &C->x_old[3 * p[2]], x43_old);
    dist = dot3(x43_old, nor);
    for (int i = 0; i < 3; ++i) nor[i] /= nor_mag * ((dist >= 0) ? 1.0 : -1.0);
    dist = fabs(dot3(x43, nor));
    w[0] = (dot3(x13, x43) * dot3(x23, x23) - dot3(x23, x43) * dot3(x13, x23)) / det;
    w[1] = (dot3(x13, x13) * dot3(x23, x43) - dot3(x13, x23) * dot3(x13, x43)) / det;
    w[2] = 1 - w[0] - w[1];
    if (fabs(w[0]) < ROUND_EPS || fabs(w[1]) < ROUND_EPS || fabs(w[2]) < ROUND_EPS) {
        double vec[3];
        for (int j = 0; j < 3; ++j) vec[j] = C->x_old[3 * p[3] + j];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) vec[j] -= w[i] * C->x_old[3 * p[i] + j];
        if (mag3(vec) > ROUND_EPS)
            for (int j = 0; j < 3; ++j) nor[j] = vec[j];
    }
    nor_mag = mag3(nor);
    if (nor_mag > ROUND_EPS) {
        for (int i = 0; i < 3; ++i) nor[i] /= nor_mag;
    } else {
        C->err = ORC_ERR_NONFINITE;              /* "nan nor vec" */
        return 0;
    }
    if (dist > h) return 0;
    double eps = C->P->round_eps;
    for (int i = 0; i < 3; ++i)
        if (w[i] > 1 + eps || w[i] < -eps) return 0;
    pt_impulse(C, p, nor, w, dist, root);
    return 1;
}

/* EdgeToEdge, dcollid3d.cpp:643-776 (edges p0p1 and p2p3) */
static int edge_to_edge(imp_ctx *C, const int32_t p[4], const double c[4][3], double h,
                        double root) {
    double x21[3], x43[3], x31[3], a, b, tmp[3], v1[3], v2[3], nor[3], nor_mag, dist;
    sub3(c[1], c[0], x21);
    sub3(c[3], c[2], x43);
    sub3(c[2], c[0], x31);
    cross3(x21, x43, tmp);
    if (mag3(tmp) < ROUND_EPS) return 0;          /* parallel edges ignored */
    a = (dot3(x43, x43) * dot3(x21, x31) - dot3(x21, x43) * dot3(x43, x31)) /
        (dot3(x21, x21) * dot3(x43, x43) - dot3(x21, x43) * dot3(x21, x43));
    b = (dot3(x21, x43) * dot3(x21, x31) - dot3(x21, x21) * dot3(x43, x31)) /
        (dot3(x21, x21) * dot3(x43, x43) - dot3(x21, x43) * dot3(x21, x43));
    a = smax(smin(a, 1.0), 0.0);
    b = smax(smin(b, 1.0), 0.0);
    for (int i = 0; i < 3; ++i) { v1[i] = a * x21[i]; v2[i] = b * x43[i]; }
    for (int i = 0; i < 3; ++i) { v1[i] = c[0][i] + v1[i]; v2[i] = c[2][i] + v2[i]; }
    sub3(v2, v1, nor);
    nor_mag = mag3(nor);
    if (nor_mag < 1000 * C->P->mach_eps) {
        const double *xo = C->x_old;
        for (int j = 0; j < 3; ++j) {
            nor[j] = (1.0 - b) * xo[3 * p[2] + j] + b * xo[3 * p[3] + j];
            nor[j] -= (1.0 - a) * xo[3 * p[0] + j] + a * xo[3 * p[1] + j];
        }
    }
    dist = dist3(v1, v2);
    if (dist > h) return 0;
    nor_mag = mag3(nor);
    if (nor_mag < C->P->mach_eps) {
        C->err = ORC_ERR_NONFINITE;              /* "Normal vector is nan" */
        return 0;
    }
    for (int i = 0; i < 3; ++i) nor[i] /= nor_mag;
    ee_impulse(C, p, nor, a, b, dist, root);
    return 1;
}

/* isCoplanar, dcollid3d.cpp:371-482 */
static int is_coplanar(int math_mode, double eps_m, const double xo[4][3],
                       const double vv4[4][3], double dt, double roots[4]) {
    double v[4][3], x[4][3];
    for (int k = 0; k < 3; ++k) { v[0][k] = vv4[0][k]; x[0][k] = xo[0][k]; }
    for (int i = 1; i < 4; ++i)
        for (int k = 0; k < 3; ++k) {
            v[i][k] = vv4[i][k] - vv4[0][k];
            x[i][k] = xo[i][k] - xo[0][k];
        }
    double a, b, c, d, vv[3], vx[3], xx[3];
    vv[0] = v[1][1] * v[2][2] - v[1][2] * v[2][1];
    vv[1] = v[1][0] * v[2][2] - v[1][2] * v[2][0];
    vv[2] = v[1][0] * v[2][1] - v[1][1] * v[2][0];
    vx[0] = v[1][1] * x[2][2] - v[1][2] * x[2][1] - v[2][1] * x[1][2] + v[2][2] * x[1][1];
    vx[1] = v[1][0] * x[2][2] - v[1][2] * x[2][0] - v[2][0] * x[1][2] + v[2][2] * x[1][0];
    vx[2] = v[1][0] * x[2][1] - v[1][1] * x[2][0] - v[2][0] * x[1][1] + v[2][1] * x[1][0];
    xx[0] = x[1][1] * x[2][2] - x[1][2] * x[2][1];
    xx[1] = x[1][0] * x[2][2] - x[1][2] * x[2][0];
    xx[2] = x[1][0] * x[2][1] - x[1][1] * x[2][0];
    a = v[3][0] * vv[0] - v[3][1] * vv[1] + v[3][2] * vv[2];
    b = x[3][0] * vv[0] - x[3][1] * vv[1] + x[3][2] * vv[2] +
        v[3][0] * vx[0] - v[3][1] * vx[1] + v[3][2] * vx[2];
    c = x[3][0] * vx[0] - x[3][1] * vx[1] + x[3][2] * vx[2] +
        v[3][0] * xx[0] - v[3][1] * xx[1] + v[3][2] * xx[2];
    d = x[3][0] * xx[0] - x[3][1] * xx[1] + x[3][2] * xx[2];
    roots[0] = roots[1] = roots[2] = -1.0;
    roots[3] = dt;
    if (fabs(a) > eps_m) {
        b /= a; c /= a; d /= a;
        a = b; b = c; c = d;
        double Q = (a * a - 3 * b) / 9;
        double R = (2 * a * a * a - 9 * a * b + 27 * c) / 54;
        double Q3 = Q * Q * Q, R2 = R * R;
        if (R2 < Q3) {
            double Qsqrt = sqrt(Q);
            double theta;
            theta = (math_mode == ORC_MATH_LIBM) ? libm_checked(acos(R / sqrt(Q3)), orc_det_acos(R / sqrt(Q3)))
                                                  : orc_det_acos(R / sqrt(Q3));
            roots[0] = -2 * Qsqrt * m_cos(math_mode, theta / 3) - a / 3;
            roots[1] = -2 * Qsqrt * m_cos(math_mode, (theta + 2 * M_PI) / 3) - a / 3;
            roots[2] = -2 * Qsqrt * m_cos(math_mode, (theta - 2 * M_PI) / 3) - a / 3;
        } else {
            double A, B;
            double sgn = (R > 0) ? 1.0 : -1.0;
            A = -sgn * m_cbrt(math_mode, fabs(R) + sqrt(R2 - Q3));
            B = (fabs(A) < ROUND_EPS) ? 0.0 : Q / A;
            roots[0] = (A + B) - a / 3.0;
            if (fabs(A - B) < ROUND_EPS) roots[1] = roots[2] = -0.5 * (A + B) - a / 3.0;
        }
    } else {
        a = b; b = c; c = d;
        double delta = b * b - 4.0 * a * c;
        if (fabs(a) > ROUND_EPS && delta > 0) {
            double delta_sqrt = sqrt(delta);
            roots[0] = (-b + delta_sqrt) / (2.0 * a);
            roots[1] = (-b - delta_sqrt) / (2.0 * a);
        } else if (fabs(a) < ROUND_EPS && fabs(b) > ROUND_EPS) {
            roots[0] = -c / b;
        }
    }
    for (int i = 0; i < 3; ++i) {
        roots[i] = roots[i] - eps_m;
        if (roots[i] < 0 || roots[i] > dt) roots[i] = -1;
    }
    double t;
    if (roots[0] > roots[1]) { t = roots[0]; roots[0] = roots[1]; roots[1] = t; }
    if (roots[0] > roots[2]) { t = roots[0]; roots[0] = roots[2]; roots[2] = t; }
    if (roots[1] > roots[2]) { t = roots[1]; roots[1] = roots[2]; roots[2] = t; }
    return (roots[0] > eps_m || roots[1] > eps_m || roots[2] > eps_m) ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* union-find with member lists, dcollid.cpp:1015-1059 */
static int32_t uf_find(orc_solver *s, int32_t p) {
    int32_t r = p;
    while (s->uf_root[r] != r) r = s->uf_root[r];
    while (s->uf_root[p] != r) { int32_t n = s->uf_root[p]; s->uf_root[p] = r; p = n; }
    return r;
}
static void uf_merge(orc_solver *s, int32_t X, int32_t Y) {
    int32_t PX = uf_find(s, X), PY = uf_find(s, Y);
    if (PX == PY) return;
    if (s->uf_w[PX] > s->uf_w[PY]) {
        s->uf_w[PX] += s->uf_w[PY];
        s->uf_root[PY] = PX;
        s->uf_next[s->uf_tail[PX]] = PY;
        s->uf_tail[PX] = s->uf_tail[PY];
    } else {
        s->uf_w[PY] += s->uf_w[PX];
        s->uf_root[PX] = PY;
        s->uf_next[s->uf_tail[PY]] = PX;
        s->uf_tail[PY] = s->uf_tail[PX];
    }
}
/* createImpZone(pts, num, first), dcollid.cpp:473-484: outside the body
 * zones built at assembly (first), pairs with a movable point are skipped */
static inline int is_movable(const orc_solver *s, int32_t v) { return s->nbody && s->vbody[v] >= 0; }
static void create_imp_zone_n(orc_solver *s, const int32_t *p, int num, int first) {
    for (int i = 0; i < num; ++i)
        for (int j = 0; j < i; ++j) {
            if (!first && (is_movable(s, p[i]) || is_movable(s, p[j]))) continue;
            uf_merge(s, p[i], p[j]);
        }
}
static void create_imp_zone(orc_solver *s, const int32_t p[4]) { create_imp_zone_n(s, p, 4, 0); }

/* makeSet (dcollid.cpp:1015-1031) and createImpZoneForRG
 * (dcollid3d.cpp:54-68): one zone per movable body, merged triangle by
 * triangle in element order */
static void make_set(orc_solver *s) {
    for (int32_t v = 0; v < s->nv; ++v) {
        s->uf_root[v] = v; s->uf_next[v] = -1; s->uf_tail[v] = v; s->uf_w[v] = 1;
    }
    if (!s->nbody) return;
    for (int32_t e = 0; e < s->nt; ++e)
        if (is_movable(s, s->ev[3 * e])) create_imp_zone_n(s, &s->ev[3 * e], 3, 1);
}

/* ------------------------------------------------------------------------ */
/* one sub-test: proximity (static coords = x_old) or CCD (MovingPointToTri /
 * MovingEdgeToEdge, dcollid3d.cpp:327-369) */
static int subtest(orc_solver *s, imp_ctx *C, int is_pt, const int32_t p[4], int ccd, double h) {
    double c[4][3];
    if (!ccd) {
        for (int q = 0; q < 4; ++q)
            for (int k = 0; k < 3; ++k) c[q][k] = s->x_old[3 * p[q] + k];
        return is_pt ? point_to_tri(C, p, (const double(*)[3])c, h, 0.0)
                     : edge_to_edge(C, p, (const double(*)[3])c, h, 0.0);
    }
    double xo[4][3], v4[4][3], roots[4];
    for (int q = 0; q < 4; ++q)
        for (int k = 0; k < 3; ++k) { xo[q][k] = s->x_old[3 * p[q] + k]; v4[q][k] = s->av[3 * p[q] + k]; }
    if (!is_coplanar(s->P.math_mode, s->P.mach_eps, (const double(*)[3])xo, (const double(*)[3])v4,
                     s->P.dt, roots))
        return 0;
    for (int i = 0; i < 4; ++i) {
        if (roots[i] < 0) continue;
        for (int q = 0; q < 4; ++q)
            for (int k = 0; k < 3; ++k) c[q][k] = xo[q][k] + roots[i] * v4[q][k];
        int hit = is_pt ? point_to_tri(C, p, (const double(*)[3])c, h, roots[i])
                        : edge_to_edge(C, p, (const double(*)[3])c, h, roots[i]);
        if (hit) return 1;
    }
    return 0;
}

/* isRigidBody(const CD_HSE*) (dcollid.cpp:1069-1099): a static or movable point */
static int elem_is_rigid(const orc_solver *s, int32_t e) {
    int np = elem_is_tri(s, e) ? 3 : 2;
    for (int i = 0; i < np; ++i) {
        int32_t v = s->ev[3 * e + i];
        if (s->fixed[v] || is_movable(s, v)) return 1;
    }
    return 0;
}

/* isProximity / isCollision dispatch (dcollid.cpp:753-836) with
 * TriToTri / TriToBond / BondToBond (dcollid3d.cpp:485-627) and the Moving*
 * variants (dcollid3d.cpp:203-325).  Returns hit mask (bit per sub-test);
 * *ran = 1 if sub-tests were evaluated. */
static int32_t narrow_pair(orc_solver *s, imp_ctx *C, int32_t A, int32_t B, int ccd, int *ran) {
    const int32_t *ev = s->ev;
    double h = ccd ? s->P.round_eps : s->P.thickness;
    int ta = elem_is_tri(s, A), tb = elem_is_tri(s, B);
    int32_t mask = 0, status = 0, p[4];
    int zone = ccd && s->zone_mode;
    *ran = 0;
#define SUBTEST(idx, ispt)                                             \
    do {                                                                \
        if (subtest(s, C, (ispt), p, ccd, h)) { status = 1; mask |= 1 << (idx); } \
        if (status && zone) create_imp_zone(s, p);                     \
    } while (0)
    if (ta && tb) {
        if (s->surf[A] == s->surf[B] && elem_is_rigid(s, A)) return 0;
        const int32_t *t1 = &ev[3 * A], *t2 = &ev[3 * B];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                if (t1[i] == t2[j]) return 0;
        *ran = 1;
        for (int k = 0; k < 2; ++k)
            for (int i = 0; i < 3; ++i) {
                /* proximity: triangle = other, point = this (TriToTri);
                 * CCD: triangle = this, point = other (MovingTriToTri) */
                const int32_t *tri_this = (k == 0) ? t1 : t2;
                const int32_t *tri_other = (k == 0) ? t2 : t1;
                const int32_t *tri = ccd ? tri_this : tri_other;
                const int32_t *pt = ccd ? tri_other : tri_this;
                p[0] = tri[0]; p[1] = tri[1]; p[2] = tri[2]; p[3] = pt[i];
                SUBTEST(3 * k + i, 1);
            }
        for (int i = 0; i < 3; ++i) {
            p[0] = t1[i]; p[1] = t1[(i + 1) % 3];
            for (int j = 0; j < 3; ++j) {
                p[2] = t2[j]; p[3] = t2[(j + 1) % 3];
                SUBTEST(6 + 3 * i + j, 0);
            }
        }
    } else if (!ta && !tb) {
        p[0] = ev[3 * A]; p[1] = ev[3 * A + 1]; p[2] = ev[3 * B]; p[3] = ev[3 * B + 1];
        for (int i = 0; i < 4; ++i)
            for (int j = i + 1; j < 4; ++j)
                if (p[i] == p[j]) return 0;
        *ran = 1;
        SUBTEST(0, 0);
    } else {
        int32_t T = ta ? A : B, Bd = ta ? B : A;
        const int32_t *t = &ev[3 * T];
        int32_t bs = ev[3 * Bd], be = ev[3 * Bd + 1];
        for (int i = 0; i < 3; ++i)
            if (t[i] == bs || t[i] == be) return 0;
        *ran = 1;
        p[0] = t[0]; p[1] = t[1]; p[2] = t[2];
        p[3] = bs; SUBTEST(0, 1);
        p[0] = t[0]; p[1] = t[1]; p[2] = t[2];
        p[3] = be; SUBTEST(1, 1);
        p[2] = bs; p[3] = be;
        for (int i = 0; i < 3; ++i) {
            p[0] = t[i]; p[1] = t[(i + 1) % 3];
            SUBTEST(2 + i, 0);
        }
    }
#undef SUBTEST
    return mask;
}

/* ------------------------------------------------------------------------ */
/* element boxes (dcollid.cpp:852-932, AABB.cpp:6-36): static boxes use x_old
 * for triangles and Coords for bonds; swept boxes cover x_old and
 * x_old + avgVel*dt; all inflated by 1e-6 */
static void element_boxes(orc_solver *s, int moving) {
    double dt = s->P.dt;
    for (int32_t e = 0; e < s->ne; ++e) {
        int np = elem_is_tri(s, e) ? 3 : 2;
        int tri = np == 3;
        box_t *bx = &s->ebox[e];
        for (int d = 0; d < 3; ++d) {
            double lo = HUGE_VAL, hi = -HUGE_VAL;
            for (int i = 0; i < np; ++i) {
                int32_t v = s->ev[3 * e + i];
                if (moving) {
                    double xo = s->x_old[3 * v + d];
                    double xn = s->x_old[3 * v + d] + s->av[3 * v + d] * dt;
                    hi = smax(hi, xo); hi = smax(hi, xn);
                    lo = smin(lo, xo); lo = smin(lo, xn);
                } else {
                    double xo = tri ? s->x_old[3 * v + d] : s->x[3 * v + d];
                    hi = smax(xo, hi);
                    lo = smin(xo, lo);
                }
            }
            bx->lo[d] = lo - BOX_EPS;
            bx->hi[d] = hi + BOX_EPS;
        }
    }
}

static inline int box_overlap(const box_t *a, const box_t *b) {
    return (a->lo[0] <= b->hi[0] && a->hi[0] >= b->lo[0]) &&
           (a->lo[1] <= b->hi[1] && a->hi[1] >= b->lo[1]) &&
           (a->lo[2] <= b->hi[2] && a->hi[2] >= b->lo[2]);
}
static inline double box_volume(const box_t *b) {
    return (b->hi[0] - b->lo[0]) * (b->hi[1] - b->lo[1]) * (b->hi[2] - b->lo[2]);
}
static inline box_t box_merge(const box_t *a, const box_t *b) {
    box_t r;
    for (int i = 0; i < 3; ++i) { r.lo[i] = smin(a->lo[i], b->lo[i]); r.hi[i] = smax(a->hi[i], b->hi[i]); }
    return r;
}

/* ---- incremental AABB tree (AABB.cpp:81-245) ---- */
static int32_t tree_new_node(tree_t *t) {
    if (t->n == t->cap) {
        int32_t nc = t->cap ? t->cap * 2 : 1024;
        t->left = (int32_t *)realloc(t->left, sizeof(int32_t) * nc);
        t->right = (int32_t *)realloc(t->right, sizeof(int32_t) * nc);
        t->parent = (int32_t *)realloc(t->parent, sizeof(int32_t) * nc);
        t->leaf = (int32_t *)realloc(t->leaf, sizeof(int32_t) * nc);
        t->has_box = (uint8_t *)realloc(t->has_box, nc);
        t->box = (box_t *)realloc(t->box, sizeof(box_t) * nc);
        t->cap = nc;
    }
    int32_t i = t->n++;
    t->left[i] = t->right[i] = t->parent[i] = -1;
    t->leaf[i] = -1;
    t->has_box[i] = 0;
    return i;
}
static void node_update_aabb(tree_t *t, const box_t *ebox, int32_t i) {
    if (t->left[i] < 0 && t->right[i] < 0) {
        t->box[i] = ebox[t->leaf[i]];
        t->has_box[i] = 1;
    } else if (!t->has_box[i]) {
        t->box[i] = box_merge(&t->box[t->left[i]], &t->box[t->right[i]]);
        t->has_box[i] = 1;
    } else {
        for (int d = 0; d < 3; ++d) {
            t->box[i].lo[d] = smin(t->box[t->left[i]].lo[d], t->box[t->right[i]].lo[d]);
            t->box[i].hi[d] = smax(t->box[t->left[i]].hi[d], t->box[t->right[i]].hi[d]);
        }
    }
}
/* insertNode(n, parent) where *slot is the reference being descended
 * (root or a child pointer), AABB.cpp:163-197; iterative with an explicit
 * path so the bottom-up updates run in the same order */
static void tree_insert(tree_t *t, const box_t *ebox, int32_t n, int32_t *path) {
    int32_t depth = 0;
    int32_t p = t->root;
    for (;;) {
        if (t->left[p] < 0 && t->right[p] < 0) {
            int32_t np_ = tree_new_node(t);
            int32_t pp = t->parent[p];
            t->parent[np_] = pp;
            if (pp >= 0) {
                if (t->left[pp] == p) t->left[pp] = np_;
                else t->right[pp] = np_;
            } else {
                t->root = np_;
            }
            t->left[np_] = n; t->right[np_] = p;
            t->parent[n] = np_; t->parent[p] = np_;
            path[depth++] = np_;
            break;
        }
        path[depth++] = p;
        const box_t *bl = &t->box[t->left[p]], *br = &t->box[t->right[p]];
        box_t ml = box_merge(bl, &t->box[n]), mr = box_merge(br, &t->box[n]);
        double vdiff1 = box_volume(&ml) - box_volume(bl);
        double vdiff2 = box_volume(&mr) - box_volume(br);
        p = (vdiff1 < vdiff2) ? t->left[p] : t->right[p];
    }
    for (int32_t k = depth - 1; k >= 0; --k) node_update_aabb(t, ebox, path[k]);
}

static int32_t *g_path_buf = NULL;
static int32_t g_path_cap = 0;

static int32_t *path_buf(int32_t need) {
    if (need > g_path_cap) {
        g_path_buf = (int32_t *)realloc(g_path_buf, sizeof(int32_t) * (size_t)need);
        g_path_cap = need;
    }
    return g_path_buf;
}

/* addAABB for every element in order (AABB.cpp:129-146), and
 * updateTreeStructure (AABB.cpp:149-161), which re-inserts the leaves in the
 * same order with their current boxes: both are "insert leaves 0..ne-1 one by
 * one into an initially empty tree".  Leaves take node ids 0..ne-1; the old
 * internal nodes are dropped (the reference leaks them). */
static void tree_build(tree_t *t, const box_t *ebox, int32_t ne) {
    if (!t->leaf_node) t->leaf_node = (int32_t *)malloc(sizeof(int32_t) * (size_t)(ne ? ne : 1));
    t->n = 0;
    t->root = -1;
    for (int32_t e = 0; e < ne; ++e) {
        int32_t nd = tree_new_node(t);
        t->leaf[nd] = e;
        t->leaf_node[e] = nd;
        node_update_aabb(t, ebox, nd);
    }
    int32_t *path = path_buf(ne + 8);
    for (int32_t e = 0; e < ne; ++e) {
        if (t->root < 0) t->root = e;
        else tree_insert(t, ebox, e, path);
    }
    t->built = 1;
}

/* updateAABBTree (AABB.cpp:212-245): refresh leaf boxes then branches
 * bottom-up (any child-before-parent order gives the same boxes) */
static void tree_refit(tree_t *t, const box_t *ebox) {
    int32_t *st = (int32_t *)malloc(sizeof(int32_t) * (size_t)(t->n + 1));
    int32_t *order = (int32_t *)malloc(sizeof(int32_t) * (size_t)(t->n + 1));
    int32_t sp = 0, no = 0;
    st[sp++] = t->root;
    while (sp) {
        int32_t i = st[--sp];
        order[no++] = i;
        if (t->left[i] >= 0) st[sp++] = t->left[i];
        if (t->right[i] >= 0) st[sp++] = t->right[i];
    }
    for (int32_t k = no - 1; k >= 0; --k) {
        int32_t i = order[k];
        if (t->left[i] < 0 && t->right[i] < 0) t->box[i] = ebox[t->leaf[i]];
        else {
            for (int d = 0; d < 3; ++d) {
                t->box[i].lo[d] = smin(t->box[t->left[i]].lo[d], t->box[t->right[i]].lo[d]);
                t->box[i].hi[d] = smax(t->box[t->left[i]].hi[d], t->box[t->right[i]].hi[d]);
            }
        }
    }
    free(st); free(order);
}

/* query (AABB.cpp:254-343): leaves in in-order; for each, a preorder walk
 * reporting overlapping, not-yet-queried leaves.  Appends (a=cur, b=n). */
static void tree_query(orc_solver *s, tree_t *t) {
    memset(s->visited, 0, (size_t)(s->ne ? s->ne : 1));
    int32_t *ist = (int32_t *)malloc(sizeof(int32_t) * (size_t)(t->n + 1));
    int32_t *qst = s->stack;
    int32_t isp = 0, cur = t->root;
    while (cur >= 0 || isp) {
        while (cur >= 0) { ist[isp++] = cur; cur = t->left[cur]; }
        cur = ist[--isp];
        if (t->left[cur] < 0 && t->right[cur] < 0) {
            int32_t n = cur;
            const box_t *nb = &t->box[n];
            int32_t qsp = 0, c2 = t->root;
            for (;;) {
                while (c2 >= 0) {
                    if (box_overlap(&t->box[c2], nb)) {
                        if (t->left[c2] < 0 && t->right[c2] < 0 && c2 != n &&
                            !s->visited[t->leaf[c2]]) {
                            int64_t pr = ((int64_t)t->leaf[c2] << 32) | (uint32_t)t->leaf[n];
                            vec_push(&s->pairs, &pr);
                        }
                        qst[qsp++] = c2;
                        c2 = t->left[c2];
                    } else break;
                }
                if (!qsp) break;
                c2 = qst[--qsp];
                c2 = t->right[c2];
            }
            s->visited[t->leaf[n]] = 1;
        }
        cur = t->right[cur];
    }
    free(ist);
}

static int cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* one broad + narrow pass; returns 1 if any pair hit */
static int run_pass(orc_solver *s, int kind) {
    int ccd = kind != ORC_PASS_PROXIMITY;
    tree_t *t = ccd ? &s->coll : &s->prox;
    element_boxes(s, ccd);
    if (s->ne == 0) return 0;
    if (!t->built) {                                 /* dcollid.cpp:368-376 */
        tree_build(t, s->ebox, s->ne);
        s->volume = box_volume(&t->box[t->root]);
    } else {                                         /* dcollid.cpp:377-385 */
        tree_refit(t, s->ebox);
        double vol = box_volume(&t->box[t->root]);
        if ((vol - s->volume) > 0.2 * s->volume) {
            tree_build(t, s->ebox, s->ne);
            s->volume = box_volume(&t->box[t->root]);
        }
    }
    s->pairs.n = 0;
    tree_query(s, t);
    int64_t *pp = (int64_t *)s->pairs.p;
    int64_t np = s->pairs.n;
    if (s->P.order_mode == ORC_ORDER_CANONICAL) {
        for (int64_t i = 0; i < np; ++i) {
            int32_t a = (int32_t)(pp[i] >> 32), b = (int32_t)(pp[i] & 0xffffffff);
            int32_t lo = a < b ? a : b, hi = a < b ? b : a;
            pp[i] = ((int64_t)lo << 32) | (uint32_t)hi;
        }
        qsort(pp, (size_t)np, sizeof(int64_t), cmp_i64);
    }
    imp_ctx C = {&s->P, s->fixed, s->nbody ? s->vbody : NULL, s->bmass, s->rg_imp, s->rg_num,
                 s->av, s->x_old, s->imp, s->fric, s->num, 0};
    orc_pass_info *pi = &s->cur;
    memset(pi, 0, sizeof(*pi));
    pi->kind = kind;
    pi->candidates = np;
    int any = 0;
    for (int64_t i = 0; i < np; ++i) {
        int32_t A, B;
        if (s->P.order_mode == ORC_ORDER_CANONICAL) {
            B = (int32_t)(pp[i] >> 32);              /* lo = query leaf */
            A = (int32_t)(pp[i] & 0xffffffff);       /* hi */
        } else {
            A = (int32_t)(pp[i] >> 32);
            B = (int32_t)(pp[i] & 0xffffffff);
        }
        int ran = 0;
        int32_t mask = narrow_pair(s, &C, A, B, ccd, &ran);
        pi->narrow_pairs += ran;
        if (mask) {
            any = 1;
            pi->hit_pairs++;
            pi->hit_subtests += __builtin_popcount((unsigned)mask);
            if (s->P.log_pairs) {
                int32_t row[4] = {s->cur_pass, A < B ? A : B, A < B ? B : A,
                                  mask | ((A < B) ? (1 << 30) : 0)};
                vec_push(&s->plog, row);
            }
        }
    }
    if (C.err) s->err = C.err;
    return any;
}

static int32_t update_impact_zone_velocity(orc_solver *s, int rg_only);

/* updateAverageVelocity (dcollid.cpp:677-751).  The RG counts are per body
 * (every point of a body holds the same count), so they are cleared after
 * all points have used them; collsnImpulse_RG is never cleared (the
 * reference resets only collsn_num_RG, dcollid.cpp:728-734). */
static void record_impulses(orc_solver *s, int any_hit) {
    if (!s->P.record_impulses) return;
    if (!s->last_imp) {
        s->last_imp = (double *)calloc(3 * (size_t)s->nv, sizeof(double));
        s->last_num = (int32_t *)calloc((size_t)s->nv, sizeof(int32_t));
    }
    if (any_hit < 0) {            /* start of a resolve */
        memset(s->last_imp, 0, 3 * (size_t)s->nv * sizeof(double));
        memset(s->last_num, 0, (size_t)s->nv * sizeof(int32_t));
    } else if (any_hit) {         /* a pass with hits */
        memcpy(s->last_imp, s->imp, 3 * (size_t)s->nv * sizeof(double));
        memcpy(s->last_num, s->num, (size_t)s->nv * sizeof(int32_t));
    }
}

static void update_average_velocity(orc_solver *s) {
    for (int32_t v = 0; v < s->nv; ++v) {
        if (s->fixed[v]) continue;
        if (s->num[v] > 0) {
            s->has[v] = 1;
            for (int k = 0; k < 3; ++k) {
                s->av[3 * v + k] += (s->imp[3 * v + k] + s->fric[3 * v + k]) / s->num[v];
                if (isinf(s->av[3 * v + k]) || isnan(s->av[3 * v + k])) s->err = ORC_ERR_NONFINITE;
                s->imp[3 * v + k] = s->fric[3 * v + k] = 0.0;
            }
            s->num[v] = 0;
        }
        int32_t b = s->nbody ? s->vbody[v] : -1;
        if (b >= 0 && s->rg_num[b] > 0) {
            s->has[v] = 1;
            for (int k = 0; k < 3; ++k) s->av[3 * v + k] += s->rg_imp[3 * b + k] / s->rg_num[b];
        }
    }
    for (int32_t b = 0; b < s->nbody; ++b) s->rg_num[b] = 0;
    /* updateImpactZoneVelocityForRG (dcollid.cpp:267-288) */
    if (s->nbody && s->P.dt > 0.0) update_impact_zone_velocity(s, 1);
}

/* Sum of per-member terms in one of two orders:
 *   lanes == 0  : sequential over members (the reference's linked-list walk)
 *   lanes == 64 : canonical order -- lane l sums members l, l+64, l+128, ...
 *                 from 0.0, then the 64 lane sums are added in lane order.
 * The canonical order is the one a 64-wide wavefront computes per zone. */
typedef void (*term_fn)(const void *ctx, int32_t p, double *out, int nout);

static void ordered_sum(int lanes, int32_t n, const int32_t *mem, term_fn f, const void *ctx,
                        int nout, double *total) {
    double t[9];
    for (int k = 0; k < nout; ++k) total[k] = 0.0;
    if (lanes == 0) {
        for (int32_t q = 0; q < n; ++q) {
            f(ctx, mem[q], t, nout);
            for (int k = 0; k < nout; ++k) total[k] = t[k] + total[k];
        }
        return;
    }
    for (int l = 0; l < lanes; ++l) {
        double part[9];
        for (int k = 0; k < nout; ++k) part[k] = 0.0;
        for (int32_t q = l; q < n; q += lanes) {
            f(ctx, mem[q], t, nout);
            for (int k = 0; k < nout; ++k) part[k] = t[k] + part[k];
        }
        for (int k = 0; k < nout; ++k) total[k] = total[k] + part[k];
    }
}

typedef struct {
    const double *x_old, *av;
    double x_cm[3], v_cm[3], m;
} zone_ctx;

static void term_pos_vel(const void *c, int32_t p, double *out, int nout) {
    const zone_ctx *z = (const zone_ctx *)c;
    (void)nout;
    for (int i = 0; i < 3; ++i) { out[i] = z->x_old[3 * p + i]; out[3 + i] = z->av[3 * p + i]; }
}
static void term_ang_mom(const void *c, int32_t p, double *out, int nout) {
    const zone_ctx *z = (const zone_ctx *)c;
    double dx[3], dv[3], Li[3];
    (void)nout;
    sub3(&z->x_old[3 * p], z->x_cm, dx);
    sub3(&z->av[3 * p], z->v_cm, dv);
    cross3(dx, dv, Li);
    for (int i = 0; i < 3; ++i) out[i] = z->m * Li[i];
}
static void term_inertia(const void *c, int32_t p, double *out, int nout) {
    const zone_ctx *z = (const zone_ctx *)c;
    double dx[3], mag_dx, tmp;
    (void)nout;
    sub3(&z->x_old[3 * p], z->x_cm, dx);
    mag_dx = mag3(dx);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            tmp = -dx[i] * dx[j];
            if (i == j) tmp += mag_dx * mag_dx;
            out[3 * i + j] = tmp * z->m;
        }
}

/* updateImpactListVelocity (dcollid3d.cpp:70-200) on members[0..n); the
 * sums x_cm, v_cm, L = sum m dx x dv, I = sum m (|dx|^2 delta - dx dx^T) are
 * accumulated in the order selected by `lanes` (see ordered_sum) */
static int zone_rigid(const orc_params *P, int32_t n, const int32_t *mem, const double *x_old,
                      double *av, const uint8_t *fixed, int lanes) {
    double m = P->mass, dt = P->dt;
    int mode = P->math_mode;
    zone_ctx z;
    z.x_old = x_old; z.av = av; z.m = m;
    double sums[9], L[3], I[3][3], tmp[3][3];
    int32_t num_pts = n;
    ordered_sum(lanes, n, mem, term_pos_vel, &z, 6, sums);
    for (int i = 0; i < 3; ++i) { z.x_cm[i] = sums[i] / num_pts; z.v_cm[i] = sums[3 + i] / num_pts; }
    ordered_sum(lanes, n, mem, term_ang_mom, &z, 3, sums);
    for (int i = 0; i < 3; ++i) L[i] = sums[i];
    ordered_sum(lanes, n, mem, term_inertia, &z, 9, sums);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) I[i][j] = sums[3 * i + j];
    const double *x_cm = z.x_cm, *v_cm = z.v_cm;
    double w[3], mag_w;
    for (int i = 0; i < 3; ++i) {
        memcpy(tmp, I, sizeof(tmp));
        for (int j = 0; j < 3; ++j) tmp[j][i] = L[j];
        if (det3((const double(*)[3])I) < ROUND_EPS) w[i] = 0.0;
        else w[i] = det3((const double(*)[3])tmp) / det3((const double(*)[3])I);
    }
    mag_w = mag3(w);
    int err = 0;
    for (int32_t q = 0; q < n; ++q) {
        int32_t p = mem[q];
        if (fixed[p]) continue;
        double x_new[3], dx[3], xF[3], xR[3], wxR[3], tmpV[3];
        sub3(&x_old[3 * p], x_cm, dx);
        if (mag_w < ROUND_EPS) {
            for (int i = 0; i < 3; ++i) { xF[i] = dx[i]; wxR[i] = 0.0; }
            sub3(dx, xF, xR);
        } else {
            double sc = dot3(dx, w) / dot3(w, w);
            for (int i = 0; i < 3; ++i) xF[i] = sc * w[i];
            sub3(dx, xF, xR);
            double sw = m_sin(mode, dt * mag_w) / mag_w;
            for (int i = 0; i < 3; ++i) tmpV[i] = sw * w[i];
            cross3(tmpV, xR, wxR);
        }
        for (int i = 0; i < 3; ++i) {
            x_new[i] = x_cm[i] + dt * v_cm[i] + xF[i] + m_cos(mode, dt * mag_w) * xR[i] + wxR[i];
            av[3 * p + i] = (x_new[i] - x_old[3 * p + i]) / dt;
            if (isnan(av[3 * p + i])) err = 1;
        }
    }
    return err;
}

/* updateImpactZoneVelocity (dcollid.cpp:290-310), every zone of two or more
 * points; with rg_only updateImpactZoneVelocityForRG (dcollid.cpp:267-288),
 * the zones holding a movable point (the bodies).  Returns #zones. */
static int32_t update_impact_zone_velocity(orc_solver *s, int rg_only) {
    int32_t nz = 0;
    int32_t *mem = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->nv);
    if (s->P.order_mode == ORC_ORDER_REFERENCE_TREE) {
        memset(s->sorted, 0, (size_t)s->nv);
        for (int32_t e = 0; e < s->ne; ++e) {
            int np = elem_is_tri(s, e) ? 3 : 2;
            for (int i = 0; i < np; ++i) {
                int32_t pt = s->ev[3 * e + i];
                if (s->sorted[pt] || s->uf_w[uf_find(s, pt)] == 1) continue;
                if (rg_only && !is_movable(s, pt)) { s->sorted[pt] = 1; continue; }
                int32_t head = uf_find(s, pt), n = 0;
                for (int32_t q = head; q >= 0; q = s->uf_next[q]) { mem[n++] = q; s->sorted[q] = 1; }
                if (zone_rigid(&s->P, n, mem, s->x_old, s->av, s->fixed, 0)) s->err = ORC_ERR_NONFINITE;
                nz++;
            }
        }
    } else {
        /* canonical: zones in order of their smallest vertex, members in
         * ascending vertex id, sums in the 64-lane order of ordered_sum */
        int32_t *root_of = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->nv);
        int32_t *cnt = (int32_t *)calloc((size_t)s->nv + 1, sizeof(int32_t));
        int32_t *start = (int32_t *)malloc(sizeof(int32_t) * ((size_t)s->nv + 1));
        int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * ((size_t)s->nv + 1));
        for (int32_t v = 0; v < s->nv; ++v) { root_of[v] = uf_find(s, v); cnt[root_of[v]]++; }
        start[0] = 0;
        for (int32_t v = 0; v < s->nv; ++v) start[v + 1] = start[v] + cnt[v];
        memcpy(fill, start, sizeof(int32_t) * ((size_t)s->nv + 1));
        for (int32_t v = 0; v < s->nv; ++v) mem[fill[root_of[v]]++] = v;
        uint8_t *done = (uint8_t *)calloc((size_t)s->nv, 1);
        for (int32_t v = 0; v < s->nv; ++v) {
            int32_t r = root_of[v];
            if (done[r] || cnt[r] < 2) continue;
            if (rg_only && !is_movable(s, v)) continue;
            done[r] = 1;
            if (zone_rigid(&s->P, cnt[r], &mem[start[r]], s->x_old, s->av, s->fixed, 64))
                s->err = ORC_ERR_NONFINITE;
            nz++;
        }
        free(root_of); free(cnt); free(start); free(fill); free(done);
    }
    free(mem);
    return nz;
}

/* detectDomainBoundaryCollision (dcollid.cpp:116-158): every point is
 * visited once per element holding it, in element order (movable points
 * are skipped).  A later visit can change a point again: when x_old lies
 * outside the box on one axis, friction from an earlier visit can shorten
 * the inward move so that the axis is touched on the next visit. */
static void domain_visit(orc_solver *s, int32_t v) {
    double dt = s->P.dt, mu = s->P.friction;
    double *avg = &s->av[3 * v];
    double dv = 0;
    for (int j = 0; j < 3; ++j) {
        double cand = s->x_old[3 * v + j] + dt * avg[j];
        double L = s->P.dom_lo[j], U = s->P.dom_hi[j];
        if (cand <= L) {
            s->has[v] = 1; dv = fabs(avg[j]); avg[j] = 0.0;
            s->x[3 * v + j] = L + s->P.thickness;
        } else if (cand >= U) {
            s->has[v] = 1; dv = fabs(avg[j]); avg[j] = 0.0;
            s->x[3 * v + j] = U - s->P.thickness;
        }
    }
    double preVt = mag3(avg);
    if (preVt > s->P.mach_eps)
        for (int j = 0; j < 3; ++j) avg[j] *= smax(1.0 - mu * dv / preVt, 0.0);
}

static void domain_boundary(orc_solver *s) {
    for (int32_t e = 0; e < s->ne; ++e) {
        int np = elem_is_tri(s, e) ? 3 : 2;
        for (int i = 0; i < np; ++i) {
            int32_t v = s->ev[3 * e + i];
            if (is_movable(s, v)) continue;
            domain_visit(s, v);
        }
    }
}

/* updateFinalForRG (dcollid.cpp:626-675): the first point of each body (in
 * element order) with has_collsn sets center_of_mass_velo to its avgVel and
 * center_of_mass to avgVel * dt + mrg_com; mrg_com then follows
 * center_of_mass.  The reference reads mrg_com of a body before its first
 * write (an empty vector); here it starts at the center of mass given by
 * orc_set_bodies. */
static void update_final_for_rg(orc_solver *s) {
    if (!s->nbody) return;
    double dt = s->P.dt;
    uint8_t *in_mrg = (uint8_t *)calloc((size_t)s->nbody, 1);
    int8_t *visited = (int8_t *)malloc((size_t)s->nbody);   /* -1: no entry */
    memset(visited, -1, (size_t)s->nbody);
    for (int32_t e = 0; e < s->ne; ++e) {
        int np = elem_is_tri(s, e) ? 3 : 2;
        for (int i = 0; i < np; ++i) {
            int32_t v = s->ev[3 * e + i];
            if (!is_movable(s, v)) continue;
            int32_t b = s->vbody[v];
            if (s->has[v] && !in_mrg[b]) {
                in_mrg[b] = 1;
                for (int j = 0; j < 3; ++j) {
                    s->bvel[3 * b + j] = s->av[3 * v + j];
                    s->bcom[3 * b + j] = s->av[3 * v + j] * dt + s->mrg_com[3 * b + j];
                }
                visited[b] = 0;
            }
            if (visited[b] <= 0) {
                for (int j = 0; j < 3; ++j) s->mrg_com[3 * b + j] = s->bcom[3 * b + j];
                visited[b] = 1;
            }
        }
    }
    free(in_mrg);
    free(visited);
}

/* reduceSuperelastOnce (dcollid.cpp:486-560) */
static int reduce_superelast_once(orc_solver *s, int32_t *num_edges, int32_t *degenerate) {
    double dt = s->P.dt;
    int has = 0;
    *num_edges = 0;
    for (int32_t e = 0; e < s->ne; ++e) {
        int np = elem_is_tri(s, e) ? 3 : 2;
        if (elem_is_rigid(s, e)) continue;
        for (int j = 0; j < ((np == 2) ? 1 : np); ++j) {
            int32_t p0 = s->ev[3 * e + (j % np)], p1 = s->ev[3 * e + ((j + 1) % np)];
            double xc[2][3];
            for (int k = 0; k < 3; ++k) {
                xc[0][k] = s->x_old[3 * p0 + k] + dt * s->av[3 * p0 + k];
                xc[1][k] = s->x_old[3 * p1 + k] + dt * s->av[3 * p1 + k];
            }
            double len_new = dist3(xc[0], xc[1]);
            double len_old = dist3(&s->x_old[3 * p0], &s->x_old[3 * p1]);
            double len0 = s->len0[3 * e + j];
            int avg = 0;
            if (len_old > ROUND_EPS && len_new > ROUND_EPS) {
                double strain_rate = (len_new - len_old) / len_old;
                double strain = (len_new - len0) / len0;
                if (fabs(strain) > SUPERELAS_TOL || fabs(strain_rate) > SUPERELAS_TOL) avg = 1;
            } else {
                avg = 1;
                (*degenerate)++;
            }
            if (avg) {
                double vt[3];
                for (int k = 0; k < 3; ++k) vt[k] = s->av[3 * p0 + k] + s->av[3 * p1 + k];
                for (int k = 0; k < 3; ++k) vt[k] = 0.5 * vt[k];
                for (int k = 0; k < 3; ++k) { s->av[3 * p0 + k] = vt[k]; s->av[3 * p1 + k] = vt[k]; }
                (*num_edges)++;
                has = 1;
            }
        }
    }
    return has;
}

static void end_pass(orc_solver *s, int kind, int index, orc_report *rep) {
    s->cur.kind = kind;
    s->cur.index = index;
    vec_push(&s->passes, &s->cur);
    rep->n_passes++;
    rep->candidates += s->cur.candidates;
    rep->narrow_pairs += s->cur.narrow_pairs;
    rep->hit_pairs += s->cur.hit_pairs;
    rep->hit_subtests += s->cur.hit_subtests;
    s->cur_pass++;
}

int orc_resolve(orc_solver *s, orc_report *rep_out) {
    orc_report rep;
    memset(&rep, 0, sizeof(rep));
    if (!s || !s->x) return ORC_ERR_ARG;
    s->err = 0;
    s->passes.n = 0;
    s->plog.n = 0;
    s->cur_pass = 0;
    double dt = s->P.dt;
    /* per-step makeSet and body zones of assembleFromInterface
     * (dcollid3d.cpp:44-45) */
    make_set(s);
    /* computeAverageVelocity (dcollid.cpp:160-220) */
    for (int64_t i = 0; i < 3 * (int64_t)s->nv; ++i) {
        if (dt > ROUND_EPS) s->av[i] = (s->x[i] - s->x_old[i]) / dt;
        else s->av[i] = 0.0;
        if (isnan(s->av[i]) || isinf(s->av[i])) s->err = ORC_ERR_NONFINITE;
    }
    if (s->err) { rep.status = s->err; if (rep_out) *rep_out = rep; return s->err; }
    memcpy(s->x, s->x_old, sizeof(double) * 3 * (size_t)s->nv);
    /* detectProximity (dcollid.cpp:390-406) */
    s->zone_mode = 0;
    record_impulses(s, -1);
    record_impulses(s, run_pass(s, ORC_PASS_PROXIMITY));
    update_average_velocity(s);
    end_pass(s, ORC_PASS_PROXIMITY, 0, &rep);
    /* detectCollision (dcollid.cpp:430-468) */
    int is_collision = 1, niter = 1, cd_count = 0;
    rep.has_collision = 0;
    while (is_collision && !s->err) {
        is_collision = run_pass(s, ORC_PASS_CCD);
        record_impulses(s, is_collision);
        if (cd_count++ == 0 && is_collision) rep.has_collision = 1;
        update_average_velocity(s);
        end_pass(s, ORC_PASS_CCD, niter - 1, &rep);
        rep.n_ccd_passes++;
        if (++niter > s->P.max_ccd_passes) break;
    }
    /* computeImpactZone (dcollid.cpp:227-265) */
    if (is_collision && !s->err) {
        s->zone_mode = 1;
        int zi = 0;
        while (is_collision && !s->err) {
            if (zi >= s->P.max_zone_passes) { rep.status = ORC_WARN_ZONE_CAP; break; }
            is_collision = run_pass(s, ORC_PASS_ZONE);
            record_impulses(s, is_collision);
            update_average_velocity(s);
            rep.n_zones = update_impact_zone_velocity(s, 0);
            end_pass(s, ORC_PASS_ZONE, zi, &rep);
            rep.n_zone_passes++;
            zi++;
        }
        s->zone_mode = 0;
    }
    /* detectDomainBoundaryCollision, updateFinalPosition, reduceSuperelast,
     * updateFinalVelocity (dcollid.cpp:347-361, 562-624) */
    domain_boundary(s);
    for (int64_t i = 0; i < 3 * (int64_t)s->nv; ++i) s->x[i] = s->x_old[i] + s->av[i] * dt;
    int has = 1, it = 0;
    int32_t nedges = 0, degenerate = 0;
    while (has && it++ < s->P.max_strain_sweeps) {
        has = reduce_superelast_once(s, &nedges, &degenerate);
        rep.strain_sweeps++;
    }
    rep.strain_edges = nedges;
    rep.degenerate_edges = degenerate;
    for (int32_t v = 0; v < s->nv; ++v)
        if (s->has[v])
            for (int k = 0; k < 3; ++k) s->vel[3 * v + k] = s->av[3 * v + k];
    update_final_for_rg(s);
    if (s->err) rep.status = s->err;
    if (rep_out) *rep_out = rep;
    return rep.status;
}

int32_t orc_pass_count(orc_solver *s) { return s ? (int32_t)s->passes.n : 0; }
int orc_pass_infos(orc_solver *s, orc_pass_info *out, int32_t cap) {
    if (!s || !out) return ORC_ERR_ARG;
    int64_t n = s->passes.n < cap ? s->passes.n : cap;
    memcpy(out, s->passes.p, (size_t)n * sizeof(orc_pass_info));
    return ORC_OK;
}
int64_t orc_pair_log_count(orc_solver *s) { return s ? s->plog.n : 0; }
int orc_pair_log(orc_solver *s, int32_t *out4, int64_t cap) {
    if (!s || !out4) return ORC_ERR_ARG;
    int64_t n = s->plog.n < cap ? s->plog.n : cap;
    memcpy(out4, s->plog.p, (size_t)n * 4 * sizeof(int32_t));
    return ORC_OK;
}
int orc_zone_labels(orc_solver *s, int32_t *out) {
    if (!s || !out) return ORC_ERR_ARG;
    int32_t *minid = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->nv);
    for (int32_t v = 0; v < s->nv; ++v) minid[v] = INT32_MAX;
    for (int32_t v = 0; v < s->nv; ++v) {
        int32_t r = uf_find(s, v);
        if (v < minid[r]) minid[r] = v;
    }
    for (int32_t v = 0; v < s->nv; ++v) out[v] = minid[uf_find(s, v)];
    free(minid);
    return ORC_OK;
}

int orc_last_impulses(orc_solver *s, double *imp, int32_t *num) {
    if (!s || !s->last_imp) return ORC_ERR_ARG;
    if (imp) memcpy(imp, s->last_imp, 3 * (size_t)s->nv * sizeof(double));
    if (num) memcpy(num, s->last_num, (size_t)s->nv * sizeof(int32_t));
    return ORC_OK;
}

/* ---- test hooks ---- */
int32_t orc_debug_pair(orc_solver *s, int32_t A, int32_t B, int32_t ccd, int32_t zone_mode) {
    if (!s || !s->x || A < 0 || B < 0 || A >= s->ne || B >= s->ne) return -1;
    imp_ctx C = {&s->P, s->fixed, s->nbody ? s->vbody : NULL, s->bmass, s->rg_imp, s->rg_num,
                 s->av, s->x_old, s->imp, s->fric, s->num, 0};
    int saved = s->zone_mode, ran = 0;
    s->zone_mode = zone_mode;
    int32_t mask = narrow_pair(s, &C, A, B, ccd, &ran);
    s->zone_mode = saved;
    if (C.err) return -1;
    return mask;
}

int orc_debug_make_set(orc_solver *s) {
    if (!s || !s->x) return ORC_ERR_ARG;
    make_set(s);
    return ORC_OK;
}

/* ---- unit entry points ---- */
int orc_is_coplanar(const double *x_old, const double *avg_vel, double dt, int32_t math_mode,
                    double *roots) {
    return is_coplanar(math_mode, DBL_EPSILON, (const double(*)[3])x_old,
                       (const double(*)[3])avg_vel, dt, roots);
}

static int unit_subtest(const orc_params *p, const double *c, const double *x_old,
                        const double *avg_vel, const uint8_t *is_fixed, double h, double *imp,
                        double *fric, int32_t *counts, int pt) {
    /* vertices 0..3 of a private 4-vertex state */
    memset(imp, 0, 12 * sizeof(double));
    memset(fric, 0, 12 * sizeof(double));
    memset(counts, 0, 4 * sizeof(int32_t));
    uint8_t fx[4] = {0, 0, 0, 0};
    if (is_fixed) memcpy(fx, is_fixed, 4);
    imp_ctx C = {p, fx, NULL, NULL, NULL, NULL, avg_vel, x_old, imp, fric, counts, 0};
    int32_t ids[4] = {0, 1, 2, 3};
    int hit = pt ? point_to_tri(&C, ids, (const double(*)[3])c, h, 0.0)
                 : edge_to_edge(&C, ids, (const double(*)[3])c, h, 0.0);
    if (C.err) return -1;
    return hit;
}
int orc_point_to_tri(const orc_params *p, const double *c, const double *x_old,
                     const double *avg_vel, const uint8_t *is_fixed, double h, double *imp,
                     double *fric, int32_t *counts) {
    return unit_subtest(p, c, x_old, avg_vel, is_fixed, h, imp, fric, counts, 1);
}
int orc_edge_to_edge(const orc_params *p, const double *c, const double *x_old,
                     const double *avg_vel, const uint8_t *is_fixed, double h, double *imp,
                     double *fric, int32_t *counts) {
    return unit_subtest(p, c, x_old, avg_vel, is_fixed, h, imp, fric, counts, 0);
}
int orc_zone_rigid_update(const orc_params *p, int32_t n, const double *x_old, double *avg_vel,
                          const uint8_t *is_fixed, int32_t lanes) {
    int32_t *mem = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    for (int32_t i = 0; i < n; ++i) mem[i] = i;
    int err = zone_rigid(p, n, mem, x_old, avg_vel, is_fixed, lanes);
    free(mem);
    return err ? ORC_ERR_NONFINITE : ORC_OK;
}
double orc_math_acos(double x) { return orc_det_acos(x); }
double orc_math_cos(double x) { return orc_det_cos(x); }
double orc_math_sin(double x) { return orc_det_sin(x); }
double orc_math_pow_third(double x) { return orc_det_pow_third(x); }

int orc_math_eval(int32_t fn, const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; ++i) {
        switch (fn) {
            case 0: y[i] = orc_det_acos(x[i]); break;
            case 1: y[i] = orc_det_cos(x[i]); break;
            case 2: y[i] = orc_det_sin(x[i]); break;
            case 3: y[i] = orc_det_pow_third(x[i]); break;
            case 10: y[i] = acos(x[i]); break;
            case 11: y[i] = cos(x[i]); break;
            case 12: y[i] = sin(x[i]); break;
            case 13: y[i] = pow(x[i], 1.0 / 3.0); break;
            default: return ORC_ERR_ARG;
        }
    }
    return ORC_OK;
}
