/*
 * oracle_impl.h -- CPU restatement of the reference's octree leaf update.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle_common.h).  This file is included by
 * one translation unit per (model, dimension) variant, which defines
 *   OD      spatial dimension (2 or 3)
 *   OMODEL  O_EULER, O_NAVIER_STOKES or O_MULTI
 *   OPFX(n) the exported-symbol prefix
 *
 * Pins (tests/test_oracle_*.py): every function below has a closed-form,
 * brute-force or invariant test -- model functions against the textbook
 * equations (Jacobian rows, HLLC states, EOS limits, limiters, viscous
 * stress), split / unsplit / LTS reassignment / boundary flux through the
 * t_* hooks at the end of this file, the flux ownership through the face
 * time coverage (every face covered for exactly the global step, except the
 * coarse/fine faces of a coarse leaf on a finer time level, which the
 * reference loses), neighbours and 2:1 flags by brute-force search, and the
 * schedule, slopes and five-equation sources by invariants of the scheme.
 * The reference ships no golden vectors, so nothing here is pinned against
 * reference output.
 *
 * Every function names the reference file:line it restates.  Arithmetic is
 * written in the reference's evaluation order and built with
 * -ffp-contract=off so that the HIP path, written to the same order, agrees
 * bit for bit.  Two deliberate restatements (both order-of-evaluation only):
 *
 *  1. Flux accumulation.  The reference scatters each face flux into both
 *     cells with `omp atomic` adds (impl/flux.hpp:82-91, core/linalg.hpp:441),
 *     so its sums have no fixed order.  Here a cell's flux total is summed
 *     face by face in a fixed order, and fluxes owed to a cell that is not
 *     updated in the current substep are kept in that cell's per-face
 *     accumulator acc[face] (summed in time order, then sub-face order).
 *     The set of flux terms, their values, and the ownership rule
 *     (impl/flux.hpp:58-59) are the reference's.
 *  2. Split slopes.  The reference recomputes the slope of each cell just
 *     before splitting it, inside a serial loop, so a cell can see a
 *     neighbour that was split earlier in the same loop
 *     (impl/adaptive_refinement.hpp:154-171).  The interpolated value of such
 *     a neighbour equals its pre-split value up to rounding; here all split
 *     slopes are evaluated on the pre-split tree.
 *
 * Cluster membership (acti_clusters / acti_buffers,
 * integrator/integrator.hpp:46-47) is kept per node (cluster, buffer) and a
 * level's cluster is visited in node-index order; every per-cell phase is
 * order independent, so this changes no value.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#include "oracle_common.h"

#define NCH (1 << OD)
#define NFACE (2 * OD)
#if OMODEL == O_EULER
#define NE (OD + 2)
#define NF NE
#define NP 0
#define NG 0
#define IU 1
#define IP (NE - 1)
#elif OMODEL == O_NAVIER_STOKES
#define NE (OD + 2)
#define NF NE
#define NP 0
#define NG (OD + 1)
#define IU 1
#define IP (NE - 1)
#define IDT OD
#else
#define NE (OD + 4)
#define NF (NE + OD + 1)
#define NP (OD + 4)
#define NG (OD + 3)
#define IU 2
#define IP (NE - 2)
#define IA (NE - 1)
#define IKN 2
#define IKAPPA (2 + OD)
#define IALPHAS (3 + OD)
#define IDT OD
#define IDA (OD + 1)
#endif
#define NPX (NP > 0 ? NP : 1)
#define NGX (NG > 0 ? NG : 1)

typedef struct {
    double size, t;
    double center[OD];
    double u[NE], p[NE], rhs[NE];
    double k[NPX], t_slope[NGX];
    double slope[OD][NE];
    double gradient[OD][NGX];
} o_cell;

typedef struct {
    o_cell cell;
    double acc[NFACE][NF]; /* received flux per face (replaces node.flux) */
    /* test instrumentation: time coverage of each face in the current global
     * step (sum of t1 - t0 over the fluxes applied to the face, weighted by
     * the evaluated face's share of this cell's face area), and the same for
     * the received fluxes still held in acc */
    double cov[NFACE], acc_cov[NFACE];
    int32_t parent, child_base;
    int32_t nbr[NFACE];
    uint8_t bc[NFACE]; /* 0 = none, 1 + domain face index, or O_BC_OBJECT (record brec) */
    int32_t brec;      /* object boundary record of the O_BC_OBJECT faces */
    int8_t acti, tree_level, cluster, buffer;
    uint8_t leaf, active, garbage, can_split, can_unsplit, should_split, is_boundary;
    double error, creation_time, max_dt, cfl;
} o_node;

typedef struct {
    o_params P;
    o_node *tree;
    int64_t n, cap;
    double global_dt;
    int acti_counter[O_MAX_LEVEL];
    double acti_time[O_MAX_LEVEL];
    int max_acti_level;
    int64_t n_substeps, n_global_steps;
    int64_t cell_updates;
    int64_t counts[O_MAX_LEVEL]; /* cluster sizes */
    int64_t acc_violations; /* received flux found where the read rule says none can be */
    int no_gc;              /* test switch: never compact (the result must not change) */
    /* scratch */
    double *tot;     /* NF per node, flux total of the current substep */
    int32_t *list;   /* current cluster list */
    double *tmp;     /* smoothing scratch */
    uint8_t *mark;   /* refinement marks */
    /* object boundary conditions (scene/scene.cpp:11-28), one record per
     * boundary node: type, unit wall normal, inflow state, fade-in */
    struct o_rec {
        int32_t type, pad;
        double normal[3];
        double p[8];
        double fade_in;
    } *recs;
    int64_t n_recs;
} o_state;

/* ------------------------------------------------------------------ */
/* std::min / std::max / std::clamp semantics (argument order matters) */
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline int iclamp(int v, int lo, int hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

/* ------------------------------------------------------------------ */
/* Model: equation of state and conversions                           */
/* ------------------------------------------------------------------ */
#if OMODEL == O_MULTI
/* model/multi.hpp:367-387 */
static double mgamma(const o_params *P, double a) {
    if (a < 1e-20) return P->gamma2;
    else if (a > 1 - 1e-20) return P->gamma1;
    else return 1 + 1 / (a / (P->gamma1 - 1) + (1 - a) / (P->gamma2 - 1));
}
static double mp0(const o_params *P, double a) {
    if (a < 1e-20) return P->P2;
    else if (a > 1 - 1e-20) return P->P1;
    else
        return (a * P->gamma1 * P->P1 / (P->gamma1 - 1) + (1 - a) * P->gamma2 * P->P2 / (P->gamma2 - 1)) /
               (1 + a / (P->gamma1 - 1) + (1 - a) / (P->gamma2 - 1));
}
/* model/multi.hpp:394-401 */
static double mtemperature(const o_params *P, double r0, double r1, double p, double a) {
    if (a < 1e-20) return (p + P->P2) / ((P->gamma2 - 1) * r1 * P->cv2);
    else if (a > 1 - 1e-20) return (p + P->P1) / ((P->gamma1 - 1) * r0 * P->cv1);
    else
        return (p / (mgamma(P, a) - 1) + a * P->P1 / (P->gamma1 - 1) + (1 - a) * P->P2 / (P->gamma2 - 1)) /
               (r0 * P->cv1 + r1 * P->cv2);
}
/* model/multi.hpp:428-430 */
static double msound_speed_2(const o_params *P, double r0, double r1, double p, double a) {
    return mgamma(P, a) * (p + mp0(P, a)) / (r0 + r1);
}
static double msound_speed(const o_params *P, double r0, double r1, double p, double a) {
    return sqrt(mgamma(P, a) * (p + mp0(P, a)) / (r0 + r1));
}
#endif

/* sound speed of a primitive state: euler.hpp:185-187, multi.hpp:420-422 */
static double sound_speed_prim(const o_params *P, const double *p) {
#if OMODEL == O_MULTI
    return sqrt(mgamma(P, p[IA]) * (p[IP] + mp0(P, p[IA])) / (p[0] + p[1]));
#else
    return sqrt(P->gamma * p[IP] / p[0]);
#endif
}

/* supersonic(): euler.hpp:197-203; multi.hpp:432-438 passes p[0] where a
 * two-component density is expected, i.e. r = {p[0], 0}. */
static int supersonic(const o_params *P, const double *p) {
    double u2 = 0;
    for (int i = 0; i < OD; ++i) u2 += p[IU + i] * p[IU + i];
#if OMODEL == O_MULTI
    return u2 > msound_speed_2(P, p[0], 0.0, p[IP], p[IA]);
#else
    return u2 > P->gamma * p[IP] / p[0];
#endif
}

/* primitive_variables / conserved_variables: euler.hpp:177-197,213-233;
 * multi.hpp:407-413,448-472 */
static void prim_from_cons(const o_params *P, const double *u, double *p) {
#if OMODEL == O_MULTI
    p[0] = u[0];
    p[1] = u[1];
    for (int i = 0; i < OD; ++i) p[IU + i] = u[IU + i] / (u[0] + u[1]);
    double ke = 0;
    for (int i = 0; i < OD; ++i) ke += 0.5 * u[IU + i] * u[IU + i];
    p[IP] = (mgamma(P, u[IA]) - 1) * (u[IP] - ke / (u[0] + u[1])) - mgamma(P, u[IA]) * mp0(P, u[IA]);
    p[IA] = u[IA];
#else
    p[0] = u[0];
    for (int i = 0; i < OD; ++i) p[IU + i] = u[IU + i] / u[0];
    double ke = 0;
    for (int i = 0; i < OD; ++i) ke += 0.5 * u[IU + i] * u[IU + i];
    p[IP] = (P->gamma - 1) * (u[IP] - ke / u[0]);
#endif
}

static void cons_from_prim(const o_params *P, const double *p, double *u) {
#if OMODEL == O_MULTI
    double ke = 0;
    u[0] = p[0];
    u[1] = p[1];
    for (int i = 0; i < OD; ++i) {
        u[IU + i] = (p[0] + p[1]) * p[IU + i];
        ke += 0.5 * p[IU + i] * p[IU + i];
    }
    u[IP] = (p[IP] + mgamma(P, p[IA]) * mp0(P, p[IA])) / (mgamma(P, p[IA]) - 1) + (p[0] + p[1]) * ke;
    u[IA] = p[IA];
#else
    double ke = 0;
    u[0] = p[0];
    for (int i = 0; i < OD; ++i) {
        u[IU + i] = p[0] * p[IU + i];
        ke += 0.5 * p[IU + i] * p[IU + i];
    }
    u[IP] = p[IP] / (P->gamma - 1) + p[0] * ke;
#endif
}

/* physical flux F(p, dim): euler.hpp:50-64; multi.hpp:78-96 */
static void flux_F(const o_params *P, const double *p, int dim, double *F) {
    int in = IU + dim;
    double ke = 0.0;
    for (int i = 0; i < NF; ++i) F[i] = 0.0;
#if OMODEL == O_MULTI
    for (int i = 0; i < OD; ++i) {
        F[IU + i] = (p[0] + p[1]) * p[in] * p[IU + i];
        ke += 0.5 * p[IU + i] * p[IU + i];
    }
    F[0] = p[in] * p[0];
    F[1] = p[in] * p[1];
    F[in] += p[IP];
    F[IP] = p[in] * (mgamma(P, p[IA]) / (mgamma(P, p[IA]) - 1) * (p[IP] + mp0(P, p[IA])) + ke * (p[0] + p[1]));
    F[IA] = p[in] * p[IA];
    F[IA + 1] = p[in];
    F[IA + 2 + dim] = p[IA];
#else
    for (int i = 0; i < OD; ++i) {
        F[IU + i] = p[0] * p[in] * p[IU + i];
        ke += 0.5 * p[IU + i] * p[IU + i];
    }
    F[0] = p[in] * p[0];
    F[in] += p[IP];
    F[IP] = p[in] * (P->gamma / (P->gamma - 1) * p[IP] + ke * p[0]);
#endif
}

/* projected flux F(p, u_projected, dim): euler.hpp:67-83; multi.hpp:101-123 */
static void flux_F_projected(const o_params *P, const double *p, double up, int dim, double *F) {
    int in = IU + dim;
    double ke = 0.0;
    for (int i = 0; i < NF; ++i) F[i] = 0.0;
#if OMODEL == O_MULTI
    for (int i = 0; i < OD; ++i) {
        F[IU + i] = (p[0] + p[1]) * up * p[IU + i];
        ke += 0.5 * p[IU + i] * p[IU + i];
    }
    F[0] = up * p[0];
    F[1] = up * p[1];
    F[in] += p[IP];
    F[IP] = up * (mgamma(P, p[IA]) / (mgamma(P, p[IA]) - 1) * (p[IP] + mp0(P, p[IA])) + ke * (p[0] + p[1]));
    F[IA] = up * p[IA];
    F[IA + 1] = up;
    F[IA + 2 + dim] = p[IA];
#else
    for (int i = 0; i < OD; ++i) {
        F[IU + i] = p[0] * up * p[IU + i];
        ke += 0.5 * p[IU + i] * p[IU + i];
    }
    F[0] = up * p[0];
    F[in] += p[IP];
    F[IP] = up * (P->gamma / (P->gamma - 1) * p[IP] + ke * p[0]);
#endif
}

/* quasi-linear Jacobian A(p, dim): euler.hpp:86-103; multi.hpp:126-145 */
static void jacobian_A(const o_params *P, const double *p, double ka, int dim, double A[NE][NE]) {
    int in = IU + dim;
    for (int i = 0; i < NE; ++i)
        for (int j = 0; j < NE; ++j) A[i][j] = (i == j) ? p[in] : 0.0;
#if OMODEL == O_MULTI
    A[0][in] = p[0];
    A[1][in] = p[1];
    A[IP][in] = (p[0] + p[1]) * msound_speed_2(P, p[0], p[1], p[IP], p[IA]);
    A[in][IP] = 1.0 / (p[0] + p[1]);
    A[in][IA] = P->sigma * ka / (p[0] + p[1]);
#else
    (void) ka;
    A[0][in] = p[0];
    A[IP][in] = p[0] * (P->gamma * p[IP] / p[0]);
    A[in][IP] = 1.0 / p[0];
#endif
}

/* compute_rhs: euler.hpp:105-116; multi.hpp:158-169 */
static void compute_rhs(const o_params *P, o_cell *c) {
    double A[NE][NE];
    for (int i = 0; i < NE; ++i) c->rhs[i] = 0.0;
#if OMODEL == O_MULTI
    double ka = c->k[IKAPPA];
#else
    double ka = 0.0;
#endif
    for (int i = 0; i < NE; ++i)
        for (int j = 0; j < OD; ++j) {
            jacobian_A(P, c->p, ka, j, A);
            for (int k = 0; k < NE; ++k) c->rhs[i] += A[i][k] * c->slope[j][k];
        }
}

/* HLLC-type star state in primitives: euler.hpp:118-153; multi.hpp:205-246 */
static void middle_state(const o_params *P, const double *pn, const double *pp, double ka, int dim, double *pm) {
    int in = IU + dim;
    for (int i = 0; i < NE; ++i) pm[i] = 0.0;
#if OMODEL == O_MULTI
    double c_n = msound_speed(P, pn[0], pn[1], pn[IP], pn[IA]);
    double c_p = msound_speed(P, pp[0], pp[1], pp[IP], pp[IA]);
    double u = 0.5 * (pn[in] + pp[in]);
    double c = 0.5 * (c_n + c_p);
    double Sn = smin(u - c, pn[in] - c_n);
    double Sp = smax(u + c, pp[in] + c_p);
    double Sm = (pp[IP] - pn[IP] + (pn[0] + pn[1]) * pn[in] * (Sn - pn[in]) - (pp[0] + pp[1]) * pp[in] * (Sp - pp[in]) -
                 P->sigma * ka * (pp[IA] - pn[IA])) /
                ((pn[0] + pn[1]) * (Sn - pn[in]) - (pp[0] + pp[1]) * (Sp - pp[in]));
    if (Sn > 0) {
        for (int i = 0; i < NE; ++i) pm[i] = pn[i];
    } else if (Sp < 0) {
        for (int i = 0; i < NE; ++i) pm[i] = pp[i];
    } else if (Sm > 0 && Sn < 0) {
        pm[0] = pn[0] * (Sn - pn[in]) / (Sn - Sm);
        pm[1] = pn[1] * (Sn - pn[in]) / (Sn - Sm);
        pm[IP] = pn[IP] + (pn[0] + pn[1]) * (Sn - pn[in]) * (Sm - pn[in]);
        pm[IA] = pn[IA];
        for (int i = 0; i < OD; ++i)
            if (i != dim) pm[IU + i] = pn[IU + i];
        pm[in] = Sm;
    } else {
        pm[0] = pp[0] * (Sp - pp[in]) / (Sp - Sm);
        pm[1] = pp[1] * (Sp - pp[in]) / (Sp - Sm);
        pm[IP] = pp[IP] + (pp[0] + pp[1]) * (Sp - pp[in]) * (Sm - pp[in]);
        pm[IA] = pp[IA];
        for (int i = 0; i < OD; ++i)
            if (i != dim) pm[IU + i] = pp[IU + i];
        pm[in] = Sm;
    }
#else
    (void) ka;
    double c_n = sqrt(P->gamma * pn[IP] / pn[0]);
    double c_p = sqrt(P->gamma * pp[IP] / pp[0]);
    double u = 0.5 * (pn[in] + pp[in]);
    double c = 0.5 * (c_n + c_p);
    double Sn = smin(u - c, pn[in] - c_n);
    double Sp = smax(u + c, pp[in] + c_p);
    double Sm = (pp[IP] - pn[IP] + pn[0] * pn[in] * (Sn - pn[in]) - pp[0] * pp[in] * (Sp - pp[in])) /
                (pn[0] * (Sn - pn[in]) - pp[0] * (Sp - pp[in]));
    if (Sn > 0) {
        for (int i = 0; i < NE; ++i) pm[i] = pn[i];
    } else if (Sp < 0) {
        for (int i = 0; i < NE; ++i) pm[i] = pp[i];
    } else if (Sm > 0 && Sn < 0) {
        pm[0] = pn[0] * (Sn - pn[in]) / (Sn - Sm);
        pm[IP] = pn[IP] + pn[0] * (Sn - pn[in]) * (Sm - pn[in]);
        for (int i = 0; i < OD; ++i) pm[IU + i] = pn[IU + i];
        pm[in] = Sm;
    } else {
        pm[0] = pp[0] * (Sp - pp[in]) / (Sp - Sm);
        pm[IP] = pp[IP] + pp[0] * (Sp - pp[in]) * (Sm - pp[in]);
        for (int i = 0; i < OD; ++i) pm[IU + i] = pp[IU + i];
        pm[in] = Sm;
    }
#endif
}

#if NG > 0
/* get_grad_input: navier_stokes.hpp:51-59; multi.hpp:148-156 */
static void grad_input(const o_params *P, const double *p, double *g) {
    for (int i = 0; i < NG; ++i) g[i] = 0.0;
    for (int i = 0; i < OD; ++i) g[i] = p[IU + i];
#if OMODEL == O_MULTI
    g[IDT] = mtemperature(P, p[0], p[1], p[IP], p[IA]);
    g[IDA] = p[IA];
#else
    g[IDT] = p[IP] / (p[0] * P->ns_R);
#endif
}

/* viscous stress vector Fv: model/common.hpp:8-24 */
static void visc_Fv(const double du[OD][NGX], const double *vel, double mu, int dim, double *fv) {
    double trace = 0.0;
    for (int i = 0; i <= OD; ++i) fv[i] = 0.0;
    for (int i = 0; i < OD; ++i) {
        fv[i] = du[dim][i] + du[i][dim];
        trace += du[i][i];
    }
    fv[dim] -= (2.0 / 3.0 * trace);
    for (int i = 0; i < OD; ++i) fv[OD] += fv[i] * vel[i];
    for (int i = 0; i <= OD; ++i) fv[i] = fv[i] * mu;
}
#endif

static double model_mu(const o_params *P, const double *p) {
#if OMODEL == O_MULTI
    return P->mu1 * p[IA] + P->mu2 * (1 - p[IA]);
#else
    (void) p;
    return P->ns_mu;
#endif
}
#if OMODEL == O_MULTI
static double model_lambda(const o_params *P, const double *p) { return P->lambda1 * p[IA] + P->lambda2 * (1 - p[IA]); }
#endif

/* max_delta_time: euler.hpp:203-215; navier_stokes.hpp:101-114; multi.hpp:322-346 */
static double max_delta_time(const o_params *P, const o_cell *c) {
    double cs = sound_speed_prim(P, c->p);
    double m = c->size / (cs + fabs(c->p[IU + 0]));
    for (int i = 1; i < OD; ++i) m = smin(m, c->size / (cs + fabs(c->p[IU + i])));
#if OMODEL == O_EULER
    return m;
#elif OMODEL == O_NAVIER_STOKES
    double k = smax(P->ns_mu, P->ns_lambda / P->ns_cp);
    if (k < 1e-15) return m;
    double dtv = c->size * c->size / k * 0.5;
    return smin(m, dtv);
#else
    double k = smax(model_mu(P, c->p), model_lambda(P, c->p) / smin(P->cp1, P->cp2));
    if (k < 1e-15) return m;
    double dtv = c->size * c->size / k / 4.0;
    return smin(m, dtv);
#endif
}

#if OMODEL == O_MULTI
/* compute_parameters: multi.hpp:174-183 */
static void compute_parameters(const o_params *P, o_cell *c) {
    c->k[0] = P->mu1 * c->p[IA] + P->mu2 * (1 - c->p[IA]);
    c->k[IALPHAS] = c->p[IA];
}
#endif

/* slope limiters: integrator/slope_limiters.hpp:7-57 */
static double limiter(int id, double s1, double s2) {
    switch (id) {
    case O_LIM_MINMOD:
        if (s1 * s2 <= 0.0) return 0.0;
        else if (fabs(s1) < fabs(s2)) return s1;
        else return s2;
    case O_LIM_UNLIMITED: return 0.5 * (s1 + s2);
    case O_LIM_VAN_LEER: {
        double r = s1 / s2;
        return (r + fabs(r)) / (1 + fabs(r)) * s2;
    }
    case O_LIM_SUPERBEE: {
        double r = s1 / s2;
        return smax(smax(0.0, smin(2.0 * r, 1.0)), smin(r, 2.0)) * s2;
    }
    case O_LIM_VON_ALBA_2: {
        double r = s1 / s2;
        return 2.0 * r / (r * r + 1) * s2;
    }
    case O_LIM_KOREN: {
        double r = (s1 + 1e-9) / (s2 + 1e-9);
        return smax(0.0, smin(2.0 * r, smin((1.0 + 2.0 * r) / 3.0, 2.0))) * s2;
    }
    default: return 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* Boundaries: integrator/boundary.hpp:75-160                          */
/* ------------------------------------------------------------------ */
static void bc_normal(int face, double *n) {
    int dim = face >> 1, dir = face & 1;
    for (int i = 0; i < OD; ++i) n[i] = 0.0;
    n[dim] = -(2.0 * dir - 1.0); /* -unit(dim)*(2dir-1): inward normal */
}

/* A boundary condition as Boundary<M> holds it (boundary.hpp:13-60): the
 * type, the wall normal (freeslip), the inflow state and fade-in time. */
typedef struct {
    int type;
    double normal[3];
    const double *p;
    double fade_in;
} o_bnd;

/* the condition of domain face `face` (config/config.cpp:47-61: the wall
 * normal of a domain face is -unit(dim)*(2dir-1)) */
static o_bnd bnd_domain(const o_params *P, int face) {
    o_bnd b;
    b.type = P->bc_type[face];
    bc_normal(face, b.normal);
    b.p = P->bc_p[face];
    b.fade_in = P->bc_fade_in[face];
    return b;
}

/* the condition of face f of a node: its domain face, or its object record */
static o_bnd bnd_of(const o_state *S, const o_node *nd, int f) {
    const uint8_t code = nd->bc[f];
    if (code == O_BC_OBJECT) {
        const struct o_rec *r = &S->recs[nd->brec];
        o_bnd b;
        b.type = r->type;
        for (int i = 0; i < 3; ++i) b.normal[i] = r->normal[i];
        b.p = r->p;
        b.fade_in = r->fade_in;
        return b;
    }
    return bnd_domain(&S->P, code ? code - 1 : f);
}

/* boundary_value: boundary.hpp:75-123 */
static void boundary_value(const o_params *P, const o_bnd *b, const o_cell *c, double *pb) {
    int type = b->type;
    const double *pc = c->p;
    if (type == O_BC_FREESLIP) {
        double dot = 0;
        for (int i = 0; i < NE; ++i) pb[i] = pc[i];
        for (int i = 0; i < OD; ++i) dot += pc[IU + i] * b->normal[i];
        for (int i = 0; i < OD; ++i) pb[IU + i] = pc[IU + i] - b->normal[i] * dot;
        return;
    }
    if (type == O_BC_NOSLIP) {
        for (int i = 0; i < NE; ++i) pb[i] = pc[i];
        for (int i = 0; i < OD; ++i) pb[IU + i] *= -1.0;
        return;
    }
    if (type == O_BC_INFLOW || type == O_BC_INFLOW_R) {
        for (int i = 0; i < NE; ++i) pb[i] = b->p[i];
        if (!supersonic(P, pc)) pb[IP] = pc[IP];
        double fi = b->fade_in;
        if (fi > 0.0) {
            double vm = smin(c->t / fi, 1.0);
            for (int i = 0; i < OD; ++i) pb[IU + i] *= vm;
        }
        return;
    }
    if (type == O_BC_DIRICHLET) {
        for (int i = 0; i < NE; ++i) pb[i] = b->p[i];
        return;
    }
    for (int i = 0; i < NE; ++i) pb[i] = pc[i]; /* outflow, zeroGradient */
}

/* boundary_flux: boundary.hpp:130-160 (result already scaled by
 * area * (t1 - t0) * direction) */
static void boundary_flux(const o_params *P, const o_bnd *b, const o_cell *c, int dim, int dir, double area,
                          double t0, double t1, double *out) {
    double pb[NE], F[NF];
    int type = b->type;
    boundary_value(P, b, c, pb);
    if (type == O_BC_FREESLIP) {
        flux_F_projected(P, c->p, pb[IU + dim], dim, F);
    } else if (type == O_BC_NOSLIP) {
        for (int i = 0; i < NF; ++i) F[i] = 0.0;
        F[IU + dim] = pb[IP];
    } else {
        double pm[NE];
        middle_state(P, pb, c->p, 0.0, dim, pm);
        flux_F(P, pm, dim, F);
    }
#if NG > 0
    {
        double vel0[OD], fv[OD + 1];
        for (int i = 0; i < OD; ++i) vel0[i] = 0.0;
        visc_Fv(c->gradient, vel0, model_mu(P, c->p), dim, fv);
        for (int i = 0; i <= OD; ++i) F[IU + i] -= fv[i];
    }
#endif
    double direction = 2 * dir - 1;
    for (int i = 0; i < NF; ++i) out[i] = F[i] * area * (t1 - t0) * direction;
}

/* ------------------------------------------------------------------ */
/* Domain: integrator/domain.hpp                                       */
/* ------------------------------------------------------------------ */
static double pow_level(double s, int e) { return pow(s, (double) e); }

/* find / _find: domain.hpp:65-82,407-417 */
static int32_t dom_find(const o_state *S, const double *pin, int max_level) {
    double p[OD], x[OD];
    for (int i = 0; i < OD; ++i) p[i] = pin[i];
    for (int i = 0; i < OD; ++i) {
        double h = 0.5 * S->P.active_size[i];
        if (p[i] < -h || p[i] > h) {
            if (S->P.periodic[i]) {
                if (p[i] < -h) p[i] += S->P.active_size[i];
                else p[i] -= S->P.active_size[i];
            } else
                return 0;
        }
    }
    double rs = S->tree[1].cell.size;
    for (int i = 0; i < OD; ++i) x[i] = 2.0 * p[i] / rs;
    int32_t idx = 1;
    for (;;) {
        const o_node *nd = &S->tree[idx];
        if (nd->leaf || nd->tree_level == max_level) return idx;
        int q = 0;
        for (int i = 0; i < OD; ++i)
            if (x[i] > 0.0) q |= 1 << i;
        for (int i = 0; i < OD; ++i) x[i] = (x[i] - 0.5 * ((q >> i) & 1 ? 1.0 : -1.0)) * 2.0;
        idx = nd->child_base + q;
    }
}

/* interpolate_p / _k / _t: domain.hpp:84-148 (recursive mean over active
 * children, in child order) */
static void interp_p(const o_state *S, int32_t idx, double *out) {
    const o_node *nd = &S->tree[idx];
    if (nd->leaf) {
        for (int i = 0; i < NE; ++i) out[i] = nd->cell.p[i];
        return;
    }
    double f[NE], g[NE];
    int n = 0;
    for (int i = 0; i < NE; ++i) f[i] = 0.0;
    for (int c = 0; c < NCH; ++c) {
        int32_t ci = nd->child_base + c;
        if (S->tree[ci].active) {
            interp_p(S, ci, g);
            for (int i = 0; i < NE; ++i) f[i] = f[i] + g[i];
            n++;
        }
    }
    for (int i = 0; i < NE; ++i) out[i] = f[i] / n;
}

static double interp_t(const o_state *S, int32_t idx) {
    const o_node *nd = &S->tree[idx];
    if (nd->leaf) return nd->cell.t;
    double t = 0.0;
    int n = 0;
    for (int c = 0; c < NCH; ++c) {
        int32_t ci = nd->child_base + c;
        if (S->tree[ci].active) {
            t += interp_t(S, ci);
            n++;
        }
    }
    return t / n;
}

#if NP > 0
static double interp_k(const o_state *S, int32_t idx, int k) {
    const o_node *nd = &S->tree[idx];
    if (nd->leaf) return nd->cell.k[k];
    double f = 0.0;
    int n = 0;
    for (int c = 0; c < NCH; ++c) {
        int32_t ci = nd->child_base + c;
        if (S->tree[ci].active) {
            f += interp_k(S, ci, k);
            n++;
        }
    }
    return f / n;
}
#endif

/* update_neighbors: domain.hpp:249-293, for one active leaf */
static void update_neighbors_leaf(o_state *S, int32_t idx) {
    o_node *nd = &S->tree[idx];
    for (int dim = 0; dim < OD; ++dim)
        for (int dir = 0; dir < 2; ++dir) {
            int f = 2 * dim + dir;
            if (nd->bc[f]) continue;
            double loc[OD];
            for (int i = 0; i < OD; ++i) loc[i] = nd->cell.center[i];
            loc[dim] = nd->cell.center[dim] + nd->cell.size * (2 * dir - 1.0);
            int32_t nb = dom_find(S, loc, nd->tree_level);
            if (!nb || !S->tree[nb].active) continue;
            nd->nbr[f] = nb;
            if (S->tree[nb].tree_level < nd->tree_level - 1) {
#pragma omp atomic write
                S->tree[nb].should_split = 1;
            }
        }
}

static void ensure_cap(o_state *S, int64_t need) {
    if (need <= S->cap) return;
    int64_t nc = S->cap ? S->cap : 1024;
    while (nc < need) nc *= 2;
    S->tree = (o_node *) realloc(S->tree, (size_t) nc * sizeof(o_node));
    memset(S->tree + S->cap, 0, (size_t) (nc - S->cap) * sizeof(o_node));
    S->tot = (double *) realloc(S->tot, (size_t) nc * NF * sizeof(double));
    S->list = (int32_t *) realloc(S->list, (size_t) nc * sizeof(int32_t));
    S->tmp = (double *) realloc(S->tmp, (size_t) nc * sizeof(double));
    S->mark = (uint8_t *) realloc(S->mark, (size_t) nc);
    memset(S->mark + S->cap, 0, (size_t) (nc - S->cap));
    S->cap = nc;
}

/* Cell::get_p(x, t): integrator/cell.hpp:23-31 (all dimensions) */
static void cell_get_p_all(const o_cell *c, const double *x, double *out) {
    for (int i = 0; i < NE; ++i) out[i] = c->p[i];
    for (int d = 0; d < OD; ++d) {
        double dx = x[d] - c->center[d];
        for (int i = 0; i < NE; ++i) out[i] = out[i] + c->slope[d][i] * dx;
    }
}

/* split: domain.hpp:192-247.  Children occupy one contiguous block. */
static void dom_split(o_state *S, int32_t pidx) {
    int32_t base = S->tree[pidx].child_base;
    int reuse = base != 0;
    if (!reuse) {
        ensure_cap(S, S->n + NCH);
        base = (int32_t) S->n;
        S->n += NCH;
    }
    o_node parent = S->tree[pidx];
    for (int i = 0; i < NCH; ++i) {
        o_node ch;
        memset(&ch, 0, sizeof ch);
        ch.parent = pidx;
        ch.leaf = 1;
        ch.acti = parent.acti;
        ch.tree_level = (int8_t) (parent.tree_level + 1);
        double q = 0.25 * parent.cell.size;
        for (int d = 0; d < OD; ++d) ch.cell.center[d] = parent.cell.center[d] + ((i >> d) & 1 ? 1.0 : -1.0) * q;
        ch.cell.size = 0.5 * parent.cell.size;
        ch.cell.t = parent.cell.t;
        cell_get_p_all(&parent.cell, ch.cell.center, ch.cell.p);
        cons_from_prim(&S->P, ch.cell.p, ch.cell.u);
        memcpy(ch.cell.k, parent.cell.k, sizeof ch.cell.k);
        memcpy(ch.cell.slope, parent.cell.slope, sizeof ch.cell.slope);
        memcpy(ch.cell.gradient, parent.cell.gradient, sizeof ch.cell.gradient);
        memcpy(ch.cell.t_slope, parent.cell.t_slope, sizeof ch.cell.t_slope);
        ch.error = -1.0;
        ch.should_split = 0;
        ch.garbage = 0;
        ch.can_split = 1;
        ch.can_unsplit = 1;
        ch.creation_time = parent.cell.t;
        ch.cluster = -1;
        ch.buffer = -1;
        ch.child_base = reuse ? S->tree[base + i].child_base : 0;
        S->tree[base + i] = ch;
    }
    S->tree[pidx].child_base = base;
    S->tree[pidx].leaf = 0;
}

/* interpolated_cell: domain.hpp:150-180 */
static void interpolated_cell(const o_state *S, int32_t idx, o_cell *out) {
    const o_node *nd = &S->tree[idx];
    if (nd->leaf) {
        *out = nd->cell;
        return;
    }
    o_cell c = nd->cell;
    c.t = 0;
    for (int i = 0; i < NGX; ++i) c.t_slope[i] = 0.0;
    for (int i = 0; i < NE; ++i) c.u[i] = 0.0;
    for (int i = 0; i < NPX; ++i) c.k[i] = 0.0;
    for (int ch = 0; ch < NCH; ++ch) {
        o_cell cc;
        interpolated_cell(S, nd->child_base + ch, &cc);
        c.t += cc.t;
        for (int i = 0; i < NE; ++i) c.u[i] = c.u[i] + cc.u[i];
        for (int i = 0; i < NPX; ++i) c.k[i] = c.k[i] + cc.k[i];
        for (int d = 0; d < OD; ++d)
            for (int i = 0; i < NGX; ++i) c.gradient[d][i] = c.gradient[d][i] + cc.gradient[d][i];
        for (int i = 0; i < NGX; ++i) c.t_slope[i] = c.t_slope[i] + cc.t_slope[i];
    }
    c.t /= NCH;
    for (int i = 0; i < NGX; ++i) c.t_slope[i] = c.t_slope[i] / NCH;
    for (int i = 0; i < NE; ++i) c.u[i] = c.u[i] / NCH;
    prim_from_cons(&S->P, c.u, c.p);
    for (int i = 0; i < NPX; ++i) c.k[i] = c.k[i] / NCH;
    for (int d = 0; d < OD; ++d)
        for (int i = 0; i < NGX; ++i) c.gradient[d][i] = c.gradient[d][i] / NCH;
    *out = c;
}

/* unsplit: domain.hpp:296-325 */
static void dom_unsplit(o_state *S, int32_t pidx) {
    o_cell c;
    interpolated_cell(S, pidx, &c);
    double err = 0.0;
    o_node *pn = &S->tree[pidx];
    for (int ch = 0; ch < NCH; ++ch) {
        o_node *cn = &S->tree[pn->child_base + ch];
        err += cn->error;
        cn->garbage = 1;
        cn->leaf = 1;
    }
    for (int f = 0; f < NFACE; ++f) pn->nbr[f] = 0;
    pn->cell = c;
    pn->leaf = 1;
    pn->error = err;
    memset(pn->acc, 0, sizeof pn->acc);
    memset(pn->acc_cov, 0, sizeof pn->acc_cov);
    pn->creation_time = c.t;
    pn->active = 1;
    pn->acti = -1;
    pn->can_split = 1;
}

/* gc: domain.hpp:327-356 + impl/acti_utils.hpp:58-74 (compaction) */
static void dom_gc_compact(o_state *S) {
    int32_t *o2n = (int32_t *) calloc((size_t) S->n, sizeof(int32_t));
    int64_t e = 1;
    for (int64_t i = 1; i < S->n; ++i) {
        if (!S->tree[i].garbage) {
            S->tree[e] = S->tree[i];
            o2n[i] = (int32_t) e;
            e++;
        } else
            o2n[i] = 0;
    }
    int64_t old_n = S->n;
    S->n = e;
    for (int64_t i = 1; i < S->n; ++i) {
        o_node *nd = &S->tree[i];
        nd->parent = o2n[nd->parent];
        nd->child_base = nd->child_base ? o2n[nd->child_base] : 0;
        for (int f = 0; f < NFACE; ++f) nd->nbr[f] = o2n[nd->nbr[f]];
    }
    (void) old_n;
    free(o2n);
}

/* ------------------------------------------------------------------ */
/* Integrator phases: integrator/impl/ (one file per phase)           */
/* ------------------------------------------------------------------ */
static double level_dt(const o_state *S, int L) { return S->global_dt / (1 << L); }

/* compute_slope: impl/slopes.hpp:15-76 */
static void compute_slope(o_state *S, int32_t idx) {
    o_node *nd = &S->tree[idx];
    double pc[NE];
    for (int i = 0; i < NE; ++i) pc[i] = nd->cell.p[i];
    for (int dim = 0; dim < OD; ++dim) {
        int32_t in = nd->nbr[2 * dim], ip = nd->nbr[2 * dim + 1];
        double pn[NE], pp[NE], xc = 0.0, xn, xp, tn, tp;
        if (in == 0) {
            o_bnd b = bnd_of(S, nd, 2 * dim);
            boundary_value(&S->P, &b, &nd->cell, pn);
            xn = -nd->cell.size;
            tn = nd->cell.t;
        } else {
            if (!S->tree[in].leaf) {
                interp_p(S, in, pn);
                tn = interp_t(S, in);
            } else {
                for (int i = 0; i < NE; ++i) pn[i] = S->tree[in].cell.p[i];
                tn = S->tree[in].cell.t;
            }
            xn = -0.5 * (nd->cell.size + S->tree[in].cell.size);
        }
        if (ip == 0) {
            o_bnd b = bnd_of(S, nd, 2 * dim + 1);
            boundary_value(&S->P, &b, &nd->cell, pp);
            xp = nd->cell.size;
            tp = nd->cell.t;
        } else {
            if (!S->tree[ip].leaf) {
                interp_p(S, ip, pp);
                tp = interp_t(S, ip);
            } else {
                for (int i = 0; i < NE; ++i) pp[i] = S->tree[ip].cell.p[i];
                tp = S->tree[ip].cell.t;
            }
            xp = 0.5 * (nd->cell.size + S->tree[ip].cell.size);
        }
        for (int i = 0; i < NE; ++i) {
            double sn = (pc[i] - pn[i]) / (xc - xn);
            double sp = (pp[i] - pc[i]) / (xp - xc);
            nd->cell.slope[dim][i] = limiter(S->P.limiter, sn, sp);
        }
#if NG > 0
        {
            double gp[NG], gn[NG];
            grad_input(&S->P, pp, gp);
            grad_input(&S->P, pn, gn);
            for (int i = 0; i < NG; ++i)
                nd->cell.gradient[dim][i] = (gp[i] - gn[i] - nd->cell.t_slope[i] * (tp - tn)) / (xp - xn);
        }
#else
        (void) tn;
        (void) tp;
#endif
    }
    compute_rhs(&S->P, &nd->cell);
}

/* Model::flux for the face (cn | cp) with the neighbour centre override:
 * euler.hpp:156-200, navier_stokes.hpp:61-82, multi.hpp:249-302.
 * fc: face centre coordinate along dim; cnc, cpc: centre coordinates along
 * dim of cn / cp (after any periodic shift). */
static void model_flux(const o_params *P, const o_cell *cn, double cnc, const o_cell *cp, double cpc, double fc,
                       double area, double t0, double t1, int dim, double *out) {
    double dt_n = 0.5 * (t1 + t0) - cn->t;
    double dt_p = 0.5 * (t1 + t0) - cp->t;
    double pn[NE], pp[NE], pm[NE], F[NF];
    double dxn = fc - cnc, dxp = fc - cpc;
    for (int i = 0; i < NE; ++i) {
        pn[i] = cn->p[i] + cn->slope[dim][i] * dxn;
        pp[i] = cp->p[i] + cp->slope[dim][i] * dxp;
    }
#if OMODEL == O_MULTI
    for (int i = 0; i < NE; ++i) {
        pn[i] = pn[i] - cn->rhs[i] * dt_n;
        pp[i] = pp[i] - cp->rhs[i] * dt_p;
    }
    double ka = 0.5 * (cn->k[IKAPPA] + cp->k[IKAPPA]);
    middle_state(P, pn, pp, ka, dim, pm);
    flux_F(P, pm, dim, F);
    if (pm[IA] > 1e-5 && pm[IA] < 1 - 1e-5) {
        double n = 0.5 * (cn->k[IKN + dim] + cp->k[IKN + dim]);
        double uc = 1.0 * fabs(pm[IU + dim]) * n;
        double ke = 0.0;
        for (int i = 0; i < OD; ++i) {
            F[IU + i] += uc * ((1 - pm[IA]) * pm[0] - pm[IA] * pm[1]) * pm[IU + i];
            ke += 0.5 * pm[IU + i] * pm[IU + i];
        }
        F[0] += uc * (1 - pm[IA]) * pm[0];
        F[1] += uc * (-pm[IA]) * pm[1];
        F[IP] += uc * ((1 - pm[IA]) * pm[0] - pm[IA] * pm[1]) * ke +
                 uc * (1 - pm[IA]) * pm[IA] *
                     (pm[IP] * (1 / (P->gamma1 - 1) - 1 / (P->gamma2 - 1)) + P->gamma1 * P->P1 / (P->gamma1 - 1) -
                      P->gamma2 * P->P2 / (P->gamma2 - 1));
        F[IA] += uc * (1 - pm[IA]) * pm[IA];
    }
    {
        double du[OD][NGX], vel[OD], fv[OD + 1];
        for (int d = 0; d < OD; ++d)
            for (int i = 0; i < NG; ++i) du[d][i] = 0.5 * (cn->gradient[d][i] + cp->gradient[d][i]);
        for (int i = 0; i < OD; ++i) vel[i] = 0.5 * (cn->p[IU + i] + cp->p[IU + i]);
        visc_Fv(du, vel, model_mu(P, pm), dim, fv);
        for (int i = 0; i <= OD; ++i) F[IU + i] -= fv[i];
        F[IU + OD] -= du[dim][IDT] * model_lambda(P, pm);
    }
    for (int i = 0; i < NF; ++i) out[i] = F[i] * area * (t1 - t0);
#else
    for (int i = 0; i < NE; ++i) {
        pn[i] = pn[i] - cn->rhs[i] * dt_n;
        pp[i] = pp[i] - cp->rhs[i] * dt_p;
    }
    middle_state(P, pn, pp, 0.0, dim, pm);
    flux_F(P, pm, dim, F);
    for (int i = 0; i < NF; ++i) out[i] = F[i] * area * (t1 - t0);
#if OMODEL == O_NAVIER_STOKES
    {
        double du[OD][NGX], vel[OD], fv[OD + 1];
        for (int d = 0; d < OD; ++d)
            for (int i = 0; i < NG; ++i) du[d][i] = 0.5 * (cn->gradient[d][i] + cp->gradient[d][i]);
        for (int i = 0; i < OD; ++i) vel[i] = 0.5 * (cn->p[IU + i] + cp->p[IU + i]);
        visc_Fv(du, vel, model_mu(P, cn->p), dim, fv);
        for (int i = 0; i <= OD; ++i) out[IU + i] -= fv[i] * area * (t1 - t0);
        double fh = P->ns_lambda * du[dim][IDT] * area * (t1 - t0);
        out[IU + OD] -= fh;
    }
#endif
#endif
}

/* The flux owner x computes for its face (dim, dir): impl/flux.hpp:32-91.
 * Returns F with the reference's sign convention: x receives -F if dir==1
 * and +F if dir==0; the neighbour receives the opposite. */
static void owner_face_flux(const o_state *S, int32_t x, int dim, int dir, double dt, double *F) {
    const o_node *nx = &S->tree[x];
    int32_t nb = nx->nbr[2 * dim + dir];
    const o_node *nn = &S->tree[nb];
    double fc = nx->cell.center[dim] + nx->cell.size * (dir - 0.5);
    double t0 = nx->cell.t, t1 = nx->cell.t + dt;
    double area = pow_level(nx->cell.size, OD - 1);
    double nbc = nn->cell.center[dim];
    if (S->P.periodic[dim]) {
        if (dir == 1) {
            if (nx->cell.center[dim] + nx->cell.size > 0.5 * S->P.active_size[dim]) nbc += S->P.active_size[dim];
        } else {
            if (nx->cell.center[dim] - nx->cell.size < -0.5 * S->P.active_size[dim]) nbc -= S->P.active_size[dim];
        }
    }
    if (dir == 1) model_flux(&S->P, &nx->cell, nx->cell.center[dim], &nn->cell, nbc, fc, area, t0, t1, dim, F);
    else model_flux(&S->P, &nn->cell, nbc, &nx->cell, nx->cell.center[dim], fc, area, t0, t1, dim, F);
}

/* active leaves of subtree `idx` that touch its face (dim, side), in
 * depth-first child order */
static int sub_face_leaves(const o_state *S, int32_t idx, int dim, int side, int32_t *out, int n) {
    const o_node *nd = &S->tree[idx];
    if (nd->leaf) {
        if (nd->active && !nd->garbage) out[n++] = idx;
        return n;
    }
    for (int c = 0; c < NCH; ++c)
        if (((c >> dim) & 1) == side) n = sub_face_leaves(S, nd->child_base + c, dim, side, out, n);
    return n;
}

#define O_MAX_SUB 4096

/* Deposit into a receiver r (not updated in this substep) the fluxes of
 * face fr of r computed at level L by its finer neighbours, in sub-face
 * order.  Called once per (r, fr) by the first computing contributor. */
static void deposit_from_fine(o_state *S, int32_t r, int fr, int L, double dt) {
    int dim = fr >> 1, dir = fr & 1;
    int32_t nb = S->tree[r].nbr[fr];
    int32_t subs[O_MAX_SUB];
    int ns = sub_face_leaves(S, nb, dim, 1 - dir, subs, 0);
    for (int s = 0; s < ns; ++s) {
        int32_t x = subs[s];
        if (S->tree[x].cluster != L) continue;
        if (S->tree[x].nbr[2 * dim + (1 - dir)] != r) continue;
        if (!(S->tree[r].acti <= L)) continue;
        double F[NF];
        owner_face_flux(S, x, dim, 1 - dir, dt, F);
        /* x's face dir is 1-dir; the receiver gets +F if x's dir==1 */
        for (int i = 0; i < NF; ++i)
            S->tree[r].acc[fr][i] = (1 - dir) == 1 ? S->tree[r].acc[fr][i] + F[i] : S->tree[r].acc[fr][i] - F[i];
        S->tree[r].acc_cov[fr] += dt * pow_level(S->tree[x].cell.size / S->tree[r].cell.size, OD - 1);
    }
}

/* first computing contributor of receiver face (r, fr) at level L */
static int32_t first_contributor(const o_state *S, int32_t r, int fr, int L) {
    int dim = fr >> 1, dir = fr & 1;
    int32_t nb = S->tree[r].nbr[fr];
    int32_t subs[O_MAX_SUB];
    int ns = sub_face_leaves(S, nb, dim, 1 - dir, subs, 0);
    for (int s = 0; s < ns; ++s) {
        int32_t x = subs[s];
        if (S->tree[x].cluster == L && S->tree[x].nbr[2 * dim + (1 - dir)] == r && S->tree[r].acti <= L) return x;
    }
    return 0;
}

/* flux phase of one substep: impl/flux.hpp:10-97 (restated, see header) */
static void flux_phase(o_state *S, const int32_t *list, int64_t nl, int L) {
    double dt = level_dt(S, L);
#pragma omp parallel for schedule(guided, 4)
    for (int64_t li = 0; li < nl; ++li) {
        int32_t y = list[li];
        o_node *ny = &S->tree[y];
        double *tot = S->tot + (size_t) y * NF;
        for (int i = 0; i < NF; ++i) tot[i] = 0.0;
        for (int dim = 0; dim < OD; ++dim)
            for (int dir = 0; dir < 2; ++dir) {
                int f = 2 * dim + dir;
                double c[NF], F[NF];
                for (int i = 0; i < NF; ++i) c[i] = ny->acc[f][i];
                for (int i = 0; i < NF; ++i) ny->acc[f][i] = 0.0;
                double cov = ny->acc_cov[f];
                ny->acc_cov[f] = 0.0;
                int32_t nb = ny->nbr[f];
                {
                    /* Received flux can only come from a face neighbour that
                     * is now on a finer time level (possibly the parent or
                     * children of the cells that deposited it); count any
                     * that appears elsewhere. */
                    int may = 0;
                    if (nb != 0 && S->tree[nb].leaf) {
                        may = S->tree[nb].acti > L;
                    } else if (nb != 0) {
                        int32_t subs[O_MAX_SUB];
                        int ns = sub_face_leaves(S, nb, dim, 1 - dir, subs, 0);
                        for (int s = 0; s < ns; ++s)
                            if (S->tree[subs[s]].acti > L) may = 1;
                    }
                    int nz = 0;
                    for (int i = 0; i < NF; ++i)
                        if (c[i] != 0.0) nz = 1;
                    if (nz && !may) {
#pragma omp atomic
                        S->acc_violations++;
                        if (getenv("ORACLE_DEBUG"))
                            fprintf(stderr,
                                    "acc violation: L=%d y=%d lev=%d acti=%d face=%d nb=%d nbleaf=%d nblev=%d "
                                    "nbacti=%d nbcl=%d nbbuf=%d\n",
                                    L, y, ny->tree_level, ny->acti, f, nb, nb ? S->tree[nb].leaf : -1,
                                    nb ? S->tree[nb].tree_level : -1, nb ? S->tree[nb].acti : -1,
                                    nb ? S->tree[nb].cluster : -1, nb ? S->tree[nb].buffer : -1);
                    }
                }
                if (nb == 0) {
                    double area = pow_level(ny->cell.size, OD - 1);
                    o_bnd b = bnd_of(S, ny, f);
                    boundary_flux(&S->P, &b, &ny->cell, dim, dir, area, ny->cell.t, ny->cell.t + dt, F);
                    for (int i = 0; i < NF; ++i) c[i] = c[i] - F[i];
                    cov += dt;
                } else if (S->tree[nb].leaf) {
                    o_node *nn = &S->tree[nb];
                    int y_computes, nb_computes = 0;
                    if (nn->tree_level < ny->tree_level) {
                        y_computes = nn->acti <= L;
                    } else {
                        y_computes = nn->acti <= L && (dir == 1 || nn->acti < L);
                        nb_computes = nn->cluster == L && nn->nbr[2 * dim + (1 - dir)] == y && dir == 0;
                    }
                    if (y_computes) {
                        owner_face_flux(S, y, dim, dir, dt, F);
                        for (int i = 0; i < NF; ++i) c[i] = dir == 1 ? c[i] - F[i] : c[i] + F[i];
                        cov += dt;
                        if (nn->cluster != L) {
                            if (nn->tree_level == ny->tree_level) {
                                int fo = 2 * dim + (1 - dir);
                                for (int i = 0; i < NF; ++i)
                                    nn->acc[fo][i] = dir == 1 ? nn->acc[fo][i] + F[i] : nn->acc[fo][i] - F[i];
                                nn->acc_cov[fo] += dt;
                            } else {
                                int fo = 2 * dim + (1 - dir);
                                if (first_contributor(S, nb, fo, L) == y) deposit_from_fine(S, nb, fo, L, dt);
                            }
                        }
                    } else if (nb_computes) {
                        owner_face_flux(S, nb, dim, 1 - dir, dt, F); /* nb's dir is 1 */
                        for (int i = 0; i < NF; ++i) c[i] = c[i] + F[i];
                        cov += dt;
                    }
                } else {
                    int32_t subs[O_MAX_SUB];
                    int ns = sub_face_leaves(S, nb, dim, 1 - dir, subs, 0);
                    for (int s = 0; s < ns; ++s) {
                        int32_t x = subs[s];
                        if (S->tree[x].cluster != L) continue;
                        if (S->tree[x].nbr[2 * dim + (1 - dir)] != y) continue;
                        owner_face_flux(S, x, dim, 1 - dir, dt, F);
                        for (int i = 0; i < NF; ++i) c[i] = (1 - dir) == 1 ? c[i] + F[i] : c[i] - F[i];
                        cov += dt * pow_level(S->tree[x].cell.size / ny->cell.size, OD - 1);
                    }
                }
                for (int i = 0; i < NF; ++i) tot[i] = tot[i] + c[i];
                ny->cov[f] += cov;
            }
    }
}

/* cell_update: impl/cell_update.hpp:9-90 (+ multi::source, multi.hpp:186-203) */
static void update_phase(o_state *S, const int32_t *list, int64_t nl, int L, int signal, int *force_break) {
    double dt = level_dt(S, L);
    int fb = 0;
#pragma omp parallel for schedule(guided, 4) reduction(| : fb)
    for (int64_t li = 0; li < nl; ++li) {
        int32_t y = list[li];
        o_node *nd = &S->tree[y];
        o_cell *c = &nd->cell;
        double vol = pow_level(c->size, OD);
        double du[NF];
#if NG > 0
        double gold[NG], gnew[NG];
        grad_input(&S->P, c->p, gold);
#endif
        for (int i = 0; i < NGX; ++i) c->t_slope[i] = 0.0;
        for (int i = 0; i < NF; ++i) du[i] = S->tot[(size_t) y * NF + i] / vol;
        for (int k = 0; k < NE; ++k) c->u[k] += du[k];
#if OMODEL == O_MULTI
        if (!signal) {
            if (c->u[IA] > 0.001)
                for (int k = IU; k < IU + OD; ++k) c->u[k] = 0.0;
        }
        {
            double s[NE];
            for (int i = 0; i < NE; ++i) s[i] = 0.0;
            for (int i = 0; i < OD; ++i) s[IU + i] = -S->P.sigma * c->k[IKAPPA] * du[IA + 2 + i];
            s[IP] = S->P.sigma * c->k[IKAPPA] * (-du[IA] + c->p[IA] * du[IA + 1]);
            s[IA] = -c->p[IA] * du[IA + 1];
            for (int i = 0; i < NE; ++i) c->u[i] = c->u[i] + s[i];
        }
#else
        (void) signal;
#endif
        prim_from_cons(&S->P, c->u, c->p);
        if (isnan(c->p[IP]) || isnan(c->p[0])) fb = 1;
#if OMODEL == O_MULTI
        c->p[0] = smax(c->p[0], 1e-3);
        c->p[1] = smax(c->p[1], 1e-6);
#else
        c->p[0] = smax(c->p[0], 1e-3);
#endif
        c->p[IP] = smax(c->p[IP], 1e-3);
        cons_from_prim(&S->P, c->p, c->u);
#if NG > 0
        grad_input(&S->P, c->p, gnew);
        for (int i = 0; i < NG; ++i) c->t_slope[i] = (gnew[i] - gold[i]) / dt;
#endif
#if OMODEL == O_MULTI
        compute_parameters(&S->P, c);
#endif
        c->t += dt;
    }
    if (fb) *force_break = 1;
    S->cell_updates += nl;
}

#if OMODEL == O_MULTI
/* neighbour value helpers for the interface chain (impl/smooth.hpp,
 * impl/normal_curvature.hpp) */
static double bc_alpha(const o_state *S, const o_node *nd, int f) {
    double pb[NE];
    o_bnd b = bnd_of(S, nd, f);
    boundary_value(&S->P, &b, &nd->cell, pb);
    return pb[IA];
}

/* smooth_alpha: impl/smooth.hpp:9-60 */
static void smooth_alpha(o_state *S, const int32_t *list, int64_t nl) {
    for (int it = 0; it < 3; ++it) {
#pragma omp parallel for schedule(guided, 4)
        for (int64_t li = 0; li < nl; ++li) {
            const o_node *nd = &S->tree[list[li]];
            double acc = 0.0;
            for (int dim = 0; dim < OD; ++dim) {
                int32_t in = nd->nbr[2 * dim], ip = nd->nbr[2 * dim + 1];
                double an, ap;
                if (in == 0) an = bc_alpha(S, nd, 2 * dim);
                else if (!S->tree[in].leaf) an = interp_k(S, in, IALPHAS);
                else an = S->tree[in].cell.k[IALPHAS];
                if (ip == 0) ap = bc_alpha(S, nd, 2 * dim + 1);
                else if (!S->tree[ip].leaf) ap = interp_k(S, ip, IALPHAS);
                else ap = S->tree[ip].cell.k[IALPHAS];
                acc += 0.5 * (an + ap) + nd->cell.k[IALPHAS];
            }
            S->tmp[li] = acc / (2 * OD);
        }
        for (int64_t li = 0; li < nl; ++li) S->tree[list[li]].cell.k[IALPHAS] = S->tmp[li];
    }
}

/* compute_normal: impl/normal_curvature.hpp:6-55 */
static void compute_normal(o_state *S, int32_t idx) {
    o_node *nd = &S->tree[idx];
    double n[OD];
    for (int dim = 0; dim < OD; ++dim) {
        int32_t in = nd->nbr[2 * dim], ip = nd->nbr[2 * dim + 1];
        double xn, xp, an, ap;
        if (in == 0) {
            an = bc_alpha(S, nd, 2 * dim);
            xn = -0.5 * nd->cell.size;
        } else {
            an = !S->tree[in].leaf ? interp_k(S, in, IALPHAS) : S->tree[in].cell.k[IALPHAS];
            xn = -0.5 * (nd->cell.size + S->tree[in].cell.size);
        }
        if (ip == 0) {
            ap = bc_alpha(S, nd, 2 * dim + 1);
            xp = 0.5 * nd->cell.size;
        } else {
            ap = !S->tree[ip].leaf ? interp_k(S, ip, IALPHAS) : S->tree[ip].cell.k[IALPHAS];
            xp = 0.5 * (nd->cell.size + S->tree[ip].cell.size);
        }
        n[dim] = (ap - an) / (xp - xn);
    }
    double d = 0;
    for (int i = 0; i < OD; ++i) d += n[i] * n[i];
    double nrm = sqrt(d);
    if (nrm < 1e-10) {
        for (int i = 0; i < OD; ++i) n[i] = 0.0;
    } else {
        for (int i = 0; i < OD; ++i) n[i] = n[i] / nrm;
    }
    for (int i = 0; i < OD; ++i) nd->cell.k[IKN + i] = n[i];
}

/* compute_curvature: impl/normal_curvature.hpp:58-102 */
static void compute_curvature(o_state *S, int32_t idx) {
    o_node *nd = &S->tree[idx];
    double kappa = 0.0;
    for (int dim = 0; dim < OD; ++dim) {
        int32_t in = nd->nbr[2 * dim], ip = nd->nbr[2 * dim + 1];
        double xn, xp, nn, np;
        if (in == 0) {
            nn = 0.0;
            xn = -0.5 * nd->cell.size;
        } else {
            nn = !S->tree[in].leaf ? interp_k(S, in, IKN + dim) : S->tree[in].cell.k[IKN + dim];
            xn = -0.5 * (nd->cell.size + S->tree[in].cell.size);
        }
        if (ip == 0) {
            np = 0.0;
            xp = 0.5 * nd->cell.size;
        } else {
            np = !S->tree[ip].leaf ? interp_k(S, ip, IKN + dim) : S->tree[ip].cell.k[IKN + dim];
            xp = 0.5 * (nd->cell.size + S->tree[ip].cell.size);
        }
        kappa -= (np - nn) / (xp - xn);
    }
    nd->cell.k[IKAPPA] = kappa;
}

/* smooth_curvature: impl/smooth.hpp:63-142.  std::pow(x, 2.0) is written
 * x*x (the correctly rounded square). */
static void smooth_curvature(o_state *S, const int32_t *list, int64_t nl) {
#pragma omp parallel for schedule(guided, 4)
    for (int64_t li = 0; li < nl; ++li) {
        const o_node *nd = &S->tree[list[li]];
        double kc = nd->cell.k[IKAPPA];
        double wa = nd->cell.k[IALPHAS] * (1 - nd->cell.k[IALPHAS]);
        double wc = wa * wa;
        double w = 0.0, ks = 0.0;
        for (int dim = 0; dim < OD; ++dim) {
            int32_t in = nd->nbr[2 * dim], ip = nd->nbr[2 * dim + 1];
            double kn, kp, an, ap;
            if (in == 0) {
                kn = 0.0;
                an = bc_alpha(S, nd, 2 * dim);
            } else if (!S->tree[in].leaf) {
                double pi[NE];
                kn = interp_k(S, in, IKAPPA);
                interp_p(S, in, pi);
                an = pi[IA];
            } else {
                kn = S->tree[in].cell.k[IKAPPA];
                an = S->tree[in].cell.p[IA];
            }
            if (ip == 0) {
                kp = 0.0;
                ap = bc_alpha(S, nd, 2 * dim + 1);
            } else if (!S->tree[ip].leaf) {
                double pi[NE];
                kp = interp_k(S, ip, IKAPPA);
                interp_p(S, ip, pi);
                ap = pi[IA];
            } else {
                kp = S->tree[ip].cell.k[IKAPPA];
                ap = S->tree[ip].cell.p[IA];
            }
            double xa = an * (1 - an), xb = ap * (1 - ap);
            double wn = xa * xa, wp = xb * xb;
            ks += 0.5 * (wn * kn + wp * kp) + kc * wc;
            w += 0.5 * (wn + wp) + wc;
        }
        S->tmp[li] = (w > 1e-6) ? ks / w : 0.0;
    }
    for (int64_t li = 0; li < nl; ++li) S->tree[list[li]].cell.k[IKAPPA] = S->tmp[li];
}

static void interface_chain(o_state *S, const int32_t *list, int64_t nl) {
    smooth_alpha(S, list, nl);
#pragma omp parallel for schedule(guided, 4)
    for (int64_t li = 0; li < nl; ++li) compute_normal(S, list[li]);
#pragma omp parallel for schedule(guided, 4)
    for (int64_t li = 0; li < nl; ++li) compute_curvature(S, list[li]);
    smooth_curvature(S, list, nl);
}
#endif

/* Scene::can_split: scene/scene.hpp:50-61 (axis-aligned zones) */
static int scene_can_split(const o_params *P, const double *x, int level) {
    if (level < P->max_refinement_level) return 1;
    for (int z = 0; z < P->n_zones; ++z) {
        double x3[3] = {x[0], OD > 1 ? x[1 % OD] : 0.0, OD > 2 ? x[2 % OD] : 0.0};
        if (level < P->zone_level[z] && x3[0] < P->zone_x1[z][0] && x3[1] < P->zone_x1[z][1] &&
            x3[2] < P->zone_x1[z][2] && x3[0] > P->zone_x0[z][0] && x3[1] > P->zone_x0[z][1] &&
            x3[2] > P->zone_x0[z][2])
            return 1;
    }
    return 0;
}

/* error estimate for one leaf: impl/adaptive_refinement.hpp:35-68 +
 * weighted_laplacian::estimate, error_estimators.hpp:19-37 */
static double estimate_error(const o_state *S, int32_t idx) {
    const o_node *nd = &S->tree[idx];
    double nbp[NFACE][NE];
    int has[NFACE];
    for (int dim = 0; dim < OD; ++dim)
        for (int dir = 0; dir < 2; ++dir) {
            int f = 2 * dim + dir;
            int32_t nb = nd->nbr[f];
            if (nb == 0) {
                has[f] = 0;
                continue;
            }
            has[f] = 1;
            if (!S->tree[nb].leaf) {
                interp_p(S, nb, nbp[f]);
            } else if (S->tree[nb].tree_level < nd->tree_level) {
                double x[OD];
                for (int i = 0; i < OD; ++i) x[i] = nd->cell.center[i];
                x[dim] = nd->cell.center[dim] + nd->cell.size * (2 * dir - 1.0);
                if (S->P.periodic[dim] && dir == 0 && x[dim] < -0.5 * S->P.active_size[dim])
                    x[dim] += S->P.active_size[dim];
                if (S->P.periodic[dim] && dir == 1 && x[dim] > 0.5 * S->P.active_size[dim])
                    x[dim] -= S->P.active_size[dim];
                cell_get_p_all(&S->tree[nb].cell, x, nbp[f]);
            } else {
                for (int i = 0; i < NE; ++i) nbp[f][i] = S->tree[nb].cell.p[i];
            }
        }
    double err = 0.0;
    for (int dim = 0; dim < OD; ++dim) {
        double d2[NE];
        for (int i = 0; i < NE; ++i) d2[i] = nd->cell.p[i] * -2.0;
        for (int dir = 0; dir < 2; ++dir) {
            int f = 2 * dim + dir;
            if (!has[f]) {
                for (int i = 0; i < NE; ++i) d2[i] = 0.0;
                break;
            }
            for (int i = 0; i < NE; ++i) d2[i] = d2[i] + nbp[f][i];
        }
        double sq = 0.0;
        for (int i = 0; i < NE; ++i) {
            double v = d2[i] * S->P.weights[i];
            sq += v * v;
        }
        err += sq;
    }
    return sqrt(err);
}

#define MK_SPLIT 1
#define MK_UNSPLIT_CAND 2
#define MK_UNSPLIT_CHILD 4

/* LTS reassignment of one entry of acti_clusters[L]:
 * impl/adaptive_refinement.hpp:264-322.  nd->max_dt is set by the caller. */
static void lts_reassign_one(const o_params *P, o_node *nd, int L, double dt, int counter) {
    nd->cfl = dt / nd->max_dt;
    if (!nd->leaf || nd->garbage) {
        nd->cluster = -1;
        return;
    }
    if (nd->cfl < 0.5 * P->max_cfl && counter == 1 && L > 0) {
        nd->cluster = -1;
        nd->buffer = (int8_t) (L - 1);
        nd->cfl *= 2;
    } else if (nd->cfl > P->max_cfl && L < O_MAX_LEVEL - 1) {
        int raise = (int) (log2(nd->cfl / P->max_cfl) + 1);
        if (L + raise < O_MAX_LEVEL) {
            nd->acti = (int8_t) (L + raise);
            nd->cfl /= 1 << raise;
            nd->cluster = (int8_t) (L + raise);
        } else {
            nd->cluster = -1;
        }
    }
}

/* adaptive_refinement: impl/adaptive_refinement.hpp:6-323 */
static void adaptive_refinement(o_state *S, int L) {
    double dt = level_dt(S, L);
    const o_params *P = &S->P;
    int64_t nl = 0;
    for (int64_t i = 1; i < S->n; ++i)
        if (S->tree[i].cluster == L) S->list[nl++] = (int32_t) i;
    int any = 0;
    /* error estimation (:18-94) */
#pragma omp parallel for schedule(guided, 4) reduction(| : any)
    for (int64_t li = 0; li < nl; ++li) {
        int32_t y = S->list[li];
        o_node *nd = &S->tree[y];
        nd->error = estimate_error(S, y);
        if (nd->can_split) nd->can_split = (uint8_t) scene_can_split(P, nd->cell.center, nd->tree_level);
        uint8_t m = 0;
        if ((nd->error > S->P.max_error && nd->can_split) || nd->should_split) m |= MK_SPLIT;
        if (nd->error < S->P.min_error && nd->can_unsplit && nd->tree_level > 0 && S->acti_counter[L] == 1)
            m |= MK_UNSPLIT_CAND;
        S->mark[y] = m;
        any |= m;
    }
    if (any) {
        /* unsplit detection (:115-149) */
        for (int64_t li = 0; li < nl; ++li) {
            int32_t y = S->list[li];
            o_node *nd = &S->tree[y];
            if (!(S->mark[y] & MK_UNSPLIT_CAND)) continue;
            if (!nd->can_unsplit) continue;
            int ok = 1;
            for (int f = 0; f < NFACE; ++f) {
                int32_t nb = nd->nbr[f];
                if (nb != 0 && (!S->tree[nb].leaf || S->tree[nb].acti > L)) ok = 0;
            }
            if (ok) S->mark[y] |= MK_UNSPLIT_CHILD;
        }
        /* split application (:154-196); slopes on the pre-split tree */
        int64_t ns = 0;
        int32_t *sl = (int32_t *) malloc((size_t) (nl + 1) * sizeof(int32_t));
        for (int64_t li = 0; li < nl; ++li) {
            int32_t y = S->list[li];
            if (!(S->mark[y] & MK_SPLIT)) continue;
            int ready = 1;
            for (int f = 0; f < NFACE; ++f) {
                int32_t nb = S->tree[y].nbr[f];
                if (nb != 0 && S->tree[nb].tree_level < S->tree[y].tree_level - 1) ready = 0;
            }
            if (ready) sl[ns++] = y;
        }
        for (int64_t s = 0; s < ns; ++s) compute_slope(S, sl[s]);
        for (int64_t s = 0; s < ns; ++s) {
            int32_t y = sl[s];
            dom_split(S, y);
            o_node *nd = &S->tree[y];
            nd->should_split = 0;
            for (int ci = 0; ci < NCH; ++ci) {
                o_node *ch = &S->tree[nd->child_base + ci];
                ch->active = 1;
                for (int dim = 0; dim < OD; ++dim)
                    for (int dir = 0; dir < 2; ++dir)
                        if (((ci >> dim) & 1) == dir) ch->bc[2 * dim + dir] = nd->bc[2 * dim + dir];
                ch->brec = nd->brec;
                ch->acti = (int8_t) L;
                ch->cluster = (int8_t) L;
            }
        }
        /* unsplit application (:198-220): parents of marked children, each once */
        int64_t nu = 0;
        for (int64_t li = 0; li < nl; ++li) {
            int32_t y = S->list[li];
            if (!(S->mark[y] & MK_UNSPLIT_CHILD)) continue;
            int32_t par = S->tree[y].parent;
            if (S->tree[par].leaf) continue; /* already unsplit */
            int ready = 1;
            for (int c = 0; c < NCH; ++c) {
                int32_t ch = S->tree[par].child_base + c;
                if (!(S->tree[ch].cluster == L && (S->mark[ch] & MK_UNSPLIT_CHILD))) ready = 0;
            }
            if (!ready) continue;
            dom_unsplit(S, par);
            S->tree[par].cluster = (int8_t) L;
            S->tree[par].acti = (int8_t) L;
            S->tree[par].active = 1;
            nu++;
        }
        for (int64_t li = 0; li < nl; ++li) S->mark[S->list[li]] = 0;
        /* neighbour update (:224-261), as a recomputation for all leaves */
        if (ns || nu)
#pragma omp parallel for schedule(guided, 64)
            for (int64_t i = 1; i < S->n; ++i) {
                o_node *nd = &S->tree[i];
                if (nd->leaf && nd->active && !nd->garbage) update_neighbors_leaf(S, (int32_t) i);
            }
        free(sl);
    }
    /* LTS reassignment (:264-322) */
    nl = 0;
    for (int64_t i = 1; i < S->n; ++i)
        if (S->tree[i].cluster == L) S->list[nl++] = (int32_t) i;
#pragma omp parallel for schedule(guided, 4)
    for (int64_t li = 0; li < nl; ++li) {
        o_node *nd = &S->tree[S->list[li]];
        nd->max_dt = max_delta_time(P, &nd->cell);
        lts_reassign_one(P, nd, L, dt, S->acti_counter[L]);
    }
}

/* assign_to_acti_clusters: impl/acti_utils.hpp:10-24 */
static void assign_all(o_state *S) {
    for (int64_t i = 1; i < S->n; ++i) {
        o_node *nd = &S->tree[i];
        if (!nd->active || !nd->leaf || nd->garbage) continue;
        nd->cfl = S->global_dt / nd->max_dt;
        nd->acti = (int8_t) iclamp((int) log2(nd->cfl / S->P.max_cfl), 0, O_MAX_LEVEL - 1);
        nd->cfl /= 1 << nd->acti;
        nd->cluster = nd->acti;
    }
}

static void recount(o_state *S) {
    for (int l = 0; l < O_MAX_LEVEL; ++l) S->counts[l] = 0;
    for (int64_t i = 1; i < S->n; ++i)
        if (S->tree[i].cluster >= 0) S->counts[S->tree[i].cluster]++;
}

/* one level substep: impl/integrate.hpp:42-64.  Returns 1 on force_break. */
static int substep(o_state *S, int L, int signal, int *force_break) {
    int64_t nl = 0;
    for (int64_t i = 1; i < S->n; ++i)
        if (S->tree[i].cluster == L) S->list[nl++] = (int32_t) i;
    flux_phase(S, S->list, nl, L);
    update_phase(S, S->list, nl, L, signal, force_break);
    if (*force_break) return 1;
#pragma omp parallel for schedule(guided, 4)
    for (int64_t li = 0; li < nl; ++li) compute_slope(S, S->list[li]);
#if OMODEL == O_MULTI
    interface_chain(S, S->list, nl);
#endif
    adaptive_refinement(S, L);
    S->acti_counter[L] = (S->acti_counter[L] + 1) % 2;
    S->acti_time[L] += S->global_dt / (1 << L);
    for (int64_t i = 1; i < S->n; ++i) {
        o_node *nd = &S->tree[i];
        if (nd->buffer == L) {
            nd->buffer = -1;
            nd->cluster = (int8_t) L;
            nd->acti = (int8_t) L;
        }
    }
    recount(S);
    S->n_substeps++;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Exported API                                                        */
/* ------------------------------------------------------------------ */

/* Integrator constructor: integrator/integrator.hpp:72-163.  The tree
 * arrives in the exchange format (state after Scene::init_domain). */
void *OPFX(create)(const o_params *P, int64_t n, const int32_t *parent, const int32_t *child_base,
                   const int8_t *level, const uint8_t *flags, const uint8_t *bc, const double *center,
                   const double *size, const double *p, const int32_t *brec, int64_t n_recs, const void *recs) {
    o_state *S = (o_state *) calloc(1, sizeof(o_state));
    S->P = *P;
    S->n_recs = n_recs;
    if (n_recs > 0) {
        S->recs = (struct o_rec *) malloc(sizeof(struct o_rec) * (size_t) n_recs);
        memcpy(S->recs, recs, sizeof(struct o_rec) * (size_t) n_recs);
    }
    ensure_cap(S, n + 16);
    S->n = n;
    for (int64_t i = 0; i < n; ++i) {
        o_node *nd = &S->tree[i];
        memset(nd, 0, sizeof *nd);
        nd->parent = parent[i];
        nd->child_base = child_base[i];
        nd->tree_level = level[i];
        nd->leaf = (flags[i] & O_F_LEAF) != 0;
        nd->active = (flags[i] & O_F_ACTIVE) != 0;
        nd->garbage = (flags[i] & O_F_GARBAGE) != 0;
        nd->can_split = (flags[i] & O_F_CAN_SPLIT) != 0;
        nd->can_unsplit = (flags[i] & O_F_CAN_UNSPLIT) != 0;
        nd->should_split = (flags[i] & O_F_SHOULD_SPLIT) != 0;
        nd->is_boundary = (flags[i] & O_F_BOUNDARY) != 0;
        for (int f = 0; f < NFACE; ++f) nd->bc[f] = bc[(size_t) i * NFACE + f];
        nd->brec = brec ? brec[i] : -1;
        for (int d = 0; d < OD; ++d) nd->cell.center[d] = center[(size_t) i * OD + d];
        nd->cell.size = size[i];
        nd->cell.t = 0.0;
        for (int k = 0; k < NE; ++k) nd->cell.p[k] = p[(size_t) i * NE + k];
        nd->cluster = -1;
        nd->buffer = -1;
        nd->acti = -1;
        nd->error = -1.0;
    }
    double gmax = -INFINITY;
    for (int64_t i = 1; i < n; ++i) {
        o_node *nd = &S->tree[i];
        if (nd->leaf && nd->active) {
            cons_from_prim(&S->P, nd->cell.p, nd->cell.u);
#if OMODEL == O_MULTI
            compute_parameters(&S->P, &nd->cell);
#endif
            for (int k = 0; k < NGX; ++k) nd->cell.t_slope[k] = 0.0;
            nd->max_dt = max_delta_time(&S->P, &nd->cell);
            gmax = smax(gmax, nd->max_dt);
        }
    }
    /* neighbours and slopes of every leaf: independent per leaf */
#pragma omp parallel for schedule(guided, 64)
    for (int64_t i = 1; i < n; ++i) {
        o_node *nd = &S->tree[i];
        if (nd->leaf && nd->active) update_neighbors_leaf(S, (int32_t) i);
    }
#pragma omp parallel for schedule(guided, 64)
    for (int64_t i = 1; i < n; ++i) {
        o_node *nd = &S->tree[i];
        if (nd->leaf && nd->active) compute_slope(S, (int32_t) i);
    }
#if OMODEL == O_MULTI
    {
        int64_t nl = 0;
        for (int64_t i = 1; i < n; ++i)
            if (S->tree[i].leaf && S->tree[i].active) S->list[nl++] = (int32_t) i;
        interface_chain(S, S->list, nl);
    }
#endif
    S->global_dt = smin(0.9 * gmax, S->P.max_dt);
    assign_all(S);
    for (int l = 0; l < O_MAX_LEVEL; ++l) {
        S->acti_counter[l] = 0;
        S->acti_time[l] = 0.0;
    }
    S->max_acti_level = 0;
    recount(S);
    for (int l = 0; l < O_MAX_LEVEL; ++l)
        if (S->counts[l] > 0 && l > S->max_acti_level) S->max_acti_level = l;
    return S;
}

void OPFX(destroy)(void *h) {
    o_state *S = (o_state *) h;
    if (!S) return;
    free(S->tree);
    free(S->tot);
    free(S->list);
    free(S->tmp);
    free(S->mark);
    free(S->recs);
    free(S);
}

/* integrate(): impl/integrate.hpp:11-87.  Returns global_dt. */
double OPFX(integrate)(void *h, int signal, int *force_break) {
    o_state *S = (o_state *) h;
    for (int64_t i = 0; i < S->n; ++i)
        for (int f = 0; f < NFACE; ++f) S->tree[i].cov[f] = 0.0;
    int64_t cnt[O_MAX_LEVEL];
    for (int l = 0; l < O_MAX_LEVEL; ++l) cnt[l] = -1;
    for (int seq = 1; seq <= (1 << (O_MAX_LEVEL - 1)); ++seq) {
        for (int l = 0; l < O_MAX_LEVEL; ++l) {
            if (seq % (1 << l) != 0) break;
            int L = O_MAX_LEVEL - l - 1;
            if (L > S->max_acti_level) {
                if (S->counts[L] > 0) {
                    S->max_acti_level = L;
                    S->acti_time[L] = S->acti_time[L - 1];
                }
            }
            if (L > S->max_acti_level) continue;
            if (substep(S, L, signal, force_break)) return S->global_dt;
        }
    }
    (void) cnt;
    int64_t nact = 0, ntree = S->n - 1;
    for (int64_t i = 1; i < S->n; ++i)
        if (S->tree[i].cluster >= 0) nact++;
    S->n_global_steps++;
    if (!S->no_gc && (double) nact / (double) ntree < 0.2) {
        dom_gc_compact(S);
        for (int64_t i = 1; i < S->n; ++i) S->tree[i].cluster = -1;
        assign_all(S);
        recount(S);
    }
    return S->global_dt;
}

/* run `steps` integrate() calls; returns number completed before a NaN break */
int64_t OPFX(run)(void *h, int64_t steps, int signal) {
    for (int64_t s = 0; s < steps; ++s) {
        int fb = 0;
        OPFX(integrate)(h, signal, &fb);
        if (fb) return s;
    }
    return steps;
}

int64_t OPFX(n_nodes)(void *h) { return ((o_state *) h)->n; }
int64_t OPFX(cell_updates)(void *h) { return ((o_state *) h)->cell_updates; }
int64_t OPFX(acc_violations)(void *h) { return ((o_state *) h)->acc_violations; }
int64_t OPFX(n_substeps)(void *h) { return ((o_state *) h)->n_substeps; }
double OPFX(global_dt)(void *h) { return ((o_state *) h)->global_dt; }
int OPFX(max_acti_level)(void *h) { return ((o_state *) h)->max_acti_level; }

/* Export every node: integer topology and FP64 state, node-index order. */
void OPFX(export_nodes)(void *h, int32_t *parent, int32_t *child_base, int32_t *nbr, int8_t *level, int8_t *acti,
                        int8_t *cluster, uint8_t *flags, double *center, double *size, double *t, double *p,
                        double *u, double *slope, double *rhs, double *k, double *grad, double *meta) {
    o_state *S = (o_state *) h;
    for (int64_t i = 0; i < S->n; ++i) {
        const o_node *nd = &S->tree[i];
        parent[i] = nd->parent;
        child_base[i] = nd->child_base;
        for (int f = 0; f < NFACE; ++f) nbr[i * NFACE + f] = nd->nbr[f];
        level[i] = nd->tree_level;
        acti[i] = nd->acti;
        cluster[i] = nd->buffer >= 0 ? (int8_t) (O_MAX_LEVEL + nd->buffer) : nd->cluster;
        flags[i] = (uint8_t) ((nd->leaf ? O_F_LEAF : 0) | (nd->active ? O_F_ACTIVE : 0) |
                              (nd->garbage ? O_F_GARBAGE : 0) | (nd->can_split ? O_F_CAN_SPLIT : 0) |
                              (nd->can_unsplit ? O_F_CAN_UNSPLIT : 0) | (nd->should_split ? O_F_SHOULD_SPLIT : 0) |
                              (nd->is_boundary ? O_F_BOUNDARY : 0));
        for (int d = 0; d < OD; ++d) center[i * OD + d] = nd->cell.center[d];
        size[i] = nd->cell.size;
        t[i] = nd->cell.t;
        for (int e = 0; e < NE; ++e) {
            p[i * NE + e] = nd->cell.p[e];
            u[i * NE + e] = nd->cell.u[e];
            rhs[i * NE + e] = nd->cell.rhs[e];
        }
        for (int d = 0; d < OD; ++d)
            for (int e = 0; e < NE; ++e) slope[(i * OD + d) * NE + e] = nd->cell.slope[d][e];
        for (int e = 0; e < NPX; ++e) k[i * NPX + e] = nd->cell.k[e];
        for (int d = 0; d < OD; ++d)
            for (int e = 0; e < NGX; ++e) grad[(i * OD + d) * NGX + e] = nd->cell.gradient[d][e];
        meta[i * 4 + 0] = nd->acti;
        meta[i * 4 + 1] = nd->error;
        meta[i * 4 + 2] = nd->max_dt;
        meta[i * 4 + 3] = nd->cfl;
    }
}


/* ---- test hooks on single phases (closed-form and brute-force pins) ---- */

/* face time coverage of the last global step, n x NFACE */
void OPFX(face_cover)(void *h, double *out) {
    o_state *S = (o_state *) h;
    for (int64_t i = 0; i < S->n; ++i)
        for (int f = 0; f < NFACE; ++f) out[i * NFACE + f] = S->tree[i].cov[f];
}
void OPFX(set_no_gc)(void *h, int on) { ((o_state *) h)->no_gc = on; }
/* the compaction integrate() runs when the active fraction is low
 * (impl/integrate.hpp:78-84), run now */
void OPFX(gc_now)(void *h) {
    o_state *S = (o_state *) h;
    dom_gc_compact(S);
    for (int64_t i = 1; i < S->n; ++i) S->tree[i].cluster = -1;
    assign_all(S);
    recount(S);
}

/* LTS reassignment of one cell (lts_reassign_one): in: max_dt, level L, dt,
 * counter, leaf; out[4] = cluster, buffer, acti, cfl (acti_in is the cell's
 * ACTI level before the decision) */
void OPFX(t_lts_reassign)(const o_params *P, double max_dt, int L, double dt, int counter, int leaf, int acti_in,
                          double *out) {
    o_node nd;
    memset(&nd, 0, sizeof nd);
    nd.max_dt = max_dt;
    nd.leaf = (uint8_t) leaf;
    nd.acti = (int8_t) acti_in;
    nd.cluster = (int8_t) L;
    nd.buffer = -1;
    lts_reassign_one(P, &nd, L, dt, counter);
    out[0] = nd.cluster;
    out[1] = nd.buffer;
    out[2] = nd.acti;
    out[3] = nd.cfl;
}

static o_state *t_scratch_state(const o_params *P, int64_t n) {
    o_state *S = (o_state *) calloc(1, sizeof(o_state));
    S->P = *P;
    ensure_cap(S, n + 16);
    S->n = n;
    memset(S->tree, 0, (size_t) S->cap * sizeof(o_node));
    return S;
}

/* Domain::split of one leaf given as [p NE, slope OD*NE, k NPX, gradient
 * OD*NGX, t_slope NGX, center OD, size, t]; children written as
 * [p NE, u NE, center OD, size, t, slope OD*NE, k NPX, gradient OD*NGX,
 * t_slope NGX] per child */
void OPFX(t_split)(const o_params *P, const double *q, double *out) {
    o_state *S = t_scratch_state(P, 2);
    o_node *pn = &S->tree[1];
    pn->leaf = 1;
    pn->active = 1;
    for (int i = 0; i < NE; ++i) pn->cell.p[i] = *q++;
    for (int d = 0; d < OD; ++d)
        for (int i = 0; i < NE; ++i) pn->cell.slope[d][i] = *q++;
    for (int i = 0; i < NPX; ++i) pn->cell.k[i] = *q++;
    for (int d = 0; d < OD; ++d)
        for (int i = 0; i < NGX; ++i) pn->cell.gradient[d][i] = *q++;
    for (int i = 0; i < NGX; ++i) pn->cell.t_slope[i] = *q++;
    for (int d = 0; d < OD; ++d) pn->cell.center[d] = *q++;
    pn->cell.size = *q++;
    pn->cell.t = *q++;
    dom_split(S, 1);
    int32_t cb = S->tree[1].child_base;
    for (int ch = 0; ch < NCH; ++ch) {
        const o_cell *c = &S->tree[cb + ch].cell;
        for (int i = 0; i < NE; ++i) *out++ = c->p[i];
        for (int i = 0; i < NE; ++i) *out++ = c->u[i];
        for (int d = 0; d < OD; ++d) *out++ = c->center[d];
        *out++ = c->size;
        *out++ = c->t;
        for (int d = 0; d < OD; ++d)
            for (int i = 0; i < NE; ++i) *out++ = c->slope[d][i];
        for (int i = 0; i < NPX; ++i) *out++ = c->k[i];
        for (int d = 0; d < OD; ++d)
            for (int i = 0; i < NGX; ++i) *out++ = c->gradient[d][i];
        for (int i = 0; i < NGX; ++i) *out++ = c->t_slope[i];
    }
    OPFX(destroy)(S);
}

/* Domain::unsplit of a parent whose NCH leaf children are given as
 * [u NE, t, k NPX, gradient OD*NGX, t_slope NGX, error] each; pgrad is the
 * parent's stored gradient (OD*NGX).  out = [u NE, p NE, t, k NPX,
 * gradient OD*NGX, t_slope NGX, error] of the parent. */
void OPFX(t_unsplit)(const o_params *P, const double *q, const double *pgrad, double *out) {
    o_state *S = t_scratch_state(P, 2 + NCH);
    o_node *pn = &S->tree[1];
    pn->child_base = 2;
    pn->active = 1;
    for (int d = 0; d < OD; ++d)
        for (int i = 0; i < NGX; ++i) pn->cell.gradient[d][i] = pgrad[d * NGX + i];
    for (int ch = 0; ch < NCH; ++ch) {
        o_node *c = &S->tree[2 + ch];
        c->leaf = 1;
        c->active = 1;
        c->parent = 1;
        for (int i = 0; i < NE; ++i) c->cell.u[i] = *q++;
        c->cell.t = *q++;
        for (int i = 0; i < NPX; ++i) c->cell.k[i] = *q++;
        for (int d = 0; d < OD; ++d)
            for (int i = 0; i < NGX; ++i) c->cell.gradient[d][i] = *q++;
        for (int i = 0; i < NGX; ++i) c->cell.t_slope[i] = *q++;
        c->error = *q++;
    }
    dom_unsplit(S, 1);
    const o_cell *c = &S->tree[1].cell;
    for (int i = 0; i < NE; ++i) *out++ = c->u[i];
    for (int i = 0; i < NE; ++i) *out++ = c->p[i];
    *out++ = c->t;
    for (int i = 0; i < NPX; ++i) *out++ = c->k[i];
    for (int d = 0; d < OD; ++d)
        for (int i = 0; i < NGX; ++i) *out++ = c->gradient[d][i];
    for (int i = 0; i < NGX; ++i) *out++ = c->t_slope[i];
    *out++ = S->tree[1].error;
    OPFX(destroy)(S);
}

/* boundary_flux of a cell with primitives p (NE), gradient (OD*NGX) and time t */
void OPFX(t_boundary_flux)(const o_params *P, int face, const double *p, const double *grad, double t, int dim,
                           int dir, double area, double t0, double t1, double *out) {
    o_cell c;
    memset(&c, 0, sizeof c);
    for (int i = 0; i < NE; ++i) c.p[i] = p[i];
    for (int d = 0; d < OD; ++d)
        for (int i = 0; i < NGX; ++i) c.gradient[d][i] = grad[d * NGX + i];
    c.t = t;
    o_bnd b = bnd_domain(P, face);
    boundary_flux(P, &b, &c, dim, dir, area, t0, t1, out);
}

/* Direct access to single-cell model functions, for closed-form tests. */
void OPFX(t_prim_from_cons)(const o_params *P, const double *u, double *p) { prim_from_cons(P, u, p); }
void OPFX(t_cons_from_prim)(const o_params *P, const double *p, double *u) { cons_from_prim(P, p, u); }
void OPFX(t_flux_F)(const o_params *P, const double *p, int dim, double *F) { flux_F(P, p, dim, F); }
void OPFX(t_middle_state)(const o_params *P, const double *pn, const double *pp, double ka, int dim, double *pm) {
    middle_state(P, pn, pp, ka, dim, pm);
}
double OPFX(t_limiter)(int id, double a, double b) { return limiter(id, a, b); }
double OPFX(t_sound_speed)(const o_params *P, const double *p) { return sound_speed_prim(P, p); }
int OPFX(t_supersonic)(const o_params *P, const double *p) { return supersonic(P, p); }
void OPFX(t_boundary_value)(const o_params *P, int face, const double *p, double t, double *pb) {
    o_cell c;
    memset(&c, 0, sizeof c);
    for (int i = 0; i < NE; ++i) c.p[i] = p[i];
    c.t = t;
    o_bnd b = bnd_domain(P, face);
    boundary_value(P, &b, &c, pb);
}
/* the boundary codes (NFACE per node) and object records of the pool */
void OPFX(export_bc)(void *h, uint8_t *bc, int32_t *brec) {
    o_state *S = (o_state *) h;
    for (int64_t i = 0; i < S->n; ++i) {
        for (int f = 0; f < NFACE; ++f) bc[(size_t) i * NFACE + f] = S->tree[i].bc[f];
        brec[i] = S->tree[i].brec;
    }
}

/* ghost state of an object boundary record (type, wall normal, inflow state, fade-in) */
void OPFX(t_boundary_value_rec)(const o_params *P, int type, const double *normal, const double *pin, double fade,
                                const double *p, double t, double *pb) {
    o_cell c;
    memset(&c, 0, sizeof c);
    for (int i = 0; i < NE; ++i) c.p[i] = p[i];
    c.t = t;
    o_bnd b;
    b.type = type;
    for (int i = 0; i < 3; ++i) b.normal[i] = normal[i];
    b.p = pin;
    b.fade_in = fade;
    boundary_value(P, &b, &c, pb);
}
double OPFX(t_max_delta_time)(const o_params *P, const double *p, double size) {
    o_cell c;
    memset(&c, 0, sizeof c);
    for (int i = 0; i < NE; ++i) c.p[i] = p[i];
    c.size = size;
    return max_delta_time(P, &c);
}
/* rhs of a cell from p, slopes (D x NE) and kappa */
void OPFX(t_compute_rhs)(const o_params *P, const double *p, const double *slope, double kappa, double *rhs) {
    o_cell c;
    memset(&c, 0, sizeof c);
    for (int i = 0; i < NE; ++i) c.p[i] = p[i];
    for (int d = 0; d < OD; ++d)
        for (int i = 0; i < NE; ++i) c.slope[d][i] = slope[d * NE + i];
#if OMODEL == O_MULTI
    c.k[IKAPPA] = kappa;
#else
    (void) kappa;
#endif
    compute_rhs(P, &c);
    for (int i = 0; i < NE; ++i) rhs[i] = c.rhs[i];
}
/* Model::flux of one face between two cells, each given as a pack
 * [p (NE), slope (OD x NE), rhs (NE), t, gradient (OD x NGX), k (NPX)] */
static void t_unpack(const double *q, o_cell *c) {
    memset(c, 0, sizeof *c);
    for (int i = 0; i < NE; ++i) c->p[i] = *q++;
    for (int d = 0; d < OD; ++d)
        for (int i = 0; i < NE; ++i) c->slope[d][i] = *q++;
    for (int i = 0; i < NE; ++i) c->rhs[i] = *q++;
    c->t = *q++;
    for (int d = 0; d < OD; ++d)
        for (int i = 0; i < NGX; ++i) c->gradient[d][i] = *q++;
    for (int i = 0; i < NPX; ++i) c->k[i] = *q++;
}
void OPFX(t_model_flux)(const o_params *P, const double *cn, double cnc, const double *cp, double cpc, double fc,
                        double area, double t0, double t1, int dim, double *out) {
    o_cell a, b;
    t_unpack(cn, &a);
    t_unpack(cp, &b);
    model_flux(P, &a, cnc, &b, cpc, fc, area, t0, t1, dim, out);
}
int OPFX(t_sizes)(int *out) {
    out[0] = OD;
    out[1] = NE;
    out[2] = NF;
    out[3] = NP;
    out[4] = NG;
    return 0;
}
