/*
 * oracle_common.h -- shared declarations of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a plain-C restatement of the
 * reference solver's octree leaf update (DACTI, src/lib/integrator/ and
 * src/lib/model/).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline.
 * The product (dacti_amd/) never links, loads or calls anything here.
 */
#ifndef DACTI_ORACLE_COMMON_H
#define DACTI_ORACLE_COMMON_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* model ids */
#define O_EULER 0
#define O_NAVIER_STOKES 1
#define O_MULTI 2

/* limiter ids (reference: integrator/slope_limiters.hpp:7-57) */
#define O_LIM_MINMOD 0
#define O_LIM_UNLIMITED 1
#define O_LIM_VAN_LEER 2
#define O_LIM_SUPERBEE 3
#define O_LIM_VON_ALBA_2 4
#define O_LIM_KOREN 5
#define O_LIM_FIRST_ORDER 6

/* boundary types (reference: integrator/boundary.hpp:21-30) */
#define O_BC_ZERO_GRADIENT 0
#define O_BC_FREESLIP 1
#define O_BC_NOSLIP 2
#define O_BC_INFLOW 3
#define O_BC_INFLOW_R 4
#define O_BC_OUTFLOW 5
#define O_BC_DIRICHLET 6
/* bc code of a face whose condition is its node's object record */
#define O_BC_OBJECT 7

/* node flag bits of the tree exchange format */
#define O_F_LEAF 1
#define O_F_ACTIVE 2
#define O_F_GARBAGE 4
#define O_F_CAN_SPLIT 8
#define O_F_CAN_UNSPLIT 16
#define O_F_SHOULD_SPLIT 32
#define O_F_BOUNDARY 64

#define O_MAX_ZONES 8
#define O_MAX_LEVEL 16

/* Case parameters: the fields of Config<M> the integrator reads
 * (integrator/integrator.hpp:80-86, impl/flux.hpp:68-76) plus the model
 * constants of model/{euler,navier_stokes,multi}.hpp. */
typedef struct o_params {
    int32_t model, dim, limiter, max_refinement_level;
    double side_length;
    double active_size[3];
    int32_t periodic[3];
    int32_t n_zones;
    double max_cfl, max_dt, min_error, max_error;
    double weights[8];
    /* euler / navier_stokes */
    double gamma, ns_mu, ns_lambda, ns_cp, ns_R;
    /* multi (five-equation) */
    double sigma, gamma1, gamma2, P1, P2, cp1, cp2, cv1, cv2, mu1, mu2, lambda1, lambda2;
    /* domain boundary conditions, face index 2*dim+dir */
    int32_t bc_type[6];
    double bc_p[6][8];
    double bc_fade_in[6];
    /* refinement zones (axis-aligned boxes, identity transform) */
    double zone_x0[O_MAX_ZONES][3];
    double zone_x1[O_MAX_ZONES][3];
    int32_t zone_level[O_MAX_ZONES];
} o_params;

#ifdef __cplusplus
}
#endif
#endif
