/*
 * dacti_gpu.h -- C ABI of the MI355X octree leaf-update engine.
 *
 * The reference solver exposes this path as a header-only C++ template,
 * dacti::integrator::Integrator<Model, Limiter, ErrorEstimator>
 * (src/lib/integrator/integrator.hpp:14-225), instantiated by each example
 * main (src/examples/multi2d.cpp:19-23).  It has no FFI; the entry points
 * below are the ones a C-ABI binding of that class needs, one per public
 * member it exercises.  All pointers are plain host pointers; every call
 * returns 0 on success or a negative DACTI_E_* code, with the message
 * available from dacti_last_error().  A handle owns its device buffers and
 * its HIP stream; handles share no state and one host thread drives each.
 */
#ifndef DACTI_GPU_H
#define DACTI_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dacti_handle dacti_handle;

/* status codes */
#define DACTI_OK 0
#define DACTI_E_ARG (-1)     /* invalid argument */
#define DACTI_E_HIP (-2)     /* HIP runtime error */
#define DACTI_E_MEMORY (-3)  /* device allocation failed */
#define DACTI_E_STATE (-4)   /* call not valid in the handle's state */

/* model ids: model/euler.hpp, model/navier_stokes.hpp, model/multi.hpp */
#define DACTI_MODEL_EULER 0
#define DACTI_MODEL_NAVIER_STOKES 1
#define DACTI_MODEL_MULTI 2

/* limiter ids: integrator/slope_limiters.hpp:7-57 */
#define DACTI_LIMITER_MINMOD 0
#define DACTI_LIMITER_UNLIMITED 1
#define DACTI_LIMITER_VAN_LEER 2
#define DACTI_LIMITER_SUPERBEE 3
#define DACTI_LIMITER_VON_ALBA_2 4
#define DACTI_LIMITER_KOREN 5
#define DACTI_LIMITER_FIRST_ORDER 6

/* boundary types: integrator/boundary.hpp:21-30 */
#define DACTI_BC_ZERO_GRADIENT 0
#define DACTI_BC_FREESLIP 1
#define DACTI_BC_NOSLIP 2
#define DACTI_BC_INFLOW 3
#define DACTI_BC_INFLOW_R 4
#define DACTI_BC_OUTFLOW 5
#define DACTI_BC_DIRICHLET 6
/* bc code of a face whose condition is its node's object record (brec) */
#define DACTI_BC_OBJECT 7

/* node flag bits (tree exchange format) */
#define DACTI_F_LEAF 1
#define DACTI_F_ACTIVE 2
#define DACTI_F_GARBAGE 4
#define DACTI_F_CAN_SPLIT 8
#define DACTI_F_CAN_UNSPLIT 16
#define DACTI_F_SHOULD_SPLIT 32
#define DACTI_F_BOUNDARY 64

#define DACTI_MAX_ZONES 8
#define DACTI_MAX_LEVEL 16

/*
 * Case parameters.  These are the Config<M> fields the integrator reads
 * (integrator/integrator.hpp:80-86, impl/flux.hpp:68-76,
 * scene/scene.hpp:50-61) and the model constants that the reference keeps
 * as process-global statics (model/multi.hpp:50-62,
 * model/navier_stokes.hpp:27-31); here they belong to one handle.
 */
typedef struct dacti_case {
    int32_t model, dim, limiter, max_refinement_level;
    double side_length;
    double active_size[3];
    int32_t periodic[3];
    int32_t n_zones;
    double max_cfl, max_dt, min_error, max_error;
    double weights[8];
    /* euler / navier_stokes: gamma (1.4), mu, lambda, cp (1005), R (287) */
    double gamma, ns_mu, ns_lambda, ns_cp, ns_R;
    /* multi: sigma, stiffened-gas pair, heat capacities, viscosities */
    double sigma, gamma1, gamma2, P1, P2, cp1, cp2, cv1, cv2, mu1, mu2, lambda1, lambda2;
    /* domain boundary conditions, face index 2*dim+dir */
    int32_t bc_type[6];
    double bc_p[6][8];
    double bc_fade_in[6];
    /* refinement zones (axis-aligned boxes) */
    double zone_x0[DACTI_MAX_ZONES][3];
    double zone_x1[DACTI_MAX_ZONES][3];
    int32_t zone_level[DACTI_MAX_ZONES];
} dacti_case;

/*
 * Initial tree: the node pool after Scene::init_domain
 * (scene/scene.cpp:369-431), in the layout of integrator/domain.hpp:14-63.
 * Node 0 is the null node, node 1 the root.  A node's 2^dim children occupy
 * indices child_base .. child_base + 2^dim - 1 (0 = no children).
 * bc[i*2*dim + f] is 0 for no boundary, 1 + the domain face index of the
 * boundary condition that face carries, or DACTI_BC_OBJECT for an object
 * boundary (the record brec[i]); p holds the primitive initial state of
 * active leaves (num_eqns values per node).  An object boundary condition
 * (integrator/boundary.hpp:13-60) is a type (DACTI_BC_*), the unit wall
 * normal for freeslip, the inflow state and fade-in time.
 */
typedef struct dacti_boundary {
    int32_t type, pad;
    double normal[3];
    double p[8];
    double fade_in;
} dacti_boundary;

typedef struct dacti_tree {
    int64_t n_nodes;
    const int32_t *parent;
    const int32_t *child_base;
    const int8_t *tree_level;
    const uint8_t *flags;
    const uint8_t *bc;
    const double *center; /* n_nodes * dim */
    const double *size;
    const double *p;      /* n_nodes * num_eqns */
    /* object boundaries (NULL / 0 without objects): brec[i] is the record
     * of node i's DACTI_BC_OBJECT faces, -1 for none */
    const int32_t *brec;
    int64_t n_records;
    const dacti_boundary *records;
} dacti_tree;

/*
 * Node-pool view filled by dacti_export_nodes (caller-allocated, n_nodes
 * entries; sizes per node: nbr 2*dim, center dim, p/u/rhs num_eqns,
 * slope dim*num_eqns, k max(num_parm,1), grad dim*max(num_grad,1), meta 4
 * = {acti_level, error, max_dt, cfl}, integrator/integrator.hpp:30).
 * cluster: LTS level whose cluster holds the node, 16 + level for a node
 * waiting in an acti buffer, -1 for none.  Any pointer may be NULL.
 */
typedef struct dacti_node_view {
    int32_t *parent, *child_base, *nbr;
    int8_t *tree_level, *acti_level, *cluster;
    uint8_t *flags;
    double *center, *size, *t, *p, *u, *slope, *rhs, *k, *grad, *meta;
} dacti_node_view;

/*
 * Case reader and initial tree (Config<M>::load_config,
 * config/config.cpp:10-243, and Scene<M>::init_domain,
 * scene/scene.cpp:369-431): reads the TOML text of a case (geometry paths
 * relative to base_dir) for a model, dimension and limiter and builds the
 * initial tree.  A "grid" initial condition (an extension) takes
 * ic_grid[shape[0]][shape[1]]([shape[2]])[num_eqns].  The case and tree
 * views stay valid until dacti_scene_destroy; errors go to
 * dacti_last_error(NULL).
 */
typedef struct dacti_scene dacti_scene;
int dacti_scene_build(const char *toml_text, const char *base_dir, int32_t model, int32_t dim, int32_t limiter,
                      const double *ic_grid, const int32_t *ic_grid_shape, dacti_scene **out);
int dacti_scene_case(const dacti_scene *s, dacti_case *out);
int dacti_scene_tree(const dacti_scene *s, dacti_tree *out);
int dacti_scene_destroy(dacti_scene *s);

/*
 * Refinement predictor on the GPU (an addition, no reference counterpart):
 * dacti_features writes, for the active leaves of h (ids ascending), their
 * feature rows: the primitives, the differences to the primitives at the
 * 3^dim - 1 same-size neighbour positions (the leaf containing the centre
 * moved by one cell width, found by descending from the root and comparing
 * with node centres; zero outside a non-periodic domain), the gradient
 * (dim x num_grad), the tree level minus min_level and log10 of the error
 * estimate; *n_features = num_eqns * 3^dim + dim * num_grad + 2.  With mu
 * and sd (device arrays of n_features floats) each value is standardised as
 * (x - mu) / sd in FP32.  feat (cap x n_features floats), ids (cap) and
 * the optional centers (cap x dim doubles) and levels (cap) of the leaves
 * are device arrays; *n receives the number of active leaves, and nothing is
 * written when it exceeds cap (call with cap 0 for the sizes).
 * dacti_locate_levels gives the tree level of the leaf containing each of n
 * points (device arrays: pts n x dim doubles, levels n; -1 outside the
 * root cell), by the same descent.
 * dacti_mark_predicted flags to split, at the next remesh, the listed
 * leaves whose three logits (n x 3 device floats: coarsen, keep, refine)
 * pick refine, plus `rings` rings of active face neighbours; *n_marked
 * counts the flagged leaves.
 */
int dacti_features(dacti_handle *h, int32_t min_level, const float *mu, const float *sd, float *feat, int32_t *ids,
                   double *centers, int8_t *levels, int64_t cap, int64_t *n, int32_t *n_features);
int dacti_locate_levels(dacti_handle *h, const double *pts, int64_t n, int8_t *levels);
int dacti_mark_predicted(dacti_handle *h, const float *logits, const int32_t *ids, int64_t n, int32_t rings,
                         int64_t *n_marked);

/* counters and timers of a handle */
typedef struct dacti_stats {
    int64_t n_nodes, n_active;
    int64_t n_substeps, n_global_steps;
    int64_t cell_updates;       /* leaf-cell updates: sum of cluster sizes per substep */
    int64_t flux_launches;      /* launches of the fused flux+update kernel */
    double flux_kernel_ms;      /* HIP-event time of those launches (when timing is on) */
    double substep_kernel_ms;   /* HIP-event time of all per-substep kernels (when timing is on) */
    double global_dt;
    int32_t max_acti_level;
    int32_t pad;
} dacti_stats;

/* sizes of a model: out[0..4] = dim, num_eqns, num_flux, num_parm, num_grad */
int dacti_model_sizes(int32_t model, int32_t dim, int32_t *out);

/*
 * Integrator constructor: integrator/integrator.hpp:72-163.  Uploads the
 * tree and, on the GPU, sets conserved state and parameters, the CFL limit
 * of each leaf and its max-reduction to global_dt, neighbours, slopes, the
 * multiphase interface fields, and the initial LTS clusters.
 */
int dacti_create(const dacti_case *c, const dacti_tree *t, dacti_handle **out);

/* Integrator::integrate(hook, signal, force_break): impl/integrate.hpp:11-87.
 * One global step; *global_dt receives the (constant) global step and
 * *force_break is set to 1 when a NaN pressure or density appeared. */
int dacti_integrate(dacti_handle *h, int32_t signal, double *global_dt, int32_t *force_break);

/* The substep_hook argument of Integrator::integrate (impl/integrate.hpp:64):
 * when set, dacti_integrate / dacti_run (and a batch stepping the handle)
 * call fn(user, h, level, acti_time[level], n_substeps, dt_level) after
 * every LTS substep the handle takes, with the handle's state complete, so
 * fn may call dacti_export_nodes / dacti_get_stats on h.  fn = NULL (the
 * default, the reference driver's no-op hook) removes it.  With a hook set
 * every substep ends on the host (as without one; a graphed whole step is
 * not possible then). */
typedef void (*dacti_substep_hook)(void *user, dacti_handle *h, int32_t level, double acti_time, int64_t n_substeps,
                                   double dt);
int dacti_set_substep_hook(dacti_handle *h, dacti_substep_hook fn, void *user);

/* `steps` calls of dacti_integrate; *completed = steps done before a break. */
int dacti_run(dacti_handle *h, int64_t steps, int32_t signal, int64_t *completed);

/* Integrator::n_active_cells(): integrator/integrator.hpp:193-202 */
int dacti_n_active(dacti_handle *h, int64_t *n);

/* size of the node pool (Integrator::get_nodes().size(), :165) */
int dacti_n_nodes(dacti_handle *h, int64_t *n);

/* Integrator::get_nodes() / for_each_active_cell(): :165-191 */
int dacti_export_nodes(dacti_handle *h, const dacti_node_view *v);

int dacti_get_stats(dacti_handle *h, dacti_stats *s);

/* LTS schedule state, the public acti_time / acti_counter arrays of
 * Integrator (integrator/integrator.hpp:48-49): DACTI_MAX_LEVEL entries
 * each; either pointer may be NULL */
int dacti_get_schedule(dacti_handle *h, double *acti_time, int32_t *acti_counter);

/* turn HIP-event timing of the per-substep kernels on or off */
int dacti_set_timing(dacti_handle *h, int32_t on);

/* Launch configuration of this handle's own stepping (see
 * dacti_batch_set_launch): block 64/128/256, persistent grid (multiple of 8,
 * 0 = one block per tile), graphs (needs grid > 0).  Results do not depend
 * on it; small adaptive cases gain from graph replay. */
int dacti_set_launch(dacti_handle *h, int32_t block, int32_t grid, int32_t graphs);

/* Flux op for dense clusters (no reference counterpart; results do not
 * depend on it): a level's cluster that fills at least half of the node pool
 * is stepped by windows of 64 consecutive node ids, each face between two
 * updating same-level leaves of a window evaluated once.  mode 0 never, 1
 * when the cluster is that dense (the default), 2 for every cluster of a
 * handle stepped alone (batches always use the per-cell op). */
int dacti_set_dense(dacti_handle *h, int32_t mode);

/* Remeshing cadence (an extension; the reference remeshes in every substep,
 * impl/integrate.hpp:59): with on = 0 the sweep sets no split or unsplit
 * marks, so the mesh stays as it is while the LTS levels still adapt. */
int dacti_set_remesh(dacti_handle *h, int32_t on);

/* Ask for the listed nodes to be split at the next remesh, as the 2:1
 * balance does through Node::should_split (integrator/node.hpp); only active
 * leaves that may still split are flagged.  Used by the learned predictor to
 * refine ahead of a moving feature. */
int dacti_mark_split(dacti_handle *h, const int32_t *ids, int64_t n);

int dacti_destroy(dacti_handle *h);

/* message of the last failed call on h (or of the last failed create when
 * h is NULL); never NULL */
const char *dacti_last_error(dacti_handle *h);

/*
 * Batches.  The reference runs one integrator per process, its model
 * constants being process-global statics (model/multi.hpp:50-62); it has no
 * batch API.  A batch steps many independent handles of one model and
 * dimension together: every op of an LTS substep is one launch over the
 * cluster lists of all instances, each instance taking part exactly when
 * its own Integrator::integrate() (impl/integrate.hpp:11-87) would run that
 * substep, so each evolves as it would alone (same results, bit for bit).
 * While a batch exists its handles must be stepped only through it; their
 * read-side calls (export, stats, n_active) stay valid.  Destroying the
 * batch leaves the handles intact; destroy a batch before any of its
 * handles.
 */
typedef struct dacti_batch dacti_batch;

int dacti_batch_create(dacti_handle *const *handles, int32_t n, dacti_batch **out);

/* one global step of every instance; *force_break = 1 when an instance
 * produced a NaN state (the batch stops in that substep) */
int dacti_batch_integrate(dacti_batch *b, int32_t signal, int32_t *force_break);
int dacti_batch_run(dacti_batch *b, int64_t steps, int32_t signal, int64_t *completed);

/* launch configuration: block = list entries per tile (64, 128 or 256),
 * grid = persistent grid in blocks (a multiple of 8; 0 = one block per
 * tile), graphs = replay the fixed kernel sequences of a substep from HIP
 * graphs (needs grid > 0).  The results do not depend on it. */
int dacti_batch_set_launch(dacti_batch *b, int32_t block, int32_t grid, int32_t graphs);
int dacti_batch_get_launch(dacti_batch *b, int32_t *block, int32_t *grid, int32_t *graphs);

/* autotuner: the next global steps each run one candidate configuration and
 * are timed; after the last candidate the fastest is kept.  Tuning runs on
 * the real steps, which advance the state as usual. */
int dacti_batch_autotune(dacti_batch *b);
/* candidates and their step times; *n: capacity in, count out */
int dacti_batch_tune_results(dacti_batch *b, int32_t *n, int32_t *block, int32_t *grid, int32_t *graphs,
                             double *ms);

/* totals over the batch (cell_updates, substeps, kernel time when timing) */
int dacti_batch_get_stats(dacti_batch *b, dacti_stats *s);
int dacti_batch_set_timing(dacti_batch *b, int32_t on);
int dacti_batch_destroy(dacti_batch *b);
const char *dacti_batch_last_error(dacti_batch *b);

#ifdef __cplusplus
}
#endif
#endif
