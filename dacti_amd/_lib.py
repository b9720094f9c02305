"""ctypes binding of the C ABI declared in include/dacti_gpu.h.

The shared library is built in-tree by dacti_amd/build.py.  There is no
fallback: loading fails with an error when the library is missing.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdacti_gpu.so")

MODEL_EULER, MODEL_NAVIER_STOKES, MODEL_MULTI = 0, 1, 2
MAX_ZONES = 8

F_LEAF, F_ACTIVE, F_GARBAGE, F_CAN_SPLIT, F_CAN_UNSPLIT, F_SHOULD_SPLIT, F_BOUNDARY = 1, 2, 4, 8, 16, 32, 64


class CaseStruct(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("dim", C.c_int32), ("limiter", C.c_int32), ("max_refinement_level", C.c_int32),
        ("side_length", C.c_double),
        ("active_size", C.c_double * 3),
        ("periodic", C.c_int32 * 3),
        ("n_zones", C.c_int32),
        ("max_cfl", C.c_double), ("max_dt", C.c_double), ("min_error", C.c_double), ("max_error", C.c_double),
        ("weights", C.c_double * 8),
        ("gamma", C.c_double), ("ns_mu", C.c_double), ("ns_lambda", C.c_double), ("ns_cp", C.c_double),
        ("ns_R", C.c_double),
        ("sigma", C.c_double), ("gamma1", C.c_double), ("gamma2", C.c_double), ("P1", C.c_double),
        ("P2", C.c_double), ("cp1", C.c_double), ("cp2", C.c_double), ("cv1", C.c_double), ("cv2", C.c_double),
        ("mu1", C.c_double), ("mu2", C.c_double), ("lambda1", C.c_double), ("lambda2", C.c_double),
        ("bc_type", C.c_int32 * 6),
        ("bc_p", (C.c_double * 8) * 6),
        ("bc_fade_in", C.c_double * 6),
        ("zone_x0", (C.c_double * 3) * MAX_ZONES),
        ("zone_x1", (C.c_double * 3) * MAX_ZONES),
        ("zone_level", C.c_int32 * MAX_ZONES),
    ]


# object boundary record (dacti_boundary): type, unit wall normal, inflow state, fade-in
RECORD_DTYPE = np.dtype([("type", np.int32), ("pad", np.int32), ("normal", np.float64, 3), ("p", np.float64, 8),
                         ("fade_in", np.float64)])
BC_OBJECT = 7  # bc code of a face whose condition is its node's record


class TreeStruct(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int64),
        ("parent", C.POINTER(C.c_int32)), ("child_base", C.POINTER(C.c_int32)),
        ("tree_level", C.POINTER(C.c_int8)), ("flags", C.POINTER(C.c_uint8)), ("bc", C.POINTER(C.c_uint8)),
        ("center", C.POINTER(C.c_double)), ("size", C.POINTER(C.c_double)), ("p", C.POINTER(C.c_double)),
        ("brec", C.POINTER(C.c_int32)), ("n_records", C.c_int64), ("records", C.c_void_p),
    ]


class NodeViewStruct(C.Structure):
    _fields_ = [
        ("parent", C.POINTER(C.c_int32)), ("child_base", C.POINTER(C.c_int32)), ("nbr", C.POINTER(C.c_int32)),
        ("tree_level", C.POINTER(C.c_int8)), ("acti_level", C.POINTER(C.c_int8)), ("cluster", C.POINTER(C.c_int8)),
        ("flags", C.POINTER(C.c_uint8)),
        ("center", C.POINTER(C.c_double)), ("size", C.POINTER(C.c_double)), ("t", C.POINTER(C.c_double)),
        ("p", C.POINTER(C.c_double)), ("u", C.POINTER(C.c_double)), ("slope", C.POINTER(C.c_double)),
        ("rhs", C.POINTER(C.c_double)), ("k", C.POINTER(C.c_double)), ("grad", C.POINTER(C.c_double)),
        ("meta", C.POINTER(C.c_double)),
    ]


class StatsStruct(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int64), ("n_active", C.c_int64), ("n_substeps", C.c_int64), ("n_global_steps", C.c_int64),
        ("cell_updates", C.c_int64), ("flux_launches", C.c_int64),
        ("flux_kernel_ms", C.c_double), ("substep_kernel_ms", C.c_double), ("global_dt", C.c_double),
        ("max_acti_level", C.c_int32), ("pad", C.c_int32),
    ]


# every entry point of include/dacti_gpu.h with its argument types
SIGNATURES = {
    "dacti_model_sizes": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "dacti_create": (C.c_int, [C.POINTER(CaseStruct), C.POINTER(TreeStruct), C.POINTER(C.c_void_p)]),
    "dacti_integrate": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "dacti_run": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_int64)]),
    "dacti_n_active": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "dacti_n_nodes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "dacti_export_nodes": (C.c_int, [C.c_void_p, C.POINTER(NodeViewStruct)]),
    "dacti_get_stats": (C.c_int, [C.c_void_p, C.POINTER(StatsStruct)]),
    "dacti_get_schedule": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "dacti_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "dacti_set_launch": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "dacti_set_remesh": (C.c_int, [C.c_void_p, C.c_int32]),
    "dacti_set_dense": (C.c_int, [C.c_void_p, C.c_int32]),
    "dacti_mark_split": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.c_int64]),
    "dacti_destroy": (C.c_int, [C.c_void_p]),
    "dacti_features": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "dacti_locate_levels": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "dacti_mark_predicted": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                       C.POINTER(C.c_int64)]),
    "dacti_scene_build": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_void_p)]),
    "dacti_scene_case": (C.c_int, [C.c_void_p, C.POINTER(CaseStruct)]),
    "dacti_scene_tree": (C.c_int, [C.c_void_p, C.POINTER(TreeStruct)]),
    "dacti_scene_destroy": (C.c_int, [C.c_void_p]),
    "dacti_last_error": (C.c_char_p, [C.c_void_p]),
    "dacti_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p)]),
    "dacti_batch_integrate": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "dacti_batch_run": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_int64)]),
    "dacti_batch_set_launch": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "dacti_batch_get_launch": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32)]),
    "dacti_batch_autotune": (C.c_int, [C.c_void_p]),
    "dacti_batch_tune_results": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                           C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "dacti_batch_get_stats": (C.c_int, [C.c_void_p, C.POINTER(StatsStruct)]),
    "dacti_batch_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "dacti_batch_destroy": (C.c_int, [C.c_void_p]),
    "dacti_batch_last_error": (C.c_char_p, [C.c_void_p]),
}

# void fn(void *user, dacti_handle *h, int32_t level, double acti_time, int64_t n_substeps, double dt)
SUBSTEP_HOOK = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_int64, C.c_double)
SIGNATURES["dacti_set_substep_hook"] = (C.c_int, [C.c_void_p, SUBSTEP_HOOK, C.c_void_p])

_lib = None


def load():
    """Load libdacti_gpu.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m dacti_amd.build`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def model_sizes(model, dim):
    """(dim, num_eqns, num_flux, num_parm, num_grad) of a model."""
    ne = dim + 4 if model == MODEL_MULTI else dim + 2
    nf = ne + dim + 1 if model == MODEL_MULTI else ne
    npar = dim + 4 if model == MODEL_MULTI else 0
    ng = 0 if model == MODEL_EULER else (dim + 1 if model == MODEL_NAVIER_STOKES else dim + 3)
    return dim, ne, nf, npar, ng


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def case_struct(case):
    """Fill a CaseStruct from a dacti_amd.case.Case."""
    s = CaseStruct()
    s.model, s.dim, s.limiter = case.model, case.dim, case.limiter
    s.max_refinement_level = case.max_refinement_level
    s.side_length = case.side_length
    for i in range(3):
        s.active_size[i] = case.active_size[i] if i < case.dim else 0.0
        s.periodic[i] = int(case.periodic[i]) if i < case.dim else 0
    s.n_zones = len(case.zones)
    s.max_cfl, s.max_dt = case.max_cfl, case.max_dt
    s.min_error, s.max_error = case.error_threshold
    for i, w in enumerate(case.weights):
        s.weights[i] = w
    for name in ("gamma", "ns_mu", "ns_lambda", "ns_cp", "ns_R", "sigma", "gamma1", "gamma2", "P1", "P2",
                 "cp1", "cp2", "cv1", "cv2", "mu1", "mu2", "lambda1", "lambda2"):
        setattr(s, name, float(case.consts[name]))
    for f in range(6):
        bc = case.bcs[f]
        s.bc_type[f] = bc.type
        for e, v in enumerate(bc.value):
            s.bc_p[f][e] = v
        s.bc_fade_in[f] = bc.fade_in
    for z, (x0, x1, lev) in enumerate(case.zones):
        for a in range(3):
            s.zone_x0[z][a] = x0[a]
            s.zone_x1[z][a] = x1[a]
        s.zone_level[z] = lev
    return s


def tree_struct(tree):
    """TreeStruct over the arrays of a dacti_amd.scene.Tree (kept alive by the caller)."""
    t = TreeStruct()
    t.n_nodes = tree.n
    t.parent = ptr(tree.parent, C.c_int32)
    t.child_base = ptr(tree.child_base, C.c_int32)
    t.tree_level = ptr(tree.level, C.c_int8)
    t.flags = ptr(tree.flags, C.c_uint8)
    t.bc = ptr(tree.bc, C.c_uint8)
    t.center = ptr(tree.center, C.c_double)
    t.size = ptr(tree.size, C.c_double)
    t.p = ptr(tree.p, C.c_double)
    if getattr(tree, "brec", None) is not None and len(tree.records):
        t.brec = ptr(tree.brec, C.c_int32)
        t.n_records = len(tree.records)
        t.records = tree.records.ctypes.data
    return t


def alloc_node_arrays(n, model, dim):
    """Host arrays for one node-pool export, in the C ABI's node-major layout."""
    _, ne, nf, npar, ng = model_sizes(model, dim)
    npx, ngx = max(npar, 1), max(ng, 1)
    return {
        "parent": np.zeros(n, np.int32), "child_base": np.zeros(n, np.int32),
        "nbr": np.zeros((n, 2 * dim), np.int32),
        "tree_level": np.zeros(n, np.int8), "acti_level": np.zeros(n, np.int8), "cluster": np.zeros(n, np.int8),
        "flags": np.zeros(n, np.uint8),
        "center": np.zeros((n, dim)), "size": np.zeros(n), "t": np.zeros(n),
        "p": np.zeros((n, ne)), "u": np.zeros((n, ne)), "slope": np.zeros((n, dim, ne)),
        "rhs": np.zeros((n, ne)), "k": np.zeros((n, npx)), "grad": np.zeros((n, dim, ngx)),
        "meta": np.zeros((n, 4)),
    }


INT_FIELDS = {"parent": C.c_int32, "child_base": C.c_int32, "nbr": C.c_int32, "tree_level": C.c_int8,
              "acti_level": C.c_int8, "cluster": C.c_int8, "flags": C.c_uint8}


def node_view(arrs):
    v = NodeViewStruct()
    for name, _ in NodeViewStruct._fields_:
        setattr(v, name, ptr(arrs[name], INT_FIELDS.get(name, C.c_double)))
    return v
