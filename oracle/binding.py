"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
PREFIX = {(0, 2): "o_euler2_", (0, 3): "o_euler3_", (1, 2): "o_ns2_", (1, 3): "o_ns3_", (2, 2): "o_multi2_",
          (2, 3): "o_multi3_"}


class OParams(C.Structure):
    """struct o_params of oracle_common.h"""
    _fields_ = [
        ("model", C.c_int32), ("dim", C.c_int32), ("limiter", C.c_int32), ("max_refinement_level", C.c_int32),
        ("side_length", C.c_double), ("active_size", C.c_double * 3), ("periodic", C.c_int32 * 3),
        ("n_zones", C.c_int32),
        ("max_cfl", C.c_double), ("max_dt", C.c_double), ("min_error", C.c_double), ("max_error", C.c_double),
        ("weights", C.c_double * 8),
        ("gamma", C.c_double), ("ns_mu", C.c_double), ("ns_lambda", C.c_double), ("ns_cp", C.c_double),
        ("ns_R", C.c_double),
        ("sigma", C.c_double), ("gamma1", C.c_double), ("gamma2", C.c_double), ("P1", C.c_double),
        ("P2", C.c_double), ("cp1", C.c_double), ("cp2", C.c_double), ("cv1", C.c_double), ("cv2", C.c_double),
        ("mu1", C.c_double), ("mu2", C.c_double), ("lambda1", C.c_double), ("lambda2", C.c_double),
        ("bc_type", C.c_int32 * 6), ("bc_p", (C.c_double * 8) * 6), ("bc_fade_in", C.c_double * 6),
        ("zone_x0", (C.c_double * 3) * 8), ("zone_x1", (C.c_double * 3) * 8), ("zone_level", C.c_int32 * 8),
    ]


CONST_NAMES = ("gamma", "ns_mu", "ns_lambda", "ns_cp", "ns_R", "sigma", "gamma1", "gamma2", "P1", "P2", "cp1",
               "cp2", "cv1", "cv2", "mu1", "mu2", "lambda1", "lambda2")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def params(case):
    """OParams from a case object (attributes as in dacti_amd.case.Case)."""
    P = OParams()
    P.model, P.dim, P.limiter = case.model, case.dim, case.limiter
    P.max_refinement_level = case.max_refinement_level
    P.side_length = case.side_length
    for i in range(3):
        P.active_size[i] = case.active_size[i] if i < case.dim else 0.0
        P.periodic[i] = int(case.periodic[i]) if i < case.dim else 0
    P.n_zones = len(case.zones)
    P.max_cfl, P.max_dt = case.max_cfl, case.max_dt
    P.min_error, P.max_error = case.error_threshold
    for i, w in enumerate(case.weights):
        P.weights[i] = w
    for name in CONST_NAMES:
        setattr(P, name, float(case.consts[name]))
    for f in range(6):
        P.bc_type[f] = case.bcs[f].type
        for e, v in enumerate(case.bcs[f].value):
            P.bc_p[f][e] = v
        P.bc_fade_in[f] = case.bcs[f].fade_in
    for z, (x0, x1, lev) in enumerate(case.zones):
        for a in range(3):
            P.zone_x0[z][a] = x0[a]
            P.zone_x1[z][a] = x1[a]
        P.zone_level[z] = lev
    return P


# object boundary record (dacti_boundary of include/dacti_gpu.h)
RECORD_DTYPE = np.dtype([("type", np.int32), ("pad", np.int32), ("normal", np.float64, 3), ("p", np.float64, 8),
                         ("fade_in", np.float64)])


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """One oracle integrator on an initial tree (arrays as in dacti_amd.scene.Tree)."""

    def __init__(self, case, tree):
        self.case = case
        self.pfx = PREFIX[(case.model, case.dim)]
        L = lib()
        self.L = L
        self.P = params(case)
        f = getattr(L, self.pfx + "create")
        f.restype = C.c_void_p
        f.argtypes = [C.POINTER(OParams), C.c_int64] + [C.c_void_p] * 9 + [C.c_int64, C.c_void_p]
        arrs = [np.ascontiguousarray(tree.parent, np.int32), np.ascontiguousarray(tree.child_base, np.int32),
                np.ascontiguousarray(tree.level, np.int8), np.ascontiguousarray(tree.flags, np.uint8),
                np.ascontiguousarray(tree.bc, np.uint8), np.ascontiguousarray(tree.center, np.float64),
                np.ascontiguousarray(tree.size, np.float64), np.ascontiguousarray(tree.p, np.float64)]
        brec = getattr(tree, "brec", None)
        brec = np.full(tree.n, -1, np.int32) if brec is None else np.ascontiguousarray(brec, np.int32)
        recs = getattr(tree, "records", None)
        recs = np.zeros(0, RECORD_DTYPE) if recs is None else np.ascontiguousarray(recs, RECORD_DTYPE)
        self._keep = (arrs, brec, recs)
        self.h = f(C.byref(self.P), tree.n, *[a.ctypes.data for a in arrs], brec.ctypes.data, len(recs),
                   recs.ctypes.data if len(recs) else None)
        self._fn("integrate", C.c_double, [C.c_void_p, C.c_int, C.POINTER(C.c_int)])
        self._fn("run", C.c_int64, [C.c_void_p, C.c_int64, C.c_int])
        for nm, rt in (("n_nodes", C.c_int64), ("cell_updates", C.c_int64), ("n_substeps", C.c_int64),
                       ("global_dt", C.c_double), ("max_acti_level", C.c_int), ("acc_violations", C.c_int64)):
            self._fn(nm, rt, [C.c_void_p])
        self._fn("destroy", None, [C.c_void_p])
        self._fn("export_nodes", None, [C.c_void_p] + [C.c_void_p] * 17)
        self._fn("face_cover", None, [C.c_void_p, C.c_void_p])
        self._fn("set_no_gc", None, [C.c_void_p, C.c_int])
        self._fn("gc_now", None, [C.c_void_p])
        self._fn("export_bc", None, [C.c_void_p, C.c_void_p, C.c_void_p])

    def _fn(self, name, res, args):
        f = getattr(self.L, self.pfx + name)
        f.restype = res
        f.argtypes = args
        setattr(self, "_" + name, f)

    def integrate(self, signal=True):
        fb = C.c_int(0)
        dt = self._integrate(self.h, int(signal), C.byref(fb))
        return dt, bool(fb.value)

    def run(self, steps, signal=True):
        return self._run(self.h, steps, int(signal))

    @property
    def n_nodes(self):
        return self._n_nodes(self.h)

    @property
    def cell_updates(self):
        return self._cell_updates(self.h)

    @property
    def n_substeps(self):
        return self._n_substeps(self.h)

    @property
    def global_dt(self):
        return self._global_dt(self.h)

    @property
    def max_acti_level(self):
        return self._max_acti_level(self.h)

    @property
    def acc_violations(self):
        return self._acc_violations(self.h)

    def export_bc(self):
        """(bc codes [n, 2D], object record of each node [n])."""
        n, D = self.n_nodes, self.case.dim
        bc = np.zeros((n, 2 * D), np.uint8)
        brec = np.zeros(n, np.int32)
        self._export_bc(self.h, bc.ctypes.data, brec.ctypes.data)
        return bc, brec

    def export(self):
        from_sizes = sizes(self.case.model, self.case.dim)
        D, NE, _, NP, NG = from_sizes
        n = self.n_nodes
        a = {
            "parent": np.zeros(n, np.int32), "child_base": np.zeros(n, np.int32),
            "nbr": np.zeros((n, 2 * D), np.int32), "tree_level": np.zeros(n, np.int8),
            "acti_level": np.zeros(n, np.int8), "cluster": np.zeros(n, np.int8), "flags": np.zeros(n, np.uint8),
            "center": np.zeros((n, D)), "size": np.zeros(n), "t": np.zeros(n), "p": np.zeros((n, NE)),
            "u": np.zeros((n, NE)), "slope": np.zeros((n, D, NE)), "rhs": np.zeros((n, NE)),
            "k": np.zeros((n, max(NP, 1))), "grad": np.zeros((n, D, max(NG, 1))), "meta": np.zeros((n, 4)),
        }
        order = ["parent", "child_base", "nbr", "tree_level", "acti_level", "cluster", "flags", "center", "size",
                 "t", "p", "u", "slope", "rhs", "k", "grad", "meta"]
        self._export_nodes(self.h, *[a[k].ctypes.data for k in order])
        return a

    def face_cover(self):
        """Time coverage of every face over the last global step (n x 2D)."""
        out = np.zeros((self.n_nodes, 2 * self.case.dim))
        self._face_cover(self.h, out.ctypes.data)
        return out

    def set_no_gc(self, on=True):
        self._set_no_gc(self.h, int(on))

    def gc_now(self):
        self._gc_now(self.h)

    def close(self):
        if self.h:
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sizes(model, dim):
    ne = dim + 4 if model == 2 else dim + 2
    nf = ne + dim + 1 if model == 2 else ne
    npar = dim + 4 if model == 2 else 0
    ng = 0 if model == 0 else (dim + 1 if model == 1 else dim + 3)
    return dim, ne, nf, npar, ng


def model_fn(model, dim, name, restype, argtypes):
    f = getattr(lib(), PREFIX[(model, dim)] + name)
    f.restype = restype
    f.argtypes = argtypes
    return f
