"""Initial tree of a case: Scene<M>::init_domain (src/lib/scene/scene.cpp:369-431).

Cases with objects ([[geometry]]: built-in shapes, STL meshes, fluid
objects and solids with their boundary conditions) are built by the C++
implementation behind the C ABI (dacti_scene_build, include/dacti/config),
from the case's TOML text.  Cases without objects are grown here, level by
level as the reference does (boundaries, split detection, split),
vectorised over each level's new nodes; the C++ builder gives the same
tree bit for bit.  Children of one split occupy consecutive indices and
split parents are taken in node order.  Scope here: domain boundary
conditions, periodic axes, uniform / jump / grid initial conditions and
axis-aligned refinement zones.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import F_ACTIVE, F_CAN_SPLIT, F_CAN_UNSPLIT, F_LEAF, RECORD_DTYPE
from .case import BC_TYPES


@dataclass
class Tree:
    n: int
    parent: np.ndarray      # int32 [n]
    child_base: np.ndarray  # int32 [n]
    level: np.ndarray       # int8  [n]
    flags: np.ndarray       # uint8 [n]
    bc: np.ndarray          # uint8 [n, 2*dim]
    center: np.ndarray      # f64   [n, dim]
    size: np.ndarray        # f64   [n]
    p: np.ndarray           # f64   [n, num_eqns]
    brec: np.ndarray = None     # int32 [n]: object boundary record of a node's object faces (-1 none)
    records: np.ndarray = None  # RECORD_DTYPE [n_records]

    def leaves(self):
        f = self.flags
        return np.nonzero((f & F_LEAF).astype(bool) & (f & F_ACTIVE).astype(bool))[0]


def _quadrant(D):
    """vecn_t::quadrant(i) for i in 0..2^D-1 (core/linalg.hpp:68-83)."""
    q = np.empty((1 << D, D))
    for i in range(1 << D):
        for d in range(D):
            q[i, d] = 1.0 if (i >> d) & 1 else -1.0
    return q


def build_tree(case):
    if case.geometry:
        return _build_tree_cpp(case)
    D = case.dim
    NE = case.num_eqns
    NCH = 1 << D
    quad = _quadrant(D)
    # growable columns
    parent = [np.zeros(2, np.int32)]
    child_base = [np.zeros(2, np.int32)]
    level = [np.array([-1, 0], np.int8)]
    center = [np.zeros((2, D))]
    size = [np.array([0.0, case.side_length])]
    leaf = [np.array([False, True])]
    can_unsplit = [np.array([False, False])]
    n = 2
    current = np.array([1], np.int64)
    # per-level state arrays are kept in flat form after each level
    P = np.concatenate(parent)
    CB = np.concatenate(child_base)
    LV = np.concatenate(level)
    CT = np.concatenate(center)
    SZ = np.concatenate(size)
    LF = np.concatenate(leaf)
    CU = np.concatenate(can_unsplit)
    BCA = np.zeros((2, 2 * D), np.uint8)
    ACT = np.array([False, True])
    inflow_r = BC_TYPES["inflow_r"]
    while current.size:
        c = CT[current]
        s = SZ[current]
        lev = LV[current].astype(np.int64)
        # compute_node_boundaries (scene.cpp:88-147): domain faces only
        has_open = np.zeros(current.size, bool)
        bcl = np.zeros((current.size, 2 * D), np.uint8)
        for dim in range(D):
            for dr in range(2):
                sgn = float(2 * dr - 1)
                nc = c[:, dim] + sgn * s
                if not case.periodic[dim]:
                    on_b = nc * sgn > case.active_size[dim] / 2.0
                else:
                    on_b = np.zeros(current.size, bool)
                bcl[on_b, 2 * dim + dr] = 1 + 2 * dim + dr
                has_open |= ~on_b
        BCA[current] = bcl
        ACT[current] = np.where(has_open, ACT[current], False)
        # detect_splits (scene.cpp:150-214), statement order kept
        ss = np.zeros(current.size, bool)
        cu = CU[current].copy()
        m = lev < case.min_refinement_level
        ss |= m
        cu[m] = False
        for f in range(2 * D):
            if case.bcs[f].type == inflow_r:
                m = (bcl[:, f] != 0) & (lev < case.max_refinement_level)
                ss |= m
                cu[m] = True
        if case.ic_type == "jump":
            m = (np.abs(c[:, case.ic_axis] - case.ic_location) < s) & (lev < case.max_refinement_level)
            ss |= m
            cu[m] = True
        for (x0, x1, zl) in case.zones:
            x3 = np.zeros((current.size, 3))
            x3[:, :D] = c
            inside = np.all((x3 < np.asarray(x1)) & (x3 > np.asarray(x0)), axis=1)
            ss |= inside & (lev < zl)
        CU[current] = cu
        # split_detected_nodes (scene.cpp:217-247), parents in node order
        sp = np.sort(current[ss & LF[current]])
        if sp.size == 0:
            break
        m_new = sp.size * NCH
        base = n + NCH * np.arange(sp.size, dtype=np.int64)
        CB[sp] = base.astype(np.int32)
        LF[sp] = False
        ps = SZ[sp]
        pc = CT[sp]
        q = 0.25 * ps
        cc = (pc[:, None, :] + quad[None, :, :] * q[:, None, None]).reshape(m_new, D)
        P = np.concatenate([P, np.repeat(sp, NCH).astype(np.int32)])
        CB = np.concatenate([CB, np.zeros(m_new, np.int32)])
        LV = np.concatenate([LV, np.repeat(LV[sp] + 1, NCH).astype(np.int8)])
        CT = np.concatenate([CT, cc])
        SZ = np.concatenate([SZ, np.repeat(0.5 * ps, NCH)])
        LF = np.concatenate([LF, np.ones(m_new, bool)])
        CU = np.concatenate([CU, np.repeat(CU[sp], NCH)])
        BCA = np.concatenate([BCA, np.zeros((m_new, 2 * D), np.uint8)])
        ACT = np.concatenate([ACT, np.ones(m_new, bool)])
        current = np.arange(n, n + m_new, dtype=np.int64)
        n += m_new
    ACT[1:] = np.where(LF[1:], True, ACT[1:])
    # fix_boundary (scene.cpp:250-335) only changes trees with solid objects.
    # cells outside the active box (scene.cpp:407-415)
    outside = np.zeros(n, bool)
    for dim in range(D):
        outside |= (CT[:, dim] > 0.5 * case.active_size[dim]) | (CT[:, dim] < -0.5 * case.active_size[dim])
    ACT[outside] = False
    ACT[0] = False
    p = np.zeros((n, NE))
    al = np.nonzero(ACT & LF)[0]
    p[al] = initial_state(case, CT[al])
    flags = (LF.astype(np.uint8) * F_LEAF) | (ACT.astype(np.uint8) * F_ACTIVE) | np.uint8(F_CAN_SPLIT) | \
        (CU.astype(np.uint8) * F_CAN_UNSPLIT)
    flags[0] = 0
    return Tree(n=n, parent=np.ascontiguousarray(P, np.int32), child_base=np.ascontiguousarray(CB, np.int32),
                level=np.ascontiguousarray(LV, np.int8), flags=np.ascontiguousarray(flags, np.uint8),
                bc=np.ascontiguousarray(BCA, np.uint8), center=np.ascontiguousarray(CT),
                size=np.ascontiguousarray(SZ), p=np.ascontiguousarray(p))


def initial_state(case, x):
    """Config::ic.x2p (config/config.cpp:194-243) at cell centres x."""
    NE = case.num_eqns
    if case.ic_type == "uniform":
        # "fixed" also returns p_ref in the reference (config.cpp:208-211)
        return np.tile(np.asarray(case.p_ref, np.float64), (x.shape[0], 1))
    if case.ic_type == "jump":
        pos = x[:, case.ic_axis] > case.ic_location
        return np.where(pos[:, None], np.asarray(case.ic_pos), np.asarray(case.ic_neg))
    if case.ic_type == "grid":
        g = np.asarray(case.ic_grid)
        D = case.dim
        shape = g.shape[:D]
        idx = []
        for d in range(D):
            h = case.active_size[d] / shape[d]
            i = np.rint((x[:, d] + 0.5 * case.active_size[d]) / h - 0.5).astype(np.int64)
            idx.append(np.clip(i, 0, shape[d] - 1))
        return g[tuple(idx)].reshape(x.shape[0], NE)
    raise ValueError(case.ic_type)


def _build_tree_cpp(case):
    """The tree of a case with objects, from its TOML text, through
    dacti_scene_build (include/dacti_gpu.h)."""
    text = getattr(case, "toml_text", None)
    if text is None:
        raise ValueError("a case with [[geometry]] objects is built from its TOML text (load_case / loads_case)")
    lib = _lib.load()
    grid = shape = None
    if case.ic_type == "grid":
        g = np.ascontiguousarray(case.ic_grid, np.float64)
        grid = g.ctypes.data
        shape = (C.c_int32 * 3)(*(list(g.shape[:case.dim]) + [0] * (3 - case.dim)))
    h = C.c_void_p()
    rc = lib.dacti_scene_build(text.encode(), str(case.base_dir).encode(), case.model, case.dim, case.limiter, grid,
                               shape, C.byref(h))
    if rc != 0:
        raise ValueError(lib.dacti_last_error(None).decode())
    try:
        t = _lib.TreeStruct()
        lib.dacti_scene_tree(h, C.byref(t))
        n, D, NE = t.n_nodes, case.dim, case.num_eqns

        def arr(ptr, dtype, cols=1):
            a = np.ctypeslib.as_array(ptr, shape=(n * cols,)).copy().astype(dtype, copy=False)
            return a.reshape(n, cols) if cols > 1 else a

        recs = np.zeros(t.n_records, RECORD_DTYPE)
        if t.n_records:
            C.memmove(recs.ctypes.data, t.records, RECORD_DTYPE.itemsize * t.n_records)
        return Tree(n=n, parent=arr(t.parent, np.int32), child_base=arr(t.child_base, np.int32),
                    level=arr(t.tree_level, np.int8), flags=arr(t.flags, np.uint8), bc=arr(t.bc, np.uint8, 2 * D),
                    center=arr(t.center, np.float64, D), size=arr(t.size, np.float64),
                    p=arr(t.p, np.float64, NE), brec=arr(t.brec, np.int32), records=recs)
    finally:
        lib.dacti_scene_destroy(h)
