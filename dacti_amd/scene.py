"""Initial tree of a case: Scene<M>::init_domain (src/lib/scene/scene.cpp:369-431).

The tree is grown level by level as the reference does (compute signed
distances, boundaries, split detection, split), vectorised over each
level's new nodes.  Children of one split occupy consecutive indices and
split parents are taken in node order.  Scope: domain boundary conditions,
periodic axes, uniform / jump / grid initial conditions, analytic
fluid objects (circle, sphere) of type "bubble" or "droplet", and
axis-aligned refinement zones.  Solid geometry (libigl signed distance,
embree ray casts, scene.cpp:47-147) is outside this engine's scope.
"""
from dataclasses import dataclass

import numpy as np

from ._lib import (F_ACTIVE, F_BOUNDARY, F_CAN_SPLIT, F_CAN_UNSPLIT, F_LEAF, F_SHOULD_SPLIT, MODEL_MULTI)
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


def _signed_distance(case, center):
    """Nearest analytic object: (distance, object index); +max without objects."""
    n = center.shape[0]
    sd = np.full(n, np.finfo(np.float64).max)
    near = np.zeros(n, np.int64)
    for mi, g in enumerate(case.geometry):
        c = np.asarray(g.center[:case.dim], dtype=np.float64)
        dist = np.sqrt(((center - c) ** 2).sum(axis=1)) - g.radius
        better = dist < sd
        sd = np.where(better, dist, sd)
        near = np.where(better, mi, near)
    return sd, near


def build_tree(case):
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
    sdist = [np.array([-1.0, 0.0])]
    nearest = [np.zeros(2, np.int64)]
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
    SD = np.concatenate(sdist)
    NR = np.concatenate(nearest)
    BCA = np.zeros((2, 2 * D), np.uint8)
    ACT = np.array([False, True])
    inflow_r = BC_TYPES["inflow_r"]
    while current.size:
        c = CT[current]
        s = SZ[current]
        lev = LV[current].astype(np.int64)
        # compute_signed_distances (scene.cpp:47-85)
        sd, near = _signed_distance(case, c)
        SD[current] = sd
        NR[current] = near
        if case.geometry:
            ACT[current] = True  # fluid objects carry no boundary condition
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
        if case.geometry:
            m = (np.abs(sd) < s * case.refinement_radius) & (lev < case.max_refinement_level)
            ss |= m
            cu[m] = True
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
        SD = np.concatenate([SD, np.zeros(m_new)])
        NR = np.concatenate([NR, np.zeros(m_new, np.int64)])
        BCA = np.concatenate([BCA, np.zeros((m_new, 2 * D), np.uint8)])
        ACT = np.concatenate([ACT, np.ones(m_new, bool)])
        current = np.arange(n, n + m_new, dtype=np.int64)
        n += m_new
    if not case.geometry:
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
    if case.model == MODEL_MULTI and D > 1 and case.geometry:
        _init_interface(case, p, al, SD, NR, NE)
    flags = (LF.astype(np.uint8) * F_LEAF) | (ACT.astype(np.uint8) * F_ACTIVE) | np.uint8(F_CAN_SPLIT) | \
        (CU.astype(np.uint8) * F_CAN_UNSPLIT)
    flags[0] = 0
    _ = F_SHOULD_SPLIT, F_BOUNDARY
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


def _init_interface(case, p, leaves, SD, NR, NE):
    """Scene::init_interface (scene.cpp:340-366); diffuse_ic is never read
    from the TOML (config.cpp:198-199), so the interface is sharp."""
    rho = case.rho_ref
    for mi, g in enumerate(case.geometry):
        sel = leaves[NR[leaves] == mi]
        sd = SD[sel]
        if g.type == "droplet":
            alpha = np.where(sd < 0, 1.0, 0.0)
        elif g.type == "bubble":
            alpha = np.where(sd > 0, 1.0, 0.0)
        else:
            continue
        p[sel, 0] = rho[0] * alpha
        p[sel, 1] = rho[1] * (1.0 - alpha)
        p[sel, NE - 1] = alpha
