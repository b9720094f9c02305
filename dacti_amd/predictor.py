"""Learned refinement predictor: data from the solver, features, labels, evaluation.

The solver's own runs are the data: at global step t every active leaf gives
one sample, its features describe its 3^D neighbourhood now, and its label is
what the estimator-driven mesh has done at that place h steps later
(impl/adaptive_refinement.hpp:16-94 decides it in the reference): 0 = the
leaf holding the cell centre is coarser (coarsened), 1 = the same level
(kept), 2 = finer (refined).  The network (dacti_amd.mlp, BF16 MFMA on the
GPU) is trained on these samples; the plain indicator stays the reference
and judge: the missed-refinement rate is the fraction of the samples the
mesh refines within h steps that the predictor does not flag.

Features of a leaf:
  * its primitive state and the differences to the primitive state at the
    3^D - 1 same-size neighbour positions (the leaf containing the centre
    moved by one cell width);
  * its gradient vector (Navier-Stokes and five-equation models);
  * its tree level above the base level and log10 of its error estimate.
The GPU computes them in place (dacti_features: one thread per leaf with a
tree descent per neighbour position); samples, labels (dacti_locate_levels
on later meshes), inference (dacti_mlp_forward_noalloc) and the split flags
(dacti_mark_predicted) stay in device memory.  `features_raw` restates the
kernel on an exported pool for the tests; `features` is the earlier host
form with each variable scaled per snapshot.
"""
import ctypes as C
import itertools

import numpy as np

from . import scene
from .integrator import Integrator

F_LEAF, F_ACTIVE, F_GARBAGE = 1, 2, 4


def _active_leaves(nodes):
    f = nodes["flags"]
    return np.nonzero(((f & F_LEAF) != 0) & ((f & F_ACTIVE) != 0) & ((f & F_GARBAGE) == 0))[0]


def locate(nodes, pts):
    """Index of the leaf containing each point (descent from the root, node
    1, by comparison with the node centres; a point on a face goes to the
    negative side), for points inside the root cell."""
    idx = np.ones(len(pts), np.int64)
    fl, cb, ctr = nodes["flags"], nodes["child_base"], nodes["center"]
    dim = ctr.shape[1]
    for _ in range(40):
        inner = (fl[idx] & F_LEAF) == 0
        if not inner.any():
            break
        q = np.zeros(len(pts), np.int64)
        for d in range(dim):
            q |= (pts[:, d] > ctr[idx, d]).astype(np.int64) << d
        idx = np.where(inner, cb[idx] + q, idx)
    return idx


def features(nodes, case):
    """(leaf ids, float32 features) of the active leaves of one node pool."""
    lv = _active_leaves(nodes)
    dim = case.dim
    p, ctr, size = nodes["p"], nodes["center"], nodes["size"]
    ne = p.shape[1]
    scale = np.maximum(np.abs(p[lv]).mean(axis=0), 1e-6)
    half = np.asarray(case.active_size[:dim]) / 2
    pc = p[lv]
    cols = [pc / scale]
    for off in itertools.product((-1, 0, 1), repeat=dim):
        if not any(off):
            continue
        pts = ctr[lv] + size[lv, None] * np.asarray(off, float)  # size is the full width
        inside = np.ones(len(lv), bool)
        for d in range(dim):
            if case.periodic[d]:
                pts[:, d] = (pts[:, d] + half[d]) % (2 * half[d]) - half[d]
            else:
                inside &= np.abs(pts[:, d]) < half[d]
        nb = locate(nodes, np.where(inside[:, None], pts, ctr[lv]))
        ps = np.where(inside[:, None], p[nb], pc)
        cols.append((ps - pc) / scale)
    if case.model != 0:  # Navier-Stokes and five-equation models carry gradients
        g = nodes["grad"][lv].reshape(len(lv), -1)
        cols.append(g / np.maximum(np.abs(g).mean(axis=0), 1e-6))
    cols.append((nodes["tree_level"][lv, None] - case.min_refinement_level).astype(float))
    cols.append(np.log10(np.maximum(nodes["meta"][lv, 1:2], 1e-12)))
    return lv, np.concatenate(cols, axis=1).astype(np.float32)


def features_raw(nodes, case):
    """(leaf ids, float32 features) as dacti_features computes them (no
    standardisation): primitives, neighbour differences, gradient, level
    above the base level, log10 error."""
    lv = _active_leaves(nodes)
    dim = case.dim
    p, ctr, size = nodes["p"], nodes["center"], nodes["size"]
    half = np.asarray(case.active_size[:dim]) / 2
    pc = p[lv]
    cols = [pc]
    for off in itertools.product((-1, 0, 1), repeat=dim):
        if not any(off):
            continue
        pts = ctr[lv] + size[lv, None] * np.asarray(off, float)
        inside = np.ones(len(lv), bool)
        for d in range(dim):
            if not case.periodic[d]:
                inside &= np.abs(pts[:, d]) < half[d]
            else:
                pts[:, d] = np.where(pts[:, d] < -half[d], pts[:, d] + 2 * half[d],
                                     np.where(pts[:, d] > half[d], pts[:, d] - 2 * half[d], pts[:, d]))
        nb = locate(nodes, np.where(inside[:, None], pts, ctr[lv]))
        cols.append(np.where(inside[:, None], p[nb] - pc, 0.0))
    if case.model != 0:
        ng = nodes["grad"].shape[2]
        cols.append(nodes["grad"][lv].reshape(len(lv), dim * ng))
    cols.append((nodes["tree_level"][lv, None] - case.min_refinement_level).astype(float))
    cols.append(np.log10(np.maximum(nodes["meta"][lv, 1:2], 1e-12)))
    return lv, np.concatenate(cols, axis=1).astype(np.float32)


def _dptr(a):
    return None if a is None else C.c_void_p(a.address)


def gpu_features(g, case, mu=None, sd=None, with_centers=False):
    """dacti_features on integrator g: (ids, features[, centres, levels]) as
    device arrays (dacti_amd.devmem); mu/sd (device float32) standardise."""
    from .devmem import DeviceArray
    n, nf = C.c_int64(), C.c_int32()
    g._check(g.lib.dacti_features(g.h, case.min_refinement_level, None, None, None, None, None, None, 0,
                                  C.byref(n), C.byref(nf)), "features")
    N, F = n.value, nf.value
    feat = DeviceArray((N, F), np.float32)
    ids = DeviceArray(N, np.int32)
    ctr = DeviceArray((N, case.dim), np.float64) if with_centers else None
    lev = DeviceArray(N, np.int8) if with_centers else None
    g._check(g.lib.dacti_features(g.h, case.min_refinement_level, _dptr(mu), _dptr(sd), _dptr(feat), _dptr(ids),
                                  _dptr(ctr), _dptr(lev), N, C.byref(n), C.byref(nf)), "features")
    return (ids, feat, ctr, lev) if with_centers else (ids, feat)


def gpu_locate_levels(g, centers):
    """Tree level of the leaf containing each point (device float64 n x dim), on the host."""
    from .devmem import DeviceArray
    lev = DeviceArray(len(centers), np.int8)
    g._check(g.lib.dacti_locate_levels(g.h, _dptr(centers), len(centers), _dptr(lev)), "locate_levels")
    return lev.numpy()


def gpu_mark(g, logits, ids, rings=0):
    """Flag the leaves whose logits (device n x 3) pick refine, plus rings of
    neighbours, to split at the next remesh; returns the number flagged."""
    n = C.c_int64()
    g._check(g.lib.dacti_mark_predicted(g.h, _dptr(logits), _dptr(ids), len(ids), rings, C.byref(n)),
             "mark_predicted")
    return n.value


def gpu_samples(case, steps, horizon, tree=None, log=None):
    """Step one case on the GPU and collect (features, window labels,
    indicator guess) per global step: features and the later meshes' levels
    at the samples' centres come from the device (dacti_features,
    dacti_locate_levels; no pool export).  The label of a sample from step t
    is 2 if the mesh at its centre is finer at any of the next `horizon`
    steps, 0 if coarser at all of them, 1 otherwise; the guess is the
    estimator's present flag.  Returns host arrays (X, y, guess) per step."""
    g = Integrator(case, tree if tree is not None else scene.build_tree(case))
    lo = np.log10(case.error_threshold[0]) if case.error_threshold[0] > 0 else -np.inf
    hi = np.log10(case.error_threshold[1])
    pending, out = [], []
    for s in range(steps + 1):
        if s > 0:
            _, fb = g.integrate()
            if log:
                log(f"step {s - 1}")
            if fb:
                break
            for w in pending:
                lv_now = gpu_locate_levels(g, w["ctr"]).astype(np.int16)
                w["finer"] |= lv_now > w["lev"]
                w["coarser"] &= lv_now < w["lev"]
                w["age"] += 1
            for w in [w for w in pending if w["age"] == horizon]:
                y = np.where(w["finer"], 2, np.where(w["coarser"], 0, 1)).astype(np.int32)
                out.append((w["X"], y, w["guess"]))
            pending = [w for w in pending if w["age"] < horizon]
        if s + horizon > steps:
            continue
        ids, Xd, ctr, lev = gpu_features(g, case, with_centers=True)
        X = Xd.numpy()
        lerr = X[:, -1].astype(np.float64)
        guess = np.where(lerr > hi, 2, np.where(lerr < lo, 0, 1)).astype(np.int32)
        pending.append({"X": X, "ctr": ctr, "lev": lev.numpy().astype(np.int16),
                        "finer": np.zeros(len(ids), bool), "coarser": np.ones(len(ids), bool),
                        "guess": guess, "age": 0})
    g.close()
    return out


def labels(nodes_t, lv, nodes_h):
    """0 / 1 / 2: the mesh h steps later is coarser / equal / finer at each leaf centre."""
    at = locate(nodes_h, nodes_t["center"][lv])
    d = nodes_h["tree_level"][at].astype(int) - nodes_t["tree_level"][lv].astype(int)
    return (np.sign(d) + 1).astype(np.int32)


def labels_window(snaps, t, lv, h):
    """What a remeshing cadence of h steps needs: 2 if the mesh is finer at
    the leaf centre at any of the next h steps, 0 if it is coarser at all of
    them, 1 otherwise."""
    x = snaps[t]["center"][lv]
    lev = snaps[t]["tree_level"][lv].astype(int)
    finer = np.zeros(len(lv), bool)
    coarser = np.ones(len(lv), bool)
    for k in range(1, h + 1):
        at = locate(snaps[t + k], x)
        d = snaps[t + k]["tree_level"][at].astype(int) - lev
        finer |= d > 0
        coarser &= d < 0
    return np.where(finer, 2, np.where(coarser, 0, 1)).astype(np.int32)


def dilate(nodes, ids, rings=1):
    """ids plus the active leaves reached through `rings` face-neighbour hops."""
    F_MASK = F_LEAF | F_ACTIVE
    cur = np.unique(np.asarray(ids, np.int64))
    out = cur
    for _ in range(rings):
        nb = nodes["nbr"][cur].ravel()
        nb = nb[nb > 0]
        nb = nb[(nodes["flags"][nb] & F_MASK) == F_MASK]
        cur = np.setdiff1d(np.unique(nb), out)
        out = np.union1d(out, cur)
    return out


def indicator_guess(nodes, lv, case):
    """The method without the network: the estimator's present flags taken as
    the forecast (refine if the error is above max_error now, coarsen if
    below min_error, keep otherwise)."""
    err = nodes["meta"][lv, 1]
    lo, hi = case.error_threshold
    return np.where(err > hi, 2, np.where(err < lo, 0, 1)).astype(np.int32)


def run_samples(case, steps, horizon, tree=None, log=None, window=False):
    """Step one case on the GPU; returns per-step (features, labels, indicator
    guess); window=True labels with labels_window (any refinement within the
    horizon) instead of the mesh exactly `horizon` steps later."""
    g = Integrator(case, tree if tree is not None else scene.build_tree(case))
    snaps = [g.get_nodes()]
    for s in range(steps):
        _, fb = g.integrate()
        snaps.append(g.get_nodes())
        if log:
            log(f"step {s}")
        if fb:
            break
    g.close()
    out = []
    for t in range(len(snaps) - horizon):
        lv, X = features(snaps[t], case)
        y = labels_window(snaps, t, lv, horizon) if window else labels(snaps[t], lv, snaps[t + horizon])
        out.append((X, y, indicator_guess(snaps[t], lv, case)))
    return out


def missed_refinement(pred, y):
    """Fraction of the samples labelled refine that pred does not flag as refine."""
    m = y == 2
    return float((pred[m] != 2).mean()) if m.any() else 0.0


class Standardizer:
    def __init__(self, X):
        self.mu = X.mean(axis=0)
        self.sd = np.maximum(X.std(axis=0), 1e-6)

    def __call__(self, X):
        return ((X - self.mu) / self.sd).astype(np.float32)


def run_cadence(case, tree, steps, every=1, model=None, std=None, log=None, rings=0):
    """Step one case with remeshing only in the global steps k with
    k % every == 0 (every = 1: the reference, which remeshes in every
    substep).  With a model, before each remeshing step the leaves it
    forecasts as refined (and, with rings > 0, their neighbours that many face
    hops out) are flagged to split (the estimator still marks as usual: it is
    the fallback).  Returns (final nodes, seconds, cell updates,
    flagged leaves)."""
    import time
    g = Integrator(case, tree)
    flagged = 0
    u0 = g.stats().cell_updates
    t0 = time.perf_counter()
    for k in range(steps):
        on = k % every == 0
        if on and model is not None:
            nodes = g.get_nodes()
            lv, X = features(nodes, case)
            ids = lv[model.predict(std(X)) == 2]
            if rings:
                ids = dilate(nodes, ids, rings)
            g.mark_split(ids)
            flagged += len(ids)
        g.set_remesh(on)
        _, fb = g.integrate()
        if log:
            log(f"step {k}")
        if fb:
            raise RuntimeError("NaN state (force_break)")
    el = time.perf_counter() - t0
    ups = g.stats().cell_updates - u0
    nodes = g.get_nodes()
    g.close()
    return nodes, el, ups, flagged


def run_cadence_gpu(case, tree, steps, every=1, model=None, mu=None, sd=None, log=None, rings=0):
    """run_cadence with the predictor in device memory: before each
    remeshing step dacti_features (standardised with mu, sd: device float32),
    the MLP's dacti_mlp_forward_noalloc and dacti_mark_predicted; nothing
    crosses to the host.  Returns (final nodes, seconds, cell updates,
    flagged leaves)."""
    import time
    from .devmem import DeviceArray
    g = Integrator(case, tree)
    flagged = 0
    u0 = g.stats().cell_updates
    t0 = time.perf_counter()
    for k in range(steps):
        on = k % every == 0
        if on and model is not None:
            ids, X = gpu_features(g, case, mu, sd)
            logits = DeviceArray((len(ids), 3), np.float32)
            if len(ids):
                model.forward_noalloc(X.address, len(ids), logits.address)
            flagged += gpu_mark(g, logits, ids, rings)
            for a in (ids, X, logits):
                a.free()
        g.set_remesh(on)
        _, fb = g.integrate()
        if log:
            log(f"step {k}")
        if fb:
            raise RuntimeError("NaN state (force_break)")
    el = time.perf_counter() - t0
    ups = g.stats().cell_updates - u0
    nodes = g.get_nodes()
    g.close()
    return nodes, el, ups, flagged


def relative_l1(ref, other):
    """Relative L1 difference of the conserved variables of `other` against
    `ref`, sampled at ref's active leaf centres and weighted by their volume."""
    lv = _active_leaves(ref)
    at = locate(other, ref["center"][lv])
    vol = ref["size"][lv] ** ref["center"].shape[1]
    num = (np.abs(other["u"][at] - ref["u"][lv]).sum(axis=1) * vol).sum()
    den = (np.abs(ref["u"][lv]).sum(axis=1) * vol).sum()
    return float(num / den)
