"""Batched stepping: each instance of a batch against its own oracle run.

A batch launches every op of a substep once for all instances; each
instance must still evolve exactly as it would alone, so its node pool is
compared with the oracle's after every global step (topology, LTS levels,
clusters and flags exactly; conserved variables to 1e-11 relative L1).  The
launch configuration (tile size, persistent grid, HIP-graph replay) must
not change any result; graphs are replayed many times with eager host work
and synchronisations in between, each replay checked against the oracle.
"""
import pytest

from dacti_amd import scene, workloads
from dacti_amd.integrator import Batch, Integrator
from oracle.binding import Oracle
from tests.helpers import amr_sod_case, compare_pools

pytestmark = pytest.mark.gpu


def _members(cases, label):
    out = []
    for k, case in enumerate(cases):
        print(f"[batch] {label}: instance {k} tree", flush=True)
        tree = scene.build_tree(case)
        out.append((Oracle(case, tree), Integrator(case, tree)))
    return out


def _check_all(pairs, label, step):
    for k, (o, g) in enumerate(pairs):
        print(f"[batch] {label}: step {step} instance {k} compare", flush=True)
        compare_pools(g.get_nodes(), o.export())


def _step(pairs, batch):
    fbs = [o.integrate()[1] for o, _ in pairs]
    assert not any(fbs)
    assert batch.integrate() is False


def test_batch_shock_bubbles_match_single_runs():
    fam = workloads.shock_bubble_family(3, seed=7, base=4, extra_range=(1, 2))
    pairs = _members([c for c, _ in fam], "shock-bubble x3")
    with Batch([g for _, g in pairs]) as b:
        for s in range(3):
            _step(pairs, b)
            _check_all(pairs, "shock-bubble x3", s)
        # persistent grid and graph replay from here on
        b.set_launch(block=128, grid=64, graphs=True)
        assert b.launch() == (128, 64, True)
        for s in range(3, 6):
            _step(pairs, b)
            _check_all(pairs, "shock-bubble x3 (graphs)", s)
        st = b.stats()
        assert st.cell_updates == sum(o.cell_updates for o, _ in pairs)
    for o, g in pairs:
        assert o.acc_violations == 0
        assert g.stats().cell_updates == o.cell_updates
        g.close()


def test_batch_uneven_euler_boxes_graphs():
    cases = [workloads.periodic_box(level=lv, dim=2, seed=s) for lv, s in ((4, 1), (6, 2), (5, 3), (3, 4))]
    pairs = _members(cases, "euler2d boxes")
    with Batch([g for _, g in pairs]) as b:
        b.set_launch(block=64, grid=256, graphs=True)
        for s in range(4):
            _step(pairs, b)
            _check_all(pairs, "euler2d boxes", s)
    for o, g in pairs:
        g.close()


def test_batch_autotune_keeps_results():
    cases = [workloads.periodic_box(level=lv, dim=2, seed=s) for lv, s in ((5, 5), (4, 6))]
    pairs = _members(cases, "autotune")
    with Batch([g for _, g in pairs]) as b:
        b.autotune()
        res = b.tune_results()
        for s in range(len(res) + 1):
            _step(pairs, b)
            _check_all(pairs, "autotune", s)
        res = b.tune_results()
        assert len(res) >= 4 and all(ms > 0 for *_, ms in res)
        best = min(res, key=lambda r: r[3])
        assert b.launch() == best[:3]
    for o, g in pairs:
        g.close()


def test_batch_small_grid_spans_instances():
    # a persistent grid much smaller than the tile count: one block walks
    # tiles of several instances (pool histogram, selection, remesh phases)
    cases = [workloads.periodic_box(level=lv, dim=2, seed=s) for lv, s in ((7, 8), (6, 9), (7, 10))]
    cases.append(amr_sod_case(base=5, top=7))  # Euler 2D with split and unsplit
    pairs = []
    for k, case in enumerate(cases):
        print(f"[batch] small grid: instance {k} tree", flush=True)
        tree = scene.build_tree(case)
        pairs.append((Oracle(case, tree), Integrator(case, tree)))
    with Batch([g for _, g in pairs]) as b:
        b.set_launch(block=64, grid=8, graphs=False)
        for s in range(2):
            _step(pairs, b)
            _check_all(pairs, "small grid", s)
        b.set_launch(block=256, grid=16, graphs=True)
        for s in range(2, 4):
            _step(pairs, b)
            _check_all(pairs, "small grid (graphs)", s)
    for o, g in pairs:
        g.close()
