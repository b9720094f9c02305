"""Substeps planned on the device (fixed grids): one wait per substep.

With a persistent grid the whole level substep runs without the host
between its phases: the deferred sweep, the remesh plan (which instances
remesh, the pool size after the splits, the tile tables of the remesh
phases and the grown pools) and the remesh phases act on the device's own
counts.  Results must not change: every case is compared with the oracle
after each global step, node pool bitwise, including instances whose state
breaks down (they stop in the same global step as the oracle's
integrate(), impl/integrate.hpp:50, and the others go on)."""
import pytest

from dacti_amd import scene, workloads
from dacti_amd.integrator import Batch, Integrator
from oracle.binding import Oracle
from tests.helpers import compare_pools

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("graphs", [False, True])
def test_device_step_shock_bubble_amr(graphs):
    case = workloads.shock_bubble(base=4, extra=2)
    tree = scene.build_tree(case)
    o, g = Oracle(case, tree), Integrator(case, tree)
    g.set_launch(block=128, grid=96, graphs=graphs)
    for s in range(4):
        print(f"[device step] shock-bubble graphs={graphs}: step {s}", flush=True)
        dto, fbo = o.integrate()
        dtg, fbg = g.integrate()
        assert dto == dtg and fbo == fbg
        compare_pools(g.get_nodes(), o.export())
    assert g.stats().cell_updates == o.cell_updates and g.stats().n_substeps == o.n_substeps
    g.close()
    o.close()


def test_device_step_shock_drop_3d():
    case = workloads.shock_drop(base=3, extra=2)
    tree = scene.build_tree(case)
    o, g = Oracle(case, tree), Integrator(case, tree)
    g.set_launch(block=64, grid=128, graphs=True)
    for s in range(3):
        print(f"[device step] shock-drop: step {s}", flush=True)
        dto, fbo = o.integrate()
        dtg, fbg = g.integrate()
        assert dto == dtg and fbo == fbg
        compare_pools(g.get_nodes(), o.export())
    g.close()
    o.close()


def test_device_step_batch_with_breakdown():
    # seed 0, 16x8 base + 2-3 levels: instance 3 breaks down in global step 3
    fam = workloads.shock_bubble_family(5, seed=0, base=4, extra_range=(2, 3))
    pairs = []
    for k, (case, _) in enumerate(fam):
        print(f"[device step] breakdown batch: instance {k} tree", flush=True)
        tree = scene.build_tree(case)
        pairs.append((Oracle(case, tree), Integrator(case, tree)))
    broken = [False] * len(pairs)
    with Batch([g for _, g in pairs]) as b:
        b.set_launch(block=128, grid=128, graphs=True)
        for s in range(5):
            print(f"[device step] breakdown batch: step {s}", flush=True)
            fbs = [None if broken[k] else o.integrate()[1] for k, (o, _) in enumerate(pairs)]
            assert b.integrate() == any(f for f in fbs if f is not None)
            for k, (o, g) in enumerate(pairs):
                if broken[k]:
                    continue
                compare_pools(g.get_nodes(), o.export())
                broken[k] = bool(fbs[k])
    assert broken == [False, False, False, True, False]
    for o, g in pairs:
        g.close()
        o.close()


def test_device_step_gc_and_pool_growth():
    # everything coarsens in the first step, so the pool is compacted after
    # it (gc re-bins the levels); with the device-planned substeps the pool
    # first grows to hold every listed cell's split
    from tests.helpers import amr_sod_case
    case = amr_sod_case(base=3, top=8, emin=1e10, emax=1e20)
    tree = scene.build_tree(case)
    n0 = tree.n
    o, g = Oracle(case, tree), Integrator(case, tree)
    g.set_launch(block=64, grid=64, graphs=True)
    for s in range(3):
        print(f"[device step] gc case: step {s}", flush=True)
        o.integrate()
        g.integrate()
        compare_pools(g.get_nodes(), o.export())
    assert g.n_nodes() < n0 // 4  # compacted
    g.close()
    o.close()


def test_device_step_refining_sod_grows_pool():
    # a jump refined from a coarse base: splits every substep, so the pool
    # grows past its initial capacity while the substeps are planned on the
    # device
    from tests.helpers import amr_sod_case
    case = amr_sod_case(base=4, top=8)
    tree = scene.build_tree(case)
    o, g = Oracle(case, tree), Integrator(case, tree)
    g.set_launch(block=128, grid=32, graphs=False)
    for s in range(4):
        print(f"[device step] refining sod: step {s}", flush=True)
        o.integrate()
        g.integrate()
        compare_pools(g.get_nodes(), o.export())
    assert g.n_nodes() > tree.n
    g.close()
    o.close()
