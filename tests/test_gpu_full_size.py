"""The 3D periodic Euler box at the sizes it is benchmarked on, against the oracle.

The 256^3 box (16,777,216 leaves, LTS levels 0-3) and the 128^3 box run
global steps on the GPU (default launch: the window ops for the dense
clusters, the per-cell ops for the others, list tiles over the whole pool)
and in the oracle from the same tree; after each global step the whole node
pool must be bitwise equal.  A size-independent property is checked as well:
with periodic boundaries and one tree level every face flux leaves one cell
and enters its neighbour within the global step (fluxes toward slower LTS
neighbours are held until the receiver updates, and every level ends the
global step synchronised), so the totals of the conserved variables over the
leaves do not change beyond rounding (relative 1e-12 for mass and energy,
1e-12 of the total absolute momentum for momentum).
"""
import sys
import time

import numpy as np
import pytest

from dacti_amd import scene, workloads
from dacti_amd.integrator import Integrator
from oracle.binding import Oracle
from tests.helpers import compare_pools, live_leaves

pytestmark = pytest.mark.gpu


def _totals(a):
    lv = live_leaves(a)
    vol = a["size"][lv] ** 3
    u = a["u"][lv]
    tot = (u * vol[:, None]).sum(axis=0)
    scale = (np.abs(u) * vol[:, None]).sum(axis=0)
    return tot, scale


def _log(t0, msg):
    print(f"[full size] {msg} ({time.perf_counter() - t0:.1f} s)", flush=True)


@pytest.mark.parametrize("level,steps", [(7, 2), (8, 1)])
def test_periodic_box_full_size(level, steps):
    t0 = time.perf_counter()
    case = workloads.periodic_box(level=level)
    tree = scene.build_tree(case)
    _log(t0, f"{1 << level}^3 box: tree of {tree.n} nodes")
    g = Integrator(case, tree)
    a = g.get_nodes()
    tot0, scale = _totals(a)
    del a
    o = Oracle(case, tree)
    del tree
    _log(t0, "integrator and oracle created")
    for s in range(steps):
        dtg, fbg = g.integrate()
        _log(t0, f"gpu step {s}")
        dto, fbo = o.integrate()
        _log(t0, f"oracle step {s}")
        assert dto == dtg and not fbg and not fbo
        a = g.get_nodes()
        rep = compare_pools(a, o.export())
        assert rep["u_bitwise"], rep
        _log(t0, f"step {s}: node pools bitwise equal")
        tot, _ = _totals(a)
        del a
        d = np.abs(tot - tot0)
        print(f"[full size] step {s}: conserved totals change {d} of {scale}", flush=True)
        assert d[0] <= 1e-12 * abs(tot0[0]) and d[-1] <= 1e-12 * abs(tot0[-1])
        assert (d[1:-1] <= 1e-12 * scale[1:-1]).all()
    st = g.stats()
    assert st.cell_updates == o.cell_updates and st.n_substeps == o.n_substeps
    assert o.acc_violations == 0
    g.close()
    o.close()
    sys.stdout.flush()
