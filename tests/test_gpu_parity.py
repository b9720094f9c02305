"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Each case starts both sides from the same initial tree and compares the
whole node pool after every global step: topology, LTS levels, clusters and
refinement flags exactly; conserved variables to 1e-11 relative L1 (the
two sides evaluate in the same order, so the state is expected bitwise).
"""
import sys

import numpy as np
import pytest

from dacti_amd import scene, workloads
from dacti_amd.integrator import Integrator
from oracle.binding import Oracle
from tests.helpers import amr_sod_case, compare_pools, sod_case, tiny_box

pytestmark = pytest.mark.gpu


def _run_pair(case, steps, label, tree=None):
    print(f"[parity] {label}: building tree", flush=True)
    if tree is None:
        tree = scene.build_tree(case)
    o = Oracle(case, tree)
    g = Integrator(case, tree)
    reps = []
    rep = compare_pools(g.get_nodes(), o.export())
    reps.append(rep)
    for s in range(steps):
        print(f"[parity] {label}: step {s}", flush=True)
        dto, fbo = o.integrate()
        dtg, fbg = g.integrate()
        assert dto == dtg and fbo == fbg
        rep = compare_pools(g.get_nodes(), o.export())
        reps.append(rep)
    assert o.acc_violations == 0
    st = g.stats()
    assert st.cell_updates == o.cell_updates
    assert st.n_substeps == o.n_substeps
    g.close()
    sys.stdout.flush()
    return reps


def test_parity_euler3d_periodic_lts():
    reps = _run_pair(tiny_box(level=4, dim=3), 3, "euler3d 16^3")
    assert all(r["u_bitwise"] for r in reps)


def test_parity_euler2d_periodic_lts():
    reps = _run_pair(tiny_box(level=5, dim=2), 4, "euler2d 32^2")
    assert all(r["u_bitwise"] for r in reps)


def test_parity_euler2d_sod_boundaries():
    _run_pair(sod_case(level=5), 4, "sod 32^2")


def test_parity_euler2d_amr():
    _run_pair(amr_sod_case(base=4, top=6), 6, "euler2d amr")


def test_parity_navier_stokes_2d_amr():
    _run_pair(amr_sod_case(base=4, top=6, model="navier_stokes"), 4, "ns2d amr")


def test_parity_multi2d_shock_bubble_small():
    case = workloads.shock_bubble(base=4, extra=2)
    _run_pair(case, 4, "multi2d shock-bubble 16x8+2")


def test_parity_multi2d_deep_subtrees():
    # five levels of refinement above the base: the far side of a neighbour
    # subtree reaches more than three levels below the neighbour, which the
    # interpolation walk must follow like the reference's recursion
    case = workloads.shock_bubble(base=4, extra=5, mach=2.0, radius=0.2, rho2=0.5)
    _run_pair(case, 2, "multi2d shock-bubble 16x8+5")


def test_parity_multi3d_shock_drop():
    # stiffened-gas water drop (P0 3.43e8 Pa) in air, SI units, AMR + LTS + interface chain in 3D
    case = workloads.shock_drop(base=3, extra=2, radius=0.002, shock_x=-0.005)
    _run_pair(case, 2, "multi3d shock-drop 8x4x4+2")


def test_parity_multi3d_shock_drop_levels_4_7():
    # the shock-drop workload itself (Mach 2.4, 1 mm drop) at octree levels
    # 4-7: ~27 M cell updates over three global steps, pools bitwise equal
    _run_pair(workloads.shock_drop(base=4, extra=3), 3, "multi3d shock-drop levels 4-7")


def test_parity_euler_static_amr_zone():
    # the conservation case of the oracle: a refinement zone one level above
    # the base, faces between tree levels and between time levels
    from tests.test_oracle_amr_conservation import _case
    for dim in (2, 3):
        _run_pair(_case(dim), 3, f"euler{dim}d static AMR zone")


def test_parity_navier_stokes_3d_box():
    from tests.helpers import ns3_box
    _run_pair(ns3_box(level=3), 2, "navier-stokes 3d 8^3")


def test_parity_gc_after_coarsening():
    # an initially refined jump that every cell wants to coarsen: the active
    # fraction falls below 0.2 in the first step and the pool is compacted
    # (impl/acti_utils.hpp:58-74), after which node ids must still match
    case = amr_sod_case(base=3, top=8, emin=1e10, emax=1e20)
    tree = scene.build_tree(case)
    n0 = tree.n
    o = Oracle(case, tree)
    g = Integrator(case, tree)
    for s in range(3):
        print(f"[parity] gc case: step {s}", flush=True)
        o.integrate()
        g.integrate()
        compare_pools(g.get_nodes(), o.export())
    assert g.n_nodes() < n0 // 4  # compacted
    g.close()
