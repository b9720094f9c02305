"""Closed-form pin of the coarse/fine flux ownership (impl/flux.hpp:58-59)
and the LTS deposits: on a static adaptive mesh with faces between tree
levels and between time levels, every face flux must enter the two cells it
separates once each with opposite signs (sub-face fluxes summed into the
coarse side), so the total mass, momentum and energy of a periodic box stay
constant to round-off.  A face computed twice, skipped on one side or
deposited with the wrong sign breaks it.  The velocity is kept small so that
no coarse leaf sits on a finer time level than its fine neighbour (the case
in which the reference loses the face's flux, kept on purpose)."""
import numpy as np

from dacti_amd import scene, workloads
from dacti_amd.case import loads_case
from oracle.binding import Oracle
from tests.helpers import live_leaves


def _case(dim):
    base = 4
    toml = workloads.PERIODIC_BOX_TOML.format(level=base).replace(
        "max_refinement_level = 4", "max_refinement_level = 5")
    zone = "[[refinement_zone]]\nstart = [-0.2, -0.3, -0.3]\nend = [0.3, 0.2, 0.3]\nmax_refinement_level = 5\n"
    if dim == 2:
        toml = toml.replace("[1.0, 1.0, 1.0]", "[1.0, 1.0]").replace("[0.0, 0.0, 0.0]", "[0.0, 0.0]")
        toml = toml.replace('[boundary_conditions.Z]\ntype = "periodic"\n', "")
    case = loads_case(toml + zone, "euler", dim)
    ref = workloads.periodic_box(level=base, dim=dim, mach_rms=0.05, seed=3)
    case.ic_grid = ref.ic_grid
    return case


def _totals(a, dim):
    lv = live_leaves(a)
    return (a["u"][lv] * a["size"][lv, None] ** dim).sum(0), lv


def test_amr_lts_conservation():
    for dim in (2, 3):
        case = _case(dim)
        o = Oracle(case, scene.build_tree(case))
        a0 = o.export()
        m0, lv0 = _totals(a0, dim)
        assert len(np.unique(a0["tree_level"][lv0])) >= 2
        for _ in range(3):
            _, fb = o.integrate()
            assert not fb
        a = o.export()
        m, lv = _totals(a, dim)
        # faces between tree levels and between time levels, on a mesh that
        # stays as it is (a split interpolates primitive variables, so it
        # conserves mass but not momentum and energy exactly)
        np.testing.assert_array_equal(a["tree_level"][lv], a0["tree_level"][lv0])
        assert len(np.unique(a["acti_level"][lv])) >= 2
        assert o.acc_violations == 0
        np.testing.assert_allclose(m, m0, rtol=1e-13, atol=1e-14, err_msg=f"dim {dim}")
        o.close()
