"""GPU parity of the flux op for dense clusters (windows of 64 node ids,
each face between two updating same-level leaves evaluated once).

Every case runs three integrators from the same tree: the window op forced
for every cluster (mode 2), never used (mode 0) and the oracle; both GPU
runs must equal the oracle bit for bit (whole node pool after each global
step), so the window op changes no result: which side evaluates a face,
the fluxes received from finer LTS neighbours, the deposits for coarser
ones and each cell's face order are those of the per-cell op
(impl/flux.hpp:32-91).  The cases cover periodic and domain boundaries,
LTS levels, tree-level faces (cells set aside for the general op), AMR,
the viscous and five-equation models and solid walls."""
import sys

import pytest

from dacti_amd import scene, workloads
from dacti_amd.integrator import Integrator
from oracle.binding import Oracle
from tests.helpers import amr_sod_case, compare_pools, ns3_box, solid_case, sod_case, tiny_box

pytestmark = pytest.mark.gpu


def _run_modes(case, steps, label):
    print(f"[dense] {label}: building tree", flush=True)
    tree = scene.build_tree(case)
    o = Oracle(case, tree)
    gs = []
    for mode in (2, 0):
        g = Integrator(case, tree)
        g.set_dense(mode)
        gs.append(g)
    for s in range(steps):
        print(f"[dense] {label}: step {s}", flush=True)
        dto, fbo = o.integrate()
        ref = o.export()
        for g in gs:
            dtg, fbg = g.integrate()
            assert dto == dtg and fbo == fbg
            rep = compare_pools(g.get_nodes(), ref)
            assert rep["u_bitwise"], rep
    for g in gs:
        st = g.stats()
        assert st.cell_updates == o.cell_updates and st.n_substeps == o.n_substeps
        g.close()
    o.close()
    sys.stdout.flush()


@pytest.mark.parametrize("dim,level,steps", [(3, 4, 3), (2, 6, 4)])
def test_dense_periodic_lts(dim, level, steps):
    _run_modes(tiny_box(level=level, dim=dim), steps, f"euler{dim}d periodic level {level}")


def test_dense_sod_boundaries():
    _run_modes(sod_case(level=5), 4, "sod 32^2")


def test_dense_euler2d_amr():
    _run_modes(amr_sod_case(base=4, top=6), 5, "euler2d amr")


def test_dense_navier_stokes():
    _run_modes(amr_sod_case(base=4, top=6, model="navier_stokes"), 3, "ns2d amr")
    _run_modes(ns3_box(level=3), 2, "ns3d 8^3")


def test_dense_static_amr_zone():
    from tests.test_oracle_amr_conservation import _case
    for dim in (2, 3):
        _run_modes(_case(dim), 3, f"euler{dim}d static AMR zone")


def test_dense_multi2d_shock_bubble():
    _run_modes(workloads.shock_bubble(base=4, extra=2), 3, "multi2d shock-bubble 16x8+2")


def test_dense_solid_cylinder():
    _run_modes(solid_case("cylinder", base=3, top=5, geom=5), 3, "cylinder")
