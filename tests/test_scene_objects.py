"""Initial trees with solid objects (scene/scene.cpp:47-335: signed
distances, object boundaries, splits to max_refinement_level_geom,
fix_boundary) and the oracle's object boundary conditions
(integrator/boundary.hpp:75-160 with a record's wall normal).

Tree invariants: every open face of an active leaf leads to an active
node; the solid's inside is inactive; leaves next to the surface sit at the
geometry level; every object face has a record whose condition is its
object's, with a unit normal pointing out of the solid.  Oracle: closed-form
ghost states of records, and a closed freeslip box around a noslip solid
conserves mass and energy to rounding (no flux crosses any boundary)."""
import ctypes as C

import numpy as np
import pytest

from dacti_amd import _lib, scene
from dacti_amd.case import ConfigError, loads_case
from oracle.binding import Oracle, model_fn, params
from tests.helpers import SOLID_GEOMETRY, SOLID_TOML, solid_case

BC = {"noslip": 2, "freeslip": 1}


def find(t, x, level, case):
    """Domain::find (integrator/domain.hpp:65-82,407-417) on a tree."""
    x = np.array(x, float)
    for d in range(case.dim):
        h = case.active_size[d] / 2
        if x[d] < -h or x[d] > h:
            if not case.periodic[d]:
                return 0
            x[d] += case.active_size[d] if x[d] < -h else -case.active_size[d]
    x = 2.0 * x / t.size[1]
    i = 1
    while not (t.flags[i] & 1) and t.level[i] != level:
        q = sum(1 << d for d in range(case.dim) if x[d] > 0)
        x = (x - 0.5 * np.where([(q >> d) & 1 for d in range(case.dim)], 1.0, -1.0)) * 2.0
        i = t.child_base[i] + q
    return i


@pytest.mark.parametrize("kind", ["cylinder", "diamond", "inlet"])
def test_solid_tree_invariants(kind):
    case = solid_case(kind, base=3, top=6, geom=7)
    t = scene.build_tree(case)
    D = case.dim
    act = (t.flags & 3) == 3
    assert len(t.records) > 0
    obj = t.bc == _lib.BC_OBJECT
    has_obj = obj.any(axis=1)
    # object faces of active leaves carry a record
    assert (t.brec[act & has_obj] >= 0).all()
    assert (t.flags[act & has_obj] & _lib.F_BOUNDARY).any()
    recs = t.records[t.brec[act & has_obj]]
    np.testing.assert_allclose(np.linalg.norm(recs["normal"][:, :D], axis=1), 1.0, rtol=1e-12)
    assert set(recs["type"]) <= set(BC.values())
    n_open = 0
    for i in np.nonzero(act)[0]:
        for f in range(2 * D):
            if t.bc[i, f]:
                continue
            x = t.center[i].copy()
            x[f >> 1] += (2 * (f & 1) - 1) * t.size[i]
            nb = find(t, x, t.level[i], case)
            # an open face never leads into the solid
            assert nb != 0 and (t.flags[nb] & 2), (i, f, nb)
            n_open += 1
    assert n_open > 100
    # leaves at the surface are at the geometry level; the solid's inside is inactive
    lv = t.leaves()
    surf = lv[has_obj[lv]]
    assert (t.level[surf] == 7).mean() > 0.5
    leaf = (t.flags & 1).astype(bool)
    inactive_leaves = np.nonzero(leaf & ~(t.flags & 2).astype(bool))[0]
    inside = [i for i in inactive_leaves if abs(t.center[i, 0]) < 1 and abs(t.center[i, 1]) < 0.5]
    assert len(inside) > 0


def test_record_normals_point_out_of_the_solid():
    case = solid_case("cylinder", base=3, top=6, geom=7)
    t = scene.build_tree(case)
    act = (t.flags & 3) == 3
    sel = act & (t.bc == _lib.BC_OBJECT).any(axis=1)
    r = t.center[sel] - [-0.4, 0.0]
    nrm = t.records[t.brec[sel]]["normal"][:, :2]
    radial = r / np.linalg.norm(r, axis=1)[:, None]
    assert np.min(np.sum(nrm * radial, axis=1)) > 0.999
    assert (t.records[t.brec[sel]]["type"] == BC["noslip"]).all()
    # active boundary cells lie outside the cylinder, within a cell of it
    d = np.linalg.norm(r, axis=1) - 0.15
    assert (d > -t.size[sel]).all() and (d < 1.5 * t.size[sel]).all()


def test_solid_needs_a_boundary_condition():
    text = SOLID_TOML.format(base=3, top=5, geom=5, emin="0.005", emax="0.02", u="0.5", v="0.0", extra="",
                             geometry=SOLID_GEOMETRY["cylinder"].replace('[boundary_conditions.object]\ntype = "noslip"\n',
                                                                         ""),
                             xneg="inflow", xpos="outflow", ybc="freeslip")
    with pytest.raises(ConfigError):
        loads_case(text, "euler", 2)
    # a fluid object needs none, and carries no boundary
    fluid = text.replace('shape = "ellipse"', 'shape = "ellipse"\ntype = "droplet"')
    t = scene.build_tree(loads_case(fluid, "euler", 2))
    act = (t.flags & 3) == 3
    assert not (t.bc[act] == _lib.BC_OBJECT).any()


@pytest.mark.parametrize("model", ["euler", "navier_stokes", "multi"])
def test_record_ghost_states_closed_forms(model):
    case = solid_case("cylinder") if model != "multi" else None
    if model == "multi":
        from tests.test_oracle_phase_pins import ABGRALL_TOML
        case = loads_case(ABGRALL_TOML.format(extra="", p1="0.0"), "multi", 2)
    D, NE = 2, case.num_eqns
    P = params(case)
    f = model_fn(case.model, 2, "t_boundary_value_rec", None,
                 [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                  C.POINTER(C.c_double), C.c_double, C.POINTER(C.c_double)])
    iu = 2 if model == "multi" else 1
    ip = NE - 2 if model == "multi" else NE - 1
    rng = np.random.default_rng(7)
    for _ in range(20):
        p = rng.uniform(0.2, 1.5, NE)
        p[iu:iu + D] = rng.normal(size=D)
        if model == "multi":
            p[NE - 1] = rng.uniform()
        th = rng.uniform(0, 2 * np.pi)
        n = np.array([np.cos(th), np.sin(th), 0.0])
        pin = rng.uniform(0.5, 2.0, 8)
        out = np.zeros(NE)
        args = lambda typ, fade, t: (C.byref(P), typ, n.ctypes.data_as(C.POINTER(C.c_double)),
                                     pin.ctypes.data_as(C.POINTER(C.c_double)), fade,
                                     p.ctypes.data_as(C.POINTER(C.c_double)), t,
                                     out.ctypes.data_as(C.POINTER(C.c_double)))
        f(*args(BC["freeslip"], 0.0, 0.0))
        vb, v = out[iu:iu + D], p[iu:iu + D]
        assert abs(vb @ n[:D]) < 1e-15 * (1 + abs(v @ n[:D]))  # no normal velocity
        t_ = np.array([-n[1], n[0]])
        assert vb @ t_ == pytest.approx(v @ t_, rel=1e-14, abs=1e-15)  # tangential kept
        np.testing.assert_array_equal(np.delete(out, range(iu, iu + D)), np.delete(p, range(iu, iu + D)))
        f(*args(BC["noslip"], 0.0, 0.0))
        np.testing.assert_array_equal(out[iu:iu + D], -p[iu:iu + D])
        # inflow: the record's state with the cell's pressure (a subsonic
        # cell), velocity faded in over fade_in
        p[iu:iu + D] *= 0.01
        f(*args(3, 0.4, 0.1))
        expect = pin[:NE].copy()
        expect[iu:iu + D] *= min(0.1 / 0.4, 1.0)
        expect[ip] = p[ip]
        np.testing.assert_array_equal(out, expect)


def test_closed_box_conserves_mass_and_energy():
    """Noslip solid in a freeslip box, static mesh: no mass or energy
    crosses a boundary face, so the totals stay to rounding."""
    case = solid_case("cylinder", base=4, top=6, geom=6, emin=-1.0, emax=1e300, closed=True, u=0.4, v=0.1)
    t = scene.build_tree(case)
    o = Oracle(case, t)

    def totals():
        e = o.export()
        lv = np.nonzero((e["flags"] & 3) == 3)[0]
        vol = e["size"][lv] ** 2
        return np.array([np.sum(e["u"][lv, 0] * vol), np.sum(e["u"][lv, 3] * vol)])

    m0 = totals()
    for _ in range(6):
        _, fb = o.integrate()
        assert not fb
    m1 = totals()
    np.testing.assert_allclose(m1, m0, rtol=1e-13)
    # and the flow did something: the momentum changed
    e = o.export()
    lv = np.nonzero((e["flags"] & 3) == 3)[0]
    assert np.abs(e["u"][lv, 1] / e["u"][lv, 0] - 0.4).max() > 1e-3


def test_split_children_keep_the_wall():
    """Remeshing next to a solid (impl/adaptive_refinement.hpp:184-189):
    children inherit their parent's outward faces, with the parent's object
    record, so after several adaptive steps every object face of an active
    leaf still has the cylinder's record, its normal pointing out of the
    cylinder."""
    case = solid_case("cylinder", base=3, top=6, geom=5, emin=0.002, emax=0.01, u=0.6)
    t = scene.build_tree(case)
    o = Oracle(case, t)
    n0 = o.n_nodes
    for _ in range(4):
        _, fb = o.integrate()
        assert not fb
    e = o.export()
    bc, brec = o.export_bc()
    assert o.n_nodes > n0  # the mesh was refined
    act = (e["flags"] & 3) == 3
    wall = act & (bc == _lib.BC_OBJECT).any(axis=1)
    assert wall.sum() > 20
    assert (brec[wall] >= 0).all() and (brec[wall] < len(t.records)).all()
    r = e["center"][wall] - [-0.4, 0.0]
    nrm = t.records[brec[wall]]["normal"][:, :2]
    assert np.min(np.sum(nrm * r / np.linalg.norm(r, axis=1)[:, None], axis=1)) > 0.99
    # most wall cells were split past the geometry level during the run
    assert (e["tree_level"][wall] == 6).sum() > (e["tree_level"][wall] == 5).sum()
