"""GPU parity on the paths the base cases do not reach: every slope
limiter, the five-equation viscous, heat-conduction and surface-tension
terms, the inflow (with fade-in), outflow and dirichlet boundaries, the lost
coarse/fine flux, the 2D shock-bubble at its full size, and the per-substep
hook of integrate().  Each case compares the whole node pool with the
oracle after every global step (topology exactly, every state field
bitwise)."""
import numpy as np
import pytest

from dacti_amd import scene, workloads
from dacti_amd.case import BC, BC_TYPES, LIMITERS, loads_case
from dacti_amd.integrator import Integrator
from tests.test_gpu_parity import _run_pair
from tests.helpers import tiny_box

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", [k for k in LIMITERS if k != "minmod"])
def test_parity_limiters(name):
    case = tiny_box(level=5, dim=2)
    case.limiter = LIMITERS[name]
    reps = _run_pair(case, 3, f"euler2d 32^2 limiter {name}")
    assert all(r["u_bitwise"] for r in reps)


def _sigma_mu_bubble():
    """Shock-bubble at 16x8 + 2 levels with surface tension (We 20),
    viscosity (Re 500, viscosity ratio 2) and heat conduction in both
    phases (the reference's conduction constants have no TOML key; they are
    set as its process-global constants would be)."""
    case = workloads.shock_bubble(base=4, extra=2)
    toml_case = loads_case(
        workloads.SHOCK_BUBBLE_TOML.format(base=4, top=6, emin="0.02", emax="0.05", rho2="0.138", shock_x="-0.7",
                                           r2=repr(case.bcs[0].value[0]), u2=repr(case.bcs[0].value[2]),
                                           p2=repr(case.bcs[0].value[4]), radius="0.25", bx="-0.3").replace(
            "alpha = 1.0\n", "alpha = 1.0\nWe = 20.0\nRe = 500.0\nmu_ratio = 2.0\n", 1), "multi", 2)
    toml_case.consts["lambda1"] = 2e-3
    toml_case.consts["lambda2"] = 5e-3
    assert toml_case.consts["sigma"] > 0 and toml_case.consts["mu1"] > 0 and toml_case.consts["mu2"] > 0
    return toml_case


def test_parity_multi2d_surface_tension_viscous_heat():
    _run_pair(_sigma_mu_bubble(), 4, "multi2d shock-bubble We 20 Re 500 with conduction")


def test_parity_multi2d_static_disc_surface_tension():
    from tests.test_oracle_phase_pins import ABGRALL_TOML
    case = loads_case(ABGRALL_TOML.format(extra="We = 10.0", p1="0.0").replace("vel = [0.5, 0.0]", "vel = [0.0, 0.0]"),
                      "multi", 2)
    tree = scene.build_tree(case)
    lf = np.nonzero((tree.flags & 3) == 3)[0]
    a = (np.hypot(tree.center[lf, 0], tree.center[lf, 1]) < 0.25).astype(float)
    tree.p[lf, 0], tree.p[lf, 1], tree.p[lf, 5] = a, 1 - a, a
    _run_pair(case, 12, "multi2d static disc We 10", tree=tree)


BC_TOML = """
[scene]
name = "bcs"
domain_size = [1.0, 1.0]
[run]
print_interval = 1.0
end_time = 1.0
[integrator]
min_refinement_level = 4
max_refinement_level = 6
error_threshold = [0.005, 0.02]
error_weights = [1.0, 0.0, 0.0, 0.1]
max_cfl = 0.5
[model.reference_values]
rho = 1.0
p = 1.0
vel = [0.6, 0.1]
[initial_condition]
type = "jump"
axis = 0
location = 0.1
value_neg = [1.0, 0.6, 0.1, 1.0]
value_pos = [0.4, 0.0, 0.0, 0.5]
[boundary_conditions.X.neg]
type = "inflow"
fade_in = 0.05
[boundary_conditions.X.pos]
type = "outflow"
[boundary_conditions.Y.neg]
type = "zeroGradient"
[boundary_conditions.Y.pos]
type = "freeslip"
"""


def test_parity_inflow_fade_outflow_dirichlet():
    """Inflow of the reference state faded in over 0.05 time units (x-),
    outflow (x+), a dirichlet state on y- (no TOML key in the reference; set
    as the boundary object would be) and freeslip (y+), with AMR.  (The
    boundary star state puts the ghost on the negative side of every face,
    so a dirichlet inflow through a positive face is not a stable setting.)"""
    case = loads_case(BC_TOML, "euler", 2)
    case.bcs[2] = BC(BC_TYPES["dirichlet"], (0.8, 0.2, 0.3, 0.9), 0.0)
    _run_pair(case, 6, "euler2d inflow/outflow/dirichlet AMR")


def test_parity_lost_coarse_fine_flux():
    from tests.test_oracle_phase_pins import a12_case
    case, tree = a12_case()
    _run_pair(case, 3, "euler2d lost coarse/fine flux", tree=tree)


def test_parity_shock_bubble_full_size():
    """The 2D shock-bubble at its own size: 64 x 32 base, four levels of
    refinement, remeshing every substep, 20 global steps."""
    _run_pair(workloads.shock_bubble(base=6, extra=4), 20, "multi2d shock-bubble 64x32+4")


def test_substep_hook_follows_schedule():
    """integrate(substep_hook) is called after every substep with the
    level's time, the running substep count and dt_L (impl/integrate.hpp:64),
    in the reference's order (finest level first at each synchronisation
    point); the hook may read the integrator's state."""
    case = workloads.shock_bubble(base=4, extra=2)
    g = Integrator(case)
    calls = []

    def hook(integ, level, acti_time, n_substeps, dt):
        calls.append((level, acti_time, n_substeps, dt, integ.n_active_cells()))

    n0 = g.stats().n_substeps
    cur = cur0 = g.max_acti_level
    gdt, fb = g.integrate(substep_hook=hook)
    assert not fb and calls
    st = g.stats()
    assert [c[2] for c in calls] == list(range(n0 + 1, st.n_substeps + 1))
    # walk the reference's order; a level above max_acti_level runs only
    # where it first gets a cluster (max_acti_level then grows)
    pos = 0
    for seq in range(1, (1 << 15) + 1):
        for l in range(16):
            if seq % (1 << l):
                break
            L = 15 - l
            if L > cur and pos < len(calls) and calls[pos][0] == L:
                cur = L
            if L <= cur:
                assert calls[pos][0] == L, (seq, L, pos, calls[pos])
                pos += 1
    assert pos == len(calls)
    # acti_time[L] starts at 0 (construction), or at acti_time[L - 1] for a
    # level that appears during the step (impl/integrate.hpp:33-38)
    t = {lv: 0.0 for lv in range(cur0 + 1)}
    for level, at, _, dt, nact in calls:
        assert dt == gdt / (1 << level)
        if level not in t:
            t[level] = t[level - 1]
        t[level] += dt
        assert at == pytest.approx(t[level], rel=1e-14)
        assert nact > 0
    with pytest.raises(ZeroDivisionError):
        g.integrate(substep_hook=lambda *a: 1 / 0)
    g.close()


# ------------------------------------------------------------ solid objects
@pytest.mark.parametrize("kind,model,u", [("cylinder", "euler", 0.5), ("diamond", "euler", 2.0),
                                          ("inlet", "euler", 1.8), ("cylinder", "navier_stokes", 0.5)])
def test_parity_solid_objects(kind, model, u):
    """Solids with noslip and freeslip walls (records with the wall normal,
    integrator/boundary.hpp:75-160), inflow/outflow/freeslip domain, AMR to
    the geometry level."""
    from tests.helpers import solid_case
    case = solid_case(kind, model=model, base=4, top=6, geom=6, u=u)
    tree = scene.build_tree(case)
    assert len(tree.records) > 0
    _run_pair(case, 5, f"{model}2d solid {kind} u {u}", tree=tree)


SPHERE3D_TOML = """
[scene]
name = "sphere"
domain_size = [2.0, 1.0, 1.0]
[run]
print_interval = 1.0
end_time = 1.0
[integrator]
min_refinement_level = 3
max_refinement_level = 5
refinement_radius = 1.5
error_threshold = [0.01, 0.05]
error_weights = [1.0, 0.0, 0.0, 0.0, 0.1]
max_cfl = 0.4
[model.reference_values]
rho = 1.0
p = 1.0
vel = [1.5, 0.0, 0.0]
[initial_condition]
type = "uniform"
[[geometry]]
name = "ball"
path = "ball.stl"
scaling = [0.3, 0.3, 0.3]
translation = [-0.3, 0.05, 0.0]
[boundary_conditions.ball]
type = "noslip"
[boundary_conditions.X.neg]
type = "inflow"
[boundary_conditions.X.pos]
type = "outflow"
[boundary_conditions.Y]
type = "periodic"
[boundary_conditions.Z.neg]
type = "freeslip"
[boundary_conditions.Z.pos]
type = "freeslip"
"""


def test_parity_solid_stl_sphere_3d(tmp_path):
    """A sphere read from an STL file (an icosphere written here), 3D Euler
    at Mach 1.3 around it, periodic in y."""
    from tests.test_geometry import icosphere, write_stl
    v, f = icosphere(3)
    write_stl(tmp_path / "ball.stl", [v[t] for t in f])
    case = loads_case(SPHERE3D_TOML, "euler", 3, base_dir=str(tmp_path))
    tree = scene.build_tree(case)
    assert len(tree.records) > 0
    _run_pair(case, 3, "euler3d STL sphere", tree=tree)
