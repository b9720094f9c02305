"""Shared helpers: small cases and node-pool comparison."""
import numpy as np

from dacti_amd import workloads
from dacti_amd.case import loads_case

TOPOLOGY = ("parent", "child_base", "nbr", "tree_level", "acti_level", "cluster", "flags")
STATE = ("t", "p", "u", "slope", "rhs", "k", "grad", "meta", "center", "size")


def live_leaves(a):
    f = a["flags"]
    return np.nonzero(((f & 1) != 0) & ((f & 2) != 0) & ((f & 4) == 0))[0]


def compare_pools(g, o, rtol_l1=1e-11):
    """Compare a GPU export with an oracle export.

    Integer topology, LTS levels, clusters and flags must match exactly.
    Every state field of every live leaf (time, primitives, conserved
    variables, slopes, rhs, parameters incl. normals and curvature,
    gradients, metadata error / max_dt / cfl, centre, size) must be bitwise
    equal: both sides evaluate in the same order.  The relative L1 of u, p
    and t is reported as well (the 1e-11 bar is implied by bitwise equality).
    Returns a dict describing the comparison."""
    rep = {}
    assert g["parent"].shape == o["parent"].shape, (g["parent"].shape, o["parent"].shape)
    for k in TOPOLOGY:
        bad = np.nonzero((g[k] != o[k]).reshape(len(g[k]), -1).any(axis=1))[0]
        assert bad.size == 0, f"{k} differs at nodes {bad[:10]}: gpu {g[k][bad[:5]]} oracle {o[k][bad[:5]]}"
    lv = live_leaves(o)
    for k in STATE:
        a, b = g[k][lv], o[k][lv]
        both_nan = np.isnan(a) & np.isnan(b)
        eq = (a == b) | both_nan
        rep[k + "_bitwise"] = bool(eq.all())
        if k in ("u", "p", "t"):
            ok = ~both_nan
            den = np.abs(b[ok]).sum()
            rel = np.abs(a[ok] - b[ok]).sum() / (den if den > 0 else 1.0)
            rep[k + "_rel_l1"] = float(rel)
            assert rel <= rtol_l1, f"{k}: relative L1 {rel}"
        if not eq.all():
            rows = np.nonzero(~eq.reshape(len(lv), -1).all(axis=1))[0]
            i = rows[0]
            raise AssertionError(f"{k} not bitwise equal at {rows.size} leaves, first node {lv[i]}: "
                                 f"gpu {a[i]} oracle {b[i]}")
    return rep


def tiny_box(level=4, dim=3):
    return workloads.periodic_box(level=level, dim=dim)


SOD_TOML = """
[scene]
name = "sod"
domain_size = [1.0, {h}]
[run]
print_interval = 1.0
end_time = 1.0
[integrator]
min_refinement_level = {lev}
max_refinement_level = {lev}
error_threshold = [-1.0, 1e300]
max_cfl = 0.5
max_dt = {maxdt}
[model.reference_values]
rho = 1.0
p = 1.0
vel = [0.0, 0.0]
[initial_condition]
type = "jump"
axis = 0
location = 0.0
value_neg = [1.0, 0.0, 0.0, 1.0]
value_pos = [0.125, 0.0, 0.0, 0.1]
[boundary_conditions.X.neg]
type = "zeroGradient"
[boundary_conditions.X.pos]
type = "zeroGradient"
[boundary_conditions.Y]
type = "periodic"
"""


def sod_case(level=7, maxdt=1e9):
    """Sod shock tube along x in a 2D strip one root-quarter high, periodic in y."""
    return loads_case(SOD_TOML.format(h=repr(1.0), lev=level, maxdt=repr(float(maxdt))), "euler", 2)


AMR_TOML = """
[scene]
name = "amr2d"
domain_size = [1.0, 1.0]
[run]
print_interval = 1.0
end_time = 1.0
[integrator]
min_refinement_level = {base}
max_refinement_level = {top}
error_threshold = [{emin}, {emax}]
error_weights = [1.0, 0.0, 0.0, 0.1]
max_cfl = 0.5
[model.reference_values]
rho = 1.0
p = 1.0
vel = [0.0, 0.0]
{extra}
[initial_condition]
type = "jump"
axis = 0
location = {loc}
value_neg = [1.0, 0.0, 0.0, 1.0]
value_pos = [0.125, 0.0, 0.0, 0.1]
[boundary_conditions.X.neg]
type = "{bcx}"
[boundary_conditions.X.pos]
type = "zeroGradient"
[boundary_conditions.Y.neg]
type = "freeslip"
[boundary_conditions.Y.pos]
type = "noslip"
"""


def amr_sod_case(base=4, top=6, model="euler", emin=0.005, emax=0.02, loc=-0.1):
    extra = ""
    if model == "navier_stokes":
        extra = "Re = 200.0\nL = 1.0\nPr = 0.72\n"
    text = AMR_TOML.format(base=base, top=top, emin=repr(emin), emax=repr(emax), extra=extra, loc=repr(loc),
                           bcx="zeroGradient")
    if model == "navier_stokes":
        text = text.replace("vel = [0.0, 0.0]", "vel = [0.3, 0.0]")
    return loads_case(text, model, 2)


def ns3_box(level=3):
    """3D periodic box with the viscous Navier-Stokes model (Re 100), the
    smooth random state of the Euler roofline case."""
    base = workloads.periodic_box(level=level)
    toml = workloads.PERIODIC_BOX_TOML.format(level=level).replace(
        "vel = [0.0, 0.0, 0.0]", "vel = [0.0, 0.0, 0.0]\nRe = 100.0\nL = 1.0\nPr = 0.72")
    case = loads_case(toml, "navier_stokes", 3)
    case.ic_grid = base.ic_grid
    return case


# ---------------------------------------------------------------- objects
SOLID_TOML = """
[scene]
name = "solid"
domain_size = [2.0, 1.0]
[run]
print_interval = 1.0
end_time = 1.0
[integrator]
min_refinement_level = {base}
max_refinement_level = {top}
max_refinement_level_geom = {geom}
refinement_radius = 2.0
error_threshold = [{emin}, {emax}]
error_weights = [1.0, 0.0, 0.0, 0.1]
max_cfl = 0.4
[model.reference_values]
rho = 1.0
p = 1.0
vel = [{u}, {v}]
{extra}
[initial_condition]
type = "uniform"
{geometry}
[boundary_conditions.X.neg]
type = "{xneg}"
[boundary_conditions.X.pos]
type = "{xpos}"
[boundary_conditions.Y.neg]
type = "{ybc}"
[boundary_conditions.Y.pos]
type = "{ybc}"
"""

SOLID_GEOMETRY = {
    # a cylinder (noslip), the reference's cylinder case in small
    "cylinder": """[[geometry]]
name = "object"
shape = "ellipse"
a = 0.15
b = 0.15
translation = [-0.4, 0.0, 0.0]
[boundary_conditions.object]
type = "noslip"
""",
    # a diamond aerofoil at 8 degrees incidence (freeslip)
    "diamond": """[[geometry]]
name = "object"
shape = "diamond"
width = 0.6
height = 0.12
rotation = [0.0, 0.0, -8.0]
translation = [-0.2, 0.02, 0.0]
[boundary_conditions.object]
type = "freeslip"
""",
    # two solids, a ramp and a cowl (freeslip), as in the reference's scramjet
    "inlet": """[[geometry]]
name = "ramp"
shape = "triangle"
v1 = [-1.0, -0.5]
v2 = [0.6, -0.5]
v3 = [0.6, -0.3]
[[geometry]]
name = "cowl"
shape = "quadrilateral"
v1 = [0.0, 0.2]
v2 = [0.8, 0.12]
v3 = [0.8, 0.2]
v4 = [0.2, 0.25]
[boundary_conditions.ramp]
type = "freeslip"
[boundary_conditions.cowl]
type = "noslip"
""",
}


def solid_case(kind, model="euler", base=4, top=6, geom=6, emin=0.005, emax=0.02, u=0.5, v=0.0, closed=False,
               limiter="minmod"):
    """A 2D case with solid objects: inflow/outflow/freeslip domain, or a
    closed freeslip box (no mass or energy crosses any face)."""
    extra = "Re = 200.0\nL = 1.0\nPr = 0.72\n" if model == "navier_stokes" else ""
    text = SOLID_TOML.format(base=base, top=top, geom=geom, emin=repr(emin), emax=repr(emax), u=repr(u), v=repr(v),
                             extra=extra, geometry=SOLID_GEOMETRY[kind],
                             xneg="freeslip" if closed else "inflow", xpos="freeslip" if closed else "outflow",
                             ybc="freeslip")
    return loads_case(text, model, 2, limiter)
