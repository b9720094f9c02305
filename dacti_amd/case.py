"""Case description and the reference's TOML case format.

`load_case(path, model, dim)` reads the keys of
src/lib/config/config.cpp:10-243 with the strict typing of
config/util.hpp:15-103 (vectors are float arrays of exact length, ints are
integers, floats are floats) and the model reference values of
model/euler.hpp:288-324, model/navier_stokes.hpp:120-134 and
model/multi.hpp:474-549.  Quirks of the reference are kept and marked.
"""
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

try:
    import tomllib as _toml
except ImportError:  # python < 3.11
    import tomli as _toml

from ._lib import MODEL_EULER, MODEL_MULTI, MODEL_NAVIER_STOKES, model_sizes

LIMITERS = {"minmod": 0, "unlimited": 1, "van_leer": 2, "superbee": 3, "von_alba_2": 4, "koren": 5,
            "first_order": 6}
BC_TYPES = {"zeroGradient": 0, "freeslip": 1, "noslip": 2, "inflow": 3, "inflow_r": 4, "outflow": 5,
            "dirichlet": 6}
MODELS = {"euler": MODEL_EULER, "navier_stokes": MODEL_NAVIER_STOKES, "multi": MODEL_MULTI}


class ConfigError(RuntimeError):
    """Raised where the reference throws std::runtime_error("Config Error")."""


@dataclass
class BC:
    type: int = 0
    value: Tuple[float, ...] = ()
    fade_in: float = 0.0


@dataclass
class Geometry:
    """A [[geometry]] entry (config/config.cpp:66-125): a built-in shape or
    an STL file ("mesh"), its type ("solid", "droplet", "bubble", ...) and
    translation.  The scene is built from the case's TOML text by the C++
    builder, which reads every key of the entry."""
    name: str
    type: str
    shape: str
    center: Tuple[float, float, float]
    radius: float = 0.0


@dataclass
class Case:
    model: int
    dim: int
    limiter: int = 0
    name: str = "case"
    active_size: Tuple[float, ...] = (1.0, 1.0, 1.0)
    side_length: float = 1.0
    periodic: Tuple[bool, ...] = (False, False, False)
    print_interval: float = 1.0
    end_time: float = 1.0
    signal_time: float = -1.0
    steady_tol: float = -1.0
    max_cfl: float = 0.5
    max_dt: float = 1.0
    min_refinement_level: int = 0
    max_refinement_level: int = 0
    max_refinement_level_geom: int = 0
    refinement_radius: float = 2.0
    error_threshold: Tuple[float, float] = (-1.0, 1e300)
    weights: Tuple[float, ...] = ()
    consts: dict = field(default_factory=dict)
    p_ref: Tuple[float, ...] = ()
    rho_ref: Optional[Tuple[float, float]] = None
    bcs: List[BC] = field(default_factory=lambda: [BC() for _ in range(6)])
    ic_type: str = "uniform"
    ic_axis: int = 0
    ic_location: float = 0.0
    ic_neg: Tuple[float, ...] = ()
    ic_pos: Tuple[float, ...] = ()
    ic_grid: object = None  # (n,..., NE) primitive array for ic_type == "grid"
    toml_text: str = None   # the case file's text (the scene builder reads the objects from it)
    base_dir: str = "."     # directory geometry paths are relative to
    geometry: List[Geometry] = field(default_factory=list)
    zones: List[Tuple[Tuple[float, float, float], Tuple[float, float, float], int]] = field(default_factory=list)

    @property
    def sizes(self):
        return model_sizes(self.model, self.dim)

    @property
    def num_eqns(self):
        return self.sizes[1]


def default_consts():
    """Model constants at their static-initialisation values."""
    gamma1, gamma2, cp1, cp2 = 6.12, 1.4, 4181.0, 1005.0
    return {
        "gamma": 1.4, "ns_mu": 0.0, "ns_lambda": 0.0, "ns_cp": 1005.0, "ns_R": 287.0,
        "sigma": 0.0, "gamma1": gamma1, "gamma2": gamma2, "P1": 343000000.0, "P2": 0.0,
        "cp1": cp1, "cp2": cp2,
        # multi.hpp:57-58: cv1/cv2 are fixed at static initialisation from the
        # default gammas; fluid_1/fluid_2 overrides do not reach them
        "cv1": cp1 / gamma1, "cv2": cp2 / gamma2,
        "mu1": 0.0, "mu2": 0.0, "lambda1": 0.0, "lambda2": 0.0,
    }


# ------------------------------------------------------------------ typing
def _at(tbl, path):
    node = tbl
    for part in path.split("."):
        if not isinstance(node, dict) or part not in node:
            return None
        node = node[part]
    return node


def get_entry(tbl, path, kind, length=None, default=None):
    """config/util.hpp:15-103"""
    v = _at(tbl, path)
    if v is None:
        if default is not None:
            return default
        raise ConfigError(f"Config: Couldn't find entry for {path}.")
    if kind == "vec":
        if not isinstance(v, list):
            raise ConfigError(f"Config: Type of {path} should be array")
        if len(v) != length:
            raise ConfigError(f"Config: Length of {path} should be {length}, but is {len(v)}")
        for i, x in enumerate(v):
            if not isinstance(x, float):
                raise ConfigError(f"Config: Type of {path}[{i}] should be floating point")
        return tuple(float(x) for x in v)
    if kind == "float":
        if not isinstance(v, float):
            raise ConfigError(f"Config: Type of {path} should be floating point")
        return v
    if kind in ("int", "uint"):
        if isinstance(v, bool) or not isinstance(v, int):
            raise ConfigError(f"Config: Type of {path} should be int")
        if kind == "uint" and v < 0:
            raise ConfigError(f"Config: {path} should be unsigned")
        return v
    if kind == "str":
        if not isinstance(v, str):
            raise ConfigError(f"Config: Type of {path} should be string")
        return v
    if kind == "bool":
        if not isinstance(v, bool):
            raise ConfigError(f"Config: Type of {path} should be boolean")
        return v
    raise ValueError(kind)


# ------------------------------------------------------- model reference
def _model_dict(case, m):
    D = case.dim
    c = case.consts
    ref = m.get("reference_values") if isinstance(m, dict) else None
    if not isinstance(ref, dict):
        raise ConfigError("Config: model.reference_values missing")
    if case.model in (MODEL_EULER, MODEL_NAVIER_STOKES):
        # euler.hpp:288-324
        rho = get_entry(ref, "rho", "float")
        prs = get_entry(ref, "p", "float")
        if "vel_magnt" in ref:
            vm = get_entry(ref, "vel_magnt", "float")
            va = get_entry(ref, "vel_angle", "float", default=0.0)
            vel = [vm * math.cos(va * math.pi / 180), vm * math.sin(va * math.pi / 180)] + [0.0] * (D - 2)
        elif "Ma" in ref:
            mach = get_entry(ref, "Ma", "float")
            va = get_entry(ref, "vel_angle", "float", default=0.0)
            vm = mach * math.sqrt(1.4 * prs / rho)
            if D == 2:
                vel = [vm * math.cos(va * math.pi / 180), vm * math.sin(va * math.pi / 180)]
            else:
                vel = [-vm * math.cos(va * math.pi / 180), 0.0, vm * math.sin(va * math.pi / 180)]
        else:
            vel = list(get_entry(ref, "vel", "vec", D))
        case.p_ref = tuple([rho] + vel + [prs])
        if case.model == MODEL_NAVIER_STOKES:
            # navier_stokes.hpp:120-134
            Re = get_entry(ref, "Re", "float")
            L = get_entry(ref, "L", "float")
            speed = math.sqrt(sum(v * v for v in vel))
            c["ns_mu"] = case.p_ref[0] * speed * L / Re
            if "Pr" in ref:
                c["ns_lambda"] = c["ns_mu"] * c["ns_cp"] / get_entry(ref, "Pr", "float")
        return
    # multi.hpp:474-549
    rho = get_entry(ref, "rho", "vec", 2)
    prs = get_entry(ref, "p", "float")
    alpha = get_entry(ref, "alpha", "float")
    case.rho_ref = rho
    if "vel_magnt" in ref:
        vm = get_entry(ref, "vel_magnt", "float")
        va = get_entry(ref, "vel_angle", "float", default=0.0)
        vel = [vm * math.cos(va * math.pi / 180), vm * math.sin(va * math.pi / 180)] + [0.0] * (D - 2)
    elif "Ma" in ref:
        mach = get_entry(ref, "Ma", "float")
        va = get_entry(ref, "vel_angle", "float", default=0.0)
        r0, r1 = rho[0] * alpha, rho[1] * (1 - alpha)
        vm = mach * _multi_sound_speed(c, r0, r1, prs, alpha)  # before fluid overrides, as the reference
        vel = [vm * math.cos(va * math.pi / 180), vm * math.sin(va * math.pi / 180)] + [0.0] * (D - 2)
    else:
        vel = list(get_entry(ref, "vel", "vec", D))
    case.p_ref = tuple([rho[0] * alpha, rho[1] * (1.0 - alpha)] + vel + [prs, alpha])
    L = get_entry(ref, "L", "float", default=1.0)
    mu_ratio = get_entry(ref, "mu_ratio", "float", default=1.0)
    We = None
    if "We" in ref:
        We = get_entry(ref, "We", "float")
        c["sigma"] = 1.0 / We
    if "Re" in ref:
        Re = get_entry(ref, "Re", "float")
        if We is None:
            raise ConfigError("Config: Re given without We (the reference reads an uninitialised We)")
        c["mu1"] = math.sqrt(rho[0] * c["sigma"] * L) * math.sqrt(We) / Re
        c["mu2"] = c["mu1"] / mu_ratio
    if "Oh" in ref:
        Oh = get_entry(ref, "Oh", "float")
        c["sigma"] = get_entry(ref, "sigma", "float")
        c["mu1"] = math.sqrt(rho[0] * c["sigma"] * L) * Oh
        c["mu2"] = c["mu1"] / mu_ratio
    if isinstance(m.get("fluid_1"), dict):
        c["P1"] = get_entry(m["fluid_1"], "P0", "float", default=c["P1"])
        c["gamma1"] = get_entry(m["fluid_1"], "gamma", "float", default=c["gamma1"])
    if isinstance(m.get("fluid_2"), dict):
        c["P2"] = get_entry(m["fluid_2"], "P0", "float", default=c["P2"])
        c["gamma2"] = get_entry(m["fluid_2"], "gamma", "float", default=c["gamma2"])


def _multi_gamma(c, a):
    if a < 1e-20:
        return c["gamma2"]
    if a > 1 - 1e-20:
        return c["gamma1"]
    return 1 + 1 / (a / (c["gamma1"] - 1) + (1 - a) / (c["gamma2"] - 1))


def _multi_p0(c, a):
    g1, g2, P1, P2 = c["gamma1"], c["gamma2"], c["P1"], c["P2"]
    if a < 1e-20:
        return P2
    if a > 1 - 1e-20:
        return P1
    return (a * g1 * P1 / (g1 - 1) + (1 - a) * g2 * P2 / (g2 - 1)) / (1 + a / (g1 - 1) + (1 - a) / (g2 - 1))


def _multi_sound_speed(c, r0, r1, p, a):
    return math.sqrt(_multi_gamma(c, a) * (p + _multi_p0(c, a)) / (r0 + r1))


# ---------------------------------------------------------------- loader
def _parse_bc(case, tbl, ne):
    """parse_boundary_condition: config/config.cpp:145-190"""
    t = get_entry(tbl, "type", "str")
    if t in ("noslip", "freeslip", "zeroGradient"):
        return BC(BC_TYPES[t], (), 0.0)
    if t == "inflow":
        return BC(BC_TYPES[t], tuple(case.p_ref), get_entry(tbl, "fade_in", "float", default=0.0))
    if t == "inflow_r":
        fade = get_entry(tbl, "fade_in", "float", default=0.0)
        return BC(BC_TYPES[t], get_entry(tbl, "value", "vec", ne), fade)
    if t == "outflow":
        return BC(BC_TYPES[t], (), 0.0)
    raise ConfigError("illegal bc type")


def load_case_table(tbl, model, dim, limiter="minmod"):
    """Config<M>::load_config (config/config.cpp:10-141) on a parsed TOML table."""
    if isinstance(model, str):
        model = MODELS[model]
    case = Case(model=model, dim=dim, limiter=LIMITERS[limiter] if isinstance(limiter, str) else limiter)
    case.consts = default_consts()
    ne = case.num_eqns
    case.name = get_entry(tbl, "scene.name", "str")
    case.active_size = get_entry(tbl, "scene.domain_size", "vec", dim)
    case.side_length = max(case.active_size)
    case.print_interval = get_entry(tbl, "run.print_interval", "float")
    case.end_time = get_entry(tbl, "run.end_time", "float")
    case.signal_time = get_entry(tbl, "run.signal_time", "float", default=-1.0)
    case.steady_tol = get_entry(tbl, "run.steady_tol", "float", default=-1.0)
    case.max_cfl = get_entry(tbl, "integrator.max_cfl", "float")
    case.max_dt = get_entry(tbl, "integrator.max_dt", "float", default=case.print_interval)
    case.min_refinement_level = get_entry(tbl, "integrator.min_refinement_level", "int")
    case.max_refinement_level = get_entry(tbl, "integrator.max_refinement_level", "int")
    case.max_refinement_level_geom = get_entry(tbl, "integrator.max_refinement_level_geom", "int",
                                               default=case.max_refinement_level)
    case.refinement_radius = get_entry(tbl, "integrator.refinement_radius", "float", default=2.0)
    case.error_threshold = get_entry(tbl, "integrator.error_threshold", "vec", 2)
    case.weights = get_entry(tbl, "integrator.error_weights", "vec", ne, default=tuple([1.0] * ne))
    _model_dict(case, tbl.get("model", {}))
    # initial condition: config/config.cpp:194-243
    ic = tbl.get("initial_condition")
    if not isinstance(ic, dict):
        raise ConfigError("Config: initial_condition missing")
    case.ic_type = get_entry(ic, "type", "str")
    if case.ic_type == "uniform":
        value = get_entry(ic, "value", "str", default="reference")
        if value == "fixed":
            get_entry(ic, "fixed_value", "vec", ne)  # read and, as the reference, ignored (:208-211)
        elif value != "reference":
            raise ConfigError(f"Unknown uniform value type {value}")
    elif case.ic_type == "jump":
        case.ic_axis = get_entry(ic, "axis", "uint")
        case.ic_location = get_entry(ic, "location", "float")
        case.ic_neg = get_entry(ic, "value_neg", "vec", ne)
        case.ic_pos = get_entry(ic, "value_pos", "vec", ne)
    elif case.ic_type == "grid":
        pass  # filled by the caller (build extension: primitive state on a uniform grid)
    else:
        raise ConfigError("illegal arg")
    # domain boundary conditions: config/config.cpp:47-61
    names = ["X", "Y", "Z"]
    periodic = [False, False, False]
    bcs = [BC() for _ in range(6)]
    bct = tbl.get("boundary_conditions", {})
    for d in range(dim):
        sub = bct.get(names[d], {}) if isinstance(bct, dict) else {}
        if isinstance(sub.get("type"), str):
            periodic[d] = sub["type"] == "periodic"
        else:
            bcs[2 * d] = _parse_bc(case, sub.get("neg", {}), ne)
            bcs[2 * d + 1] = _parse_bc(case, sub.get("pos", {}), ne)
    case.periodic = tuple(periodic)
    case.bcs = bcs
    # geometry: config/config.cpp:66-125 (shapes, meshes and their boundary
    # conditions are read by the C++ scene builder from the same text)
    for g in tbl.get("geometry", []) or []:
        name = get_entry(g, "name", "str")
        gtype = get_entry(g, "type", "str", default="solid")
        tr = get_entry(g, "translation", "vec", 3, default=(0.0, 0.0, 0.0))
        if "path" in g:
            shape = "mesh"
        elif "shape" in g:
            shape = get_entry(g, "shape", "str")
        else:
            raise ConfigError(f"Geometry '{name}' needs to either have an associated file path or a shape.")
        if gtype == "solid" and not (isinstance(bct, dict) and name in bct):
            raise ConfigError(f"Must define boundary condition for solid geometry '{name}'")
        radius = g.get("radius", 0.0)
        case.geometry.append(Geometry(name, gtype, shape, tuple(tr), float(radius) if radius else 0.0))
    for z in tbl.get("refinement_zone", []) or []:
        case.zones.append((get_entry(z, "start", "vec", 3), get_entry(z, "end", "vec", 3),
                           get_entry(z, "max_refinement_level", "uint")))
    return case


def load_case(path, model, dim, limiter="minmod"):
    with open(path, "r") as f:
        text = f.read()
    case = load_case_table(_toml.loads(text), model, dim, limiter)
    case.toml_text, case.base_dir = text, os.path.dirname(os.path.abspath(path))
    return case


def loads_case(text, model, dim, limiter="minmod", base_dir="."):
    case = load_case_table(_toml.loads(text), model, dim, limiter)
    case.toml_text, case.base_dir = text, base_dir
    return case
