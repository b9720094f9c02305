"""Case builders for the benchmark workloads (TOML text + generated fields).

periodic_box:  3D single-phase Euler in a periodic unit box on a uniform
               octree (level 8 = 256^3 leaves at full size); density
               1 + 0.1 g with g white noise smoothed by a periodic Gaussian
               (sigma = 4 cells) and renormalised to unit std; velocity a
               smoothed, divergence-free random field with RMS Mach 0.5;
               p = 1, gamma = 1.4; no refinement.
shock_bubble:  2D five-equation shock-bubble on [-1,1]x[-0.5,0.5] (the
               [0,2]x[0,1] domain centred at the origin, as the reference
               centres every domain): Mach 1.22 shock in air hitting a
               helium-like bubble.
"""
import math

import numpy as np

from .case import loads_case


def smoothed_noise(n, dim, seed, sigma_cells):
    """White N(0,1) noise on an n^dim periodic grid, Gaussian-smoothed."""
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n,) * dim)
    k = 2.0 * np.pi * np.fft.fftfreq(n)
    ks = np.meshgrid(*([k] * dim), indexing="ij")
    k2 = sum(kk * kk for kk in ks)
    f = np.fft.fftn(w) * np.exp(-0.5 * sigma_cells ** 2 * k2)
    return np.real(np.fft.ifftn(f))


def solenoidal_velocity(n, dim, seeds, sigma_cells, rms):
    k = 2.0 * np.pi * np.fft.fftfreq(n)
    ks = np.meshgrid(*([k] * dim), indexing="ij")
    k2 = sum(kk * kk for kk in ks)
    k2[(0,) * dim] = 1.0
    uh = [np.fft.fftn(smoothed_noise(n, dim, s, sigma_cells)) for s in seeds]
    div = sum(kk * u for kk, u in zip(ks, uh)) / k2
    uh = [u - kk * div for kk, u in zip(ks, uh)]
    for u in uh:
        u[(0,) * dim] = 0.0
    vel = np.stack([np.real(np.fft.ifftn(u)) for u in uh], axis=-1)
    cur = math.sqrt(float((vel ** 2).sum(axis=-1).mean()))
    return vel * (rms / cur)


PERIODIC_BOX_TOML = """
[scene]
name = "periodic_box"
domain_size = [1.0, 1.0, 1.0]

[run]
print_interval = 1.0
end_time = 1.0

[integrator]
min_refinement_level = {level}
max_refinement_level = {level}
error_threshold = [-1.0, 1e300]
max_cfl = 0.5

[model.reference_values]
rho = 1.0
p = 1.0
vel = [0.0, 0.0, 0.0]

[initial_condition]
type = "grid"

[boundary_conditions.X]
type = "periodic"
[boundary_conditions.Y]
type = "periodic"
[boundary_conditions.Z]
type = "periodic"
"""


def periodic_box(level=8, sigma_cells=4.0, mach_rms=0.5, seed=1, dim=3):
    """The 3D periodic Euler roofline case at 2^level cells per axis."""
    n = 1 << level
    toml = PERIODIC_BOX_TOML.format(level=level)
    if dim == 2:
        toml = toml.replace("[1.0, 1.0, 1.0]", "[1.0, 1.0]").replace("[0.0, 0.0, 0.0]", "[0.0, 0.0]")
        toml = toml.replace('[boundary_conditions.Z]\ntype = "periodic"\n', "")
    case = loads_case(toml, "euler", dim)
    g = smoothed_noise(n, dim, seed, sigma_cells)
    g = (g - g.mean()) / g.std()
    rho = 1.0 + 0.1 * g
    vel = solenoidal_velocity(n, dim, [seed + 1 + i for i in range(dim)], sigma_cells, mach_rms * math.sqrt(1.4))
    prim = np.empty((n,) * dim + (dim + 2,))
    prim[..., 0] = rho
    prim[..., 1:1 + dim] = vel
    prim[..., dim + 1] = 1.0
    case.ic_grid = prim
    return case


def normal_shock(rho1, p1, mach, gamma=1.4, P0=0.0):
    """Post-shock state behind a shock moving into gas at rest
    (config/util.hpp:135-150)."""
    c1 = math.sqrt(gamma * (p1 + P0) / rho1)
    r_ratio = 1.0 - 2.0 / (gamma + 1) * (1 - 1 / (mach * mach))
    r2 = rho1 / r_ratio
    u2 = 2.0 / (gamma + 1) * (mach - 1 / mach) * c1
    p2 = p1 * (((2 * P0 / p1 + 1) * (1 - r_ratio) * gamma + 1 + r_ratio) / ((r_ratio - 1) * gamma + 1 + r_ratio))
    return r2, u2, p2


SHOCK_BUBBLE_TOML = """
[scene]
name = "shock_bubble"
domain_size = [2.0, 1.0]

[run]
print_interval = 1.0
end_time = 1.0

[integrator]
min_refinement_level = {base}
max_refinement_level = {top}
refinement_radius = 2.0
error_threshold = [{emin}, {emax}]
error_weights = [1.0, 1.0, 0.0, 0.0, 0.0, 1.0]
max_cfl = 0.5

[model.reference_values]
rho = [1.0, {rho2}]
vel = [0.0, 0.0]
p = 1.0
alpha = 1.0

[model.fluid_1]
gamma = 1.4
P0 = 0.0

[model.fluid_2]
gamma = 1.67
P0 = 0.0

[initial_condition]
type = "jump"
axis = 0
location = {shock_x}
value_neg = [{r2}, 0.0, {u2}, 0.0, {p2}, 1.0]
value_pos = [1.0, 0.0, 0.0, 0.0, 1.0, 1.0]

[[geometry]]
name = "bubble"
shape = "circle"
type = "bubble"
radius = {radius}
translation = [{bx}, 0.0, 0.0]

[boundary_conditions.X.neg]
type = "inflow_r"
value = [{r2}, 0.0, {u2}, 0.0, {p2}, 1.0]

[boundary_conditions.X.pos]
type = "zeroGradient"

[boundary_conditions.Y.neg]
type = "zeroGradient"

[boundary_conditions.Y.pos]
type = "zeroGradient"
"""


def shock_bubble(base=6, extra=4, mach=1.22, radius=0.25, bubble_x=-0.3, shock_x=-0.7, rho2=0.138,
                 error_threshold=(0.02, 0.05)):
    """2D shock-bubble, five-equation model (fluid 1 = air, alpha = 1 outside)."""
    r2, u2, p2 = normal_shock(1.0, 1.0, mach)
    toml = SHOCK_BUBBLE_TOML.format(base=base, top=base + extra, emin=repr(float(error_threshold[0])),
                                    emax=repr(float(error_threshold[1])), rho2=repr(float(rho2)),
                                    shock_x=repr(float(shock_x)), r2=repr(r2), u2=repr(u2), p2=repr(p2),
                                    radius=repr(float(radius)), bx=repr(float(bubble_x)))
    return loads_case(toml, "multi", 2)


def shock_bubble_family(n, seed=0, base=6, extra_range=(4, 7), mach_range=(1.1, 3.0), radius_range=(0.1, 0.3),
                        ratio_range=(0.1, 10.0)):
    """n independent 2D shock-bubble cases for batched runs.

    Instance k draws, from a generator seeded with (seed, k): the shock Mach
    number uniform in mach_range, the bubble radius uniform in radius_range,
    the bubble/ambient density ratio log-uniform in ratio_range and the number
    of refinement levels above the base uniform in extra_range (inclusive).
    Returns a list of (case, params) pairs.
    """
    out = []
    for k in range(n):
        g = np.random.default_rng([seed, k])
        mach = float(g.uniform(*mach_range))
        radius = float(g.uniform(*radius_range))
        ratio = float(np.exp(g.uniform(np.log(ratio_range[0]), np.log(ratio_range[1]))))
        extra = int(g.integers(extra_range[0], extra_range[1] + 1))
        case = shock_bubble(base=base, extra=extra, mach=mach, radius=radius, rho2=ratio)
        out.append((case, {"mach": mach, "radius": radius, "ratio": ratio, "extra": extra}))
    return out


SHOCK_DROP_TOML = """
[scene]
name = "shock_drop"
domain_size = [0.016, 0.008, 0.008]

[run]
print_interval = 1.0
end_time = 1.0

[integrator]
min_refinement_level = {base}
max_refinement_level = {top}
error_threshold = [{emin}, {emax}]
error_weights = [1.0e-3, 1.0, 0.0, 0.0, 0.0, 1.0e-5, 1.0]
max_cfl = 0.5

[model.reference_values]
rho = [1000.0, 1.2]
vel = [0.0, 0.0, 0.0]
p = 1.0e5
alpha = 0.0

[model.fluid_1]
gamma = 6.12
P0 = 3.43e8

[model.fluid_2]
gamma = 1.4
P0 = 0.0

[initial_condition]
type = "jump"
axis = 0
location = {shock_x}
value_neg = [0.0, {r2}, {u2}, 0.0, 0.0, {p2}, 0.0]
value_pos = [0.0, 1.2, 0.0, 0.0, 0.0, 1.0e5, 0.0]

[[geometry]]
name = "drop"
shape = "sphere"
type = "droplet"
radius = {radius}
translation = [{dx}, 0.0, 0.0]

[boundary_conditions.X.neg]
type = "inflow_r"
value = [0.0, {r2}, {u2}, 0.0, 0.0, {p2}, 0.0]

[boundary_conditions.X.pos]
type = "zeroGradient"

[boundary_conditions.Y]
type = "zeroGradient"

[boundary_conditions.Z]
type = "zeroGradient"
"""


def shock_drop(base=5, extra=5, mach=2.4, radius=0.001, drop_x=0.0, shock_x=-0.003, error_threshold=(0.02, 0.05),
               max_dt=None):
    """3D shock-drop, five-equation stiffened gas in SI units: a Mach `mach`
    shock in air (gamma 1.4, 1.2 kg/m^3, 1e5 Pa) hits a water drop (gamma
    6.12, P0 3.43e8 Pa, 1000 kg/m^3) in a 16 x 8 x 8 mm box (fluid 1 = water,
    alpha = 1 inside the drop).  max_dt (integrator.max_dt) caps the global
    step, and with it the spread of LTS levels."""
    r2, u2, p2 = normal_shock(1.2, 1.0e5, mach)
    toml = SHOCK_DROP_TOML.format(base=base, top=base + extra, emin=repr(float(error_threshold[0])),
                                  emax=repr(float(error_threshold[1])), shock_x=repr(float(shock_x)), r2=repr(r2),
                                  u2=repr(u2), p2=repr(p2), radius=repr(float(radius)), dx=repr(float(drop_x)))
    if max_dt is not None:
        toml = toml.replace("max_cfl = 0.5\n", f"max_cfl = 0.5\nmax_dt = {float(max_dt)!r}\n", 1)
    return loads_case(toml, "multi", 3)
