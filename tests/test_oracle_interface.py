"""Closed-form pins of the oracle's interface chain (impl/smooth.hpp:9-60,
impl/normal_curvature.hpp:6-101): five-equation fluid at rest at uniform
pressure (an exact steady state: no velocity, so no flux but the pressure's,
which cancels), with a prescribed volume-fraction field; after one global
step the chain has run on every cell.

* a linear ramp of alpha along one axis: the Jacobi smoothing keeps it, the
  normal is the unit vector along the ramp (pointing to larger alpha) and the
  curvature is zero, away from the domain boundaries;
* a smoothed disc of radius R (alpha = 1 inside): the normal points to the
  centre and the curvature -div(n) near the interface is 1/R, up to the
  discretisation error.
"""
import numpy as np
import pytest

from dacti_amd import scene
from dacti_amd.case import loads_case
from oracle.binding import Oracle
from tests.helpers import live_leaves

TOML = """
[scene]
name = "interface_pin"
domain_size = [1.0, 1.0]

[run]
print_interval = 1.0
end_time = 1.0

[integrator]
min_refinement_level = {level}
max_refinement_level = {level}
error_threshold = [-1.0, 1e300]
max_cfl = 0.5

[model.reference_values]
rho = [1.0, 0.5]
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
type = "grid"

[boundary_conditions.X]
type = "zeroGradient"
[boundary_conditions.Y]
type = "zeroGradient"
"""
LEVEL = 5
N = 1 << LEVEL
IKN, IKAPPA, IALPHAS = 2, 4, 5           # k layout of the 2D five-equation model


def _run(alpha):
    case = loads_case(TOML.format(level=LEVEL), "multi", 2)
    prim = np.zeros((N, N, 6))
    prim[..., 0], prim[..., 1], prim[..., 4] = 1.0, 0.5, 1.0
    prim[..., 5] = alpha
    case.ic_grid = prim
    o = Oracle(case, scene.build_tree(case))
    _, fb = o.integrate()
    assert not fb
    a = o.export()
    o.close()
    lv = live_leaves(a)
    return a, lv


def _centres():
    x = (np.arange(N) + 0.5) / N - 0.5
    return np.meshgrid(x, x, indexing="ij")


@pytest.mark.parametrize("axis", [0, 1])
def test_linear_ramp_normal_and_zero_curvature(axis):
    X = _centres()
    alpha = 0.5 + 0.4 * X[axis]
    a, lv = _run(alpha)
    ctr = a["center"][lv]
    inner = (np.abs(ctr) < 0.5 - 5.0 / N).all(axis=1)   # four cells clear of the boundaries
    k = a["k"][lv][inner]
    p = a["p"][lv][inner]
    np.testing.assert_allclose(a["p"][lv][:, :5], np.broadcast_to([1.0, 0.5, 0.0, 0.0, 1.0], (len(lv), 5)),
                               rtol=0, atol=1e-14)                                # still at rest
    np.testing.assert_allclose(k[:, IALPHAS], p[:, 5], rtol=0, atol=1e-14)        # smoothing keeps a ramp
    np.testing.assert_allclose(k[:, IKN + axis], 1.0, rtol=0, atol=1e-12)         # along the ramp, up-gradient
    np.testing.assert_array_equal(k[:, IKN + 1 - axis], 0.0)
    np.testing.assert_allclose(k[:, IKAPPA], 0.0, rtol=0, atol=1e-9)


def test_disc_curvature_is_one_over_radius():
    X, Y = _centres()
    R = 0.25
    r = np.hypot(X, Y)
    alpha = 0.5 * (1.0 - np.tanh((r - R) / (1.5 / N)))
    a, lv = _run(alpha)
    ctr = a["center"][lv]
    k = a["k"][lv]
    al = k[:, IALPHAS]
    band = (al > 0.2) & (al < 0.8)
    assert band.sum() > 20
    rr = np.hypot(ctr[band, 0], ctr[band, 1])
    # the normal points to the centre: n . r_hat = -1
    ndot = (k[band, IKN] * ctr[band, 0] + k[band, IKN + 1] * ctr[band, 1]) / rr
    np.testing.assert_allclose(ndot, -1.0, atol=0.05)
    kap = k[band, IKAPPA]
    assert abs(np.mean(kap * rr) - 1.0) < 0.1, np.mean(kap * rr)   # kappa ~ 1/r at the interface
