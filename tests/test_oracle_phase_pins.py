"""Pins of the oracle's per-phase functions by closed forms, brute force and
invariants of the scheme.

Each test names the reference code it restates and checks it against what
the mathematics fixes, so that a dropped term, a wrong sign or index, or a
level off by one fails:

* compute_rhs (model/euler.hpp:86-116, model/multi.hpp:126-169): the
  quasi-linear primitive equations written out term by term;
* middle_state with surface tension (model/multi.hpp:205-246): a static
  interface in Laplace equilibrium has zero contact speed;
* Domain::split (integrator/domain.hpp:192-247): children of a cell with a
  linear reconstruction hold the reconstruction at their centres;
* Domain::unsplit / interpolated_cell (domain.hpp:150-180,296-325): the
  parent holds the mean of its children's conserved state;
* LTS reassignment (impl/adaptive_refinement.hpp:264-322): brute-force
  search of the smallest stable level;
* boundary_flux (integrator/boundary.hpp:130-160): wall, outflow and
  supersonic-inflow fluxes in closed form;
* flux ownership (impl/flux.hpp:58-59): every face is covered for the
  whole global time step from both sides, except the coarse/fine faces whose
  coarse side sits on a finer time level, which the reference loses;
* update_neighbors / should_split (domain.hpp:249-293): brute-force
  neighbour search over all nodes and the 2:1 flag;
* compute_slope + interpolate_p (impl/slopes.hpp:15-76,
  domain.hpp:84-148): exact gradient of a linear field across tree levels;
* the MUSCL-Hancock predictor (model/euler.hpp:156-200): second-order
  convergence on a smooth advected wave;
* the five-equation update and sources (impl/cell_update.hpp,
  model/multi.hpp:186-203): pressure and velocity equilibrium across an
  advected material interface, and the Laplace pressure jump of a disc
  under surface tension;
* gc (domain.hpp:327-356, impl/acti_utils.hpp:58-74): compaction keeps the
  leaf set and re-bins the LTS levels from the global dt;
* the LTS schedule (impl/integrate.hpp:11-87): every leaf ends a global step
  at the same time with a stable CFL number.
"""
import ctypes as C
import math

import numpy as np
import pytest

from dacti_amd import scene, workloads
from dacti_amd.case import _multi_gamma, _multi_p0, loads_case
from oracle.binding import OParams, Oracle, model_fn, params
from tests.helpers import amr_sod_case, live_leaves

D3 = C.POINTER(C.c_double)
PP = C.POINTER(OParams)


def arr(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def ptr(a):
    return a.ctypes.data_as(D3)


def euler_params(dim):
    return params(workloads.periodic_box(level=2, dim=dim))


def multi_params(dim=2, sigma=0.0):
    P = params(workloads.shock_bubble(base=3, extra=1)) if dim == 2 else params(
        workloads.shock_drop(base=2, extra=1))
    P.sigma = sigma
    return P


def pconsts(P):
    return {k: getattr(P, k) for k in ("gamma1", "gamma2", "P1", "P2")}


# ------------------------------------------------------------ compute_rhs
@pytest.mark.parametrize("dim", [2, 3])
def test_rhs_euler_quasilinear_closed_form(oracle_lib, dim):
    """rhs = sum_j A_j(p) slope_j for the primitive Euler equations:
    rho_t + u.grad rho + rho div u, u_t + (u.grad) u + grad p / rho,
    p_t + u.grad p + gamma p div u."""
    P = euler_params(dim)
    ne = dim + 2
    f = model_fn(0, dim, "t_compute_rhs", None, [PP, D3, D3, C.c_double, D3])
    rng = np.random.default_rng(11)
    for _ in range(60):
        p = arr([rng.uniform(0.3, 3), *rng.normal(size=dim), rng.uniform(0.3, 3)])
        sl = arr(rng.normal(size=(dim, ne)))
        rhs = np.zeros(ne)
        f(C.byref(P), ptr(p), ptr(sl), 0.0, ptr(rhs))
        rho, u, pr = p[0], p[1:1 + dim], p[-1]
        want = np.zeros(ne)
        for j in range(dim):
            s = sl[j]
            want[0] += u[j] * s[0] + rho * s[1 + j]
            for i in range(dim):
                want[1 + i] += u[j] * s[1 + i]
            want[1 + j] += s[-1] / rho
            want[-1] += u[j] * s[-1] + 1.4 * pr * s[1 + j]
        np.testing.assert_allclose(rhs, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_rhs_multi_quasilinear_closed_form(oracle_lib, dim):
    """Five-equation model in primitives (r0, r1, u, p, alpha): partial
    densities advected and compressed, momentum with pressure and the
    surface-tension term sigma kappa grad(alpha) / rho, pressure with
    rho c^2 div u (mixture sound speed), alpha advected."""
    sig, kap = 0.7, -1.3
    P = multi_params(dim, sigma=sig)
    ne = dim + 4
    ip, ia = ne - 2, ne - 1
    f = model_fn(2, dim, "t_compute_rhs", None, [PP, D3, D3, C.c_double, D3])
    rng = np.random.default_rng(12)
    cst = pconsts(P)
    for _ in range(60):
        a = rng.uniform(0.05, 0.95)
        r0, r1 = rng.uniform(0.1, 2.0) * a, rng.uniform(0.1, 2.0) * (1 - a)
        pr = rng.uniform(1e5, 1e6) if dim == 3 else rng.uniform(0.5, 3)
        p = arr([r0, r1, *rng.normal(size=dim), pr, a])
        sl = arr(rng.normal(size=(dim, ne)))
        rhs = np.zeros(ne)
        f(C.byref(P), ptr(p), ptr(sl), kap, ptr(rhs))
        rho = r0 + r1
        c2 = _multi_gamma(cst, a) * (pr + _multi_p0(cst, a)) / rho
        u = p[2:2 + dim]
        want = np.zeros(ne)
        for j in range(dim):
            s = sl[j]
            want[0] += u[j] * s[0] + r0 * s[2 + j]
            want[1] += u[j] * s[1] + r1 * s[2 + j]
            for i in range(dim):
                want[2 + i] += u[j] * s[2 + i]
            want[2 + j] += (s[ip] + sig * kap * s[ia]) / rho
            want[ip] += u[j] * s[ip] + rho * c2 * s[2 + j]
            want[ia] += u[j] * s[ia]
        np.testing.assert_allclose(rhs, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())


# ------------------------------------------------------ star state, sigma
@pytest.mark.parametrize("dim", [0, 1])
def test_star_state_laplace_equilibrium(oracle_lib, dim):
    """A static interface whose pressure jump balances surface tension,
    p_+ - p_- = sigma kappa (alpha_+ - alpha_-), has zero contact speed:
    no mass or volume crosses it.  Without the sigma kappa term of the
    contact speed it would move."""
    sig, kap = 0.5, 2.0
    P = multi_params(2, sigma=sig)
    f = model_fn(2, 2, "t_middle_state", None, [PP, D3, D3, C.c_double, C.c_int, D3])
    for an, ap, pn_ in ((1.0, 0.0, 2.0), (0.0, 1.0, 1.0), (0.75, 0.25, 3.0)):
        pp_ = pn_ + sig * kap * (ap - an)
        pn = arr([1.0 * an, 0.5 * (1 - an), 0.0, 0.0, pn_, an])
        pr = arr([1.0 * ap, 0.5 * (1 - ap), 0.0, 0.0, pp_, ap])
        pm = np.zeros(6)
        f(C.byref(P), ptr(pn), ptr(pr), kap, dim, ptr(pm))
        assert pm[2 + dim] == 0.0
        # the same states without surface tension are not in equilibrium
        f(C.byref(P), ptr(pn), ptr(pr), 0.0, dim, ptr(pm))
        assert abs(pm[2 + dim]) > 1e-3


# ----------------------------------------------------------- split / unsplit
def _split(model, dim, parent):
    f = model_fn(model, dim, "t_split", None, [PP, D3, D3])
    P = euler_params(dim) if model == 0 else multi_params(dim)
    from oracle.binding import sizes
    _, ne, _, npar, ng = sizes(model, dim)
    npx, ngx = max(npar, 1), max(ng, 1)
    q = arr(np.concatenate([parent["p"], parent["slope"].ravel(), parent["k"], parent["grad"].ravel(),
                            parent["t_slope"], parent["center"], [parent["size"], parent["t"]]]))
    per = ne + ne + dim + 2 + dim * ne + npx + dim * ngx + ngx
    out = np.zeros((1 << dim) * per)
    f(C.byref(P), ptr(q), ptr(out))
    kids = []
    for ch in range(1 << dim):
        o = out[ch * per:(ch + 1) * per]
        k = {}
        pos = 0
        for name, n in (("p", ne), ("u", ne), ("center", dim), ("size", 1), ("t", 1), ("slope", dim * ne),
                        ("k", npx), ("grad", dim * ngx), ("t_slope", ngx)):
            k[name] = o[pos:pos + n]
            pos += n
        kids.append(k)
    return P, kids, ne, npx, ngx


@pytest.mark.parametrize("model,dim", [(0, 2), (0, 3), (2, 2), (2, 3)])
def test_split_children_hold_linear_reconstruction(oracle_lib, model, dim):
    """Child c of a split sits at centre x_c = x + q s_c (s_c in {-1,1}^D by
    the bits of c, q = size/4) and holds the parent's linear reconstruction
    p + sum_d slope_d (x_c - x)_d in primitives; the rest is inherited."""
    rng = np.random.default_rng(20 + dim + model)
    from oracle.binding import sizes
    _, ne, _, npar, ng = sizes(model, dim)
    npx, ngx = max(npar, 1), max(ng, 1)
    if model == 0:
        p = np.array([1.3, *rng.normal(size=dim) * 0.3, 2.1])
    else:
        p = np.array([0.6, 0.2, *rng.normal(size=dim) * 0.3, 2.1 if dim == 2 else 2e5, 0.7])
    parent = {"p": p, "slope": rng.normal(size=(dim, ne)) * 0.2, "k": rng.normal(size=npx),
              "grad": rng.normal(size=(dim, ngx)), "t_slope": rng.normal(size=ngx),
              "center": rng.uniform(-0.4, 0.4, size=dim), "size": 0.125, "t": 0.375}
    P, kids, _, _, _ = _split(model, dim, parent)
    cons = model_fn(model, dim, "t_cons_from_prim", None, [PP, D3, D3])
    for ch, k in enumerate(kids):
        sgn = np.array([1.0 if (ch >> d) & 1 else -1.0 for d in range(dim)])
        xc = parent["center"] + sgn * 0.25 * parent["size"]
        np.testing.assert_array_equal(k["center"], xc)
        want = parent["p"] + (parent["slope"] * (xc - parent["center"])[:, None]).sum(0)
        np.testing.assert_allclose(k["p"], want, rtol=1e-14, atol=1e-15)
        u = np.zeros(ne)
        cons(C.byref(P), ptr(arr(k["p"])), ptr(u))
        np.testing.assert_array_equal(k["u"], u)
        assert k["size"][0] == 0.5 * parent["size"] and k["t"][0] == parent["t"]
        np.testing.assert_array_equal(k["slope"], parent["slope"].ravel())
        np.testing.assert_array_equal(k["k"], parent["k"])
        np.testing.assert_array_equal(k["grad"], parent["grad"].ravel())
        np.testing.assert_array_equal(k["t_slope"], parent["t_slope"])


@pytest.mark.parametrize("model,dim", [(0, 2), (0, 3), (2, 2), (1, 3)])
def test_unsplit_parent_is_mean_of_children(oracle_lib, model, dim):
    """The merged parent holds the mean conserved state, time, parameters and
    time slopes of its children, its primitives follow from the mean
    conserved state, its error is the children's sum, and its gradient is
    (stored parent gradient + children's gradients) / 2^D (the reference
    starts the gradient sum from the parent's stale value)."""
    from oracle.binding import sizes
    _, ne, _, npar, ng = sizes(model, dim)
    npx, ngx = max(npar, 1), max(ng, 1)
    nch = 1 << dim
    rng = np.random.default_rng(30 + dim + model)
    P = euler_params(dim) if model == 0 else (multi_params(dim) if model == 2 else params(
        amr_sod_case(model="navier_stokes")) if dim == 2 else euler_params(dim))
    if model == 1:
        from tests.helpers import ns3_box
        P = params(ns3_box(level=2))
    prim = model_fn(model, dim, "t_prim_from_cons", None, [PP, D3, D3])
    cons = model_fn(model, dim, "t_cons_from_prim", None, [PP, D3, D3])
    kids = []
    for _ in range(nch):
        if model == 2:
            a = rng.uniform(0.1, 0.9)
            p = arr([a * rng.uniform(0.5, 1.5), (1 - a) * rng.uniform(0.5, 1.5), *rng.normal(size=dim) * 0.3,
                     rng.uniform(1, 3) if dim == 2 else rng.uniform(1e5, 3e5), a])
        else:
            p = arr([rng.uniform(0.5, 2), *rng.normal(size=dim) * 0.3, rng.uniform(1, 3)])
        u = np.zeros(ne)
        cons(C.byref(P), ptr(p), ptr(u))
        kids.append({"u": u, "t": rng.uniform(0, 1), "k": rng.normal(size=npx), "grad": rng.normal(size=(dim, ngx)),
                     "t_slope": rng.normal(size=ngx), "error": rng.uniform(0, 1)})
    pgrad = arr(rng.normal(size=(dim, ngx)))
    q = arr(np.concatenate([np.concatenate([k["u"], [k["t"]], k["k"], k["grad"].ravel(), k["t_slope"],
                                            [k["error"]]]) for k in kids]))
    out = np.zeros(ne + ne + 1 + npx + dim * ngx + ngx + 1)
    f = model_fn(model, dim, "t_unsplit", None, [PP, D3, D3, D3])
    f(C.byref(P), ptr(q), ptr(pgrad), ptr(out))
    u, p, t = out[:ne], out[ne:2 * ne], out[2 * ne]
    k = out[2 * ne + 1:2 * ne + 1 + npx]
    g = out[2 * ne + 1 + npx:2 * ne + 1 + npx + dim * ngx]
    ts = out[2 * ne + 1 + npx + dim * ngx:-1]
    np.testing.assert_allclose(u, np.mean([c["u"] for c in kids], axis=0), rtol=1e-14)
    pw = np.zeros(ne)
    prim(C.byref(P), ptr(arr(u)), ptr(pw))
    np.testing.assert_array_equal(p, pw)
    assert t == pytest.approx(np.mean([c["t"] for c in kids]), rel=1e-14)
    np.testing.assert_allclose(k, np.mean([c["k"] for c in kids], axis=0), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(g, (pgrad + sum(c["grad"] for c in kids)).ravel() / nch, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(ts, np.mean([c["t_slope"] for c in kids], axis=0), rtol=1e-13, atol=1e-15)
    assert out[-1] == pytest.approx(sum(c["error"] for c in kids), rel=1e-14)


# ------------------------------------------------------------ LTS decision
def test_lts_reassignment_brute_force(oracle_lib):
    """After a substep of level L a leaf with CFL number c = dt_L / max_dt:
    * c > max_cfl: moves to the smallest finer level L + k (k >= 1) on which
      c / 2^k < max_cfl, i.e. it is stable there and c / 2^k >= max_cfl / 2;
      dropped from every level when that level would exceed 15;
    * c < max_cfl / 2, on every second substep (counter 1) and L > 0: waits
      in the buffer of level L - 1 with c doubled;
    * otherwise, or not a leaf (dropped): unchanged."""
    P = euler_params(2)
    mc = P.max_cfl
    f = model_fn(0, 2, "t_lts_reassign", None, [PP, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, D3])
    rng = np.random.default_rng(40)
    cfls = list(rng.uniform(0.0, 3.0, 60) * mc) + list(np.exp(rng.uniform(0, 12, 40)) * mc) + \
        [mc, 2 * mc, 4 * mc, 0.5 * mc, 0.25 * mc, 1e6 * mc]
    out = np.zeros(4)
    for L in range(16):
        dt = 0.5 ** L
        for c in cfls:
            if c <= 0:
                continue
            for counter in (0, 1):
                for leaf in (0, 1):
                    mdt = dt / c
                    f(C.byref(P), mdt, L, dt, counter, leaf, L, ptr(out))
                    cl, buf, acti, cfl = int(out[0]), int(out[1]), int(out[2]), out[3]
                    c0 = dt / mdt
                    if not leaf:
                        assert cl == -1 and buf == -1 and acti == L
                    elif c0 < 0.5 * mc and counter == 1 and L > 0:
                        assert (cl, buf, acti) == (-1, L - 1, L) and cfl == 2 * c0
                    elif c0 > mc and L < 15:
                        k = 1
                        while c0 / 2 ** k >= mc:
                            k += 1
                        if L + k < 16:
                            assert (cl, buf, acti) == (L + k, -1, L + k), (L, c0 / mc, k, out)
                            assert cfl == c0 / 2 ** k and 0.5 * mc <= cfl < mc
                        else:
                            assert (cl, buf, acti) == (-1, -1, L)
                    else:
                        assert (cl, buf, acti) == (L, -1, L) and cfl == c0


# ----------------------------------------------------------- boundary flux
def _bflux(model, dim, P, face, p, dim_, dir_, t=0.0, grad=None, area=0.5, t0=0.25, t1=0.5):
    from oracle.binding import sizes
    _, ne, nf, _, ng = sizes(model, dim)
    g = arr(np.zeros((dim, max(ng, 1))) if grad is None else grad)
    f = model_fn(model, dim, "t_boundary_flux", None,
                 [PP, C.c_int, D3, D3, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, D3])
    out = np.zeros(nf)
    f(C.byref(P), face, ptr(arr(p)), ptr(g), t, dim_, dir_, area, t0, t1, ptr(out))
    return out


def _euler_F(p, d, gamma=1.4):
    rho, u, pr = p[0], np.asarray(p[1:-1]), p[-1]
    un = u[d]
    F = np.zeros(len(p))
    F[0] = rho * un
    F[1:-1] = rho * un * u
    F[1 + d] += pr
    F[-1] = un * (gamma / (gamma - 1) * pr + 0.5 * rho * (u @ u))
    return F


def test_boundary_flux_closed_forms(oracle_lib):
    """Domain-face fluxes, scaled by area (t1 - t0) and the face direction
    (-1 on a negative face, +1 on a positive one):
    freeslip and noslip walls carry only the wall pressure; outflow and
    zeroGradient the physical flux of the cell state; a supersonic inflow
    through a negative face the physical flux of the inflow state."""
    P = euler_params(2)
    scale = 0.5 * 0.25
    p = np.array([1.2, 0.3, -0.4, 2.0])
    for face in range(4):
        d, dr = face >> 1, face & 1
        sgn = 2 * dr - 1
        for bct in (1, 2):  # freeslip, noslip
            P.bc_type[face] = bct
            out = _bflux(0, 2, P, face, p, d, dr)
            want = np.zeros(4)
            want[1 + d] = p[-1]
            np.testing.assert_allclose(out, sgn * scale * want, rtol=1e-14, atol=1e-15)
        for bct in (0, 5):  # zeroGradient, outflow
            P.bc_type[face] = bct
            out = _bflux(0, 2, P, face, p, d, dr)
            np.testing.assert_allclose(out, sgn * scale * _euler_F(p, d), rtol=1e-14, atol=1e-15)
    # dirichlet inflow, supersonic into the domain through the x- face
    P.bc_type[0] = 6
    pin = [1.0, 5.0, 0.2, 1.0]
    for e, v in enumerate(pin):
        P.bc_p[0][e] = v
    cell = np.array([0.9, 4.0, 0.1, 1.1])
    out = _bflux(0, 2, P, 0, cell, 0, 0)
    np.testing.assert_allclose(out, -scale * _euler_F(np.array(pin), 0), rtol=1e-14)
    # the same face with an inflow boundary faded in over 2 time units: at
    # t = 0.5 the ghost velocity is a quarter of the inflow's; a subsonic
    # cell keeps its own pressure in the ghost
    P.bc_type[0] = 3
    P.bc_fade_in[0] = 2.0
    pb = np.zeros(4)
    bv = model_fn(0, 2, "t_boundary_value", None, [PP, C.c_int, D3, C.c_double, D3])
    sub = arr([1.0, 0.1, 0.0, 1.5])
    bv(C.byref(P), 0, ptr(sub), 0.5, ptr(pb))
    np.testing.assert_allclose(pb, [1.0, 1.25, 0.05, 1.5], rtol=1e-15)


def test_boundary_flux_viscous_wall(oracle_lib):
    """Navier-Stokes noslip wall: pressure minus the Newtonian stress of the
    cell's velocity gradient, mu (du_i/dx_d + du_d/dx_i - 2/3 delta_id div u);
    no heat flux and no stress work at the wall (zero wall velocity)."""
    case = amr_sod_case(model="navier_stokes")
    P = params(case)
    mu = P.ns_mu
    assert mu > 0
    p = np.array([1.2, 0.3, -0.4, 2.0])
    g = np.array([[0.7, -0.2, 0.05], [0.4, 0.9, -0.1]])  # g[a][e]: d(u_e)/dx_a, e < D; e = D temperature
    for face in range(4):
        d, dr = face >> 1, face & 1
        P.bc_type[face] = 2
        out = _bflux(1, 2, P, face, p, d, dr, grad=g)
        tr = g[0][0] + g[1][1]
        want = np.zeros(4)
        want[1 + d] = p[-1]
        for i in range(2):
            want[1 + i] -= mu * (g[d][i] + g[i][d] - (2.0 / 3.0 * tr if i == d else 0.0))
        np.testing.assert_allclose(out, (2 * dr - 1) * 0.125 * want, rtol=1e-13, atol=1e-16)


def test_limiter_von_alba(oracle_lib):
    """von_alba_2 (integrator/slope_limiters.hpp:38-43): phi(r) = 2r/(r^2+1)."""
    f = model_fn(0, 3, "t_limiter", C.c_double, [C.c_int, C.c_double, C.c_double])
    rng = np.random.default_rng(5)
    for _ in range(300):
        s1, s2 = rng.normal(size=2) * 10 ** rng.uniform(-2, 2, size=2)
        r = s1 / s2
        assert f(4, s1, s2) == pytest.approx(2 * r / (r * r + 1) * s2, rel=1e-12, abs=1e-300)


# ------------------------------------------------ flux ownership coverage
def _static_zone_case(dim, mach=0.05, seed=3):
    from tests.test_oracle_amr_conservation import _case
    return _case(dim)


@pytest.mark.parametrize("dim", [2, 3])
def test_every_face_covered_once_per_global_step(oracle_lib, dim):
    """On a static mesh with faces between tree levels and between time
    levels (no coarse leaf on a finer time level than its fine neighbour),
    the fluxes applied to every face of every leaf over one global step span
    exactly the global time step: each face is evaluated once per substep
    of its faster side and reaches the slower side through the deposits,
    sub-faces weighted by their share of the face.  A face evaluated twice,
    skipped, or deposited to the wrong face changes the sum."""
    case = _static_zone_case(dim)
    o = Oracle(case, scene.build_tree(case))
    for step in range(2):
        dt, fb = o.integrate()
        assert not fb
        a = o.export()
        lv = live_leaves(a)
        assert len(np.unique(a["acti_level"][lv])) >= 2 and len(np.unique(a["tree_level"][lv])) >= 2
        cov = o.face_cover()[lv]
        # sums of dt / 2^L in different orders: equal up to rounding
        bad = np.argwhere(np.abs(cov - dt) > 1e-14 * dt)
        assert bad.size == 0, (step, bad[:5], cov[tuple(bad[0])] / dt)
    o.close()


A12_TOML = """
[scene]
name = "lost_flux"
domain_size = [1.0, 1.0]
[run]
print_interval = 1.0
end_time = 1.0
[integrator]
min_refinement_level = 4
max_refinement_level = 4
error_threshold = [-1.0, 1e300]
max_cfl = 0.5
[model.reference_values]
rho = 1.0
p = 1.0
vel = [0.0, 0.0]
[initial_condition]
type = "uniform"
[boundary_conditions.X]
type = "periodic"
[boundary_conditions.Y]
type = "periodic"
[[refinement_zone]]
start = [-0.25, -1.0, -1.0]
end = [0.25, 1.0, 1.0]
max_refinement_level = 5
"""


def a12_case():
    """Contact at rest between a refined strip of cold dense gas (|x| < 1/4,
    level 5) and hot light gas (rho 0.04, level 4) at uniform pressure: the
    coarse leaves' sound speed is five times larger, so they sit on a finer
    time level than their fine neighbours."""
    case = loads_case(A12_TOML, "euler", 2)
    tree = scene.build_tree(case)
    leaves = np.nonzero(tree.flags & 1)[0]
    coarse = leaves[tree.level[leaves] == 4]
    tree.p[coarse, 0] = 0.04
    return case, tree


def test_lost_flux_of_coarse_leaf_on_finer_time_level(oracle_lib):
    """The coarse/fine faces of a coarse leaf on a finer time level than its
    fine neighbours are evaluated by neither side (the coarse side only
    evaluates sub-faces of fine leaves in its own cluster, the fine side
    only neighbours on its own or a slower time level): those faces have
    zero time coverage on both sides, every other face the full step."""
    case, tree = a12_case()
    o = Oracle(case, tree)
    dt, fb = o.integrate()
    assert not fb
    a = o.export()
    lv = live_leaves(a)
    cov = o.face_cover()
    lev, acti, nbr = a["tree_level"], a["acti_level"], a["nbr"]
    assert acti[lv][lev[lv] == 4].min() > acti[lv][lev[lv] == 5].max()
    lost = np.zeros(cov.shape, bool)
    for y in lv:
        for f in range(4):
            nb = nbr[y, f]
            if nb == 0:
                continue
            if lev[y] == 5 and (a["flags"][nb] & 1) and lev[nb] == 4:
                lost[y, f] = True          # fine side: its neighbour is the coarse leaf
            if lev[y] == 4 and not (a["flags"][nb] & 1):
                lost[y, f] = True          # coarse side: its neighbour is refined
    xs = a["center"][:, 0]
    assert lost[lv].sum() == 2 * 16 + 2 * 32  # both strip edges, 16 coarse and 32 fine faces each
    assert np.all(np.abs(np.abs(xs[lv][lost[lv].any(1)]) - 0.25) < 1.0 / 16)
    np.testing.assert_array_equal(cov[lv][lost[lv]], 0.0)
    np.testing.assert_allclose(cov[lv][~lost[lv]], dt, rtol=1e-14)
    # so the interface leaves feel only the pressure of their other side and
    # start to move; the middle of the slow strip, out of reach of one
    # substep's stencil, stays at rest
    far = lv[np.abs(xs[lv]) < 0.1]
    assert np.abs(a["p"][far, 1]).max() == 0.0
    assert np.abs(a["p"][lv[lost[lv].any(1)], 1]).min() > 0.0
    o.close()


# ------------------------------------------------------- neighbours, 2:1
def _expected_neighbors(a, case, y):
    D = case.dim
    fl = a["flags"]
    live = np.nonzero((fl & 4) == 0)[0]
    live = live[live >= 1]
    cs, ss, ls = a["center"][live], a["size"][live], a["tree_level"][live]
    out = []
    for f in range(2 * D):
        d, dr = f >> 1, f & 1
        pt = a["center"][y].copy()
        pt[d] += a["size"][y] * (2 * dr - 1)
        h = 0.5 * case.active_size[d]
        if pt[d] < -h or pt[d] > h:
            if not case.periodic[d]:
                out.append(0)
                continue
            pt[d] += -case.active_size[d] if pt[d] > h else case.active_size[d]
        inside = np.all(np.abs(cs - pt) < 0.5 * ss[:, None], axis=1) & (ls <= a["tree_level"][y])
        cand = live[inside]
        out.append(int(cand[np.argmax(a["tree_level"][cand])]) if cand.size else 0)
    return out


def test_neighbors_and_two_to_one_brute_force(oracle_lib):
    """After every global step of an adaptive run, each face neighbour of a
    leaf is the deepest live node at most at the leaf's tree level that
    contains the point one cell width across the face (brute force over all
    nodes), and a leaf with a face neighbour more than one level finer is
    flagged should_split; tree levels of face neighbours never differ by
    more than two."""
    case = amr_sod_case(base=3, top=6)
    o = Oracle(case, scene.build_tree(case))
    for step in range(4):
        _, fb = o.integrate()
        assert not fb
        a = o.export()
        lv = live_leaves(a)
        assert len(np.unique(a["tree_level"][lv])) >= 3
        for y in lv:
            want = _expected_neighbors(a, case, y)
            np.testing.assert_array_equal(a["nbr"][y], want, err_msg=f"step {step} leaf {y}")
            for nb in want:
                if nb and (a["flags"][nb] & 1):
                    assert abs(int(a["tree_level"][nb]) - int(a["tree_level"][y])) <= 2
                    if a["tree_level"][nb] < a["tree_level"][y] - 1:
                        assert a["flags"][nb] & 32, (step, y, nb)
    o.close()


# ------------------------------------------------------ slopes, interp_p
def test_slopes_of_linear_field_across_levels(oracle_lib):
    """A linear primitive field sampled at the leaf centres of a mesh with a
    refined strip: every interior slope is the exact gradient, also where
    the neighbour is a refined subtree (its interpolated value is the mean
    of its children, the field at its centre) or a coarser leaf (spacing
    (size + neighbour size) / 2); the zeroGradient domain faces give zero
    slopes in the boundary leaves (minmod of a zero one-sided slope)."""
    toml = A12_TOML.replace('type = "periodic"\n[boundary_conditions.Y]', 'type = "zeroGradient"\n'
                            '[boundary_conditions.Y]').replace(
        '[boundary_conditions.X]\ntype = "zeroGradient"', '[boundary_conditions.X.neg]\ntype = "zeroGradient"\n'
        '[boundary_conditions.X.pos]\ntype = "zeroGradient"')
    case = loads_case(toml, "euler", 2)
    tree = scene.build_tree(case)
    lf = np.nonzero((tree.flags & 3) == 3)[0]
    # linear along x only: the coarse/fine faces are normal to x (a coarse
    # neighbour's value is taken at its own centre, off in y)
    gx, gy = 0.25, 0.0
    x = tree.center[lf]
    tree.p[lf, 0] = 1.0 + gx * x[:, 0] + gy * x[:, 1]
    tree.p[lf, 1] = 0.0
    tree.p[lf, 2] = 0.0
    tree.p[lf, 3] = 1.0
    o = Oracle(case, tree)
    a = o.export()
    sl = a["slope"]
    inner = lf[np.abs(a["center"][lf, 0]) < 0.5 - a["size"][lf]]
    assert np.unique(a["tree_level"][inner]).size == 2
    np.testing.assert_allclose(sl[inner, 0, 0], gx, rtol=1e-12)
    np.testing.assert_array_equal(sl[lf, 1, 0], 0.0)  # periodic along y
    edge = lf[np.abs(a["center"][lf, 0]) > 0.5 - a["size"][lf]]
    np.testing.assert_array_equal(sl[edge, 0, 0], 0.0)
    np.testing.assert_array_equal(sl[lf, :, 1:], 0.0)
    o.close()


# --------------------------------------------- MUSCL-Hancock convergence
HANCOCK_TOML = """
[scene]
name = "wave"
domain_size = [1.0, 1.0]
[run]
print_interval = 1.0
end_time = 1.0
[integrator]
min_refinement_level = {lev}
max_refinement_level = {lev}
error_threshold = [-1.0, 1e300]
max_cfl = 0.5
max_dt = {dt}
[model.reference_values]
rho = 1.0
p = 1.0
vel = [1.0, 0.0]
[initial_condition]
type = "uniform"
[boundary_conditions.X]
type = "periodic"
[boundary_conditions.Y]
type = "periodic"
"""


def _wave_error(lev, T=0.25):
    h = 1.0 / (1 << lev)
    dt = h / 8
    case = loads_case(HANCOCK_TOML.format(lev=lev, dt=repr(dt)), "euler", 2, limiter="unlimited")
    tree = scene.build_tree(case)
    lf = np.nonzero((tree.flags & 3) == 3)[0]
    tree.p[lf, 0] = 1.0 + 0.2 * np.sin(2 * np.pi * tree.center[lf, 0])
    o = Oracle(case, tree)
    assert o.global_dt == dt
    steps = int(round(T / dt))
    assert o.run(steps) == steps
    a = o.export()
    lv = live_leaves(a)
    exact = 1.0 + 0.2 * np.sin(2 * np.pi * (a["center"][lv, 0] - T))
    o.close()
    return np.abs(a["p"][lv, 0] - exact).mean()


def test_muscl_hancock_second_order_on_smooth_wave(oracle_lib):
    """A density wave advected at unit speed through uniform pressure, with
    the unlimited (central) slope: the MUSCL-Hancock scheme is second order
    in space and time.  Without the half-step predictor p - dt/2 A grad p
    (or with a wrong face extrapolation) it drops to first order."""
    e = [_wave_error(lev) for lev in (4, 5, 6)]
    orders = [math.log2(e[i] / e[i + 1]) for i in range(2)]
    assert orders[-1] > 1.8, (e, orders)


# ------------------------------------------- five-equation equilibrium
ABGRALL_TOML = """
[scene]
name = "abgrall"
domain_size = [1.0, 1.0]
[run]
print_interval = 1.0
end_time = 1.0
[integrator]
min_refinement_level = 5
max_refinement_level = 5
error_threshold = [-1.0, 1e300]
max_cfl = 0.5
[model.reference_values]
rho = [1.0, 0.2]
vel = [0.5, 0.0]
p = 1.0
alpha = 1.0
{extra}
[model.fluid_1]
gamma = 1.4
P0 = {p1}
[model.fluid_2]
gamma = 3.0
P0 = 0.0
[initial_condition]
type = "uniform"
[boundary_conditions.X]
type = "periodic"
[boundary_conditions.Y]
type = "periodic"
"""


def test_five_equation_pressure_velocity_equilibrium(oracle_lib):
    """Abgrall's test: a material interface (alpha 1 -> 0, density ratio 5,
    different stiffened-gas laws) advected through uniform pressure and
    velocity keeps both uniform, to round-off, across the interface
    (partial densities, mixture EOS, the alpha equation's div u source and
    the interface compression are consistent)."""
    for p1 in ("0.0", "0.5"):
        case = loads_case(ABGRALL_TOML.format(extra="", p1=p1), "multi", 2)
        tree = scene.build_tree(case)
        lf = np.nonzero((tree.flags & 3) == 3)[0]
        a = (np.abs(tree.center[lf, 0]) < 0.25).astype(float)
        tree.p[lf, 0] = 1.0 * a
        tree.p[lf, 1] = 0.2 * (1 - a)
        tree.p[lf, 5] = a
        o = Oracle(case, tree)
        assert o.run(12) == 12
        s = o.export()
        lv = live_leaves(s)
        al = s["p"][lv, 5]
        assert ((al > 0.01) & (al < 0.99)).sum() > 0  # the interface moved and diffused
        np.testing.assert_allclose(s["p"][lv, 4], 1.0, rtol=1e-12, err_msg=f"P0 {p1}")
        np.testing.assert_allclose(s["p"][lv, 2], 0.5, rtol=1e-12, err_msg=f"P0 {p1}")
        np.testing.assert_allclose(s["p"][lv, 3], 0.0, atol=1e-12, err_msg=f"P0 {p1}")
        o.close()


def test_surface_tension_laplace_pressure_of_disc(oracle_lib):
    """A disc of fluid 1 (radius R) at rest in fluid 2 under surface tension
    sigma = 1/We: the pressure inside rises above the outside by about the
    Laplace jump sigma / R (2D).  Pins the sign and size of the sigma kappa
    terms of the contact speed, the Jacobian and the momentum and energy
    sources together with the curvature chain."""
    We = 10.0
    case = loads_case(ABGRALL_TOML.format(extra=f"We = {We}", p1="0.0").replace(
        "vel = [0.5, 0.0]", "vel = [0.0, 0.0]"), "multi", 2)
    tree = scene.build_tree(case)
    lf = np.nonzero((tree.flags & 3) == 3)[0]
    R = 0.25
    r = np.hypot(tree.center[lf, 0], tree.center[lf, 1])
    a = (r < R).astype(float)
    tree.p[lf, 0] = 1.0 * a
    tree.p[lf, 1] = 1.0 * (1 - a)
    tree.p[lf, 5] = a
    o = Oracle(case, tree)
    # the inviscid disc rings (acoustic waves bounce between the interface
    # and the periodic images): average the jump over several periods
    jumps = []
    for _ in range(20):
        assert o.run(20) == 20
        s = o.export()
        lv = live_leaves(s)
        rr = np.hypot(s["center"][lv, 0], s["center"][lv, 1])
        jumps.append(s["p"][lv, 4][rr < 0.5 * R].mean() - s["p"][lv, 4][rr > 1.6 * R].mean())
    lap = (1.0 / We) / R
    assert 0.75 * lap < np.mean(jumps) < 1.5 * lap, (jumps, lap)
    o.close()


# --------------------------------------------------------------- gc
def test_gc_compacts_and_rebins(oracle_lib):
    """Compaction (domain.hpp:327-356) removes the garbage nodes and keeps
    every other node, its state and its neighbour/child links under the new
    numbering; the ACTI levels are then re-binned from the global dt as at
    construction, level = clamp(floor(log2(global_dt / max_dt / max_cfl)),
    0, 15) (impl/acti_utils.hpp:10-24,58-74)."""
    case = amr_sod_case(base=3, top=8, emin=1e10, emax=1e20)
    o = Oracle(case, scene.build_tree(case))
    o.set_no_gc(True)
    o.integrate()
    a = o.export()
    keep = np.nonzero((a["flags"] & 4) == 0)[0]
    assert keep[0] == 0 and keep.size < a["flags"].size // 2
    o2n = np.zeros(a["flags"].size, np.int64)
    o2n[keep] = np.arange(keep.size)
    o.gc_now()
    b = o.export()
    assert b["flags"].size == keep.size
    for k in ("center", "size", "t", "p", "u", "slope", "rhs", "tree_level", "flags"):
        np.testing.assert_array_equal(b[k], a[k][keep], err_msg=k)
    np.testing.assert_array_equal(b["parent"][1:], o2n[a["parent"][keep[1:]]])
    cb = a["child_base"][keep]
    np.testing.assert_array_equal(b["child_base"], np.where(cb > 0, o2n[cb], 0))
    lv = live_leaves(b)
    nb = a["nbr"][keep][lv]
    np.testing.assert_array_equal(b["nbr"][lv], o2n[nb])
    cfl = o.global_dt / b["meta"][lv, 2]
    want = np.clip(np.floor(np.log2(cfl / case.max_cfl)), 0, 15).astype(int)
    np.testing.assert_array_equal(b["acti_level"][lv], want)
    np.testing.assert_allclose(b["meta"][lv, 3], cfl / 2.0 ** want, rtol=1e-15)
    o.close()


# ------------------------------------------------------- LTS schedule
def test_schedule_synchronises_leaves_and_keeps_cfl_stable(oracle_lib):
    """After every global step of adaptive runs (splits, unsplits, LTS level
    changes), every live leaf has advanced by exactly the global time step
    and carries a stable CFL number (<= max_cfl) on its time level."""
    for case in (amr_sod_case(base=3, top=6), workloads.shock_bubble(base=4, extra=2)):
        o = Oracle(case, scene.build_tree(case))
        for step in range(1, 4):
            dt, fb = o.integrate()
            assert not fb
            a = o.export()
            lv = live_leaves(a)
            np.testing.assert_allclose(a["t"][lv], step * dt, rtol=1e-13)
            assert (a["meta"][lv, 3] <= case.max_cfl).all()
            assert (a["meta"][lv, 3] > 0).all()
        o.close()
