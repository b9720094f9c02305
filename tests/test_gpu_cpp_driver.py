"""The reference-shaped C++ interface on the GPU (tests/cpp/run_case.cpp):
config::Config + scene::Scene + integrator::Integrator<Model, Limiter,
ErrorEstimator> built by class-template argument deduction, driven by
driver::default_driver with io::vtu_writer and observer::DropEnergy, and
integrate(hook, signal, force_break) with a per-substep hook.  The frames'
cell data, the observables CSV and the hook schedule are checked against
the oracle run of the same case (same TOML, same initial tree)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from dacti_amd import _lib, scene, workloads
from dacti_amd.case import loads_case
from oracle.binding import Oracle
from tests.test_cpp_driver import read_vtu

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAX_LEVEL = 16


@pytest.fixture(scope="module")
def run_case_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = tmp_path_factory.mktemp("cpp") / "run_case"
    inc = os.path.join(ROOT, "include")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(inc, "dacti"), "-I", inc,
                    os.path.join(ROOT, "tests", "cpp", "run_case.cpp"), "-L", libdir, "-ldacti_gpu",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    return exe


def bubble_text(end_time, print_interval):
    """The shock-bubble case at 16 x 8 + 2 levels with surface tension
    (We 20), as TOML text."""
    case = workloads.shock_bubble(base=4, extra=2)
    text = workloads.SHOCK_BUBBLE_TOML.format(
        base=4, top=6, emin="0.02", emax="0.05", rho2="0.138", shock_x="-0.7", r2=repr(case.bcs[0].value[0]),
        u2=repr(case.bcs[0].value[2]), p2=repr(case.bcs[0].value[4]), radius="0.25", bx="-0.3")
    text = text.replace("alpha = 1.0\n", "alpha = 1.0\nWe = 20.0\n", 1)
    text = text.replace("max_cfl = 0.5\n", "max_cfl = 0.5\nmax_dt = 1.0\n", 1)  # dt independent of the frames
    text = text.replace("print_interval = 1.0", f"print_interval = {print_interval!r}")
    return text.replace("end_time = 1.0", f"end_time = {end_time!r}")


def order_active(e):
    """Node indices in for_each_active_cell order: cluster keys 0..31, node order inside."""
    return np.concatenate([np.nonzero(e["cluster"] == k)[0] for k in range(2 * MAX_LEVEL)])


def energies(e, sigma):
    idx = order_active(e)
    ke = et = 0.0
    for i in idx:
        vol = e["size"][i] ** 2
        p = e["p"][i]
        kin = 0.5 * (p[0] + p[1]) * (p[2] ** 2 + p[3] ** 2)
        g = e["grad"][i]
        ke += vol * kin
        et += vol * (kin + sigma * np.sqrt(g[0, 3] ** 2 + g[1, 3] ** 2))
    return ke, et


def test_default_driver_frames_and_observables(run_case_exe, tmp_path):
    probe = loads_case(bubble_text(1.0, 1.0), "multi", 2)
    o = Oracle(probe, scene.build_tree(probe))
    dt, _ = o.integrate()
    del o
    steps = 3
    text = bubble_text(dt * (steps - 0.5), dt * 0.5)  # a frame after every step, three steps
    (tmp_path / "cases").mkdir()
    path = tmp_path / "cases" / "bubble.toml"
    path.write_text(text)
    print("[cpp-driver] shock-bubble 16x8+2 We 20, default_driver, 3 steps", flush=True)
    r = subprocess.run([str(run_case_exe), str(path), "multi2", "driver"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    out = tmp_path / "out" / loads_case(text, "multi", 2).name
    csv = (out / f"{out.name}_observables.csv").read_text().splitlines()
    assert csv[0] == "time,frame,step,n_cells,ke,Etot" and len(csv) == 1 + steps
    case = loads_case(text, "multi", 2)
    sigma = case.consts["sigma"]
    assert sigma == 1.0 / 20.0
    o = Oracle(case, scene.build_tree(case))
    for s in range(steps):
        o.integrate()
        e = o.export()
        row = [float(x) for x in csv[1 + s].split(",")]
        assert row[1:4] == [s + 1, s, len(order_active(e))]
        np.testing.assert_allclose(row[0], dt * (s + 1), rtol=1e-5)
        np.testing.assert_allclose(row[4:], energies(e, sigma), rtol=1e-5)
        arrays, npts, ncells = read_vtu(out / f"{out.name}_{s + 1:04d}.vtu")
        idx = order_active(e)
        assert ncells == len(idx) and npts == 8 * len(idx)
        # cell data: parameters and metadata, bit for bit
        for q in range(6):
            np.testing.assert_array_equal(arrays[f"k_{q}"].ravel(), e["k"][idx, q])
        for k, name in enumerate(["acti_level", "error", "max_dt", "cfl"]):
            np.testing.assert_array_equal(arrays[name].ravel(), e["meta"][idx, k])
        # point data: the linear reconstruction at the corners
        pts = arrays["Points"].reshape(-1, 8, 3)
        c = e["center"][idx]
        pc = e["p"][idx][:, None, :] + e["slope"][idx][:, None, 0, :] * (pts[:, :, 0:1] - c[:, None, 0:1]) + \
            e["slope"][idx][:, None, 1, :] * (pts[:, :, 1:2] - c[:, None, 1:2])
        got = np.stack([arrays[f"p_{q}"].ravel() for q in range(6)], axis=1).reshape(-1, 8, 6)
        np.testing.assert_allclose(got, pc, rtol=1e-14, atol=1e-300)
        half = 0.5 * e["size"][idx]
        np.testing.assert_array_equal(np.abs(pts[:, :, 0] - c[:, None, 0]), np.repeat(half[:, None], 8, 1))
    assert sorted(f for f in os.listdir(out) if f.endswith(".vtu")) == [f"{out.name}_{f:04d}.vtu"
                                                                       for f in range(steps + 1)]


def test_integrate_hook_and_counters(run_case_exe, tmp_path):
    text = bubble_text(1.0, 1.0)
    path = tmp_path / "bubble.toml"
    path.write_text(text)
    print("[cpp-driver] shock-bubble 16x8+2, integrate(hook) 2 steps", flush=True)
    r = subprocess.run([str(run_case_exe), str(path), "multi2", "hook", "2"], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    hooks = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("hook ")]
    steps = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("step ")]
    assert len(steps) == 2 and hooks
    case = loads_case(text, "multi", 2)
    o = Oracle(case, scene.build_tree(case))
    for s in range(2):
        dt, _ = o.integrate()
        fields = steps[s]
        assert float(fields[3]) == dt
        assert int(fields[7]) == s + 1 and int(fields[9]) == o.max_acti_level
    # one hook call per substep, numbered 1..n, dt_L = global dt / 2^L, and
    # the counters of the step equal the oracle's
    assert [int(h[3]) for h in hooks] == list(range(1, len(hooks) + 1))
    assert int(steps[-1][5]) == len(hooks) == o.n_substeps
    assert int(steps[-1][11]) == o.cell_updates
    gdt = float(steps[0][3])
    for h in hooks:
        assert float(h[4]) == gdt / (1 << int(h[1]))
    assert int(steps[-1][13]) == len(order_active(o.export()))


def test_default_driver_aero_force_on_a_solid(run_case_exe, tmp_path):
    """euler<2> around a noslip cylinder through default_driver with
    AeroForce: the CSV carries Cl and Cd, the frames the Cp and
    ObjectSurface fields, nonzero exactly on the cells flagged as object
    boundaries (the oracle's is_boundary)."""
    from tests.helpers import solid_case
    probe = solid_case("cylinder", u=0.5)
    o = Oracle(probe, scene.build_tree(probe))
    dt, _ = o.integrate()
    del o
    text = probe.toml_text.replace("end_time = 1.0", f"end_time = {dt * 2.5!r}").replace(
        "print_interval = 1.0", f"print_interval = {dt * 0.5!r}").replace("max_cfl = 0.4", "max_cfl = 0.4\nmax_dt = 1.0")
    (tmp_path / "cases").mkdir()
    path = tmp_path / "cases" / "cyl.toml"
    path.write_text(text)
    print("[cpp-driver] euler2d cylinder, default_driver + AeroForce, 3 steps", flush=True)
    r = subprocess.run([str(run_case_exe), str(path), "euler2", "driver"], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    out = tmp_path / "out" / "solid"
    csv = (out / "solid_observables.csv").read_text().splitlines()
    assert csv[0] == "time,frame,step,n_cells,Cl,Cd" and len(csv) == 4
    case = loads_case(text, "euler", 2)
    o = Oracle(case, scene.build_tree(case))
    for _ in range(3):
        o.integrate()
    e = o.export()
    idx = order_active(e)
    arrays, _, ncells = read_vtu(out / "solid_0003.vtu")
    assert ncells == len(idx)
    surf = arrays["Aerodynamic Forces/ObjectSurface"].ravel()
    np.testing.assert_array_equal(surf, ((e["flags"][idx] & 64) != 0).astype(float))
    cp = arrays["Aerodynamic Forces/Cp"].ravel()
    p_ref, rho_ref, u2 = 1.0, 1.0, 0.25
    np.testing.assert_array_equal(cp[surf == 1], (e["p"][idx][surf == 1, 3] - p_ref) / (0.5 * rho_ref * u2))
    assert (cp[surf == 0] == 0).all() and surf.sum() > 10
