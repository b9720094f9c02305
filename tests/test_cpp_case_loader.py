"""The C++ TOML case loader and initial-tree builder (include/dacti/config)
against the Python ones (dacti_amd/case.py, dacti_amd/scene.py), which the
GPU parity tests use: the same case structure field by field and the same
initial tree bit for bit, over every case text the tests and workloads
load, plus the typing errors of config/util.hpp:15-103."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from dacti_amd import _lib, scene, workloads
from dacti_amd.case import ConfigError, loads_case
from tests import helpers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELS = {0: "euler", 1: "navier_stokes", 2: "multi"}


@pytest.fixture(scope="module")
def dump_tool(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = tmp_path_factory.mktemp("cpp") / "case_tree_dump"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "case_tree_dump.cpp"), "-o", str(exe)], check=True)
    return exe


def _captured_texts(monkeypatch, fn, *args, **kw):
    """The (text, model, dim, limiter) a workload builder passes to loads_case."""
    seen = []
    orig = workloads.loads_case

    def cap(text, model, dim, limiter="minmod"):
        seen.append((text, model, dim, limiter))
        return orig(text, model, dim, limiter)

    monkeypatch.setattr(workloads, "loads_case", cap)
    case = fn(*args, **kw)
    monkeypatch.setattr(workloads, "loads_case", orig)
    return seen[-1], case


def _run(dump_tool, tmp_path, text, model, dim, limiter, grid=None):
    src = tmp_path / "case.toml"
    src.write_text(text)
    out = tmp_path / "out.bin"
    cmd = [str(dump_tool), str(src), model, str(dim), limiter, str(out)]
    if grid is not None:
        gpath = tmp_path / "grid.bin"
        np.ascontiguousarray(grid, np.float64).tofile(gpath)
        shape = list(grid.shape[:dim]) + [0] * (3 - dim)
        cmd += [str(gpath)] + [str(s) for s in shape]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=60)
    return r, out


def _compare(dump_tool, tmp_path, text, model, dim, limiter="minmod", case=None):
    py = case if case is not None else loads_case(text, model, dim, limiter)
    grid = py.ic_grid if py.ic_type == "grid" else None
    r, out = _run(dump_tool, tmp_path, text, model, dim, limiter, grid)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes()
    ncs = C.sizeof(_lib.CaseStruct)
    cc = _lib.CaseStruct.from_buffer_copy(raw[:ncs])
    pc = _lib.case_struct(py)
    for name, _ in _lib.CaseStruct._fields_:
        a, b = getattr(cc, name), getattr(pc, name)
        if hasattr(a, "__len__"):
            a = np.array([list(x) if hasattr(x, "__len__") else x for x in a])
            b = np.array([list(x) if hasattr(x, "__len__") else x for x in b])
            np.testing.assert_array_equal(a, b, err_msg=name)
        else:
            assert a == b, name
    tree = scene.build_tree(py)
    n = int(np.frombuffer(raw[ncs:ncs + 8], np.int64)[0])
    assert n == tree.n
    off = ncs + 8
    ne = py.num_eqns
    for attr, dt, width in (("parent", np.int32, 1), ("child_base", np.int32, 1), ("level", np.int8, 1),
                            ("flags", np.uint8, 1), ("bc", np.uint8, 2 * dim), ("center", np.float64, dim),
                            ("size", np.float64, 1), ("p", np.float64, ne)):
        cnt = n * width
        got = np.frombuffer(raw[off:off + cnt * np.dtype(dt).itemsize], dt)
        off += cnt * np.dtype(dt).itemsize
        want = np.ascontiguousarray(getattr(tree, attr), dt).reshape(-1)
        np.testing.assert_array_equal(got, want, err_msg=attr)
    assert off == len(raw)
    return tree


def test_helpers_cases(dump_tool, tmp_path):
    _compare(dump_tool, tmp_path, helpers.SOD_TOML.format(h=repr(1.0), lev=5, maxdt=repr(1e9)), "euler", 2)
    for model in ("euler", "navier_stokes", "multi"):
        if model == "multi":
            continue
        extra = "Re = 200.0\nL = 1.0\nPr = 0.72\n" if model == "navier_stokes" else ""
        text = helpers.AMR_TOML.format(base=3, top=6, emin="0.005", emax="0.02", extra=extra, loc="-0.1",
                                       bcx="zeroGradient")
        if model == "navier_stokes":
            text = text.replace("vel = [0.0, 0.0]", "vel = [0.3, 0.0]")
        t = _compare(dump_tool, tmp_path, text, model, 2)
        assert (t.level > 3).any()  # the jump refines the tree


@pytest.mark.parametrize("limiter", ["minmod", "van_leer", "koren", "first_order"])
def test_limiter_ids(dump_tool, tmp_path, limiter):
    _compare(dump_tool, tmp_path, helpers.SOD_TOML.format(h=repr(1.0), lev=3, maxdt=repr(1.0)), "euler", 2, limiter)


def test_workload_cases(dump_tool, tmp_path, monkeypatch):
    (text, model, dim, lim), _ = _captured_texts(monkeypatch, workloads.shock_bubble, base=4, extra=3)
    t = _compare(dump_tool, tmp_path, text, model, dim, lim)
    assert t.n > 1000
    (text, model, dim, lim), _ = _captured_texts(monkeypatch, workloads.shock_drop, base=3, extra=2)
    _compare(dump_tool, tmp_path, text, model, dim, lim)
    seen = []
    orig = workloads.loads_case
    monkeypatch.setattr(workloads, "loads_case", lambda *a: seen.append(a) or orig(*a))
    workloads.shock_bubble_family(3, seed=5, base=4, extra_range=(1, 3))  # seeded Mach, radius, density ratio
    monkeypatch.setattr(workloads, "loads_case", orig)
    assert len(seen) == 3
    for text, model, dim in (a[:3] for a in seen):
        _compare(dump_tool, tmp_path, text, model, dim)
    (text, model, dim, lim), case = _captured_texts(monkeypatch, workloads.periodic_box, level=3)
    _compare(dump_tool, tmp_path, text, model, dim, lim, case=case)
    (text, model, dim, lim), case = _captured_texts(monkeypatch, workloads.periodic_box, level=4, dim=2)
    _compare(dump_tool, tmp_path, text, model, dim, lim, case=case)


def test_pin_cases(dump_tool, tmp_path):
    from tests.test_gpu_parity_paths import BC_TOML
    from tests.test_oracle_phase_pins import A12_TOML, ABGRALL_TOML, HANCOCK_TOML
    from tests.test_oracle_pins import LAPLACE_TOML
    _compare(dump_tool, tmp_path, BC_TOML, "euler", 2)
    _compare(dump_tool, tmp_path, A12_TOML, "euler", 2)
    _compare(dump_tool, tmp_path, HANCOCK_TOML.format(lev=4, dt=repr(0.001)), "euler", 2, "unlimited")
    _compare(dump_tool, tmp_path, ABGRALL_TOML.format(extra="We = 10.0", p1="0.5"), "multi", 2)
    case = loads_case(LAPLACE_TOML, "euler", 2)
    case.ic_grid = np.random.default_rng(3).uniform(0.5, 1.5, (16, 16, 4))  # one grid cell per leaf
    _compare(dump_tool, tmp_path, LAPLACE_TOML, "euler", 2, case=case)
    case.ic_grid = np.random.default_rng(4).uniform(0.5, 1.5, (5, 7, 4))  # nearest of a coarser grid
    _compare(dump_tool, tmp_path, LAPLACE_TOML, "euler", 2, case=case)


def _expect_error(dump_tool, tmp_path, text, model="euler", dim=2):
    with pytest.raises((ConfigError, ValueError, KeyError)):
        loads_case(text, model, dim)
    r, _ = _run(dump_tool, tmp_path, text, model, dim, "minmod")
    assert r.returncode == 3 and r.stdout.startswith("ConfigError"), r.stdout + r.stderr


def test_typing_errors(dump_tool, tmp_path):
    base = helpers.SOD_TOML.format(h=repr(1.0), lev=3, maxdt=repr(1.0))
    _expect_error(dump_tool, tmp_path, base.replace("max_cfl = 0.5", "max_cfl = 1"))          # int for float
    _expect_error(dump_tool, tmp_path, base.replace("vel = [0.0, 0.0]", "vel = [0.0, 0.0, 0.0]"))  # length
    _expect_error(dump_tool, tmp_path, base.replace("vel = [0.0, 0.0]", "vel = [0, 0.0]"))    # int element
    _expect_error(dump_tool, tmp_path, base.replace('type = "zeroGradient"', 'type = "wall"', 1))  # bc type
    _expect_error(dump_tool, tmp_path, base.replace('type = "jump"', 'type = "ramp"'))         # ic type
    _expect_error(dump_tool, tmp_path, base.replace("min_refinement_level = 3", "min_refinement_level = 3.0"))
    _expect_error(dump_tool, tmp_path, base.replace("[run]\n", "[run]\nend_time = 2.0\n"))     # duplicate key
    _expect_error(dump_tool, tmp_path, base.replace("end_time = 1.0\n", ""))                   # missing key


def test_toml_syntax(dump_tool, tmp_path):
    """Comments, dotted and quoted keys, inline tables, multi-line arrays,
    exponents and underscores parse as TOML defines them."""
    text = """
# a comment
[scene]
"name" = 'sod'   # literal string
domain_size = [
    1.0,  # x
    1.0,
]
[run]
print_interval = 1_0.0
end_time = 1e0
[integrator]
min_refinement_level = 3
max_refinement_level = 3
error_threshold = [-1.0, +1e300]
max_cfl = 5E-1
model.reference_values = { rho = 1.0, p = 1.0, vel = [0.0, 0.0] }
[initial_condition]
type = "uniform"
[boundary_conditions]
X.type = "periodic"
Y = { type = "periodic" }
"""
    # model.reference_values sits under [integrator] here, so it must be
    # missing at the top level: the loader reports it
    _expect_error(dump_tool, tmp_path, text)
    fixed = text.replace("model.reference_values = { rho = 1.0, p = 1.0, vel = [0.0, 0.0] }\n", "") + \
        "[model]\nreference_values = { rho = 1.0, p = 1.0, vel = [0.0, 0.0] }\n"
    _compare(dump_tool, tmp_path, fixed, "euler", 2)
    c = loads_case(fixed, "euler", 2)
    assert c.print_interval == 10.0 and c.max_cfl == 0.5 and c.periodic[:2] == (True, True)
