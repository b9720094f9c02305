"""CPU checks of the product: the C ABI library and the host-side loaders.

No compute call is made through the ABI here (there is no GPU); the
library must load and export every entry point include/dacti_gpu.h declares.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from dacti_amd import _lib, scene, workloads
from dacti_amd.case import ConfigError, loads_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="dacti_gpu.h"):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(dacti_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), "binding signatures out of sync with the header"
    from dacti_amd import mlp
    mnames = declared_functions("dacti_mlp.h")
    assert len(mnames) >= 8
    for n in mnames:
        assert hasattr(lib, n), n
    assert set(mnames) == set(mlp.SIGNATURES), "MLP binding signatures out of sync with the header"


def test_struct_layouts_match_header():
    # dacti_case: 4 ints, 1 double, 3 doubles, 3 ints, 1 int, 4 doubles, 8 doubles, 5 + 13 doubles,
    # 6 ints, 48 doubles, 6 doubles, 24 + 24 doubles, 8 ints
    assert C.sizeof(_lib.CaseStruct) == 16 + 8 + 24 + 12 + 4 + 32 + 64 + 144 + 24 + 384 + 48 + 384 + 32
    # dacti_tree: n_nodes, 8 array pointers, brec, n_records, records
    assert C.sizeof(_lib.TreeStruct) == 8 + 8 * 8 + 8 + 8 + 8
    # dacti_boundary: type, pad, normal[3], p[8], fade_in
    assert _lib.RECORD_DTYPE.itemsize == 8 + 24 + 64 + 8
    assert C.sizeof(_lib.NodeViewStruct) == 17 * 8


def test_model_sizes():
    assert _lib.model_sizes(0, 3) == (3, 5, 5, 0, 0)
    assert _lib.model_sizes(1, 2) == (2, 4, 4, 0, 3)
    assert _lib.model_sizes(2, 2) == (2, 6, 9, 6, 5)
    assert _lib.model_sizes(2, 3) == (3, 7, 11, 7, 6)


def test_uniform_tree_is_morton_ordered():
    case = workloads.periodic_box(level=3)
    t = scene.build_tree(case)
    assert t.n == 1 + 1 + 8 + 64 + 512
    leaves = t.leaves()
    assert len(leaves) == 512 and (np.diff(leaves) == 1).all()
    h = 1.0 / 8
    ijk = np.rint((t.center[leaves] + 0.5) / h - 0.5).astype(int)
    morton = np.zeros(len(leaves), int)
    for b in range(3):
        for d in range(3):
            morton |= ((ijk[:, d] >> b) & 1) << (3 * b + d)
    assert (morton == np.arange(512)).all()
    assert (t.bc == 0).all()  # fully periodic


def test_shock_bubble_tree():
    case = workloads.shock_bubble(base=4, extra=2)
    t = scene.build_tree(case)
    lv = t.leaves()
    lev = t.level[lv]
    assert lev.min() == 4 and lev.max() == 6
    # active region is [-1,1] x [-0.5,0.5]
    assert np.all(np.abs(t.center[lv, 1]) < 0.5) and np.all(np.abs(t.center[lv, 0]) < 1.0)
    # boundaries on the domain faces only, each carrying its face's condition
    for f in range(4):
        dim, dr = f // 2, f % 2
        has = t.bc[lv, f] != 0
        edge = np.abs(t.center[lv, dim] + (2 * dr - 1) * t.size[lv] - (2 * dr - 1) * case.active_size[dim] / 2) < 1e-12
        edge |= (2 * dr - 1) * (t.center[lv, dim] + (2 * dr - 1) * t.size[lv]) > case.active_size[dim] / 2
        assert (has == edge).all()
        assert (t.bc[lv, f][has] == 1 + f).all()
    # finest cells near the bubble interface and the shock
    fine = lv[lev == 6]
    r = np.hypot(t.center[fine, 0] + 0.3, t.center[fine, 1])
    near = (np.abs(r - 0.25) < 0.2) | (np.abs(t.center[fine, 0] + 0.7) < 0.2) | (t.center[fine, 0] < -0.9)
    assert near.all()
    # alpha = 0 inside the bubble, 1 outside; partial densities from rho_ref
    rr = np.hypot(t.center[lv, 0] + 0.3, t.center[lv, 1])
    a = t.p[lv, 5]
    assert ((rr < 0.25) == (a == 0.0)).all()
    np.testing.assert_allclose(t.p[lv, 0], 1.0 * a)
    np.testing.assert_allclose(t.p[lv, 1], 0.138 * (1 - a))


TOML_BASE = """
[scene]
name = "t"
domain_size = [1.0, 1.0]
[run]
print_interval = 1.0
end_time = 1.0
[integrator]
min_refinement_level = 3
max_refinement_level = 3
error_threshold = [0.1, 0.2]
max_cfl = {cfl}
[model.reference_values]
rho = 1.0
p = 1.0
vel = [0.0, 0.0]
[initial_condition]
type = "uniform"
value = "fixed"
fixed_value = [2.0, 0.0, 0.0, 2.0]
[boundary_conditions.X]
type = "periodic"
[boundary_conditions.Y.neg]
type = "inflow_r"
value = [1.0, 0.0, 0.5, 1.0]
fade_in = 0.5
[boundary_conditions.Y.pos]
type = "outflow"
"""


def test_toml_strict_types_and_quirks():
    case = loads_case(TOML_BASE.format(cfl="0.5"), "euler", 2)
    assert case.max_dt == case.print_interval  # integrator.max_dt defaults to run.print_interval
    assert case.periodic == (True, False, False)
    assert case.bcs[2].type == 4 and case.bcs[2].value == (1.0, 0.0, 0.5, 1.0) and case.bcs[2].fade_in == 0.5
    # uniform + fixed returns the reference state (config.cpp:208-211)
    t = scene.build_tree(case)
    np.testing.assert_array_equal(t.p[t.leaves()], [[1.0, 0.0, 0.0, 1.0]] * len(t.leaves()))
    with pytest.raises(ConfigError):
        loads_case(TOML_BASE.format(cfl="1"), "euler", 2)  # integer where a float is required
    with pytest.raises(ConfigError):
        loads_case(TOML_BASE.format(cfl="0.5").replace("vel = [0.0, 0.0]", "vel = [0.0, 0.0, 0.0]"), "euler", 2)
    with pytest.raises(ConfigError):
        loads_case(TOML_BASE.format(cfl="0.5").replace('type = "outflow"', 'type = "wall"'), "euler", 2)


def test_multi_reference_values():
    case = workloads.shock_bubble(base=3, extra=1)
    assert case.consts["gamma1"] == 1.4 and case.consts["gamma2"] == 1.67
    assert case.consts["cv1"] == 4181.0 / 6.12  # fixed at static initialisation (multi.hpp:57)
    assert case.p_ref == (1.0, 0.0, 0.0, 0.0, 1.0, 1.0)


def test_loading_requires_built_library(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError):
        _lib.load()


CPP_PROBE = r"""
#include "dacti/integrator/integrator.hpp"
#include <cstdio>
using namespace dacti;
static const char *CASE = R"(
[scene]
name = "probe"
domain_size = [1.0, 1.0]
[run]
print_interval = 1.0
end_time = 1.0
[integrator]
min_refinement_level = 3
max_refinement_level = 4
error_threshold = [0.01, 0.1]
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
)";
int probe_integrator(const char *out_root) {
    // the reference's construction by class-template argument deduction;
    // without a GPU the handle cannot be created and this throws
    auto config = config::Config<model::euler<2>>::from_string(CASE, out_root);
    scene::Scene<model::euler<2>> scene(config);
    try {
        integrator::Integrator integ{integrator::slope::minmod{},
                                     integrator::error::weighted_laplacian{config->error_weights}, scene, config};
        std::printf("created %zu\n", integ.n_active_cells());
    } catch (const dacti::Error &e) {
        std::printf("create failed %d\n", e.code);
    }
    return (int) scene.init_tree().n;
}
int main(int, char **argv) {
    std::printf("tree %d\n", probe_integrator(argv[1]));
    auto s = dacti::model_sizes(DACTI_MODEL_MULTI, 3);
    std::printf("%d %d %d %d %d\n", s.dim, s.num_eqns, s.num_flux, s.num_parm, s.num_grad);
    try {
        dacti::model_sizes(7, 3);
    } catch (const dacti::Error &e) {
        std::printf("error %d\n", e.code);
    }
    return 0;
}
"""


def test_cpp_facade_compiles_and_links(tmp_path):
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "probe.cpp"
    src.write_text(CPP_PROBE)
    exe = tmp_path / "probe"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                    str(src), "-L", libdir, "-ldacti_gpu", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), str(tmp_path)], check=True, capture_output=True, text=True,
                         timeout=120).stdout.split("\n")
    assert out[0] == "create failed -2" and out[1] == f"tree {2 + 4 + 16 + 64}"  # null node, root, levels 1-3
    assert out[2] == "3 7 11 7 6"
    assert out[3] == "error -1"
