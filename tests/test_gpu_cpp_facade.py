"""The untyped C++ owner of a GPU handle (include/dacti/integrator/gpu_solver.hpp) driving a case on the GPU.

A small C++ driver, compiled here against the header and libdacti_gpu.so,
reads a case and its initial tree written by Python, runs two global steps
through dacti::GpuSolver and writes the node pool's conserved state; it must
equal the oracle's bit for bit (the same path as the Python facade)."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from dacti_amd import _lib, scene, workloads
from oracle.binding import Oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include "dacti/integrator/gpu_solver.hpp"
#include <cstdio>
#include <vector>
template <typename T> std::vector<T> rd(FILE *f, size_t n) { std::vector<T> v(n); if (fread(v.data(), sizeof(T), n, f) != n) throw 1; return v; }
int main(int argc, char **argv) {
    FILE *f = std::fopen(argv[1], "rb");
    dacti_case c;
    if (fread(&c, sizeof c, 1, f) != 1) return 2;
    int64_t n; if (fread(&n, 8, 1, f) != 1) return 2;
    auto sz = dacti::model_sizes(c.model, c.dim);
    auto parent = rd<int32_t>(f, n), cb = rd<int32_t>(f, n);
    auto lev = rd<int8_t>(f, n);
    auto fl = rd<uint8_t>(f, n), bc = rd<uint8_t>(f, n * 2 * c.dim);
    auto ctr = rd<double>(f, n * c.dim), size = rd<double>(f, n), p = rd<double>(f, n * sz.num_eqns);
    auto brec = rd<int32_t>(f, n);
    int64_t nrec; if (fread(&nrec, 8, 1, f) != 1) return 2;
    auto recs = rd<dacti_boundary>(f, nrec);
    std::fclose(f);
    dacti_tree t{n, parent.data(), cb.data(), lev.data(), fl.data(), bc.data(), ctr.data(), size.data(), p.data(),
                 brec.data(), nrec, recs.data()};
    dacti::GpuSolver g(c, t);
    bool brk = false;
    for (int s = 0; s < 2; ++s) g.integrate(brk);
    dacti::NodePool P = g.get_nodes();
    FILE *o = std::fopen(argv[2], "wb");
    std::fwrite(&P.n, 8, 1, o);
    std::fwrite(P.u.data(), sizeof(double), P.u.size(), o);
    std::fwrite(P.flags.data(), 1, P.flags.size(), o);
    std::fclose(o);
    std::printf("cpp driver: %lld nodes, %lld active\n", (long long) P.n, (long long) g.n_active_cells());
    return brk ? 3 : 0;
}
"""


def test_cpp_facade_matches_oracle(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    case = workloads.shock_bubble(base=4, extra=2)
    tree = scene.build_tree(case)
    cs = _lib.case_struct(case)
    inp = tmp_path / "case.bin"
    with open(inp, "wb") as f:
        f.write(bytes(cs))
        f.write(np.int64(tree.n).tobytes())
        for a, dt in ((tree.parent, np.int32), (tree.child_base, np.int32), (tree.level, np.int8),
                      (tree.flags, np.uint8), (tree.bc, np.uint8), (tree.center, np.float64),
                      (tree.size, np.float64), (tree.p, np.float64), (tree.brec, np.int32)):
            f.write(np.ascontiguousarray(a, dtype=dt).tobytes())
        f.write(np.int64(len(tree.records)).tobytes())
        f.write(np.ascontiguousarray(tree.records).tobytes())
    src = tmp_path / "driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "driver"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir,
                    "-ldacti_gpu", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = tmp_path / "out.bin"
    print("[cpp] running the C++ driver", flush=True)
    r = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    o = Oracle(case, tree)
    o.integrate()
    o.integrate()
    e = o.export()
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:8], np.int64)[0])
    ne = e["u"].shape[1]
    u = np.frombuffer(raw[8:8 + 8 * n * ne], np.float64).reshape(n, ne)
    flags = np.frombuffer(raw[8 + 8 * n * ne:], np.uint8)
    assert n == len(e["flags"])
    np.testing.assert_array_equal(flags, e["flags"])
    lv = np.nonzero((e["flags"] & 3) == 3)[0]
    np.testing.assert_array_equal(u[lv], e["u"][lv])
