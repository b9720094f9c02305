"""Object geometry of the scene (include/dacti/scene/geometry.hpp) against
closed forms: the adaptive sampling of built-in curves
(scene/geometry_data.cpp:26-99), signed distances and normals of circles,
ellipses and polygons (point-segment distances and even-odd insideness
computed here), open outlines, segment intersection, and STL meshes
(normalised to unit extent, split into components) against the spheres
they approximate."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = tmp_path_factory.mktemp("cpp") / "geometry_probe"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "geometry_probe.cpp"), "-o", str(exe)], check=True)
    return Probe(exe)


class Probe:
    def __init__(self, exe):
        self.exe = exe

    def run(self, cmds):
        r = subprocess.run([str(self.exe)], input="\n".join(cmds) + "\n", capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        return r.stdout.splitlines()

    def sd(self, setup, pts):
        out = self.run(setup + [f"sd {float(x)!r} {float(y)!r} {float(z)!r}" for x, y, z in pts])[len(setup):]
        a = np.array([[float(v) for v in ln.split()] for ln in out])
        return a[:, 0], a[:, 1:4], a[:, 4].astype(int)


def seg_dist(p, a, b):
    ab = b - a
    t = np.clip(((p - a) @ ab) / (ab @ ab), 0.0, 1.0)
    return np.linalg.norm(p - (a + t[:, None] * ab), axis=1)


def poly_dist(p, verts, closed=True):
    n = len(verts)
    segs = range(n) if closed else range(n - 1)
    return np.min([seg_dist(p, verts[i], verts[(i + 1) % n]) for i in segs], axis=0)


def inside_poly(p, verts):
    inside = np.zeros(len(p), bool)
    n = len(verts)
    for i in range(n):
        a, b = verts[i], verts[i - 1]
        cross = (a[1] > p[:, 1]) != (b[1] > p[:, 1])
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = (b[0] - a[0]) * (p[:, 1] - a[1]) / (b[1] - a[1]) + a[0]
        inside ^= cross & (p[:, 0] < xint)
    return inside


def points(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(lo, hi, (n, 2))
    return p, [(x, y, 0.0) for x, y in p]


def test_adaptive_sampling_levels(probe):
    # a circle turns its tangent by the interval's length: every interval is
    # split to the finest level 13 (2 pi / 2^13 < 1e-3), 8192 samples at the
    # interval midpoints
    out = probe.run(["samples circle 0.25", "samples ellipse 0.2 0.02", "samples naca 0.02 0.4 0.12",
                     "samples joukowski 0.05 0.12 0.0"])
    n, t0, t1, *_ = [float(v) for v in out[0].split()]
    assert n == 8192 and t0 == pytest.approx(np.pi / 8192, rel=1e-15) and t1 == pytest.approx(3 * np.pi / 8192)
    # a flat ellipse turns slowly along its flanks: fewer samples there
    ne = int(out[1].split()[0])
    assert 16 <= ne < 8192
    # an open curve keeps its end points
    v = [float(x) for x in out[2].split()]
    assert v[1] == -1.0 and v[-1] == 1.0 and 16 + 2 <= v[0] <= 8192 + 2
    assert 16 <= int(out[3].split()[0]) <= 8192


def test_circle_distance_sign_normal(probe):
    r, c = 0.25, np.array([-0.3, 0.1])
    p, pts = points(3000, -0.8, 0.4, 1)
    sd, n, _ = probe.sd([f"shape circle {r} {c[0]} {c[1]} 0.0"], pts)
    exact = np.linalg.norm(p - c, axis=1) - r
    sag = r * (1 - np.cos(np.pi / 8192))  # chord deviation of the 8192-gon
    assert np.max(np.abs(sd - exact)) <= sag + 1e-15
    far = np.abs(exact) > 2 * sag
    assert (np.sign(sd[far]) == np.sign(exact[far])).all()
    # smooth outline: the blended vertex normal is the radial direction
    radial = (p - c) / np.linalg.norm(p - c, axis=1)[:, None]
    nn = n[:, :2] / np.linalg.norm(n[:, :2], axis=1)[:, None]
    assert np.min(np.sum(nn * radial, axis=1)) > 1 - 1e-6
    assert np.max(np.abs(n[:, 2])) < 1e-12


def test_ellipse_against_dense_outline(probe):
    a, b = 0.2, 0.1
    p, pts = points(400, -0.35, 0.35, 2)
    sd, _, _ = probe.sd([f"shape ellipse {a} {b} 0.0 0.0 0.0"], pts)
    t = np.linspace(0, 2 * np.pi, 200001)
    ring = np.stack([a * np.cos(t), b * np.sin(t)], axis=1)
    d = np.array([np.min(np.linalg.norm(ring - q, axis=1)) for q in p])
    inside = (p[:, 0] / a) ** 2 + (p[:, 1] / b) ** 2 < 1
    exact = np.where(inside, -d, d)
    np.testing.assert_allclose(sd, exact, atol=2e-6)


@pytest.mark.parametrize("rot", [0.0, 30.0, -75.0])
def test_diamond_exact(probe, rot):
    w, h, tx, ty = 1.0, 0.176, 0.1, -0.05
    base = np.array([[0, h / 2], [w / 2, 0], [0, -h / 2], [-w / 2, 0]])
    th = np.radians(rot)
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    verts = base @ R.T + [tx, ty]
    p, pts = points(2000, -0.7, 0.7, 3)
    sd, n, _ = probe.sd([f"shape diamond {w} {h} {tx} {ty} {rot}"], pts)
    d = poly_dist(p, verts)
    exact = np.where(inside_poly(p, verts), -d, d)
    np.testing.assert_allclose(sd, exact, atol=1e-14)
    # flat faces: the normal of the nearest edge, pointing out of the
    # (clockwise-listed) diamond
    e = np.argmin([seg_dist(p, verts[i], verts[(i + 1) % 4]) for i in range(4)], axis=0)
    ab = verts[(e + 1) % 4] - verts[e]
    out = np.stack([ab[:, 1], -ab[:, 0]], axis=1) / np.linalg.norm(ab, axis=1)[:, None]
    centroid_side = np.sum((verts[e] - [tx, ty]) * out, axis=1)
    out[centroid_side < 0] *= -1
    # points whose nearest feature is an edge's interior (a corner belongs to two faces)
    a = verts[e]
    t = np.sum((p - a) * ab, axis=1) / np.sum(ab * ab, axis=1)
    inner = (t > 1e-9) & (t < 1 - 1e-9)
    assert inner.sum() > 1000
    np.testing.assert_allclose(n[inner, :2], out[inner], atol=1e-12)


def test_polygon_orientation_and_open_line(probe):
    tri = [(-0.4, 0.3), (0.4, -0.02), (0.4, 0.3)]
    p, pts = points(500, -0.6, 0.6, 4)
    cw = " ".join(f"{x} {y}" for x, y in tri)
    ccw = " ".join(f"{x} {y}" for x, y in tri[::-1])
    s1, _, _ = probe.sd([f"shape triangle {cw} 0.0 0.0 0.0"], pts)
    s2, _, _ = probe.sd([f"shape triangle {ccw} 0.0 0.0 0.0"], pts)
    np.testing.assert_allclose(s1, s2, rtol=0, atol=1e-15)  # same distances; rounding follows the vertex order
    v = np.array(tri)
    np.testing.assert_allclose(s1, np.where(inside_poly(p, v), -1, 1) * poly_dist(p, v), atol=1e-14)
    # a line has no inside: the distance to the segment, always positive
    s3, _, _ = probe.sd(["shape line -0.3 0.1 0.4 -0.2 0.0 0.0 0.0"], pts)
    np.testing.assert_allclose(s3, seg_dist(p, np.array([-0.3, 0.1]), np.array([0.4, -0.2])), atol=1e-14)


def test_segment_intersection(probe):
    setup = ["shape quad -0.2 -0.2 0.2 -0.2 0.2 0.2 -0.2 0.2 0.0 0.0 0.0"]
    out = probe.run(setup + ["ray -0.5 0 0 0.5 0 0", "ray -0.5 0 0 -0.3 0 0", "ray -0.1 0 0 0.1 0.05 0",
                             "ray -0.5 0.3 0 0.5 0.3 0", "ray 0.1 -0.5 0 0.1 -0.1 0"])[1:]
    assert out == ["1", "0", "0", "0", "1"]


def icosphere(level):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(x, float) / np.linalg.norm(x) for x in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                x = v[a] + v[b]
                v.append(x / np.linalg.norm(x))
                mid[k] = len(v) - 1
            return mid[k]

        nf = []
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v), np.array(f)


def write_stl(path, tris, binary=True):
    if binary:
        with open(path, "wb") as fh:
            fh.write(b"\0" * 80 + struct.pack("<I", len(tris)))
            for t in tris:
                fh.write(struct.pack("<3f", 0, 0, 0) + b"".join(struct.pack("<3f", *x) for x in t) + b"\0\0")
    else:
        with open(path, "w") as fh:
            fh.write("solid s\n")
            for t in tris:
                fh.write(" facet normal 0 0 0\n  outer loop\n")
                for x in t:
                    fh.write(f"   vertex {x[0]:.9e} {x[1]:.9e} {x[2]:.9e}\n")
                fh.write("  endloop\n endfacet\n")
            fh.write("endsolid s\n")


@pytest.mark.parametrize("binary", [True, False])
def test_stl_sphere_and_components(probe, tmp_path, binary):
    v, f = icosphere(4)
    one = [v[t] for t in f]
    path = tmp_path / "sphere.stl"
    write_stl(path, one, binary)
    # unit extent after normalisation (radius 0.5), scaled by 0.4: radius 0.2
    rng = np.random.default_rng(5)
    p = rng.uniform(-0.4, 0.4, (500, 3)) + [0.1, 0, 0]
    sd, n, _ = probe.sd([f"stl {path} 0.4 0.4 0.4 0.1 0.0 0.0"], [tuple(x) for x in p])
    exact = np.linalg.norm(p - [0.1, 0, 0], axis=1) - 0.2
    assert np.max(np.abs(sd - exact)) < 2e-3  # chord deviation of the level-4 icosphere
    far = np.abs(exact) > 4e-3
    assert (np.sign(sd[far]) == np.sign(exact[far])).all()
    radial = (p - [0.1, 0, 0]) / np.linalg.norm(p - [0.1, 0, 0], axis=1)[:, None]
    nn = n / np.linalg.norm(n, axis=1)[:, None]
    assert np.min(np.sum(nn * radial, axis=1)) > 0.99
    # two spheres in one file: two objects, each the nearest on its side
    two = one + [t + [3.0, 0, 0] for t in one]
    path2 = tmp_path / "two.stl"
    write_stl(path2, two, binary)
    out = probe.run([f"stl {path2} 1 1 1 0 0 0", "count"])
    assert out[1].split()[0] == "2"
    # extent 5 -> normalised: radius 0.2, centres at x = -0.3 and 0.3
    sd2, _, mi = probe.sd([f"stl {path2} 1 1 1 0 0 0"], [(-0.3, 0, 0), (0.3, 0.1, 0), (-0.3, 0.25, 0)])
    np.testing.assert_allclose(sd2, [-0.2, -0.1, 0.05], atol=2e-3)
    assert list(mi) == [0, 1, 0]  # components in the order of their first face
