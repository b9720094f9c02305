// dacti/scene/geometry.hpp -- object geometry of a case: the reference's
// built-in 2D shapes and STL meshes as triangle meshes, with the queries the
// initial tree needs (signed distance, nearest point and normal, segment
// intersection).
//
// Shapes follow scene/geometry.hpp and scene/geometry_data.{hpp,cpp} of the
// reference: ellipse, circle, naca_airfoil and Joukowski_airfoil curves are
// sampled adaptively (a binary refinement of the parameter interval, levels
// 4..13, split while the spread of unit tangents over 101 probes exceeds
// 1e-3; samples at the leaf midpoints, plus the end points of an open
// curve), line, diamond, quadrilateral and triangle are polygons.  A 2D
// outline is extruded to z in [-1, 1] (side faces only, two triangles per
// segment, geometry_data.cpp:154-181) and placed by the object's transform
// p' = R S p + t, R = Rx Ry Rz from `rotation` in degrees
// (object_mesh.hpp:34-55, object_mesh.cpp:93-104).  STL files (ASCII or
// binary) are normalised to unit extent first, their duplicate vertices
// merged (1e-12) and split into connected components, one object each
// (object_mesh.cpp:13-30,93-179).
//
// Queries (replacing libigl's AABB distance, winding-number sign and
// embree's segment test, scene/scene.cpp:47-85,11-45, which are not part of
// this build; parity with them is unpinned):
//  * distance: exact closest point over the triangles, through a bounding
//    volume hierarchy;
//  * sign: negative inside, from the pseudonormal at the closest feature
//    (exact for closed, consistently oriented meshes and closed outlines);
//    open outlines (line, NACA) have no inside;
//  * normal at the closest point: the face normal, or for smooth (curved)
//    outlines and STL meshes the barycentric blend of angle-weighted vertex
//    normals (object_mesh.hpp:68-111);
//  * segment intersection: exact, over the triangles.
// Face orientation is made outward (the reference orients by libigl's
// orient_outward): a 2D outline by its signed area, a 3D component by its
// signed volume.
#ifndef DACTI_GEOMETRY_HPP
#define DACTI_GEOMETRY_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../core/types.hpp"

namespace dacti::geom {

// ------------------------------------------------------------ vector helpers
inline vec3_t cross(const vec3_t &a, const vec3_t &b) {
    return vec3_t(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}

// ------------------------------------------------------------ transform
struct Transform {
    vec3_t translation, rotation, scaling = vec3_t::One();
    scalar_t normalization = 1.0;

    // rotation matrix Rx(rx) Ry(ry) Rz(rz), angles in degrees
    std::array<std::array<double, 3>, 3> rotate() const {
        const double k = M_PI / 180.0;
        auto rx = [](double a) {
            return std::array<std::array<double, 3>, 3>{{{1, 0, 0}, {0, std::cos(a), -std::sin(a)}, {0, std::sin(a), std::cos(a)}}};
        };
        auto ry = [](double a) {
            return std::array<std::array<double, 3>, 3>{{{std::cos(a), 0, std::sin(a)}, {0, 1, 0}, {-std::sin(a), 0, std::cos(a)}}};
        };
        auto rz = [](double a) {
            return std::array<std::array<double, 3>, 3>{{{std::cos(a), -std::sin(a), 0}, {std::sin(a), std::cos(a), 0}, {0, 0, 1}}};
        };
        auto mul = [](const std::array<std::array<double, 3>, 3> &A, const std::array<std::array<double, 3>, 3> &B) {
            std::array<std::array<double, 3>, 3> C{};
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int l = 0; l < 3; ++l) C[i][j] += A[i][l] * B[l][j];
            return C;
        };
        return mul(mul(rx(rotation[0] * k), ry(rotation[1] * k)), rz(rotation[2] * k));
    }

    // p' = R S p + t
    vec3_t apply(const vec3_t &p) const {
        const auto R = rotate();
        vec3_t sp = p.cwiseMul(scaling), r;
        for (int i = 0; i < 3; ++i) r[i] = R[i][0] * sp[0] + R[i][1] * sp[1] + R[i][2] * sp[2];
        return r + translation;
    }
};

// ------------------------------------------------------------ curves
using Curve = std::function<vec2_t(double)>;

inline Curve ellipse(double a, double b) {
    return [a, b](double t) { return vec2_t(a * std::cos(t), b * std::sin(t)); };
}

inline Curve circle(double r) {
    return [r](double t) { return vec2_t(r * std::cos(t), r * std::sin(t)); };
}

// NACA 4-digit airfoil over x in [-1, 1] (lower side for x < 0), chord
// centred at 0 (scene/geometry.hpp:58-101)
inline Curve naca4(double m, double p, double t) {
    return [m, p, t](double x) {
        const double sgn = x >= 0 ? 1.0 : -1.0;
        x *= sgn;
        const double yt = 5 * t * (0.2969 * std::sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x -
                                   0.1036 * x * x * x * x);
        const double yc = x < p ? m * (2 * p * x - x * x) / (p * p) : m * (1 - 2 * p + 2 * p * x - x * x) / ((1 - p) * (1 - p));
        const double dyc = x < p ? 2 * m * (p - x) / (p * p) : 2 * m * (p - x) / ((1 - p) * (1 - p));
        const double th = std::atan(dyc);
        return vec2_t(x - sgn * yt * std::sin(th) - 0.5, yc + sgn * yt * std::cos(th));
    };
}

// Joukowski airfoil of camber m, thickness t and trailing-edge angle delta
// (degrees), unit chord centred at 0 (scene/geometry.hpp:104-161)
inline Curve joukowski(double m, double t, double delta) {
    using cd = std::complex<double>;
    const cd I(0.0, 1.0);
    const double xc = -1 / 1.299 * t, yc = 2.0 * m;
    const cd zeta0(xc, yc);
    const double beta = std::asin(yc), a = std::cos(beta) + xc, n = 2.0 - delta / 180.0;
    auto tr = [a, n](cd z) { return 0.5 * n * a * (std::pow(z + a, n) + std::pow(z - a, n)) / (std::pow(z + a, n) - std::pow(z - a, n)); };
    const double LE = std::real(tr(std::exp(I * (beta + M_PI)) + zeta0));
    const double TE = std::real(tr(std::exp(I * (-beta)) + zeta0));
    const double chord = TE - LE;
    return [=](double theta) {
        const cd z = tr(std::exp(I * theta) + zeta0) - (LE + TE) * 0.5;
        return vec2_t(std::real(z) / chord, std::imag(z) / chord);
    };
}

// adaptive parameter sampling (geometry_data.cpp:26-99)
inline std::vector<double> adaptive_sampling(const Curve &curve, double t0, double t1, bool closed, int min_level = 4,
                                             int max_level = 13, double max_error = 1e-3) {
    struct N {
        double a, len;
        int level;
        size_t left = 0, right = 0;
    };
    const double eps = 1e-7;
    std::vector<N> tree{{NAN, NAN, -1}, {0.5 * (t0 + t1), t1 - t0, 0}};
    std::vector<size_t> queue{1};  // front at the back: the reference pops the back and pushes the front
    size_t head = 0;
    while (head < queue.size()) {
        const size_t i = queue[head++];
        const double a0 = tree[i].a - tree[i].len / 2.0, a1 = tree[i].a + tree[i].len / 2.0;
        const int n = 100;
        vec2_t mn = vec2_t::constant(std::numeric_limits<double>::max()), mx = vec2_t::constant(std::numeric_limits<double>::lowest());
        for (int k = 0; k <= n; ++k) {
            const double a = a0 + (a1 - a0) / n * k;
            const vec2_t pn = curve(a - eps), pp = curve(a + eps);
            const vec2_t d = (pp - pn) / (pp - pn).norm();
            for (int q = 0; q < 2; ++q) {
                mn[q] = std::min(mn[q], d[q]);
                mx[q] = std::max(mx[q], d[q]);
            }
        }
        const double err = (mx - mn).norm();
        if ((err > max_error && tree[i].level < max_level) || tree[i].level < min_level) {
            const N cur = tree[i];
            tree[i].left = tree.size();
            tree[i].right = tree.size() + 1;
            tree.push_back({cur.a - 0.25 * cur.len, 0.5 * cur.len, cur.level + 1});
            tree.push_back({cur.a + 0.25 * cur.len, 0.5 * cur.len, cur.level + 1});
            queue.push_back(tree[i].left);
            queue.push_back(tree[i].right);
        }
    }
    std::vector<double> out;
    if (!closed) out.push_back(t0);
    std::function<void(size_t)> walk = [&](size_t i) {
        if (tree[i].left && tree[i].right) {
            walk(tree[i].left);
            walk(tree[i].right);
        } else {
            out.push_back(tree[i].a);
        }
    };
    walk(1);
    if (!closed) out.push_back(t1);
    return out;
}

// ------------------------------------------------------------ mesh
struct Mesh {
    std::vector<vec3_t> V;
    std::vector<std::array<int, 3>> F;
    std::vector<vec3_t> FN, VN;  // unit face normals, angle-weighted vertex normals
    std::vector<std::array<int, 3>> adj;  // face across edge k = (F[k], F[k+1]), -1 at a border
    bool is_smooth = false;
    // 2D outline (extruded meshes): the polygon in its plane
    bool planar = false, closed = false;
    std::vector<vec2_t> outline;

    // bounding volume hierarchy over the faces
    struct BNode {
        vec3_t lo, hi;
        int left = -1, right = -1, first = 0, count = 0;
    };
    std::vector<BNode> bvh;
    std::vector<int> order;

    void finish() {
        compute_normals();
        build_bvh();
        std::map<std::pair<int, int>, std::vector<std::pair<int, int>>> edges;
        for (size_t f = 0; f < F.size(); ++f)
            for (int k = 0; k < 3; ++k) {
                const int a = F[f][k], b = F[f][(k + 1) % 3];
                edges[{std::min(a, b), std::max(a, b)}].push_back({(int) f, k});
            }
        adj.assign(F.size(), {-1, -1, -1});
        for (const auto &e : edges)
            if (e.second.size() == 2) {
                adj[(size_t) e.second[0].first][(size_t) e.second[0].second] = e.second[1].first;
                adj[(size_t) e.second[1].first][(size_t) e.second[1].second] = e.second[0].first;
            }
    }

    void compute_normals() {
        FN.assign(F.size(), vec3_t());
        VN.assign(V.size(), vec3_t());
        for (size_t f = 0; f < F.size(); ++f) {
            const vec3_t &a = V[F[f][0]], &b = V[F[f][1]], &c = V[F[f][2]];
            vec3_t n = cross(b - a, c - a);
            const double l = n.norm();
            FN[f] = l > 0 ? n / l : vec3_t();
            for (int k = 0; k < 3; ++k) {
                const vec3_t &p = V[F[f][k]], &q = V[F[f][(k + 1) % 3]], &r = V[F[f][(k + 2) % 3]];
                const vec3_t u = q - p, w = r - p;
                const double nu = u.norm(), nw = w.norm();
                if (nu == 0 || nw == 0) continue;
                const double ang = std::acos(std::max(-1.0, std::min(1.0, u.dot(w) / (nu * nw))));
                VN[F[f][k]] += FN[f] * ang;
            }
        }
        for (auto &n : VN) {
            const double l = n.norm();
            if (l > 0) n = n / l;
        }
    }

    void build_bvh() {
        order.resize(F.size());
        std::iota(order.begin(), order.end(), 0);
        bvh.clear();
        if (!F.empty()) build(0, (int) F.size());
    }

    int build(int first, int count) {
        BNode nd;
        nd.lo = vec3_t::constant(std::numeric_limits<double>::max());
        nd.hi = vec3_t::constant(std::numeric_limits<double>::lowest());
        for (int i = first; i < first + count; ++i)
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) {
                    nd.lo[a] = std::min(nd.lo[a], V[F[order[i]][k]][a]);
                    nd.hi[a] = std::max(nd.hi[a], V[F[order[i]][k]][a]);
                }
        const int id = (int) bvh.size();
        bvh.push_back(nd);
        if (count <= 4) {
            bvh[id].first = first;
            bvh[id].count = count;
            return id;
        }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (nd.hi[a] - nd.lo[a] > nd.hi[axis] - nd.lo[axis]) axis = a;
        auto cen = [&](int f) { return V[F[f][0]][axis] + V[F[f][1]][axis] + V[F[f][2]][axis]; };
        const int mid = first + count / 2;
        std::nth_element(order.begin() + first, order.begin() + mid, order.begin() + first + count,
                         [&](int x, int y) { return cen(x) < cen(y); });
        const int l = build(first, mid - first);
        const int r = build(mid, first + count - mid);
        bvh[id].left = l;
        bvh[id].right = r;
        return id;
    }
};

// closest point on triangle abc to p (Ericson, Real-Time Collision
// Detection 5.1.5); region: 0 face, 1..3 vertex a/b/c, 4..6 edge ab/bc/ca
inline vec3_t closest_on_triangle(const vec3_t &p, const vec3_t &a, const vec3_t &b, const vec3_t &c, int &region,
                                  vec3_t &bary) {
    const vec3_t ab = b - a, ac = c - a, ap = p - a;
    const double d1 = ab.dot(ap), d2 = ac.dot(ap);
    if (d1 <= 0 && d2 <= 0) { region = 1; bary = vec3_t(1, 0, 0); return a; }
    const vec3_t bp = p - b;
    const double d3 = ab.dot(bp), d4 = ac.dot(bp);
    if (d3 >= 0 && d4 <= d3) { region = 2; bary = vec3_t(0, 1, 0); return b; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double v = d1 / (d1 - d3);
        region = 4; bary = vec3_t(1 - v, v, 0);
        return a + v * ab;
    }
    const vec3_t cp = p - c;
    const double d5 = ab.dot(cp), d6 = ac.dot(cp);
    if (d6 >= 0 && d5 <= d6) { region = 3; bary = vec3_t(0, 0, 1); return c; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const double w = d2 / (d2 - d6);
        region = 6; bary = vec3_t(1 - w, 0, w);
        return a + w * ac;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        region = 5; bary = vec3_t(0, 1 - w, w);
        return b + w * (c - b);
    }
    const double den = 1.0 / (va + vb + vc), v = vb * den, w = vc * den;
    region = 0; bary = vec3_t(1 - v - w, v, w);
    return a + ab * v + ac * w;
}

struct Closest {
    double sq = std::numeric_limits<double>::max();
    int face = -1, region = 0;
    vec3_t point, bary;
};

inline double box_sq(const vec3_t &p, const vec3_t &lo, const vec3_t &hi) {
    double s = 0;
    for (int a = 0; a < 3; ++a) {
        const double d = p[a] < lo[a] ? lo[a] - p[a] : (p[a] > hi[a] ? p[a] - hi[a] : 0.0);
        s += d * d;
    }
    return s;
}

inline Closest closest_point(const Mesh &m, const vec3_t &p) {
    Closest best;
    if (m.bvh.empty()) return best;
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const int id = stack.back();
        stack.pop_back();
        const auto &nd = m.bvh[(size_t) id];
        if (box_sq(p, nd.lo, nd.hi) > best.sq) continue;
        if (nd.left < 0) {
            for (int i = nd.first; i < nd.first + nd.count; ++i) {
                const int f = m.order[(size_t) i];
                int reg;
                vec3_t bary;
                const vec3_t q = closest_on_triangle(p, m.V[m.F[f][0]], m.V[m.F[f][1]], m.V[m.F[f][2]], reg, bary);
                const double s = (q - p).squaredNorm();
                if (s < best.sq) best = {s, f, reg, q, bary};
            }
        } else {
            const double dl = box_sq(p, m.bvh[(size_t) nd.left].lo, m.bvh[(size_t) nd.left].hi);
            const double dr = box_sq(p, m.bvh[(size_t) nd.right].lo, m.bvh[(size_t) nd.right].hi);
            if (dl < dr) {
                stack.push_back(nd.right);
                stack.push_back(nd.left);
            } else {
                stack.push_back(nd.left);
                stack.push_back(nd.right);
            }
        }
    }
    return best;
}

// normal of the mesh at a closest point: face normal, or the blend of vertex
// normals for smooth meshes (not normalised, object_mesh.hpp:86-101)
inline vec3_t normal_at(const Mesh &m, const Closest &c) {
    if (!m.is_smooth) return m.FN[(size_t) c.face];
    const auto &f = m.F[(size_t) c.face];
    return c.bary[0] * m.VN[(size_t) f[0]] + c.bary[1] * m.VN[(size_t) f[1]] + c.bary[2] * m.VN[(size_t) f[2]];
}

// signed distance (negative inside) of p to the mesh; c receives the
// closest point.  The sign is that of (p - c) . n with n the pseudonormal of
// the closest feature: the face normal, the sum of the two face normals at
// an edge, the angle-weighted normal at a vertex.  Open outlines have no
// inside.
inline double signed_distance(const Mesh &m, const vec3_t &p, Closest &c) {
    c = closest_point(m, p);
    const double d = std::sqrt(c.sq);
    if (c.face < 0 || (m.planar && !m.closed)) return d;
    vec3_t n;
    const auto &f = m.F[(size_t) c.face];
    if (c.region == 0) {
        n = m.FN[(size_t) c.face];
    } else if (c.region <= 3) {
        n = m.VN[(size_t) f[c.region - 1]];
    } else {
        n = m.FN[(size_t) c.face];
        const int g = m.adj[(size_t) c.face][c.region - 4];
        if (g >= 0) n += m.FN[(size_t) g];
    }
    return (p - c.point).dot(n) < 0 ? -d : d;
}

// does the segment a -> b cross a face (Moller-Trumbore, t in [0, 1])?
inline bool intersect_segment(const Mesh &m, const vec3_t &a, const vec3_t &b, double &t_hit) {
    if (m.bvh.empty()) return false;
    const vec3_t dir = b - a;
    vec3_t lo, hi;
    for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(a[k], b[k]);
        hi[k] = std::max(a[k], b[k]);
    }
    bool hit = false;
    t_hit = std::numeric_limits<double>::max();
    std::vector<int> stack{0};
    while (!stack.empty()) {
        const auto &nd = m.bvh[(size_t) stack.back()];
        stack.pop_back();
        bool overlap = true;
        for (int k = 0; k < 3; ++k) overlap = overlap && nd.lo[k] <= hi[k] && nd.hi[k] >= lo[k];
        if (!overlap) continue;
        if (nd.left >= 0) {
            stack.push_back(nd.left);
            stack.push_back(nd.right);
            continue;
        }
        for (int i = nd.first; i < nd.first + nd.count; ++i) {
            const auto &f = m.F[(size_t) m.order[(size_t) i]];
            const vec3_t &v0 = m.V[(size_t) f[0]], e1 = m.V[(size_t) f[1]] - v0, e2 = m.V[(size_t) f[2]] - v0;
            const vec3_t pv = cross(dir, e2);
            const double det = e1.dot(pv);
            if (std::abs(det) < 1e-300) continue;
            const double inv = 1.0 / det;
            const vec3_t tv = a - v0;
            const double u = tv.dot(pv) * inv;
            if (u < 0 || u > 1) continue;
            const vec3_t qv = cross(tv, e1);
            const double v = dir.dot(qv) * inv;
            if (v < 0 || u + v > 1) continue;
            const double t = e2.dot(qv) * inv;
            if (t < 0 || t > 1) continue;
            hit = true;
            t_hit = std::min(t_hit, t);
        }
    }
    return hit;
}

// ------------------------------------------------------------ builders
// extruded outline (z in [-1, 1]); closed outlines wrap
inline Mesh extrude(const std::vector<vec2_t> &pts, bool closed, bool smooth, const Transform &tf) {
    Mesh m;
    m.planar = true;
    m.closed = closed && pts.size() >= 3;
    m.is_smooth = smooth;
    for (const vec2_t &p : pts) {
        m.V.push_back(tf.apply(vec3_t(p[0], p[1], -1.0)));
        m.V.push_back(tf.apply(vec3_t(p[0], p[1], 1.0)));
        const vec3_t q = tf.apply(vec3_t(p[0], p[1], 0.0));
        m.outline.push_back(vec2_t(q[0], q[1]));
    }
    const size_t n = pts.size(), segs = closed ? n : n - 1;
    for (size_t i = 0; i < segs; ++i) {
        const int j = (int) ((i + 1) % n), k = (int) i;
        m.F.push_back({2 * k, 2 * j, 2 * k + 1});
        m.F.push_back({2 * j, 2 * j + 1, 2 * k + 1});
    }
    // outward faces: for a counter-clockwise outline the side faces above
    // already point outward; a clockwise one is flipped
    double area = 0;
    for (size_t i = 0; i < m.outline.size(); ++i) {
        const vec2_t &a = m.outline[i], &b = m.outline[(i + 1) % m.outline.size()];
        area += a[0] * b[1] - b[0] * a[1];
    }
    if (area < 0)
        for (auto &f : m.F) std::swap(f[1], f[2]);
    m.finish();
    return m;
}

inline Mesh curve_mesh(const Curve &c, double t0, double t1, bool closed, const Transform &tf) {
    std::vector<vec2_t> pts;
    for (double a : adaptive_sampling(c, t0, t1, closed)) pts.push_back(c(a));
    return extrude(pts, closed, true, tf);
}

inline Mesh polygon_mesh(const std::vector<vec2_t> &verts, const Transform &tf) {
    return extrude(verts, verts.size() >= 3, false, tf);
}

// STL file (ASCII or binary) as raw triangles
inline std::vector<std::array<vec3_t, 3>> read_stl(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("cannot open geometry file " + path);
    std::string data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::vector<std::array<vec3_t, 3>> tris;
    if (data.size() >= 84) {
        uint32_t n;
        std::memcpy(&n, data.data() + 80, 4);
        if (84 + (size_t) n * 50 == data.size()) {
            for (uint32_t i = 0; i < n; ++i) {
                std::array<vec3_t, 3> t;
                for (int k = 0; k < 3; ++k) {
                    float xyz[3];
                    std::memcpy(xyz, data.data() + 84 + (size_t) i * 50 + 12 + 12 * k, 12);
                    t[k] = vec3_t(xyz[0], xyz[1], xyz[2]);
                }
                tris.push_back(t);
            }
            return tris;
        }
    }
    std::istringstream ss(data);
    std::string tok;
    std::array<vec3_t, 3> t;
    int k = 0;
    while (ss >> tok) {
        if (tok == "vertex") {
            double x, y, z;
            ss >> x >> y >> z;
            t[k++] = vec3_t(x, y, z);
            if (k == 3) {
                tris.push_back(t);
                k = 0;
            }
        }
    }
    if (tris.empty()) throw std::runtime_error("no triangles in " + path);
    return tris;
}

// meshes of an STL file: normalised to unit extent, transformed, duplicate
// vertices merged, one mesh per connected component, faces outward
inline std::vector<Mesh> stl_meshes(const std::string &path, Transform tf) {
    const auto tris = read_stl(path);
    vec3_t lo = vec3_t::constant(std::numeric_limits<double>::max()), hi = -lo;
    for (const auto &t : tris)
        for (const auto &v : t)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], v[a]);
                hi[a] = std::max(hi[a], v[a]);
            }
    const vec3_t center = 0.5 * (lo + hi);
    double ext = 0;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, hi[a] - lo[a]);
    tf.normalization /= ext;
    tf.translation -= center * tf.normalization;
    // merge vertices closer than 1e-12 (grid hashing)
    std::vector<vec3_t> V;
    std::vector<std::array<int, 3>> F;
    std::map<std::array<long long, 3>, std::vector<int>> grid;
    auto key = [](const vec3_t &p) {
        return std::array<long long, 3>{(long long) std::floor(p[0] * 1e10), (long long) std::floor(p[1] * 1e10),
                                        (long long) std::floor(p[2] * 1e10)};
    };
    auto vid = [&](const vec3_t &p) {
        const auto k0 = key(p);
        for (long long dx = -1; dx <= 1; ++dx)
            for (long long dy = -1; dy <= 1; ++dy)
                for (long long dz = -1; dz <= 1; ++dz) {
                    auto it = grid.find({k0[0] + dx, k0[1] + dy, k0[2] + dz});
                    if (it == grid.end()) continue;
                    for (int i : it->second)
                        if ((V[(size_t) i] - p).norm() < 1e-12) return i;
                }
        V.push_back(p);
        grid[k0].push_back((int) V.size() - 1);
        return (int) V.size() - 1;
    };
    for (const auto &t : tris) {
        std::array<int, 3> f;
        for (int k = 0; k < 3; ++k) f[k] = vid(tf.apply(t[k] * tf.normalization));
        if (f[0] != f[1] && f[1] != f[2] && f[0] != f[2]) F.push_back(f);
    }
    // connected components over shared vertices
    std::vector<int> comp(V.size());
    std::iota(comp.begin(), comp.end(), 0);
    std::function<int(int)> root = [&](int x) { return comp[(size_t) x] == x ? x : comp[(size_t) x] = root(comp[(size_t) x]); };
    for (const auto &f : F)
        for (int k = 1; k < 3; ++k) comp[(size_t) root(f[k])] = root(f[0]);
    std::map<int, size_t> cid;
    std::vector<Mesh> out;
    std::vector<std::map<int, int>> remap;
    for (const auto &f : F) {
        const int r = root(f[0]);
        if (!cid.count(r)) {
            cid[r] = out.size();
            out.emplace_back();
            remap.emplace_back();
        }
        const size_t c = cid[r];
        std::array<int, 3> g;
        for (int k = 0; k < 3; ++k) {
            auto it = remap[c].find(f[k]);
            if (it == remap[c].end()) {
                it = remap[c].emplace(f[k], (int) out[c].V.size()).first;
                out[c].V.push_back(V[(size_t) f[k]]);
            }
            g[k] = it->second;
        }
        out[c].F.push_back(g);
    }
    for (auto &m : out) {
        double vol = 0;
        for (const auto &f : m.F) vol += m.V[(size_t) f[0]].dot(cross(m.V[(size_t) f[1]], m.V[(size_t) f[2]]));
        if (vol < 0)
            for (auto &f : m.F) std::swap(f[1], f[2]);
        m.is_smooth = true;
        m.finish();
    }
    return out;
}

}  // namespace dacti::geom

#endif
