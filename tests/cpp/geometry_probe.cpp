// Queries of include/dacti/scene/geometry.hpp for the closed-form checks in
// tests/test_geometry.py.  Reads commands from stdin, one per line, and
// prints one result line each:
//   shape <kind> <params...> <tx> <ty> <rz>   build a 2D shape (replaces the current meshes)
//   stl <path> <sx> <sy> <sz> <tx> <ty> <tz>  load an STL file
//   count                                     -> "<meshes> <outline points or faces of mesh 0>"
//   sd <x> <y> <z>                            -> "<signed distance> <nx> <ny> <nz> <mesh>" (nearest mesh)
//   ray <ax> <ay> <az> <bx> <by> <bz>         -> "<0|1>"
//   samples <kind> <params...>                -> "<n> <t0> <t1> ... " (first 4) of the adaptive sampling
// kinds: circle r | ellipse a b | naca m p t | joukowski m t delta | diamond w h |
//        triangle x1 y1 x2 y2 x3 y3 | quad x1 y1 .. x4 y4 | line x1 y1 x2 y2
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "dacti/scene/geometry.hpp"

using namespace dacti;

static geom::Curve curve_of(std::istream &in, const std::string &kind, double &t0, double &t1, bool &closed) {
    t0 = 0.0, t1 = 2 * M_PI, closed = true;
    double a, b, c;
    if (kind == "circle") {
        in >> a;
        return geom::circle(a);
    }
    if (kind == "ellipse") {
        in >> a >> b;
        return geom::ellipse(a, b);
    }
    if (kind == "naca") {
        in >> a >> b >> c;
        t0 = -1.0, t1 = 1.0, closed = false;
        return geom::naca4(a, b, c);
    }
    in >> a >> b >> c;
    return geom::joukowski(a, b, c);
}

int main() {
    std::vector<geom::Mesh> meshes;
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string cmd;
        in >> cmd;
        if (cmd == "shape") {
            std::string kind;
            in >> kind;
            std::vector<vec2_t> poly;
            geom::Curve cv;
            double t0 = 0, t1 = 0;
            bool closed = true, is_curve = false;
            if (kind == "circle" || kind == "ellipse" || kind == "naca" || kind == "joukowski") {
                cv = curve_of(in, kind, t0, t1, closed);
                is_curve = true;
            } else {
                int nv = kind == "diamond" ? 0 : kind == "triangle" ? 3 : kind == "quad" ? 4 : 2;
                if (kind == "diamond") {
                    double w, h;
                    in >> w >> h;
                    poly = {vec2_t(0, h / 2), vec2_t(w / 2, 0), vec2_t(0, -h / 2), vec2_t(-w / 2, 0)};
                }
                for (int k = 0; k < nv; ++k) {
                    double x, y;
                    in >> x >> y;
                    poly.push_back(vec2_t(x, y));
                }
            }
            geom::Transform tf;
            double tx, ty, rz;
            in >> tx >> ty >> rz;
            tf.translation = vec3_t(tx, ty, 0.0);
            tf.rotation = vec3_t(0.0, 0.0, rz);
            meshes = {is_curve ? geom::curve_mesh(cv, t0, t1, closed, tf) : geom::polygon_mesh(poly, tf)};
            std::printf("ok\n");
        } else if (cmd == "stl") {
            std::string path;
            geom::Transform tf;
            double sx, sy, sz, tx, ty, tz;
            in >> path >> sx >> sy >> sz >> tx >> ty >> tz;
            tf.scaling = vec3_t(sx, sy, sz);
            tf.translation = vec3_t(tx, ty, tz);
            meshes = geom::stl_meshes(path, tf);
            std::printf("ok\n");
        } else if (cmd == "count") {
            std::printf("%zu %zu\n", meshes.size(), meshes[0].planar ? meshes[0].outline.size() : meshes[0].F.size());
        } else if (cmd == "sd") {
            double x, y, z;
            in >> x >> y >> z;
            double best = 1e300, bd = 0;
            vec3_t bn;
            size_t bm = 0;
            for (size_t m = 0; m < meshes.size(); ++m) {
                geom::Closest c;
                const double sd = geom::signed_distance(meshes[m], vec3_t(x, y, z), c);
                if (sd < best) {
                    best = sd;
                    bd = sd;
                    bn = geom::normal_at(meshes[m], c);
                    bm = m;
                }
            }
            std::printf("%.17g %.17g %.17g %.17g %zu\n", bd, bn[0], bn[1], bn[2], bm);
        } else if (cmd == "ray") {
            double a[3], b[3];
            in >> a[0] >> a[1] >> a[2] >> b[0] >> b[1] >> b[2];
            bool hit = false;
            for (const auto &m : meshes) {
                double t;
                hit = hit || geom::intersect_segment(m, vec3_t(a[0], a[1], a[2]), vec3_t(b[0], b[1], b[2]), t);
            }
            std::printf("%d\n", hit ? 1 : 0);
        } else if (cmd == "samples") {
            std::string kind;
            in >> kind;
            double t0, t1;
            bool closed;
            const geom::Curve cv = curve_of(in, kind, t0, t1, closed);
            const auto s = geom::adaptive_sampling(cv, t0, t1, closed);
            std::printf("%zu", s.size());
            for (size_t k = 0; k < 4 && k < s.size(); ++k) std::printf(" %.17g", s[k]);
            std::printf(" %.17g\n", s.back());
        }
        std::fflush(stdout);
    }
    return 0;
}
