// dacti/config/case.hpp -- the reference's TOML case format and initial tree, in C++.
//
// load_case() reads the keys of Config<M>::load_config
// (src/lib/config/config.cpp:10-141) with the strict typing of get_entry
// (config/util.hpp:15-103: vectors are float arrays of exact length, floats
// must be written as floats, ints as ints), the model reference values
// (model/euler.hpp:288-324, model/navier_stokes.hpp:120-134,
// model/multi.hpp:474-549), the boundary conditions (config.cpp:145-190)
// and initial conditions (config.cpp:194-243).  build_tree() grows the
// initial octree level by level as Scene<M>::init_domain does
// (scene/scene.cpp:369-431).  to_dacti_case() / Tree::view() give the
// structures dacti_create takes.  Quirks of the reference are kept and
// marked.  Scope as in the Python loader (dacti_amd/case.py,
// dacti_amd/scene.py), which this header reproduces bit for bit: domain
// boundary conditions, periodic axes, uniform and jump initial conditions,
// the objects of [[geometry]] (built-in 2D shapes and STL meshes, see
// scene/geometry.hpp; an analytic "sphere" of `radius` is an extension of
// this build), fluid objects ("bubble", "droplet") and solids with their
// boundary conditions, axis-aligned refinement zones.  Refinement zones
// with transform_like keep their axis-aligned box.
#ifndef DACTI_CASE_HPP
#define DACTI_CASE_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <filesystem>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../dacti_gpu.h"
#include "../scene/geometry.hpp"
#include "toml.hpp"

namespace dacti {

struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct BC {
    int type = DACTI_BC_ZERO_GRADIENT;
    std::vector<double> value;
    double fade_in = 0.0;
};

// An object of the scene (config/config.hpp:69-77): one connected component
// of a geometry's mesh, or an analytic sphere, with the geometry's type
// ("solid" by default, "droplet", "bubble", ...) and boundary condition.
struct Object {
    std::string name, type = "solid", shape;
    bool has_bc = false;
    BC bc;
    bool analytic = false;  // sphere of `radius` around `center`
    std::array<double, 3> center{0, 0, 0};
    double radius = 0;
    std::shared_ptr<geom::Mesh> mesh;
};

struct Zone {
    std::array<double, 3> x0, x1;
    int level;
};

struct Case {
    int model = DACTI_MODEL_EULER, dim = 2, limiter = DACTI_LIMITER_MINMOD;
    std::string name = "case";
    std::array<double, 3> active_size{1, 1, 1};
    double side_length = 1;
    std::array<bool, 3> periodic{false, false, false};
    double print_interval = 1, end_time = 1, signal_time = -1, steady_tol = -1;
    double max_cfl = 0.5, max_dt = 1;
    int min_refinement_level = 0, max_refinement_level = 0, max_refinement_level_geom = 0;
    double refinement_radius = 2.0;
    std::array<double, 2> error_threshold{-1.0, 1e300};
    std::vector<double> weights;
    std::map<std::string, double> consts;
    std::vector<double> p_ref, vel_ref;
    bool has_rho_ref = false;
    std::array<double, 2> rho_ref{0, 0};
    std::array<BC, 6> bcs;
    std::string ic_type = "uniform";
    int ic_axis = 0;
    double ic_location = 0;
    std::vector<double> ic_neg, ic_pos;
    // ic_type "grid" (an extension of this build: a primitive state on a
    // uniform grid over the active box, filled by the caller): row-major
    // [shape[0]][shape[1]]([shape[2]])[num_eqns]
    std::vector<double> ic_grid;
    std::array<int, 3> ic_grid_shape{0, 0, 0};
    std::vector<Object> objects;
    std::vector<Zone> zones;

    int num_eqns() const { return model == DACTI_MODEL_MULTI ? dim + 4 : dim + 2; }
};

namespace detail {

inline const toml::Value *at(const toml::Value &t, const std::string &path) { return t.at_path(path); }

[[noreturn]] inline void missing(const std::string &path) {
    throw ConfigError("Config: Couldn't find entry for " + path + ".");
}

inline double get_float(const toml::Value &t, const std::string &path, const double *dflt = nullptr) {
    const toml::Value *v = at(t, path);
    if (!v) {
        if (dflt) return *dflt;
        missing(path);
    }
    if (v->kind != toml::Value::FLOAT) throw ConfigError("Config: Type of " + path + " should be floating point");
    return v->f;
}

inline int64_t get_int(const toml::Value &t, const std::string &path, bool unsig, const int64_t *dflt = nullptr) {
    const toml::Value *v = at(t, path);
    if (!v) {
        if (dflt) return *dflt;
        missing(path);
    }
    if (v->kind != toml::Value::INTEGER) throw ConfigError("Config: Type of " + path + " should be int");
    if (unsig && v->i < 0) throw ConfigError("Config: " + path + " should be unsigned");
    return v->i;
}

inline std::string get_str(const toml::Value &t, const std::string &path, const char *dflt = nullptr) {
    const toml::Value *v = at(t, path);
    if (!v) {
        if (dflt) return dflt;
        missing(path);
    }
    if (v->kind != toml::Value::STRING) throw ConfigError("Config: Type of " + path + " should be string");
    return v->s;
}

inline std::vector<double> get_vec(const toml::Value &t, const std::string &path, size_t n,
                                   const std::vector<double> *dflt = nullptr) {
    const toml::Value *v = at(t, path);
    if (!v) {
        if (dflt) return *dflt;
        missing(path);
    }
    if (v->kind != toml::Value::ARRAY) throw ConfigError("Config: Type of " + path + " should be array");
    if (v->a.size() != n)
        throw ConfigError("Config: Length of " + path + " should be " + std::to_string(n) + ", but is " +
                          std::to_string(v->a.size()));
    std::vector<double> out;
    for (size_t i = 0; i < n; ++i) {
        if (v->a[i]->kind != toml::Value::FLOAT)
            throw ConfigError("Config: Type of " + path + "[" + std::to_string(i) + "] should be floating point");
        out.push_back(v->a[i]->f);
    }
    return out;
}

inline double multi_gamma(const std::map<std::string, double> &c, double a) {
    if (a < 1e-20) return c.at("gamma2");
    if (a > 1 - 1e-20) return c.at("gamma1");
    return 1 + 1 / (a / (c.at("gamma1") - 1) + (1 - a) / (c.at("gamma2") - 1));
}

inline double multi_p0(const std::map<std::string, double> &c, double a) {
    const double g1 = c.at("gamma1"), g2 = c.at("gamma2"), P1 = c.at("P1"), P2 = c.at("P2");
    if (a < 1e-20) return P2;
    if (a > 1 - 1e-20) return P1;
    return (a * g1 * P1 / (g1 - 1) + (1 - a) * g2 * P2 / (g2 - 1)) / (1 + a / (g1 - 1) + (1 - a) / (g2 - 1));
}

// model constants at their static-initialisation values
inline std::map<std::string, double> default_consts() {
    const double gamma1 = 6.12, gamma2 = 1.4, cp1 = 4181.0, cp2 = 1005.0;
    return {{"gamma", 1.4},   {"ns_mu", 0.0},    {"ns_lambda", 0.0}, {"ns_cp", 1005.0}, {"ns_R", 287.0},
            {"sigma", 0.0},   {"gamma1", gamma1}, {"gamma2", gamma2}, {"P1", 343000000.0}, {"P2", 0.0},
            {"cp1", cp1},     {"cp2", cp2},
            // multi.hpp:57-58: cv1/cv2 are fixed at static initialisation from
            // the default gammas; fluid_1/fluid_2 overrides do not reach them
            {"cv1", cp1 / gamma1}, {"cv2", cp2 / gamma2},
            {"mu1", 0.0},     {"mu2", 0.0},      {"lambda1", 0.0},   {"lambda2", 0.0}};
}

// Model::dict: euler.hpp:288-324, navier_stokes.hpp:120-134, multi.hpp:474-549
inline void model_dict(Case &cs, const toml::Value *m) {
    const int D = cs.dim;
    auto &c = cs.consts;
    const toml::Value *ref = m ? m->get("reference_values") : nullptr;
    if (!ref || !ref->is_table()) throw ConfigError("Config: model.reference_values missing");
    auto angle_vel = [&](double vm, double va) {
        std::vector<double> vel{vm * std::cos(va * M_PI / 180), vm * std::sin(va * M_PI / 180)};
        while ((int) vel.size() < D) vel.push_back(0.0);
        return vel;
    };
    const double zero = 0.0;
    std::vector<double> vel;
    if (cs.model != DACTI_MODEL_MULTI) {
        const double rho = get_float(*ref, "rho"), prs = get_float(*ref, "p");
        if (ref->get("vel_magnt")) {
            vel = angle_vel(get_float(*ref, "vel_magnt"), get_float(*ref, "vel_angle", &zero));
        } else if (ref->get("Ma")) {
            const double mach = get_float(*ref, "Ma"), va = get_float(*ref, "vel_angle", &zero);
            const double vm = mach * std::sqrt(1.4 * prs / rho);
            if (D == 2) vel = {vm * std::cos(va * M_PI / 180), vm * std::sin(va * M_PI / 180)};
            else vel = {-vm * std::cos(va * M_PI / 180), 0.0, vm * std::sin(va * M_PI / 180)};
        } else {
            vel = get_vec(*ref, "vel", D);
        }
        cs.p_ref = {rho};
        cs.p_ref.insert(cs.p_ref.end(), vel.begin(), vel.end());
        cs.p_ref.push_back(prs);
        cs.vel_ref = vel;
        if (cs.model == DACTI_MODEL_NAVIER_STOKES) {
            const double Re = get_float(*ref, "Re"), L = get_float(*ref, "L");
            double sp = 0;
            for (double v : vel) sp += v * v;
            const double speed = std::sqrt(sp);
            c["ns_mu"] = cs.p_ref[0] * speed * L / Re;
            if (ref->get("Pr")) c["ns_lambda"] = c["ns_mu"] * c["ns_cp"] / get_float(*ref, "Pr");
        }
        return;
    }
    const std::vector<double> rho = get_vec(*ref, "rho", 2);
    const double prs = get_float(*ref, "p"), alpha = get_float(*ref, "alpha");
    cs.has_rho_ref = true;
    cs.rho_ref = {rho[0], rho[1]};
    if (ref->get("vel_magnt")) {
        vel = angle_vel(get_float(*ref, "vel_magnt"), get_float(*ref, "vel_angle", &zero));
    } else if (ref->get("Ma")) {
        const double mach = get_float(*ref, "Ma"), va = get_float(*ref, "vel_angle", &zero);
        const double r0 = rho[0] * alpha, r1 = rho[1] * (1 - alpha);
        // the sound speed before the fluid overrides, as the reference
        const double vm = mach * std::sqrt(multi_gamma(c, alpha) * (prs + multi_p0(c, alpha)) / (r0 + r1));
        vel = angle_vel(vm, va);
    } else {
        vel = get_vec(*ref, "vel", D);
    }
    cs.p_ref = {rho[0] * alpha, rho[1] * (1.0 - alpha)};
    cs.p_ref.insert(cs.p_ref.end(), vel.begin(), vel.end());
    cs.p_ref.push_back(prs);
    cs.p_ref.push_back(alpha);
    cs.vel_ref = vel;
    const double one = 1.0;
    const double L = get_float(*ref, "L", &one), mu_ratio = get_float(*ref, "mu_ratio", &one);
    bool has_we = false;
    double We = 0;
    if (ref->get("We")) {
        We = get_float(*ref, "We");
        has_we = true;
        c["sigma"] = 1.0 / We;
    }
    if (ref->get("Re")) {
        const double Re = get_float(*ref, "Re");
        if (!has_we) throw ConfigError("Config: Re given without We (the reference reads an uninitialised We)");
        c["mu1"] = std::sqrt(rho[0] * c["sigma"] * L) * std::sqrt(We) / Re;
        c["mu2"] = c["mu1"] / mu_ratio;
    }
    if (ref->get("Oh")) {
        const double Oh = get_float(*ref, "Oh");
        c["sigma"] = get_float(*ref, "sigma");
        c["mu1"] = std::sqrt(rho[0] * c["sigma"] * L) * Oh;
        c["mu2"] = c["mu1"] / mu_ratio;
    }
    for (int k = 1; k <= 2; ++k) {
        const toml::Value *fl = m->get(k == 1 ? "fluid_1" : "fluid_2");
        if (fl && fl->is_table()) {
            const std::string P = k == 1 ? "P1" : "P2", G = k == 1 ? "gamma1" : "gamma2";
            const double p0 = c[P], g0 = c[G];
            c[P] = get_float(*fl, "P0", &p0);
            c[G] = get_float(*fl, "gamma", &g0);
        }
    }
}

// parse_boundary_condition: config/config.cpp:145-190
inline BC parse_bc(const Case &cs, const toml::Value *t) {
    static const toml::Value empty;
    const toml::Value &tb = t ? *t : empty;
    const std::string type = get_str(tb, "type");
    const double zero = 0.0;
    if (type == "noslip") return {DACTI_BC_NOSLIP, {}, 0.0};
    if (type == "freeslip") return {DACTI_BC_FREESLIP, {}, 0.0};
    if (type == "zeroGradient") return {DACTI_BC_ZERO_GRADIENT, {}, 0.0};
    if (type == "inflow") return {DACTI_BC_INFLOW, cs.p_ref, get_float(tb, "fade_in", &zero)};
    if (type == "inflow_r") {
        const double fade = get_float(tb, "fade_in", &zero);
        return {DACTI_BC_INFLOW_R, get_vec(tb, "value", cs.num_eqns()), fade};
    }
    if (type == "outflow") return {DACTI_BC_OUTFLOW, {}, 0.0};
    throw ConfigError("illegal bc type");
}

}  // namespace detail

inline int limiter_id(const std::string &name) {
    static const std::map<std::string, int> ids = {{"minmod", 0}, {"unlimited", 1}, {"van_leer", 2}, {"superbee", 3},
                                                   {"von_alba_2", 4}, {"koren", 5}, {"first_order", 6}};
    auto it = ids.find(name);
    if (it == ids.end()) throw ConfigError("unknown limiter " + name);
    return it->second;
}

// Config<M>::load_config on a parsed table (config/config.cpp:10-141)
inline Case load_case(const toml::Value &tbl, int model, int dim, int limiter = DACTI_LIMITER_MINMOD,
                      const std::string &base_dir = ".") {
    using namespace detail;
    Case cs;
    cs.model = model;
    cs.dim = dim;
    cs.limiter = limiter;
    cs.consts = default_consts();
    const int ne = cs.num_eqns();
    cs.name = get_str(tbl, "scene.name");
    const std::vector<double> as = get_vec(tbl, "scene.domain_size", dim);
    for (int i = 0; i < dim; ++i) cs.active_size[i] = as[i];
    cs.side_length = *std::max_element(as.begin(), as.end());
    cs.print_interval = get_float(tbl, "run.print_interval");
    cs.end_time = get_float(tbl, "run.end_time");
    const double m1 = -1.0, two = 2.0;
    cs.signal_time = get_float(tbl, "run.signal_time", &m1);
    cs.steady_tol = get_float(tbl, "run.steady_tol", &m1);
    cs.max_cfl = get_float(tbl, "integrator.max_cfl");
    cs.max_dt = get_float(tbl, "integrator.max_dt", &cs.print_interval);
    cs.min_refinement_level = (int) get_int(tbl, "integrator.min_refinement_level", false);
    cs.max_refinement_level = (int) get_int(tbl, "integrator.max_refinement_level", false);
    const int64_t mrl = cs.max_refinement_level;
    cs.max_refinement_level_geom = (int) get_int(tbl, "integrator.max_refinement_level_geom", false, &mrl);
    cs.refinement_radius = get_float(tbl, "integrator.refinement_radius", &two);
    const std::vector<double> et = get_vec(tbl, "integrator.error_threshold", 2);
    cs.error_threshold = {et[0], et[1]};
    const std::vector<double> ones((size_t) ne, 1.0);
    cs.weights = get_vec(tbl, "integrator.error_weights", ne, &ones);
    model_dict(cs, tbl.get("model"));
    // initial condition: config/config.cpp:194-243
    const toml::Value *ic = tbl.get("initial_condition");
    if (!ic || !ic->is_table()) throw ConfigError("Config: initial_condition missing");
    cs.ic_type = get_str(*ic, "type");
    if (cs.ic_type == "uniform") {
        const std::string value = get_str(*ic, "value", "reference");
        if (value == "fixed") get_vec(*ic, "fixed_value", ne);  // read and ignored, as the reference (:208-211)
        else if (value != "reference") throw ConfigError("Unknown uniform value type " + value);
    } else if (cs.ic_type == "jump") {
        cs.ic_axis = (int) get_int(*ic, "axis", true);
        cs.ic_location = get_float(*ic, "location");
        cs.ic_neg = get_vec(*ic, "value_neg", ne);
        cs.ic_pos = get_vec(*ic, "value_pos", ne);
    } else if (cs.ic_type == "grid") {
        // filled by the caller (ic_grid, ic_grid_shape)
    } else {
        throw ConfigError("illegal arg");
    }
    // domain boundary conditions: config/config.cpp:47-61
    static const char *names[3] = {"X", "Y", "Z"};
    const toml::Value *bct = tbl.get("boundary_conditions");
    for (int d = 0; d < dim; ++d) {
        const toml::Value *sub = bct ? bct->get(names[d]) : nullptr;
        const toml::Value *ty = sub ? sub->get("type") : nullptr;
        if (ty && ty->kind == toml::Value::STRING) {
            cs.periodic[d] = ty->s == "periodic";
        } else {
            cs.bcs[2 * d] = parse_bc(cs, sub ? sub->get("neg") : nullptr);
            cs.bcs[2 * d + 1] = parse_bc(cs, sub ? sub->get("pos") : nullptr);
        }
    }
    // geometry: config/config.cpp:62-125, scene/geometry_data.hpp:34-146
    if (const toml::Value *geo = tbl.get("geometry")) {
        const std::vector<double> z3{0.0, 0.0, 0.0}, o3{1.0, 1.0, 1.0};
        for (const auto &g : geo->a) {
            const std::string name = get_str(*g, "name");
            const std::string gtype = get_str(*g, "type", "solid");
            geom::Transform tf;
            const auto rot = get_vec(*g, "rotation", 3, &z3), sc = get_vec(*g, "scaling", 3, &o3),
                       tr = get_vec(*g, "translation", 3, &z3);
            tf.rotation = vec3_t(rot[0], rot[1], rot[2]);
            tf.scaling = vec3_t(sc[0], sc[1], sc[2]);
            tf.translation = vec3_t(tr[0], tr[1], tr[2]);
            std::vector<std::shared_ptr<geom::Mesh>> meshes;
            Object base;
            base.name = name;
            base.type = gtype;
            if (g->get("path")) {
                std::filesystem::path gp = std::filesystem::path(base_dir) / get_str(*g, "path");
                bool smooth = true;
                if (const toml::Value *sm = g->get("smooth")) {
                    if (sm->kind != toml::Value::BOOLEAN) throw ConfigError("Config: Type of smooth should be bool");
                    smooth = sm->b;
                }
                for (auto &m : geom::stl_meshes(gp.string(), tf)) {
                    m.is_smooth = smooth;
                    meshes.push_back(std::make_shared<geom::Mesh>(std::move(m)));
                }
                base.shape = "mesh";
            } else if (g->get("shape")) {
                const std::string shape = get_str(*g, "shape");
                base.shape = shape;
                auto v2 = [&](const char *k) {
                    const auto v = get_vec(*g, k, 2);
                    return vec2_t(v[0], v[1]);
                };
                geom::Mesh m;
                bool have = true;
                if (shape == "ellipse") {
                    m = geom::curve_mesh(geom::ellipse(get_float(*g, "a"), get_float(*g, "b")), 0.0, 2 * M_PI, true, tf);
                } else if (shape == "circle") {
                    m = geom::curve_mesh(geom::circle(get_float(*g, "radius")), 0.0, 2 * M_PI, true, tf);
                } else if (shape == "naca_airfoil") {
                    m = geom::curve_mesh(geom::naca4(get_float(*g, "camber"), get_float(*g, "position"),
                                                     get_float(*g, "thickness")), -1.0, 1.0, false, tf);
                } else if (shape == "Joukowski_airfoil") {
                    const double zero = 0.0;
                    m = geom::curve_mesh(geom::joukowski(get_float(*g, "camber"), get_float(*g, "thickness"),
                                                         get_float(*g, "trailing_edge_angle", &zero)),
                                         0.0, 2 * M_PI, true, tf);
                } else if (shape == "line") {
                    const double zero = 0.0, len = get_float(*g, "length", &zero);
                    m = len > 0.0 ? geom::polygon_mesh({vec2_t(-len / 2.0, 0), vec2_t(len / 2.0, 0)}, tf)
                                  : geom::polygon_mesh({v2("v1"), v2("v2")}, tf);
                } else if (shape == "diamond") {
                    const double a = get_float(*g, "width"), b = get_float(*g, "height");
                    m = geom::polygon_mesh({vec2_t(0, b / 2), vec2_t(a / 2, 0), vec2_t(0, -b / 2), vec2_t(-a / 2, 0)}, tf);
                } else if (shape == "quadrilateral") {
                    m = geom::polygon_mesh({v2("v1"), v2("v2"), v2("v3"), v2("v4")}, tf);
                } else if (shape == "triangle") {
                    m = geom::polygon_mesh({v2("v1"), v2("v2"), v2("v3")}, tf);
                } else if (shape == "sphere") {
                    have = false;
                    base.analytic = true;
                    base.radius = get_float(*g, "radius");
                    base.center = {tr[0], tr[1], tr[2]};
                } else {
                    throw ConfigError("Unknown shape type " + shape);
                }
                if (have) meshes.push_back(std::make_shared<geom::Mesh>(std::move(m)));
            } else {
                throw ConfigError("Geometry '" + name + "' needs to either have an associated file path or a shape.");
            }
            const toml::Value *gbc = bct ? bct->get(name) : nullptr;
            if (gtype == "solid" && !gbc)
                throw ConfigError("Must define boundary condition for solid geometry '" + name + "'");
            if (gbc) {
                base.has_bc = true;
                base.bc = parse_bc(cs, gbc);
            }
            if (base.analytic) {
                cs.objects.push_back(base);
            } else {
                for (auto &m : meshes) {
                    Object o = base;
                    o.mesh = m;
                    cs.objects.push_back(o);
                }
            }
        }
    }
    if (const toml::Value *zs = tbl.get("refinement_zone")) {
        for (const auto &z : zs->a) {
            const std::vector<double> s = get_vec(*z, "start", 3), e = get_vec(*z, "end", 3);
            cs.zones.push_back({{s[0], s[1], s[2]}, {e[0], e[1], e[2]}, (int) get_int(*z, "max_refinement_level", true)});
        }
    }
    return cs;
}

inline Case load_case_file(const std::string &path, int model, int dim, int limiter = DACTI_LIMITER_MINMOD) {
    const std::string dir = std::filesystem::absolute(path).parent_path().string();
    return load_case(*toml::parse_file(path), model, dim, limiter, dir);
}

// the C-ABI case structure (dacti_create)
inline dacti_case to_dacti_case(const Case &cs) {
    if ((int) cs.zones.size() > DACTI_MAX_ZONES) throw ConfigError("too many refinement zones");
    dacti_case c{};
    c.model = cs.model;
    c.dim = cs.dim;
    c.limiter = cs.limiter;
    c.max_refinement_level = cs.max_refinement_level;
    c.side_length = cs.side_length;
    for (int i = 0; i < 3; ++i) {
        c.active_size[i] = i < cs.dim ? cs.active_size[i] : 0.0;
        c.periodic[i] = i < cs.dim ? (int) cs.periodic[i] : 0;
    }
    c.n_zones = (int) cs.zones.size();
    c.max_cfl = cs.max_cfl;
    c.max_dt = cs.max_dt;
    c.min_error = cs.error_threshold[0];
    c.max_error = cs.error_threshold[1];
    for (size_t i = 0; i < cs.weights.size() && i < 8; ++i) c.weights[i] = cs.weights[i];
    const auto &k = cs.consts;
    c.gamma = k.at("gamma"), c.ns_mu = k.at("ns_mu"), c.ns_lambda = k.at("ns_lambda"), c.ns_cp = k.at("ns_cp");
    c.ns_R = k.at("ns_R"), c.sigma = k.at("sigma"), c.gamma1 = k.at("gamma1"), c.gamma2 = k.at("gamma2");
    c.P1 = k.at("P1"), c.P2 = k.at("P2"), c.cp1 = k.at("cp1"), c.cp2 = k.at("cp2"), c.cv1 = k.at("cv1");
    c.cv2 = k.at("cv2"), c.mu1 = k.at("mu1"), c.mu2 = k.at("mu2"), c.lambda1 = k.at("lambda1");
    c.lambda2 = k.at("lambda2");
    for (int f = 0; f < 6; ++f) {
        c.bc_type[f] = cs.bcs[f].type;
        for (size_t e = 0; e < cs.bcs[f].value.size() && e < 8; ++e) c.bc_p[f][e] = cs.bcs[f].value[e];
        c.bc_fade_in[f] = cs.bcs[f].fade_in;
    }
    for (size_t z = 0; z < cs.zones.size(); ++z) {
        for (int a = 0; a < 3; ++a) {
            c.zone_x0[z][a] = cs.zones[z].x0[a];
            c.zone_x1[z][a] = cs.zones[z].x1[a];
        }
        c.zone_level[z] = cs.zones[z].level;
    }
    return c;
}

// Initial tree (the layout of dacti_tree), owning its arrays.  Faces with an
// object boundary carry bc code DACTI_BC_OBJECT; their condition is the
// node's record brec (the nearest object boundary of the node's centre,
// scene/scene.cpp:11-28), shared by all the node's object faces.
struct Tree {
    int64_t n = 0;
    int dim = 2, num_eqns = 4;
    std::vector<int32_t> parent, child_base, brec;
    std::vector<int8_t> level;
    std::vector<uint8_t> flags, bc;
    std::vector<double> center, size, p;
    std::vector<dacti_boundary> records;

    dacti_tree view() const {
        return {n, parent.data(), child_base.data(), level.data(), flags.data(), bc.data(), center.data(), size.data(),
                p.data(), brec.data(), (int64_t) records.size(), records.data()};
    }
};

// Config::ic.x2p (config/config.cpp:194-243) at a cell centre
inline std::vector<double> initial_state(const Case &cs, const double *x) {
    if (cs.ic_type == "uniform") return cs.p_ref;  // "fixed" also returns p_ref (config.cpp:208-211)
    if (cs.ic_type == "jump") return x[cs.ic_axis] > cs.ic_location ? cs.ic_pos : cs.ic_neg;
    if (cs.ic_type == "grid") {
        // nearest grid cell centre (ties to even), clamped to the grid
        const int NE = cs.num_eqns();
        size_t flat = 0, total = 1;
        for (int d = 0; d < cs.dim; ++d) {
            const int n = cs.ic_grid_shape[d];
            if (n <= 0) throw ConfigError("grid initial condition without a grid");
            const double h = cs.active_size[d] / n;
            long i = std::lrint((x[d] + 0.5 * cs.active_size[d]) / h - 0.5);
            i = std::min<long>(std::max<long>(i, 0), n - 1);
            flat = flat * (size_t) n + (size_t) i;
            total *= (size_t) n;
        }
        if (cs.ic_grid.size() != total * (size_t) NE) throw ConfigError("grid initial condition of the wrong size");
        return std::vector<double>(cs.ic_grid.begin() + (long) (flat * NE), cs.ic_grid.begin() + (long) ((flat + 1) * NE));
    }
    throw ConfigError("unsupported initial condition " + cs.ic_type);
}

namespace detail {

// signed distance of x (dim components) to an object and, when asked, the
// nearest point's normal (unit, dim components)
inline double object_distance(const Object &o, const double *x, int dim, double *normal = nullptr,
                              double *abs_dist = nullptr) {
    if (o.analytic) {
        double r2 = 0.0;
        for (int d = 0; d < dim; ++d) {
            const double dx = x[d] - o.center[d];
            r2 += dx * dx;
        }
        const double r = std::sqrt(r2), sd = r - o.radius;
        if (normal)
            for (int d = 0; d < dim; ++d) normal[d] = r > 0 ? (x[d] - o.center[d]) / r : (d == 0 ? 1.0 : 0.0);
        if (abs_dist) *abs_dist = std::abs(sd);
        return sd;
    }
    vec3_t q;
    for (int d = 0; d < dim; ++d) q[d] = x[d];
    geom::Closest c;
    const double sd = geom::signed_distance(*o.mesh, q, c);
    if (normal) {
        const vec3_t n = geom::normal_at(*o.mesh, c);
        double l2 = 0;
        for (int d = 0; d < dim; ++d) l2 += n[d] * n[d];
        const double l = std::sqrt(l2);
        for (int d = 0; d < dim; ++d) normal[d] = n[d] / l;
    }
    if (abs_dist) *abs_dist = std::sqrt(c.sq);
    return sd;
}

// does the segment a -> b cross the object's surface?
inline bool object_ray(const Object &o, const double *a, const double *b, int dim) {
    if (o.analytic) {
        double da = 0, db = 0;
        for (int d = 0; d < dim; ++d) {
            da += (a[d] - o.center[d]) * (a[d] - o.center[d]);
            db += (b[d] - o.center[d]) * (b[d] - o.center[d]);
        }
        return (std::sqrt(da) - o.radius) * (std::sqrt(db) - o.radius) <= 0;
    }
    vec3_t A, B;
    for (int d = 0; d < dim; ++d) {
        A[d] = a[d];
        B[d] = b[d];
    }
    double t;
    return geom::intersect_segment(*o.mesh, A, B, t);
}

}  // namespace detail

// Scene<M>::init_domain (scene/scene.cpp:369-431): the tree grows level by
// level (signed distances, boundaries, split detection, splits in node
// order; children of a split occupy consecutive indices), then
// fix_boundary runs ten times (:250-335), cells outside the active box are
// deactivated and the active leaves get the initial state (with the
// interface of droplet and bubble objects, :340-366).  Where the reference
// reads neighbours' activity while other threads change it
// (compute_node_boundaries, :88-147), this reads the values from before the
// pass.
inline Tree build_tree(const Case &cs) {
    const int D = cs.dim, NE = cs.num_eqns(), NCH = 1 << D, NF = 2 * D;
    const bool has_obj = !cs.objects.empty();
    Tree T;
    T.dim = D;
    T.num_eqns = NE;
    std::vector<int32_t> &P = T.parent, &CB = T.child_base, &BR = T.brec;
    std::vector<int8_t> &LV = T.level;
    std::vector<double> CT((size_t) 2 * D, 0.0), SZ{0.0, cs.side_length}, SD{-1.0, 0.0};
    std::vector<char> LF{0, 1}, CU{0, 0}, ACT{0, 1}, ISB{0, 0};
    std::vector<int64_t> NR{0, 0};
    std::vector<uint8_t> BCA((size_t) 2 * NF, 0);
    P = {0, 0};
    CB = {0, 0};
    LV = {-1, 0};
    BR = {-1, -1};
    int64_t n = 2;

    // Domain::find (integrator/domain.hpp:65-82,407-417)
    auto find = [&](std::vector<double> x, int max_level) -> int64_t {
        for (int d = 0; d < D; ++d) {
            if (x[(size_t) d] < -0.5 * cs.active_size[d] || x[(size_t) d] > 0.5 * cs.active_size[d]) {
                if (!cs.periodic[d]) return 0;
                x[(size_t) d] += x[(size_t) d] < -0.5 * cs.active_size[d] ? cs.active_size[d] : -cs.active_size[d];
            }
        }
        for (int d = 0; d < D; ++d) x[(size_t) d] = 2.0 * x[(size_t) d] / SZ[1];
        int64_t idx = 1;
        while (!LF[(size_t) idx] && LV[(size_t) idx] != max_level) {
            int q = 0;
            for (int d = 0; d < D; ++d)
                if (x[(size_t) d] > 0.0) q |= 1 << d;
            for (int d = 0; d < D; ++d) x[(size_t) d] = (x[(size_t) d] - 0.5 * (((q >> d) & 1) ? 1.0 : -1.0)) * 2.0;
            idx = CB[(size_t) idx] + q;
        }
        return idx;
    };
    auto nb_center = [&](int64_t i, int f) {
        std::vector<double> x(&CT[(size_t) i * D], &CT[(size_t) i * D] + D);
        x[(size_t) (f >> 1)] += (2 * (f & 1) - 1) * SZ[(size_t) i];
        return x;
    };
    // Scene::raycast (:31-45): false at the first object without a boundary condition
    auto raycast = [&](int64_t i, const std::vector<double> &b) {
        for (const Object &o : cs.objects) {
            if (!o.has_bc) return false;
            if (detail::object_ray(o, &CT[(size_t) i * D], b.data(), D)) return true;
        }
        return false;
    };
    // Scene::nearest_boundary (:11-28): the condition of the nearest object
    // with one, with its unit normal there; a new record
    auto nearest_boundary = [&](int64_t i) {
        dacti_boundary r{};
        r.type = DACTI_BC_ZERO_GRADIENT;
        double best = std::numeric_limits<double>::max();
        for (const Object &o : cs.objects) {
            if (!o.has_bc) continue;
            double nrm[3] = {0, 0, 0}, dist = 0;
            detail::object_distance(o, &CT[(size_t) i * D], D, nrm, &dist);
            if (dist < best) {
                best = dist;
                r = dacti_boundary{};
                r.type = o.bc.type;
                for (int d = 0; d < D; ++d) r.normal[d] = nrm[d];
                for (size_t e = 0; e < o.bc.value.size() && e < 8; ++e) r.p[e] = o.bc.value[e];
                r.fade_in = o.bc.fade_in;
            }
        }
        T.records.push_back(r);
        return (int32_t) T.records.size() - 1;
    };
    // Domain::split (domain.hpp:192-247) of a leaf: 2^D consecutive children
    auto split = [&](int64_t par) {
        const int64_t base = n;
        CB[(size_t) par] = (int32_t) base;
        LF[(size_t) par] = 0;
        const double ps = SZ[(size_t) par], q = 0.25 * ps;
        for (int ch = 0; ch < NCH; ++ch) {
            P.push_back((int32_t) par);
            CB.push_back(0);
            LV.push_back((int8_t) (LV[(size_t) par] + 1));
            for (int d = 0; d < D; ++d) CT.push_back(CT[(size_t) par * D + d] + (((ch >> d) & 1) ? 1.0 : -1.0) * q);
            SZ.push_back(0.5 * ps);
            LF.push_back(1);
            CU.push_back(1);
            SD.push_back(0.0);
            NR.push_back(0);
            for (int f = 0; f < NF; ++f) BCA.push_back(0);
            ACT.push_back(1);
            ISB.push_back(0);
            BR.push_back(-1);
            ++n;
        }
        return base;
    };

    std::vector<int64_t> current{1};
    while (!current.empty()) {
        // compute_signed_distances (scene.cpp:47-85)
        for (int64_t i : current) {
            double sd = std::numeric_limits<double>::max();
            int64_t near = 0;
            for (size_t mi = 0; mi < cs.objects.size(); ++mi) {
                const double dist = detail::object_distance(cs.objects[mi], &CT[(size_t) i * D], D);
                if (dist < sd) {
                    sd = dist;
                    near = (int64_t) mi;
                }
            }
            SD[(size_t) i] = sd;
            NR[(size_t) i] = near;
            if (has_obj) ACT[(size_t) i] = sd < 0.0 ? !cs.objects[(size_t) near].has_bc : 1;
        }
        // compute_node_boundaries (scene.cpp:88-147)
        const std::vector<char> act0 = ACT;
        for (int64_t i : current) {
            bool open = false, obj = false;
            for (int f = 0; f < NF; ++f) {
                const int d = f >> 1, dir = f & 1;
                const std::vector<double> nc = nb_center(i, f);
                uint8_t code = 0;
                if (!cs.periodic[d] && nc[(size_t) d] * (2 * dir - 1) > cs.active_size[d] / 2.0) {
                    code = (uint8_t) (1 + f);
                } else {
                    // without objects every other face is open
                    const int64_t nb = has_obj ? find(nc, LV[(size_t) i]) : 0;
                    if (has_obj && (act0[(size_t) nb] != act0[(size_t) i] || raycast(i, nc))) {
                        code = DACTI_BC_OBJECT;
                        obj = true;
                    } else {
                        open = true;
                    }
                }
                BCA[(size_t) i * NF + f] = code;
            }
            if (!open) ACT[(size_t) i] = 0;
            if (obj) {
                CU[(size_t) i] = 0;
                ISB[(size_t) i] = 1;
                BR[(size_t) i] = nearest_boundary(i);
            }
        }
        // detect_splits (scene.cpp:150-214), statement order kept
        std::vector<int64_t> sp;
        for (int64_t i : current) {
            const double s = SZ[(size_t) i];
            const int lev = LV[(size_t) i];
            const double *c = &CT[(size_t) i * D];
            bool split_it = false;
            char &cu = CU[(size_t) i];
            if (lev < cs.min_refinement_level) {
                split_it = true;
                cu = 0;
            }
            if (has_obj && std::abs(SD[(size_t) i]) < s * cs.refinement_radius) {
                if (cs.objects[(size_t) NR[(size_t) i]].type == "solid") {
                    if (lev < cs.max_refinement_level_geom) {
                        split_it = true;
                        cu = 0;
                    }
                } else if (lev < cs.max_refinement_level) {
                    split_it = true;
                    cu = 1;
                }
            }
            for (int f = 0; f < NF; ++f) {
                const uint8_t code = BCA[(size_t) i * NF + f];
                const int type = code == DACTI_BC_OBJECT ? T.records[(size_t) BR[(size_t) i]].type
                                 : code ? cs.bcs[code - 1].type : -1;
                if (type == DACTI_BC_INFLOW_R && lev < cs.max_refinement_level) {
                    split_it = true;
                    cu = 1;
                }
            }
            if (cs.ic_type == "jump" && std::abs(c[cs.ic_axis] - cs.ic_location) < s && lev < cs.max_refinement_level) {
                split_it = true;
                cu = 1;
            }
            for (const Zone &z : cs.zones) {
                double x3[3] = {0, 0, 0};
                for (int d = 0; d < D; ++d) x3[d] = c[d];
                bool inside = true;
                for (int a = 0; a < 3; ++a) inside = inside && x3[a] < z.x1[a] && x3[a] > z.x0[a];
                if (inside && lev < z.level) split_it = true;
            }
            if (split_it && LF[(size_t) i]) sp.push_back(i);
        }
        // split_detected_nodes (scene.cpp:217-247), parents in node order
        std::sort(sp.begin(), sp.end());
        std::vector<int64_t> next;
        for (int64_t par : sp) {
            const int64_t base = split(par);
            for (int ch = 0; ch < NCH; ++ch) {
                CU[(size_t) (base + ch)] = CU[(size_t) par];
                next.push_back(base + ch);
            }
        }
        current.swap(next);
    }
    if (!has_obj)
        for (int64_t i = 1; i < n; ++i)
            if (LF[(size_t) i]) ACT[(size_t) i] = 1;

    // fix_boundary (scene.cpp:250-335), ten times
    for (int it = 0; it < 10 && has_obj; ++it) {
        std::vector<int64_t> fresh;  // new_leaves, in insertion order
        std::vector<char> is_fresh((size_t) n, 0);
        for (int64_t i = 0; i < n; ++i) {
            if (ACT[(size_t) i] || !LF[(size_t) i]) continue;
            for (int f = 0; f < NF; ++f) {
                const std::vector<double> nc = nb_center(i, f);
                while (true) {
                    const int64_t nb = find(nc, LV[(size_t) i]);
                    if (!(nb && ACT[(size_t) nb] && LV[(size_t) nb] < LV[(size_t) i])) break;
                    const int64_t base = split(nb);
                    is_fresh.resize((size_t) n, 0);
                    if (is_fresh[(size_t) nb]) {
                        is_fresh[(size_t) nb] = 0;
                        fresh.erase(std::find(fresh.begin(), fresh.end(), nb));
                    }
                    for (int ch = 0; ch < NCH; ++ch) {
                        const int64_t x = base + ch;
                        CU[(size_t) x] = CU[(size_t) nb];
                        for (int g = 0; g < NF; ++g) {
                            const int a = g >> 1, dir = g & 1;
                            if (((ch >> a) & 1) == dir) {
                                BCA[(size_t) x * NF + g] = BCA[(size_t) nb * NF + g];
                                if (BCA[(size_t) x * NF + g] == DACTI_BC_OBJECT) BR[(size_t) x] = BR[(size_t) nb];
                            }
                        }
                        ACT[(size_t) x] = 1;
                        is_fresh[(size_t) x] = 1;
                        fresh.push_back(x);
                    }
                }
            }
        }
        is_fresh.resize((size_t) n, 0);
        for (int64_t i : fresh) {
            int32_t rec = -1;
            for (int f = 0; f < NF; ++f) {
                BCA[(size_t) i * NF + f] = 0;
                const std::vector<double> nc = nb_center(i, f);
                const int64_t nb = find(nc, LV[(size_t) i]);
                if (ACT[(size_t) nb] != ACT[(size_t) i] || raycast(i, nc)) {
                    if (rec < 0) rec = nearest_boundary(i);
                    BCA[(size_t) i * NF + f] = DACTI_BC_OBJECT;
                }
            }
            BR[(size_t) i] = rec;
        }
        for (int64_t i = 0; i < n; ++i) {
            if (!(ACT[(size_t) i] && LF[(size_t) i])) continue;
            for (int f = 0; f < NF; ++f) {
                if (!BCA[(size_t) i * NF + f]) continue;
                const int64_t nb = find(nb_center(i, f), LV[(size_t) i]);
                if (nb && !BCA[(size_t) nb * NF + (f ^ 1)]) BCA[(size_t) i * NF + f] = 0;
            }
        }
    }

    // cells outside the active box (scene.cpp:407-415)
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < D; ++d) {
            const double c = CT[(size_t) i * D + d];
            if (c > 0.5 * cs.active_size[d] || c < -0.5 * cs.active_size[d]) ACT[(size_t) i] = 0;
        }
    ACT[0] = 0;
    T.n = n;
    T.center = CT;
    T.size = SZ;
    T.bc = BCA;
    T.p.assign((size_t) n * NE, 0.0);
    T.flags.assign((size_t) n, 0);
    for (int64_t i = 1; i < n; ++i) {
        if (ACT[(size_t) i] && LF[(size_t) i]) {
            const std::vector<double> v = initial_state(cs, &CT[(size_t) i * D]);
            for (int e = 0; e < NE; ++e) T.p[(size_t) i * NE + e] = v[(size_t) e];
            // Scene::init_interface (scene.cpp:340-366); diffuse_ic is never
            // read from the TOML (config.cpp:198-199), so the interface is sharp
            if (cs.model == DACTI_MODEL_MULTI && D > 1 && has_obj) {
                const Object &o = cs.objects[(size_t) NR[(size_t) i]];
                const double sd = SD[(size_t) i];
                if (o.type == "droplet" || o.type == "bubble") {
                    const double alpha = o.type == "droplet" ? (sd < 0 ? 1.0 : 0.0) : (sd > 0 ? 1.0 : 0.0);
                    T.p[(size_t) i * NE + 0] = cs.rho_ref[0] * alpha;
                    T.p[(size_t) i * NE + 1] = cs.rho_ref[1] * (1.0 - alpha);
                    T.p[(size_t) i * NE + NE - 1] = alpha;
                }
            }
        }
        T.flags[(size_t) i] = (uint8_t) ((LF[(size_t) i] ? DACTI_F_LEAF : 0) | (ACT[(size_t) i] ? DACTI_F_ACTIVE : 0) |
                                         DACTI_F_CAN_SPLIT | (CU[(size_t) i] ? DACTI_F_CAN_UNSPLIT : 0) |
                                         (ISB[(size_t) i] ? DACTI_F_BOUNDARY : 0));
    }
    return T;
}

}  // namespace dacti

#endif
