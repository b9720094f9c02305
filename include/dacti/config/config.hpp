// dacti/config/config.hpp -- config::Config<M>, the case of a run.
//
// Config<M>(path) reads the TOML case file (config/config.hpp:91-103 and
// config/config.cpp:10-141 of the reference) and exposes the same public
// fields: case_name, active_size, side_length, io_interval, end_time,
// signal_time, convergence_tol, max_cfl, max_dt, refinement levels,
// error_threshold, error_weights, p_ref, vel_ref, rho_ref, IsPeriodic,
// visualize_mesh, the output directory result_out_dir = <project>/out/<name>
// (created, as the reference does) and the refinement zones.  Loading also
// sets the model's case constants (model::multi<D>::sigma, gamma1, ...,
// model::navier_stokes<D>::mu_, lambda_), which the reference keeps as
// process-global statics; each GPU handle takes its own copy at creation.
// The parsed case (dacti::Case) stays available as `cs`; it is what the
// scene and the integrator hand to the C ABI.
#ifndef DACTI_CONFIG_HPP
#define DACTI_CONFIG_HPP

#include <array>
#include <filesystem>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "../model/common.hpp"
#include "case.hpp"
#include "toml.hpp"

namespace dacti {

// the vertices of an object's mesh, the part of ObjMesh the observers read
// (a 2D outline extruded to z = -1, +1: rows 2i and 2i+1 share x and y;
// empty for an analytic sphere)
struct ObjMesh {
    struct Vertices {
        std::vector<vec3_t> rows_;
        size_t rows() const { return rows_.size(); }
        scalar_t operator()(size_t i, size_t j) const { return rows_[i][j]; }
    } V;
};

namespace config {

template <typename M>
class Config {
public:
    using vecn_t = typename M::vecn_t;
    using prim_t = typename M::prim_t;

    std::filesystem::path config_file_path, config_file_dir, prj_dir;
    std::string result_out_dir;

    std::string case_name;
    bool visualize_mesh = false;
    vecn_t active_size;
    scalar_t side_length = 1;

    scalar_t io_interval = 1, end_time = 1, signal_time = -1, convergence_tol = -1, max_cfl = 0.5, max_dt = 1;

    int min_refinement_level = 0, max_refinement_level = 0, max_refinement_level_geom = 0;
    scalar_t refinement_radius = 2;
    vec2_t error_threshold;
    prim_t error_weights;

    prim_t p_ref;
    vecn_t vel_ref;
    std::optional<vec2_t> rho_ref;

    std::array<bool, M::NumDims> IsPeriodic{};

    struct RefineZone {
        vec3_t x0, x1;
        size_t refinement_level;
        bool contains(const vec3_t &x) const {
            for (int a = 0; a < 3; ++a)
                if (!(x[a] < x1[a] && x[a] > x0[a])) return false;
            return true;
        }
    };
    std::vector<RefineZone> refine_zones;

    struct Object {
        ObjMesh mesh;
        std::string type;
        bool has_bc = false;
    };
    std::vector<Object> objects;

    Case cs;  // the parsed case (C-ABI fields, initial condition, geometry)

    explicit Config(const std::string &config_file) {
        config_file_path = std::filesystem::canonical(config_file);
        config_file_dir = config_file_path.parent_path();
        prj_dir = config_file_dir.parent_path();
        load(*toml::parse_file(config_file_path.string()));
    }

    // a case given as TOML text; results go to <out_root>/out/<name>
    static std::shared_ptr<Config> from_string(const std::string &text, const std::string &out_root) {
        auto c = std::shared_ptr<Config>(new Config());
        c->prj_dir = std::filesystem::absolute(out_root);
        c->config_file_dir = c->prj_dir;
        c->load(*toml::parse(text));
        return c;
    }

    // the parsed case with the public fields above, which a caller may have
    // changed after loading, written back
    Case effective_case() const {
        Case c = cs;
        for (size_t d = 0; d < M::NumDims; ++d) {
            c.active_size[d] = active_size[d];
            c.periodic[d] = IsPeriodic[d];
        }
        c.side_length = side_length;
        c.print_interval = io_interval;
        c.end_time = end_time;
        c.signal_time = signal_time;
        c.steady_tol = convergence_tol;
        c.max_cfl = max_cfl;
        c.max_dt = max_dt;
        c.min_refinement_level = min_refinement_level;
        c.max_refinement_level = max_refinement_level;
        c.max_refinement_level_geom = max_refinement_level_geom;
        c.refinement_radius = refinement_radius;
        c.error_threshold = {error_threshold[0], error_threshold[1]};
        c.weights.assign(error_weights.begin(), error_weights.end());
        c.zones.clear();
        for (const auto &z : refine_zones)
            c.zones.push_back({{z.x0[0], z.x0[1], z.x0[2]}, {z.x1[0], z.x1[1], z.x1[2]}, (int) z.refinement_level});
        return c;
    }

private:
    Config() = default;

    void load(const toml::Value &tbl) {
        cs = load_case(tbl, M::id, (int) M::NumDims);
        case_name = cs.name;
        if (const toml::Value *v = tbl.at_path("scene.visualize_mesh"))
            if (v->kind == toml::Value::BOOLEAN) visualize_mesh = v->b;
        for (size_t d = 0; d < M::NumDims; ++d) {
            active_size[d] = cs.active_size[d];
            IsPeriodic[d] = cs.periodic[d];
        }
        side_length = cs.side_length;
        io_interval = cs.print_interval;
        end_time = cs.end_time;
        signal_time = cs.signal_time;
        convergence_tol = cs.steady_tol;
        max_cfl = cs.max_cfl;
        max_dt = cs.max_dt;
        min_refinement_level = cs.min_refinement_level;
        max_refinement_level = cs.max_refinement_level;
        max_refinement_level_geom = cs.max_refinement_level_geom;
        refinement_radius = cs.refinement_radius;
        error_threshold = vec2_t(cs.error_threshold[0], cs.error_threshold[1]);
        for (size_t e = 0; e < M::NumEqns; ++e) {
            error_weights[e] = cs.weights[e];
            p_ref[e] = cs.p_ref[e];
        }
        for (size_t d = 0; d < M::NumDims; ++d) vel_ref[d] = cs.vel_ref[d];
        if (cs.has_rho_ref) rho_ref = vec2_t(cs.rho_ref[0], cs.rho_ref[1]);
        for (const Zone &z : cs.zones)
            refine_zones.push_back({vec3_t(z.x0[0], z.x0[1], z.x0[2]), vec3_t(z.x1[0], z.x1[1], z.x1[2]),
                                    (size_t) z.level});
        for (const auto &o : cs.objects) {
            ObjMesh m;
            if (o.mesh) m.V.rows_ = o.mesh->V;
            objects.push_back({m, o.type, o.has_bc});
        }
        set_model_constants();
        // output directory (config/config.cpp:28-36)
        result_out_dir = (prj_dir / "out" / case_name).string();
        std::filesystem::create_directories(result_out_dir);
    }

    void set_model_constants() {
        const auto &k = cs.consts;
        if constexpr (M::id == DACTI_MODEL_MULTI) {
            M::sigma = k.at("sigma"), M::gamma1 = k.at("gamma1"), M::gamma2 = k.at("gamma2");
            M::P1 = k.at("P1"), M::P2 = k.at("P2"), M::cp1 = k.at("cp1"), M::cp2 = k.at("cp2");
            M::cv1 = k.at("cv1"), M::cv2 = k.at("cv2"), M::mu1 = k.at("mu1"), M::mu2 = k.at("mu2");
            M::lambda1 = k.at("lambda1"), M::lambda2 = k.at("lambda2");
        } else if constexpr (M::id == DACTI_MODEL_NAVIER_STOKES) {
            M::mu_ = k.at("ns_mu"), M::lambda_ = k.at("ns_lambda");
        }
    }
};

}  // namespace config
}  // namespace dacti

#endif
