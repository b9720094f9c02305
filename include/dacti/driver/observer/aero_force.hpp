// dacti/driver/observer/aero_force.hpp -- AeroForce: lift and drag
// coefficients from the pressure on object outlines, 2D
// (driver/observer/aero_force.hpp:8-131 of the reference).
//
// Sample points: every outline segment a -> b of every object (vertex
// rows 2i and 2i+2 of the extruded outline, wrapping) is cut into
// max(int(10 |b - a|), 1) pieces; each piece's midpoint, moved 0.01 against
// the segment's normal n = rot-90(b - a) / |b - a|, samples the pressure of
// the leaf that contains it.  Cp = (p - p_ref) 2 / (rho_ref |u_ref|^2),
// Cl = sum Cp l n_y, Cd = sum Cp l n_x.  Field variables: Cp on cells that
// carry an object boundary (0 elsewhere) and the boundary flag, one value
// per active cell (the reference's field arrays also keep n_active_cells
// leading zeros from a resize before the push_back, :110-113; not kept).
#ifndef DACTI_OBSERVER_AERO_FORCE_HPP
#define DACTI_OBSERVER_AERO_FORCE_HPP

#include <algorithm>
#include <memory>
#include <vector>

#include "../../config/config.hpp"
#include "common.hpp"

namespace dacti::observer {

class AeroForce : public Observer {
    mutable int step = 0;
    mutable bool mesh_initialized = false;
    mutable int n_samples = -1;
    mutable std::vector<vec2_t> mesh_line_centers;
    mutable std::vector<scalar_t> mesh_areas;
    mutable std::vector<vec2_t> mesh_normals;

public:
    AeroForce() : Observer("Aerodynamic Forces", {"Cl", "Cd"}, {"Cp", "ObjectSurface"}) {}

    template <typename ConfigPtr>
    void ensure_sample_points(const ConfigPtr &config, int sample_resolution = 10) const {
        if (mesh_initialized) return;
        n_samples = 0;
        for (const auto &obj : config->objects) {
            const auto &V = obj.mesh.V;
            const size_t N = V.rows();
            for (size_t i = 0; i < N; i += 2) {
                const vec2_t a(V(i, 0), V(i, 1));
                const vec2_t b(V((i + 2) % N, 0), V((i + 2) % N, 1));
                const vec2_t ab = b - a;
                const scalar_t len = ab.norm();
                const vec2_t nrm = (ab / len).rotaten90();
                const int k = std::max(static_cast<int>(sample_resolution * len), 1);
                for (int j = 0; j < k; ++j) {
                    vec2_t x = a + (j + 0.5) * ab / (scalar_t) k;
                    x -= nrm * 0.01;
                    mesh_line_centers.push_back(x);
                    mesh_areas.push_back(len / k);
                    mesh_normals.push_back(nrm);
                }
                n_samples += k;
            }
        }
        mesh_initialized = true;
    }

    template <typename Integrator, typename Model>
    void observe(const Integrator &integrator, const std::shared_ptr<config::Config<Model>> config,
                 const bool observe_FieldVars = true) const {
        static_assert(Model::NumDims == 2, "AeroForce samples 2D outlines");
        ensure_sample_points(config);
        const scalar_t rho_ref = config->p_ref[0], p_ref = config->p_ref[Model::ip];
        const scalar_t q_r = 2.0 / (rho_ref * config->vel_ref.squaredNorm());
        const auto dom = integrator.get_domain();
        scalar_t Cl = 0, Cd = 0;
        for (int i = 0; i < n_samples; ++i) {
            const size_t idx = dom.find(mesh_line_centers[(size_t) i], dom.max_depth);
            const scalar_t cp = (dom.pool().p[idx * Model::NumEqns + Model::ip] - p_ref) * q_r;
            Cl += cp * mesh_areas[(size_t) i] * mesh_normals[(size_t) i][1];
            Cd += cp * mesh_areas[(size_t) i] * mesh_normals[(size_t) i][0];
        }
        ++step;
        _IntegVars = {Cl, Cd};
        if (!observe_FieldVars) return;
        for (auto &f : _FieldVars) f.clear();
        integrator.for_each_active_cell([&](int, const auto &node) {
            const bool surf = node.is_boundary;
            _FieldVars[0].push_back(surf ? (node.cell.p[Model::ip] - p_ref) / (0.5 * rho_ref * config->vel_ref.squaredNorm())
                                         : 0.0);
            _FieldVars[1].push_back(surf ? 1.0 : 0.0);
        });
    }
};

}  // namespace dacti::observer

#endif
