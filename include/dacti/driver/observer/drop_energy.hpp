// dacti/driver/observer/drop_energy.hpp -- DropEnergy: kinetic energy and
// kinetic plus surface energy of a two-phase run
// (driver/observer/drop_energy.hpp:8-34 of the reference).
//
//   ke   = sum over active cells of vol * rho |u|^2 / 2
//   Etot = sum of vol * (rho |u|^2 / 2 + sigma |d alpha|)
//
// with rho = alpha1 rho1 + alpha2 rho2 (the two partial densities), and
// |d alpha| the length of the first two components of the volume-fraction
// gradient, as in the reference (also in 3D).
#ifndef DACTI_OBSERVER_DROP_ENERGY_HPP
#define DACTI_OBSERVER_DROP_ENERGY_HPP

#include <cmath>
#include <memory>

#include "../../config/config.hpp"
#include "common.hpp"

namespace dacti::observer {

class DropEnergy : public Observer {
public:
    DropEnergy() : Observer("Droplet Energy", {"ke", "Etot"}) {}

    template <typename Integrator, typename Model>
    void observe(const Integrator &integrator, const std::shared_ptr<config::Config<Model>> config,
                 const bool observe_FieldVars = true) const {
        (void) config;
        (void) observe_FieldVars;
        scalar_t ke = 0, Etot = 0;
        integrator.for_each_active_cell([&](int, const auto &node) {
            const auto &c = node.cell;
            const scalar_t vol = c.volume();
            const scalar_t rho = c.get_p()[0] + c.get_p()[1];
            const scalar_t kin = 0.5 * rho * c.get_velocity().squaredNorm();
            const auto g = c.get_gradient();
            const scalar_t ga = std::sqrt(g[0][Model::da] * g[0][Model::da] + g[1][Model::da] * g[1][Model::da]);
            ke += vol * kin;
            Etot += vol * (kin + Model::sigma * ga);
        });
        _IntegVars = {ke, Etot};
    }
};

}  // namespace dacti::observer

#endif
