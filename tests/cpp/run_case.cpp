// A case run through the reference-shaped C++ interface on the GPU:
// config::Config, scene::Scene, integrator::Integrator (class template
// argument deduction from the limiter and estimator tags), and either
//   driver  -- driver::default_driver with io::vtu_writer and the
//              observers of the model (DropEnergy for multi<2>, AeroForce
//              for euler<2>), or
//   hook    -- `steps` global steps of integrate(hook, signal, force_break),
//              printing one line per hook call ("hook level acti_time
//              n_substeps dt n_active") and the counters after each step.
// usage: run_case <case.toml> <multi2|euler2|ns3> <driver|hook> [steps]
#include <cstdio>
#include <cstdlib>
#include <string>
#include <tuple>

#include "config/config.hpp"
#include "driver/default_driver.hpp"
#include "driver/io/vtu_writer.hpp"
#include "driver/observer/aero_force.hpp"
#include "driver/observer/drop_energy.hpp"
#include "integrator/error_estimators.hpp"
#include "integrator/integrator.hpp"
#include "integrator/slope_limiters.hpp"
#include "model/euler.hpp"
#include "model/multi.hpp"
#include "model/navier_stokes.hpp"
#include "scene/scene.hpp"

using namespace dacti;

template <typename Model, typename Observers>
int run(const std::string &path, const std::string &mode, int steps, Observers obs) {
    auto config = std::make_shared<config::Config<Model>>(path);
    scene::Scene<Model> scene(config);
    integrator::Integrator integrator{integrator::slope::minmod{},
                                      integrator::error::weighted_laplacian{config->error_weights}, scene, config};
    std::printf("created: %zu active cells, global dt %.17g\n", integrator.n_active_cells(), integrator.global_dt);
    std::fflush(stdout);
    if (mode == "driver") {
        io::vtu_writer writer(config->case_name, config->result_out_dir);
        auto driver = driver::default_driver{integrator, writer, config, obs};
        driver.run();
        return driver.summary.nan_break ? 3 : 0;
    }
    for (int s = 0; s < steps; ++s) {
        bool brk = false;
        auto hook = [](auto &integ, int level, scalar_t acti_time, int n_substeps, scalar_t dt) {
            std::printf("hook %d %.17g %d %.17g %zu\n", level, acti_time, n_substeps, dt, integ.n_active_cells());
        };
        const scalar_t dt = integrator.integrate(hook, true, brk);
        std::printf("step %d dt %.17g substeps %d global_steps %d max_level %d updates %llu cells %zu\n", s, dt,
                    integrator.n_substeps, integrator.n_global_steps, integrator.max_acti_level,
                    integrator.total_cell_updates, integrator.n_active_cells());
        if (brk) return 3;
    }
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const std::string model = argv[2], mode = argv[3];
    const int steps = argc > 4 ? std::atoi(argv[4]) : 2;
    try {
        if (model == "multi2") return run<model::multi<2>>(argv[1], mode, steps, std::tuple{observer::DropEnergy()});
        if (model == "euler2") return run<model::euler<2>>(argv[1], mode, steps, std::tuple{observer::AeroForce()});
        if (model == "ns3") return run<model::navier_stokes<3>>(argv[1], mode, steps, std::tuple<>{});
    } catch (const std::exception &e) {
        std::printf("error: %s\n", e.what());
        return 4;
    }
    return 2;
}
