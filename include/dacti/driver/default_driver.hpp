// dacti/driver/default_driver.hpp -- driver::default_driver, the run loop.
//
// default_driver{integrator, writer, config, std::tuple{observers...}}.run()
// keeps the loop contract of driver/default_driver.hpp:77-217 of the
// reference:
//  * <result_out_dir>/<case_name>_observables.csv starts with the header
//    "time,frame,step,n_cells" followed by every observer's integral
//    variable names; the initial observation writes no row, frame 0 is
//    written with the initial state;
//  * while time < end_time and not converged: one global step
//    (signal = time > signal_time), time += global_dt;
//  * with steady_tol > 0, every step appends a CSV row (time, frame, step,
//    cells, integral variables) and stops once max_i |dI_i| / dt < tol;
//  * a frame (CSV row with the integral variables, the writer's file with
//    the field variables) is written when time - last frame time >=
//    print_interval, on convergence and on a NaN state (force_break), which
//    then ends the run with the message "Forcing break due to nan";
//  * step counts from 0 and is incremented after the frame check.
// The step runs without a per-substep hook (the reference passes a no-op).
// Progress is logged to stdout.  Observers take (integrator, config,
// observe_field_vars) as in driver/observer/*.hpp.
#ifndef DACTI_DEFAULT_DRIVER_HPP
#define DACTI_DEFAULT_DRIVER_HPP

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../config/config.hpp"

namespace dacti::driver {

template <typename Integrator, typename Model, typename Writer, typename... Observers>
class default_driver {
    Integrator integrator;
    Writer writer;
    std::shared_ptr<config::Config<Model>> config;
    std::tuple<Observers...> observers;
    std::vector<std::vector<scalar_t>> all_FieldVars;
    std::vector<scalar_t> all_IntegVars;
    std::vector<std::string> all_FieldVar_names, all_IntegVar_names;

public:
    default_driver(const Integrator &integrator, const Writer &writer, std::shared_ptr<config::Config<Model>> config,
                   std::tuple<Observers...> observers)
        : integrator(integrator), writer(writer), config(config), observers(observers) {}

    // run summary, for callers and tests
    struct Summary {
        double time = 0;
        int frames = 0, steps = 0;
        bool converged = false, nan_break = false;
    };
    Summary summary;

    void observe_all_variables(std::ofstream &csv, bool init_observation = true, bool observe_FieldVars = true) {
        std::apply(
            [&](const auto &...obs) {
                (observe_one(obs, csv, init_observation, observe_FieldVars), ...);
            },
            observers);
    }

    void run() {
        info("Starting simulation \"" + config->case_name + "\"");
        const auto start = std::chrono::steady_clock::now();
        const scalar_t end_time = config->end_time, print_interval = config->io_interval;
        double time = 0.0, last_print_time = 0.0;
        int frame_index = 0, step_index = 0, last_print_step = 0;
        bool converged = false, force_break = false;

        const std::string csv_path = config->result_out_dir + "/" + config->case_name + "_observables.csv";
        std::ofstream csv(csv_path);
        if (!csv) throw std::runtime_error("default_driver: cannot open " + csv_path);
        csv << "time,frame,step,n_cells";
        observe_all_variables(csv);
        csv << "\n";
        writer.write_data(integrator, 0.0, 0, all_FieldVars, all_FieldVar_names);

        auto try_write = [&](scalar_t acti_time, scalar_t dt) {
            if (config->convergence_tol > 0) {
                if (all_IntegVar_names.empty())
                    throw std::runtime_error("default_driver: the convergence check needs an observed variable");
                csv << time << "," << frame_index << "," << step_index << "," << integrator.n_active_cells();
                const std::vector<scalar_t> old = all_IntegVars;
                all_IntegVars.clear();
                observe_all_variables(csv, false, false);
                csv << "\n";
                csv.flush();
                scalar_t max_delta = 0.0;
                for (size_t i = 0; i < all_IntegVars.size(); ++i)
                    max_delta = std::max(max_delta, std::abs(all_IntegVars[i] - old[i]) / dt);
                char buf[128];
                std::snprintf(buf, sizeof buf, "Max Delta: %.2e (tol: %.2e)", max_delta, config->convergence_tol);
                info(buf);
                if (max_delta < config->convergence_tol) {
                    info("Converged!");
                    converged = true;
                }
            }
            if (acti_time - last_print_time >= print_interval || converged || force_break) {
                ++frame_index;
                const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
                char buf[256];
                std::snprintf(buf, sizeof buf,
                              "[Case \"%s\" | Frame %d] steps %d, sim. time %.2e / %.2e (%.1f%%), cells %zu, "
                              "global dt %.2e, wall %.2fs",
                              config->case_name.c_str(), frame_index, step_index - last_print_step, time, end_time,
                              100.0 * time / end_time, (size_t) integrator.n_active_cells(), integrator.global_dt,
                              secs);
                info(buf);
                all_FieldVars.clear();
                csv << time << "," << frame_index << "," << step_index << "," << integrator.n_active_cells();
                observe_all_variables(csv, false);
                csv << "\n";
                csv.flush();
                last_print_time = acti_time;
                last_print_step = step_index;
                writer.write_data(integrator, time, frame_index, all_FieldVars, all_FieldVar_names);
            }
        };

        while (time < end_time && !converged) {
            const scalar_t global_dt = integrator.integrate(time > config->signal_time, force_break);
            time += global_dt;
            try_write(time, global_dt);
            if (force_break) {
                std::fprintf(stderr, "[error] Forcing break due to nan\n");
                break;
            }
            ++step_index;
        }
        summary = {time, frame_index, step_index, converged, force_break};
    }

    const Integrator &get_integrator() const { return integrator; }

private:
    static void info(const std::string &m) {
        std::printf("[info] %s\n", m.c_str());
        std::fflush(stdout);
    }

    template <typename Obs>
    void observe_one(const Obs &obs, std::ofstream &csv, bool init_observation, bool observe_FieldVars) {
        if (init_observation) {
            for (const auto &n : obs.FieldVar_names()) all_FieldVar_names.push_back(obs.name() + "/" + n);
            for (const auto &n : obs.IntegVar_names()) {
                all_IntegVar_names.push_back(n);
                csv << "," << n;
            }
        }
        obs.observe(integrator, config, observe_FieldVars);
        const auto vals = obs.IntegVars();
        for (size_t i = 0; i < obs.NumIntegVars(); ++i) {
            all_IntegVars.push_back(vals[i]);
            if (!init_observation) csv << "," << vals[i];
        }
        if (observe_FieldVars)
            for (auto &f : obs.FieldVars()) all_FieldVars.push_back(std::move(f));
    }
};

template <typename Integrator, typename Model, typename Writer, typename... Observers>
default_driver(const Integrator &, const Writer &, std::shared_ptr<config::Config<Model>>, std::tuple<Observers...>)
    -> default_driver<Integrator, Model, Writer, Observers...>;

}  // namespace dacti::driver

#endif
