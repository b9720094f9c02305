// driver::default_driver, io::vtu_writer and observer::DropEnergy over a
// host-side stand-in for the integrator (no GPU): a fixed set of 2D
// five-equation cells whose velocity is scaled by `decay` every step and
// whose state turns NaN at step `nan_step` (force_break).  Writes the
// observables CSV and the frames under <out_root>/out/<case name>/ and
// prints one line per integrate() call: "step <signal> <time>".
// usage: driver_mock <case.toml> <out_root> <decay> <nan_step> <dt>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "config/config.hpp"
#include "driver/default_driver.hpp"
#include "driver/io/vtu_writer.hpp"
#include "driver/observer/drop_energy.hpp"
#include "integrator/integrator.hpp"

using namespace dacti;
using M = model::multi<2>;

struct MockIntegrator {
    using model_type = M;
    using node_t = integrator::_internal::Node<M>;
    static constexpr size_t NumDims = 2, NumMeta = 4;
    static constexpr int MAX_LEVEL = DACTI_MAX_LEVEL;
    static constexpr std::array<const char *, NumMeta> METADATA_NAMES = {"acti_level", "error", "max_dt", "cfl"};

    std::vector<node_t> cells;
    double decay = 0.5, dt = 0.1, global_dt = 0.1;
    int step = 0, nan_step = -1;

    MockIntegrator(double decay, int nan_step, double dt) : decay(decay), dt(dt), global_dt(dt), nan_step(nan_step) {
        // a 4 x 2 patch of level-2 cells and two level-3 cells
        for (int k = 0; k < 10; ++k) {
            node_t n;
            auto &c = n.cell;
            c.size = k < 8 ? 0.25 : 0.125;
            c.center = k < 8 ? vec2_t(-0.375 + 0.25 * (k % 4), -0.125 + 0.25 * (k / 4)) : vec2_t(0.3125 + 0.125 * (k - 8), 0.4375);
            c.p[0] = 0.8 + 0.01 * k;   // alpha1 rho1
            c.p[1] = 0.1 * (k % 3);    // alpha2 rho2
            c.p[2] = 0.3 + 0.02 * k;   // u
            c.p[3] = -0.1 * (k % 2);   // v
            c.p[4] = 1.0 + 0.05 * k;   // p
            c.p[5] = (k % 4) / 3.0;    // alpha
            for (size_t e = 0; e < M::NumEqns; ++e) {
                c.slope[0][e] = 0.01 * (double) (e + 1) * (k % 2 ? 1 : -1);
                c.slope[1][e] = 0.002 * (double) e;
            }
            for (size_t q = 0; q < M::NumParm; ++q) c.k[q] = 0.1 * (double) q + k;
            c.gradient[0][M::da] = 0.5 * k;
            c.gradient[1][M::da] = -0.25;
            n.acti_level = k % 3;
            n.error = 1e-3 * k;
            n.max_dt = 0.2 + k;
            n.cfl = 0.01 * k;
            n.is_leaf = n.is_active = true;
            cells.push_back(n);
        }
    }

    double integrate(bool signal, bool &force_break) {
        std::printf("step %d %.17g\n", signal ? 1 : 0, (double) step);
        ++step;
        for (auto &n : cells) {
            n.cell.p[2] *= decay;
            n.cell.p[3] *= decay;
            n.cell.t += dt;
        }
        if (step == nan_step) {
            cells[3].cell.p[4] = NAN;
            force_break = true;
        }
        return dt;
    }
    size_t n_active_cells() const { return cells.size(); }
    template <typename F>
    void for_each_active_cell(F &&f) const {
        int i = 0;
        for (const auto &n : cells) f(i++, n);
    }
};

int main(int argc, char **argv) {
    if (argc != 6) return 2;
    std::ifstream in(argv[1]);
    std::stringstream ss;
    ss << in.rdbuf();
    auto config = config::Config<M>::from_string(ss.str(), argv[2]);
    MockIntegrator integ(std::atof(argv[3]), std::atoi(argv[4]), std::atof(argv[5]));
    io::vtu_writer writer(config->case_name, config->result_out_dir);
    auto driver = driver::default_driver{integ, writer, config, std::tuple{observer::DropEnergy()}};
    driver.run();
    std::printf("summary %.17g %d %d %d %d\n", driver.summary.time, driver.summary.frames, driver.summary.steps,
                (int) driver.summary.converged, (int) driver.summary.nan_break);
    return 0;
}
