// observer::AeroForce over a host stand-in for the integrator: a uniform
// grid of n x n cells over [-1, 1] x [-0.5, 0.5] with the pressure
// p_ref + a x + b y at the cell centres.  Prints "Cl Cd samples".
// usage: aero_mock <case.toml> <out_root> <a> <b> <n>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <vector>

#include "config/config.hpp"
#include "driver/observer/aero_force.hpp"
#include "integrator/integrator.hpp"

using namespace dacti;
using M = model::euler<2>;

struct Pool {
    std::vector<double> p;
};

struct MockDomain {
    int max_depth = 16, n = 0;
    Pool P;
    size_t find(vec2_t x, int) const {
        const int i = std::min(n - 1, std::max(0, (int) ((x[0] + 1.0) / 2.0 * n)));
        const int j = std::min(n / 2 - 1, std::max(0, (int) ((x[1] + 0.5) * n / 2.0)));
        return (size_t) (j * n + i);
    }
    const Pool &pool() const { return P; }
};

struct MockIntegrator {
    MockDomain dom;
    MockDomain get_domain() const { return dom; }
    size_t n_active_cells() const { return 0; }
    template <typename F>
    void for_each_active_cell(F &&) const {}
};

int main(int argc, char **argv) {
    if (argc != 6) return 2;
    std::ifstream in(argv[1]);
    std::stringstream ss;
    ss << in.rdbuf();
    auto config = config::Config<M>::from_string(ss.str(), argv[2]);
    const double a = std::atof(argv[3]), b = std::atof(argv[4]);
    MockIntegrator integ;
    const int n = std::atoi(argv[5]);
    integ.dom.n = n;
    integ.dom.P.p.assign((size_t) n * n / 2 * M::NumEqns, 0.0);
    const double h = 2.0 / n;
    for (int j = 0; j < n / 2; ++j)
        for (int i = 0; i < n; ++i) {
            const double x = -1.0 + (i + 0.5) * h, y = -0.5 + (j + 0.5) * h;
            double *p = &integ.dom.P.p[((size_t) j * n + i) * M::NumEqns];
            p[0] = 1.0;
            p[M::ip] = config->p_ref[M::ip] + a * x + b * y;
        }
    observer::AeroForce obs;
    obs.observe(integ, config, false);
    const auto v = obs.IntegVars();
    std::printf("%.17g %.17g\n", v[0], v[1]);
    return 0;
}
