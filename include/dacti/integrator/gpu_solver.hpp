// dacti/integrator/gpu_solver.hpp -- untyped C++ owner of one GPU handle.
//
// GpuSolver wraps the C ABI (dacti_gpu.h): construction runs the
// integrator's initialisation (integrator/integrator.hpp:72-163 of the
// reference) on the GPU, integrate() is one global step
// (impl/integrate.hpp:11-87), get_nodes() copies the node pool to the host.
// Model, limiter and estimator weights come from the dacti_case fields.
// dacti::integrator::Integrator<Model, Limiter, ErrorEstimator>
// (integrator/integrator.hpp in this directory) is the reference-shaped
// interface on top of it.  Errors throw dacti::Error; a NaN state is
// reported through force_break as in impl/cell_update.hpp:52-57.
#ifndef DACTI_GPU_SOLVER_HPP
#define DACTI_GPU_SOLVER_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../dacti_gpu.h"

namespace dacti {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// Sizes of a model variant (Model::NumEqns etc.).
struct ModelSizes {
    int dim, num_eqns, num_flux, num_parm, num_grad;
};

inline ModelSizes model_sizes(int model, int dim) {
    int32_t s[5];
    int rc = dacti_model_sizes(model, dim, s);
    if (rc != DACTI_OK) throw Error(rc, dacti_last_error(nullptr));
    return {s[0], s[1], s[2], s[3], s[4]};
}

// Host copy of the node pool, node-major (Integrator::get_nodes()).
struct NodePool {
    int64_t n = 0;
    ModelSizes sz{};
    std::vector<int32_t> parent, child_base, nbr;
    std::vector<int8_t> tree_level, acti_level, cluster;
    std::vector<uint8_t> flags;
    std::vector<double> center, size, t, p, u, slope, rhs, k, grad, meta;

    bool is_leaf(int64_t i) const { return flags[i] & DACTI_F_LEAF; }
    bool is_active(int64_t i) const { return flags[i] & DACTI_F_ACTIVE; }
    const double *prim(int64_t i) const { return &p[i * sz.num_eqns]; }
    const double *cons(int64_t i) const { return &u[i * sz.num_eqns]; }
};

class GpuSolver {
public:
    static constexpr int NumMeta = 4;  // acti_level, error, max_dt, cfl (integrator.hpp:30)
    static constexpr int MAX_LEVEL = DACTI_MAX_LEVEL;

    GpuSolver(const dacti_case &c, const dacti_tree &t) : sz_(model_sizes(c.model, c.dim)) {
        int rc = dacti_create(&c, &t, &h_);
        if (rc != DACTI_OK) throw Error(rc, dacti_last_error(nullptr));
    }
    GpuSolver(const GpuSolver &) = delete;
    GpuSolver &operator=(const GpuSolver &) = delete;
    GpuSolver(GpuSolver &&o) noexcept : h_(std::exchange(o.h_, nullptr)), sz_(o.sz_) {}
    ~GpuSolver() {
        if (h_) dacti_destroy(h_);
    }

    // One global step: returns the global dt; force_break is set on a NaN state.
    double integrate(bool &force_break, bool signal = true) {
        double dt = 0.0;
        int32_t fb = 0;
        check(dacti_integrate(h_, signal ? 1 : 0, &dt, &fb));
        force_break = fb != 0;
        return dt;
    }

    // `steps` global steps; returns the number completed before a break.
    int64_t run(int64_t steps, bool signal = true) {
        int64_t done = 0;
        check(dacti_run(h_, steps, signal ? 1 : 0, &done));
        return done;
    }

    int64_t n_active_cells() const {
        int64_t n = 0;
        check(dacti_n_active(h_, &n));
        return n;
    }

    int64_t n_nodes() const {
        int64_t n = 0;
        check(dacti_n_nodes(h_, &n));
        return n;
    }

    NodePool get_nodes() const {
        NodePool P;
        P.n = n_nodes();
        P.sz = sz_;
        const size_t n = size_t(P.n), D = sz_.dim, NE = sz_.num_eqns;
        const size_t NP = sz_.num_parm > 0 ? sz_.num_parm : 1, NG = sz_.num_grad > 0 ? sz_.num_grad : 1;
        P.parent.resize(n), P.child_base.resize(n), P.nbr.resize(n * 2 * D);
        P.tree_level.resize(n), P.acti_level.resize(n), P.cluster.resize(n), P.flags.resize(n);
        P.center.resize(n * D), P.size.resize(n), P.t.resize(n);
        P.p.resize(n * NE), P.u.resize(n * NE), P.slope.resize(n * D * NE), P.rhs.resize(n * NE);
        P.k.resize(n * NP), P.grad.resize(n * D * NG), P.meta.resize(n * NumMeta);
        dacti_node_view v{P.parent.data(), P.child_base.data(), P.nbr.data(),
                          P.tree_level.data(), P.acti_level.data(), P.cluster.data(),
                          P.flags.data(), P.center.data(), P.size.data(), P.t.data(),
                          P.p.data(), P.u.data(), P.slope.data(), P.rhs.data(),
                          P.k.data(), P.grad.data(), P.meta.data()};
        check(dacti_export_nodes(h_, &v));
        return P;
    }

    // action(i, node_index, pool) over the leaves of the LTS clusters and then
    // the acti buffers, level by level (integrator.hpp:172-191).
    template <typename F>
    void for_each_active_cell(F &&action) const {
        NodePool P = get_nodes();
        int64_t i = 0;
        for (int key = 0; key < 2 * MAX_LEVEL; ++key)
            for (int64_t idx = 0; idx < P.n; ++idx)
                if (P.cluster[idx] == key) action(i++, idx, P);
    }

    dacti_stats stats() const {
        dacti_stats s{};
        check(dacti_get_stats(h_, &s));
        return s;
    }
    // LTS schedule state: acti_time[L], acti_counter[L] (integrator.hpp:48-49)
    void schedule(double *acti_time, int32_t *acti_counter) const {
        check(dacti_get_schedule(h_, acti_time, acti_counter));
    }
    // one global step with fn(user, handle, level, acti_time, n_substeps, dt)
    // called after every substep (impl/integrate.hpp:64)
    double integrate_hooked(dacti_substep_hook fn, void *user, bool &force_break, bool signal = true) {
        check(dacti_set_substep_hook(h_, fn, user));
        double dt = 0.0;
        int32_t fb = 0;
        const int rc = dacti_integrate(h_, signal ? 1 : 0, &dt, &fb);
        dacti_set_substep_hook(h_, nullptr, nullptr);
        check(rc);
        force_break = fb != 0;
        return dt;
    }
    double global_dt() const { return stats().global_dt; }
    int max_acti_level() const { return stats().max_acti_level; }
    void set_timing(bool on) { check(dacti_set_timing(h_, on ? 1 : 0)); }
    dacti_handle *handle() const { return h_; }

private:
    void check(int rc) const {
        if (rc != DACTI_OK) throw Error(rc, dacti_last_error(h_));
    }
    dacti_handle *h_ = nullptr;
    ModelSizes sz_;
};

}  // namespace dacti

#endif
