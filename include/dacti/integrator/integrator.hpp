// dacti/integrator/integrator.hpp -- integrator::Integrator<Model, Limiter,
// ErrorEstimator>, the reference's solver interface over the GPU path.
//
// Same template name, constructor and class-template argument deduction as
// integrator/integrator.hpp:14-15,72-76 of the reference:
//
//   integrator::Integrator integrator{integrator::slope::minmod{},
//       integrator::error::weighted_laplacian{config->error_weights},
//       scene, config};
//
// Construction builds the scene's initial tree and runs the integrator's
// initialisation (integrator.hpp:72-163) on the GPU.  integrate(hook, signal,
// force_break) is one global step (impl/integrate.hpp:11-87) and calls
// hook(integrator, level, acti_time, n_substeps, dt) after every substep
// (impl/integrate.hpp:64); integrate(signal, force_break) is the same step
// without a hook, so the GPU never waits on the host between substeps.
// The read side keeps the reference's names: n_active_cells,
// for_each_active_cell(action(i, node)) in cluster-then-buffer level order
// (integrator.hpp:172-191; inside one level the order is the node order),
// get_node, get_nodes, get_cell_index, get_domain and the public fields
// global_dt, acti_time, acti_counter, n_substeps, n_global_steps,
// max_acti_level, total_cell_updates, runtime_duration, METADATA_NAMES.
// Nodes are host copies taken from the GPU pool when first read after a
// step (one copy per step, reused until the next step).
//
// Differences from the reference, by design: copies of an Integrator share
// one GPU state (the reference's default_driver copies the integrator by
// value, driver/default_driver.hpp:20,38, and then only uses its copy);
// public counters are refreshed by the copy that steps.  The model's
// case constants are per GPU handle.  Errors throw dacti::Error.
#ifndef DACTI_INTEGRATOR_INTEGRATOR_HPP
#define DACTI_INTEGRATOR_INTEGRATOR_HPP

#include <array>
#include <chrono>
#include <cmath>
#include <exception>
#include <memory>
#include <utility>
#include <vector>

#include "../config/config.hpp"
#include "../scene/scene.hpp"
#include "error_estimators.hpp"
#include "gpu_solver.hpp"
#include "slope_limiters.hpp"

namespace dacti::integrator {

namespace _internal {

// integrator/cell.hpp:8-107: one cell's state with the reference's accessors
template <typename Model>
class Cell {
public:
    using vecn_t = typename Model::vecn_t;
    using prim_t = typename Model::prim_t;
    using conv_t = typename Model::conv_t;
    using parm_t = typename Model::parm_t;
    using grad_t = typename Model::grad_t;
    using matS_t = typename Model::matS_t;
    using matG_t = typename Model::matG_t;

    scalar_t size = 0, t = 0;
    vecn_t center;
    conv_t u, rhs;
    prim_t p;
    parm_t k;
    grad_t t_slope;  // not kept on the host side (zero)
    matS_t slope{};
    matG_t gradient{};

    prim_t get_p(const vecn_t &x, scalar_t) const {
        prim_t r = p;
        for (size_t i = 0; i < Model::NumDims; ++i) r += slope[i] * (x[i] - center[i]);
        return r;
    }
    prim_t get_p(const vecn_t &x, scalar_t, size_t dim) const { return p + slope[dim] * (x[dim] - center[dim]); }
    conv_t get_u() const { return u; }
    prim_t get_p() const { return p; }
    prim_t get_slope(size_t dim) const { return slope[dim]; }
    matS_t get_slope() const { return slope; }
    parm_t get_k(const vecn_t &, scalar_t) const { return k; }
    parm_t get_k() const { return k; }
    matG_t get_gradient() const { return gradient; }
    vecn_t get_velocity() const {
        vecn_t v;
        for (size_t i = 0; i < Model::NumDims; ++i) v[i] = p[Model::iu + i];
        return v;
    }
    vecn_t get_center() const { return center; }
    vecn_t get_size() const { return vecn_t::constant(size); }
    scalar_t get_t() const { return t; }
    scalar_t volume() const { return std::pow(size, (double) Model::NumDims); }
    conv_t get_rhs() const { return rhs; }
    vecn_t face_position(int dim, int dir) const {
        vecn_t r = center;
        r[dim] += (2 * dir - 1) * size / 2.0;
        return r;
    }
};

// integrator/node.hpp:17-67: a node of the pool
template <typename Model>
class Node {
public:
    static constexpr size_t NumChildren = size_t(1) << Model::NumDims;
    Cell<Model> cell;
    std::array<size_t, NumChildren> child_indices{};
    std::array<size_t, 2 * Model::NumDims> neighbors{};  // index 2*dim+dir
    int acti_level = 0, tree_level = -1;
    size_t parent_index = 0;
    scalar_t error = 0, max_dt = 0, cfl = 0;
    bool is_leaf = false, is_active = false, is_garbage = false, can_split = false, can_unsplit = false,
         should_split = false, is_boundary = false;

    std::array<double, 4> get_metadata() const { return {(scalar_t) acti_level, error, max_dt, cfl}; }
    size_t find_child(size_t child_index) const {
        for (size_t i = 0; i < NumChildren; ++i)
            if (child_index == child_indices[i]) return i;
        return 0;
    }
};

// Node i of a host copy of the pool
template <typename Model>
Node<Model> make_node(const NodePool &P, size_t i) {
    constexpr size_t D = Model::NumDims, NE = Model::NumEqns, NP = Model::NumParm, NG = Model::NumGrad;
    const size_t NPX = NP > 0 ? NP : 1, NGX = NG > 0 ? NG : 1;
    Node<Model> n;
    auto &c = n.cell;
    c.size = P.size[i];
    c.t = P.t[i];
    for (size_t d = 0; d < D; ++d) c.center[d] = P.center[i * D + d];
    for (size_t e = 0; e < NE; ++e) {
        c.p[e] = P.p[i * NE + e];
        c.u[e] = P.u[i * NE + e];
        c.rhs[e] = P.rhs[i * NE + e];
        for (size_t d = 0; d < D; ++d) c.slope[d][e] = P.slope[(i * D + d) * NE + e];
    }
    for (size_t q = 0; q < NP; ++q) c.k[q] = P.k[i * NPX + q];
    for (size_t d = 0; d < D; ++d)
        for (size_t q = 0; q < NG; ++q) c.gradient[d][q] = P.grad[(i * D + d) * NGX + q];
    const uint8_t f = P.flags[i];
    n.is_leaf = f & DACTI_F_LEAF;
    n.is_active = f & DACTI_F_ACTIVE;
    n.is_garbage = f & DACTI_F_GARBAGE;
    n.can_split = f & DACTI_F_CAN_SPLIT;
    n.can_unsplit = f & DACTI_F_CAN_UNSPLIT;
    n.should_split = f & DACTI_F_SHOULD_SPLIT;
    n.is_boundary = f & DACTI_F_BOUNDARY;
    n.tree_level = P.tree_level[i];
    n.acti_level = P.acti_level[i];
    n.parent_index = (size_t) P.parent[i];
    if (P.child_base[i] > 0)
        for (size_t ch = 0; ch < Node<Model>::NumChildren; ++ch) n.child_indices[ch] = (size_t) P.child_base[i] + ch;
    for (size_t f2 = 0; f2 < 2 * D; ++f2) n.neighbors[f2] = (size_t) std::max(P.nbr[i * 2 * D + f2], 0);
    n.error = P.meta[i * 4 + 1];
    n.max_dt = P.meta[i * 4 + 2];
    n.cfl = P.meta[i * 4 + 3];
    return n;
}

// Host view of the pool with the reference's tree queries
// (integrator/domain.hpp:14-82,407-417)
template <typename Model>
class Domain {
public:
    using vecn_t = typename Model::vecn_t;
    static constexpr size_t root_index = 1;
    int max_depth = DACTI_MAX_LEVEL;

    Domain(std::shared_ptr<const NodePool> pool, vecn_t active_size, std::array<bool, Model::NumDims> periodic)
        : pool_(std::move(pool)), active_size_(active_size), periodic_(periodic) {}

    size_t size() const { return (size_t) pool_->n; }
    Node<Model> operator[](size_t i) const { return make_node<Model>(*pool_, i); }
    const NodePool &pool() const { return *pool_; }

    // leaf (or node at max_level) containing p, after one periodic wrap; 0
    // outside a non-periodic domain
    size_t find(vecn_t p, int max_level) const {
        for (size_t i = 0; i < Model::NumDims; ++i) {
            if (p[i] < -0.5 * active_size_[i] || p[i] > 0.5 * active_size_[i]) {
                if (!periodic_[i]) return 0;
                p[i] += p[i] < -0.5 * active_size_[i] ? active_size_[i] : -active_size_[i];
            }
        }
        vecn_t x = 2.0 * p / pool_->size[root_index];
        size_t idx = root_index;
        const NodePool &P = *pool_;
        while (!(P.flags[idx] & DACTI_F_LEAF) && P.tree_level[idx] != max_level) {
            size_t q = 0;
            for (size_t d = 0; d < Model::NumDims; ++d)
                if (x[d] > 0.0) q |= size_t(1) << d;
            x = (x - 0.5 * vecn_t::quadrant(q)) * 2.0;
            idx = (size_t) P.child_base[idx] + q;
        }
        return idx;
    }

private:
    std::shared_ptr<const NodePool> pool_;
    vecn_t active_size_;
    std::array<bool, Model::NumDims> periodic_;
};

}  // namespace _internal

template <typename Model, typename Limiter, typename ErrorEstimator>
class Integrator {
public:
    using model_type = Model;
    using vecn_t = typename Model::vecn_t;
    using node_t = _internal::Node<Model>;

    static constexpr size_t NumDims = Model::NumDims;
    static constexpr size_t NumMeta = 4;
    static constexpr int MAX_LEVEL = DACTI_MAX_LEVEL;
    static constexpr std::array<const char *, NumMeta> METADATA_NAMES = {"acti_level", "error", "max_dt", "cfl"};

    scene::Scene<Model> &scene;
    std::shared_ptr<config::Config<Model>> config;
    Limiter limiter;
    ErrorEstimator error_estimator;

    scalar_t max_error, min_error;
    size_t max_tree_depth;
    scalar_t max_dt, max_cfl;

    std::array<int, MAX_LEVEL> acti_counter{};
    std::array<scalar_t, MAX_LEVEL> acti_time{};
    int n_substeps = 0, n_global_steps = 0, max_acti_level = 0;
    unsigned long long total_flux_computations = 0, total_cell_updates = 0;
    scalar_t global_dt = -INFINITY, runtime_duration = 0.0;

    Integrator(Limiter limiter, ErrorEstimator error_estimator, scene::Scene<Model> &scene,
               std::shared_ptr<config::Config<Model>> config)
        : scene(scene), config(config), limiter(limiter), error_estimator(error_estimator) {
        max_error = config->error_threshold[1];
        min_error = config->error_threshold[0];
        max_tree_depth = (size_t) config->max_refinement_level;
        max_dt = config->max_dt;
        max_cfl = config->max_cfl;
        const Tree tree = scene.init_tree();
        const dacti_case c = make_case();
        const dacti_tree tv = tree.view();
        st_ = std::make_shared<State>();
        st_->solver = std::make_unique<GpuSolver>(c, tv);
        refresh();
    }

    // one global step with the per-substep hook (impl/integrate.hpp:11-87)
    template <typename Hook>
    scalar_t integrate(Hook &&substep_hook, bool signal, bool &force_break) {
        HookCtx<std::remove_reference_t<Hook>> ctx{this, &substep_hook, nullptr};
        const auto t0 = std::chrono::steady_clock::now();
        st_->snap.reset();
        const double dt = st_->solver->integrate_hooked(&HookCtx<std::remove_reference_t<Hook>>::call, &ctx,
                                                        force_break, signal);
        finish(t0);
        if (ctx.err) std::rethrow_exception(ctx.err);
        return dt;
    }

    // one global step without a hook
    scalar_t integrate(bool signal, bool &force_break) {
        const auto t0 = std::chrono::steady_clock::now();
        st_->snap.reset();
        const double dt = st_->solver->integrate(force_break, signal);
        finish(t0);
        return dt;
    }

    size_t n_active_cells() const { return (size_t) st_->solver->n_active_cells(); }

    template <typename F>
    void for_each_active_cell(F &&action) const {
        const NodePool &P = *snapshot();
        std::vector<std::vector<size_t>> keys(2 * MAX_LEVEL);
        for (int64_t idx = 0; idx < P.n; ++idx) {
            const int key = P.cluster[(size_t) idx];
            if (key >= 0 && key < 2 * MAX_LEVEL) keys[(size_t) key].push_back((size_t) idx);
        }
        int i = 0;
        for (const auto &list : keys)
            for (size_t idx : list) {
                const node_t node = _internal::make_node<Model>(P, idx);
                action(i, node);
                ++i;
            }
    }

    node_t get_node(size_t index) const { return _internal::make_node<Model>(*snapshot(), index); }
    std::vector<node_t> get_nodes() const {
        const auto s = snapshot();
        std::vector<node_t> out;
        out.reserve((size_t) s->n);
        for (int64_t i = 0; i < s->n; ++i) out.push_back(_internal::make_node<Model>(*s, (size_t) i));
        return out;
    }
    _internal::Domain<Model> get_domain() const { return {snapshot(), config->active_size, config->IsPeriodic}; }
    size_t get_cell_index(vecn_t pos) const { return get_domain().find(pos, DACTI_MAX_LEVEL); }

    // the untyped GPU solver behind this integrator (C-ABI handle, stats)
    GpuSolver &solver() const { return *st_->solver; }

private:
    struct State {
        std::unique_ptr<GpuSolver> solver;
        std::shared_ptr<const NodePool> snap;
    };
    std::shared_ptr<State> st_;

    template <typename Hook>
    struct HookCtx {
        Integrator *self;
        Hook *hook;
        std::exception_ptr err;
        static void call(void *user, dacti_handle *, int32_t level, double acti_time, int64_t n_sub, double dt) {
            auto *c = static_cast<HookCtx *>(user);
            if (c->err) return;  // the first exception ends the hook calls of this step
            try {
                c->self->st_->snap.reset();
                c->self->refresh();
                (*c->hook)(*c->self, (int) level, acti_time, (int) n_sub, dt);
            } catch (...) {
                c->err = std::current_exception();
            }
        }
    };

    std::shared_ptr<const NodePool> snapshot() const {
        if (!st_->snap) st_->snap = std::make_shared<const NodePool>(st_->solver->get_nodes());
        return st_->snap;
    }

    void finish(std::chrono::steady_clock::time_point t0) {
        runtime_duration += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        refresh();
    }

    void refresh() {
        const dacti_stats s = st_->solver->stats();
        n_substeps = (int) s.n_substeps;
        n_global_steps = (int) s.n_global_steps;
        max_acti_level = s.max_acti_level;
        total_cell_updates = (unsigned long long) s.cell_updates;
        global_dt = s.global_dt;
        double t[DACTI_MAX_LEVEL];
        int32_t c[DACTI_MAX_LEVEL];
        st_->solver->schedule(t, c);
        for (int l = 0; l < MAX_LEVEL; ++l) {
            acti_time[(size_t) l] = t[l];
            acti_counter[(size_t) l] = c[l];
        }
    }

    // the case as the C ABI takes it: the parsed case with the Config's
    // public fields (which a caller may have changed), the limiter's id and
    // the estimator's weights
    dacti_case make_case() const {
        Case cs = config->effective_case();
        cs.limiter = Limiter::id;
        cs.max_cfl = max_cfl;
        cs.max_dt = max_dt;
        cs.error_threshold = {min_error, max_error};
        cs.weights.assign(error_estimator.weights.begin(), error_estimator.weights.end());
        return to_dacti_case(cs);
    }
};

}  // namespace dacti::integrator

#endif
