// scene.hip -- the TOML case reader and initial-tree builder behind the C ABI.
//
// dacti_scene_build runs the header implementation of include/dacti/config
// (Config<M>::load_config, config/config.cpp:10-243 of the reference) and
// build_tree (Scene<M>::init_domain, scene/scene.cpp:369-431) on the host, so
// that callers of the C ABI (and the Python binding) read a TOML case and get
// the case structure and initial tree that dacti_create takes.  Host code
// only.
#include <string>

#include "../../include/dacti/config/case.hpp"
#include "../../include/dacti_gpu.h"

namespace dacti_rt {
extern thread_local std::string g_create_error;
}

struct dacti_scene {
    dacti::Case cs;
    dacti::Tree tree;
};

extern "C" {

int dacti_scene_build(const char *toml_text, const char *base_dir, int32_t model, int32_t dim, int32_t limiter,
                      const double *ic_grid, const int32_t *ic_grid_shape, dacti_scene **out) {
    if (!toml_text || !out) {
        dacti_rt::g_create_error = "dacti_scene_build: null argument";
        return DACTI_E_ARG;
    }
    *out = nullptr;
    if (model < 0 || model > 2 || dim < 2 || dim > 3 || limiter < 0 || limiter > 6) {
        dacti_rt::g_create_error = "dacti_scene_build: unsupported model, dimension or limiter";
        return DACTI_E_ARG;
    }
    try {
        auto s = new dacti_scene();
        s->cs = dacti::load_case(*dacti::toml::parse(toml_text), model, dim, limiter, base_dir ? base_dir : ".");
        if (s->cs.ic_type == "grid") {
            if (!ic_grid || !ic_grid_shape) {
                delete s;
                dacti_rt::g_create_error = "dacti_scene_build: a grid initial condition needs its grid";
                return DACTI_E_ARG;
            }
            size_t total = (size_t) s->cs.num_eqns();
            for (int d = 0; d < dim; ++d) {
                s->cs.ic_grid_shape[d] = ic_grid_shape[d];
                total *= (size_t) std::max(ic_grid_shape[d], 0);
            }
            s->cs.ic_grid.assign(ic_grid, ic_grid + total);
        }
        s->tree = dacti::build_tree(s->cs);
        *out = s;
    } catch (const std::exception &e) {
        dacti_rt::g_create_error = std::string("dacti_scene_build: ") + e.what();
        return DACTI_E_ARG;
    }
    return DACTI_OK;
}

int dacti_scene_case(const dacti_scene *s, dacti_case *out) {
    if (!s || !out) return DACTI_E_ARG;
    *out = dacti::to_dacti_case(s->cs);
    return DACTI_OK;
}

int dacti_scene_tree(const dacti_scene *s, dacti_tree *out) {
    if (!s || !out) return DACTI_E_ARG;
    *out = s->tree.view();
    return DACTI_OK;
}

int dacti_scene_destroy(dacti_scene *s) {
    delete s;
    return DACTI_OK;
}

}  // extern "C"
