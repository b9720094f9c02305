// dacti/scene/scene.hpp -- scene::Scene<M>: the initial tree of a case.
//
// Scene<M>(config) holds the case; init_tree() grows the initial octree as
// Scene<M>::init_domain does (scene/scene.cpp:369-431 of the reference; see
// build_tree in config/case.hpp for the scope), and can_split() is
// scene/scene.hpp:50-61.  The integrator calls init_tree() at construction
// and hands the tree to the GPU.
#ifndef DACTI_SCENE_HPP
#define DACTI_SCENE_HPP

#include <memory>
#include <stdexcept>

#include "../config/config.hpp"

namespace dacti::scene {

template <typename M>
class Scene {
public:
    const std::shared_ptr<config::Config<M>> config;

    explicit Scene(std::shared_ptr<config::Config<M>> config) : config(config) {}

    Tree init_tree() const { return build_tree(config->effective_case()); }

    bool can_split(const typename M::vecn_t &x, size_t current_level) const {
        if ((int) current_level < config->max_refinement_level) return true;
        const vec3_t x3 = x.template project<3>();
        for (const auto &zone : config->refine_zones)
            if (current_level < zone.refinement_level && zone.contains(x3)) return true;
        return false;
    }

    // the reference opens an interactive OpenGL view of the initial mesh
    // (scene/scene.cpp:434-506); there is no viewer here
    template <typename Domain>
    void visualize_mesh(const Domain &) const {
        throw std::runtime_error("visualize_mesh: no interactive viewer in this build; write a frame instead");
    }
};

}  // namespace dacti::scene

#endif
