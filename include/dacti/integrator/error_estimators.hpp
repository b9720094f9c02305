// dacti/integrator/error_estimators.hpp -- the refinement error estimator.
#ifndef DACTI_ERROR_ESTIMATORS_HPP
#define DACTI_ERROR_ESTIMATORS_HPP

#include <cmath>

#include "../core/types.hpp"

namespace dacti {

namespace integrator::error {

// error = sqrt(sum_d |w (p- + p+ - 2p)|^2), a dimension's term zero when a
// neighbour is missing (integrator/error_estimators.hpp:10-38).  The GPU
// sweep evaluates it with these weights.
template <size_t NumEqns>
struct weighted_laplacian {
    static constexpr bool RequiresNeighborhood = true;
    vec_t<NumEqns> weights;
    weighted_laplacian(vec_t<NumEqns> weights) : weights(weights) {}

    // host evaluation on primitive states: p, then the neighbours' states
    // nb[dim][dir] (nullptr = no neighbour)
    scalar_t estimate(const vec_t<NumEqns> &p, const vec_t<NumEqns> *const nb[][2], size_t dims) const {
        scalar_t error = 0.0;
        for (size_t d = 0; d < dims; ++d) {
            if (!nb[d][0] || !nb[d][1]) continue;
            vec_t<NumEqns> d2 = -2.0 * p;
            d2 += *nb[d][0];
            d2 += *nb[d][1];
            error += d2.cwiseMul(weights).squaredNorm();
        }
        return std::sqrt(error);
    }
};

}  // namespace integrator::error

}  // namespace dacti

#endif
