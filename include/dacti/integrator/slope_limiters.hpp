// dacti/integrator/slope_limiters.hpp -- the seven slope limiters as tags.
//
// limiter(s1, s2) of integrator/slope_limiters.hpp:7-57 on the host; the
// integrator passes the tag's id to the GPU sweep, which evaluates the
// same expression per slope.
#ifndef DACTI_SLOPE_LIMITERS_HPP
#define DACTI_SLOPE_LIMITERS_HPP

#include <cmath>

#include "../../dacti_gpu.h"
#include "../core/types.hpp"

namespace dacti {

namespace integrator::slope {

struct minmod {
    static constexpr int id = DACTI_LIMITER_MINMOD;
    scalar_t operator()(scalar_t s1, scalar_t s2) const {
        if (s1 * s2 <= 0.0) return 0.0;
        return std::abs(s1) < std::abs(s2) ? s1 : s2;
    }
};
struct unlimited {
    static constexpr int id = DACTI_LIMITER_UNLIMITED;
    scalar_t operator()(scalar_t s1, scalar_t s2) const { return 0.5 * (s1 + s2); }
};
struct van_leer {
    static constexpr int id = DACTI_LIMITER_VAN_LEER;
    scalar_t operator()(scalar_t s1, scalar_t s2) const {
        const scalar_t r = s1 / s2;
        return (r + std::abs(r)) / (1 + std::abs(r)) * s2;
    }
};
struct superbee {
    static constexpr int id = DACTI_LIMITER_SUPERBEE;
    scalar_t operator()(scalar_t s1, scalar_t s2) const {
        const scalar_t r = s1 / s2;
        const scalar_t a = (2.0 * r < 1.0) ? 2.0 * r : 1.0, b = (r < 2.0) ? r : 2.0;
        const scalar_t m = (0.0 < a) ? a : 0.0;
        return ((m < b) ? b : m) * s2;
    }
};
struct von_alba_2 {
    static constexpr int id = DACTI_LIMITER_VON_ALBA_2;
    scalar_t operator()(scalar_t s1, scalar_t s2) const {
        const scalar_t r = s1 / s2;
        return 2.0 * r / (r * r + 1) * s2;
    }
};
struct koren {
    static constexpr int id = DACTI_LIMITER_KOREN;
    scalar_t operator()(scalar_t s1, scalar_t s2) const {
        const scalar_t r = (s1 + 1e-9) / (s2 + 1e-9);
        const scalar_t a = (1.0 + 2.0 * r) / 3.0 < 2.0 ? (1.0 + 2.0 * r) / 3.0 : 2.0;
        const scalar_t b = 2.0 * r < a ? 2.0 * r : a;
        return (0.0 < b ? b : 0.0) * s2;
    }
};
struct first_order {
    static constexpr int id = DACTI_LIMITER_FIRST_ORDER;
    scalar_t operator()(scalar_t, scalar_t) const { return 0.0; }
};

}  // namespace integrator::slope

}  // namespace dacti

#endif
