// dacti/core/types.hpp -- scalar and vector aliases (core/types.hpp:10-27 of
// the reference): FP64 scalars, vec_t<N> value vectors.
#ifndef DACTI_CORE_TYPES_HPP
#define DACTI_CORE_TYPES_HPP

#include <cmath>
#include <limits>

#include "linalg.hpp"

namespace dacti {

template <size_t N>
using vec_t = linalg::vec<scalar_t, N>;
using vec4_t = vec_t<4>;

namespace core {
using dacti::scalar_t;
using dacti::vec2_t;
using dacti::vec3_t;
using dacti::vec4_t;
using dacti::vec_t;
constexpr scalar_t epsilon() { return std::numeric_limits<scalar_t>::epsilon(); }
}  // namespace core

}  // namespace dacti

#endif
