// dacti/model/euler.hpp -- model::euler<D> (src/lib/model/euler.hpp of the reference).
#ifndef DACTI_MODEL_EULER_HPP
#define DACTI_MODEL_EULER_HPP
#include "common.hpp"
#endif
