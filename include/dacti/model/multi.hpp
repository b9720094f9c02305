// dacti/model/multi.hpp -- model::multi<D> (src/lib/model/multi.hpp of the reference).
#ifndef DACTI_MODEL_MULTI_HPP
#define DACTI_MODEL_MULTI_HPP
#include "common.hpp"
#endif
