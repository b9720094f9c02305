// dacti/model/navier_stokes.hpp -- model::navier_stokes<D> (src/lib/model/navier_stokes.hpp of the reference).
#ifndef DACTI_MODEL_NAVIER_STOKES_HPP
#define DACTI_MODEL_NAVIER_STOKES_HPP
#include "common.hpp"
#endif
