// variant_navier_stokes3.hip -- the leaf-update kernels and the batched substep driver for one
// model (navier_stokes, 3-D), compiled as a translation unit of its own.
#include "runtime.hpp"

namespace dacti_rt {
int variant_init_NAVIER_STOKES_3(dacti_handle *h) { return Ops<NAVIER_STOKES, 3>::init(h); }
int variant_integrate_NAVIER_STOKES_3(dacti_batch *b, int signal, int *force_break) {
    return BOps<NAVIER_STOKES, 3>::integrate(b, signal, force_break);
}
}  // namespace dacti_rt
