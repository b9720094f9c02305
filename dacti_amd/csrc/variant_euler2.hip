// variant_euler2.hip -- the leaf-update kernels and the batched substep driver for one
// model (euler, 2-D), compiled as a translation unit of its own.
#include "runtime.hpp"

namespace dacti_rt {
int variant_init_EULER_2(dacti_handle *h) { return Ops<EULER, 2>::init(h); }
int variant_integrate_EULER_2(dacti_batch *b, int signal, int *force_break) {
    return BOps<EULER, 2>::integrate(b, signal, force_break);
}
}  // namespace dacti_rt
