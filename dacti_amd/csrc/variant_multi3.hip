// variant_multi3.hip -- the leaf-update kernels and the batched substep driver for one
// model (multi, 3-D), compiled as a translation unit of its own.
#include "runtime.hpp"

namespace dacti_rt {
int variant_init_MULTI_3(dacti_handle *h) { return Ops<MULTI, 3>::init(h); }
int variant_integrate_MULTI_3(dacti_batch *b, int signal, int *force_break) {
    return BOps<MULTI, 3>::integrate(b, signal, force_break);
}
}  // namespace dacti_rt
