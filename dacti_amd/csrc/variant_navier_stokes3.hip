// variant_navier_stokes3.hip -- the leaf-update kernels and the batched substep driver for one
// model (navier_stokes, 3-D), compiled as a translation unit of its own.
#include "runtime.hpp"

namespace dacti_rt {
int variant_init_NAVIER_STOKES_3(dacti_handle *h) { return Ops<NAVIER_STOKES, 3>::init(h); }
int variant_integrate_NAVIER_STOKES_3(dacti_batch *b, int signal, int *force_break) {
    return BOps<NAVIER_STOKES, 3>::integrate(b, signal, force_break);
}
int variant_features_NAVIER_STOKES_3(dacti_handle *h, int min_level, const float *mu, const float *sd, float *feat,
                               int32_t *ids, double *centers, int8_t *levels, int64_t cap, int64_t *n_out,
                               int32_t *n_feat) {
    return Ops<NAVIER_STOKES, 3>::features(h, min_level, mu, sd, feat, ids, centers, levels, cap, n_out, n_feat);
}
int variant_locate_NAVIER_STOKES_3(dacti_handle *h, const double *pts, int64_t n, int8_t *lev) {
    return Ops<NAVIER_STOKES, 3>::locate(h, pts, n, lev);
}
int variant_mark_predicted_NAVIER_STOKES_3(dacti_handle *h, const float *logits, const int32_t *ids, int64_t n, int rings,
                                     int64_t *n_marked) {
    return Ops<NAVIER_STOKES, 3>::mark_predicted(h, logits, ids, n, rings, n_marked);
}
}  // namespace dacti_rt
