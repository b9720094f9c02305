// variant_multi2.hip -- the leaf-update kernels and the batched substep driver for one
// model (multi, 2-D), compiled as a translation unit of its own.
#include "runtime.hpp"

namespace dacti_rt {
int variant_init_MULTI_2(dacti_handle *h) { return Ops<MULTI, 2>::init(h); }
int variant_integrate_MULTI_2(dacti_batch *b, int signal, int *force_break) {
    return BOps<MULTI, 2>::integrate(b, signal, force_break);
}
int variant_features_MULTI_2(dacti_handle *h, int min_level, const float *mu, const float *sd, float *feat,
                               int32_t *ids, double *centers, int8_t *levels, int64_t cap, int64_t *n_out,
                               int32_t *n_feat) {
    return Ops<MULTI, 2>::features(h, min_level, mu, sd, feat, ids, centers, levels, cap, n_out, n_feat);
}
int variant_locate_MULTI_2(dacti_handle *h, const double *pts, int64_t n, int8_t *lev) {
    return Ops<MULTI, 2>::locate(h, pts, n, lev);
}
int variant_mark_predicted_MULTI_2(dacti_handle *h, const float *logits, const int32_t *ids, int64_t n, int rings,
                                     int64_t *n_marked) {
    return Ops<MULTI, 2>::mark_predicted(h, logits, ids, n, rings, n_marked);
}
}  // namespace dacti_rt
