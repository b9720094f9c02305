// runtime.hpp -- host runtime of the MI355X leaf-update engine (shared by
// the per-model translation units and the C ABI).
//
// The host side restates Integrator (src/lib/integrator/integrator.hpp):
// the handle (one integrator's state), the batch (instances stepped
// together, batch.hpp) and the per-model construction (Ops::init).  Every
// per-cell phase runs as a HIP kernel over the level's cluster list
// (kernels.hpp); the host keeps only scalars per instance.
#pragma once
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dacti_gpu.h"
#include "kernels.hpp"

static_assert(sizeof(dacti::BndRec) == sizeof(dacti_boundary) && offsetof(dacti::BndRec, p) == offsetof(dacti_boundary, p) &&
                  offsetof(dacti::BndRec, fade_in) == offsetof(dacti_boundary, fade_in),
              "BndRec mirrors dacti_boundary");

using namespace dacti;

namespace dacti_rt {

struct Sizes {
    int D, NE, NF, NP, NG, NPX, NGX, NCH, NFACE;
};

inline Sizes model_sizes(int model, int D) {
    Sizes s;
    s.D = D;
    s.NE = model == MULTI ? D + 4 : D + 2;
    s.NF = model == MULTI ? s.NE + D + 1 : s.NE;
    s.NP = model == MULTI ? D + 4 : 0;
    s.NG = model == EULER ? 0 : (model == NAVIER_STOKES ? D + 1 : D + 3);
    s.NPX = s.NP > 0 ? s.NP : 1;
    s.NGX = s.NG > 0 ? s.NG : 1;
    s.NCH = 1 << D;
    s.NFACE = 2 * D;
    return s;
}

// one device field: element size and number of components per node
struct Field {
    void **ptr;
    size_t esz;
    int ncomp;
};

}  // namespace dacti_rt

using dacti_rt::Sizes;

struct dacti_handle {
    int model = 0, D = 0;
    Sizes sz{};
    MConst mc{};
    double side_length = 0, max_dt_cfg = 0, max_cfl = 0;
    Dev d{};
    int64_t n = 0;  // nodes in the pool (incl. null node 0)
    double global_dt = 0;
    int acti_counter[16] = {0};
    double acti_time[16] = {0};
    int max_acti_level = 0;
    int64_t counts[16] = {0};
    int64_t pending[16] = {0};
    int64_t n_substeps = 0, n_global_steps = 0, cell_updates = 0;
    int64_t flux_launches = 0;
    double flux_ms = 0, substep_ms = 0;
    bool timing = false;
    hipStream_t stream = nullptr;
    int32_t *d_list = nullptr;
    int32_t *d_list2 = nullptr;
    int32_t *d_off = nullptr;
    uint8_t *d_rn = nullptr;
    int32_t *d_nsel = nullptr;
    BndRec *d_brecs = nullptr;  // object boundary records (fixed at construction)
    void *d_tmp = nullptr;
    size_t tmp_bytes = 0;
    unsigned long long *d_hist = nullptr;
    Status *d_status = nullptr;
    Status *d_status_home = nullptr;  // d_status as allocated (a batch swaps d_status while it steps)
    Status *h_status = nullptr;
    int32_t *h_nsel = nullptr;
    unsigned long long *h_hist = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    struct dacti_batch *solo = nullptr;    // batch of this handle alone
    int remesh = 1;  // 0: refinement marks paused (host-set cadence)
    int dense = 1;   // window flux op for dense clusters: 0 never, 1 when the cluster fills half the pool, 2 always
    int walign = 0;  // window alignment: first node id of the initial tree's deepest level, mod 64
    bool static_mesh = false;  // one tree level, no cell may split or unsplit: the mesh never changes
    dacti_substep_hook hook = nullptr;  // called after every substep (impl/integrate.hpp:64)
    void *hook_user = nullptr;
    std::string err;
};

// A batch: instances stepped together, one launch per op and LTS substep
// for all of them (see kernels.hpp, batched launches).  A handle alone is
// stepped as its own batch of one.
struct dacti_batch {
    std::vector<dacti_handle *> hs;
    int model = 0, D = 0, n = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    Item *h_items = nullptr, *d_items = nullptr;
    MConst *d_mcs = nullptr;                     // constants of each instance
    int *h_toff = nullptr, *d_toff = nullptr;    // list tiles, n + 1 entries
    int *h_toffp = nullptr, *d_toffp = nullptr;  // pool tiles, n + 1 entries
    int *h_toffr = nullptr, *d_toffr = nullptr;  // remesh: 256-entry list tiles of the meshing instances
    int *h_toff2 = nullptr, *d_toff2 = nullptr;  // remesh: tiles of the rebuilt lists (upper bounds)
    int *d_tilecnt = nullptr;
    int64_t tilecnt_cap = 0;
    Status *h_st = nullptr, *d_st = nullptr;
    unsigned long long *h_hist = nullptr, *d_hist = nullptr;
    // the blocks the pointers above live in: items and tile tables (one
    // upload per substep), status and histograms (one download)
    void *d_up = nullptr, *h_up = nullptr, *d_down = nullptr, *h_down = nullptr;
    size_t up_bytes = 0, down_bytes = 0;
    // launch configuration: threads per list tile, persistent grid (0 = one
    // block per tile), HIP graphs for the fixed kernel sequences
    int block = 256, grid = 0;
    bool graphs = false;
    hipGraphExec_t g_head = nullptr, g_tail = nullptr;
    hipGraphExec_t g_head_dense = nullptr;  // head with the window flux op (a dense cluster of a batch of one)
    hipGraphExec_t g_step = nullptr;        // a whole substep planned on the device
    hipGraphExec_t g_static = nullptr;      // head + tail of meshes that cannot change
    bool fixed_mesh = false;                // the substep being launched is one of those
    hipGraphExec_t g_static_dense = nullptr;
    bool dense = false;                     // this substep takes the window flux op
    bool timing = false;
    // head start/end, tail start/end, remesh start/end
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double flux_ms = 0, substep_ms = 0;
    int64_t cell_updates = 0, n_substeps = 0, n_global_steps = 0, launches = 0;
    int top = 15;  // no cluster above this level in any instance
    // instances whose update produced a NaN state: they stop where the
    // reference's integrate() returns (impl/integrate.hpp:50) and are not
    // stepped again (a batch of one restarts them on the next call, as the
    // reference does)
    std::vector<char> broken;
    // autotuner: candidate launch configurations, one timed global step each
    std::vector<int> tune_block, tune_grid, tune_graphs;
    std::vector<double> tune_ms;
    size_t tune_next = 0;
    bool tuning = false;
    std::string err;
};

namespace dacti_rt {

#define HIPCHK(h, call)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);              \
            return DACTI_E_HIP;                                                          \
        }                                                                                \
    } while (0)

inline unsigned blocks(int64_t n, int b = 256) { return (unsigned) ((n + b - 1) / b); }

int alloc_pool(dacti_handle *h, int64_t C, Dev *old, int64_t n);
int ensure_capacity(dacti_handle *h, int64_t need);
int refresh_counts(dacti_handle *h);
int gc(dacti_handle *h);
int batch_ensure_tiles(dacti_batch *b, int64_t tiles);
void batch_drop_graphs(dacti_batch *b);

// Makes a handle use its batch's stream and status slot while the batch
// steps it (remesh and gc run on the handle).
struct Bind {
    dacti_handle *h;
    hipStream_t s0;
    Status *d0, *h0;
    Bind(dacti_batch *b, int i) : h(b->hs[(size_t) i]), s0(h->stream), d0(h->d_status), h0(h->h_status) {
        h->stream = b->stream;
        h->d_status = b->d_st + i;
        h->h_status = b->h_st + i;
    }
    ~Bind() {
        h->stream = s0;
        h->d_status = d0;
        h->h_status = h0;
    }
};

inline Item base_item(dacti_batch *b, int i) {
    dacti_handle *h = b->hs[(size_t) i];
    Item it{};
    it.d = h->d;
    it.d.err = (GMEM int32_t *) &b->d_st[i].dev_error;
    it.list = h->d_list;
    it.rest = h->d_list2;
    it.st = b->d_st + i;
    it.hist = b->d_hist + 16 * i;
    it.n = h->n;
    it.remesh = h->remesh;
    it.walign = h->walign;
    return it;
}

// per-model entry points, one translation unit each (variant_*.hip)
#define DACTI_VARIANTS(X) X(EULER, 2) X(EULER, 3) X(NAVIER_STOKES, 2) X(NAVIER_STOKES, 3) X(MULTI, 2) X(MULTI, 3)
#define DACTI_DECLARE_VARIANT(M, Dm) \
    int variant_init_##M##_##Dm(dacti_handle *h); \
    int variant_integrate_##M##_##Dm(dacti_batch *b, int signal, int *force_break); \
    int variant_features_##M##_##Dm(dacti_handle *h, int min_level, const float *mu, const float *sd, float *feat, \
                                    int32_t *ids, double *centers, int8_t *levels, int64_t cap, int64_t *n_out, \
                                    int32_t *n_feat); \
    int variant_locate_##M##_##Dm(dacti_handle *h, const double *pts, int64_t n, int8_t *lev); \
    int variant_mark_predicted_##M##_##Dm(dacti_handle *h, const float *logits, const int32_t *ids, int64_t n, \
                                          int rings, int64_t *n_marked);
DACTI_VARIANTS(DACTI_DECLARE_VARIANT)

template <int MODEL, int D>
struct Ops {
    using S = Sz<MODEL, D>;

    static int init(dacti_handle *h) {
        Dev &d = h->d;
        k_init_cells<MODEL, D><<<blocks(d.C), 256, 0, h->stream>>>(d, h->mc, h->n);
        HIPCHK(h, hipGetLastError());
        size_t bytes = h->tmp_bytes;
        double *d_max = (double *) d.tot;  // free until the first substep
        HIPCHK(h, hipcub::DeviceReduce::Max(h->d_tmp, bytes, (double *) d.scratch, d_max, (int) d.C, h->stream));
        double gmax = 0;
        HIPCHK(h, hipMemcpyAsync(&gmax, d_max, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        k_neighbors<D><<<blocks(h->n), 256, 0, h->stream>>>(d, h->mc, h->n);
        HIPCHK(h, hipGetLastError());
        // every active leaf, in node order
        struct LeafActive {
            const uint8_t *flags;
            __host__ __device__ bool operator()(const int32_t &i) const {
                return (flags[i] & F_LEAF) && (flags[i] & F_ACTIVE);
            }
        };
        hipcub::CountingInputIterator<int32_t> it(1);
        bytes = h->tmp_bytes;
        HIPCHK(h, hipcub::DeviceSelect::If(h->d_tmp, bytes, it, h->d_list, h->d_nsel, (int) (h->n - 1),
                                           LeafActive{d.flags}, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->h_nsel, h->d_nsel, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        const int nl = *h->h_nsel;
        if (nl > 0) {
            k_slope<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(d, h->mc, h->d_list, nl);
            HIPCHK(h, hipGetLastError());
            if constexpr (MODEL == MULTI) {
                int rc = chain(h, nl);
                if (rc) return rc;
            }
        }
        // global_dt = std::min(0.9 * max_dt_leaves, integrator.max_dt): integrator.hpp:150
        h->global_dt = (h->max_dt_cfg < 0.9 * gmax) ? h->max_dt_cfg : 0.9 * gmax;
        k_assign<<<blocks(h->n), 256, 0, h->stream>>>(d, h->n, h->global_dt, h->max_cfl);
        HIPCHK(h, hipGetLastError());
        int rc = refresh_counts(h);
        if (rc) return rc;
        h->max_acti_level = 0;
        for (int l = 0; l < 16; ++l) {
            h->acti_counter[l] = 0;
            h->acti_time[l] = 0.0;
            if (h->counts[l] > 0) h->max_acti_level = std::max(h->max_acti_level, l);
        }
        return DACTI_OK;
    }

    // multiphase interface fields over h->d_list[0, nl) (constructor,
    // integrator.hpp:130-140), through the batched ops of the handle's batch
    static int chain(dacti_handle *h, int nl) {
        if constexpr (MODEL == MULTI) {
            dacti_batch *b = h->solo;
            b->h_items[0] = base_item(b, 0);
            b->h_items[0].nl = nl;
            b->h_toff[0] = 0;
            b->h_toff[1] = (int) blocks(nl, b->block);
            HIPCHK(h, hipMemcpyAsync(b->d_items, b->h_items, sizeof(Item), hipMemcpyHostToDevice, h->stream));
            HIPCHK(h, hipMemcpyAsync(b->d_toff, b->h_toff, 2 * sizeof(int), hipMemcpyHostToDevice, h->stream));
            HIPCHK(h, hipMemsetAsync(b->d_st, 0, sizeof(Status), h->stream));
            const unsigned g = std::max(1u, (unsigned) b->h_toff[1]);
            const unsigned ga = g * (unsigned) (b->block / 64);  // k_axes: 64 cells a block
            for (int it = 0; it < 3; ++it) {
                k_axes<D><<<ga, 64 * D, 0, h->stream>>>(b->d_items, b->d_mcs, b->d_toff, 1, b->block, OpSmoothAlpha<MODEL, D>{});
                k_batch<<<g, b->block, 0, h->stream>>>(b->d_items, b->d_mcs, b->d_toff, 1, OpStoreK{S::IALPHAS});
            }
            k_axes<D><<<ga, 64 * D, 0, h->stream>>>(b->d_items, b->d_mcs, b->d_toff, 1, b->block, OpNormal<MODEL, D>{});
            k_axes<D><<<ga, 64 * D, 0, h->stream>>>(b->d_items, b->d_mcs, b->d_toff, 1, b->block, OpCurvature<MODEL, D>{});
            k_axes<D><<<ga, 64 * D, 0, h->stream>>>(b->d_items, b->d_mcs, b->d_toff, 1, b->block, OpSmoothCurvature<MODEL, D>{});
            k_batch<<<g, b->block, 0, h->stream>>>(b->d_items, b->d_mcs, b->d_toff, 1, OpStoreK{S::IKAPPA});
            HIPCHK(h, hipGetLastError());
            HIPCHK(h, hipStreamSynchronize(h->stream));
        }
        return DACTI_OK;
    }

    // predictor features of the active leaves (ids ascending): k_features
    static int features(dacti_handle *h, int min_level, const float *mu, const float *sd, float *feat, int32_t *ids,
                        double *centers, int8_t *levels, int64_t cap, int64_t *n_out, int32_t *n_feat) {
        *n_feat = n_features<MODEL, D>();
        const int64_t n = h->n;
        k_active_flags<<<blocks(n), 256, 0, h->stream>>>(h->d, n, h->d_rn);
        size_t bytes = h->tmp_bytes;
        HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(h->d_tmp, bytes, h->d_rn, h->d_off, (int) n, h->stream));
        int32_t last_off = 0;
        uint8_t last_fl = 0;
        HIPCHK(h, hipMemcpyAsync(&last_off, h->d_off + n - 1, 4, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipMemcpyAsync(&last_fl, h->d_rn + n - 1, 1, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        *n_out = (int64_t) last_off + last_fl;
        if (*n_out > cap || !feat || !ids) return DACTI_OK;  // sizes only
        k_active_ids<<<blocks(n), 256, 0, h->stream>>>(n, h->d_rn, h->d_off, ids, cap);
        if (*n_out > 0)
            k_features<MODEL, D><<<blocks(*n_out, 128), 128, 0, h->stream>>>(h->d, h->mc, ids, *n_out, min_level, mu, sd,
                                                                            feat, centers, levels);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return DACTI_OK;
    }

    static int locate(dacti_handle *h, const double *pts, int64_t n, int8_t *lev) {
        if (n > 0) k_locate_levels<D><<<blocks(n), 256, 0, h->stream>>>(h->d, h->mc, pts, n, lev);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipStreamSynchronize(h->stream));
        return DACTI_OK;
    }

    static int mark_predicted(dacti_handle *h, const float *logits, const int32_t *ids, int64_t n, int rings,
                              int64_t *n_marked) {
        const int64_t nn = h->n;
        HIPCHK(h, hipMemsetAsync(h->d_rn, 0, (size_t) nn, h->stream));
        if (n > 0) k_pred_seed<<<blocks(n), 256, 0, h->stream>>>(logits, ids, n, h->d_rn);
        for (int r = 1; r <= rings && r < 255; ++r)
            k_pred_ring<D><<<blocks(nn), 256, 0, h->stream>>>(h->d, nn, h->d_rn, (uint8_t) r);
        unsigned long long *cnt = (unsigned long long *) h->d_hist;
        HIPCHK(h, hipMemsetAsync(cnt, 0, sizeof(unsigned long long), h->stream));
        k_pred_mark<<<blocks(nn), 256, 0, h->stream>>>(h->d, nn, h->d_rn, cnt);
        unsigned long long c = 0;
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipMemcpyAsync(&c, cnt, sizeof c, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        *n_marked = (int64_t) c;
        return DACTI_OK;
    }
};

#include "batch.hpp"

template <int MODEL, int D>
int dispatch_init(dacti_handle *h) {
    return Ops<MODEL, D>::init(h);
}


}  // namespace dacti_rt
