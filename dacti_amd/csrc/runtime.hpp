// runtime.hpp -- host runtime of the MI355X leaf-update engine (shared by
// the per-model translation units and the C ABI).
//
// The host side restates Integrator (src/lib/integrator/integrator.hpp) and
// its integrate() schedule (impl/integrate.hpp:11-87); every per-cell phase
// runs as a HIP kernel over the level's cluster list (kernels.hpp).  The
// host keeps only scalars: the LTS counters and times, the per-level
// cluster sizes and the node count.
#pragma once
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dacti_gpu.h"
#include "kernels.hpp"

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
    void *d_tmp = nullptr;
    size_t tmp_bytes = 0;
    unsigned long long *d_hist = nullptr;
    Status *d_status = nullptr;
    Status *h_status = nullptr;
    int32_t *h_nsel = nullptr;
    unsigned long long *h_hist = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
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

struct ClusterIs {
    const int8_t *cluster;
    int level;
    __host__ __device__ bool operator()(const int32_t &i) const { return cluster[i] == level; }
};

int alloc_pool(dacti_handle *h, int64_t C, Dev *old, int64_t n);
int ensure_capacity(dacti_handle *h, int64_t need);
int build_list(dacti_handle *h, int L, int64_t *nl, int32_t *out = nullptr);
int build_list_async(dacti_handle *h, int L, int32_t *out, int32_t *count);
int read_status(dacti_handle *h);
int refresh_counts(dacti_handle *h);
int finish_substep(dacti_handle *h);
int gc(dacti_handle *h);

// per-model entry points, one translation unit each (variant_*.hip)
#define DACTI_VARIANTS(X) X(EULER, 2) X(EULER, 3) X(NAVIER_STOKES, 2) X(NAVIER_STOKES, 3) X(MULTI, 2) X(MULTI, 3)
#define DACTI_DECLARE_VARIANT(M, Dm) \
    int variant_init_##M##_##Dm(dacti_handle *h); \
    int variant_integrate_##M##_##Dm(dacti_handle *h, int signal, int *force_break);
DACTI_VARIANTS(DACTI_DECLARE_VARIANT)

template <int MODEL, int D>
struct Ops {
    using S = Sz<MODEL, D>;

    static int init(dacti_handle *h) {
        Dev &d = h->d;
        k_init_cells<MODEL, D><<<blocks(d.C), 256, 0, h->stream>>>(d, h->mc, h->n);
        HIPCHK(h, hipGetLastError());
        size_t bytes = h->tmp_bytes;
        double *d_max = d.tot;  // free until the first substep
        HIPCHK(h, hipcub::DeviceReduce::Max(h->d_tmp, bytes, d.scratch, d_max, (int) d.C, h->stream));
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

    // multiphase interface fields: impl/integrate.hpp:49-54
    static int chain(dacti_handle *h, int nl) {
        if constexpr (MODEL == MULTI) {
            Dev &d = h->d;
            for (int it = 0; it < 3; ++it) {
                k_smooth_alpha<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(d, h->mc, h->d_list, nl);
                k_store_k<0><<<blocks(nl), 256, 0, h->stream>>>(d, h->d_list, nl, S::IALPHAS);
            }
            k_normal<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(d, h->mc, h->d_list, nl);
            k_curvature<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(d, h->mc, h->d_list, nl);
            k_smooth_curvature<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(d, h->mc, h->d_list, nl);
            k_store_k<0><<<blocks(nl), 256, 0, h->stream>>>(d, h->d_list, nl, S::IKAPPA);
            HIPCHK(h, hipGetLastError());
        }
        return DACTI_OK;
    }

    // one level substep: impl/integrate.hpp:42-64; returns 1 on force_break
    static int substep(dacti_handle *h, int L, int signal, int *force_break) {
        const double dt = h->global_dt / (1 << L);
        // the cluster size is known from the level histogram of the last substep
        const int nl = (int) h->counts[L];
        int rc = 0;
        bool changed = false;
        if (nl > 0) {
            HIPCHK(h, hipMemsetAsync(h->d_status, 0, sizeof(Status), h->stream));
            rc = build_list_async(h, L, h->d_list, &h->d_status->n_list);
            if (rc) return rc;
            if (h->timing) HIPCHK(h, hipEventRecord(h->ev[0], h->stream));
            k_flux_simple<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(h->d, h->mc, h->d_list, nl, L, dt, h->d_list2,
                                                                       &h->d_status->n_rest);
            k_flux<MODEL, D><<<std::min<unsigned>(blocks(nl), 1024), 256, 0, h->stream>>>(h->d, h->mc, h->d_list2,
                                                                                      &h->d_status->n_rest, L, dt);
            k_update<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(h->d, h->mc, h->d_list, nl, dt, signal, h->d_status);
            HIPCHK(h, hipGetLastError());
            if (h->timing) HIPCHK(h, hipEventRecord(h->ev[1], h->stream));
            h->flux_launches++;
            h->cell_updates += nl;
            if constexpr (MODEL == MULTI) {
                // the interface chain is not guarded by the NaN flag: stop here
                rc = read_status(h);
                if (rc) return rc;
                if (h->h_status->nan_flag) {
                    *force_break = 1;
                    return 1;
                }
            }
            // slopes, refinement marks and LTS decisions in one sweep (a
            // no-op when the update produced a NaN state)
            k_sweep<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(h->d, h->mc, h->d_list, nl, L, dt,
                                                                 h->acti_counter[L], 0, h->d_status);
            HIPCHK(h, hipGetLastError());
            rc = chain(h, nl);
            if (rc) return rc;
            rc = read_status(h);
            if (rc) return rc;
            if (h->h_status->nan_flag) {
                *force_break = 1;
                return 1;
            }
            if (h->h_status->n_list != nl) {
                h->err = "internal: cluster list size differs from the level histogram";
                return DACTI_E_STATE;
            }
            if (h->h_status->n_deferred > 0) {
                k_sweep<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(h->d, h->mc, h->d_list, nl, L, dt,
                                                                     h->acti_counter[L], 1, h->d_status);
                HIPCHK(h, hipGetLastError());
                rc = read_status(h);
                if (rc) return rc;
            }
            if (h->h_status->n_marked > 0) {
                rc = remesh(h, L, nl, dt);
                if (rc) return rc;
            }
            k_apply<<<blocks(nl), 256, 0, h->stream>>>(h->d, h->d_list, nl, L);
            HIPCHK(h, hipGetLastError());
            if (h->timing) HIPCHK(h, hipEventRecord(h->ev[2], h->stream));
            changed = true;
        }
        h->acti_counter[L] = (h->acti_counter[L] + 1) % 2;
        h->acti_time[L] += h->global_dt / (1 << L);
        if (h->pending[L] > 0) {
            k_flush<<<blocks(h->n), 256, 0, h->stream>>>(h->d, h->n, L);
            HIPCHK(h, hipGetLastError());
            h->pending[L] = 0;
            changed = true;
        }
        if (changed) {
            // one round trip: level histogram, status (dev_error, lowered)
            rc = finish_substep(h);
            if (rc) return rc;
        }
        if (nl > 0) {
            if (h->h_status->dev_error) {
                h->err = "a neighbour subtree is more than three levels deeper than the cell (2:1 balance broken)";
                return DACTI_E_STATE;
            }
            if (h->timing) {
                float a = 0, b = 0;
                HIPCHK(h, hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
                HIPCHK(h, hipEventElapsedTime(&b, h->ev[0], h->ev[2]));
                h->flux_ms += a;
                h->substep_ms += b;
            }
            if (L > 0) h->pending[L - 1] += h->h_status->lowered;
        }
        h->n_substeps++;
        return 0;
    }

    // split / unsplit application and neighbour update (:97-262), then the
    // LTS decision of the cells that were marked or created
    static int remesh(dacti_handle *h, int L, int nl, double dt) {
        k_mark2<D><<<blocks(nl), 256, 0, h->stream>>>(h->d, h->d_list, nl, L, h->d_rn, h->d_status);
        HIPCHK(h, hipGetLastError());
        int rc = read_status(h);
        if (rc) return rc;
        const int n_ready = h->h_status->n_ready_split;
        const int n_new = h->h_status->n_new_blocks;
        bool changed = false;
        if (n_ready > 0) {
            size_t bytes = h->tmp_bytes;
            HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(h->d_tmp, bytes, h->d_rn, h->d_off, nl, h->stream));
            const int64_t need = h->n + (int64_t) S::NCH * n_new;
            if (need > h->d.C) {
                // keep the list across the reallocation
                std::vector<int32_t> keep((size_t) nl), offs((size_t) nl);
                HIPCHK(h, hipMemcpyAsync(keep.data(), h->d_list, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, h->stream));
                HIPCHK(h, hipMemcpyAsync(offs.data(), h->d_off, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, h->stream));
                HIPCHK(h, hipStreamSynchronize(h->stream));
                rc = ensure_capacity(h, need + need / 4);
                if (rc) return rc;
                HIPCHK(h, hipMemcpyAsync(h->d_list, keep.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, h->stream));
                HIPCHK(h, hipMemcpyAsync(h->d_off, offs.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, h->stream));
                HIPCHK(h, hipStreamSynchronize(h->stream));
            }
            k_split_slopes<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(h->d, h->mc, h->d_list, nl);
            k_split<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(h->d, h->mc, h->d_list, nl, h->d_off, h->n, L);
            HIPCHK(h, hipGetLastError());
            h->n += (int64_t) S::NCH * n_new;
            changed = true;
        }
        k_unsplit<MODEL, D><<<blocks(nl), 256, 0, h->stream>>>(h->d, h->mc, h->d_list, nl, L, h->d_status);
        HIPCHK(h, hipGetLastError());
        rc = read_status(h);
        if (rc) return rc;
        if (h->h_status->n_unsplit > 0) changed = true;
        if (changed) {
            k_neighbors<D><<<blocks(h->n), 256, 0, h->stream>>>(h->d, h->mc, h->n);
            HIPCHK(h, hipGetLastError());
        }
        // cluster L now: the listed cells (split parents and coarsened
        // children still in it, to be dropped), new children, unsplit parents
        int64_t n2 = 0;
        rc = build_list(h, L, &n2, h->d_list2);
        if (rc) return rc;
        if (n2 > 0) {
            k_reassign<MODEL, D><<<blocks(n2), 256, 0, h->stream>>>(h->d, h->mc, h->d_list2, (int) n2, L, dt,
                                                                     h->acti_counter[L], h->d_status);
            HIPCHK(h, hipGetLastError());
        }
        return DACTI_OK;
    }
};

template <int MODEL, int D>
int integrate_impl(dacti_handle *h, int signal, int *force_break) {
    *force_break = 0;
    for (int seq = 1; seq <= (1 << 15); ++seq) {
        for (int l = 0; l < 16; ++l) {
            if (seq % (1 << l) != 0) break;
            const int L = 15 - l;
            if (L > h->max_acti_level && h->counts[L] > 0) {
                h->max_acti_level = L;
                h->acti_time[L] = h->acti_time[L - 1];
            }
            if (L > h->max_acti_level) continue;
            const int rc = Ops<MODEL, D>::substep(h, L, signal, force_break);
            if (rc < 0) return rc;
            if (rc == 1) return DACTI_OK;
        }
    }
    int64_t nact = 0;
    for (int l = 0; l < 16; ++l) nact += h->counts[l];
    h->n_global_steps++;
    if ((double) nact / (double) (h->n - 1) < 0.2) return gc(h);
    return DACTI_OK;
}

template <int MODEL, int D>
int dispatch_init(dacti_handle *h) {
    return Ops<MODEL, D>::init(h);
}


}  // namespace dacti_rt
