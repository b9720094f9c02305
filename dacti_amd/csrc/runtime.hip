// runtime.hip -- host runtime and C ABI of the MI355X leaf-update engine.
//
// Pool allocation, construction, garbage collection, batches and the C ABI.
// The integrate() schedule itself (impl/integrate.hpp:11-87) is in
// batch.hpp: a single integrator is stepped as a batch of one.  The host
// keeps only scalars per instance: the LTS counters and times, the per-level
// cluster sizes and the node count.  The model-specific code is compiled once
// per model in variant_*.hip.
#include "runtime.hpp"

namespace dacti_rt {

thread_local std::string g_create_error = "";

std::vector<Field> fields_of(dacti_handle *h, Dev &d) {
    const Sizes &s = h->sz;
    return {
        {(void **) &d.parent, 4, 1},         {(void **) &d.child_base, 4, 1},
        {(void **) &d.nbr, 4, s.NFACE},      {(void **) &d.level, 1, 1},
        {(void **) &d.acti, 1, 1},           {(void **) &d.cluster, 1, 1},
        {(void **) &d.buffer, 1, 1},         {(void **) &d.flags, 1, 1},
        {(void **) &d.bc, 1, s.NFACE},       {(void **) &d.mark, 1, 1},
        {(void **) &d.nextc, 1, 1},
        {(void **) &d.center, 8, s.D},       {(void **) &d.size, 8, 1},
        {(void **) &d.t, 8, 1},              {(void **) &d.p, 8, s.NE},
        {(void **) &d.u, 8, s.NE},           {(void **) &d.rhs, 8, s.NE},
        {(void **) &d.tot, 8, s.NF},         {(void **) &d.k, 8, s.NPX},
        {(void **) &d.t_slope, 8, s.NGX},    {(void **) &d.slope, 8, s.D * s.NE},
        {(void **) &d.grad, 8, s.D * s.NGX}, {(void **) &d.acc, 8, s.NFACE * s.NF},
        {(void **) &d.dep, 8, s.D * s.NF},
        {(void **) &d.error, 8, 1},          {(void **) &d.ctime, 8, 1},
        {(void **) &d.max_dt, 8, 1},         {(void **) &d.cfl, 8, 1},
        {(void **) &d.scratch, 8, 1},        {(void **) &d.brec, 4, 1},
        {(void **) &d.meta, 4, 1},
    };
}

void free_dev(dacti_handle *h, Dev &d) {
    for (auto &f : fields_of(h, d)) {
        if (*f.ptr) (void) hipFree(*f.ptr);
        *f.ptr = nullptr;
    }
}

// allocate a pool of capacity C; when `old` is given, copy its first n nodes
int alloc_pool(dacti_handle *h, int64_t C, Dev *old, int64_t n) {
    Dev nd{};
    nd.C = C;
    std::vector<Field> nf = fields_of(h, nd);
    std::vector<Field> of;
    if (old) of = fields_of(h, *old);
    for (size_t k = 0; k < nf.size(); ++k) {
        const size_t bytes = nf[k].esz * (size_t) nf[k].ncomp * (size_t) C;
        hipError_t e = hipMalloc(nf[k].ptr, bytes);
        if (e != hipSuccess) {
            h->err = "device allocation of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e);
            free_dev(h, nd);
            return DACTI_E_MEMORY;
        }
        HIPCHK(h, hipMemsetAsync(*nf[k].ptr, 0, bytes, h->stream));
        if (old && n > 0) {
            for (int cmp = 0; cmp < nf[k].ncomp; ++cmp) {
                char *dst = (char *) *nf[k].ptr + nf[k].esz * (size_t) cmp * (size_t) C;
                const char *src = (const char *) *of[k].ptr + of[k].esz * (size_t) cmp * (size_t) old->C;
                HIPCHK(h, hipMemcpyAsync(dst, src, nf[k].esz * (size_t) n, hipMemcpyDeviceToDevice, h->stream));
            }
        }
    }
    // initialise the per-node bytes that mean "none"
    HIPCHK(h, hipMemsetAsync(nd.cluster + n, 0xff, (size_t) (C - n), h->stream));
    HIPCHK(h, hipMemsetAsync(nd.buffer + n, 0xff, (size_t) (C - n), h->stream));
    HIPCHK(h, hipMemsetAsync(nd.acti + n, 0xff, (size_t) (C - n), h->stream));
    if (old) {
        HIPCHK(h, hipStreamSynchronize(h->stream));
        free_dev(h, *old);
    }
    // scratch buffers sized by capacity
    if (h->d_list) (void) hipFree(h->d_list);
    if (h->d_list2) (void) hipFree(h->d_list2);
    h->d_list2 = nullptr;
    if (h->d_off) (void) hipFree(h->d_off);
    if (h->d_rn) (void) hipFree(h->d_rn);
    if (h->d_tmp) (void) hipFree(h->d_tmp);
    h->d_list = nullptr;
    h->d_off = nullptr;
    h->d_rn = nullptr;
    h->d_tmp = nullptr;
    HIPCHK(h, hipMalloc(&h->d_list, sizeof(int32_t) * (size_t) C));
    HIPCHK(h, hipMalloc(&h->d_list2, sizeof(int32_t) * (size_t) C));
    HIPCHK(h, hipMalloc(&h->d_off, sizeof(int32_t) * (size_t) C));
    HIPCHK(h, hipMalloc(&h->d_rn, (size_t) C));
    size_t b1 = 0, b2 = 0, b3 = 0;
    hipcub::CountingInputIterator<int32_t> it(1);
    struct Always {
        __host__ __device__ bool operator()(const int32_t &) const { return true; }
    };
    HIPCHK(h, hipcub::DeviceSelect::If(nullptr, b1, it, h->d_list, h->d_nsel, (int) C, Always{}, h->stream));
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(nullptr, b2, h->d_rn, h->d_off, (int) C, h->stream));
    HIPCHK(h, hipcub::DeviceReduce::Max(nullptr, b3, (double *) nd.scratch, (double *) nd.tot, (int) C, h->stream));
    h->tmp_bytes = std::max(b1, std::max(b2, b3)) + 1024;
    HIPCHK(h, hipMalloc(&h->d_tmp, h->tmp_bytes));
    nd.err = (GMEM int32_t *) &h->d_status_home->dev_error;
    nd.brecs = (const GMEM BndRec *) h->d_brecs;
    h->d = nd;
    return DACTI_OK;
}

int ensure_capacity(dacti_handle *h, int64_t need) {
    if (need <= h->d.C) return DACTI_OK;
    int64_t C = h->d.C;
    while (C < need) C = C + C / 2 + 4096;
    Dev old = h->d;
    return alloc_pool(h, C, &old, h->n);
}



int refresh_counts(dacti_handle *h) {
    HIPCHK(h, hipMemsetAsync(h->d_hist, 0, sizeof(unsigned long long) * 16, h->stream));
    const int64_t nb = std::min<int64_t>(blocks(h->n, 256), 4096);
    k_level_hist<<<(unsigned) std::max<int64_t>(nb, 1), 256, 0, h->stream>>>(h->d.cluster, h->n, h->d_hist);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(h->h_hist, h->d_hist, sizeof(unsigned long long) * 16, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int l = 0; l < 16; ++l) h->counts[l] = (int64_t) h->h_hist[l];
    return DACTI_OK;
}

// ------------------------------------------------------------------
// per-model dispatch
// garbage collection: domain.hpp:327-356, impl/acti_utils.hpp:58-74
__global__ void k_gc_keep(Dev d, int64_t n, uint8_t *keep) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keep[i] = (i == 0) ? 1 : ((d.flags[i] & F_GARBAGE) ? 0 : 1);
}

template <typename T>
__global__ void k_gc_move(const T *__restrict__ src, T *__restrict__ dst, int64_t n, int64_t Cs, int64_t Cd, int ncomp,
                          const uint8_t *keep, const int32_t *o2n) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const int64_t j = o2n[i];
    for (int c = 0; c < ncomp; ++c) dst[c * Cd + j] = src[c * Cs + i];
}

__global__ void k_gc_remap(Dev d, int64_t n, int nface, const uint8_t *keep_old, const int32_t *o2n) {
    const int64_t j = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (j < 1 || j >= n) return;
    auto m = [&](int32_t o) -> int32_t { return (o > 0 && keep_old[o]) ? o2n[o] : 0; };
    d.parent[j] = m(d.parent[j]);
    d.child_base[j] = d.child_base[j] ? m(d.child_base[j]) : 0;
    for (int f = 0; f < nface; ++f) d.nbr[f * d.C + j] = m(d.nbr[f * d.C + j]);
    d.cluster[j] = -1;
}

int gc(dacti_handle *h) {
    const int64_t n = h->n;
    uint8_t *keep = nullptr;
    int32_t *o2n = nullptr;
    HIPCHK(h, hipMalloc(&keep, (size_t) n));
    HIPCHK(h, hipMalloc(&o2n, sizeof(int32_t) * (size_t) n));
    k_gc_keep<<<blocks(n), 256, 0, h->stream>>>(h->d, n, keep);
    size_t bytes = 0;
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, keep, o2n, (int) n, h->stream));
    void *tmp = nullptr;
    HIPCHK(h, hipMalloc(&tmp, bytes));
    HIPCHK(h, hipcub::DeviceScan::ExclusiveSum(tmp, bytes, keep, o2n, (int) n, h->stream));
    int32_t last_off = 0;
    uint8_t last_keep = 0;
    HIPCHK(h, hipMemcpyAsync(&last_off, o2n + n - 1, 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(&last_keep, keep + n - 1, 1, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int64_t n_new = last_off + last_keep;
    Dev nd{};
    nd.C = h->d.C;
    std::vector<Field> nf = fields_of(h, nd), of = fields_of(h, h->d);
    for (size_t k = 0; k < nf.size(); ++k) {
        const size_t bytes_k = nf[k].esz * (size_t) nf[k].ncomp * (size_t) nd.C;
        hipError_t e = hipMalloc(nf[k].ptr, bytes_k);
        if (e != hipSuccess) {
            h->err = "gc allocation failed";
            free_dev(h, nd);
            return DACTI_E_MEMORY;
        }
        HIPCHK(h, hipMemsetAsync(*nf[k].ptr, 0, bytes_k, h->stream));
        switch (nf[k].esz) {
        case 1:
            k_gc_move<uint8_t><<<blocks(n), 256, 0, h->stream>>>((const uint8_t *) *of[k].ptr, (uint8_t *) *nf[k].ptr, n,
                                                                 h->d.C, nd.C, nf[k].ncomp, keep, o2n);
            break;
        case 4:
            k_gc_move<int32_t><<<blocks(n), 256, 0, h->stream>>>((const int32_t *) *of[k].ptr, (int32_t *) *nf[k].ptr, n,
                                                                 h->d.C, nd.C, nf[k].ncomp, keep, o2n);
            break;
        default:
            k_gc_move<double><<<blocks(n), 256, 0, h->stream>>>((const double *) *of[k].ptr, (double *) *nf[k].ptr, n,
                                                                h->d.C, nd.C, nf[k].ncomp, keep, o2n);
        }
    }
    HIPCHK(h, hipMemsetAsync(nd.cluster + n_new, 0xff, (size_t) (nd.C - n_new), h->stream));
    HIPCHK(h, hipMemsetAsync(nd.buffer + n_new, 0xff, (size_t) (nd.C - n_new), h->stream));
    HIPCHK(h, hipMemsetAsync(nd.acti + n_new, 0xff, (size_t) (nd.C - n_new), h->stream));
    nd.err = (GMEM int32_t *) &h->d_status_home->dev_error;
    nd.brecs = (const GMEM BndRec *) h->d_brecs;
    k_gc_remap<<<blocks(n_new), 256, 0, h->stream>>>(nd, n_new, h->sz.NFACE, keep, o2n);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    free_dev(h, h->d);
    h->d = nd;
    h->n = n_new;
    (void) hipFree(keep);
    (void) hipFree(o2n);
    (void) hipFree(tmp);
    k_assign<<<blocks(h->n), 256, 0, h->stream>>>(h->d, h->n, h->global_dt, h->max_cfl);
    HIPCHK(h, hipGetLastError());
    return refresh_counts(h);
}


int init_any(dacti_handle *h) {
#define DACTI_CASE_INIT(M, Dm) \
    if (h->model == M && h->D == Dm) return variant_init_##M##_##Dm(h);
    DACTI_VARIANTS(DACTI_CASE_INIT)
    h->err = "unsupported model/dimension";
    return DACTI_E_ARG;
}

int batch_integrate_any(dacti_batch *b, int signal, int *fb) {
#define DACTI_CASE_INT(M, Dm) \
    if (b->model == M && b->D == Dm) return variant_integrate_##M##_##Dm(b, signal, fb);
    DACTI_VARIANTS(DACTI_CASE_INT)
    b->err = "unsupported model/dimension";
    return DACTI_E_ARG;
}

int integrate_any(dacti_handle *h, int signal, int *fb) {
    const int rc = batch_integrate_any(h->solo, signal, fb);
    if (rc) h->err = h->solo->err;
    return rc;
}

// ------------------------------------------------------------------ batches
void batch_drop_graphs(dacti_batch *b) {
    if (b->g_head) (void) hipGraphExecDestroy(b->g_head);
    if (b->g_tail) (void) hipGraphExecDestroy(b->g_tail);
    if (b->g_head_dense) (void) hipGraphExecDestroy(b->g_head_dense);
    if (b->g_step) (void) hipGraphExecDestroy(b->g_step);
    if (b->g_static) (void) hipGraphExecDestroy(b->g_static);
    if (b->g_static_dense) (void) hipGraphExecDestroy(b->g_static_dense);
    b->g_head = b->g_tail = b->g_head_dense = b->g_step = b->g_static = b->g_static_dense = nullptr;
}

void batch_free(dacti_batch *b) {
    if (!b) return;
    if (b->stream) (void) hipStreamSynchronize(b->stream);
    batch_drop_graphs(b);
    for (void *p : {b->d_up, (void *) b->d_mcs, (void *) b->d_tilecnt, b->d_down})
        if (p) (void) hipFree(p);
    for (void *p : {b->h_up, b->h_down})
        if (p) (void) hipHostFree(p);
    for (auto &e : b->ev)
        if (e) (void) hipEventDestroy(e);
    if (b->own_stream && b->stream) (void) hipStreamDestroy(b->stream);
    delete b;
}

int batch_ensure_tiles(dacti_batch *b, int64_t tiles) {
    if (tiles > INT32_MAX / 2) {
        b->err = "batch too large";
        return DACTI_E_ARG;
    }
    if (tiles <= b->tilecnt_cap) return DACTI_OK;
    // the captured graphs hold the old table
    if (hipStreamSynchronize(b->stream) != hipSuccess) return DACTI_E_HIP;
    batch_drop_graphs(b);
    if (b->d_tilecnt) (void) hipFree(b->d_tilecnt);
    b->d_tilecnt = nullptr;
    const int64_t cap = tiles + tiles / 2 + 64;
    if (hipMalloc(&b->d_tilecnt, sizeof(int) * (size_t) cap) != hipSuccess) {
        b->tilecnt_cap = 0;
        b->err = "tile table allocation failed";
        return DACTI_E_MEMORY;
    }
    b->tilecnt_cap = cap;
    return DACTI_OK;
}

// a batch over handles of one model; stream: the stream to run on (a new
// one when null)
int batch_make(dacti_handle *const *hs, int n, hipStream_t stream, dacti_batch **out, std::string &err) {
    dacti_batch *b = new dacti_batch();
    b->hs.assign(hs, hs + n);
    b->broken.assign((size_t) n, 0);
    b->n = n;
    b->model = hs[0]->model;
    b->D = hs[0]->D;
    if (stream) {
        b->stream = stream;
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
            err = "stream creation failed";
            delete b;
            return DACTI_E_HIP;
        }
        b->own_stream = true;
    }
    const size_t n1 = (size_t) n + 1;
    // the per-substep uploads (items and four tile tables) and downloads
    // (status and level histograms) each sit in one block with the same
    // layout on both sides, so each direction is one copy
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t o_toff = al(sizeof(Item) * (size_t) n), o_toffp = o_toff + al(sizeof(int) * n1),
                 o_toffr = o_toffp + al(sizeof(int) * n1), o_toff2 = o_toffr + al(sizeof(int) * n1);
    b->up_bytes = o_toff2 + sizeof(int) * n1;
    const size_t o_hist = al(sizeof(Status) * (size_t) n);
    b->down_bytes = o_hist + sizeof(unsigned long long) * 16 * (size_t) n;
    bool ok = hipMalloc(&b->d_up, b->up_bytes) == hipSuccess && hipHostMalloc(&b->h_up, b->up_bytes) == hipSuccess &&
              hipMalloc(&b->d_down, b->down_bytes) == hipSuccess &&
              hipHostMalloc(&b->h_down, b->down_bytes) == hipSuccess &&
              hipMalloc(&b->d_mcs, sizeof(MConst) * (size_t) n) == hipSuccess;
    if (ok) {
        char *du = (char *) b->d_up, *hu = (char *) b->h_up, *dd = (char *) b->d_down, *hd = (char *) b->h_down;
        b->d_items = (Item *) du;
        b->h_items = (Item *) hu;
        b->d_toff = (int *) (du + o_toff);
        b->h_toff = (int *) (hu + o_toff);
        b->d_toffp = (int *) (du + o_toffp);
        b->h_toffp = (int *) (hu + o_toffp);
        b->d_toffr = (int *) (du + o_toffr);
        b->h_toffr = (int *) (hu + o_toffr);
        b->d_toff2 = (int *) (du + o_toff2);
        b->h_toff2 = (int *) (hu + o_toff2);
        b->d_st = (Status *) dd;
        b->h_st = (Status *) hd;
        b->d_hist = (unsigned long long *) (dd + o_hist);
        b->h_hist = (unsigned long long *) (hd + o_hist);
        std::memset(b->h_up, 0, b->up_bytes);
    }
    for (auto &e : b->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    if (!ok) {
        err = "batch allocation failed";
        batch_free(b);
        return DACTI_E_MEMORY;
    }
    for (int i = 0; i < n && ok; ++i)
        ok = hipMemcpy(b->d_mcs + i, &hs[i]->mc, sizeof(MConst), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        err = "constant upload failed";
        batch_free(b);
        return DACTI_E_HIP;
    }
    std::memset(b->h_toff, 0, sizeof(int) * n1);
    std::memset(b->h_toffp, 0, sizeof(int) * n1);
    std::memset(b->h_toffr, 0, sizeof(int) * n1);
    std::memset(b->h_toff2, 0, sizeof(int) * n1);
    std::memset((void *) b->h_st, 0, sizeof(Status) * (size_t) n);
    *out = b;
    return DACTI_OK;
}

void destroy_handle(dacti_handle *h) {
    if (!h) return;
    if (h->stream) (void) hipStreamSynchronize(h->stream);
    batch_free(h->solo);
    free_dev(h, h->d);
    for (void *p : {(void *) h->d_list, (void *) h->d_list2, (void *) h->d_off, (void *) h->d_rn, (void *) h->d_nsel, h->d_tmp,
                    (void *) h->d_hist, (void *) h->d_status, (void *) h->d_brecs})
        if (p) (void) hipFree(p);
    if (h->h_status) (void) hipHostFree(h->h_status);
    if (h->h_nsel) (void) hipHostFree(h->h_nsel);
    if (h->h_hist) (void) hipHostFree(h->h_hist);
    for (auto &e : h->ev)
        if (e) (void) hipEventDestroy(e);
    if (h->stream) (void) hipStreamDestroy(h->stream);
    delete h;
}

template <typename T>
int upload(dacti_handle *h, T *dst, const T *src, int64_t n, int ncomp_src_major, int64_t C) {
    // src is node-major [i * ncomp + c]; dst is component-major [c * C + i]
    std::vector<T> tmp((size_t) n);
    for (int c = 0; c < ncomp_src_major; ++c) {
        for (int64_t i = 0; i < n; ++i) tmp[(size_t) i] = src[(size_t) i * ncomp_src_major + c];
        HIPCHK(h, hipMemcpy(dst + (size_t) c * C, tmp.data(), sizeof(T) * (size_t) n, hipMemcpyHostToDevice));
    }
    return DACTI_OK;
}

template <typename T>
int download(dacti_handle *h, T *dst, const T *src, int64_t n, int ncomp, int64_t C) {
    if (!dst) return DACTI_OK;
    std::vector<T> tmp((size_t) n);
    for (int c = 0; c < ncomp; ++c) {
        HIPCHK(h, hipMemcpy(tmp.data(), src + (size_t) c * C, sizeof(T) * (size_t) n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) dst[(size_t) i * ncomp + c] = tmp[(size_t) i];
    }
    return DACTI_OK;
}

}  // namespace dacti_rt

using namespace dacti_rt;

// ======================================================================
// C ABI
// ======================================================================
extern "C" {

int dacti_model_sizes(int32_t model, int32_t dim, int32_t *out) {
    if (!out || model < 0 || model > 2 || dim < 2 || dim > 3) return DACTI_E_ARG;
    const Sizes s = model_sizes(model, dim);
    out[0] = s.D;
    out[1] = s.NE;
    out[2] = s.NF;
    out[3] = s.NP;
    out[4] = s.NG;
    return DACTI_OK;
}

int dacti_create(const dacti_case *c, const dacti_tree *t, dacti_handle **out) {
    if (!c || !t || !out) {
        g_create_error = "dacti_create: null argument";
        return DACTI_E_ARG;
    }
    *out = nullptr;
    if (c->model < 0 || c->model > 2 || c->dim < 2 || c->dim > 3) {
        g_create_error = "dacti_create: unsupported model/dimension";
        return DACTI_E_ARG;
    }
    if (t->n_nodes < 2 || !t->parent || !t->child_base || !t->tree_level || !t->flags || !t->bc || !t->center ||
        !t->size || !t->p) {
        g_create_error = "dacti_create: incomplete tree";
        return DACTI_E_ARG;
    }
    if (c->n_zones < 0 || c->n_zones > DACTI_MAX_ZONES || c->max_cfl <= 0) {
        g_create_error = "dacti_create: invalid case parameters";
        return DACTI_E_ARG;
    }
    {
        // bc codes: 0, 1 + a domain face, or DACTI_BC_OBJECT with a valid record
        const int nface = 2 * c->dim;
        if (t->n_records < 0 || (t->n_records > 0 && !t->records)) {
            g_create_error = "dacti_create: invalid boundary records";
            return DACTI_E_ARG;
        }
        for (int64_t i = 0; i < t->n_nodes; ++i)
            for (int f = 0; f < nface; ++f) {
                const uint8_t b = t->bc[(size_t) i * nface + f];
                const bool ok = b <= nface || (b == DACTI_BC_OBJECT && t->brec && t->brec[i] >= 0 &&
                                               t->brec[i] < t->n_records);
                if (!ok) {
                    g_create_error = "dacti_create: node " + std::to_string(i) + " face " + std::to_string(f) +
                                     " has bc code " + std::to_string(b) + " without a valid record";
                    return DACTI_E_ARG;
                }
            }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_error = "dacti_create: no HIP device available";
        return DACTI_E_HIP;
    }
    dacti_handle *h = new dacti_handle();
    {
        // windows of the dense flux op start where the deepest level of the
        // initial tree starts (a breadth-first level is in Morton order)
        int lmax = -1;
        int64_t first = 1;
        for (int64_t i = 1; i < t->n_nodes; ++i)
            if (t->tree_level[i] > lmax) lmax = t->tree_level[i], first = i;
        h->walign = (int) (first & 63);
        // a mesh that cannot change: every active leaf on one tree level, at
        // or above every level the case refines to (Scene::can_split,
        // scene/scene.hpp:50-61, is false for all of them), and none that
        // may unsplit or is flagged for a split.  Splits and unsplits are
        // the only way leaves gain those flags or another level.
        bool fixed = true;
        int lev = -1;
        for (int64_t i = 1; i < t->n_nodes && fixed; ++i) {
            const uint8_t f = t->flags[i];
            if (!(f & F_LEAF) || !(f & F_ACTIVE) || (f & F_GARBAGE)) continue;
            if (f & (F_CAN_UNSPLIT | F_SHOULD_SPLIT)) fixed = false;
            if (lev < 0) lev = t->tree_level[i];
            else if (t->tree_level[i] != lev) fixed = false;
        }
        if (lev < c->max_refinement_level) fixed = false;
        for (int z = 0; z < c->n_zones; ++z)
            if (lev < c->zone_level[z]) fixed = false;
        h->static_mesh = fixed;
    }
    h->model = c->model;
    h->D = c->dim;
    h->sz = model_sizes(c->model, c->dim);
    h->side_length = c->side_length;
    h->max_dt_cfg = c->max_dt;
    h->max_cfl = c->max_cfl;
    MConst &m = h->mc;
    m.gamma = c->gamma;
    m.ns_mu = c->ns_mu;
    m.ns_lambda = c->ns_lambda;
    m.ns_cp = c->ns_cp;
    m.ns_R = c->ns_R;
    m.sigma = c->sigma;
    m.gamma1 = c->gamma1;
    m.gamma2 = c->gamma2;
    m.P1 = c->P1;
    m.P2 = c->P2;
    m.cp1 = c->cp1;
    m.cp2 = c->cp2;
    m.cv1 = c->cv1;
    m.cv2 = c->cv2;
    m.mu1 = c->mu1;
    m.mu2 = c->mu2;
    m.lambda1 = c->lambda1;
    m.lambda2 = c->lambda2;
    m.max_cfl = c->max_cfl;
    m.min_error = c->min_error;
    m.max_error = c->max_error;
    for (int i = 0; i < 8; ++i) m.weights[i] = c->weights[i];
    for (int i = 0; i < 3; ++i) {
        m.active_size[i] = c->active_size[i];
        m.periodic[i] = c->periodic[i];
    }
    m.limiter = c->limiter;
    m.max_refinement_level = c->max_refinement_level;
    m.n_zones = c->n_zones;
    for (int f = 0; f < 6; ++f) {
        m.bc_type[f] = c->bc_type[f];
        m.bc_fade_in[f] = c->bc_fade_in[f];
        for (int e = 0; e < 8; ++e) m.bc_p[f][e] = c->bc_p[f][e];
    }
    for (int z = 0; z < 8; ++z) {
        for (int a = 0; a < 3; ++a) {
            m.zone_x0[z][a] = c->zone_x0[z][a];
            m.zone_x1[z][a] = c->zone_x1[z][a];
        }
        m.zone_level[z] = c->zone_level[z];
    }
    // Cell::volume() and CellFace::area() (integrator/cell.hpp:88-90,135-137):
    // a node at tree level l has size side_length * 2^-l exactly.
    for (int l = 0; l < 20; ++l) {
        const double s = std::ldexp(c->side_length, -l);
        m.vol[l] = std::pow(s, (double) c->dim);
        m.area[l] = std::pow(s, (double) (c->dim - 1));
    }
    int rc = DACTI_OK;
    auto fail = [&](int code) {
        g_create_error = h->err;
        destroy_handle(h);
        return code;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        h->err = "stream creation failed";
        return fail(DACTI_E_HIP);
    }
    for (auto &e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            h->err = "event creation failed";
            return fail(DACTI_E_HIP);
        }
    {
        dacti_handle *one[1] = {h};
        if ((rc = batch_make(one, 1, h->stream, &h->solo, h->err))) return fail(rc);
    }
    if (hipMalloc(&h->d_nsel, sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&h->d_hist, sizeof(unsigned long long) * 16) != hipSuccess ||
        hipMalloc(&h->d_status, sizeof(Status)) != hipSuccess ||
        hipHostMalloc(&h->h_status, sizeof(Status)) != hipSuccess ||
        hipHostMalloc(&h->h_nsel, sizeof(int32_t)) != hipSuccess ||
        hipHostMalloc(&h->h_hist, sizeof(unsigned long long) * 16) != hipSuccess) {
        h->err = "small allocation failed";
        return fail(DACTI_E_MEMORY);
    }
    h->d_status_home = h->d_status;
    if (t->n_records > 0) {
        if (hipMalloc(&h->d_brecs, sizeof(BndRec) * (size_t) t->n_records) != hipSuccess ||
            hipMemcpy(h->d_brecs, t->records, sizeof(BndRec) * (size_t) t->n_records, hipMemcpyHostToDevice) !=
                hipSuccess) {
            h->err = "boundary record upload failed";
            return fail(DACTI_E_MEMORY);
        }
    }
    const int64_t n = t->n_nodes;
    h->n = n;
    const int64_t C = n + std::max<int64_t>(n / 2, 65536);
    rc = alloc_pool(h, C, nullptr, 0);
    if (rc) return fail(rc);
    const Sizes &s = h->sz;
    Dev &d = h->d;
    if ((rc = upload<int32_t>(h, d.parent, t->parent, n, 1, C)) ||
        (rc = upload<int32_t>(h, d.child_base, t->child_base, n, 1, C)) ||
        (rc = upload<int8_t>(h, d.level, t->tree_level, n, 1, C)) ||
        (rc = upload<uint8_t>(h, d.flags, t->flags, n, 1, C)) ||
        (rc = upload<uint8_t>(h, d.bc, t->bc, n, s.NFACE, C)) ||
        (rc = upload<double>(h, d.center, t->center, n, s.D, C)) ||
        (rc = upload<double>(h, d.size, t->size, n, 1, C)) || (rc = upload<double>(h, d.p, t->p, n, s.NE, C)))
        return fail(rc);
    {
        std::vector<int32_t> br((size_t) n, -1);
        if (t->brec) std::copy(t->brec, t->brec + n, br.begin());
        if ((rc = upload<int32_t>(h, d.brec, br.data(), n, 1, C))) return fail(rc);
    }
    if (hipMemset(d.cluster, 0xff, (size_t) C) != hipSuccess || hipMemset(d.buffer, 0xff, (size_t) C) != hipSuccess ||
        hipMemset(d.acti, 0xff, (size_t) C) != hipSuccess) {
        h->err = "memset failed";
        return fail(DACTI_E_HIP);
    }
    std::vector<double> minus1((size_t) n, -1.0);
    if (hipMemcpy(d.error, minus1.data(), sizeof(double) * (size_t) n, hipMemcpyHostToDevice) != hipSuccess) {
        h->err = "upload failed";
        return fail(DACTI_E_HIP);
    }
    rc = init_any(h);
    if (rc) return fail(rc);
    if (hipStreamSynchronize(h->stream) != hipSuccess) {
        h->err = "init failed";
        return fail(DACTI_E_HIP);
    }
    *out = h;
    return DACTI_OK;
}

int dacti_integrate(dacti_handle *h, int32_t signal, double *global_dt, int32_t *force_break) {
    if (!h) return DACTI_E_ARG;
    int fb = 0;
    const int rc = integrate_any(h, signal, &fb);
    if (force_break) *force_break = fb;
    if (global_dt) *global_dt = h->global_dt;
    return rc;
}

int dacti_run(dacti_handle *h, int64_t steps, int32_t signal, int64_t *completed) {
    if (!h || steps < 0) return DACTI_E_ARG;
    int64_t s = 0;
    for (; s < steps; ++s) {
        int fb = 0;
        const int rc = integrate_any(h, signal, &fb);
        if (rc) {
            if (completed) *completed = s;
            return rc;
        }
        if (fb) break;
    }
    if (completed) *completed = s;
    return DACTI_OK;
}

int dacti_n_active(dacti_handle *h, int64_t *n) {
    if (!h || !n) return DACTI_E_ARG;
    int64_t a = 0;
    for (int l = 0; l < 16; ++l) a += h->counts[l] + h->pending[l];
    *n = a;
    return DACTI_OK;
}

int dacti_n_nodes(dacti_handle *h, int64_t *n) {
    if (!h || !n) return DACTI_E_ARG;
    *n = h->n;
    return DACTI_OK;
}

int dacti_export_nodes(dacti_handle *h, const dacti_node_view *v) {
    if (!h || !v) return DACTI_E_ARG;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const Sizes &s = h->sz;
    const Dev &d = h->d;
    const int64_t n = h->n, C = d.C;
    int rc;
    if ((rc = download<int32_t>(h, v->parent, d.parent, n, 1, C)) ||
        (rc = download<int32_t>(h, v->child_base, d.child_base, n, 1, C)) ||
        (rc = download<int32_t>(h, v->nbr, d.nbr, n, s.NFACE, C)) ||
        (rc = download<int8_t>(h, v->tree_level, d.level, n, 1, C)) ||
        (rc = download<int8_t>(h, v->acti_level, d.acti, n, 1, C)) ||
        (rc = download<uint8_t>(h, v->flags, d.flags, n, 1, C)) ||
        (rc = download<double>(h, v->center, d.center, n, s.D, C)) ||
        (rc = download<double>(h, v->size, d.size, n, 1, C)) || (rc = download<double>(h, v->t, d.t, n, 1, C)) ||
        (rc = download<double>(h, v->p, d.p, n, s.NE, C)) || (rc = download<double>(h, v->u, d.u, n, s.NE, C)) ||
        (rc = download<double>(h, v->slope, d.slope, n, s.D * s.NE, C)) ||
        (rc = download<double>(h, v->rhs, d.rhs, n, s.NE, C)) || (rc = download<double>(h, v->k, d.k, n, s.NPX, C)) ||
        (rc = download<double>(h, v->grad, d.grad, n, s.D * s.NGX, C)))
        return rc;
    if (v->cluster) {
        std::vector<int8_t> cl((size_t) n), bf((size_t) n);
        HIPCHK(h, hipMemcpy(cl.data(), d.cluster, (size_t) n, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(bf.data(), d.buffer, (size_t) n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) v->cluster[i] = bf[(size_t) i] >= 0 ? (int8_t) (16 + bf[(size_t) i]) : cl[(size_t) i];
    }
    if (v->meta) {
        std::vector<int8_t> ac((size_t) n);
        std::vector<double> er((size_t) n), md((size_t) n), cf((size_t) n);
        HIPCHK(h, hipMemcpy(ac.data(), d.acti, (size_t) n, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(er.data(), d.error, 8 * (size_t) n, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(md.data(), d.max_dt, 8 * (size_t) n, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(cf.data(), d.cfl, 8 * (size_t) n, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) {
            v->meta[i * 4 + 0] = ac[(size_t) i];
            v->meta[i * 4 + 1] = er[(size_t) i];
            v->meta[i * 4 + 2] = md[(size_t) i];
            v->meta[i * 4 + 3] = cf[(size_t) i];
        }
    }
    return DACTI_OK;
}

int dacti_batch_create(dacti_handle *const *handles, int32_t n, dacti_batch **out) {
    if (!handles || n < 1 || !out) {
        g_create_error = "dacti_batch_create: invalid arguments";
        return DACTI_E_ARG;
    }
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!handles[i] || handles[i]->model != handles[0]->model || handles[i]->D != handles[0]->D) {
            g_create_error = "dacti_batch_create: handles must be valid and share model and dimension";
            return DACTI_E_ARG;
        }
        for (int j = 0; j < i; ++j)
            if (handles[j] == handles[i]) {
                g_create_error = "dacti_batch_create: a handle appears twice";
                return DACTI_E_ARG;
            }
        if (hipStreamSynchronize(handles[i]->stream) != hipSuccess) {
            g_create_error = "dacti_batch_create: a handle's stream failed";
            return DACTI_E_HIP;
        }
    }
    std::string err;
    const int rc = batch_make(handles, n, nullptr, out, err);
    if (rc) g_create_error = "dacti_batch_create: " + err;
    return rc;
}

int dacti_batch_integrate(dacti_batch *b, int32_t signal, int32_t *force_break) {
    if (!b) return DACTI_E_ARG;
    int fb = 0;
    const int rc = batch_integrate_any(b, signal, &fb);
    if (force_break) *force_break = fb;
    return rc;
}

int dacti_batch_run(dacti_batch *b, int64_t steps, int32_t signal, int64_t *completed) {
    if (!b || steps < 0) return DACTI_E_ARG;
    int64_t s = 0;
    for (; s < steps; ++s) {
        int fb = 0;
        const int rc = batch_integrate_any(b, signal, &fb);
        if (rc) {
            if (completed) *completed = s;
            return rc;
        }
        if (fb) break;
    }
    if (completed) *completed = s;
    return DACTI_OK;
}

int dacti_batch_set_launch(dacti_batch *b, int32_t block, int32_t grid, int32_t graphs) {
    if (!b || (block != 64 && block != 128 && block != 256) || grid < 0 || (grid % 8) != 0 || (graphs && grid == 0)) {
        if (b) b->err = "dacti_batch_set_launch: block must be 64, 128 or 256, grid a multiple of 8 (non-zero with graphs)";
        return DACTI_E_ARG;
    }
    if (block != b->block || grid != b->grid || (graphs != 0) != b->graphs) batch_drop_graphs(b);
    b->block = block;
    b->grid = grid;
    b->graphs = graphs != 0;
    b->tuning = false;
    return DACTI_OK;
}

int dacti_batch_get_launch(dacti_batch *b, int32_t *block, int32_t *grid, int32_t *graphs) {
    if (!b) return DACTI_E_ARG;
    if (block) *block = b->block;
    if (grid) *grid = b->grid;
    if (graphs) *graphs = b->graphs ? 1 : 0;
    return DACTI_OK;
}

int dacti_batch_autotune(dacti_batch *b) {
    if (!b) return DACTI_E_ARG;
    int cus = 256;
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
        cus = prop.multiProcessorCount;
    const int g8 = (cus + 7) / 8 * 8;
    b->tune_block = {256, 256, 256, 128, 128, 64};
    b->tune_grid = {0, 4 * g8, 16 * g8, 8 * g8, 32 * g8, 16 * g8};
    b->tune_graphs = {0, 1, 1, 1, 1, 1};
    b->tune_ms.assign(b->tune_block.size(), 0.0);
    b->tune_next = 0;
    b->tuning = true;
    return DACTI_OK;
}

int dacti_batch_tune_results(dacti_batch *b, int32_t *n, int32_t *block, int32_t *grid, int32_t *graphs,
                             double *ms) {
    if (!b || !n) return DACTI_E_ARG;
    const int32_t cap = *n;
    *n = (int32_t) b->tune_block.size();
    for (int32_t k = 0; k < *n && k < cap; ++k) {
        if (block) block[k] = b->tune_block[(size_t) k];
        if (grid) grid[k] = b->tune_grid[(size_t) k];
        if (graphs) graphs[k] = b->tune_graphs[(size_t) k];
        if (ms) ms[k] = b->tune_ms[(size_t) k];
    }
    return DACTI_OK;
}

int dacti_batch_get_stats(dacti_batch *b, dacti_stats *s) {
    if (!b || !s) return DACTI_E_ARG;
    std::memset(s, 0, sizeof *s);
    for (dacti_handle *h : b->hs) {
        s->n_nodes += h->n;
        for (int l = 0; l < 16; ++l) s->n_active += h->counts[l] + h->pending[l];
        s->max_acti_level = std::max(s->max_acti_level, h->max_acti_level);
    }
    s->n_substeps = b->n_substeps;
    s->n_global_steps = b->n_global_steps;
    s->cell_updates = b->cell_updates;
    s->flux_launches = b->launches;
    s->flux_kernel_ms = b->flux_ms;
    s->substep_kernel_ms = b->substep_ms;
    s->global_dt = b->hs[0]->global_dt;
    return DACTI_OK;
}

int dacti_batch_set_timing(dacti_batch *b, int32_t on) {
    if (!b) return DACTI_E_ARG;
    b->timing = on != 0;
    return DACTI_OK;
}

int dacti_batch_destroy(dacti_batch *b) {
    batch_free(b);
    return DACTI_OK;
}

const char *dacti_batch_last_error(dacti_batch *b) {
    if (!b) return g_create_error.c_str();
    return b->err.c_str();
}

int dacti_set_launch(dacti_handle *h, int32_t block, int32_t grid, int32_t graphs) {
    if (!h) return DACTI_E_ARG;
    const int rc = dacti_batch_set_launch(h->solo, block, grid, graphs);
    if (rc) h->err = h->solo->err;
    return rc;
}

int dacti_set_substep_hook(dacti_handle *h, dacti_substep_hook fn, void *user) {
    if (!h) return DACTI_E_ARG;
    h->hook = fn;
    h->hook_user = user;
    return DACTI_OK;
}

int dacti_set_dense(dacti_handle *h, int32_t mode) {
    if (!h) return DACTI_E_ARG;
    if (mode < 0 || mode > 2) {
        h->err = "dacti_set_dense: mode must be 0, 1 or 2";
        return DACTI_E_ARG;
    }
    h->dense = mode;
    return DACTI_OK;
}

int dacti_set_remesh(dacti_handle *h, int32_t on) {
    if (!h) return DACTI_E_ARG;
    h->remesh = on != 0;
    return DACTI_OK;
}

int dacti_mark_split(dacti_handle *h, const int32_t *ids, int64_t n) {
    if (!h || (n > 0 && !ids) || n < 0) return DACTI_E_ARG;
    if (n == 0) return DACTI_OK;
    h->static_mesh = false;  // flagged leaves split at the next remesh
    int32_t *d = nullptr;
    HIPCHK(h, hipMalloc(&d, sizeof(int32_t) * (size_t) n));
    hipError_t e = hipMemcpy(d, ids, sizeof(int32_t) * (size_t) n, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        k_mark_split<<<blocks(n), 256, 0, h->stream>>>(h->d, d, n, h->n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void) hipFree(d);
    if (e != hipSuccess) {
        h->err = std::string("dacti_mark_split: ") + hipGetErrorString(e);
        return DACTI_E_HIP;
    }
    return DACTI_OK;
}

int dacti_features(dacti_handle *h, int32_t min_level, const float *mu, const float *sd, float *feat, int32_t *ids,
                   double *centers, int8_t *levels, int64_t cap, int64_t *n, int32_t *n_features) {
    if (!h || !n || !n_features || cap < 0 || (!mu) != (!sd)) return DACTI_E_ARG;
    int rc = DACTI_E_ARG;
#define DACTI_CASE_FEAT(M, Dm) \
    if (h->model == M && h->D == Dm) rc = variant_features_##M##_##Dm(h, min_level, mu, sd, feat, ids, centers, levels, cap, n, \
                                                          n_features);
    DACTI_VARIANTS(DACTI_CASE_FEAT)
    return rc;
}

int dacti_locate_levels(dacti_handle *h, const double *pts, int64_t n, int8_t *levels) {
    if (!h || n < 0 || (n > 0 && (!pts || !levels))) return DACTI_E_ARG;
    int rc = DACTI_E_ARG;
#define DACTI_CASE_LOC(M, Dm) \
    if (h->model == M && h->D == Dm) rc = variant_locate_##M##_##Dm(h, pts, n, levels);
    DACTI_VARIANTS(DACTI_CASE_LOC)
    return rc;
}

int dacti_mark_predicted(dacti_handle *h, const float *logits, const int32_t *ids, int64_t n, int32_t rings,
                         int64_t *n_marked) {
    if (!h || n < 0 || rings < 0 || !n_marked || (n > 0 && (!logits || !ids))) return DACTI_E_ARG;
    h->static_mesh = false;  // flagged leaves split at the next remesh
    int rc = DACTI_E_ARG;
#define DACTI_CASE_MARK(M, Dm) \
    if (h->model == M && h->D == Dm) rc = variant_mark_predicted_##M##_##Dm(h, logits, ids, n, rings, n_marked);
    DACTI_VARIANTS(DACTI_CASE_MARK)
    return rc;
}

int dacti_get_stats(dacti_handle *h, dacti_stats *s) {
    if (!h || !s) return DACTI_E_ARG;
    std::memset(s, 0, sizeof *s);
    s->n_nodes = h->n;
    int64_t a = 0;
    for (int l = 0; l < 16; ++l) a += h->counts[l] + h->pending[l];
    s->n_active = a;
    s->n_substeps = h->n_substeps;
    s->n_global_steps = h->n_global_steps;
    s->cell_updates = h->cell_updates;
    s->flux_launches = h->flux_launches;
    s->flux_kernel_ms = h->solo->flux_ms;
    s->substep_kernel_ms = h->solo->substep_ms;
    s->global_dt = h->global_dt;
    s->max_acti_level = h->max_acti_level;
    return DACTI_OK;
}

int dacti_get_schedule(dacti_handle *h, double *acti_time, int32_t *acti_counter) {
    if (!h) return DACTI_E_ARG;
    for (int l = 0; l < DACTI_MAX_LEVEL; ++l) {
        if (acti_time) acti_time[l] = h->acti_time[l];
        if (acti_counter) acti_counter[l] = h->acti_counter[l];
    }
    return DACTI_OK;
}

int dacti_set_timing(dacti_handle *h, int32_t on) {
    if (!h) return DACTI_E_ARG;
    h->timing = on != 0;
    h->solo->timing = on != 0;
    return DACTI_OK;
}

int dacti_destroy(dacti_handle *h) {
    destroy_handle(h);
    return DACTI_OK;
}

const char *dacti_last_error(dacti_handle *h) {
    if (!h) return g_create_error.c_str();
    return h->err.c_str();
}

}  // extern "C"
