// batch.hpp -- the LTS schedule of integrate() (impl/integrate.hpp:11-87)
// for a batch of instances, one launch per op and substep for all of them.
// Included by runtime.hpp inside namespace dacti_rt.
//
// Each instance keeps the reference's per-integrator state (global_dt, the
// acti counters and times, max_acti_level, the per-level cluster sizes) and
// takes part in a substep exactly when its own integrate() would run one, so
// every instance of a batch evolves as it would alone.  Per substep the host
// writes one Item per instance and the tile tables, launches the fixed head
// sequence (list selection, fluxes, update, sweep, interface chain, status
// copy), waits once, runs the rare data-dependent steps (deferred sweep,
// remesh) eagerly, then launches the fixed tail (LTS apply, buffer flush,
// level histogram, copies) and waits again.  With graphs on, head and tail
// are replayed from HIP graphs captured once per launch configuration.

template <int MODEL, int D>
struct BOps {
    using S = Sz<MODEL, D>;

    static unsigned list_grid(const dacti_batch *b) {
        return b->grid > 0 ? (unsigned) b->grid : std::max(1u, (unsigned) b->h_toff[b->n]);
    }
    static unsigned pool_grid(const dacti_batch *b) {
        return b->grid > 0 ? (unsigned) b->grid : std::max(1u, (unsigned) b->h_toffp[b->n]);
    }

    // the window flux op (one instance): a persistent grid, or one block per
    // window of the pool (the kernel strides over the windows of the
    // cluster's id range; a grid that fits on the GPU at once left a tail of
    // idle SIMDs and was 3% slower)
    static unsigned window_grid(const dacti_batch *b) {
        if (b->grid > 0) return 4u * (unsigned) b->grid;
        const dacti_handle *h = b->hs[0];
        const int64_t w = ((h->n - 1 - h->walign) >> 6) + 2;
                return (unsigned) std::min<int64_t>((w + 7) / 8 * 8, int64_t(1) << 22);
    }

    static int launch_head(dacti_batch *b) { return launch_head_ops(b, true); }

    // the sweep with a cell's axes on separate wavefronts (k_sweep_axes) and
    // the coarse deposits with a (cell, face) a thread, for lists of up to
    // 1 M cells and on fixed grids (a captured sequence takes one form for
    // every list); one thread per cell for longer lists, which fill the GPU
    // without it (the 3D shock-drop: sweep 2.5% faster overall, deposits 25%)
    static bool split_small(const dacti_batch *b) {
        return b->grid > 0 || (int64_t) b->h_toff[b->n] * b->block < (int64_t(1) << 20);
    }

    static int launch_head_ops(dacti_batch *b, bool copy_status) {
        const hipStream_t st = b->stream;
        const unsigned gl = list_grid(b), gp = pool_grid(b);
        const int B = b->block, n = b->n;
        Item *it = b->d_items;
        if (b->dense) k_sel_count<0, true><<<gp, 256, 0, st>>>(it, b->d_toffp, n, b->d_tilecnt);
        else k_sel_count<0><<<gp, 256, 0, st>>>(it, b->d_toffp, n, b->d_tilecnt);
        k_sel_scan<0><<<1, 1024, 0, st>>>(it, b->d_toffp, n, b->d_tilecnt);
        k_sel_scatter<0><<<gp, 256, 0, st>>>(it, b->d_toffp, n, b->d_tilecnt);
        // the ops over the set-aside cells stride over the list tiles; next to
        // a window op (one instance, few cells set aside) they take the
        // set-aside cells directly on a bounded grid
        // (a mesh that cannot change has one tree level: nothing is set aside
        // and nothing marked, which the host checks in the status)
        const unsigned gr = std::min(gl, 1024u);
        const bool fixed = b->fixed_mesh;
        if (b->dense) {
            k_flux_window<MODEL, D><<<window_grid(b), 64, 0, st>>>(it, b->d_mcs);
            if (!fixed) {
                k_rest<<<gr, B, 0, st>>>(it, b->d_mcs, n, OpFluxRest<MODEL, D>{});
                k_rest<<<gr, B, 0, st>>>(it, b->d_mcs, n, OpDepositCoarse<MODEL, D>{});
            }
        } else {
            // 64 listed cells a block (the block size is 64, 128 or 256)
            const unsigned ga = b->grid > 0 ? (unsigned) b->grid : std::max(1u, (unsigned) b->h_toff[n] * (unsigned) (B / 64));
            k_flux_axes<MODEL, D><<<ga, 64 * D, 0, st>>>(it, b->d_mcs, b->d_toff, n, B);
            if (!fixed) {
                k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpFluxRest<MODEL, D>{});
                // one instance: a (set-aside cell, face) a thread
                if (n == 1 && split_small(b)) k_rest<<<gl, B, 0, st>>>(it, b->d_mcs, n, OpDepositCoarse<MODEL, D>{});
                else k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpDepositCoarse<MODEL, D>{});
            }
        }
        k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpUpdate<MODEL, D>{});
        if (b->dense) {
            k_sweep_window<MODEL, D><<<window_grid(b), 64, 0, st>>>(it, b->d_mcs);
            if (!fixed) k_rest<<<gr, B, 0, st>>>(it, b->d_mcs, n, OpSweepRest<MODEL, D>{});
        } else {
            if (split_small(b)) {
                const unsigned ga = b->grid > 0 ? (unsigned) b->grid : std::max(1u, (unsigned) b->h_toff[n] * (unsigned) (B / 64));
                k_sweep_axes<MODEL, D><<<ga, 64 * D, 0, st>>>(it, b->d_mcs, b->d_toff, n, B, 0);
            } else {
                k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpSweep<MODEL, D>{0});
            }
        }
        if constexpr (MODEL == MULTI) {
            const unsigned ga = b->grid > 0 ? (unsigned) b->grid : std::max(1u, (unsigned) b->h_toff[n] * (unsigned) (B / 64));
            for (int r = 0; r < 3; ++r) {
                k_axes<D><<<ga, 64 * D, 0, st>>>(it, b->d_mcs, b->d_toff, n, B, OpSmoothAlpha<MODEL, D>{});
                k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpStoreK{S::IALPHAS});
            }
            k_axes<D><<<ga, 64 * D, 0, st>>>(it, b->d_mcs, b->d_toff, n, B, OpNormal<MODEL, D>{});
            k_axes<D><<<ga, 64 * D, 0, st>>>(it, b->d_mcs, b->d_toff, n, B, OpCurvature<MODEL, D>{});
            k_axes<D><<<ga, 64 * D, 0, st>>>(it, b->d_mcs, b->d_toff, n, B, OpSmoothCurvature<MODEL, D>{});
            k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpStoreK{S::IKAPPA});
        }
        // split readiness and unsplit detection of the marked cells
        if (!fixed) k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpMark2<D>{});
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            b->err = std::string("head launch: ") + hipGetErrorString(e);
            return DACTI_E_HIP;
        }
        if (!copy_status) return DACTI_OK;
        e = hipMemcpyAsync(b->h_st, b->d_st, sizeof(Status) * (size_t) n, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) {
            b->err = std::string("status copy: ") + hipGetErrorString(e);
            return DACTI_E_HIP;
        }
        return DACTI_OK;
    }

    static int launch_tail(dacti_batch *b) { return launch_tail_ops(b, true); }

    // zero: clear the level histograms here (k_sel_scan<0> of a head in the
    // same sequence clears them otherwise)
    static int launch_tail_ops(dacti_batch *b, bool zero) {
        const hipStream_t st = b->stream;
        const unsigned gl = list_grid(b), gp = pool_grid(b);
        const int n = b->n;
        Item *it = b->d_items;
        k_batch<<<gl, b->block, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpApply{});
        k_pool_flush<<<gp, 256, 0, st>>>(it, b->d_toffp, n);
        if (zero) k_zero_hist<<<(unsigned) n, 64, 0, st>>>(it, n);
        k_pool_hist<<<std::min(gp, 1024u), 256, 0, st>>>(it, b->d_toffp, n);
        hipError_t e = hipGetLastError();
        // status and level histograms: one block, one copy
        if (e == hipSuccess) e = hipMemcpyAsync(b->h_down, b->d_down, b->down_bytes, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) {
            b->err = std::string("tail launch: ") + hipGetErrorString(e);
            return DACTI_E_HIP;
        }
        return DACTI_OK;
    }

    // A whole level substep with every decision on the device: head, the
    // deferred sweep (acts only where cells were deferred), the remesh plan
    // and phases (act only on instances with marks), tail.  Needs fixed
    // grids (b->grid > 0) and pools with room for every listed cell to split
    // (device_step checks); one wait at the end instead of two or three.
    static int launch_step(dacti_batch *b) {
        const hipStream_t st = b->stream;
        const int n = b->n, B = b->block;
        const unsigned G = (unsigned) b->grid;
        Item *it = b->d_items;
        int rc = launch_head_ops(b, false);
        if (rc) return rc;
        k_sweep_axes<MODEL, D><<<G, 64 * D, 0, st>>>(it, b->d_mcs, b->d_toff, n, B, 1);
        k_clear_split_counts<true><<<(unsigned) ((n + 63) / 64), 64, 0, st>>>(it, n);
        k_batch<<<G, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpMark2<D, true>{});
        k_remesh_plan<<<1, 64, 0, st>>>(it, n, S::NCH, B, b->d_toffr, b->d_toff2, b->d_toffp);
        k_rn_count<<<G, 256, 0, st>>>(it, b->d_toffr, n, b->d_tilecnt);
        k_sel_scan<2><<<1, 1024, 0, st>>>(it, b->d_toffr, n, b->d_tilecnt);
        k_rn_scatter<<<G, 256, 0, st>>>(it, b->d_toffr, n, b->d_tilecnt);
        k_axes<D><<<G, 64 * D, 0, st>>>(it, b->d_mcs, b->d_toff, n, B, OpSplitSlopesAxes<MODEL, D>{});
        k_batch<<<G, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpSplit<MODEL, D>{});
        k_batch<<<G, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpUnsplit<MODEL, D>{});
        k_pool_neighbors<D><<<G, 256, 0, st>>>(it, b->d_mcs, b->d_toffp, n);
        k_sel_count<1><<<G, 256, 0, st>>>(it, b->d_toffp, n, b->d_tilecnt);
        k_sel_scan<1><<<1, 1024, 0, st>>>(it, b->d_toffp, n, b->d_tilecnt);
        k_sel_scatter<1><<<G, 256, 0, st>>>(it, b->d_toffp, n, b->d_tilecnt);
        k_batch<<<G, B, 0, st>>>(it, b->d_mcs, b->d_toff2, n, OpReassign<MODEL, D>{});
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            b->err = std::string("remesh launch: ") + hipGetErrorString(e);
            return DACTI_E_HIP;
        }
        return launch_tail_ops(b, false);
    }

    // Whether a substep can run as launch_step: fixed grids and each
    // participating pool with room for all its listed cells to split (grown
    // here when that growth is small); the tile-count table sized for the
    // pools and lists either way.
    static bool device_step(dacti_batch *b, const std::vector<char> &part, int L, int &rc) {
        rc = 0;
        if (b->grid <= 0) return false;
        int64_t tiles = 0;
        for (int i = 0; i < b->n; ++i) {
            dacti_handle *h = b->hs[(size_t) i];
            if (!part[(size_t) i]) continue;
            const int64_t grow = (int64_t) S::NCH * h->counts[L];
            if (grow > (int64_t(1) << 22)) return false;
            tiles += (h->n + grow + POOL_TILE - 1) / POOL_TILE + (h->counts[L] + 255) / 256 + 1;
        }
        for (int i = 0; i < b->n; ++i) {
            if (!part[(size_t) i]) continue;
            dacti_handle *h = b->hs[(size_t) i];
            const int64_t need = h->n + (int64_t) S::NCH * h->counts[L];
            if (need > h->d.C) {
                Bind bind(b, i);
                rc = ensure_capacity(h, need + need / 2);
                if (rc) {
                    fail_from(b, i);
                    return false;
                }
            }
        }
        rc = batch_ensure_tiles(b, tiles + 64);
        return rc == 0;
    }

    // run a fixed sequence, from a graph captured on first use when graphs are on

    static int run_seq(dacti_batch *b, hipGraphExec_t *g, int (*seq)(dacti_batch *)) {
        if (!b->graphs || b->grid <= 0) return seq(b);
        if (!*g) {
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) {
                b->err = std::string("begin capture: ") + hipGetErrorString(e);
                return DACTI_E_HIP;
            }
            const int rc = seq(b);
            e = hipStreamEndCapture(b->stream, &graph);
            if (rc) return rc;
            if (e != hipSuccess) {
                b->err = std::string("end capture: ") + hipGetErrorString(e);
                return DACTI_E_HIP;
            }
            e = hipGraphInstantiate(g, graph, nullptr, nullptr, 0);
            (void) hipGraphDestroy(graph);
            if (e != hipSuccess) {
                *g = nullptr;
                b->err = std::string("graph instantiate: ") + hipGetErrorString(e);
                return DACTI_E_HIP;
            }
        }
        const hipError_t e = hipGraphLaunch(*g, b->stream);
        if (e != hipSuccess) {
            b->err = std::string("graph launch: ") + hipGetErrorString(e);
            return DACTI_E_HIP;
        }
        return DACTI_OK;
    }

    static int sync(dacti_batch *b) {
        const hipError_t e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) {
            b->err = std::string("stream synchronize: ") + hipGetErrorString(e);
            return DACTI_E_HIP;
        }
        return DACTI_OK;
    }

    // pool tile table (instances whose clusters may change) and the items
    static int upload_tables(dacti_batch *b, const std::vector<char> &changed) {
        int64_t tp = 0;
        for (int i = 0; i < b->n; ++i) {
            dacti_handle *h = b->hs[(size_t) i];
            b->h_items[i].d = h->d;
            b->h_items[i].d.err = (GMEM int32_t *) &b->d_st[i].dev_error;
            b->h_items[i].n = h->n;
            b->h_items[i].list = h->d_list;
            b->h_items[i].rest = h->d_list2;
            b->h_items[i].rn = h->d_rn;
            b->h_items[i].off = h->d_off;
            b->h_toffp[i] = (int) tp;
            if (changed[(size_t) i]) tp += (h->n + POOL_TILE - 1) / POOL_TILE;
        }
        b->h_toffp[b->n] = (int) tp;
        int rc = batch_ensure_tiles(b, tp);
        if (rc) return rc;
        // items and the four tile tables: one block, one copy
        const hipError_t e = hipMemcpyAsync(b->d_up, b->h_up, b->up_bytes, hipMemcpyHostToDevice, b->stream);
        if (e != hipSuccess) {
            b->err = std::string("table upload: ") + hipGetErrorString(e);
            return DACTI_E_HIP;
        }
        return DACTI_OK;
    }

    static int copy_status(dacti_batch *b) {
        if (hipMemcpyAsync(b->h_st, b->d_st, sizeof(Status) * (size_t) b->n, hipMemcpyDeviceToHost, b->stream) !=
            hipSuccess) {
            b->err = "status copy failed";
            return DACTI_E_HIP;
        }
        return sync(b);
    }

    static unsigned grid_for(const dacti_batch *b, int64_t tiles) {
        return b->grid > 0 ? (unsigned) b->grid : (unsigned) std::max<int64_t>(tiles, 1);
    }

    // Split / unsplit application and neighbour update
    // (impl/adaptive_refinement.hpp:97-262), then the LTS decision of the
    // cells that were marked or created, for every instance in ids, each
    // phase one launch over all of them.  One wait: the split counts decide
    // the pool growth.
    static int remesh_all(dacti_batch *b, int L, const std::vector<int> &ids, const std::vector<char> &changed) {
        const hipStream_t st = b->stream;
        const int n = b->n, B = b->block;
        std::vector<char> meshing((size_t) n, 0);
        for (int i : ids) meshing[(size_t) i] = 1;
        int64_t tr = 0;
        for (int i = 0; i < n; ++i) {
            b->h_items[i].meshing = meshing[(size_t) i];
            b->h_toffr[i] = (int) tr;
            if (meshing[(size_t) i]) tr += (b->h_items[i].nl + 255) / 256;
            b->h_toff2[i] = 0;
        }
        b->h_toffr[n] = (int) tr;
        b->h_toff2[n] = 0;
        int rc = 0;
        // split counts from the head (OpMark2); pool growth on the host (rare
        // reallocation), pool size before the splits
        int64_t t2 = 0;
        for (int i = 0; i < n; ++i) {
            b->h_toff2[i] = (int) t2;
            if (!meshing[(size_t) i]) continue;
            dacti_handle *h = b->hs[(size_t) i];
            const int nl = b->h_items[i].nl;
            const int n_ready = b->h_st[i].n_ready_split, n_new = b->h_st[i].n_new_blocks;
            const int64_t need = h->n + (int64_t) S::NCH * n_new;
            if (need > h->d.C) {
                // keep the list and the new-block flags across the reallocation
                Bind bind(b, i);
                std::vector<int32_t> keep((size_t) nl);
                std::vector<uint8_t> rn((size_t) nl);
                HIPCHK(b, hipMemcpyAsync(keep.data(), h->d_list, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, st));
                HIPCHK(b, hipMemcpyAsync(rn.data(), h->d_rn, (size_t) nl, hipMemcpyDeviceToHost, st));
                HIPCHK(b, hipStreamSynchronize(st));
                rc = ensure_capacity(h, need + need / 4);
                if (rc) {
                    fail_from(b, i);
                    return rc;
                }
                HIPCHK(b, hipMemcpyAsync(h->d_list, keep.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, st));
                HIPCHK(b, hipMemcpyAsync(h->d_rn, rn.data(), (size_t) nl, hipMemcpyHostToDevice, st));
                HIPCHK(b, hipStreamSynchronize(st));
            }
            b->h_items[i].n_before = h->n;
            h->n += (int64_t) S::NCH * n_new;
            // the rebuilt list: the listed cells, the new children and at
            // most one unsplit parent per listed cell
            t2 += (2 * (int64_t) nl + (int64_t) S::NCH * n_ready + B - 1) / B;
        }
        b->h_toff2[n] = (int) t2;
        rc = batch_ensure_tiles(b, std::max<int64_t>(tr, 1));
        if (rc) return rc;
        rc = upload_tables(b, changed);  // new pool sizes and pool tiles
        if (rc) return rc;
        // new-block offsets, splits, unsplits
        if (b->timing && hipEventRecord(b->ev[4], st) != hipSuccess) return DACTI_E_HIP;
        const unsigned gr = grid_for(b, tr);
        k_rn_count<<<gr, 256, 0, st>>>(b->d_items, b->d_toffr, n, b->d_tilecnt);
        k_sel_scan<2><<<1, 1024, 0, st>>>(b->d_items, b->d_toffr, n, b->d_tilecnt);
        k_rn_scatter<<<gr, 256, 0, st>>>(b->d_items, b->d_toffr, n, b->d_tilecnt);
        const unsigned gs = b->grid > 0 ? (unsigned) b->grid : std::max(1u, (unsigned) b->h_toff[n] * (unsigned) (B / 64));
        k_axes<D><<<gs, 64 * D, 0, st>>>(b->d_items, b->d_mcs, b->d_toff, n, B, OpSplitSlopesAxes<MODEL, D>{});
        k_batch<<<list_grid(b), B, 0, st>>>(b->d_items, b->d_mcs, b->d_toff, n, OpSplit<MODEL, D>{});
        k_batch<<<list_grid(b), B, 0, st>>>(b->d_items, b->d_mcs, b->d_toff, n, OpUnsplit<MODEL, D>{});
        // neighbours, the rebuilt cluster list, its LTS decisions
        const unsigned gp = pool_grid(b);
        const unsigned gn = b->grid > 0 ? gp : gp * (unsigned) (POOL_TILE / 256 * 2 * D);  // a (node, face) a thread
        k_pool_neighbors<D><<<gn, 256, 0, st>>>(b->d_items, b->d_mcs, b->d_toffp, n);
        k_sel_count<1><<<gp, 256, 0, st>>>(b->d_items, b->d_toffp, n, b->d_tilecnt);
        k_sel_scan<1><<<1, 1024, 0, st>>>(b->d_items, b->d_toffp, n, b->d_tilecnt);
        k_sel_scatter<1><<<gp, 256, 0, st>>>(b->d_items, b->d_toffp, n, b->d_tilecnt);
        k_batch<<<grid_for(b, t2), B, 0, st>>>(b->d_items, b->d_mcs, b->d_toff2, n, OpReassign<MODEL, D>{});
        HIPCHK(b, hipGetLastError());
        if (b->timing && hipEventRecord(b->ev[5], st) != hipSuccess) return DACTI_E_HIP;
        return DACTI_OK;
    }

    static void fail_from(dacti_batch *b, int i) {
        b->err = "instance " + std::to_string(i) + ": " + b->hs[(size_t) i]->err;
    }

    // one level substep of the instances marked in part: impl/integrate.hpp:42-64.
    // An instance whose update produced a NaN state sets *force_break and
    // leaves part (it stops as the reference's integrate() returns).
    static int substep(dacti_batch *b, int L, std::vector<char> &part, int signal, int *force_break) {
        const int n = b->n;
        std::vector<char> changed((size_t) n, 0);
        int64_t tl = 0;
        bool any_list = false, any_changed = false;
        for (int i = 0; i < n; ++i) {
            dacti_handle *h = b->hs[(size_t) i];
            Item it = base_item(b, i);
            it.L = L;
            it.dt = h->global_dt / (1 << L);
            it.counter = h->acti_counter[L];
            it.signal = signal;
            b->h_toff[i] = (int) tl;
            if (part[(size_t) i]) {
                it.nl = (int) h->counts[L];
                it.flush = h->pending[L] > 0;
                tl += (it.nl + b->block - 1) / b->block;
                changed[(size_t) i] = it.nl > 0 || it.flush;
                any_list |= it.nl > 0;
                any_changed |= (bool) changed[(size_t) i];
            }
            b->h_items[i] = it;
        }
        b->h_toff[n] = (int) tl;
        int rc = 0;
        bool remeshed = false;
        bool fixed = any_list;
        for (int i = 0; i < n && fixed; ++i) fixed = b->hs[(size_t) i]->static_mesh;
        b->fixed_mesh = fixed;
        if (fixed) return substep_static(b, L, part, changed, force_break);
        if (any_list && !(n == 1 && b->hs[0]->dense == 2) && device_step(b, part, L, rc))
            return substep_device(b, L, part, changed, force_break);
        if (rc) return rc;
        if (any_changed) {
            rc = upload_tables(b, changed);
            if (rc) return rc;
        }
        if (any_list) {
            if (b->timing && hipEventRecord(b->ev[0], b->stream) != hipSuccess) return DACTI_E_HIP;
            b->dense = false;
            if (n == 1 && b->h_items[0].nl > 0) {
                const dacti_handle *h0 = b->hs[0];
                b->dense = h0->dense == 2 || (h0->dense == 1 && 2 * (int64_t) b->h_items[0].nl >= h0->n);
            }
            rc = run_seq(b, b->dense ? &b->g_head_dense : &b->g_head, launch_head);
            if (rc) return rc;
            if (b->timing && hipEventRecord(b->ev[1], b->stream) != hipSuccess) return DACTI_E_HIP;
            rc = sync(b);
            if (rc) return rc;
            bool deferred = false, nan = false;
            for (int i = 0; i < n; ++i) {
                Item &it = b->h_items[i];
                if (it.nl == 0) continue;
                dacti_handle *h = b->hs[(size_t) i];
                const Status &s = b->h_st[i];
                h->flux_launches++;
                h->cell_updates += it.nl;
                b->cell_updates += it.nl;
                if (s.n_list != it.nl) {
                    b->err = "internal: cluster list size differs from the level histogram";
                    return DACTI_E_STATE;
                }
                if (s.nan_flag) {
                    // this instance stops here: no sweep (the ops skipped it),
                    // no remesh, no LTS apply, no counters
                    nan = true;
                    b->broken[(size_t) i] = 1;
                    part[(size_t) i] = 0;
                    changed[(size_t) i] = 0;
                    it.nl = 0;
                    it.flush = 0;
                    continue;
                }
                deferred |= s.n_deferred > 0;
            }
            b->launches++;
            if (nan) {
                *force_break = 1;
                any_changed = false;
                for (int i = 0; i < n; ++i) any_changed |= (bool) changed[(size_t) i];
                rc = upload_tables(b, changed);
                if (rc) return rc;
            }
            if (deferred) {
                if (split_small(b)) {
                    const unsigned ga = b->grid > 0 ? (unsigned) b->grid
                                                    : std::max(1u, (unsigned) b->h_toff[n] * (unsigned) (b->block / 64));
                    k_sweep_axes<MODEL, D><<<ga, 64 * D, 0, b->stream>>>(b->d_items, b->d_mcs, b->d_toff, n, b->block, 1);
                } else {
                    k_batch<<<list_grid(b), b->block, 0, b->stream>>>(b->d_items, b->d_mcs, b->d_toff, n, OpSweep<MODEL, D>{1});
                }
                k_clear_split_counts<<<(unsigned) ((n + 63) / 64), 64, 0, b->stream>>>(b->d_items, n);
                k_batch<<<list_grid(b), b->block, 0, b->stream>>>(b->d_items, b->d_mcs, b->d_toff, n, OpMark2<D>{});
                if (hipGetLastError() != hipSuccess ||
                    hipMemcpyAsync(b->h_st, b->d_st, sizeof(Status) * (size_t) n, hipMemcpyDeviceToHost, b->stream) !=
                        hipSuccess) {
                    b->err = "deferred sweep launch failed";
                    return DACTI_E_HIP;
                }
                rc = sync(b);
                if (rc) return rc;
            }
            std::vector<int> marked;
            for (int i = 0; i < n; ++i)
                if (b->h_items[i].nl > 0 && b->h_st[i].n_marked > 0) marked.push_back(i);
            if (!marked.empty()) {
                // uploads the grown pools' items and pool tiles itself
                rc = remesh_all(b, L, marked, changed);
                if (rc) return rc;
                remeshed = true;
            }
            any_list = false;
            for (int i = 0; i < n; ++i) any_list |= b->h_items[i].nl > 0;
        }
        if (any_changed) {
            if (b->timing && hipEventRecord(b->ev[2], b->stream) != hipSuccess) return DACTI_E_HIP;
            rc = run_seq(b, &b->g_tail, launch_tail);
            if (rc) return rc;
            if (b->timing && hipEventRecord(b->ev[3], b->stream) != hipSuccess) return DACTI_E_HIP;
            rc = sync(b);
            if (rc) return rc;
        }
        if (b->timing && any_list && any_changed) {
            float a = 0, c = 0, m = 0;
            if (hipEventElapsedTime(&a, b->ev[0], b->ev[1]) != hipSuccess ||
                hipEventElapsedTime(&c, b->ev[2], b->ev[3]) != hipSuccess ||
                (remeshed && hipEventElapsedTime(&m, b->ev[4], b->ev[5]) != hipSuccess))
                return DACTI_E_HIP;
            b->flux_ms += a;
            b->substep_ms += a + c + m;
        }
        return finish_substep(b, L, part, changed);
    }

    // one substep as launch_step (the items were prepared by substep())
    static int substep_device(dacti_batch *b, int L, std::vector<char> &part, std::vector<char> &changed,
                              int *force_break) {
        const int n = b->n;
        int rc = upload_tables(b, changed);
        if (rc) return rc;
        b->dense = false;
        if (b->timing && hipEventRecord(b->ev[0], b->stream) != hipSuccess) return DACTI_E_HIP;
        rc = run_seq(b, &b->g_step, launch_step);
        if (rc) return rc;
        if (b->timing && hipEventRecord(b->ev[1], b->stream) != hipSuccess) return DACTI_E_HIP;
        rc = sync(b);
        if (rc) return rc;
        if (b->timing) {
            float a = 0;
            if (hipEventElapsedTime(&a, b->ev[0], b->ev[1]) != hipSuccess) return DACTI_E_HIP;
            b->flux_ms += a;
            b->substep_ms += a;
        }
        b->launches++;
        for (int i = 0; i < n; ++i) {
            Item &it = b->h_items[i];
            if (it.nl == 0) continue;
            dacti_handle *h = b->hs[(size_t) i];
            const Status &s = b->h_st[i];
            h->flux_launches++;
            h->cell_updates += it.nl;
            b->cell_updates += it.nl;
            if (s.n_list != it.nl) {
                b->err = "internal: cluster list size differs from the level histogram";
                return DACTI_E_STATE;
            }
            if (s.nan_flag) {
                // this instance stops here (the device skipped its sweep,
                // remesh and LTS apply, and its level histogram)
                *force_break = 1;
                b->broken[(size_t) i] = 1;
                part[(size_t) i] = 0;
                changed[(size_t) i] = 0;
                it.nl = 0;
                it.flush = 0;
                continue;
            }
            h->n = s.n_pool;
        }
        return finish_substep(b, L, part, changed);
    }

    // A substep of meshes that cannot change (every instance on one tree
    // level with no cell that may split or unsplit): no cell is deferred
    // (that needs a coarser neighbour) or marked, so head and tail run as
    // one sequence with one wait, on the launch grids of the host-planned
    // path (the window ops for a dense cluster included).
    static int launch_static(dacti_batch *b) {
        const int rc = launch_head_ops(b, false);
        if (rc) return rc;
        return launch_tail_ops(b, false);
    }

    static int substep_static(dacti_batch *b, int L, std::vector<char> &part, std::vector<char> &changed,
                              int *force_break) {
        const int n = b->n;
        int rc = upload_tables(b, changed);
        if (rc) return rc;
        b->dense = false;
        if (n == 1 && b->h_items[0].nl > 0) {
            const dacti_handle *h0 = b->hs[0];
            b->dense = h0->dense == 2 || (h0->dense == 1 && 2 * (int64_t) b->h_items[0].nl >= h0->n);
        }
        if (b->timing && hipEventRecord(b->ev[0], b->stream) != hipSuccess) return DACTI_E_HIP;
        rc = run_seq(b, b->dense ? &b->g_static_dense : &b->g_static, launch_static);
        if (rc) return rc;
        if (b->timing && hipEventRecord(b->ev[1], b->stream) != hipSuccess) return DACTI_E_HIP;
        rc = sync(b);
        if (rc) return rc;
        if (b->timing) {
            float a = 0;
            if (hipEventElapsedTime(&a, b->ev[0], b->ev[1]) != hipSuccess) return DACTI_E_HIP;
            b->flux_ms += a;
            b->substep_ms += a;
        }
        b->launches++;
        for (int i = 0; i < n; ++i) {
            Item &it = b->h_items[i];
            if (it.nl == 0) continue;
            dacti_handle *h = b->hs[(size_t) i];
            const Status &s = b->h_st[i];
            h->flux_launches++;
            h->cell_updates += it.nl;
            b->cell_updates += it.nl;
            if (s.n_list != it.nl) {
                b->err = "internal: cluster list size differs from the level histogram";
                return DACTI_E_STATE;
            }
            if (s.nan_flag) {
                // this instance stops here (its sweep and LTS apply were
                // skipped on the device, its level histogram is not taken)
                *force_break = 1;
                b->broken[(size_t) i] = 1;
                part[(size_t) i] = 0;
                changed[(size_t) i] = 0;
                it.nl = 0;
                it.flush = 0;
                continue;
            }
            if (s.n_deferred > 0 || s.n_marked > 0 || s.n_rest > 0 || s.n_rest2 > 0) {
                b->err = "internal: cells set aside, deferred or marked in a mesh that cannot change";
                return DACTI_E_STATE;
            }
        }
        return finish_substep(b, L, part, changed);
    }

    // the host side of a substep's end: cluster sizes, counters, LTS times,
    // the caller's hook
    static int finish_substep(dacti_batch *b, int L, std::vector<char> &part, std::vector<char> &changed) {
        const int n = b->n;
        for (int i = 0; i < n; ++i) {
            dacti_handle *h = b->hs[(size_t) i];
            if (changed[(size_t) i])
                for (int l = 0; l < 16; ++l) h->counts[l] = (int64_t) b->h_hist[16 * i + l];
            if (part[(size_t) i]) {
                const Item &it = b->h_items[i];
                if (it.nl > 0) {
                    const Status &s = b->h_st[i];
                    if (s.dev_error) {
                        h->err = "a subtree next to a face is more than three levels deeper than the cell (2:1 balance broken)";
                        fail_from(b, i);
                        return DACTI_E_STATE;
                    }
                    if (L > 0) h->pending[L - 1] += s.lowered;
                }
                h->acti_counter[L] = (h->acti_counter[L] + 1) % 2;
                h->acti_time[L] += h->global_dt / (1 << L);
                if (it.flush) h->pending[L] = 0;
                h->n_substeps++;
            }
        }
        update_top(b);
        b->n_substeps++;
        // the caller's per-substep hook (impl/integrate.hpp:64), after the
        // substep of each instance that took part, with its state on the host
        // side consistent (the stream is idle here)
        for (int i = 0; i < n; ++i) {
            dacti_handle *h = b->hs[(size_t) i];
            if (part[(size_t) i] && h->hook) h->hook(h->hook_user, h, L, h->acti_time[L], h->n_substeps, h->global_dt / (1 << L));
        }
        return 0;
    }

    // highest level any instance may run: a level with a cluster, a pending
    // buffer, or at most the instance's max_acti_level

    static void update_top(dacti_batch *b) {
        int top = 0;
        for (int i = 0; i < b->n; ++i) {
            const dacti_handle *h = b->hs[(size_t) i];
            for (int l = 15; l > top; --l)
                if (h->counts[l] > 0 || h->pending[l] > 0 || l <= h->max_acti_level) {
                    top = l;
                    break;
                }
        }
        b->top = top;
    }

    static void tune_begin(dacti_batch *b) {
        if (!b->tuning) return;
        const size_t k = b->tune_next;
        const int blk = b->tune_block[k], grd = b->tune_grid[k];
        const bool gr = b->tune_graphs[k] != 0;
        if (blk != b->block || grd != b->grid || gr != b->graphs) batch_drop_graphs(b);
        b->block = blk;
        b->grid = grd;
        b->graphs = gr;
    }

    static void tune_end(dacti_batch *b, double ms) {
        if (!b->tuning) return;
        b->tune_ms[b->tune_next] = ms;
        if (++b->tune_next < b->tune_block.size()) return;
        size_t best = 0;
        for (size_t k = 1; k < b->tune_ms.size(); ++k)
            if (b->tune_ms[k] < b->tune_ms[best]) best = k;
        b->tuning = false;
        if (b->tune_block[best] != b->block || b->tune_grid[best] != b->grid ||
            (b->tune_graphs[best] != 0) != b->graphs)
            batch_drop_graphs(b);
        b->block = b->tune_block[best];
        b->grid = b->tune_grid[best];
        b->graphs = b->tune_graphs[best] != 0;
    }

    // one global step of every instance: impl/integrate.hpp:11-87
    static int integrate(dacti_batch *b, int signal, int *force_break) {
        *force_break = 0;
        if (b->n == 1) b->broken[0] = 0;
        tune_begin(b);
        const auto c0 = std::chrono::steady_clock::now();
        std::vector<char> part((size_t) b->n, 0);
        for (int seq = 1; seq <= (1 << 15); ++seq) {
            for (int l = 0; l < 16; ++l) {
                if (seq % (1 << l) != 0) break;
                const int L = 15 - l;
                if (L > b->top) continue;
                bool any = false;
                for (int i = 0; i < b->n; ++i) {
                    dacti_handle *h = b->hs[(size_t) i];
                    if (L > h->max_acti_level && h->counts[L] > 0) {
                        h->max_acti_level = L;
                        h->acti_time[L] = h->acti_time[L - 1];
                    }
                    part[(size_t) i] = L <= h->max_acti_level && !b->broken[(size_t) i];
                    any |= (bool) part[(size_t) i];
                }
                if (!any) continue;
                const int rc = substep(b, L, part, signal, force_break);
                if (rc < 0) return rc;
            }
        }
        for (int i = 0; i < b->n; ++i) {
            if (b->broken[(size_t) i]) continue;  // returned before gc (or frozen)
            dacti_handle *h = b->hs[(size_t) i];
            int64_t nact = 0;
            for (int l = 0; l < 16; ++l) nact += h->counts[l];
            h->n_global_steps++;
            if ((double) nact / (double) (h->n - 1) < 0.2) {
                Bind bind(b, i);
                const int rc = gc(h);
                if (rc) {
                    fail_from(b, i);
                    return rc;
                }
            }
        }
        // gc re-binned the levels from the global dt (a leaf dropped by the
        // raise rule comes back on level 15): the next step must see them
        update_top(b);
        b->n_global_steps++;
        const auto c1 = std::chrono::steady_clock::now();
        tune_end(b, std::chrono::duration<double, std::milli>(c1 - c0).count());
        return DACTI_OK;
    }
};
