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

    static int launch_head(dacti_batch *b) {
        const hipStream_t st = b->stream;
        const unsigned gl = list_grid(b), gp = pool_grid(b);
        const int B = b->block, n = b->n;
        Item *it = b->d_items;
        k_sel_count<<<gp, 256, 0, st>>>(it, b->d_toffp, n, b->d_tilecnt);
        k_sel_scan<<<1, 1024, 0, st>>>(it, b->d_toffp, n, b->d_tilecnt);
        k_sel_scatter<<<gp, 256, 0, st>>>(it, b->d_toffp, n, b->d_tilecnt);
        k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpFluxSimple<MODEL, D>{});
        k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpFluxRest<MODEL, D>{});
        k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpUpdate<MODEL, D>{});
        k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpSweep<MODEL, D>{0});
        if constexpr (MODEL == MULTI) {
            for (int r = 0; r < 3; ++r) {
                k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpSmoothAlpha<MODEL, D>{});
                k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpStoreK{S::IALPHAS});
            }
            k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpNormal<MODEL, D>{});
            k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpCurvature<MODEL, D>{});
            k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpSmoothCurvature<MODEL, D>{});
            k_batch<<<gl, B, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpStoreK{S::IKAPPA});
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            b->err = std::string("head launch: ") + hipGetErrorString(e);
            return DACTI_E_HIP;
        }
        e = hipMemcpyAsync(b->h_st, b->d_st, sizeof(Status) * (size_t) n, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) {
            b->err = std::string("status copy: ") + hipGetErrorString(e);
            return DACTI_E_HIP;
        }
        return DACTI_OK;
    }

    static int launch_tail(dacti_batch *b) {
        const hipStream_t st = b->stream;
        const unsigned gl = list_grid(b), gp = pool_grid(b);
        const int n = b->n;
        Item *it = b->d_items;
        k_batch<<<gl, b->block, 0, st>>>(it, b->d_mcs, b->d_toff, n, OpApply{});
        k_pool_flush<<<gp, 256, 0, st>>>(it, b->d_toffp, n);
        k_zero_hist<<<(unsigned) n, 64, 0, st>>>(it, n);
        k_pool_hist<<<std::min(gp, 1024u), 256, 0, st>>>(it, b->d_toffp, n);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->h_hist, b->d_hist, sizeof(unsigned long long) * 16 * (size_t) n,
                               hipMemcpyDeviceToHost, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->h_st, b->d_st, sizeof(Status) * (size_t) n, hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) {
            b->err = std::string("tail launch: ") + hipGetErrorString(e);
            return DACTI_E_HIP;
        }
        return DACTI_OK;
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
            b->h_items[i].n = h->n;
            b->h_items[i].list = h->d_list;
            b->h_items[i].rest = h->d_list2;
            b->h_toffp[i] = (int) tp;
            if (changed[(size_t) i]) tp += (h->n + POOL_TILE - 1) / POOL_TILE;
        }
        b->h_toffp[b->n] = (int) tp;
        int rc = batch_ensure_tiles(b, tp);
        if (rc) return rc;
        const size_t n1 = (size_t) b->n + 1;
        hipError_t e = hipMemcpyAsync(b->d_items, b->h_items, sizeof(Item) * (size_t) b->n, hipMemcpyHostToDevice,
                                      b->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->d_toff, b->h_toff, sizeof(int) * n1, hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->d_toffp, b->h_toffp, sizeof(int) * n1, hipMemcpyHostToDevice, b->stream);
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

    // Split / unsplit application and neighbour update
    // (impl/adaptive_refinement.hpp:97-262), then the LTS decision of the
    // cells that were marked or created, for every instance in ids.  Each
    // phase is launched for all instances before the one wait that its
    // decisions need.
    static int remesh_all(dacti_batch *b, int L, const std::vector<int> &ids) {
        const hipStream_t st = b->stream;
        for (int i : ids) {
            dacti_handle *h = b->hs[(size_t) i];
            const int nl = b->h_items[i].nl;
            k_mark2<D><<<blocks(nl), 256, 0, st>>>(h->d, h->d_list, nl, L, h->d_rn, b->d_st + i);
        }
        if (hipGetLastError() != hipSuccess) {
            b->err = "remesh launch failed";
            return DACTI_E_HIP;
        }
        int rc = copy_status(b);
        if (rc) return rc;
        std::vector<char> topo((size_t) b->n, 0);
        for (int i : ids) {
            Bind bind(b, i);
            dacti_handle *h = bind.h;
            const int nl = b->h_items[i].nl;
            const int n_ready = b->h_st[i].n_ready_split;
            const int n_new = b->h_st[i].n_new_blocks;
            if (n_ready > 0) {
                size_t bytes = h->tmp_bytes;
                HIPCHK(b, hipcub::DeviceScan::ExclusiveSum(h->d_tmp, bytes, h->d_rn, h->d_off, nl, st));
                const int64_t need = h->n + (int64_t) S::NCH * n_new;
                if (need > h->d.C) {
                    // keep the list across the reallocation
                    std::vector<int32_t> keep((size_t) nl), offs((size_t) nl);
                    HIPCHK(b, hipMemcpyAsync(keep.data(), h->d_list, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, st));
                    HIPCHK(b, hipMemcpyAsync(offs.data(), h->d_off, sizeof(int32_t) * nl, hipMemcpyDeviceToHost, st));
                    HIPCHK(b, hipStreamSynchronize(st));
                    rc = ensure_capacity(h, need + need / 4);
                    if (rc) {
                        fail_from(b, i);
                        return rc;
                    }
                    HIPCHK(b, hipMemcpyAsync(h->d_list, keep.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, st));
                    HIPCHK(b, hipMemcpyAsync(h->d_off, offs.data(), sizeof(int32_t) * nl, hipMemcpyHostToDevice, st));
                    HIPCHK(b, hipStreamSynchronize(st));
                }
                k_split_slopes<MODEL, D><<<blocks(nl), 256, 0, st>>>(h->d, h->mc, h->d_list, nl);
                k_split<MODEL, D><<<blocks(nl), 256, 0, st>>>(h->d, h->mc, h->d_list, nl, h->d_off, h->n, L);
                h->n += (int64_t) S::NCH * n_new;
                topo[(size_t) i] = 1;
            }
            k_unsplit<MODEL, D><<<blocks(nl), 256, 0, st>>>(h->d, h->mc, h->d_list, nl, L, b->d_st + i);
            HIPCHK(b, hipGetLastError());
        }
        rc = copy_status(b);
        if (rc) return rc;
        // cluster L now: the listed cells (split parents and coarsened
        // children still in it, to be dropped), new children, unsplit parents
        for (int i : ids) {
            dacti_handle *h = b->hs[(size_t) i];
            if (topo[(size_t) i] || b->h_st[i].n_unsplit > 0)
                k_neighbors<D><<<blocks(h->n), 256, 0, st>>>(h->d, h->mc, h->n);
            hipcub::CountingInputIterator<int32_t> it(1);
            size_t bytes = h->tmp_bytes;
            HIPCHK(b, hipcub::DeviceSelect::If(h->d_tmp, bytes, it, h->d_list2, h->d_nsel, (int) (h->n - 1),
                                               ClusterIs{h->d.cluster, L}, st));
            HIPCHK(b, hipMemcpyAsync(h->h_nsel, h->d_nsel, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        }
        rc = sync(b);
        if (rc) return rc;
        for (int i : ids) {
            dacti_handle *h = b->hs[(size_t) i];
            const int n2 = *h->h_nsel;
            if (n2 > 0)
                k_reassign<MODEL, D><<<blocks(n2), 256, 0, st>>>(h->d, h->mc, h->d_list2, n2, L, b->h_items[i].dt,
                                                                 h->acti_counter[L], b->d_st + i);
        }
        HIPCHK(b, hipGetLastError());
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
        if (any_changed) {
            rc = upload_tables(b, changed);
            if (rc) return rc;
        }
        if (any_list) {
            if (b->timing && hipEventRecord(b->ev[0], b->stream) != hipSuccess) return DACTI_E_HIP;
            rc = run_seq(b, &b->g_head, launch_head);
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
                k_batch<<<list_grid(b), b->block, 0, b->stream>>>(b->d_items, b->d_mcs, b->d_toff, n, OpSweep<MODEL, D>{1});
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
                rc = remesh_all(b, L, marked);
                if (rc) return rc;
                // pools may have grown or moved: new items and pool tiles
                rc = sync(b);
                if (rc) return rc;
                rc = upload_tables(b, changed);
                if (rc) return rc;
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
            float a = 0, c = 0;
            if (hipEventElapsedTime(&a, b->ev[0], b->ev[1]) != hipSuccess ||
                hipEventElapsedTime(&c, b->ev[2], b->ev[3]) != hipSuccess)
                return DACTI_E_HIP;
            b->flux_ms += a;
            b->substep_ms += a + c;
        }
        int top = 0;
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
            for (int l = 15; l > top; --l)
                if (h->counts[l] > 0 || h->pending[l] > 0 || l <= h->max_acti_level) {
                    top = l;
                    break;
                }
        }
        b->top = top;
        b->n_substeps++;
        return 0;
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
        b->n_global_steps++;
        const auto c1 = std::chrono::steady_clock::now();
        tune_end(b, std::chrono::duration<double, std::milli>(c1 - c0).count());
        return DACTI_OK;
    }
};
