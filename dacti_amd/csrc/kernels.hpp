// kernels.hpp -- HIP kernels of the octree leaf update for gfx950.
//
// The node pool lives in HBM as structure-of-arrays indexed by node id
// (stride C = capacity): a level-L substep touches only that level's
// cluster, given as a sorted list of node ids, and every array access of a
// wavefront is a gather over consecutive ids.  One thread owns one leaf.
//
// Flux accumulation is a gather: a cell sums its own faces in a fixed
// order; fluxes owed to a cell that is not updated in the current substep
// go to that cell's per-face accumulator acc[face], written by exactly one
// thread (no atomics), so results do not depend on scheduling.
#pragma once

#include "model.hpp"


namespace dacti {

enum : uint8_t {
    F_LEAF = 1, F_ACTIVE = 2, F_GARBAGE = 4, F_CAN_SPLIT = 8, F_CAN_UNSPLIT = 16, F_SHOULD_SPLIT = 32, F_BOUNDARY = 64
};
enum : uint8_t { MK_SPLIT = 1, MK_UNSPLIT_CAND = 2, MK_UNSPLIT_CHILD = 4, MK_READY = 8, MK_DEFER = 16, MK_RA_DONE = 32 };

struct Dev {
    int64_t C;
    GMEM int32_t *parent, *child_base, *nbr;  // nbr[f*C + i]
    GMEM int8_t *level, *acti, *cluster, *buffer;
    GMEM uint8_t *flags, *bc, *mark;          // bc[f*C + i]
    GMEM int8_t *nextc;                       // cluster after this substep's LTS decision
    GMEM double *center;                      // center[d*C + i]
    GMEM double *size, *t;
    GMEM double *p, *u, *rhs, *tot;           // [e*C + i]
    GMEM double *k, *t_slope;                 // [e*C + i]
    GMEM double *slope;                       // [(d*NE + e)*C + i]
    GMEM double *grad;                        // [(d*NGX + e)*C + i]
    GMEM double *acc;                         // [(f*NF + e)*C + i]
    GMEM double *dep;                         // [(d*NF + e)*C + i] flux of a fine cell's face toward a coarser receiver
    GMEM double *error, *ctime, *max_dt, *cfl;
    GMEM double *scratch;                     // C doubles (smoothing / reductions)
    GMEM int32_t *err;                        // device error word (subtree deeper than supported)
    GMEM int32_t *brec;                       // object boundary record of a node's BC_OBJECT faces
    GMEM uint32_t *meta;                      // flags | level << 8 | acti << 16 | cluster << 24, written by the
                                              // list selection of a dense substep for the window flux op
    const GMEM BndRec *brecs;                 // the records (fixed at construction)
};

// device status block read back by the host after a substep phase
struct Status {
    int32_t dev_error;      // set when a subtree is deeper than the unrolled walks
    int32_t nan_flag;
    int32_t n_marked;       // cells with a split or unsplit-candidate mark
    int32_t n_ready_split;  // split cells passing the 2:1 check
    int32_t n_new_blocks;   // of those, without a reusable child block
    int32_t n_unsplit;
    int32_t n_deferred;
    int32_t n_rest;         // cells left to the general flux kernel
    int32_t n_rest2;        // cells of a dense cluster left to the per-cell sweep
    int32_t kept;
    int32_t lowered;
    int32_t dropped;
    int32_t raised[16];
    int32_t n_list;         // size of the substep's cluster list as selected on the device
    int32_t n_list2;        // size of the cluster list rebuilt after a remesh
    int64_t n_pool;         // pool size after this substep's splits (substeps planned on the device)
};

__device__ __forceinline__ bool is_leaf(const Dev &d, int32_t i) { return d.flags[i] & F_LEAF; }

// a[k] for a runtime k < N, as a chain of selects over values passed through
// an empty asm, so that the chain cannot be folded back into an indexed load
// (which would keep the array in private memory)
template <int N>
__device__ __forceinline__ int pick_i(const int *a, int k) {
    int w[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
        w[q] = a[q];
        asm volatile("" : "+v"(w[q]));
    }
    int v = w[0];
#pragma unroll
    for (int q = 1; q < N; ++q) v = k == q ? w[q] : v;
    return v;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Give XCD x a contiguous run of logical blocks, so a run of
// consecutive list entries (spatial neighbours in Morton order) shares one
// L2.  A bijection of [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// byte-wide atomic OR through the enclosing aligned 32-bit word
__device__ __forceinline__ void atomic_or_u8(GMEM uint8_t *addr, uint8_t v) {
    const uintptr_t a = (uintptr_t) addr;
    unsigned int *w = (unsigned int *) (a & ~uintptr_t(3));
    atomicOr(w, (unsigned int) v << (8 * (a & 3)));
}

// Domain::find / _find: integrator/domain.hpp:65-82,407-417
template <int D>
__device__ inline int32_t dom_find(const Dev &d, const MConst &c, double *p, int max_level) {
    for (int i = 0; i < D; ++i) {
        const double h = 0.5 * c.active_size[i];
        if (p[i] < -h || p[i] > h) {
            if (c.periodic[i]) {
                if (p[i] < -h) p[i] += c.active_size[i];
                else p[i] -= c.active_size[i];
            } else {
                return 0;
            }
        }
    }
    const double rs = d.size[1];
    double x[D];
    for (int i = 0; i < D; ++i) x[i] = 2.0 * p[i] / rs;
    int32_t idx = 1;
    for (int guard = 0; guard < 64; ++guard) {
        if ((d.flags[idx] & F_LEAF) || d.level[idx] == max_level) return idx;
        int q = 0;
        for (int i = 0; i < D; ++i)
            if (x[i] > 0.0) q |= 1 << i;
        for (int i = 0; i < D; ++i) x[i] = (x[i] - 0.5 * (((q >> i) & 1) ? 1.0 : -1.0)) * 2.0;
        idx = d.child_base[idx] + q;
    }
    return idx;
}

// Domain::interpolate_p / _t / _k: integrator/domain.hpp:84-148: the mean
// over a node's active children, each child contributing its leaf value or,
// recursively, its own children's mean.  The 2:1 balance bounds a subtree's
// depth only next to the face it shares with a leaf; on its far side the
// subtree can reach the finest level.  The walk keeps an explicit stack of
// up to 16 levels in private memory (a non-inlined, rarely taken path) and
// adds the child means in the reference's order.
template <int NV, typename Leaf>
__device__ __forceinline__ void interp_walk(const GMEM uint8_t *flags, const GMEM int32_t *child_base, GMEM int32_t *err,
                                            int32_t idx, double *out, int nch, Leaf &&leaf) {
    if (flags[idx] & F_LEAF) {
        leaf(idx, out);
        return;
    }
    constexpr int MAXD = 16;
    double acc[MAXD][NV];
    int cnt[MAXD], pos[MAXD];
    int32_t base[MAXD];
    int dp = 0;
    for (int e = 0; e < NV; ++e) acc[0][e] = 0.0;
    cnt[0] = 0;
    pos[0] = 0;
    base[0] = child_base[idx];
    while (true) {
        if (pos[dp] == nch) {
            double m[NV];
            for (int e = 0; e < NV; ++e) m[e] = acc[dp][e] / cnt[dp];
            if (dp == 0) {
                for (int e = 0; e < NV; ++e) out[e] = m[e];
                return;
            }
            --dp;
            for (int e = 0; e < NV; ++e) acc[dp][e] = acc[dp][e] + m[e];
            cnt[dp]++;
            pos[dp]++;
            continue;
        }
        const int32_t i = base[dp] + pos[dp];
        const uint8_t fl = flags[i];
        if (!(fl & F_ACTIVE)) {
            pos[dp]++;
            continue;
        }
        if (fl & F_LEAF) {
            double g[NV];
            leaf(i, g);
            for (int e = 0; e < NV; ++e) acc[dp][e] = acc[dp][e] + g[e];
            cnt[dp]++;
            pos[dp]++;
            continue;
        }
        if (dp + 1 == MAXD) {  // deeper than any tree of 16 levels
            atomicOr((int32_t *) err, 1);
            pos[dp]++;
            continue;
        }
        ++dp;
        for (int e = 0; e < NV; ++e) acc[dp][e] = 0.0;
        cnt[dp] = 0;
        pos[dp] = 0;
        base[dp] = child_base[i];
    }
}

// The non-inlined walks take plain pointers: a reference to the kernel's
// by-value Dev argument would make every thread copy it to scratch.
template <int NE, int NCH>
__device__ __noinline__ void interp_vec(const GMEM uint8_t *flags, const GMEM int32_t *child_base, GMEM int32_t *err,
                                        const GMEM double *field, int64_t C, int32_t idx, double *out) {
    interp_walk<NE>(flags, child_base, err, idx, out, NCH, [&](int32_t i, double *o) {
        for (int e = 0; e < NE; ++e) o[e] = field[e * C + i];
    });
}

template <int NCH>
__device__ __noinline__ double interp_scalar(const GMEM uint8_t *flags, const GMEM int32_t *child_base, GMEM int32_t *err,
                                             const GMEM double *field, int32_t idx) {
    double o[1];
    interp_walk<1>(flags, child_base, err, idx, o, NCH, [&](int32_t i, double *r) { r[0] = field[i]; });
    return o[0];
}

template <int MODEL, int D>
__device__ __forceinline__ void interp_p(const Dev &d, int32_t idx, double *out) {
    interp_vec<Sz<MODEL, D>::NE, (1 << D)>(d.flags, d.child_base, d.err, d.p, d.C, idx, out);
}

template <int D>
__device__ __forceinline__ double interp_t(const Dev &d, int32_t idx) {
    return interp_scalar<(1 << D)>(d.flags, d.child_base, d.err, d.t, idx);
}

template <int D>
__device__ __forceinline__ double interp_k(const Dev &d, int32_t idx, int kk) {
    return interp_scalar<(1 << D)>(d.flags, d.child_base, d.err, d.k + kk * d.C, idx);
}

// Active leaves of subtree `root` touching its face (dim, side), visited in
// depth-first child order, down to three levels below root (deeper
// subtrees set the device error word).
template <int D, typename Fn>
__device__ inline void for_sub_face_leaves(const Dev &d, int32_t root, int dim, int side, Fn &&fn) {
    auto visit_leaf = [&](int32_t idx) {
        const uint8_t fl = d.flags[idx];
        if ((fl & F_ACTIVE) && !(fl & F_GARBAGE)) fn(idx);
    };
    if (d.flags[root] & F_LEAF) {
        visit_leaf(root);
        return;
    }
    const int32_t c1 = d.child_base[root];
    for (int a = 0; a < (1 << D); ++a) {
        if (((a >> dim) & 1) != side) continue;
        const int32_t i1 = c1 + a;
        if (d.flags[i1] & F_LEAF) {
            visit_leaf(i1);
            continue;
        }
        const int32_t c2 = d.child_base[i1];
        for (int b = 0; b < (1 << D); ++b) {
            if (((b >> dim) & 1) != side) continue;
            const int32_t i2 = c2 + b;
            if (d.flags[i2] & F_LEAF) {
                visit_leaf(i2);
                continue;
            }
            const int32_t c3 = d.child_base[i2];
            for (int c = 0; c < (1 << D); ++c) {
                if (((c >> dim) & 1) != side) continue;
                const int32_t i3 = c3 + c;
                if (!(d.flags[i3] & F_LEAF)) atomicOr((int32_t *) d.err, 1);
                visit_leaf(i3);
            }
        }
    }
}

// count of lanes with pred, added once per wavefront
__device__ __forceinline__ void wave_count_add(bool pred, int32_t *counter) {
    const unsigned long long m = __ballot(pred);
    if (m && (threadIdx.x & 63) == (unsigned) (__ffsll((long long) m) - 1)) atomicAdd(counter, (int32_t) __popcll(m));
}

// Cell data one side of a face needs (integrator/cell.hpp:8-38).  The
// gradient (NS, multi) enters only the viscous and heat terms at the end of
// the face flux, so a side carries where it lives (the node's first
// gradient component and the component stride) and the flux reads it there:
// registers are not held for it through the Riemann solve.
template <int MODEL, int D>
struct Side {
    using S = Sz<MODEL, D>;
    double p[S::NE], s[S::NE], r[S::NE];
    double t;
    double kap, kn;               // multi: k[kappa], k[ikn + dim]
    const GMEM double *g;         // NS / multi: &grad[node]; component (a, e) at g[(a * NGX + e) * gC]
    int64_t gC;
    __device__ __forceinline__ double grad(int a, int e) const { return g[(a * S::NGX + e) * gC]; }
};

template <int MODEL, int D>
__device__ __forceinline__ void load_side(const Dev &d, int32_t i, int dim, Side<MODEL, D> &s) {
    using S = Sz<MODEL, D>;
    const int64_t C = d.C;
#pragma unroll
    for (int e = 0; e < S::NE; ++e) {
        s.p[e] = d.p[e * C + i];
        s.s[e] = d.slope[(dim * S::NE + e) * C + i];
        s.r[e] = d.rhs[e * C + i];
    }
    s.t = d.t[i];
    if constexpr (MODEL == MULTI) {
        s.kap = d.k[S::IKAPPA * C + i];
        s.kn = d.k[(S::IKN + dim) * C + i];
    }
    if constexpr (S::NG > 0) {
        s.g = d.grad + i;
        s.gC = C;
    }
}

// Model::flux of the face (cn | cp): euler.hpp:156-200,
// navier_stokes.hpp:61-82, multi.hpp:249-302.  fc: face centre along dim;
// cnc, cpc: cell centres along dim after any periodic shift.
// one side's MUSCL-Hancock state at the face: p + slope dx - rhs (tm - t),
// tm = (t0 + t1) / 2 (euler.hpp:168-181)
template <int MODEL, int D>
__device__ __forceinline__ void face_state(const double *p, const double *s, const double *r, double t, double dx,
                                           double tm, double *q) {
    constexpr int NE = Sz<MODEL, D>::NE;
    const double dta = tm - t;
#pragma unroll
    for (int e = 0; e < NE; ++e) q[e] = p[e] + s[e] * dx;
#pragma unroll
    for (int e = 0; e < NE; ++e) q[e] = q[e] - r[e] * dta;
}

// The face flux from the two face states pn | pp (euler.hpp:182-199,
// navier_stokes.hpp:61-82, multi.hpp:262-302).  The cell sides n and pz
// enter only through means (curvature, normal, gradients, velocity), so
// their order does not matter.
template <int MODEL, int D>
__device__ __forceinline__ void riemann_flux(const MConst &c, const double *pn, const double *pp,
                                             const Side<MODEL, D> &n, const Side<MODEL, D> &pz, double area, double t0,
                                             double t1, int dim, double *out) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE, NF = S::NF, IU = S::IU, IP = S::IP;
    double pm[NE], F[NF];
    if constexpr (MODEL == MULTI) {
        constexpr int IA = S::IA;
        const double ka = 0.5 * (n.kap + pz.kap);
        middle_state<MODEL, D>(c, pn, pp, ka, dim, pm);
        flux_F<MODEL, D, false>(c, pm, 0.0, dim, F);
        if (pm[IA] > 1e-5 && pm[IA] < 1 - 1e-5) {
            const double nn = 0.5 * (n.kn + pz.kn);
            const double uc = 1.0 * fabs(pm[IU + dim]) * nn;
            double ke = 0.0;
            for (int i = 0; i < D; ++i) {
                F[IU + i] += uc * ((1 - pm[IA]) * pm[0] - pm[IA] * pm[1]) * pm[IU + i];
                ke += 0.5 * pm[IU + i] * pm[IU + i];
            }
            F[0] += uc * (1 - pm[IA]) * pm[0];
            F[1] += uc * (-pm[IA]) * pm[1];
            F[IP] += uc * ((1 - pm[IA]) * pm[0] - pm[IA] * pm[1]) * ke +
                     uc * (1 - pm[IA]) * pm[IA] *
                         (pm[IP] * (1 / (c.gamma1 - 1) - 1 / (c.gamma2 - 1)) + c.gamma1 * c.P1 / (c.gamma1 - 1) -
                          c.gamma2 * c.P2 / (c.gamma2 - 1));
            F[IA] += uc * (1 - pm[IA]) * pm[IA];
        }
        double du[D][S::NGX], vel[D], fv[D + 1];
        for (int a = 0; a < D; ++a)
            for (int e = 0; e < S::NG; ++e) du[a][e] = 0.5 * (n.grad(a, e) + pz.grad(a, e));
        for (int i = 0; i < D; ++i) vel[i] = 0.5 * (n.p[IU + i] + pz.p[IU + i]);
        visc_Fv<D, S::NGX>(du, vel, model_mu<MODEL, D>(c, pm), dim, fv);
        for (int i = 0; i <= D; ++i) F[IU + i] -= fv[i];
        F[IU + D] -= du[dim][S::IDT] * model_lambda<MODEL, D>(c, pm);
#pragma unroll
        for (int e = 0; e < NF; ++e) out[e] = F[e] * area * (t1 - t0);
    } else {
        middle_state<MODEL, D>(c, pn, pp, 0.0, dim, pm);
        flux_F<MODEL, D, false>(c, pm, 0.0, dim, F);
#pragma unroll
        for (int e = 0; e < NF; ++e) out[e] = F[e] * area * (t1 - t0);
        if constexpr (MODEL == NAVIER_STOKES) {
            double du[D][S::NGX], vel[D], fv[D + 1];
            for (int a = 0; a < D; ++a)
                for (int e = 0; e < S::NG; ++e) du[a][e] = 0.5 * (n.grad(a, e) + pz.grad(a, e));
            for (int i = 0; i < D; ++i) vel[i] = 0.5 * (n.p[IU + i] + pz.p[IU + i]);
            visc_Fv<D, S::NGX>(du, vel, c.ns_mu, dim, fv);
            for (int i = 0; i <= D; ++i) out[IU + i] -= fv[i] * area * (t1 - t0);
            out[IU + D] -= c.ns_lambda * du[dim][S::IDT] * area * (t1 - t0);
        }
    }
}

template <int MODEL, int D>
__device__ __forceinline__ void model_flux(const MConst &c, const Side<MODEL, D> &n, double cnc,
                                           const Side<MODEL, D> &pz, double cpc, double fc, double area, double t0,
                                           double t1, int dim, double *out) {
    constexpr int NE = Sz<MODEL, D>::NE;
    const double tm = 0.5 * (t1 + t0);
    double pn[NE], pp[NE];
    face_state<MODEL, D>(n.p, n.s, n.r, n.t, fc - cnc, tm, pn);
    face_state<MODEL, D>(pz.p, pz.s, pz.r, pz.t, fc - cpc, tm, pp);
    riemann_flux<MODEL, D>(c, pn, pp, n, pz, area, t0, t1, dim, out);
}

// Flux of owner x's face (dim, dir) from the two loaded sides: impl/flux.hpp:32-91.
// sx is x's side, sn the neighbour's; nbc the neighbour centre along dim.
// Returns F with the reference's sign convention (x receives -F if dir==1,
// +F if dir==0; the neighbour the opposite).

// model_flux with the face axis as a constant for the five-equation model:
// each case is inlined with dim fixed, so the per-axis selections index the
// state arrays statically (with a runtime axis the compiler keeps them in
// private memory)
template <int MODEL, int D>
__device__ __forceinline__ void model_flux_axis(const MConst &c, const Side<MODEL, D> &n, double cnc,
                                                const Side<MODEL, D> &pz, double cpc, double fc, double area,
                                                double t0, double t1, int dim, double *out) {
    if constexpr (MODEL == MULTI) {
        if (dim == 0) model_flux<MODEL, D>(c, n, cnc, pz, cpc, fc, area, t0, t1, 0, out);
        else if (D == 2 || dim == 1) model_flux<MODEL, D>(c, n, cnc, pz, cpc, fc, area, t0, t1, 1, out);
        else model_flux<MODEL, D>(c, n, cnc, pz, cpc, fc, area, t0, t1, D - 1, out);
    } else {
        model_flux<MODEL, D>(c, n, cnc, pz, cpc, fc, area, t0, t1, dim, out);
    }
}

template <int MODEL, int D>
__device__ __forceinline__ void face_flux_sides_body(const MConst &c, const Side<MODEL, D> &sx, double xc, double xs, int xlev,
                                       const Side<MODEL, D> &sn, double nbc, int dim, int dir, double dt,
                                       double *F) {
    const double fc = xc + xs * (dir - 0.5);
    const double t0 = sx.t;
    const double t1 = t0 + dt;
    const double area = c.area[xlev];
    if (c.periodic[dim]) {
        if (dir == 1) {
            if (xc + xs > 0.5 * c.active_size[dim]) nbc += c.active_size[dim];
        } else {
            if (xc - xs < -0.5 * c.active_size[dim]) nbc -= c.active_size[dim];
        }
    }
    if (dir == 1) model_flux_axis<MODEL, D>(c, sx, xc, sn, nbc, fc, area, t0, t1, dim, F);
    else model_flux_axis<MODEL, D>(c, sn, nbc, sx, xc, fc, area, t0, t1, dim, F);
}

template <int MODEL, int D>
__device__ __noinline__ void face_flux_sides(const MConst &c, const Side<MODEL, D> &sx, double xc, double xs, int xlev,
                                             const Side<MODEL, D> &sn, double nbc, int dim, int dir, double dt,
                                             double *F) {
    face_flux_sides_body<MODEL, D>(c, sx, xc, xs, xlev, sn, nbc, dim, dir, dt, F);
}

// The general flux path inlines the face flux, except for the 3D
// five-equation model, whose inlined sides would exceed the register file
// (measured: one wave per SIMD and 13% slower).
template <int MODEL, int D>
__device__ __forceinline__ void face_flux_pick(const MConst &c, const Side<MODEL, D> &sx, double xc, double xs, int xlev,
                                               const Side<MODEL, D> &sn, double nbc, int dim, int dir, double dt,
                                               double *F) {
    if constexpr (MODEL == MULTI && D == 3) face_flux_sides<MODEL, D>(c, sx, xc, xs, xlev, sn, nbc, dim, dir, dt, F);
    else face_flux_sides_body<MODEL, D>(c, sx, xc, xs, xlev, sn, nbc, dim, dir, dt, F);
}

// Same, loading x's side from memory; sn is x's neighbour across the face.
template <int MODEL, int D>
__device__ __noinline__ void owner_face_flux_with(const Dev &d, const MConst &c, int32_t x, const Side<MODEL, D> &sn,
                                                  int32_t nb, int dim, int dir, double dt, double *F) {
    const int64_t C = d.C;
    Side<MODEL, D> sx;
    load_side<MODEL, D>(d, x, dim, sx);
    face_flux_sides<MODEL, D>(c, sx, d.center[dim * C + x], d.size[x], d.level[x], sn, d.center[dim * C + nb], dim,
                              dir, dt, F);
}

// ghost state of face f of cell y: the domain condition the face carries,
// or the cell's object record (integrator/boundary.hpp:75-123)
template <int MODEL, int D>
__device__ inline void boundary_value_at(const Dev &d, const MConst &c, int32_t y, int f, const double *pc, double t,
                                         double *pb) {
    const uint8_t b = d.bc[f * d.C + y];
    if (b == BC_OBJECT) boundary_value_rec<MODEL, D>(c, &d.brecs[d.brec[y]], pc, t, pb);
    else boundary_value<MODEL, D>(c, b ? b - 1 : f, pc, t, pb);
}

// boundary_flux: integrator/boundary.hpp:130-160 (already scaled), face f
// of cell y
template <int MODEL, int D>
__device__ inline void boundary_flux(const Dev &d, const MConst &c, int32_t y, int f, int dim, int dir, double dt,
                                     double *out) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE, NF = S::NF, IU = S::IU, IP = S::IP;
    const int64_t C = d.C;
    double pc[NE], pb[NE], F[NF];
    for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
    const double t0 = d.t[y], t1 = t0 + dt;
    boundary_value_at<MODEL, D>(d, c, y, f, pc, t0, pb);
    const uint8_t b = d.bc[f * C + y];
    const int type = b == BC_OBJECT ? d.brecs[d.brec[y]].type : c.bc_type[b ? b - 1 : f];
    if (type == BC_FREESLIP) {
        flux_F<MODEL, D, true>(c, pc, pb[IU + dim], dim, F);
    } else if (type == BC_NOSLIP) {
        for (int e = 0; e < NF; ++e) F[e] = 0.0;
        F[IU + dim] = pb[IP];
    } else {
        double pm[NE];
        middle_state<MODEL, D>(c, pb, pc, 0.0, dim, pm);
        flux_F<MODEL, D, false>(c, pm, 0.0, dim, F);
    }
    if constexpr (S::NG > 0) {
        double g[D][S::NGX], vel0[D], fv[D + 1];
        for (int a = 0; a < D; ++a)
            for (int e = 0; e < S::NGX; ++e) g[a][e] = d.grad[(a * S::NGX + e) * C + y];
        for (int i = 0; i < D; ++i) vel0[i] = 0.0;
        visc_Fv<D, S::NGX>(g, vel0, model_mu<MODEL, D>(c, pc), dim, fv);
        for (int i = 0; i <= D; ++i) F[IU + i] -= fv[i];
    }
    const double direction = 2 * dir - 1;
    const double area = c.area[d.level[y]];
    for (int e = 0; e < NF; ++e) out[e] = F[e] * area * (t1 - t0) * direction;
}

// ------------------------------------------------------------------ flux
// One level-L substep's flux totals: impl/flux.hpp:10-97 (restated as a
// gather, see the file comment).  Writes tot[NF] of every listed cell and
// deposits fluxes owed to cells outside the cluster.  The listed cell's own
// data is loaded once; one neighbour side is loaded per face.
template <int MODEL, int D>
__device__ void flux_cell(const Dev &d, const MConst &c, int32_t y, int L, double dt) {
    using S = Sz<MODEL, D>;
    constexpr int NF = S::NF, NE = S::NE;
    const int64_t C = d.C;
    const int ylev = d.level[y];
    const double ys = d.size[y];
    Side<MODEL, D> sy;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        sy.p[e] = d.p[e * C + y];
        sy.r[e] = d.rhs[e * C + y];
    }
    sy.t = d.t[y];
    if constexpr (MODEL == MULTI) sy.kap = d.k[S::IKAPPA * C + y];
    if constexpr (S::NG > 0) {
        sy.g = d.grad + y;
        sy.gC = C;
    }
    double tot[NF];
#pragma unroll
    for (int e = 0; e < NF; ++e) tot[e] = 0.0;
#pragma unroll 1
    for (int dim = 0; dim < D; ++dim) {
#pragma unroll
        for (int e = 0; e < NE; ++e) sy.s[e] = d.slope[(dim * NE + e) * C + y];
        if constexpr (MODEL == MULTI) sy.kn = d.k[(S::IKN + dim) * C + y];
        const double yc = d.center[dim * C + y];
#pragma unroll 1
        for (int dir = 0; dir < 2; ++dir) {
            const int f = 2 * dim + dir;
            const int32_t nb = d.nbr[f * C + y];
            double cc[NF], F[NF];
#pragma unroll
            for (int e = 0; e < NF; ++e) cc[e] = 0.0;
            if (nb == 0) {
                boundary_flux<MODEL, D>(d, c, y, f, dim, dir, dt, F);
#pragma unroll
                for (int e = 0; e < NF; ++e) cc[e] = cc[e] - F[e];
            } else if (d.flags[nb] & F_LEAF) {
                const int nlev = d.level[nb];
                const int nacti = d.acti[nb];
                const int ncl = d.cluster[nb];
                if (nacti > L) {
                    // deposits of a neighbour on a finer time level (or of the
                    // cells it was split from or coarsened from)
#pragma unroll
                    for (int e = 0; e < NF; ++e) {
                        cc[e] = d.acc[(f * NF + e) * C + y];
                        d.acc[(f * NF + e) * C + y] = 0.0;
                    }
                }
                bool y_computes, nb_computes = false;
                if (nlev < ylev) {
                    y_computes = nacti <= L;
                } else {
                    y_computes = nacti <= L && (dir == 1 || nacti < L);
                    nb_computes = ncl == L && dir == 0 && d.nbr[(2 * dim + 1) * C + nb] == y;
                }
                if (y_computes || nb_computes) {
                    Side<MODEL, D> sn;
                    load_side<MODEL, D>(d, nb, dim, sn);
                    const double nbc = d.center[dim * C + nb];
                    if (y_computes) {
                        face_flux_pick<MODEL, D>(c, sy, yc, ys, ylev, sn, nbc, dim, dir, dt, F);
#pragma unroll
                        for (int e = 0; e < NF; ++e) cc[e] = dir == 1 ? cc[e] - F[e] : cc[e] + F[e];
                    } else {
                        face_flux_pick<MODEL, D>(c, sn, nbc, d.size[nb], nlev, sy, yc, dim, 1, dt, F);
#pragma unroll
                        for (int e = 0; e < NF; ++e) cc[e] = cc[e] + F[e];
                    }
                    if (y_computes && ncl != L) {
                        const int fo = 2 * dim + (1 - dir);
                        if (nlev == ylev) {
#pragma unroll
                            for (int e = 0; e < NF; ++e) {
                                GMEM double &a = d.acc[(fo * NF + e) * C + nb];
                                a = dir == 1 ? a + F[e] : a - F[e];
                            }
                        } else {
                            // coarser receiver: kept here and deposited by
                            // deposit_coarse (the first computing sibling adds
                            // all siblings' fluxes in sub-face order)
#pragma unroll
                            for (int e = 0; e < NF; ++e) d.dep[(dim * NF + e) * C + y] = F[e];
                        }
                    }
                }
            } else {
                // finer neighbours: received deposits plus this substep's sub-faces
                bool any_fast = false;
                for_sub_face_leaves<D>(d, nb, dim, 1 - dir, [&](int32_t x) {
                    if (d.acti[x] > L) any_fast = true;
                });
                if (any_fast) {
#pragma unroll
                    for (int e = 0; e < NF; ++e) {
                        cc[e] = d.acc[(f * NF + e) * C + y];
                        d.acc[(f * NF + e) * C + y] = 0.0;
                    }
                }
                for_sub_face_leaves<D>(d, nb, dim, 1 - dir, [&](int32_t x) {
                    if (d.cluster[x] == L && d.nbr[(2 * dim + (1 - dir)) * C + x] == y) {
                        double G[NF];
                        owner_face_flux_with<MODEL, D>(d, c, x, sy, y, dim, 1 - dir, dt, G);
                        for (int e = 0; e < NF; ++e) cc[e] = (1 - dir) == 1 ? cc[e] + G[e] : cc[e] - G[e];
                    }
                });
            }
#pragma unroll
            for (int e = 0; e < NF; ++e) tot[e] = tot[e] + cc[e];
        }
    }
#pragma unroll
    for (int e = 0; e < NF; ++e) d.tot[e * C + y] = tot[e];
}

// The deposits of a set-aside cell toward coarser receivers outside the
// cluster (impl/flux.hpp:82-91 with the oracle's order): for each such face
// the first computing sibling of the sub-face group adds the siblings'
// fluxes (stored by flux_cell) to the receiver's face accumulator in
// sub-face order.  Runs after every flux of the substep is computed.
// One face f of y: the faces are independent (each writes only the
// receiver's slot on that face, and one sibling of a group writes it).
template <int MODEL, int D>
__device__ void deposit_coarse_face(const Dev &d, int32_t y, int f, int L) {
    using S = Sz<MODEL, D>;
    constexpr int NF = S::NF;
    const int64_t C = d.C;
    const int ylev = d.level[y];
    const int dim = f >> 1, dir = f & 1;
    const int32_t nb = d.nbr[f * C + y];
    if (nb == 0 || !(d.flags[nb] & F_LEAF)) return;
    if (d.level[nb] >= ylev || d.acti[nb] > L || d.cluster[nb] == L) return;
    const int fo = 2 * dim + (1 - dir);
    const int32_t rn = d.nbr[fo * C + nb];
    int32_t first = 0;
    for_sub_face_leaves<D>(d, rn, dim, dir, [&](int32_t x) {
        if (first == 0 && d.cluster[x] == L && d.nbr[f * C + x] == nb) first = x;
    });
    if (first != y) return;
    for_sub_face_leaves<D>(d, rn, dim, dir, [&](int32_t x) {
        if (d.cluster[x] == L && d.nbr[f * C + x] == nb) {
            for (int e = 0; e < NF; ++e) {
                GMEM double &a = d.acc[(fo * NF + e) * C + nb];
                const double g = d.dep[(dim * NF + e) * C + x];
                a = dir == 1 ? a + g : a - g;
            }
        }
    });
}

template <int MODEL, int D>
__device__ void deposit_coarse(const Dev &d, int32_t y, int L) {
#pragma unroll 1
    for (int f = 0; f < 2 * D; ++f) deposit_coarse_face<MODEL, D>(d, y, f, L);
}

// ---------------------------------------------------------------- update
// cell_update: impl/cell_update.hpp:9-90, multi::source multi.hpp:186-203
template <int MODEL, int D>
__device__ __forceinline__ void update_body(const Dev &d, const MConst &c, const int32_t *__restrict__ list, int nl, int li, double dt,
                                                 int signal, Status *st) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE, NF = S::NF, IU = S::IU, IP = S::IP;
    if (li >= nl) return;
    const int64_t C = d.C;
    const int32_t y = list[li];
    double u[NE], p[NE], du[NF];
    const double vol = c.vol[d.level[y]];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        u[e] = d.u[e * C + y];
        p[e] = d.p[e * C + y];
    }
    double gold[S::NGX], gnew[S::NGX];
    if constexpr (S::NG > 0) grad_input<MODEL, D>(c, p, gold);
    double rv;
    if (pow2_recip(vol, rv)) {
#pragma unroll
        for (int e = 0; e < NF; ++e) du[e] = d.tot[e * C + y] * rv;
    } else {
#pragma unroll
        for (int e = 0; e < NF; ++e) du[e] = d.tot[e * C + y] / vol;
    }
#pragma unroll
    for (int e = 0; e < NE; ++e) u[e] += du[e];
    if constexpr (MODEL == MULTI) {
        constexpr int IA = S::IA;
        if (!signal) {
            if (u[IA] > 0.001)
                for (int e = IU; e < IU + D; ++e) u[e] = 0.0;
        }
        const double kap = d.k[S::IKAPPA * C + y];
        double s[NE];
        for (int e = 0; e < NE; ++e) s[e] = 0.0;
        for (int i = 0; i < D; ++i) s[IU + i] = -c.sigma * kap * du[IA + 2 + i];
        s[IP] = c.sigma * kap * (-du[IA] + p[IA] * du[IA + 1]);
        s[IA] = -p[IA] * du[IA + 1];
        for (int e = 0; e < NE; ++e) u[e] = u[e] + s[e];
    }
    prim_from_cons<MODEL, D>(c, u, p);
    if (isnan(p[IP]) || isnan(p[0])) atomicOr(&st->nan_flag, 1);
    p[0] = smax(p[0], 1e-3);
    if constexpr (MODEL == MULTI) p[1] = smax(p[1], 1e-6);
    p[IP] = smax(p[IP], 1e-3);
    cons_from_prim<MODEL, D>(c, p, u);
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        d.u[e * C + y] = u[e];
        d.p[e * C + y] = p[e];
    }
    if constexpr (S::NG > 0) {
        grad_input<MODEL, D>(c, p, gnew);
        for (int e = 0; e < S::NG; ++e) d.t_slope[e * C + y] = (gnew[e] - gold[e]) / dt;
    }
    if constexpr (MODEL == MULTI) {
        d.k[0 * C + y] = c.mu1 * p[S::IA] + c.mu2 * (1 - p[S::IA]);
        d.k[S::IALPHAS * C + y] = p[S::IA];
    }
    d.t[y] = d.t[y] + dt;
}

// ----------------------------------------------------------------- slope
// compute_slope: impl/slopes.hpp:15-76
// compute_slope's axis dim (slopes.hpp:15-76): the limited slope (stored and
// returned in sl) and the gradient of that axis; the axes are independent
template <int MODEL, int D>
__device__ inline void compute_slope_axis(const Dev &d, const MConst &c, int32_t y, int dim, double *sl) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE;
    const int64_t C = d.C;
    double pc[NE];
    for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
    const double ys = d.size[y], yt = d.t[y];
    double ts[S::NGX];
    if constexpr (S::NG > 0)
        for (int e = 0; e < S::NG; ++e) ts[e] = d.t_slope[e * C + y];
    {
        const int32_t in = d.nbr[(2 * dim) * C + y], ip = d.nbr[(2 * dim + 1) * C + y];
        double pn[NE], pp[NE], xn, xp, tn, tp;
        const double xc = 0.0;
        if (in == 0) {
            boundary_value_at<MODEL, D>(d, c, y, 2 * dim, pc, yt, pn);
            xn = -ys;
            tn = yt;
        } else {
            if (!(d.flags[in] & F_LEAF)) {
                // the walk writes through a pointer: a temporary keeps pn in
                // registers, and the own state is read again after the calls
                double tmp[NE];
                interp_p<MODEL, D>(d, in, tmp);
                for (int e = 0; e < NE; ++e) pn[e] = tmp[e];
                tn = interp_t<D>(d, in);
                for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
            } else {
                for (int e = 0; e < NE; ++e) pn[e] = d.p[e * C + in];
                tn = d.t[in];
            }
            xn = -0.5 * (ys + d.size[in]);
        }
        if (ip == 0) {
            boundary_value_at<MODEL, D>(d, c, y, 2 * dim + 1, pc, yt, pp);
            xp = ys;
            tp = yt;
        } else {
            if (!(d.flags[ip] & F_LEAF)) {
                double tmp[NE];
                interp_p<MODEL, D>(d, ip, tmp);
                for (int e = 0; e < NE; ++e) pp[e] = tmp[e];
                tp = interp_t<D>(d, ip);
                for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
            } else {
                for (int e = 0; e < NE; ++e) pp[e] = d.p[e * C + ip];
                tp = d.t[ip];
            }
            xp = 0.5 * (ys + d.size[ip]);
        }
        for (int e = 0; e < NE; ++e) {
            const double sn = (pc[e] - pn[e]) / (xc - xn);
            const double sp = (pp[e] - pc[e]) / (xp - xc);
            sl[e] = limiter(c.limiter, sn, sp);
            d.slope[(dim * NE + e) * C + y] = sl[e];
        }
        if constexpr (S::NG > 0) {
            double gp[S::NG], gn[S::NG];
            grad_input<MODEL, D>(c, pp, gp);
            grad_input<MODEL, D>(c, pn, gn);
            for (int e = 0; e < S::NG; ++e)
                d.grad[(dim * S::NGX + e) * C + y] = (gp[e] - gn[e] - ts[e] * (tp - tn)) / (xp - xn);
        }
    }
}

template <int MODEL, int D>
__device__ inline void compute_slope_dev(const Dev &d, const MConst &c, int32_t y) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE;
    const int64_t C = d.C;
    double pc[NE], slope[D][NE];
#pragma unroll
    for (int dim = 0; dim < D; ++dim) compute_slope_axis<MODEL, D>(d, c, y, dim, slope[dim]);
    for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
    double rhs[NE];
    const double kap = MODEL == MULTI ? d.k[S::IKAPPA * C + y] : 0.0;
    compute_rhs<MODEL, D>(c, pc, slope, kap, rhs);
    for (int e = 0; e < NE; ++e) d.rhs[e * C + y] = rhs[e];
}

template <int MODEL, int D>
__global__ void __launch_bounds__(256) k_slope(Dev d, MConst c, const int32_t *__restrict__ list, int nl) {
    const int li = blockIdx.x * blockDim.x + threadIdx.x;
    if (li >= nl) return;
    compute_slope_dev<MODEL, D>(d, c, list[li]);
}

// Scene::can_split: scene/scene.hpp:50-61
template <int D>
__device__ inline bool scene_can_split(const MConst &c, const double *x, int level) {
    if (level < c.max_refinement_level) return true;
    for (int z = 0; z < c.n_zones; ++z) {
        const double x3[3] = {x[0], D > 1 ? x[1 % D] : 0.0, D > 2 ? x[2 % D] : 0.0};
        if (level < c.zone_level[z] && x3[0] < c.zone_x1[z][0] && x3[1] < c.zone_x1[z][1] && x3[2] < c.zone_x1[z][2] &&
            x3[0] > c.zone_x0[z][0] && x3[1] > c.zone_x0[z][1] && x3[2] > c.zone_x0[z][2])
            return true;
    }
    return false;
}

// ------------------------------------------------------------ sweep
// After the update of level L: slopes (impl/slopes.hpp:15-76), the error
// estimate and split / unsplit marks (impl/adaptive_refinement.hpp:18-94)
// and, for unmarked cells, the LTS decision (:264-322) in one neighbour
// sweep.  The decision is stored (nextc, buffer) and applied by k_apply
// after any remesh, so the remesh sees the levels the reference sees.  A
// cell whose error needs the new slope of a coarser neighbour updated in
// the same substep is deferred to a second pass (pass = 1).
template <int MODEL, int D>
__device__ inline void lts_decide(const Dev &d, const MConst &c, int32_t y, const double *p, double size, int L,
                                  double dt, int counter, bool &lowered) {
    const double mdt = max_delta_time<MODEL, D>(c, p, size);
    double cfl = dt / mdt;
    int8_t nc = (int8_t) L;
    if (cfl < 0.5 * c.max_cfl && counter == 1 && L > 0) {
        d.buffer[y] = (int8_t) (L - 1);
        nc = -1;
        cfl *= 2;
        lowered = true;
    } else if (cfl > c.max_cfl && L < 15) {
        const int raise = (int) (log2(cfl / c.max_cfl) + 1);
        if (L + raise < 16) {
            nc = (int8_t) (L + raise);
            cfl /= 1 << raise;
        } else {
            nc = -1;
        }
    }
    d.max_dt[y] = mdt;
    d.cfl[y] = cfl;
    d.nextc[y] = nc;
}

// One neighbour of a cell for the sweep (slopes.hpp:21-55 and
// adaptive_refinement.hpp:38-64): pv is the value the slope uses, qv the
// value the error estimator uses, h false for a boundary face.  DIR is a
// template parameter so the per-side arrays stay in registers.
template <int MODEL, int D, int DIR>
__device__ __forceinline__ void sweep_side(const Dev &d, const MConst &c, int32_t y, int dim, double *pc,
                                           double *ctr, double ys, double yt, int ylev, int L, int pass,
                                           double *pv, double *qv, double &tv, double &xv, bool &hv,
                                           bool &deferred) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE;
    const int64_t C = d.C;
    const int32_t nb = d.nbr[(2 * dim + DIR) * C + y];
    hv = true;
    if (nb == 0) {
        boundary_value_at<MODEL, D>(d, c, y, 2 * dim + DIR, pc, yt, pv);
        xv = DIR ? ys : -ys;
        tv = yt;
        hv = false;
        return;
    }
    if (!(d.flags[nb] & F_LEAF)) {
        // the walk writes through a pointer: give it a temporary, so that pv
        // itself can stay in registers on the common path
        double tmp[NE];
        interp_p<MODEL, D>(d, nb, tmp);
        for (int e = 0; e < NE; ++e) pv[e] = tmp[e];
        tv = interp_t<D>(d, nb);
        for (int e = 0; e < NE; ++e) qv[e] = pv[e];
        // the cell's own state and centre are read again after the calls
        // (same values), so they are not live across them: otherwise their
        // registers are saved to private memory for every cell, walk or not
        for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
        for (int a = 0; a < D; ++a) ctr[a] = d.center[a * C + y];
    } else {
        for (int e = 0; e < NE; ++e) pv[e] = d.p[e * C + nb];
        tv = d.t[nb];
        if (d.level[nb] < ylev) {
            if (pass == 0 && d.cluster[nb] == L) {
                deferred = true;
            } else {
                // the coarser leaf's linear reconstruction at the neighbour
                // position (adaptive_refinement.hpp:48-59)
                // (dim is a runtime value: select per axis, no indexed
                // private arrays)
                double xd = pick<D>(ctr, dim) + ys * (2 * DIR - 1.0);
                if (c.periodic[dim] && DIR == 0 && xd < -0.5 * c.active_size[dim]) xd += c.active_size[dim];
                if (c.periodic[dim] && DIR == 1 && xd > 0.5 * c.active_size[dim]) xd -= c.active_size[dim];
                for (int e = 0; e < NE; ++e) qv[e] = pv[e];
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    const double dx = (a == dim ? xd : ctr[a]) - d.center[a * C + nb];
                    for (int e = 0; e < NE; ++e) qv[e] = qv[e] + d.slope[(a * NE + e) * C + nb] * dx;
                }
            }
        } else {
            for (int e = 0; e < NE; ++e) qv[e] = pv[e];
        }
    }
    const double ns = d.level[nb] == ylev ? ys : d.size[nb];
    xv = DIR ? 0.5 * (ys + ns) : -0.5 * (ys + ns);
}

// the end of a cell's sweep: error, split / unsplit marks
// (impl/adaptive_refinement.hpp:18-94) or, for an unmarked cell, the LTS
// decision (:264-322); a deferred cell only keeps its mark
template <int MODEL, int D>
__device__ __forceinline__ void sweep_decide(const Dev &d, const MConst &c, int32_t y, const double *pc, const double *ctr,
                                             double ys, int ylev, int L, double dt, int counter, bool deferred, double err,
                                             Status *st, int remesh, uint8_t fl) {
    bool lowered = false, marked = false;
    if (deferred) {
        d.mark[y] = MK_DEFER;
    } else {
        err = sqrt(err);
        d.error[y] = err;
        if (fl & F_CAN_SPLIT) {
            if (!scene_can_split<D>(c, ctr, ylev)) fl &= ~F_CAN_SPLIT;
            d.flags[y] = fl;
        }
        uint8_t m = 0;
        // with remeshing paused (a cadence set by the host, not a reference
        // behaviour) every cell takes the LTS decision
        if (remesh) {
            if ((err > c.max_error && (fl & F_CAN_SPLIT)) || (fl & F_SHOULD_SPLIT)) m |= MK_SPLIT;
            if (err < c.min_error && (fl & F_CAN_UNSPLIT) && ylev > 0 && counter == 1) m |= MK_UNSPLIT_CAND;
        }
        if (m == 0) {
            lts_decide<MODEL, D>(d, c, y, pc, ys, L, dt, counter, lowered);
            m = MK_RA_DONE;
        } else {
            marked = true;
        }
        d.mark[y] = m;
    }
    wave_count_add(deferred, &st->n_deferred);
    wave_count_add(marked, &st->n_marked);
    wave_count_add(lowered, &st->lowered);
}

// One axis of a cell's sweep from its two neighbour values (pn | pp: the
// values the slope uses, qn | qp: those the error estimator uses, hn / hp:
// false at a domain face): the limited slope (impl/slopes.hpp:15-76), the
// rhs terms of this axis, the gradient and the Laplacian error term
// (error_estimators.hpp:19-37).
// sweep_axis_core: the limited slope sl (stored with the gradient in pass
// 0) and the axis's Laplacian error term sq; sweep_axis adds the axis's rhs
// terms (pass 0) and sq to the cell's sums
template <int MODEL, int D>
__device__ __forceinline__ void sweep_axis_core(const Dev &d, const MConst &c, int32_t y, int dim, int pass,
                                                const double *pc, const double *pn, const double *pp, const double *qn,
                                                const double *qp, double xn, double xp, double tn, double tp, bool hn,
                                                bool hp, const double *ts, double *sl, double &sq_out) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE;
    const int64_t C = d.C;
    const double xc = 0.0;
    double sn[NE], sp[NE];
    double rn, rp;
    const double dn = xc - xn, dp = xp - xc;
    if (pow2_recip(dn, rn) && pow2_recip(dp, rp)) {
        for (int e = 0; e < NE; ++e) {
            sn[e] = (pc[e] - pn[e]) * rn;
            sp[e] = (pp[e] - pc[e]) * rp;
        }
    } else {
        for (int e = 0; e < NE; ++e) {
            sn[e] = (pc[e] - pn[e]) / dn;
            sp[e] = (pp[e] - pc[e]) / dp;
        }
    }
    for (int e = 0; e < NE; ++e) sl[e] = limiter(c.limiter, sn[e], sp[e]);
    if (pass == 0) {
        for (int e = 0; e < NE; ++e) __builtin_nontemporal_store(sl[e], &d.slope[(dim * NE + e) * C + y]);
        if constexpr (S::NG > 0) {
            double gp[S::NG], gn[S::NG];
            grad_input<MODEL, D>(c, pp, gp);
            grad_input<MODEL, D>(c, pn, gn);
            double rg;
            const double qg = xp - xn;
            if (pow2_recip(qg, rg)) {
                for (int e = 0; e < S::NG; ++e)
                    d.grad[(dim * S::NGX + e) * C + y] = (gp[e] - gn[e] - ts[e] * (tp - tn)) * rg;
            } else {
                for (int e = 0; e < S::NG; ++e)
                    d.grad[(dim * S::NGX + e) * C + y] = (gp[e] - gn[e] - ts[e] * (tp - tn)) / qg;
            }
        }
    }
    // weighted_laplacian::estimate, error_estimators.hpp:19-37
    double sq = 0.0;
    for (int e = 0; e < NE; ++e) {
        double d2 = pc[e] * -2.0;
        d2 = hn ? d2 + qn[e] : 0.0;
        d2 = (hn && hp) ? d2 + qp[e] : 0.0;
        const double v = d2 * c.weights[e];
        sq += v * v;
    }
    sq_out = sq;
}

template <int MODEL, int D>
__device__ __forceinline__ void sweep_axis(const Dev &d, const MConst &c, int32_t y, int dim, int pass, const double *pc,
                                           const double *pn, const double *pp, const double *qn, const double *qp,
                                           double xn, double xp, double tn, double tp, bool hn, bool hp,
                                           const double *ts, double kap, double a_ip_in, double a_in_ip, double a_in_ia,
                                           double *rhs, double &err) {
    double sl[Sz<MODEL, D>::NE], sq;
    sweep_axis_core<MODEL, D>(d, c, y, dim, pass, pc, pn, pp, qn, qp, xn, xp, tn, tp, hn, hp, ts, sl, sq);
    // rhs[i] gains the dim-j terms in the reference's (i, j, k) order
    if (pass == 0) rhs_add_dim<MODEL, D>(c, pc, sl, dim, kap, a_ip_in, a_in_ip, a_in_ia, rhs);
    err += sq;
}

template <int MODEL, int D>
__device__ __forceinline__ void sweep_body(const Dev &d, const MConst &c, const int32_t *__restrict__ list, int nl, int li, int L,
                                                double dt, int counter, int pass, Status *st, int remesh) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE;
    if (li >= nl) return;
    const int64_t C = d.C;
    const int32_t y = list[li];
    if (pass == 1 && !(d.mark[y] & MK_DEFER)) return;
    if (st->nan_flag) return;  // integrate() stops after the update (impl/integrate.hpp:50)
    double pc[NE], ctr[D], rhs[NE];
    for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
    for (int a = 0; a < D; ++a) ctr[a] = d.center[a * C + y];
    const double ys = d.size[y], yt = d.t[y];
    const int ylev = d.level[y];
    const double kap = MODEL == MULTI ? d.k[S::IKAPPA * C + y] : 0.0;
    double a_ip_in, a_in_ip, a_in_ia;
    rhs_coeffs<MODEL, D>(c, pc, kap, a_ip_in, a_in_ip, a_in_ia);
    for (int e = 0; e < NE; ++e) rhs[e] = 0.0;
    double ts[S::NGX];
    if constexpr (S::NG > 0)
        for (int e = 0; e < S::NG; ++e) ts[e] = d.t_slope[e * C + y];
    bool deferred = false;
    double err = 0.0;
#pragma unroll 1
    for (int dim = 0; dim < D; ++dim) {
        double pn[NE], pp[NE], qn[NE], qp[NE], xn, xp, tn, tp;
        bool hn, hp;
        const double xc = 0.0;
        sweep_side<MODEL, D, 0>(d, c, y, dim, pc, ctr, ys, yt, ylev, L, pass, pn, qn, tn, xn, hn, deferred);
        sweep_side<MODEL, D, 1>(d, c, y, dim, pc, ctr, ys, yt, ylev, L, pass, pp, qp, tp, xp, hp, deferred);
        sweep_axis<MODEL, D>(d, c, y, dim, pass, pc, pn, pp, qn, qp, xn, xp, tn, tp, hn, hp, ts, kap, a_ip_in, a_in_ip,
                             a_in_ia, rhs, err);
    }
    if (pass == 0)
        for (int e = 0; e < NE; ++e) __builtin_nontemporal_store(rhs[e], &d.rhs[e * C + y]);
    sweep_decide<MODEL, D>(d, c, y, pc, ctr, ys, ylev, L, dt, counter, deferred, err, st, remesh, d.flags[y]);
}

// apply the stored LTS decisions (impl/adaptive_refinement.hpp:308-321) and
// clear the marks of the listed cells
__device__ __forceinline__ void apply_body(const Dev &d, const int32_t *__restrict__ list, int nl, int li, int L) {
    if (li >= nl) return;
    const int32_t y = list[li];
    if (d.mark[y] & MK_RA_DONE) {
        const int8_t nc = d.nextc[y];
        d.cluster[y] = nc;
        if (nc > L) d.acti[y] = nc;
    }
    d.mark[y] = 0;
}

// ------------------------------------------------- multiphase interface
template <int MODEL, int D>
__device__ inline double bc_alpha(const Dev &d, const MConst &c, int32_t y, int f) {
    using S = Sz<MODEL, D>;
    double pc[S::NE], pb[S::NE];
    for (int e = 0; e < S::NE; ++e) pc[e] = d.p[e * d.C + y];
    boundary_value_at<MODEL, D>(d, c, y, f, pc, d.t[y], pb);
    return pb[S::IA];
}

// The interface chain's per-cell sums are one term per axis, added in axis
// order; each op gives the term of one axis (axis) and the sum (finish), so
// the axes can run on separate wavefronts (k_axes).
// smooth_alpha, one Jacobi sweep (impl/smooth.hpp:9-60): the axis's term
// 0.5 (a- + a+) + a
template <int MODEL, int D>
__device__ __forceinline__ double smooth_alpha_axis(const Dev &d, const MConst &c, int32_t y, int dim) {
    using S = Sz<MODEL, D>;
    const int64_t C = d.C;
    const double ka = d.k[S::IALPHAS * C + y];
    const int32_t in = d.nbr[(2 * dim) * C + y], ip = d.nbr[(2 * dim + 1) * C + y];
    double an, ap;
    if (in == 0) an = bc_alpha<MODEL, D>(d, c, y, 2 * dim);
    else if (!(d.flags[in] & F_LEAF)) an = interp_k<D>(d, in, S::IALPHAS);
    else an = d.k[S::IALPHAS * C + in];
    if (ip == 0) ap = bc_alpha<MODEL, D>(d, c, y, 2 * dim + 1);
    else if (!(d.flags[ip] & F_LEAF)) ap = interp_k<D>(d, ip, S::IALPHAS);
    else ap = d.k[S::IALPHAS * C + ip];
    return 0.5 * (an + ap) + ka;
}

__device__ __forceinline__ void store_k_body(const Dev &d, const int32_t *__restrict__ list, int nl, int li, int kk) {
    if (li >= nl) return;
    d.k[kk * d.C + list[li]] = d.scratch[li];
}

// compute_normal: impl/normal_curvature.hpp:6-55
// the axis's component of the alpha gradient, (a+ - a-) / (x+ - x-)
template <int MODEL, int D>
__device__ __forceinline__ double normal_axis(const Dev &d, const MConst &c, int32_t y, int dim) {
    using S = Sz<MODEL, D>;
    const int64_t C = d.C;
    const double ys = d.size[y];
    const int32_t in = d.nbr[(2 * dim) * C + y], ip = d.nbr[(2 * dim + 1) * C + y];
    double xn, xp, an, ap;
    if (in == 0) {
        an = bc_alpha<MODEL, D>(d, c, y, 2 * dim);
        xn = -0.5 * ys;
    } else {
        an = !(d.flags[in] & F_LEAF) ? interp_k<D>(d, in, S::IALPHAS) : d.k[S::IALPHAS * C + in];
        xn = -0.5 * (ys + d.size[in]);
    }
    if (ip == 0) {
        ap = bc_alpha<MODEL, D>(d, c, y, 2 * dim + 1);
        xp = 0.5 * ys;
    } else {
        ap = !(d.flags[ip] & F_LEAF) ? interp_k<D>(d, ip, S::IALPHAS) : d.k[S::IALPHAS * C + ip];
        xp = 0.5 * (ys + d.size[ip]);
    }
    return (ap - an) / (xp - xn);
}

// compute_curvature: impl/normal_curvature.hpp:58-102
// the axis's term of the curvature, (n+ - n-) / (x+ - x-) of the normal's
// axis component (subtracted)
template <int MODEL, int D>
__device__ __forceinline__ double curvature_axis(const Dev &d, const MConst &c, int32_t y, int dim) {
    using S = Sz<MODEL, D>;
    const int64_t C = d.C;
    const double ys = d.size[y];
    const int32_t in = d.nbr[(2 * dim) * C + y], ip = d.nbr[(2 * dim + 1) * C + y];
    double xn, xp, nn, np;
    if (in == 0) {
        nn = 0.0;
        xn = -0.5 * ys;
    } else {
        nn = !(d.flags[in] & F_LEAF) ? interp_k<D>(d, in, S::IKN + dim) : d.k[(S::IKN + dim) * C + in];
        xn = -0.5 * (ys + d.size[in]);
    }
    if (ip == 0) {
        np = 0.0;
        xp = 0.5 * ys;
    } else {
        np = !(d.flags[ip] & F_LEAF) ? interp_k<D>(d, ip, S::IKN + dim) : d.k[(S::IKN + dim) * C + ip];
        xp = 0.5 * (ys + d.size[ip]);
    }
    return (np - nn) / (xp - xn);
}

// smooth_curvature (one iteration): impl/smooth.hpp:63-142
// the axis's terms of the weighted sums: v[0] of curvature x weight, v[1]
// of the weights
template <int MODEL, int D>
__device__ __forceinline__ void smooth_curvature_axis(const Dev &d, const MConst &c, int32_t y, int dim, double *v) {
    using S = Sz<MODEL, D>;
    const int64_t C = d.C;
    const double kc = d.k[S::IKAPPA * C + y];
    const double al = d.k[S::IALPHAS * C + y];
    const double wa = al * (1 - al);
    const double wc = wa * wa;
    const int32_t in = d.nbr[(2 * dim) * C + y], ip = d.nbr[(2 * dim + 1) * C + y];
    double kn, kp, an, ap;
    if (in == 0) {
        kn = 0.0;
        an = bc_alpha<MODEL, D>(d, c, y, 2 * dim);
    } else if (!(d.flags[in] & F_LEAF)) {
        double pi[S::NE];
        kn = interp_k<D>(d, in, S::IKAPPA);
        interp_p<MODEL, D>(d, in, pi);
        an = pi[S::IA];
    } else {
        kn = d.k[S::IKAPPA * C + in];
        an = d.p[S::IA * C + in];
    }
    if (ip == 0) {
        kp = 0.0;
        ap = bc_alpha<MODEL, D>(d, c, y, 2 * dim + 1);
    } else if (!(d.flags[ip] & F_LEAF)) {
        double pi[S::NE];
        kp = interp_k<D>(d, ip, S::IKAPPA);
        interp_p<MODEL, D>(d, ip, pi);
        ap = pi[S::IA];
    } else {
        kp = d.k[S::IKAPPA * C + ip];
        ap = d.p[S::IA * C + ip];
    }
    const double xa = an * (1 - an), xb = ap * (1 - ap);
    const double wn = xa * xa, wp = xb * xb;
    v[0] = 0.5 * (wn * kn + wp * kp) + kc * wc;
    v[1] = 0.5 * (wn + wp) + wc;
}

// --------------------------------------------------- refinement flagging
// unsplit detection (:115-149) and split readiness (:154-165)
template <int D>
__device__ __forceinline__ void mark2_body(const Dev &d, const int32_t *__restrict__ list, int nl, int li, int L,
                                           uint8_t *ready_new, Status *st) {
    if (li >= nl) return;
    const int64_t C = d.C;
    const int32_t y = list[li];
    uint8_t m = d.mark[y];
    uint8_t rn = 0;
    const int ylev = d.level[y];
    if ((m & MK_UNSPLIT_CAND) && (d.flags[y] & F_CAN_UNSPLIT)) {
        bool ok = true;
        for (int f = 0; f < 2 * D; ++f) {
            const int32_t nb = d.nbr[f * C + y];
            if (nb != 0 && (!(d.flags[nb] & F_LEAF) || d.acti[nb] > L)) ok = false;
        }
        if (ok) m |= MK_UNSPLIT_CHILD;
    }
    if (m & MK_SPLIT) {
        bool ready = true;
        for (int f = 0; f < 2 * D; ++f) {
            const int32_t nb = d.nbr[f * C + y];
            if (nb != 0 && d.level[nb] < ylev - 1) ready = false;
        }
        if (ready) {
            m |= MK_READY;
            rn = d.child_base[y] == 0 ? 1 : 0;
            atomicAdd(&st->n_ready_split, 1);
            if (rn) atomicAdd(&st->n_new_blocks, 1);
        }
    }
    d.mark[y] = m;
    ready_new[li] = rn;
}


// Domain::split (domain.hpp:192-247) + child set-up (:174-195).  new_off is
// the exclusive scan of "needs a new child block" over the list.
template <int MODEL, int D>
__device__ __forceinline__ void split_body(const Dev &d, const MConst &c, const int32_t *__restrict__ list, int nl,
                                           int li, const int32_t *new_off, int64_t n_before, int L) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE;
    if (li >= nl) return;
    const int64_t C = d.C;
    const int32_t y = list[li];
    if (!(d.mark[y] & MK_READY)) return;
    int32_t base = d.child_base[y];
    const bool reuse = base != 0;
    if (!reuse) base = (int32_t) (n_before + (int64_t) S::NCH * new_off[li]);
    double pc[NE], sl[D][NE], ctr[D];
    for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
    for (int a = 0; a < D; ++a)
        for (int e = 0; e < NE; ++e) sl[a][e] = d.slope[(a * NE + e) * C + y];
    for (int a = 0; a < D; ++a) ctr[a] = d.center[a * C + y];
    const double ps = d.size[y], pt = d.t[y];
    const int plev = d.level[y];
    for (int ch = 0; ch < S::NCH; ++ch) {
        const int32_t x = base + ch;
        double cc[D], p[NE], u[NE];
        const double q = 0.25 * ps;
        for (int a = 0; a < D; ++a) cc[a] = ctr[a] + (((ch >> a) & 1) ? 1.0 : -1.0) * q;
        for (int e = 0; e < NE; ++e) p[e] = pc[e];
        for (int a = 0; a < D; ++a) {
            const double dx = cc[a] - ctr[a];
            for (int e = 0; e < NE; ++e) p[e] = p[e] + sl[a][e] * dx;
        }
        cons_from_prim<MODEL, D>(c, p, u);
        d.parent[x] = y;
        if (!reuse) d.child_base[x] = 0;
        for (int f = 0; f < 2 * D; ++f) {
            d.nbr[f * C + x] = 0;
            const int a = f >> 1, dir = f & 1;
            d.bc[f * C + x] = (((ch >> a) & 1) == dir) ? d.bc[f * C + y] : 0;
            for (int e = 0; e < S::NF; ++e) d.acc[(f * S::NF + e) * C + x] = 0.0;
        }
        d.brec[x] = d.brec[y];
        d.level[x] = (int8_t) (plev + 1);
        d.acti[x] = (int8_t) L;
        d.cluster[x] = (int8_t) L;
        d.buffer[x] = -1;
        d.flags[x] = F_LEAF | F_ACTIVE | F_CAN_SPLIT | F_CAN_UNSPLIT;
        d.mark[x] = 0;
        for (int a = 0; a < D; ++a) d.center[a * C + x] = cc[a];
        d.size[x] = 0.5 * ps;
        d.t[x] = pt;
        for (int e = 0; e < NE; ++e) {
            d.p[e * C + x] = p[e];
            d.u[e * C + x] = u[e];
            d.rhs[e * C + x] = 0.0;
        }
        for (int a = 0; a < D; ++a)
            for (int e = 0; e < NE; ++e) d.slope[(a * NE + e) * C + x] = sl[a][e];
        for (int e = 0; e < S::NPX; ++e) d.k[e * C + x] = d.k[e * C + y];
        for (int e = 0; e < S::NGX; ++e) d.t_slope[e * C + x] = d.t_slope[e * C + y];
        for (int a = 0; a < D; ++a)
            for (int e = 0; e < S::NGX; ++e) d.grad[(a * S::NGX + e) * C + x] = d.grad[(a * S::NGX + e) * C + y];
        d.error[x] = -1.0;
        d.ctime[x] = pt;
        d.max_dt[x] = 0.0;
        d.cfl[x] = 0.0;
    }
    d.child_base[y] = base;
    d.flags[y] = (uint8_t) (d.flags[y] & ~(F_LEAF | F_SHOULD_SPLIT));
}

// Domain::unsplit (domain.hpp:296-325) of the parents whose children are
// all marked (:198-220); one thread per parent (the one of child 0).
template <int MODEL, int D>
__device__ void interpolated_cell_acc(const Dev &d, int32_t idx, double &t, double *u, double *k, double (*g)[8],
                                      double *ts, const MConst &c);

template <int MODEL, int D>
__device__ __forceinline__ void unsplit_body(const Dev &d, const MConst &c, const int32_t *__restrict__ list, int nl,
                                             int li, int L, Status *st) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE, NCH = S::NCH;
    if (li >= nl) return;
    const int64_t C = d.C;
    const int32_t y = list[li];
    if (!(d.mark[y] & MK_UNSPLIT_CHILD)) return;
    const int32_t par = d.parent[y];
    const int32_t cb = d.child_base[par];
    if (cb != y) return;
    for (int ch = 0; ch < NCH; ++ch) {
        const int32_t x = cb + ch;
        if (!(d.cluster[x] == L && (d.mark[x] & MK_UNSPLIT_CHILD))) return;
    }
    // interpolated_cell(par): domain.hpp:150-180 (children are reached
    // through the recursion; gradient starts from the parent's stored one)
    double t = 0.0, u[NE], k[S::NPX], g[D][8], ts[S::NGX];
    for (int e = 0; e < NE; ++e) u[e] = 0.0;
    for (int e = 0; e < S::NPX; ++e) k[e] = 0.0;
    for (int e = 0; e < S::NGX; ++e) ts[e] = 0.0;
    for (int a = 0; a < D; ++a)
        for (int e = 0; e < S::NGX; ++e) g[a][e] = d.grad[(a * S::NGX + e) * C + par];
    for (int ch = 0; ch < NCH; ++ch) interpolated_cell_acc<MODEL, D>(d, cb + ch, t, u, k, g, ts, c);
    t /= NCH;
    for (int e = 0; e < S::NGX; ++e) ts[e] = ts[e] / NCH;
    for (int e = 0; e < NE; ++e) u[e] = u[e] / NCH;
    double p[NE];
    prim_from_cons<MODEL, D>(c, u, p);
    for (int e = 0; e < S::NPX; ++e) k[e] = k[e] / NCH;
    for (int a = 0; a < D; ++a)
        for (int e = 0; e < S::NGX; ++e) g[a][e] = g[a][e] / NCH;
    double err = 0.0;
    for (int ch = 0; ch < NCH; ++ch) {
        const int32_t x = cb + ch;
        err += d.error[x];
        d.flags[x] = (uint8_t) (d.flags[x] | F_GARBAGE | F_LEAF);
    }
    for (int f = 0; f < 2 * D; ++f) {
        d.nbr[f * C + par] = 0;
        for (int e = 0; e < S::NF; ++e) d.acc[(f * S::NF + e) * C + par] = 0.0;
    }
    d.t[par] = t;
    for (int e = 0; e < NE; ++e) {
        d.u[e * C + par] = u[e];
        d.p[e * C + par] = p[e];
    }
    for (int e = 0; e < S::NPX; ++e) d.k[e * C + par] = k[e];
    for (int e = 0; e < S::NGX; ++e) d.t_slope[e * C + par] = ts[e];
    for (int a = 0; a < D; ++a)
        for (int e = 0; e < S::NGX; ++e) d.grad[(a * S::NGX + e) * C + par] = g[a][e];
    d.error[par] = err;
    d.ctime[par] = t;
    d.flags[par] = (uint8_t) (d.flags[par] | F_LEAF | F_ACTIVE | F_CAN_SPLIT);
    d.acti[par] = (int8_t) L;
    d.cluster[par] = (int8_t) L;
    atomicAdd(&st->n_unsplit, 1);
}

// Accumulates interpolated_cell(idx) of one child into the running sums
// (domain.hpp:162-169); non-leaf children recurse.
template <int MODEL, int D>
__device__ void interpolated_cell_acc(const Dev &d, int32_t idx, double &t, double *u, double *k, double (*g)[8],
                                      double *ts, const MConst &c) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE;
    const int64_t C = d.C;
    double ct, cu[NE], ck[S::NPX], cg[D][8], cts[S::NGX];
    if (d.flags[idx] & F_LEAF) {
        ct = d.t[idx];
        for (int e = 0; e < NE; ++e) cu[e] = d.u[e * C + idx];
        for (int e = 0; e < S::NPX; ++e) ck[e] = d.k[e * C + idx];
        for (int a = 0; a < D; ++a)
            for (int e = 0; e < S::NGX; ++e) cg[a][e] = d.grad[(a * S::NGX + e) * C + idx];
        for (int e = 0; e < S::NGX; ++e) cts[e] = d.t_slope[e * C + idx];
    } else {
        ct = 0.0;
        for (int e = 0; e < NE; ++e) cu[e] = 0.0;
        for (int e = 0; e < S::NPX; ++e) ck[e] = 0.0;
        for (int e = 0; e < S::NGX; ++e) cts[e] = 0.0;
        for (int a = 0; a < D; ++a)
            for (int e = 0; e < S::NGX; ++e) cg[a][e] = d.grad[(a * S::NGX + e) * C + idx];
        const int32_t cb = d.child_base[idx];
        for (int ch = 0; ch < S::NCH; ++ch) interpolated_cell_acc<MODEL, D>(d, cb + ch, ct, cu, ck, cg, cts, c);
        ct /= S::NCH;
        for (int e = 0; e < S::NGX; ++e) cts[e] = cts[e] / S::NCH;
        for (int e = 0; e < NE; ++e) cu[e] = cu[e] / S::NCH;
        for (int e = 0; e < S::NPX; ++e) ck[e] = ck[e] / S::NCH;
        for (int a = 0; a < D; ++a)
            for (int e = 0; e < S::NGX; ++e) cg[a][e] = cg[a][e] / S::NCH;
    }
    t += ct;
    for (int e = 0; e < NE; ++e) u[e] = u[e] + cu[e];
    for (int e = 0; e < S::NPX; ++e) k[e] = k[e] + ck[e];
    for (int a = 0; a < D; ++a)
        for (int e = 0; e < S::NGX; ++e) g[a][e] = g[a][e] + cg[a][e];
    for (int e = 0; e < S::NGX; ++e) ts[e] = ts[e] + cts[e];
}


// dom_find started at a node h: the root descent passes through every live
// (non-garbage) node that contains the point, and stops at the first leaf or
// node of level max_level, so starting at such an h (after climbing out of
// garbage blocks and above max_level) gives the same node.  A point a cell
// centre lies at least half a cell away from every coarser node's faces and
// centre, so comparing it with h's centre picks the same children as the
// scaled coordinates of dom_find.  Falls back to dom_find when h does not
// contain the point.
template <int D>
__device__ inline int32_t dom_find_from(const Dev &d, const MConst &c, double *p, int max_level, int32_t h, int64_t n) {
    for (int i = 0; i < D; ++i) {
        const double hw = 0.5 * c.active_size[i];
        if (p[i] < -hw || p[i] > hw) {
            if (c.periodic[i]) {
                if (p[i] < -hw) p[i] += c.active_size[i];
                else p[i] -= c.active_size[i];
            } else {
                return 0;
            }
        }
    }
    if (h >= 1 && h < n) {
        const int64_t C = d.C;
        for (int g = 0; g < 64 && h > 1 && ((d.flags[h] & F_GARBAGE) || d.level[h] > max_level); ++g) h = d.parent[h];
        bool in = h >= 1 && !(d.flags[h] & F_GARBAGE) && d.level[h] <= max_level;
        if (in) {
            const double hs = 0.5 * d.size[h];
            for (int i = 0; i < D; ++i) in = in && fabs(p[i] - d.center[i * C + h]) < hs;
        }
        if (in) {
            for (int guard = 0; guard < 64; ++guard) {
                if ((d.flags[h] & F_LEAF) || d.level[h] == max_level) return h;
                int q = 0;
                for (int i = 0; i < D; ++i)
                    if (p[i] > d.center[i * C + h]) q |= 1 << i;
                h = d.child_base[h] + q;
            }
            return h;
        }
    }
    return dom_find<D>(d, c, p, max_level);
}

// update_neighbors (domain.hpp:249-293) for every active leaf.  The search
// for each face's neighbour starts at the face's previous neighbour (or, for
// a new child, at its parent for an inner face and at the parent's neighbour
// for an outer one): after a remesh it is the answer or one level off.
// One face f of node i (the faces of a node are independent: each reads and
// writes only its own link, and the parent's links are not written here).
template <int D>
__device__ __forceinline__ void neighbors_face(const Dev &d, const MConst &c, int64_t i, int f, int64_t n) {
    if (i < 1 || i >= n) return;
    const uint8_t fl = d.flags[i];
    if (!(fl & F_LEAF) || !(fl & F_ACTIVE) || (fl & F_GARBAGE)) return;
    const int64_t C = d.C;
    if (d.bc[f * C + i]) return;
    const int dim = f >> 1, dir = f & 1;
    const int lev = d.level[i];
    const double s = d.size[i];
    const int32_t par = d.parent[i];
    double loc[D];
    for (int a = 0; a < D; ++a) loc[a] = d.center[a * C + i];
    int32_t hint = d.nbr[f * C + i];
    if (hint == 0 && par >= 1 && par < n) {
        const bool inner = (dir == 1) == (loc[dim] < d.center[dim * C + par]);
        hint = inner ? par : d.nbr[f * C + par];
    }
    loc[dim] = loc[dim] + s * (2 * dir - 1.0);
    const int32_t nb = dom_find_from<D>(d, c, loc, lev, hint, n);
    if (!nb || !(d.flags[nb] & F_ACTIVE)) return;
    d.nbr[f * C + i] = nb;
    if (d.level[nb] < lev - 1) atomic_or_u8(&d.flags[nb], F_SHOULD_SPLIT);
}

template <int D>
__device__ __forceinline__ void neighbors_node(const Dev &d, const MConst &c, int64_t i, int64_t n) {
    for (int f = 0; f < 2 * D; ++f) neighbors_face<D>(d, c, i, f, n);
}

template <int D>
__global__ void __launch_bounds__(256) k_neighbors(Dev d, MConst c, int64_t n) {
    neighbors_node<D>(d, c, (int64_t) blockIdx.x * blockDim.x + threadIdx.x, n);
}

// ------------------------------------------------------- LTS reassignment
// impl/adaptive_refinement.hpp:264-322
template <int MODEL, int D>
__device__ __forceinline__ void reassign_body(const Dev &d, const MConst &c, const int32_t *__restrict__ list, int nl,
                                              int li0, int L, double dt, int counter, Status *st) {
    using S = Sz<MODEL, D>;
    if (li0 - (int) (threadIdx.x & 63) >= nl) return;  // whole wavefront past the end
    const bool live = li0 < nl;
    const int li = live ? li0 : nl - 1;
    const int64_t C = d.C;
    const int32_t y = list[li];
    if (!live || (d.mark[y] & MK_RA_DONE)) return;
    double p[S::NE];
    for (int e = 0; e < S::NE; ++e) p[e] = d.p[e * C + y];
    const double mdt = max_delta_time<MODEL, D>(c, p, d.size[y]);
    double cfl = dt / mdt;
    d.max_dt[y] = mdt;
    const uint8_t fl = d.flags[y];
    bool lowered = false;
    if (!(fl & F_LEAF) || (fl & F_GARBAGE)) {
        d.cluster[y] = -1;
    } else if (cfl < 0.5 * c.max_cfl && counter == 1 && L > 0) {
        d.cluster[y] = -1;
        d.buffer[y] = (int8_t) (L - 1);
        cfl *= 2;
        lowered = true;
    } else if (cfl > c.max_cfl && L < 15) {
        const int raise = (int) (log2(cfl / c.max_cfl) + 1);
        if (L + raise < 16) {
            d.acti[y] = (int8_t) (L + raise);
            cfl /= 1 << raise;
            d.cluster[y] = (int8_t) (L + raise);
        } else {
            d.cluster[y] = -1;
        }
    }
    d.cfl[y] = cfl;
    wave_count_add(lowered, &st->lowered);
}

// flush_acti_buffer: impl/acti_utils.hpp:43-50
__device__ __forceinline__ void flush_node(const Dev &d, int64_t i, int64_t n, int L) {
    if (i < 1 || i >= n) return;
    if (d.buffer[i] == L) {
        d.buffer[i] = -1;
        d.cluster[i] = (int8_t) L;
        d.acti[i] = (int8_t) L;
    }
}

// ------------------------------------------------------------------ init
// Integrator constructor, first loop: integrator/integrator.hpp:100-122
template <int MODEL, int D>
__global__ void k_init_cells(Dev d, MConst c, int64_t n) {
    using S = Sz<MODEL, D>;
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.C) return;
    d.scratch[i] = -INFINITY;
    if (i < 1 || i >= n) return;
    const uint8_t fl = d.flags[i];
    if (!(fl & F_LEAF) || !(fl & F_ACTIVE)) return;
    const int64_t C = d.C;
    double p[S::NE], u[S::NE];
    for (int e = 0; e < S::NE; ++e) p[e] = d.p[e * C + i];
    cons_from_prim<MODEL, D>(c, p, u);
    for (int e = 0; e < S::NE; ++e) d.u[e * C + i] = u[e];
    if constexpr (MODEL == MULTI) {
        d.k[0 * C + i] = c.mu1 * p[S::IA] + c.mu2 * (1 - p[S::IA]);
        d.k[S::IALPHAS * C + i] = p[S::IA];
    }
    for (int e = 0; e < S::NGX; ++e) d.t_slope[e * C + i] = 0.0;
    const double m = max_delta_time<MODEL, D>(c, p, d.size[i]);
    d.max_dt[i] = m;
    d.scratch[i] = m;
}

// assign_to_acti_clusters: impl/acti_utils.hpp:10-24
static __global__ void k_assign(Dev d, int64_t n, double global_dt, double max_cfl) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 1 || i >= n) return;
    const uint8_t fl = d.flags[i];
    if (!(fl & F_ACTIVE) || !(fl & F_LEAF) || (fl & F_GARBAGE)) return;
    double cfl = global_dt / d.max_dt[i];
    int lv = (int) log2(cfl / max_cfl);
    lv = lv < 0 ? 0 : (15 < lv ? 15 : lv);
    d.acti[i] = (int8_t) lv;
    d.cfl[i] = cfl / (1 << lv);
    d.cluster[i] = (int8_t) lv;
}

static __global__ void k_level_hist(const int8_t *__restrict__ cluster, int64_t n, unsigned long long *hist) {
    __shared__ unsigned int h[32];
    if (threadIdx.x < 32) h[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) {
        const int cl = cluster[i];
        if (i >= 1 && cl >= 0 && cl < 16) atomicAdd(&h[cl], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 16 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long) h[threadIdx.x]);
}

// ------------------------------------------------------------ batched launches
// One launch of an op covers one substep of every instance of a batch (a
// single integrator is a batch of one).  Item i describes an instance's node
// pool, its constants (device copy), its cluster list and its dt, counter
// and flags for this substep.  The list work of item i is cut into tiles of
// blockDim.x entries, its pool work into tiles of POOL_TILE nodes; toff[i] is
// the first tile of item i and toff[n_items] the total.  A persistent grid
// strides over the tiles; with a grid that is a multiple of 8 blocks the
// tile index is remapped so that each XCD takes a contiguous range.  The
// instances' constants come through their own restrict-qualified array, so
// the compiler may keep them in scalar registers across the ops' stores.
struct Item {
    Dev d;
    int32_t *list;
    int32_t *rest;
    Status *st;
    unsigned long long *hist;  // 16 per-level cluster counts
    int64_t n;                 // pool size
    double dt;
    int32_t nl, L, counter, signal, flush, remesh;
    // remesh of this substep: ready/new-block flags and new-block offsets per
    // list entry, pool size before the splits, whether this instance remeshes
    uint8_t *rn;
    int32_t *off;
    int64_t n_before;
    int32_t meshing;
    int32_t walign;  // dense clusters: node ids of window starts are walign mod 64
};
constexpr int POOL_TILE = 2048;

__device__ __forceinline__ int find_item(const int *__restrict__ toff, int n_items, int tile) {
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (toff[mid] <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <class Op>
__global__ void __launch_bounds__(256) k_batch(const Item *__restrict__ items, const MConst *__restrict__ mcs,
                                               const int *__restrict__ toff, int n_items, Op op) {
    const int total = toff[n_items];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int tile = xcd_block(t, total);
        const int i = find_item(toff, n_items, tile);
        op(items[i], mcs[i], (tile - toff[i]) * (int) blockDim.x + (int) threadIdx.x);
    }
}


// The set-aside cells are few and each is a long chain of dependent
// loads: spread them round-robin over all of the item's tiles (blocks), so
// that every CU takes some instead of the first few taking all.  Launched
// over the list tiles (k_batch, list index li) or over the set-aside cells
// themselves (k_rest, r).
template <int MODEL, int D>
struct OpFluxRest {
    __device__ __forceinline__ int count(const Item &it) const { return it.st->n_rest; }
    __device__ __forceinline__ void cell(const Item &it, const MConst &c, int r) const {
        flux_cell<MODEL, D>(it.d, c, it.rest[r], it.L, it.dt);
    }
    __device__ __forceinline__ void operator()(const Item &it, const MConst &c, int li) const {
        const int B = (int) blockDim.x, T = (it.nl + B - 1) / B;
        const int r = (li % B) * T + li / B;
        if (r >= it.st->n_rest) return;
        cell(it, c, r);
    }
};
template <int MODEL, int D>
struct OpDepositCoarse {
    // k_rest: a (set-aside cell, face) a thread
    __device__ __forceinline__ int count(const Item &it) const { return it.st->n_rest * 2 * D; }
    __device__ __forceinline__ void cell(const Item &it, const MConst &, int r) const {
        deposit_coarse_face<MODEL, D>(it.d, it.rest[r / (2 * D)], r % (2 * D), it.L);
    }
    __device__ __forceinline__ void operator()(const Item &it, const MConst &c, int li) const {
        const int B = (int) blockDim.x, T = (it.nl + B - 1) / B;
        const int r = (li % B) * T + li / B;
        if (r >= it.st->n_rest) return;
        deposit_coarse<MODEL, D>(it.d, it.rest[r], it.L);
    }
};

// An op over the set-aside cells of each item with a list, r = 0 ..
// op.count(item) - 1 (a cell, or a (cell, face) pair), spread round-robin
// over the blocks.  For a single instance (each block reads every item's
// count): next to the window ops the grid does not stride over the list's
// tiles (a dense cluster has tens of thousands) to find the few cells, and
// the coarse deposits of small lists take a thread per (cell, face).
template <class Op>
__global__ void __launch_bounds__(256) k_rest(const Item *__restrict__ items, const MConst *__restrict__ mcs,
                                              int n_items, Op op) {
    const int G = (int) gridDim.x, B = (int) blockDim.x;
    for (int i = 0; i < n_items; ++i) {
        const Item &it = items[i];
        if (it.nl == 0) continue;
        const int nr = op.count(it);
        for (int r = (int) blockIdx.x + G * (int) threadIdx.x; r < nr; r += G * B) op.cell(it, mcs[i], r);
    }
}

// Lean flux op for the cells of a cluster whose faces are all domain
// boundaries or same-tree-level leaves (every cell of a uniform tree, most
// cells of an adaptive one): one flux evaluation site per face, no sub-face
// walks (impl/flux.hpp:32-91).  A cell with any other face is appended to
// `rest` (order irrelevant: cells are independent) for OpFluxRest.
// A cell's axes run on D wavefronts of one block: 64 listed cells a block,
// wavefront w takes axis w, classifies the axis's two neighbours and
// evaluates or receives the cell's - and + face of that axis.  A face
// between two listed cells of the wavefront (consecutive ids sit in
// consecutive lanes: siblings, runs of a level) is evaluated once, by its
// owner as its + face, and handed to the other cell by a lane permute.  The
// first wavefront adds the axes' contributions in the cell's face order (-
// then + per axis), so the sums are those of one thread evaluating every
// face of the cell in order.  Each wavefront's chain is one axis long and
// there are D times as many as with one thread per cell (measured 1-2%
// faster on every workload, most on small clusters).
// Grid: tile_cells / 64 blocks for each list tile of tile_cells cells (a
// multiple of 64), blockDim = 64 * D.
template <int MODEL, int D>
__global__ void __launch_bounds__(64 * D) k_flux_axes(const Item *__restrict__ items, const MConst *__restrict__ mcs,
                                                      const int *__restrict__ toff, int n_items, int tile_cells) {
    using S = Sz<MODEL, D>;
    constexpr int NF = S::NF, NE = S::NE;
    __shared__ double P[D - 1][2][NF][64];  // the - and + contributions of axes 1 .. D-1
    __shared__ uint8_t SIMPLE[D][64];
    const int per_tile = tile_cells >> 6;
    const int total = toff[n_items] * per_tile;
    const int lane = (int) (threadIdx.x & 63), dim = (int) (threadIdx.x >> 6);
    const int fm = 2 * dim, fp = 2 * dim + 1;
    for (int g = (int) blockIdx.x; g < total; g += (int) gridDim.x) {
        const int gg = xcd_block(g, total);
        const int tile = gg / per_tile;
        const int i = find_item(toff, n_items, tile);
        const Item &it = items[i];
        const MConst &c = mcs[i];
        const Dev &d = it.d;
        const int64_t C = d.C;
        const int L = it.L;
        const double dt = it.dt;
        const int li = (tile - toff[i]) * tile_cells + (gg % per_tile) * 64 + lane;
        const bool valid = li < it.nl;
        const int32_t y = valid ? it.list[li] : -1;
        const int ylev = valid ? d.level[y] : 0;
        // this axis's neighbours, their LTS level and cluster
        const int32_t nbm = valid ? d.nbr[fm * C + y] : 0;
        const int32_t nbp = valid ? d.nbr[fp * C + y] : 0;
        const int nacm = nbm ? d.acti[nbm] : 0, nclm = nbm ? d.cluster[nbm] : -1;
        const int nacp = nbp ? d.acti[nbp] : 0, nclp = nbp ? d.cluster[nbp] : -1;
        bool simple = true;
        if (nbm != 0 && (!(d.flags[nbm] & F_LEAF) || d.level[nbm] != ylev)) simple = false;
        if (nbp != 0 && (!(d.flags[nbp] & F_LEAF) || d.level[nbp] != ylev)) simple = false;
        SIMPLE[dim][lane] = simple;
        __syncthreads();
#pragma unroll
        for (int a = 0; a < D; ++a) simple = simple && SIMPLE[a][lane];
        if (dim == 0 && valid && !simple) it.rest[atomicAdd(&it.st->n_rest, 1)] = y;
        const bool act = valid && simple;
        if (!__syncthreads_or(act)) continue;  // uniform over the block
        const double ys = act ? d.size[y] : 0.0;
        const double area = c.area[ylev];
        Side<MODEL, D> sy;
        double yc = 0.0;
        if (act) {
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                sy.p[e] = d.p[e * C + y];
                sy.r[e] = d.rhs[e * C + y];
                sy.s[e] = d.slope[(dim * NE + e) * C + y];
            }
            sy.t = d.t[y];
            if constexpr (MODEL == MULTI) {
                sy.kap = d.k[S::IKAPPA * C + y];
                sy.kn = d.k[(S::IKN + dim) * C + y];
            }
            if constexpr (S::NG > 0) {
                sy.g = d.grad + y;
                sy.gC = C;
            }
            yc = d.center[dim * C + y];
        }
        // the face flux of (dim, dir) from the owner's side: y's own face
        // (y_computes) or the neighbour's face toward y (impl/flux.hpp:58-91)
        auto eval = [&](int dir, int32_t nb, bool y_computes, double *F) {
            Side<MODEL, D> sn;
            load_side<MODEL, D>(d, nb, dim, sn);
            const double nbc = d.center[dim * C + nb];
            const double oc = y_computes ? yc : nbc;
            const int odir = y_computes ? dir : 1;
            const double fc = oc + ys * (odir - 0.5);
            const double t0 = y_computes ? sy.t : sn.t;
            const double t1 = t0 + dt;
            double other = y_computes ? nbc : yc;
            if (c.periodic[dim]) {
                if (odir == 1) {
                    if (oc + ys > 0.5 * c.active_size[dim]) other += c.active_size[dim];
                } else {
                    if (oc - ys < -0.5 * c.active_size[dim]) other -= c.active_size[dim];
                }
            }
            const double cy = y_computes ? yc : other;
            const double cn = y_computes ? other : nbc;
            if (dir == 1) model_flux_axis<MODEL, D>(c, sy, cy, sn, cn, fc, area, t0, t1, dim, F);
            else model_flux_axis<MODEL, D>(c, sn, cn, sy, cy, fc, area, t0, t1, dim, F);
        };
        // + face first: y evaluates it unless the neighbour is on a finer
        // LTS level (then its deposit is received)
        double ccp[NF], F[NF];
#pragma unroll
        for (int e = 0; e < NF; ++e) ccp[e] = 0.0, F[e] = 0.0;
        bool shares = false;  // F holds y's + face flux toward a neighbour of the cluster
        if (!act) {
        } else if (nbp == 0) {
            boundary_flux<MODEL, D>(d, c, y, fp, dim, 1, dt, F);
#pragma unroll
            for (int e = 0; e < NF; ++e) ccp[e] = ccp[e] - F[e];
        } else if (nacp > L) {
#pragma unroll
            for (int e = 0; e < NF; ++e) {
                ccp[e] = d.acc[(fp * NF + e) * C + y];
                d.acc[(fp * NF + e) * C + y] = 0.0;
            }
        } else {
            eval(1, nbp, true, F);
#pragma unroll
            for (int e = 0; e < NF; ++e) ccp[e] = ccp[e] - F[e];
            if (nclp != L) {
#pragma unroll
                for (int e = 0; e < NF; ++e) {
                    GMEM double &a = d.acc[(fm * NF + e) * C + nbp];
                    a = a + F[e];
                }
            } else {
                shares = true;
            }
        }
        // the - neighbour's + face flux, when the - neighbour is listed in
        // this group and evaluated that face toward y (every lane takes part)
        const int64_t dl = (int64_t) y - nbm;
        const int src = nbm != 0 && dl > 0 && dl <= lane ? lane - (int) dl : lane;
        const int32_t src_id = __shfl(y, src);
        const int32_t src_nbp = __shfl(nbp, src);
        const int src_shares = __shfl((int) shares, src);
        double X[NF];
#pragma unroll
        for (int e = 0; e < NF; ++e) X[e] = __shfl(F[e], src);
        const bool handed = src != lane && src_id == nbm && src_nbp == y && src_shares;
        double ccm[NF];
#pragma unroll
        for (int e = 0; e < NF; ++e) ccm[e] = 0.0;
        if (!act) {
        } else if (nbm == 0) {
            boundary_flux<MODEL, D>(d, c, y, fm, dim, 0, dt, F);
#pragma unroll
            for (int e = 0; e < NF; ++e) ccm[e] = ccm[e] - F[e];
        } else {
            if (nacm > L) {
#pragma unroll
                for (int e = 0; e < NF; ++e) {
                    ccm[e] = d.acc[(fm * NF + e) * C + y];
                    d.acc[(fm * NF + e) * C + y] = 0.0;
                }
            }
            // nbm is a same-level leaf (y is simple), so its + neighbour is y
            // and it evaluates the face when it updates in this substep
            const bool y_computes = nacm < L;
            const bool nb_computes = !y_computes && nclm == L;
            if (nb_computes && handed) {
#pragma unroll
                for (int e = 0; e < NF; ++e) ccm[e] = ccm[e] + X[e];
            } else if (y_computes || nb_computes) {
                eval(0, nbm, y_computes, F);
#pragma unroll
                for (int e = 0; e < NF; ++e) ccm[e] = ccm[e] + F[e];
                if (y_computes && nclm != L) {
#pragma unroll
                    for (int e = 0; e < NF; ++e) {
                        GMEM double &a = d.acc[(fp * NF + e) * C + nbm];
                        a = a - F[e];
                    }
                }
            }
        }
        if (dim > 0) {
#pragma unroll
            for (int e = 0; e < NF; ++e) {
                P[dim - 1][0][e][lane] = ccm[e];
                P[dim - 1][1][e][lane] = ccp[e];
            }
        }
        __syncthreads();
        if (dim == 0 && act) {
            // y's faces in their order: - then + of each axis
            double tot[NF];
#pragma unroll
            for (int e = 0; e < NF; ++e) tot[e] = 0.0;
#pragma unroll
            for (int e = 0; e < NF; ++e) tot[e] = tot[e] + ccm[e];
#pragma unroll
            for (int e = 0; e < NF; ++e) tot[e] = tot[e] + ccp[e];
#pragma unroll
            for (int a = 1; a < D; ++a) {
#pragma unroll
                for (int e = 0; e < NF; ++e) tot[e] = tot[e] + P[a - 1][0][e][lane];
#pragma unroll
                for (int e = 0; e < NF; ++e) tot[e] = tot[e] + P[a - 1][1][e][lane];
            }
#pragma unroll
            for (int e = 0; e < NF; ++e) d.tot[e * C + y] = tot[e];
        }
        __syncthreads();  // SIMPLE and P are reused by the next group
    }
}

// ------------------------------------------------------- dense clusters
// The lean flux op for a cluster that fills most of the pool (a uniform or
// nearly uniform level), as one wavefront per aligned window of 64
// consecutive node ids: a 4x4x4 (8x8) Morton block of a breadth-first
// level.  A face between two same-level leaves of the window that both
// update in this substep is evaluated once, by its left cell, and the
// right cell takes the flux from the left cell's lane; every other face a
// cell needs is evaluated in one compacted pass over the window.  Which
// side evaluates a face, the fluxes received from finer neighbours, the
// deposits for neighbours on other LTS levels and the per-cell face order
// of the sum are k_flux_axes's (impl/flux.hpp:32-91), so the results
// equal that op's bit for bit.  Cells with a face to another tree level go
// to the rest list, as there.  One instance (items[0]); the windows cover
// the ids from the cluster list's first to its last entry, aligned to
// it.walign.
enum : int { WK_NONE = 0, WK_BOUNDARY = 1, WK_OWN = 2, WK_NEIGHBOUR = 3, WK_SHARED = 4 };

// a window lane's cell side from lane o (every lane takes part)
template <int MODEL, int D>
__device__ __forceinline__ void shfl_side(const Side<MODEL, D> &a, int o, Side<MODEL, D> &b) {
    using S = Sz<MODEL, D>;
#pragma unroll
    for (int e = 0; e < S::NE; ++e) {
        b.p[e] = __shfl(a.p[e], o);
        b.s[e] = __shfl(a.s[e], o);
        b.r[e] = __shfl(a.r[e], o);
    }
    b.t = __shfl(a.t, o);
    if constexpr (MODEL == MULTI) {
        b.kap = __shfl(a.kap, o);
        b.kn = __shfl(a.kn, o);
    }
    if constexpr (S::NG > 0) {
        b.g = (const GMEM double *) (uintptr_t) __shfl((unsigned long long) (uintptr_t) a.g, o);
        b.gC = a.gC;
    }
}

// Every lane holds its own node's state, time and (per axis) slope, so the
// side of a neighbour inside the window comes from that neighbour's lane
// (lane permutes), and so do its metadata word and the + face fluxes; only
// neighbours outside the window are read from memory.  LDS holds only the
// compacted pass (queue and results).  Three waves per SIMD: the register
// budget of a fourth (128 VGPRs) spills and ran 10% slower.  Two for the
// five-equation model, whose sides spill heavily in 168 VGPRs.
template <int MODEL, int D>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MODEL == MULTI ? 2 : 3))) k_flux_window(const Item *__restrict__ items, const MConst *__restrict__ mcs) {
    using S = Sz<MODEL, D>;
    constexpr int NF = S::NF, NE = S::NE;
    __shared__ double M[D][NF][64];  // - face flux evaluated in the compacted pass
    __shared__ int32_t Q[D * 64];    // compacted pass: lane | dim << 6 | kind << 8 | deposit << 11
    const Item &it = items[0];
    const MConst &c = mcs[0];
    const int nl = it.nl;
    if (nl == 0) return;
    const Dev &d = it.d;
    const int64_t C = d.C;
    const int L = it.L;
    const double dt = it.dt;
    const int64_t al = it.walign;
    const int64_t w0 = ((int64_t) it.list[0] - al) >> 6, w1 = ((int64_t) it.list[nl - 1] - al) >> 6;
    const int nw = (int) (w1 - w0 + 1);
    const int lane = (int) threadIdx.x;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int k = (int) blockIdx.x; k < nw; k += (int) gridDim.x) {
        const int64_t base = al + (w0 + xcd_block(k, nw)) * 64;
        const int64_t yid = base + lane;
        const bool valid = yid >= 1 && yid < it.n;
        const int32_t y = (int32_t) yid;
        // the lane's own node: metadata, state, time
        const uint32_t ymeta = valid ? d.meta[y] : 0u;
        const bool member = valid && ((ymeta >> 24) & 0xff) == (uint32_t) L;
        const int ylev = (int) ((ymeta >> 8) & 0xff);
        Side<MODEL, D> sy;
        double ys = 0.0;
        if (valid) {
            ys = d.size[y];
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                sy.p[e] = d.p[e * C + y];
                sy.r[e] = d.rhs[e * C + y];
            }
            sy.t = d.t[y];
            if constexpr (MODEL == MULTI) sy.kap = d.k[S::IKAPPA * C + y];
            if constexpr (S::NG > 0) {
                sy.g = d.grad + y;
                sy.gC = C;
            }
        }
        // the neighbour ids are loaded with the node's own data (for every
        // valid lane: one round of memory latency fewer) and kept for members
        int32_t nbs[2 * D];
#pragma unroll
        for (int f = 0; f < 2 * D; ++f) {
            const int32_t nb = valid ? d.nbr[f * C + y] : 0;
            nbs[f] = member ? nb : 0;
        }
        // face f's neighbour: ACTI level in bits 4f..4f+3, cluster + 1 in
        // bits 5f..5f+4 (cluster is -1..15); its metadata from its lane when
        // it lies in the window
        uint32_t nacp = 0, nclp = 0;
        bool simple = true;
#pragma unroll
        for (int f = 0; f < 2 * D; ++f) {
            const int32_t nb = nbs[f];
            const int64_t o64 = (int64_t) nb - base;
            const bool inr = nb != 0 && o64 >= 0 && o64 < 64;
            uint32_t mw = __shfl(ymeta, inr ? (int) o64 : lane);
            if (nb != 0 && !inr) mw = d.meta[nb];
            if (nb) {
                nacp |= ((mw >> 16) & 15u) << (4 * f);
                nclp |= ((((mw >> 24) & 0xff) + 1) & 31u) << (5 * f);
                if (!(mw & F_LEAF) || (int) ((mw >> 8) & 0xff) != ylev) simple = false;
            }
        }
        if (member && !simple) it.rest[atomicAdd(&it.st->n_rest, 1)] = y;
        const bool part = member && simple;
        const uint64_t pm = __ballot(part);
        // a + face is evaluated by the lane's cell unless the neighbour is on
        // a finer LTS level (whose deposit is received in the sum)
        int cp = 0;
#pragma unroll
        for (int dim = 0; dim < D; ++dim)
            if (part && (nbs[2 * dim + 1] == 0 || (int) ((nacp >> (4 * (2 * dim + 1))) & 15) <= L)) cp |= 1 << dim;
        // - faces: the left cell's evaluation when it is a window lane that
        // evaluates the face, else queued for the compacted pass
        int kinds = 0;  // 3 bits per axis
        int qn = 0;
#pragma unroll
        for (int dim = 0; dim < D; ++dim) {
            const int f = 2 * dim;
            const int32_t nbm = nbs[f];
            const int64_t o64 = (int64_t) nbm - base;
            const bool inr = nbm != 0 && o64 >= 0 && o64 < 64;
            const int o = inr ? (int) o64 : lane;
            // the window neighbour's + face decisions (every lane takes part)
            const int cp_w = __shfl(cp, o);
            int kind = WK_NONE, dep = 0;
            if (part) {
                const int nac = (int) ((nacp >> (4 * f)) & 15), ncl = (int) ((nclp >> (5 * f)) & 31) - 1;
                if (nbm == 0) {
                    kind = WK_BOUNDARY;
                } else if (nac < L) {
                    kind = WK_OWN;
                    dep = ncl != L;
                } else if (ncl == L) {
                    // nbm is a same-level leaf (y takes part), so its + neighbour
                    // is y: it evaluates the face
                    const bool inwin = inr && ((pm >> o) & 1);
                    kind = inwin && ((cp_w >> dim) & 1) ? WK_SHARED : WK_NEIGHBOUR;
                }
            }
            kinds |= kind << (3 * dim);
            const bool q = kind == WK_BOUNDARY || kind == WK_OWN || kind == WK_NEIGHBOUR;
            const uint64_t qm = __ballot(q);
            if (q) Q[qn + __popcll(qm & lt)] = lane | dim << 6 | kind << 8 | dep << 11;
            qn += __popcll(qm);
        }
        __syncthreads();
        for (int j0 = 0; j0 < qn; j0 += 64) {
            const int j = j0 + lane;
            const bool act = j < qn;
            const int q = act ? Q[j] : lane;
            const int zl = q & 63, dim = (q >> 6) & 3;
            // the queued cell's state, time, size, level and - neighbours
            // from its lane (every lane takes part)
            Side<MODEL, D> sz;
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                sz.p[e] = __shfl(sy.p[e], zl);
                sz.r[e] = __shfl(sy.r[e], zl);
            }
            sz.t = __shfl(sy.t, zl);
            const double zs = __shfl(ys, zl);
            const int zlev = __shfl(ylev, zl);
            int zm[D];
#pragma unroll
            for (int a = 0; a < D; ++a) zm[a] = __shfl(nbs[2 * a], zl);
            if (act) {
                const int kind = (q >> 8) & 7;
                const int32_t z = (int32_t) (base + zl);
                double F[NF];
                if (kind == WK_BOUNDARY) {
                    boundary_flux<MODEL, D>(d, c, z, 2 * dim, dim, 0, dt, F);
                } else {
                    const int32_t nb = pick_i<D>(zm, dim);
#pragma unroll
                    for (int e = 0; e < NE; ++e) sz.s[e] = d.slope[(dim * NE + e) * C + z];
                    if constexpr (MODEL == MULTI) {
                        sz.kap = d.k[S::IKAPPA * C + z];
                        sz.kn = d.k[(S::IKN + dim) * C + z];
                    }
                    if constexpr (S::NG > 0) {
                        sz.g = d.grad + z;
                        sz.gC = C;
                    }
                    Side<MODEL, D> sn;
                    load_side<MODEL, D>(d, nb, dim, sn);
                    const double zc = d.center[dim * C + z], nc = d.center[dim * C + nb];
                    if (kind == WK_OWN) {
                        face_flux_sides_body<MODEL, D>(c, sz, zc, zs, zlev, sn, nc, dim, 0, dt, F);
                        if ((q >> 11) & 1) {
#pragma unroll
                            for (int e = 0; e < NF; ++e) {
                                GMEM double &a = d.acc[((2 * dim + 1) * NF + e) * C + nb];
                                a = a - F[e];
                            }
                        }
                    } else {
                        face_flux_sides_body<MODEL, D>(c, sn, nc, zs, zlev, sz, zc, dim, 1, dt, F);
                    }
                }
#pragma unroll
                for (int e = 0; e < NF; ++e) M[dim][e][zl] = F[e];
            }
        }
        __syncthreads();
        // + faces axis by axis; after each axis the cell adds its - and +
        // face in its face order
        double tot[NF];
#pragma unroll
        for (int e = 0; e < NF; ++e) tot[e] = 0.0;
#pragma unroll 1
        for (int dim = 0; dim < D; ++dim) {
            const int fm = 2 * dim, fp = 2 * dim + 1;
            double yc = 0.0;
            if (valid) {
#pragma unroll
                for (int e = 0; e < NE; ++e) sy.s[e] = d.slope[(dim * NE + e) * C + y];
                if constexpr (MODEL == MULTI) sy.kn = d.k[(S::IKN + dim) * C + y];
                yc = d.center[dim * C + y];
            }
            const int32_t nbp = pick_i<2 * D>(nbs, fp);
            const int64_t op64 = (int64_t) nbp - base;
            const bool inrp = nbp != 0 && op64 >= 0 && op64 < 64;
            // the + neighbour's side and centre from its lane (every lane takes part)
            const int op = inrp ? (int) op64 : lane;
            Side<MODEL, D> sn;
            shfl_side<MODEL, D>(sy, op, sn);
            double cn = __shfl(yc, op);
            double F[NF];
#pragma unroll
            for (int e = 0; e < NF; ++e) F[e] = 0.0;
            if ((cp >> dim) & 1) {
                if (nbp == 0) {
                    boundary_flux<MODEL, D>(d, c, y, fp, dim, 1, dt, F);
                } else {
                    if (!inrp) {
                        load_side<MODEL, D>(d, nbp, dim, sn);
                        cn = d.center[dim * C + nbp];
                    }
                    face_flux_sides_body<MODEL, D>(c, sy, yc, ys, ylev, sn, cn, dim, 1, dt, F);
                    if ((int) ((nclp >> (5 * fp)) & 31) - 1 != L) {
#pragma unroll
                        for (int e = 0; e < NF; ++e) {
                            GMEM double &a = d.acc[(fm * NF + e) * C + nbp];
                            a = a + F[e];
                        }
                    }
                }
            }
            // the - neighbour's + face flux, from its lane (every lane takes part)
            const int32_t nbm = pick_i<2 * D>(nbs, fm);
            const int kind = (kinds >> (3 * dim)) & 7;
            const int om = kind == WK_SHARED ? (int) ((int64_t) nbm - base) : lane;
            double X[NF];
#pragma unroll
            for (int e = 0; e < NF; ++e) X[e] = __shfl(F[e], om);
            if (part) {
                const bool takem = nbm != 0 && (int) ((nacp >> (4 * fm)) & 15) > L;
                double cc[NF];
#pragma unroll
                for (int e = 0; e < NF; ++e) {
                    cc[e] = 0.0;
                    if (takem) {
                        cc[e] = d.acc[(fm * NF + e) * C + y];
                        d.acc[(fm * NF + e) * C + y] = 0.0;
                    }
                }
                if (kind == WK_SHARED) {
#pragma unroll
                    for (int e = 0; e < NF; ++e) cc[e] = cc[e] + X[e];
                } else if (kind == WK_BOUNDARY) {
#pragma unroll
                    for (int e = 0; e < NF; ++e) cc[e] = cc[e] - M[dim][e][lane];
                } else if (kind != WK_NONE) {
#pragma unroll
                    for (int e = 0; e < NF; ++e) cc[e] = cc[e] + M[dim][e][lane];
                }
#pragma unroll
                for (int e = 0; e < NF; ++e) tot[e] = tot[e] + cc[e];
                const bool takep = nbp != 0 && (int) ((nacp >> (4 * fp)) & 15) > L;
#pragma unroll
                for (int e = 0; e < NF; ++e) {
                    cc[e] = 0.0;
                    if (takep) {
                        cc[e] = d.acc[(fp * NF + e) * C + y];
                        d.acc[(fp * NF + e) * C + y] = 0.0;
                    }
                }
                if ((cp >> dim) & 1) {
#pragma unroll
                    for (int e = 0; e < NF; ++e) cc[e] = cc[e] - F[e];
                }
#pragma unroll
                for (int e = 0; e < NF; ++e) tot[e] = tot[e] + cc[e];
            }
        }
        if (part) {
#pragma unroll
            for (int e = 0; e < NF; ++e) d.tot[e * C + y] = tot[e];
        }
        // the next window's queue and results reuse the LDS
        __syncthreads();
    }
}

// The sweep (slopes, rhs, error, marks, LTS decision) of a dense cluster by
// the windows of k_flux_window.  A window stages the states of its 64 ids in
// LDS and loads those of its members' out-of-window neighbours (the halo)
// into LDS in one batch, so a cell's neighbour reads wait for two rounds of
// memory latency instead of one per face.  Cells with a neighbour that is
// not a same-level leaf go to the rest list (counted in n_rest2; the flux
// ops are done with it) for the per-cell sweep.  Both evaluate sweep_axis
// and sweep_decide on the same values, so the results are the same bit for
// bit.
template <int MODEL, int D>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MODEL == MULTI ? 2 : 3))) k_sweep_window(const Item *__restrict__ items, const MConst *__restrict__ mcs) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE, HC = 128;
    __shared__ double P[NE][64];   // states of the window's ids
    __shared__ double T[64];       // their times (gradients only)
    __shared__ double H[NE][HC];   // halo states
    __shared__ double HT[HC];
    __shared__ int32_t HN[HC];     // halo node ids
    __shared__ int32_t FLV[64];    // flags | level << 8 of the window's ids
    __shared__ int32_t HFLV[HC];
    const Item &it = items[0];
    const MConst &c = mcs[0];
    const int nl = it.nl;
    if (nl == 0 || it.st->nan_flag) return;  // integrate() stops after the update (impl/integrate.hpp:50)
    const Dev &d = it.d;
    const int64_t C = d.C;
    const int L = it.L;
    const int64_t al = it.walign;
    const int64_t w0 = ((int64_t) it.list[0] - al) >> 6, w1 = ((int64_t) it.list[nl - 1] - al) >> 6;
    const int nw = (int) (w1 - w0 + 1);
    const int lane = (int) threadIdx.x;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int k = (int) blockIdx.x; k < nw; k += (int) gridDim.x) {
        const int64_t base = al + (w0 + xcd_block(k, nw)) * 64;
        const int64_t yid = base + lane;
        const bool valid = yid >= 1 && yid < it.n;
        const int32_t y = (int32_t) yid;
        bool member = false;
        int flv = 0;
        if (valid) {
            // the metadata word the dense list selection stored this substep
            const uint32_t m = d.meta[y];
            member = ((m >> 24) & 0xff) == (uint32_t) L;
            flv = (int) (m & 0xffff);
#pragma unroll
            for (int e = 0; e < NE; ++e) P[e][lane] = d.p[e * C + y];
            if constexpr (S::NG > 0) T[lane] = d.t[y];
        }
        FLV[lane] = flv;
        // the member's own size and flags, in flight during the halo phase
        double ys = 0.0;
        uint8_t yfl = 0;
        if (member) {
            ys = d.size[y];
            yfl = d.flags[y];
        }
        int32_t nbs[2 * D];
        int hs[2 * D];
        int qn = 0;
#pragma unroll
        for (int f = 0; f < 2 * D; ++f) {
            // loaded for every valid lane, with the node's own data
            const int32_t nv = valid ? d.nbr[f * C + y] : 0;
            const int32_t nb = member ? nv : 0;
            nbs[f] = nb;
            const int64_t o = (int64_t) nb - base;
            const bool out = nb != 0 && !(o >= 0 && o < 64);
            const uint64_t m = __ballot(out);
            const int pos = qn + __popcll(m & lt);
            hs[f] = out ? pos : -1;
            if (out && pos < HC) HN[pos] = nb;
            qn += __popcll(m);
        }
        __syncthreads();
        const int qh = qn < HC ? qn : HC;
        for (int j = lane; j < qh; j += 64) {
            const int32_t nb = HN[j];
#pragma unroll
            for (int e = 0; e < NE; ++e) H[e][j] = d.p[e * C + nb];
            if constexpr (S::NG > 0) HT[j] = d.t[nb];
            HFLV[j] = (int) (d.meta[nb] & 0xffff);
        }
        __syncthreads();
        if (member) {
            const int ylev = (int8_t) (flv >> 8);
            bool fast = true;
#pragma unroll
            for (int f = 0; f < 2 * D; ++f) {
                const int32_t nb = nbs[f];
                if (nb == 0) continue;
                const int h = hs[f];
                const int v = h < 0 ? FLV[nb - base] : h < HC ? HFLV[h] : (int) (d.meta[nb] & 0xffff);
                if (!(v & F_LEAF) || ((v >> 8) & 0xff) != (flv >> 8)) fast = false;
            }
            if (!fast) {
                it.rest[atomicAdd(&it.st->n_rest2, 1)] = y;
            } else {
                double pc[NE], ctr[D], rhs[NE];
#pragma unroll
                for (int e = 0; e < NE; ++e) pc[e] = P[e][lane];
                // the centre only serves the refinement-zone test of a cell
                // that may still split
#pragma unroll
                for (int a = 0; a < D; ++a) ctr[a] = (flv & F_CAN_SPLIT) ? d.center[a * C + y] : 0.0;
                const double yt = S::NG > 0 ? T[lane] : 0.0;  // the time enters the gradients only
                const double kap = MODEL == MULTI ? d.k[S::IKAPPA * C + y] : 0.0;
                double a_ip_in, a_in_ip, a_in_ia;
                rhs_coeffs<MODEL, D>(c, pc, kap, a_ip_in, a_in_ip, a_in_ia);
#pragma unroll
                for (int e = 0; e < NE; ++e) rhs[e] = 0.0;
                double ts[S::NGX];
                if constexpr (S::NG > 0)
                    for (int e = 0; e < S::NG; ++e) ts[e] = d.t_slope[e * C + y];
                // a neighbour's state: a domain face's ghost, the window, the
                // halo, or (beyond the halo's capacity) global memory
                auto fetch = [&](int f, int dir, double *pv, double &tv, double &xv, bool &hv) {
                    const int32_t nb = pick_i<2 * D>(nbs, f);
                    if (nb == 0) {
                        const double tb = S::NG > 0 ? yt : d.t[y];
                        boundary_value_at<MODEL, D>(d, c, y, f, pc, tb, pv);
                        xv = dir ? ys : -ys;
                        tv = tb;
                        hv = false;
                        return;
                    }
                    hv = true;
                    const int h = pick_i<2 * D>(hs, f);
                    tv = 0.0;
                    if (h < 0) {
                        const int o = (int) (nb - base);
#pragma unroll
                        for (int e = 0; e < NE; ++e) pv[e] = P[e][o];
                        if constexpr (S::NG > 0) tv = T[o];
                    } else if (h < HC) {
#pragma unroll
                        for (int e = 0; e < NE; ++e) pv[e] = H[e][h];
                        if constexpr (S::NG > 0) tv = HT[h];
                    } else {
#pragma unroll
                        for (int e = 0; e < NE; ++e) pv[e] = d.p[e * C + nb];
                        tv = d.t[nb];
                    }
                    xv = dir ? 0.5 * (ys + ys) : -0.5 * (ys + ys);
                };
                double err = 0.0;
#pragma unroll 1
                for (int dim = 0; dim < D; ++dim) {
                    double pn[NE], pp[NE], xn, xp, tn, tp;
                    bool hn, hp;
                    fetch(2 * dim, 0, pn, tn, xn, hn);
                    fetch(2 * dim + 1, 1, pp, tp, xp, hp);
                    sweep_axis<MODEL, D>(d, c, y, dim, 0, pc, pn, pp, pn, pp, xn, xp, tn, tp, hn, hp, ts, kap, a_ip_in,
                                         a_in_ip, a_in_ia, rhs, err);
                }
#pragma unroll
                for (int e = 0; e < NE; ++e) __builtin_nontemporal_store(rhs[e], &d.rhs[e * C + y]);
                sweep_decide<MODEL, D>(d, c, y, pc, ctr, ys, ylev, L, it.dt, it.counter, false, err, it.st, it.remesh, yfl);
            }
        }
        __syncthreads();
    }
}

template <int MODEL, int D>
struct OpSweepRest {
    __device__ __forceinline__ int count(const Item &it) const { return it.st->n_rest2; }
    __device__ __forceinline__ void cell(const Item &it, const MConst &c, int r) const {
        sweep_body<MODEL, D>(it.d, c, it.rest, it.st->n_rest2, r, it.L, it.dt, it.counter, 0, it.st, it.remesh);
    }
    __device__ __forceinline__ void operator()(const Item &it, const MConst &c, int li) const { cell(it, c, li); }
};

template <int MODEL, int D>
struct OpUpdate {
    __device__ __forceinline__ void operator()(const Item &it, const MConst &c, int li) const {
        update_body<MODEL, D>(it.d, c, it.list, it.nl, li, it.dt, it.signal, it.st);
    }
};
template <int MODEL, int D>
struct OpSweep {
    int pass;
    __device__ __forceinline__ void operator()(const Item &it, const MConst &c, int li) const {
        if (pass == 1 && it.st->n_deferred == 0) return;
        sweep_body<MODEL, D>(it.d, c, it.list, it.nl, li, it.L, it.dt, it.counter, pass, it.st, it.remesh);
    }
};

// The sweep of a cluster's listed cells (sweep_body) with a cell's axes on
// D wavefronts of one block, as k_flux_axes: wavefront w fetches the cell's
// two axis-w neighbours and computes the limited slope, gradient and error
// term of axis w (slope and gradient stored in pass 0); the first wavefront
// then adds the axes' rhs terms and error terms in axis order, exactly as
// sweep_body's axis loop adds them, and takes the cell's decision.  Pass 1
// re-sweeps the deferred cells only (nothing when none was deferred).
template <int MODEL, int D>
__global__ void __launch_bounds__(64 * D) k_sweep_axes(const Item *__restrict__ items, const MConst *__restrict__ mcs,
                                                       const int *__restrict__ toff, int n_items, int tile_cells,
                                                       int pass) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE;
    __shared__ double SL[D - 1][NE][64];  // the limited slopes of axes 1 .. D-1
    __shared__ double SQ[D - 1][64];      // their error terms
    __shared__ uint8_t DEF[D][64];        // deferred by the axis
    const int per_tile = tile_cells >> 6;
    const int total = toff[n_items] * per_tile;
    const int lane = (int) (threadIdx.x & 63), dim = (int) (threadIdx.x >> 6);
    for (int g = (int) blockIdx.x; g < total; g += (int) gridDim.x) {
        const int gg = xcd_block(g, total);
        const int tile = gg / per_tile;
        const int i = find_item(toff, n_items, tile);
        const Item &it = items[i];
        const MConst &c = mcs[i];
        const Dev &d = it.d;
        const int64_t C = d.C;
        const int L = it.L;
        const int li = (tile - toff[i]) * tile_cells + (gg % per_tile) * 64 + lane;
        // an instance whose update produced a NaN stops (impl/integrate.hpp:50)
        bool act = li < it.nl && !it.st->nan_flag && !(pass == 1 && it.st->n_deferred == 0);
        const int32_t y = act ? it.list[li] : 0;
        if (act && pass == 1 && !(d.mark[y] & MK_DEFER)) act = false;
        if (!__syncthreads_or(act)) continue;  // uniform over the block
        double pc[NE], ctr[D], ys = 0.0, yt = 0.0, kap = 0.0;
        int ylev = 0;
        double ts[S::NGX];
#pragma unroll
        for (int e = 0; e < S::NGX; ++e) ts[e] = 0.0;
        bool deferred = false;
        double sl[NE], sq = 0.0;
        if (act) {
            for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
            for (int a = 0; a < D; ++a) ctr[a] = d.center[a * C + y];
            ys = d.size[y];
            yt = d.t[y];
            ylev = d.level[y];
            if constexpr (MODEL == MULTI) kap = d.k[S::IKAPPA * C + y];
            if constexpr (S::NG > 0)
                for (int e = 0; e < S::NG; ++e) ts[e] = d.t_slope[e * C + y];
            double pn[NE], pp[NE], qn[NE], qp[NE], xn, xp, tn, tp;
            bool hn, hp;
            sweep_side<MODEL, D, 0>(d, c, y, dim, pc, ctr, ys, yt, ylev, L, pass, pn, qn, tn, xn, hn, deferred);
            sweep_side<MODEL, D, 1>(d, c, y, dim, pc, ctr, ys, yt, ylev, L, pass, pp, qp, tp, xp, hp, deferred);
            sweep_axis_core<MODEL, D>(d, c, y, dim, pass, pc, pn, pp, qn, qp, xn, xp, tn, tp, hn, hp, ts, sl, sq);
        }
        if (dim > 0) {
#pragma unroll
            for (int e = 0; e < NE; ++e) SL[dim - 1][e][lane] = sl[e];
            SQ[dim - 1][lane] = sq;
        }
        DEF[dim][lane] = deferred;
        __syncthreads();
        if (dim == 0 && act) {
            double a_ip_in, a_in_ip, a_in_ia;
            rhs_coeffs<MODEL, D>(c, pc, kap, a_ip_in, a_in_ip, a_in_ia);
            double rhs[NE], err = 0.0;
#pragma unroll
            for (int e = 0; e < NE; ++e) rhs[e] = 0.0;
            if (pass == 0) rhs_add_dim<MODEL, D>(c, pc, sl, 0, kap, a_ip_in, a_in_ip, a_in_ia, rhs);
            err += sq;
            bool def = deferred;
#pragma unroll
            for (int a = 1; a < D; ++a) {
                double sa[NE];
#pragma unroll
                for (int e = 0; e < NE; ++e) sa[e] = SL[a - 1][e][lane];
                if (pass == 0) rhs_add_dim<MODEL, D>(c, pc, sa, a, kap, a_ip_in, a_in_ip, a_in_ia, rhs);
                err += SQ[a - 1][lane];
                def = def || DEF[a][lane];
            }
            if (pass == 0)
                for (int e = 0; e < NE; ++e) __builtin_nontemporal_store(rhs[e], &d.rhs[e * C + y]);
            sweep_decide<MODEL, D>(d, c, y, pc, ctr, ys, ylev, L, it.dt, it.counter, def, err, it.st, it.remesh,
                                   d.flags[y]);
        }
        __syncthreads();  // SL, SQ and DEF are reused by the next group
    }
}
// the interface chain (impl/integrate.hpp:53-58) is skipped for an instance
// whose update produced a NaN state
// The interface chain's ops (impl/smooth.hpp, impl/normal_curvature.hpp)
// as an axis term and an in-order sum, for k_axes; an instance whose update
// produced a NaN state skips them
// smooth_alpha (one iteration, impl/smooth.hpp:9-60): the mean into scratch
template <int MODEL, int D>
struct OpSmoothAlpha {
    static constexpr int NV = 1;
    __device__ __forceinline__ bool take(const Item &, int32_t) const { return true; }
    __device__ __forceinline__ void axis(const Item &it, const MConst &c, int32_t y, int dim, double *v) const {
        v[0] = smooth_alpha_axis<MODEL, D>(it.d, c, y, dim);
    }
    __device__ __forceinline__ void finish(const Item &it, const MConst &, int32_t, int li, const double (*w)[NV]) const {
        double acc = 0.0;
        for (int a = 0; a < D; ++a) acc += w[a][0];
        it.d.scratch[li] = acc / (2 * D);
    }
};
struct OpStoreK {
    int kk;
    __device__ __forceinline__ void operator()(const Item &it, const MConst &c, int li) const {
        if (it.st->nan_flag) return;
        store_k_body(it.d, it.list, it.nl, li, kk);
    }
};
// compute_normal (impl/normal_curvature.hpp:6-55): the unit normal
template <int MODEL, int D>
struct OpNormal {
    static constexpr int NV = 1;
    __device__ __forceinline__ bool take(const Item &, int32_t) const { return true; }
    __device__ __forceinline__ void axis(const Item &it, const MConst &c, int32_t y, int dim, double *v) const {
        v[0] = normal_axis<MODEL, D>(it.d, c, y, dim);
    }
    __device__ __forceinline__ void finish(const Item &it, const MConst &, int32_t y, int, const double (*w)[NV]) const {
        using S = Sz<MODEL, D>;
        double dd = 0;
        for (int i = 0; i < D; ++i) dd += w[i][0] * w[i][0];
        const double nrm = sqrt(dd);
        for (int i = 0; i < D; ++i) it.d.k[(S::IKN + i) * it.d.C + y] = (nrm < 1e-10) ? 0.0 : w[i][0] / nrm;
    }
};
// compute_curvature (impl/normal_curvature.hpp:58-102): minus the normal's
// divergence
template <int MODEL, int D>
struct OpCurvature {
    static constexpr int NV = 1;
    __device__ __forceinline__ bool take(const Item &, int32_t) const { return true; }
    __device__ __forceinline__ void axis(const Item &it, const MConst &c, int32_t y, int dim, double *v) const {
        v[0] = curvature_axis<MODEL, D>(it.d, c, y, dim);
    }
    __device__ __forceinline__ void finish(const Item &it, const MConst &, int32_t y, int, const double (*w)[NV]) const {
        using S = Sz<MODEL, D>;
        double kappa = 0.0;
        for (int a = 0; a < D; ++a) kappa -= w[a][0];
        it.d.k[S::IKAPPA * it.d.C + y] = kappa;
    }
};
// smooth_curvature (one iteration, impl/smooth.hpp:63-142): the weighted
// mean into scratch
template <int MODEL, int D>
struct OpSmoothCurvature {
    static constexpr int NV = 2;
    __device__ __forceinline__ bool take(const Item &, int32_t) const { return true; }
    __device__ __forceinline__ void axis(const Item &it, const MConst &c, int32_t y, int dim, double *v) const {
        smooth_curvature_axis<MODEL, D>(it.d, c, y, dim, v);
    }
    __device__ __forceinline__ void finish(const Item &it, const MConst &, int32_t, int li, const double (*w)[NV]) const {
        double ks = 0.0, wt = 0.0;
        for (int a = 0; a < D; ++a) {
            ks += w[a][0];
            wt += w[a][1];
        }
        it.d.scratch[li] = (wt > 1e-6) ? ks / wt : 0.0;
    }
};

// The slopes (and rhs) of the cells about to split, on the pre-split tree
// (adaptive_refinement.hpp:169): compute_slope per axis, the rhs from the
// axes' slopes in axis order (as compute_slope_dev)
template <int MODEL, int D>
struct OpSplitSlopesAxes {
    static constexpr int NV = Sz<MODEL, D>::NE;
    __device__ __forceinline__ bool take(const Item &it, int32_t y) const {
        return it.meshing && (it.d.mark[y] & MK_READY);
    }
    __device__ __forceinline__ void axis(const Item &it, const MConst &c, int32_t y, int dim, double *v) const {
        compute_slope_axis<MODEL, D>(it.d, c, y, dim, v);
    }
    __device__ __forceinline__ void finish(const Item &it, const MConst &c, int32_t y, int, const double (*w)[NV]) const {
        using S = Sz<MODEL, D>;
        constexpr int NE = S::NE;
        const Dev &d = it.d;
        const int64_t C = d.C;
        double pc[NE], rhs[NE];
        for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
        const double kap = MODEL == MULTI ? d.k[S::IKAPPA * C + y] : 0.0;
        compute_rhs<MODEL, D>(c, pc, w, kap, rhs);
        for (int e = 0; e < NE; ++e) d.rhs[e * C + y] = rhs[e];
    }
};

// A per-cell op whose result is a sum of one term per axis, added in axis
// order (Op::axis gives axis w's NV terms, Op::finish adds them): a cell's
// axes on D wavefronts of one block, 64 listed cells a block (as
// k_flux_axes), the first wavefront adding the terms.  Skipped for an
// instance whose update produced a NaN state.
template <int D, class Op>
__global__ void __launch_bounds__(64 * D) k_axes(const Item *__restrict__ items, const MConst *__restrict__ mcs,
                                                 const int *__restrict__ toff, int n_items, int tile_cells, Op op) {
    constexpr int NV = Op::NV;
    __shared__ double V[D - 1][NV][64];
    const int per_tile = tile_cells >> 6;
    const int total = toff[n_items] * per_tile;
    const int lane = (int) (threadIdx.x & 63), dim = (int) (threadIdx.x >> 6);
    for (int g = (int) blockIdx.x; g < total; g += (int) gridDim.x) {
        const int gg = xcd_block(g, total);
        const int tile = gg / per_tile;
        const int i = find_item(toff, n_items, tile);
        const Item &it = items[i];
        const MConst &c = mcs[i];
        const int li = (tile - toff[i]) * tile_cells + (gg % per_tile) * 64 + lane;
        const bool listed = li < it.nl && !it.st->nan_flag;
        const int32_t y = listed ? it.list[li] : 0;
        const bool act = listed && op.take(it, y);
        if (!__syncthreads_or(act)) continue;  // uniform over the block
        double v[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = 0.0;
        if (act) op.axis(it, c, y, dim, v);
        if (dim > 0) {
#pragma unroll
            for (int k = 0; k < NV; ++k) V[dim - 1][k][lane] = v[k];
        }
        __syncthreads();
        if (dim == 0 && act) {
            double w[D][NV];
#pragma unroll
            for (int k = 0; k < NV; ++k) w[0][k] = v[k];
#pragma unroll
            for (int a = 1; a < D; ++a)
#pragma unroll
                for (int k = 0; k < NV; ++k) w[a][k] = V[a - 1][k][lane];
            op.finish(it, c, y, li, w);
        }
        __syncthreads();  // V is reused by the next group
    }
}

struct OpApply {
    __device__ __forceinline__ void operator()(const Item &it, const MConst &c, int li) const {
        if (it.st->nan_flag) return;  // an instance with a NaN state stops before the LTS apply
        apply_body(it.d, it.list, it.nl, li, it.L);
    }
};

// exclusive scan of one value per thread over a block of 1024 threads
__device__ __forceinline__ int block_exclusive_scan_1024(int v, int *lds, int &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    if (threadIdx.x < 16) {
        int a = lds[threadIdx.x];
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(a, o, 16);
            if ((int) threadIdx.x >= o) a += y;
        }
        lds[16 + threadIdx.x] = a;
    }
    __syncthreads();
    total = lds[31];
    const int r = (w ? lds[16 + w - 1] : 0) + x - v;
    __syncthreads();
    return r;
}

// Cluster-L lists of the batch, in node order (the order of the reference's
// acti_clusters after a rebuild is irrelevant to the results; node order keeps
// the sweeps coalesced): per pool tile a count, one scan per item, a scatter.
// a node's flags, tree level, ACTI level and cluster in one word
__device__ __forceinline__ uint32_t node_meta(const Dev &d, int64_t i) {
    return (uint32_t) d.flags[i] | (uint32_t) (uint8_t) d.level[i] << 8 | (uint32_t) (uint8_t) d.acti[i] << 16 |
           (uint32_t) (uint8_t) d.cluster[i] << 24;
}

// META: also store every node's metadata word (d.meta) for the window flux op
// A pool tile (POOL_TILE = 2048 consecutive node ids) is taken by a block
// of 256 threads, 8 consecutive ids per thread: one 8-byte load per
// per-node byte array instead of eight 1-byte loads.  Ids at and beyond the
// pool size count as absent.
static_assert(POOL_TILE == 256 * 8, "a pool tile is 8 ids per thread of a 256-thread block");

// the 8 bytes of a per-node byte array at ids base8 .. base8 + 7 (base8 a
// multiple of 8); bytes at ids >= n read as `fill`
__device__ __forceinline__ uint64_t load8(const GMEM int8_t *a, int64_t base8, int64_t n, uint8_t fill) {
    if (base8 + 8 <= n) return *(const GMEM uint64_t *) (a + base8);
    uint64_t v = 0;
    for (int j = 0; j < 8; ++j) v |= (uint64_t) (base8 + j < n ? (uint8_t) a[base8 + j] : fill) << (8 * j);
    return v;
}
__device__ __forceinline__ uint64_t load8(const GMEM uint8_t *a, int64_t base8, int64_t n, uint8_t fill) {
    return load8((const GMEM int8_t *) a, base8, n, fill);
}

// bit j set where byte j of v equals b
__device__ __forceinline__ unsigned bytes_eq(uint64_t v, uint8_t b) {
    unsigned m = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) m |= (unsigned) (((v >> (8 * j)) & 0xff) == b) << j;
    return m;
}

// the ids of a thread's 8 that belong to the cluster list of level L (id 0,
// the null node, never does)
__device__ __forceinline__ unsigned tile_members(const Dev &d, int64_t base8, int64_t n, int L) {
    unsigned m = bytes_eq(load8(d.cluster, base8, n, 0xff), (uint8_t) L);
    if (base8 == 0) m &= ~1u;
    return m;
}

// META: also store every node's metadata word (d.meta) for the window flux op
template <int OUT, bool META = false>
__global__ void __launch_bounds__(256) k_sel_count(const Item *__restrict__ items, const int *__restrict__ toff,
                                                   int n_items, int *tilecnt) {
    __shared__ int wc[4];
    const int total = toff[n_items];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int tile = xcd_block(t, total);
        const int i = find_item(toff, n_items, tile);
        const Item &it = items[i];
        int cnt = 0;
        if (OUT == 0 ? it.nl > 0 : it.meshing != 0) {
            const int64_t base8 = (int64_t) (tile - toff[i]) * POOL_TILE + 8 * (int64_t) threadIdx.x;
            cnt = __popc(tile_members(it.d, base8, it.n, it.L));
            if (META && base8 < it.n) {
                const uint64_t fl = load8(it.d.flags, base8, it.n, 0), lv = load8(it.d.level, base8, it.n, 0),
                               ac = load8(it.d.acti, base8, it.n, 0), cl = load8(it.d.cluster, base8, it.n, 0);
                uint32_t w[8];
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    w[j] = (uint32_t) ((fl >> (8 * j)) & 0xff) | (uint32_t) ((lv >> (8 * j)) & 0xff) << 8 |
                           (uint32_t) ((ac >> (8 * j)) & 0xff) << 16 | (uint32_t) ((cl >> (8 * j)) & 0xff) << 24;
                if (base8 + 8 <= it.n) {
                    *(GMEM uint4 *) (it.d.meta + base8) = make_uint4(w[0], w[1], w[2], w[3]);
                    *(GMEM uint4 *) (it.d.meta + base8 + 4) = make_uint4(w[4], w[5], w[6], w[7]);
                } else {
                    for (int j = 0; j < 8; ++j)
                        if (base8 + j < it.n) it.d.meta[base8 + j] = w[j];
                }
            }
        }
        // block sum: wave sums, then the first thread adds the four
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
        if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) tilecnt[tile] = wc[0] + wc[1] + wc[2] + wc[3];
        __syncthreads();
    }
}

// one block of 1024 threads: tile offsets within each item and the item's
// list size (OUT 0: the substep's list, with a cleared status; OUT 1: the
// list rebuilt after a remesh; OUT 2: offsets only)
template <int OUT>
__global__ void __launch_bounds__(1024) k_sel_scan(const Item *__restrict__ items, const int *__restrict__ toff,
                                                   int n_items, int *tilecnt) {
    constexpr int K = 16;  // counts per thread in one pass
    __shared__ int lds[32];
    for (int i = 0; i < n_items; ++i) {
        const int a = toff[i], b = toff[i + 1];
        int running = 0;
        if (b - a <= 1024 * K) {
            // one pass: each thread loads its run of consecutive counts at
            // once, one block scan of the run sums
            const int per = (b - a + 1023) / 1024;
            const int s0 = a + (int) threadIdx.x * per;
            int v[K], sum = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                v[k] = k < per && s0 + k < b ? tilecnt[s0 + k] : 0;
                sum += v[k];
            }
            int tot = 0;
            int r = block_exclusive_scan_1024(sum, lds, tot);
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k < per && s0 + k < b) {
                    tilecnt[s0 + k] = r;
                    r += v[k];
                }
            running = tot;
        } else {
            for (int s = a; s < b; s += 1024) {
                const int idx = s + threadIdx.x;
                const int v = idx < b ? tilecnt[idx] : 0;
                int tot = 0;
                const int ex = block_exclusive_scan_1024(v, lds, tot);
                if (idx < b) tilecnt[idx] = running + ex;
                running += tot;
            }
        }
        if (threadIdx.x == 0) {
            Status *st = items[i].st;
            if (OUT == 0) {
                *st = Status{};
                st->n_list = running;
            } else if (OUT == 1) {
                st->n_list2 = running;
            }
        }
        // the level histogram the substep's tail gathers starts at zero
        if (OUT == 0 && threadIdx.x < 16) items[i].hist[threadIdx.x] = 0;
    }
}

template <int OUT>
__global__ void __launch_bounds__(256) k_sel_scatter(const Item *__restrict__ items, const int *__restrict__ toff,
                                                     int n_items, const int *__restrict__ tileoff) {
    __shared__ int wc[4];
    const int total = toff[n_items];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int tile = xcd_block(t, total);
        const int i = find_item(toff, n_items, tile);
        const Item &it = items[i];
        if (OUT == 0 ? it.nl == 0 : it.meshing == 0) continue;  // uniform over the block
        int32_t *out = OUT == 0 ? it.list : it.rest;
        const int64_t base8 = (int64_t) (tile - toff[i]) * POOL_TILE + 8 * (int64_t) threadIdx.x;
        const unsigned m = tile_members(it.d, base8, it.n, it.L);
        const int c = __popc(m);
        // inclusive scan of c over the wave, then over the waves: ids keep
        // their order (thread t holds ids 8t .. 8t + 7 of the tile)
        int x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wc[w] = x;
        __syncthreads();
        int pos = tileoff[tile] + x - c;
        for (int q = 0; q < w; ++q) pos += wc[q];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if ((m >> j) & 1) out[pos++] = (int32_t) (base8 + j);
        __syncthreads();
    }
}

// flush_acti_buffer of the items that have a pending buffer, then the level
// histogram of every item in the pool table
static __global__ void __launch_bounds__(256) k_pool_flush(const Item *__restrict__ items, const int *__restrict__ toff,
                                                    int n_items) {
    const int total = toff[n_items];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int tile = xcd_block(t, total);
        const int i = find_item(toff, n_items, tile);
        const Item &it = items[i];
        // an instance whose state turned NaN stops before the flush
        if (!it.flush || it.st->nan_flag) continue;
        // 8 ids per thread; only ids whose buffer holds level L change
        const int64_t base8 = (int64_t) (tile - toff[i]) * POOL_TILE + 8 * (int64_t) threadIdx.x;
        const unsigned m = bytes_eq(load8(it.d.buffer, base8, it.n, 0xff), (uint8_t) it.L);
        if (m)
            for (int j = 0; j < 8; ++j)
                if ((m >> j) & 1) flush_node(it.d, base8 + j, it.n, it.L);
    }
}

// flag listed leaves for splitting at the next remesh (the split keeps the
// reference's conditions: an active leaf that may still split)
// ------------------------------------------------- refinement predictor
// Number of predictor features of a model: the leaf's primitives, the
// differences to the primitives at the 3^D - 1 same-size neighbour positions,
// the gradient (D x NG), the tree level above the base level and log10 of
// the error estimate.
template <int MODEL, int D>
__host__ __device__ constexpr int n_features() {
    using S = Sz<MODEL, D>;
    return S::NE * (D == 2 ? 9 : 27) + D * S::NG + 2;
}

// the leaf containing point x (inside the domain): descent from the root
// choosing each child by comparing x with the node's centre (a point on a
// face goes to the negative side), so that ties do not depend on rounding
// of scaled coordinates
template <int D>
__device__ inline int32_t leaf_at(const Dev &d, const double *x) {
    int32_t idx = 1;
    for (int guard = 0; guard < 64 && !(d.flags[idx] & F_LEAF); ++guard) {
        int q = 0;
        for (int a = 0; a < D; ++a)
            if (x[a] > d.center[a * d.C + idx]) q |= 1 << a;
        idx = d.child_base[idx] + q;
    }
    return idx;
}

static __global__ void k_active_flags(Dev d, int64_t n, uint8_t *__restrict__ fl) {
    const int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t f = d.flags[i];
    fl[i] = (i > 0 && (f & F_LEAF) && (f & F_ACTIVE) && !(f & F_GARBAGE)) ? 1 : 0;
}

static __global__ void k_active_ids(int64_t n, const uint8_t *__restrict__ fl, const int32_t *__restrict__ off,
                             int32_t *__restrict__ ids, int64_t cap) {
    const int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (i >= n || !fl[i] || off[i] >= cap) return;
    ids[off[i]] = (int32_t) i;
}

// features of the listed leaves (one thread per leaf), standardised as
// (x - mu) / sd in FP32 when mu is given.  A neighbour position is the
// centre moved by one cell width; outside a non-periodic domain the leaf's
// own state stands in (difference 0), a periodic axis wraps.
template <int MODEL, int D>
__global__ void k_features(Dev d, MConst c, const int32_t *__restrict__ ids, int64_t n, int min_level,
                           const float *__restrict__ mu, const float *__restrict__ sd, float *__restrict__ out,
                           double *__restrict__ centers, int8_t *__restrict__ levels) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE, NG = S::NG, NGX = S::NGX, NOFF = D == 2 ? 9 : 27, NFEAT = n_features<MODEL, D>();
    const int64_t k = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t C = d.C;
    const int32_t y = ids[k];
    float *o = out + k * NFEAT;
    int col = 0;
    auto put = [&](double v) {
        float f = (float) v;
        if (mu) f = (f - mu[col]) / sd[col];
        o[col++] = f;
    };
    double pc[NE], ctr[D];
    for (int e = 0; e < NE; ++e) pc[e] = d.p[e * C + y];
    for (int a = 0; a < D; ++a) ctr[a] = d.center[a * C + y];
    if (centers)
        for (int a = 0; a < D; ++a) centers[k * D + a] = ctr[a];
    if (levels) levels[k] = d.level[y];
    const double sz = d.size[y];
    for (int e = 0; e < NE; ++e) put(pc[e]);
    for (int m = 0; m < NOFF; ++m) {
        int off[D], r = m;
        bool zero = true;
        for (int a = D - 1; a >= 0; --a) {
            off[a] = r % 3 - 1;
            r /= 3;
            zero = zero && off[a] == 0;
        }
        if (zero) continue;
        double x[D];
        bool inside = true;
        for (int a = 0; a < D; ++a) {
            x[a] = ctr[a] + sz * (double) off[a];
            const double h = 0.5 * c.active_size[a];
            if (!c.periodic[a]) inside = inside && fabs(x[a]) < h;
            else if (x[a] < -h) x[a] += c.active_size[a];
            else if (x[a] > h) x[a] -= c.active_size[a];
        }
        const int32_t nb = inside ? leaf_at<D>(d, x) : 0;
        for (int e = 0; e < NE; ++e) put(nb > 0 ? d.p[e * C + nb] - pc[e] : 0.0);
    }
    for (int a = 0; a < D; ++a)
        for (int e = 0; e < NG; ++e) put(d.grad[(a * NGX + e) * C + y]);
    put((double) (d.level[y] - min_level));
    put(log10(fmax(d.error[y], 1e-12)));
}

// tree level of the leaf containing each point (-1 outside the root cell)
template <int D>
__global__ void k_locate_levels(Dev d, MConst c, const double *__restrict__ pts, int64_t n, int8_t *__restrict__ lev) {
    const int64_t k = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (k >= n) return;
    double x[D];
    bool inside = true;
    for (int a = 0; a < D; ++a) {
        x[a] = pts[k * D + a];
        inside = inside && fabs(x[a]) <= 0.5 * d.size[1];
    }
    lev[k] = inside ? d.level[leaf_at<D>(d, x)] : (int8_t) -1;
}

// leaves whose logits pick class 2 (refine): rn[y] = 1
static __global__ void k_pred_seed(const float *__restrict__ logits, const int32_t *__restrict__ ids, int64_t n,
                            uint8_t *__restrict__ rn) {
    const int64_t k = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float *l = logits + 3 * k;
    if (l[2] > l[0] && l[2] > l[1]) rn[ids[k]] = 1;
}

// one ring of face neighbours: active leaves next to a cell flagged r get r + 1
template <int D>
__global__ void k_pred_ring(Dev d, int64_t n, uint8_t *__restrict__ rn, uint8_t r) {
    const int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (i >= n || rn[i] != r) return;
    for (int f = 0; f < 2 * D; ++f) {
        const int32_t nb = d.nbr[f * d.C + i];
        if (nb <= 0) continue;
        const uint8_t fl = d.flags[nb];
        if ((fl & F_LEAF) && (fl & F_ACTIVE) && !(fl & F_GARBAGE) && rn[nb] == 0) rn[nb] = (uint8_t) (r + 1);
    }
}

// flag the leaves with rn > 0 to split at the next remesh (k_mark_split's rule)
static __global__ void k_pred_mark(Dev d, int64_t n, const uint8_t *__restrict__ rn, unsigned long long *__restrict__ count) {
    const int64_t y = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (y >= n || y < 1 || !rn[y]) return;
    const uint8_t fl = d.flags[y];
    if ((fl & F_LEAF) && (fl & F_ACTIVE) && !(fl & F_GARBAGE) && (fl & F_CAN_SPLIT)) {
        d.flags[y] = fl | F_SHOULD_SPLIT;
        atomicAdd(count, 1ull);
    }
}

static __global__ void k_mark_split(Dev d, const int32_t *__restrict__ ids, int64_t n, int64_t n_nodes) {
    const int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t y = ids[i];
    if (y < 1 || y >= n_nodes) return;
    const uint8_t fl = d.flags[y];
    if ((fl & F_LEAF) && (fl & F_ACTIVE) && !(fl & F_GARBAGE) && (fl & F_CAN_SPLIT)) d.flags[y] = fl | F_SHOULD_SPLIT;
}

static __global__ void k_zero_hist(const Item *__restrict__ items, int n_items) {
    const int i = blockIdx.x;
    if (i < n_items && threadIdx.x < 16) items[i].hist[threadIdx.x] = 0;
}

// the number of bytes of v equal to b (exact zero-byte test on v ^ b..b)
__device__ __forceinline__ unsigned count_bytes_eq(uint64_t v, uint8_t b) {
    const uint64_t x = v ^ (0x0101010101010101ull * b);
    const uint64_t t = ~(((x & 0x7f7f7f7f7f7f7f7full) + 0x7f7f7f7f7f7f7f7full) | x | 0x7f7f7f7f7f7f7f7full);
    return (unsigned) __popcll(t);
}

// Per-level counts of the listed cells' clusters over the pool: a streaming
// pass.  Each thread keeps its own 16 level counts over the tiles its block
// takes (8 ids per tile, four tiles' loads in flight); the wave adds them up
// and the block's counts go to the item's global counters when the item
// changes and at the end (few global atomics).
static __global__ void __launch_bounds__(256) k_pool_hist(const Item *__restrict__ items, const int *__restrict__ toff,
                                                          int n_items) {
    constexpr int U = 4;
    __shared__ unsigned int h[16];
    const int total = toff[n_items];
    if ((int) blockIdx.x >= total) return;  // no tile for this block
    const int lane = (int) threadIdx.x & 63;
    int cur = -1;
    unsigned cnt[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) cnt[l] = 0;
    if (threadIdx.x < 16) h[threadIdx.x] = 0;
    __syncthreads();
    // the thread counts into the block's, the block's into item cur's
    auto flush = [&]() {
        // the levels counted in the wave, then one sum per level
        unsigned pm = 0;
#pragma unroll
        for (int l = 0; l < 16; ++l) pm |= (unsigned) (cnt[l] != 0) << l;
        for (int o = 32; o > 0; o >>= 1) pm |= (unsigned) __shfl_xor((int) pm, o);
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            if ((pm >> l) & 1) {
                unsigned c = cnt[l];
                for (int o = 32; o > 0; o >>= 1) c += (unsigned) __shfl_xor((int) c, o);
                if (lane == 0) atomicAdd(&h[l], c);
            }
            cnt[l] = 0;
        }
        __syncthreads();
        if (cur >= 0 && threadIdx.x < 16 && h[threadIdx.x])
            atomicAdd(&items[cur].hist[threadIdx.x], (unsigned long long) h[threadIdx.x]);
        __syncthreads();
        if (threadIdx.x < 16) h[threadIdx.x] = 0;
        __syncthreads();
    };
    for (int t0 = (int) blockIdx.x; t0 < total; t0 += U * (int) gridDim.x) {
        uint64_t v[U];
        int ti[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int t = t0 + u * (int) gridDim.x;
            ti[u] = -1;
            v[u] = ~0ull;
            if (t < total) {
                const int tile = xcd_block(t, total);
                const int i = find_item(toff, n_items, tile);
                const Item &it = items[i];
                const int64_t base8 = (int64_t) (tile - toff[i]) * POOL_TILE + 8 * (int64_t) threadIdx.x;
                // clusters of the thread's 8 ids (absent ids and id 0 as -1)
                v[u] = load8(it.d.cluster, base8, it.n, 0xff);
                if (base8 == 0) v[u] |= 0xff;
                ti[u] = i;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (ti[u] < 0) break;
            if (ti[u] != cur) {
                if (cur >= 0) flush();
                cur = ti[u];
            }
#pragma unroll
            for (int l = 0; l < 16; ++l) cnt[l] += count_bytes_eq(v[u], (uint8_t) l);
        }
    }
    flush();
}

// ------------------------------------------------- batched remesh phases
// (impl/adaptive_refinement.hpp:97-322) for the instances with marks
// (Item::meshing); the others return at once.
// Launched with the substep's head for every instance that has marks (the
// sweep counted them), so the remesh needs no wait of its own; run again,
// after clearing its counts, when deferred cells received marks later.
// AGAIN: the repeat after the deferred sweep, which the substeps planned
// on the device launch unconditionally (it acts only where cells were
// deferred)
template <int D, bool AGAIN = false>
struct OpMark2 {
    __device__ __forceinline__ void operator()(const Item &it, const MConst &, int li) const {
        if (it.st->n_marked == 0 || it.st->nan_flag) return;
        if (AGAIN && it.st->n_deferred == 0) return;
        mark2_body<D>(it.d, it.list, it.nl, li, it.L, it.rn, it.st);
    }
};

template <bool AGAIN = false>
__global__ void k_clear_split_counts(const Item *__restrict__ items, int n_items) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    if (AGAIN && items[i].st->n_deferred == 0) return;
    items[i].st->n_ready_split = 0;
    items[i].st->n_new_blocks = 0;
}

// The remesh plan of a substep, on the device (what the host computes
// between the head and the remesh otherwise): which instances remesh, their
// pool size before and after the splits (the host guaranteed the capacity),
// and the tile tables of the remesh phases (toffr: list tiles of 256,
// toff2: the rebuilt lists in tiles of `block`) and of the grown pools
// (toffp).  One thread; instances in order.
static __global__ void k_remesh_plan(Item *__restrict__ items, int n_items, int nch, int block,
                                     int *__restrict__ toffr, int *__restrict__ toff2, int *__restrict__ toffp) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int tr = 0, t2 = 0, tp = 0;
    for (int i = 0; i < n_items; ++i) {
        Item &it = items[i];
        Status *st = it.st;
        toffr[i] = tr;
        toff2[i] = t2;
        toffp[i] = tp;
        const bool nan = st->nan_flag != 0;
        const bool mesh = it.nl > 0 && st->n_marked > 0 && !nan;
        it.meshing = mesh ? 1 : 0;
        if (mesh) {
            tr += (it.nl + 255) / 256;
            it.n_before = it.n;
            it.n = it.n + (int64_t) nch * st->n_new_blocks;
            t2 += (int) ((2 * (int64_t) it.nl + (int64_t) nch * st->n_ready_split + block - 1) / block);
        }
        st->n_pool = it.n;
        // an instance that broke down takes no part in the tail's pool passes
        if ((it.nl > 0 || it.flush) && !nan) tp += (int) ((it.n + POOL_TILE - 1) / POOL_TILE);
    }
    toffr[n_items] = tr;
    toff2[n_items] = t2;
    toffp[n_items] = tp;
}

// exclusive scan of the new-block flags over each item's list: per list
// tile a count (tiles of 256 entries), one scan per item, then the offsets
static __global__ void __launch_bounds__(256) k_rn_count(const Item *__restrict__ items, const int *__restrict__ toff,
                                                  int n_items, int *tilecnt) {
    __shared__ int wc[4];
    const int total = toff[n_items];
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = find_item(toff, n_items, t);
        const Item &it = items[i];
        const int li = (t - toff[i]) * 256 + (int) threadIdx.x;
        const bool pr = it.meshing && li < it.nl && it.rn[li];
        const int c = __popcll(__ballot(pr));
        if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = c;
        __syncthreads();
        if (threadIdx.x == 0) tilecnt[t] = wc[0] + wc[1] + wc[2] + wc[3];
        __syncthreads();
    }
}

static __global__ void __launch_bounds__(256) k_rn_scatter(const Item *__restrict__ items, const int *__restrict__ toff,
                                                    int n_items, const int *__restrict__ tileoff) {
    __shared__ int wc[4];
    const int total = toff[n_items];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int i = find_item(toff, n_items, t);
        const Item &it = items[i];
        const int li = (t - toff[i]) * 256 + (int) threadIdx.x;
        const bool pr = it.meshing && li < it.nl && it.rn[li];
        const uint64_t m = __ballot(pr);
        if (lane == 0) wc[w] = __popcll(m);
        __syncthreads();
        int before = 0;
        for (int q = 0; q < w; ++q) before += wc[q];
        if (it.meshing && li < it.nl) it.off[li] = tileoff[t] + before + __popcll(m & lt);
        __syncthreads();
    }
}

template <int MODEL, int D>
struct OpSplit {
    __device__ __forceinline__ void operator()(const Item &it, const MConst &c, int li) const {
        if (!it.meshing) return;
        split_body<MODEL, D>(it.d, c, it.list, it.nl, li, it.off, it.n_before, it.L);
    }
};
template <int MODEL, int D>
struct OpUnsplit {
    __device__ __forceinline__ void operator()(const Item &it, const MConst &c, int li) const {
        if (!it.meshing) return;
        unsplit_body<MODEL, D>(it.d, c, it.list, it.nl, li, it.L, it.st);
    }
};
// LTS decision of the cells of the rebuilt list (size on the device)
template <int MODEL, int D>
struct OpReassign {
    __device__ __forceinline__ void operator()(const Item &it, const MConst &c, int li) const {
        if (!it.meshing) return;
        reassign_body<MODEL, D>(it.d, c, it.rest, it.st->n_list2, li, it.L, it.dt, it.counter, it.st);
    }
};

// The neighbour rebuild over the pool: one thread per (node, face), so a
// thread's chain is one face search; POOL_TILE / 256 * 2D blocks a pool tile.
template <int D>
__global__ void __launch_bounds__(256) k_pool_neighbors(const Item *__restrict__ items,
                                                        const MConst *__restrict__ mcs, const int *__restrict__ toff,
                                                        int n_items) {
    constexpr int PARTS = POOL_TILE / 256 * 2 * D;
    const int total = toff[n_items] * PARTS;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int gg = xcd_block(t, total);
        const int tile = gg / PARTS, part = gg % PARTS;
        const int i = find_item(toff, n_items, tile);
        const Item &it = items[i];
        // only where the topology changed (a split or an unsplit happened)
        if (!it.meshing || (it.st->n_ready_split == 0 && it.st->n_unsplit == 0)) continue;
        const int64_t node = (int64_t) (tile - toff[i]) * POOL_TILE + (part / (2 * D)) * 256 + threadIdx.x;
        neighbors_face<D>(it.d, mcs[i], node, part % (2 * D), it.n);
    }
}

}  // namespace dacti
