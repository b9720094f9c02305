// model.hpp -- device-side physics of the leaf update (FP64 throughout).
//
// Restates src/lib/model/{euler,navier_stokes,multi,common}.hpp and
// src/lib/integrator/{slope_limiters,boundary}.hpp.  Expressions keep the
// reference's evaluation order and the library is built with
// -ffp-contract=off, so every multiply and add is rounded separately, as in
// the reference's plain C++.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Pool arrays are global-memory pointers (address space 1): loads through
// them are global_load instructions, not flat ones (a flat load also counts
// against the scalar-load counter, so every wait for a scalar load would
// drain the vector loads in flight).
#if defined(__HIP_DEVICE_COMPILE__)
#define GMEM __attribute__((address_space(1)))
#else
#define GMEM
#endif

namespace dacti {

enum { EULER = 0, NAVIER_STOKES = 1, MULTI = 2 };

// Sizes and variable indices of a model
// (euler.hpp:18-47, navier_stokes.hpp:18-39, multi.hpp:18-48).
template <int MODEL, int D>
struct Sz {
    static constexpr int NE = MODEL == MULTI ? D + 4 : D + 2;
    static constexpr int NF = MODEL == MULTI ? NE + D + 1 : NE;
    static constexpr int NP = MODEL == MULTI ? D + 4 : 0;
    static constexpr int NG = MODEL == EULER ? 0 : (MODEL == NAVIER_STOKES ? D + 1 : D + 3);
    static constexpr int IU = MODEL == MULTI ? 2 : 1;
    static constexpr int IP = MODEL == MULTI ? NE - 2 : NE - 1;
    static constexpr int IA = NE - 1;  // multi only
    static constexpr int IKN = 2, IKAPPA = 2 + D, IALPHAS = 3 + D;
    static constexpr int IDT = D, IDA = D + 1;
    static constexpr int NPX = NP > 0 ? NP : 1;
    static constexpr int NGX = NG > 0 ? NG : 1;
    static constexpr int NCH = 1 << D;
    static constexpr int NFACE = 2 * D;
};

// Per-handle constants (the reference keeps the model ones process-global).
struct MConst {
    double gamma, ns_mu, ns_lambda, ns_cp, ns_R;
    double sigma, gamma1, gamma2, P1, P2, cp1, cp2, cv1, cv2, mu1, mu2, lambda1, lambda2;
    double max_cfl, min_error, max_error;
    double weights[8];
    double active_size[3];
    int periodic[3];
    int limiter, max_refinement_level, n_zones;
    int bc_type[6];
    double bc_p[6][8];
    double bc_fade_in[6];
    double zone_x0[8][3], zone_x1[8][3];
    int zone_level[8];
    double vol[20], area[20];  // pow(size, D), pow(size, D-1) per tree level
};

enum { BC_ZERO_GRADIENT = 0, BC_FREESLIP, BC_NOSLIP, BC_INFLOW, BC_INFLOW_R, BC_OUTFLOW, BC_DIRICHLET };
// bc code of a face whose condition is its node's object record
enum : uint8_t { BC_OBJECT = 7 };

__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }  // std::max

// Division by an exact power of two: x / 2^k and x * 2^-k are the same
// correctly rounded value (also for zeros, infinities and subnormal
// results), so where a divisor is a power of two -- every cell size and
// volume of a domain whose root cell has a power-of-two side -- a product
// with the exact reciprocal replaces the division without changing a bit.
// Returns whether q is such a divisor (normal, with a normal reciprocal)
// and sets r = 1 / q.
__device__ __forceinline__ bool pow2_recip(double q, double &r) {
    const unsigned long long b = (unsigned long long) __double_as_longlong(q);
    const unsigned long long e = (b >> 52) & 0x7ffull;
    r = __longlong_as_double((long long) ((b & 0x8000000000000000ull) | ((2046ull - e) << 52)));
    return (b & 0xfffffffffffffull) == 0 && e >= 2 && e <= 2045;
}

// ---------------------------------------------------------------- EOS
// multi.hpp:367-387
__device__ inline double mgamma(const MConst &c, double a) {
    if (a < 1e-20) return c.gamma2;
    else if (a > 1 - 1e-20) return c.gamma1;
    return 1 + 1 / (a / (c.gamma1 - 1) + (1 - a) / (c.gamma2 - 1));
}
__device__ inline double mp0(const MConst &c, double a) {
    if (a < 1e-20) return c.P2;
    else if (a > 1 - 1e-20) return c.P1;
    return (a * c.gamma1 * c.P1 / (c.gamma1 - 1) + (1 - a) * c.gamma2 * c.P2 / (c.gamma2 - 1)) /
           (1 + a / (c.gamma1 - 1) + (1 - a) / (c.gamma2 - 1));
}
// multi.hpp:394-401
__device__ inline double mtemperature(const MConst &c, double r0, double r1, double p, double a) {
    if (a < 1e-20) return (p + c.P2) / ((c.gamma2 - 1) * r1 * c.cv2);
    else if (a > 1 - 1e-20) return (p + c.P1) / ((c.gamma1 - 1) * r0 * c.cv1);
    return (p / (mgamma(c, a) - 1) + a * c.P1 / (c.gamma1 - 1) + (1 - a) * c.P2 / (c.gamma2 - 1)) /
           (r0 * c.cv1 + r1 * c.cv2);
}
__device__ inline double msound_speed_2(const MConst &c, double r0, double r1, double p, double a) {
    return mgamma(c, a) * (p + mp0(c, a)) / (r0 + r1);
}

template <int MODEL, int D>
__device__ inline double sound_speed_prim(const MConst &c, const double *p) {
    using S = Sz<MODEL, D>;
    if constexpr (MODEL == MULTI)
        return sqrt(mgamma(c, p[S::IA]) * (p[S::IP] + mp0(c, p[S::IA])) / (p[0] + p[1]));
    else
        return sqrt(c.gamma * p[S::IP] / p[0]);
}

// euler.hpp:197-203; multi.hpp:432-438 (density argument is {p[0], 0})
template <int MODEL, int D>
__device__ inline bool supersonic(const MConst &c, const double *p) {
    using S = Sz<MODEL, D>;
    double u2 = 0;
    for (int i = 0; i < D; ++i) u2 += p[S::IU + i] * p[S::IU + i];
    if constexpr (MODEL == MULTI) return u2 > msound_speed_2(c, p[0], 0.0, p[S::IP], p[S::IA]);
    else return u2 > c.gamma * p[S::IP] / p[0];
}

// primitive_variables / conserved_variables
template <int MODEL, int D>
__device__ inline void prim_from_cons(const MConst &c, const double *u, double *p) {
    using S = Sz<MODEL, D>;
    double ke = 0;
    if constexpr (MODEL == MULTI) {
        p[0] = u[0];
        p[1] = u[1];
        for (int i = 0; i < D; ++i) p[S::IU + i] = u[S::IU + i] / (u[0] + u[1]);
        for (int i = 0; i < D; ++i) ke += 0.5 * u[S::IU + i] * u[S::IU + i];
        p[S::IP] = (mgamma(c, u[S::IA]) - 1) * (u[S::IP] - ke / (u[0] + u[1])) - mgamma(c, u[S::IA]) * mp0(c, u[S::IA]);
        p[S::IA] = u[S::IA];
    } else {
        p[0] = u[0];
        for (int i = 0; i < D; ++i) p[S::IU + i] = u[S::IU + i] / u[0];
        for (int i = 0; i < D; ++i) ke += 0.5 * u[S::IU + i] * u[S::IU + i];
        p[S::IP] = (c.gamma - 1) * (u[S::IP] - ke / u[0]);
    }
}

template <int MODEL, int D>
__device__ inline void cons_from_prim(const MConst &c, const double *p, double *u) {
    using S = Sz<MODEL, D>;
    double ke = 0;
    if constexpr (MODEL == MULTI) {
        u[0] = p[0];
        u[1] = p[1];
        for (int i = 0; i < D; ++i) {
            u[S::IU + i] = (p[0] + p[1]) * p[S::IU + i];
            ke += 0.5 * p[S::IU + i] * p[S::IU + i];
        }
        u[S::IP] = (p[S::IP] + mgamma(c, p[S::IA]) * mp0(c, p[S::IA])) / (mgamma(c, p[S::IA]) - 1) + (p[0] + p[1]) * ke;
        u[S::IA] = p[S::IA];
    } else {
        u[0] = p[0];
        for (int i = 0; i < D; ++i) {
            u[S::IU + i] = p[0] * p[S::IU + i];
            ke += 0.5 * p[S::IU + i] * p[S::IU + i];
        }
        u[S::IP] = p[S::IP] / (c.gamma - 1) + p[0] * ke;
    }
}

// physical flux F(p, dim) (projected variant when PROJ): euler.hpp:50-83, multi.hpp:78-123
template <int MODEL, int D, bool PROJ>
__device__ inline void flux_F(const MConst &c, const double *p, double up, int dim, double *F) {
    using S = Sz<MODEL, D>;
    const int in = S::IU + dim;
    const double un = PROJ ? up : p[in];
    double ke = 0.0;
    for (int i = 0; i < S::NF; ++i) F[i] = 0.0;
    if constexpr (MODEL == MULTI) {
        for (int i = 0; i < D; ++i) {
            F[S::IU + i] = (p[0] + p[1]) * un * p[S::IU + i];
            ke += 0.5 * p[S::IU + i] * p[S::IU + i];
        }
        F[0] = un * p[0];
        F[1] = un * p[1];
        F[in] += p[S::IP];
        const double g = mgamma(c, p[S::IA]);
        F[S::IP] = un * (g / (g - 1) * (p[S::IP] + mp0(c, p[S::IA])) + ke * (p[0] + p[1]));
        F[S::IA] = un * p[S::IA];
        F[S::IA + 1] = un;
        F[S::IA + 2 + dim] = p[S::IA];
    } else {
        for (int i = 0; i < D; ++i) {
            F[S::IU + i] = PROJ ? p[0] * un * p[S::IU + i] : p[0] * p[in] * p[S::IU + i];
            ke += 0.5 * p[S::IU + i] * p[S::IU + i];
        }
        F[0] = un * p[0];
        F[in] += p[S::IP];
        F[S::IP] = un * (c.gamma / (c.gamma - 1) * p[S::IP] + ke * p[0]);
    }
}


// a[k] for a runtime k < N, as a chain of selects (an indexed private array
// would be placed in scratch memory); pick_reg below also keeps the chain
// from being folded back into an indexed load
template <int N>
__device__ __forceinline__ double pick(const double *a, int k) {
    double v = a[0];
#pragma unroll
    for (int q = 1; q < N; ++q) v = k == q ? a[q] : v;
    return v;
}

// pick for an array the caller keeps in registers: the values pass through
// an empty asm first, so the select chain cannot be folded back into an
// indexed load (which would put the whole array in private memory)
template <int N>
__device__ __forceinline__ double pick_reg(const double *a, int k) {
    double v[N];
#pragma unroll
    for (int q = 0; q < N; ++q) {
        v[q] = a[q];
        asm volatile("" : "+v"(v[q]));
    }
    double r = v[0];
#pragma unroll
    for (int q = 1; q < N; ++q) r = k == q ? v[q] : r;
    return r;
}

// rhs = sum_j A_j(p) slope_j (Hancock time derivative): euler.hpp:86-116,
// multi.hpp:126-169.  Only the structurally non-zero Jacobian entries are
// summed, in the reference's (i, j, k) order; the skipped products are
// zero and leave the running sum unchanged.  rhs_add_dim adds the terms of
// one dimension j to every row, so a caller can stream the slopes.
template <int MODEL, int D>
__device__ inline void rhs_coeffs(const MConst &c, const double *p, double kappa, double &a_ip_in, double &a_in_ip,
                                  double &a_in_ia) {
    using S = Sz<MODEL, D>;
    a_in_ia = 0.0;
    if constexpr (MODEL == MULTI) {
        a_ip_in = (p[0] + p[1]) * msound_speed_2(c, p[0], p[1], p[S::IP], p[S::IA]);
        a_in_ip = 1.0 / (p[0] + p[1]);
        a_in_ia = c.sigma * kappa / (p[0] + p[1]);
    } else {
        a_ip_in = p[0] * (c.gamma * p[S::IP] / p[0]);
        a_in_ip = 1.0 / p[0];
    }
}

template <int MODEL, int D>
__device__ inline void rhs_add_dim(const MConst &c, const double *p, const double *sl, int j, double kappa,
                                   double a_ip_in, double a_in_ip, double a_in_ia, double *rhs) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE, IU = S::IU, IP = S::IP;
    const int in = IU + j;
    // the velocity component j of p and sl, selected without indexing
    const double p_in = pick_reg<D>(p + IU, j), s_in = pick_reg<D>(sl + IU, j);
#pragma unroll
    for (int i = 0; i < NE; ++i) {
        double r = rhs[i];
        if (i == 0) {
            r += p_in * sl[0];
            r += p[0] * s_in;
        } else if (MODEL == MULTI && i == 1) {
            r += p_in * sl[1];
            r += p[1] * s_in;
        } else if (i == in) {
            r += p_in * s_in;
            r += a_in_ip * sl[IP];
            if constexpr (MODEL == MULTI) r += a_in_ia * sl[S::IA];
        } else if (i == IP) {
            r += a_ip_in * s_in;
            r += p_in * sl[IP];
        } else {
            r += p_in * sl[i];
        }
        rhs[i] = r;
    }
    // The reference forms the dense product (every A[i][k] * slope[k],
    // euler.hpp:105-116), so a structurally zero entry times a NaN or
    // infinite slope (a limiter's 0/0, slope_limiters.hpp:26-27,41-42)
    // still reaches the sum; add those products when some slope is not
    // finite (for finite slopes they are zeros and change nothing).
    bool fin = true;
#pragma unroll
    for (int k = 0; k < NE; ++k) fin = fin && isfinite(sl[k]);
    if (!fin) {
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            double r = rhs[i];
#pragma unroll
            for (int k = 0; k < NE; ++k) {
                bool nz = (k == i) || (i == 0 && k == in) || (i == in && k == IP) || (i == IP && k == in);
                if constexpr (MODEL == MULTI) nz = nz || (i == 1 && k == in) || (i == in && k == S::IA);
                if (!nz) r += 0.0 * sl[k];
            }
            rhs[i] = r;
        }
    }
}

template <int MODEL, int D>
__device__ inline void compute_rhs(const MConst &c, const double *p, const double (*slope)[Sz<MODEL, D>::NE],
                                   double kappa, double *rhs) {
    double a_ip_in, a_in_ip, a_in_ia;
    rhs_coeffs<MODEL, D>(c, p, kappa, a_ip_in, a_in_ip, a_in_ia);
    for (int i = 0; i < Sz<MODEL, D>::NE; ++i) rhs[i] = 0.0;
    for (int j = 0; j < D; ++j) rhs_add_dim<MODEL, D>(c, p, slope[j], j, kappa, a_ip_in, a_in_ip, a_in_ia, rhs);
}

// HLLC-type star state in primitives: euler.hpp:118-153, multi.hpp:205-246
template <int MODEL, int D>
__device__ inline void middle_state(const MConst &c, const double *pn, const double *pp, double ka, int dim,
                                    double *pm) {
    using S = Sz<MODEL, D>;
    constexpr int NE = S::NE, IU = S::IU, IP = S::IP;
    const int in = IU + dim;
    for (int i = 0; i < NE; ++i) pm[i] = 0.0;
    double c_n, c_p, rn, rp;
    if constexpr (MODEL == MULTI) {
        c_n = sqrt(mgamma(c, pn[S::IA]) * (pn[IP] + mp0(c, pn[S::IA])) / (pn[0] + pn[1]));
        c_p = sqrt(mgamma(c, pp[S::IA]) * (pp[IP] + mp0(c, pp[S::IA])) / (pp[0] + pp[1]));
        rn = pn[0] + pn[1];
        rp = pp[0] + pp[1];
    } else {
        c_n = sqrt(c.gamma * pn[IP] / pn[0]);
        c_p = sqrt(c.gamma * pp[IP] / pp[0]);
        rn = pn[0];
        rp = pp[0];
    }
    const double u = 0.5 * (pn[in] + pp[in]);
    const double cc = 0.5 * (c_n + c_p);
    const double Sn = smin(u - cc, pn[in] - c_n);
    const double Sp = smax(u + cc, pp[in] + c_p);
    double num = pp[IP] - pn[IP] + rn * pn[in] * (Sn - pn[in]) - rp * pp[in] * (Sp - pp[in]);
    if constexpr (MODEL == MULTI) num = num - c.sigma * ka * (pp[S::IA] - pn[S::IA]);
    const double Sm = num / (rn * (Sn - pn[in]) - rp * (Sp - pp[in]));
    if (Sn > 0) {
        for (int i = 0; i < NE; ++i) pm[i] = pn[i];
    } else if (Sp < 0) {
        for (int i = 0; i < NE; ++i) pm[i] = pp[i];
    } else if (Sm > 0 && Sn < 0) {
        pm[0] = pn[0] * (Sn - pn[in]) / (Sn - Sm);
        if constexpr (MODEL == MULTI) pm[1] = pn[1] * (Sn - pn[in]) / (Sn - Sm);
        pm[IP] = pn[IP] + rn * (Sn - pn[in]) * (Sm - pn[in]);
        if constexpr (MODEL == MULTI) pm[S::IA] = pn[S::IA];
        for (int i = 0; i < D; ++i)
            if (MODEL != MULTI || i != dim) pm[IU + i] = pn[IU + i];
        pm[in] = Sm;
    } else {
        pm[0] = pp[0] * (Sp - pp[in]) / (Sp - Sm);
        if constexpr (MODEL == MULTI) pm[1] = pp[1] * (Sp - pp[in]) / (Sp - Sm);
        pm[IP] = pp[IP] + rp * (Sp - pp[in]) * (Sm - pp[in]);
        if constexpr (MODEL == MULTI) pm[S::IA] = pp[S::IA];
        for (int i = 0; i < D; ++i)
            if (MODEL != MULTI || i != dim) pm[IU + i] = pp[IU + i];
        pm[in] = Sm;
    }
}

// get_grad_input: navier_stokes.hpp:51-59, multi.hpp:148-156
template <int MODEL, int D>
__device__ inline void grad_input(const MConst &c, const double *p, double *g) {
    using S = Sz<MODEL, D>;
    if constexpr (S::NG > 0) {
        for (int i = 0; i < S::NG; ++i) g[i] = 0.0;
        for (int i = 0; i < D; ++i) g[i] = p[S::IU + i];
        if constexpr (MODEL == MULTI) {
            g[S::IDT] = mtemperature(c, p[0], p[1], p[S::IP], p[S::IA]);
            g[S::IDA] = p[S::IA];
        } else {
            g[S::IDT] = p[S::IP] / (p[0] * c.ns_R);
        }
    }
}

// viscous stress vector: model/common.hpp:8-24
template <int D, int NGX>
__device__ inline void visc_Fv(const double (*du)[NGX], const double *vel, double mu, int dim, double *fv) {
    double trace = 0.0;
    for (int i = 0; i <= D; ++i) fv[i] = 0.0;
    for (int i = 0; i < D; ++i) {
        fv[i] = du[dim][i] + du[i][dim];
        trace += du[i][i];
    }
    fv[dim] -= (2.0 / 3.0 * trace);
    for (int i = 0; i < D; ++i) fv[D] += fv[i] * vel[i];
    for (int i = 0; i <= D; ++i) fv[i] = fv[i] * mu;
}

template <int MODEL, int D>
__device__ inline double model_mu(const MConst &c, const double *p) {
    if constexpr (MODEL == MULTI) return c.mu1 * p[Sz<MODEL, D>::IA] + c.mu2 * (1 - p[Sz<MODEL, D>::IA]);
    else return c.ns_mu;
}
template <int MODEL, int D>
__device__ inline double model_lambda(const MConst &c, const double *p) {
    return c.lambda1 * p[Sz<MODEL, D>::IA] + c.lambda2 * (1 - p[Sz<MODEL, D>::IA]);
}

// CFL / viscous limit: euler.hpp:203-215, navier_stokes.hpp:101-114, multi.hpp:322-346
template <int MODEL, int D>
__device__ inline double max_delta_time(const MConst &c, const double *p, double size) {
    using S = Sz<MODEL, D>;
    const double cs = sound_speed_prim<MODEL, D>(c, p);
    double m = size / (cs + fabs(p[S::IU]));
    for (int i = 1; i < D; ++i) m = smin(m, size / (cs + fabs(p[S::IU + i])));
    if constexpr (MODEL == EULER) {
        return m;
    } else if constexpr (MODEL == NAVIER_STOKES) {
        const double k = smax(c.ns_mu, c.ns_lambda / c.ns_cp);
        if (k < 1e-15) return m;
        return smin(m, size * size / k * 0.5);
    } else {
        const double k = smax(model_mu<MODEL, D>(c, p), model_lambda<MODEL, D>(c, p) / smin(c.cp1, c.cp2));
        if (k < 1e-15) return m;
        return smin(m, size * size / k / 4.0);
    }
}

// slope limiters: integrator/slope_limiters.hpp:7-57
__device__ inline double limiter(int id, double s1, double s2) {
    switch (id) {
    case 0:
        if (s1 * s2 <= 0.0) return 0.0;
        else if (fabs(s1) < fabs(s2)) return s1;
        else return s2;
    case 1: return 0.5 * (s1 + s2);
    case 2: {
        const double r = s1 / s2;
        return (r + fabs(r)) / (1 + fabs(r)) * s2;
    }
    case 3: {
        const double r = s1 / s2;
        return smax(smax(0.0, smin(2.0 * r, 1.0)), smin(r, 2.0)) * s2;
    }
    case 4: {
        const double r = s1 / s2;
        return 2.0 * r / (r * r + 1) * s2;
    }
    case 5: {
        const double r = (s1 + 1e-9) / (s2 + 1e-9);
        return smax(0.0, smin(2.0 * r, smin((1.0 + 2.0 * r) / 3.0, 2.0))) * s2;
    }
    default: return 0.0;
    }
}

// an object boundary condition: the C ABI's dacti_boundary
struct BndRec {
    int32_t type, pad;
    double normal[3];
    double p[8];
    double fade_in;
};

// ghost state of a domain boundary face: integrator/boundary.hpp:75-123
template <int MODEL, int D>
__device__ inline void boundary_value(const MConst &c, int face, const double *pc, double t, double *pb) {
    using S = Sz<MODEL, D>;
    const int type = c.bc_type[face];
    if (type == BC_FREESLIP) {
        const int fd = face >> 1;
        const double nrm = -(2.0 * (face & 1) - 1.0);
        double dot = 0;
        for (int i = 0; i < S::NE; ++i) pb[i] = pc[i];
        for (int i = 0; i < D; ++i) dot += pc[S::IU + i] * (i == fd ? nrm : 0.0);
        for (int i = 0; i < D; ++i) pb[S::IU + i] = pc[S::IU + i] - (i == fd ? nrm : 0.0) * dot;
        return;
    }
    if (type == BC_NOSLIP) {
        for (int i = 0; i < S::NE; ++i) pb[i] = pc[i];
        for (int i = 0; i < D; ++i) pb[S::IU + i] *= -1.0;
        return;
    }
    if (type == BC_INFLOW || type == BC_INFLOW_R) {
        for (int i = 0; i < S::NE; ++i) pb[i] = c.bc_p[face][i];
        if (!supersonic<MODEL, D>(c, pc)) pb[S::IP] = pc[S::IP];
        const double fi = c.bc_fade_in[face];
        if (fi > 0.0) {
            const double vm = smin(t / fi, 1.0);
            for (int i = 0; i < D; ++i) pb[S::IU + i] *= vm;
        }
        return;
    }
    if (type == BC_DIRICHLET) {
        for (int i = 0; i < S::NE; ++i) pb[i] = c.bc_p[face][i];
        return;
    }
    for (int i = 0; i < S::NE; ++i) pb[i] = pc[i];
}

// ghost state of an object boundary (record r): integrator/boundary.hpp:75-123
// with the record's wall normal and inflow state
template <int MODEL, int D>
__device__ inline void boundary_value_rec(const MConst &c, const GMEM BndRec *r, const double *pc, double t,
                                          double *pb) {
    using S = Sz<MODEL, D>;
    const int type = r->type;
    if (type == BC_FREESLIP) {
        double dot = 0;
        for (int i = 0; i < S::NE; ++i) pb[i] = pc[i];
        for (int i = 0; i < D; ++i) dot += pc[S::IU + i] * r->normal[i];
        for (int i = 0; i < D; ++i) pb[S::IU + i] = pc[S::IU + i] - r->normal[i] * dot;
        return;
    }
    if (type == BC_NOSLIP) {
        for (int i = 0; i < S::NE; ++i) pb[i] = pc[i];
        for (int i = 0; i < D; ++i) pb[S::IU + i] *= -1.0;
        return;
    }
    if (type == BC_INFLOW || type == BC_INFLOW_R) {
        for (int i = 0; i < S::NE; ++i) pb[i] = r->p[i];
        if (!supersonic<MODEL, D>(c, pc)) pb[S::IP] = pc[S::IP];
        const double fi = r->fade_in;
        if (fi > 0.0) {
            const double vm = smin(t / fi, 1.0);
            for (int i = 0; i < D; ++i) pb[S::IU + i] *= vm;
        }
        return;
    }
    if (type == BC_DIRICHLET) {
        for (int i = 0; i < S::NE; ++i) pb[i] = r->p[i];
        return;
    }
    for (int i = 0; i < S::NE; ++i) pb[i] = pc[i];
}

}  // namespace dacti
