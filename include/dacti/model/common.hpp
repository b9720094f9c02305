// dacti/model/common.hpp -- model tags of the reference's solver interface.
//
// The reference selects its physics with a template argument:
// model::euler<D>, model::navier_stokes<D>, model::multi<D>
// (src/lib/model/euler.hpp:18-47, navier_stokes.hpp:14-39,
// multi.hpp:18-75).  These tags carry the same static constants and value
// types, so writers and observers written against the reference compile;
// the device implementations behind them are the GPU kernels, selected
// through the C ABI's model id.
#ifndef DACTI_MODEL_COMMON_HPP
#define DACTI_MODEL_COMMON_HPP

#include <array>
#include <cstddef>

#include "../../dacti_gpu.h"
#include "../core/types.hpp"

namespace dacti {

namespace model {

// value types of a model (model/euler.hpp:20-32): conv_t / prim_t NumEqns,
// flux_t NumFlux, parm_t NumParm, grad_t NumGrad, matS_t / matG_t one row
// per dimension
template <size_t D, size_t NE, size_t NF, size_t NP, size_t NG>
struct types {
    using vecn_t = linalg::vec<scalar_t, D>;
    using conv_t = linalg::vec<scalar_t, NE>;
    using prim_t = linalg::vec<scalar_t, NE>;
    using flux_t = linalg::vec<scalar_t, NF>;
    using parm_t = linalg::vec<scalar_t, NP>;
    using grad_t = linalg::vec<scalar_t, NG>;
    using matS_t = std::array<prim_t, D>;
    using matG_t = std::array<grad_t, D>;
};

template <size_t D>
struct euler : types<D, D + 2, D + 2, 0, 0> {
    static constexpr int id = DACTI_MODEL_EULER;
    static constexpr bool Multiphase = false, HasSource = false, Viscosity = false, HLLC_solver = true;
    static constexpr size_t NumDims = D, NumEqns = D + 2, NumFlux = D + 2, NumParm = 0, NumGrad = 0;
    static constexpr size_t iu = 1, ip = D + 1;
    static constexpr scalar_t gamma = 1.4;
};

template <size_t D>
struct navier_stokes : types<D, D + 2, D + 2, 0, D + 1> {
    static constexpr int id = DACTI_MODEL_NAVIER_STOKES;
    static constexpr bool Multiphase = false, HasSource = false, Viscosity = true, HLLC_solver = true;
    static constexpr bool HeatDiffusion = true;
    static constexpr size_t NumDims = D, NumEqns = D + 2, NumFlux = D + 2, NumParm = 0, NumGrad = D + 1;
    static constexpr size_t iu = 1, ip = D + 1, dT = D;
    static constexpr scalar_t gamma = 1.4, cp = 1005.0, R = 287.0;
    // viscosity and conductivity of the case (navier_stokes.hpp:30-31)
    static inline scalar_t mu_ = 0.0, lambda_ = 0.0;
};

template <size_t D>
struct multi : types<D, D + 4, D + 4 + D + 1, D + 4, D + 3> {
    static constexpr int id = DACTI_MODEL_MULTI;
    static constexpr bool Multiphase = true, HasSource = true, Viscosity = true, HLLC_solver = true;
    static constexpr bool HeatDiffusion = true;
    static constexpr size_t NumDims = D, NumEqns = D + 4, NumFlux = D + 4 + D + 1, NumParm = D + 4, NumGrad = D + 3;
    static constexpr size_t iu = 2, ip = D + 2, ia = D + 3;
    // parameters: k[0] viscosity, k[ikn..ikn+D) interface normal, k[kappa]
    // curvature, k[alphas] smoothed volume fraction
    static constexpr size_t ikn = 2, kappa = 2 + D, alphas = 3 + D;
    // gradient inputs: velocity[D], then temperature and volume fraction
    static constexpr size_t dT = D, da = D + 1;
    // constants of the case; the reference keeps them process-global
    // (multi.hpp:50-62), here Config<multi<D>> sets them when it loads a
    // case and each integrator handle carries its own copy
    static inline scalar_t sigma = 0.0, gamma1 = 6.12, gamma2 = 1.4, P1 = 343000000.0, P2 = 0.0;
    static inline scalar_t cp1 = 4181.0, cp2 = 1005.0, cv1 = 4181.0 / 6.12, cv2 = 1005.0 / 1.4;
    static inline scalar_t mu1 = 0.0, mu2 = 0.0, lambda1 = 0.0, lambda2 = 0.0;
};

}  // namespace model


}  // namespace dacti

#endif
