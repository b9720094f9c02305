// dacti/core/linalg.hpp -- small fixed-size vectors for host-side callers.
//
// The reference's writers and observers use linalg::vec<S, N>
// (src/lib/core/linalg.hpp:37-559) through the cell accessors
// (get_p, get_velocity, get_center, ...).  This is the part of that
// interface they touch: element access, arithmetic, norms, projections and
// the quadrant/unit constructors.  A default-constructed vec is zero, as
// in the reference (core/linalg.hpp:47-50).
#ifndef DACTI_LINALG_HPP
#define DACTI_LINALG_HPP

#include <array>
#include <cmath>
#include <cstddef>
#include <type_traits>

namespace dacti {

using scalar_t = double;

namespace linalg {

template <typename S, size_t N>
struct vec {
    std::array<S, N> v{};

    vec() = default;
    template <typename... A, typename = std::enable_if_t<sizeof...(A) == N && (N > 1)>>
    vec(A... a) : v{S(a)...} {}
    explicit vec(const std::array<S, N> &a) : v(a) {}

    static vec constant(S c) {
        vec r;
        r.v.fill(c);
        return r;
    }
    static vec Zero() { return vec{}; }
    static vec One() { return constant(S(1)); }
    static vec unit(size_t d) {
        vec r;
        r.v[d] = S(1);
        return r;
    }
    // the corner of the unit cube selected by the bits of i: -1 or +1 per axis
    static vec quadrant(size_t i) {
        vec r;
        for (size_t d = 0; d < N; ++d) r.v[d] = ((i >> d) & 1) ? S(1) : S(-1);
        return r;
    }

    static constexpr size_t size() { return N; }
    S &operator[](size_t i) { return v[i]; }
    const S &operator[](size_t i) const { return v[i]; }
    S *data() { return v.data(); }
    const S *data() const { return v.data(); }
    auto begin() { return v.begin(); }
    auto end() { return v.end(); }
    auto begin() const { return v.begin(); }
    auto end() const { return v.end(); }

    vec &operator+=(const vec &o) {
        for (size_t i = 0; i < N; ++i) v[i] += o.v[i];
        return *this;
    }
    vec &operator-=(const vec &o) {
        for (size_t i = 0; i < N; ++i) v[i] -= o.v[i];
        return *this;
    }
    vec &operator*=(S s) {
        for (size_t i = 0; i < N; ++i) v[i] *= s;
        return *this;
    }
    friend vec operator+(vec a, const vec &b) { return a += b; }
    friend vec operator-(vec a, const vec &b) { return a -= b; }
    friend vec operator-(vec a) {
        for (size_t i = 0; i < N; ++i) a.v[i] = -a.v[i];
        return a;
    }
    friend vec operator*(vec a, S s) { return a *= s; }
    friend vec operator*(S s, vec a) { return a *= s; }
    friend vec operator/(vec a, S s) {
        for (size_t i = 0; i < N; ++i) a.v[i] /= s;
        return a;
    }

    S dot(const vec &o) const {
        S s = S(0);
        for (size_t i = 0; i < N; ++i) s += v[i] * o.v[i];
        return s;
    }
    S squaredNorm() const { return dot(*this); }
    S norm() const { return std::sqrt(squaredNorm()); }
    vec cwiseMul(const vec &o) const {
        vec r;
        for (size_t i = 0; i < N; ++i) r.v[i] = v[i] * o.v[i];
        return r;
    }
    // first M components, zero-padded
    template <size_t M>
    vec<S, M> project() const {
        vec<S, M> r;
        for (size_t i = 0; i < M && i < N; ++i) r[i] = v[i];
        return r;
    }
    // first M components, padded with d
    template <size_t M>
    vec<S, M> project_with_default(S d) const {
        vec<S, M> r = vec<S, M>::constant(d);
        for (size_t i = 0; i < M && i < N; ++i) r[i] = v[i];
        return r;
    }
    // 2D: rotation by -90 degrees (x, y) -> (y, -x)
    vec rotaten90() const {
        static_assert(N == 2, "rotaten90 is defined for 2D vectors");
        return vec(v[1], -v[0]);
    }
};

}  // namespace linalg

using vec2_t = linalg::vec<scalar_t, 2>;
using vec3_t = linalg::vec<scalar_t, 3>;

}  // namespace dacti

#endif
