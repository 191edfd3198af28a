// dare::utils::Vector<N, T>: the fixed-size vector every index, coordinate and stencil entry of
// the reference is made of (Utilities/Vector.h:44-840, Vector_traits.h:56-230).
// Kept alike because the call surface depends on it: element-wise +, -, *, / with another
// vector or a scalar, scalar * v == v * scalar and scalar / v == v / scalar (the reference's
// definition, Vector.h:818-834), missing constructor arguments default to zero, and the named
// accessors i(), j(), k() (integral) and x(), y(), z() (floating point). Storage is a
// std::array, so the type also works wherever the mirror's code used one.
#pragma once

#include <array>
#include <cmath>
#include <cstddef>
#include <type_traits>

namespace dare::utils {

template <std::size_t N, typename T>
class Vector : public std::array<T, N> {
public:
    using Base = std::array<T, N>;

    Vector() : Base{} {}
    // Vector(a, b, ...): the first values in order, the rest zero (Vector.inl:28-34, 362-379).
    // With every value given it converts implicitly (so {1, 2, 3} initialises an argument);
    // with fewer it is explicit, as in the reference, so a scalar never turns into a vector.
    template <typename A, typename... Ts,
              std::enable_if_t<std::is_convertible_v<A, T> && (std::is_convertible_v<Ts, T> && ...) &&
                                   sizeof...(Ts) + 1 == N,
                               int> = 0>
    Vector(const A& a, const Ts&... rest) : Base{} {     // NOLINT
        set(a, rest...);
    }
    template <typename A, typename... Ts,
              std::enable_if_t<std::is_convertible_v<A, T> && (std::is_convertible_v<Ts, T> && ...) &&
                                   (sizeof...(Ts) + 1 < N),
                               int> = 0>
    explicit Vector(const A& a, const Ts&... rest) : Base{} {
        set(a, rest...);
    }
    template <typename A, typename = std::enable_if_t<std::is_convertible_v<A, T> && !std::is_same_v<A, T>>>
    Vector(const Vector<N, A>& o) : Base{} {              // NOLINT: converting copy, Vector.inl:36-42
        for (std::size_t n = 0; n < N; n++) (*this)[n] = static_cast<T>(o[n]);
    }
    Vector(const Base& a) : Base(a) {}                    // NOLINT

    template <typename A>
    void SetAllValues(const A& v) {
        for (auto& e : *this) e = static_cast<T>(v);
    }
    template <typename... Ts>
    void SetValues(const Ts&... v) {
        static_assert(sizeof...(Ts) <= N, "too many values");
        Base::fill(T{});
        set(v...);
    }

    // named accessors (Vector_traits.h:122-230)
    T& i() { return (*this)[0]; }
    T i() const { return (*this)[0]; }
    T& j() { static_assert(N > 1); return (*this)[1]; }
    T j() const { static_assert(N > 1); return (*this)[1]; }
    T& k() { static_assert(N > 2); return (*this)[2]; }
    T k() const { static_assert(N > 2); return (*this)[2]; }
    T& x() { return (*this)[0]; }
    T x() const { return (*this)[0]; }
    T& y() { static_assert(N > 1); return (*this)[1]; }
    T y() const { static_assert(N > 1); return (*this)[1]; }
    T& z() { static_assert(N > 2); return (*this)[2]; }
    T z() const { static_assert(N > 2); return (*this)[2]; }

private:
    template <typename... Ts>
    void set(const Ts&... v) {
        const T vals[] = {static_cast<T>(v)...};
        for (std::size_t n = 0; n < sizeof...(Ts) && n < N; n++) (*this)[n] = vals[n];
    }

public:
#define DARE_VECTOR_OP(OP)                                                   \
    Vector operator OP(const Vector& o) const {                              \
        Vector r(*this);                                                     \
        for (std::size_t n = 0; n < N; n++) r[n] = (*this)[n] OP o[n];       \
        return r;                                                            \
    }                                                                        \
    Vector operator OP(const T& v) const {                                   \
        Vector r(*this);                                                     \
        for (std::size_t n = 0; n < N; n++) r[n] = (*this)[n] OP v;          \
        return r;                                                            \
    }                                                                        \
    void operator OP##=(const Vector& o) {                                   \
        for (std::size_t n = 0; n < N; n++) (*this)[n] = (*this)[n] OP o[n]; \
    }                                                                        \
    void operator OP##=(const T& v) {                                        \
        for (std::size_t n = 0; n < N; n++) (*this)[n] = (*this)[n] OP v;    \
    }
    DARE_VECTOR_OP(+)
    DARE_VECTOR_OP(-)
    DARE_VECTOR_OP(*)
    DARE_VECTOR_OP(/)
#undef DARE_VECTOR_OP

    bool operator==(const Vector& o) const {
        for (std::size_t n = 0; n < N; n++)
            if (!((*this)[n] == o[n])) return false;
        return true;
    }
    bool operator!=(const Vector& o) const { return !(*this == o); }

    T dot(const Vector& o) const {
        T s{};
        for (std::size_t n = 0; n < N; n++) s += (*this)[n] * o[n];
        return s;
    }
    T length() const { return std::sqrt(dot(*this)); }
    T AllSum() const {
        T s{};
        for (const auto& e : *this) s += e;
        return s;
    }
    T AllAbsSum() const {
        T s{};
        for (const auto& e : *this) s += std::abs(e);
        return s;
    }
};

// scalar (op) vector: Vector.h:818-834 defines both as vector (op) scalar
template <std::size_t N, typename T>
Vector<N, T> operator*(const T& s, const Vector<N, T>& v) {
    return v * s;
}
template <std::size_t N, typename T>
Vector<N, T> operator/(const T& s, const Vector<N, T>& v) {
    return v / s;
}

}  // namespace dare::utils
