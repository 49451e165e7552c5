// Small fixed-size vectors standing in for the Eigen types the reference's
// grid containers use (Vector3i, Vector3d; core/Common.h:58-59). Only the
// members the drop-in Pressure / MagneticField and their tests need.
#pragma once
#include <array>
#include <cmath>

namespace Pivot {

template <class T>
struct Vector3 {
    std::array<T, 3> v{T(0), T(0), T(0)};
    Vector3() = default;
    Vector3(T x, T y, T z) : v{x, y, z} {}
    T &operator[](int a) { return v[a]; }
    T const &operator[](int a) const { return v[a]; }
    T &x() { return v[0]; }
    T &y() { return v[1]; }
    T &z() { return v[2]; }
    T x() const { return v[0]; }
    T y() const { return v[1]; }
    T z() const { return v[2]; }
    static Vector3 Zero() { return {T(0), T(0), T(0)}; }
    static Vector3 Ones() { return {T(1), T(1), T(1)}; }
    static Vector3 Constant(T c) { return {c, c, c}; }
    static Vector3 Unit(int a) {
        Vector3 u = Zero();
        u[a] = T(1);
        return u;
    }
    template <class U>
    Vector3<U> cast() const { return {U(v[0]), U(v[1]), U(v[2])}; }
    T prod() const { return v[0] * v[1] * v[2]; }
    T dot(Vector3 const &o) const { return v[0] * o[0] + v[1] * o[1] + v[2] * o[2]; }
    T squaredNorm() const { return dot(*this); }
    T norm() const { return std::sqrt(squaredNorm()); }
    bool operator==(Vector3 const &o) const { return v == o.v; }
    bool operator!=(Vector3 const &o) const { return v != o.v; }
    Vector3 operator+(Vector3 const &o) const { return {v[0] + o[0], v[1] + o[1], v[2] + o[2]}; }
    Vector3 operator-(Vector3 const &o) const { return {v[0] - o[0], v[1] - o[1], v[2] - o[2]}; }
    Vector3 operator*(T s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
    Vector3 operator/(T s) const { return {v[0] / s, v[1] / s, v[2] / s}; }
};

using Vector3i = Vector3<int>;
using Vector3d = Vector3<double>;

}  // namespace Pivot
