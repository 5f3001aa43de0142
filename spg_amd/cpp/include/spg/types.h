// Small fixed-size types standing in for the Eigen aliases of spg/include/spg/types.h:8-41
// (Real = double, Vector3, Matrix2/3, Int2/3/4). Only what the VBD path uses.
#pragma once

#include <array>
#include <cmath>
#include <vector>

namespace spg {

using Real = double;
using Int2 = std::array<int, 2>;
using Int3 = std::array<int, 3>;
using Int4 = std::array<int, 4>;

template <int R, int C>
struct Matrix {
    std::array<Real, R * C> m{};
    Real &operator()(int r, int c) { return m[r * C + c]; }
    Real operator()(int r, int c) const { return m[r * C + c]; }
    static Matrix Zero() { return Matrix{}; }
    static Matrix Identity()
    {
        Matrix a{};
        for (int i = 0; i < (R < C ? R : C); ++i) a(i, i) = 1;
        return a;
    }
};
using Matrix1 = Matrix<1, 1>;
using Matrix2 = Matrix<2, 2>;
using Matrix3 = Matrix<3, 3>;

struct Vector3 {
    std::array<Real, 3> v{};
    Vector3() = default;
    Vector3(Real x, Real y, Real z) : v{x, y, z} {}
    Real &operator[](int i) { return v[i]; }
    Real operator[](int i) const { return v[i]; }
    Real x() const { return v[0]; }
    Real y() const { return v[1]; }
    Real z() const { return v[2]; }
    static Vector3 Zero() { return {}; }
    Vector3 operator+(const Vector3 &o) const { return {v[0] + o[0], v[1] + o[1], v[2] + o[2]}; }
    Vector3 operator-(const Vector3 &o) const { return {v[0] - o[0], v[1] - o[1], v[2] - o[2]}; }
    Vector3 operator*(Real s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
    Vector3 &operator+=(const Vector3 &o)
    {
        for (int i = 0; i < 3; ++i) v[i] += o[i];
        return *this;
    }
    Real dot(const Vector3 &o) const { return v[0] * o[0] + v[1] * o[1] + v[2] * o[2]; }
    Vector3 cross(const Vector3 &o) const
    {
        return {v[1] * o[2] - v[2] * o[1], v[2] * o[0] - v[0] * o[2], v[0] * o[1] - v[1] * o[0]};
    }
    Real squaredNorm() const { return dot(*this); }
    Real norm() const { return std::sqrt(squaredNorm()); }
    Vector3 normalized() const
    {
        const Real n = norm();
        return n > 0 ? Vector3{v[0] / n, v[1] / n, v[2] / n} : *this;
    }
    void setZero() { v = {0, 0, 0}; }
};

using VectorX = std::vector<Real>;

}  // namespace spg
