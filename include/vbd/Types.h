// Eigen-free value types for the source-compatible VBD drop-in (header-only).
//
// Mirrors the names the reference hot path uses from Types.h (Types.h:18-70, :83-120):
// Real/FloatingType, IdType, Vec3, Vec3I, Mat3/Matrix3r, TVerticesMat (3 x N column-major,
// i.e. xyz contiguous per vertex), VecDynamic, TetVIdsArr, EPSILON and AnalyticSolver3x3.
// Only the operations the VBD path needs are provided; Eigen is not required.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <vector>

using FloatingType = float;
using Real = FloatingType;

namespace VBD {

using IdType = int32_t;

#ifndef SQR
#define SQR(x) ((x) * (x))
#endif
#ifndef CUBE
#define CUBE(x) ((x) * (x) * (x))
#endif
#ifndef EPSILON
#define EPSILON 1e-8
#endif

struct Vec3 {
    Real v[3];
    Vec3() : v{0, 0, 0} {}
    Vec3(Real a, Real b, Real c) : v{a, b, c} {}
    Real& operator[](int i) { return v[i]; }
    Real operator[](int i) const { return v[i]; }
    Real& operator()(int i) { return v[i]; }
    Real operator()(int i) const { return v[i]; }
    Real x() const { return v[0]; }
    Real y() const { return v[1]; }
    Real z() const { return v[2]; }
    Vec3 operator+(const Vec3& o) const { return Vec3(v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]); }
    Vec3 operator-(const Vec3& o) const { return Vec3(v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]); }
    Vec3 operator*(Real s) const { return Vec3(v[0] * s, v[1] * s, v[2] * s); }
    Vec3 operator/(Real s) const { return Vec3(v[0] / s, v[1] / s, v[2] / s); }
    Vec3& operator+=(const Vec3& o) {
        v[0] += o.v[0]; v[1] += o.v[1]; v[2] += o.v[2];
        return *this;
    }
    Vec3& operator-=(const Vec3& o) {
        v[0] -= o.v[0]; v[1] -= o.v[1]; v[2] -= o.v[2];
        return *this;
    }
    Real dot(const Vec3& o) const { return v[0] * o.v[0] + (v[1] * o.v[1] + v[2] * o.v[2]); }
    Vec3 cross(const Vec3& o) const {
        return Vec3(v[1] * o.v[2] - v[2] * o.v[1], v[2] * o.v[0] - v[0] * o.v[2], v[0] * o.v[1] - v[1] * o.v[0]);
    }
    void setZero() { v[0] = v[1] = v[2] = 0; }
};
inline Vec3 operator*(Real s, const Vec3& a) { return a * s; }

struct Vec3I {
    IdType v[3];
    Vec3I() : v{-1, -1, -1} {}
    Vec3I(IdType a, IdType b, IdType c) : v{a, b, c} {}
    IdType& operator[](int i) { return v[i]; }
    IdType operator[](int i) const { return v[i]; }
};

// 3x3 matrix accessed as (row, col); stored row-major.
struct Mat3 {
    Real m[9];
    Mat3() { setZero(); }
    Real& operator()(int r, int c) { return m[3 * r + c]; }
    Real operator()(int r, int c) const { return m[3 * r + c]; }
    void setZero() {
        for (Real& a : m) a = 0;
    }
    static Mat3 Zero() { return Mat3(); }
    static Mat3 Identity() {
        Mat3 I;
        I(0, 0) = I(1, 1) = I(2, 2) = 1;
        return I;
    }
    Real determinant() const {
        const Mat3& A = *this;
        return A(0, 0) * (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) - A(0, 1) * (A(1, 0) * A(2, 2) - A(1, 2) * A(2, 0)) +
               A(0, 2) * (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0));
    }
    Mat3& operator+=(const Mat3& o) {
        for (int i = 0; i < 9; ++i) m[i] += o.m[i];
        return *this;
    }
};
using Matrix3r = Mat3;

// 3 x N column-major vertex matrix (TVerticesMat, Types.h:53): column i = vertex i.
class TVerticesMat {
   public:
    struct Col {
        Real* p;
        Real& operator[](int i) { return p[i]; }
        Real& operator()(int i) { return p[i]; }
        Col& operator=(const Vec3& a) {
            p[0] = a[0]; p[1] = a[1]; p[2] = a[2];
            return *this;
        }
        Col& operator+=(const Vec3& a) {
            p[0] += a[0]; p[1] += a[1]; p[2] += a[2];
            return *this;
        }
        operator Vec3() const { return Vec3(p[0], p[1], p[2]); }
    };
    void resize(size_t rows, size_t cols) {
        if (rows != 3) throw std::invalid_argument("TVerticesMat has 3 rows");
        d_.assign(3 * cols, Real(0));
        n_ = cols;
    }
    void setZero() { std::fill(d_.begin(), d_.end(), Real(0)); }
    size_t cols() const { return n_; }
    size_t rows() const { return 3; }
    Real* data() { return d_.data(); }
    const Real* data() const { return d_.data(); }
    Col col(size_t i) { return Col{d_.data() + 3 * i}; }
    Vec3 col(size_t i) const { return Vec3(d_[3 * i], d_[3 * i + 1], d_[3 * i + 2]); }
    Real& operator()(int r, size_t c) { return d_[3 * c + r]; }
    Real operator()(int r, size_t c) const { return d_[3 * c + r]; }

   private:
    std::vector<Real> d_;
    size_t n_ = 0;
};

// Dynamic vector with the accessors the reference uses on VecDynamic (Types.h:54).
class VecDynamic {
   public:
    void resize(size_t n) { d_.assign(n, Real(0)); }
    void setZero() { std::fill(d_.begin(), d_.end(), Real(0)); }
    size_t size() const { return d_.size(); }
    Real& operator()(size_t i) { return d_[i]; }
    Real operator()(size_t i) const { return d_[i]; }
    Real& operator[](size_t i) { return d_[i]; }
    Real operator[](size_t i) const { return d_[i]; }
    Real* data() { return d_.data(); }
    const Real* data() const { return d_.data(); }

   private:
    std::vector<Real> d_;
};

using TetVIdsArr = std::array<IdType, 4>;
using TTetVIdsArr = TetVIdsArr;

// AnalyticSolver3x3 (Types.h:83-120): adjugate solve; zero and false when |det| < 1e-8.
class AnalyticSolver3x3 {
   public:
    static inline bool solve(const Mat3& A, const Vec3& b, Vec3& x) {
        const FloatingType det = A.determinant();
        if (std::abs(det) < EPSILON) {
            x.setZero();
            return false;
        }
        const FloatingType invDet = 1.0f / det;
        Mat3 adj;
        adj(0, 0) = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
        adj(0, 1) = A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2);
        adj(0, 2) = A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1);
        adj(1, 0) = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2);
        adj(1, 1) = A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0);
        adj(1, 2) = A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2);
        adj(2, 0) = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
        adj(2, 1) = A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1);
        adj(2, 2) = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
        for (int r = 0; r < 3; ++r) x[r] = invDet * (adj(r, 0) * b[0] + (adj(r, 1) * b[1] + adj(r, 2) * b[2]));
        return true;
    }
    static inline bool solveWithRegularization(const Mat3& A, const Vec3& b, Vec3& x, FloatingType reg = 1e-6f) {
        Mat3 R = A;
        R(0, 0) += reg; R(1, 1) += reg; R(2, 2) += reg;
        return solve(R, b, x);
    }
};

}  // namespace VBD
