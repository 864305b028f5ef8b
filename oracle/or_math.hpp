// oracle/or_math.hpp — small fixed-size vector/matrix types for the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP product; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// The arithmetic mirrors the reference's own value semantics so that an FP32
// instantiation rounds like the reference CPU path:
//   * vector / scalar multiplies by the reciprocal   (src/Vector.cuh:134-140)
//   * dot / norm2 accumulate from 0 in index order     (src/Vector.cuh:153-167)
//   * matrix products accumulate k ascending from 0    (src/Matrix.cuh:146-153)
//   * min/max/abs are the plain ternaries of src/MathHelper.cuh:15-29
// Build with -ffp-contract=off so no FMA changes the rounding.
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

template <class T> inline T tmin(T a, T b) { return a < b ? a : b; }
template <class T> inline T tmax(T a, T b) { return a > b ? a : b; }
template <class T> inline T tabs(T x) { return x < T(0) ? -x : x; }
template <class T> inline T tclamp(T x, T a, T b) { return x < a ? a : (x > b ? b : x); }
template <class T> inline T tsqr(T x) { return x * x; }
template <class T> inline T tmin3(T a, T b, T c) { return tmin(a, tmin(b, c)); }
template <class T> inline T tmin4(T a, T b, T c, T d) { return tmin(tmin(a, b), tmin(c, d)); }
template <class T> inline int tsign(T x) { return x < T(0) ? -1 : 1; }

template <class T, int N> struct Vec {
    T v[N];
    Vec() { for (int i = 0; i < N; i++) v[i] = T(0); }
    T& operator[](int i) { return v[i]; }
    const T& operator[](int i) const { return v[i]; }
    Vec operator+(const Vec& o) const { Vec r; for (int i = 0; i < N; i++) r.v[i] = v[i] + o.v[i]; return r; }
    Vec operator-(const Vec& o) const { Vec r; for (int i = 0; i < N; i++) r.v[i] = v[i] - o.v[i]; return r; }
    Vec operator-() const { Vec r; for (int i = 0; i < N; i++) r.v[i] = -v[i]; return r; }
    Vec operator*(T s) const { Vec r; for (int i = 0; i < N; i++) r.v[i] = v[i] * s; return r; }
    friend Vec operator*(T s, const Vec& a) { Vec r; for (int i = 0; i < N; i++) r.v[i] = s * a.v[i]; return r; }
    Vec operator/(T s) const { T inv = T(1) / s; Vec r; for (int i = 0; i < N; i++) r.v[i] = v[i] * inv; return r; }
    void operator+=(const Vec& o) { for (int i = 0; i < N; i++) v[i] += o.v[i]; }
    void operator-=(const Vec& o) { for (int i = 0; i < N; i++) v[i] -= o.v[i]; }
    void operator*=(T s) { for (int i = 0; i < N; i++) v[i] *= s; }
    T dot(const Vec& o) const { T a = T(0); for (int i = 0; i < N; i++) a += v[i] * o.v[i]; return a; }
    T norm2() const { T a = T(0); for (int i = 0; i < N; i++) a += v[i] * v[i]; return a; }
    T norm() const { return std::sqrt(norm2()); }
    Vec normalized() const { return *this / norm(); }
    void normalize() { *this = *this / norm(); }
};

template <class T> using V2 = Vec<T, 2>;
template <class T> using V3 = Vec<T, 3>;

template <class T> inline V3<T> mk3(T a, T b, T c) { V3<T> r; r[0] = a; r[1] = b; r[2] = c; return r; }
template <class T> inline V2<T> mk2(T a, T b) { V2<T> r; r[0] = a; r[1] = b; return r; }
template <class T> inline V3<T> cross(const V3<T>& a, const V3<T>& b) {
    return mk3<T>(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
template <class T> inline T cross2(const V2<T>& a, const V2<T>& b) { return a[0] * b[1] - a[1] * b[0]; }
// a . (b x c)   (src/MathHelper.cuh:43-45)
template <class T> inline T mixed(const V3<T>& a, const V3<T>& b, const V3<T>& c) { return a.dot(cross(b, c)); }

template <class T, int R, int C> struct Mat {
    T m[R][C];
    Mat() { for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) m[i][j] = T(0); }
    static Mat ident(T s) { Mat r; for (int i = 0; i < R && i < C; i++) r.m[i][i] = s; return r; }
    T& operator()(int i, int j) { return m[i][j]; }
    const T& operator()(int i, int j) const { return m[i][j]; }
    Mat operator+(const Mat& o) const { Mat r; for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) r.m[i][j] = m[i][j] + o.m[i][j]; return r; }
    Mat operator-(const Mat& o) const { Mat r; for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) r.m[i][j] = m[i][j] - o.m[i][j]; return r; }
    Mat operator-() const { Mat r; for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) r.m[i][j] = -m[i][j]; return r; }
    friend Mat operator*(T s, const Mat& a) { Mat r; for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) r.m[i][j] = s * a.m[i][j]; return r; }
    Mat operator/(T s) const { T inv = T(1) / s; Mat r; for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) r.m[i][j] = m[i][j] * inv; return r; }
    void operator+=(const Mat& o) { for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) m[i][j] += o.m[i][j]; }
    void operator-=(const Mat& o) { for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) m[i][j] -= o.m[i][j]; }
    Vec<T, R> operator*(const Vec<T, C>& x) const {
        Vec<T, R> r;
        for (int i = 0; i < R; i++) for (int j = 0; j < C; j++) r.v[i] += m[i][j] * x.v[j];
        return r;
    }
    template <int Q> Mat<T, R, Q> operator*(const Mat<T, C, Q>& o) const {
        Mat<T, R, Q> r;
        for (int i = 0; i < R; i++)
            for (int j = 0; j < Q; j++)
                for (int k = 0; k < C; k++) r.m[i][j] += m[i][k] * o.m[k][j];
        return r;
    }
    Mat<T, C, R> transpose() const { Mat<T, C, R> r; for (int i = 0; i < C; i++) for (int j = 0; j < R; j++) r.m[i][j] = m[j][i]; return r; }
    Vec<T, C> row(int i) const { Vec<T, C> r; for (int j = 0; j < C; j++) r.v[j] = m[i][j]; return r; }
    Vec<T, R> col(int j) const { Vec<T, R> r; for (int i = 0; i < R; i++) r.v[i] = m[i][j]; return r; }
};

template <class T, int N, int P> inline Mat<T, N, P> outer(const Vec<T, N>& a, const Vec<T, P>& b) {
    Mat<T, N, P> r;
    for (int i = 0; i < N; i++) for (int j = 0; j < P; j++) r.m[i][j] = a.v[i] * b.v[j];
    return r;
}

// 2x2 inverse by the adjugate (src/Matrix.cuh:182-196)
template <class T> inline Mat<T, 2, 2> inverse2(const Mat<T, 2, 2>& a) {
    Mat<T, 2, 2> r;
    T invDet = T(1) / (a.m[0][0] * a.m[1][1] - a.m[0][1] * a.m[1][0]);
    r.m[0][0] = a.m[1][1] * invDet;
    r.m[0][1] = -a.m[0][1] * invDet;
    r.m[1][0] = -a.m[1][0] * invDet;
    r.m[1][1] = a.m[0][0] * invDet;
    return r;
}

}  // namespace orc
