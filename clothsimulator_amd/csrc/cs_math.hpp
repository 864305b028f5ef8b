// cs_math.hpp — scalar/vector helpers shared by the gfx950 kernels and the host side.
//
// Rounding conventions follow the reference's value types so FP32 results
// agree with its CPU path to the last bit where the operation sequence is the
// same (bounds, Morton codes, distances, node geometry):
//   * vector / scalar multiplies by the reciprocal   (src/Vector.cuh:134-140)
//   * dot / norm2 accumulate x, y, z in order from 0 (src/Vector.cuh:153-167)
//   * min/max/abs are the plain ternaries             (src/MathHelper.cuh:15-29)
// All sources are compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define CS_HD __host__ __device__ __forceinline__

namespace cs {

// Pointers read from memory (job tables, per-cloth descriptors) are generic, and
// generic accesses compile to flat instructions; device code casts them to the
// global address space so they load and store as global memory.
template <class T> using G1 = __attribute__((address_space(1))) T;
template <class T> __device__ __forceinline__ G1<T>* g1(T* p) { return (G1<T>*)p; }
// float4 as a builtin vector (copyable from any address space, unlike the float4 class)
typedef float f4n __attribute__((ext_vector_type(4)));


template <class T> CS_HD T tmin(T a, T b) { return a < b ? a : b; }
template <class T> CS_HD T tmax(T a, T b) { return a > b ? a : b; }
template <class T> CS_HD T tabs(T x) { return x < T(0) ? -x : x; }
template <class T> CS_HD T tclamp(T x, T a, T b) { return x < a ? a : (x > b ? b : x); }
template <class T> CS_HD T tsqr(T x) { return x * x; }

template <class T> struct V3 {
    T x, y, z;
};
using V3f = V3<float>;
using V3d = V3<double>;

template <class T> CS_HD V3<T> mk(T a, T b, T c) { V3<T> r; r.x = a; r.y = b; r.z = c; return r; }
template <class T> CS_HD V3<T> operator+(V3<T> a, V3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> CS_HD V3<T> operator-(V3<T> a, V3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> CS_HD V3<T> operator-(V3<T> a) { return mk(-a.x, -a.y, -a.z); }
template <class T> CS_HD V3<T> operator*(T s, V3<T> a) { return mk(s * a.x, s * a.y, s * a.z); }
template <class T> CS_HD V3<T> operator*(V3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <class T> CS_HD V3<T> operator/(V3<T> a, T s) { T inv = T(1) / s; return mk(a.x * inv, a.y * inv, a.z * inv); }
template <class T> CS_HD void operator+=(V3<T>& a, V3<T> b) { a.x += b.x; a.y += b.y; a.z += b.z; }
template <class T> CS_HD T dot(V3<T> a, V3<T> b) { T s = T(0); s += a.x * b.x; s += a.y * b.y; s += a.z * b.z; return s; }
template <class T> CS_HD T norm2(V3<T> a) { return dot(a, a); }
template <class T> CS_HD T norm(V3<T> a) { return sqrt(norm2(a)); }
template <class T> CS_HD V3<T> normalized(V3<T> a) { return a / norm(a); }
template <class T> CS_HD V3<T> cross(V3<T> a, V3<T> b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> CS_HD T mixed(V3<T> a, V3<T> b, V3<T> c) { return dot(a, cross(b, c)); }
template <class T> CS_HD T comp(V3<T> a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

CS_HD V3f ld3(const float4& a) { return mk(a.x, a.y, a.z); }
CS_HD float4 st4(V3f a, float w = 0.0f) { return make_float4(a.x, a.y, a.z, w); }
CS_HD V3d tod(V3f a) { return mk((double)a.x, (double)a.y, (double)a.z); }

// Signed vertex-face / edge-edge distance with barycentric weights
// (src/MathHelper.cu:38-69).  Returns false in the degenerate case where the
// reference returns INFINITY with w unset (a case no caller can accept).
template <class T>
CS_HD bool signed_vf(V3<T> x, V3<T> y0, V3<T> y1, V3<T> y2, V3<T>& n, T w[4], T& h) {
    n = cross(normalized(y1 - y0), normalized(y2 - y0));
    if (norm2(n) < T(1e-6f)) return false;
    n = normalized(n);
    h = dot(x - y0, n);
    T b0 = mixed(y1 - x, y2 - x, n);
    T b1 = mixed(y2 - x, y0 - x, n);
    T b2 = mixed(y0 - x, y1 - x, n);
    w[0] = T(1);
    w[1] = -b0 / (b0 + b1 + b2);
    w[2] = -b1 / (b0 + b1 + b2);
    w[3] = -b2 / (b0 + b1 + b2);
    return true;
}

template <class T>
CS_HD bool signed_ee(V3<T> x0, V3<T> x1, V3<T> y0, V3<T> y1, V3<T>& n, T w[4], T& h) {
    n = cross(normalized(x1 - x0), normalized(y1 - y0));
    if (norm2(n) < T(1e-6f)) return false;
    n = normalized(n);
    h = dot(x0 - y0, n);
    T a0 = mixed(y1 - x1, y0 - x1, n);
    T a1 = mixed(y0 - x0, y1 - x0, n);
    T b0 = mixed(x0 - y1, x1 - y1, n);
    T b1 = mixed(x1 - y0, x0 - y0, n);
    w[0] = a0 / (a0 + a1);
    w[1] = a1 / (a0 + a1);
    w[2] = -b0 / (b0 + b1);
    w[3] = -b1 / (b0 + b1);
    return true;
}

// Axis-aligned box with the reference's empty value (FLT_MAX, -FLT_MAX) (src/Bounds.cu:3-6).
struct Box {
    float lo[3], hi[3];
};
CS_HD Box box_empty() {
    Box b;
    for (int i = 0; i < 3; i++) { b.lo[i] = 3.402823466e+38f; b.hi[i] = -3.402823466e+38f; }
    return b;
}
CS_HD void box_add(Box& b, V3f p) {
    float c[3] = {p.x, p.y, p.z};
    for (int i = 0; i < 3; i++) { b.lo[i] = tmin(b.lo[i], c[i]); b.hi[i] = tmax(b.hi[i], c[i]); }
}
CS_HD void box_add(Box& b, const Box& o) {
    for (int i = 0; i < 3; i++) { b.lo[i] = tmin(b.lo[i], o.lo[i]); b.hi[i] = tmax(b.hi[i], o.hi[i]); }
}
// a.overlap(b.dilate(th)) (src/Bounds.cu:63-87)
CS_HD bool box_overlap(const Box& a, const Box& b, float th) {
    for (int i = 0; i < 3; i++) {
        if (a.lo[i] > b.hi[i] + th) return false;
        if (a.hi[i] < b.lo[i] - th) return false;
    }
    return true;
}

// 3x3 inverse by the adjugate (row-major blocks)
__device__ __forceinline__ void inv3(const float a[9], float o[9]) {
    float c00 = a[4] * a[8] - a[5] * a[7];
    float c01 = a[5] * a[6] - a[3] * a[8];
    float c02 = a[3] * a[7] - a[4] * a[6];
    float det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    float id = 1.0f / det;
    o[0] = c00 * id;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    o[3] = c01 * id;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    o[6] = c02 * id;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

}  // namespace cs
