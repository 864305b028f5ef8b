// cs_motion.hpp — rigid transformations and keyframe motions (src/Transformation.cu,
// src/Quaternion.cu, src/Motion.cu).  The component arithmetic follows the reference
// operation by operation (scalar division by a reciprocal for vectors, quaternions not
// renormalised), so host and device evaluate bit-identical transforms: every
// translation unit is compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/clothsim.h"

namespace cs {

// Transformation = (scaling, translation, rotation quaternion (s, v)); the cs_transform layout.
struct Xf {
    float sc, tx, ty, tz, qs, qx, qy, qz;
};

__host__ __device__ inline Xf xf_identity() { return Xf{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f}; }

// Quaternion::rotate (src/Quaternion.cu:73-75):
//   x (s^2 - |v|^2) + (2 v)(v.x) + (2 (v x x)) s
__host__ __device__ inline void xf_rotate(const Xf& T, float x, float y, float z, float& ox, float& oy, float& oz) {
    const float n2 = T.qx * T.qx + T.qy * T.qy + T.qz * T.qz;
    const float c0 = T.qs * T.qs - n2;
    const float d = T.qx * x + T.qy * y + T.qz * z;
    const float cx = T.qy * z - T.qz * y, cy = T.qz * x - T.qx * z, cz = T.qx * y - T.qy * x;
    ox = (x * c0 + (2.0f * T.qx) * d) + (2.0f * cx) * T.qs;
    oy = (y * c0 + (2.0f * T.qy) * d) + (2.0f * cy) * T.qs;
    oz = (z * c0 + (2.0f * T.qz) * d) + (2.0f * cz) * T.qs;
}

// Transformation::applyToPoint (src/Transformation.cu:82-84): scaling * rotate(p) + translation
__host__ __device__ inline void xf_point(const Xf& T, float x, float y, float z, float& ox, float& oy, float& oz) {
    float rx, ry, rz;
    xf_rotate(T, x, y, z, rx, ry, rz);
    ox = T.sc * rx + T.tx;
    oy = T.sc * ry + T.ty;
    oz = T.sc * rz + T.tz;
}

// Transformation::inverse (src/Transformation.cu:74-80) with Quaternion::inverse (src/Quaternion.cu:66-71)
inline Xf xf_inverse(const Xf& T) {
    Xf r;
    r.sc = 1.0f / T.sc;
    const float n2 = T.qs * T.qs + (T.qx * T.qx + T.qy * T.qy + T.qz * T.qz);
    const float inv = 1.0f / n2;
    r.qs = T.qs / n2;
    r.qx = -T.qx * inv;
    r.qy = -T.qy * inv;
    r.qz = -T.qz * inv;
    float rx, ry, rz;
    xf_rotate(r, r.sc * T.tx, r.sc * T.ty, r.sc * T.tz, rx, ry, rz);
    r.tx = 0.0f - rx;
    r.ty = 0.0f - ry;
    r.tz = 0.0f - rz;
    return r;
}

// Transformation::operator*(Transformation) (src/Transformation.cu:58-64) with the quaternion
// product (src/Quaternion.cu:51-56): the rigid map "apply b, then a".
inline Xf xf_compose(const Xf& a, const Xf& b) {
    Xf r;
    r.sc = a.sc * b.sc;
    float rx, ry, rz;
    xf_rotate(a, b.sc * b.tx, b.sc * b.ty, b.sc * b.tz, rx, ry, rz);
    r.tx = a.tx + rx;
    r.ty = a.ty + ry;
    r.tz = a.tz + rz;
    r.qs = a.qs * b.qs - (a.qx * b.qx + a.qy * b.qy + a.qz * b.qz);
    r.qx = (a.qs * b.qx + b.qs * a.qx) + (a.qy * b.qz - a.qz * b.qy);
    r.qy = (a.qs * b.qy + b.qs * a.qy) + (a.qz * b.qx - a.qx * b.qz);
    r.qz = (a.qs * b.qz + b.qs * a.qz) + (a.qx * b.qy - a.qy * b.qx);
    return r;
}

inline Xf xf_from(const cs_transform& t) {
    return Xf{t.scale, t.translate[0], t.translate[1], t.translate[2], t.rotate[0], t.rotate[1], t.rotate[2], t.rotate[3]};
}
inline cs_transform xf_to(const Xf& x) {
    cs_transform t;
    t.scale = x.sc;
    t.translate[0] = x.tx; t.translate[1] = x.ty; t.translate[2] = x.tz;
    t.rotate[0] = x.qs; t.rotate[1] = x.qx; t.rotate[2] = x.qy; t.rotate[3] = x.qz;
    return t;
}

// Keyframes of one motion with their Hermite tangents (src/Motion.cu:3-20,45-51).
struct Motion {
    std::vector<float> t;
    std::vector<Xf> x, v;
};

// Builds a motion from n keys (CS_E_ARG unless n >= 1, times finite and strictly increasing).
int motion_build(const float* times, const cs_transform* keys, int n, Motion& M);
// Motion::computeTransformation (src/Motion.cu:55-72) at `time`.
Xf motion_at(const Motion& M, float time);

// The transformations of every motion of a context at one time, passed to kernels by value.
struct XfTable {
    Xf m[CS_MAX_MOTIONS];
};

}  // namespace cs
