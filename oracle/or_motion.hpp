// oracle/or_motion.hpp — CPU restatement of the reference's rigid transformations and
// keyframe motions (TEST INFRASTRUCTURE ONLY, see or_math.hpp).
//
// A transformation is 8 floats: scaling, translation (3), rotation quaternion (s, v)
// (src/Transformation.cuh:12-16).  Keyframe arithmetic is FP32 as in the reference;
// applying a transformation to a point runs in the scene's precision T.
// Pinned in tests/test_oracle_motion.py by closed forms: the two-key Hermite is
// x0 + (x1 - x0)(3s^2 - 2s^3); interpolation hits every key; the interior tangent
// equals the central difference; rotations by unit quaternions match
// scipy.spatial.transform.Rotation; T^-1(T(p)) = p; composition against rotations composed
// by scipy (application in turn for unit scales).
#pragma once
#include <array>
#include <vector>

#include "or_math.hpp"

namespace orc {

using Tf = std::array<float, 8>;

inline Tf tf_identity() { return Tf{1.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f, 0.0f}; }

// src/Transformation.cu:24-72 (componentwise; scaling and the quaternion's scalar divide,
// the vector parts multiply by the reciprocal, src/Vector.cuh:126-132, src/Quaternion.cu:58-64)
inline Tf tf_add(const Tf& a, const Tf& b) { Tf r; for (int i = 0; i < 8; i++) r[i] = a[i] + b[i]; return r; }
inline Tf tf_sub(const Tf& a, const Tf& b) { Tf r; for (int i = 0; i < 8; i++) r[i] = a[i] - b[i]; return r; }
inline Tf tf_scale(float s, const Tf& a) { Tf r; for (int i = 0; i < 8; i++) r[i] = s * a[i]; return r; }
inline Tf tf_div(const Tf& a, float s) {
    float inv = 1.0f / s;
    Tf r;
    r[0] = a[0] / s;
    for (int i = 1; i < 4; i++) r[i] = a[i] * inv;
    r[4] = a[4] / s;
    for (int i = 5; i < 8; i++) r[i] = a[i] * inv;
    return r;
}

// Quaternion::rotate (src/Quaternion.cu:73-75) in precision T
template <class T> V3<T> tf_rotate(const Tf& X, const V3<T>& p) {
    T s = T(X[4]);
    V3<T> q = mk3<T>(T(X[5]), T(X[6]), T(X[7]));
    return p * (s * s - q.norm2()) + T(2) * q * q.dot(p) + T(2) * cross(q, p) * s;
}
// Transformation::applyToPoint (src/Transformation.cu:82-84)
template <class T> V3<T> tf_point(const Tf& X, const V3<T>& p) {
    return T(X[0]) * tf_rotate(X, p) + mk3<T>(T(X[1]), T(X[2]), T(X[3]));
}

// Transformation::inverse (src/Transformation.cu:74-80), Quaternion::inverse (src/Quaternion.cu:66-71)
inline Tf tf_inverse(const Tf& X) {
    Tf r;
    r[0] = 1.0f / X[0];
    V3<float> q = mk3<float>(X[5], X[6], X[7]);
    float n2 = X[4] * X[4] + q.norm2();
    r[4] = X[4] / n2;
    V3<float> qi = (-q) / n2;
    r[5] = qi[0]; r[6] = qi[1]; r[7] = qi[2];
    V3<float> t = V3<float>() - tf_rotate(r, r[0] * mk3<float>(X[1], X[2], X[3]));
    r[1] = t[0]; r[2] = t[1]; r[3] = t[2];
    return r;
}

// Transformation::operator*(Transformation) (src/Transformation.cu:58-64), quaternion
// product (src/Quaternion.cu:51-56)
inline Tf tf_compose(const Tf& a, const Tf& b) {
    Tf r;
    r[0] = a[0] * b[0];
    V3<float> t = mk3<float>(a[1], a[2], a[3]) + tf_rotate(a, b[0] * mk3<float>(b[1], b[2], b[3]));
    r[1] = t[0]; r[2] = t[1]; r[3] = t[2];
    V3<float> av = mk3<float>(a[5], a[6], a[7]), bv = mk3<float>(b[5], b[6], b[7]);
    r[4] = a[4] * b[4] - av.dot(bv);
    V3<float> v = a[4] * bv + b[4] * av + cross(av, bv);
    r[5] = v[0]; r[6] = v[1]; r[7] = v[2];
    return r;
}

// Motion (src/Motion.cu:3-20,39-72)
struct Motion {
    std::vector<float> t;
    std::vector<Tf> x, v;

    bool build(const float* times, const Tf* keys, int n) {
        if (n < 1) return false;
        for (int i = 1; i < n; i++)
            if (!(times[i] > times[i - 1])) return false;
        t.assign(times, times + n);
        x.assign(keys, keys + n);
        if (t[0] > 0.0f) {
            t.insert(t.begin(), 0.0f);
            x.insert(x.begin(), tf_identity());
        }
        int m = (int)t.size();
        v.assign(m, tf_scale(0.0f, tf_identity()));
        for (int i = 1; i + 1 < m; i++) v[i] = tf_div(tf_sub(x[i + 1], x[i - 1]), t[i + 1] - t[i - 1]);
        return true;
    }

    Tf at(float time) const {
        int n = (int)t.size(), k = 0;
        while (k < n - 1 && t[k + 1] < time) k++;
        if (k == n - 1) return x[n - 1];
        float h = t[k + 1] - t[k];
        Tf dx = tf_sub(x[k + 1], x[k]);
        Tf a = tf_sub(tf_scale(h, tf_add(v[k + 1], v[k])), tf_scale(2.0f, dx));
        Tf b = tf_sub(tf_scale(3.0f, dx), tf_scale(h, tf_add(v[k + 1], tf_scale(2.0f, v[k]))));
        Tf c = tf_scale(h, v[k]);
        float s = (time - t[k]) / (t[k + 1] - t[k]);
        float s2 = s * s, s3 = s * s2;
        return tf_add(tf_add(tf_add(tf_scale(s3, a), tf_scale(s2, b)), tf_scale(s, c)), x[k]);
    }
};

}  // namespace orc
