// cs_motion.hip — keyframe motions and what they move: obstacles (src/Obstacle.cu:44-85,
// src/ObstacleHelper.cu:10-26) and handle anchors (src/ClothHelper.cu:3-14,
// src/PhysicsHelper.cu:379-396).
//
// Motions are evaluated on the host once per physics step (a few keyframe lookups,
// as the reference does in Simulator::updateTransformations, src/Simulator.cu:565-569)
// and reach the kernels by value in an XfTable, so a step needs no copy.  Obstacle
// velocities are written by one launch over every obstacle node (48 B per moving
// node: base and x read, v written); handle anchors by one launch per cloth with
// moving handles.
#include <cmath>

#include "cs_kernels.hpp"
#include "cs_math.hpp"
#include "cs_motion.hpp"

namespace cs {

namespace {

// Componentwise arithmetic of Transformation (src/Transformation.cu:24-72): scaling and the
// quaternion's scalar divide directly, vectors multiply by the reciprocal (src/Vector.cuh:126-132).
Xf add(const Xf& a, const Xf& b) {
    return Xf{a.sc + b.sc, a.tx + b.tx, a.ty + b.ty, a.tz + b.tz, a.qs + b.qs, a.qx + b.qx, a.qy + b.qy, a.qz + b.qz};
}
Xf sub(const Xf& a, const Xf& b) {
    return Xf{a.sc - b.sc, a.tx - b.tx, a.ty - b.ty, a.tz - b.tz, a.qs - b.qs, a.qx - b.qx, a.qy - b.qy, a.qz - b.qz};
}
Xf mul(float s, const Xf& a) {
    return Xf{s * a.sc, s * a.tx, s * a.ty, s * a.tz, s * a.qs, s * a.qx, s * a.qy, s * a.qz};
}
Xf divide(const Xf& a, float s) {
    const float inv = 1.0f / s;
    return Xf{a.sc / s, a.tx * inv, a.ty * inv, a.tz * inv, a.qs / s, a.qx * inv, a.qy * inv, a.qz * inv};
}

__global__ void __launch_bounds__(256) k_obstacle_motion(int nc, int n, const float4* __restrict__ base,
                                                         const int* __restrict__ motion, XfTable T, float invDt,
                                                         const float4* __restrict__ x, float4* __restrict__ v) {
    const int i = nc + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = motion[i - nc];
    if (m < 0) return;
    const float4 b = base[i - nc];
    float qx, qy, qz;
    xf_point(T.m[m], b.x, b.y, b.z, qx, qy, qz);
    const float4 p = x[i];
    v[i] = make_float4((qx - p.x) * invDt, (qy - p.y) * invDt, (qz - p.z) * invDt, 0.0f);
}

// Obstacle::transform (src/Obstacle.cu:44-64, src/ObstacleHelper.cu:10-17): x = x0 = T(base).
__global__ void __launch_bounds__(256) k_obstacle_place(int nc, int n, const float4* __restrict__ base,
                                                        const int* __restrict__ motion, XfTable T,
                                                        float4* __restrict__ x, float4* __restrict__ x0) {
    const int i = nc + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = motion[i - nc];
    if (m < 0) return;
    const float4 b = base[i - nc];
    float qx, qy, qz;
    xf_point(T.m[m], b.x, b.y, b.z, qx, qy, qz);
    const float4 p = make_float4(qx, qy, qz, 0.0f);
    x[i] = p;
    x0[i] = p;
}

// One thread per handled row: the sum of its handles' current anchors, in insertion order.
__global__ void __launch_bounds__(256) k_handle_anchors(int ng, const int* __restrict__ ptr,
                                                        const int* __restrict__ row,
                                                        const float4* __restrict__ anchor, XfTable T,
                                                        float4* __restrict__ hpos) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (int h = ptr[g]; h < ptr[g + 1]; h++) {
        const float4 a = anchor[h];
        const int m = __float_as_int(a.w);
        float px = a.x, py = a.y, pz = a.z;
        if (m >= 0) xf_point(T.m[m], a.x, a.y, a.z, px, py, pz);
        sx += px;
        sy += py;
        sz += pz;
    }
    hpos[row[g]] = make_float4(sx, sy, sz, 0.0f);
}

}  // namespace

int motion_build(const float* times, const cs_transform* keys, int n, Motion& M) {
    if (n < 1 || !times || !keys) return CS_E_ARG;
    for (int i = 0; i < n; i++) {
        if (!std::isfinite(times[i])) return CS_E_ARG;
        if (i > 0 && !(times[i] > times[i - 1])) return CS_E_ARG;
    }
    M.t.assign(times, times + n);
    M.x.resize(n);
    for (int i = 0; i < n; i++) M.x[i] = xf_from(keys[i]);
    if (M.t[0] > 0.0f) {  // identity key at t = 0 (src/Motion.cu:13-18)
        M.t.insert(M.t.begin(), 0.0f);
        M.x.insert(M.x.begin(), xf_identity());
    }
    const int m = (int)M.t.size();
    M.v.assign(m, mul(0.0f, xf_identity()));
    for (int i = 1; i < m - 1; i++) M.v[i] = divide(sub(M.x[i + 1], M.x[i - 1]), M.t[i + 1] - M.t[i - 1]);
    return CS_OK;
}

Xf motion_at(const Motion& M, float time) {
    const int n = (int)M.t.size();
    int i = 0;
    while (i < n - 1 && M.t[i + 1] < time) i++;
    if (i == n - 1) return M.x[n - 1];
    // Hermite coefficients of segment i (src/Motion.cu:45-51)
    const float dt = M.t[i + 1] - M.t[i];
    const Xf dx = sub(M.x[i + 1], M.x[i]);
    const Xf a = sub(mul(dt, add(M.v[i + 1], M.v[i])), mul(2.0f, dx));
    const Xf b = sub(mul(3.0f, dx), mul(dt, add(M.v[i + 1], mul(2.0f, M.v[i]))));
    const Xf c = mul(dt, M.v[i]);
    const Xf& d = M.x[i];
    const float s = (time - M.t[i]) / (M.t[i + 1] - M.t[i]);
    const float s2 = s * s;
    const float s3 = s * s2;
    return add(add(add(mul(s3, a), mul(s2, b)), mul(s, c)), d);
}

// Simulator::simulateStep's counter and updateTransformations, then Obstacle::step and the
// handle anchors of the new time (src/Simulator.cu:838-843,565-573).
void motion_step(cs_ctx* C) {
    C->steps++;
    if (C->motions.empty()) return;
    const float time = (float)C->steps * C->P.dt;
    XfTable T;
    for (size_t m = 0; m < C->motions.size(); m++) T.m[m] = motion_at(C->motions[m], time);
    if (C->anyMovingObstacle) {
        const int n = C->N - C->NC;
        hipLaunchKernelGGL(k_obstacle_motion, dim3(cdiv(n, 256)), dim3(256), 0, C->stream, C->NC, C->N,
                           (const float4*)C->obsBase.ptr<float4>(), (const int*)C->obsMotion.ptr<int>(), T,
                           1.0f / C->P.dt, (const float4*)C->x.ptr<float4>(), C->v.ptr<float4>());
        CS_LAUNCH_CHECK();
    }
    for (ClothSys& S : C->sys) {
        if (S.nHandleRows == 0) continue;
        hipLaunchKernelGGL(k_handle_anchors, dim3(cdiv(S.nHandleRows, 256)), dim3(256), 0, C->stream, S.nHandleRows,
                           (const int*)S.hgPtr.ptr<int>(), (const int*)S.hgRow.ptr<int>(),
                           (const float4*)S.hAnchor.ptr<float4>(), T, S.hpos.ptr<float4>());
        CS_LAUNCH_CHECK();
    }
}

void motion_place(cs_ctx* C) {
    if (!C->anyMovingObstacle) return;
    if (!C->finalized) {  // host state: re-place from the stored bases
        const int first = C->meshes[C->nCloths].n0;
        for (const MeshInfo& M : C->meshes) {
            if (M.cloth || M.motion < 0) continue;
            const Xf T = motion_now(C, M.motion);
            for (int i = M.n0; i < M.n0 + M.nn; i++) {
                const float* b = &C->hObsBase[3 * (size_t)(i - first)];
                float* x = &C->hx[3 * (size_t)i];
                xf_point(T, b[0], b[1], b[2], x[0], x[1], x[2]);
            }
        }
        return;
    }
    const float time = (float)C->steps * C->P.dt;
    XfTable T;
    for (size_t m = 0; m < C->motions.size(); m++) T.m[m] = motion_at(C->motions[m], time);
    const int n = C->N - C->NC;
    hipLaunchKernelGGL(k_obstacle_place, dim3(cdiv(n, 256)), dim3(256), 0, C->stream, C->NC, C->N,
                       (const float4*)C->obsBase.ptr<float4>(), (const int*)C->obsMotion.ptr<int>(), T,
                       C->x.ptr<float4>(), C->x0.ptr<float4>());
    CS_LAUNCH_CHECK();
    for (const MeshInfo& M : C->meshes) {
        if (M.cloth || M.motion < 0) continue;
        launch_face_geometry(C, M.f0, M.nf);
        launch_node_geometry(C, M.n0, M.nn);
    }
    CS_HIP(hipStreamSynchronize(C->stream));
}

Xf motion_now(const cs_ctx* C, int motion) { return motion_at(C->motions[motion], (float)C->steps * C->P.dt); }

}  // namespace cs

extern "C" cs_status cs_motion_eval(const float* times, const cs_transform* keys, int n, float time,
                                    cs_transform* out) {
    if (!out) return CS_E_ARG;
    cs::Motion M;
    const int st = cs::motion_build(times, keys, n, M);
    if (st != CS_OK) return (cs_status)st;
    *out = cs::xf_to(cs::motion_at(M, time));
    return CS_OK;
}
