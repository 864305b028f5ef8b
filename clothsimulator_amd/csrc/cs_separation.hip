// cs_separation.hip — the separation step after remeshing (Simulator::separationStep GPU branch,
// src/Simulator.cu:745-799, with findIntersections, src/Simulator.cu:531-562, and
// src/SeparationHelper.cu).
//
// Per iteration: candidate face pairs from the trees (non-CCD boxes, collision thickness), one
// thread per pair tests whether the two faces cross (checkIntersectionMidpoint,
// src/SeparationHelper.cu:7-70) and the crossing pairs are compacted at scan offsets.  Each new
// intersection gets the direction that separates the two points' pre-remeshing positions (the
// frame's backup faces hold them in material space) and the deepest vertex-face or edge-edge
// point along it (farthestPoint, src/SeparationHelper.cu:150-219); that host-side part touches a
// handful of records.  The accumulated constraints go to the augmented-Lagrangian solve shared
// with the impact response (separation_response in cs_response.hip), then the faces and trees
// are updated.  A pass that leaves no crossing pair ends the step; without deformable obstacles
// first, then with them (obstacle area halved per iteration), as the reference does.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

CS_HD int sgn(float x) { return x < 0.0f ? -1 : 1; }  // sign (src/MathHelper.cuh:47-49)

CS_HD int major_axis(V3f v) {  // src/SeparationHelper.cu:3-5
    const float a = tabs(v.x), b = tabs(v.y), c = tabs(v.z);
    return a > b && a > c ? 0 : (b > c ? 1 : 2);
}

CS_HD V3f face_point(const V3f* y, const float* b) { return (b[0] * y[0] + b[1] * y[1]) + b[2] * y[2]; }

// facePlaneIntersection (src/SeparationHelper.cu:7-35): the segment where `plane`'s plane cuts
// face y, as barycentric end points b0, b1 of y
CS_HD bool face_plane(const V3f* y, V3f p0, V3f n, float* b0, float* b1) {
    float h[3];
    int ss = 0;
    for (int i = 0; i < 3; i++) {
        h[i] = dot(y[i] - p0, n);
        ss += sgn(h[i]);
    }
    if (ss == -3 || ss == 3) return false;
    int v0 = -1;
    for (int i = 0; i < 3; i++)
        if (sgn(h[i]) == -ss) v0 = i;
    const int v1 = (v0 + 1) % 3, v2 = (v0 + 2) % 3;
    const float t0 = h[v0] / (h[v0] - h[v1]);
    const float t1 = h[v0] / (h[v0] - h[v2]);
    b0[v0] = 1.0f - t0;
    b0[v1] = t0;
    b0[v2] = 0.0f;
    b1[v0] = 1.0f - t1;
    b1[v2] = t1;
    b1[v1] = 0.0f;
    return true;
}

// checkIntersectionMidpoint (src/SeparationHelper.cu:37-70)
CS_HD bool intersect_midpoint(const int* na, const V3f* ya, V3f nA, const int* nb, const V3f* yb, V3f nB, float* b0,
                              float* b1) {
    for (int i = 0; i < 3; i++)  // Face::adjacent: a shared node
        for (int j = 0; j < 3; j++)
            if (na[i] == nb[j]) return false;
    const V3f c = cross(nA, nB);
    if (norm2(c) < 1e-12f) return false;
    float b00[3], b01[3], b10[3], b11[3];
    if (!face_plane(ya, yb[0], nB, b00, b01)) return false;
    if (!face_plane(yb, ya[0], nA, b10, b11)) return false;
    const int ax = major_axis(c);
    const float a00 = comp(face_point(ya, b00), ax), a01 = comp(face_point(ya, b01), ax);
    const float a10 = comp(face_point(yb, b10), ax), a11 = comp(face_point(yb, b11), ax);
    const float aMin = tmax(tmin(a00, a01), tmin(a10, a11));
    const float aMax = tmin(tmax(a00, a01), tmax(a10, a11));
    if (aMin > aMax) return false;
    const float aMid = 0.5f * (aMin + aMax);
    if (a00 == a01) {
        for (int k = 0; k < 3; k++) b0[k] = b00[k];
    } else {
        const float t = (aMid - a00) / (a01 - a00);
        for (int k = 0; k < 3; k++) b0[k] = b00[k] + t * (b01[k] - b00[k]);
    }
    if (a10 == a11) {
        for (int k = 0; k < 3; k++) b1[k] = b10[k];
    } else {
        const float t = (aMid - a10) / (a11 - a10);
        for (int k = 0; k < 3; k++) b1[k] = b10[k] + t * (b11[k] - b10[k]);
    }
    return true;
}

// out: flag[i] (crossing), rec[2i] = (b0, face0 bits), rec[2i+1] = (b1, face1 bits)
__global__ void __launch_bounds__(256) k_intersect(long n, const int2* __restrict__ pairs,
                                                   const int4* __restrict__ fnodes, const float4* __restrict__ x,
                                                   const float4* __restrict__ fnrm, uint32_t* __restrict__ flag,
                                                   float4* __restrict__ rec) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int2 p = pairs[i];
    const int4 fa = fnodes[p.x], fb = fnodes[p.y];
    const int na[3] = {fa.x, fa.y, fa.z}, nb[3] = {fb.x, fb.y, fb.z};
    V3f ya[3], yb[3];
    for (int k = 0; k < 3; k++) {
        ya[k] = ld3(x[na[k]]);
        yb[k] = ld3(x[nb[k]]);
    }
    float b0[3], b1[3];
    const bool hit = intersect_midpoint(na, ya, ld3(fnrm[p.x]), nb, yb, ld3(fnrm[p.y]), b0, b1);
    flag[i] = hit ? 1u : 0u;
    if (hit) {
        rec[2 * i] = make_float4(b0[0], b0[1], b0[2], __int_as_float(p.x));
        rec[2 * i + 1] = make_float4(b1[0], b1[1], b1[2], __int_as_float(p.y));
    }
}

__global__ void k_intersect_compact(long n, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                    const float4* __restrict__ rec, float4* __restrict__ out) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint32_t k = pos[i];
    out[2 * k] = rec[2 * i];
    out[2 * k + 1] = rec[2 * i + 1];
}

// ---------------------------------------------------------------- host side

struct HostGeo {  // current positions and face normals, downloaded once per iteration
    std::vector<float4> x, fn;
};

V3f hx(const HostGeo& g, int node) { return ld3(g.x[node]); }

// clearVertexFaceDistance (src/SeparationHelper.cu:150-182): vertices of face A behind face B
// along d
void clear_vf(const int* na, const int* nb, V3f nB, const HostGeo& g, V3f d, float& maxDist, float* bA, float* bB) {
    const V3f x0 = hx(g, nb[0]), x1 = hx(g, nb[1]), x2 = hx(g, nb[2]);
    for (int i = 0; i < 3; i++) {
        const V3f x = hx(g, na[i]);
        const float h = dot(x - x0, nB);
        const float dh = dot(d, nB);
        if (h * dh >= 0.0f) continue;
        const float a0 = mixed(x2 - x1, x - x1, d);
        const float a1 = mixed(x0 - x2, x - x2, d);
        const float a2 = mixed(x1 - x0, x - x0, d);
        if (a0 <= 0.0f || a1 <= 0.0f || a2 <= 0.0f) continue;
        const float dist = -h / dh;
        if (dist > maxDist) {
            maxDist = dist;
            bA[0] = bA[1] = bA[2] = 0.0f;
            bA[i] = 1.0f;
            bB[0] = a0 / (a0 + a1 + a2);
            bB[1] = a1 / (a0 + a1 + a2);
            bB[2] = a2 / (a0 + a1 + a2);
        }
    }
}

// clearEdgeEdgeDistance (src/SeparationHelper.cu:184-212)
void clear_ee(const int* na, const int* nb, const HostGeo& g, V3f d, float& maxDist, float* b0, float* b1) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const V3f x00 = hx(g, na[i]), x01 = hx(g, na[(i + 1) % 3]);
            const V3f x10 = hx(g, nb[j]), x11 = hx(g, nb[(j + 1) % 3]);
            const V3f n = cross(normalized(x01 - x00), normalized(x11 - x10));
            const float h = dot(x00 - x10, n);
            const float dh = dot(d, n);
            if (h * dh >= 0.0f) continue;
            const float a00 = mixed(x01 - x10, x11 - x10, d);
            const float a01 = mixed(x11 - x10, x00 - x10, d);
            const float a10 = mixed(x01 - x00, x11 - x00, d);
            const float a11 = mixed(x10 - x00, x01 - x00, d);
            if (a00 * a01 <= 0.0f || a10 * a11 <= 0.0f) continue;
            const float dist = -h / dh;
            if (dist > maxDist) {
                maxDist = dist;
                b0[0] = b0[1] = b0[2] = 0.0f;
                b0[i] = a00 / (a00 + a01);
                b0[(i + 1) % 3] = a01 / (a00 + a01);
                b1[0] = b1[1] = b1[2] = 0.0f;
                b1[j] = a10 / (a10 + a11);
                b1[(j + 1) % 3] = a11 / (a10 + a11);
            }
        }
}

// BackupFace::barycentricCoordinates / position (src/BackupFace.cu:7-15)
bool backup_position(const std::vector<BackupFace>& faces, float u0, float u1, V3f& out) {
    for (const BackupFace& f : faces) {
        const float d1x = f.u[0][0] - f.u[2][0], d1y = f.u[0][1] - f.u[2][1];
        const float d2x = f.u[1][0] - f.u[2][0], d2y = f.u[1][1] - f.u[2][1];
        const float invDet = 1.0f / (d1x * d2y - d2x * d1y);
        const float i00 = d2y * invDet, i01 = -d2x * invDet, i10 = -d1y * invDet, i11 = d1x * invDet;
        const float rx = u0 - f.u[2][0], ry = u1 - f.u[2][1];
        const float b0 = i00 * rx + i01 * ry, b1 = i10 * rx + i11 * ry;
        const float b2 = 1.0f - b0 - b1;
        if (b0 >= -1e-6f && b1 >= -1e-6f && b2 >= -1e-5f) {
            const V3f y0 = mk(f.x[0][0], f.x[0][1], f.x[0][2]), y1 = mk(f.x[1][0], f.x[1][1], f.x[1][2]);
            const V3f y2 = mk(f.x[2][0], f.x[2][1], f.x[2][2]);
            out = (b0 * y0 + b1 * y1) + b2 * y2;
            return true;
        }
    }
    return false;
}

}  // namespace

// The position a face point had before remeshing (Simulator::oldPosition, src/Simulator.cu:489-505):
// obstacle faces are not remeshed, so their current position; a cloth face point goes through
// its material-space coordinate to the backup face that encloses it (the first in face order).
static V3f old_position(const cs_ctx* C, const std::vector<std::vector<BackupFace>>& backup, const HostGeo& g, int f,
                        const float* b) {
    const int na[3] = {C->hfn[3 * f], C->hfn[3 * f + 1], C->hfn[3 * f + 2]};
    V3f y[3] = {hx(g, na[0]), hx(g, na[1]), hx(g, na[2])};
    const V3f here = face_point(y, b);
    if (f >= C->FC) return here;
    int c = 0;
    while (c + 1 < C->nCloths && f >= C->meshes[c + 1].f0) c++;
    float u0 = 0.0f, u1 = 0.0f;
    for (int k = 0; k < 3; k++) {  // b0 u0 + b1 u1 + b2 u2, left to right
        const int vt = C->hfv[3 * f + k];
        u0 = u0 + b[k] * C->huv[2 * vt];
        u1 = u1 + b[k] * C->huv[2 * vt + 1];
    }
    V3f out;
    if ((size_t)c < backup.size() && backup_position(backup[c], u0, u1, out)) return out;
    return here;  // no enclosing backup face (the reference leaves this undefined)
}

int separation_step(cs_ctx* C, const std::vector<std::vector<BackupFace>>& backup) {
    hipStream_t s = C->stream;
    const float th = C->P.collision_thickness;
    float obstacleArea = 1e3f;
    bool success = false;
    int deform = 0;
    std::vector<SepRec> cons;
    HostGeo g;
    bvh_build_all(C, false);
    for (deform = 0; deform < 2; deform++) {
        cons.clear();
        for (int it = 0; it < 100; it++) {  // MAX_SEPARATION_ITERATION (src/SeparationHelper.cuh:15)
            const long np = bvh_all_pairs(C, th);
            int nNew = 0;
            float4* rec = nullptr;
            if (np > 0) {
                uint32_t* flag = C->sepFlag.as<uint32_t>(np + 1);
                uint32_t* pos = C->sepPos.as<uint32_t>(np + 1);
                float4* raw = C->sepRaw.as<float4>(2 * np);
                hipLaunchKernelGGL(k_intersect, dim3(cdiv(np, 256)), dim3(256), 0, s, np, C->pairs.ptr<int2>(),
                                   C->fnodes.ptr<int4>(), C->x.ptr<float4>(), C->fnrm.ptr<float4>(), flag, raw);
                CS_LAUNCH_CHECK();
                scan_exclusive(flag, pos, (int)np, C->ws, s);
                uint32_t total = 0;
                CS_HIP(hipMemcpyAsync(&total, pos + np, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                CS_HIP(hipStreamSynchronize(s));
                nNew = (int)total;
                if (nNew > 0) {
                    rec = C->sepRec.as<float4>(2 * (size_t)nNew);
                    hipLaunchKernelGGL(k_intersect_compact, dim3(cdiv(np, 256)), dim3(256), 0, s, np, flag, pos, raw,
                                       rec);
                    CS_LAUNCH_CHECK();
                }
            }
            if (nNew == 0) {
                success = true;
                break;
            }
            std::vector<float4> h(2 * (size_t)nNew);
            g.x.resize(C->N);
            g.fn.resize(C->F);
            CS_HIP(hipMemcpyAsync(h.data(), rec, sizeof(float4) * h.size(), hipMemcpyDeviceToHost, s));
            CS_HIP(hipMemcpyAsync(g.x.data(), C->x.p, sizeof(float4) * C->N, hipMemcpyDeviceToHost, s));
            CS_HIP(hipMemcpyAsync(g.fn.data(), C->fnrm.p, sizeof(float4) * C->F, hipMemcpyDeviceToHost, s));
            CS_HIP(hipStreamSynchronize(s));
            for (int k = 0; k < nNew; k++) {
                const float4 r0 = h[2 * k], r1 = h[2 * k + 1];
                int f0, f1;
                std::memcpy(&f0, &r0.w, sizeof(int));
                std::memcpy(&f1, &r1.w, sizeof(int));
                float b0[3] = {r0.x, r0.y, r0.z}, b1[3] = {r1.x, r1.y, r1.z};
                const V3f p0 = old_position(C, backup, g, f0, b0), p1 = old_position(C, backup, g, f1, b1);
                const V3f d = normalized(p0 - p1);
                const int na[3] = {C->hfn[3 * f0], C->hfn[3 * f0 + 1], C->hfn[3 * f0 + 2]};
                const int nb[3] = {C->hfn[3 * f1], C->hfn[3 * f1 + 1], C->hfn[3 * f1 + 2]};
                float maxDist = 0.0f;  // farthestPoint (src/SeparationHelper.cu:214-219)
                clear_vf(na, nb, ld3(g.fn[f1]), g, d, maxDist, b0, b1);
                clear_vf(nb, na, ld3(g.fn[f0]), g, -d, maxDist, b1, b0);
                clear_ee(na, nb, g, d, maxDist, b0, b1);
                SepRec R;
                for (int j = 0; j < 3; j++) {
                    R.nodes[j] = na[j];
                    R.nodes[3 + j] = nb[j];
                    R.w[j] = b0[j];
                    R.w[3 + j] = -b1[j];
                }
                R.n[0] = d.x;
                R.n[1] = d.y;
                R.n[2] = d.z;
                R.pad = 0.0f;
                cons.push_back(R);
            }
            SepRec* dcons = C->sepCons.as<SepRec>(cons.size());
            CS_HIP(hipMemcpyAsync(dcons, cons.data(), sizeof(SepRec) * cons.size(), hipMemcpyHostToDevice, s));
            separation_response(C, dcons, (int)cons.size(), deform, obstacleArea);
            launch_face_geometry(C, 0, C->FC);
            bvh_refit_range(C, 0, C->nCloths);
            if (deform == 1) {
                launch_face_geometry(C, C->FC, C->F - C->FC);
                bvh_refit_range(C, C->nCloths, (int)C->meshes.size());
                obstacleArea *= 0.5f;
            }
            C->sepIters++;
        }
        if (success) break;
    }
    if (!success) {
        C->err = "Separation step failed: intersections remain after 100 iterations in both deform passes";
        return CS_E_COLLISION_FAILED;
    }
    launch_node_geometry(C, 0, C->NC);
    launch_update_velocities(C, 0, C->NC);
    if (deform == 1) {
        launch_node_geometry(C, C->NC, C->N - C->NC);
        launch_update_velocities(C, C->NC, C->N - C->NC);
    }
    CS_HIP(hipStreamSynchronize(s));
    return CS_OK;
}

}  // namespace cs
