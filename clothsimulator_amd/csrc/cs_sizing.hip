// cs_sizing.hip — the sizing field of the remeshing step on the device.
//
//   k_near_planes   per cloth node, the nearest obstacle point closer than `thickness`
//                   (Cloth::findNearestPlane, src/Cloth.cu:200-233, checkNearestPoint /
//                   setNearestPlane, src/RemeshingHelper.cu:17-39): obstacle faces streamed
//                   through LDS in tiles of 256, every thread tests its node against the tile.
//   k_face_sizing   per cloth face, the sizing tensor (faceSizing, src/RemeshingHelper.cu:41-186)
//                   from the face's uv, its nodes' x, v and normals, its edges' dihedral angles
//                   and the nodes' planes.
//
// Both are FP32 in the reference's operation order.  The vertex average (Cloth::computeSizing,
// src/Cloth.cu:235-256) runs on the host in face order, where the operators need it.
#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

struct M2f {  // row-major 2x2
    float a, b, c, d;
};
CS_HD M2f m2(float a, float b, float c, float d) { M2f m; m.a = a; m.b = b; m.c = c; m.d = d; return m; }
CS_HD M2f operator+(M2f p, M2f q) { return m2(p.a + q.a, p.b + q.b, p.c + q.c, p.d + q.d); }
CS_HD M2f operator-(M2f p, M2f q) { return m2(p.a - q.a, p.b - q.b, p.c - q.c, p.d - q.d); }
CS_HD M2f operator*(float s, M2f p) { return m2(s * p.a, s * p.b, s * p.c, s * p.d); }
CS_HD M2f mdiv(M2f p, float s) { float i = 1.0f / s; return m2(p.a * i, p.b * i, p.c * i, p.d * i); }
CS_HD M2f mmul(M2f p, M2f q) {
    return m2(p.a * q.a + p.b * q.c, p.a * q.b + p.b * q.d, p.c * q.a + p.d * q.c, p.c * q.b + p.d * q.d);
}
CS_HD M2f mt(M2f p) { return m2(p.a, p.c, p.b, p.d); }
CS_HD M2f outer2(float x, float y) { return m2(x * x, x * y, y * x, y * y); }

// 3x2 (columns c0, c1) as three rows
struct M32 {
    float r[3][2];
};
// A^T A of a 3x2 matrix
CS_HD M2f ata(const M32& A) {
    M2f m = m2(0, 0, 0, 0);
    for (int i = 0; i < 3; i++) {
        m.a += A.r[i][0] * A.r[i][0];
        m.b += A.r[i][0] * A.r[i][1];
        m.c += A.r[i][1] * A.r[i][0];
        m.d += A.r[i][1] * A.r[i][1];
    }
    return m;
}
// Face::derivative: [q1 - q0, q2 - q0] (3x2) times the rest inverse (src/Face.cu:109-111)
CS_HD M32 derivative(V3f q0, V3f q1, V3f q2, M2f inv) {
    const V3f d1 = q1 - q0, d2 = q2 - q0;
    M32 D;
    const float c0[3] = {d1.x, d1.y, d1.z}, c1[3] = {d2.x, d2.y, d2.z};
    for (int i = 0; i < 3; i++) {
        D.r[i][0] = c0[i] * inv.a + c1[i] * inv.c;
        D.r[i][1] = c0[i] * inv.b + c1[i] * inv.d;
    }
    return D;
}

// eigenvalueDecomposition (src/MathHelper.cu:3-36): A symmetric, columns of Q the eigenvectors
CS_HD void eig(M2f A, M2f& Q, float l[2]) {
    const float a = A.a, b = A.c, d = A.d;
    const float amd = a - d, apd = a + d, b2 = b * b;
    const float det = sqrtf(4.0f * b2 + amd * amd);
    l[0] = 0.5f * (apd + det);
    l[1] = 0.5f * (apd - det);
    if (b != 0.0f) {
        float v0 = l[0] - d, vn = sqrtf(v0 * v0 + b2);
        Q.a = v0 / vn;
        Q.c = b / vn;
        v0 = l[1] - d;
        vn = sqrtf(v0 * v0 + b2);
        Q.b = v0 / vn;
        Q.d = b / vn;
    } else if (a >= d) {
        Q = m2(1.0f, 0.0f, 0.0f, 1.0f);
    } else {
        Q = m2(0.0f, 1.0f, 1.0f, 0.0f);
    }
}
// Q diag(l) Q^T
CS_HD M2f qlq(M2f Q, const float l[2]) { return mmul(mmul(Q, m2(l[0], 0.0f, 0.0f, l[1])), mt(Q)); }

CS_HD M2f msqrt(M2f A) {  // src/RemeshingHelper.cu:48-55
    M2f Q;
    float l[2];
    eig(A, Q, l);
    for (int i = 0; i < 2; i++) l[i] = l[i] >= 0.0f ? sqrtf(l[i]) : -sqrtf(-l[i]);
    return qlq(Q, l);
}

CS_HD M2f mmax(M2f A, float v) {  // src/RemeshingHelper.cu:57-64
    M2f Q;
    float l[2];
    eig(A, Q, l);
    for (int i = 0; i < 2; i++) l[i] = tmax(l[i], v);
    return qlq(Q, l);
}

struct Disk {
    float ox, oy, r;
};
CS_HD bool encloses(Disk p, Disk q) {  // Disk::enclose (src/Disk.cu:13-15)
    const float dx = p.ox - q.ox, dy = p.oy - q.oy;
    return p.r >= q.r && tsqr(p.r - q.r) >= dx * dx + dy * dy;
}
CS_HD Disk circ2(Disk d0, Disk d1) {  // src/RemeshingHelper.cu:87-92
    const float dx = d0.ox - d1.ox, dy = d0.oy - d1.oy;
    const float d = sqrtf(dx * dx + dy * dy);
    const float r = 0.5f * (d0.r + d + d1.r);
    const float t = (r - d0.r) / d;
    Disk o;
    o.ox = d0.ox + t * (d1.ox - d0.ox);
    o.oy = d0.oy + t * (d1.oy - d0.oy);
    o.r = r;
    return o;
}
CS_HD Disk circ3(Disk d0, Disk d1, Disk d2) {  // the disk tangent to three (src/RemeshingHelper.cu:94-126)
    const float x0 = d0.ox, y0 = d0.oy, r0 = d0.r, x1 = d1.ox, y1 = d1.oy, r1 = d1.r, x2 = d2.ox, y2 = d2.oy, r2 = d2.r;
    const float v11 = 2.0f * x1 - 2.0f * x0, v12 = 2.0f * y1 - 2.0f * y0;
    const float v13 = x0 * x0 - x1 * x1 + y0 * y0 - y1 * y1 - r0 * r0 + r1 * r1;
    const float v14 = 2.0f * r1 - 2.0f * r0;
    const float v21 = 2.0f * x2 - 2.0f * x1, v22 = 2.0f * y2 - 2.0f * y1;
    const float v23 = x1 * x1 - x2 * x2 + y1 * y1 - y2 * y2 - r1 * r1 + r2 * r2;
    const float v24 = 2.0f * r2 - 2.0f * r1;
    const float w12 = v12 / v11, w13 = v13 / v11, w14 = v14 / v11;
    const float w22 = v22 / v21 - w12, w23 = v23 / v21 - w13, w24 = v24 / v21 - w14;
    const float P = -w23 / w22, Q = w24 / w22;
    const float M = -w12 * P - w13, N = w14 - w12 * Q;
    const float a = N * N + Q * Q - 1.0f;
    const float b = 2.0f * M * N - 2.0f * N * x0 + 2.0f * P * Q - 2.0f * Q * y0 + 2.0f * r0;
    const float c = x0 * x0 + M * M - 2.0f * M * x0 + P * P + y0 * y0 - 2.0f * P * y0 - r0 * r0;
    const float D = b * b - 4.0f * a * c;
    const float rs = (-b - sqrtf(D)) / (2.0f * a);
    Disk o;
    o.ox = M + N * rs;
    o.oy = P + Q * rs;
    o.r = rs;
    return o;
}

// maxTensor (src/RemeshingHelper.cu:128-157)
CS_HD M2f max_tensor(const M2f* M) {
    Disk d[5];
    int n = 0;
    for (int i = 0; i < 5; i++)
        if (M[i].a + M[i].d != 0.0f) {
            d[n].ox = 0.5f * (M[i].a - M[i].d);
            d[n].oy = 0.5f * (M[i].b + M[i].c);
            d[n].r = 0.5f * (M[i].a + M[i].d);
            n++;
        }
    Disk k;
    if (n == 0) {
        k.ox = k.oy = k.r = 0.0f;
    } else {
        k = d[0];
    }
    for (int i = 1; i < n; i++)
        if (!encloses(k, d[i])) {
            k = d[i];
            for (int j = 0; j < i; j++)
                if (!encloses(k, d[j])) {
                    k = circ2(d[i], d[j]);
                    for (int q = 0; q < j; q++)
                        if (!encloses(k, d[q])) k = circ3(d[i], d[j], d[q]);
                }
        }
    return m2(k.r + k.ox, k.oy, k.oy, k.r - k.ox);
}

// unsignedVertexEdgeDistance (src/MathHelper.cu:71-80), its weights written through
CS_HD float ve_dist(V3f x, V3f y0, V3f y1, float& wy0, float& wy1) {
    const V3f e = y1 - y0;
    const float t = tclamp(dot(x - y0, e) / dot(e, e), 0.0f, 1.0f);
    const V3f y = y0 + t * e;
    wy0 = 1.0f - t;
    wy1 = t;
    return norm(x - y);
}

// unsignedVertexFaceDistance (src/MathHelper.cu:82-119): returns d; w[1..3] as the reference
// leaves them (the edge branches write their two weights whether or not they are closer)
CS_HD float vf_dist(V3f x, V3f y0, V3f y1, V3f y2, float w[4]) {
    const V3f nt = normalized(cross(y1 - y0, y2 - y0));
    float d = tabs(dot(x - y0, nt));
    const float b0 = mixed(y1 - x, y2 - x, nt);
    const float b1 = mixed(y2 - x, y0 - x, nt);
    const float b2 = mixed(y0 - x, y1 - x, nt);
    if (b0 >= 0.0f && b1 >= 0.0f && b2 >= 0.0f) {
        w[1] = -b0 / (b0 + b1 + b2);
        w[2] = -b1 / (b0 + b1 + b2);
        w[3] = -b2 / (b0 + b1 + b2);
        return d;
    }
    d = INFINITY;
    if (b0 < 0.0f) {
        const float dt = ve_dist(x, y1, y2, w[2], w[3]);
        if (dt < d) { d = dt; w[1] = 0.0f; }
    }
    if (b1 < 0.0f) {
        const float dt = ve_dist(x, y2, y0, w[3], w[1]);
        if (dt < d) { d = dt; w[2] = 0.0f; }
    }
    if (b2 < 0.0f) {
        const float dt = ve_dist(x, y0, y1, w[1], w[2]);
        if (dt < d) { d = dt; w[3] = 0.0f; }
    }
    return d;
}

constexpr int TILE = 256;

// planes: float4 p, float4 n per cloth node (n = 0: no plane)
__global__ void __launch_bounds__(256) k_near_planes(int n0, int nn, const float4* __restrict__ x,
                                                     const int4* __restrict__ fnodes, int fo0, int nfo,
                                                     float thickness, float4* __restrict__ planes) {
    __shared__ float ty[TILE][9];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool live = i < nn;
    const V3f xi = live ? ld3(x[n0 + i]) : mk(0.0f, 0.0f, 0.0f);
    float best = thickness;
    V3f pt = xi;
    for (int t0 = 0; t0 < nfo; t0 += TILE) {
        __syncthreads();
        const int f = t0 + (int)threadIdx.x;
        if (f < nfo) {
            const int4 fn = fnodes[fo0 + f];
            const float4 a = x[fn.x], b = x[fn.y], c = x[fn.z];
            ty[threadIdx.x][0] = a.x; ty[threadIdx.x][1] = a.y; ty[threadIdx.x][2] = a.z;
            ty[threadIdx.x][3] = b.x; ty[threadIdx.x][4] = b.y; ty[threadIdx.x][5] = b.z;
            ty[threadIdx.x][6] = c.x; ty[threadIdx.x][7] = c.y; ty[threadIdx.x][8] = c.z;
        }
        __syncthreads();
        const int m = tmin(TILE, nfo - t0);
        if (live)
            for (int k = 0; k < m; k++) {
                const V3f y0 = mk(ty[k][0], ty[k][1], ty[k][2]);
                const V3f y1 = mk(ty[k][3], ty[k][4], ty[k][5]);
                const V3f y2 = mk(ty[k][6], ty[k][7], ty[k][8]);
                float w[4] = {1.0f, 0.0f, 0.0f, 0.0f};
                const float d = vf_dist(xi, y0, y1, y2, w);
                if (d < best) {  // checkNearestPoint: x = -(w1 y0 + w2 y1 + w3 y2)
                    best = d;
                    pt = -((w[1] * y0 + w[2] * y1) + w[3] * y2);
                }
            }
    }
    if (!live) return;
    const V3f n = xi - pt;
    if (norm2(n) > 1e-8f) {
        planes[2 * i] = st4(pt);
        planes[2 * i + 1] = st4(normalized(n));
    } else {
        planes[2 * i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        planes[2 * i + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

struct SizingArgs {
    int f0, nf, n0;  // the cloth's faces and first node (global ids)
    const float* fuv;  // 6 per cloth face (local)
    const int4 *fnodes, *fedges;
    const float4 *finv, *fnrm, *x, *v, *nrm, *planes;
    const float* farea;
    const int2 *enodes, *eadj;
    float refineAngle, refineVelocity, refineCompression, ribStiffening, sizeMin, sizeMax, aspectMin;
    float* out;  // 4 per face
};

// Edge::angle (src/Edge.cu:114-124)
__device__ float edge_angle(const SizingArgs& A, int e) {
    const int2 ad = A.eadj[e];
    if (ad.x < 0 || ad.y < 0) return 0.0f;
    const int2 en = A.enodes[e];
    const V3f t = normalized(ld3(A.x[en.x]) - ld3(A.x[en.y]));
    const V3f n0 = ld3(A.fnrm[ad.x]), n1 = ld3(A.fnrm[ad.y]);
    return atan2f(dot(t, cross(n0, n1)), dot(n0, n1));
}

__global__ void __launch_bounds__(256) k_face_sizing(SizingArgs A) {
    const int lf = blockIdx.x * 256 + threadIdx.x;
    if (lf >= A.nf) return;
    const int f = A.f0 + lf;
    const int4 fn = A.fnodes[f], fe = A.fedges[f];
    const float4 iv = A.finv[f];
    const M2f inv = m2(iv.x, iv.y, iv.z, iv.w);
    const float area = A.farea[f];
    const int nd[3] = {fn.x, fn.y, fn.z}, ed[3] = {fe.x, fe.y, fe.z};
    M2f M[5];
    // Face::curvature (src/Face.cu:113-122)
    M2f Sw1 = m2(0, 0, 0, 0);
    for (int i = 0; i < 3; i++) {
        const int j = (i + 1) % 3;
        const float ex = A.fuv[6 * lf + 2 * j] - A.fuv[6 * lf + 2 * i];
        const float ey = A.fuv[6 * lf + 2 * j + 1] - A.fuv[6 * lf + 2 * i + 1];
        const float en = sqrtf(ex * ex + ey * ey);
        const float ie = 1.0f / en;
        const float tx = -ey * ie, ty = ex * ie;
        Sw1 = Sw1 - (0.5f * edge_angle(A, ed[i]) * en) * outer2(tx, ty);
    }
    Sw1 = mdiv(Sw1, area);
    const float ra2 = tsqr(A.refineAngle);
    M[0] = mdiv(mmul(mt(Sw1), Sw1), ra2);
    const M32 Sw2 = derivative(ld3(A.nrm[nd[0]]), ld3(A.nrm[nd[1]]), ld3(A.nrm[nd[2]]), inv);
    const M2f S2 = ata(Sw2);
    M[1] = mdiv(S2, ra2);
    const M32 V = derivative(ld3(A.v[nd[0]]), ld3(A.v[nd[1]]), ld3(A.v[nd[2]]), inv);
    M[2] = mdiv(ata(V), tsqr(A.refineVelocity));
    const M32 Fd = derivative(ld3(A.x[nd[0]]), ld3(A.x[nd[1]]), ld3(A.x[nd[2]]), inv);
    const M2f G = ata(Fd) - m2(1.0f, 0.0f, 0.0f, 1.0f);
    {  // compressionMetric (src/RemeshingHelper.cu:66-70)
        const M2f P = m2(S2.d, -S2.b, -S2.c, S2.a);
        const float c2 = tsqr(A.refineCompression);
        const M2f T = (4.0f * c2) * P;
        const M2f D = mmul(mt(G), G) - m2(T.a * A.ribStiffening, T.b * A.ribStiffening, T.c * A.ribStiffening,
                                          T.d * A.ribStiffening);
        M[3] = mdiv(mmax(msqrt(D) - G, 0.0f), 2.0f * c2);
    }
    {  // obstacleMetric (src/RemeshingHelper.cu:72-85)
        M2f O = m2(0, 0, 0, 0);
        for (int i = 0; i < 3; i++) {
            const float4 pn = A.planes[2 * (nd[i] - A.n0) + 1];
            const V3f n = ld3(pn);
            if (norm2(n) == 0.0f) continue;
            const V3f p = ld3(A.planes[2 * (nd[i] - A.n0)]);
            float h[3];
            for (int j = 0; j < 3; j++) h[j] = dot(ld3(A.x[nd[j]]) - p, n);
            const float h1 = h[1] - h[0], h2 = h[2] - h[0];
            const float dx = inv.a * h1 + inv.c * h2, dy = inv.b * h1 + inv.d * h2;
            O = O + mdiv(outer2(dx, dy), tsqr(h[i]));
        }
        M[4] = mdiv(O, 3.0f);
    }
    const M2f S = max_tensor(M);
    M2f Q;
    float l[2];
    eig(S, Q, l);
    for (int i = 0; i < 2; i++) l[i] = tclamp(l[i], 1.0f / tsqr(A.sizeMax), 1.0f / tsqr(A.sizeMin));
    const float lMax = tmax(l[0], l[1]);
    const float lMin = lMax * tsqr(A.aspectMin);
    for (int i = 0; i < 2; i++) l[i] = tmax(l[i], lMin);
    const M2f R = qlq(Q, l);
    A.out[4 * lf] = R.a;
    A.out[4 * lf + 1] = R.b;
    A.out[4 * lf + 2] = R.c;
    A.out[4 * lf + 3] = R.d;
}

}  // namespace

// Face sizing of cloth c at the current state (4 floats per face of the cloth, in face order)
// and, when planesOut is given, the nodes' planes (8 floats per node: p, 0, n, 0).
void remesh_sizing(cs_ctx* C, int c, const RemeshParams& r, std::vector<float>& out, std::vector<float>* planesOut) {
    const MeshInfo& M = C->meshes[c];
    hipStream_t s = C->stream;
    launch_face_geometry(C, 0, C->F);
    launch_node_geometry(C, 0, C->N);
    std::vector<float> fuv(6 * (size_t)M.nf);
    for (int lf = 0; lf < M.nf; lf++)
        for (int k = 0; k < 3; k++) {
            const int vt = C->hfv[3 * (M.f0 + lf) + k];
            fuv[6 * lf + 2 * k] = C->huv[2 * vt];
            fuv[6 * lf + 2 * k + 1] = C->huv[2 * vt + 1];
        }
    float* dfuv = C->rmUv.as<float>(fuv.size());
    CS_HIP(hipMemcpyAsync(dfuv, fuv.data(), sizeof(float) * fuv.size(), hipMemcpyHostToDevice, s));
    float4* planes = C->rmPlanes.as<float4>(2 * (size_t)M.nn);
    const int fo0 = C->FC, nfo = C->F - C->FC;
    hipLaunchKernelGGL(k_near_planes, dim3(cdiv(M.nn, 256)), dim3(256), 0, s, M.n0, M.nn, C->x.ptr<float4>(),
                       C->fnodes.ptr<int4>(), fo0, nfo, 10.0f * C->P.repulsion_thickness, planes);
    CS_LAUNCH_CHECK();
    SizingArgs A;
    A.f0 = M.f0;
    A.nf = M.nf;
    A.n0 = M.n0;
    A.fuv = dfuv;
    A.fnodes = C->fnodes.ptr<int4>();
    A.fedges = C->fedges.ptr<int4>();
    A.finv = C->finv.ptr<float4>();
    A.fnrm = C->fnrm.ptr<float4>();
    A.x = C->x.ptr<float4>();
    A.v = C->v.ptr<float4>();
    A.nrm = C->nrm.ptr<float4>();
    A.planes = planes;
    A.farea = C->farea.ptr<float>();
    A.enodes = C->enodes.ptr<int2>();
    A.eadj = C->eadj.ptr<int2>();
    A.refineAngle = r.refineAngle;
    A.refineVelocity = r.refineVelocity;
    A.refineCompression = r.refineCompression;
    A.ribStiffening = r.ribStiffening;
    A.sizeMin = r.sizeMin;
    A.sizeMax = r.sizeMax;
    A.aspectMin = r.aspectMin;
    A.out = C->rmSizing.as<float>(4 * (size_t)M.nf);
    hipLaunchKernelGGL(k_face_sizing, dim3(cdiv(M.nf, 256)), dim3(256), 0, s, A);
    CS_LAUNCH_CHECK();
    out.resize(4 * (size_t)M.nf);
    CS_HIP(hipMemcpyAsync(out.data(), A.out, sizeof(float) * out.size(), hipMemcpyDeviceToHost, s));
    if (planesOut) {
        planesOut->resize(8 * (size_t)M.nn);
        CS_HIP(hipMemcpyAsync(planesOut->data(), planes, sizeof(float) * planesOut->size(), hipMemcpyDeviceToHost, s));
    }
    CS_HIP(hipStreamSynchronize(s));
}

}  // namespace cs
