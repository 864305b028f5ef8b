// cs_assemble.hip — implicit-Euler system A dv = b of one cloth
// (reference: Cloth::physicsStep GPU, src/Cloth.cu:695-764, force kernels of
// src/PhysicsHelper.cu).
//
//   A = M + dt^2 k_h I (handles) - dt^2 J_stretch - dt^2 J_bend - dt^2 J_impulse - dt J_friction
//   b = dt m g + wind + dt (f + dt J v) (elastic, impulse) + dt f (friction) + handle springs
//
// Instead of the reference's 156M-triplet COO sort at 513^2, the system is
// built in three passes:
//  1. element kernels (one thread per face / interior edge) write a compact
//     record: for stretching fu, fv and six scalars from which every 3x3 block
//     of the 9x9 Jacobian follows, plus the per-node RHS; for bending the
//     dihedral gradient (12 floats), one scalar and the per-node RHS;
//  2. a row kernel (one thread per node row, one wavefront per 64-row slice)
//     gathers the records of the faces/hinges incident to its node, forms its
//     blocks in LDS and writes its sliced-ELL row once, coalesced;
//  3. proximity blocks (dynamic pattern) are emitted as (row, col) keyed
//     records, radix-sorted (stable) and reduced into a contact CSR.
// No float atomics anywhere: the result is deterministic run to run.
#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

constexpr int FACE_REC = 6;  // float4 per face record
constexpr int EDGE_REC = 7;  // float4 per edge record

// Trilinear stretching-stiffness lookup (src/Material.cu:91-124)
__device__ __forceinline__ void stretch_k(const float* __restrict__ tab, float g00, float g01, float g11, float k[4]) {
    const float lim = (float)(MAT_N - 1) - 1e-6f;
    float x = tclamp((g00 + 0.25f) * MAT_N, 0.0f, lim);
    int xi = (int)x;
    x -= xi;
    float y = tclamp((g11 + 0.25f) * MAT_N, 0.0f, lim);
    int yi = (int)y;
    y -= yi;
    float z = tclamp(tabs(g01) * MAT_N, 0.0f, lim);
    int zi = (int)z;
    z -= zi;
    k[0] = k[1] = k[2] = k[3] = 0.0f;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int l = 0; l < 2; l++) {
                float w = (i ? x : 1.0f - x) * (j ? y : 1.0f - y) * (l ? z : 1.0f - z);
                float4 s = *reinterpret_cast<const float4*>(tab + ((((xi + i) * MAT_N) + (yi + j)) * MAT_N + (zi + l)) * 4);
                k[0] += s.x * w;
                k[1] += s.y * w;
                k[2] += s.z * w;
                k[3] += s.w * w;
            }
}

// Bending stiffness (src/Material.cu:126-155) with the bias weights fixed at set-up.
__device__ __forceinline__ float bend_k(const float* __restrict__ bend, float length, float angle, float area,
                                        float biasId, float biasW) {
    float curv = length * angle / area;
    float alpha = 0.5f * curv;
    float vw = 0.2f * alpha;
    vw = tmin(vw, 4.0f - 1e-6f);
    int vid = (int)vw;
    vid = tmax(vid, 0);
    vw -= vid;
    int bid = (int)biasId;
    float ans = 0.0f;
    ans += bend[5 * bid + vid] * ((1.0f - biasW) * (1.0f - vw));
    ans += bend[5 * bid + vid + 1] * ((1.0f - biasW) * vw);
    ans += bend[5 * (bid + 1) + vid] * (biasW * (1.0f - vw));
    ans += bend[5 * (bid + 1) + vid + 1] * (biasW * vw);
    return tmax(ans, 0.0f);
}

struct FaceCoef {
    V3f fu, fv;
    float A0, A1, A2, A3, Su, Sv;
};

// 3x3 block (a,b) of dt^2 * area * hess of the stretching energy
// (src/PhysicsHelper.cu:264-291 expanded per node pair):
//   du_a du_b (A0 fu fu^T + Su I) + dv_a dv_b (A2 fv fv^T + Sv I)
//   + A1 (du_a dv_b fu fv^T + dv_a du_b fv fu^T) + 2 A3 fuv_a fuv_b^T
__device__ __forceinline__ void stretch_block(const FaceCoef& c, float dua, float dva, float dub, float dvb,
                                              float out[9]) {
    const float uu = dua * dub, vv = dva * dvb, uv = dua * dvb, vu = dva * dub;
    V3f pa = 0.5f * (dua * c.fv + dva * c.fu);
    V3f pb = 0.5f * (dub * c.fv + dvb * c.fu);
    const float fu[3] = {c.fu.x, c.fu.y, c.fu.z}, fv[3] = {c.fv.x, c.fv.y, c.fv.z};
    const float qa[3] = {pa.x, pa.y, pa.z}, qb[3] = {pb.x, pb.y, pb.z};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float s = uu * c.A0 * fu[i] * fu[j] + vv * c.A2 * fv[i] * fv[j] +
                      c.A1 * (uv * fu[i] * fv[j] + vu * fv[i] * fu[j]) + 2.0f * c.A3 * qa[i] * qb[j];
            if (i == j) s += uu * c.Su + vv * c.Sv;
            out[3 * i + j] = s;
        }
}

// One thread per cloth face: compact stretching record + per-node RHS (with wind).
__global__ void __launch_bounds__(256) k_face_records(int f0, int nf, float dt, float wdens, float wdrag, float wvx,
                                                      float wvy, float wvz, const float4* __restrict__ x,
                                                      const float4* __restrict__ v, const int4* __restrict__ fnodes,
                                                      const float4* __restrict__ finv, const float* __restrict__ farea,
                                                      const float4* __restrict__ fnrm, const int* __restrict__ fmat,
                                                      const float* __restrict__ tables, float4* __restrict__ rec) {
    int f = f0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= f0 + nf) return;
    int4 fn = fnodes[f];
    V3f x0 = ld3(x[fn.x]), x1 = ld3(x[fn.y]), x2 = ld3(x[fn.z]);
    V3f v0 = ld3(v[fn.x]), v1 = ld3(v[fn.y]), v2 = ld3(v[fn.z]);
    float4 Y = finv[f];  // (Y00, Y01, Y10, Y11)
    float area = farea[f];
    V3f d1 = x1 - x0, d2 = x2 - x0;
    // F = [d1 d2] Y  (Face::derivative, src/Face.cu:109-111)
    V3f fu = Y.x * d1 + Y.z * d2;
    V3f fv = Y.y * d1 + Y.w * d2;
    float g00 = 0.5f * (dot(fu, fu) - 1.0f);
    float g11 = 0.5f * (dot(fv, fv) - 1.0f);
    float g01 = 0.5f * dot(fu, fv);
    float k[4];
    stretch_k(tables + (size_t)fmat[f] * (MAT_N * MAT_N * MAT_N * 4), g00, g01, g11, k);
    const float s = dt * dt * area;
    const float g00p = tmax(g00, 0.0f), g11p = tmax(g11, 0.0f);
    FaceCoef c;
    c.fu = fu; c.fv = fv;
    c.A0 = s * k[0]; c.A1 = s * k[1]; c.A2 = s * k[2]; c.A3 = s * k[3];
    c.Su = s * (k[0] * g00p + k[1] * g11p);
    c.Sv = s * (k[2] * g11p + k[1] * g00p);
    const float du[3] = {-Y.x - Y.z, Y.x, Y.z};
    const float dv[3] = {-Y.y - Y.w, Y.y, Y.w};
    // elastic force f_a = -area * grad_a
    const float cu = k[0] * g00 + k[1] * g11, cv = k[2] * g11 + k[1] * g00, cuv = 2.0f * k[3] * g01;
    // wind and air drag (src/PhysicsHelper.cu:178-200), the same on the three nodes
    V3f n = ld3(fnrm[f]);
    V3f avg = (v0 + v1 + v2) / 3.0f;
    V3f rel = mk(wvx, wvy, wvz) - avg;
    float vn = dot(n, rel);
    V3f vt = rel - vn * n;
    V3f fw = dt * ((area * (wdens * tabs(vn) * vn * n + wdrag * vt)) / 3.0f);
    const V3f vv[3] = {v0, v1, v2};
    float4* r = rec + (size_t)(f - f0) * FACE_REC;
    r[0] = make_float4(fu.x, fu.y, fu.z, c.A0);
    r[1] = make_float4(fv.x, fv.y, fv.z, c.A1);
    r[2] = make_float4(c.A2, c.A3, c.Su, c.Sv);
#pragma unroll
    for (int a = 0; a < 3; a++) {
        V3f grad = (cu * du[a] + 0.5f * cuv * dv[a]) * fu + (cv * dv[a] + 0.5f * cuv * du[a]) * fv;
        // b_a = dt f_a - sum_b A_ab v_b   with A_ab = -dt^2 J_ab
        V3f ba = (-dt * area) * grad;
#pragma unroll
        for (int b = 0; b < 3; b++) {
            float blk[9];
            stretch_block(c, du[a], dv[a], du[b], dv[b], blk);
            ba.x -= blk[0] * vv[b].x + blk[1] * vv[b].y + blk[2] * vv[b].z;
            ba.y -= blk[3] * vv[b].x + blk[4] * vv[b].y + blk[5] * vv[b].z;
            ba.z -= blk[6] * vv[b].x + blk[7] * vv[b].y + blk[8] * vv[b].z;
        }
        ba += fw;
        r[3 + a] = st4(ba);
    }
}

__device__ __forceinline__ float line_dist(V3f p, V3f a, V3f b) {  // src/PhysicsHelper.cu:314-319
    V3f e = b - a;
    V3f t = p - a;
    V3f r = e * dot(e, t) / norm2(e);
    return norm(t - r);
}
__device__ __forceinline__ float line_w1(V3f p, V3f a, V3f b) {  // src/PhysicsHelper.cu:321-325
    V3f e = b - a;
    return dot(e, p - a) / norm2(e);
}

// One thread per interior cloth edge: dihedral gradient, scale and per-node RHS
// (src/PhysicsHelper.cu:327-353, Edge::angle src/Edge.cu:104-114).
__global__ void __launch_bounds__(256) k_edge_records(int e0, int ne, float dt, const float4* __restrict__ x,
                                                      const float4* __restrict__ v, const int2* __restrict__ enodes,
                                                      const int2* __restrict__ eopp, const int2* __restrict__ eadj,
                                                      const float* __restrict__ farea,
                                                      const float4* __restrict__ fnrm, const float4* __restrict__ ebias,
                                                      const int* __restrict__ fmat, const float* __restrict__ bends,
                                                      float4* __restrict__ rec) {
    int e = e0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= e0 + ne) return;
    int2 ad = eadj[e];
    float4* r = rec + (size_t)(e - e0) * EDGE_REC;
    if (ad.x < 0 || ad.y < 0) return;  // boundary edges have no hinge
    int2 en = enodes[e], eo = eopp[e];
    V3f p0 = ld3(x[en.x]), p1 = ld3(x[en.y]), p2 = ld3(x[eo.x]), p3 = ld3(x[eo.y]);
    V3f n0 = ld3(fnrm[ad.x]), n1 = ld3(fnrm[ad.y]);
    float length = norm(p1 - p0);
    V3f ed = normalized(p0 - p1);
    float angle = atan2f(dot(ed, cross(n0, n1)), dot(n0, n1));
    float area = farea[ad.x] + farea[ad.y];
    float h0 = line_dist(p2, p0, p1), h1 = line_dist(p3, p0, p1);
    float t0 = line_w1(p2, p0, p1), t1 = line_w1(p3, p0, p1);
    V3f g[4];
    g[0] = -(1.0f - t0) * n0 / h0 - (1.0f - t1) * n1 / h1;
    g[1] = -t0 * n0 / h0 - t1 * n1 / h1;
    g[2] = n0 / h0;
    g[3] = n1 / h1;
    float4 bi = ebias[e];
    const float* bend = bends + 15 * fmat[ad.x];
    float k0 = bend_k(bend, length, angle, area, bi.x, bi.y);
    float k1 = bend_k(bend, length, angle, area, bi.z, bi.w);
    float coef = -0.25f * tmin(k0, k1) * (length * length) / area;
    // A_ab = -dt^2 coef g_a g_b^T ; b_a = dt coef g_a (theta + dt sum_b g_b . v_b)
    float gv = dot(g[0], ld3(v[en.x])) + dot(g[1], ld3(v[en.y])) + dot(g[2], ld3(v[eo.x])) + dot(g[3], ld3(v[eo.y]));
    float bs = dt * coef * (angle + dt * gv);
    r[0] = st4(g[0], -dt * dt * coef);
    r[1] = st4(g[1]);
    r[2] = st4(g[2]);
    r[3] = st4(g[3]);
    V3f b0 = bs * g[0], b1 = bs * g[1], b2 = bs * g[2], b3 = bs * g[3];
    r[4] = make_float4(b0.x, b0.y, b0.z, b1.x);
    r[5] = make_float4(b1.y, b1.z, b2.x, b2.y);
    r[6] = make_float4(b2.z, b3.x, b3.y, b3.z);
}

__device__ __forceinline__ V3f edge_rhs(const float4* r, int a) {
    float4 q4 = r[4], q5 = r[5], q6 = r[6];
    if (a == 0) return mk(q4.x, q4.y, q4.z);
    if (a == 1) return mk(q4.w, q5.x, q5.y);
    if (a == 2) return mk(q5.z, q5.w, q6.x);
    return mk(q6.y, q6.z, q6.w);
}

// The rows of one cloth's system.
struct RowArgs {
    int n0, nn;
    const int *sliceOff, *sliceW, *incPtr, *diagSlot;
    const int2* inc;
    const float* hcount;
    const float4* hpos;
    float *vals, *b;
};

// Element records and per-node data shared by every cloth.
struct RowShared {
    float dt, gx, gy, gz, hk;
    const float4 *faceRec, *finv, *edgeRec;  // records indexed by global face / edge id
    const float* mass;
    const float4 *x, *v;
};

// One thread per row (node) of one cloth; one wavefront per 64-row slice.
// acc[(slot*9 + c)*64 + lane] accumulates the row's blocks in LDS.
__device__ __forceinline__ void assemble_slice(const RowArgs& R, const RowShared& G, int s, float* acc) {
    const int nn = R.nn, n0 = R.n0;
    const float dt = G.dt, hk = G.hk;
    const int* __restrict__ incPtr = R.incPtr;
    const int2* __restrict__ inc = R.inc;
    const float4* __restrict__ faceRec = G.faceRec;
    const float4* __restrict__ finv = G.finv;
    const float4* __restrict__ edgeRec = G.edgeRec;
    const int lane = threadIdx.x, row = s * SLICE + lane;
    const int W = R.sliceW[s];
    for (int k = 0; k < W * 9; k++) acc[k * SLICE + lane] = 0.0f;
    V3f bacc = mk(0.0f, 0.0f, 0.0f);
    if (row < nn) {
        const int pe = incPtr[row + 1];
        for (int p = incPtr[row]; p < pe; p++) {
            int2 ic = inc[p];
            const int a = ic.y & 3;
            if (ic.x >= 0) {
                const int f = ic.x;
                const float4* r = faceRec + (size_t)f * FACE_REC;
                float4 r0 = r[0], r1 = r[1], r2 = r[2];
                float4 Y = finv[f];
                FaceCoef c;
                c.fu = ld3(r0); c.A0 = r0.w;
                c.fv = ld3(r1); c.A1 = r1.w;
                c.A2 = r2.x; c.A3 = r2.y; c.Su = r2.z; c.Sv = r2.w;
                const float du[3] = {-Y.x - Y.z, Y.x, Y.z};
                const float dv[3] = {-Y.y - Y.w, Y.y, Y.w};
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    float blk[9];
                    stretch_block(c, du[a], dv[a], du[b], dv[b], blk);
                    const int slot = (ic.y >> (2 + 6 * b)) & 63;
#pragma unroll
                    for (int q = 0; q < 9; q++) acc[(slot * 9 + q) * SLICE + lane] += blk[q];
                }
                bacc += ld3(r[3 + a]);
            } else {
                const int e = ic.x & 0x7fffffff;
                const float4* r = edgeRec + (size_t)e * EDGE_REC;
                float4 ga4 = r[a];
                V3f ga = ld3(ga4);
                float cc = r[0].w;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    V3f gb = ld3(r[b]);
                    const int slot = (ic.y >> (2 + 6 * b)) & 63;
                    const float ga_[3] = {cc * ga.x, cc * ga.y, cc * ga.z}, gb_[3] = {gb.x, gb.y, gb.z};
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) acc[(slot * 9 + 3 * i + j) * SLICE + lane] += ga_[i] * gb_[j];
                }
                bacc += edge_rhs(r, a);
            }
        }
        // mass, gravity and handle springs (src/PhysicsHelper.cu:152-176,379-396)
        const int g = n0 + row;
        const float m = G.mass[g];
        const float hc = R.hcount[row];
        const int ds = R.diagSlot[row];
        const float dg = m + hc * (dt * dt * hk);
        acc[(ds * 9 + 0) * SLICE + lane] += dg;
        acc[(ds * 9 + 4) * SLICE + lane] += dg;
        acc[(ds * 9 + 8) * SLICE + lane] += dg;
        bacc += (dt * m) * mk(G.gx, G.gy, G.gz);
        if (hc > 0.0f) {
            V3f hp = ld3(R.hpos[row]);
            bacc += (dt * ((hp - hc * ld3(G.x[g])) - (hc * dt) * ld3(G.v[g]))) * hk;
        }
        R.b[3 * row + 0] = bacc.x;
        R.b[3 * row + 1] = bacc.y;
        R.b[3 * row + 2] = bacc.z;
    }
    float* out = R.vals + (size_t)R.sliceOff[s] * 9 * SLICE;
    for (int k = 0; k < W * 9; k++) out[k * SLICE + lane] = acc[k * SLICE + lane];
}

__global__ void __launch_bounds__(64) k_assemble_rows(RowArgs R, RowShared G) {
    extern __shared__ float acc[];
    assemble_slice(R, G, blockIdx.x, acc);
}

// Per-cloth view of the batched assembly of independent instances.
struct AsmJob {
    RowArgs R;
    int slice0, blk0;  // first slice / first 256-row block of the cloth in the batch grids
    int u0, nu;        // its contact blocks in the batch's sorted unique blocks
    const int* cDiagSrc;
    float *bc, *Minv, *cVal;
    int *cRowPtr, *cDiag, *cCol;
};

// the job whose range [first(j), first(j + 1)) holds v (first() ascending)
template <class F> __device__ __forceinline__ int find_job(int njobs, int v, F&& first) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first(mid) <= v) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(64) k_assemble_rows_batch(const AsmJob* __restrict__ jobs, int njobs, RowShared G) {
    extern __shared__ float acc[];
    const int j = find_job(njobs, (int)blockIdx.x, [&](int m) { return jobs[m].slice0; });
    assemble_slice(jobs[j].R, G, (int)blockIdx.x - jobs[j].slice0, acc);
}

// ---------------------------------------------------------------- contacts
// The impulse and friction blocks of proximity p (src/PhysicsHelper.cu:398-468) as 16
// records keyed by keyOf(node a, node b), the block's (row, col) key or `invalid` for a
// block of no system; the RHS of row a rides on the (a, a) record.
template <class KeyFn>
__device__ __forceinline__ void emit_contact(int p, const cs_proximity_rec* __restrict__ prox, int NC, float dt,
                                             float thickness, const float4* __restrict__ x,
                                             const float4* __restrict__ v, const float* __restrict__ mass,
                                             uint64_t* __restrict__ keys, uint32_t* __restrict__ idx,
                                             float* __restrict__ recs, uint64_t invalid, KeyFn&& keyOf) {
    const cs_proximity_rec P = prox[p];
    V3f xs[4], vs[4];
    V3f n = mk(P.n[0], P.n[1], P.n[2]);
    float d = -thickness;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int nd = P.nodes[k];
        xs[k] = ld3(x[nd]);
        vs[k] = ld3(v[nd]);
        d += P.w[k] * dot(n, xs[k]);
    }
    d = tmax(-d, 0.0f);
    const float kk = P.stiffness;
    // impulse: f = 0.5 k/th d^2 N, J = -k/th d N N^T
    const float fimp = 0.5f * kk / thickness * (d * d);
    const float jimp = -kk / thickness * d;
    // friction
    const float Fn = 0.5f * kk / thickness * (d * d);
    V3f vr = mk(0.0f, 0.0f, 0.0f);
    float invMass = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        vr += P.w[k] * vs[k];
        if (P.nodes[k] < NC) invMass += (P.w[k] * P.w[k]) / mass[P.nodes[k]];
    }
    V3f vt = vr - dot(n, vr) * n;  // (I - n n^T) v
    const float fv = tmin(P.mu * Fn / norm(vt), 1.0f / (dt * invMass));
    // N . v for the impulse RHS: sum_l w_l n . v_l
    float nv = 0.0f;
#pragma unroll
    for (int l = 0; l < 4; l++) nv += P.w[l] * dot(n, vs[l]);
    const float nn3[3] = {n.x, n.y, n.z};
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int slot = 16 * p + 4 * a + b;
            const uint64_t key = keyOf(P.nodes[a], P.nodes[b]);
            keys[slot] = key;
            idx[slot] = slot;
            if (key == invalid) continue;
            float* rr = recs + (size_t)slot * 12;
            const float wab = P.w[a] * P.w[b];
            const float ci = -dt * dt * jimp * wab;  // impulse: -dt^2 J
            const float cf = dt * fv * wab;          // friction: -dt J = dt f_v w w T
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    float T = (i == j ? 1.0f : 0.0f) - nn3[i] * nn3[j];
                    rr[3 * i + j] = ci * nn3[i] * nn3[j] + cf * T;
                }
            if (a == b) {
                // dt (f_imp + dt J_imp v) + dt f_fric
                V3f ba = (dt * (fimp * P.w[a] + dt * jimp * P.w[a] * nv)) * n + (-dt * fv * P.w[a]) * vt;
                rr[9] = ba.x; rr[10] = ba.y; rr[11] = ba.z;
            } else {
                rr[9] = rr[10] = rr[11] = 0.0f;
            }
        }
}

// One thread per proximity: impulse and friction blocks (src/PhysicsHelper.cu:398-468) of
// the cloth whose rows are [n0, n0 + nn), keyed by local (row, col).
__global__ void __launch_bounds__(256) k_contact_emit(int nP, const cs_proximity_rec* __restrict__ prox, int n0, int nn,
                                                      int NC, float dt, float thickness, const float4* __restrict__ x,
                                                      const float4* __restrict__ v, const float* __restrict__ mass,
                                                      uint64_t* __restrict__ keys, uint32_t* __restrict__ idx,
                                                      float* __restrict__ recs) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nP) return;
    const uint64_t invalid = (uint64_t)nn * (uint64_t)nn;
    emit_contact(p, prox, NC, dt, thickness, x, v, mass, keys, idx, recs, invalid, [&](int na, int nb) {
        const bool in = na >= n0 && na < n0 + nn && nb >= n0 && nb < n0 + nn;
        return in ? (uint64_t)(na - n0) * (uint64_t)nn + (uint64_t)(nb - n0) : invalid;
    });
}

// The same for every cloth of a batch of independent instances at once: keyed by the
// global (row, col) node pair when both nodes are rows of one cloth.
__global__ void __launch_bounds__(256) k_contact_emit_all(int nP, const cs_proximity_rec* __restrict__ prox,
                                                          const AsmJob* __restrict__ jobs, int njobs, int NC,
                                                          float dt, float thickness, const float4* __restrict__ x,
                                                          const float4* __restrict__ v, const float* __restrict__ mass,
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ idx,
                                                          float* __restrict__ recs) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nP) return;
    const uint64_t invalid = (uint64_t)NC * (uint64_t)NC;
    emit_contact(p, prox, NC, dt, thickness, x, v, mass, keys, idx, recs, invalid, [&](int na, int nb) {
        if (na >= NC || nb >= NC) return invalid;
        const int ca = find_job(njobs, na, [&](int m) { return jobs[m].R.n0; });
        const int cb = find_job(njobs, nb, [&](int m) { return jobs[m].R.n0; });
        return ca == cb ? (uint64_t)na * (uint64_t)NC + (uint64_t)nb : invalid;
    });
}

__global__ void k_contact_heads(int n, const uint64_t* __restrict__ keys, uint64_t invalid, uint32_t* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = keys[i];
    flag[i] = (k != invalid && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
}

__global__ void k_contact_reduce(int n, int nn, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ idx,
                                 const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                 const float* __restrict__ recs, int* __restrict__ col, float* __restrict__ val,
                                 uint64_t* __restrict__ ukeys, float* __restrict__ bc, int* __restrict__ cdiag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const uint64_t k = keys[i];
    const int u = (int)pos[i];
    float s[12];
#pragma unroll
    for (int q = 0; q < 12; q++) s[q] = 0.0f;
    for (int j = i; j < n && keys[j] == k; j++) {
        const float* r = recs + (size_t)idx[j] * 12;
#pragma unroll
        for (int q = 0; q < 12; q++) s[q] += r[q];
    }
    const int rw = (int)(k / (uint64_t)nn), cl = (int)(k % (uint64_t)nn);
    col[u] = cl;
    ukeys[u] = k;
#pragma unroll
    for (int q = 0; q < 9; q++) val[(size_t)u * 9 + q] = s[q];
    if (rw == cl) {
        bc[3 * rw + 0] = s[9];
        bc[3 * rw + 1] = s[10];
        bc[3 * rw + 2] = s[11];
        cdiag[rw] = u;
    }
}

// cRowPtr[r] = first unique block with row >= r
__global__ void k_contact_rowptr(int nn, int nu, const uint64_t* __restrict__ ukeys, int* __restrict__ rowPtr) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nn) return;
    const uint64_t target = (uint64_t)r * (uint64_t)nn;
    int lo = 0, hi = nu;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (ukeys[mid] < target) lo = mid + 1; else hi = mid;
    }
    rowPtr[r] = lo;
}

__device__ __forceinline__ void block_jacobi_row(int row, const int* __restrict__ sliceOff,
                                                 const int* __restrict__ diagSlot, const float* __restrict__ vals,
                                                 const float* __restrict__ cblk, float* __restrict__ Minv) {
    const int s = row / SLICE, lane = row % SLICE;
    const float* base = vals + ((size_t)sliceOff[s] + diagSlot[row]) * 9 * SLICE + lane;
    float a[9];
#pragma unroll
    for (int q = 0; q < 9; q++) a[q] = base[q * SLICE];
    if (cblk) {
#pragma unroll
        for (int q = 0; q < 9; q++) a[q] += cblk[q];
    }
    float o[9];
    inv3(a, o);
#pragma unroll
    for (int q = 0; q < 9; q++) Minv[(size_t)row * 9 + q] = o[q];
}

__global__ void k_block_jacobi(int nn, const int* __restrict__ sliceOff, const int* __restrict__ diagSlot,
                               const float* __restrict__ vals, const int* __restrict__ cdiag,
                               const float* __restrict__ cval, float* __restrict__ Minv) {
    int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nn) return;
    const int cd = cdiag[row];
    block_jacobi_row(row, sliceOff, diagSlot, vals, cd >= 0 ? cval + (size_t)cd * 9 : nullptr, Minv);
}

// Batched instances, one thread per cloth row: the row's slice of its cloth's contact CSR
// (row pointer, diagonal block, contact RHS) from the batch-wide reduction, and its
// block-Jacobi inverse.  nu == 0: no contact anywhere in the batch.
__global__ void __launch_bounds__(256) k_contact_rows_all(const AsmJob* __restrict__ jobs, int njobs, int NC, int nu,
                                                          const uint64_t* __restrict__ ukeys,
                                                          const int* __restrict__ gdiag, const float* __restrict__ gval,
                                                          const float* __restrict__ gbc) {
    const int j = find_job(njobs, (int)blockIdx.x, [&](int m) { return jobs[m].blk0; });
    const AsmJob& J = jobs[j];
    const int r = ((int)blockIdx.x - J.blk0) * 256 + threadIdx.x;
    if (r >= J.R.nn) return;
    const int g = J.R.n0 + r;
    const float* cblk = nullptr;
    if (nu > 0) {
        // first unique block of the cloth with row >= r
        const uint64_t target = (uint64_t)g * (uint64_t)NC;
        int lo = J.u0, hi = J.u0 + J.nu;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ukeys[mid] < target) lo = mid + 1; else hi = mid;
        }
        J.cRowPtr[r] = lo - J.u0;
        if (r == J.R.nn - 1) J.cRowPtr[J.R.nn] = J.nu;
        const int cd = gdiag[g];
        J.cDiag[r] = cd >= 0 ? cd - J.u0 : -1;
        J.bc[3 * r + 0] = gbc[3 * (size_t)g + 0];
        J.bc[3 * r + 1] = gbc[3 * (size_t)g + 1];
        J.bc[3 * r + 2] = gbc[3 * (size_t)g + 2];
        if (cd >= 0) cblk = gval + (size_t)cd * 9;
    } else {
        J.cRowPtr[r] = 0;
        if (r == J.R.nn - 1) J.cRowPtr[J.R.nn] = 0;
        J.cDiag[r] = -1;
        J.bc[3 * r + 0] = 0.0f;
        J.bc[3 * r + 1] = 0.0f;
        J.bc[3 * r + 2] = 0.0f;
    }
    block_jacobi_row(r, J.R.sliceOff, J.R.diagSlot, J.R.vals, cblk, J.Minv);
}

// Batched instances, one thread per unique contact block: its cloth-local column and value.
__global__ void __launch_bounds__(256) k_contact_blocks_all(const AsmJob* __restrict__ jobs, int njobs, int NC, int nu,
                                                            const uint64_t* __restrict__ ukeys,
                                                            const int* __restrict__ gcol,
                                                            const float* __restrict__ gval) {
    const int u = blockIdx.x * 256 + threadIdx.x;
    if (u >= nu) return;
    const int j = find_job(njobs, u, [&](int m) { return jobs[m].u0; });
    const AsmJob& J = jobs[j];
    const int k = u - J.u0;
    J.cCol[k] = gcol[u] - J.R.n0;
#pragma unroll
    for (int q = 0; q < 9; q++) J.cVal[(size_t)k * 9 + q] = gval[(size_t)u * 9 + q];
}

// u0[c]: first unique block of cloth c (rows from n0[c]); u0[njobs] = nu = pos[n].
__global__ void k_contact_bounds(const AsmJob* __restrict__ jobs, int njobs, int NC, const uint32_t* __restrict__ total,
                                 const uint64_t* __restrict__ ukeys, int* __restrict__ u0) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > njobs) return;
    const int nu = (int)*total;
    if (c == njobs) { u0[c] = nu; return; }
    const uint64_t target = (uint64_t)jobs[c].R.n0 * (uint64_t)NC;
    int lo = 0, hi = nu;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ukeys[mid] < target) lo = mid + 1; else hi = mid;
    }
    u0[c] = lo;
}

}  // namespace

void launch_element_records(cs_ctx* C, int c) {
    const MeshInfo& M = C->meshes[c];
    const cs_params& P = C->P;
    hipStream_t s = C->stream;
    // records are indexed by global face / edge id (cloth faces and edges come first)
    const int tok = ktimer_begin(C, CS_KERNEL_FACE_RECORDS);
    hipLaunchKernelGGL(k_face_records, dim3(cdiv(M.nf, 256)), dim3(256), 0, s, M.f0, M.nf, P.dt, P.wind_density,
                       P.wind_drag, P.wind_velocity[0], P.wind_velocity[1], P.wind_velocity[2], C->x.ptr<float4>(),
                       C->v.ptr<float4>(), C->fnodes.ptr<int4>(), C->finv.ptr<float4>(), C->farea.ptr<float>(),
                       C->fnrm.ptr<float4>(), C->fmat.ptr<int>(), C->matTable.ptr<float>(),
                       C->faceRec.ptr<float4>() + (size_t)M.f0 * FACE_REC);
    CS_LAUNCH_CHECK();
    ktimer_end(C, tok);
    hipLaunchKernelGGL(k_edge_records, dim3(cdiv(M.ne, 256)), dim3(256), 0, s, M.e0, M.ne, P.dt, C->x.ptr<float4>(),
                       C->v.ptr<float4>(), C->enodes.ptr<int2>(), C->eopp.ptr<int2>(), C->eadj.ptr<int2>(),
                       C->farea.ptr<float>(), C->fnrm.ptr<float4>(), C->ebias.ptr<float4>(), C->fmat.ptr<int>(),
                       C->matBend.ptr<float>(), C->edgeRec.ptr<float4>() + (size_t)M.e0 * EDGE_REC);
    CS_LAUNCH_CHECK();
}

static RowArgs row_args(cs_ctx* C, int c) {
    const MeshInfo& M = C->meshes[c];
    ClothSys& S = C->sys[c];
    RowArgs R;
    R.n0 = M.n0;
    R.nn = M.nn;
    R.sliceOff = S.sliceOff.ptr<int>();
    R.sliceW = S.sliceW.ptr<int>();
    R.incPtr = S.incPtr.ptr<int>();
    R.diagSlot = S.diagSlot.ptr<int>();
    R.inc = S.inc.ptr<int2>();
    R.hcount = S.hcount.ptr<float>();
    R.hpos = S.hpos.ptr<float4>();
    R.vals = S.vals.ptr<float>();
    R.b = S.b.ptr<float>();
    return R;
}

static RowShared row_shared(cs_ctx* C) {
    const cs_params& P = C->P;
    RowShared G;
    G.dt = P.dt;
    G.gx = P.gravity[0];
    G.gy = P.gravity[1];
    G.gz = P.gravity[2];
    G.hk = P.handle_stiffness;
    G.faceRec = C->faceRec.ptr<float4>();
    G.finv = C->finv.ptr<float4>();
    G.edgeRec = C->edgeRec.ptr<float4>();
    G.mass = C->mass.ptr<float>();
    G.x = C->x.ptr<float4>();
    G.v = C->v.ptr<float4>();
    return G;
}

void launch_assemble_rows(cs_ctx* C, int c) {
    ClothSys& S = C->sys[c];
    S.solved = false;
    S.asmEpoch++;
    size_t lds = (size_t)S.maxWidth * 9 * SLICE * sizeof(float);
    const int tok = ktimer_begin(C, CS_KERNEL_ASSEMBLE_ROWS);
    hipLaunchKernelGGL(k_assemble_rows, dim3(S.nslices), dim3(SLICE), lds, C->stream, row_args(C, c), row_shared(C));
    CS_LAUNCH_CHECK();
    ktimer_end(C, tok);
}

// Every cloth's system of a batch of independent instances (each cloth's rows, contacts and
// block-Jacobi inverse equal those of launch_element_records + launch_assemble_rows +
// assemble_contacts + launch_block_jacobi on that cloth): one launch per pass over all
// cloths and one host synchronisation.  Returns the number of contact blocks.
int assemble_all(cs_ctx* C) {
    hipStream_t s = C->stream;
    const cs_params& P = C->P;
    const int ncl = C->nCloths;
    // element records of every cloth face and edge (cloths come first in both arrays)
    hipLaunchKernelGGL(k_face_records, dim3(cdiv(C->FC, 256)), dim3(256), 0, s, 0, C->FC, P.dt, P.wind_density,
                       P.wind_drag, P.wind_velocity[0], P.wind_velocity[1], P.wind_velocity[2], C->x.ptr<float4>(),
                       C->v.ptr<float4>(), C->fnodes.ptr<int4>(), C->finv.ptr<float4>(), C->farea.ptr<float>(),
                       C->fnrm.ptr<float4>(), C->fmat.ptr<int>(), C->matTable.ptr<float>(), C->faceRec.ptr<float4>());
    CS_LAUNCH_CHECK();
    if (C->EC > 0) {
        hipLaunchKernelGGL(k_edge_records, dim3(cdiv(C->EC, 256)), dim3(256), 0, s, 0, C->EC, P.dt, C->x.ptr<float4>(),
                           C->v.ptr<float4>(), C->enodes.ptr<int2>(), C->eopp.ptr<int2>(), C->eadj.ptr<int2>(),
                           C->farea.ptr<float>(), C->fnrm.ptr<float4>(), C->ebias.ptr<float4>(), C->fmat.ptr<int>(),
                           C->matBend.ptr<float>(), C->edgeRec.ptr<float4>());
        CS_LAUNCH_CHECK();
    }
    std::vector<AsmJob> jobs(ncl);
    int slices = 0, blocks = 0, maxW = 1;
    for (int c = 0; c < ncl; c++) {
        ClothSys& S = C->sys[c];
        S.solved = false;
        S.asmEpoch++;
        AsmJob& J = jobs[c];
        J.R = row_args(C, c);
        J.slice0 = slices;
        J.blk0 = blocks;
        J.u0 = J.nu = 0;
        J.cDiagSrc = nullptr;
        J.bc = S.bc.ptr<float>();
        J.Minv = S.Minv.ptr<float>();
        J.cRowPtr = S.cRowPtr.ptr<int>();
        J.cDiag = S.cDiag.ptr<int>();
        J.cCol = nullptr;
        J.cVal = nullptr;
        slices += S.nslices;
        blocks += cdiv(S.nn, 256);
        maxW = std::max(maxW, S.maxWidth);
    }
    AsmJob* dj = C->asmJobs.as<AsmJob>(ncl);
    CS_HIP(hipMemcpyAsync(dj, jobs.data(), sizeof(AsmJob) * ncl, hipMemcpyHostToDevice, s));
    const int tok = ktimer_begin(C, CS_KERNEL_ASSEMBLE_ROWS);
    hipLaunchKernelGGL(k_assemble_rows_batch, dim3(slices), dim3(SLICE), (size_t)maxW * 9 * SLICE * sizeof(float), s,
                       (const AsmJob*)dj, ncl, row_shared(C));
    CS_LAUNCH_CHECK();
    ktimer_end(C, tok);
    const int NC = C->NC;
    const int nP = C->nProx;
    int nu = 0;
    const uint64_t* ukeys = nullptr;
    const int* gdiag = nullptr;
    const float *gval = nullptr, *gbc = nullptr;
    if (nP > 0) {
        const long nr = 16L * nP;
        uint64_t* keys = C->cKeys.as<uint64_t>(nr);
        uint64_t* keysTmp = C->cKeysTmp.as<uint64_t>(nr);
        uint32_t* idx = C->cIdx.as<uint32_t>(nr);
        uint32_t* idxTmp = C->cIdxTmp.as<uint32_t>(nr);
        float* recs = C->cRec.as<float>(nr * 12);
        uint32_t* flag = C->cFlag.as<uint32_t>(nr + 1);
        uint32_t* pos = C->cPos.as<uint32_t>(nr + 1);
        int* col = C->gcCol.as<int>(nr);
        float* val = C->gcVal.as<float>(9 * nr);
        float* bc = C->gcBc.as<float>(3L * NC);
        int* diag = C->gcDiag.as<int>(NC);
        int* u0 = C->gcU0.as<int>(ncl + 1);
        fill_f32(bc, 0.0f, 3L * NC, s);
        fill_u32(reinterpret_cast<uint32_t*>(diag), 0xffffffffu, NC, s);
        hipLaunchKernelGGL(k_contact_emit_all, dim3(cdiv(nP, 256)), dim3(256), 0, s, nP,
                           C->prox.ptr<cs_proximity_rec>(), (const AsmJob*)dj, ncl, NC, P.dt, P.repulsion_thickness,
                           C->x.ptr<float4>(), C->v.ptr<float4>(), C->mass.ptr<float>(), keys, idx, recs);
        CS_LAUNCH_CHECK();
        const uint64_t invalid = (uint64_t)NC * (uint64_t)NC;
        int bits = 1;
        while (bits < 64 && (invalid >> bits) != 0) bits++;
        radix_sort_pairs(keys, idx, keysTmp, idxTmp, (int)nr, 0, bits, C->ws, s);
        hipLaunchKernelGGL(k_contact_heads, dim3(cdiv(nr, 256)), dim3(256), 0, s, (int)nr, keys, invalid, flag);
        CS_LAUNCH_CHECK();
        scan_exclusive(flag, pos, (int)nr, C->ws, s);
        hipLaunchKernelGGL(k_contact_reduce, dim3(cdiv(nr, 256)), dim3(256), 0, s, (int)nr, NC, keys, idx, flag, pos,
                           recs, col, val, keysTmp, bc, diag);
        CS_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_contact_bounds, dim3(cdiv(ncl + 1, 256)), dim3(256), 0, s, (const AsmJob*)dj, ncl, NC,
                           pos + nr, keysTmp, u0);
        CS_LAUNCH_CHECK();
        int* hu = reinterpret_cast<int*>(pinned_batch(C, sizeof(int) * (ncl + 1)));
        CS_HIP(hipMemcpyAsync(hu, u0, sizeof(int) * (ncl + 1), hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        nu = hu[ncl];
        for (int c = 0; c < ncl; c++) {
            ClothSys& S = C->sys[c];
            AsmJob& J = jobs[c];
            J.u0 = hu[c];
            J.nu = hu[c + 1] - hu[c];
            S.nContact = J.nu;
            J.cCol = S.cCol.as<int>(std::max(J.nu, 1));
            J.cVal = S.cVal.as<float>(9L * std::max(J.nu, 1));
        }
        CS_HIP(hipMemcpyAsync(dj, jobs.data(), sizeof(AsmJob) * ncl, hipMemcpyHostToDevice, s));
        ukeys = keysTmp;
        gdiag = diag;
        gval = val;
        gbc = bc;
        if (nu > 0) {
            hipLaunchKernelGGL(k_contact_blocks_all, dim3(cdiv(nu, 256)), dim3(256), 0, s, (const AsmJob*)dj, ncl, NC,
                               nu, ukeys, (const int*)col, gval);
            CS_LAUNCH_CHECK();
        }
    } else {
        for (int c = 0; c < ncl; c++) C->sys[c].nContact = 0;
    }
    hipLaunchKernelGGL(k_contact_rows_all, dim3(blocks), dim3(256), 0, s, (const AsmJob*)dj, ncl, NC, nu, ukeys, gdiag,
                       gval, gbc);
    CS_LAUNCH_CHECK();
    // the job table's host copy must outlive its copies to the device
    CS_HIP(hipStreamSynchronize(s));
    return nu;
}

int assemble_contacts(cs_ctx* C, int c) {
    const MeshInfo& M = C->meshes[c];
    ClothSys& S = C->sys[c];
    hipStream_t s = C->stream;
    // clear the previous step's contact part
    fill_f32(S.bc.ptr<float>(), 0.0f, 3L * M.nn, s);
    fill_u32(reinterpret_cast<uint32_t*>(S.cDiag.ptr<int>()), 0xffffffffu, M.nn, s);
    fill_u32(reinterpret_cast<uint32_t*>(S.cRowPtr.ptr<int>()), 0u, M.nn + 1, s);
    S.nContact = 0;
    const int nP = C->nProx;
    if (nP == 0) return 0;
    const long nr = 16L * nP;
    uint64_t* keys = C->cKeys.as<uint64_t>(nr);
    uint64_t* keysTmp = C->cKeysTmp.as<uint64_t>(nr);
    uint32_t* idx = C->cIdx.as<uint32_t>(nr);
    uint32_t* idxTmp = C->cIdxTmp.as<uint32_t>(nr);
    float* recs = C->cRec.as<float>(nr * 12);
    uint32_t* flag = C->cFlag.as<uint32_t>(nr + 1);
    uint32_t* pos = C->cPos.as<uint32_t>(nr + 1);
    hipLaunchKernelGGL(k_contact_emit, dim3(cdiv(nP, 256)), dim3(256), 0, s, nP, C->prox.ptr<cs_proximity_rec>(), M.n0,
                       M.nn, C->NC, C->P.dt, C->P.repulsion_thickness, C->x.ptr<float4>(), C->v.ptr<float4>(),
                       C->mass.ptr<float>(), keys, idx, recs);
    CS_LAUNCH_CHECK();
    const uint64_t invalid = (uint64_t)M.nn * (uint64_t)M.nn;
    int bits = 1;
    while (bits < 64 && (invalid >> bits) != 0) bits++;
    radix_sort_pairs(keys, idx, keysTmp, idxTmp, (int)nr, 0, bits, C->ws, s);
    hipLaunchKernelGGL(k_contact_heads, dim3(cdiv(nr, 256)), dim3(256), 0, s, (int)nr, keys, invalid, flag);
    CS_LAUNCH_CHECK();
    scan_exclusive(flag, pos, (int)nr, C->ws, s);
    CS_HIP(hipMemcpyAsync(C->hCount, pos + nr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    CS_HIP(hipStreamSynchronize(s));
    const int nu = (int)C->hCount[0];
    S.nContact = nu;
    if (nu == 0) return 0;
    int* col = S.cCol.as<int>(nu);
    float* val = S.cVal.as<float>(9L * nu);
    uint64_t* ukeys = keysTmp;  // sorted keys live in `keys`; the temporary is free again
    hipLaunchKernelGGL(k_contact_reduce, dim3(cdiv(nr, 256)), dim3(256), 0, s, (int)nr, M.nn, keys, idx, flag, pos,
                       recs, col, val, ukeys, S.bc.ptr<float>(), S.cDiag.ptr<int>());
    CS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_contact_rowptr, dim3(cdiv(M.nn + 1, 256)), dim3(256), 0, s, M.nn, nu, ukeys,
                       S.cRowPtr.ptr<int>());
    CS_LAUNCH_CHECK();
    return nu;
}

void launch_block_jacobi(cs_ctx* C, int c) {
    ClothSys& S = C->sys[c];
    const float* cval = S.nContact > 0 ? S.cVal.ptr<float>() : S.vals.ptr<float>();
    hipLaunchKernelGGL(k_block_jacobi, dim3(cdiv(S.nn, 256)), dim3(256), 0, C->stream, S.nn, S.sliceOff.ptr<int>(),
                       S.diagSlot.ptr<int>(), S.vals.ptr<float>(), S.cDiag.ptr<int>(), cval, S.Minv.ptr<float>());
    CS_LAUNCH_CHECK();
}

}  // namespace cs
