// cs_proximity.hip — repulsion/friction proximities (Simulator::findProximitiesGpu,
// src/Simulator.cu:300-349; tests src/PhysicsHelper.cu:3-150; records src/Proximity.cu).
//
// Candidate face pairs come from the LBVH at 2 x repulsion thickness.  One
// thread per pair runs its 15 tests (6 vertex-face, 9 edge-edge).  The nearest candidate per (primitive, side) is kept with a
// 64-bit atomicMin on (distance bits << 32 | other id): distances are >= 0, so
// the unsigned order of the packed word is the reference's
// Pair<float, X*>::operator< with pointer order = global id order, and the
// minimum does not depend on the order threads arrive in.  A scan then turns
// every slot with distance < 2 x thickness into a record, in (primitive, side)
// order: nodes, then edges, then faces.
#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

__device__ __forceinline__ uint64_t pack(float d, int other) {
    return ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)other;
}

// inEdge (src/PhysicsHelper.cu:3-18)
__device__ __forceinline__ bool in_edge(float w, int2 a, int2 e1n, int2 e1adj, const float4* __restrict__ x,
                                        const float4* __restrict__ fnrm) {
    V3f p = (1.0f - w) * ld3(x[a.x]) + w * ld3(x[a.y]);
    bool in = true;
    const int nodes[2] = {e1n.x, e1n.y}, adj[2] = {e1adj.x, e1adj.y};
#pragma unroll
    for (int i = 0; i < 2; i++) {
        if (adj[i] < 0) continue;
        V3f q0 = ld3(x[nodes[i]]), q1 = ld3(x[nodes[1 - i]]);
        in &= (mixed(q1 - q0, ld3(fnrm[adj[i]]), p - q0) >= 0.0f);
    }
    return in;
}

struct ProxArgs {
    const int2* pairs;
    long nPairs;
    int NC, EC, FC;
    const float4 *x, *nrm, *fnrm;
    const int4 *fnodes, *fedges;
    const int2 *enodes, *eadj;
    uint64_t *bestNode, *bestEdge, *bestFace;
};

// checkVertexFaceProximityGpu (src/PhysicsHelper.cu:20-49): node (at xn) against face f
// (nodes fn at xf)
__device__ __forceinline__ void vf_test(const ProxArgs& A, int node, V3f xn, int f, int4 fn, const V3f (&xf)[3]) {
    if (node == fn.x || node == fn.y || node == fn.z) return;
    V3f n;
    float w[4], h;
    if (!signed_vf(xn, xf[0], xf[1], xf[2], n, w, h)) return;
    const float d = fabsf(h);
    if (!(tmin(-w[1], tmin(-w[2], -w[3])) >= 1e-6f)) return;
    if (node < A.NC) {
        int side = dot(n, ld3(A.nrm[node])) >= 0.0f ? 0 : 1;
        atomicMin((unsigned long long*)&A.bestNode[2 * node + side], (unsigned long long)pack(d, f));
    }
    if (f < A.FC) {
        int side = -dot(n, ld3(A.fnrm[f])) >= 0.0f ? 0 : 1;
        atomicMin((unsigned long long*)&A.bestFace[2 * f + side], (unsigned long long)pack(d, node));
    }
}

// checkEdgeEdgeProximityGpu (src/PhysicsHelper.cu:51-81): edge e0 (nodes a at xa) against
// edge e1 (nodes b at xb)
__device__ __forceinline__ void ee_test(const ProxArgs& A, int e0, int2 a, V3f xa0, V3f xa1, int e1, int2 b, V3f xb0,
                                        V3f xb1) {
    if (a.x == b.x || a.x == b.y || a.y == b.x || a.y == b.y) return;
    V3f n;
    float w[4], h;
    if (!signed_ee(xa0, xa1, xb0, xb1, n, w, h)) return;
    const float d = fabsf(h);
    if (!(tmin(tmin(w[0], w[1]), tmin(-w[2], -w[3])) >= 1e-6f)) return;
    if (!in_edge(w[1], a, b, A.eadj[e1], A.x, A.fnrm)) return;
    if (!in_edge(-w[3], b, a, A.eadj[e0], A.x, A.fnrm)) return;
    if (e0 < A.EC) {
        int side = dot(n, ld3(A.nrm[a.x]) + ld3(A.nrm[a.y])) >= 0.0f ? 0 : 1;
        atomicMin((unsigned long long*)&A.bestEdge[2 * e0 + side], (unsigned long long)pack(d, e1));
    }
    if (e1 < A.EC) {
        int side = -dot(n, ld3(A.nrm[b.x]) + ld3(A.nrm[b.y])) >= 0.0f ? 0 : 1;
        atomicMin((unsigned long long*)&A.bestEdge[2 * e1 + side], (unsigned long long)pack(d, e0));
    }
}

// position of node id among a face's three nodes (every edge node of the face is one of them)
__device__ __forceinline__ V3f pick(int id, int4 fn, const V3f (&xf)[3]) {
    return id == fn.x ? xf[0] : (id == fn.y ? xf[1] : xf[2]);
}

// checkProximitiesGpu (src/PhysicsHelper.cu:88-111), one thread per face pair: the two faces'
// six node positions are loaded once for all 15 tests (the 3 + 3 vertex-face tests, then the
// 3 x 3 edge-edge tests).  The atomicMin results do not depend on which thread runs a test.
__global__ void __launch_bounds__(256) k_prox_tests(ProxArgs A) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.nPairs) return;
    const int2 pr = A.pairs[p];
    const int4 fa = A.fnodes[pr.x], fb = A.fnodes[pr.y];
    const V3f xa[3] = {ld3(A.x[fa.x]), ld3(A.x[fa.y]), ld3(A.x[fa.z])};
    const V3f xb[3] = {ld3(A.x[fb.x]), ld3(A.x[fb.y]), ld3(A.x[fb.z])};
    vf_test(A, fa.x, xa[0], pr.y, fb, xb);
    vf_test(A, fa.y, xa[1], pr.y, fb, xb);
    vf_test(A, fa.z, xa[2], pr.y, fb, xb);
    vf_test(A, fb.x, xb[0], pr.x, fa, xa);
    vf_test(A, fb.y, xb[1], pr.x, fa, xa);
    vf_test(A, fb.z, xb[2], pr.x, fa, xa);
    const int4 ea = A.fedges[pr.x], eb = A.fedges[pr.y];
    const int ia[3] = {ea.x, ea.y, ea.z}, ib[3] = {eb.x, eb.y, eb.z};
    int2 na[3], nb[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        na[j] = A.enodes[ia[j]];
        nb[j] = A.enodes[ib[j]];
    }
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
        for (int l = 0; l < 3; l++)
            ee_test(A, ia[j], na[j], pick(na[j].x, fa, xa), pick(na[j].y, fa, xa), ib[l], nb[l],
                    pick(nb[l].x, fb, xb), pick(nb[l].y, fb, xb));
}

__global__ void k_prox_flags(long nTotal, long nNodeSlots, long nEdgeSlots, const uint64_t* __restrict__ bestNode,
                             const uint64_t* __restrict__ bestEdge, const uint64_t* __restrict__ bestFace, float lim,
                             uint32_t* __restrict__ flag) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nTotal) return;
    uint64_t v = k < nNodeSlots ? bestNode[k] : (k < nNodeSlots + nEdgeSlots ? bestEdge[k - nNodeSlots] : bestFace[k - nNodeSlots - nEdgeSlots]);
    uint32_t ok = 0;
    if (v != ~0ull) {
        float d = __uint_as_float((uint32_t)(v >> 32));
        ok = d < lim ? 1u : 0u;
    }
    flag[k] = ok;
}

// Proximity::Proximity (src/Proximity.cu:5-21)
__device__ void make_vf(int node, int f, const float4* x, const int4* fnodes, const float* area, const float* farea,
                        float kColl, float mu, float muObs, int NC, int FC, cs_proximity_rec& r) {
    int4 fn = fnodes[f];
    r.nodes[0] = node; r.nodes[1] = fn.x; r.nodes[2] = fn.y; r.nodes[3] = fn.z;
    r.stiffness = kColl * tmin(area[node], farea[f]);
    V3f n;
    float w[4], h = 0.0f;
    signed_vf(ld3(x[node]), ld3(x[fn.x]), ld3(x[fn.y]), ld3(x[fn.z]), n, w, h);
    if (h < 0.0f) n = -n;
    r.n[0] = n.x; r.n[1] = n.y; r.n[2] = n.z;
    for (int k = 0; k < 4; k++) r.w[k] = w[k];
    r.mu = (node < NC && f < FC) ? mu : muObs;
}

__device__ float edge_area(int e, const int2* eadj, const float* farea) {  // Edge::area (src/Edge.cu:96-102)
    int2 a = eadj[e];
    float s = 0.0f;
    if (a.x >= 0) s += farea[a.x];
    if (a.y >= 0) s += farea[a.y];
    return s;
}

__device__ void make_ee(int e0, int e1, const float4* x, const int2* enodes, const int2* eadj, const float* farea,
                        float kColl, float mu, float muObs, int EC, cs_proximity_rec& r) {
    int2 a = enodes[e0], b = enodes[e1];
    r.nodes[0] = a.x; r.nodes[1] = a.y; r.nodes[2] = b.x; r.nodes[3] = b.y;
    r.stiffness = kColl * tmin(edge_area(e0, eadj, farea), edge_area(e1, eadj, farea));
    V3f n;
    float w[4], h = 0.0f;
    signed_ee(ld3(x[a.x]), ld3(x[a.y]), ld3(x[b.x]), ld3(x[b.y]), n, w, h);
    if (h < 0.0f) n = -n;
    r.n[0] = n.x; r.n[1] = n.y; r.n[2] = n.z;
    for (int k = 0; k < 4; k++) r.w[k] = w[k];
    r.mu = (e0 < EC && e1 < EC) ? mu : muObs;
}

__global__ void k_prox_records(long nTotal, long nNodeSlots, long nEdgeSlots, const uint64_t* __restrict__ bestNode,
                               const uint64_t* __restrict__ bestEdge, const uint64_t* __restrict__ bestFace,
                               const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                               const float4* __restrict__ x, const int4* __restrict__ fnodes,
                               const int2* __restrict__ enodes, const int2* __restrict__ eadj,
                               const float* __restrict__ area, const float* __restrict__ farea, float kColl, float mu,
                               float muObs, int NC, int EC, int FC, cs_proximity_rec* __restrict__ out) {
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nTotal || !flag[k]) return;
    cs_proximity_rec r;
    if (k < nNodeSlots) {
        int other = (int)(uint32_t)bestNode[k];
        make_vf((int)(k / 2), other, x, fnodes, area, farea, kColl, mu, muObs, NC, FC, r);
    } else if (k < nNodeSlots + nEdgeSlots) {
        long kk = k - nNodeSlots;
        int other = (int)(uint32_t)bestEdge[kk];
        make_ee((int)(kk / 2), other, x, enodes, eadj, farea, kColl, mu, muObs, EC, r);
    } else {
        long kk = k - nNodeSlots - nEdgeSlots;
        int other = (int)(uint32_t)bestFace[kk];
        make_vf(other, (int)(kk / 2), x, fnodes, area, farea, kColl, mu, muObs, NC, FC, r);
    }
    out[pos[k]] = r;
}

}  // namespace

int find_proximities(cs_ctx* C, bool build_trees) {
    hipStream_t s = C->stream;
    const float rep = C->P.repulsion_thickness;
    if (build_trees) bvh_build_all(C, false);
    long nPairs = bvh_all_pairs(C, 2.0f * rep);
    fill_u64(C->bestNode.ptr<uint64_t>(), ~0ull, 2L * C->N, s);
    fill_u64(C->bestEdge.ptr<uint64_t>(), ~0ull, 2L * C->E, s);
    fill_u64(C->bestFace.ptr<uint64_t>(), ~0ull, 2L * C->F, s);
    if (nPairs > 0) {
        ProxArgs A;
        A.pairs = C->pairs.ptr<int2>();
        A.nPairs = nPairs;
        A.NC = C->NC; A.EC = C->EC; A.FC = C->FC;
        A.x = C->x.ptr<float4>(); A.nrm = C->nrm.ptr<float4>(); A.fnrm = C->fnrm.ptr<float4>();
        A.fnodes = C->fnodes.ptr<int4>(); A.fedges = C->fedges.ptr<int4>();
        A.enodes = C->enodes.ptr<int2>(); A.eadj = C->eadj.ptr<int2>();
        A.bestNode = C->bestNode.ptr<uint64_t>(); A.bestEdge = C->bestEdge.ptr<uint64_t>();
        A.bestFace = C->bestFace.ptr<uint64_t>();
        hipLaunchKernelGGL(k_prox_tests, dim3(cdiv(nPairs, 256)), dim3(256), 0, s, A);
        CS_LAUNCH_CHECK();
    }
    const long nNode = 2L * C->N, nEdge = 2L * C->E, nFace = 2L * C->F;
    const long nTotal = nNode + nEdge + nFace;
    uint32_t* flag = C->proxFlag.as<uint32_t>(nTotal + 1);
    uint32_t* pos = C->proxPos.as<uint32_t>(nTotal + 1);
    hipLaunchKernelGGL(k_prox_flags, dim3(cdiv(nTotal, 256)), dim3(256), 0, s, nTotal, nNode, nEdge,
                       C->bestNode.ptr<uint64_t>(), C->bestEdge.ptr<uint64_t>(), C->bestFace.ptr<uint64_t>(),
                       2.0f * rep, flag);
    CS_LAUNCH_CHECK();
    scan_exclusive(flag, pos, (int)nTotal, C->ws, s);
    CS_HIP(hipMemcpyAsync(C->hCount, pos + nTotal, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    CS_HIP(hipStreamSynchronize(s));
    const int nP = (int)C->hCount[0];
    C->nProx = nP;
    cs_proximity_rec* out = C->prox.as<cs_proximity_rec>(std::max(nP, 1));
    if (nP > 0) {
        hipLaunchKernelGGL(k_prox_records, dim3(cdiv(nTotal, 256)), dim3(256), 0, s, nTotal, nNode, nEdge,
                           C->bestNode.ptr<uint64_t>(), C->bestEdge.ptr<uint64_t>(), C->bestFace.ptr<uint64_t>(), flag,
                           pos, C->x.ptr<float4>(), C->fnodes.ptr<int4>(), C->enodes.ptr<int2>(), C->eadj.ptr<int2>(),
                           C->area.ptr<float>(), C->farea.ptr<float>(), C->P.collision_stiffness, C->P.friction,
                           C->P.obstacle_friction, C->NC, C->EC, C->FC, out);
        CS_LAUNCH_CHECK();
    }
    return nP;
}

}  // namespace cs
