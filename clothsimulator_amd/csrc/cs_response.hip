// cs_response.hip — impact response: the augmented-Lagrangian gradient solve of
// CollisionOptimization (src/optimization/CollisionOptimization.cu:3-145,
// src/optimization/Optimization.cu:98-181), kept to reference semantics.
//
//   minimise   1/2 invMass sum_u m_u |x_u - x1_u|^2
//   subject to c_i = sum_j w_ij n_i . x_ij - thickness >= 0       (one per impact)
// mu = 1e3, clampViolation(x, +1) = min(x, 0), backtracking factor 0.7,
// <= 100 iterations, stops on s < 1e-12 or |f - ft| < 1e-10.
//
// The impact nodes are de-duplicated by a stable radix sort of (node, slot),
// which also gives every node the list of its constraint slots in ascending
// order, so gradients are gathered (no atomics).  The whole iteration, line
// search included, runs in one 1024-thread workgroup: the reference's host
// round trip per line-search evaluation is gone.
#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

constexpr int OB = 1024;

// The constraints c_i = sum_j w_ij n_i . x_ij - thickness as the two stages store them: impacts
// (4 nodes, FP64 record fields) and separation constraints (6 nodes: the two faces' points).
struct ImpView {
    static constexpr int K = 4;
    const cs_impact_rec* r;
    __device__ int node(int i, int j) const { return r[i].nodes[j]; }
    __device__ float w(int i, int j) const { return (float)r[i].w[j]; }
    __device__ V3f n(int i) const { return mk((float)r[i].n[0], (float)r[i].n[1], (float)r[i].n[2]); }
};
struct SepView {
    static constexpr int K = 6;
    const SepRec* r;
    __device__ int node(int i, int j) const { return r[i].nodes[j]; }
    __device__ float w(int i, int j) const { return r[i].w[j]; }
    __device__ V3f n(int i) const { return mk(r[i].n[0], r[i].n[1], r[i].n[2]); }
};

template <class View>
__global__ void k_opt_collect(int m, View V, int deform, int NC, int N, uint64_t* __restrict__ keys,
                              uint32_t* __restrict__ vals) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m) return;
    int nd = V.node(s / View::K, s % View::K);
    keys[s] = (deform == 1 || nd < NC) ? (uint64_t)nd : (uint64_t)N;
    vals[s] = (uint32_t)s;
}

__global__ void k_opt_heads(int m, int N, const uint64_t* __restrict__ keys, uint32_t* __restrict__ flag) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    flag[k] = (keys[k] < (uint64_t)N && (k == 0 || keys[k - 1] != keys[k])) ? 1u : 0u;
}

__global__ void k_opt_map(int m, int N, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                          const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos, int* __restrict__ nodes,
                          int* __restrict__ start, int* __restrict__ end, int* __restrict__ slotMap) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    uint64_t key = keys[k];
    if (key >= (uint64_t)N) return;
    int u = (int)(pos[k] + flag[k]) - 1;
    slotMap[vals[k]] = u;
    if (flag[k]) { nodes[u] = (int)key; start[u] = k; }
    if (k + 1 == m || keys[k + 1] != key) end[u] = k + 1;
}

__device__ float wg_sum(float v, float* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int w = 0; w < OB / 64; w++) s += lds[w];
        lds[OB / 64] = s;
    }
    __syncthreads();
    return lds[OB / 64];
}

template <class View>
struct OptArgs {
    int nC;
    const uint32_t* nNptr;  // unique node count (scan total)
    View V;
    const int* nodes;
    const int* start;
    const int* end;
    const int* slotMap;
    const uint32_t* sortedSlots;
    float4* x;
    const float4* x1;
    const float* mass;  // per node: the objective's weight (mass for impacts, area for separation)
    int NC;
    float obstacleMass, thickness;  // obstacle nodes' weight
    float4 *X, *Xn, *G;
    float *lambda, *coef;
};

template <class View>
__device__ __forceinline__ float node_mass(const OptArgs<View>& A, int u) {
    int nd = A.nodes[u];
    return nd < A.NC ? A.mass[nd] : A.obstacleMass;
}

template <class View>
__device__ float constraint(const OptArgs<View>& A, const float4* Xp, int i) {
    const V3f n = A.V.n(i);
    float c = -A.thickness;
    for (int j = 0; j < View::K; j++) {
        int u = A.slotMap[View::K * i + j];
        V3f p = u > -1 ? ld3(Xp[u]) : ld3(A.x[A.V.node(i, j)]);
        c += A.V.w(i, j) * dot(n, p);
    }
    return c;
}

// value(x) (src/optimization/Optimization.cu:20-31): objective + penalty terms
template <class View>
__device__ float value(const OptArgs<View>& A, int nN, const float4* Xp, float invMass, float mu, float* lds,
                       bool storeCoef) {
    float a = 0.0f;
    for (int u = threadIdx.x; u < nN; u += OB) {
        V3f d = ld3(Xp[u]) - ld3(A.x1[A.nodes[u]]);
        a += node_mass(A, u) * norm2(d);
    }
    float obj = 0.5f * wg_sum(a, lds) * invMass;
    float pen = 0.0f;
    for (int i = threadIdx.x; i < A.nC; i += OB) {
        float coef = tmin(constraint(A, Xp, i) + A.lambda[i] / mu, 0.0f);
        if (storeCoef) A.coef[i] = coef;
        if (coef != 0.0f) pen += 0.5f * mu * (coef * coef);
    }
    return obj + wg_sum(pen, lds);
}

template <class View>
__global__ void __launch_bounds__(OB) k_opt_solve(OptArgs<View> A, int* iterOut) {
    __shared__ float lds[OB / 64 + 1];
    const int nN = (int)*A.nNptr;
    if (nN == 0) return;
    float im = 0.0f;
    for (int u = threadIdx.x; u < nN; u += OB) im += 1.0f / node_mass(A, u);
    const float invMass = wg_sum(im, lds) / (float)nN;
    for (int u = threadIdx.x; u < nN; u += OB) A.X[u] = A.x[A.nodes[u]];
    for (int i = threadIdx.x; i < A.nC; i += OB) A.lambda[i] = 0.0f;
    __syncthreads();
    const float mu = 1e3f;
    float s = 1e-3f * (float)nN;
    int iter = 0;
    for (; iter < 100; iter++) {
        // valueAndGradient (src/optimization/Optimization.cu:33-49)
        const float f = value(A, nN, A.X, invMass, mu, lds, true);
        __syncthreads();
        float g2 = 0.0f;
        for (int u = threadIdx.x; u < nN; u += OB) {
            V3f g = (invMass * node_mass(A, u)) * (ld3(A.X[u]) - ld3(A.x1[A.nodes[u]]));
            for (int k = A.start[u]; k < A.end[u]; k++) {
                const int slot = (int)A.sortedSlots[k];
                const int i = slot / View::K, j = slot % View::K;
                const float coef = A.coef[i];
                if (coef != 0.0f) g += ((mu * coef) * A.V.w(i, j)) * A.V.n(i);
            }
            A.G[u] = st4(g);
            g2 += norm2(g);
        }
        const float norm2g = wg_sum(g2, lds);
        s /= 0.7f;
        float ft;
        do {
            s *= 0.7f;
            for (int u = threadIdx.x; u < nN; u += OB) A.Xn[u] = st4(ld3(A.X[u]) - s * ld3(A.G[u]));
            __syncthreads();
            ft = value(A, nN, A.Xn, invMass, mu, lds, false);
        } while (ft >= f - 0.5f * s * norm2g && s >= 1e-12f && tabs(f - ft) >= 1e-10f);
        if (s < 1e-12f || tabs(f - ft) < 1e-10f) break;
        for (int u = threadIdx.x; u < nN; u += OB) A.X[u] = A.Xn[u];
        __syncthreads();
        // updateMultiplier (src/optimization/Optimization.cu:51-57)
        for (int i = threadIdx.x; i < A.nC; i += OB) A.lambda[i] = tmin(A.lambda[i] + mu * constraint(A, A.X, i), 0.0f);
        __syncthreads();
    }
    for (int u = threadIdx.x; u < nN; u += OB) {
        float4 p = A.X[u];
        A.x[A.nodes[u]] = make_float4(p.x, p.y, p.z, 0.0f);
    }
    if (threadIdx.x == 0 && iterOut) *iterOut = iter;
}

template <class View>
void opt_response(cs_ctx* C, View V, int nImp, int deform, const float* weight, float obstacleWeight) {
    if (nImp <= 0) return;
    hipStream_t s = C->stream;
    const int m = View::K * nImp;
    uint64_t* keys = C->optKeys.as<uint64_t>(m);
    uint64_t* ktmp = C->optKeysTmp.as<uint64_t>(m);
    uint32_t* vals = C->optVals.as<uint32_t>(m);
    uint32_t* vtmp = C->optValsTmp.as<uint32_t>(m);
    uint32_t* flag = C->optFlag.as<uint32_t>(m + 1);
    uint32_t* pos = C->optPos.as<uint32_t>(m + 1);
    int* nodes = C->optNodes.as<int>(m);
    int* start = C->optStart.as<int>(2 * (size_t)m);
    int* end = start + m;
    int* slotMap = C->optSlot.as<int>(m);
    hipLaunchKernelGGL(k_opt_collect<View>, dim3(cdiv(m, 256)), dim3(256), 0, s, m, V, deform, C->NC, C->N, keys,
                       vals);
    CS_LAUNCH_CHECK();
    int bits = 1;
    while ((1L << bits) <= C->N) bits++;
    radix_sort_pairs(keys, vals, ktmp, vtmp, m, 0, bits, C->ws, s);
    hipLaunchKernelGGL(k_opt_heads, dim3(cdiv(m, 256)), dim3(256), 0, s, m, C->N, keys, flag);
    CS_LAUNCH_CHECK();
    scan_exclusive(flag, pos, m, C->ws, s);
    fill_u32(reinterpret_cast<uint32_t*>(slotMap), 0xffffffffu, m, s);
    hipLaunchKernelGGL(k_opt_map, dim3(cdiv(m, 256)), dim3(256), 0, s, m, C->N, keys, vals, flag, pos, nodes, start,
                       end, slotMap);
    CS_LAUNCH_CHECK();
    OptArgs<View> A;
    A.nC = nImp;
    A.nNptr = pos + m;
    A.V = V;
    A.nodes = nodes; A.start = start; A.end = end; A.slotMap = slotMap; A.sortedSlots = vals;
    A.x = C->x.ptr<float4>(); A.x1 = C->x1.ptr<float4>(); A.mass = weight;
    A.NC = C->NC;
    A.obstacleMass = obstacleWeight;
    A.thickness = C->P.collision_thickness;
    A.X = C->optX.as<float4>(m);
    A.Xn = C->optNext.as<float4>(m);
    A.G = C->optGrad.as<float4>(m);
    A.lambda = C->optLambda.as<float>(nImp);
    A.coef = C->optCoef.as<float>(nImp);
    int* iterOut = C->optScal.as<int>(4);
    hipLaunchKernelGGL(k_opt_solve<View>, dim3(1), dim3(OB), 0, s, A, iterOut);
    CS_LAUNCH_CHECK();
}

}  // namespace

void collision_response(cs_ctx* C, int nImp, int deform, float obstacleMass) {
    ImpView V;
    V.r = C->impAcc.ptr<cs_impact_rec>();
    opt_response(C, V, nImp, deform, C->mass.ptr<float>(), obstacleMass);
}

// SeparationOptimization (src/optimization/SeparationOptimization.cu:3-158): the same solve with
// per-node area weights (obstacle nodes: obstacleArea) and 6-node constraints.
void separation_response(cs_ctx* C, const SepRec* recs, int n, int deform, float obstacleArea) {
    SepView V;
    V.r = recs;
    opt_response(C, V, n, deform, C->area.ptr<float>(), obstacleArea);
}

}  // namespace cs
