// cs_amg.hip — aggregation multigrid preconditioner of the PCG (CS_PRECOND_AMG).
//
// The reference preconditions its CG with the Jacobi diagonal
// (src/Cloth.cu:766-809); at 513^2 nodes that takes thousands of iterations,
// because the iteration count of a one-level preconditioner grows with the
// mesh size.  This preconditioner is one V(1,1) cycle of aggregation
// multigrid; it changes the iteration count, not the system solved
// (||r|| <= tol ||b|| on the same A and b).
//
//   aggregation graph of level 0: cloth nodes adjacent through a face;
//   aggregation (host, once per cloth; the topology never changes): in
//     ascending order a node whose neighbours are all unaggregated founds an
//     aggregate with all of them; then each remaining node joins the aggregate
//     of its first aggregated neighbour (ascending) or founds a singleton;
//     the coarse graph links aggregates holding adjacent nodes; levels are
//     added until one has <= 64 nodes;
//   prolongation: piecewise constant, a 3x3 identity block per node;
//   operators: level 0 is the full system (sliced ELL + contact blocks);
//     level l+1 = P^T A_l P of the static system (no contact blocks), summed
//     per step on the device from precomputed contributor lists, in a fixed
//     order;
//   cycle: x = w D^-1 b; t = b - A x; coarse correction e from P^T t;
//     result = x' + w D^-1 (b - A x') with x' = x + P e; w = 0.6; D the 3x3
//     diagonal blocks (level 0: the block-Jacobi Minv of the full system);
//     the coarsest level is solved with its dense inverse (Gauss-Jordan in
//     FP64, one workgroup, once per solve).
// The cycle is symmetric (same smoother before and after, exact coarse solve),
// so the preconditioner is SPD and CG applies.
#include <algorithm>
#include <cstring>
#include <map>

#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

constexpr int AMG_COARSEST = 64;
constexpr int AB = 256;

// ---------------------------------------------------------------- host topology

void aggregate(const std::vector<int>& gp, const std::vector<int>& gi, std::vector<int>& agg, int& na) {
    const int n = (int)gp.size() - 1;
    agg.assign(n, -1);
    na = 0;
    for (int i = 0; i < n; i++) {
        if (agg[i] >= 0) continue;
        bool free = true;
        for (int k = gp[i]; k < gp[i + 1]; k++) free = free && agg[gi[k]] < 0;
        if (!free) continue;
        agg[i] = na;
        for (int k = gp[i]; k < gp[i + 1]; k++) agg[gi[k]] = na;
        na++;
    }
    for (int i = 0; i < n; i++) {
        if (agg[i] >= 0) continue;
        int a = -1;
        for (int k = gp[i]; k < gp[i + 1] && a < 0; k++) a = agg[gi[k]];
        agg[i] = a >= 0 ? a : na++;
    }
}

void coarse_graph(const std::vector<int>& gp, const std::vector<int>& gi, const std::vector<int>& agg, int na,
                  std::vector<int>& cp, std::vector<int>& cidx) {
    std::vector<std::vector<int>> adj(na);
    for (int i = 0; i + 1 < (int)gp.size(); i++)
        for (int k = gp[i]; k < gp[i + 1]; k++)
            if (agg[i] != agg[gi[k]]) adj[agg[i]].push_back(agg[gi[k]]);
    cp.assign(na + 1, 0);
    cidx.clear();
    for (int a = 0; a < na; a++) {
        std::sort(adj[a].begin(), adj[a].end());
        adj[a].erase(std::unique(adj[a].begin(), adj[a].end()), adj[a].end());
        cidx.insert(cidx.end(), adj[a].begin(), adj[a].end());
        cp[a + 1] = (int)cidx.size();
    }
}

template <class V> void upload_vec(DBuf& d, const std::vector<V>& h, hipStream_t s) {
    V* p = d.as<V>(std::max<size_t>(1, h.size()));
    if (!h.empty()) CS_HIP(hipMemcpyAsync(p, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice, s));
}

// ---------------------------------------------------------------- numeric setup kernels

// coarse block k = sum of its contributors; FROM_ELL: contributor ids are value
// offsets in the level-0 sliced ELL (component c at id + 64 c), else block ids of
// the finer block CSR (component c at 9 id + c)
template <bool FROM_ELL>
__device__ __forceinline__ void galerkin_block(int k, const int* __restrict__ cptr, const int* __restrict__ cidx,
                                               const float* __restrict__ fine, float* __restrict__ out) {
    float s[9];
#pragma unroll
    for (int c = 0; c < 9; c++) s[c] = 0.0f;
    for (int t = cptr[k]; t < cptr[k + 1]; t++) {
        const size_t id = (size_t)cidx[t];
#pragma unroll
        for (int c = 0; c < 9; c++) s[c] += FROM_ELL ? fine[id + (size_t)c * SLICE] : fine[id * 9 + c];
    }
#pragma unroll
    for (int c = 0; c < 9; c++) out[(size_t)k * 9 + c] = s[c];
}

template <bool FROM_ELL>
__global__ void __launch_bounds__(AB) k_amg_galerkin(int nnzb, const int* __restrict__ cptr,
                                                     const int* __restrict__ cidx, const float* __restrict__ fine,
                                                     float* __restrict__ out) {
    const int k = blockIdx.x * AB + threadIdx.x;
    if (k < nnzb) galerkin_block<FROM_ELL>(k, cptr, cidx, fine, out);
}

// Level-1 diagonal blocks += the contact diagonal blocks of their level-0 nodes (in
// ascending node order): obstacle contacts and each node's share of a self-contact stiffen
// its aggregate, so the coarse correction does not move what the contacts hold.
__device__ __forceinline__ void contact_diag_row(int a, const int* __restrict__ mptr, const int* __restrict__ midx,
                                                 const int* __restrict__ diag1, const int* __restrict__ cDiag,
                                                 const float* __restrict__ cVal, float* __restrict__ val1) {
    float s[9];
#pragma unroll
    for (int c = 0; c < 9; c++) s[c] = 0.0f;
    bool any = false;
    for (int t = mptr[a]; t < mptr[a + 1]; t++) {
        const int k = cDiag[midx[t]];
        if (k < 0) continue;
        any = true;
#pragma unroll
        for (int c = 0; c < 9; c++) s[c] += cVal[(size_t)k * 9 + c];
    }
    if (!any) return;
    float* d = val1 + (size_t)diag1[a] * 9;
#pragma unroll
    for (int c = 0; c < 9; c++) d[c] += s[c];
}

__global__ void __launch_bounds__(AB) k_amg_contact_diag(int n1, const int* __restrict__ mptr,
                                                         const int* __restrict__ midx, const int* __restrict__ diag1,
                                                         const int* __restrict__ cDiag, const float* __restrict__ cVal,
                                                         float* __restrict__ val1) {
    const int a = blockIdx.x * AB + threadIdx.x;
    if (a < n1) contact_diag_row(a, mptr, midx, diag1, cDiag, cVal, val1);
}

__device__ __forceinline__ void dinv_row(int i, const int* __restrict__ diag, const float* __restrict__ val,
                                         float* __restrict__ Dinv) {
    float a[9], o[9];
#pragma unroll
    for (int c = 0; c < 9; c++) a[c] = val[(size_t)diag[i] * 9 + c];
    inv3(a, o);
#pragma unroll
    for (int c = 0; c < 9; c++) Dinv[(size_t)i * 9 + c] = o[c];
}

__global__ void __launch_bounds__(AB) k_amg_dinv(int n, const int* __restrict__ diag, const float* __restrict__ val,
                                                 float* __restrict__ Dinv) {
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i < n) dinv_row(i, diag, val, Dinv);
}

// Dense inverse of the coarsest operator (m = 3n <= 192, symmetric positive
// definite) by the symmetric sweep operator in FP64 (Gauss-Jordan without
// pivoting that keeps the matrix symmetric), the lower triangle held in
// registers: entry e = r (r + 1) / 2 + k of the triangle is slot e / 1024 of
// thread e % 1024.  Per pivot c, column c goes through LDS and
//   a_cc = -1/p,  a_rc = a_rc/p (r != c),  a_rk -= a_rc (a_kc/p) (r, k != c);
// after every pivot the triangle holds -A^-1.  The operator is SPD, so the
// pivots are positive Schur-complement diagonals.
constexpr int CB = 1024;
constexpr int CT = (3 * AMG_COARSEST * (3 * AMG_COARSEST + 1) / 2 + CB - 1) / CB;  // triangle slots per thread
__device__ __forceinline__ void coarse_inverse(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                               const float* __restrict__ val, double* __restrict__ work,
                                               float* __restrict__ inv) {
    __shared__ double colc[3 * AMG_COARSEST];
    const int m = 3 * n, tid = threadIdx.x;
    const int T = m * (m + 1) / 2;
    for (int e = tid; e < m * m; e += CB) work[e] = 0.0;
    __syncthreads();
    for (int i = tid; i < n; i += CB)
        for (int q = rp[i]; q < rp[i + 1]; q++)
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) work[(size_t)(3 * i + a) * m + 3 * ci[q] + b] = val[(size_t)q * 9 + 3 * a + b];
    __syncthreads();
    double v[CT];
    int rk[CT];  // r << 8 | k (-1: no entry)
#pragma unroll
    for (int u = 0; u < CT; u++) {
        const int e = tid + u * CB;
        int r = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
        while (r * (r + 1) / 2 > e) r--;
        while ((r + 1) * (r + 2) / 2 <= e) r++;
        const int k = e - r * (r + 1) / 2;
        const bool ok = e < T;
        rk[u] = ok ? (r << 8 | k) : -1;
        v[u] = ok ? work[(size_t)r * m + k] : 0.0;
    }
    for (int c = 0; c < m; c++) {
#pragma unroll
        for (int u = 0; u < CT; u++) {
            int q = rk[u];
            asm volatile("" : "+v"(q));  // keeps the derived LDS addresses out of registers across pivots
            if (q < 0) continue;
            const int r = q >> 8, k = q & 255;
            if (k == c) colc[r] = v[u];       // (r, c), r >= c
            else if (r == c) colc[k] = v[u];  // (c, k), k < c
        }
        __syncthreads();
        const double d = 1.0 / colc[c];
#pragma unroll
        for (int u = 0; u < CT; u++) {
            int q = rk[u];
            asm volatile("" : "+v"(q));
            if (q < 0) continue;
            const int r = q >> 8, k = q & 255;
            double nv;
            if (r == c && k == c) nv = -d;
            else if (r == c) nv = colc[k] * d;
            else if (k == c) nv = colc[r] * d;
            else nv = v[u] - colc[r] * (colc[k] * d);
            v[u] = nv;
            asm volatile("" ::: "memory");  // one slot's LDS reads at a time (register pressure)
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < CT; u++) {
        if (rk[u] < 0) continue;
        const int r = rk[u] >> 8, k = rk[u] & 255;
        const float x = (float)(-v[u]);
        inv[(size_t)r * m + k] = x;
        inv[(size_t)k * m + r] = x;
    }
}

__global__ void __launch_bounds__(CB) k_amg_coarse_inverse(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                           const float* __restrict__ val, double* __restrict__ work,
                                                           float* __restrict__ inv) {
    coarse_inverse(n, rp, ci, val, work, inv);
}

// ---------------------------------------------------------------- cycle kernels

__device__ __forceinline__ V3f ld3v(const float* p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void st3v(float* p, int i, V3f a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }
__device__ __forceinline__ V3f mat3v(const float* m, V3f a) {
    return mk(m[0] * a.x + m[1] * a.y + m[2] * a.z, m[3] * a.x + m[4] * a.y + m[5] * a.z,
              m[6] * a.x + m[7] * a.y + m[8] * a.z);
}

// x = w D^-1 b
__global__ void __launch_bounds__(AB) k_amg_pre(int n, const float* __restrict__ Dinv, const float* __restrict__ b,
                                                float* __restrict__ x) {
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= n) return;
    st3v(x, i, AMG_OMEGA * mat3v(Dinv + 9 * (size_t)i, ld3v(b, i)));
}

// (A y)_i over level 0 (sliced ELL + contact CSR), y_j = x_j (+ e_agg(j) when PROLONG)
template <bool PROLONG>
__device__ __forceinline__ V3f row0(int i, const int* __restrict__ sliceOff, const int* __restrict__ sliceW,
                                    const int* __restrict__ cols, const float* __restrict__ vals,
                                    const int* __restrict__ cRowPtr, const int* __restrict__ cCol,
                                    const float* __restrict__ cVal, const float* __restrict__ x,
                                    const int* __restrict__ agg, const float* __restrict__ e) {
    const int s = i / SLICE, lane = i % SLICE;
    const int W = sliceW[s];
    const size_t off = (size_t)sliceOff[s];
    const int* cp = cols + off * SLICE + lane;
    const float* vp = vals + off * 9 * SLICE + lane;
    float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
    for (int k = 0; k < W; k++) {
        const int j = cp[k * SLICE];
        V3f xj = ld3v(x, j);
        if (PROLONG) xj = xj + ld3v(e, agg[j]);
        const float* a = vp + (size_t)k * 9 * SLICE;
        y0 += a[0 * SLICE] * xj.x + a[1 * SLICE] * xj.y + a[2 * SLICE] * xj.z;
        y1 += a[3 * SLICE] * xj.x + a[4 * SLICE] * xj.y + a[5 * SLICE] * xj.z;
        y2 += a[6 * SLICE] * xj.x + a[7 * SLICE] * xj.y + a[8 * SLICE] * xj.z;
    }
    for (int k = cRowPtr[i], ce = cRowPtr[i + 1]; k < ce; k++) {
        const int j = cCol[k];
        V3f xj = ld3v(x, j);
        if (PROLONG) xj = xj + ld3v(e, agg[j]);
        const float* a = cVal + (size_t)k * 9;
        y0 += a[0] * xj.x + a[1] * xj.y + a[2] * xj.z;
        y1 += a[3] * xj.x + a[4] * xj.y + a[5] * xj.z;
        y2 += a[6] * xj.x + a[7] * xj.y + a[8] * xj.z;
    }
    return mk(y0, y1, y2);
}

// (A y)_i over a block-CSR level
// Row i of a block-CSR level times x (PROLONG: times x + e gathered through agg),
// blocks in ascending order; the column, value and gather loads of four blocks
// are issued together (the sum is formed in the same order as one at a time).
template <bool PROLONG>
__device__ __forceinline__ V3f rowc(int i, const int* __restrict__ rp, const int* __restrict__ ci,
                                    const float* __restrict__ val, const float* __restrict__ x,
                                    const int* __restrict__ agg, const float* __restrict__ e) {
    float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
    const int k0 = rp[i], k1 = rp[i + 1];
    for (int k = k0; k < k1; k += 4) {
        int j[4];
        float a[4][9];
        V3f xj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int kk = min(k + u, k1 - 1);
            j[u] = ci[kk];
#pragma unroll
            for (int c = 0; c < 9; c++) a[u][c] = val[(size_t)kk * 9 + c];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) xj[u] = ld3v(x, j[u]);
        if (PROLONG) {
            int g[4];
#pragma unroll
            for (int u = 0; u < 4; u++) g[u] = agg[j[u]];
#pragma unroll
            for (int u = 0; u < 4; u++) xj[u] = xj[u] + ld3v(e, g[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (k + u < k1) {
                y0 += a[u][0] * xj[u].x + a[u][1] * xj[u].y + a[u][2] * xj[u].z;
                y1 += a[u][3] * xj[u].x + a[u][4] * xj[u].y + a[u][5] * xj[u].z;
                y2 += a[u][6] * xj[u].x + a[u][7] * xj[u].y + a[u][8] * xj[u].z;
            }
    }
    return mk(y0, y1, y2);
}

struct Level0 {
    const int *sliceOff, *sliceW, *cols, *cRowPtr, *cCol;
    const float *vals, *cVal, *Minv;
};

// t = b - A x on level 0
__global__ void __launch_bounds__(AB) k_amg_resid0(int nn, Level0 A, const float* __restrict__ b,
                                                   const float* __restrict__ x, float* __restrict__ t) {
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= nn) return;
    const V3f ax = row0<false>(i, A.sliceOff, A.sliceW, A.cols, A.vals, A.cRowPtr, A.cCol, A.cVal, x, nullptr, nullptr);
    st3v(t, i, ld3v(b, i) - ax);
}

// Pre-smoothing and residual of a block-CSR level in one pass: x = w D^-1 b
// (stored for the own row, formed on the fly for the gathered ones) and
// t = b - A x.  Coarse levels are small, so the extra gathers hit the cache and
// one launch is saved.
__device__ __forceinline__ void presid_row(int i, const int* __restrict__ rp, const int* __restrict__ ci,
                                           const float* __restrict__ val, const float* __restrict__ Dinv,
                                           const float* __restrict__ b, float* __restrict__ x, float* __restrict__ t) {
    float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
    const int k0 = rp[i], k1 = rp[i + 1];
    for (int k = k0; k < k1; k += 4) {  // four blocks' loads at a time, summed in order
        int j[4];
        float a[4][9];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int kk = min(k + u, k1 - 1);
            j[u] = ci[kk];
#pragma unroll
            for (int c = 0; c < 9; c++) a[u][c] = val[(size_t)kk * 9 + c];
        }
        V3f xj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) xj[u] = AMG_OMEGA * mat3v(Dinv + 9 * (size_t)j[u], ld3v(b, j[u]));
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (k + u < k1) {
                y0 += a[u][0] * xj[u].x + a[u][1] * xj[u].y + a[u][2] * xj[u].z;
                y1 += a[u][3] * xj[u].x + a[u][4] * xj[u].y + a[u][5] * xj[u].z;
                y2 += a[u][6] * xj[u].x + a[u][7] * xj[u].y + a[u][8] * xj[u].z;
            }
    }
    const V3f bi = ld3v(b, i);
    st3v(x, i, AMG_OMEGA * mat3v(Dinv + 9 * (size_t)i, bi));
    st3v(t, i, bi - mk(y0, y1, y2));
}

__global__ void __launch_bounds__(AB) k_amg_presid(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const float* __restrict__ val, const float* __restrict__ Dinv,
                                                   const float* __restrict__ b, float* __restrict__ x,
                                                   float* __restrict__ t) {
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i < n) presid_row(i, rp, ci, val, Dinv, b, x, t);
}

// bc_I = sum of t over the members of aggregate I (ascending)
__device__ __forceinline__ void restrict_row(int I, const int* __restrict__ mptr, const int* __restrict__ midx,
                                             const float* __restrict__ t, float* __restrict__ bc) {
    V3f s = mk(0.0f, 0.0f, 0.0f);
    const int k0 = mptr[I], k1 = mptr[I + 1];
    for (int k = k0; k < k1; k += 4) {  // four members' loads at a time, summed in order
        V3f m[4];
#pragma unroll
        for (int u = 0; u < 4; u++) m[u] = ld3v(t, midx[min(k + u, k1 - 1)]);
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (k + u < k1) s = s + m[u];
    }
    st3v(bc, I, s);
}

__global__ void __launch_bounds__(AB) k_amg_restrict(int nc, const int* __restrict__ mptr, const int* __restrict__ midx,
                                                     const float* __restrict__ t, float* __restrict__ bc) {
    const int I = blockIdx.x * AB + threadIdx.x;
    if (I < nc) restrict_row(I, mptr, midx, t, bc);
}

// x = inv b on the coarsest level: one workgroup of 16 waves, a wave per row,
// the lanes of a wave over the row (coalesced) and a shuffle reduction
constexpr int SB = 1024;
__device__ __forceinline__ void coarse_solve(int m, const float* __restrict__ inv, const float* __restrict__ b,
                                             float* __restrict__ x) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int r = wave; r < m; r += SB / 64) {
        float s = 0.0f;
        for (int k = lane; k < m; k += 64) s += inv[(size_t)r * m + k] * b[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) x[r] = s;
    }
}

// out = x' + w D^-1 (b - A x'), x' = x + e_agg, on a block-CSR level
__device__ __forceinline__ void post_row(int i, const int* __restrict__ rp, const int* __restrict__ ci,
                                         const float* __restrict__ val, const float* __restrict__ Dinv,
                                         const int* __restrict__ agg, const float* __restrict__ e,
                                         const float* __restrict__ b, const float* __restrict__ x,
                                         float* __restrict__ out) {
    const V3f ax = rowc<true>(i, rp, ci, val, x, agg, e);
    const V3f xi = ld3v(x, i) + ld3v(e, agg[i]);
    st3v(out, i, xi + AMG_OMEGA * mat3v(Dinv + 9 * (size_t)i, ld3v(b, i) - ax));
}

__global__ void __launch_bounds__(AB) k_amg_post(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                 const float* __restrict__ val, const float* __restrict__ Dinv,
                                                 const int* __restrict__ agg, const float* __restrict__ e,
                                                 const float* __restrict__ b, const float* __restrict__ x,
                                                 float* __restrict__ out) {
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i < n) post_row(i, rp, ci, val, Dinv, agg, e, b, x, out);
}

// ---- the small levels in one workgroup ----
// Levels of at most SMALL nodes cost one launch latency per kernel but little
// work; one workgroup runs their whole sub-cycle (pre-smooth + residual,
// restriction, coarsest solve, prolongation + post-smooth) with barriers
// between the phases, in the same order and arithmetic as the per-level kernels.
// Every vector of the sub-cycle lives in LDS (SMAXN nodes over all its levels),
// so a row's gathers are LDS reads, and each row issues the column and value
// loads of eight blocks at a time.
constexpr int SMALL = 512;  // measured: one workgroup is slower than per-level launches above this
constexpr int MAXSMALL = 8;
constexpr int SMAXN = 1024;
struct SmallLevel {
    int n, off;  // nodes, first node in the LDS vectors
    int nnzb, nNext;  // blocks; nodes of the next level (0 for the coarsest)
    const int *rp, *ci, *agg, *mptr, *midx;
    const float *val, *Dinv;
};
struct SmallCycle {
    int ns;  // levels in the sub-cycle; the last is the coarsest
    int m;   // 3 n of the coarsest
    const float* inv;
    // the first level's right-hand side is the restriction of the finer level's residual:
    // bc_I = sum of rt over its members rmidx[rmptr[I] .. rmptr[I + 1]) (ascending), nFine nodes
    int nFine;
    const int *rmptr, *rmidx;
    const float* rt;
    float* out0;      // its result (the coarsest's solution when it is the only level)
    // unit mode (op != nullptr): workgroup j runs the sub-cycle on the unit right-hand side e_j
    // of the first level instead of the restriction, and writes its result to column j of the
    // row-major dense operator op (dim x dim, dim = 3 n of the first level)
    float* op;
    int dim;
    SmallLevel L[MAXSMALL];
};

// sum over the blocks k0..k1-1 (ascending) of val_k v[ci_k], v an LDS vector
constexpr int RW = 8;  // blocks whose loads are in flight together
__device__ __forceinline__ V3f row_lds(int k0, int k1, const int* __restrict__ ci, const float* __restrict__ val,
                                       const float* v) {
    float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
    for (int k = k0; k < k1; k += RW) {
        int j[RW];
        float a[RW][9];
#pragma unroll
        for (int u = 0; u < RW; u++) {
            const int kk = min(k + u, k1 - 1);
            j[u] = ci[kk];
#pragma unroll
            for (int c = 0; c < 9; c++) a[u][c] = val[(size_t)kk * 9 + c];
        }
#pragma unroll
        for (int u = 0; u < RW; u++)
            if (k + u < k1) {
                const float* x = v + 3 * j[u];
                const float x0 = x[0], x1 = x[1], x2 = x[2];
                y0 += a[u][0] * x0 + a[u][1] * x1 + a[u][2] * x2;
                y1 += a[u][3] * x0 + a[u][4] * x1 + a[u][5] * x2;
                y2 += a[u][6] * x0 + a[u][7] * x1 + a[u][8] * x2;
            }
    }
    return mk(y0, y1, y2);
}
__device__ __forceinline__ V3f ld3s(const float* v, int i) { return mk(v[3 * i], v[3 * i + 1], v[3 * i + 2]); }
__device__ __forceinline__ void st3s(float* v, int i, V3f a) { v[3 * i] = a.x; v[3 * i + 1] = a.y; v[3 * i + 2] = a.z; }

// One load per 128-byte line of [p, p + bytes): the lines are in L2 when the phases read them.
__device__ __forceinline__ int touch(const void* p, size_t bytes, int tid) {
    int acc = 0;
    if (!p) return acc;
    for (size_t o = (size_t)tid * 128; o < bytes; o += (size_t)SB * 128) acc ^= *(const int*)((const char*)p + o);
    return acc;
}

__global__ void __launch_bounds__(SB) k_amg_small_cycle(SmallCycle Q) {
    __shared__ float vb[3 * SMAXN], vx[3 * SMAXN], vt[3 * SMAXN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool unit = Q.op != nullptr;
    // The sub-cycle's levels (about 0.25 MB) were last read an iteration ago; their lines
    // come back from HBM in one round of independent loads instead of one round per phase.
    if (!unit) {
        int acc = touch(Q.inv, sizeof(float) * (size_t)Q.m * Q.m, tid);
        acc ^= touch(Q.rmptr, sizeof(int) * (size_t)(Q.L[0].n + 1), tid);
        acc ^= touch(Q.rmidx, sizeof(int) * (size_t)Q.nFine, tid);
        acc ^= touch(Q.rt, sizeof(float) * 3 * (size_t)Q.nFine, tid);
        for (int l = 0; l + 1 < Q.ns; l++) {
            const SmallLevel& L = Q.L[l];
            acc ^= touch(L.rp, sizeof(int) * (size_t)(L.n + 1), tid);
            acc ^= touch(L.ci, sizeof(int) * (size_t)L.nnzb, tid);
            acc ^= touch(L.val, sizeof(float) * 9 * (size_t)L.nnzb, tid);
            acc ^= touch(L.Dinv, sizeof(float) * 9 * (size_t)L.n, tid);
            acc ^= touch(L.agg, sizeof(int) * (size_t)L.n, tid);
            acc ^= touch(L.mptr, sizeof(int) * (size_t)(L.nNext + 1), tid);
            acc ^= touch(L.midx, sizeof(int) * (size_t)L.n, tid);
        }
        asm volatile("" ::"v"(acc));
    }
    if (unit) {
        for (int e = tid; e < 3 * Q.L[0].n; e += SB) vb[3 * Q.L[0].off + e] = e == (int)blockIdx.x ? 1.0f : 0.0f;
    }
    // restriction into the first level (as k_amg_restrict: members in order)
    for (int I = tid; !unit && I < Q.L[0].n; I += SB) {
        V3f sum = mk(0.0f, 0.0f, 0.0f);
        const int k0 = Q.rmptr[I], k1 = Q.rmptr[I + 1];
        for (int k = k0; k < k1; k += 8) {
            int j[8];
#pragma unroll
            for (int u = 0; u < 8; u++) j[u] = Q.rmidx[min(k + u, k1 - 1)];
            V3f m[8];
#pragma unroll
            for (int u = 0; u < 8; u++) m[u] = ld3v(Q.rt, j[u]);
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (k + u < k1) sum = sum + m[u];
        }
        st3s(vb + 3 * Q.L[0].off, I, sum);
    }
    __syncthreads();
    for (int l = 0; l + 1 < Q.ns; l++) {
        const SmallLevel& L = Q.L[l];
        float *b = vb + 3 * L.off, *x = vx + 3 * L.off, *t = vt + 3 * L.off;
        for (int i = tid; i < L.n; i += SB) st3s(x, i, AMG_OMEGA * mat3v(L.Dinv + 9 * (size_t)i, ld3s(b, i)));
        __syncthreads();
        for (int i = tid; i < L.n; i += SB) st3s(t, i, ld3s(b, i) - row_lds(L.rp[i], L.rp[i + 1], L.ci, L.val, x));
        __syncthreads();
        const SmallLevel& N = Q.L[l + 1];
        float* nb = vb + 3 * N.off;
        for (int I = tid; I < N.n; I += SB) {  // members' ids eight at a time, summed in order
            V3f sum = mk(0.0f, 0.0f, 0.0f);
            const int k0 = L.mptr[I], k1 = L.mptr[I + 1];
            for (int k = k0; k < k1; k += 8) {
                int j[8];
#pragma unroll
                for (int u = 0; u < 8; u++) j[u] = L.midx[min(k + u, k1 - 1)];
#pragma unroll
                for (int u = 0; u < 8; u++)
                    if (k + u < k1) sum = sum + ld3s(t, j[u]);
            }
            st3s(nb, I, sum);
        }
        __syncthreads();
    }
    {
        // x = inv b on the coarsest level.  inv is stored symmetric, so column k is row k and
        // lanes read it coalesced along r; each wave sums its own run of k for every r, and
        // the 16 partial sums of a row are added in wave order.
        const SmallLevel& C = Q.L[Q.ns - 1];
        const float* cb = vb + 3 * C.off;
        __shared__ float part[SB / 64][3 * AMG_COARSEST];
        const int m = Q.m, kc = (m + SB / 64 - 1) / (SB / 64);
        const int k0 = wave * kc, k1 = min(m, k0 + kc);
        for (int r = lane; r < m; r += 64) {
            float sum = 0.0f;
            for (int k = k0; k < k1; k++) sum += Q.inv[(size_t)k * m + r] * cb[k];
            part[wave][r] = sum;
        }
        __syncthreads();
        for (int r = tid; r < m; r += SB) {
            float sum = part[0][r];
#pragma unroll
            for (int w = 1; w < SB / 64; w++) sum += part[w][r];
            vx[3 * C.off + r] = sum;
        }
        __syncthreads();
    }
    // the result of every level is left in vx
    for (int l = Q.ns - 2; l >= 0; l--) {
        const SmallLevel& L = Q.L[l];
        const float* e = vx + 3 * Q.L[l + 1].off;
        float *b = vb + 3 * L.off, *x = vx + 3 * L.off, *xp = vt + 3 * L.off;
        for (int i = tid; i < L.n; i += SB) st3s(xp, i, ld3s(x, i) + ld3s(e, L.agg[i]));
        __syncthreads();
        for (int i = tid; i < L.n; i += SB) {
            const V3f ax = row_lds(L.rp[i], L.rp[i + 1], L.ci, L.val, xp);
            st3s(x, i, ld3s(xp, i) + AMG_OMEGA * mat3v(L.Dinv + 9 * (size_t)i, ld3s(b, i) - ax));
        }
        __syncthreads();
    }
    if (unit) {
        for (int e = tid; e < Q.dim; e += SB) Q.op[(size_t)e * Q.dim + blockIdx.x] = vx[3 * Q.L[0].off + e];
        return;
    }
    for (int e = tid; e < 3 * Q.L[0].n; e += SB) Q.out0[e] = vx[3 * Q.L[0].off + e];
}

// The small levels' sub-cycle applied as its dense operator: every workgroup forms the first
// small level's right-hand side (the restriction of the finer residual, members summed in
// order, as the sub-cycle does) in LDS, then each wave produces one row of op . b, its lanes
// striding along the row and the row sum reduced over the wave.
constexpr int DB = 256;  // threads per workgroup, one row per wave
__global__ void __launch_bounds__(DB) k_amg_small_dense(int n0, int dim, const int* __restrict__ rmptr,
                                                        const int* __restrict__ rmidx, const float* __restrict__ rt,
                                                        const float* __restrict__ op, float* __restrict__ out) {
    __shared__ float b[3 * SMAXN];
    constexpr int NK = SMAXN * 3 / 64;  // row entries per lane at most
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = blockIdx.x * (DB / 64) + wave;
    const int nk = (dim + 63) / 64;
    // the row's loads go out first: they do not depend on the right-hand side
    float a[NK];
    const float* row = op + (size_t)min(r, dim - 1) * dim;
#pragma unroll
    for (int u = 0; u < NK; u++) a[u] = u < nk && lane + 64 * u < dim ? row[lane + 64 * u] : 0.0f;
    for (int I = tid; I < n0; I += DB) {  // members' ids, then their values, eight at a time
        V3f sum = mk(0.0f, 0.0f, 0.0f);
        const int k0 = rmptr[I], k1 = rmptr[I + 1];
        for (int k = k0; k < k1; k += 8) {
            int j[8];
#pragma unroll
            for (int u = 0; u < 8; u++) j[u] = rmidx[min(k + u, k1 - 1)];
            V3f m[8];
#pragma unroll
            for (int u = 0; u < 8; u++) m[u] = ld3v(rt, j[u]);
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (k + u < k1) sum = sum + m[u];
        }
        st3s(b, I, sum);
    }
    __syncthreads();
    if (r >= dim) return;
    float acc = 0.0f;
#pragma unroll
    for (int u = 0; u < NK; u++)
        if (u < nk && lane + 64 * u < dim) acc += a[u] * b[lane + 64 * u];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) out[r] = acc;
}

// AP = A_static P of level 0, per step: each AP slot sums the level-0 blocks of
// its row whose column lies in its aggregate (slot order, fixed).
__device__ __forceinline__ void ap_build_row(int i, const int* __restrict__ sliceOff, const int* __restrict__ sliceW,
                                             const float* __restrict__ vals, const uint8_t* __restrict__ target,
                                             const int* __restrict__ apSliceOff, const int* __restrict__ apSliceW,
                                             float* __restrict__ apVals) {
    const int s = i / SLICE, lane = i % SLICE;
    const int W = sliceW[s], WA = apSliceW[s];
    const size_t off = (size_t)sliceOff[s], aoff = (size_t)apSliceOff[s];
    for (int m = 0; m < WA; m++) {
        float a[9];
#pragma unroll
        for (int c = 0; c < 9; c++) a[c] = 0.0f;
        for (int k = 0; k < W; k++) {
            if (target[(off + k) * SLICE + lane] != m) continue;
            const float* v = vals + (off + k) * 9 * SLICE + lane;
#pragma unroll
            for (int c = 0; c < 9; c++) a[c] += v[c * SLICE];
        }
        float* o = apVals + (aoff + m) * 9 * SLICE + lane;
#pragma unroll
        for (int c = 0; c < 9; c++) o[c * SLICE] = a[c];
    }
}

__global__ void __launch_bounds__(AB) k_amg_ap_build(int nn, const int* __restrict__ sliceOff,
                                                     const int* __restrict__ sliceW, const float* __restrict__ vals,
                                                     const uint8_t* __restrict__ target,
                                                     const int* __restrict__ apSliceOff,
                                                     const int* __restrict__ apSliceW, float* __restrict__ apVals) {
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i < nn) ap_build_row(i, sliceOff, sliceW, vals, target, apSliceOff, apSliceW, apVals);
}

// Level 0 post-smoothing from the stored residual t = r - A x of the same cycle:
// r - A (x + P e) = t - (A P) e, with A P = AP (static part) plus the contact
// blocks gathered through agg.  z = x + e_agg + w Minv (t - (A P) e); one partial
// of r.z per block.
__global__ void __launch_bounds__(AB) k_amg_post0ap(int nn, Level0 A, const int* __restrict__ apSliceOff,
                                                    const int* __restrict__ apSliceW, const int* __restrict__ apCols,
                                                    const float* __restrict__ apVals, const int* __restrict__ agg,
                                                    const float* __restrict__ e, const float* __restrict__ r,
                                                    const float* __restrict__ x, const float* __restrict__ t,
                                                    float4* __restrict__ z, float* __restrict__ partRZ) {
    __shared__ float lds[AB / 64];
    float prz = 0.0f;
    for (int i = blockIdx.x * AB + threadIdx.x; i < nn; i += gridDim.x * AB) {
        const int s = i / SLICE, lane = i % SLICE;
        const int WA = apSliceW[s];
        const size_t aoff = (size_t)apSliceOff[s];
        float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
        for (int m = 0; m < WA; m++) {
            const V3f ej = ld3v(e, apCols[(aoff + m) * SLICE + lane]);
            const float* a = apVals + (aoff + m) * 9 * SLICE + lane;
            y0 += a[0 * SLICE] * ej.x + a[1 * SLICE] * ej.y + a[2 * SLICE] * ej.z;
            y1 += a[3 * SLICE] * ej.x + a[4 * SLICE] * ej.y + a[5 * SLICE] * ej.z;
            y2 += a[6 * SLICE] * ej.x + a[7 * SLICE] * ej.y + a[8 * SLICE] * ej.z;
        }
        for (int k = A.cRowPtr[i], ce = A.cRowPtr[i + 1]; k < ce; k++) {
            const V3f ej = ld3v(e, agg[A.cCol[k]]);
            const float* a = A.cVal + (size_t)k * 9;
            y0 += a[0] * ej.x + a[1] * ej.y + a[2] * ej.z;
            y1 += a[3] * ej.x + a[4] * ej.y + a[5] * ej.z;
            y2 += a[6] * ej.x + a[7] * ej.y + a[8] * ej.z;
        }
        const V3f xi = ld3v(x, i) + ld3v(e, agg[i]);
        const V3f zi = xi + AMG_OMEGA * mat3v(A.Minv + 9 * (size_t)i, ld3v(t, i) - mk(y0, y1, y2));
        z[i] = make_float4(zi.x, zi.y, zi.z, 0.0f);
        prz += dot(ld3v(r, i), zi);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) prz += __shfl_xor(prz, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = prz;
    __syncthreads();
    if (threadIdx.x == 0) partRZ[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// Level 0 with the learned smoother, before its sweeps: the residual after the coarse
// correction, res = t - A_static P e - C P e (into t), and x' = x + P e (into x).
template <bool STREAM>
__global__ void __launch_bounds__(AB) k_amg_post0ap_res(int nn, Level0 A, const int* __restrict__ apSliceOff,
                                                        const int* __restrict__ apSliceW,
                                                        const int* __restrict__ apCols,
                                                        const float* __restrict__ apVals, const int* __restrict__ agg,
                                                        const float* __restrict__ e, float* __restrict__ x,
                                                        float* __restrict__ t) {
    const int i = blockIdx.x * AB + threadIdx.x;
    if (i >= nn) return;
    const int s = i / SLICE, lane = i % SLICE;
    const int WA = apSliceW[s];
    const size_t aoff = (size_t)apSliceOff[s];
    float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
    for (int m = 0; m < WA; m++) {
        const int* cp = apCols + (aoff + m) * SLICE + lane;
        const V3f ej = ld3v(e, STREAM ? __builtin_nontemporal_load(cp) : *cp);
        const float* ap = apVals + (aoff + m) * 9 * SLICE + lane;
        float a[9];
#pragma unroll
        for (int c = 0; c < 9; c++) a[c] = STREAM ? __builtin_nontemporal_load(ap + c * SLICE) : ap[c * SLICE];
        y0 += a[0] * ej.x + a[1] * ej.y + a[2] * ej.z;
        y1 += a[3] * ej.x + a[4] * ej.y + a[5] * ej.z;
        y2 += a[6] * ej.x + a[7] * ej.y + a[8] * ej.z;
    }
    for (int k = A.cRowPtr[i], ce = A.cRowPtr[i + 1]; k < ce; k++) {
        const V3f ej = ld3v(e, agg[A.cCol[k]]);
        const float* a = A.cVal + (size_t)k * 9;
        y0 += a[0] * ej.x + a[1] * ej.y + a[2] * ej.z;
        y1 += a[3] * ej.x + a[4] * ej.y + a[5] * ej.z;
        y2 += a[6] * ej.x + a[7] * ej.y + a[8] * ej.z;
    }
    st3v(x, i, ld3v(x, i) + ld3v(e, agg[i]));
    st3v(t, i, ld3v(t, i) - mk(y0, y1, y2));
}

}  // namespace

// Host topology of cloth c's hierarchy and its device index arrays (once; the mesh never changes).
static void amg_build(cs_ctx* C, int c) {
    ClothSys& S = C->sys[c];
    AmgSys& G = S.amg;
    const MeshInfo& M = C->meshes[S.mesh];
    hipStream_t st = C->stream;
    const int nn = M.nn;
    // level-0 graph: nodes adjacent through a face
    std::vector<std::vector<int>> adj(nn);
    for (int f = M.f0; f < M.f0 + M.nf; f++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                if (a != b) adj[C->hfn[3 * f + a] - M.n0].push_back(C->hfn[3 * f + b] - M.n0);
    std::vector<int> gp(nn + 1, 0), gi;
    for (int i = 0; i < nn; i++) {
        std::sort(adj[i].begin(), adj[i].end());
        adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
        gi.insert(gi.end(), adj[i].begin(), adj[i].end());
        gp[i + 1] = (int)gi.size();
    }
    std::vector<int> sizes{nn};
    std::vector<std::vector<int>> aggs;
    int n = nn;
    while (n > AMG_COARSEST) {
        std::vector<int> agg, cp, cidx;
        int na = 0;
        aggregate(gp, gi, agg, na);
        if (na >= n) break;
        coarse_graph(gp, gi, agg, na, cp, cidx);
        aggs.push_back(agg);
        sizes.push_back(na);
        gp.swap(cp);
        gi.swap(cidx);
        n = na;
    }
    G.built = true;
    G.nl = (int)sizes.size();
    if (sizes.back() > AMG_COARSEST || G.nl == 1) {  // too small, or cannot be coarsened: not usable
        G.nl = 0;
        return;
    }
    G.L.clear();
    G.L.resize(G.nl);
    for (int l = 0; l < G.nl; l++) G.L[l].n = sizes[l];
    // level-0 static pattern with the value offset of each block in the sliced ELL
    std::vector<int> frp = S.hRowPtr, fci = S.hCol, fid(S.hCol.size());
    for (int r = 0; r < nn; r++) {
        const int s = r / SLICE, lane = r % SLICE;
        for (int k = S.hRowPtr[r]; k < S.hRowPtr[r + 1]; k++)
            fid[k] = (int)(((size_t)S.hSliceOff[s] + (k - S.hRowPtr[r])) * 9 * SLICE + lane);
    }
    for (int l = 0; l + 1 < G.nl; l++) {
        const std::vector<int>& agg = aggs[l];
        const int nf = sizes[l], nc = sizes[l + 1];
        // restriction lists and the prolongation map
        std::vector<int> mptr(nc + 1, 0), midx(nf);
        for (int i = 0; i < nf; i++) mptr[agg[i] + 1]++;
        for (int a = 0; a < nc; a++) mptr[a + 1] += mptr[a];
        std::vector<int> fill(mptr.begin(), mptr.end() - 1);
        for (int i = 0; i < nf; i++) midx[fill[agg[i]]++] = i;
        upload_vec(G.L[l].agg, agg, st);
        upload_vec(G.L[l].mptr, mptr, st);
        upload_vec(G.L[l].midx, midx, st);
        // Galerkin pattern of level l+1 with contributor lists (fine order: row, then slot)
        std::map<uint64_t, std::vector<int>> blocks;
        for (int i = 0; i < nf; i++)
            for (int k = frp[i]; k < frp[i + 1]; k++)
                blocks[(uint64_t)agg[i] * (uint64_t)nc + (uint64_t)agg[fci[k]]].push_back(fid[k]);
        std::vector<int> crp(nc + 1, 0), cci, cptr(1, 0), cidx, diag(nc, -1), cid;
        for (const auto& kv : blocks) {
            const int I = (int)(kv.first / nc), J = (int)(kv.first % nc);
            crp[I + 1]++;
            if (I == J) diag[I] = (int)cci.size();
            cid.push_back((int)cci.size());
            cci.push_back(J);
            cidx.insert(cidx.end(), kv.second.begin(), kv.second.end());
            cptr.push_back((int)cidx.size());
        }
        for (int a = 0; a < nc; a++) crp[a + 1] += crp[a];
        AmgLevel& Lc = G.L[l + 1];
        Lc.nnzb = (int)cci.size();
        upload_vec(Lc.rp, crp, st);
        upload_vec(Lc.ci, cci, st);
        upload_vec(Lc.diag, diag, st);
        upload_vec(Lc.cptr, cptr, st);
        upload_vec(Lc.cidx, cidx, st);
        Lc.val.as<float>(9 * (size_t)std::max(1, Lc.nnzb));
        Lc.Dinv.as<float>(9 * (size_t)nc);
        frp.swap(crp);
        fci.swap(cci);
        fid.swap(cid);
    }
    // level-0 A P: per row, one slot per distinct aggregate of its columns (ascending)
    {
        const std::vector<int>& agg = aggs[0];
        const int ns = S.nslices;
        std::vector<uint8_t> target((size_t)S.totalSlots * SLICE, 255);
        std::vector<int> apOff(ns + 1, 0), apW(ns, 0);
        std::vector<std::vector<int>> rowAggs(nn);
        for (int r = 0; r < nn; r++) {
            std::vector<int>& v = rowAggs[r];
            for (int k = S.hRowPtr[r]; k < S.hRowPtr[r + 1]; k++) v.push_back(agg[S.hCol[k]]);
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            if (v.size() > 254) {  // a row this long has no uint8 slot index: not usable
                G.nl = 0;
                return;
            }
            apW[r / SLICE] = std::max(apW[r / SLICE], (int)v.size());
        }
        for (int s = 0; s < ns; s++) apOff[s + 1] = apOff[s] + apW[s];
        std::vector<int> apCols((size_t)std::max(1, apOff[ns]) * SLICE, 0);
        for (int r = 0; r < nn; r++) {
            const int s = r / SLICE, lane = r % SLICE;
            const std::vector<int>& v = rowAggs[r];
            for (int m = 0; m < apW[s]; m++)
                apCols[((size_t)apOff[s] + m) * SLICE + lane] = m < (int)v.size() ? v[m] : agg[r];
            for (int k = S.hRowPtr[r]; k < S.hRowPtr[r + 1]; k++) {
                const int m = (int)(std::lower_bound(v.begin(), v.end(), agg[S.hCol[k]]) - v.begin());
                target[((size_t)S.hSliceOff[s] + (k - S.hRowPtr[r])) * SLICE + lane] = (uint8_t)m;
            }
        }
        upload_vec(G.apTarget, target, st);
        upload_vec(G.apSliceOff, apOff, st);
        upload_vec(G.apSliceW, apW, st);
        upload_vec(G.apCols, apCols, st);
        G.apVals.as<float>((size_t)std::max(1, apOff[ns]) * SLICE * 9);
        fill_f32(G.apVals.ptr<float>(), 0.0f, (long)std::max(1, apOff[ns]) * SLICE * 9, st);
    }
    for (int l = 0; l < G.nl; l++) {
        AmgLevel& L = G.L[l];
        if (l > 0) L.b.as<float>(3 * (size_t)L.n);
        L.x.as<float>(3 * (size_t)L.n);
        L.t.as<float>(3 * (size_t)L.n);
    }
    const int m = 3 * sizes.back();
    G.work.as<double>(2 * (size_t)m * m);
    G.inv.as<float>((size_t)m * m);
}

// Number of levels of cloth c's hierarchy (sizes[l] filled up to cap); 0 when the
// cloth cannot use the preconditioner (<= 64 nodes, or no coarsening possible).
int amg_levels(cs_ctx* C, int c, int* sizes, int cap) {
    AmgSys& G = C->sys[c].amg;
    if (!G.built) amg_build(C, c);
    for (int l = 0; l < G.nl && l < cap; l++) sizes[l] = G.L[l].n;
    return G.nl;
}

// The first of the small levels (at most SMALL nodes, their vectors within SMAXN nodes, at
// most MAXSMALL of them), which run as one sub-cycle.
static int small_first(const AmgSys& G) {
    int l0 = G.nl - 1;
    for (int l = 1; l < G.nl; l++)
        if (G.L[l].n <= SMALL) { l0 = l; break; }
    l0 = std::max(l0, G.nl - MAXSMALL);
    for (;;) {  // the small levels' vectors must fit the workgroup's LDS
        int tot = 0;
        for (int l = l0; l < G.nl; l++) tot += G.L[l].n;
        if (tot <= SMAXN || l0 == G.nl - 1) break;
        l0++;
    }
    return l0;
}

// The sub-cycle's arguments for the small levels l0.. (the first restricts level l0 - 1's residual).
static SmallCycle small_args(AmgSys& G, int l0) {
    SmallCycle Q;
    Q.ns = G.nl - l0;
    Q.m = 3 * G.L[G.nl - 1].n;
    Q.inv = G.inv.ptr<float>();
    Q.nFine = G.L[l0 - 1].n;
    Q.rmptr = G.L[l0 - 1].mptr.ptr<int>();
    Q.rmidx = G.L[l0 - 1].midx.ptr<int>();
    Q.rt = G.L[l0 - 1].t.ptr<float>();
    Q.out0 = l0 == G.nl - 1 ? G.L[l0].x.ptr<float>() : G.L[l0].t.ptr<float>();
    Q.op = nullptr;
    Q.dim = 3 * G.L[l0].n;
    int off = 0;
    for (int k = 0; k < Q.ns; k++) {
        AmgLevel& L = G.L[l0 + k];
        const bool last = l0 + k == G.nl - 1;
        SmallLevel& q = Q.L[k];
        q.n = L.n;
        q.off = off;
        q.nnzb = L.nnzb;
        q.nNext = last ? 0 : G.L[l0 + k + 1].n;
        off += L.n;
        q.rp = L.rp.ptr<int>();
        q.ci = L.ci.ptr<int>();
        q.val = L.val.ptr<float>();
        q.Dinv = L.Dinv.ptr<float>();
        q.agg = last ? nullptr : L.agg.ptr<int>();
        q.mptr = last ? nullptr : L.mptr.ptr<int>();
        q.midx = last ? nullptr : L.midx.ptr<int>();
    }
    return Q;
}

// Per-solve numeric setup: Galerkin operators of the static system, their diagonal inverses,
// the coarsest dense inverse, and the small levels' sub-cycle as a dense operator (the
// sub-cycle run on every unit right-hand side at once, one workgroup each): the cycles of the
// solve then apply it as one matrix-vector product spread over the chip instead of running the
// sub-cycle's dependent phases in one workgroup.
void amg_prepare(cs_ctx* C, int c, bool fork) {
    ClothSys& S = C->sys[c];
    AmgSys& G = S.amg;
    if (!G.built) amg_build(C, c);
    hipStream_t s = C->stream;
    // a previous fork whose caller failed before its join may still be running on the side
    // stream: the operators below are rewritten only after it has finished
    CS_HIP(hipStreamWaitEvent(s, C->evJoin, 0));
    for (int l = 1; l < G.nl; l++) {
        AmgLevel& L = G.L[l];
        if (l == 1)
            hipLaunchKernelGGL(k_amg_galerkin<true>, dim3(cdiv(L.nnzb, AB)), dim3(AB), 0, s, L.nnzb, L.cptr.ptr<int>(),
                               L.cidx.ptr<int>(), S.vals.ptr<float>(), L.val.ptr<float>());
        else
            hipLaunchKernelGGL(k_amg_galerkin<false>, dim3(cdiv(L.nnzb, AB)), dim3(AB), 0, s, L.nnzb,
                               L.cptr.ptr<int>(), L.cidx.ptr<int>(), G.L[l - 1].val.ptr<float>(), L.val.ptr<float>());
        if (l == 1 && S.nContact > 0)
            hipLaunchKernelGGL(k_amg_contact_diag, dim3(cdiv(L.n, AB)), dim3(AB), 0, s, L.n, G.L[0].mptr.ptr<int>(),
                               G.L[0].midx.ptr<int>(), L.diag.ptr<int>(), S.cDiag.ptr<int>(), S.cVal.ptr<float>(),
                               L.val.ptr<float>());
        CS_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_amg_dinv, dim3(cdiv(L.n, AB)), dim3(AB), 0, s, L.n, L.diag.ptr<int>(), L.val.ptr<float>(),
                           L.Dinv.ptr<float>());
        CS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_amg_ap_build, dim3(cdiv(S.nn, AB)), dim3(AB), 0, s, S.nn, S.sliceOff.ptr<int>(),
                       S.sliceW.ptr<int>(), S.vals.ptr<float>(), G.apTarget.ptr<uint8_t>(), G.apSliceOff.ptr<int>(),
                       G.apSliceW.ptr<int>(), G.apVals.ptr<float>());
    CS_LAUNCH_CHECK();
    // The coarsest inverse and the small levels' operator occupy one workgroup and then one
    // workgroup per unit vector for about 0.6 ms at 513^2; with `fork` they run on the side
    // stream while the caller's next setup (the network's inference) runs on the main stream,
    // and amg_join must precede the first cycle.
    if (fork) {
        CS_HIP(hipEventRecord(C->evFork, s));
        CS_HIP(hipStreamWaitEvent(C->side, C->evFork, 0));
        s = C->side;
    }
    const AmgLevel& Lc = G.L[G.nl - 1];
    hipLaunchKernelGGL(k_amg_coarse_inverse, dim3(1), dim3(CB), 0, s, Lc.n, Lc.rp.ptr<int>(), Lc.ci.ptr<int>(),
                       Lc.val.ptr<float>(), G.work.ptr<double>(), G.inv.ptr<float>());
    CS_LAUNCH_CHECK();
    const int l0 = small_first(G);
    SmallCycle Q = small_args(G, l0);
    G.smallL0 = l0;
    G.smallDim = Q.dim;
    Q.op = G.smallOp.as<float>((size_t)Q.dim * Q.dim);
    hipLaunchKernelGGL(k_amg_small_cycle, dim3(Q.dim), dim3(SB), 0, s, Q);
    CS_LAUNCH_CHECK();
    if (fork) CS_HIP(hipEventRecord(C->evJoin, s));
}

// The main stream waits for the setup amg_prepare forked to the side stream.
void amg_join(cs_ctx* C) { CS_HIP(hipStreamWaitEvent(C->stream, C->evJoin, 0)); }

// z = M^-1 r (one V(1,1) cycle); partRZ[b] = block b's share of r.z for the
// nb blocks of the PCG grid (nb * 256 threads, rows in grid-stride order).
float* amg_level0_x(cs_ctx* C, int c) { return C->sys[c].amg.L[0].x.ptr<float>(); }

void amg_apply(cs_ctx* C, int c, const float* r, float4* z, float* partRZ, int nb, bool presmoothed, bool learned) {
    ClothSys& S = C->sys[c];
    AmgSys& G = S.amg;
    hipStream_t s = C->stream;
    Level0 A0;
    A0.sliceOff = S.sliceOff.ptr<int>();
    A0.sliceW = S.sliceW.ptr<int>();
    A0.cols = S.cols.ptr<int>();
    A0.vals = S.vals.ptr<float>();
    A0.cRowPtr = S.cRowPtr.ptr<int>();
    A0.cCol = S.nContact > 0 ? S.cCol.ptr<int>() : S.cRowPtr.ptr<int>();
    A0.cVal = S.nContact > 0 ? S.cVal.ptr<float>() : S.vals.ptr<float>();
    A0.Minv = S.Minv.ptr<float>();
    // levels l0.. (the first with at most SMALL nodes, l0 >= 1): their sub-cycle is applied as
    // the dense operator amg_prepare built
    const int l0 = G.smallL0;
    // down: smooth, residual, restrict
    for (int l = 0; l < l0; l++) {
        AmgLevel& L = G.L[l];
        const float* b = l == 0 ? r : L.b.ptr<float>();
        const float* Dinv = l == 0 ? S.Minv.ptr<float>() : L.Dinv.ptr<float>();
        const int g = cdiv(L.n, AB);
        if (l == 0) {
            if (!presmoothed) hipLaunchKernelGGL(k_amg_pre, dim3(g), dim3(AB), 0, s, L.n, Dinv, b, L.x.ptr<float>());
            hipLaunchKernelGGL(k_amg_resid0, dim3(g), dim3(AB), 0, s, L.n, A0, b, L.x.ptr<float>(), L.t.ptr<float>());
        } else {
            hipLaunchKernelGGL(k_amg_presid, dim3(g), dim3(AB), 0, s, L.n, L.rp.ptr<int>(), L.ci.ptr<int>(),
                               L.val.ptr<float>(), Dinv, b, L.x.ptr<float>(), L.t.ptr<float>());
        }
        AmgLevel& N = G.L[l + 1];
        if (l + 1 < l0)  // the small-level cycle restricts into level l0 itself
            hipLaunchKernelGGL(k_amg_restrict, dim3(cdiv(N.n, AB)), dim3(AB), 0, s, N.n, L.mptr.ptr<int>(),
                               L.midx.ptr<int>(), L.t.ptr<float>(), N.b.ptr<float>());
    }
    {
        const AmgLevel& F = G.L[l0 - 1];
        float* out0 = l0 == G.nl - 1 ? G.L[l0].x.ptr<float>() : G.L[l0].t.ptr<float>();
        hipLaunchKernelGGL(k_amg_small_dense, dim3(cdiv(G.smallDim, DB / 64)), dim3(DB), 0, s, G.L[l0].n, G.smallDim,
                           F.mptr.ptr<int>(), F.midx.ptr<int>(), F.t.ptr<float>(), G.smallOp.ptr<float>(), out0);
    }
    // up: prolongate and post-smooth; the result of level l is in L[l].t (coarsest: x)
    for (int l = l0 - 1; l >= 0; l--) {
        AmgLevel& L = G.L[l];
        const AmgLevel& N = G.L[l + 1];
        const float* e = (l + 1 == G.nl - 1) ? N.x.ptr<float>() : N.t.ptr<float>();
        if (l == 0 && learned) {
            hipLaunchKernelGGL((learned ? k_amg_post0ap_res<true> : k_amg_post0ap_res<false>), dim3(cdiv(L.n, AB)),
                               dim3(AB), 0, s, L.n, A0,
                               G.apSliceOff.ptr<int>(), G.apSliceW.ptr<int>(), G.apCols.ptr<int>(),
                               G.apVals.ptr<float>(), L.agg.ptr<int>(), e, L.x.ptr<float>(), L.t.ptr<float>());
            lpc_smooth_post(C, c, L.t.ptr<float>(), L.x.ptr<float>(), r, z, partRZ, nb);
        } else if (l == 0)
            hipLaunchKernelGGL(k_amg_post0ap, dim3(nb), dim3(AB), 0, s, L.n, A0, G.apSliceOff.ptr<int>(),
                               G.apSliceW.ptr<int>(), G.apCols.ptr<int>(), G.apVals.ptr<float>(), L.agg.ptr<int>(), e,
                               r, L.x.ptr<float>(), L.t.ptr<float>(), z, partRZ);
        else
            hipLaunchKernelGGL(k_amg_post, dim3(cdiv(L.n, AB)), dim3(AB), 0, s, L.n, L.rp.ptr<int>(), L.ci.ptr<int>(),
                               L.val.ptr<float>(), L.Dinv.ptr<float>(), L.agg.ptr<int>(), e, L.b.ptr<float>(),
                               L.x.ptr<float>(), L.t.ptr<float>());
    }
    CS_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- batched hierarchies
// Independent instances (CS_FLAG_INSTANCES) solve together.  Every cloth keeps its own
// hierarchy (amg_build); one launch covers level l of every cloth that has one, so a cycle
// costs a few launches per level whatever the number of cloths.  Job = (cloth, level l >= 1);
// a block map gives each workgroup its (job, block).  Each cloth's coarsest level is solved
// with its own dense inverse (one workgroup per cloth), and every level's result, the
// coarsest's too, is left in its t.  Level 0 (the cloth's own system, sliced ELL) belongs to
// the batched PCG kernels.

namespace {

struct BLev {
    int n, nnzb, coarsest, pad;
    const int *rp, *ci, *diag, *cptr, *cidx;
    const float* fine;         // Galerkin source: the level-0 sliced-ELL values (l = 1) or level l-1's blocks
    float *val, *Dinv;
    const int *fmptr, *fmidx;  // level l-1's members of each node here
    const float* ft;           // level l-1's residual (the restriction's source)
    float *b, *x, *t;          // right-hand side, pre-smoothed iterate, residual / result
    const int* agg;            // here -> level l+1 (not coarsest)
    const float* e;            // level l+1's result (not coarsest)
    const int* cDiag;          // level 1: each level-0 node's contact diagonal block (-1: none)
    const float* cVal;         // level 1: the cloth's contact blocks (null: none)
    double* work;              // coarsest: Gauss-Jordan workspace
    float* inv;                // coarsest: dense inverse
};

struct BAp {  // level 0 of one cloth: A_static P
    int nn, pad;
    const int *sliceOff, *sliceW, *apSliceOff, *apSliceW;
    const float* vals;
    const uint8_t* target;
    float* apVals;
};

__device__ __forceinline__ int2 job_block(const int2* __restrict__ map) {
    const int2 m = map[blockIdx.x];
    return make_int2(__builtin_amdgcn_readfirstlane(m.x), __builtin_amdgcn_readfirstlane(m.y));
}

template <bool FROM_ELL>
__global__ void __launch_bounds__(AB) k_bamg_galerkin(const BLev* __restrict__ jobs, const int2* __restrict__ map) {
    const int2 m = job_block(map);
    const BLev& L = jobs[m.x];
    const int k = m.y * AB + threadIdx.x;
    if (k < L.nnzb) galerkin_block<FROM_ELL>(k, L.cptr, L.cidx, L.fine, L.val);
}

__global__ void __launch_bounds__(AB) k_bamg_contact_diag(const BLev* __restrict__ jobs, const int2* __restrict__ map) {
    const int2 m = job_block(map);
    const BLev& L = jobs[m.x];
    const int a = m.y * AB + threadIdx.x;
    if (a < L.n && L.cVal) contact_diag_row(a, L.fmptr, L.fmidx, L.diag, L.cDiag, L.cVal, L.val);
}

__global__ void __launch_bounds__(AB) k_bamg_dinv(const BLev* __restrict__ jobs, const int2* __restrict__ map) {
    const int2 m = job_block(map);
    const BLev& L = jobs[m.x];
    const int i = m.y * AB + threadIdx.x;
    if (i < L.n) dinv_row(i, L.diag, L.val, L.Dinv);
}

__global__ void __launch_bounds__(AB) k_bamg_ap_build(const BAp* __restrict__ jobs, const int2* __restrict__ map) {
    const int2 m = job_block(map);
    const BAp& J = jobs[m.x];
    const int i = m.y * AB + threadIdx.x;
    if (i < J.nn) ap_build_row(i, J.sliceOff, J.sliceW, J.vals, J.target, J.apSliceOff, J.apSliceW, J.apVals);
}

__global__ void __launch_bounds__(CB) k_bamg_coarse_inverse(const BLev* __restrict__ jobs,
                                                            const int2* __restrict__ map) {
    const BLev& L = jobs[job_block(map).x];
    coarse_inverse(L.n, L.rp, L.ci, L.val, L.work, L.inv);
}

__global__ void __launch_bounds__(AB) k_bamg_restrict(const BLev* __restrict__ jobs, const int2* __restrict__ map) {
    const int2 m = job_block(map);
    const BLev& L = jobs[m.x];
    const int I = m.y * AB + threadIdx.x;
    if (I < L.n) restrict_row(I, L.fmptr, L.fmidx, L.ft, L.b);
}

__global__ void __launch_bounds__(AB) k_bamg_presid(const BLev* __restrict__ jobs, const int2* __restrict__ map) {
    const int2 m = job_block(map);
    const BLev& L = jobs[m.x];
    const int i = m.y * AB + threadIdx.x;
    if (i < L.n) presid_row(i, L.rp, L.ci, L.val, L.Dinv, L.b, L.x, L.t);
}

__global__ void __launch_bounds__(SB) k_bamg_coarse_solve(const BLev* __restrict__ jobs, const int2* __restrict__ map) {
    const BLev& L = jobs[job_block(map).x];
    coarse_solve(3 * L.n, L.inv, L.b, L.t);
}

__global__ void __launch_bounds__(AB) k_bamg_post(const BLev* __restrict__ jobs, const int2* __restrict__ map) {
    const int2 m = job_block(map);
    const BLev& L = jobs[m.x];
    const int i = m.y * AB + threadIdx.x;
    if (i < L.n) post_row(i, L.rp, L.ci, L.val, L.Dinv, L.agg, L.e, L.b, L.x, L.t);
}

}  // namespace

// The job maps of the cloths' hierarchies (once per context: the meshes never change).
const std::vector<int>& bamg_build(cs_ctx* C) {
    auto& B = C->bamg;
    if (B.ncl == C->nCloths) return B.usable;
    const int ncl = C->nCloths;
    B.usable.assign(ncl, 0);
    B.nl = 0;
    int sizes[32];
    for (int c = 0; c < ncl; c++) {
        const int nl = amg_levels(C, c, sizes, 32);
        B.usable[c] = nl > 1;
        if (nl > 1) B.nl = std::max(B.nl, nl);
    }
    std::vector<int2> map;
    auto seg = [&](auto&& emit) {
        cs_ctx::BatchAmg::Seg g;
        g.off = (int)map.size();
        emit();
        g.n = (int)map.size() - g.off;
        return g;
    };
    B.jobOff.assign(std::max(B.nl, 1) + 1, 0);
    B.rows.assign(B.nl, {});
    B.sparse.assign(B.nl, {});
    B.blocks.assign(B.nl, {});
    int job = 0;
    std::vector<int> coarsestJob(ncl, -1);
    for (int l = 1; l < B.nl; l++) {
        B.jobOff[l] = job;
        std::vector<int> cl;  // cloths with a level l, in order: job jobOff[l] + k is cl[k]
        for (int c = 0; c < ncl; c++)
            if (B.usable[c] && C->sys[c].amg.nl > l) cl.push_back(c);
        B.rows[l] = seg([&] {
            for (int k = 0; k < (int)cl.size(); k++)
                for (int b = 0; b < cdiv(C->sys[cl[k]].amg.L[l].n, AB); b++) map.push_back(make_int2(job + k, b));
        });
        B.sparse[l] = seg([&] {
            for (int k = 0; k < (int)cl.size(); k++)
                if (C->sys[cl[k]].amg.nl - 1 > l)
                    for (int b = 0; b < cdiv(C->sys[cl[k]].amg.L[l].n, AB); b++) map.push_back(make_int2(job + k, b));
        });
        B.blocks[l] = seg([&] {
            for (int k = 0; k < (int)cl.size(); k++)
                for (int b = 0; b < cdiv(C->sys[cl[k]].amg.L[l].nnzb, AB); b++) map.push_back(make_int2(job + k, b));
        });
        for (int k = 0; k < (int)cl.size(); k++)
            if (C->sys[cl[k]].amg.nl - 1 == l) coarsestJob[cl[k]] = job + k;
        job += (int)cl.size();
    }
    B.jobOff[std::max(B.nl, 1)] = job;
    if (B.nl >= 1) B.jobOff[0] = 0;
    B.jobCloth.assign(job, -1);
    for (int l = 1; l < B.nl; l++) {
        int k = B.jobOff[l];
        for (int c = 0; c < ncl; c++)
            if (B.usable[c] && C->sys[c].amg.nl > l) B.jobCloth[k++] = c;
    }
    int u = 0;
    B.ap = seg([&] {
        for (int c = 0; c < ncl; c++)
            if (B.usable[c]) {
                for (int b = 0; b < cdiv(C->sys[c].nn, AB); b++) map.push_back(make_int2(u, b));
                u++;
            }
    });
    B.coarse = seg([&] {
        for (int c = 0; c < ncl; c++)
            if (B.usable[c]) map.push_back(make_int2(coarsestJob[c], 0));
    });
    int2* dm = B.maps.as<int2>(std::max<size_t>(1, map.size()));
    if (!map.empty())
        CS_HIP(hipMemcpyAsync(dm, map.data(), sizeof(int2) * map.size(), hipMemcpyHostToDevice, C->stream));
    CS_HIP(hipStreamSynchronize(C->stream));
    B.hmap.swap(map);
    B.selected = false;
    B.ncl = ncl;
    return B.usable;
}

// Restricts the cycle's launches to the cloths with live[c] != 0 (null: every cloth again).
// The host copy of the selection must outlive its upload: call after a stream synchronisation.
void bamg_select(cs_ctx* C, const std::vector<char>* live) {
    auto& B = C->bamg;
    if (B.nl < 2) return;
    if (!live) {
        B.selected = false;
        return;
    }
    B.hsel.clear();
    auto pick = [&](const cs_ctx::BatchAmg::Seg& g) {
        cs_ctx::BatchAmg::Seg o;
        o.off = (int)B.hsel.size();
        for (int e = g.off; e < g.off + g.n; e++)
            if ((*live)[B.jobCloth[B.hmap[e].x]]) B.hsel.push_back(B.hmap[e]);
        o.n = (int)B.hsel.size() - o.off;
        return o;
    };
    B.crows.assign(B.nl, {});
    B.csparse.assign(B.nl, {});
    for (int l = 1; l < B.nl; l++) {
        B.crows[l] = pick(B.rows[l]);
        B.csparse[l] = pick(B.sparse[l]);
    }
    B.ccoarse = pick(B.coarse);
    int2* dm = B.cmaps.as<int2>(std::max<size_t>(1, B.hsel.size()));
    if (!B.hsel.empty())
        CS_HIP(hipMemcpyAsync(dm, B.hsel.data(), sizeof(int2) * B.hsel.size(), hipMemcpyHostToDevice, C->stream));
    B.selected = true;
}

// Per solve: every usable cloth's Galerkin operators, their diagonal inverses, level 0's
// A P and the coarsest inverses, as in amg_prepare, one launch per level for all cloths.
void bamg_prepare(cs_ctx* C) {
    auto& B = C->bamg;
    bamg_build(C);
    if (B.nl < 2) return;
    hipStream_t s = C->stream;
    const int ncl = C->nCloths;
    // job tables, rebuilt per solve (the contact blocks move between steps)
    std::vector<BLev> jobs(B.jobOff[B.nl]);
    std::vector<BAp> ap;
    for (int l = 1; l < B.nl; l++) {
        int k = B.jobOff[l];
        for (int c = 0; c < ncl; c++) {
            ClothSys& S = C->sys[c];
            AmgSys& G = S.amg;
            if (!B.usable[c] || G.nl <= l) continue;
            AmgLevel& L = G.L[l];
            AmgLevel& F = G.L[l - 1];
            BLev& J = jobs[k++];
            std::memset(&J, 0, sizeof(J));
            J.n = L.n;
            J.nnzb = L.nnzb;
            J.coarsest = l == G.nl - 1;
            J.rp = L.rp.ptr<int>();
            J.ci = L.ci.ptr<int>();
            J.diag = L.diag.ptr<int>();
            J.cptr = L.cptr.ptr<int>();
            J.cidx = L.cidx.ptr<int>();
            J.fine = l == 1 ? S.vals.ptr<float>() : F.val.ptr<float>();
            J.val = L.val.ptr<float>();
            J.Dinv = L.Dinv.ptr<float>();
            J.fmptr = F.mptr.ptr<int>();
            J.fmidx = F.midx.ptr<int>();
            J.ft = F.t.ptr<float>();
            J.b = L.b.ptr<float>();
            J.x = L.x.ptr<float>();
            J.t = L.t.ptr<float>();
            J.agg = J.coarsest ? nullptr : L.agg.ptr<int>();
            J.e = J.coarsest ? nullptr : G.L[l + 1].t.ptr<float>();
            J.cDiag = S.cDiag.ptr<int>();
            J.cVal = S.nContact > 0 ? S.cVal.ptr<float>() : nullptr;
            J.work = G.work.ptr<double>();
            J.inv = G.inv.ptr<float>();
        }
    }
    for (int c = 0; c < ncl; c++) {
        if (!B.usable[c]) continue;
        ClothSys& S = C->sys[c];
        AmgSys& G = S.amg;
        BAp J;
        std::memset(&J, 0, sizeof(J));
        J.nn = S.nn;
        J.sliceOff = S.sliceOff.ptr<int>();
        J.sliceW = S.sliceW.ptr<int>();
        J.apSliceOff = G.apSliceOff.ptr<int>();
        J.apSliceW = G.apSliceW.ptr<int>();
        J.vals = S.vals.ptr<float>();
        J.target = G.apTarget.ptr<uint8_t>();
        J.apVals = G.apVals.ptr<float>();
        ap.push_back(J);
    }
    BLev* dj = B.jobs.as<BLev>(jobs.size());
    BAp* da = B.apJobs.as<BAp>(ap.size());
    CS_HIP(hipMemcpyAsync(dj, jobs.data(), sizeof(BLev) * jobs.size(), hipMemcpyHostToDevice, s));
    CS_HIP(hipMemcpyAsync(da, ap.data(), sizeof(BAp) * ap.size(), hipMemcpyHostToDevice, s));
    const int2* map = B.maps.ptr<int2>();
    for (int l = 1; l < B.nl; l++) {
        if (B.blocks[l].n) {
            if (l == 1)
                hipLaunchKernelGGL(k_bamg_galerkin<true>, dim3(B.blocks[l].n), dim3(AB), 0, s, dj, map + B.blocks[l].off);
            else
                hipLaunchKernelGGL(k_bamg_galerkin<false>, dim3(B.blocks[l].n), dim3(AB), 0, s, dj, map + B.blocks[l].off);
        }
        if (l == 1 && B.rows[l].n)
            hipLaunchKernelGGL(k_bamg_contact_diag, dim3(B.rows[l].n), dim3(AB), 0, s, dj, map + B.rows[l].off);
        if (B.rows[l].n) hipLaunchKernelGGL(k_bamg_dinv, dim3(B.rows[l].n), dim3(AB), 0, s, dj, map + B.rows[l].off);
        CS_LAUNCH_CHECK();
    }
    if (B.ap.n) hipLaunchKernelGGL(k_bamg_ap_build, dim3(B.ap.n), dim3(AB), 0, s, da, map + B.ap.off);
    if (B.coarse.n) hipLaunchKernelGGL(k_bamg_coarse_inverse, dim3(B.coarse.n), dim3(CB), 0, s, dj, map + B.coarse.off);
    CS_LAUNCH_CHECK();
    // the host tables are released on return
    CS_HIP(hipStreamSynchronize(s));
}

// Levels >= 1 of one cycle for every usable cloth: from level 0's residual (its t, written
// by the batched PCG) to level 1's result (its t), which level 0's post-smoothing prolongs.
void bamg_cycle(cs_ctx* C) {
    auto& B = C->bamg;
    hipStream_t s = C->stream;
    const BLev* dj = B.jobs.ptr<BLev>();
    const int2* map = B.selected ? B.cmaps.ptr<int2>() : B.maps.ptr<int2>();
    const auto& rows = B.selected ? B.crows : B.rows;
    const auto& sparse = B.selected ? B.csparse : B.sparse;
    const auto& coarse = B.selected ? B.ccoarse : B.coarse;
    for (int l = 1; l < B.nl; l++) {
        if (rows[l].n) hipLaunchKernelGGL(k_bamg_restrict, dim3(rows[l].n), dim3(AB), 0, s, dj, map + rows[l].off);
        if (sparse[l].n) hipLaunchKernelGGL(k_bamg_presid, dim3(sparse[l].n), dim3(AB), 0, s, dj, map + sparse[l].off);
    }
    if (coarse.n) hipLaunchKernelGGL(k_bamg_coarse_solve, dim3(coarse.n), dim3(SB), 0, s, dj, map + coarse.off);
    for (int l = B.nl - 2; l >= 1; l--)
        if (sparse[l].n) hipLaunchKernelGGL(k_bamg_post, dim3(sparse[l].n), dim3(AB), 0, s, dj, map + sparse[l].off);
    CS_LAUNCH_CHECK();
}

}  // namespace cs
