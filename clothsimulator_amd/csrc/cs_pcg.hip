// cs_pcg.hip — 3x3 block-Jacobi preconditioned CG on one cloth's system.
//
// Same iteration as the reference GPU loop (src/Cloth.cu:766-809): x = 0,
// r = b, z = M^-1 r, d = z; per iteration q = A d, alpha = (r.z)/(d.q),
// x += alpha d, r -= alpha q, z = M^-1 r, beta = (r.z)'/(r.z), d = z + beta d.
// Differences: M is the inverse 3x3 diagonal block (not the scalar diagonal)
// and the loop stops when ||r|| <= tol ||b|| (not ||r|| < 1e-6 absolute).
//
// Two kernels per iteration, every scalar on the device and no host round
// trip inside a chunk of iterations:
//   SpMV    s = A z and t = A d_prev in one sweep over the matrix, then beta
//           from the r.z partials, q = s + beta t (= A d exactly: no
//           recurrence), d = z + beta d_prev for the block's own rows, partial
//           d.q; it also takes the stop decision from the r.r partials;
//   update  alpha, x += alpha d, r -= alpha q, z = M^-1 r, partials r.z, r.r.
// Each kernel writes one partial sum per block; the next kernel's blocks each
// reduce those partials themselves in a fixed order, so no grid-wide fence or
// atomic is needed and the result is deterministic.  The partials are loaded
// at kernel entry and reduced after the kernel's memory sweep (SpMV) or after
// its row loads (update), so their latency hides behind the sweep.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>

#include <hip/hip_cooperative_groups.h>

#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

constexpr int PB = 256;      // threads per block
constexpr int PPT = 8;       // partials held per thread by the fast reduction path
constexpr int PMAX = 16384;  // at most this many blocks (grid-stride over rows beyond), a multiple of 8
constexpr int PCG_RESTARTS = 4;  // continuations from the true residual before a solve gives up

// partial-sum slots (each PMAX floats): dq, rz[2] (ping-pong by iteration parity), rr, bb
enum { P_DQ = 0, P_RZ0 = 1, P_RZ1 = 2, P_RR = 3, P_BB = 4, P_COUNT = 5 };

struct PcgDev {
    int iters;  // iterations completed when the loop stopped
    int done;
    int maxit;
    int last;   // iteration whose stop test produced rrcur
    float tol2bb, bb, rr, rz;
    float rrcur;  // r.r seen by the latest stop test
    int restartAt;  // batched solve: the iteration that restarts the direction from z (-1: none)
    int skip;       // batched solve: batch iterations this cloth sat out (its own count is it - skip)
    float stop2bb;  // r.r that stops the recursion: tol^2 b.b, a quarter of it after two missed checks
};

// Workgroups are dealt to the 8 XCDs round-robin.  Logical block L gives each
// XCD one contiguous range of rows, so the column gathers of neighbouring rows
// and the next kernel's reads of the same rows hit that XCD's L2.
// gridDim.x is a multiple of 8.
__device__ __forceinline__ int logical_block() {
    const int per = gridDim.x >> 3;
    return (blockIdx.x & 7) * per + (blockIdx.x >> 3);
}

// Sum of NS values over the block (every thread gets the sums).
template <int NS>
__device__ __forceinline__ void block_sums(float (&v)[NS], float* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[s] += __shfl_xor(v[s], o, 64);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int s = 0; s < NS; s++) lds[4 * s + wave] = v[s];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; s++) v[s] = (lds[4 * s] + lds[4 * s + 1]) + (lds[4 * s + 2] + lds[4 * s + 3]);
}

__device__ __forceinline__ float block_sum(float v, float* lds) {
    float a[1] = {v};
    block_sums<1>(a, lds);
    return a[0];
}

// Partials of NS slots: issue() loads this thread's share (nb <= PB * PPT),
// sums() reduces them over the block in a fixed order.  Larger grids take the
// looped path.
template <int NS>
struct Partials {
    float raw[NS][PPT];
    // unconditional loads (clamped index, masked in sums) so that nothing waits for them here
    __device__ __forceinline__ void issue(const float* __restrict__ part, const int (&slot)[NS], int nb) {
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int u = 0; u < PPT; u++) raw[s][u] = part[slot[s] * PMAX + min((int)threadIdx.x + u * PB, nb - 1)];
    }
    __device__ __forceinline__ void sums(const float* __restrict__ part, const int (&slot)[NS], int nb, float* lds,
                                         float (&out)[NS]) {
        // keeps the compiler from summing (and so waiting for) the loads at issue()
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int u = 0; u < PPT; u++) asm volatile("" : "+v"(raw[s][u]));
#pragma unroll
        for (int s = 0; s < NS; s++) {
            float a = 0.0f;
#pragma unroll
            for (int u = 0; u < PPT; u++) a += (int)threadIdx.x + u * PB < nb ? raw[s][u] : 0.0f;
            for (int b = threadIdx.x + PPT * PB; b < nb; b += PB) a += part[slot[s] * PMAX + b];
            out[s] = a;
        }
        block_sums<NS>(out, lds);
    }
};

template <class P> __device__ __forceinline__ V3f mat3v(P m, V3f a) {
    return mk(m[0] * a.x + m[1] * a.y + m[2] * a.z, m[3] * a.x + m[4] * a.y + m[5] * a.z,
              m[6] * a.x + m[7] * a.y + m[8] * a.z);
}
template <class P> __device__ __forceinline__ V3f ld3v(P p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
template <class P> __device__ __forceinline__ void st3v(P p, int i, V3f a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }
__device__ __forceinline__ V3f v3(float4 a) { return mk(a.x, a.y, a.z); }
__device__ __forceinline__ V3f v3(f4n a) { return mk(a.x, a.y, a.z); }
__device__ __forceinline__ f4n f4g(V3f a) { return f4n{a.x, a.y, a.z, 0.0f}; }
__device__ __forceinline__ float4 f4(V3f a) { return make_float4(a.x, a.y, a.z, 0.0f); }

__device__ __forceinline__ void acc3(float& y0, float& y1, float& y2, const float (&a)[9], V3f v) {
    y0 += a[0] * v.x + a[1] * v.y + a[2] * v.z;
    y1 += a[3] * v.x + a[4] * v.y + a[5] * v.z;
    y2 += a[6] * v.x + a[7] * v.y + a[8] * v.z;
}

// How z = M^-1 r is formed after the update of r:
//   PC_JACOBI    z = Minv r and the r.z partials, in the same kernel;
//   PC_SMOOTH    the multigrid cycle that follows computes z and r.z; this kernel
//                does its first step, the damped Jacobi pre-smoothing xs = w Minv r;
//   PC_EXTERNAL  a separate preconditioner pass computes z and r.z.
enum { PC_JACOBI = 0, PC_SMOOTH = 1, PC_EXTERNAL = 2 };

// x = 0, r = b, and the first preconditioner step as in k_pcg_update.
template <int MODE>
__global__ void __launch_bounds__(PB) k_pcg_init(int nn, const float* __restrict__ b, const float* __restrict__ bc,
                                                 const float* __restrict__ Minv, float* __restrict__ x,
                                                 float* __restrict__ r, float4* __restrict__ z,
                                                 float* __restrict__ xs, float* __restrict__ part) {
    __shared__ float lds[8];
    const int L = logical_block();
    float acc[2] = {0.0f, 0.0f};
    for (int i = L * PB + threadIdx.x; i < nn; i += gridDim.x * PB) {
        V3f bi = ld3v(b, i) + ld3v(bc, i);
        st3v(x, i, mk(0.0f, 0.0f, 0.0f));
        st3v(r, i, bi);
        if (MODE != PC_EXTERNAL) {
            const V3f zi = mat3v(Minv + 9 * (size_t)i, bi);
            if (MODE == PC_JACOBI) {
                z[i] = f4(zi);
                acc[0] += dot(bi, zi);
            } else {
                st3v(xs, i, AMG_OMEGA * zi);
            }
        }
        acc[1] += dot(bi, bi);
    }
    block_sums<2>(acc, lds);
    if (threadIdx.x == 0) {
        if (MODE == PC_JACOBI) part[P_RZ0 * PMAX + L] = acc[0];
        part[P_BB * PMAX + L] = acc[1];
    }
}

__global__ void __launch_bounds__(PB) k_pcg_setup(int nb, const float* __restrict__ part, PcgDev* sc, float tol2,
                                                  int maxit) {
    __shared__ float lds[8];
    float acc[2] = {0.0f, 0.0f};
    for (int b = threadIdx.x; b < nb; b += PB) {
        acc[0] += part[P_BB * PMAX + b];
        acc[1] += part[P_RZ0 * PMAX + b];
    }
    block_sums<2>(acc, lds);
    if (threadIdx.x == 0) {
        const float bb = acc[0];
        sc->iters = 0;
        sc->maxit = maxit;
        sc->bb = bb;
        sc->rr = bb;
        sc->rz = acc[1];
        sc->tol2bb = tol2 * bb;
        sc->stop2bb = tol2 * bb;
        sc->done = (bb <= tol2 * bb) || maxit <= 0;
    }
}

// Iteration `it` over the sliced-ELL static part plus the contact CSR.  For
// it == 0, d = z (stored to dcur) and q = A d.  For it > 0 one sweep accumulates s = A z and
// t = A d_prev; then the stop test on r.r of the current residual, beta, and
// q = s + beta t, d = z + beta d_prev (stored to dcur for the block's rows).
// The grid covers every row in one pass (no second round for a few blocks).
// STREAM: the matrix is read with non-temporal loads.  Between two SpMVs of an iteration with the
// learned-smoother multigrid cycle, its four factor sweeps and the level-0 residual move more data
// than the 256 MB last-level cache keeps for the matrix, and streaming the SpMV's matrix reads past
// it leaves the gathered vectors more room (5.85 against 5.41 TB/s, 39.5 against 39.8 ms per step
// on the 513^2 crumple).  With block Jacobi, the learned factor alone or the plain multigrid the
// cached matrix is reused by the next SpMV and streaming measured slower (90.7 / 70.9 / 33.5 ms
// against 91.3 / 66.6 / 32.6), so those read it normally.
template <bool FIRST, bool STREAM>
__global__ void __launch_bounds__(PB) k_pcg_spmv(int it, int nn, const int* __restrict__ sliceOff,
                                                 const int* __restrict__ sliceW, const int* __restrict__ cols,
                                                 const float* __restrict__ vals, const int* __restrict__ cRowPtr,
                                                 const int* __restrict__ cCol, const float* __restrict__ cVal,
                                                 const float4* __restrict__ z, const float4* __restrict__ dprev,
                                                 float4* __restrict__ dcur, float* __restrict__ q,
                                                 float* __restrict__ part, PcgDev* sc) {
    __shared__ float lds[16];
    // Only the first iteration reads `done` up front (an all-zero right-hand side).  Later
    // iterations start their sweep at once and take the stop decision from the r.r partials,
    // which no kernel changes after convergence: a launch past the stop sweeps and returns
    // without writing, which costs less than every launch waiting for a scalar load.
    if (FIRST && sc->done) return;
    const int L = logical_block();
    const int nb = gridDim.x;
    const int p = it & 1;
    const int slot[3] = {P_RR, P_RZ0 + p, P_RZ0 + (1 - p)};
    Partials<3> pr;
    if (!FIRST) pr.issue(part, slot, nb);
    // (s0,s1,s2) = A z; for it > 0 also (t0,t1,t2) = A d_prev.
    const float4* __restrict__ g0 = z;
    const int i0 = L * PB + threadIdx.x;
    const int stride = gridDim.x * PB;
    float pdq = 0.0f;
    float beta = 0.0f;
    bool first_row = true;
    for (int i = i0; first_row || i < nn; i += stride) {
        const bool live = i < nn;
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
        if (live) {
            // One row per lane, slots in order; the compiler's own scheduling of this
            // plain loop keeps the column, value and gather loads of several slots in
            // flight (measured faster on gfx950 than hand-grouped variants).
            const int s = i / SLICE, lane = i % SLICE;
            const int W = sliceW[s];
            const size_t off = (size_t)sliceOff[s];
            const int* cp = cols + off * SLICE + lane;
            const float* vp = vals + off * 9 * SLICE + lane;
            for (int k = 0; k < W; k++) {
                const int j = STREAM ? __builtin_nontemporal_load(cp + k * SLICE) : cp[k * SLICE];
                const float* ap = vp + (size_t)k * 9 * SLICE;
                float a[9];
#pragma unroll
                for (int c = 0; c < 9; c++) a[c] = STREAM ? __builtin_nontemporal_load(ap + c * SLICE) : ap[c * SLICE];
                acc3(s0, s1, s2, a, v3(g0[j]));
                if (!FIRST) acc3(t0, t1, t2, a, v3(dprev[j]));
            }
        }
        if (live) {
            const int ce = cRowPtr[i + 1];
            for (int kc = cRowPtr[i]; kc < ce; kc++) {
                const int jj = cCol[kc];
                float a[9];
#pragma unroll
                for (int c = 0; c < 9; c++) a[c] = cVal[(size_t)kc * 9 + c];
                acc3(s0, s1, s2, a, v3(g0[jj]));
                if (!FIRST) acc3(t0, t1, t2, a, v3(dprev[jj]));
            }
        }
        if (!FIRST && first_row) {
            float tot[3];
            pr.sums(part, slot, nb, lds, tot);
            if (L == 0 && threadIdx.x == 0) {  // progress for the host's choice of chunk length
                sc->rrcur = tot[0];
                sc->last = it;
            }
            // stop on the tolerance, the budget, or a non-finite r.r (NaN fails every comparison)
            if (tot[0] <= sc->stop2bb || it >= sc->maxit || !(tot[0] <= FLT_MAX)) {  // the same in every block
                if (L == 0 && threadIdx.x == 0 && !sc->done) {
                    sc->iters = it;
                    sc->rr = tot[0];
                    sc->done = 1;
                }
                return;
            }
            beta = tot[1] / tot[2];
        }
        first_row = false;
        if (!live) continue;
        V3f di, qi;
        if (FIRST) {
            di = v3(z[i]);
            qi = mk(s0, s1, s2);
            dcur[i] = f4(di);
        } else {
            const float4 zi = z[i], dp = dprev[i];
            di = mk(zi.x + beta * dp.x, zi.y + beta * dp.y, zi.z + beta * dp.z);
            qi = mk(s0 + beta * t0, s1 + beta * t1, s2 + beta * t2);
            dcur[i] = f4(di);
        }
        st3v(q, i, qi);
        pdq += dot(di, qi);
    }
    pdq = block_sum(pdq, lds);
    if (threadIdx.x == 0) part[P_DQ * PMAX + L] = pdq;
}

// alpha, x += alpha d, r -= alpha q, partial r.r, and the preconditioner step of MODE.
template <int MODE>
__global__ void __launch_bounds__(PB) k_pcg_update(int it, int nn, const float* __restrict__ Minv,
                                                   float* __restrict__ x, float* __restrict__ r,
                                                   float4* __restrict__ z, float* __restrict__ xs,
                                                   const float4* __restrict__ d, const float* __restrict__ q,
                                                   float* __restrict__ part, const PcgDev* sc) {
    __shared__ float lds[8];
    const int L = logical_block();
    const int nb = gridDim.x;
    const int p = it & 1;
    const int slot[2] = {P_DQ, P_RZ0 + p};
    Partials<2> pr;
    pr.issue(part, slot, nb);
    float acc[2] = {0.0f, 0.0f};
    float alpha = 0.0f;
    bool first_row = true;
    for (int i = L * PB + threadIdx.x; first_row || i < nn; i += gridDim.x * PB) {
        const bool live = i < nn;
        V3f di, qi, xi, ri;
        float m[9];
        if (live) {  // row loads before the scalar reduction
            di = v3(d[i]);
            qi = ld3v(q, i);
            xi = ld3v(x, i);
            ri = ld3v(r, i);
            if (MODE != PC_EXTERNAL)
#pragma unroll
                for (int c = 0; c < 9; c++) m[c] = Minv[9 * (size_t)i + c];
        }
        if (first_row) {
            if (sc->done) return;  // read after the row loads are in flight; the same in every thread
            float tot[2];
            pr.sums(part, slot, nb, lds, tot);
            alpha = tot[1] / tot[0];
            first_row = false;
        }
        if (!live) continue;
        xi = xi + alpha * di;
        ri = ri - alpha * qi;
        st3v(x, i, xi);
        st3v(r, i, ri);
        if (MODE != PC_EXTERNAL) {
            const V3f zi = mk(m[0] * ri.x + m[1] * ri.y + m[2] * ri.z, m[3] * ri.x + m[4] * ri.y + m[5] * ri.z,
                              m[6] * ri.x + m[7] * ri.y + m[8] * ri.z);
            if (MODE == PC_JACOBI) {
                z[i] = f4(zi);
                acc[0] += dot(ri, zi);
            } else {
                st3v(xs, i, AMG_OMEGA * zi);
            }
        }
        acc[1] += dot(ri, ri);
    }
    block_sums<2>(acc, lds);
    if (threadIdx.x == 0) {
        if (MODE == PC_JACOBI) part[(P_RZ0 + (1 - p)) * PMAX + L] = acc[0];
        part[P_RR * PMAX + L] = acc[1];
    }
}

// The true residual of the current iterate, r = b + bc - A x, with every row sum in FP64 (the
// recursive r of the loop drifts from it in FP32), stored as the new r; partial r.r, and the
// preconditioner step of MODE on it (partials in the r.z slot of iteration `it`).
template <int MODE>
__global__ void __launch_bounds__(PB) k_pcg_true_residual(int it, int nn, const int* __restrict__ sliceOff,
                                                          const int* __restrict__ sliceW, const int* __restrict__ cols,
                                                          const float* __restrict__ vals,
                                                          const int* __restrict__ cRowPtr, const int* __restrict__ cCol,
                                                          const float* __restrict__ cVal, const float* __restrict__ b,
                                                          const float* __restrict__ bc, const float* __restrict__ Minv,
                                                          const float* __restrict__ x, float* __restrict__ r,
                                                          float4* __restrict__ z, float* __restrict__ xs,
                                                          float* __restrict__ part) {
    __shared__ float lds[8];
    const int L = logical_block();
    float acc[2] = {0.0f, 0.0f};
    for (int i = L * PB + threadIdx.x; i < nn; i += gridDim.x * PB) {
        double y[3] = {0.0, 0.0, 0.0};
        auto add = [&](const float (&a)[9], int j) {
            const double xj[3] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
#pragma unroll
            for (int u = 0; u < 3; u++) y[u] += (double)a[3 * u] * xj[0] + (double)a[3 * u + 1] * xj[1] +
                                                 (double)a[3 * u + 2] * xj[2];
        };
        const int s = i / SLICE, lane = i % SLICE;
        const int W = sliceW[s];
        const size_t off = (size_t)sliceOff[s];
        for (int k = 0; k < W; k++) {
            const float* ap = vals + (off + k) * 9 * SLICE + lane;
            float a[9];
#pragma unroll
            for (int c = 0; c < 9; c++) a[c] = ap[c * SLICE];
            add(a, cols[(off + k) * SLICE + lane]);
        }
        for (int kc = cRowPtr[i]; kc < cRowPtr[i + 1]; kc++) {
            float a[9];
#pragma unroll
            for (int c = 0; c < 9; c++) a[c] = cVal[(size_t)kc * 9 + c];
            add(a, cCol[kc]);
        }
        V3f ri;
        ri.x = (float)((double)b[3 * i] + (double)bc[3 * i] - y[0]);
        ri.y = (float)((double)b[3 * i + 1] + (double)bc[3 * i + 1] - y[1]);
        ri.z = (float)((double)b[3 * i + 2] + (double)bc[3 * i + 2] - y[2]);
        st3v(r, i, ri);
        if (MODE != PC_EXTERNAL) {
            const V3f zi = mat3v(Minv + 9 * (size_t)i, ri);
            if (MODE == PC_JACOBI) {
                z[i] = f4(zi);
                acc[0] += dot(ri, zi);
            } else {
                st3v(xs, i, AMG_OMEGA * zi);
            }
        }
        acc[1] += dot(ri, ri);
    }
    block_sums<2>(acc, lds);
    if (threadIdx.x == 0) {
        if (MODE == PC_JACOBI) part[(P_RZ0 + (it & 1)) * PMAX + L] = acc[0];
        part[P_RR * PMAX + L] = acc[1];
    }
}

// The recursive stop of a continuation after the true residual missed the tolerance: the first
// continuation stops at tol again (a few iterations from the true residual usually land under
// it); a later one aims at half the tolerance, for the systems whose true residual keeps
// trailing the recursive one by more than the last few iterations gain.
__device__ __forceinline__ float continuation_stop(float tol2bb, int restarts) {
    return restarts >= 1 ? 0.25f * tol2bb : tol2bb;
}

// The stop decision on the true residual: accept (done) when it meets tol^2 b.b or the
// iteration budget is spent, otherwise reopen the loop from iteration `it`.
__global__ void __launch_bounds__(PB) k_pcg_recheck(int nb, int it, int restarts, const float* __restrict__ part,
                                                   PcgDev* sc) {
    __shared__ float lds[8];
    float acc[1] = {0.0f};
    for (int k = threadIdx.x; k < nb; k += PB) acc[0] += part[P_RR * PMAX + k];
    block_sums<1>(acc, lds);
    if (threadIdx.x == 0) {
        sc->rr = acc[0];
        sc->rrcur = acc[0];
        sc->last = it;
        sc->iters = it;
        sc->done = (acc[0] <= sc->tol2bb || it >= sc->maxit || !(acc[0] <= FLT_MAX)) ? 1 : 0;
        sc->stop2bb = continuation_stop(sc->tol2bb, restarts);
    }
}

// ---------------------------------------------------------------- batched solve
// Independent cloths solved together: every launch covers the rows of all
// cloths (256 rows per block, a block never spans two cloths), and each cloth
// keeps its own partial sums, alpha, beta and stop flag, so the iterates are
// those of separate solves while a whole batch costs two launches per iteration.

struct BCloth {
    int nn, nb, amg, pad1;  // rows, blocks, preconditioned by its multigrid cycle
    const int *sliceOff, *sliceW, *cols, *cRowPtr, *cCol;
    const float *vals, *cVal, *Minv, *b, *bc;
    float *x, *r, *q;
    float4 *z, *d0, *d1;
    float* part;  // P_COUNT slots of nb partials: slot s of block k at part[s * nb + k]
    PcgDev* sc;
    // multigrid (amg): level 0's pre-smoothed iterate and residual, its aggregates, level 1's
    // result, and A_static P in sliced ELL
    float *xs, *t0;
    const int* agg0;
    const float* e1;
    const int *apSliceOff, *apSliceW, *apCols;
    const float* apVals;
};

// A BCloth's pointers in the global address space (loads and stores through them
// compile to global instructions, not flat ones).
struct BClothG {
    int nn, nb, amg;
    const G1<int> *sliceOff, *sliceW, *cols, *cRowPtr, *cCol;
    const G1<float> *vals, *cVal, *Minv, *b, *bc;
    G1<float> *x, *r, *q;
    G1<f4n> *z, *d0, *d1;
    G1<float>* part;
    G1<PcgDev>* sc;
    G1<float> *xs, *t0;
    const G1<int>* agg0;
    const G1<float>* e1;
    const G1<int> *apSliceOff, *apSliceW, *apCols;
    const G1<float>* apVals;
};
__device__ __forceinline__ BClothG gview(const BCloth& c) {
    BClothG g;
    g.nn = c.nn;
    g.nb = c.nb;
    g.sliceOff = g1(c.sliceOff); g.sliceW = g1(c.sliceW); g.cols = g1(c.cols);
    g.cRowPtr = g1(c.cRowPtr); g.cCol = g1(c.cCol);
    g.vals = g1(c.vals); g.cVal = g1(c.cVal); g.Minv = g1(c.Minv); g.b = g1(c.b); g.bc = g1(c.bc);
    g.x = g1(c.x); g.r = g1(c.r); g.q = g1(c.q);
    g.z = (G1<f4n>*)c.z; g.d0 = (G1<f4n>*)c.d0; g.d1 = (G1<f4n>*)c.d1;
    g.part = g1(c.part);
    g.sc = g1(c.sc);
    g.amg = c.amg;
    g.xs = g1(c.xs); g.t0 = g1(c.t0); g.agg0 = g1(c.agg0); g.e1 = g1(c.e1);
    g.apSliceOff = g1(c.apSliceOff); g.apSliceW = g1(c.apSliceW); g.apCols = g1(c.apCols); g.apVals = g1(c.apVals);
    return g;
}

// Sum of NS values over a BP-thread block (every thread gets the sums).
template <int NS, int BP>
__device__ __forceinline__ void bsums(float (&v)[NS], float* lds) {
    constexpr int NW = BP / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[s] += __shfl_xor(v[s], o, 64);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int s = 0; s < NS; s++) lds[NW * s + wave] = v[s];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; s++) {  // pairwise tree over the waves (block_sums' order at 4 waves)
        float a[NW];
#pragma unroll
        for (int w = 0; w < NW; w++) a[w] = lds[NW * s + w];
#pragma unroll
        for (int h = NW / 2; h > 0; h >>= 1)
#pragma unroll
            for (int w = 0; w < h; w++) a[w] = a[2 * w] + a[2 * w + 1];
        v[s] = a[0];
    }
}

// Partials of NS slots of one cloth (nb <= 4 BP blocks in the fast path):
// issue() loads this thread's share unconditionally, sums() reduces them.
template <int NS, int BP>
struct BPartials {
    float raw[NS][4];
    template <class P> __device__ __forceinline__ void issue(P part, const int (&slot)[NS], int nb) {
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int u = 0; u < 4; u++) raw[s][u] = part[slot[s] * nb + min((int)threadIdx.x + u * BP, nb - 1)];
    }
    template <class P> __device__ __forceinline__ void sums(P part, const int (&slot)[NS], int nb, float* lds,
                                                         float (&out)[NS]) {
#pragma unroll
        for (int s = 0; s < NS; s++)
#pragma unroll
            for (int u = 0; u < 4; u++) asm volatile("" : "+v"(raw[s][u]));
#pragma unroll
        for (int s = 0; s < NS; s++) {
            float a = 0.0f;
#pragma unroll
            for (int u = 0; u < 4; u++) a += (int)threadIdx.x + u * BP < nb ? raw[s][u] : 0.0f;
            for (int b = threadIdx.x + 4 * BP; b < nb; b += BP) a += part[slot[s] * nb + b];
            out[s] = a;
        }
        bsums<NS, BP>(out, lds);
    }
};

// The batch grid covers the unfinished cloths only (rebuilt at every host
// poll), padded to a multiple of 8 blocks and walked in logical order
// (logical_block): each XCD gets one contiguous run of the cloths' blocks, so
// neighbouring rows' gathers share that XCD's L2.  bmap[L] = (cloth, block of
// the cloth) of logical block L (cloth -1: padding).

// (cloth, block) of this workgroup, as wave-uniform values (scalar loads of the cloth's table)
__device__ __forceinline__ int2 batch_block(const int2* __restrict__ bmap) {
    const int2 m = bmap[logical_block()];
    return make_int2(__builtin_amdgcn_readfirstlane(m.x), __builtin_amdgcn_readfirstlane(m.y));
}

template <int BP>
__global__ void __launch_bounds__(BP) k_bpcg_init(const BCloth* __restrict__ cl, const int2* __restrict__ bmap) {
    __shared__ float lds[2 * BP / 64];
    const int2 m = batch_block(bmap);
    if (m.x < 0) return;
    const BClothG c = gview(cl[m.x]);
    const int lb = m.y;
    const int i = lb * BP + threadIdx.x;
    float acc[2] = {0.0f, 0.0f};
    if (i < c.nn) {
        const V3f bi = ld3v(c.b, i) + ld3v(c.bc, i);
        st3v(c.x, i, mk(0.0f, 0.0f, 0.0f));
        st3v(c.r, i, bi);
        const V3f zi = mat3v(c.Minv + 9 * (size_t)i, bi);
        if (c.amg) {
            st3v(c.xs, i, AMG_OMEGA * zi);  // the cycle's pre-smoothing; it writes z and r.z
        } else {
            c.z[i] = f4g(zi);
            acc[0] = dot(bi, zi);
        }
        acc[1] = dot(bi, bi);
    }
    bsums<2, BP>(acc, lds);
    if (threadIdx.x == 0) {
        if (!c.amg) c.part[P_RZ0 * c.nb + lb] = acc[0];
        c.part[P_BB * c.nb + lb] = acc[1];
    }
}

// one block per cloth
template <int BP>
__global__ void __launch_bounds__(BP) k_bpcg_setup(const BCloth* __restrict__ cl, float tol2, int maxit) {
    __shared__ float lds[2 * BP / 64];
    const BClothG c = gview(cl[blockIdx.x]);
    const int slot[2] = {P_BB, P_RZ0};
    BPartials<2, BP> pr;
    pr.issue(c.part, slot, c.nb);
    float v[2];
    pr.sums(c.part, slot, c.nb, lds, v);
    if (threadIdx.x == 0) {
        c.sc->iters = 0;
        c.sc->maxit = maxit;
        c.sc->bb = v[0];
        c.sc->rr = v[0];
        c.sc->rz = v[1];
        c.sc->tol2bb = tol2 * v[0];
        c.sc->stop2bb = tol2 * v[0];
        c.sc->done = (v[0] <= tol2 * v[0]) || maxit <= 0;
        c.sc->restartAt = -1;
        c.sc->skip = 0;
    }
}

// A matrix load, non-temporal under STREAM (a batch whose matrix cannot stay in the
// last-level cache between iterations leaves that cache to the gathered vectors).
template <bool STREAM, class T> __device__ __forceinline__ T ldm(const G1<T>* p) {
    return STREAM ? __builtin_nontemporal_load(p) : *p;
}

// Iteration `it` of every unfinished cloth, as k_pcg_spmv: the sweep s = A z,
// t = A d_prev first, then the stop test from the r.r partials (loaded at entry),
// beta, q = s + beta t, d = z + beta d_prev and the partial d.q.  A stopped
// cloth's partials no longer change, so its blocks keep deciding to stop.
// The body of one (cloth, block) m; lds holds 3 * BP / 64 floats.  Its early returns are uniform
// over the workgroup (per-cloth values and block sums).
template <int BP, bool FIRST, bool STREAM>
__device__ __forceinline__ void bpcg_spmv_block(const BCloth* __restrict__ cl, int2 m, int it, float* lds) {
    const BClothG c = gview(cl[m.x]);
    const int lb = m.y;
    const int p = it & 1;
    const int slot[3] = {P_RR, P_RZ0 + p, P_RZ0 + (1 - p)};
    BPartials<3, BP> pr;
    if (!FIRST) pr.issue(c.part, slot, c.nb);
    if (FIRST && c.sc->done) return;
    const G1<f4n>* dprev = p ? c.d0 : c.d1;
    const int i = lb * BP + threadIdx.x;
    const bool live = i < c.nn;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
    if (live) {
        const int sl = i / SLICE, lane = i % SLICE;
        const int W = c.sliceW[sl];
        const size_t off = (size_t)c.sliceOff[sl];
        const G1<int>* cp = c.cols + off * SLICE + lane;
        const G1<float>* vp = c.vals + off * 9 * SLICE + lane;
        // Two slots per step: both columns, then both slots' values and gathers, so a wave
        // has two slots' loads in flight behind one column latency (the batch's matrix
        // streams from HBM).  The sums still run slot by slot.
        int k = 0;
        for (; k + 1 < W; k += 2) {
            const int j0 = ldm<STREAM>(cp + k * SLICE), j1 = ldm<STREAM>(cp + (k + 1) * SLICE);
            const G1<float>* ap = vp + (size_t)k * 9 * SLICE;
            float a0[9], a1[9];
#pragma unroll
            for (int q = 0; q < 9; q++) {
                a0[q] = ldm<STREAM>(ap + q * SLICE);
                a1[q] = ldm<STREAM>(ap + (9 + q) * SLICE);
            }
            const f4n z0 = c.z[j0], z1 = c.z[j1];
            f4n e0, e1;
            if (!FIRST) {
                e0 = dprev[j0];
                e1 = dprev[j1];
            }
            acc3(s0, s1, s2, a0, v3(z0));
            if (!FIRST) acc3(t0, t1, t2, a0, v3(e0));
            acc3(s0, s1, s2, a1, v3(z1));
            if (!FIRST) acc3(t0, t1, t2, a1, v3(e1));
        }
        if (k < W) {
            const int j = ldm<STREAM>(cp + k * SLICE);
            const G1<float>* ap = vp + (size_t)k * 9 * SLICE;
            float a[9];
#pragma unroll
            for (int q = 0; q < 9; q++) a[q] = ldm<STREAM>(ap + q * SLICE);
            acc3(s0, s1, s2, a, v3(c.z[j]));
            if (!FIRST) acc3(t0, t1, t2, a, v3(dprev[j]));
        }
        for (int kc = c.cRowPtr[i], ce = c.cRowPtr[i + 1]; kc < ce; kc++) {
            const int jj = c.cCol[kc];
            float a[9];
#pragma unroll
            for (int q = 0; q < 9; q++) a[q] = c.cVal[(size_t)kc * 9 + q];
            acc3(s0, s1, s2, a, v3(c.z[jj]));
            if (!FIRST) acc3(t0, t1, t2, a, v3(dprev[jj]));
        }
    }
    float beta = 0.0f;
    // a cloth continued from its true residual starts its direction from z again at `restartAt`
    const bool restart = !FIRST && c.sc->restartAt == it;
    if (!FIRST && !restart) {
        float v[3];
        pr.sums(c.part, slot, c.nb, lds, v);
        const int own = it - c.sc->skip;  // this cloth's own iteration count
        if (v[0] <= c.sc->stop2bb || own >= c.sc->maxit || !(v[0] <= FLT_MAX)) {  // the same in every block
            if (lb == 0 && threadIdx.x == 0 && !c.sc->done) {
                c.sc->iters = own;
                c.sc->rr = v[0];
                c.sc->done = 1;
            }
            return;
        }
        beta = v[1] / v[2];
    }
    float pdq = 0.0f;
    if (live) {
        G1<f4n>* dcur = FIRST ? c.d0 : (p ? c.d1 : c.d0);
        const f4n zi = c.z[i];
        V3f di, qi;
        if (FIRST || restart) {
            di = v3(zi);
            qi = mk(s0, s1, s2);
        } else {
            const f4n dp = dprev[i];
            di = mk(zi.x + beta * dp.x, zi.y + beta * dp.y, zi.z + beta * dp.z);
            qi = mk(s0 + beta * t0, s1 + beta * t1, s2 + beta * t2);
        }
        dcur[i] = f4g(di);
        st3v(c.q, i, qi);
        pdq = dot(di, qi);
    }
    {
        float t1[1] = {pdq};
        bsums<1, BP>(t1, lds);
        pdq = t1[0];
    }
    if (threadIdx.x == 0) c.part[P_DQ * c.nb + lb] = pdq;
}

template <int BP, bool FIRST, bool STREAM>
__global__ void __launch_bounds__(BP) k_bpcg_spmv(const BCloth* __restrict__ cl, const int2* __restrict__ bmap,
                                                  int it) {
    __shared__ float lds[3 * BP / 64];
    const int2 m = batch_block(bmap);
    if (m.x < 0) return;
    bpcg_spmv_block<BP, FIRST, STREAM>(cl, m, it, lds);
}

// alpha, x += alpha d, r -= alpha q, z = Minv r, partials r.z and r.r (the row
// loads are issued before the partials are reduced)
template <int BP>
__device__ __forceinline__ void bpcg_update_block(const BCloth* __restrict__ cl, int2 m, int it, float* lds) {
    const BClothG c = gview(cl[m.x]);
    const int lb = m.y;
    const int p = it & 1;
    const int slot[2] = {P_DQ, P_RZ0 + p};
    BPartials<2, BP> pr;
    pr.issue(c.part, slot, c.nb);
    const G1<f4n>* d = it == 0 ? c.d0 : (p ? c.d1 : c.d0);
    const int i = lb * BP + threadIdx.x;
    const bool live = i < c.nn;
    V3f di = mk(0.0f, 0.0f, 0.0f), xi = di, ri = di, qi = di;
    if (live) {
        di = v3(d[i]);
        xi = ld3v(c.x, i);
        ri = ld3v(c.r, i);
        qi = ld3v(c.q, i);
    }
    if (c.sc->done) return;
    float v[2];
    pr.sums(c.part, slot, c.nb, lds, v);
    const float alpha = v[1] / v[0];
    float acc[2] = {0.0f, 0.0f};
    if (live) {
        xi = xi + alpha * di;
        ri = ri - alpha * qi;
        st3v(c.x, i, xi);
        st3v(c.r, i, ri);
        const V3f zi = mat3v(c.Minv + 9 * (size_t)i, ri);
        if (c.amg) {
            st3v(c.xs, i, AMG_OMEGA * zi);
        } else {
            c.z[i] = f4g(zi);
            acc[0] = dot(ri, zi);
        }
        acc[1] = dot(ri, ri);
    }
    bsums<2, BP>(acc, lds);
    if (threadIdx.x == 0) {
        if (!c.amg) c.part[(P_RZ0 + (1 - p)) * c.nb + lb] = acc[0];
        c.part[P_RR * c.nb + lb] = acc[1];
    }
}

template <int BP>
__global__ void __launch_bounds__(BP) k_bpcg_update(const BCloth* __restrict__ cl, const int2* __restrict__ bmap,
                                                    int it) {
    __shared__ float lds[2 * BP / 64];
    const int2 m = batch_block(bmap);
    if (m.x < 0) return;
    bpcg_update_block<BP>(cl, m, it, lds);
}

// Iterations [it0, it0 + n) of the block-Jacobi solve of every cloth of the map in one
// cooperative launch (a grid no larger than the workgroups the device holds at once): each
// workgroup walks its share of the map, the two phases of an iteration separated by grid
// barriers.  Workgroups are dealt to the XCDs round-robin; XCD x's workgroups take the x-th
// eighth of the map (nmap and gridDim.x are multiples of 8), the rows logical_block gives
// that XCD in a full grid.  Every workgroup runs all n iterations and meets every barrier; the
// per-block bodies are the launched kernels', so the results are theirs bit for bit.
template <int BP>
__global__ void __launch_bounds__(BP) k_bpcg_persist(const BCloth* __restrict__ cl, const int2* __restrict__ bmap,
                                                     int nmap, int it0, int n) {
    __shared__ float lds[3 * BP / 64];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int per = gridDim.x >> 3, xcd = blockIdx.x & 7, span = nmap >> 3;
    const int2* __restrict__ mine = bmap + xcd * span;
    for (int k = 0; k < n; k++) {
        const int it = it0 + k;
        for (int L = blockIdx.x >> 3; L < span; L += per) {
            const int2 mm = mine[L];
            const int2 m = make_int2(__builtin_amdgcn_readfirstlane(mm.x), __builtin_amdgcn_readfirstlane(mm.y));
            if (m.x < 0) continue;
            if (it == 0) bpcg_spmv_block<BP, true, false>(cl, m, it, lds);
            else bpcg_spmv_block<BP, false, false>(cl, m, it, lds);
        }
        __threadfence();
        grid.sync();
        for (int L = blockIdx.x >> 3; L < span; L += per) {
            const int2 mm = mine[L];
            const int2 m = make_int2(__builtin_amdgcn_readfirstlane(mm.x), __builtin_amdgcn_readfirstlane(mm.y));
            if (m.x < 0) continue;
            bpcg_update_block<BP>(cl, m, it, lds);
        }
        __threadfence();
        grid.sync();
    }
}

// The true residual of every listed cloth (the map covers them), r = b + bc - A x with FP64 row
// sums, z = Minv r; partials r.z (slot of iteration `it`) and r.r.
template <int BP>
__global__ void __launch_bounds__(BP) k_bpcg_true(const BCloth* __restrict__ cl, const int2* __restrict__ bmap,
                                                  int it) {
    __shared__ float lds[2 * BP / 64];
    const int2 m = batch_block(bmap);
    if (m.x < 0) return;
    const BClothG c = gview(cl[m.x]);
    const int lb = m.y;
    const int i = lb * BP + threadIdx.x;
    float acc[2] = {0.0f, 0.0f};
    if (i < c.nn) {
        double y[3] = {(double)c.b[3 * i] + (double)c.bc[3 * i], (double)c.b[3 * i + 1] + (double)c.bc[3 * i + 1],
                       (double)c.b[3 * i + 2] + (double)c.bc[3 * i + 2]};
        auto sub = [&](const float (&a)[9], int j) {
            const double xj[3] = {c.x[3 * j], c.x[3 * j + 1], c.x[3 * j + 2]};
#pragma unroll
            for (int u = 0; u < 3; u++)
                y[u] -= (double)a[3 * u] * xj[0] + (double)a[3 * u + 1] * xj[1] + (double)a[3 * u + 2] * xj[2];
        };
        const int sl = i / SLICE, lane = i % SLICE;
        const int W = c.sliceW[sl];
        const size_t off = (size_t)c.sliceOff[sl];
        for (int k = 0; k < W; k++) {
            float a[9];
#pragma unroll
            for (int q = 0; q < 9; q++) a[q] = c.vals[(off + k) * 9 * SLICE + q * SLICE + lane];
            sub(a, c.cols[(off + k) * SLICE + lane]);
        }
        for (int kc = c.cRowPtr[i], ce = c.cRowPtr[i + 1]; kc < ce; kc++) {
            float a[9];
#pragma unroll
            for (int q = 0; q < 9; q++) a[q] = c.cVal[(size_t)kc * 9 + q];
            sub(a, c.cCol[kc]);
        }
        const V3f ri = mk((float)y[0], (float)y[1], (float)y[2]);
        st3v(c.r, i, ri);
        const V3f zi = mat3v(c.Minv + 9 * (size_t)i, ri);
        if (c.amg) {
            st3v(c.xs, i, AMG_OMEGA * zi);
        } else {
            c.z[i] = f4g(zi);
            acc[0] = dot(ri, zi);
        }
        acc[1] = dot(ri, ri);
    }
    bsums<2, BP>(acc, lds);
    if (threadIdx.x == 0) {
        if (!c.amg) c.part[(P_RZ0 + (it & 1)) * c.nb + lb] = acc[0];
        c.part[P_RR * c.nb + lb] = acc[1];
    }
}

// Level 0 of the multigrid cycle of every listed cloth that uses one: t = r - A xs (sliced
// ELL, then the contact blocks, in slot order).  liveOnly: cloths already stopped are skipped.
template <int BP, bool STREAM>
__global__ void __launch_bounds__(BP) k_bpcg_resid0(const BCloth* __restrict__ cl, const int2* __restrict__ bmap,
                                                    int liveOnly) {
    const int2 m = batch_block(bmap);
    if (m.x < 0) return;
    const BClothG c = gview(cl[m.x]);
    if (!c.amg || (liveOnly && c.sc->done)) return;
    const int i = m.y * BP + threadIdx.x;
    if (i >= c.nn) return;
    const int sl = i / SLICE, lane = i % SLICE;
    const int W = c.sliceW[sl];
    const size_t off = (size_t)c.sliceOff[sl];
    float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
    for (int k = 0; k < W; k++) {
        const int j = ldm<STREAM>(c.cols + (off + k) * SLICE + lane);
        float a[9];
#pragma unroll
        for (int q = 0; q < 9; q++) a[q] = ldm<STREAM>(c.vals + (off + k) * 9 * SLICE + q * SLICE + lane);
        acc3(y0, y1, y2, a, ld3v(c.xs, j));
    }
    for (int kc = c.cRowPtr[i], ce = c.cRowPtr[i + 1]; kc < ce; kc++) {
        float a[9];
#pragma unroll
        for (int q = 0; q < 9; q++) a[q] = c.cVal[(size_t)kc * 9 + q];
        acc3(y0, y1, y2, a, ld3v(c.xs, c.cCol[kc]));
    }
    st3v(c.t0, i, ld3v(c.r, i) - mk(y0, y1, y2));
}

// Level 0 post-smoothing from the stored residual t = r - A xs: with e = level 1's result,
// r - A (xs + P e) = t - (A P) e (A P = the static A P plus the contact blocks gathered through
// the aggregates), z = xs + P e + w Minv (t - (A P) e); the partial of r.z into `slot`.
template <int BP, bool STREAM>
__global__ void __launch_bounds__(BP) k_bpcg_post0(const BCloth* __restrict__ cl, const int2* __restrict__ bmap,
                                                   int liveOnly, int slot) {
    __shared__ float lds[BP / 64];
    const int2 m = batch_block(bmap);
    if (m.x < 0) return;
    const BClothG c = gview(cl[m.x]);
    if (!c.amg || (liveOnly && c.sc->done)) return;
    const int lb = m.y;
    const int i = lb * BP + threadIdx.x;
    float prz = 0.0f;
    if (i < c.nn) {
        const int sl = i / SLICE, lane = i % SLICE;
        const int WA = c.apSliceW[sl];
        const size_t aoff = (size_t)c.apSliceOff[sl];
        float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
        for (int k = 0; k < WA; k++) {
            float a[9];
#pragma unroll
            for (int q = 0; q < 9; q++) a[q] = ldm<STREAM>(c.apVals + (aoff + k) * 9 * SLICE + q * SLICE + lane);
            acc3(y0, y1, y2, a, ld3v(c.e1, ldm<STREAM>(c.apCols + (aoff + k) * SLICE + lane)));
        }
        for (int kc = c.cRowPtr[i], ce = c.cRowPtr[i + 1]; kc < ce; kc++) {
            float a[9];
#pragma unroll
            for (int q = 0; q < 9; q++) a[q] = c.cVal[(size_t)kc * 9 + q];
            acc3(y0, y1, y2, a, ld3v(c.e1, c.agg0[c.cCol[kc]]));
        }
        const V3f xi = ld3v(c.xs, i) + ld3v(c.e1, c.agg0[i]);
        const V3f zi = xi + AMG_OMEGA * mat3v(c.Minv + 9 * (size_t)i, ld3v(c.t0, i) - mk(y0, y1, y2));
        c.z[i] = f4g(zi);
        prz = dot(ld3v(c.r, i), zi);
    }
    float v[1] = {prz};
    bsums<1, BP>(v, lds);
    if (threadIdx.x == 0) c.part[slot * c.nb + lb] = v[0];
}

// One block per listed cloth: accept it when the true residual meets tol^2 b.b (or its
// iterations are spent), otherwise reopen it with its direction restarting at `it`.
template <int BP>
__global__ void __launch_bounds__(BP) k_bpcg_recheck(const BCloth* __restrict__ cl, const int* __restrict__ list,
                                                     int it, int restarts) {
    __shared__ float lds[BP / 64];
    const BClothG c = gview(cl[list[blockIdx.x]]);
    const int slot[1] = {P_RR};
    BPartials<1, BP> pr;
    pr.issue(c.part, slot, c.nb);
    float v[1];
    pr.sums(c.part, slot, c.nb, lds, v);
    if (threadIdx.x == 0) {
        c.sc->rr = v[0];
        if (v[0] <= c.sc->tol2bb || c.sc->iters >= c.sc->maxit || !(v[0] <= FLT_MAX)) {
            c.sc->done = 1;
        } else {  // reopened from its true residual
            c.sc->done = 0;
            c.sc->skip = it - c.sc->iters;
            c.sc->restartAt = it;
            c.sc->stop2bb = continuation_stop(c.sc->tol2bb, restarts);
        }
    }
}

}  // namespace

int pcg_solve(cs_ctx* C, int c, float* relres, int* converged) {
    ClothSys& S = C->sys[c];
    hipStream_t s = C->stream;
    const int nn = S.nn;
    const int nb = std::min(PMAX, (cdiv(nn, PB) + 7) & ~7);
    float* part = C->pcgPartial.as<float>((size_t)P_COUNT * PMAX);
    PcgDev* sc = reinterpret_cast<PcgDev*>(C->pcgScal.as<PcgScalars>(1));
    static_assert(sizeof(PcgDev) <= sizeof(PcgScalars), "scalar block");
    const float tol = C->P.pcg_rel_tol;
    const int maxit = C->P.pcg_max_iters;
    const int* cRowPtr = S.cRowPtr.ptr<int>();
    const int* cCol = S.nContact > 0 ? S.cCol.ptr<int>() : cRowPtr;
    const float* cVal = S.nContact > 0 ? S.cVal.ptr<float>() : S.vals.ptr<float>();
    float4* z = S.z.ptr<float4>();
    float4* dgen[2] = {S.d.ptr<float4>(), S.d.ptr<float4>() + nn};
    int levels[16];
    const bool amgl = C->P.precond == CS_PRECOND_AMG_LEARNED && amg_levels(C, c, levels, 16) > 1;
    const bool amg = C->P.precond == CS_PRECOND_AMG && amg_levels(C, c, levels, 16) > 1;
    const bool learned = C->P.precond == CS_PRECOND_LEARNED;
    float* xs = nullptr;  // level-0 pre-smoothing iterate of the multigrid cycle
    if (amgl) {
        amg_prepare(C, c, true);  // its single-workgroup part overlaps the network's inference
        lpc_prepare(C, c);
        lpc_smoother_prepare(C, c);
        xs = amg_level0_x(C, c);
        hipLaunchKernelGGL(k_pcg_init<PC_EXTERNAL>, dim3(nb), dim3(PB), 0, s, nn, S.b.ptr<float>(),
                           S.bc.ptr<float>(), S.Minv.ptr<float>(), S.x.ptr<float>(), S.r.ptr<float>(), z, xs, part);
        CS_LAUNCH_CHECK();
        lpc_smooth_pre(C, c, S.r.ptr<float>(), xs);
        amg_join(C);
        amg_apply(C, c, S.r.ptr<float>(), z, part + P_RZ0 * PMAX, nb, true, true);
    } else if (amg) {
        amg_prepare(C, c);
        xs = amg_level0_x(C, c);
        hipLaunchKernelGGL(k_pcg_init<PC_SMOOTH>, dim3(nb), dim3(PB), 0, s, nn, S.b.ptr<float>(), S.bc.ptr<float>(),
                           S.Minv.ptr<float>(), S.x.ptr<float>(), S.r.ptr<float>(), z, xs, part);
        CS_LAUNCH_CHECK();
        amg_apply(C, c, S.r.ptr<float>(), z, part + P_RZ0 * PMAX, nb, true);
    } else if (learned) {
        lpc_prepare(C, c);
        hipLaunchKernelGGL(k_pcg_init<PC_EXTERNAL>, dim3(nb), dim3(PB), 0, s, nn, S.b.ptr<float>(),
                           S.bc.ptr<float>(), S.Minv.ptr<float>(), S.x.ptr<float>(), S.r.ptr<float>(), z, xs, part);
        CS_LAUNCH_CHECK();
        lpc_apply(C, c, S.r.ptr<float>(), z, part + P_RZ0 * PMAX, nb);
    } else {
        hipLaunchKernelGGL(k_pcg_init<PC_JACOBI>, dim3(nb), dim3(PB), 0, s, nn, S.b.ptr<float>(), S.bc.ptr<float>(),
                           S.Minv.ptr<float>(), S.x.ptr<float>(), S.r.ptr<float>(), z, xs, part);
        CS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_pcg_setup, dim3(1), dim3(PB), 0, s, nb, part, sc, tol * tol, maxit);
    CS_LAUNCH_CHECK();
    // iterations per host poll: a multigrid iteration costs several times a Jacobi one
    // Iterations per host poll.  The first chunk has a fixed length; later ones
    // aim at the stop: the host extrapolates the r.r decrease between its last two
    // polls to the iteration that reaches tol^2 b.b (between 1 and 4x the first
    // length).  Each
    // iteration launched past the stop costs a full multigrid or learned
    // iteration (only the SpMV returns early), each poll one host round trip.
    const int chunk0 = (amg || amgl) ? 4 : learned ? 8 : 16;
    int chunk = chunk0;
    // The multigrid cycle of an iteration (with the learned smoother: its two pre-smoothing
    // sweeps too) is the same launch sequence every iteration up to the parity of the
    // r.z partial slot; with CS_FLAG_GRAPH_CYCLES it is captured once per parity and solve
    // into a HIP graph and replayed (about 15 launches become one).  Measured on the 513^2
    // crumple it saves nothing: the two instantiations per solve cost what the replays
    // save, so direct launches are the default.
    hipGraphExec_t cycle[2] = {nullptr, nullptr};
    const bool graphs = (amg || amgl) && (C->P.flags & CS_FLAG_GRAPH_CYCLES);
    auto launch_cycle = [&](int par) {
        if (amgl) lpc_smooth_pre(C, c, S.r.ptr<float>(), xs);
        amg_apply(C, c, S.r.ptr<float>(), z, part + (P_RZ0 + (1 - par)) * PMAX, nb, true, amgl);
    };
    auto run_cycle = [&](int par) {
        if (!graphs) {
            launch_cycle(par);
            return;
        }
        if (!cycle[par]) {
            hipGraph_t g = nullptr;
            CS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            launch_cycle(par);
            CS_HIP(hipStreamEndCapture(s, &g));
            const hipError_t e = hipGraphInstantiate(&cycle[par], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            CS_HIP(e);
        }
        CS_HIP(hipGraphLaunch(cycle[par], s));
    };
    struct GraphRelease {
        hipGraphExec_t* e;
        ~GraphRelease() {
            for (int k = 0; k < 2; k++)
                if (e[k]) (void)hipGraphExecDestroy(e[k]);
        }
    } release{cycle};
    int done = 0;
    PcgDev h{};
    int it = 0;
    int first = 0;      // the iteration that starts its direction from z alone (0, or a restart)
    int restarts = 0;
    float rr0 = -1.0f;  // r.r at the previous poll, and its iteration
    int it0 = 0;
    auto poll = [&]() {
        CS_HIP(hipMemcpyAsync(C->hScal, sc, sizeof(PcgDev), hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        std::memcpy(&h, C->hScal, sizeof(PcgDev));
        ktimer_collect(C);
    };
    for (;;) {
        while (!done && it <= maxit + 8 * chunk0) {
            for (int k = 0; k < chunk; k++, it++) {
                // sampled: every 8th SpMV of the solve, each after an empty event pair that measures
                // the markers' own latency
                const bool sample = (it - first) % 8 == 1;
                if (sample) ktimer_end(C, ktimer_begin(C, CS_KERNEL_EVENT_PAIR));
                const int tok = sample ? ktimer_begin(C, CS_KERNEL_PCG_SPMV) : -1;
                float4* dcur = dgen[it & 1];
                const bool stream = amgl;
                auto spmv = it == first ? (stream ? k_pcg_spmv<true, true> : k_pcg_spmv<true, false>)
                                        : (stream ? k_pcg_spmv<false, true> : k_pcg_spmv<false, false>);
                hipLaunchKernelGGL(spmv, dim3(nb), dim3(PB), 0, s, it, nn, S.sliceOff.ptr<int>(), S.sliceW.ptr<int>(),
                                   S.cols.ptr<int>(), S.vals.ptr<float>(), cRowPtr, cCol, cVal, z,
                                   it == first ? dcur : dgen[1 - (it & 1)], dcur, S.q.ptr<float>(), part, sc);
                ktimer_end(C, tok);
                if (amgl) {
                    hipLaunchKernelGGL(k_pcg_update<PC_EXTERNAL>, dim3(nb), dim3(PB), 0, s, it, nn,
                                       S.Minv.ptr<float>(), S.x.ptr<float>(), S.r.ptr<float>(), z, xs, dcur,
                                       S.q.ptr<float>(), part, sc);
                    run_cycle(it & 1);
                } else if (amg) {
                    hipLaunchKernelGGL(k_pcg_update<PC_SMOOTH>, dim3(nb), dim3(PB), 0, s, it, nn, S.Minv.ptr<float>(),
                                       S.x.ptr<float>(), S.r.ptr<float>(), z, xs, dcur, S.q.ptr<float>(), part, sc);
                    run_cycle(it & 1);
                } else if (learned) {
                    hipLaunchKernelGGL(k_pcg_update<PC_EXTERNAL>, dim3(nb), dim3(PB), 0, s, it, nn,
                                       S.Minv.ptr<float>(), S.x.ptr<float>(), S.r.ptr<float>(), z, xs, dcur,
                                       S.q.ptr<float>(), part, sc);
                    lpc_apply(C, c, S.r.ptr<float>(), z, part + (P_RZ0 + (1 - (it & 1))) * PMAX, nb);
                } else {
                    hipLaunchKernelGGL(k_pcg_update<PC_JACOBI>, dim3(nb), dim3(PB), 0, s, it, nn, S.Minv.ptr<float>(),
                                       S.x.ptr<float>(), S.r.ptr<float>(), z, xs, dcur, S.q.ptr<float>(), part, sc);
                }
            }
            CS_LAUNCH_CHECK();
            poll();
            done = h.done;
            if (done) break;
            chunk = chunk0;
            if (rr0 > 0.0f && h.rrcur > 0.0f && rr0 > h.rrcur && h.last > it0) {
                // geometric extrapolation of the latest r.r decrease
                const double rate = std::log((double)h.rrcur / rr0) / (h.last - it0);  // < 0
                const double need = std::log((double)h.stop2bb / h.rrcur) / rate;      // iterations to go
                const int left = (int)std::ceil(need) - (it - 1 - h.last);             // beyond those launched
                chunk = std::max(1, std::min(4 * chunk0, left));
            }
            rr0 = h.rrcur;
            it0 = h.last;
        }
        if (!done) throw StatusError(CS_E_HIP, "PCG: the device never reported its stop");
        // The loop stopped on its recursive residual.  Accept only what the true residual
        // b - A x (row sums in FP64) confirms; otherwise continue from it, the search direction
        // restarted from its preconditioned residual, at most PCG_RESTARTS times.
        it = h.iters;
        const int* so = S.sliceOff.ptr<int>();
        const int* sw = S.sliceW.ptr<int>();
        const int* co = S.cols.ptr<int>();
        const float* va = S.vals.ptr<float>();
        if (amgl || learned)
            hipLaunchKernelGGL(k_pcg_true_residual<PC_EXTERNAL>, dim3(nb), dim3(PB), 0, s, it, nn, so, sw, co, va,
                               cRowPtr, cCol, cVal, S.b.ptr<float>(), S.bc.ptr<float>(), S.Minv.ptr<float>(),
                               S.x.ptr<float>(), S.r.ptr<float>(), z, xs, part);
        else if (amg)
            hipLaunchKernelGGL(k_pcg_true_residual<PC_SMOOTH>, dim3(nb), dim3(PB), 0, s, it, nn, so, sw, co, va,
                               cRowPtr, cCol, cVal, S.b.ptr<float>(), S.bc.ptr<float>(), S.Minv.ptr<float>(),
                               S.x.ptr<float>(), S.r.ptr<float>(), z, xs, part);
        else
            hipLaunchKernelGGL(k_pcg_true_residual<PC_JACOBI>, dim3(nb), dim3(PB), 0, s, it, nn, so, sw, co, va,
                               cRowPtr, cCol, cVal, S.b.ptr<float>(), S.bc.ptr<float>(), S.Minv.ptr<float>(),
                               S.x.ptr<float>(), S.r.ptr<float>(), z, xs, part);
        hipLaunchKernelGGL(k_pcg_recheck, dim3(1), dim3(PB), 0, s, nb, it, restarts, part, sc);
        CS_LAUNCH_CHECK();
        poll();
        if (h.done || restarts == PCG_RESTARTS) break;
        restarts++;
        float* slot = part + (P_RZ0 + (it & 1)) * PMAX;  // where iteration `it` reads r.z
        if (amgl) {
            lpc_smooth_pre(C, c, S.r.ptr<float>(), xs);
            amg_apply(C, c, S.r.ptr<float>(), z, slot, nb, true, true);
        } else if (amg) {
            amg_apply(C, c, S.r.ptr<float>(), z, slot, nb, true);
        } else if (learned) {
            lpc_apply(C, c, S.r.ptr<float>(), z, slot, nb);
        }
        first = it;
        done = 0;
        chunk = chunk0;
        rr0 = -1.0f;
    }
    if (relres) *relres = h.bb > 0.0f ? sqrtf(h.rr / h.bb) : 0.0f;
    if (converged) *converged = (h.rr <= h.tol2bb);
    S.solved = h.rr <= h.tol2bb;  // S.x solves the assembled system
    return h.iters;
}

// The batched solve with BP rows per block and `chunk` iterations per host poll.  amg: the
// cloths with a usable hierarchy run one multigrid cycle per iteration (level 0 in these
// kernels, the coarser levels of all cloths in bamg_cycle's launches).
// Batches whose static matrices exceed this read them with non-temporal loads in the SpMV: half
// the 256 MB last-level cache, which also holds the vectors and the other kernels' data.
constexpr size_t kStreamBytes = 128u << 20;

// Workgroups of k_bpcg_persist<BP> the device holds at once, a multiple of 8 (the grid of a
// cooperative launch may not exceed it).
template <int BP>
static int persist_grid(cs_ctx* C) {
    int coop = 0, cus = 0, perCU = 0;
    CS_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, C->device));
    if (!coop) throw StatusError(CS_E_HIP, "CS_FLAG_PERSISTENT_BATCH: the device has no cooperative launch");
    CS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, C->device));
    CS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_bpcg_persist<BP>, BP, 0));
    const int g = (perCU * cus) & ~7;
    if (g < 8) throw StatusError(CS_E_HIP, "CS_FLAG_PERSISTENT_BATCH: too few resident workgroups");
    return g;
}

template <int BP>
static void solve_batched(cs_ctx* C, int chunk, int* iters, float* relres, int* converged, bool amg) {
    hipStream_t s = C->stream;
    const int ncl = C->nCloths;
    std::vector<int> useAmg(ncl, 0);
    if (amg) {
        useAmg = bamg_build(C);
        bamg_prepare(C);
        bamg_select(C, nullptr);
    }
    std::vector<char> live(ncl, 1);
    // the coarse levels of the cycle follow the level-0 map: only the cloths still iterating
    auto select_live = [&](const PcgDev* st) {
        if (!amg) return;
        for (int c = 0; c < ncl; c++) live[c] = !st[c].done;
        bamg_select(C, &live);
    };
    std::vector<BCloth> tab(ncl);
    size_t partTotal = 0;
    for (int c = 0; c < ncl; c++) {
        tab[c].nb = std::max(1, cdiv(C->sys[c].nn, BP));
        partTotal += (size_t)P_COUNT * tab[c].nb;
    }
    float* part = C->bPart.as<float>(partTotal);
    PcgDev* sc = C->bScal.as<PcgDev>(ncl);
    size_t po = 0;
    for (int c = 0; c < ncl; c++) {
        ClothSys& S = C->sys[c];
        BCloth& b = tab[c];
        b.nn = S.nn;
        b.sliceOff = S.sliceOff.ptr<int>();
        b.sliceW = S.sliceW.ptr<int>();
        b.cols = S.cols.ptr<int>();
        b.vals = S.vals.ptr<float>();
        b.cRowPtr = S.cRowPtr.ptr<int>();
        b.cCol = S.nContact > 0 ? S.cCol.ptr<int>() : S.cRowPtr.ptr<int>();
        b.cVal = S.nContact > 0 ? S.cVal.ptr<float>() : S.vals.ptr<float>();
        b.Minv = S.Minv.ptr<float>();
        b.b = S.b.ptr<float>();
        b.bc = S.bc.ptr<float>();
        b.x = S.x.ptr<float>();
        b.r = S.r.ptr<float>();
        b.q = S.q.ptr<float>();
        b.z = S.z.ptr<float4>();
        b.d0 = S.d.ptr<float4>();
        b.d1 = S.d.ptr<float4>() + S.nn;
        b.part = part + po;
        po += (size_t)P_COUNT * b.nb;
        b.sc = sc + c;
        b.amg = useAmg[c];
        if (b.amg) {
            AmgSys& G = S.amg;
            b.xs = G.L[0].x.ptr<float>();
            b.t0 = G.L[0].t.ptr<float>();
            b.agg0 = G.L[0].agg.ptr<int>();
            b.e1 = G.L[1].t.ptr<float>();
            b.apSliceOff = G.apSliceOff.ptr<int>();
            b.apSliceW = G.apSliceW.ptr<int>();
            b.apCols = G.apCols.ptr<int>();
            b.apVals = G.apVals.ptr<float>();
        } else {
            b.xs = b.t0 = nullptr;
            b.agg0 = b.apSliceOff = b.apSliceW = b.apCols = nullptr;
            b.e1 = b.apVals = nullptr;
        }
    }
    const bool anyAmg = std::find(useAmg.begin(), useAmg.end(), 1) != useAmg.end();
    // block Jacobi only: the multigrid cycle's launches stay launches
    const bool persist = (C->P.flags & CS_FLAG_PERSISTENT_BATCH) && !anyAmg;
    const int persistMax = persist ? persist_grid<BP>(C) : 0;
    size_t matBytes = 0;  // the static matrices of all cloths: 36 B of values + 4 B of column per slot
    for (int c = 0; c < ncl; c++) matBytes += (size_t)C->sys[c].totalSlots * SLICE * 40;
    const bool streamA = matBytes > kStreamBytes;
    // one multigrid cycle for the cloths of the map (their r and xs written): z and the partial
    // r.z into `slot`
    auto cycle = [&](const int2* map, int nblk, int liveOnly, int slot) {
        if (!anyAmg) return;
        hipLaunchKernelGGL((streamA ? k_bpcg_resid0<BP, true> : k_bpcg_resid0<BP, false>), dim3(nblk), dim3(BP), 0, s,
                           (const BCloth*)C->bTab.p, map, liveOnly);
        bamg_cycle(C);
        hipLaunchKernelGGL((streamA ? k_bpcg_post0<BP, true> : k_bpcg_post0<BP, false>), dim3(nblk), dim3(BP), 0, s,
                           (const BCloth*)C->bTab.p, map, liveOnly, slot);
    };
    // (cloth, block) of every logical block of the unfinished cloths, padded to a multiple of 8
    std::vector<int2> bmap;
    auto build_map = [&](const PcgDev* done) {
        bmap.clear();
        for (int c = 0; c < ncl; c++)
            if (!done || !done[c].done)
                for (int k = 0; k < tab[c].nb; k++) bmap.push_back(make_int2(c, k));
        while (bmap.size() % 8) bmap.push_back(make_int2(-1, 0));
    };
    build_map(nullptr);
    int NB = (int)bmap.size();
    const int NBfull = NB;
    BCloth* dtab = C->bTab.as<BCloth>(ncl);
    int2* dmap = C->bMap.as<int2>(NB);
    CS_HIP(hipMemcpyAsync(dtab, tab.data(), sizeof(BCloth) * ncl, hipMemcpyHostToDevice, s));
    CS_HIP(hipMemcpyAsync(dmap, bmap.data(), sizeof(int2) * NB, hipMemcpyHostToDevice, s));
    const float tol = C->P.pcg_rel_tol;
    const int maxit = C->P.pcg_max_iters;
    hipLaunchKernelGGL(k_bpcg_init<BP>, dim3(NB), dim3(BP), 0, s, dtab, dmap);
    cycle(dmap, NB, 0, P_RZ0);
    hipLaunchKernelGGL(k_bpcg_setup<BP>, dim3(ncl), dim3(BP), 0, s, dtab, tol * tol, maxit);
    CS_LAUNCH_CHECK();
    const PcgDev* h = reinterpret_cast<const PcgDev*>(pinned_batch(C, sizeof(PcgDev) * ncl));
    int it = 0;
    bool all = false;
    std::vector<char> verified(ncl, 0);  // accepted on the true residual
    int* dlist = C->bList.as<int>(ncl);
    for (int restarts = 0;; restarts++) {
        const int passStart = it;
        while (!all && it - passStart <= maxit + 2 * chunk) {
            if (persist) {  // the poll's iterations in one cooperative launch
                int nmap = NB, it0 = it, n = chunk;
                const BCloth* ctab = dtab;
                const int2* cmap = dmap;
                void* args[] = {(void*)&ctab, (void*)&cmap, (void*)&nmap, (void*)&it0, (void*)&n};
                CS_HIP(hipLaunchCooperativeKernel((const void*)k_bpcg_persist<BP>, dim3(std::min(NB, persistMax)),
                                                  dim3(BP), args, 0, s));
                it += chunk;
            }
            for (int k = 0; k < chunk && !persist; k++, it++) {
                // sampled: every 8th batch SpMV while every cloth is still iterating (the launch
                // then covers the whole batch), each after an empty event pair that measures the
                // markers' own latency
                const bool sample = it % 8 == 1 && NB == NBfull;
                if (sample) ktimer_end(C, ktimer_begin(C, CS_KERNEL_EVENT_PAIR));
                const int tok = sample ? ktimer_begin(C, CS_KERNEL_PCG_SPMV) : -1;
                auto spmv = it == 0 ? (streamA ? k_bpcg_spmv<BP, true, true> : k_bpcg_spmv<BP, true, false>)
                                    : (streamA ? k_bpcg_spmv<BP, false, true> : k_bpcg_spmv<BP, false, false>);
                hipLaunchKernelGGL(spmv, dim3(NB), dim3(BP), 0, s, dtab, dmap, it);
                ktimer_end(C, tok);
                hipLaunchKernelGGL(k_bpcg_update<BP>, dim3(NB), dim3(BP), 0, s, dtab, dmap, it);
                cycle(dmap, NB, 1, P_RZ0 + (1 - (it & 1)));
            }
            CS_LAUNCH_CHECK();
            CS_HIP(hipMemcpyAsync(C->hBatch, sc, sizeof(PcgDev) * ncl, hipMemcpyDeviceToHost, s));
            CS_HIP(hipStreamSynchronize(s));
            ktimer_collect(C);
            all = true;
            int active = 0;
            for (int c = 0; c < ncl; c++) {
                all = all && h[c].done;
                active += h[c].done ? 0 : tab[c].nb;
            }
            if (!all && active + 8 <= NB) {  // drop the finished cloths from the grid
                build_map(h);
                NB = (int)bmap.size();
                CS_HIP(hipMemcpyAsync(dmap, bmap.data(), sizeof(int2) * NB, hipMemcpyHostToDevice, s));
                select_live(h);
            }
        }
        if (!all) break;
        // Every cloth stopped on its recursive residual: the ones not yet confirmed are checked on
        // their true residual (FP64 row sums) and reopened from it when it misses the tolerance,
        // at most PCG_RESTARTS times.
        std::vector<int> list;
        bmap.clear();
        for (int c = 0; c < ncl; c++)
            if (!verified[c]) {
                list.push_back(c);
                for (int k = 0; k < tab[c].nb; k++) bmap.push_back(make_int2(c, k));
            }
        while (bmap.size() % 8) bmap.push_back(make_int2(-1, 0));
        NB = (int)bmap.size();
        CS_HIP(hipMemcpyAsync(dmap, bmap.data(), sizeof(int2) * NB, hipMemcpyHostToDevice, s));
        CS_HIP(hipMemcpyAsync(dlist, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, s));
        if (amg) {
            for (int c = 0; c < ncl; c++) live[c] = !verified[c];
            bamg_select(C, &live);
        }
        hipLaunchKernelGGL(k_bpcg_true<BP>, dim3(NB), dim3(BP), 0, s, dtab, dmap, it);
        cycle(dmap, NB, 0, P_RZ0 + (it & 1));
        hipLaunchKernelGGL(k_bpcg_recheck<BP>, dim3((unsigned)list.size()), dim3(BP), 0, s, dtab, dlist, it,
                           restarts);
        CS_LAUNCH_CHECK();
        CS_HIP(hipMemcpyAsync(C->hBatch, sc, sizeof(PcgDev) * ncl, hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        all = true;
        for (int c : list) {
            if (h[c].done) verified[c] = 1;
            all = all && h[c].done;
        }
        if (all || restarts == PCG_RESTARTS) {
            all = true;  // a cloth still open here keeps its (true) residual: reported as not converged
            break;
        }
        build_map(h);  // the reopened cloths
        NB = (int)bmap.size();
        CS_HIP(hipMemcpyAsync(dmap, bmap.data(), sizeof(int2) * NB, hipMemcpyHostToDevice, s));
        select_live(h);
    }
    if (!all) throw StatusError(CS_E_HIP, "batched PCG: the device never reported every stop");
    for (int c = 0; c < ncl; c++) {
        iters[c] = h[c].iters;
        relres[c] = h[c].bb > 0.0f ? sqrtf(h[c].rr / h[c].bb) : 0.0f;
        converged[c] = h[c].rr <= h[c].tol2bb;
        C->sys[c].solved = converged[c] != 0;
    }
}

// Batched-solve autotuner: the first solves of a context try every (rows per
// block, iterations per poll) candidate twice, round robin, and score each by
// host wall time per node-iteration (sum over cloths of iterations x rows); the
// best is kept for the context's later solves.  The rows per block set the
// grouping of the dot-product partials, so results of different candidates agree
// to rounding; CS_FLAG_FIXED_TUNING keeps (256, 16) for bit-reproducible runs.
static const int kTuneRows[] = {256, 128, 256, 128};
static const int kTuneChunk[] = {16, 16, 32, 32};
constexpr int kTuneN = 4, kTuneRounds = 2;

// All cloths' systems (assembled, Minv current) solved together with PCG (block Jacobi, or
// each cloth's multigrid cycle).  Per cloth: iterations, relative residual and convergence,
// as pcg_solve.  A multigrid iteration costs several Jacobi ones, so it polls a quarter as often.
void pcg_solve_batched(cs_ctx* C, int* iters, float* relres, int* converged, bool amg) {
    auto& T = C->batchTune;
    int cand = T.choice;
    const bool tuning = !(C->P.flags & CS_FLAG_FIXED_TUNING) && T.trials < kTuneN * kTuneRounds;
    if (C->P.flags & CS_FLAG_FIXED_TUNING) cand = 0;
    else if (tuning) cand = T.trials % kTuneN;
    const auto t0 = std::chrono::steady_clock::now();
    const int chunk = amg ? kTuneChunk[cand] / 4 : kTuneChunk[cand];
    if (kTuneRows[cand] == 128) solve_batched<128>(C, chunk, iters, relres, converged, amg);
    else solve_batched<256>(C, chunk, iters, relres, converged, amg);
    if (tuning) {
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        double work = 0.0;
        for (int c = 0; c < C->nCloths; c++) work += (double)iters[c] * C->sys[c].nn;
        T.score[cand] += work > 0.0 ? sec / work : 0.0;
        if (++T.trials == kTuneN * kTuneRounds) {
            int best = 0;
            for (int k = 1; k < kTuneN; k++)
                if (T.score[k] < T.score[best]) best = k;
            T.choice = best;
        }
    }
    T.rows = kTuneRows[T.choice];
    T.chunk = amg ? kTuneChunk[T.choice] / 4 : kTuneChunk[T.choice];
}

}  // namespace cs
