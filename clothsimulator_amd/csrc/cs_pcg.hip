// cs_pcg.hip — 3x3 block-Jacobi preconditioned CG on one cloth's system.
//
// Same iteration as the reference GPU loop (src/Cloth.cu:766-809): x = 0,
// r = b, z = M^-1 r, d = z; per iteration q = A d, alpha = (r.z)/(d.q),
// x += alpha d, r -= alpha q, z = M^-1 r, beta = (r.z)'/(r.z), d = z + beta d.
// Differences: M is the inverse 3x3 diagonal block (not the scalar diagonal)
// and the loop stops when ||r|| <= tol ||b|| (not ||r|| < 1e-6 absolute).
//
// Three kernels per iteration; every scalar stays on the device.  Dot
// products are reduced per 256-thread block and finished by the last block to
// arrive (fixed summation order, so the result is deterministic).  A kernel
// whose `done` flag is set returns immediately, so the host can queue a chunk
// of iterations and poll the flag once per chunk.
#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

constexpr int PB = 256;

__device__ __forceinline__ float block_sum(float v, float* lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    float s = 0.0f;
    if (threadIdx.x == 0) s = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    __syncthreads();
    return s;
}

// Last-block-done reduction of K partial sums per block.  Returns true in the
// block that finishes; that block's thread 0 gets the totals in `tot`.
template <int K>
__device__ __forceinline__ bool finish_sums(const float (&part)[K], float* partial, unsigned* counter, float* tot,
                                            float* lds) {
    __shared__ int amLast;
    float s[K];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = block_sum(part[k], lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) partial[k * gridDim.x + blockIdx.x] = s[k];
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        amLast = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!amLast) return false;
    __threadfence();
#pragma unroll
    for (int k = 0; k < K; k++) {
        float a = 0.0f;
        for (int b = threadIdx.x; b < (int)gridDim.x; b += PB)
            a += __hip_atomic_load(&partial[k * gridDim.x + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        tot[k] = block_sum(a, lds);
    }
    if (threadIdx.x == 0) *counter = 0u;
    return true;
}

__device__ __forceinline__ V3f mat3v(const float* __restrict__ m, V3f a) {
    return mk(m[0] * a.x + m[1] * a.y + m[2] * a.z, m[3] * a.x + m[4] * a.y + m[5] * a.z,
              m[6] * a.x + m[7] * a.y + m[8] * a.z);
}

__device__ __forceinline__ V3f ld3v(const float* p, int i) { return mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
__device__ __forceinline__ void st3v(float* p, int i, V3f a) { p[3 * i] = a.x; p[3 * i + 1] = a.y; p[3 * i + 2] = a.z; }

__global__ void __launch_bounds__(PB) k_pcg_init(int nn, const float* __restrict__ b, const float* __restrict__ bc,
                                                 const float* __restrict__ Minv, float* __restrict__ x,
                                                 float* __restrict__ r, float* __restrict__ z, float* __restrict__ d,
                                                 float* partial, unsigned* counter, PcgScalars* sc, float tol2,
                                                 int maxit) {
    __shared__ float lds[4];
    int i = blockIdx.x * PB + threadIdx.x;
    float prz = 0.0f, pbb = 0.0f;
    if (i < nn) {
        V3f bi = ld3v(b, i) + ld3v(bc, i);
        V3f zi = mat3v(Minv + 9 * (size_t)i, bi);
        st3v(x, i, mk(0.0f, 0.0f, 0.0f));
        st3v(r, i, bi);
        st3v(z, i, zi);
        st3v(d, i, zi);
        prz = dot(bi, zi);
        pbb = dot(bi, bi);
    }
    float part[2] = {prz, pbb}, tot[2];
    if (finish_sums<2>(part, partial, counter, tot, lds) && threadIdx.x == 0) {
        sc->rz = tot[0];
        sc->bb = tot[1];
        sc->rr = tot[1];
        sc->tol2bb = tol2 * tot[1];
        sc->iters = 0;
        sc->maxit = maxit;
        sc->done = (tot[1] <= tol2 * tot[1]) || maxit <= 0;
    }
}

// q = A d over the sliced-ELL static part plus the contact CSR; partial d.q.
__global__ void __launch_bounds__(PB) k_pcg_spmv(int nn, const int* __restrict__ sliceOff, const int* __restrict__ sliceW,
                                                 const int* __restrict__ cols, const float* __restrict__ vals,
                                                 const int* __restrict__ cRowPtr, const int* __restrict__ cCol,
                                                 const float* __restrict__ cVal, const float* __restrict__ d,
                                                 float* __restrict__ q, float* partial, unsigned* counter,
                                                 PcgScalars* sc) {
    __shared__ float lds[4];
    if (sc->done) return;
    const int i = blockIdx.x * PB + threadIdx.x;
    const int s = i / SLICE, lane = i % SLICE;
    float pdq = 0.0f;
    if (i < nn) {
        const int W = sliceW[s];
        const size_t off = (size_t)sliceOff[s];
        const int* cp = cols + off * SLICE + lane;
        const float* vp = vals + off * 9 * SLICE + lane;
        float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
        for (int k = 0; k < W; k++) {
            const int j = cp[k * SLICE];
            const float d0 = d[3 * j], d1 = d[3 * j + 1], d2 = d[3 * j + 2];
            const float* a = vp + (size_t)k * 9 * SLICE;
            y0 += a[0 * SLICE] * d0 + a[1 * SLICE] * d1 + a[2 * SLICE] * d2;
            y1 += a[3 * SLICE] * d0 + a[4 * SLICE] * d1 + a[5 * SLICE] * d2;
            y2 += a[6 * SLICE] * d0 + a[7 * SLICE] * d1 + a[8 * SLICE] * d2;
        }
        const int ce = cRowPtr[i + 1];
        for (int k = cRowPtr[i]; k < ce; k++) {
            const int j = cCol[k];
            const float d0 = d[3 * j], d1 = d[3 * j + 1], d2 = d[3 * j + 2];
            const float* a = cVal + (size_t)k * 9;
            y0 += a[0] * d0 + a[1] * d1 + a[2] * d2;
            y1 += a[3] * d0 + a[4] * d1 + a[5] * d2;
            y2 += a[6] * d0 + a[7] * d1 + a[8] * d2;
        }
        q[3 * i] = y0;
        q[3 * i + 1] = y1;
        q[3 * i + 2] = y2;
        pdq = d[3 * i] * y0 + d[3 * i + 1] * y1 + d[3 * i + 2] * y2;
    }
    float part[1] = {pdq}, tot[1];
    if (finish_sums<1>(part, partial, counter, tot, lds) && threadIdx.x == 0) {
        sc->dq = tot[0];
        sc->alpha = sc->rz / tot[0];
    }
}

__global__ void __launch_bounds__(PB) k_pcg_update(int nn, const float* __restrict__ Minv, float* __restrict__ x,
                                                   float* __restrict__ r, float* __restrict__ z,
                                                   const float* __restrict__ d, const float* __restrict__ q,
                                                   float* partial, unsigned* counter, PcgScalars* sc) {
    __shared__ float lds[4];
    if (sc->done) return;
    const float alpha = sc->alpha;
    const int i = blockIdx.x * PB + threadIdx.x;
    float prz = 0.0f, prr = 0.0f;
    if (i < nn) {
        V3f di = ld3v(d, i), qi = ld3v(q, i);
        V3f xi = ld3v(x, i) + alpha * di;
        V3f ri = ld3v(r, i) - alpha * qi;
        V3f zi = mat3v(Minv + 9 * (size_t)i, ri);
        st3v(x, i, xi);
        st3v(r, i, ri);
        st3v(z, i, zi);
        prz = dot(ri, zi);
        prr = dot(ri, ri);
    }
    float part[2] = {prz, prr}, tot[2];
    if (finish_sums<2>(part, partial, counter, tot, lds) && threadIdx.x == 0) {
        sc->beta = tot[0] / sc->rz;
        sc->rz = tot[0];
        sc->rr = tot[1];
        sc->iters += 1;
        sc->done = (tot[1] <= sc->tol2bb) || (sc->iters >= sc->maxit);
    }
}

__global__ void __launch_bounds__(PB) k_pcg_dir(int nn, const float* __restrict__ z, float* __restrict__ d,
                                                const PcgScalars* sc) {
    if (sc->done) return;
    const float beta = sc->beta;
    const int i = blockIdx.x * PB + threadIdx.x;
    if (i >= 3 * nn) return;
    d[i] = z[i] + beta * d[i];
}

}  // namespace

int pcg_solve(cs_ctx* C, int c, float* relres, int* converged) {
    ClothSys& S = C->sys[c];
    hipStream_t s = C->stream;
    const int nn = S.nn;
    const int nb = cdiv(nn, PB);
    float* partial = C->pcgPartial.as<float>(2 * (size_t)nb + 8);
    unsigned* counter = C->pcgCounter.ptr<unsigned>();
    PcgScalars* sc = C->pcgScal.ptr<PcgScalars>();
    const float tol = C->P.pcg_rel_tol;
    const int maxit = C->P.pcg_max_iters;
    const int* cRowPtr = S.cRowPtr.ptr<int>();
    const int* cCol = S.nContact > 0 ? S.cCol.ptr<int>() : cRowPtr;
    const float* cVal = S.nContact > 0 ? S.cVal.ptr<float>() : S.vals.ptr<float>();
    hipLaunchKernelGGL(k_pcg_init, dim3(nb), dim3(PB), 0, s, nn, S.b.ptr<float>(), S.bc.ptr<float>(),
                       S.Minv.ptr<float>(), S.x.ptr<float>(), S.r.ptr<float>(), S.z.ptr<float>(), S.d.ptr<float>(),
                       partial, counter, sc, tol * tol, maxit);
    CS_LAUNCH_CHECK();
    const int chunk = 16;
    int done = 0;
    for (int it = 0; it < maxit + chunk && !done; it += chunk) {
        for (int k = 0; k < chunk; k++) {
            // the first SpMV of a chunk is sampled: the host saw `done == 0` after the previous chunk,
            // so this launch always does the full product
            const int tok = k == 0 ? ktimer_begin(C, CS_KERNEL_PCG_SPMV) : -1;
            hipLaunchKernelGGL(k_pcg_spmv, dim3(nb), dim3(PB), 0, s, nn, S.sliceOff.ptr<int>(), S.sliceW.ptr<int>(),
                               S.cols.ptr<int>(), S.vals.ptr<float>(), cRowPtr, cCol, cVal, S.d.ptr<float>(),
                               S.q.ptr<float>(), partial, counter, sc);
            ktimer_end(C, tok);
            hipLaunchKernelGGL(k_pcg_update, dim3(nb), dim3(PB), 0, s, nn, S.Minv.ptr<float>(), S.x.ptr<float>(),
                               S.r.ptr<float>(), S.z.ptr<float>(), S.d.ptr<float>(), S.q.ptr<float>(), partial,
                               counter, sc);
            hipLaunchKernelGGL(k_pcg_dir, dim3(cdiv(3L * nn, PB)), dim3(PB), 0, s, nn, S.z.ptr<float>(),
                               S.d.ptr<float>(), sc);
        }
        CS_LAUNCH_CHECK();
        CS_HIP(hipMemcpyAsync(C->hScal, sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        done = C->hScal->done;
        ktimer_collect(C);
    }
    const PcgScalars& h = *C->hScal;
    if (relres) *relres = h.bb > 0.0f ? sqrtf(h.rr / h.bb) : 0.0f;
    if (converged) *converged = (h.rr <= h.tol2bb);
    return h.iters;
}

}  // namespace cs
