// cs_noalloc.hip — cs_solve_bsr_noalloc: block-Jacobi PCG on a system held in the caller's
// device memory, allocating nothing.
//
// The same algorithm and stop rule as the block-Jacobi solve of cs_solve (x = 0, r = b,
// z = d = M^-1 r; per iteration q = A d, alpha = r.z / d.q, x += alpha d, r -= alpha q, stop when
// r.r <= tol^2 b.b or at max_iters, z = M^-1 r, beta = r.z_new / r.z, d = z + beta d; a stop
// confirmed by the true residual b - A x in FP64, else continued from it), on a
// block-CSR matrix (3x3 row-major blocks, every row holding its diagonal block) instead of the
// context's sliced ELL.  Every temporary (M^-1, r, z, d, q, partial sums, scalars) lives in
// `scratch`, one file-scope device array of fixed size: each GPU holds its own copy for the
// life of the process, and every call on that GPU reuses it, so calls that overlap on one GPU
// race.  The host reads back one scalar record every NA_CHUNK iterations.
#include <cfloat>
#include <mutex>

#include "cs_common.hpp"
#include "cs_math.hpp"
#include "../../include/clothsim.h"

namespace cs {

namespace {

constexpr int NA_MAX = CS_NOALLOC_MAX_NODES;
constexpr int NA_TB = 256;                    // threads per workgroup, one row each
constexpr int NA_MAXB = NA_MAX / NA_TB;       // workgroups (partial-sum slots) at most
constexpr int NA_CHUNK = 16;                  // iterations between host reads
constexpr int NA_RESTARTS = 4;                // continuations from the true residual at most

struct NaScal {
    float bb, rz, rr, alpha, beta, tol2bb;
    int done, iters, maxit, bad;
};

// float offsets into the scratch array
constexpr size_t NA_MINV = 0;
constexpr size_t NA_R = NA_MINV + 9 * (size_t)NA_MAX;
constexpr size_t NA_Z = NA_R + 3 * (size_t)NA_MAX;
constexpr size_t NA_D = NA_Z + 3 * (size_t)NA_MAX;
constexpr size_t NA_Q = NA_D + 3 * (size_t)NA_MAX;
constexpr size_t NA_PART = NA_Q + 3 * (size_t)NA_MAX;  // 2 floats per workgroup
constexpr size_t NA_SCAL = NA_PART + 2 * (size_t)NA_MAXB;
constexpr size_t NA_FLOATS = NA_SCAL + 64;
static_assert(sizeof(NaScal) <= 64 * sizeof(float), "scalar record");

// The temporaries of every call on this GPU (about 5.5 MB).  Host code never passes this name
// to a kernel: it stands for the device array only in device code; the host uses the address
// hipGetSymbolAddress returns for the current device.
__device__ __align__(256) float scratch[NA_FLOATS];

struct NaPtr {
    float *Minv, *r, *z, *d, *q, *part;
    NaScal* sc;
};

NaPtr carve(float* base) {
    NaPtr p;
    p.Minv = base + NA_MINV;
    p.r = base + NA_R;
    p.z = base + NA_Z;
    p.d = base + NA_D;
    p.q = base + NA_Q;
    p.part = base + NA_PART;
    p.sc = reinterpret_cast<NaScal*>(base + NA_SCAL);
    return p;
}

__device__ __forceinline__ void mul3(const float* M, const float* v, float* o) {
    o[0] = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
    o[1] = M[3] * v[0] + M[4] * v[1] + M[5] * v[2];
    o[2] = M[6] * v[0] + M[7] * v[1] + M[8] * v[2];
}

// Fixed-order workgroup sum of two values (wave shuffles, then the waves in order).
__device__ void wg_sum2(float a, float b, float* out) {
    __shared__ float lds[2][NA_TB / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if (lane == 0) { lds[0][w] = a; lds[1][w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float s0 = 0.0f, s1 = 0.0f;
        for (int k = 0; k < NA_TB / 64; k++) { s0 += lds[0][k]; s1 += lds[1][k]; }
        out[0] = s0;
        out[1] = s1;
    }
}

// M^-1 from the diagonal blocks, x = 0, r = b, z = d = M^-1 b; partials b.b and r.z.  A row
// without its diagonal block or with a column outside [0, n) raises `bad`: the set-up reduction
// then ends the solve before any SpMV, and the call returns CS_E_ARG.
__global__ void __launch_bounds__(NA_TB) k_na_init(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const float* __restrict__ vals, const float* __restrict__ b,
                                                   float* __restrict__ x, NaPtr P) {
    const int i = blockIdx.x * NA_TB + threadIdx.x;
    float bb = 0.0f, rz = 0.0f;
    if (i < n) {
        int diag = -1;
        bool cols_ok = true;  // a column outside [0, n) would make the SpMV read outside the scratch vectors
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            const int j = ci[k];
            cols_ok = cols_ok && j >= 0 && j < n;
            if (j == i && diag < 0) diag = k;
        }
        if (!cols_ok) P.sc->bad = 1;
        float a[9], m[9];
        if (diag >= 0) {
            for (int c = 0; c < 9; c++) a[c] = vals[(size_t)diag * 9 + c];
            inv3(a, m);
        } else {
            P.sc->bad = 1;  // the caller's matrix lacks a diagonal block
            for (int c = 0; c < 9; c++) m[c] = (c % 4 == 0) ? 1.0f : 0.0f;
        }
        for (int c = 0; c < 9; c++) P.Minv[(size_t)i * 9 + c] = m[c];
        float bi[3] = {b[3 * i], b[3 * i + 1], b[3 * i + 2]}, zi[3];
        mul3(m, bi, zi);
        for (int c = 0; c < 3; c++) {
            x[3 * i + c] = 0.0f;
            P.r[3 * i + c] = bi[c];
            P.z[3 * i + c] = zi[c];
            P.d[3 * i + c] = zi[c];
            bb += bi[c] * bi[c];
            rz += bi[c] * zi[c];
        }
    }
    wg_sum2(bb, rz, P.part + 2 * blockIdx.x);
}

// One workgroup: the partial sums in workgroup order.  mode 0: b.b and r.z of the start;
// mode 1: alpha = r.z / d.q; mode 2: r.r and the new r.z, the stop test and beta.
__global__ void __launch_bounds__(NA_TB) k_na_reduce(int nb, int mode, float tol2, int maxit, NaPtr P) {
    NaScal* sc = P.sc;
    if (mode != 0 && sc->done) return;
    float a = 0.0f, b = 0.0f;
    for (int k = threadIdx.x; k < nb; k += NA_TB) {
        a += P.part[2 * k];
        b += P.part[2 * k + 1];
    }
    __shared__ float tot[2];
    wg_sum2(a, b, tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    if (mode == 0) {
        sc->bb = tot[0];
        sc->rr = tot[0];
        sc->rz = tot[1];
        sc->tol2bb = tol2 * tot[0];
        sc->iters = 0;
        sc->maxit = maxit;
        sc->done = sc->bad || tot[0] <= tol2 * tot[0] || maxit <= 0;
    } else if (mode == 1) {
        sc->alpha = sc->rz / tot[0];
    } else {
        sc->rr = tot[0];
        sc->iters += 1;
        if (tot[0] <= sc->tol2bb || sc->iters >= sc->maxit || !(tot[0] <= FLT_MAX)) {  // NaN: breakdown
            sc->done = 1;
            return;
        }
        sc->beta = tot[1] / sc->rz;
        sc->rz = tot[1];
    }
}

// q = A d, partial d.q
__global__ void __launch_bounds__(NA_TB) k_na_spmv(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const float* __restrict__ vals, NaPtr P) {
    if (P.sc->done) return;
    const int i = blockIdx.x * NA_TB + threadIdx.x;
    float dq = 0.0f;
    if (i < n) {
        float acc[3] = {0.0f, 0.0f, 0.0f};
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            const int j = ci[k];
            const float* A = vals + (size_t)k * 9;
            const float dj[3] = {P.d[3 * j], P.d[3 * j + 1], P.d[3 * j + 2]};
            float t[3];
            mul3(A, dj, t);
            acc[0] += t[0];
            acc[1] += t[1];
            acc[2] += t[2];
        }
        for (int c = 0; c < 3; c++) {
            P.q[3 * i + c] = acc[c];
            dq += P.d[3 * i + c] * acc[c];
        }
    }
    wg_sum2(dq, 0.0f, P.part + 2 * blockIdx.x);
}

// x += alpha d, r -= alpha q, z = M^-1 r; partials r.r and r.z
__global__ void __launch_bounds__(NA_TB) k_na_update(int n, float* __restrict__ x, NaPtr P) {
    if (P.sc->done) return;
    const int i = blockIdx.x * NA_TB + threadIdx.x;
    float rr = 0.0f, rz = 0.0f;
    if (i < n) {
        const float alpha = P.sc->alpha;
        float ri[3], zi[3];
        for (int c = 0; c < 3; c++) {
            x[3 * i + c] += alpha * P.d[3 * i + c];
            ri[c] = P.r[3 * i + c] - alpha * P.q[3 * i + c];
            P.r[3 * i + c] = ri[c];
        }
        mul3(P.Minv + (size_t)i * 9, ri, zi);
        for (int c = 0; c < 3; c++) {
            P.z[3 * i + c] = zi[c];
            rr += ri[c] * ri[c];
            rz += ri[c] * zi[c];
        }
    }
    wg_sum2(rr, rz, P.part + 2 * blockIdx.x);
}

// d = z + beta d
__global__ void __launch_bounds__(NA_TB) k_na_dir(int n, NaPtr P) {
    if (P.sc->done) return;
    const int i = blockIdx.x * NA_TB + threadIdx.x;
    if (i >= n) return;
    const float beta = P.sc->beta;
    for (int c = 0; c < 3; c++) P.d[3 * i + c] = P.z[3 * i + c] + beta * P.d[3 * i + c];
}

// The true residual r = b - A x with FP64 row sums, z = M^-1 r and d = z (the direction
// restarted from it); partials r.r and r.z.
__global__ void __launch_bounds__(NA_TB) k_na_true(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const float* __restrict__ vals, const float* __restrict__ b,
                                                   const float* __restrict__ x, NaPtr P) {
    const int i = blockIdx.x * NA_TB + threadIdx.x;
    float rr = 0.0f, rz = 0.0f;
    if (i < n) {
        double y[3] = {b[3 * i], b[3 * i + 1], b[3 * i + 2]};
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            const int j = ci[k];
            const float* A = vals + (size_t)k * 9;
            for (int u = 0; u < 3; u++)
                y[u] -= (double)A[3 * u] * x[3 * j] + (double)A[3 * u + 1] * x[3 * j + 1] +
                        (double)A[3 * u + 2] * x[3 * j + 2];
        }
        const float ri[3] = {(float)y[0], (float)y[1], (float)y[2]};
        float zi[3];
        mul3(P.Minv + (size_t)i * 9, ri, zi);
        for (int c = 0; c < 3; c++) {
            P.r[3 * i + c] = ri[c];
            P.z[3 * i + c] = zi[c];
            P.d[3 * i + c] = zi[c];
            rr += ri[c] * ri[c];
            rz += ri[c] * zi[c];
        }
    }
    wg_sum2(rr, rz, P.part + 2 * blockIdx.x);
}

// Accept when the true residual meets the tolerance (or the budget is spent); otherwise reopen
// the loop with r.z of the replaced residual.
__global__ void __launch_bounds__(NA_TB) k_na_recheck(int nb, NaPtr P) {
    float a = 0.0f, b = 0.0f;
    for (int k = threadIdx.x; k < nb; k += NA_TB) {
        a += P.part[2 * k];
        b += P.part[2 * k + 1];
    }
    __shared__ float tot[2];
    wg_sum2(a, b, tot);
    __syncthreads();
    if (threadIdx.x != 0) return;
    NaScal* sc = P.sc;
    sc->rr = tot[0];
    sc->rz = tot[1];
    sc->done = (tot[0] <= sc->tol2bb || sc->iters >= sc->maxit || !(tot[0] <= FLT_MAX)) ? 1 : 0;
}

__global__ void k_na_clear(NaScal* sc) {
    if (threadIdx.x == 0) *sc = NaScal{};
}

// The scratch address on each device, looked up once per device with that device current.
std::mutex gNaMutex;
float* gNaBase[64] = {};

float* scratch_base() {
    int dev = 0;
    CS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) throw StatusError(CS_E_NO_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lock(gNaMutex);
    if (!gNaBase[dev]) {
        void* p = nullptr;
        CS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        gNaBase[dev] = static_cast<float*>(p);
    }
    return gNaBase[dev];
}

int solve_noalloc(hipStream_t s, int n, const int* rp, const int* ci, const float* vals, const float* b, float* x,
                  float tol, int maxit, int* iters, float* relres) {
    NaPtr P = carve(scratch_base());
    const int nb = cdiv(n, NA_TB);
    // the scalar record starts cleared: `bad` is only ever raised by the set-up kernel
    hipLaunchKernelGGL(k_na_clear, dim3(1), dim3(64), 0, s, P.sc);
    CS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_na_init, dim3(nb), dim3(NA_TB), 0, s, n, rp, ci, vals, b, x, P);
    CS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_na_reduce, dim3(1), dim3(NA_TB), 0, s, nb, 0, tol * tol, maxit, P);
    CS_LAUNCH_CHECK();
    NaScal h{};
    auto read = [&]() {
        CS_HIP(hipMemcpyAsync(&h, P.sc, sizeof(NaScal), hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
    };
    // The loop stops on its recursive residual; as in cs_solve, only the true residual (FP64 row
    // sums) accepts the result, otherwise the loop goes on from it, at most NA_RESTARTS times.
    for (int restart = 0;; restart++) {
        for (int done = 0; done <= maxit; done += NA_CHUNK) {
            read();
            if (h.done) break;
            for (int k = 0; k < NA_CHUNK; k++) {
                hipLaunchKernelGGL(k_na_spmv, dim3(nb), dim3(NA_TB), 0, s, n, rp, ci, vals, P);
                hipLaunchKernelGGL(k_na_reduce, dim3(1), dim3(NA_TB), 0, s, nb, 1, 0.0f, 0, P);
                hipLaunchKernelGGL(k_na_update, dim3(nb), dim3(NA_TB), 0, s, n, x, P);
                hipLaunchKernelGGL(k_na_reduce, dim3(1), dim3(NA_TB), 0, s, nb, 2, 0.0f, 0, P);
                hipLaunchKernelGGL(k_na_dir, dim3(nb), dim3(NA_TB), 0, s, n, P);
            }
            CS_LAUNCH_CHECK();
        }
        read();
        if (h.bad) return CS_E_ARG;
        if (!h.done) throw StatusError(CS_E_HIP, "no-allocation PCG: the device never reported its stop");
        hipLaunchKernelGGL(k_na_true, dim3(nb), dim3(NA_TB), 0, s, n, rp, ci, vals, b, x, P);
        hipLaunchKernelGGL(k_na_recheck, dim3(1), dim3(NA_TB), 0, s, nb, P);
        CS_LAUNCH_CHECK();
        read();
        if (h.done || restart == NA_RESTARTS) break;
    }
    if (iters) *iters = h.iters;
    if (relres) *relres = h.bb > 0.0f ? sqrtf(h.rr / h.bb) : 0.0f;
    return h.rr <= h.tol2bb ? CS_OK : CS_E_NOT_CONVERGED;
}

}  // namespace

}  // namespace cs

extern "C" cs_status cs_solve_bsr_noalloc(void* stream, int n_nodes, const int* row_ptr, const int* col_idx,
                                          const float* vals, const float* b, float* x, float rel_tol, int max_iters,
                                          int* iters, float* rel_residual) {
    if (n_nodes > CS_NOALLOC_MAX_NODES) return CS_E_CAPACITY;
    if (n_nodes < 1 || !row_ptr || !col_idx || !vals || !b || !x || !(rel_tol > 0.0f) || !(rel_tol < 1.0f) ||
        max_iters < 0)
        return CS_E_ARG;
    try {
        return static_cast<cs_status>(cs::solve_noalloc(static_cast<hipStream_t>(stream), n_nodes, row_ptr, col_idx,
                                                        vals, b, x, rel_tol, max_iters, iters, rel_residual));
    } catch (const cs::StatusError& e) {
        return static_cast<cs_status>(e.status);
    } catch (const std::exception&) {
        return CS_E_HIP;
    }
}
