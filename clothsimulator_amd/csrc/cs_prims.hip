// cs_prims.hip — exclusive scan and stable LSD radix sort for gfx950.
//
// Radix sort: 8-bit digits, 2048-element tiles (256 threads x 8 items, striped so
// item k of thread t is element k*256+t).  Per pass: a tile histogram, a scan of
// the digit-major histogram table, and a scatter whose in-tile rank comes from
// 64-lane ballot matching (8 ballots give the lanes holding the same digit).
// The pass is stable, so sorting bits [b0,b1) of (key, value) pairs whose values
// were emitted in index order yields the order of (key, index).
#include "cs_prims.hpp"

namespace cs {

namespace {

constexpr int SC_BLOCK = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_BLOCK * SC_ITEMS;

// Inclusive scan of one value per thread across a 256-thread block; returns the
// block total through `total`.
__device__ __forceinline__ uint32_t block_scan_incl(uint32_t v, uint32_t* lds /*4*/, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    if (lane == 63) lds[wave] = v;
    __syncthreads();
    uint32_t off = 0;
    for (int w = 0; w < wave; w++) off += lds[w];
    total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return v + off;
}

__global__ void __launch_bounds__(SC_BLOCK) k_scan_reduce(const uint32_t* __restrict__ in, int n,
                                                          uint32_t* __restrict__ sums) {
    __shared__ uint32_t lds[4];
    long base = (long)blockIdx.x * SC_TILE + (long)threadIdx.x * SC_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        long i = base + k;
        if (i < n) s += in[i];
    }
    uint32_t total;
    block_scan_incl(s, lds, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Single block: exclusive scan of nb block sums in place, sums[nb] = total.
__global__ void __launch_bounds__(SC_BLOCK) k_scan_top(uint32_t* sums, int nb) {
    __shared__ uint32_t lds[4];
    uint32_t carry = 0;
    for (int base = 0; base < nb; base += SC_BLOCK) {
        int i = base + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0;
        uint32_t total;
        uint32_t incl = block_scan_incl(v, lds, total);
        if (i < nb) sums[i] = carry + incl - v;
        carry += total;
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void __launch_bounds__(SC_BLOCK) k_scan_down(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                        int n, const uint32_t* __restrict__ sums, int nb) {
    __shared__ uint32_t lds[4];
    long base = (long)blockIdx.x * SC_TILE + (long)threadIdx.x * SC_ITEMS;
    uint32_t v[SC_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        long i = base + k;
        v[k] = i < n ? in[i] : 0;
        s += v[k];
    }
    uint32_t total;
    uint32_t incl = block_scan_incl(s, lds, total);
    uint32_t run = sums[blockIdx.x] + incl - s;
#pragma unroll
    for (int k = 0; k < SC_ITEMS; k++) {
        long i = base + k;
        if (i < n) out[i] = run;
        run += v[k];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == SC_BLOCK - 1) out[n] = sums[nb];
}

constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;

__global__ void __launch_bounds__(RS_BLOCK) k_radix_hist(const uint64_t* __restrict__ keys, int n, int shift,
                                                         int ntiles, uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    long base = (long)blockIdx.x * RS_TILE;
#pragma unroll
    for (int k = 0; k < RS_ITEMS; k++) {
        long i = base + k * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_BLOCK) k_radix_scatter(const uint64_t* __restrict__ kin,
                                                            const uint32_t* __restrict__ vin,
                                                            uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                            int n, int shift, int ntiles,
                                                            const uint32_t* __restrict__ offs) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t wc[4][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    base[tid] = offs[(long)tid * ntiles + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    long tile = (long)blockIdx.x * RS_TILE;
    for (int k = 0; k < RS_ITEMS; k++) {
        wc[0][tid] = 0; wc[1][tid] = 0; wc[2][tid] = 0; wc[3][tid] = 0;
        __syncthreads();
        long i = tile + k * RS_BLOCK + tid;
        bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0ull;
        uint32_t d = valid ? ((uint32_t)(key >> shift) & 255u) : 0u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            int bit = (d >> b) & 1;
            uint64_t bl = __ballot(bit);
            m &= bit ? bl : ~bl;
        }
        uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && rank == 0) wc[wave][d] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t t0 = wc[0][tid], t1 = wc[1][tid], t2 = wc[2][tid], t3 = wc[3][tid];
        __syncthreads();
        wc[0][tid] = 0;
        wc[1][tid] = t0;
        wc[2][tid] = t0 + t1;
        wc[3][tid] = t0 + t1 + t2;
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + wc[wave][d] + rank;
            kout[pos] = key;
            vout[pos] = vin[i];
        }
        __syncthreads();
        base[tid] += t0 + t1 + t2 + t3;
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f32(float* p, float v, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_iota_u32(uint32_t* p, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

}  // namespace

void scan_exclusive(const uint32_t* in, uint32_t* out, int n, Scratch& ws, hipStream_t s) {
    if (n <= 0) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, out, 0u, 1L);
        CS_LAUNCH_CHECK();
        return;
    }
    int nb = cdiv(n, SC_TILE);
    uint32_t* sums = ws.d.as<uint32_t>(nb + 1);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SC_BLOCK), 0, s, in, n, sums);
    CS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(SC_BLOCK), 0, s, sums, nb);
    CS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(SC_BLOCK), 0, s, in, out, n, sums, nb);
    CS_LAUNCH_CHECK();
}

void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp, uint32_t* vals_tmp, int n, int bit0,
                      int bit1, Scratch& ws, hipStream_t s) {
    if (n <= 1 || bit1 <= bit0) return;
    int ntiles = cdiv(n, RS_TILE);
    long hn = 256L * ntiles;
    uint32_t* hist = ws.c.as<uint32_t>(hn + 1);
    uint64_t* ksrc = keys;
    uint32_t* vsrc = vals;
    uint64_t* kdst = keys_tmp;
    uint32_t* vdst = vals_tmp;
    int passes = 0;
    for (int shift = bit0; shift < bit1; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(ntiles), dim3(RS_BLOCK), 0, s, ksrc, n, shift, ntiles, hist);
        CS_LAUNCH_CHECK();
        scan_exclusive(hist, hist, (int)hn, ws, s);
        hipLaunchKernelGGL(k_radix_scatter, dim3(ntiles), dim3(RS_BLOCK), 0, s, ksrc, vsrc, kdst, vdst, n, shift,
                           ntiles, hist);
        CS_LAUNCH_CHECK();
        uint64_t* tk = ksrc; ksrc = kdst; kdst = tk;
        uint32_t* tv = vsrc; vsrc = vdst; vdst = tv;
        passes++;
    }
    if (passes & 1) {
        CS_HIP(hipMemcpyAsync(keys, ksrc, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, s));
        CS_HIP(hipMemcpyAsync(vals, vsrc, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s));
    }
}

void fill_u32(uint32_t* p, uint32_t v, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(cdiv(n, 256)), dim3(256), 0, s, p, v, n);
    CS_LAUNCH_CHECK();
}
void fill_u64(uint64_t* p, uint64_t v, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(n, 256)), dim3(256), 0, s, p, v, n);
    CS_LAUNCH_CHECK();
}
void fill_f32(float* p, float v, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_f32, dim3(cdiv(n, 256)), dim3(256), 0, s, p, v, n);
    CS_LAUNCH_CHECK();
}
void iota_u32(uint32_t* p, long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_iota_u32, dim3(cdiv(n, 256)), dim3(256), 0, s, p, n);
    CS_LAUNCH_CHECK();
}

}  // namespace cs
