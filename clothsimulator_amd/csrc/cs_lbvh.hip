// cs_lbvh.hip — linear BVH over the faces of one mesh and leaf-vs-tree traversal.
//
// Build (reference: BVH::BVH GPU, src/BVH.cu:18-50; src/BVHHelper.cu):
//   leaf boxes over x (and x0 for CCD), scene box, 30-bit Morton code of each
//   box centre with the reference's FP32 operation order, a stable LSD radix
//   sort of the codes carrying the face index (equivalent to sorting the
//   reference's 64-bit key code<<32|face), Karras 2012 internal nodes.  An
//   internal node covers a contiguous range [first, last] of sorted leaves, so
//   its box is the union of that range and its maxIndex (largest sorted leaf
//   index in the subtree) is `last`.  Boxes come from range-union tables
//   (unions of 2^k consecutive leaves for k <= 5, and of 2^k consecutive
//   64-leaf blocks), at most six table reads per node: no thread waits for
//   another and no grid-wide fence is needed.  min/max are exact, so the boxes
//   equal the reference's bottom-up unions bit for bit.
// Traversal (src/BVHHelper.cu:151-307): one thread per query leaf, explicit
//   stack in LDS; a leaf pair is reported when the query box overlaps the
//   target leaf box dilated by `thickness`; for a mesh against itself only
//   leaves with a larger sorted index are reported.
#include <cstring>
#include <string>

#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

constexpr int TB = 64;          // traversal block (one wavefront)
constexpr int STACK = 64;       // traversal stack depth (the reference uses 64)

__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ Box face_box(const float4* __restrict__ x, const float4* __restrict__ x0, int4 fn, bool ccd) {
    Box b = box_empty();
    const int ids[3] = {fn.x, fn.y, fn.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        box_add(b, ld3(x[ids[k]]));
        if (ccd) box_add(b, ld3(x0[ids[k]]));
    }
    return b;
}

__global__ void __launch_bounds__(256) k_leaf_boxes(int f0, int nf, bool ccd, const float4* __restrict__ x,
                                                    const float4* __restrict__ x0, const int4* __restrict__ fnodes,
                                                    float4* __restrict__ fbox, float4* __restrict__ partial) {
    __shared__ float red[6][256];
    int f = blockIdx.x * 256 + threadIdx.x;
    Box b = box_empty();
    if (f < nf) {
        b = face_box(x, x0, fnodes[f0 + f], ccd);
        fbox[2 * f] = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.0f);
        fbox[2 * f + 1] = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.0f);
    }
    for (int k = 0; k < 3; k++) { red[k][threadIdx.x] = b.lo[k]; red[3 + k][threadIdx.x] = b.hi[k]; }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; k++) {
                red[k][threadIdx.x] = tmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
                red[3 + k][threadIdx.x] = tmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = make_float4(red[0][0], red[1][0], red[2][0], 0.0f);
        partial[2 * blockIdx.x + 1] = make_float4(red[3][0], red[4][0], red[5][0], 0.0f);
    }
}

// scene box of the leaf boxes: min/max are exact, so the reduction order is free.
__global__ void __launch_bounds__(256) k_scene_box(int nb, const float4* __restrict__ partial, float4* __restrict__ scene) {
    __shared__ float red[6][256];
    Box b = box_empty();
    for (int i = threadIdx.x; i < nb; i += 256) {
        Box p;
        float4 lo = partial[2 * i], hi = partial[2 * i + 1];
        p.lo[0] = lo.x; p.lo[1] = lo.y; p.lo[2] = lo.z;
        p.hi[0] = hi.x; p.hi[1] = hi.y; p.hi[2] = hi.z;
        box_add(b, p);
    }
    for (int k = 0; k < 3; k++) { red[k][threadIdx.x] = b.lo[k]; red[3 + k][threadIdx.x] = b.hi[k]; }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; k++) {
                red[k][threadIdx.x] = tmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
                red[3 + k][threadIdx.x] = tmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float p0 = red[0][0], p1 = red[1][0], p2 = red[2][0];
        scene[0] = make_float4(p0, p1, p2, 0.0f);
        scene[1] = make_float4(red[3][0] - p0, red[4][0] - p1, red[5][0] - p2, 0.0f);
    }
}

// mortonCode / initializeLeafNodes (src/BVHHelper.cu:18-42)
__global__ void __launch_bounds__(256) k_morton(int nf, const float4* __restrict__ fbox, const float4* __restrict__ scene,
                                                uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= nf) return;
    float4 lo = fbox[2 * f], hi = fbox[2 * f + 1];
    float4 p = scene[0], d = scene[1];
    const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    const float pp[3] = {p.x, p.y, p.z}, dd[3] = {d.x, d.y, d.z};
    uint32_t e[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float s = dd[i] == 0.0f ? 0.0f : (c[i] - pp[i]) / dd[i];
        s = fminf(fmaxf(s * 1024.0f, 0.0f), 1023.0f);
        e[i] = expand_bits((uint32_t)s);
    }
    keys[f] = (uint64_t)((e[0] << 2) + (e[1] << 1) + e[2]);
    vals[f] = (uint32_t)f;
}

// leaves in sorted order; keys become the reference's 64-bit code<<32 | face
__global__ void __launch_bounds__(256) k_fill_leaves(int nf, int f0, int mesh, uint64_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ vals, const float4* __restrict__ fbox,
                                                     BNode* __restrict__ nodes, float4* __restrict__ faceBox) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nf) return;
    uint32_t f = vals[i];
    keys[i] = (keys[i] << 32) | (uint64_t)f;
    float4 lo = fbox[2 * f], hi = fbox[2 * f + 1];
    faceBox[2 * (f0 + (int)f)] = make_float4(lo.x, lo.y, lo.z, __int_as_float(mesh));
    faceBox[2 * (f0 + (int)f) + 1] = make_float4(hi.x, hi.y, hi.z, __int_as_float(i));
    BNode n;
    n.lo[0] = lo.x; n.lo[1] = lo.y; n.lo[2] = lo.z;
    n.hi[0] = hi.x; n.hi[1] = hi.y; n.hi[2] = hi.z;
    n.maxIndex = i;
    n.face = f0 + (int)f;
    nodes[nf - 1 + i] = n;
}

__device__ __forceinline__ int clz64(uint64_t a) { return __clzll((long long)a); }

// findRange / findSplit / initializeInternalNodes (src/BVHHelper.cu:48-103)
__global__ void __launch_bounds__(256) k_karras(int ni, const uint64_t* __restrict__ k, int2* __restrict__ child,
                                                int2* __restrict__ range) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ni) return;
    int left, right;
    if (i == 0) {
        left = 0;
        right = ni;
    } else {
        const uint64_t code = k[i];
        const int dl = clz64(code ^ k[i - 1]);
        const int dr = clz64(code ^ k[i + 1]);
        if (dl > dr) {
            int l = -1, r = i - 1;
            while (r - l > 1) { int mid = (l + r) >> 1; if (clz64(code ^ k[mid]) > dr) r = mid; else l = mid; }
            left = r; right = i;
        } else {
            int l = i + 1, r = ni + 1;
            while (r - l > 1) { int mid = (l + r) >> 1; if (clz64(code ^ k[mid]) > dl) l = mid; else r = mid; }
            left = i; right = l;
        }
    }
    const uint64_t code = k[left];
    const int dd = clz64(code ^ k[right]);
    int l = left, r = right;
    while (r - l > 1) { int mid = (l + r) >> 1; if (clz64(code ^ k[mid]) > dd) l = mid; else r = mid; }
    const int mid = l;
    const int L = (mid == left) ? ni + mid : mid;        // unified index: leaves start at ni = nf-1
    const int R = (mid + 1 == right) ? ni + mid + 1 : mid + 1;
    child[i] = make_int2(L, R);
    range[i] = make_int2(left, right);
}

// A box as six floats (24 B) in the range-union tables.
struct LBox {
    float lo[3], hi[3];
};
__device__ __forceinline__ LBox lbox_empty() {
    LBox b;
    const Box e = box_empty();
    for (int k = 0; k < 3; k++) { b.lo[k] = e.lo[k]; b.hi[k] = e.hi[k]; }
    return b;
}
__device__ __forceinline__ LBox lbox_union(const LBox& a, const LBox& b) {
    LBox c;
    for (int k = 0; k < 3; k++) { c.lo[k] = tmin(a.lo[k], b.lo[k]); c.hi[k] = tmax(a.hi[k], b.hi[k]); }
    return c;
}
__device__ __forceinline__ LBox leaf_lbox(const BNode* nodes, int nf, int j) {
    LBox b;
    if (j >= nf) return lbox_empty();
    const BNode& n = nodes[nf - 1 + j];
    for (int k = 0; k < 3; k++) { b.lo[k] = n.lo[k]; b.hi[k] = n.hi[k]; }
    return b;
}

constexpr int LEAF_LEVELS = 5;  // leaf table levels 1..5 (runs of 2..32 leaves)

// One 128-thread workgroup per 64-leaf block: leaves [64b, 64b + 128) in LDS,
// levels 1..5 of the leaf table for the block's 64 leaves, and the block's own
// box (level 0 of the block table).
__global__ void __launch_bounds__(128) k_leaf_table(int nf, const BNode* __restrict__ nodes, LBox* __restrict__ leafTab,
                                                    LBox* __restrict__ blockTab) {
    __shared__ LBox buf[2][128];
    const int t = threadIdx.x;
    const int j0 = blockIdx.x * 64;
    buf[0][t] = leaf_lbox(nodes, nf, j0 + t);
    __syncthreads();
    int cur = 0;
    for (int k = 1; k <= LEAF_LEVELS + 1; k++) {
        const int h = 1 << (k - 1);
        const LBox v = t + h < 128 ? lbox_union(buf[cur][t], buf[cur][t + h]) : buf[cur][t];
        buf[cur ^ 1][t] = v;
        __syncthreads();
        cur ^= 1;
        if (k <= LEAF_LEVELS) {
            if (t < 64 && j0 + t < nf) leafTab[(size_t)(k - 1) * nf + j0 + t] = v;
        } else if (t == 0) {
            blockTab[blockIdx.x] = v;  // 64 leaves
        }
    }
}

// Level k of the block table from level k - 1.
__global__ void __launch_bounds__(256) k_block_table(int nblk, int k, LBox* __restrict__ blockTab) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i + (1 << k) > nblk) return;
    const LBox* prev = blockTab + (size_t)(k - 1) * nblk;
    blockTab[(size_t)k * nblk + i] = lbox_union(prev[i], prev[i + (1 << (k - 1))]);
}

__device__ __forceinline__ int floor_log2(int v) { return 31 - __clz(v); }

// union of sorted leaves [a, b], b - a < 63 (two overlapping runs of 2^k leaves, k <= 5)
__device__ __forceinline__ LBox short_run(const BNode* nodes, const LBox* leafTab, int nf, int a, int b) {
    const int len = b - a + 1;
    const int k = floor_log2(len);
    if (k == 0) return leaf_lbox(nodes, nf, a);
    const LBox* tab = leafTab + (size_t)(k - 1) * nf;
    return lbox_union(tab[a], tab[b - (1 << k) + 1]);
}

// internal node boxes from their leaf ranges
__device__ __forceinline__ void internal_box(int i, int nf, int nblk, const int2* __restrict__ range,
                                             const LBox* __restrict__ leafTab, const LBox* __restrict__ blockTab,
                                             BNode* __restrict__ nodes) {
    const int2 rg = range[i];
    const int a = rg.x, b = rg.y;
    LBox box;
    if (b - a < 63) {
        box = short_run(nodes, leafTab, nf, a, b);
    } else {
        const int A = (a + 63) & ~63;  // first block boundary at or after a
        const int B = (b + 1) & ~63;   // last block boundary at or before b + 1
        box = lbox_empty();
        if (A > a) box = lbox_union(box, short_run(nodes, leafTab, nf, a, A - 1));
        if (b >= B) box = lbox_union(box, short_run(nodes, leafTab, nf, B, b));
        const int nb = (B - A) >> 6;  // whole blocks A/64 .. B/64 - 1
        if (nb > 0) {
            const int k = floor_log2(nb);
            const LBox* tab = blockTab + (size_t)k * nblk;
            box = lbox_union(box, lbox_union(tab[A >> 6], tab[(B >> 6) - (1 << k)]));
        }
    }
    BNode n;
    for (int k = 0; k < 3; k++) { n.lo[k] = box.lo[k]; n.hi[k] = box.hi[k]; }
    n.maxIndex = b;
    n.face = -1;
    nodes[i] = n;
}

__global__ void __launch_bounds__(256) k_internal_boxes(int nf, int nblk, const int2* __restrict__ range,
                                                        const LBox* __restrict__ leafTab,
                                                        const LBox* __restrict__ blockTab, BNode* __restrict__ nodes) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nf - 1) return;
    internal_box(i, nf, nblk, range, leafTab, blockTab, nodes);
}

// ---------------------------------------------------------------- batched build of small meshes
// One 1024-thread workgroup builds one mesh's whole tree, every small mesh in
// one launch: leaf boxes, scene box, Morton codes, a bitonic sort of the unique
// keys code<<32 | face (the order of the stable radix sort above), leaves,
// Karras nodes, then the same range-union tables and internal boxes as
// k_leaf_table / k_block_table / k_internal_boxes, level by level behind
// workgroup barriers.  Every table entry is the union of a fixed run of sorted
// leaves and min/max are exact, so the trees equal those of the per-mesh
// kernels bit for bit.  Refit of small meshes runs the same way.
constexpr int BB = 1024;

struct BuildJob {
    int f0, nf, mesh, pow2;  // faces, first face, mesh id, sort length (power of two >= nf)
    int nblk, levels;        // 64-leaf blocks and block-table levels
    int ccd;                 // leaf boxes span x and x0
    uint64_t* keys;          // nf: code<<32 | face of the sorted leaves
    uint64_t* tmp;           // pow2: sort buffer
    BNode* nodes;
    int2* child;
    int2* range;
    float4* fbox;
    LBox* leafTab;
    LBox* blockTab;
};

// range-union tables and internal boxes of job J's tree from its current leaf
// boxes and stored ranges (the whole workgroup; ends behind a barrier)
__device__ void tree_boxes(const BuildJob& J, int t) {
    const int nf = J.nf;
    if (nf <= 1) return;
    for (int k = 1; k <= LEAF_LEVELS; k++) {
        const int h = 1 << (k - 1);
        LBox* cur = J.leafTab + (size_t)(k - 1) * nf;
        const LBox* prev = k == 1 ? nullptr : J.leafTab + (size_t)(k - 2) * nf;
        for (int j = t; j < nf; j += BB) {
            const LBox a = k == 1 ? leaf_lbox(J.nodes, nf, j) : prev[j];
            const LBox b = j + h >= nf ? lbox_empty() : (k == 1 ? leaf_lbox(J.nodes, nf, j + h) : prev[j + h]);
            cur[j] = lbox_union(a, b);
        }
        __syncthreads();
    }
    const LBox* run32 = J.leafTab + (size_t)(LEAF_LEVELS - 1) * nf;
    for (int b = t; b < J.nblk; b += BB) {
        const int j = 64 * b;
        J.blockTab[b] = lbox_union(run32[j], j + 32 < nf ? run32[j + 32] : lbox_empty());
    }
    __syncthreads();
    for (int k = 1; k < J.levels; k++) {
        const LBox* prev = J.blockTab + (size_t)(k - 1) * J.nblk;
        LBox* cur = J.blockTab + (size_t)k * J.nblk;
        for (int i = t; i + (1 << k) <= J.nblk; i += BB) cur[i] = lbox_union(prev[i], prev[i + (1 << (k - 1))]);
        __syncthreads();
    }
    for (int i = t; i < nf - 1; i += BB) internal_box(i, nf, J.nblk, J.range, J.leafTab, J.blockTab, J.nodes);
    __syncthreads();
}

__global__ void __launch_bounds__(BB) k_bvh_build_batch(const BuildJob* __restrict__ jobs, bool ccd,
                                                        const float4* __restrict__ x, const float4* __restrict__ x0,
                                                        const int4* __restrict__ fnodes,
                                                        float4* __restrict__ faceBox) {
    __shared__ float red[6][BB / 64];
    __shared__ float scene[6];
    const BuildJob J = jobs[blockIdx.x];
    const int t = threadIdx.x;
    // leaf boxes and the scene box
    Box acc = box_empty();
    for (int f = t; f < J.nf; f += BB) {
        const Box b = face_box(x, x0, fnodes[J.f0 + f], ccd);
        J.fbox[2 * f] = make_float4(b.lo[0], b.lo[1], b.lo[2], 0.0f);
        J.fbox[2 * f + 1] = make_float4(b.hi[0], b.hi[1], b.hi[2], 0.0f);
        box_add(acc, b);
    }
    for (int k = 0; k < 3; k++)
        for (int o = 32; o > 0; o >>= 1) {
            acc.lo[k] = tmin(acc.lo[k], __shfl_xor(acc.lo[k], o, 64));
            acc.hi[k] = tmax(acc.hi[k], __shfl_xor(acc.hi[k], o, 64));
        }
    if ((t & 63) == 0)
        for (int k = 0; k < 3; k++) { red[k][t >> 6] = acc.lo[k]; red[3 + k][t >> 6] = acc.hi[k]; }
    __syncthreads();
    if (t == 0) {
        Box b = box_empty();
        for (int w = 0; w < BB / 64; w++)
            for (int k = 0; k < 3; k++) { b.lo[k] = tmin(b.lo[k], red[k][w]); b.hi[k] = tmax(b.hi[k], red[3 + k][w]); }
        for (int k = 0; k < 3; k++) { scene[k] = b.lo[k]; scene[3 + k] = b.hi[k] - b.lo[k]; }
    }
    __syncthreads();
    // Morton codes (as k_morton) into the sort buffer, padded with the largest key
    for (int f = t; f < J.pow2; f += BB) {
        uint64_t key = ~0ull;
        if (f < J.nf) {
            const float4 lo = J.fbox[2 * f], hi = J.fbox[2 * f + 1];
            const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
            uint32_t e[3];
            for (int i = 0; i < 3; i++) {
                float q = scene[3 + i] == 0.0f ? 0.0f : (c[i] - scene[i]) / scene[3 + i];
                q = fminf(fmaxf(q * 1024.0f, 0.0f), 1023.0f);
                e[i] = expand_bits((uint32_t)q);
            }
            key = ((uint64_t)((e[0] << 2) + (e[1] << 1) + e[2]) << 32) | (uint64_t)f;
        }
        J.tmp[f] = key;
    }
    __syncthreads();
    // bitonic sort, ascending
    for (int k = 2; k <= J.pow2; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = t; i < J.pow2; i += BB) {
                const int l = i ^ j;
                if (l > i) {
                    const uint64_t a = J.tmp[i], b = J.tmp[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        J.tmp[i] = b;
                        J.tmp[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    // leaves (as k_fill_leaves)
    for (int i = t; i < J.nf; i += BB) {
        const uint64_t key = J.tmp[i];
        const uint32_t f = (uint32_t)key;
        J.keys[i] = key;
        const float4 lo = J.fbox[2 * f], hi = J.fbox[2 * f + 1];
        faceBox[2 * (J.f0 + (int)f)] = make_float4(lo.x, lo.y, lo.z, __int_as_float(J.mesh));
        faceBox[2 * (J.f0 + (int)f) + 1] = make_float4(hi.x, hi.y, hi.z, __int_as_float(i));
        BNode n;
        n.lo[0] = lo.x; n.lo[1] = lo.y; n.lo[2] = lo.z;
        n.hi[0] = hi.x; n.hi[1] = hi.y; n.hi[2] = hi.z;
        n.maxIndex = i;
        n.face = J.f0 + (int)f;
        J.nodes[J.nf - 1 + i] = n;
    }
    __syncthreads();
    // internal nodes (as k_karras)
    const int ni = J.nf - 1;
    const uint64_t* k = J.keys;
    for (int i = t; i < ni; i += BB) {
        int left, right;
        if (i == 0) {
            left = 0;
            right = ni;
        } else {
            const uint64_t code = k[i];
            const int dl = clz64(code ^ k[i - 1]);
            const int dr = clz64(code ^ k[i + 1]);
            if (dl > dr) {
                int l = -1, r = i - 1;
                while (r - l > 1) { int mid = (l + r) >> 1; if (clz64(code ^ k[mid]) > dr) r = mid; else l = mid; }
                left = r; right = i;
            } else {
                int l = i + 1, r = ni + 1;
                while (r - l > 1) { int mid = (l + r) >> 1; if (clz64(code ^ k[mid]) > dl) l = mid; else r = mid; }
                left = i; right = l;
            }
        }
        const uint64_t code = k[left];
        const int dd = clz64(code ^ k[right]);
        int l = left, r = right;
        while (r - l > 1) { int mid = (l + r) >> 1; if (clz64(code ^ k[mid]) > dd) l = mid; else r = mid; }
        const int mid = l;
        J.child[i] = make_int2(mid == left ? ni + mid : mid, mid + 1 == right ? ni + mid + 1 : mid + 1);
        J.range[i] = make_int2(left, right);
    }
    __syncthreads();
    tree_boxes(J, t);
}

// Refit of small meshes, one workgroup each (as k_refit_leaves + internal_boxes).
__global__ void __launch_bounds__(BB) k_bvh_refit_batch(const BuildJob* __restrict__ jobs, const float4* __restrict__ x,
                                                        const float4* __restrict__ x0, const int4* __restrict__ fnodes,
                                                        float4* __restrict__ faceBox) {
    const BuildJob J = jobs[blockIdx.x];
    const int t = threadIdx.x;
    for (int i = t; i < J.nf; i += BB) {
        BNode& n = J.nodes[J.nf - 1 + i];
        const Box b = face_box(x, x0, fnodes[n.face], J.ccd != 0);
        for (int c = 0; c < 3; c++) { n.lo[c] = b.lo[c]; n.hi[c] = b.hi[c]; }
        float4* r = faceBox + 2 * (size_t)n.face;
        r[0] = make_float4(b.lo[0], b.lo[1], b.lo[2], r[0].w);
        r[1] = make_float4(b.hi[0], b.hi[1], b.hi[2], r[1].w);
    }
    __syncthreads();
    tree_boxes(J, t);
}

// refit: recompute leaf boxes in place (sorted order, maxIndex unchanged)
__global__ void __launch_bounds__(256) k_refit_leaves(int nf, bool ccd, const float4* __restrict__ x,
                                                      const float4* __restrict__ x0, const int4* __restrict__ fnodes,
                                                      BNode* __restrict__ nodes, float4* __restrict__ faceBox) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nf) return;
    BNode& n = nodes[nf - 1 + i];
    Box b = face_box(x, x0, fnodes[n.face], ccd);
    for (int k = 0; k < 3; k++) { n.lo[k] = b.lo[k]; n.hi[k] = b.hi[k]; }
    float4* r = faceBox + 2 * (size_t)n.face;  // mesh and sorted index (.w) are unchanged
    r[0] = make_float4(b.lo[0], b.lo[1], b.lo[2], r[0].w);
    r[1] = make_float4(b.hi[0], b.hi[1], b.hi[2], r[1].w);
}

__device__ __forceinline__ Box node_box(const BNode& n) {
    Box b;
    for (int k = 0; k < 3; k++) { b.lo[k] = n.lo[k]; b.hi[k] = n.hi[k]; }
    return b;
}

// A tree node (32 B) read through a global-address-space pointer: two 16-byte loads.
typedef int i2n __attribute__((ext_vector_type(2)));
__device__ __forceinline__ BNode load_node(const G1<f4n>* t, int i) {
    const f4n a = t[2 * i], b = t[2 * i + 1];
    BNode n;
    n.lo[0] = a.x; n.lo[1] = a.y; n.lo[2] = a.z; n.maxIndex = __float_as_int(a.w);
    n.hi[0] = b.x; n.hi[1] = b.y; n.hi[2] = b.z; n.face = __float_as_int(b.w);
    return n;
}

// Leaf-vs-tree traversal for query leaf qi.  emit(face) for each reported leaf.
constexpr uint32_t LEAF_PAIRS = 1u << 16;  // candidate pairs of one query face at most

// Returns false if the stack overflowed or emit() refused a leaf (the query is then incomplete).
template <bool SELF, class Emit>
__device__ __forceinline__ bool traverse(const G1<f4n>* __restrict__ tn, const G1<i2n>* __restrict__ tchild, int tnf,
                                         const Box& qb, int qi, float th, int* stk, Emit&& emit) {
    if (tnf == 1) {  // a one-face target has no internal node
        const BNode t0 = load_node(tn, 0);
        if (!SELF && box_overlap(qb, node_box(t0), th)) return emit(t0.face);
        return true;
    }
    const int ni = tnf - 1;
    int top = 0;
    int node = 0;
    for (;;) {
        const i2n ch = tchild[node];
        const BNode L = load_node(tn, ch.x), R = load_node(tn, ch.y);
        const bool lLeaf = ch.x >= ni, rLeaf = ch.y >= ni;
        const bool ovL = (!SELF || qi < L.maxIndex) && box_overlap(qb, node_box(L), th);
        const bool ovR = (!SELF || qi < R.maxIndex) && box_overlap(qb, node_box(R), th);
        if (ovL && lLeaf && !emit(L.face)) return false;
        if (ovR && rLeaf && !emit(R.face)) return false;
        const bool tL = ovL && !lLeaf, tR = ovR && !rLeaf;
        if (!tL && !tR) {
            if (top == 0) break;
            node = stk[(--top) * TB];
        } else {
            node = tL ? ch.x : ch.y;
            if (tL && tR) {
                if (top >= STACK) return false;
                stk[(top++) * TB] = ch.y;
            }
        }
    }
    return true;
}

// One traversal of a batch: the leaves of query tree q against tree t.
struct TravJob {
    int qnf, tnf, self, blk0;  // query leaves, target leaves, q == t, first block of the batch grid
    long off;                  // position of query leaf 0 in the concatenated per-leaf arrays
    const BNode* qn;
    const BNode* tn;
    const int2* tchild;
};

// Every traversal of the batch in one grid (each block belongs to one job, found
// by binary search over the jobs' first blocks).  MODE 0: count each query
// leaf's pairs into num[off + qi]; MODE 1: emit them at offs[off + qi].
template <int MODE>
__global__ void __launch_bounds__(TB) k_traverse_batch(const TravJob* __restrict__ jobs, int njobs, float th,
                                                       uint32_t* __restrict__ num, const uint32_t* __restrict__ offs,
                                                       int2* __restrict__ out, unsigned* __restrict__ overflow) {
    __shared__ int stk[STACK * TB];
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].blk0 <= (int)blockIdx.x) lo = mid; else hi = mid - 1;
    }
    const TravJob J = jobs[lo];
    const int qi = ((int)blockIdx.x - J.blk0) * TB + threadIdx.x;
    if (qi >= J.qnf) return;
    const BNode q = load_node((const G1<f4n>*)J.qn, J.qnf - 1 + qi);
    const Box qb = node_box(q);
    const int qface = q.face;
    uint32_t cnt = 0;
    const uint32_t base = MODE == 1 ? offs[J.off + qi] : 0;
    auto emit = [&](int f) {
        if (cnt >= LEAF_PAIRS) return false;  // far beyond any real contact: positions not finite
        if (MODE == 1) out[base + cnt] = make_int2(qface, f);
        cnt++;
        return true;
    };
    const G1<f4n>* tn = (const G1<f4n>*)J.tn;
    const G1<i2n>* tchild = (const G1<i2n>*)J.tchild;
    const bool ok = J.self ? traverse<true>(tn, tchild, J.tnf, qb, qi, th, stk + threadIdx.x, emit)
                           : traverse<false>(tn, tchild, J.tnf, qb, qi, th, stk + threadIdx.x, emit);
    if (MODE == 0) num[J.off + qi] = cnt;
    if (!ok) atomicOr(overflow, cnt >= LEAF_PAIRS ? 2u : 1u);
}

// The exact 64-bit sum of the per-leaf pair counts (one atomic per block: an integer sum, so the
// order does not matter).  The exclusive scan that places the pairs is 32-bit: a total at or
// above 2^32 would wrap it, so the host checks this sum before the emit pass.
__global__ void __launch_bounds__(256) k_count_total(const uint32_t* __restrict__ num, long n,
                                                     unsigned long long* __restrict__ total) {
    __shared__ unsigned long long part[4];
    unsigned long long v = 0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) v += num[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(total, part[0] + part[1] + part[2] + part[3]);
}

static int floor_log2_host(int v) {
    int k = 0;
    while ((2L << k) <= v) k++;
    return k;
}

void ensure_bvh(cs_ctx* C, Bvh& B, int nf) {
    B.nf = nf;
    B.keys.as<uint64_t>(nf);
    B.keysTmp.as<uint64_t>(nf);
    B.vals.as<uint32_t>(nf);
    B.valsTmp.as<uint32_t>(nf);
    B.nodes.as<BNode>(2L * nf - 1);
    B.child.as<int2>(nf > 1 ? nf - 1 : 1);
    B.range.as<int2>(nf > 1 ? nf - 1 : 1);
    B.leafTab.as<LBox>((size_t)LEAF_LEVELS * nf);
    const int nblk = cdiv(nf, 64);
    B.blockTab.as<LBox>((size_t)(floor_log2_host(nblk) + 1) * nblk);
    B.fbox.as<float4>(2L * nf);
    B.partial.as<float4>(2L * cdiv(nf, 256));
    B.scene.as<float4>(2);
}

// internal boxes and maxIndex of the built tree from its current leaf boxes
void internal_boxes(cs_ctx* C, Bvh& B) {
    if (B.nf <= 1) return;
    hipStream_t s = C->stream;
    const int nblk = cdiv(B.nf, 64);
    const int levels = floor_log2_host(nblk) + 1;
    hipLaunchKernelGGL(k_leaf_table, dim3(nblk), dim3(128), 0, s, B.nf, B.nodes.ptr<BNode>(), B.leafTab.ptr<LBox>(),
                       B.blockTab.ptr<LBox>());
    CS_LAUNCH_CHECK();
    for (int k = 1; k < levels; k++) {
        hipLaunchKernelGGL(k_block_table, dim3(cdiv(nblk - (1 << k) + 1, 256)), dim3(256), 0, s, nblk, k,
                           B.blockTab.ptr<LBox>());
        CS_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_internal_boxes, dim3(cdiv(B.nf - 1, 256)), dim3(256), 0, s, B.nf, nblk, B.range.ptr<int2>(),
                       B.leafTab.ptr<LBox>(), B.blockTab.ptr<LBox>(), B.nodes.ptr<BNode>());
    CS_LAUNCH_CHECK();
}

}  // namespace

void bvh_build(cs_ctx* C, int mesh, bool ccd) {
    const MeshInfo& M = C->meshes[mesh];
    Bvh& B = C->bvh[mesh];
    B.mesh = mesh;
    B.ccd = ccd;
    ensure_bvh(C, B, M.nf);
    hipStream_t s = C->stream;
    const int nb = cdiv(M.nf, 256);
    hipLaunchKernelGGL(k_leaf_boxes, dim3(nb), dim3(256), 0, s, M.f0, M.nf, ccd, C->x.ptr<float4>(),
                       C->x0.ptr<float4>(), C->fnodes.ptr<int4>(), B.fbox.ptr<float4>(), B.partial.ptr<float4>());
    CS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scene_box, dim3(1), dim3(256), 0, s, nb, B.partial.ptr<float4>(), B.scene.ptr<float4>());
    CS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_morton, dim3(nb), dim3(256), 0, s, M.nf, B.fbox.ptr<float4>(), B.scene.ptr<float4>(),
                       B.keys.ptr<uint64_t>(), B.vals.ptr<uint32_t>());
    CS_LAUNCH_CHECK();
    radix_sort_pairs(B.keys.ptr<uint64_t>(), B.vals.ptr<uint32_t>(), B.keysTmp.ptr<uint64_t>(),
                     B.valsTmp.ptr<uint32_t>(), M.nf, 0, 30, C->ws, s);
    float4* faceBox = C->faceBox.as<float4>(2 * (size_t)C->F);  // one array over all meshes: fixed size
    hipLaunchKernelGGL(k_fill_leaves, dim3(nb), dim3(256), 0, s, M.nf, M.f0, mesh, B.keys.ptr<uint64_t>(),
                       B.vals.ptr<uint32_t>(), B.fbox.ptr<float4>(), B.nodes.ptr<BNode>(), faceBox);
    CS_LAUNCH_CHECK();
    if (M.nf > 1) {
        hipLaunchKernelGGL(k_karras, dim3(cdiv(M.nf - 1, 256)), dim3(256), 0, s, M.nf - 1, B.keys.ptr<uint64_t>(),
                           B.child.ptr<int2>(), B.range.ptr<int2>());
        CS_LAUNCH_CHECK();
        internal_boxes(C, B);
    }
}

constexpr long MAX_PAIRS = 1L << 28;  // candidate pairs per traversal pass at most
constexpr int SMALL_BVH = 1 << 15;  // meshes up to this many faces build and refit in one batched launch
// One workgroup builds a tree in time growing with its size (2 ms at 32,768 faces, latency-bound
// on one CU); the per-mesh kernels spread a tree over the chip for about 35 launches.  Trees
// above ALONE_FACES take the per-mesh path when at most ALONE_MAX of them are in the stage.
constexpr int ALONE_FACES = 1 << 12;
constexpr int ALONE_MAX = 4;

// Meshes of [m0, m1) that build and refit with the per-mesh kernels.
static std::vector<bool> per_mesh(const cs_ctx* C, int m0, int m1) {
    int big = 0;
    for (int m = m0; m < m1; m++) big += C->meshes[m].nf > ALONE_FACES && C->meshes[m].nf <= SMALL_BVH;
    std::vector<bool> alone(m1 - m0);
    for (int m = m0; m < m1; m++) {
        const int nf = C->meshes[m].nf;
        alone[m - m0] = nf > SMALL_BVH || (nf > ALONE_FACES && big <= ALONE_MAX);
    }
    return alone;
}

static BuildJob small_job(const MeshInfo& M, Bvh& B) {
    BuildJob j;
    j.f0 = M.f0;
    j.nf = M.nf;
    j.mesh = B.mesh;
    j.pow2 = 1;
    while (j.pow2 < M.nf) j.pow2 <<= 1;
    j.nblk = cdiv(M.nf, 64);
    j.levels = floor_log2_host(j.nblk) + 1;
    j.ccd = B.ccd ? 1 : 0;
    j.keys = B.keys.ptr<uint64_t>();
    j.tmp = nullptr;
    j.nodes = B.nodes.ptr<BNode>();
    j.child = B.child.ptr<int2>();
    j.range = B.range.ptr<int2>();
    j.fbox = B.fbox.ptr<float4>();
    j.leafTab = B.leafTab.ptr<LBox>();
    j.blockTab = B.blockTab.ptr<LBox>();
    return j;
}

// Launches kern over the job table (one workgroup per job) and waits for it:
// the host copy of the table must outlive the copy to the device.
template <class K, class... A>
static void launch_jobs(cs_ctx* C, const std::vector<BuildJob>& jobs, K kern, A... args) {
    hipStream_t s = C->stream;
    BuildJob* dj = C->bvhJobs.as<BuildJob>(jobs.size());
    CS_HIP(hipMemcpyAsync(dj, jobs.data(), sizeof(BuildJob) * jobs.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(kern, dim3((unsigned)jobs.size()), dim3(BB), 0, s, dj, args...);
    CS_LAUNCH_CHECK();
    CS_HIP(hipStreamSynchronize(s));
}

// Every mesh's tree: small meshes in one batched launch, larger ones with the
// per-mesh kernels.
void bvh_build_all(cs_ctx* C, bool ccd) {
    std::vector<BuildJob> jobs;
    size_t tmpLen = 0;
    const std::vector<bool> alone = per_mesh(C, 0, (int)C->meshes.size());
    for (size_t m = 0; m < C->meshes.size(); m++) {
        const MeshInfo& M = C->meshes[m];
        if (alone[m] || M.nf < 1) {
            bvh_build(C, (int)m, ccd);
            continue;
        }
        Bvh& B = C->bvh[m];
        B.mesh = (int)m;
        B.ccd = ccd;
        ensure_bvh(C, B, M.nf);
        jobs.push_back(small_job(M, B));
        tmpLen += jobs.back().pow2;
    }
    if (jobs.empty()) return;
    uint64_t* tmp = C->bvhSortTmp.as<uint64_t>(tmpLen);
    size_t o = 0;
    for (BuildJob& j : jobs) { j.tmp = tmp + o; o += j.pow2; }
    float4* faceBox = C->faceBox.as<float4>(2 * (size_t)C->F);
    launch_jobs(C, jobs, k_bvh_build_batch, ccd, (const float4*)C->x.ptr<float4>(), (const float4*)C->x0.ptr<float4>(),
                (const int4*)C->fnodes.ptr<int4>(), faceBox);
}

// Refit of meshes [m0, m1) (each built before): small meshes batched.
void bvh_refit_range(cs_ctx* C, int m0, int m1) {
    std::vector<BuildJob> jobs;
    const std::vector<bool> alone = per_mesh(C, 0, (int)C->meshes.size());
    for (int m = m0; m < m1; m++) {
        Bvh& B = C->bvh[m];
        if (B.nf < 1) continue;
        if (alone[m]) {
            bvh_refit(C, m);
            continue;
        }
        jobs.push_back(small_job(C->meshes[m], B));
    }
    if (jobs.empty()) return;
    launch_jobs(C, jobs, k_bvh_refit_batch, (const float4*)C->x.ptr<float4>(), (const float4*)C->x0.ptr<float4>(),
                (const int4*)C->fnodes.ptr<int4>(), C->faceBox.ptr<float4>());
}

void bvh_refit(cs_ctx* C, int mesh) {
    Bvh& B = C->bvh[mesh];
    hipLaunchKernelGGL(k_refit_leaves, dim3(cdiv(B.nf, 256)), dim3(256), 0, C->stream, B.nf, B.ccd,
                       C->x.ptr<float4>(), C->x0.ptr<float4>(), C->fnodes.ptr<int4>(), B.nodes.ptr<BNode>(),
                       C->faceBox.ptr<float4>());
    CS_LAUNCH_CHECK();
    internal_boxes(C, B);
}

// Simulator::traverse order (src/Simulator.cu:179-196): for each cloth, itself,
// then every earlier cloth, then every obstacle.  Independent instances
// (CS_FLAG_INSTANCES): cloth c against itself and obstacle c only.
template <class F> static void for_each_traversal(cs_ctx* C, F&& f) {
    if (C->P.flags & CS_FLAG_INSTANCES) {
        for (int c = 0; c < C->nCloths; c++) {
            f(c, c, true);
            if (C->nCloths + c < (int)C->meshes.size()) f(c, C->nCloths + c, false);
        }
        return;
    }
    for (int c = 0; c < C->nCloths; c++) {
        f(c, c, true);
        for (int j = 0; j < c; j++) f(c, j, false);
        for (int o = C->nCloths; o < (int)C->meshes.size(); o++) f(c, o, false);
    }
}

// All candidate pairs of the traversal order, written to C->pairs without any
// atomic: a count pass per traversal into one concatenated array of per-leaf
// counts, one exclusive scan over all of it (so the offsets are global pair
// positions), one host read of the total, and an emit pass per traversal at the
// scanned offsets.  The order is deterministic: traversal, then query leaf,
// then tree order.  One host synchronisation per call, however many meshes.
long bvh_all_pairs(cs_ctx* C, float th) {
    hipStream_t s = C->stream;
    unsigned* flag = C->pairCount.as<unsigned>(4);  // [1] stack overflow, [2..3] 64-bit pair total
    fill_u32(flag, 0u, 4, s);
    std::vector<TravJob> jobs;
    long len = 0;
    int blocks = 0;
    for_each_traversal(C, [&](int q, int t, bool self) {
        const Bvh& Q = C->bvh[q];
        const Bvh& T = C->bvh[t];
        if (Q.nf < 1 || T.nf < 1) return;
        TravJob j;
        j.qnf = Q.nf;
        j.tnf = T.nf;
        j.self = self ? 1 : 0;
        j.blk0 = blocks;
        j.off = len;
        j.qn = Q.nodes.ptr<BNode>();
        j.tn = T.nodes.ptr<BNode>();
        j.tchild = T.child.ptr<int2>();
        jobs.push_back(j);
        len += Q.nf;
        blocks += cdiv(Q.nf, TB);
    });
    if (len == 0) return 0;
    uint32_t* num = C->ws.a.as<uint32_t>(len + 1);
    uint32_t* offs = C->ws.b.as<uint32_t>(len + 1);
    TravJob* dj = C->travJobs.as<TravJob>(jobs.size());
    CS_HIP(hipMemcpyAsync(dj, jobs.data(), sizeof(TravJob) * jobs.size(), hipMemcpyHostToDevice, s));
    const int nj = (int)jobs.size();
    hipLaunchKernelGGL(k_traverse_batch<0>, dim3(blocks), dim3(TB), 0, s, (const TravJob*)dj, nj, th, num,
                       (const uint32_t*)nullptr, (int2*)nullptr, flag + 1);
    CS_LAUNCH_CHECK();
    unsigned long long* total64 = reinterpret_cast<unsigned long long*>(flag + 2);
    hipLaunchKernelGGL(k_count_total, dim3(std::min(cdiv(len, 256), 1024)), dim3(256), 0, s, num, len, total64);
    CS_LAUNCH_CHECK();
    scan_exclusive(num, offs, (int)len, C->ws, s);
    CS_HIP(hipMemcpyAsync(C->hCount, offs + len, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    CS_HIP(hipMemcpyAsync(C->hCount + 1, flag + 1, 3 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    CS_HIP(hipStreamSynchronize(s));  // also ends the job table's host-to-device copy
    if (C->hCount[1] & 2u)
        throw StatusError(CS_E_CAPACITY, "LBVH traversal: a face has more than 65536 candidate pairs "
                                         "(positions not finite?)");
    if (C->hCount[1]) throw StatusError(CS_E_CAPACITY, "LBVH traversal stack overflow (tree deeper than 64)");
    unsigned long long pairs64 = 0;
    std::memcpy(&pairs64, C->hCount + 2, sizeof(pairs64));
    // Every box overlaps every other once a position is not finite (NaN compares false), which
    // puts the pair count in the tens of billions at 513^2; the cap keeps the pair list, the
    // 32-bit scans over it and the CCD's 15 tests per pair in range.
    if (pairs64 > (unsigned long long)MAX_PAIRS)
        throw StatusError(CS_E_CAPACITY, "LBVH traversal: " + std::to_string(pairs64) +
                                             " candidate pairs exceed the limit of 2^28 (positions not finite?)");
    const long total = C->hCount[0];
    if (total == 0) return 0;
    C->pairs.grow_keep(sizeof(int2) * (size_t)total, 0, s);
    C->pairCap = C->pairs.bytes / sizeof(int2);
    hipLaunchKernelGGL(k_traverse_batch<1>, dim3(blocks), dim3(TB), 0, s, (const TravJob*)dj, nj, th,
                       (uint32_t*)nullptr, (const uint32_t*)offs, C->pairs.ptr<int2>(), flag + 1);
    CS_LAUNCH_CHECK();
    return total;
}

long bvh_pairs_ordered(cs_ctx* C, float th, int* host_out, long cap) {
    long total = bvh_all_pairs(C, th);
    if (host_out && total > 0) {
        long m = std::min(total, cap);
        CS_HIP(hipMemcpyAsync(host_out, C->pairs.ptr<int2>(), sizeof(int2) * m, hipMemcpyDeviceToHost, C->stream));
        CS_HIP(hipStreamSynchronize(C->stream));
    }
    return total;
}

}  // namespace cs
