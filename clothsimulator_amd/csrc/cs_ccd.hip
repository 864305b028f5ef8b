// cs_ccd.hip — continuous collision detection in FP64 and the independent impact set.
//
// findImpacts (src/Simulator.cu:383-392) / checkImpactsGpu
// (src/CollisionHelper.cu:137-162): for each candidate face pair of the CCD
// LBVH (boxes over x0 and x), 6 vertex-face and 9 edge-edge tests, each with a
// swept-box prefilter and the coplanarity cubic of checkImpact
// (src/CollisionHelper.cu:62-109) solved in FP64.  An edge-edge pair is always
// evaluated with the lower global edge id first, so a test reached from
// several face pairs yields one bit-identical result.  Two stages: a per-pair
// filter keeps every distinct test once (filters + ownership, below), then a
// per-test kernel runs the FP64 cubic; impacts are sorted by (type, ids),
// de-duplicated, and sorted by (t, type, ids).
//
// independentImpacts GPU (src/Simulator.cu:422-461, kernels
// src/CollisionHelper.cu:164-237): rounds in which a candidate is accepted when
// it is the lowest-sorted live candidate on each of its eligible nodes; this
// equals the sequential greedy pass in sorted order.
#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

// Roots of a t^3 + b t^2 + c t + d on [0, 1], ascending, by isolation: [0, 1]
// is cut at the critical points, each piece with a sign change holds one root
// found by safeguarded Newton (bisection fallback) to |dt| < 1e-12, and exact
// zeros at cut points are roots.  Only roots in [0, 1] matter to the caller, so
// no iteration is spent on the reference's whole-line Newton search
// (src/CollisionHelper.cu:3-60).
__device__ __forceinline__ double cubic_eval(double a, double b, double c, double d, double t) {
    return d + t * (c + t * (b + t * a));
}
__device__ int cubic_roots_unit(double a, double b, double c, double d, double out[4]) {
    double cut[4];
    int nc = 0;
    cut[nc++] = 0.0;
    const double qa = 3.0 * a, qb = 2.0 * b, qc = c;
    double cr[2];
    int ncr = 0;
    if (qa != 0.0) {
        double disc = qb * qb - 4.0 * qa * qc;
        if (disc >= 0.0) {
            double sq = sqrt(disc);
            double q = -0.5 * (qb + (qb < 0.0 ? -sq : sq));
            double r1 = q / qa;
            double r2 = q != 0.0 ? qc / q : r1;
            if (r2 < r1) { double tt = r1; r1 = r2; r2 = tt; }
            cr[ncr++] = r1;
            if (r2 != r1) cr[ncr++] = r2;
        }
    } else if (qb != 0.0) {
        cr[ncr++] = -qc / qb;
    }
    for (int i = 0; i < ncr; i++)
        if (cr[i] > 0.0 && cr[i] < 1.0) cut[nc++] = cr[i];
    cut[nc++] = 1.0;
    int n = 0;
    for (int s = 0; s + 1 < nc; s++) {
        double l = cut[s], r = cut[s + 1];
        double fl = cubic_eval(a, b, c, d, l), fr = cubic_eval(a, b, c, d, r);
        if (fl == 0.0) {
            if (n == 0 || out[n - 1] != l) out[n++] = l;
            continue;
        }
        if (fr == 0.0 || (fl < 0.0) == (fr < 0.0)) continue;
        double t = 0.5 * (l + r);
        for (int it = 0; it < 100; it++) {
            double ft = cubic_eval(a, b, c, d, t);
            if (ft == 0.0) break;
            if ((ft < 0.0) == (fl < 0.0)) { l = t; fl = ft; } else { r = t; }
            double df = c + t * (2.0 * b + 3.0 * t * a);
            double tn = df != 0.0 ? t - ft / df : 0.5 * (l + r);
            if (!(tn > l && tn < r)) tn = 0.5 * (l + r);
            double step = tn - t;
            t = tn;
            if (tabs(step) < 1e-12 || r - l < 1e-12) break;
        }
        out[n++] = t;
    }
    double f1 = cubic_eval(a, b, c, d, 1.0);
    if (f1 == 0.0 && (n == 0 || out[n - 1] != 1.0)) out[n++] = 1.0;
    return n;
}

// checkImpact (src/CollisionHelper.cu:62-109) in FP64
__device__ bool check_impact(int type, const V3d X0[4], const V3d X1[4], cs_impact_rec& imp) {
    V3d x0 = X0[0];
    V3d v0 = X1[0] - x0;
    V3d x1 = X0[1] - x0, x2 = X0[2] - x0, x3 = X0[3] - x0;
    V3d v1 = (X1[1] - X0[1]) - v0, v2 = (X1[2] - X0[2]) - v0, v3 = (X1[3] - X0[3]) - v0;
    double a0 = mixed(x1, x2, x3);
    double a1 = mixed(v1, x2, x3) + mixed(x1, v2, x3) + mixed(x1, x2, v3);
    double a2 = mixed(x1, v2, v3) + mixed(v1, x2, v3) + mixed(v1, v2, x3);
    double a3 = mixed(v1, v2, v3);
    double t[4];
    int ns = cubic_roots_unit(a3, a2, a1, a0, t);
    for (int i = 0; i < ns; i++) {
        if (t[i] < 0.0 || t[i] > 1.0) continue;
        V3d p[4];
        for (int k = 0; k < 4; k++) p[k] = X0[k] + t[i] * (X1[k] - X0[k]);
        V3d n;
        double w[4], d;
        bool ok, inside;
        if (type == 0) {
            ok = signed_vf(p[0], p[1], p[2], p[3], n, w, d);
            inside = ok && (tmin(-w[1], tmin(-w[2], -w[3])) >= -1e-6);
        } else {
            ok = signed_ee(p[0], p[1], p[2], p[3], n, w, d);
            inside = ok && (tmin(tmin(w[0], w[1]), tmin(-w[2], -w[3])) >= -1e-6);
        }
        if (!ok) continue;
        if (dot(n, w[1] * v1 + w[2] * v2 + w[3] * v3) > 0.0) n = -n;
        if (tabs(d) < 1e-6 && inside) {
            imp.t = t[i] + 0.0;  // a root of -0.0 is reported as +0.0 (t keys are sorted by their bits)
            for (int k = 0; k < 4; k++) imp.w[k] = w[k];
            imp.n[0] = n.x; imp.n[1] = n.y; imp.n[2] = n.z;
            return true;
        }
    }
    return false;
}

// ---- stage 1: per candidate face pair, the cheap filters and ownership ----
//
// A vertex-face test (n, f) or an edge-edge test (e0, e1) can be reached from
// several face pairs of the list (every pair of faces holding its primitives).
// Its filters (no shared node; swept node/edge box overlapping the dilated
// face/edge box) and its FP64 evaluation depend on the primitives only, so all
// those evaluations are bit-identical.  The filter stage keeps each test once:
// from the first pair, in ascending face order, that the list holds.  Whether
// the list holds a face pair is decided from the pair predicate itself (the
// query face's box against the target face's box dilated by `thickness`, the
// query being the lower sorted leaf of a mesh against itself, the cloth face
// against an obstacle, the later cloth against an earlier one), so the kept set
// is exactly the set of distinct tests the list reaches.

struct FilterArgs {
    const int2* pairs;
    long nPairs;
    float th;
    int keyBits;
    int nCloths;
    const float4 *x, *x0;
    const int4 *fnodes, *fedges;
    const int2 *enodes, *eadj;
    const int *nfPtr, *nfIdx;
    const float4* fbox;  // per face: (leaf box lo, mesh bits), (leaf box hi, sorted index bits)
    uint16_t* mask;      // per pair: survivors (bit t for test t)
    uint32_t *blkVF, *blkEE;  // per 256-pair block: survivors of each kind
};

__device__ __forceinline__ void swept_box(Box& b, const float4* x, const float4* x0, int n) {
    box_add(b, ld3(x[n]));
    box_add(b, ld3(x0[n]));
}
__device__ __forceinline__ Box node_sbox(const FilterArgs& A, int n) {
    Box b = box_empty();
    swept_box(b, A.x, A.x0, n);
    return b;
}
// A face's leaf box (the swept box of its three nodes), mesh and sorted leaf index.
struct FaceRec {
    Box b;
    int mesh, rank;
};
__device__ __forceinline__ FaceRec face_rec(const FilterArgs& A, int f) {
    const float4 lo = A.fbox[2 * (size_t)f], hi = A.fbox[2 * (size_t)f + 1];
    FaceRec r;
    r.b.lo[0] = lo.x; r.b.lo[1] = lo.y; r.b.lo[2] = lo.z;
    r.b.hi[0] = hi.x; r.b.hi[1] = hi.y; r.b.hi[2] = hi.z;
    r.mesh = __float_as_int(lo.w);
    r.rank = __float_as_int(hi.w);
    return r;
}

// Does the candidate list hold the face pair {g, f} (g != f)?
__device__ bool in_list(const FilterArgs& A, const FaceRec& g, const FaceRec& f) {
    const bool og = g.mesh >= A.nCloths, of = f.mesh >= A.nCloths;
    if (og && of) return false;  // obstacles are never traversed against each other
    bool gQuery;
    if (g.mesh == f.mesh) gQuery = g.rank < f.rank;
    else if (og != of) gQuery = !og;  // the cloth face queries the obstacle's tree
    else gQuery = g.mesh > f.mesh;    // the later cloth queries the earlier one
    return gQuery ? box_overlap(g.b, f.b, A.th) : box_overlap(f.b, g.b, A.th);
}

// vertex-face test (n, f) reached from the pair (p, f), n in p
__device__ bool vf_owner(const FilterArgs& A, int n, int p, const FaceRec& f) {
    for (int k = A.nfPtr[n], e = A.nfPtr[n + 1]; k < e; k++) {
        const int g = A.nfIdx[k] >> 2;  // ascending
        if (g == p) return true;
        if (in_list(A, face_rec(A, g), f)) return false;
    }
    return true;
}

// edge-edge test (e0, e1), e0 < e1, reached from the pair (p0, p1), e0 in p0, e1 in p1
__device__ bool ee_owner(const FilterArgs& A, int e0, int e1, int p0, int p1) {
    const int2 a = A.eadj[e0], b = A.eadj[e1];
    int g0[2], g1[2], n0 = 0, n1 = 0;
    if (a.x >= 0) g0[n0++] = a.x;
    if (a.y >= 0) g0[n0++] = a.y;
    if (b.x >= 0) g1[n1++] = b.x;
    if (b.y >= 0) g1[n1++] = b.y;
    if (n0 == 2 && g0[1] < g0[0]) { int t = g0[0]; g0[0] = g0[1]; g0[1] = t; }
    if (n1 == 2 && g1[1] < g1[0]) { int t = g1[0]; g1[0] = g1[1]; g1[1] = t; }
    for (int i = 0; i < n0; i++) {
        if (g0[i] == p0 && g1[0] == p1) return true;
        const FaceRec r0 = face_rec(A, g0[i]);
        for (int j = 0; j < n1; j++) {
            if (g0[i] == p0 && g1[j] == p1) return true;
            if (g0[i] == g1[j]) continue;
            if (in_list(A, r0, face_rec(A, g1[j]))) return false;
        }
    }
    return true;
}

__device__ __forceinline__ int pick3(int4 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

__device__ __forceinline__ Box box_union(const Box& a, const Box& b) {
    Box c;
    for (int k = 0; k < 3; k++) { c.lo[k] = tmin(a.lo[k], b.lo[k]); c.hi[k] = tmax(a.hi[k], b.hi[k]); }
    return c;
}
// the swept box of node id, one of the face's nodes fn whose swept boxes are b[0..2]
__device__ __forceinline__ Box pick_box(int id, int4 fn, const Box b[3]) {
    return id == fn.x ? b[0] : (id == fn.y ? b[1] : b[2]);
}

// Pass 1, one thread per face pair: the 15 tests (t < 3: node t of fa against
// fb; t < 6: node t - 3 of fb against fa; else edge (t - 6) / 3 of fa against
// edge (t - 6) % 3 of fb) through the filters and ownership, with the six
// nodes' swept boxes loaded once.  Writes the pair's survivor mask and each
// block's survivor counts; no atomics.
__global__ void __launch_bounds__(256) k_ccd_filter(FilterArgs A) {
    __shared__ uint32_t red[2][4];
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    unsigned m = 0;
    if (p < A.nPairs) {
        const int2 pr = A.pairs[p];
        const int fa = pr.x, fb = pr.y;
        const int4 na = A.fnodes[fa], nb = A.fnodes[fb];
        const FaceRec ra = face_rec(A, fa), rb = face_rec(A, fb);
        Box ba[3], bb[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            ba[k] = node_sbox(A, pick3(na, k));
            bb[k] = node_sbox(A, pick3(nb, k));
        }
#pragma unroll
        for (int t = 0; t < 6; t++) {
            const bool fromA = t < 3;
            const int n = fromA ? pick3(na, t) : pick3(nb, t - 3);
            const int4 ff = fromA ? nb : na;
            if (n == ff.x || n == ff.y || n == ff.z) continue;
            const FaceRec& rf = fromA ? rb : ra;
            if (!box_overlap(fromA ? ba[t] : bb[t - 3], rf.b, A.th)) continue;
            if (vf_owner(A, n, fromA ? fa : fb, rf)) m |= 1u << t;
        }
        const int4 ea = A.fedges[fa], eb = A.fedges[fb];
#pragma unroll
        for (int t = 6; t < 15; t++) {
            int e0 = pick3(ea, (t - 6) / 3), e1 = pick3(eb, (t - 6) % 3);
            const int2 a = A.enodes[e0], b = A.enodes[e1];
            if (a.x == b.x || a.x == b.y || a.y == b.x || a.y == b.y) continue;
            Box b0 = box_union(pick_box(a.x, na, ba), pick_box(a.y, na, ba));
            Box b1 = box_union(pick_box(b.x, nb, bb), pick_box(b.y, nb, bb));
            int p0 = fa, p1 = fb;
            if (e1 < e0) {  // the lower edge id first
                int sw = e0; e0 = e1; e1 = sw;
                sw = p0; p0 = p1; p1 = sw;
                const Box tb = b0; b0 = b1; b1 = tb;
            }
            if (!box_overlap(b0, b1, A.th)) continue;
            if (ee_owner(A, e0, e1, p0, p1)) m |= 1u << t;
        }
        A.mask[p] = (uint16_t)m;
    }
    unsigned cv = __popc(m & 63u), ce = __popc(m >> 6);
    for (int o = 32; o > 0; o >>= 1) {
        cv += __shfl_xor(cv, o, 64);
        ce += __shfl_xor(ce, o, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][wave] = cv; red[1][wave] = ce; }
    __syncthreads();
    if (threadIdx.x == 0) {
        A.blkVF[blockIdx.x] = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        A.blkEE[blockIdx.x] = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    }
}

struct EmitArgs {
    const int2* pairs;
    long nPairs;
    int keyBits;
    const int4 *fnodes, *fedges;
    const uint16_t* mask;
    const uint32_t *offVF, *offEE;  // exclusive scans of the block counts
    uint64_t *candVF, *candEE;
};

// inclusive scan of v over the 64 lanes of the wave
__device__ __forceinline__ unsigned wave_incl_scan(unsigned v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// Pass 2: the survivors' keys at their positions in pair order.
__global__ void __launch_bounds__(256) k_ccd_emit(EmitArgs A) {
    __shared__ uint32_t wsum[2][4];
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    const unsigned m = p < A.nPairs ? A.mask[p] : 0u;
    const unsigned cv = __popc(m & 63u), ce = __popc(m >> 6);
    const unsigned iv = wave_incl_scan(cv), ie = wave_incl_scan(ce);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) { wsum[0][wave] = iv; wsum[1][wave] = ie; }
    __syncthreads();
    if (m == 0) return;
    unsigned ov = A.offVF[blockIdx.x] + iv - cv, oe = A.offEE[blockIdx.x] + ie - ce;
    for (int w = 0; w < wave; w++) { ov += wsum[0][w]; oe += wsum[1][w]; }
    const int2 pr = A.pairs[p];
    const int fa = pr.x, fb = pr.y;
    const int4 na = A.fnodes[fa], nb = A.fnodes[fb];
#pragma unroll
    for (int t = 0; t < 6; t++)
        if (m >> t & 1u) {
            const int n = t < 3 ? pick3(na, t) : pick3(nb, t - 3);
            const int f = t < 3 ? fb : fa;
            A.candVF[ov++] = ((uint64_t)n << A.keyBits) | (uint64_t)f;
        }
    if (m >> 6) {
        const int4 ea = A.fedges[fa], eb = A.fedges[fb];
#pragma unroll
        for (int t = 6; t < 15; t++)
            if (m >> t & 1u) {
                int e0 = pick3(ea, (t - 6) / 3), e1 = pick3(eb, (t - 6) % 3);
                if (e1 < e0) { const int sw = e0; e0 = e1; e1 = sw; }
                A.candEE[oe++] = (1ull << (2 * A.keyBits)) | ((uint64_t)e0 << A.keyBits) | (uint64_t)e1;
            }
    }
}

// ---- stage 2: the FP64 test of every distinct candidate ----

struct CheckArgs {
    const uint64_t* cand;
    long n;
    int keyBits;
    const float4 *x, *x0;
    const int4* fnodes;
    const int2* enodes;
    cs_impact_rec* out;
    uint64_t* keys;
    unsigned* count;
    long cap;
};

// checkVertexFaceImpact / checkEdgeEdgeImpact (src/CollisionHelper.cu:111-135)
// after their filters: the cubic, its roots and the proximity test at each root.
template <int TYPE>
__global__ void __launch_bounds__(256) k_ccd_check(CheckArgs A) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= A.n) return;
    const uint64_t key = A.cand[i];
    const uint64_t mask = (1ull << A.keyBits) - 1;
    const int a = (int)((key >> A.keyBits) & mask), b = (int)(key & mask);
    cs_impact_rec imp;
    imp.type = TYPE;
    if (TYPE == 0) {
        const int4 fn = A.fnodes[b];
        imp.nodes[0] = a; imp.nodes[1] = fn.x; imp.nodes[2] = fn.y; imp.nodes[3] = fn.z;
    } else {
        const int2 ea = A.enodes[a], eb = A.enodes[b];
        imp.nodes[0] = ea.x; imp.nodes[1] = ea.y; imp.nodes[2] = eb.x; imp.nodes[3] = eb.y;
    }
    V3d X0[4], X1[4];
    for (int k = 0; k < 4; k++) {
        X0[k] = tod(ld3(A.x0[imp.nodes[k]]));
        X1[k] = tod(ld3(A.x[imp.nodes[k]]));
    }
    if (!check_impact(TYPE, X0, X1, imp)) return;
    const unsigned slot = atomicAdd(A.count, 1u);
    if ((long)slot < A.cap) {
        A.out[slot] = imp;
        A.keys[slot] = key;
    }
}

__global__ void k_unique_heads(int n, const uint64_t* __restrict__ keys, uint32_t* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// unique impacts, keyed by their t bits (t >= 0, so the bit order is the value order)
__global__ void k_unique_compact(int n, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ pos,
                                 const uint32_t* __restrict__ idx, const cs_impact_rec* __restrict__ imps,
                                 uint64_t* __restrict__ tkeys, uint32_t* __restrict__ tidx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    uint32_t u = pos[i];
    uint32_t j = idx[i];
    tkeys[u] = (uint64_t)__double_as_longlong(imps[j].t);
    tidx[u] = j;
}

__global__ void k_gather_impacts(int n, const uint32_t* __restrict__ idx, const cs_impact_rec* __restrict__ in,
                                 cs_impact_rec* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = in[idx[i]];
}

// One workgroup: rounds of "minimum sorted index on every eligible node".
// state: 0 live candidate, 1 accepted, 2 excluded.
constexpr int IB = 1024;
__global__ void __launch_bounds__(IB) k_independent(int n, const cs_impact_rec* __restrict__ imps, int deform, int NC,
                                                    int* nodeMin, int* nodeRemoved, int* state,
                                                    cs_impact_rec* __restrict__ out, int* nOut) {
    __shared__ int live;
    __shared__ int wsum[IB / 64];
    auto eligible = [&](int nd) { return deform == 1 || nd < NC; };
    for (int i = threadIdx.x; i < n; i += IB) {
        state[i] = 0;
        for (int k = 0; k < 4; k++) {
            int nd = imps[i].nodes[k];
            if (eligible(nd)) { __hip_atomic_store(&nodeRemoved[nd], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
        }
    }
    __syncthreads();
    for (;;) {
        for (int i = threadIdx.x; i < n; i += IB)
            if (state[i] == 0)
                for (int k = 0; k < 4; k++) {
                    int nd = imps[i].nodes[k];
                    if (eligible(nd)) __hip_atomic_store(&nodeMin[nd], 0x7fffffff, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += IB)
            if (state[i] == 0)
                for (int k = 0; k < 4; k++) {
                    int nd = imps[i].nodes[k];
                    if (eligible(nd)) atomicMin(&nodeMin[nd], i);
                }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += IB)
            if (state[i] == 0) {
                bool ok = true;
                for (int k = 0; k < 4; k++) {
                    int nd = imps[i].nodes[k];
                    if (eligible(nd) && __hip_atomic_load(&nodeMin[nd], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != i)
                        ok = false;
                }
                if (ok) state[i] = 3;  // accepted this round
            }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += IB)
            if (state[i] == 3) {
                state[i] = 1;
                for (int k = 0; k < 4; k++) {
                    int nd = imps[i].nodes[k];
                    if (eligible(nd)) __hip_atomic_store(&nodeRemoved[nd], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        if (threadIdx.x == 0) live = 0;
        __syncthreads();
        int mine = 0;
        for (int i = threadIdx.x; i < n; i += IB)
            if (state[i] == 0) {
                bool rem = false;
                for (int k = 0; k < 4; k++) {
                    int nd = imps[i].nodes[k];
                    if (eligible(nd) && __hip_atomic_load(&nodeRemoved[nd], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                        rem = true;
                }
                if (rem) state[i] = 2; else mine++;
            }
        if (mine) atomicAdd(&live, mine);
        __syncthreads();
        if (live == 0) break;
        __syncthreads();
    }
    // ordered compaction of the accepted impacts
    int carry = 0;
    for (int base = 0; base < n; base += IB) {
        int i = base + threadIdx.x;
        int f = (i < n && state[i] == 1) ? 1 : 0;
        int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        uint64_t bal = __ballot(f);
        int before = __popcll(bal & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
        if (lane == 0) wsum[wave] = __popcll(bal);
        __syncthreads();
        int off = carry;
        for (int w = 0; w < wave; w++) off += wsum[w];
        int tot = 0;
        for (int w = 0; w < IB / 64; w++) tot += wsum[w];
        if (f) out[off + before] = imps[i];
        __syncthreads();
        carry += tot;
    }
    if (threadIdx.x == 0) *nOut = carry;
}

}  // namespace

int find_impacts(cs_ctx* C, float th) {
    hipStream_t s = C->stream;
    long nPairs = bvh_all_pairs(C, th);
    if (nPairs == 0) return 0;
    int keyBits = 1;
    long mx = std::max(std::max((long)C->N, (long)C->F), (long)C->E);
    while ((1L << keyBits) <= mx) keyBits++;
    unsigned* count = C->impCount.as<unsigned>(4);
    // stage 1: distinct candidate tests (vertex-face and edge-edge lists), in pair order
    const int nblk = (int)cdiv(nPairs, 256);
    uint16_t* mask = C->ccdMask.as<uint16_t>(nPairs);
    uint32_t* blk = C->ccdBlk.as<uint32_t>(4L * (nblk + 1));
    uint32_t *blkVF = blk, *blkEE = blk + (nblk + 1), *offVF = blk + 2L * (nblk + 1), *offEE = blk + 3L * (nblk + 1);
    {
        FilterArgs A;
        A.pairs = C->pairs.ptr<int2>();
        A.nPairs = nPairs;
        A.th = th;
        A.keyBits = keyBits;
        A.nCloths = C->nCloths;
        A.x = C->x.ptr<float4>(); A.x0 = C->x0.ptr<float4>();
        A.fnodes = C->fnodes.ptr<int4>(); A.fedges = C->fedges.ptr<int4>();
        A.enodes = C->enodes.ptr<int2>(); A.eadj = C->eadj.ptr<int2>();
        A.nfPtr = C->nfPtr.ptr<int>(); A.nfIdx = C->nfIdx.ptr<int>(); A.fbox = C->faceBox.ptr<float4>();
        A.mask = mask;
        A.blkVF = blkVF;
        A.blkEE = blkEE;
        hipLaunchKernelGGL(k_ccd_filter, dim3(nblk), dim3(256), 0, s, A);
        CS_LAUNCH_CHECK();
    }
    scan_exclusive(blkVF, offVF, nblk, C->ws, s);
    scan_exclusive(blkEE, offEE, nblk, C->ws, s);
    CS_HIP(hipMemcpyAsync(C->hCount, offVF + nblk, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    CS_HIP(hipMemcpyAsync(C->hCount + 1, offEE + nblk, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    CS_HIP(hipStreamSynchronize(s));
    const long nVF = C->hCount[0], nEE = C->hCount[1];
    const long ccap = std::max(1L, std::max(nVF, nEE));
    if (nVF + nEE > 0) {
        uint64_t* cand = C->ccdCand.as<uint64_t>(2 * ccap);
        EmitArgs E;
        E.pairs = C->pairs.ptr<int2>();
        E.nPairs = nPairs;
        E.keyBits = keyBits;
        E.fnodes = C->fnodes.ptr<int4>(); E.fedges = C->fedges.ptr<int4>();
        E.mask = mask;
        E.offVF = offVF;
        E.offEE = offEE;
        E.candVF = cand;
        E.candEE = cand + ccap;
        hipLaunchKernelGGL(k_ccd_emit, dim3(nblk), dim3(256), 0, s, E);
        CS_LAUNCH_CHECK();
    }
    // stage 2: FP64 tests; impacts are appended with their keys
    long cap = std::max<long>(4096, (nVF + nEE) / 8);
    for (int attempt = 0; attempt < 2; attempt++) {
        cs_impact_rec* imps = C->impacts.as<cs_impact_rec>(cap);
        uint64_t* keys = C->impKeys.as<uint64_t>(cap);
        fill_u32(count + 2, 0u, 1, s);
        CheckArgs A;
        A.keyBits = keyBits;
        A.x = C->x.ptr<float4>(); A.x0 = C->x0.ptr<float4>();
        A.fnodes = C->fnodes.ptr<int4>(); A.enodes = C->enodes.ptr<int2>();
        A.out = imps; A.keys = keys; A.count = count + 2; A.cap = cap;
        const uint64_t* cand = C->ccdCand.ptr<uint64_t>();
        if (nVF > 0) {
            A.cand = cand;
            A.n = nVF;
            hipLaunchKernelGGL(k_ccd_check<0>, dim3(cdiv(nVF, 256)), dim3(256), 0, s, A);
            CS_LAUNCH_CHECK();
        }
        if (nEE > 0) {
            A.cand = cand + ccap;
            A.n = nEE;
            hipLaunchKernelGGL(k_ccd_check<1>, dim3(cdiv(nEE, 256)), dim3(256), 0, s, A);
            CS_LAUNCH_CHECK();
        }
        CS_HIP(hipMemcpyAsync(C->hCount, count + 2, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        if ((long)C->hCount[0] <= cap) break;
        cap = (long)C->hCount[0] + 1024;
    }
    const int n = (int)C->hCount[0];
    if (n == 0) return 0;
    cs_impact_rec* imps = C->impacts.ptr<cs_impact_rec>();
    uint64_t* keys = C->impKeys.ptr<uint64_t>();
    uint64_t* ktmp = C->impKeysTmp.as<uint64_t>(n);
    uint32_t* idx = C->impIdx.as<uint32_t>(n);
    uint32_t* itmp = C->impIdxTmp.as<uint32_t>(n);
    iota_u32(idx, n, s);
    radix_sort_pairs(keys, idx, ktmp, itmp, n, 0, 2 * keyBits + 1, C->ws, s);
    uint32_t* flag = C->ws.a.as<uint32_t>(n + 1);
    uint32_t* pos = C->ws.b.as<uint32_t>(n + 1);
    hipLaunchKernelGGL(k_unique_heads, dim3(cdiv(n, 256)), dim3(256), 0, s, n, keys, flag);
    CS_LAUNCH_CHECK();
    scan_exclusive(flag, pos, n, C->ws, s);
    CS_HIP(hipMemcpyAsync(C->hCount, pos + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    CS_HIP(hipStreamSynchronize(s));
    const int nu = (int)C->hCount[0];
    // ktmp/itmp are free after the sort: reuse them for the t keys
    hipLaunchKernelGGL(k_unique_compact, dim3(cdiv(n, 256)), dim3(256), 0, s, n, flag, pos, idx, imps, ktmp, itmp);
    CS_LAUNCH_CHECK();
    radix_sort_pairs(ktmp, itmp, keys, idx, nu, 0, 62, C->ws, s);
    cs_impact_rec* sorted = C->impSorted.as<cs_impact_rec>(nu);
    hipLaunchKernelGGL(k_gather_impacts, dim3(cdiv(nu, 256)), dim3(256), 0, s, nu, itmp, imps, sorted);
    CS_LAUNCH_CHECK();
    return nu;
}

int independent_impacts(cs_ctx* C, const cs_impact_rec* dIn, int n, int deform, cs_impact_rec* dOut) {
    if (n <= 0) return 0;
    hipStream_t s = C->stream;
    int* nodeMin = C->nodeMin.as<int>(C->N);
    int* removed = C->nodeRemoved.as<int>(C->N);
    int* state = C->impState.as<int>(n + 1);
    hipLaunchKernelGGL(k_independent, dim3(1), dim3(IB), 0, s, n, dIn, deform, C->NC, nodeMin, removed, state, dOut,
                       state + n);
    CS_LAUNCH_CHECK();
    CS_HIP(hipMemcpyAsync(C->hCount, state + n, sizeof(int), hipMemcpyDeviceToHost, s));
    CS_HIP(hipStreamSynchronize(s));
    return (int)C->hCount[0];
}

}  // namespace cs
