// cs_lpc.hip — learned sparse-approximate-inverse preconditioner on gfx950.
//
// The reference preconditions its CG with the inverse diagonal
// (src/PhysicsHelper.cu:494-499, used by src/Cloth.cu:766-809).  Here a
// mesh message-passing network reads the cloth's assembled system and emits a
// block lower-triangular factor G on the static pattern; the preconditioner is
// M^-1 = G^T G (symmetric positive definite whenever G's diagonal blocks are
// nonsingular, which their exp-parametrised diagonals guarantee).
//
// Network (H = 64 features, 3 message-passing layers):
//   S_i = chol(Minv_i)^T, Ahat_ij = S_i A_ij S_j^T (so Ahat_ii = I)
//   node input xn = [sum_j |Ahat_ij|^2, degree/8, sum_contact |S_i Ac_ij S_j^T|^2, 1, 0..] (16)
//   edge input xe = [Ahat_ij (9), |Ahat_ij|, 1, 0..] (16) per directed edge i -> j
//   h = relu(xn Wn + bn), e = relu(xe We + be)
//   layer: P = h Wa, Q = h Wb, z = relu(P_i + Q_j + e Wc + bc), m = z W2 + b2,
//          e += m, h += relu([h, sum_j m] U1 + bu1) U2 + bu2
//   o = relu(h Dn1 + bdn1) Dn2 + bdn2 -> Ghat_ii = [[e^o0,0,0],[o3,e^o1,0],[o4,o5,e^o2]]
//   q = relu(e De1 + bde1) De2 + bde2 (lower edges) -> Ghat_ij = q[0..8] (3x3 row-major)
//   G_ij = Ghat_ij S_j.  Ghat = I is exactly block Jacobi.
// Training minimises Kaporin's condition number of Ghat Ahat Ghat^T:
//   loss = log(trace(Ghat Ahat Ghat^T) / n) - (2/n) sum_i (o_i0 + o_i1 + o_i2).
//
// Every matrix product runs on bf16 MFMA (v_mfma_f32_16x16x16_bf16) with fp32
// accumulation.  A wave owns a tile of 16 rows (edges or nodes) x 64 features
// held transposed: lane (g = lane/16, n = lane%16) keeps row n, features
// 16 mb + 4 g + i (mb, i = 0..3).  That is both the MFMA result layout of
// Y^T = W^T X^T and the B-operand layout of the next product, so chained
// layers never shuffle.  64-feature rows are stored permuted (position
// 16 g + 4 mb + i) so each lane loads its 16 values contiguously.
//
// Edges are the slots of the static sliced ELL: a workgroup is one 64-row
// slice (4 waves x 16 rows) and walks the slice's slot columns, so the sum
// over a row's messages stays in registers and the node update runs in the
// same kernel.  Non-edge slots (diagonal, padding) carry zero messages.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

constexpr int HF = 64;  // hidden features
constexpr int NLAYER = 3;
constexpr int LB = 256;  // workgroup of the slice kernels: 4 waves x 16 rows

// flat weight vector (float offsets); matrices are [in][out] row-major
constexpr int O_WN = 0, O_BN = 1024, O_WE = 1088, O_BE = 2112;
constexpr int O_L0 = 2176, L_SIZE = 28928;
constexpr int L_WA = 0, L_WB = 4096, L_WC = 8192, L_BC = 12288, L_W2 = 12352, L_B2 = 16448, L_U1 = 16512,
              L_BU1 = 24704, L_U2 = 24768, L_BU2 = 28864;
constexpr int O_DEC = O_L0 + NLAYER * L_SIZE;
constexpr int O_DN1 = O_DEC + 0, O_BDN1 = O_DEC + 4096, O_DN2 = O_DEC + 4160, O_BDN2 = O_DEC + 5184,
              O_DE1 = O_DEC + 5200, O_BDE1 = O_DEC + 9296, O_DE2 = O_DEC + 9360, O_BDE2 = O_DEC + 10384;
constexpr int W_COUNT = O_DEC + 10400;
static_assert(W_COUNT == 99360, "weight layout");
__host__ __device__ constexpr int lw(int l, int o) { return O_L0 + l * L_SIZE + o; }

struct MatDesc {
    int off, K, N;
};
// every matrix of the network (fragments are built for these)
const MatDesc kMats[] = {{O_WN, 16, 64}, {O_WE, 16, 64},
                         {lw(0, L_WA), 64, 64}, {lw(0, L_WB), 64, 64}, {lw(0, L_WC), 64, 64}, {lw(0, L_W2), 64, 64},
                         {lw(0, L_U1), 128, 64}, {lw(0, L_U2), 64, 64},
                         {lw(1, L_WA), 64, 64}, {lw(1, L_WB), 64, 64}, {lw(1, L_WC), 64, 64}, {lw(1, L_W2), 64, 64},
                         {lw(1, L_U1), 128, 64}, {lw(1, L_U2), 64, 64},
                         {lw(2, L_WA), 64, 64}, {lw(2, L_WB), 64, 64}, {lw(2, L_WC), 64, 64}, {lw(2, L_W2), 64, 64},
                         {lw(2, L_U1), 128, 64}, {lw(2, L_U2), 64, 64},
                         {O_DN1, 64, 64}, {O_DN2, 64, 16}, {O_DE1, 64, 64}, {O_DE2, 64, 16}};

// ---------------------------------------------------------------- bf16 tiles

__device__ __forceinline__ bf16_t bf_rne(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf_val(bf16_t b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float((uint32_t)b << 16);
#else
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
#endif
}
__device__ __forceinline__ s4 to_bf(f4v a) {
    s4 r;
    r.x = (short)bf_rne(a.x);
    r.y = (short)bf_rne(a.y);
    r.z = (short)bf_rne(a.z);
    r.w = (short)bf_rne(a.w);
    return r;
}
__device__ __forceinline__ f4v from_bf(s4 a) {
    f4v r;
    r.x = bf_val((bf16_t)a.x);
    r.y = bf_val((bf16_t)a.y);
    r.z = bf_val((bf16_t)a.z);
    r.w = bf_val((bf16_t)a.w);
    return r;
}

struct Tile {
    f4v v[4];
};
struct BTile {
    s4 v[4];
};

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ f4v zero4() { return f4v{0.0f, 0.0f, 0.0f, 0.0f}; }
__device__ __forceinline__ f4v relu4(f4v a) {
    return f4v{fmaxf(a.x, 0.0f), fmaxf(a.y, 0.0f), fmaxf(a.z, 0.0f), fmaxf(a.w, 0.0f)};
}
__device__ __forceinline__ f4v mask4(f4v a, s4 ref) {  // a where the stored activation is positive
    const f4v r = from_bf(ref);
    return f4v{r.x > 0.0f ? a.x : 0.0f, r.y > 0.0f ? a.y : 0.0f, r.z > 0.0f ? a.z : 0.0f, r.w > 0.0f ? a.w : 0.0f};
}
__device__ __forceinline__ BTile to_bf(const Tile& t) {
    BTile b;
#pragma unroll
    for (int m = 0; m < 4; m++) b.v[m] = to_bf(t.v[m]);
    return b;
}
__device__ __forceinline__ Tile from_bf(const BTile& b) {
    Tile t;
#pragma unroll
    for (int m = 0; m < 4; m++) t.v[m] = from_bf(b.v[m]);
    return t;
}
// bias of the tile's features: lane (g, n) holds features 16 mb + 4 g + i
__device__ __forceinline__ Tile bias_tile(const float* __restrict__ b) {
    const int g = lane_id() >> 4;
    Tile t;
#pragma unroll
    for (int m = 0; m < 4; m++) t.v[m] = *reinterpret_cast<const f4v*>(b + 16 * m + 4 * g);
    return t;
}
__device__ __forceinline__ f4v bias16(const float* __restrict__ b) {
    return *reinterpret_cast<const f4v*>(b + 4 * (lane_id() >> 4));
}

// 64-feature rows, permuted storage (position 16 g + 4 mb + i)
__device__ __forceinline__ Tile ld_tile(const float* __restrict__ p, long row) {
    const f4v* q = reinterpret_cast<const f4v*>(p + row * HF + 16 * (lane_id() >> 4));
    Tile t;
#pragma unroll
    for (int m = 0; m < 4; m++) t.v[m] = q[m];
    return t;
}
__device__ __forceinline__ void st_tile(float* __restrict__ p, long row, const Tile& t) {
    f4v* q = reinterpret_cast<f4v*>(p + row * HF + 16 * (lane_id() >> 4));
#pragma unroll
    for (int m = 0; m < 4; m++) q[m] = t.v[m];
}
__device__ __forceinline__ BTile ld_btile(const bf16_t* __restrict__ p, long row) {
    const s4* q = reinterpret_cast<const s4*>(p + row * HF + 16 * (lane_id() >> 4));
    BTile t;
#pragma unroll
    for (int m = 0; m < 4; m++) t.v[m] = q[m];
    return t;
}
__device__ __forceinline__ void st_btile(bf16_t* __restrict__ p, long row, const BTile& t) {
    s4* q = reinterpret_cast<s4*>(p + row * HF + 16 * (lane_id() >> 4));
#pragma unroll
    for (int m = 0; m < 4; m++) q[m] = t.v[m];
}
// 16-feature rows, natural order: lane (g, n) holds features 4 g .. 4 g + 3
__device__ __forceinline__ void st16(bf16_t* __restrict__ p, long row, s4 v) {
    *reinterpret_cast<s4*>(p + row * 16 + 4 * (lane_id() >> 4)) = v;
}

#define MFMA(a, b, c) __builtin_amdgcn_mfma_f32_16x16x16bf16_1k((a), (b), (c), 0, 0, 0)

// y += X W for X given as KS k-steps (bf16 tiles of 16 features); W has 64 outputs.
// wf: forward fragments of W, [mb][ks][lane].
template <int KS>
__device__ __forceinline__ void mm64(Tile& y, const s4 (&x)[KS], const s4* wf) {
    const int lane = lane_id();
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int k = 0; k < KS; k++) y.v[m] = MFMA(wf[(m * KS + k) * 64 + lane], x[k], y.v[m]);
}
// y += X W for W with 16 outputs
template <int KS>
__device__ __forceinline__ f4v mm16(f4v y, const s4 (&x)[KS], const s4* wf) {
    const int lane = lane_id();
#pragma unroll
    for (int k = 0; k < KS; k++) y = MFMA(wf[k * 64 + lane], x[k], y);
    return y;
}
// dx (KB blocks of 16 input features) += dY W^T for dY given as MS blocks; wt: transposed
// fragments of W, [kb][ms][lane].
template <int KB, int MS>
__device__ __forceinline__ void mmT(f4v (&dx)[KB], const s4 (&dy)[MS], const s4* wt) {
    const int lane = lane_id();
#pragma unroll
    for (int k = 0; k < KB; k++)
#pragma unroll
        for (int m = 0; m < MS; m++) dx[k] = MFMA(wt[(k * MS + m) * 64 + lane], dy[m], dx[k]);
}

__device__ __forceinline__ long vidx(long slot, int c) { return (slot >> 6) * (9 * 64) + c * 64 + (slot & 63); }

// ---------------------------------------------------------------- set-up kernels

// Forward and transposed MFMA fragments of one K x N matrix (bf16, round to nearest even).
__global__ void k_lpc_frag(const float* __restrict__ w, bf16_t* __restrict__ frag, bf16_t* __restrict__ fragT,
                           int off, int K, int N) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;  // one thread per 4-element fragment
    if (t >= K * N / 4) return;
    const int lane = t & 63, g = lane >> 4, n = lane & 15;
    const int KS = K / 16, MB = N / 16;
    {   // forward: index (mb KS + ks) 64 + lane -> W[16 ks + 4 g + j][16 mb + n]
        const int blk = t >> 6, mb = blk / KS, ks = blk % KS;
        for (int j = 0; j < 4; j++) frag[off + 4 * t + j] = bf_rne(w[off + (16 * ks + 4 * g + j) * N + 16 * mb + n]);
    }
    {   // transposed: index (kb MB + ms) 64 + lane -> W[16 kb + n][16 ms + 4 g + j]
        const int blk = t >> 6, kb = blk / MB, ms = blk % MB;
        for (int j = 0; j < 4; j++) fragT[off + 4 * t + j] = bf_rne(w[off + (16 * kb + n) * N + 16 * ms + 4 * g + j]);
    }
}

// S_i = chol(Minv_i)^T (lower triangle of Minv read)
__global__ void k_lpc_scale(int nn, const float* __restrict__ Minv, float* __restrict__ S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const float* m = Minv + 9 * (size_t)i;
    const float l00 = sqrtf(m[0]);
    const float l10 = m[3] / l00, l20 = m[6] / l00;
    const float l11 = sqrtf(m[4] - l10 * l10);
    const float l21 = (m[7] - l20 * l10) / l11;
    const float l22 = sqrtf(m[8] - l20 * l20 - l21 * l21);
    float* s = S + 9 * (size_t)i;
    s[0] = l00; s[1] = l10; s[2] = l20;
    s[3] = 0.0f; s[4] = l11; s[5] = l21;
    s[6] = 0.0f; s[7] = 0.0f; s[8] = l22;
}

// ---------------------------------------------------------------- forward

// Activations kept for the backward pass (training only; null in inference).
struct Stash {
    bf16_t* xn;        // 16 per node row
    bf16_t* xe;        // 16 per slot
    float* Ah;         // 9 per slot (Ahat of edge slots)
    bf16_t* e[NLAYER + 1];   // edge states: inputs of the layers, then the final states
    bf16_t* z[NLAYER];       // post-relu edge hidden
    bf16_t* hb[NLAYER + 1];  // bf16(h) entering each layer, then the final h
    bf16_t* aggb[NLAYER];
    bf16_t* u[NLAYER];       // post-relu node hidden
    bf16_t* dnh;             // node decoder hidden (per node row)
    bf16_t* deh;             // edge decoder hidden (per slot)
    float* o;                // 16 per node row
    float* gh;               // Ghat, 9 per slot (lower edge slots and the diagonal slot)
};

struct Topo {
    int nn;
    const int* sliceOff;
    const int* sliceW;
    const int* cols;
    const int* rowLen;
    const int* diagSlot;
    const int* mirror;
};

struct EncArgs {
    Topo t;
    const float* vals;  // A, value layout
    const int* cRowPtr;
    const int* cCol;
    const float* cVal;
    const float* S;
    const float* w;
    const s4* wf;
    bf16_t* e;
    float *h, *P, *Q;
    Stash st;
};

template <bool TRAIN>
__global__ void __launch_bounds__(LB) k_lpc_encode(EncArgs a) {
    const int s = blockIdx.x;
    const int lane = lane_id(), g = lane >> 4, n = lane & 15;
    const int li = 16 * (threadIdx.x >> 6) + n;
    const int row = s * 64 + li;
    const bool live = row < a.t.nn;
    const int W = a.t.sliceW[s];
    const long off = a.t.sliceOff[s];
    const int rl = live ? a.t.rowLen[row] : 0;
    float Si[9];
#pragma unroll
    for (int c = 0; c < 9; c++) Si[c] = live ? a.S[9 * (size_t)row + c] : 0.0f;
    const s4* fWe = a.wf + O_WE / 4;
    float nf0 = 0.0f, deg = 0.0f, nf2 = 0.0f;
    for (int k = 0; k < W; k++) {
        const long sig = (off + k) * 64 + li;
        const int col = a.t.cols[sig];
        const bool edge = live && k < rl && col != row;
        float Ah[9];
#pragma unroll
        for (int c = 0; c < 9; c++) Ah[c] = 0.0f;
        float nrm2 = 0.0f;
        if (edge) {
            float A[9], Sj[9], T[9];
#pragma unroll
            for (int c = 0; c < 9; c++) A[c] = a.vals[vidx(sig, c)];
#pragma unroll
            for (int c = 0; c < 9; c++) Sj[c] = a.S[9 * (size_t)col + c];
            // T = Si A, Ah = T Sj^T
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int c = 0; c < 3; c++) T[3 * r + c] = Si[3 * r] * A[c] + Si[3 * r + 1] * A[3 + c] + Si[3 * r + 2] * A[6 + c];
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int c = 0; c < 3; c++)
                    Ah[3 * r + c] = T[3 * r] * Sj[3 * c] + T[3 * r + 1] * Sj[3 * c + 1] + T[3 * r + 2] * Sj[3 * c + 2];
#pragma unroll
            for (int c = 0; c < 9; c++) nrm2 += Ah[c] * Ah[c];
            nf0 += nrm2;
            deg += 1.0f;
        }
        f4v f;
        if (g == 0) f = f4v{Ah[0], Ah[1], Ah[2], Ah[3]};
        else if (g == 1) f = f4v{Ah[4], Ah[5], Ah[6], Ah[7]};
        else if (g == 2) f = f4v{Ah[8], sqrtf(nrm2), edge ? 1.0f : 0.0f, 0.0f};
        else f = zero4();
        const s4 xb[1] = {to_bf(f)};
        if (TRAIN) {
            st16(a.st.xe, sig, xb[0]);
            if (g == 0)
#pragma unroll
                for (int c = 0; c < 9; c++) a.st.Ah[sig * 9 + c] = Ah[c];
        }
        Tile e0 = bias_tile(a.w + O_BE);
        mm64<1>(e0, xb, fWe);
#pragma unroll
        for (int m = 0; m < 4; m++) e0.v[m] = relu4(e0.v[m]);
        st_btile(a.e, sig, to_bf(e0));
    }
    if (live) {
        for (int kc = a.cRowPtr[row], ce = a.cRowPtr[row + 1]; kc < ce; kc++) {
            const int cc = a.cCol[kc];
            if (cc == row) continue;
            float A[9], Sj[9], T[9];
#pragma unroll
            for (int c = 0; c < 9; c++) A[c] = a.cVal[(size_t)kc * 9 + c];
#pragma unroll
            for (int c = 0; c < 9; c++) Sj[c] = a.S[9 * (size_t)cc + c];
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int c = 0; c < 3; c++) T[3 * r + c] = Si[3 * r] * A[c] + Si[3 * r + 1] * A[3 + c] + Si[3 * r + 2] * A[6 + c];
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const float v = T[3 * r] * Sj[3 * c] + T[3 * r + 1] * Sj[3 * c + 1] + T[3 * r + 2] * Sj[3 * c + 2];
                    nf2 += v * v;
                }
        }
    }
    const f4v fn = (g == 0 && live) ? f4v{nf0, deg * 0.125f, nf2, 1.0f} : zero4();
    const s4 xn[1] = {to_bf(fn)};
    if (TRAIN) st16(a.st.xn, row, xn[0]);
    Tile h = bias_tile(a.w + O_BN);
    mm64<1>(h, xn, a.wf + O_WN / 4);
#pragma unroll
    for (int m = 0; m < 4; m++) h.v[m] = relu4(h.v[m]);
    st_tile(a.h, row, h);
    const BTile hb = to_bf(h);
    Tile P{}, Q{};
#pragma unroll
    for (int m = 0; m < 4; m++) { P.v[m] = zero4(); Q.v[m] = zero4(); }
    mm64<4>(P, hb.v, a.wf + lw(0, L_WA) / 4);
    mm64<4>(Q, hb.v, a.wf + lw(0, L_WB) / 4);
    st_tile(a.P, row, P);
    st_tile(a.Q, row, Q);
}

struct LayerArgs {
    Topo t;
    int layer;
    const float* S;
    const float *P, *Q;
    float *Pn, *Qn;
    float* h;
    const bf16_t* ein;
    bf16_t* eout;
    bf16_t* G;
    const float* w;
    const s4* wf;
    Stash st;
};

// Ghat (row-major 3x3) from q / o and G = Ghat S, written to slot `sig` (and its transpose to `msig`)
__device__ __forceinline__ void write_g(bf16_t* __restrict__ G, long sig, long msig, const float (&gh)[9],
                                        const float* __restrict__ Sj) {
    float r[9];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) r[3 * a + b] = gh[3 * a] * Sj[b] + gh[3 * a + 1] * Sj[3 + b] + gh[3 * a + 2] * Sj[6 + b];
#pragma unroll
    for (int c = 0; c < 9; c++) G[vidx(sig, c)] = bf_rne(r[c]);
    if (msig >= 0)
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) G[vidx(msig, 3 * a + b)] = bf_rne(r[3 * b + a]);
}

template <bool TRAIN, bool LAST>
__global__ void __launch_bounds__(LB) k_lpc_layer(LayerArgs a) {
    __shared__ s4 wl[2 * 16 * 64];     // Wc and W2 forward fragments
    __shared__ float xch[4][16][17];  // per-wave exchange for the decoders
    const int s = blockIdx.x;
    const int wv = threadIdx.x >> 6;
    const int lane = lane_id(), g = lane >> 4, n = lane & 15;
    const int li = 16 * wv + n;
    const int row = s * 64 + li;
    const bool live = row < a.t.nn;
    const int l = a.layer;
    {
        const s4* fc = a.wf + lw(l, L_WC) / 4;
        const s4* f2 = a.wf + lw(l, L_W2) / 4;
        for (int t = threadIdx.x; t < 1024; t += LB) {
            wl[t] = fc[t];
            wl[1024 + t] = f2[t];
        }
    }
    __syncthreads();
    const int W = a.t.sliceW[s];
    const long off = a.t.sliceOff[s];
    const int rl = live ? a.t.rowLen[row] : 0;
    const Tile P = ld_tile(a.P, row);
    const Tile bc = bias_tile(a.w + lw(l, L_BC));
    const Tile b2 = bias_tile(a.w + lw(l, L_B2));
    Tile agg;
#pragma unroll
    for (int m = 0; m < 4; m++) agg.v[m] = zero4();
    for (int k = 0; k < W; k++) {
        const long sig = (off + k) * 64 + li;
        const int col = a.t.cols[sig];
        const bool edge = live && k < rl && col != row;
        const Tile Q = ld_tile(a.Q, col);
        const BTile eb = ld_btile(a.ein, sig);
        Tile z;
#pragma unroll
        for (int m = 0; m < 4; m++) z.v[m] = P.v[m] + Q.v[m] + bc.v[m];
        mm64<4>(z, eb.v, wl);
#pragma unroll
        for (int m = 0; m < 4; m++) z.v[m] = relu4(z.v[m]);
        const BTile zb = to_bf(z);
        if (TRAIN) st_btile(a.st.z[l], sig, zb);
        Tile msg = b2;
        mm64<4>(msg, zb.v, wl + 1024);
        if (!edge)
#pragma unroll
            for (int m = 0; m < 4; m++) msg.v[m] = zero4();
        Tile eo;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            eo.v[m] = from_bf(eb.v[m]) + msg.v[m];
            agg.v[m] += msg.v[m];
        }
        const BTile eob = to_bf(eo);
        st_btile(a.eout, sig, eob);
        if (LAST) {
            Tile dh = bias_tile(a.w + O_BDE1);
            mm64<4>(dh, eob.v, a.wf + O_DE1 / 4);
#pragma unroll
            for (int m = 0; m < 4; m++) dh.v[m] = relu4(dh.v[m]);
            const BTile dhb = to_bf(dh);
            if (TRAIN) st_btile(a.st.deh, sig, dhb);
            const f4v q = mm16<4>(bias16(a.w + O_BDE2), dhb.v, a.wf + O_DE2 / 4);
#pragma unroll
            for (int i = 0; i < 4; i++) xch[wv][n][4 * g + i] = q[i];
            __syncthreads();
            if (g == 0 && edge && col < row) {
                float gh[9];
#pragma unroll
                for (int c = 0; c < 9; c++) gh[c] = xch[wv][n][c];
                write_g(a.G, sig, a.t.mirror[sig], gh, a.S + 9 * (size_t)col);
                if (TRAIN)
#pragma unroll
                    for (int c = 0; c < 9; c++) a.st.gh[sig * 9 + c] = gh[c];
            }
            __syncthreads();
        }
    }
    // node update
    const Tile h = ld_tile(a.h, row);
    const BTile hb = to_bf(h);
    const BTile ab = to_bf(agg);
    if (TRAIN) {
        st_btile(a.st.hb[l], row, hb);
        st_btile(a.st.aggb[l], row, ab);
    }
    const s4 x8[8] = {hb.v[0], hb.v[1], hb.v[2], hb.v[3], ab.v[0], ab.v[1], ab.v[2], ab.v[3]};
    Tile u = bias_tile(a.w + lw(l, L_BU1));
    mm64<8>(u, x8, a.wf + lw(l, L_U1) / 4);
#pragma unroll
    for (int m = 0; m < 4; m++) u.v[m] = relu4(u.v[m]);
    const BTile ub = to_bf(u);
    if (TRAIN) st_btile(a.st.u[l], row, ub);
    Tile hn = bias_tile(a.w + lw(l, L_BU2));
    mm64<4>(hn, ub.v, a.wf + lw(l, L_U2) / 4);
#pragma unroll
    for (int m = 0; m < 4; m++) hn.v[m] += h.v[m];
    st_tile(a.h, row, hn);
    const BTile hnb = to_bf(hn);
    if (!LAST) {
        Tile Pn, Qn;
#pragma unroll
        for (int m = 0; m < 4; m++) { Pn.v[m] = zero4(); Qn.v[m] = zero4(); }
        mm64<4>(Pn, hnb.v, a.wf + lw(l + 1, L_WA) / 4);
        mm64<4>(Qn, hnb.v, a.wf + lw(l + 1, L_WB) / 4);
        st_tile(a.Pn, row, Pn);
        st_tile(a.Qn, row, Qn);
    } else {
        if (TRAIN) st_btile(a.st.hb[NLAYER], row, hnb);
        Tile dn = bias_tile(a.w + O_BDN1);
        mm64<4>(dn, hnb.v, a.wf + O_DN1 / 4);
#pragma unroll
        for (int m = 0; m < 4; m++) dn.v[m] = relu4(dn.v[m]);
        const BTile dnb = to_bf(dn);
        if (TRAIN) st_btile(a.st.dnh, row, dnb);
        const f4v o = mm16<4>(bias16(a.w + O_BDN2), dnb.v, a.wf + O_DN2 / 4);
        if (TRAIN) *reinterpret_cast<f4v*>(a.st.o + (long)row * 16 + 4 * g) = o;
#pragma unroll
        for (int i = 0; i < 4; i++) xch[wv][n][4 * g + i] = o[i];
        __syncthreads();
        if (g == 0 && live) {
            const float* ov = xch[wv][n];
            const float gh[9] = {expf(ov[0]), 0.0f, 0.0f, ov[3], expf(ov[1]), 0.0f, ov[4], ov[5], expf(ov[2])};
            const long sig = (off + a.t.diagSlot[row]) * 64 + li;
            write_g(a.G, sig, -1, gh, a.S + 9 * (size_t)row);
            if (TRAIN)
#pragma unroll
                for (int c = 0; c < 9; c++) a.st.gh[sig * 9 + c] = gh[c];
        }
    }
}

// ---------------------------------------------------------------- PCG sweeps

// y = G r (lower slots and the diagonal)
__global__ void __launch_bounds__(256) k_lpc_lower(int nn, const int* __restrict__ sliceOff,
                                                   const int* __restrict__ cols, const int* __restrict__ diagSlot,
                                                   const bf16_t* __restrict__ G, const float* __restrict__ r,
                                                   float4* __restrict__ y) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nn; i += gridDim.x * 256) {
        const int s = i / 64, lane = i % 64;
        const size_t off = (size_t)sliceOff[s];
        const int kd = diagSlot[i];
        const int* cp = cols + off * 64 + lane;
        const bf16_t* gp = G + off * 9 * 64 + lane;
        float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
        for (int k = 0; k <= kd; k++) {
            const int j = cp[k * 64];
            const float r0 = r[3 * j], r1 = r[3 * j + 1], r2 = r[3 * j + 2];
            const bf16_t* a = gp + (size_t)k * 9 * 64;
            y0 += bf_val(a[0 * 64]) * r0 + bf_val(a[1 * 64]) * r1 + bf_val(a[2 * 64]) * r2;
            y1 += bf_val(a[3 * 64]) * r0 + bf_val(a[4 * 64]) * r1 + bf_val(a[5 * 64]) * r2;
            y2 += bf_val(a[6 * 64]) * r0 + bf_val(a[7 * 64]) * r1 + bf_val(a[8 * 64]) * r2;
        }
        y[i] = make_float4(y0, y1, y2, 0.0f);
    }
}

// z = G^T y (diagonal transposed, upper slots hold the transposed lower blocks), then by MODE:
//   UP_PCG   z -> zout (float4), one partial of r.z per block
//   UP_POW   z -> zout (float4), one partial of |z|^2 per block
//   UP_PRE   w z -> xout (3 floats per node); w = *omega
//   UP_POST  base + w z -> zout (float4), one partial of r.z per block
enum { UP_PCG = 0, UP_POW = 1, UP_PRE = 2, UP_POST = 3 };

template <int MODE>
__global__ void __launch_bounds__(256) k_lpc_upper(int nn, const int* __restrict__ sliceOff,
                                                   const int* __restrict__ cols, const int* __restrict__ diagSlot,
                                                   const int* __restrict__ rowLen, const bf16_t* __restrict__ G,
                                                   const float4* __restrict__ y, const float* __restrict__ r,
                                                   const float* __restrict__ base, const float* __restrict__ omega,
                                                   float4* __restrict__ zout, float* __restrict__ xout,
                                                   float* __restrict__ part) {
    __shared__ float lds[4];
    float acc = 0.0f;
    const float w = (MODE == UP_PRE || MODE == UP_POST) ? *omega : 1.0f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nn; i += gridDim.x * 256) {
        const int s = i / 64, lane = i % 64;
        const size_t off = (size_t)sliceOff[s];
        const int kd = diagSlot[i], rl = rowLen[i];
        const int* cp = cols + off * 64 + lane;
        const bf16_t* gp = G + off * 9 * 64 + lane;
        const float4 yi = y[i];
        const bf16_t* d = gp + (size_t)kd * 9 * 64;
        float z0 = bf_val(d[0 * 64]) * yi.x + bf_val(d[3 * 64]) * yi.y + bf_val(d[6 * 64]) * yi.z;
        float z1 = bf_val(d[1 * 64]) * yi.x + bf_val(d[4 * 64]) * yi.y + bf_val(d[7 * 64]) * yi.z;
        float z2 = bf_val(d[2 * 64]) * yi.x + bf_val(d[5 * 64]) * yi.y + bf_val(d[8 * 64]) * yi.z;
        for (int k = kd + 1; k < rl; k++) {
            const float4 yj = y[cp[k * 64]];
            const bf16_t* a = gp + (size_t)k * 9 * 64;
            z0 += bf_val(a[0 * 64]) * yj.x + bf_val(a[1 * 64]) * yj.y + bf_val(a[2 * 64]) * yj.z;
            z1 += bf_val(a[3 * 64]) * yj.x + bf_val(a[4 * 64]) * yj.y + bf_val(a[5 * 64]) * yj.z;
            z2 += bf_val(a[6 * 64]) * yj.x + bf_val(a[7 * 64]) * yj.y + bf_val(a[8 * 64]) * yj.z;
        }
        if (MODE == UP_PRE) {
            xout[3 * i] = w * z0;
            xout[3 * i + 1] = w * z1;
            xout[3 * i + 2] = w * z2;
            continue;
        }
        if (MODE == UP_POST) {
            z0 = base[3 * i] + w * z0;
            z1 = base[3 * i + 1] + w * z1;
            z2 = base[3 * i + 2] + w * z2;
        }
        zout[i] = make_float4(z0, z1, z2, 0.0f);
        if (MODE == UP_POW) acc += z0 * z0 + z1 * z1 + z2 * z2;
        else acc += r[3 * i] * z0 + r[3 * i + 1] * z1 + r[3 * i + 2] * z2;
    }
    if (MODE == UP_PRE) return;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// ---------------------------------------------------------------- smoother damping

// Sum of n partials, every thread of the block gets it.
__device__ __forceinline__ float sum_partials(const float* __restrict__ part, int n, float* lds) {
    float a = 0.0f;
    for (int b = threadIdx.x; b < n; b += 256) a += part[b];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = a;
    __syncthreads();
    return (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// Power-iteration start: v_i = (+-1, +-1, +-1) from a hash of i.
__global__ void k_lpc_pow_init(int nn, float4* __restrict__ v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    uint32_t h = (uint32_t)i * 2654435761u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    v[i] = make_float4((h & 1) ? 1.0f : -1.0f, (h & 2) ? 1.0f : -1.0f, (h & 4) ? 1.0f : -1.0f, 0.0f);
}

// w = A (v / |v|) with |v|^2 the sum of the previous sweep's partials (none: |v| = 1);
// A = the level-0 system (static sliced ELL + contact CSR)
__global__ void __launch_bounds__(256) k_lpc_pow_spmv(int nn, const int* __restrict__ sliceOff,
                                                      const int* __restrict__ sliceW, const int* __restrict__ cols,
                                                      const float* __restrict__ vals, const int* __restrict__ cRowPtr,
                                                      const int* __restrict__ cCol, const float* __restrict__ cVal,
                                                      const float4* __restrict__ v, const float* __restrict__ part,
                                                      int npart, float* __restrict__ w) {
    __shared__ float lds[4];
    const float sc = part ? rsqrtf(sum_partials(part, npart, lds)) : 1.0f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nn; i += gridDim.x * 256) {
        const int s = i / 64, lane = i % 64;
        const size_t off = (size_t)sliceOff[s];
        const int W = sliceW[s];
        const int* cp = cols + off * 64 + lane;
        const float* vp = vals + off * 9 * 64 + lane;
        float y0 = 0.0f, y1 = 0.0f, y2 = 0.0f;
        for (int k = 0; k < W; k++) {
            const float4 vj = v[cp[k * 64]];
            const float* a = vp + (size_t)k * 9 * 64;
            y0 += a[0 * 64] * vj.x + a[1 * 64] * vj.y + a[2 * 64] * vj.z;
            y1 += a[3 * 64] * vj.x + a[4 * 64] * vj.y + a[5 * 64] * vj.z;
            y2 += a[6 * 64] * vj.x + a[7 * 64] * vj.y + a[8 * 64] * vj.z;
        }
        for (int k = cRowPtr[i], ce = cRowPtr[i + 1]; k < ce; k++) {
            const float4 vj = v[cCol[k]];
            const float* a = cVal + (size_t)k * 9;
            y0 += a[0] * vj.x + a[1] * vj.y + a[2] * vj.z;
            y1 += a[3] * vj.x + a[4] * vj.y + a[5] * vj.z;
            y2 += a[6] * vj.x + a[7] * vj.y + a[8] * vj.z;
        }
        w[3 * i] = sc * y0;
        w[3 * i + 1] = sc * y1;
        w[3 * i + 2] = sc * y2;
    }
}

// omega = c / lambda, lambda = |G^T G A v| for the last (unit) v
__global__ void __launch_bounds__(256) k_lpc_omega(const float* __restrict__ part, int npart, float c,
                                                   float* __restrict__ omega) {
    __shared__ float lds[4];
    const float lam = sqrtf(sum_partials(part, npart, lds));
    if (threadIdx.x == 0) *omega = (lam > 0.0f && lam < INFINITY) ? c / lam : 0.0f;
}

// ---------------------------------------------------------------- training: loss

// Per row i: Z_ij = sum_{k in P_i} Ghat_ik Ahat_kj for the lower pattern P_i (j <= i),
// and the partials of trace(Ghat Ahat Ghat^T) = sum <Z_ij, Ghat_ij> and of sum_i (o_i0 + o_i1 + o_i2).
struct LossArgs {
    Topo t;
    const float* gh;
    const float* Ah;
    const float* o;
    float* Z;
    double* part;
};

__global__ void __launch_bounds__(256) k_lpc_loss(LossArgs a) {
    __shared__ double lds[2][4];
    double pt = 0.0, po = 0.0;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < a.t.nn) {
        const int li = i % 64;
        const long off = a.t.sliceOff[i / 64];
        const int kd = a.t.diagSlot[i];
        for (int p = 0; p <= kd; p++) {
            const long sp = (off + p) * 64 + li;
            const int j = a.t.cols[sp];
            float Z[9];
#pragma unroll
            for (int c = 0; c < 9; c++) Z[c] = 0.0f;
            for (int q = 0; q <= kd; q++) {
                const long sq = (off + q) * 64 + li;
                const int k = a.t.cols[sq];
                const float* Gk = a.gh + sq * 9;
                if (k == j) {  // Ahat_jj = I
#pragma unroll
                    for (int c = 0; c < 9; c++) Z[c] += Gk[c];
                    continue;
                }
                const long offk = a.t.sliceOff[k / 64];
                const int lk = k % 64, rlk = a.t.rowLen[k];
                for (int t = 0; t < rlk; t++) {
                    const long st = (offk + t) * 64 + lk;
                    if (a.t.cols[st] != j) continue;
                    const float* A = a.Ah + st * 9;
#pragma unroll
                    for (int r = 0; r < 3; r++)
#pragma unroll
                        for (int c = 0; c < 3; c++)
                            Z[3 * r + c] += Gk[3 * r] * A[c] + Gk[3 * r + 1] * A[3 + c] + Gk[3 * r + 2] * A[6 + c];
                    break;
                }
            }
            const float* Gj = a.gh + sp * 9;
            float tr = 0.0f;
#pragma unroll
            for (int c = 0; c < 9; c++) {
                a.Z[sp * 9 + c] = Z[c];
                tr += Z[c] * Gj[c];
            }
            pt += tr;
        }
        po = (double)a.o[(long)i * 16] + a.o[(long)i * 16 + 1] + a.o[(long)i * 16 + 2];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        pt += __shfl_xor(pt, o, 64);
        po += __shfl_xor(po, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        lds[0][threadIdx.x >> 6] = pt;
        lds[1][threadIdx.x >> 6] = po;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a.part[2 * blockIdx.x] = (lds[0][0] + lds[0][1]) + (lds[0][2] + lds[0][3]);
        a.part[2 * blockIdx.x + 1] = (lds[1][0] + lds[1][1]) + (lds[1][2] + lds[1][3]);
    }
}

// scal = [loss, 2 / T, 2 / n, T]
__global__ void __launch_bounds__(256) k_lpc_loss_final(int nblk, int n3, const double* __restrict__ part,
                                                        float* __restrict__ scal) {
    __shared__ double lds[2][256];
    double t = 0.0, so = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) {
        t += part[2 * b];
        so += part[2 * b + 1];
    }
    lds[0][threadIdx.x] = t;
    lds[1][threadIdx.x] = so;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            lds[0][threadIdx.x] += lds[0][threadIdx.x + h];
            lds[1][threadIdx.x] += lds[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double T = lds[0][0];
        scal[0] = (float)(log(T / n3) - 2.0 * lds[1][0] / n3);
        scal[1] = (float)(2.0 / T);
        scal[2] = (float)(2.0 / n3);
        scal[3] = (float)T;
    }
}

// ---------------------------------------------------------------- training: solution term

// bt = b + bc (the right-hand side the PCG solves)
__global__ void k_lpc_rhs(int n3, const float* __restrict__ b, const float* __restrict__ bc, float* __restrict__ bt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n3) bt[i] = b[i] + bc[i];
}

// partials of u.p and b.x (u float4 per node, p / b / x 3 per node)
__global__ void __launch_bounds__(256) k_lpc_sol_dots(int nn, const float4* __restrict__ u, const float* __restrict__ p,
                                                      const float* __restrict__ b, const float* __restrict__ x,
                                                      float* __restrict__ part) {
    __shared__ float lds[2][4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float a0 = 0.0f, a1 = 0.0f;
    if (i < nn) {
        const float4 ui = u[i];
        a0 = ui.x * p[3 * i] + ui.y * p[3 * i + 1] + ui.z * p[3 * i + 2];
        a1 = b[3 * i] * x[3 * i] + b[3 * i + 1] * x[3 * i + 1] + b[3 * i + 2] * x[3 * i + 2];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o, 64);
        a1 += __shfl_xor(a1, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        lds[0][threadIdx.x >> 6] = a0;
        lds[1][threadIdx.x >> 6] = a1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (lds[0][0] + lds[0][1]) + (lds[0][2] + lds[0][3]);
        part[2 * blockIdx.x + 1] = (lds[1][0] + lds[1][1]) + (lds[1][2] + lds[1][3]);
    }
}

// L_sol = 1 - alpha^2 / (beta gamma) with alpha = b^T G^T G b, beta = u^T A u, gamma = b^T x*;
// sc = [c1, c2] of dL/dG_ij = c1 y_i b_j^T + c2 (y_i p_j^T + w_i b_j^T); adds lam L_sol to the loss.
__global__ void __launch_bounds__(256) k_lpc_sol_final(int nblk, const float* __restrict__ palpha,
                                                       const float* __restrict__ pdots, float lam,
                                                       float* __restrict__ scal, float* __restrict__ sc) {
    __shared__ double lds[3][256];
    double a = 0.0, bt = 0.0, gm = 0.0;
    for (int k = threadIdx.x; k < nblk; k += 256) {
        a += palpha[k];
        bt += pdots[2 * k];
        gm += pdots[2 * k + 1];
    }
    lds[0][threadIdx.x] = a;
    lds[1][threadIdx.x] = bt;
    lds[2][threadIdx.x] = gm;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            for (int q = 0; q < 3; q++) lds[q][threadIdx.x] += lds[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double al = lds[0][0], be = lds[1][0], ga = lds[2][0];
        const bool ok = be > 0.0 && ga > 0.0;
        const double L = ok ? 1.0 - al * al / (be * ga) : 0.0;
        sc[0] = ok ? (float)(-4.0 * al / (be * ga)) : 0.0f;
        sc[1] = ok ? (float)(2.0 * al * al / (be * be * ga)) : 0.0f;
        sc[2] = (float)L;
        scal[0] += lam * (float)L;
    }
}

// ---------------------------------------------------------------- training: smoothing term
// L_sm = |(I - w G^T G A) v|_A^2 / |v|_A^2 for a probe v: how much of the probe's energy-norm
// error one smoothing sweep leaves.  Its gradient dL/dG_ij = c2 (y_i p_j^T + w_i u_j^T) with
// u = A v, y = G u, e = v - w G^T y, p = A e, w = G p, c2 = -2 w / (v.u) goes through
// k_lpc_sol_grad (c1 = 0).

// v_i = (+-1, +-1, +-1) from a hash of (i, seed)
__global__ void k_lpc_probe(int nn, uint32_t seed, float4* __restrict__ v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    uint32_t h = (uint32_t)i * 2654435761u + seed * 2246822519u;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    v[i] = make_float4((h & 1) ? 1.0f : -1.0f, (h & 2) ? 1.0f : -1.0f, (h & 4) ? 1.0f : -1.0f, 0.0f);
}

// e = v - w s
__global__ void k_lpc_sub(int nn, const float4* __restrict__ v, const float4* __restrict__ s, float w,
                          float4* __restrict__ e) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    const float4 a = v[i], b = s[i];
    e[i] = make_float4(a.x - w * b.x, a.y - w * b.y, a.z - w * b.z, 0.0f);
}

// partials of e.p and v.u (e, v float4; p, u 3 per node)
__global__ void __launch_bounds__(256) k_lpc_smooth_dots(int nn, const float4* __restrict__ e,
                                                         const float* __restrict__ p, const float4* __restrict__ v,
                                                         const float* __restrict__ u, float* __restrict__ part) {
    __shared__ float lds[2][4];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float a0 = 0.0f, a1 = 0.0f;
    if (i < nn) {
        const float4 ei = e[i], vi = v[i];
        a0 = ei.x * p[3 * i] + ei.y * p[3 * i + 1] + ei.z * p[3 * i + 2];
        a1 = vi.x * u[3 * i] + vi.y * u[3 * i + 1] + vi.z * u[3 * i + 2];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o, 64);
        a1 += __shfl_xor(a1, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        lds[0][threadIdx.x >> 6] = a0;
        lds[1][threadIdx.x >> 6] = a1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = (lds[0][0] + lds[0][1]) + (lds[0][2] + lds[0][3]);
        part[2 * blockIdx.x + 1] = (lds[1][0] + lds[1][1]) + (lds[1][2] + lds[1][3]);
    }
}

// L_sm = (e.p) / (v.u); sc = [0, -2 w / (v.u), L_sm]; adds lam L_sm to the loss
__global__ void __launch_bounds__(256) k_lpc_smooth_final(int nblk, const float* __restrict__ pdots, float lam,
                                                          float w, float* __restrict__ scal, float* __restrict__ sc) {
    __shared__ double lds[2][256];
    double num = 0.0, den = 0.0;
    for (int k = threadIdx.x; k < nblk; k += 256) {
        num += pdots[2 * k];
        den += pdots[2 * k + 1];
    }
    lds[0][threadIdx.x] = num;
    lds[1][threadIdx.x] = den;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            lds[0][threadIdx.x] += lds[0][threadIdx.x + h];
            lds[1][threadIdx.x] += lds[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double n = lds[0][0], d = lds[1][0];
        const bool ok = d > 0.0;
        const double L = ok ? n / d : 0.0;
        sc[0] = 0.0f;
        sc[1] = ok ? (float)(-2.0 * w / d) : 0.0f;
        sc[2] = (float)L;
        scal[0] += lam * (float)L;
    }
}

// Z_ij += lam (T / 2) dL_sol/dGhat_ij for the lower and diagonal slots of row i, where
// dL_sol/dGhat_ij = (c1 y_i b_j^T + c2 (y_i p_j^T + w_i b_j^T)) S_j^T; the decoders' backward
// multiplies Z by 2 / T.
__global__ void __launch_bounds__(256) k_lpc_sol_grad(Topo t, const float4* __restrict__ y,
                                                      const float4* __restrict__ w, const float* __restrict__ b,
                                                      const float* __restrict__ p, const float* __restrict__ S,
                                                      const float* __restrict__ scal, const float* __restrict__ sc,
                                                      float lam, float* __restrict__ Z) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= t.nn) return;
    const float f = lam * scal[3] * 0.5f;
    const float c1 = sc[0] * f, c2 = sc[1] * f;
    const float4 yi = y[i], wi = w[i];
    const float ya[3] = {yi.x, yi.y, yi.z}, wa[3] = {wi.x, wi.y, wi.z};
    const int li = i % 64;
    const long off = t.sliceOff[i / 64];
    const int kd = t.diagSlot[i];
    for (int k = 0; k <= kd; k++) {
        const long sg = (off + k) * 64 + li;
        const int j = t.cols[sg];
        const float bj[3] = {b[3 * j], b[3 * j + 1], b[3 * j + 2]};
        const float pj[3] = {p[3 * j], p[3 * j + 1], p[3 * j + 2]};
        float D[9];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int c = 0; c < 3; c++) D[3 * a + c] = c1 * ya[a] * bj[c] + c2 * (ya[a] * pj[c] + wa[a] * bj[c]);
        const float* Sj = S + 9 * (size_t)j;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int c = 0; c < 3; c++)  // (D S_j^T)_ac = sum_k D_ak S_j[c][k]
                Z[sg * 9 + 3 * a + c] += D[3 * a] * Sj[3 * c] + D[3 * a + 1] * Sj[3 * c + 1] + D[3 * a + 2] * Sj[3 * c + 2];
    }
}

// ---------------------------------------------------------------- training: backward

struct BackBufs {
    float* Z;
    const float* scal;
    float* dh;      // fp32 64 per node row: dL/dh of the layer being processed
    float* de;      // fp32 64 per slot
    bf16_t* dob;    // 16 per node row
    bf16_t* ddnb;   // 64 per node row
    bf16_t* dqb;    // 16 per slot
    bf16_t* ddeb;   // 64 per slot
    bf16_t* dhb[NLAYER + 1];  // dL/dh_l (l = 1..3)
    bf16_t* dub[NLAYER];
    bf16_t* dmb[NLAYER];
    bf16_t* dzb[NLAYER];
    bf16_t* dPb[NLAYER];
    bf16_t* dQb[NLAYER];
    bf16_t* dh0b;   // encoder: dL/d(pre-activation) of h0, per node row
    bf16_t* de0b;   // encoder: of e0, per slot
};

struct BackArgs {
    Topo t;
    int layer;
    const float* w;
    const s4* wt;  // transposed fragments
    Stash st;
    BackBufs bb;
};

// Decoders: dL/do and dL/dq from Z, back through both decoder MLPs.
__global__ void __launch_bounds__(LB) k_lpc_bdec(BackArgs a) {
    const int s = blockIdx.x;
    const int lane = lane_id(), g = lane >> 4, n = lane & 15;
    const int li = 16 * (threadIdx.x >> 6) + n;
    const int row = s * 64 + li;
    const bool live = row < a.t.nn;
    const float sT = a.bb.scal[1], sn = a.bb.scal[2];
    const long off = a.t.sliceOff[s];
    const int W = a.t.sliceW[s];
    const int rl = live ? a.t.rowLen[row] : 0;
    f4v dov = zero4();
    if (live) {
        const long sd = (off + a.t.diagSlot[row]) * 64 + li;
        const float* Z = a.bb.Z + sd * 9;
        const float* o = a.st.o + (long)row * 16;
        if (g == 0) dov = f4v{sT * Z[0] * expf(o[0]) - sn, sT * Z[4] * expf(o[1]) - sn, sT * Z[8] * expf(o[2]) - sn, sT * Z[3]};
        else if (g == 1) dov = f4v{sT * Z[6], sT * Z[7], 0.0f, 0.0f};
    }
    const s4 dob[1] = {to_bf(dov)};
    st16(a.bb.dob, row, dob[0]);
    f4v dn[4] = {zero4(), zero4(), zero4(), zero4()};
    mmT<4, 1>(dn, dob, a.wt + O_DN2 / 4);
    const BTile dnh = ld_btile(a.st.dnh, row);
    BTile dnb;
#pragma unroll
    for (int m = 0; m < 4; m++) dnb.v[m] = to_bf(mask4(dn[m], dnh.v[m]));
    st_btile(a.bb.ddnb, row, dnb);
    Tile dh;
#pragma unroll
    for (int m = 0; m < 4; m++) dh.v[m] = zero4();
    mmT<4, 4>(dh.v, dnb.v, a.wt + O_DN1 / 4);
    st_tile(a.bb.dh, row, dh);
    for (int k = 0; k < W; k++) {
        const long sig = (off + k) * 64 + li;
        const int col = a.t.cols[sig];
        const bool lower = live && k < rl && col < row;
        f4v dq = zero4();
        if (lower) {
            const float* Z = a.bb.Z + sig * 9;
            if (g == 0) dq = f4v{sT * Z[0], sT * Z[1], sT * Z[2], sT * Z[3]};
            else if (g == 1) dq = f4v{sT * Z[4], sT * Z[5], sT * Z[6], sT * Z[7]};
            else if (g == 2) dq = f4v{sT * Z[8], 0.0f, 0.0f, 0.0f};
        }
        const s4 dqb[1] = {to_bf(dq)};
        st16(a.bb.dqb, sig, dqb[0]);
        f4v dd[4] = {zero4(), zero4(), zero4(), zero4()};
        mmT<4, 1>(dd, dqb, a.wt + O_DE2 / 4);
        const BTile deh = ld_btile(a.st.deh, sig);
        BTile ddb;
#pragma unroll
        for (int m = 0; m < 4; m++) ddb.v[m] = to_bf(mask4(dd[m], deh.v[m]));
        st_btile(a.bb.ddeb, sig, ddb);
        Tile de;
#pragma unroll
        for (int m = 0; m < 4; m++) de.v[m] = zero4();
        mmT<4, 4>(de.v, ddb.v, a.wt + O_DE1 / 4);
        st_tile(a.bb.de, sig, de);
    }
}

// dh += dP_l Wa_l^T + dQ_l Wb_l^T, with dQ_l of a row gathered from the dz of the transposed
// slots of its edges; stores dQ_l (bf16).
__device__ __forceinline__ void finish_dh(const BackArgs& a, int l, int row, bool live, int li, long off, int W,
                                          int rl, Tile& dh) {
    f4v dq[4] = {zero4(), zero4(), zero4(), zero4()};
    for (int k = 0; k < W; k++) {
        const long sig = (off + k) * 64 + li;
        const int col = a.t.cols[sig];
        if (!(live && k < rl && col != row)) continue;
        const BTile t = ld_btile(a.bb.dzb[l], a.t.mirror[sig]);
#pragma unroll
        for (int m = 0; m < 4; m++) dq[m] += from_bf(t.v[m]);
    }
    BTile dqb;
#pragma unroll
    for (int m = 0; m < 4; m++) dqb.v[m] = to_bf(dq[m]);
    st_btile(a.bb.dQb[l], row, dqb);
    const BTile dpb = ld_btile(a.bb.dPb[l], row);
    mmT<4, 4>(dh.v, dpb.v, a.wt + lw(l, L_WA) / 4);
    mmT<4, 4>(dh.v, dqb.v, a.wt + lw(l, L_WB) / 4);
}

// Layer l backward: node update, then the edge messages (dm, dz, de) and dP.
__global__ void __launch_bounds__(LB) k_lpc_blayer(BackArgs a) {
    __shared__ s4 wl[2 * 16 * 64];  // transposed fragments of W2 and Wc
    const int s = blockIdx.x;
    const int lane = lane_id(), n = lane & 15;
    const int li = 16 * (threadIdx.x >> 6) + n;
    const int row = s * 64 + li;
    const bool live = row < a.t.nn;
    const int l = a.layer;
    {
        const s4* f2 = a.wt + lw(l, L_W2) / 4;
        const s4* fc = a.wt + lw(l, L_WC) / 4;
        for (int t = threadIdx.x; t < 1024; t += LB) {
            wl[t] = f2[t];
            wl[1024 + t] = fc[t];
        }
    }
    __syncthreads();
    const long off = a.t.sliceOff[s];
    const int W = a.t.sliceW[s];
    const int rl = live ? a.t.rowLen[row] : 0;
    Tile dh = ld_tile(a.bb.dh, row);
    if (l + 1 < NLAYER) finish_dh(a, l + 1, row, live, li, off, W, rl, dh);
    if (!live)
#pragma unroll
        for (int m = 0; m < 4; m++) dh.v[m] = zero4();
    const BTile dhb = to_bf(dh);
    st_btile(a.bb.dhb[l + 1], row, dhb);
    f4v du[4] = {zero4(), zero4(), zero4(), zero4()};
    mmT<4, 4>(du, dhb.v, a.wt + lw(l, L_U2) / 4);
    const BTile u = ld_btile(a.st.u[l], row);
    BTile dub;
#pragma unroll
    for (int m = 0; m < 4; m++) dub.v[m] = to_bf(mask4(du[m], u.v[m]));
    st_btile(a.bb.dub[l], row, dub);
    f4v dhin[8];
#pragma unroll
    for (int m = 0; m < 8; m++) dhin[m] = zero4();
    mmT<8, 4>(dhin, dub.v, a.wt + lw(l, L_U1) / 4);
    Tile dres;
#pragma unroll
    for (int m = 0; m < 4; m++) dres.v[m] = dh.v[m] + dhin[m];
    st_tile(a.bb.dh, row, dres);
    f4v dP[4] = {zero4(), zero4(), zero4(), zero4()};
    for (int k = 0; k < W; k++) {
        const long sig = (off + k) * 64 + li;
        const int col = a.t.cols[sig];
        const bool edge = live && k < rl && col != row;
        Tile de = ld_tile(a.bb.de, sig);
        BTile dmb;
#pragma unroll
        for (int m = 0; m < 4; m++) dmb.v[m] = to_bf(edge ? de.v[m] + dhin[4 + m] : zero4());
        st_btile(a.bb.dmb[l], sig, dmb);
        f4v dz[4] = {zero4(), zero4(), zero4(), zero4()};
        mmT<4, 4>(dz, dmb.v, wl);
        const BTile z = ld_btile(a.st.z[l], sig);
        BTile dzb;
#pragma unroll
        for (int m = 0; m < 4; m++) dzb.v[m] = to_bf(mask4(dz[m], z.v[m]));
        st_btile(a.bb.dzb[l], sig, dzb);
        mmT<4, 4>(de.v, dzb.v, wl + 1024);
        st_tile(a.bb.de, sig, de);
#pragma unroll
        for (int m = 0; m < 4; m++) dP[m] += from_bf(dzb.v[m]);
    }
    BTile dpb;
#pragma unroll
    for (int m = 0; m < 4; m++) dpb.v[m] = to_bf(dP[m]);
    st_btile(a.bb.dPb[l], row, dpb);
}

// Layer 0's dP/dQ terms, then the encoders' pre-activation gradients.
__global__ void __launch_bounds__(LB) k_lpc_bfinal(BackArgs a) {
    const int s = blockIdx.x;
    const int lane = lane_id(), n = lane & 15;
    const int li = 16 * (threadIdx.x >> 6) + n;
    const int row = s * 64 + li;
    const bool live = row < a.t.nn;
    const long off = a.t.sliceOff[s];
    const int W = a.t.sliceW[s];
    const int rl = live ? a.t.rowLen[row] : 0;
    Tile dh = ld_tile(a.bb.dh, row);
    finish_dh(a, 0, row, live, li, off, W, rl, dh);
    const BTile h0 = ld_btile(a.st.hb[0], row);
    BTile d0;
#pragma unroll
    for (int m = 0; m < 4; m++) d0.v[m] = to_bf(live ? mask4(dh.v[m], h0.v[m]) : zero4());
    st_btile(a.bb.dh0b, row, d0);
    for (int k = 0; k < W; k++) {
        const long sig = (off + k) * 64 + li;
        const Tile de = ld_tile(a.bb.de, sig);
        const BTile e0 = ld_btile(a.st.e[0], sig);
        BTile db;
#pragma unroll
        for (int m = 0; m < 4; m++) db.v[m] = to_bf(mask4(de.v[m], e0.v[m]));
        st_btile(a.bb.de0b, sig, db);
    }
}

// dW += X^T dY and db += sum dY over a job's rows (bf16 rows, fp32 MFMA accumulation),
// reduced over the workgroup and added to the global gradient.
struct WJob {
    const bf16_t* x0;  // 16 (KB == 1) or 64 features per row
    const bf16_t* x1;  // the second 64 features (KB == 8)
    const bf16_t* dy;  // 16 (MB == 1) or 64 per row
    long rows;
    int gOff, bOff;    // flat offsets of the matrix and bias (-1: none) gradients
};

template <int KB, int MB>
__global__ void __launch_bounds__(256) k_lpc_wgrad(const WJob* __restrict__ jobs, float* __restrict__ grad,
                                                   int chunk) {
    constexpr int KX = 16 * KB, NY = 16 * MB;
    __shared__ bf16_t xs[4][16][KX + 2];
    __shared__ bf16_t ys[4][16][NY + 2];
    __shared__ float red[KX * NY + NY];
    const WJob J = jobs[blockIdx.y];
    const long r0 = (long)blockIdx.x * chunk;
    if (r0 >= J.rows) return;
    const long r1 = r0 + chunk < J.rows ? r0 + chunk : J.rows;
    const int wv = threadIdx.x >> 6, lane = lane_id(), g = lane >> 4, n = lane & 15;
    f4v acc[KB][MB], accb[MB];
#pragma unroll
    for (int k = 0; k < KB; k++)
#pragma unroll
        for (int m = 0; m < MB; m++) acc[k][m] = zero4();
#pragma unroll
    for (int m = 0; m < MB; m++) accb[m] = zero4();
    const short one = (short)0x3F80;
    const s4 aones = n == 0 ? s4{one, one, one, one} : s4{0, 0, 0, 0};
    for (long t0 = r0; t0 < r1; t0 += 64) {  // the same trip count in every wave
        const long rr = t0 + 16 * wv + n;
        const bool ok = rr < r1;
        // stage X and dY rows into natural feature order
        if (KB == 1) {
            const s4 v = ok ? *reinterpret_cast<const s4*>(J.x0 + rr * 16 + 4 * g) : s4{0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < 4; i++) xs[wv][n][4 * g + i] = (bf16_t)v[i];
        } else {
#pragma unroll
            for (int h = 0; h < KB / 4; h++) {
                const bf16_t* src = h == 0 ? J.x0 : J.x1;
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const s4 v = ok ? *reinterpret_cast<const s4*>(src + rr * 64 + 16 * g + 4 * m) : s4{0, 0, 0, 0};
#pragma unroll
                    for (int i = 0; i < 4; i++) xs[wv][n][64 * h + 16 * m + 4 * g + i] = (bf16_t)v[i];
                }
            }
        }
        if (MB == 1) {
            const s4 v = ok ? *reinterpret_cast<const s4*>(J.dy + rr * 16 + 4 * g) : s4{0, 0, 0, 0};
#pragma unroll
            for (int i = 0; i < 4; i++) ys[wv][n][4 * g + i] = (bf16_t)v[i];
        } else {
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const s4 v = ok ? *reinterpret_cast<const s4*>(J.dy + rr * 64 + 16 * g + 4 * m) : s4{0, 0, 0, 0};
#pragma unroll
                for (int i = 0; i < 4; i++) ys[wv][n][16 * m + 4 * g + i] = (bf16_t)v[i];
            }
        }
        __syncthreads();
        s4 b[MB];
#pragma unroll
        for (int m = 0; m < MB; m++)
            b[m] = s4{(short)ys[wv][4 * g][16 * m + n], (short)ys[wv][4 * g + 1][16 * m + n],
                      (short)ys[wv][4 * g + 2][16 * m + n], (short)ys[wv][4 * g + 3][16 * m + n]};
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const s4 av = s4{(short)xs[wv][4 * g][16 * k + n], (short)xs[wv][4 * g + 1][16 * k + n],
                             (short)xs[wv][4 * g + 2][16 * k + n], (short)xs[wv][4 * g + 3][16 * k + n]};
#pragma unroll
            for (int m = 0; m < MB; m++) acc[k][m] = MFMA(av, b[m], acc[k][m]);
        }
#pragma unroll
        for (int m = 0; m < MB; m++) accb[m] = MFMA(aones, b[m], accb[m]);
        __syncthreads();
    }
    for (int w2 = 0; w2 < 4; w2++) {
        if (wv == w2) {
#pragma unroll
            for (int k = 0; k < KB; k++)
#pragma unroll
                for (int m = 0; m < MB; m++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int idx = (16 * k + 4 * g + i) * NY + 16 * m + n;
                        red[idx] = (w2 ? red[idx] : 0.0f) + acc[k][m][i];
                    }
            if (g == 0)
#pragma unroll
                for (int m = 0; m < MB; m++) {
                    const int idx = KX * NY + 16 * m + n;
                    red[idx] = (w2 ? red[idx] : 0.0f) + accb[m][0];
                }
        }
        __syncthreads();
    }
    for (int q = threadIdx.x; q < KX * NY; q += 256) atomicAdd(grad + J.gOff + q, red[q]);
    if (J.bOff >= 0)
        for (int q = threadIdx.x; q < NY; q += 256) atomicAdd(grad + J.bOff + q, red[KX * NY + q]);
}

__global__ void k_lpc_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

// ---------------------------------------------------------------- host: weights and inference

int lpc_weight_count() { return W_COUNT; }

void lpc_set_weights(cs_ctx* C, const float* w) {
    LpcNet& N = C->lpcNet;
    hipStream_t st = C->stream;
    N.w.as<float>(W_COUNT);
    N.frag.as<bf16_t>(W_COUNT);
    N.fragT.as<bf16_t>(W_COUNT);
    CS_HIP(hipMemcpyAsync(N.w.p, w, sizeof(float) * W_COUNT, hipMemcpyHostToDevice, st));
    lpc_refresh_fragments(C);
    // a new set of weights restarts the optimiser
    N.m.as<float>(W_COUNT);
    N.v.as<float>(W_COUNT);
    fill_f32(N.m.ptr<float>(), 0.0f, W_COUNT, st);
    fill_f32(N.v.ptr<float>(), 0.0f, W_COUNT, st);
    N.t = 0;
    CS_HIP(hipStreamSynchronize(st));
    N.loaded = true;
}

void lpc_refresh_fragments(cs_ctx* C) {
    LpcNet& N = C->lpcNet;
    N.version++;  // every weight change ends here
    for (const MatDesc& m : kMats) {
        const int nt = m.K * m.N / 4;
        hipLaunchKernelGGL(k_lpc_frag, dim3(cdiv(nt, 256)), dim3(256), 0, C->stream, N.w.ptr<float>(),
                           N.frag.ptr<bf16_t>(), N.fragT.ptr<bf16_t>(), m.off, m.K, m.N);
    }
    CS_LAUNCH_CHECK();
}

void lpc_get_weights(cs_ctx* C, float* w) {
    if (!C->lpcNet.loaded) throw StatusError(CS_E_STATE, "no learned-preconditioner weights are set");
    CS_HIP(hipMemcpyAsync(w, C->lpcNet.w.p, sizeof(float) * W_COUNT, hipMemcpyDeviceToHost, C->stream));
    CS_HIP(hipStreamSynchronize(C->stream));
}

namespace {

void lpc_alloc(cs_ctx* C, int c) {
    ClothSys& S = C->sys[c];
    LpcSys& L = S.lpc;
    if (L.built) return;
    hipStream_t st = C->stream;
    const size_t slots = (size_t)S.totalSlots * SLICE;
    const size_t rows = (size_t)S.nslices * SLICE;
    // mirror slot of every edge slot
    std::vector<int> mirror(slots, -1);
    for (int i = 0; i < S.nn; i++) {
        const int si = i / SLICE, li = i % SLICE;
        for (int k = 0; k < S.hRowLen[i]; k++) {
            const int j = S.hCol[S.hRowPtr[i] + k];
            if (j == i) continue;
            const int* b = S.hCol.data() + S.hRowPtr[j];
            const int kj = (int)(std::lower_bound(b, b + S.hRowLen[j], i) - b);
            mirror[((size_t)S.hSliceOff[si] + k) * SLICE + li] =
                (int)(((size_t)S.hSliceOff[j / SLICE] + kj) * SLICE + j % SLICE);
        }
    }
    L.mirror.as<int>(slots);
    CS_HIP(hipMemcpyAsync(L.mirror.p, mirror.data(), sizeof(int) * slots, hipMemcpyHostToDevice, st));
    L.S.as<float>(9 * (size_t)S.nn);
    L.G.as<bf16_t>(9 * slots);
    fill_u32(L.G.ptr<uint32_t>(), 0u, (long)(9 * slots / 2), st);  // padding slots stay zero
    L.y.as<float4>(S.nn);
    for (DBuf* b : {&L.h, &L.P, &L.Q, &L.P2, &L.Q2}) b->as<float>(rows * HF);
    L.e.as<bf16_t>(slots * HF);
    CS_HIP(hipStreamSynchronize(st));
    L.built = true;
}

Topo topo_of(ClothSys& S) {
    Topo t;
    t.nn = S.nn;
    t.sliceOff = S.sliceOff.ptr<int>();
    t.sliceW = S.sliceW.ptr<int>();
    t.cols = S.cols.ptr<int>();
    t.rowLen = S.rowLen.ptr<int>();
    t.diagSlot = S.diagSlot.ptr<int>();
    t.mirror = S.lpc.mirror.ptr<int>();
    return t;
}

// Runs the network on cloth c's assembled system (Minv current) and writes G.
template <bool TRAIN>
void lpc_forward(cs_ctx* C, int c, const Stash& st) {
    ClothSys& S = C->sys[c];
    LpcSys& L = S.lpc;
    LpcNet& N = C->lpcNet;
    hipStream_t s = C->stream;
    hipLaunchKernelGGL(k_lpc_scale, dim3(cdiv(S.nn, 256)), dim3(256), 0, s, S.nn, S.Minv.ptr<float>(),
                       L.S.ptr<float>());
    const s4* wf = reinterpret_cast<const s4*>(N.frag.ptr<bf16_t>());
    EncArgs ea;
    ea.t = topo_of(S);
    ea.vals = S.vals.ptr<float>();
    ea.cRowPtr = S.cRowPtr.ptr<int>();
    ea.cCol = S.nContact > 0 ? S.cCol.ptr<int>() : S.cRowPtr.ptr<int>();
    ea.cVal = S.nContact > 0 ? S.cVal.ptr<float>() : S.vals.ptr<float>();
    ea.S = L.S.ptr<float>();
    ea.w = N.w.ptr<float>();
    ea.wf = wf;
    ea.e = TRAIN ? st.e[0] : L.e.ptr<bf16_t>();
    ea.h = L.h.ptr<float>();
    ea.P = L.P.ptr<float>();
    ea.Q = L.Q.ptr<float>();
    ea.st = st;
    hipLaunchKernelGGL(k_lpc_encode<TRAIN>, dim3(S.nslices), dim3(LB), 0, s, ea);
    float* PQ[2][2] = {{L.P.ptr<float>(), L.Q.ptr<float>()}, {L.P2.ptr<float>(), L.Q2.ptr<float>()}};
    for (int l = 0; l < NLAYER; l++) {
        LayerArgs la;
        la.t = topo_of(S);
        la.layer = l;
        la.S = L.S.ptr<float>();
        la.P = PQ[l & 1][0];
        la.Q = PQ[l & 1][1];
        la.Pn = PQ[(l + 1) & 1][0];
        la.Qn = PQ[(l + 1) & 1][1];
        la.h = L.h.ptr<float>();
        la.ein = TRAIN ? st.e[l] : L.e.ptr<bf16_t>();
        la.eout = TRAIN ? st.e[l + 1] : L.e.ptr<bf16_t>();
        la.G = L.G.ptr<bf16_t>();
        la.w = N.w.ptr<float>();
        la.wf = wf;
        la.st = st;
        if (l + 1 < NLAYER)
            hipLaunchKernelGGL((k_lpc_layer<TRAIN, false>), dim3(S.nslices), dim3(LB), 0, s, la);
        else
            hipLaunchKernelGGL((k_lpc_layer<TRAIN, true>), dim3(S.nslices), dim3(LB), 0, s, la);
    }
    CS_LAUNCH_CHECK();
}

}  // namespace

void lpc_prepare(cs_ctx* C, int c) {
    if (!C->lpcNet.loaded) throw StatusError(CS_E_STATE, "CS_PRECOND_LEARNED needs weights (cs_set_learned_weights)");
    lpc_alloc(C, c);
    lpc_forward<false>(C, c, Stash{});
}

void lpc_apply(cs_ctx* C, int c, const float* r, float4* z, float* partRZ, int nb) {
    ClothSys& S = C->sys[c];
    LpcSys& L = S.lpc;
    hipStream_t s = C->stream;
    const int g = std::min(nb, cdiv(S.nn, 256));
    hipLaunchKernelGGL(k_lpc_lower, dim3(g), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(), S.cols.ptr<int>(),
                       S.diagSlot.ptr<int>(), L.G.ptr<bf16_t>(), r, L.y.ptr<float4>());
    // the r.z partials are summed over nb slots by the next kernel: write all of them
    hipLaunchKernelGGL(k_lpc_upper<UP_PCG>, dim3(nb), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(),
                       S.cols.ptr<int>(), S.diagSlot.ptr<int>(), S.rowLen.ptr<int>(), L.G.ptr<bf16_t>(),
                       L.y.ptr<float4>(), r, nullptr, nullptr, z, nullptr, partRZ);
    CS_LAUNCH_CHECK();
}

namespace {

void launch_lower(cs_ctx* C, ClothSys& S, const float* r, int g) {
    hipLaunchKernelGGL(k_lpc_lower, dim3(g), dim3(256), 0, C->stream, S.nn, S.sliceOff.ptr<int>(), S.cols.ptr<int>(),
                       S.diagSlot.ptr<int>(), S.lpc.G.ptr<bf16_t>(), r, S.lpc.y.ptr<float4>());
}

}  // namespace

// Damping of the learned smoother: omega = 1.3 / lambda_max(G^T G A), lambda from 12 power
// iterations (an estimate from below, so omega lambda_max stays under 2 unless the estimate
// is off by more than a third).
void lpc_smoother_prepare(cs_ctx* C, int c) {
    ClothSys& S = C->sys[c];
    LpcSys& L = S.lpc;
    hipStream_t s = C->stream;
    const int g = cdiv(S.nn, 256);
    float4* v = L.pv.as<float4>(S.nn);
    float* w = L.pw.as<float>(3 * (size_t)S.nn);
    float* part = L.ppart.as<float>(g);
    L.omega.as<float>(1);
    const int* cCol = S.nContact > 0 ? S.cCol.ptr<int>() : S.cRowPtr.ptr<int>();
    const float* cVal = S.nContact > 0 ? S.cVal.ptr<float>() : S.vals.ptr<float>();
    // 12 sweeps from the hashed start, once per assembled system and set of weights.  (A
    // warm start from the previous step's vector with fewer sweeps misses the new
    // dominant modes that fresh contact blocks bring, and the smoother then diverges.)
    hipLaunchKernelGGL(k_lpc_pow_init, dim3(g), dim3(256), 0, s, S.nn, v);
    for (int it = 0; it < 12; it++) {
        hipLaunchKernelGGL(k_lpc_pow_spmv, dim3(g), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(),
                           S.sliceW.ptr<int>(), S.cols.ptr<int>(), S.vals.ptr<float>(), S.cRowPtr.ptr<int>(), cCol,
                           cVal, v, it ? part : nullptr, g, w);
        launch_lower(C, S, w, g);
        hipLaunchKernelGGL(k_lpc_upper<UP_POW>, dim3(g), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(),
                           S.cols.ptr<int>(), S.diagSlot.ptr<int>(), S.rowLen.ptr<int>(), L.G.ptr<bf16_t>(),
                           L.y.ptr<float4>(), nullptr, nullptr, nullptr, v, nullptr, part);
    }
    hipLaunchKernelGGL(k_lpc_omega, dim3(1), dim3(256), 0, s, part, g, 1.3f, L.omega.ptr<float>());
    CS_LAUNCH_CHECK();
    L.omegaAsm = S.asmEpoch;
    L.omegaNet = C->lpcNet.version;
}

void lpc_smooth_pre(cs_ctx* C, int c, const float* r, float* x) {
    ClothSys& S = C->sys[c];
    LpcSys& L = S.lpc;
    const int g = cdiv(S.nn, 256);
    launch_lower(C, S, r, g);
    hipLaunchKernelGGL(k_lpc_upper<UP_PRE>, dim3(g), dim3(256), 0, C->stream, S.nn, S.sliceOff.ptr<int>(),
                       S.cols.ptr<int>(), S.diagSlot.ptr<int>(), S.rowLen.ptr<int>(), L.G.ptr<bf16_t>(),
                       L.y.ptr<float4>(), nullptr, nullptr, L.omega.ptr<float>(), nullptr, x, nullptr);
    CS_LAUNCH_CHECK();
}

void lpc_smooth_post(cs_ctx* C, int c, const float* res, const float* base, const float* r, float4* z,
                     float* partRZ, int nb) {
    ClothSys& S = C->sys[c];
    LpcSys& L = S.lpc;
    launch_lower(C, S, res, std::min(nb, cdiv(S.nn, 256)));
    hipLaunchKernelGGL(k_lpc_upper<UP_POST>, dim3(nb), dim3(256), 0, C->stream, S.nn, S.sliceOff.ptr<int>(),
                       S.cols.ptr<int>(), S.diagSlot.ptr<int>(), S.rowLen.ptr<int>(), L.G.ptr<bf16_t>(),
                       L.y.ptr<float4>(), r, base, L.omega.ptr<float>(), z, nullptr, partRZ);
    CS_LAUNCH_CHECK();
}

float lpc_smoother_omega(cs_ctx* C, int c) {
    float w = 0.0f;
    CS_HIP(hipMemcpyAsync(&w, C->sys[c].lpc.omega.p, sizeof(float), hipMemcpyDeviceToHost, C->stream));
    CS_HIP(hipStreamSynchronize(C->stream));
    return w;
}

void lpc_factor_blocks(cs_ctx* C, int c, float* diag, float* lower) {
    ClothSys& S = C->sys[c];
    LpcSys& L = S.lpc;
    const size_t slots = (size_t)S.totalSlots * SLICE;
    std::vector<bf16_t> G(9 * slots);
    CS_HIP(hipMemcpyAsync(G.data(), L.G.p, sizeof(bf16_t) * 9 * slots, hipMemcpyDeviceToHost, C->stream));
    CS_HIP(hipStreamSynchronize(C->stream));
    long t = 0;
    for (int i = 0; i < S.nn; i++) {
        const int si = i / SLICE, li = i % SLICE;
        for (int k = 0; k < S.hRowLen[i]; k++) {
            const int j = S.hCol[S.hRowPtr[i] + k];
            if (j > i) break;
            const size_t sig = ((size_t)S.hSliceOff[si] + k) * SLICE + li;
            float* out = j == i ? diag + 9 * (size_t)i : lower + 9 * t++;
            for (int q = 0; q < 9; q++) out[q] = bf_val(G[(sig >> 6) * 576 + q * 64 + (sig & 63)]);
        }
    }
}

// ---------------------------------------------------------------- host: training

namespace {

// Carves 256-byte aligned sub-buffers out of one arena.
struct Carver {
    size_t used = 0;
    char* base = nullptr;
    template <class T> T* take(size_t n) {
        T* p = reinterpret_cast<T*>(base + used);
        used += (n * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

size_t carve_all(char* base, size_t slots, size_t rows, int nblk, Stash& st, BackBufs& bb, double*& part,
                 float*& scal) {
    Carver c;
    c.base = base;
    const size_t E64 = slots * HF, N64 = rows * HF;
    st.xn = c.take<bf16_t>(rows * 16);
    st.xe = c.take<bf16_t>(slots * 16);
    st.Ah = c.take<float>(slots * 9);
    for (auto& p : st.e) p = c.take<bf16_t>(E64);
    for (auto& p : st.z) p = c.take<bf16_t>(E64);
    for (auto& p : st.hb) p = c.take<bf16_t>(N64);
    for (auto& p : st.aggb) p = c.take<bf16_t>(N64);
    for (auto& p : st.u) p = c.take<bf16_t>(N64);
    st.dnh = c.take<bf16_t>(N64);
    st.deh = c.take<bf16_t>(E64);
    st.o = c.take<float>(rows * 16);
    st.gh = c.take<float>(slots * 9);
    bb.Z = c.take<float>(slots * 9);
    bb.dh = c.take<float>(N64);
    bb.de = c.take<float>(E64);
    bb.dob = c.take<bf16_t>(rows * 16);
    bb.ddnb = c.take<bf16_t>(N64);
    bb.dqb = c.take<bf16_t>(slots * 16);
    bb.ddeb = c.take<bf16_t>(E64);
    bb.dhb[0] = nullptr;
    for (int l = 1; l <= NLAYER; l++) bb.dhb[l] = c.take<bf16_t>(N64);
    for (auto& p : bb.dub) p = c.take<bf16_t>(N64);
    for (auto& p : bb.dmb) p = c.take<bf16_t>(E64);
    for (auto& p : bb.dzb) p = c.take<bf16_t>(E64);
    for (auto& p : bb.dPb) p = c.take<bf16_t>(N64);
    for (auto& p : bb.dQb) p = c.take<bf16_t>(N64);
    bb.dh0b = c.take<bf16_t>(N64);
    bb.de0b = c.take<bf16_t>(E64);
    part = c.take<double>(2 * (size_t)nblk);
    scal = c.take<float>(4);
    bb.scal = scal;
    return c.used;
}

struct SolBufs {
    float* bt;    // 3 per node row: b + bc
    float4* u;    // G^T G b
    float* p;     // A u
    float4* y;    // G b
    float4* w;    // G p
    float* palpha;  // nblk partials of b.u
    float* pdots;   // 2 nblk partials of u.p and b.x
    float* sc;      // c1, c2, L_sol
};

size_t carve_sol(char* base, size_t rows, int nblk, SolBufs& sb) {
    Carver c;
    c.base = base;
    sb.bt = c.take<float>(3 * rows);
    sb.u = c.take<float4>(rows);
    sb.p = c.take<float>(3 * rows);
    sb.y = c.take<float4>(rows);
    sb.w = c.take<float4>(rows);
    sb.palpha = c.take<float>(nblk);
    sb.pdots = c.take<float>(2 * (size_t)nblk);
    sb.sc = c.take<float>(4);
    return c.used;
}

struct SmoothBufs {
    float4 *v, *y, *s, *e, *w;  // probe, G u, G^T y, v - w s, G p
    float *u, *p;               // A v, A e (3 per node row)
    float* part;                // nblk partials of the G^T sweep (unused)
    float* pdots;               // 2 nblk partials of e.p and v.u
    float* sc;                  // c1, c2, L_sm
};

size_t carve_smooth(char* base, size_t rows, int nblk, SmoothBufs& m) {
    Carver c;
    c.base = base;
    m.v = c.take<float4>(rows);
    m.y = c.take<float4>(rows);
    m.s = c.take<float4>(rows);
    m.e = c.take<float4>(rows);
    m.w = c.take<float4>(rows);
    m.u = c.take<float>(3 * rows);
    m.p = c.take<float>(3 * rows);
    m.part = c.take<float>(nblk);
    m.pdots = c.take<float>(2 * (size_t)nblk);
    m.sc = c.take<float>(4);
    return c.used;
}

template <int KB, int MB>
void launch_wgrad(cs_ctx* C, const std::vector<WJob>& jobs, WJob* djobs, long maxRows) {
    if (jobs.empty()) return;
    const int chunk = 4096;
    hipLaunchKernelGGL((k_lpc_wgrad<KB, MB>), dim3(cdiv(maxRows, chunk), (unsigned)jobs.size()), dim3(256), 0,
                       C->stream, djobs, C->lpcNet.grad.ptr<float>(), chunk);
}

}  // namespace

float lpc_train_step(cs_ctx* C, int c, float lr, float solWeight, float smoothWeight, uint32_t seed,
                     float* gradOut) {
    LpcNet& N = C->lpcNet;
    if (!N.loaded) throw StatusError(CS_E_STATE, "training needs initial weights (cs_set_learned_weights)");
    lpc_alloc(C, c);
    ClothSys& S = C->sys[c];
    hipStream_t s = C->stream;
    const size_t slots = (size_t)S.totalSlots * SLICE;
    const size_t rows = (size_t)S.nslices * SLICE;
    const int nblk = cdiv(S.nn, 256);
    Stash st{};
    BackBufs bb{};
    double* part = nullptr;
    float* scal = nullptr;
    SolBufs sb{};
    SmoothBufs mb{};
    const size_t need1 = carve_all(nullptr, slots, rows, nblk, st, bb, part, scal);
    const size_t need2 = need1 + carve_sol(nullptr, rows, nblk, sb);
    const size_t need = need2 + carve_smooth(nullptr, rows, nblk, mb);
    char* base = N.arena.as<char>(need);
    carve_all(base, slots, rows, nblk, st, bb, part, scal);
    carve_sol(base + need1, rows, nblk, sb);
    carve_smooth(base + need2, rows, nblk, mb);
    lpc_forward<true>(C, c, st);
    Topo t = topo_of(S);
    LossArgs la{t, st.gh, st.Ah, st.o, bb.Z, part};
    hipLaunchKernelGGL(k_lpc_loss, dim3(nblk), dim3(256), 0, s, la);
    hipLaunchKernelGGL(k_lpc_loss_final, dim3(1), dim3(256), 0, s, nblk, 3 * S.nn, part, scal);
    if (S.solved && solWeight > 0.0f) {
        // the solution term: how well M^-1 b points at the solution in the energy norm
        const float lam = solWeight;
        LpcSys& L = S.lpc;
        const int* cCol = S.nContact > 0 ? S.cCol.ptr<int>() : S.cRowPtr.ptr<int>();
        const float* cVal = S.nContact > 0 ? S.cVal.ptr<float>() : S.vals.ptr<float>();
        hipLaunchKernelGGL(k_lpc_rhs, dim3(cdiv(3 * S.nn, 256)), dim3(256), 0, s, 3 * S.nn, S.b.ptr<float>(),
                           S.bc.ptr<float>(), sb.bt);
        hipLaunchKernelGGL(k_lpc_lower, dim3(nblk), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(), S.cols.ptr<int>(),
                           S.diagSlot.ptr<int>(), L.G.ptr<bf16_t>(), sb.bt, sb.y);
        hipLaunchKernelGGL(k_lpc_upper<UP_PCG>, dim3(nblk), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(),
                           S.cols.ptr<int>(), S.diagSlot.ptr<int>(), S.rowLen.ptr<int>(), L.G.ptr<bf16_t>(), sb.y,
                           sb.bt, nullptr, nullptr, sb.u, nullptr, sb.palpha);
        hipLaunchKernelGGL(k_lpc_pow_spmv, dim3(nblk), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(),
                           S.sliceW.ptr<int>(), S.cols.ptr<int>(), S.vals.ptr<float>(), S.cRowPtr.ptr<int>(), cCol,
                           cVal, sb.u, nullptr, 0, sb.p);
        hipLaunchKernelGGL(k_lpc_lower, dim3(nblk), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(), S.cols.ptr<int>(),
                           S.diagSlot.ptr<int>(), L.G.ptr<bf16_t>(), sb.p, sb.w);
        hipLaunchKernelGGL(k_lpc_sol_dots, dim3(nblk), dim3(256), 0, s, S.nn, sb.u, sb.p, sb.bt, S.x.ptr<float>(),
                           sb.pdots);
        hipLaunchKernelGGL(k_lpc_sol_final, dim3(1), dim3(256), 0, s, nblk, sb.palpha, sb.pdots, lam, scal, sb.sc);
        hipLaunchKernelGGL(k_lpc_sol_grad, dim3(nblk), dim3(256), 0, s, t, sb.y, sb.w, sb.bt, sb.p, L.S.ptr<float>(),
                           scal, sb.sc, lam, bb.Z);
        CS_LAUNCH_CHECK();
    }
    if (smoothWeight > 0.0f) {
        // the smoothing term, on the probe of `seed`, for one sweep at damping 1
        const float lam = smoothWeight, w = 1.0f;
        LpcSys& L = S.lpc;
        const int* cCol = S.nContact > 0 ? S.cCol.ptr<int>() : S.cRowPtr.ptr<int>();
        const float* cVal = S.nContact > 0 ? S.cVal.ptr<float>() : S.vals.ptr<float>();
        hipLaunchKernelGGL(k_lpc_probe, dim3(nblk), dim3(256), 0, s, S.nn, seed, mb.v);
        hipLaunchKernelGGL(k_lpc_pow_spmv, dim3(nblk), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(),
                           S.sliceW.ptr<int>(), S.cols.ptr<int>(), S.vals.ptr<float>(), S.cRowPtr.ptr<int>(), cCol,
                           cVal, mb.v, nullptr, 0, mb.u);
        hipLaunchKernelGGL(k_lpc_lower, dim3(nblk), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(), S.cols.ptr<int>(),
                           S.diagSlot.ptr<int>(), L.G.ptr<bf16_t>(), mb.u, mb.y);
        hipLaunchKernelGGL(k_lpc_upper<UP_PCG>, dim3(nblk), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(),
                           S.cols.ptr<int>(), S.diagSlot.ptr<int>(), S.rowLen.ptr<int>(), L.G.ptr<bf16_t>(), mb.y,
                           mb.u, nullptr, nullptr, mb.s, nullptr, mb.part);
        hipLaunchKernelGGL(k_lpc_sub, dim3(nblk), dim3(256), 0, s, S.nn, mb.v, mb.s, w, mb.e);
        hipLaunchKernelGGL(k_lpc_pow_spmv, dim3(nblk), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(),
                           S.sliceW.ptr<int>(), S.cols.ptr<int>(), S.vals.ptr<float>(), S.cRowPtr.ptr<int>(), cCol,
                           cVal, mb.e, nullptr, 0, mb.p);
        hipLaunchKernelGGL(k_lpc_lower, dim3(nblk), dim3(256), 0, s, S.nn, S.sliceOff.ptr<int>(), S.cols.ptr<int>(),
                           S.diagSlot.ptr<int>(), L.G.ptr<bf16_t>(), mb.p, mb.w);
        hipLaunchKernelGGL(k_lpc_smooth_dots, dim3(nblk), dim3(256), 0, s, S.nn, mb.e, mb.p, mb.v, mb.u, mb.pdots);
        hipLaunchKernelGGL(k_lpc_smooth_final, dim3(1), dim3(256), 0, s, nblk, mb.pdots, lam, w, scal, mb.sc);
        hipLaunchKernelGGL(k_lpc_sol_grad, dim3(nblk), dim3(256), 0, s, t, mb.y, mb.w, mb.u, mb.p, L.S.ptr<float>(),
                           scal, mb.sc, lam, bb.Z);
        CS_LAUNCH_CHECK();
    }
    N.grad.as<float>(W_COUNT);
    fill_f32(N.grad.ptr<float>(), 0.0f, W_COUNT, s);
    BackArgs ba;
    ba.t = t;
    ba.layer = 0;
    ba.w = N.w.ptr<float>();
    ba.wt = reinterpret_cast<const s4*>(N.fragT.ptr<bf16_t>());
    ba.st = st;
    ba.bb = bb;
    hipLaunchKernelGGL(k_lpc_bdec, dim3(S.nslices), dim3(LB), 0, s, ba);
    for (int l = NLAYER - 1; l >= 0; l--) {
        ba.layer = l;
        hipLaunchKernelGGL(k_lpc_blayer, dim3(S.nslices), dim3(LB), 0, s, ba);
    }
    hipLaunchKernelGGL(k_lpc_bfinal, dim3(S.nslices), dim3(LB), 0, s, ba);
    CS_LAUNCH_CHECK();
    // weight gradients
    const long R = (long)rows, E = (long)slots;
    std::vector<WJob> j14, j44, j84, j41;
    j14.push_back({st.xn, nullptr, bb.dh0b, R, O_WN, O_BN});
    j14.push_back({st.xe, nullptr, bb.de0b, E, O_WE, O_BE});
    for (int l = 0; l < NLAYER; l++) {
        j44.push_back({st.hb[l], nullptr, bb.dPb[l], R, lw(l, L_WA), -1});
        j44.push_back({st.hb[l], nullptr, bb.dQb[l], R, lw(l, L_WB), -1});
        j44.push_back({st.e[l], nullptr, bb.dzb[l], E, lw(l, L_WC), lw(l, L_BC)});
        j44.push_back({st.z[l], nullptr, bb.dmb[l], E, lw(l, L_W2), lw(l, L_B2)});
        j44.push_back({st.u[l], nullptr, bb.dhb[l + 1], R, lw(l, L_U2), lw(l, L_BU2)});
        j84.push_back({st.hb[l], st.aggb[l], bb.dub[l], R, lw(l, L_U1), lw(l, L_BU1)});
    }
    j44.push_back({st.hb[NLAYER], nullptr, bb.ddnb, R, O_DN1, O_BDN1});
    j44.push_back({st.e[NLAYER], nullptr, bb.ddeb, E, O_DE1, O_BDE1});
    j41.push_back({st.dnh, nullptr, bb.dob, R, O_DN2, O_BDN2});
    j41.push_back({st.deh, nullptr, bb.dqb, E, O_DE2, O_BDE2});
    std::vector<WJob> all;
    for (auto* v : {&j14, &j44, &j84, &j41}) all.insert(all.end(), v->begin(), v->end());
    WJob* dj = N.jobs.as<WJob>(all.size());
    CS_HIP(hipMemcpyAsync(dj, all.data(), sizeof(WJob) * all.size(), hipMemcpyHostToDevice, s));
    const long maxRows = std::max(R, E);
    launch_wgrad<1, 4>(C, j14, dj, maxRows);
    launch_wgrad<4, 4>(C, j44, dj + j14.size(), maxRows);
    launch_wgrad<8, 4>(C, j84, dj + j14.size() + j44.size(), maxRows);
    launch_wgrad<4, 1>(C, j41, dj + j14.size() + j44.size() + j84.size(), maxRows);
    CS_LAUNCH_CHECK();
    float hs[4];
    CS_HIP(hipMemcpyAsync(hs, scal, sizeof(hs), hipMemcpyDeviceToHost, s));
    if (gradOut) CS_HIP(hipMemcpyAsync(gradOut, N.grad.p, sizeof(float) * W_COUNT, hipMemcpyDeviceToHost, s));
    if (lr > 0.0f) {
        N.t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - powf(b1, (float)N.t), c2 = 1.0f - powf(b2, (float)N.t);
        hipLaunchKernelGGL(k_lpc_adam, dim3(cdiv(W_COUNT, 256)), dim3(256), 0, s, W_COUNT, N.w.ptr<float>(),
                           N.grad.ptr<float>(), N.m.ptr<float>(), N.v.ptr<float>(), lr, b1, b2, 1e-8f, c1, c2);
        lpc_refresh_fragments(C);
    }
    CS_HIP(hipStreamSynchronize(s));
    return hs[0];
}

}  // namespace cs
