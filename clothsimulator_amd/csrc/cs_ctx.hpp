// cs_ctx.hpp — the context behind the opaque cs_ctx handle: scene topology,
// device-resident state (HBM layout) and the per-stage working buffers.
//
// HBM layout (all meshes concatenated: cloths first, then obstacles):
//   nodes  x, x0, x1, v, n         float4 each (16 B, one dwordx4 per gather)
//          area, mass              float
//          (free <=> global id < nClothNodes)
//   faces  fnodes int4 (3 node ids + mesh id), fedges int4, fnrm float4,
//          finv float4 (2x2 rest inverse), farea, fmass float
//   edges  enodes int2 (min,max), eopp int2 (opposite node or -1), eadj int2 (face or -1)
// Per cloth, the static implicit-Euler pattern is a sliced-ELL BSR: slices of
// 64 rows (one wavefront), each slice padded to its longest row; within a slice
// block slot k of value c for row r sits at vals[off + (k*9+c)*64 + r%64], so a
// wave reads and writes 256 contiguous bytes per (k, c).  Padding slots hold
// col = row and zero values.
#pragma once
#include <string>
#include <vector>

#include "../../include/clothsim.h"
#include "cs_common.hpp"
#include "cs_motion.hpp"
#include "cs_prims.hpp"
#include "cs_remesh.hpp"

namespace cs {

static const int MAT_N = 30;
static const int SLICE = 64;
struct HostMaterial {
    float density = 0.0f;
    std::vector<float> table;  // 30^3 x 4
    float bend[15];
};

struct MeshInfo {
    bool cloth;
    int n0, nn, f0, nf, e0, ne, v0, nv;
    int material;
    int motion = -1;  // obstacles: the motion moving it (-1: static)
};

// Per-mesh LBVH over faces.  Unified node array: internal i at i (i < nf-1),
// leaf j at nf-1+j.  BNode is 32 B: lo[3], maxIndex, hi[3], face (-1 internal).
struct BNode {
    float lo[3];
    int maxIndex;
    float hi[3];
    int face;
};

struct Bvh {
    int mesh = -1;
    int nf = 0;
    bool ccd = false;
    DBuf keys, keysTmp;    // uint64 (morton) keys
    DBuf vals, valsTmp;    // uint32 local face ids
    DBuf nodes;            // BNode[2nf-1]
    DBuf child;            // int2[nf-1]
    DBuf range;            // int2[nf-1]: first and last sorted leaf under each internal node
    DBuf leafTab;          // LBox[5][nf]: union of sorted leaves [j, j + 2^k), k = 1..5
    DBuf blockTab;         // LBox[levels][nblk]: union of 64-leaf blocks [i, i + 2^k)
    DBuf fbox;             // float4 x2 per face (unsorted leaf boxes)
    DBuf partial;          // block partial boxes
    DBuf scene;            // float4 x2 (p, d)
};

// One level of the aggregation multigrid preconditioner.  Level 0 is the
// cloth's own system (sliced ELL + contacts, block-Jacobi Minv); levels >= 1
// hold block-CSR Galerkin operators of the static system.
struct AmgLevel {
    int n = 0, nnzb = 0;
    DBuf rp, ci, diag;   // (l >= 1) block CSR, sorted columns; diagonal block index per row
    DBuf val, Dinv;      // (l >= 1) float 9 per block / per node
    DBuf cptr, cidx;     // (l >= 1) Galerkin contributors of each block: value ids in the finer level
    DBuf agg;            // (not coarsest) node -> node of the next level
    DBuf mptr, midx;     // (not coarsest) node of the next level -> its nodes here (ascending)
    DBuf b, x, t;        // float 3n: right-hand side, smoothed iterate, residual / result
};

struct AmgSys {
    bool built = false;
    int nl = 0;
    std::vector<AmgLevel> L;
    DBuf work;  // double 2 m^2: Gauss-Jordan of the coarsest operator (m = 3 n of the last level)
    DBuf inv;   // float m^2: its inverse
    // The small levels' sub-cycle (levels l0.. of at most 512 nodes, from the restricted
    // residual of level l0 to level l0's result) as one dense matrix, rebuilt per solve:
    // row-major, dim = 3 n(l0) rows and columns
    DBuf smallOp;
    int smallDim = 0, smallL0 = 0;
    // A_static P of level 0 in sliced ELL (row i, one slot per distinct aggregate of
    // its columns): apTarget[level-0 slot] = the AP slot it sums into (255: none)
    DBuf apTarget, apSliceOff, apSliceW, apCols, apVals;
};

// Learned preconditioner of one cloth: M^-1 = G^T G with G block lower
// triangular on the static pattern, stored in bfloat16 in the value layout of
// the sliced ELL (9 per slot): slot (i, j < i) holds G_ij, the diagonal
// slot G_ii and slot (i, j > i) the transposed block G_ji^T, so both sweeps
// only gather.  Network activations in 64-feature rows (see cs_lpc.hip).
struct LpcSys {
    bool built = false;
    DBuf S;       // float 9 per node: S_i = chol(Minv_i)^T, row-major
    DBuf mirror;  // int per slot: the slot of the transposed block (-1: padding or diagonal)
    DBuf G;       // bf16 9 per slot, value layout
    DBuf y;       // float4 per node: G r
    DBuf h, P, Q, P2, Q2;  // float 64 per node row (rows padded to whole slices)
    DBuf e;       // bf16 64 per slot: edge states
    // learned smoother of the multigrid cycle: power-iteration vectors and partials, damping
    DBuf pv, pw, ppart, omega;
    unsigned omegaAsm = ~0u, omegaNet = ~0u;  // assembly and weights the current omega was estimated for
};

struct ClothSys {
    int mesh;
    unsigned asmEpoch = 0;  // incremented by every assembly of the system
    int nn, nslices, totalSlots, maxWidth;
    // static sliced-ELL pattern
    DBuf sliceOff;  // int[nslices+1] (in slots)
    DBuf sliceW;    // int[nslices]
    DBuf cols;      // int[totalSlots*64]
    DBuf vals;      // float[totalSlots*64*9]
    DBuf rowLen;    // int[nn]
    DBuf diagSlot;  // int[nn]
    DBuf incPtr;    // int[nn+1]
    DBuf inc;       // int2 per incidence: x = element (face id, or edge id | 1<<31), y = a | slots<<2 (6 bits each)
    DBuf hcount;    // float[nn]  handles on the node
    DBuf hpos;      // float4[nn] sum of handle anchors
    // rows with a moving handle: all of the cloth's handled rows, re-anchored every step
    int nHandleRows = 0;
    DBuf hgPtr, hgRow;  // int[rows+1] into hAnchor, int[rows] row
    DBuf hAnchor;       // float4 per handle: anchor, motion index bits in w (-1: fixed)
    DBuf b;         // float[3nn] static right-hand side
    DBuf bc;        // float[3nn] contact right-hand side
    DBuf Minv;      // float[9nn]
    // dynamic contact blocks (CSR over rows)
    int nContact = 0;
    DBuf cRowPtr, cCol, cVal, cDiag;
    // PCG vectors: x, r, q float[3nn]; z float4[nn]; d float4[2nn] (two generations)
    DBuf x, r, z, d, q;
    // host copies used by cs_assemble
    std::vector<int> hRowPtr, hCol;  // exact (unpadded) static pattern
    std::vector<int> hSliceOff, hSliceW, hRowLen;
    int nHandles = 0;
    bool solved = false;  // x holds the converged solution of the current assembled system
    AmgSys amg;
    LpcSys lpc;
};

// Weights of the preconditioner network: float master copy and bfloat16
// MFMA operand fragments (forward and transposed), each indexed like the flat
// weight vector (cs_lpc.hip).
struct LpcNet {
    bool loaded = false;
    DBuf w, frag, fragT;
    // training: gradient, Adam moments and step count, activation arena, job table
    DBuf grad, m, v, arena, jobs;
    int t = 0;
    unsigned version = 0;  // incremented whenever the weights change
};

// A face as it was before remeshing: material-space and world positions of its corners
// (BackupFace, src/BackupFace.cuh; Mesh::backupFaces, src/Mesh.cu:116-128).
struct BackupFace {
    float u[3][2];
    float x[3][3];
};

// One separation constraint: sum_j w_j n . x_j >= thickness over the two faces' six nodes, w the
// barycentric weights of the first face's point and minus those of the second's.
struct SepRec {
    int nodes[6];
    float w[6];
    float n[3];
    float pad;
};

struct PcgScalars {
    float rz, dq, alpha, beta, rr, bb, tol2bb, pad;
    int iters, done, maxit, pad2;
};

// Sampled HIP-event timing of hot kernels (enabled by CS_FLAG_TIME_KERNELS).
struct KernelTimer {
    std::vector<hipEvent_t> events;              // pairs (start, stop)
    std::vector<std::pair<int, int>> pending;    // (kernel id, pair index) awaiting collection
    double total[CS_KERNEL_COUNT] = {};
    int count[CS_KERNEL_COUNT] = {};
};

}  // namespace cs

struct cs_ctx {
    cs_params P;
    hipStream_t stream = nullptr;
    // a second stream (default priority) for per-solve setup that one workgroup runs while the
    // main stream is busy with other setup: forked from and joined to `stream` by events
    hipStream_t side = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    std::string err;
    bool finalized = false;
    int device = 0;

    // host staging (kept for topology and I/O)
    std::vector<cs::MeshInfo> meshes;
    int nCloths = 0, nObstacles = 0;
    std::vector<float> hx, hv, huv;  // 3N, 3N, 2V
    std::vector<int> hvnode;         // uv vertex -> node
    std::vector<int> hfv, hfn;       // 3F uv-vertex ids, node ids
    std::vector<int> hfe;            // 3F edge ids
    std::vector<int> hen, heopp, headj;  // 2E each (eopp: uv vertex id)
    std::vector<cs::HostMaterial> materials;
    struct HandleH { int cloth, node; float p[3]; int motion; };
    std::vector<HandleH> handles;

    int N = 0, F = 0, E = 0, NC = 0, FC = 0, EC = 0;  // totals and cloth totals

    // motions (cs_motion.hip): keyframes, the step counter they are evaluated at, and the
    // obstacles' base shapes (host: 3 floats per obstacle node, set by cs_set_obstacle_motion)
    std::vector<cs::Motion> motions;
    int steps = 0;
    std::vector<float> hObsBase;
    bool anyMovingObstacle = false;
    cs::DBuf obsBase, obsMotion;  // per obstacle node: float4 base, int motion (-1: static)

    // device state
    cs::DBuf x, x0, x1, v, nrm, area, mass;  // nodes
    cs::DBuf fnodes, fedges, fnrm, finv, farea, fmass;  // faces
    cs::DBuf enodes, eopp, eadj, ebias;  // edges (ebias: float4 (id0, w0, id1, w1) per cloth edge)
    cs::DBuf nfPtr, nfIdx;  // node -> incident faces (face*4 + local), faces ascending
    cs::DBuf matTable, matBend;  // all materials concatenated
    cs::DBuf fmat;  // int per cloth face: material index
    cs::DBuf faceRec, edgeRec;   // per-step element records (cloth faces / edges)
    std::vector<cs::ClothSys> sys;
    std::vector<cs::Bvh> bvh;     // one per mesh (rebuilt per stage)
    cs::DBuf bvhJobs, bvhSortTmp, travJobs; // job tables of the batched tree builds and traversals
    cs::LpcNet lpcNet;            // learned preconditioner weights (CS_PRECOND_LEARNED)

    // proximity
    cs::DBuf pairs, pairCount;           // int2 pairs, uint32 counter(s)
    size_t pairCap = 0;
    cs::DBuf bestNode, bestEdge, bestFace;  // uint64 keys
    cs::DBuf proxFlag, proxPos, prox;     // uint32, uint32, cs_proximity_rec
    int nProx = 0;
    // contact assembly scratch
    cs::DBuf cKeys, cKeysTmp, cIdx, cIdxTmp, cRec, cFlag, cPos;
    // PCG
    cs::DBuf pcgPartial, pcgScal, pcgCounter;
    cs::PcgScalars* hScal = nullptr;  // pinned
    // batched solve of independent cloths: per-cloth table, block -> cloth map, partials, scalars
    cs::DBuf bTab, bMap, bPart, bScal, bList;
    struct BatchTune {
        int trials = 0, choice = 0;    // solves tried so far, best candidate
        double score[4] = {0, 0, 0, 0};  // seconds per node-iteration, summed over trials
        int rows = 256, chunk = 16;      // the setting in use
    } batchTune;
    void* hBatch = nullptr;  // pinned readback of per-cloth data (scalars, contact bounds)
    size_t hBatchCap = 0;    // bytes
    // asynchronous frame download (cs_frame_begin / cs_frame_end): two slots, each a device
    // snapshot of every cloth node's x then v (float4) and its pinned host copy, downloaded on
    // the copy stream
    struct FrameSlot {
        cs::DBuf stage;
        float4* host = nullptr;
        size_t cap = 0;              // nodes the pinned copy holds
        hipEvent_t done = nullptr;   // recorded on the copy stream after the download
        bool begun = false;
        int total = 0;               // cloth nodes in the snapshot
        std::vector<int> n0, nn;     // per cloth at the snapshot
    } frame[2];
    hipStream_t copy = nullptr;
    hipEvent_t evSnap = nullptr;
    // batched multigrid of the independent cloths (cs_amg.hip): every cloth keeps its own
    // hierarchy; one launch covers level l of every cloth that has one
    struct BatchAmg {
        int ncl = -1;                  // cloth count the maps were built for (-1: not built)
        int nl = 0;                    // most levels of any cloth
        std::vector<int> usable;       // per cloth: 1 if its hierarchy is used
        std::vector<int> jobOff;       // first job of level l (l >= 1), jobOff[nl] = job count
        struct Seg { int off = 0, n = 0; };
        std::vector<Seg> rows, sparse, blocks;  // per level: block-map segments (all jobs / non-coarsest / Galerkin blocks)
        Seg ap, coarse;                // level-0 A P of every cloth; coarsest job of every cloth
        cs::DBuf jobs, apJobs, maps;   // BLev[jobOff[nl]], BAp[usable], int2 map entries
        // the cycle's maps restricted to the cloths still iterating (bamg_select)
        std::vector<int> jobCloth;     // cloth of each job
        std::vector<int2> hmap, hsel;  // host copies: every map, the selected cycle maps
        std::vector<Seg> crows, csparse;
        Seg ccoarse;
        bool selected = false;         // the cycle uses the selected maps
        cs::DBuf cmaps;
    } bamg;
    cs::DBuf asmJobs;        // per-cloth table of the batched assembly
    cs::DBuf gcCol, gcVal, gcBc, gcDiag, gcU0;  // batch-wide contact reduction
    // CCD / impacts
    cs::DBuf impacts, impCount, impKeys, impKeysTmp, impIdx, impIdxTmp, impSorted;
    // per face, from its mesh's last-built LBVH: float4 (leaf box lo, mesh id bits),
    // float4 (leaf box hi, sorted leaf index bits)
    cs::DBuf faceBox;
    cs::DBuf ccdCand;   // uint64 candidate keys of the CCD prefilter (one per distinct test)
    cs::DBuf ccdMask, ccdBlk;  // per pair survivor masks; per block counts and their scans
    cs::DBuf impAcc;  // accumulated impacts of the deform pass
    int nImpAcc = 0;
    cs::DBuf nodeMin, nodeRemoved, impState;
    // response (optimization) scratch
    cs::DBuf optKeys, optKeysTmp, optVals, optValsTmp, optFlag, optPos, optNodes, optStart, optSlot,
        optX, optNext, optGrad, optLambda, optCoef, optScal;
    uint32_t* hCount = nullptr;  // pinned small readback
    cs::Scratch ws;
    cs::KernelTimer kt;
    // remeshing: per cloth the "remeshing" block (has = 0: the cloth is never remeshed), per node
    // the preserve flag (boundary, seam and handle nodes of the loaded mesh; set on the first
    // remesh), the size of the last flip pass (shared by every cloth, as the reference's static)
    struct RemeshCfg { bool has = false; cs::RemeshParams r; };
    std::vector<RemeshCfg> remeshCfg;
    std::vector<unsigned char> hpres;
    bool presReady = false;
    int flipCount = 0;
    cs::DBuf rmUv, rmPlanes, rmSizing;
    float remeshMs[4] = {0.0f, 0.0f, 0.0f, 0.0f};  // last cs_remesh: sizing, operators, rebuild, separation
    // separation step after remeshing (cs_separation.hip)
    cs::DBuf sepFlag, sepPos, sepRaw, sepRec, sepCons;
    int sepIters = 0;  // solves of the last separation step
};
