// cs_host.cpp — host-side set-up: material tables, mesh topology, the static
// sliced-ELL pattern with its element->slot maps, and the upload to HBM.
//
// Topology follows Mesh::load (src/Mesh.cu:353-452, src/MeshHelper.cu:34-102):
// edges are the sorted unique (min,max) node pairs of a mesh, and each face
// side initialises the edge side that matches its orientation
// (Edge::initialize, src/Edge.cu:13-31).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cs_ctx.hpp"
#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

// ---------------------------------------------------------------- material
// Stretching samples from a 2x2 eigen-decomposition of 2G, with strain/angle
// bilinear weights over the 2x5 measured rows, doubled (src/Material.cu:45-85,
// src/MathHelper.cu:3-36); bending is the 3x5 table times its multiplier.
static void eig2(float a, float b, float d, float Q[2][2], float l[2]) {
    float amd = a - d, apd = a + d, b2 = b * b;
    float det = std::sqrt(4.0f * b2 + amd * amd);
    l[0] = 0.5f * (apd + det);
    l[1] = 0.5f * (apd - det);
    if (b != 0.0f) {
        float v0 = l[0] - d, vn = std::sqrt(v0 * v0 + b2);
        Q[0][0] = v0 / vn;
        Q[1][0] = b / vn;
        v0 = l[1] - d;
        vn = std::sqrt(v0 * v0 + b2);
        Q[0][1] = v0 / vn;
        Q[1][1] = b / vn;
    } else if (a >= d) {
        Q[0][0] = 1.0f; Q[1][0] = 0.0f; Q[0][1] = 0.0f; Q[1][1] = 1.0f;
    } else {
        Q[0][0] = 0.0f; Q[1][0] = 1.0f; Q[0][1] = 1.0f; Q[1][1] = 0.0f;
    }
}

void material_build(const cs_material_in& in, HostMaterial& m) {
    const float th = in.thicken;
    m.density = in.density * (in.density_mult * th);
    const float smul = in.stretching_mult * th;
    const float bmul = in.bending_mult * th;
    float rows[2][5][4];
    for (int j = 0; j < 5; j++)
        for (int c = 0; c < 4; c++) {
            rows[0][j][c] = in.stretching[0][c];
            rows[1][j][c] = in.stretching[j + 1][c];
        }
    m.table.assign(MAT_N * MAT_N * MAT_N * 4, 0.0f);
    for (int i = 0; i < MAT_N; i++)
        for (int j = 0; j < MAT_N; j++)
            for (int k = 0; k < MAT_N; k++) {
                float g00 = static_cast<float>(i) / MAT_N - 0.25f;
                float g11 = static_cast<float>(j) / MAT_N - 0.25f;
                float g01 = static_cast<float>(k) / MAT_N;
                float Q[2][2], l[2];
                eig2(2.0f * g00, 2.0f * g01, 2.0f * g11, Q, l);
                if (l[0] == l[1] && Q[0][0] == 0.0f && Q[1][1] == 0.0f) {
                    Q[0][0] = 1.0f; Q[0][1] = 0.0f; Q[1][0] = 0.0f; Q[1][1] = 1.0f;
                }
                if (Q[0][0] < 0.0f) { Q[0][0] = -Q[0][0]; Q[1][0] = -Q[1][0]; }
                if (Q[1][1] < 0.0f) { Q[0][1] = -Q[0][1]; Q[1][1] = -Q[1][1]; }
                float sw = (std::sqrt(l[0] + 1.0f) - 1.0f) * 6.0f;
                sw = tclamp(sw, 0.0f, 1.0f - 1e-6f);
                int si = static_cast<int>(sw);
                sw -= si;
                float aw = static_cast<float>(static_cast<double>(tabs(std::atan2(Q[1][0], Q[0][0]))) / M_PI * 8.0f);
                aw = tclamp(aw, 0.0f, 4.0f - 1e-6f);
                int ai = static_cast<int>(aw);
                aw -= ai;
                const float w00 = (1.0f - sw) * (1.0f - aw), w01 = (1.0f - sw) * aw;
                const float w10 = sw * (1.0f - aw), w11 = sw * aw;
                float* out = &m.table[(((i * MAT_N) + j) * MAT_N + k) * 4];
                for (int c = 0; c < 4; c++) {
                    float s = 0.0f;
                    s += rows[si][ai][c] * w00;
                    s += rows[si][ai + 1][c] * w01;
                    s += rows[si + 1][ai][c] * w10;
                    s += rows[si + 1][ai + 1][c] * w11;
                    out[c] = tmax(s, 0.0f) * 2.0f * smul;
                }
            }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 5; j++) m.bend[5 * i + j] = in.bending[i][j] * bmul;
}

// Bias-angle interpolation of the bending table for one hinge side: depends on
// the material-space edge direction only, so it is fixed at set-up
// (src/Material.cu:130-137).
static void bending_bias(float d0, float d1, float& id, float& w) {
    float bw = static_cast<float>(static_cast<double>(tabs(std::atan2(d1, d0)) * 4.0f) / M_PI);
    if (bw > 2.0f) bw = 4.0f - bw;
    bw = tclamp(bw, 0.0f, 2.0f - 1e-6f);
    int bi = static_cast<int>(bw);
    id = static_cast<float>(bi);
    w = bw - bi;
}

// ---------------------------------------------------------------- topology
int host_add_mesh(cs_ctx* C, const cs_mesh_in& in, bool cloth, int material) {
    if (in.n_nodes <= 0 || in.n_faces <= 0 || in.n_verts <= 0 || !in.x || !in.uv || !in.face_nodes || !in.face_verts)
        return CS_E_ARG;
    for (long k = 0; k < 3L * in.n_faces; k++) {
        if (in.face_nodes[k] < 0 || in.face_nodes[k] >= in.n_nodes) return CS_E_FORMAT;
        if (in.face_verts[k] < 0 || in.face_verts[k] >= in.n_verts) return CS_E_FORMAT;
    }
    MeshInfo M;
    M.cloth = cloth;
    M.n0 = (int)(C->hx.size() / 3);
    M.nn = in.n_nodes;
    M.v0 = (int)(C->huv.size() / 2);
    M.nv = in.n_verts;
    M.f0 = (int)(C->hfn.size() / 3);
    M.nf = in.n_faces;
    M.e0 = (int)(C->hen.size() / 2);
    M.material = material;
    C->hx.insert(C->hx.end(), in.x, in.x + 3L * in.n_nodes);
    if (in.v) C->hv.insert(C->hv.end(), in.v, in.v + 3L * in.n_nodes);
    else C->hv.insert(C->hv.end(), 3L * in.n_nodes, 0.0f);
    C->huv.insert(C->huv.end(), in.uv, in.uv + 2L * in.n_verts);
    C->hvnode.insert(C->hvnode.end(), in.n_verts, -1);
    for (long k = 0; k < 3L * in.n_faces; k++) {
        int nd = M.n0 + in.face_nodes[k], vt = M.v0 + in.face_verts[k];
        C->hfn.push_back(nd);
        C->hfv.push_back(vt);
        C->hfe.push_back(-1);
        C->hvnode[vt] = nd;
    }
    // edge occurrences keyed by (min,max) node pair, in face order; stable sort
    struct Occ { int a, b, opp, face; };
    std::vector<Occ> occ((size_t)3 * in.n_faces);
    for (int f = 0; f < in.n_faces; f++) {
        int fg = M.f0 + f;
        for (int k = 0; k < 3; k++) {
            int p = C->hfn[3 * fg + k], q = C->hfn[3 * fg + (k + 1) % 3];
            occ[3 * f + k] = {std::min(p, q), std::max(p, q), C->hfv[3 * fg + (k + 2) % 3], fg};
        }
    }
    std::stable_sort(occ.begin(), occ.end(), [](const Occ& p, const Occ& q) { return p.a < q.a || (p.a == q.a && p.b < q.b); });
    for (size_t k = 0; k < occ.size(); k++) {
        if (k == 0 || occ[k].a != occ[k - 1].a || occ[k].b != occ[k - 1].b) {
            C->hen.push_back(occ[k].a); C->hen.push_back(occ[k].b);
            for (int s = 0; s < 2; s++) { C->heopp.push_back(-1); C->headj.push_back(-1); }
        }
        int e = (int)C->hen.size() / 2 - 1;
        int f = occ[k].face;
        for (int i = 0; i < 3; i++) {
            int va = C->hfv[3 * f + i], vb = C->hfv[3 * f + (i + 1) % 3];
            int na = C->hvnode[va], nb = C->hvnode[vb];
            if (C->hen[2 * e] == na && C->hen[2 * e + 1] == nb) {
                C->heopp[2 * e] = occ[k].opp; C->headj[2 * e] = f; break;
            } else if (C->hen[2 * e + 1] == na && C->hen[2 * e] == nb) {
                C->heopp[2 * e + 1] = occ[k].opp; C->headj[2 * e + 1] = f; break;
            }
        }
        for (int i = 0; i < 3; i++) {  // Face::setEdge (src/Face.cu:22-33)
            int n0 = C->hfn[3 * f + i], n1 = C->hfn[3 * f + (i + 1) % 3];
            if ((n0 == C->hen[2 * e] && n1 == C->hen[2 * e + 1]) || (n1 == C->hen[2 * e] && n0 == C->hen[2 * e + 1])) {
                C->hfe[3 * f + i] = e; break;
            }
        }
    }
    M.ne = (int)C->hen.size() / 2 - M.e0;
    C->meshes.push_back(M);
    if (cloth) C->nCloths++; else C->nObstacles++;
    return CS_OK;
}

// uv vertex pair (v0, v1) of edge e on side s: the face's vertices ordered so
// that node(v0) = en[0], node(v1) = en[1] (Edge::vertices, src/Edge.cu:13-31).
static void edge_side_verts(const cs_ctx* C, int e, int s, int& v0, int& v1) {
    int f = C->headj[2 * e + s];
    v0 = v1 = -1;
    for (int i = 0; i < 3; i++) {
        int va = C->hfv[3 * f + i], vb = C->hfv[3 * f + (i + 1) % 3];
        int na = C->hvnode[va], nb = C->hvnode[vb];
        if (s == 0 && na == C->hen[2 * e] && nb == C->hen[2 * e + 1]) { v0 = va; v1 = vb; return; }
        if (s == 1 && nb == C->hen[2 * e] && na == C->hen[2 * e + 1]) { v0 = vb; v1 = va; return; }
    }
}

template <class T> static T* upload(DBuf& b, const std::vector<T>& h, hipStream_t s) {
    T* p = b.as<T>(h.size() ? h.size() : 1);
    if (!h.empty()) CS_HIP(hipMemcpyAsync(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
    return p;
}

// Builds the static block pattern of one cloth, its sliced-ELL layout and the
// per-row incidence lists that the assembly kernel walks.
static void build_cloth_system(cs_ctx* C, int c, ClothSys& S) {
    const MeshInfo& M = C->meshes[c];
    S.mesh = c;
    S.nn = M.nn;
    std::vector<std::vector<int>> adj(M.nn);
    auto link = [&](int a, int b) { adj[a - M.n0].push_back(b - M.n0); };
    for (int i = M.n0; i < M.n0 + M.nn; i++) link(i, i);
    for (int f = M.f0; f < M.f0 + M.nf; f++)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) link(C->hfn[3 * f + a], C->hfn[3 * f + b]);
    auto hinge = [&](int e, int st[4]) {
        st[0] = C->hen[2 * e]; st[1] = C->hen[2 * e + 1];
        st[2] = C->hvnode[C->heopp[2 * e]]; st[3] = C->hvnode[C->heopp[2 * e + 1]];
    };
    auto interior = [&](int e) { return C->headj[2 * e] >= 0 && C->headj[2 * e + 1] >= 0; };
    for (int e = M.e0; e < M.e0 + M.ne; e++) {
        if (!interior(e)) continue;
        int st[4];
        hinge(e, st);
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) link(st[a], st[b]);
    }
    S.hRowPtr.assign(M.nn + 1, 0);
    S.hRowLen.assign(M.nn, 0);
    S.hCol.clear();
    for (int i = 0; i < M.nn; i++) {
        std::vector<int>& r = adj[i];
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
        if (r.size() > 63) throw StatusError(CS_E_CAPACITY, "node valence too high for the 63-block row limit");
        S.hRowLen[i] = (int)r.size();
        S.hRowPtr[i + 1] = S.hRowPtr[i] + (int)r.size();
        S.hCol.insert(S.hCol.end(), r.begin(), r.end());
    }
    S.nslices = cdiv(M.nn, SLICE);
    S.hSliceOff.assign(S.nslices + 1, 0);
    S.hSliceW.assign(S.nslices, 0);
    S.maxWidth = 0;
    for (int s = 0; s < S.nslices; s++) {
        int w = 0;
        for (int r = s * SLICE; r < std::min(M.nn, (s + 1) * SLICE); r++) w = std::max(w, S.hRowLen[r]);
        S.hSliceW[s] = w;
        S.hSliceOff[s + 1] = S.hSliceOff[s] + w;
        S.maxWidth = std::max(S.maxWidth, w);
    }
    S.totalSlots = S.hSliceOff[S.nslices];
    std::vector<int> cols((size_t)S.totalSlots * SLICE, 0), diag(M.nn, 0);
    for (int s = 0; s < S.nslices; s++)
        for (int l = 0; l < SLICE; l++) {
            int r = s * SLICE + l;
            for (int k = 0; k < S.hSliceW[s]; k++) {
                int col = 0;
                if (r < M.nn) col = k < S.hRowLen[r] ? adj[r][k] : r;
                cols[((size_t)S.hSliceOff[s] + k) * SLICE + l] = col;
            }
            if (r < M.nn)
                for (int k = 0; k < S.hRowLen[r]; k++)
                    if (adj[r][k] == r) diag[r] = k;
        }
    auto slotOf = [&](int row, int colLocal) {
        const std::vector<int>& r = adj[row];
        return (int)(std::lower_bound(r.begin(), r.end(), colLocal) - r.begin());
    };
    // incidences: faces (ascending) then interior edges (ascending)
    std::vector<std::vector<std::pair<int, int>>> inc(M.nn);
    for (int f = M.f0; f < M.f0 + M.nf; f++)
        for (int a = 0; a < 3; a++) {
            int row = C->hfn[3 * f + a] - M.n0;
            int y = a;
            for (int b = 0; b < 3; b++) y |= slotOf(row, C->hfn[3 * f + b] - M.n0) << (2 + 6 * b);
            inc[row].push_back({f, y});
        }
    for (int e = M.e0; e < M.e0 + M.ne; e++) {
        if (!interior(e)) continue;
        int st[4];
        hinge(e, st);
        for (int a = 0; a < 4; a++) {
            int row = st[a] - M.n0;
            int y = a;
            for (int b = 0; b < 4; b++) y |= slotOf(row, st[b] - M.n0) << (2 + 6 * b);
            inc[row].push_back({e | (1 << 31), y});
        }
    }
    std::vector<int> incPtr(M.nn + 1, 0), incFlat;
    for (int i = 0; i < M.nn; i++) {
        incPtr[i + 1] = incPtr[i] + (int)inc[i].size();
        for (auto& p : inc[i]) { incFlat.push_back(p.first); incFlat.push_back(p.second); }
    }
    std::vector<float> hcount(M.nn, 0.0f), hpos(4 * (size_t)M.nn, 0.0f);
    S.nHandles = 0;
    bool moving = false;
    std::vector<std::vector<int>> byRow(M.nn);  // handles of each row, in insertion order
    for (size_t k = 0; k < C->handles.size(); k++) {
        const auto& h = C->handles[k];
        if (h.cloth != c) continue;
        int i = h.node - M.n0;
        hcount[i] += 1.0f;
        for (int a = 0; a < 3; a++) hpos[4 * i + a] += h.p[a];
        S.nHandles++;
        moving |= h.motion >= 0;
        byRow[i].push_back((int)k);
    }
    S.nHandleRows = 0;
    if (moving) {  // every handled row is re-anchored each step (src/PhysicsHelper.cu:386)
        std::vector<int> ptr(1, 0), rows;
        std::vector<float4> anchor;
        for (int i = 0; i < M.nn; i++) {
            if (byRow[i].empty()) continue;
            for (int k : byRow[i]) {
                const auto& h = C->handles[k];
                float4 a = make_float4(h.p[0], h.p[1], h.p[2], 0.0f);
                std::memcpy(&a.w, &h.motion, sizeof(int));
                anchor.push_back(a);
            }
            rows.push_back(i);
            ptr.push_back((int)anchor.size());
        }
        S.nHandleRows = (int)rows.size();
        upload(S.hgPtr, ptr, C->stream);
        upload(S.hgRow, rows, C->stream);
        upload(S.hAnchor, anchor, C->stream);
    }
    hipStream_t st = C->stream;
    upload(S.sliceOff, S.hSliceOff, st);
    upload(S.sliceW, S.hSliceW, st);
    upload(S.cols, cols, st);
    upload(S.rowLen, S.hRowLen, st);
    upload(S.diagSlot, diag, st);
    upload(S.incPtr, incPtr, st);
    upload(S.inc, incFlat, st);
    upload(S.hcount, hcount, st);
    upload(S.hpos, hpos, st);
    S.vals.as<float>((size_t)S.totalSlots * SLICE * 9);
    fill_f32(S.vals.ptr<float>(), 0.0f, (long)S.totalSlots * SLICE * 9, st);  // rows past nn stay zero
    S.b.as<float>(3 * (size_t)M.nn);
    S.bc.as<float>(3 * (size_t)M.nn);
    S.Minv.as<float>(9 * (size_t)M.nn);
    for (DBuf* v : {&S.x, &S.r, &S.q}) v->as<float>(3 * (size_t)M.nn);
    S.z.as<float4>((size_t)M.nn);      // gathered by the SpMV: one 16-byte load per node
    S.d.as<float4>(2 * (size_t)M.nn);  // search direction, two generations (ping-pong)
    S.cRowPtr.as<int>(M.nn + 1);
    S.cDiag.as<int>(M.nn);
    fill_u32(S.cRowPtr.ptr<uint32_t>(), 0u, M.nn + 1, st);
    fill_u32(S.cDiag.ptr<uint32_t>(), 0xffffffffu, M.nn, st);
    fill_f32(S.bc.ptr<float>(), 0.0f, 3L * M.nn, st);
    S.nContact = 0;
}

void host_finalize(cs_ctx* C) {
    if (C->finalized) return;
    hipStream_t st = C->stream;
    C->N = (int)C->hx.size() / 3;
    C->F = (int)C->hfn.size() / 3;
    C->E = (int)C->hen.size() / 2;
    C->NC = C->FC = C->EC = 0;
    for (auto& M : C->meshes)
        if (M.cloth) { C->NC += M.nn; C->FC += M.nf; C->EC += M.ne; }
    const int N = C->N, F = C->F, E = C->E;
    // nodes
    std::vector<float4> x4(N), v4(N);
    for (int i = 0; i < N; i++) {
        x4[i] = make_float4(C->hx[3 * i], C->hx[3 * i + 1], C->hx[3 * i + 2], 0.0f);
        v4[i] = make_float4(C->hv[3 * i], C->hv[3 * i + 1], C->hv[3 * i + 2], 0.0f);
    }
    upload(C->x, x4, st);
    upload(C->x0, x4, st);
    upload(C->x1, x4, st);
    upload(C->v, v4, st);
    if (C->anyMovingObstacle) {
        const int no = N - C->NC;
        std::vector<float4> base(no, make_float4(0, 0, 0, 0));
        std::vector<int> mot(no, -1);
        for (const MeshInfo& M : C->meshes) {
            if (M.cloth || M.motion < 0) continue;
            for (int i = M.n0; i < M.n0 + M.nn; i++) {
                const float* b = &C->hObsBase[3 * (size_t)(i - C->NC)];
                base[i - C->NC] = make_float4(b[0], b[1], b[2], 0.0f);
                mot[i - C->NC] = M.motion;
            }
        }
        upload(C->obsBase, base, st);
        upload(C->obsMotion, mot, st);
    }
    C->nrm.as<float4>(N);
    C->area.as<float>(N);
    C->mass.as<float>(N);
    // faces
    std::vector<int4> fn4(F), fe4(F);
    std::vector<float4> finv(F, make_float4(0, 0, 0, 0));
    std::vector<float> farea(F, 0.0f), fmass(F, 0.0f);
    std::vector<int> fmat(std::max(C->FC, 1), 0);
    for (size_t m = 0; m < C->meshes.size(); m++) {
        const MeshInfo& M = C->meshes[m];
        for (int f = M.f0; f < M.f0 + M.nf; f++) {
            fn4[f] = make_int4(C->hfn[3 * f], C->hfn[3 * f + 1], C->hfn[3 * f + 2], (int)m);
            fe4[f] = make_int4(C->hfe[3 * f], C->hfe[3 * f + 1], C->hfe[3 * f + 2], 0);
            if (M.cloth) {  // Face::initialize (src/Face.cu:12-20)
                const float* u0 = &C->huv[2 * C->hfv[3 * f]];
                const float* u1 = &C->huv[2 * C->hfv[3 * f + 1]];
                const float* u2 = &C->huv[2 * C->hfv[3 * f + 2]];
                float d1x = u1[0] - u0[0], d1y = u1[1] - u0[1];
                float d2x = u2[0] - u0[0], d2y = u2[1] - u0[1];
                // inverse of [d1 d2] (columns) by the adjugate
                float invDet = 1.0f / (d1x * d2y - d2x * d1y);
                finv[f] = make_float4(d2y * invDet, -d2x * invDet, -d1y * invDet, d1x * invDet);
                float a = 0.5f * tabs(d1x * d2y - d1y * d2x);
                farea[f] = a;
                fmass[f] = a * C->materials[M.material].density;
                fmat[f] = M.material;
            }
        }
    }
    upload(C->fnodes, fn4, st);
    upload(C->fedges, fe4, st);
    upload(C->finv, finv, st);
    upload(C->farea, farea, st);
    upload(C->fmass, fmass, st);
    upload(C->fmat, fmat, st);
    C->fnrm.as<float4>(F);
    // edges
    std::vector<int2> en2(E), eo2(E), ea2(E);
    std::vector<float4> bias(std::max(C->EC, 1), make_float4(0, 0, 0, 0));
    for (int e = 0; e < E; e++) {
        en2[e] = make_int2(C->hen[2 * e], C->hen[2 * e + 1]);
        int o0 = C->heopp[2 * e] >= 0 ? C->hvnode[C->heopp[2 * e]] : -1;
        int o1 = C->heopp[2 * e + 1] >= 0 ? C->hvnode[C->heopp[2 * e + 1]] : -1;
        eo2[e] = make_int2(o0, o1);
        ea2[e] = make_int2(C->headj[2 * e], C->headj[2 * e + 1]);
        if (e < C->EC && C->headj[2 * e] >= 0 && C->headj[2 * e + 1] >= 0) {
            float id[2], w[2];
            for (int s = 0; s < 2; s++) {
                int v0, v1;
                edge_side_verts(C, e, s, v0, v1);
                float d0 = C->huv[2 * v1] - C->huv[2 * v0];
                float d1 = C->huv[2 * v1 + 1] - C->huv[2 * v0 + 1];
                bending_bias(d0, d1, id[s], w[s]);
            }
            bias[e] = make_float4(id[0], w[0], id[1], w[1]);
        }
    }
    upload(C->enodes, en2, st);
    upload(C->eopp, eo2, st);
    upload(C->eadj, ea2, st);
    upload(C->ebias, bias, st);
    // node -> incident faces (ascending face order)
    std::vector<int> cnt(N + 1, 0), idx(3 * (size_t)F);
    for (int f = 0; f < F; f++) for (int a = 0; a < 3; a++) cnt[C->hfn[3 * f + a] + 1]++;
    for (int i = 0; i < N; i++) cnt[i + 1] += cnt[i];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int f = 0; f < F; f++) for (int a = 0; a < 3; a++) idx[fill[C->hfn[3 * f + a]]++] = 4 * f + a;
    upload(C->nfPtr, cnt, st);
    upload(C->nfIdx, idx, st);
    // materials
    std::vector<float> tab, bend;
    for (auto& m : C->materials) {
        tab.insert(tab.end(), m.table.begin(), m.table.end());
        bend.insert(bend.end(), m.bend, m.bend + 15);
    }
    if (tab.empty()) { tab.assign(4, 0.0f); bend.assign(15, 0.0f); }
    upload(C->matTable, tab, st);
    upload(C->matBend, bend, st);
    // element records
    C->faceRec.as<float4>(6 * (size_t)std::max(C->FC, 1));
    C->edgeRec.as<float4>(7 * (size_t)std::max(C->EC, 1));
    // cloth systems
    C->sys.clear();
    C->sys.resize(C->nCloths);
    for (int c = 0; c < C->nCloths; c++) build_cloth_system(C, c, C->sys[c]);
    C->bvh.clear();
    C->bvh.resize(C->meshes.size());
    // proximity candidates
    C->bestNode.as<uint64_t>(2 * (size_t)N);
    C->bestEdge.as<uint64_t>(2 * (size_t)E);
    C->bestFace.as<uint64_t>(2 * (size_t)F);
    C->pcgPartial.as<float>(4096);
    C->pcgScal.as<PcgScalars>(1);
    C->pcgCounter.as<uint32_t>(4);
    fill_u32(C->pcgCounter.ptr<uint32_t>(), 0u, 4, st);
    // geometry of every mesh (tail of Mesh::load, src/Mesh.cu:449-451)
    launch_face_geometry(C, 0, F);
    launch_node_geometry(C, 0, N);
    CS_HIP(hipStreamSynchronize(st));
    C->finalized = true;
}

}  // namespace cs
