// cs_remesh_ctx.cpp — the remeshing step on a context (Simulator::remeshingStep,
// src/Simulator.cu:686-695, with Cloth::remeshingStep, src/Cloth.cu:815-833).
//
// Per remeshed cloth: the sizing field on the device (cs_sizing.hip), the operators on the host
// (cs_remesh.cpp) over the context's own topology lists, then every mesh's host arrays are
// re-concatenated and the context is finalised again: a topology change rebuilds what the
// design keeps static per cloth (the sliced-ELL pattern and its slot maps, the incidence lists,
// the multigrid hierarchy, the learned factor's pattern, the trees).  Node states come from the
// device; handles follow their nodes, which are preserved and so never removed.
#include <algorithm>
#include <chrono>

#include "cs_kernels.hpp"

namespace cs {

namespace {

// One mesh in local ids.
struct Local {
    std::vector<float> x, x0, v, uv;
    std::vector<unsigned char> pres;
    std::vector<int> vnode, fv, fe, en, eopp, eadj;
};

Local extract(const cs_ctx* C, const MeshInfo& M) {
    Local L;
    L.x.assign(C->hx.begin() + 3L * M.n0, C->hx.begin() + 3L * (M.n0 + M.nn));
    L.v.assign(C->hv.begin() + 3L * M.n0, C->hv.begin() + 3L * (M.n0 + M.nn));
    L.pres.assign(C->hpres.begin() + M.n0, C->hpres.begin() + M.n0 + M.nn);
    L.uv.assign(C->huv.begin() + 2L * M.v0, C->huv.begin() + 2L * (M.v0 + M.nv));
    for (int k = 0; k < M.nv; k++) L.vnode.push_back(C->hvnode[M.v0 + k] - M.n0);
    for (long k = 3L * M.f0; k < 3L * (M.f0 + M.nf); k++) {
        L.fv.push_back(C->hfv[k] - M.v0);
        L.fe.push_back(C->hfe[k] - M.e0);
    }
    for (long k = 2L * M.e0; k < 2L * (M.e0 + M.ne); k++) {
        L.en.push_back(C->hen[k] - M.n0);
        L.eopp.push_back(C->heopp[k] >= 0 ? C->heopp[k] - M.v0 : -1);
        L.eadj.push_back(C->headj[k] >= 0 ? C->headj[k] - M.f0 : -1);
    }
    return L;
}

// Nodes of boundary and seam edges (src/MeshHelper.cu:93-102) and handle nodes
// (src/Cloth.cu:649-673), as the loaded meshes have them.
void init_preserve(cs_ctx* C) {
    C->hpres.assign(C->hx.size() / 3, 0);
    for (const MeshInfo& M : C->meshes) {
        if (!M.cloth) continue;
        for (int e = M.e0; e < M.e0 + M.ne; e++) {
            const int a0 = C->headj[2 * e], a1 = C->headj[2 * e + 1];
            bool mark = a0 < 0 || a1 < 0;
            if (!mark) {  // a seam: the two sides name different uv vertices
                int s0[2] = {-1, -1}, s1[2] = {-1, -1};
                for (int side = 0; side < 2; side++) {
                    const int f = C->headj[2 * e + side];
                    for (int i = 0; i < 3; i++) {
                        const int va = C->hfv[3 * f + i], vb = C->hfv[3 * f + (i + 1) % 3];
                        const int na = C->hvnode[va], nb = C->hvnode[vb];
                        int* s = side == 0 ? s0 : s1;
                        if (na == C->hen[2 * e] && nb == C->hen[2 * e + 1]) { s[0] = va; s[1] = vb; }
                        else if (nb == C->hen[2 * e] && na == C->hen[2 * e + 1]) { s[0] = vb; s[1] = va; }
                    }
                }
                mark = s0[0] != s1[0] || s0[1] != s1[1];
            }
            if (mark) C->hpres[C->hen[2 * e]] = C->hpres[C->hen[2 * e + 1]] = 1;
        }
    }
    for (const auto& h : C->handles) C->hpres[h.node] = 1;  // global node id
    C->presReady = true;
}

}  // namespace

int remesh_context(cs_ctx* C) {
    using clk = std::chrono::steady_clock;
    auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<float, std::milli>(b - a).count(); };
    float tSizing = 0.0f, tOps = 0.0f;
    if (!C->finalized) host_finalize(C);
    hipStream_t s = C->stream;
    const int N = C->N;
    std::vector<float4> x4(N), x04(N), v4(N);
    CS_HIP(hipMemcpyAsync(x4.data(), C->x.p, sizeof(float4) * N, hipMemcpyDeviceToHost, s));
    CS_HIP(hipMemcpyAsync(x04.data(), C->x0.p, sizeof(float4) * N, hipMemcpyDeviceToHost, s));
    CS_HIP(hipMemcpyAsync(v4.data(), C->v.p, sizeof(float4) * N, hipMemcpyDeviceToHost, s));
    CS_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < N; i++) {
        C->hx[3 * i] = x4[i].x; C->hx[3 * i + 1] = x4[i].y; C->hx[3 * i + 2] = x4[i].z;
        C->hv[3 * i] = v4[i].x; C->hv[3 * i + 1] = v4[i].y; C->hv[3 * i + 2] = v4[i].z;
    }
    if (!C->presReady) init_preserve(C);
    const int nm = (int)C->meshes.size();
    // the frame's faces before remeshing (Simulator::simulateStep, src/Simulator.cu:858-866)
    std::vector<std::vector<BackupFace>> backup(C->nCloths);
    for (int c = 0; c < C->nCloths; c++) {
        const MeshInfo& M = C->meshes[c];
        backup[c].resize(M.nf);
        for (int f = 0; f < M.nf; f++)
            for (int k = 0; k < 3; k++) {
                const int vt = C->hfv[3 * (M.f0 + f) + k], nd = C->hfn[3 * (M.f0 + f) + k];
                backup[c][f].u[k][0] = C->huv[2 * vt];
                backup[c][f].u[k][1] = C->huv[2 * vt + 1];
                for (int a = 0; a < 3; a++) backup[c][f].x[k][a] = C->hx[3 * nd + a];
            }
    }
    std::vector<Local> loc(nm);
    std::vector<std::vector<int>> newIndex(nm);  // remeshed cloths: old local node -> new local node
    bool any = false;
    for (int m = 0; m < nm; m++) {
        const MeshInfo& M = C->meshes[m];
        loc[m] = extract(C, M);
        for (int i = M.n0; i < M.n0 + M.nn; i++) {
            loc[m].x0.push_back(x04[i].x);
            loc[m].x0.push_back(x04[i].y);
            loc[m].x0.push_back(x04[i].z);
        }
        if (!M.cloth || m >= (int)C->remeshCfg.size() || !C->remeshCfg[m].has) continue;
        const RemeshParams& r = C->remeshCfg[m].r;
        std::vector<float> fs;
        const auto t0 = clk::now();
        remesh_sizing(C, m, r, fs, nullptr);
        const auto t1 = clk::now();
        tSizing += ms(t0, t1);
        std::vector<float> x0(3L * M.nn);
        for (int i = 0; i < M.nn; i++) {
            x0[3 * i] = x04[M.n0 + i].x; x0[3 * i + 1] = x04[M.n0 + i].y; x0[3 * i + 2] = x04[M.n0 + i].z;
        }
        const Local& L = loc[m];
        RMesh rm;
        if (!rmesh_build(rm, M.nn, L.x.data(), x0.data(), L.v.data(), L.pres.data(), M.nv, L.uv.data(),
                         L.vnode.data(), M.nf, L.fv.data(), M.ne, L.en.data()))
            throw StatusError(CS_E_FORMAT, "remeshing: the cloth's topology is not a consistently oriented "
                                           "manifold (an edge side with two faces)");
        rm.flipCount = C->flipCount;
        rmesh_set_face_sizing(rm, fs.data());
        rmesh_remesh(rm, r);
        C->flipCount = rm.flipCount;
        RMeshOut o;
        rmesh_output(rm, M.nn, o);
        tOps += ms(t1, clk::now());
        newIndex[m].assign(M.nn, -1);
        for (size_t i = 0; i < o.nodeOrigin.size(); i++)
            if (o.nodeOrigin[i] >= 0) newIndex[m][o.nodeOrigin[i]] = (int)i;
        Local R;
        R.x = std::move(o.x);
        R.x0 = std::move(o.x0);
        R.v = std::move(o.v);
        R.uv = std::move(o.uv);
        R.pres = std::move(o.preserve);
        R.vnode = std::move(o.vnode);
        R.fv = std::move(o.fv);
        R.fe = std::move(o.fe);
        R.en = std::move(o.en);
        R.eopp = std::move(o.eopp);
        R.eadj = std::move(o.eadj);
        loc[m] = std::move(R);
        any = true;
    }
    C->remeshMs[0] = tSizing;
    C->remeshMs[1] = tOps;
    C->remeshMs[2] = C->remeshMs[3] = 0.0f;
    C->sepIters = 0;
    if (!any) return CS_OK;
    const auto tb = clk::now();
    // handles follow their (preserved) nodes: local ids now, global once the meshes are placed
    std::vector<int> handleLocal;
    for (auto& h : C->handles) {
        int k = h.node - C->meshes[h.cloth].n0;
        if (!newIndex[h.cloth].empty()) k = newIndex[h.cloth][k];
        if (k < 0) throw StatusError(CS_E_STATE, "remeshing removed a handle node");
        handleLocal.push_back(k);
    }
    // re-concatenate every mesh
    C->hx.clear(); C->hv.clear(); C->huv.clear(); C->hpres.clear();
    C->hvnode.clear(); C->hfv.clear(); C->hfn.clear(); C->hfe.clear();
    C->hen.clear(); C->heopp.clear(); C->headj.clear();
    for (int m = 0; m < nm; m++) {
        MeshInfo& M = C->meshes[m];
        const Local& L = loc[m];
        M.n0 = (int)C->hx.size() / 3;
        M.nn = (int)L.x.size() / 3;
        M.v0 = (int)C->huv.size() / 2;
        M.nv = (int)L.uv.size() / 2;
        M.f0 = (int)C->hfv.size() / 3;
        M.nf = (int)L.fv.size() / 3;
        M.e0 = (int)C->hen.size() / 2;
        M.ne = (int)L.en.size() / 2;
        C->hx.insert(C->hx.end(), L.x.begin(), L.x.end());
        C->hv.insert(C->hv.end(), L.v.begin(), L.v.end());
        C->huv.insert(C->huv.end(), L.uv.begin(), L.uv.end());
        C->hpres.insert(C->hpres.end(), L.pres.begin(), L.pres.end());
        for (int k : L.vnode) C->hvnode.push_back(k + M.n0);
        for (int k : L.fv) {
            C->hfv.push_back(k + M.v0);
            C->hfn.push_back(L.vnode[k] + M.n0);
        }
        for (int k : L.fe) C->hfe.push_back(k + M.e0);
        for (int k : L.en) C->hen.push_back(k + M.n0);
        for (int k : L.eopp) C->heopp.push_back(k >= 0 ? k + M.v0 : -1);
        for (int k : L.eadj) C->headj.push_back(k >= 0 ? k + M.f0 : -1);
    }
    for (size_t k = 0; k < C->handles.size(); k++)
        C->handles[k].node = C->meshes[C->handles[k].cloth].n0 + handleLocal[k];
    C->finalized = false;
    C->bamg.ncl = -1;
    C->bamg.selected = false;
    host_finalize(C);
    // the nodes' previous positions (the step's start), so that the velocities the separation
    // step recomputes, (x - x0) / dt, are the step's
    std::vector<float4> x0n(C->N);
    for (int m = 0; m < nm; m++) {
        const MeshInfo& M = C->meshes[m];
        for (int i = 0; i < M.nn; i++)
            x0n[M.n0 + i] = make_float4(loc[m].x0[3 * i], loc[m].x0[3 * i + 1], loc[m].x0[3 * i + 2], 0.0f);
    }
    CS_HIP(hipMemcpyAsync(C->x0.p, x0n.data(), sizeof(float4) * C->N, hipMemcpyHostToDevice, C->stream));
    CS_HIP(hipStreamSynchronize(C->stream));
    const auto ts = clk::now();
    C->remeshMs[2] = ms(tb, ts);
    const int st = separation_step(C, backup);
    C->remeshMs[3] = ms(ts, clk::now());
    return st;
}

}  // namespace cs
