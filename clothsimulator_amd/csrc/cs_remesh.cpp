// cs_remesh.cpp — flips, splits and collapses of the remeshing step on the host.
//
// Each pass evaluates its predicate for every edge, takes the maximal independent set of the
// minimum-index rule of the reference's device path (src/RemeshingHelper.cu:249-338, 560-647,
// driven by src/Cloth.cu:286-349, 382-432, 567-646; see independent()),
// applies the operators of the chosen edges (src/Operator.cu:7-217) and updates the element lists
// as Mesh::apply does (src/Mesh.cu:138-165): removed entries leave in order, new ones are
// appended in operator order.  The operators of one set touch disjoint faces, so applying them
// one after another gives what the reference's one-thread-per-edge launch gives.
//
// Arithmetic is FP32 in a fixed order (no contraction); the CPU restatement follows the same
// order, so both choose the same edges.
#include "cs_remesh.hpp"

#include <algorithm>
#include <cmath>

namespace cs {

namespace {

using F2 = std::array<float, 2>;
using M2 = std::array<float, 4>;

inline F2 sub2(const F2& a, const F2& b) { return {a[0] - b[0], a[1] - b[1]}; }
inline float dot2(const F2& a, const F2& b) { return a[0] * b[0] + a[1] * b[1]; }
inline F2 m2v(const M2& M, const F2& v) { return {M[0] * v[0] + M[1] * v[1], M[2] * v[0] + M[3] * v[1]}; }
inline float norm2(const F2& a) { return std::sqrt(dot2(a, a)); }

// Edge::initialize: the side whose orientation the face has (src/Edge.cu:13-31)
void init_edge(RMesh& m, int e, int opp, int f) {
    const int n0 = m.en[e][0], n1 = m.en[e][1];
    for (int i = 0; i < 3; i++) {
        const int a = m.fv[f][i], b = m.fv[f][(i + 1) % 3];
        if (m.vnode[a] == n0 && m.vnode[b] == n1) {
            m.ev[e][0] = a;
            m.ev[e][1] = b;
            m.eopp[e][0] = opp;
            m.eadj[e][0] = f;
            return;
        }
        if (m.vnode[a] == n1 && m.vnode[b] == n0) {
            m.ev[e][2] = b;
            m.ev[e][3] = a;
            m.eopp[e][1] = opp;
            m.eadj[e][1] = f;
            return;
        }
    }
}

int new_edge(RMesh& m, int n0, int n1) {
    m.en.push_back({n0, n1});
    m.ev.push_back({-1, -1, -1, -1});
    m.eopp.push_back({-1, -1});
    m.eadj.push_back({-1, -1});
    return (int)m.en.size() - 1;
}

int new_face(RMesh& m, int a, int b, int c) {
    m.fv.push_back({a, b, c});
    m.fe.push_back({-1, -1, -1});
    return (int)m.fv.size() - 1;
}

int new_vertex(RMesh& m, const F2& u, const M2& s, int node) {
    m.U.push_back(u);
    m.S.push_back(s);
    m.vnode.push_back(node);
    return (int)m.U.size() - 1;
}

// Face::findEdge (src/Face.cu:76-85)
int find_edge(const RMesh& m, int f, int a, int b) {
    for (int i = 0; i < 3; i++) {
        const int p = m.fv[f][i], q = m.fv[f][(i + 1) % 3];
        if ((p == a && q == b) || (q == a && p == b)) return m.fe[f][i];
    }
    return -1;
}

inline bool boundary(const RMesh& m, int e) { return m.eadj[e][0] < 0 || m.eadj[e][1] < 0; }
inline bool seam(const RMesh& m, int e) {
    return !boundary(m, e) && (m.ev[e][0] != m.ev[e][2] || m.ev[e][1] != m.ev[e][3]);
}

// Face::initialize's material-space area (src/Face.cu:12-20)
float farea(const RMesh& m, int f) {
    const F2 &a = m.U[m.fv[f][0]], &b = m.U[m.fv[f][1]], &c = m.U[m.fv[f][2]];
    const F2 d1 = sub2(b, a), d2 = sub2(c, a);
    return 0.5f * std::fabs(d1[0] * d2[1] - d1[1] * d2[0]);
}

// shouldFlip (src/RemeshingHelper.cu:224-238)
bool should_flip(const RMesh& m, int e, const RemeshParams& r) {
    if (boundary(m, e) || seam(m, e)) return false;
    const int v0 = m.ev[e][0], v1 = m.ev[e][3], v2 = m.eopp[e][0], v3 = m.eopp[e][1];
    const F2 &x = m.U[v0], &y = m.U[v1], &z = m.U[v2], &w = m.U[v3];
    M2 M;
    for (int c = 0; c < 4; c++) M[c] = 0.25f * (((m.S[v0][c] + m.S[v1][c]) + m.S[v2][c]) + m.S[v3][c]);
    const float a0 = farea(m, m.eadj[e][0]), a1 = farea(m, m.eadj[e][1]);
    const float lhs = a1 * dot2(sub2(x, z), m2v(M, sub2(y, z))) + a0 * dot2(sub2(y, w), m2v(M, sub2(x, w)));
    return lhs < -r.flipThreshold * (a0 + a1);
}

// edgeMetric (src/RemeshingHelper.cu:379-389)
float vmetric(const RMesh& m, int a, int b) {
    if (a < 0 || b < 0) return 0.0f;
    const F2 du = sub2(m.U[a], m.U[b]);
    return std::sqrt(0.5f * (dot2(du, m2v(m.S[a], du)) + dot2(du, m2v(m.S[b], du))));
}

float emetric(const RMesh& m, int e) {
    const float a = vmetric(m, m.ev[e][0], m.ev[e][1]), b = vmetric(m, m.ev[e][2], m.ev[e][3]);
    return b > a ? b : a;
}

struct Rings {  // per node: incident edges and faces, in list order (Cloth::buildAdjacents, src/Cloth.cu:418-452)
    std::vector<int> eptr, eidx, fptr, fidx;
};

void build_rings(const RMesh& m, Rings& g) {
    const size_t N = m.X.size();
    g.eptr.assign(N + 1, 0);
    g.fptr.assign(N + 1, 0);
    for (int e : m.edges)
        for (int s = 0; s < 2; s++) g.eptr[m.en[e][s] + 1]++;
    for (int f : m.faces)
        for (int k = 0; k < 3; k++) g.fptr[m.vnode[m.fv[f][k]] + 1]++;
    for (size_t i = 0; i < N; i++) {
        g.eptr[i + 1] += g.eptr[i];
        g.fptr[i + 1] += g.fptr[i];
    }
    g.eidx.resize(g.eptr[N]);
    g.fidx.resize(g.fptr[N]);
    std::vector<int> fe(g.eptr.begin(), g.eptr.end() - 1), ff(g.fptr.begin(), g.fptr.end() - 1);
    for (int e : m.edges)
        for (int s = 0; s < 2; s++) g.eidx[fe[m.en[e][s]]++] = e;
    for (int f : m.faces)
        for (int k = 0; k < 3; k++) g.fidx[ff[m.vnode[m.fv[f][k]]]++] = f;
}

// shouldCollapseGpu (src/RemeshingHelper.cu:438-485)
bool should_collapse(const RMesh& m, int e, int side, const Rings& g, const RemeshParams& r) {
    const int n0 = m.en[e][side], n1 = m.en[e][1 - side];
    if (m.pres[n0]) return false;
    bool flag = false;
    for (int k = g.eptr[n0]; k < g.eptr[n0 + 1]; k++)
        if (boundary(m, g.eidx[k]) || seam(m, g.eidx[k])) {
            flag = true;
            break;
        }
    if (flag && !boundary(m, e) && !seam(m, e)) return false;
    const int v00 = m.ev[e][side], v01 = m.ev[e][1 - side];
    const int v10 = m.ev[e][2 + side], v11 = m.ev[e][3 - side];
    for (int k = g.fptr[n0]; k < g.fptr[n0 + 1]; k++) {
        const int f = g.fidx[k];
        int vs[3] = {m.fv[f][0], m.fv[f][1], m.fv[f][2]};
        if (m.vnode[vs[0]] == n1 || m.vnode[vs[1]] == n1 || m.vnode[vs[2]] == n1) continue;
        for (int j = 0; j < 3; j++) {
            if (vs[j] == v00) vs[j] = v01;
            else if (vs[j] == v10) vs[j] = v11;
        }
        const F2 &u0 = m.U[vs[0]], &u1 = m.U[vs[1]], &u2 = m.U[vs[2]];
        const F2 d1 = sub2(u1, u0), d2 = sub2(u2, u0);
        const float a = 0.5f * (d1[0] * d2[1] - d1[1] * d2[0]);
        const float p = (norm2(sub2(u0, u1)) + norm2(sub2(u1, u2))) + norm2(sub2(u2, u0));
        const float aspect = 12.0f * std::sqrt(3.0f) * a / (p * p);
        if (a < 1e-6f || aspect < r.aspectMin) return false;
        for (int j = 0; j < 3; j++)
            if (vs[j] != v01 && vs[j] != v11 && vmetric(m, vs[(j + 1) % 3], vs[(j + 2) % 3]) > 0.9f) return false;
    }
    return true;
}

// Maximal independent set by the minimum-index rule over candidates whose faces are listed in CSR
// form (faces of candidate i: cf[cp[i] .. cp[i+1])).  The reference's device rounds (every open
// candidate writes its index into its faces' minimum; one holding the minimum on all of its faces
// is chosen and claims them; candidates touching a claimed face drop out) end with exactly the
// candidates that no smaller chosen candidate shares a face with: the smallest open candidate is
// always chosen, and a candidate waits only on smaller ones.  That set is the greedy pass in index
// order, taken here in one sweep.
std::vector<char> independent(const std::vector<int>& cp, const std::vector<int>& cf, size_t nFacePool) {
    const size_t n = cp.size() - 1;
    std::vector<char> removed(nFacePool, 0), sel(n, 0);
    for (size_t i = 0; i < n; i++) {
        bool free = true;
        for (int k = cp[i]; k < cp[i + 1] && free; k++) free = !removed[cf[k]];
        if (!free) continue;
        sel[i] = 1;
        for (int k = cp[i]; k < cp[i + 1]; k++) removed[cf[k]] = 1;
    }
    return sel;
}

// The faces an edge operator changes: the edge's one or two faces.
void edge_faces_csr(const RMesh& m, const std::vector<int>& edges, std::vector<int>& cp, std::vector<int>& cf) {
    cp.assign(1, 0);
    cf.clear();
    for (int e : edges) {
        for (int s = 0; s < 2; s++)
            if (m.eadj[e][s] >= 0) cf.push_back(m.eadj[e][s]);
        cp.push_back((int)cf.size());
    }
}

void remove_from(std::vector<int>& list, const std::vector<int>& rem, size_t pool) {
    std::vector<char> mark(pool, 0);
    for (int k : rem)
        if (k >= 0) mark[k] = 1;
    size_t w = 0;
    for (size_t i = 0; i < list.size(); i++)
        if (!mark[list[i]]) list[w++] = list[i];
    list.resize(w);
}

void append(std::vector<int>& list, const std::vector<int>& add) { list.insert(list.end(), add.begin(), add.end()); }

// ---------------------------------------------------------------- flip (src/RemeshingHelper.cu:320-375)
bool flip_some(RMesh& m, const RemeshParams& r) {
    std::vector<int> cand;
    for (int e : m.edges)
        if (should_flip(m, e, r)) cand.push_back(e);
    std::vector<int> cp, cf;
    edge_faces_csr(m, cand, cp, cf);
    std::vector<char> sel = independent(cp, cf, m.fv.size());
    std::vector<int> ind;
    for (size_t i = 0; i < cand.size(); i++)
        if (sel[i]) ind.push_back(cand[i]);
    if (ind.empty() || (int)ind.size() == m.flipCount) return false;
    m.flipCount = (int)ind.size();
    std::vector<int> ae, re, af, rf;
    for (int e : ind) {
        const int v0 = m.ev[e][0], v1 = m.ev[e][3], v2 = m.eopp[e][0], v3 = m.eopp[e][1];
        const int f0 = m.eadj[e][0], f1 = m.eadj[e][1];
        const int e0 = find_edge(m, f0, v1, v2), e1 = find_edge(m, f0, v2, v0);
        const int e2 = find_edge(m, f1, v0, v3), e3 = find_edge(m, f1, v3, v1);
        const int ne = new_edge(m, m.vnode[v2], m.vnode[v3]);
        const int nf0 = new_face(m, v0, v3, v2);
        const int nf1 = new_face(m, v1, v2, v3);
        init_edge(m, ne, v0, nf0);
        init_edge(m, ne, v1, nf1);
        m.fe[nf0] = {e2, ne, e1};
        m.fe[nf1] = {e0, ne, e3};
        init_edge(m, e0, v3, nf1);
        init_edge(m, e1, v3, nf0);
        init_edge(m, e2, v2, nf0);
        init_edge(m, e3, v2, nf1);
        ae.push_back(ne);
        re.push_back(e);
        af.push_back(nf0);
        af.push_back(nf1);
        rf.push_back(f0);
        rf.push_back(f1);
    }
    remove_from(m.edges, re, m.en.size());
    append(m.edges, ae);
    remove_from(m.faces, rf, m.fv.size());
    append(m.faces, af);
    return true;
}

void flip_edges(RMesh& m, const RemeshParams& r) {
    const size_t n = m.edges.size();
    for (size_t i = 0; i < 2 * n; i++)
        if (!flip_some(m, r)) break;
}

// ---------------------------------------------------------------- split (src/RemeshingHelper.cu:391-436)
bool split_some(RMesh& m, const RemeshParams& r) {
    struct C { float metric; int e; };
    std::vector<C> cs;
    for (int e : m.edges) {
        const float mt = emetric(m, e);
        if (mt > 1.0f) cs.push_back({mt, e});
    }
    std::stable_sort(cs.begin(), cs.end(), [](const C& a, const C& b) { return a.metric > b.metric; });
    std::vector<int> ce, cp, cf;
    for (const C& c : cs) ce.push_back(c.e);
    edge_faces_csr(m, ce, cp, cf);
    std::vector<char> sel = independent(cp, cf, m.fv.size());
    std::vector<int> ind;
    for (size_t i = 0; i < cs.size(); i++)
        if (sel[i]) ind.push_back(cs[i].e);
    if (ind.empty()) return false;
    std::vector<int> an, av, ae, re, af, rf;
    for (int e : ind) {
        const int n0 = m.en[e][0], n1 = m.en[e][1];
        const int nn = (int)m.X.size();
        std::array<float, 3> px, p0, pv;
        for (int c = 0; c < 3; c++) {
            px[c] = 0.5f * (m.X[n0][c] + m.X[n1][c]);
            p0[c] = 0.5f * (m.X0[n0][c] + m.X0[n1][c]);
            pv[c] = 0.5f * (m.V[n0][c] + m.V[n1][c]);
        }
        m.X.push_back(px);
        m.X0.push_back(p0);
        m.V.push_back(pv);
        m.pres.push_back(0);
        const int ne[2] = {new_edge(m, nn, n0), new_edge(m, nn, n1)};
        an.push_back(nn);
        re.push_back(e);
        auto mid = [&](int a, int b, F2& u, M2& s) {
            for (int c = 0; c < 2; c++) u[c] = 0.5f * (m.U[a][c] + m.U[b][c]);
            for (int c = 0; c < 4; c++) s[c] = 0.5f * (m.S[a][c] + m.S[b][c]);
        };
        int nv[2];
        F2 u;
        M2 s;
        if (seam(m, e)) {
            mid(m.ev[e][0], m.ev[e][1], u, s);
            nv[0] = new_vertex(m, u, s, nn);
            mid(m.ev[e][2], m.ev[e][3], u, s);
            nv[1] = new_vertex(m, u, s, nn);
            av.push_back(nv[0]);
            av.push_back(nv[1]);
        } else {
            const int j = m.eopp[e][0] >= 0 ? 0 : 1;
            mid(m.ev[e][2 * j], m.ev[e][2 * j + 1], u, s);
            nv[0] = nv[1] = new_vertex(m, u, s, nn);
            av.push_back(nv[0]);
        }
        ae.push_back(ne[0]);
        ae.push_back(ne[1]);
        for (int j = 0; j < 2; j++) {
            if (m.eopp[e][j] < 0) continue;
            const int v0 = m.ev[e][2 * j + j], v1 = m.ev[e][2 * j + 1 - j], v2 = m.eopp[e][j];
            const int f = m.eadj[e][j];
            const int e0 = find_edge(m, f, v1, v2), e1 = find_edge(m, f, v2, v0);
            const int w = nv[j];
            const int ne0 = ne[j], ne1 = ne[1 - j];
            const int ne2 = new_edge(m, nn, m.vnode[v2]);
            const int nf0 = new_face(m, v0, w, v2);
            const int nf1 = new_face(m, v2, w, v1);
            init_edge(m, ne0, v2, nf0);
            init_edge(m, ne1, v2, nf1);
            init_edge(m, ne2, v0, nf0);
            init_edge(m, ne2, v1, nf1);
            m.fe[nf0] = {ne0, ne2, e1};
            m.fe[nf1] = {ne2, ne1, e0};
            init_edge(m, e0, w, nf1);
            init_edge(m, e1, w, nf0);
            ae.push_back(ne2);
            af.push_back(nf0);
            af.push_back(nf1);
            rf.push_back(f);
        }
    }
    append(m.nodes, an);
    append(m.verts, av);
    remove_from(m.edges, re, m.en.size());
    append(m.edges, ae);
    remove_from(m.faces, rf, m.fv.size());
    append(m.faces, af);
    return true;
}

// ---------------------------------------------------------------- collapse (src/RemeshingHelper.cu:649-757)
bool collapse_some(RMesh& m, const RemeshParams& r) {
    Rings g;
    build_rings(m, g);
    std::vector<std::array<int, 2>> cand;
    for (int e : m.edges) {
        if (should_collapse(m, e, 0, g, r)) cand.push_back({e, 0});
        else if (should_collapse(m, e, 1, g, r)) cand.push_back({e, 1});
    }
    std::vector<int> cp(1, 0), cf;  // the faces around the node that goes
    for (const auto& c : cand) {
        const int n0 = m.en[c[0]][c[1]];
        cf.insert(cf.end(), g.fidx.begin() + g.fptr[n0], g.fidx.begin() + g.fptr[n0 + 1]);
        cp.push_back((int)cf.size());
    }
    std::vector<char> sel = independent(cp, cf, m.fv.size());
    std::vector<std::array<int, 2>> ind;
    for (size_t i = 0; i < cand.size(); i++)
        if (sel[i]) ind.push_back(cand[i]);
    if (ind.empty()) return false;
    std::vector<int> rn, rv, re, rf;
    for (const auto& c : ind) {
        const int e = c[0], side = c[1];
        const int n0 = m.en[e][side], n1 = m.en[e][1 - side];
        const int v00 = m.ev[e][side], v01 = m.ev[e][1 - side];
        const int v10 = m.ev[e][2 + side], v11 = m.ev[e][3 - side];
        rn.push_back(n0);
        re.push_back(e);
        if (seam(m, e)) {
            rv.push_back(v00);
            rv.push_back(v10);
        } else {
            rv.push_back(v00 >= 0 ? v00 : v10);
        }
        for (int j = 0; j < 2; j++) {
            if (m.eopp[e][j] < 0) continue;
            const int v0 = m.ev[e][2 * j + side], v1 = m.ev[e][2 * j + 1 - side], v2 = m.eopp[e][j];
            const int f = m.eadj[e][j];
            const int e0 = find_edge(m, f, v1, v2), e1 = find_edge(m, f, v2, v0);
            int v = -1, h = -1;
            if (!boundary(m, e1)) {
                const int s = m.eopp[e1][0] != v1 ? 0 : 1;
                v = m.eopp[e1][s];
                h = m.eadj[e1][s];
                for (int i = 0; i < 3; i++)  // Face::replaceEdge: the first match
                    if (m.fe[h][i] == e1) {
                        m.fe[h][i] = e0;
                        break;
                    }
            }
            for (int s = 0; s < 2; s++)  // Edge::replaceOpposite
                if (m.eopp[e0][s] == v0) {
                    m.eopp[e0][s] = v;
                    break;
                }
            for (int s = 0; s < 2; s++)  // Edge::replaceAdjacent
                if (m.eadj[e0][s] == f) {
                    m.eadj[e0][s] = h;
                    break;
                }
            re.push_back(e1);
            rf.push_back(f);
        }
        for (int k = g.eptr[n0]; k < g.eptr[n0 + 1]; k++) {
            const int a = g.eidx[k];
            if (a == e) continue;
            for (int s = 0; s < 2; s++)  // Edge::replaceNode
                if (m.en[a][s] == n0) {
                    m.en[a][s] = n1;
                    break;
                }
            const int from[2] = {v00, v10}, to[2] = {v01, v11};
            for (int q = 0; q < 2; q++)  // Edge::replaceVertex: the first match on each side
                for (int s = 0; s < 2; s++)
                    for (int t = 0; t < 2; t++)
                        if (m.ev[a][2 * s + t] == from[q]) {
                            m.ev[a][2 * s + t] = to[q];
                            break;
                        }
        }
        const int adj0 = m.eadj[e][0], adj1 = m.eadj[e][1];
        for (int k = g.fptr[n0]; k < g.fptr[n0 + 1]; k++) {
            const int h = g.fidx[k];
            if (h == adj0 || h == adj1) continue;
            int opp = -1;
            for (int i = 0; i < 3; i++)  // Face::findOpposite
                if (m.vnode[m.fv[h][i]] == n0) {
                    opp = m.fe[h][(i + 1) % 3];
                    break;
                }
            const int from[2] = {v00, v10}, to[2] = {v01, v11};
            for (int q = 0; q < 2; q++)
                for (int s = 0; s < 2; s++)
                    if (m.eopp[opp][s] == from[q]) {
                        m.eopp[opp][s] = to[q];
                        break;
                    }
            for (int q = 0; q < 2; q++)  // Face::replaceVertex
                for (int i = 0; i < 3; i++)
                    if (m.fv[h][i] == from[q]) {
                        m.fv[h][i] = to[q];
                        break;
                    }
        }
    }
    remove_from(m.nodes, rn, m.X.size());
    remove_from(m.verts, rv, m.U.size());
    remove_from(m.edges, re, m.en.size());
    remove_from(m.faces, rf, m.fv.size());
    return true;
}

}  // namespace

bool rmesh_build(RMesh& m, int nn, const float* x, const float* x0, const float* v, const unsigned char* preserve,
                 int nv, const float* uv, const int* vnode, int nf, const int* fv, int ne, const int* en) {
    m = RMesh();
    m.X.resize(nn);
    m.X0.resize(nn);
    m.V.resize(nn);
    m.pres.resize(nn);
    for (int i = 0; i < nn; i++) {
        for (int c = 0; c < 3; c++) {
            m.X[i][c] = x[3 * i + c];
            m.X0[i][c] = x0[3 * i + c];
            m.V[i][c] = v[3 * i + c];
        }
        m.pres[i] = preserve[i] ? 1 : 0;
    }
    m.U.resize(nv);
    m.S.assign(nv, M2{0.0f, 0.0f, 0.0f, 0.0f});
    m.vnode.assign(vnode, vnode + nv);
    for (int k = 0; k < nv; k++) m.U[k] = {uv[2 * k], uv[2 * k + 1]};
    m.fv.resize(nf);
    m.fe.assign(nf, {-1, -1, -1});
    for (int f = 0; f < nf; f++) m.fv[f] = {fv[3 * f], fv[3 * f + 1], fv[3 * f + 2]};
    m.en.resize(ne);
    m.ev.assign(ne, {-1, -1, -1, -1});
    m.eopp.assign(ne, {-1, -1});
    m.eadj.assign(ne, {-1, -1});
    // edge of each node pair, through a sorted key table
    std::vector<std::pair<long long, int>> key(ne);
    for (int e = 0; e < ne; e++) {
        m.en[e] = {en[2 * e], en[2 * e + 1]};
        const long long a = std::min(en[2 * e], en[2 * e + 1]), b = std::max(en[2 * e], en[2 * e + 1]);
        key[e] = {a * nn + b, e};
    }
    std::sort(key.begin(), key.end());
    for (int f = 0; f < nf; f++)
        for (int i = 0; i < 3; i++) {
            const long long p = m.vnode[m.fv[f][i]], q = m.vnode[m.fv[f][(i + 1) % 3]];
            const long long k = std::min(p, q) * nn + std::max(p, q);
            auto it = std::lower_bound(key.begin(), key.end(), std::make_pair(k, -1));
            if (it == key.end() || it->first != k) return false;
            const int e = it->second;
            m.fe[f][i] = e;
            // a manifold mesh fills each edge side once: a second face on the same side (a third
            // face on the edge, or two faces with the same orientation) is malformed input
            const int side = m.vnode[m.fv[f][i]] == m.en[e][0] ? 0 : 1;
            if (m.eadj[e][side] >= 0) return false;
            init_edge(m, e, m.fv[f][(i + 2) % 3], f);
        }
    for (int i = 0; i < nn; i++) m.nodes.push_back(i);
    for (int k = 0; k < nv; k++) m.verts.push_back(k);
    for (int e = 0; e < ne; e++) m.edges.push_back(e);
    for (int f = 0; f < nf; f++) m.faces.push_back(f);
    return true;
}

void rmesh_set_face_sizing(RMesh& m, const float* fs) {
    std::vector<M2> acc(m.U.size(), M2{0.0f, 0.0f, 0.0f, 0.0f});
    std::vector<float> ar(m.U.size(), 0.0f);
    for (size_t i = 0; i < m.faces.size(); i++) {
        const int f = m.faces[i];
        const float a = farea(m, f);
        for (int k = 0; k < 3; k++) {
            const int vt = m.fv[f][k];
            ar[vt] = ar[vt] + a;
            for (int c = 0; c < 4; c++) acc[vt][c] = acc[vt][c] + a * fs[4 * i + c];
        }
    }
    for (int vt : m.verts) {
        const float inv = 1.0f / ar[vt];
        for (int c = 0; c < 4; c++) m.S[vt][c] = acc[vt][c] * inv;
    }
}

void rmesh_remesh(RMesh& m, const RemeshParams& r) {
    flip_edges(m, r);
    while (split_some(m, r)) flip_edges(m, r);
    while (collapse_some(m, r)) flip_edges(m, r);
}

void rmesh_output(const RMesh& m, int nInputNodes, RMeshOut& o) {
    std::vector<int> ni(m.X.size(), -1), vi(m.U.size(), -1);
    for (size_t i = 0; i < m.nodes.size(); i++) ni[m.nodes[i]] = (int)i;
    for (size_t i = 0; i < m.verts.size(); i++) vi[m.verts[i]] = (int)i;
    const size_t N = m.nodes.size(), V = m.verts.size(), Fc = m.faces.size(), E = m.edges.size();
    o.x.resize(3 * N);
    o.x0.resize(3 * N);
    o.v.resize(3 * N);
    o.preserve.resize(N);
    o.nodeOrigin.resize(N);
    for (size_t i = 0; i < N; i++) {
        const int n = m.nodes[i];
        for (int c = 0; c < 3; c++) {
            o.x[3 * i + c] = m.X[n][c];
            o.x0[3 * i + c] = m.X0[n][c];
            o.v[3 * i + c] = m.V[n][c];
        }
        o.preserve[i] = m.pres[n];
        o.nodeOrigin[i] = n < nInputNodes ? n : -1;
    }
    o.uv.resize(2 * V);
    o.sizing.resize(4 * V);
    o.vnode.resize(V);
    for (size_t i = 0; i < V; i++) {
        const int k = m.verts[i];
        o.uv[2 * i] = m.U[k][0];
        o.uv[2 * i + 1] = m.U[k][1];
        for (int c = 0; c < 4; c++) o.sizing[4 * i + c] = m.S[k][c];
        o.vnode[i] = ni[m.vnode[k]];
    }
    o.fv.resize(3 * Fc);
    for (size_t i = 0; i < Fc; i++)
        for (int c = 0; c < 3; c++) o.fv[3 * i + c] = vi[m.fv[m.faces[i]][c]];
    std::vector<int> ei(m.en.size(), -1), fi(m.fv.size(), -1);
    for (size_t i = 0; i < E; i++) ei[m.edges[i]] = (int)i;
    for (size_t i = 0; i < Fc; i++) fi[m.faces[i]] = (int)i;
    o.en.resize(2 * E);
    o.eopp.resize(2 * E);
    o.eadj.resize(2 * E);
    for (size_t i = 0; i < E; i++) {
        const int e = m.edges[i];
        for (int c = 0; c < 2; c++) {
            o.en[2 * i + c] = ni[m.en[e][c]];
            o.eopp[2 * i + c] = m.eopp[e][c] >= 0 ? vi[m.eopp[e][c]] : -1;
            o.eadj[2 * i + c] = m.eadj[e][c] >= 0 ? fi[m.eadj[e][c]] : -1;
        }
    }
    o.fe.resize(3 * Fc);
    for (size_t i = 0; i < Fc; i++)
        for (int c = 0; c < 3; c++) o.fe[3 * i + c] = ei[m.fe[m.faces[i]][c]];
}

}  // namespace cs
