// or_amg.hpp — CPU restatement of the aggregation multigrid preconditioner
// (test infrastructure: the checker of the product's CS_PRECOND_AMG path).
//
// The reference has no multigrid: its PCG uses a Jacobi preconditioner
// (src/Cloth.cu:766-809).  This preconditioner is this project's addition; it
// changes how fast PCG reaches ||r|| <= tol ||b||, not the system solved.
// Definition (shared with the product, which restates it on the GPU):
//   aggregation graph of level 0: cloth nodes adjacent through a face;
//   aggregation: (1) in ascending node order, a node whose neighbours are all
//   unaggregated founds an aggregate with all of them; (2) in ascending order,
//   each remaining node joins the aggregate of its first aggregated neighbour
//   (ascending), or founds a singleton; aggregates are numbered in creation
//   order; the coarse graph links aggregates holding adjacent nodes;
//   levels are added until a level has <= AMG_COARSEST nodes;
//   prolongation: piecewise constant (3x3 identity block per node);
//   level-0 operator: the full system; coarse operators: Galerkin products
//   P^T A P of the static system (mass, forces, handles: no proximity blocks);
//   V(1,1) cycle: x = w D^-1 b; x += P V(P^T (b - A x)); x += w D^-1 (b - A x),
//   w = AMG_OMEGA, D the 3x3 diagonal blocks; exact solve on the coarsest level.
#pragma once
#include <map>
#include <vector>

#include "or_physics.hpp"

namespace orc {

static const int AMG_COARSEST = 64;
static const double AMG_OMEGA = 0.6;

// graph: CSR adjacency (sorted, no self loops)
inline void amg_aggregate(const vector<int>& gp, const vector<int>& gi, vector<int>& agg, int& na) {
    const int n = (int)gp.size() - 1;
    agg.assign(n, -1);
    na = 0;
    for (int i = 0; i < n; i++) {
        if (agg[i] >= 0) continue;
        bool free = true;
        for (int k = gp[i]; k < gp[i + 1]; k++) free = free && agg[gi[k]] < 0;
        if (!free) continue;
        agg[i] = na;
        for (int k = gp[i]; k < gp[i + 1]; k++) agg[gi[k]] = na;
        na++;
    }
    for (int i = 0; i < n; i++) {
        if (agg[i] >= 0) continue;
        int a = -1;
        for (int k = gp[i]; k < gp[i + 1] && a < 0; k++) a = agg[gi[k]];
        agg[i] = a >= 0 ? a : na++;
    }
}

inline void amg_coarse_graph(const vector<int>& gp, const vector<int>& gi, const vector<int>& agg, int na,
                             vector<int>& cp, vector<int>& cidx) {
    vector<vector<int>> adj(na);
    for (int i = 0; i + 1 < (int)gp.size(); i++)
        for (int k = gp[i]; k < gp[i + 1]; k++)
            if (agg[i] != agg[gi[k]]) adj[agg[i]].push_back(agg[gi[k]]);
    cp.assign(na + 1, 0);
    cidx.clear();
    for (int a = 0; a < na; a++) {
        std::sort(adj[a].begin(), adj[a].end());
        adj[a].erase(std::unique(adj[a].begin(), adj[a].end()), adj[a].end());
        cidx.insert(cidx.end(), adj[a].begin(), adj[a].end());
        cp[a + 1] = (int)cidx.size();
    }
}

// node graph of a triangle mesh (local ids)
inline void amg_face_graph(int n, const vector<array<int, 3>>& faces, vector<int>& gp, vector<int>& gi) {
    vector<vector<int>> adj(n);
    for (const auto& f : faces)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                if (a != b) adj[f[a]].push_back(f[b]);
    gp.assign(n + 1, 0);
    gi.clear();
    for (int i = 0; i < n; i++) {
        std::sort(adj[i].begin(), adj[i].end());
        adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
        gi.insert(gi.end(), adj[i].begin(), adj[i].end());
        gp[i + 1] = (int)gi.size();
    }
}

// aggregate maps of every level (agg[l]: level-l node -> level-(l+1) node) and level sizes
struct AmgTopo {
    vector<int> sizes;
    vector<vector<int>> agg;
};

inline void amg_topology(int n, vector<int> gp, vector<int> gi, AmgTopo& T) {
    T.sizes.assign(1, n);
    T.agg.clear();
    while (n > AMG_COARSEST) {
        vector<int> agg, cp, ci;
        int na = 0;
        amg_aggregate(gp, gi, agg, na);
        if (na >= n) break;  // no coarsening possible
        amg_coarse_graph(gp, gi, agg, na, cp, ci);
        T.agg.push_back(agg);
        T.sizes.push_back(na);
        gp.swap(cp);
        gi.swap(ci);
        n = na;
    }
}

template <class T> void amg_cloth_topology(const Scene<T>& S, int cloth, AmgTopo& topo) {
    const MeshRange& R = S.meshes[cloth];
    vector<array<int, 3>> faces(R.nf);
    for (int f = 0; f < R.nf; f++)
        for (int a = 0; a < 3; a++) faces[f][a] = S.fn[R.f0 + f][a] - R.n0;
    vector<int> gp, gi;
    amg_face_graph(R.nn, faces, gp, gi);
    amg_topology(R.nn, gp, gi, topo);
}

struct AmgLevel {
    int n = 0;
    vector<int> rp, ci;  // block CSR, sorted columns
    vector<double> val;  // 9 per block
    vector<double> Dinv;  // 9 per node
};

struct Amg {
    AmgTopo topo;
    vector<AmgLevel> L;          // L[0]: the full system; L[l >= 1]: Galerkin of the static system
    vector<double> coarseInv;   // dense inverse of the coarsest operator
};

inline void amg_dinv(AmgLevel& a) {
    a.Dinv.assign(9 * (size_t)a.n, 0.0);
    for (int i = 0; i < a.n; i++)
        for (int k = a.rp[i]; k < a.rp[i + 1]; k++)
            if (a.ci[k] == i) inverse3(&a.val[9 * (size_t)k], &a.Dinv[9 * (size_t)i]);
}

inline void amg_galerkin(const AmgLevel& f, const vector<int>& agg, int nc, AmgLevel& c) {
    std::map<uint64_t, array<double, 9>> blocks;
    for (int i = 0; i < f.n; i++)
        for (int k = f.rp[i]; k < f.rp[i + 1]; k++) {
            uint64_t key = (uint64_t)agg[i] * (uint64_t)nc + (uint64_t)agg[f.ci[k]];
            auto it = blocks.find(key);
            if (it == blocks.end()) { array<double, 9> z; z.fill(0.0); it = blocks.emplace(key, z).first; }
            for (int e = 0; e < 9; e++) it->second[e] += f.val[9 * (size_t)k + e];
        }
    c.n = nc;
    c.rp.assign(nc + 1, 0);
    c.ci.clear();
    c.val.clear();
    for (auto& kv : blocks) {
        c.rp[kv.first / nc + 1]++;
        c.ci.push_back((int)(kv.first % nc));
        c.val.insert(c.val.end(), kv.second.begin(), kv.second.end());
    }
    for (int i = 0; i < nc; i++) c.rp[i + 1] += c.rp[i];
}

// dense inverse by Gauss-Jordan with partial pivoting (m x m, row-major)
inline void dense_inverse(vector<double> a, int m, vector<double>& inv) {
    inv.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; i++) inv[(size_t)i * m + i] = 1.0;
    for (int c = 0; c < m; c++) {
        int p = c;
        for (int r = c + 1; r < m; r++) if (std::fabs(a[(size_t)r * m + c]) > std::fabs(a[(size_t)p * m + c])) p = r;
        if (p != c)
            for (int k = 0; k < m; k++) { std::swap(a[(size_t)p * m + k], a[(size_t)c * m + k]); std::swap(inv[(size_t)p * m + k], inv[(size_t)c * m + k]); }
        const double d = 1.0 / a[(size_t)c * m + c];
        for (int k = 0; k < m; k++) { a[(size_t)c * m + k] *= d; inv[(size_t)c * m + k] *= d; }
        for (int r = 0; r < m; r++) {
            if (r == c) continue;
            const double f = a[(size_t)r * m + c];
            if (f == 0.0) continue;
            for (int k = 0; k < m; k++) { a[(size_t)r * m + k] -= f * a[(size_t)c * m + k]; inv[(size_t)r * m + k] -= f * inv[(size_t)c * m + k]; }
        }
    }
}

// full: the system being solved (level 0); stat: its static part (same n)
template <class T>
void amg_setup(const AmgTopo& topo, const vector<int>& rp, const vector<int>& ci, const vector<T>& val,
               const vector<int>& srp, const vector<int>& sci, const vector<T>& sval, Amg& M) {
    M.topo = topo;
    const int nl = (int)topo.sizes.size();
    M.L.assign(nl, AmgLevel());
    AmgLevel& L0 = M.L[0];
    L0.n = topo.sizes[0];
    L0.rp = rp;
    L0.ci = ci;
    L0.val.assign(val.begin(), val.end());
    amg_dinv(L0);
    // coarse operators: Galerkin products of the static system plus the contact blocks'
    // diagonal part (a contact with an obstacle, or a node's own share of a self-contact, stiffens
    // its aggregate; contact couplings between distant nodes are left out of the coarse levels)
    AmgLevel stat;
    stat.n = L0.n;
    stat.rp = srp;
    stat.ci = sci;
    stat.val.assign(sval.begin(), sval.end());
    for (int i = 0; i < L0.n; i++) {
        int ds = -1, df = -1;
        for (int k = srp[i]; k < srp[i + 1]; k++)
            if (sci[k] == i) ds = k;
        for (int k = rp[i]; k < rp[i + 1]; k++)
            if (ci[k] == i) df = k;
        if (ds >= 0 && df >= 0)
            for (int c = 0; c < 9; c++) stat.val[9 * (size_t)ds + c] = val[9 * (size_t)df + c];
    }
    const AmgLevel* src = &stat;
    for (int l = 1; l < nl; l++) {
        amg_galerkin(*src, topo.agg[l - 1], topo.sizes[l], M.L[l]);
        amg_dinv(M.L[l]);
        src = &M.L[l];
    }
    const AmgLevel& C = nl > 1 ? M.L[nl - 1] : stat;
    const int m = 3 * C.n;
    vector<double> dense((size_t)m * m, 0.0);
    for (int i = 0; i < C.n; i++)
        for (int k = C.rp[i]; k < C.rp[i + 1]; k++)
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) dense[(size_t)(3 * i + a) * m + 3 * C.ci[k] + b] = C.val[9 * (size_t)k + 3 * a + b];
    dense_inverse(dense, m, M.coarseInv);
}

inline void amg_spmv(const AmgLevel& a, const vector<double>& x, vector<double>& y) {
    y.assign(3 * (size_t)a.n, 0.0);
    for (int i = 0; i < a.n; i++)
        for (int k = a.rp[i]; k < a.rp[i + 1]; k++)
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++) y[3 * i + r] += a.val[9 * (size_t)k + 3 * r + c] * x[3 * a.ci[k] + c];
}

inline void amg_smooth(const AmgLevel& a, const vector<double>& res, vector<double>& x) {
    for (int i = 0; i < a.n; i++)
        for (int r = 0; r < 3; r++) {
            double s = 0.0;
            for (int c = 0; c < 3; c++) s += a.Dinv[9 * (size_t)i + 3 * r + c] * res[3 * i + c];
            x[3 * i + r] += AMG_OMEGA * s;
        }
}

inline void amg_vcycle(const Amg& M, int l, const vector<double>& b, vector<double>& x) {
    const AmgLevel& a = M.L[l];
    const int nl = (int)M.L.size();
    x.assign(3 * (size_t)a.n, 0.0);
    if (l == nl - 1) {
        const int m = 3 * a.n;
        for (int i = 0; i < m; i++) {
            double s = 0.0;
            for (int k = 0; k < m; k++) s += M.coarseInv[(size_t)i * m + k] * b[k];
            x[i] = s;
        }
        return;
    }
    amg_smooth(a, b, x);
    vector<double> ax, r(b.size());
    amg_spmv(a, x, ax);
    for (size_t i = 0; i < r.size(); i++) r[i] = b[i] - ax[i];
    const vector<int>& agg = M.topo.agg[l];
    vector<double> bc(3 * (size_t)M.L[l + 1].n, 0.0), ec;
    for (int i = 0; i < a.n; i++)
        for (int c = 0; c < 3; c++) bc[3 * agg[i] + c] += r[3 * i + c];
    amg_vcycle(M, l + 1, bc, ec);
    for (int i = 0; i < a.n; i++)
        for (int c = 0; c < 3; c++) x[3 * i + c] += ec[3 * agg[i] + c];
    amg_spmv(a, x, ax);
    for (size_t i = 0; i < r.size(); i++) r[i] = b[i] - ax[i];
    amg_smooth(a, r, x);
}

}  // namespace orc
