// oracle/oracle.cpp — C entry points of the CPU oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The product
// (libclothsim.so) never links or calls it.
//
// The entry points mirror include/clothsim.h so a parity test drives the
// oracle and the HIP product through the same call sequence.  A context holds
// either an FP32 scene (the reference's FP32 CPU path restated) or an FP64 scene
// (the FP64 oracle).  CCD roots are FP64 in both.
//
// Parity status of each function (how it is pinned) is listed in the header of
// tests/test_oracle.py.
#include <memory>

#include "or_collide.hpp"

using namespace orc;

namespace {

struct Handle {
    int fp64 = 0;
    std::unique_ptr<Scene<float>> s32;
    std::unique_ptr<Scene<double>> s64;
};

template <class T> Scene<T>& scene(Handle* h);
template <> Scene<float>& scene<float>(Handle* h) { return *h->s32; }
template <> Scene<double>& scene<double>(Handle* h) { return *h->s64; }

template <class F> int dispatch(void* hv, F&& f) {
    Handle* h = static_cast<Handle*>(hv);
    if (!h) return CS_E_ARG;
    if (h->fp64) return f(*h->s64);
    return f(*h->s32);
}

template <class T> int add_mesh(Scene<T>& S, const cs_mesh_in* m, const cs_material_in* mat, int* id) {
    if (!m) return CS_E_ARG;
    bool cloth = mat != nullptr;
    if (S.started) return CS_E_STATE;
    if (cloth && S.nObstacles > 0) return CS_E_STATE;  // cloths are created before obstacles (src/Simulator.cu:85-100)
    int matIndex = -1;
    if (cloth) {
        S.materials.push_back(Material());
        material_build(*mat, S.materials.back());
        matIndex = (int)S.materials.size() - 1;
    }
    int st = scene_add_mesh(S, *m, cloth, matIndex);
    if (st != CS_OK) return st;
    const MeshRange& R = S.meshes.back();
    update_face_geometry(S, R);  // Mesh::load tail (src/Mesh.cu:449-451)
    update_node_geometry(S, R);
    if (cloth) { if (id) *id = S.nCloths; S.nCloths++; }
    else { if (id) *id = S.nObstacles; S.nObstacles++; }
    return CS_OK;
}

template <class T> const MeshRange* mesh_of(Scene<T>& S, int kind, int id) {
    int m = kind == CS_MESH_CLOTH ? id : S.nCloths + id;
    if (id < 0 || (kind == CS_MESH_CLOTH && id >= S.nCloths) || (kind == CS_MESH_OBSTACLE && id >= S.nObstacles)) return nullptr;
    return &S.meshes[m];
}

template <class T> void prox_out(const Prox<T>& p, cs_proximity_rec& r) {
    for (int k = 0; k < 4; k++) { r.nodes[k] = p.nodes[k]; r.w[k] = (float)p.w[k]; }
    for (int c = 0; c < 3; c++) r.n[c] = (float)p.n[c];
    r.mu = (float)p.mu; r.stiffness = (float)p.stiffness;
}

void impact_out(const Impact& i, cs_impact_rec& r) {
    r.type = i.type;
    for (int k = 0; k < 4; k++) { r.nodes[k] = i.nodes[k]; r.w[k] = i.w[k]; }
    for (int c = 0; c < 3; c++) r.n[c] = i.n[c];
    r.t = i.t;
}

}  // namespace

extern "C" {

void* or_create(const cs_params* p, int fp64) {
    if (!p) return nullptr;
    Handle* h = new Handle();
    h->fp64 = fp64;
    if (fp64) { h->s64.reset(new Scene<double>()); h->s64->P = *p; }
    else { h->s32.reset(new Scene<float>()); h->s32->P = *p; }
    return h;
}

void or_destroy(void* h) { delete static_cast<Handle*>(h); }

int or_set_params(void* h, const cs_params* p) {
    return dispatch(h, [&](auto& S) { S.P = *p; return (int)CS_OK; });
}

int or_add_cloth(void* h, const cs_mesh_in* m, const cs_material_in* mat, int* id) {
    if (!mat) return CS_E_ARG;
    return dispatch(h, [&](auto& S) { return add_mesh(S, m, mat, id); });
}

int or_add_obstacle(void* h, const cs_mesh_in* m, int* id) {
    return dispatch(h, [&](auto& S) { return add_mesh(S, m, nullptr, id); });
}

int or_add_moving_handles(void* h, int cloth, const int* nodes, int n, int motion) {
    return dispatch(h, [&](auto& S) {
        if (cloth < 0 || cloth >= S.nCloths || n < 0) return (int)CS_E_ARG;
        if (motion < -1 || motion >= (int)S.motions.size()) return (int)CS_E_ARG;
        const MeshRange& R = S.meshes[cloth];
        for (int i = 0; i < n; i++)
            if (nodes[i] < 0 || nodes[i] >= R.nn) return (int)CS_E_ARG;
        // anchor = T^-1 (x) for a moving handle (src/Cloth.cu:653-665)
        Tf inv = motion >= 0 ? tf_inverse(S.motion_now(motion)) : tf_identity();
        for (int i = 0; i < n; i++) {
            typename std::decay_t<decltype(S)>::Handle hd;
            hd.cloth = cloth; hd.node = R.n0 + nodes[i]; hd.motion = motion;
            hd.pos = motion >= 0 ? tf_point(inv, S.x[hd.node]) : S.x[hd.node];
            S.handles.push_back(hd);
        }
        return (int)CS_OK;
    });
}

int or_add_handles(void* h, int cloth, const int* nodes, int n) { return or_add_moving_handles(h, cloth, nodes, n, -1); }

static Tf tf_in(const float* k) { Tf t; for (int i = 0; i < 8; i++) t[i] = k[i]; return t; }

// keys: 8 floats per key (scale, translate xyz, rotate s x y z)
int or_add_motion(void* h, const float* times, const float* keys, int n, int* id) {
    std::vector<Tf> k(n > 0 ? n : 0);
    for (int i = 0; i < n; i++) k[i] = tf_in(keys + 8 * i);
    return dispatch(h, [&](auto& S) {
        Motion M;
        if (!M.build(times, k.data(), n)) return (int)CS_E_ARG;
        S.motions.push_back(M);
        if (id) *id = (int)S.motions.size() - 1;
        return (int)CS_OK;
    });
}

int or_motion_eval(const float* times, const float* keys, int n, float time, float* out) {
    std::vector<Tf> k(n > 0 ? n : 0);
    for (int i = 0; i < n; i++) k[i] = tf_in(keys + 8 * i);
    Motion M;
    if (!M.build(times, k.data(), n)) return CS_E_ARG;
    Tf r = M.at(time);
    for (int i = 0; i < 8; i++) out[i] = r[i];
    return CS_OK;
}

// op 0: T(p); 1: T^-1 (p); 2: A(B(p)) through the composed transformation A * B (b = B)
int or_transform_points(const float* a, const float* b, int op, const double* p, int n, double* out) {
    Tf A = tf_in(a);
    if (op == 1) A = tf_inverse(A);
    if (op == 2) A = tf_compose(A, tf_in(b));
    for (int i = 0; i < n; i++) {
        V3<double> q = tf_point(A, mk3<double>(p[3 * i], p[3 * i + 1], p[3 * i + 2]));
        for (int c = 0; c < 3; c++) out[3 * i + c] = q[c];
    }
    return CS_OK;
}

int or_transform_compose(const float* a, const float* b, float* out) {
    Tf r = tf_compose(tf_in(a), tf_in(b));
    for (int i = 0; i < 8; i++) out[i] = r[i];
    return CS_OK;
}

int or_transform_inverse(const float* a, float* out) {
    Tf r = tf_inverse(tf_in(a));
    for (int i = 0; i < 8; i++) out[i] = r[i];
    return CS_OK;
}

int or_set_obstacle_motion(void* h, int obstacle, int motion) {
    return dispatch(h, [&](auto& S) {
        if (S.started || obstacle < 0 || obstacle >= S.nObstacles) return (int)CS_E_ARG;
        if (motion < 0 || motion >= (int)S.motions.size()) return (int)CS_E_ARG;
        MeshRange& R = S.meshes[S.nCloths + obstacle];
        if (R.motion >= 0) return (int)CS_E_STATE;
        if (S.base.size() < S.x.size()) S.base.resize(S.x.size());
        // Obstacle::initialize + transform (src/Obstacle.cu:21-64)
        Tf T = S.motion_now(motion);
        for (int i = R.n0; i < R.n0 + R.nn; i++) {
            S.base[i] = S.x[i];
            S.x[i] = S.x0[i] = tf_point(T, S.base[i]);
        }
        R.motion = motion;
        update_face_geometry(S, R);
        update_node_geometry(S, R);
        return (int)CS_OK;
    });
}

// Simulator::resume: nSteps, then Obstacle::transform at the new time (src/Simulator.cu:1045-1053)
int or_set_step_count(void* h, int steps) {
    return dispatch(h, [&](auto& S) {
        S.steps = steps;
        for (int o = S.nCloths; o < (int)S.meshes.size(); o++) {
            const MeshRange& R = S.meshes[o];
            if (R.motion < 0) continue;
            Tf T = S.motion_now(R.motion);
            for (int i = R.n0; i < R.n0 + R.nn; i++) S.x[i] = S.x0[i] = tf_point(T, S.base[i]);
            update_face_geometry(S, R);
            update_node_geometry(S, R);
        }
        return (int)CS_OK;
    });
}

int or_physics_step(void* h, int direct, cs_stats* st) {
    return dispatch(h, [&](auto& S) { return physics_step(S, direct, st); });
}

int or_collision_step(void* h, cs_stats* st) {
    return dispatch(h, [&](auto& S) { return collision_step(S, st); });
}

int or_step(void* h, int direct, cs_stats* st) {
    return dispatch(h, [&](auto& S) {
        int a = physics_step(S, direct, st);
        if (a != CS_OK && a != CS_E_NOT_CONVERGED) return a;
        if (S.P.flags & CS_FLAG_NO_COLLISION) return a;
        int b = collision_step(S, st);
        return b != CS_OK ? b : a;
    });
}

int or_get_counts(void* h, int kind, int id, int* nn, int* nf, int* ne) {
    return dispatch(h, [&](auto& S) {
        const MeshRange* R = mesh_of(S, kind, id);
        if (!R) return (int)CS_E_ARG;
        if (nn) *nn = R->nn;
        if (nf) *nf = R->nf;
        if (ne) *ne = R->ne;
        return (int)CS_OK;
    });
}

int or_get_state(void* h, int kind, int id, float* x, float* x0, float* v) {
    return dispatch(h, [&](auto& S) {
        const MeshRange* R = mesh_of(S, kind, id);
        if (!R) return (int)CS_E_ARG;
        for (int i = 0; i < R->nn; i++)
            for (int c = 0; c < 3; c++) {
                if (x) x[3 * i + c] = (float)S.x[R->n0 + i][c];
                if (x0) x0[3 * i + c] = (float)S.x0[R->n0 + i][c];
                if (v) v[3 * i + c] = (float)S.v[R->n0 + i][c];
            }
        return (int)CS_OK;
    });
}

int or_get_state64(void* h, int kind, int id, double* x, double* x0, double* v) {
    return dispatch(h, [&](auto& S) {
        const MeshRange* R = mesh_of(S, kind, id);
        if (!R) return (int)CS_E_ARG;
        for (int i = 0; i < R->nn; i++)
            for (int c = 0; c < 3; c++) {
                if (x) x[3 * i + c] = (double)S.x[R->n0 + i][c];
                if (x0) x0[3 * i + c] = (double)S.x0[R->n0 + i][c];
                if (v) v[3 * i + c] = (double)S.v[R->n0 + i][c];
            }
        return (int)CS_OK;
    });
}

// Sets node state and refreshes face and node geometry of that mesh (as Mesh::load does).
int or_set_state(void* h, int kind, int id, const float* x, const float* x0, const float* v) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        const MeshRange* R = mesh_of(S, kind, id);
        if (!R) return (int)CS_E_ARG;
        for (int i = 0; i < R->nn; i++)
            for (int c = 0; c < 3; c++) {
                if (x) S.x[R->n0 + i][c] = (T)x[3 * i + c];
                if (x0) S.x0[R->n0 + i][c] = (T)x0[3 * i + c];
                if (v) S.v[R->n0 + i][c] = (T)v[3 * i + c];
            }
        update_face_geometry(S, *R);
        update_node_geometry(S, *R);
        return (int)CS_OK;
    });
}

int or_get_geometry(void* h, int kind, int id, float* normals, float* area, float* mass) {
    return dispatch(h, [&](auto& S) {
        const MeshRange* R = mesh_of(S, kind, id);
        if (!R) return (int)CS_E_ARG;
        for (int i = 0; i < R->nn; i++) {
            for (int c = 0; c < 3; c++) if (normals) normals[3 * i + c] = (float)S.nrm[R->n0 + i][c];
            if (area) area[i] = (float)S.area[R->n0 + i];
            if (mass) mass[i] = (float)S.mass[R->n0 + i];
        }
        return (int)CS_OK;
    });
}

// Edge topology of one mesh (global ids): nodes[2], opposite nodes[2] (-1 = none), adjacent faces[2].
int or_get_edges(void* h, int kind, int id, int* nodes, int* opp, int* adj) {
    return dispatch(h, [&](auto& S) {
        const MeshRange* R = mesh_of(S, kind, id);
        if (!R) return (int)CS_E_ARG;
        for (int e = 0; e < R->ne; e++) {
            int g = R->e0 + e;
            for (int s = 0; s < 2; s++) {
                if (nodes) nodes[2 * e + s] = S.en[g][s];
                if (opp) opp[2 * e + s] = S.eopp[g][s] >= 0 ? S.vnode[S.eopp[g][s]] : -1;
                if (adj) adj[2 * e + s] = S.eadj[g][s];
            }
        }
        return (int)CS_OK;
    });
}

int or_material_table(const cs_material_in* mat, float* table, float* density, float* bending) {
    if (!mat) return CS_E_ARG;
    Material m;
    material_build(*mat, m);
    if (table) std::memcpy(table, m.table.data(), sizeof(float) * m.table.size());
    if (density) *density = m.density;
    if (bending) std::memcpy(bending, m.bend, sizeof(m.bend));
    return CS_OK;
}

// Stretching stiffness lookup of one cloth's material at strain G (FP64 scene: double math).
int or_stretching_stiffness(void* h, int cloth, double g00, double g01, double g11, double* k) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        if (cloth < 0 || cloth >= S.nCloths) return (int)CS_E_ARG;
        Vec<T, 4> kk = stretching_stiffness<T>(S.materials[S.meshes[cloth].material], (T)g00, (T)g01, (T)g11);
        for (int c = 0; c < 4; c++) k[c] = (double)kk[c];
        return (int)CS_OK;
    });
}

// Element force / Jacobian of one face (stretching) or one interior edge (bending).
int or_stretching_force(void* h, int face, double* f9, double* J81) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        if (face < 0 || face >= (int)S.fn.size() || !S.ffree[face]) return (int)CS_E_ARG;
        const Material& m = S.materials[S.meshes[S.nodeMesh[S.fn[face][0]]].material];
        V9<T> f; M9<T> J;
        stretching_force(S, face, m, f, J);
        for (int i = 0; i < 9; i++) { f9[i] = (double)f[i]; for (int j = 0; j < 9; j++) J81[9 * i + j] = (double)J(i, j); }
        return (int)CS_OK;
    });
}

int or_bending_force(void* h, int edge, double* f12, double* J144, double* angle) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        if (edge < 0 || edge >= (int)S.en.size() || S.edgeBoundary(edge)) return (int)CS_E_ARG;
        const MeshRange& R = S.meshes[S.nodeMesh[S.en[edge][0]]];
        if (!R.cloth) return (int)CS_E_ARG;
        V12<T> f; M12<T> J;
        bending_force(S, edge, S.materials[R.material], f, J);
        for (int i = 0; i < 12; i++) { f12[i] = (double)f[i]; for (int j = 0; j < 12; j++) J144[12 * i + j] = (double)J(i, j); }
        if (angle) *angle = (double)edge_angle(S, edge);
        return (int)CS_OK;
    });
}

// Sorted LBVH keys (morton << 32 | local face) of one mesh (global mesh index).
int or_lbvh_keys(void* h, int mesh, int ccd, uint64_t* keys) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        if (mesh < 0 || mesh >= (int)S.meshes.size()) return (int)CS_E_ARG;
        BVH<T> B;
        bvh_build(S, mesh, ccd != 0, B);
        for (int i = 0; i < B.nLeaves; i++) keys[i] = B.keys[i];
        return (int)CS_OK;
    });
}

// Candidate face pairs of Simulator::traverse; brute != 0 enumerates all leaf pairs instead of the tree.
int or_lbvh_pairs(void* h, int ccd, float thickness, int brute, int* pairs, int cap, int* n) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        vector<BVH<T>> bvhs;
        build_all(S, ccd != 0, bvhs);
        long cnt = 0;
        auto emit = [&](int a, int b) {
            if (cnt < cap && pairs) { pairs[2 * cnt] = a; pairs[2 * cnt + 1] = b; }
            cnt++;
        };
        if (brute) traverse_bruteforce(S, bvhs, (T)thickness, emit);
        else traverse_all(S, bvhs, (T)thickness, emit);
        *n = (int)cnt;
        return cnt > cap ? (int)CS_E_CAPACITY : (int)CS_OK;
    });
}

// brute != 0: over every face pair of the traversed mesh pairs instead of the tree's pairs
int or_find_proximities_brute(void* h, int brute, cs_proximity_rec* out, int cap, int* n) {
    return dispatch(h, [&](auto& S) {
        auto prox = find_proximities(S, brute != 0);
        *n = (int)prox.size();
        if ((int)prox.size() > cap) return (int)CS_E_CAPACITY;
        for (size_t i = 0; i < prox.size(); i++) prox_out(prox[i], out[i]);
        return (int)CS_OK;
    });
}

int or_find_proximities(void* h, cs_proximity_rec* out, int cap, int* n) {
    return or_find_proximities_brute(h, 0, out, cap, n);
}

// Assembled system of one cloth on the current state (proximities detected
// first unless CS_FLAG_NO_PROXIMITY).  Values are returned as double.
int or_assemble(void* h, int cloth, int* rowPtr, int* col, double* vals, int cap, int* nnzb, double* b) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        if (cloth < 0 || cloth >= S.nCloths) return (int)CS_E_ARG;
        vector<Prox<T>> prox;
        if (!(S.P.flags & CS_FLAG_NO_PROXIMITY)) prox = find_proximities(S);
        System<T> sys;
        assemble(S, cloth, prox, sys);
        vector<int> rp, ci; vector<T> val;
        sys.to_bsr(rp, ci, val);
        *nnzb = (int)ci.size();
        if (!rowPtr) return (int)CS_OK;
        if ((int)ci.size() > cap) return (int)CS_E_CAPACITY;
        for (size_t i = 0; i < rp.size(); i++) rowPtr[i] = rp[i];
        for (size_t i = 0; i < ci.size(); i++) col[i] = ci[i];
        for (size_t i = 0; i < val.size(); i++) vals[i] = (double)val[i];
        for (size_t i = 0; i < sys.b.size(); i++) b[i] = (double)sys.b[i];
        return (int)CS_OK;
    });
}

// The aggregation multigrid preconditioner of cloth `cloth` at the current state:
// level sizes (sizes != NULL: up to cap entries, *nlevels set) and, when r != NULL,
// z = one V(1,1) cycle applied to r (3 * n_nodes doubles).
int or_amg(void* h, int cloth, int* sizes, int cap, int* nlevels, const double* r, double* z) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        if (cloth < 0 || cloth >= S.nCloths) return (int)CS_E_ARG;
        AmgTopo topo;
        amg_cloth_topology(S, cloth, topo);
        *nlevels = (int)topo.sizes.size();
        if (sizes)
            for (int l = 0; l < *nlevels && l < cap; l++) sizes[l] = topo.sizes[l];
        if (!r) return (int)CS_OK;
        vector<Prox<T>> prox;
        if (!(S.P.flags & CS_FLAG_NO_PROXIMITY)) prox = find_proximities(S);
        System<T> sys, ss;
        assemble(S, cloth, prox, sys);
        assemble(S, cloth, vector<Prox<T>>(), ss);
        vector<int> rp, ci, srp, sci;
        vector<T> val, sval;
        sys.to_bsr(rp, ci, val);
        ss.to_bsr(srp, sci, sval);
        Amg M;
        amg_setup(topo, rp, ci, val, srp, sci, sval, M);
        vector<double> bi(r, r + 3 * (size_t)topo.sizes[0]), xo;
        amg_vcycle(M, 0, bi, xo);
        for (size_t i = 0; i < xo.size(); i++) z[i] = xo[i];
        return (int)CS_OK;
    });
}

// Block-Jacobi PCG (direct == 0), the banded LL^T (direct == 1) or the nested-dissection
// multifrontal LL^T (direct == 2) on a caller BSR system.
int or_solve_bsr(int n, const int* rowPtr, const int* col, const double* vals, const double* b, int direct,
                 int fp64, double tol, int maxIters, double* x, int* iters, double* relres) {
    auto run = [&](auto zero) {
        using T = decltype(zero);
        vector<int> rp(rowPtr, rowPtr + n + 1), ci(col, col + rowPtr[n]);
        vector<T> val(9 * (size_t)rowPtr[n]), bb(3 * (size_t)n), xx;
        for (size_t i = 0; i < val.size(); i++) val[i] = (T)vals[i];
        for (size_t i = 0; i < bb.size(); i++) bb[i] = (T)b[i];
        int it = 0; T rr = T(0);
        if (direct == 2) { if (!direct_solve_nd(rp, ci, val, bb, xx)) return (int)CS_E_NOT_CONVERGED; }
        else if (direct) { if (!direct_solve(rp, ci, val, bb, xx)) return (int)CS_E_NOT_CONVERGED; }
        else it = pcg_block_jacobi(rp, ci, val, bb, (T)tol, maxIters, xx, &rr);
        for (size_t i = 0; i < xx.size(); i++) x[i] = (double)xx[i];
        if (iters) *iters = it;
        if (relres) *relres = (double)rr;
        return (int)CS_OK;
    };
    return fp64 ? run(0.0) : run(0.0f);
}

int or_find_impacts(void* h, float thickness, cs_impact_rec* out, int cap, int* n) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        vector<BVH<T>> bvhs;
        build_all(S, true, bvhs);
        auto imps = find_impacts_with(S, bvhs, (T)thickness);
        *n = (int)imps.size();
        if ((int)imps.size() > cap) return (int)CS_E_CAPACITY;
        for (size_t i = 0; i < imps.size(); i++) impact_out(imps[i], out[i]);
        return (int)CS_OK;
    });
}

// The impacts of the reference's FP32 CCD (float arithmetic, its whole-line Newton root finder,
// src/CollisionHelper.cu:3-109) on the same candidates: for the FP32-vs-FP64 divergence measure.
int or_find_impacts_ref32(void* h, float thickness, cs_impact_rec* out, int cap, int* n) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        vector<BVH<T>> bvhs;
        build_all(S, true, bvhs);
        auto imps = find_impacts_with(S, bvhs, (T)thickness, true);
        *n = (int)imps.size();
        if ((int)imps.size() > cap) return (int)CS_E_CAPACITY;
        for (size_t i = 0; i < imps.size(); i++) impact_out(imps[i], out[i]);
        return (int)CS_OK;
    });
}

int or_independent_impacts(void* h, const cs_impact_rec* in, int n, int deform, cs_impact_rec* out, int* nout) {
    return dispatch(h, [&](auto& S) {
        vector<Impact> v(n);
        for (int i = 0; i < n; i++) {
            v[i].type = in[i].type; v[i].t = in[i].t;
            for (int k = 0; k < 4; k++) { v[i].nodes[k] = in[i].nodes[k]; v[i].w[k] = in[i].w[k]; }
            for (int c = 0; c < 3; c++) v[i].n[c] = in[i].n[c];
            v[i].key = 0;
        }
        auto sel = independent_impacts(S, v, deform);
        *nout = (int)sel.size();
        for (size_t i = 0; i < sel.size(); i++) impact_out(sel[i], out[i]);
        return (int)CS_OK;
    });
}

// CollisionOptimization(impacts, thickness, deform, obstacleMass) + Optimization::solve on the
// scene's current x (start) and x1 (objective anchor): src/optimization/CollisionOptimization.cu:37-145,
// src/optimization/Optimization.cu:98-145.
int or_collision_optimize(void* h, const cs_impact_rec* in, int n, float thickness, int deform, float obstacleMass,
                          int* iters) {
    return dispatch(h, [&](auto& S) {
        using T = std::decay_t<decltype(S.x[0][0])>;
        vector<Impact> v(n);
        for (int i = 0; i < n; i++) {
            v[i].type = in[i].type; v[i].t = in[i].t;
            for (int k = 0; k < 4; k++) {
                if (in[i].nodes[k] < 0 || in[i].nodes[k] >= (int)S.x.size()) return (int)CS_E_ARG;
                v[i].nodes[k] = in[i].nodes[k]; v[i].w[k] = in[i].w[k];
            }
            for (int c = 0; c < 3; c++) v[i].n[c] = in[i].n[c];
            v[i].key = 0;
        }
        int it = n > 0 ? collision_optimize(S, v, (T)thickness, deform, (T)obstacleMass) : 0;
        if (iters) *iters = it;
        return (int)CS_OK;
    });
}

// The reference's whole-line cubic solver restated in FP64 (src/CollisionHelper.cu:3-60).
int or_solve_cubic(double a, double b, double c, double d, double* roots) { return solve_cubic(a, b, c, d, roots); }

// The isolated root finder on [0, 1] used by the CCD (roots must hold 4 entries).
int or_cubic_roots_unit(double a, double b, double c, double d, double* roots) {
    return cubic_roots_unit(a, b, c, d, roots);
}

}  // extern "C"
