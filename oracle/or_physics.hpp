// oracle/or_physics.hpp — forces, implicit-Euler assembly and solvers of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (see or_math.hpp).  Restates the reference CPU path
// src/Cloth.cu:42-192,676-693 and the force kernels of src/PhysicsHelper.cu.
#pragma once
#include <map>
#include <queue>

#include "or_direct.hpp"
#include "or_scene.hpp"

namespace orc {

using std::array;
using std::vector;

template <class T> using V9 = Vec<T, 9>;
template <class T> using V12 = Vec<T, 12>;
template <class T> using M9 = Mat<T, 9, 9>;
template <class T> using M12 = Mat<T, 12, 12>;

// One proximity record (src/Proximity.cuh:10-21)
template <class T> struct Prox {
    int nodes[4];
    V3<T> n;
    T mu, stiffness, w[4];
};

// ---------------------------------------------------------------- stretching
// src/PhysicsHelper.cu:264-291 and Face::derivative src/Face.cu:109-111
template <class T> void stretching_force(const Scene<T>& S, int f, const Material& mat, V9<T>& F, M9<T>& J) {
    const V3<T>& x0 = S.x[S.fn[f][0]];
    const V3<T>& x1 = S.x[S.fn[f][1]];
    const V3<T>& x2 = S.x[S.fn[f][2]];
    Mat<T, 3, 2> Dx;
    V3<T> d1 = x1 - x0, d2 = x2 - x0;
    for (int i = 0; i < 3; i++) { Dx(i, 0) = d1[i]; Dx(i, 1) = d2[i]; }
    const Mat<T, 2, 2>& Y = S.finv[f];
    Mat<T, 3, 2> Fm = Dx * Y;
    Mat<T, 2, 2> G = T(0.5f) * (Fm.transpose() * Fm - Mat<T, 2, 2>::ident(T(1)));
    Mat<T, 2, 3> D;
    V2<T> r0 = Y.row(0), r1 = Y.row(1);
    V2<T> c0 = -r0 - r1;
    for (int i = 0; i < 2; i++) { D(i, 0) = c0[i]; D(i, 1) = r0[i]; D(i, 2) = r1[i]; }
    Mat<T, 3, 9> Du, Dv;
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < 3; i++) { Du(i, 3 * a + i) = D(0, a); Dv(i, 3 * a + i) = D(1, a); }
    V3<T> fu = Fm.col(0), fv = Fm.col(1);
    Mat<T, 9, 3> DuT = Du.transpose(), DvT = Dv.transpose();
    V9<T> fuu = DuT * fu;
    V9<T> fvv = DvT * fv;
    V9<T> fuv = T(0.5f) * (DuT * fv + DvT * fu);
    Vec<T, 4> k = stretching_stiffness<T>(mat, G(0, 0), G(0, 1), G(1, 1));
    V9<T> grad = k[0] * G(0, 0) * fuu + k[2] * G(1, 1) * fvv + k[1] * (G(0, 0) * fvv + G(1, 1) * fuu) +
                 T(2) * k[3] * G(0, 1) * fuv;
    T g00 = tmax(G(0, 0), T(0)), g11 = tmax(G(1, 1), T(0));
    M9<T> hess = k[0] * (outer(fuu, fuu) + (g00 * DuT) * Du) + k[2] * (outer(fvv, fvv) + (g11 * DvT) * Dv) +
                 k[1] * (outer(fuu, fvv) + (g00 * DvT) * Dv + outer(fvv, fuu) + (g11 * DuT) * Du) +
                 T(2) * k[3] * outer(fuv, fuv);
    T a = S.farea[f];
    F = -a * grad;
    J = -a * hess;
}

// ---------------------------------------------------------------- bending
// src/PhysicsHelper.cu:314-325
template <class T> T point_line_distance(const V3<T>& x, const V3<T>& a, const V3<T>& b) {
    V3<T> e = b - a;
    V3<T> t = x - a;
    V3<T> r = e * e.dot(t) / e.norm2();
    return (t - r).norm();
}
template <class T> V2<T> line_weights(const V3<T>& x, const V3<T>& a, const V3<T>& b) {
    V3<T> e = b - a;
    T t = e.dot(x - a) / e.norm2();
    return mk2<T>(T(1) - t, t);
}
// Edge::angle (src/Edge.cu:104-114)
template <class T> T edge_angle(const Scene<T>& S, int e) {
    if (S.edgeBoundary(e)) return T(0);
    V3<T> ed = (S.x[S.en[e][0]] - S.x[S.en[e][1]]).normalized();
    const V3<T>& n0 = S.fnrm[S.eadj[e][0]];
    const V3<T>& n1 = S.fnrm[S.eadj[e][1]];
    T sine = ed.dot(cross(n0, n1));
    T cosine = n0.dot(n1);
    return std::atan2(sine, cosine);
}
// src/PhysicsHelper.cu:327-353
template <class T> void bending_force(const Scene<T>& S, int e, const Material& mat, V12<T>& F, M12<T>& J) {
    const V3<T>& x0 = S.x[S.en[e][0]];
    const V3<T>& x1 = S.x[S.en[e][1]];
    const V3<T>& x2 = S.x[S.vnode[S.eopp[e][0]]];
    const V3<T>& x3 = S.x[S.vnode[S.eopp[e][1]]];
    int a0 = S.eadj[e][0], a1 = S.eadj[e][1];
    const V3<T>& n0 = S.fnrm[a0];
    const V3<T>& n1 = S.fnrm[a1];
    T length = (x1 - x0).norm();
    T angle = edge_angle(S, e);
    T area = S.farea[a0] + S.farea[a1];
    T h0 = point_line_distance(x2, x0, x1);
    T h1 = point_line_distance(x3, x0, x1);
    V2<T> w0 = line_weights(x2, x0, x1);
    V2<T> w1 = line_weights(x3, x0, x1);
    V3<T> g0 = -w0[0] * n0 / h0 - w1[0] * n1 / h1;
    V3<T> g1 = -w0[1] * n0 / h0 - w1[1] * n1 / h1;
    V3<T> g2 = n0 / h0;
    V3<T> g3 = n1 / h1;
    V12<T> dtheta;
    for (int i = 0; i < 3; i++) { dtheta[i] = g0[i]; dtheta[3 + i] = g1[i]; dtheta[6 + i] = g2[i]; dtheta[9 + i] = g3[i]; }
    V2<T> du0 = S.u[S.ev[e][0][1]] - S.u[S.ev[e][0][0]];
    V2<T> du1 = S.u[S.ev[e][1][1]] - S.u[S.ev[e][1][0]];
    T k0 = bending_stiffness<T>(mat, length, angle, area, du0[0], du0[1]);
    T k1 = bending_stiffness<T>(mat, length, angle, area, du1[0], du1[1]);
    T coefficient = T(-0.25f) * tmin(k0, k1) * tsqr(length) / area;
    F = coefficient * angle * dtheta;
    J = coefficient * outer(dtheta, dtheta);
}

// ---------------------------------------------------------------- proximity forces
// src/PhysicsHelper.cu:398-406
template <class T> void impulse_force(const Prox<T>& p, T d, T thickness, V12<T>& F, M12<T>& J) {
    V12<T> N;
    for (int i = 0; i < 4; i++) for (int c = 0; c < 3; c++) N[3 * i + c] = p.w[i] * p.n[c];
    F = T(0.5f) * p.stiffness / thickness * tsqr(d) * N;
    J = -p.stiffness / thickness * d * outer(N, N);
}
// src/PhysicsHelper.cu:408-436
template <class T> void friction_force(const Scene<T>& S, const Prox<T>& p, T d, T thickness, T dt, V12<T>& F, M12<T>& J) {
    T Fn = T(0.5f) * p.stiffness / thickness * tsqr(d);
    V3<T> vv;
    T invMass = T(0);
    for (int i = 0; i < 4; i++) {
        vv += p.w[i] * S.v[p.nodes[i]];
        if (S.nodeFree(p.nodes[i])) invMass += tsqr(p.w[i]) / S.mass[p.nodes[i]];
    }
    Mat<T, 3, 3> Tm = Mat<T, 3, 3>::ident(T(1)) - outer(p.n, p.n);
    V3<T> vt = Tm * vv;
    T f_v = tmin(p.mu * Fn / vt.norm(), T(1) / (dt * invMass));
    V12<T> Vv;
    for (int i = 0; i < 4; i++) for (int c = 0; c < 3; c++) Vv[3 * i + c] = p.w[i] * vt[c];
    F = -f_v * Vv;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            Mat<T, 3, 3> Tij = -f_v * p.w[i] * p.w[j] * Tm;
            for (int k = 0; k < 3; k++) for (int h = 0; h < 3; h++) J(3 * i + k, 3 * j + h) = Tij(k, h);
        }
}

// ---------------------------------------------------------------- system
// Sparse 3x3-block matrix keyed by (row, col) with scalar accumulation order
// equal to the reference CPU's coeffRef += sequence (src/Cloth.cu:42-76).
template <class T> struct System {
    int n = 0;  // nodes
    std::unordered_map<uint64_t, array<T, 9>> A;
    vector<T> b;
    array<T, 9>& blk(int r, int c) {
        uint64_t key = (uint64_t)r * (uint64_t)n + (uint64_t)c;
        auto it = A.find(key);
        if (it == A.end()) { array<T, 9> z; z.fill(T(0)); it = A.emplace(key, z).first; }
        return it->second;
    }
    template <int K> void add(const Mat<T, 3 * K, 3 * K>& B, const Vec<T, 3 * K>& bb, const int (&idx)[K]) {
        for (int i = 0; i < K; i++) {
            int xr = idx[i];
            if (xr > -1) {
                for (int j = 0; j < K; j++) {
                    int yc = idx[j];
                    if (yc > -1) {
                        array<T, 9>& a = blk(xr, yc);
                        for (int k = 0; k < 3; k++) for (int h = 0; h < 3; h++) a[3 * k + h] += B(3 * i + k, 3 * j + h);
                    }
                }
                for (int j = 0; j < 3; j++) b[3 * xr + j] += bb[3 * i + j];
            }
        }
    }
    // sorted BSR view
    void to_bsr(vector<int>& rowPtr, vector<int>& col, vector<T>& val) const {
        vector<std::pair<uint64_t, const array<T, 9>*>> items;
        items.reserve(A.size());
        for (auto& kv : A) items.push_back({kv.first, &kv.second});
        std::sort(items.begin(), items.end(), [](const std::pair<uint64_t, const array<T, 9>*>& p, const std::pair<uint64_t, const array<T, 9>*>& q) { return p.first < q.first; });
        rowPtr.assign(n + 1, 0);
        col.resize(items.size());
        val.resize(9 * items.size());
        for (size_t k = 0; k < items.size(); k++) {
            int r = (int)(items[k].first / (uint64_t)n);
            col[k] = (int)(items[k].first % (uint64_t)n);
            rowPtr[r + 1]++;
            for (int c = 0; c < 9; c++) val[9 * k + c] = (*items[k].second)[c];
        }
        for (int r = 0; r < n; r++) rowPtr[r + 1] += rowPtr[r];
    }
};

// Cloth::physicsStep CPU assembly (src/Cloth.cu:78-192,676-683): mass, gravity,
// wind, stretching, bending, handles, proximities — in that order.
template <class T> void assemble(const Scene<T>& S, int cloth, const vector<Prox<T>>& prox, System<T>& sys) {
    const MeshRange& R = S.meshes[cloth];
    const Material& mat = S.materials[R.material];
    const cs_params& P = S.P;
    T dt = T(P.dt);
    sys.n = R.nn;
    sys.A.clear();
    sys.A.reserve(16 * (size_t)R.nn);
    sys.b.assign(3 * R.nn, T(0));
    // initializeForces (src/Cloth.cu:78-92)
    for (int i = 0; i < R.nn; i++) {
        array<T, 9>& a = sys.blk(i, i);
        for (int j = 0; j < 3; j++) a[4 * j] = S.mass[R.n0 + i];
    }
    // addExternalForces (src/Cloth.cu:94-126)
    V3<T> grav = mk3<T>(P.gravity[0], P.gravity[1], P.gravity[2]);
    for (int i = 0; i < R.nn; i++) {
        V3<T> g = dt * S.mass[R.n0 + i] * grav;
        for (int j = 0; j < 3; j++) sys.b[3 * i + j] += g[j];
    }
    V3<T> wv = mk3<T>(P.wind_velocity[0], P.wind_velocity[1], P.wind_velocity[2]);
    T wdens = T(P.wind_density), wdrag = T(P.wind_drag);
    for (int f = R.f0; f < R.f0 + R.nf; f++) {
        const V3<T>& n = S.fnrm[f];
        T area = S.farea[f];
        int a = S.fn[f][0], bnode = S.fn[f][1], c = S.fn[f][2];
        V3<T> average = (S.v[a] + S.v[bnode] + S.v[c]) / T(3);
        V3<T> relative = wv - average;
        T vn = n.dot(relative);
        V3<T> vt = relative - vn * n;
        V3<T> force = area * (wdens * tabs(vn) * vn * n + wdrag * vt) / T(3);
        V3<T> fw = dt * force;
        for (int i = 0; i < 3; i++) {
            sys.b[3 * (a - R.n0) + i] += fw[i];
            sys.b[3 * (bnode - R.n0) + i] += fw[i];
            sys.b[3 * (c - R.n0) + i] += fw[i];
        }
    }
    // addInternalForces (src/Cloth.cu:128-159)
    for (int f = R.f0; f < R.f0 + R.nf; f++) {
        int idx[3];
        V9<T> vel;
        for (int k = 0; k < 3; k++) {
            idx[k] = S.fn[f][k] - R.n0;
            for (int c = 0; c < 3; c++) vel[3 * k + c] = S.v[S.fn[f][k]][c];
        }
        V9<T> Fv; M9<T> J;
        stretching_force(S, f, mat, Fv, J);
        sys.template add<3>(-dt * dt * J, dt * (Fv + dt * J * vel), idx);
    }
    for (int e = R.e0; e < R.e0 + R.ne; e++) {
        if (S.edgeBoundary(e)) continue;
        int nodes[4] = {S.en[e][0], S.en[e][1], S.vnode[S.eopp[e][0]], S.vnode[S.eopp[e][1]]};
        int idx[4];
        V12<T> vel;
        for (int k = 0; k < 4; k++) {
            idx[k] = nodes[k] - R.n0;
            for (int c = 0; c < 3; c++) vel[3 * k + c] = S.v[nodes[k]][c];
        }
        V12<T> Fv; M12<T> J;
        bending_force(S, e, mat, Fv, J);
        sys.template add<4>(-dt * dt * J, dt * (Fv + dt * J * vel), idx);
    }
    // addHandleForces (src/Cloth.cu:161-173)
    T hk = T(P.handle_stiffness);
    for (const auto& h : S.handles) {
        if (h.cloth != cloth) continue;
        int i = h.node - R.n0;
        array<T, 9>& a = sys.blk(i, i);
        for (int j = 0; j < 3; j++) a[4 * j] += dt * dt * hk;
        // the anchor moves with its motion (src/PhysicsHelper.cu:386)
        V3<T> pos = h.motion >= 0 ? tf_point(S.motion_now(h.motion), h.pos) : h.pos;
        V3<T> fh = dt * ((pos - S.x[h.node]) - dt * S.v[h.node]) * hk;
        for (int j = 0; j < 3; j++) sys.b[3 * i + j] += fh[j];
    }
    // addProximityForces (src/Cloth.cu:175-192)
    T thickness = T(P.repulsion_thickness);
    for (const Prox<T>& p : prox) {
        V12<T> vel;
        int idx[4];
        for (int k = 0; k < 4; k++) {
            int nd = p.nodes[k];
            idx[k] = (nd >= R.n0 && nd < R.n0 + R.nn) ? nd - R.n0 : -1;
            for (int c = 0; c < 3; c++) vel[3 * k + c] = S.v[nd][c];
        }
        T d = -thickness;
        for (int k = 0; k < 4; k++) d += p.w[k] * p.n.dot(S.x[p.nodes[k]]);
        d = tmax(-d, T(0));
        V12<T> Fv; M12<T> J;
        impulse_force(p, d, thickness, Fv, J);
        sys.template add<4>(-dt * dt * J, dt * (Fv + dt * J * vel), idx);
        friction_force(S, p, d, thickness, dt, Fv, J);
        sys.template add<4>(-dt * J, dt * Fv, idx);
    }
}

// ---------------------------------------------------------------- solvers
template <class T> void inverse3(const T a[9], T out[9]) {
    T c00 = a[4] * a[8] - a[5] * a[7];
    T c01 = a[5] * a[6] - a[3] * a[8];
    T c02 = a[3] * a[7] - a[4] * a[6];
    T det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    T id = T(1) / det;
    out[0] = c00 * id;
    out[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    out[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    out[3] = c01 * id;
    out[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    out[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    out[6] = c02 * id;
    out[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    out[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

// 3x3 block-Jacobi preconditioned CG with a relative residual test.  Same
// iteration as the reference GPU loop (src/Cloth.cu:766-809), except the
// preconditioner is the inverse 3x3 diagonal block and the stop test is
// ||r|| <= tol*||b|| (instead of scalar Jacobi and ||r|| < 1e-6).
// Dot products are summed per fixed 4096-entry chunk and the chunk sums in
// chunk order, so the result does not depend on the OpenMP thread count.
template <class T> T chunked_dot(const vector<T>& p, const vector<T>& q) {
    const long m = (long)p.size(), CH = 4096, nc = (m + CH - 1) / CH;
    vector<T> part(nc);
#pragma omp parallel for schedule(static)
    for (long c = 0; c < nc; c++) {
        T s = T(0);
        for (long i = c * CH; i < std::min(m, (c + 1) * CH); i++) s += p[i] * q[i];
        part[c] = s;
    }
    T s = T(0);
    for (long c = 0; c < nc; c++) s += part[c];
    return s;
}

// PCG with the preconditioner applyM(in, out) (out = M^-1 in).
template <class T, class ApplyM>
int pcg_with(const vector<int>& rowPtr, const vector<int>& col, const vector<T>& val, const vector<T>& b, T tol,
             int maxIters, vector<T>& x, T* relres, ApplyM&& applyM) {
    int n = (int)rowPtr.size() - 1;
    int m = 3 * n;
    vector<T> r(b), z(m), d(m), q(m);
    x.assign(m, T(0));
    T bnorm = std::sqrt(chunked_dot(b, b));
    applyM(r, z);
    d = z;
    T rz = chunked_dot(r, z);
    int it = 0;
    for (; it < maxIters; it++) {
        T rn = std::sqrt(chunked_dot(r, r));
        if (rn <= tol * bnorm) break;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) {
                T s = T(0);
                for (int k = rowPtr[i]; k < rowPtr[i + 1]; k++)
                    for (int c = 0; c < 3; c++) s += val[9 * k + 3 * a + c] * d[3 * col[k] + c];
                q[3 * i + a] = s;
            }
        T alpha = rz / chunked_dot(d, q);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < m; i++) { x[i] += alpha * d[i]; r[i] -= alpha * q[i]; }
        applyM(r, z);
        T rzn = chunked_dot(r, z);
        T beta = rzn / rz;
        rz = rzn;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < m; i++) d[i] = z[i] + beta * d[i];
    }
    if (relres) *relres = bnorm > T(0) ? std::sqrt(chunked_dot(r, r)) / bnorm : T(0);
    return it;
}

// 3x3 block-Jacobi preconditioner of a BSR matrix
template <class T> void block_jacobi(const vector<int>& rowPtr, const vector<int>& col, const vector<T>& val,
                                     vector<T>& Minv) {
    int n = (int)rowPtr.size() - 1;
    Minv.assign(9 * (size_t)n, T(0));
    for (int i = 0; i < n; i++) {
        const T* dg = nullptr;
        for (int k = rowPtr[i]; k < rowPtr[i + 1]; k++) if (col[k] == i) dg = &val[9 * k];
        inverse3(dg, &Minv[9 * i]);
    }
}

template <class T> int pcg_block_jacobi(const vector<int>& rowPtr, const vector<int>& col, const vector<T>& val,
                                        const vector<T>& b, T tol, int maxIters, vector<T>& x, T* relres) {
    const int n = (int)rowPtr.size() - 1;
    vector<T> Minv;
    block_jacobi(rowPtr, col, val, Minv);
    return pcg_with(rowPtr, col, val, b, tol, maxIters, x, relres, [&](const vector<T>& in, vector<T>& out) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) {
                T s = T(0);
                for (int c = 0; c < 3; c++) s += Minv[9 * i + 3 * a + c] * in[3 * i + c];
                out[3 * i + a] = s;
            }
    });
}

// Direct sparse LL^T (the reference CPU solve is Eigen SimplicialLLT,
// src/Cloth.cu:686-688).  Banded Cholesky in FP64 after a reverse Cuthill-McKee
// ordering of the block graph; exact up to rounding, so it serves as the
// direct-solve anchor for small and medium systems.
template <class T> bool direct_solve(const vector<int>& rowPtr, const vector<int>& col, const vector<T>& val,
                                     const vector<T>& b, vector<T>& x) {
    int n = (int)rowPtr.size() - 1;
    // RCM ordering on blocks
    vector<int> deg(n), perm, inv(n, -1);
    for (int i = 0; i < n; i++) deg[i] = rowPtr[i + 1] - rowPtr[i];
    vector<char> seen(n, 0);
    for (int s0 = 0; s0 < n; s0++) {
        if (seen[s0]) continue;
        int start = s0;
        for (int i = s0; i < n; i++) if (!seen[i] && deg[i] < deg[start]) start = i;
        std::queue<int> qu;
        qu.push(start); seen[start] = 1;
        while (!qu.empty()) {
            int i = qu.front(); qu.pop();
            perm.push_back(i);
            vector<int> nb;
            for (int k = rowPtr[i]; k < rowPtr[i + 1]; k++) if (!seen[col[k]]) { nb.push_back(col[k]); seen[col[k]] = 1; }
            std::sort(nb.begin(), nb.end(), [&](int a, int c) { return deg[a] < deg[c]; });
            for (int j : nb) qu.push(j);
        }
    }
    std::reverse(perm.begin(), perm.end());
    for (int i = 0; i < n; i++) inv[perm[i]] = i;
    int bwBlocks = 0;
    for (int i = 0; i < n; i++)
        for (int k = rowPtr[i]; k < rowPtr[i + 1]; k++) bwBlocks = std::max(bwBlocks, std::abs(inv[i] - inv[col[k]]));
    int m = 3 * n;
    int bw = 3 * bwBlocks + 2;
    // lower band storage: L[i][i-j] for j in [i-bw, i]
    vector<double> Lb((size_t)m * (bw + 1), 0.0);
    auto at = [&](int i, int j) -> double& { return Lb[(size_t)i * (bw + 1) + (i - j)]; };
    for (int i = 0; i < n; i++)
        for (int k = rowPtr[i]; k < rowPtr[i + 1]; k++) {
            int pi = inv[i], pj = inv[col[k]];
            for (int a = 0; a < 3; a++)
                for (int c = 0; c < 3; c++) {
                    int r = 3 * pi + a, cc = 3 * pj + c;
                    if (cc <= r) at(r, cc) = (double)val[9 * k + 3 * a + c];
                }
        }
    for (int j = 0; j < m; j++) {
        double s = at(j, j);
        int k0 = std::max(0, j - bw);
        for (int k = k0; k < j; k++) s -= at(j, k) * at(j, k);
        if (!(s > 0.0)) return false;
        double djj = std::sqrt(s);
        at(j, j) = djj;
        int iMax = std::min(m - 1, j + bw);
        for (int i = j + 1; i <= iMax; i++) {
            double t = at(i, j);
            int kk0 = std::max(k0, i - bw);
            for (int k = kk0; k < j; k++) t -= at(i, k) * at(j, k);
            at(i, j) = t / djj;
        }
    }
    vector<double> y(m);
    for (int i = 0; i < n; i++) for (int a = 0; a < 3; a++) y[3 * inv[i] + a] = (double)b[3 * i + a];
    for (int i = 0; i < m; i++) {
        double s = y[i];
        for (int k = std::max(0, i - bw); k < i; k++) s -= at(i, k) * y[k];
        y[i] = s / at(i, i);
    }
    for (int i = m - 1; i >= 0; i--) {
        double s = y[i];
        for (int k = i + 1; k <= std::min(m - 1, i + bw); k++) s -= at(k, i) * y[k];
        y[i] = s / at(i, i);
    }
    x.assign(m, T(0));
    for (int i = 0; i < n; i++) for (int a = 0; a < 3; a++) x[3 * i + a] = (T)y[3 * inv[i] + a];
    return true;
}

// The same LL^T with the nested-dissection multifrontal factorisation of or_direct.hpp
// (fill-reducing order: the large systems).
template <class T> bool direct_solve_nd(const vector<int>& rowPtr, const vector<int>& col, const vector<T>& val,
                                        const vector<T>& b, vector<T>& x) {
    NdCholesky ch;
    if (!ch.factor(rowPtr, col, val)) return false;
    ch.solve(b, x);
    return true;
}

}  // namespace orc
