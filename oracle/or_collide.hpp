// oracle/or_collide.hpp — LBVH, proximities, CCD, independent impacts and the
// impact response of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (see or_math.hpp).  Restates src/BVH.cu,
// src/BVHHelper.cu, src/Simulator.cu:179-461,571-684, src/PhysicsHelper.cu:3-150,
// src/CollisionHelper.cu, src/Proximity.cu and src/optimization/*.
#pragma once
#include <functional>
#include <set>

#include "or_amg.hpp"
#include "or_physics.hpp"

namespace orc {

// ---------------------------------------------------------------- bounds
// src/Bounds.cu:3-87.  The empty box is (FLT_MAX, -FLT_MAX).
template <class T> struct Box {
    T lo[3], hi[3];
    Box() { for (int i = 0; i < 3; i++) { lo[i] = T(FLT_MAX); hi[i] = T(-FLT_MAX); } }
    void add(const V3<T>& p) { for (int i = 0; i < 3; i++) { lo[i] = tmin(lo[i], p[i]); hi[i] = tmax(hi[i], p[i]); } }
    void add(const Box& b) { for (int i = 0; i < 3; i++) { lo[i] = tmin(lo[i], b.lo[i]); hi[i] = tmax(hi[i], b.hi[i]); } }
    Box dilate(T th) const { Box r = *this; for (int i = 0; i < 3; i++) { r.lo[i] -= th; r.hi[i] += th; } return r; }
    bool overlap(const Box& b) const {
        for (int i = 0; i < 3; i++) {
            if (lo[i] > b.hi[i]) return false;
            if (hi[i] < b.lo[i]) return false;
        }
        return true;
    }
    bool overlap(const Box& b, T th) const { return overlap(b.dilate(th)); }
};

template <class T> Box<T> face_box(const Scene<T>& S, int f, bool ccd) {  // Face::bounds src/Face.cu:94-103
    Box<T> b;
    for (int i = 0; i < 3; i++) { b.add(S.x[S.fn[f][i]]); if (ccd) b.add(S.x0[S.fn[f][i]]); }
    return b;
}
template <class T> Box<T> node_box(const Scene<T>& S, int n, bool ccd) {  // Node::bounds src/Node.cu:18-24
    Box<T> b;
    b.add(S.x[n]);
    if (ccd) b.add(S.x0[n]);
    return b;
}
template <class T> Box<T> edge_box(const Scene<T>& S, int e, bool ccd) {  // Edge::bounds src/Edge.cu:85-94
    Box<T> b;
    for (int i = 0; i < 2; i++) { b.add(S.x[S.en[e][i]]); if (ccd) b.add(S.x0[S.en[e][i]]); }
    return b;
}

// ---------------------------------------------------------------- LBVH
// src/BVHHelper.cu:10-29
inline uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
template <class T> uint32_t morton_code(const V3<T>& c, const V3<T>& p, const V3<T>& d) {
    T q[3];
    for (int i = 0; i < 3; i++) {
        T s = d[i] == T(0) ? T(0) : (c[i] - p[i]) / d[i];
        q[i] = std::fmin(std::fmax(s * T(1024), T(0)), T(1023));
    }
    uint32_t xx = expand_bits((uint32_t)q[0]);
    uint32_t yy = expand_bits((uint32_t)q[1]);
    uint32_t zz = expand_bits((uint32_t)q[2]);
    return (xx << 2) + (yy << 1) + zz;
}
inline int clz64(uint64_t a) { return a == 0 ? 64 : __builtin_clzll(a); }

template <class T> struct BVH {
    int mesh = -1;
    bool ccd = false;
    int nLeaves = 0;
    vector<uint64_t> keys;    // sorted (morton << 32 | local face)
    vector<int> leafFace;     // sorted leaf -> global face
    vector<Box<T>> leafBox;
    // internal node i: children (index, isLeaf), box, maxIndex
    vector<int> left, right;
    vector<uint8_t> leftLeaf, rightLeaf;
    vector<Box<T>> ibox;
    vector<int> imax;
};

// src/BVHHelper.cu:48-89 (Karras 2012 range / split on the 64-bit keys)
inline void karras_range(int nInternal, const vector<uint64_t>& k, int i, int& left, int& right) {
    if (i == 0) { left = 0; right = nInternal; return; }
    uint64_t code = k[i];
    int dl = clz64(code ^ k[i - 1]);
    int dr = clz64(code ^ k[i + 1]);
    if (dl > dr) {
        int l = -1, r = i - 1;
        while (r - l > 1) { int mid = (l + r) >> 1; if (clz64(code ^ k[mid]) > dr) r = mid; else l = mid; }
        left = r; right = i;
    } else {
        int l = i + 1, r = nInternal + 1;
        while (r - l > 1) { int mid = (l + r) >> 1; if (clz64(code ^ k[mid]) > dl) l = mid; else r = mid; }
        left = i; right = l;
    }
}
inline int karras_split(const vector<uint64_t>& k, int left, int right) {
    uint64_t code = k[left];
    int d = clz64(code ^ k[right]);
    int l = left, r = right;
    while (r - l > 1) { int mid = (l + r) >> 1; if (clz64(code ^ k[mid]) > d) l = mid; else r = mid; }
    return l;
}

template <class T> void bvh_refit(const Scene<T>& S, BVH<T>& B) {
    for (int i = 0; i < B.nLeaves; i++) B.leafBox[i] = face_box(S, B.leafFace[i], B.ccd);
    int ni = B.nLeaves - 1;
    // children have larger indices than nothing in particular: do a post-order pass
    std::function<void(int)> rec = [&](int i) {
        Box<T> b; int mx = -1;
        if (B.leftLeaf[i]) { b.add(B.leafBox[B.left[i]]); mx = std::max(mx, B.left[i]); }
        else { rec(B.left[i]); b.add(B.ibox[B.left[i]]); mx = std::max(mx, B.imax[B.left[i]]); }
        if (B.rightLeaf[i]) { b.add(B.leafBox[B.right[i]]); mx = std::max(mx, B.right[i]); }
        else { rec(B.right[i]); b.add(B.ibox[B.right[i]]); mx = std::max(mx, B.imax[B.right[i]]); }
        B.ibox[i] = b; B.imax[i] = mx;
    };
    if (ni > 0) rec(0);
}

// BVH::BVH GPU (src/BVH.cu:18-50)
template <class T> void bvh_build(const Scene<T>& S, int mesh, bool ccd, BVH<T>& B) {
    const MeshRange& R = S.meshes[mesh];
    B.mesh = mesh; B.ccd = ccd; B.nLeaves = R.nf;
    vector<Box<T>> boxes(R.nf);
    Box<T> scene;
    for (int f = 0; f < R.nf; f++) { boxes[f] = face_box(S, R.f0 + f, ccd); scene.add(boxes[f]); }
    V3<T> p, d;
    for (int i = 0; i < 3; i++) { p[i] = scene.lo[i]; d[i] = scene.hi[i] - scene.lo[i]; }
    B.keys.resize(R.nf);
    for (int f = 0; f < R.nf; f++) {
        V3<T> c;
        for (int i = 0; i < 3; i++) c[i] = T(0.5f) * (boxes[f].lo[i] + boxes[f].hi[i]);
        B.keys[f] = ((uint64_t)morton_code(c, p, d) << 32) | (uint64_t)f;
    }
    std::sort(B.keys.begin(), B.keys.end());
    B.leafFace.resize(R.nf); B.leafBox.resize(R.nf);
    for (int i = 0; i < R.nf; i++) {
        int f = (int)(B.keys[i] & 0xffffffffu);
        B.leafFace[i] = R.f0 + f; B.leafBox[i] = boxes[f];
    }
    int ni = R.nf - 1;
    B.left.assign(std::max(ni, 0), 0); B.right.assign(std::max(ni, 0), 0);
    B.leftLeaf.assign(std::max(ni, 0), 0); B.rightLeaf.assign(std::max(ni, 0), 0);
    B.ibox.assign(std::max(ni, 0), Box<T>()); B.imax.assign(std::max(ni, 0), -1);
    for (int i = 0; i < ni; i++) {  // initializeInternalNodes src/BVHHelper.cu:91-103
        int l, r;
        karras_range(ni, B.keys, i, l, r);
        int mid = karras_split(B.keys, l, r);
        B.left[i] = mid; B.leftLeaf[i] = (mid == l);
        B.right[i] = mid + 1; B.rightLeaf[i] = (mid + 1 == r);
    }
    bvh_refit(S, B);
}

// Leaf-vs-tree stack traversal (src/BVHHelper.cu:151-307).  self: the query is
// sorted leaf qi of B and only leaves with larger sorted index are reported.
template <class T> void bvh_query(const BVH<T>& B, const Box<T>& qb, int qi, bool self, T th, const std::function<void(int)>& emit) {
    if (B.nLeaves == 1) {  // a one-face mesh has no internal node: test the leaf itself
        if (!self && qb.overlap(B.leafBox[0], th)) emit(B.leafFace[0]);
        return;
    }
    vector<int> stack;
    int node = 0;
    for (;;) {
        int L = B.left[node], R = B.right[node];
        int lmax = B.leftLeaf[node] ? L : B.imax[L];
        int rmax = B.rightLeaf[node] ? R : B.imax[R];
        const Box<T>& lb = B.leftLeaf[node] ? B.leafBox[L] : B.ibox[L];
        const Box<T>& rb = B.rightLeaf[node] ? B.leafBox[R] : B.ibox[R];
        bool ovL = (!self || qi < lmax) && qb.overlap(lb, th);
        bool ovR = (!self || qi < rmax) && qb.overlap(rb, th);
        if (ovL && B.leftLeaf[node]) emit(B.leafFace[L]);
        if (ovR && B.rightLeaf[node]) emit(B.leafFace[R]);
        bool tL = ovL && !B.leftLeaf[node];
        bool tR = ovR && !B.rightLeaf[node];
        if (!tL && !tR) {
            if (stack.empty()) break;
            node = stack.back(); stack.pop_back();
        } else {
            node = tL ? L : R;
            if (tL && tR) stack.push_back(R);
        }
    }
}

// Simulator::traverse (src/Simulator.cu:179-196): cloth self, cloth x earlier
// cloth, cloth x obstacle.  Emits (query face, other face).
template <class T> void traverse_all(const Scene<T>& S, const vector<BVH<T>>& bvhs, T th,
                                     const std::function<void(int, int)>& emit) {
    for (int c = 0; c < S.nCloths; c++) {
        const BVH<T>& B = bvhs[c];
        for (int i = 0; i < B.nLeaves; i++) {
            int f0 = B.leafFace[i];
            bvh_query(B, B.leafBox[i], i, true, th, [&](int f1) { emit(f0, f1); });
        }
        for (int j = 0; j < c; j++)
            for (int i = 0; i < B.nLeaves; i++) {
                int f0 = B.leafFace[i];
                bvh_query(bvhs[j], B.leafBox[i], i, false, th, [&](int f1) { emit(f0, f1); });
            }
        for (int o = S.nCloths; o < (int)S.meshes.size(); o++)
            for (int i = 0; i < B.nLeaves; i++) {
                int f0 = B.leafFace[i];
                bvh_query(bvhs[o], B.leafBox[i], i, false, th, [&](int f1) { emit(f0, f1); });
            }
    }
}

// The same pair set by brute force over all leaf pairs (used to pin the tree).
template <class T> void traverse_bruteforce(const Scene<T>& S, const vector<BVH<T>>& bvhs, T th,
                                            const std::function<void(int, int)>& emit) {
    for (int c = 0; c < S.nCloths; c++) {
        const BVH<T>& B = bvhs[c];
        for (int i = 0; i < B.nLeaves; i++)
            for (int j = i + 1; j < B.nLeaves; j++)
                if (B.leafBox[i].overlap(B.leafBox[j], th)) emit(B.leafFace[i], B.leafFace[j]);
        for (int o = 0; o < (int)S.meshes.size(); o++) {
            if (o == c || (o < S.nCloths && o > c)) continue;
            const BVH<T>& C = bvhs[o];
            for (int i = 0; i < B.nLeaves; i++)
                for (int j = 0; j < C.nLeaves; j++)
                    if (B.leafBox[i].overlap(C.leafBox[j], th)) emit(B.leafFace[i], C.leafFace[j]);
        }
    }
}

// All candidate pairs of traverse_all, in the same order, computed over leaf
// chunks in parallel (each chunk's pairs are kept in leaf order and the
// chunks concatenated in order).
template <class T> vector<std::pair<int, int>> collect_pairs(const Scene<T>& S, const vector<BVH<T>>& bvhs, T th) {
    vector<std::pair<int, int>> all;
    auto run = [&](int c, int target, bool self) {
        const BVH<T>& B = bvhs[c];
        const BVH<T>& Tg = bvhs[target];
        const int CH = 1024;
        const int nch = (B.nLeaves + CH - 1) / CH;
        vector<vector<std::pair<int, int>>> parts(nch);
#pragma omp parallel for schedule(dynamic, 4)
        for (int ch = 0; ch < nch; ch++)
            for (int i = ch * CH; i < std::min(B.nLeaves, (ch + 1) * CH); i++) {
                int f0 = B.leafFace[i];
                bvh_query(Tg, B.leafBox[i], i, self, th, [&](int f1) { parts[ch].push_back({f0, f1}); });
            }
        for (auto& p : parts) all.insert(all.end(), p.begin(), p.end());
    };
    for (int c = 0; c < S.nCloths; c++) {
        run(c, c, true);
        for (int j = 0; j < c; j++) run(c, j, false);
        for (int o = S.nCloths; o < (int)S.meshes.size(); o++) run(c, o, false);
    }
    return all;
}

template <class T> void build_all(const Scene<T>& S, bool ccd, vector<BVH<T>>& bvhs) {
    bvhs.assign(S.meshes.size(), BVH<T>());
    for (int m = 0; m < (int)S.meshes.size(); m++) bvh_build(S, m, ccd, bvhs[m]);
}

// ---------------------------------------------------------------- distances
// src/MathHelper.cu:38-69.  Returns false for the degenerate case (the
// reference returns INFINITY and leaves w unset, which can never pass the
// callers' distance tests).
template <class T> bool signed_vf_distance(const V3<T>& x, const V3<T>& y0, const V3<T>& y1, const V3<T>& y2, V3<T>& n, T w[4], T& h) {
    n = cross((y1 - y0).normalized(), (y2 - y0).normalized());
    if (n.norm2() < T(1e-6f)) return false;
    n.normalize();
    h = (x - y0).dot(n);
    T b0 = mixed(y1 - x, y2 - x, n);
    T b1 = mixed(y2 - x, y0 - x, n);
    T b2 = mixed(y0 - x, y1 - x, n);
    w[0] = T(1);
    w[1] = -b0 / (b0 + b1 + b2);
    w[2] = -b1 / (b0 + b1 + b2);
    w[3] = -b2 / (b0 + b1 + b2);
    return true;
}
template <class T> bool signed_ee_distance(const V3<T>& x0, const V3<T>& x1, const V3<T>& y0, const V3<T>& y1, V3<T>& n, T w[4], T& h) {
    n = cross((x1 - x0).normalized(), (y1 - y0).normalized());
    if (n.norm2() < T(1e-6f)) return false;
    n.normalize();
    h = (x0 - y0).dot(n);
    T a0 = mixed(y1 - x1, y0 - x1, n);
    T a1 = mixed(y0 - x0, y1 - x0, n);
    T b0 = mixed(x0 - y1, x1 - y1, n);
    T b1 = mixed(x1 - y0, x0 - y0, n);
    w[0] = a0 / (a0 + a1);
    w[1] = a1 / (a0 + a1);
    w[2] = -b0 / (b0 + b1);
    w[3] = -b1 / (b0 + b1);
    return true;
}

// ---------------------------------------------------------------- proximities
template <class T> struct Best {
    T d; int other; bool has = false;
    void offer(T dd, int o) {  // Pair<float, X*>::operator< (src/Pair.cuh:38-40), X* order = global index order
        if (!has || dd < d || (dd == d && o < other)) { d = dd; other = o; has = true; }
    }
};

// inEdge (src/PhysicsHelper.cu:3-18)
template <class T> bool in_edge(const Scene<T>& S, T w, int e0, int e1) {
    V3<T> x = (T(1) - w) * S.x[S.en[e0][0]] + w * S.x[S.en[e0][1]];
    bool in = true;
    for (int i = 0; i < 2; i++) {
        int f = S.eadj[e1][i];
        if (f < 0) continue;
        int n0 = S.en[e1][i], n1 = S.en[e1][1 - i];
        V3<T> e = S.x[n1] - S.x[n0];
        V3<T> r = x - S.x[n0];
        in &= (mixed(e, S.fnrm[f], r) >= T(0));
    }
    return in;
}

template <class T> struct ProxCand {
    vector<Best<T>> node, edge, face;  // [2*count], key = 2*primitive + side
    void offer(int cat, int slot, T d, int other) {
        (cat == 0 ? node : (cat == 1 ? edge : face))[slot].offer(d, other);
    }
};
// candidates of a chunk of pairs, merged into ProxCand afterwards
template <class T> struct CandList {
    struct C4 { int cat, slot; T d; int other; };
    vector<C4> v;
    void offer(int cat, int slot, T d, int other) { v.push_back({cat, slot, d, other}); }
};

// checkVertexFaceProximityGpu (src/PhysicsHelper.cu:20-49); d = |h| with -0 read as +0
template <class T, class Sink> void vf_proximity(const Scene<T>& S, int node, int f, Sink& C) {
    int n0 = S.fn[f][0], n1 = S.fn[f][1], n2 = S.fn[f][2];
    if (node == n0 || node == n1 || node == n2) return;
    V3<T> n; T w[4], h;
    if (!signed_vf_distance(S.x[node], S.x[n0], S.x[n1], S.x[n2], n, w, h)) return;
    T d = std::fabs(h);
    bool inside = (tmin3(-w[1], -w[2], -w[3]) >= T(1e-6f));
    if (!inside) return;
    if (S.nodeFree(node)) {
        int side = n.dot(S.nrm[node]) >= T(0) ? 0 : 1;
        C.offer(0, 2 * node + side, d, f);
    }
    if (S.faceFree(f)) {
        int side = -n.dot(S.fnrm[f]) >= T(0) ? 0 : 1;
        C.offer(2, 2 * f + side, d, node);
    }
}
// checkEdgeEdgeProximityGpu (src/PhysicsHelper.cu:51-81)
template <class T, class Sink> void ee_proximity(const Scene<T>& S, int e0, int e1, Sink& C) {
    int a = S.en[e0][0], b = S.en[e0][1], c = S.en[e1][0], d4 = S.en[e1][1];
    if (a == c || a == d4 || b == c || b == d4) return;
    V3<T> n; T w[4], h;
    if (!signed_ee_distance(S.x[a], S.x[b], S.x[c], S.x[d4], n, w, h)) return;
    T d = std::fabs(h);
    bool inside = (tmin4(w[0], w[1], -w[2], -w[3]) >= T(1e-6f) && in_edge(S, w[1], e0, e1) && in_edge(S, -w[3], e1, e0));
    if (!inside) return;
    if (S.edgeFree(e0)) {
        int side = n.dot(S.nrm[a] + S.nrm[b]) >= T(0) ? 0 : 1;
        C.offer(1, 2 * e0 + side, d, e1);
    }
    if (S.edgeFree(e1)) {
        int side = -n.dot(S.nrm[c] + S.nrm[d4]) >= T(0) ? 0 : 1;
        C.offer(1, 2 * e1 + side, d, e0);
    }
}

// Proximity::Proximity (src/Proximity.cu:5-21)
template <class T> Prox<T> make_vf_prox(const Scene<T>& S, int node, int f) {
    Prox<T> p;
    p.nodes[0] = node; p.nodes[1] = S.fn[f][0]; p.nodes[2] = S.fn[f][1]; p.nodes[3] = S.fn[f][2];
    p.stiffness = T(S.P.collision_stiffness) * tmin(S.area[node], S.farea[f]);
    T h;
    signed_vf_distance(S.x[p.nodes[0]], S.x[p.nodes[1]], S.x[p.nodes[2]], S.x[p.nodes[3]], p.n, p.w, h);
    if (h < T(0)) p.n = -p.n;
    p.mu = (S.nodeFree(node) && S.faceFree(f)) ? T(S.P.friction) : T(S.P.obstacle_friction);
    return p;
}
template <class T> Prox<T> make_ee_prox(const Scene<T>& S, int e0, int e1) {
    Prox<T> p;
    p.nodes[0] = S.en[e0][0]; p.nodes[1] = S.en[e0][1]; p.nodes[2] = S.en[e1][0]; p.nodes[3] = S.en[e1][1];
    p.stiffness = T(S.P.collision_stiffness) * tmin(S.edgeArea(e0), S.edgeArea(e1));
    T h;
    signed_ee_distance(S.x[p.nodes[0]], S.x[p.nodes[1]], S.x[p.nodes[2]], S.x[p.nodes[3]], p.n, p.w, h);
    if (h < T(0)) p.n = -p.n;
    p.mu = (S.edgeFree(e0) && S.edgeFree(e1)) ? T(S.P.friction) : T(S.P.obstacle_friction);
    return p;
}

// Every face pair of the mesh pairs Simulator::traverse visits (a mesh against
// itself, a cloth against earlier cloths and against obstacles), with no box
// test: the tree-free candidate set that pins find_proximities.  A proximity
// closer than 2 thickness lies between faces whose dilated boxes overlap, so
// the nearest candidate of every slot below the limit is the same in both.
template <class T> vector<std::pair<int, int>> all_face_pairs(const Scene<T>& S) {
    vector<std::pair<int, int>> all;
    for (int c = 0; c < S.nCloths; c++) {
        const int f0 = S.meshes[c].f0, nf = S.meshes[c].nf;
        for (int i = 0; i < nf; i++)
            for (int j = i + 1; j < nf; j++) all.push_back({f0 + i, f0 + j});
        for (int o = 0; o < (int)S.meshes.size(); o++) {
            if (o == c || (o < S.nCloths && o > c)) continue;
            for (int i = 0; i < nf; i++)
                for (int j = 0; j < S.meshes[o].nf; j++) all.push_back({f0 + i, S.meshes[o].f0 + j});
        }
    }
    return all;
}

// Simulator::findProximitiesGpu (src/Simulator.cu:300-349); brute: over all_face_pairs
// instead of the tree's candidate pairs
template <class T> vector<Prox<T>> find_proximities(const Scene<T>& S, bool brute = false) {
    T thick = T(S.P.repulsion_thickness);
    ProxCand<T> C;
    C.node.assign(2 * S.x.size(), Best<T>());
    C.edge.assign(2 * S.en.size(), Best<T>());
    C.face.assign(2 * S.fn.size(), Best<T>());
    vector<std::pair<int, int>> pairs;
    if (brute) {
        pairs = all_face_pairs(S);
    } else {
        vector<BVH<T>> bvhs;
        build_all(S, false, bvhs);
        pairs = collect_pairs(S, bvhs, T(2) * thick);
    }
    const long np = (long)pairs.size(), CH = 4096, nch = (np + CH - 1) / CH;
    vector<CandList<T>> parts(nch);
#pragma omp parallel for schedule(dynamic, 4)
    for (long ch = 0; ch < nch; ch++)
        for (long p = ch * CH; p < std::min(np, (ch + 1) * CH); p++) {
            int f0 = pairs[p].first, f1 = pairs[p].second;
            for (int j = 0; j < 3; j++) vf_proximity(S, S.fn[f0][j], f1, parts[ch]);
            for (int j = 0; j < 3; j++) vf_proximity(S, S.fn[f1][j], f0, parts[ch]);
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) ee_proximity(S, S.fe[f0][j], S.fe[f1][k], parts[ch]);
        }
    for (auto& part : parts)  // the minimum per slot does not depend on the merge order
        for (auto& c4 : part.v) C.offer(c4.cat, c4.slot, c4.d, c4.other);
    vector<Prox<T>> out;
    T lim = T(2) * thick;
    for (size_t k = 0; k < C.node.size(); k++)
        if (C.node[k].has && C.node[k].d < lim) out.push_back(make_vf_prox(S, (int)(k / 2), C.node[k].other));
    for (size_t k = 0; k < C.edge.size(); k++)
        if (C.edge[k].has && C.edge[k].d < lim) out.push_back(make_ee_prox(S, (int)(k / 2), C.edge[k].other));
    for (size_t k = 0; k < C.face.size(); k++)
        if (C.face[k].has && C.face[k].d < lim) out.push_back(make_vf_prox(S, C.face[k].other, (int)(k / 2)));
    return out;
}

// ---------------------------------------------------------------- CCD (FP64)
struct Impact {
    int type;  // 0 VF, 1 EE
    int nodes[4];
    double t, w[4], n[3];
    int64_t key;  // type<<62 | a<<31 | b  (VF: node, face; EE: edge, edge)
};

// src/CollisionHelper.cu:3-19
inline double newton_root(double a, double b, double c, double d, double x0, int dir) {
    if (dir != 0) {
        double y0 = d + x0 * (c + x0 * (b + x0 * a));
        double ddy0 = 2.0 * b + x0 * (6.0 * a);
        x0 += dir * std::sqrt(tabs(2.0 * y0 / ddy0));
    }
    for (int iter = 0; iter < 100; iter++) {
        double y = d + x0 * (c + x0 * (b + x0 * a));
        double dy = c + x0 * (2.0 * b + 3.0 * x0 * a);
        if (dy == 0) return x0;
        double x1 = x0 - y / dy;
        if (tabs(x0 - x1) < 1e-6) return x0;
        x0 = x1;
    }
    return x0;
}
// src/CollisionHelper.cu:21-37
inline int solve_quadratic(double a, double b, double c, double x[2]) {
    double d = b * b - 4.0 * a * c;
    if (d < 0.0) { x[0] = -b / (2.0 * a); return 0; }
    double q = -(b + tsign(b) * std::sqrt(d)) * 0.5;
    int i = 0;
    if (tabs(a) > 1e-12 * tabs(q)) x[i++] = q / a;
    if (tabs(q) > 1e-12 * tabs(c)) x[i++] = c / q;
    if (i == 2 && x[0] > x[1]) std::swap(x[0], x[1]);
    return i;
}
// src/CollisionHelper.cu:39-60
inline int solve_cubic(double a, double b, double c, double d, double x[3]) {
    double xc[2] = {0.0, 0.0};  // the reference leaves xc[0] unset when 3a = 2b = 0; start Newton at 0
    int n = solve_quadratic(3.0 * a, 2.0 * b, c, xc);
    if (n == 0) { x[0] = newton_root(a, b, c, d, xc[0], 0); return 1; }
    if (n == 1) return solve_quadratic(b, c, d, x);
    double yc[2] = {d + xc[0] * (c + xc[0] * (b + xc[0] * a)), d + xc[1] * (c + xc[1] * (b + xc[1] * a))};
    int i = 0;
    if (yc[0] * a >= 0.0) x[i++] = newton_root(a, b, c, d, xc[0], -1);
    if (yc[0] * yc[1] <= 0.0) {
        int closer = tabs(yc[0]) < tabs(yc[1]) ? 0 : 1;
        x[i++] = newton_root(a, b, c, d, xc[closer], closer == 0 ? 1 : -1);
    }
    if (yc[1] * a <= 0.0) x[i++] = newton_root(a, b, c, d, xc[1], 1);
    return i;
}

// Roots of f(t) = a t^3 + b t^2 + c t + d on [0, 1] in ascending order, by
// isolation: [0, 1] is cut at the critical points of f, and each piece whose
// end values have opposite signs holds exactly one root, found by safeguarded
// Newton (bisection when a Newton step leaves the bracket) to |dt| < 1e-12.
// Exact zeros at the cut points are roots too.  This replaces the
// reference's whole-line Newton search (src/CollisionHelper.cu:3-60), whose
// roots outside [0, 1] are discarded by its caller anyway; see the design
// notes for the cases where the two differ (grazing double roots).
inline double cubic_eval(double a, double b, double c, double d, double t) { return d + t * (c + t * (b + t * a)); }
inline int cubic_roots_unit(double a, double b, double c, double d, double out[3]) {
    double cut[4];
    int nc = 0;
    cut[nc++] = 0.0;
    const double qa = 3.0 * a, qb = 2.0 * b, qc = c;
    double cr[2];
    int ncr = 0;
    if (qa != 0.0) {
        double disc = qb * qb - 4.0 * qa * qc;
        if (disc >= 0.0) {
            double sq = std::sqrt(disc);
            double q = -0.5 * (qb + (qb < 0.0 ? -sq : sq));
            double r1 = q / qa;
            double r2 = q != 0.0 ? qc / q : r1;
            if (r2 < r1) std::swap(r1, r2);
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

// checkImpact (src/CollisionHelper.cu:62-109), promoted to FP64.
inline bool check_impact(int type, const V3<double> X0[4], const V3<double> X1[4], Impact& imp) {
    V3<double> x0 = X0[0];
    V3<double> v0 = X1[0] - x0;
    V3<double> x1 = X0[1] - x0, x2 = X0[2] - x0, x3 = X0[3] - x0;
    V3<double> v1 = (X1[1] - X0[1]) - v0, v2 = (X1[2] - X0[2]) - v0, v3 = (X1[3] - X0[3]) - v0;
    double a0 = mixed(x1, x2, x3);
    double a1 = mixed(v1, x2, x3) + mixed(x1, v2, x3) + mixed(x1, x2, v3);
    double a2 = mixed(x1, v2, v3) + mixed(v1, x2, v3) + mixed(v1, v2, x3);
    double a3 = mixed(v1, v2, v3);
    double t[4];
    int ns = cubic_roots_unit(a3, a2, a1, a0, t);
    for (int i = 0; i < ns; i++) {
        if (t[i] < 0.0 || t[i] > 1.0) continue;
        V3<double> p[4];
        for (int k = 0; k < 4; k++) p[k] = X0[k] + t[i] * (X1[k] - X0[k]);  // Node::position src/Node.cu:26-28
        V3<double> n; double w[4], d;
        bool ok, inside;
        if (type == 0) {
            ok = signed_vf_distance(p[0], p[1], p[2], p[3], n, w, d);
            inside = ok && (tmin3(-w[1], -w[2], -w[3]) >= -1e-6);
        } else {
            ok = signed_ee_distance(p[0], p[1], p[2], p[3], n, w, d);
            inside = ok && (tmin4(w[0], w[1], -w[2], -w[3]) >= -1e-6);
        }
        if (!ok) continue;
        if (n.dot(w[1] * v1 + w[2] * v2 + w[3] * v3) > 0.0) n = -n;
        if (tabs(d) < 1e-6 && inside) {
            imp.t = t[i] + 0.0;  // a root of -0.0 is reported as +0.0
            for (int k = 0; k < 4; k++) imp.w[k] = w[k];
            for (int k = 0; k < 3; k++) imp.n[k] = n[k];
            return true;
        }
    }
    return false;
}

// ---- the reference's own FP32 CCD (src/CollisionHelper.cu:3-109), kept in float.  The
// product and the checks above run the cubic in FP64 with isolated roots; this restatement only
// measures how often the FP32 reference path decides differently (a divergence statistic).
inline float newton_root_f(float a, float b, float c, float d, float x0, int dir) {
    if (dir != 0) {
        float y0 = d + x0 * (c + x0 * (b + x0 * a));
        float ddy0 = 2.0f * b + x0 * (6.0f * a);
        x0 += dir * std::sqrt(tabs(2.0f * y0 / ddy0));
    }
    for (int iter = 0; iter < 100; iter++) {
        float y = d + x0 * (c + x0 * (b + x0 * a));
        float dy = c + x0 * (2.0f * b + 3.0f * x0 * a);
        if (dy == 0) return x0;
        float x1 = x0 - y / dy;
        if (tabs(x0 - x1) < 1e-6f) return x0;
        x0 = x1;
    }
    return x0;
}
inline int solve_quadratic_f(float a, float b, float c, float x[2]) {
    float d = b * b - 4.0f * a * c;
    if (d < 0.0f) { x[0] = -b / (2.0f * a); return 0; }
    float q = -(b + tsign(b) * std::sqrt(d)) * 0.5f;
    int i = 0;
    if (tabs(a) > 1e-12f * tabs(q)) x[i++] = q / a;
    if (tabs(q) > 1e-12f * tabs(c)) x[i++] = c / q;
    if (i == 2 && x[0] > x[1]) std::swap(x[0], x[1]);
    return i;
}
inline int solve_cubic_f(float a, float b, float c, float d, float x[3]) {
    float xc[2] = {0.0f, 0.0f};
    int n = solve_quadratic_f(3.0f * a, 2.0f * b, c, xc);
    if (n == 0) { x[0] = newton_root_f(a, b, c, d, xc[0], 0); return 1; }
    if (n == 1) return solve_quadratic_f(b, c, d, x);
    float yc[2] = {d + xc[0] * (c + xc[0] * (b + xc[0] * a)), d + xc[1] * (c + xc[1] * (b + xc[1] * a))};
    int i = 0;
    if (yc[0] * a >= 0.0f) x[i++] = newton_root_f(a, b, c, d, xc[0], -1);
    if (yc[0] * yc[1] <= 0.0f) {
        int closer = tabs(yc[0]) < tabs(yc[1]) ? 0 : 1;
        x[i++] = newton_root_f(a, b, c, d, xc[closer], closer == 0 ? 1 : -1);
    }
    if (yc[1] * a <= 0.0f) x[i++] = newton_root_f(a, b, c, d, xc[1], 1);
    return i;
}
inline bool check_impact_ref32(int type, const V3<float> X0[4], const V3<float> X1[4], Impact& imp) {
    V3<float> x0 = X0[0];
    V3<float> v0 = X1[0] - x0;
    V3<float> x1 = X0[1] - x0, x2 = X0[2] - x0, x3 = X0[3] - x0;
    V3<float> v1 = (X1[1] - X0[1]) - v0, v2 = (X1[2] - X0[2]) - v0, v3 = (X1[3] - X0[3]) - v0;
    float a0 = mixed(x1, x2, x3);
    float a1 = mixed(v1, x2, x3) + mixed(x1, v2, x3) + mixed(x1, x2, v3);
    float a2 = mixed(x1, v2, v3) + mixed(v1, x2, v3) + mixed(v1, v2, x3);
    float a3 = mixed(v1, v2, v3);
    float t[3];
    int ns = solve_cubic_f(a3, a2, a1, a0, t);
    for (int i = 0; i < ns; i++) {
        if (t[i] < 0.0f || t[i] > 1.0f) continue;
        V3<float> p[4];
        for (int k = 0; k < 4; k++) p[k] = X0[k] + t[i] * (X1[k] - X0[k]);
        V3<float> n; float w[4], d;
        bool ok, inside;
        if (type == 0) {
            ok = signed_vf_distance(p[0], p[1], p[2], p[3], n, w, d);
            inside = ok && (tmin3(-w[1], -w[2], -w[3]) >= -1e-6f);
        } else {
            ok = signed_ee_distance(p[0], p[1], p[2], p[3], n, w, d);
            inside = ok && (tmin4(w[0], w[1], -w[2], -w[3]) >= -1e-6f);
        }
        if (!ok) continue;
        if (n.dot(w[1] * v1 + w[2] * v2 + w[3] * v3) > 0.0f) n = -n;
        if (tabs(d) < 1e-6f && inside) {
            imp.t = t[i];
            for (int k = 0; k < 4; k++) imp.w[k] = w[k];
            for (int k = 0; k < 3; k++) imp.n[k] = n[k];
            return true;
        }
    }
    return false;
}
template <class T> void load4f(const Scene<T>& S, const int nd[4], V3<float> X0[4], V3<float> X1[4]) {
    for (int k = 0; k < 4; k++)
        for (int c = 0; c < 3; c++) { X0[k][c] = (float)S.x0[nd[k]][c]; X1[k][c] = (float)S.x[nd[k]][c]; }
}

template <class T> void load4(const Scene<T>& S, const int nd[4], V3<double> X0[4], V3<double> X1[4]) {
    for (int k = 0; k < 4; k++)
        for (int c = 0; c < 3; c++) { X0[k][c] = (double)S.x0[nd[k]][c]; X1[k][c] = (double)S.x[nd[k]][c]; }
}

inline int64_t impact_key(int type, int a, int b) { return ((int64_t)type << 62) | ((int64_t)a << 31) | (int64_t)b; }

// checkVertexFaceImpact (src/CollisionHelper.cu:111-122)
template <class T> void vf_impact(const Scene<T>& S, int node, int f, T th, vector<Impact>& out, bool ref32 = false) {
    int n0 = S.fn[f][0], n1 = S.fn[f][1], n2 = S.fn[f][2];
    if (node == n0 || node == n1 || node == n2) return;
    if (!node_box(S, node, true).overlap(face_box(S, f, true), th)) return;
    Impact imp; imp.type = 0;
    imp.nodes[0] = node; imp.nodes[1] = n0; imp.nodes[2] = n1; imp.nodes[3] = n2;
    bool hit;
    if (ref32) {
        V3<float> X0[4], X1[4];
        load4f(S, imp.nodes, X0, X1);
        hit = check_impact_ref32(0, X0, X1, imp);
    } else {
        V3<double> X0[4], X1[4];
        load4(S, imp.nodes, X0, X1);
        hit = check_impact(0, X0, X1, imp);
    }
    if (hit) { imp.key = impact_key(0, node, f); out.push_back(imp); }
}
// checkEdgeEdgeImpact (src/CollisionHelper.cu:124-135); the pair is evaluated in
// canonical order (lower global edge id first) so that one edge pair has one result.
template <class T> void ee_impact(const Scene<T>& S, int e0, int e1, T th, vector<Impact>& out, bool ref32 = false) {
    if (e1 < e0) std::swap(e0, e1);
    int a = S.en[e0][0], b = S.en[e0][1], c = S.en[e1][0], d = S.en[e1][1];
    if (a == c || a == d || b == c || b == d) return;
    if (!edge_box(S, e0, true).overlap(edge_box(S, e1, true), th)) return;
    Impact imp; imp.type = 1;
    imp.nodes[0] = a; imp.nodes[1] = b; imp.nodes[2] = c; imp.nodes[3] = d;
    bool hit;
    if (ref32) {
        V3<float> X0[4], X1[4];
        load4f(S, imp.nodes, X0, X1);
        hit = check_impact_ref32(1, X0, X1, imp);
    } else {
        V3<double> X0[4], X1[4];
        load4(S, imp.nodes, X0, X1);
        hit = check_impact(1, X0, X1, imp);
    }
    if (hit) { imp.key = impact_key(1, e0, e1); out.push_back(imp); }
}

inline bool impact_less(const Impact& p, const Impact& q) { return p.t < q.t || (p.t == q.t && p.key < q.key); }

// Simulator::findImpacts (src/Simulator.cu:383-392) + checkImpactsGpu
// (src/CollisionHelper.cu:137-162), deduplicated and sorted by (t, key).
template <class T> vector<Impact> find_impacts_with(const Scene<T>& S, const vector<BVH<T>>& bvhs, T th,
                                                    bool ref32 = false) {
    vector<std::pair<int, int>> pairs = collect_pairs(S, bvhs, th);
    const long np = (long)pairs.size(), CH = 4096, nch = (np + CH - 1) / CH;
    vector<vector<Impact>> parts(nch);
#pragma omp parallel for schedule(dynamic, 4)
    for (long ch = 0; ch < nch; ch++)
        for (long p = ch * CH; p < std::min(np, (ch + 1) * CH); p++) {
            int f0 = pairs[p].first, f1 = pairs[p].second;
            for (int j = 0; j < 3; j++) vf_impact(S, S.fn[f0][j], f1, th, parts[ch], ref32);
            for (int j = 0; j < 3; j++) vf_impact(S, S.fn[f1][j], f0, th, parts[ch], ref32);
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) ee_impact(S, S.fe[f0][j], S.fe[f1][k], th, parts[ch], ref32);
        }
    vector<Impact> all;
    for (auto& part : parts) all.insert(all.end(), part.begin(), part.end());
    std::sort(all.begin(), all.end(), [](const Impact& p, const Impact& q) { return p.key < q.key; });
    vector<Impact> uniq;
    for (size_t i = 0; i < all.size(); i++)
        if (i == 0 || all[i].key != all[i - 1].key) uniq.push_back(all[i]);
    std::sort(uniq.begin(), uniq.end(), impact_less);
    return uniq;
}

// Simulator::independentImpacts CPU (src/Simulator.cu:394-420): greedy in sorted order.
template <class T> vector<Impact> independent_impacts(const Scene<T>& S, const vector<Impact>& sorted, int deform) {
    std::set<int> taken;
    vector<Impact> ans;
    for (const Impact& imp : sorted) {
        bool ok = true;
        for (int i = 0; i < 4; i++) {
            int nd = imp.nodes[i];
            if ((deform == 1 || S.nodeFree(nd)) && taken.count(nd)) { ok = false; break; }
        }
        if (ok) {
            ans.push_back(imp);
            for (int i = 0; i < 4; i++) {
                int nd = imp.nodes[i];
                if (deform == 1 || S.nodeFree(nd)) taken.insert(nd);
            }
        }
    }
    return ans;
}

// ---------------------------------------------------------------- impact response
// CollisionOptimization + Optimization::solve CPU (src/optimization/CollisionOptimization.cu:3-145,
// src/optimization/Optimization.cu:98-145, clampViolation OptimizationHelper.cu:3-5 with sign = 1).
template <class T> int collision_optimize(Scene<T>& S, const vector<Impact>& impacts, T thickness, int deform, T obstacleMass) {
    int nC = (int)impacts.size();
    vector<std::pair<int, int>> in;
    for (int i = 0; i < nC; i++)
        for (int j = 0; j < 4; j++) {
            int nd = impacts[i].nodes[j];
            if (deform == 1 || S.nodeFree(nd)) in.push_back({nd, 4 * i + j});
        }
    std::sort(in.begin(), in.end());
    vector<int> nodes, idx(4 * nC, -1);
    T invMass = T(0);
    for (size_t i = 0; i < in.size(); i++) {
        if (i == 0 || in[i].first != in[i - 1].first) {
            int nd = in[i].first;
            T m = S.nodeFree(nd) ? S.mass[nd] : obstacleMass;
            invMass += T(1) / m;
            nodes.push_back(nd);
        }
        idx[in[i].second] = (int)nodes.size() - 1;
    }
    int nN = (int)nodes.size();
    invMass /= T(nN);
    vector<T> W(4 * nC);
    vector<V3<T>> Nn(nC);
    for (int i = 0; i < nC; i++) {
        for (int j = 0; j < 4; j++) W[4 * i + j] = (T)(float)impacts[i].w[j];
        for (int c = 0; c < 3; c++) Nn[i][c] = (T)(float)impacts[i].n[c];
    }
    auto nmass = [&](int k) { int nd = nodes[k]; return S.nodeFree(nd) ? S.mass[nd] : obstacleMass; };
    auto objective = [&](const vector<V3<T>>& X) {
        T a = T(0);
        for (int k = 0; k < nN; k++) a += nmass(k) * (X[k] - S.x1[nodes[k]]).norm2();
        return T(0.5f) * a * invMass;
    };
    auto constraint = [&](const vector<V3<T>>& X, int i) {
        T c = -thickness;
        for (int j = 0; j < 4; j++) {
            int k = idx[4 * i + j];
            if (k > -1) c += W[4 * i + j] * Nn[i].dot(X[k]);
            else c += W[4 * i + j] * Nn[i].dot(S.x[impacts[i].nodes[j]]);
        }
        return c;
    };
    T mu = T(1e3f);
    vector<T> lambda(nC, T(0));
    auto value = [&](const vector<V3<T>>& X) {
        T a = objective(X);
        for (int i = 0; i < nC; i++) {
            T coef = tmin(constraint(X, i) + lambda[i] / mu, T(0));
            if (coef != T(0)) a += T(0.5f) * mu * tsqr(coef);
        }
        return a;
    };
    vector<V3<T>> cur(nN), nxt(nN), grad(nN);
    for (int k = 0; k < nN; k++) cur[k] = S.x[nodes[k]];
    T s = T(1e-3f) * T(nN), f = T(0), ft = T(0);
    int iter = 0;
    for (; iter < 100; iter++) {
        f = objective(cur);
        for (int k = 0; k < nN; k++) grad[k] = invMass * nmass(k) * (cur[k] - S.x1[nodes[k]]);
        for (int i = 0; i < nC; i++) {
            T coef = tmin(constraint(cur, i) + lambda[i] / mu, T(0));
            if (coef != T(0)) {
                f += T(0.5f) * mu * tsqr(coef);
                T factor = mu * coef;
                for (int j = 0; j < 4; j++) {
                    int k = idx[4 * i + j];
                    if (k > -1) grad[k] += factor * W[4 * i + j] * Nn[i];
                }
            }
        }
        T norm2 = T(0);
        for (int k = 0; k < nN; k++) norm2 += grad[k].norm2();
        s /= T(0.7f);
        do {
            s *= T(0.7f);
            for (int k = 0; k < nN; k++) nxt[k] = cur[k] - s * grad[k];
            ft = value(nxt);
        } while (ft >= f - T(0.5f) * s * norm2 && s >= T(1e-12f) && tabs(f - ft) >= T(1e-10f));
        if (s < T(1e-12f) || tabs(f - ft) < T(1e-10f)) break;
        cur = nxt;
        for (int i = 0; i < nC; i++) lambda[i] = tmin(lambda[i] + mu * constraint(cur, i), T(0));
    }
    for (int k = 0; k < nN; k++) S.x[nodes[k]] = cur[k];
    return iter;
}

// ---------------------------------------------------------------- steps
// Simulator::physicsStep (src/Simulator.cu:571-600), one cloth system at a time.
template <class T> int physics_step(Scene<T>& S, int directSolver, cs_stats* st) {
    S.started = true;
    // Simulator::simulateStep: nSteps + 1, updateTransformations, Obstacle::step
    // (src/Simulator.cu:838-843,565-573, src/Obstacle.cu:66-85)
    S.steps++;
    if (!S.motions.empty()) {
        T invDt = T(1) / T(S.P.dt);
        for (int o = S.nCloths; o < (int)S.meshes.size(); o++) {
            const MeshRange& R = S.meshes[o];
            if (R.motion < 0) continue;
            Tf X = S.motion_now(R.motion);
            for (int i = R.n0; i < R.n0 + R.nn; i++) S.v[i] = (tf_point(X, S.base[i]) - S.x[i]) * invDt;
        }
    }
    vector<Prox<T>> prox;
    if (!(S.P.flags & CS_FLAG_NO_PROXIMITY)) prox = find_proximities(S);
    if (st) { st->n_proximities = (int)prox.size(); st->pcg_iters = 0; st->pcg_rel_residual = 0; st->pcg_converged = 1; }
    int status = CS_OK;
    for (int c = 0; c < S.nCloths; c++) {
        const MeshRange& R = S.meshes[c];
        System<T> sys;
        assemble(S, c, prox, sys);
        vector<int> rp, ci; vector<T> val, dv;
        sys.to_bsr(rp, ci, val);
        if (directSolver == 2) {
            if (!direct_solve_nd(rp, ci, val, sys.b, dv)) return CS_E_NOT_CONVERGED;
        } else if (directSolver) {
            if (!direct_solve(rp, ci, val, sys.b, dv)) return CS_E_NOT_CONVERGED;
        } else {
            T rr = T(0);
            int it;
            if (S.P.precond == CS_PRECOND_AMG) {
                System<T> ss;  // the static system (no proximity blocks) for the coarse operators
                ss.n = sys.n;
                ss.b.assign(sys.b.size(), T(0));
                assemble(S, c, vector<Prox<T>>(), ss);
                vector<int> srp, sci;
                vector<T> sval;
                ss.to_bsr(srp, sci, sval);
                AmgTopo topo;
                amg_cloth_topology(S, c, topo);
                Amg M;
                amg_setup(topo, rp, ci, val, srp, sci, sval, M);
                it = pcg_with(rp, ci, val, sys.b, T(S.P.pcg_rel_tol), S.P.pcg_max_iters, dv, &rr,
                              [&](const vector<T>& in, vector<T>& out) {
                                  vector<double> bi(in.begin(), in.end()), xo;
                                  amg_vcycle(M, 0, bi, xo);
                                  out.assign(xo.begin(), xo.end());
                              });
            } else {
                it = pcg_block_jacobi(rp, ci, val, sys.b, T(S.P.pcg_rel_tol), S.P.pcg_max_iters, dv, &rr);
            }
            if (st) { st->pcg_iters += it; st->pcg_rel_residual = std::max(st->pcg_rel_residual, (float)rr); }
            if (it >= S.P.pcg_max_iters) { status = CS_E_NOT_CONVERGED; if (st) st->pcg_converged = 0; }
        }
        for (int i = 0; i < R.nn; i++)  // updateNodes (src/PhysicsHelper.cu:501-506)
            for (int a = 0; a < 3; a++) S.v[R.n0 + i][a] += dv[3 * i + a];
    }
    T dt = T(S.P.dt);
    for (size_t i = 0; i < S.x.size(); i++) { S.x0[i] = S.x[i]; S.x[i] += dt * S.v[i]; }  // Mesh::updatePositions
    update_geometry(S, true, false);
    update_geometry(S, false, true);
    return status;
}

// Simulator::collisionStep (src/Simulator.cu:602-684), the GPU control flow.
template <class T> int collision_step(Scene<T>& S, cs_stats* st) {
    vector<BVH<T>> bvhs;
    build_all(S, true, bvhs);
    T th = T(S.P.collision_thickness);
    T obstacleMass = T(1e3f);
    bool success = false;
    int deform = 0, iters = 0, nImp = 0;
    vector<Impact> impacts;
    for (deform = 0; deform < 2; deform++) {
        impacts.clear();
        for (int i = 0; i < S.P.max_collision_iters; i++) {
            vector<Impact> fresh = find_impacts_with(S, bvhs, th);
            if (fresh.empty()) { success = true; break; }
            fresh = independent_impacts(S, fresh, deform);
            impacts.insert(impacts.end(), fresh.begin(), fresh.end());
            collision_optimize(S, impacts, th, deform, obstacleMass);
            iters++;
            for (int m = 0; m < S.nCloths; m++) bvh_refit(S, bvhs[m]);
            if (deform == 1) {
                for (int m = S.nCloths; m < (int)S.meshes.size(); m++) bvh_refit(S, bvhs[m]);
                obstacleMass *= T(0.5f);
            }
        }
        if (success) break;
    }
    nImp = (int)impacts.size();
    if (st) { st->collision_iters = iters; st->collision_deform = deform; st->n_impacts = nImp; }
    if (!success) return CS_E_COLLISION_FAILED;
    update_geometry(S, true, false);
    T invDt = T(1) / T(S.P.dt);
    for (int c = 0; c < S.nCloths; c++) {  // Mesh::updateVelocities (src/Mesh.cu:238-247)
        const MeshRange& R = S.meshes[c];
        for (int i = R.n0; i < R.n0 + R.nn; i++) S.v[i] = (S.x[i] - S.x0[i]) * invDt;
    }
    if (deform == 1) {
        update_geometry(S, false, true);
        for (int o = S.nCloths; o < (int)S.meshes.size(); o++) {
            const MeshRange& R = S.meshes[o];
            for (int i = R.n0; i < R.n0 + R.nn; i++) S.v[i] = (S.x[i] - S.x0[i]) * invDt;
        }
    }
    return CS_OK;
}

}  // namespace orc
