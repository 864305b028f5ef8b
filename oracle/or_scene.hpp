// oracle/or_scene.hpp — CPU restatement of the reference cloth step: data model,
// material, geometry, forces and the implicit-Euler system.
//
// TEST INFRASTRUCTURE ONLY (see or_math.hpp).  Scene<float> restates the
// reference's FP32 CPU path; Scene<double> is the FP64 oracle the product's
// errors are stated against.  Every function cites the reference file:line it
// follows (paths relative to the reference root).
#pragma once
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "../include/clothsim.h"
#include "or_math.hpp"
#include "or_motion.hpp"

namespace orc {

static const int MAT_N = 30;  // src/Material.cuh:16

// ---------------------------------------------------------------- material
struct Material {
    float density = 0.0f;
    std::vector<float> table;  // [30][30][30][4]
    float bend[3][5];
    float& t(int i, int j, int k, int c) { return table[(((i * MAT_N) + j) * MAT_N + k) * 4 + c]; }
    float t(int i, int j, int k, int c) const { return table[(((i * MAT_N) + j) * MAT_N + k) * 4 + c]; }
};

// src/MathHelper.cu:3-36 (2x2 symmetric eigen-decomposition, FP32)
inline void eigen2(float a, float b, float d, float Q[2][2], float l[2]) {
    float amd = a - d;
    float apd = a + d;
    float b2 = b * b;
    float det = std::sqrt(4.0f * b2 + tsqr(amd));
    float l1 = 0.5f * (apd + det);
    float l2 = 0.5f * (apd - det);
    l[0] = l1;
    l[1] = l2;
    if (b != 0.0f) {
        float v0 = l1 - d, v1 = b;
        float vn = std::sqrt(tsqr(v0) + b2);
        Q[0][0] = v0 / vn;
        Q[1][0] = v1 / vn;
        v0 = l2 - d;
        vn = std::sqrt(tsqr(v0) + b2);
        Q[0][1] = v0 / vn;
        Q[1][1] = v1 / vn;
    } else if (a >= d) {
        Q[0][0] = 1.0f; Q[1][0] = 0.0f; Q[0][1] = 0.0f; Q[1][1] = 1.0f;
    } else {
        Q[0][0] = 0.0f; Q[1][0] = 1.0f; Q[0][1] = 1.0f; Q[1][1] = 0.0f;
    }
}

// src/Material.cu:45-85 (one stretching sample, FP32)
inline void stretching_sample(float G00, float G01, float G11, const float data[2][5][4], float out[4]) {
    float Q[2][2], l[2];
    eigen2(2.0f * G00, 2.0f * G01, 2.0f * G11, Q, l);
    if (l[0] == l[1] && Q[0][0] == 0.0f && Q[1][1] == 0.0f) {
        Q[0][0] = 1.0f; Q[0][1] = 0.0f; Q[1][0] = 0.0f; Q[1][1] = 1.0f;
    }
    if (Q[0][0] < 0.0f) { Q[0][0] = -Q[0][0]; Q[1][0] = -Q[1][0]; }
    if (Q[1][1] < 0.0f) { Q[0][1] = -Q[0][1]; Q[1][1] = -Q[1][1]; }
    float strainWeight = (std::sqrt(l[0] + 1.0f) - 1.0f) * 6.0f;
    strainWeight = tclamp(strainWeight, 0.0f, 1.0f - 1e-6f);
    int strainId = static_cast<int>(strainWeight);
    strainWeight -= strainId;
    // the reference divides a float by the double M_PI here
    float angleWeight = static_cast<float>(static_cast<double>(tabs(std::atan2(Q[1][0], Q[0][0]))) / M_PI * 8.0f);
    angleWeight = tclamp(angleWeight, 0.0f, 4.0f - 1e-6f);
    int angleId = static_cast<int>(angleWeight);
    angleWeight -= angleId;
    float w[2][2];
    w[0][0] = (1.0f - strainWeight) * (1.0f - angleWeight);
    w[0][1] = (1.0f - strainWeight) * angleWeight;
    w[1][0] = strainWeight * (1.0f - angleWeight);
    w[1][1] = strainWeight * angleWeight;
    float ans[4] = {0, 0, 0, 0};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int c = 0; c < 4; c++) ans[c] += data[strainId + i][angleId + j][c] * w[i][j];
    for (int c = 0; c < 4; c++) out[c] = tmax(ans[c], 0.0f) * 2.0f;
}

// src/Material.cu:3-41
inline void material_build(const cs_material_in& in, Material& m) {
    float thicken = in.thicken;
    float densityMult = in.density_mult * thicken;
    float stretchMult = in.stretching_mult * thicken;
    float bendMult = in.bending_mult * thicken;
    m.density = in.density * densityMult;
    float data[2][5][4];
    for (int c = 0; c < 4; c++) data[0][0][c] = in.stretching[0][c];
    for (int i = 1; i < 5; i++) for (int c = 0; c < 4; c++) data[0][i][c] = data[0][0][c];
    for (int i = 0; i < 5; i++) for (int c = 0; c < 4; c++) data[1][i][c] = in.stretching[i + 1][c];
    m.table.assign(MAT_N * MAT_N * MAT_N * 4, 0.0f);
    for (int i = 0; i < MAT_N; i++)
        for (int j = 0; j < MAT_N; j++)
            for (int k = 0; k < MAT_N; k++) {
                float G00 = static_cast<float>(i) / MAT_N - 0.25f;
                float G11 = static_cast<float>(j) / MAT_N - 0.25f;
                float G01 = static_cast<float>(k) / MAT_N;
                float s[4];
                stretching_sample(G00, G01, G11, data, s);
                for (int c = 0; c < 4; c++) m.t(i, j, k, c) = s[c] * stretchMult;
            }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 5; j++) m.bend[i][j] = in.bending[i][j] * bendMult;
}

// src/Material.cu:91-124 (trilinear lookup)
template <class T> inline Vec<T, 4> stretching_stiffness(const Material& m, T G00, T G01, T G11) {
    const T lim = T(MAT_N - 1) - T(1e-6f);
    T x = (G00 + T(0.25f)) * T(MAT_N);
    x = tclamp(x, T(0), lim);
    int xi = static_cast<int>(x);
    x -= xi;
    T y = (G11 + T(0.25f)) * T(MAT_N);
    y = tclamp(y, T(0), lim);
    int yi = static_cast<int>(y);
    y -= yi;
    T z = tabs(G01) * T(MAT_N);
    z = tclamp(z, T(0), lim);
    int zi = static_cast<int>(z);
    z -= zi;
    T w[2][2][2];
    w[0][0][0] = (T(1) - x) * (T(1) - y) * (T(1) - z);
    w[0][0][1] = (T(1) - x) * (T(1) - y) * z;
    w[0][1][0] = (T(1) - x) * y * (T(1) - z);
    w[0][1][1] = (T(1) - x) * y * z;
    w[1][0][0] = x * (T(1) - y) * (T(1) - z);
    w[1][0][1] = x * (T(1) - y) * z;
    w[1][1][0] = x * y * (T(1) - z);
    w[1][1][1] = x * y * z;
    Vec<T, 4> ans;
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 2; k++)
                for (int c = 0; c < 4; c++) ans[c] += T(m.t(xi + i, yi + j, zi + k, c)) * w[i][j][k];
    return ans;
}

// src/Material.cu:126-155 (bilinear bending lookup)
template <class T> inline T bending_stiffness(const Material& m, T length, T angle, T area, T du0, T du1) {
    T curv = length * angle / area;
    T alpha = T(0.5f) * curv;
    T biasWeight = static_cast<T>(static_cast<double>(tabs(std::atan2(du1, du0)) * T(4)) / M_PI);
    if (biasWeight > T(2)) biasWeight = T(4) - biasWeight;
    biasWeight = tclamp(biasWeight, T(0), T(2) - T(1e-6f));
    int biasId = static_cast<int>(biasWeight);
    biasWeight -= biasId;
    T valueWeight = T(0.2f) * alpha;
    valueWeight = tmin(valueWeight, T(4) - T(1e-6f));
    int valueId = static_cast<int>(valueWeight);
    valueId = tmax(valueId, 0);
    valueWeight -= valueId;
    T w[2][2];
    w[0][0] = (T(1) - biasWeight) * (T(1) - valueWeight);
    w[0][1] = (T(1) - biasWeight) * valueWeight;
    w[1][0] = biasWeight * (T(1) - valueWeight);
    w[1][1] = biasWeight * valueWeight;
    T ans = T(0);
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) ans += T(m.bend[biasId + i][valueId + j]) * w[i][j];
    return tmax(ans, T(0));
}

// ---------------------------------------------------------------- scene
struct MeshRange {
    bool cloth;
    int n0, nn;  // nodes
    int v0, nv;  // uv vertices
    int f0, nf;  // faces
    int e0, ne;  // edges
    int material;
    int motion = -1;  // obstacles: the motion moving it (src/Obstacle.cuh:19)
};

template <class T> struct Scene {
    cs_params P;
    std::vector<MeshRange> meshes;  // cloths first, then obstacles
    int nCloths = 0, nObstacles = 0;
    bool started = false;
    // nodes
    std::vector<V3<T>> x, x0, x1, v, nrm;
    std::vector<T> area, mass;
    std::vector<uint8_t> isFree;
    std::vector<int> nodeMesh;
    // uv vertices
    std::vector<V2<T>> u;
    std::vector<int> vnode;
    // faces
    std::vector<std::array<int, 3>> fv, fn, fe;
    std::vector<Mat<T, 2, 2>> finv;
    std::vector<T> farea, fmass;
    std::vector<V3<T>> fnrm;
    std::vector<uint8_t> ffree;
    // edges: nodes (min,max), per side: uv vertices (2), opposite uv vertex, adjacent face
    std::vector<std::array<int, 2>> en, eopp, eadj;
    std::vector<std::array<std::array<int, 2>, 2>> ev;
    // handles (src/Handle.cuh:9-16): anchor in the motion's frame (motion -1: world)
    struct Handle { int cloth; int node; V3<T> pos; int motion; };
    std::vector<Handle> handles;
    // motions, the step counter they are evaluated at (src/Simulator.cu:565-569) and the
    // base shape of every moving obstacle's nodes (src/Obstacle.cu:21-36)
    std::vector<Motion> motions;
    int steps = 0;
    std::vector<V3<T>> base;
    Tf motion_now(int m) const { return motions[m].at(float(steps) * P.dt); }
    std::vector<Material> materials;

    bool nodeFree(int i) const { return isFree[i] != 0; }
    bool faceFree(int f) const { return isFree[fn[f][0]] && isFree[fn[f][1]] && isFree[fn[f][2]]; }  // Face.cu:50-52
    bool edgeFree(int e) const { return isFree[en[e][0]] && isFree[en[e][1]]; }  // Edge.cu:77-79
    bool edgeBoundary(int e) const { return eadj[e][0] < 0 || eadj[e][1] < 0; }  // Edge.cu:81-83
    T edgeArea(int e) const {  // Edge.cu:96-102
        T a = T(0);
        for (int s = 0; s < 2; s++) if (eadj[e][s] >= 0) a += farea[eadj[e][s]];
        return a;
    }
    int meshOfNode(int i) const { return nodeMesh[i]; }
};

// Mesh::load topology (src/Mesh.cu:353-452, src/MeshHelper.cu:34-102, src/Edge.cu:13-31,
// src/Face.cu:12-20).  Edges are the sorted unique (min,max) node pairs; each
// face-side initialises the edge side that matches its orientation.
template <class T> int scene_add_mesh(Scene<T>& S, const cs_mesh_in& in, bool cloth, int materialIndex) {
    if (in.n_nodes <= 0 || in.n_faces <= 0 || in.n_verts <= 0 || !in.x || !in.uv || !in.face_nodes || !in.face_verts)
        return CS_E_ARG;
    for (int k = 0; k < 3 * in.n_faces; k++) {
        if (in.face_nodes[k] < 0 || in.face_nodes[k] >= in.n_nodes) return CS_E_FORMAT;
        if (in.face_verts[k] < 0 || in.face_verts[k] >= in.n_verts) return CS_E_FORMAT;
    }
    MeshRange R;
    R.cloth = cloth;
    R.n0 = (int)S.x.size(); R.nn = in.n_nodes;
    R.v0 = (int)S.u.size(); R.nv = in.n_verts;
    R.f0 = (int)S.fv.size(); R.nf = in.n_faces;
    R.e0 = (int)S.en.size();
    R.material = materialIndex;
    int meshId = (int)S.meshes.size();
    for (int i = 0; i < in.n_nodes; i++) {
        V3<T> p = mk3<T>(in.x[3 * i], in.x[3 * i + 1], in.x[3 * i + 2]);
        V3<T> vel = in.v ? mk3<T>(in.v[3 * i], in.v[3 * i + 1], in.v[3 * i + 2]) : V3<T>();
        S.x.push_back(p); S.x0.push_back(p); S.x1.push_back(p); S.v.push_back(vel);
        S.nrm.push_back(V3<T>()); S.area.push_back(T(0)); S.mass.push_back(T(0));
        S.isFree.push_back(cloth ? 1 : 0); S.nodeMesh.push_back(meshId);
    }
    for (int i = 0; i < in.n_verts; i++) { S.u.push_back(mk2<T>(in.uv[2 * i], in.uv[2 * i + 1])); S.vnode.push_back(-1); }
    const Material* mat = cloth ? &S.materials[materialIndex] : nullptr;
    struct EdgeOcc { int a, b, opp, face; };
    std::vector<EdgeOcc> occ;
    occ.reserve(3 * in.n_faces);
    for (int f = 0; f < in.n_faces; f++) {
        std::array<int, 3> nodes, verts;
        for (int k = 0; k < 3; k++) {
            nodes[k] = R.n0 + in.face_nodes[3 * f + k];
            verts[k] = R.v0 + in.face_verts[3 * f + k];
            S.vnode[verts[k]] = nodes[k];
        }
        S.fv.push_back(verts); S.fn.push_back(nodes); S.fe.push_back({-1, -1, -1});
        Mat<T, 2, 2> inv; T a = T(0), m = T(0);
        if (mat) {  // Face::initialize, src/Face.cu:12-20
            V2<T> d1 = S.u[verts[1]] - S.u[verts[0]];
            V2<T> d2 = S.u[verts[2]] - S.u[verts[0]];
            Mat<T, 2, 2> D; D(0, 0) = d1[0]; D(1, 0) = d1[1]; D(0, 1) = d2[0]; D(1, 1) = d2[1];
            inv = inverse2(D);
            a = T(0.5f) * tabs(cross2(d1, d2));
            m = a * T(mat->density);
        }
        S.finv.push_back(inv); S.farea.push_back(a); S.fmass.push_back(m);
        S.fnrm.push_back(V3<T>()); S.ffree.push_back(cloth ? 1 : 0);
        int fg = R.f0 + f;
        for (int k = 0; k < 3; k++) {
            int a0 = nodes[k], a1 = nodes[(k + 1) % 3];
            EdgeOcc o; o.a = tmin(a0, a1); o.b = tmax(a0, a1); o.opp = verts[(k + 2) % 3]; o.face = fg;
            occ.push_back(o);
        }
    }
    std::stable_sort(occ.begin(), occ.end(), [](const EdgeOcc& p, const EdgeOcc& q) {
        return p.a < q.a || (p.a == q.a && p.b < q.b);
    });
    for (size_t k = 0; k < occ.size(); k++) {
        if (k == 0 || occ[k].a != occ[k - 1].a || occ[k].b != occ[k - 1].b) {
            S.en.push_back({occ[k].a, occ[k].b});
            S.eopp.push_back({-1, -1}); S.eadj.push_back({-1, -1});
            S.ev.push_back({std::array<int, 2>{-1, -1}, std::array<int, 2>{-1, -1}});
        }
        int e = (int)S.en.size() - 1;
        int f = occ[k].face;
        // Edge::initialize (src/Edge.cu:13-31)
        for (int i = 0; i < 3; i++) {
            int vert0 = S.fv[f][i], vert1 = S.fv[f][(i + 1) % 3];
            if (S.en[e][0] == S.vnode[vert0] && S.en[e][1] == S.vnode[vert1]) {
                S.ev[e][0] = {vert0, vert1}; S.eopp[e][0] = occ[k].opp; S.eadj[e][0] = f; break;
            } else if (S.en[e][1] == S.vnode[vert0] && S.en[e][0] == S.vnode[vert1]) {
                S.ev[e][1] = {vert1, vert0}; S.eopp[e][1] = occ[k].opp; S.eadj[e][1] = f; break;
            }
        }
        // Face::setEdge (src/Face.cu:22-33)
        for (int i = 0; i < 3; i++) {
            int n0 = S.fn[f][i], n1 = S.fn[f][(i + 1) % 3];
            if ((n0 == S.en[e][0] && n1 == S.en[e][1]) || (n1 == S.en[e][0] && n0 == S.en[e][1])) {
                S.fe[f][i] = e; break;
            }
        }
    }
    R.ne = (int)S.en.size() - R.e0;
    S.meshes.push_back(R);
    return CS_OK;
}

// Face::update (src/Face.cu:124-132)
template <class T> void update_face_geometry(Scene<T>& S, const MeshRange& R) {
    for (int f = R.f0; f < R.f0 + R.nf; f++) {
        V3<T> d1 = S.x[S.fn[f][1]] - S.x[S.fn[f][0]];
        V3<T> d2 = S.x[S.fn[f][2]] - S.x[S.fn[f][0]];
        V3<T> c = cross(d1, d2);
        S.fnrm[f] = c.normalized();
        if (!S.faceFree(f)) {
            S.farea[f] = T(0.5f) * c.norm();
            S.fmass[f] = T(0);
        }
    }
}

// Mesh::updateNodeGeometries CPU (src/Mesh.cu:182-214): faces visited in order.
template <class T> void update_node_geometry(Scene<T>& S, const MeshRange& R) {
    for (int i = R.n0; i < R.n0 + R.nn; i++) {
        S.x1[i] = S.x[i]; S.nrm[i] = V3<T>(); S.area[i] = T(0); S.mass[i] = T(0);
    }
    for (int f = R.f0; f < R.f0 + R.nf; f++) {
        T a = S.farea[f] / T(3);
        T m = S.fmass[f] / T(3);
        for (int i = 0; i < 3; i++) {
            int node = S.fn[f][i];
            V3<T> e0 = S.x[S.fn[f][(i + 1) % 3]] - S.x[node];
            V3<T> e1 = S.x[S.fn[f][(i + 2) % 3]] - S.x[node];
            S.nrm[node] += cross(e0, e1) / (e0.norm2() * e1.norm2());
            S.area[node] += a;
            S.mass[node] += m;
        }
    }
    for (int i = R.n0; i < R.n0 + R.nn; i++) S.nrm[i].normalize();
}

template <class T> void update_geometry(Scene<T>& S, bool cloths, bool obstacles) {
    for (const MeshRange& R : S.meshes)
        if ((R.cloth && cloths) || (!R.cloth && obstacles)) update_face_geometry(S, R);
    for (const MeshRange& R : S.meshes)
        if ((R.cloth && cloths) || (!R.cloth && obstacles)) update_node_geometry(S, R);
}

}  // namespace orc
