// cs_remesh.hpp — the edge operators of the remeshing step on one cloth's topology (host).
//
// The mesh keeps the reference's element records (Node, Vertex, Edge, Face; src/Node.cuh,
// src/Vertex.cuh, src/Edge.cuh, src/Face.cuh) as integer ids into growing pools, plus the four
// element lists whose order the operators depend on.  An edge has nodes (n0, n1) and, per side s,
// the side face's vertices at n0 and n1, the opposite vertex and the face (Edge::initialize,
// src/Edge.cu:13-31); -1 stands for the reference's null pointer.
#pragma once
#include <array>
#include <vector>

namespace cs {

struct RemeshParams {  // the cloth's "remeshing" block (src/Remeshing.cu:3-19)
    float refineAngle, refineVelocity, refineCompression, ribStiffening;
    float sizeMin, sizeMax, aspectMin, flipThreshold;
};

struct RMesh {
    // nodes
    std::vector<std::array<float, 3>> X, X0, V;
    std::vector<char> pres;
    // vertices
    std::vector<std::array<float, 2>> U;
    std::vector<std::array<float, 4>> S;  // sizing, row-major 2x2
    std::vector<int> vnode;
    // edges
    std::vector<std::array<int, 2>> en, eopp, eadj;
    std::vector<std::array<int, 4>> ev;  // side s: ev[2s] at n0, ev[2s+1] at n1
    // faces
    std::vector<std::array<int, 3>> fv, fe;
    // element lists
    std::vector<int> nodes, verts, edges, faces;
    // independent-set size of the last flip pass (a function-level static in the reference,
    // src/Cloth.cu:321; kept per context here)
    int flipCount = 0;
};

// Builds the records from compact arrays; edges are taken in the given order and orientation,
// their sides from the faces.  Returns false on a face edge missing from `en`.
bool rmesh_build(RMesh& m, int nn, const float* x, const float* x0, const float* v, const unsigned char* preserve,
                 int nv, const float* uv, const int* vnode, int nf, const int* fv, int ne, const int* en);
// Vertex sizing from per-face sizing (4 floats per face of the face list, in list order).
void rmesh_set_face_sizing(RMesh& m, const float* fs);
// flip, split, collapse (Cloth::remeshingStep, src/Cloth.cu:815-833).
void rmesh_remesh(RMesh& m, const RemeshParams& r);

struct RMeshOut {
    std::vector<float> x, x0, v, uv, sizing;
    std::vector<unsigned char> preserve;
    std::vector<int> vnode, fv, en;
    std::vector<int> fe, eopp, eadj;  // face edges; per edge side the opposite vertex and the face (-1: none)
    std::vector<int> nodeOrigin;  // per output node: its input node id, or -1 for a new node
};
void rmesh_output(const RMesh& m, int nInputNodes, RMeshOut& o);

}  // namespace cs
