// cs_remesh_api.cpp — C-ABI of the host-side remeshing operators (include/clothsim.h).
#include <cmath>
#include <algorithm>
#include <new>

#include "../../include/clothsim.h"
#include "cs_remesh.hpp"

struct cs_remesh_result {
    cs::RMeshOut o;
};

namespace {
template <class T> void put(T* dst, const std::vector<T>& src) {
    if (dst) std::copy(src.begin(), src.end(), dst);
}
}  // namespace

extern "C" {

void cs_default_remeshing(cs_remeshing* r) {
    if (!r) return;
    r->refine_angle = r->refine_velocity = r->refine_compression = INFINITY;
    r->rib_stiffening = 1.0f;
    r->size_min = -INFINITY;
    r->size_max = INFINITY;
    r->aspect_min = -INFINITY;
    r->flip_threshold = 1e-2f;
}

cs_status cs_remesh_mesh(const cs_remeshing* r, int n_nodes, const float* x, const float* x0, const float* v,
                         const unsigned char* preserve, int n_verts, const float* uv, const int* vert_node,
                         int n_faces, const int* face_verts, int n_edges, const int* edge_nodes,
                         const float* face_sizing, int* flip_count, cs_remesh_result** out) {
    if (!r || !x || !x0 || !v || !preserve || !uv || !vert_node || !face_verts || !edge_nodes || !face_sizing ||
        !flip_count || !out || n_nodes <= 0 || n_verts <= 0 || n_faces <= 0 || n_edges <= 0)
        return CS_E_ARG;
    *out = nullptr;
    for (int k = 0; k < n_verts; k++)
        if (vert_node[k] < 0 || vert_node[k] >= n_nodes) return CS_E_FORMAT;
    for (long k = 0; k < 3L * n_faces; k++)
        if (face_verts[k] < 0 || face_verts[k] >= n_verts) return CS_E_FORMAT;
    for (long k = 0; k < 2L * n_edges; k++)
        if (edge_nodes[k] < 0 || edge_nodes[k] >= n_nodes) return CS_E_FORMAT;
    try {
        cs::RMesh m;
        if (!cs::rmesh_build(m, n_nodes, x, x0, v, preserve, n_verts, uv, vert_node, n_faces, face_verts, n_edges,
                             edge_nodes))
            return CS_E_FORMAT;
        m.flipCount = *flip_count;
        const cs::RemeshParams p{r->refine_angle, r->refine_velocity, r->refine_compression, r->rib_stiffening,
                                 r->size_min, r->size_max, r->aspect_min, r->flip_threshold};
        cs::rmesh_set_face_sizing(m, face_sizing);
        cs::rmesh_remesh(m, p);
        *flip_count = m.flipCount;
        cs_remesh_result* res = new cs_remesh_result;
        cs::rmesh_output(m, n_nodes, res->o);
        *out = res;
        return CS_OK;
    } catch (const std::bad_alloc&) {
        return CS_E_OOM;
    }
}

cs_status cs_remesh_result_counts(const cs_remesh_result* res, int* n_nodes, int* n_verts, int* n_faces,
                                  int* n_edges) {
    if (!res) return CS_E_ARG;
    if (n_nodes) *n_nodes = (int)res->o.preserve.size();
    if (n_verts) *n_verts = (int)res->o.vnode.size();
    if (n_faces) *n_faces = (int)(res->o.fv.size() / 3);
    if (n_edges) *n_edges = (int)(res->o.en.size() / 2);
    return CS_OK;
}

cs_status cs_remesh_result_get(const cs_remesh_result* res, float* x, float* x0, float* v, unsigned char* preserve,
                               float* uv, int* vert_node, int* face_verts, int* edge_nodes, float* vert_sizing,
                               int* node_origin) {
    if (!res) return CS_E_ARG;
    put(x, res->o.x);
    put(x0, res->o.x0);
    put(v, res->o.v);
    put(preserve, res->o.preserve);
    put(uv, res->o.uv);
    put(vert_node, res->o.vnode);
    put(face_verts, res->o.fv);
    put(edge_nodes, res->o.en);
    put(vert_sizing, res->o.sizing);
    put(node_origin, res->o.nodeOrigin);
    return CS_OK;
}

void cs_remesh_result_free(cs_remesh_result* res) { delete res; }

}  // extern "C"
