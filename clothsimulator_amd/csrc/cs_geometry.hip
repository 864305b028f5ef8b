// cs_geometry.hip — per-step geometry of all meshes.
//
// Face normals (and areas of obstacle faces), node normals/areas/masses, the
// explicit position update and the velocity update.  Node quantities are
// gathered over the node's incident faces in ascending face order, so the sums
// run in the same order as the reference CPU loop over faces
// (src/Mesh.cu:182-214) without any float atomics.
#include "cs_kernels.hpp"
#include "cs_math.hpp"

namespace cs {

namespace {

// Face::update (src/Face.cu:124-132)
__global__ void k_face_geometry(int f0, int nf, int nClothFaces, const float4* __restrict__ x,
                                const int4* __restrict__ fnodes, float4* __restrict__ fnrm,
                                float* __restrict__ farea, float* __restrict__ fmass) {
    int f = f0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= f0 + nf) return;
    int4 fn = fnodes[f];
    V3f p0 = ld3(x[fn.x]), p1 = ld3(x[fn.y]), p2 = ld3(x[fn.z]);
    V3f c = cross(p1 - p0, p2 - p0);
    fnrm[f] = st4(normalized(c));
    if (f >= nClothFaces) {  // obstacle faces (not free): current area, zero mass
        farea[f] = 0.5f * norm(c);
        fmass[f] = 0.0f;
    }
}

// Mesh::updateNodeGeometries (src/Mesh.cu:182-214)
__global__ void k_node_geometry(int n0, int nn, const float4* __restrict__ x, const int4* __restrict__ fnodes,
                                const float* __restrict__ farea, const float* __restrict__ fmass,
                                const int* __restrict__ nfPtr, const int* __restrict__ nfIdx,
                                float4* __restrict__ x1, float4* __restrict__ nrm, float* __restrict__ area,
                                float* __restrict__ mass) {
    int i = n0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n0 + nn) return;
    float4 xi4 = x[i];
    V3f xi = ld3(xi4);
    V3f n = mk(0.0f, 0.0f, 0.0f);
    float a = 0.0f, m = 0.0f;
    for (int k = nfPtr[i]; k < nfPtr[i + 1]; k++) {
        int code = nfIdx[k];
        int f = code >> 2, loc = code & 3;
        int4 fn = fnodes[f];
        int o1 = loc == 0 ? fn.y : (loc == 1 ? fn.z : fn.x);
        int o2 = loc == 0 ? fn.z : (loc == 1 ? fn.x : fn.y);
        V3f e0 = ld3(x[o1]) - xi;
        V3f e1 = ld3(x[o2]) - xi;
        n += cross(e0, e1) / (norm2(e0) * norm2(e1));
        a += farea[f] / 3.0f;
        m += fmass[f] / 3.0f;
    }
    x1[i] = xi4;
    nrm[i] = st4(normalized(n));
    area[i] = a;
    mass[i] = m;
}

// Mesh::updatePositions (src/Mesh.cu:226-236)
__global__ void k_update_positions(int n0, int nn, float dt, float4* __restrict__ x, float4* __restrict__ x0,
                                   const float4* __restrict__ v) {
    int i = n0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n0 + nn) return;
    float4 p = x[i];
    V3f vi = ld3(v[i]);
    x0[i] = p;
    x[i] = st4(ld3(p) + dt * vi);
}

// Mesh::updateVelocities (src/Mesh.cu:238-247)
__global__ void k_update_velocities(int n0, int nn, float invDt, const float4* __restrict__ x,
                                    const float4* __restrict__ x0, float4* __restrict__ v) {
    int i = n0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n0 + nn) return;
    v[i] = st4((ld3(x[i]) - ld3(x0[i])) * invDt);
}

// updateNodes (src/PhysicsHelper.cu:501-506): v += dv
__global__ void k_apply_dv(int n0, int nn, const float* __restrict__ dv, float4* __restrict__ v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    float4 a = v[n0 + i];
    a.x += dv[3 * i];
    a.y += dv[3 * i + 1];
    a.z += dv[3 * i + 2];
    v[n0 + i] = a;
}

}  // namespace

void launch_face_geometry(cs_ctx* C, int f0, int nf) {
    if (nf <= 0) return;
    hipLaunchKernelGGL(k_face_geometry, dim3(cdiv(nf, 256)), dim3(256), 0, C->stream, f0, nf, C->FC,
                       C->x.ptr<float4>(), C->fnodes.ptr<int4>(), C->fnrm.ptr<float4>(), C->farea.ptr<float>(),
                       C->fmass.ptr<float>());
    CS_LAUNCH_CHECK();
}

void launch_node_geometry(cs_ctx* C, int n0, int nn) {
    if (nn <= 0) return;
    hipLaunchKernelGGL(k_node_geometry, dim3(cdiv(nn, 256)), dim3(256), 0, C->stream, n0, nn, C->x.ptr<float4>(),
                       C->fnodes.ptr<int4>(), C->farea.ptr<float>(), C->fmass.ptr<float>(), C->nfPtr.ptr<int>(),
                       C->nfIdx.ptr<int>(), C->x1.ptr<float4>(), C->nrm.ptr<float4>(), C->area.ptr<float>(),
                       C->mass.ptr<float>());
    CS_LAUNCH_CHECK();
}

void launch_update_positions(cs_ctx* C, int n0, int nn) {
    if (nn <= 0) return;
    hipLaunchKernelGGL(k_update_positions, dim3(cdiv(nn, 256)), dim3(256), 0, C->stream, n0, nn, C->P.dt,
                       C->x.ptr<float4>(), C->x0.ptr<float4>(), C->v.ptr<float4>());
    CS_LAUNCH_CHECK();
}

void launch_update_velocities(cs_ctx* C, int n0, int nn) {
    if (nn <= 0) return;
    float invDt = 1.0f / C->P.dt;
    hipLaunchKernelGGL(k_update_velocities, dim3(cdiv(nn, 256)), dim3(256), 0, C->stream, n0, nn, invDt,
                       C->x.ptr<float4>(), C->x0.ptr<float4>(), C->v.ptr<float4>());
    CS_LAUNCH_CHECK();
}

void launch_apply_dv(cs_ctx* C, int c) {
    const MeshInfo& M = C->meshes[c];
    hipLaunchKernelGGL(k_apply_dv, dim3(cdiv(M.nn, 256)), dim3(256), 0, C->stream, M.n0, M.nn,
                       C->sys[c].x.ptr<float>(), C->v.ptr<float4>());
    CS_LAUNCH_CHECK();
}

}  // namespace cs
