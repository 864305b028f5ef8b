// cs_api.cpp — the C-ABI of libclothsim.so (include/clothsim.h).
//
// Every entry point runs inside guard(): HIP failures and capacity/state
// errors come back as cs_status with a message for cs_last_error, never as an
// exception or exit().  Stage order follows Simulator::physicsStep /
// collisionStep (src/Simulator.cu:571-684).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "cs_kernels.hpp"

using namespace cs;

namespace cs {

int ktimer_begin(cs_ctx* C, int kernel) {
    if (!(C->P.flags & CS_FLAG_TIME_KERNELS)) return -1;
    KernelTimer& K = C->kt;
    int pair = (int)K.pending.size();
    while ((int)K.events.size() < 2 * (pair + 1)) {
        hipEvent_t e;
        CS_HIP(hipEventCreate(&e));
        K.events.push_back(e);
    }
    K.pending.push_back({kernel, pair});
    CS_HIP(hipEventRecord(K.events[2 * pair], C->stream));
    return pair;
}

void ktimer_end(cs_ctx* C, int token) {
    if (token < 0) return;
    CS_HIP(hipEventRecord(C->kt.events[2 * token + 1], C->stream));
}

void ktimer_collect(cs_ctx* C) {
    KernelTimer& K = C->kt;
    for (auto& p : K.pending) {
        float ms = 0.0f;
        CS_HIP(hipEventSynchronize(K.events[2 * p.second + 1]));
        CS_HIP(hipEventElapsedTime(&ms, K.events[2 * p.second], K.events[2 * p.second + 1]));
        K.total[p.first] += ms;
        K.count[p.first] += 1;
    }
    K.pending.clear();
}

void* pinned_batch(cs_ctx* C, size_t bytes) {
    if (C->hBatchCap < bytes) {
        if (C->hBatch) CS_HIP(hipHostFree(C->hBatch));
        C->hBatch = nullptr;
        C->hBatchCap = 0;
        CS_HIP(hipHostMalloc(&C->hBatch, bytes, hipHostMallocDefault));
        C->hBatchCap = bytes;
    }
    return C->hBatch;
}

}  // namespace cs

namespace {

template <class F> cs_status guard(cs_ctx* C, F&& f) {
    try {
        return static_cast<cs_status>(f());
    } catch (const StatusError& e) {
        if (C) C->err = e.what();
        return static_cast<cs_status>(e.status);
    } catch (const HipError& e) {
        if (C) C->err = e.what();
        return CS_E_HIP;
    } catch (const std::bad_alloc&) {
        if (C) C->err = "host allocation failed";
        return CS_E_OOM;
    } catch (const std::exception& e) {
        if (C) C->err = e.what();
        return CS_E_HIP;
    }
}

bool valid_params(const cs_params* p) {
    return p && p->dt > 0.0f && p->pcg_rel_tol > 0.0f && p->pcg_rel_tol < 1.0f && p->pcg_max_iters >= 0 &&
           p->max_collision_iters >= 0 && p->repulsion_thickness > 0.0f && p->collision_thickness >= 0.0f &&
           (p->precond == CS_PRECOND_BLOCK_JACOBI || p->precond == CS_PRECOND_LEARNED || p->precond == CS_PRECOND_AMG ||
            p->precond == CS_PRECOND_AMG_LEARNED);
}

const MeshInfo* mesh_of(cs_ctx* C, int kind, int id) {
    if (kind == CS_MESH_CLOTH) return (id >= 0 && id < C->nCloths) ? &C->meshes[id] : nullptr;
    if (kind == CS_MESH_OBSTACLE) return (id >= 0 && id < C->nObstacles) ? &C->meshes[C->nCloths + id] : nullptr;
    return nullptr;
}

struct StageTimer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    float lap(hipStream_t s) {
        CS_HIP(hipStreamSynchronize(s));
        auto t1 = std::chrono::steady_clock::now();
        float ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
        t0 = t1;
        return ms;
    }
};

int physics_step(cs_ctx* C, cs_stats* st) {
    host_finalize(C);
    hipStream_t s = C->stream;
    StageTimer T;
    cs_stats local{};
    motion_step(C);
    if (!(C->P.flags & CS_FLAG_NO_PROXIMITY)) {
        bvh_build_all(C, false);
        local.ms_bvh = T.lap(s);
        find_proximities(C, false);
    } else {
        C->nProx = 0;
    }
    local.n_proximities = C->nProx;
    local.ms_proximity = T.lap(s);
    int status = CS_OK;
    local.pcg_converged = 1;
    const bool batched = (C->P.flags & CS_FLAG_INSTANCES) && C->nCloths > 1 &&
                         (C->P.precond == CS_PRECOND_BLOCK_JACOBI || C->P.precond == CS_PRECOND_AMG);
    if (batched) {
        // independent instances: assemble every cloth, then one batched solve
        local.n_contact_blocks += assemble_all(C);
        local.ms_assemble += T.lap(s);
        std::vector<int> its(C->nCloths), conv(C->nCloths);
        std::vector<float> rel(C->nCloths);
        pcg_solve_batched(C, its.data(), rel.data(), conv.data(), C->P.precond == CS_PRECOND_AMG);
        for (int c = 0; c < C->nCloths; c++) {
            local.pcg_iters += its[c];
            local.pcg_rel_residual = std::max(local.pcg_rel_residual, rel[c]);
            if (!conv[c]) { local.pcg_converged = 0; status = CS_E_NOT_CONVERGED; }
            launch_apply_dv(C, c);
        }
        local.ms_solve += T.lap(s);
    } else {
        for (int c = 0; c < C->nCloths; c++) {
            launch_element_records(C, c);
            launch_assemble_rows(C, c);
            local.n_contact_blocks += assemble_contacts(C, c);
            launch_block_jacobi(C, c);
            local.ms_assemble += T.lap(s);
            float rel = 0.0f;
            int conv = 0;
            local.pcg_iters += pcg_solve(C, c, &rel, &conv);
            local.pcg_rel_residual = std::max(local.pcg_rel_residual, rel);
            if (!conv) { local.pcg_converged = 0; status = CS_E_NOT_CONVERGED; }
            launch_apply_dv(C, c);
            local.ms_solve += T.lap(s);
        }
    }
    launch_update_positions(C, 0, C->N);
    launch_face_geometry(C, 0, C->F);
    launch_node_geometry(C, 0, C->N);
    local.ms_update = T.lap(s);
    ktimer_collect(C);
    if (st) {
        st->pcg_iters = local.pcg_iters;
        st->pcg_rel_residual = local.pcg_rel_residual;
        st->pcg_converged = local.pcg_converged;
        st->n_proximities = local.n_proximities;
        st->n_contact_blocks = local.n_contact_blocks;
        st->ms_bvh = local.ms_bvh;
        st->ms_proximity = local.ms_proximity;
        st->ms_assemble = local.ms_assemble;
        st->ms_solve = local.ms_solve;
        st->ms_update = local.ms_update;
    }
    return status;
}

int collision_step(cs_ctx* C, cs_stats* st) {
    host_finalize(C);
    hipStream_t s = C->stream;
    StageTimer T;
    bvh_build_all(C, true);
    const float msBvh = T.lap(s);
    const float th = C->P.collision_thickness;
    float obstacleMass = 1e3f;
    bool success = false;
    int deform = 0, iters = 0;
    for (deform = 0; deform < 2; deform++) {
        C->nImpAcc = 0;
        for (int i = 0; i < C->P.max_collision_iters; i++) {
            int n = find_impacts(C, th);
            if (n == 0) { success = true; break; }
            C->impAcc.grow_keep(sizeof(cs_impact_rec) * (size_t)(C->nImpAcc + n),
                                sizeof(cs_impact_rec) * (size_t)C->nImpAcc, s);
            int m = independent_impacts(C, C->impSorted.ptr<cs_impact_rec>(), n, deform,
                                        C->impAcc.ptr<cs_impact_rec>() + C->nImpAcc);
            C->nImpAcc += m;
            collision_response(C, C->nImpAcc, deform, obstacleMass);
            iters++;
            bvh_refit_range(C, 0, C->nCloths);
            if (deform == 1) {
                bvh_refit_range(C, C->nCloths, (int)C->meshes.size());
                obstacleMass *= 0.5f;
            }
        }
        if (success) break;
    }
    if (st) {
        st->collision_iters = iters;
        st->collision_deform = deform;
        st->n_impacts = C->nImpAcc;
        st->ms_bvh += msBvh;
    }
    if (!success) {
        C->err = "Collision step failed: impacts remain after max_collision_iters in both deform passes";
        if (st) st->ms_collision = T.lap(s);
        return CS_E_COLLISION_FAILED;
    }
    // cloths come first in the node and face arrays: one launch each over all of them
    launch_face_geometry(C, 0, C->FC);
    launch_node_geometry(C, 0, C->NC);
    launch_update_velocities(C, 0, C->NC);
    if (deform == 1)
        for (int o = C->nCloths; o < (int)C->meshes.size(); o++) {
            const MeshInfo& M = C->meshes[o];
            launch_face_geometry(C, M.f0, M.nf);
            launch_node_geometry(C, M.n0, M.nn);
            launch_update_velocities(C, M.n0, M.nn);
        }
    if (st) st->ms_collision = T.lap(s);
    else CS_HIP(hipStreamSynchronize(s));
    return CS_OK;
}

int step(cs_ctx* C, cs_stats* st) {
    int a = physics_step(C, st);
    if (a != CS_OK && a != CS_E_NOT_CONVERGED) return a;
    if (C->P.flags & CS_FLAG_NO_COLLISION) return a;
    int b = collision_step(C, st);
    return b != CS_OK ? b : a;
}

}  // namespace

extern "C" {

const char* cs_version(void) { return "clothsim-mi355x 0.1 (gfx950)"; }

void cs_default_params(cs_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->dt = 0.005f;
    p->gravity[2] = -9.8f;
    p->wind_density = 1.0f;
    p->handle_stiffness = 1e3f;
    p->collision_stiffness = 1e9f;
    p->repulsion_thickness = 1e-3f;
    p->collision_thickness = 1e-4f;
    p->friction = 0.6f;
    p->obstacle_friction = 0.3f;
    p->pcg_rel_tol = 1e-5f;
    p->pcg_max_iters = 4000;
    p->precond = CS_PRECOND_BLOCK_JACOBI;
    p->max_collision_iters = 100;
}

cs_status cs_create(const cs_params* params, cs_ctx** ctx_out) {
    if (!ctx_out) return CS_E_ARG;
    *ctx_out = nullptr;
    if (!valid_params(params)) return CS_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CS_E_NO_DEVICE;
    cs_ctx* C = new cs_ctx();
    cs_status st = guard(C, [&]() {
        CS_HIP(hipGetDevice(&C->device));
        hipDeviceProp_t prop;
        CS_HIP(hipGetDeviceProperties(&prop, C->device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw StatusError(CS_E_NO_DEVICE, std::string("device is ") + prop.gcnArchName + ", library built for gfx950");
        CS_HIP(hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking));
        CS_HIP(hipStreamCreateWithFlags(&C->side, hipStreamNonBlocking));
        CS_HIP(hipEventCreateWithFlags(&C->evFork, hipEventDisableTiming));
        CS_HIP(hipEventCreateWithFlags(&C->evJoin, hipEventDisableTiming));
        CS_HIP(hipHostMalloc(reinterpret_cast<void**>(&C->hScal), sizeof(PcgScalars), hipHostMallocDefault));
        CS_HIP(hipHostMalloc(reinterpret_cast<void**>(&C->hCount), 64, hipHostMallocDefault));
        C->P = *params;
        return (int)CS_OK;
    });
    if (st != CS_OK) {
        cs_destroy(C);
        return st;
    }
    *ctx_out = C;
    return CS_OK;
}

void cs_destroy(cs_ctx* C) {
    if (!C) return;
    if (C->stream) (void)hipStreamSynchronize(C->stream);
    if (C->side) (void)hipStreamSynchronize(C->side);
    if (C->evFork) (void)hipEventDestroy(C->evFork);
    if (C->evJoin) (void)hipEventDestroy(C->evJoin);
    if (C->hScal) (void)hipHostFree(C->hScal);
    if (C->hBatch) (void)hipHostFree(C->hBatch);
    if (C->hCount) (void)hipHostFree(C->hCount);
    for (hipEvent_t e : C->kt.events) (void)hipEventDestroy(e);
    if (C->copy) (void)hipStreamSynchronize(C->copy);
    for (auto& F : C->frame) {
        if (F.done) (void)hipEventDestroy(F.done);
        if (F.host) (void)hipHostFree(F.host);
    }
    if (C->evSnap) (void)hipEventDestroy(C->evSnap);
    cs_ctx* c = C;
    // device buffers are released by the DBuf destructors
    hipStream_t s = c->stream, s2 = c->side, s3 = c->copy;
    delete c;
    if (s) (void)hipStreamDestroy(s);
    if (s2) (void)hipStreamDestroy(s2);
    if (s3) (void)hipStreamDestroy(s3);
}

const char* cs_last_error(const cs_ctx* C) { return C ? C->err.c_str() : "null context"; }

cs_status cs_set_params(cs_ctx* C, const cs_params* p) {
    if (!C || !valid_params(p)) return CS_E_ARG;
    C->P = *p;
    return CS_OK;
}

cs_status cs_add_cloth(cs_ctx* C, const cs_mesh_in* mesh, const cs_material_in* material, int* cloth_id) {
    if (!C || !mesh || !material) return CS_E_ARG;
    if (C->finalized || C->nObstacles > 0) return CS_E_STATE;
    return guard(C, [&]() {
        HostMaterial m;
        material_build(*material, m);
        C->materials.push_back(m);
        int st = host_add_mesh(C, *mesh, true, (int)C->materials.size() - 1);
        if (st != CS_OK) { C->materials.pop_back(); return st; }
        if (cloth_id) *cloth_id = C->nCloths - 1;
        return (int)CS_OK;
    });
}

cs_status cs_add_obstacle(cs_ctx* C, const cs_mesh_in* mesh, int* obstacle_id) {
    if (!C || !mesh) return CS_E_ARG;
    if (C->finalized) return CS_E_STATE;
    return guard(C, [&]() {
        int st = host_add_mesh(C, *mesh, false, -1);
        if (st != CS_OK) return st;
        if (obstacle_id) *obstacle_id = C->nObstacles - 1;
        return (int)CS_OK;
    });
}

cs_status cs_add_handles(cs_ctx* C, int cloth_id, const int* nodes, int n) {
    return cs_add_moving_handles(C, cloth_id, nodes, n, -1);
}

cs_status cs_add_moving_handles(cs_ctx* C, int cloth_id, const int* nodes, int n, int motion_id) {
    if (!C || n < 0 || (n > 0 && !nodes)) return CS_E_ARG;
    if (C->finalized) return CS_E_STATE;
    if (cloth_id < 0 || cloth_id >= C->nCloths) return CS_E_ARG;
    if (motion_id < -1 || motion_id >= (int)C->motions.size()) return CS_E_ARG;
    const MeshInfo& M = C->meshes[cloth_id];
    for (int i = 0; i < n; i++)
        if (nodes[i] < 0 || nodes[i] >= M.nn) return CS_E_ARG;
    // anchor = T^-1 (x) for a moving handle (src/ClothHelper.cu:12, src/Cloth.cu:663)
    const Xf inv = motion_id >= 0 ? xf_inverse(motion_now(C, motion_id)) : xf_identity();
    for (int i = 0; i < n; i++) {
        cs_ctx::HandleH h;
        h.cloth = cloth_id;
        h.node = M.n0 + nodes[i];
        h.motion = motion_id;
        const float* x = &C->hx[3 * (size_t)h.node];
        if (motion_id >= 0) xf_point(inv, x[0], x[1], x[2], h.p[0], h.p[1], h.p[2]);
        else for (int k = 0; k < 3; k++) h.p[k] = x[k];
        C->handles.push_back(h);
    }
    return CS_OK;
}

cs_status cs_add_motion(cs_ctx* C, const float* times, const cs_transform* keys, int n, int* motion_id) {
    if (!C) return CS_E_ARG;
    if ((int)C->motions.size() >= CS_MAX_MOTIONS) {
        C->err = "at most CS_MAX_MOTIONS motions per context";
        return CS_E_CAPACITY;
    }
    Motion M;
    const int st = motion_build(times, keys, n, M);
    if (st != CS_OK) {
        C->err = "motion keys need n >= 1 and finite, strictly increasing times";
        return (cs_status)st;
    }
    C->motions.push_back(std::move(M));
    if (motion_id) *motion_id = (int)C->motions.size() - 1;
    return CS_OK;
}

cs_status cs_set_obstacle_motion(cs_ctx* C, int obstacle_id, int motion_id) {
    if (!C) return CS_E_ARG;
    if (C->finalized) return CS_E_STATE;
    if (obstacle_id < 0 || obstacle_id >= C->nObstacles) return CS_E_ARG;
    if (motion_id < 0 || motion_id >= (int)C->motions.size()) return CS_E_ARG;
    MeshInfo& M = C->meshes[C->nCloths + obstacle_id];
    if (M.motion >= 0) return CS_E_STATE;  // its base shape is already taken
    const int first = C->meshes[C->nCloths].n0;  // obstacle nodes follow every cloth node
    const size_t need = 3 * (size_t)(M.n0 + M.nn - first);
    if (C->hObsBase.size() < need) C->hObsBase.resize(need, 0.0f);
    // base = the loaded shape; x = x0 = T(base) (src/Obstacle.cu:21-64)
    const Xf T = motion_now(C, motion_id);
    for (int i = M.n0; i < M.n0 + M.nn; i++) {
        float* b = &C->hObsBase[3 * (size_t)(i - first)];
        float* x = &C->hx[3 * (size_t)i];
        for (int k = 0; k < 3; k++) b[k] = x[k];
        xf_point(T, b[0], b[1], b[2], x[0], x[1], x[2]);
    }
    M.motion = motion_id;
    C->anyMovingObstacle = true;
    return CS_OK;
}

cs_status cs_set_step_count(cs_ctx* C, int steps) {
    if (!C || steps < 0) return CS_E_ARG;
    return guard(C, [&]() {
        C->steps = steps;
        motion_place(C);
        return (int)CS_OK;
    });
}

cs_status cs_get_topology(cs_ctx* C, int kind, int id, int* face_nodes, int* face_verts, float* uv, int* n_verts) {
    if (!C) return CS_E_ARG;
    const MeshInfo* M = mesh_of(C, kind, id);
    if (!M) return CS_E_ARG;
    if (n_verts) *n_verts = M->nv;
    for (int f = 0; f < M->nf; f++)
        for (int k = 0; k < 3; k++) {
            const size_t g = 3 * (size_t)(M->f0 + f) + k;
            if (face_nodes) face_nodes[3 * f + k] = C->hfn[g] - M->n0;
            if (face_verts) face_verts[3 * f + k] = C->hfv[g] - M->v0;
        }
    if (uv) std::copy(C->huv.begin() + 2 * (size_t)M->v0, C->huv.begin() + 2 * (size_t)(M->v0 + M->nv), uv);
    return CS_OK;
}

cs_status cs_get_step_count(const cs_ctx* C, int* steps) {
    if (!C || !steps) return CS_E_ARG;
    *steps = C->steps;
    return CS_OK;
}

cs_status cs_physics_step(cs_ctx* C, cs_stats* st) {
    if (!C) return CS_E_ARG;
    if (st) *st = cs_stats{};
    if (C->nCloths == 0) return CS_E_STATE;
    return guard(C, [&]() { return physics_step(C, st); });
}

cs_status cs_collision_step(cs_ctx* C, cs_stats* st) {
    if (!C) return CS_E_ARG;
    if (st) *st = cs_stats{};
    if (C->nCloths == 0) return CS_E_STATE;
    return guard(C, [&]() { return collision_step(C, st); });
}

cs_status cs_step(cs_ctx* C, cs_stats* st) {
    if (!C) return CS_E_ARG;
    if (st) *st = cs_stats{};
    if (C->nCloths == 0) return CS_E_STATE;
    return guard(C, [&]() { return step(C, st); });
}

cs_status cs_run(cs_ctx* C, int steps, cs_stats* acc) {
    if (!C || steps < 0) return CS_E_ARG;
    if (C->nCloths == 0) return CS_E_STATE;
    return guard(C, [&]() {
        int worst = CS_OK;
        cs_stats sum{};
        sum.pcg_converged = 1;
        for (int i = 0; i < steps; i++) {
            cs_stats one{};
            int r = step(C, &one);
            if (r != CS_OK && r != CS_E_NOT_CONVERGED) return r;
            if (r != CS_OK) worst = r;
            sum.pcg_iters += one.pcg_iters;
            sum.pcg_rel_residual = std::max(sum.pcg_rel_residual, one.pcg_rel_residual);
            sum.pcg_converged &= one.pcg_converged;
            sum.n_proximities += one.n_proximities;
            sum.n_contact_blocks += one.n_contact_blocks;
            sum.collision_iters += one.collision_iters;
            sum.collision_deform = std::max(sum.collision_deform, one.collision_deform);
            sum.n_impacts += one.n_impacts;
            sum.ms_bvh += one.ms_bvh;
            sum.ms_proximity += one.ms_proximity;
            sum.ms_assemble += one.ms_assemble;
            sum.ms_solve += one.ms_solve;
            sum.ms_update += one.ms_update;
            sum.ms_collision += one.ms_collision;
        }
        if (acc) *acc = sum;
        return worst;
    });
}

cs_status cs_get_counts(cs_ctx* C, int kind, int id, int* nn, int* nf, int* ne) {
    if (!C) return CS_E_ARG;
    const MeshInfo* M = mesh_of(C, kind, id);
    if (!M) return CS_E_ARG;
    if (nn) *nn = M->nn;
    if (nf) *nf = M->nf;
    if (ne) *ne = M->ne;
    return CS_OK;
}

cs_status cs_get_state(cs_ctx* C, int kind, int id, float* x, float* x0, float* v) {
    if (!C) return CS_E_ARG;
    const MeshInfo* M = mesh_of(C, kind, id);
    if (!M) return CS_E_ARG;
    return guard(C, [&]() {
        host_finalize(C);
        std::vector<float4> h(M->nn);
        float* outs[3] = {x, x0, v};
        DBuf* bufs[3] = {&C->x, &C->x0, &C->v};
        for (int k = 0; k < 3; k++) {
            if (!outs[k]) continue;
            CS_HIP(hipMemcpyAsync(h.data(), bufs[k]->ptr<float4>() + M->n0, sizeof(float4) * M->nn,
                                  hipMemcpyDeviceToHost, C->stream));
            CS_HIP(hipStreamSynchronize(C->stream));
            for (int i = 0; i < M->nn; i++) {
                outs[k][3 * i] = h[i].x;
                outs[k][3 * i + 1] = h[i].y;
                outs[k][3 * i + 2] = h[i].z;
            }
        }
        return (int)CS_OK;
    });
}

cs_status cs_frame_begin(cs_ctx* C, int slot) {
    if (!C || slot < 0 || slot > 1) return CS_E_ARG;
    return guard(C, [&]() {
        host_finalize(C);
        auto& F = C->frame[slot];
        if (!C->copy) {
            CS_HIP(hipStreamCreateWithFlags(&C->copy, hipStreamNonBlocking));
            CS_HIP(hipEventCreateWithFlags(&C->evSnap, hipEventDisableTiming));
        }
        if (!F.done) CS_HIP(hipEventCreateWithFlags(&F.done, hipEventDisableTiming));
        int total = 0;
        std::vector<int> n0, nn;
        for (int c = 0; c < C->nCloths; c++) {
            const MeshInfo* M = mesh_of(C, CS_MESH_CLOTH, c);
            n0.push_back(M->n0);
            nn.push_back(M->nn);
            total = std::max(total, M->n0 + M->nn);
        }
        if (F.begun) {
            // the slot's previous download still reads its snapshot and writes its host copy
            if (F.cap < (size_t)total) CS_HIP(hipEventSynchronize(F.done));
            else CS_HIP(hipStreamWaitEvent(C->stream, F.done, 0));
        }
        if (F.cap < (size_t)total) {  // both buffers grow together (device buffers never shrink)
            if (F.host) CS_HIP(hipHostFree(F.host));
            F.host = nullptr;
            F.cap = 0;
            CS_HIP(hipHostMalloc(reinterpret_cast<void**>(&F.host), sizeof(float4) * 2 * (size_t)total,
                                 hipHostMallocDefault));
            F.cap = total;
        }
        F.begun = false;
        if (total > 0) {
            float4* st = F.stage.as<float4>(2 * (size_t)total);
            CS_HIP(hipMemcpyAsync(st, C->x.p, sizeof(float4) * total, hipMemcpyDeviceToDevice, C->stream));
            CS_HIP(hipMemcpyAsync(st + total, C->v.p, sizeof(float4) * total, hipMemcpyDeviceToDevice, C->stream));
            CS_HIP(hipEventRecord(C->evSnap, C->stream));
            CS_HIP(hipStreamWaitEvent(C->copy, C->evSnap, 0));
            CS_HIP(hipMemcpyAsync(F.host, st, sizeof(float4) * 2 * (size_t)total, hipMemcpyDeviceToHost, C->copy));
        }
        CS_HIP(hipEventRecord(F.done, C->copy));
        F.total = total;
        F.n0 = n0;
        F.nn = nn;
        F.begun = true;
        return (int)CS_OK;
    });
}

cs_status cs_frame_end(cs_ctx* C, int slot, int cloth, float* x, float* v, int* n_nodes) {
    if (!C || slot < 0 || slot > 1) return CS_E_ARG;
    auto& F = C->frame[slot];
    if (!F.begun) return CS_E_STATE;
    if (cloth < 0 || cloth >= (int)F.nn.size()) return CS_E_ARG;
    // no guard(): this call may run on another host thread while the context steps, and touches
    // only the slot's event and pinned copy
    if (hipEventSynchronize(F.done) != hipSuccess) return CS_E_HIP;
    const int n = F.nn[cloth];
    const float4* hx = F.host + F.n0[cloth];
    const float4* hv = F.host + F.total + F.n0[cloth];
    for (int i = 0; i < n && x; i++) {
        x[3 * i] = hx[i].x;
        x[3 * i + 1] = hx[i].y;
        x[3 * i + 2] = hx[i].z;
    }
    for (int i = 0; i < n && v; i++) {
        v[3 * i] = hv[i].x;
        v[3 * i + 1] = hv[i].y;
        v[3 * i + 2] = hv[i].z;
    }
    if (n_nodes) *n_nodes = n;
    return CS_OK;
}

cs_status cs_set_state(cs_ctx* C, int kind, int id, const float* x, const float* x0, const float* v) {
    if (!C) return CS_E_ARG;
    const MeshInfo* M = mesh_of(C, kind, id);
    if (!M) return CS_E_ARG;
    return guard(C, [&]() {
        host_finalize(C);
        std::vector<float4> h(M->nn);
        const float* ins[3] = {x, x0, v};
        DBuf* bufs[3] = {&C->x, &C->x0, &C->v};
        for (int k = 0; k < 3; k++) {
            if (!ins[k]) continue;
            for (int i = 0; i < M->nn; i++) h[i] = make_float4(ins[k][3 * i], ins[k][3 * i + 1], ins[k][3 * i + 2], 0.0f);
            CS_HIP(hipMemcpyAsync(bufs[k]->ptr<float4>() + M->n0, h.data(), sizeof(float4) * M->nn,
                                  hipMemcpyHostToDevice, C->stream));
            CS_HIP(hipStreamSynchronize(C->stream));
        }
        launch_face_geometry(C, M->f0, M->nf);
        launch_node_geometry(C, M->n0, M->nn);
        CS_HIP(hipStreamSynchronize(C->stream));
        return (int)CS_OK;
    });
}

cs_status cs_get_geometry(cs_ctx* C, int kind, int id, float* normals, float* area, float* mass) {
    if (!C) return CS_E_ARG;
    const MeshInfo* M = mesh_of(C, kind, id);
    if (!M) return CS_E_ARG;
    return guard(C, [&]() {
        host_finalize(C);
        hipStream_t s = C->stream;
        if (normals) {
            std::vector<float4> h(M->nn);
            CS_HIP(hipMemcpyAsync(h.data(), C->nrm.ptr<float4>() + M->n0, sizeof(float4) * M->nn, hipMemcpyDeviceToHost, s));
            CS_HIP(hipStreamSynchronize(s));
            for (int i = 0; i < M->nn; i++) { normals[3 * i] = h[i].x; normals[3 * i + 1] = h[i].y; normals[3 * i + 2] = h[i].z; }
        }
        if (area) CS_HIP(hipMemcpyAsync(area, C->area.ptr<float>() + M->n0, sizeof(float) * M->nn, hipMemcpyDeviceToHost, s));
        if (mass) CS_HIP(hipMemcpyAsync(mass, C->mass.ptr<float>() + M->n0, sizeof(float) * M->nn, hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        return (int)CS_OK;
    });
}

cs_status cs_material_table(const cs_material_in* material, float* table, float* density, float* bending) {
    if (!material) return CS_E_ARG;
    try {
        HostMaterial m;
        material_build(*material, m);
        if (table) std::memcpy(table, m.table.data(), sizeof(float) * m.table.size());
        if (density) *density = m.density;
        if (bending) std::memcpy(bending, m.bend, sizeof(m.bend));
    } catch (...) {
        return CS_E_OOM;
    }
    return CS_OK;
}

cs_status cs_lbvh_pairs(cs_ctx* C, int mode, float thickness, int* pairs, int cap, int* n_out) {
    if (!C || !n_out || cap < 0 || (mode & ~(CS_LBVH_CCD | CS_LBVH_REFIT))) return CS_E_ARG;
    if (C->nCloths == 0) return CS_E_STATE;
    const bool ccd = (mode & CS_LBVH_CCD) != 0;
    if (mode & CS_LBVH_REFIT)
        for (size_t m = 0; m < C->meshes.size(); m++)
            if (!C->finalized || C->bvh.size() != C->meshes.size() || C->bvh[m].mesh != (int)m ||
                C->bvh[m].nf != C->meshes[m].nf || C->bvh[m].ccd != ccd) {
                C->err = "cs_lbvh_pairs: CS_LBVH_REFIT needs trees built by a previous call in the same mode";
                return CS_E_STATE;
            }
    return guard(C, [&]() {
        host_finalize(C);
        if (mode & CS_LBVH_REFIT)
            bvh_refit_range(C, 0, (int)C->meshes.size());
        else
            bvh_build_all(C, ccd);
        long n = bvh_pairs_ordered(C, thickness, pairs, pairs ? cap : 0);
        *n_out = (int)n;
        return (pairs && n > cap) ? (int)CS_E_CAPACITY : (int)CS_OK;
    });
}

cs_status cs_find_proximities(cs_ctx* C, cs_proximity_rec* out, int cap, int* n_out) {
    if (!C || !n_out || cap < 0) return CS_E_ARG;
    if (C->nCloths == 0) return CS_E_STATE;
    return guard(C, [&]() {
        host_finalize(C);
        int n = find_proximities(C);
        *n_out = n;
        if (!out) return (int)CS_OK;
        if (n > cap) return (int)CS_E_CAPACITY;
        if (n > 0) {
            CS_HIP(hipMemcpyAsync(out, C->prox.ptr<cs_proximity_rec>(), sizeof(cs_proximity_rec) * n,
                                  hipMemcpyDeviceToHost, C->stream));
            CS_HIP(hipStreamSynchronize(C->stream));
        }
        return (int)CS_OK;
    });
}

cs_status cs_assemble(cs_ctx* C, int cloth_id, int* row_ptr, int* col_idx, float* vals, int cap_blocks, int* nnzb_out,
                      float* b) {
    if (!C || !nnzb_out) return CS_E_ARG;
    if (cloth_id < 0 || cloth_id >= C->nCloths) return CS_E_ARG;
    return guard(C, [&]() {
        host_finalize(C);
        hipStream_t s = C->stream;
        if (!(C->P.flags & CS_FLAG_NO_PROXIMITY)) find_proximities(C);
        else C->nProx = 0;
        launch_element_records(C, cloth_id);
        launch_assemble_rows(C, cloth_id);
        assemble_contacts(C, cloth_id);
        launch_block_jacobi(C, cloth_id);
        ClothSys& S = C->sys[cloth_id];
        const int nn = S.nn;
        std::vector<float> sv((size_t)S.totalSlots * SLICE * 9), hb(3 * (size_t)nn), hbc(3 * (size_t)nn);
        std::vector<int> crp(nn + 1, 0), ccol;
        std::vector<float> cval;
        CS_HIP(hipMemcpyAsync(sv.data(), S.vals.ptr<float>(), sizeof(float) * sv.size(), hipMemcpyDeviceToHost, s));
        CS_HIP(hipMemcpyAsync(hb.data(), S.b.ptr<float>(), sizeof(float) * hb.size(), hipMemcpyDeviceToHost, s));
        CS_HIP(hipMemcpyAsync(hbc.data(), S.bc.ptr<float>(), sizeof(float) * hbc.size(), hipMemcpyDeviceToHost, s));
        CS_HIP(hipMemcpyAsync(crp.data(), S.cRowPtr.ptr<int>(), sizeof(int) * (nn + 1), hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        if (S.nContact > 0) {
            ccol.resize(S.nContact);
            cval.resize(9 * (size_t)S.nContact);
            CS_HIP(hipMemcpyAsync(ccol.data(), S.cCol.ptr<int>(), sizeof(int) * ccol.size(), hipMemcpyDeviceToHost, s));
            CS_HIP(hipMemcpyAsync(cval.data(), S.cVal.ptr<float>(), sizeof(float) * cval.size(), hipMemcpyDeviceToHost, s));
            CS_HIP(hipStreamSynchronize(s));
        }
        // merge the static and contact parts of each row (columns ascending)
        std::vector<int> rp(nn + 1, 0), cols;
        std::vector<float> vv;
        for (int r = 0; r < nn; r++) {
            const int sl = r / SLICE, lane = r % SLICE;
            std::vector<std::pair<int, std::array<float, 9>>> row;
            for (int k = 0; k < S.hRowLen[r]; k++) {
                std::array<float, 9> a;
                for (int q = 0; q < 9; q++) a[q] = sv[(((size_t)S.hSliceOff[sl] + k) * 9 + q) * SLICE + lane];
                row.push_back({S.hCol[S.hRowPtr[r] + k], a});
            }
            for (int k = crp[r]; k < crp[r + 1]; k++) {
                std::array<float, 9> a;
                for (int q = 0; q < 9; q++) a[q] = cval[9 * (size_t)k + q];
                auto it = std::find_if(row.begin(), row.end(), [&](const std::pair<int, std::array<float, 9>>& e) { return e.first == ccol[k]; });
                if (it != row.end()) for (int q = 0; q < 9; q++) it->second[q] += a[q];
                else row.push_back({ccol[k], a});
            }
            std::sort(row.begin(), row.end(), [](const std::pair<int, std::array<float, 9>>& p, const std::pair<int, std::array<float, 9>>& q) { return p.first < q.first; });
            for (auto& e : row) { cols.push_back(e.first); vv.insert(vv.end(), e.second.begin(), e.second.end()); }
            rp[r + 1] = (int)cols.size();
        }
        *nnzb_out = (int)cols.size();
        if (!row_ptr) return (int)CS_OK;
        if ((int)cols.size() > cap_blocks) return (int)CS_E_CAPACITY;
        std::copy(rp.begin(), rp.end(), row_ptr);
        std::copy(cols.begin(), cols.end(), col_idx);
        std::copy(vv.begin(), vv.end(), vals);
        if (b)
            for (size_t i = 0; i < hb.size(); i++) b[i] = hb[i] + hbc[i];
        return (int)CS_OK;
    });
}

cs_status cs_solve(cs_ctx* C, int cloth_id, float* dv, cs_stats* st) {
    if (!C || !dv) return CS_E_ARG;
    if (cloth_id < 0 || cloth_id >= C->nCloths) return CS_E_ARG;
    if (!C->finalized) return CS_E_STATE;
    return guard(C, [&]() {
        float rel = 0.0f;
        int conv = 0;
        int it = pcg_solve(C, cloth_id, &rel, &conv);
        ClothSys& S = C->sys[cloth_id];
        CS_HIP(hipMemcpyAsync(dv, S.x.ptr<float>(), sizeof(float) * 3 * (size_t)S.nn, hipMemcpyDeviceToHost, C->stream));
        CS_HIP(hipStreamSynchronize(C->stream));
        if (st) {
            st->pcg_iters = it;
            st->pcg_rel_residual = rel;
            st->pcg_converged = conv;
        }
        return conv ? (int)CS_OK : (int)CS_E_NOT_CONVERGED;
    });
}

cs_status cs_precondition(cs_ctx* C, int cloth_id, const float* r, float* z, int* level_sizes, int cap,
                          int* n_levels) {
    if (!C || !n_levels || cap < 0 || (cap > 0 && !level_sizes) || (r && !z)) return CS_E_ARG;
    if (cloth_id < 0 || cloth_id >= C->nCloths) return CS_E_ARG;
    if (C->P.precond == CS_PRECOND_BLOCK_JACOBI) return CS_E_ARG;
    if (r && !C->finalized) return CS_E_STATE;  // z needs an assembled system (cs_assemble)
    if (C->P.precond == CS_PRECOND_LEARNED) {
        *n_levels = 0;
        if (!r) return CS_OK;
        return guard(C, [&]() {
            ClothSys& S = C->sys[cloth_id];
            hipStream_t s = C->stream;
            const size_t n3 = 3 * (size_t)S.nn;
            CS_HIP(hipMemcpyAsync(S.r.ptr<float>(), r, sizeof(float) * n3, hipMemcpyHostToDevice, s));
            lpc_prepare(C, cloth_id);
            const int nb = std::max(1, cdiv(S.nn, 256));
            lpc_apply(C, cloth_id, S.r.ptr<float>(), S.z.ptr<float4>(), C->ws.a.as<float>(nb), nb);
            std::vector<float> z4(4 * (size_t)S.nn);
            CS_HIP(hipMemcpyAsync(z4.data(), S.z.ptr<float4>(), sizeof(float) * z4.size(), hipMemcpyDeviceToHost, s));
            CS_HIP(hipStreamSynchronize(s));
            for (int i = 0; i < S.nn; i++)
                for (int k = 0; k < 3; k++) z[3 * i + k] = z4[4 * (size_t)i + k];
            return (int)CS_OK;
        });
    }
    return guard(C, [&]() {
        host_finalize(C);
        std::vector<int> sizes(64);
        const int nl = amg_levels(C, cloth_id, sizes.data(), 64);
        *n_levels = nl;
        for (int l = 0; l < nl && l < cap; l++) level_sizes[l] = sizes[l];
        if (!r) return (int)CS_OK;
        if (nl <= 1) return (int)CS_E_STATE;
        ClothSys& S = C->sys[cloth_id];
        hipStream_t s = C->stream;
        const size_t n3 = 3 * (size_t)S.nn;
        CS_HIP(hipMemcpyAsync(S.r.ptr<float>(), r, sizeof(float) * n3, hipMemcpyHostToDevice, s));
        amg_prepare(C, cloth_id);
        const int nb = std::max(1, cdiv(S.nn, 256));
        if (C->P.precond == CS_PRECOND_AMG_LEARNED) {
            lpc_prepare(C, cloth_id);
            lpc_smoother_prepare(C, cloth_id);
            lpc_smooth_pre(C, cloth_id, S.r.ptr<float>(), amg_level0_x(C, cloth_id));
            amg_apply(C, cloth_id, S.r.ptr<float>(), S.z.ptr<float4>(), C->ws.a.as<float>(nb), nb, true, true);
        } else {
            amg_apply(C, cloth_id, S.r.ptr<float>(), S.z.ptr<float4>(), C->ws.a.as<float>(nb), nb, false);
        }
        std::vector<float> z4(4 * (size_t)S.nn);
        CS_HIP(hipMemcpyAsync(z4.data(), S.z.ptr<float4>(), sizeof(float) * z4.size(), hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < S.nn; i++)
            for (int k = 0; k < 3; k++) z[3 * i + k] = z4[4 * (size_t)i + k];
        return (int)CS_OK;
    });
}

int cs_lpc_weight_count(void) { return lpc_weight_count(); }

cs_status cs_set_learned_weights(cs_ctx* C, const float* w, int n) {
    if (!C || !w || n != lpc_weight_count()) return CS_E_ARG;
    for (int i = 0; i < n; i++)
        if (!std::isfinite(w[i])) return CS_E_ARG;
    return guard(C, [&]() {
        lpc_set_weights(C, w);  // the factor is rebuilt from these at the next solve
        return (int)CS_OK;
    });
}

cs_status cs_get_learned_weights(cs_ctx* C, float* w, int n) {
    if (!C || !w || n != lpc_weight_count()) return CS_E_ARG;
    return guard(C, [&]() {
        lpc_get_weights(C, w);
        return (int)CS_OK;
    });
}

cs_status cs_lpc_factor(cs_ctx* C, int cloth_id, float* diag, float* lower, int cap_lower, int* n_lower) {
    if (!C || !n_lower || cap_lower < 0) return CS_E_ARG;
    if (cloth_id < 0 || cloth_id >= C->nCloths) return CS_E_ARG;
    if (!C->finalized) return CS_E_STATE;  // needs an assembled system
    if (!C->lpcNet.loaded) return CS_E_STATE;
    return guard(C, [&]() {
        ClothSys& S = C->sys[cloth_id];
        const int nl = ((int)S.hCol.size() - S.nn) / 2;
        *n_lower = nl;
        if (!diag || !lower) return (int)CS_OK;
        if (nl > cap_lower) return (int)CS_E_CAPACITY;
        lpc_prepare(C, cloth_id);
        lpc_factor_blocks(C, cloth_id, diag, lower);
        return (int)CS_OK;
    });
}

cs_status cs_lpc_train_step(cs_ctx* C, int cloth_id, float lr, float sol_weight, float* loss, float* grad) {
    if (!C || !loss || !(lr >= 0.0f) || !(sol_weight >= 0.0f)) return CS_E_ARG;
    if (cloth_id < 0 || cloth_id >= C->nCloths) return CS_E_ARG;
    if (!C->finalized) return CS_E_STATE;
    if (!C->lpcNet.loaded) return CS_E_STATE;
    return guard(C, [&]() {
        *loss = lpc_train_step(C, cloth_id, lr, sol_weight, 0.0f, 0u, grad);
        return (int)CS_OK;
    });
}

cs_status cs_lpc_train_step_smooth(cs_ctx* C, int cloth_id, float lr, float sol_weight, float smooth_weight,
                                   unsigned int seed, float* loss, float* grad) {
    if (!C || !loss || !(lr >= 0.0f) || !(sol_weight >= 0.0f) || !(smooth_weight >= 0.0f)) return CS_E_ARG;
    if (cloth_id < 0 || cloth_id >= C->nCloths) return CS_E_ARG;
    if (!C->finalized) return CS_E_STATE;
    if (!C->lpcNet.loaded) return CS_E_STATE;
    return guard(C, [&]() {
        *loss = lpc_train_step(C, cloth_id, lr, sol_weight, smooth_weight, seed, grad);
        return (int)CS_OK;
    });
}

cs_status cs_set_remeshing(cs_ctx* C, int cloth_id, const cs_remeshing* r) {
    if (!C || cloth_id < 0 || cloth_id >= C->nCloths) return CS_E_ARG;
    if ((int)C->remeshCfg.size() < C->nCloths) C->remeshCfg.resize(C->nCloths);
    cs_ctx::RemeshCfg& R = C->remeshCfg[cloth_id];
    R.has = r != nullptr;
    if (r)
        R.r = RemeshParams{r->refine_angle, r->refine_velocity, r->refine_compression, r->rib_stiffening,
                           r->size_min, r->size_max, r->aspect_min, r->flip_threshold};
    return CS_OK;
}

cs_status cs_remesh(cs_ctx* C) {
    if (!C) return CS_E_ARG;
    return guard(C, [&]() { return remesh_context(C); });
}

cs_status cs_separation_step(cs_ctx* C, const int* faces_per_cloth, const float* backup_uv, const float* backup_x) {
    if (!C || !faces_per_cloth || !backup_uv || !backup_x) return CS_E_ARG;
    for (int c = 0; c < C->nCloths; c++)
        if (faces_per_cloth[c] < 0) return CS_E_ARG;
    return guard(C, [&]() {
        host_finalize(C);
        std::vector<std::vector<BackupFace>> backup(C->nCloths);
        size_t k = 0;
        for (int c = 0; c < C->nCloths; c++) {
            backup[c].resize(faces_per_cloth[c]);
            for (BackupFace& f : backup[c]) {
                for (int j = 0; j < 3; j++) {
                    f.u[j][0] = backup_uv[6 * k + 2 * j];
                    f.u[j][1] = backup_uv[6 * k + 2 * j + 1];
                    for (int a = 0; a < 3; a++) f.x[j][a] = backup_x[9 * k + 3 * j + a];
                }
                k++;
            }
        }
        C->sepIters = 0;
        return separation_step(C, backup);
    });
}

cs_status cs_remesh_times(const cs_ctx* C, float* sizing_ms, float* operators_ms, float* rebuild_ms,
                          float* separation_ms, int* separation_solves) {
    if (!C) return CS_E_ARG;
    if (sizing_ms) *sizing_ms = C->remeshMs[0];
    if (operators_ms) *operators_ms = C->remeshMs[1];
    if (rebuild_ms) *rebuild_ms = C->remeshMs[2];
    if (separation_ms) *separation_ms = C->remeshMs[3];
    if (separation_solves) *separation_solves = C->sepIters;
    return CS_OK;
}

cs_status cs_remesh_sizing(cs_ctx* C, int cloth_id, const cs_remeshing* r, float* face_sizing, float* planes) {
    if (!C || !r || !face_sizing || cloth_id < 0 || cloth_id >= C->nCloths) return CS_E_ARG;
    return guard(C, [&]() {
        if (!C->finalized) host_finalize(C);
        const RemeshParams p{r->refine_angle, r->refine_velocity, r->refine_compression, r->rib_stiffening,
                             r->size_min, r->size_max, r->aspect_min, r->flip_threshold};
        std::vector<float> fs, pl;
        remesh_sizing(C, cloth_id, p, fs, planes ? &pl : nullptr);
        std::copy(fs.begin(), fs.end(), face_sizing);
        if (planes) std::copy(pl.begin(), pl.end(), planes);
        return (int)CS_OK;
    });
}

cs_status cs_batch_tuning(const cs_ctx* C, int* rows_per_block, int* chunk, int* trials) {
    if (!C || !rows_per_block || !chunk || !trials) return CS_E_ARG;
    *rows_per_block = C->batchTune.rows;
    *chunk = C->batchTune.chunk;
    *trials = C->batchTune.trials;
    return CS_OK;
}

cs_status cs_lpc_smoother_omega(cs_ctx* C, int cloth_id, float* omega) {
    if (!C || !omega) return CS_E_ARG;
    if (cloth_id < 0 || cloth_id >= C->nCloths) return CS_E_ARG;
    if (!C->finalized || !C->lpcNet.loaded) return CS_E_STATE;
    return guard(C, [&]() {
        lpc_prepare(C, cloth_id);
        lpc_smoother_prepare(C, cloth_id);
        *omega = lpc_smoother_omega(C, cloth_id);
        return (int)CS_OK;
    });
}

cs_status cs_find_impacts(cs_ctx* C, float thickness, cs_impact_rec* out, int cap, int* n_out) {
    if (!C || !n_out || cap < 0) return CS_E_ARG;
    if (C->nCloths == 0) return CS_E_STATE;
    return guard(C, [&]() {
        host_finalize(C);
        bvh_build_all(C, true);
        int n = find_impacts(C, thickness);
        *n_out = n;
        if (!out) return (int)CS_OK;
        if (n > cap) return (int)CS_E_CAPACITY;
        if (n > 0) {
            CS_HIP(hipMemcpyAsync(out, C->impSorted.ptr<cs_impact_rec>(), sizeof(cs_impact_rec) * n,
                                  hipMemcpyDeviceToHost, C->stream));
            CS_HIP(hipStreamSynchronize(C->stream));
        }
        return (int)CS_OK;
    });
}

cs_status cs_independent_impacts(cs_ctx* C, const cs_impact_rec* in, int n, int deform, cs_impact_rec* out, int* n_out) {
    if (!C || n < 0 || !n_out || (n > 0 && (!in || !out)) || (deform != 0 && deform != 1)) return CS_E_ARG;
    return guard(C, [&]() {
        host_finalize(C);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 4; k++)
                if (in[i].nodes[k] < 0 || in[i].nodes[k] >= C->N) return (int)CS_E_ARG;
        *n_out = 0;
        if (n == 0) return (int)CS_OK;
        hipStream_t s = C->stream;
        cs_impact_rec* din = C->impSorted.as<cs_impact_rec>(n);
        cs_impact_rec* dout = C->impacts.as<cs_impact_rec>(n);
        CS_HIP(hipMemcpyAsync(din, in, sizeof(cs_impact_rec) * n, hipMemcpyHostToDevice, s));
        int m = independent_impacts(C, din, n, deform, dout);
        CS_HIP(hipMemcpyAsync(out, dout, sizeof(cs_impact_rec) * m, hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        *n_out = m;
        return (int)CS_OK;
    });
}

cs_status cs_impact_response(cs_ctx* C, const cs_impact_rec* in, int n, int deform, float obstacle_mass, int* iters) {
    if (!C || n < 0 || (n > 0 && !in) || (deform != 0 && deform != 1) || !(obstacle_mass > 0.0f)) return CS_E_ARG;
    return guard(C, [&]() {
        host_finalize(C);
        for (int i = 0; i < n; i++)
            for (int k = 0; k < 4; k++)
                if (in[i].nodes[k] < 0 || in[i].nodes[k] >= C->N) return (int)CS_E_ARG;
        if (iters) *iters = 0;
        if (n == 0) return (int)CS_OK;
        hipStream_t s = C->stream;
        cs_impact_rec* d = C->impAcc.as<cs_impact_rec>(n);
        CS_HIP(hipMemcpyAsync(d, in, sizeof(cs_impact_rec) * n, hipMemcpyHostToDevice, s));
        fill_u32(C->optScal.as<uint32_t>(4), 0u, 4, s);
        collision_response(C, n, deform, obstacle_mass);
        int it = 0;
        CS_HIP(hipMemcpyAsync(C->hCount, C->optScal.ptr<uint32_t>(), sizeof(unsigned), hipMemcpyDeviceToHost, s));
        CS_HIP(hipStreamSynchronize(s));
        it = (int)C->hCount[0];
        if (iters) *iters = it;
        return (int)CS_OK;
    });
}

cs_status cs_kernel_time(cs_ctx* C, int kernel_id, double* total_ms, int* count, int reset) {
    if (!C || kernel_id < 0 || kernel_id >= CS_KERNEL_COUNT) return CS_E_ARG;
    return guard(C, [&]() {
        ktimer_collect(C);
        if (total_ms) *total_ms = C->kt.total[kernel_id];
        if (count) *count = C->kt.count[kernel_id];
        if (reset) {
            C->kt.total[kernel_id] = 0.0;
            C->kt.count[kernel_id] = 0;
        }
        return (int)CS_OK;
    });
}

}  // extern "C"
