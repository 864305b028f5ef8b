// cs_kernels.hpp — host launchers of the gfx950 kernels, grouped by stage.
#pragma once
#include "cs_ctx.hpp"

namespace cs {

// sampled kernel timing (cs_api.cpp): begin returns a token (-1 when disabled)
int ktimer_begin(cs_ctx* C, int kernel);
void ktimer_end(cs_ctx* C, int token);
void ktimer_collect(cs_ctx* C);  // call after a stream synchronisation

// host set-up (cs_host.cpp)
void material_build(const cs_material_in& in, HostMaterial& m);
int host_add_mesh(cs_ctx* C, const cs_mesh_in& in, bool cloth, int material);
void host_finalize(cs_ctx* C);

// geometry (cs_geometry.hip)
void launch_face_geometry(cs_ctx* C, int f0, int nf);
// remeshing (cs_sizing.hip, cs_remesh_ctx.cpp)
void remesh_sizing(cs_ctx* C, int c, const RemeshParams& r, std::vector<float>& faceSizing,
                   std::vector<float>* planes);
int remesh_context(cs_ctx* C);  // returns a cs_status (the separation step can fail)
// separation after remeshing (cs_separation.hip); backup: per cloth, its faces before remeshing
int separation_step(cs_ctx* C, const std::vector<std::vector<BackupFace>>& backup);
void launch_node_geometry(cs_ctx* C, int n0, int nn);
void launch_update_positions(cs_ctx* C, int n0, int nn);
void launch_update_velocities(cs_ctx* C, int n0, int nn);
void launch_apply_dv(cs_ctx* C, int c);

// assembly (cs_assemble.hip)
void launch_element_records(cs_ctx* C, int c);
void launch_assemble_rows(cs_ctx* C, int c);
int assemble_contacts(cs_ctx* C, int c);  // returns number of contact blocks
void launch_block_jacobi(cs_ctx* C, int c);
int assemble_all(cs_ctx* C);  // batched instances: every cloth's system, contacts and block Jacobi
void* pinned_batch(cs_ctx* C, size_t bytes);  // the context's pinned readback buffer, grown to `bytes`

// PCG (cs_pcg.hip)
int pcg_solve(cs_ctx* C, int c, float* relres, int* converged);
// PCG of every cloth at once (independent instances); per-cloth results.  amg: each cloth whose
// hierarchy is usable is preconditioned by its multigrid cycle, the others by block Jacobi.
void pcg_solve_batched(cs_ctx* C, int* iters, float* relres, int* converged, bool amg = false);
// aggregation multigrid preconditioner (cs_amg.hip)
constexpr float AMG_OMEGA = 0.6f;  // damping of the block-Jacobi smoother
// batched multigrid of the independent cloths: bamg_build returns, per cloth, 1 if its
// hierarchy is used; bamg_prepare is the per-solve setup of all of them; bamg_cycle runs
// levels >= 1 of one cycle (level 0's residual in, level 1's result out: amg.L[0].t, L[1].t)
const std::vector<int>& bamg_build(cs_ctx* C);
void bamg_prepare(cs_ctx* C);
void bamg_cycle(cs_ctx* C);
// restricts bamg_cycle to the cloths with live[c] != 0 (null: all); call after a stream sync
void bamg_select(cs_ctx* C, const std::vector<char>* live);
int amg_levels(cs_ctx* C, int c, int* sizes, int cap);  // 0: not usable for this cloth
// fork: the coarsest inverse and the small-level operator run on the side stream; call
// amg_join before the first cycle
void amg_prepare(cs_ctx* C, int c, bool fork = false);
void amg_join(cs_ctx* C);
// z = M^-1 r.  presmoothed: the caller already wrote x = w Minv r into amg_level0_x.
// learned: the level-0 smoother is the learned factor (lpc_smooth_*; the caller pre-smooths)
void amg_apply(cs_ctx* C, int c, const float* r, float4* z, float* partRZ, int nb, bool presmoothed,
               bool learned = false);
float* amg_level0_x(cs_ctx* C, int c);
// learned preconditioner (cs_lpc.hip)
int lpc_weight_count();
void lpc_set_weights(cs_ctx* C, const float* w);  // host array of lpc_weight_count() floats
void lpc_get_weights(cs_ctx* C, float* w);
void lpc_refresh_fragments(cs_ctx* C);            // after the float weights changed on the device
void lpc_prepare(cs_ctx* C, int c);               // network on cloth c's system -> G
// z = G^T G r and one r.z partial per block (nb blocks)
void lpc_apply(cs_ctx* C, int c, const float* r, float4* z, float* partRZ, int nb);
// The learned factor as the level-0 smoother of the multigrid cycle (after lpc_prepare):
// damping from a power iteration, x = w G^T G r, and z = base + w G^T G res with r.z partials.
void lpc_smoother_prepare(cs_ctx* C, int c);
void lpc_smooth_pre(cs_ctx* C, int c, const float* r, float* x);
void lpc_smooth_post(cs_ctx* C, int c, const float* res, const float* base, const float* r, float4* z,
                     float* partRZ, int nb);
float lpc_smoother_omega(cs_ctx* C, int c);
// G's diagonal blocks (9 nn) and lower blocks (9 per lower entry, static-pattern row order)
void lpc_factor_blocks(cs_ctx* C, int c, float* diag, float* lower);
// one training step on cloth c's system: loss (Kaporin, plus solWeight times the solution term
// when the system has been solved), gradient (optional host copy), Adam when lr > 0
float lpc_train_step(cs_ctx* C, int c, float lr, float solWeight, float smoothWeight, uint32_t seed,
                     float* grad);

// motions (cs_motion.hip)
void motion_step(cs_ctx* C);                // step counter + 1, obstacle velocities and handle anchors
Xf motion_now(const cs_ctx* C, int motion);  // a motion at the context's current time
void motion_place(cs_ctx* C);               // moving obstacles to x = x0 = T(now)(base)

// LBVH (cs_lbvh.hip)
void bvh_build(cs_ctx* C, int mesh, bool ccd);
void bvh_build_all(cs_ctx* C, bool ccd);  // every mesh; small meshes in one batched launch
void bvh_refit(cs_ctx* C, int mesh);
void bvh_refit_range(cs_ctx* C, int m0, int m1);  // meshes [m0, m1); small meshes in one batched launch
// Candidate face pairs of all traversals (cloth self, cloth x earlier cloth,
// cloth x obstacle) appended to C->pairs; returns the count.
long bvh_all_pairs(cs_ctx* C, float thickness);
// Deterministic pair list (count / scan / emit), for cs_lbvh_pairs.
long bvh_pairs_ordered(cs_ctx* C, float thickness, int* host_out, long cap);

// proximities (cs_proximity.hip)
int find_proximities(cs_ctx* C, bool build_trees = true);  // build_trees false: the stage's trees are built

// CCD and impact response (cs_ccd.hip, cs_response.hip)
int find_impacts(cs_ctx* C, float thickness);  // sorted, unique impacts in C->impSorted (cs_impact_rec)
int independent_impacts(cs_ctx* C, const cs_impact_rec* dIn, int n, int deform, cs_impact_rec* dOut);
void collision_response(cs_ctx* C, int nImpacts, int deform, float obstacleMass);
void separation_response(cs_ctx* C, const SepRec* recs, int n, int deform, float obstacleArea);

}  // namespace cs
