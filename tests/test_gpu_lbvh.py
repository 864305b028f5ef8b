"""GPU parity of the collision pipeline on meshes above 4,096 faces, and between separate cloths.

A stage builds (and refits) a tree with the per-mesh kernels (Morton codes, multi-pass radix
sort, Karras hierarchy, range tables, internal boxes) when the mesh has more than 32,768 faces,
or more than 4,096 faces while the stage holds at most four such meshes; every smaller mesh is
built by the one-workgroup batched kernel.  The drape, fold and crumple meshes here (49^2 to
129^2 nodes: 4,608 to 32,768 faces, one such mesh per scene) therefore take the per-mesh path,
the one the 129^2 and 513^2 workloads use; the five-cloth scene takes the batched path with the
same mesh, and its pairs must be identical.

Bars: candidate pairs, proximity selection and impact lists bit-exact against the FP32 oracle
(and pairs against brute force over every face pair); post-step positions within 1e-5 m of the
oracle's direct solve.  Reference: src/BVH.cu:18-50,167-177, src/BVHHelper.cu:151-307,
src/Simulator.cu:179-196,300-349,383-461,602-684.
"""
import numpy as np
import pytest

import oracle
from clothsimulator_amd import CS_MESH_CLOTH, CS_MESH_OBSTACLE, Simulator, default_params, material_struct
from clothsimulator_amd._cabi import CS_E_NOT_CONVERGED, CS_OK, MeshArrays
from clothsimulator_amd.scenes import ClothSpec, SceneSpec, crumple, drape, fold, grid_mesh

pytestmark = pytest.mark.gpu


def _key(p):
    return np.sort(p[:, 0].astype(np.int64) * (1 << 32) + p[:, 1])


def _pair(spec):
    return Simulator(spec), oracle.OracleScene(spec)


def _set_both(sim, orc, kind, idx, x, x0, v):
    sim.set_state(kind, idx, x, x0, v)
    orc.set_state(kind, idx, x, x0, v)


def _lowered_drape(n):
    spec = drape(n)
    sim, orc = _pair(spec)
    x, x0, v = orc.state()
    x = x.copy()
    x[:, 2] -= 0.098  # onto the sphere top (z = 0.3)
    _set_both(sim, orc, CS_MESH_CLOTH, 0, x, x, v)
    return spec, sim, orc, x, v


def _check_pairs(tag, sim, orc, cases, brute=True, refit=False):
    for ccd, th in cases:
        g = sim.lbvh_pairs(ccd, th, refit=refit)
        o = orc.lbvh_pairs(ccd, th)
        print("%s pairs ccd=%d th=%g refit=%d: gpu %d oracle %d" % (tag, ccd, th, refit, len(g), len(o)), flush=True)
        assert len(o) > 0
        if brute:
            assert np.array_equal(_key(o), _key(orc.lbvh_pairs(ccd, th, brute=True)))
        assert np.array_equal(_key(g), _key(o))


def test_drape65_per_mesh_pairs_bitexact():
    spec, sim, orc, x, v = _lowered_drape(65)
    assert sim.counts(CS_MESH_CLOTH, 0)[1] == 8192
    _check_pairs("drape65", sim, orc, ((False, 2e-3), (True, 1e-4), (False, 5e-2)))


def test_drape65_per_mesh_refit_bitexact():
    """Trees built on one state and refit to another give the pairs a fresh build gives.  A refit
    keeps the build's leaf order, and a self pair is emitted with the face of lower sorted index
    as the query (src/BVHHelper.cu:167-168), so self pairs are compared unordered and their
    orientation against the sorted order of the build state (the oracle's Morton keys there)."""
    spec, sim, orc, x, v = _lowered_drape(65)
    F = 8192
    rng = np.random.default_rng(5)
    for ccd, th in ((False, 2e-3), (True, 1e-4)):
        sim.lbvh_pairs(ccd, th)  # build on the lowered state
        rank = np.empty(F, np.int64)
        rank[(orc.lbvh_keys(0, ccd) & np.uint64(0xFFFFFFFF)).astype(np.int64)] = np.arange(F)
        x1 = (x + rng.normal(0.0, 4e-3, x.shape)).astype(np.float32)
        x1[:, 2] -= 0.01
        _set_both(sim, orc, CS_MESH_CLOTH, 0, x1, x, v)
        g = sim.lbvh_pairs(ccd, th, refit=True)
        o = orc.lbvh_pairs(ccd, th)
        assert np.array_equal(_key(o), _key(orc.lbvh_pairs(ccd, th, brute=True)))
        print("drape65 refit ccd=%d: gpu %d oracle %d" % (ccd, len(g), len(o)), flush=True)
        selfg = (g < F).all(1)
        assert len(g) == len(o) and selfg.sum() > 0
        und = lambda p: np.sort(np.sort(p, axis=1)[:, 0].astype(np.int64) * (1 << 32) + np.sort(p, axis=1)[:, 1])
        assert np.array_equal(und(g[selfg]), und(o[(o < F).all(1)]))
        assert np.array_equal(_key(g[~selfg]), _key(o[~(o < F).all(1)]))
        assert (rank[g[selfg, 0]] < rank[g[selfg, 1]]).all()
        _set_both(sim, orc, CS_MESH_CLOTH, 0, x, x, v)


def test_refit_without_trees_is_a_state_error():
    sim = Simulator(drape(17))
    from clothsimulator_amd import ClothError
    with pytest.raises(ClothError) as e:
        sim.lbvh_pairs(True, 1e-4, refit=True)
    assert e.value.status == 7  # CS_E_STATE
    sim.lbvh_pairs(False, 2e-3)
    with pytest.raises(ClothError):
        sim.lbvh_pairs(True, 1e-4, refit=True)  # trees built without the swept boxes


def test_per_mesh_and_batched_trees_give_identical_pairs():
    """One 65^2 cloth over its sphere (per-mesh trees) against the same cloth as cloth 0 of five
    (five meshes above 4,096 faces: every tree from the batched kernel)."""
    spec, sim, orc, x, v = _lowered_drape(65)
    x = x.astype(np.float32)
    cloth = MeshArrays(x, spec.cloths[0].mesh.faces, uv=spec.cloths[0].mesh.uv,
                       face_verts=spec.cloths[0].mesh.face_verts)
    far = [MeshArrays(x + np.array([5.0 * (k + 1), 0.0, 0.0], np.float32), cloth.faces, uv=cloth.uv,
                      face_verts=cloth.face_verts) for k in range(4)]
    spec5 = SceneSpec("five", spec.params, [ClothSpec(cloth, material_struct())] +
                      [ClothSpec(m, material_struct()) for m in far], spec.obstacles, 1)
    sim5 = Simulator(spec5)
    F = 8192
    for ccd, th in ((False, 2e-3), (True, 1e-4), (False, 5e-2)):
        a = sim.lbvh_pairs(ccd, th)
        b = sim5.lbvh_pairs(ccd, th)
        # cloth 0 and the sphere only: map the sphere's face ids (after 1 vs 5 cloths)
        keep = (b < F).all(1) | ((b[:, 0] < F) & (b[:, 1] >= 5 * F))
        b0 = b[keep].copy()
        b0[b0 >= 5 * F] -= 4 * F
        print("per-mesh %d, batched %d (cloth 0: %d)" % (len(a), len(b), len(b0)), flush=True)
        assert len(a) > 0 and np.array_equal(_key(a), _key(b0))
        # the far cloths see only themselves, and no pair crosses cloths
        rest = b[~keep]
        assert (rest[:, 0] // F == rest[:, 1] // F).all()


def test_drape129_pairs_match_oracle():
    spec, sim, orc, x, v = _lowered_drape(129)
    assert sim.counts(CS_MESH_CLOTH, 0)[1] == 32768
    _check_pairs("drape129", sim, orc, ((False, 2e-3), (True, 1e-4)), brute=False)
    _check_pairs("drape129", sim, orc, ((False, 2e-3),), brute=True)


def _prox_equal(g, o):
    assert len(o) > 0 and len(g) == len(o)
    assert np.array_equal(g["nodes"], o["nodes"])  # same (primitive, side) order
    assert np.allclose(g["n"], o["n"], atol=1e-6)
    assert np.allclose(g["w"], o["w"], atol=1e-5)
    assert np.allclose(g["stiffness"], o["stiffness"], rtol=1e-6)
    assert np.array_equal(g["mu"], o["mu"])


def _impacts_equal(sim, orc, g, o):
    assert len(o) > 0 and len(g) == len(o)
    for f in ("type", "nodes", "t", "w", "n"):
        assert np.array_equal(g[f], o[f]), f
    for deform in (0, 1):
        gi = sim.independent_impacts(g, deform)
        oi = orc.independent_impacts(o, deform)
        assert len(gi) > 0 and np.array_equal(gi["nodes"], oi["nodes"])


def _crossing(x, n, radius, depth):
    """The upper layer's nodes within `radius` of the sheet centre moved down by `depth`."""
    x1 = x.copy()
    m = n // 2
    up = np.zeros(len(x), bool)
    up[m * n:] = True
    c = np.array([0.5, 0.5])
    sel = up & (np.linalg.norm(x[:, :2] - c, axis=1) < radius)
    x1[sel, 2] -= depth
    return x1, int(sel.sum())


def test_fold65_self_contact_pairs_proximities_impacts():
    spec = fold(65)
    sim, orc = _pair(spec)
    _check_pairs("fold65", sim, orc, ((False, 2e-3), (True, 1e-4)))
    g, o = sim.proximities(), orc.proximities()
    print("fold65 proximities: gpu %d oracle %d" % (len(g), len(o)), flush=True)
    _prox_equal(g, o)
    assert len(o) > 2000  # every upper node is within the repulsion range of the lower half
    x = orc.state()[0]
    x1, moved = _crossing(x, 65, 0.2, 3e-3)
    _set_both(sim, orc, CS_MESH_CLOTH, 0, x1, x, np.zeros_like(x))
    _check_pairs("fold65 crossing", sim, orc, ((True, 1e-4),), brute=False)
    g, o = sim.impacts(1e-4), orc.impacts(1e-4)
    print("fold65 impacts: gpu %d oracle %d (%d nodes moved)" % (len(g), len(o), moved), flush=True)
    _impacts_equal(sim, orc, g, o)


def test_fold49_collision_step_matches_oracle():
    spec = fold(49)
    sim, orc = _pair(spec)
    assert sim.counts(CS_MESH_CLOTH, 0)[1] == 4608
    x = orc.state()[0]
    x1, moved = _crossing(x, 49, 0.08, 2.5e-3)
    _set_both(sim, orc, CS_MESH_CLOTH, 0, x1, x, np.zeros_like(x))
    s, st = sim.collision_step()
    r, ost = orc.collision_step()
    print("fold49 collision (%d nodes crossing): gpu status %d iters %d impacts %d; oracle status %d iters %d "
          "impacts %d" % (moved, s, st.collision_iters, st.n_impacts, r, ost.collision_iters, ost.n_impacts),
          flush=True)
    assert s == CS_OK and r == CS_OK and st.n_impacts > 0
    assert len(sim.impacts(1e-4)) == 0
    d = np.abs(sim.state()[0] - orc.state()[0]).max()
    print("fold49 collision: max |dx| %.3g" % d)
    assert d <= 1e-5


def test_fold49_stages_match_direct_solve():
    """Every upper node in contact (about 20,000 contact blocks of stiffness 1e9 x area): each stage
    from the oracle's state, the physics step against the oracle's direct solve, then the
    collision step on the oracle's post-physics state.  The PCG runs to 1e-7 here: at the default
    1e-5 two solutions that both meet the residual differ by up to 3e-4 m on this state.  Below
    1e-6 the difference stops shrinking: it is then set by the FP32 rounding of the assembled
    contact blocks (GPU and oracle sum them in different orders) times the conditioning they give
    the system, 1.6e-5 m at step 1 (41 impacts, 20,484 contact blocks), under 3e-6 m at the others.
    Bars: physics 2e-5 m, collision 1e-5 m."""
    spec = fold(49)
    spec.params.pcg_rel_tol = 1e-7
    sim = Simulator(spec)
    ref = oracle.OracleScene(spec)
    worst_p, worst_c = 0.0, 0.0
    for k in range(6):
        x, x0, v = ref.state()
        sim.set_state(CS_MESH_CLOTH, 0, x, x0, v)
        s, st = sim.physics_step()
        assert s == CS_OK and st.pcg_converged == 1
        r, ost = ref.physics_step(direct=True)
        assert r == CS_OK and st.n_proximities == ost.n_proximities
        dp = float(np.abs(sim.state()[0] - ref.state()[0]).max())
        x, x0, v = ref.state()
        sim.set_state(CS_MESH_CLOTH, 0, x, x0, v)
        s, cst = sim.collision_step()
        r, ocst = ref.collision_step()
        assert s == CS_OK and r == CS_OK and cst.n_impacts == ocst.n_impacts
        dc = float(np.abs(sim.state()[0] - ref.state()[0]).max())
        worst_p, worst_c = max(worst_p, dp), max(worst_c, dc)
        print("fold49 step %d: pcg %d prox %d contact blocks %d impacts %d: physics max|dx| %.3g, collision %.3g" % (
            k, st.pcg_iters, st.n_proximities, st.n_contact_blocks, cst.n_impacts, dp, dc), flush=True)
    assert worst_p <= 2e-5 and worst_c <= 1e-5


def test_fold49_solution_meets_its_residual_in_fp64():
    """The GPU's solution of the contact-heavy fold system at the default tolerance, checked in
    FP64 on the host against the assembled system it solved: ||b - A dv|| / ||b||."""
    import scipy.sparse as sp
    spec = fold(49)
    sim = Simulator(spec)
    ref = oracle.OracleScene(spec)
    for k in range(3):
        sim.set_state(CS_MESH_CLOTH, 0, *ref.state())
        rp, ci, vals, b = sim.assemble(0)
        dv, st = sim.solve(0)
        A = sp.bsr_matrix((vals.astype(np.float64), ci, rp), shape=(len(b), len(b)))
        res = np.linalg.norm(b.astype(np.float64) - A @ dv.astype(np.float64)) / np.linalg.norm(b)
        print("fold49 step %d: %d blocks, pcg %d, recursive %.3g, FP64 true residual %.3g" % (
            k, len(ci), st.pcg_iters, st.pcg_rel_residual, res), flush=True)
        assert st.pcg_converged == 1 and res <= 2e-5
        ref.step(direct=True)


def _crumple_contact_steps(tol):
    """The crumple generator (z perturbed by N(0, 1 mm), seed 0) dropped from 2 cm: free fall on
    the GPU alone to step 22, then 16 full steps, each from the oracle's state, against the
    oracle's direct solve; contact with the ground plane starts at step 25."""
    spec = crumple(49, height=0.02)
    sim = Simulator(spec)
    for _ in range(22):
        s, st = sim.step()
        assert s == CS_OK
    x, x0, v = sim.state()
    spec2 = crumple(49, height=0.02)
    spec2.params.pcg_rel_tol = tol
    ref = oracle.OracleScene(spec2)
    ref.set_state(CS_MESH_CLOTH, 0, x, x0, v)
    sim2 = Simulator(spec2)
    drift, prox = [], 0
    for k in range(16):
        x, x0, v = ref.state()
        sim2.set_state(CS_MESH_CLOTH, 0, x, x0, v)
        s, st = sim2.step()
        assert s in (CS_OK, CS_E_NOT_CONVERGED) and st.pcg_converged == 1
        r, ost = ref.step(direct=True)
        assert r == CS_OK and st.n_proximities == ost.n_proximities and st.n_impacts == ost.n_impacts
        drift.append(float(np.abs(sim2.state()[0] - ref.state()[0]).max()))
        prox = max(prox, st.n_proximities)
        print("crumple49 tol %g step %d: pcg %d prox %d impacts %d max|dx| %.3g" % (
            tol, 22 + k, st.pcg_iters, st.n_proximities, st.n_impacts, drift[-1]), flush=True)
    assert prox > 1000
    return np.array(drift)


def test_crumple49_ground_contact_steps_within_1e5_of_direct_solve():
    """At pcg_rel_tol 1e-6 every step is within 1e-5 m of the direct solve.  At the default 1e-5
    the contact steps drift up to about 4e-5 m (the residual bounds the velocity error only up to
    the conditioning the ground contacts add), and the drift shrinks with the tolerance: it is
    the solve's tolerance, not a different system."""
    tight = _crumple_contact_steps(1e-6)
    assert tight.max() <= 1e-5
    loose = _crumple_contact_steps(1e-5)
    assert loose.max() <= 1e-4
    assert tight.max() < 0.3 * loose.max()


def _two_cloths(n, gap):
    """Two separate cloths (two meshes): an n x n sheet and a second one above it, offset by a
    third of a cell; cloth 1 against cloth 0 is the cloth x cloth traversal."""
    size = 0.3
    h = size / (n - 1)
    xa, fa = grid_mesh(n, size)
    xb, fb = grid_mesh(n, size, origin=(h / 3, h / 3, gap))
    spec = SceneSpec("two_cloths", default_params(dt=0.002),
                     [ClothSpec(MeshArrays(xa, fa), material_struct()), ClothSpec(MeshArrays(xb, fb), material_struct())],
                     [], 1)
    return spec


@pytest.mark.parametrize("n", [9, 49])
def test_two_cloths_pairs_proximities_impacts(n):
    spec = _two_cloths(n, 1.5e-3)
    sim, orc = _pair(spec)
    F = 2 * (n - 1) ** 2
    g = sim.lbvh_pairs(False, 2e-3)
    cross = (g[:, 0] >= F) & (g[:, 1] < F)
    assert cross.sum() > 0  # cloth 1 (query) against cloth 0
    _check_pairs("two_cloths%d" % n, sim, orc, ((False, 2e-3), (True, 1e-4)), brute=True)
    _prox_equal(sim.proximities(), orc.proximities())
    # cloth 1 moves down through cloth 0
    xb, x0b, vb = orc.state(CS_MESH_CLOTH, 1)
    x1 = xb.copy()
    x1[:, 2] -= 3e-3
    _set_both(sim, orc, CS_MESH_CLOTH, 1, x1, xb, np.zeros_like(xb))
    gi, oi = sim.impacts(1e-4), orc.impacts(1e-4)
    print("two_cloths%d impacts: gpu %d oracle %d" % (n, len(gi), len(oi)), flush=True)
    _impacts_equal(sim, orc, gi, oi)
