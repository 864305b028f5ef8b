"""GPU parity of the impact response and of the host-forced failure paths.

* cs_impact_response (the augmented-Lagrangian response of
  src/optimization/CollisionOptimization.cu:37-145, src/optimization/Optimization.cu:137-181) on
  the cases whose oracle results are pinned by closed forms in test_oracle_response.py: one
  vertex below a static face, two impacts sharing the vertex, deform = 1 with light obstacle
  nodes.  Bar: every node within 1e-6 m of the FP32 oracle, the same iteration count.
* CS_E_NOT_CONVERGED: PCG capped at 3 iterations (the reference stops silently at its cap,
  src/Cloth.cu:784-809); the step still completes, as the oracle's does.
* CS_E_COLLISION_FAILED: impacts left after max_collision_iters in both deform passes (the
  reference prints "Collision step failed!" and exits, src/Simulator.cu:668-671); the library
  returns the status and a message instead.
"""
import numpy as np
import pytest

import oracle
from clothsimulator_amd import CS_MESH_CLOTH, CS_MESH_OBSTACLE, Simulator
from clothsimulator_amd._cabi import CS_E_COLLISION_FAILED, CS_E_NOT_CONVERGED, CS_OK
from clothsimulator_amd.scenes import fold, hang
from test_oracle_response import response_scene, vf_impact

pytestmark = pytest.mark.gpu


def _positions(s):
    return np.vstack([s.state(CS_MESH_CLOTH, 0)[0], s.state(CS_MESH_OBSTACLE, 0)[0]]).astype(np.float64)


@pytest.mark.parametrize("case", ["one_vertex", "two_impacts", "deform1"])
def test_impact_response_matches_oracle(case):
    spec = response_scene()
    sim, orc = Simulator(spec), oracle.OracleScene(spec)
    x1 = _positions(sim)
    assert np.array_equal(x1, _positions(orc))
    imps = vf_impact(x1, 0, (3, 4, 5), [0, 3, 4, 5])
    deform, M = 0, 1e3
    if case == "two_impacts":
        imps = np.concatenate([imps, vf_impact(x1, 0, (6, 7, 8), [0, 6, 7, 8])])
    if case == "deform1":
        deform, M = 1, 3.0 * float(orc.geometry()[2][0])
    ig = sim.impact_response(imps, deform, obstacle_mass=M)
    io = orc.impact_response(imps, deform, obstacle_mass=M)
    xg, xo = _positions(sim), _positions(orc)
    d = np.abs(xg - xo).max()
    print("%s: gpu %d iterations, oracle %d; moved %.4g m; max |gpu - oracle| %.3g m" % (
        case, ig, io, np.abs(xo - x1).max(), d))
    assert ig == io
    assert d <= 1e-6
    moved = np.abs(xo - x1).max(axis=1) > 0
    assert np.array_equal(np.abs(xg - x1).max(axis=1) > 0, moved)


def test_impact_response_rejects_bad_input():
    sim = Simulator(response_scene())
    x1 = _positions(sim)
    imp = vf_impact(x1, 0, (3, 4, 5), [0, 3, 4, 99])
    from clothsimulator_amd import ClothError
    with pytest.raises(ClothError):
        sim.impact_response(imp, 0)
    assert sim.impact_response(imp[:0], 0) == 0


def test_pcg_cap_returns_not_converged_and_completes_the_step():
    spec = hang(17)
    orc = oracle.OracleScene(spec)
    for _ in range(4):  # from rest two iterations suffice; four steps in, about 130 are needed
        orc.step(direct=True)
    sim = Simulator(spec)
    sim.set_state(CS_MESH_CLOTH, 0, *orc.state())
    p = spec.params
    p.pcg_max_iters = 3
    sim.set_params(p)
    orc.set_params(p)
    x0 = sim.state()[0].copy()
    s, st = sim.step()
    r, ost = orc.step(direct=False)
    print("capped PCG: gpu status %d iters %d converged %d rel %.3g; oracle status %d iters %d" % (
        s, st.pcg_iters, st.pcg_converged, st.pcg_rel_residual, r, ost.pcg_iters))
    assert s == CS_E_NOT_CONVERGED and r == CS_E_NOT_CONVERGED
    assert st.pcg_converged == 0 and st.pcg_iters == 3 and st.pcg_rel_residual > spec.params.pcg_rel_tol
    x = sim.state()[0]
    assert np.isfinite(x).all() and np.abs(x - x0).max() > 0  # the step still moved the cloth
    assert np.abs(x - orc.state()[0]).max() <= 1e-5  # three iterations of the same algorithm


@pytest.mark.parametrize("iters", [0, 1])
def test_collision_failure_is_reported_not_fatal(iters):
    """fold49 with 34 upper nodes moved through the lower layer needs 4 response iterations;
    with max_collision_iters 1 both deform passes end with impacts left (with 0 they never
    look)."""
    spec = fold(49)
    spec.params.max_collision_iters = iters
    sim, orc = Simulator(spec), oracle.OracleScene(spec)
    x = orc.state()[0]
    m = 49 // 2
    x1 = x.copy()
    sel = np.zeros(len(x), bool)
    sel[m * 49:] = np.linalg.norm(x[m * 49:, :2] - 0.5, axis=1) < 0.08
    x1[sel, 2] -= 2.5e-3
    for o in (sim, orc):
        o.set_state(CS_MESH_CLOTH, 0, x1, x, np.zeros_like(x))
    s, st = sim.collision_step()
    r, ost = orc.collision_step()
    msg = sim.L.cs_last_error(sim.h).decode()
    print("max_collision_iters %d: gpu status %d (%s) iters %d; oracle status %d iters %d" % (
        iters, s, msg, st.collision_iters, r, ost.collision_iters))
    assert s == CS_E_COLLISION_FAILED and r == CS_E_COLLISION_FAILED
    assert "Collision step failed" in msg
    assert st.collision_iters == ost.collision_iters == 2 * iters
    # the context stays usable: a normal step afterwards succeeds
    p = spec.params
    p.max_collision_iters = 100
    sim.set_params(p)
    s, _ = sim.collision_step()
    assert s == CS_OK
