"""The aggregation multigrid preconditioner (CS_PRECOND_AMG) against its CPU
restatement in the oracle: identical hierarchy (integer topology), the V-cycle
operator within FP32 tolerance, symmetric positive definiteness of the
preconditioner, and PCG solutions within 1e-5 m of the direct solve."""
import numpy as np
import pytest

import oracle
from clothsimulator_amd import CS_MESH_CLOTH, Simulator, scenes
from clothsimulator_amd._cabi import CS_PRECOND_AMG, CS_PRECOND_BLOCK_JACOBI

pytestmark = pytest.mark.gpu


def perturbed(spec, seed=0, amp=0.01, vel=0.1):
    rng = np.random.default_rng(seed)
    x = spec.cloths[0].mesh.x.astype(np.float64)
    x = x + rng.normal(0, amp, x.shape)
    v = rng.normal(0, vel, x.shape)
    return x.astype(np.float32), v.astype(np.float32)


def pair(spec):
    spec.params.precond = CS_PRECOND_AMG
    sim = Simulator(spec)
    o = oracle.OracleScene(spec, fp64=True)
    return sim, o


@pytest.mark.parametrize("name", ["hang33", "drape33"])
def test_hierarchy_matches_oracle(name):
    spec = scenes.by_name(name)
    sim, o = pair(spec)
    sizes, _ = sim.precondition(0)
    assert sizes == o.amg_levels(0)
    assert len(sizes) >= 2 and sizes[-1] <= 64


@pytest.mark.parametrize("name", ["hang33", "drape33"])
def test_vcycle_matches_oracle(name):
    spec = scenes.by_name(name)
    sim, o = pair(spec)
    x, v = perturbed(spec)
    for s in (sim, o):
        s.set_state(CS_MESH_CLOTH, 0, x, x, v)
    sim.assemble(0)
    rng = np.random.default_rng(1)
    n = x.shape[0]
    for trial in range(3):
        r = rng.normal(0, 1, (n, 3)).astype(np.float32)
        _, zg = sim.precondition(0, r)
        zo = o.amg_apply(r.astype(np.float64), 0)
        err = np.abs(zg - zo).max() / np.abs(zo).max()
        assert err < 2e-3, (trial, err)


def test_preconditioner_is_spd():
    spec = scenes.by_name("hang33")
    sim, _ = pair(spec)
    x, v = perturbed(spec)
    sim.set_state(CS_MESH_CLOTH, 0, x, x, v)
    sim.assemble(0)
    rng = np.random.default_rng(2)
    n = x.shape[0]
    r1 = rng.normal(0, 1, (n, 3)).astype(np.float32)
    r2 = rng.normal(0, 1, (n, 3)).astype(np.float32)
    _, z1 = sim.precondition(0, r1)
    _, z2 = sim.precondition(0, r2)
    a, b = float((r2 * z1).sum()), float((r1 * z2).sum())
    assert abs(a - b) <= 1e-4 * max(abs(a), abs(b))   # symmetric
    assert float((r1 * z1).sum()) > 0 and float((r2 * z2).sum()) > 0


@pytest.mark.parametrize("name,steps", [("hang33", 4), ("drape33", 4)])
def test_amg_steps_match_direct_solve(name, steps):
    """Each physics step from the oracle's state: GPU AMG-PCG vs the oracle's direct solve."""
    spec = scenes.by_name(name)
    sim, o = pair(spec)
    iters = []
    for k in range(steps):
        x, x0, v = o.state(fp64=False)
        sim.set_state(CS_MESH_CLOTH, 0, x, x0, v)
        s, st = sim.physics_step()
        assert s == 0, s
        iters.append(st.pcg_iters)
        o.physics_step(direct=True)
        xg, _, _ = sim.state(CS_MESH_CLOTH, 0)
        xo, _, _ = o.state(fp64=False)
        assert np.abs(xg - xo).max() <= 1e-5, (k, np.abs(xg - xo).max())
    # the same iteration on the oracle (FP64 restatement) needs about as many iterations
    spec2 = scenes.by_name(name)
    spec2.params.precond = CS_PRECOND_AMG
    o2 = oracle.OracleScene(spec2, fp64=True)
    s, st = o2.physics_step()
    assert abs(st.pcg_iters - iters[0]) <= max(3, 0.2 * st.pcg_iters)


def test_amg_needs_far_fewer_iterations_than_jacobi():
    spec = scenes.by_name("drape65")
    x, v = perturbed(spec, amp=0.002, vel=0.05)
    counts = {}
    for pre in (CS_PRECOND_BLOCK_JACOBI, CS_PRECOND_AMG):
        spec.params.precond = pre
        sim = Simulator(spec)
        sim.set_state(CS_MESH_CLOTH, 0, x, x, v)
        s, st = sim.physics_step()
        assert s == 0
        counts[pre] = st.pcg_iters
    assert counts[CS_PRECOND_AMG] * 2 < counts[CS_PRECOND_BLOCK_JACOBI], counts


def test_vcycle_matches_oracle_with_contacts():
    """A drape resting on its sphere: contact blocks enter level 0 and, through their diagonal
    part, the coarse operators; the GPU cycle against the oracle's on the same state."""
    spec = scenes.by_name("drape33")
    spec.params.precond = CS_PRECOND_BLOCK_JACOBI
    run = Simulator(spec)
    for _ in range(60):
        run.step()
    x, x0, v = run.state(CS_MESH_CLOTH, 0)
    spec = scenes.by_name("drape33")
    sim, o = pair(spec)
    for s in (sim, o):
        s.set_state(CS_MESH_CLOTH, 0, x, x0, v)
    sim.assemble(0)
    assert len(sim.proximities()) > 0
    rng = np.random.default_rng(3)
    n = x.shape[0]
    for trial in range(2):
        r = rng.normal(0, 1, (n, 3)).astype(np.float32)
        _, zg = sim.precondition(0, r)
        zo = o.amg_apply(r.astype(np.float64), 0)
        err = np.abs(zg - zo).max() / np.abs(zo).max()
        assert err < 2e-3, (trial, err)


@pytest.mark.parametrize("precond", ["amg", "amg-learned"])
def test_graph_replayed_cycles_equal_direct_launches(precond):
    """The multigrid cycle replayed from captured HIP graphs (CS_FLAG_GRAPH_CYCLES) gives
    bit-identical steps to the same kernels launched one by one, over several steps with
    eager work between the replays (every step reassembles and re-solves)."""
    from clothsimulator_amd._cabi import CS_FLAG_GRAPH_CYCLES, CS_PRECOND_AMG_LEARNED
    runs = []
    for flags in (CS_FLAG_GRAPH_CYCLES, 0):
        spec = scenes.crumple(33, steps=8)
        spec.params.precond = CS_PRECOND_AMG if precond == "amg" else CS_PRECOND_AMG_LEARNED
        spec.params.flags |= flags
        sim = Simulator(spec)
        its = []
        for _ in range(8):
            s, st = sim.step()
            assert s == 0, s
            its.append(st.pcg_iters)
        runs.append((sim.state(CS_MESH_CLOTH, 0)[0].copy(), its))
    print("%s iterations: graphs %s, direct %s" % (precond, runs[0][1], runs[1][1]))
    assert runs[0][1] == runs[1][1]
    assert np.array_equal(runs[0][0], runs[1][0])
