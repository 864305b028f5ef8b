"""Independent instances with the multigrid preconditioner (CS_FLAG_INSTANCES + CS_PRECOND_AMG):
every cloth keeps its own aggregation hierarchy, the batched PCG runs one V(1,1) cycle per cloth
and iteration with each level launched once for all cloths.  The cycle is the single-cloth cycle
(same aggregates, Galerkin operators, smoother and exact coarsest solve); only the coarsest levels
differ in form: the single-cloth solve applies its small levels as one dense operator, the batch
runs them level by level.  So each instance must match its own single-cloth multigrid run to the
solve tolerance, need a multigrid's iteration count, and meet 1e-5 on its true residual."""
import numpy as np
import pytest

from clothsimulator_amd import CS_MESH_CLOTH, Simulator, scenes
from clothsimulator_amd._cabi import CS_FLAG_FIXED_TUNING, CS_PRECOND_AMG, CS_PRECOND_BLOCK_JACOBI

pytestmark = pytest.mark.gpu


def single(spec, k):
    p = scenes.default_params(dt=spec.params.dt)
    p.precond = spec.params.precond
    return scenes.SceneSpec("inst%d" % k, p, [spec.cloths[k]], [spec.obstacles[k]], spec.steps)


def _batch(count, seed, lo, hi, gap=None, precond=CS_PRECOND_AMG):
    spec = scenes.instances(count=count, seed=seed, side_lo=lo, side_hi=hi, gap=gap)
    spec.params.flags |= CS_FLAG_FIXED_TUNING
    spec.params.precond = precond
    return spec


def _compare(spec, steps):
    batch = Simulator(spec)
    its_b, worst_res = [], 0.0
    for _ in range(steps):
        s, st = batch.step()
        assert s == 0, s
        its_b.append(st.pcg_iters)
        worst_res = max(worst_res, st.pcg_rel_residual)
    its_s = np.zeros(steps, dtype=int)
    worst = 0.0
    for k in range(len(spec.cloths)):
        sim = Simulator(single(spec, k))
        for j in range(steps):
            s, st = sim.step()
            assert s == 0, s
            its_s[j] += st.pcg_iters
        d = float(np.abs(batch.state(CS_MESH_CLOTH, k)[0] - sim.state()[0]).max())
        worst = max(worst, d)
        sim.close()
    print("batched multigrid vs separate: max|dx| %.3g, iterations per step %s vs %s, worst true residual %.3g"
          % (worst, its_b, list(its_s), worst_res))
    return worst, np.array(its_b), its_s, worst_res


def test_batched_multigrid_matches_separate_multigrid_runs():
    spec = _batch(6, 3, 17, 65)
    worst, its_b, its_s, res = _compare(spec, 6)
    assert res <= 1e-5
    assert worst <= 1e-5
    assert np.all(np.abs(its_b - its_s) <= np.maximum(3, 0.1 * its_s)), (its_b, its_s)


def test_streamed_batched_multigrid_matches_separate_runs():
    """24 instances of sides 113..129: about 183 MB of static blocks (40 B each, about 13 per
    node), above the 128 MB from which the batched SpMV, level-0 residual and post-smoothing read
    their matrices with non-temporal loads; every instance still equals its separate multigrid
    run to 1e-5 m."""
    spec = _batch(24, 5, 113, 129)
    mb = sum(c.mesh.n_nodes for c in spec.cloths) * 13 * 40 / 2 ** 20
    assert mb > 160, mb
    worst, its_b, its_s, res = _compare(spec, 3)
    assert res <= 1e-5
    assert worst <= 1e-5
    assert np.all(np.abs(its_b - its_s) <= np.maximum(3, 0.1 * its_s)), (its_b, its_s)


def test_batched_multigrid_with_contacts_matches_separate_runs():
    """Cloths resting on their spheres: level 1 of each cloth adds its own contact diagonal
    blocks (the batched contact assembly's per-cloth views).  In contact, a difference at the
    solve tolerance can change which impacts a later collision step finds, so a free run of
    12 steps drifts apart by more than one step's bar (3.2e-4 m measured): each separate run
    restarts every step from the batch's state, and every step must agree within 1e-5 m."""
    spec = _batch(4, 5, 17, 41, gap=0.004)
    batch = Simulator(spec)
    sims = [Simulator(single(spec, k)) for k in range(4)]
    worst, contacts, its_b, its_s = 0.0, 0, [], []
    for _ in range(12):
        state = [batch.state(CS_MESH_CLOTH, k) for k in range(4)]
        s, st = batch.step()
        assert s == 0, s
        assert st.pcg_rel_residual <= 1e-5
        contacts += st.n_contact_blocks
        its_b.append(st.pcg_iters)
        n = 0
        for k, sim in enumerate(sims):
            sim.set_state(CS_MESH_CLOTH, 0, *state[k])
            s, st1 = sim.step()
            assert s == 0, s
            n += st1.pcg_iters
            worst = max(worst, float(np.abs(batch.state(CS_MESH_CLOTH, k)[0] - sim.state()[0]).max()))
        its_s.append(n)
    print("contact steps: max|dx| per step %.3g, contact blocks %d, iterations %s vs %s" % (
        worst, contacts, its_b, its_s))
    assert contacts > 0
    assert worst <= 1e-5
    its_b, its_s = np.array(its_b), np.array(its_s)
    assert np.all(np.abs(its_b - its_s) <= np.maximum(3, 0.1 * its_s)), (its_b, its_s)


def test_batch_mixes_multigrid_and_block_jacobi_cloths():
    """Cloths of 64 nodes or fewer (sides 5, 7 and 8 here) have no usable hierarchy: they run
    block Jacobi inside the same batched solve, as their single-cloth runs do."""
    spec = _batch(6, 3, 5, 33)
    sides = [int(round(np.sqrt(c.mesh.n_nodes))) for c in spec.cloths]
    print("sides", sides)
    assert sum(n <= 8 for n in sides) >= 2 and max(sides) >= 17
    worst, its_b, its_s, res = _compare(spec, 4)
    assert res <= 1e-5 and worst <= 1e-5
    assert np.all(np.abs(its_b - its_s) <= np.maximum(3, 0.1 * its_s)), (its_b, its_s)


def test_batched_multigrid_beats_block_jacobi_iterations():
    """32 instances up to side 129, resting on their spheres: the batched multigrid needs a
    fraction of block Jacobi's iterations per cloth, and both meet the tolerance on every cloth."""
    counts = {}
    for pre in (CS_PRECOND_BLOCK_JACOBI, CS_PRECOND_AMG):
        spec = _batch(32, 11, 17, 129, gap=0.004, precond=pre)
        sim = Simulator(spec)
        its = 0
        for _ in range(8):
            s, st = sim.step()
            assert s == 0, s
            assert st.pcg_converged == 1 and st.pcg_rel_residual <= 1e-5
            its += st.pcg_iters
        counts[pre] = its
        sim.close()
    print("iterations over 8 steps, 32 cloths: block Jacobi %d, multigrid %d" % (
        counts[CS_PRECOND_BLOCK_JACOBI], counts[CS_PRECOND_AMG]))
    assert counts[CS_PRECOND_AMG] * 3 < counts[CS_PRECOND_BLOCK_JACOBI]
