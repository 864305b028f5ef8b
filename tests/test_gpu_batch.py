"""Independent instances (CS_FLAG_INSTANCES): the batched block-Jacobi solve of all
cloths gives the same steps as each instance in its own context, and the
collision pairs stay inside their instance."""
import numpy as np
import pytest

from clothsimulator_amd import CS_MESH_CLOTH, Simulator, scenes
from clothsimulator_amd._cabi import CS_FLAG_FIXED_TUNING, CS_FLAG_INSTANCES

pytestmark = pytest.mark.gpu


def single(spec, k):
    """Instance k of a batch scene as a scene of its own."""
    p = scenes.default_params(dt=spec.params.dt)
    return scenes.SceneSpec("inst%d" % k, p, [spec.cloths[k]], [spec.obstacles[k]], spec.steps)


def test_batched_instances_match_separate_runs():
    spec = scenes.instances(count=4, seed=3, side_lo=17, side_hi=29)
    spec.params.flags |= CS_FLAG_FIXED_TUNING  # the single-cloth grouping of partial sums: comparable to 1e-5
    batch = Simulator(spec)
    its_batch = 0
    for _ in range(6):
        s, st = batch.step()
        assert s == 0, s
        its_batch += st.pcg_iters
    its_single = 0
    for k in range(4):
        sim = Simulator(single(spec, k))
        for _ in range(6):
            s, st = sim.step()
            assert s == 0, s
            its_single += st.pcg_iters
        xs, _, vs = sim.state(CS_MESH_CLOTH, 0)
        xb, _, vb = batch.state(CS_MESH_CLOTH, k)
        assert np.abs(xb - xs).max() <= 1e-5, (k, np.abs(xb - xs).max())
        assert np.abs(vb - vs).max() <= 2e-3, (k, np.abs(vb - vs).max())
    # the same iteration per cloth; only the grouping of partial sums differs
    assert abs(its_batch - its_single) <= max(4, 0.02 * its_single), (its_batch, its_single)


def test_batched_contacts_match_separate_runs():
    """Instances resting on their spheres: the batched contact assembly (one sort over every
    cloth's blocks) gives each cloth the contact blocks of its own run."""
    spec = scenes.instances(count=3, seed=5, side_lo=17, side_hi=25, gap=0.004)
    spec.params.flags |= CS_FLAG_FIXED_TUNING
    batch = Simulator(spec)
    contacts_batch = 0
    for _ in range(12):
        s, st = batch.step()
        assert s == 0, s
        contacts_batch += st.n_contact_blocks
    assert contacts_batch > 0
    contacts_single = 0
    for k in range(3):
        sim = Simulator(single(spec, k))
        for _ in range(12):
            s, st = sim.step()
            assert s == 0, s
            contacts_single += st.n_contact_blocks
        xs, _, vs = sim.state(CS_MESH_CLOTH, 0)
        xb, _, vb = batch.state(CS_MESH_CLOTH, k)
        print("instance %d: max|dx| %.3g max|dv| %.3g" % (k, np.abs(xb - xs).max(), np.abs(vb - vs).max()))
        assert np.abs(xb - xs).max() <= 1e-5, (k, np.abs(xb - xs).max())
        assert np.abs(vb - vs).max() <= 2e-3, (k, np.abs(vb - vs).max())
    assert contacts_batch == contacts_single, (contacts_batch, contacts_single)


def test_instance_pairs_stay_inside_their_instance():
    spec = scenes.instances(count=3, seed=4, side_lo=17, side_hi=21, spacing=1.0)   # neighbouring cloths touch
    sim = Simulator(spec)
    for _ in range(2):
        sim.step()
    faces = [c.mesh.n_faces for c in spec.cloths] + [o.n_faces for o in spec.obstacles]
    owner = np.repeat(np.r_[np.arange(3), np.arange(3)], faces)   # cloth k and obstacle k: instance k
    pairs = sim.lbvh_pairs(True, 0.02)
    assert len(pairs) > 0
    assert np.all(owner[pairs[:, 0]] == owner[pairs[:, 1]])
    # without the flag the same scene finds pairs across instances
    spec.params.flags &= ~CS_FLAG_INSTANCES
    sim2 = Simulator(spec)
    for _ in range(2):
        sim2.step()
    p2 = sim2.lbvh_pairs(True, 0.02)
    assert np.any(owner[p2[:, 0]] != owner[p2[:, 1]])


def test_batched_solve_autotuner_settles():
    """The batched solve times its candidate launch shapes over the first 8 solves and keeps
    one; with CS_FLAG_FIXED_TUNING it keeps (256 rows, 16 iterations per poll)."""
    spec = scenes.instances(count=6, seed=7, side_lo=17, side_hi=33)
    sim = Simulator(spec)
    for k in range(10):
        s, st = sim.step()
        assert s == 0, s
    rows, chunk, trials = sim.batch_tuning()
    print("tuned: %d rows per block, %d iterations per poll after %d timed solves" % (rows, chunk, trials))
    assert trials == 8 and rows in (128, 256) and chunk in (16, 32)
    spec.params.flags |= CS_FLAG_FIXED_TUNING
    fixed = Simulator(spec)
    for k in range(2):
        s, st = fixed.step()
        assert s == 0, s
    assert fixed.batch_tuning() == (256, 16, 0)
