"""Independent instances (CS_FLAG_INSTANCES): the batched block-Jacobi solve of all
cloths gives the same steps as each instance in its own context, and the
collision pairs stay inside their instance."""
import numpy as np
import pytest

from clothsimulator_amd import CS_MESH_CLOTH, Simulator, scenes
from clothsimulator_amd._cabi import CS_FLAG_FIXED_TUNING, CS_FLAG_INSTANCES, CS_FLAG_PERSISTENT_BATCH

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


@pytest.mark.parametrize("count,side_lo,side_hi,steps", [(3, 17, 25, 8), (256, 33, 65, 3)])
def test_persistent_batched_solve_equals_launched(count, side_lo, side_hi, steps):
    """CS_FLAG_PERSISTENT_BATCH: the iterations between two polls as one cooperative launch with
    grid barriers give the launched kernels' steps bit for bit (the same per-block bodies in the
    same order), on instances starting 4 mm above their spheres.  256 instances of sides 33..65
    make about 2,500 (cloth, 256-row block) pairs, more than the roughly 1,500 workgroups the
    device holds at once, so each workgroup walks several; their static matrices (over 300 MB) also
    put the launched SpMV on non-temporal loads (above 128 MB), the persistent one on cached
    ones."""
    states = []
    for flags in (CS_FLAG_PERSISTENT_BATCH, 0):
        spec = scenes.instances(count=count, seed=5, side_lo=side_lo, side_hi=side_hi, gap=0.004)
        spec.params.flags |= CS_FLAG_FIXED_TUNING | flags
        sim = Simulator(spec)
        its = []
        for _ in range(steps):
            s, st = sim.step()
            assert s == 0, s
            its.append(st.pcg_iters)
        print("flags %d: iterations %s" % (flags, its), flush=True)
        states.append(([sim.state(CS_MESH_CLOTH, k) for k in range(count)], its))
    (a, ia), (b, ib) = states
    assert ia == ib
    for k in range(count):
        for u, v in zip(a[k], b[k]):
            assert np.array_equal(u, v), k


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


def test_batch32_up_to_side129_matches_separate_runs_and_oracle():
    """32 instances with sides drawn from 17..129 (most above 4,096 faces): the batch builds every
    tree with the one-workgroup kernel, each separate run builds its large tree with the
    per-mesh kernels.  The smallest and the largest instance, resynchronised with the oracle
    before every step, are within 1e-5 m of the oracle's PCG step (1e-7 relative) after each of
    5 steps; then every instance, run 3 more steps alone from the batch's state, is within
    1e-5 m of the batch's own 3 steps."""
    import oracle
    spec = scenes.instances(count=32, seed=11, side_lo=17, side_hi=129)
    spec.params.flags |= CS_FLAG_FIXED_TUNING
    sides = [int(round(np.sqrt(c.mesh.n_nodes))) for c in spec.cloths]
    print("sides: min %d max %d, %d above 4,096 faces" % (min(sides), max(sides),
                                                         sum(2 * (n - 1) ** 2 > 4096 for n in sides)))
    assert max(sides) >= 120 and sum(2 * (n - 1) ** 2 > 4096 for n in sides) > 4
    batch = Simulator(spec)
    picks = [int(np.argmin(sides)), int(np.argmax(sides))]
    refs = {}
    for k in picks:
        one = single(spec, k)
        one.params.pcg_rel_tol = 1e-7
        refs[k] = oracle.OracleScene(one)
    worst_o = 0.0
    for step in range(5):
        for k in picks:
            batch.set_state(CS_MESH_CLOTH, k, *refs[k].state())
        s, st = batch.step()
        assert s == 0, s
        for k in picks:
            r, _ = refs[k].step(direct=False)
            assert r == 0
            d = float(np.abs(batch.state(CS_MESH_CLOTH, k)[0] - refs[k].state()[0]).max())
            worst_o = max(worst_o, d)
            print("step %d instance %d (side %d): max|dx| vs oracle %.3g" % (step, k, sides[k], d), flush=True)
    assert worst_o <= 1e-5
    # separate runs of every instance from the batch's current state
    state = [batch.state(CS_MESH_CLOTH, k) for k in range(32)]
    for _ in range(3):
        s, st = batch.step()
        assert s == 0, s
    worst = 0.0
    for k in range(32):
        sim = Simulator(single(spec, k))
        sim.set_state(CS_MESH_CLOTH, 0, *state[k])
        for _ in range(3):
            s, st = sim.step()
            assert s == 0, s
        worst = max(worst, float(np.abs(batch.state(CS_MESH_CLOTH, k)[0] - sim.state()[0]).max()))
        sim.close()
    print("32 instances: max|dx| batch vs separate %.3g" % worst)
    assert worst <= 1e-5
