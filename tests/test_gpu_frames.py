"""The asynchronous frame download (cs_frame_begin / cs_frame_end): each slot returns the cloth
states of the moment it was begun, whatever ran after, for every cloth; slots are reused; a
remesh between begin and end does not change the snapshot.  It replaces the synchronous copy in
Simulator::save (src/Simulator.cu:950-957, src/Cloth.cu:865-887)."""
import numpy as np
import pytest

from clothsimulator_amd import CS_MESH_CLOTH, ClothError, Simulator, scenes
from clothsimulator_amd._cabi import CS_E_STATE

pytestmark = pytest.mark.gpu


def step(sim, k):
    for _ in range(k):
        s, _ = sim.step()
        assert s == 0, s


def test_slots_hold_the_state_at_their_begin():
    spec = scenes.instances(count=2, seed=2, side_lo=17, side_hi=21, gap=0.01)
    sim = Simulator(spec)
    with pytest.raises(ClothError) as e:
        sim.frame_end(1, 0)
    assert e.value.status == CS_E_STATE
    step(sim, 2)
    want0 = [sim.state(CS_MESH_CLOTH, c) for c in range(2)]
    sim.frame_begin(0)
    step(sim, 3)
    want1 = [sim.state(CS_MESH_CLOTH, c) for c in range(2)]
    sim.frame_begin(1)
    step(sim, 2)
    for c in range(2):
        x, v = sim.frame_end(0, c)
        assert np.array_equal(x, want0[c][0]) and np.array_equal(v, want0[c][2])
        x, v = sim.frame_end(1, c)
        assert np.array_equal(x, want1[c][0]) and np.array_equal(v, want1[c][2])
    assert not np.array_equal(want0[0][0], want1[0][0])
    # slot 0 again, begun while slot 1's snapshot is kept
    want2 = [sim.state(CS_MESH_CLOTH, c) for c in range(2)]
    sim.frame_begin(0)
    step(sim, 1)
    for c in range(2):
        x, v = sim.frame_end(0, c)
        assert np.array_equal(x, want2[c][0]) and np.array_equal(v, want2[c][2])
        x, _ = sim.frame_end(1, c)
        assert np.array_equal(x, want1[c][0])


def test_snapshot_survives_a_remesh():
    spec = scenes.drape_remesh(17)
    sim = Simulator(spec)
    step(sim, 40)
    x_before, _, v_before = sim.state(CS_MESH_CLOTH, 0)
    sim.frame_begin(1)
    sim.remesh()
    step(sim, 1)
    n_after = sim.counts(CS_MESH_CLOTH, 0)[0]
    x, v = sim.frame_end(1, 0)
    print("remeshed %d -> %d nodes; the snapshot keeps %d" % (len(x_before), n_after, len(x)))
    assert n_after != len(x_before)
    assert np.array_equal(x, x_before) and np.array_equal(v, v_before)
