"""Moving obstacles and moving handles on the GPU (cs_set_obstacle_motion,
cs_add_moving_handles, motions from scene files) against the CPU oracle.

  * a sheet held by two handles that translate and turn, lifted by a sphere moving up into
    it (proximities with a moving obstacle, its velocity in the friction terms, impacts):
    cloth positions within 1e-5 m per step of the FP32 oracle's direct solve, obstacle
    positions bit-exact with the oracle;
  * scene files (keyframe motions, a .mot file, an obstacle sequence): every moving
    obstacle stays the rigid image T(t)(base) of its mesh, T from the host evaluation;
  * the step counter the motions are evaluated at: incremented per physics step, and a
    resumed context (cs_set_step_count) continues the motion from there.
"""
import os

import numpy as np
import pytest

import clothsimulator_amd as cs
import oracle
from clothsimulator_amd import CS_E_NOT_CONVERGED, CS_MESH_CLOTH, CS_MESH_OBSTACLE, CS_OK, Simulator
from clothsimulator_amd._cabi import MeshArrays, default_params, material_struct
from clothsimulator_amd.scenes import ClothSpec, SceneSpec, grid_mesh, icosphere

pytestmark = pytest.mark.gpu

IDENT = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]


def lift_scene(n=17):
    """A 1 m sheet at z = 0.5 whose two y = 0.5 corners are handles moving by (0.2, 0, 0.1) with a
    10 degree turn over 1 s, and a sphere of radius 0.25 starting 1.5 mm under it and rising
    2 cm in 0.1 s."""
    x, f = grid_mesh(n, 1.0, origin=(-0.5, -0.5, 0.5))
    sx, sf = icosphere(2, 0.25)
    sx = sx + np.array([0.0, 0.0, 0.5 - 0.25 - 1.5e-3])
    handles = [(n - 1) * n + 0, (n - 1) * n + (n - 1)]
    hkeys = [IDENT, [1.0, 0.2, 0.0, 0.1, *cs.axis_angle((0, 0, 1), 10.0)]]
    okeys = [IDENT, [1.0, 0.0, 0.0, 0.02, 1.0, 0.0, 0.0, 0.0]]
    return SceneSpec("lift%d" % n, default_params(dt=0.005),
                     [ClothSpec(MeshArrays(x, f), material_struct(), handles, handle_motion=0)],
                     [MeshArrays(sx, sf)], 12, motions=[([0.0, 1.0], hkeys), ([0.0, 0.1], okeys)],
                     obstacle_motions=[1])


def test_moving_obstacle_and_handles_match_oracle():
    spec = lift_scene()
    sim = Simulator(spec)
    ref = oracle.OracleScene(spec)
    xo0 = sim.state(CS_MESH_OBSTACLE, 0)[0]
    assert np.array_equal(xo0, ref.state(CS_MESH_OBSTACLE, 0)[0])
    worst, prox, imp = 0.0, 0, 0
    for k in range(spec.steps):
        x, x0, v = ref.state()
        sim.set_state(CS_MESH_CLOTH, 0, x, x0, v)
        s, st = sim.step()
        assert s in (CS_OK, CS_E_NOT_CONVERGED) and st.pcg_converged == 1
        r, ost = ref.step(direct=True)
        assert r == CS_OK
        assert st.collision_deform == 0 and ost.collision_deform == 0
        d = np.abs(sim.state()[0] - ref.state()[0]).max()
        go, _, gv = sim.state(CS_MESH_OBSTACLE, 0)
        oo, _, ov = ref.state(CS_MESH_OBSTACLE, 0)
        assert np.array_equal(go, oo) and np.array_equal(gv, ov), k
        worst = max(worst, d)
        prox += st.n_proximities
        imp += st.n_impacts
        print("lift17 step %d: prox %d (oracle %d) impacts %d max|dx| %.3g obstacle top %.6f" % (
            k, st.n_proximities, ost.n_proximities, st.n_impacts, d, go[:, 2].max()), flush=True)
    assert sim.step_count() == spec.steps
    assert worst <= 1e-5
    assert prox > 0 and imp > 0
    # the sphere rose by the motion's full 2 cm (t = 0.06 s of 0.1 s: the smoothstep 0.648)
    rise = sim.state(CS_MESH_OBSTACLE, 0)[0][:, 2].max() - xo0[:, 2].max()
    assert rise == pytest.approx(0.02 * (3 * 0.6 ** 2 - 2 * 0.6 ** 3), abs=2e-6)


TRI = "v -1 -1 0\nv 1 -1 0\nv 0 1 0\nf 1 2 3\n"
BALL_KEYS = """[{"time": 0, "transform": {"translate": [0, 0, -3]}},
                {"time": 0.05, "transform": {"translate": [0.5, 0, -3], "rotate": [30, 1, 1, 0]}},
                {"time": 0.1, "transform": {"scale": 1.5, "translate": [1, 0, -4]}}]"""


def write_motion_scene(d):
    from test_oracle_motion import motion_scene, write, write_mot
    rng = np.random.default_rng(9)
    frames = [[np.concatenate([rng.normal(size=3) * 0.1 + [0, 0, -3], rng.normal(size=4)]) for _ in range(2)]
              for _ in range(5)]
    write_mot(d, frames, 2)
    write(d, "body0000.obj", TRI)
    write(d, "body0001.obj", TRI)
    mot = '{"motfile": "walk.mot", "fps": 30, "transform": {"translate": [0, 0, -1]}}'
    return motion_scene(d, mot, '"body%04d.obj"'), motion_scene_keys(d)


def motion_scene_keys(d):
    from test_oracle_motion import motion_scene
    sub = os.path.join(d, "keys")
    os.makedirs(sub, exist_ok=True)
    return motion_scene(sub, "[%s]" % BALL_KEYS, '[{"mesh": "tri.obj", "motion": 0}]')


@pytest.mark.parametrize("which", ["mot", "keys"])
def test_scene_file_obstacles_move_rigidly(tmp_path, which):
    d = str(tmp_path)
    mot_path, key_path = write_motion_scene(d)
    path, root = (mot_path, d) if which == "mot" else (key_path, os.path.join(d, "keys"))
    sim = Simulator.from_config(path, root)
    base = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], dtype=np.float64)
    dt = np.float32(0.04) / np.float32(8)
    n_obs = sim.info.n_obstacles
    for k in range(1, 9):
        s, st = sim.step()
        assert s in (CS_OK, CS_E_NOT_CONVERGED)
        time = np.float32(k) * dt
        for o in range(n_obs):
            T = cs.scene_motion(path, o, time, root).as_array()
            want = oracle.transform_points(T, base)
            got = sim.state(CS_MESH_OBSTACLE, o)[0]
            err = np.abs(got - want).max()
            print("%s obstacle %d step %d: |x - T(t)(base)| %.3g" % (which, o, k, err), flush=True)
            assert err <= 1e-5 * max(1.0, np.abs(want).max())
    assert sim.step_count() == 8


def test_resumed_step_count_continues_the_motion(tmp_path):
    d = str(tmp_path)
    _, key_path = write_motion_scene(d)
    root = os.path.join(d, "keys")
    sim = Simulator.from_config(key_path, root)
    base = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], dtype=np.float64)
    dt = np.float32(0.04) / np.float32(8)
    sim.state(CS_MESH_OBSTACLE, 0)  # the device state exists before the counter moves
    sim.set_step_count(12)
    # placed at T(12 dt)(base) with x0 = x, as the reference's resume transforms its obstacles
    T12 = cs.scene_motion(key_path, 0, np.float32(12) * dt, root).as_array()
    x, x0, _ = sim.state(CS_MESH_OBSTACLE, 0)
    assert np.abs(x - oracle.transform_points(T12, base)).max() <= 1e-6 and np.array_equal(x, x0)
    s, _ = sim.physics_step()
    assert s in (CS_OK, CS_E_NOT_CONVERGED)
    assert sim.step_count() == 13
    T = cs.scene_motion(key_path, 0, np.float32(13) * dt, root).as_array()
    assert np.abs(sim.state(CS_MESH_OBSTACLE, 0)[0] - oracle.transform_points(T, base)).max() <= 1e-5
