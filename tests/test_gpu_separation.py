"""The separation step on the device (cs_separation_step, the second half of cs_remesh) against
the oracle (oracle/separation.py): two sheets, one node of the upper one pushed through the
lower, with the frame's backup faces holding them apart.  Reference: Simulator::separationStep
(src/Simulator.cu:745-799), src/SeparationHelper.cu, src/optimization/SeparationOptimization.cu."""
import numpy as np
import pytest

from oracle import separation as S
from clothsimulator_amd import CS_MESH_CLOTH, MeshArrays, Simulator, scenes
from clothsimulator_amd._cabi import default_params, material_struct
from test_oracle_separation import crossing_sheets

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dip", [0.005, 0.02])
def test_separation_matches_oracle(dip):
    x, faces, face_uv, mof, backup = crossing_sheets(dip=dip)
    n0 = 25
    fa, fb = faces[mof == 0], faces[mof == 1] - n0
    ua = face_uv[mof == 0].reshape(-1, 2)
    ub = face_uv[mof == 1].reshape(-1, 2)
    p = default_params(dt=0.005)
    spec = scenes.SceneSpec("crossing", p, [
        scenes.ClothSpec(MeshArrays(x[:n0], fa, uv=ua, face_verts=np.arange(3 * len(fa)).reshape(-1, 3)),
                         material_struct()),
        scenes.ClothSpec(MeshArrays(x[n0:], fb, uv=ub, face_verts=np.arange(3 * len(fb)).reshape(-1, 3)),
                         material_struct())], [], 1)
    sim = Simulator(spec)
    x0 = x.copy()
    for c, sl in ((0, slice(0, n0)), (1, slice(n0, None))):
        sim.set_state(CS_MESH_CLOTH, c, x[sl], x0[sl], np.zeros_like(x[sl]))
    area = np.concatenate([sim.geometry(CS_MESH_CLOTH, c)[1] for c in (0, 1)]).astype(np.float64)
    st = sim.separation_step([(u.reshape(-1, 3, 2), xx.reshape(-1, 3, 3)) for u, xx in backup])
    assert st == 0
    xd = np.concatenate([sim.state(CS_MESH_CLOTH, c)[0] for c in (0, 1)]).astype(np.float64)
    vd = np.concatenate([sim.state(CS_MESH_CLOTH, c)[2] for c in (0, 1)]).astype(np.float64)
    xo, vo, ok, solves = S.separation_step(x, x0, x, area, faces, face_uv, mof, 2, len(x), backup,
                                           p.collision_thickness, p.dt)
    assert ok
    assert not S.find_intersections(xd, faces, mof, 2)  # the device left no crossing pair
    moved = float(np.abs(xo - x).max())
    err = float(np.abs(xd - xo).max())
    dev_solves = sim.remesh_times()[4]
    print("dip %g: solves %d (oracle %d), largest move %.3g m, device vs oracle %.3g m" % (
        dip, dev_solves, solves, moved, err))
    # the penalty solve stops on an absolute change of the objective (1e-10), which FP32 reaches
    # at a different iterate than FP64 (3 and 2 solves against 5 and 3): measured 2.6 % (dip 5 mm)
    # and 5.5 % (20 mm) of the move
    assert err <= 0.1 * moved + 2e-5
    assert np.allclose(vd, (xd - x0) / p.dt, rtol=1e-4, atol=1e-4)
