"""Writes tests/golden/small_cases.npz: answers of the CPU oracle on small seeded scenes, the
fixtures the product is checked against (tests/test_golden.py).

    python tests/golden/make_golden.py

  hang9       FP64 assembled system of a 9 x 9 sheet on two handles (BSR row_ptr, col_idx,
              3x3 blocks, b) and the FP32 oracle's positions after each of 3 direct-solve steps
  sheets      two parallel 9 x 9 sheets 1.5 mm apart: LBVH candidate pairs (proximity
              thickness), proximity records; 1 mm apart with the lower sheet moved 2 mm up
              through the upper one: CCD impacts
  motion      a three-key motion evaluated at nine times (8 floats each)
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from clothsimulator_amd._cabi import CS_MESH_CLOTH, MeshArrays, default_params, material_struct  # noqa: E402
from clothsimulator_amd.scenes import ClothSpec, SceneSpec, grid_mesh, hang  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "small_cases.npz")

MOTION_T = np.array([0.0, 0.5, 1.25], dtype=np.float32)
MOTION_K = np.array([[1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0],
                     [1.5, 0.2, -0.1, 0.3, 0.9238795, 0.0, 0.3826834, 0.0],
                     [0.5, -0.4, 0.6, 0.0, 0.7071068, 0.7071068, 0.0, 0.0]], dtype=np.float32)
MOTION_AT = np.array([0.0, 0.1, 0.25, 0.5, 0.6, 0.9, 1.25, 1.3, 2.0], dtype=np.float32)


def two_sheets(n=9, gap=1e-3, offset=0.3):
    """One cloth mesh made of two parallel n x n sheets 0.1 m wide, the upper one shifted by
    `offset` cells in x and y."""
    xa, fa = grid_mesh(n, 0.1, origin=(0.0, 0.0, 0.0))
    h = 0.1 / (n - 1)
    xb, fb = grid_mesh(n, 0.1, origin=(offset * h, offset * h, gap))
    x = np.vstack([xa, xb])
    f = np.vstack([fa, fb + n * n])
    return SceneSpec("two_sheets", default_params(dt=0.005), [ClothSpec(MeshArrays(x, f), material_struct())],
                     [], 1), x


def crossing_state(x):
    """x0 = the rest positions, x1 = the lower sheet moved 2 mm up through the upper one."""
    nn = x.shape[0] // 2
    x0 = x.astype(np.float32).copy()
    x1 = x0.copy()
    x1[:nn, 2] += 2e-3
    return x1, x0, np.zeros_like(x0)


def main():
    out = {}
    spec = hang(9)
    o64 = oracle.OracleScene(spec, fp64=True)
    rp, ci, vals, b = o64.assemble(0)
    out.update(hang9_row_ptr=rp, hang9_col_idx=ci, hang9_vals=vals, hang9_b=b)
    o32 = oracle.OracleScene(spec)
    out["hang9_x0"] = o32.state()[0]
    xs = []
    for _ in range(3):
        r, _ = o32.step(direct=True)
        assert r == 0
        xs.append(o32.state()[0])
    out["hang9_x_steps"] = np.stack(xs)

    spec, x = two_sheets(9, gap=1.5e-3)
    o = oracle.OracleScene(spec)
    out["sheets_pairs"] = np.asarray(o.lbvh_pairs(False, 2e-3), dtype=np.int32)
    out["sheets_prox"] = o.proximities()
    spec, x = two_sheets(9, gap=1e-3)
    o = oracle.OracleScene(spec)
    o.set_state(CS_MESH_CLOTH, 0, *crossing_state(x))
    out["sheets_impacts"] = o.impacts(1e-4)

    out["motion_t"], out["motion_k"], out["motion_at"] = MOTION_T, MOTION_K, MOTION_AT
    out["motion_out"] = np.stack([oracle.motion_eval(MOTION_T, MOTION_K, t) for t in MOTION_AT])
    np.savez_compressed(OUT, **out)
    print("%s: %d arrays, %d bytes" % (OUT, len(out), os.path.getsize(OUT)))


if __name__ == "__main__":
    main()
