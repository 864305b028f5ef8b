"""Pins of the oracle's material tables and lookups for NON-constant materials (no GPU).

A constant material makes every interpolation weight cancel, so these checks use tables whose
entries differ cell by cell, and reach each result by a route other than the oracle's code:

  stretching table  every entry of the 30^3 table from numpy's symmetric eigendecomposition
                    of 2G (largest eigenvalue -> strain weight, its eigenvector's angle ->
                    angle weight; src/Material.cu:45-85), plus the diagonal closed forms
                    (G01 = 0: Q = I gives angle weight 0; G11 > G00: the swapped Q gives the
                    last angle sample);
  stretching lookup scipy's RegularGridInterpolator (linear) over the same table, inside the
                    grid and clamped outside it (src/Material.cu:91-124);
  bending lookup    a table affine in its two indices is reproduced exactly by bilinear
                    interpolation, so k = c0 + c1 * bias + c2 * value for hinges at several
                    material-space angles and curvatures (src/Material.cu:126-155), read back
                    through the bending force f = -1/4 k |e|^2 / (A0 + A1) * theta * dtheta
                    (src/PhysicsHelper.cu:327-353).
"""
import numpy as np
import pytest

import oracle
from clothsimulator_amd._cabi import MeshArrays, default_params, material_struct
from clothsimulator_amd.scenes import ClothSpec, SceneSpec, hang

N = 30

# an anisotropic material with negative entries (the table clamps samples at 0)
ROWS = [[16.6, -14.7, 34.5, 36.9],
        [46.4, 53.7, 261.0, -29.7],
        [49.3, 87.6, 206.4, 22.8],
        [124.0, 105.2, 366.0, 44.2],
        [139.7, 73.6, 413.6, -174.4],
        [127.4, 85.8, 405.9, 32.2]]
STRETCH_MULT = 1.7


def _sample_eigh(g00, g01, g11, rows):
    """One table entry from numpy's eigendecomposition (FP64)."""
    rows = np.asarray(rows, dtype=np.float64)
    if g01 == 0.0:
        lam = 2.0 * max(g00, g11)
        theta = 0.0 if g00 >= g11 else np.pi / 2  # the eigenvector of the larger entry
    else:
        w, V = np.linalg.eigh(2.0 * np.array([[g00, g01], [g01, g11]]))
        lam = w[1]
        v = V[:, 1]
        theta = np.arctan2(abs(v[1]), abs(v[0]))  # angle of the principal axis, folded into [0, pi/2]
    sw = np.clip((np.sqrt(lam + 1.0) - 1.0) * 6.0, 0.0, 1.0 - 1e-6)  # strain sample 0 or 1
    aw = np.clip(theta / np.pi * 8.0, 0.0, 4.0 - 1e-6)               # angle samples 0..4
    ai = int(aw)
    aw -= ai
    # strain sample 0 is row 0 at every angle; strain sample 1 is rows 1..5 over the angles
    s = (1.0 - sw) * rows[0] + sw * ((1.0 - aw) * rows[1 + ai] + aw * rows[2 + ai])
    return np.maximum(s, 0.0) * 2.0


def test_stretching_table_from_eigendecomposition():
    table, _, _ = oracle.material_table(material_struct(stretching=ROWS, stretching_mult=STRETCH_MULT))
    ref = np.zeros((N, N, N, 4))
    for i in range(N):
        for j in range(N):
            for k in range(N):
                ref[i, j, k] = _sample_eigh(i / N - 0.25, k / N, j / N - 0.25, ROWS) * STRETCH_MULT
    scale = np.abs(ref).max()
    err = np.abs(table - ref).max()
    print("stretching table vs eigh: max err %.3g of %.3g" % (err, scale))
    assert err <= 1e-5 * scale
    assert (table >= 0).all() and (ref == 0).any()  # the clamp is exercised


def test_stretching_table_diagonal_closed_forms():
    table, _, _ = oracle.material_table(material_struct(stretching=ROWS))
    r = np.asarray(ROWS, dtype=np.float64)
    for i in range(N):
        for j in range(N):
            g00, g11 = i / N - 0.25, j / N - 0.25
            if g00 >= g11:
                # Q = I: angle weight 0; strain weight from the eigenvalue 2 G00
                sw = min(max((np.sqrt(2 * g00 + 1) - 1) * 6, 0.0), 1 - 1e-6)
                want = np.maximum((1 - sw) * r[0] + sw * r[1], 0) * 2
            else:
                # swapped Q: angle weight 4 - 1e-6, i.e. all but 1e-6 of the last angle sample
                sw = min(max((np.sqrt(2 * g11 + 1) - 1) * 6, 0.0), 1 - 1e-6)
                aw = 1 - 1e-6
                want = np.maximum((1 - sw) * r[0] + sw * ((1 - aw) * r[4] + aw * r[5]), 0) * 2
            assert np.allclose(table[i, j, 0], want, rtol=2e-5, atol=1e-4), (i, j, table[i, j, 0], want)


def test_stretching_lookup_is_scipy_trilinear():
    from scipy.interpolate import RegularGridInterpolator
    mat = material_struct(stretching=ROWS)
    spec = hang(5)
    spec.cloths[0].material = mat
    o = oracle.OracleScene(spec, fp64=True)
    table, _, _ = oracle.material_table(mat)
    axis = np.arange(N, dtype=np.float64)
    interp = RegularGridInterpolator((axis, axis, axis), table.astype(np.float64), method="linear")
    rng = np.random.default_rng(3)
    g = np.column_stack([rng.uniform(-0.35, 0.8, 300), rng.uniform(-1.0, 1.0, 300), rng.uniform(-0.35, 0.8, 300)])
    for g00, g01, g11 in g:
        # table coordinates, clamped to the grid as the lookup does
        c = np.clip([(g00 + 0.25) * N, (g11 + 0.25) * N, abs(g01) * N], 0.0, N - 1 - 1e-6)
        want = interp(c[None, :])[0]
        got = o.stretching_stiffness(0, g00, g01, g11)
        assert np.allclose(got, want, rtol=1e-6, atol=1e-6 * np.abs(table).max()), ((g00, g01, g11), got, want)


def _hinge(phi, beta, mat):
    """Two faces on edge (0, 1) folded by about beta, material space rotated by phi."""
    x = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.04, 0.08, 0.0],
                  [0.06, -0.07 * np.cos(beta), 0.07 * np.sin(beta)]])
    flat = np.array([[0.0, 0.0], [0.1, 0.0], [0.04, 0.08], [0.06, -0.07]])
    R = np.array([[np.cos(phi), -np.sin(phi)], [np.sin(phi), np.cos(phi)]])
    uv = flat @ R.T
    faces = np.array([[0, 1, 2], [1, 0, 3]])
    mesh = MeshArrays(x, faces, uv=uv, face_verts=faces)
    spec = SceneSpec("hinge", default_params(), [ClothSpec(mesh, mat)], [], 1)
    o = oracle.OracleScene(spec, fp64=True)
    nodes, opp, adj = o.edges()
    e = int(np.where((adj >= 0).all(1))[0][0])
    return o, e, nodes[e], opp[e], mesh.uv.astype(np.float64)


@pytest.mark.parametrize("phi", [0.3, 0.9, 1.3, 2.2])
@pytest.mark.parametrize("beta", [0.4, 1.2, 2.0])
def test_bending_lookup_affine_table(phi, beta):
    c0, c1, c2, mult = 5e-6, 3e-6, 2e-6, 0.7
    bend = [[c0 + c1 * i + c2 * j for j in range(5)] for i in range(3)]
    o, e, en, eo, uv = _hinge(phi, beta, material_struct(bending=bend, bending_mult=mult))
    f, J, angle = o.bending_force(e)
    P = o.state(fp64=True)[0]
    x0, x1, x2, x3 = P[en[0]], P[en[1]], P[eo[0]], P[eo[1]]
    tri = lambda a, b, c: 0.5 * abs((b - a)[0] * (c - a)[1] - (b - a)[1] * (c - a)[0])
    area = tri(uv[0], uv[1], uv[2]) + tri(uv[0], uv[1], uv[3])  # rest areas from material space
    length = np.linalg.norm(x1 - x0)                               # current edge length
    d = uv[1] - uv[0]
    bias = abs(np.arctan2(d[1], d[0])) * 4 / np.pi
    if bias > 2:
        bias = 4 - bias
    value = 0.2 * 0.5 * length * angle / area
    assert 0.05 < bias < 1.95 and 0.2 < value < 3.9  # strictly inside the table
    k = mult * (c0 + c1 * bias + c2 * value)
    coef = -0.25 * k * length ** 2 / area
    # dtheta = f / (coef * angle), and J = coef * dtheta dtheta^T
    dth = f / (coef * angle)
    assert np.allclose(J, coef * np.outer(dth, dth), rtol=1e-6, atol=1e-9 * np.abs(J).max())
    # the gradient itself is pinned by finite differences in test_oracle.py; here its norm must be
    # that of d(angle)/dx: |dtheta|^2 = sum over the 4 nodes
    n0 = np.cross(x1 - x0, x2 - x0); h0 = np.linalg.norm(n0) / np.linalg.norm(x1 - x0)
    n1 = np.cross(x0 - x1, x3 - x1); h1 = np.linalg.norm(n1) / np.linalg.norm(x1 - x0)
    # the two opposite nodes move the angle by 1/h each along their face normals
    assert np.linalg.norm(dth[6:9]) == pytest.approx(1 / h0, rel=1e-6)
    assert np.linalg.norm(dth[9:12]) == pytest.approx(1 / h1, rel=1e-6)
