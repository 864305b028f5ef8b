"""Pins of the oracle's impact response (CollisionOptimization + Optimization::solve,
src/optimization/CollisionOptimization.cu:37-145, src/optimization/Optimization.cu:137-181),
which solves

    minimise 1/2 invMass sum_u m_u |x_u - x1_u|^2   s.t.  c_i = sum_j w_ij n_i . x_ij - thickness >= 0

by an augmented Lagrangian (mu = 1e3) with ONE backtracking gradient step per multiplier update
and at most 100 iterations.  That budget does not reach the minimiser: for one vertex the result
overshoots the closed-form projection by a fixed ratio (1.085 once the cap is hit, 1.161 when
the f-change test stops it after 3 iterations), the same ratio at every violation size.  So the
pins are what holds exactly for this algorithm, plus bounds against the closed form:

  exact    the displacement of every moved node is parallel to the impact normal; nodes outside
           the impacts, and obstacle nodes when deform = 0, do not move; a satisfied impact moves
           nothing; the result is linear in the violation (every iterate is homogeneous in it);
           a rigid translation of the scene translates the result; scaling every mass by k
           changes nothing (invMass normalises the objective); with deform = 1, heavier
           obstacle nodes move less and the vertex more (the gradient steps are not mass-scaled,
           so an infinitely heavy obstacle stalls the line search instead of reproducing deform 0)
  bounded  the displacement lies between 1 and 1.2 times the closed-form minimiser's
           x = x1 + W^-1 A^T lam, (A W^-1 A^T) lam = thickness - A x1 (W: node masses), for one
           vertex, for two impacts sharing the vertex, and for deform = 1 with light obstacle
           nodes; the constraints end satisfied.
"""
import numpy as np
import pytest

import oracle
from clothsimulator_amd import CS_MESH_CLOTH, CS_MESH_OBSTACLE
from clothsimulator_amd._cabi import IMPACT_DTYPE, MeshArrays, default_params, material_struct
from clothsimulator_amd.scenes import ClothSpec, SceneSpec

TH = 1e-4  # collision thickness (the default)


def response_scene(vz=-2e-3, tilt=(0.3, -0.2)):
    """A free cloth triangle (nodes 0-2) whose vertex 0 sits vz below a static obstacle triangle
    (nodes 3-5) in the plane z = tx x + ty y, and below a second one (nodes 6-8) in the mirror
    plane z = -tx x + ty y: the vertex lies over the valley line x = 0 where they cross, so
    projecting it onto either plane alone leaves it below the other."""
    tx, ty = tilt
    cloth = np.array([[0.0, 0.03, 0.0], [0.3, 0.0, 0.05], [0.0, 0.3, 0.05]])
    plane = lambda p, a, b: np.array([p[0], p[1], a * p[0] + b * p[1]])
    ob = np.array([plane(p, tx, ty) for p in ([-0.5, -0.5], [0.6, -0.4], [-0.3, 0.7])] +
                  [plane(p, -tx, ty) for p in ([-0.4, -0.6], [0.7, -0.3], [-0.5, 0.5])])
    cloth[0, 2] = tx * cloth[0, 0] + ty * cloth[0, 1] + vz
    spec = SceneSpec("response", default_params(), [ClothSpec(MeshArrays(cloth, np.array([[0, 1, 2]])),
                                                            material_struct())],
                     [MeshArrays(ob, np.array([[0, 1, 2], [3, 4, 5]]))], 1)
    return spec


def vf_impact(x, v, face, rec_nodes):
    """A vertex-face impact record of vertex v against triangle face (global ids): w = (1, -b),
    n = the face normal, b = barycentric coordinates of v's projection."""
    a, b, c = (x[k] for k in face)
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    T = np.column_stack([b - a, c - a])
    uv = np.linalg.lstsq(T, x[v] - a, rcond=None)[0]
    bary = np.array([1 - uv.sum(), uv[0], uv[1]])
    imp = np.zeros(1, dtype=IMPACT_DTYPE)
    imp["type"] = 0
    imp["nodes"] = [rec_nodes]
    imp["w"] = [[1.0, -bary[0], -bary[1], -bary[2]]]
    imp["n"] = [n]
    imp["t"] = 0.5
    return imp


def all_positions(o):
    xc = o.state(CS_MESH_CLOTH, 0, fp64=True)[0]
    xo = o.state(CS_MESH_OBSTACLE, 0, fp64=True)[0]
    return np.vstack([xc, xo])


def constraint(imp, x):
    return sum(imp["w"][0][j] * imp["n"][0] @ x[imp["nodes"][0][j]] for j in range(4)) - TH


def qp_active(x1, imps, movable, mass):
    """Minimiser with every constraint active: x = x1 + W^-1 A^T lam."""
    nodes = sorted(movable)
    col = {u: k for k, u in enumerate(nodes)}
    A = np.zeros((len(imps), 3 * len(nodes)))
    rhs = np.zeros(len(imps))
    for i, imp in enumerate(imps):
        c0 = constraint(imp, x1)
        rhs[i] = -c0
        for j in range(4):
            u = imp["nodes"][0][j]
            if u in col:
                A[i, 3 * col[u]:3 * col[u] + 3] += imp["w"][0][j] * imp["n"][0]
    Winv = np.repeat([1.0 / mass[u] for u in nodes], 3)
    lam = np.linalg.solve((A * Winv) @ A.T, rhs)
    assert (lam >= 0).all()  # every constraint active
    x = x1.copy()
    d = (Winv * (A.T @ lam)).reshape(-1, 3)
    for u in nodes:
        x[u] += d[col[u]]
    return x


def _one(vz=-2e-3, fp64=True, deform=0, M=1e3, shift=None, mass_scale=None):
    spec = response_scene(vz=vz)
    if shift is not None:
        for m in [c.mesh for c in spec.cloths] + spec.obstacles:
            m.x += np.asarray(shift, np.float32)
    if mass_scale is not None:
        spec.cloths[0].material.density *= mass_scale
    o = oracle.OracleScene(spec, fp64=fp64)
    x1 = all_positions(o)
    imp = vf_impact(x1, 0, (3, 4, 5), [0, 3, 4, 5])
    m0 = float(o.geometry()[2][0])
    it = o.impact_response(imp, deform, obstacle_mass=M)
    return x1, all_positions(o), imp, it, m0


@pytest.mark.parametrize("fp64", [False, True])
def test_one_vertex_moves_along_normal(fp64):
    x1, x, imp, it, _ = _one(fp64=fp64)
    viol = -constraint(imp, x1)
    d = x - x1
    n = imp["n"][0].astype(np.float32).astype(np.float64)  # the response reads n in FP32
    t = d[0] @ n
    print("one vertex (fp64=%d): %d iterations, violation %.4g m, moved %.4g m (ratio %.4f)" % (fp64, it, viol, t,
                                                                                           t / viol))
    assert np.linalg.norm(d[0] - t * n) <= 1e-5 * t  # along n
    assert 1.0 <= t / viol <= 1.2  # the closed-form projection moves it by exactly viol
    assert constraint(imp, x) >= 0
    assert np.array_equal(x[1:], x1[1:])  # nothing else moves


def test_response_is_linear_in_the_violation():
    base = None
    for k, vz in enumerate((-1e-3, -2e-3, -4e-3)):
        x1, x, imp, it, _ = _one(vz=vz)
        t = (x - x1)[0] @ imp["n"][0]
        v = -constraint(imp, x1)
        base = base or t / v
        assert t / v == pytest.approx(base, rel=1e-9)


def test_response_translation_and_mass_scale_invariance():
    x1, x, imp, _, _ = _one()
    d = x - x1
    x1s, xs, _, _, _ = _one(shift=(0.25, -0.5, 0.125))
    assert np.abs((xs - x1s) - d).max() <= 5e-6 * np.abs(d).max()  # inputs re-rounded to FP32 after the shift
    x1m, xm, _, _, _ = _one(mass_scale=7.0)
    assert np.abs((xm - x1m) - d).max() <= 1e-12


def test_two_impacts_sharing_the_vertex():
    spec = response_scene()
    o = oracle.OracleScene(spec, fp64=True)
    x1 = all_positions(o)
    i1 = vf_impact(x1, 0, (3, 4, 5), [0, 3, 4, 5])
    i2 = vf_impact(x1, 0, (6, 7, 8), [0, 6, 7, 8])
    assert constraint(i1, x1) < 0 and constraint(i2, x1) < 0
    imps = np.concatenate([i1, i2])
    mass = o.geometry()[2]
    want = qp_active(x1, [i1, i2], {0}, {0: float(mass[0])})
    o.impact_response(imps, 0)
    x = all_positions(o)
    dw, d = want[0] - x1[0], x[0] - x1[0]
    print("two impacts: closed form %s, result %s" % (dw, d))
    # the displacement stays in the plane of the two normals and near the minimiser's
    nrm = np.cross(i1["n"][0].astype(np.float32), i2["n"][0].astype(np.float32)).astype(np.float64)
    assert abs(d @ nrm) <= 1e-6 * np.linalg.norm(d)
    assert np.linalg.norm(d - dw) <= 0.2 * np.linalg.norm(dw)
    assert min(constraint(i1, x), constraint(i2, x)) >= 0
    assert np.array_equal(x[1:], x1[1:])


def test_deform1_weights_nodes_by_obstacle_mass():
    x1, _, _, _, m0 = _one()
    M = 3.0 * m0  # obstacle nodes three times heavier than the vertex: everything moves
    x1, x, imp, it, _ = _one(deform=1, M=M)
    want = qp_active(x1, [imp], {0, 3, 4, 5}, {0: m0, 3: M, 4: M, 5: M})
    n = imp["n"][0].astype(np.float32).astype(np.float64)  # the response reads n in FP32
    d, dw = x - x1, want - x1
    print("deform 1: closed form along n %s, result %s" % (dw[[0, 3, 4, 5]] @ n, d[[0, 3, 4, 5]] @ n))
    for u in (0, 3, 4, 5):
        assert np.linalg.norm(np.cross(d[u], n)) <= 1e-6 * np.abs(d).max()  # along n
    assert d[0] @ n > 0 and (d[[3, 4, 5]] @ n < 0).all()  # vertex and face pushed apart
    assert np.abs(d - dw).max() <= 0.2 * np.abs(dw).max()
    assert constraint(imp, x) >= 0
    assert np.array_equal(x[[1, 2, 6, 7, 8]], x1[[1, 2, 6, 7, 8]])
    # heavier obstacle nodes move less and leave more of the move to the vertex
    moves = []
    for k in (1.0, 10.0, 100.0):
        x1k, xk, _, _, _ = _one(deform=1, M=k * m0)
        dk = xk - x1k
        moves.append((dk[0] @ n, np.abs(dk[[3, 4, 5]] @ n).max()))
    assert moves[0][0] < moves[1][0] < moves[2][0]
    assert moves[0][1] > moves[1][1] > moves[2][1] > 0


def test_satisfied_impact_moves_nothing():
    spec = response_scene(vz=+5e-3)
    o = oracle.OracleScene(spec, fp64=True)
    x1 = all_positions(o)
    imp = vf_impact(x1, 0, (3, 4, 5), [0, 3, 4, 5])
    assert constraint(imp, x1) > 0
    o.impact_response(imp, 0)
    assert np.abs(all_positions(o) - x1).max() <= 1e-12
