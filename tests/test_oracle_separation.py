"""The separation step's CPU restatement (oracle/separation.py) against closed forms and its own
post-condition.  Reference: src/SeparationHelper.cu, src/Simulator.cu:697-799,
src/optimization/SeparationOptimization.cu."""
import numpy as np

from oracle import separation as S


def test_crossing_point_of_two_triangles():
    """A in z = 0, B in x = 0.5 crossing it: their shared segment is x = 0.5, z = 0, y in [0.2, 1.2]
    (B's cut) within A's [0, 1.5]; the midpoint (0.5, 0.7, 0) on both faces."""
    ya = np.array([[0.0, 0, 0], [2, 0, 0], [0, 2, 0]])
    yb = np.array([[0.5, 0.2, -1], [0.5, 0.2, 1], [0.5, 1.2, 0]])
    nA = np.array([0.0, 0, 1])
    nB = np.cross(yb[1] - yb[0], yb[2] - yb[0])
    nB /= np.linalg.norm(nB)
    b0, b1 = S.intersect_midpoint([0, 1, 2], ya, nA, [3, 4, 5], yb, nB)
    assert np.allclose(b0 @ ya, [0.5, 0.7, 0]) and np.allclose(b1 @ yb, [0.5, 0.7, 0])
    assert np.isclose(b0.sum(), 1) and np.isclose(b1.sum(), 1)
    # sharing a node, parallel, or apart: no crossing
    assert S.intersect_midpoint([0, 1, 2], ya, nA, [2, 4, 5], yb, nB) is None
    assert S.intersect_midpoint([0, 1, 2], ya, nA, [3, 4, 5], ya + [0, 0, 0.1], nA) is None
    assert S.intersect_midpoint([0, 1, 2], ya, nA, [3, 4, 5], yb + [5, 0, 0], nB) is None


def test_deepest_point_of_a_vertex_through_a_face():
    """B's vertex p is depth delta below A; with d from B's old point to A's pointing down (B was
    above), the deepest point is p against its projection on A, at depth delta."""
    delta = 0.01
    ya = np.array([[0.0, 0, 0], [2, 0, 0], [0, 2, 0]])
    yb = np.array([[0.5, 0.5, -delta], [1.5, 0.5, 0.5], [0.5, 1.5, 0.5]])
    nA = np.array([0.0, 0, 1])
    nB = np.cross(yb[1] - yb[0], yb[2] - yb[0])
    nB /= np.linalg.norm(nB)
    b0, b1 = np.zeros(3), np.zeros(3)
    depth = S.farthest_point(ya, nA, yb, nB, np.array([0.0, 0, -1]), b0, b1)
    assert np.isclose(depth, delta)
    assert np.allclose(b1, [1, 0, 0]) and np.allclose(b0 @ ya, [0.5, 0.5, 0])


def test_solve_against_the_weighted_projection():
    """One constraint n . (X_a - X_b) >= th on two free nodes of weights wa, wb: the minimiser of
    1/2 invW sum w |X - x1|^2 moves them apart along n by the violation in total, the heavier one
    less.  The penalty solve (mu = 1e3) ends within a few percent of the total; it stops on the
    objective's change (1e-10) before the split between the two reaches the weights' ratio, as
    the reference's solve does."""
    n = np.array([0.0, 0, 1])
    x1 = np.array([[0.0, 0, -0.004], [0.0, 0, 0.0]])
    wt = np.array([2.0, 1.0])
    th = 1e-3
    X = S.augmented_lagrangian(x1.copy(), x1, wt, [([0, 1], np.array([1.0, -1.0]), n)], [], th)
    viol = th - (x1[0, 2] - x1[1, 2])
    da, db = X[0, 2] - x1[0, 2], x1[1, 2] - X[1, 2]
    assert 0.97 * viol <= da + db <= 1.05 * viol
    assert 0 < da < db
    assert np.allclose(X[:, :2], 0)


def sheet(n, z, L=1.0, origin=(0.0, 0.0)):
    xs = np.linspace(0, L, n)
    X, Y = np.meshgrid(xs, xs)
    uv = np.stack([X.ravel(), Y.ravel()], 1)
    x = np.concatenate([uv + origin, np.full((n * n, 1), z)], 1)
    f = []
    for j in range(n - 1):
        for i in range(n - 1):
            a = j * n + i
            f += [[a, a + 1, a + n + 1], [a, a + n + 1, a + n]]
    return x, uv, np.array(f)


def crossing_sheets(n=5, dip=0.02):
    """Sheet 0 flat at z = 0; sheet 1 above it at z = 0.01 with its middle node pushed to z = -dip
    (through sheet 0).  Backup: both sheets flat (the pushed node back at 0.01)."""
    xa, ua, fa = sheet(n, 0.0)
    xb, ub, fb = sheet(n, 0.01, L=0.5, origin=(0.25, 0.25))
    back_b = xb.copy()
    xb[(n * n) // 2, 2] = -dip
    x = np.concatenate([xa, xb])
    faces = np.concatenate([fa, fb + len(xa)])
    mesh_of_face = np.array([0] * len(fa) + [1] * len(fb))
    face_uv = np.concatenate([ua[fa], ub[fb]])
    backup = [(ua[fa], xa[fa]), (ub[fb], back_b[fb])]
    return x, faces, face_uv, mesh_of_face, backup


def test_step_separates_crossing_sheets():
    x, faces, face_uv, mof, backup = crossing_sheets()
    assert S.find_intersections(x, faces, mof, 2)
    area = np.full(len(x), 0.01)
    xs, v, ok, solves = S.separation_step(x, x, x, area, faces, face_uv, mof, 2, len(x), backup, 1e-4, 0.005)
    assert ok and solves >= 1
    assert not S.find_intersections(xs, faces, mof, 2)
    # the pushed node went back up, sheet 0 moved down locally; far nodes stay
    k = 25 + 12
    assert xs[k, 2] > x[k, 2]
    moved = np.linalg.norm(xs - x, axis=1)
    assert moved.max() < 0.05 and np.count_nonzero(moved > 1e-9) < 20
