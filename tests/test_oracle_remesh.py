"""The remeshing step: the oracle's sizing terms against closed forms, its enclosing disk against
brute force, its operators against invariants and hand-worked cases, and the library's host-side
operators (cs_remesh_mesh) against the oracle index for index (no device needed).

Reference: Cloth::remeshingStep (src/Cloth.cu:815-833), faceSizing and the operators'
predicates (src/RemeshingHelper.cu:41-485), Operator (src/Operator.cu:7-217)."""
import math

import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import remesh as R
from clothsimulator_amd import remesh_mesh, remeshing_struct


def grid(n, L=1.0, jitter=0.0, seed=0, z=None):
    xs = np.linspace(0, L, n)
    X, Y = np.meshgrid(xs, xs)
    uv = np.stack([X.ravel(), Y.ravel()], 1)
    if jitter:
        rng = np.random.default_rng(seed)
        inner = (uv[:, 0] > 0) & (uv[:, 0] < L) & (uv[:, 1] > 0) & (uv[:, 1] < L)
        uv[inner] += rng.uniform(-jitter, jitter, (inner.sum(), 2)) * L / (n - 1)
    x = np.concatenate([uv, np.zeros((n * n, 1)) if z is None else z(uv)[:, None]], 1)
    f = []
    for j in range(n - 1):
        for i in range(n - 1):
            a = j * n + i
            f.append([a, a + 1, a + n + 1])
            f.append([a, a + n + 1, a + n])
    return x, uv, np.array(f)


def cylinder(nc, nr, R0=0.3):
    """A tube: (nc + 1) x nr uv vertices over nc x nr nodes, the last uv column on the first node
    column, so the wrap edges are seams."""
    uv, vnode, x = [], [], []
    for j in range(nr):
        for i in range(nc + 1):
            uv.append([i / nc * 2 * np.pi * R0, j / (nr - 1)])
            vnode.append(j * nc + (i % nc))
    for j in range(nr):
        for i in range(nc):
            th = 2 * np.pi * i / nc
            x.append([R0 * np.cos(th), R0 * np.sin(th), j / (nr - 1)])
    f, W = [], nc + 1
    for j in range(nr - 1):
        for i in range(nc):
            a = j * W + i
            f.append([a, a + 1, a + W + 1])
            f.append([a, a + W + 1, a + W])
    return np.array(x), np.array(uv), np.array(vnode), np.array(f)


def random_sizing(nf, seed, hmin=0.04, hmax=0.3):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nf):
        th = rng.uniform(0, np.pi)
        lam = 1 / rng.uniform(hmin, hmax, 2) ** 2
        Q = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
        out.append((Q * lam) @ Q.T)
    return np.array(out, np.float32).reshape(-1, 4)


def iso(nf, h):
    return np.tile(np.array([1 / h ** 2, 0, 0, 1 / h ** 2], np.float32), (nf, 1))


PARAMS = dict(size=(0.05, 0.2), aspect_min=0.2)


def oracle_mesh(x, uv, vnode, fv, seed=1):
    en = R.sorted_edges(vnode, fv)
    pres = R.load_preserve(len(x), vnode, fv, en)
    v = np.random.default_rng(seed).normal(0, 0.1, np.shape(x))
    return R.Mesh(x, np.asarray(x) * 0.5, v, pres, uv, vnode, fv, en), (x, np.asarray(x) * 0.5, v, pres, uv, vnode,
                                                                         fv, en)


def check_valid(a, area0):
    """A valid manifold with the material area kept: every edge side names a face that holds both
    nodes in that orientation; uv areas positive and summing to area0; no repeated node pair."""
    fv, en, uv, vn = a["fv"], a["en"], a["uv"], a["vnode"]
    A = 0.0
    for f in fv:
        u0, u1, u2 = uv[f].astype(np.float64)
        ar = 0.5 * ((u1 - u0)[0] * (u2 - u0)[1] - (u1 - u0)[1] * (u2 - u0)[0])
        assert ar > 0
        A += ar
    assert abs(A - area0) <= 1e-5 * area0
    pairs = {}
    for f in fv:
        for i in range(3):
            p, q = int(vn[f[i]]), int(vn[f[(i + 1) % 3]])
            pairs.setdefault((min(p, q), max(p, q)), []).append((p, q))
    assert len(pairs) == len(en)
    for e in en:
        k = (min(e), max(e))
        assert k in pairs and 1 <= len(pairs[k]) <= 2
        if len(pairs[k]) == 2:  # two faces: opposite orientations
            assert pairs[k][0] == pairs[k][1][::-1]


# ---------------------------------------------------------------- sizing pins

def test_max_tensor_is_the_smallest_enclosing_disk():
    rng = np.random.default_rng(0)
    for trial in range(40):
        k = int(rng.integers(1, 6))
        Ms = []
        for _ in range(k):
            B = rng.normal(0, 1, (2, 2))
            Ms.append(B @ B.T + 0.1 * np.eye(2))
        S = R.max_tensor(Ms)
        o, r = R._disk(S)
        disks = [R._disk(M) for M in Ms]
        for o2, r2 in disks:  # encloses every disk
            assert r - r2 >= -1e-9 and (r - r2) ** 2 >= float(np.dot(o - o2, o - o2)) - 1e-9 * (1 + r * r)

        def rad(c):
            return max(np.linalg.norm(c - o2) + r2 for o2, r2 in disks)
        best = min((minimize(rad, o2, method="Nelder-Mead", options=dict(xatol=1e-12, fatol=1e-12, maxiter=4000))
                    for o2, _ in disks), key=lambda s: s.fun)
        assert r <= best.fun * (1 + 1e-6) + 1e-12, (trial, r, best.fun)
    # one tensor maps to itself; none gives zero
    M = np.array([[3.0, 0.5], [0.5, 1.0]])
    assert np.allclose(R.max_tensor([M]), M)
    assert np.all(R.max_tensor([np.zeros((2, 2))]) == 0)


def _geo(x, uv, fv):
    vnode = np.arange(len(uv))
    return R.Geometry(x, uv, vnode, fv, R.sorted_edges(vnode, fv))


def test_flat_rest_sheet_takes_the_largest_size():
    x, uv, fv = grid(6)
    r = R.Remeshing(refine_angle=0.3, refine_velocity=0.5, refine_compression=0.01, size=(0.01, 0.2), aspect_min=0.2)
    S = R.face_sizing(_geo(x, uv, fv), np.zeros_like(x), (np.zeros_like(x), np.zeros_like(x)), r)
    assert np.allclose(S, np.eye(2) / 0.2 ** 2, rtol=1e-12)


def test_curvature_of_a_cylinder():
    """x = (Rc sin(u/Rc), v, Rc (1 - cos(u/Rc))): interior faces see |curvature| 1/Rc across u and
    0 along v, from the dihedral angles and from the normal gradient alike."""
    Rc, n = 0.5, 21
    x, uv, fv = grid(n, z=lambda uv: 0 * uv[:, 0])
    x[:, 0] = Rc * np.sin(uv[:, 0] / Rc)
    x[:, 2] = Rc * (1 - np.cos(uv[:, 0] / Rc))
    g = _geo(x, uv, fv)
    h = 1.0 / (n - 1)
    for f in range(len(fv)):
        u = uv[fv[f]]
        if u.min() < 2 * h or u.max() > 1 - 2 * h:
            continue  # away from the boundary, where the one-sided angles are zero
        K = g.curvature(f)
        assert abs(abs(K[0, 0]) - 1 / Rc) <= 0.08 / Rc and abs(K[1, 1]) <= 0.08 / Rc and abs(K[0, 1]) <= 0.08 / Rc
        D = g.derivative(f, g.nrm)
        M = D.T @ D
        assert abs(M[0, 0] - 1 / Rc ** 2) <= 0.05 / Rc ** 2 and abs(M[1, 1]) <= 0.05 / Rc ** 2


def test_velocity_and_compression_metrics_closed_form():
    x, uv, fv = grid(5)
    # (refine_compression is given: left at its default of +inf the reference's compression
    # metric is inf * 0)
    r = R.Remeshing(refine_velocity=0.5, refine_compression=1.0, size=(1e-3, 10.0), aspect_min=1e-3)
    A = np.array([[0.3, -0.2], [0.1, 0.6], [0.4, 0.0]])  # v = A (u, v)
    v = uv @ A.T
    S = R.face_sizing(_geo(x, uv, fv), v, (np.zeros_like(x), np.zeros_like(x)), r)
    assert np.allclose(S, A.T @ A / 0.25, rtol=1e-10)
    # uniform compression by s along u: (1 - s^2) / c^2 across u, nothing along v
    s, c = 0.8, 0.05
    xc = x.copy()
    xc[:, 0] *= s
    M = R.compression_metric(np.diag([s * s - 1, 0.0]), np.zeros((2, 2)), R.Remeshing(refine_compression=c))
    assert np.allclose(M, np.diag([(1 - s * s) / c ** 2, 0.0]), rtol=1e-12)
    r2 = R.Remeshing(refine_compression=c, size=(1e-3, 10.0), aspect_min=1e-3)
    S2 = R.face_sizing(_geo(xc, uv, fv), np.zeros_like(x), (np.zeros_like(x), np.zeros_like(x)), r2)
    expect = np.diag([(1 - s * s) / c ** 2, 0.0])
    expect = np.maximum(expect, 1 / 10.0 ** 2)
    expect[0, 1] = expect[1, 0] = 0
    assert np.allclose(S2, expect, rtol=1e-9)
    # stretching refines nothing
    xs = x.copy()
    xs[:, 0] *= 1.2
    S3 = R.face_sizing(_geo(xs, uv, fv), np.zeros_like(x), (np.zeros_like(x), np.zeros_like(x)), r2)
    assert np.allclose(S3, np.eye(2) / 100.0, rtol=1e-12)


def test_obstacle_plane_and_metric_closed_form():
    """A sheet tilted as z = h0 + a u over a ground square: each node's plane is its projection,
    and the obstacle metric is diag(a^2, 0) / 3 * sum_i 1 / h_i^2."""
    h0, a = 0.05, 0.1
    x, uv, fv = grid(5, z=lambda uv: h0 + a * uv[:, 0])
    ground = (np.array([[-2.0, -2, 0], [3, -2, 0], [3, 3, 0], [-2, 3, 0]]), np.array([[0, 1, 2], [0, 2, 3]]))
    P, N = R.nearest_planes(x, [ground], thickness=0.5)
    assert np.allclose(P[:, :2], x[:, :2]) and np.allclose(P[:, 2], 0) and np.allclose(N, [0, 0, 1])
    r = R.Remeshing(refine_compression=1.0, size=(1e-3, 10.0), aspect_min=1e-3)
    S = R.face_sizing(_geo(x, uv, fv), np.zeros_like(x), (P, N), r)
    for f in range(len(fv)):
        hs = x[fv[f], 2]
        assert np.allclose(S[f], np.diag([a * a / 3 * np.sum(1 / hs ** 2), 1 / 100.0]), rtol=1e-9)
    # beyond the thickness: no plane
    _, N2 = R.nearest_planes(x + [0, 0, 1.0], [ground], thickness=0.5)
    assert not np.any(N2)


def test_nearest_point_off_the_face_is_negated_as_the_reference_computes_it():
    """On the edge branch the distance routine's weights are positive, so the reference's point
    -(w1 y0 + w2 y1 + w3 y2) is the negated closest point (src/RemeshingHelper.cu:17-29)."""
    tri = (np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]]), np.array([[0, 1, 2]]))
    x = np.array([[0.5, -0.1, 0.05]])  # closest point (0.5, 0, 0), on edge y0 y1
    P, N = R.nearest_planes(x, [tri], thickness=1.0)
    assert np.allclose(P[0], [-0.5, 0.0, 0.0])
    assert np.allclose(N[0], (x[0] - P[0]) / np.linalg.norm(x[0] - P[0]))


# ---------------------------------------------------------------- operator pins

def test_flip_of_a_long_diagonal():
    """A parallelogram split along its long diagonal flips to the short one."""
    uv = np.array([[0.0, 0], [2, 0], [3, 1], [1, 1]])
    x = np.concatenate([uv, np.zeros((4, 1))], 1)
    fv = np.array([[0, 1, 2], [0, 2, 3]])
    m, _ = oracle_mesh(x, uv, np.arange(4), fv)
    m.set_face_sizing(iso(2, 10.0))
    r = R.Remeshing(size=(0.01, 100.0), aspect_min=0.01)
    m.flip_edges(r)
    a = m.arrays()
    assert sorted((min(e), max(e)) for e in a["en"].tolist()) == [(0, 1), (0, 3), (1, 2), (1, 3), (2, 3)]
    assert a["en"].tolist()[-1] == [3, 1]  # (opposite on side 0, opposite on side 1), appended last
    assert a["fv"].tolist() == [[0, 1, 3], [2, 3, 1]]  # (v0, v3, v2), (v1, v2, v3) of the flipped edge


def test_split_of_the_longest_edge():
    """One triangle whose hypotenuse alone exceeds metric 1: a node at its midpoint (x, x0, v
    averaged), two faces."""
    uv = np.array([[0.0, 0], [1, 0], [0, 1]])
    x = np.array([[0.0, 0, 0], [1, 0, 0.2], [0, 1, 0.4]])
    m = R.Mesh(x, x + 1, x * 2, [1, 1, 1], uv, np.arange(3), [[0, 1, 2]], R.sorted_edges(np.arange(3), [[0, 1, 2]]))
    m.set_face_sizing(iso(1, 1.2))  # metric 1/1.2 on the legs, sqrt(2)/1.2 on the hypotenuse
    r = R.Remeshing(size=(0.01, 100.0), aspect_min=0.01)
    assert m.split_some(r)
    assert not m.split_some(r)
    a = m.arrays()
    assert np.array_equal(a["x"][3], np.float32([0.5, 0.5, 0.3]))
    assert np.array_equal(a["x0"][3], np.float32([1.5, 1.5, 1.3]))
    assert np.array_equal(a["v"][3], np.float32([1, 1, 0.6]))
    assert np.array_equal(a["uv"][3], np.float32([0.5, 0.5]))
    assert a["fv"].tolist() == [[1, 3, 0], [0, 3, 2]] and not a["preserve"][3]
    check_valid(a, 0.5)


def test_collapse_of_an_interior_node():
    """A node at the centre of a square fan, under a sizing that wants one cell: it collapses
    onto a corner and the square becomes two triangles."""
    uv = np.array([[0.0, 0], [1, 0], [1, 1], [0, 1], [0.5, 0.5]])
    x = np.concatenate([uv, np.zeros((5, 1))], 1)
    fv = np.array([[0, 1, 4], [1, 2, 4], [2, 3, 4], [3, 0, 4]])
    m, _ = oracle_mesh(x, uv, np.arange(5), fv)
    assert list(m.pres) == [True, True, True, True, False]
    m.set_face_sizing(iso(4, 2.0))
    r = R.Remeshing(size=(0.01, 100.0), aspect_min=0.01)
    m.collapse_edges(r)
    a = m.arrays()
    assert len(a["x"]) == 4 and len(a["fv"]) == 2 and len(a["en"]) == 5
    check_valid(a, 1.0)


@pytest.mark.parametrize("case", ["refine", "coarsen", "aniso0", "aniso1", "tube"])
def test_oracle_invariants(case):
    r = R.Remeshing(**PARAMS)
    if case == "tube":
        x, uv, vnode, fv = cylinder(8, 6)
        fs = random_sizing(len(fv), 11, 0.05, 0.4)
        area = 2 * np.pi * 0.3
    else:
        n, h = {"refine": (5, 0.1), "coarsen": (17, 0.3)}.get(case, (9, None))
        x, uv, fv = grid(n, jitter=0.3 if h is None else 0.0, seed=int(case[-1]) if h is None else 0)
        vnode = np.arange(len(uv))
        fs = iso(len(fv), h) if h else random_sizing(len(fv), int(case[-1]))
        area = 1.0
    m, _ = oracle_mesh(x, uv, vnode, fv)
    m.set_face_sizing(fs)
    m.flip_edges(r)
    m.split_edges(r)
    assert max(m.emetric(e) for e in m.edges) <= 1.0  # the split loop ends with every edge within 1
    check_valid(m.arrays(), area)
    m.collapse_edges(r)
    a = m.arrays()
    check_valid(a, area)
    # preserved nodes survive, at their place
    pres0 = R.load_preserve(len(x), vnode, fv, R.sorted_edges(vnode, fv))
    kept = {tuple(p) for p in a["x"][a["preserve"]]}
    assert all(tuple(np.float32(p)) in kept for p in np.asarray(x)[pres0])
    if case == "tube":
        assert len(a["uv"]) > len(a["x"])  # the seam is still there


# ---------------------------------------------------------------- the library against the oracle

@pytest.mark.parametrize("case", ["refine", "coarsen", "aniso0", "aniso1", "aniso2", "tube0", "tube1"])
def test_host_operators_match_oracle(case):
    r = R.Remeshing(**PARAMS)
    rr = remeshing_struct(**PARAMS)
    if case.startswith("tube"):
        x, uv, vnode, fv = cylinder(8, 6)
        fs = random_sizing(len(fv), 10 + int(case[-1]), 0.05, 0.4)
    else:
        n, h = {"refine": (5, 0.1), "coarsen": (17, 0.3)}.get(case, (9, None))
        x, uv, fv = grid(n, jitter=0.3 if h is None else 0.0, seed=int(case[-1]) if h is None else 0)
        vnode = np.arange(len(uv))
        fs = iso(len(fv), h) if h else random_sizing(len(fv), int(case[-1]))
    m, args = oracle_mesh(x, uv, vnode, fv)
    m.flip_count = 3
    m.remesh(fs, r)
    a = m.arrays()
    b, fc = remesh_mesh(rr, *args, fs, flip_count=3)
    for k in ("x", "x0", "v", "preserve", "uv", "vnode", "fv", "en", "sizing"):
        assert np.array_equal(a[k], b[k]), k
    assert fc == m.flip_count
    # node origins: input nodes keep their positions
    o = b["node_origin"]
    assert np.array_equal(b["x"][o >= 0], np.asarray(x, np.float32)[o[o >= 0]])
    print("%s: %d -> %d nodes, %d faces" % (case, len(x), len(a["x"]), len(a["fv"])))


def test_host_operators_reject_malformed_input():
    x, uv, fv = grid(3)
    vnode = np.arange(len(uv))
    en = R.sorted_edges(vnode, fv)
    rr = remeshing_struct(**PARAMS)
    from clothsimulator_amd import ClothError
    with pytest.raises(ClothError):
        remesh_mesh(rr, x, x, x, np.zeros(len(x)), uv, vnode, fv, en[:-1], iso(len(fv), 0.1))  # an edge missing
    bad = fv.copy()
    bad[0, 0] = len(uv)
    with pytest.raises(ClothError):
        remesh_mesh(rr, x, x, x, np.zeros(len(x)), uv, vnode, bad, en, iso(len(fv), 0.1))
    flipped = fv.copy()
    flipped[0] = flipped[0][::-1]  # one face against its neighbours' orientation
    with pytest.raises(ClothError):
        remesh_mesh(rr, x, x, x, np.zeros(len(x)), uv, vnode, flipped, en, iso(len(fv), 0.1))


def test_minimum_index_rounds_equal_the_greedy_pass():
    """The reference's device rounds (min index per face, chosen when minimal on all its faces)
    choose exactly what a greedy pass in index order chooses; the library takes the greedy pass."""
    rng = np.random.default_rng(5)
    for trial in range(200):
        n, nf = int(rng.integers(1, 60)), int(rng.integers(1, 40))
        faces = [sorted(set(rng.integers(0, nf, int(rng.integers(1, 7))).tolist())) for _ in range(n)]
        rounds = R.Mesh.independent(list(range(n)), lambda c: faces[c])
        taken, greedy = set(), []
        for c in range(n):
            if not taken & set(faces[c]):
                greedy.append(c)
                taken |= set(faces[c])
        assert rounds == greedy, trial
