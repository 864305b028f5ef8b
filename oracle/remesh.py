"""CPU restatement of the reference's remeshing step (Cloth::remeshingStep, src/Cloth.cu:815-833;
driven once per frame by Simulator::remeshingStep, src/Simulator.cu:686-695) -- test
infrastructure only: the tests, smoke() and the bench's CPU column use it as the checker, the
library never calls it.

Two parts:

* Sizing, in FP64 (compared with the device's FP32 sizing by tolerance):
  nearest obstacle planes (Cloth::findNearestPlane, src/Cloth.cu:200-233, with
  checkNearestPoint / setNearestPlane, src/RemeshingHelper.cu:17-39, and the vertex-face distance
  of src/MathHelper.cu:71-119), the face sizing tensor (faceSizing, src/RemeshingHelper.cu:41-186:
  curvature, normal and velocity gradients, the compression metric, the obstacle metric, the
  smallest disk enclosing the five tensors' disks, clamped eigenvalues) and the area-weighted
  vertex sizing (Cloth::computeSizing, src/Cloth.cu:235-256).
* The edge operators, in FP32 with the library's own operation order, so that the resulting
  topology can be compared index for index: flips (shouldFlip, src/RemeshingHelper.cu:224-238),
  splits by the edge metric (:379-389), collapses (shouldCollapseGpu, :438-485), each applied to
  a maximal independent set chosen by the minimum-index rule of the reference's device path
  (src/RemeshingHelper.cu:249-338, 560-647, driven by src/Cloth.cu:286-349, 382-432, 567-646),
  with the operators of src/Operator.cu:7-217 / src/RemeshingHelper.cu:320-757 and the list
  update of Mesh::apply (src/Mesh.cu:138-165: removed entries leave in order, new ones are
  appended in operator order).

Pinned by tests/test_oracle_remesh.py: closed forms of the sizing terms (a cylinder's curvature,
a linear velocity field, uniform compression, a plane at a known height), the enclosing disk
against brute force, and the operators against invariants (a valid manifold, material area kept,
no edge above metric 1 after splitting, no flippable edge left) and hand-worked small cases.
"""
import math

import numpy as np

F = np.float32
INF = math.inf


class Remeshing:
    """The cloth's "remeshing" block (src/Remeshing.cu:3-19)."""

    def __init__(self, refine_angle=INF, refine_velocity=INF, refine_compression=INF, rib_stiffening=1.0,
                 size=None, aspect_min=-INF, flip_threshold=1e-2):
        self.refine_angle = refine_angle
        self.refine_velocity = refine_velocity
        self.refine_compression = refine_compression
        self.rib_stiffening = rib_stiffening
        self.size_min, self.size_max = (-INF, INF) if size is None else (size[0], size[1])
        self.aspect_min = aspect_min
        self.flip_threshold = flip_threshold


# ---------------------------------------------------------------- sizing (FP64)

def _vf_nearest(x, y0, y1, y2):
    """unsignedVertexFaceDistance (src/MathHelper.cu:71-119) with its weight writes as they are:
    on the edge branches each call writes its two weights whether or not it is the closer one.
    Returns (d, w[1..3])."""
    w = [1.0, 0.0, 0.0, 0.0]
    nt = np.cross(y1 - y0, y2 - y0)
    nt = nt / np.linalg.norm(nt)
    d = abs(float(np.dot(x - y0, nt)))
    b0 = float(np.dot(np.cross(y1 - x, y2 - x), nt))
    b1 = float(np.dot(np.cross(y2 - x, y0 - x), nt))
    b2 = float(np.dot(np.cross(y0 - x, y1 - x), nt))
    if b0 >= 0 and b1 >= 0 and b2 >= 0:
        s = b0 + b1 + b2
        return d, [-b0 / s, -b1 / s, -b2 / s]

    def ve(a, b):
        t = min(max(float(np.dot(x - a, b - a) / np.dot(b - a, b - a)), 0.0), 1.0)
        return float(np.linalg.norm(x - (a + t * (b - a)))), 1.0 - t, t

    d = INF
    if b0 < 0:
        dt, w[2], w[3] = ve(y1, y2)
        if dt < d:
            d, w[1] = dt, 0.0
    if b1 < 0:
        dt, w[3], w[1] = ve(y2, y0)
        if dt < d:
            d, w[2] = dt, 0.0
    if b2 < 0:
        dt, w[1], w[2] = ve(y0, y1)
        if dt < d:
            d, w[3] = dt, 0.0
    return d, w[1:]


def nearest_planes(x, obstacles, thickness):
    """Per cloth node: the plane (p, n) through the nearest obstacle point closer than
    `thickness`, n = (x - p) / |x - p|, or n = 0 (src/Cloth.cu:200-233).  The point is
    -(w1 y0 + w2 y1 + w3 y2) with the weights the distance routine leaves (positive on the edge
    branches, so there the point is the negated closest point, as the reference computes it).
    obstacles: list of (y (m x 3), faces (k x 3)).  Faces are visited in order; the first of
    equal distances wins."""
    x = np.asarray(x, np.float64)
    P = np.zeros_like(x)
    N = np.zeros_like(x)
    for i in range(len(x)):
        best, pt = thickness, x[i].copy()
        for y, f in obstacles:
            y = np.asarray(y, np.float64)
            for a, b, c in f:
                d, w = _vf_nearest(x[i], y[a], y[b], y[c])
                if d < best:
                    best, pt = d, -(w[0] * y[a] + w[1] * y[b] + w[2] * y[c])
        n = x[i] - pt
        if float(np.dot(n, n)) > 1e-8:
            P[i], N[i] = pt, n / np.linalg.norm(n)
    return P, N


def nearest_plane_candidates(x, obstacles, thickness, rtol=1e-5):
    """Per node, every plane (p, n) a nearest-point search could return: those of the faces within
    rtol (relative) of the smallest distance below `thickness`.  Faces that share the closest
    edge or vertex tie up to rounding, and the point the reference keeps depends on which of them
    its traversal meets first (with the edge-branch weights, the tied faces give different
    points)."""
    x = np.asarray(x, np.float64)
    out = []
    for i in range(len(x)):
        ds = []
        for y, f in obstacles:
            y = np.asarray(y, np.float64)
            for a, b, c in f:
                d, w = _vf_nearest(x[i], y[a], y[b], y[c])
                if d < thickness:
                    ds.append((d, -(w[0] * y[a] + w[1] * y[b] + w[2] * y[c])))
        c = []
        if ds:
            dmin = min(d for d, _ in ds)
            for d, pt in ds:
                n = x[i] - pt
                if d <= dmin * (1 + rtol) and float(np.dot(n, n)) > 1e-8:
                    c.append((pt, n / np.linalg.norm(n)))
        out.append(c)
    return out


def _sym_fn(A, fn):
    l, Q = np.linalg.eigh(0.5 * (A + A.T))
    return (Q * fn(l)) @ Q.T


def compression_metric(G, S2, r):
    """compressionMetric (src/RemeshingHelper.cu:66-70)."""
    P = np.array([[S2[1, 1], -S2[1, 0]], [-S2[0, 1], S2[0, 0]]])
    c2 = r.refine_compression ** 2
    D = G.T @ G - 4.0 * c2 * P * r.rib_stiffening
    sq = _sym_fn(D, lambda l: np.where(l >= 0, np.sqrt(np.abs(l)), -np.sqrt(np.abs(l))))
    return _sym_fn(-G + sq, lambda l: np.maximum(l, 0.0)) / (2.0 * c2)


def _disk(M):
    return np.array([0.5 * (M[0, 0] - M[1, 1]), 0.5 * (M[0, 1] + M[1, 0])]), 0.5 * (M[0, 0] + M[1, 1])


def _encloses(o, r, o2, r2):
    return r >= r2 and (r - r2) ** 2 >= float(np.dot(o - o2, o - o2))


def _circ2(d0, d1):
    (o0, r0), (o1, r1) = d0, d1
    d = float(np.linalg.norm(o0 - o1))
    r = 0.5 * (r0 + d + r1)
    t = (r - r0) / d
    return o0 + t * (o1 - o0), r


def _circ3(d0, d1, d2):
    """The disk tangent to three disks from outside (src/RemeshingHelper.cu:94-126)."""
    (x0, y0), r0 = d0
    (x1, y1), r1 = d1
    (x2, y2), r2 = d2
    v11, v12 = 2 * x1 - 2 * x0, 2 * y1 - 2 * y0
    v13 = x0 * x0 - x1 * x1 + y0 * y0 - y1 * y1 - r0 * r0 + r1 * r1
    v14 = 2 * r1 - 2 * r0
    v21, v22 = 2 * x2 - 2 * x1, 2 * y2 - 2 * y1
    v23 = x1 * x1 - x2 * x2 + y1 * y1 - y2 * y2 - r1 * r1 + r2 * r2
    v24 = 2 * r2 - 2 * r1
    w12, w13, w14 = v12 / v11, v13 / v11, v14 / v11
    w22, w23, w24 = v22 / v21 - w12, v23 / v21 - w13, v24 / v21 - w14
    P, Q = -w23 / w22, w24 / w22
    M, N = -w12 * P - w13, w14 - w12 * Q
    a = N * N + Q * Q - 1.0
    b = 2 * M * N - 2 * N * x0 + 2 * P * Q - 2 * Q * y0 + 2 * r0
    c = x0 * x0 + M * M - 2 * M * x0 + P * P + y0 * y0 - 2 * P * y0 - r0 * r0
    rs = (-b - math.sqrt(b * b - 4 * a * c)) / (2 * a)
    return np.array([M + N * rs, P + Q * rs]), rs


def max_tensor(Ms):
    """maxTensor (src/RemeshingHelper.cu:128-157): each tensor with a non-zero trace is the disk
    (center ((a - d) / 2, (b + c) / 2), radius (a + d) / 2); the smallest disk enclosing them
    (incremental construction, the reference's order) maps back to a tensor."""
    d = [_disk(M) for M in Ms if np.trace(M) != 0.0]
    if not d:
        return np.zeros((2, 2))
    o, r = d[0]
    for i in range(1, len(d)):
        if not _encloses(o, r, *d[i]):
            o, r = d[i]
            for j in range(i):
                if not _encloses(o, r, *d[j]):
                    o, r = _circ2(d[i], d[j])
                    for k in range(j):
                        if not _encloses(o, r, *d[k]):
                            o, r = _circ3(d[i], d[j], d[k])
    return np.array([[r + o[0], o[1]], [o[1], r - o[0]]])


class Geometry:
    """Face and node geometry of one cloth (src/Face.cu:109-131, src/Mesh.cu:182-214,
    src/Edge.cu:114-124) in FP64."""

    def __init__(self, x, uv, vnode, fv, en):
        self.x = np.asarray(x, np.float64)
        self.uv = np.asarray(uv, np.float64)
        self.vnode = np.asarray(vnode)
        self.fv = np.asarray(fv)
        fn = self.vnode[self.fv]
        self.fn = fn
        y0, y1, y2 = self.x[fn[:, 0]], self.x[fn[:, 1]], self.x[fn[:, 2]]
        c = np.cross(y1 - y0, y2 - y0)
        self.fnrm = c / np.linalg.norm(c, axis=1)[:, None]
        u0, u1, u2 = self.uv[self.fv[:, 0]], self.uv[self.fv[:, 1]], self.uv[self.fv[:, 2]]
        d1, d2 = u1 - u0, u2 - u0
        self.farea = 0.5 * np.abs(d1[:, 0] * d2[:, 1] - d1[:, 1] * d2[:, 0])
        Dm = np.stack([d1, d2], axis=2)  # columns d1, d2
        self.finv = np.linalg.inv(Dm)
        n = np.zeros_like(self.x)
        for f in range(len(fn)):
            for i in range(3):
                a = self.x[fn[f, i]]
                e0, e1 = self.x[fn[f, (i + 1) % 3]] - a, self.x[fn[f, (i + 2) % 3]] - a
                n[fn[f, i]] += np.cross(e0, e1) / (np.dot(e0, e0) * np.dot(e1, e1))
        self.nrm = n / np.linalg.norm(n, axis=1)[:, None]
        # face edges: the edge joining face vertices i and i + 1
        key = {}
        for e, (a, b) in enumerate(np.asarray(en)):
            key[(int(min(a, b)), int(max(a, b)))] = e
        self.en = np.asarray(en)
        m = Mesh(np.zeros((len(x), 3)), np.zeros((len(x), 3)), np.zeros((len(x), 3)), np.zeros(len(x)), uv, vnode,
                 fv, en)
        self.eadj = np.array(m.eadj, dtype=np.int64).reshape(-1, 2)
        self.fe = np.array([[key[(int(min(fn[f, i], fn[f, (i + 1) % 3])), int(max(fn[f, i], fn[f, (i + 1) % 3])))]
                             for i in range(3)] for f in range(len(fn))])

    def angle(self, e):
        a0, a1 = self.eadj[e]
        if a0 < 0 or a1 < 0:
            return 0.0
        t = self.x[self.en[e, 0]] - self.x[self.en[e, 1]]
        t = t / np.linalg.norm(t)
        n0, n1 = self.fnrm[a0], self.fnrm[a1]
        return math.atan2(float(np.dot(t, np.cross(n0, n1))), float(np.dot(n0, n1)))

    def curvature(self, f):
        S = np.zeros((2, 2))
        for i in range(3):
            e = self.uv[self.fv[f, (i + 1) % 3]] - self.uv[self.fv[f, i]]
            t = np.array([-e[1], e[0]]) / np.linalg.norm(e)
            S -= 0.5 * self.angle(self.fe[f, i]) * np.linalg.norm(e) * np.outer(t, t)
        return S / self.farea[f]

    def derivative(self, f, q):
        n = self.fn[f]
        return np.stack([q[n[1]] - q[n[0]], q[n[2]] - q[n[0]]], axis=1) @ self.finv[f]


def face_sizing(geo, v, planes, r):
    """faceSizing for every face (F x 2 x 2), src/RemeshingHelper.cu:159-186."""
    v = np.asarray(v, np.float64)
    P, Nn = planes
    out = np.zeros((len(geo.fv), 2, 2))
    for f in range(len(geo.fv)):
        Sw1 = geo.curvature(f)
        M0 = Sw1.T @ Sw1 / r.refine_angle ** 2
        Sw2 = geo.derivative(f, geo.nrm)
        M1 = Sw2.T @ Sw2 / r.refine_angle ** 2
        V = geo.derivative(f, v)
        M2 = V.T @ V / r.refine_velocity ** 2
        Fx = geo.derivative(f, geo.x)
        M3 = compression_metric(Fx.T @ Fx - np.eye(2), Sw2.T @ Sw2, r)
        M4 = np.zeros((2, 2))
        nodes = geo.fn[f]
        for i in range(3):
            n = Nn[nodes[i]]
            if not np.any(n):
                continue
            h = [float(np.dot(geo.x[nodes[j]] - P[nodes[i]], n)) for j in range(3)]
            dh = geo.finv[f].T @ np.array([h[1] - h[0], h[2] - h[0]])
            M4 += np.outer(dh, dh) / h[i] ** 2
        M4 /= 3.0
        S = max_tensor([M0, M1, M2, M3, M4])
        lo = 1.0 / r.size_max ** 2
        hi = 1.0 / r.size_min ** 2

        def clampl(l):
            l = np.minimum(np.maximum(l, lo), hi)
            return np.maximum(l, l.max() * r.aspect_min ** 2)
        out[f] = _sym_fn(S, clampl)
    return out


def vertex_sizing(uv, fv, fsizing, nverts):
    """Area-weighted vertex sizing (src/Cloth.cu:235-256), FP64."""
    uv = np.asarray(uv, np.float64)
    S = np.zeros((nverts, 2, 2))
    A = np.zeros(nverts)
    for f, (a, b, c) in enumerate(fv):
        d1, d2 = uv[b] - uv[a], uv[c] - uv[a]
        ar = 0.5 * abs(d1[0] * d2[1] - d1[1] * d2[0])
        for k in (a, b, c):
            A[k] += ar
            S[k] += ar * fsizing[f]
    return S / A[:, None, None]


# ---------------------------------------------------------------- operators (FP32)

def _m2v(M, v):
    return (M[0] * v[0] + M[1] * v[1], M[2] * v[0] + M[3] * v[1])


def _dot2(a, b):
    return a[0] * b[0] + a[1] * b[1]


def _sub2(a, b):
    return (a[0] - b[0], a[1] - b[1])


class Mesh:
    """One cloth's topology with the reference's element records (Node, Vertex, Edge, Face) as
    integer ids into pools, and the four element lists whose order the operators depend on.

    Edges: nodes (n0, n1); per side s the face's vertices (v_s0 at n0, v_s1 at n1), the opposite
    vertex and the face (Edge::initialize, src/Edge.cu:13-31)."""

    def __init__(self, x, x0, v, preserve, uv, vnode, fv, en):
        self.X = [tuple(F(c) for c in p) for p in np.asarray(x, np.float32)]
        self.X0 = [tuple(F(c) for c in p) for p in np.asarray(x0, np.float32)]
        self.V = [tuple(F(c) for c in p) for p in np.asarray(v, np.float32)]
        self.pres = [bool(p) for p in preserve]
        self.U = [tuple(F(c) for c in p) for p in np.asarray(uv, np.float32)]
        self.vnode = [int(k) for k in vnode]
        self.S = [(F(0), F(0), F(0), F(0))] * len(self.U)
        self.fv = [list(map(int, f)) for f in fv]
        self.fe = [[-1, -1, -1] for _ in self.fv]
        self.en = [list(map(int, e)) for e in en]
        self.ev = [[[-1, -1], [-1, -1]] for _ in self.en]
        self.eopp = [[-1, -1] for _ in self.en]
        self.eadj = [[-1, -1] for _ in self.en]
        key = {(min(a, b), max(a, b)): e for e, (a, b) in enumerate(self.en)}
        for f, fv3 in enumerate(self.fv):
            for i in range(3):
                a, b = self.vnode[fv3[i]], self.vnode[fv3[(i + 1) % 3]]
                e = key[(min(a, b), max(a, b))]
                self.fe[f][i] = e
                self.init_edge(e, fv3[(i + 2) % 3], f)
        self.nodes = list(range(len(self.X)))
        self.verts = list(range(len(self.U)))
        self.edges = list(range(len(self.en)))
        self.faces = list(range(len(self.fv)))
        self.flip_count = 0  # the reference keeps this in a function-level static (src/Cloth.cu:321)

    # -- records
    def init_edge(self, e, opp, f):
        n0, n1 = self.en[e]
        fv3 = self.fv[f]
        for i in range(3):
            a, b = fv3[i], fv3[(i + 1) % 3]
            if self.vnode[a] == n0 and self.vnode[b] == n1:
                self.ev[e][0] = [a, b]
                self.eopp[e][0] = opp
                self.eadj[e][0] = f
                return
            if self.vnode[a] == n1 and self.vnode[b] == n0:
                self.ev[e][1] = [b, a]
                self.eopp[e][1] = opp
                self.eadj[e][1] = f
                return

    def new_edge(self, n0, n1):
        self.en.append([n0, n1])
        self.ev.append([[-1, -1], [-1, -1]])
        self.eopp.append([-1, -1])
        self.eadj.append([-1, -1])
        return len(self.en) - 1

    def new_face(self, a, b, c):
        self.fv.append([a, b, c])
        self.fe.append([-1, -1, -1])
        return len(self.fv) - 1

    def new_vertex(self, u, S, node):
        self.U.append(u)
        self.S.append(S)
        self.vnode.append(node)
        return len(self.U) - 1

    def find_edge(self, f, a, b):
        fv3 = self.fv[f]
        for i in range(3):
            p, q = fv3[i], fv3[(i + 1) % 3]
            if (p == a and q == b) or (q == a and p == b):
                return self.fe[f][i]
        return -1

    def boundary(self, e):
        return self.eadj[e][0] < 0 or self.eadj[e][1] < 0

    def seam(self, e):
        return (not self.boundary(e)) and (self.ev[e][0][0] != self.ev[e][1][0] or self.ev[e][0][1] != self.ev[e][1][1])

    def farea(self, f):
        a, b, c = (self.U[k] for k in self.fv[f])
        d1, d2 = _sub2(b, a), _sub2(c, a)
        return F(0.5) * abs(d1[0] * d2[1] - d1[1] * d2[0])

    # -- sizing
    def set_face_sizing(self, fs):
        """Vertex sizing = area-weighted mean of the faces' sizing, accumulated in face-list order
        (src/Cloth.cu:235-256); fs: per face of the current face list, 4 floats."""
        acc = {k: [F(0), F(0), F(0), F(0)] for k in self.verts}
        ar = {k: F(0) for k in self.verts}
        fs = np.asarray(fs, np.float32).reshape(-1, 4)
        for i, f in enumerate(self.faces):
            a = self.farea(f)
            for k in self.fv[f]:
                ar[k] = ar[k] + a
                for c in range(4):
                    acc[k][c] = acc[k][c] + a * fs[i][c]
        for k in self.verts:
            inv = F(1) / ar[k]
            self.S[k] = tuple(acc[k][c] * inv for c in range(4))

    # -- predicates
    def should_flip(self, e, r):
        if self.boundary(e) or self.seam(e):
            return False
        v0, v1 = self.ev[e][0][0], self.ev[e][1][1]
        v2, v3 = self.eopp[e]
        x, y, z, w = self.U[v0], self.U[v1], self.U[v2], self.U[v3]
        S0, S1, S2, S3 = self.S[v0], self.S[v1], self.S[v2], self.S[v3]
        M = tuple(F(0.25) * (((S0[c] + S1[c]) + S2[c]) + S3[c]) for c in range(4))
        a0, a1 = self.farea(self.eadj[e][0]), self.farea(self.eadj[e][1])
        lhs = a1 * _dot2(_sub2(x, z), _m2v(M, _sub2(y, z))) + a0 * _dot2(_sub2(y, w), _m2v(M, _sub2(x, w)))
        return lhs < -F(r.flip_threshold) * (a0 + a1)

    def vmetric(self, a, b):
        if a < 0 or b < 0:
            return F(0)
        du = _sub2(self.U[a], self.U[b])
        return np.sqrt(F(0.5) * (_dot2(du, _m2v(self.S[a], du)) + _dot2(du, _m2v(self.S[b], du))))

    def emetric(self, e):
        return max(self.vmetric(*self.ev[e][0]), self.vmetric(*self.ev[e][1]))

    def should_collapse(self, e, side, ring_e, ring_f, r):
        n0, n1 = self.en[e][side], self.en[e][1 - side]
        if self.pres[n0]:
            return False
        if any(self.boundary(k) or self.seam(k) for k in ring_e[n0]) and not (self.boundary(e) or self.seam(e)):
            return False
        v00, v01 = self.ev[e][0][side], self.ev[e][0][1 - side]
        v10, v11 = self.ev[e][1][side], self.ev[e][1][1 - side]
        for f in ring_f[n0]:
            vs = list(self.fv[f])
            if any(self.vnode[k] == n1 for k in vs):
                continue
            vs = [v01 if k == v00 else (v11 if k == v10 else k) for k in vs]
            u0, u1, u2 = (self.U[k] for k in vs)
            d1, d2 = _sub2(u1, u0), _sub2(u2, u0)
            a = F(0.5) * (d1[0] * d2[1] - d1[1] * d2[0])

            def nrm(p):
                return np.sqrt(_dot2(p, p))
            p = (nrm(_sub2(u0, u1)) + nrm(_sub2(u1, u2))) + nrm(_sub2(u2, u0))
            aspect = F(12) * np.sqrt(F(3)) * a / (p * p)
            if a < F(1e-6) or aspect < F(r.aspect_min):
                return False
            for j in range(3):
                if vs[j] != v01 and vs[j] != v11 and self.vmetric(vs[(j + 1) % 3], vs[(j + 2) % 3]) > F(0.9):
                    return False
        return True

    # -- independent sets (minimum-index rule)
    @staticmethod
    def independent(cands, faces_of):
        n = len(cands)
        removed = {}
        for c in cands:
            for f in faces_of(c):
                removed[f] = False
        sel = [False] * n
        new = n
        while True:
            num = new
            mi = {}
            for i, c in enumerate(cands):
                fs = faces_of(c)
                if not any(removed[f] for f in fs):
                    for f in fs:
                        mi[f] = min(mi.get(f, n), i)
            for i, c in enumerate(cands):
                fs = faces_of(c)
                if all((not removed[f]) and mi.get(f, n) == i for f in fs):
                    sel[i] = True
                    for f in fs:
                        removed[f] = True
            new = sel.count(False)
            if not num > new:
                break
        return [c for i, c in enumerate(cands) if sel[i]]

    def apply(self, rn=(), an=(), rv=(), av=(), re=(), ae=(), rf=(), af=()):
        for lst, rem, add in ((self.nodes, rn, an), (self.verts, rv, av), (self.edges, re, ae), (self.faces, rf, af)):
            rs = set(rem)
            lst[:] = [k for k in lst if k not in rs] + list(add)

    # -- operators
    def flip_some(self, r):
        cands = [e for e in self.edges if self.should_flip(e, r)]
        ind = self.independent(cands, lambda e: [f for f in self.eadj[e] if f >= 0])
        if not ind or len(ind) == self.flip_count:
            return False
        self.flip_count = len(ind)
        ae, re, af, rf = [], [], [], []
        for e in ind:
            v0, v1 = self.ev[e][0][0], self.ev[e][1][1]
            v2, v3 = self.eopp[e]
            f0, f1 = self.eadj[e]
            e0, e1 = self.find_edge(f0, v1, v2), self.find_edge(f0, v2, v0)
            e2, e3 = self.find_edge(f1, v0, v3), self.find_edge(f1, v3, v1)
            ne = self.new_edge(self.vnode[v2], self.vnode[v3])
            nf0 = self.new_face(v0, v3, v2)
            nf1 = self.new_face(v1, v2, v3)
            self.init_edge(ne, v0, nf0)
            self.init_edge(ne, v1, nf1)
            self.fe[nf0] = [e2, ne, e1]
            self.fe[nf1] = [e0, ne, e3]
            self.init_edge(e0, v3, nf1)
            self.init_edge(e1, v3, nf0)
            self.init_edge(e2, v2, nf0)
            self.init_edge(e3, v2, nf1)
            ae.append(ne)
            re.append(e)
            af += [nf0, nf1]
            rf += [f0, f1]
        self.apply(re=re, ae=ae, rf=rf, af=af)
        return True

    def flip_edges(self, r):
        for _ in range(2 * len(self.edges)):
            if not self.flip_some(r):
                break

    def split_some(self, r):
        cands = [(self.emetric(e), e) for e in self.edges]
        cands = [c for c in cands if c[0] > F(1)]
        cands.sort(key=lambda c: -c[0])  # stable: equal metrics keep list order
        ind = self.independent([e for _, e in cands], lambda e: [f for f in self.eadj[e] if f >= 0])
        if not ind:
            return False
        an, av, ae, re, af, rf = [], [], [], [], [], []
        for e in ind:
            n0, n1 = self.en[e]
            half = F(0.5)
            nn = len(self.X)
            self.X.append(tuple(half * (self.X[n0][c] + self.X[n1][c]) for c in range(3)))
            self.X0.append(tuple(half * (self.X0[n0][c] + self.X0[n1][c]) for c in range(3)))
            self.V.append(tuple(half * (self.V[n0][c] + self.V[n1][c]) for c in range(3)))
            self.pres.append(False)
            ne = [self.new_edge(nn, n0), self.new_edge(nn, n1)]
            an.append(nn)
            re.append(e)

            def mid(a, b):
                return (tuple(half * (self.U[a][c] + self.U[b][c]) for c in range(2)),
                        tuple(half * (self.S[a][c] + self.S[b][c]) for c in range(4)))
            if self.seam(e):
                nv = [self.new_vertex(*mid(*self.ev[e][0]), nn), self.new_vertex(*mid(*self.ev[e][1]), nn)]
                av += nv
            else:
                j = 0 if self.eopp[e][0] >= 0 else 1
                k = self.new_vertex(*mid(*self.ev[e][j]), nn)
                nv = [k, k]
                av.append(k)
            added_e, added_f = [ne[0], ne[1]], []
            for j in range(2):
                if self.eopp[e][j] < 0:
                    continue
                v0, v1 = self.ev[e][j][j], self.ev[e][j][1 - j]
                v2 = self.eopp[e][j]
                f = self.eadj[e][j]
                e0, e1 = self.find_edge(f, v1, v2), self.find_edge(f, v2, v0)
                w = nv[j]
                ne0, ne1 = ne[j], ne[1 - j]
                ne2 = self.new_edge(nn, self.vnode[v2])
                nf0 = self.new_face(v0, w, v2)
                nf1 = self.new_face(v2, w, v1)
                self.init_edge(ne0, v2, nf0)
                self.init_edge(ne1, v2, nf1)
                self.init_edge(ne2, v0, nf0)
                self.init_edge(ne2, v1, nf1)
                self.fe[nf0] = [ne0, ne2, e1]
                self.fe[nf1] = [ne2, ne1, e0]
                self.init_edge(e0, w, nf1)
                self.init_edge(e1, w, nf0)
                added_e.append(ne2)
                added_f += [nf0, nf1]
                rf.append(f)
            ae += added_e
            af += added_f
        self.apply(an=an, av=av, re=re, ae=ae, rf=rf, af=af)
        return True

    def rings(self):
        ring_e = {n: [] for n in self.nodes}
        for e in self.edges:
            for n in self.en[e]:
                ring_e[n].append(e)
        ring_f = {n: [] for n in self.nodes}
        for f in self.faces:
            for k in self.fv[f]:
                ring_f[self.vnode[k]].append(f)
        return ring_e, ring_f

    def collapse_some(self, r):
        ring_e, ring_f = self.rings()
        cands = []
        for e in self.edges:
            if self.should_collapse(e, 0, ring_e, ring_f, r):
                cands.append((e, 0))
            elif self.should_collapse(e, 1, ring_e, ring_f, r):
                cands.append((e, 1))
        ind = self.independent(cands, lambda c: ring_f[self.en[c[0]][c[1]]])
        if not ind:
            return False
        rn, rv, re, rf = [], [], [], []
        for e, side in ind:
            n0, n1 = self.en[e][side], self.en[e][1 - side]
            v00, v01 = self.ev[e][0][side], self.ev[e][0][1 - side]
            v10, v11 = self.ev[e][1][side], self.ev[e][1][1 - side]
            rn.append(n0)
            re.append(e)
            if self.seam(e):
                rv += [v00, v10]
            else:
                rv.append(v00 if v00 >= 0 else v10)
            for j in range(2):
                if self.eopp[e][j] < 0:
                    continue
                v0, v1 = self.ev[e][j][side], self.ev[e][j][1 - side]
                v2 = self.eopp[e][j]
                f = self.eadj[e][j]
                e0, e1 = self.find_edge(f, v1, v2), self.find_edge(f, v2, v0)
                if not self.boundary(e1):
                    s = 0 if self.eopp[e1][0] != v1 else 1
                    v, g = self.eopp[e1][s], self.eadj[e1][s]
                    for i in range(3):  # Face::replaceEdge: the first match
                        if self.fe[g][i] == e1:
                            self.fe[g][i] = e0
                            break
                else:
                    v, g = -1, -1
                # replaceOpposite / replaceAdjacent: the first side that matches
                for s in range(2):
                    if self.eopp[e0][s] == v0:
                        self.eopp[e0][s] = v
                        break
                for s in range(2):
                    if self.eadj[e0][s] == f:
                        self.eadj[e0][s] = g
                        break
                re.append(e1)
                rf.append(f)
            for k in ring_e[n0]:
                if k == e:
                    continue
                for s in range(2):
                    if self.en[k][s] == n0:
                        self.en[k][s] = n1
                        break
                for a, b in ((v00, v01), (v10, v11)):
                    for s in range(2):  # Edge::replaceVertex: the first match on each side
                        for t in range(2):
                            if self.ev[k][s][t] == a:
                                self.ev[k][s][t] = b
                                break
            adj = self.eadj[e]
            for g in ring_f[n0]:
                if g == adj[0] or g == adj[1]:
                    continue
                fv3 = self.fv[g]
                opp = -1
                for i in range(3):
                    if self.vnode[fv3[i]] == n0:
                        opp = self.fe[g][(i + 1) % 3]
                        break
                for a, b in ((v00, v01), (v10, v11)):
                    for s in range(2):
                        if self.eopp[opp][s] == a:
                            self.eopp[opp][s] = b
                            break
                for a, b in ((v00, v01), (v10, v11)):
                    for i in range(3):
                        if self.fv[g][i] == a:
                            self.fv[g][i] = b
                            break
        self.apply(rn=rn, rv=[k for k in rv if k >= 0], re=re, rf=rf)
        return True

    def split_edges(self, r):
        while self.split_some(r):
            self.flip_edges(r)

    def collapse_edges(self, r):
        while self.collapse_some(r):
            self.flip_edges(r)

    def remesh(self, face_sizing, r):
        """flip, split, collapse (src/Cloth.cu:815-833) with the given per-face sizing."""
        self.set_face_sizing(face_sizing)
        self.flip_edges(r)
        self.split_edges(r)
        self.collapse_edges(r)

    # -- output
    def arrays(self):
        """Compact arrays in list order: x, x0, v, preserve, uv, vnode, fv (vertex ids), en."""
        ni = {n: i for i, n in enumerate(self.nodes)}
        vi = {k: i for i, k in enumerate(self.verts)}
        x = np.array([self.X[n] for n in self.nodes], np.float32).reshape(-1, 3)
        x0 = np.array([self.X0[n] for n in self.nodes], np.float32).reshape(-1, 3)
        v = np.array([self.V[n] for n in self.nodes], np.float32).reshape(-1, 3)
        pres = np.array([self.pres[n] for n in self.nodes], bool)
        uv = np.array([self.U[k] for k in self.verts], np.float32).reshape(-1, 2)
        vnode = np.array([ni[self.vnode[k]] for k in self.verts], np.int32)
        fv = np.array([[vi[k] for k in self.fv[f]] for f in self.faces], np.int32).reshape(-1, 3)
        en = np.array([[ni[n] for n in self.en[e]] for e in self.edges], np.int32).reshape(-1, 2)
        S = np.array([self.S[k] for k in self.verts], np.float32).reshape(-1, 4)
        return dict(x=x, x0=x0, v=v, preserve=pres, uv=uv, vnode=vnode, fv=fv, en=en, sizing=S)


def sorted_edges(vnode, fv):
    """The edge list a mesh loads with: the distinct (min, max) node pairs, ascending
    (src/MeshHelper.cu:34-102)."""
    vnode = np.asarray(vnode)
    s = set()
    for f in fv:
        for i in range(3):
            a, b = int(vnode[f[i]]), int(vnode[f[(i + 1) % 3]])
            s.add((min(a, b), max(a, b)))
    return np.array(sorted(s), np.int32).reshape(-1, 2)


def load_preserve(nn, vnode, fv, en, handles=()):
    """Nodes preserved at load: those of boundary and seam edges (src/MeshHelper.cu:93-102),
    and handle nodes (src/Cloth.cu:650-662)."""
    m = Mesh(np.zeros((nn, 3)), np.zeros((nn, 3)), np.zeros((nn, 3)), np.zeros(nn), np.zeros((len(vnode), 2)),
             vnode, fv, en)
    p = np.zeros(nn, bool)
    for e in m.edges:
        if m.boundary(e) or m.seam(e):
            p[m.en[e][0]] = p[m.en[e][1]] = True
    p[list(handles)] = True
    return p
