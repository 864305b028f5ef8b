"""CPU restatement of the reference's separation step after remeshing (Simulator::separationStep,
src/Simulator.cu:697-799; findIntersections / checkIntersection, src/Simulator.cu:489-562;
src/SeparationHelper.cu; SeparationOptimization, src/optimization/SeparationOptimization.cu:3-158,
solved by Optimization::solve, src/optimization/Optimization.cu:98-181) -- test infrastructure
only: the tests use it as the checker, the library never calls it.

FP64 throughout.  Candidate pairs are every pair of faces of the traversal's mesh pairs (cloth
self, cloth x earlier cloth, cloth x obstacle): a crossing pair's boxes always overlap, so the
crossing set is the reference's; the order of the constraints (traversal order there, face
order here) changes only rounding.

Pinned by tests/test_oracle_separation.py: the crossing point of two triangles against the
closed-form line of their planes, the deepest point against a vertex pushed through a face by a
known depth, the solve against the closed-form weighted projection for one constraint, and the
whole step leaving no crossing pair.
"""
import numpy as np


def _sign(x):
    return -1 if x < 0 else 1


def _face_plane(y, p0, n):
    """facePlaneIntersection (src/SeparationHelper.cu:7-35)."""
    h = [float(np.dot(y[i] - p0, n)) for i in range(3)]
    ss = sum(_sign(v) for v in h)
    if ss in (-3, 3):
        return None
    v0 = -1
    for i in range(3):
        if _sign(h[i]) == -ss:
            v0 = i
    v1, v2 = (v0 + 1) % 3, (v0 + 2) % 3
    t0 = h[v0] / (h[v0] - h[v1])
    t1 = h[v0] / (h[v0] - h[v2])
    b0, b1 = np.zeros(3), np.zeros(3)
    b0[v0], b0[v1] = 1 - t0, t0
    b1[v0], b1[v2] = 1 - t1, t1
    return b0, b1


def intersect_midpoint(na, ya, nA, nb, yb, nB):
    """checkIntersectionMidpoint (src/SeparationHelper.cu:37-70): barycentric points (b0 on face a,
    b1 on face b) at the middle of the segment the two faces share, or None."""
    if set(na) & set(nb):
        return None
    c = np.cross(nA, nB)
    if float(np.dot(c, c)) < 1e-12:
        return None
    r0, r1 = _face_plane(ya, yb[0], nB), _face_plane(yb, ya[0], nA)
    if r0 is None or r1 is None:
        return None
    (b00, b01), (b10, b11) = r0, r1
    ac = np.abs(c)
    ax = 0 if ac[0] > ac[1] and ac[0] > ac[2] else (1 if ac[1] > ac[2] else 2)
    a00, a01 = (b00 @ ya)[ax], (b01 @ ya)[ax]
    a10, a11 = (b10 @ yb)[ax], (b11 @ yb)[ax]
    lo = max(min(a00, a01), min(a10, a11))
    hi = min(max(a00, a01), max(a10, a11))
    if lo > hi:
        return None
    mid = 0.5 * (lo + hi)
    b0 = b00 if a00 == a01 else b00 + (mid - a00) / (a01 - a00) * (b01 - b00)
    b1 = b10 if a10 == a11 else b10 + (mid - a10) / (a11 - a10) * (b11 - b10)
    return b0, b1


def _clear_vf(ya, yb, nB, d, best, bA, bB):
    """clearVertexFaceDistance (src/SeparationHelper.cu:150-182)."""
    x0, x1, x2 = yb
    for i in range(3):
        x = ya[i]
        h = float(np.dot(x - x0, nB))
        dh = float(np.dot(d, nB))
        if h * dh >= 0:
            continue
        a0 = float(np.dot(x2 - x1, np.cross(x - x1, d)))
        a1 = float(np.dot(x0 - x2, np.cross(x - x2, d)))
        a2 = float(np.dot(x1 - x0, np.cross(x - x0, d)))
        if a0 <= 0 or a1 <= 0 or a2 <= 0:
            continue
        dist = -h / dh
        if dist > best[0]:
            best[0] = dist
            bA[:] = 0
            bA[i] = 1
            bB[:] = np.array([a0, a1, a2]) / (a0 + a1 + a2)


def _clear_ee(ya, yb, d, best, b0, b1):
    """clearEdgeEdgeDistance (src/SeparationHelper.cu:184-212)."""
    for i in range(3):
        for j in range(3):
            x00, x01, x10, x11 = ya[i], ya[(i + 1) % 3], yb[j], yb[(j + 1) % 3]
            n = np.cross((x01 - x00) / np.linalg.norm(x01 - x00), (x11 - x10) / np.linalg.norm(x11 - x10))
            h = float(np.dot(x00 - x10, n))
            dh = float(np.dot(d, n))
            if h * dh >= 0:
                continue
            a00 = float(np.dot(x01 - x10, np.cross(x11 - x10, d)))
            a01 = float(np.dot(x11 - x10, np.cross(x00 - x10, d)))
            a10 = float(np.dot(x01 - x00, np.cross(x11 - x00, d)))
            a11 = float(np.dot(x10 - x00, np.cross(x01 - x00, d)))
            if a00 * a01 <= 0 or a10 * a11 <= 0:
                continue
            dist = -h / dh
            if dist > best[0]:
                best[0] = dist
                b0[:] = 0
                b0[i], b0[(i + 1) % 3] = a00 / (a00 + a01), a01 / (a00 + a01)
                b1[:] = 0
                b1[j], b1[(j + 1) % 3] = a10 / (a10 + a11), a11 / (a10 + a11)


def farthest_point(ya, nA, yb, nB, d, b0, b1):
    """farthestPoint (src/SeparationHelper.cu:214-219): the deepest vertex-face or edge-edge point
    pair along d, if any (b0, b1 updated in place); returns its depth (0: none)."""
    best = [0.0]
    _clear_vf(ya, yb, nB, d, best, b0, b1)
    _clear_vf(yb, ya, nA, -d, best, b1, b0)
    _clear_ee(ya, yb, d, best, b0, b1)
    return best[0]


def backup_position(faces_u, faces_x, u):
    """Simulator::oldPosition (src/Simulator.cu:489-495): the first backup face whose barycentric
    coordinates of u are within (-1e-6, -1e-6, -1e-5), or None."""
    for U, X in zip(faces_u, faces_x):
        M = np.stack([U[0] - U[2], U[1] - U[2]], axis=1)
        b = np.linalg.solve(M, u - U[2])
        bb = np.array([b[0], b[1], 1 - b[0] - b[1]])
        if bb[0] >= -1e-6 and bb[1] >= -1e-6 and bb[2] >= -1e-5:
            return bb @ X
    return None


def augmented_lagrangian(X0, x1, wt, cons, fixed, thickness):
    """Optimization::solve (src/optimization/Optimization.cu:98-181) for
    min 1/2 invW sum_u wt_u |X_u - x1_u|^2  s.t.  sum_j w_j n . p_j - thickness >= 0, where
    p_j = X[slot] for a free slot, fixed[slot] otherwise.  cons: list of (slots, w, n), slots index
    X (>= 0) or fixed (< 0 as -1 - k).  invW = mean(1 / wt).  Returns X."""
    X = np.array(X0, np.float64)
    nN = len(X)
    invW = float(np.mean(1.0 / wt))
    mu = 1e3
    lam = np.zeros(len(cons))

    def cval(Xp, i):
        slots, w, n = cons[i]
        return sum(w[j] * np.dot(n, Xp[s] if s >= 0 else fixed[-1 - s]) for j, s in enumerate(slots)) - thickness

    def value(Xp):
        f = 0.5 * invW * float(np.sum(wt * np.sum((Xp - x1) ** 2, axis=1)))
        coef = np.array([min(cval(Xp, i) + lam[i] / mu, 0.0) for i in range(len(cons))])
        return f + float(np.sum(0.5 * mu * coef ** 2)), coef

    s = 1e-3 * nN
    for _ in range(100):
        f, coef = value(X)
        g = invW * wt[:, None] * (X - x1)
        for i, (slots, w, n) in enumerate(cons):
            if coef[i] != 0:
                for j, sl in enumerate(slots):
                    if sl >= 0:
                        g[sl] += mu * coef[i] * w[j] * n
        n2 = float(np.sum(g * g))
        s /= 0.7
        while True:
            s *= 0.7
            Xn = X - s * g
            ft, _ = value(Xn)
            if not (ft >= f - 0.5 * s * n2 and s >= 1e-12 and abs(f - ft) >= 1e-10):
                break
        if s < 1e-12 or abs(f - ft) < 1e-10:
            break
        X = Xn
        lam = np.minimum(lam + mu * np.array([cval(X, i) for i in range(len(cons))]), 0.0)
    return X


def face_normals(x, faces):
    y = x[faces]
    c = np.cross(y[:, 1] - y[:, 0], y[:, 2] - y[:, 0])
    return c / np.linalg.norm(c, axis=1)[:, None]


def find_intersections(x, faces, mesh_of_face, n_cloth_meshes):
    """Every crossing face pair of the traversal's mesh pairs, (f0, f1, b0, b1), f0 the face of the
    later (or same) cloth, obstacles second."""
    fn = face_normals(x, faces)
    out = []
    F = len(faces)
    for a in range(F):
        ma = mesh_of_face[a]
        if ma >= n_cloth_meshes:
            continue
        for b in range(F):
            mb = mesh_of_face[b]
            if mb > ma and mb < n_cloth_meshes:
                continue  # cloth pairs once: a later cloth against an earlier one
            if mb == ma and b <= a:
                continue
            r = intersect_midpoint(faces[a], x[faces[a]], fn[a], faces[b], x[faces[b]], fn[b])
            if r is not None:
                out.append((a, b, r[0], r[1]))
    return out


def separation_step(x, x0, x1, area, faces, face_uv, mesh_of_face, n_cloth_meshes, n_cloth_nodes, backup, thickness,
                    dt):
    """The whole step on global arrays (cloth nodes first).  backup: per cloth mesh (U (F, 3, 2),
    X (F, 3, 3)).  Returns (x, v of the cloth nodes, success, solves)."""
    x = np.array(x, np.float64)
    obstacle_area = 1e3
    solves = 0
    success = False
    for deform in range(2):
        cons_raw = []
        for _ in range(100):
            new = find_intersections(x, faces, mesh_of_face, n_cloth_meshes)
            if not new:
                success = True
                break
            fn = face_normals(x, faces)
            for f0, f1, b0, b1 in new:
                def old(f, b):
                    if mesh_of_face[f] >= n_cloth_meshes:
                        return b @ x[faces[f]]
                    u = b @ face_uv[f]
                    p = backup_position(*backup[mesh_of_face[f]], u)
                    return b @ x[faces[f]] if p is None else p
                d = old(f0, b0) - old(f1, b1)
                d = d / np.linalg.norm(d)
                b0, b1 = b0.copy(), b1.copy()
                farthest_point(x[faces[f0]], fn[f0], x[faces[f1]], fn[f1], d, b0, b1)
                cons_raw.append((list(faces[f0]) + list(faces[f1]), np.concatenate([b0, -b1]), d))
            # nodes of the solve: free ones (deform 0) or all (deform 1), sorted
            nodes = sorted({int(k) for nd, _, _ in cons_raw for k in nd if deform == 1 or k < n_cloth_nodes})
            idx = {k: i for i, k in enumerate(nodes)}
            fixed = []
            cons = []
            for nd, w, d in cons_raw:
                slots = []
                for k in nd:
                    if k in idx:
                        slots.append(idx[k])
                    else:
                        fixed.append(x[k].copy())
                        slots.append(-len(fixed))
                cons.append((slots, w, d))
            wt = np.array([area[k] if k < n_cloth_nodes else obstacle_area for k in nodes])
            X = augmented_lagrangian(x[nodes], x1[nodes], wt, cons, fixed, thickness)
            x[nodes] = X
            solves += 1
            if deform == 1:
                obstacle_area *= 0.5
        if success:
            break
    v = (x[:n_cloth_nodes] - x0[:n_cloth_nodes]) / dt
    return x, v, success, solves
