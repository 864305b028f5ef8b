"""Pins of the learned preconditioner's CPU restatement (oracle/lpc.py), no GPU:
its hand-written backward against central finite differences of its own
forward, the Kaporin trace and gradient against dense matrix algebra, the
untrained network reducing to block Jacobi, G^T G being the symmetric
positive definite operator it claims, and bfloat16 rounding on known values."""
import numpy as np
import pytest

from oracle import lpc


def small_system(n=14, seed=0):
    """A random SPD block system on a banded node graph (pattern includes the diagonal)."""
    rng = np.random.default_rng(seed)
    adj = {i: {i} for i in range(n)}
    for i in range(n):
        for j in (i + 1, i + 4):
            if j < n:
                adj[i].add(j)
                adj[j].add(i)
    rp, ci = [0], []
    for i in range(n):
        ci += sorted(adj[i])
        rp.append(len(ci))
    rp, ci = np.array(rp), np.array(ci)
    row = np.repeat(np.arange(n), np.diff(rp))
    vals = np.zeros((len(ci), 3, 3))
    off = {}
    for i, j in zip(row, ci):
        if i < j:
            off[(i, j)] = rng.normal(0, 0.3, (3, 3))
    for k, (i, j) in enumerate(zip(row, ci)):
        if i < j:
            vals[k] = off[(i, j)]
        elif i > j:
            vals[k] = off[(j, i)].T
        else:
            m = rng.normal(0, 0.3, (3, 3))
            vals[k] = 3.0 * np.eye(3) + m @ m.T
    Minv = np.linalg.inv(vals[row == ci])
    return lpc.Graph(rp, ci, vals, Minv), vals, rp, ci


def weights_with_outputs(seed=1, scale=0.05):
    W = lpc.unpack(lpc.init_weights(seed))
    rng = np.random.default_rng(seed + 10)
    for k in ("Dn2", "bdn2", "De2", "bde2"):
        W[k] = rng.normal(0, scale, W[k].shape)
    return lpc.pack(W)


def dense(rp, ci, blocks, n):
    A = np.zeros((3 * n, 3 * n))
    row = np.repeat(np.arange(n), np.diff(rp))
    for k, (i, j) in enumerate(zip(row, ci)):
        A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blocks[k]
    return A


def test_weight_layout():
    assert lpc.COUNT == 99360
    w = lpc.init_weights(0)
    assert np.array_equal(lpc.pack(lpc.unpack(w)), w)


def test_backward_matches_finite_differences():
    g, _, _, _ = small_system()
    w = weights_with_outputs()
    loss, grad = lpc.loss_and_grad(w, g)
    rng = np.random.default_rng(2)
    # one coordinate in every parameter block, plus random ones
    idx, off = [], 0
    for name, shape in lpc.PARAMS:
        k = int(np.prod(shape))
        idx.append(off + int(rng.integers(k)))
        off += k
    idx += list(rng.choice(lpc.COUNT, 40, replace=False))
    eps = 1e-6
    for t in idx:
        wp, wm = w.copy(), w.copy()
        wp[t] += eps
        wm[t] -= eps
        fd = (lpc.loss_and_grad(wp, g)[0] - lpc.loss_and_grad(wm, g)[0]) / (2 * eps)
        assert abs(fd - grad[t]) <= 1e-6 * max(1.0, abs(fd)) + 2e-9, (t, fd, grad[t])


def test_kaporin_trace_and_gradient_against_dense_algebra():
    g, vals, rp, ci = small_system(seed=3)
    o, q, c = lpc.forward(weights_with_outputs(seed=4, scale=0.2), g)
    loss, T, do, dq = lpc.kaporin(g, o, q, c["Ah"])
    D, L = lpc.ghat_blocks(o, q)
    n = g.n
    Gh = np.zeros((3 * n, 3 * n))
    for i in range(n):
        Gh[3 * i:3 * i + 3, 3 * i:3 * i + 3] = D[i]
    for t, e in enumerate(g.low):
        i, j = g.src[e], g.dst[e]
        Gh[3 * i:3 * i + 3, 3 * j:3 * j + 3] = L[t]
    # Ahat = S A S^T with S = blockdiag(chol(Minv)^T): unit diagonal blocks
    S = np.zeros((3 * n, 3 * n))
    for i in range(n):
        S[3 * i:3 * i + 3, 3 * i:3 * i + 3] = c["S"][i]
    Ahat = S @ dense(rp, ci, vals, n) @ S.T
    assert np.allclose(Ahat[:3, :3], np.eye(3), atol=1e-12)
    Td = np.trace(Gh @ Ahat @ Gh.T)
    assert abs(T - Td) <= 1e-10 * Td
    # the loss is Kaporin's number of Ghat Ahat Ghat^T up to the constant log det(Ahat) / n
    M = Gh @ Ahat @ Gh.T
    kap = np.log(np.trace(M) / (3 * n)) - np.linalg.slogdet(M)[1] / (3 * n)
    assert abs(loss - (kap + np.linalg.slogdet(Ahat)[1] / (3 * n))) <= 1e-10
    # d trace / d Ghat = 2 Ghat Ahat on the pattern: dq is (2/T) of that for the lower blocks
    GA = 2.0 * Gh @ Ahat / Td
    for t, e in enumerate(g.low):
        i, j = g.src[e], g.dst[e]
        assert np.allclose(dq[t, :9], GA[3 * i:3 * i + 3, 3 * j:3 * j + 3].ravel(), rtol=1e-9, atol=1e-12)
    i = 5
    Gi = GA[3 * i:3 * i + 3, 3 * i:3 * i + 3]
    assert np.isclose(do[i, 3], Gi[1, 0]) and np.isclose(do[i, 4], Gi[2, 0]) and np.isclose(do[i, 5], Gi[2, 1])
    assert np.isclose(do[i, 0], Gi[0, 0] * np.exp(o[i, 0]) - 2.0 / (3 * n))


def test_untrained_network_is_block_jacobi():
    g, vals, rp, ci = small_system(seed=5)
    o, q, c = lpc.forward(lpc.init_weights(6), g)
    assert np.all(o == 0) and np.all(q == 0)
    Gd, Gl = lpc.factor(g, o, q, c["S"])
    r = np.random.default_rng(7).normal(0, 1, (g.n, 3))
    z = lpc.apply(g, Gd, Gl, r)
    assert np.allclose(z, np.einsum("nab,nb->na", g.Minv, r), rtol=1e-12, atol=1e-12)
    loss, T, _, _ = lpc.kaporin(g, o, q, c["Ah"])
    assert loss == 0.0 and T == 3 * g.n


def test_preconditioner_is_GtG():
    g, vals, rp, ci = small_system(seed=8)
    o, q, c = lpc.forward(weights_with_outputs(seed=9, scale=0.3), g)
    Gd, Gl = lpc.factor(g, o, q, c["S"])
    n = g.n
    G = np.zeros((3 * n, 3 * n))
    for i in range(n):
        G[3 * i:3 * i + 3, 3 * i:3 * i + 3] = Gd[i]
    for t, e in enumerate(g.low):
        i, j = g.src[e], g.dst[e]
        G[3 * i:3 * i + 3, 3 * j:3 * j + 3] = Gl[t]
    Mi = G.T @ G
    r = np.random.default_rng(10).normal(0, 1, (n, 3))
    assert np.allclose(lpc.apply(g, Gd, Gl, r).ravel(), Mi @ r.ravel(), rtol=1e-10, atol=1e-12)
    assert np.linalg.eigvalsh(Mi).min() > 0
    # block lower triangular: no block above the diagonal, lower-triangular diagonal blocks
    for i in range(n):
        assert np.all(G[3 * i:3 * i + 3, 3 * i + 3:] == 0)


@pytest.mark.parametrize("x,expect", [(1.0, 1.0), (1.0 + 2.0 ** -8, 1.0), (1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7),
                                      (1.0 + 2.0 ** -7 + 2.0 ** -8, 1.0 + 2.0 ** -6), (-2.5, -2.5),
                                      (3.0e38, 226 * 2.0 ** 120)])
def test_bf16_rounding(x, expect):
    got = float(lpc.bf16(np.array([x]))[0])
    assert abs(got - expect) <= 1e-4 * abs(expect)


def test_package_weight_layout_matches_oracle():
    from clothsimulator_amd import lpc_layout
    assert lpc_layout.PARAMS == lpc.PARAMS and lpc_layout.COUNT == lpc.COUNT
    assert np.array_equal(lpc_layout.init_weights(3), lpc.init_weights(3).astype(np.float32))


def test_solution_term_gradient_and_zero_at_eigenvectors():
    g, vals, rp, ci = small_system(seed=11)
    A = dense(rp, ci, vals, g.n)
    rng = np.random.default_rng(1)
    xs = rng.normal(0, 1, 3 * g.n)
    b = A @ xs
    w = weights_with_outputs(seed=3, scale=0.1)
    loss, grad = lpc.loss_and_grad(w, g, b=b, x=xs)
    assert loss > lpc.loss_and_grad(w, g)[0]   # the term is 1 - cos^2 >= 0
    for t in rng.choice(lpc.COUNT, 25, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[t] += 1e-6
        wm[t] -= 1e-6
        fd = (lpc.loss_and_grad(wp, g, b=b, x=xs)[0] - lpc.loss_and_grad(wm, g, b=b, x=xs)[0]) / 2e-6
        assert abs(fd - grad[t]) <= 1e-6 * max(1.0, abs(fd)) + 2e-9, (t, fd, grad[t])
    # untrained network (M = block Jacobi): an eigenvector x of M A gives M b = lambda x, cos = 1
    o, q, c = lpc.forward(lpc.init_weights(2), g)
    Minv = np.zeros_like(A)
    for i in range(g.n):
        Minv[3 * i:3 * i + 3, 3 * i:3 * i + 3] = g.Minv[i]
    lam, V = np.linalg.eig(Minv @ A)
    x = V[:, np.argmax(lam.real)].real
    L, _, _ = lpc.solution_term(g, o, q, c["S"], A @ x, x)
    assert abs(L) < 1e-10


def test_smoothing_term_gradient_and_dense_value():
    """L_sm = |(I - w G^T G A) v|_A^2 / |v|_A^2 against dense algebra, and its gradient against
    central finite differences; the device's probe hash restated (+-1 entries, reproducible)."""
    g, vals, rp, ci = small_system(seed=12)
    A = dense(rp, ci, vals, g.n)
    v = lpc.probe(g.n, 7)
    assert set(np.unique(v)) == {-1.0, 1.0} and np.array_equal(v, lpc.probe(g.n, 7))
    assert not np.array_equal(v, lpc.probe(g.n, 8))
    w = weights_with_outputs(seed=4, scale=0.1)
    o, q, c = lpc.forward(w, g)
    Gd, Gl = lpc.factor(g, o, q, c["S"])
    G = np.zeros_like(A)
    for i in range(g.n):
        G[3 * i:3 * i + 3, 3 * i:3 * i + 3] = Gd[i]
    for t, (i, j) in enumerate(zip(g.src[g.low], g.dst[g.low])):
        G[3 * i:3 * i + 3, 3 * j:3 * j + 3] = Gl[t]
    for omega in (1.0, 0.7):
        e = v.ravel() - omega * G.T @ G @ A @ v.ravel()
        dense_L = (e @ A @ e) / (v.ravel() @ A @ v.ravel())
        L, _, _ = lpc.smoothing_term(g, o, q, c["S"], v, omega)
        assert abs(L - dense_L) <= 1e-12 * max(1.0, abs(dense_L)), (omega, L, dense_L)
    base = lpc.loss_and_grad(w, g)[0]
    loss, grad = lpc.loss_and_grad(w, g, v=v, lam_s=0.5)
    assert abs(loss - base - 0.5 * lpc.smoothing_term(g, o, q, c["S"], v)[0]) <= 1e-12 * max(1.0, abs(loss))
    rng = np.random.default_rng(2)
    for t in rng.choice(lpc.COUNT, 25, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[t] += 1e-6
        wm[t] -= 1e-6
        fd = (lpc.loss_and_grad(wp, g, v=v, lam_s=0.5)[0] - lpc.loss_and_grad(wm, g, v=v, lam_s=0.5)[0]) / 2e-6
        assert abs(fd - grad[t]) <= 1e-6 * max(1.0, abs(fd)) + 2e-9, (t, fd, grad[t])
