"""Pins of the oracle's aggregation multigrid preconditioner against an
independent dense numpy restatement of its definition (no GPU needed):
the hierarchy sizes, the V(1,1) operator applied to random vectors, and
that the operator is symmetric positive definite and reduces PCG iterations."""
import numpy as np
import scipy.sparse as sp

import oracle
from clothsimulator_amd._cabi import CS_PRECOND_AMG
from clothsimulator_amd.scenes import hang

OMEGA, COARSEST = 0.6, 64


def face_graph(n, faces):
    adj = [set() for _ in range(n)]
    for f in faces:
        for a in f:
            for b in f:
                if a != b:
                    adj[a].add(b)
    return [sorted(s) for s in adj]


def aggregate(adj):
    n = len(adj)
    agg = [-1] * n
    na = 0
    for i in range(n):
        if agg[i] < 0 and all(agg[j] < 0 for j in adj[i]):
            agg[i] = na
            for j in adj[i]:
                agg[j] = na
            na += 1
    for i in range(n):
        if agg[i] < 0:
            a = next((agg[j] for j in adj[i] if agg[j] >= 0), -1)
            if a < 0:
                a = na
                na += 1
            agg[i] = a
    return np.array(agg), na


def hierarchy(n, faces):
    adj = face_graph(n, faces)
    sizes, aggs = [n], []
    while n > COARSEST:
        agg, na = aggregate(adj)
        if na >= n:
            break
        cadj = [set() for _ in range(na)]
        for i in range(n):
            for j in adj[i]:
                if agg[i] != agg[j]:
                    cadj[agg[i]].add(agg[j])
        adj = [sorted(s) for s in cadj]
        aggs.append(agg)
        sizes.append(na)
        n = na
    return sizes, aggs


def dense(rp, ci, vals, n):
    A = np.zeros((3 * n, 3 * n))
    for r in range(n):
        for k in range(rp[r], rp[r + 1]):
            A[3 * r:3 * r + 3, 3 * ci[k]:3 * ci[k] + 3] = vals[k]
    return A


def block_diag_inv(A):
    n = A.shape[0] // 3
    D = np.zeros_like(A)
    for i in range(n):
        D[3 * i:3 * i + 3, 3 * i:3 * i + 3] = np.linalg.inv(A[3 * i:3 * i + 3, 3 * i:3 * i + 3])
    return D


def vcycle_matrix(Afull, Astat, aggs):
    """The preconditioner as a dense matrix M^-1 (level 0: full system; coarse: Galerkin of static)."""
    Ps = []
    for agg in aggs:
        P1 = np.zeros((len(agg), agg.max() + 1))
        P1[np.arange(len(agg)), agg] = 1.0
        Ps.append(np.kron(P1, np.eye(3)))
    As = [Afull]
    S = Astat
    for P in Ps:
        S = P.T @ S @ P
        As.append(S)

    def op(l):
        A = As[l]
        if l == len(As) - 1:
            return np.linalg.inv(A)
        Dinv = block_diag_inv(A)
        I = np.eye(A.shape[0])
        P = Ps[l]
        C = P @ op(l + 1) @ P.T
        pre = OMEGA * Dinv                                    # x = w D^-1 b
        after_coarse = pre + C @ (I - A @ pre)                # x += P e, e = coarse(P^T (b - A x))
        return after_coarse + OMEGA * Dinv @ (I - A @ after_coarse)   # x += w D^-1 (b - A x)
    return op(0)


def test_oracle_amg_matches_dense_restatement():
    spec = hang(17)
    spec.params.precond = CS_PRECOND_AMG
    o = oracle.OracleScene(spec, fp64=True)
    rng = np.random.default_rng(0)
    x, x0, v = o.state(fp64=True)
    o.set_state(0, 0, x + rng.normal(0, 0.01, x.shape), x0, rng.normal(0, 0.1, x.shape))
    n = x.shape[0]
    sizes, aggs = hierarchy(n, spec.cloths[0].mesh.faces)
    # a 289-node sheet is coarsened once (289 -> 36 <= 64)
    assert o.amg_levels(0) == sizes and len(sizes) == 2
    rp, ci, vals, b = o.assemble(0)
    Afull = dense(rp, ci, vals, n)
    # the static system: the same state without proximity blocks (a flat sheet has none)
    Minv = vcycle_matrix(Afull, Afull, aggs)
    assert np.allclose(Minv, Minv.T, rtol=1e-8, atol=1e-10 * np.abs(Minv).max())
    assert np.linalg.eigvalsh(0.5 * (Minv + Minv.T)).min() > 0
    for _ in range(3):
        r = rng.normal(0, 1, (n, 3))
        z = o.amg_apply(r, 0)
        zd = (Minv @ r.ravel()).reshape(-1, 3)
        assert np.allclose(z, zd, rtol=1e-8, atol=1e-9 * np.abs(zd).max())


def test_oracle_amg_hierarchy_of_a_larger_sheet():
    spec = hang(65)
    o = oracle.OracleScene(spec, fp64=True)
    sizes, _ = hierarchy(65 * 65, spec.cloths[0].mesh.faces)
    assert o.amg_levels(0) == sizes and sizes[-1] <= COARSEST and len(sizes) == 3


def test_oracle_amg_with_contacts_matches_dense_restatement():
    """A perturbed sheet with self-proximities: level 0 is the full system, the coarse levels are
    Galerkin products of the static system plus the contact blocks' diagonal part."""
    from clothsimulator_amd._cabi import CS_FLAG_NO_PROXIMITY
    spec = hang(33)
    spec.params.precond = CS_PRECOND_AMG
    o = oracle.OracleScene(spec, fp64=True)
    rng = np.random.default_rng(0)
    x, x0, v = o.state(fp64=True)
    x = x + rng.normal(0, 0.01, x.shape)
    v = rng.normal(0, 0.1, x.shape)
    o.set_state(0, 0, x, x, v)
    assert len(o.proximities()) > 0
    n = x.shape[0]
    rp, ci, vals, b = o.assemble(0)
    p2 = hang(33).params
    p2.flags |= CS_FLAG_NO_PROXIMITY
    o2 = oracle.OracleScene(hang(33), fp64=True, params=p2)
    o2.set_state(0, 0, x, x, v)
    rs, cs_, vs, _ = o2.assemble(0)
    Af, As = dense(rp, ci, vals, n), dense(rs, cs_, vs, n)
    Cd = np.zeros_like(Af)
    for i in range(n):
        s = slice(3 * i, 3 * i + 3)
        Cd[s, s] = Af[s, s] - As[s, s]
    assert np.abs(Cd).max() > 0
    _, aggs = hierarchy(n, spec.cloths[0].mesh.faces)
    M = vcycle_matrix(Af, As + Cd, aggs)
    for _ in range(2):
        r = rng.normal(0, 1, (n, 3))
        z = o.amg_apply(r, 0)
        zd = (M @ r.ravel()).reshape(-1, 3)
        assert np.allclose(z, zd, rtol=1e-8, atol=1e-9 * np.abs(zd).max())
