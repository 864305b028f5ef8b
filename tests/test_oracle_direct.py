"""Pins of the oracle's nested-dissection multifrontal LL^T (oracle/or_direct.hpp), the direct
solve behind the CPU reference column at large sizes (the reference's CPU path factors every
step's system with Eigen::SimplicialLLT, src/Cloth.cu:686-688).

It must give the solution of the same system as the RCM-ordered banded Cholesky and as a dense
numpy solve, on systems with contact blocks, with several disconnected cloths (a forest of
dissection trees) and with an irregular (random) block pattern; and it must refuse a matrix that
is not positive definite."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from clothsimulator_amd import scenes


def _system(spec):
    # the FP64 twin: its assembled system is symmetric to rounding (the FP32 one carries
    # asymmetries of FP32 size in its contact blocks, which a Cholesky, reading one triangle,
    # does not see)
    o = oracle.OracleScene(spec, fp64=True)
    rp, ci, vals, b = o.assemble()
    return rp, ci, vals.reshape(-1, 9), b


def _dense(rp, ci, v, b):
    A = sp.bsr_matrix((v.reshape(-1, 3, 3), ci, rp), shape=(len(b), len(b))).toarray()
    return A, np.linalg.solve(A, b)


@pytest.mark.parametrize("name", ["hang17", "fold17", "crumple33"])
def test_nd_matches_banded_and_dense(name):
    spec = {"hang17": lambda: scenes.hang(17), "fold17": lambda: scenes.fold(17),
            "crumple33": lambda: scenes.crumple(33)}[name]()
    rp, ci, v, b = _system(spec)
    A, xd = _dense(rp, ci, v, b)
    x1, _, _ = oracle.solve_bsr(rp, ci, v, b, direct=1)
    x2, _, _ = oracle.solve_bsr(rp, ci, v, b, direct=2)
    cond = np.linalg.cond(A)
    e1 = np.linalg.norm(x1 - xd) / np.linalg.norm(xd)
    e2 = np.linalg.norm(x2 - xd) / np.linalg.norm(xd)
    print("%s: cond %.3g, banded vs dense %.3g, nested dissection vs dense %.3g" % (name, cond, e1, e2))
    # backward stable (normwise backward error at rounding level) and as close to the dense
    # solution as the banded factorisation of the same system
    berr = np.linalg.norm(b - A @ x2) / (np.linalg.norm(A, 2) * np.linalg.norm(x2) + np.linalg.norm(b))
    print("  backward error %.3g" % berr)
    assert berr <= 1e-14
    assert e2 <= 2 * e1 + 1e-14


def test_nd_two_disconnected_cloths():
    """Two cloths in one system (block-diagonal: two dissection trees)."""
    rpa, cia, va, ba = _system(scenes.hang(9))
    rpb, cib, vb, bb = _system(scenes.crumple(13))
    na = len(rpa) - 1
    rp = np.concatenate([rpa, rpb[1:] + rpa[-1]])
    ci = np.concatenate([cia, cib + na])
    v = np.vstack([va, vb])
    b = np.concatenate([ba, bb])
    A, xd = _dense(rp, ci, v, b)
    x2, _, _ = oracle.solve_bsr(rp, ci, v, b, direct=2)
    assert np.linalg.norm(x2 - xd) <= 1e-10 * np.linalg.norm(xd)


def test_nd_random_pattern():
    """An irregular symmetric block pattern (random graph over 400 nodes), SPD by diagonal
    dominance."""
    rng = np.random.default_rng(3)
    n = 400
    edges = set()
    for i in range(n):
        for j in rng.choice(n, 4, replace=False):
            if i != j:
                edges.add((min(i, j), max(i, j)))
    rows = [[i] for i in range(n)]
    for i, j in edges:
        rows[i].append(j)
        rows[j].append(i)
    blocks = {}
    for i, j in edges:
        B = rng.normal(0, 1, (3, 3))
        blocks[(i, j)] = B
        blocks[(j, i)] = B.T
    rp, ci, v = [0], [], []
    for i in range(n):
        cols = sorted(rows[i])
        off = sum(np.abs(blocks[(i, j)]).sum() for j in cols if j != i)
        for j in cols:
            if j == i:
                D = rng.normal(0, 1, (3, 3))
                v.append((D @ D.T + (off + 1.0) * np.eye(3)).ravel())
            else:
                v.append(blocks[(i, j)].ravel())
            ci.append(j)
        rp.append(len(ci))
    rp, ci, v = np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)
    b = rng.normal(0, 1, 3 * n)
    A, xd = _dense(rp, ci, v, b)
    x2, _, _ = oracle.solve_bsr(rp, ci, v, b, direct=2)
    assert np.linalg.norm(x2 - xd) <= 1e-11 * np.linalg.norm(xd)


def test_nd_refuses_indefinite():
    rp, ci, v, b = _system(scenes.hang(9))
    v = v.copy()
    d = [k for i in range(len(rp) - 1) for k in range(rp[i], rp[i + 1]) if ci[k] == i]
    v[d[40]] = -v[d[40]]  # one negative diagonal block
    with pytest.raises(Exception):
        oracle.solve_bsr(rp, ci, v, b, direct=2)


def test_nd_step_equals_banded_step():
    """A full oracle step with either direct solver gives the same positions."""
    spec = scenes.fold(17)
    a, c = oracle.OracleScene(spec), oracle.OracleScene(spec)
    for _ in range(2):
        assert a.step(direct=1)[0] == 0 and c.step(direct=2)[0] == 0
    assert np.abs(a.state()[0] - c.state()[0]).max() <= 1e-6
