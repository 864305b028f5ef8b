"""The committed fixtures of tests/golden/small_cases.npz (written by
tests/golden/make_golden.py from the CPU oracle):

  * without a GPU: the oracle built here still reproduces every fixture (so a change to the
    oracle that moves an answer is seen);
  * on the GPU: the product matches the fixtures to the parity bar -- the assembled pattern
    exactly and its values within 2e-5 of each row's largest entry, positions within 1e-5 m
    after each of 3 steps, candidate pairs as a set and proximity primitives exactly, CCD
    impacts bit for bit, motion evaluation bit for bit.
"""
import os
import sys

import numpy as np
import pytest

import clothsimulator_amd as cs
import oracle
from clothsimulator_amd import CS_E_NOT_CONVERGED, CS_MESH_CLOTH, CS_OK, Simulator
from clothsimulator_amd.scenes import hang

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_golden import crossing_state, two_sheets  # noqa: E402

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "small_cases.npz"))


def pair_key(p):
    return np.sort(p[:, 0].astype(np.int64) * (1 << 32) + p[:, 1])


def test_oracle_reproduces_fixtures():
    spec = hang(9)
    rp, ci, vals, b = oracle.OracleScene(spec, fp64=True).assemble(0)
    assert np.array_equal(rp, G["hang9_row_ptr"]) and np.array_equal(ci, G["hang9_col_idx"])
    assert np.allclose(vals, G["hang9_vals"], rtol=1e-12, atol=1e-14) and np.allclose(b, G["hang9_b"], rtol=1e-12, atol=1e-14)
    o = oracle.OracleScene(spec)
    for k in range(3):
        assert o.step(direct=True)[0] == CS_OK
        assert np.allclose(o.state()[0], G["hang9_x_steps"][k], atol=1e-7), k
    spec, x = two_sheets(9, gap=1.5e-3)
    o = oracle.OracleScene(spec)
    assert np.array_equal(np.asarray(o.lbvh_pairs(False, 2e-3), dtype=np.int32), G["sheets_pairs"])
    assert np.array_equal(o.proximities(), G["sheets_prox"])
    spec, x = two_sheets(9, gap=1e-3)
    o = oracle.OracleScene(spec)
    o.set_state(CS_MESH_CLOTH, 0, *crossing_state(x))
    assert np.array_equal(o.impacts(1e-4), G["sheets_impacts"])
    for t, want in zip(G["motion_at"], G["motion_out"]):
        assert np.array_equal(oracle.motion_eval(G["motion_t"], G["motion_k"], t), want)


def test_product_motion_eval_matches_fixture():
    for t, want in zip(G["motion_at"], G["motion_out"]):
        assert np.array_equal(cs.motion_eval(G["motion_t"], G["motion_k"], t).as_array(), want)


@pytest.mark.gpu
def test_product_assembly_and_steps_match_fixtures():
    sim = Simulator(hang(9))
    rp, ci, vals, b = sim.assemble(0)
    assert np.array_equal(rp, G["hang9_row_ptr"]) and np.array_equal(ci, G["hang9_col_idx"])
    ov = G["hang9_vals"]
    for r in range(len(rp) - 1):
        s, e = rp[r], rp[r + 1]
        assert np.abs(vals[s:e] - ov[s:e]).max() <= 2e-5 * np.abs(ov[s:e]).max(), r
    assert np.abs(b - G["hang9_b"]).max() <= 2e-5 * np.abs(G["hang9_b"]).max()
    sim = Simulator(hang(9))
    for k in range(3):
        s, st = sim.step()
        assert s in (CS_OK, CS_E_NOT_CONVERGED) and st.pcg_converged == 1
        d = np.abs(sim.state()[0] - G["hang9_x_steps"][k]).max()
        print("hang9 step %d: max |x - fixture| %.3g" % (k, d))
        assert d <= 1e-5


@pytest.mark.gpu
def test_product_pairs_proximities_impacts_match_fixtures():
    spec, x = two_sheets(9, gap=1.5e-3)
    sim = Simulator(spec)
    assert np.array_equal(pair_key(sim.lbvh_pairs(False, 2e-3)), pair_key(G["sheets_pairs"]))
    p, want = sim.proximities(), G["sheets_prox"]
    assert np.array_equal(p["nodes"], want["nodes"]) and np.array_equal(p["mu"], want["mu"])
    assert np.allclose(p["n"], want["n"], atol=1e-6) and np.allclose(p["w"], want["w"], atol=1e-5)
    spec, x = two_sheets(9, gap=1e-3)
    sim = Simulator(spec)
    sim.set_state(CS_MESH_CLOTH, 0, *crossing_state(x))
    got, want = sim.impacts(1e-4), G["sheets_impacts"]
    print("sheets: %d pairs, %d proximities, %d impacts" % (len(G["sheets_pairs"]), len(p), len(got)))
    assert len(want) > 0 and len(got) == len(want)
    for f in ("type", "nodes", "t", "w", "n"):
        assert np.array_equal(got[f], want[f]), f
