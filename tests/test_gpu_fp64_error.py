"""Each result of the GPU step measured against the FP64 oracle (the same templates in double).

The product computes in FP32 except the CCD cubic, which is FP64 on the device.  Per result, the
error against the FP64 oracle on the same input state, and the bar the test holds it to (fold33:
a 33^2 sheet folded over itself, proximities and contact blocks from the first step; hang33: the
same sheet hanging from two corners, no contact):

  result                          measure                                              bar
  proximity records (fold33)      records found by one side only (a distance within    <= 1 %
                                  FP32 rounding of the thickness)
                                  common records: max |n - n64|, |w - w64|,            1e-5, 1e-5,
                                  stiffness relative                                   1e-4
  assembled A, b (fold33)         rows no one-sided record touches: max |A - A64| /    5e-5, 2e-5
                                  max |row of A64|; max |b - b64| / max |b64|
  PCG solution (hang33, fold33;   |dv - dv64| / |dv64|, dv64 the FP64 direct solve     cond(A) * 1e-6
  tol 1e-6)                       of the same FP32 system
  physics step x (tol 1e-6)       hang33: max |x - x64|                                1e-5 m
                                  fold33 (includes the one-sided records' blocks)      1e-4 m
  impacts (FP64 CCD on device)    max |t - t64|, |w - w64|, |n - n64|                  1e-9
  impact response x               max |x - x64| (100 gradient iterations each)         5e-6 m
  full step x (fold33, tol 1e-6)  max |x - x64| per step from the FP64 state           1e-4 m

(The solves run to 1e-6 rather than 1e-5 so that the figures show the FP32 arithmetic more than
the tolerance; at 1e-7 an FP32 PCG can stall on its true residual.)

Each test prints the measured value; the figures are quoted where the parity status is listed.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from clothsimulator_amd import CS_MESH_CLOTH, Simulator
from clothsimulator_amd.scenes import fold, hang
from test_gpu_lbvh import _crossing

pytestmark = pytest.mark.gpu


def _scene(kind="fold", tol=1e-5):
    spec = fold(33) if kind == "fold" else hang(33)
    spec.params.pcg_rel_tol = tol
    return spec


def _keyed(recs):
    return {tuple(int(v) for v in r["nodes"]): r for r in recs}


def test_proximity_error_vs_fp64():
    spec = _scene()
    sim, o64 = Simulator(spec), oracle.OracleScene(spec, fp64=True)
    kg, ko = _keyed(sim.proximities()), _keyed(o64.proximities())
    both = set(kg) & set(ko)
    print("fold33 proximities: GPU %d, FP64 oracle %d, common %d" % (len(kg), len(ko), len(both)))
    assert len(set(kg) ^ set(ko)) <= 0.01 * len(both) and len(both) > 100
    en = max(float(np.abs(kg[k]["n"] - ko[k]["n"]).max()) for k in both)
    ew = max(float(np.abs(kg[k]["w"] - ko[k]["w"]).max()) for k in both)
    es = max(float(abs(kg[k]["stiffness"] - ko[k]["stiffness"]) / abs(ko[k]["stiffness"])) for k in both)
    print("  common records: max |n - n64| %.3g, |w - w64| %.3g, stiffness relative %.3g" % (en, ew, es))
    assert en <= 1e-5 and ew <= 1e-5 and es <= 1e-4


def test_assembly_error_vs_fp64():
    spec = _scene()
    sim, o64 = Simulator(spec), oracle.OracleScene(spec, fp64=True)
    kg, ko = _keyed(sim.proximities()), _keyed(o64.proximities())
    touched = {n for k in set(kg) ^ set(ko) for n in k if n >= 0}
    rp, ci, vals, b = sim.assemble(0)
    orp, oci, ovals, ob = o64.assemble(0)
    clean, dirty = 0.0, 0.0
    for r in range(len(rp) - 1):
        if orp[r + 1] - orp[r] != rp[r + 1] - rp[r] or not np.array_equal(ci[rp[r]:rp[r + 1]], oci[orp[r]:orp[r + 1]]):
            assert r in touched, r  # a pattern difference only where a one-sided record adds a block
            continue
        s, e = rp[r], rp[r + 1]
        err = float(np.abs(vals[s:e] - ovals[s:e]).max() / np.abs(ovals[s:e]).max())
        if r in touched:
            dirty = max(dirty, err)
        else:
            clean = max(clean, err)
    eb = float(np.abs(b - ob).max() / np.abs(ob).max())
    print("fold33 assembly vs FP64: A %.3g of the row maximum (%d rows touched by one-sided records: %.3g), "
          "b %.3g of max|b|" % (clean, len(touched), dirty, eb))
    assert clean <= 5e-5 and eb <= 2e-5


@pytest.mark.parametrize("kind", ["hang", "fold"])
def test_solve_error_vs_fp64(kind):
    spec = _scene(kind, 1e-6)
    sim = Simulator(spec)
    rp, ci, vals, b = sim.assemble(0)
    dv, st = sim.solve(0)
    v64 = vals.reshape(-1, 9).astype(np.float64)
    dv64, _, _ = oracle.solve_bsr(rp, ci, v64, b.astype(np.float64), direct=True)
    dv64 = np.asarray(dv64).reshape(-1)
    err = float(np.linalg.norm(np.asarray(dv, np.float64) - dv64) / np.linalg.norm(dv64))
    A = sp.bsr_matrix((vals.astype(np.float64), ci, rp), shape=(len(b), len(b))).toarray()
    ev = np.linalg.eigvalsh(A)
    cond = float(ev[-1] / ev[0])
    print("%s33 PCG (tol 1e-6, %d iterations) vs FP64 direct solve of the same system: relative %.3g, "
          "cond(A) %.3g" % (kind, st.pcg_iters, err, cond))
    assert st.pcg_converged == 1 and err <= cond * 1e-6


@pytest.mark.parametrize("kind,bar", [("hang", 1e-5), ("fold", 1e-4)])
def test_physics_step_error_vs_fp64(kind, bar):
    spec = _scene(kind, 1e-6)
    sim, o64 = Simulator(spec), oracle.OracleScene(spec, fp64=True)
    s, _ = sim.physics_step()
    assert s == 0
    r, _ = o64.physics_step(direct=True)
    assert r == 0
    d = float(np.abs(sim.state()[0] - o64.state(fp64=True)[0]).max())
    print("%s33 physics step vs FP64 (direct solve): max |dx| %.3g m" % (kind, d))
    assert d <= bar


def _crossed(n=33):
    spec = _scene()
    sim, o64 = Simulator(spec), oracle.OracleScene(spec, fp64=True)
    x = sim.state()[0]
    x1, moved = _crossing(x, n, 0.2, 3e-3)
    assert moved > 10
    for s in (sim, o64):
        s.set_state(CS_MESH_CLOTH, 0, x1, x, np.zeros_like(x))
    return sim, o64


def test_impact_error_vs_fp64():
    sim, o64 = _crossed()
    g, o = sim.impacts(1e-4), o64.impacts(1e-4)
    kg = {(int(r["type"]),) + tuple(int(v) for v in r["nodes"]): r for r in g}
    ko = {(int(r["type"]),) + tuple(int(v) for v in r["nodes"]): r for r in o}
    print("fold33 crossing impacts: GPU %d, FP64 oracle %d, common %d" % (len(kg), len(ko), len(set(kg) & set(ko))))
    assert set(kg) == set(ko) and len(kg) > 100
    et = max(abs(float(kg[k]["t"]) - float(ko[k]["t"])) for k in kg)
    ew = max(float(np.abs(kg[k]["w"] - ko[k]["w"]).max()) for k in kg)
    en = max(float(np.abs(kg[k]["n"] - ko[k]["n"]).max()) for k in kg)
    print("  max |t - t64| %.3g, |w - w64| %.3g, |n - n64| %.3g" % (et, ew, en))
    assert et <= 1e-9 and ew <= 1e-9 and en <= 1e-9


def test_impact_response_error_vs_fp64():
    sim, o64 = _crossed()
    imps = sim.independent_impacts(sim.impacts(1e-4), 0)
    assert len(imps) > 10
    ig = sim.impact_response(imps, 0)
    io = o64.impact_response(imps, 0)
    d = float(np.abs(sim.state()[0].astype(np.float64) - o64.state(fp64=True)[0]).max())
    print("fold33 impact response (%d impacts): GPU %d iterations, FP64 oracle %d; max |dx| %.3g m" % (
        len(imps), ig, io, d))
    assert d <= 5e-6


def test_full_step_error_vs_fp64():
    spec = _scene("fold", 1e-6)
    sim, o64 = Simulator(spec), oracle.OracleScene(spec, fp64=True)
    worst = 0.0
    for k in range(3):
        x, x0, v = o64.state(fp64=True)
        sim.set_state(CS_MESH_CLOTH, 0, x.astype(np.float32), x0.astype(np.float32), v.astype(np.float32))
        s, st = sim.step()
        assert s == 0
        r, _ = o64.step(direct=True)
        assert r == 0
        d = float(np.abs(sim.state()[0] - o64.state(fp64=True)[0]).max())
        worst = max(worst, d)
        print("fold33 step %d (physics + collision) vs FP64, from the FP64 state: max |dx| %.3g m" % (k, d))
    assert worst <= 1e-4
