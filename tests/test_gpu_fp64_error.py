"""Each result of the GPU step measured against the FP64 oracle (the same templates in double).

The product computes in FP32 except the CCD cubic, which is FP64 on the device.  Per result,
on fold33 (a sheet folded over itself: proximities, contact blocks and, after a crossing, impacts
from the first step), the error against the FP64 oracle on the same input state, and the bar
the test holds it to:

  result                          measure                                         bar
  assembled A (blocks)            max |A - A64| / max |row of A64|                2e-5
  assembled b                     max |b - b64| / max |b64|                       2e-5
  PCG solution dv (tol 1e-7)      |dv - dv64| / |dv64|, dv64 the FP64 direct      1e-4
                                  solve of the FP64 system
  physics step x (tol 1e-7)       max |x - x64| (m)                               1e-5
  proximity records               max |n - n64|, |w - w64|, |stiffness| rel.      1e-5, 1e-5, 1e-4
  impacts (FP64 CCD on device)    max |t - t64|, |w - w64|, |n - n64|             1e-9
  impact response x               max |x - x64| (m)                               1e-6
  full step x (tol 1e-7)          max |x - x64| (m), physics + collision          1e-5

Each test prints the measured value; the figures are quoted where the parity status is listed.
"""
import numpy as np
import pytest

import oracle
from clothsimulator_amd import CS_MESH_CLOTH, Simulator
from clothsimulator_amd.scenes import fold
from test_gpu_lbvh import _crossing

pytestmark = pytest.mark.gpu


def _scene(tol=1e-5):
    spec = fold(33)
    spec.params.pcg_rel_tol = tol
    return spec


def test_assembly_error_vs_fp64():
    spec = _scene()
    sim, o64 = Simulator(spec), oracle.OracleScene(spec, fp64=True)
    rp, ci, vals, b = sim.assemble(0)
    orp, oci, ovals, ob = o64.assemble(0)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    worst = 0.0
    for r in range(len(rp) - 1):
        s, e = rp[r], rp[r + 1]
        worst = max(worst, float(np.abs(vals[s:e] - ovals[s:e]).max() / np.abs(ovals[s:e]).max()))
    eb = float(np.abs(b - ob).max() / np.abs(ob).max())
    print("fold33 assembly vs FP64: A %.3g of the row maximum, b %.3g of max|b|" % (worst, eb))
    assert worst <= 2e-5 and eb <= 2e-5


def test_solve_error_vs_fp64():
    spec = _scene(1e-7)
    sim, o64 = Simulator(spec), oracle.OracleScene(spec, fp64=True)
    sim.assemble(0)
    dv, st = sim.solve(0)
    orp, oci, ovals, ob = o64.assemble(0)
    dv64, _, _ = oracle.solve_bsr(orp, oci, ovals.reshape(-1, 9), ob, direct=True)
    dv64 = np.asarray(dv64).reshape(-1)
    err = float(np.linalg.norm(np.asarray(dv, np.float64).reshape(-1) - dv64) / np.linalg.norm(dv64))
    print("fold33 PCG (tol 1e-7, %d iterations) vs FP64 direct solve of the FP64 system: relative %.3g"
          % (st.pcg_iters, err))
    assert st.pcg_converged == 1 and err <= 1e-4


def test_physics_step_error_vs_fp64():
    spec = _scene(1e-7)  # the FP32 arithmetic, not the solve tolerance
    sim, o64 = Simulator(spec), oracle.OracleScene(spec, fp64=True)
    s = sim.physics_step()
    assert s[0] == 0
    r, _ = o64.physics_step(direct=True)
    assert r == 0
    d = float(np.abs(sim.state()[0] - o64.state(fp64=True)[0]).max())
    print("fold33 physics step vs FP64 (direct solve): max |dx| %.3g m" % d)
    assert d <= 1e-5


def test_proximity_error_vs_fp64():
    spec = _scene()
    sim, o64 = Simulator(spec), oracle.OracleScene(spec, fp64=True)
    g, o = sim.proximities(), o64.proximities()
    kg = {tuple(int(v) for v in r["nodes"]): r for r in g}
    ko = {tuple(int(v) for v in r["nodes"]): r for r in o}
    both = set(kg) & set(ko)
    print("fold33 proximities: GPU %d, FP64 oracle %d, common %d" % (len(kg), len(ko), len(both)))
    assert len(both) >= 0.99 * max(len(kg), len(ko)) and len(both) > 100
    en = max(float(np.abs(kg[k]["n"] - ko[k]["n"]).max()) for k in both)
    ew = max(float(np.abs(kg[k]["w"] - ko[k]["w"]).max()) for k in both)
    es = max(float(abs(kg[k]["stiffness"] - ko[k]["stiffness"]) / abs(ko[k]["stiffness"])) for k in both)
    print("  max |n - n64| %.3g, |w - w64| %.3g, stiffness relative %.3g" % (en, ew, es))
    assert en <= 1e-5 and ew <= 1e-5 and es <= 1e-4


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
    assert d <= 1e-6


def test_full_step_error_vs_fp64():
    spec = _scene(1e-7)
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
    assert worst <= 1e-5
