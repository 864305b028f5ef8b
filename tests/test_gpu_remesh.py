"""The remeshing step on the device against the oracle (oracle/remesh.py).

* sizing: the device's nearest obstacle planes and face sizing tensors (FP32) against the FP64
  restatement on the same state;
* operators: the context's remesh (cs_remesh) against the oracle's operators fed the device's own
  sizing, node for node and face for face;
* the rebuilt context: a physics step after remeshing against the oracle's direct solve on the
  remeshed mesh, and a hanging sheet remeshed every frame keeping its handles.

Reference: Simulator::remeshingStep (src/Simulator.cu:686-695), Cloth::remeshingStep
(src/Cloth.cu:815-833)."""
import numpy as np
import pytest

import oracle
from oracle import remesh as R
from clothsimulator_amd import CS_MESH_CLOTH, CS_MESH_OBSTACLE, MeshArrays, Simulator, remeshing_struct, scenes
from clothsimulator_amd._cabi import material_struct

pytestmark = pytest.mark.gpu

# the reference's conf/hang.json block
BLOCK = dict(refine_angle=0.3, refine_compression=0.005, refine_velocity=0.5, size=(10e-3, 200e-3), aspect_min=0.2)


def stepped_drape(n=17, steps=25):
    spec = scenes.drape(n)
    sim = Simulator(spec)
    for _ in range(steps):
        s, _ = sim.step()
        assert s == 0
    return spec, sim


def oracle_inputs(sim, cloth=0):
    x, x0, v = sim.state(CS_MESH_CLOTH, cloth)
    fn, fv, uv = sim.topology(CS_MESH_CLOTH, cloth)
    vnode = np.zeros(len(uv), np.int64)
    vnode[fv.ravel()] = fn.ravel()
    return x, x0, v, uv, vnode, fv


def obstacles_of(sim, spec):
    out = []
    for k in range(len(spec.obstacles)):
        y = sim.state(CS_MESH_OBSTACLE, k)[0]
        fn, _, _ = sim.topology(CS_MESH_OBSTACLE, k)
        out.append((y.astype(np.float64), fn))
    return out


def test_sizing_matches_oracle():
    spec, sim = stepped_drape()
    r = remeshing_struct(**BLOCK)
    fs, pl = sim.remesh_sizing(0, r)
    x, x0, v, uv, vnode, fv = oracle_inputs(sim)
    thick = 10 * spec.params.repulsion_thickness
    P, N = R.nearest_planes(x.astype(np.float64), obstacles_of(sim, spec), thick)
    has = np.any(N != 0, axis=1)
    assert np.array_equal(has, np.any(pl[:, 4:7] != 0, axis=1))
    assert has.sum() > 10
    ep = float(np.abs(pl[has, :3] - P[has]).max())
    en = float(np.abs(pl[has, 4:7] - N[has]).max())
    geo = R.Geometry(x, uv, vnode, fv, R.sorted_edges(vnode, fv))
    S = R.face_sizing(geo, v.astype(np.float64), (P, N), R.Remeshing(**BLOCK))
    S = S.reshape(-1, 4)
    scale = np.abs(S).max(axis=1)
    rel = np.abs(fs - S).max(axis=1) / scale
    print("drape17 after 25 steps: %d nodes with a plane (max |p - p64| %.3g, |n - n64| %.3g); face sizing vs "
          "FP64: median %.3g, 99th percentile %.3g, max %.3g of the tensor" % (
              has.sum(), ep, en, np.median(rel), np.percentile(rel, 99), rel.max()))
    assert ep <= 1e-5 and en <= 1e-4
    assert np.percentile(rel, 99) <= 1e-3 and rel.max() <= 1e-2


def test_context_remesh_matches_oracle_operators():
    spec, sim = stepped_drape()
    r = remeshing_struct(**BLOCK)
    fs, _ = sim.remesh_sizing(0, r)
    x, x0, v, uv, vnode, fv = oracle_inputs(sim)
    en = R.sorted_edges(vnode, fv)
    m = R.Mesh(x, x0, v, R.load_preserve(len(x), vnode, fv, en), uv, vnode, fv, en)
    m.remesh(fs, R.Remeshing(**BLOCK))
    a = m.arrays()
    sim.set_remeshing(0, r)
    sim.remesh()
    nn, nf, ne = sim.counts(CS_MESH_CLOTH, 0)
    print("drape17 remeshed: %d -> %d nodes, %d faces, %d edges (oracle %d, %d, %d)" % (
        len(x), nn, nf, ne, len(a["x"]), len(a["fv"]), len(a["en"])))
    assert (nn, nf, ne) == (len(a["x"]), len(a["fv"]), len(a["en"]))
    x2, x02, v2 = sim.state(CS_MESH_CLOTH, 0)
    fn2, fv2, uv2 = sim.topology(CS_MESH_CLOTH, 0)
    assert np.array_equal(x2, a["x"]) and np.array_equal(v2, a["v"])
    assert np.array_equal(fv2, a["fv"]) and np.array_equal(uv2, a["uv"])
    assert np.array_equal(fn2, a["vnode"][a["fv"]])
    # the obstacle is untouched
    assert sim.counts(CS_MESH_OBSTACLE, 0) == (len(spec.obstacles[0].x), len(spec.obstacles[0].faces),
                                               sim.counts(CS_MESH_OBSTACLE, 0)[2])


def test_step_after_remesh_matches_oracle():
    """The rebuilt context (pattern, incidence lists, trees) steps like the oracle on the new mesh."""
    spec, sim = stepped_drape()
    sim.set_remeshing(0, remeshing_struct(**BLOCK))
    sim.remesh()
    x, x0, v = sim.state(CS_MESH_CLOTH, 0)
    fn, fv, uv = sim.topology(CS_MESH_CLOTH, 0)
    y = sim.state(CS_MESH_OBSTACLE, 0)[0]
    yf, _, _ = sim.topology(CS_MESH_OBSTACLE, 0)
    spec2 = scenes.SceneSpec("drape17-remeshed", spec.params, [scenes.ClothSpec(
        MeshArrays(x, fn, uv=uv, face_verts=fv, v=v), material_struct())], [MeshArrays(y, yf)], 1)
    o = oracle.OracleScene(spec2, fp64=True)
    worst = 0.0
    for k in range(3):
        xs, x0s, vs = sim.state(CS_MESH_CLOTH, 0)
        o.set_state(CS_MESH_CLOTH, 0, xs, x0s, vs)
        s, st = sim.physics_step()
        assert s == 0 and st.pcg_converged
        r, _ = o.physics_step(direct=True)
        assert r == 0
        d = float(np.abs(sim.state(CS_MESH_CLOTH, 0)[0] - o.state(fp64=True)[0]).max())
        worst = max(worst, d)
        s, _ = sim.collision_step()
        assert s == 0
    print("remeshed drape17 (%d nodes): physics steps vs FP64 direct solve, max |dx| %.3g m" % (len(x), worst))
    assert worst <= 1e-5


def test_hang_remeshed_every_frame_keeps_its_handles():
    spec = scenes.hang(9)
    sim = Simulator(spec)
    sim.set_remeshing(0, remeshing_struct(**BLOCK))
    corners = spec.cloths[0].mesh.x[spec.cloths[0].handles]
    counts = []
    for frame in range(6):
        for _ in range(8):
            s, _ = sim.step()
            assert s == 0
        sim.remesh()
        x, _, v = sim.state(CS_MESH_CLOTH, 0)
        assert np.isfinite(x).all() and np.isfinite(v).all()
        counts.append(sim.counts(CS_MESH_CLOTH, 0)[0])
        # the handled corners are preserved nodes: still there, still (nearly) in place
        d = np.min(np.linalg.norm(x[None, :, :] - corners[:, None, :], axis=2), axis=1)
        assert d.max() <= 1e-3, d
    print("hang9 remeshed each frame: node counts %s" % counts)
    assert counts[-1] != 81  # the sheet was remeshed
