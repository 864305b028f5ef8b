"""Scene files of the reference (configuration JSON + OBJ + material JSON) read
through cs_load_scene.  CPU tests parse only (no device); the GPU test checks
the placed meshes against an explicit construction of the same scene."""
import math
import os
import re

import numpy as np
import pytest

import clothsimulator_amd as cs
from clothsimulator_amd import CS_MESH_CLOTH, CS_MESH_OBSTACLE, ClothError
from clothsimulator_amd._cabi import MeshArrays, default_params, material_struct
from clothsimulator_amd.scenes import ClothSpec, SceneSpec

REF = "/root/reference"

MATERIAL = """{
    // a material file in the reference's format (comments allowed)
    "density": 0.2,
    "stretching": [[10, 1, 20, 5], [11, 1, 21, 5], [12, 2, 22, 6], [13, 2, 23, 6], [14, 3, 24, 7], [15, 3, 25, 7]],
    "bending": [[1e-6, 2e-6, 3e-6, 4e-6, 5e-6], [1e-6, 2e-6, 3e-6, 4e-6, 5e-6], [1e-6, 2e-6, 3e-6, 4e-6, 5e-6]]
}
"""

# a 3x2-quad sheet written with quads (triangulated by the loader), velocities and no 'vt'
SHEET = """# sheet
v 0 0 0
v 0.5 0 0
v 1 0 0
v 0 0.5 0
v 0.5 0.5 0
v 1 0.5 0
nv 0 0 1
nv 0 0 1
nv 0 0 1
nv 0 0 1
nv 0 0 1
nv 0 0 1
f 1 2 5 4
f 2 3 6 5
"""

TRI = """v -1 -1 0
v 1 -1 0
v 0 1 0
vt 0 0
vt 1 0
vt 0 1
f 1/1 2/2 3/3
"""

CONFIG = """// test scene
{
    "frame_time": 0.04,
    "frame_steps": 8,
    "end_time": 1.5,
    "cloths": [{
        "mesh": "sheet.obj",
        "transform": {"translate": [0.1, -0.2, 0.5], "rotate": [90, 0, 0, 1], "scale": 2},
        "materials": {"data": "mat.json", "thicken": 2, "bending_mult": 0.5},
        "remeshing": {"refine_angle": 0.3}
    }],
    "obstacles": [{"mesh": "tri.obj", "transform": {"scale": 0.5}}],
    "handles": [{"nodes": [0, 2]}],
    "gravity": [0, 0, -9.8],
    "wind": {"velocity": [1, 0, 0]},
    "magic": {"repulsion_thickness": 5e-3, "collision_stiffness": 1e6},
    "disable": ["remeshing"]
}
"""


def write_scene(d, config=CONFIG):
    for name, text in (("mat.json", MATERIAL), ("sheet.obj", SHEET), ("tri.obj", TRI), ("scene.json", config)):
        with open(os.path.join(d, name), "w") as f:
            f.write(text)
    return os.path.join(d, "scene.json")


def test_scene_info_of_written_scene(tmp_path):
    path = write_scene(str(tmp_path))
    info = cs.scene_info(path, str(tmp_path))
    assert (info.n_cloths, info.n_obstacles, info.n_handles) == (1, 1, 1)
    assert info.n_cloth_nodes == 6 and info.n_cloth_faces == 4  # two quads -> four triangles
    assert info.frame_steps == 8 and info.frame_time == pytest.approx(0.04)
    assert info.end_time == pytest.approx(1.5) and info.end_frame == 2 ** 31 - 1


REMESH_BLOCK = ('"remeshing": {"refine_angle": 0.3, "refine_compression": 0.005, "refine_velocity": 0.5, '
                '"size": [10e-3, 200e-3], "aspect_min": 0.2}')


def remeshing_config():
    return CONFIG.replace('"remeshing": {"refine_angle": 0.3}', REMESH_BLOCK).replace(
        ',\n    "disable": ["remeshing"]', '')


def test_remeshing_flag(tmp_path):
    """"disable": ["remeshing"] turns the per-frame remeshing off (src/Simulator.cu:858)."""
    assert cs.scene_info(write_scene(str(tmp_path)), str(tmp_path)).remeshing == 0
    assert '"disable"' not in remeshing_config()
    assert cs.scene_info(write_scene(str(tmp_path), remeshing_config()), str(tmp_path)).remeshing == 1


@pytest.mark.parametrize("bad, status", [
    ('"obstacles": [{"mesh": "tri.obj", "motion": 0}]', "CS_E_FORMAT.*missing motion"),
    ('"obstacles": "obs_%s.obj"', "CS_E_FORMAT.*%d"),
    ('"handles": [{"nodes": [99]}]', "CS_E_FORMAT"),
    ('"handles": [{"nodes": [0], "motion": 1}]', "CS_E_FORMAT.*missing motion"),
])
def test_malformed_scenes(tmp_path, bad, status):
    cfg = CONFIG.replace('"obstacles": [{"mesh": "tri.obj", "transform": {"scale": 0.5}}]', '"obstacles": []')
    cfg = cfg.replace('"handles": [{"nodes": [0, 2]}]', '"handles": []')
    key = bad.split(":")[0]
    cfg = cfg.replace('%s: []' % key, bad)
    path = write_scene(str(tmp_path), cfg)
    with pytest.raises(ClothError, match=status):
        cs.scene_info(path, str(tmp_path))


def test_json_syntax_error_is_reported(tmp_path):
    path = write_scene(str(tmp_path), CONFIG.replace('"frame_steps": 8,', '"frame_steps": 8'))
    with pytest.raises(ClothError, match="CS_E_FORMAT.*line"):
        cs.scene_info(path, str(tmp_path))


def test_missing_mesh_file(tmp_path):
    path = write_scene(str(tmp_path), CONFIG.replace("sheet.obj", "nothere.obj"))
    with pytest.raises(ClothError, match="CS_E_ARG.*nothere.obj"):
        cs.scene_info(path, str(tmp_path))


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
def test_reference_configurations_parse():
    """Every configuration of the reference loads (with its motions, .mot files and obstacle
    sequences), except where a file it names is not in the reference tree: conf/karate.json's
    motions/karate-smooth.mot is absent, and that is reported by name."""
    loaded = {}
    for name in sorted(os.listdir(os.path.join(REF, "conf"))):
        try:
            info = cs.scene_info(os.path.join(REF, "conf", name), REF)
            assert info.n_cloths >= 1 and info.n_cloth_faces > 0
            loaded[name] = info
        except ClothError as e:
            missing = re.search(r"failed to open \w+ file: (\S+)", str(e))
            assert "CS_E_ARG" in str(e) and missing and not os.path.exists(missing.group(1)), (name, str(e))
    assert len(loaded) >= 11
    assert (loaded["funnel.json"].n_motions, loaded["funnel.json"].n_moving_obstacles) == (1, 1)
    assert (loaded["dress-blue.json"].n_motions, loaded["dress-blue.json"].n_moving_obstacles) == (16, 16)
    assert loaded["sphere.json"].n_moving_obstacles == 1 and loaded["sphere.json"].n_handles == 1
    assert loaded["hang.json"].remeshing == 1 and loaded["hang32.json"].remeshing == 0


def expected_scene():
    """The written scene built explicitly: transform = scale * rotate + translate,
    quads fanned from the corner with the largest minimum angle (corner 0 on ties)."""
    raw = np.array([[0, 0, 0], [0.5, 0, 0], [1, 0, 0], [0, 0.5, 0], [0.5, 0.5, 0], [1, 0.5, 0]], dtype=np.float64)
    rot = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], dtype=np.float64)  # 90 degrees about z
    x = 2.0 * raw @ rot.T + np.array([0.1, -0.2, 0.5])
    return x, rot


@pytest.mark.gpu
def test_loaded_scene_places_meshes(tmp_path):
    path = write_scene(str(tmp_path))
    sim = cs.Simulator.from_config(path, str(tmp_path))
    x, x0, v = sim.state(CS_MESH_CLOTH, 0)
    xe, rot = expected_scene()
    assert np.allclose(x, xe, atol=1e-6)
    assert np.allclose(v, np.tile([0, 0, 1.0], (6, 1)), atol=1e-7)  # rotation about z keeps +z
    xo, _, _ = sim.state(CS_MESH_OBSTACLE, 0)
    assert np.allclose(xo, 0.5 * np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]]), atol=1e-7)
    n, area, mass = sim.geometry(CS_MESH_CLOTH, 0)
    # uv = scale * raw (x, y): rest area 2^2 * 0.5 m^2, density 0.2 * thicken 2
    assert area.sum() == pytest.approx(4 * 0.5, rel=1e-5)
    assert mass.sum() == pytest.approx(0.2 * 2 * 4 * 0.5, rel=1e-5)
    s, st = sim.step()
    assert s in (0, 4)


def build_host_mirror(out_dir):
    """Compile tests/cpp/host_mirror.cpp (include/clothsim.hpp over libclothsim.so) with g++."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib_dir = os.path.join(root, "clothsimulator_amd")
    exe = os.path.join(out_dir, "host_mirror")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(root, "include"),
                           os.path.join(root, "tests", "cpp", "host_mirror.cpp"), "-o", exe, "-L", lib_dir,
                           "-lclothsim", "-Wl,-rpath," + lib_dir, "-pthread"])
    return exe


def test_mesh_save_text_is_the_references(tmp_path):
    """Mesh::save writes v / nv / vt / f rows as the reference's stream at precision 20 does
    (src/Mesh.cu:454-507): printf's %.20g of each float (here from Python's exact formatter) and
    1-based "node/vertex" face indices, including zeros, tiny, huge and inexact values."""
    import subprocess
    vals = np.array([0.0, -0.0, 1e-30, 3.4e38, 0.1, -2.5, 1.0 / 3.0, 1e-45, 123456.789], dtype=np.float32)
    src = tmp_path / "save.cpp"
    src.write_text('#include "clothsim.hpp"\nint main(int, char** argv) {\n  clothsim::Mesh m;\n'
                   '  m.x = {%s};\n  m.v = {%s};\n  m.uv = {%s};\n  m.faces = {0, 1, 2};\n'
                   '  m.faceVerts = {2, 0, 1};\n  m.save(argv[1]);\n}\n' % (
                       ", ".join("%.9ef" % v for v in vals), ", ".join("%.9ef" % v for v in vals[::-1]),
                       ", ".join("%.9ef" % v for v in vals[:6])))
    exe = tmp_path / "save"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lib_dir = os.path.join(root, "clothsimulator_amd")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-I", os.path.join(root, "include"), str(src), "-o", str(exe),
                           "-L", lib_dir, "-lclothsim", "-Wl,-rpath," + lib_dir])
    out = tmp_path / "m.obj"
    subprocess.check_call([str(exe), str(out)])
    g = lambda a: " ".join("%.20g" % float(t) for t in a)
    want = []
    for i in range(3):
        want += ["v " + g(vals[3 * i:3 * i + 3]), "nv " + g(vals[::-1][3 * i:3 * i + 3])]
    want += ["vt " + g(vals[2 * k:2 * k + 2]) for k in range(3)] + ["f 1/3 2/1 3/2"]
    assert out.read_text().splitlines() == want


def test_write_config_round_trip(tmp_path):
    """cs_write_config keeps the configuration, points cloth 0 at a new mesh and drops its
    transform; the written file loads to the same scene."""
    import json
    d = str(tmp_path)
    path = write_scene(d)
    with open(os.path.join(d, "frame7_cloth0.obj"), "w") as f:
        f.write(open(os.path.join(d, "sheet.obj")).read())
    out = os.path.join(d, "saved.json")
    cs.write_config(path, out, [os.path.join(d, "frame7_cloth0.obj")])
    j = json.load(open(out))
    assert j["cloths"][0]["mesh"].endswith("frame7_cloth0.obj") and "transform" not in j["cloths"][0]
    assert j["frame_steps"] == 8 and j["obstacles"][0]["transform"]["scale"] == 0.5
    a, b = cs.scene_info(path, d), cs.scene_info(out, d)
    for f in ("n_cloths", "n_obstacles", "n_handles", "n_cloth_nodes", "n_cloth_faces", "frame_steps"):
        assert getattr(a, f) == getattr(b, f), f
    with pytest.raises(ClothError, match="CS_E_ARG"):
        cs.write_config(path, out, ["a.obj", "b.obj"])


def test_cpp_mirror_compiles(tmp_path):
    cs.load_library()
    assert os.path.exists(build_host_mirror(str(tmp_path)))


@pytest.mark.gpu
def test_cpp_mirror_runs_frames(tmp_path):
    import subprocess
    exe = build_host_mirror(str(tmp_path))
    path = write_scene(str(tmp_path))
    out = tmp_path / "frames"
    out.mkdir()
    r = subprocess.run([exe, path, str(tmp_path), str(out), "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    files = sorted(os.listdir(out))
    assert files == ["config.json", "frame0_cloth0.obj", "frame1_cloth0.obj", "frame2_cloth0.obj"]
    x0 = obj_rows(out / "frame0_cloth0.obj", "v")
    assert np.allclose(x0, expected_scene()[0], atol=1e-6)
    x2 = obj_rows(out / "frame2_cloth0.obj", "v")
    assert np.all(np.isfinite(x2)) and not np.allclose(x2, x0)  # 16 steps of gravity moved it
    # the reference's OBJ: v / nv per node, vt per material-space vertex, f a/b faces
    assert len(obj_rows(out / "frame2_cloth0.obj", "nv")) == 6
    # the sheet has no vt lines: one material-space vertex per polygon corner at scale * (x, y)
    vt, raw = obj_rows(out / "frame2_cloth0.obj", "vt"), obj_rows(tmp_path / "sheet.obj", "v")
    assert len(vt) == 8 and all(np.any(np.all(2 * raw[:, :2] == u, axis=1)) for u in vt)
    assert all("/" in l for l in open(out / "frame2_cloth0.obj") if l.startswith("f "))


@pytest.mark.gpu
def test_cpp_mirror_remeshes_at_frame_ends(tmp_path):
    """With a full "remeshing" block and no "disable", every frame ends with cs_remesh: the saved
    frames hold the refined sheet (a valid manifold, the material area kept)."""
    import subprocess
    exe = build_host_mirror(str(tmp_path))
    path = write_scene(str(tmp_path), remeshing_config())
    out = tmp_path / "frames"
    out.mkdir()
    r = subprocess.run([exe, path, str(tmp_path), str(out), "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n = [len(obj_rows(out / ("frame%d_cloth0.obj" % k), "v")) for k in range(3)]
    print("remeshed sheet: nodes per saved frame %s" % n)
    assert n[0] == 6 and n[1] > 6 and n[2] > 6
    vt = obj_rows(out / "frame2_cloth0.obj", "vt")
    faces = [[int(t.split("/")[1]) - 1 for t in l.split()[1:]] for l in open(out / "frame2_cloth0.obj")
             if l.startswith("f ")]
    area = sum(0.5 * abs(np.cross(vt[b] - vt[a], vt[c] - vt[a])) for a, b, c in faces)
    assert area == pytest.approx(2.0, rel=1e-5)  # the 1 x 0.5 sheet scaled by 2
    assert np.all(np.isfinite(obj_rows(out / "frame2_cloth0.obj", "v")))


@pytest.mark.gpu
def test_cpp_mirror_overlapped_saves_equal_direct_ones(tmp_path):
    """Simulator::saveOverlapped (device snapshot, files written on a writer thread while the
    next frame steps) writes byte for byte the files of Simulator::save, here on the remeshing
    sheet, whose node count changes at every frame end."""
    import subprocess
    exe = build_host_mirror(str(tmp_path))
    path = write_scene(str(tmp_path), remeshing_config())
    outs = []
    for mode in ("direct", "overlapped"):
        out = tmp_path / mode
        out.mkdir()
        r = subprocess.run([exe, path, str(tmp_path), str(out), "3", mode], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(out)
    names = sorted(os.listdir(outs[0]))
    assert names == sorted(os.listdir(outs[1])) and len(names) == 5
    for n in names:
        a, b = open(outs[0] / n).read(), open(outs[1] / n).read()
        if n == "config.json":
            a, b = a.replace(str(outs[0]), "OUT"), b.replace(str(outs[1]), "OUT")
        assert a == b, n


@pytest.mark.gpu
def test_cpp_mirror_resume_continues_the_run(tmp_path):
    """Two frames saved, resumed for a third (Resume mode reads the saved config.json and frame
    files), against three frames straight through.  Without handles (a resumed run anchors its
    handles at the loaded positions, as the reference's does) the two agree."""
    import subprocess
    exe = build_host_mirror(str(tmp_path))
    path = write_scene(str(tmp_path), CONFIG.replace('"handles": [{"nodes": [0, 2]}]', '"handles": []'))
    out, straight = tmp_path / "frames", tmp_path / "straight"
    out.mkdir()
    straight.mkdir()
    r = subprocess.run([exe, path, str(tmp_path), str(out), "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe, "--resume", str(out), str(tmp_path), "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "resumed at frame 2 step 16" in r.stdout
    r = subprocess.run([exe, path, str(tmp_path), str(straight), "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for tag in ("v", "nv"):
        a, b = obj_rows(out / "frame3_cloth0.obj", tag), obj_rows(straight / "frame3_cloth0.obj", tag)
        print("resumed vs straight frame 3 %s: max diff %.3g" % (tag, np.abs(a - b).max()))
        assert np.abs(a - b).max() <= 1e-6


def obj_rows(path, tag):
    return np.array([[float(t) for t in l.split()[1:]] for l in open(path) if l.split() and l.split()[0] == tag])


def test_instances_scene_generator():
    from clothsimulator_amd import scenes
    from clothsimulator_amd._cabi import CS_FLAG_INSTANCES
    spec = scenes.instances(count=256, seed=0)
    sides = np.array([int(round(np.sqrt(c.mesh.n_nodes))) for c in spec.cloths])
    assert len(spec.cloths) == len(spec.obstacles) == 256
    assert sides.min() >= 17 and sides.max() <= 129 and len(np.unique(sides)) > 80
    assert spec.params.flags & CS_FLAG_INSTANCES and spec.params.dt == np.float32(0.005)
    # every cloth sits above its own sphere and away from the others
    for k in (0, 17, 255):
        c = spec.cloths[k].mesh.x.mean(axis=0)
        o = spec.obstacles[k].x
        r = (o.max(axis=0) - o.min(axis=0))[0] / 2
        assert 0.2 - 1e-9 <= r <= 0.4 + 1e-9
        assert np.linalg.norm(o.mean(axis=0)[:2] - c[:2]) <= 0.2 * np.sqrt(2) + 1e-9
        assert spec.cloths[k].mesh.x[:, 2].min() > o[:, 2].max()
    assert scenes.instances(count=256, seed=0).n_cloth_nodes == spec.n_cloth_nodes   # deterministic
