// cs_scene.cpp — the reference's scene files over the C-ABI.
//
// Reads a configuration JSON of the reference (src/Simulator.cu:3-135), its
// OBJ meshes (src/Mesh.cu:305-347, with polygon fans chosen by the largest
// minimum angle, src/Mesh.cu:27-54), "transform" blocks
// (src/Transformation.cu:9-14,82-88, quaternion rotation src/Quaternion.cu:7-16,
// 76-78) and material JSON files (src/Material.cu:3-41), then creates a
// context and adds the cloths, motions (keyframe lists or a .mot file,
// src/Simulator.cu:39-79), obstacles (listed, or the "obstacles": "<pattern>"
// sequence, each moving with the motion of its index, src/Simulator.cu:90-101)
// and handles in the reference's order; each cloth's "remeshing" block is set on
// the context unless the "disable" list names "remeshing".
#include <cmath>
#include <climits>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/clothsim.h"
#include "cs_json.hpp"
#include "cs_motion.hpp"

namespace {

thread_local std::string g_scene_error;

struct F3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
F3 f3(float a, float b, float c) { F3 r; r.x = a; r.y = b; r.z = c; return r; }
F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
F3 operator*(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
float norm(F3 a) { return std::sqrt(dot(a, a)); }
F3 normalized(F3 a) { return a * (1.0f / norm(a)); }

struct SceneError : std::runtime_error {
    int status;
    SceneError(int st, const std::string& m) : std::runtime_error(m), status(st) {}
};

float num(const cs::json::Value& v, float d, const char* what) { return v.is_null() ? d : (float)v.as_number(what); }
int inum(const cs::json::Value& v, int d, const char* what) { return v.is_null() ? d : (int)v.as_number(what); }
F3 vec3(const cs::json::Value& v, const char* what) {
    if (v.is_null()) return F3();
    return f3((float)v[0].as_number(what), (float)v[1].as_number(what), (float)v[2].as_number(what));
}

// "transform": {"scale": s, "translate": [x,y,z], "rotate": [deg, ax, ay, az]}
struct Transform {
    float scaling = 1.0f;
    F3 translation;
    float qs = 1.0f;  // rotation quaternion (s, v)
    F3 qv;
    explicit Transform(const cs::json::Value& j) {
        scaling = num(j["scale"], 1.0f, "transform.scale");
        translation = vec3(j["translate"], "transform.translate");
        const cs::json::Value& r = j["rotate"];
        float deg = 0.0f;
        F3 axis;
        if (!r.is_null()) {
            deg = (float)r[0].as_number("transform.rotate");
            axis = f3((float)r[1].as_number("transform.rotate"), (float)r[2].as_number("transform.rotate"),
                      (float)r[3].as_number("transform.rotate"));
        }
        const float angle = (float)(deg * M_PI / 180.0f);
        if (angle == 0.0f) {
            qs = 1.0f;
            qv = F3();
        } else {
            qs = std::cos(0.5f * angle);
            qv = std::sin(0.5f * angle) * normalized(axis);
        }
    }
    F3 rotate(F3 x) const { return x * (qs * qs - dot(qv, qv)) + 2.0f * qv * dot(qv, x) + 2.0f * cross(qv, x) * qs; }
    F3 point(F3 p) const { return scaling * rotate(p) + translation; }
    F3 vector(F3 v) const { return rotate(v); }
    cs::Xf xf() const { return cs::Xf{scaling, translation.x, translation.y, translation.z, qs, qv.x, qv.y, qv.z}; }
};

std::string read_file(const std::string& path, const char* what) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw SceneError(CS_E_ARG, std::string("failed to open ") + what + " file: " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::string resolve(const std::string& root, const std::string& p) {
    if (p.empty() || p[0] == '/' || root.empty()) return p;
    return root.back() == '/' ? root + p : root + "/" + p;
}

cs::json::Value parse_file(const std::string& path, const char* what) {
    try {
        return cs::json::parse(read_file(path, what));
    } catch (const SceneError&) {
        throw;
    } catch (const std::exception& e) {
        throw SceneError(CS_E_FORMAT, std::string(what) + " " + path + ": " + e.what());
    }
}

struct ObjMesh {
    std::vector<float> x, v, uv;
    std::vector<int> fx, fu;
};

float corner_angle(F3 x0, F3 x1, F3 x2) {
    F3 e1 = normalized(x1 - x0), e2 = normalized(x2 - x0);
    float c = dot(e1, e2);
    c = c < -1.0f ? -1.0f : (c > 1.0f ? 1.0f : c);
    return std::acos(c);
}

// fan from the corner whose fan has the largest minimum angle (first on ties)
void triangulate(const std::vector<F3>& x, const std::vector<int>& px, const std::vector<int>& pu, ObjMesh& m) {
    const int n = (int)px.size();
    float best = -INFINITY;
    int index = 0;
    for (int i = 0; i < n; i++) {
        float mn = INFINITY;
        F3 x0 = x[px[i]];
        for (int j = 2; j < n; j++) {
            F3 x1 = x[px[(i + j - 1) % n]], x2 = x[px[(i + j) % n]];
            mn = std::min(mn, std::min(corner_angle(x0, x1, x2), std::min(corner_angle(x1, x2, x0), corner_angle(x2, x0, x1))));
        }
        if (mn > best) {
            best = mn;
            index = i;
        }
    }
    for (int i = 2; i < n; i++) {
        const int a = index, b = (index + i - 1) % n, c = (index + i) % n;
        m.fx.push_back(px[a]); m.fx.push_back(px[b]); m.fx.push_back(px[c]);
        m.fu.push_back(pu[a]); m.fu.push_back(pu[b]); m.fu.push_back(pu[c]);
    }
}

ObjMesh load_obj(const std::string& path, const Transform& T) {
    std::istringstream in(read_file(path, "mesh"));
    ObjMesh m;
    std::vector<F3> raw, xs, vs;
    std::vector<float> uv;
    std::string line;
    int lineNo = 0;
    auto bad = [&](const char* what) {
        throw SceneError(CS_E_FORMAT, "mesh " + path + " line " + std::to_string(lineNo) + ": " + what);
    };
    while (std::getline(in, line)) {
        lineNo++;
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag)) continue;
        if (tag == "v") {
            float a, b, c;
            if (!(ls >> a >> b >> c)) bad("bad vertex");
            raw.push_back(f3(a, b, c));
            xs.push_back(T.point(f3(a, b, c)));
        } else if (tag == "nv") {
            float a, b, c;
            if (!(ls >> a >> b >> c)) bad("bad velocity");
            vs.push_back(T.vector(f3(a, b, c)));
        } else if (tag == "vt") {
            float a, b;
            if (!(ls >> a >> b)) bad("bad texture vertex");
            uv.push_back(T.scaling * a);
            uv.push_back(T.scaling * b);
        } else if (tag == "f") {
            const bool slashed = line.find('/') != std::string::npos;
            std::vector<int> px, pu;
            std::string tok;
            while (ls >> tok) {
                if (slashed) {
                    size_t s = tok.find('/');
                    if (s == std::string::npos) bad("mixed face formats");
                    px.push_back(std::atoi(tok.substr(0, s).c_str()) - 1);
                    pu.push_back(std::atoi(tok.substr(s + 1).c_str()) - 1);
                } else {
                    const int xi = std::atoi(tok.c_str()) - 1;
                    if (xi < 0 || xi >= (int)raw.size()) bad("face index out of range");
                    uv.push_back(T.scaling * raw[xi].x);  // uv = raw (x, y) of the node (src/Mesh.cu:334-338)
                    uv.push_back(T.scaling * raw[xi].y);
                    px.push_back(xi);
                    pu.push_back((int)(uv.size() / 2) - 1);
                }
            }
            if (px.size() < 3) bad("face with fewer than 3 corners");
            for (size_t k = 0; k < px.size(); k++)
                if (px[k] < 0 || px[k] >= (int)xs.size() || pu[k] < 0) bad("face index out of range");
            triangulate(xs, px, pu, m);
        }
    }
    for (size_t k = 0; k < m.fu.size(); k++)
        if (m.fu[k] >= (int)(uv.size() / 2)) throw SceneError(CS_E_FORMAT, "mesh " + path + ": texture index out of range");
    for (const F3& p : xs) { m.x.push_back(p.x); m.x.push_back(p.y); m.x.push_back(p.z); }
    m.v.assign(m.x.size(), 0.0f);
    for (size_t i = 0; i < vs.size() && i < xs.size(); i++) {
        m.v[3 * i] = vs[i].x; m.v[3 * i + 1] = vs[i].y; m.v[3 * i + 2] = vs[i].z;
    }
    m.uv = std::move(uv);
    if (m.fx.empty()) throw SceneError(CS_E_FORMAT, "mesh " + path + " has no faces");
    return m;
}

cs_mesh_in mesh_in(const ObjMesh& m) {
    cs_mesh_in mi;
    mi.n_nodes = (int)(m.x.size() / 3);
    mi.x = m.x.data();
    mi.v = m.v.data();
    mi.n_verts = (int)(m.uv.size() / 2);
    mi.uv = m.uv.data();
    mi.n_faces = (int)(m.fx.size() / 3);
    mi.face_nodes = m.fx.data();
    mi.face_verts = m.fu.data();
    return mi;
}

cs_material_in material_in(const cs::json::Value& j, const std::string& root) {
    cs_material_in mat;
    std::memset(&mat, 0, sizeof(mat));
    mat.thicken = num(j["thicken"], 1.0f, "materials.thicken");
    mat.density_mult = num(j["density_mult"], 1.0f, "materials.density_mult");
    mat.stretching_mult = num(j["stretching_mult"], 1.0f, "materials.stretching_mult");
    mat.bending_mult = num(j["bending_mult"], 1.0f, "materials.bending_mult");
    if (!j["data"].is_string()) throw SceneError(CS_E_FORMAT, "cloth materials need a \"data\" file");
    const std::string path = resolve(root, j["data"].as_string("materials.data"));
    cs::json::Value d = parse_file(path, "material");
    mat.density = (float)d["density"].as_number("density");
    for (int i = 0; i < 6; i++)
        for (int k = 0; k < 4; k++) mat.stretching[i][k] = (float)d["stretching"][i][k].as_number("stretching");
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 5; k++) mat.bending[i][k] = (float)d["bending"][i][k].as_number("bending");
    return mat;
}

struct MotionKeys {
    std::vector<float> t;
    std::vector<cs_transform> x;
};

// "motions" (src/Simulator.cu:39-79): a list of keyframe lists (src/Motion.cu:3-20), or a
// {"motfile", "fps", "transform"} block naming a .mot file of per-body frames whose keys are
// transform * (position, rotation quaternion w x y z) at i / fps (src/Motion.cu:22-37).
std::vector<MotionKeys> read_motions(const cs::json::Value& j, const std::string& root) {
    std::vector<MotionKeys> out;
    const cs::json::Value& mj = j["motions"];
    if (mj.is_null()) return out;
    if (mj.is_array()) {
        for (size_t m = 0; m < mj.size(); m++) {
            MotionKeys K;
            for (size_t k = 0; k < mj[m].size(); k++) {
                K.t.push_back((float)mj[m][k]["time"].as_number("motion time"));
                K.x.push_back(cs::xf_to(Transform(mj[m][k]["transform"]).xf()));
            }
            if (K.t.empty()) throw SceneError(CS_E_FORMAT, "motion " + std::to_string(m) + " has no keyframes");
            out.push_back(std::move(K));
        }
        return out;
    }
    if (!mj.is_object() || !mj["motfile"].is_string())
        throw SceneError(CS_E_FORMAT, "\"motions\" is neither a list nor a {\"motfile\"} block");
    const float fps = num(mj["fps"], 30.0f, "motions.fps");
    if (!(fps > 0.0f)) throw SceneError(CS_E_FORMAT, "motions.fps must be positive");
    const cs::Xf T = Transform(mj["transform"]).xf();
    const std::string path = resolve(root, mj["motfile"].as_string("motions.motfile"));
    std::istringstream in(read_file(path, "motion"));
    std::string a, b;
    int frames = 0;
    if (!(in >> a >> frames) || frames < 1) throw SceneError(CS_E_FORMAT, "motion file " + path + ": bad frame count");
    const float dt = 1.0f / fps;
    while (in >> a >> b) {  // "body[k] position", frames x (x y z), "body[k] rotation", frames x (w x y z)
        std::vector<float> pos(3 * (size_t)frames), rot(4 * (size_t)frames);
        for (float& f : pos)
            if (!(in >> f)) throw SceneError(CS_E_FORMAT, "motion file " + path + ": truncated positions");
        if (!(in >> a >> b)) throw SceneError(CS_E_FORMAT, "motion file " + path + ": missing rotations");
        for (float& f : rot)
            if (!(in >> f)) throw SceneError(CS_E_FORMAT, "motion file " + path + ": truncated rotations");
        MotionKeys K;
        for (int i = 0; i < frames; i++) {
            const cs::Xf body{1.0f, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2],
                              rot[4 * i], rot[4 * i + 1], rot[4 * i + 2], rot[4 * i + 3]};
            K.t.push_back((float)i * dt);
            K.x.push_back(cs::xf_to(cs::xf_compose(T, body)));
        }
        out.push_back(std::move(K));
    }
    if (out.empty()) throw SceneError(CS_E_FORMAT, "motion file " + path + " holds no bodies");
    return out;
}

// The reference's stringFormat(pattern, i) (src/Simulator.cu:134-141) for the one integer
// conversion a sequence pattern may hold ("%d", "%04d", ...); anything else is refused.
std::string format_index(const std::string& pat, int i) {
    std::string out;
    int conversions = 0;
    for (size_t k = 0; k < pat.size(); k++) {
        if (pat[k] != '%') { out += pat[k]; continue; }
        if (k + 1 < pat.size() && pat[k + 1] == '%') { out += '%'; k++; continue; }
        size_t e = k + 1;
        bool zero = false;
        if (e < pat.size() && pat[e] == '0') { zero = true; e++; }
        int width = 0;
        while (e < pat.size() && pat[e] >= '0' && pat[e] <= '9' && width < 100) width = 10 * width + (pat[e++] - '0');
        if (e >= pat.size() || pat[e] != 'd' || ++conversions > 1)
            throw SceneError(CS_E_FORMAT, "obstacle pattern \"" + pat + "\" needs exactly one %d conversion");
        std::string d = std::to_string(i);
        while ((int)d.size() < width) d.insert(d.begin(), zero ? '0' : ' ');
        out += d;
        k = e;
    }
    if (conversions != 1) throw SceneError(CS_E_FORMAT, "obstacle pattern \"" + pat + "\" needs exactly one %d conversion");
    return out;
}

bool file_exists(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    return (bool)f;
}

struct Scene {
    cs_scene_info info;
    cs_params params;
    std::vector<ObjMesh> cloths, obstacles;
    std::vector<int> obstacleMotion;
    std::vector<MotionKeys> motions;
    std::vector<cs_material_in> materials;
    struct Handles { int cloth, motion; std::vector<int> nodes; };
    std::vector<Handles> handles;
    std::vector<cs_remeshing> remeshing;  // per cloth: its "remeshing" block (defaults when absent)
};

// Remeshing ctor (src/Remeshing.cu:3-19)
cs_remeshing remeshing_in(const cs::json::Value& r) {
    cs_remeshing o;
    cs_default_remeshing(&o);
    o.refine_angle = num(r["refine_angle"], INFINITY, "remeshing.refine_angle");
    o.refine_velocity = num(r["refine_velocity"], INFINITY, "remeshing.refine_velocity");
    o.refine_compression = num(r["refine_compression"], INFINITY, "remeshing.refine_compression");
    o.rib_stiffening = num(r["rib_stiffening"], 1.0f, "remeshing.rib_stiffening");
    if (!r["size"].is_null()) {
        o.size_min = (float)r["size"][0].as_number("remeshing.size");
        o.size_max = (float)r["size"][1].as_number("remeshing.size");
    }
    o.aspect_min = num(r["aspect_min"], -INFINITY, "remeshing.aspect_min");
    o.flip_threshold = num(r["flip_threshold"], 1e-2f, "remeshing.flip_threshold");
    return o;
}

Scene read_scene(const std::string& path, const std::string& root, const cs_params* base) {
    Scene S;
    std::memset(&S.info, 0, sizeof(S.info));
    cs::json::Value j = parse_file(path, "configuration");
    if (!j.is_object()) throw SceneError(CS_E_FORMAT, "configuration " + path + " is not a JSON object");
    S.info.frame_time = (float)j["frame_time"].as_number("frame_time");
    S.info.frame_steps = (int)j["frame_steps"].as_number("frame_steps");
    if (S.info.frame_steps <= 0 || !(S.info.frame_time > 0.0f)) throw SceneError(CS_E_FORMAT, "frame_time and frame_steps must be positive");
    S.info.end_time = num(j["end_time"], INFINITY, "end_time");
    S.info.end_frame = inum(j["end_frame"], INT_MAX, "end_frame");
    if (base) S.params = *base; else cs_default_params(&S.params);
    cs_params& P = S.params;
    P.dt = S.info.frame_time / S.info.frame_steps;
    P.friction = num(j["friction"], 0.6f, "friction");
    P.obstacle_friction = num(j["obs_friction"], 0.3f, "obs_friction");
    F3 g = vec3(j["gravity"], "gravity");
    P.gravity[0] = g.x; P.gravity[1] = g.y; P.gravity[2] = g.z;
    const cs::json::Value& w = j["wind"];
    P.wind_density = num(w["density"], 1.0f, "wind.density");
    P.wind_drag = num(w["drag"], 0.0f, "wind.drag");
    F3 wv = vec3(w["velocity"], "wind.velocity");
    P.wind_velocity[0] = wv.x; P.wind_velocity[1] = wv.y; P.wind_velocity[2] = wv.z;
    const cs::json::Value& mg = j["magic"];
    P.handle_stiffness = num(mg["handle_stiffness"], 1e3f, "magic.handle_stiffness");
    P.collision_stiffness = num(mg["collision_stiffness"], 1e9f, "magic.collision_stiffness");
    P.repulsion_thickness = num(mg["repulsion_thickness"], 1e-3f, "magic.repulsion_thickness");
    P.collision_thickness = num(mg["projection_thickness"], 0.1f * P.repulsion_thickness, "magic.projection_thickness");

    const cs::json::Value& cl = j["cloths"];
    if (!cl.is_array() || cl.size() == 0) throw SceneError(CS_E_FORMAT, "configuration has no \"cloths\" array");
    for (size_t i = 0; i < cl.size(); i++) {
        const cs::json::Value& c = cl[i];
        if (!c["mesh"].is_string()) throw SceneError(CS_E_FORMAT, "cloth " + std::to_string(i) + " has no mesh");
        Transform T(c["transform"]);
        S.cloths.push_back(load_obj(resolve(root, c["mesh"].as_string("mesh")), T));
        S.materials.push_back(material_in(c["materials"].is_array() ? c["materials"][0] : c["materials"], root));
        S.remeshing.push_back(remeshing_in(c["remeshing"]));
    }
    // every cloth remeshes once per frame unless "disable" lists "remeshing" (src/Simulator.cu:858)
    S.info.remeshing = 1;
    const cs::json::Value& dis = j["disable"];
    for (size_t i = 0; i < dis.size(); i++)
        if (dis[i].is_string() && dis[i].as_string("disable") == "remeshing") S.info.remeshing = 0;
    S.motions = read_motions(j, root);
    const int nm = (int)S.motions.size();
    if (nm > CS_MAX_MOTIONS) throw SceneError(CS_E_CAPACITY, "more than CS_MAX_MOTIONS motions");
    const cs::json::Value& ob = j["obstacles"];
    if (ob.is_string()) {  // obstacle i = pattern % i, moving with motion i (src/Simulator.cu:95-101)
        const std::string pat = ob.as_string("obstacles");
        for (int i = 0;; i++) {
            const std::string p = resolve(root, format_index(pat, i));
            if (!file_exists(p)) break;
            if (i >= nm) throw SceneError(CS_E_FORMAT, "obstacle sequence file " + p + " has no motion " + std::to_string(i));
            S.obstacles.push_back(load_obj(p, Transform(cs::json::Value())));
            S.obstacleMotion.push_back(i);
        }
    } else {
        for (size_t i = 0; i < ob.size(); i++) {
            const cs::json::Value& o = ob[i];
            const int m = inum(o["motion"], -1, "obstacle motion");
            if (m < -1 || m >= nm) throw SceneError(CS_E_FORMAT, "obstacle " + std::to_string(i) + " names a missing motion");
            if (!o["mesh"].is_string()) throw SceneError(CS_E_FORMAT, "obstacle " + std::to_string(i) + " has no mesh");
            S.obstacles.push_back(load_obj(resolve(root, o["mesh"].as_string("mesh")), Transform(o["transform"])));
            S.obstacleMotion.push_back(m);
        }
    }
    const cs::json::Value& hs = j["handles"];
    for (size_t i = 0; i < hs.size(); i++) {
        const cs::json::Value& h = hs[i];
        const int cloth = inum(h["cloth"], 0, "handle cloth");
        if (cloth < 0 || cloth >= (int)S.cloths.size()) throw SceneError(CS_E_FORMAT, "handle on a missing cloth");
        const int motion = inum(h["motion"], -1, "handle motion");
        if (motion < -1 || motion >= nm) throw SceneError(CS_E_FORMAT, "handle " + std::to_string(i) + " names a missing motion");
        std::vector<int> nodes;
        const int nn = (int)(S.cloths[cloth].x.size() / 3);
        for (size_t k = 0; k < h["nodes"].size(); k++) {
            const int n = (int)h["nodes"][k].as_number("handle node");
            if (n < 0 || n >= nn) throw SceneError(CS_E_FORMAT, "handle node " + std::to_string(n) + " out of range");
            nodes.push_back(n);
        }
        S.handles.push_back(Scene::Handles{cloth, motion, std::move(nodes)});
    }
    S.info.n_cloths = (int)S.cloths.size();
    S.info.n_obstacles = (int)S.obstacles.size();
    S.info.n_handles = (int)S.handles.size();
    S.info.n_motions = nm;
    for (int m : S.obstacleMotion) S.info.n_moving_obstacles += m >= 0;
    for (const ObjMesh& m : S.cloths) {
        S.info.n_cloth_nodes += (int)(m.x.size() / 3);
        S.info.n_cloth_faces += (int)(m.fx.size() / 3);
    }
    return S;
}

}  // namespace

extern "C" cs_status cs_load_scene(const char* config_path, const char* root_dir, const cs_params* base,
                                   cs_ctx** ctx_out, cs_scene_info* info) {
    g_scene_error.clear();
    if (ctx_out) *ctx_out = nullptr;
    if (!config_path) return CS_E_ARG;
    cs_ctx* C = nullptr;
    try {
        Scene S = read_scene(config_path, root_dir ? root_dir : "", base);
        if (info) *info = S.info;
        if (!ctx_out) return CS_OK;
        cs_status st = cs_create(&S.params, &C);
        if (st != CS_OK) throw SceneError(st, "cs_create failed");
        auto check = [&](cs_status s, const char* what) {
            if (s != CS_OK) throw SceneError(s, std::string(what) + ": " + cs_last_error(C));
        };
        std::vector<int> ids;
        for (size_t i = 0; i < S.cloths.size(); i++) {
            cs_mesh_in mi = mesh_in(S.cloths[i]);
            int id = -1;
            check(cs_add_cloth(C, &mi, &S.materials[i], &id), "cs_add_cloth");
            ids.push_back(id);
        }
        for (const MotionKeys& K : S.motions)
            check(cs_add_motion(C, K.t.data(), K.x.data(), (int)K.t.size(), nullptr), "cs_add_motion");
        for (size_t i = 0; i < S.obstacles.size(); i++) {
            cs_mesh_in mi = mesh_in(S.obstacles[i]);
            int id = -1;
            check(cs_add_obstacle(C, &mi, &id), "cs_add_obstacle");
            if (S.obstacleMotion[i] >= 0) check(cs_set_obstacle_motion(C, id, S.obstacleMotion[i]), "cs_set_obstacle_motion");
        }
        for (const auto& h : S.handles)
            check(cs_add_moving_handles(C, ids[h.cloth], h.nodes.data(), (int)h.nodes.size(), h.motion),
                  "cs_add_moving_handles");
        if (S.info.remeshing)
            for (size_t i = 0; i < S.cloths.size(); i++)
                check(cs_set_remeshing(C, ids[i], &S.remeshing[i]), "cs_set_remeshing");
        *ctx_out = C;
        return CS_OK;
    } catch (const SceneError& e) {
        g_scene_error = e.what();
        if (C) cs_destroy(C);
        return (cs_status)e.status;
    } catch (const std::exception& e) {
        g_scene_error = e.what();
        if (C) cs_destroy(C);
        return CS_E_FORMAT;
    }
}

extern "C" const char* cs_scene_error(void) { return g_scene_error.c_str(); }

extern "C" cs_status cs_write_config(const char* config_path, const char* out_path, const char* const* cloth_meshes,
                                     int n_cloths) {
    g_scene_error.clear();
    if (!config_path || !out_path || n_cloths < 0 || (n_cloths > 0 && !cloth_meshes)) return CS_E_ARG;
    try {
        cs::json::Value j = parse_file(config_path, "configuration");
        if (!j.is_object()) throw SceneError(CS_E_FORMAT, std::string("configuration ") + config_path + " is not a JSON object");
        cs::json::Value* cloths = nullptr;
        for (auto& kv : j.obj)
            if (kv.first == "cloths") cloths = &kv.second;
        if (!cloths || !cloths->is_array() || (int)cloths->arr.size() != n_cloths)
            throw SceneError(CS_E_ARG, "the configuration's \"cloths\" do not match the cloth count");
        for (int i = 0; i < n_cloths; i++) {
            cs::json::Value& c = cloths->arr[i];
            if (!c.is_object() || !cloth_meshes[i]) throw SceneError(CS_E_ARG, "cloth " + std::to_string(i));
            std::vector<std::pair<std::string, cs::json::Value>> kept;
            bool mesh = false;
            for (auto& kv : c.obj) {
                if (kv.first == "transform") continue;
                if (kv.first == "mesh") { kv.second = cs::json::Value(); kv.second.kind = cs::json::Value::String;
                                          kv.second.str = cloth_meshes[i]; mesh = true; }
                kept.push_back(kv);
            }
            if (!mesh) {
                cs::json::Value m;
                m.kind = cs::json::Value::String;
                m.str = cloth_meshes[i];
                kept.emplace_back("mesh", m);
            }
            c.obj = std::move(kept);
        }
        std::ofstream out(out_path, std::ios::binary);
        if (!out) throw SceneError(CS_E_ARG, std::string("cannot write ") + out_path);
        out << cs::json::dump(j);
        if (!out) throw SceneError(CS_E_ARG, std::string("cannot write ") + out_path);
        return CS_OK;
    } catch (const SceneError& e) {
        g_scene_error = e.what();
        return (cs_status)e.status;
    } catch (const std::exception& e) {
        g_scene_error = e.what();
        return CS_E_FORMAT;
    }
}

extern "C" cs_status cs_scene_motion(const char* config_path, const char* root_dir, int motion_id, float time,
                                     cs_transform* out) {
    g_scene_error.clear();
    if (!config_path || !out) return CS_E_ARG;
    try {
        cs::json::Value j = parse_file(config_path, "configuration");
        std::vector<MotionKeys> ms = read_motions(j, root_dir ? root_dir : "");
        if (motion_id < 0 || motion_id >= (int)ms.size()) throw SceneError(CS_E_ARG, "no motion " + std::to_string(motion_id));
        return cs_motion_eval(ms[motion_id].t.data(), ms[motion_id].x.data(), (int)ms[motion_id].t.size(), time, out);
    } catch (const SceneError& e) {
        g_scene_error = e.what();
        return (cs_status)e.status;
    } catch (const std::exception& e) {
        g_scene_error = e.what();
        return CS_E_FORMAT;
    }
}
