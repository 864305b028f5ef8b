// clothsim.hpp — header-only C++ mirror of the reference's host classes over the C-ABI.
//
// Names follow the reference so that host code reads the same:
//   Simulator(path)               Simulator::Simulator (src/Simulator.cu:3-135), scene file only
//   Simulator::physicsStep()      src/Simulator.cu:571-600
//   Simulator::collisionStep()    src/Simulator.cu:602-684
//   Simulator::simulateStep()     src/Simulator.cu:838-897 (a frame's last step remeshes, separation
//                                 step included, unless the scene disables remeshing)
//   Simulator::getCloths() / getObstacles(), Cloth::getMesh(), Obstacle::getMesh()
//   Mesh::save(path)              OBJ with "v"/"nv" per node, "vt" per uv vertex and "f a/b" faces at
//                                 precision 20 (src/Mesh.cu:454-507), readable by cs_load_scene
//   Simulator::save(directory)    frame<N>_cloth<i>.obj + config.json (src/Simulator.cu:950-957)
//   Simulator::simulateOffline    src/Simulator.cu:1018-1043 (a save per frame until finished())
//   Simulator(Resume, directory)  Resume mode: directory/config.json, nFrames = lastFrame(),
//   Simulator::resumeOffline      nSteps = nFrames * frameSteps (src/Simulator.cu:1045-1083)
//   Simulator::saveOverlapped     save() with the frame's nodes snapshotted on the device and its files
//                                 written on a writer thread while the steps go on (no counterpart)
// Errors throw clothsim::Error (the reference prints and exit(1)s instead).
#pragma once
#include <charconv>
#include <cstdio>
#include <fstream>
#include <future>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "clothsim.h"

namespace clothsim {

struct Error : std::runtime_error {
    cs_status status;
    Error(cs_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

// A snapshot of one mesh's node state (positions, previous positions, velocities) and faces.
struct Mesh {
    std::vector<float> x, x0, v;  // 3 per node
    std::vector<int> faces;       // 3 node ids per face (mesh-local)
    std::vector<int> faceVerts;   // 3 uv-vertex ids per face (mesh-local)
    std::vector<float> uv;        // 2 per uv vertex (material space)
    int nodes() const { return (int)(x.size() / 3); }
    // Numbers are formatted by std::to_chars, general format at precision 20: the text of a
    // stream at setprecision(20) (printf's %.20g of the float), without the stream's cost.
    void save(const std::string& path) const {
        std::string o;
        o.reserve(nodes() * 150 + faces.size() * 12 + uv.size() * 25);
        auto num = [&o](float f) {
            char b[64];
            o.append(b, std::to_chars(b, b + sizeof b, (double)f, std::chars_format::general, 20).ptr);
        };
        auto idx = [&o](int k) {
            char b[16];
            o.append(b, std::to_chars(b, b + sizeof b, k).ptr);
        };
        auto row = [&](const char* tag, const float* p, int n) {
            o += tag;
            for (int c = 0; c < n; c++) {
                o += ' ';
                num(p[c]);
            }
            o += '\n';
        };
        for (int i = 0; i < nodes(); i++) {
            row("v", &x[3 * (size_t)i], 3);
            row("nv", &v[3 * (size_t)i], 3);
        }
        for (size_t k = 0; k + 1 < uv.size(); k += 2) row("vt", &uv[k], 2);
        for (size_t f = 0; f + 2 < faces.size(); f += 3) {
            o += 'f';
            for (int c = 0; c < 3; c++) {
                o += ' ';
                idx(faces[f + c] + 1);
                o += '/';
                idx(faceVerts[f + c] + 1);
            }
            o += '\n';
        }
        std::ofstream out(path, std::ios::binary);
        if (!out) throw Error(CS_E_ARG, "cannot write " + path);
        out.write(o.data(), (std::streamsize)o.size());
        if (!out) throw Error(CS_E_ARG, "cannot write " + path);
    }
};

class Simulator;

// One cloth or obstacle of a Simulator (kind = CS_MESH_CLOTH or CS_MESH_OBSTACLE).
class MeshHandle {
  public:
    MeshHandle(cs_ctx* ctx, int kind, int id) : ctx_(ctx), kind_(kind), id_(id) {}
    Mesh getMesh() const {
        Mesh m = topology();
        const size_t nn = 3 * (size_t)nodes();
        m.x.resize(nn);
        m.x0.resize(nn);
        m.v.resize(nn);
        check(cs_get_state(ctx_, kind_, id_, m.x.data(), m.x0.data(), m.v.data()), "cs_get_state");
        return m;
    }
    // the faces and material-space vertices alone (host data)
    Mesh topology() const {
        int nn = 0, nf = 0, ne = 0;
        check(cs_get_counts(ctx_, kind_, id_, &nn, &nf, &ne), "cs_get_counts");
        Mesh m;
        int nv = 0;
        check(cs_get_topology(ctx_, kind_, id_, nullptr, nullptr, nullptr, &nv), "cs_get_topology");
        m.faces.resize(3 * (size_t)nf);
        m.faceVerts.resize(3 * (size_t)nf);
        m.uv.resize(2 * (size_t)nv);
        check(cs_get_topology(ctx_, kind_, id_, m.faces.data(), m.faceVerts.data(), m.uv.data(), &nv),
              "cs_get_topology");
        return m;
    }
    int nodes() const {
        int nn = 0, nf = 0, ne = 0;
        check(cs_get_counts(ctx_, kind_, id_, &nn, &nf, &ne), "cs_get_counts");
        return nn;
    }
    void setState(const Mesh& m) {
        check(cs_set_state(ctx_, kind_, id_, m.x.data(), m.x0.empty() ? nullptr : m.x0.data(),
                           m.v.empty() ? nullptr : m.v.data()),
              "cs_set_state");
    }
    int id() const { return id_; }

  protected:
    void check(cs_status s, const char* what) const {
        if (s != CS_OK) throw Error(s, std::string(what) + ": " + cs_last_error(ctx_));
    }
    cs_ctx* ctx_;
    int kind_, id_;
};

struct Cloth : MeshHandle {
    Cloth(cs_ctx* ctx, int id) : MeshHandle(ctx, CS_MESH_CLOTH, id) {}
};
struct Obstacle : MeshHandle {
    Obstacle(cs_ctx* ctx, int id) : MeshHandle(ctx, CS_MESH_OBSTACLE, id) {}
};

struct Resume {};  // tag of the resuming constructor

class Simulator {
  public:
    // Loads a reference configuration file; relative mesh / material paths resolve against rootDir.
    explicit Simulator(const std::string& path, const std::string& rootDir = "", const cs_params* base = nullptr)
        : path_(path), root_(rootDir) {
        load(base);
    }
    // Resume mode: the configuration a save wrote into `directory` (its cloths are the last saved
    // frame), continued from that frame (src/Simulator.cu:3-16,1045-1053).
    Simulator(Resume, const std::string& directory, const std::string& rootDir = "", const cs_params* base = nullptr)
        : path_(directory + "/config.json"), root_(rootDir) {
        load(base);
        nFrames_ = lastFrame(directory, info_.n_cloths);
        nSteps_ = nFrames_ * info_.frame_steps;
        cs_status s = cs_set_step_count(ctx_, nSteps_);
        if (s != CS_OK) throw Error(s, std::string("cs_set_step_count: ") + cs_last_error(ctx_));
    }
    ~Simulator() {
        try {
            flush();
        } catch (...) {
        }
        cs_destroy(ctx_);
    }
    Simulator(const Simulator&) = delete;
    Simulator& operator=(const Simulator&) = delete;

    cs_stats physicsStep() { return run(cs_physics_step, "cs_physics_step"); }
    cs_stats collisionStep() { return run(cs_collision_step, "cs_collision_step"); }
    cs_stats simulateStep() {
        cs_stats st = run(cs_step, "cs_step");
        nSteps_++;
        if (nSteps_ % info_.frame_steps == 0) {
            nFrames_++;
            if (info_.remeshing) {  // Simulator::remeshingStep (src/Simulator.cu:857-871)
                cs_status s = cs_remesh(ctx_);
                if (s != CS_OK) throw Error(s, std::string("cs_remesh: ") + cs_last_error(ctx_));
            }
        }
        return st;
    }
    // Simulator::finished (src/Simulator.cu:979-981)
    bool finished() const {
        const float dt = info_.frame_time / info_.frame_steps;
        return nSteps_ * dt >= info_.end_time || nFrames_ >= info_.end_frame;
    }
    // simulateOffline: save, then steps until finished() with a save after every frame
    // (src/Simulator.cu:1018-1043)
    void simulateOffline(const std::string& directory) {
        save(directory);
        resumeOffline(directory);
    }
    // resumeOffline's loop (src/Simulator.cu:1073-1083), for a context made with Resume; with
    // `overlapped`, each frame is written by saveOverlapped while the next one steps
    void resumeOffline(const std::string& directory, bool overlapped = false) {
        while (!finished()) {
            simulateStep();
            if (nSteps_ % info_.frame_steps != 0) continue;
            if (overlapped) saveOverlapped(directory);
            else save(directory);
        }
        flush();
    }
    // Simulator::save (src/Simulator.cu:950-957, src/Cloth.cu:865-887): every cloth as
    // <directory>/frame<nFrames>_cloth<i>.obj and <directory>/config.json naming those meshes.
    void save(const std::string& directory) {
        flush();  // an overlapped frame's config.json first
        std::vector<std::string> paths;
        for (size_t c = 0; c < cloths_.size(); c++) {
            paths.push_back(framePath(directory, nFrames_, (int)c));
            cloths_[c].getMesh().save(paths.back());
        }
        writeConfig(directory, paths);
    }
    // save() overlapped with the steps that follow: every cloth's nodes are snapshotted on the
    // device into one of two download slots (cs_frame_begin) and the frame's files are written on
    // a writer thread, its config.json after the previous frame's.  The same files as save();
    // flush() (also run by resumeOffline and the destructor) waits for the pending writes and
    // rethrows their errors.
    void saveOverlapped(const std::string& directory) {
        const int slot = slot_;
        slot_ ^= 1;
        if (pending_[slot].valid()) pending_[slot].get();  // its writer has read the slot
        cs_status s = cs_frame_begin(ctx_, slot);
        if (s != CS_OK) throw Error(s, std::string("cs_frame_begin: ") + cs_last_error(ctx_));
        std::vector<Mesh> meshes;
        for (const Cloth& c : cloths_) meshes.push_back(c.topology());
        std::shared_future<void> prev = last_;
        const int frame = nFrames_;
        last_ = pending_[slot] =
            std::async(std::launch::async, [this, directory, slot, frame, prev, meshes]() mutable {
                std::vector<std::string> paths;
                for (size_t c = 0; c < meshes.size(); c++) {
                    int n = 0;
                    cs_status st = cs_frame_end(ctx_, slot, (int)c, nullptr, nullptr, &n);
                    if (st == CS_OK) {
                        meshes[c].x.resize(3 * (size_t)n);
                        meshes[c].v.resize(3 * (size_t)n);
                        st = cs_frame_end(ctx_, slot, (int)c, meshes[c].x.data(), meshes[c].v.data(), &n);
                    }
                    if (st != CS_OK) throw Error(st, "cs_frame_end failed");
                    paths.push_back(framePath(directory, frame, (int)c));
                    meshes[c].save(paths.back());
                }
                if (prev.valid()) prev.get();
                writeConfig(directory, paths);
            }).share();
    }
    void flush() {
        for (auto& f : pending_)
            if (f.valid()) {
                std::shared_future<void> g = f;
                f = std::shared_future<void>();
                g.get();
            }
        last_ = std::shared_future<void>();
    }
    // Simulator::lastFrame (src/Simulator.cu:962-977): the last frame whose every cloth file exists
    static int lastFrame(const std::string& directory, int nCloths) {
        int f = 0;
        for (;; f++) {
            bool all = true;
            for (int c = 0; c < nCloths && all; c++) {
                std::ifstream in(directory + "/frame" + std::to_string(f + 1) + "_cloth" + std::to_string(c) + ".obj");
                all = (bool)in;
            }
            if (!all) return f;
        }
    }

    std::vector<Cloth>& getCloths() { return cloths_; }
    std::vector<Obstacle>& getObstacles() { return obstacles_; }
    const cs_scene_info& info() const { return info_; }
    int steps() const { return nSteps_; }
    int frames() const { return nFrames_; }
    cs_ctx* handle() { return ctx_; }

  private:
    static std::string framePath(const std::string& directory, int frame, int cloth) {
        return directory + "/frame" + std::to_string(frame) + "_cloth" + std::to_string(cloth) + ".obj";
    }
    // config.json naming the frame's cloth files, written to a temporary and renamed
    void writeConfig(const std::string& directory, const std::vector<std::string>& paths) const {
        std::vector<const char*> p;
        for (const std::string& q : paths) p.push_back(q.c_str());
        const std::string cfg = directory + "/config.json";
        const std::string tmp = cfg + ".tmp";
        cs_status s = cs_write_config(path_.c_str(), tmp.c_str(), p.data(), (int)p.size());
        if (s != CS_OK) throw Error(s, std::string("cs_write_config: ") + cs_scene_error());
        if (std::rename(tmp.c_str(), cfg.c_str()) != 0) throw Error(CS_E_ARG, "cannot write " + cfg);
    }
    void load(const cs_params* base) {
        cs_status s = cs_load_scene(path_.c_str(), root_.c_str(), base, &ctx_, &info_);
        if (s != CS_OK) throw Error(s, std::string("cs_load_scene: ") + cs_scene_error());
        for (int i = 0; i < info_.n_cloths; i++) cloths_.emplace_back(ctx_, i);
        for (int i = 0; i < info_.n_obstacles; i++) obstacles_.emplace_back(ctx_, i);
    }
    cs_stats run(cs_status (*f)(cs_ctx*, cs_stats*), const char* what) {
        cs_stats st{};
        cs_status s = f(ctx_, &st);
        if (s != CS_OK && s != CS_E_NOT_CONVERGED) throw Error(s, std::string(what) + ": " + cs_last_error(ctx_));
        return st;
    }
    std::string path_, root_;
    cs_ctx* ctx_ = nullptr;
    cs_scene_info info_{};
    std::vector<Cloth> cloths_;
    std::vector<Obstacle> obstacles_;
    int nSteps_ = 0, nFrames_ = 0;
    int slot_ = 0;                          // download slot of the next saveOverlapped
    std::shared_future<void> pending_[2];   // the writer of each slot's last frame
    std::shared_future<void> last_;         // the latest writer
};

}  // namespace clothsim
