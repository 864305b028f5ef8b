// clothsim.hpp — header-only C++ mirror of the reference's host classes over the C-ABI.
//
// Names follow the reference so that host code reads the same:
//   Simulator(path)               Simulator::Simulator (src/Simulator.cu:3-135), scene file only
//   Simulator::physicsStep()      src/Simulator.cu:571-600
//   Simulator::collisionStep()    src/Simulator.cu:602-684
//   Simulator::simulateStep()     src/Simulator.cu:838-897 (remeshing and separation disabled)
//   Simulator::getCloths() / getObstacles(), Cloth::getMesh(), Obstacle::getMesh()
//   Mesh::save(path)              OBJ with "v" and "nv" lines at precision 20 (src/Mesh.cu:454-470)
// Errors throw clothsim::Error (the reference prints and exit(1)s instead).
#pragma once
#include <cstdio>
#include <fstream>
#include <iomanip>
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
    int nodes() const { return (int)(x.size() / 3); }
    void save(const std::string& path) const {
        std::ofstream out(path);
        if (!out) throw Error(CS_E_ARG, "cannot write " + path);
        out << std::setprecision(20);
        for (int i = 0; i < nodes(); i++) {
            out << "v " << x[3 * i] << " " << x[3 * i + 1] << " " << x[3 * i + 2] << "\n";
            out << "nv " << v[3 * i] << " " << v[3 * i + 1] << " " << v[3 * i + 2] << "\n";
        }
        for (size_t f = 0; f + 2 < faces.size(); f += 3)
            out << "f " << faces[f] + 1 << " " << faces[f + 1] + 1 << " " << faces[f + 2] + 1 << "\n";
    }
};

class Simulator;

// One cloth or obstacle of a Simulator (kind = CS_MESH_CLOTH or CS_MESH_OBSTACLE).
class MeshHandle {
  public:
    MeshHandle(cs_ctx* ctx, int kind, int id) : ctx_(ctx), kind_(kind), id_(id) {}
    Mesh getMesh() const {
        int nn = 0, nf = 0, ne = 0;
        check(cs_get_counts(ctx_, kind_, id_, &nn, &nf, &ne), "cs_get_counts");
        Mesh m;
        m.x.resize(3 * (size_t)nn);
        m.x0.resize(3 * (size_t)nn);
        m.v.resize(3 * (size_t)nn);
        check(cs_get_state(ctx_, kind_, id_, m.x.data(), m.x0.data(), m.v.data()), "cs_get_state");
        return m;
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

class Simulator {
  public:
    // Loads a reference configuration file; relative mesh / material paths resolve against rootDir.
    explicit Simulator(const std::string& path, const std::string& rootDir = "", const cs_params* base = nullptr) {
        cs_status s = cs_load_scene(path.c_str(), rootDir.c_str(), base, &ctx_, &info_);
        if (s != CS_OK) throw Error(s, std::string("cs_load_scene: ") + cs_scene_error());
        for (int i = 0; i < info_.n_cloths; i++) cloths_.emplace_back(ctx_, i);
        for (int i = 0; i < info_.n_obstacles; i++) obstacles_.emplace_back(ctx_, i);
    }
    ~Simulator() { cs_destroy(ctx_); }
    Simulator(const Simulator&) = delete;
    Simulator& operator=(const Simulator&) = delete;

    cs_stats physicsStep() { return run(cs_physics_step, "cs_physics_step"); }
    cs_stats collisionStep() { return run(cs_collision_step, "cs_collision_step"); }
    cs_stats simulateStep() {
        cs_stats st = run(cs_step, "cs_step");
        nSteps_++;
        if (nSteps_ % info_.frame_steps == 0) nFrames_++;
        return st;
    }
    // Steps until the configuration's end_time / end_frame, saving every cloth once per frame
    // as <directory>/frame<NNNNN>_<cloth>.obj (frame 0 is the initial state).
    void simulateOffline(const std::string& directory) {
        saveFrame(directory);
        const float dt = info_.frame_time / info_.frame_steps;
        while ((nSteps_ + 1) * dt <= info_.end_time + 0.5f * dt && nFrames_ < info_.end_frame) {
            const int frames = nFrames_;
            simulateStep();
            if (nFrames_ != frames) saveFrame(directory);
        }
    }
    void saveFrame(const std::string& directory) const {
        for (size_t c = 0; c < cloths_.size(); c++) {
            char name[64];
            std::snprintf(name, sizeof(name), "/frame%05d_%02d.obj", nFrames_, (int)c);
            cloths_[c].getMesh().save(directory + name);
        }
    }

    std::vector<Cloth>& getCloths() { return cloths_; }
    std::vector<Obstacle>& getObstacles() { return obstacles_; }
    const cs_scene_info& info() const { return info_; }
    int steps() const { return nSteps_; }
    int frames() const { return nFrames_; }
    cs_ctx* handle() { return ctx_; }

  private:
    cs_stats run(cs_status (*f)(cs_ctx*, cs_stats*), const char* what) {
        cs_stats st{};
        cs_status s = f(ctx_, &st);
        if (s != CS_OK && s != CS_E_NOT_CONVERGED) throw Error(s, std::string(what) + ": " + cs_last_error(ctx_));
        return st;
    }
    cs_ctx* ctx_ = nullptr;
    cs_scene_info info_{};
    std::vector<Cloth> cloths_;
    std::vector<Obstacle> obstacles_;
    int nSteps_ = 0, nFrames_ = 0;
};

}  // namespace clothsim
