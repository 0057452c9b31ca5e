// A scene = one mesh + one particle system + one integrator with its FEM
// dynamic and constraints, advanced by step(dt, substeps).  This is the
// composition the reference's script components build:
//   Cuda_FEM_Dynamic::init                 include/Script/Dynamic/Cuda_FEM_Dynamic.h:18-25
//   Cuda_VBD_FEM_Dynamic                   src/Script/Dynamic/Cuda_VBD_FEM_Dynamic.cpp:13-28
//   Cuda_XPBD_FEM_Dynamic                  src/Script/Dynamic/Cuda_XPBD_FEM_Dynamic.cpp:16-37
//   Cuda_FEM_Explicit_Dynamic              include/Script/Dynamic/Cuda_FEM_Explicit_Dynamic.h:13-21
//   Cuda_ParticleSystem_Dynamics::update   src/Script/Dynamic/Cuda_ParticleSystem_Dynamic.cpp:17-31
//   Cuda_Constraint_Rigid_Controller       src/Script/Dynamic/Cuda_Constraint_Rigid_Controller.cpp:9-23
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "oe_gpu.hpp"

namespace oe {

class Server;
struct Batch;

enum Solver : int { SolverVBD = 0, SolverXPBD = 1, SolverExplicit = 2 };

struct SceneConfig {
    int solver = SolverVBD;
    Material material = Stable_NeoHooke;
    float young = 1e6f, poisson = 0.45f, density = 1000.f;
    Mass_Distribution distribution = FemShape;
    int iterations = 10;          // VBD iterations per substep
    float rho = 0.f;              // VBD Chebyshev spectral radius (0 = off)
    float damping = 0.f;          // VBD / explicit stiffness damping
    float global_damping = 0.f;   // XPBD velocity damping
    int vbd_version = Reduction_Symmetry;
    int target_color = -1;
    int xpbd_version = OptiShared;
    float gravity[3] = {0.f, -9.81f, 0.f};
    uint64_t shuffle_seed = 0;    // 0: ascending colour order every sweep
};

struct Scene {
    Scene(Engine* eng, HostMesh mesh, const SceneConfig& cfg);
    ~Scene();
    // fixed vertices = ids (GPU_Fix_Constraint) ; at most one fixation per scene
    void fix_vertices(const std::vector<int>& ids);
    // fixed vertices = {x : dot(x - o, n) > 0} (reference Plane area, include/Tools/Area.h:50-54)
    std::vector<int> select_plane(const float o[3], const float n[3]) const;
    void box_limit(const float pmin[3], const float pmax[3]);
    void step(float dt, int substeps);
    void set_graph(bool on);
    // Drop the captured substep graph (kept enabled): the next step captures
    // the launch sequence again, e.g. after a tile change.
    void invalidate_graph() { drop_graph(); }
    size_t device_bytes() const;

    Engine* eng;
    HostMesh mesh;
    SceneConfig cfg;
    std::vector<float> masses;
    std::unique_ptr<GPU_ParticleSystem> ps;
    std::unique_ptr<GPU_Integrator> integrator;
    GPU_FEM* fem = nullptr;            // owned by the integrator
    GPU_VBD_FEM* vbd_fem = nullptr;
    GPU_PBD_FEM* pbd_fem = nullptr;
    GPU_FEM_Explicit* explicit_fem = nullptr;
    GPU_Fix_Constraint* fix = nullptr;
    int64_t steps_done = 0;
    // Host-side links to the batches and the server that hold this scene.
    // A server steps the scene on its own streams, so while `server` is set
    // direct steps, state writes and constraint changes are refused and reads
    // first wait for the scene's submitted requests.  Batches copy the
    // scene's constraints into their device tables when they are built, so
    // constraint and tile changes are refused while `batches` is not empty.
    Server* server = nullptr;
    int server_slot = -1;
    std::vector<Batch*> batches;

private:
    void run_substeps(float dt, int substeps);
    bool use_graph = false;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    float graph_dt = 0.f;
    int graph_sub = 0;
    void drop_graph();
};

}  // namespace oe
