// GPU dynamics object model.  It keeps the reference's plugin interface for the
// GPU soft-body path so that the same composition works:
//   GPU_ParticleSystem        include/GPU/GPU_ParticleSystem.h:54-73
//   GPU_Dynamic               include/GPU/GPU_Dynamic.h:6-11   (start / step / active)
//   GPU_Integrator            include/GPU/GPU_Integrator.h:11-32 (dynamics + constraints)
//   GPU_SemiExplicit          include/GPU/GPU_Integrator.h:34-36
//   GPU_VBD                   include/GPU/VBD/GPU_VBD.h:13-48
//   GPU_PBD                   include/GPU/PBD/GPU_PBD.h
//   GPU_FEM                   include/GPU/GPU_FEM.h:61-101
//   GPU_VBD_FEM               include/GPU/VBD/GPU_VBD_FEM.h:62-130
//   GPU_PBD_FEM               include/GPU/PBD/GPU_PBD_FEM.h:15-51
//   GPU_FEM_Explicit          include/GPU/Explicit/GPU_FEM_Explicit.h:12-33
//   GPU_Fix_Constraint        include/GPU/GPU_Constraint.h:11-18
//   GPU_Box_Limit             include/GPU/GPU_Constraint.h:20-24
// The state lives in SoA device arrays owned by the engine's memory pool and
// every launch goes to the engine's stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include <map>

#include "oe_device.hpp"
#include "oe_elementwise.hpp"
#include "oe_gather.hpp"
#include "oe_host.hpp"
#include "oe_q2.hpp"

namespace oe {

struct Engine {
    int device = 0;
    hipStream_t stream = nullptr;
    MemoryPool pool;
    bool autotune = false;   // tune the layout of every VBD / explicit scene created
    explicit Engine(int dev, size_t budget_bytes = SIZE_MAX);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;
};

struct GPU_ParticleSystem {
    GPU_ParticleSystem(Engine* eng, const std::vector<float>& xyz, const std::vector<float>& masses);
    int nb_particles() const { return n; }
    PSView get_parameters() const;
    void get_position(std::vector<float>& xyz) const;
    void get_prev_position(std::vector<float>& xyz) const;
    void get_init_position(std::vector<float>& xyz) const;
    void get_velocity(std::vector<float>& xyz) const;
    void get_forces(std::vector<float>& xyz) const;
    void get_masks(std::vector<int>& m) const;
    void get_mass(std::vector<float>& m) const;
    void get_inv_mass(std::vector<float>& w) const;
    void set_position(const std::vector<float>& xyz);
    void set_velocity(const std::vector<float>& xyz);
    void set_state(const float* p, const float* last_p, const float* v);
    size_t device_bytes() const;
    std::vector<float> to_displacement(const float* xyz) const;
    // Keep low-order parts of the displacement records in their w lanes (see
    // dpack): set by the VBD integrator, whose state is then hi + lo per vertex.
    void enable_compensation();
    bool compensated() const { return comp; }

    Engine* eng;
    int n;
    std::vector<float> rest;   // rest positions X (host copy, 3n)
    DBuf<float4> p, last_p, init_p, v, f;
    DBuf<float> m, w;
    DBuf<int> mask;

private:
    bool comp = false;
    float lo_unit = 1.f;   // quantum of the low-order parts (lo_unit_for the rest bounding box)
    // displacement of every vertex, hi + lo in double (lo = 0 without compensation)
    std::vector<double> full(const DBuf<float4>& rec) const;
    void upload_split(const float* xyz, DBuf<float4>& rec);
};

struct GPU_Dynamic {
    GPU_Dynamic() = default;
    virtual ~GPU_Dynamic() = default;
    virtual void start(GPU_ParticleSystem*, float) {}            // once per step
    virtual void step(GPU_ParticleSystem* ps, float dt) = 0;      // once per iteration / substep
    bool active = true;
};

struct GPU_Integrator : GPU_Dynamic {
    explicit GPU_Integrator(bool apply = true) : apply_constraints(apply) {}
    virtual void add_dynamics(GPU_Dynamic* d) { _dynamics.emplace_back(d); }
    virtual void add_constraint(GPU_Dynamic* c) { _constraints.emplace_back(c); }
    virtual void init_dynamics(GPU_ParticleSystem* ps, float dt);
    virtual void eval_dynamics(GPU_ParticleSystem* ps, float dt);
    virtual void eval_constraints(GPU_ParticleSystem* ps, float dt);

    bool apply_constraints;
    float gravity[3] = {0.f, -9.81f, 0.f};
    hipStream_t stream = nullptr;
    std::vector<std::unique_ptr<GPU_Dynamic>> _dynamics;
    std::vector<std::unique_ptr<GPU_Dynamic>> _constraints;
};

struct GPU_SemiExplicit final : GPU_Integrator {
    void step(GPU_ParticleSystem* ps, float dt) override;
};

struct GPU_VBD final : GPU_Integrator {
    GPU_VBD(GPU_ParticleSystem* ps, int it, float rho);
    void step(GPU_ParticleSystem* ps, float dt) override;
    void add_dynamics(GPU_Dynamic* d) override;
    int iteration;
    float _rho;
    DBuf<float4> y, prev_it_p, prev_it2_p;
};

struct GPU_PBD final : GPU_Integrator {
    explicit GPU_PBD(float global_damping = 0.f) : _global_damping(global_damping) {}
    void step(GPU_ParticleSystem* ps, float dt) override;
    float _global_damping;
};

struct GPU_FEM : GPU_Dynamic {
    GPU_FEM(Engine* eng, Element element, const HostMesh& mesh, float young, float poisson, Material material);
    FemView fem_view() const;
    // Per-element stress * V, volume, volume difference and inverted flag of
    // the current state (E floats each, any may be null; synchronising).
    void diagnostics(GPU_ParticleSystem* ps, float* stress, float* volume, float* volume_diff, float* inverted);
    Engine* eng;
    Element element;
    int S, Q, E;
    MaterialView mt;
    bool affine = false;   // element records stored once per element
    DBuf<int> topo;
    DBuf<float> jxv, dN, wq;
    std::vector<double> elem_volume;
};

enum VBD_Version : int { Base = 0, Threaded_Quadrature = 1, Reduction_Symmetry = 2, Better_Coloration = 3 };

// How a VBD colour pass runs, per element type:
//   gather  one kernel, L lanes per vertex, QL quadrature points per lane;
//   cache   the same with the displacement gradients of every (element,
//           quadrature point) kept in device memory between colours (QL = 1);
//   tasks   element tasks, one per incidence of the colour (L, QL unused), in
//           one kernel whose blocks put the partials of a run of the colour's
//           vertices in LDS, then solve those vertices (sum-factorised
//           gradients for 27-node hexahedra).
enum PassMode : int { PASS_GATHER = 0, PASS_CACHE = 1, PASS_TASKS = 2 };
struct GatherTile {
    int L = 16;
    int QL = 1;
    int mode = PASS_GATHER;
    // element tasks: incidences per throughput-form block (the elements one
    // workgroup stages in LDS; 0 = the kernel's maximum), and the number of
    // throughput-form blocks below which a colour takes the latency form
    int inc = 0;
    int64_t lat_blocks = kLatencyBlocks;
};
GatherTile default_tile(Element e);
// Largest incidences per block the element-task kernels hold.
inline int max_task_inc(Element e) { return e == Hexa27 ? kQ2PencilMaxInc : kFusedIncPerBlock; }

// Vertex blocks of the fused element-task kernels: int4 {first slot, slot
// count, first owners index, incidence count} per block, colour-major;
// color_begin[c] = first block of colour c.
struct TaskTables {
    DBuf<int4> blocks;
    std::vector<int32_t> color_begin;
    int32_t color_blocks(int c) const { return color_begin[c + 1] - color_begin[c]; }
};

struct GPU_VBD_FEM final : GPU_FEM {
    GPU_VBD_FEM(Engine* eng, Element element, const HostMesh& mesh, Material material, float young, float poisson,
                float damping, VBD_Version v = Reduction_Symmetry, int target_color = -1);
    void step(GPU_ParticleSystem* ps, float dt) override;
    // Colour order of every sweep: ascending (seed 0), or a Fisher-Yates
    // shuffle of 0..nc-1 per sweep (i = nc-1 down to 1 swaps i with
    // SplitMix64() % (i + 1)) driven by one SplitMix64 stream seeded once.
    // The reference shuffles with a time-seeded generator every iteration
    // (src/GPU/VBD/GPU_VBD_FEM.cu:713-716); a fixed seed makes it repeatable.
    void set_shuffle_seed(uint64_t seed);
    void next_color_order(std::vector<int>& order);
    int nb_colors() const { return colors.nb_color; }

    Coloration colors;
    OwnersCSR owners;   // host copy (colour ranges are read at launch time)
    VBD_Version version;
    int target_color;
    float damping;
    GatherTile tile;
    bool shuffle = false;
    uint64_t rng_state = 0;   // SplitMix64 state
    const float4* y = nullptr;   // bound by GPU_VBD::add_dynamics
    DBuf<int> d_slot_vid;
    DBuf<int2> d_slot_nbo;
    DBuf<int> d_own;
    DBuf<float4> p_new;
    DBuf<float> ju;           // cached displacement gradients, E * Q * 9 (allocated on first use)
    float* ensure_ju();
    // Residual force of every free vertex's local problem at the current
    // positions (3N floats; fixed vertices keep 0; synchronising).
    void residual(GPU_ParticleSystem* ps, float dt, std::vector<float>& r);
    // Enqueue the residual kernel only: float4 per vertex into out4, or
    // packed xyz into out3 when out3 is set; fixed vertices are not written.
    // Allocates nothing.
    void launch_residual(GPU_ParticleSystem* ps, float dt, float4* out4, float* out3);
    DBuf<float4> d_residual;
    bool cached() const { return tile.mode == PASS_CACHE && version != Better_Coloration; }
    // Vertex blocks of the throughput form, one table per block size
    // (incidences per block), built on first use and kept while the scene
    // lives: batches and servers hold pointers into them.
    std::map<int, TaskTables> thr_tables;
    // Vertex blocks of the kernels' latency form (one task per thread or
    // half-wave); empty when a vertex has more incidences than one holds.
    TaskTables lat_tables;
    bool lat_built = false;
    // Build (if needed) and return the throughput tables for `inc`
    // incidences per block (0 = the kernel's maximum); throws
    // std::invalid_argument when a vertex has more incidences than a block.
    const TaskTables& throughput_tables(int inc);
    // the incidences per throughput-form block a tile of inc gets on this mesh
    // (q2_pencil_fit for 27-node hexahedra; 0 = the kernel's maximum)
    int fit_inc(int inc) const;
    int max_incidences = 0;   // most incident elements of one vertex
    // The tables already built for `inc` (std::logic_error if not): for
    // step(), which may run under HIP graph capture and must not allocate.
    const TaskTables& built_tables(int inc) const;
    const TaskTables* latency_tables();   // null when absent
    void prepare_tasks();                 // the tables of the current tile
    bool tasks() const {
        return tile.mode == PASS_TASKS && version != Better_Coloration && (element != Hexa27 || has_q2_basis);
    }
    bool has_q2_basis = false;
    q2::Basis1D q2_basis{};
    int64_t conflicts = 0;
    // optional per-launch timing of the gather kernel (events around each colour)
    bool time_gather = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double gather_ms = 0.0;
    int64_t gather_launches = 0;
    void collect_gather_timing();
    ~GPU_VBD_FEM() override;
};

enum XPBD_FEM_VERSION : int { XPBD = 0, Quadrature = 1, OptiShared = 2, OptiGroup = 3 };

struct GPU_PBD_FEM final : GPU_FEM {
    GPU_PBD_FEM(Engine* eng, Element element, const HostMesh& mesh, float young, float poisson, Material material,
                XPBD_FEM_VERSION v = OptiShared);
    void step(GPU_ParticleSystem* ps, float dt) override;
    Coloration colors;   // element colours
    ElementSchedule sched;
    XPBD_FEM_VERSION version;
    DBuf<int> d_eids;
};

struct GPU_FEM_Explicit final : GPU_FEM {
    GPU_FEM_Explicit(Engine* eng, Element element, const HostMesh& mesh, float young, float poisson,
                     Material material, float damping);
    void step(GPU_ParticleSystem* ps, float dt) override;
    // The element-centric pass followed by the semi-implicit update in its
    // vertex pass; false (nothing launched) in the gather mode.
    bool step_integrate(GPU_ParticleSystem* ps, float dt, const float gravity[3]);
    float _damping;
    // mode PASS_TASKS (the default): element-centric force pass (one element
    // per 32 lanes, then a per-vertex sum of the element partials);
    // PASS_GATHER: the per-vertex gather with tile L x QL
    GatherTile tile;
    int n_vertices;
    DBuf<int> d_slot_vid;
    DBuf<int2> d_slot_nbo;
    DBuf<int> d_own;
    DBuf<int> d_inc_at;      // E * S: owners position of (element, local node)
    DBuf<float> d_partial;   // E * S * 9: per (element, local node) force and Hessian block, owners order
};

struct GPU_Fix_Constraint final : GPU_Dynamic {
    GPU_Fix_Constraint(Engine* eng, const std::vector<float>& xyz, const std::vector<int>& ids);
    void step(GPU_ParticleSystem* ps, float dt) override;
    FixTransform t;
    std::vector<int> ids;
    DBuf<int> d_ids;
    DBuf<unsigned char> d_flags;   // per vertex: 1 when fixed (the fused fix + box pass)
    hipStream_t stream;
};

struct GPU_Box_Limit final : GPU_Dynamic {
    GPU_Box_Limit(hipStream_t s, const float pmin_[3], const float pmax_[3]);
    void step(GPU_ParticleSystem* ps, float dt) override;
    float pmin[3], pmax[3];
    hipStream_t stream;
};

}  // namespace oe
