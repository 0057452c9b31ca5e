// Batched VBD stepping of many independent scenes on one GPU.
//
// The reference steps each scene (entity) on its own, one launch per colour
// (src/GPU/VBD/GPU_VBD_FEM.cu:712-756, driven scene after scene by
// src/Core/Engine.cpp:15-21).  Here a launch covers colour c of every scene
// that has a colour c: scenes are independent, so colour c of scene A and
// colour c of scene B can be solved concurrently, and a launch carries enough
// vertices to fill the chip.  Per step the launch sequence is
//   predictor (all vertices of all scenes)
//   for each iteration:  refresh of the cached displacement gradients (per
//                          element type whose tile caches them)
//                        for each global colour c, for each element type:
//                          gather + local solve of colour c
//                        fixation + box limit (all vertices, one fused kernel)
//   velocity update
// Each scene still sees the reference's order: ascending colours, constraints
// after every sweep.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <vector>

#include "oe_gather.hpp"
#include "oe_scene.hpp"

namespace oe {

// Launch plan for a set of scenes.  Scene indices refer to a SceneDev table.
// Every launch covers a contiguous run of segments whose block_begin values
// are launch-local; kernels find their segment by binary search, so the plan
// is O(segments), not O(blocks), to build and upload.
struct Schedule {
    struct GLaunch {
        int color, element, S, Q, L, QL;
        int mode;                         // PassMode
        bool latency;                     // task pass in the kernels' latency form
        int inc = 0;                      // task pass: the tile's incidences per throughput-form block
        int seg_first, seg_count, nblocks;   // gather: slot ranges; tasks: vertex-block ranges
        int64_t slots;   // vertex slots covered (all scenes)
        int64_t incidences;
    };
    // refresh of cached gradients: one launch per element type, over jsegs
    struct JLaunch {
        int S, Q;
        int seg_first, seg_count, nblocks;
    };
    std::vector<BatchSeg> segs;
    std::vector<GLaunch> glaunches;   // ascending colour, then element type
    std::vector<VSeg> vsegs;          // every vertex of every scene, one launch
    int vblocks = 0;
    std::vector<JLaunch> jlaunches;
    std::vector<VSeg> jsegs;          // count = the scene's elements
    q2::Basis1D q2_basis{};
    int max_iterations = 0;
    int64_t vertex_updates_per_substep = 0;
    double algorithmic_bytes_per_substep = 0.0;   // gather kernels, all iterations
    std::map<int, double> algorithmic_bytes_by_element;    // the same per element type
    std::map<int, int64_t> vertex_updates_by_element;

    void clear();
    // scenes[i] is placed at SceneDev table index table_index[i]; want_latency
    // lets task passes too small to fill the GPU (as a server's ticks often
    // are) take the latency form
    void build(const std::vector<Scene*>& scenes, const std::vector<int>& table_index,
               const std::map<int, GatherTile>& tiles, bool want_latency = false);
};

// Device copies of a schedule's tables (grown on demand, reused across uploads).
struct DeviceSchedule {
    DBuf<BatchSeg> segs;
    DBuf<VSeg> vsegs;
    DBuf<VSeg> jsegs;
    void upload(Engine* eng, const Schedule& s, hipStream_t stream, bool sync);
};

// Validates that a scene can be stepped in a batch and describes it for the
// batched kernels.  `fixed` is the scene's per-vertex fixation flag array; the
// scene's gradient cache is allocated when its element type's tile uses it.
SceneDev make_scene_dev(Scene& s, const unsigned char* fixed, const std::map<int, GatherTile>& tiles);
std::vector<unsigned char> fixed_flags(const Scene& s);
void check_batchable(const Scene& s);

// HIP events around colour-pass launches; durations are summed in total and
// per tag (the launch's element type).
struct GatherTimer {
    bool enabled = false;
    std::vector<hipEvent_t> pool;
    std::vector<int> tags;   // one per event pair
    size_t used = 0;
    double ms = 0.0;
    int64_t launches = 0;
    std::map<int, double> tag_ms;
    std::map<int, int64_t> tag_launches;
    void before(hipStream_t s, int tag = 0);
    void after(hipStream_t s);
    void collect(hipStream_t s);
    ~GatherTimer();
};

// A second stream for the Q2 colour passes of a schedule that also holds other
// element types: the two chains of colour passes are independent within an
// iteration, so they run side by side and fill each other's idle issue slots
// (the Q2 pass is VALU-bound, the P2 pass waits on loads).  The fork and join
// events order them against the per-iteration passes on the main stream.
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    void create();
    ~SideStream();
};

// Enqueue `substeps` substeps of dt/substeps for every scene of the schedule.
// With `side` (and no timer) a schedule with Q2 and other element types runs
// its Q2 colour passes on side->stream.
void launch_schedule(const Schedule& sch, const DeviceSchedule& dev, const SceneDev* d_scenes, float dt, int substeps,
                     hipStream_t stream, GatherTimer* timer, SideStream* side = nullptr);

// The default tile of every element type: a scene's (default_tile) for a
// server's ticks, and for a batch the same but 28-incidence Q2 pencil blocks.
std::map<int, GatherTile> default_tiles(bool served = false);

// A batch registers itself with its scenes (Scene::batches).  A scene may be
// in several batches but appear once per batch, and a scene held by a server
// cannot join one.  Destroying a scene first detaches it: the batch then
// refuses to step (std::logic_error).
struct Batch {
    Batch(Engine* eng, std::vector<Scene*> scenes);
    ~Batch();
    void step(float dt, int substeps);
    void set_graph(bool on);
    void set_streams(int n);   // 1 or 2 (the Q2 passes on a second stream)
    void set_tile(Element e, GatherTile t);
    void rebuild();
    void scene_destroyed(Scene* s);

    Engine* eng;
    std::vector<Scene*> scenes;
    std::vector<DBuf<unsigned char>> flags;
    DBuf<SceneDev> d_scenes;   // rewritten by rebuild()
    std::map<int, GatherTile> tiles;
    Schedule sched;
    DeviceSchedule dsched;
    GatherTimer timer;
    SideStream side;

private:
    bool broken = false;   // a scene of the batch was destroyed
    bool use_graph = false;
    int streams = 2;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    float graph_dt = 0.f;
    int graph_sub = 0;
    void drop_graph();
};

}  // namespace oe
