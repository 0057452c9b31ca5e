// Launch interface of the element-task VBD pass for 27-node hexahedra (see
// oe_q2.hip): per colour, a task kernel over the colour's incidences writes a
// 9-float partial per incidence, then a vertex kernel sums them and solves.
#pragma once

#include <hip/hip_runtime.h>

#include "oe_device.hpp"
#include "oe_gather.hpp"
#include "oe_q2_tensor.hpp"

namespace oe {

struct Q2TaskArgs {
    PSView ps;
    FemView fem;
    MaterialView mt;
    q2::Basis1D basis;
    const int* own;    // owners array
    float* partial;    // 9 floats per owners entry
    int first;         // first owners index of the colour
    int count;         // incidences of the colour
};

struct VertexSolveArgs {
    PSView ps;
    OwnersView own;
    const float4* y;
    const float* partial;
    float damping;
    float dt;
    int first_slot;
    int count;
};

// Batched tasks: BatchSeg.first / count are an owners range of one scene.
struct BatchQ2TaskArgs {
    const SceneDev* scenes;
    const BatchSeg* segs;
    int seg_first;
    int seg_count;
    int iteration;
    q2::Basis1D basis;
};

int q2_tasks_per_block();
void launch_q2_tasks(const Q2TaskArgs& a, hipStream_t s);
void launch_batch_q2_tasks(const BatchQ2TaskArgs& b, int nblocks, hipStream_t s);
void launch_vertex_solve(const VertexSolveArgs& a, hipStream_t s);
// BatchSeg over slots, 256 slots per block
void launch_batch_vertex_solve(const BatchGatherArgs& b, int nblocks, hipStream_t s);

}  // namespace oe
