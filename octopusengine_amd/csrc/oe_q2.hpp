// Launch interface of the fused element-task VBD pass for 27-node hexahedra
// (see oe_q2.hip): per colour, one kernel whose blocks compute the partials of
// their vertices' incidences in LDS and then solve those vertices.
#pragma once

#include <hip/hip_runtime.h>

#include "oe_device.hpp"
#include "oe_gather.hpp"
#include "oe_q2_tensor.hpp"

namespace oe {

// One block per vertex block (int4 {first slot, slot count, first owners
// index, incidence count}, colour-major) with at most kQ2FusedIncPerBlock
// incidences and kQ2FusedMaxSlots slots.
// The latency form (launches too small to fill the GPU) takes one task per
// half-wave and loads the vertex state along the element chain; it has its own
// vertex blocks of at most kQ2FusedIncPerBlockLat incidences.
constexpr int kQ2FusedThreads = 256;
// 4 tasks per half-wave.  8 are 2% faster on a large batch but double the
// critical path of a block, which the served latency of large scenes pays.
constexpr int kQ2FusedIncPerBlock = 32;
constexpr int kQ2FusedIncPerBlockLat = 8;
constexpr int kQ2FusedMaxSlots = 64;
// The pencil form of the throughput pass (oe_q2p.hip, the default): one-wave
// blocks of 7 tasks at a time, in up to 5 rounds; chosen by a tile of 14, 21,
// 28 or 35 incidences per block.  Other tile sizes (up to 32) take the 27-lane
// throughput form.
//  * 28 (4 rounds): batches, whose launches hold enough blocks to cover the
//    longer blocks (256-scene step 47.2 ms; 51.4 ms with the 27-lane form).
//  * 14 (2 rounds): single scenes and the server, whose launches are short of
//    blocks (served p99 6.6 ms; 7.8 ms at 28 and 7.0 ms with the 27-lane form).
// A scene whose vertices have more incident elements than the tile's blocks
// hold takes the next size up (q2_pencil_fit): the kernel runs blocks of any
// of these sizes, so scenes of one launch may differ.
constexpr int kQ2PencilIncPerBlock = 28;
constexpr int kQ2PencilIncLatency = 14;
constexpr int kQ2PencilMaxInc = 35;
constexpr int kQ2PencilMaxSlots = kQ2PencilMaxInc;
inline bool q2_pencil(int inc) { return inc >= 14 && inc <= kQ2PencilMaxInc && inc % 7 == 0; }
// the pencil block size for a tile of inc when a vertex has up to max_incidences
inline int q2_pencil_fit(int inc, int max_incidences) {
    if (!q2_pencil(inc)) return inc;
    while (inc < max_incidences && inc < kQ2PencilMaxInc) inc += 7;
    return inc;
}

struct Q2FusedArgs {
    PSView ps;
    FemView fem;
    OwnersView own;
    MaterialView mt;
    q2::Basis1D basis;
    const float4* y;
    const int4* blocks;   // vertex blocks of the scene (of the form used)
    float damping;
    float dt;
    int first_block;      // the colour's blocks: [first_block, first_block + count)
    int count;
    int latency;          // 1: latency form (blocks are the latency-form vertex blocks)
};

// Batched: BatchSeg.first / count are a range of one scene's vertex blocks,
// one block each.
struct BatchQ2FusedArgs {
    const SceneDev* scenes;
    const BatchSeg* segs;
    int seg_first;
    int seg_count;
    int iteration;
    float dt;
    q2::Basis1D basis;
    int latency;   // 1: latency form and the scenes' latency vertex blocks
};

void launch_q2_fused(const Q2FusedArgs& a, hipStream_t s);
void launch_batch_q2_fused(const BatchQ2FusedArgs& b, int nblocks, hipStream_t s);
// throughput form with pencil tasks (blocks of at most kQ2PencilIncPerBlock incidences)
void launch_q2_pencil(const Q2FusedArgs& a, hipStream_t s);
void launch_batch_q2_pencil(const BatchQ2FusedArgs& b, int nblocks, hipStream_t s);

}  // namespace oe
