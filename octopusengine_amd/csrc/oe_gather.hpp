// Launch interface of the per-vertex gather kernel (see oe_gather.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "oe_device.hpp"
#include "oe_elementwise.hpp"

namespace oe {

enum GatherMode : int {
    GATHER_VBD_INPLACE = 0,     // Gauss-Seidel inside a proper colour (reference v3)
    GATHER_VBD_JACOBI = 1,      // write p_new, copied after the colour (reference v4)
    GATHER_EXPLICIT_FORCE = 2,  // internal force only, overwrite f (explicit integrator)
    GATHER_VBD_RESIDUAL = 3,    // residual force of the local VBD problem into p_new (no update)
};

struct GatherArgs {
    PSView ps;
    FemView fem;
    OwnersView own;
    MaterialView mt;
    const float4* y;   // inertia target (VBD)
    float4* p_new;     // Jacobi output (VBD v4) or residual forces
    // Displacement gradients Ju(e, q) = sum_k u_k (x) dN_k(q), 9 floats per
    // (element, quadrature point), or null to rebuild them from the nodes.
    // When set (in-place VBD only) the gather reads Ju instead of the S node
    // positions and, after its vertex moved by du, adds du (x) dN_r(q) to the
    // records of its incident elements.  Within one colour no two vertices
    // share an element, so every record has a single writer.
    float* ju;
    float damping;
    float dt;
    int first_slot;
    int count;
    int mode;
    // Residual mode only: when set, the residual of vertex v goes to
    // out3[3v .. 3v + 2] (packed xyz) instead of p_new[v].
    float* out3;
};

// Everything a batched kernel needs to know about one scene; an array of these
// lives in device memory and blocks index it by scene.
struct SceneDev {
    PSView ps;
    FemView fem;
    OwnersView own;
    MaterialView mt;
    float4* y;
    float* ju;        // cached displacement gradients, or null
    const int4* fblocks;       // vertex blocks of the fused element tasks (throughput form), or null
    const int4* fblocks_lat;   // the same for the latency form
    float damping;
    int iterations;
    float g[3];
    // constraints: fixed vertices (per-vertex flag, may be null) then box limit
    const unsigned char* fixed;
    float fix_com[3];
    float fix_offset[3];
    float fix_R[9];
    int has_box;
    float box_min[3];
    float box_max[3];
};

// A run of `count` consecutive slots (first = first slot) of one scene, covered
// by the launch-local blocks [block_begin, next segment's block_begin).
struct BatchSeg {
    int scene;
    int first;
    int count;
    int block_begin;
};

struct BatchGatherArgs {
    const SceneDev* scenes;
    const BatchSeg* segs;
    int seg_first;          // this launch's segments: segs[seg_first .. seg_first + seg_count)
    int seg_count;
    int iteration;          // scenes with fewer iterations skip the launch
    float dt;
};

// Fused element tasks: the task form of a VBD colour pass in one kernel for
// element types other than the 27-node hexahedron.  A block takes a run of
// consecutive slots of one colour (a "vertex block", int4 {first slot, slot
// count, first owners index, incidence count}) whose incidences fit
// kFusedIncPerBlock.  Its threads first take one incidence (element e, local
// node r) each, kFusedThreads apart, and put the force / Hessian partial of
// node r over all Q points of e in LDS; then one thread per slot adds its
// vertex's partials in the owners order and runs the local solve.
// Two forms share the kernels: the throughput form (each thread takes up to
// kFusedRuns incidences in turn) and the latency form for launches too small
// to fill the GPU (one incidence per thread, so a block's critical path holds
// one element; the vertex state is loaded along the element chain).  Each
// form has its own vertex blocks.  One run per thread measured fastest for the
// throughput form too (P2 batch: 11.25 ms per step against 11.50 ms for two,
// with no loads run ahead: those held 12 VGPRs and cost a wave per SIMD), so
// the two forms differ only in when the vertex state is loaded.
constexpr int kFusedThreads = 256;
constexpr int kFusedRuns = 1;
constexpr int kFusedIncPerBlock = kFusedThreads * kFusedRuns;
constexpr int kFusedIncPerBlockLat = kFusedThreads;
constexpr int kFusedMaxSlots = kFusedThreads;
// Below this many throughput-form blocks (2 per CU) a task-pass launch cannot
// fill the GPU: single scenes, and batches built for latency, use the latency
// form there.
constexpr int64_t kLatencyBlocks = 512;

struct FusedTaskArgs {
    PSView ps;
    FemView fem;
    OwnersView own;
    MaterialView mt;
    const float4* y;
    const int4* blocks;   // vertex blocks of the scene, colour-major (of the form used)
    float damping;
    float dt;
    int first_block;      // the colour's vertex blocks: [first_block, first_block + count)
    int count;
    int latency;          // 1: latency form (blocks are the latency-form vertex blocks)
};
// Batched: BatchSeg.first / count are a range of one scene's vertex blocks,
// one block each.
struct BatchFusedArgs {
    const SceneDev* scenes;
    const BatchSeg* segs;
    int seg_first;
    int seg_count;
    int iteration;
    float dt;
    int latency;   // 1: latency form and the scenes' latency vertex blocks
};

// Batched refresh of the cached displacement gradients: VSeg.count is the
// scene's element count; each block takes refresh_elements_per_block(S, Q)
// elements.
struct BatchRefreshArgs {
    const SceneDev* scenes;
    const VSeg* segs;
    int seg_first;
    int seg_count;
    int iteration;
};

// Element-centric internal forces of the explicit integrator: per element and
// local node r, the 9 floats (force, then the Hessian block xx xy xz yy yz zz)
// of node r over the element's Q points at partial[inc_at[e S + r] 9], where
// inc_at is the owners position of (e, r).
struct ExplicitElemArgs {
    PSView ps;
    FemView fem;
    MaterialView mt;
    const int* inc_at;   // E * S
    float* partial;      // E * S * 9, owners order
};

#define OE_GATHER_DECL(S, Q)                                                                                  \
    bool launch_gather_##S##_##Q(int L, int QL, const GatherArgs& a, hipStream_t s);                          \
    bool launch_batch_gather_##S##_##Q(int L, int QL, bool cached, const BatchGatherArgs& b, int nblocks,     \
                                       hipStream_t s);                                                        \
    void launch_ju_refresh_##S##_##Q(const PSView& ps, const FemView& f, float* ju, hipStream_t s);          \
    void launch_batch_ju_refresh_##S##_##Q(const BatchRefreshArgs& b, int nblocks, hipStream_t s);          \
    int refresh_elements_per_block_##S##_##Q();                                                               \
    void launch_fused_tasks_##S##_##Q(const FusedTaskArgs& a, hipStream_t s);                                 \
    void launch_batch_fused_tasks_##S##_##Q(const BatchFusedArgs& b, int nblocks, hipStream_t s);            \
    void launch_explicit_elem_##S##_##Q(const ExplicitElemArgs& a, hipStream_t s);
OE_GATHER_DECL(4, 1)
OE_GATHER_DECL(8, 8)
OE_GATHER_DECL(10, 4)
OE_GATHER_DECL(20, 11)
OE_GATHER_DECL(27, 27)
#undef OE_GATHER_DECL

inline bool launch_batch_gather(int S, int Q, int L, int QL, bool cached, const BatchGatherArgs& b, int nblocks,
                                hipStream_t s) {
    if (S == 4 && Q == 1) return launch_batch_gather_4_1(L, QL, cached, b, nblocks, s);
    if (S == 8 && Q == 8) return launch_batch_gather_8_8(L, QL, cached, b, nblocks, s);
    if (S == 10 && Q == 4) return launch_batch_gather_10_4(L, QL, cached, b, nblocks, s);
    if (S == 20 && Q == 11) return launch_batch_gather_20_11(L, QL, cached, b, nblocks, s);
    if (S == 27 && Q == 27) return launch_batch_gather_27_27(L, QL, cached, b, nblocks, s);
    return false;
}

inline bool launch_ju_refresh(int S, int Q, const PSView& ps, const FemView& f, float* ju, hipStream_t s) {
    if (S == 4 && Q == 1) launch_ju_refresh_4_1(ps, f, ju, s);
    else if (S == 8 && Q == 8) launch_ju_refresh_8_8(ps, f, ju, s);
    else if (S == 10 && Q == 4) launch_ju_refresh_10_4(ps, f, ju, s);
    else if (S == 20 && Q == 11) launch_ju_refresh_20_11(ps, f, ju, s);
    else if (S == 27 && Q == 27) launch_ju_refresh_27_27(ps, f, ju, s);
    else return false;
    return true;
}

inline bool launch_fused_tasks(int S, int Q, const FusedTaskArgs& a, hipStream_t s) {
    if (S == 4 && Q == 1) launch_fused_tasks_4_1(a, s);
    else if (S == 8 && Q == 8) launch_fused_tasks_8_8(a, s);
    else if (S == 10 && Q == 4) launch_fused_tasks_10_4(a, s);
    else if (S == 20 && Q == 11) launch_fused_tasks_20_11(a, s);
    else return false;
    return true;
}

inline bool launch_batch_fused_tasks(int S, int Q, const BatchFusedArgs& b, int nblocks, hipStream_t s) {
    if (S == 4 && Q == 1) launch_batch_fused_tasks_4_1(b, nblocks, s);
    else if (S == 8 && Q == 8) launch_batch_fused_tasks_8_8(b, nblocks, s);
    else if (S == 10 && Q == 4) launch_batch_fused_tasks_10_4(b, nblocks, s);
    else if (S == 20 && Q == 11) launch_batch_fused_tasks_20_11(b, nblocks, s);
    else return false;
    return true;
}

inline int refresh_elements_per_block(int S, int Q) {
    if (S == 4 && Q == 1) return refresh_elements_per_block_4_1();
    if (S == 8 && Q == 8) return refresh_elements_per_block_8_8();
    if (S == 10 && Q == 4) return refresh_elements_per_block_10_4();
    if (S == 20 && Q == 11) return refresh_elements_per_block_20_11();
    if (S == 27 && Q == 27) return refresh_elements_per_block_27_27();
    return 0;
}

inline bool launch_batch_ju_refresh(int S, int Q, const BatchRefreshArgs& b, int nblocks, hipStream_t s) {
    if (S == 4 && Q == 1) launch_batch_ju_refresh_4_1(b, nblocks, s);
    else if (S == 8 && Q == 8) launch_batch_ju_refresh_8_8(b, nblocks, s);
    else if (S == 10 && Q == 4) launch_batch_ju_refresh_10_4(b, nblocks, s);
    else if (S == 20 && Q == 11) launch_batch_ju_refresh_20_11(b, nblocks, s);
    else if (S == 27 && Q == 27) launch_batch_ju_refresh_27_27(b, nblocks, s);
    else return false;
    return true;
}

inline bool launch_explicit_elem(int S, int Q, const ExplicitElemArgs& a, hipStream_t s) {
    if (S == 4 && Q == 1) launch_explicit_elem_4_1(a, s);
    else if (S == 8 && Q == 8) launch_explicit_elem_8_8(a, s);
    else if (S == 10 && Q == 4) launch_explicit_elem_10_4(a, s);
    else if (S == 20 && Q == 11) launch_explicit_elem_20_11(a, s);
    else if (S == 27 && Q == 27) launch_explicit_elem_27_27(a, s);
    else return false;
    return true;
}

// Dispatch on (S, Q); returns false for an unsupported tile shape.
inline bool launch_gather(int S, int Q, int L, int QL, const GatherArgs& a, hipStream_t s) {
    if (S == 4 && Q == 1) return launch_gather_4_1(L, QL, a, s);
    if (S == 8 && Q == 8) return launch_gather_8_8(L, QL, a, s);
    if (S == 10 && Q == 4) return launch_gather_10_4(L, QL, a, s);
    if (S == 20 && Q == 11) return launch_gather_20_11(L, QL, a, s);
    if (S == 27 && Q == 27) return launch_gather_27_27(L, QL, a, s);
    return false;
}

}  // namespace oe
