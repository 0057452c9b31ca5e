// Launchers of the per-vertex integrator and constraint kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "oe_device.hpp"

namespace oe {

// Rigid target of the fixed vertices: target = offset + com + R (x0 - com).
struct FixTransform {
    float com[3];
    float offset[3];
    float R[9];   // column-major
};

void launch_vbd_predict(const PSView& ps, float4* y, float4* prev_it_p, float dt, const float g[3], hipStream_t s);
void launch_vbd_velocity(const PSView& ps, float dt, hipStream_t s);
void launch_chebyshev(const PSView& ps, float4* prev_it_p, float4* prev_it2_p, int it, float omega, hipStream_t s);
void launch_semi_implicit(const PSView& ps, float dt, const float g[3], hipStream_t s);
void launch_pbd_velocity(const PSView& ps, float dt, float global_damping, hipStream_t s);
void launch_fix(const PSView& ps, const int* ids, int n, const FixTransform& t, hipStream_t s);
void launch_box(const PSView& ps, const float pmin[3], const float pmax[3], hipStream_t s);
// launch_fix then launch_box in one pass over all vertices; fixed = per-vertex flag
void launch_fix_box(const PSView& ps, const unsigned char* fixed, const FixTransform& t, const float pmin[3],
                    const float pmax[3], hipStream_t s);
// Explicit integrator: each vertex's internal force from the element partials
// (9 floats per (element, local node), owners order) and the damping term.
// With gravity set, the semi-implicit update (launch_semi_implicit) of the
// vertex follows in the same thread.
void launch_explicit_vertex(const PSView& ps, const OwnersView& own, const float* partial, int count, float damping,
                            const float* gravity, float dt, hipStream_t s);
void launch_copy_colour(const PSView& ps, const int* slot_vid, int first, int count, const float4* p_new,
                        hipStream_t s);

struct SceneDev;
// All vertices of one scene, covered by blocks [block_begin, next segment's
// block_begin) of a launch (block_begin is launch-local).
struct VSeg {
    int scene;
    int count;
    int block_begin;
};
struct BatchVArgs {
    const SceneDev* scenes;
    const VSeg* segs;   // one launch covers segs[0 .. nsegs)
    int nsegs;
    float dt;
    int iteration;
};
void launch_batch_predict(const BatchVArgs& b, int nblocks, hipStream_t s);
void launch_batch_fix_box(const BatchVArgs& b, int nblocks, hipStream_t s);
void launch_batch_velocity(const BatchVArgs& b, int nblocks, hipStream_t s);

}  // namespace oe
