// Launch interface of the per-vertex gather kernel (see oe_gather.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "oe_device.hpp"

namespace oe {

enum GatherMode : int {
    GATHER_VBD_INPLACE = 0,     // Gauss-Seidel inside a proper colour (reference v3)
    GATHER_VBD_JACOBI = 1,      // write p_new, copied after the colour (reference v4)
    GATHER_EXPLICIT_FORCE = 2,  // internal force only, overwrite f (explicit integrator)
};

struct GatherArgs {
    PSView ps;
    FemView fem;
    OwnersView own;
    MaterialView mt;
    const float4* y;   // inertia target (VBD)
    float4* p_new;     // Jacobi output (VBD v4)
    float damping;
    float dt;
    int first_slot;
    int count;
    int mode;
};

bool launch_gather_4_1(int L, int QL, const GatherArgs& a, hipStream_t s);
bool launch_gather_8_8(int L, int QL, const GatherArgs& a, hipStream_t s);
bool launch_gather_10_4(int L, int QL, const GatherArgs& a, hipStream_t s);
bool launch_gather_20_11(int L, int QL, const GatherArgs& a, hipStream_t s);
bool launch_gather_27_27(int L, int QL, const GatherArgs& a, hipStream_t s);

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
