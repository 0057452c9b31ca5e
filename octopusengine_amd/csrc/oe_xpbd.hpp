// Launch interface of the XPBD per-colour projection kernels (oe_xpbd.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "oe_device.hpp"

namespace oe {

// Material ids as seen by device code (same values as oe::Material).
constexpr int NeoHooke_id = 2;
constexpr int Stable_NeoHooke_id = 3;

struct XpbdArgs {
    PSView ps;
    FemView fem;
    MaterialView mt;
    const int* eids;   // colour-major element list
    int first;         // first element of the colour
    int count;         // elements in the colour
    float dt;
};

// v0_form selects the sqrt-constraint formulation with the fixed-vertex mask
// (reference kernel_XPBD_V0); otherwise the energy formulation of V1/V2.
bool launch_xpbd(int S, int Q, bool v0_form, const XpbdArgs& a, hipStream_t s);

}  // namespace oe
