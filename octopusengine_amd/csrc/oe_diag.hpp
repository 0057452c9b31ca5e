// Launcher of the per-element diagnostics kernel (see oe_diag.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "oe_device.hpp"

namespace oe {

// Any output pointer may be null.  Returns false for an unsupported element.
bool launch_element_diagnostics(const PSView& ps, const FemView& f, const MaterialView& mt, float* stress,
                                float* volume, float* volume_diff, float* inverted, hipStream_t s);

}  // namespace oe
