// Static device scratch of the allocation-free residual entry point (see
// oe_scratch.hip).
#pragma once

#include <cstdint>

namespace oe {

struct Scene;

// Largest scene (in vertices) the static scratch holds: 3 floats each, 6 MiB.
constexpr int kScratchMaxVertices = 1 << 19;

// Once per device, with no allocation: look up the device address of the
// static scratch array.  Called when an engine is created on the device.
void scratch_setup(int device);

// Residual force of every free vertex of a VBD scene (3 floats per vertex,
// fixed vertices 0) into the host array r, through the static scratch;
// synchronising.  Calls no hipMalloc / hipFree.
void residual_scratch(Scene& sc, float dt, float* r);

}  // namespace oe
