// The residual entry point that allocates nothing: its kernel writes into a
// static device array instead of a buffer taken from the engine's pool, and
// the result is copied straight into the caller's array.  For callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// `scratch` holds 3 floats per vertex for scenes of up to kScratchMaxVertices
// vertices (6 MiB, held on every device the library runs on until the process
// ends).  Every call on one device shares that device's copy and leaves the
// last call's residuals in it; calls that overlap on one device (two host
// threads, or two engines' streams) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>

#include "oe_scene.hpp"
#include "oe_scratch.hpp"

static __device__ __align__(256) float scratch[3 * oe::kScratchMaxVertices];

namespace oe {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
float* g_scratch[kMaxDevices] = {};   // device address of `scratch` on each device, once set up
}  // namespace

void scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device id out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return;
    int current = -1;
    OE_HIP(hipGetDevice(&current));
    if (current != device) OE_HIP(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (current != device) (void)hipSetDevice(current);
    if (e != hipSuccess || !ptr)
        throw HipError(e, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<float*>(ptr);
}

void residual_scratch(Scene& sc, float dt, float* r) {
    if (!sc.vbd_fem) throw std::invalid_argument("the residual is defined for VBD scenes");
    const int n = sc.ps->nb_particles();
    if (n > kScratchMaxVertices)
        throw std::invalid_argument("oe_scene_residual_scratch: the scene has more vertices than the static "
                                    "scratch holds (524288)");
    const int dev = sc.eng->device;
    float* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (dev >= 0 && dev < kMaxDevices) ptr = g_scratch[dev];
    }
    if (!ptr) throw std::logic_error("static scratch not set up on this device (no engine created on it)");
    const hipStream_t s = sc.eng->stream;
    const size_t bytes = sizeof(float) * 3 * static_cast<size_t>(n);
    // fixed vertices are not written by the kernel and must read 0
    OE_HIP(hipMemsetAsync(ptr, 0, bytes, s));
    sc.vbd_fem->launch_residual(sc.ps.get(), dt, nullptr, ptr);
    OE_HIP(hipMemcpyAsync(r, ptr, bytes, hipMemcpyDeviceToHost, s));
    OE_HIP(hipStreamSynchronize(s));
}

}  // namespace oe
