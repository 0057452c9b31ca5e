// HIP plumbing: error checking, an owning device buffer with engine-level
// memory accounting, and the plain-old-data views passed to kernels by value.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace oe {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};

struct BudgetError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define OE_HIP(expr)                                                                             \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess)                                                                    \
            throw ::oe::HipError(_e, std::string(#expr " failed: ") + hipGetErrorString(_e) +  \
                                          " at " __FILE__ ":" + std::to_string(__LINE__));      \
    } while (0)

#define OE_LAUNCH_CHECK() OE_HIP(hipGetLastError())

// Device memory accounting shared by every buffer of one engine: allocation
// beyond `budget` bytes is refused before hipMalloc is called.
struct MemoryPool {
    std::atomic<size_t> used{0};
    size_t budget = SIZE_MAX;
    void reserve(size_t bytes) {
        size_t cur = used.load();
        while (true) {
            if (cur + bytes > budget)
                throw BudgetError("device memory budget exceeded (" + std::to_string(cur + bytes) + " > " +
                                  std::to_string(budget) + " bytes)");
            if (used.compare_exchange_weak(cur, cur + bytes)) return;
        }
    }
    void release(size_t bytes) { used.fetch_sub(bytes); }
};

template <typename T>
class DBuf {
public:
    DBuf() = default;
    DBuf(MemoryPool* pool, size_t n) { allocate(pool, n); }
    DBuf(MemoryPool* pool, const std::vector<T>& h) {
        allocate(pool, h.size());
        upload(h);
    }
    ~DBuf() { reset(); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept { swap(o); }
    DBuf& operator=(DBuf&& o) noexcept {
        reset();
        swap(o);
        return *this;
    }

    void allocate(MemoryPool* pool, size_t n) {
        reset();
        const size_t bytes = (n ? n : 1) * sizeof(T);
        if (pool) pool->reserve(bytes);
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            if (pool) pool->release(bytes);
            throw HipError(e, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        }
        ptr_ = static_cast<T*>(p);
        n_ = n;
        bytes_ = bytes;
        pool_ = pool;
    }
    // Set-up path: synchronous, the host vector may be freed on return.
    void upload(const std::vector<T>& h, hipStream_t s = nullptr) {
        if (h.size() > n_) throw std::out_of_range("upload larger than buffer");
        if (!h.empty()) OE_HIP(hipMemcpyAsync(ptr_, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
        OE_HIP(hipStreamSynchronize(s));
    }
    void upload_raw(const T* h, size_t n, hipStream_t s) {
        if (n > n_) throw std::out_of_range("upload larger than buffer");
        if (n) OE_HIP(hipMemcpyAsync(ptr_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(std::vector<T>& h, hipStream_t s) const {
        h.resize(n_);
        if (n_) OE_HIP(hipMemcpyAsync(h.data(), ptr_, n_ * sizeof(T), hipMemcpyDeviceToHost, s));
        OE_HIP(hipStreamSynchronize(s));
    }
    void reset() {
        if (ptr_) {
            (void)hipFree(ptr_);
            if (pool_) pool_->release(bytes_);
        }
        ptr_ = nullptr;
        n_ = 0;
        bytes_ = 0;
        pool_ = nullptr;
    }
    T* get() const { return ptr_; }
    size_t size() const { return n_; }
    size_t bytes() const { return bytes_; }

private:
    void swap(DBuf& o) {
        std::swap(ptr_, o.ptr_);
        std::swap(n_, o.n_);
        std::swap(bytes_, o.bytes_);
        std::swap(pool_, o.pool_);
    }
    T* ptr_ = nullptr;
    size_t n_ = 0;
    size_t bytes_ = 0;
    MemoryPool* pool_ = nullptr;
};

// Particle state, structure of arrays.  Each field is its own array; a
// position is one 16-byte record so that a gather of a node's position is one
// vector load.  p, last_p (and the VBD inertia target y) hold displacements
// u = x - X from the rest position X = init_p: in FP32 a displacement of a
// millimetre keeps ~1e-10 m resolution where an absolute coordinate of a metre
// keeps ~6e-8 m, and deformation gradients are built from differences of
// these values.
struct PSView {
    int n;
    float4* p;        // displacement
    float4* last_p;   // displacement at the start of the step
    float4* init_p;   // rest (absolute) position
    float4* v;
    float4* f;
    const float* m;
    const float* w;
    int* mask;
    // 1 when the displacement records carry low-order parts in w (the VBD
    // state, see dpack); 0 when w is zero and the state is plain FP32.
    int comp;
    // quantum of the low-order parts (a power of two) and its inverse
    float lo_unit, lo_inv;
};

// Compensated displacement records.  The VBD integrator keeps p, last_p, its
// inertia target y and its Chebyshev history to ~34 bits over the scene's
// range at no extra memory traffic: a record holds hi = float(u) in x, y, z
// and the remainder u - hi, rounded to a multiple of the scene's quantum D,
// as three 10-bit two's-complement integers in the bits of w (x in bits 0-9,
// y in 10-19, z in 20-29).  D = U 2^-34, U the power of two at or above 8
// times the scene's rest bounding-box diagonal (lo_unit_for): for |u| < U,
// |u - hi| <= U 2^-25 = 512 D, stored clamped to +-511 D; larger
// displacements keep the clamped part, and their precision falls back toward
// float's.  Bits 30-31 stay zero, so w always reads as a finite non-negative
// float, and w = 0 is the plain float value.  The element kernels read the
// node values as (hi - ref) + lo in FP32 (rel_node); the predictor, the
// constraints and the per-vertex solve work on hi + lo in FP64 and store the
// result packed again, so that the state is not rounded to float at every
// update.
struct dvec3 {
    double x, y, z;
};
__host__ __device__ __forceinline__ unsigned lo_code(double r, double inv) {
    const double q = fmin(fmax(rint(r * inv), -511.0), 511.0);
    return static_cast<unsigned>(static_cast<int>(q)) & 0x3ffu;
}
// the 10-bit field of w at bit `shift`, sign-extended
__host__ __device__ __forceinline__ int lo_int(unsigned w, int shift) {
    return static_cast<int>(w << (22 - shift)) >> 22;
}
__host__ __device__ __forceinline__ float4 dpack(const dvec3& u, float inv_unit) {
    const double inv = inv_unit;
    const float hx = static_cast<float>(u.x), hy = static_cast<float>(u.y), hz = static_cast<float>(u.z);
    const unsigned w = lo_code(u.x - hx, inv) | lo_code(u.y - hy, inv) << 10 | lo_code(u.z - hz, inv) << 20;
    return make_float4(hx, hy, hz, __builtin_bit_cast(float, w));
}
__host__ __device__ __forceinline__ dvec3 dunpack(float4 h, float unit) {
    const unsigned w = __builtin_bit_cast(unsigned, h.w);
    const double d = unit;
    return dvec3{h.x + lo_int(w, 0) * d, h.y + lo_int(w, 10) * d, h.z + lo_int(w, 20) * d};
}
// D for a scene whose rest bounding box has diagonal `diag` (metres)
inline float lo_unit_for(double diag) {
    const double L = fmin(fmax(diag, 1e-6), 1e9);
    return static_cast<float>(ldexp(1.0, static_cast<int>(ceil(log2(8.0 * L))) - 34));
}
// A node's value for an element's deformation gradient: (hi - ref) + lo in
// FP32, where ref is the high part of a displacement common to the element's
// nodes.  The gradient depends only on differences of the node values (an
// element's shape-function derivatives sum to zero), so ref drops out; the
// difference of two nearby large displacements is exact (Sterbenz) where the
// sum over the nodes of the large values themselves would cancel in FP32, and
// the low-order part restores the bits the high part rounded away.  On
// records without low-order parts (w = 0) it is hi - ref.
__device__ __forceinline__ void rel_node(float4 u, float4 ref, float unit, float out[3]) {
    const unsigned w = __float_as_uint(u.w);
    out[0] = fmaf(static_cast<float>(lo_int(w, 0)), unit, u.x - ref.x);
    out[1] = fmaf(static_cast<float>(lo_int(w, 10)), unit, u.y - ref.y);
    out[2] = fmaf(static_cast<float>(lo_int(w, 20)), unit, u.z - ref.z);
}

// Per-mesh FEM data (reference GPU_FEM_Pameters, include/GPU/GPU_FEM.h:49-59).
// Two layouts of the element records:
//   affine == 0: E * Q * 10 floats, JX_inv (column-major) then V = |det J_q| w_q
//                for every (element, quadrature point);
//   affine == 1: E * 10 floats, JX_inv then |det J| once per element, and
//                V_q = |det J| * wq[q].  Used when every element's map from
//                the reference element is affine (straight-sided tetrahedra,
//                parallelepiped hexahedra), where J does not depend on q.
// wq holds the quadrature weights in the affine layout and ones otherwise, so
// that V = rec[9] * wq[q] in both.
struct FemView {
    int E, S, Q;
    const int* topo;     // E * S
    const float* jxv;
    const float* dN;     // Q * S * 3
    int affine;
    const float* wq;     // Q
};

// Loads through pointers known to address device (global) memory.  Every
// buffer a view points at is a device allocation; a pointer read from a table
// in memory is otherwise a generic one, whose loads the hardware counts
// together with LDS operations, so that a wait for one waits for all.
typedef float oe_float4_vec __attribute__((ext_vector_type(4)));
template <class T>
__device__ __forceinline__ T gload(const T* p, size_t i) {
    return ((const __attribute__((address_space(1))) T*)p)[i];
}
__device__ __forceinline__ float4 gload(const float4* p, size_t i) {
    const oe_float4_vec v = ((const __attribute__((address_space(1))) oe_float4_vec*)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
typedef int oe_int2_vec __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int2 gload(const int2* p, size_t i) {
    const oe_int2_vec v = ((const __attribute__((address_space(1))) oe_int2_vec*)p)[i];
    return make_int2(v.x, v.y);
}
typedef int oe_int4_vec __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 gload(const int4* p, size_t i) {
    const oe_int4_vec v = ((const __attribute__((address_space(1))) oe_int4_vec*)p)[i];
    return make_int4(v.x, v.y, v.z, v.w);
}
template <class T>
__device__ __forceinline__ void gstore(T* p, size_t i, T v) {
    ((__attribute__((address_space(1))) T*)p)[i] = v;
}
__device__ __forceinline__ void gstore(float4* p, size_t i, float4 v) {
    const oe_float4_vec w = {v.x, v.y, v.z, v.w};
    ((__attribute__((address_space(1))) oe_float4_vec*)p)[i] = w;
}

// Element record of quadrature point q of element e in either layout
// (branch-free: the layout only changes the record index).
// The record index is formed in 32 bits (the engine refuses meshes with
// E * Q >= 2^32): a 64-bit q held across the Q2 task loop costs registers.
__device__ __forceinline__ void load_element_record(const FemView& f, int e, int q, float JXi[9], float& V) {
    const size_t idx = f.affine ? static_cast<unsigned>(e)
                                : static_cast<unsigned>(e) * static_cast<unsigned>(f.Q) + static_cast<unsigned>(q);
#pragma unroll
    for (int i = 0; i < 9; ++i) JXi[i] = gload(f.jxv, idx * 10 + i);
    V = gload(f.jxv, idx * 10 + 9) * gload(f.wq, q);
}

// Colour-major owners CSR (reference GPU_Owners_Parameters,
// include/GPU/VBD/GPU_VBD_FEM.h:36-42).
struct OwnersView {
    const int* slot_vid;
    const int2* slot_nbo;   // (incident element count, offset into own)
    const int* own;         // (eid << 5) | local node
};

// What the vertex solve of one slot reads.  The latency forms of the fused
// task kernels load it during the element work of their block, so that the
// loads overlap that work (a slot's vertex is written only by the block that
// solves it).
struct SlotState {
    int vid;
    int n_inc, first_inc;   // incident elements, first owners index
    int free;               // mask (0: fixed, not solved)
    float4 p, last_p, y;
    float m;
};

// The latency forms load a slot's state in two steps, each issued right after
// a step of the block's own task chain (owners entry, then node ids), so that
// no wait of that chain waits for the state: vector-memory loads complete in
// order, and a wait for one load waits for every load issued before it.
__device__ __forceinline__ void slot_state_level1(const OwnersView& own, int slot, SlotState& st) {
    st.vid = gload(own.slot_vid, slot);
    const int2 nbo = gload(own.slot_nbo, slot);
    st.n_inc = nbo.x;
    st.first_inc = nbo.y;
}
__device__ __forceinline__ void slot_state_level2(const PSView& ps, const float4* y, SlotState& st) {
    st.free = gload(ps.mask, st.vid);
    st.p = gload(static_cast<const float4*>(ps.p), st.vid);
    st.last_p = gload(static_cast<const float4*>(ps.last_p), st.vid);
    st.y = gload(y, st.vid);
    st.m = gload(ps.m, st.vid);
}
// Where the state is first used: the empty asm takes the values as inputs and
// outputs, so the compiler cannot convert them (and wait for their loads)
// before this point.
__device__ __forceinline__ void slot_state_settle(SlotState& st) {
    asm volatile(""
                 : "+v"(st.p.x), "+v"(st.p.y), "+v"(st.p.z), "+v"(st.p.w), "+v"(st.last_p.x), "+v"(st.last_p.y),
                   "+v"(st.last_p.z), "+v"(st.last_p.w), "+v"(st.y.x), "+v"(st.y.y), "+v"(st.y.z), "+v"(st.y.w),
                   "+v"(st.m), "+v"(st.free), "+v"(st.vid), "+v"(st.n_inc), "+v"(st.first_inc));
}

__device__ __forceinline__ SlotState load_slot_state(const PSView& ps, const OwnersView& own, const float4* y,
                                                     int slot) {
    SlotState st;
    st.vid = gload(own.slot_vid, slot);
    const int2 nbo = gload(own.slot_nbo, slot);
    st.n_inc = nbo.x;
    st.first_inc = nbo.y;
    st.free = gload(ps.mask, st.vid);
    st.p = gload(static_cast<const float4*>(ps.p), st.vid);
    st.last_p = gload(static_cast<const float4*>(ps.last_p), st.vid);
    st.y = gload(y, st.vid);
    st.m = gload(ps.m, st.vid);
    return st;
}

struct MaterialView {
    int material;
    float lambda, mu, alpha;   // alpha = 1 (NeoHooke) or 1 + mu/lambda (Stable NeoHooke)
};

}  // namespace oe
