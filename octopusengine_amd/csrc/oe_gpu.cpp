// Host side of the GPU dynamics objects: buffers, colour schedules and the
// launch sequence of every step.  The step sequences follow
//   GPU_VBD::step            src/GPU/VBD/GPU_VBD.cu:38-61
//   GPU_VBD_FEM::step        src/GPU/VBD/GPU_VBD_FEM.cu:712-756
//   GPU_PBD::step            src/GPU/PBD/GPU_PBD.cu:20-26
//   GPU_PBD_FEM::step        src/GPU/PBD/GPU_PBD_FEM.cu:640-717
//   GPU_SemiExplicit::step   src/GPU/GPU_Integrator.cu:36-42
//   GPU_FEM_Explicit::step   src/GPU/Explicit/GPU_FEM_Explicit.cu:109-115
//   GPU_Fix_Constraint       src/GPU/GPU_Constraint.cu:94-114
#include "oe_gpu.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>

#include "oe_diag.hpp"
#include "oe_gather.hpp"
#include "oe_q2.hpp"
#include "oe_xpbd.hpp"

namespace oe {

// ---------------------------------------------------------------- Engine
Engine::Engine(int dev, size_t budget_bytes) : device(dev) {
    OE_HIP(hipSetDevice(dev));
    // default-priority stream (no priority other than the default anywhere)
    OE_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    pool.budget = budget_bytes;
}

Engine::~Engine() {
    if (stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}

// ------------------------------------------------------ GPU_ParticleSystem
namespace {
std::vector<float4> to_f4(const std::vector<float>& xyz) {
    const size_t n = xyz.size() / 3;
    std::vector<float4> out(n);
    for (size_t i = 0; i < n; ++i) out[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.f);
    return out;
}
void from_f4(const std::vector<float4>& v, std::vector<float>& xyz) {
    xyz.resize(v.size() * 3);
    for (size_t i = 0; i < v.size(); ++i) {
        xyz[3 * i] = v[i].x;
        xyz[3 * i + 1] = v[i].y;
        xyz[3 * i + 2] = v[i].z;
    }
}
}  // namespace

GPU_ParticleSystem::GPU_ParticleSystem(Engine* e, const std::vector<float>& xyz, const std::vector<float>& masses)
    : eng(e), n(static_cast<int>(xyz.size() / 3)) {
    if (static_cast<int>(masses.size()) != n) throw std::invalid_argument("mass count != vertex count");
    const std::vector<float4> X = to_f4(xyz);
    const std::vector<float4> zero(n, make_float4(0.f, 0.f, 0.f, 0.f));
    std::vector<float> inv(n);
    for (int i = 0; i < n; ++i) inv[i] = 1.f / masses[i];   // w = 1/m (GPU_ParticleSystem.cu:11-12)
    MemoryPool* pool = &eng->pool;
    rest = xyz;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            const double x = xyz[3 * static_cast<size_t>(i) + k];
            lo[k] = i ? std::min(lo[k], x) : x;
            hi[k] = i ? std::max(hi[k], x) : x;
        }
    lo_unit = lo_unit_for(std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                    (hi[2] - lo[2]) * (hi[2] - lo[2])));
    p = DBuf<float4>(pool, zero);        // displacement from the rest position
    last_p = DBuf<float4>(pool, zero);
    init_p = DBuf<float4>(pool, X);
    v = DBuf<float4>(pool, std::vector<float4>(n, make_float4(0.f, 0.f, 0.f, 0.f)));
    f = DBuf<float4>(pool, std::vector<float4>(n, make_float4(0.f, 0.f, 0.f, 0.f)));
    m = DBuf<float>(pool, masses);
    w = DBuf<float>(pool, inv);
    mask = DBuf<int>(pool, std::vector<int>(n, 1));   // mask starts at 1 (GPU_ParticleSystem.cu:21)
}

PSView GPU_ParticleSystem::get_parameters() const {
    return PSView{n,          p.get(),    last_p.get(), init_p.get(), v.get(), f.get(), m.get(), w.get(), mask.get(),
                  comp ? 1 : 0, lo_unit,    1.f / lo_unit};
}

// the records' w lanes are zero until a compensated integrator writes them
void GPU_ParticleSystem::enable_compensation() { comp = true; }

std::vector<double> GPU_ParticleSystem::full(const DBuf<float4>& rec) const {
    std::vector<float4> h;
    rec.download(h, eng->stream);
    std::vector<double> u(3 * h.size());
    for (size_t i = 0; i < h.size(); ++i) {
        const dvec3 d = comp ? dunpack(h[i], lo_unit) : dvec3{h[i].x, h[i].y, h[i].z};
        u[3 * i] = d.x;
        u[3 * i + 1] = d.y;
        u[3 * i + 2] = d.z;
    }
    return u;
}

void GPU_ParticleSystem::get_position(std::vector<float>& xyz) const {
    const std::vector<double> u = full(p);
    xyz.resize(u.size());
    for (size_t i = 0; i < xyz.size(); ++i) xyz[i] = static_cast<float>(rest[i] + u[i]);
}
void GPU_ParticleSystem::get_prev_position(std::vector<float>& xyz) const {
    const std::vector<double> u = full(last_p);
    xyz.resize(u.size());
    for (size_t i = 0; i < xyz.size(); ++i) xyz[i] = static_cast<float>(rest[i] + u[i]);
}

// xyz (absolute, 3n floats) -> displacement records from the double
// difference: packed with its low-order part when the state is compensated,
// rounded to float with w = 0 otherwise
void GPU_ParticleSystem::upload_split(const float* xyz, DBuf<float4>& rec) {
    std::vector<float4> h(n);
    for (int i = 0; i < n; ++i) {
        const dvec3 u{static_cast<double>(xyz[3 * i]) - rest[3 * i], static_cast<double>(xyz[3 * i + 1]) - rest[3 * i + 1],
                      static_cast<double>(xyz[3 * i + 2]) - rest[3 * i + 2]};
        h[i] = comp ? dpack(u, 1.f / lo_unit)
                    : make_float4(static_cast<float>(u.x), static_cast<float>(u.y), static_cast<float>(u.z), 0.f);
    }
    rec.upload(h, eng->stream);
}
void GPU_ParticleSystem::get_init_position(std::vector<float>& xyz) const {
    std::vector<float4> h;
    init_p.download(h, eng->stream);
    from_f4(h, xyz);
}
void GPU_ParticleSystem::get_velocity(std::vector<float>& xyz) const {
    std::vector<float4> h;
    v.download(h, eng->stream);
    from_f4(h, xyz);
}
void GPU_ParticleSystem::get_forces(std::vector<float>& xyz) const {
    std::vector<float4> h;
    f.download(h, eng->stream);
    from_f4(h, xyz);
}
void GPU_ParticleSystem::get_masks(std::vector<int>& out) const { mask.download(out, eng->stream); }
void GPU_ParticleSystem::get_mass(std::vector<float>& out) const { m.download(out, eng->stream); }
void GPU_ParticleSystem::get_inv_mass(std::vector<float>& out) const { w.download(out, eng->stream); }
std::vector<float> GPU_ParticleSystem::to_displacement(const float* xyz) const {
    std::vector<float> u(3 * static_cast<size_t>(n));
    for (size_t i = 0; i < u.size(); ++i) u[i] = static_cast<float>(static_cast<double>(xyz[i]) - rest[i]);
    return u;
}
void GPU_ParticleSystem::set_position(const std::vector<float>& xyz) { upload_split(xyz.data(), p); }
void GPU_ParticleSystem::set_velocity(const std::vector<float>& xyz) { v.upload(to_f4(xyz), eng->stream); }
void GPU_ParticleSystem::set_state(const float* pp, const float* lp, const float* vv) {
    if (pp) upload_split(pp, p);
    if (lp) upload_split(lp, last_p);
    if (vv) v.upload(to_f4(std::vector<float>(vv, vv + 3 * static_cast<size_t>(n))), eng->stream);
}
size_t GPU_ParticleSystem::device_bytes() const {
    return p.bytes() + last_p.bytes() + init_p.bytes() + v.bytes() + f.bytes() + m.bytes() + w.bytes() + mask.bytes();
}

// ------------------------------------------------------------ Integrators
void GPU_Integrator::init_dynamics(GPU_ParticleSystem* ps, float dt) {
    for (auto& d : _dynamics)
        if (d->active) d->start(ps, dt);
}
void GPU_Integrator::eval_dynamics(GPU_ParticleSystem* ps, float dt) {
    for (auto& d : _dynamics)
        if (d->active) d->step(ps, dt);
}
// A fixation directly followed by a box limit (the order a scene adds them in)
// runs as one fused pass; everything else in insertion order.
void GPU_Integrator::eval_constraints(GPU_ParticleSystem* ps, float dt) {
    for (size_t i = 0; i < _constraints.size(); ++i) {
        GPU_Dynamic* c = _constraints[i].get();
        if (!c->active) continue;
        if (i + 1 < _constraints.size() && _constraints[i + 1]->active) {
            auto* fix = dynamic_cast<GPU_Fix_Constraint*>(c);
            auto* box = dynamic_cast<GPU_Box_Limit*>(_constraints[i + 1].get());
            if (fix && box) {
                launch_fix_box(ps->get_parameters(), fix->d_flags.get(), fix->t, box->pmin, box->pmax, fix->stream);
                ++i;
                continue;
            }
        }
        c->step(ps, dt);
    }
}

void GPU_SemiExplicit::step(GPU_ParticleSystem* ps, float dt) {
    // A lone element-centric explicit force pass integrates in its own vertex
    // pass (the same update as launch_semi_implicit, one launch fewer).
    if (_dynamics.size() == 1 && _dynamics[0]->active)
        if (auto* fe = dynamic_cast<GPU_FEM_Explicit*>(_dynamics[0].get()); fe && fe->step_integrate(ps, dt, gravity)) {
            if (apply_constraints) eval_constraints(ps, dt);
            return;
        }
    eval_dynamics(ps, dt);
    launch_semi_implicit(ps->get_parameters(), dt, gravity, stream);
    if (apply_constraints) eval_constraints(ps, dt);
}

GPU_VBD::GPU_VBD(GPU_ParticleSystem* ps, int it, float rho) : iteration(it), _rho(rho) {
    if (it < 0) throw std::invalid_argument("iteration count must be >= 0");
    MemoryPool* pool = &ps->eng->pool;
    stream = ps->eng->stream;
    ps->enable_compensation();   // the VBD state keeps low-order parts in the records' w lanes
    y.allocate(pool, ps->n);
    OE_HIP(hipMemcpyAsync(y.get(), ps->p.get(), sizeof(float4) * ps->n, hipMemcpyDeviceToDevice, stream));
    if (rho > 0.f) {
        prev_it_p.allocate(pool, ps->n);
        prev_it2_p.allocate(pool, ps->n);
        OE_HIP(hipMemcpyAsync(prev_it_p.get(), ps->p.get(), sizeof(float4) * ps->n, hipMemcpyDeviceToDevice, stream));
        OE_HIP(hipMemcpyAsync(prev_it2_p.get(), ps->p.get(), sizeof(float4) * ps->n, hipMemcpyDeviceToDevice, stream));
    }
    OE_HIP(hipStreamSynchronize(stream));
}

void GPU_VBD::add_dynamics(GPU_Dynamic* d) {
    if (auto* fem = dynamic_cast<GPU_VBD_FEM*>(d)) fem->y = y.get();
    GPU_Integrator::add_dynamics(d);
}

void GPU_VBD::step(GPU_ParticleSystem* ps, float dt) {
    const PSView v = ps->get_parameters();
    // rho > eps with the reference's eps = FLT_MIN * 1e6 (include/Core/Base.h:11)
    const bool cheb = _rho > std::numeric_limits<float>::min() * 1e6f;
    init_dynamics(ps, dt);
    launch_vbd_predict(v, y.get(), cheb ? prev_it_p.get() : nullptr, dt, gravity, stream);
    float omega = 1.f;
    for (int j = 0; j < iteration; ++j) {
        eval_dynamics(ps, dt);
        eval_constraints(ps, dt);
        if (cheb) {
            if (j == 1) omega = 2.f / (2.f - _rho * _rho);
            else if (j > 1) omega = 4.f / (4.f - _rho * _rho * omega);
            launch_chebyshev(v, prev_it_p.get(), prev_it2_p.get(), j, omega, stream);
        }
    }
    launch_vbd_velocity(v, dt, stream);
}

void GPU_PBD::step(GPU_ParticleSystem* ps, float dt) {
    const PSView v = ps->get_parameters();
    launch_semi_implicit(v, dt, gravity, stream);
    eval_dynamics(ps, dt);
    eval_constraints(ps, dt);
    launch_pbd_velocity(v, dt, _global_damping, stream);
}

// ---------------------------------------------------------------- GPU_FEM
GPU_FEM::GPU_FEM(Engine* e, Element el, const HostMesh& mesh, float young, float poisson, Material material)
    : eng(e), element(el) {
    if (mesh.element != el) throw std::invalid_argument("mesh element type mismatch");
    if (material != NeoHooke && material != Stable_NeoHooke)
        throw std::invalid_argument("material not supported on the GPU path (NeoHooke / Stable_NeoHooke)");
    FemConstants fc = build_fem_constants(mesh);
    S = fc.S;
    Q = fc.Q;
    E = fc.E;
    if (static_cast<uint64_t>(E) * static_cast<uint64_t>(Q) >= (uint64_t{1} << 32))
        throw std::length_error("mesh too large: elements x quadrature points must stay below 2^32");
    mt.material = material;
    lame_parameters(young, poisson, &mt.lambda, &mt.mu);
    mt.alpha = (material == Stable_NeoHooke) ? 1.f + mt.mu / mt.lambda : 1.f;
    topo = DBuf<int>(&eng->pool, std::vector<int>(mesh.topo.begin(), mesh.topo.end()));
    std::vector<float> rec, w;
    affine = build_affine_records(mesh, rec, w);
    if (!affine) w.assign(Q, 1.f);   // V is stored whole in the per-point records
    jxv = DBuf<float>(&eng->pool, affine ? rec : fc.jxv);
    wq = DBuf<float>(&eng->pool, w);
    dN = DBuf<float>(&eng->pool, fc.dN);
    elem_volume = std::move(fc.elem_volume);
}

void GPU_FEM::diagnostics(GPU_ParticleSystem* ps, float* stress, float* volume, float* volume_diff, float* inverted) {
    DBuf<float> out(&eng->pool, static_cast<size_t>(E) * 4);
    float* base = out.get();
    if (!launch_element_diagnostics(ps->get_parameters(), fem_view(), mt, base, base + E, base + 2 * static_cast<size_t>(E),
                                    base + 3 * static_cast<size_t>(E), eng->stream))
        throw std::invalid_argument("unsupported element");
    OE_LAUNCH_CHECK();
    std::vector<float> h;
    out.download(h, eng->stream);
    float* dst[4] = {stress, volume, volume_diff, inverted};
    for (int k = 0; k < 4; ++k)
        if (dst[k]) std::memcpy(dst[k], h.data() + static_cast<size_t>(k) * E, sizeof(float) * E);
}

FemView GPU_FEM::fem_view() const {
    return FemView{E, S, Q, topo.get(), jxv.get(), dN.get(), affine ? 1 : 0, wq.get()};
}

GatherTile default_tile(Element e) {
    switch (e) {
        case Tetra: return {8, 1, PASS_GATHER};
        case Hexa: return {16, 1, PASS_GATHER};
        case Tetra10: return {64, 1, PASS_TASKS};
        case Tetra20: return {64, 1, PASS_GATHER};
        case Hexa27: return {64, 1, PASS_TASKS, kQ2PencilIncLatency};
        default: return {16, 1, PASS_GATHER};
    }
}

namespace {
void upload_owners(Engine* eng, const OwnersCSR& o, DBuf<int>& slot_vid, DBuf<int2>& slot_nbo, DBuf<int>& own) {
    std::vector<int2> nbo(o.slot_vid.size());
    for (size_t i = 0; i < nbo.size(); ++i) nbo[i] = make_int2(o.slot_nb[i], o.slot_off[i]);
    slot_vid = DBuf<int>(&eng->pool, std::vector<int>(o.slot_vid.begin(), o.slot_vid.end()));
    slot_nbo = DBuf<int2>(&eng->pool, nbo);
    own = DBuf<int>(&eng->pool, std::vector<int>(o.own.begin(), o.own.end()));
}
}  // namespace

// ------------------------------------------------------------ GPU_VBD_FEM
GPU_VBD_FEM::GPU_VBD_FEM(Engine* e, Element el, const HostMesh& mesh, Material material, float young, float poisson,
                         float damp, VBD_Version v, int target)
    : GPU_FEM(e, el, mesh, young, poisson, material), version(v), target_color(target), damping(damp),
      tile(default_tile(el)) {
    const MultiGraph g = build_primal_graph(el, mesh.topo, mesh.nb_vertices());
    if (g.n != mesh.nb_vertices()) throw std::invalid_argument("mesh has vertices used by no element");
    colors = color_dsat(g);
    if (version == Better_Coloration) color_reduction(g, colors, target_color);   // throws for target <= 0
    conflicts = count_conflicts(g, colors);
    if (conflicts != 0 && version != Better_Coloration)
        throw std::logic_error("DSAT produced a conflicting colouring");
    owners = build_owners(mesh, colors, true);
    upload_owners(eng, owners, d_slot_vid, d_slot_nbo, d_own);
    for (int32_t nb : owners.slot_nb) max_incidences = std::max(max_incidences, static_cast<int>(nb));
    if (version == Better_Coloration) p_new.allocate(&eng->pool, mesh.nb_vertices());
    if (el == Hexa27) has_q2_basis = q2_basis_1d(q2_basis);
    // allocated here, not in step(), which may run under HIP graph capture
    if (cached()) ensure_ju();
    if (tasks()) {
        try {
            prepare_tasks();
        } catch (const std::invalid_argument&) {
            // a vertex with more incidences than a task block holds: the
            // default pass falls back to the gather kernel
            tile = element == Hexa27 ? GatherTile{64, 1, PASS_GATHER} : GatherTile{16, 2, PASS_GATHER};
        }
    }
}

GPU_VBD_FEM::~GPU_VBD_FEM() {
    for (hipEvent_t ev : ev_pool) (void)hipEventDestroy(ev);
}

namespace {
void upload_vertex_blocks(Engine* eng, const OwnersCSR& o, int max_inc, int max_slots, DBuf<int4>& d,
                          std::vector<int32_t>& color_begin) {
    std::vector<int32_t> blocks;
    std::vector<int32_t> begin;
    build_vertex_blocks(o, max_inc, max_slots, blocks, begin);
    std::vector<int4> b4(blocks.size() / 4);
    for (size_t i = 0; i < b4.size(); ++i)
        b4[i] = make_int4(blocks[4 * i], blocks[4 * i + 1], blocks[4 * i + 2], blocks[4 * i + 3]);
    d = DBuf<int4>(&eng->pool, b4);
    color_begin = std::move(begin);
}
}  // namespace

int GPU_VBD_FEM::fit_inc(int inc) const {
    const int n = inc > 0 ? inc : (element == Hexa27 ? kQ2FusedIncPerBlock : max_task_inc(element));
    return element == Hexa27 ? q2_pencil_fit(n, max_incidences) : n;
}

const TaskTables& GPU_VBD_FEM::throughput_tables(int inc) {
    const int n = fit_inc(inc);
    if (n > max_task_inc(element) || (element == Hexa27 && !q2_pencil(n) && n > kQ2FusedIncPerBlock))
        throw std::invalid_argument("more incidences per block than the task kernel holds");
    auto it = thr_tables.find(n);
    if (it != thr_tables.end()) return it->second;
    TaskTables t;
    upload_vertex_blocks(eng, owners, n, element == Hexa27 ? kQ2FusedMaxSlots : kFusedMaxSlots, t.blocks,
                         t.color_begin);
    return thr_tables.emplace(n, std::move(t)).first->second;
}

const TaskTables& GPU_VBD_FEM::built_tables(int inc) const {
    const auto it = thr_tables.find(fit_inc(inc));
    if (it == thr_tables.end()) throw std::logic_error("element-task vertex blocks not prepared");
    return it->second;
}

const TaskTables* GPU_VBD_FEM::latency_tables() {
    if (!lat_built) {
        lat_built = true;
        const bool q2 = element == Hexa27;
        try {
            upload_vertex_blocks(eng, owners, q2 ? kQ2FusedIncPerBlockLat : kFusedIncPerBlockLat,
                                 q2 ? kQ2FusedMaxSlots : kFusedMaxSlots, lat_tables.blocks, lat_tables.color_begin);
        } catch (const std::invalid_argument&) {
            // a vertex with too many incidences for a latency block: throughput form only
            lat_tables = TaskTables();
        }
    }
    return lat_tables.blocks.get() ? &lat_tables : nullptr;
}

void GPU_VBD_FEM::prepare_tasks() {
    throughput_tables(tile.inc);
    latency_tables();
}

float* GPU_VBD_FEM::ensure_ju() {
    if (!ju.get()) ju.allocate(&eng->pool, static_cast<size_t>(E) * Q * 9);
    return ju.get();
}

void GPU_VBD_FEM::launch_residual(GPU_ParticleSystem* ps, float dt, float4* out4, float* out3) {
    if (!y) throw std::logic_error("GPU_VBD_FEM used outside a GPU_VBD integrator");
    if (!(dt > 0.f)) throw std::invalid_argument("residual needs dt > 0");
    GatherArgs a{};
    a.ps = ps->get_parameters();
    a.fem = fem_view();
    a.own = OwnersView{d_slot_vid.get(), d_slot_nbo.get(), d_own.get()};
    a.mt = mt;
    a.y = y;
    a.p_new = out4;
    a.out3 = out3;
    a.ju = nullptr;
    a.damping = damping;
    a.dt = dt;
    a.first_slot = 0;
    a.count = ps->nb_particles();
    a.mode = GATHER_VBD_RESIDUAL;
    if (!launch_gather(S, Q, tile.L, tile.QL, a, eng->stream)) throw std::invalid_argument("unsupported gather tile");
    OE_LAUNCH_CHECK();
}

void GPU_VBD_FEM::residual(GPU_ParticleSystem* ps, float dt, std::vector<float>& r) {
    if (!(dt > 0.f)) throw std::invalid_argument("residual needs dt > 0");
    const int n = ps->nb_particles();
    if (!d_residual.get()) d_residual.allocate(&eng->pool, n);
    // fixed vertices are not written: clear every call (the fixed set may
    // have changed since the last one)
    OE_HIP(hipMemsetAsync(d_residual.get(), 0, sizeof(float4) * n, eng->stream));
    launch_residual(ps, dt, d_residual.get(), nullptr);
    std::vector<float4> h;
    d_residual.download(h, eng->stream);
    r.resize(3 * static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        r[3 * i] = h[i].x;
        r[3 * i + 1] = h[i].y;
        r[3 * i + 2] = h[i].z;
    }
}

void GPU_VBD_FEM::set_shuffle_seed(uint64_t seed) {
    shuffle = seed != 0;
    rng_state = seed;
}

namespace {
// SplitMix64 (Steele, Lea and Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014): add the golden-ratio increment, then mix.
uint64_t splitmix64(uint64_t& state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

void GPU_VBD_FEM::next_color_order(std::vector<int>& order) {
    order.resize(colors.nb_color);
    std::iota(order.begin(), order.end(), 0);
    if (!shuffle) return;
    for (int i = colors.nb_color - 1; i > 0; --i) {
        const int j = static_cast<int>(splitmix64(rng_state) % static_cast<uint64_t>(i + 1));
        std::swap(order[i], order[j]);
    }
}

void GPU_VBD_FEM::step(GPU_ParticleSystem* ps, float dt) {
    if (!y) throw std::logic_error("GPU_VBD_FEM used outside a GPU_VBD integrator");
    std::vector<int> order;
    next_color_order(order);
    GatherArgs a{};
    a.ps = ps->get_parameters();
    a.fem = fem_view();
    a.own = OwnersView{d_slot_vid.get(), d_slot_nbo.get(), d_own.get()};
    a.mt = mt;
    a.y = y;
    a.p_new = p_new.get();
    a.damping = damping;
    a.dt = dt;
    const bool jacobi = (version == Better_Coloration);
    a.mode = jacobi ? GATHER_VBD_JACOBI : GATHER_VBD_INPLACE;
    hipStream_t s = eng->stream;
    a.ju = nullptr;
    if (cached()) {
        // gradients of this iteration's starting positions
        if (!ju.get()) throw std::logic_error("gradient cache not allocated");
        a.ju = ju.get();
        if (!launch_ju_refresh(S, Q, a.ps, a.fem, a.ju, s)) throw std::invalid_argument("unsupported element");
    }
    const bool by_tasks = tasks();
    const TaskTables* thr = by_tasks ? &built_tables(tile.inc) : nullptr;
    const TaskTables* lt = by_tasks && lat_tables.blocks.get() ? &lat_tables : nullptr;
    for (int c : order) {
        a.first_slot = owners.color_begin[c];
        a.count = owners.color_begin[c + 1] - a.first_slot;
        if (time_gather) {
            while (ev_pool.size() < ev_used + 2) {
                hipEvent_t ev;
                OE_HIP(hipEventCreate(&ev));
                ev_pool.push_back(ev);
            }
            OE_HIP(hipEventRecord(ev_pool[ev_used], s));
        }
        if (by_tasks) {
            // the latency form for colours with few throughput-form blocks
            // (by default under 2 per CU)
            const bool lat = lt && thr->color_blocks(c) < tile.lat_blocks;
            const TaskTables& tt = lat ? *lt : *thr;
            const int4* blocks = tt.blocks.get();
            const int b0 = tt.color_begin[c], nb = tt.color_blocks(c);
            if (element == Hexa27) {
                const Q2FusedArgs qa{a.ps, a.fem, a.own, a.mt, q2_basis, y, blocks, damping, dt, b0, nb, lat ? 1 : 0};
                if (!lat && q2_pencil(tile.inc))
                    launch_q2_pencil(qa, s);
                else
                    launch_q2_fused(qa, s);
            }
            else if (!launch_fused_tasks(S, Q, FusedTaskArgs{a.ps, a.fem, a.own, a.mt, y, blocks, damping, dt, b0, nb,
                                                             lat ? 1 : 0},
                                         s))
                throw std::invalid_argument("unsupported element");
        } else if (!launch_gather(S, Q, tile.L, tile.QL, a, s)) {
            throw std::invalid_argument("unsupported gather tile shape");
        }
        if (time_gather) {
            OE_HIP(hipEventRecord(ev_pool[ev_used + 1], s));
            ev_used += 2;
        }
        if (jacobi) launch_copy_colour(a.ps, d_slot_vid.get(), a.first_slot, a.count, p_new.get(), s);
    }
}

void GPU_VBD_FEM::collect_gather_timing() {
    if (ev_used == 0) return;
    OE_HIP(hipStreamSynchronize(eng->stream));
    for (size_t i = 0; i + 1 < ev_used; i += 2) {
        float ms = 0.f;
        OE_HIP(hipEventElapsedTime(&ms, ev_pool[i], ev_pool[i + 1]));
        gather_ms += ms;
        gather_launches++;
    }
    ev_used = 0;
}

// ------------------------------------------------------------ GPU_PBD_FEM
GPU_PBD_FEM::GPU_PBD_FEM(Engine* e, Element el, const HostMesh& mesh, float young, float poisson, Material material,
                         XPBD_FEM_VERSION v)
    : GPU_FEM(e, el, mesh, young, poisson, material), version(v) {
    const MultiGraph g = build_dual_vertex_graph(el, mesh.topo, mesh.nb_vertices());
    colors = color_greedy_slf(g);
    if (count_conflicts(g, colors) != 0) throw std::logic_error("element colouring has conflicts");
    sched = build_element_schedule(colors, E);
    d_eids = DBuf<int>(&eng->pool, std::vector<int>(sched.eids.begin(), sched.eids.end()));
}

void GPU_PBD_FEM::step(GPU_ParticleSystem* ps, float dt) {
    const bool v0 = (S == 4) || version == XPBD;
    // OptiGroup launches nothing for elements other than P1 in the reference
    // (src/GPU/PBD/GPU_PBD_FEM.cu:671-692); kept as is.
    if (!v0 && version == OptiGroup) return;
    XpbdArgs a{};
    a.ps = ps->get_parameters();
    a.fem = fem_view();
    a.mt = mt;
    a.eids = d_eids.get();
    a.dt = dt;
    for (int c = 0; c < colors.nb_color; ++c) {
        a.first = sched.color_begin[c];
        a.count = sched.color_begin[c + 1] - a.first;
        if (!launch_xpbd(S, Q, v0, a, eng->stream)) throw std::invalid_argument("unsupported element for XPBD");
    }
}

// ------------------------------------------------------- GPU_FEM_Explicit
GPU_FEM_Explicit::GPU_FEM_Explicit(Engine* e, Element el, const HostMesh& mesh, float young, float poisson,
                                   Material material, float damp)
    : GPU_FEM(e, el, mesh, young, poisson, material), _damping(damp), tile{64, 1, PASS_TASKS},
      n_vertices(mesh.nb_vertices()) {
    Coloration one;
    one.nb_color = 1;
    one.colors.assign(n_vertices, 0);
    const OwnersCSR o = build_owners(mesh, one);   // every vertex, ascending id
    upload_owners(eng, o, d_slot_vid, d_slot_nbo, d_own);
    if (o.own.size() != static_cast<size_t>(mesh.nb_elements()) * S)
        throw std::invalid_argument("explicit force pass: an element lists a vertex twice");
    std::vector<int> inc_at(o.own.size(), 0);
    for (size_t j = 0; j < o.own.size(); ++j)
        inc_at[static_cast<size_t>(o.own[j] >> 5) * S + (o.own[j] & 31)] = static_cast<int>(j);
    d_inc_at = DBuf<int>(&eng->pool, inc_at);
    d_partial = DBuf<float>(&eng->pool, o.own.size() * 9);
}

void GPU_FEM_Explicit::step(GPU_ParticleSystem* ps, float dt) {
    GatherArgs a{};
    a.ps = ps->get_parameters();
    a.fem = fem_view();
    a.own = OwnersView{d_slot_vid.get(), d_slot_nbo.get(), d_own.get()};
    a.mt = mt;
    a.damping = _damping;
    a.dt = dt;
    a.ju = nullptr;
    a.first_slot = 0;
    a.count = n_vertices;
    a.mode = GATHER_EXPLICIT_FORCE;
    if (tile.mode == PASS_TASKS) {
        ExplicitElemArgs ea{a.ps, a.fem, a.mt, d_inc_at.get(), d_partial.get()};
        if (!launch_explicit_elem(S, Q, ea, eng->stream)) throw std::invalid_argument("unsupported element");
        launch_explicit_vertex(a.ps, a.own, d_partial.get(), n_vertices, _damping, nullptr, dt, eng->stream);
        return;
    }
    if (!launch_gather(S, Q, tile.L, tile.QL, a, eng->stream)) throw std::invalid_argument("unsupported gather tile");
}

bool GPU_FEM_Explicit::step_integrate(GPU_ParticleSystem* ps, float dt, const float gravity[3]) {
    if (tile.mode != PASS_TASKS || ps->n != n_vertices) return false;
    const PSView v = ps->get_parameters();
    ExplicitElemArgs ea{v, fem_view(), mt, d_inc_at.get(), d_partial.get()};
    if (!launch_explicit_elem(S, Q, ea, eng->stream)) throw std::invalid_argument("unsupported element");
    launch_explicit_vertex(v, OwnersView{d_slot_vid.get(), d_slot_nbo.get(), d_own.get()}, d_partial.get(), n_vertices,
                           _damping, gravity, dt, eng->stream);
    return true;
}

// ------------------------------------------------------------ constraints
GPU_Fix_Constraint::GPU_Fix_Constraint(Engine* e, const std::vector<float>& xyz, const std::vector<int>& sel)
    : ids(sel), stream(e->stream) {
    float com[3] = {0.f, 0.f, 0.f};
    for (int i : ids) {
        if (i < 0 || static_cast<size_t>(i) >= xyz.size() / 3) throw std::invalid_argument("fixed id out of range");
        com[0] += xyz[3 * i]; com[1] += xyz[3 * i + 1]; com[2] += xyz[3 * i + 2];
    }
    const float cnt = static_cast<float>(ids.size());
    for (int d = 0; d < 3; ++d) {
        t.com[d] = ids.empty() ? 0.f : com[d] / cnt;
        t.offset[d] = 0.f;
    }
    const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    std::memcpy(t.R, I, sizeof(I));
    d_ids = DBuf<int>(&e->pool, ids);
    std::vector<unsigned char> flags(xyz.size() / 3, 0);
    for (int i : ids) flags[i] = 1;
    d_flags = DBuf<unsigned char>(&e->pool, flags);
}

void GPU_Fix_Constraint::step(GPU_ParticleSystem* ps, float) {
    launch_fix(ps->get_parameters(), d_ids.get(), static_cast<int>(ids.size()), t, stream);
}

GPU_Box_Limit::GPU_Box_Limit(hipStream_t s, const float pmin_[3], const float pmax_[3]) : stream(s) {
    for (int d = 0; d < 3; ++d) { pmin[d] = pmin_[d]; pmax[d] = pmax_[d]; }
}

void GPU_Box_Limit::step(GPU_ParticleSystem* ps, float) {
    launch_box(ps->get_parameters(), pmin, pmax, stream);
}

}  // namespace oe
