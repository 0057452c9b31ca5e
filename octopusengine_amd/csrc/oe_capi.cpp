// extern "C" boundary.  Every entry point catches exceptions and maps them to
// an oe_status; the message is kept per thread for oe_last_error().
#include "../../include/octopus_engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "oe_batch.hpp"
#include "oe_server.hpp"
#include "oe_scene.hpp"
#include "oe_scratch.hpp"
#include "oe_tune.hpp"

struct oe_engine {
    std::unique_ptr<oe::Engine> eng;
};
struct oe_scene {
    std::unique_ptr<oe::Scene> scene;
    oe_engine* owner;
};
struct oe_batch {
    std::unique_ptr<oe::Batch> batch;
};
struct oe_server {
    std::unique_ptr<oe::Server> server;
    oe_engine* owner;
};
struct oe_record {
    std::string path;
    oe::JsonRecord rec;
};
struct oe_vtk {
    std::unique_ptr<oe::VtkWriter> w;
};
struct oe_mesh {
    oe::LoadedMesh m;
};

namespace {

thread_local std::string g_last_error;

oe_status fail(oe_status code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

template <typename F>
oe_status guard(F&& fn) {
    try {
        fn();
        return OE_OK;
    } catch (const oe::BudgetError& e) {
        return fail(OE_ERR_BUDGET, e.what());
    } catch (const oe::HipError& e) {
        return fail(OE_ERR_HIP, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(OE_ERR_INVALID, e.what());
    } catch (const std::out_of_range& e) {
        return fail(OE_ERR_INVALID, e.what());
    } catch (const std::length_error& e) {
        return fail(OE_ERR_INVALID, e.what());
    } catch (const std::logic_error& e) {
        return fail(OE_ERR_STATE, e.what());
    } catch (const std::bad_alloc&) {
        return fail(OE_ERR_INTERNAL, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(OE_ERR_INTERNAL, e.what());
    } catch (...) {
        return fail(OE_ERR_INTERNAL, "unknown error");
    }
}

#define OE_REQUIRE(cond, msg) \
    do {                      \
        if (!(cond)) throw std::invalid_argument(msg); \
    } while (0)

oe::HostMesh mesh_from(int32_t element, const float* xyz, int32_t nv, const int32_t* topo, int32_t ne) {
    OE_REQUIRE(oe::element_supported(element), "unsupported element type");
    OE_REQUIRE(nv > 0 && ne > 0 && xyz && topo, "empty mesh");
    oe::HostMesh m;
    m.element = static_cast<oe::Element>(element);
    const int S = oe::elem_nb_vertices(m.element);
    m.xyz.assign(xyz, xyz + 3 * static_cast<size_t>(nv));
    m.topo.assign(topo, topo + static_cast<size_t>(S) * ne);
    for (int32_t v : m.topo) OE_REQUIRE(v >= 0 && v < nv, "topology index out of range");
    return m;
}

void copy3(std::vector<float>& src, float* dst) { std::memcpy(dst, src.data(), src.size() * sizeof(float)); }

// Reads of a served scene first wait for every request submitted for it.
void wait_served(oe::Scene& sc) {
    if (sc.server) sc.server->wait_slot(sc.server_slot);
}

// A server steps its scenes on its own streams: nothing else may step them or
// write their state.
void require_not_served(const oe::Scene& sc, const char* what) {
    if (sc.server) throw std::logic_error(std::string(what) + ": the scene is held by a server");
}

// Batches and servers copy a scene's constraints and task tables when they
// take the scene in.
void require_unshared(const oe::Scene& sc, const char* what) {
    require_not_served(sc, what);
    if (!sc.batches.empty()) throw std::logic_error(std::string(what) + ": the scene belongs to a batch");
}

}  // namespace

extern "C" {

void oe_scene_desc_default(oe_scene_desc* d) {
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->element = OE_TETRA10;
    d->material = OE_STABLE_NEOHOOKE;
    d->young = 1e6f;
    d->poisson = 0.45f;
    d->density = 1000.f;
    d->mass_distribution = OE_MASS_FEMSHAPE;
    d->solver = OE_SOLVER_VBD;
    d->iterations = 10;
    d->vbd_version = OE_VBD_REDUCTION_SYMMETRY;
    d->target_color = -1;
    d->xpbd_version = OE_XPBD_OPTISHARED;
    d->gravity[0] = 0.f;
    d->gravity[1] = -9.81f;
    d->gravity[2] = 0.f;
}

const char* oe_last_error(void) { return g_last_error.c_str(); }
const char* oe_version(void) { return "octopusengine_amd 0.1 (gfx950)"; }

oe_status oe_engine_create(int32_t device, uint64_t budget_bytes, oe_engine** out) {
    return guard([&] {
        OE_REQUIRE(out, "null output");
        auto* e = new oe_engine();
        try {
            e->eng = std::make_unique<oe::Engine>(device, budget_bytes ? budget_bytes : SIZE_MAX);
            oe::scratch_setup(device);
        } catch (...) {
            delete e;
            throw;
        }
        *out = e;
    });
}

oe_status oe_engine_destroy(oe_engine* e) {
    return guard([&] { delete e; });
}

oe_status oe_engine_synchronize(oe_engine* e) {
    return guard([&] {
        OE_REQUIRE(e, "null engine");
        OE_HIP(hipStreamSynchronize(e->eng->stream));
        OE_HIP(hipDeviceSynchronize());
    });
}

oe_status oe_engine_memory(oe_engine* e, uint64_t* used, uint64_t* budget) {
    return guard([&] {
        OE_REQUIRE(e, "null engine");
        if (used) *used = e->eng->pool.used.load();
        if (budget) *budget = e->eng->pool.budget;
    });
}

oe_status oe_scene_create(oe_engine* e, const oe_scene_desc* d, oe_scene** out) {
    return guard([&] {
        OE_REQUIRE(e && d && out, "null argument");
        oe::HostMesh m = mesh_from(d->element, d->xyz, d->n_vertices, d->topology, d->n_elements);
        oe::SceneConfig c;
        OE_REQUIRE(d->material == OE_NEOHOOKE || d->material == OE_STABLE_NEOHOOKE, "unsupported material");
        OE_REQUIRE(d->mass_distribution == OE_MASS_UNIFORM || d->mass_distribution == OE_MASS_FEMSHAPE,
                   "unknown mass distribution");
        OE_REQUIRE(d->solver >= OE_SOLVER_VBD && d->solver <= OE_SOLVER_EXPLICIT, "unknown solver");
        OE_REQUIRE(d->vbd_version >= OE_VBD_BASE && d->vbd_version <= OE_VBD_BETTER_COLORATION, "unknown VBD version");
        OE_REQUIRE(d->xpbd_version >= OE_XPBD_XPBD && d->xpbd_version <= OE_XPBD_OPTIGROUP, "unknown XPBD version");
        OE_REQUIRE(d->iterations >= 0, "iterations must be >= 0");
        c.solver = d->solver;
        c.material = static_cast<oe::Material>(d->material);
        c.young = d->young;
        c.poisson = d->poisson;
        c.density = d->density;
        c.distribution = static_cast<oe::Mass_Distribution>(d->mass_distribution);
        c.iterations = d->iterations;
        c.rho = d->rho;
        c.damping = d->damping;
        c.global_damping = d->global_damping;
        c.vbd_version = d->vbd_version;
        c.target_color = d->target_color;
        c.xpbd_version = d->xpbd_version;
        for (int k = 0; k < 3; ++k) c.gravity[k] = d->gravity[k];
        c.shuffle_seed = d->shuffle_seed;
        auto* s = new oe_scene();
        try {
            s->scene = std::make_unique<oe::Scene>(e->eng.get(), std::move(m), c);
        } catch (...) {
            delete s;
            throw;
        }
        s->owner = e;
        if (e->eng->autotune && (s->scene->vbd_fem || s->scene->explicit_fem)) {
            try {
                oe::autotune_scene(*s->scene, false);
            } catch (...) {
                delete s;
                throw;
            }
        }
        *out = s;
    });
}

oe_status oe_scene_destroy(oe_scene* s) {
    return guard([&] {
        if (!s) return;
        oe::Scene& sc = *s->scene;
        // a served scene leaves its server once its requests are done; its
        // batches refuse to step from now on
        if (sc.server) sc.server->remove_scene(sc.server_slot);
        const std::vector<oe::Batch*> batches = sc.batches;
        for (oe::Batch* b : batches) b->scene_destroyed(&sc);
        sc.batches.clear();
        delete s;
    });
}

oe_status oe_scene_fix_vertices(oe_scene* s, const int32_t* ids, int32_t n) {
    return guard([&] {
        OE_REQUIRE(s && (ids || n == 0) && n >= 0, "bad argument");
        require_unshared(*s->scene, "fix_vertices");
        s->scene->fix_vertices(std::vector<int>(ids, ids + n));
    });
}

oe_status oe_scene_fix_plane(oe_scene* s, const float o[3], const float n[3], int32_t* n_fixed) {
    return guard([&] {
        OE_REQUIRE(s && o && n, "bad argument");
        require_unshared(*s->scene, "fix_plane");
        const std::vector<int> ids = s->scene->select_plane(o, n);
        s->scene->fix_vertices(ids);
        if (n_fixed) *n_fixed = static_cast<int32_t>(ids.size());
    });
}

oe_status oe_scene_box_limit(oe_scene* s, const float pmin[3], const float pmax[3]) {
    return guard([&] {
        OE_REQUIRE(s && pmin && pmax, "bad argument");
        require_unshared(*s->scene, "box_limit");
        s->scene->box_limit(pmin, pmax);
    });
}

oe_status oe_scene_step(oe_scene* s, float dt, int32_t substeps) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        require_not_served(*s->scene, "step");
        s->scene->step(dt, substeps);
    });
}

oe_status oe_scene_set_graph(oe_scene* s, int32_t enable) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        s->scene->set_graph(enable != 0);
    });
}

namespace {
void check_tile(int element, int32_t lanes, int32_t qpl, int32_t mode) {
    const int Q = oe::elem_nb_quadrature(static_cast<oe::Element>(element));
    OE_REQUIRE(lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64, "lanes must be 4..64, power of 2");
    OE_REQUIRE(qpl >= 1 && Q % qpl == 0, "quadrature points per lane must divide Q");
    OE_REQUIRE(qpl == 1 || qpl == Q || (Q == 4 && qpl == 2) || (Q == 27 && (qpl == 3 || qpl == 9)),
               "tile not instantiated");
    OE_REQUIRE(mode == oe::PASS_GATHER || mode == oe::PASS_CACHE || mode == oe::PASS_TASKS,
               "mode must be 0 (gather), 1 (cached gradients) or 2 (element tasks)");
    OE_REQUIRE(mode != oe::PASS_CACHE || qpl == 1, "the gradient cache takes one quadrature point per lane");
}
}  // namespace

oe_status oe_scene_set_tile(oe_scene* s, int32_t lanes, int32_t qpl, int32_t mode) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        oe::Scene& sc = *s->scene;
        require_unshared(sc, "set_tile");
        check_tile(sc.mesh.element, lanes, qpl, mode);
        OE_REQUIRE(mode == oe::PASS_GATHER || sc.vbd_fem || (mode == oe::PASS_TASKS && sc.explicit_fem),
                   "cached gradients are a VBD option; element tasks a VBD or explicit one");
        OE_REQUIRE(mode != oe::PASS_TASKS || !sc.vbd_fem || sc.mesh.element != oe::Hexa27 || sc.vbd_fem->has_q2_basis,
                   "no tensor factors for this 27-node hexahedron");
        sc.set_graph(false);
        if (sc.vbd_fem) {
            const oe::GatherTile before = sc.vbd_fem->tile;
            sc.vbd_fem->tile = oe::GatherTile{lanes, qpl, mode};
            try {
                if (sc.vbd_fem->cached()) sc.vbd_fem->ensure_ju();
                if (sc.vbd_fem->tasks()) sc.vbd_fem->prepare_tasks();
            } catch (...) {
                sc.vbd_fem->tile = before;
                throw;
            }
        }
        if (sc.explicit_fem) sc.explicit_fem->tile = oe::GatherTile{lanes, qpl, mode};
    });
}

namespace {
oe::GatherTile tile_from(const oe_tile* t) {
    oe::GatherTile g{t->lanes, t->qpl, t->mode};
    g.inc = t->inc_per_block;
    g.lat_blocks = t->latency_blocks;
    return g;
}
void tile_to(const oe::GatherTile& g, oe_tile* t) {
    t->lanes = g.L;
    t->qpl = g.QL;
    t->mode = g.mode;
    t->inc_per_block = g.inc;
    t->latency_blocks = g.lat_blocks;
}
void check_tile_ex(int element, const oe_tile* t) {
    check_tile(element, t->lanes, t->qpl, t->mode);
    OE_REQUIRE(t->inc_per_block >= 0 && t->inc_per_block <= oe::max_task_inc(static_cast<oe::Element>(element)),
               "incidences per block must be 0 (the kernel's maximum) or at most the kernel's maximum");
    OE_REQUIRE(t->latency_blocks >= 0, "latency_blocks must be >= 0");
}
}  // namespace

oe_status oe_scene_set_tile_ex(oe_scene* s, const oe_tile* t) {
    return guard([&] {
        OE_REQUIRE(s && t, "bad argument");
        oe::Scene& sc = *s->scene;
        require_unshared(sc, "set_tile");
        check_tile_ex(sc.mesh.element, t);
        OE_REQUIRE(t->mode == oe::PASS_GATHER || sc.vbd_fem || (t->mode == oe::PASS_TASKS && sc.explicit_fem),
                   "cached gradients are a VBD option; element tasks a VBD or explicit one");
        OE_REQUIRE(t->mode != oe::PASS_TASKS || !sc.vbd_fem || sc.mesh.element != oe::Hexa27 ||
                       sc.vbd_fem->has_q2_basis,
                   "no tensor factors for this 27-node hexahedron");
        oe::apply_scene_tile(sc, tile_from(t));
    });
}

oe_status oe_scene_get_tile_ex(oe_scene* s, oe_tile* t) {
    return guard([&] {
        OE_REQUIRE(s && t, "bad argument");
        oe::GatherTile g{0, 0, oe::PASS_GATHER};
        if (s->scene->vbd_fem) g = s->scene->vbd_fem->tile;
        if (s->scene->explicit_fem) g = s->scene->explicit_fem->tile;
        tile_to(g, t);
    });
}

oe_status oe_scene_autotune(oe_scene* s, int32_t force, oe_tile* chosen, double* sweep_ms, int32_t* candidates) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        require_unshared(*s->scene, "autotune");
        const oe::TuneRecord r = oe::autotune_scene(*s->scene, force != 0);
        if (chosen) tile_to(r.tile, chosen);
        if (sweep_ms) *sweep_ms = r.ms;
        if (candidates) *candidates = r.candidates;
    });
}

oe_status oe_engine_set_autotune(oe_engine* e, int32_t enable) {
    return guard([&] {
        OE_REQUIRE(e, "null engine");
        e->eng->autotune = enable != 0;
    });
}

oe_status oe_autotune_cache(int32_t* n, int32_t* device, int32_t* element, int32_t* bucket, int32_t* form,
                            oe_tile* tiles, double* ms, int32_t* candidates) {
    return guard([&] {
        OE_REQUIRE(n && *n >= 0, "bad capacity");
        const std::vector<oe::TuneRecord> recs = oe::tune_cache_records();
        const int32_t cap = *n;
        *n = static_cast<int32_t>(recs.size());
        for (int32_t i = 0; i < cap && i < static_cast<int32_t>(recs.size()); ++i) {
            const oe::TuneRecord& r = recs[i];
            if (device) device[i] = r.device;
            if (element) element[i] = r.element;
            if (bucket) bucket[i] = r.bucket;
            if (form) form[i] = r.form;
            if (tiles) tile_to(r.tile, &tiles[i]);
            if (ms) ms[i] = r.ms;
            if (candidates) candidates[i] = r.candidates;
        }
    });
}

oe_status oe_autotune_clear(void) {
    return guard([&] { oe::tune_cache_clear(); });
}

oe_status oe_scene_get_tile(oe_scene* s, int32_t* lanes, int32_t* qpl, int32_t* mode) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        oe::GatherTile t{0, 0, oe::PASS_GATHER};
        if (s->scene->vbd_fem) t = s->scene->vbd_fem->tile;
        if (s->scene->explicit_fem) t = s->scene->explicit_fem->tile;
        if (lanes) *lanes = t.L;
        if (qpl) *qpl = t.QL;
        if (mode) {
            *mode = oe::PASS_GATHER;
            if (s->scene->vbd_fem && s->scene->vbd_fem->cached()) *mode = oe::PASS_CACHE;
            if (s->scene->vbd_fem && s->scene->vbd_fem->tasks()) *mode = oe::PASS_TASKS;
            if (s->scene->explicit_fem) *mode = t.mode;
        }
    });
}

oe_status oe_scene_get_layout(oe_scene* s, int32_t* affine) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        if (affine) *affine = s->scene->fem->affine ? 1 : 0;
    });
}

oe_status oe_scene_info(oe_scene* s, int32_t* nv, int32_t* ne, int32_t* S, int32_t* Q, int32_t* nc, uint64_t* bytes) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        oe::Scene& sc = *s->scene;
        if (nv) *nv = sc.mesh.nb_vertices();
        if (ne) *ne = sc.mesh.nb_elements();
        if (S) *S = sc.fem->S;
        if (Q) *Q = sc.fem->Q;
        if (nc) {
            *nc = 0;
            if (sc.vbd_fem) *nc = sc.vbd_fem->colors.nb_color;
            if (sc.pbd_fem) *nc = sc.pbd_fem->colors.nb_color;
        }
        if (bytes) *bytes = sc.device_bytes();
    });
}

#define OE_GETTER3(name, method)                             \
    oe_status name(oe_scene* s, float* xyz) {                \
        return guard([&] {                                   \
            OE_REQUIRE(s && xyz, "bad argument");            \
            wait_served(*s->scene);                          \
            std::vector<float> h;                            \
            s->scene->ps->method(h);                         \
            copy3(h, xyz);                                   \
        });                                                  \
    }

OE_GETTER3(oe_scene_get_positions, get_position)
OE_GETTER3(oe_scene_get_last_positions, get_prev_position)
OE_GETTER3(oe_scene_get_init_positions, get_init_position)
OE_GETTER3(oe_scene_get_velocities, get_velocity)
OE_GETTER3(oe_scene_get_forces, get_forces)

oe_status oe_scene_get_masses(oe_scene* s, float* m) {
    return guard([&] {
        OE_REQUIRE(s && m, "bad argument");
        wait_served(*s->scene);
        std::vector<float> h;
        s->scene->ps->get_mass(h);
        std::memcpy(m, h.data(), h.size() * sizeof(float));
    });
}

oe_status oe_scene_get_masks(oe_scene* s, int32_t* mask) {
    return guard([&] {
        OE_REQUIRE(s && mask, "bad argument");
        wait_served(*s->scene);
        std::vector<int> h;
        s->scene->ps->get_masks(h);
        std::memcpy(mask, h.data(), h.size() * sizeof(int32_t));
    });
}

oe_status oe_scene_get_colors(oe_scene* s, int32_t* colors) {
    return guard([&] {
        OE_REQUIRE(s && colors, "bad argument");
        const oe::Scene& sc = *s->scene;
        const std::vector<int32_t>* c = nullptr;
        if (sc.vbd_fem) c = &sc.vbd_fem->colors.colors;
        if (sc.pbd_fem) c = &sc.pbd_fem->colors.colors;
        if (!c) throw std::logic_error("scene has no colouring (explicit solver)");
        std::memcpy(colors, c->data(), c->size() * sizeof(int32_t));
    });
}

oe_status oe_scene_set_state(oe_scene* s, const float* p, const float* last_p, const float* v) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        require_not_served(*s->scene, "set_state");
        s->scene->ps->set_state(p, last_p, v);
    });
}

oe_status oe_scene_set_timing(oe_scene* s, int32_t enable) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        if (!s->scene->vbd_fem) throw std::logic_error("gather timing is available on VBD scenes");
        s->scene->vbd_fem->time_gather = enable != 0;
    });
}

oe_status oe_scene_get_timing(oe_scene* s, double* total_ms, int64_t* launches) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        if (!s->scene->vbd_fem) throw std::logic_error("gather timing is available on VBD scenes");
        oe::GPU_VBD_FEM* f = s->scene->vbd_fem;
        f->collect_gather_timing();
        if (total_ms) *total_ms = f->gather_ms;
        if (launches) *launches = f->gather_launches;
        f->gather_ms = 0.0;
        f->gather_launches = 0;
    });
}

// ------------------------------------------------------------------ batches
oe_status oe_batch_create(oe_engine* e, oe_scene* const* scenes, int32_t n, oe_batch** out) {
    return guard([&] {
        OE_REQUIRE(e && scenes && n > 0 && out, "bad argument");
        std::vector<oe::Scene*> v;
        for (int32_t i = 0; i < n; ++i) {
            OE_REQUIRE(scenes[i], "null scene");
            OE_REQUIRE(scenes[i]->owner == e, "scene belongs to another engine");
            v.push_back(scenes[i]->scene.get());
        }
        auto* b = new oe_batch();
        try {
            b->batch = std::make_unique<oe::Batch>(e->eng.get(), std::move(v));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

oe_status oe_batch_destroy(oe_batch* b) {
    return guard([&] { delete b; });
}

oe_status oe_batch_step(oe_batch* b, float dt, int32_t substeps) {
    return guard([&] {
        OE_REQUIRE(b, "null batch");
        b->batch->step(dt, substeps);
    });
}

oe_status oe_batch_set_graph(oe_batch* b, int32_t enable) {
    return guard([&] {
        OE_REQUIRE(b, "null batch");
        b->batch->set_graph(enable != 0);
    });
}

oe_status oe_batch_set_streams(oe_batch* b, int32_t n) {
    return guard([&] {
        OE_REQUIRE(b, "null batch");
        OE_REQUIRE(n == 1 || n == 2, "a batch steps on 1 or 2 streams");
        b->batch->set_streams(n);
    });
}

oe_status oe_batch_set_tile(oe_batch* b, int32_t element, int32_t lanes, int32_t qpl, int32_t mode) {
    return guard([&] {
        OE_REQUIRE(b && oe::element_supported(element), "bad argument");
        check_tile(element, lanes, qpl, mode);
        b->batch->set_tile(static_cast<oe::Element>(element), oe::GatherTile{lanes, qpl, mode});
    });
}

oe_status oe_batch_set_tile_ex(oe_batch* b, int32_t element, const oe_tile* t) {
    return guard([&] {
        OE_REQUIRE(b && t && oe::element_supported(element), "bad argument");
        check_tile_ex(element, t);
        b->batch->set_tile(static_cast<oe::Element>(element), tile_from(t));
    });
}

oe_status oe_batch_get_tile_ex(oe_batch* b, int32_t element, oe_tile* t) {
    return guard([&] {
        OE_REQUIRE(b && t && oe::element_supported(element), "bad argument");
        const auto it = b->batch->tiles.find(element);
        tile_to(it != b->batch->tiles.end() ? it->second : oe::default_tile(static_cast<oe::Element>(element)), t);
    });
}

oe_status oe_batch_autotune(oe_batch* b, int32_t force) {
    return guard([&] {
        OE_REQUIRE(b, "null batch");
        (void)oe::autotune_batch(*b->batch, force != 0);
    });
}

oe_status oe_batch_set_timing(oe_batch* b, int32_t enable) {
    return guard([&] {
        OE_REQUIRE(b, "null batch");
        b->batch->timer.enabled = enable != 0;
    });
}

oe_status oe_batch_get_timing(oe_batch* b, double* gather_ms, int64_t* launches) {
    return guard([&] {
        OE_REQUIRE(b, "null batch");
        oe::GatherTimer& t = b->batch->timer;
        t.collect(b->batch->eng->stream);
        if (gather_ms) *gather_ms = t.ms;
        if (launches) *launches = t.launches;
        t.ms = 0.0;
        t.launches = 0;
    });
}

oe_status oe_batch_get_timing_element(oe_batch* b, int32_t element, double* gather_ms, int64_t* launches) {
    return guard([&] {
        OE_REQUIRE(b, "null batch");
        oe::GatherTimer& t = b->batch->timer;
        t.collect(b->batch->eng->stream);
        if (gather_ms) *gather_ms = t.tag_ms[element];
        if (launches) *launches = t.tag_launches[element];
        t.tag_ms.erase(element);
        t.tag_launches.erase(element);
    });
}

oe_status oe_batch_info_element(oe_batch* b, int32_t element, int32_t* gather_launches, int64_t* vertex_updates,
                                double* gather_bytes) {
    return guard([&] {
        OE_REQUIRE(b, "null batch");
        const oe::Schedule& s = b->batch->sched;
        int32_t n = 0;
        for (const auto& gl : s.glaunches) n += gl.element == element ? 1 : 0;
        const auto vu = s.vertex_updates_by_element.find(element);
        const auto by = s.algorithmic_bytes_by_element.find(element);
        if (gather_launches) *gather_launches = n * s.max_iterations;
        if (vertex_updates) *vertex_updates = vu == s.vertex_updates_by_element.end() ? 0 : vu->second;
        if (gather_bytes) *gather_bytes = by == s.algorithmic_bytes_by_element.end() ? 0.0 : by->second;
    });
}

oe_status oe_batch_info(oe_batch* b, int32_t* gather_launches, int64_t* vertex_updates, double* gather_bytes) {
    return guard([&] {
        OE_REQUIRE(b, "null batch");
        const oe::Schedule& s = b->batch->sched;
        if (gather_launches) *gather_launches = static_cast<int32_t>(s.glaunches.size()) * s.max_iterations;
        if (vertex_updates) *vertex_updates = s.vertex_updates_per_substep;
        if (gather_bytes) *gather_bytes = s.algorithmic_bytes_per_substep;
    });
}

// ------------------------------------------------------------------ server
oe_status oe_server_create(oe_engine* e, int32_t max_batch, int32_t streams, oe_server** out) {
    return guard([&] {
        OE_REQUIRE(e && out, "bad argument");
        auto* s = new oe_server();
        try {
            s->server = std::make_unique<oe::Server>(e->eng.get(), max_batch, streams);
        } catch (...) {
            delete s;
            throw;
        }
        s->owner = e;
        *out = s;
    });
}

oe_status oe_server_destroy(oe_server* sv) {
    return guard([&] { delete sv; });
}

oe_status oe_server_add_scene(oe_server* sv, oe_scene* s, int32_t* slot) {
    return guard([&] {
        OE_REQUIRE(sv && s, "bad argument");
        OE_REQUIRE(s->owner == sv->owner, "scene belongs to another engine");
        const int id = sv->server->add_scene(s->scene.get());
        if (slot) *slot = id;
    });
}

oe_status oe_server_set_tile_ex(oe_server* sv, int32_t element, const oe_tile* t) {
    return guard([&] {
        OE_REQUIRE(sv && t && oe::element_supported(element), "bad argument");
        check_tile_ex(element, t);
        sv->server->set_tile(static_cast<oe::Element>(element), tile_from(t));
    });
}

oe_status oe_server_get_tile_ex(oe_server* sv, int32_t element, oe_tile* t) {
    return guard([&] {
        OE_REQUIRE(sv && t && oe::element_supported(element), "bad argument");
        tile_to(sv->server->tile(static_cast<oe::Element>(element)), t);
    });
}

oe_status oe_server_submit(oe_server* sv, int32_t slot, float dt, int32_t substeps, uint64_t* ticket) {
    return guard([&] {
        OE_REQUIRE(sv, "null server");
        const uint64_t t = sv->server->submit(slot, dt, substeps);
        if (ticket) *ticket = t;
    });
}

oe_status oe_server_drain(oe_server* sv) {
    return guard([&] {
        OE_REQUIRE(sv, "null server");
        sv->server->drain();
    });
}

oe_status oe_server_wait(oe_server* sv, uint64_t ticket) {
    return guard([&] {
        OE_REQUIRE(sv, "null server");
        sv->server->wait(ticket);
    });
}

oe_status oe_server_poll(oe_server* sv, uint64_t ticket, int32_t* done) {
    return guard([&] {
        OE_REQUIRE(sv && done, "bad argument");
        *done = sv->server->poll(ticket) ? 1 : 0;
    });
}

oe_status oe_server_stats(oe_server* sv, int64_t* completed, double* p50, double* p99, double* mx, int64_t* vu,
                          int64_t* batches, double* mean_batch) {
    return guard([&] {
        OE_REQUIRE(sv, "null server");
        const oe::Server::Stats s = sv->server->stats();
        if (completed) *completed = s.completed;
        if (p50) *p50 = s.p50_ms;
        if (p99) *p99 = s.p99_ms;
        if (mx) *mx = s.max_ms;
        if (vu) *vu = s.vertex_updates;
        if (batches) *batches = s.batches;
        if (mean_batch) *mean_batch = s.mean_batch;
    });
}

oe_status oe_server_latencies(oe_server* sv, double* out, int64_t cap, int64_t* n) {
    return guard([&] {
        OE_REQUIRE(sv && cap >= 0, "bad argument");
        const std::vector<double> v = sv->server->latencies();
        const int64_t k = std::min<int64_t>(cap, static_cast<int64_t>(v.size()));
        if (out && k > 0) std::memcpy(out, v.data(), static_cast<size_t>(k) * sizeof(double));
        if (n) *n = static_cast<int64_t>(v.size());
    });
}

oe_status oe_server_latency_slots(oe_server* sv, int32_t* out, int64_t cap, int64_t* n) {
    return guard([&] {
        OE_REQUIRE(sv && cap >= 0, "bad argument");
        const std::vector<int> v = sv->server->latency_slots();
        const int64_t k = std::min<int64_t>(cap, static_cast<int64_t>(v.size()));
        if (out && k > 0) std::memcpy(out, v.data(), static_cast<size_t>(k) * sizeof(int32_t));
        if (n) *n = static_cast<int64_t>(v.size());
    });
}

// ------------------------------------------------------------------ host side
oe_status oe_host_beam_mesh(int32_t element, int32_t nx, int32_t ny, int32_t nz, float sx, float sy, float sz,
                            int32_t* n_vertices, int32_t* n_elements, float* xyz, int32_t* topology) {
    return guard([&] {
        OE_REQUIRE(oe::element_supported(element), "unsupported element type");
        const oe::HostMesh m = oe::beam_mesh(static_cast<oe::Element>(element), nx, ny, nz, sx, sy, sz);
        if (n_vertices) *n_vertices = m.nb_vertices();
        if (n_elements) *n_elements = m.nb_elements();
        if (xyz) std::memcpy(xyz, m.xyz.data(), m.xyz.size() * sizeof(float));
        if (topology) std::memcpy(topology, m.topo.data(), m.topo.size() * sizeof(int32_t));
    });
}

oe_status oe_host_shape(int32_t element, int32_t* S, int32_t* Q, double* dN, double* weights, double* qpoints) {
    return guard([&] {
        OE_REQUIRE(oe::element_supported(element), "unsupported element type");
        const oe::ShapeTable& t = oe::shape_table(static_cast<oe::Element>(element));
        if (S) *S = t.S;
        if (Q) *Q = t.Q;
        if (dN) std::memcpy(dN, t.dN.data(), t.dN.size() * sizeof(double));
        if (weights) std::memcpy(weights, t.weights.data(), t.weights.size() * sizeof(double));
        if (qpoints) std::memcpy(qpoints, t.qpoints.data(), t.qpoints.size() * sizeof(double));
    });
}

oe_status oe_host_color_primal(int32_t element, const int32_t* topology, int32_t ne, int32_t nv, int32_t target,
                               int32_t* colors, int32_t* n_colors, int64_t* conflicts) {
    return guard([&] {
        OE_REQUIRE(oe::element_supported(element) && topology && ne > 0, "bad argument");
        const oe::Element el = static_cast<oe::Element>(element);
        const int S = oe::elem_nb_vertices(el);
        std::vector<int32_t> topo(topology, topology + static_cast<size_t>(S) * ne);
        const oe::MultiGraph g = oe::build_primal_graph(el, topo, nv);
        oe::Coloration c = oe::color_dsat(g);
        if (target > 0) oe::color_reduction(g, c, target);
        if (colors) std::memcpy(colors, c.colors.data(), c.colors.size() * sizeof(int32_t));
        if (n_colors) *n_colors = c.nb_color;
        if (conflicts) *conflicts = oe::count_conflicts(g, c);
    });
}

oe_status oe_host_color_dual(int32_t element, const int32_t* topology, int32_t ne, int32_t nv, int32_t* colors,
                             int32_t* n_colors, int64_t* conflicts) {
    return guard([&] {
        OE_REQUIRE(oe::element_supported(element) && topology && ne > 0, "bad argument");
        const oe::Element el = static_cast<oe::Element>(element);
        const int S = oe::elem_nb_vertices(el);
        std::vector<int32_t> topo(topology, topology + static_cast<size_t>(S) * ne);
        const oe::MultiGraph g = oe::build_dual_vertex_graph(el, topo, nv);
        const oe::Coloration c = oe::color_greedy_slf(g);
        if (colors) std::memcpy(colors, c.colors.data(), c.colors.size() * sizeof(int32_t));
        if (n_colors) *n_colors = c.nb_color;
        if (conflicts) *conflicts = oe::count_conflicts(g, c);
    });
}

oe_status oe_host_fem_constants(int32_t element, const float* xyz, int32_t nv, const int32_t* topology, int32_t ne,
                                float* jxv) {
    return guard([&] {
        const oe::HostMesh m = mesh_from(element, xyz, nv, topology, ne);
        const oe::FemConstants fc = oe::build_fem_constants(m);
        if (jxv) std::memcpy(jxv, fc.jxv.data(), fc.jxv.size() * sizeof(float));
    });
}

oe_status oe_host_affine_records(int32_t element, const float* xyz, int32_t nv, const int32_t* topology, int32_t ne,
                                 float* records, float* weights, int32_t* affine) {
    return guard([&] {
        OE_REQUIRE(affine, "bad argument");
        const oe::HostMesh m = mesh_from(element, xyz, nv, topology, ne);
        std::vector<float> rec, wq;
        *affine = oe::build_affine_records(m, rec, wq) ? 1 : 0;
        if (*affine && records) std::memcpy(records, rec.data(), rec.size() * sizeof(float));
        if (*affine && weights) std::memcpy(weights, wq.data(), wq.size() * sizeof(float));
    });
}

oe_status oe_host_masses(int32_t element, const float* xyz, int32_t nv, const int32_t* topology, int32_t ne,
                         float density, int32_t distribution, float* masses) {
    return guard([&] {
        const oe::HostMesh m = mesh_from(element, xyz, nv, topology, ne);
        const oe::FemConstants fc = oe::build_fem_constants(m);
        const std::vector<float> ms =
            oe::compute_fem_mass(m, fc, density, static_cast<oe::Mass_Distribution>(distribution));
        if (masses) std::memcpy(masses, ms.data(), ms.size() * sizeof(float));
    });
}

// ------------------------------------------------------------------ diagnostics
oe_status oe_scene_element_diagnostics(oe_scene* s, float* stress, float* volume, float* volume_diff,
                                       float* inverted) {
    return guard([&] {
        OE_REQUIRE(s, "null scene");
        wait_served(*s->scene);
        s->scene->fem->diagnostics(s->scene->ps.get(), stress, volume, volume_diff, inverted);
    });
}

oe_status oe_scene_residual(oe_scene* s, float dt, float* r) {
    return guard([&] {
        OE_REQUIRE(s && r, "bad argument");
        OE_REQUIRE(s->scene->vbd_fem, "the residual is defined for VBD scenes");
        wait_served(*s->scene);
        std::vector<float> h;
        s->scene->vbd_fem->residual(s->scene->ps.get(), dt, h);
        std::memcpy(r, h.data(), h.size() * sizeof(float));
    });
}

oe_status oe_scene_residual_scratch(oe_scene* s, float dt, float* r) {
    return guard([&] {
        OE_REQUIRE(s && r, "bad argument");
        OE_REQUIRE(s->scene->vbd_fem, "the residual is defined for VBD scenes");
        wait_served(*s->scene);
        oe::residual_scratch(*s->scene, dt, r);
    });
}

// ------------------------------------------------------------------ mesh files
oe_status oe_mesh_load(const char* path, oe_mesh** out) {
    return guard([&] {
        OE_REQUIRE(path && out, "bad argument");
        auto m = std::make_unique<oe_mesh>();
        m->m = oe::load_mesh(path);
        *out = m.release();
    });
}

oe_status oe_mesh_info(oe_mesh* m, int32_t* n_vertices, int32_t* counts) {
    return guard([&] {
        OE_REQUIRE(m, "bad argument");
        if (n_vertices) *n_vertices = static_cast<int32_t>(m->m.xyz.size() / 3);
        if (counts) {
            for (int e = 0; e < 10; ++e) counts[e] = 0;
            for (const auto& kv : m->m.topologies)
                counts[kv.first] = static_cast<int32_t>(kv.second.size() / oe::elem_nb_vertices(kv.first));
        }
    });
}

oe_status oe_mesh_vertices(oe_mesh* m, float* xyz) {
    return guard([&] {
        OE_REQUIRE(m && xyz, "bad argument");
        std::memcpy(xyz, m->m.xyz.data(), m->m.xyz.size() * sizeof(float));
    });
}

oe_status oe_mesh_topology(oe_mesh* m, int32_t element, int32_t* topology) {
    return guard([&] {
        OE_REQUIRE(m && topology && oe::element_supported(element), "bad argument");
        const auto it = m->m.topologies.find(static_cast<oe::Element>(element));
        if (it != m->m.topologies.end())
            std::memcpy(topology, it->second.data(), it->second.size() * sizeof(int32_t));
    });
}

oe_status oe_mesh_destroy(oe_mesh* m) {
    return guard([&] { delete m; });
}

// ------------------------------------------------------------------ VTK I/O
oe_status oe_vtk_open(const char* path, oe_vtk** out) {
    return guard([&] {
        OE_REQUIRE(path && out, "bad argument");
        auto* v = new oe_vtk();
        try {
            v->w = std::make_unique<oe::VtkWriter>(path);
        } catch (...) {
            delete v;
            throw;
        }
        *out = v;
    });
}

oe_status oe_vtk_mesh(oe_vtk* v, int32_t element, const float* xyz, int32_t nv, const int32_t* topology, int32_t ne) {
    return guard([&] {
        OE_REQUIRE(v && v->w && oe::element_supported(element), "bad argument");
        OE_REQUIRE((xyz || nv == 0) && (topology || ne == 0), "bad argument");
        v->w->mesh(static_cast<oe::Element>(element), xyz, nv, topology, ne);
    });
}

oe_status oe_vtk_point_data(oe_vtk* v) {
    return guard([&] {
        OE_REQUIRE(v && v->w, "bad argument");
        v->w->point_data();
    });
}

oe_status oe_vtk_cell_data(oe_vtk* v) {
    return guard([&] {
        OE_REQUIRE(v && v->w, "bad argument");
        v->w->cell_data();
    });
}

oe_status oe_vtk_scalars(oe_vtk* v, const char* name, const float* data, int32_t n) {
    return guard([&] {
        OE_REQUIRE(v && v->w && name && (data || n == 0), "bad argument");
        v->w->scalars(name, data, n);
    });
}

oe_status oe_vtk_vectors(oe_vtk* v, const char* name, const float* data, int32_t n) {
    return guard([&] {
        OE_REQUIRE(v && v->w && name && (data || n == 0), "bad argument");
        v->w->vectors(name, data, n);
    });
}

oe_status oe_vtk_close(oe_vtk* v) {
    return guard([&] {
        OE_REQUIRE(v, "bad argument");
        std::unique_ptr<oe_vtk> own(v);
        if (own->w) own->w->close();
    });
}

oe_status oe_scene_save_vtk(oe_scene* s, const char* path) {
    return guard([&] {
        OE_REQUIRE(s && path, "bad argument");
        oe::Scene& sc = *s->scene;
        wait_served(sc);
        std::vector<float> rest, pos, m;
        sc.ps->get_init_position(rest);
        sc.ps->get_position(pos);
        sc.ps->get_mass(m);
        std::vector<float> u(pos.size());
        for (size_t i = 0; i < u.size(); ++i) u[i] = pos[i] - rest[i];
        const int nv = sc.mesh.nb_vertices();
        const int ne = sc.mesh.nb_elements();
        const int S = oe::elem_nb_vertices(sc.mesh.element);
        std::vector<float> u_norm(nv), stress(ne), volume(ne), volume_diff(ne), inverted(ne);
        for (int i = 0; i < nv; ++i)
            u_norm[i] = std::sqrt(u[3 * i] * u[3 * i] + u[3 * i + 1] * u[3 * i + 1] + u[3 * i + 2] * u[3 * i + 2]);
        sc.fem->diagnostics(sc.ps.get(), stress.data(), volume.data(), volume_diff.data(), inverted.data());
        // Cuda_FEM_Dynamic::get_stress_vertices (include/Script/Dynamic/Cuda_FEM_Dynamic.h:96-117)
        std::vector<float> smooth(nv, 0.f);
        for (int e = 0; e < ne; ++e)
            for (int j = 0; j < S; ++j) smooth[sc.mesh.topo[static_cast<size_t>(e) * S + j]] += stress[e];
        for (int i = 0; i < nv; ++i) smooth[i] /= m[i];
        oe::VtkWriter w(path);
        w.mesh(sc.mesh.element, rest.data(), nv, sc.mesh.topo.data(), ne);
        w.point_data();
        w.scalars("weights", m.data(), nv);
        w.scalars("u_norm", u_norm.data(), nv);
        w.vectors("u", u.data(), nv);
        w.scalars("smooth_stress", smooth.data(), nv);
        w.cell_data();
        w.scalars("stress", stress.data(), ne);
        w.scalars("volume", volume.data(), ne);
        w.scalars("volume_diff", volume_diff.data(), ne);
        w.scalars("inverted", inverted.data(), ne);
        w.close();
    });
}

// ------------------------------------------------------------------ JSON record
oe_status oe_record_open(const char* path, oe_record** out) {
    return guard([&] {
        OE_REQUIRE(path && out, "bad argument");
        auto r = std::make_unique<oe_record>();
        r->path = path;
        *out = r.release();
    });
}

oe_status oe_record_time(oe_record* r, int32_t frame, float time, float dt, float fixed_time, float fixed_dt) {
    return guard([&] {
        OE_REQUIRE(r, "bad argument");
        r->rec.time(frame, time, dt, fixed_time, fixed_dt);
    });
}

oe_status oe_record_mesh(oe_record* r, const int32_t* counts, int32_t n_vertices) {
    return guard([&] {
        OE_REQUIRE(r && counts && n_vertices >= 0, "bad argument");
        r->rec.mesh(counts, n_vertices);
    });
}

oe_status oe_record_material(oe_record* r, int32_t xpbd, int32_t material, float poisson, float young,
                             int32_t iterations, int32_t sub_iterations, float density) {
    return guard([&] {
        OE_REQUIRE(r, "bad argument");
        OE_REQUIRE(material >= 0 && material <= 3, "unknown material");
        r->rec.material(xpbd != 0, material, poisson, young, iterations, sub_iterations, density);
    });
}

oe_status oe_record_string(oe_record* r, const char* name, const char* value) {
    return guard([&] {
        OE_REQUIRE(r && name && value, "bad argument");
        r->rec.string(name, value);
    });
}

oe_status oe_record_flexion(oe_record* r, float error, const float* p, const float* target) {
    return guard([&] {
        OE_REQUIRE(r && p && target, "bad argument");
        r->rec.flexion(error, p, target);
    });
}

oe_status oe_record_rotation(oe_record* r, float rot_max, float beam_length, const float* dist, const float* angles,
                             int32_t n) {
    return guard([&] {
        OE_REQUIRE(r && n >= 0 && (n == 0 || (dist && angles)), "bad argument");
        r->rec.rotation(rot_max, beam_length, dist, angles, n);
    });
}

oe_status oe_record_text(oe_record* r, char* buf, int64_t cap, int64_t* len) {
    return guard([&] {
        OE_REQUIRE(r && len, "bad argument");
        const std::string t = r->rec.text();
        *len = static_cast<int64_t>(t.size());
        if (buf) {
            OE_REQUIRE(cap > static_cast<int64_t>(t.size()), "buffer too small");
            std::memcpy(buf, t.c_str(), t.size() + 1);
        }
    });
}

oe_status oe_record_close(oe_record* r) {
    return guard([&] {
        OE_REQUIRE(r, "bad argument");
        std::unique_ptr<oe_record> own(r);
        own->rec.save(own->path);
    });
}

oe_status oe_flexion_error(const float* rest, const float* pos, int32_t n_vertices, const float* follow,
                           const float* target, float* error, float* p, int32_t* vertex) {
    return guard([&] {
        OE_REQUIRE(rest && pos && follow && target && error && p && vertex && n_vertices >= 0, "bad argument");
        *vertex = oe::flexion_error(rest, pos, n_vertices, follow, target, *error, p);
    });
}

oe_status oe_rotation_error(int32_t element, const int32_t* topology, int32_t n_elements, const float* rest,
                            const float* pos, int32_t n_vertices, int32_t samples_per_edge, float beam_length,
                            float* dist, float* angles, int32_t* n) {
    return guard([&] {
        OE_REQUIRE(topology && rest && pos && dist && angles && n && n_elements >= 0, "bad argument");
        const oe::Element el = static_cast<oe::Element>(element);
        OE_REQUIRE(el == oe::Tetra || el == oe::Hexa || el == oe::Tetra10 || el == oe::Tetra20 || el == oe::Hexa27,
                   "unsupported element");
        const int S = oe::elem_nb_vertices(el);
        for (int64_t i = 0; i < static_cast<int64_t>(n_elements) * S; ++i)
            OE_REQUIRE(topology[i] >= 0 && topology[i] < n_vertices, "topology index out of range");
        std::vector<float> d, a;
        oe::rotation_error(el, topology, n_elements, rest, pos, samples_per_edge, beam_length, d, a);
        OE_REQUIRE(static_cast<int64_t>(d.size()) <= *n, "output capacity too small");
        std::copy(d.begin(), d.end(), dist);
        std::copy(a.begin(), a.end(), angles);
        *n = static_cast<int32_t>(d.size());
    });
}

}  // extern "C"
