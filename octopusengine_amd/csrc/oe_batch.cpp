#include "oe_batch.hpp"

#include "oe_q2.hpp"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>

namespace oe {

namespace {
constexpr int kBlock = 256;
GPU_Box_Limit* find_box(const Scene& s) {
    GPU_Box_Limit* box = nullptr;
    for (const auto& c : s.integrator->_constraints) {
        if (auto* b = dynamic_cast<GPU_Box_Limit*>(c.get())) {
            if (box) throw std::invalid_argument("batched scenes support one box limit");
            box = b;
        }
    }
    return box;
}
}  // namespace

std::map<int, GatherTile> default_tiles(bool served) {
    std::map<int, GatherTile> t;
    for (Element e : {Tetra, Hexa, Tetra10, Tetra20, Hexa27}) t[e] = default_tile(e);
    // a batch's launches hold many blocks: the longer pencil blocks
    if (!served) t[Hexa27].inc = kQ2PencilIncPerBlock;
    return t;
}

void check_batchable(const Scene& s) {
    if (!s.vbd_fem) throw std::invalid_argument("batched stepping needs VBD scenes");
    if (s.vbd_fem->version == Better_Coloration)
        throw std::invalid_argument("batched stepping needs a proper colouring (not Better_Coloration)");
    if (s.cfg.rho > 0.f) throw std::invalid_argument("batched stepping does not run the Chebyshev acceleration");
    if (s.vbd_fem->shuffle) throw std::invalid_argument("batched stepping uses the ascending colour order");
    for (const auto& c : s.integrator->_constraints)
        if (!dynamic_cast<GPU_Fix_Constraint*>(c.get()) && !dynamic_cast<GPU_Box_Limit*>(c.get()))
            throw std::invalid_argument("unsupported constraint in a batched scene");
    (void)find_box(s);
}

std::vector<unsigned char> fixed_flags(const Scene& s) {
    std::vector<unsigned char> f(s.ps->n, 0);
    if (s.fix)
        for (int id : s.fix->ids) f[id] = 1;
    return f;
}

namespace {
GatherTile tile_of(const std::map<int, GatherTile>& tiles, Element el) {
    const auto it = tiles.find(el);
    return it != tiles.end() ? it->second : default_tile(el);
}
}  // namespace

SceneDev make_scene_dev(Scene& s, const unsigned char* fixed, const std::map<int, GatherTile>& tiles) {
    check_batchable(s);
    SceneDev d{};
    const GatherTile t = tile_of(tiles, s.mesh.element);
    if (t.mode == PASS_TASKS && s.mesh.element == Hexa27 && !s.vbd_fem->has_q2_basis)
        throw std::invalid_argument("no tensor factors for this 27-node hexahedron");
    d.ju = t.mode == PASS_CACHE ? s.vbd_fem->ensure_ju() : nullptr;
    d.fblocks = d.fblocks_lat = nullptr;
    if (t.mode == PASS_TASKS) {
        d.fblocks = s.vbd_fem->throughput_tables(t.inc).blocks.get();
        const TaskTables* lt = s.vbd_fem->latency_tables();
        d.fblocks_lat = lt ? lt->blocks.get() : nullptr;
    }
    d.ps = s.ps->get_parameters();
    d.fem = s.fem->fem_view();
    d.own = OwnersView{s.vbd_fem->d_slot_vid.get(), s.vbd_fem->d_slot_nbo.get(), s.vbd_fem->d_own.get()};
    d.mt = s.fem->mt;
    auto* vbd = dynamic_cast<GPU_VBD*>(s.integrator.get());
    if (!vbd) throw std::logic_error("VBD scene without a GPU_VBD integrator");
    d.y = vbd->y.get();
    d.damping = s.vbd_fem->damping;
    d.iterations = vbd->iteration;
    for (int k = 0; k < 3; ++k) d.g[k] = vbd->gravity[k];
    d.fixed = s.fix ? fixed : nullptr;
    if (s.fix) {
        std::memcpy(d.fix_com, s.fix->t.com, sizeof(d.fix_com));
        std::memcpy(d.fix_offset, s.fix->t.offset, sizeof(d.fix_offset));
        std::memcpy(d.fix_R, s.fix->t.R, sizeof(d.fix_R));
    }
    if (GPU_Box_Limit* box = find_box(s)) {
        d.has_box = 1;
        for (int k = 0; k < 3; ++k) {
            d.box_min[k] = box->pmin[k];
            d.box_max[k] = box->pmax[k];
        }
    }
    return d;
}

void Schedule::clear() {
    segs.clear();
    glaunches.clear();
    vsegs.clear();
    vblocks = 0;
    jlaunches.clear();
    jsegs.clear();
    max_iterations = 0;
    vertex_updates_per_substep = 0;
    algorithmic_bytes_per_substep = 0.0;
    algorithmic_bytes_by_element.clear();
    vertex_updates_by_element.clear();
}

void Schedule::build(const std::vector<Scene*>& scenes, const std::vector<int>& table_index,
                     const std::map<int, GatherTile>& tiles, bool want_latency) {
    clear();
    int max_colors = 0;
    for (Scene* s : scenes) {
        max_colors = std::max(max_colors, s->vbd_fem->nb_colors());
        max_iterations = std::max(max_iterations, s->cfg.iterations);
        vertex_updates_per_substep += static_cast<int64_t>(s->ps->n) * s->cfg.iterations;
        const int S = s->fem->S, Q = s->fem->Q;
        const double bytes =
            (static_cast<double>(s->fem->E) * S * (40.0 * Q + 16.0 * S + 8.0) + 56.0 * s->ps->n) * s->cfg.iterations;
        algorithmic_bytes_per_substep += bytes;
        algorithmic_bytes_by_element[s->mesh.element] += bytes;
        vertex_updates_by_element[s->mesh.element] += static_cast<int64_t>(s->ps->n) * s->cfg.iterations;
    }
    static const Element order[] = {Tetra, Tetra10, Tetra20, Hexa, Hexa27};
    for (int c = 0; c < max_colors; ++c) {
        for (Element el : order) {
            GLaunch gl{};
            gl.color = c;
            gl.element = el;
            gl.seg_first = static_cast<int>(segs.size());
            // Task passes take the latency form when it was asked for, every
            // scene of the launch has its vertex blocks, and the throughput
            // form would give fewer than 2 blocks per CU.
            const GatherTile et = tile_of(tiles, el);
            if (want_latency && et.mode == PASS_TASKS) {
                bool all = true;
                int64_t through_blocks = 0;
                for (Scene* s : scenes) {
                    if (s->mesh.element != el || c >= s->vbd_fem->nb_colors()) continue;
                    if (!s->vbd_fem->latency_tables()) all = false;
                    through_blocks += s->vbd_fem->throughput_tables(et.inc).color_blocks(c);
                }
                gl.latency = all && through_blocks < et.lat_blocks;
            }
            bool any = false;
            for (size_t i = 0; i < scenes.size(); ++i) {
                Scene* s = scenes[i];
                if (s->mesh.element != el || c >= s->vbd_fem->nb_colors()) continue;
                const GatherTile t = tile_of(tiles, el);
                const int first = s->vbd_fem->owners.color_begin[c];
                const int count = s->vbd_fem->owners.color_begin[c + 1] - first;
                if (count <= 0) continue;
                if (!any) {
                    gl.S = s->fem->S;
                    gl.Q = s->fem->Q;
                    gl.L = t.L;
                    gl.QL = t.QL;
                    gl.mode = t.mode;
                    gl.inc = t.inc;
                    any = true;
                    if (gl.mode == PASS_TASKS) q2_basis = s->vbd_fem->q2_basis;
                }
                if (gl.mode == PASS_TASKS) {
                    // fused tasks: one block per vertex block of the colour
                    const TaskTables* tt =
                        gl.latency ? s->vbd_fem->latency_tables() : &s->vbd_fem->throughput_tables(t.inc);
                    if (!tt) throw std::logic_error("vertex blocks not prepared");
                    const int nfb = tt->color_blocks(c);
                    segs.push_back(BatchSeg{table_index[i], tt->color_begin[c], nfb, gl.nblocks});
                    gl.nblocks += nfb;
                } else {
                    // gather: L lanes per slot
                    const int nb = static_cast<int>((static_cast<int64_t>(count) * gl.L + kBlock - 1) / kBlock);
                    segs.push_back(BatchSeg{table_index[i], first, count, gl.nblocks});
                    gl.nblocks += nb;
                }
                gl.slots += count;
                gl.incidences += s->vbd_fem->owners.color_incidences[c];
            }
            gl.seg_count = static_cast<int>(segs.size()) - gl.seg_first;
            if (any) glaunches.push_back(gl);
        }
    }
    for (size_t i = 0; i < scenes.size(); ++i) {
        const int n = scenes[i]->ps->n;
        vsegs.push_back(VSeg{table_index[i], n, vblocks});
        vblocks += (n + kBlock - 1) / kBlock;
    }
    for (Element el : order) {
        if (tile_of(tiles, el).mode != PASS_CACHE) continue;
        JLaunch jl{};
        jl.seg_first = static_cast<int>(jsegs.size());
        bool any = false;
        for (size_t i = 0; i < scenes.size(); ++i) {
            Scene* s = scenes[i];
            if (s->mesh.element != el) continue;
            const int E = s->fem->E;
            if (E <= 0) continue;
            if (static_cast<int64_t>(E) * s->fem->Q > INT32_MAX)
                throw std::invalid_argument("scene too large for the gradient cache");
            jl.S = s->fem->S;
            jl.Q = s->fem->Q;
            any = true;
            const int epb = refresh_elements_per_block(jl.S, jl.Q);
            if (epb <= 0) throw std::invalid_argument("unsupported element");
            jsegs.push_back(VSeg{table_index[i], E, jl.nblocks});
            jl.nblocks += (E + epb - 1) / epb;
        }
        jl.seg_count = static_cast<int>(jsegs.size()) - jl.seg_first;
        if (any) jlaunches.push_back(jl);
    }
}

namespace {
template <typename T>
void upload_grow(Engine* eng, DBuf<T>& buf, const std::vector<T>& h, hipStream_t s) {
    if (buf.size() < h.size() || !buf.get()) buf.allocate(&eng->pool, std::max<size_t>(h.size() + h.size() / 2, 16));
    if (!h.empty()) OE_HIP(hipMemcpyAsync(buf.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
}
}  // namespace

void DeviceSchedule::upload(Engine* eng, const Schedule& s, hipStream_t stream, bool sync) {
    upload_grow(eng, segs, s.segs, stream);
    upload_grow(eng, vsegs, s.vsegs, stream);
    upload_grow(eng, jsegs, s.jsegs, stream);
    if (sync) OE_HIP(hipStreamSynchronize(stream));
}

void GatherTimer::before(hipStream_t s, int tag) {
    if (!enabled) return;
    while (pool.size() < used + 2) {
        hipEvent_t ev;
        OE_HIP(hipEventCreate(&ev));
        pool.push_back(ev);
    }
    tags.resize(used / 2 + 1);
    tags[used / 2] = tag;
    OE_HIP(hipEventRecord(pool[used], s));
}

void GatherTimer::after(hipStream_t s) {
    if (!enabled) return;
    OE_HIP(hipEventRecord(pool[used + 1], s));
    used += 2;
}

void GatherTimer::collect(hipStream_t s) {
    if (used == 0) return;
    OE_HIP(hipStreamSynchronize(s));
    for (size_t i = 0; i + 1 < used; i += 2) {
        float t = 0.f;
        OE_HIP(hipEventElapsedTime(&t, pool[i], pool[i + 1]));
        ms += t;
        launches++;
        tag_ms[tags[i / 2]] += t;
        tag_launches[tags[i / 2]]++;
    }
    used = 0;
}

GatherTimer::~GatherTimer() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

void SideStream::create() {
    if (stream) return;
    OE_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    OE_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    OE_HIP(hipEventCreateWithFlags(&join, hipEventDisableTiming));
}

SideStream::~SideStream() {
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (stream) (void)hipStreamDestroy(stream);
}

void launch_schedule(const Schedule& sch, const DeviceSchedule& dev, const SceneDev* d_scenes, float dt, int substeps,
                     hipStream_t stream, GatherTimer* timer, SideStream* side) {
    if (sch.vsegs.empty()) return;
    auto is_q2 = [](const Schedule::GLaunch& gl) { return gl.S == 27 && gl.Q == 27; };
    bool any_q2 = false, any_other = false;
    for (const Schedule::GLaunch& gl : sch.glaunches) (is_q2(gl) ? any_q2 : any_other) = true;
    const bool split = side && side->stream && !timer && any_q2 && any_other;
    const float h = dt / static_cast<float>(substeps);
    BatchVArgs v{d_scenes, dev.vsegs.get(), static_cast<int>(sch.vsegs.size()), h, 0};
    BatchGatherArgs g{d_scenes, dev.segs.get(), 0, 0, 0, h};
    BatchRefreshArgs r{d_scenes, dev.jsegs.get(), 0, 0, 0};
    BatchQ2FusedArgs tk{d_scenes, dev.segs.get(), 0, 0, 0, h, sch.q2_basis, 0};
    BatchFusedArgs ft{d_scenes, dev.segs.get(), 0, 0, 0, h, 0};
    for (int sub = 0; sub < substeps; ++sub) {
        launch_batch_predict(v, sch.vblocks, stream);
        for (int it = 0; it < sch.max_iterations; ++it) {
            g.iteration = r.iteration = tk.iteration = ft.iteration = it;
            for (const Schedule::JLaunch& jl : sch.jlaunches) {
                r.seg_first = jl.seg_first;
                r.seg_count = jl.seg_count;
                if (!launch_batch_ju_refresh(jl.S, jl.Q, r, jl.nblocks, stream))
                    throw std::invalid_argument("unsupported element in batch");
            }
            if (split) {
                OE_HIP(hipEventRecord(side->fork, stream));
                OE_HIP(hipStreamWaitEvent(side->stream, side->fork, 0));
            }
            for (const Schedule::GLaunch& gl : sch.glaunches) {
                const hipStream_t ls = split && is_q2(gl) ? side->stream : stream;
                g.seg_first = gl.seg_first;
                g.seg_count = gl.seg_count;
                if (timer) timer->before(ls, gl.element);
                if (gl.mode == PASS_TASKS && is_q2(gl)) {
                    tk.seg_first = gl.seg_first;
                    tk.seg_count = gl.seg_count;
                    tk.latency = gl.latency ? 1 : 0;
                    if (!gl.latency && q2_pencil(gl.inc))
                        launch_batch_q2_pencil(tk, gl.nblocks, ls);
                    else
                        launch_batch_q2_fused(tk, gl.nblocks, ls);
                } else if (gl.mode == PASS_TASKS) {
                    ft.seg_first = gl.seg_first;
                    ft.seg_count = gl.seg_count;
                    ft.latency = gl.latency ? 1 : 0;
                    if (!launch_batch_fused_tasks(gl.S, gl.Q, ft, gl.nblocks, ls))
                        throw std::invalid_argument("unsupported element in batch");
                } else if (!launch_batch_gather(gl.S, gl.Q, gl.L, gl.QL, gl.mode == PASS_CACHE, g, gl.nblocks, ls)) {
                    throw std::invalid_argument("unsupported gather tile in batch");
                }
                if (timer) timer->after(ls);
            }
            if (split) {
                OE_HIP(hipEventRecord(side->join, side->stream));
                OE_HIP(hipStreamWaitEvent(stream, side->join, 0));
            }
            v.iteration = it;
            launch_batch_fix_box(v, sch.vblocks, stream);
        }
        launch_batch_velocity(v, sch.vblocks, stream);
    }
}

Batch::Batch(Engine* e, std::vector<Scene*> sc) : eng(e), scenes(std::move(sc)), tiles(default_tiles()) {
    if (scenes.empty()) throw std::invalid_argument("empty batch");
    for (size_t i = 0; i < scenes.size(); ++i) {
        Scene* s = scenes[i];
        if (!s) throw std::invalid_argument("null scene in batch");
        if (s->eng != eng) throw std::invalid_argument("scene belongs to another engine");
        // two segments of one launch would update the same vertices
        if (std::find(scenes.begin(), scenes.begin() + static_cast<long>(i), s) != scenes.begin() + static_cast<long>(i))
            throw std::invalid_argument("a scene appears twice in the batch");
        if (s->server) throw std::logic_error("the scene is held by a server");
        check_batchable(*s);
    }
    flags.reserve(scenes.size());
    for (Scene* s : scenes) flags.emplace_back(&eng->pool, fixed_flags(*s));
    d_scenes.allocate(&eng->pool, scenes.size());
    side.create();
    rebuild();
    for (Scene* s : scenes) s->batches.push_back(this);
}

Batch::~Batch() {
    drop_graph();
    for (Scene* s : scenes) {
        if (!s) continue;
        auto& v = s->batches;
        v.erase(std::remove(v.begin(), v.end(), this), v.end());
    }
}

void Batch::scene_destroyed(Scene* s) {
    for (Scene*& p : scenes)
        if (p == s) p = nullptr;
    broken = true;
    drop_graph();
}

void Batch::rebuild() {
    drop_graph();
    std::vector<SceneDev> host;
    for (size_t i = 0; i < scenes.size(); ++i) host.push_back(make_scene_dev(*scenes[i], flags[i].get(), tiles));
    d_scenes.upload(host, eng->stream);
    std::vector<int> idx(scenes.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = static_cast<int>(i);
    sched.build(scenes, idx, tiles);
    dsched.upload(eng, sched, eng->stream, true);
}

void Batch::set_tile(Element e, GatherTile t) {
    if (broken) throw std::logic_error("a scene of this batch was destroyed");
    tiles[e] = t;
    rebuild();
}

void Batch::set_graph(bool on) {
    if (!on) drop_graph();
    use_graph = on;
}

void Batch::set_streams(int n) {
    if (n != streams) drop_graph();   // a captured graph holds the old launch structure
    streams = n;
}

void Batch::drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
}

void Batch::step(float dt, int substeps) {
    if (broken) throw std::logic_error("a scene of this batch was destroyed");
    if (!(dt > 0.f) || substeps <= 0) throw std::invalid_argument("step needs dt > 0 and substeps > 0");
    if (!use_graph || timer.enabled) {
        launch_schedule(sched, dsched, d_scenes.get(), dt, substeps, eng->stream, timer.enabled ? &timer : nullptr,
                        streams == 2 ? &side : nullptr);
    } else {
        if (!graph_exec || graph_dt != dt || graph_sub != substeps) {
            drop_graph();
            OE_HIP(hipStreamBeginCapture(eng->stream, hipStreamCaptureModeThreadLocal));
            try {
                launch_schedule(sched, dsched, d_scenes.get(), dt, substeps, eng->stream, nullptr,
                                streams == 2 ? &side : nullptr);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(eng->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            OE_HIP(hipStreamEndCapture(eng->stream, &graph));
            OE_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
            graph_dt = dt;
            graph_sub = substeps;
        }
        OE_HIP(hipGraphLaunch(graph_exec, eng->stream));
    }
    OE_LAUNCH_CHECK();
    for (Scene* s : scenes) s->steps_done++;
}

}  // namespace oe
