#include "oe_tune.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <mutex>
#include <stdexcept>
#include <tuple>

namespace oe {

namespace {

// The time step of the timing runs.  A colour sweep's cost does not depend on
// it (the same kernels run the same arithmetic), only the state it leaves.
constexpr float kTuneDt = 0.005f;

std::mutex g_cache_mu;
std::map<std::tuple<int, int, int, int>, TuneRecord> g_cache;

bool cache_find(const TuneRecord& key, TuneRecord& out) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    const auto it = g_cache.find({key.device, key.element, key.bucket, key.form});
    if (it == g_cache.end()) return false;
    out = it->second;
    return true;
}

void cache_store(const TuneRecord& r) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache[{r.device, r.element, r.bucket, r.form}] = r;
}

// The state a timing run moves, copied on the device: p, last_p, v, f (with
// the low-order parts in the w lanes) and the VBD inertia target y of every
// scene.
class Snapshot {
public:
    Snapshot(Engine* eng, const std::vector<Scene*>& scenes) : eng_(eng) {
        for (Scene* sc : scenes) {
            GPU_ParticleSystem* ps = sc->ps.get();
            for (DBuf<float4>* b : {&ps->p, &ps->last_p, &ps->v, &ps->f})
                if (b->get()) fields_.push_back(b);
            if (auto* vbd = dynamic_cast<GPU_VBD*>(sc->integrator.get())) fields_.push_back(&vbd->y);
        }
        for (DBuf<float4>* src : fields_) {
            DBuf<float4> copy(&eng_->pool, src->size());
            OE_HIP(hipMemcpyAsync(copy.get(), src->get(), src->bytes(), hipMemcpyDeviceToDevice, eng_->stream));
            saved_.push_back(std::move(copy));
        }
        OE_HIP(hipStreamSynchronize(eng_->stream));
    }
    void restore() {
        for (size_t k = 0; k < fields_.size(); ++k)
            OE_HIP(hipMemcpyAsync(fields_[k]->get(), saved_[k].get(), fields_[k]->bytes(), hipMemcpyDeviceToDevice,
                                  eng_->stream));
        OE_HIP(hipStreamSynchronize(eng_->stream));
    }

private:
    Engine* eng_;
    std::vector<DBuf<float4>*> fields_;
    std::vector<DBuf<float4>> saved_;
};

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    Events() {
        OE_HIP(hipEventCreate(&a));
        OE_HIP(hipEventCreate(&b));
    }
    ~Events() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

struct Timed {
    GatherTile tile;
    double ms;
};

// Time every candidate once, then the best third (at least 3) twice more;
// each candidate's best time counts.  `run(t, reps)` returns milliseconds per
// repetition, or a negative value when the candidate cannot run.
template <class Run>
std::vector<Timed> race(const std::vector<GatherTile>& cands, Run run) {
    std::vector<Timed> res;
    for (const GatherTile& t : cands) {
        const double ms = run(t, 1);
        if (ms >= 0.0) res.push_back(Timed{t, ms});
    }
    if (res.empty()) throw std::runtime_error("autotune: no candidate layout could run");
    std::sort(res.begin(), res.end(), [](const Timed& x, const Timed& y) { return x.ms < y.ms; });
    const size_t keep = std::min(res.size(), std::max<size_t>(3, (res.size() + 2) / 3));
    res.resize(keep);
    for (int pass = 0; pass < 2; ++pass)
        for (Timed& r : res) {
            const double ms = run(r.tile, 3);
            if (ms >= 0.0) r.ms = std::min(r.ms, ms);
        }
    std::sort(res.begin(), res.end(), [](const Timed& x, const Timed& y) { return x.ms < y.ms; });
    return res;
}

bool tasks_possible(const Scene& sc) {
    const GPU_VBD_FEM* f = sc.vbd_fem;
    return f && f->version != Better_Coloration && (f->element != Hexa27 || f->has_q2_basis);
}

// Which colours of a scene take the latency form under a task tile: layouts
// that differ only in the threshold but make the same choice for every colour
// are the same launch sequence and are timed once.
std::vector<bool> latency_pattern(GPU_VBD_FEM& f, const GatherTile& t) {
    const TaskTables& thr = f.throughput_tables(t.inc);
    const bool has_lat = f.latency_tables() != nullptr;
    std::vector<bool> pat(f.nb_colors());
    for (int c = 0; c < f.nb_colors(); ++c) pat[c] = has_lat && thr.color_blocks(c) < t.lat_blocks;
    return pat;
}

void release_unused(GPU_VBD_FEM& f) {
    if (!f.cached()) f.ju.reset();
    const int keep = f.tasks() ? f.fit_inc(f.tile.inc) : -1;
    for (auto it = f.thr_tables.begin(); it != f.thr_tables.end();)
        it = it->first == keep ? std::next(it) : f.thr_tables.erase(it);
}

}  // namespace

int size_bucket(int64_t vertices) {
    int b = 0;
    while (vertices > 1) {
        vertices >>= 1;
        ++b;
    }
    return b;
}

std::vector<GatherTile> tile_candidates(Element el, bool vbd, bool tasks_ok) {
    const int Q = elem_nb_quadrature(el);
    std::vector<int> qls{1};
    if (Q > 1) qls.push_back(Q);
    if (Q == 4) qls.push_back(2);
    if (Q == 27) {
        qls.push_back(3);
        qls.push_back(9);
    }
    std::vector<GatherTile> out;
    for (int L : {4, 8, 16, 32, 64})
        for (int ql : qls) out.push_back(GatherTile{L, ql, PASS_GATHER});
    if (!vbd) {   // the explicit force pass: the element-centric form
        out.push_back(GatherTile{64, 1, PASS_TASKS});
        return out;
    }
    for (int L : {4, 8, 16, 32, 64}) out.push_back(GatherTile{L, 1, PASS_CACHE});
    if (!tasks_ok) return out;
    const int maxi = el == Hexa27 ? kQ2FusedIncPerBlock : max_task_inc(el);
    std::vector<int> incs{maxi, maxi / 2, maxi / 4};
    if (el == Hexa27)   // the pencil form of the throughput pass
        for (int inc : {kQ2PencilIncLatency, 21, kQ2PencilIncPerBlock, kQ2PencilMaxInc}) incs.push_back(inc);
    for (int inc : incs)
        for (int64_t lat : {int64_t{0}, int64_t{256}, kLatencyBlocks, int64_t{2048},
                            std::numeric_limits<int64_t>::max()}) {
            GatherTile t{64, 1, PASS_TASKS};
            t.inc = inc;
            t.lat_blocks = lat;
            out.push_back(t);
        }
    return out;
}

void apply_scene_tile(Scene& sc, const GatherTile& t) {
    if (sc.vbd_fem) {
        const GatherTile before = sc.vbd_fem->tile;
        sc.vbd_fem->tile = t;
        try {
            if (sc.vbd_fem->cached()) sc.vbd_fem->ensure_ju();
            if (sc.vbd_fem->tasks()) sc.vbd_fem->prepare_tasks();
        } catch (...) {
            sc.vbd_fem->tile = before;
            throw;
        }
    }
    if (sc.explicit_fem)
        sc.explicit_fem->tile = t.mode == PASS_TASKS ? GatherTile{64, 1, PASS_TASKS} : GatherTile{t.L, t.QL, PASS_GATHER};
    sc.invalidate_graph();
}

TuneRecord autotune_scene(Scene& sc, bool force) {
    if (!sc.vbd_fem && !sc.explicit_fem) throw std::invalid_argument("autotune: VBD or explicit scenes only");
    if (sc.server || !sc.batches.empty())
        throw std::logic_error("autotune: the scene belongs to a batch or a server");
    TuneRecord rec;
    rec.device = sc.eng->device;
    rec.element = sc.mesh.element;
    rec.bucket = size_bucket(sc.ps->n);
    rec.form = TUNE_SCENE;
    TuneRecord hit;
    if (!force && cache_find(rec, hit)) {
        try {
            apply_scene_tile(sc, hit.tile);
            if (sc.vbd_fem) release_unused(*sc.vbd_fem);
            hit.candidates = 0;
            return hit;
        } catch (const std::invalid_argument&) {
            // this scene cannot take the cached layout (a vertex with more
            // incidences than its blocks hold): tune it
        }
    }
    const bool vbd = sc.vbd_fem != nullptr;
    std::vector<GatherTile> cands = tile_candidates(sc.mesh.element, vbd, vbd && tasks_possible(sc));
    const GatherTile before = vbd ? sc.vbd_fem->tile : sc.explicit_fem->tile;
    Snapshot snap(sc.eng, {&sc});
    Events ev;
    std::vector<std::pair<int, std::vector<bool>>> seen;
    std::vector<GatherTile> unique;
    for (const GatherTile& t : cands) {
        if (t.mode != PASS_TASKS || !vbd) {
            unique.push_back(t);
            continue;
        }
        try {
            std::pair<int, std::vector<bool>> sig{t.inc, latency_pattern(*sc.vbd_fem, t)};
            if (std::find(seen.begin(), seen.end(), sig) != seen.end()) continue;
            seen.push_back(std::move(sig));
            unique.push_back(t);
        } catch (const std::invalid_argument&) {
            // a vertex with more incidences than blocks of this size hold
        }
    }
    auto run = [&](const GatherTile& t, int reps) -> double {
        try {
            apply_scene_tile(sc, t);
        } catch (const std::invalid_argument&) {
            return -1.0;
        } catch (const BudgetError&) {
            return -1.0;
        }
        GPU_Dynamic* d = vbd ? static_cast<GPU_Dynamic*>(sc.vbd_fem) : static_cast<GPU_Dynamic*>(sc.explicit_fem);
        d->step(sc.ps.get(), kTuneDt);   // warm-up
        OE_HIP(hipEventRecord(ev.a, sc.eng->stream));
        for (int r = 0; r < reps; ++r) d->step(sc.ps.get(), kTuneDt);
        OE_HIP(hipEventRecord(ev.b, sc.eng->stream));
        OE_HIP(hipEventSynchronize(ev.b));
        OE_LAUNCH_CHECK();
        float ms = 0.f;
        OE_HIP(hipEventElapsedTime(&ms, ev.a, ev.b));
        return ms / reps;
    };
    std::vector<Timed> res;
    try {
        res = race(unique, run);
    } catch (...) {
        apply_scene_tile(sc, before);
        snap.restore();
        throw;
    }
    snap.restore();
    apply_scene_tile(sc, res.front().tile);
    if (vbd) release_unused(*sc.vbd_fem);
    rec.tile = res.front().tile;
    rec.ms = res.front().ms;
    rec.candidates = static_cast<int>(unique.size());
    cache_store(rec);
    return rec;
}

std::vector<TuneRecord> autotune_batch(Batch& b, bool force) {
    std::map<int, int64_t> vertices;
    std::map<int, bool> tasks_ok;
    std::vector<int64_t> steps_before;
    for (Scene* s : b.scenes) {
        if (!s) throw std::logic_error("a scene of this batch was destroyed");
        vertices[s->mesh.element] += s->ps->n;
        const auto it = tasks_ok.find(s->mesh.element);
        const bool ok = tasks_possible(*s);
        tasks_ok[s->mesh.element] = it == tasks_ok.end() ? ok : (it->second && ok);
        steps_before.push_back(s->steps_done);
    }
    const std::map<int, GatherTile> tiles_before = b.tiles;
    auto tile_before = [&](int el) {
        const auto it = tiles_before.find(el);
        return it != tiles_before.end() ? it->second : default_tile(static_cast<Element>(el));
    };
    std::vector<TuneRecord> out;
    std::vector<int> to_time;
    for (const auto& kv : vertices) {
        TuneRecord rec;
        rec.device = b.eng->device;
        rec.element = kv.first;
        rec.bucket = size_bucket(kv.second);
        rec.form = TUNE_BATCH;
        TuneRecord hit;
        if (!force && cache_find(rec, hit)) {
            try {
                b.set_tile(static_cast<Element>(kv.first), hit.tile);
                hit.candidates = 0;
                out.push_back(hit);
                continue;
            } catch (const std::invalid_argument&) {
                b.set_tile(static_cast<Element>(kv.first), tile_before(kv.first));
            }
        }
        to_time.push_back(kv.first);
    }
    if (to_time.empty()) return out;
    // the batch's timer is borrowed: its state is put back afterwards
    const bool timer_enabled = b.timer.enabled;
    b.timer.collect(b.eng->stream);
    const double timer_ms = b.timer.ms;
    const int64_t timer_launches = b.timer.launches;
    const std::map<int, double> timer_tag_ms = b.timer.tag_ms;
    const std::map<int, int64_t> timer_tag_launches = b.timer.tag_launches;
    auto put_back = [&] {
        b.timer.collect(b.eng->stream);
        b.timer.enabled = timer_enabled;
        b.timer.ms = timer_ms;
        b.timer.launches = timer_launches;
        b.timer.tag_ms = timer_tag_ms;
        b.timer.tag_launches = timer_tag_launches;
    };
    Snapshot snap(b.eng, b.scenes);
    b.timer.enabled = true;
    try {
        for (int el : to_time) {
            std::vector<GatherTile> cands;
            // batches run the throughput form only: one threshold per block size
            for (const GatherTile& t : tile_candidates(static_cast<Element>(el), true, tasks_ok[el]))
                if (t.mode != PASS_TASKS || t.lat_blocks == kLatencyBlocks) cands.push_back(t);
            auto run = [&](const GatherTile& t, int reps) -> double {
                try {
                    b.set_tile(static_cast<Element>(el), t);
                } catch (const std::invalid_argument&) {
                    return -1.0;
                } catch (const BudgetError&) {
                    return -1.0;
                }
                b.step(kTuneDt, 1);   // warm-up
                b.timer.collect(b.eng->stream);
                b.timer.tag_ms[el] = 0.0;
                for (int r = 0; r < reps; ++r) b.step(kTuneDt, 1);
                b.timer.collect(b.eng->stream);
                return b.timer.tag_ms[el] / reps;
            };
            std::vector<Timed> res;
            try {
                res = race(cands, run);
            } catch (...) {
                b.set_tile(static_cast<Element>(el), tile_before(el));
                throw;
            }
            b.set_tile(static_cast<Element>(el), res.front().tile);
            TuneRecord rec;
            rec.device = b.eng->device;
            rec.element = el;
            rec.bucket = size_bucket(vertices[el]);
            rec.form = TUNE_BATCH;
            rec.tile = res.front().tile;
            rec.ms = res.front().ms;
            rec.candidates = static_cast<int>(cands.size());
            cache_store(rec);
            out.push_back(rec);
        }
    } catch (...) {
        snap.restore();
        put_back();
        throw;
    }
    snap.restore();
    for (size_t i = 0; i < b.scenes.size(); ++i) b.scenes[i]->steps_done = steps_before[i];
    put_back();
    return out;
}

std::vector<TuneRecord> tune_cache_records() {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    std::vector<TuneRecord> out;
    for (const auto& kv : g_cache) out.push_back(kv.second);
    return out;
}

void tune_cache_clear() {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache.clear();
}

}  // namespace oe
