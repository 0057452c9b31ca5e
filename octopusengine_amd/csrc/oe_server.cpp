#include "oe_server.hpp"

#include <algorithm>
#include <stdexcept>

namespace oe {

Server::Server(Engine* eng, int max_batch, int nstreams) : eng_(eng), max_batch_(max_batch), tiles_(default_tiles(true)) {
    if (max_batch < 1 || nstreams < 1 || nstreams > 8) throw std::invalid_argument("max_batch >= 1, 1 <= streams <= 8");
    slots_.resize(nstreams);
    for (StreamSlot& s : slots_) {
        OE_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
        OE_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    }
    for (int i = 0; i < nstreams; ++i) threads_.emplace_back([this, i] { worker(i); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (std::thread& t : threads_)
        if (t.joinable()) t.join();
    for (StreamSlot& s : slots_) {
        if (s.stream) {
            (void)hipStreamSynchronize(s.stream);
            (void)hipStreamDestroy(s.stream);
        }
        if (s.done) (void)hipEventDestroy(s.done);
    }
    // requests still pending are dropped; the scenes are free again
    for (Scene* s : scenes_)
        if (s) {
            s->server = nullptr;
            s->server_slot = -1;
        }
}

int Server::add_scene(Scene* s) {
    if (!s || s->eng != eng_) throw std::invalid_argument("scene belongs to another engine");
    // one slot per scene: two slots would let two streams step the same buffers
    if (s->server == this) throw std::invalid_argument("the scene is already held by this server");
    if (s->server) throw std::logic_error("the scene is held by another server");
    if (!s->batches.empty()) throw std::logic_error("the scene belongs to a batch");
    check_batchable(*s);
    std::unique_lock<std::mutex> lk(mu_);
    if (!error_.empty()) throw std::runtime_error(error_);
    // the device table is reallocated only while nothing is in flight
    idle_cv_.wait(lk, [&] { return (inflight_.empty() && launching_ == 0) || !error_.empty(); });
    if (!error_.empty()) throw std::runtime_error(error_);
    flags_.emplace_back(&eng_->pool, fixed_flags(*s));
    host_table_.push_back(make_scene_dev(*s, flags_.back().get(), tiles_));
    scenes_.push_back(s);
    scene_busy_.push_back(0);
    slot_open_.push_back(0);
    grow_table();
    const int slot = static_cast<int>(scenes_.size()) - 1;
    s->server = this;
    s->server_slot = slot;
    return slot;
}

void Server::wait(uint64_t ticket) {
    std::unique_lock<std::mutex> lk(mu_);
    if (ticket == 0 || ticket >= next_ticket_) throw std::invalid_argument("unknown ticket");
    done_cv_.wait(lk, [&] { return !open_tickets_.count(ticket) || !error_.empty(); });
    if (!error_.empty()) throw std::runtime_error(error_);
}

bool Server::poll(uint64_t ticket) {
    std::lock_guard<std::mutex> lk(mu_);
    if (ticket == 0 || ticket >= next_ticket_) throw std::invalid_argument("unknown ticket");
    if (!error_.empty()) throw std::runtime_error(error_);
    return !open_tickets_.count(ticket);
}

void Server::wait_slot(int slot) {
    std::unique_lock<std::mutex> lk(mu_);
    if (slot < 0 || slot >= static_cast<int>(scenes_.size()) || !scenes_[slot])
        throw std::invalid_argument("unknown scene slot");
    done_cv_.wait(lk, [&] { return slot_open_[slot] == 0 || !error_.empty(); });
    if (!error_.empty()) throw std::runtime_error(error_);
}

void Server::remove_scene(int slot) {
    wait_slot(slot);
    std::lock_guard<std::mutex> lk(mu_);
    Scene* s = scenes_[slot];
    scenes_[slot] = nullptr;   // submit() refuses the slot from now on
    s->server = nullptr;
    s->server_slot = -1;
}

void Server::set_tile(Element e, GatherTile t) {
    std::lock_guard<std::mutex> lk(mu_);
    if (!error_.empty()) throw std::runtime_error(error_);
    if (!pending_.empty() || !inflight_.empty() || launching_ != 0)
        throw std::logic_error("the server's layout changes only while no request is pending or in flight");
    std::map<int, GatherTile> tiles = tiles_;
    tiles[e] = t;
    // every scene of the type first, so that a refused layout leaves the server as it was
    std::vector<std::pair<size_t, SceneDev>> changed;
    for (size_t i = 0; i < scenes_.size(); ++i)
        if (scenes_[i] && scenes_[i]->mesh.element == e)
            changed.emplace_back(i, make_scene_dev(*scenes_[i], flags_[i].get(), tiles));
    tiles_ = std::move(tiles);
    for (const auto& c : changed) host_table_[c.first] = c.second;
    table_uploaded_ = 0;
    grow_table();
}

GatherTile Server::tile(Element e) {
    std::lock_guard<std::mutex> lk(mu_);
    const auto it = tiles_.find(e);
    return it != tiles_.end() ? it->second : default_tile(e);
}

void Server::grow_table() {
    if (table_.size() < host_table_.size()) {
        DBuf<SceneDev> bigger(&eng_->pool, std::max<size_t>(16, host_table_.size() * 2));
        table_ = std::move(bigger);
        table_uploaded_ = 0;
    }
    const size_t n = host_table_.size() - table_uploaded_;
    OE_HIP(hipMemcpy(table_.get() + table_uploaded_, host_table_.data() + table_uploaded_, n * sizeof(SceneDev),
                     hipMemcpyHostToDevice));
    table_uploaded_ = host_table_.size();
}

uint64_t Server::submit(int slot, float dt, int substeps) {
    if (!(dt > 0.f) || substeps <= 0) throw std::invalid_argument("step needs dt > 0 and substeps > 0");
    uint64_t t;
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (!error_.empty()) throw std::runtime_error(error_);
        if (slot < 0 || slot >= static_cast<int>(scenes_.size()) || !scenes_[slot])
            throw std::invalid_argument("unknown scene slot");
        t = next_ticket_++;
        pending_.push_back(Request{t, slot, dt, substeps, Clock::now()});
        open_tickets_.insert(t);
        slot_open_[slot]++;
        submitted_++;
    }
    cv_.notify_one();
    return t;
}

void Server::drain() {
    std::unique_lock<std::mutex> lk(mu_);
    idle_cv_.wait(lk, [&] { return (pending_.empty() && inflight_.empty() && launching_ == 0) || !error_.empty(); });
    if (!error_.empty()) throw std::runtime_error(error_);
}

void Server::retire(bool) {
    std::vector<Tick> done;
    {
        std::lock_guard<std::mutex> lk(mu_);
        for (size_t i = 0; i < inflight_.size();) {
            const hipError_t q = hipEventQuery(inflight_[i].done);
            if (q == hipErrorNotReady) {
                ++i;
                continue;
            }
            if (q != hipSuccess) throw HipError(q, std::string("tick failed: ") + hipGetErrorString(q));
            done.push_back(std::move(inflight_[i]));
            inflight_.erase(inflight_.begin() + static_cast<long>(i));
        }
        if (done.empty()) return;
        const Clock::time_point now = Clock::now();
        for (Tick& t : done) {
            for (const Request& r : t.reqs) {
                lat_ms_.push_back(std::chrono::duration<double, std::milli>(now - r.arrival).count());
                lat_slot_.push_back(r.slot);
                scene_busy_[r.slot] = 0;
                slot_open_[r.slot]--;
                open_tickets_.erase(r.ticket);
            }
            completed_ += static_cast<int64_t>(t.reqs.size());
            vertex_updates_ += t.vertex_updates;
            slots_[t.stream].busy = false;
        }
        if (pending_.empty() && inflight_.empty() && launching_ == 0) idle_cv_.notify_all();
    }
    done_cv_.notify_all();
}

bool Server::try_launch(std::unique_lock<std::mutex>& lk, int si) {
    if (slots_[si].busy || pending_.empty()) return false;
    Tick tick;
    tick.stream = si;
    std::vector<char> picked(scenes_.size(), 0);
    float dt = 0.f;
    int sub = 0;
    int element = -1;
    std::deque<Request> rest;
    // A tick holds scenes of one element type: every scene of a tick finishes
    // with the tick's last launch, and a tick's launch count grows with the
    // colour counts of its types (about 14 for P2 beams, 29 for Q2), so mixing
    // types would make every request wait for the other type's colours.
    for (Request& r : pending_) {
        const int el = scenes_[r.slot]->mesh.element;
        const bool fits = static_cast<int>(tick.reqs.size()) < max_batch_ && !scene_busy_[r.slot] && !picked[r.slot] &&
                          (tick.reqs.empty() || (r.dt == dt && r.substeps == sub && el == element));
        if (fits) {
            if (tick.reqs.empty()) {
                dt = r.dt;
                sub = r.substeps;
                element = el;
            }
            picked[r.slot] = 1;
            tick.reqs.push_back(r);
        } else {
            // the scene's later requests wait behind this one
            picked[r.slot] = 1;
            rest.push_back(r);
        }
    }
    if (tick.reqs.empty()) return false;
    pending_.swap(rest);
    for (const Request& r : tick.reqs) scene_busy_[r.slot] = 1;
    StreamSlot& slot = slots_[si];
    slot.busy = true;
    std::vector<Scene*> sc;
    std::vector<int> idx;
    for (const Request& r : tick.reqs) {
        sc.push_back(scenes_[r.slot]);
        idx.push_back(r.slot);
    }
    batches_++;
    batch_sum_ += static_cast<int64_t>(tick.reqs.size());
    max_batch_seen_ = std::max<int64_t>(max_batch_seen_, static_cast<int64_t>(tick.reqs.size()));
    const SceneDev* table = table_.get();
    launching_++;   // holds add_scene() off the table until the tick is in flight
    lk.unlock();
    // plan + launch outside the lock so that submit() is never blocked by it
    slot.sched.build(sc, idx, tiles_, true);
    slot.dev.upload(eng_, slot.sched, slot.stream, false);
    launch_schedule(slot.sched, slot.dev, table, dt, sub, slot.stream, nullptr);
    OE_HIP(hipEventRecord(slot.done, slot.stream));
    tick.done = slot.done;
    tick.vertex_updates = slot.sched.vertex_updates_per_substep * sub;
    for (Scene* s : sc) s->steps_done++;
    lk.lock();
    launching_--;
    inflight_.push_back(std::move(tick));
    return true;
}

void Server::worker(int si) {
    try {
        OE_HIP(hipSetDevice(eng_->device));
        while (true) {
            retire(false);
            std::unique_lock<std::mutex> lk(mu_);
            if (stop_ || !error_.empty()) break;
            if (try_launch(lk, si)) continue;
            if (inflight_.empty() && launching_ == 0) {
                // nothing on the device to retire: sleep until a request comes
                cv_.wait(lk, [&] { return stop_.load() || !pending_.empty(); });
            } else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu_);
        if (error_.empty()) error_ = e.what();
        idle_cv_.notify_all();
        cv_.notify_all();
        done_cv_.notify_all();
    }
}

Server::Stats Server::stats() {
    std::lock_guard<std::mutex> lk(mu_);
    Stats s;
    s.completed = completed_;
    s.vertex_updates = vertex_updates_;
    s.batches = batches_;
    s.max_batch_seen = max_batch_seen_;
    s.mean_batch = batches_ ? static_cast<double>(batch_sum_) / batches_ : 0.0;
    if (!lat_ms_.empty()) {
        std::vector<double> v = lat_ms_;
        std::sort(v.begin(), v.end());
        auto pct = [&](double p) {
            const double x = p * (v.size() - 1);
            const size_t i = static_cast<size_t>(x);
            const double f = x - i;
            return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
        };
        s.p50_ms = pct(0.50);
        s.p99_ms = pct(0.99);
        s.max_ms = v.back();
    }
    return s;
}

std::vector<double> Server::latencies() {
    std::lock_guard<std::mutex> lk(mu_);
    return lat_ms_;
}

std::vector<int> Server::latency_slots() {
    std::lock_guard<std::mutex> lk(mu_);
    return lat_slot_;
}

}  // namespace oe
