// Continuous-batching request loop over HBM-resident scenes.
//
// There is no counterpart in the reference (it steps every scene once per
// frame from Engine::update, src/Core/Engine.cpp:15-21).  A request asks for
// one scene to be advanced by step(dt, substeps).  Each stream has a worker
// thread that repeatedly
//   * retires finished ticks (HIP event query) and records each request's
//     latency = completion time - arrival time,
//   * when its stream is free, forms a tick from the oldest pending requests
//     whose scenes are not already on the device (one request per scene per
//     tick, up to max_batch scenes, same dt/substeps, one element type),
//   * launches the tick on its stream as one batched step (the launch plan of
//     oe_batch) - so requests that arrive while others run are admitted as
//     soon as a stream and their scene are free, ticks on different streams
//     overlap on the device, and the host-side launch work of different
//     streams runs in parallel.
// Scene state never leaves HBM; scenes are admitted against the engine's
// memory budget when they are created.
// Requests of one scene run in submission order: once a scene's oldest
// pending request is passed over while a tick is formed, none of its later
// requests joins that tick.  Every request gets a ticket; wait(ticket) and
// poll(ticket) report its completion, wait_slot() that of all requests
// submitted for one scene so far (scene reads use it).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_set>
#include <vector>

#include "oe_batch.hpp"

namespace oe {

class Server {
public:
    Server(Engine* eng, int max_batch, int nstreams);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    int add_scene(Scene* s);                                 // returns the slot id
    uint64_t submit(int slot, float dt, int substeps);       // thread-safe
    void drain();                                            // wait for every submitted request
    void wait(uint64_t ticket);                              // wait for one request
    bool poll(uint64_t ticket);                              // true once the request is complete
    void wait_slot(int slot);                                // wait for every request of one scene
    void remove_scene(int slot);                             // wait_slot, then release the scene
    // layout of one element type's passes; refused while requests are pending or in flight
    void set_tile(Element e, GatherTile t);
    GatherTile tile(Element e);
    struct Stats {
        int64_t completed = 0, vertex_updates = 0, batches = 0, max_batch_seen = 0;
        double p50_ms = 0, p99_ms = 0, max_ms = 0, mean_batch = 0;
    };
    Stats stats();
    std::vector<double> latencies();
    std::vector<int> latency_slots();   // slot of each entry of latencies()

private:
    using Clock = std::chrono::steady_clock;
    struct Request {
        uint64_t ticket;
        int slot;
        float dt;
        int substeps;
        Clock::time_point arrival;
    };
    struct Tick {
        int stream = -1;
        std::vector<Request> reqs;
        hipEvent_t done = nullptr;
        int64_t vertex_updates = 0;
    };
    struct StreamSlot {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        Schedule sched;
        DeviceSchedule dev;
        bool busy = false;
    };

    void worker(int si);
    bool try_launch(std::unique_lock<std::mutex>& lk, int si);
    void retire(bool block);
    void grow_table();

    Engine* eng_;
    int max_batch_;
    std::vector<Scene*> scenes_;
    std::vector<DBuf<unsigned char>> flags_;
    std::vector<SceneDev> host_table_;
    DBuf<SceneDev> table_;
    size_t table_uploaded_ = 0;
    std::map<int, GatherTile> tiles_;
    std::vector<StreamSlot> slots_;
    std::vector<char> scene_busy_;
    std::vector<int64_t> slot_open_;           // submitted and not yet completed, per slot
    std::unordered_set<uint64_t> open_tickets_;
    std::condition_variable done_cv_;          // a request completed (or an error)
    std::deque<Request> pending_;
    std::vector<Tick> inflight_;
    int launching_ = 0;   // ticks taken from pending_ and not yet in inflight_
    std::mutex mu_;
    std::condition_variable cv_;
    std::condition_variable idle_cv_;
    std::atomic<bool> stop_{false};
    uint64_t next_ticket_ = 1;
    int64_t submitted_ = 0, completed_ = 0, vertex_updates_ = 0, batches_ = 0, batch_sum_ = 0, max_batch_seen_ = 0;
    std::vector<double> lat_ms_;
    std::vector<int> lat_slot_;
    std::string error_;
    std::vector<std::thread> threads_;
};

}  // namespace oe
