// Serving loop: continuous batching of power-flow requests over grids resident on the GPU.
//
// Each resident grid has a worker thread with its own engine context (its own HIP
// stream), so launches for different grids overlap on the device. A worker takes
// every request queued for its grid (FIFO, up to an instance cap and what the HBM
// budget admits) into one batched launch; requests that arrive while a launch is
// in flight join the next one. The grid's Ybus values, base injections and
// warm-start voltages stay in device memory; device buffers only grow.
//
// Request kinds: a single power flow (optionally with its own Sbus), a
// contingency batch (per-outage Ybus deltas, warm start from the resident V), a
// time-series chunk (one Sbus row per step, each warm-started from the resident V).
//
// Device memory: each grid's launch buffers grow with the largest launch it ran and are kept
// (counted in in_use() until close); a launch is admitted when what it would add fits the budget.
// add_grid runs a warm-up launch (one instance per compute unit, in each precision form a launch
// may take), so that the first requests do not pay for the launch shape's set-up.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "host.hpp"

namespace ls2g {

struct ServeResult {
    int ninst = 0;
    CplxVect V;          // ninst x n
    IntVect iters, status;
    double latency_s = 0;  // submit -> result available
    double queue_s = 0;    // submit -> launch start
    int launch_instances = 0;  // size of the launch it rode in
    bool ok = false;
    std::string error;
};

class ServeLoop {
public:
    // hbm_budget: bytes the loop may use for batch workspaces (0: 25% of free memory)
    ServeLoop(size_t hbm_budget, int max_instances_per_launch, int precision);
    ~ServeLoop();
    ServeLoop(const ServeLoop&) = delete;
    ServeLoop& operator=(const ServeLoop&) = delete;

    int add_grid(const CscC& Ybus, const CplxVect& Sbus, const CplxVect& V0, const IntVect& slack_ids,
                 const RealVect& slack_weights, const IntVect& pv, const IntVect& pq, int max_iter, double tol);
    // kind 0: single PF (sbus empty -> resident Sbus); kind 1: contingencies (dptr/dpos/dval, ninst = dptr.size()-1);
    // kind 2: time-series chunk (sbus = ninst rows)
    uint64_t submit(int grid, int kind, int ninst, const CplxVect& sbus, const IntVect& dptr, const IntVect& dpos,
                    const CplxVect& dval);
    bool ready(uint64_t id);
    // waits up to timeout_s (<0: forever); returns false on timeout; throws for an id that was
    // never issued or whose result was already taken
    bool wait(uint64_t id, double timeout_s, ServeResult& out);
    // rows (instances) and bus count of a pending request, for callers that size result buffers
    bool request_shape(uint64_t id, int& ninst, int& n);
    void close();
    int n_grids() const { return (int)grids_.size(); }
    int grid_buses(int grid) const {
        if (grid < 0 || grid >= (int)grids_.size()) throw std::runtime_error("ServeLoop: bad grid");
        return grids_[grid]->n;
    }
    size_t budget() const { return budget_; }
    size_t in_use() const { return in_use_.load(); }
    int launches() const { return launches_.load(); }

private:
    struct Req {
        uint64_t id;
        int kind, ninst;
        CplxVect sbus;
        IntVect dptr, dpos;
        CplxVect dval;
        std::chrono::steady_clock::time_point t_submit, t_start;
    };
    struct Grid {
        int n = 0, nnzY = 0, max_iter = 10;
        double tol = 1e-8;
        CplxVect Yx, Sbus, V0;
        pfgpu_ctx* ctx = nullptr;
        pfgpu_grid* grid = nullptr;
        size_t bytes_per_inst = 0;
        size_t held = 0;  // bytes of this grid's grow-only device buffers (counted in in_use_)
        // resident device data
        double *dY = nullptr, *dS0 = nullptr, *dV0 = nullptr;
        // grow-only launch buffers
        size_t cap = 0;
        double *dS = nullptr, *dV = nullptr, *ddv = nullptr;
        int *ddp = nullptr, *ddq = nullptr, *dIt = nullptr, *dSt = nullptr;
        size_t cap_d = 0;
        std::deque<Req> queue;
        std::thread worker;
    };
    void run(int gi);
    void ensure_capacity(Grid& g, size_t ninst, size_t ndelta);
    static size_t buffer_bytes(const Grid& g, size_t ninst, size_t ndelta);

    size_t budget_;
    int max_inst_;
    int precision_;
    std::vector<std::unique_ptr<Grid>> grids_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::map<uint64_t, ServeResult> done_;
    std::map<uint64_t, std::pair<int, int>> pending_;  // id -> (grid, ninst) until its result is taken
    std::atomic<size_t> in_use_{0};
    std::atomic<int> launches_{0};
    uint64_t next_id_ = 1;
    bool stop_ = false;
};

}  // namespace ls2g
