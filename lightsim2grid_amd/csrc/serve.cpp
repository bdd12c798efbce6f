// Serving loop (see serve.hpp).
#include "serve.hpp"

#include <algorithm>
#include <cstring>
#include <sstream>
#include <stdexcept>

namespace ls2g {

namespace {
using clk = std::chrono::steady_clock;
double secs(clk::duration d) { return std::chrono::duration<double>(d).count(); }
}  // namespace

ServeLoop::ServeLoop(size_t hbm_budget, int max_instances_per_launch, int precision)
    : budget_(hbm_budget), max_inst_(std::max(1, max_instances_per_launch)), precision_(precision) {
    if (precision != PFGPU_FP64 && precision != PFGPU_FP32) throw std::runtime_error("ServeLoop: bad precision");
    if (budget_ == 0) {
        int ncu = 0, lds = 0;
        size_t fr = 0;
        check_rc(pfgpu_ctx_device_info(engine_ctx(), &ncu, &lds, &fr), "pfgpu_ctx_device_info");
        budget_ = fr / 4;
    }
}

ServeLoop::~ServeLoop() { close(); }

void ServeLoop::close() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) return;
        stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& g : grids_)
        if (g->worker.joinable()) g->worker.join();
    {
        // requests still queued fail, so that no caller waits for them forever
        std::lock_guard<std::mutex> lk(mu_);
        for (auto& g : grids_) {
            for (auto& r : g->queue) {
                ServeResult o;
                o.ninst = r.ninst;
                o.ok = false;
                o.error = "ServeLoop closed before the request ran";
                done_[r.id] = std::move(o);
            }
            g->queue.clear();
        }
    }
    cv_done_.notify_all();
    for (auto& g : grids_) {
        for (void* p : {(void*)g->dY, (void*)g->dS0, (void*)g->dV0, (void*)g->dS, (void*)g->dV,
                        (void*)g->ddv, (void*)g->ddp, (void*)g->ddq, (void*)g->dIt, (void*)g->dSt})
            if (p) pfgpu_dev_free(g->ctx, p);
        if (g->grid) pfgpu_grid_destroy(g->grid);
        if (g->ctx) pfgpu_ctx_destroy(g->ctx);
    }
    grids_.clear();
}

int ServeLoop::add_grid(const CscC& Y, const CplxVect& Sbus, const CplxVect& V0, const IntVect& slack_ids,
                        const RealVect& w, const IntVect& pv, const IntVect& pq, int max_iter, double tol) {
    const int n = Y.ncol;
    if ((int)Sbus.size() != n || (int)V0.size() != n) throw std::runtime_error("ServeLoop::add_grid: size mismatch");
    auto g = std::make_unique<Grid>();
    g->n = n;
    g->nnzY = (int)Y.x.size();
    g->max_iter = max_iter;
    g->tol = tol;
    g->Yx = Y.x;
    g->Sbus = Sbus;
    g->V0 = V0;
    check_rc(pfgpu_ctx_create(0, budget_, &g->ctx), "pfgpu_ctx_create");
    check_rc(pfgpu_grid_load(g->ctx, n, Y.p.data(), Y.i.data(), slack_ids.data(), (int)slack_ids.size(), w.data(),
                             pv.data(), (int)pv.size(), pq.data(), (int)pq.size(), 0, &g->grid),
             "pfgpu_grid_load");
    pfgpu_grid_info info{};
    check_rc(pfgpu_grid_get_info(g->grid, &info), "pfgpu_grid_get_info");
    // per-instance device bytes of a launch: Ybus values (contingencies), Va/Vm/Ibus, spilled
    // value slots (upper bound), Sbus / V in / V out, iterations and status
    // transient device bytes one instance of a launch may take in the engine (at most): Va, Vm and
    // Ibus rows, FP64 spilled value slots, flags and fallback list
    g->bytes_per_inst = (size_t)n * 32 + (size_t)info.n_slots * 8 + 64;
    auto alloc = [&](double** p, size_t bytes) {
        void* q = nullptr;
        check_rc(pfgpu_dev_alloc(g->ctx, bytes, &q), "pfgpu_dev_alloc");
        *p = static_cast<double*>(q);
    };
    alloc(&g->dY, Y.x.size() * 16);
    alloc(&g->dS0, (size_t)n * 16);
    alloc(&g->dV0, (size_t)n * 16);
    check_rc(pfgpu_memcpy_h2d(g->ctx, g->dY, Y.x.data(), Y.x.size() * 16), "h2d");
    check_rc(pfgpu_memcpy_h2d(g->ctx, g->dS0, Sbus.data(), (size_t)n * 16), "h2d");
    check_rc(pfgpu_memcpy_h2d(g->ctx, g->dV0, V0.data(), (size_t)n * 16), "h2d");
    // warm-up: a launch of one instance per compute unit at the resident injections, in each
    // precision form a launch of this loop may take (the FP32 form re-solves stalled rows in FP64),
    // so that the first requests do not pay for the launch shape's lane schedule, the kernel's
    // first load and the growth of the engine's workspace to every resident workgroup
    {
        int ncu = 0, lds = 0;
        size_t fr = 0;
        check_rc(pfgpu_ctx_device_info(g->ctx, &ncu, &lds, &fr), "pfgpu_ctx_device_info");
        const int M = std::max(1, std::min(max_inst_, ncu));
        double* wV = nullptr;
        alloc(&wV, (size_t)M * n * 16);
        void *wIt = nullptr, *wSt = nullptr;
        check_rc(pfgpu_dev_alloc(g->ctx, (size_t)M * 4, &wIt), "pfgpu_dev_alloc");
        check_rc(pfgpu_dev_alloc(g->ctx, (size_t)M * 4, &wSt), "pfgpu_dev_alloc");
        std::vector<int> forms{precision_};
        if (precision_ == PFGPU_FP32) forms.push_back(PFGPU_FP64);
        int rc = 0;
        for (int prec : forms) {
            rc = pfgpu_solve_batch_device(g->ctx, g->grid, M, g->dY, 0, nullptr, nullptr, nullptr, g->dS0, 0, g->dV0, 0,
                                          wV, max_iter, tol, prec, 1, static_cast<int*>(wIt), static_cast<int*>(wSt),
                                          nullptr);
            if (rc != 0) break;
        }
        pfgpu_dev_free(g->ctx, wV);
        pfgpu_dev_free(g->ctx, wIt);
        pfgpu_dev_free(g->ctx, wSt);
        check_rc(rc, "pfgpu_solve_batch_device (warm-up)");
    }
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw std::runtime_error("ServeLoop is closed");
    const int gi = (int)grids_.size();
    grids_.push_back(std::move(g));
    grids_[gi]->worker = std::thread(&ServeLoop::run, this, gi);
    return gi;
}

uint64_t ServeLoop::submit(int grid, int kind, int ninst, const CplxVect& sbus, const IntVect& dptr,
                           const IntVect& dpos, const CplxVect& dval) {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw std::runtime_error("ServeLoop is closed");
    if (grid < 0 || grid >= (int)grids_.size()) throw std::runtime_error("ServeLoop::submit: bad grid");
    Grid& g = *grids_[grid];
    if (kind < 0 || kind > 2 || ninst < 1) throw std::runtime_error("ServeLoop::submit: bad request");
    if (kind == 0 && ninst != 1) throw std::runtime_error("ServeLoop::submit: a single power flow has one instance");
    if (kind == 0 && !sbus.empty() && (int)sbus.size() != g.n) throw std::runtime_error("ServeLoop::submit: Sbus size");
    if (kind == 1) {
        if ((int)dptr.size() != ninst + 1 || dptr[0] != 0 || dptr.back() != (int)dpos.size() || dpos.size() != dval.size())
            throw std::runtime_error("ServeLoop::submit: bad contingency deltas");
        for (int q : dpos)
            if (q < 0 || q >= g.nnzY) throw std::runtime_error("ServeLoop::submit: delta position out of range");
    }
    if (kind == 2 && (int)sbus.size() != ninst * g.n) throw std::runtime_error("ServeLoop::submit: Sbus rows");
    if ((size_t)ninst * g.bytes_per_inst > budget_) throw std::runtime_error("ServeLoop::submit: request exceeds the HBM budget");
    Req r;
    r.id = next_id_++;
    r.kind = kind;
    r.ninst = ninst;
    r.sbus = sbus;
    r.dptr = dptr;
    r.dpos = dpos;
    r.dval = dval;
    r.t_submit = clk::now();
    pending_[r.id] = {grid, ninst};
    g.queue.push_back(std::move(r));
    cv_work_.notify_all();
    return next_id_ - 1;
}

bool ServeLoop::request_shape(uint64_t id, int& ninst, int& n) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = pending_.find(id);
    if (it == pending_.end()) return false;
    ninst = it->second.second;
    n = it->second.first < (int)grids_.size() ? grids_[it->second.first]->n : 0;
    return true;
}

bool ServeLoop::ready(uint64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    return done_.count(id) != 0;
}

bool ServeLoop::wait(uint64_t id, double timeout_s, ServeResult& out) {
    std::unique_lock<std::mutex> lk(mu_);
    if (!pending_.count(id)) throw std::runtime_error("ServeLoop: unknown request id (never issued, or its result was taken)");
    auto pred = [&] { return done_.count(id) != 0; };
    if (timeout_s < 0) cv_done_.wait(lk, pred);
    else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return false;
    out = std::move(done_[id]);
    done_.erase(id);
    pending_.erase(id);
    return true;
}

size_t ServeLoop::buffer_bytes(const Grid& g, size_t ninst, size_t ndelta) {
    return ninst * ((size_t)g.n * 32 + 12) + 4 + ndelta * 20;
}

void ServeLoop::ensure_capacity(Grid& g, size_t ninst, size_t ndelta) {
    auto grow = [&](void** p, size_t bytes) {
        if (*p) check_rc(pfgpu_dev_free(g.ctx, *p), "pfgpu_dev_free");
        *p = nullptr;
        check_rc(pfgpu_dev_alloc(g.ctx, bytes, p), "pfgpu_dev_alloc");
    };
    if (ninst > g.cap) {
        const size_t c = ninst;
        grow((void**)&g.dS, c * g.n * 16);
        grow((void**)&g.dV, c * g.n * 16);
        grow((void**)&g.ddp, (c + 1) * 4);
        grow((void**)&g.dIt, c * 4);
        grow((void**)&g.dSt, c * 4);
        g.cap = c;
    }
    if (ndelta > g.cap_d || !g.ddq) {
        const size_t c = std::max<size_t>(ndelta, 16);
        grow((void**)&g.ddq, c * 4);
        grow((void**)&g.ddv, c * 16);
        g.cap_d = c;
    }
    g.held = buffer_bytes(g, g.cap, g.cap_d);
}

void ServeLoop::run(int gi) {
    Grid& g = *grids_[gi];
    const int n = g.n;
    for (;;) {
        std::vector<Req> batch;
        size_t reserved = 0;
        {
            std::unique_lock<std::mutex> lk(mu_);
            for (;;) {
                if (stop_) return;
                if (!g.queue.empty()) {
                    // admit FIFO requests while they fit the instance cap and the HBM budget: the
                    // launch's transient workspace plus what its buffers would grow by
                    int total = 0;
                    size_t nd = 0, need = 0;
                    while (!g.queue.empty()) {
                        const Req& r = g.queue.front();
                        if (!batch.empty() && total + r.ninst > max_inst_) break;
                        const size_t grow = std::max(buffer_bytes(g, std::max(g.cap, (size_t)(total + r.ninst)),
                                                                  std::max(g.cap_d, nd + r.dpos.size())),
                                                     g.held) - g.held;
                        const size_t b = (size_t)(total + r.ninst) * g.bytes_per_inst + grow;
                        if (in_use_.load() + b > budget_) break;
                        total += r.ninst;
                        nd += r.dpos.size();
                        need = b;
                        batch.push_back(std::move(g.queue.front()));
                        g.queue.pop_front();
                    }
                    if (!batch.empty()) {
                        in_use_ += need;
                        reserved = need;
                        break;
                    }
                }
                cv_work_.wait(lk);  // new work, or memory released by another grid
            }
        }
        const auto t_start = clk::now();
        int M = 0, nd = 0;
        bool any_delta = false;
        for (auto& r : batch) {
            r.t_start = t_start;
            M += r.ninst;
            nd += (int)r.dpos.size();
            any_delta |= r.kind == 1;
        }
        ServeResult err_res;
        std::vector<ServeResult> res(batch.size());
        try {
            const size_t held0 = g.held;
            ensure_capacity(g, M, nd);
            // Sbus rows (the request's own, or the resident base row) packed on the host, one copy
            int row = 0;
            std::vector<int> dptr(1, 0);
            IntVect dpos;
            CplxVect dval;
            CplxVect Srows((size_t)M * n);
            for (auto& r : batch) {
                for (int k = 0; k < r.ninst; ++k, ++row) {
                    const cplx* src = r.kind == 2 ? r.sbus.data() + (size_t)k * n
                                      : (r.kind == 0 && !r.sbus.empty()) ? r.sbus.data() : g.Sbus.data();
                    std::copy(src, src + n, Srows.begin() + (size_t)row * n);
                    if (r.kind == 1)
                        for (int q = r.dptr[k]; q < r.dptr[k + 1]; ++q) { dpos.push_back(r.dpos[q]); dval.push_back(r.dval[q]); }
                    dptr.push_back((int)dpos.size());
                }
            }
            check_rc(pfgpu_memcpy_h2d(g.ctx, g.dS, Srows.data(), Srows.size() * 16), "h2d");
            {
                std::lock_guard<std::mutex> lk(mu_);
                in_use_ += g.held - held0;  // kept until close
                in_use_ -= std::min(reserved, g.held - held0);
                reserved -= std::min(reserved, g.held - held0);
            }
            if (any_delta) {
                check_rc(pfgpu_memcpy_h2d(g.ctx, g.ddp, dptr.data(), dptr.size() * 4), "h2d");
                if (!dpos.empty()) {
                    check_rc(pfgpu_memcpy_h2d(g.ctx, g.ddq, dpos.data(), dpos.size() * 4), "h2d");
                    check_rc(pfgpu_memcpy_h2d(g.ctx, g.ddv, dval.data(), dval.size() * 16), "h2d");
                }
            }
            pfgpu_stats st{};
            // one instance per workgroup whatever the launch size: the lane schedule, hence the
            // rounding, of a request does not depend on what else rides in its launch
            check_rc(pfgpu_solve_batch_device(g.ctx, g.grid, M, g.dY, 0, any_delta ? g.ddp : nullptr,
                                              any_delta ? g.ddq : nullptr, any_delta ? g.ddv : nullptr, g.dS, n, g.dV0,
                                              0, g.dV, g.max_iter, g.tol, precision_, 1, g.dIt, g.dSt, &st),
                     "pfgpu_solve_batch_device");
            CplxVect Vall((size_t)M * n);
            IntVect it(M), stt(M);
            check_rc(pfgpu_memcpy_d2h(g.ctx, Vall.data(), g.dV, (size_t)M * n * 16), "d2h");
            check_rc(pfgpu_memcpy_d2h(g.ctx, it.data(), g.dIt, (size_t)M * 4), "d2h");
            check_rc(pfgpu_memcpy_d2h(g.ctx, stt.data(), g.dSt, (size_t)M * 4), "d2h");
            row = 0;
            for (size_t b = 0; b < batch.size(); ++b) {
                const Req& r = batch[b];
                ServeResult& o = res[b];
                o.ninst = r.ninst;
                o.V.assign(Vall.begin() + (size_t)row * n, Vall.begin() + (size_t)(row + r.ninst) * n);
                o.iters.assign(it.begin() + row, it.begin() + row + r.ninst);
                o.status.assign(stt.begin() + row, stt.begin() + row + r.ninst);
                o.launch_instances = M;
                o.ok = true;
                row += r.ninst;
            }
        } catch (const std::exception& e) {
            for (auto& o : res) { o.ok = false; o.error = e.what(); }
        }
        ++launches_;
        const auto t_end = clk::now();
        {
            std::lock_guard<std::mutex> lk(mu_);
            in_use_ -= reserved;
            for (size_t b = 0; b < batch.size(); ++b) {
                res[b].latency_s = secs(t_end - batch[b].t_submit);
                res[b].queue_s = secs(batch[b].t_start - batch[b].t_submit);
                done_[batch[b].id] = std::move(res[b]);
            }
        }
        cv_done_.notify_all();
        cv_work_.notify_all();  // memory released: other grids may admit work
    }
}

}  // namespace ls2g
