// C ABI of the serving loop (pfgpu_server_* / pfgpu_submit / pfgpu_poll in include/pfgpu.h): a
// thin layer over ls2g::ServeLoop, for callers that bind C (ctypes, cgo, JNI) rather than the
// pybind11 module. It replaces nothing in the reference, whose batch entry points are the
// synchronous ContingencyAnalysisCPP.compute / TimeSeriesCPP.compute_Vs (src/main.cpp:1186,1234):
// requests are submitted and their results polled, and requests that overlap in time share launches.
#include <cstring>
#include <memory>
#include <string>

#include "engine.hpp"
#include "host.hpp"
#include "serve.hpp"

struct pfgpu_server {
    std::unique_ptr<ls2g::ServeLoop> loop;
};

namespace {
// exceptions become PFGPU_E_* codes with the message in pfgpu_last_error(), as for the engine
template <class F>
int serve_guarded(F&& f) {
    return pfgpu_engine::guarded(std::forward<F>(f));
}
}  // namespace

extern "C" {

int pfgpu_server_create(size_t hbm_budget, int max_instances_per_launch, int precision, pfgpu_server** out) {
    return serve_guarded([&]() -> int {
        if (!out) throw std::runtime_error("out is NULL");
        *out = nullptr;
        auto s = std::make_unique<pfgpu_server>();
        s->loop = std::make_unique<ls2g::ServeLoop>(hbm_budget, max_instances_per_launch, precision);
        *out = s.release();
        return 0;
    });
}

int pfgpu_server_destroy(pfgpu_server* srv) {
    return serve_guarded([&]() -> int {
        if (!srv) return 0;
        srv->loop->close();
        delete srv;
        return 0;
    });
}

int pfgpu_server_add_grid(pfgpu_server* srv, int n, const int* Yp, const int* Yi, const double* Yx, const double* Sbus,
                          const double* V0, const int* slack_ids, int nslack, const double* slack_weights, const int* pv,
                          int npv, const int* pq, int npq, int max_iter, double tol, int* grid_id) {
    return serve_guarded([&]() -> int {
        if (!srv || !Yp || !Yi || !Yx || !Sbus || !V0 || !slack_ids || !slack_weights || !grid_id || n <= 0 ||
            nslack < 1 || npv < 0 || npq < 0 || (npv && !pv) || (npq && !pq))
            throw std::runtime_error("pfgpu_server_add_grid: invalid arguments");
        ls2g::CscC Y;
        Y.nrow = Y.ncol = n;
        Y.p.assign(Yp, Yp + n + 1);
        Y.i.assign(Yi, Yi + Yp[n]);
        const auto* yx = reinterpret_cast<const ls2g::cplx*>(Yx);
        Y.x.assign(yx, yx + Yp[n]);
        const auto* sb = reinterpret_cast<const ls2g::cplx*>(Sbus);
        const auto* v0 = reinterpret_cast<const ls2g::cplx*>(V0);
        *grid_id = srv->loop->add_grid(Y, ls2g::CplxVect(sb, sb + n), ls2g::CplxVect(v0, v0 + n),
                                       ls2g::IntVect(slack_ids, slack_ids + nslack),
                                       ls2g::RealVect(slack_weights, slack_weights + n), ls2g::IntVect(pv, pv + npv),
                                       ls2g::IntVect(pq, pq + npq), max_iter, tol);
        return 0;
    });
}

int pfgpu_submit(pfgpu_server* srv, int grid_id, int kind, int ninst, const double* Sbus, const int* dptr,
                 const int* dpos, const double* dval, uint64_t* request_id) {
    return serve_guarded([&]() -> int {
        if (!srv || !request_id) throw std::runtime_error("pfgpu_submit: invalid arguments");
        int n = 0;
        ls2g::CplxVect S;
        ls2g::IntVect p, q;
        ls2g::CplxVect v;
        if (kind == PFGPU_REQ_CONTINGENCIES) {
            if (!dptr || !dpos || !dval || ninst < 1) throw std::runtime_error("pfgpu_submit: contingencies need deltas");
            p.assign(dptr, dptr + ninst + 1);
            q.assign(dpos, dpos + dptr[ninst]);
            const auto* dv = reinterpret_cast<const ls2g::cplx*>(dval);
            v.assign(dv, dv + dptr[ninst]);
        }
        // the grid's bus count gives the Sbus row length
        if (Sbus && (kind == PFGPU_REQ_PF || kind == PFGPU_REQ_TIME_SERIES)) {
            n = srv->loop->grid_buses(grid_id);
            const int nr = kind == PFGPU_REQ_PF ? 1 : ninst;
            const auto* sb = reinterpret_cast<const ls2g::cplx*>(Sbus);
            S.assign(sb, sb + (size_t)nr * n);
        }
        if (kind == PFGPU_REQ_TIME_SERIES && !Sbus) throw std::runtime_error("pfgpu_submit: a time-series chunk needs Sbus rows");
        *request_id = srv->loop->submit(grid_id, kind, kind == PFGPU_REQ_PF ? 1 : ninst, S, p, q, v);
        return 0;
    });
}

int pfgpu_poll(pfgpu_server* srv, uint64_t request_id, double timeout_s, int* ready, double* V, int* iters, int* status,
               pfgpu_request_info* info) {
    return serve_guarded([&]() -> int {
        if (!srv || !ready) throw std::runtime_error("pfgpu_poll: invalid arguments");
        *ready = 0;
        ls2g::ServeResult r;
        if (!srv->loop->wait(request_id, timeout_s, r)) return 0;
        *ready = 1;
        if (info) {
            info->ninst = r.ninst;
            info->launch_instances = r.launch_instances;
            info->latency_s = r.latency_s;
            info->queue_s = r.queue_s;
        }
        if (!r.ok) throw std::runtime_error("request failed: " + r.error);
        if (V) std::memcpy(V, r.V.data(), r.V.size() * sizeof(ls2g::cplx));
        if (iters) std::memcpy(iters, r.iters.data(), r.iters.size() * sizeof(int));
        if (status) std::memcpy(status, r.status.data(), r.status.size() * sizeof(int));
        return 0;
    });
}

}  // extern "C"
