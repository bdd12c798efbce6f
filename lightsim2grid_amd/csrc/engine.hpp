// Internals shared by the engine's translation units (not part of the C ABI): the error
// plumbing behind the int status codes and the context structure.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/pfgpu.h"

namespace pfgpu_engine {

// message of the last failing call on this thread (pfgpu_last_error)
std::string& last_error();

struct ApiError : std::runtime_error {
    int code;
    ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Runs f, turning exceptions into a PFGPU_E_* code and the thread's message: no C++
// exception crosses the C ABI.
template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const ApiError& e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        last_error() = "out of host memory";
        return PFGPU_E_NOMEM;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return PFGPU_E_INVALID;
    } catch (...) {
        last_error() = "unknown error";
        return PFGPU_E_INTERNAL;
    }
}

}  // namespace pfgpu_engine

#define HIP_OK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            throw pfgpu_engine::ApiError(PFGPU_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct pfgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0;
    int n_cu = 0;
    int lds_max = 0;
    double wall_hz = 1e8;  // frequency of the device wall clock read by the phase timers
    // grow-only workspaces: the NR launch's per-workgroup area, and the per-call rows (flags,
    // fallback list, counters, phase clocks) that outlive one launch of the call
    void* ws = nullptr;
    size_t ws_bytes = 0;
    void* aux = nullptr;
    size_t aux_bytes = 0;
    // kept initial voltages of an in-place FP32 solve (grow-only)
    void* io = nullptr;
    size_t io_bytes = 0;
    // host-buffer entry points: device copies of the caller's arrays (grow-only)
    void* stage = nullptr;
    size_t stage_bytes = 0;
    // two-pass single chains: the first pass's rows, the second pass's starts, flags (grow-only)
    void* chain = nullptr;
    size_t chain_bytes = 0;
    long long last_rows = 0;            // rows of the last solve call
    const int* last_flags = nullptr;    // and their flags (device)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::mutex mu;
};

