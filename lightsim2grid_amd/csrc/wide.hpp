// Wide Newton-Raphson launches: every workgroup of the device works on one instance at a time.
//
// The batched kernel gives each instance one workgroup, which fills the device when a launch holds
// hundreds of instances. A single large grid (or a single time-series chain) then runs on one
// compute unit. The wide kernel spreads each phase of one NR iteration over all resident
// workgroups instead: the J fill, the mismatch and the voltage update by entry / bus, and the
// factorisation and the backward solve level by level, each level's tasks split across the whole
// grid (a team of lanes per task). Grid-wide barriers separate the levels (a cooperative launch
// keeps every workgroup resident).
// The dense tail of the factor (the top chain of the elimination tree) runs on workgroup 0.
// The value array lives in global memory (one per launch, reused by the instances in turn).
// Reference: the same per-iteration phases as BaseNRAlgo::compute_pf (BaseNRAlgo.tpp:106-167).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pfgpu_wide {

constexpr int kThreads = 512;  // lanes per workgroup

struct WArgs {
    int n, N, nJ, nzero, nlev, nblev, m;
    int single_slack, max_iter, chain, fresh, nblocks;
    long long nunits, nrows;
    double tol, stall_tol;
    // Ybus row view (mismatch), F slots, per-bus step slots, slack step slot
    const int *yr_ptr, *yr_col, *yr_pos, *fslot_p, *fslot_q;
    const int2* bus_dx;
    int slack_dx;
    // J entries (reference CSC order): Ybus position, row bus, column bus, selector, constant,
    // value slot (bit 30: store the reciprocal)
    const int *j_ypos, *j_rbus, *j_cbus, *j_slot, *zero_slots;
    const uint8_t* j_sel;
    const double* j_cval;
    // factor + forward tasks {out, pair begin, pair end, kind} by level, pairs {l, u, 1/d, -},
    // backward tasks {x, pair begin, pair end, 1/d} by level, pairs {u, x'}, and per level the
    // log2 of the team size
    const int* f_lvl;
    const int4* tasks;
    const int4* pairs;
    const int* f_team;
    const int* b_lvl;
    const int4* btasks;
    const int2* bpairs;
    const int* b_team;
    const double* w;
    // the caller's batch
    const double* Yx;
    long long y_stride;
    const int *dptr, *dpos;
    const double* dval;
    const double* S;
    long long s_stride;
    double* V;
    int *iters, *status, *flags;
    const int* imap;
    // workspace: value slots (FT), Va, Vm, Ibus per bus, this instance's Ybus values, reduction
    // partials (4 * nblocks doubles), control words (pivot flags)
    void* X;
    double *Va, *Vm, *Ib, *Yw, *red;
    int* ctl;
    unsigned int* bar;  // kBarrierWords words of grid-barrier state (cleared by launch())
};

constexpr int kBarrierWords = 10 * 32;

// Launches the wide kernel (cooperative launch) for FT = float (fp32 != 0) or double.
hipError_t launch(const WArgs& a, bool fp32, size_t lds_bytes, hipStream_t stream);

// Workgroups that stay resident together at lds_bytes of dynamic LDS (at most one per compute
// unit), or 0 when the kernel cannot be launched cooperatively.
int resident_blocks(int device, size_t lds_bytes, bool fp32);

// Dynamic LDS of a launch: the dense tail tile and its vector (m x (m + 1) values).
inline size_t lds_bytes(int m, bool fp32) { return (size_t)m * (m + 1) * (fp32 ? 4 : 8) + 64; }

}  // namespace pfgpu_wide
