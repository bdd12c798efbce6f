// MI355X batched Newton-Raphson power flow: HIP kernels + the C ABI (include/pfgpu.h).
//
// One workgroup (256 lanes) runs the whole NR loop for G interleaved
// instances (a "persistent" NR kernel: no host round trip per iteration).
// Lane t serves instance g = t % G as slot s = t / G of nslot = 256 / G.
// Per iteration, all phases run inside the kernel, separated by workgroup
// barriers:
//   M  mismatch   Ibus = Ybus V (row view), F = V.conj(Ibus) - Sbus (+ slack
//                 share), scattered to the forward-solve slots, inf-norm
//                 (BaseAlgo.cpp:56-118, :120-129)
//   F  factor     level-scheduled Crout tasks: each entry of the static-pivot
//                 LU (and of the forward-solved right-hand side) is
//                 a - sum X[l]*X[u]*X[1/d]; the J value a is computed in the
//                 task from Ybus, V and Ibus through the scatter map
//                 (_dSbus_dV + known-pattern fill, BaseNRAlgo.tpp:202-260,
//                 :637-688; klu_refactor, KLUSolver.cpp:47-59)
//   B  backward   level-scheduled x_j = (y_j - sum u_ji x_i) / d_j (klu_solve)
//   U  update     slack_absorbed, Va, Vm, V (BaseNRAlgo.tpp:138-158)
// Per level, a team of lanes (power of two) shares each task's dot product
// and reduces it with cross-lane shuffles.
// The value array X of an instance lives in LDS for its first `nlds` slots
// (the hottest, by construction of the plan) and in a global spill buffer for
// the rest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pfgpu.h"
#include "analysis.hpp"
#include "engine.hpp"
#include "wide.hpp"

std::string& pfgpu_engine::last_error() {
    thread_local std::string msg;
    return msg;
}

// Scratch of the no-allocation entry point (pfgpu_solve_batch_device_noalloc): a device array that
// every GPU the library runs on holds until the process ends. Its name is not a device address in
// host code: the address is looked up once per device when the device's first context is created
// (scratch_init) and kept per device; host code uses only that address.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};
std::string g_scratch_err[kMaxDevices];
std::mutex g_noalloc_mu[kMaxDevices];  // no-allocation calls on one device run one at a time

// With `device` current: the address of this device's copy of the scratch array.
void scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    g_scratch_done[device] = true;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    else
        g_scratch_ptr[device] = p;
}
}  // namespace

namespace {

using pfgpu_engine::ApiError;
using pfgpu_engine::guarded;

constexpr int kMaxBlock = 1024;
constexpr int kMaxGraphs = 8;  // captured launch sequences kept per grid
constexpr int kMaxEdits = 8;  // Ybus edits per instance kept in LDS
constexpr int kMaxWaves = kMaxBlock / 64;
constexpr int kMaxG = 32;
constexpr int kEll = 4;  // Ybus row entries read at a time in the mismatch (column groups of its table)
typedef __attribute__((address_space(3))) double lds_dbl;
typedef __attribute__((address_space(3))) int lds_int;
typedef __attribute__((address_space(3))) unsigned char lds_u8;

// Per-workgroup state of the NR kernel, in dynamic LDS after the value slots and the level
// table. The kernel declares no static LDS, so its value slots start at LDS address 0.
// Layout: part[kWaves][G], slack[G], flast[G], dval[G][kMaxEdits][2] (doubles), then done,
// it, status, piv, grow [G], dpos[G][kMaxEdits], nd, d0 [G], ndone, grp (ints).
__host__ __device__ constexpr int wg_state_bytes(int G, int NT) {
    return 8 * G * (NT / 64 + 2 + 2 * kMaxEdits) + 4 * (G * (7 + kMaxEdits) + 2);
}

// Per-row diagnostic flags written by the NR kernel (pfgpu_stats / the fallback passes):
constexpr int kFlagStall = 1;   // not converged, but the last mismatch is below the stall level
constexpr int kFlagGrowth = 2;  // a pivot lost more than the growth threshold to cancellation
constexpr int kStatusNotRun = 8;  // chain step after the chain's first failure (NotInitError)
// Host re-pivot threshold: a diagonal pivot is kept while it holds at least this fraction of its
// column's largest candidate (klu_defaults' tol)
constexpr double kRepivotTau = 1e-3;

// ---------------------------------------------------------------------------
// Kernel arguments
// ---------------------------------------------------------------------------
struct KArgs {
    int n, N, nJ, nzero, nlev, nblev, nlds, nspill, single_slack, ninst, max_iter, m;
    double tol;
    const int *yr_ptr, *yr_col, *yr_pos, *fslot_p, *fslot_q, *ux_bus, *ux_slot;
    const int2* yr_pc;     // per Ybus row entry: {CSC position, column bus}
    const int2* yr_ell;    // [ell_w][n]: the entries of every row as in yr_pc, {-1, row} past its end
    // J fill by Ybus entry: position, {row bus, column bus}, 4 target slots, 4 J indices
    const int* ye_pos;
    const int2* ye_rc;
    const int4* ye_slot;
    const int4* ye_jidx;
    int nye;
    const int *jc_slot, *jc_jidx;  // constant J entries
    const double* jc_val;
    int njc;
    // per-instance Ybus edits applied on the fly: Y_inst = Y - dval at dpos, entries
    // [dptr[i], dptr[i+1]) of instance i (dptr == nullptr: no edits)
    const int *dptr, *dpos;
    const double* dval;
    const int8_t* ux_kind;
    const int2* bus_dx;  // per bus: {slot of its Va step, slot of its Vm step}, -1 if not an unknown
    int slack_dx;        // slot of the slack_absorbed step, -1 for the single-slack form
    const int *j_ypos, *j_rbus, *j_cbus, *j_slot, *zero_slots;
    const uint8_t* j_sel;
    const double* j_cval;
    // lane streams (per launch shape): item k of lane s < U_l at level l sits at
    // base_l + k * U_l + s, base_l = sum over earlier levels of K * U
    const int* lv_K;     // {items per lane, used lanes}: nlev factor levels, then nblev backward levels
    const int4* fs;      // {l, u, 1/d, end}: l < 0 no pair; end >= 0: slot | log2(team) << 26 | recip << 30
    const int4* bs;      // {u, x', end, 1/d}: u < 0 no pair; end >= 0: slot | log2(team) << 26
    int lvk_off;         // byte offset of the level table in dynamic LDS
    int sh_off;          // byte offset of the per-workgroup state in dynamic LDS
    int v_off;           // lean launches: byte offset of the instances' V in dynamic LDS ([bus][G] complex)
    const double* Yell;  // Ybus values in yr_ell order ([ell_w][n] complex, 0 past a row's end), or null
    const double* Yye;   // Ybus values in ye_pos order, or null
    const double* w;
    const double* Yx;
    long long y_stride;  // doubles between instances (0 = shared)
    const double* S;
    long long s_stride;  // doubles between instances (0 = shared)
    double* V;           // [ninst][n] complex
    double* Va;          // [ninst][n]
    double* Vm;          // [ninst][n]
    double* Ib;          // [ninst][nyd] complex: Ibus of the bus of diagonal J-fill entry e < nyd
    const int* ye_diag_of;  // per bus: its diagonal J-fill entry, -1 if none
    int nyd;             // diagonal J-fill entries (they come first)
    void* Xg;            // spill [nblocks][nspill][G]
    int* iters;
    int* status;
    double* Jx_last;     // optional: the J values (reference CSC order) of row jx_row at its J fills
    long long jx_row;
    // launch instance i solves the caller's instance imap[i] (null: i): the index of its Ybus
    // values / edits, its V0 and V rows, its Sbus rows and its outputs
    const int* imap;
    // time-series chains: instance c runs the steps c * chain + k, k < chain, one after the
    // other, each warm-started from the previous step's V (TimeSeries.cpp:80-103); the rows
    // of V, Sbus, iters, status and flags are steps; a chain stops at its first failure and
    // its later steps get kStatusNotRun. chain == 1: one solve per instance.
    int chain;
    long long nrows;     // rows of V / iters / status (ninst * chain, the last chain may be short)
    int fresh;           // 1: the first iteration factors a fresh solver (pivot failure -> SolverFactor)
    double stall_tol;    // kFlagStall: the last mismatch of a failed solve is below this
    double growth_tau;   // kFlagGrowth: |pivot| < growth_tau * |its value before elimination|
    int* flags;          // per row (null: not written)
    unsigned long long* phase_t;  // workgroup 0's wall clock per phase (null: not measured)
    int* queue;          // work-queue counter (zeroed before the launch)
    long long ngroups;   // instance groups of G: ceil(ninst / G)
    // DC power flow (BaseDCAlgo::compute_pf, BaseDCAlgo.tpp:15-179) on the same factor and
    // solve phases: the grid is loaded in the single-slack layout with every non-slack bus
    // carrying only its angle, J = Re(Ybus) without the slack rows / columns, F = J theta -
    // Re(Sbus) with theta = Va - Va(slack0); one Newton step solves the linear system, then
    // the status is SolverSolve (7) if theta is not finite or |theta| >= 1e6
    int dc;
    int slack0;
    // Fast-decoupled power flow (BaseFDPFAlgo::compute_pf, BaseFDPFAlgo.tpp:12-160) on the same
    // phases: the grid is loaded in the single-slack layout with every J entry a constant, B' on
    // the P-row / angle-column block, B'' on the Q-row / magnitude-column block and 0 elsewhere, so
    // that the factor is that of the block-diagonal [B' 0; 0 B'']. F = (V conj(Ybus V) - Sbus +
    // Re(sum Sbus) w) / |V|. The loop alternates half-steps: odd ones apply only the angle steps
    // (P half), even ones only the magnitude steps (Q half), with the convergence test after each;
    // s_it counts half-steps, the reported iterations are P half-steps, max_iter bounds them.
    int fdpf;
};

struct c64 {
    double re, im;
};
__device__ __forceinline__ c64 cmul(c64 a, c64 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ c64 cconj(c64 a) { return {a.re, -a.im}; }
__device__ __forceinline__ c64 csub(c64 a, c64 b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ c64 cadd(c64 a, c64 b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ c64 ld_c(const double* p, int i) { return {p[2 * i], p[2 * i + 1]}; }

// Reduce over the nslot lanes of one instance: shuffles inside the wave
// (lanes g, g+G, ... of a wave), then the kWaves partials through LDS.
template <int G, int NT, bool IS_MAX>
__device__ __forceinline__ double inst_reduce(double v, lds_dbl* part) {
    constexpr int kWaves = NT / 64;
#pragma unroll
    for (int o = G; o < 64; o <<= 1) {
        const double u = __shfl_xor(v, o);
        v = IS_MAX ? fmax(v, u) : fmin(v, u);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane < G) part[wave * G + lane] = v;
    __syncthreads();
    const int g = threadIdx.x % G;
    double r = part[g];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) r = IS_MAX ? fmax(r, part[k * G + g]) : fmin(r, part[k * G + g]);
    return r;
}

// J value at the current V / Ibus (dS_dVa / dS_dVm, BaseNRAlgo.tpp:202-260) of a J entry
// with selector sel (0/1: Re/Im dS_dVa, 2/3: Re/Im dS_dVm, 4: constant cv) built on the
// Ybus entry y = Y(r, c); diag says r == c.
__device__ __forceinline__ double jvalue_of(int sel, bool diag, c64 y, c64 vr, c64 vc, c64 ib, double cv) {
    if (sel == 4) return cv;
    if (sel >= 2) {
        const double ic = 1.0 / sqrt(vc.re * vc.re + vc.im * vc.im);
        c64 d = cmul(cconj(cmul(y, {vc.re * ic, vc.im * ic})), vr);
        if (diag) {
            const double ir = 1.0 / sqrt(vr.re * vr.re + vr.im * vr.im);
            d = cadd(d, cmul(cconj(ib), {vr.re * ir, vr.im * ir}));
        }
        return sel == 2 ? d.re : d.im;
    }
    c64 t = cmul(y, vc);
    if (diag) t = csub(t, ib);
    const c64 d = cmul(cconj({-t.re, -t.im}), {-vr.im, vr.re});  // * (j V_r)
    return sel == 0 ? d.re : d.im;
}

// Orders LDS accesses between the lanes of one wave (they run in lockstep; the
// fence makes the compiler wait for outstanding LDS operations and not reorder).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double inf_if_nonfinite(double a) { return (a <= 1.7976931348623157e308) ? a : INFINITY; }

// Polar <-> rectangular conversions of one lane's buses (b0, b0 + step, ...).
// Kept out of line: inlined FP64 atan2 / sincos raise the NR kernel's register
// count past 128 VGPRs, which at 1024 lanes per workgroup means scratch spills
// in the factor loop.
// Va = arg(V), Vm = |V| (BaseNRAlgo.tpp:77-79, BaseNRSingleSlackAlgo.tpp:131-132)
// V holds bus b at V[2 b vs] (vs = 1 in global memory, G in LDS).
template <typename VP>
__device__ __noinline__ void polar_of(VP V, int vs, double* __restrict__ Va, double* __restrict__ Vm, int b0, int n,
                                      int step) {
    for (int b = b0; b < n; b += step) {
        const double vr = V[2 * b * vs], vi = V[2 * b * vs + 1];
        Va[b] = atan2(vi, vr);
        Vm[b] = sqrt(vr * vr + vi * vi);
    }
}
// V = Vm * exp(j Va) (BaseNRAlgo.tpp:150-151); returns min Vm over the lane's buses
template <typename VP>
__device__ __noinline__ double rect_of(VP V, int vs, const double* __restrict__ Va, const double* __restrict__ Vm,
                                       int b0, int n, int step) {
    double vmin = INFINITY;
    for (int b = b0; b < n; b += step) {
        double sn, cs;
        sincos(Va[b], &sn, &cs);
        const double m = Vm[b];
        V[2 * b * vs] = m * cs;
        V[2 * b * vs + 1] = m * sn;
        vmin = fmin(vmin, m);
    }
    return vmin;
}

// One bus of the voltage update (BaseNRAlgo.tpp:138-151): Va and Vm (taken from V first in
// the single-slack form, BaseNRSingleSlackAlgo.tpp:131-132) minus the Newton step, then
// V = Vm exp(j Va); returns Vm.
template <typename VP>
__device__ __noinline__ double update_bus(VP V, int vs, double* __restrict__ Va, double* __restrict__ Vm, int b,
                                          double dva, double dvm, bool polar_first) {
    double va, vm;
    if (polar_first) {
        const double vr = V[2 * b * vs], vi = V[2 * b * vs + 1];
        va = atan2(vi, vr);
        vm = sqrt(vr * vr + vi * vi);
    } else {
        va = Va[b];
        vm = Vm[b];
    }
    va -= dva;
    vm -= dvm;
    Va[b] = va;
    Vm[b] = vm;
    double sn, cs;
    sincos(va, &sn, &cs);
    V[2 * b * vs] = vm * cs;
    V[2 * b * vs + 1] = vm * sn;
    return vm;
}

// Lane streams. Level l uses lanes [0, U) with K items per lane, item k of lane s
// at base + k * U + s. End word (forward: item.w, backward: item.z): bits 0-24
// slot, 25 assign (a fill-in entry: its slot held another value before), 26-28
// log2(team), 29 reciprocal, 30 end of task, 31 the lane's last item of the level
// (the lane loads nothing after it).
constexpr int kSC = 2;  // items per prefetch group of the general walk
constexpr int kHB = 4;  // items of a level loaded ahead, during the level before it (lean walk)
constexpr int kEnd = 1 << 30;
__device__ __forceinline__ int4 no_item() { return int4{-1, -1, 0, 0}; }

// Moves v from lane + D to each lane (D < 16: within a row of 16 lanes by DPP row_shl,
// no LDS traffic; lanes whose source is outside the row get 0).
template <int D>
__device__ __forceinline__ float lane_down(float v) {
    if constexpr (D < 16) {
        return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x100 + D, 0xF, 0xF, true));
    } else {
        return __shfl_down(v, D);
    }
}
template <int D>
__device__ __forceinline__ double lane_down(double v) {
    if constexpr (D < 16) {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x100 + D, 0xF, 0xF, true);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x100 + D, 0xF, 0xF, true);
        return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
    } else {
        return __shfl_down(v, D);
    }
}

// Sum over a team of 2^lt lanes (stride G, aligned to 2^lt * G) into its lowest lane.
// maxlt (uniform) bounds the steps; lanes of smaller teams drop the later steps by a
// select, not a branch. All lanes of a team share lt, so the lanes a leader reads are
// active and in its team.
template <int G, typename FT>
__device__ __forceinline__ FT team_sum(FT acc, int lt, int maxlt) {
    if (maxlt > 0) { const FT t = lane_down<1 * G>(acc); acc += lt > 0 ? t : (FT)0; }
    if constexpr (2 * G <= 32) if (maxlt > 1) { const FT t = lane_down<2 * G>(acc); acc += lt > 1 ? t : (FT)0; }
    if constexpr (4 * G <= 32) if (maxlt > 2) { const FT t = lane_down<4 * G>(acc); acc += lt > 2 ? t : (FT)0; }
    if constexpr (8 * G <= 32) if (maxlt > 3) { const FT t = lane_down<8 * G>(acc); acc += lt > 3 ? t : (FT)0; }
    if constexpr (16 * G <= 32) if (maxlt > 4) { const FT t = lane_down<16 * G>(acc); acc += lt > 4 ? t : (FT)0; }
    if constexpr (32 * G <= 32) if (maxlt > 5) { const FT t = lane_down<32 * G>(acc); acc += lt > 5 ? t : (FT)0; }
    return acc;
}

// Reciprocal of a pivot: the FP32 factor form uses the hardware reciprocal (1 ulp),
// the FP64 form an IEEE division.
__device__ __forceinline__ float pivot_recip(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double pivot_recip(double v) { return 1.0 / v; }
// v of lane `lane` (uniform over the wave) in every lane
__device__ __forceinline__ float lane_read(float v, int lane) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
__device__ __forceinline__ double lane_read(double v, int lane) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// zero or non-finite pivot (klu_refactor failure)
template <typename FT>
__device__ __forceinline__ bool bad_pivot(FT v) { return !(fabs(v) > (FT)0 && fabs(v) < (FT)INFINITY); }

template <typename FT, int G, int NT, bool LEAN>
__global__ void __launch_bounds__(NT) k_nr(const KArgs a) {
    constexpr int kWaves = NT / 64;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    // explicit LDS address space: every access is a ds_read/ds_write, never a flat
    // pointer into LDS (flat pointers into LDS mixed with global ones trip the gfx950 backend)
    typedef __attribute__((address_space(3))) FT lds_ft;
    lds_ft* Xl = (lds_ft*)(lds_raw);
    lds_u8* const shb = (lds_u8*)lds_raw + a.sh_off;
    lds_dbl* const s_part = (lds_dbl*)shb;  // [kWaves][G]
    lds_dbl* const s_slack = s_part + kWaves * G;
    lds_dbl* const s_flast = s_slack + G;  // mismatch of the current solve's last iteration
    lds_dbl* const s_dval = s_flast + G;  // [G][kMaxEdits][2]: this instance's Ybus edits
    lds_int* const s_done = (lds_int*)(s_dval + G * kMaxEdits * 2);
    lds_int* const s_it = s_done + G;
    lds_int* const s_status = s_it + G;
    lds_int* const s_piv = s_status + G;
    lds_int* const s_grow = s_piv + G;
    lds_int* const s_dpos = s_grow + G;  // [G][kMaxEdits]
    lds_int* const s_nd = s_dpos + G * kMaxEdits;
    lds_int* const s_d0 = s_nd + G;
    lds_int* const s_ndone = s_d0 + G;
    lds_int* const s_grp = s_ndone + 1;
    // {items per lane, used lanes} of every factor level, then of every backward level
    // (in LDS after the value slots)
    lds_int* lvK = (lds_int*)(lds_raw + a.lvk_off);
    // the same table read through the scalar cache (uniform loads from the constant address space)
    typedef __attribute__((address_space(4))) const int cst_int;
    const cst_int* const lv_s = (const cst_int*)a.lv_K;
    for (int i = threadIdx.x; i < 2 * (a.nlev + a.nblev); i += NT) lvK[i] = a.lv_K[i];

    const int tid = threadIdx.x;
    const int g = tid % G, s = tid / G;
    constexpr int nslot = NT / G;
    const int n = a.n;
    // phase clocks of workgroup 0 (J fill, factor + forward, dense tail, backward, update, mismatch)
    const bool timing = a.phase_t != nullptr && blockIdx.x == 0 && tid == 0;
    unsigned long long t_ph = timing ? wall_clock64() : 0ull;
    auto tmark = [&](int ph) {
        if (timing) {
            const unsigned long long t = wall_clock64();
            a.phase_t[ph] += t - t_ph;
            t_ph = t;
        }
    };
    // Persistent workgroups: a launch has as many workgroups as can be resident; each takes
    // groups of G instances from a work queue (a counter zeroed by the launch's first kernel)
    // until none is left, so the per-workgroup workspace does not grow with the batch.
    for (;;) {
    if (tid == 0) *s_grp = atomicAdd(a.queue, 1);
    __syncthreads();
    const long long grp = *s_grp;
    __syncthreads();
    if (grp >= a.ngroups) break;
    const long long inst = grp * G + g;
    const bool valid = inst < a.ninst;
    const long long ii = (long long)blockIdx.x * G + g;  // workspace row of this lane's instance
    const long long oi = valid ? (a.imap ? (long long)a.imap[inst] : inst) : 0;  // the caller's instance
    FT* Xgb = reinterpret_cast<FT*>(a.Xg) + (size_t)blockIdx.x * (size_t)a.nspill * G;
    const int nlds = a.nlds;
    auto XL = [&](int slot) -> FT {
        return slot < nlds ? Xl[slot * G + g] : Xgb[(size_t)(slot - nlds) * G + g];
    };
    // Operand read for the level loops: always an LDS read (clamped slot) and, only for a
    // spilled slot, a global load into a separate register. A shared destination register
    // would make every LDS read wait for all outstanding global loads (the prefetched items).
    // lean launches (every slot in LDS): items carry byte offsets of instance 0's slots, which
    // are their LDS addresses (the value slots start at LDS address 0: the kernel has no static
    // LDS, checked by the host before a lean launch), so an operand read takes no address add
    const int xg = g * (int)sizeof(FT);
    auto XB = [&](int off) -> FT { return *(lds_ft*)(size_t)(unsigned)(off + xg); };
    auto XBset = [&](int off, FT v) { *(lds_ft*)(size_t)(unsigned)(off + xg) = v; };
    auto XLo = [&](int slot) -> FT {
        const bool in_lds = slot < nlds;
        const FT vl = Xl[min(slot, nlds - 1) * G + g];
        FT vg = (FT)0;
        if (!in_lds) vg = Xgb[(size_t)(slot - nlds) * G + g];
        return (in_lds ? vl : (FT)0) + vg;  // a sum, not a select: keeps vl and vg apart
    };
    auto XS = [&](int slot, FT v) {
        if (slot < nlds) Xl[slot * G + g] = v;
        else Xgb[(size_t)(slot - nlds) * G + g] = v;
    };
    const double* Y = a.Yx + oi * a.y_stride;
    // the current step's row (chain mode) or the instance's row
    long long row = oi * a.chain;
    bool run = valid;  // this instance solves the current step (uniform over its lanes)
    const double* S = a.S + row * a.s_stride;
    double* V = a.V + row * 2 * (long long)n;
    double* Va = a.Va + ii * (long long)n;
    double* Vm = a.Vm + ii * (long long)n;
    double* Ib = a.Ib + ii * 2 * (long long)a.nyd;
    // V of this instance: in LDS for lean launches (loaded here, stored back at the end), else global
    lds_dbl* const Vl = (lds_dbl*)(lds_raw + a.v_off) + 2 * g;
    auto ldV = [&](int b) -> c64 {
        if constexpr (LEAN) return {Vl[2 * b * G], Vl[2 * b * G + 1]};
        else return ld_c(V, b);
    };
    if constexpr (LEAN)
        if (valid)
            for (int b = s; b < n; b += nslot) {
                Vl[2 * b * G] = V[2 * b];
                Vl[2 * b * G + 1] = V[2 * b + 1];
            }
    // this instance's Ybus edits: the first kMaxEdits in LDS, any further ones read from memory
    if (s == 0) {
        s_nd[g] = 0;
        s_d0[g] = 0;
        if (a.dptr != nullptr && valid) {
            s_d0[g] = a.dptr[oi];
            s_nd[g] = a.dptr[oi + 1] - a.dptr[oi];
        }
    }
    __syncthreads();
    if (s < kMaxEdits && s < s_nd[g]) {
        const int k = s_d0[g] + s;
        s_dpos[g * kMaxEdits + s] = a.dpos[k];
        s_dval[(g * kMaxEdits + s) * 2] = a.dval[2 * k];
        s_dval[(g * kMaxEdits + s) * 2 + 1] = a.dval[2 * k + 1];
    }
    const int nd = s_nd[g], d0 = s_d0[g];
    // Ybus value at CSC position pos for this instance
    // The first kMaxEdits edit positions are read into registers at the start of a phase
    // (edit_pos), so that an entry is checked against them with compares only.
    typedef int EditPos[kMaxEdits];
    auto edit_pos = [&](EditPos& dp) {
#pragma unroll
        for (int d = 0; d < kMaxEdits; ++d) dp[d] = d < nd ? s_dpos[g * kMaxEdits + d] : INT_MIN;
    };
    auto edits = [&](c64 y, int pos, const EditPos& dp) -> c64 {
        bool hit = false;
#pragma unroll
        for (int d = 0; d < kMaxEdits; ++d) hit |= dp[d] == pos;
        if (hit)
#pragma unroll
            for (int d = 0; d < kMaxEdits; ++d)
                if (dp[d] == pos) {
                    y.re -= s_dval[(g * kMaxEdits + d) * 2];
                    y.im -= s_dval[(g * kMaxEdits + d) * 2 + 1];
                }
        for (int d = kMaxEdits; d < nd; ++d)
            if (a.dpos[d0 + d] == pos) {
                y.re -= a.dval[2 * (d0 + d)];
                y.im -= a.dval[2 * (d0 + d) + 1];
            }
        return y;
    };
    auto Yat = [&](int pos, const EditPos& dp) -> c64 { return edits(ld_c(Y, pos), pos, dp); };
    // Ybus values gathered in the order of the mismatch / J-fill tables (shared Ybus only):
    // loaded alongside the table entries instead of after them
    const double* const Yell = a.Yell;
    const double* const Yye = a.Yye;

    bool alive = valid;  // chain mode: the chain has not failed yet
    for (int step = 0; step < a.chain; ++step) {
    row = oi * a.chain + step;
    run = alive && row < a.nrows;
    S = a.S + (run ? row : 0) * a.s_stride;
    V = a.V + (run ? row : 0) * 2 * (long long)n;
    if (!LEAN && run && step > 0)  // warm start from the previous step's V (own buses: no barrier)
        for (int b = s; b < n; b += nslot) {
            V[2 * b] = V[2 * b - 2 * (long long)n];
            V[2 * b + 1] = V[2 * b + 1 - 2 * (long long)n];
        }
    // ---- init: Va/Vm from V, slack_absorbed = Re(sum Sbus) (BaseNRAlgo.tpp:60,77-79)
    double ssum = 0.0;
    if (run) {
        if constexpr (LEAN) polar_of(Vl, G, Va, Vm, s, n, nslot);
        else polar_of(V, 1, Va, Vm, s, n, nslot);
        for (int b = s; b < n; b += nslot) ssum += S[2 * b];
    }
    // sum-reduce (as max of a single value is not a sum, do a dedicated pass)
#pragma unroll
    for (int o = G; o < 64; o <<= 1) ssum += __shfl_xor(ssum, o);
    {
        const int lane = tid & 63, wave = tid >> 6;
        if (lane < G) s_part[wave * G + lane] = ssum;
        __syncthreads();
        if (s == 0) {
            double t = 0;
            for (int k = 0; k < kWaves; ++k) t += s_part[k * G + g];
            s_slack[g] = t;
            s_it[g] = 0;
            s_status[g] = 0;
            s_piv[g] = 0;
            s_grow[g] = 0;
            s_flast[g] = INFINITY;
            s_done[g] = run ? 0 : 1;
        }
        __syncthreads();
    }

    // ---- mismatch phase ------------------------------------------------------
    // Ibus = Y V by rows, from a column-major table of every row's entries (ell_w columns,
    // padded with {-1, row}): a row's entries are read kEll at a time with no row-pointer
    // load before them, coalesced across lanes, and summed in row order.
    auto mismatch = [&]() -> double {
        double fm = 0.0;
        if (run && !s_done[g] && a.dc) {
            // DC: F = Re(Y) theta - Re(S) over the non-slack columns; fm = max |theta| of the rows
            EditPos dp;
            edit_pos(dp);
            const double vas = Va[a.slack0];
            for (int r = s; r < n; r += nslot) {
                const int len = a.yr_ptr[r + 1] - a.yr_ptr[r];
                const int fp = a.fslot_p[r];
                double acc = 0.0;
                for (int k0 = 0; k0 == 0 || k0 < len; k0 += kEll) {
#pragma unroll
                    for (int k = 0; k < kEll; ++k) {
                        const int2 pc = a.yr_ell[(size_t)(k0 + k) * n + r];
                        if (pc.x < 0 || a.bus_dx[pc.y].x < 0) continue;
                        const c64 yv = Yell ? edits(ld_c(Yell, (k0 + k) * n + r), pc.x, dp) : Yat(pc.x, dp);
                        acc += yv.re * (Va[pc.y] - vas);
                    }
                }
                if (fp >= 0) {
                    XS(fp, (FT)(acc - S[2 * r]));
                    fm = fmax(fm, inf_if_nonfinite(fabs(Va[r] - vas)));
                }
            }
        } else if (run && !s_done[g]) {
            const double sa = s_slack[g];
            EditPos dp;
            edit_pos(dp);
            for (int r = s; r < n; r += nslot) {
                c64 acc = {0.0, 0.0};
                // the row's own data, loaded with its first entries
                const int len = a.yr_ptr[r + 1] - a.yr_ptr[r];
                const c64 sr = ld_c(S, r);
                const double wr = (a.single_slack && !a.fdpf) ? 0.0 : a.w[r];
                const int fp = a.fslot_p[r], fq = a.fslot_q[r], ed = a.ye_diag_of[r];
                for (int k0 = 0; k0 == 0 || k0 < len; k0 += kEll) {
                    int2 pc[kEll];
#pragma unroll
                    for (int k = 0; k < kEll; ++k) pc[k] = a.yr_ell[(size_t)(k0 + k) * n + r];
                    c64 yv[kEll], vv[kEll];
#pragma unroll
                    for (int k = 0; k < kEll; ++k) {
                        yv[k] = Yell ? edits(ld_c(Yell, (k0 + k) * n + r), pc[k].x, dp) : Yat(max(pc[k].x, 0), dp);
                        vv[k] = ldV(pc[k].y);
                    }
#pragma unroll
                    for (int k = 0; k < kEll; ++k)
                        if (pc[k].x >= 0) acc = cadd(acc, cmul(yv[k], vv[k]));
                }
                if (ed >= 0) {
                    Ib[2 * ed] = acc.re;
                    Ib[2 * ed + 1] = acc.im;
                }
                c64 mis = csub(cmul(ldV(r), cconj(acc)), sr);
                if (!a.single_slack || a.fdpf) mis.re += sa * wr;
                if (a.fdpf) {  // P and Q of the fast-decoupled method: the mismatch divided by |V|
                    const c64 vr = ldV(r);
                    const double ivm = 1.0 / sqrt(vr.re * vr.re + vr.im * vr.im);
                    mis.re *= ivm;
                    mis.im *= ivm;
                }
                if (fp >= 0) { XS(fp, (FT)mis.re); fm = fmax(fm, inf_if_nonfinite(fabs(mis.re))); }
                if (fq >= 0) { XS(fq, (FT)mis.im); fm = fmax(fm, inf_if_nonfinite(fabs(mis.im))); }
            }
        }
        return inst_reduce<G, NT, true>(fm, s_part);
    };

    double fmx = mismatch();
    if (s == 0 && run && !s_done[g] && !a.dc) {  // DC: always one linear solve
        s_flast[g] = fmx;
        if (fmx < a.tol) s_done[g] = 1;  // converged at iteration 0
        else if (a.max_iter <= 0) { s_status[g] = 2; s_done[g] = 1; }
    }
    __syncthreads();
    tmark(5);

    for (;;) {
        if (tid == 0) {
            int nd = 0;
            for (int k = 0; k < G; ++k) nd += s_done[k];
            *s_ndone = nd;
        }
        __syncthreads();
        if (*s_ndone == G) break;
        bool act = run && !s_done[g];
        if (s == 0 && act) s_it[g] += 1;

        // ---- J fill: zero the tile fill-in slots, write every J entry ------------------
        // (dS_dVa / dS_dVm, BaseNRAlgo.tpp:202-260, per Ybus entry: one set of loads feeds the
        // up to 4 J entries built on it; V / |V| of the column bus computed on the spot)
        if (act) {
            auto put = [&](int slot, int jidx, double jv) {
                if (slot < 0) return;
                if (a.Jx_last != nullptr && row == a.jx_row) a.Jx_last[jidx] = jv;
                FT v = (FT)jv;
                if (slot & (1 << 30)) {  // pivot with nothing to eliminate
                    if (bad_pivot(v)) s_piv[g] = 1;
                    v = (FT)1 / v;
                }
                XS(slot & ((1 << 30) - 1), v);
            };
            for (int i = s; i < a.nzero; i += nslot) XS(a.zero_slots[i], (FT)0);
            for (int i = s; i < a.njc; i += nslot) put(a.jc_slot[i], a.jc_jidx[i], a.jc_val[i]);
            // kJE entries per lane at a time: the table loads of all of them, then their Y, V,
            // V/|V| and Ibus loads, then the arithmetic (two memory round trips per batch)
            constexpr int kJE = 2;
            EditPos dp;
            edit_pos(dp);
            for (int e0 = s; e0 < a.nye; e0 += kJE * nslot) {
                int pos[kJE];
                int2 rc[kJE];
                int4 sl[kJE];
                c64 y[kJE], vr[kJE], vc[kJE], vnc[kJE], ib[kJE];
                // loads that need nothing loaded first: the table entries, the gathered Ybus values
                // and the Ibus of diagonal entries (they come first)
#pragma unroll
                for (int u = 0; u < kJE; ++u) {
                    const int e = min(e0 + u * nslot, a.nye - 1);
                    pos[u] = a.ye_pos[e];
                    rc[u] = a.ye_rc[e];
                    sl[u] = a.ye_slot[e];
                    if (Yye) y[u] = ld_c(Yye, e);
                    ib[u] = e0 + u * nslot < a.nyd ? ld_c(Ib, e) : c64{0.0, 0.0};
                }
#pragma unroll
                for (int u = 0; u < kJE; ++u) {
                    y[u] = Yye ? edits(y[u], pos[u], dp) : Yat(pos[u], dp);
                    vr[u] = ldV(rc[u].x);
                    vc[u] = ldV(rc[u].y);
                    const double ic = 1.0 / sqrt(vc[u].re * vc[u].re + vc[u].im * vc[u].im);
                    vnc[u] = {vc[u].re * ic, vc[u].im * ic};
                }
#pragma unroll
                for (int u = 0; u < kJE; ++u) {
                    const int e = e0 + u * nslot;
                    if (e >= a.nye) break;
                    if (a.dc) {  // DC: the J entry of a P row and an angle column is Re(Y)
                        put(sl[u].x, (a.Jx_last != nullptr && row == a.jx_row) ? a.ye_jidx[e].x : 0, y[u].re);
                        continue;
                    }
                    const bool diag = e < a.nyd;
                    c64 t = cmul(y[u], vc[u]);
                    c64 dvm = cmul(cconj(cmul(y[u], vnc[u])), vr[u]);
                    if (diag) {
                        t = csub(t, ib[u]);
                        dvm = cadd(dvm, cmul(cconj(ib[u]), vnc[u]));
                    }
                    const c64 dva = cmul(cconj({-t.re, -t.im}), {-vr[u].im, vr[u].re});  // * (j V_r)
                    int4 jx = {0, 0, 0, 0};
                    if (a.Jx_last != nullptr && row == a.jx_row) jx = a.ye_jidx[e];
                    put(sl[u].x, jx.x, dva.re);
                    put(sl[u].y, jx.y, dva.im);
                    put(sl[u].z, jx.z, dvm.re);
                    put(sl[u].w, jx.w, dvm.im);
                }
            }
        }
        __syncthreads();
        tmark(0);

        // ---- factor + forward solve --------------------------------------------
        // General walk (launches that spill): the kSC-item groups of all levels form one
        // sequence, walked with three register buffers in turn: while one group is worked on,
        // the next two (possibly of the next level, across the barrier) are loading. The
        // operands of a group are read before its prefetch is issued, so waiting for them
        // leaves the prefetches in flight. The lean walk is described at its own code.
        {
            bool bad = false;   // a zero or non-finite pivot met by this lane
            bool grow = false;  // a pivot that lost more than growth_tau of its value to cancellation
            const FT gtau = (FT)a.growth_tau;
            if constexpr (LEAN) {
            // Level table word 0: K | Lc << 16. A lane's K items of a level form K / Lc
            // chunks, and every task ends at the last item of a chunk, at item indices that are
            // the same for all lanes: the items are summed with no test, and the team reduction
            // and the leader's update run once per chunk. The first kHB items of a level are
            // loaded during the level before it (two register buffers, by level parity); the
            // rest of a longer level is loaded in the level. Every lane issues every load: items
            // past a level's end read the padding item 0, lanes of a wave idle at a level read a
            // valid item of it they do not use.
            // Level sizes come from scalar loads of the table in memory, and item addresses are
            // a scalar base plus the lane (lanes past U read items they do not use; the stream
            // ends with nslot padding items): the loop computes no address in a vector register,
            // so no prefetch in flight is waited for before its level.
            auto lvk = [&](int l, int w) -> int { return l < a.nlev ? lv_s[2 * l + w] : 0; };
            const int wave_s0 = (threadIdx.x & ~63) / G;
            auto load_head = [&](int base, int K, int U, int4 (&q)[kHB]) {
#pragma unroll
                for (int u = 0; u < kHB; ++u) q[u] = a.fs[(size_t)base + (size_t)min(u, K - 1) * U + s];
            };
            auto ld_item = [&](int base, int K, int U, int k) { return a.fs[(size_t)base + (size_t)min(k, K - 1) * U + s]; };
            FT acc = (FT)0;
            // end of a chunk: the team's lanes reduce, the leader updates the task's slot
            auto chunk_end = [&](const int e, const int T) {
                const int lt = (e >> 26) & 7;
                const FT sum = team_sum<G>(acc, lt, T);
                if ((e & kEnd) && (s & ((1 << lt) - 1)) == 0) {
                    const int off = e & ((1 << 25) - 1);
                    const FT a0 = ((e >> 25) & 1) ? (FT)0 : XB(off);  // fill-in: assign
                    FT v = a0 - sum;
                    if ((e >> 29) & 1) {
                        bad |= bad_pivot(v);
                        grow |= fabs(v) < gtau * fabs(a0);
                        v = pivot_recip(v);
                    }
                    XBset(off, v);
                }
                acc = (FT)0;
            };
            auto prod = [&](const int4 q) -> FT { return XB(q.x) * XB(q.y) * XB(q.z); };
            auto level = [&](int base, int K, int Lc, int U, int T, const int4 (&h)[kHB]) {
                if (wave_s0 < U) {  // waves with no lane in use at this level only keep the barrier
                    FT p[kHB];
#pragma unroll
                    for (int u = 0; u < kHB; ++u) p[u] = prod(h[u]);
                    int kend = Lc - 1;
#pragma unroll
                    for (int u = 0; u < kHB; ++u)
                        if (u < K) {
                            acc += p[u];
                            if (u == kend) {
                                chunk_end(h[u].w, T);
                                kend += Lc;
                            }
                        }
                    // the items past the head, two at a time; the loads of the next two are
                    // issued before the current two are used (two register pairs in turn)
                    auto step = [&](int k, const int4 r0, const int4 r1) {
                        const FT p0 = prod(r0), p1 = prod(r1);
                        acc += p0;
                        if (k == kend) {
                            chunk_end(r0.w, T);
                            kend += Lc;
                        }
                        acc += k + 1 < K ? p1 : (FT)0;  // r1 repeats item K - 1 past the level's end
                        if (k + 1 == kend) {
                            chunk_end(r1.w, T);
                            kend += Lc;
                        }
                    };
                    if (K > kHB) {
                        int4 A0 = ld_item(base, K, U, kHB), A1 = ld_item(base, K, U, kHB + 1);
                        for (int k = kHB; k < K; k += 4) {
                            const int4 B0 = ld_item(base, K, U, k + 2), B1 = ld_item(base, K, U, k + 3);
                            step(k, A0, A1);
                            if (k + 2 >= K) break;
                            A0 = ld_item(base, K, U, k + 4);
                            A1 = ld_item(base, K, U, k + 5);
                            step(k + 2, B0, B1);
                        }
                    }
                }
                __syncthreads();
            };
            int4 H0[kHB], H1[kHB];
            int base = 1;  // level 0 follows the leading padding item
            int w0 = lvk(0, 0), w1 = lvk(0, 1);
            load_head(base, w0 & 0xffff, w1 & 0xffffff, H0);
            for (int l = 0; l < a.nlev; l += 2) {
                int nb = base + (w0 & 0xffff) * (w1 & 0xffffff);
                int n0 = lvk(l + 1, 0), n1 = lvk(l + 1, 1);
                load_head(nb, n0 & 0xffff, n1 & 0xffffff, H1);
                level(base, w0 & 0xffff, w0 >> 16, w1 & 0xffffff, w1 >> 24, H0);
                if (l + 1 >= a.nlev) break;
                base = nb;
                w0 = n0;
                w1 = n1;
                nb = base + (w0 & 0xffff) * (w1 & 0xffffff);
                n0 = lvk(l + 2, 0);
                n1 = lvk(l + 2, 1);
                load_head(nb, n0 & 0xffff, n1 & 0xffffff, H0);
                level(base, w0 & 0xffff, w0 >> 16, w1 & 0xffffff, w1 >> 24, H1);
                base = nb;
                w0 = n0;
                w1 = n1;
            }
            } else {
            // level sizes of the load cursor (lK, lU) and of the process cursor (pK, pU, pT:
            // log2 of its largest team), uniform
            auto lvk = [&](int l, int w) -> int { return l < a.nlev ? lv_s[2 * l + w] : 0; };
            int ll = 0, lk = 0, lbase = 0;   // load cursor: level, item, level base
            int lK = lvk(0, 0), lU = lvk(0, 1) & 0xffffff;
            bool more = true;                 // the lane has items left in the load cursor's level
            // Every lane issues every prefetch load (a lane with nothing to load reads item 0
            // as a stand-in) and keeps a validity bit per item, applied when the item is used:
            // a load skipped by some lanes would make the compiler wait for all loads in flight.
            auto prefetch = [&](int4 (&q)[kSC], int& vbits) {
                const bool on = act && more && s < lU;
                vbits = 0;
#pragma unroll
                for (int u = 0; u < kSC; ++u) {
                    const bool ok = on && lk + u < lK;
                    q[u] = a.fs[ok ? (size_t)lbase + (size_t)(lk + u) * lU + s : (size_t)0];
                    vbits |= (int)ok << u;
                }
                lk += kSC;
                if (lk >= lK) {
                    lbase += lK * lU;
                    ++ll;
                    lk = 0;
                    lK = lvk(ll, 0);
                    lU = lvk(ll, 1) & 0xffffff;
                    more = true;
                }
            };
            int pl = 0, pk = 0, pK = lK, pU = lU, pT = lvk(0, 1) >> 24;  // process cursor
            const int wave_s0 = (threadIdx.x & ~63) / G;  // first instance slot of this wave
            FT acc = (FT)0;
            auto process = [&](const int4 (&cur_raw)[kSC], int cur_v, int4 (&nxt)[kSC], int& nxt_v) {
                int4 cur[kSC];
#pragma unroll
                for (int u = 0; u < kSC; ++u) cur[u] = (cur_v >> u) & 1 ? cur_raw[u] : no_item();
                FT prod[kSC];
                bool last = false;
                // a wave none of whose lanes is used at this level only keeps the cursors and the barrier
                const bool wave_on = wave_s0 < pU;
#pragma unroll
                for (int u = 0; u < kSC; ++u) {
                    if (!wave_on) break;
                    const int4 q = cur[u];
                    prod[u] = (FT)0;
                    if (max(max((unsigned)q.x, (unsigned)q.y), (unsigned)q.z) < (unsigned)nlds) prod[u] = Xl[q.x * G + g] * Xl[q.y * G + g] * Xl[q.z * G + g];  // all operands in LDS
                    else if (q.x >= 0) prod[u] = XL(q.x) * XL(q.y) * XL(q.z);
                    last |= q.w < 0;
                }
                // the lane's last item: later chunks of the load cursor's level are skipped
                if (last && pl == ll) more = false;
                prefetch(nxt, nxt_v);
                if (act && wave_on) {
#pragma unroll
                    for (int u = 0; u < kSC; ++u) {
                        const int e = cur[u].w;
                        acc += prod[u];
                        // end of a task: the team's lanes reduce, the leader writes
                        const bool end = e & kEnd;
                        if (end) {
                        const int lt = (e >> 26) & 7;
                        const FT sum = team_sum<G>(acc, lt, pT);
                        const bool lead = (s & ((1 << lt) - 1)) == 0;
                            if (lead) {
                                const int slot = e & ((1 << 25) - 1);
                                const bool rc = (e >> 29) & 1;
                                // one explicit read-modify-write per address space (a merged generic
                                // pointer for this slot trips the gfx950 backend)
                                const bool as = (e >> 25) & 1;  // fill-in: the slot's old content is not ours
                                if (slot < nlds) {
                                    const FT a0 = as ? (FT)0 : Xl[slot * G + g];
                                    FT v = a0 - sum;
                                    if (rc) { bad |= bad_pivot(v); grow |= fabs(v) < gtau * fabs(a0); v = pivot_recip(v); }
                                    Xl[slot * G + g] = v;
                                } else {
                                    const FT a0 = as ? (FT)0 : Xgb[(size_t)(slot - nlds) * G + g];
                                    FT v = a0 - sum;
                                    if (rc) { bad |= bad_pivot(v); grow |= fabs(v) < gtau * fabs(a0); v = pivot_recip(v); }
                                    Xgb[(size_t)(slot - nlds) * G + g] = v;
                                }
                            }
                        }
                        acc = end ? (FT)0 : acc;
                    }
                }
                pk += kSC;
                if (pk >= pK) {  // the level's last chunk: its writes are visible to the next level
                    __syncthreads();
                    ++pl;
                    pk = 0;
                    pK = lvk(pl, 0);
                    const int w1 = lvk(pl, 1);
                    pU = w1 & 0xffffff;
                    pT = w1 >> 24;
                }
            };
            int4 A[kSC], B[kSC], C[kSC];
            int Av = 0, Bv = 0, Cv = 0;
            prefetch(A, Av);
            prefetch(B, Bv);
            while (pl < a.nlev) {
                process(A, Av, C, Cv);
                if (pl >= a.nlev) break;
                process(B, Bv, A, Av);
                if (pl >= a.nlev) break;
                process(C, Cv, B, Bv);
            }
            }
            if (bad) s_piv[g] = 1;
            if (grow) s_grow[g] = 1;
        }
        tmark(1);
        // ---- dense tail: the m x m tile of every instance ---------------------------------
        // Tile at slots [0, m*(m+1)): row-major with row stride m + 1, the right-hand side
        // held as column m. Right-looking LU with diagonal pivots, four columns per pass (one
        // barrier per pass): pass k eliminates columns k .. k+c-1 (c = min(4, m - k)) from the
        // trailing block (rows >= k+c, columns >= k+c, the vector included) with the rank-c update
        //   T(i,j) -= sum_q L(i,k+q) U(k+q,j),
        //   U(k+q,j) = T(k+q,j) - sum_{r<q} L(k+q,k+r) U(k+r,j),
        //   L(i,k+q) = (T(i,k+q) - sum_{r<q} L(i,k+r) U(k+r,k+q)) / d_k+q,
        // every lane computing the c x c diagonal block's factors and the L and U factors of its
        // entry from the tile. Rows k+1 .. k+c-1 of U and the c reciprocal pivots are stored one
        // pass later (so that no pass writes what it reads); L itself is not stored, nothing
        // after the forward sweep reads it.
        // Then the backward sweep on the first wave alone (its lanes hold s < 64 / G of
        // every instance; LDS accesses of one wave complete in issue order, so its passes
        // need no barrier): x_k = y_k / d_k, y_i -= U(i,k) x_k for i < k; with one instance per
        // workgroup, in registers.
        if (a.m > 0) {
            const int m = a.m, ld = m + 1;
            auto T = [&](int i, int j) -> int { return (i * ld + j) * G + g; };  // LDS index
            // the previous pass's block: its columns, reciprocal pivots and L factors
            int pc = 0;
            FT pr0 = (FT)0, pr1 = (FT)0, pr2 = (FT)0, pr3 = (FT)0;
            FT pl10 = (FT)0, pl20 = (FT)0, pl21 = (FT)0, pl30 = (FT)0, pl31 = (FT)0, pl32 = (FT)0;
            bool bad = false;
            for (int k = 0;; k += 4) {
                if (act) {
                    const int c = min(4, m - k);  // columns of this pass's block (<= 0: none)
                    FT r0 = (FT)0, r1 = (FT)0, r2 = (FT)0, r3 = (FT)0;
                    FT l10 = (FT)0, l20 = (FT)0, l21 = (FT)0, l30 = (FT)0, l31 = (FT)0, l32 = (FT)0;
                    if (c > 0) {
                        // the block's LU, in the same order in every lane
                        FT u01 = (FT)0, u02 = (FT)0, u03 = (FT)0, u12 = (FT)0, u13 = (FT)0, u23 = (FT)0;
                        const FT d0 = Xl[T(k, k)];
                        bad |= bad_pivot(d0);
                        r0 = pivot_recip(d0);
                        if (c > 1) {
                            u01 = Xl[T(k, k + 1)];
                            l10 = Xl[T(k + 1, k)] * r0;
                            const FT d1 = Xl[T(k + 1, k + 1)] - l10 * u01;
                            bad |= bad_pivot(d1);
                            r1 = pivot_recip(d1);
                        }
                        if (c > 2) {
                            u02 = Xl[T(k, k + 2)];
                            u12 = Xl[T(k + 1, k + 2)] - l10 * u02;
                            l20 = Xl[T(k + 2, k)] * r0;
                            l21 = (Xl[T(k + 2, k + 1)] - l20 * u01) * r1;
                            const FT d2 = Xl[T(k + 2, k + 2)] - l20 * u02 - l21 * u12;
                            bad |= bad_pivot(d2);
                            r2 = pivot_recip(d2);
                        }
                        if (c > 3) {
                            u03 = Xl[T(k, k + 3)];
                            u13 = Xl[T(k + 1, k + 3)] - l10 * u03;
                            u23 = Xl[T(k + 2, k + 3)] - l20 * u03 - l21 * u13;
                            l30 = Xl[T(k + 3, k)] * r0;
                            l31 = (Xl[T(k + 3, k + 1)] - l30 * u01) * r1;
                            l32 = (Xl[T(k + 3, k + 2)] - l30 * u02 - l31 * u12) * r2;
                            const FT d3 = Xl[T(k + 3, k + 3)] - l30 * u03 - l31 * u13 - l32 * u23;
                            bad |= bad_pivot(d3);
                            r3 = pivot_recip(d3);
                        }
                        // entries (i, j), i in [k+c, m), j in [k+c, m]: w rows of w + 1
                        const int w = m - k - c, wc = w + 1;
                        // row = floor((idx + 1/2) / wc): the fractional part of the exact
                        // quotient stays >= 1/(2 wc) from an integer, far above the float error
                        const float inv = __builtin_amdgcn_rcpf((float)wc);
                        for (int idx = s; idx < w * wc; idx += nslot) {
                            const int rr = (int)(((float)idx + 0.5f) * inv);
                            const int i = k + c + rr, j = k + c + (idx - rr * wc);
                            const FT li0 = Xl[T(i, k)] * r0;
                            const FT u0j = Xl[T(k, j)];
                            FT upd = li0 * u0j;
                            if (c > 1) {
                                const FT li1 = (Xl[T(i, k + 1)] - li0 * u01) * r1;
                                const FT u1j = Xl[T(k + 1, j)] - l10 * u0j;
                                upd += li1 * u1j;
                                if (c > 2) {
                                    const FT li2 = (Xl[T(i, k + 2)] - li0 * u02 - li1 * u12) * r2;
                                    const FT u2j = Xl[T(k + 2, j)] - l20 * u0j - l21 * u1j;
                                    upd += li2 * u2j;
                                    if (c > 3) {
                                        const FT li3 = (Xl[T(i, k + 3)] - li0 * u03 - li1 * u13 - li2 * u23) * r3;
                                        const FT u3j = Xl[T(k + 3, j)] - l30 * u0j - l31 * u1j - l32 * u2j;
                                        upd += li3 * u3j;
                                    }
                                }
                            }
                            Xl[T(i, j)] -= upd;
                        }
                    }
                    if (k >= 4 && pc > 1) {  // rows p+1 .. p+pc-1 of U of the previous pass's block
                        const int p = k - 4;
                        for (int j = p + 2 + s; j <= m; j += nslot) {
                            const FT u0j = Xl[T(p, j)];
                            const FT u1j = Xl[T(p + 1, j)] - pl10 * u0j;
                            Xl[T(p + 1, j)] = u1j;
                            if (pc > 2 && j > p + 2) {
                                const FT u2j = Xl[T(p + 2, j)] - pl20 * u0j - pl21 * u1j;
                                Xl[T(p + 2, j)] = u2j;
                                if (pc > 3 && j > p + 3)
                                    Xl[T(p + 3, j)] = Xl[T(p + 3, j)] - pl30 * u0j - pl31 * u1j - pl32 * u2j;
                            }
                        }
                    }
                    if (k >= 4 && pc > 0 && s == 0) {  // and its reciprocal pivots
                        const int p = k - 4;
                        Xl[T(p, p)] = pr0;
                        if (pc > 1) Xl[T(p + 1, p + 1)] = pr1;
                        if (pc > 2) Xl[T(p + 2, p + 2)] = pr2;
                        if (pc > 3) Xl[T(p + 3, p + 3)] = pr3;
                    }
                    pc = max(c, 0);
                    pr0 = r0; pr1 = r1; pr2 = r2; pr3 = r3;
                    pl10 = l10; pl20 = l20; pl21 = l21; pl30 = l30; pl31 = l31; pl32 = l32;
                }
                __syncthreads();
                if (k >= m) break;
            }
            if (bad && s == 0) s_piv[g] = 1;
            if (G == 1 && m <= 128) {
                // One instance per workgroup: lane s of the first wave holds y_s and y_s+64 in
                // registers, and x_k = y_k / d_k reaches every lane by a lane read of its owner,
                // so a step is a few dependent register operations. U(i, k) and 1/d_k are not
                // written by the sweep: each step's reads are issued one step ahead.
                if (threadIdx.x < 64 && act) {
                    FT y0 = s < m ? Xl[T(s, m)] : (FT)0, y1 = s + 64 < m ? Xl[T(s + 64, m)] : (FT)0;
                    // column k's reads, clamped into the tile (a lane's U entries of rows i >= k
                    // are read but not used)
                    auto col = [&](int k, FT& u0, FT& u1, FT& dk) {
                        const int c = max(k, 0);
                        dk = Xl[T(c, c)];
                        u0 = Xl[T(min(s, m - 1), c)];
                        u1 = Xl[T(min(s + 64, m - 1), c)];
                    };
                    FT u0, u1, dk;
                    col(m - 1, u0, u1, dk);
                    for (int k = m - 1; k >= 0; --k) {
                        FT nu0, nu1, ndk;
                        col(k - 1, nu0, nu1, ndk);
                        const FT xk = lane_read(k < 64 ? y0 : y1, k & 63) * dk;
                        if (s == (k & 63)) {
                            if (k < 64) y0 = xk;
                            else y1 = xk;
                        }
                        if (s < k) y0 -= u0 * xk;  // rows i < k only: a non-finite x_k leaves the others
                        if (s + 64 < k) y1 -= u1 * xk;
                        u0 = nu0; u1 = nu1; dk = ndk;
                    }
                    if (s < m) Xl[T(s, m)] = y0;
                    if (s + 64 < m) Xl[T(s + 64, m)] = y1;
                }
            } else if (threadIdx.x < 64 && act) {
                constexpr int ws = 64 / G;  // lanes of one instance in the first wave
                for (int k = m - 1; k >= 0; --k) {
                    const FT xk = Xl[T(k, m)] * Xl[T(k, k)];
                    for (int i = s; i < k; i += ws) Xl[T(i, m)] -= Xl[T(i, k)] * xk;
                    if (s == 0) Xl[T(k, m)] = xk;
                    // keep this pass's LDS writes ahead of the next pass's reads in the code
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
            __syncthreads();
        }
        tmark(2);
        // zero / non-finite pivot: klu_factor failure on the first factorisation of a fresh
        // solver (KLUSolver.cpp:38-44), klu_refactor failure afterwards (:50-57)
        if (s == 0 && act && s_piv[g]) {
            s_status[g] = (a.fresh && s_it[g] == 1) ? 5 : 6;  // (FDPF: the first half-step)
            s_done[g] = 1;
        }
        __syncthreads();
        act = run && !s_done[g];

        // ---- backward solve (same pipelined walk as the factor phase) -----------------
        {
            if constexpr (LEAN) {
            // chunked items as in the factor walk; x = (x - sum u x') / d at a chunk's end
            auto lvk = [&](int l, int w) -> int { return l < a.nblev ? lv_s[2 * (a.nlev + l) + w] : 0; };
            const int wave_s0 = (threadIdx.x & ~63) / G;
            auto load_head = [&](int base, int K, int U, int4 (&q)[kHB]) {
#pragma unroll
                for (int u = 0; u < kHB; ++u) q[u] = a.bs[(size_t)base + (size_t)min(u, K - 1) * U + s];
            };
            auto ld_item = [&](int base, int K, int U, int k) { return a.bs[(size_t)base + (size_t)min(k, K - 1) * U + s]; };
            FT acc = (FT)0;
            auto chunk_end = [&](const int4 q, const int T) {
                const int e = q.z;
                const int lt = (e >> 26) & 7;
                const FT sum = team_sum<G>(acc, lt, T);
                if ((e & kEnd) && (s & ((1 << lt) - 1)) == 0) {
                    const int off = e & ((1 << 25) - 1);
                    XBset(off, (XB(off) - sum) * XB(q.w));
                }
                acc = (FT)0;
            };
            auto prod = [&](const int4 q) -> FT { return XB(q.x) * XB(q.y); };
            auto level = [&](int base, int K, int Lc, int U, int T, const int4 (&h)[kHB]) {
                if (wave_s0 < U) {  // waves with no lane in use at this level only keep the barrier
                    FT p[kHB];
#pragma unroll
                    for (int u = 0; u < kHB; ++u) p[u] = prod(h[u]);
                    int kend = Lc - 1;
#pragma unroll
                    for (int u = 0; u < kHB; ++u)
                        if (u < K) {
                            acc += p[u];
                            if (u == kend) {
                                chunk_end(h[u], T);
                                kend += Lc;
                            }
                        }
                    // the items past the head, two at a time; the loads of the next two are
                    // issued before the current two are used (two register pairs in turn)
                    auto step = [&](int k, const int4 r0, const int4 r1) {
                        const FT p0 = prod(r0), p1 = prod(r1);
                        acc += p0;
                        if (k == kend) {
                            chunk_end(r0, T);
                            kend += Lc;
                        }
                        acc += k + 1 < K ? p1 : (FT)0;  // r1 repeats item K - 1 past the level's end
                        if (k + 1 == kend) {
                            chunk_end(r1, T);
                            kend += Lc;
                        }
                    };
                    if (K > kHB) {
                        int4 A0 = ld_item(base, K, U, kHB), A1 = ld_item(base, K, U, kHB + 1);
                        for (int k = kHB; k < K; k += 4) {
                            const int4 B0 = ld_item(base, K, U, k + 2), B1 = ld_item(base, K, U, k + 3);
                            step(k, A0, A1);
                            if (k + 2 >= K) break;
                            A0 = ld_item(base, K, U, k + 4);
                            A1 = ld_item(base, K, U, k + 5);
                            step(k + 2, B0, B1);
                        }
                    }
                }
                __syncthreads();
            };
            int4 H0[kHB], H1[kHB];
            int base = 1;  // level 0 follows the leading padding item
            int w0 = lvk(0, 0), w1 = lvk(0, 1);
            load_head(base, w0 & 0xffff, w1 & 0xffffff, H0);
            for (int l = 0; l < a.nblev; l += 2) {
                int nb = base + (w0 & 0xffff) * (w1 & 0xffffff);
                int n0 = lvk(l + 1, 0), n1 = lvk(l + 1, 1);
                load_head(nb, n0 & 0xffff, n1 & 0xffffff, H1);
                level(base, w0 & 0xffff, w0 >> 16, w1 & 0xffffff, w1 >> 24, H0);
                if (l + 1 >= a.nblev) break;
                base = nb;
                w0 = n0;
                w1 = n1;
                nb = base + (w0 & 0xffff) * (w1 & 0xffffff);
                n0 = lvk(l + 2, 0);
                n1 = lvk(l + 2, 1);
                load_head(nb, n0 & 0xffff, n1 & 0xffffff, H0);
                level(base, w0 & 0xffff, w0 >> 16, w1 & 0xffffff, w1 >> 24, H1);
                base = nb;
                w0 = n0;
                w1 = n1;
            }
            } else {
            // level sizes of the load cursor (lK, lU) and of the process cursor (pK, pU, pT:
            // log2 of its largest team), uniform
            auto lvk = [&](int l, int w) -> int { return l < a.nblev ? lv_s[2 * (a.nlev + l) + w] : 0; };
            int ll = 0, lk = 0, lbase = 0;   // load cursor: level, item, level base
            int lK = lvk(0, 0), lU = lvk(0, 1) & 0xffffff;
            bool more = true;                 // the lane has items left in the load cursor's level
            // Every lane issues every prefetch load (a lane with nothing to load reads item 0
            // as a stand-in) and keeps a validity bit per item, applied when the item is used:
            // a load skipped by some lanes would make the compiler wait for all loads in flight.
            auto prefetch = [&](int4 (&q)[kSC], int& vbits) {
                const bool on = act && more && s < lU;
                vbits = 0;
#pragma unroll
                for (int u = 0; u < kSC; ++u) {
                    const bool ok = on && lk + u < lK;
                    q[u] = a.bs[ok ? (size_t)lbase + (size_t)(lk + u) * lU + s : (size_t)0];
                    vbits |= (int)ok << u;
                }
                lk += kSC;
                if (lk >= lK) {
                    lbase += lK * lU;
                    ++ll;
                    lk = 0;
                    lK = lvk(ll, 0);
                    lU = lvk(ll, 1) & 0xffffff;
                    more = true;
                }
            };
            int pl = 0, pk = 0, pK = lK, pU = lU, pT = lvk(0, 1) >> 24;  // process cursor
            const int wave_s0 = (threadIdx.x & ~63) / G;  // first instance slot of this wave
            FT acc = (FT)0;
            // U entries and pivots are final once the factor phase is done (no value starts in the
            // backward phase, so no recycled slot changes): the U operands of the next group and
            // the 1/d of its tasks are read while the current group is worked on, and a spilled
            // one's memory latency overlaps that work instead of following its item's arrival.
            auto operands = [&](const int4 (&raw)[kSC], int v, FT (&pu)[kSC], FT (&pd)[kSC]) {
#pragma unroll
                for (int u = 0; u < kSC; ++u) {
                    const int4 q = (v >> u) & 1 ? raw[u] : no_item();
                    const FT xu = XLo(max(q.x, 0)), xd = XLo(max(q.w, 0));
                    pu[u] = q.x >= 0 ? xu : (FT)0;
                    pd[u] = (q.z & kEnd) ? xd : (FT)0;
                }
            };
            auto process = [&](const int4 (&cur_raw)[kSC], int cur_v, const FT (&pu)[kSC], const FT (&pd)[kSC],
                               int4 (&nxt)[kSC], int& nxt_v, const int4 (&pend)[kSC], int pend_v, FT (&pend_u)[kSC],
                               FT (&pend_d)[kSC]) {
                int4 cur[kSC];
#pragma unroll
                for (int u = 0; u < kSC; ++u) cur[u] = (cur_v >> u) & 1 ? cur_raw[u] : no_item();
                FT prod[kSC];
                bool last = false;
                // a wave none of whose lanes is used at this level only keeps the cursors and the barrier
                const bool wave_on = wave_s0 < pU;
#pragma unroll
                for (int u = 0; u < kSC; ++u) {
                    if (!wave_on) break;
                    const int4 q = cur[u];
                    FT xv = (FT)0;  // x' is written by earlier backward levels: read after their barriers
                    if ((unsigned)q.y < (unsigned)nlds) xv = Xl[q.y * G + g];
                    else if (q.y >= 0) xv = Xgb[(size_t)(q.y - nlds) * G + g];
                    prod[u] = pu[u] * xv;
                    last |= q.z < 0;
                }
                // the lane's last item: later chunks of the load cursor's level are skipped
                if (last && pl == ll) more = false;
                prefetch(nxt, nxt_v);
                operands(pend, pend_v, pend_u, pend_d);
                if (act && wave_on) {
#pragma unroll
                    for (int u = 0; u < kSC; ++u) {
                        const int e = cur[u].z;
                        acc += prod[u];
                        // end of a task: the team's lanes reduce, the leader writes
                        const bool end = e & kEnd;
                        if (end) {
                        const int lt = (e >> 26) & 7;
                        const FT sum = team_sum<G>(acc, lt, pT);
                        const bool lead = (s & ((1 << lt) - 1)) == 0;
                            if (lead) {
                                const int slot = e & ((1 << 25) - 1);
                                const FT rd = pd[u];
                                if (slot < nlds) {
                                    Xl[slot * G + g] = (Xl[slot * G + g] - sum) * rd;
                                } else {
                                    FT* xp = &Xgb[(size_t)(slot - nlds) * G + g];
                                    *xp = (*xp - sum) * rd;
                                }
                            }
                        }
                        acc = end ? (FT)0 : acc;
                    }
                }
                pk += kSC;
                if (pk >= pK) {  // the level's last chunk: its writes are visible to the next level
                    __syncthreads();
                    ++pl;
                    pk = 0;
                    pK = lvk(pl, 0);
                    const int w1 = lvk(pl, 1);
                    pU = w1 & 0xffffff;
                    pT = w1 >> 24;
                }
            };
            int4 A[kSC], B[kSC], C[kSC];
            int Av = 0, Bv = 0, Cv = 0;
            FT Au[kSC], Ad[kSC], Bu[kSC], Bd[kSC], Cu[kSC], Cd[kSC];
            prefetch(A, Av);
            prefetch(B, Bv);
            operands(A, Av, Au, Ad);
            while (pl < a.nblev) {
                process(A, Av, Au, Ad, C, Cv, B, Bv, Bu, Bd);
                if (pl >= a.nblev) break;
                process(B, Bv, Bu, Bd, A, Av, C, Cv, Cu, Cd);
                if (pl >= a.nblev) break;
                process(C, Cv, Cu, Cd, B, Bv, A, Av, Au, Ad);
            }
            }
        }

        tmark(3);
        // ---- voltage update ------------------------------------------------------------
        // per bus, by the lane that owns it: its Va and Vm steps are read from their slots
        double vmin = INFINITY;
        if (act) {
            if (s == 0 && a.slack_dx >= 0) s_slack[g] -= (double)XL(a.slack_dx);
            for (int b = s; b < n; b += nslot) {
                const int2 us = a.bus_dx[b];
                // FDPF: the P half-step (odd s_it) moves the angles only, the Q half the magnitudes
                const bool pstep = !a.fdpf || (s_it[g] & 1) != 0, qstep = !a.fdpf || (s_it[g] & 1) == 0;
                const double dva = (us.x >= 0 && pstep) ? (double)XL(us.x) : 0.0;
                const double dvm = (us.y >= 0 && qstep) ? (double)XL(us.y) : 0.0;
                double vm;
                if constexpr (LEAN) vm = update_bus(Vl, G, Va, Vm, b, dva, dvm, a.single_slack != 0);
                else vm = update_bus(V, 1, Va, Vm, b, dva, dvm, a.single_slack != 0);
                vmin = fmin(vmin, vm);
            }
        }
        vmin = inst_reduce<G, NT, false>(vmin, s_part);
        if (act && vmin < 0.0) {
            if constexpr (LEAN) polar_of(Vl, G, Va, Vm, s, n, nslot);
            else polar_of(V, 1, Va, Vm, s, n, nslot);
        }
        __syncthreads();
        tmark(4);

        // ---- mismatch + convergence ----------------------------------------------------
        fmx = mismatch();
        if (s == 0 && act && a.dc) {
            s_status[g] = fmx < 1e6 ? 0 : 7;  // max |theta| (inf if not finite)
            s_done[g] = 1;
        } else if (s == 0 && act) {
            s_flast[g] = fmx;
            if (!(fmx <= 1.7976931348623157e308)) { s_status[g] = 3; s_done[g] = 1; }
            else if (fmx < a.tol) { s_status[g] = 0; s_done[g] = 1; }
            else if (a.fdpf ? ((s_it[g] & 1) == 0 && s_it[g] / 2 >= a.max_iter) : s_it[g] >= a.max_iter) {
                s_status[g] = 2;
                s_done[g] = 1;
            }
        }
        __syncthreads();
        tmark(5);
    }
    // ---- end of the step: results of this row; a failed chain stops ---------------------
    if (run) {
        if constexpr (LEAN)
            for (int b = s; b < n; b += nslot) {
                V[2 * b] = Vl[2 * b * G];
                V[2 * b + 1] = Vl[2 * b * G + 1];
            }
        if (s == 0) {
            const int stt = s_status[g];
            a.iters[row] = a.fdpf ? (s_it[g] + 1) / 2 : s_it[g];
            a.status[row] = stt;
            if (a.flags)
                a.flags[row] = (stt != 0 && s_flast[g] < a.stall_tol ? kFlagStall : 0) | (s_grow[g] ? kFlagGrowth : 0);
        }
        alive = s_status[g] == 0;
    } else if (valid && row < a.nrows) {  // a step after its chain's failure
        for (int b = s; b < n; b += nslot) {
            double* Vr = a.V + row * 2 * (long long)n;
            Vr[2 * b] = 0.0;
            Vr[2 * b + 1] = 0.0;
        }
        if (s == 0) {
            a.iters[row] = 0;
            a.status[row] = kStatusNotRun;
            if (a.flags) a.flags[row] = 0;
        }
    }
    __syncthreads();  // the next step rewrites the workgroup state
    }  // step
    }  // instance groups
}

// Per-instance initial V (broadcast or per instance).
// Ybus values in the order of the mismatch table (0 past a row's end) and of the J-fill table.
__global__ void k_gather_y(const double* __restrict__ Y, const int2* __restrict__ ell, int nell,
                           const int* __restrict__ ye_pos, int nye, double* __restrict__ Yell, double* __restrict__ Yye) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nell) {
        const int p = ell[i].x;
        Yell[2 * i] = p >= 0 ? Y[2 * p] : 0.0;
        Yell[2 * i + 1] = p >= 0 ? Y[2 * p + 1] : 0.0;
    } else if (i < nell + nye) {
        const int e = i - nell, p = ye_pos[e];
        Yye[2 * e] = Y[2 * p];
        Yye[2 * e + 1] = Y[2 * p + 1];
    }
}

// The first kernel of a solve launch: zeroes the NR kernel's work-queue counter and, unless
// the voltages are solved in place, writes the initial V of every launch instance (its first
// row: the chain head in chain mode) from V0 (one row per instance, or one shared row).
__global__ void k_init_v(const double* V0, long long v0_stride, double* V, int n, int ninst, const int* imap,
                         int chain, int* queue, unsigned long long* phase, int copy) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *queue = 0;
    if (blockIdx.x == 0 && threadIdx.x < 8 && phase) phase[threadIdx.x] = 0;
    if (!copy) return;
    const long long total = (long long)ninst * n;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const long long i = k / n;
        const int b = (int)(k - i * n);
        const long long oi = imap ? (long long)imap[i] : i;
        const long long r = oi * chain;
        V[2 * (r * n + b)] = V0[oi * v0_stride + 2 * b];
        V[2 * (r * n + b) + 1] = V0[oi * v0_stride + 2 * b + 1];
    }
}

// Compaction of the rows to re-solve in FP64 after an FP32-factor launch, in row order, by one
// workgroup of 1024 lanes: rows that did not converge and stalled (their last mismatch is below
// the stall level: the FP32 steps, not the Newton iteration, stopped the progress), or met a zero
// / non-finite pivot or a non-finite mismatch. A row that diverged is not re-solved.
// Chain mode takes whole chains (a chain is re-run from its head). Writes the list and its
// length.
// Lists the units (instances, or chains) whose first failure calls for another pass: mode 0, the
// FP32 form's FP64 re-solve (breakdown or stall); mode 1, the host re-pivot (a pivot failure on the
// first factorisation of a fresh solve, SolverFactor).
__global__ void __launch_bounds__(1024) k_collect_fallback(const int* status, const int* flags, long long nrows,
                                                            int chain, int mode, int* list, int* count) {
    __shared__ int wsum[16];
    __shared__ int base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) base = 0;
    __syncthreads();
    const long long nunits = (nrows + chain - 1) / chain;
    for (long long u0 = 0; u0 < nunits; u0 += 1024) {
        const long long u = u0 + tid;
        bool take = false;
        if (u < nunits)
            for (int k = 0; k < chain; ++k) {
                const long long r = u * chain + k;
                if (r >= nrows) break;
                const int st = status[r];
                if (st == 0 || st == kStatusNotRun) continue;
                take = mode == 0 ? (st == 3 || st == 5 || st == 6 || (flags[r] & kFlagStall) != 0) : st == 5;
                break;  // a chain's first failure decides
            }
        const unsigned long long m = __ballot(take);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; ++w) off += wsum[w];
        if (take) list[off + before] = (int)u;
        __syncthreads();
        if (tid == 0) {
            int t = 0;
            for (int w = 0; w < 16; ++w) t += wsum[w];
            base += t;
        }
        __syncthreads();
    }
    if (tid == 0) *count = base;
}

// Two-pass single chains (see solve_single_chain). Starts of the second pass: row 0 from the
// chain's head, row k from the first pass's row k - 1.
__global__ void k_chain_starts(const double* head, const double* V1, double* V0p, long long rows, int n) {
    const long long total = rows * n;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const long long r = k / n;
        const int b = (int)(k - r * n);
        const double* src = r == 0 ? head + 2 * b : V1 + 2 * ((r - 1) * n + b);
        V0p[2 * k] = src[0];
        V0p[2 * k + 1] = src[1];
    }
}

// agree[r] = 1 when row r converged in both passes to the same voltages (max |V1 - V2| <= eps),
// one workgroup of 256 lanes per row
__global__ void __launch_bounds__(256) k_rows_agree(const double* V1, const double* V2, const int* st1,
                                                    const int* st2, long long rows, int n, double eps, int* agree) {
    __shared__ int bad;
    for (long long r = blockIdx.x; r < rows; r += gridDim.x) {
        if (threadIdx.x == 0) bad = (st1[r] != 0 || st2[r] != 0) ? 1 : 0;
        __syncthreads();
        int mine = 0;
        for (int b = threadIdx.x; b < n; b += blockDim.x) {
            const double dr = V1[2 * (r * n + b)] - V2[2 * (r * n + b)];
            const double di = V1[2 * (r * n + b) + 1] - V2[2 * (r * n + b) + 1];
            mine |= !(dr * dr + di * di <= eps * eps);  // a NaN disagrees
        }
        if (mine) atomicOr(&bad, 1);
        __syncthreads();
        if (threadIdx.x == 0) agree[r] = bad ? 0 : 1;
        __syncthreads();
    }
}

// Rows [r0, r1) of a chain after its first failure: not run, 0 iterations, zero voltages and flags.
__global__ void k_chain_not_run(double* V, int* iters, int* status, int* flags, long long r0, long long r1, int n) {
    const long long total = (r1 - r0) * n;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        V[2 * (r0 * n + k)] = 0.0;
        V[2 * (r0 * n + k) + 1] = 0.0;
        if (k % n == 0) {
            const long long r = r0 + k / n;
            iters[r] = 0;
            status[r] = kStatusNotRun;
            if (flags) flags[r] = 0;
        }
    }
}

// Work-queue order for the next launch of a batch of the same size: the instances by decreasing
// iteration count of this launch (longest first; ties in any order). Taking the longest solves
// first shortens the end of a launch, where a few long solves (diverging outages run max_iter
// iterations) would otherwise start late and leave the other workgroups idle. One workgroup of
// 1024 lanes: a counting sort (a histogram of the counts in LDS, offsets by decreasing count, then
// every instance placed at its count's next offset), two passes over the counts.
__global__ void __launch_bounds__(1024) k_order_by_iters(const int* iters, long long n, int nkeys, int* perm) {
    __shared__ int cnt[64], cur[64];
    const int tid = threadIdx.x;
    auto key = [&](long long u) {
        const int it = iters[u];
        return it >= nkeys - 1 ? nkeys - 1 : (it < 0 ? 0 : it);
    };
    if (tid < 64) cnt[tid] = 0;
    __syncthreads();
    for (long long u = tid; u < n; u += 1024) atomicAdd(&cnt[key(u)], 1);
    __syncthreads();
    if (tid == 0) {
        int off = 0;
        for (int k = nkeys - 1; k >= 0; --k) {
            cur[k] = off;
            off += cnt[k];
        }
    }
    __syncthreads();
    for (long long u = tid; u < n; u += 1024) perm[atomicAdd(&cur[key(u)], 1)] = (int)u;
}

// Branch flows at the origin side, one (instance, branch) per thread.
__global__ void k_branch_flows(int ninst, int n, const double* __restrict__ V, int nbr, const int* __restrict__ f,
                               const int* __restrict__ t, const double* __restrict__ yff,
                               const double* __restrict__ yft, const double* __restrict__ vn, double sn_mva,
                               double* __restrict__ amps, double* __restrict__ p) {
    const long long total = (long long)ninst * nbr;
    const double inv_sqrt3 = 0.57735026918962576451;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const long long i = k / nbr;
        const int b = (int)(k - i * nbr);
        const double* Vi = V + 2 * i * (long long)n;
        const c64 ef = ld_c(Vi, f[b]), et = ld_c(Vi, t[b]);
        const c64 I = cadd(cmul(ld_c(yff, b), ef), cmul(ld_c(yft, b), et));
        const c64 S = cmul(ef, cconj(I));
        if (p) p[k] = S.re * sn_mva;
        if (amps) {
            const double sabs = sqrt(S.re * S.re + S.im * S.im);
            const double eabs = sqrt(ef.re * ef.re + ef.im * ef.im);
            amps[k] = sabs * sn_mva * inv_sqrt3 / (eabs * vn[b]);
        }
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// A graph capture on one thread's stream and a call that synchronises with the device
// (hipFree, hipMalloc, a copy on the legacy stream) on another thread must not overlap: HIP fails
// the call with hipErrorStreamCaptureImplicit (it would wait on the capturing stream). Several
// contexts run on threads of their own (the serving loop), so every such call of this library and
// every capture hold this lock.
std::recursive_mutex& capture_mu() {
    static std::recursive_mutex m;
    return m;
}
hipError_t dev_malloc(void** p, size_t bytes) {
    std::lock_guard<std::recursive_mutex> lk(capture_mu());
    return hipMalloc(p, bytes);
}
hipError_t dev_free(void* p) {
    std::lock_guard<std::recursive_mutex> lk(capture_mu());
    return hipFree(p);
}

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void upload(const std::vector<T>& v) {
        release();
        n = v.size();
        if (n == 0) return;
        std::lock_guard<std::recursive_mutex> lk(capture_mu());
        HIP_OK(hipMalloc(&p, n * sizeof(T)));
        HIP_OK(hipMemcpy(p, v.data(), n * sizeof(T), hipMemcpyHostToDevice));
    }
    void release() {
        if (p) (void)dev_free(p);
        p = nullptr;
        n = 0;
    }
    T* get_or_alloc(size_t count) {
        if (n < count) {
            release();
            HIP_OK(dev_malloc((void**)&p, count * sizeof(T)));
            n = count;
        }
        return p;
    }
    ~DevBuf() { release(); }
};

}  // namespace


struct pfgpu_grid {
    pfgpu_ctx* ctx = nullptr;
    pfgpu::YbusPattern Y;
    pfgpu::JPattern J;
    pfgpu::DevicePlan P;
    pfgpu_grid_info info{};
    DevBuf<int> yr_ptr, yr_col, yr_pos, fslot_p, fslot_q, ux_bus, ux_slot, j_ypos, j_rbus, j_cbus, j_slot, zero_slots;
    DevBuf<int> yr_pc, yr_ell, ye_pos, ye_rc, ye_slot, ye_jidx, ye_diag_of, jc_slot, jc_jidx;
    int ell_w = 0;  // columns of yr_ell
    DevBuf<double> jc_val;
    DevBuf<int8_t> ux_kind;
    DevBuf<int> bus_dx;
    int slack_dx = -1;
    DevBuf<uint8_t> j_sel;
    DevBuf<double> j_cval, w;
    // per-(G, block size) lane streams of the factor/forward and backward schedules
    struct Streams {
        DevBuf<int> lv_K, fs, bs;
    };
    std::vector<std::unique_ptr<Streams>> streams;  // by (G, block size, value size, lean)
    // autotuned launch shape per precision (G, block), 0 = not tuned
    int tuned_G[2] = {0, 0}, tuned_NT[2] = {0, 0};
    // the next solve's first iteration is the solver's first factorisation (KLULinearSolver::
    // initialize, KLUSolver.cpp:24-45); cleared by a solve, set by load and pfgpu_grid_reset
    bool fresh = true;
    bool dc = false;  // a DC grid (pfgpu_grid_load_dc): one linear solve per instance
    bool fdpf = false;  // a fast-decoupled grid (pfgpu_grid_load_fdpf): constant J, half-steps
    std::vector<double> w_h;  // slack weights (host copy)
    // Pivots chosen on the numbers (host re-pivot): a plan of the same grid whose factor takes
    // its pivots off the diagonal where the first factorisation of a fresh solve met a zero or
    // non-finite pivot; once built it serves every solve until pfgpu_grid_reset, as klu_refactor
    // reuses the pivots of klu_factor (KLUSolver.cpp:24-59)
    std::unique_ptr<pfgpu_grid> alt;
    bool alt_active = false;
    DevBuf<double> jx_probe;  // J values of the instance the re-pivot chooses its pivots on
    // work-queue order of the next batch of order_n instances (k_order_by_iters of the last one)
    DevBuf<int> order;
    long long order_n = 0;
    // Wide launches (wide.hpp): the plan's tasks and pairs as flat tables and the per-level team
    // sizes for `blocks` workgroups (uploaded by the first wide launch)
    struct Wide {
        DevBuf<int> f_lvl, b_lvl, f_team, b_team, tasks, pairs, btasks, bpairs;
        int blocks = 0;
    } wide;
    // Launch sequences (k_init_v, k_gather_y, k_nr) captured as HIP graphs, by the exact launch
    // arguments: a sequence seen a second time is captured, and replayed from then on.
    struct GraphEntry {
        std::vector<unsigned char> key;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        unsigned long long used = 0;
    };
    std::vector<GraphEntry> graphs;                 // at most kMaxGraphs, least recently used replaced
    std::vector<std::vector<unsigned char>> seen;   // sequences launched once, not captured yet
    unsigned long long clock = 0;
    ~pfgpu_grid() {
        for (auto& e : graphs) {
            if (e.exec) (void)hipGraphExecDestroy(e.exec);
            if (e.graph) (void)hipGraphDestroy(e.graph);
        }
    }
};

namespace {

// (the dynamic LDS limit of the variant is raised by resident_blocks before every launch, so that
// a launch under stream capture issues nothing but the launch)
template <typename FT, int G, int NT, bool LEAN>
void launch_nr(const KArgs& a, int nblocks, size_t lds, hipStream_t st) {
    auto kfn = k_nr<FT, G, NT, LEAN>;
    hipLaunchKernelGGL(kfn, dim3(nblocks), dim3(NT), lds, st, a);
    HIP_OK(hipGetLastError());
}

template <typename FT, int NT, bool LEAN>
void launch_nr_g(int G, const KArgs& a, int nblocks, size_t lds, hipStream_t st) {
    switch (G) {
        case 1: launch_nr<FT, 1, NT, LEAN>(a, nblocks, lds, st); break;
        case 2: launch_nr<FT, 2, NT, LEAN>(a, nblocks, lds, st); break;
        case 4: launch_nr<FT, 4, NT, LEAN>(a, nblocks, lds, st); break;
        case 8: launch_nr<FT, 8, NT, LEAN>(a, nblocks, lds, st); break;
        case 16: launch_nr<FT, 16, NT, LEAN>(a, nblocks, lds, st); break;
        case 32: launch_nr<FT, 32, NT, LEAN>(a, nblocks, lds, st); break;
        default: throw ApiError(PFGPU_E_INVALID, "instances_per_wg must be 1, 2, 4, 8, 16 or 32");
    }
}

template <typename FT>
void launch_nr_any(int G, int NT, bool lean, const KArgs& a, int nblocks, size_t lds, hipStream_t st) {
    if (lean) {
        if (NT == 1024) launch_nr_g<FT, 1024, true>(G, a, nblocks, lds, st);
        else launch_nr_g<FT, 256, true>(G, a, nblocks, lds, st);
    } else {
        if (NT == 1024) launch_nr_g<FT, 1024, false>(G, a, nblocks, lds, st);
        else launch_nr_g<FT, 256, false>(G, a, nblocks, lds, st);
    }
}

int log2i(int g) {
    int l = 0;
    while ((1 << l) < g) ++l;
    return l;
}

struct LaunchChoice {
    int G, NT, nlds, nspill, nblocks;
    bool lean;       // every value slot in LDS: the lean kernel variant and streams
    size_t lds;      // dynamic LDS bytes: value slots, the level table, the workgroup state
    size_t lvk_off;  // byte offset of the level table
    size_t sh_off;   // byte offset of the workgroup state
    size_t v_off;    // lean launches: byte offset of the instances' V
};

template <typename FT, int NT, bool LEAN>
const void* nr_kernel_g(int G) {
    switch (G) {
        case 1: return (const void*)k_nr<FT, 1, NT, LEAN>;
        case 2: return (const void*)k_nr<FT, 2, NT, LEAN>;
        case 4: return (const void*)k_nr<FT, 4, NT, LEAN>;
        case 8: return (const void*)k_nr<FT, 8, NT, LEAN>;
        case 16: return (const void*)k_nr<FT, 16, NT, LEAN>;
        case 32: return (const void*)k_nr<FT, 32, NT, LEAN>;
        default: throw ApiError(PFGPU_E_INVALID, "instances_per_wg must be 1, 2, 4, 8, 16 or 32");
    }
}

// Static LDS of one k_nr variant (per-instance scalars and the reduction buffer);
// the larger of the lean and the general variant.
size_t static_lds(int fsize, int G, int NT) {
    static std::mutex mu;
    static std::map<int, size_t> cache;
    const int key = fsize * 100000 + G * 10000 + NT;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    size_t mx = 0;
    for (bool lean : {false, true}) {
        const void* fn;
        if (fsize == 4)
            fn = NT == 1024 ? (lean ? nr_kernel_g<float, 1024, true>(G) : nr_kernel_g<float, 1024, false>(G))
                            : (lean ? nr_kernel_g<float, 256, true>(G) : nr_kernel_g<float, 256, false>(G));
        else
            fn = NT == 1024 ? (lean ? nr_kernel_g<double, 1024, true>(G) : nr_kernel_g<double, 1024, false>(G))
                            : (lean ? nr_kernel_g<double, 256, true>(G) : nr_kernel_g<double, 256, false>(G));
        hipFuncAttributes attr{};
        HIP_OK(hipFuncGetAttributes(&attr, fn));
        mx = std::max(mx, (size_t)attr.sharedSizeBytes);
    }
    cache[key] = mx;
    return mx;
}

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// Launch heuristic: the largest G whose value arrays fit entirely in LDS while
// keeping at least two workgroups per CU; otherwise G = 1 with the hottest
// slots in LDS and the rest spilled to global memory, or, for a large batch of
// a plan of which more than half spills even at G = 1, G = 4 (FP32) / 2 (FP64).
LaunchChoice choose_launch(const pfgpu_ctx* ctx, const pfgpu::DevicePlan& P, int ninst, int fsize, int G_req,
                           int NT_req = 0) {
    // per workgroup: static LDS of the variant + value slots + the level table
    const size_t table = align16((size_t)(P.nlev + P.nblev) * 8);
    auto avail = [&](int G, int NT) -> size_t {
        const size_t fixed = static_lds(fsize, G, NT) + table + align16((size_t)wg_state_bytes(G, NT)) + 32;
        return (size_t)ctx->lds_max > fixed ? (size_t)ctx->lds_max - fixed : 0;
    };
    LaunchChoice c{};
    if (G_req > 0) {
        c.G = G_req;
    } else {
        c.G = 1;
        for (int G = kMaxG; G >= 2; G >>= 1) {
            const size_t need = (size_t)P.nvals * G * fsize + (size_t)P.n * G * 16;
            const long long blocks = (ninst + G - 1) / G;
            if (need <= avail(G, 1024) && blocks >= 2LL * std::max(1, ctx->n_cu)) { c.G = G; break; }
        }
        if (c.G == 1) {
            // small batches: still allow G>1 if it fits and there is more than one instance
            for (int G = std::min(kMaxG, 8); G >= 2; G >>= 1) {
                const size_t need = (size_t)P.nvals * G * fsize + (size_t)P.n * G * 16;
                if (need <= avail(G, 1024) && ninst >= 4 * G * std::max(1, ctx->n_cu)) { c.G = G; break; }
            }
        }
        // a plan of which more than half spills even at one instance per workgroup: large batches
        // take several instances per workgroup (4 with FP32 factors, 2 with FP64), which hide the
        // spilled slots' memory latency better (the autotuner's choice on the 10 000-bus N-1 sweep in
        // both forms; on the 2000-bus FP64 plan, a third of which spills, it keeps 1)
        const int Gs = fsize == 4 ? 4 : 2;
        if (c.G == 1 && (size_t)P.nvals * fsize > 2 * avail(1, 1024) && ninst >= 2LL * Gs * std::max(1, ctx->n_cu))
            c.G = Gs;
    }
    // the dense tail must be LDS-resident
    while (c.G > 1 && (size_t)(P.m * P.m + P.m) * c.G * fsize > avail(c.G, 1024)) c.G >>= 1;
    // one workgroup per CU when the value arrays take more than half of the LDS: give it 16 waves
    const size_t full = (size_t)P.nvals * c.G * fsize;
    c.NT = full > avail(c.G, 1024) / 2 ? 1024 : 256;
    if (NT_req == 256 || NT_req == 1024) c.NT = NT_req;
    const size_t cap = avail(c.G, c.NT) / ((size_t)c.G * fsize);
    c.nlds = (int)std::min<size_t>((size_t)P.nvals, cap);
    if (c.nlds < std::max(1, P.m * P.m + P.m))
        throw ApiError(PFGPU_E_INVALID, "the grid's dense tail does not fit in LDS");
    c.nspill = P.nvals - c.nlds;
    c.lvk_off = align16((size_t)c.nlds * c.G * fsize);
    c.sh_off = c.lvk_off + table;
    c.v_off = c.sh_off + align16((size_t)wg_state_bytes(c.G, c.NT));
    // lean: every value slot in LDS, and the instances' V beside them
    c.lean = c.nspill == 0 && c.v_off + (size_t)P.n * c.G * 16 <= (size_t)ctx->lds_max;
    // lean items carry LDS addresses: the value slots must start at LDS address 0
    if (c.lean && static_lds(fsize, c.G, c.NT) != 0)
        throw ApiError(PFGPU_E_INTERNAL, "the NR kernel has static LDS: lean item offsets would not be LDS addresses");
    c.nblocks = (ninst + c.G - 1) / c.G;
    c.lds = c.lean ? c.v_off + (size_t)P.n * c.G * 16 : c.v_off;
    return c;
}

int shape_index(int G, int NT, int fsize, bool lean) {
    return ((2 * log2i(G) + (NT == 1024 ? 1 : 0)) * 2 + (fsize == 8 ? 1 : 0)) * 2 + (lean ? 1 : 0);
}

void ensure_streams(pfgpu_grid* gr, int G, int NT, int fsize, bool lean) {
    const int ix = shape_index(G, NT, fsize, lean);
    if ((int)gr->streams.size() <= ix) gr->streams.resize(ix + 1);
    if (gr->streams[ix]) return;
    pfgpu::LeanFormat fmt;
    fmt.elem_bytes = G * fsize;
    fmt.lane_round = std::max(1, 64 / G);
    const pfgpu::LaneStreams L = pfgpu::build_lane_streams(gr->P, G, NT / G, lean ? &fmt : nullptr);
    auto t = std::make_unique<pfgpu_grid::Streams>();
    auto up = [](DevBuf<int>& d, const std::vector<int>& v) { d.upload(v.empty() ? std::vector<int>(4, -1) : v); };
    std::vector<int> K;
    // per level {K | Lc << 16 (lean streams: chunk length), U | log2(largest team) << 24}
    auto lg = [](int t) { int l = 0; while ((1 << l) < t) ++l; return l; };
    for (size_t l = 0; l < L.f_lv.size() / 4; ++l) {
        K.push_back(L.f_lv[4 * l + 1] | (lean ? L.f_lc[l] << 16 : 0));
        K.push_back(L.f_lv[4 * l + 3] | (lg(L.f_lv[4 * l + 2]) << 24));
    }
    for (size_t l = 0; l < L.b_lv.size() / 4; ++l) {
        K.push_back(L.b_lv[4 * l + 1] | (lean ? L.b_lc[l] << 16 : 0));
        K.push_back(L.b_lv[4 * l + 3] | (lg(L.b_lv[4 * l + 2]) << 24));
    }
    up(t->lv_K, K);
    up(t->fs, L.fs);
    up(t->bs, L.bs);
    gr->streams[ix] = std::move(t);
}

void* grow_buffer(void*& p, size_t& cap, size_t bytes) {
    if (bytes > cap) {
        if (p) HIP_OK(dev_free(p));
        p = nullptr;
        cap = 0;
        HIP_OK(dev_malloc(&p, bytes));
        cap = bytes;
    }
    return p;
}

void* ws_get(pfgpu_ctx* ctx, size_t bytes) {
    if (bytes > ctx->budget) throw ApiError(PFGPU_E_BUDGET, "batch workspace exceeds the context's HBM budget");
    return grow_buffer(ctx->ws, ctx->ws_bytes, bytes);
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// One solve call as the NR launches see it (device pointers).
struct SolveArgs {
    int ninst = 0;          // launch units: instances, or chains in chain mode
    int chain = 1;          // steps per chain (1: one solve per instance)
    long long nrows = 0;    // rows of V / Sbus / iters / status (ninst * chain, the last chain may be short)
    const double* Yx = nullptr;
    int Yx_stride = 0;
    const int *dptr = nullptr, *dpos = nullptr;
    const double* dval = nullptr;
    const double* S = nullptr;
    int sbus_stride = 0;
    const double* V0 = nullptr;
    int v0_stride = 0;
    double* V = nullptr;
    int max_iter = 10;
    double tol = 1e-8;
    int precision = PFGPU_FP64;  // PFGPU_FP64, PFGPU_FP32 or PFGPU_FP32_STRICT
    int G_req = 0, NT_req = 0;
    int* iters = nullptr;
    int* status = nullptr;
    double* Jx_last = nullptr;
    long long jx_row = 0;   // the row whose J values Jx_last receives
};

// Resident workgroups of a launch shape (occupancy of the kernel variant at its LDS size).
int resident_blocks(const pfgpu_ctx* ctx, int fsize, int G, int NT, bool lean, size_t lds) {
    const void* fn;
    if (fsize == 4)
        fn = NT == 1024 ? (lean ? nr_kernel_g<float, 1024, true>(G) : nr_kernel_g<float, 1024, false>(G))
                        : (lean ? nr_kernel_g<float, 256, true>(G) : nr_kernel_g<float, 256, false>(G));
    else
        fn = NT == 1024 ? (lean ? nr_kernel_g<double, 1024, true>(G) : nr_kernel_g<double, 1024, false>(G))
                        : (lean ? nr_kernel_g<double, 256, true>(G) : nr_kernel_g<double, 256, false>(G));
    if (lds > 65536) HIP_OK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, NT, lds));
    if (per_cu < 1) throw ApiError(PFGPU_E_INTERNAL, "the NR launch shape cannot be resident on a CU");
    return per_cu * std::max(1, ctx->n_cu);
}

// Per-call rows in ctx->aux: [queue counter | fallback count | pad] [phase clocks: 8 x u64]
// [flags: nrows ints] [fallback list: ninst ints]
struct AuxLayout {
    size_t o_queue = 0, o_count = 4, o_phase = 64, o_flags = 128, o_list = 0, bytes = 0;
};
AuxLayout aux_layout(long long nrows, long long ninst) {
    AuxLayout L;
    L.o_list = align_up(L.o_flags + (size_t)std::max(nrows, 1LL) * 4);
    L.bytes = align_up(L.o_list + (size_t)std::max(ninst, 1LL) * 4);
    return L;
}

// A workspace given to the call instead of allocated by it (the no-allocation entry point): the
// per-call rows at aux, the kept initial voltages at io, the launch area at ws (ws_bytes).
struct FixedScratch {
    char* aux = nullptr;
    char* io = nullptr;
    char* ws = nullptr;
    size_t ws_bytes = 0;
};

// Wide launches: one instance at a time over every workgroup of the device (wide.hpp). Taken for
// instances_per_wg == -1, and by default for a single launch unit (one instance, or one chain of
// time steps) on a plan of at least kWidePairs factor pairs, where one workgroup per instance would
// leave the device idle. Not for DC grids, J captures or the allocation-free entry point.
constexpr long long kWidePairs = 200000;
constexpr int kWideBlocks = 32;  // workgroups of a wide launch (PFGPU_WIDE_BLOCKS overrides)

bool wide_wanted(const pfgpu_grid* gr, const SolveArgs& q, int G_req, int nunits, const FixedScratch* fx) {
    if (gr->dc || gr->fdpf || q.Jx_last || fx) return false;
    if (G_req == -1) return true;
    return G_req == 0 && nunits == 1 && (long long)gr->P.pairs.size() / 4 >= kWidePairs;
}

// log2 of the team size of each level: the largest power of two (at most 64, at most the level's
// longest task rounded up) with which the level's tasks still take one lane team each
std::vector<int> level_teams(const std::vector<int>& lvl, const std::vector<int>& tasks, long long lanes) {
    std::vector<int> out;
    for (size_t l = 0; l + 1 < lvl.size(); ++l) {
        const long long nt = lvl[l + 1] - lvl[l];
        int maxp = 1;
        for (int t = lvl[l]; t < lvl[l + 1]; ++t) maxp = std::max(maxp, tasks[4 * t + 2] - tasks[4 * t + 1]);
        int lt = 0;
        while (lt < 6 && (1 << lt) < maxp && nt * (2LL << lt) <= lanes) ++lt;
        out.push_back(lt);
    }
    return out;
}

void launch_wide(pfgpu_ctx* ctx, pfgpu_grid* gr, const SolveArgs& q, int precision, int nunits, const int* imap,
                 const AuxLayout& AL, bool fresh, pfgpu_stats* st) {
    const pfgpu::DevicePlan& P = gr->P;
    const int n = P.n;
    const bool fp32 = precision != PFGPU_FP64;
    const int fsize = fp32 ? 4 : 8;
    const size_t lds = pfgpu_wide::lds_bytes(P.m, fp32);
    if (lds > (size_t)ctx->lds_max) throw ApiError(PFGPU_E_INVALID, "the dense tail does not fit a wide launch's LDS");
    int nb = pfgpu_wide::resident_blocks(ctx->device, lds, fp32);
    if (nb <= 0) throw ApiError(PFGPU_E_DEVICE, "the device cannot run a cooperative (wide) launch");
    // a grid barrier costs about 0.1 us per workgroup (one device-scope atomic each): a few dozen
    // workgroups balance barrier cost against the lanes a wide level needs
    static const int wide_blocks = [] {
        const char* e = std::getenv("PFGPU_WIDE_BLOCKS");
        return e ? std::max(1, std::atoi(e)) : kWideBlocks;
    }();
    nb = std::min(nb, wide_blocks);
    auto& W = gr->wide;
    if (W.blocks != nb) {
        const long long lanes = (long long)nb * pfgpu_wide::kThreads;
        W.f_lvl.upload(P.f_lvl);
        W.b_lvl.upload(P.b_lvl);
        W.f_team.upload(level_teams(P.f_lvl, P.tasks, lanes));
        W.b_team.upload(level_teams(P.b_lvl, P.btasks, lanes));
        auto nonempty = [](const std::vector<int>& v, size_t k) { return v.empty() ? std::vector<int>(k, 0) : v; };
        W.tasks.upload(nonempty(P.tasks, 4));
        W.pairs.upload(nonempty(P.pairs, 4));
        W.btasks.upload(nonempty(P.btasks, 4));
        W.bpairs.upload(nonempty(P.bpairs, 2));
        W.blocks = nb;
    }
    size_t off = 0;
    const size_t o_x = off; off += align_up((size_t)P.nvals * fsize);
    const size_t o_va = off; off += align_up((size_t)n * 8);
    const size_t o_vm = off; off += align_up((size_t)n * 8);
    const size_t o_ib = off; off += align_up((size_t)n * 16);
    const size_t o_yw = off; off += align_up((size_t)P.nnzY * 16);
    const size_t o_red = off; off += align_up((size_t)4 * nb * 8);
    const size_t o_ctl = off; off += align_up(64 * sizeof(int));
    const size_t o_bar = off; off += align_up(pfgpu_wide::kBarrierWords * sizeof(unsigned int));
    char* ws = static_cast<char*>(ws_get(ctx, off));
    char* aux = static_cast<char*>(ctx->aux);
    hipStream_t stream = ctx->stream;
    const bool copy = !(q.V0 == q.V && q.v0_stride != 0 && q.chain == 1);
    pfgpu_wide::WArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n = n; a.N = P.N; a.nJ = P.nJ; a.nzero = (int)P.zero_slots.size(); a.nlev = P.nlev; a.nblev = P.nblev; a.m = P.m;
    a.single_slack = gr->info.single_slack; a.max_iter = q.max_iter; a.chain = q.chain; a.fresh = fresh ? 1 : 0;
    a.nblocks = nb; a.nunits = nunits; a.nrows = q.nrows; a.tol = q.tol; a.stall_tol = std::sqrt(q.tol);
    a.yr_ptr = gr->yr_ptr.p; a.yr_col = gr->yr_col.p; a.yr_pos = gr->yr_pos.p;
    a.fslot_p = gr->fslot_p.p; a.fslot_q = gr->fslot_q.p;
    a.bus_dx = reinterpret_cast<const int2*>(gr->bus_dx.p); a.slack_dx = gr->slack_dx;
    a.j_ypos = gr->j_ypos.p; a.j_rbus = gr->j_rbus.p; a.j_cbus = gr->j_cbus.p; a.j_slot = gr->j_slot.p;
    a.zero_slots = gr->zero_slots.p; a.j_sel = gr->j_sel.p; a.j_cval = gr->j_cval.p;
    a.f_lvl = W.f_lvl.p; a.tasks = reinterpret_cast<const int4*>(W.tasks.p); a.pairs = reinterpret_cast<const int4*>(W.pairs.p);
    a.f_team = W.f_team.p;
    a.b_lvl = W.b_lvl.p; a.btasks = reinterpret_cast<const int4*>(W.btasks.p);
    a.bpairs = reinterpret_cast<const int2*>(W.bpairs.p); a.b_team = W.b_team.p;
    a.w = gr->w.p;
    a.Yx = q.Yx; a.y_stride = (long long)q.Yx_stride * 2;
    a.dptr = q.dptr; a.dpos = q.dpos; a.dval = q.dval;
    a.S = q.S; a.s_stride = (long long)q.sbus_stride * 2;
    a.V = q.V; a.iters = q.iters; a.status = q.status; a.flags = reinterpret_cast<int*>(aux + AL.o_flags);
    a.imap = imap;
    a.X = ws + o_x;
    a.Va = reinterpret_cast<double*>(ws + o_va); a.Vm = reinterpret_cast<double*>(ws + o_vm);
    a.Ib = reinterpret_cast<double*>(ws + o_ib); a.Yw = reinterpret_cast<double*>(ws + o_yw);
    a.red = reinterpret_cast<double*>(ws + o_red); a.ctl = reinterpret_cast<int*>(ws + o_ctl);
    a.bar = reinterpret_cast<unsigned int*>(ws + o_bar);
    HIP_OK(hipEventRecord(ctx->ev0, stream));
    hipLaunchKernelGGL(k_init_v, dim3(copy ? 1024 : 1), dim3(256), 0, stream, q.V0, (long long)q.v0_stride * 2, q.V, n,
                       nunits, imap, q.chain, reinterpret_cast<int*>(aux + AL.o_queue), nullptr, copy ? 1 : 0);
    HIP_OK(hipGetLastError());
    HIP_OK(pfgpu_wide::launch(a, fp32, lds, stream));
    HIP_OK(hipEventRecord(ctx->ev1, stream));
    HIP_OK(hipStreamSynchronize(stream));
    if (st) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        st->kernel_s += ms * 1e-3;
        st->instances_per_wg = -1;
        st->block_size = pfgpu_wide::kThreads;
        st->lds_slots = 0;
        st->workgroups = nb;
        st->resident_workgroups = nb;
    }
}

// One NR launch (k_init_v, k_gather_y for a shared Ybus, k_nr) over `nunits` launch units;
// imap (device, or null) maps them to the caller's instances / chains. With a fixed workspace the
// launch allocates nothing: the schedule of its launch shape must already be on the device, and
// the number of workgroups is capped to what the workspace holds (each workgroup takes groups of
// instances from the work queue, so fewer workgroups only means less concurrency).
void launch_pass(pfgpu_ctx* ctx, pfgpu_grid* gr, const SolveArgs& q, int precision, int G_req, int NT_req,
                 int nunits, const int* imap, const AuxLayout& AL, bool fresh, pfgpu_stats* st,
                 const FixedScratch* fx = nullptr) {
    if (wide_wanted(gr, q, G_req, nunits, fx)) {
        launch_wide(ctx, gr, q, precision, nunits, imap, AL, fresh, st);
        return;
    }
    const pfgpu::DevicePlan& P = gr->P;
    const int n = P.n;
    const int fsize = precision == PFGPU_FP64 ? 8 : 4;
    const int pslot = precision == PFGPU_FP64 ? 0 : 1;
    if (G_req == -1) G_req = 0;  // a wide launch was not possible here (DC grid, J capture, fixed scratch)
    if (G_req == 0 && NT_req == 0 && gr->tuned_G[pslot]) {
        G_req = gr->tuned_G[pslot];
        NT_req = gr->tuned_NT[pslot];
    }
    LaunchChoice lc = choose_launch(ctx, P, nunits, fsize, G_req, NT_req);
    if (fx) {
        const int ix = shape_index(lc.G, lc.NT, fsize, lc.lean);
        if ((int)gr->streams.size() <= ix || !gr->streams[ix])
            throw ApiError(PFGPU_E_INVALID, "the schedule of this launch shape is not on the device yet: solve a batch "
                                            "of this shape once with pfgpu_solve_batch_device first");
    } else {
        ensure_streams(gr, lc.G, lc.NT, fsize, lc.lean);
    }
    const long long ngroups = ((long long)nunits + lc.G - 1) / lc.G;
    int nres = (int)std::min<long long>(ngroups, resident_blocks(ctx, fsize, lc.G, lc.NT, lc.lean, lc.lds));
    const bool gather = q.Yx_stride == 0;
    const size_t gather_bytes = gather ? align_up((size_t)gr->ell_w * n * 16) + align_up(P.ye_pos.size() * 16) : 0;
    if (fx) {
        // workgroups whose rows fit the fixed workspace (4 * 256 bytes covers the roundings)
        const size_t per_block = (size_t)lc.G * ((size_t)n * 16 + (size_t)std::max(P.nyd, 1) * 16) +
                                 (size_t)lc.nspill * lc.G * fsize;
        const size_t room = fx->ws_bytes > gather_bytes + 1024 ? fx->ws_bytes - gather_bytes - 1024 : 0;
        const long long cap = per_block ? (long long)(room / per_block) : nres;
        if (cap < 1) throw ApiError(PFGPU_E_BUDGET, "the grid's per-workgroup rows do not fit the fixed scratch");
        nres = (int)std::min<long long>(nres, cap);
    }
    const long long slots = (long long)nres * lc.G;  // workspace rows: one per resident instance
    size_t off = 0;
    const size_t o_va = off; off += align_up((size_t)slots * n * 8);
    const size_t o_vm = off; off += align_up((size_t)slots * n * 8);
    const size_t o_ib = off; off += align_up((size_t)slots * std::max(P.nyd, 1) * 16);
    const size_t o_x = off; off += align_up((size_t)nres * lc.nspill * lc.G * fsize);
    const size_t o_yell = off; off += gather ? align_up((size_t)gr->ell_w * n * 16) : 0;
    const size_t o_yye = off; off += gather ? align_up(P.ye_pos.size() * 16) : 0;
    if (fx && off > fx->ws_bytes) throw ApiError(PFGPU_E_INTERNAL, "fixed scratch layout overflow");
    char* ws = fx ? fx->ws : static_cast<char*>(ws_get(ctx, std::max<size_t>(off, 256)));
    char* aux = fx ? fx->aux : static_cast<char*>(ctx->aux);
    hipStream_t stream = ctx->stream;
    int* queue = reinterpret_cast<int*>(aux + AL.o_queue);
    unsigned long long* phase = st ? reinterpret_cast<unsigned long long*>(aux + AL.o_phase) : nullptr;
    // in place (V0 == V, one row per instance): nothing to copy, rows are read where they are
    const bool copy = !(q.V0 == q.V && q.v0_stride != 0 && q.chain == 1);
    KArgs a;
    std::memset(&a, 0, sizeof(a));  // padding too: the struct's bytes key the graph cache
    a.n = n; a.N = P.N; a.nJ = P.nJ; a.nzero = (int)P.zero_slots.size(); a.nlev = P.nlev; a.nblev = P.nblev;
    a.m = P.m;
    if (lc.nlds < P.m * P.m + P.m) throw ApiError(PFGPU_E_INTERNAL, "dense tail does not fit in LDS");
    a.nlds = lc.nlds; a.nspill = lc.nspill;
    a.single_slack = gr->info.single_slack; a.ninst = nunits; a.max_iter = q.max_iter; a.tol = q.tol;
    a.yr_ptr = gr->yr_ptr.p; a.yr_col = gr->yr_col.p; a.yr_pos = gr->yr_pos.p;
    a.yr_pc = reinterpret_cast<const int2*>(gr->yr_pc.p);
    a.yr_ell = reinterpret_cast<const int2*>(gr->yr_ell.p);
    a.ye_pos = gr->ye_pos.p; a.ye_rc = reinterpret_cast<const int2*>(gr->ye_rc.p);
    a.ye_slot = reinterpret_cast<const int4*>(gr->ye_slot.p); a.ye_jidx = reinterpret_cast<const int4*>(gr->ye_jidx.p);
    a.nye = (int)P.ye_pos.size();
    a.jc_slot = gr->jc_slot.p; a.jc_jidx = gr->jc_jidx.p; a.jc_val = gr->jc_val.p; a.njc = (int)P.jc_slot.size();
    a.dptr = q.dptr; a.dpos = q.dpos; a.dval = q.dval;
    a.fslot_p = gr->fslot_p.p; a.fslot_q = gr->fslot_q.p;
    a.ux_bus = gr->ux_bus.p; a.ux_slot = gr->ux_slot.p; a.ux_kind = gr->ux_kind.p;
    a.bus_dx = reinterpret_cast<const int2*>(gr->bus_dx.p); a.slack_dx = gr->slack_dx;
    a.j_ypos = gr->j_ypos.p; a.j_rbus = gr->j_rbus.p; a.j_cbus = gr->j_cbus.p; a.j_sel = gr->j_sel.p; a.j_cval = gr->j_cval.p;
    a.j_slot = gr->j_slot.p; a.zero_slots = gr->zero_slots.p;
    const auto& ls = *gr->streams[shape_index(lc.G, lc.NT, fsize, lc.lean)];
    a.lv_K = ls.lv_K.p; a.lvk_off = (int)lc.lvk_off; a.sh_off = (int)lc.sh_off;
    a.fs = reinterpret_cast<const int4*>(ls.fs.p); a.bs = reinterpret_cast<const int4*>(ls.bs.p);
    a.w = gr->w.p;
    a.Yx = q.Yx; a.y_stride = (long long)q.Yx_stride * 2;
    a.S = q.S; a.s_stride = (long long)q.sbus_stride * 2;
    a.V = q.V; a.Va = reinterpret_cast<double*>(ws + o_va); a.Vm = reinterpret_cast<double*>(ws + o_vm);
    a.Ib = reinterpret_cast<double*>(ws + o_ib); a.Xg = ws + o_x;
    a.ye_diag_of = gr->ye_diag_of.p; a.nyd = P.nyd;
    a.iters = q.iters; a.status = q.status; a.Jx_last = q.Jx_last; a.jx_row = q.jx_row;
    a.v_off = (int)lc.v_off;
    a.Yell = gather ? reinterpret_cast<const double*>(ws + o_yell) : nullptr;
    a.Yye = gather ? reinterpret_cast<const double*>(ws + o_yye) : nullptr;
    a.imap = imap;
    a.chain = q.chain;
    a.nrows = q.nrows;
    a.fresh = fresh ? 1 : 0;
    a.stall_tol = std::sqrt(q.tol);
    a.growth_tau = fsize == 4 ? 1.0 / 1024.0 : 1e-10;
    a.flags = reinterpret_cast<int*>(aux + AL.o_flags);
    a.phase_t = phase;
    a.queue = queue;
    a.ngroups = ngroups;
    a.dc = gr->dc ? 1 : 0;
    a.slack0 = P.slack0;
    a.fdpf = gr->fdpf ? 1 : 0;
    if (gr->dc) a.max_iter = 1;
    const int nell = gr->ell_w * n, nye = (int)P.ye_pos.size();
    auto issue = [&]() {
        hipLaunchKernelGGL(k_init_v, dim3(copy ? 1024 : 1), dim3(256), 0, stream, q.V0, (long long)q.v0_stride * 2, q.V,
                           n, nunits, imap, q.chain, queue, phase, copy ? 1 : 0);
        HIP_OK(hipGetLastError());
        if (gather) {
            hipLaunchKernelGGL(k_gather_y, dim3((nell + nye + 255) / 256), dim3(256), 0, stream, q.Yx, a.yr_ell, nell,
                               a.ye_pos, nye, reinterpret_cast<double*>(ws + o_yell),
                               reinterpret_cast<double*>(ws + o_yye));
            HIP_OK(hipGetLastError());
        }
        if (fsize == 4) launch_nr_any<float>(lc.G, lc.NT, lc.lean, a, nres, lc.lds, stream);
        else launch_nr_any<double>(lc.G, lc.NT, lc.lean, a, nres, lc.lds, stream);
    };
    // the launch sequence's key: every argument of the three launches
    struct Extra {
        const double* V0;
        long long v0s;
        int copy, nres, G, NT, lean, fsize, gather;
        size_t lds;
    } ex;
    std::memset(&ex, 0, sizeof(ex));
    ex.V0 = q.V0; ex.v0s = (long long)q.v0_stride * 2; ex.copy = copy; ex.nres = nres; ex.G = lc.G; ex.NT = lc.NT;
    ex.lean = lc.lean; ex.fsize = fsize; ex.gather = gather; ex.lds = lc.lds;
    std::vector<unsigned char> key(sizeof(KArgs) + sizeof(Extra));
    std::memcpy(key.data(), &a, sizeof(KArgs));
    std::memcpy(key.data() + sizeof(KArgs), &ex, sizeof(Extra));
    static const bool graphs_on = [] {
        const char* e = std::getenv("PFGPU_GRAPH");
        return !(e && e[0] == '0');
    }();
    bool replayed = false;
    HIP_OK(hipEventRecord(ctx->ev0, stream));
    if (graphs_on && !fx) {  // capturing allocates: not for the fixed-workspace entry point
        pfgpu_grid::GraphEntry* hit = nullptr;
        for (auto& e : gr->graphs)
            if (e.key == key) hit = &e;
        if (!hit && std::find(gr->seen.begin(), gr->seen.end(), key) != gr->seen.end()) {
            // second launch of this sequence: capture it
            hipGraph_t g = nullptr;
            std::lock_guard<std::recursive_mutex> lk(capture_mu());  // no device-wide call of another thread meanwhile
            HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                issue();
            } catch (...) {
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_OK(hipStreamEndCapture(stream, &g));
            hipGraphExec_t exg = nullptr;
            const hipError_t ie = hipGraphInstantiate(&exg, g, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(g);
                throw ApiError(PFGPU_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
            }
            if ((int)gr->graphs.size() >= kMaxGraphs) {  // replace the least recently used
                auto lru = std::min_element(gr->graphs.begin(), gr->graphs.end(),
                                            [](const auto& x, const auto& y) { return x.used < y.used; });
                (void)hipGraphExecDestroy(lru->exec);
                (void)hipGraphDestroy(lru->graph);
                gr->graphs.erase(lru);
            }
            gr->graphs.push_back({key, g, exg, 0});
            hit = &gr->graphs.back();
            gr->seen.erase(std::find(gr->seen.begin(), gr->seen.end(), key));
        }
        if (hit) {
            hit->used = ++gr->clock;
            HIP_OK(hipGraphLaunch(hit->exec, stream));
            replayed = true;
        } else {
            if (gr->seen.size() >= 2 * kMaxGraphs) gr->seen.erase(gr->seen.begin());
            gr->seen.push_back(key);
        }
    }
    if (!replayed) issue();
    HIP_OK(hipEventRecord(ctx->ev1, stream));
    HIP_OK(hipStreamSynchronize(stream));
    if (st) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        st->kernel_s += ms * 1e-3;
        st->instances_per_wg = lc.G;
        st->block_size = lc.NT;
        st->lds_slots = lc.nlds;
        st->workgroups = (int)std::min<long long>(ngroups, INT_MAX);
        st->resident_workgroups = nres;
        st->graph_launches += replayed ? 1 : 0;
        unsigned long long ph[8];
        HIP_OK(hipMemcpyAsync(ph, phase, sizeof(ph), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        for (int k = 0; k < 6; ++k) st->phase_s[k] += (double)ph[k] / ctx->wall_hz;
    }
}

// Computes the grid's symbolic factor, schedule and device plan from g->Y and g->J (elimination
// order `perm`, or the bus-block minimum degree) and uploads its tables.
void build_plan(pfgpu_grid* g, const std::vector<int>* perm);

// Core solve on device buffers: the launch in the requested form; with FP32 factors (unless
// strict), the rows that stalled or broke down are re-solved from their initial voltages with
// FP64 factors (chains from their heads).
void solve_device(pfgpu_ctx* ctx, pfgpu_grid* gr, const SolveArgs& q0, pfgpu_stats* st,
                  const FixedScratch* fx = nullptr) {
    const pfgpu::DevicePlan& P = gr->P;
    const int n = P.n;
    if (q0.ninst <= 0 || q0.nrows <= 0) return;
    SolveArgs q = q0;
    if (q.precision != PFGPU_FP64 && q.precision != PFGPU_FP32 && q.precision != PFGPU_FP32_STRICT)
        throw ApiError(PFGPU_E_INVALID, "bad precision");
    if (q.Yx_stride != 0 && q.Yx_stride != P.nnzY) throw ApiError(PFGPU_E_INVALID, "Yx_stride must be 0 or nnz_ybus");
    if (q.sbus_stride != 0 && q.sbus_stride != n) throw ApiError(PFGPU_E_INVALID, "sbus_stride must be 0 or n");
    if (q.v0_stride != 0 && q.v0_stride != n) throw ApiError(PFGPU_E_INVALID, "v0_stride must be 0 or n");
    if (q.max_iter < 0) throw ApiError(PFGPU_E_INVALID, "max_iter must be >= 0");
    if (q.chain < 1) throw ApiError(PFGPU_E_INVALID, "chain length must be >= 1");
    if (q.chain > 1 && q.V0 == q.V) throw ApiError(PFGPU_E_INVALID, "chains need V0 apart from V");
    if (q.dptr && q.Yx_stride != 0) throw ApiError(PFGPU_E_INVALID, "Ybus deltas need a shared base (Yx_stride 0)");
    if (gr->dc && q.precision != PFGPU_FP64) throw ApiError(PFGPU_E_INVALID, "a DC grid is solved with FP64 factors");
    if (gr->fdpf && q.precision != PFGPU_FP64)
        throw ApiError(PFGPU_E_INVALID, "a fast-decoupled grid is solved with FP64 factors");
    const AuxLayout AL = aux_layout(q.nrows, q.ninst);
    char* aux = fx ? fx->aux : static_cast<char*>(grow_buffer(ctx->aux, ctx->aux_bytes, AL.bytes));
    const bool fp32 = q.precision != PFGPU_FP64;
    const bool fallback = q.precision == PFGPU_FP32;
    // in place with an FP64 re-solve possible: keep the initial voltages
    const double* V0_keep = q.V0;
    if (fallback && q.V0 == q.V && q.v0_stride != 0) {
        const size_t bytes = (size_t)q.ninst * n * 16;
        double* keep = fx ? reinterpret_cast<double*>(fx->io)
                          : static_cast<double*>(grow_buffer(ctx->io, ctx->io_bytes, bytes));
        if (!keep) throw ApiError(PFGPU_E_INTERNAL, "no room kept for the initial voltages");
        HIP_OK(hipMemcpyAsync(keep, q.V0, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        V0_keep = keep;
    }
    const bool fresh = gr->fresh;
    pfgpu_grid* pg = gr->alt_active ? gr->alt.get() : gr;  // the plan in use
    // large batches take their instances longest first, by the iterations of the previous batch of
    // the same size (any order gives the same results: instances are independent)
    const bool ordered = !fx && q.chain == 1 && q.ninst >= 2LL * std::max(1, ctx->n_cu) &&
                         !wide_wanted(pg, q, q.G_req, q.ninst, fx);
    const int* qorder = ordered && gr->order_n == q.ninst ? gr->order.p : nullptr;
    launch_pass(ctx, pg, q, fp32 ? PFGPU_FP32 : PFGPU_FP64, q.G_req, q.NT_req, q.ninst, qorder, AL, fresh, st, fx);
    if (ordered) {
        int* perm = gr->order.get_or_alloc((size_t)q.ninst);
        hipLaunchKernelGGL(k_order_by_iters, dim3(1), dim3(1024), 0, ctx->stream, q.iters, (long long)q.ninst,
                           std::min(std::max(q.max_iter, 0), 62) + 1, perm);
        HIP_OK(hipGetLastError());
        gr->order_n = q.ninst;
    }
    gr->fresh = false;
    ctx->last_rows = q.nrows;
    ctx->last_flags = reinterpret_cast<const int*>(aux + AL.o_flags);
    int* list = reinterpret_cast<int*>(aux + AL.o_list);
    int* count = reinterpret_cast<int*>(aux + AL.o_count);
    if (fallback) {
        hipLaunchKernelGGL(k_collect_fallback, dim3(1), dim3(1024), 0, ctx->stream, q.status,
                           reinterpret_cast<const int*>(aux + AL.o_flags), q.nrows, q.chain, 0, list, count);
        HIP_OK(hipGetLastError());
        int nfb = 0;
        HIP_OK(hipMemcpyAsync(&nfb, count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        if (nfb > 0) {
            SolveArgs r = q;
            r.V0 = V0_keep;
            r.Jx_last = nullptr;
            pfgpu_stats s2{};
            // the caller's launch shape when it fixed one (results independent of which rows re-solve)
            launch_pass(ctx, pg, r, PFGPU_FP64, q.G_req, q.NT_req, nfb, list, AL, fresh, st ? &s2 : nullptr, fx);
            if (st) {
                st->fallback_fp64 = nfb;
                st->kernel_s += s2.kernel_s;
                for (int k = 0; k < 6; ++k) st->phase_s[k] += s2.phase_s[k];
            }
        }
    }
    // host re-pivot: only a fresh solve's first factorisation chooses pivots (klu_factor); the
    // strict FP32 form and the allocation-free entry point leave their SolverFactor rows as they are
    if (!fresh || fx || gr->alt_active || q.precision == PFGPU_FP32_STRICT) return;
    hipLaunchKernelGGL(k_collect_fallback, dim3(1), dim3(1024), 0, ctx->stream, q.status,
                       reinterpret_cast<const int*>(aux + AL.o_flags), q.nrows, q.chain, 1, list, count);
    HIP_OK(hipGetLastError());
    int head[2] = {0, 0};
    HIP_OK(hipMemcpyAsync(&head[0], count, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipMemcpyAsync(&head[1], list, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    const int nrp = head[0];
    if (nrp <= 0) return;
    // J at the first listed unit's initial voltages (one FP64 iteration): the matrix whose first
    // factorisation failed (of a chain: its first step's)
    double* jx = gr->jx_probe.get_or_alloc((size_t)std::max(P.nJ, 1));
    SolveArgs p = q;
    p.V0 = V0_keep;
    p.max_iter = 1;
    p.Jx_last = jx;
    p.jx_row = (long long)head[1] * q.chain;
    launch_pass(ctx, gr, p, PFGPU_FP64, q.G_req, q.NT_req, 1, list, AL, true, nullptr, nullptr);
    std::vector<double> Jx((size_t)P.nJ);
    HIP_OK(hipMemcpyAsync(Jx.data(), jx, Jx.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    const std::vector<int> order = pfgpu::column_order(gr->J);
    std::vector<int> rowpos = pfgpu::threshold_pivot_rows(gr->J, Jx, order, kRepivotTau);
    if (rowpos.empty()) return;  // numerically singular: SolverFactor stands, as klu_factor's failure
    auto alt = std::make_unique<pfgpu_grid>();
    alt->ctx = ctx;
    alt->Y = gr->Y;
    alt->J = gr->J;
    alt->J.rowpos = std::move(rowpos);
    alt->dc = gr->dc;
    alt->fdpf = gr->fdpf;
    alt->info = gr->info;
    alt->w_h = gr->w_h;
    build_plan(alt.get(), &order);
    gr->alt = std::move(alt);
    gr->alt_active = true;
    // the listed units again, from their initial voltages, on the re-pivoted plan
    SolveArgs r = q;
    r.V0 = V0_keep;
    r.Jx_last = nullptr;
    pfgpu_stats s3{};
    launch_pass(ctx, gr->alt.get(), r, PFGPU_FP64, q.G_req, q.NT_req, nrp, list, AL, true, st ? &s3 : nullptr,
                nullptr);
    if (st) {
        st->repivot_rows = nrp;
        st->kernel_s += s3.kernel_s;
        for (int k = 0; k < 6; ++k) st->phase_s[k] += s3.phase_s[k];
    }
}

}  // namespace

// A single warm-started chain (TimeSeries.cpp:80-103) in batched passes instead of one step after
// the other. Pass 1 solves every step from the chain's head. Pass 2 solves step k from pass 1's
// step k - 1 (step 0 from the head). Step 1 of pass 2 is then the sequential chain's step 1 exactly,
// and by induction step k is the chain's step k as long as every earlier step converged in both
// passes to the same voltages (to within kChainAgree): its start differs from the chain's by what
// two converged solves of one step differ, far below the tolerance, so it takes the same NR
// iterations to the same solution. At the first step that fails in pass 2 (with a valid start) the
// chain stops, as the reference's loop does; at the first step whose passes disagree (pass 1
// diverged from the head, or reached another solution), the rest is solved again the same way with
// that step's pass-2 voltages as the head. After kChainRounds rounds, or for a short rest, the rest
// runs as a sequential chain. A long chain of a large grid is taken in windows of steps whose pass
// buffers fit a quarter of the context's HBM budget (up to kChainWindowBytes), each window's head
// being the previous window's last step. Per-row flags are gathered for pfgpu_ctx_last_flags.
constexpr long long kChainTwoPassMin = 4;  // shorter chains run sequentially
constexpr int kChainRounds = 8;
constexpr double kChainAgree = 1e-6;       // pu: two converged solves of one step agree
constexpr size_t kChainWindowBytes = (size_t)1 << 30;

bool chain_two_pass_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("PFGPU_TS_SEQUENTIAL");
        return !(e && e[0] == '1');
    }();
    return on;
}

void add_stats(pfgpu_stats* st, const pfgpu_stats& s) {
    if (!st) return;
    st->kernel_s += s.kernel_s;
    st->fallback_fp64 += s.fallback_fp64;
    st->repivot_rows += s.repivot_rows;
    st->graph_launches += s.graph_launches;
    for (int k = 0; k < 6; ++k) st->phase_s[k] += s.phase_s[k];
    st->instances_per_wg = s.instances_per_wg;
    st->lds_slots = s.lds_slots;
    st->workgroups = s.workgroups;
    st->block_size = s.block_size;
    st->resident_workgroups = s.resident_workgroups;
}

void solve_single_chain(pfgpu_ctx* ctx, pfgpu_grid* gr, const SolveArgs& q, pfgpu_stats* st) {
    const int n = gr->P.n;
    const long long T = q.nrows;
    const size_t row = (size_t)n * 2;  // doubles per voltage row
    // steps per window: V1 and the pass-2 starts take 32 n bytes per step
    const size_t cap = std::min(kChainWindowBytes, ctx->budget / 4);
    const long long W = std::max<long long>(kChainTwoPassMin, std::min<long long>(T, (long long)(cap / ((size_t)n * 32 + 16))));
    // workspace: V1 and the pass-2 starts [W][n] complex, pass-1 iterations / statuses, agreement
    // [W] ints, the chain's flags [T] ints
    const size_t vbytes = align_up((size_t)W * n * 16), ibytes = align_up((size_t)W * 4);
    const size_t fbytes = align_up((size_t)T * 4);
    char* w = static_cast<char*>(grow_buffer(ctx->chain, ctx->chain_bytes, 2 * vbytes + 3 * ibytes + fbytes));
    double* V1 = reinterpret_cast<double*>(w);
    double* V0p = reinterpret_cast<double*>(w + vbytes);
    int* it1 = reinterpret_cast<int*>(w + 2 * vbytes);
    int* st1 = reinterpret_cast<int*>(w + 2 * vbytes + ibytes);
    int* agree = reinterpret_cast<int*>(w + 2 * vbytes + 2 * ibytes);
    int* flagsT = reinterpret_cast<int*>(w + 2 * vbytes + 3 * ibytes);
    hipStream_t s = ctx->stream;
    const double* head = q.V0;  // the chain's head row
    long long r0 = 0;
    int rounds = 0;
    auto keep_flags = [&](long long at, long long rows) {
        if (ctx->last_flags && rows > 0)
            HIP_OK(hipMemcpyAsync(flagsT + at, ctx->last_flags, (size_t)rows * 4, hipMemcpyDeviceToDevice, s));
    };
    auto grid_of = [](long long total) { return (int)std::min<long long>(std::max<long long>((total + 255) / 256, 1), 4096); };
    while (r0 < T) {
        if (T - r0 < kChainTwoPassMin || rounds >= kChainRounds) {
            const long long R = T - r0;
            // the rest as a sequential chain from the head
            SolveArgs c = q;
            c.ninst = 1; c.chain = (int)R; c.nrows = R;
            c.S = q.S + r0 * row; c.V0 = head; c.v0_stride = 0;
            c.V = q.V + r0 * row; c.iters = q.iters + r0; c.status = q.status + r0;
            pfgpu_stats sc{};
            solve_device(ctx, gr, c, st ? &sc : nullptr);
            add_stats(st, sc);
            keep_flags(r0, R);
            break;
        }
        const long long R = std::min(T - r0, W);  // this window's steps
        // pass 1: every step of the window from the head
        SolveArgs p1 = q;
        p1.ninst = (int)R; p1.chain = 1; p1.nrows = R; p1.Yx_stride = 0;  // every step on the chain's Ybus
        p1.S = q.S + r0 * row; p1.sbus_stride = n; p1.V0 = head; p1.v0_stride = 0;
        p1.V = V1; p1.iters = it1; p1.status = st1;
        pfgpu_stats s1{};
        solve_device(ctx, gr, p1, st ? &s1 : nullptr);
        add_stats(st, s1);
        // pass 2: step k from pass 1's step k - 1, step 0 from the head
        hipLaunchKernelGGL(k_chain_starts, dim3(grid_of(R * n)), dim3(256), 0, s, head, V1, V0p, R, n);
        HIP_OK(hipGetLastError());
        SolveArgs p2 = p1;
        p2.V0 = V0p; p2.v0_stride = n;
        p2.V = q.V + r0 * row; p2.iters = q.iters + r0; p2.status = q.status + r0;
        pfgpu_stats s2{};
        solve_device(ctx, gr, p2, st ? &s2 : nullptr);
        add_stats(st, s2);
        keep_flags(r0, R);
        hipLaunchKernelGGL(k_rows_agree, dim3((unsigned)std::min<long long>(R, 4096)), dim3(256), 0, s, V1, p2.V, st1,
                           p2.status, R, n, kChainAgree, agree);
        HIP_OK(hipGetLastError());
        std::vector<int> st2h((size_t)R), agh((size_t)R);
        HIP_OK(hipMemcpyAsync(st2h.data(), p2.status, (size_t)R * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(agh.data(), agree, (size_t)R * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        long long k = 0;
        while (k < R && st2h[k] == 0 && agh[k]) ++k;
        if (k == R) {
            // every step of the window agrees: pass 2 is the chain; the next window starts from its last
            // step
            head = q.V + (r0 + R - 1) * row;
            r0 += R;
            continue;
        }
        if (st2h[k] != 0) {
            // the chain's first failure (its start is valid): the later steps are not run
            if (r0 + k + 1 < T) {
                hipLaunchKernelGGL(k_chain_not_run, dim3(grid_of((T - r0 - k - 1) * n)), dim3(256), 0, s, q.V, q.iters,
                                   q.status, flagsT, r0 + k + 1, T, n);
                HIP_OK(hipGetLastError());
            }
            break;
        }
        // step k converged in pass 2 but its passes disagree: steps after it start again from it
        head = q.V + (r0 + k) * row;
        r0 += k + 1;
        ++rounds;
    }
    ctx->last_rows = T;
    ctx->last_flags = flagsT;
}

// The solve of a call: a single chain of enough steps in two batched passes, anything else as is.
// The passes need a tolerance tight enough that two converged solves of one step are equivalent
// starts for the next one (kChainTwoPassTol); with a looser one the chain runs step by step.
constexpr double kChainTwoPassTol = 1e-6;
void solve_call(pfgpu_ctx* ctx, pfgpu_grid* gr, const SolveArgs& q, pfgpu_stats* st) {
    if (q.ninst == 1 && q.chain > 1 && q.nrows >= kChainTwoPassMin && !q.Jx_last && !q.dptr &&
        q.tol <= kChainTwoPassTol && !gr->dc && chain_two_pass_enabled())
        solve_single_chain(ctx, gr, q, st);
    else
        solve_device(ctx, gr, q, st);
}

namespace {
// B' and B'' of a fast-decoupled grid (n x n, CSC)
struct FdpfB {
    const int *Bp_p, *Bp_i;
    const double* Bp_x;
    const int *Bpp_p, *Bpp_i;
    const double* Bpp_x;
};
int load_grid(pfgpu_ctx* ctx, int n, const int* Yp, const int* Yi, const int* slack_ids, int nslack,
              const double* slack_weights, const int* pv, int npv, const int* pq, int npq, int single_slack, bool dc,
              pfgpu_grid** out, const FdpfB* fd = nullptr);
}  // namespace

extern "C" {

const char* pfgpu_last_error(void) { return pfgpu_engine::last_error().c_str(); }
int pfgpu_abi_version(void) { return PFGPU_ABI_VERSION; }

int pfgpu_ctx_create(int device, size_t hbm_budget_bytes, pfgpu_ctx** out) {
    return guarded([&]() -> int {
        if (!out) throw ApiError(PFGPU_E_INVALID, "out is NULL");
        *out = nullptr;
        int ndev = 0;
        HIP_OK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw ApiError(PFGPU_E_INVALID, "no such device");
        HIP_OK(hipSetDevice(device));
        std::lock_guard<std::recursive_mutex> lk(capture_mu());
        scratch_init(device);
        auto c = std::make_unique<pfgpu_ctx>();
        c->device = device;
        HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_OK(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
        HIP_OK(hipDeviceGetAttribute(&c->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
        int wall_khz = 0;
        if (hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && wall_khz > 0)
            c->wall_hz = 1e3 * wall_khz;
        size_t fr = 0, tot = 0;
        HIP_OK(hipMemGetInfo(&fr, &tot));
        c->budget = hbm_budget_bytes ? hbm_budget_bytes : (size_t)(0.9 * (double)fr);
        HIP_OK(hipEventCreate(&c->ev0));
        HIP_OK(hipEventCreate(&c->ev1));
        *out = c.release();
        return 0;
    });
}

int pfgpu_ctx_destroy(pfgpu_ctx* ctx) {
    return guarded([&]() -> int {
        if (!ctx) return 0;
        (void)hipSetDevice(ctx->device);
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        std::lock_guard<std::recursive_mutex> lk(capture_mu());
        for (void* p : {ctx->ws, ctx->aux, ctx->io, ctx->stage, ctx->chain})
            if (p) (void)dev_free(p);
        if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
        if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return 0;
    });
}

int pfgpu_ctx_device_info(pfgpu_ctx* ctx, int* n_cu, int* lds_bytes, size_t* free_bytes) {
    return guarded([&]() -> int {
        if (!ctx) throw ApiError(PFGPU_E_INVALID, "ctx is NULL");
        HIP_OK(hipSetDevice(ctx->device));
        if (n_cu) *n_cu = ctx->n_cu;
        if (lds_bytes) *lds_bytes = ctx->lds_max;
        if (free_bytes) {
            size_t fr = 0, tot = 0;
            HIP_OK(hipMemGetInfo(&fr, &tot));
            *free_bytes = fr;
        }
        return 0;
    });
}

int pfgpu_grid_load(pfgpu_ctx* ctx, int n, const int* Yp, const int* Yi, const int* slack_ids, int nslack,
                    const double* slack_weights, const int* pv, int npv, const int* pq, int npq, int single_slack,
                    pfgpu_grid** out) {
    return load_grid(ctx, n, Yp, Yi, slack_ids, nslack, slack_weights, pv, npv, pq, npq, single_slack, false, out);
}

int pfgpu_grid_load_dc(pfgpu_ctx* ctx, int n, const int* Yp, const int* Yi, const int* slack_ids, int nslack,
                       const int* pv, int npv, const int* pq, int npq, pfgpu_grid** out) {
    return guarded([&]() -> int {
        if (!ctx || !out || !slack_ids || n <= 0 || nslack < 1 || npv < 0 || npq < 0 || (npv > 0 && !pv) ||
            (npq > 0 && !pq))
            throw ApiError(PFGPU_E_INVALID, "invalid grid arguments");
        // BaseDCAlgo::compute_pf (BaseDCAlgo.tpp:61-72): my_pv = slack_ids[1:] ++ pv; the buses in
        // neither my_pv nor pq are the slack buses (their angle is fixed), every other bus carries
        // its angle: the single-slack layout with all of them as PV and no PQ bus
        std::vector<char> unk(n, 0);
        for (int k = 1; k < nslack; ++k)
            if (slack_ids[k] >= 0 && slack_ids[k] < n) unk[slack_ids[k]] = 1;
        for (int k = 0; k < npv; ++k)
            if (pv[k] >= 0 && pv[k] < n) unk[pv[k]] = 1;
        for (int k = 0; k < npq; ++k)
            if (pq[k] >= 0 && pq[k] < n) unk[pq[k]] = 1;
        std::vector<int> ang, slack;
        for (int b = 0; b < n; ++b) (unk[b] ? ang : slack).push_back(b);
        if (slack.empty()) throw ApiError(PFGPU_E_INVALID, "All buses are tagged as PV or PQ, there can be no slack.");
        std::vector<double> w(n, 0.0);
        w[slack[0]] = 1.0;
        return load_grid(ctx, n, Yp, Yi, slack.data(), (int)slack.size(), w.data(), ang.data(), (int)ang.size(),
                         nullptr, 0, 1, true, out);
    });
}

int pfgpu_grid_load_fdpf(pfgpu_ctx* ctx, int n, const int* Yp, const int* Yi, const int* slack_ids, int nslack,
                         const double* slack_weights, const int* pv, int npv, const int* pq, int npq, const int* Bp_p,
                         const int* Bp_i, const double* Bp_x, const int* Bpp_p, const int* Bpp_i, const double* Bpp_x,
                         pfgpu_grid** out) {
    return guarded([&]() -> int {
        if (!ctx || !out || !slack_ids || !slack_weights || n <= 0 || nslack < 1 || npv < 0 || npq < 0 ||
            (npv > 0 && !pv) || (npq > 0 && !pq) || !Bp_p || !Bp_i || !Bp_x || !Bpp_p || !Bpp_i || !Bpp_x)
            throw ApiError(PFGPU_E_INVALID, "invalid grid arguments");
        // my_pv = slack_ids[1:] ++ pv (retrieve_pv_with_slack, BaseAlgo.hpp:201-216): the other slack
        // buses carry their angle, the first one is the reference; the single-slack layout
        std::vector<int> mypv(slack_ids + 1, slack_ids + nslack);
        mypv.insert(mypv.end(), pv, pv + npv);
        const FdpfB fd{Bp_p, Bp_i, Bp_x, Bpp_p, Bpp_i, Bpp_x};
        return load_grid(ctx, n, Yp, Yi, slack_ids, 1, slack_weights, mypv.data(), (int)mypv.size(), pq, npq, 1, false,
                         out, &fd);
    });
}

namespace {
void build_plan(pfgpu_grid* g, const std::vector<int>* perm) {
    pfgpu::SymbolicLU S = pfgpu::symbolic_lu(g->J, 96, perm);
    pfgpu::Schedule sch = pfgpu::build_schedule(S);
    g->P = pfgpu::build_device_plan(g->Y, g->J, S, sch);
    const auto& P = g->P;
    g->info.n_bus = g->Y.n;
    g->info.n_jac = P.N;
    g->info.nnz_ybus = P.nnzY;
    g->info.nnz_jac = P.nJ;
    g->info.nnz_lu = P.nnzLU;
    g->info.n_slots = P.nvals;
    g->info.n_levels = P.nlev;
    g->info.n_back_levels = P.nblev;
    g->info.n_pairs = (int64_t)P.pairs.size() / 4;
    g->info.flops = P.flops;
    g->yr_ptr.upload(P.yr_ptr); g->yr_col.upload(P.yr_col); g->yr_pos.upload(P.yr_pos);
    {
        std::vector<int> pc(2 * P.yr_pos.size());
        for (size_t k = 0; k < P.yr_pos.size(); ++k) { pc[2 * k] = P.yr_pos[k]; pc[2 * k + 1] = P.yr_col[k]; }
        g->yr_pc.upload(pc.empty() ? std::vector<int>(2, 0) : pc);
        // column-major row table, width the longest row rounded up to kEll columns
        const int n = (int)P.yr_ptr.size() - 1;
        int w = 1;
        for (int r = 0; r < n; ++r) w = std::max(w, P.yr_ptr[r + 1] - P.yr_ptr[r]);
        w = (w + kEll - 1) / kEll * kEll;
        g->ell_w = w;
        std::vector<int> ell(2 * (size_t)w * std::max(n, 1), 0);
        for (int r = 0; r < n; ++r)
            for (int k = 0; k < w; ++k) {
                const int q = P.yr_ptr[r] + k;
                const bool in = q < P.yr_ptr[r + 1];
                ell[2 * ((size_t)k * n + r)] = in ? P.yr_pos[q] : -1;
                ell[2 * ((size_t)k * n + r) + 1] = in ? P.yr_col[q] : r;
            }
        g->yr_ell.upload(ell);
    }
    auto up_or = [](auto& d, const auto& v, size_t minsz) {
        auto w = v;
        if (w.size() < minsz) w.resize(minsz);
        d.upload(w);
    };
    up_or(g->ye_pos, P.ye_pos, 1); up_or(g->ye_rc, P.ye_rc, 2); up_or(g->ye_slot, P.ye_slot, 4);
    up_or(g->ye_jidx, P.ye_jidx, 4); up_or(g->ye_diag_of, P.ye_diag_of, 1); up_or(g->jc_slot, P.jc_slot, 1); up_or(g->jc_jidx, P.jc_jidx, 1);
    up_or(g->jc_val, P.jc_val, 1);
    g->fslot_p.upload(P.fslot_p); g->fslot_q.upload(P.fslot_q);
    g->ux_bus.upload(P.ux_bus); g->ux_slot.upload(P.ux_slot); g->ux_kind.upload(P.ux_kind);
    {
        std::vector<int> bd(2 * (size_t)std::max(P.n, 1), -1);
        for (size_t j = 0; j < P.ux_bus.size(); ++j) {
            if (P.ux_kind[j] == 0) g->slack_dx = P.ux_slot[j];
            else bd[2 * (size_t)P.ux_bus[j] + (P.ux_kind[j] == 1 ? 0 : 1)] = P.ux_slot[j];
        }
        g->bus_dx.upload(bd);
    }
    g->j_ypos.upload(P.j_ypos); g->j_rbus.upload(P.j_rbus); g->j_cbus.upload(P.j_cbus); g->j_sel.upload(P.j_sel);
    g->j_cval.upload(P.j_cval); g->j_slot.upload(P.j_slot);
    g->zero_slots.upload(P.zero_slots.empty() ? std::vector<int>(1, 0) : P.zero_slots);
    g->w.upload(g->w_h);
}

// The J entries of a fast-decoupled grid become constants: B'(row bus, column bus) on the P-row /
// angle-column block, B''(row bus, column bus) on the Q-row / magnitude-column block, 0 on the two
// coupling blocks (fill_sparse_matrices, BaseFDPFAlgo.tpp:164-214: Bp = grid_Bp[pvpq, pvpq], Bpp =
// grid_Bpp[pq, pq]). Every entry of those restrictions must lie in the Ybus pattern.
void fdpf_constants(pfgpu::JPattern& J, int n, const FdpfB& fd) {
    const int npvpq = (int)J.pvpq.size(), N = J.N;
    auto bus = [&](int k) { return k < npvpq ? J.pvpq[k] : J.pq[k - npvpq]; };
    auto lookup = [&](const int* Ap, const int* Ai, const double* Ax, int r, int c) -> double {
        double v = 0.0;  // duplicates are summed, as setFromTriplets does
        for (int q = Ap[c]; q < Ap[c + 1]; ++q)
            if (Ai[q] == r) v += Ax[q];
        return v;
    };
    for (int c = 0; c < N; ++c)
        for (int q = J.Jp[c]; q < J.Jp[c + 1]; ++q) {
            const int r = J.Ji[q];
            const bool prow = r < npvpq, acol = c < npvpq;
            double v = 0.0;
            if (prow && acol) v = lookup(fd.Bp_p, fd.Bp_i, fd.Bp_x, bus(r), bus(c));
            if (!prow && !acol) v = lookup(fd.Bpp_p, fd.Bpp_i, fd.Bpp_x, bus(r), bus(c));
            J.sel[q] = pfgpu::SEL_CONST;
            J.src[q] = -1;
            J.cval[q] = v;
        }
    // every stored entry of the restrictions has a J position
    std::vector<int> kpos(n, -1);  // bus -> J index of its angle (P rows) / magnitude (Q rows)
    std::vector<long long> keys;
    for (int c = 0; c < N; ++c)
        for (int q = J.Jp[c]; q < J.Jp[c + 1]; ++q) keys.push_back((long long)J.Ji[q] * N + c);
    std::sort(keys.begin(), keys.end());
    auto check = [&](const int* Ap, const int* Ai, int k0, int k1) {
        std::fill(kpos.begin(), kpos.end(), -1);
        for (int k = k0; k < k1; ++k) kpos[bus(k)] = k;
        for (int c = 0; c < n; ++c) {
            if (kpos[c] < 0) continue;
            for (int q = Ap[c]; q < Ap[c + 1]; ++q) {
                if (Ai[q] < 0 || Ai[q] >= n) throw ApiError(PFGPU_E_INVALID, "B' / B'' row index out of range");
                if (kpos[Ai[q]] < 0) continue;
                if (!std::binary_search(keys.begin(), keys.end(), (long long)kpos[Ai[q]] * N + kpos[c]))
                    throw ApiError(PFGPU_E_INVALID, "B' / B'' hold an entry outside the Ybus pattern");
            }
        }
    };
    check(fd.Bp_p, fd.Bp_i, 0, npvpq);
    check(fd.Bpp_p, fd.Bpp_i, npvpq, N);
}

int load_grid(pfgpu_ctx* ctx, int n, const int* Yp, const int* Yi, const int* slack_ids, int nslack,
              const double* slack_weights, const int* pv, int npv, const int* pq, int npq, int single_slack, bool dc,
              pfgpu_grid** out, const FdpfB* fd) {
    return guarded([&]() -> int {
        if (!ctx || !out || !Yp || !Yi || !slack_ids || !slack_weights || n <= 0 || nslack < 1 || npv < 0 || npq < 0 ||
            (npv > 0 && !pv) || (npq > 0 && !pq))
            throw ApiError(PFGPU_E_INVALID, "invalid grid arguments");
        *out = nullptr;
        HIP_OK(hipSetDevice(ctx->device));
        auto g = std::make_unique<pfgpu_grid>();
        g->ctx = ctx;
        g->Y.n = n;
        g->Y.p.assign(Yp, Yp + n + 1);
        if (g->Y.p[0] != 0) throw ApiError(PFGPU_E_INVALID, "Yp[0] must be 0");
        for (int c = 0; c < n; ++c)
            if (g->Y.p[c + 1] < g->Y.p[c]) throw ApiError(PFGPU_E_INVALID, "Yp must be non-decreasing");
        g->Y.i.assign(Yi, Yi + Yp[n]);
        for (int c = 0; c < n; ++c)
            for (int q = Yp[c]; q < Yp[c + 1]; ++q) {
                if (Yi[q] < 0 || Yi[q] >= n) throw ApiError(PFGPU_E_INVALID, "Ybus row index out of range");
                if (q > Yp[c] && Yi[q] <= Yi[q - 1]) throw ApiError(PFGPU_E_INVALID, "Ybus rows must be sorted and unique");
            }
        for (int c = 0; c < n; ++c) {
            bool diag = false;
            for (int q = Yp[c]; q < Yp[c + 1]; ++q) diag |= (Yi[q] == c);
            if (!diag) throw ApiError(PFGPU_E_INVALID, "Ybus must hold every diagonal entry");
        }
        for (int k = 0; k < nslack; ++k)
            if (slack_ids[k] < 0 || slack_ids[k] >= n) throw ApiError(PFGPU_E_INVALID, "slack bus out of range");
        for (int k = 0; k < npq; ++k)
            if (pq[k] < 0 || pq[k] >= n) throw ApiError(PFGPU_E_INVALID, "pq bus out of range");
        for (int k = 0; k < npv; ++k)
            if (pv[k] < 0 || pv[k] >= n) throw ApiError(PFGPU_E_INVALID, "pv bus out of range");
        g->J = pfgpu::build_jacobian_pattern(g->Y, slack_ids, nslack, slack_weights, pv, npv, pq, npq, single_slack != 0);
        if (fd) fdpf_constants(g->J, n, *fd);
        g->w_h.assign(slack_weights, slack_weights + n);
        g->info.single_slack = single_slack ? 1 : 0;
        g->info.dc = dc ? 1 : 0;
        g->info.fdpf = fd ? 1 : 0;
        g->dc = dc;
        g->fdpf = fd != nullptr;
        build_plan(g.get(), nullptr);
        *out = g.release();
        return 0;
    });
}
}  // namespace

int pfgpu_grid_destroy(pfgpu_grid* grid) {
    return guarded([&]() -> int {
        if (!grid) return 0;
        (void)hipSetDevice(grid->ctx->device);
        delete grid;
        return 0;
    });
}

int pfgpu_grid_get_info(const pfgpu_grid* grid, pfgpu_grid_info* info) {
    return guarded([&]() -> int {
        if (!grid || !info) throw ApiError(PFGPU_E_INVALID, "NULL argument");
        *info = grid->info;
        return 0;
    });
}

int pfgpu_grid_get_J_pattern(const pfgpu_grid* grid, int* Jp, int* Ji) {
    return guarded([&]() -> int {
        if (!grid || !Jp || !Ji) throw ApiError(PFGPU_E_INVALID, "NULL argument");
        std::copy(grid->J.Jp.begin(), grid->J.Jp.end(), Jp);
        std::copy(grid->J.Ji.begin(), grid->J.Ji.end(), Ji);
        return 0;
    });
}

int pfgpu_grid_reset(pfgpu_grid* grid) {
    return guarded([&]() -> int {
        if (!grid) throw ApiError(PFGPU_E_INVALID, "grid is NULL");
        // the next solve factors afresh: pivots are chosen again (the re-pivoted plan is dropped)
        HIP_OK(hipSetDevice(grid->ctx->device));
        HIP_OK(hipStreamSynchronize(grid->ctx->stream));
        grid->fresh = true;
        grid->alt_active = false;
        grid->alt.reset();
        grid->order_n = 0;  // and the next batch takes its instances in index order
        return 0;
    });
}

int pfgpu_autotune_device(pfgpu_ctx* ctx, pfgpu_grid* grid, int ninst, const double* d_Yx, int Yx_stride,
                          const int* d_dptr, const int* d_dpos, const double* d_dval, const double* d_Sbus,
                          int sbus_stride, const double* d_V0, int v0_stride, double* d_V, int max_iter, double tol,
                          int precision, int* d_iters, int* d_status, int* best_G, int* best_block,
                          double* best_seconds) {
    return guarded([&]() -> int {
        if (!ctx || !grid || !d_Yx || !d_Sbus || !d_V0 || !d_V || !d_iters || !d_status || ninst <= 0)
            throw ApiError(PFGPU_E_INVALID, "invalid autotune arguments");
        if (precision != PFGPU_FP64 && precision != PFGPU_FP32 && precision != PFGPU_FP32_STRICT)
            throw ApiError(PFGPU_E_INVALID, "bad precision");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        const int fsize = precision == PFGPU_FP64 ? 8 : 4;
        const int pslot = precision == PFGPU_FP64 ? 0 : 1;
        double best = 1e300;
        int bG = 1, bNT = 256;
        SolveArgs q;
        q.ninst = ninst; q.nrows = ninst; q.Yx = d_Yx; q.Yx_stride = Yx_stride; q.dptr = d_dptr; q.dpos = d_dpos;
        q.dval = d_dval; q.S = d_Sbus; q.sbus_stride = sbus_stride; q.V0 = d_V0; q.v0_stride = v0_stride; q.V = d_V;
        q.max_iter = max_iter; q.tol = tol; q.iters = d_iters; q.status = d_status;
        // the kernel alone is timed: no FP64 re-solve pass
        q.precision = precision == PFGPU_FP64 ? PFGPU_FP64 : PFGPU_FP32_STRICT;
        pfgpu_grid* pg = grid->alt_active ? grid->alt.get() : grid;  // the plan in use
        for (int G = 1; G <= kMaxG; G <<= 1) {
            for (int NT : {256, 1024}) {
                LaunchChoice lc = choose_launch(ctx, pg->P, ninst, fsize, G, NT);
                if (lc.G != G) continue;  // the dense tail did not fit at this G
                q.G_req = G;
                q.NT_req = NT;
                double tmin = 1e300;
                for (int rep = 0; rep < 2; ++rep) {
                    pfgpu_stats st{};
                    solve_device(ctx, grid, q, &st);
                    tmin = std::min(tmin, st.kernel_s);
                }
                if (tmin < best) { best = tmin; bG = G; bNT = NT; }
            }
        }
        pg = grid->alt_active ? grid->alt.get() : grid;
        pg->tuned_G[pslot] = bG;
        pg->tuned_NT[pslot] = bNT;
        if (best_G) *best_G = bG;
        if (best_block) *best_block = bNT;
        if (best_seconds) *best_seconds = best;
        return 0;
    });
}

int pfgpu_branch_flows_device(pfgpu_ctx* ctx, int ninst, int n, const double* d_V, int nbr, const int* d_f,
                              const int* d_t, const double* d_yff, const double* d_yft, const double* d_vn_kv_from,
                              double sn_mva, double* d_amps, double* d_p) {
    return guarded([&]() -> int {
        if (!ctx || ninst < 0 || n <= 0 || nbr < 0 || (nbr > 0 && (!d_V || !d_f || !d_t || !d_yff || !d_yft || !d_vn_kv_from)))
            throw ApiError(PFGPU_E_INVALID, "invalid branch-flow arguments");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        const long long total = (long long)ninst * nbr;
        if (total == 0 || (!d_amps && !d_p)) return 0;
        const int blocks = (int)std::min<long long>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(k_branch_flows, dim3(blocks), dim3(256), 0, ctx->stream, ninst, n, d_V, nbr, d_f, d_t, d_yff,
                           d_yft, d_vn_kv_from, sn_mva, d_amps, d_p);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int pfgpu_dev_alloc(pfgpu_ctx* ctx, size_t bytes, void** dptr) {
    return guarded([&]() -> int {
        if (!ctx || !dptr) throw ApiError(PFGPU_E_INVALID, "NULL argument");
        HIP_OK(hipSetDevice(ctx->device));
        *dptr = nullptr;
        HIP_OK(dev_malloc(dptr, std::max<size_t>(bytes, 1)));
        return 0;
    });
}
int pfgpu_dev_free(pfgpu_ctx* ctx, void* dptr) {
    return guarded([&]() -> int {
        if (!ctx) throw ApiError(PFGPU_E_INVALID, "NULL ctx");
        HIP_OK(hipSetDevice(ctx->device));
        if (dptr) HIP_OK(dev_free(dptr));
        return 0;
    });
}
int pfgpu_memcpy_h2d(pfgpu_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&]() -> int {
        if (!ctx) throw ApiError(PFGPU_E_INVALID, "NULL ctx");
        HIP_OK(hipSetDevice(ctx->device));
        if (bytes) HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}
int pfgpu_memcpy_d2h(pfgpu_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&]() -> int {
        if (!ctx) throw ApiError(PFGPU_E_INVALID, "NULL ctx");
        HIP_OK(hipSetDevice(ctx->device));
        if (bytes) HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int pfgpu_memcpy_d2d(pfgpu_ctx* ctx, void* dst, const void* src, size_t bytes) {
    return guarded([&]() -> int {
        if (!ctx) throw ApiError(PFGPU_E_INVALID, "NULL ctx");
        HIP_OK(hipSetDevice(ctx->device));
        if (bytes) HIP_OK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int pfgpu_ctx_last_flags(pfgpu_ctx* ctx, int* flags, long long nrows) {
    return guarded([&]() -> int {
        if (!ctx || (!flags && nrows > 0) || nrows < 0) throw ApiError(PFGPU_E_INVALID, "invalid arguments");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        if (nrows > ctx->last_rows) throw ApiError(PFGPU_E_INVALID, "more rows than the last solve call had");
        if (nrows == 0) return 0;
        HIP_OK(hipMemcpyAsync(flags, ctx->last_flags, (size_t)nrows * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

namespace {
void check_deltas_host(const int* dptr, const int* dpos, int ninst, int nnzY) {
    if (dptr[0] != 0) throw ApiError(PFGPU_E_INVALID, "dptr[0] must be 0");
    for (int i = 0; i < ninst; ++i)
        if (dptr[i + 1] < dptr[i]) throw ApiError(PFGPU_E_INVALID, "dptr must be non-decreasing");
    for (int k = 0; k < dptr[ninst]; ++k)
        if (dpos[k] < 0 || dpos[k] >= nnzY) throw ApiError(PFGPU_E_INVALID, "delta position out of range");
}
}  // namespace

int pfgpu_solve_batch_device(pfgpu_ctx* ctx, pfgpu_grid* grid, int ninst, const double* d_Yx, int Yx_stride,
                             const int* d_dptr, const int* d_dpos, const double* d_dval, const double* d_Sbus,
                             int sbus_stride, const double* d_V0, int v0_stride, double* d_V, int max_iter, double tol,
                             int precision, int instances_per_wg, int* d_iters, int* d_status, pfgpu_stats* stats) {
    return guarded([&]() -> int {
        if (!ctx || !grid || !d_Yx || !d_Sbus || !d_V0 || !d_V || !d_iters || !d_status || ninst < 0)
            throw ApiError(PFGPU_E_INVALID, "invalid solve arguments");
        if ((d_dptr == nullptr) != (d_dpos == nullptr) || (d_dptr == nullptr) != (d_dval == nullptr))
            throw ApiError(PFGPU_E_INVALID, "dptr/dpos/dval must be all set or all NULL");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        auto t0 = std::chrono::steady_clock::now();
        if (stats) *stats = pfgpu_stats{};
        SolveArgs q;
        q.ninst = ninst; q.nrows = ninst; q.Yx = d_Yx; q.Yx_stride = Yx_stride; q.dptr = d_dptr; q.dpos = d_dpos;
        q.dval = d_dval; q.S = d_Sbus; q.sbus_stride = sbus_stride; q.V0 = d_V0; q.v0_stride = v0_stride; q.V = d_V;
        q.max_iter = max_iter; q.tol = tol; q.precision = precision; q.G_req = instances_per_wg;
        q.iters = d_iters; q.status = d_status;
        solve_device(ctx, grid, q, stats);
        if (stats) stats->total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    });
}

int pfgpu_solve_batch_device_noalloc(pfgpu_ctx* ctx, pfgpu_grid* grid, int ninst, const double* d_Yx, int Yx_stride,
                                     const int* d_dptr, const int* d_dpos, const double* d_dval,
                                     const double* d_Sbus, int sbus_stride, const double* d_V0, int v0_stride,
                                     double* d_V, int max_iter, double tol, int precision, int instances_per_wg,
                                     int* d_iters, int* d_status, pfgpu_stats* stats) {
    return guarded([&]() -> int {
        if (!ctx || !grid || !d_Yx || !d_Sbus || !d_V0 || !d_V || !d_iters || !d_status || ninst < 0)
            throw ApiError(PFGPU_E_INVALID, "invalid solve arguments");
        if ((d_dptr == nullptr) != (d_dpos == nullptr) || (d_dptr == nullptr) != (d_dval == nullptr))
            throw ApiError(PFGPU_E_INVALID, "dptr/dpos/dval must be all set or all NULL");
        if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch_ptr[ctx->device])
            throw ApiError(PFGPU_E_DEVICE, "no scratch address for this device" +
                                               (ctx->device >= 0 && ctx->device < kMaxDevices
                                                    ? ": " + g_scratch_err[ctx->device] : std::string()));
        std::lock_guard<std::mutex> lk_dev(g_noalloc_mu[ctx->device]);
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        auto t0 = std::chrono::steady_clock::now();
        if (stats) *stats = pfgpu_stats{};
        if (ninst == 0) return 0;
        SolveArgs q;
        q.ninst = ninst; q.nrows = ninst; q.Yx = d_Yx; q.Yx_stride = Yx_stride; q.dptr = d_dptr; q.dpos = d_dpos;
        q.dval = d_dval; q.S = d_Sbus; q.sbus_stride = sbus_stride; q.V0 = d_V0; q.v0_stride = v0_stride; q.V = d_V;
        q.max_iter = max_iter; q.tol = tol; q.precision = precision; q.G_req = instances_per_wg;
        q.iters = d_iters; q.status = d_status;
        // scratch layout: per-call rows, the kept initial voltages (in-place FP32 only), the launch area
        const AuxLayout AL = aux_layout(ninst, ninst);
        const bool keep = precision == PFGPU_FP32 && d_V0 == d_V && v0_stride != 0;
        const size_t io_bytes = keep ? align_up((size_t)ninst * grid->P.n * 16) : 0;
        if (AL.bytes + io_bytes + 4096 > kScratchBytes)
            throw ApiError(PFGPU_E_BUDGET, "the batch does not fit the 64 MiB no-allocation scratch: fewer instances");
        char* base = static_cast<char*>(g_scratch_ptr[ctx->device]);
        FixedScratch fx;
        fx.aux = base;
        fx.io = keep ? base + AL.bytes : nullptr;
        fx.ws = base + AL.bytes + io_bytes;
        fx.ws_bytes = kScratchBytes - AL.bytes - io_bytes;
        solve_device(ctx, grid, q, stats, &fx);
        if (stats) stats->total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    });
}

int pfgpu_solve_chains_device(pfgpu_ctx* ctx, pfgpu_grid* grid, int nrows, int chain_len, const double* d_Yx,
                              int Yx_stride, const double* d_Sbus, const double* d_V0, int v0_stride, double* d_V,
                              int max_iter, double tol, int precision, int instances_per_wg, int* d_iters,
                              int* d_status, pfgpu_stats* stats) {
    return guarded([&]() -> int {
        if (!ctx || !grid || !d_Yx || !d_Sbus || !d_V0 || !d_V || !d_iters || !d_status || nrows < 0 || chain_len < 1)
            throw ApiError(PFGPU_E_INVALID, "invalid chain arguments");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        auto t0 = std::chrono::steady_clock::now();
        if (stats) *stats = pfgpu_stats{};
        SolveArgs q;
        q.chain = std::min(chain_len, std::max(nrows, 1));
        q.ninst = (nrows + q.chain - 1) / q.chain;
        q.nrows = nrows;
        q.Yx = d_Yx; q.Yx_stride = Yx_stride; q.S = d_Sbus; q.sbus_stride = grid->P.n; q.V0 = d_V0;
        q.v0_stride = v0_stride; q.V = d_V; q.max_iter = max_iter; q.tol = tol; q.precision = precision;
        q.G_req = instances_per_wg; q.iters = d_iters; q.status = d_status;
        solve_call(ctx, grid, q, stats);
        if (stats) stats->total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    });
}

namespace {
// Host-buffer solve: the caller's arrays are copied into the context's staging buffer, solved
// there, and the results copied back.
int solve_host(pfgpu_ctx* ctx, pfgpu_grid* grid, int ninst, int chain, long long nrows, const double* Yx,
               int Yx_stride, const int* dptr, const int* dpos, const double* dval, const double* Sbus, int sbus_stride,
               const double* V0, int v0_stride, double* V, int max_iter, double tol, int precision,
               int instances_per_wg, int* iters, int* status, double* Jx_last, pfgpu_stats* stats) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    pfgpu_stats st{};
    const pfgpu::DevicePlan& P = grid->P;
    const int n = P.n;
    const int nd = dptr ? dptr[ninst] : 0;
    const size_t ny = (size_t)(Yx_stride ? ninst : 1) * P.nnzY * 16;
    const size_t ns = (size_t)(sbus_stride ? nrows : 1) * n * 16;
    const size_t nv0 = (size_t)(v0_stride ? ninst : 1) * n * 16;
    const size_t nv = (size_t)nrows * n * 16;
    size_t off = 0;
    const size_t o_y = off; off += align_up(ny);
    const size_t o_s = off; off += align_up(ns);
    const size_t o_v0 = off; off += align_up(nv0);
    const size_t o_v = off; off += align_up(nv);
    const size_t o_it = off; off += align_up((size_t)nrows * 4);
    const size_t o_st = off; off += align_up((size_t)nrows * 4);
    const size_t o_j = off; off += Jx_last ? align_up((size_t)P.nJ * 8) : 0;
    const size_t o_dp = off; off += dptr ? align_up((size_t)(ninst + 1) * 4) : 0;
    const size_t o_dq = off; off += dptr ? align_up((size_t)std::max(nd, 1) * 4) : 0;
    const size_t o_dv = off; off += dptr ? align_up((size_t)std::max(nd, 1) * 16) : 0;
    char* b = static_cast<char*>(grow_buffer(ctx->stage, ctx->stage_bytes, std::max<size_t>(off, 256)));
    hipStream_t s = ctx->stream;
    HIP_OK(hipMemcpyAsync(b + o_y, Yx, ny, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(b + o_s, Sbus, ns, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(b + o_v0, V0, nv0, hipMemcpyHostToDevice, s));
    if (dptr) {
        HIP_OK(hipMemcpyAsync(b + o_dp, dptr, (size_t)(ninst + 1) * 4, hipMemcpyHostToDevice, s));
        if (nd) {
            HIP_OK(hipMemcpyAsync(b + o_dq, dpos, (size_t)nd * 4, hipMemcpyHostToDevice, s));
            HIP_OK(hipMemcpyAsync(b + o_dv, dval, (size_t)nd * 16, hipMemcpyHostToDevice, s));
        }
    }
    HIP_OK(hipStreamSynchronize(s));
    auto t1 = clk::now();
    st.h2d_s = std::chrono::duration<double>(t1 - t0).count();
    SolveArgs q;
    q.ninst = ninst; q.chain = chain; q.nrows = nrows;
    q.Yx = reinterpret_cast<const double*>(b + o_y); q.Yx_stride = Yx_stride;
    if (dptr) {
        q.dptr = reinterpret_cast<const int*>(b + o_dp);
        q.dpos = reinterpret_cast<const int*>(b + o_dq);
        q.dval = reinterpret_cast<const double*>(b + o_dv);
    }
    q.S = reinterpret_cast<const double*>(b + o_s); q.sbus_stride = sbus_stride;
    q.V0 = reinterpret_cast<const double*>(b + o_v0); q.v0_stride = v0_stride;
    q.V = reinterpret_cast<double*>(b + o_v);
    q.max_iter = max_iter; q.tol = tol; q.precision = precision; q.G_req = instances_per_wg;
    q.iters = reinterpret_cast<int*>(b + o_it); q.status = reinterpret_cast<int*>(b + o_st);
    q.Jx_last = Jx_last ? reinterpret_cast<double*>(b + o_j) : nullptr;
    solve_call(ctx, grid, q, &st);
    auto t2 = clk::now();
    HIP_OK(hipMemcpyAsync(V, b + o_v, nv, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(iters, b + o_it, (size_t)nrows * 4, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(status, b + o_st, (size_t)nrows * 4, hipMemcpyDeviceToHost, s));
    if (Jx_last) HIP_OK(hipMemcpyAsync(Jx_last, b + o_j, (size_t)P.nJ * 8, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    st.d2h_s = std::chrono::duration<double>(clk::now() - t2).count();
    st.total_s = std::chrono::duration<double>(clk::now() - t0).count();
    if (stats) *stats = st;
    return 0;
}
}  // namespace

int pfgpu_solve_batch(pfgpu_ctx* ctx, pfgpu_grid* grid, int ninst, const double* Yx, int Yx_stride, const int* dptr,
                      const int* dpos, const double* dval, const double* Sbus, int sbus_stride, double* V, int v0_stride,
                      int max_iter, double tol, int precision, int instances_per_wg, int* iters, int* status,
                      double* Jx_last, pfgpu_stats* stats) {
    return guarded([&]() -> int {
        if (!ctx || !grid || !Yx || !Sbus || !V || !iters || !status || ninst < 0)
            throw ApiError(PFGPU_E_INVALID, "invalid solve arguments");
        if ((dptr == nullptr) != (dpos == nullptr) || (dptr == nullptr) != (dval == nullptr))
            throw ApiError(PFGPU_E_INVALID, "dptr/dpos/dval must be all set or all NULL");
        if (stats) *stats = pfgpu_stats{};
        if (ninst == 0) return 0;
        if (dptr) check_deltas_host(dptr, dpos, ninst, grid->P.nnzY);
        return solve_host(ctx, grid, ninst, 1, ninst, Yx, Yx_stride, dptr, dpos, dval, Sbus, sbus_stride, V, v0_stride,
                          V, max_iter, tol, precision, instances_per_wg, iters, status, Jx_last, stats);
    });
}

int pfgpu_solve_chains(pfgpu_ctx* ctx, pfgpu_grid* grid, int nrows, int chain_len, const double* Yx, const double* Sbus,
                       const double* V0, int v0_stride, double* V, int max_iter, double tol, int precision,
                       int instances_per_wg, int* iters, int* status, pfgpu_stats* stats) {
    return guarded([&]() -> int {
        if (!ctx || !grid || !Yx || !Sbus || !V0 || !V || !iters || !status || nrows < 0 || chain_len < 1)
            throw ApiError(PFGPU_E_INVALID, "invalid chain arguments");
        if (v0_stride != 0 && v0_stride != grid->P.n) throw ApiError(PFGPU_E_INVALID, "v0_stride must be 0 or n");
        if (stats) *stats = pfgpu_stats{};
        if (nrows == 0) return 0;
        const int L = std::min(chain_len, nrows);
        const int nch = (nrows + L - 1) / L;
        return solve_host(ctx, grid, nch, L, nrows, Yx, 0, nullptr, nullptr, nullptr, Sbus, grid->P.n, V0, v0_stride, V,
                          max_iter, tol, precision, instances_per_wg, iters, status, nullptr, stats);
    });
}

}  // extern "C"
