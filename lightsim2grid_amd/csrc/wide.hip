// Wide Newton-Raphson kernel: all resident workgroups solve one instance at a time (wide.hpp).
//
// Per NR iteration (BaseNRAlgo.tpp:106-167), each phase spread over the whole grid and closed by
// a grid-wide barrier:
//   J fill     one lane per J entry (dS_dVa / dS_dVm through the scatter map, BaseNRAlgo.tpp:202-260)
//   factor     level by level: the level's Crout tasks (entries of the static-pivot LU and of the
//              forward-solved right-hand side) split over the grid, a team of 2^t lanes per task
//              sharing its dot product (klu_refactor, KLUSolver.cpp:47-59)
//   tail       the dense m x m tail on workgroup 0, in LDS (right-looking LU, forward, backward)
//   backward   level by level, x_j = (y_j - sum u_ji x_i) / d_j (klu_solve)
//   update     one lane per bus (BaseNRAlgo.tpp:138-158)
//   mismatch   one lane per bus: Ibus, F, max |F| (BaseAlgo.cpp:56-129)
// The value slots, the voltages and the instance's Ybus values live in global memory; a grid
// barrier (cooperative launch) makes each phase's writes visible to the next. The grid's reduced
// values (max |F|, sums) are combined in a fixed order, so every workgroup takes the same
// decisions and the results do not depend on timing.
#include <hip/hip_runtime.h>

#include <cmath>

#include "wide.hpp"

namespace pfgpu_wide {
namespace {

constexpr int kWaves = kThreads / 64;
constexpr int kFlagStall = 1;
constexpr int kStatusNotRun = 8;

struct c64 {
    double re, im;
};
__device__ __forceinline__ c64 cmul(c64 a, c64 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ c64 cconj(c64 a) { return {a.re, -a.im}; }
__device__ __forceinline__ c64 csub(c64 a, c64 b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ c64 cadd(c64 a, c64 b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ c64 ld_c(const double* p, long long i) { return {p[2 * i], p[2 * i + 1]}; }
__device__ __forceinline__ double inf_if_nonfinite(double a) { return (a <= 1.7976931348623157e308) ? a : INFINITY; }

__device__ __forceinline__ float pivot_recip(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ __forceinline__ double pivot_recip(double v) { return 1.0 / v; }
template <typename FT>
__device__ __forceinline__ bool bad_pivot(FT v) { return !(fabs(v) > (FT)0 && fabs(v) < (FT)INFINITY); }

// Grid-wide barrier in two levels: workgroup k arrives on the counter of group k % 8, the last
// arrival of a group on the top counter, and the last of those advances the generation word that
// every workgroup waits on (the arrivals are device-scope acquire-release atomics, the wait an
// acquire load, so the writes before the barrier are visible after it). Spreading the arrivals
// over eight counters halves the cost of one counter per grid (about 3 us instead of 5 at 32
// workgroups). Every workgroup of the cooperative launch is resident, so each wait ends; the
// counters are back at 0 after each barrier.
struct GridBarrier {
    unsigned int* w;  // [8 group counters | top | generation], 32 words apart (zeroed before the launch)
    unsigned int gen;
    int nb;

    __device__ void sync() {
        __syncthreads();
        if (threadIdx.x == 0) {
            const int grp = blockIdx.x & 7;
            const unsigned int gsize = (unsigned int)((nb - grp + 7) / 8);
            const unsigned int ng = (unsigned int)(nb < 8 ? nb : 8);
            unsigned int* top = w + 8 * 32;
            unsigned int* genw = w + 9 * 32;
            const unsigned int old = __hip_atomic_fetch_add(w + grp * 32, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (old == gsize - 1) {
                __hip_atomic_store(w + grp * 32, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned int t = __hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                if (t == ng - 1) {
                    __hip_atomic_store(top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(genw, gen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            while (__hip_atomic_load(genw, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) __builtin_amdgcn_s_sleep(1);
        }
        ++gen;  // every lane keeps the count (uniform)
        __syncthreads();
    }
};

// Reduction over the grid: the workgroup's lanes, then one partial per workgroup into red[],
// a grid barrier, then every workgroup combines the partials in index order. Two buffers
// alternate so that a workgroup never overwrites partials another one may still read.
struct GridReduce {
    double* red;
    int nb;
    int par;
    double* s;  // LDS: kWaves + 1 doubles

    template <bool MAX>
    __device__ double run(double v, GridBarrier& grid) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int o = 32; o > 0; o >>= 1) {
            const double t = __shfl_xor(v, o);
            v = MAX ? fmax(v, t) : v + t;
        }
        if (lane == 0) s[wave] = v;
        __syncthreads();
        double* buf = red + (size_t)par * nb;
        // the partials and the control words are written and read with device-scope atomics: every
        // workgroup must read the same values, since they decide the control flow around the
        // grid barriers
        if (threadIdx.x == 0) {
            double t = s[0];
            for (int w = 1; w < kWaves; ++w) t = MAX ? fmax(t, s[w]) : t + s[w];
            __hip_atomic_store(&buf[blockIdx.x], t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        grid.sync();
        if (wave == 0) {
            double t = MAX ? -INFINITY : 0.0;
            for (int i = lane; i < nb; i += 64) {
                const double b = __hip_atomic_load(&buf[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                t = MAX ? fmax(t, b) : t + b;
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double u = __shfl_xor(t, o);
                t = MAX ? fmax(t, u) : t + u;
            }
            if (lane == 0) s[kWaves] = t;
        }
        __syncthreads();
        const double r = s[kWaves];
        __syncthreads();
        par ^= 1;
        return r;
    }
};

// One factor / forward level: the level's tasks in turn over the grid's teams of 2^lt lanes.
template <typename FT>
__device__ void factor_level(const WArgs& a, FT* X, int l, long long gt, long long GT, int* pflag) {
    const int lt = a.f_team[l], T = 1 << lt;
    const int me = (int)(gt & (T - 1));
    const int t1 = a.f_lvl[l + 1];
    for (long long t = a.f_lvl[l] + (gt >> lt); t < t1; t += GT >> lt) {
        const int4 tk = a.tasks[t];
        FT acc = (FT)0;
        for (int p = tk.y + me; p < tk.z; p += T) {
            const int4 pr = a.pairs[p];
            acc += X[pr.x] * X[pr.y] * X[pr.z];
        }
        for (int o = T >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        if (me == 0) {
            const FT v = ((tk.w & 4) ? (FT)0 : X[tk.x]) - acc;
            if ((tk.w & 3) == 1) {
                if (bad_pivot(v)) atomicOr(pflag, 1);
                X[tk.x] = pivot_recip(v);
            } else {
                X[tk.x] = v;
            }
        }
    }
}

// One backward level: x_j = (x_j - sum u x') * (1/d_j).
template <typename FT>
__device__ void backward_level(const WArgs& a, FT* X, int l, long long gt, long long GT) {
    const int lt = a.b_team[l], T = 1 << lt;
    const int me = (int)(gt & (T - 1));
    const int t1 = a.b_lvl[l + 1];
    for (long long t = a.b_lvl[l] + (gt >> lt); t < t1; t += GT >> lt) {
        const int4 bt = a.btasks[t];
        FT acc = (FT)0;
        for (int p = bt.y + me; p < bt.z; p += T) {
            const int2 pr = a.bpairs[p];
            acc += X[pr.x] * X[pr.y];
        }
        for (int o = T >> 1; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
        if (me == 0) X[bt.x] = (X[bt.x] - acc) * X[bt.w];
    }
}

// Dense tail on workgroup 0: the m x m tile (slots [0, m (m + 1)), row stride m + 1, its vector
// in column m) factored right-looking without pivoting, forward on the vector, then backward; the
// tile diagonal ends up holding 1/d.
template <typename FT>
__device__ void dense_tail(const WArgs& a, FT* X, FT* Tl, int* pflag) {
    const int m = a.m, ld = m + 1, tid = threadIdx.x;
    for (int i = tid; i < m * ld; i += kThreads) Tl[i] = X[i];
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        const FT dk = Tl[k * ld + k];
        if (tid == 0 && bad_pivot(dk)) atomicOr(pflag, 1);
        const FT rd = pivot_recip(dk);
        const int w = m - k;  // columns k+1 .. m (the vector last)
        for (int idx = tid; idx < (m - 1 - k) * w; idx += kThreads) {
            const int i = k + 1 + idx / w, j = k + 1 + idx % w;
            Tl[i * ld + j] -= (Tl[i * ld + k] * rd) * Tl[k * ld + j];
        }
        __syncthreads();
        for (int i = k + 1 + tid; i < m; i += kThreads) Tl[i * ld + k] *= rd;
        if (tid == 0) Tl[k * ld + k] = rd;
        __syncthreads();
    }
    for (int k = m - 1; k >= 0; --k) {
        const FT x = Tl[k * ld + m] * Tl[k * ld + k];
        for (int i = tid; i < k; i += kThreads) Tl[i * ld + m] -= Tl[i * ld + k] * x;
        __syncthreads();
        if (tid == 0) Tl[k * ld + m] = x;
        __syncthreads();
    }
    for (int i = tid; i < m * ld; i += kThreads) X[i] = Tl[i];
}

template <typename FT>
__global__ void __launch_bounds__(kThreads) k_wide(const WArgs a) {
    GridBarrier grid{a.bar, 0u, (int)gridDim.x};
    extern __shared__ unsigned char lds_raw[];
    __shared__ double s_red[kWaves + 1];
    FT* const Tl = reinterpret_cast<FT*>(lds_raw);
    FT* const X = static_cast<FT*>(a.X);
    const long long GT = (long long)gridDim.x * kThreads;
    const long long gt = (long long)blockIdx.x * kThreads + threadIdx.x;
    const int n = a.n;
    GridReduce gr{a.red, (int)gridDim.x, 0, s_red};
    double* const Va = a.Va;
    double* const Vm = a.Vm;
    double* const Ib = a.Ib;
    double* const Yw = a.Yw;

    for (long long u = 0; u < a.nunits; ++u) {
        const long long oi = a.imap ? (long long)a.imap[u] : u;
        // this instance's Ybus values: shared or own row, minus its edits (in list order)
        {
            const double* Y = a.Yx + oi * a.y_stride;
            const int d0 = a.dptr ? a.dptr[oi] : 0, d1 = a.dptr ? a.dptr[oi + 1] : 0;
            const int nnzY = a.yr_ptr[n];
            for (long long p = gt; p < nnzY; p += GT) {
                double re = Y[2 * p], im = Y[2 * p + 1];
                for (int d = d0; d < d1; ++d)
                    if (a.dpos[d] == p) {
                        re -= a.dval[2 * d];
                        im -= a.dval[2 * d + 1];
                    }
                Yw[2 * p] = re;
                Yw[2 * p + 1] = im;
            }
        }
        bool alive = true;
        for (int step = 0; step < a.chain; ++step) {
            const long long row = oi * a.chain + step;
            if (row >= a.nrows) break;
            double* const V = a.V + row * 2 * (long long)n;
            const double* const S = a.S + row * a.s_stride;
            if (!alive) {  // a step after its chain's failure
                for (long long b = gt; b < n; b += GT) {
                    V[2 * b] = 0.0;
                    V[2 * b + 1] = 0.0;
                }
                if (gt == 0) {
                    a.iters[row] = 0;
                    a.status[row] = kStatusNotRun;
                    if (a.flags) a.flags[row] = 0;
                }
                continue;
            }
            // warm start from the previous step's V (each lane copies the buses it owns below)
            if (step > 0)
                for (long long b = gt; b < n; b += GT) {
                    V[2 * b] = V[2 * b - 2 * (long long)n];
                    V[2 * b + 1] = V[2 * b + 1 - 2 * (long long)n];
                }
            if (gt == 0) {
                __hip_atomic_store(&a.ctl[0], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&a.ctl[1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            // Va, Vm; slack_absorbed = Re(sum Sbus) (BaseNRAlgo.tpp:60, 77-79)
            double ssum = 0.0;
            for (long long b = gt; b < n; b += GT) {
                const double vr = V[2 * b], vi = V[2 * b + 1];
                Va[b] = atan2(vi, vr);
                Vm[b] = sqrt(vr * vr + vi * vi);
                ssum += S[2 * b];
            }
            double slack = gr.run<false>(ssum, grid);

            auto mismatch = [&]() -> double {
                double fm = 0.0;
                for (long long r = gt; r < n; r += GT) {
                    c64 acc = {0.0, 0.0};
                    for (int q = a.yr_ptr[r]; q < a.yr_ptr[r + 1]; ++q)
                        acc = cadd(acc, cmul(ld_c(Yw, a.yr_pos[q]), ld_c(V, a.yr_col[q])));
                    Ib[2 * r] = acc.re;
                    Ib[2 * r + 1] = acc.im;
                    c64 mis = csub(cmul(ld_c(V, r), cconj(acc)), ld_c(S, r));
                    if (!a.single_slack) mis.re += slack * a.w[r];
                    const int fp = a.fslot_p[r], fq = a.fslot_q[r];
                    if (fp >= 0) { X[fp] = (FT)mis.re; fm = fmax(fm, inf_if_nonfinite(fabs(mis.re))); }
                    if (fq >= 0) { X[fq] = (FT)mis.im; fm = fmax(fm, inf_if_nonfinite(fabs(mis.im))); }
                }
                return gr.run<true>(fm, grid);
            };

            double fm = mismatch();
            double flast = fm;
            int it = 0, st = 0;
            bool done = false;
            if (fm < a.tol) done = true;
            else if (a.max_iter <= 0) { st = 2; done = true; }
            while (!done) {
                ++it;
                const int fi = it & 1;
                if (gt == 0)  // the next iteration's pivot flag
                    __hip_atomic_store(&a.ctl[fi ^ 1], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                // ---- J fill -----------------------------------------------------------------
                for (long long i = gt; i < a.nzero; i += GT) X[a.zero_slots[i]] = (FT)0;
                for (long long q = gt; q < a.nJ; q += GT) {
                    const int sel = a.j_sel[q];
                    double jv;
                    if (sel == 4) {
                        jv = a.j_cval[q];
                    } else {
                        const int r = a.j_rbus[q], c = a.j_cbus[q];
                        const c64 y = ld_c(Yw, a.j_ypos[q]);
                        const c64 vr = ld_c(V, r), vc = ld_c(V, c);
                        if (sel >= 2) {  // dS/dVm: conj(Y V/|V|) V_r (+ conj(Ibus) V/|V| on the diagonal)
                            const double ic = 1.0 / sqrt(vc.re * vc.re + vc.im * vc.im);
                            const c64 vn = {vc.re * ic, vc.im * ic};
                            c64 d = cmul(cconj(cmul(y, vn)), vr);
                            if (r == c) d = cadd(d, cmul(cconj(ld_c(Ib, r)), vn));
                            jv = sel == 2 ? d.re : d.im;
                        } else {  // dS/dVa: conj(-(Y V_c - Ibus on the diagonal)) * j V_r
                            c64 t = cmul(y, vc);
                            if (r == c) t = csub(t, ld_c(Ib, r));
                            const c64 d = cmul(cconj({-t.re, -t.im}), {-vr.im, vr.re});
                            jv = sel == 0 ? d.re : d.im;
                        }
                    }
                    const int js = a.j_slot[q];
                    FT v = (FT)jv;
                    if (js & (1 << 30)) {  // a pivot with nothing to eliminate
                        if (bad_pivot(v)) atomicOr(&a.ctl[fi], 1);
                        v = pivot_recip(v);
                    }
                    X[js & ((1 << 30) - 1)] = v;
                }
                grid.sync();
                // ---- factor + forward, level by level --------------------------------------
                for (int l = 0; l < a.nlev; ++l) {
                    factor_level<FT>(a, X, l, gt, GT, &a.ctl[fi]);
                    grid.sync();
                }
                if (a.m > 0) {
                    if (blockIdx.x == 0) dense_tail<FT>(a, X, Tl, &a.ctl[fi]);
                    grid.sync();
                }
                // zero / non-finite pivot: klu_factor failure on the first factorisation of a
                // fresh solver (KLUSolver.cpp:38-44), klu_refactor failure afterwards (:50-57)
                if (__hip_atomic_load(&a.ctl[fi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    st = (a.fresh && it == 1) ? 5 : 6;
                    break;
                }
                // ---- backward -------------------------------------------------------------
                for (int l = 0; l < a.nblev; ++l) {
                    backward_level<FT>(a, X, l, gt, GT);
                    grid.sync();
                }
                // ---- voltage update -----------------------------------------------------
                if (a.slack_dx >= 0) slack -= (double)X[a.slack_dx];
                double vmin = INFINITY;
                for (long long b = gt; b < n; b += GT) {
                    const int2 us = a.bus_dx[b];
                    const double dva = us.x >= 0 ? (double)X[us.x] : 0.0;
                    const double dvm = us.y >= 0 ? (double)X[us.y] : 0.0;
                    double va, vm;
                    if (a.single_slack) {  // BaseNRSingleSlackAlgo.tpp:131-132
                        const double vr = V[2 * b], vi = V[2 * b + 1];
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
                    V[2 * b] = vm * cs;
                    V[2 * b + 1] = vm * sn;
                    vmin = fmin(vmin, vm);
                }
                vmin = -gr.run<true>(-vmin, grid);
                if (vmin < 0.0)
                    for (long long b = gt; b < n; b += GT) {
                        const double vr = V[2 * b], vi = V[2 * b + 1];
                        Va[b] = atan2(vi, vr);
                        Vm[b] = sqrt(vr * vr + vi * vi);
                    }
                // ---- mismatch + convergence ------------------------------------------------
                fm = mismatch();
                flast = fm;
                if (!(fm <= 1.7976931348623157e308)) { st = 3; break; }
                if (fm < a.tol) break;
                if (it >= a.max_iter) { st = 2; break; }
            }
            if (gt == 0) {
                a.iters[row] = it;
                a.status[row] = st;
                if (a.flags) a.flags[row] = (st != 0 && flast < a.stall_tol) ? kFlagStall : 0;
            }
            alive = st == 0;
            grid.sync();  // the next step reuses the workspace and the control words
        }
        grid.sync();  // the next instance rewrites its Ybus values
    }
}

__global__ void k_clear_words(unsigned int* w, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) w[i] = 0u;
}

}  // namespace

hipError_t launch(const WArgs& a, bool fp32, size_t lds_bytes, hipStream_t stream) {
    hipLaunchKernelGGL(k_clear_words, dim3(1), dim3(256), 0, stream, a.bar, kBarrierWords);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    WArgs arg = a;
    void* params[] = {&arg};
    if (fp32)
        return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_wide<float>), dim3(a.nblocks), dim3(kThreads),
                                          params, (unsigned)lds_bytes, stream);
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_wide<double>), dim3(a.nblocks), dim3(kThreads),
                                      params, (unsigned)lds_bytes, stream);
}

int resident_blocks(int device, size_t lds_bytes, bool fp32) {
    int coop = 0, ncu = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess || !coop) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    const void* f = fp32 ? reinterpret_cast<const void*>(&k_wide<float>) : reinterpret_cast<const void*>(&k_wide<double>);
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kThreads, lds_bytes) != hipSuccess || per_cu < 1)
        return 0;
    return ncu;  // one workgroup per compute unit
}

}  // namespace pfgpu_wide
