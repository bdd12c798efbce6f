// Device kernels on either side of the batched NR solve: what the batch drivers do before it
// (Ybus values from branch parameters, time-series Sbus rows, contingency Ybus edits and the
// connectivity check) and after it (branch flows of every solved row, with the contingency
// clean-up), so that a batch does not round-trip through the host between these steps.
//   Ybus stamps           GridModel::pre_process_solver subset (GridModel.cpp:450-571),
//                         TwoSidesContainer_rxh_A / LineContainer.cpp:13-34
//   Sbus rows             TimeSeries::compute_Vs (TimeSeries.cpp:45-53), fill_SBus_real/imag
//                         (TimeSeries.hpp:80-151)
//   contingency edits     ContingencyAnalysis::init_li_coeffs / remove_from_Ybus (ContingencyAnalysis.cpp:47-110)
//   connectivity          ContingencyAnalysis::check_invertible (ContingencyAnalysis.cpp:14-45)
//   flows + clean-up      BaseBatchSolverSynch::compute_amps_flows / compute_active_power_flows
//                         (BaseBatchSolverSynch.hpp:86-192), ContingencyAnalysis::clean_flows (:237-250)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/pfgpu.h"
#include "engine.hpp"

namespace {

using pfgpu_engine::ApiError;
using pfgpu_engine::guarded;

// Branch flows at the origin side of every (row, branch): I = yff Ef + yft Et, S = Ef conj(I),
// amps = (|S| sn_mva) / (sqrt(3) (|Ef| vn_kv)), p = Re(S) sn_mva, evaluated in that order.
// Row r takes the voltages of solved row row_src[r] (identity if row_src is null; -1: a row of
// zero voltages, whose amps are 0/0 = NaN). A branch with f < 0 is out of service: 0. Then the
// branches listed for row r (cptr/cbr) get 0 where the value is finite (clean_flows).
__global__ void k_batch_flows(long long nrows, int n, const double* __restrict__ V, const int* __restrict__ row_src,
                              int nbr, const int* __restrict__ f, const int* __restrict__ t,
                              const double* __restrict__ yff, const double* __restrict__ yft,
                              const double* __restrict__ vn, double sn_mva, const int* __restrict__ cptr,
                              const int* __restrict__ cbr, double* __restrict__ amps, double* __restrict__ p, int dc) {
    const long long total = nrows * nbr;
    const double sqrt3 = 1.7320508075688772;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const long long r = k / nbr;
        const int b = (int)(k - r * nbr);
        const int fb = f[b];
        double a = 0.0, pw = 0.0;
        if (fb >= 0) {
            const long long src = row_src ? (long long)row_src[r] : r;
            double efr = 0.0, efi = 0.0, etr = 0.0, eti = 0.0;
            if (src >= 0) {
                const double* Vr = V + 2 * src * (long long)n;
                efr = Vr[2 * fb];
                efi = Vr[2 * fb + 1];
                etr = Vr[2 * t[b]];
                eti = Vr[2 * t[b] + 1];
            }
            const double gr = yff[2 * b], gi = yff[2 * b + 1], hr = yft[2 * b], hi = yft[2 * b + 1];
            if (dc) {
                // P = (Re(yff) arg(Ef) + Re(yft) arg(Et)) sn_mva, A = |P| / (sqrt(3) |Ef| vn_kv)
                pw = (gr * atan2(efi, efr) + hr * atan2(eti, etr)) * sn_mva;
                a = fabs(pw) / (sqrt3 * (hypot(efr, efi) * vn[b]));
            } else {
                // I = yff Ef + yft Et
                const double ir = (gr * efr - gi * efi) + (hr * etr - hi * eti);
                const double ii = (gr * efi + gi * efr) + (hr * eti + hi * etr);
                // S = Ef conj(I)
                const double sr = efr * ir + efi * ii;
                const double si = efi * ir - efr * ii;
                a = (hypot(sr, si) * sn_mva) / (sqrt3 * (hypot(efr, efi) * vn[b]));
                pw = sr * sn_mva;
            }
        }
        if (cptr)
            for (int q = cptr[r]; q < cptr[r + 1]; ++q)
                if (cbr[q] == b) {
                    if (isfinite(a)) a = 0.0;
                    if (isfinite(pw)) pw = 0.0;
                }
        if (amps) amps[k] = a;
        if (p) p[k] = pw;
    }
}

// Sbus rows: per (step t, solver bus b) the real part adds the generators' p, then the static
// generators' p and subtracts the loads' p, the imaginary part subtracts the loads' q, each in
// element order (the order of the reference's element loops for one bus); then both are divided
// by sn_mva unless sn_div is 0. Element lists per bus are CSR (ptr[ns + 1], element ids); the
// injection matrices are row-major [T x elements].
__global__ void k_sbus_rows(int T, int ns, const int* __restrict__ gptr, const int* __restrict__ gel, int ngen,
                            const double* __restrict__ gen_p, const int* __restrict__ sptr,
                            const int* __restrict__ sel, int nsgen, const double* __restrict__ sgen_p,
                            const int* __restrict__ lptr, const int* __restrict__ lel, int nload,
                            const double* __restrict__ load_p, const double* __restrict__ load_q, double sn_div,
                            double* __restrict__ S) {
    const long long total = (long long)T * ns;
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < total;
         k += (long long)gridDim.x * blockDim.x) {
        const long long t = k / ns;
        const int b = (int)(k - t * ns);
        double re = 0.0, im = 0.0;
        for (int q = gptr[b]; q < gptr[b + 1]; ++q) re += gen_p[t * ngen + gel[q]];
        for (int q = sptr[b]; q < sptr[b + 1]; ++q) re += sgen_p[t * nsgen + sel[q]];
        for (int q = lptr[b]; q < lptr[b + 1]; ++q) re -= load_p[t * nload + lel[q]];
        for (int q = lptr[b]; q < lptr[b + 1]; ++q) im -= load_q[t * nload + lel[q]];
        if (sn_div != 0.0) {
            re = re / sn_div;
            im = im / sn_div;
        }
        S[2 * k] = re;
        S[2 * k + 1] = im;
    }
}

// The four stamps of branch k (pi model, line charging split in halves): ys = 1 / (r + j x)
// as (r - j x) / (r^2 + x^2); yff = ytt = ys + h / 2, yft = ytf = -ys. Rounded operation by
// operation with no fused multiply-add, as the host mirror evaluates them.
__device__ __forceinline__ void branch_stamp(double r, double x, double hr, double hi, int which, double& yr,
                                             double& yi) {
#pragma clang fp contract(off)
    const double d = r * r + x * x;
    const double sr = r / d, si = -x / d;
    if (which < 2) {
        yr = sr + 0.5 * hr;
        yi = si + 0.5 * hi;
    } else {
        yr = -sr;
        yi = -si;
    }
}

// Ybus values on a fixed CSC pattern: entry e sums its contributions cidx[cptr[e] .. cptr[e+1])
// in list order, contribution c = 4 * branch + {0: ff, 1: tt, 2: ft, 3: tf}; a branch with
// on == 0 contributes nothing (its entries stay as explicit zeros).
__global__ void k_ybus_values(int nnz, const int* __restrict__ cptr, const int* __restrict__ cidx,
                              const double* __restrict__ r, const double* __restrict__ x,
                              const double* __restrict__ h, const int* __restrict__ on, double* __restrict__ Yx) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += gridDim.x * blockDim.x) {
        double yr = 0.0, yi = 0.0;
        for (int q = cptr[e]; q < cptr[e + 1]; ++q) {
            const int c = cidx[q], k = c >> 2;
            if (!on[k]) continue;
            double vr, vi;
            branch_stamp(r[k], x[k], h[2 * k], h[2 * k + 1], c & 3, vr, vi);
            yr += vr;
            yi += vi;
        }
        Yx[2 * e] = yr;
        Yx[2 * e + 1] = yi;
    }
}

// Contingency edits: contingency c removes the branches cbr[cptr[c] .. cptr[c+1]); branch k has
// CSC positions pos[4k .. 4k+3] (ff, tt, ft, tf; -1: out of service, no edit) and values
// y[4k .. 4k+3]. Edits of contingency c fill [dptr[c], dptr[c+1]) with dptr from the host
// (4 per in-service branch), in branch order: Y_c = Y - dval at dpos.
__global__ void k_cont_edits(int ncont, const int* __restrict__ cptr, const int* __restrict__ cbr,
                             const int* __restrict__ pos, const double* __restrict__ y, const int* __restrict__ dptr,
                             int* __restrict__ dpos, double* __restrict__ dval) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncont; c += gridDim.x * blockDim.x) {
        int o = dptr[c];
        for (int q = cptr[c]; q < cptr[c + 1]; ++q) {
            const int k = cbr[q];
            if (pos[4 * k] < 0) continue;
            for (int u = 0; u < 4; ++u, ++o) {
                dpos[o] = pos[4 * k + u];
                dval[2 * o] = y[2 * (4 * k + u)];
                dval[2 * o + 1] = y[2 * (4 * k + u) + 1];
            }
        }
    }
}

constexpr int kBfsMaxEdits = 64;
constexpr int kBfsThreads = 256;

// Connectivity after each contingency (check_invertible): breadth-first search from bus 0 over
// the Ybus entries whose edited value has |y| > 1e-8; connected[c] = every bus reached. One
// workgroup per contingency (grid-stride over contingencies), level-synchronous: the frontier of
// a level is scanned by all lanes, a bus is claimed with an LDS atomic on the visited bitmap and
// appended to the next frontier. n <= max_bus (the frontiers and the bitmap live in LDS).
__global__ void __launch_bounds__(kBfsThreads) k_connectivity(int ncont, int n, const int* __restrict__ Yp,
                                                              const int* __restrict__ Yi,
                                                              const double* __restrict__ Yx,
                                                              const int* __restrict__ dptr,
                                                              const int* __restrict__ dpos,
                                                              const double* __restrict__ dval,
                                                              int* __restrict__ connected) {
    extern __shared__ __align__(16) unsigned char lds[];
    unsigned int* vis = reinterpret_cast<unsigned int*>(lds);  // [(n + 31) / 32]
    const int nw = (n + 31) / 32;
    int* fr0 = reinterpret_cast<int*>(vis + ((nw + 3) & ~3));
    int* fr1 = fr0 + n;
    __shared__ int s_len[2];
    __shared__ int s_nvis;
    __shared__ int s_epos[kBfsMaxEdits];
    __shared__ double s_eval[kBfsMaxEdits][2];
    __shared__ int s_ne;
    const int tid = threadIdx.x;
    for (int c = blockIdx.x; c < ncont; c += gridDim.x) {
        for (int w = tid; w < nw; w += blockDim.x) vis[w] = 0u;
        if (tid == 0) {
            s_ne = dptr ? dptr[c + 1] - dptr[c] : 0;
            s_len[0] = 1;
            s_len[1] = 0;
            s_nvis = 1;
            fr0[0] = 0;
        }
        __syncthreads();
        const int ne = s_ne, d0 = dptr ? dptr[c] : 0;
        for (int q = tid; q < ne && q < kBfsMaxEdits; q += blockDim.x) {
            s_epos[q] = dpos[d0 + q];
            s_eval[q][0] = dval[2 * (d0 + q)];
            s_eval[q][1] = dval[2 * (d0 + q) + 1];
        }
        if (tid == 0) vis[0] = 1u;
        __syncthreads();
        int cur = 0;
        for (;;) {
            const int len = s_len[cur];
            if (len == 0) break;
            int* in = cur ? fr1 : fr0;
            int* out = cur ? fr0 : fr1;
            // entries of the frontier's columns, one (column, entry) per lane at a time
            for (int i = tid; i < len; i += blockDim.x) {
                const int col = in[i];
                for (int q = Yp[col]; q < Yp[col + 1]; ++q) {
                    const int r = Yi[q];
                    if (vis[r >> 5] & (1u << (r & 31))) continue;
                    double yr = Yx[2 * q], yi = Yx[2 * q + 1];
                    for (int e = 0; e < ne; ++e) {
                        const bool lds_e = e < kBfsMaxEdits;
                        const int ep = lds_e ? s_epos[e] : dpos[d0 + e];
                        if (ep == q) {
                            yr -= lds_e ? s_eval[e][0] : dval[2 * (d0 + e)];
                            yi -= lds_e ? s_eval[e][1] : dval[2 * (d0 + e) + 1];
                        }
                    }
                    if (!(hypot(yr, yi) > 1e-8)) continue;
                    const unsigned int bit = 1u << (r & 31);
                    const unsigned int old = atomicOr(&vis[r >> 5], bit);
                    if (!(old & bit)) out[atomicAdd(&s_len[cur ^ 1], 1)] = r;
                }
            }
            __syncthreads();
            if (tid == 0) {
                s_nvis += s_len[cur ^ 1];
                s_len[cur] = 0;
            }
            __syncthreads();
            cur ^= 1;
        }
        if (tid == 0) connected[c] = s_nvis == n ? 1 : 0;
        __syncthreads();
    }
}

size_t bfs_lds(int n) { return (size_t)(((n + 31) / 32 + 3) & ~3) * 4 + (size_t)2 * n * 4; }

int grid_of(long long work, int block) { return (int)std::max<long long>(1, std::min<long long>((work + block - 1) / block, 65536)); }

int batch_flows(pfgpu_ctx* ctx, long long nrows, int n, const double* d_V, const int* d_row_src, int nbr,
                const int* d_f, const int* d_t, const double* d_yff, const double* d_yft, const double* d_vn_kv_from,
                double sn_mva, const int* d_clean_ptr, const int* d_clean_br, double* d_amps, double* d_p, int dc) {
    return guarded([&]() -> int {
        if (!ctx || nrows < 0 || n <= 0 || nbr < 0 || (d_clean_ptr == nullptr) != (d_clean_br == nullptr) ||
            (nrows > 0 && nbr > 0 && (!d_V || !d_f || !d_t || !d_yff || !d_yft || !d_vn_kv_from)))
            throw ApiError(PFGPU_E_INVALID, "invalid batch-flow arguments");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        const long long total = nrows * nbr;
        if (total == 0 || (!d_amps && !d_p)) return 0;
        hipLaunchKernelGGL(k_batch_flows, dim3(grid_of(total, 256)), dim3(256), 0, ctx->stream, nrows, n, d_V, d_row_src,
                           nbr, d_f, d_t, d_yff, d_yft, d_vn_kv_from, sn_mva, d_clean_ptr, d_clean_br, d_amps, d_p, dc);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}
}  // namespace

extern "C" {

int pfgpu_batch_flows_device(pfgpu_ctx* ctx, long long nrows, int n, const double* d_V, const int* d_row_src, int nbr,
                             const int* d_f, const int* d_t, const double* d_yff, const double* d_yft,
                             const double* d_vn_kv_from, double sn_mva, const int* d_clean_ptr, const int* d_clean_br,
                             double* d_amps, double* d_p) {
    return batch_flows(ctx, nrows, n, d_V, d_row_src, nbr, d_f, d_t, d_yff, d_yft, d_vn_kv_from, sn_mva, d_clean_ptr,
                       d_clean_br, d_amps, d_p, 0);
}

int pfgpu_batch_flows_dc_device(pfgpu_ctx* ctx, long long nrows, int n, const double* d_V, const int* d_row_src,
                                int nbr, const int* d_f, const int* d_t, const double* d_yff, const double* d_yft,
                                const double* d_vn_kv_from, double sn_mva, const int* d_clean_ptr,
                                const int* d_clean_br, double* d_amps, double* d_p) {
    return batch_flows(ctx, nrows, n, d_V, d_row_src, nbr, d_f, d_t, d_yff, d_yft, d_vn_kv_from, sn_mva, d_clean_ptr,
                       d_clean_br, d_amps, d_p, 1);
}

int pfgpu_sbus_rows_device(pfgpu_ctx* ctx, int T, int ns, const int* d_gen_ptr, const int* d_gen_el, int ngen,
                           const double* d_gen_p, const int* d_sgen_ptr, const int* d_sgen_el, int nsgen,
                           const double* d_sgen_p, const int* d_load_ptr, const int* d_load_el, int nload,
                           const double* d_load_p, const double* d_load_q, double sn_mva, double* d_Sbus) {
    return guarded([&]() -> int {
        if (!ctx || T < 0 || ns <= 0 || ngen < 0 || nsgen < 0 || nload < 0 || !d_gen_ptr || !d_sgen_ptr ||
            !d_load_ptr || !d_Sbus || (ngen > 0 && (!d_gen_el || !d_gen_p)) ||
            (nsgen > 0 && (!d_sgen_el || !d_sgen_p)) || (nload > 0 && (!d_load_el || !d_load_p || !d_load_q)))
            throw ApiError(PFGPU_E_INVALID, "invalid Sbus-row arguments");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        const long long total = (long long)T * ns;
        if (total == 0) return 0;
        // TimeSeries.cpp:53 divides unless sn_mva equals 1 within 1e-7 (BaseConstants::_tol_equal_float)
        const double sn_div = std::fabs(sn_mva - 1.0) > 1e-7 ? sn_mva : 0.0;
        hipLaunchKernelGGL(k_sbus_rows, dim3(grid_of(total, 256)), dim3(256), 0, ctx->stream, T, ns, d_gen_ptr, d_gen_el,
                           ngen, d_gen_p, d_sgen_ptr, d_sgen_el, nsgen, d_sgen_p, d_load_ptr, d_load_el, nload, d_load_p,
                           d_load_q, sn_div, d_Sbus);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int pfgpu_ybus_values_device(pfgpu_ctx* ctx, int nnz, const int* d_contrib_ptr, const int* d_contrib, int nbr,
                             const double* d_r, const double* d_x, const double* d_h, const int* d_on, double* d_Yx) {
    return guarded([&]() -> int {
        if (!ctx || nnz < 0 || nbr < 0 || !d_contrib_ptr || !d_Yx ||
            (nbr > 0 && (!d_contrib || !d_r || !d_x || !d_h || !d_on)))
            throw ApiError(PFGPU_E_INVALID, "invalid Ybus-value arguments");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        if (nnz == 0) return 0;
        hipLaunchKernelGGL(k_ybus_values, dim3(grid_of(nnz, 256)), dim3(256), 0, ctx->stream, nnz, d_contrib_ptr,
                           d_contrib, d_r, d_x, d_h, d_on, d_Yx);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int pfgpu_contingency_edits_device(pfgpu_ctx* ctx, int ncont, const int* d_cont_ptr, const int* d_cont_br,
                                   const int* d_br_pos, const double* d_br_y, const int* d_dptr, int* d_dpos,
                                   double* d_dval) {
    return guarded([&]() -> int {
        if (!ctx || ncont < 0 || (ncont > 0 && (!d_cont_ptr || !d_cont_br || !d_br_pos || !d_br_y || !d_dptr ||
                                                !d_dpos || !d_dval)))
            throw ApiError(PFGPU_E_INVALID, "invalid contingency-edit arguments");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        if (ncont == 0) return 0;
        hipLaunchKernelGGL(k_cont_edits, dim3(grid_of(ncont, 256)), dim3(256), 0, ctx->stream, ncont, d_cont_ptr,
                           d_cont_br, d_br_pos, d_br_y, d_dptr, d_dpos, d_dval);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

int pfgpu_connectivity_device(pfgpu_ctx* ctx, int ncont, int n, const int* d_Yp, const int* d_Yi, const double* d_Yx,
                              const int* d_dptr, const int* d_dpos, const double* d_dval, int* d_connected) {
    return guarded([&]() -> int {
        if (!ctx || ncont < 0 || n <= 0 || !d_Yp || !d_Yi || !d_Yx || !d_connected ||
            (d_dptr == nullptr) != (d_dpos == nullptr) || (d_dptr == nullptr) != (d_dval == nullptr))
            throw ApiError(PFGPU_E_INVALID, "invalid connectivity arguments");
        std::lock_guard<std::mutex> lk(ctx->mu);
        HIP_OK(hipSetDevice(ctx->device));
        if (ncont == 0) return 0;
        const size_t lds = bfs_lds(n);
        // static LDS of the kernel: lengths, counters and the staged edits (about 1.3 KB)
        if (lds + 2048 > (size_t)ctx->lds_max)
            throw ApiError(PFGPU_E_INVALID, "grid too large for the device connectivity check");
        if (lds > 65536) HIP_OK(hipFuncSetAttribute((const void*)k_connectivity,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const int blocks = std::min(ncont, 8 * std::max(1, ctx->n_cu));
        hipLaunchKernelGGL(k_connectivity, dim3(blocks), dim3(kBfsThreads), lds, ctx->stream, ncont, n, d_Yp, d_Yi, d_Yx,
                           d_dptr, d_dpos, d_dval, d_connected);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(ctx->stream));
        return 0;
    });
}

}  // extern "C"
