/*
 * pfgpu — C ABI of the MI355X batched Newton-Raphson power-flow engine.
 *
 * Plain C types only (no HIP, torch or C++ types cross this boundary). All
 * functions return an int status: 0 on success, a positive ErrorType code of
 * the reference (src/Utils.hpp:47-56) for numerical failures, or a negative
 * PFGPU_E_* code for API misuse / device errors; pfgpu_last_error() gives a
 * message for the calling thread. No C++ exception crosses the ABI.
 *
 * Complex arrays are interleaved (re, im) doubles, as std::complex<double>.
 * Ybus is CSC with sorted row indices, the layout the reference's solvers take
 * (Eigen::SparseMatrix<cplx_type>, src/powerflow_algorithm/BaseAlgo.hpp:126-136).
 *
 * Which reference interface each entry point replaces:
 *   pfgpu_grid_load      BaseNRAlgo's first-call work: index prep (BaseNRAlgo.tpp:59-75),
 *                        J pattern + value map (BaseNRAlgo.tpp:396-635) and
 *                        KLULinearSolver::initialize / klu_analyze (KLUSolver.cpp:24-45)
 *   pfgpu_solve_batch    BaseNRAlgo<KLULinearSolver>::compute_pf (BaseNRAlgo.tpp:15-186) or
 *                        BaseNRSingleSlackAlgo::compute_pf (BaseNRSingleSlackAlgo.tpp:12-183)
 *                        run over many instances at once; with Ybus deltas it replaces the
 *                        contingency loop ContingencyAnalysis::compute (ContingencyAnalysis.cpp:204-231),
 *                        with per-instance Sbus the step loop TimeSeries::compute_Vs (TimeSeries.cpp:80-103)
 *   pfgpu_grid_get_J     BaseNRAlgo::get_J (BaseNRAlgo.hpp:36-45)
 */
#ifndef PFGPU_H
#define PFGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFGPU_ABI_VERSION 1

/* API error codes (negative). Positive codes are the reference's ErrorType values:
 * 0 NoError, 1 SingularMatrix, 2 TooManyIterations, 3 InifiniteValue, 4 SolverAnalyze,
 * 5 SolverFactor, 6 SolverReFactor, 7 SolverSolve, 8 NotInitError, 9 LicenseError. */
#define PFGPU_E_INVALID  (-1)
#define PFGPU_E_DEVICE   (-2)
#define PFGPU_E_NOMEM    (-3)
#define PFGPU_E_BUDGET   (-4)
#define PFGPU_E_INTERNAL (-5)

/* precision of the linear solve (factor + triangular solves); the mismatch,
 * the Jacobian entries, the voltage update and the convergence test are FP64 */
#define PFGPU_FP64 0
#define PFGPU_FP32 1

typedef struct pfgpu_ctx pfgpu_ctx;
typedef struct pfgpu_grid pfgpu_grid;

typedef struct pfgpu_grid_info {
    int n_bus;          /* buses (Ybus dimension) */
    int n_jac;          /* J dimension (N+1 for the distributed-slack layout) */
    int nnz_ybus;
    int nnz_jac;
    int64_t nnz_lu;     /* diag + L + U entries of the static-pivot factor */
    int n_slots;        /* per-instance value slots (LU + forward-solve vector) */
    int n_levels;       /* factor + forward dependency levels */
    int n_back_levels;  /* backward-solve levels */
    int64_t n_pairs;    /* multiply pairs of one refactorisation + forward solve */
    double flops;       /* flops of one refactorisation */
    int single_slack;
} pfgpu_grid_info;

/* Per-call timings (seconds) and launch choices. */
typedef struct pfgpu_stats {
    double total_s;       /* whole call, host wall clock */
    double h2d_s;         /* host -> device copies (0 for the device entry point) */
    double kernel_s;      /* NR kernel time (HIP events on the context stream) */
    double d2h_s;         /* device -> host copies */
    int instances_per_wg; /* G chosen for the launch */
    int lds_slots;        /* value slots kept in LDS per instance */
    int workgroups;
    int fallback_fp64;    /* FP32-form instances re-solved in FP64 */
} pfgpu_stats;

/* Thread-local message of the last failing call. */
const char* pfgpu_last_error(void);
int pfgpu_abi_version(void);

/* Context = one device + one HIP stream + an HBM budget for batch workspaces.
 * hbm_budget_bytes == 0 means 90% of the device's free memory at creation. */
int pfgpu_ctx_create(int device, size_t hbm_budget_bytes, pfgpu_ctx** out);
int pfgpu_ctx_destroy(pfgpu_ctx* ctx);
/* Device properties used by the launch heuristics. */
int pfgpu_ctx_device_info(pfgpu_ctx* ctx, int* n_cu, int* lds_bytes_per_wg, size_t* free_bytes);

/* Load a grid: builds the J pattern / scatter map, the ordering, the symbolic
 * static-pivot LU and the level schedule on the host, and uploads them.
 * Yx (base values) may be NULL; values are given per solve call anyway. */
int pfgpu_grid_load(pfgpu_ctx* ctx, int n, const int* Yp, const int* Yi,
                    const int* slack_ids, int nslack, const double* slack_weights,
                    const int* pv, int npv, const int* pq, int npq,
                    int single_slack, pfgpu_grid** out);
int pfgpu_grid_destroy(pfgpu_grid* grid);
int pfgpu_grid_get_info(const pfgpu_grid* grid, pfgpu_grid_info* info);
/* J pattern in the reference layout: Jp[n_jac+1], Ji[nnz_jac]. */
int pfgpu_grid_get_J_pattern(const pfgpu_grid* grid, int* Jp, int* Ji);

/*
 * Batched NR solve, host buffers (copied in and out by the call).
 *   Yx        base Ybus values [nnz_ybus complex]
 *   Yx_stride 0: every instance uses Yx; nnz_ybus: Yx holds ninst value sets
 *   dptr/dpos/dval  optional per-instance deltas (NULL for none): instance i
 *             uses Yx[dpos[k]] - dval[k] for k in [dptr[i], dptr[i+1]) — the
 *             contingency Ybus edit of ContingencyAnalysis::remove_from_Ybus
 *   Sbus      [ninst x n complex] if sbus_stride == n, or one vector if 0
 *   V         [ninst x n complex] in: initial voltages, out: final voltages
 *   v0_stride n: V holds one initial vector per instance; 0: V[0:n] is the
 *             initial vector of every instance (V still receives ninst rows)
 *   precision PFGPU_FP64 or PFGPU_FP32 (FP32 factor/solve, FP64 elsewhere; an
 *             instance that does not converge in FP32 is re-solved in FP64)
 *   instances_per_wg  0 = automatic, else G in {1,2,4,8,16,32}
 *   iters/status      [ninst] outputs: NR iterations (linear solves) and the
 *             ErrorType (0 converged)
 *   Jx_last   optional [nnz_jac] J values of instance 0's last iteration (or NULL)
 */
int pfgpu_solve_batch(pfgpu_ctx* ctx, pfgpu_grid* grid, int ninst,
                      const double* Yx, int Yx_stride,
                      const int* dptr, const int* dpos, const double* dval,
                      const double* Sbus, int sbus_stride,
                      double* V, int v0_stride,
                      int max_iter, double tol, int precision, int instances_per_wg,
                      int* iters, int* status, double* Jx_last, pfgpu_stats* stats);

/*
 * Same as pfgpu_solve_batch with every array already resident in device
 * memory (pointers from pfgpu_dev_alloc or any device allocation on the
 * context's device); the call runs on the context's stream and returns after
 * the stream is synchronised. d_V0 (read only, one row per instance if
 * v0_stride == n, one shared row if 0) gives the initial voltages and d_V
 * [ninst x n complex] receives the results. No host copies and no FP64
 * re-solve pass: the status of FP32 instances that did not converge is
 * reported as is.
 */
int pfgpu_solve_batch_device(pfgpu_ctx* ctx, pfgpu_grid* grid, int ninst,
                             const double* d_Yx, int Yx_stride,
                             const int* d_dptr, const int* d_dpos, const double* d_dval,
                             const double* d_Sbus, int sbus_stride,
                             const double* d_V0, int v0_stride, double* d_V,
                             int max_iter, double tol, int precision, int instances_per_wg,
                             int* d_iters, int* d_status, pfgpu_stats* stats);

/*
 * Launch autotuner: runs the batch described by the arguments (same meaning as
 * pfgpu_solve_batch_device; outputs are overwritten) once per candidate launch
 * shape — instances per workgroup G in {1, 2, 4, 8, 16, 32} x block size in
 * {256, 1024} — keeps the fastest for this grid and precision, and returns it.
 * Later solves with instances_per_wg == 0 use the tuned shape.
 */
int pfgpu_autotune_device(pfgpu_ctx* ctx, pfgpu_grid* grid, int ninst,
                          const double* d_Yx, int Yx_stride,
                          const int* d_dptr, const int* d_dpos, const double* d_dval,
                          const double* d_Sbus, int sbus_stride,
                          const double* d_V0, int v0_stride, double* d_V,
                          int max_iter, double tol, int precision,
                          int* d_iters, int* d_status, int* best_G, int* best_block, double* best_seconds);

/*
 * Branch flows at the origin side for every instance, on device-resident voltages
 * (the post-processing of BaseBatchSolverSynch::compute_flows_from_Vs,
 * src/batch_algorithm/BaseBatchSolverSynch.hpp:86-192):
 *   I = yff Ef + yft Et,  S = Ef conj(I),
 *   amps = |S| sn_mva / (sqrt(3) |Ef| vn_kv_from)  (kA),  p = Re(S) sn_mva  (MW).
 * d_V is [ninst x n] complex; d_f/d_t/d_yff/d_yft/d_vn_kv_from describe nbr branches
 * (complex admittances interleaved); d_amps / d_p are [ninst x nbr] (either may be NULL).
 * A row of zero voltages gives NaN amps, as the reference's division does.
 */
int pfgpu_branch_flows_device(pfgpu_ctx* ctx, int ninst, int n, const double* d_V, int nbr,
                              const int* d_f, const int* d_t, const double* d_yff, const double* d_yft,
                              const double* d_vn_kv_from, double sn_mva, double* d_amps, double* d_p);

/* Device memory helpers bound to the context's device and stream. */
int pfgpu_dev_alloc(pfgpu_ctx* ctx, size_t bytes, void** dptr);
int pfgpu_dev_free(pfgpu_ctx* ctx, void* dptr);
int pfgpu_memcpy_h2d(pfgpu_ctx* ctx, void* dst, const void* src, size_t bytes);
int pfgpu_memcpy_d2h(pfgpu_ctx* ctx, void* dst, const void* src, size_t bytes);
int pfgpu_memcpy_d2d(pfgpu_ctx* ctx, void* dst, const void* src, size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* PFGPU_H */
