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
 * Which reference interface each entry point replaces (paths under src/):
 *   pfgpu_grid_load      BaseNRAlgo's first-call work: index prep (BaseNRAlgo.tpp:59-75),
 *                        J pattern + value map (BaseNRAlgo.tpp:396-635) and
 *                        KLULinearSolver::initialize / klu_analyze (KLUSolver.cpp:24-45)
 *   pfgpu_grid_reset     BaseNRAlgo::reset / KLULinearSolver::reset (BaseNRAlgo.tpp:188-200,
 *                        KLUSolver.cpp:14-22): the next solve factors afresh
 *   pfgpu_solve_batch    BaseNRAlgo<KLULinearSolver>::compute_pf (BaseNRAlgo.tpp:15-186) or
 *                        BaseNRSingleSlackAlgo::compute_pf (BaseNRSingleSlackAlgo.tpp:12-183)
 *                        run over many instances at once; with Ybus deltas it replaces the
 *                        contingency loop ContingencyAnalysis::compute (ContingencyAnalysis.cpp:204-231)
 *   pfgpu_solve_chains   the warm-started step loop of TimeSeries::compute_Vs (TimeSeries.cpp:80-103):
 *                        one chain is the reference's loop; several chains run side by side
 *   pfgpu_grid_get_J_pattern + the Jx_last output of pfgpu_solve_batch
 *                        BaseNRAlgo::get_J (BaseNRAlgo.hpp:36-45)
 */
#ifndef PFGPU_H
#define PFGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFGPU_ABI_VERSION 2

/* API error codes (negative). Positive codes are the reference's ErrorType values:
 * 0 NoError, 1 SingularMatrix, 2 TooManyIterations, 3 InifiniteValue, 4 SolverAnalyze,
 * 5 SolverFactor, 6 SolverReFactor, 7 SolverSolve, 8 NotInitError, 9 LicenseError. */
#define PFGPU_E_INVALID  (-1)
#define PFGPU_E_DEVICE   (-2)
#define PFGPU_E_NOMEM    (-3)
#define PFGPU_E_BUDGET   (-4)
#define PFGPU_E_INTERNAL (-5)

/* precision of the linear solve (factor + triangular solves); the mismatch,
 * the Jacobian entries, the voltage update and the convergence test are FP64.
 * PFGPU_FP32: rows that do not converge with FP32 factors and stalled (their last mismatch is
 * below sqrt(tol): the inexact FP32 steps, not the Newton iteration, stopped the progress), or met
 * a zero / non-finite pivot or a non-finite mismatch, are re-solved with FP64 factors from their
 * initial voltages (chains from their heads); rows that diverged are not. pfgpu_stats.fallback_fp64
 * counts them. PFGPU_FP32_STRICT: no re-solve. */
#define PFGPU_FP64 0
#define PFGPU_FP32 1
#define PFGPU_FP32_STRICT 2

/* per-row flags (pfgpu_ctx_last_flags) */
#define PFGPU_FLAG_STALL 1   /* not converged, last mismatch below sqrt(tol) */
#define PFGPU_FLAG_GROWTH 2  /* a pivot kept less than 2^-10 (FP32) / 1e-10 (FP64) of its value */

/* status of a chain step after the chain's first failure (not solved): NotInitError */
#define PFGPU_STATUS_NOT_RUN 8

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
    int dc;             /* loaded with pfgpu_grid_load_dc */
    int fdpf;           /* loaded with pfgpu_grid_load_fdpf */
} pfgpu_grid_info;

/* Per-call timings (seconds) and launch choices. */
typedef struct pfgpu_stats {
    double total_s;       /* whole call, host wall clock */
    double h2d_s;         /* host -> device copies (0 for the device entry points) */
    double kernel_s;      /* NR launches (HIP events on the context stream), FP64 re-solve included */
    double d2h_s;         /* device -> host copies */
    int instances_per_wg; /* G chosen for the launch */
    int lds_slots;        /* value slots kept in LDS per instance */
    int workgroups;       /* instance groups of G */
    int fallback_fp64;    /* FP32-form rows (or chains) re-solved in FP64 */
    /* wall clock of workgroup 0 per NR phase, summed over its iterations and instances:
     * J fill, factor + forward solve, dense tail, backward solve, voltage update, mismatch */
    double phase_s[6];
    int block_size;           /* lanes per workgroup */
    int resident_workgroups;  /* workgroups launched (each loops over instance groups) */
    int graph_launches;       /* launch sequences replayed from a captured HIP graph (see below) */
    int repivot_rows;         /* rows (or chains) re-solved on pivots chosen on the numbers (see below) */
} pfgpu_stats;

/* Launch sequences: every solve call issues k_init_v, k_gather_y (shared Ybus) and the NR kernel on
 * the context's stream. A grid keeps the last 8 distinct sequences (by their exact arguments: device
 * pointers, sizes, launch shape) as instantiated HIP graphs: a sequence issued a second time is
 * captured, and later identical calls replay the graph (one launch instead of three argument
 * setups). PFGPU_GRAPH=0 in the environment turns this off. */

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
 * static-pivot LU and the level schedule on the host, and uploads them. Only the Ybus
 * pattern (Yp, Yi) is taken here; its values are given to every solve call. */
int pfgpu_grid_load(pfgpu_ctx* ctx, int n, const int* Yp, const int* Yi,
                    const int* slack_ids, int nslack, const double* slack_weights,
                    const int* pv, int npv, const int* pq, int npq,
                    int single_slack, pfgpu_grid** out);
/* DC power flow grid (BaseDCAlgo::compute_pf, src/powerflow_algorithm/BaseDCAlgo.tpp:15-179):
 * the solve calls then take the DC Ybus values, factor B = Re(Ybus) without the slack rows and
 * columns (slack buses: in neither slack_ids[1:] ++ pv nor pq) and solve B theta = Re(Sbus) once
 * per instance on the NR kernel's factor and solve phases; V = |V0| exp(j (theta + arg V0[first
 * slack])), 1 iteration, status SolverSolve (7) if theta is not finite or |theta| >= 1e6.
 * max_iter and tol are ignored, as in the reference; PFGPU_FP64 only. */
int pfgpu_grid_load_dc(pfgpu_ctx* ctx, int n, const int* Yp, const int* Yi, const int* slack_ids, int nslack,
                       const int* pv, int npv, const int* pq, int npq, pfgpu_grid** out);
int pfgpu_grid_destroy(pfgpu_grid* grid);
/* The next solve on the grid factors afresh: a zero / non-finite pivot in its first
 * iteration is SolverFactor (5), later ones SolverReFactor (6), as with klu_factor then
 * klu_refactor. A loaded grid starts fresh.
 * Pivots chosen on the numbers (host re-pivot, the partial pivoting of klu_factor,
 * KLUSolver.cpp:31-45): when the first factorisation of a fresh solve meets a zero or
 * non-finite diagonal pivot, the library takes J at the first such instance's initial voltages,
 * chooses its pivot rows on the host (threshold partial pivoting: the diagonal row while it
 * holds at least 1e-3 of its column's largest candidate), re-solves those instances in FP64 on
 * a plan with these pivots (pfgpu_stats.repivot_rows) and keeps that plan for the grid's later
 * solves, as klu_refactor keeps klu_factor's pivots, until this call drops it. A numerically
 * singular J keeps SolverFactor. PFGPU_FP32_STRICT and pfgpu_solve_batch_device_noalloc do not
 * re-pivot.
 * Work-queue order: a batch of at least 2 x (compute units) instances (one per row, no chains)
 * takes its instances longest first, by their iteration counts in the previous batch of the same
 * size on this grid (index order for the first one and after this call). The results do not
 * depend on the order. */
int pfgpu_grid_reset(pfgpu_grid* grid);

/* A fast-decoupled grid (BaseFDPFAlgo<LinearSolver, XB_BX>::compute_pf, BaseFDPFAlgo.tpp:12-160;
 * the reference builds grid_Bp / grid_Bpp with GridModel::fillBp_Bpp, GridModel.cpp:1037-1060):
 * Bp_* / Bpp_* are the n x n matrices B' and B'' (CSC, any row order, duplicates summed). The
 * grid takes my_pv = slack_ids[1:] followed by pv as angle unknowns (retrieve_pv_with_slack) and
 * factors the block-diagonal [B'[pvpq, pvpq] 0; 0 B''[pq, pq]] with the NR kernels' factor and
 * solve phases. A solve alternates the P half-step (Va[pvpq] -= B'^-1 P) and the Q half-step
 * (Vm[pq] -= B''^-1 Q), P and Q being the mismatch (slack_absorbed = Re(sum Sbus) kept fixed)
 * divided by |V|, with the convergence test (max |P|, |Q| < tol) after each; iters counts P
 * half-steps, max_iter bounds them. Solved with PFGPU_FP64 only. */
int pfgpu_grid_load_fdpf(pfgpu_ctx* ctx, int n, const int* Yp, const int* Yi, const int* slack_ids, int nslack,
                         const double* slack_weights, const int* pv, int npv, const int* pq, int npq, const int* Bp_p,
                         const int* Bp_i, const double* Bp_x, const int* Bpp_p, const int* Bpp_i, const double* Bpp_x,
                         pfgpu_grid** out);
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
 *   precision PFGPU_FP64, PFGPU_FP32 or PFGPU_FP32_STRICT (see above)
 *   instances_per_wg  0 = automatic, else G in {1,2,4,8,16,32}; -1 = wide launch: the
 *             instances one after the other, each spread over 32 workgroups that run the NR
 *             phases and each elimination level together, separated by grid-wide barriers
 *             (cooperative launch). Automatic for a single instance (or a single chain) of a
 *             plan with at least 200000 factor pairs; pfgpu_stats.instances_per_wg is -1 then.
 *             Not for DC grids; PFGPU_WIDE_BLOCKS sets the workgroup count.
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
 * [ninst x n complex] receives the results (d_V0 == d_V solves in place).
 * The FP64 re-solve of PFGPU_FP32 runs on the device too.
 */
int pfgpu_solve_batch_device(pfgpu_ctx* ctx, pfgpu_grid* grid, int ninst,
                             const double* d_Yx, int Yx_stride,
                             const int* d_dptr, const int* d_dpos, const double* d_dval,
                             const double* d_Sbus, int sbus_stride,
                             const double* d_V0, int v0_stride, double* d_V,
                             int max_iter, double tol, int precision, int instances_per_wg,
                             int* d_iters, int* d_status, pfgpu_stats* stats);

/*
 * pfgpu_solve_batch_device without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Same arguments and results; the workspace is a 64 MiB device
 * array the library holds per GPU until the process ends (per-call rows: 8 bytes per instance,
 * plus 16 n per instance for an in-place FP32 call; the rest bounds how many workgroups run at
 * once, not the batch). A batch whose rows do not fit fails with PFGPU_E_BUDGET. The launch
 * shape's schedule must already be on the device: one pfgpu_solve_batch_device (or autotune) call
 * with a batch of the same shape puts it there, else PFGPU_E_INVALID. Calls on one GPU share the
 * array: the library runs them one at a time within the process (calls that overlapped on one
 * GPU would race on it); each call keeps the array's last contents and clears what must start at
 * zero itself.
 */
int pfgpu_solve_batch_device_noalloc(pfgpu_ctx* ctx, pfgpu_grid* grid, int ninst, const double* d_Yx, int Yx_stride,
                                     const int* d_dptr, const int* d_dpos, const double* d_dval,
                                     const double* d_Sbus, int sbus_stride, const double* d_V0, int v0_stride,
                                     double* d_V, int max_iter, double tol, int precision, int instances_per_wg,
                                     int* d_iters, int* d_status, pfgpu_stats* stats);

/*
 * Time series (TimeSeries::compute_Vs, TimeSeries.cpp:80-103): nrows steps, each with its own
 * Sbus row, cut into chains of chain_len consecutive steps (chain_len >= nrows: the reference's
 * single chain). The head of chain c starts from V0 (row c if v0_stride == n, the shared row if
 * 0); every later step starts from the previous step's converged V. A chain stops at its first
 * step that does not converge; its later steps get PFGPU_STATUS_NOT_RUN, 0 iterations and zero
 * voltages. Several chains run in one launch, the steps of a chain one after the other inside it.
 * A single chain of 4 steps or more (tol <= 1e-6) runs in batched passes instead: pass 1 solves
 * every step from the head, pass 2 solves step k from pass 1's step k - 1. Up to the first step
 * whose two passes
 * did not converge to the same voltages (within 1e-6 pu), each pass-2 step starts where the chain's
 * own step would, to within what two converged solves of one step differ, so it takes the same
 * iterations to the same solution (within the tolerance); the rest is solved again from that
 * step, and after 8 such rounds as a sequential chain (a long chain of a large grid is taken in
 * windows whose pass buffers fit a quarter of the context's HBM budget). Results equal the
 * step-by-step chain's in converged flags and iterations, and in voltages to within the
 * tolerance; the step-by-step chain runs instead when PFGPU_TS_SEQUENTIAL=1 is set in the
 * environment.
 *   Yx      shared Ybus values [nnz_ybus complex]
 *   Sbus    [nrows x n complex]; V [nrows x n complex] out; iters / status [nrows] out
 */
int pfgpu_solve_chains(pfgpu_ctx* ctx, pfgpu_grid* grid, int nrows, int chain_len, const double* Yx,
                       const double* Sbus, const double* V0, int v0_stride, double* V, int max_iter, double tol,
                       int precision, int instances_per_wg, int* iters, int* status, pfgpu_stats* stats);
/* The same on device-resident arrays (d_V0 must not alias d_V; Yx_stride 0 or nnz_ybus per chain). */
int pfgpu_solve_chains_device(pfgpu_ctx* ctx, pfgpu_grid* grid, int nrows, int chain_len, const double* d_Yx,
                              int Yx_stride, const double* d_Sbus, const double* d_V0, int v0_stride, double* d_V,
                              int max_iter, double tol, int precision, int instances_per_wg, int* d_iters,
                              int* d_status, pfgpu_stats* stats);

/* Per-row PFGPU_FLAG_* of the last solve call on the context (of its first pass: the FP32
 * launch for PFGPU_FP32; for a single chain in batched passes, of the pass that gave each row);
 * nrows must not exceed that call's rows. */
int pfgpu_ctx_last_flags(pfgpu_ctx* ctx, int* flags, long long nrows);

/*
 * Launch autotuner: runs the batch described by the arguments (same meaning as
 * pfgpu_solve_batch_device; outputs are overwritten) once per candidate launch
 * shape — instances per workgroup G in {1, 2, 4, 8, 16, 32} x block size in
 * {256, 1024} — keeps the fastest for this grid and precision, and returns it.
 * Later solves with instances_per_wg == 0 use the tuned shape. The FP32 forms are
 * timed without the FP64 re-solve pass.
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

/*
 * Batched branch flows of solved rows with the contingency clean-up, on device arrays
 * (BaseBatchSolverSynch::compute_amps_flows / compute_active_power_flows,
 * BaseBatchSolverSynch.hpp:86-192, then ContingencyAnalysis::clean_flows,
 * ContingencyAnalysis.cpp:237-250). Output row r (of nrows) uses row d_row_src[r] of d_V
 * ([rows x n] complex; d_row_src NULL: row r; -1: a row of zero voltages, whose amps are
 * NaN as in the reference). Branch b has solver buses d_f[b] / d_t[b] (d_f[b] < 0: out of
 * service, flows 0), complex yff / yft and vn_kv of its origin bus.
 *   amps = (|S| sn_mva) / (sqrt(3) (|Ef| vn_kv)), p = Re(S) sn_mva, S = Ef conj(yff Ef + yft Et).
 * d_clean_ptr[nrows + 1] / d_clean_br (or both NULL) list per row the branches whose finite
 * flows are set to 0 (the outaged branches of a contingency).
 */
int pfgpu_batch_flows_device(pfgpu_ctx* ctx, long long nrows, int n, const double* d_V, const int* d_row_src, int nbr,
                             const int* d_f, const int* d_t, const double* d_yff, const double* d_yft,
                             const double* d_vn_kv_from, double sn_mva, const int* d_clean_ptr, const int* d_clean_br,
                             double* d_amps, double* d_p);

/*
 * The same for rows solved by the DC solver (the !is_ac branches of compute_amps_flows /
 * compute_active_power_flows, BaseBatchSolverSynch.hpp:96-135 and :151-186): d_yff / d_yft are the
 * branches' DC coefficients (1/x and -1/x, ydc_11 / ydc_12), and
 *   p = (Re(yff) arg(Ef) + Re(yft) arg(Et)) sn_mva,  amps = |p| / (sqrt(3) (|Ef| vn_kv)).
 * A row of zero voltages gives p = 0 and NaN amps, as in the reference.
 */
int pfgpu_batch_flows_dc_device(pfgpu_ctx* ctx, long long nrows, int n, const double* d_V, const int* d_row_src,
                                int nbr, const int* d_f, const int* d_t, const double* d_yff, const double* d_yft,
                                const double* d_vn_kv_from, double sn_mva, const int* d_clean_ptr,
                                const int* d_clean_br, double* d_amps, double* d_p);

/*
 * Time-series Sbus rows on the device (TimeSeries::compute_Vs, TimeSeries.cpp:45-53):
 *   Sbus[t, b] = sum gen_p[t, g] + sum sgen_p[t, s] - sum load_p[t, l] - j sum load_q[t, l]
 * over the elements at solver bus b (CSR lists ptr[ns + 1] / element ids, in element order),
 * divided by sn_mva unless it is 1 (within 1e-7). Injection matrices row-major [T x elements];
 * d_Sbus [T x ns] complex.
 */
int pfgpu_sbus_rows_device(pfgpu_ctx* ctx, int T, int ns, const int* d_gen_ptr, const int* d_gen_el, int ngen,
                           const double* d_gen_p, const int* d_sgen_ptr, const int* d_sgen_el, int nsgen,
                           const double* d_sgen_p, const int* d_load_ptr, const int* d_load_el, int nload,
                           const double* d_load_p, const double* d_load_q, double sn_mva, double* d_Sbus);

/*
 * Ybus values from branch parameters on a fixed CSC pattern (the Ybus part of
 * GridModel::pre_process_solver, GridModel.cpp:450-571; line stamps of LineContainer.cpp:13-34):
 * branch k: ys = (r - j x) / (r^2 + x^2), yff = ytt = ys + h/2, yft = ytf = -ys. Entry e sums
 * the contributions d_contrib[d_contrib_ptr[e] .. d_contrib_ptr[e+1]) in list order,
 * contribution c = 4 k + {0 ff, 1 tt, 2 ft, 3 tf}; branches with d_on[k] == 0 add nothing.
 * d_h complex [nbr]; d_Yx [nnz] complex out.
 */
int pfgpu_ybus_values_device(pfgpu_ctx* ctx, int nnz, const int* d_contrib_ptr, const int* d_contrib, int nbr,
                             const double* d_r, const double* d_x, const double* d_h, const int* d_on, double* d_Yx);

/*
 * Contingency Ybus edits on the device (ContingencyAnalysis::init_li_coeffs + remove_from_Ybus,
 * ContingencyAnalysis.cpp:47-110): contingency c removes the branches
 * d_cont_br[d_cont_ptr[c] .. d_cont_ptr[c+1]); branch k has CSC positions d_br_pos[4k .. 4k+3]
 * (ff, tt, ft, tf; -1: out of service, no edit) and complex values d_br_y[4k .. 4k+3]. Writes the
 * edits of c at [d_dptr[c], d_dptr[c+1]) of d_dpos / d_dval (d_dptr given: 4 per in-service
 * branch), the delta format of the solve entry points.
 */
int pfgpu_contingency_edits_device(pfgpu_ctx* ctx, int ncont, const int* d_cont_ptr, const int* d_cont_br,
                                   const int* d_br_pos, const double* d_br_y, const int* d_dptr, int* d_dpos,
                                   double* d_dval);

/*
 * Connectivity of the grid after each contingency (ContingencyAnalysis::check_invertible,
 * ContingencyAnalysis.cpp:14-45): breadth-first search from bus 0 over the Ybus entries whose
 * edited value (Yx minus the contingency's deltas, NULL deltas: none) has |y| > 1e-8;
 * d_connected[c] = 1 when every bus is reached. One workgroup per contingency; the grid must
 * fit the LDS (n up to about 16000 buses).
 */
int pfgpu_connectivity_device(pfgpu_ctx* ctx, int ncont, int n, const int* d_Yp, const int* d_Yi, const double* d_Yx,
                              const int* d_dptr, const int* d_dpos, const double* d_dval, int* d_connected);

/*
 * Serving loop (continuous batching over grids resident on the device). Requests on one grid
 * that are queued while a launch runs join the next launch, up to max_instances_per_launch and
 * what the HBM budget admits; each grid has its own worker, stream and context, so launches of
 * different grids overlap on the device. No counterpart in the reference, whose batch entry
 * points (ContingencyAnalysisCPP.compute / TimeSeriesCPP.compute_Vs, src/main.cpp:1186,1234)
 * are synchronous calls.
 *   hbm_budget 0: a quarter of the free device memory; precision as for the solve calls
 */
typedef struct pfgpu_server pfgpu_server;
#define PFGPU_REQ_PF 0            /* one power flow (Sbus NULL: the grid's resident Sbus) */
#define PFGPU_REQ_CONTINGENCIES 1 /* ninst outages as Ybus deltas (dptr[ninst + 1], dpos, dval) */
#define PFGPU_REQ_TIME_SERIES 2   /* ninst Sbus rows, each warm-started from the resident V */
typedef struct pfgpu_request_info {
    int ninst;             /* rows of the result */
    int launch_instances;  /* instances of the launch the request rode in */
    double latency_s;      /* submit -> result */
    double queue_s;        /* submit -> launch start */
} pfgpu_request_info;
int pfgpu_server_create(size_t hbm_budget, int max_instances_per_launch, int precision, pfgpu_server** out);
/* Stops the workers; requests still queued complete with an error. */
int pfgpu_server_destroy(pfgpu_server* srv);
/* Loads a grid (pattern analysis, schedules) and keeps its Ybus values, base Sbus and warm-start
 * V0 resident; grid_id receives its id. Runs a warm-up launch (one instance per compute unit at
 * the resident injections, in each precision form a launch may take) before it returns. */
int pfgpu_server_add_grid(pfgpu_server* srv, int n, const int* Yp, const int* Yi, const double* Yx, const double* Sbus,
                          const double* V0, const int* slack_ids, int nslack, const double* slack_weights, const int* pv,
                          int npv, const int* pq, int npq, int max_iter, double tol, int* grid_id);
int pfgpu_submit(pfgpu_server* srv, int grid_id, int kind, int ninst, const double* Sbus, const int* dptr,
                 const int* dpos, const double* dval, uint64_t* request_id);
/* Waits up to timeout_s (0: no wait, < 0: until done). ready = 1: the result was copied into V
 * [ninst x n complex], iters and status [ninst] (any may be NULL) and the request is retired;
 * ready = 0: not done yet. A failed request returns an error code once ready. */
int pfgpu_poll(pfgpu_server* srv, uint64_t request_id, double timeout_s, int* ready, double* V, int* iters, int* status,
               pfgpu_request_info* info);

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
