// Host-side analysis for the batched GPU Newton-Raphson engine.
//
// Everything here is computed once per (grid pattern, slack/pv/pq set) and is
// shared by every instance of a batch:
//   * the Jacobian pattern in the reference's layout and the scatter map from
//     each J entry to a Ybus entry + {dVa,dVm}x{Re,Im} selector
//     (reference: src/powerflow_algorithm/BaseNRAlgo.tpp:396-635, single slack
//     BaseNRSingleSlackAlgo.tpp:247-...);
//   * a fill-reducing symmetric ordering of J (minimum degree on the bus graph,
//     each bus carrying its 1 or 2 unknowns as one block);
//   * the symbolic LU with static diagonal pivots on the symmetric pattern
//     (this plays the role of klu_analyze, src/linear_solvers/KLUSolver.cpp:31);
//   * the level schedule that the HIP kernels execute: Crout-style dot-product
//     tasks for the numeric refactorisation fused with the forward solve, and
//     the backward-solve levels.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace pfgpu {

using cplx = std::complex<double>;

// Jacobian entry sources: J value = Re/Im of dS_dVa or dS_dVm at a Ybus entry.
enum JSel : uint8_t { SEL_DVA_RE = 0, SEL_DVA_IM = 1, SEL_DVM_RE = 2, SEL_DVM_IM = 3, SEL_CONST = 4 };

struct JPattern {
    int n_bus = 0;
    int N = 0;       // J dimension
    int off = 0;     // 1 for the distributed-slack layout (row/col 0 = slack), 0 for single slack
    int slack0 = 0;  // reference slack bus
    std::vector<int> my_pv, pq, pvpq;
    std::vector<int> Jp, Ji;     // CSC, rows sorted (reference layout)
    std::vector<int> src;        // Ybus CSC value index per J entry (-1 for constants)
    std::vector<uint8_t> sel;    // JSel per J entry
    std::vector<double> cval;    // constant value per J entry (slack-weight column)
    // Row of the factored matrix that J row r becomes (a row permutation chosen by pivoting on
    // the numbers, threshold_pivot_rows); empty: J is factored with its own diagonal.
    std::vector<int> rowpos;
    int row_of(int r) const { return rowpos.empty() ? r : rowpos[r]; }
};

// Ybus in CSC (rows sorted per column), as the reference takes it.
struct YbusPattern {
    int n = 0;
    std::vector<int> p, i;
};

JPattern build_jacobian_pattern(const YbusPattern& Y, const int* slack_ids, int nslack, const double* slack_w,
                                const int* pv, int npv, const int* pq, int npq, bool single_slack);

// One Crout task computes one entry of the factor (or of the forward-solved
// right-hand side) as  v = a - sum_p  X[lp] * X[up] * X[dp]  over its pair range,
// where X is the per-instance value array, X[dp] holds 1/d_k for the pivot k of
// the pair (L is kept unscaled). Diagonal tasks store 1/v instead of v.
struct SymbolicLU {
    int N = 0;
    std::vector<int> perm;      // perm[new] = old J index
    std::vector<int> pinv;      // pinv[old] = new
    std::vector<int> parent;    // elimination tree (new indices), -1 = root
    // column "arms": for new column k, sorted rows R_k (> k) of L(:,k) == cols of U(k,:)
    std::vector<int> cp, ci;    // CSC of strictly-lower L pattern (new indices)
    // value array layout (per instance): [ diag(0..N-1) | Lvals(nnzL) | Uvals(nnzL) | y(0..N-1) ]
    int64_t nnzL = 0;
    int64_t nvals = 0;          // total value slots per instance
    int64_t off_L = 0, off_U = 0, off_y = 0;
    // J entry t -> value slot
    std::vector<int> jslot;
    // dense tail: the trailing m = N - k0 columns (the top chain of the elimination tree) are
    // factored as one dense m x m block; the Crout tasks only apply pivots k < k0 to them
    int k0 = 0, m = 0;
    // flops of one refactorisation (multiply-adds counted as 2, divisions 1)
    double flops = 0;
};

// max_tail: largest dense tail (0 disables it); perm: a given elimination order (perm[k] = J
// column eliminated k-th) instead of the bus-block minimum degree
SymbolicLU symbolic_lu(const JPattern& J, int max_tail = 96, const std::vector<int>* perm = nullptr);

// The bus-block minimum-degree elimination order of J (perm[k] = J column eliminated k-th).
std::vector<int> column_order(const JPattern& J);

// Host re-pivot: a left-looking sparse LU with threshold partial pivoting (Gilbert and Peierls'
// algorithm, the pivot rule of KLU: the diagonal row while its magnitude is at least tau times the
// column's largest candidate, else the largest) of J's values Jx, columns in the order `perm`.
// Returns rowpos (J row -> the J column it is the pivot of), or an empty vector when a column has
// no nonzero candidate (numerically singular).
std::vector<int> threshold_pivot_rows(const JPattern& J, const std::vector<double>& Jx, const std::vector<int>& perm,
                                      double tau = 1e-3);

constexpr int kTaskAssign = 4;  // DevicePlan task kind flag: the task assigns its value

struct Schedule {
    // Factor + forward-solve tasks, grouped by level.
    std::vector<int> lvl_ptr;          // level l covers tasks [lvl_ptr[l], lvl_ptr[l+1])
    std::vector<int> task_out;         // value slot written
    std::vector<int> task_pbeg;        // pair range begin (pair ranges are contiguous: end = next begin)
    std::vector<int> task_pend;
    std::vector<uint8_t> task_diag;    // 1 => store reciprocal
    std::vector<int> pair_l, pair_u, pair_d;  // value slots of the unscaled L entry, the U/y entry, and 1/d
    // Backward solve: x_j = (y_j - sum_i u_ji x_i) / d_j, levels from the root down.
    std::vector<int> blvl_ptr;
    std::vector<int> bcol;             // column j per backward task
    // statistics
    int max_pairs_per_task = 0;
    int64_t npairs = 0;
};

Schedule build_schedule(const SymbolicLU& S);

// Flattened arrays uploaded to the device. Value slots are renumbered by their
// accesses per NR iteration (the most read and written first, after the dense
// tile), so that a launch can keep the first `nlds` slots of every instance in
// LDS and spill the rest to a per-instance global buffer.
struct DevicePlan {
    int n = 0, N = 0, off = 0, nnzY = 0, nJ = 0, nvals = 0, slack0 = 0;
    // dense tail: slots [0, m*(m+1)) hold the m x m block row-major with row stride m + 1,
    // its forward/backward vector as column m
    int m = 0;
    int zslot = 0;  // a slot that always holds 0 (zeroed by the J fill, never written otherwise)
    int nlev = 0, nblev = 0;
    // Ybus row view for Ibus = Y V: row r covers [yr_ptr[r], yr_ptr[r+1]) with column yr_col and CSC value pos yr_pos
    std::vector<int> yr_ptr, yr_col, yr_pos;
    // per bus: value slot of the P-row / Q-row of F (-1 if the bus has no such equation)
    std::vector<int> fslot_p, fslot_q;
    // per J unknown (original J column order): bus, kind (0 slack_absorbed, 1 Va, 2 Vm), value slot of dx
    std::vector<int> ux_bus, ux_slot;
    std::vector<int8_t> ux_kind;
    // J entries (reference CSC order): Ybus CSC pos, row bus, col bus, selector, constant,
    // and the value slot the J-fill phase writes (bit 30 set: store the reciprocal, for a
    // pivot whose column has nothing to eliminate)
    std::vector<int> j_ypos, j_rbus, j_cbus, j_slot;
    std::vector<uint8_t> j_sel;
    std::vector<double> j_cval;
    // J fill by Ybus entry: each listed Ybus entry (CSC position ye_pos, row bus / column
    // bus in ye_rc) feeds up to 4 J entries: Re and Im of dS/dVa, Re and Im of dS/dVm.
    // ye_slot holds their value slots (-1: absent; bit 30: store the reciprocal) and
    // ye_jidx their J entry indices. The constant J entries (slack-weight column) are
    // listed apart in jc_slot / jc_val.
    // The diagonal entries come first (nyd of them, in bus order); ye_diag_of[bus] is the
    // index of the bus's diagonal entry, -1 if it has none.
    std::vector<int> ye_pos, ye_rc, ye_slot, ye_jidx;
    int nyd = 0;
    std::vector<int> ye_diag_of;
    std::vector<int> jc_slot, jc_jidx;
    std::vector<double> jc_val;
    // dense-tile slots that hold no J entry (fill-in): zeroed by the J-fill phase
    std::vector<int> zero_slots;
    // factor + forward tasks (level-major), 4 ints each: {out slot, pair begin, pair end, kind};
    // kind & 3: 0 = L/U entry, 1 = pivot (stores the reciprocal), 2 = forward-solve entry;
    // kind & kTaskAssign: a fill-in entry outside the dense tile (no J entry: it assigns -sum,
    // its slot is recycled; under a row permutation this can be a pivot as well).
    // Other tasks start from the value already in their slot (J entry, 0 for tile
    // fill-in, or F) and subtract sum X[l]*X[u]*X[d] over their pairs. Tasks without
    // pairs are not listed. Value slots are shared by values with disjoint lifetimes.
    std::vector<int> f_lvl;                 // nlev + 1
    std::vector<int> f_maxp;                // per level max pair count
    std::vector<int> tasks;                 // 4 * ntask
    std::vector<int> pairs;                 // 4 * npairs: {l, u, d, 0}
    // backward tasks (level-major), 4 ints each: {x slot, pair begin, pair end, slot of 1/d};
    // x = (x - sum X[u]*X[x']) * X[1/d]
    std::vector<int> b_lvl, b_maxp;
    std::vector<int> btasks;                // 4 * N
    std::vector<int> bpairs;                // 2 * nbpairs: {u, x'}
    double flops = 0;
    int64_t nnzLU = 0;
};

DevicePlan build_device_plan(const YbusPattern& Y, const JPattern& J, const SymbolicLU& S, const Schedule& sch);

// Per-level team sizes (lanes per task, power of two, team * G <= 64) for a
// launch with G interleaved instances and `nslot` lanes per instance.
std::vector<int> level_teams(const std::vector<int>& lvl_ptr, const std::vector<int>& maxp, int G, int nslot);

// Lane streams: the factor/forward and backward task lists of a plan laid out
// for one launch shape (G interleaved instances, nslot lanes per instance).
// Level l uses lanes [0, U_l); item k of lane s sits at base_l + k * U_l + s, so
// the address of every work item a lane reads depends on no earlier load: a
// level's items can be fetched before the barrier that precedes it.
// Each task is shared by a team of T lanes (power of two, T * G <= 64, aligned
// to T); every lane of the team holds ceil(pairs / T) consecutive items for it,
// the last of which marks the task's end and carries log2(T). Per level, the
// smallest item count K per lane for which the tasks fit is searched (each task
// takes the smallest team that holds it in K items). A lane's last item of a
// level is flagged, so that it loads nothing after it.
// End word (forward: item[3], backward: item[2]):
//   bits 0-24 slot, 25 assign (forward fill-in: the slot takes -sum, its old
//   content belongs to another value), 26-28 log2(T), 29 reciprocal (forward
//   pivots), 30 end of task, 31 the lane's last item of the level.
constexpr int kStreamTeamShift = 26;
constexpr int kStreamSlotMask = (1 << 25) - 1;
constexpr int kStreamAssign = 1 << 25;
constexpr int kStreamRecip = 1 << 29;
constexpr int kStreamEnd = 1 << 30;
constexpr int kStreamLast = (int)(1u << 31);

struct LaneStreams {
    int G = 0, nslot = 0;
    // per level: {base item, K items per lane, largest team size, U used lanes}
    std::vector<int> f_lv, b_lv;
    // lean streams only, per level: the chunk length Lc (see LeanFormat)
    std::vector<int> f_lc, b_lc;
    // forward items {l, u, d, end word}: l < 0 = no pair
    std::vector<int> fs;
    // backward items {u, x', end word, 1/d slot}: u < 0 = no pair
    std::vector<int> bs;
};

// Lean encoding, for launches that keep every value slot in LDS: items hold LDS byte
// offsets (slot * elem_bytes) instead of slot ids, items without a pair point at the
// zero slot (no per-item checks), the lanes of a level are rounded up to lane_round
// (a whole wave), and no last-item flags are set (every lane of a used wave walks all K
// items; the padding items add zeros). Task ends are aligned: a level's K items per lane
// form K / Lc chunks of Lc items, each task takes one chunk of a team of T lanes (T the
// smallest power of two with ceil(pairs / T) <= Lc), so every lane's task ends fall on
// the items Lc - 1, 2 Lc - 1, ..., the same for all lanes, and the kernel accumulates
// the other items without any test. The end word of a chunk of a lane without a task
// is 0. A lean stream starts with one padding item and ends with nslot of them, so that
// lanes may read past their level's lanes and levels.
struct LeanFormat {
    int elem_bytes = 4;  // bytes between consecutive slots of one instance (G * value size)
    int lane_round = 1;  // lanes per wave of one instance (64 / G)
    int head_items = 4;  // items per lane of a level that the kernel loads during the level before
};

LaneStreams build_lane_streams(const DevicePlan& P, int G, int nslot, const LeanFormat* lean = nullptr);

std::string describe(const JPattern& J, const SymbolicLU& S, const Schedule& sch);

}  // namespace pfgpu
