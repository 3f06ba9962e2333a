// eq_internal.h -- host-side batch state shared by the translation units of
// libeq_field.so.  Not part of the ABI.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <array>
#include <string>
#include <vector>

#include "../../include/eq_field.h"
#include "eq_common.h"

namespace eqk {

// tile table entry: the tile's problem and place, precomputed so that a block
// has its node addresses and its problem's state index after one load
struct TileRef {
    int32_t pi;          // problem index in the level table
    int32_t prob;        // global problem index (Krylov state)
    int32_t gt;          // tile index in the level's full table (partial sums)
    int32_t local;       // tile index inside the problem
    int32_t i0, j0;      // first row and column of the tile
    int32_t pitch, nx, ny;
    int32_t flags;       // bit 0: tile inside the grid and off the walls; bit 1: so is the ring around it
    int64_t base;        // element offset of node (i0, j0) in the level pool
};

// A tile of the fused level kernels (k_level_down, k_level_up): kFusedTileW
// columns by kFusedTileH0 rows on level 0, kFusedTileHc rows on the coarse
// levels (measured: the coarse levels' short launches favour more, shorter
// workgroups, level 0 fewer halo rows), the coarse nodes it owns and their
// restriction support, and both levels' geometry, so a workgroup's first
// loads wait on this record alone.
constexpr int kFusedTileW = 56, kFusedTileH0 = 16, kFusedTileHc = 8;
struct FusedTile {
    int32_t pi, prob;             // problem in the level's table, global problem (Krylov state)
    int32_t gt;                   // level 0: slot of the tile's partial sums (after the level's own tiles')
    int32_t i0, j0;               // first fine row and column
    int32_t Ia, Ib, Ja, Jb;       // coarse rows [Ia, Ib) and columns [Ja, Jb) owned
    int32_t ilo, ihi, jlo, jhi;   // fine rows and columns of their restriction support
    int32_t inner;                // bit 0: the way down's region holds interior nodes only; bit 1: the way up's
    int32_t nx, ny, pitch, dmask; // fine level
    int32_t cnx, cny, cpitch, cdmask; // coarse level
    float rnxf, rnyf, crnxf, crnyf;
    double rnx, rny;
    int64_t origin, corigin;
};

struct LevelView {
    const LevelProb* probs;
    const TileRef* tiles;
    int32_t ntiles;
};

// per-problem Krylov scalars, device resident
struct CgState {
    double rz, pq, rr, bb, alpha, beta, rr_true, rzold;
    double tol2;            // rtol^2 (times bb when compared)
    int32_t iter, done, max_iter, flexible;
    int32_t restart;        // next beta is 0 (after a residual replacement)
    int32_t have_prev;      // the previous solution is kept: start CG from it
    int32_t have_prev2;     // so is the one before: start from their linear extrapolation
    int32_t extrap;         // extrapolation allowed (solver setting)
    // BiCGStab: rho = (r^, r), omega, and the pieces of omega
    double rho, omega, ts, tt;
};

// One group of the batch's problems as the CG launches see it: global
// problems [p0, p1); level-0 problem records [pi0, pi0 + npi) (the finish
// tables); the level-0 tiles; per level, the fused tiles [f0, f0 + nf) and
// the bottom workgroups [r0, r0 + nr).  Every range is contiguous: the level
// tables are built problem by problem.
struct GroupView {
    int32_t p0 = 0, p1 = 0, pi0 = 0, npi = 0;
    int32_t t0 = 0, nt = 0;
    std::vector<int32_t> f0, nf, r0, nr;
    double nodes = 0.0;
    // per level, the group's smoothed, active and bottom nodes and its share
    // of the coarsest inverses' bytes (the launches' algorithmic bytes)
    std::vector<double> nsm, nact, nbot, dense;
};

// one bottom V-cycle workgroup: its problem's level records (bottom level,
// then each coarser level down to the coarsest) start at rec in the level's
// bottom record table; nl coarser levels follow the bottom one
struct BottomRef {
    int32_t rec, nl, prob;   // first record of the chain, levels under the bottom level, global problem
    int32_t nx, ny, pitch;   // the bottom level (right-hand side and correction in the level pool)
    int32_t lds_total;       // LDS elements of the chain's level arrays
    int32_t nz;              // nodes of the coarsest level
    int64_t origin;          // the bottom level's node (0,0) in the pool
    int64_t dense_off;       // the coarsest level's inverse in the dense pool
};

// one coarsest-level inverse rebuilt on the device (cell-driven tensor)
struct InvJob {
    int32_t lev, idx;    // the problem's coarsest level and its index in that level's table
    int64_t scratch;     // offset (doubles) of its assembly scratch: A (n*n), X (n*n), tensor (3n)
};

// one boundary well (Simulation::computeBoundaryWell, src/simulation.cpp:581-604)
struct WellDev {
    int32_t prob;        // problem index (its level-0 table entry)
    int32_t pad_;
    double scaling;      // wellScaling
    double decay;        // boundaryDecayRate
    double dt;
    double w;            // boundaryWellConcentration
    double flux;         // the last totalBoundaryFlux it took
    int64_t underflows;  // boundaryUnderFlow
};

struct Level {
    std::vector<LevelProb> probs;     // host copy
    std::vector<TileRef> tiles_all;   // every problem at this level
    std::vector<TileRef> tiles_smooth;// problems smoothed here by grid launches (above their bottom level)
    std::vector<TileRef> tiles_bottom;// problems whose bottom level is this one
    std::vector<TileRef> tiles_active;// problems at or above their bottom level (restriction targets)
    std::vector<FusedTile> fused;     // problems smoothed here, in the fused kernels' tiles
    std::vector<LevelProb> probs_fin; // level 0: host copy of the table behind d_probs_fin
    std::vector<int32_t> bottom;      // problems whose bottom level is this one
    std::vector<BottomRef> bottom_refs;   // one per bottom problem
    std::vector<LevelProb> bottom_recs;   // copies of their level records, chain by chain
    LevelProb* d_probs = nullptr;
    TileRef* d_tiles_all = nullptr;
    TileRef* d_tiles_smooth = nullptr;
    TileRef* d_tiles_bottom = nullptr;
    TileRef* d_tiles_active = nullptr;
    FusedTile* d_fused = nullptr;
    LevelProb* d_probs_fin = nullptr; // level 0: the level table with each fused problem's partial-sum slots
    BottomRef* d_bottom = nullptr;
    LevelProb* d_bottom_recs = nullptr;
    int R = 1;                        // rows per thread of the level's tiles (tile height 4R)
    bool nested_down = true;          // every problem smoothed here halves exactly to the next level
    int64_t pool = 0;                 // elements per vector on this level
    double nodes_all = 0.0;           // grid nodes summed over the level's problems
    double nodes_smooth = 0.0;        // ... over the problems that smooth here
    double nodes_active = 0.0;        // ... over the problems at or above their bottom level
    double nodes_bottom = 0.0;        // ... over the problems whose bottom level this is
    double bottom_dense_bytes = 0.0;  // their coarsest inverses (read by the bottom V-cycle)
    size_t bottom_lds_bytes = 0;      // dynamic LDS of the bottom V-cycle launched from this level
    // multigrid vectors (element type = smoother precision)
    void* b = nullptr;  // level >= 1, or level 0 FP32 copy of r
    void* xa = nullptr;
    void* xb = nullptr;
    void* r = nullptr;  // residual (levels that restrict it in a separate launch)
    void* tens = nullptr; // tensor field D11|D22|D12 in smoother precision (allocated on first use)
    LevelView all() const { return LevelView{d_probs, d_tiles_all, (int32_t)tiles_all.size()}; }
    LevelView smooth() const { return LevelView{d_probs, d_tiles_smooth, (int32_t)tiles_smooth.size()}; }
    LevelView bottom_tiles() const { return LevelView{d_probs, d_tiles_bottom, (int32_t)tiles_bottom.size()}; }
    LevelView active() const { return LevelView{d_probs, d_tiles_active, (int32_t)tiles_active.size()}; }
};

// algorithmic bytes and launches of a launch sequence
struct Acct {
    double bytes = 0.0;
    int64_t launches = 0;
};

struct ProfileSlot {
    bool enabled = false;
    int64_t launches = 0;
    double total_ms = 0.0;
    double bytes = 0.0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    std::vector<double> pending_bytes;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
};

} // namespace eqk

struct eq_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t nprob = 0, nspecies = 1, ntraps = 0;
    std::vector<eq_field_desc> descs;
    eq_solver_desc solver{};
    int32_t mixed = 0;                 // 1: FP32 V-cycle
    bool any_tensor = false;           // some problem has per-node rows (tensor field or advection)
    std::vector<std::vector<std::vector<double>>> vel_host; // per problem, per level: row velocities
    std::vector<void*> vel_bufs;       // device copies of the row velocities: [(prob * levels cap + level) * 2 + {0: V-cycle precision, 1: FP64 on level 0}], allocated on first use and reused
    int64_t lds_budget = 64 * 1024;    // LDS bytes the bottom V-cycle may use per workgroup
    int64_t lds_budget_max = 64 * 1024; // ... at most (a coarsest level above lds_budget takes this)
    std::vector<eqk::Level> levels;
    // level 0 FP64 Krylov vectors (pooled over problems)
    double* u = nullptr;   // field / iterate x
    double* rhs = nullptr;
    double* r = nullptr;
    double* p = nullptr;   // search direction of the previous iteration
    double* p2 = nullptr;  // search direction being formed (swapped with p every iteration)
    double* q = nullptr;
    double* z = nullptr;
    double* xprev = nullptr;    // last solution, the next solve's starting iterate
    double* xprev2 = nullptr;   // the solution before it (extrapolated start)
    // BiCGStab vectors (allocated with the batch for the BiCGStab kinds)
    double* bs_rhat = nullptr;
    double* bs_p = nullptr;
    double* bs_v = nullptr;
    double* bs_s = nullptr;
    double* bs_t = nullptr;
    hipGraphExec_t g_bicg = nullptr;
    double* tens0_d = nullptr;  // level-0 tensor field D11|D22|D12 in FP64 (allocated on first use)
    double* partial = nullptr;  // 4 slots per level-0 tile
    void* dense = nullptr;      // dense inverses of coarsest levels (smoother precision)
    eqk::CgState* d_state = nullptr;
    eqk::CgState* h_state = nullptr; // pinned mirror
    // iterative refinement (EQ_SOLVER_IR_GMG_CG): the inner FP32 solve's
    // Krylov state, its correction and residual (FP32), and the refinement's
    // scale per problem (FP64 residual norm before the inner solve)
    eqk::CgState* d_state_in = nullptr;
    eqk::CgState* h_state_in = nullptr;
    float* ir_d = nullptr;
    float* ir_r = nullptr;
    double* ir_scale = nullptr;
    hipGraphExec_t g_ir2 = nullptr;
    eqk::Acct g_ir2_acct;
    int32_t last_iters = 0;
    // captured launch sequences (created on first use): the set-up, and CG
    // iterations: g_cg[g][f][d][n][t] runs, for problem group g, the solve's
    // set-up (f = 1), then n iterations (n a power of two up to kCgGraphMax) from search
    // direction p (d = 0) or p2 (d = 1), then (t = 1) the convergence
    // confirmation
    static constexpr int kCgGraphMax = 16;
    static constexpr int kMaxGroups = 4;
    hipGraphExec_t g_cg[kMaxGroups][2][2][kCgGraphMax + 1][2] = {};
    eqk::Acct g_cg_acct[kMaxGroups][2][2][kCgGraphMax + 1][2];
    // CG problem groups solved concurrently (solver option `streams`): the
    // views of each (built for groups_for groups on first use; empty when the
    // batch cannot be split), the streams of groups 1.. (group 0 runs on the
    // batch's stream), and the events that order the streams around a solve
    std::vector<eqk::GroupView> groups;
    int32_t groups_for = 0;
    hipStream_t gstream[kMaxGroups] = {};
    hipEvent_t grp_start = nullptr;
    hipEvent_t grp_done[kMaxGroups] = {};
    int32_t last_iters_g[kMaxGroups] = {};
    hipGraphExec_t g_init = nullptr;
    std::vector<int32_t> iters;
    std::vector<double> relres;
    // cell coupling scratch (grown on demand)
    void* cell_scratch = nullptr;
    size_t cell_scratch_bytes = 0;
    double** d_field_ptrs = nullptr; // device table: node (0,0) of every problem's field
    void* d_cell_traps = nullptr;    // coupling tile grid of every trap (device)
    std::vector<int32_t> cell_tile_base; // first coupling tile of every trap, and the total
    int64_t* h_cell_off = nullptr;   // pinned staging of a call's cell offsets
    // pair buffers of the tile coupling, sized before a call's pair count is
    // known: pair_cap slots (0: size the next call from its count, read back);
    // the last call's count arrives in h_pair_total (pinned, written by a
    // kernel) once pair_ev has completed
    int64_t pair_cap = 0;
    int64_t* h_pair_total = nullptr;
    hipEvent_t pair_ev = nullptr;
    bool pair_pending = false;
    hipEvent_t cell_off_ev = nullptr; // the last upload of h_cell_off (reused only once it has run)
    int64_t* d_cell_off = nullptr;
    double* flux_out = nullptr;      // device scalar
    double* h_flux = nullptr;        // pinned
    eqk::ProfileSlot prof[EQ_KERNEL_COUNT_];
    // per-node wall values and flow channels (allocated on first use)
    std::vector<std::array<double*, 4>> wall_vals;        // per problem and wall
    std::vector<std::array<eq_channel_desc, 2>> chan_desc; // per problem: top, bottom (substeps 0: none)
    void* d_chans = nullptr;                              // device channel table
    int32_t chan_count = 0;
    std::vector<double*> chan_scratch;
    // boundary wells (allocated on first use): the device table is the live
    // state; the host keeps which problems have one and their parameters
    std::vector<eqk::WellDev> wells;
    eqk::WellDev* d_wells = nullptr;
    int32_t d_wells_cap = 0;
    // algorithmic traffic: launches add to *acct_to, which is `acct` except
    // while a graph is captured (then the graph's own count, added to `acct`
    // at every replay)
    eqk::Acct acct;
    eqk::Acct* acct_to = &acct;
    eqk::Acct g_init_acct, g_bicg_acct;
    // tensors set from the cells' angles (allocated on first use): per problem,
    // node (0,0) of its finest-level tensor in FP64 and (FP32 V-cycle) FP32;
    // the coarsest inverses to rebuild after each such coupling
    double** d_tens64 = nullptr;
    float** d_tens32 = nullptr;
    std::vector<int32_t> cell_tensor_traps;
    std::vector<eqk::InvJob> inv_jobs;
    eqk::InvJob* d_inv_jobs = nullptr;
    double* inv_scratch = nullptr;
    eqk::LevelProb** d_level_tables = nullptr;
};

namespace eqk {
int set_error(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);
// flow channels: update after a solve; release on destroy
int channels_after_step(eq_batch* b);
void free_channels(eq_batch* b);
// the level-0 Dirichlet values a well wrote on the device, copied back into
// the host records before the host uploads its copy of the table
int pull_well_values(eq_batch* b);
// boundary wells: update after a step
int wells_after_step(eq_batch* b);
// cell-driven tensors: after the coupling wrote the finest level, carry it to
// the coarse levels and rebuild the coarsest inverses (eq_solver.hip)
int refresh_cell_tensors(eq_batch* b);
// the coupling's trap table entry of `trap` (eq_cells.hip)
int set_trap_tensor(eq_batch* b, int trap, double Dx, double Dy, int on);
// the coupling's tables and the fixed scratch's device address (at batch creation)
int init_coupling(eq_batch* b);
} // namespace eqk

namespace eqk {
// Every kernel launch of the library: counts its algorithmic bytes, and when
// timing is on for its family `kid`, carries start and stop events
// (hipExtLaunchKernelGGL) that bracket the kernel's own execution, not the
// stream's gaps around it.
template <typename K, typename... Args>
inline void launch(eq_batch* b, int kid, double bytes, K kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                   Args... args)
{
    b->acct_to->bytes += bytes;
    b->acct_to->launches += 1;
    if (kid < 0 || kid >= EQ_KERNEL_COUNT_ || !b->prof[kid].enabled) {
        hipExtLaunchKernelGGL(kernel, grid, block, lds, s, nullptr, nullptr, 0, args...);
        return;
    }
    ProfileSlot& ps = b->prof[kid];
    std::pair<hipEvent_t, hipEvent_t> ev;
    if (!ps.pool.empty()) {
        ev = ps.pool.back();
        ps.pool.pop_back();
    } else {
        hipEventCreate(&ev.first);
        hipEventCreate(&ev.second);
    }
    hipExtLaunchKernelGGL(kernel, grid, block, lds, s, ev.first, ev.second, 0, args...);
    ps.pending.push_back(ev);
    ps.pending_bytes.push_back(bytes);
}
// a replayed graph's traffic
inline void replayed(eq_batch* b, const Acct& g)
{
    b->acct.bytes += g.bytes;
    b->acct.launches += g.launches;
}
} // namespace eqk

#define EQ_HIP(call)                                                  \
    do {                                                              \
        hipError_t e_ = (call);                                       \
        if (e_ != hipSuccess) return eqk::hip_check(e_, #call);       \
    } while (0)
