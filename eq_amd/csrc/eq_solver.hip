// eq_solver.hip -- batched implicit P1 reaction-diffusion solve on gfx950.
//
// Replaces fenicsInterface::stepDiffusion's LinearVariationalSolver::solve
// (src/fHSL.cpp:98-108): for every problem (species x trap) of a batch,
//   b = M u0 (+ dt f int(phi) + Robin)      hslD linear form  (fenics/hslD.h:3466-3689)
//   (M(1+dt gamma) + dt D K + Robin) u+ = b  hslD bilinear form (fenics/hslD.h:3123-3441)
//   u+ = g on Dirichlet walls                DirichletBC::apply ([ext], src/fHSL.cpp:467-537)
// The Dirichlet rows are eliminated symmetrically (x_D = g held fixed, the
// Krylov vectors live on the free nodes), which gives the solution of the
// row-replaced system the reference factorises, and keeps the operator SPD
// for CG.
//
// Layout in HBM: per level, one pool per vector; problem p owns rows
// [-1, ny+1] x columns [-1, pitch-1] with pitch = roundup(nx+2, 64) elements,
// node (i,j) at origin + i*pitch + j, zero ghost frame, rows 64-element
// aligned.  Kernels tile 64 columns x 4 rows per 256-thread block (one
// wavefront per row segment, coalesced), one launch covers every problem.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "eq_internal.h"

namespace eqk {

constexpr int kMaxLev = 16;          // multigrid levels per problem
constexpr int64_t kBottomLdsMax = 150 * 1024; // cap on the LDS bytes of the levels one workgroup runs
constexpr int64_t kBottomLdsSmall = 40 * 1024; // ... for batches of fewer than kBottomSmallBatch problems
constexpr int kBottomSmallBatch = 64;
constexpr int kBottomThreads = 1024;
constexpr int kFinishThreads = 1024;
constexpr double kFuseMaxNodes = 512.0 * 1024.0; // levels below this many nodes use the fused multigrid kernels

// ------------------------------------------------------------ errors

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int hip_check(hipError_t e, const char* what)
{
    return set_error(EQ_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------ device helpers

__device__ __forceinline__ int tile_index()
{
    // Workgroups are dealt round-robin to the 8 XCDs; hand each XCD a
    // contiguous run of tiles so vertical neighbours share its L2.
    const int per = gridDim.x >> 3;
    const int b = blockIdx.x;
    return (b & 7) * per + (b >> 3);
}

inline int grid_for(int ntiles) { return 8 * ((ntiles + 7) / 8); }

struct NodeCtx {
    const LevelProb* P;
    int t;      // index in the launched tile table
    int gt;     // tile index in the level's full table (partial sums)
    int pi, i, j;
    int prob;          // global problem index (Krylov state)
    int pitch, nx, ny; // of the problem, in registers for each_row
    int64_t idx;
    bool in;
    bool full;  // the block's whole tile is inside the grid, off the walls
    bool full2; // ... and so is the one-node ring around it
};

// A thread owns R consecutive rows of one column of its tile (the level's
// tile height is kWaves * R): decode gives the first, each_row visits them
// all (unrolled, so the loads of all rows are in flight together and the
// block's set-up is paid once per R nodes).

template <int R>
__device__ __forceinline__ bool decode(const LevelView& L, NodeCtx& c)
{
    c.t = tile_index();
    if (c.t >= L.ntiles) return false;
    const TileRef tr = L.tiles[c.t];
    c.pi = tr.pi;
    c.prob = tr.prob;
    c.P = &L.probs[tr.pi];
    c.gt = tr.gt;
    const int dj = (int)(threadIdx.x & (kTileW - 1)), di = (int)(threadIdx.x / kTileW) * R;
    c.j = tr.j0 + dj;
    c.i = tr.i0 + di;
    c.pitch = tr.pitch;
    c.nx = tr.nx;
    c.ny = tr.ny;
    c.in = (c.i <= c.ny) && (c.j <= c.nx);
    c.full = (tr.flags & 1) != 0;
    c.full2 = (tr.flags & 2) != 0;
    c.idx = tr.base + (int64_t)di * tr.pitch + dj;
    return true;
}

template <int R, typename F>
__device__ __forceinline__ void each_row(const NodeCtx& c, F&& f)
{
#pragma unroll
    for (int r = 0; r < R; ++r) {
        NodeCtx n = c;
        n.i = c.i + r;
        n.idx = c.idx + (int64_t)r * c.pitch;
        n.in = (n.i <= c.ny) && (c.j <= c.nx);
        f(n);
    }
}

// f(0) .. f(N-1) with each index a compile-time constant (decltype(r)::value),
// so per-row register arrays are never indexed at run time
template <int N, typename F>
__device__ __forceinline__ void static_rows(F&& f)
{
    if constexpr (N > 0) {
        static_rows<N - 1>(f);
        f(std::integral_constant<int, N - 1>{});
    }
}

// The R+2 rows i-1 .. i+R of columns j-1, j, j+1 around a thread's R nodes,
// one load per row and column (the rows the R stencils share are read once).
template <int R, typename T>
struct Window {
    T w[R + 2], m[R + 2], e[R + 2];
    __device__ __forceinline__ void load(const T* __restrict__ x, int64_t k0, int pitch)
    {
#pragma unroll
        for (int q = 0; q < R + 2; ++q) {
            const T* p = x + k0 + (int64_t)(q - 1) * pitch;
            w[q] = p[-1];
            m[q] = p[0];
            e[q] = p[1];
        }
    }
    // the same where rows past ny + 1 (outside the problem's frame) may be
    // among them: those read the frame row ny + 1 instead (their values
    // belong to no node of the grid and are not used)
    __device__ __forceinline__ void load_rows(const T* __restrict__ x, int64_t k0, int pitch, int i0, int ny)
    {
#pragma unroll
        for (int q = 0; q < R + 2; ++q) {
            const T* p = x + k0 + (int64_t)(min(i0 - 1 + q, ny + 1) - i0) * pitch;
            w[q] = p[-1];
            m[q] = p[0];
            e[q] = p[1];
        }
    }
    // A x at row r of the R (interior constant row): centre, W, E, S, N, SW, NE
    template <bool TENS>
    __device__ __forceinline__ T apply(const Row<T>& rc, int r) const
    {
        return row_apply<TENS>(rc, m[r + 1], w[r + 1], e[r + 1], m[r], m[r + 2], w[r], e[r + 2]);
    }
};

// Rows of the interior nodes i .. i+R-1 (column j) of a problem without a
// tensor field: the level's constant row, plus the grid row's convection
// with a velocity (what row_coefs gives there).  load() reads the R+2 row
// velocities up front, with the other loads.
template <int R, typename T, bool TENS>
struct PlainRows {
    T v[R + 2];
    __device__ __forceinline__ void load(const LevelProb& P, int i)
    {
        if (TENS && P.vel) {
            const T* vb = vel_base<T>(P);
#pragma unroll
            for (int q = 0; q < R + 2; ++q) v[q] = vb[i - 1 + q];
        }
    }
    __device__ __forceinline__ Row<T> row(const LevelProb& P, int r, int i, int j) const
    {
        Row<T> rc = interior_row<T>(P);
        if (TENS && P.vel) add_convection_v<T>(P, i + r, j, v[r], v[r + 1], v[r + 2], rc);
        return rc;
    }
};
// the inverse diagonal of such a row (row_dinv)
template <typename T, bool TENS>
__device__ __forceinline__ T plain_dinv(const LevelProb& P, const Row<T>& rc)
{
    return (TENS && P.vel) ? (T)1 / row_diag<TENS>(rc) : interior_dinv<T>(P);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// deterministic block sum over NT threads; result valid in thread 0
template <int NV, int NT>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double red[NT / 64][NV];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int q = 0; q < NT / 64; ++q) s += red[q][k];
            v[k] = s;
        }
    __syncthreads();
}

// Write this tile's partial sums (plain stores).  The per-problem totals are
// formed by k_finish in the next launch: the kernel boundary makes the
// partials visible across XCDs, where an in-kernel device-scope fence per
// block would write back the XCD's whole L2 every time.
template <int NV>
__device__ __forceinline__ void tile_partial(double (&v)[NV], double* partial, int gt)
{
    block_sum<NV, kBlock>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) partial[(int64_t)gt * 4 + k] = v[k];
}

enum FinishOp {
    kFinRhs = 0, kFinInit = 1, kFinAp = 2, kFinUpdate = 3, kFinRz = 4, kFinTrue = 5,
    // BiCGStab
    kFinBInit = 6, kFinBAlpha = 7, kFinBOmega = 8, kFinBUpdate = 9, kFinBRestart = 10,
    // iterative refinement: the inner solve's start, the outer residual check
    kFinIrInit = 11, kFinIrOuter = 12,
    // CG confirmation on the device: kFinTrue for the problems that report
    // convergence only; kFinRz for the problems it restarted only
    kFinTrueDone = 13, kFinRzRestart = 14
};

template <typename T>
__device__ __forceinline__ void load7(const T* __restrict__ v, int64_t idx, int pitch, T& c, T& w, T& e, T& s, T& n,
                                      T& sw, T& ne)
{
    c = v[idx];
    w = v[idx - 1];
    e = v[idx + 1];
    s = v[idx - pitch];
    n = v[idx + pitch];
    sw = v[idx - pitch - 1];
    ne = v[idx + pitch + 1];
}

// ------------------------------------------------------------ level-0 Krylov kernels (FP64)

// b = M u0 + sources; ||b||^2 of the row-replaced system (b_D = g)
template <int R0>
__global__ __launch_bounds__(kBlock) void k_rhs(LevelView L, const double* __restrict__ u, double* __restrict__ b,
                                                double* partial)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    const LevelProb& P = *c.P; // reads P.gw[wall]: a register copy would go to scratch
    double v[1] = {0.0};
    if (c.full) {
        // interior tile: no wall terms, no Dirichlet rows; every load first
        // (Q: a register copy, no wall lookups here)
        const LevelProb Q = *c.P;
        Window<R0, double> U;
        U.load(u, c.idx, c.pitch);
        static_rows<R0>([&](auto rr) {
            constexpr int r = decltype(rr)::value;
            const double bv = rhs_row(Q, c.i + r, c.j, U.m[r + 1], U.w[r + 1], U.e[r + 1], U.m[r], U.m[r + 2], U.w[r],
                                      U.e[r + 2]);
            b[c.idx + (int64_t)r * c.pitch] = bv;
            v[0] += bv * bv;
        });
        tile_partial<1>(v, partial, c.gt);
        return;
    }
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        double uc, uw, ue, us, un, usw, une;
        load7(u, n.idx, P.pitch, uc, uw, ue, us, un, usw, une);
        const double bv = rhs_row(P, n.i, n.j, uc, uw, ue, us, un, usw, une);
        b[n.idx] = bv;
        if (is_dirichlet(P, n.i, n.j)) {
            const double g = dirichlet_value(P, n.i, n.j);
            v[0] += g * g;
        } else {
            v[0] += bv * bv;
        }
    });
    tile_partial<1>(v, partial, c.gt);
}

// x_D = g; r = b - A x; Jacobi: z = r / diag(A), r.z.  The iterate starts
// from the previous step's solution when it is kept (xprev), else from the
// field itself: the field carries this step's cell deposits, whose sharp
// edges would otherwise make the first residual ~1e4 times larger.
template <int R0, bool TENS>
__global__ __launch_bounds__(kBlock) void k_init(LevelView L, double* x, const double* __restrict__ xprev,
                                                 const double* __restrict__ xprev2, const double* __restrict__ b,
                                                 double* __restrict__ r, double* __restrict__ z, double* partial,
                                                 const CgState* st, int jacobi)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    const LevelProb& P = *c.P; // reads P.gw[wall]: a register copy would go to scratch
    const bool prev = st[c.prob].have_prev != 0;
    // the starting iterate: the last solution, or the line through the last two
    const bool prev2 = prev && st[c.prob].have_prev2 != 0 && st[c.prob].extrap != 0;
    const double* __restrict__ src = prev ? xprev : x;
    double v[2] = {0.0, 0.0};
    if (c.full2) {
        // the tile and its ring are interior: no Dirichlet node or neighbour;
        // every load first (Q: a register copy, no wall lookups here)
        const LevelProb Q = *c.P;
        Window<R0, double> X;
        X.load(src, c.idx, c.pitch);
        if (prev2) {
            Window<R0, double> X2;
            X2.load(xprev2, c.idx, c.pitch);
#pragma unroll
            for (int q = 0; q < R0 + 2; ++q) {
                X.w[q] = 2.0 * X.w[q] - X2.w[q];
                X.m[q] = 2.0 * X.m[q] - X2.m[q];
                X.e[q] = 2.0 * X.e[q] - X2.e[q];
            }
        }
        double bl[R0];
        static_rows<R0>([&](auto rr) {
            constexpr int w = decltype(rr)::value;
            bl[w] = b[c.idx + (int64_t)w * c.pitch];
        });
        static_rows<R0>([&](auto rr) {
            constexpr int w = decltype(rr)::value;
            const int i = c.i + w;
            const int64_t k = c.idx + (int64_t)w * c.pitch;
            if (prev) x[k] = X.m[w + 1];
            const Row<double> rc = row_coefs<double, TENS>(Q, i, c.j, k);
            const double rv = bl[w] - X.template apply<TENS>(rc, w);
            r[k] = rv;
            v[0] += rv * rv;
            if (jacobi) {
                const double zv = rv * row_dinv<TENS>(Q, i, c.j, rc);
                z[k] = zv;
                v[1] += rv * zv;
            }
        });
        tile_partial<2>(v, partial, c.gt);
        return;
    }
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        if (is_dirichlet(P, n.i, n.j)) {
            x[n.idx] = dirichlet_value(P, n.i, n.j);
            r[n.idx] = 0.0;
            if (jacobi) z[n.idx] = 0.0;
        } else {
            const int i = n.i, j = n.j, pt = P.pitch;
            const int64_t k = n.idx;
            // Dirichlet neighbours take their wall value (x is being set to g there)
            auto start = [&](int64_t kk) -> double { return prev2 ? 2.0 * xprev[kk] - xprev2[kk] : src[kk]; };
            auto val = [&](int ii, int jj, int64_t kk) -> double {
                return is_dirichlet_nb(P, ii, jj) ? dirichlet_value(P, ii, jj) : start(kk);
            };
            const double xc = start(k);
            const double xw = val(i, j - 1, k - 1), xe = val(i, j + 1, k + 1);
            const double xs = val(i - 1, j, k - pt), xn = val(i + 1, j, k + pt);
            const double xsw = val(i - 1, j - 1, k - pt - 1), xne = val(i + 1, j + 1, k + pt + 1);
            if (prev) x[k] = xc;
            const Row<double> rc = row_coefs<double, TENS>(P, i, j, k);
            const double rv = b[k] - row_apply<TENS>(rc, xc, xw, xe, xs, xn, xsw, xne);
            r[k] = rv;
            v[0] += rv * rv;
            if (jacobi) {
                const double zv = rv * row_dinv<TENS>(P, i, j, rc);
                z[k] = zv;
                v[1] += rv * zv;
            }
        }
    });
    tile_partial<2>(v, partial, c.gt);
}

// keep the solution as the next step's starting iterate, and the one before
// it when the start is extrapolated
template <int R0>
__global__ __launch_bounds__(kBlock) void k_keep(LevelView L, const double* __restrict__ x,
                                                 double* __restrict__ xprev, double* __restrict__ xprev2, CgState* st)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    const bool two = st[c.prob].extrap != 0;
    double xv[R0], pv[R0];
    static_rows<R0>([&](auto rr) {
        constexpr int r = decltype(rr)::value;
        if (c.i + r <= c.ny && c.j <= c.nx) {
            xv[r] = x[c.idx + (int64_t)r * c.pitch];
            if (two) pv[r] = xprev[c.idx + (int64_t)r * c.pitch];
        }
    });
    static_rows<R0>([&](auto rr) {
        constexpr int r = decltype(rr)::value;
        if (c.i + r <= c.ny && c.j <= c.nx) {
            xprev[c.idx + (int64_t)r * c.pitch] = xv[r];
            if (two) xprev2[c.idx + (int64_t)r * c.pitch] = pv[r];
        }
    });
    if (c.gt == c.P->tile_begin && threadIdx.x == 0) {
        st[c.prob].have_prev2 = st[c.prob].have_prev;
        st[c.prob].have_prev = 1;
    }
}

// p_new = z + beta p_old (formed at the seven stencil points and rounded to
// the storage type TP), q = A p_new in FP64 from the stored values, p_new.q.
// q is exactly A times the direction x will move along, so the recurrence
// residual tracks b - A x whatever TP is.  p_old and p_new are different buffers.
template <int R0, typename TP, bool TENS, typename TV = double>
__global__ __launch_bounds__(kBlock) void k_ap_dir(LevelView L, const TP* __restrict__ z, const TP* __restrict__ pold,
                                                   TP* __restrict__ pnew, TV* __restrict__ q, double* partial,
                                                   const CgState* st)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    // a register copy of the problem record: its fields are read once, not
    // again after every store (the compiler cannot tell the two apart)
    const LevelProb P = *c.P;
    const CgState& s = st[c.prob];
    if (s.done) return;
    const double beta = s.beta;
    double v[1] = {0.0};
    auto dir = [beta](TP zz, TP pp) -> TV { return (TV)(TP)((double)zz + beta * (double)pp); };
    if (c.full && !(TENS && P.aniso)) {
        // interior tile without a tensor field: every load first, then the R rows
        Window<R0, TP> Z, Q;
        Z.load(z, c.idx, P.pitch);
        Q.load(pold, c.idx, P.pitch);
        PlainRows<R0, TV, TENS> V;
        V.load(P, c.i);
        TV dw[R0 + 2], dm[R0 + 2], de[R0 + 2];
#pragma unroll
        for (int q = 0; q < R0 + 2; ++q) {
            dw[q] = dir(Z.w[q], Q.w[q]);
            dm[q] = dir(Z.m[q], Q.m[q]);
            de[q] = dir(Z.e[q], Q.e[q]);
        }
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int64_t k = c.idx + (int64_t)r * P.pitch;
            const TV nc = dm[r + 1];
            const Row<TV> ri = V.row(P, r, c.i, c.j);
            const TV qv = row_apply<TENS>(ri, nc, dw[r + 1], de[r + 1], dm[r], dm[r + 2], dw[r], de[r + 2]);
            pnew[k] = (TP)nc;
            q[k] = qv;
            v[0] += (double)nc * (double)qv;
        }
        tile_partial<1>(v, partial, c.gt);
        return;
    }
    // the other tiles: every row's loads first as well, then the rows
    Window<R0, TP> Z, Q;
    Z.load_rows(z, c.idx, c.pitch, c.i, c.ny);
    Q.load_rows(pold, c.idx, c.pitch, c.i, c.ny);
    static_rows<R0>([&](auto rr) {
        constexpr int r = decltype(rr)::value;
        if (!((c.i + r <= c.ny) && (c.j <= c.nx))) return;
        const int i = c.i + r;
        const int64_t k = c.idx + (int64_t)r * c.pitch;
        const TV nc = dir(Z.m[r + 1], Q.m[r + 1]);
        TV qv = (TV)0;
        if (!is_dirichlet(P, i, c.j)) {
            const Row<TV> rc = row_coefs<TV, TENS>(P, i, c.j, k);
            qv = row_apply<TENS>(rc, nc, dir(Z.w[r + 1], Q.w[r + 1]), dir(Z.e[r + 1], Q.e[r + 1]), dir(Z.m[r], Q.m[r]),
                                 dir(Z.m[r + 2], Q.m[r + 2]), dir(Z.w[r], Q.w[r]), dir(Z.e[r + 2], Q.e[r + 2]));
        }
        pnew[k] = (TP)nc;
        q[k] = qv;
        v[0] += (double)nc * (double)qv;
    });
    tile_partial<1>(v, partial, c.gt);
}

// x += alpha p; r -= alpha q; Jacobi: z = r/diag, r.z; flexible: r.z_old
template <int R0, typename TP, bool TENS, typename TV = double>
__global__ __launch_bounds__(kBlock) void k_update(LevelView L, TV* __restrict__ x, TV* __restrict__ r,
                                                   TP* __restrict__ z, const TP* __restrict__ p,
                                                   const TV* __restrict__ q, double* partial, const CgState* st,
                                                   int jacobi, TP* __restrict__ x1, TP* __restrict__ bcopy, TP omega)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    const LevelProb P = *c.P;
    const CgState& s0 = st[c.prob];
    if (s0.done) return;
    const TV alpha = (TV)s0.alpha;
    const int flexible = s0.flexible;
    // one accumulator for whichever r.z this mode forms (two accumulators
    // picked by the branch end up addressed through a select, in scratch)
    double rr = 0.0, rzx = 0.0;
    if (c.full && !jacobi) {
        // interior tile: every load first, then the R rows
        TV xv[R0], rv0[R0], qv[R0];
        TP pv[R0], zv[R0];
        PlainRows<R0, TP, TENS> V;
        if (x1) V.load(P, c.i);
#pragma unroll
        for (int w = 0; w < R0; ++w) {
            const int64_t k = c.idx + (int64_t)w * P.pitch;
            xv[w] = x[k];
            rv0[w] = r[k];
            qv[w] = q[k];
            pv[w] = p[k];
            zv[w] = flexible ? z[k] : (TP)0;
        }
#pragma unroll
        for (int w = 0; w < R0; ++w) {
            const int64_t k = c.idx + (int64_t)w * P.pitch;
            x[k] = xv[w] + alpha * (TV)pv[w];
            const TV rv = rv0[w] - alpha * qv[w];
            r[k] = rv;
            rr += (double)rv * (double)rv;
            rzx += flexible ? (double)rv * (double)zv[w] : 0.0;
            if (x1) {
                // the preconditioner's first sweep from zero, as k_smooth forms it
                TP wd;
                if (!(TENS && P.aniso)) {
                    wd = omega * plain_dinv<TP, TENS>(P, V.row(P, w, c.i, c.j));
                } else {
                    const Row<TP> rc = row_coefs<TP, TENS>(P, c.i + w, c.j, k);
                    wd = omega * row_dinv<TENS>(P, c.i + w, c.j, rc);
                }
                const TP bv = (TP)rv;
                x1[k] = wd * bv;
                if (bcopy) bcopy[k] = bv;
            }
        }
        double v[3] = {rr, 0.0, rzx};
        tile_partial<3>(v, partial, c.gt);
        return;
    }
    // every row's loads first, then the rows
    bool act[R0];
    TV xv[R0], rv0[R0], qv[R0];
    TP pv[R0], zv0[R0];
    static_rows<R0>([&](auto ww) {
        constexpr int w = decltype(ww)::value;
        const int64_t k = c.idx + (int64_t)w * c.pitch;
        act[w] = (c.i + w <= c.ny) && (c.j <= c.nx) && !is_dirichlet(P, c.i + w, c.j);
        if (act[w]) {
            xv[w] = x[k];
            rv0[w] = r[k];
            qv[w] = q[k];
            pv[w] = p[k];
            zv0[w] = (!jacobi && flexible) ? z[k] : (TP)0;
        }
    });
    static_rows<R0>([&](auto ww) {
        constexpr int w = decltype(ww)::value;
        if (!act[w]) return;
        const int i = c.i + w;
        const int64_t k = c.idx + (int64_t)w * c.pitch;
        x[k] = xv[w] + alpha * (TV)pv[w];
        const TV rv = rv0[w] - alpha * qv[w];
        r[k] = rv;
        rr += (double)rv * (double)rv;
        double t = 0.0;
        if (jacobi) {
            const Row<double> rc = row_coefs<double, TENS>(P, i, c.j, k);
            const TP zv = (TP)((double)rv * row_dinv<TENS>(P, i, c.j, rc));
            z[k] = zv;
            t = (double)rv * (double)zv;
        } else if (flexible) {
            t = (double)rv * (double)zv0[w];
        }
        rzx += t;
        if (x1) {
            const Row<TP> rc = row_coefs<TP, TENS>(P, i, c.j, k);
            const TP wd = omega * row_dinv<TENS>(P, i, c.j, rc);
            const TP bv = (TP)rv;
            x1[k] = wd * bv;
            if (bcopy) bcopy[k] = bv;
        }
    });
    // slot 1: r.z (Jacobi), slot 2: r_new.z_old (flexible)
    double v[3] = {rr, jacobi ? rzx : 0.0, jacobi ? 0.0 : rzx};
    tile_partial<3>(v, partial, c.gt);
}

// true residual r = b - A x (replaces the recurrence residual); final stop test.
// done_only: the problems that report convergence only (the others are still
// iterating and keep their recurrence residual).  keep (optional): x is also
// kept as the next solve's start (the confirmation that passes is the last
// read of the final x), and the problem marked as having one.
template <int R0, bool TENS>
__global__ __launch_bounds__(kBlock) void k_true_residual(LevelView L, const double* __restrict__ x,
                                                          const double* __restrict__ b, double* __restrict__ r,
                                                          double* partial, CgState* st, int done_only,
                                                          double* __restrict__ keep)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    if (done_only && !st[c.prob].done) return;
    const LevelProb P = *c.P;
    if (keep && c.gt == P.tile_begin && threadIdx.x == 0) st[c.prob].have_prev = 1;
    double v[1] = {0.0};
    if (c.full) {
        // interior tile: every load first
        Window<R0, double> X;
        X.load(x, c.idx, c.pitch);
        double bl[R0];
        static_rows<R0>([&](auto rr) {
            constexpr int w = decltype(rr)::value;
            bl[w] = b[c.idx + (int64_t)w * c.pitch];
        });
        static_rows<R0>([&](auto rr) {
            constexpr int w = decltype(rr)::value;
            const int64_t k = c.idx + (int64_t)w * c.pitch;
            const Row<double> rc = row_coefs<double, TENS>(P, c.i + w, c.j, k);
            const double rv = bl[w] - X.template apply<TENS>(rc, w);
            r[k] = rv;
            if (keep) keep[k] = X.m[w + 1];
            v[0] += rv * rv;
        });
        tile_partial<1>(v, partial, c.gt);
        return;
    }
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        if (is_dirichlet(P, n.i, n.j)) {
            if (keep) keep[n.idx] = x[n.idx];
            return;
        }
        double xc, xw, xe, xs, xn, xsw, xne;
        load7(x, n.idx, P.pitch, xc, xw, xe, xs, xn, xsw, xne);
        const Row<double> rc = row_coefs<double, TENS>(P, n.i, n.j, n.idx);
        const double rv = b[n.idx] - row_apply<TENS>(rc, xc, xw, xe, xs, xn, xsw, xne);
        r[n.idx] = rv;
        if (keep) keep[n.idx] = xc;
        v[0] += rv * rv;
    });
    tile_partial<1>(v, partial, c.gt);
}

// Jacobi preconditioner on its own (restart after a residual replacement): z = r/diag, r.z
// Iterative refinement, start of an inner solve: the FP64 residual scaled by
// 1/||r|| into FP32 (so the FP32 solve works near unit size whatever the
// residual's magnitude), the correction zeroed, the scale kept, the inner
// Krylov state reset; problems whose refinement has converged sit out.
template <int R0>
__global__ __launch_bounds__(kBlock) void k_ir_begin(LevelView L, const double* __restrict__ r64,
                                                     float* __restrict__ r32, float* __restrict__ d32,
                                                     double* __restrict__ scale, const CgState* outer, CgState* inner,
                                                     double* partial, double inner_rtol)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    const CgState& so = outer[c.prob];
    const bool first = c.gt == c.P->tile_begin && threadIdx.x == 0;
    if (so.done) {
        if (first) inner[c.prob].done = 1;
        return;
    }
    const double nrm = sqrt(so.rr);
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    double v[1] = {0.0};
    double rv[R0];
    static_rows<R0>([&](auto rr) {
        constexpr int r = decltype(rr)::value;
        if (c.i + r <= c.ny && c.j <= c.nx) rv[r] = r64[c.idx + (int64_t)r * c.pitch];
    });
    static_rows<R0>([&](auto rr) {
        constexpr int r = decltype(rr)::value;
        if (!(c.i + r <= c.ny && c.j <= c.nx)) return;
        const int64_t k = c.idx + (int64_t)r * c.pitch;
        const float f = (float)(rv[r] * inv);
        r32[k] = f;
        d32[k] = 0.0f;
        v[0] += (double)f * (double)f;
    });
    if (first) {
        scale[c.prob] = nrm;
        CgState& si = inner[c.prob];
        // reduce the residual by inner_rtol, or by no more than the outer
        // tolerance still asks (with a margin of 3): the last correction is
        // not solved past what the stop test needs
        const double need = (so.rr > 0.0) ? 0.3 * sqrt(so.tol2 * so.bb / so.rr) : 1.0;
        const double tgt = fmin(0.5, fmax(inner_rtol, need));
        si.tol2 = tgt * tgt;
        si.iter = 0;
        si.beta = 0.0;
        si.restart = 0;
        si.rzold = 0.0;
        si.done = 0;
    }
    tile_partial<1>(v, partial, c.gt);
}

// Iterative refinement, end of an inner solve: x += ||r|| d in FP64
template <int R0>
__global__ __launch_bounds__(kBlock) void k_ir_apply(LevelView L, double* __restrict__ x, const float* __restrict__ d32,
                                                     const double* __restrict__ scale, const CgState* outer)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    if (outer[c.prob].done) return;
    const double sc = scale[c.prob];
    double xv[R0];
    float dv[R0];
    static_rows<R0>([&](auto rr) {
        constexpr int r = decltype(rr)::value;
        if (c.i + r <= c.ny && c.j <= c.nx) {
            xv[r] = x[c.idx + (int64_t)r * c.pitch];
            dv[r] = d32[c.idx + (int64_t)r * c.pitch];
        }
    });
    static_rows<R0>([&](auto rr) {
        constexpr int r = decltype(rr)::value;
        if (c.i + r <= c.ny && c.j <= c.nx) x[c.idx + (int64_t)r * c.pitch] = xv[r] + sc * (double)dv[r];
    });
}

template <int R0, bool TENS>
__global__ __launch_bounds__(kBlock) void k_jacobi_z(LevelView L, const double* __restrict__ r,
                                                     double* __restrict__ z, double* partial, const CgState* st,
                                                     int restart_only)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    const LevelProb P = *c.P;
    if (st[c.prob].done || (restart_only && !st[c.prob].restart)) return;
    double v[1] = {0.0};
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        double zv = 0.0;
        if (!is_dirichlet(P, n.i, n.j)) {
            const Row<double> rc = row_coefs<double, TENS>(P, n.i, n.j, n.idx);
            zv = r[n.idx] * row_dinv<TENS>(P, n.i, n.j, rc);
        }
        z[n.idx] = zv;
        v[0] += r[n.idx] * zv;
    });
    tile_partial<1>(v, partial, c.gt);
}

// Per-problem totals of the tile partials, summed in tile order
// (deterministic), and the Krylov scalar updates that follow from them.
// One block per problem; problems the producer skipped (done) are skipped.
// slots OFF.. OFF+NV-1 of v += the problem's tile partials 0..NV-1 (ntiles
// tiles from base), in a fixed order: every total is deterministic
template <int NV, int OFF, int NT>
__device__ __forceinline__ void sum_tile_partials(const double* base, int ntiles, double* v)
{
    int q = threadIdx.x;
    for (; q + 3 * NT < ntiles; q += 4 * NT) {
        const double* p0 = base + (int64_t)q * 4;
        const double* p1 = p0 + 4 * NT;
        const double* p2 = p1 + 4 * NT;
        const double* p3 = p2 + 4 * NT;
#pragma unroll
        for (int k = 0; k < NV; ++k) v[OFF + k] += (p0[k] + p1[k]) + (p2[k] + p3[k]);
    }
    // at most three more per thread: loaded together (indices clamped to the
    // last tile), added in the same order as one at a time
    if (q < ntiles) {
        const int last = ntiles - 1;
        const double* p0 = base + (int64_t)q * 4;
        const double* p1 = base + (int64_t)min(q + NT, last) * 4;
        const double* p2 = base + (int64_t)min(q + 2 * NT, last) * 4;
        double a0[NV], a1[NV], a2[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            a0[k] = p0[k];
            a1[k] = p1[k];
            a2[k] = p2[k];
        }
#pragma unroll
        for (int k = 0; k < NV; ++k) v[OFF + k] += a0[k];
        if (q + NT < ntiles)
#pragma unroll
            for (int k = 0; k < NV; ++k) v[OFF + k] += a1[k];
        if (q + 2 * NT < ntiles)
#pragma unroll
            for (int k = 0; k < NV; ++k) v[OFF + k] += a2[k];
    }
}

// CG scalars after k_update: v[0] = r.r, v[1] = r.z (Jacobi), v[2] = r_new.z_old
__device__ __forceinline__ void fin_update(CgState& s, const double* v, int jacobi)
{
    s.iter += 1;
    s.rr = v[0];
    s.rzold = v[2];
    if (jacobi) {
        s.beta = v[1] / s.rz;
        s.rz = v[1];
    }
    s.done = (v[0] <= s.tol2 * s.bb || s.iter >= s.max_iter) ? 1 : 0;
}

// CG scalars after the preconditioner: rz = r.z, beta (Polak-Ribiere when flexible)
__device__ __forceinline__ void fin_rz(CgState& s, double rz)
{
    if (s.iter == 0 || s.restart) {
        s.beta = 0.0;
        s.restart = 0;
    } else {
        const double num = s.flexible ? (rz - s.rzold) : rz;
        s.beta = num / s.rz;
    }
    s.rz = rz;
}

__global__ __launch_bounds__(kFinishThreads) void k_finish(const LevelProb* __restrict__ probs,
                                                           const double* __restrict__ partial, CgState* st, int op,
                                                           int jacobi)
{
    const LevelProb& P = probs[blockIdx.x];
    CgState& s = st[P.prob];
    // the partials are summed before the state is tested, so their loads
    // and the state's are in flight together (a finished problem's partials
    // are read and dropped)
    double v[3] = {0.0, 0.0, 0.0};
    sum_tile_partials<3, 0, kFinishThreads>(partial + (int64_t)P.tile_begin * 4, P.ntiles, v);
    if ((op == kFinAp || op == kFinUpdate || op == kFinRz || op == kFinBAlpha || op == kFinBOmega ||
         op == kFinBUpdate || op == kFinIrInit) && s.done)
        return;
    if ((op == kFinTrueDone && !s.done) || (op == kFinRzRestart && (s.done || !s.restart))) return;
    block_sum<3, kFinishThreads>(v);
    if (threadIdx.x != 0) return;
    switch (op) {
    case kFinRhs:
        s.bb = v[0];
        break;
    case kFinInit:
        s.rr = v[0];
        s.rz = v[1];
        s.iter = 0;
        s.beta = 0.0;
        s.restart = 0;
        s.done = (v[0] <= s.tol2 * s.bb) ? 1 : 0;
        break;
    case kFinAp:
        s.pq = v[0];
        if (v[0] > 0.0 && s.rz > 0.0) {
            s.alpha = s.rz / v[0];
        } else {
            s.alpha = 0.0;
            s.iter += 1; // counted, so repeated breakdowns end at max_iter
            s.done = 1;  // exact solution reached (or breakdown): stop this problem
        }
        break;
    case kFinUpdate:
        fin_update(s, v, jacobi);
        break;
    case kFinRz:
    case kFinRzRestart:
        fin_rz(s, v[0]);
        break;
    case kFinTrue:
    case kFinTrueDone:
        s.rr_true = v[0];
        s.rr = v[0];
        s.done = (v[0] <= s.tol2 * s.bb || s.iter >= s.max_iter) ? 1 : 0;
        s.restart = s.done ? 0 : 1;
        break;
    case kFinIrInit: // inner FP32 solve from the scaled residual: its norm is the reference
        s.bb = v[0];
        s.rr = v[0];
        s.done = (v[0] > 0.0) ? 0 : 1;
        break;
    case kFinIrOuter: // the FP64 residual after a correction: one refinement step done
        s.rr_true = v[0];
        s.rr = v[0];
        s.iter += 1;
        s.done = (v[0] <= s.tol2 * s.bb || s.iter >= s.max_iter) ? 1 : 0;
        break;
    case kFinBRestart: // after kFinTrue: restarted problems begin again from r^ = r
        if (s.restart) {
            s.rho = s.rr;
            s.beta = 0.0;
            s.alpha = 1.0;
            s.omega = 1.0;
            s.restart = 0;
        }
        break;
    // BiCGStab (right preconditioned, van der Vorst 1992): r^ = r at the
    // start and after every restart, so (r^, r) = (r, r)
    case kFinBInit:
        s.rr = v[0];
        s.rho = v[0];
        s.iter = 0;
        s.alpha = 1.0;
        s.omega = 1.0;
        s.beta = 0.0;
        s.restart = 0;
        s.done = (v[0] <= s.tol2 * s.bb) ? 1 : 0;
        break;
    case kFinBAlpha: // v[0] = (r^, v)
        if (v[0] != 0.0 && s.rho != 0.0) {
            s.alpha = s.rho / v[0];
        } else {
            s.alpha = 0.0;
            s.iter += 1; // counted, so repeated breakdowns end at max_iter
            s.done = 1;  // breakdown: the true-residual check restarts it
        }
        break;
    case kFinBOmega: // v[0] = (t, s), v[1] = (t, t)
        s.ts = v[0];
        s.tt = v[1];
        s.omega = (v[1] > 0.0) ? v[0] / v[1] : 0.0;
        break;
    case kFinBUpdate: { // v[0] = (r^, r_new), v[1] = (r_new, r_new)
        s.iter += 1;
        s.rr = v[1];
        const double rho_new = v[0];
        if (s.omega != 0.0 && s.rho != 0.0) {
            s.beta = (rho_new / s.rho) * (s.alpha / s.omega);
        } else {
            s.beta = 0.0;
            s.done = 1; // stagnation: the true-residual check restarts it
        }
        s.rho = rho_new;
        if (v[1] <= s.tol2 * s.bb || s.iter >= s.max_iter) s.done = 1;
        break;
    }
    }
}

// kFinUpdate for the last k_update (its partials in the tile slots of the
// level-0 table upd) and then, for the problems still iterating, kFinRz for
// the V-cycle that followed it (r.z in the slots of table rz, which must not
// overlap the update's): one launch where the two take one each.  The
// V-cycle ran before the update's stop test; a problem that stopped there
// drops its r.z, as kFinRz does for a finished problem.
__global__ __launch_bounds__(kFinishThreads) void k_finish_update_rz(const LevelProb* __restrict__ rz,
                                                                     const LevelProb* __restrict__ upd,
                                                                     const double* __restrict__ partial, CgState* st)
{
    const LevelProb& P = upd[blockIdx.x];
    const LevelProb& Q = rz[blockIdx.x];
    CgState& s = st[P.prob];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    sum_tile_partials<3, 0, kFinishThreads>(partial + (int64_t)P.tile_begin * 4, P.ntiles, v);
    sum_tile_partials<1, 3, kFinishThreads>(partial + (int64_t)Q.tile_begin * 4, Q.ntiles, v);
    if (s.done) return;
    block_sum<4, kFinishThreads>(v);
    if (threadIdx.x != 0) return;
    fin_update(s, v, 0);
    if (!s.done) fin_rz(s, v[3]);
}

// ------------------------------------------------------------ BiCGStab kernels (FP64 vectors)

// r^ = r (start and restarts)
template <int R0>
__global__ __launch_bounds__(kBlock) void k_copy(LevelView L, const double* __restrict__ src, double* __restrict__ dst,
                                                 const CgState* st)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (n.in) dst[n.idx] = src[n.idx];
    });
}

// p = r + beta (p - omega v)   (beta = 0: p = r)
template <int R0>
__global__ __launch_bounds__(kBlock) void k_bicg_p(LevelView L, const double* __restrict__ r, double* __restrict__ p,
                                                   const double* __restrict__ v, const CgState* st)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    const CgState& s0 = st[c.prob];
    if (s0.done) return;
    const double beta = s0.beta, omega = s0.omega;
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        p[n.idx] = (beta == 0.0) ? r[n.idx] : r[n.idx] + beta * (p[n.idx] - omega * v[n.idx]);
    });
}

// y = A xh on the free rows (0 on Dirichlet rows), xh in the preconditioner's
// precision; partials: mode 0: (a, y); mode 1: (y, a) and (y, y)
template <int R0, typename T, bool TENS>
__global__ __launch_bounds__(kBlock) void k_bicg_apply(LevelView L, const T* __restrict__ xh, double* __restrict__ y,
                                                       const double* __restrict__ a, int mode, double* partial,
                                                       const CgState* st)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    const LevelProb P = *c.P;
    if (st[c.prob].done) return;
    double d0 = 0.0, d1 = 0.0;
    if (c.full && !(TENS && P.aniso)) {
        // interior tile without a tensor field: every load first
        Window<R0, T> X;
        X.load(xh, c.idx, P.pitch);
        double av[R0];
#pragma unroll
        for (int r = 0; r < R0; ++r) av[r] = a[c.idx + (int64_t)r * P.pitch];
        PlainRows<R0, double, TENS> V;
        V.load(P, c.i);
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const Row<double> rc = V.row(P, r, c.i, c.j);
            const double yv = row_apply<TENS>(rc, (double)X.m[r + 1], (double)X.w[r + 1], (double)X.e[r + 1],
                                              (double)X.m[r], (double)X.m[r + 2], (double)X.w[r], (double)X.e[r + 2]);
            y[c.idx + (int64_t)r * P.pitch] = yv;
            d0 += yv * av[r];
            if (mode) d1 += yv * yv;
        }
        double v[2] = {d0, d1};
        tile_partial<2>(v, partial, c.gt);
        return;
    }
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        double yv = 0.0;
        if (!is_dirichlet(P, n.i, n.j)) {
            T xc, xw, xe, xs, xn, xsw, xne;
            load7(xh, n.idx, P.pitch, xc, xw, xe, xs, xn, xsw, xne);
            const Row<double> rc = row_coefs<double, TENS>(P, n.i, n.j, n.idx);
            yv = row_apply<TENS>(rc, (double)xc, (double)xw, (double)xe, (double)xs, (double)xn, (double)xsw,
                                 (double)xne);
        }
        y[n.idx] = yv;
        d0 += yv * a[n.idx];
        if (mode) d1 += yv * yv;
    });
    double v[2] = {d0, d1};
    tile_partial<2>(v, partial, c.gt);
}

// s = r - alpha v
template <int R0>
__global__ __launch_bounds__(kBlock) void k_bicg_s(LevelView L, const double* __restrict__ r,
                                                   const double* __restrict__ v, double* __restrict__ sv,
                                                   const CgState* st)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    const CgState& s0 = st[c.prob];
    if (s0.done) return;
    const double alpha = s0.alpha;
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (n.in) sv[n.idx] = r[n.idx] - alpha * v[n.idx];
    });
}

// x += alpha p^ + omega s^;  r = s - omega t;  partials (r^, r), (r, r)
template <int R0, typename T>
__global__ __launch_bounds__(kBlock) void k_bicg_x(LevelView L, double* __restrict__ x, double* __restrict__ r,
                                                   const T* __restrict__ ph, const T* __restrict__ sh,
                                                   const double* __restrict__ sv, const double* __restrict__ t,
                                                   const double* __restrict__ rhat, double* partial, const CgState* st)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    const CgState& s0 = st[c.prob];
    if (s0.done) return;
    const double alpha = s0.alpha, omega = s0.omega;
    double d0 = 0.0, d1 = 0.0;
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        const int64_t k = n.idx;
        x[k] += alpha * (double)ph[k] + omega * (double)sh[k];
        const double rv = sv[k] - omega * t[k];
        r[k] = rv;
        d0 += rhat[k] * rv;
        d1 += rv * rv;
    });
    double v[2] = {d0, d1};
    tile_partial<2>(v, partial, c.gt);
}

// ------------------------------------------------------------ multigrid node operations
// Each takes the array base, the element index k of node (i,j) and the row
// pitch, so the same code runs on the pitched HBM pools and on the compact
// LDS copies of the bottom levels.

// damped Jacobi value at a free node: x + omega (b - A x) / diag  (x = 0 when zero_init)
// (k indexes b and x with row pitch `pitch`; kt is the node's element in the
// level pool, where the tensor arrays live)
template <typename T, typename TB, bool TENS>
__device__ __forceinline__ T smooth_val(const LevelProb& P, int i, int j, int64_t k, int pitch,
                                        const TB* __restrict__ b, const T* __restrict__ xin, T omega,
                                        bool zero_init, T& bv, int64_t kt)
{
    const Row<T> rc = row_coefs<T, TENS>(P, i, j, kt);
    const T wd = omega * row_dinv<TENS>(P, i, j, rc);
    bv = (T)b[k];
    if (zero_init) return wd * bv;
    T xc, xw, xe, xs, xn, xsw, xne;
    load7(xin, k, pitch, xc, xw, xe, xs, xn, xsw, xne);
    return xc + wd * (bv - row_apply<TENS>(rc, xc, xw, xe, xs, xn, xsw, xne));
}

template <typename T, bool TENS>
__device__ __forceinline__ T residual_val(const LevelProb& P, int i, int j, int64_t k, int pitch,
                                          const T* __restrict__ b, const T* __restrict__ x, int64_t kt)
{
    T xc, xw, xe, xs, xn, xsw, xne;
    load7(x, k, pitch, xc, xw, xe, xs, xn, xsw, xne);
    const Row<T> rc = row_coefs<T, TENS>(P, i, j, kt);
    return b[k] - row_apply<TENS>(rc, xc, xw, xe, xs, xn, xsw, xne);
}

// b_c(I,J) = sum_f phi_IJ(f) r_f over the fine nodes within one coarse cell of
// (I,J); rf(i, j, k) gives r at fine node (i,j), k = i*fpitch + j, and must
// give 0 one node outside the grid (the nested form reads there).  FETCH:
// rf is a plain load (fetch the whole support first); otherwise rf computes
// (visit only the support's nodes).
template <typename T, bool FETCH, typename RF>
__device__ __forceinline__ T restrict_gen(const LevelProb& Pc, const LevelProb& Pf, int I, int J, int fpitch, RF&& rf)
{
    const int nxf = Pf.nx, nyf = Pf.ny, nxc = Pc.nx, nyc = Pc.ny;
    if (nxf == 2 * nxc && nyf == 2 * nyc) {
        // nested: weight 1 at the coincident node, 1/2 at its six stencil
        // neighbours, 0 at NW/SE
        const int i0 = 2 * I, j0 = 2 * J;
        const int64_t k0 = (int64_t)i0 * fpitch + j0;
        return rf(i0, j0, k0) + (T)0.5 * (((rf(i0, j0 - 1, k0 - 1) + rf(i0, j0 + 1, k0 + 1)) +
                                           (rf(i0 - 1, j0, k0 - fpitch) + rf(i0 + 1, j0, k0 + fpitch))) +
                                          (rf(i0 - 1, j0 - 1, k0 - fpitch - 1) + rf(i0 + 1, j0 + 1, k0 + fpitch + 1)));
    }
    // fine nodes strictly inside the hat's support: (J-1) nxf/nxc < j < (J+1) nxf/nxc
    const int jlo = J == 0 ? 0 : div_floor_pos((J - 1) * nxf, nxc, Pc.rnxf) + 1;
    const int jhi = min(nxf, div_floor_pos((J + 1) * nxf - 1, nxc, Pc.rnxf));
    const int ilo = I == 0 ? 0 : div_floor_pos((I - 1) * nyf, nyc, Pc.rnyf) + 1;
    const int ihi = min(nyf, div_floor_pos((I + 1) * nyf - 1, nyc, Pc.rnyf));
    const T inx = recip_nx<T>(Pf), iny = recip_ny<T>(Pf);
    if constexpr (!FETCH) {
        T sum = (T)0;
        for (int i = ilo; i <= ihi; ++i) {
            const int dv = i * nyc - I * nyf;
            for (int j = jlo; j <= jhi; ++j)
                sum += hat<T>(j * nxc - J * nxf, dv, inx, iny) * rf(i, j, (int64_t)i * fpitch + j);
        }
        return sum;
    }
    // n_f <= 2 n_c on every axis (the hierarchy never coarsens by more), so
    // the support holds at most kSup fine nodes per axis: all values are
    // fetched first, then summed row by row in the order of the loops above
    constexpr int kSup = 4;
    T v[kSup][kSup];
#pragma unroll
    for (int a = 0; a < kSup; ++a)
#pragma unroll
        for (int q = 0; q < kSup; ++q) {
            const int i = ilo + a, j = jlo + q;
            v[a][q] = (i <= ihi && j <= jhi) ? rf(i, j, (int64_t)i * fpitch + j) : (T)0;
        }
    T sum = (T)0;
#pragma unroll
    for (int a = 0; a < kSup; ++a) {
        const int i = ilo + a, dv = i * nyc - I * nyf;
#pragma unroll
        for (int q = 0; q < kSup; ++q) {
            const int j = jlo + q;
            if (i <= ihi && j <= jhi) sum += hat<T>(j * nxc - J * nxf, dv, inx, iny) * v[a][q];
        }
    }
    return sum;
}

// the same from a stored residual: fine node (i,j) at rf[forg + i*fpitch + j],
// zero frame around the grid.  FETCH for residuals in memory (the loads'
// latency is what counts); the bottom solve's LDS copies take the loop.
template <typename T, bool FETCH = true>
__device__ __forceinline__ T restrict_val(const LevelProb& Pc, const LevelProb& Pf, int I, int J,
                                          const T* __restrict__ rf, int64_t forg, int fpitch)
{
    return restrict_gen<T, FETCH>(Pc, Pf, I, J, fpitch, [&](int, int, int64_t k) { return rf[forg + k]; });
}

// prolong_val's nested case with the four corners of the coarse cell loaded
// unconditionally (indices clamped into the coarse grid) and the one needed
// selected: the same value, with no load under a branch
template <typename T>
__device__ __forceinline__ T prolong_nested(int i, int j, int nxc, int nyc, const T* __restrict__ xc, int64_t corg,
                                            int cpitch)
{
    const int I = i >> 1, J = j >> 1, a = i & 1, bb = j & 1;
    const int I1 = min(I + 1, nyc), J1 = min(J + 1, nxc);
    const T* r0 = xc + corg + (int64_t)I * cpitch;
    const T* r1 = xc + corg + (int64_t)I1 * cpitch;
    const T x00 = r0[J], x01 = r0[J1], x10 = r1[J], x11 = r1[J1];
    const T xo = a ? (bb ? x11 : x10) : (bb ? x01 : x00);
    return (a | bb) ? (T)0.5 * (x00 + xo) : x00;
}

// (P x_c) at fine node (i,j): the four corners of the coarse cell holding it;
// coarse node (I,J) lives at xc[corg + I*cpitch + J]
template <typename T>
__device__ __forceinline__ T prolong_val(const LevelProb& Pf, const LevelProb& Pc, int i, int j,
                                         const T* __restrict__ xc, int64_t corg, int cpitch)
{
    const int nxf = Pf.nx, nyf = Pf.ny, nxc = Pc.nx, nyc = Pc.ny;
    if (nxf == 2 * nxc && nyf == 2 * nyc) {
        // nested: coincident node, or the mean of the two ends of the coarse
        // edge (horizontal, vertical or the SW-NE diagonal) the node halves
        const int a = i & 1, bb = j & 1;
        const T* c0 = xc + corg + (int64_t)(i >> 1) * cpitch + (j >> 1);
        const T x00 = c0[0];
        const T xo = a ? (bb ? c0[cpitch + 1] : c0[cpitch]) : (bb ? c0[1] : x00);
        return (a | bb) ? (T)0.5 * (x00 + xo) : x00;
    }
    const int J0 = min(nxc - 1, div_floor_pos(j * nxc, nxf, Pf.rnxf));
    const int I0 = min(nyc - 1, div_floor_pos(i * nyc, nyf, Pf.rnyf));
    const T inx = recip_nx<T>(Pf), iny = recip_ny<T>(Pf);
    const int du = j * nxc - J0 * nxf, dv = i * nyc - I0 * nyf;
    const int64_t k = corg + (int64_t)I0 * cpitch + J0;
    return hat<T>(du, dv, inx, iny) * xc[k] + hat<T>(du - nxf, dv, inx, iny) * xc[k + 1] +
           hat<T>(du, dv - nyf, inx, iny) * xc[k + cpitch] +
           hat<T>(du - nxf, dv - nyf, inx, iny) * xc[k + cpitch + 1];
}

// ------------------------------------------------------------ multigrid grid kernels

// Damped Jacobi sweep over the launched tiles.  bcopy (optional) keeps b in
// the smoother precision; dot (last level-0 sweep) adds the partial r.z with
// b the FP64 residual and x_out the preconditioned residual.  zero_init 2:
// from x = 0 for the problems a confirmation has just restarted only.
template <int R, typename T, typename TB, typename TO, bool TENS>
__global__ __launch_bounds__(kBlock) void k_smooth(LevelView L, const TB* __restrict__ b, const T* __restrict__ xin,
                                                   TO* __restrict__ xout, T* __restrict__ bcopy, T omega,
                                                   int zero_init, int dot, double* partial, const CgState* st)
{
    NodeCtx c;
    if (!decode<R>(L, c)) return;
    const LevelProb P = *c.P;
    if (st[c.prob].done || (zero_init == 2 && !st[c.prob].restart)) return;
    double v[1] = {0.0};
    if (R > 1 && c.full && !(TENS && P.aniso)) {
        // interior tile without a tensor field: the same arithmetic as
        // smooth_val with no per-node tests and one load per row and column;
        // every load before the first store, so no row waits on the one before
        TB bl[R];
#pragma unroll
        for (int r = 0; r < R; ++r) bl[r] = b[c.idx + (int64_t)r * P.pitch];
        Window<R, T> X;
        if (!zero_init) X.load(xin, c.idx, P.pitch);
        PlainRows<R, T, TENS> V;
        V.load(P, c.i);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t k = c.idx + (int64_t)r * P.pitch;
            const Row<T> ri = V.row(P, r, c.i, c.j);
            const T wd = omega * plain_dinv<T, TENS>(P, ri);
            const T bv = (T)bl[r];
            const T out = zero_init ? wd * bv : X.m[r + 1] + wd * (bv - X.template apply<TENS>(ri, r));
            if (bcopy) bcopy[k] = bv;
            const TO o = (TO)out;
            xout[k] = o;
            if (dot) v[0] += (double)bl[r] * (double)o;
        }
        if (dot) tile_partial<1>(v, partial, c.gt);
        return;
    }
    if (R > 1) {
        // the other tiles: every row's loads first as well, then the rows
        // (smooth_val's arithmetic)
        bool act[R], dirn[R];
        TB bl[R];
        Window<R, T> X;
        if (!zero_init) X.load_rows(xin, c.idx, c.pitch, c.i, c.ny);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = c.i + r;
            act[r] = (i <= c.ny) && (c.j <= c.nx);
            dirn[r] = act[r] && is_dirichlet(P, i, c.j);
            bl[r] = (act[r] && !dirn[r]) ? b[c.idx + (int64_t)r * c.pitch] : (TB)0;
        }
        static_rows<R>([&](auto rr) {
            constexpr int r = decltype(rr)::value;
            if (!act[r]) return;
            const int i = c.i + r;
            const int64_t k = c.idx + (int64_t)r * c.pitch;
            if (dirn[r]) {
                xout[k] = (TO)0;
                if (bcopy) bcopy[k] = (T)0;
                return;
            }
            const Row<T> rc = row_coefs<T, TENS>(P, i, c.j, k);
            const T wd = omega * row_dinv<TENS>(P, i, c.j, rc);
            const T bv = (T)bl[r];
            const T out = zero_init ? wd * bv : X.m[r + 1] + wd * (bv - X.template apply<TENS>(rc, r));
            if (bcopy) bcopy[k] = bv;
            const TO o = (TO)out;
            xout[k] = o;
            if (dot) v[0] += (double)bl[r] * (double)o;
        });
        if (dot) tile_partial<1>(v, partial, c.gt);
        return;
    }
    each_row<R>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        if (is_dirichlet(P, n.i, n.j)) {
            xout[n.idx] = (TO)0;
            if (bcopy) bcopy[n.idx] = (T)0;
        } else {
            T bv;
            const T out = smooth_val<T, TB, TENS>(P, n.i, n.j, n.idx, P.pitch, b, xin, omega, zero_init != 0, bv, n.idx);
            if (bcopy) bcopy[n.idx] = bv;
            const TO o = (TO)out;
            xout[n.idx] = o;
            if (dot) v[0] += (double)b[n.idx] * (double)o;
        }
    });
    if (dot) tile_partial<1>(v, partial, c.gt);
}

template <int R, typename T, bool TENS>
__global__ __launch_bounds__(kBlock) void k_residual(LevelView L, const T* __restrict__ b, const T* __restrict__ x,
                                                     T* __restrict__ r, const CgState* st)
{
    NodeCtx c;
    if (!decode<R>(L, c)) return;
    const LevelProb P = *c.P;
    if (st[c.prob].done) return;
    if (R > 1 && c.full && !(TENS && P.aniso)) {
        T bl[R];
#pragma unroll
        for (int q = 0; q < R; ++q) bl[q] = b[c.idx + (int64_t)q * P.pitch];
        Window<R, T> X;
        X.load(x, c.idx, P.pitch);
        PlainRows<R, T, TENS> V;
        V.load(P, c.i);
#pragma unroll
        for (int q = 0; q < R; ++q) r[c.idx + (int64_t)q * P.pitch] = bl[q] - X.template apply<TENS>(V.row(P, q, c.i, c.j), q);
        return;
    }
    if (R > 1) {
        // the other tiles: every row's loads first as well (residual_val's arithmetic)
        bool act[R], dirn[R];
        T bl[R];
        Window<R, T> X;
        X.load_rows(x, c.idx, c.pitch, c.i, c.ny);
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int i = c.i + q;
            act[q] = (i <= c.ny) && (c.j <= c.nx);
            dirn[q] = act[q] && is_dirichlet(P, i, c.j);
            bl[q] = (act[q] && !dirn[q]) ? b[c.idx + (int64_t)q * c.pitch] : (T)0;
        }
        static_rows<R>([&](auto qq) {
            constexpr int q = decltype(qq)::value;
            if (!act[q]) return;
            const int64_t k = c.idx + (int64_t)q * c.pitch;
            if (dirn[q]) {
                r[k] = (T)0;
                return;
            }
            const Row<T> rc = row_coefs<T, TENS>(P, c.i + q, c.j, k);
            r[k] = bl[q] - X.template apply<TENS>(rc, q);
        });
        return;
    }
    each_row<R>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        r[n.idx] = is_dirichlet(P, n.i, n.j) ? (T)0 : residual_val<T, TENS>(P, n.i, n.j, n.idx, P.pitch, b, x, n.idx);
    });
}

// RC: rows per thread of the coarse level's tiles; all RC values are gathered
// before the first store
template <int RC, typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(LevelView Lc, const LevelProb* __restrict__ fine,
                                                     const T* __restrict__ rf, T* __restrict__ bc, const CgState* st)
{
    NodeCtx c;
    if (!decode<RC>(Lc, c)) return;
    const LevelProb Pc = *c.P;
    if (st[Pc.prob].done) return;
    const LevelProb Pf = fine[Pc.fidx];
    bool act[RC];
    T val[RC];
    static_rows<RC>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        const int i = c.i + q;
        act[q] = (i <= c.ny) && (c.j <= c.nx);
        val[q] = (act[q] && !is_dirichlet(Pc, i, c.j)) ? restrict_val<T>(Pc, Pf, i, c.j, rf, Pf.origin, Pf.pitch) : (T)0;
    });
    static_rows<RC>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        if (act[q]) bc[c.idx + (int64_t)q * c.pitch] = val[q];
    });
}

template <int R, typename T>
__global__ __launch_bounds__(kBlock) void k_prolong(LevelView Lf, const LevelProb* __restrict__ coarse,
                                                    const T* __restrict__ xc, T* __restrict__ xf, const CgState* st)
{
    NodeCtx c;
    if (!decode<R>(Lf, c)) return;
    const LevelProb Pf = *c.P;
    if (st[Pf.prob].done) return;
    const LevelProb Pc = coarse[Pf.cidx];
    if (R > 1 && c.full) {
        // interior tile: the coarse values and the fine x first, then the R
        // rows (the transfer's kind picked once, so no load sits under a branch)
        T add[R], xv[R];
        auto rows = [&](auto&& pval) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                add[r] = pval(c.i + r);
                xv[r] = xf[c.idx + (int64_t)r * Pf.pitch];
            }
        };
        if (Pf.nx == 2 * Pc.nx && Pf.ny == 2 * Pc.ny)
            rows([&](int i) { return prolong_nested<T>(i, c.j, Pc.nx, Pc.ny, xc, Pc.origin, Pc.pitch); });
        else
            rows([&](int i) { return prolong_val<T>(Pf, Pc, i, c.j, xc, Pc.origin, Pc.pitch); });
#pragma unroll
        for (int r = 0; r < R; ++r) xf[c.idx + (int64_t)r * Pf.pitch] = xv[r] + add[r];
        return;
    }
    each_row<R>(c, [&](const NodeCtx& n) {
        if (!n.in || is_dirichlet(Pf, n.i, n.j)) return;
        xf[n.idx] += prolong_val<T>(Pf, Pc, n.i, n.j, xc, Pc.origin, Pc.pitch);
    });
}

// ------------------------------------------------------------ fused multigrid kernels
// Each gives bit for bit what the two launches it replaces give: the values
// the first launch would have stored are recomputed, in the same precision
// and order, at the stencil points the second one reads.

// x after the first sweep from x = 0: omega b / diag (0 on Dirichlet nodes
// and outside the grid)
template <typename T, typename TB, bool TENS>
__device__ __forceinline__ T first_sweep_at(const LevelProb& P, int i, int j, int64_t k, const TB* __restrict__ b,
                                            T omega)
{
    if (i < 0 || j < 0 || i > P.ny || j > P.nx || is_dirichlet(P, i, j)) return (T)0;
    const Row<T> rc = row_coefs<T, TENS>(P, i, j, k);
    const T wd = omega * row_dinv<TENS>(P, i, j, rc);
    return wd * (T)b[k];
}

// pre-smoothing sweeps 1 and 2 (sweep 1 from x = 0); bcopy as in k_smooth
template <int R, typename T, typename TB, bool TENS>
__global__ __launch_bounds__(kBlock) void k_presmooth2(LevelView L, const TB* __restrict__ b, T* __restrict__ xout,
                                                       T* __restrict__ bcopy, T omega, T omega2, const CgState* st)
{
    NodeCtx c;
    if (!decode<R>(L, c)) return;
    const LevelProb P = *c.P;
    if (st[c.prob].done) return;
    if (R > 1 && c.full2 && !(TENS && P.aniso)) {
        // the tile and its ring are interior, no tensor field: b over the
        // R+2 rows of the three columns, the first sweep's x there (its row's
        // damped inverse diagonal times b), then the second sweep -- the
        // values first_sweep_at and the loop below give, with every load first
        Window<R, TB> B;
        B.load(b, c.idx, P.pitch);
        T vv[R + 4];
        if (TENS && P.vel) {
            const T* vb = vel_base<T>(P);
#pragma unroll
            for (int q = 0; q < R + 4; ++q) vv[q] = vb[c.i - 2 + q];
        }
        auto row_q = [&](int q) -> Row<T> { // grid row c.i - 1 + q
            Row<T> rc = interior_row<T>(P);
            if (TENS && P.vel) add_convection_v<T>(P, c.i - 1 + q, c.j, vv[q], vv[q + 1], vv[q + 2], rc);
            return rc;
        };
        T x1w[R + 2], x1m[R + 2], x1e[R + 2];
#pragma unroll
        for (int q = 0; q < R + 2; ++q) {
            const T wq = omega * plain_dinv<T, TENS>(P, row_q(q));
            x1w[q] = wq * (T)B.w[q];
            x1m[q] = wq * (T)B.m[q];
            x1e[q] = wq * (T)B.e[q];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t k = c.idx + (int64_t)r * P.pitch;
            const Row<T> rc = row_q(r + 1);
            const T wd = omega2 * plain_dinv<T, TENS>(P, rc);
            const T bv = (T)B.m[r + 1];
            xout[k] = x1m[r + 1] + wd * (bv - row_apply<TENS>(rc, x1m[r + 1], x1w[r + 1], x1e[r + 1], x1m[r], x1m[r + 2],
                                                               x1w[r], x1e[r + 2]));
            if (bcopy) bcopy[k] = bv;
        }
        return;
    }
    each_row<R>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        const int i = n.i, j = n.j, pt = P.pitch;
        const int64_t k = n.idx;
        if (is_dirichlet(P, i, j)) {
            xout[k] = (T)0;
            if (bcopy) bcopy[k] = (T)0;
            return;
        }
        const T xc = first_sweep_at<T, TB, TENS>(P, i, j, k, b, omega);
        const T xw = first_sweep_at<T, TB, TENS>(P, i, j - 1, k - 1, b, omega);
        const T xe = first_sweep_at<T, TB, TENS>(P, i, j + 1, k + 1, b, omega);
        const T xs = first_sweep_at<T, TB, TENS>(P, i - 1, j, k - pt, b, omega);
        const T xn = first_sweep_at<T, TB, TENS>(P, i + 1, j, k + pt, b, omega);
        const T xsw = first_sweep_at<T, TB, TENS>(P, i - 1, j - 1, k - pt - 1, b, omega);
        const T xne = first_sweep_at<T, TB, TENS>(P, i + 1, j + 1, k + pt + 1, b, omega);
        const Row<T> rc = row_coefs<T, TENS>(P, i, j, k);
        const T wd = omega2 * row_dinv<TENS>(P, i, j, rc);
        const T bv = (T)b[k];
        xout[k] = xc + wd * (bv - row_apply<TENS>(rc, xc, xw, xe, xs, xn, xsw, xne));
        if (bcopy) bcopy[k] = bv;
    });
}

// residual of the fine level and its restriction, RC coarse nodes per thread
template <int RC, typename T, bool TENS>
__global__ __launch_bounds__(kBlock) void k_residual_restrict(LevelView Lc, const LevelProb* __restrict__ fine,
                                                              const T* __restrict__ bf, const T* __restrict__ xf,
                                                              T* __restrict__ bc, const CgState* st)
{
    NodeCtx c;
    if (!decode<RC>(Lc, c)) return;
    const LevelProb Pc = *c.P;
    if (st[Pc.prob].done) return;
    const LevelProb Pf = fine[Pc.fidx];
    auto racc = [&](int i, int j, int64_t k) -> T {
        if (i < 0 || j < 0 || i > Pf.ny || j > Pf.nx || is_dirichlet(Pf, i, j)) return (T)0;
        return residual_val<T, TENS>(Pf, i, j, Pf.origin + k, Pf.pitch, bf, xf, Pf.origin + k);
    };
    bool act[RC];
    T val[RC];
    static_rows<RC>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        const int i = c.i + q;
        act[q] = (i <= c.ny) && (c.j <= c.nx);
        val[q] = (act[q] && !is_dirichlet(Pc, i, c.j)) ? restrict_gen<T, false>(Pc, Pf, i, c.j, Pf.pitch, racc) : (T)0;
    });
    static_rows<RC>([&](auto qq) {
        constexpr int q = decltype(qq)::value;
        if (act[q]) bc[c.idx + (int64_t)q * c.pitch] = val[q];
    });
}

// coarse-grid correction x += P x_c followed by the first post-smoothing
// sweep (xin holds x, xout receives the sweep; dot as in k_smooth)
template <int R, typename T, typename TB, typename TO, bool TENS>
__global__ __launch_bounds__(kBlock) void k_prolong_smooth(LevelView Lf, const LevelProb* __restrict__ coarse,
                                                           const T* __restrict__ xc, const TB* __restrict__ b,
                                                           const T* __restrict__ xin, TO* __restrict__ xout, T omega,
                                                           int dot, double* partial, const CgState* st)
{
    NodeCtx c;
    if (!decode<R>(Lf, c)) return;
    const LevelProb P = *c.P;
    if (st[c.prob].done) return;
    const LevelProb Pc = coarse[P.cidx];
    auto y_at = [&](int i, int j, int64_t k) -> T {
        if (i < 0 || j < 0 || i > P.ny || j > P.nx || is_dirichlet(P, i, j)) return xin[k];
        return xin[k] + prolong_val<T>(P, Pc, i, j, xc, Pc.origin, Pc.pitch);
    };
    double v[1] = {0.0};
    each_row<R>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        const int i = n.i, j = n.j, pt = P.pitch;
        const int64_t k = n.idx;
        if (is_dirichlet(P, i, j)) {
            xout[k] = (TO)0;
            return;
        }
        const T yc = y_at(i, j, k), yw = y_at(i, j - 1, k - 1), ye = y_at(i, j + 1, k + 1);
        const T ys = y_at(i - 1, j, k - pt), yn = y_at(i + 1, j, k + pt);
        const T ysw = y_at(i - 1, j - 1, k - pt - 1), yne = y_at(i + 1, j + 1, k + pt + 1);
        const Row<T> rc = row_coefs<T, TENS>(P, i, j, k);
        const T wd = omega * row_dinv<TENS>(P, i, j, rc);
        const T bv = (T)b[k];
        const TO o = (TO)(yc + wd * (bv - row_apply<TENS>(rc, yc, yw, ye, ys, yn, ysw, yne)));
        xout[k] = o;
        if (dot) v[0] += (double)b[k] * (double)o;
    });
    if (dot) tile_partial<1>(v, partial, c.gt);
}

// level-0 problems whose whole V-cycle runs in one workgroup: cast the FP64
// residual in, and the result out (with the partial r.z)
template <int R0, typename T, typename TI = double>
__global__ __launch_bounds__(kBlock) void k_cast_in(LevelView L, const TI* __restrict__ r, T* __restrict__ b,
                                                    const CgState* st)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    if (st[c.prob].done) return;
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (n.in) b[n.idx] = (T)r[n.idx];
    });
}

template <int R0, typename T, typename TZ, typename TI = double>
__global__ __launch_bounds__(kBlock) void k_cast_out(LevelView L, const T* __restrict__ x, const TI* __restrict__ r,
                                                     TZ* __restrict__ z, double* partial, const CgState* st)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    if (st[c.prob].done) return;
    double v[1] = {0.0};
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        const TZ zv = (TZ)x[n.idx];
        z[n.idx] = zv;
        v[0] += (double)r[n.idx] * (double)zv;
    });
    tile_partial<1>(v, partial, c.gt);
}

// ------------------------------------------------------------ bottom V-cycle (one workgroup per problem, in LDS)

// byte offset of a bottom chain's level records in the dynamic LDS: after
// the level arrays, 16-byte aligned
template <typename T>
__host__ __device__ inline int64_t bottom_rec_offset(const LevelProb& P0)
{
    return ((int64_t)P0.lds_total * (int64_t)sizeof(T) + 15) / 16 * 16;
}

constexpr int kRecWords = (int)(sizeof(LevelProb) / 4);

// A level record read from LDS into uniform (scalar) registers, word by word:
// a plain copy would hold every field in vector registers.  Words of fields
// the caller never reads are dropped by the compiler.
__device__ __forceinline__ LevelProb uniform_rec(const LevelProb* p)
{
    LevelProb r;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(p);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
    for (int w = 0; w < kRecWords; ++w) dst[w] = __builtin_amdgcn_readfirstlane(src[w]);
    return r;
}
static_assert(sizeof(LevelProb) % 8 == 0, "level records are copied to LDS by 4-byte words and stay 8-byte aligned");
static_assert(kMaxLev * kRecWords <= 3 * kBottomThreads, "a chain's records load in three words per thread");

// LDS arrays of a bottom level: x0, x1, b, each (ny+3) x (nx+3) with a zero frame
template <typename T>
struct LdsLevel {
    T* x0;
    T* x1;
    T* b;
    int pitch;
    __device__ int64_t at(int i, int j) const { return (int64_t)(i + 1) * pitch + (j + 1); }
    __device__ int ati(int i, int j) const { return (i + 1) * pitch + (j + 1); }
    __device__ T* x(int k) const { return k ? x1 : x0; } // select, not an indexed array (no scratch)
};

template <typename T>
__device__ __forceinline__ LdsLevel<T> lds_level(T* sm, const LevelProb& P)
{
    LdsLevel<T> L;
    L.pitch = P.nx + 3;
    const int64_t sz = (int64_t)(P.ny + 3) * L.pitch;
    L.x0 = sm + P.lds_off;
    L.x1 = L.x0 + sz;
    L.b = L.x1 + sz;
    return L;
}

// The boundary ring of a bottom level, one node per thread.
template <typename FB>
__device__ __forceinline__ void for_ring(const LevelProb& P, FB&& fb)
{
    const int nw = P.nx + 1, nr = 2 * nw + 2 * (P.ny - 1);
    for (int t = threadIdx.x; t < nr; t += kBottomThreads) {
        int i, j;
        if (t < nw) {
            i = 0;
            j = t;
        } else if (t < 2 * nw) {
            i = P.ny;
            j = t - nw;
        } else if (t < 2 * nw + P.ny - 1) {
            i = 1 + t - 2 * nw;
            j = 0;
        } else {
            i = 1 + t - 2 * nw - (P.ny - 1);
            j = P.nx;
        }
        fb(i, j);
    }
}

// Visit every node of a bottom level once: interior nodes by rows (whole
// waves take the interior row, no boundary lanes mixed in), then the
// boundary ring, one node per thread.
template <typename FI, typename FB>
__device__ __forceinline__ void for_nodes_split(const LevelProb& P, FI&& fi, FB&& fb)
{
    for (int i = 1 + (int)(threadIdx.x >> 6); i < P.ny; i += kBottomThreads / 64)
        for (int j = 1 + (int)(threadIdx.x & 63); j < P.nx; j += 64) fi(i, j);
    for_ring(P, fb);
}

template <typename T, bool TENS>
__device__ __forceinline__ void lds_sweep(const LevelProb& P, const LdsLevel<T>& S, int from, int to, T omega,
                                          bool zero_init)
{
    const Row<T> ri = interior_row<T>(P);
    const T wdi = omega * interior_dinv<T>(P);
    T* __restrict__ xo = S.x(to);
    const T* __restrict__ xi = S.x(from);
    auto general = [&](int i, int j) {
        if (!is_dirichlet(P, i, j)) {
            T bv;
            const int64_t k = S.at(i, j);
            xo[k] = smooth_val<T, T, TENS>(P, i, j, k, S.pitch, S.b, xi, omega, zero_init, bv,
                                     P.origin + (int64_t)i * P.pitch + j);
        }
    };
    for_nodes_split(
        P,
        [&](int i, int j) {
            if (P.aniso || P.vel) { // no common interior row
                general(i, j);
                return;
            }
            const int64_t k = S.at(i, j);
            const T bv = S.b[k];
            if (zero_init) {
                xo[k] = wdi * bv;
            } else {
                T xc, xw, xe, xs, xn, xsw, xne;
                load7(xi, k, S.pitch, xc, xw, xe, xs, xn, xsw, xne);
                xo[k] = xc + wdi * (bv - row_apply<TENS>(ri, xc, xw, xe, xs, xn, xsw, xne));
            }
        },
        general);
    __syncthreads();
}

// ---- bottom levels without a tensor field or velocity (the common case)
//
// A free ring node's row depends only on the walls it lies on (row_coefs_at:
// the element counts around the node and the Robin edges of its walls), so
// one table of 16 rows, indexed by wall_code, holds every ring row of a level
// together with the reciprocal of its diagonal -- the values row_coefs and
// row_dinv give at those nodes.  The interior runs in strips of kBotRows
// rows per thread from a register window (one load per row and column), with
// 32-bit LDS indices; the level's strips are dealt round-robin to the waves
// and its ring nodes start at the first wave with one strip fewer.

constexpr int kBotWaves = kBottomThreads / 64;
constexpr int kBotRows = 4;

__device__ __forceinline__ int wall_code(const LevelProb& P, int i, int j)
{
    return (j == 0 ? 1 : 0) | (j == P.nx ? 2 : 0) | (i == 0 ? 4 : 0) | (i == P.ny ? 8 : 0);
}

template <typename T>
struct RingTable {
    static constexpr int kRowWords = (int)(sizeof(Row<T>) / sizeof(T));
    static constexpr int kWords = 16 * (kRowWords + 1);
    T* mem;
    __device__ Row<T>* rows() const { return reinterpret_cast<Row<T>*>(mem); }
    __device__ T* dinv() const { return mem + 16 * kRowWords; }
    // threads 0..15 fill the table (entry c from a node with wall code c);
    // the caller synchronises before the first use
    template <bool TENS>
    __device__ void fill(const LevelProb& P) const
    {
        const int c = (int)threadIdx.x;
        if (c == 0 || c >= 16 || (c & 3) == 3 || (c & 12) == 12) return;
        const int j = (c & 1) ? 0 : (c & 2) ? P.nx : min(1, P.nx);
        const int i = (c & 4) ? 0 : (c & 8) ? P.ny : min(1, P.ny);
        const Row<T> rc = row_coefs_at<T>(P, i, j);
        rows()[c] = rc;
        dinv()[c] = (T)1 / row_diag<TENS>(rc);
    }
};
static_assert(sizeof(Row<float>) == 13 * sizeof(float) && sizeof(Row<double>) == 13 * sizeof(double),
              "ring rows are stored as 13 coefficients");

struct BotSplit {
    int ncb, items, ring0;
};
__device__ __forceinline__ BotSplit bot_split(const LevelProb& P)
{
    BotSplit s;
    s.ncb = (P.nx - 1 + 63) / 64; // column blocks over the interior columns 1 .. nx-1
    const int nst = (P.ny - 1 + kBotRows - 1) / kBotRows;
    s.items = (P.nx > 1 && P.ny > 1) ? s.ncb * nst : 0;
    s.ring0 = (s.items % kBotWaves) * 64;
    return s;
}

// f(i0, j, a0): interior strip of this thread's item (rows i0 .. i0+kBotRows-1,
// those below ny; column j < nx; a0 the LDS index of (i0, j))
template <typename T, typename F>
__device__ __forceinline__ void for_strips(const LevelProb& P, const LdsLevel<T>& S, const BotSplit& sp, F&& f)
{
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = (int)(threadIdx.x & 63);
    for (int it = wave; it < sp.items; it += kBotWaves) {
        const int cb = it % sp.ncb, st = it / sp.ncb;
        const int j = 1 + cb * 64 + lane;
        const int i0 = 1 + st * kBotRows;
        if (j < P.nx) f(i0, j, S.ati(i0, j));
    }
}

// g(i, j, a): this thread's ring nodes
template <typename T, typename G>
__device__ __forceinline__ void for_ring_nodes(const LevelProb& P, const LdsLevel<T>& S, const BotSplit& sp, G&& g)
{
    const int nw = P.nx + 1, nr = 2 * nw + 2 * (P.ny - 1);
    const int t0 = (int)((threadIdx.x + kBottomThreads - sp.ring0) & (kBottomThreads - 1));
    for (int t = t0; t < nr; t += kBottomThreads) {
        int i, j;
        if (t < nw) {
            i = 0;
            j = t;
        } else if (t < 2 * nw) {
            i = P.ny;
            j = t - nw;
        } else if (t < 2 * nw + P.ny - 1) {
            i = 1 + t - 2 * nw;
            j = 0;
        } else {
            i = 1 + t - 2 * nw - (P.ny - 1);
            j = P.nx;
        }
        g(i, j, S.ati(i, j));
    }
}

// the kBotRows + 2 rows i0-1 .. i0+kBotRows of columns j-1, j, j+1 (rows past
// ny + 1 read the frame row ny + 1; their values are not used)
template <typename T>
struct BotWindow {
    T w[kBotRows + 2], m[kBotRows + 2], e[kBotRows + 2];
    __device__ __forceinline__ void load(const T* __restrict__ x, int a0, int pitch, int i0, int ny)
    {
#pragma unroll
        for (int q = 0; q < kBotRows + 2; ++q) {
            const int a = a0 + (min(i0 - 1 + q, ny + 1) - i0) * pitch;
            w[q] = x[a - 1];
            m[q] = x[a];
            e[q] = x[a + 1];
        }
    }
};

template <typename T>
__device__ __forceinline__ void lds_load7(const T* __restrict__ v, int a, int pitch, T& c, T& w, T& e, T& s, T& n,
                                          T& sw, T& ne)
{
    c = v[a];
    w = v[a - 1];
    e = v[a + 1];
    s = v[a - pitch];
    n = v[a + pitch];
    sw = v[a - pitch - 1];
    ne = v[a + pitch + 1];
}

// damped Jacobi sweep of a plain bottom level (lds_sweep's values)
template <typename T, bool TENS>
__device__ __forceinline__ void lds_sweep_plain(const LevelProb& P, const LdsLevel<T>& S, const RingTable<T>& rt,
                                                int from, int to, T omega, bool zero_init)
{
    const BotSplit sp = bot_split(P);
    const Row<T> ri = interior_row<T>(P);
    const T wdi = omega * interior_dinv<T>(P);
    T* __restrict__ xo = S.x(to);
    const T* __restrict__ xi = S.x(from);
    const int pt = S.pitch;
    for_strips(P, S, sp, [&](int i0, int j, int a0) {
        T bl[kBotRows];
#pragma unroll
        for (int r = 0; r < kBotRows; ++r) bl[r] = S.b[a0 + (min(i0 + r, P.ny) - i0) * pt];
        if (zero_init) {
#pragma unroll
            for (int r = 0; r < kBotRows; ++r)
                if (i0 + r < P.ny) xo[a0 + r * pt] = wdi * bl[r];
            return;
        }
        BotWindow<T> X;
        X.load(xi, a0, pt, i0, P.ny);
#pragma unroll
        for (int r = 0; r < kBotRows; ++r)
            if (i0 + r < P.ny)
                xo[a0 + r * pt] = X.m[r + 1] + wdi * (bl[r] - row_apply<TENS>(ri, X.m[r + 1], X.w[r + 1], X.e[r + 1],
                                                                              X.m[r], X.m[r + 2], X.w[r], X.e[r + 2]));
    });
    for_ring_nodes(P, S, sp, [&](int i, int j, int a) {
        if (is_dirichlet(P, i, j)) return;
        const int c = wall_code(P, i, j);
        const Row<T> rc = rt.rows()[c];
        const T wd = omega * rt.dinv()[c];
        const T bv = S.b[a];
        if (zero_init) {
            xo[a] = wd * bv;
            return;
        }
        T xc, xw, xe, xs, xn, xsw, xne;
        lds_load7(xi, a, pt, xc, xw, xe, xs, xn, xsw, xne);
        xo[a] = xc + wd * (bv - row_apply<TENS>(rc, xc, xw, xe, xs, xn, xsw, xne));
    });
    __syncthreads();
}

// residual of a plain bottom level into r (0 on Dirichlet nodes)
template <typename T, bool TENS>
__device__ __forceinline__ void lds_residual_plain(const LevelProb& P, const LdsLevel<T>& S, const RingTable<T>& rt,
                                                   const T* __restrict__ x, T* __restrict__ r)
{
    const BotSplit sp = bot_split(P);
    const Row<T> ri = interior_row<T>(P);
    const int pt = S.pitch;
    for_strips(P, S, sp, [&](int i0, int j, int a0) {
        T bl[kBotRows];
#pragma unroll
        for (int q = 0; q < kBotRows; ++q) bl[q] = S.b[a0 + (min(i0 + q, P.ny) - i0) * pt];
        BotWindow<T> X;
        X.load(x, a0, pt, i0, P.ny);
#pragma unroll
        for (int q = 0; q < kBotRows; ++q)
            if (i0 + q < P.ny)
                r[a0 + q * pt] = bl[q] - row_apply<TENS>(ri, X.m[q + 1], X.w[q + 1], X.e[q + 1], X.m[q], X.m[q + 2],
                                                         X.w[q], X.e[q + 2]);
    });
    for_ring_nodes(P, S, sp, [&](int i, int j, int a) {
        if (is_dirichlet(P, i, j)) {
            r[a] = (T)0;
            return;
        }
        const Row<T> rc = rt.rows()[wall_code(P, i, j)];
        T xc, xw, xe, xs, xn, xsw, xne;
        lds_load7(x, a, pt, xc, xw, xe, xs, xn, xsw, xne);
        r[a] = S.b[a] - row_apply<TENS>(rc, xc, xw, xe, xs, xn, xsw, xne);
    });
}

// The V-cycle from the problem's bottom level down, entirely in LDS: the
// bottom level's right-hand side is read from HBM, its correction is written
// to `out` (the buffer its finer level's prolongation reads).  Dirichlet
// entries stay at the zero the LDS is cleared to.  The chain's level records
// are copied into LDS behind the level arrays with the right-hand side, in
// one batch of loads, so no level waits on a record load from memory.
template <typename T, bool TENS>
__global__ __launch_bounds__(kBottomThreads) void k_bottom(const BottomRef* __restrict__ refs,
                                                           const LevelProb* __restrict__ recs, const T* __restrict__ bin,
                                                           T* __restrict__ out, T omega, T omega2, int pre, int post,
                                                           const T* __restrict__ dense, const CgState* st)
{
    extern __shared__ __align__(16) unsigned char eq_smem[];
    __shared__ int fin[kMaxLev];
    __shared__ T red[kBottomThreads]; // coarsest-level partial products
    // the ring rows of a level share red's storage (red is used only by the
    // coarsest solve, between barriers): two tables, level k's at rtab(k),
    // so the next level's is filled while the current one is read
    static_assert(2 * RingTable<T>::kWords <= kBottomThreads, "ring tables fit the partials' array");
    auto rtab = [&](int k) { return RingTable<T>{red + (k & 1) * RingTable<T>::kWords}; };
    const BottomRef ref = refs[blockIdx.x];
    const int nl = ref.nl;
    T* sm = reinterpret_cast<T*>(eq_smem);
    // Every load of the set-up waits on ref alone, and all are issued before
    // the first use: the coarsest level's dense inverse (this thread's first
    // kPre entries, used after the down-sweep, so the solve there waits on no
    // load from memory), the bottom level's right-hand side (kLoad nodes per
    // thread), the chain's level records and the problem's state.
    constexpr int kPre = (sizeof(T) == 4 && !TENS) ? 32 : 0;
    T dpre[kPre > 0 ? kPre : 1];
    {
        const int n = ref.nz, nch = max(1, kBottomThreads / n);
        const int t = (int)threadIdx.x, ch = t / n, k = t - ch * n;
        const T* D = dense + ref.dense_off;
#pragma unroll
        for (int u = 0; u < kPre; ++u) {
            const int m = ch + u * nch;
            dpre[u] = (t < nch * n && m < n) ? D[(int64_t)m * n + k] : (T)0;
        }
    }
    constexpr int kLoad = 12;
    const int rw = ref.nx + 1, nb = rw * (ref.ny + 1);
    const float rrw = 1.0f / (float)rw;
    const T* __restrict__ bs = bin + ref.origin;
    T bv[kLoad];
    int bi[kLoad], bj[kLoad];
#pragma unroll
    for (int q = 0; q < kLoad; ++q) {
        const int m = q * kBottomThreads + (int)threadIdx.x;
        bv[q] = (T)0;
        bi[q] = -1;
        bj[q] = 0;
        if (m < nb) {
            const int i = div_floor_pos(m, rw, rrw), j = m - i * rw;
            bv[q] = bs[(int64_t)i * ref.pitch + j];
            bi[q] = i;
            bj[q] = j;
        }
    }
    const uint32_t* __restrict__ ws = reinterpret_cast<const uint32_t*>(recs + ref.rec);
    const int nwords = (nl + 1) * kRecWords;
    uint32_t wv[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int t = (int)threadIdx.x + q * kBottomThreads;
        wv[q] = t < nwords ? ws[t] : 0u;
    }
    const LevelProb P0 = recs[ref.rec];
    if (st[ref.prob].done) return;
    const LevelProb* lrec = reinterpret_cast<const LevelProb*>(eq_smem + bottom_rec_offset<T>(P0));
    {
        const LdsLevel<T> S = lds_level(sm, P0);
        for (int m = threadIdx.x; m < ref.lds_total; m += kBottomThreads) sm[m] = (T)0;
        const bool plain0 = !(TENS && (P0.aniso || P0.vel));
        if (plain0) rtab(0).template fill<TENS>(P0);
        __syncthreads();
        // b, and on a plain level the first sweep from zero (omega b / diag
        // on free nodes) with it
        const T wdi0 = omega * interior_dinv<T>(P0);
        auto put = [&](int i, int j, T v) {
            const int a = S.ati(i, j);
            S.b[a] = v;
            if (!plain0) return;
            if (interior(P0, i, j))
                S.x0[a] = wdi0 * v;
            else if (!is_dirichlet(P0, i, j))
                S.x0[a] = (omega * rtab(0).dinv()[wall_code(P0, i, j)]) * v;
        };
#pragma unroll
        for (int q = 0; q < kLoad; ++q)
            if (bi[q] >= 0) put(bi[q], bj[q], bv[q]);
        // levels of more nodes than one batch holds
        for (int m = kLoad * kBottomThreads + (int)threadIdx.x; m < nb; m += kBottomThreads) {
            const int i = div_floor_pos(m, rw, rrw), j = m - i * rw;
            put(i, j, bs[(int64_t)i * ref.pitch + j]);
        }
        uint32_t* wd = reinterpret_cast<uint32_t*>(eq_smem + bottom_rec_offset<T>(P0));
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const int t = (int)threadIdx.x + q * kBottomThreads;
            if (t < nwords) wd[t] = wv[q];
        }
        __syncthreads();
    }
    for (int k = 0; k < nl; ++k) {
        const LevelProb P = uniform_rec(lrec + k);
        const LevelProb Pc = uniform_rec(lrec + k + 1);
        const LdsLevel<T> S = lds_level(sm, P), C = lds_level(sm, Pc);
        // a plain level's ring rows and first sweep from zero were written by
        // the phase that formed its b (the load, or the restriction above)
        const bool plain = !(TENS && (P.aniso || P.vel));
        const RingTable<T> rt = rtab(k);
        // pre-sweep s is damped by omega (s even) or omega2 (s odd)
        auto sweep = [&](int from, int to, bool zero_init, T w) {
            if (plain)
                lds_sweep_plain<T, TENS>(P, S, rt, from, to, w, zero_init);
            else
                lds_sweep<T, TENS>(P, S, from, to, w, zero_init);
        };
        if (!plain) sweep(0, 0, true, omega);
        int cur = 0;
        for (int s = 1; s < pre; ++s) {
            sweep(cur, cur ^ 1, false, (s & 1) ? omega2 : omega);
            cur ^= 1;
        }
        if (threadIdx.x == 0) fin[k] = cur;
        // residual into the free buffer, then restrict it into the coarser b
        // (with the coarser level's first sweep when it is plain and smoothed)
        T* r = S.x(cur ^ 1);
        const bool cplain = !Pc.coarsest && !(TENS && (Pc.aniso || Pc.vel));
        const RingTable<T> rtc = rtab(k + 1);
        if (cplain) rtc.template fill<TENS>(Pc); // read after the barrier below
        if (plain) {
            lds_residual_plain<T, TENS>(P, S, rt, S.x(cur), r);
        } else {
            const Row<T> ri = interior_row<T>(P);
            const T* __restrict__ xc_ = S.x(cur);
            auto general = [&](int i, int j) {
                const int64_t kk = S.at(i, j);
                r[kk] = is_dirichlet(P, i, j) ? (T)0 : residual_val<T, TENS>(P, i, j, kk, S.pitch, S.b, xc_,
                                                                          P.origin + (int64_t)i * P.pitch + j);
            };
            for_nodes_split(
                P,
                [&](int i, int j) {
                    if (P.aniso || P.vel) {
                        general(i, j);
                        return;
                    }
                    const int64_t kk = S.at(i, j);
                    T xc, xw, xe, xs, xn, xsw, xne;
                    load7(xc_, kk, S.pitch, xc, xw, xe, xs, xn, xsw, xne);
                    r[kk] = S.b[kk] - row_apply<TENS>(ri, xc, xw, xe, xs, xn, xsw, xne);
                },
                general);
        }
        __syncthreads();
        {
            // interior coarse nodes are never Dirichlet
            const int64_t f0 = S.at(0, 0);
            const T wdic = omega * interior_dinv<T>(Pc);
            for_nodes_split(
                Pc,
                [&](int i, int j) {
                    const int a = C.ati(i, j);
                    const T v = restrict_val<T, false>(Pc, P, i, j, r, f0, S.pitch);
                    C.b[a] = v;
                    if (cplain) C.x0[a] = wdic * v;
                },
                [&](int i, int j) {
                    if (is_dirichlet(Pc, i, j)) return;
                    const int a = C.ati(i, j);
                    const T v = restrict_val<T, false>(Pc, P, i, j, r, f0, S.pitch);
                    C.b[a] = v;
                    if (cplain) C.x0[a] = (omega * rtc.dinv()[wall_code(Pc, i, j)]) * v;
                });
        }
        __syncthreads();
    }
    // coarsest level: x = A^{-1} b with the dense inverse, stored column by
    // column.  The columns are dealt round-robin to nch groups of n threads,
    // so a group reads whole columns (coalesced) and every load is
    // independent; the group partials are summed in a fixed order.
    {
        const LevelProb Pc = uniform_rec(lrec + nl);
        const LdsLevel<T> C = lds_level(sm, Pc);
        const int nw = Pc.nx + 1, n = nw * (Pc.ny + 1);
        const T* D = dense + Pc.dense_off;
        T* cb = C.x1; // b in natural order (x1 is free on the coarsest level)
        for (int m = threadIdx.x; m < n; m += kBottomThreads) {
            const int mi = m / nw, mj = m - mi * nw;
            cb[m] = C.b[C.at(mi, mj)];
        }
        __syncthreads();
        // n <= 1024 nodes, so nch * n <= kBottomThreads: one pass, t = threadIdx.x
        const int nch = max(1, kBottomThreads / n);
        for (int t = threadIdx.x; t < nch * n; t += kBottomThreads) {
            const int ch = t / n, k = t - ch * n;
            T s = (T)0;
#pragma unroll
            for (int u = 0; u < kPre; ++u) {
                const int m = ch + u * nch;
                if (m < n) s += dpre[u] * cb[m];
            }
            // the rest of the inverse comes from L2 or HBM: kBatch
            // independent loads in flight per batch, then the products in
            // the same m order
            constexpr int kBatch = 16;
            for (int m0 = ch + kPre * nch; m0 < n; m0 += kBatch * nch) {
                T dv[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    const int m = m0 + u * nch;
                    dv[u] = (m < n) ? D[(int64_t)m * n + k] : (T)0;
                }
#pragma unroll
                for (int u = 0; u < kBatch; ++u) {
                    const int m = m0 + u * nch;
                    if (m < n) s += dv[u] * cb[m];
                }
            }
            red[t] = s;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < n; k += kBottomThreads) {
            T s = (T)0;
            for (int ch = 0; ch < nch; ++ch) s += red[ch * n + k];
            const int i = k / nw, j = k - i * nw;
            C.x(0)[C.at(i, j)] = s;
        }
        __syncthreads();
    }
    int from = 0; // buffer of the coarser level's result
    for (int k = nl - 1; k >= 0; --k) {
        const LevelProb P = uniform_rec(lrec + k);
        const LevelProb Pc = uniform_rec(lrec + k + 1);
        const LdsLevel<T> S = lds_level(sm, P), C = lds_level(sm, Pc);
        int cur = fin[k];
        const bool plain = !(TENS && (P.aniso || P.vel));
        const RingTable<T> rt = rtab(k);
        if (plain) rt.template fill<TENS>(P); // read after the barrier below
        {
            // interior fine nodes are never Dirichlet
            T* __restrict__ xf = S.x(cur);
            const T* __restrict__ xcs = C.x(from);
            const int64_t c0 = C.at(0, 0);
            if (P.nx == 2 * Pc.nx && P.ny == 2 * Pc.ny) {
                // nested: the interior in strips (strip rows i0 .. i0+3 start
                // odd, so they read coarse rows (i0-1)/2 .. (i0-1)/2 + 2 at
                // columns j/2 and j/2 + 1, all loaded first), prolong_val's
                // values; the ring node by node
                const BotSplit sp = bot_split(P);
                for_strips(P, S, sp, [&](int i0, int j, int a0) {
                    const int R0 = (i0 - 1) >> 1, jc = j >> 1, jo = j & 1;
                    const int ca = C.ati(R0, jc);
                    T cv[3][2], xv[kBotRows];
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        const int ar = ca + min(q, Pc.ny - R0) * C.pitch;
                        cv[q][0] = xcs[ar];
                        cv[q][1] = xcs[ar + 1];
                    }
#pragma unroll
                    for (int r = 0; r < kBotRows; ++r) xv[r] = xf[a0 + (min(i0 + r, P.ny) - i0) * S.pitch];
#pragma unroll
                    for (int r = 0; r < kBotRows; ++r) {
                        if (i0 + r >= P.ny) continue;
                        // fine row i0 + r: coarse row R0 + (r + 1) / 2, odd when r is even
                        const int q = (r + 1) >> 1;
                        const T x00 = cv[q][0];
                        T add;
                        if ((r & 1) == 0) { // odd fine row: the vertical or diagonal coarse edge
                            const T xo = jo ? cv[q + 1][1] : cv[q + 1][0];
                            add = (T)0.5 * (x00 + xo);
                        } else { // even fine row: the coincident node or the horizontal edge
                            add = jo ? (T)0.5 * (x00 + cv[q][1]) : x00;
                        }
                        xf[a0 + r * S.pitch] = xv[r] + add;
                    }
                });
                for_ring_nodes(P, S, sp, [&](int i, int j, int a) {
                    if (!is_dirichlet(P, i, j)) xf[a] += prolong_val<T>(P, Pc, i, j, xcs, c0, C.pitch);
                });
            } else {
                for_nodes_split(
                    P, [&](int i, int j) { xf[S.ati(i, j)] += prolong_val<T>(P, Pc, i, j, xcs, c0, C.pitch); },
                    [&](int i, int j) {
                        if (!is_dirichlet(P, i, j)) xf[S.ati(i, j)] += prolong_val<T>(P, Pc, i, j, xcs, c0, C.pitch);
                    });
            }
        }
        __syncthreads();
        for (int s = 0; s < post; ++s) {
            // the pre-sweeps' dampings in reverse order
            const T w = ((post - 1 - s) & 1) ? omega2 : omega;
            if (plain)
                lds_sweep_plain<T, TENS>(P, S, rt, cur, cur ^ 1, w, false);
            else
                lds_sweep<T, TENS>(P, S, cur, cur ^ 1, w, false);
            cur ^= 1;
        }
        from = cur;
    }
    {
        const LevelProb P0 = uniform_rec(lrec);
        const LdsLevel<T> S = lds_level(sm, P0);
        const T* __restrict__ xs = S.x(from);
        auto put = [&](int i, int j) { out[P0.origin + (int64_t)i * P0.pitch + j] = xs[S.ati(i, j)]; };
        for_nodes_split(P0, put, put);
    }
}

// ------------------------------------------------------------ fused coarse levels
//
// One launch for a coarse level's way down (pre-sweeps 1 and 2 from x = 0,
// the residual, its restriction) and one for its way up (the coarse-grid
// correction, post-sweeps 1 and 2), in place of the four to seven launches of
// the kernels above; for launches without a tensor field or velocity.  A
// workgroup takes a tile of kFW x kFH fine nodes and stages the region within
// kFM nodes around it in LDS.  The way down owns the coarse nodes (I, J)
// whose fine position (floor(I nyf/nyc), floor(J nxf/nxc)) lies in its tile;
// with a coarsening ratio in (1, 2] their restriction support reaches at most
// one fine node before the tile and two past it, and the two sweeps' halos
// two more, so every value the tile needs lies in the region.  Values outside
// the grid are zero, as in the pools' frames.  Each node's arithmetic is that
// of smooth_val, residual_val, restrict_gen and prolong_val.  The tile record
// (FusedTile) carries the owned ranges and both levels' geometry, so the
// first loads of a workgroup wait on that record alone; the level record is
// read alongside them, for the rows.

constexpr int kFW = kFusedTileW, kFM = 3;
constexpr int kFRW = 64; // LDS row of a region: columns j0-3 .. j0+kFW+3
static_assert(kFW + 2 * kFM + 1 <= kFRW, "a region row fits one wave");
// a tile of TH rows: region rows i0-3 .. i0+TH+3, and region rows per wave
#define EQ_FUSED_DIMS(TH)                                                                                              \
    constexpr int kFH = (TH);                                                                                          \
    constexpr int kFRH = kFH + 2 * kFM + 1;                                                                            \
    constexpr int kFRows = (kFRH + kWaves - 1) / kWaves;

// the two levels' geometry as level records (the fields the wall tests and
// the grid transfers read)
__device__ __forceinline__ LevelProb fused_fine(const FusedTile& f)
{
    LevelProb G{};
    G.nx = f.nx;
    G.ny = f.ny;
    G.pitch = f.pitch;
    G.dmask = f.dmask;
    G.origin = f.origin;
    G.rnx = f.rnx;
    G.rny = f.rny;
    G.rnxf = f.rnxf;
    G.rnyf = f.rnyf;
    return G;
}
__device__ __forceinline__ LevelProb fused_coarse(const FusedTile& f)
{
    LevelProb G{};
    G.nx = f.cnx;
    G.ny = f.cny;
    G.pitch = f.cpitch;
    G.dmask = f.cdmask;
    G.origin = f.corigin;
    G.rnxf = f.crnxf;
    G.rnyf = f.crnyf;
    return G;
}

// A x at a free node of a plain problem (its row: the level's interior row,
// or the ring table's), with its damped inverse diagonal
template <typename T>
__device__ __forceinline__ T fused_apply(const LevelProb& P, const RingTable<T>& rt, int i, int j, T& wd, T omega,
                                         T vc, T vw, T ve, T vs, T vn, T vsw, T vne)
{
    if (interior(P, i, j)) {
        wd = omega * interior_dinv<T>(P);
        return row_apply<false>(interior_row<T>(P), vc, vw, ve, vs, vn, vsw, vne);
    }
    const int c = wall_code(P, i, j);
    wd = omega * rt.dinv()[c];
    return row_apply<false>(rt.rows()[c], vc, vw, ve, vs, vn, vsw, vne);
}

template <typename T>
__device__ __forceinline__ T fused_wd(const LevelProb& P, const RingTable<T>& rt, int i, int j, T omega)
{
    return interior(P, i, j) ? omega * interior_dinv<T>(P) : omega * rt.dinv()[wall_code(P, i, j)];
}

// Region rows ra .. rb of this lane's column, K consecutive rows per wave:
// the wave's 3 x (K+2) window of src is loaded first, then f(k, row, c, w,
// e, s, n, sw, ne) runs for each of its rows, k = 0 .. K-1 (rows past rb + 1
// read row rb + 1; their values are not used).
template <int K, typename T, typename F>
__device__ __forceinline__ void region_strips(const T* __restrict__ src, int ra, int rb, int lane, F&& f)
{
    const int r0 = ra + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * K;
    if (r0 > rb) return;
    T wv[K + 2], mv[K + 2], ev[K + 2];
#pragma unroll
    for (int q = 0; q < K + 2; ++q) {
        const int a = min(r0 - 1 + q, rb + 1) * kFRW + lane;
        wv[q] = src[a - 1];
        mv[q] = src[a];
        ev[q] = src[a + 1];
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (r0 + k <= rb) f(k, r0 + k, mv[k + 1], wv[k + 1], ev[k + 1], mv[k], mv[k + 2], wv[k], ev[k + 2]);
}

// X1: sweep 1 is already in x1 (formed by the kernel that formed b);
// otherwise it is formed here from b, read in type TB and rounded to T (the
// finest level reads the Krylov residual itself).  TH: tile rows.
template <typename T, int TH, bool X1, typename TB = T>
__global__ __launch_bounds__(kBlock) void k_level_down(const FusedTile* __restrict__ tiles, int ntiles,
                                                       const LevelProb* __restrict__ probs, const TB* __restrict__ b,
                                                       T* __restrict__ xout, T* __restrict__ bc, T omega, T omega2,
                                                       const CgState* st, const T* __restrict__ x1)
{
    EQ_FUSED_DIMS(TH)
    __shared__ T sb[kFRH * kFRW], s1[kFRH * kFRW], s2[kFRH * kFRW];
    __shared__ T ring[RingTable<T>::kWords];
    const int t = tile_index();
    if (t >= ntiles) return;
    const FusedTile f = tiles[t];
    const LevelProb G = fused_fine(f), Gc = fused_coarse(f);
    const int i0 = f.i0, j0 = f.j0, ri = i0 - kFM, rj = j0 - kFM;
    const int ilo = f.ilo, ihi = f.ihi, jlo = f.jlo, jhi = f.jhi;
    const int lane = (int)(threadIdx.x & 63), w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int j = rj + lane;
    const bool jin = j >= 0 && j <= G.nx;
    // b (and x1) over the support and two nodes around it, every load first:
    // each lane loads from its node clamped into that region (a load under a
    // branch would wait for its value at the branch's end, one row at a time)
    // and keeps zero where it is outside; b is rounded to T where it is staged
    TB bv[kFRows];
    T xv[kFRows];
    uint32_t okm = 0;
    {
        const int ia = max(ilo - 2, 0), ib = min(ihi + 2, G.ny);
        const int jc = min(max(j, max(jlo - 2, 0)), min(jhi + 2, G.nx));
        const bool jok = j >= jlo - 2 && j <= jhi + 2 && jin;
#pragma unroll
        for (int q = 0; q < kFRows; ++q) {
            const int i = ilo - 2 + w + q * kWaves;
            const int64_t g = G.origin + (int64_t)min(max(i, ia), ib) * G.pitch + jc;
            bv[q] = b[g];
            xv[q] = X1 ? x1[g] : (T)0;
            if (jok && i <= ihi + 2 && i >= 0 && i <= G.ny) okm |= 1u << q;
        }
#pragma unroll
        for (int q = 0; q < kFRows; ++q)
            if (!((okm >> q) & 1u)) {
                bv[q] = (TB)0;
                xv[q] = (T)0;
            }
    }
    const LevelProb P = probs[f.pi];
    if (st[f.prob].done) return;
    const RingTable<T> rt{ring};
    if (f.inner & 1) {
        // every node of the region interior (none on a wall, none outside the
        // grid): the level's shared row, no per-node tests, rows in strips
        const Row<T> rin = interior_row<T>(P);
        const T wdi = omega * interior_dinv<T>(P), wdi2 = omega2 * interior_dinv<T>(P);
#pragma unroll
        for (int q = 0; q < kFRows; ++q) {
            const int i = ilo - 2 + w + q * kWaves;
            if (i > ihi + 2 || j < jlo - 2 || j > jhi + 2) continue;
            const int a = (i - ri) * kFRW + lane;
            const T bq = (T)bv[q];
            sb[a] = bq;
            s1[a] = X1 ? xv[q] : wdi * bq;
        }
        __syncthreads();
        const bool jt = j >= j0 && j < j0 + kFW; // a column of the tile
        if (j >= jlo - 1 && j <= jhi + 1)
            region_strips<(kFH + 5 + kWaves - 1) / kWaves>(
                s1, ilo - 1 - ri, ihi + 1 - ri, lane, [&](int, int r, T c, T vw, T ve, T vs, T vn, T vsw, T vne) {
                    const int a = r * kFRW + lane, i = r + ri;
                    const T x2 = c + wdi2 * (sb[a] - row_apply<false>(rin, c, vw, ve, vs, vn, vsw, vne));
                    s2[a] = x2;
                    if (i >= i0 && i < i0 + kFH) {
                        T* __restrict__ row = xout + (G.origin + (int64_t)i * G.pitch);
                        if (jt) row[j] = x2;
                    }
                });
        __syncthreads();
        if (j >= jlo && j <= jhi)
            region_strips<(kFH + 3 + kWaves - 1) / kWaves>(
                s2, ilo - ri, ihi - ri, lane, [&](int, int r, T c, T vw, T ve, T vs, T vn, T vsw, T vne) {
                    const int a = r * kFRW + lane;
                    s1[a] = sb[a] - row_apply<false>(rin, c, vw, ve, vs, vn, vsw, vne);
                });
        __syncthreads();
    } else {
        rt.template fill<false>(P);
        __syncthreads();
        // sweep 1 from x = 0 there
#pragma unroll
        for (int q = 0; q < kFRows; ++q) {
            const int i = ilo - 2 + w + q * kWaves;
            if (i > ihi + 2 || j < jlo - 2 || j > jhi + 2) continue;
            const int a = (i - ri) * kFRW + lane;
            const bool in = jin && i >= 0 && i <= G.ny;
            const T bq = (T)bv[q];
            sb[a] = bq;
            s1[a] = (in && !is_dirichlet(G, i, j)) ? (X1 ? xv[q] : fused_wd(P, rt, i, j, omega) * bq) : (T)0;
        }
        __syncthreads();
        // sweep 2 one node around the support; the tile's own nodes go to xout
#pragma unroll
        for (int q = 0; q < kFRows; ++q) {
            const int i = ilo - 1 + w + q * kWaves;
            if (i > ihi + 1 || j < jlo - 1 || j > jhi + 1) continue;
            const int a = (i - ri) * kFRW + lane;
            const bool in = jin && i >= 0 && i <= G.ny;
            T x2 = (T)0;
            if (in && !is_dirichlet(G, i, j)) {
                T wd;
                const T xc = s1[a];
                const T ax = fused_apply(P, rt, i, j, wd, omega2, xc, s1[a - 1], s1[a + 1], s1[a - kFRW], s1[a + kFRW],
                                         s1[a - kFRW - 1], s1[a + kFRW + 1]);
                x2 = xc + wd * (sb[a] - ax);
            }
            s2[a] = x2;
            if (in && i >= i0 && i < i0 + kFH && j >= j0 && j < j0 + kFW)
                xout[G.origin + (int64_t)i * G.pitch + j] = x2;
        }
        __syncthreads();
        // residual over the support (zero outside the grid, where the nested
        // restriction reads one node past the edge); it replaces x1
#pragma unroll
        for (int q = 0; q < kFRows; ++q) {
            const int i = ilo - 1 + w + q * kWaves;
            if (i > ihi + 1 || j < jlo - 1 || j > jhi + 1) continue;
            const int a = (i - ri) * kFRW + lane;
            const bool in = jin && i >= 0 && i <= G.ny;
            if (!in) {
                s1[a] = (T)0;
                continue;
            }
            if (i < ilo || i > ihi || j < jlo || j > jhi) continue;
            T rv = (T)0;
            if (!is_dirichlet(G, i, j)) {
                T wd;
                const T ax = fused_apply(P, rt, i, j, wd, omega, s2[a], s2[a - 1], s2[a + 1], s2[a - kFRW],
                                         s2[a + kFRW], s2[a - kFRW - 1], s2[a + kFRW + 1]);
                rv = sb[a] - ax;
            }
            s1[a] = rv;
        }
        __syncthreads();
    }
    // the owned coarse nodes: the restricted residual (0 on Dirichlet nodes)
    auto rf = [&](int fi, int fj, int64_t) -> T { return s1[(fi - ri) * kFRW + (fj - rj)]; };
    // (two coarse rows a pass per wave: lanes 0-31 and 32-63, 32 columns each)
    for (int J0 = f.Ja; J0 < f.Jb; J0 += 32) {
        const int J = J0 + (lane & 31);
        for (int I = f.Ia + 2 * w + (lane >> 5); I < f.Ib; I += 2 * kWaves) {
            if (J >= f.Jb) continue;
            const T v = is_dirichlet(Gc, I, J) ? (T)0 : restrict_gen<T, false>(Gc, G, I, J, 0, rf);
            bc[Gc.origin + (int64_t)I * Gc.pitch + J] = v;
        }
    }
}

// DOT (the finest level): the last sweep's b is r (type TI) rounded to T, as
// in k_smooth, and the tile adds r.x_out to its partial sum in slot gt.  b is
// read in type TB and rounded to T (the finest level: r itself).
template <typename T, typename TI, int TH, bool DOT, typename TB = T>
__global__ __launch_bounds__(kBlock) void k_level_up(const FusedTile* __restrict__ tiles, int ntiles,
                                                     const LevelProb* __restrict__ probs, const T* __restrict__ xc,
                                                     const TB* __restrict__ b, const T* __restrict__ xin,
                                                     T* __restrict__ xout, T omega, T omega2, const CgState* st,
                                                     const TI* __restrict__ rin, double* partial)
{
    EQ_FUSED_DIMS(TH)
    __shared__ T sb[kFRH * kFRW], sy[kFRH * kFRW], ss[kFRH * kFRW];
    __shared__ T ring[RingTable<T>::kWords];
    const int t = tile_index();
    if (t >= ntiles) return;
    const FusedTile f = tiles[t];
    const LevelProb G = fused_fine(f), Gc = fused_coarse(f);
    const int i0 = f.i0, j0 = f.j0, ri = i0 - kFM, rj = j0 - kFM;
    const int lane = (int)(threadIdx.x & 63), w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int j = rj + lane;
    const bool jq = j >= j0 - 2 && j <= j0 + kFW + 1 && j >= 0 && j <= G.nx;
    // x + P x_c two nodes around the tile (x alone on Dirichlet nodes), and b
    constexpr int kUpRows = (kFH + 4 + kWaves - 1) / kWaves;
    // Every lane loads from its node clamped into that region, with no branch
    // around the loads (a load under a branch waits for its value at the
    // branch's end, one row at a time), and keeps zero outside it; b is
    // rounded to T where it is staged.
    T xv[kUpRows], cv[kUpRows];
    TB bv[kUpRows];
    {
        const int ia = max(i0 - 2, 0), ib = min(i0 + kFH + 1, G.ny);
        const int jc = min(max(j, max(j0 - 2, 0)), min(j0 + kFW + 1, G.nx));
        uint32_t okm = 0, dm = 0;
        // the rows' loads in one straight run (the transfer's kind is uniform)
        auto rows = [&](auto&& pval) {
#pragma unroll
            for (int q = 0; q < kUpRows; ++q) {
                const int i = i0 - 2 + w + q * kWaves;
                const int ic = min(max(i, ia), ib);
                const int64_t g = G.origin + (int64_t)ic * G.pitch + jc;
                xv[q] = xin[g];
                bv[q] = b[g];
                cv[q] = pval(ic, jc);
                if (jq && i <= i0 + kFH + 1 && i >= 0 && i <= G.ny) {
                    okm |= 1u << q;
                    if (is_dirichlet(G, i, j)) dm |= 1u << q;
                }
            }
        };
        if (G.nx == 2 * Gc.nx && G.ny == 2 * Gc.ny)
            rows([&](int ic, int jc2) { return prolong_nested<T>(ic, jc2, Gc.nx, Gc.ny, xc, Gc.origin, Gc.pitch); });
        else
            rows([&](int ic, int jc2) { return prolong_val<T>(G, Gc, ic, jc2, xc, Gc.origin, Gc.pitch); });
#pragma unroll
        for (int q = 0; q < kUpRows; ++q) {
            if (!((okm >> q) & 1u)) {
                xv[q] = (T)0;
                bv[q] = (TB)0;
            }
            if (!((okm >> q) & 1u) || ((dm >> q) & 1u)) cv[q] = (T)0;
        }
    }
    const bool inner = (f.inner & 2) != 0;
    constexpr int kUp = (kFH + kWaves - 1) / kWaves; // tile rows per wave in the inner path's last sweep
    const bool jt = j >= j0 && j < j0 + kFW;
    const LevelProb P = probs[f.pi];
    if (st[f.prob].done) return;
    const RingTable<T> rt{ring};
    if (!inner) rt.template fill<false>(P);
#pragma unroll
    for (int q = 0; q < kUpRows; ++q) {
        const int i = i0 - 2 + w + q * kWaves;
        if (i > i0 + kFH + 1 || j < j0 - 2 || j > j0 + kFW + 1) continue;
        const int a = (i - ri) * kFRW + lane;
        sy[a] = xv[q] + cv[q];
        sb[a] = (T)bv[q];
    }
    __syncthreads();
    double dot[1] = {0.0};
    if (inner) {
        // every node of the region interior: the shared row, rows in strips
        const Row<T> rw = interior_row<T>(P);
        const T wdi = omega * interior_dinv<T>(P), wdi2 = omega2 * interior_dinv<T>(P);
        if (j >= j0 - 1 && j <= j0 + kFW)
            region_strips<(kFH + 2 + kWaves - 1) / kWaves>(
                sy, i0 - 1 - ri, i0 + kFH - ri, lane, [&](int, int r, T c, T vw, T ve, T vs, T vn, T vsw, T vne) {
                    const int a = r * kFRW + lane;
                    ss[a] = c + wdi * (sb[a] - row_apply<false>(rw, c, vw, ve, vs, vn, vsw, vne));
                });
        __syncthreads();
        if (jt)
            region_strips<kUp>(
                ss, i0 - ri, i0 + kFH - 1 - ri, lane, [&](int, int r, T c, T vw, T ve, T vs, T vn, T vsw, T vne) {
                    const int a = r * kFRW + lane, i = r + ri;
                    const int64_t g0 = G.origin + (int64_t)i * G.pitch;
                    T bb = sb[a];
                    TI rv = (TI)0;
                    if (DOT) {
                        rv = (rin + g0)[j];
                        bb = (T)rv;
                    }
                    const T o = c + wdi2 * (bb - row_apply<false>(rw, c, vw, ve, vs, vn, vsw, vne));
                    (xout + g0)[j] = o;
                    if (DOT) dot[0] += (double)rv * (double)o;
                });
        if (DOT) tile_partial<1>(dot, partial, f.gt);
        return;
    }
    // post-sweep 1 one node around the tile
#pragma unroll
    for (int q = 0; q < kUpRows; ++q) {
        const int i = i0 - 1 + w + q * kWaves;
        if (i > i0 + kFH || j < j0 - 1 || j > j0 + kFW) continue;
        const int a = (i - ri) * kFRW + lane;
        const bool in = j >= 0 && j <= G.nx && i >= 0 && i <= G.ny;
        T s = (T)0;
        if (in && !is_dirichlet(G, i, j)) {
            T wd;
            const T yc = sy[a];
            const T ay = fused_apply(P, rt, i, j, wd, omega, yc, sy[a - 1], sy[a + 1], sy[a - kFRW], sy[a + kFRW],
                                     sy[a - kFRW - 1], sy[a + kFRW + 1]);
            s = yc + wd * (sb[a] - ay);
        }
        ss[a] = s;
    }
    __syncthreads();
    // post-sweep 2 on the tile
#pragma unroll
    for (int q = 0; q < (kFH + kWaves - 1) / kWaves; ++q) {
        const int i = i0 + w + q * kWaves;
        if (i >= i0 + kFH || i > G.ny || j < j0 || j >= j0 + kFW || j > G.nx) continue;
        const int a = (i - ri) * kFRW + lane;
        const int64_t g = G.origin + (int64_t)i * G.pitch + j;
        T o = (T)0;
        if (!is_dirichlet(G, i, j)) {
            T wd;
            const T sc = ss[a];
            const T as = fused_apply(P, rt, i, j, wd, omega2, sc, ss[a - 1], ss[a + 1], ss[a - kFRW], ss[a + kFRW],
                                     ss[a - kFRW - 1], ss[a + kFRW + 1]);
            T bb = sb[a];
            TI rv = (TI)0;
            if (DOT) {
                rv = rin[g];
                bb = (T)rv;
            }
            o = sc + wd * (bb - as);
            if (DOT) dot[0] += (double)rv * (double)o;
        }
        xout[g] = o;
    }
    if (DOT) tile_partial<1>(dot, partial, f.gt);
}

// ------------------------------------------------------------ diagnostics

// mode 0: y = A x (Dirichlet rows identity); mode 1: y = b(x) (Dirichlet rows = g)
template <int R0, bool TENS>
__global__ __launch_bounds__(kBlock) void k_apply_full(LevelView L, int pi_only, const double* __restrict__ x,
                                                       double* __restrict__ y, int mode)
{
    NodeCtx c;
    if (!decode<R0>(L, c)) return;
    if (c.pi != pi_only) return;
    const LevelProb& P = *c.P; // reads P.gw[wall]: a register copy would go to scratch
    each_row<R0>(c, [&](const NodeCtx& n) {
        if (!n.in) return;
        double vc, vw, ve, vs, vn, vsw, vne;
        load7(x, n.idx, P.pitch, vc, vw, ve, vs, vn, vsw, vne);
        if (is_dirichlet(P, n.i, n.j)) {
            y[n.idx] = (mode == 0) ? vc : dirichlet_value(P, n.i, n.j);
        } else if (mode == 0) {
            const Row<double> rc = row_coefs<double, TENS>(P, n.i, n.j, n.idx);
            y[n.idx] = row_apply<TENS>(rc, vc, vw, ve, vs, vn, vsw, vne);
        } else {
            y[n.idx] = rhs_row(P, n.i, n.j, vc, vw, ve, vs, vn, vsw, vne);
        }
    });
}

// ------------------------------------------------------------ boundary flux

// -int grad(u).n ds over all walls with the P1 gradient of the triangle that
// owns each boundary edge (fenics/boundary.h:2652-2742), times dt*D
// (src/fHSL.cpp:156-160).  One block, one problem.
// Block sum of the boundary terms; the total is valid in thread 0.
__device__ double flux_sum(const LevelProb& P, const double* __restrict__ u)
{
    const int nx = P.nx, ny = P.ny, pt = P.pitch;
    const double* U = u + P.origin;
    double v[1] = {0.0};
    for (int j = threadIdx.x; j < nx; j += blockDim.x) {
        // bottom edge (0,j)-(0,j+1) in the lower-right triangle of cell (0,j)
        v[0] += P.hx * (U[pt + j + 1] - U[j + 1]) / P.hy;
        // top edge (ny,j)-(ny,j+1) in the upper-left triangle of cell (ny-1,j)
        v[0] += P.hx * (U[(int64_t)(ny - 1) * pt + j] - U[(int64_t)ny * pt + j]) / P.hy;
    }
    for (int i = threadIdx.x; i < ny; i += blockDim.x) {
        // left edge (i,0)-(i+1,0) in the upper-left triangle of cell (i,0)
        v[0] += P.hy * (U[(int64_t)(i + 1) * pt + 1] - U[(int64_t)(i + 1) * pt]) / P.hx;
        // right edge (i,nx)-(i+1,nx) in the lower-right triangle of cell (i,nx-1)
        v[0] += P.hy * (U[(int64_t)i * pt + nx - 1] - U[(int64_t)i * pt + nx]) / P.hx;
    }
    block_sum<1, 256>(v);
    return v[0];
}

__global__ __launch_bounds__(256) void k_flux(const LevelProb* __restrict__ probs, int pi,
                                              const double* __restrict__ u, double* out)
{
    const LevelProb& P = probs[pi];
    const double v = flux_sum(P, u);
    if (threadIdx.x == 0) *out = P.Ddt * v;
}

// One block per boundary well: the flux of the new field, the well update of
// Simulation::computeBoundaryWell (src/simulation.cpp:581-604) in the
// reference's operation order (w += flux/scaling; w -= (dt*rate)*w; clamp at
// 0), and setBoundaryValues(w) on every Dirichlet wall of the problem's level-0
// record (src/fHSL.cpp:601-603), read by the next solve's kernels.
__global__ __launch_bounds__(256) void k_wells(LevelProb* probs, WellDev* wells, const double* __restrict__ u)
{
    WellDev& W = wells[blockIdx.x];
    LevelProb& P = probs[W.prob];
    const double v = flux_sum(P, u);
    if (threadIdx.x != 0) return;
    const double flux = P.Ddt * v;
    double w = __dadd_rn(W.w, flux / W.scaling);
    w = __dsub_rn(w, __dmul_rn(__dmul_rn(W.dt, W.decay), w));
    if (w < 0.0) {
        w = 0.0;
        W.underflows += 1;
    }
    W.w = w;
    W.flux = flux;
    for (int k = 0; k < 4; ++k)
        if ((P.dmask >> k) & 1) P.g[k] = w;
}

// Cell-driven tensor, coarse level: each node takes the tensor of the nearest
// node of the next finer level (the injection set_tensor applies on the host),
// for the problems whose tensor comes from the cells.  One block per tile.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tensor_inject(LevelView L, const LevelProb* __restrict__ fine, int R)
{
    const TileRef t = L.tiles[blockIdx.x];
    const LevelProb& P = L.probs[t.pi];
    if (!P.celltens) return;
    const LevelProb& F = fine[P.fidx];
    const T* src = std::is_same<T, float>::value ? (const T*)F.tens_f : (const T*)F.tens_d;
    T* dst = std::is_same<T, float>::value ? (T*)P.tens_f : (T*)P.tens_d;
    const int j = t.j0 + (int)(threadIdx.x & 63);
    if (j > P.nx) return;
    const int jf = (int)((2LL * j * F.nx + P.nx) / (2LL * P.nx));
    for (int r = 0; r < R; ++r) {
        const int i = t.i0 + (int)(threadIdx.x >> 6) * R + r;
        if (i > P.ny) break;
        const int i_f = (int)((2LL * i * F.ny + P.ny) / (2LL * P.ny));
        const int64_t kf = F.origin + (int64_t)i_f * F.pitch + jf, kc = P.origin + (int64_t)i * P.pitch + j;
        for (int c = 0; c < 3; ++c) dst[kc + c * P.tstride] = src[kf + c * F.tstride];
    }
}

// Cell-driven tensor, coarsest level: the dense inverse of the level's
// operator (Dirichlet rows identity), as the host builds it for set_tensor:
// assemble in FP64 from the level's tensor, banded LU without pivoting, then
// one column of the inverse per thread, stored column-major in the smoother's
// precision.  One block per problem; scratch A (n*n), X (n*n), tensor (3n).
template <typename T>
__global__ __launch_bounds__(256) void k_coarse_inverse(LevelProb* const* __restrict__ levels,
                                                        const InvJob* __restrict__ jobs, double* scratch,
                                                        T* __restrict__ dense)
{
    const InvJob J = jobs[blockIdx.x];
    const LevelProb& P = levels[J.lev][J.idx];
    const int nw = P.nx + 1, n = nw * (P.ny + 1), bw = nw + 1;
    double* A = scratch + J.scratch;
    double* X = A + (int64_t)n * n;
    double* t11 = X + (int64_t)n * n;
    double* t22 = t11 + n;
    double* t12 = t22 + n;
    const T* ts = std::is_same<T, float>::value ? (const T*)P.tens_f : (const T*)P.tens_d;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int i = k / nw, j = k - i * nw;
        const int64_t q = P.origin + (int64_t)i * P.pitch + j;
        t11[k] = (double)ts[q];
        t22[k] = (double)ts[q + P.tstride];
        t12[k] = (double)ts[q + 2 * P.tstride];
    }
    for (int64_t k = threadIdx.x; k < (int64_t)n * n; k += blockDim.x) A[k] = 0.0;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int i = k / nw, j = k - i * nw;
        if (is_dirichlet(P, i, j)) {
            A[(int64_t)k * n + k] = 1.0;
            continue;
        }
        Row<double> r = row_coefs_at<double>(P, i, j);
        aniso_stiffness<double>(P, i, j, t11 + k, t22 + k, t12 + k, nw, r);
        double* row = A + (int64_t)k * n;
        row[k] += row_diag(r);
        auto put = [&](int ii, int jj, double v) {
            if (ii < 0 || jj < 0 || ii > P.ny || jj > P.nx || v == 0.0) return;
            if (is_dirichlet(P, ii, jj)) return;
            row[jj + ii * nw] += v;
        };
        put(i, j - 1, r.mw - r.kw);
        put(i, j + 1, r.me - r.ke);
        put(i - 1, j, r.ms - r.ks);
        put(i + 1, j, r.mn - r.kn);
        put(i - 1, j - 1, r.msw - r.ksw);
        put(i + 1, j + 1, r.mne - r.kne);
    }
    __syncthreads();
    // banded LU in place, multipliers below the diagonal
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    for (int c = 0; c < n; ++c) {
        const double d = A[(int64_t)c * n + c];
        if (!(d > 0.0)) {
            if (threadIdx.x == 0) bad = 1;
            break; // uniform: every thread reads the same pivot
        }
        const int rmax = min(n - 1, c + bw);
        for (int r = c + 1 + (int)threadIdx.x; r <= rmax; r += blockDim.x) A[(int64_t)r * n + c] /= d;
        __syncthreads();
        const int w = rmax - c;
        for (int e = threadIdx.x; e < w * w; e += blockDim.x) {
            const int r = c + 1 + e / w, q = c + 1 + e % w;
            const double f = A[(int64_t)r * n + c];
            if (f != 0.0) A[(int64_t)r * n + q] -= f * A[(int64_t)c * n + q];
        }
        __syncthreads();
    }
    __syncthreads();
    if (bad) return; // cannot happen for Dx, Dy > 0 (D11 D22 - D12^2 = Dx Dy): the old inverse stays
    // column m of the inverse: L y = e_m, U x = y
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        double* x = X + (int64_t)m * n;
        for (int r = 0; r < n; ++r) x[r] = (r == m) ? 1.0 : 0.0;
        for (int r = m + 1; r < n; ++r) {
            double v = x[r];
            for (int q = max(m, r - bw); q < r; ++q) v -= A[(int64_t)r * n + q] * x[q];
            x[r] = v;
        }
        for (int r = n - 1; r >= 0; --r) {
            double v = x[r];
            const int qmax = min(n - 1, r + bw);
            for (int q = r + 1; q <= qmax; ++q) v -= A[(int64_t)r * n + q] * x[q];
            x[r] = v / A[(int64_t)r * n + r];
        }
        T* out = dense + P.dense_off + (int64_t)m * n;
        for (int r = 0; r < n; ++r) out[r] = (T)x[r];
    }
}

// setBoundaryValues: every Dirichlet wall of one problem takes `value`
__global__ void k_set_walls(LevelProb* probs, int pi, double value)
{
    if (threadIdx.x != 0) return;
    LevelProb& P = probs[pi];
    for (int k = 0; k < 4; ++k)
        if ((P.dmask >> k) & 1) P.g[k] = value;
}

// the wells after a step (called by eq_batch_step, after the channels)
int wells_after_step(eq_batch* b)
{
    if (b->wells.empty()) return EQ_OK;
    double nodes = 0.0; // the boundary rows and columns each well's flux reads
    for (const WellDev& W : b->wells) {
        const LevelProb& P = b->levels[0].probs[W.prob];
        nodes += 4.0 * (P.nx + P.ny);
    }
    launch(b, EQ_KERNEL_BOUNDARY, nodes * sizeof(double), k_wells, dim3((int)b->wells.size()), dim3(256), 0, b->stream,
           b->levels[0].d_probs, b->d_wells, (const double*)b->u);
    EQ_HIP(hipGetLastError());
    return EQ_OK;
}

int pull_well_values(eq_batch* b)
{
    if (b->wells.empty()) return EQ_OK;
    Level& L = b->levels[0];
    std::vector<LevelProb> dev(L.probs.size());
    EQ_HIP(hipStreamSynchronize(b->stream));
    EQ_HIP(hipMemcpy(dev.data(), L.d_probs, sizeof(LevelProb) * dev.size(), hipMemcpyDeviceToHost));
    EQ_HIP(hipMemcpy(b->wells.data(), b->d_wells, sizeof(WellDev) * b->wells.size(), hipMemcpyDeviceToHost));
    for (const WellDev& W : b->wells)
        for (int k = 0; k < 4; ++k) L.probs[W.prob].g[k] = dev[W.prob].g[k];
    return EQ_OK;
}

// ------------------------------------------------------------ profiling

// launch on the standard tile block: eqk::launch with kBlock threads
template <typename K, typename... Args>
static void launch_tiles(eq_batch* b, int kid, double bytes, K kernel, int grid, hipStream_t s, Args... args)
{
    launch(b, kid, bytes, kernel, dim3(grid), dim3(kBlock), 0, s, args...);
}

static void prof_collect(eq_batch* b)
{
    for (int k = 0; k < EQ_KERNEL_COUNT_; ++k) {
        ProfileSlot& s = b->prof[k];
        for (size_t e = 0; e < s.pending.size(); ++e) {
            float ms = 0.0f;
            hipEventSynchronize(s.pending[e].second);
            hipEventElapsedTime(&ms, s.pending[e].first, s.pending[e].second);
            s.total_ms += ms;
            s.launches += 1;
            s.bytes += s.pending_bytes[e];
            s.pool.push_back(s.pending[e]);
        }
        s.pending.clear();
        s.pending_bytes.clear();
    }
}

// ------------------------------------------------------------ hierarchy set-up

// ------------------------------------------------------------ hierarchy set-up

static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

static void fill_level_prob(LevelProb& P, const eq_field_desc& d, int nx, int ny, bool level0)
{
    std::memset(&P, 0, sizeof(P));
    P.nx = nx;
    P.ny = ny;
    P.pitch = (int32_t)round_up(nx + 2, kRowPad);
    const double hx = d.width / nx, hy = d.height / ny;
    P.hx = hx;
    P.hy = hy;
    P.Ddt = d.dt * d.D;
    const double area6 = hx * hy / 12.0;
    P.mh = (1.0 + d.dt * d.gamma) * area6;
    P.mo = (1.0 + d.dt * d.gamma) * (hx * hy / 24.0);
    P.sa = d.dt * d.D * hy / (2.0 * hx);
    P.sb = d.dt * d.D * hx / (2.0 * hy);
    P.sxy = d.dt * d.D / 2.0;
    P.cw = d.dt * hy / 24.0;
    if (d.scheme == EQ_SCHEME_FD) {
        // diffusionPETSc rows times hx*hy (the wall/corner halving is per node)
        P.fd = 1;
        P.fdm = (1.0 + d.dt * d.gamma) * hx * hy;
        P.fkx = d.dt * d.D * hy / hx;
        P.fky = d.dt * d.D * hx / hy;
        const double seg[4] = {hy, hy, hx, hx}; // the wall-normal ghost term times hx*hy
        for (int w = 0; w < 4; ++w)
            if (d.wall[w].kind == EQ_BC_ROBIN) {
                P.frob[w] = 2.0 * d.dt * d.wall[w].r * seg[w];
                P.fsrc[w] = 2.0 * d.dt * d.wall[w].r * d.wall[w].s * seg[w];
            }
        P.fsrcf = d.dt * d.f * hx * hy;
    }
    const double elen[4] = {hy, hy, hx, hx};
    P.dmask = 0;
    for (int w = 0; w < 4; ++w) {
        if (d.wall[w].kind == EQ_BC_DIRICHLET) P.dmask |= (1 << w);
        if (d.wall[w].kind == EQ_BC_ROBIN) {
            P.rob[w] = d.dt * d.wall[w].r * elen[w];
            if (level0) P.robs[w] = d.dt * d.wall[w].r * d.wall[w].s * elen[w] * 0.5;
        }
        P.g[w] = (d.wall[w].kind == EQ_BC_DIRICHLET) ? d.wall[w].value : 0.0;
    }
    if (level0) {
        P.mh0 = P.fd ? hx * hy : area6; // FD: the scaled identity
        P.mo0 = hx * hy / 24.0;
        P.src = d.dt * d.f * hx * hy / 6.0;
    }
    P.cidx = -1;
    P.fidx = -1;
    set_interior_row(P);
}

// Inverse of the eliminated operator on a coarsest level, column by column
// (inv[m*n + k] = (A^-1)_{k,m}).  A is SPD (Dirichlet rows and columns are
// identity) and banded (half-bandwidth nx+2 in natural order), so LU without
// pivoting inside the band is stable and costs O(n bw^2), and the n
// solves O(n^2 bw), instead of O(n^3) for a dense elimination.
static bool dense_inverse(const LevelProb& P, const double* h11, const double* h22, const double* h12,
                          const double* hv, std::vector<double>& inv)
{
    const int nw = P.nx + 1, n = nw * (P.ny + 1);
    const int bw = nw + 1;
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i <= P.ny; ++i)
        for (int j = 0; j <= P.nx; ++j) {
            const int k = j + i * nw;
            if (is_dirichlet(P, i, j)) {
                A[(size_t)k * n + k] = 1.0;
                continue;
            }
            // h11/h22/h12: the level's tensor in natural order (NULL: constant coefficients)
            Row<double> r = row_coefs_at<double>(P, i, j);
            if (h11) aniso_stiffness<double>(P, i, j, h11 + k, h22 + k, h12 + k, nw, r);
            if (hv) add_convection<double>(P, i, j, hv, r); // hv: the level's row velocities
            auto put = [&](int ii, int jj, double v) {
                if (ii < 0 || jj < 0 || ii > P.ny || jj > P.nx || v == 0.0) return;
                if (is_dirichlet(P, ii, jj)) return;
                A[(size_t)k * n + (jj + ii * nw)] += v;
            };
            A[(size_t)k * n + k] += row_diag(r);
            put(i, j - 1, r.mw - r.kw);
            put(i, j + 1, r.me - r.ke);
            put(i - 1, j, r.ms - r.ks);
            put(i + 1, j, r.mn - r.kn);
            put(i - 1, j - 1, r.msw - r.ksw);
            put(i + 1, j + 1, r.mne - r.kne);
        }
    // in-place banded LU (multipliers below the diagonal)
    for (int c = 0; c < n; ++c) {
        const double d = A[(size_t)c * n + c];
        if (!(d > 0.0)) return false;
        const int rmax = std::min(n - 1, c + bw);
        for (int r = c + 1; r <= rmax; ++r) {
            const double f = A[(size_t)r * n + c] / d;
            if (f == 0.0) continue;
            A[(size_t)r * n + c] = f;
            for (int q = c + 1; q <= rmax; ++q) A[(size_t)r * n + q] -= f * A[(size_t)c * n + q];
        }
    }
    inv.assign((size_t)n * n, 0.0);
    std::vector<double> x(n);
    for (int m = 0; m < n; ++m) {
        std::fill(x.begin(), x.end(), 0.0);
        x[m] = 1.0;
        for (int r = m + 1; r < n; ++r) { // L y = e_m (y_r = 0 for r < m)
            double v = x[r];
            for (int q = std::max(m, r - bw); q < r; ++q) v -= A[(size_t)r * n + q] * x[q];
            x[r] = v;
        }
        for (int r = n - 1; r >= 0; --r) { // U x = y
            double v = x[r];
            const int qmax = std::min(n - 1, r + bw);
            for (int q = r + 1; q <= qmax; ++q) v -= A[(size_t)r * n + q] * x[q];
            x[r] = v / A[(size_t)r * n + r];
        }
        std::copy(x.begin(), x.end(), inv.begin() + (size_t)m * n);
    }
    return true;
}

template <typename T>
static int alloc_zero(T** p, int64_t n, hipStream_t s)
{
    if (n <= 0) n = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
    if (e != hipSuccess) return set_error(EQ_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    e = hipMemsetAsync(*p, 0, sizeof(T) * (size_t)n, s);
    if (e != hipSuccess) return hip_check(e, "hipMemsetAsync");
    return EQ_OK;
}

template <typename V>
static int upload_vec(V** d, const std::vector<V>& h, hipStream_t s)
{
    EQ_HIP(hipMalloc((void**)d, sizeof(V) * std::max<size_t>(1, h.size())));
    if (!h.empty()) EQ_HIP(hipMemcpyAsync(*d, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice, s));
    return EQ_OK;
}

static int validate(const eq_field_desc& d, std::string& why)
{
    if (d.nx < 1 || d.ny < 1) { why = "nx and ny must be >= 1"; return 1; }
    if (d.nx > 65536 || d.ny > 65536 || (int64_t)(d.nx + 1) * (d.ny + 1) > (int64_t)1 << 30) {
        why = "grid too large (nx, ny <= 65536, nodes <= 2^30)";
        return 1;
    }
    if (!(d.width > 0) || !(d.height > 0)) { why = "width and height must be > 0"; return 1; }
    if (d.scheme != EQ_SCHEME_P1 && d.scheme != EQ_SCHEME_FD) { why = "unknown scheme"; return 1; }
    if (d.marking != EQ_WALLS_GRID && d.marking != EQ_WALLS_COORD) { why = "unknown wall marking"; return 1; }
    if (!(d.dt > 0) || !(d.D >= 0) || !(d.gamma >= 0) || !std::isfinite(d.f)) {
        why = "need dt > 0, D >= 0, gamma >= 0, finite f";
        return 1;
    }
    for (int w = 0; w < 4; ++w) {
        const int k = d.wall[w].kind;
        if (k < EQ_BC_NEUMANN || k > EQ_BC_ROBIN) { why = "unknown boundary kind"; return 1; }
        if (k == EQ_BC_ROBIN && !(d.wall[w].r >= 0)) { why = "Robin rate must be >= 0"; return 1; }
    }
    return 0;
}

// The reference's wall test on DOLFIN's vertex coordinates: near(x_j, W) with
// x_j = 0 + (j*(W - 0))/nx, likewise in y (RectangleMesh [ext], src/fHSL.cpp:169-172;
// DirichletBoundary_TrapEdge::inside, src/fHSL.h:575-590).  The left and bottom
// walls sit at exact zeros.  Every facet of a wall has the wall's coordinate at
// both ends, so a wall is found whole or not at all.
static void walls_found(const eq_field_desc& d, int32_t found[4])
{
    const double kDolfinEps = 3.0e-16;
    const double xr = 0.0 + ((double)d.nx * (d.width - 0.0)) / (double)d.nx;
    const double yt = 0.0 + ((double)d.ny * (d.height - 0.0)) / (double)d.ny;
    found[EQ_WALL_LEFT] = 1;
    found[EQ_WALL_BOTTOM] = 1;
    found[EQ_WALL_RIGHT] = std::fabs(xr - d.width) < kDolfinEps ? 1 : 0;
    found[EQ_WALL_TOP] = std::fabs(yt - d.height) < kDolfinEps ? 1 : 0;
}

// Each bottom problem's chain of level records (its bottom level, then every
// coarser level) copied out of the level tables, so that the bottom V-cycle
// reads them in one batch of loads.  Built with the hierarchy and rebuilt,
// and uploaded again when upload is set, after a level record changes.
static int sync_bottom_recs(eq_batch* b, bool upload)
{
    for (size_t l = 0; l < b->levels.size(); ++l) {
        Level& L = b->levels[l];
        L.bottom_refs.clear();
        L.bottom_recs.clear();
        for (int32_t k : L.bottom) {
            const LevelProb& P0 = L.probs[k];
            BottomRef r{};
            r.rec = (int32_t)L.bottom_recs.size();
            r.prob = P0.prob;
            r.nx = P0.nx;
            r.ny = P0.ny;
            r.pitch = P0.pitch;
            r.lds_total = P0.lds_total;
            r.origin = P0.origin;
            size_t lev = l;
            int idx = k;
            L.bottom_recs.push_back(P0);
            while (!b->levels[lev].probs[idx].coarsest) {
                idx = b->levels[lev].probs[idx].cidx;
                ++lev;
                L.bottom_recs.push_back(b->levels[lev].probs[idx]);
                ++r.nl;
            }
            const LevelProb& Pz = b->levels[lev].probs[idx];
            r.nz = (Pz.nx + 1) * (Pz.ny + 1);
            r.dense_off = Pz.dense_off;
            L.bottom_refs.push_back(r);
        }
        if (upload && !L.bottom_recs.empty())
            EQ_HIP(hipMemcpyAsync(L.d_bottom_recs, L.bottom_recs.data(), sizeof(LevelProb) * L.bottom_recs.size(),
                                  hipMemcpyHostToDevice, b->stream));
    }
    if (upload) EQ_HIP(hipStreamSynchronize(b->stream));
    return EQ_OK;
}

// Cells of the coarse grid over n fine cells when the grid cannot be halved
// exactly: about n/2, rounded up to a multiple of the largest power of two
// M <= (n/2)/8 (at most 1/8 more cells), so the levels under it halve exactly
// log2(M) more times and keep the cheap nested transfers.
static int coarse_cells(int n)
{
    if (n <= 2) return n;
    const int h = (n + 1) / 2;
    int m = 1;
    while (2 * m * 8 <= h) m *= 2;
    return (h + m - 1) / m * m;
}

// coarse nodes [Ia, Ib) owned by the fine nodes [i0, i0 + n) -- those whose
// fine position floor(I nf / nc) lies there -- and the fine nodes [lo, hi]
// of their restriction support (restrict_gen's ranges)
static void fused_owned(int i0, int n, int nf, int nc, int32_t& Ia, int32_t& Ib, int32_t& lo, int32_t& hi)
{
    const int64_t f = nf, c = nc;
    Ia = (int32_t)(((int64_t)i0 * c + f - 1) / f);
    Ib = (int32_t)std::min<int64_t>(c + 1, ((int64_t)(i0 + n) * c + f - 1) / f);
    lo = Ia == 0 ? 0 : (int32_t)(((int64_t)(Ia - 1) * f) / c) + 1;
    hi = (int32_t)std::min<int64_t>(f, ((int64_t)Ib * f - 1) / c);
}

template <typename T>
static int build_hierarchy(eq_batch* b)
{
    const eq_solver_desc& S = b->solver;
    const bool mg = S.kind != EQ_SOLVER_JACOBI_CG;
    const int maxlev = mg ? std::min(kMaxLev, std::max(1, S.mg_max_levels)) : 1;
    // sizes per problem per level; bottom level = first level at or below kBottomNodes
    std::vector<std::vector<std::pair<int, int>>> dims(b->nprob);
    std::vector<int> lbot(b->nprob, 0);
    int nlev = 1;
    for (int p = 0; p < b->nprob; ++p) {
        int nx = b->descs[p].nx, ny = b->descs[p].ny;
        dims[p].push_back({nx, ny});
        while (mg && (int)dims[p].size() < maxlev &&
               (int64_t)(nx + 1) * (ny + 1) > (int64_t)S.mg_coarse_max_nodes) {
            int cx, cy;
            if (nx > 2 && ny > 2 && nx % 2 == 0 && ny % 2 == 0) {
                // nested: every coarse node is a fine node
                cx = nx / 2;
                cy = ny / 2;
            } else {
                cx = coarse_cells(nx);
                cy = coarse_cells(ny);
            }
            // every transfer relies on n_f <= 2 n_c per axis (at most four
            // fine nodes per axis in a coarse hat's support)
            if (cx == nx && cy == ny) break;
            if (2 * cx < nx || 2 * cy < ny) return set_error(EQ_ERR_INVALID, "internal: coarsening ratio above 2");
            nx = cx;
            ny = cy;
            dims[p].push_back({nx, ny});
        }
        if (mg && (int64_t)(nx + 1) * (ny + 1) > 1024)
            return set_error(EQ_ERR_INVALID, "coarsest multigrid level exceeds 1024 nodes; raise mg_max_levels");
        if (mg) {
            // bottom level: the first level from which every coarser level's
            // three LDS arrays (x0, x1, b with a zero frame) and its level
            // record fit in the LDS budget
            // (a coarsest level too large for the batch's cap takes the device's)
            auto level_foot = [&](int k) {
                return 3 * (int64_t)(dims[p][k].first + 3) * (dims[p][k].second + 3) * (int64_t)sizeof(T) +
                       (int64_t)sizeof(LevelProb);
            };
            const int nlp = (int)dims[p].size();
            const int64_t budget = 16 + level_foot(nlp - 1) > b->lds_budget ? b->lds_budget_max : b->lds_budget;
            int64_t foot = 16; // alignment of the records after the arrays
            int l = nlp;
            while (l > 0) {
                const int64_t f = level_foot(l - 1);
                if (foot + f > budget) break;
                foot += f;
                --l;
            }
            if (l == (int)dims[p].size())
                return set_error(EQ_ERR_INVALID, "coarsest multigrid level does not fit the bottom solver's LDS");
            lbot[p] = l;
        }
        nlev = std::max(nlev, (int)dims[p].size());
    }
    b->levels.assign(nlev, Level());
    for (int l = 0; l < nlev; ++l) {
        double nodes = 0.0;
        for (int p = 0; p < b->nprob; ++p)
            if ((int)dims[p].size() > l) nodes += (double)(dims[p][l].first + 1) * (double)(dims[p][l].second + 1);
        const double min4 = S.tile4_min_nodes > 0 ? (double)S.tile4_min_nodes : kWideRowsMinNodes;
        const double min8 = S.tile8_min_nodes > 0 ? (double)S.tile8_min_nodes : kWideRows8MinNodes;
        b->levels[l].R = nodes >= min8 ? kRowsWide : (l == 0 || nodes >= min4) ? kRows0 : kRowsC;
    }
    std::vector<std::vector<int>> where(b->nprob, std::vector<int>(nlev, -1));
    for (int l = 0; l < nlev; ++l) {
        Level& L = b->levels[l];
        int64_t off = 0;
        int32_t tile = 0;
        for (int p = 0; p < b->nprob; ++p) {
            if ((int)dims[p].size() <= l) continue;
            LevelProb P;
            fill_level_prob(P, b->descs[p], dims[p][l].first, dims[p][l].second, l == 0);
            P.prob = p;
            P.coarsest = ((int)dims[p].size() == l + 1) ? 1 : 0;
            P.bottom = (mg && l == lbot[p]) ? 1 : 0;
            P.below = (mg && l > lbot[p]) ? 1 : 0;
            if (mg && l >= lbot[p]) {
                int64_t off = 0, total = 0;
                for (int k = lbot[p]; k < (int)dims[p].size(); ++k) {
                    const int64_t sz = 3 * (int64_t)(dims[p][k].first + 3) * (dims[p][k].second + 3);
                    if (k < l) off += sz;
                    total += sz;
                }
                P.lds_off = (int32_t)off;
                P.lds_total = (int32_t)total;
            }
            if (!P.coarsest) {
                P.nxc = dims[p][l + 1].first;
                P.nyc = dims[p][l + 1].second;
            }
            const int64_t rows = P.ny + 3;
            P.origin = off + kRowPad + P.pitch;
            off += round_up(kRowPad + rows * P.pitch + kRowPad, kRowPad);
            P.tiles_x = (P.nx + 1 + kTileW - 1) / kTileW;
            const int th = kWaves * L.R; // tile height on this level
            P.tiles_y = (P.ny + 1 + th - 1) / th;
            P.ntiles = P.tiles_x * P.tiles_y;
            P.tile_begin = tile;
            tile += P.ntiles;
            where[p][l] = (int)L.probs.size();
            if (l > 0) P.fidx = where[p][l - 1];
            L.probs.push_back(P);
        }
        L.pool = off;
    }
    // cross-level links, tile tables, dense inverses
    std::vector<T> dense_all;
    for (int l = 0; l < nlev; ++l) {
        Level& L = b->levels[l];
        for (size_t k = 0; k < L.probs.size(); ++k) {
            LevelProb& P = L.probs[k];
            if (!P.coarsest) P.cidx = where[P.prob][l + 1];
            const bool smooth = mg && !P.bottom && !P.below;
            const bool active = !P.below;
            const int th = kWaves * L.R; // tile height on this level
            for (int t = 0; t < P.ntiles; ++t) {
                TileRef r{};
                r.pi = (int32_t)k;
                r.prob = P.prob;
                r.local = t;
                r.gt = P.tile_begin + t;
                const int ty = t / P.tiles_x, tx = t - ty * P.tiles_x;
                r.i0 = ty * th;
                r.j0 = tx * kTileW;
                r.pitch = P.pitch;
                r.nx = P.nx;
                r.ny = P.ny;
                r.base = P.origin + (int64_t)r.i0 * P.pitch + r.j0;
                // rows i0 .. i0+th-1 and columns j0 .. j0+kTileW-1 all in 1 .. n-1
                const bool full = r.i0 >= 1 && r.j0 >= 1 && r.i0 + th <= P.ny && r.j0 + kTileW <= P.nx;
                const bool full2 = r.i0 >= 2 && r.j0 >= 2 && r.i0 + th < P.ny && r.j0 + kTileW < P.nx;
                r.flags = (full ? 1 : 0) | (full2 ? 2 : 0);
                L.tiles_all.push_back(r);
                if (smooth) L.tiles_smooth.push_back(r);
                if (active && P.bottom) L.tiles_bottom.push_back(r);
                if (active) L.tiles_active.push_back(r);
            }
            if (smooth && !P.coarsest) {
                // the fused coarse-level kernels' tiles
                const LevelProb& Pc = b->levels[l + 1].probs[P.cidx];
                const int fx = (P.nx + 1 + kFusedTileW - 1) / kFusedTileW;
                const int th = l == 0 ? kFusedTileH0 : kFusedTileHc;
                const int fy = (P.ny + 1 + th - 1) / th;
                for (int t = 0; t < fx * fy; ++t) {
                    FusedTile f{};
                    f.pi = (int32_t)k;
                    f.prob = P.prob;
                    f.gt = -1;
                    f.i0 = (t / fx) * th;
                    f.j0 = (t % fx) * kFusedTileW;
                    fused_owned(f.i0, th, P.ny, Pc.ny, f.Ia, f.Ib, f.ilo, f.ihi);
                    fused_owned(f.j0, kFusedTileW, P.nx, Pc.nx, f.Ja, f.Jb, f.jlo, f.jhi);
                    if (f.ilo - 2 >= 1 && f.ihi + 2 <= P.ny - 1 && f.jlo - 2 >= 1 && f.jhi + 2 <= P.nx - 1) f.inner |= 1;
                    if (f.i0 - 2 >= 1 && f.i0 + th + 1 <= P.ny - 1 && f.j0 - 2 >= 1 &&
                        f.j0 + kFusedTileW + 1 <= P.nx - 1)
                        f.inner |= 2;
                    f.nx = P.nx;
                    f.ny = P.ny;
                    f.pitch = P.pitch;
                    f.dmask = P.dmask;
                    f.cnx = Pc.nx;
                    f.cny = Pc.ny;
                    f.cpitch = Pc.pitch;
                    f.cdmask = Pc.dmask;
                    f.rnxf = P.rnxf;
                    f.rnyf = P.rnyf;
                    f.crnxf = Pc.rnxf;
                    f.crnyf = Pc.rnyf;
                    f.rnx = P.rnx;
                    f.rny = P.rny;
                    f.origin = P.origin;
                    f.corigin = Pc.origin;
                    L.fused.push_back(f);
                }
            }
            const double nodes = (double)(P.nx + 1) * (double)(P.ny + 1);
            L.nodes_all += nodes;
            if (smooth) L.nodes_smooth += nodes;
            if (active) L.nodes_active += nodes;
            if (active && P.bottom) L.nodes_bottom += nodes;
            if (smooth && !(P.nx == 2 * P.nxc && P.ny == 2 * P.nyc)) L.nested_down = false;
            if (P.bottom) L.bottom.push_back((int32_t)k);
            if (P.coarsest && mg) {
                std::vector<double> inv;
                if (!dense_inverse(P, nullptr, nullptr, nullptr, nullptr, inv))
                    return set_error(EQ_ERR_INVALID, "singular coarse operator");
                P.dense_off = (int64_t)dense_all.size();
                for (double v : inv) dense_all.push_back((T)v); // column-major already
            }
        }
    }
    // level 0: the fused tiles' partial sums follow the level's own tiles',
    // and the finish of r.z after a fused V-cycle reads them there
    {
        Level& L0 = b->levels[0];
        L0.probs_fin = L0.probs;
        const int32_t base = (int32_t)L0.tiles_all.size();
        for (size_t k = 0; k < L0.fused.size(); ++k) {
            FusedTile& f = L0.fused[k];
            f.gt = base + (int32_t)k;
            LevelProb& Q = L0.probs_fin[f.pi];
            if (Q.tile_begin < base) {
                Q.tile_begin = f.gt;
                Q.ntiles = 0;
            }
            Q.ntiles += 1;
        }
    }
    // bottom chains; dynamic LDS = the level arrays, then the level records
    sync_bottom_recs(b, false);
    for (Level& L : b->levels)
        for (const BottomRef& r : L.bottom_refs) {
            const LevelProb& Pc = L.bottom_recs[r.rec + r.nl];
            const double nc = (double)(Pc.nx + 1) * (double)(Pc.ny + 1);
            L.bottom_dense_bytes += nc * nc * sizeof(T);
        }
    for (Level& L : b->levels)
        for (const BottomRef& r : L.bottom_refs)
            L.bottom_lds_bytes = std::max<size_t>(
                L.bottom_lds_bytes, (size_t)(bottom_rec_offset<T>(L.bottom_recs[r.rec]) + (r.nl + 1) * (int64_t)sizeof(LevelProb)));
    // device tables and vectors
    hipStream_t s = b->stream;
    int rc;
    for (int l = 0; l < nlev; ++l) {
        Level& L = b->levels[l];
        if ((rc = upload_vec(&L.d_probs, L.probs, s)) || (rc = upload_vec(&L.d_tiles_all, L.tiles_all, s)) ||
            (rc = upload_vec(&L.d_tiles_smooth, L.tiles_smooth, s)) ||
            (rc = upload_vec(&L.d_tiles_bottom, L.tiles_bottom, s)) ||
            (rc = upload_vec(&L.d_tiles_active, L.tiles_active, s)) || (rc = upload_vec(&L.d_fused, L.fused, s)) ||
            (l == 0 && (rc = upload_vec(&L.d_probs_fin, L.probs_fin, s))) ||
            (rc = upload_vec(&L.d_bottom, L.bottom_refs, s)) ||
            (rc = upload_vec(&L.d_bottom_recs, L.bottom_recs, s)))
            return rc;
        if (mg) {
            if ((rc = alloc_zero((T**)&L.b, L.pool, s)) || (rc = alloc_zero((T**)&L.xa, L.pool, s)) ||
                (rc = alloc_zero((T**)&L.xb, L.pool, s)) || (rc = alloc_zero((T**)&L.r, L.pool, s)))
                return rc;
        }
    }
    if (!dense_all.empty()) {
        if ((rc = alloc_zero((T**)&b->dense, (int64_t)dense_all.size(), s))) return rc;
        EQ_HIP(hipMemcpyAsync(b->dense, dense_all.data(), sizeof(T) * dense_all.size(), hipMemcpyHostToDevice, s));
    }
    size_t lds_max = 0;
    for (const Level& L : b->levels) lds_max = std::max(lds_max, L.bottom_lds_bytes);
    if (lds_max > 64 * 1024) {
        EQ_HIP(hipFuncSetAttribute((const void*)k_bottom<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds_max));
        EQ_HIP(hipFuncSetAttribute((const void*)k_bottom<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds_max));
    }
    const int64_t n0 = b->levels[0].pool;
    const bool bicg = S.kind == EQ_SOLVER_GMG_BICGSTAB || S.kind == EQ_SOLVER_MIXED_GMG_BICGSTAB;
    if ((rc = alloc_zero(&b->u, n0, s)) || (rc = alloc_zero(&b->rhs, n0, s)) || (rc = alloc_zero(&b->r, n0, s)) ||
        (rc = alloc_zero(&b->p, n0, s)) || (rc = alloc_zero(&b->p2, n0, s)) || (rc = alloc_zero(&b->q, n0, s)) ||
        (rc = alloc_zero(&b->z, n0, s)) || (rc = alloc_zero(&b->xprev, n0, s)) ||
        (S.extrapolate && (rc = alloc_zero(&b->xprev2, n0, s))) ||
        (bicg && ((rc = alloc_zero(&b->bs_rhat, n0, s)) || (rc = alloc_zero(&b->bs_p, n0, s)) ||
                  (rc = alloc_zero(&b->bs_v, n0, s)) || (rc = alloc_zero(&b->bs_s, n0, s)) ||
                  (rc = alloc_zero(&b->bs_t, n0, s)))) ||
        (rc = alloc_zero(&b->partial, 4 * (int64_t)(b->levels[0].tiles_all.size() + b->levels[0].fused.size()), s)))
        return rc;
    EQ_HIP(hipStreamSynchronize(s));
    return EQ_OK;
}

// ------------------------------------------------------------ V-cycle

template <typename T, bool TENS, typename TI = double>
struct VCycle {
    eq_batch* b;
    const TI* in = nullptr;     // level-0 right-hand side: the vector to precondition (FP64, or FP32 when
                                // it is already in the smoother's precision)
    T* out = nullptr;           // level-0 result
    bool pre1 = false;          // level 0's first sweep from zero is already in xa (and its b in L.b)
    bool fuse_top = false;      // level 0 may run fused (then it forms its first sweep from `in` itself)
    const GroupView* grp = nullptr; // a problem group: its fused tiles and bottom workgroups only (every level fused)
    hipStream_t strm = nullptr;     // the stream the launches go to (default: the batch's)
    const CgState* st_ = nullptr; // Krylov state whose `done` the launches test (default: the batch's)
    // an input in the smoother's precision is read directly; an FP64 one is
    // copied to L.b by the first sweep when the smoother runs in FP32
    static constexpr bool kDirect = std::is_same<TI, T>::value;

    explicit VCycle(eq_batch* bb) : b(bb), in(initial_in(bb)), out((T*)bb->z) {}
    static const TI* initial_in(eq_batch* bb)
    {
        if constexpr (std::is_same<TI, double>::value) return bb->r;
        else return nullptr;
    }

    // z = V(r) on level 0 (and the partial r.z for the next k_finish)
    void run0() { run(0); }

    // Level l: its right-hand side is in L.b (level 0: the FP64 residual
    // b->r); on return the level's correction is in final_buffer(l)
    // (level 0: b->z).
    T* final_buffer(int l) const
    {
        const eq_solver_desc& S = b->solver;
        const Level& L = b->levels[l];
        if (fused_level(l)) return (T*)L.xa;
        const int f = ((S.mg_pre - 1) + S.mg_post) & 1;
        return f ? (T*)L.xb : (T*)L.xa;
    }

    // A level goes down in one launch (k_level_down: pre-sweeps into xb, the
    // restricted residual into the next level's b) and up in another
    // (k_level_up: xb plus the correction, post-sweeps into xa; level 0: into
    // `out`, with the partial r.z in the fused tiles' slots).  Level 0 does
    // so when its solver allows it (fuse_top), reading `in` itself.
    bool fused_level(int l) const
    {
        const eq_solver_desc& S = b->solver;
        if (TENS || S.mg_level_fused == 0 || S.mg_pre != 2 || S.mg_post != 2 || b->levels[l].fused.empty()) return false;
        return l > 0 || fuse_top;
    }
    // the level table whose tile ranges locate the partial r.z of the last run0()
    const LevelProb* rz_table() const
    {
        const Level& L0 = b->levels[0];
        return (fused_level(0) ? L0.d_probs_fin : L0.d_probs) + (grp ? grp->pi0 : 0);
    }

    void run(int l)
    {
        const int R = b->levels[l].R;
        if (R == kRowsWide)
            run_level<kRowsWide>(l);
        else if (R == kRows0)
            run_level<kRows0>(l);
        else
            run_level<kRowsC>(l);
    }

    template <int R>
    void run_level(int l)
    {
        const eq_solver_desc& S = b->solver;
        Level& L = b->levels[l];
        // pre-sweep k is damped by mg_omega (k even) or mg_omega2 (k odd);
        // post-sweep k by pre-sweep (post - 1 - k)'s damping
        const T om1 = (T)S.mg_omega, om2 = (T)(S.mg_omega2 > 0 ? S.mg_omega2 : S.mg_omega);
        auto wpre = [&](int k) { return (k & 1) ? om2 : om1; };
        auto wpost = [&](int k) { return wpre(S.mg_post - 1 - k); };
        const T omega = om1;
        hipStream_t s = strm ? strm : b->stream;
        const CgState* st = st_ ? st_ : b->d_state;
        const bool top = (l == 0);
        const int ns = (int)L.tiles_smooth.size();
        const int gs = grid_for(ns);
        // the fused tiles this cycle launches (a group's, or all), and the
        // nodes its launches count: a group's, or the level's
        const FusedTile* ft = (const FusedTile*)L.d_fused + (grp ? grp->f0[l] : 0);
        const int nf = grp ? grp->nf[l] : (int)L.fused.size(), gf = grid_for(nf);
        const double lv_smooth = grp ? grp->nsm[l] : L.nodes_smooth;
        const double lv_bottom = grp ? grp->nbot[l] : L.nodes_bottom;
        const double lv_dense = grp ? grp->dense[l] : L.bottom_dense_bytes;
        auto next_active = [&]() { return grp ? grp->nact[l + 1] : b->levels[l + 1].nodes_active; };
        T* X[2] = {(T*)L.xa, (T*)L.xb};
        int cur = 0;
        if (ns > 0 && top && fused_level(0)) {
            Level& C = b->levels[1];
            const double nodes = lv_smooth, cnodes = next_active();
            // b: the Krylov residual `in`, rounded to T as it is read (the
            // first sweep from zero formed from it here)
            launch_tiles(b, EQ_KERNEL_TRANSFER0, nodes * (sizeof(TI) + sizeof(T)) + cnodes * sizeof(T),
                         k_level_down<T, kFusedTileH0, false, TI>, gf, s, ft, nf,
                         (const LevelProb*)L.d_probs, in, (T*)L.xb, (T*)C.b, wpre(0), wpre(1), st, (const T*)nullptr);
            run(1);
            launch_tiles(b, EQ_KERNEL_SMOOTH, nodes * (2 * sizeof(T) + sizeof(TI)) + cnodes * sizeof(T),
                         k_level_up<T, TI, kFusedTileH0, true, TI>, gf, s, ft, nf,
                         (const LevelProb*)L.d_probs, (const T*)final_buffer(1), in, (const T*)L.xb, out, wpost(0),
                         wpost(1), st, in, b->partial);
        } else if (ns > 0 && fused_level(l)) {
            // (a group with no problem smoothed here has none below either;
            // its problems whose bottom is this level follow)
            if (nf > 0) {
                Level& C = b->levels[l + 1];
                const double nodes = lv_smooth, cnodes = next_active();
                // b read, x written; the coarse b written
                launch_tiles(b, EQ_KERNEL_TRANSFER_COARSE, nodes * 2 * sizeof(T) + cnodes * sizeof(T),
                             k_level_down<T, kFusedTileHc, false>, gf, s, ft, nf, (const LevelProb*)L.d_probs,
                             (const T*)L.b, (T*)L.xb, (T*)C.b, wpre(0), wpre(1), st, (const T*)nullptr);
                run(l + 1);
                // x and b read, x written; the coarse correction read
                launch_tiles(b, EQ_KERNEL_SMOOTH_COARSE, nodes * 3 * sizeof(T) + cnodes * sizeof(T),
                             k_level_up<T, T, kFusedTileHc, false>, gf, s, ft, nf, (const LevelProb*)L.d_probs,
                             (const T*)final_buffer(l + 1), (const T*)L.b, (const T*)L.xb, (T*)L.xa, wpost(0),
                             wpost(1), st, (const T*)nullptr, (double*)nullptr);
            }
        } else if (ns > 0) {
            const double nodes = L.nodes_smooth;
            // pre-smoothing; the first sweep starts from x = 0; level 0 reads
            // the FP64 residual and (mixed precision) keeps an FP32 copy of it
            const T* bl = top ? ((kDirect || !b->mixed) ? (const T*)in : (const T*)L.b) : (const T*)L.b;
            T* bcopy = (top && b->mixed && !kDirect) ? (T*)L.b : nullptr;
            const double bytes_b = top ? sizeof(TI) : sizeof(T);
            // Small levels run at the launch-latency floor: there the fused
            // kernels (two launches' work in one, the first launch's values
            // recomputed at the stencil points) win; large levels are
            // bandwidth- and issue-bound and keep the separate launches.
            const double fmax = S.mg_fuse_max_nodes > 0 ? (double)S.mg_fuse_max_nodes : kFuseMaxNodes;
            // a non-nested transfer recomputes general interpolation weights
            // at every stencil point when fused: it stays in separate launches
            const bool fuse = nodes < fmax && L.nested_down;
            const bool fuse_pre =
                top ? S.mg_fuse_pre_top != 0 : nodes < (S.mg_fuse_pre_max_nodes > 0 ? (double)S.mg_fuse_pre_max_nodes : fmax);
            // pre1: the kernel that formed the level-0 residual also wrote
            // the first sweep into X[0] (and the FP32 copy of the residual)
            const bool first_done = top && pre1;
            const bool pair = !first_done && S.mg_pre >= 2 && fuse_pre;
            // sweeps time as the finest level's or the coarse levels' family
            const int slot = top ? EQ_KERNEL_SMOOTH : EQ_KERNEL_SMOOTH_COARSE;
            const int xfer = top ? EQ_KERNEL_TRANSFER0 : EQ_KERNEL_TRANSFER_COARSE;
            if (!first_done) {
                const double by = nodes * (bytes_b + sizeof(T) * (bcopy ? 2 : 1));
                if (pair) {
                    // sweeps 1 and 2 in one launch
                    if (top)
                        launch_tiles(b, slot, by, k_presmooth2<R, T, TI, TENS>, gs, s, L.smooth(), in, X[1], bcopy,
                                    wpre(0), wpre(1), st);
                    else
                        launch_tiles(b, slot, by, k_presmooth2<R, T, T, TENS>, gs, s, L.smooth(), bl, X[1], (T*)nullptr,
                                    wpre(0), wpre(1), st);
                    cur = 1;
                } else if (top) {
                    launch_tiles(b, slot, by, k_smooth<R, T, TI, T, TENS>, gs, s, L.smooth(), in, (const T*)nullptr, X[0],
                                bcopy, omega, 1, 0, (double*)nullptr, st);
                } else {
                    launch_tiles(b, slot, by, k_smooth<R, T, T, T, TENS>, gs, s, L.smooth(), bl, (const T*)nullptr, X[0],
                                (T*)nullptr, omega, 1, 0, (double*)nullptr, st);
                }
            }
            for (int k = pair ? 2 : 1; k < S.mg_pre; ++k) {
                launch_tiles(b, slot, nodes * 3 * sizeof(T), k_smooth<R, T, T, T, TENS>, gs, s, L.smooth(), bl,
                            (const T*)X[cur], X[cur ^ 1], (T*)nullptr, wpre(k), 0, 0, (double*)nullptr, st);
                cur ^= 1;
            }
            // residual and its restriction over the coarse nodes
            Level& C = b->levels[l + 1];
            const int na = (int)C.tiles_active.size();
            const double cbytes = C.nodes_active * sizeof(T);
            if (fuse) {
                const double by = nodes * 2 * sizeof(T) + cbytes; // b (the FP32 copy on level 0) and x
                if (C.R == kRowsWide)
                    launch_tiles(b, xfer, by, k_residual_restrict<kRowsWide, T, TENS>, grid_for(na), s, C.active(),
                                 (const LevelProb*)L.d_probs, bl, (const T*)X[cur], (T*)C.b, st);
                else if (C.R == kRows0)
                    launch_tiles(b, xfer, by, k_residual_restrict<kRows0, T, TENS>, grid_for(na), s, C.active(),
                                 (const LevelProb*)L.d_probs, bl, (const T*)X[cur], (T*)C.b, st);
                else
                    launch_tiles(b, xfer, by, k_residual_restrict<kRowsC, T, TENS>, grid_for(na), s, C.active(),
                                 (const LevelProb*)L.d_probs, bl, (const T*)X[cur], (T*)C.b, st);
            } else {
                launch_tiles(b, xfer, nodes * 3 * sizeof(T), k_residual<R, T, TENS>, gs, s, L.smooth(), bl,
                             (const T*)X[cur], (T*)L.r, st);
                const double by = nodes * sizeof(T) + cbytes;
                if (C.R == kRowsWide)
                    launch_tiles(b, xfer, by, k_restrict<kRowsWide, T>, grid_for(na), s, C.active(),
                                 (const LevelProb*)L.d_probs, (const T*)L.r, (T*)C.b, st);
                else if (C.R == kRows0)
                    launch_tiles(b, xfer, by, k_restrict<kRows0, T>, grid_for(na), s, C.active(), (const LevelProb*)L.d_probs,
                                 (const T*)L.r, (T*)C.b, st);
                else
                    launch_tiles(b, xfer, by, k_restrict<kRowsC, T>, grid_for(na), s, C.active(), (const LevelProb*)L.d_probs,
                                 (const T*)L.r, (T*)C.b, st);
            }
            run(l + 1);
            const T* xc = final_buffer(l + 1);
            // coarse-grid correction fused with the first post-smoothing sweep;
            // the last level-0 sweep reads the FP64 residual, writes z and the
            // partial r.z
            for (int k = 0; k < S.mg_post; ++k) {
                const bool last = (k == S.mg_post - 1);
                const bool first = (k == 0) && fuse;
                if (k == 0 && !fuse)
                    launch_tiles(b, xfer, nodes * 2 * sizeof(T) + cbytes, k_prolong<R, T>, gs, s, L.smooth(),
                                 (const LevelProb*)C.d_probs, xc, X[cur], st);
                // the last level-0 sweep reads the FP64 residual, the others b in T
                const double by = nodes * (sizeof(T) * 2 + (last ? bytes_b : sizeof(T))) + (first ? cbytes : 0.0);
                if (top && last) {
                    if (first)
                        launch_tiles(b, slot, by, k_prolong_smooth<R, T, TI, T, TENS>, gs, s, L.smooth(), C.d_probs, xc,
                                    in, (const T*)X[cur], out, wpost(k), 1, b->partial, st);
                    else
                        launch_tiles(b, slot, by, k_smooth<R, T, TI, T, TENS>, gs, s, L.smooth(), in, (const T*)X[cur],
                                    out, (T*)nullptr, wpost(k), 0, 1, b->partial, st);
                } else {
                    if (first)
                        launch_tiles(b, slot, by, k_prolong_smooth<R, T, T, T, TENS>, gs, s, L.smooth(), C.d_probs, xc, bl,
                                    (const T*)X[cur], X[cur ^ 1], wpost(k), 0, (double*)nullptr, st);
                    else
                        launch_tiles(b, slot, by, k_smooth<R, T, T, T, TENS>, gs, s, L.smooth(), bl, (const T*)X[cur],
                                    X[cur ^ 1], (T*)nullptr, wpost(k), 0, 0, (double*)nullptr, st);
                    cur ^= 1;
                }
            }
        }
        const int r0 = grp ? grp->r0[l] : 0, nr = grp ? grp->nr[l] : (int)L.bottom.size();
        if (nr > 0) {
            // the bottom solve leaves the level's correction in xl; on level 0
            // it is copied to `out` with the partial (in, out)
            T* xl = final_buffer(l);
            const int nb = (int)L.tiles_bottom.size();
            if (top)
                launch_tiles(b, EQ_KERNEL_BOTTOM, L.nodes_bottom * (sizeof(TI) + sizeof(T)), k_cast_in<R, T, TI>,
                             grid_for(nb), s, L.bottom_tiles(), in, (T*)L.b, st);
            // the bottom levels live in LDS: HBM sees the level's b and x and the coarsest inverses
            launch(b, EQ_KERNEL_BOTTOM, lv_bottom * 2 * sizeof(T) + lv_dense, k_bottom<T, TENS>,
                   dim3(nr), dim3(kBottomThreads), L.bottom_lds_bytes, s,
                   (const BottomRef*)L.d_bottom + r0, (const LevelProb*)L.d_bottom_recs, (const T*)L.b, xl, om1, om2, S.mg_pre,
                   S.mg_post, (const T*)b->dense, st);
            if (top)
                launch_tiles(b, EQ_KERNEL_BOTTOM, L.nodes_bottom * (2 * sizeof(T) + sizeof(TI)), k_cast_out<R, T, T, TI>,
                             grid_for(nb), s, L.bottom_tiles(), (const T*)xl, in, out, b->partial, st);
        }
    }
};

// ------------------------------------------------------------ solve driver

// Captured launch sequences hold the batch's buffers, cycle shape and kernel
// specialisations: a change of any of them drops them all (recaptured on use).
static void drop_graphs(eq_batch* b)
{
    for (auto& gr : b->g_cg)
        for (auto& f : gr)
            for (auto& d : f)
                for (auto& n : d)
                    for (hipGraphExec_t& g : n) {
                        if (g) hipGraphExecDestroy(g);
                        g = nullptr;
                    }
    for (hipGraphExec_t* g : {&b->g_init, &b->g_bicg, &b->g_ir2}) {
        if (*g) hipGraphExecDestroy(*g);
        *g = nullptr;
    }
}

// The CG problem groups: n runs of consecutive problems (group k: problems
// [k P / n, (k+1) P / n)), each with the ranges of every level table its
// launches read.  No groups (fewer than two problems, or a table whose entries
// are not ordered by group) leaves the batch one group.
static void build_groups(eq_batch* b, int n)
{
    b->groups.clear();
    b->groups_for = n;
    const int np = b->nprob;
    n = std::min(n, np);
    if (n < 2) return;
    std::vector<int> gof(np);
    for (int g = 0; g < n; ++g)
        for (int p = (int)((int64_t)g * np / n); p < (int)((int64_t)(g + 1) * np / n); ++p) gof[p] = g;
    auto grp_of = [&](int prob) { return gof[prob]; };
    // a table's runs per group, or false when its entries are not ordered by group
    std::vector<int32_t> bg(n), ct(n);
    auto ranges = [&](int m, auto&& group_at) -> bool {
        std::fill(bg.begin(), bg.end(), 0);
        std::fill(ct.begin(), ct.end(), 0);
        int prev = 0;
        for (int k = 0; k < m; ++k) {
            const int g = group_at(k);
            if (g < prev) return false;
            if (ct[g] == 0) bg[g] = k;
            ++ct[g];
            prev = g;
        }
        return true;
    };
    std::vector<GroupView> G(n);
    for (int g = 0; g < n; ++g) {
        G[g].p0 = (int32_t)((int64_t)g * np / n);
        G[g].p1 = (int32_t)((int64_t)(g + 1) * np / n);
    }
    const Level& L0 = b->levels[0];
    if (!ranges((int)L0.probs.size(), [&](int k) { return grp_of(L0.probs[k].prob); })) return;
    for (int g = 0; g < n; ++g) {
        G[g].pi0 = bg[g];
        G[g].npi = ct[g];
    }
    if (!ranges((int)L0.tiles_all.size(), [&](int k) { return grp_of(L0.tiles_all[k].prob); })) return;
    for (int g = 0; g < n; ++g) {
        G[g].t0 = bg[g];
        G[g].nt = ct[g];
    }
    for (const LevelProb& P : L0.probs) G[grp_of(P.prob)].nodes += (double)(P.nx + 1) * (P.ny + 1);
    for (const Level& L : b->levels) {
        std::vector<double> nsm(n, 0.0), nact(n, 0.0), nbot(n, 0.0);
        for (const LevelProb& P : L.probs) {
            const double v = (double)(P.nx + 1) * (P.ny + 1);
            const int g = grp_of(P.prob);
            if (!P.bottom && !P.below) nsm[g] += v;
            if (!P.below) nact[g] += v;
            if (P.bottom && !P.below) nbot[g] += v;
        }
        for (int g = 0; g < n; ++g) {
            G[g].nsm.push_back(nsm[g]);
            G[g].nact.push_back(nact[g]);
            G[g].nbot.push_back(nbot[g]);
        }
        if (!ranges((int)L.fused.size(), [&](int k) { return grp_of(L.fused[k].prob); })) return;
        for (int g = 0; g < n; ++g) {
            G[g].f0.push_back(bg[g]);
            G[g].nf.push_back(ct[g]);
        }
        if (!ranges((int)L.bottom.size(), [&](int k) { return grp_of(L.probs[L.bottom[k]].prob); })) return;
        for (int g = 0; g < n; ++g) {
            G[g].r0.push_back(bg[g]);
            G[g].nr.push_back(ct[g]);
            G[g].dense.push_back(L.bottom.empty() ? 0.0 : L.bottom_dense_bytes * ct[g] / (double)L.bottom.size());
        }
    }
    for (int g = 0; g < n; ++g)
        if (G[g].npi == 0 || G[g].nt == 0) return;
    b->groups = G;
}

static int read_states(eq_batch* b)
{
    EQ_HIP(hipMemcpyAsync(b->h_state, b->d_state, sizeof(CgState) * b->nprob, hipMemcpyDeviceToHost, b->stream));
    EQ_HIP(hipStreamSynchronize(b->stream));
    return EQ_OK;
}

template <typename T, bool TENS, int R0>
static int solve_all(eq_batch* b, eq_step_stats* stats)
{
    const auto t0 = std::chrono::steady_clock::now();
    const eq_solver_desc& S = b->solver;
    Level& L0 = b->levels[0];
    const bool jacobi = (S.kind == EQ_SOLVER_JACOBI_CG);
    // Launch sequences are replayed from hipGraphs unless per-launch event
    // timing is on (events need eager launches).
    bool profiling = false;
    for (int k = 0; k < EQ_KERNEL_COUNT_; ++k) profiling = profiling || b->prof[k].enabled;
    const bool graphs = S.use_graphs && !profiling;
    // A fused finest level reads the residual and forms the V-cycle's first
    // sweep (x = omega D^-1 r from x = 0) itself.  Otherwise that sweep is
    // pointwise in the new residual: k_update forms it with r, and after
    // k_init or the true residual a sweep launch does; the V-cycle then
    // starts from its second sweep.
    constexpr int kMG = eq_batch::kMaxGroups;
    VCycle<T, TENS> vcs[kMG] = {VCycle<T, TENS>(b), VCycle<T, TENS>(b), VCycle<T, TENS>(b), VCycle<T, TENS>(b)};
    for (auto& v : vcs) {
        v.fuse_top = !jacobi;
        v.pre1 = !jacobi && !v.fused_level(0);
    }
    const bool pre1 = vcs[0].pre1;
    T* x1 = pre1 ? (T*)L0.xa : nullptr;
    T* bcopy = (pre1 && b->mixed) ? (T*)L0.b : nullptr;
    const T omega = (T)S.mg_omega;
    // Problem groups (solver option streams), each on its own stream and
    // stopping at its own iteration count: one group's latency-bound launches
    // (coarse levels, bottom solve, finishes) overlap another's bandwidth-
    // bound ones.  Only where every level runs on the fused kernels, whose
    // launches take a group's tiles; the per-problem arithmetic is the same.
    if (b->groups_for != S.streams) build_groups(b, S.streams);
    bool split = S.streams >= 2 && graphs && !jacobi && !b->groups.empty() && vcs[0].fused_level(0) &&
                 L0.bottom.empty();
    for (size_t l = 0; split && l < b->levels.size(); ++l)
        if (!b->levels[l].tiles_smooth.empty() && !vcs[0].fused_level((int)l)) split = false;
    const int ng = split ? (int)b->groups.size() : 1;
    if (split) {
        if (!b->grp_start) EQ_HIP(hipEventCreateWithFlags(&b->grp_start, hipEventDisableTiming));
        for (int g = 1; g < ng; ++g) {
            if (!b->gstream[g]) EQ_HIP(hipStreamCreateWithFlags(&b->gstream[g], hipStreamNonBlocking));
            if (!b->grp_done[g]) EQ_HIP(hipEventCreateWithFlags(&b->grp_done[g], hipEventDisableTiming));
        }
    }
    auto gstream = [&](int g) { return g ? b->gstream[g] : b->stream; };
    if (split)
        for (int g = 0; g < ng; ++g) {
            vcs[g].grp = &b->groups[g];
            vcs[g].strm = gstream(g);
        }
    // the group the lambdas below launch for (the whole batch when unsplit)
    hipStream_t s = b->stream;
    LevelView A = L0.all();
    int g0 = grid_for(A.ntiles);
    double fine_nodes = L0.nodes_all;
    int fin_n = b->nprob;                  // finish blocks: the group's level-0 records
    const LevelProb* fin_tab = L0.d_probs; // from here
    int gp0 = 0, gp1 = b->nprob;           // the group's problems
    VCycle<T, TENS>* vc = &vcs[0];
    auto use_group = [&](int g) {
        vc = &vcs[g];
        if (ng == 1) return;
        const GroupView& G = b->groups[g];
        s = gstream(g);
        A = LevelView{L0.d_probs, L0.d_tiles_all + G.t0, G.nt};
        g0 = grid_for(A.ntiles);
        fine_nodes = G.nodes;
        fin_n = G.npi;
        fin_tab = L0.d_probs + G.pi0;
        gp0 = G.p0;
        gp1 = G.p1;
    };
    auto fin_bytes = [&]() { return (double)A.ntiles * 3 * sizeof(double) + 2.0 * fin_n * sizeof(CgState); };
    auto finish = [&](int op) {
        launch(b, EQ_KERNEL_FINISH, fin_bytes(), k_finish, dim3(fin_n), dim3(kFinishThreads), 0, s, fin_tab,
               (const double*)b->partial, b->d_state, op, jacobi ? 1 : 0);
    };
    // zero_init 2: for the problems a confirmation restarted only
    auto first_sweep = [&](int zero_init) {
        const int ns = (int)L0.tiles_smooth.size();
        if (!pre1 || ns == 0) return;
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, L0.nodes_smooth * (sizeof(double) + sizeof(T) * (bcopy ? 2 : 1)),
                     k_smooth<R0, T, double, T, TENS>, grid_for(ns), s, L0.smooth(), (const double*)b->r,
                     (const T*)nullptr, x1, bcopy, omega, zero_init, 0, (double*)nullptr, (const CgState*)b->d_state);
    };
    // The finest level's r.z has slots of its own when that level is fused:
    // the update's scalars can then wait for the finish after the next
    // V-cycle (merged), and a chunk of n iterations takes 2n + 1 finishes.
    const bool merge = !jacobi && vcs[0].fused_level(0);
    // one CG iteration from search direction pold to pnew; merged: the
    // previous update's finish is still to do
    auto iteration = [&](double* pold, double* pnew, bool merged) {
        if (!jacobi) {
            vc->run0();
            if (merged)
                launch(b, EQ_KERNEL_FINISH, fin_bytes() + (double)L0.fused.size() * sizeof(double) / ng,
                       k_finish_update_rz, dim3(fin_n), dim3(kFinishThreads), 0, s, vc->rz_table(), fin_tab,
                       (const double*)b->partial, b->d_state);
            else
                launch(b, EQ_KERNEL_FINISH, fin_bytes(), k_finish, dim3(fin_n), dim3(kFinishThreads), 0, s,
                       vc->rz_table(), (const double*)b->partial, b->d_state, (int)kFinRz, 0);
        }
        // reads z, p_old; writes p_new, q (FP64)
        launch_tiles(b, EQ_KERNEL_STENCIL64, fine_nodes * (3 * sizeof(T) + sizeof(double)), k_ap_dir<R0, T, TENS>, g0, s, A,
                    (const T*)b->z, (const T*)pold, (T*)pnew, b->q, b->partial, (const CgState*)b->d_state);
        finish(kFinAp);
        // reads x, r, q (FP64) and p (+ z when flexible); writes x, r (+ z for
        // Jacobi; + the first sweep, and the FP32 residual when mixed, when the
        // finest level is not fused)
        launch_tiles(b, EQ_KERNEL_CG_UPDATE,
                    fine_nodes * (5 * sizeof(double) + 2 * sizeof(T) + (x1 ? sizeof(T) : 0) + (bcopy ? sizeof(T) : 0)),
                    k_update<R0, T, TENS>, g0, s, A, b->u, b->r, (T*)b->z, (const T*)pnew, (const double*)b->q, b->partial,
                    (const CgState*)b->d_state, jacobi ? 1 : 0, x1, bcopy, omega);
        if (!merge) finish(kFinUpdate);
    };
    // Without extrapolation the confirmation that passes keeps the solution
    // as the next solve's start; with it, k_keep shifts the two kept ones
    // once the solve is over.
    double* keep = S.extrapolate ? nullptr : b->xprev;
    // The confirmation after every chunk of iterations: problems that report
    // convergence get their true residual (the others are left to iterate);
    // those above rtol restart CG from it and form their first sweep (or
    // Jacobi z) again.  Repeating it on a confirmed problem changes nothing.
    auto confirm = [&]() {
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, fine_nodes * (keep ? 32.0 : 24.0), k_true_residual<R0, TENS>, g0, s, A,
                     (const double*)b->u, (const double*)b->rhs, b->r, b->partial, b->d_state, 1, keep);
        finish(kFinTrueDone);
        first_sweep(2);
        if (jacobi) {
            launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, fine_nodes * 16.0, k_jacobi_z<R0, TENS>, g0, s, A,
                         (const double*)b->r, b->z, b->partial, (const CgState*)b->d_state, 1);
            finish(kFinRzRestart);
        }
    };
    // n iterations from search direction p (dir 0) or p2 (dir 1): the two
    // buffers swap roles every iteration
    auto iterations = [&](int dir, int n, bool tail) {
        for (int k = 0; k < n; ++k) {
            const bool odd = ((dir + k) & 1) != 0;
            iteration(odd ? b->p2 : b->p, odd ? b->p : b->p2, merge && k > 0);
        }
        if (merge) finish(kFinUpdate);
        if (tail) confirm();
    };
    auto setup = [&]() {
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, fine_nodes * 16.0, k_rhs<R0>, g0, s, A, (const double*)b->u, b->rhs,
                     b->partial);
        finish(kFinRhs);
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, fine_nodes * (S.extrapolate ? 48.0 : 40.0), k_init<R0, TENS>, g0, s, A,
                     b->u, (const double*)b->xprev, (const double*)b->xprev2, (const double*)b->rhs, b->r, b->z,
                     b->partial, (const CgState*)b->d_state, jacobi ? 1 : 0);
        finish(kFinInit);
        first_sweep(1);
    };
    auto capture = [&](hipGraphExec_t* exec, Acct* acct, auto&& body) -> int {
        if (*exec) return EQ_OK;
        hipGraph_t g = nullptr;
        EQ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        *acct = Acct{};
        b->acct_to = acct; // the captured launches count toward the graph
        body();
        b->acct_to = &b->acct;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) return hip_check(e, "hipStreamEndCapture");
        const hipError_t e2 = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            *exec = nullptr;
            return hip_check(e2, "hipGraphInstantiate");
        }
        return EQ_OK;
    };
    int rc;
    // The set-up, each chunk of iterations and the confirmation after it go
    // out together, then the group's states to the host; the first chunk is
    // the group's last iteration count, so a solve that needs as many again
    // reads the states once per group.  Every group's launches are queued
    // before any group's states are waited for.
    constexpr int kMaxG = eq_batch::kCgGraphMax;
    int chunk[kMG], dir[kMG];
    bool gdone[kMG];
    for (int g = 0; g < kMG; ++g) {
        const int last = ng > 1 ? b->last_iters_g[g] : b->last_iters;
        chunk[g] = std::max(1, last > 0 ? last : S.check_every);
        dir[g] = 0;
        gdone[g] = g >= ng;
    }
    if (ng > 1) {
        // the other groups start behind everything already on the batch's stream
        EQ_HIP(hipEventRecord(b->grp_start, b->stream));
        for (int g = 1; g < ng; ++g) EQ_HIP(hipStreamWaitEvent(b->gstream[g], b->grp_start, 0));
    }
    auto all_done = [&]() {
        for (int g = 0; g < ng; ++g)
            if (!gdone[g]) return false;
        return true;
    };
    for (int guard = 0; guard < 1000000 && !all_done(); ++guard) {
        for (int g = 0; g < ng; ++g) {
            if (gdone[g]) continue;
            use_group(g);
            for (int left = chunk[g]; left > 0;) {
                // graphs of 16, 8, 4, 2 and 1 iterations: the few a solve can
                // need are captured once, whatever its iteration count
                int n = 1;
                while (2 * n <= std::min(left, kMaxG)) n *= 2;
                const bool tail = (n == left);
                const int first = (guard == 0 && left == chunk[g]) ? 1 : 0;
                const int d = dir[g];
                auto body = [&]() {
                    if (first) setup();
                    iterations(d, n, tail);
                };
                if (graphs) {
                    hipGraphExec_t* ge = &b->g_cg[g][first][d][n][tail ? 1 : 0];
                    Acct* ac = &b->g_cg_acct[g][first][d][n][tail ? 1 : 0];
                    if ((rc = capture(ge, ac, body))) return rc;
                    EQ_HIP(hipGraphLaunch(*ge, s));
                    replayed(b, *ac);
                } else {
                    body();
                }
                dir[g] = (d + n) & 1;
                left -= n;
            }
            EQ_HIP(hipMemcpyAsync(b->h_state + gp0, b->d_state + gp0, sizeof(CgState) * (gp1 - gp0),
                                  hipMemcpyDeviceToHost, s));
        }
        EQ_HIP(hipGetLastError());
        for (int g = 0; g < ng; ++g) {
            if (gdone[g]) continue;
            use_group(g);
            EQ_HIP(hipStreamSynchronize(s));
            bool all = true;
            for (int p = gp0; p < gp1; ++p) all = all && b->h_state[p].done;
            gdone[g] = all;
            chunk[g] = std::max(1, S.check_every);
        }
    }
    if (ng > 1) {
        // the batch's stream goes on behind the other groups' launches
        for (int g = 1; g < ng; ++g) {
            EQ_HIP(hipEventRecord(b->grp_done[g], b->gstream[g]));
            EQ_HIP(hipStreamWaitEvent(b->stream, b->grp_done[g], 0));
        }
        for (int g = 0; g < ng; ++g) {
            int32_t m = 0;
            for (int p = b->groups[g].p0; p < b->groups[g].p1; ++p) m = std::max(m, b->h_state[p].iter);
            b->last_iters_g[g] = m;
        }
    }
    if (!keep)
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, L0.nodes_all * 32.0, k_keep<R0>, grid_for((int)L0.tiles_all.size()),
                     b->stream, L0.all(), (const double*)b->u, b->xprev, b->xprev2, b->d_state);
    EQ_HIP(hipGetLastError());
    prof_collect(b);
    int32_t itmax = 0, nconv = 0;
    double relmax = 0.0;
    for (int p = 0; p < b->nprob; ++p) {
        const CgState& c = b->h_state[p];
        const double rel = c.bb > 0 ? std::sqrt(c.rr_true / c.bb) : std::sqrt(c.rr_true);
        b->iters[p] = c.iter;
        b->relres[p] = rel;
        itmax = std::max(itmax, c.iter);
        relmax = std::max(relmax, rel);
        if (c.rr_true <= c.tol2 * c.bb) ++nconv;
    }
    b->last_iters = itmax;
    if (stats) {
        stats->iterations_max = itmax;
        stats->nconverged = nconv;
        stats->relres_max = relmax;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (nconv < b->nprob) {
        char msg[128];
        std::snprintf(msg, sizeof(msg), "max_iter reached above rtol (relres %.3e)", relmax);
        return set_error(EQ_ERR_NOT_CONVERGED, msg);
    }
    return EQ_OK;
}

} // namespace eqk

using namespace eqk;

// ------------------------------------------------------------ C ABI

// Right-preconditioned BiCGStab for the nonsymmetric (advection) operator:
// two V-cycles and two operator applications per iteration, all vectors FP64
// except the preconditioned directions p^, s^ (the V-cycle's precision).
template <typename T, bool TENS, int R0>
static int solve_bicgstab(eq_batch* b, eq_step_stats* stats)
{
    const auto t0 = std::chrono::steady_clock::now();
    const eq_solver_desc& S = b->solver;
    Level& L0 = b->levels[0];
    const LevelView A = L0.all();
    const int g0 = grid_for(A.ntiles);
    hipStream_t s = b->stream;
    CgState* st = b->d_state;
    VCycle<T, TENS> vp(b), vsc(b);
    vp.in = b->bs_p;
    vp.out = (T*)b->p;   // p^
    vsc.in = b->bs_s;
    vsc.out = (T*)b->p2; // s^
    const double fin_bytes = (double)L0.tiles_all.size() * 3 * sizeof(double) + 2.0 * b->nprob * sizeof(CgState);
    const double n0 = L0.nodes_all;
    auto finish = [&](int op) {
        launch(b, EQ_KERNEL_FINISH, fin_bytes, k_finish, dim3(b->nprob), dim3(kFinishThreads), 0, s,
               (const LevelProb*)L0.d_probs, (const double*)b->partial, st, op, 0);
    };
    auto iteration = [&]() {
        launch_tiles(b, EQ_KERNEL_BICG_VEC, n0 * 32.0, k_bicg_p<R0>, g0, s, A, (const double*)b->r, b->bs_p,
                     (const double*)b->bs_v, (const CgState*)st);
        vp.run0();
        launch_tiles(b, EQ_KERNEL_STENCIL64, n0 * (sizeof(T) + 2 * sizeof(double)), k_bicg_apply<R0, T, TENS>, g0, s,
                    A, (const T*)b->p, b->bs_v, (const double*)b->bs_rhat, 0, b->partial, (const CgState*)st);
        finish(kFinBAlpha);
        launch_tiles(b, EQ_KERNEL_BICG_VEC, n0 * 24.0, k_bicg_s<R0>, g0, s, A, (const double*)b->r,
                     (const double*)b->bs_v, b->bs_s, (const CgState*)st);
        vsc.run0();
        launch_tiles(b, EQ_KERNEL_STENCIL64, n0 * (sizeof(T) + 2 * sizeof(double)), k_bicg_apply<R0, T, TENS>, g0, s,
                     A, (const T*)b->p2, b->bs_t, (const double*)b->bs_s, 1, b->partial, (const CgState*)st);
        finish(kFinBOmega);
        launch_tiles(b, EQ_KERNEL_CG_UPDATE, n0 * (2 * sizeof(T) + 6 * sizeof(double)), k_bicg_x<R0, T>, g0, s, A,
                    b->u, b->r, (const T*)b->p, (const T*)b->p2, b->bs_s, b->bs_t, b->bs_rhat, b->partial, st);
        finish(kFinBUpdate);
    };
    auto setup = [&]() {
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * 16.0, k_rhs<R0>, g0, s, A, (const double*)b->u, b->rhs, b->partial);
        finish(kFinRhs);
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * (S.extrapolate ? 48.0 : 40.0), k_init<R0, TENS>, g0, s, A, b->u,
                     (const double*)b->xprev, (const double*)b->xprev2, (const double*)b->rhs, b->r, b->z, b->partial,
                     (const CgState*)st, 0);
        finish(kFinBInit);
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * 16.0, k_copy<R0>, g0, s, A, (const double*)b->r, b->bs_rhat,
                     (const CgState*)st);
    };
    bool profiling = false;
    for (int k = 0; k < EQ_KERNEL_COUNT_; ++k) profiling = profiling || b->prof[k].enabled;
    const bool graphs = S.use_graphs && !profiling;
    auto capture = [&](hipGraphExec_t* exec, Acct* acct, auto&& body) -> int {
        if (*exec) return EQ_OK;
        hipGraph_t g = nullptr;
        EQ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        *acct = Acct{};
        b->acct_to = acct; // the captured launches count toward the graph
        body();
        b->acct_to = &b->acct;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) return hip_check(e, "hipStreamEndCapture");
        const hipError_t e2 = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            *exec = nullptr;
            return hip_check(e2, "hipGraphInstantiate");
        }
        return EQ_OK;
    };
    int rc;
    if (graphs) {
        if ((rc = capture(&b->g_init, &b->g_init_acct, setup)) || (rc = capture(&b->g_bicg, &b->g_bicg_acct, iteration)))
            return rc;
        EQ_HIP(hipGraphLaunch(b->g_init, s));
        replayed(b, b->g_init_acct);
    } else {
        setup();
    }
    EQ_HIP(hipGetLastError());
    int chunk = std::max(1, b->last_iters > 0 ? b->last_iters : S.check_every);
    for (int guard = 0; guard < 1000000; ++guard) {
        for (int k = 0; k < chunk; ++k) {
            if (graphs) {
                EQ_HIP(hipGraphLaunch(b->g_bicg, s));
                replayed(b, b->g_bicg_acct);
            } else {
                iteration();
            }
        }
        EQ_HIP(hipGetLastError());
        if ((rc = read_states(b))) return rc;
        bool all = true;
        for (int p = 0; p < b->nprob; ++p) all = all && b->h_state[p].done;
        if (all) {
            // confirm on the true residual; the others restart from r^ = r
            launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * 24.0, k_true_residual<R0, TENS>, g0, s, A, (const double*)b->u,
                         (const double*)b->rhs, b->r, b->partial, st, 0, (double*)nullptr);
            finish(kFinTrue);
            launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * 16.0, k_copy<R0>, g0, s, A, (const double*)b->r, b->bs_rhat,
                         (const CgState*)st);
            finish(kFinBRestart);
            if ((rc = read_states(b))) return rc;
            all = true;
            for (int p = 0; p < b->nprob; ++p) all = all && b->h_state[p].done;
            if (all) break;
        }
        chunk = std::max(1, S.check_every);
    }
    launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * (S.extrapolate ? 32.0 : 16.0), k_keep<R0>, g0, s, A,
                 (const double*)b->u, b->xprev, b->xprev2, st);
    EQ_HIP(hipGetLastError());
    prof_collect(b);
    int32_t itmax = 0, nconv = 0;
    double relmax = 0.0;
    for (int p = 0; p < b->nprob; ++p) {
        const CgState& c = b->h_state[p];
        const double rel = c.bb > 0 ? std::sqrt(c.rr_true / c.bb) : std::sqrt(c.rr_true);
        b->iters[p] = c.iter;
        b->relres[p] = rel;
        itmax = std::max(itmax, c.iter);
        relmax = std::max(relmax, rel);
        if (c.rr_true <= c.tol2 * c.bb) ++nconv;
    }
    b->last_iters = itmax;
    if (stats) {
        stats->iterations_max = itmax;
        stats->nconverged = nconv;
        stats->relres_max = relmax;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (nconv < b->nprob) {
        char msg[128];
        std::snprintf(msg, sizeof(msg), "max_iter reached above rtol (relres %.3e)", relmax);
        return set_error(EQ_ERR_NOT_CONVERGED, msg);
    }
    return EQ_OK;
}

// Iterative refinement (EQ_SOLVER_IR_GMG_CG): FP64 holds only the residual
// b - A x and the iterate x; each correction A d = r is solved in FP32 by
// flexible CG preconditioned with the FP32 V-cycle (FP32 vectors, FP32 stencil
// apply), from the residual scaled to unit norm, until its residual falls by
// ir_inner_rtol; then x += ||r|| d and the FP64 residual is formed again.
template <int R0, bool TENS>
static int solve_ir(eq_batch* b, eq_step_stats* stats)
{
    using T = float;
    const auto t0 = std::chrono::steady_clock::now();
    const eq_solver_desc& S = b->solver;
    Level& L0 = b->levels[0];
    const LevelView A = L0.all();
    const int g0 = grid_for(A.ntiles);
    hipStream_t s = b->stream;
    const double n0 = L0.nodes_all;
    CgState* so = b->d_state;
    CgState* si = b->d_state_in;
    const double fin_bytes = (double)L0.tiles_all.size() * 3 * sizeof(double) + 2.0 * b->nprob * sizeof(CgState);
    auto finish = [&](CgState* st, int op) {
        launch(b, EQ_KERNEL_FINISH, fin_bytes, k_finish, dim3(b->nprob), dim3(kFinishThreads), 0, s,
               (const LevelProb*)L0.d_probs, (const double*)b->partial, st, op, 0);
    };
    const T omega = (T)S.mg_omega;
    VCycle<T, TENS, T> vc(b);
    vc.in = b->ir_r;
    vc.out = (T*)b->z;
    vc.st_ = si;
    // the fused finest level forms its first sweep itself; otherwise the
    // inner update (and the launch after k_ir_begin) forms it
    vc.fuse_top = true;
    vc.pre1 = !vc.fused_level(0);
    T* x1 = vc.pre1 ? (T*)L0.xa : nullptr;
    // one inner iteration from direction pold to pnew (FP32 throughout)
    auto iteration = [&](T* pold, T* pnew) {
        vc.run0();
        launch(b, EQ_KERNEL_FINISH, fin_bytes, k_finish, dim3(b->nprob), dim3(kFinishThreads), 0, s, vc.rz_table(),
               (const double*)b->partial, si, (int)kFinRz, 0);
        launch_tiles(b, EQ_KERNEL_STENCIL64, n0 * 4.0 * sizeof(T), k_ap_dir<R0, T, TENS, T>, g0, s, A, (const T*)b->z,
                     (const T*)pold, pnew, (T*)b->q, b->partial, (const CgState*)si);
        finish(si, kFinAp);
        launch_tiles(b, EQ_KERNEL_CG_UPDATE, n0 * (x1 ? 7.0 : 6.0) * sizeof(T), k_update<R0, T, TENS, T>, g0, s, A, b->ir_d, b->ir_r,
                     (T*)b->z, (const T*)pnew, (const T*)b->q, b->partial, (const CgState*)si, 0, x1, (T*)nullptr, omega);
        finish(si, kFinUpdate);
    };
    auto two_iterations = [&]() {
        iteration((T*)b->p, (T*)b->p2);
        iteration((T*)b->p2, (T*)b->p);
    };
    auto setup = [&]() {
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * 16.0, k_rhs<R0>, g0, s, A, (const double*)b->u, b->rhs, b->partial);
        finish(so, kFinRhs);
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * (S.extrapolate ? 48.0 : 40.0), k_init<R0, TENS>, g0, s, A, b->u,
                     (const double*)b->xprev, (const double*)b->xprev2, (const double*)b->rhs, b->r, b->z, b->partial,
                     (const CgState*)so, 0);
        finish(so, kFinInit);
    };
    bool profiling = false;
    for (int k = 0; k < EQ_KERNEL_COUNT_; ++k) profiling = profiling || b->prof[k].enabled;
    const bool graphs = S.use_graphs && !profiling;
    auto capture = [&](hipGraphExec_t* exec, Acct* acct, auto&& body) -> int {
        if (*exec) return EQ_OK;
        hipGraph_t g = nullptr;
        EQ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        *acct = Acct{};
        b->acct_to = acct;
        body();
        b->acct_to = &b->acct;
        const hipError_t e = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) return hip_check(e, "hipStreamEndCapture");
        const hipError_t e2 = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            *exec = nullptr;
            return hip_check(e2, "hipGraphInstantiate");
        }
        return EQ_OK;
    };
    int rc;
    if (graphs) {
        if ((rc = capture(&b->g_init, &b->g_init_acct, setup)) || (rc = capture(&b->g_ir2, &b->g_ir2_acct, two_iterations)))
            return rc;
        EQ_HIP(hipGraphLaunch(b->g_init, s));
        replayed(b, b->g_init_acct);
    } else {
        setup();
    }
    EQ_HIP(hipGetLastError());
    auto read_both = [&]() -> int {
        EQ_HIP(hipMemcpyAsync(b->h_state, so, sizeof(CgState) * b->nprob, hipMemcpyDeviceToHost, s));
        EQ_HIP(hipMemcpyAsync(b->h_state_in, si, sizeof(CgState) * b->nprob, hipMemcpyDeviceToHost, s));
        EQ_HIP(hipStreamSynchronize(s));
        return EQ_OK;
    };
    if ((rc = read_both())) return rc;
    std::vector<int32_t> inner_total(b->nprob, 0);
    const int ns = (int)L0.tiles_smooth.size();
    for (int outer = 0; outer < 1000; ++outer) {
        bool all = true;
        for (int p = 0; p < b->nprob; ++p) all = all && b->h_state[p].done;
        if (all) break;
        // FP32 correction problem from the scaled FP64 residual
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * (8.0 + 2.0 * sizeof(T)), k_ir_begin<R0>, g0, s, A,
                     (const double*)b->r, b->ir_r, b->ir_d, b->ir_scale, (const CgState*)so, si, b->partial,
                     S.ir_inner_rtol);
        finish(si, kFinIrInit);
        if (ns > 0 && x1)
            launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, L0.nodes_smooth * 2.0 * sizeof(T), k_smooth<R0, T, T, T, TENS>,
                         grid_for(ns), s, L0.smooth(), (const T*)b->ir_r, (const T*)nullptr, x1, (T*)nullptr, omega, 1, 0,
                         (double*)nullptr, (const CgState*)si);
        int chunk = std::max(2, b->last_iters > 0 ? b->last_iters : S.check_every);
        for (int guard = 0; guard < 1000000; ++guard) {
            for (int k = 0; k < chunk; k += 2) {
                if (graphs) {
                    EQ_HIP(hipGraphLaunch(b->g_ir2, s));
                    replayed(b, b->g_ir2_acct);
                } else {
                    two_iterations();
                }
            }
            EQ_HIP(hipGetLastError());
            if ((rc = read_both())) return rc;
            bool inner_done = true;
            for (int p = 0; p < b->nprob; ++p) inner_done = inner_done && b->h_state_in[p].done;
            if (inner_done) break;
            chunk = std::max(2, S.check_every);
        }
        int32_t imax = 0;
        for (int p = 0; p < b->nprob; ++p) {
            if (!b->h_state[p].done) inner_total[p] += b->h_state_in[p].iter;
            imax = std::max(imax, b->h_state_in[p].iter);
        }
        b->last_iters = imax;
        // x += ||r|| d, then the FP64 residual and the outer test
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * (16.0 + sizeof(T)), k_ir_apply<R0>, g0, s, A, b->u,
                     (const T*)b->ir_d, (const double*)b->ir_scale, (const CgState*)so);
        launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * 24.0, k_true_residual<R0, TENS>, g0, s, A, (const double*)b->u,
                     (const double*)b->rhs, b->r, b->partial, so, 0, (double*)nullptr);
        finish(so, kFinIrOuter);
        if ((rc = read_both())) return rc;
    }
    launch_tiles(b, EQ_KERNEL_SOLVE_ONCE, n0 * (S.extrapolate ? 32.0 : 16.0), k_keep<R0>, g0, s, A,
                 (const double*)b->u, b->xprev, b->xprev2, so);
    EQ_HIP(hipGetLastError());
    prof_collect(b);
    int32_t itmax = 0, nconv = 0;
    double relmax = 0.0;
    for (int p = 0; p < b->nprob; ++p) {
        const CgState& c = b->h_state[p];
        const double rel = c.bb > 0 ? std::sqrt(c.rr_true / c.bb) : std::sqrt(c.rr_true);
        b->iters[p] = inner_total[p];
        b->relres[p] = rel;
        itmax = std::max(itmax, inner_total[p]);
        relmax = std::max(relmax, rel);
        if (c.rr_true <= c.tol2 * c.bb) ++nconv;
    }
    if (stats) {
        stats->iterations_max = itmax;
        stats->nconverged = nconv;
        stats->relres_max = relmax;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (nconv < b->nprob) {
        char msg[128];
        std::snprintf(msg, sizeof(msg), "refinement steps exhausted above rtol (relres %.3e)", relmax);
        return set_error(EQ_ERR_NOT_CONVERGED, msg);
    }
    return EQ_OK;
}

// precision of the multigrid x whether any problem carries a tensor field
static int run_solve(eq_batch* b, eq_step_stats* stats)
{
    const bool bicg = b->solver.kind == EQ_SOLVER_GMG_BICGSTAB || b->solver.kind == EQ_SOLVER_MIXED_GMG_BICGSTAB;
    if (!bicg)
        for (const LevelProb& P : b->levels[0].probs)
            if (P.vel) return set_error(EQ_ERR_INVALID, "advection makes the system nonsymmetric: use a BiCGStab solver");
    if (b->solver.kind == EQ_SOLVER_IR_GMG_CG) {
        if (b->levels[0].R == kRowsWide)
            return b->any_tensor ? solve_ir<kRowsWide, true>(b, stats) : solve_ir<kRowsWide, false>(b, stats);
        return b->any_tensor ? solve_ir<kRows0, true>(b, stats) : solve_ir<kRows0, false>(b, stats);
    }
    // level 0's rows per thread (tile height) is a template argument of every level-0 kernel
    if (b->levels[0].R == kRowsWide) {
        if (bicg) {
            if (b->mixed)
                return b->any_tensor ? solve_bicgstab<float, true, kRowsWide>(b, stats)
                                     : solve_bicgstab<float, false, kRowsWide>(b, stats);
            return b->any_tensor ? solve_bicgstab<double, true, kRowsWide>(b, stats)
                                 : solve_bicgstab<double, false, kRowsWide>(b, stats);
        }
        if (b->mixed)
            return b->any_tensor ? solve_all<float, true, kRowsWide>(b, stats) : solve_all<float, false, kRowsWide>(b, stats);
        return b->any_tensor ? solve_all<double, true, kRowsWide>(b, stats) : solve_all<double, false, kRowsWide>(b, stats);
    }
    if (bicg) {
        if (b->mixed)
            return b->any_tensor ? solve_bicgstab<float, true, kRows0>(b, stats)
                                 : solve_bicgstab<float, false, kRows0>(b, stats);
        return b->any_tensor ? solve_bicgstab<double, true, kRows0>(b, stats)
                             : solve_bicgstab<double, false, kRows0>(b, stats);
    }
    if (b->mixed) return b->any_tensor ? solve_all<float, true, kRows0>(b, stats) : solve_all<float, false, kRows0>(b, stats);
    return b->any_tensor ? solve_all<double, true, kRows0>(b, stats) : solve_all<double, false, kRows0>(b, stats);
}

extern "C" {

const char* eq_last_error(void) { return g_last_error.c_str(); }

int eq_walls_found(const eq_field_desc* desc, int32_t found[4])
{
    if (!desc || !found) return set_error(EQ_ERR_INVALID, "NULL argument");
    std::string why;
    if (validate(*desc, why)) return set_error(EQ_ERR_INVALID, why);
    walls_found(*desc, found);
    return EQ_OK;
}

void eq_solver_default(eq_solver_desc* s)
{
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->kind = EQ_SOLVER_MIXED_GMG_CG;
    s->max_iter = 20000;
    s->rtol = 1e-10;
    s->mg_pre = 2;
    s->mg_post = 2;
    // Two sweeps damped by the reciprocals of the roots of the degree-2
    // Chebyshev polynomial on [0.5, 2] (D^-1 A of the 7-point rows has its
    // spectrum in (0, 2]): 1 / (1.25 +- 0.75 cos(pi/4)).  Measured on the
    // 2000x1000 trap: 9.0 CG iterations per step against 9.3 with 0.8, 0.8.
    s->mg_omega = 1.0 / (1.25 + 0.75 * 0.70710678118654752);
    s->mg_coarse_max_nodes = 256;
    s->mg_max_levels = 16;
    s->check_every = 4;
    s->use_graphs = 1;
    s->mg_fuse_max_nodes = (int32_t)kFuseMaxNodes;
    s->mg_fuse_pre_max_nodes = (int32_t)kFuseMaxNodes;
    s->mg_bottom_lds_bytes = 0; // by batch size (eq_batch_create)
    s->tile4_min_nodes = (int32_t)kWideRowsMinNodes;
    s->tile8_min_nodes = (int32_t)kWideRows8MinNodes;
    s->extrapolate = 0; // measured: no fewer iterations on the 2000x1000 workload (10 either way)
    s->ir_max_outer = 20;
    s->ir_inner_rtol = 1e-4;
    s->mg_level_fused = 1;
    s->mg_omega2 = 1.0 / (1.25 - 0.75 * 0.70710678118654752);
    s->streams = 2;
}

int eq_batch_create(const eq_field_desc* descs, int32_t nprob, int32_t nspecies, const eq_solver_desc* solver,
                    eq_batch** out)
{
    if (!out) return set_error(EQ_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!descs || nprob < 1) return set_error(EQ_ERR_INVALID, "need at least one problem");
    if (nspecies < 1 || nprob % nspecies) return set_error(EQ_ERR_INVALID, "nprob must be a multiple of nspecies");
    for (int p = 0; p < nprob; ++p) {
        std::string why;
        if (validate(descs[p], why)) return set_error(EQ_ERR_INVALID, "problem " + std::to_string(p) + ": " + why);
        const eq_field_desc& t0 = descs[(p / nspecies) * nspecies];
        if (descs[p].nx != t0.nx || descs[p].ny != t0.ny || descs[p].width != t0.width ||
            descs[p].height != t0.height)
            return set_error(EQ_ERR_INVALID, "species of one trap must share its grid");
    }
    eq_solver_desc S;
    if (solver) S = *solver; else eq_solver_default(&S);
    if (S.kind < EQ_SOLVER_JACOBI_CG || S.kind > EQ_SOLVER_IR_GMG_CG) return set_error(EQ_ERR_INVALID, "unknown solver");
    if (S.kind == EQ_SOLVER_IR_GMG_CG && (!(S.ir_inner_rtol > 0 && S.ir_inner_rtol < 1) || S.ir_max_outer < 1))
        return set_error(EQ_ERR_INVALID, "refinement needs 0 < ir_inner_rtol < 1 and ir_max_outer >= 1");
    if (!(S.rtol > 0) || S.max_iter < 1 || S.check_every < 1) return set_error(EQ_ERR_INVALID, "bad solver tolerances");
    if (S.streams < 1 || S.streams > eq_batch::kMaxGroups) return set_error(EQ_ERR_INVALID, "streams must be 1 to 4");
    if (S.kind != EQ_SOLVER_JACOBI_CG &&
        (S.mg_pre < 1 || S.mg_post < 1 || !(S.mg_omega > 0) || !(S.mg_omega2 >= 0) || S.mg_coarse_max_nodes < 4 ||
         S.mg_max_levels < 1))
        return set_error(EQ_ERR_INVALID, "bad multigrid parameters");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(EQ_ERR_NO_DEVICE, "no HIP device");
    eq_batch* b = new (std::nothrow) eq_batch();
    if (!b) return set_error(EQ_ERR_NOMEM, "host allocation");
    hipGetDevice(&b->device);
    {
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device) != hipSuccess || lds <= 0)
            lds = 64 * 1024;
        // leave room for the kernel's static shared memory (tables and the
        // coarsest-level partials)
        // 0: by the batch's size.  The bottom solve runs one workgroup per
        // problem: a batch of few problems leaves almost every CU idle during
        // it, so its levels go to the grid kernels sooner (measured: 2000x1000
        // trap, 2 problems, 40 KiB 1.5 % faster per step than 150 KiB; the
        // 512-trap batch, 1024 problems, 0.6 % slower)
        const int64_t cap = S.mg_bottom_lds_bytes > 0 ? (int64_t)S.mg_bottom_lds_bytes
                            : nprob < kBottomSmallBatch ? kBottomLdsSmall
                                                        : kBottomLdsMax;
        b->lds_budget_max = std::min<int64_t>(kBottomLdsMax, (int64_t)lds - 1024 - kBottomThreads * (int64_t)sizeof(double));
        b->lds_budget = std::min<int64_t>(cap, b->lds_budget_max);
    }
    b->nprob = nprob;
    b->nspecies = nspecies;
    b->ntraps = nprob / nspecies;
    b->descs.assign(descs, descs + nprob);
    for (eq_field_desc& d : b->descs) {
        if (d.marking != EQ_WALLS_COORD) continue;
        int32_t found[4];
        walls_found(d, found);
        for (int w = 0; w < 4; ++w)
            if (!found[w]) d.wall[w] = eq_bc{EQ_BC_NEUMANN, 0, 0.0, 0.0, 0.0};
    }
    b->solver = S;
    b->mixed = (S.kind == EQ_SOLVER_MIXED_GMG_CG || S.kind == EQ_SOLVER_MIXED_GMG_BICGSTAB ||
                S.kind == EQ_SOLVER_IR_GMG_CG) ? 1 : 0;
    b->iters.assign(nprob, 0);
    b->relres.assign(nprob, 0.0);
    int rc = EQ_OK;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return set_error(EQ_ERR_HIP, "hipStreamCreate");
    }
    rc = b->mixed ? build_hierarchy<float>(b) : build_hierarchy<double>(b);
    if (rc == EQ_OK) {
        std::vector<CgState> st(nprob);
        for (int p = 0; p < nprob; ++p) {
            std::memset(&st[p], 0, sizeof(CgState));
            st[p].tol2 = S.rtol * S.rtol;
            st[p].max_iter = S.max_iter;
            st[p].flexible = b->mixed;
            st[p].extrap = S.extrapolate ? 1 : 0;
            if (S.kind == EQ_SOLVER_IR_GMG_CG) st[p].max_iter = S.ir_max_outer; // refinement steps
        }
        if (S.kind == EQ_SOLVER_IR_GMG_CG && rc == EQ_OK) {
            // the inner FP32 solves: flexible CG to ir_inner_rtol (relative to the scaled residual)
            std::vector<CgState> si(nprob);
            for (int p = 0; p < nprob; ++p) {
                std::memset(&si[p], 0, sizeof(CgState));
                si[p].tol2 = S.ir_inner_rtol * S.ir_inner_rtol;
                si[p].max_iter = S.max_iter;
                si[p].flexible = 1;
            }
            const int64_t n0 = b->levels[0].pool;
            if (hipMalloc((void**)&b->d_state_in, sizeof(CgState) * nprob) != hipSuccess ||
                hipHostMalloc((void**)&b->h_state_in, sizeof(CgState) * nprob, hipHostMallocDefault) != hipSuccess ||
                hipMalloc((void**)&b->ir_d, sizeof(float) * n0) != hipSuccess ||
                hipMalloc((void**)&b->ir_r, sizeof(float) * n0) != hipSuccess ||
                hipMalloc((void**)&b->ir_scale, sizeof(double) * nprob) != hipSuccess)
                rc = set_error(EQ_ERR_NOMEM, "refinement buffers");
            else if (hipMemcpy(b->d_state_in, si.data(), sizeof(CgState) * nprob, hipMemcpyHostToDevice) != hipSuccess ||
                     hipMemset(b->ir_d, 0, sizeof(float) * n0) != hipSuccess ||
                     hipMemset(b->ir_r, 0, sizeof(float) * n0) != hipSuccess)
                rc = set_error(EQ_ERR_HIP, "refinement state upload");
        }
        if (hipMalloc((void**)&b->d_state, sizeof(CgState) * nprob) != hipSuccess ||
            hipHostMalloc((void**)&b->h_state, sizeof(CgState) * nprob, hipHostMallocDefault) != hipSuccess ||
            hipMalloc((void**)&b->flux_out, sizeof(double)) != hipSuccess ||
            hipHostMalloc((void**)&b->h_flux, sizeof(double), hipHostMallocDefault) != hipSuccess) {
            rc = set_error(EQ_ERR_NOMEM, "state allocation");
        } else if (hipMemcpy(b->d_state, st.data(), sizeof(CgState) * nprob, hipMemcpyHostToDevice) != hipSuccess) {
            rc = set_error(EQ_ERR_HIP, "state upload");
        } else {
            std::vector<double*> fp(nprob);
            for (int p = 0; p < nprob; ++p) fp[p] = b->u + b->levels[0].probs[p].origin;
            if (hipMalloc((void**)&b->d_field_ptrs, sizeof(double*) * nprob) != hipSuccess ||
                hipMemcpy(b->d_field_ptrs, fp.data(), sizeof(double*) * nprob, hipMemcpyHostToDevice) != hipSuccess)
                rc = set_error(EQ_ERR_NOMEM, "field table");
        }
    }
    if (rc == EQ_OK) rc = init_coupling(b);
    if (rc != EQ_OK) {
        std::string keep = g_last_error;
        eq_batch_destroy(b);
        g_last_error = keep;
        return rc;
    }
    *out = b;
    return EQ_OK;
}

void eq_batch_destroy(eq_batch* b)
{
    if (!b) return;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (Level& L : b->levels) {
        hipFree(L.d_probs);
        hipFree(L.d_tiles_all);
        hipFree(L.d_tiles_smooth);
        hipFree(L.d_bottom);
        hipFree(L.d_bottom_recs);
        hipFree(L.d_tiles_bottom);
        hipFree(L.d_tiles_active);
        hipFree(L.d_fused);
        hipFree(L.d_probs_fin);
        hipFree(L.b);
        hipFree(L.xa);
        hipFree(L.xb);
        hipFree(L.r);
        hipFree(L.tens);
    }
    hipFree(b->u);
    hipFree(b->rhs);
    hipFree(b->r);
    hipFree(b->p);
    hipFree(b->p2);
    hipFree(b->xprev);
    hipFree(b->xprev2);
    hipFree(b->d_state_in);
    if (b->h_state_in) hipHostFree(b->h_state_in);
    hipFree(b->ir_d);
    hipFree(b->ir_r);
    hipFree(b->ir_scale);
    if (b->g_ir2) hipGraphExecDestroy(b->g_ir2);
    hipFree(b->bs_rhat);
    hipFree(b->bs_p);
    hipFree(b->bs_v);
    hipFree(b->bs_s);
    hipFree(b->bs_t);
    if (b->g_bicg) hipGraphExecDestroy(b->g_bicg);
    if (b->mixed) hipFree(b->tens0_d); // else it is level 0's own tensor array
    for (void* v : b->vel_bufs) hipFree(v);
    hipFree(b->q);
    hipFree(b->z);
    hipFree(b->partial);
    hipFree(b->dense);
    hipFree(b->d_state);
    hipFree(b->flux_out);
    hipFree(b->d_field_ptrs);
    hipFree(b->cell_scratch);
    hipFree(b->d_wells);
    hipFree(b->d_tens64);
    hipFree(b->d_tens32);
    hipFree(b->d_level_tables);
    hipFree(b->d_inv_jobs);
    hipFree(b->inv_scratch);
    free_channels(b);
    hipFree(b->d_cell_traps);
    hipFree(b->d_cell_off);
    if (b->h_cell_off) hipHostFree(b->h_cell_off);
    if (b->h_pair_total) hipHostFree(b->h_pair_total);
    if (b->pair_ev) hipEventDestroy(b->pair_ev);
    if (b->cell_off_ev) hipEventDestroy(b->cell_off_ev);
    if (b->h_state) hipHostFree(b->h_state);
    if (b->h_flux) hipHostFree(b->h_flux);
    for (int k = 0; k < EQ_KERNEL_COUNT_; ++k) {
        for (auto& e : b->prof[k].pool) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
        for (auto& e : b->prof[k].pending) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    }
    if (b->grp_start) hipEventDestroy(b->grp_start);
    for (hipEvent_t e : b->grp_done)
        if (e) hipEventDestroy(e);
    for (hipStream_t gs : b->gstream)
        if (gs) hipStreamDestroy(gs);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

int32_t eq_batch_num_problems(const eq_batch* b) { return b ? b->nprob : 0; }

int64_t eq_batch_num_nodes(const eq_batch* b, int32_t prob)
{
    if (!b || prob < 0 || prob >= b->nprob) return -1;
    return (int64_t)(b->descs[prob].nx + 1) * (b->descs[prob].ny + 1);
}

void* eq_batch_stream(const eq_batch* b) { return b ? (void*)b->stream : nullptr; }

int eq_batch_sync(eq_batch* b)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    EQ_HIP(hipStreamSynchronize(b->stream));
    return EQ_OK;
}

static int copy_field(eq_batch* b, int32_t prob, const double* src, double* dst, size_t n, hipMemcpyKind kind,
                      bool sync)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    const LevelProb& P = b->levels[0].probs[prob];
    const size_t nw = (size_t)P.nx + 1, nh = (size_t)P.ny + 1;
    if (n != nw * nh) return set_error(EQ_ERR_INVALID, "length must be (nx+1)*(ny+1)");
    if (!src && !dst) return set_error(EQ_ERR_INVALID, "NULL data");
    double* dev = b->u + P.origin;
    const bool to_dev = (kind == hipMemcpyHostToDevice) || (dst == nullptr);
    if (to_dev) {
        // a new state from the host: the kept solution is no starting point for it
        EQ_HIP(hipMemsetAsync(&b->d_state[prob].have_prev, 0, 2 * sizeof(int32_t), b->stream)); // and have_prev2
        EQ_HIP(hipMemcpy2DAsync(dev, sizeof(double) * P.pitch, src, sizeof(double) * nw, sizeof(double) * nw, nh,
                                kind, b->stream));
    } else {
        EQ_HIP(hipMemcpy2DAsync(dst, sizeof(double) * nw, dev, sizeof(double) * P.pitch, sizeof(double) * nw, nh,
                                kind, b->stream));
    }
    if (sync) EQ_HIP(hipStreamSynchronize(b->stream));
    return EQ_OK;
}

int eq_batch_upload(eq_batch* b, int32_t prob, const double* u, size_t n)
{
    if (!u) return set_error(EQ_ERR_INVALID, "NULL data");
    return copy_field(b, prob, u, nullptr, n, hipMemcpyHostToDevice, true);
}

int eq_batch_download(eq_batch* b, int32_t prob, double* u, size_t n)
{
    if (!u) return set_error(EQ_ERR_INVALID, "NULL data");
    return copy_field(b, prob, nullptr, u, n, hipMemcpyDeviceToHost, true);
}

int eq_batch_upload_dev(eq_batch* b, int32_t prob, const double* d_u, size_t n)
{
    if (!d_u) return set_error(EQ_ERR_INVALID, "NULL data");
    return copy_field(b, prob, d_u, nullptr, n, hipMemcpyDeviceToDevice, false);
}

int eq_batch_download_dev(eq_batch* b, int32_t prob, double* d_u, size_t n)
{
    if (!d_u) return set_error(EQ_ERR_INVALID, "NULL data");
    // device-to-device: the destination is natural order
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    const LevelProb& P = b->levels[0].probs[prob];
    const size_t nw = (size_t)P.nx + 1, nh = (size_t)P.ny + 1;
    if (n != nw * nh) return set_error(EQ_ERR_INVALID, "length must be (nx+1)*(ny+1)");
    EQ_HIP(hipMemcpy2DAsync(d_u, sizeof(double) * nw, b->u + P.origin, sizeof(double) * P.pitch,
                            sizeof(double) * nw, nh, hipMemcpyDeviceToDevice, b->stream));
    return EQ_OK;
}

int eq_batch_step(eq_batch* b, eq_step_stats* stats)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    hipSetDevice(b->device);
    int rc = run_solve(b, stats);
    if (rc != EQ_OK) return rc;
    if ((rc = channels_after_step(b))) return rc; // flow channels take the new trap fields' flux
    return wells_after_step(b);                   // then the boundary wells
}

// ------------------------------------------------------------ tensor field

// Per-node diffusion tensor of one problem (D11, D22, D12 in natural order,
// fenicsInterface's D11/D22/D12 vectors, src/fHSL.cpp:313-323, 387-392),
// or constant coefficients again when all three are NULL.  Coarse levels
// take the value of the nearest fine node; the coarsest level's inverse is
// rebuilt from its tensor.
extern "C++" {
template <typename T>
static int set_tensor_impl(eq_batch* b, int32_t prob, const double* d11, const double* d22, const double* d12)
{
    const int nlev = (int)b->levels.size();
    hipStream_t s = b->stream;
    EQ_HIP(hipStreamSynchronize(s)); // no kernel may be reading the tables being replaced
    int prc = pull_well_values(b);
    if (prc) return prc;
    const bool on = d11 != nullptr;
    if (on) {
        for (int l = 0; l < nlev; ++l) {
            Level& L = b->levels[l];
            if (!L.tens) {
                int rc = alloc_zero((T**)&L.tens, 3 * L.pool, s);
                if (rc) return rc;
            }
        }
        if (!b->tens0_d) {
            if (b->mixed) {
                int rc = alloc_zero(&b->tens0_d, 3 * b->levels[0].pool, s);
                if (rc) return rc;
            } else {
                b->tens0_d = (double*)b->levels[0].tens;
            }
        }
    }
    // host copies level by level, natural order
    std::vector<double> cur[3], nxt[3];
    const LevelProb* Pf = nullptr;
    std::vector<double> inv;
    for (int l = 0; l < nlev; ++l) {
        Level& L = b->levels[l];
        int pi = -1;
        for (int k = 0; k < (int)L.probs.size(); ++k)
            if (L.probs[k].prob == prob) pi = k;
        if (pi < 0) break;
        LevelProb& P = L.probs[pi];
        const int nw = P.nx + 1, nh = P.ny + 1;
        if (on) {
            if (l == 0) {
                for (int c = 0; c < 3; ++c) cur[c].assign((c == 0 ? d11 : c == 1 ? d22 : d12),
                                                          (c == 0 ? d11 : c == 1 ? d22 : d12) + (size_t)nw * nh);
            } else {
                const int fw = Pf->nx + 1;
                for (int c = 0; c < 3; ++c) {
                    nxt[c].resize((size_t)nw * nh);
                    for (int I = 0; I < nh; ++I) {
                        const int i = (int)((2LL * I * Pf->ny + P.ny) / (2LL * P.ny)); // nearest fine row
                        for (int J = 0; J < nw; ++J) {
                            const int j = (int)((2LL * J * Pf->nx + P.nx) / (2LL * P.nx));
                            nxt[c][(size_t)I * nw + J] = cur[c][(size_t)i * fw + j];
                        }
                    }
                    cur[c].swap(nxt[c]);
                }
            }
            // upload in the level's precision (and in FP64 on level 0)
            std::vector<T> rowsT((size_t)nw * nh);
            for (int c = 0; c < 3; ++c) {
                for (size_t q = 0; q < rowsT.size(); ++q) rowsT[q] = (T)cur[c][q];
                T* dst = (T*)L.tens + (int64_t)c * L.pool + P.origin;
                EQ_HIP(hipMemcpy2DAsync(dst, sizeof(T) * P.pitch, rowsT.data(), sizeof(T) * nw, sizeof(T) * nw, nh,
                                        hipMemcpyHostToDevice, s));
                EQ_HIP(hipStreamSynchronize(s)); // rowsT is reused
                if (l == 0 && b->mixed) {
                    double* dd = b->tens0_d + (int64_t)c * L.pool + P.origin;
                    EQ_HIP(hipMemcpy2DAsync(dd, sizeof(double) * P.pitch, cur[c].data(), sizeof(double) * nw,
                                            sizeof(double) * nw, nh, hipMemcpyHostToDevice, s));
                }
            }
            P.aniso = 1;
            P.tstride = L.pool;
            P.tens_d = (l == 0) ? b->tens0_d : (b->mixed ? nullptr : (const double*)L.tens);
            P.tens_f = b->mixed ? (const float*)L.tens : nullptr;
        } else {
            P.aniso = 0;
            P.tens_d = nullptr;
            P.tens_f = nullptr;
        }
        if (P.coarsest && b->solver.kind != EQ_SOLVER_JACOBI_CG && b->dense) {
            if (!dense_inverse(P, on ? cur[0].data() : nullptr, on ? cur[1].data() : nullptr,
                               on ? cur[2].data() : nullptr,
                               (P.vel && (int)b->vel_host[prob].size() > l) ? b->vel_host[prob][l].data() : nullptr,
                               inv))
                return set_error(EQ_ERR_INVALID, "singular coarse operator (tensor field not positive definite?)");
            std::vector<T> invT(inv.size());
            for (size_t q = 0; q < inv.size(); ++q) invT[q] = (T)inv[q];
            EQ_HIP(hipMemcpyAsync((T*)b->dense + P.dense_off, invT.data(), sizeof(T) * invT.size(),
                                  hipMemcpyHostToDevice, s));
            EQ_HIP(hipStreamSynchronize(s));
        }
        EQ_HIP(hipMemcpyAsync(L.d_probs, L.probs.data(), sizeof(LevelProb) * L.probs.size(), hipMemcpyHostToDevice,
                              s));
        EQ_HIP(hipStreamSynchronize(s));
        Pf = &P;
    }
    return sync_bottom_recs(b, true);
}

} // extern "C++"

// The kernels specialise on whether any problem has per-node rows (tensor
// field or advection); a change of that drops the captured graphs.
static void refresh_general(eq_batch* b)
{
    bool any = false;
    for (const LevelProb& P : b->levels[0].probs) any = any || P.aniso || P.vel;
    if (any != b->any_tensor) {
        b->any_tensor = any;
        drop_graphs(b);
    }
}

static int set_tensor_checked(eq_batch* b, int32_t prob, const double* D11, const double* D22, const double* D12,
                              size_t n)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    const bool any = D11 || D22 || D12, all = D11 && D22 && D12;
    if (any && !all) return set_error(EQ_ERR_INVALID, "D11, D22, D12: all or none");
    const LevelProb& P = b->levels[0].probs[prob];
    if (all && n != (size_t)(P.nx + 1) * (P.ny + 1)) return set_error(EQ_ERR_INVALID, "length must be (nx+1)*(ny+1)");
    if (all && P.fd) return set_error(EQ_ERR_INVALID, "the finite-difference scheme has no tensor field");
    if (all && P.vel) return set_error(EQ_ERR_INVALID, "advection and a tensor field together are not supported");
    hipSetDevice(b->device);
    return b->mixed ? set_tensor_impl<float>(b, prob, D11, D22, D12) : set_tensor_impl<double>(b, prob, D11, D22, D12);
}

} // extern "C"

namespace eqk {

template <typename T>
static int refresh_cell_tensors_t(eq_batch* b)
{
    hipStream_t s = b->stream;
    for (size_t l = 1; l < b->levels.size(); ++l) {
        Level& L = b->levels[l];
        const Level& Fn = b->levels[l - 1];
        const int nt = (int)L.tiles_all.size();
        if (nt == 0) continue;
        // each coarse node reads and writes the three entries
        launch(b, EQ_KERNEL_COUPLE_SETUP, L.nodes_all * 6.0 * sizeof(T), k_tensor_inject<T>, dim3(nt), dim3(kBlock), 0, s,
               L.all(), (const LevelProb*)Fn.d_probs, L.R);
    }
    if (b->inv_jobs.empty()) return EQ_OK;
    double by = 0.0;
    for (const InvJob& J : b->inv_jobs) {
        const LevelProb& P = b->levels[J.lev].probs[J.idx];
        const double n = (double)(P.nx + 1) * (P.ny + 1);
        by += n * n * (3.0 * sizeof(double) + sizeof(T));
    }
    launch(b, EQ_KERNEL_COUPLE_SETUP, by, k_coarse_inverse<T>, dim3((int)b->inv_jobs.size()), dim3(256), 0, s,
           (LevelProb* const*)b->d_level_tables, (const InvJob*)b->d_inv_jobs, b->inv_scratch, (T*)b->dense);
    return EQ_OK;
}

int refresh_cell_tensors(eq_batch* b)
{
    return b->mixed ? refresh_cell_tensors_t<float>(b) : refresh_cell_tensors_t<double>(b);
}

} // namespace eqk

extern "C" {

int eq_batch_cell_tensor(eq_batch* b, int32_t trap, double Dx, double Dy, int32_t enable)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (trap < 0 || trap >= b->ntraps) return set_error(EQ_ERR_INVALID, "trap index out of range");
    if (enable && (!(Dx > 0) || !(Dy > 0) || !std::isfinite(Dx) || !std::isfinite(Dy)))
        return set_error(EQ_ERR_INVALID, "Dx and Dy must be finite and > 0");
    const int ns = b->nspecies, p0 = trap * ns;
    for (int p = p0; p < p0 + ns; ++p) {
        const LevelProb& P = b->levels[0].probs[p];
        if (enable && P.fd) return set_error(EQ_ERR_INVALID, "the finite-difference scheme has no tensor field");
        if (enable && P.vel) return set_error(EQ_ERR_INVALID, "advection and a tensor field together are not supported");
    }
    hipSetDevice(b->device);
    const bool was = std::find(b->cell_tensor_traps.begin(), b->cell_tensor_traps.end(), trap) !=
                     b->cell_tensor_traps.end();
    int rc;
    if ((rc = set_trap_tensor(b, trap, Dx, Dy, enable ? 1 : 0))) return rc;
    const LevelProb& Q = b->levels[0].probs[p0];
    const size_t n = (size_t)(Q.nx + 1) * (Q.ny + 1);
    if (enable) {
        // start from the reference's isotropic grids; this allocates the tensor
        // arrays of every level and marks the problems anisotropic
        std::vector<double> one(n, 1.0), zero(n, 0.0);
        for (int p = p0; p < p0 + ns; ++p)
            if ((rc = set_tensor_checked(b, p, one.data(), one.data(), zero.data(), n))) return rc;
    } else {
        for (int p = p0; p < p0 + ns; ++p)
            if ((rc = set_tensor_checked(b, p, nullptr, nullptr, nullptr, 0))) return rc;
    }
    // mark (or clear) the problems on every level, and rebuild the job list
    for (size_t l = 0; l < b->levels.size(); ++l) {
        Level& L = b->levels[l];
        for (LevelProb& P : L.probs)
            if (P.prob >= p0 && P.prob < p0 + ns) P.celltens = enable ? 1 : 0;
        EQ_HIP(hipMemcpy(L.d_probs, L.probs.data(), sizeof(LevelProb) * L.probs.size(), hipMemcpyHostToDevice));
    }
    if ((rc = sync_bottom_recs(b, true))) return rc;
    if (enable && !was) b->cell_tensor_traps.push_back(trap);
    if (!enable && was) b->cell_tensor_traps.erase(std::find(b->cell_tensor_traps.begin(), b->cell_tensor_traps.end(), trap));
    // device tables: per problem, node (0,0) of its finest-level tensor
    if (!b->d_tens64) {
        if (hipMalloc((void**)&b->d_tens64, sizeof(double*) * b->nprob) != hipSuccess ||
            hipMalloc((void**)&b->d_tens32, sizeof(float*) * b->nprob) != hipSuccess ||
            hipMalloc((void**)&b->d_level_tables, sizeof(LevelProb*) * b->levels.size()) != hipSuccess)
            return set_error(EQ_ERR_NOMEM, "tensor tables");
        std::vector<LevelProb*> lt;
        for (Level& L : b->levels) lt.push_back(L.d_probs);
        EQ_HIP(hipMemcpy(b->d_level_tables, lt.data(), sizeof(LevelProb*) * lt.size(), hipMemcpyHostToDevice));
    }
    {
        std::vector<double*> t64(b->nprob, nullptr);
        std::vector<float*> t32(b->nprob, nullptr);
        for (int p = 0; p < b->nprob; ++p) {
            const LevelProb& P = b->levels[0].probs[p];
            if (!P.celltens) continue;
            t64[p] = b->tens0_d + P.origin;
            if (b->mixed) t32[p] = (float*)b->levels[0].tens + P.origin;
        }
        EQ_HIP(hipMemcpy(b->d_tens64, t64.data(), sizeof(double*) * b->nprob, hipMemcpyHostToDevice));
        EQ_HIP(hipMemcpy(b->d_tens32, t32.data(), sizeof(float*) * b->nprob, hipMemcpyHostToDevice));
    }
    // coarsest inverses of every cell-driven problem (multigrid kinds only)
    b->inv_jobs.clear();
    int64_t need = 0;
    if (b->solver.kind != EQ_SOLVER_JACOBI_CG && b->dense)
        for (size_t l = 0; l < b->levels.size(); ++l) {
            const Level& L = b->levels[l];
            for (size_t k = 0; k < L.probs.size(); ++k) {
                const LevelProb& P = L.probs[k];
                if (!P.celltens || !P.coarsest) continue;
                const int64_t nn = (int64_t)(P.nx + 1) * (P.ny + 1);
                b->inv_jobs.push_back(InvJob{(int32_t)l, (int32_t)k, need});
                need += 2 * nn * nn + 3 * nn;
            }
        }
    hipFree(b->d_inv_jobs);
    hipFree(b->inv_scratch);
    b->d_inv_jobs = nullptr;
    b->inv_scratch = nullptr;
    if (!b->inv_jobs.empty()) {
        if (hipMalloc((void**)&b->d_inv_jobs, sizeof(InvJob) * b->inv_jobs.size()) != hipSuccess ||
            hipMalloc((void**)&b->inv_scratch, sizeof(double) * need) != hipSuccess)
            return set_error(EQ_ERR_NOMEM, "coarsest inverse scratch");
        EQ_HIP(hipMemcpy(b->d_inv_jobs, b->inv_jobs.data(), sizeof(InvJob) * b->inv_jobs.size(), hipMemcpyHostToDevice));
    }
    refresh_general(b);
    return EQ_OK;
}

int eq_batch_tensor_values(eq_batch* b, int32_t prob, double* D11, double* D22, double* D12, size_t n)
{
    if (!b || !D11 || !D22 || !D12) return set_error(EQ_ERR_INVALID, "NULL argument");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    const LevelProb& P = b->levels[0].probs[prob];
    const size_t nw = (size_t)P.nx + 1, nh = (size_t)P.ny + 1;
    if (n != nw * nh) return set_error(EQ_ERR_INVALID, "length must be (nx+1)*(ny+1)");
    hipSetDevice(b->device);
    EQ_HIP(hipStreamSynchronize(b->stream));
    double* out[3] = {D11, D22, D12};
    if (!P.aniso) {
        std::fill(D11, D11 + n, 1.0);
        std::fill(D22, D22 + n, 1.0);
        std::fill(D12, D12 + n, 0.0);
        return EQ_OK;
    }
    for (int c = 0; c < 3; ++c)
        EQ_HIP(hipMemcpy2D(out[c], sizeof(double) * nw, b->tens0_d + (int64_t)c * b->levels[0].pool + P.origin,
                           sizeof(double) * P.pitch, sizeof(double) * nw, nh, hipMemcpyDeviceToHost));
    return EQ_OK;
}

int eq_batch_set_tensor(eq_batch* b, int32_t prob, const double* D11, const double* D22, const double* D12,
                        size_t n)
{
    int rc = set_tensor_checked(b, prob, D11, D22, D12, n);
    if (rc != EQ_OK || !b) return rc;
    refresh_general(b);
    return EQ_OK;
}


// x-velocity per grid row of one problem (advection, Galerkin P1), or none.
// Coarse levels take the nearest fine row; the coarsest inverse is rebuilt.
extern "C++" {
template <typename T>
static int set_velocity_impl(eq_batch* b, int32_t prob, const double* vx)
{
    hipStream_t s = b->stream;
    EQ_HIP(hipStreamSynchronize(s));
    int prc = pull_well_values(b);
    if (prc) return prc;
    const int nlev = (int)b->levels.size();
    if ((int)b->vel_host.size() < b->nprob) b->vel_host.resize(b->nprob);
    std::vector<std::vector<double>>& rows = b->vel_host[prob];
    rows.clear();
    const LevelProb* Pf = nullptr;
    std::vector<double> inv;
    for (int l = 0; l < nlev; ++l) {
        Level& L = b->levels[l];
        int pi = -1;
        for (int k = 0; k < (int)L.probs.size(); ++k)
            if (L.probs[k].prob == prob) pi = k;
        if (pi < 0) break;
        LevelProb& P = L.probs[pi];
        if (vx) {
            std::vector<double> r((size_t)P.ny + 1);
            for (int I = 0; I <= P.ny; ++I)
                r[I] = l == 0 ? vx[I] : rows[l - 1][(size_t)((2LL * I * Pf->ny + P.ny) / (2LL * P.ny))];
            rows.push_back(r);
            std::vector<T> rt(r.begin(), r.end());
            // one buffer per (problem, level, precision), kept for later calls
            const size_t slot = ((size_t)prob * kMaxLev + (size_t)l) * 2;
            if (b->vel_bufs.size() < (size_t)b->nprob * kMaxLev * 2) b->vel_bufs.resize((size_t)b->nprob * kMaxLev * 2, nullptr);
            if (!b->vel_bufs[slot] && hipMalloc(&b->vel_bufs[slot], sizeof(T) * rt.size()) != hipSuccess) {
                b->vel_bufs[slot] = nullptr;
                return set_error(EQ_ERR_NOMEM, "velocity rows");
            }
            void* dT = b->vel_bufs[slot];
            EQ_HIP(hipMemcpy(dT, rt.data(), sizeof(T) * rt.size(), hipMemcpyHostToDevice));
            double* dD = nullptr;
            if (l == 0 && b->mixed) {
                if (!b->vel_bufs[slot + 1] &&
                    hipMalloc(&b->vel_bufs[slot + 1], sizeof(double) * r.size()) != hipSuccess) {
                    b->vel_bufs[slot + 1] = nullptr;
                    return set_error(EQ_ERR_NOMEM, "velocity rows");
                }
                dD = (double*)b->vel_bufs[slot + 1];
                EQ_HIP(hipMemcpy(dD, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice));
            }
            P.vel = 1;
            P.vel_f = b->mixed ? (const float*)dT : nullptr;
            P.vel_d = b->mixed ? dD : (const double*)dT;
        } else {
            P.vel = 0;
            P.vel_f = nullptr;
            P.vel_d = nullptr;
        }
        if (P.coarsest && b->solver.kind != EQ_SOLVER_JACOBI_CG && b->dense) {
            // the tensor field, if any, is not kept on the host: velocity and
            // tensor together keep the tensor's coarsest inverse out of date
            if (!dense_inverse(P, nullptr, nullptr, nullptr, vx ? rows.back().data() : nullptr, inv))
                return set_error(EQ_ERR_INVALID, "singular coarse operator");
            std::vector<T> invT(inv.begin(), inv.end());
            EQ_HIP(hipMemcpy((T*)b->dense + P.dense_off, invT.data(), sizeof(T) * invT.size(), hipMemcpyHostToDevice));
        }
        EQ_HIP(hipMemcpy(L.d_probs, L.probs.data(), sizeof(LevelProb) * L.probs.size(), hipMemcpyHostToDevice));
        Pf = &P;
    }
    return sync_bottom_recs(b, true);
}
} // extern "C++"

int eq_batch_set_velocity(eq_batch* b, int32_t prob, const double* vx, size_t n)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    const LevelProb& P = b->levels[0].probs[prob];
    if (vx && n != (size_t)P.ny + 1) return set_error(EQ_ERR_INVALID, "length must be ny+1");
    if (vx && P.fd) return set_error(EQ_ERR_INVALID, "the finite-difference scheme has no advection");
    if (vx && P.aniso) return set_error(EQ_ERR_INVALID, "advection and a tensor field together are not supported");
    hipSetDevice(b->device);
    const int rc = b->mixed ? set_velocity_impl<float>(b, prob, vx) : set_velocity_impl<double>(b, prob, vx);
    if (rc) return rc;
    refresh_general(b);
    return EQ_OK;
}

int eq_field_set_tensor(eq_field* f, const double* D11, const double* D22, const double* D12, size_t n)
{
    return eq_batch_set_tensor(f, 0, D11, D22, D12, n);
}

// Candidate V-cycle shapes (pre, post sweeps) for eq_batch_autotune.
static const int kTuneShapes[][2] = {{1, 1}, {2, 1}, {2, 2}, {3, 3}};

int eq_batch_autotune(eq_batch* b, int32_t repeats, eq_solver_desc* chosen)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (b->solver.kind == EQ_SOLVER_JACOBI_CG) {
        if (chosen) *chosen = b->solver;
        return EQ_OK; // nothing to tune
    }
    hipSetDevice(b->device);
    hipStream_t s = b->stream;
    const int64_t n0 = b->levels[0].pool;
    // snapshot of everything a step changes: field, kept solution, Krylov state
    double* save = nullptr;
    eqk::CgState* save_st = nullptr;
    if (hipMalloc((void**)&save, sizeof(double) * 3 * n0) != hipSuccess ||
        hipMalloc((void**)&save_st, sizeof(eqk::CgState) * b->nprob) != hipSuccess) {
        hipFree(save);
        return set_error(EQ_ERR_NOMEM, "autotune snapshot");
    }
    auto copy_state = [&](bool to_save) -> int {
        double* u = to_save ? save : b->u;
        const double* uf = to_save ? b->u : save;
        double* xp = to_save ? save + n0 : b->xprev;
        const double* xpf = to_save ? b->xprev : save + n0;
        EQ_HIP(hipMemcpyAsync(u, uf, sizeof(double) * n0, hipMemcpyDeviceToDevice, s));
        EQ_HIP(hipMemcpyAsync(xp, xpf, sizeof(double) * n0, hipMemcpyDeviceToDevice, s));
        if (b->xprev2)
            EQ_HIP(hipMemcpyAsync(to_save ? save + 2 * n0 : b->xprev2, to_save ? b->xprev2 : save + 2 * n0,
                                  sizeof(double) * n0, hipMemcpyDeviceToDevice, s));
        EQ_HIP(hipMemcpyAsync(to_save ? save_st : b->d_state, to_save ? b->d_state : save_st,
                              sizeof(eqk::CgState) * b->nprob, hipMemcpyDeviceToDevice, s));
        EQ_HIP(hipStreamSynchronize(s));
        return EQ_OK;
    };
    const eq_solver_desc orig = b->solver;
    const int32_t orig_last = b->last_iters;
    int32_t orig_last_g[eq_batch::kMaxGroups];
    std::copy(b->last_iters_g, b->last_iters_g + eq_batch::kMaxGroups, orig_last_g);
    int rc = copy_state(true);
    double best_t = 1e300;
    int best = -1, best_fuse = orig.mg_fuse_max_nodes;
    const int reps = std::max(1, repeats);
    int best_top = orig.mg_fuse_pre_top;
    // candidates: every cycle shape with the batch's fusion settings, then
    // the fastest shape with every level fused, then the fastest so far with
    // the finest level's pre-sweeps fused the other way
    const int nshapes = (int)(sizeof(kTuneShapes) / sizeof(kTuneShapes[0]));
    for (int k = 0; rc == EQ_OK && k <= nshapes + 1; ++k) {
        if (k >= nshapes && best < 0) break;
        const int shape = k < nshapes ? k : best;
        const int32_t fuse = k < nshapes ? orig.mg_fuse_max_nodes : k == nshapes ? INT32_MAX : best_fuse;
        const int32_t top = k <= nshapes ? orig.mg_fuse_pre_top : !best_top;
        if (k == nshapes && orig.mg_fuse_max_nodes == INT32_MAX) continue;
        if (k == nshapes + 1 && kTuneShapes[shape][0] < 2) break; // one pre-sweep: nothing to fuse
        b->solver.mg_pre = kTuneShapes[shape][0];
        b->solver.mg_post = kTuneShapes[shape][1];
        b->solver.mg_fuse_max_nodes = fuse;
        b->solver.mg_fuse_pre_top = top;
        drop_graphs(b);
        double t = 0.0;
        bool ok = true;
        for (int r = 0; r <= reps && rc == EQ_OK; ++r) { // r = 0 captures the graphs, untimed
            b->last_iters = orig_last;
            std::copy(orig_last_g, orig_last_g + eq_batch::kMaxGroups, b->last_iters_g);
            eq_step_stats st{};
            const auto t0 = std::chrono::steady_clock::now();
            const int src = run_solve(b, &st);
            const hipError_t e = hipStreamSynchronize(s);
            if (e != hipSuccess) {
                rc = hip_check(e, "autotune step");
                break;
            }
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (src != EQ_OK || st.nconverged != b->nprob) ok = false;
            if (r > 0) t += dt;
            rc = copy_state(false);
        }
        if (rc == EQ_OK && ok && t < best_t) {
            best_t = t;
            best = shape;
            best_fuse = fuse;
            best_top = top;
        }
    }
    b->solver = orig;
    if (best >= 0) {
        b->solver.mg_pre = kTuneShapes[best][0];
        b->solver.mg_post = kTuneShapes[best][1];
        b->solver.mg_fuse_max_nodes = best_fuse;
        b->solver.mg_fuse_pre_top = best_top;
    }
    b->last_iters = orig_last;
    std::copy(orig_last_g, orig_last_g + eq_batch::kMaxGroups, b->last_iters_g);
    drop_graphs(b);
    hipFree(save);
    hipFree(save_st);
    if (rc != EQ_OK) return rc;
    if (chosen) *chosen = b->solver;
    return EQ_OK;
}

int eq_batch_problem_stats(const eq_batch* b, int32_t* iterations, double* relres)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    for (int p = 0; p < b->nprob; ++p) {
        if (iterations) iterations[p] = b->iters[p];
        if (relres) relres[p] = b->relres[p];
    }
    return EQ_OK;
}

static int apply_diag(eq_batch* b, int32_t prob, const double* x, double* y, size_t n, int mode)
{
    if (!b || !x || !y) return set_error(EQ_ERR_INVALID, "NULL argument");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    const LevelProb& P = b->levels[0].probs[prob];
    const size_t nw = (size_t)P.nx + 1, nh = (size_t)P.ny + 1;
    if (n != nw * nh) return set_error(EQ_ERR_INVALID, "length must be (nx+1)*(ny+1)");
    hipSetDevice(b->device);
    hipStream_t s = b->stream;
    EQ_HIP(hipMemcpy2DAsync(b->q + P.origin, sizeof(double) * P.pitch, x, sizeof(double) * nw, sizeof(double) * nw,
                            nh, hipMemcpyHostToDevice, s));
    const LevelView A = b->levels[0].all();
    if (b->any_tensor)
        if (b->levels[0].R == kRowsWide)
            launch_tiles(b, -1, 0.0, k_apply_full<kRowsWide, true>, grid_for(A.ntiles), s, A, (int)prob, (const double*)b->q, b->z, mode);
        else
            launch_tiles(b, -1, 0.0, k_apply_full<kRows0, true>, grid_for(A.ntiles), s, A, (int)prob, (const double*)b->q, b->z, mode);
    else
        if (b->levels[0].R == kRowsWide)
            launch_tiles(b, -1, 0.0, k_apply_full<kRowsWide, false>, grid_for(A.ntiles), s, A, (int)prob, (const double*)b->q, b->z, mode);
        else
            launch_tiles(b, -1, 0.0, k_apply_full<kRows0, false>, grid_for(A.ntiles), s, A, (int)prob, (const double*)b->q, b->z, mode);
    EQ_HIP(hipGetLastError());
    EQ_HIP(hipMemcpy2DAsync(y, sizeof(double) * nw, b->z + P.origin, sizeof(double) * P.pitch, sizeof(double) * nw,
                            nh, hipMemcpyDeviceToHost, s));
    EQ_HIP(hipStreamSynchronize(s));
    return EQ_OK;
}

int eq_batch_apply(eq_batch* b, int32_t prob, const double* x, double* y, size_t n)
{
    return apply_diag(b, prob, x, y, n, 0);
}

int eq_batch_rhs(eq_batch* b, int32_t prob, const double* u0, double* rhs, size_t n)
{
    return apply_diag(b, prob, u0, rhs, n, 1);
}

int eq_batch_boundary_flux(eq_batch* b, int32_t prob, double* total)
{
    if (!b || !total) return set_error(EQ_ERR_INVALID, "NULL argument");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    const LevelProb& P = b->levels[0].probs[prob];
    launch(b, EQ_KERNEL_BOUNDARY, 4.0 * (P.nx + P.ny) * sizeof(double), k_flux, dim3(1), dim3(256), 0, b->stream,
           (const LevelProb*)b->levels[0].d_probs, (int)prob, (const double*)b->u, b->flux_out);
    EQ_HIP(hipGetLastError());
    EQ_HIP(hipMemcpyAsync(b->h_flux, b->flux_out, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    EQ_HIP(hipStreamSynchronize(b->stream));
    *total = *b->h_flux;
    return EQ_OK;
}

int eq_batch_set_boundary_value(eq_batch* b, int32_t prob, double value)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    if (!std::isfinite(value)) return set_error(EQ_ERR_INVALID, "boundary value must be finite");
    LevelProb& P = b->levels[0].probs[prob];
    if (!P.dmask) return set_error(EQ_ERR_INVALID, "the problem has no Dirichlet wall");
    hipSetDevice(b->device);
    for (int k = 0; k < 4; ++k)
        if ((P.dmask >> k) & 1) P.g[k] = value; // the host copy stays current
    launch(b, EQ_KERNEL_BOUNDARY, 0.0, k_set_walls, dim3(1), dim3(64), 0, b->stream, b->levels[0].d_probs, (int)prob,
           value);
    EQ_HIP(hipGetLastError());
    return EQ_OK;
}

int eq_batch_set_well(eq_batch* b, int32_t prob, const eq_well_desc* well)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    const LevelProb& P = b->levels[0].probs[prob];
    if (well) {
        if (P.dmask != 15) return set_error(EQ_ERR_INVALID, "a boundary well needs all four walls Dirichlet");
        for (int k = 0; k < 4; ++k)
            if (P.gw[k]) return set_error(EQ_ERR_INVALID, "a boundary well's walls cannot carry per-node values");
        if (!(well->well_scaling > 0) || !std::isfinite(well->well_scaling) || !(well->decay_rate >= 0) ||
            !std::isfinite(well->decay_rate) || !std::isfinite(well->w0))
            return set_error(EQ_ERR_INVALID, "need well_scaling > 0, decay_rate >= 0, finite w0");
    }
    hipSetDevice(b->device);
    int rc = pull_well_values(b); // also brings the device's well states into b->wells
    if (rc) return rc;
    for (size_t k = 0; k < b->wells.size(); ++k)
        if (b->wells[k].prob == prob) {
            b->wells.erase(b->wells.begin() + (ptrdiff_t)k);
            break;
        }
    if (well) {
        WellDev W{};
        W.prob = prob;
        W.scaling = well->well_scaling;
        W.decay = well->decay_rate;
        W.dt = b->descs[prob].dt;
        W.w = well->w0;
        b->wells.push_back(W);
    }
    if (!b->d_wells && !b->wells.empty()) {
        if (hipMalloc((void**)&b->d_wells, sizeof(WellDev) * b->nprob) != hipSuccess) {
            b->d_wells = nullptr;
            return set_error(EQ_ERR_NOMEM, "well table");
        }
        b->d_wells_cap = b->nprob;
    }
    if (!b->wells.empty())
        EQ_HIP(hipMemcpy(b->d_wells, b->wells.data(), sizeof(WellDev) * b->wells.size(), hipMemcpyHostToDevice));
    // the first solve sees the starting concentration on the walls
    return well ? eq_batch_set_boundary_value(b, prob, well->w0) : EQ_OK;
}

int eq_batch_well_state(eq_batch* b, int32_t prob, double* w, double* last_flux, int64_t* underflows)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    hipSetDevice(b->device);
    int rc = pull_well_values(b);
    if (rc) return rc;
    for (const WellDev& W : b->wells)
        if (W.prob == prob) {
            if (w) *w = W.w;
            if (last_flux) *last_flux = W.flux;
            if (underflows) *underflows = W.underflows;
            return EQ_OK;
        }
    return set_error(EQ_ERR_INVALID, "the problem has no boundary well");
}

int eq_batch_vertex_map(const eq_batch* b, int32_t prob, double* coords, int32_t* dof_from_vertex)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (prob < 0 || prob >= b->nprob) return set_error(EQ_ERR_INVALID, "problem index out of range");
    const eq_field_desc& d = b->descs[prob];
    const int64_t nw = d.nx + 1, n = nw * (d.ny + 1);
    for (int64_t v = 0; v < n; ++v) {
        const int64_t i = v / nw, j = v - i * nw;
        // RectangleMesh vertex coordinates, natural order
        if (coords) {
            coords[2 * v] = ((double)j * d.width) / (double)d.nx;
            coords[2 * v + 1] = ((double)i * d.height) / (double)d.ny;
        }
        if (dof_from_vertex) dof_from_vertex[v] = (int32_t)v;
    }
    return EQ_OK;
}

int eq_batch_profile(eq_batch* b, int32_t kernel, int32_t enable)
{
    if (!b || kernel < 0 || kernel >= EQ_KERNEL_COUNT_) return set_error(EQ_ERR_INVALID, "bad kernel id");
    b->prof[kernel].enabled = enable != 0;
    b->prof[kernel].launches = 0;
    b->prof[kernel].total_ms = 0.0;
    b->prof[kernel].bytes = 0.0;
    return EQ_OK;
}

int eq_batch_profile_read(eq_batch* b, int32_t kernel, int64_t* launches, double* total_ms, double* bytes)
{
    if (!b || kernel < 0 || kernel >= EQ_KERNEL_COUNT_) return set_error(EQ_ERR_INVALID, "bad kernel id");
    hipStreamSynchronize(b->stream);
    prof_collect(b);
    if (launches) *launches = b->prof[kernel].launches;
    if (total_ms) *total_ms = b->prof[kernel].total_ms;
    if (bytes) *bytes = b->prof[kernel].bytes;
    return EQ_OK;
}

int eq_batch_traffic(eq_batch* b, double* bytes, int64_t* launches, int32_t reset)
{
    if (!b) return set_error(EQ_ERR_INVALID, "NULL batch");
    if (bytes) *bytes = b->acct.bytes;
    if (launches) *launches = b->acct.launches;
    if (reset) b->acct = Acct{};
    return EQ_OK;
}

// single-field API
int eq_field_create(const eq_field_desc* desc, const eq_solver_desc* solver, eq_field** out)
{
    return eq_batch_create(desc, 1, 1, solver, out);
}
void eq_field_destroy(eq_field* f) { eq_batch_destroy(f); }
int eq_field_upload(eq_field* f, const double* u, size_t n) { return eq_batch_upload(f, 0, u, n); }
int eq_field_download(eq_field* f, double* u, size_t n) { return eq_batch_download(f, 0, u, n); }
int eq_field_step(eq_field* f, eq_step_stats* stats) { return eq_batch_step(f, stats); }
int eq_field_vertex_map(const eq_field* f, double* coords, int32_t* dof_from_vertex)
{
    return eq_batch_vertex_map(f, 0, coords, dof_from_vertex);
}
int eq_field_boundary_flux(eq_field* f, double* total) { return eq_batch_boundary_flux(f, 0, total); }
int eq_cells_couple(eq_field* f, const eq_cell* cells, size_t nc, double npm, const double* q, double* sampled,
                    int32_t* counts, uint32_t flags)
{
    return eq_batch_couple(f, 0, cells, nc, npm, q, sampled, counts, flags);
}

} // extern "C"
