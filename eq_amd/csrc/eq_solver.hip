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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "eq_internal.h"

namespace eqk {

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
    int t, pi, i, j;
    int64_t idx;
    bool in;
};

__device__ __forceinline__ bool decode(const LevelView& L, NodeCtx& c)
{
    c.t = tile_index();
    if (c.t >= L.ntiles) return false;
    const TileRef tr = L.tiles[c.t];
    c.pi = tr.pi;
    c.P = &L.probs[tr.pi];
    const int ty = tr.local / c.P->tiles_x, tx = tr.local - ty * c.P->tiles_x;
    c.j = tx * kTileW + (int)(threadIdx.x & (kTileW - 1));
    c.i = ty * kTileH + (int)(threadIdx.x / kTileW);
    c.in = (c.i <= c.P->ny) && (c.j <= c.P->nx);
    c.idx = c.P->origin + (int64_t)c.i * c.P->pitch + c.j;
    return true;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// deterministic block sum; result valid in thread 0
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double red[kBlock / 64][NV];
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
            for (int q = 0; q < (int)(blockDim.x / 64); ++q) s += red[q][k];
            v[k] = s;
        }
    __syncthreads();
}

// Write this tile's partial sums (plain stores).  The per-problem totals are
// formed by k_finish in the next launch: the kernel boundary makes the
// partials visible across XCDs, where an in-kernel device-scope fence per
// block would write back the XCD's whole L2 every time.
template <int NV>
__device__ __forceinline__ void tile_partial(double (&v)[NV], double* partial, int t)
{
    block_sum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) partial[(int64_t)t * 4 + k] = v[k];
}

enum FinishOp { kFinRhs = 0, kFinInit = 1, kFinAp = 2, kFinUpdate = 3, kFinRz = 4, kFinTrue = 5 };

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

// ------------------------------------------------------------ level-0 kernels (FP64)

// b = M u0 + sources; ||b||^2 of the row-replaced system (b_D = g)
__global__ __launch_bounds__(kBlock) void k_rhs(LevelView L, const double* __restrict__ u, double* __restrict__ b,
                                                double* partial, CgState* st)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    const LevelProb& P = *c.P;
    double v[1] = {0.0};
    if (c.in) {
        double uc, uw, ue, us, un, usw, une;
        load7(u, c.idx, P.pitch, uc, uw, ue, us, un, usw, une);
        const double bv = rhs_row(P, c.i, c.j, uc, uw, ue, us, un, usw, une);
        b[c.idx] = bv;
        if (is_dirichlet(P, c.i, c.j)) {
            const double g = dirichlet_value(P, c.i, c.j);
            v[0] = g * g;
        } else {
            v[0] = bv * bv;
        }
    }
    tile_partial<1>(v, partial, c.t);
}

// x_D = g; r = b - A x; Jacobi: z = p = r / diag(A)
__global__ __launch_bounds__(kBlock) void k_init(LevelView L, double* x, const double* __restrict__ b,
                                                 double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ p, double* partial, CgState* st, int jacobi)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    const LevelProb& P = *c.P;
    double v[2] = {0.0, 0.0};
    if (c.in) {
        if (is_dirichlet(P, c.i, c.j)) {
            x[c.idx] = dirichlet_value(P, c.i, c.j);
            r[c.idx] = 0.0;
            z[c.idx] = 0.0;
            p[c.idx] = 0.0;
        } else {
            const int i = c.i, j = c.j, pt = P.pitch;
            const int64_t k = c.idx;
            auto val = [&](int ii, int jj, int64_t kk) -> double {
                return is_dirichlet(P, ii, jj) ? dirichlet_value(P, ii, jj) : x[kk];
            };
            const double xc = x[k];
            const double xw = val(i, j - 1, k - 1), xe = val(i, j + 1, k + 1);
            const double xs = val(i - 1, j, k - pt), xn = val(i + 1, j, k + pt);
            const double xsw = val(i - 1, j - 1, k - pt - 1), xne = val(i + 1, j + 1, k + pt + 1);
            const Row<double> rc = row_coefs<double>(P, i, j);
            const double rv = b[k] - row_apply(rc, xc, xw, xe, xs, xn, xsw, xne);
            r[k] = rv;
            v[0] = rv * rv;
            if (jacobi) {
                const double zv = rv / row_diag(rc);
                z[k] = zv;
                p[k] = zv;
                v[1] = rv * zv;
            }
        }
    }
    tile_partial<2>(v, partial, c.t);
}

// q = A p; alpha = rz / (p.q)
__global__ __launch_bounds__(kBlock) void k_ap(LevelView L, const double* __restrict__ p, double* __restrict__ q,
                                               double* partial, CgState* st)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    const LevelProb& P = *c.P;
    if (st[P.prob].done) return;
    double v[1] = {0.0};
    if (c.in) {
        double pc, pw, pe, ps, pn, psw, pne;
        load7(p, c.idx, P.pitch, pc, pw, pe, ps, pn, psw, pne);
        double qv = 0.0;
        if (!is_dirichlet(P, c.i, c.j)) {
            const Row<double> rc = row_coefs<double>(P, c.i, c.j);
            qv = row_apply(rc, pc, pw, pe, ps, pn, psw, pne);
        }
        q[c.idx] = qv;
        v[0] = pc * qv;
    }
    tile_partial<1>(v, partial, c.t);
}

// x += alpha p; r -= alpha q; Jacobi: z = r/diag, beta; flexible: r.z_old
__global__ __launch_bounds__(kBlock) void k_update(LevelView L, double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ z, const double* __restrict__ p,
                                                   const double* __restrict__ q, double* partial, CgState* st,
                                                   int jacobi)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    const LevelProb& P = *c.P;
    const CgState& s0 = st[P.prob];
    if (s0.done) return;
    const double alpha = s0.alpha;
    const int flexible = s0.flexible;
    double v[3] = {0.0, 0.0, 0.0};
    if (c.in && !is_dirichlet(P, c.i, c.j)) {
        const int64_t k = c.idx;
        x[k] += alpha * p[k];
        const double rv = r[k] - alpha * q[k];
        r[k] = rv;
        v[0] = rv * rv;
        if (jacobi) {
            const Row<double> rc = row_coefs<double>(P, c.i, c.j);
            const double zv = rv / row_diag(rc);
            z[k] = zv;
            v[1] = rv * zv;
        } else if (flexible) {
            v[2] = rv * z[k];
        }
    }
    tile_partial<3>(v, partial, c.t);
}

// multigrid path: rz = r.z (after the V-cycle), beta (Fletcher-Reeves or
// Polak-Ribiere for a variable preconditioner)
__global__ __launch_bounds__(kBlock) void k_rz(LevelView L, const double* __restrict__ r,
                                               const double* __restrict__ z, double* partial, CgState* st)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    const LevelProb& P = *c.P;
    if (st[P.prob].done) return;
    double v[1] = {0.0};
    if (c.in) v[0] = r[c.idx] * z[c.idx];
    tile_partial<1>(v, partial, c.t);
}

// p = z + beta p
__global__ __launch_bounds__(kBlock) void k_dir(LevelView L, const double* __restrict__ z, double* __restrict__ p,
                                                const CgState* st)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    const LevelProb& P = *c.P;
    const CgState& s = st[P.prob];
    if (s.done) return;
    if (c.in) p[c.idx] = z[c.idx] + s.beta * p[c.idx];
}

// true residual r = b - A x (replaces the recurrence residual); final stop test
__global__ __launch_bounds__(kBlock) void k_true_residual(LevelView L, const double* __restrict__ x,
                                                          const double* __restrict__ b, double* __restrict__ r,
                                                          double* partial, CgState* st)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    const LevelProb& P = *c.P;
    double v[1] = {0.0};
    if (c.in && !is_dirichlet(P, c.i, c.j)) {
        double xc, xw, xe, xs, xn, xsw, xne;
        load7(x, c.idx, P.pitch, xc, xw, xe, xs, xn, xsw, xne);
        const Row<double> rc = row_coefs<double>(P, c.i, c.j);
        const double rv = b[c.idx] - row_apply(rc, xc, xw, xe, xs, xn, xsw, xne);
        r[c.idx] = rv;
        v[0] = rv * rv;
    }
    tile_partial<1>(v, partial, c.t);
}

// Per-problem totals of the tile partials, summed in tile order
// (deterministic), and the Krylov scalar updates that follow from them.
// One block per problem; problems the producer skipped (done) are skipped.
__global__ __launch_bounds__(kBlock) void k_finish(const LevelProb* __restrict__ probs,
                                                   const double* __restrict__ partial, CgState* st, int op,
                                                   int jacobi)
{
    const LevelProb& P = probs[blockIdx.x];
    CgState& s = st[P.prob];
    if ((op == kFinAp || op == kFinUpdate || op == kFinRz) && s.done) return;
    double v[3] = {0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < P.ntiles; q += blockDim.x) {
        const double* pp = partial + (int64_t)(P.tile_begin + q) * 4;
        v[0] += pp[0];
        v[1] += pp[1];
        v[2] += pp[2];
    }
    block_sum<3>(v);
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
            s.done = 1; // exact solution reached (or breakdown): stop this problem
        }
        break;
    case kFinUpdate:
        s.iter += 1;
        s.rr = v[0];
        s.rzold = v[2];
        if (jacobi) {
            s.beta = v[1] / s.rz;
            s.rz = v[1];
        }
        s.done = (v[0] <= s.tol2 * s.bb || s.iter >= s.max_iter) ? 1 : 0;
        break;
    case kFinRz:
        if (s.iter == 0 || s.restart) {
            s.beta = 0.0;
            s.restart = 0;
        } else {
            const double num = s.flexible ? (v[0] - s.rzold) : v[0];
            s.beta = num / s.rz;
        }
        s.rz = v[0];
        break;
    case kFinTrue:
        s.rr_true = v[0];
        s.rr = v[0];
        s.done = (v[0] <= s.tol2 * s.bb || s.iter >= s.max_iter) ? 1 : 0;
        s.restart = s.done ? 0 : 1;
        break;
    }
}

// Jacobi preconditioner on its own (restart after a residual replacement)
__global__ __launch_bounds__(kBlock) void k_jacobi_z(LevelView L, const double* __restrict__ r,
                                                     double* __restrict__ z, const CgState* st)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    const LevelProb& P = *c.P;
    if (st[P.prob].done) return;
    if (!c.in) return;
    if (is_dirichlet(P, c.i, c.j)) {
        z[c.idx] = 0.0;
        return;
    }
    const Row<double> rc = row_coefs<double>(P, c.i, c.j);
    z[c.idx] = r[c.idx] / row_diag(rc);
}

// ------------------------------------------------------------ multigrid kernels

// Damped Jacobi sweep  x_out = x_in + omega (b - A x_in) / diag  (x_in = 0
// when zero_init).  bcopy (optional) keeps b in the smoother precision.
template <typename T, typename TB, typename TO>
__global__ __launch_bounds__(kBlock) void k_smooth(LevelView L, const TB* __restrict__ b, const T* __restrict__ xin,
                                                   TO* __restrict__ xout, T* __restrict__ bcopy, T omega,
                                                   int zero_init, const CgState* st)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    const LevelProb& P = *c.P;
    if (st[P.prob].done) return;
    if (!c.in) return;
    if (is_dirichlet(P, c.i, c.j)) {
        xout[c.idx] = (TO)0;
        if (bcopy) bcopy[c.idx] = (T)0;
        return;
    }
    const Row<T> rc = row_coefs<T>(P, c.i, c.j);
    const T d = row_diag(rc);
    const T bv = (T)b[c.idx];
    if (bcopy) bcopy[c.idx] = bv;
    T out;
    if (zero_init) {
        out = omega * bv / d;
    } else {
        T xc, xw, xe, xs, xn, xsw, xne;
        load7(xin, c.idx, P.pitch, xc, xw, xe, xs, xn, xsw, xne);
        out = xc + omega * (bv - row_apply(rc, xc, xw, xe, xs, xn, xsw, xne)) / d;
    }
    xout[c.idx] = (TO)out;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_residual(LevelView L, const T* __restrict__ b, const T* __restrict__ x,
                                                     T* __restrict__ r, const CgState* st)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    const LevelProb& P = *c.P;
    if (st[P.prob].done) return;
    if (!c.in) return;
    if (is_dirichlet(P, c.i, c.j)) {
        r[c.idx] = (T)0;
        return;
    }
    T xc, xw, xe, xs, xn, xsw, xne;
    load7(x, c.idx, P.pitch, xc, xw, xe, xs, xn, xsw, xne);
    const Row<T> rc = row_coefs<T>(P, c.i, c.j);
    r[c.idx] = b[c.idx] - row_apply(rc, xc, xw, xe, xs, xn, xsw, xne);
}

// b_c = P^T r_f, gathered per coarse node over the fine nodes whose P1
// interpolation uses it (no atomics).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(LevelView Lc, const LevelProb* __restrict__ fine,
                                                     const T* __restrict__ rf, T* __restrict__ bc, const CgState* st)
{
    NodeCtx c;
    if (!decode(Lc, c)) return;
    const LevelProb& Pc = *c.P;
    if (st[Pc.prob].done) return;
    if (!c.in) return;
    if (is_dirichlet(Pc, c.i, c.j)) {
        bc[c.idx] = (T)0;
        return;
    }
    const LevelProb& Pf = fine[Pc.fidx];
    const int nxf = Pf.nx, nyf = Pf.ny, nxc = Pc.nx, nyc = Pc.ny;
    const int I = c.i, J = c.j;
    const int jlo = (J >= 1) ? (int)(((int64_t)(J - 1) * nxf + nxc - 1) / nxc) : 0;
    const int jhi = (int)std::min<int64_t>((int64_t)nxf, ((int64_t)(J + 1) * nxf) / nxc);
    const int ilo = (I >= 1) ? (int)(((int64_t)(I - 1) * nyf + nyc - 1) / nyc) : 0;
    const int ihi = (int)std::min<int64_t>((int64_t)nyf, ((int64_t)(I + 1) * nyf) / nyc);
    T sum = (T)0;
    for (int i = ilo; i <= ihi; ++i) {
        const T* row = rf + Pf.origin + (int64_t)i * Pf.pitch;
        for (int j = jlo; j <= jhi; ++j) {
            const Interp w = interp_weights(i, j, nxf, nyf, nxc, nyc);
            double wt = 0.0;
            if (w.I0 == I) {
                if (w.J0 == J) wt += w.wbl;
                else if (w.J0 + 1 == J) wt += w.wbr;
            } else if (w.I0 + 1 == I) {
                if (w.J0 == J) wt += w.wtl;
                else if (w.J0 + 1 == J) wt += w.wtr;
            }
            if (wt != 0.0) sum += (T)wt * row[j];
        }
    }
    bc[c.idx] = sum;
}

// x_f += P x_c  (fine Dirichlet nodes stay 0)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_prolong(LevelView Lf, const LevelProb* __restrict__ coarse,
                                                    const T* __restrict__ xc, T* __restrict__ xf, const CgState* st)
{
    NodeCtx c;
    if (!decode(Lf, c)) return;
    const LevelProb& Pf = *c.P;
    if (st[Pf.prob].done) return;
    if (!c.in || is_dirichlet(Pf, c.i, c.j)) return;
    const LevelProb& Pc = coarse[Pf.cidx];
    const Interp w = interp_weights(c.i, c.j, Pf.nx, Pf.ny, Pc.nx, Pc.ny);
    const int64_t k = Pc.origin + (int64_t)w.I0 * Pc.pitch + w.J0;
    const T v = (T)w.wbl * xc[k] + (T)w.wbr * xc[k + 1] + (T)w.wtl * xc[k + Pc.pitch] +
                (T)w.wtr * xc[k + Pc.pitch + 1];
    xf[c.idx] += v;
}

// coarsest level: x = A^{-1} b with a precomputed dense inverse (one block per problem)
template <typename T, typename TO>
__global__ __launch_bounds__(256) void k_coarse_dense(const LevelProb* __restrict__ probs,
                                                      const int32_t* __restrict__ list, const T* __restrict__ dense,
                                                      const T* __restrict__ b, TO* __restrict__ x, const CgState* st)
{
    __shared__ T sb[1024];
    const LevelProb& P = probs[list[blockIdx.x]];
    if (st[P.prob].done) return;
    const int nw = P.nx + 1, n = nw * (P.ny + 1);
    for (int m = threadIdx.x; m < n; m += blockDim.x) {
        const int i = m / nw, j = m - i * nw;
        sb[m] = b[P.origin + (int64_t)i * P.pitch + j];
    }
    __syncthreads();
    const T* A = dense + P.dense_off;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        T s = (T)0;
        const T* row = A + (int64_t)k * n;
        for (int m = 0; m < n; ++m) s += row[m] * sb[m];
        const int i = k / nw, j = k - i * nw;
        x[P.origin + (int64_t)i * P.pitch + j] = (TO)s;
    }
}

// ------------------------------------------------------------ diagnostics

// mode 0: y = A x (Dirichlet rows identity); mode 1: y = b(x) (Dirichlet rows = g)
__global__ __launch_bounds__(kBlock) void k_apply_full(LevelView L, int pi_only, const double* __restrict__ x,
                                                       double* __restrict__ y, int mode)
{
    NodeCtx c;
    if (!decode(L, c)) return;
    if (c.pi != pi_only || !c.in) return;
    const LevelProb& P = *c.P;
    double vc, vw, ve, vs, vn, vsw, vne;
    load7(x, c.idx, P.pitch, vc, vw, ve, vs, vn, vsw, vne);
    if (is_dirichlet(P, c.i, c.j)) {
        y[c.idx] = (mode == 0) ? vc : dirichlet_value(P, c.i, c.j);
        return;
    }
    if (mode == 0) {
        const Row<double> rc = row_coefs<double>(P, c.i, c.j);
        y[c.idx] = row_apply(rc, vc, vw, ve, vs, vn, vsw, vne);
    } else {
        y[c.idx] = rhs_row(P, c.i, c.j, vc, vw, ve, vs, vn, vsw, vne);
    }
}

// ------------------------------------------------------------ boundary flux

// -int grad(u).n ds over all walls with the P1 gradient of the triangle that
// owns each boundary edge (fenics/boundary.h:2652-2742), times dt*D
// (src/fHSL.cpp:156-160).  One block, one problem.
__global__ __launch_bounds__(256) void k_flux(const LevelProb* __restrict__ probs, int pi,
                                              const double* __restrict__ u, double* out)
{
    const LevelProb& P = probs[pi];
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
    block_sum<1>(v);
    if (threadIdx.x == 0) *out = P.Ddt * v[0];
}

// ------------------------------------------------------------ profiling

void prof_begin(eq_batch* b, int kernel)
{
    ProfileSlot& s = b->prof[kernel];
    if (!s.enabled) return;
    std::pair<hipEvent_t, hipEvent_t> ev;
    if (!s.pool.empty()) {
        ev = s.pool.back();
        s.pool.pop_back();
    } else {
        hipEventCreate(&ev.first);
        hipEventCreate(&ev.second);
    }
    hipEventRecord(ev.first, b->stream);
    s.pending.push_back(ev);
}

void prof_end(eq_batch* b, int kernel, double bytes)
{
    ProfileSlot& s = b->prof[kernel];
    if (!s.enabled) return;
    hipEventRecord(s.pending.back().second, b->stream);
    s.pending_bytes.push_back(bytes);
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
        P.mh0 = area6;
        P.mo0 = hx * hy / 24.0;
        P.src = d.dt * d.f * hx * hy / 6.0;
    }
    P.cidx = -1;
    P.fidx = -1;
}

// Dense inverse of the eliminated operator on a coarsest level (FP64 Gauss-Jordan
// with partial pivoting; SPD so it is well posed).
static bool dense_inverse(const LevelProb& P, std::vector<double>& inv)
{
    const int nw = P.nx + 1, n = nw * (P.ny + 1);
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i <= P.ny; ++i)
        for (int j = 0; j <= P.nx; ++j) {
            const int k = j + i * nw;
            if (is_dirichlet(P, i, j)) {
                A[(size_t)k * n + k] = 1.0;
                continue;
            }
            const Row<double> r = row_coefs<double>(P, i, j);
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
            put(i - 1, j - 1, r.msw);
            put(i + 1, j + 1, r.mne);
        }
    inv.assign((size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k) inv[(size_t)k * n + k] = 1.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(A[(size_t)r * n + col]) > std::fabs(A[(size_t)piv * n + col])) piv = r;
        if (A[(size_t)piv * n + col] == 0.0) return false;
        if (piv != col)
            for (int c = 0; c < n; ++c) {
                std::swap(A[(size_t)piv * n + c], A[(size_t)col * n + c]);
                std::swap(inv[(size_t)piv * n + c], inv[(size_t)col * n + c]);
            }
        const double d = A[(size_t)col * n + col];
        for (int c = 0; c < n; ++c) {
            A[(size_t)col * n + c] /= d;
            inv[(size_t)col * n + c] /= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = A[(size_t)r * n + col];
            if (f == 0.0) continue;
            for (int c = 0; c < n; ++c) {
                A[(size_t)r * n + c] -= f * A[(size_t)col * n + c];
                inv[(size_t)r * n + c] -= f * inv[(size_t)col * n + c];
            }
        }
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

static int validate(const eq_field_desc& d, std::string& why)
{
    if (d.nx < 1 || d.ny < 1) { why = "nx and ny must be >= 1"; return 1; }
    if ((int64_t)(d.nx + 1) * (d.ny + 1) > (int64_t)1 << 30) { why = "grid too large"; return 1; }
    if (!(d.width > 0) || !(d.height > 0)) { why = "width and height must be > 0"; return 1; }
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

template <typename T>
static int build_hierarchy(eq_batch* b)
{
    const eq_solver_desc& S = b->solver;
    const bool mg = S.kind != EQ_SOLVER_JACOBI_CG;
    const int maxlev = mg ? std::max(1, S.mg_max_levels) : 1;
    // sizes per problem per level
    std::vector<std::vector<std::pair<int, int>>> dims(b->nprob);
    int nlev = 1;
    for (int p = 0; p < b->nprob; ++p) {
        int nx = b->descs[p].nx, ny = b->descs[p].ny;
        dims[p].push_back({nx, ny});
        while (mg && (int)dims[p].size() < maxlev &&
               (int64_t)(nx + 1) * (ny + 1) > (int64_t)S.mg_coarse_max_nodes) {
            const int cx = nx > 2 ? (nx + 1) / 2 : nx, cy = ny > 2 ? (ny + 1) / 2 : ny;
            if (cx == nx && cy == ny) break;
            nx = cx;
            ny = cy;
            dims[p].push_back({nx, ny});
        }
        if (mg && (int64_t)(nx + 1) * (ny + 1) > 1024)
            return set_error(EQ_ERR_INVALID, "coarsest multigrid level exceeds 1024 nodes; raise mg_max_levels");
        nlev = std::max(nlev, (int)dims[p].size());
    }
    b->levels.assign(nlev, Level());
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
            if (!P.coarsest) {
                P.nxc = dims[p][l + 1].first;
                P.nyc = dims[p][l + 1].second;
            }
            const int64_t rows = P.ny + 3;
            P.origin = off + kRowPad + P.pitch;
            off += round_up(kRowPad + rows * P.pitch + kRowPad, kRowPad);
            P.tiles_x = (P.nx + 1 + kTileW - 1) / kTileW;
            P.tiles_y = (P.ny + 1 + kTileH - 1) / kTileH;
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
            for (int t = 0; t < P.ntiles; ++t) {
                L.tiles_all.push_back(TileRef{(int32_t)k, t});
                if (!P.coarsest) L.tiles_smooth.push_back(TileRef{(int32_t)k, t});
            }
            const double nodes = (double)(P.nx + 1) * (double)(P.ny + 1);
            L.nodes_all += nodes;
            if (!P.coarsest) L.nodes_smooth += nodes;
            if (P.coarsest && mg) {
                std::vector<double> inv;
                if (!dense_inverse(P, inv)) return set_error(EQ_ERR_INVALID, "singular coarse operator");
                P.dense_off = (int64_t)dense_all.size();
                for (double v : inv) dense_all.push_back((T)v);
                L.coarse.push_back((int32_t)k);
                L.max_dense = std::max<int64_t>(L.max_dense, (int64_t)(P.nx + 1) * (P.ny + 1));
            }
        }
    }
    // device tables and vectors
    hipStream_t s = b->stream;
    for (int l = 0; l < nlev; ++l) {
        Level& L = b->levels[l];
        EQ_HIP(hipMalloc((void**)&L.d_probs, sizeof(LevelProb) * L.probs.size()));
        EQ_HIP(hipMemcpyAsync(L.d_probs, L.probs.data(), sizeof(LevelProb) * L.probs.size(), hipMemcpyHostToDevice, s));
        EQ_HIP(hipMalloc((void**)&L.d_tiles_all, sizeof(TileRef) * std::max<size_t>(1, L.tiles_all.size())));
        EQ_HIP(hipMemcpyAsync(L.d_tiles_all, L.tiles_all.data(), sizeof(TileRef) * L.tiles_all.size(),
                              hipMemcpyHostToDevice, s));
        EQ_HIP(hipMalloc((void**)&L.d_tiles_smooth, sizeof(TileRef) * std::max<size_t>(1, L.tiles_smooth.size())));
        if (!L.tiles_smooth.empty())
            EQ_HIP(hipMemcpyAsync(L.d_tiles_smooth, L.tiles_smooth.data(), sizeof(TileRef) * L.tiles_smooth.size(),
                                  hipMemcpyHostToDevice, s));
        EQ_HIP(hipMalloc((void**)&L.d_coarse, sizeof(int32_t) * std::max<size_t>(1, L.coarse.size())));
        if (!L.coarse.empty())
            EQ_HIP(hipMemcpyAsync(L.d_coarse, L.coarse.data(), sizeof(int32_t) * L.coarse.size(),
                                  hipMemcpyHostToDevice, s));
        if (mg) {
            int rc;
            if (l > 0 || b->mixed) {
                if ((rc = alloc_zero((T**)&L.b, L.pool, s))) return rc;
            }
            if ((rc = alloc_zero((T**)&L.xa, L.pool, s))) return rc;
            if ((rc = alloc_zero((T**)&L.xb, L.pool, s))) return rc;
            if ((rc = alloc_zero((T**)&L.r, L.pool, s))) return rc;
        }
    }
    if (!dense_all.empty()) {
        int rc;
        if ((rc = alloc_zero((T**)&b->dense, (int64_t)dense_all.size(), s))) return rc;
        EQ_HIP(hipMemcpyAsync(b->dense, dense_all.data(), sizeof(T) * dense_all.size(), hipMemcpyHostToDevice, s));
    }
    const int64_t n0 = b->levels[0].pool;
    int rc;
    if ((rc = alloc_zero(&b->u, n0, s)) || (rc = alloc_zero(&b->rhs, n0, s)) || (rc = alloc_zero(&b->r, n0, s)) ||
        (rc = alloc_zero(&b->p, n0, s)) || (rc = alloc_zero(&b->q, n0, s)) || (rc = alloc_zero(&b->z, n0, s)) ||
        (rc = alloc_zero(&b->partial, 4 * (int64_t)b->levels[0].tiles_all.size(), s)))
        return rc;
    EQ_HIP(hipStreamSynchronize(s));
    return EQ_OK;
}

// ------------------------------------------------------------ V-cycle

template <typename T>
struct VCycle {
    eq_batch* b;
    // returns the buffer holding the level's result (level 0: z)
    void* run(int l, const void* bl_in)
    {
        const eq_solver_desc& S = b->solver;
        Level& L = b->levels[l];
        const T omega = (T)S.mg_omega;
        hipStream_t s = b->stream;
        const CgState* st = b->d_state;
        const bool top = (l == 0);
        const int ns = (int)L.tiles_smooth.size();
        const int gs = grid_for(ns);
        // final buffer of this level after post-smoothing (z on level 0)
        T* X[2] = {(T*)L.xa, (T*)L.xb};
        int cur = 0;
        const T* bl = (const T*)bl_in;
        if (ns > 0) {
            // pre-smoothing (first sweep from x = 0); level 0 reads the FP64
            // residual and, in mixed precision, keeps an FP32 copy of it
            const double fine_bytes = L.nodes_smooth; // nodes; times bytes per node below
            if (top) {
                prof_begin(b, EQ_KERNEL_SMOOTH);
                if (b->mixed)
                    k_smooth<T, double, T><<<gs, kBlock, 0, s>>>(L.smooth(), (const double*)bl_in, nullptr, X[0],
                                                                 (T*)L.b, omega, 1, st);
                else
                    k_smooth<T, double, T><<<gs, kBlock, 0, s>>>(L.smooth(), (const double*)bl_in, nullptr, X[0],
                                                                 nullptr, omega, 1, st);
                prof_end(b, EQ_KERNEL_SMOOTH, fine_bytes * (sizeof(double) + sizeof(T) * (b->mixed ? 2 : 1)));
                bl = b->mixed ? (const T*)L.b : (const T*)bl_in;
            } else {
                k_smooth<T, T, T><<<gs, kBlock, 0, s>>>(L.smooth(), bl, nullptr, X[0], nullptr, omega, 1, st);
            }
            for (int k = 1; k < S.mg_pre; ++k) {
                if (top) prof_begin(b, EQ_KERNEL_SMOOTH);
                k_smooth<T, T, T><<<gs, kBlock, 0, s>>>(L.smooth(), bl, X[cur], X[cur ^ 1], nullptr, omega, 0, st);
                if (top) prof_end(b, EQ_KERNEL_SMOOTH, fine_bytes * 3 * sizeof(T));
                cur ^= 1;
            }
            k_residual<T><<<gs, kBlock, 0, s>>>(L.smooth(), bl, X[cur], (T*)L.r, st);
            Level& C = b->levels[l + 1];
            const int nc = (int)C.tiles_all.size();
            k_restrict<T><<<grid_for(nc), kBlock, 0, s>>>(C.all(), L.d_probs, (const T*)L.r, (T*)C.b, st);
            const T* xc = (const T*)run(l + 1, C.b);
            k_prolong<T><<<gs, kBlock, 0, s>>>(L.smooth(), C.d_probs, xc, X[cur], st);
            for (int k = 0; k < S.mg_post; ++k) {
                const bool last = (k == S.mg_post - 1);
                if (top && last) {
                    prof_begin(b, EQ_KERNEL_SMOOTH);
                    k_smooth<T, T, double><<<gs, kBlock, 0, s>>>(L.smooth(), bl, X[cur], b->z, nullptr, omega, 0, st);
                    prof_end(b, EQ_KERNEL_SMOOTH, fine_bytes * (2 * sizeof(T) + sizeof(double)));
                } else {
                    if (top) prof_begin(b, EQ_KERNEL_SMOOTH);
                    k_smooth<T, T, T><<<gs, kBlock, 0, s>>>(L.smooth(), bl, X[cur], X[cur ^ 1], nullptr, omega, 0,
                                                            st);
                    if (top) prof_end(b, EQ_KERNEL_SMOOTH, fine_bytes * 3 * sizeof(T));
                    cur ^= 1;
                }
            }
        }
        void* out = top ? (void*)b->z : (void*)X[cur];
        if (!L.coarse.empty()) {
            if (top) {
                // single-level problems: the dense solve reads the residual directly
                if (b->mixed) {
                    // FP32 copy of the residual for the coarse problems
                    k_smooth<T, double, T><<<grid_for((int)L.tiles_all.size()), kBlock, 0, s>>>(
                        L.all(), (const double*)bl_in, nullptr, X[1], (T*)L.b, (T)0, 1, st);
                    k_coarse_dense<T, double><<<(int)L.coarse.size(), 256, 0, s>>>(
                        L.d_probs, L.d_coarse, (const T*)b->dense, (const T*)L.b, b->z, st);
                } else {
                    k_coarse_dense<T, double><<<(int)L.coarse.size(), 256, 0, s>>>(
                        L.d_probs, L.d_coarse, (const T*)b->dense, (const T*)bl_in, b->z, st);
                }
            } else {
                k_coarse_dense<T, T><<<(int)L.coarse.size(), 256, 0, s>>>(L.d_probs, L.d_coarse, (const T*)b->dense,
                                                                          (const T*)bl_in, (T*)out, st);
            }
        }
        return out;
    }
};

// ------------------------------------------------------------ solve driver

static int read_states(eq_batch* b)
{
    EQ_HIP(hipMemcpyAsync(b->h_state, b->d_state, sizeof(CgState) * b->nprob, hipMemcpyDeviceToHost, b->stream));
    EQ_HIP(hipStreamSynchronize(b->stream));
    return EQ_OK;
}

template <typename T>
static int solve_all(eq_batch* b, eq_step_stats* stats)
{
    const auto t0 = std::chrono::steady_clock::now();
    const eq_solver_desc& S = b->solver;
    Level& L0 = b->levels[0];
    const LevelView A = L0.all();
    const int g0 = grid_for(A.ntiles);
    hipStream_t s = b->stream;
    const bool jacobi = (S.kind == EQ_SOLVER_JACOBI_CG);
    const double fine_nodes = L0.nodes_all;
    VCycle<T> vc{b};
    auto finish = [&](int op) {
        k_finish<<<b->nprob, kBlock, 0, s>>>(L0.d_probs, b->partial, b->d_state, op, jacobi ? 1 : 0);
    };
    auto precondition = [&]() {
        vc.run(0, b->r);
        k_rz<<<g0, kBlock, 0, s>>>(A, b->r, b->z, b->partial, b->d_state);
        finish(kFinRz);
    };
    auto iteration = [&]() {
        prof_begin(b, EQ_KERNEL_STENCIL64);
        k_ap<<<g0, kBlock, 0, s>>>(A, b->p, b->q, b->partial, b->d_state);
        prof_end(b, EQ_KERNEL_STENCIL64, fine_nodes * 2 * sizeof(double));
        finish(kFinAp);
        prof_begin(b, EQ_KERNEL_CG_UPDATE);
        k_update<<<g0, kBlock, 0, s>>>(A, b->u, b->r, b->z, b->p, b->q, b->partial, b->d_state, jacobi ? 1 : 0);
        prof_end(b, EQ_KERNEL_CG_UPDATE, fine_nodes * sizeof(double) * (jacobi ? 7 : (b->mixed ? 7 : 6)));
        finish(kFinUpdate);
        if (!jacobi) precondition();
        k_dir<<<g0, kBlock, 0, s>>>(A, b->z, b->p, b->d_state);
    };

    k_rhs<<<g0, kBlock, 0, s>>>(A, b->u, b->rhs, b->partial, b->d_state);
    finish(kFinRhs);
    k_init<<<g0, kBlock, 0, s>>>(A, b->u, b->rhs, b->r, b->z, b->p, b->partial, b->d_state, jacobi ? 1 : 0);
    finish(kFinInit);
    if (!jacobi) {
        precondition();
        k_dir<<<g0, kBlock, 0, s>>>(A, b->z, b->p, b->d_state);
    }
    EQ_HIP(hipGetLastError());

    int done_iters = 0;
    int chunk = std::max(1, b->last_iters > 0 ? b->last_iters : S.check_every);
    int rc;
    for (int guard = 0; guard < 1000000; ++guard) {
        for (int k = 0; k < chunk; ++k) iteration();
        done_iters += chunk;
        EQ_HIP(hipGetLastError());
        if ((rc = read_states(b))) return rc;
        bool all = true;
        for (int p = 0; p < b->nprob; ++p) all = all && b->h_state[p].done;
        if (all) {
            // confirm on the true residual; problems above rtol resume
            k_true_residual<<<g0, kBlock, 0, s>>>(A, b->u, b->rhs, b->r, b->partial, b->d_state);
            finish(kFinTrue);
            if ((rc = read_states(b))) return rc;
            all = true;
            for (int p = 0; p < b->nprob; ++p) all = all && b->h_state[p].done;
            if (all) break;
            // restart CG from the replaced residual: p = M^-1 r
            if (jacobi) {
                k_jacobi_z<<<g0, kBlock, 0, s>>>(A, b->r, b->z, b->d_state);
                k_rz<<<g0, kBlock, 0, s>>>(A, b->r, b->z, b->partial, b->d_state);
                finish(kFinRz);
            } else {
                precondition();
            }
            k_dir<<<g0, kBlock, 0, s>>>(A, b->z, b->p, b->d_state);
        }
        chunk = std::max(1, S.check_every);
    }
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
    (void)done_iters;
    if (stats) {
        stats->iterations_max = itmax;
        stats->nconverged = nconv;
        stats->relres_max = relmax;
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (nconv < b->nprob)
        {
        char msg[128];
        std::snprintf(msg, sizeof(msg), "max_iter reached above rtol (relres %.3e)", relmax);
        return set_error(EQ_ERR_NOT_CONVERGED, msg);
    }
    return EQ_OK;
}

} // namespace eqk

using namespace eqk;

// ------------------------------------------------------------ C ABI

extern "C" {

const char* eq_last_error(void) { return g_last_error.c_str(); }

void eq_solver_default(eq_solver_desc* s)
{
    if (!s) return;
    std::memset(s, 0, sizeof(*s));
    s->kind = EQ_SOLVER_MIXED_GMG_CG;
    s->max_iter = 20000;
    s->rtol = 1e-10;
    s->mg_pre = 2;
    s->mg_post = 2;
    s->mg_omega = 0.8;
    s->mg_coarse_max_nodes = 256;
    s->mg_max_levels = 16;
    s->check_every = 4;
    s->use_graphs = 0;
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
    if (S.kind < EQ_SOLVER_JACOBI_CG || S.kind > EQ_SOLVER_MIXED_GMG_CG) return set_error(EQ_ERR_INVALID, "unknown solver");
    if (!(S.rtol > 0) || S.max_iter < 1 || S.check_every < 1) return set_error(EQ_ERR_INVALID, "bad solver tolerances");
    if (S.kind != EQ_SOLVER_JACOBI_CG &&
        (S.mg_pre < 1 || S.mg_post < 1 || !(S.mg_omega > 0) || S.mg_coarse_max_nodes < 4 || S.mg_max_levels < 1))
        return set_error(EQ_ERR_INVALID, "bad multigrid parameters");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(EQ_ERR_NO_DEVICE, "no HIP device");
    eq_batch* b = new (std::nothrow) eq_batch();
    if (!b) return set_error(EQ_ERR_NOMEM, "host allocation");
    hipGetDevice(&b->device);
    b->nprob = nprob;
    b->nspecies = nspecies;
    b->ntraps = nprob / nspecies;
    b->descs.assign(descs, descs + nprob);
    b->solver = S;
    b->mixed = (S.kind == EQ_SOLVER_MIXED_GMG_CG) ? 1 : 0;
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
    for (Level& L : b->levels) {
        hipFree(L.d_probs);
        hipFree(L.d_tiles_all);
        hipFree(L.d_tiles_smooth);
        hipFree(L.d_coarse);
        hipFree(L.b);
        hipFree(L.xa);
        hipFree(L.xb);
        hipFree(L.r);
    }
    hipFree(b->u);
    hipFree(b->rhs);
    hipFree(b->r);
    hipFree(b->p);
    hipFree(b->q);
    hipFree(b->z);
    hipFree(b->partial);
    hipFree(b->dense);
    hipFree(b->d_state);
    hipFree(b->flux_out);
    hipFree(b->d_field_ptrs);
    hipFree(b->cell_scratch);
    if (b->h_state) hipHostFree(b->h_state);
    if (b->h_flux) hipHostFree(b->h_flux);
    for (int k = 0; k < EQ_KERNEL_COUNT_; ++k) {
        for (auto& e : b->prof[k].pool) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
        for (auto& e : b->prof[k].pending) { hipEventDestroy(e.first); hipEventDestroy(e.second); }
    }
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
    return b->mixed ? solve_all<float>(b, stats) : solve_all<double>(b, stats);
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
    k_apply_full<<<grid_for(A.ntiles), kBlock, 0, s>>>(A, prob, b->q, b->z, mode);
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
    k_flux<<<1, 256, 0, b->stream>>>(b->levels[0].d_probs, prob, b->u, b->flux_out);
    EQ_HIP(hipGetLastError());
    EQ_HIP(hipMemcpyAsync(b->h_flux, b->flux_out, sizeof(double), hipMemcpyDeviceToHost, b->stream));
    EQ_HIP(hipStreamSynchronize(b->stream));
    *total = *b->h_flux;
    return EQ_OK;
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
