// eq_common.h -- shared host/device definitions of the MI355X field step.
//
// The operator is the P1 form of eQ's hslD problem (fenics/hslD.ufl:31-49)
// on the structured "right"-diagonal RectangleMesh (src/fHSL.cpp:164-173),
// written as a matrix-free 7-point stencil whose entries are the closed form
// of the FFC element tensors (fenics/hslD.h:3123-3259, 3284-3441) summed over
// the (up to) six triangles around a node:
//
//   square cell (i,j) has lower-right triangle (BL,BR,TR) and upper-left
//   triangle (BL,TL,TR); per triangle of legs hx, hy:
//     mass  |T|/6 on the diagonal, |T|/12 off it          (|T| = hx*hy/2)
//     stiff dt*D*hy/(2hx) along x legs, dt*D*hx/(2hy) along y legs, and 0
//           between the two ends of the hypotenuse (the right angle makes
//           that cotangent vanish).
//
// The stiffness part is evaluated in difference form
//   sum_nb k_nb * (v_c - v_nb)
// (K has zero row sums), so a residual b - A v keeps its accuracy when
// dt*D/h^2 is large: the plain form loses ~log10(8*dt*D/h^2) digits to
// cancellation, the difference form only what the differences carry.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define EQ_HD __host__ __device__
#else
#define EQ_HD
#endif

namespace eqk {

constexpr int kTileW = 64;  // one wavefront spans the tile's columns
constexpr int kWaves = 4;   // waves per 256-thread block
constexpr int kBlock = kTileW * kWaves;
// consecutive rows per thread: four on the finest level (amortises each
// block's set-up over more nodes and keeps more loads in flight), one on the
// coarse levels (too few nodes there to give up parallelism)
constexpr int kRows0 = 4;
constexpr int kRowsC = 1;
// coarse levels with at least this many nodes (over the batch) take level 0's
// tile shape: they are bandwidth-bound, the smaller ones latency-bound
constexpr double kWideRowsMinNodes = 8.0 * 1024.0 * 1024.0;
// levels with at least this many nodes take 8 rows per thread (64 x 32 tiles):
// more reuse per load and one round of blocks, where the launch is large
// enough for that to pay (config 3 and 4's two finest levels, not a 2000 x 1000 trap)
constexpr int kRowsWide = 8;
constexpr double kWideRows8MinNodes = 32.0 * 1024.0 * 1024.0;
constexpr int kRowPad = 64; // pitch granularity in elements

// walls in eQ::boundaryCondition::wall order (src/eQ.h:335-340)
enum { kLeft = 0, kRight = 1, kTop = 2, kBottom = 3 };
enum { kNeumann = 0, kDirichlet = 1, kRobin = 2 };

template <typename T>
struct Row {
    T mc, mw, me, ms, mn, msw, mne; // mass-like couplings (mass, reaction, Robin)
    T kw, ke, ks, kn;               // stiffness difference weights
    T ksw, kne;                     // diagonal stiffness weights (0 unless D12 != 0)
};

// One problem (species x trap) on one multigrid level.
struct LevelProb {
    int32_t nx, ny;        // cells
    int32_t pitch;         // elements between rows
    int32_t prob;          // global problem index
    int64_t origin;        // element offset of node (0,0) inside the level pool
    int64_t dense_off;     // coarsest only: element offset of the dense inverse
    int32_t tile_begin;    // first tile of this problem in the level's tile table
    int32_t tiles_x, tiles_y, ntiles;
    int32_t dmask;         // bit w set: wall w is Dirichlet
    int32_t coarsest;      // 1 if this level is the problem's coarsest
    int32_t cidx;          // index of the problem in the next level's table (-1 if none)
    int32_t nxc, nyc;      // next level's cells (0 if none)
    int32_t fidx;          // index of the problem in the previous (finer) level's table (-1 on level 0)
    int32_t bottom;        // 1: from this level down, one workgroup runs the V-cycle
    int32_t below;         // 1: level lies under the problem's bottom level (no grid launches)
    int32_t lds_off;       // bottom levels: element offset of this level's LDS arrays
    int32_t lds_total;     // bottom level: LDS elements the problem's bottom V-cycle needs
    // left-hand side (all levels)
    double mh, mo;         // (1+dt*gamma)*hx*hy/12 and /24: mass per triangle
    double sa, sb;         // dt*D*hy/(2*hx), dt*D*hx/(2*hy): stiffness per triangle
    double rob[4];         // dt*r*|e| per Robin wall edge, 0 elsewhere
    // right-hand side (level 0 only)
    double mh0, mo0;       // hx*hy/12, hx*hy/24 (no gamma on the right)
    double src;            // dt*f*hx*hy/6 per adjacent triangle
    double robs[4];        // dt*r*s*|e|/2 per Robin wall edge endpoint
    double g[4];           // Dirichlet value per wall
    const double* gw[4];   // per-node Dirichlet values along a wall (level 0; NULL: g)
    double hx, hy;         // spacing (boundary flux)
    double Ddt;            // dt*D (boundary flux scaling)
    // the row every interior node shares (0 < i < ny, 0 < j < nx), in both
    // precisions, and the reciprocal of its diagonal
    Row<double> rin;
    Row<float> rinf;
    double rin_dinv;
    float rinf_dinv;
    // 1/nx, 1/ny (grid transfers: hat arguments and index division)
    double rnx, rny;
    float rnxf, rnyf;
    // per-node diffusion tensor (D11, D22, D12 in natural order, P1 in each
    // triangle; fenics/hslD.ufl:31): arrays of the level's pool layout,
    // D11 at tens[k], D22 at tens[k + tstride], D12 at tens[k + 2 tstride]
    int32_t aniso;         // 1: the tensor arrays below define the stiffness
    int32_t fd;            // 1: 5-point finite-difference scheme (diffusionPETSc)
    double sxy;            // dt*D/2: stiffness per triangle for the mixed derivative
    double fdm;            // FD: (1+dt*gamma)*hx*hy, the scaled identity row
    double fkx, fky;       // FD: dt*D*hy/hx, dt*D*hx/hy (x and y couplings, interior)
    double frob[4];        // FD: 2*dt*r*(length of the wall's node segment) per Robin wall
    double fsrc[4];        // FD: the matching right-hand side 2*dt*r*s*(...)
    double fsrcf;          // FD: dt*f*hx*hy
    const double* tens_d;  // FP64 copy (level 0)
    const float* tens_f;   // FP32 copy (levels smoothed in FP32)
    int64_t tstride;
    // advection: x-velocity per grid row (ny+1 values, P1 in y), Galerkin
    // convection dt * int (v du/dx) phi; cw = dt*hx*hy/(24 hx) = dt*hy/24
    int32_t vel;           // 1: the row velocities below are set
    int32_t celltens;      // 1: the tensor comes from the cells' angles (refreshed by the coupling)
    double cw;
    const double* vel_d;   // FP64 rows (level 0)
    const float* vel_f;    // FP32 rows (levels smoothed in FP32)
};

// (i,j) must be a node of the grid
EQ_HD inline bool is_dirichlet(const LevelProb& P, int i, int j)
{
    return ((P.dmask & 1) && j == 0) || ((P.dmask & 2) && j == P.nx) ||
           ((P.dmask & 4) && i == P.ny) || ((P.dmask & 8) && i == 0);
}

// The same for a stencil neighbour that may lie one node outside the grid:
// those are never Dirichlet (dirichlet_value indexes per-node wall arrays by
// the position along the wall, which must lie on the wall).
EQ_HD inline bool is_dirichlet_nb(const LevelProb& P, int i, int j)
{
    if (i < 0 || j < 0 || i > P.ny || j > P.nx) return false;
    return is_dirichlet(P, i, j);
}

// Dirichlet value of a node; walls are applied left, right, top, bottom as
// fenicsInterface::createHSL pushes them (src/fHSL.cpp:476-537), so a corner
// on two Dirichlet walls keeps the later wall's value.
EQ_HD inline double dirichlet_value(const LevelProb& P, int i, int j)
{
    double v = 0.0;
    if (P.fd) {
        // diffusionPETSc::ApplyBoundaryConditions sets top/bottom first, then
        // right/left (diffuclass.cpp:221-290): left/right win at corners
        if ((P.dmask & 4) && i == P.ny) v = P.gw[kTop] ? P.gw[kTop][j] : P.g[kTop];
        if ((P.dmask & 8) && i == 0) v = P.gw[kBottom] ? P.gw[kBottom][j] : P.g[kBottom];
        if ((P.dmask & 2) && j == P.nx) v = P.gw[kRight] ? P.gw[kRight][i] : P.g[kRight];
        if ((P.dmask & 1) && j == 0) v = P.gw[kLeft] ? P.gw[kLeft][i] : P.g[kLeft];
        return v;
    }
    if ((P.dmask & 1) && j == 0) v = P.gw[kLeft] ? P.gw[kLeft][i] : P.g[kLeft];
    if ((P.dmask & 2) && j == P.nx) v = P.gw[kRight] ? P.gw[kRight][i] : P.g[kRight];
    if ((P.dmask & 4) && i == P.ny) v = P.gw[kTop] ? P.gw[kTop][j] : P.g[kTop];
    if ((P.dmask & 8) && i == 0) v = P.gw[kBottom] ? P.gw[kBottom][j] : P.g[kBottom];
    return v;
}

// FD row of a free node (diffusionPETSc's ghost-reflection rows times
// hx*hy*wx*wy, wx = 1/2 on the left/right walls, wy = 1/2 on top/bottom:
// the scaled system is symmetric and has the same solution)
template <typename T>
EQ_HD inline Row<T> fd_row(const LevelProb& P, int i, int j)
{
    const T one = (T)1, half = (T)0.5, zero = (T)0;
    const bool L = j == 0, R = j == P.nx, B = i == 0, Tp = i == P.ny;
    const T wx = (L || R) ? half : one, wy = (B || Tp) ? half : one;
    Row<T> r;
    double m = P.fdm;
    if (L) m += P.frob[kLeft];
    if (R) m += P.frob[kRight];
    if (B) m += P.frob[kBottom];
    if (Tp) m += P.frob[kTop];
    r.mc = wx * wy * (T)m;
    r.mw = r.me = r.ms = r.mn = r.msw = r.mne = zero;
    const T kx = wy * (T)P.fkx, ky = wx * (T)P.fky;
    r.kw = L ? zero : kx;
    r.ke = R ? zero : kx;
    r.ks = B ? zero : ky;
    r.kn = Tp ? zero : ky;
    r.ksw = r.kne = zero;
    return r;
}


EQ_HD inline bool interior(const LevelProb& P, int i, int j) { return i > 0 && j > 0 && i < P.ny && j < P.nx; }

template <typename T> EQ_HD inline const Row<T>& interior_row(const LevelProb& P);
template <> EQ_HD inline const Row<double>& interior_row<double>(const LevelProb& P) { return P.rin; }
template <> EQ_HD inline const Row<float>& interior_row<float>(const LevelProb& P) { return P.rinf; }
template <typename T> EQ_HD inline T interior_dinv(const LevelProb& P);
template <> EQ_HD inline double interior_dinv<double>(const LevelProb& P) { return P.rin_dinv; }
template <> EQ_HD inline float interior_dinv<float>(const LevelProb& P) { return P.rinf_dinv; }

// Row of A = (1+dt*gamma)M + dt*D*K + Robin at node (i,j) from the element
// counts around it (Dirichlet rows are handled by the callers), evaluated in
// the precision T of the caller (the FP32 smoother builds FP32 rows).
template <typename T>
EQ_HD inline Row<T> row_coefs_at(const LevelProb& P, int i, int j)
{
    if (P.fd) return fd_row<T>(P, i, j);
    const T one = (T)1, zero = (T)0, two = (T)2;
    const T mh = (T)P.mh, mo = (T)P.mo;
    const T sw = (i > 0 && j > 0) ? one : zero;
    const T se = (i > 0 && j < P.nx) ? one : zero;
    const T nw = (i < P.ny && j > 0) ? one : zero;
    const T ne = (i < P.ny && j < P.nx) ? one : zero;
    Row<T> r;
    r.mc = (two * sw + se + nw + two * ne) * mh;
    r.mw = (sw + nw) * mo;
    r.me = (se + ne) * mo;
    r.ms = (sw + se) * mo;
    r.mn = (nw + ne) * mo;
    r.msw = two * sw * mo;
    r.mne = two * ne * mo;
    r.kw = (sw + nw) * (T)P.sa;
    r.ke = (se + ne) * (T)P.sa;
    r.ks = (sw + se) * (T)P.sb;
    r.kn = (nw + ne) * (T)P.sb;
    r.ksw = zero;
    r.kne = zero;
    // Robin wall edges: dt*r*|e|*[[1/3,1/6],[1/6,1/3]]  (fenics/hslD.h:3284-3441)
    const T third = (T)(1.0 / 3.0), sixth = (T)(1.0 / 6.0);
    const T dn = (i > 0) ? one : zero, up = (i < P.ny) ? one : zero;
    const T lf = (j > 0) ? one : zero, rt = (j < P.nx) ? one : zero;
    if (j == 0 && P.rob[kLeft] != 0.0) {
        const T q = (T)P.rob[kLeft];
        r.mc += (dn + up) * q * third;
        r.ms += dn * q * sixth;
        r.mn += up * q * sixth;
    }
    if (j == P.nx && P.rob[kRight] != 0.0) {
        const T q = (T)P.rob[kRight];
        r.mc += (dn + up) * q * third;
        r.ms += dn * q * sixth;
        r.mn += up * q * sixth;
    }
    if (i == 0 && P.rob[kBottom] != 0.0) {
        const T q = (T)P.rob[kBottom];
        r.mc += (lf + rt) * q * third;
        r.mw += lf * q * sixth;
        r.me += rt * q * sixth;
    }
    if (i == P.ny && P.rob[kTop] != 0.0) {
        const T q = (T)P.rob[kTop];
        r.mc += (lf + rt) * q * third;
        r.mw += lf * q * sixth;
        r.me += rt * q * sixth;
    }
    return r;
}

// TENS = false: no problem of the launch has a tensor field, so the diagonal
// stiffness weights are zero and their terms are skipped (same values)
template <bool TENS = true, typename T>
EQ_HD inline T row_diag(const Row<T>& r)
{
    if constexpr (TENS) return r.mc + ((r.kw + r.ke) + (r.ks + r.kn)) + (r.ksw + r.kne);
    return r.mc + ((r.kw + r.ke) + (r.ks + r.kn));
}

// Stiffness of node (i,j) from a per-node tensor field: the six triangles
// around the node, each with the mean of its three vertex tensors (the P1
// tensor integrated exactly, fenics/hslD.h:3123-3259).  Per triangle of legs
// hx, hy (area hx*hy/2) the element matrix dt*D*|T| g_a^T Dbar g_b gives, in
// difference form k = -K_ab,
//   x legs  dt*D*(hy/(2hx)) d11 - (dt*D/2) d12     y legs  dt*D*(hx/(2hy)) d22 - (dt*D/2) d12
//   SW-NE hypotenuse (dt*D/2) d12
// d11, d22, d12 point at node (i,j) of arrays with row pitch `pitch`; the
// entries one node outside the grid are read only by triangles that exist.
template <typename T>
EQ_HD inline void aniso_stiffness(const LevelProb& P, int i, int j, const T* d11, const T* d22, const T* d12,
                                  int pitch, Row<T>& r)
{
    const bool sw = i > 0 && j > 0, se = i > 0 && j < P.nx, nw = i < P.ny && j > 0, ne = i < P.ny && j < P.nx;
    const T sa = (T)P.sa, sb = (T)P.sb, sx = (T)P.sxy, third = (T)(1.0 / 3.0);
    // vertex offsets: c, W, E, S, N, SW, NE
    const int oW = -1, oE = 1, oS = -pitch, oN = pitch, oSW = -pitch - 1, oNE = pitch + 1;
    auto mean = [&](const T* a, int o1, int o2) { return (a[0] + a[o1] + a[o2]) * third; };
    T kw = (T)0, ke = (T)0, ks = (T)0, kn = (T)0, ksw = (T)0, kne = (T)0;
    if (sw) {
        // lower-right triangle (SW, S, c) and upper-left (SW, W, c) of the SW cell
        const T b12 = mean(d12, oSW, oS), b22 = mean(d22, oSW, oS);
        ksw += sx * b12;
        ks += sb * b22 - sx * b12;
        const T c12 = mean(d12, oSW, oW), c11 = mean(d11, oSW, oW);
        ksw += sx * c12;
        kw += sa * c11 - sx * c12;
    }
    if (se) {
        // upper-left triangle (S, c, E) of the SE cell
        const T a12 = mean(d12, oS, oE);
        ks += sb * mean(d22, oS, oE) - sx * a12;
        ke += sa * mean(d11, oS, oE) - sx * a12;
    }
    if (nw) {
        // lower-right triangle (W, c, N) of the NW cell
        const T a12 = mean(d12, oW, oN);
        kw += sa * mean(d11, oW, oN) - sx * a12;
        kn += sb * mean(d22, oW, oN) - sx * a12;
    }
    if (ne) {
        // lower-right (c, E, NE) and upper-left (c, N, NE) of the node's own cell
        const T b12 = mean(d12, oE, oNE);
        ke += sa * mean(d11, oE, oNE) - sx * b12;
        kne += sx * b12;
        const T c12 = mean(d12, oN, oNE);
        kn += sb * mean(d22, oN, oNE) - sx * c12;
        kne += sx * c12;
    }
    r.kw = kw; r.ke = ke; r.ks = ks; r.kn = kn; r.ksw = ksw; r.kne = kne;
}

// Convection rows of node (i,j) added to r in difference form (k_nb += -dt
// C[c][nb]; C has zero row sums).  Per triangle around the node with
// vertices (c, a, b): int_T v phi_c = |T|/12 (2 v_c + v_a + v_b) times
// dphi_nb/dx = -+1/hx; with v constant along rows the six triangles give
//   SW cell  lower (SW,S,c): C[SW] -= I/hx, C[S] += I/hx   I ~ 2v_i + 2v_{i-1}
//            upper (SW,W,c): C[W]  -= I/hx                 I ~ 3v_i +  v_{i-1}
//   SE cell  upper (S,c,E):  C[E]  += I/hx                 I ~ 3v_i +  v_{i-1}
//   NW cell  lower (W,c,N):  C[W]  -= I/hx                 I ~ 3v_i +  v_{i+1}
//   own cell lower (c,E,NE): C[E]  += I/hx                 I ~ 3v_i +  v_{i+1}
//            upper (c,N,NE): C[N]  -= I/hx, C[NE] += I/hx  I ~ 2v_i + 2v_{i+1}
// (I = |T|/12 * (...), |T|/(12 hx) = hy/24).  v holds the level's rows.
template <typename T>
EQ_HD inline void add_convection_v(const LevelProb& P, int i, int j, T vs, T vi, T vn, Row<T>& r)
{
    const bool sw = i > 0 && j > 0, se = i > 0 && j < P.nx, nw = i < P.ny && j > 0, ne = i < P.ny && j < P.nx;
    const T cw = (T)P.cw, three = (T)3, two = (T)2;
    if (sw) {
        const T I1 = cw * (two * vi + two * vs), I2 = cw * (three * vi + vs);
        r.ksw += I1;
        r.ks -= I1;
        r.kw += I2;
    }
    if (se) r.ke -= cw * (three * vi + vs);
    if (nw) r.kw += cw * (three * vi + vn);
    if (ne) {
        const T I5 = cw * (three * vi + vn), I6 = cw * (two * vi + two * vn);
        r.ke -= I5;
        r.kn += I6;
        r.kne -= I6;
    }
}

// the same with the level's row velocities v (rows i-1, i, i+1 as they exist)
template <typename T>
EQ_HD inline void add_convection(const LevelProb& P, int i, int j, const T* v, Row<T>& r)
{
    const T vs = i > 0 ? v[i - 1] : (T)0, vn = i < P.ny ? v[i + 1] : (T)0;
    add_convection_v<T>(P, i, j, vs, v[i], vn, r);
}

template <typename T> EQ_HD inline const T* vel_base(const LevelProb& P);
template <> EQ_HD inline const double* vel_base<double>(const LevelProb& P) { return P.vel_d; }
template <> EQ_HD inline const float* vel_base<float>(const LevelProb& P) { return P.vel_f; }

template <typename T> EQ_HD inline const T* tens_base(const LevelProb& P);
template <> EQ_HD inline const double* tens_base<double>(const LevelProb& P) { return P.tens_d; }
template <> EQ_HD inline const float* tens_base<float>(const LevelProb& P) { return P.tens_f; }

// Row of node (i,j), element k of the level pool.  Constant-coefficient
// problems take the interior rows from the per-level table and build the
// rest in place; problems with a tensor field build every row from it.
template <typename T, bool TENS = true>
EQ_HD inline Row<T> row_coefs(const LevelProb& P, int i, int j, int64_t k)
{
    if (TENS && (P.aniso || P.vel)) {
        if (!P.aniso && interior(P, i, j)) {
            // velocity only: the level's constant interior row plus this
            // grid row's convection (no element counting needed)
            Row<T> r = interior_row<T>(P);
            add_convection<T>(P, i, j, vel_base<T>(P), r);
            return r;
        }
        Row<T> r = row_coefs_at<T>(P, i, j);
        if (P.aniso) {
            const T* t = tens_base<T>(P) + k;
            aniso_stiffness<T>(P, i, j, t, t + P.tstride, t + 2 * P.tstride, P.pitch, r);
        }
        if (P.vel) add_convection<T>(P, i, j, vel_base<T>(P), r);
        return r;
    }
    if (interior(P, i, j)) return interior_row<T>(P);
    return row_coefs_at<T>(P, i, j);
}

// 1 / diag(A) at node (i,j) whose row is rc
template <bool TENS = true, typename T>
EQ_HD inline T row_dinv(const LevelProb& P, int i, int j, const Row<T>& rc)
{
    if ((!TENS || !(P.aniso || P.vel)) && interior(P, i, j)) return interior_dinv<T>(P);
    return (T)1 / row_diag<TENS>(rc);
}

// fills the interior row of a level (host)
inline void set_interior_row(LevelProb& P)
{
    LevelProb Q = P;
    Q.nx = 2;
    Q.ny = 2;
    P.rin = row_coefs_at<double>(Q, 1, 1);
    P.rinf = row_coefs_at<float>(Q, 1, 1);
    P.rin_dinv = 1.0 / row_diag(P.rin);
    P.rinf_dinv = 1.0f / row_diag(P.rinf);
    P.rnx = 1.0 / P.nx;
    P.rny = 1.0 / P.ny;
    P.rnxf = 1.0f / (float)P.nx;
    P.rnyf = 1.0f / (float)P.ny;
}

template <typename T> EQ_HD inline T recip_nx(const LevelProb& P);
template <> EQ_HD inline double recip_nx<double>(const LevelProb& P) { return P.rnx; }
template <> EQ_HD inline float recip_nx<float>(const LevelProb& P) { return P.rnxf; }
template <typename T> EQ_HD inline T recip_ny(const LevelProb& P);
template <> EQ_HD inline double recip_ny<double>(const LevelProb& P) { return P.rny; }
template <> EQ_HD inline float recip_ny<float>(const LevelProb& P) { return P.rnyf; }

// floor(a / b) for a >= 0, b > 0, a / b < 2^22 (and a + b < 2^31), from the
// FP32 reciprocal rb ~ 1/b and one exact correction step each way (the FP32
// quotient is off by less than one)
EQ_HD inline int div_floor_pos(int a, int b, float rb)
{
    int q = (int)((float)a * rb);
    if (q * b > a) --q;
    if ((q + 1) * b <= a) ++q;
    return q;
}

// y_c = (A v)_c from the seven values around the node
template <bool TENS = true, typename T>
EQ_HD inline T row_apply(const Row<T>& r, T vc, T vw, T ve, T vs, T vn, T vsw, T vne)
{
    T stiff = (r.kw * (vc - vw) + r.ke * (vc - ve)) + (r.ks * (vc - vs) + r.kn * (vc - vn));
    if constexpr (TENS) stiff = stiff + (r.ksw * (vc - vsw) + r.kne * (vc - vne));
    const T mass = r.mc * vc + ((r.mw * vw + r.me * ve) + (r.ms * vs + r.mn * vn)) + (r.msw * vsw + r.mne * vne);
    return stiff + mass;
}

// Right-hand side row of level 0: b = M u0 + dt*f*int(phi) + Robin dt*r*s*int(phi)
// (fenics/hslD.h:3466-3529, 3554-3689).
EQ_HD inline double rhs_row(const LevelProb& P, int i, int j, double uc, double uw, double ue, double us,
                            double un, double usw, double une)
{
    if (P.fd) {
        // u_old plus the Robin ghost sources, scaled as the FD rows
        const bool L = j == 0, R = j == P.nx, B = i == 0, Tp = i == P.ny;
        const double wx = (L || R) ? 0.5 : 1.0, wy = (B || Tp) ? 0.5 : 1.0;
        double b = P.mh0 * uc + P.fsrcf;
        if (L) b += P.fsrc[kLeft];
        if (R) b += P.fsrc[kRight];
        if (B) b += P.fsrc[kBottom];
        if (Tp) b += P.fsrc[kTop];
        return wx * wy * b;
    }
    const double sw = (i > 0 && j > 0) ? 1.0 : 0.0;
    const double se = (i > 0 && j < P.nx) ? 1.0 : 0.0;
    const double nw = (i < P.ny && j > 0) ? 1.0 : 0.0;
    const double ne = (i < P.ny && j < P.nx) ? 1.0 : 0.0;
    const double ntri = 2.0 * sw + se + nw + 2.0 * ne;
    double b = ntri * P.mh0 * uc + P.mo0 * (((sw + nw) * uw + (se + ne) * ue) + ((sw + se) * us + (nw + ne) * un)) +
               2.0 * P.mo0 * (sw * usw + ne * une);
    b += ntri * P.src;
    const double dn = (i > 0) ? 1.0 : 0.0, up = (i < P.ny) ? 1.0 : 0.0;
    const double lf = (j > 0) ? 1.0 : 0.0, rt = (j < P.nx) ? 1.0 : 0.0;
    if (j == 0) b += (dn + up) * P.robs[kLeft];
    if (j == P.nx) b += (dn + up) * P.robs[kRight];
    if (i == 0) b += (lf + rt) * P.robs[kBottom];
    if (i == P.ny) b += (lf + rt) * P.robs[kTop];
    return b;
}

// Coarse-to-fine transfer: P1 interpolation on the coarse level's "right"
// triangulation at fine node coordinates s = j*nxc/nxf, t = i*nyc/nyf (in
// coarse cells).  The hat function of coarse node (I,J) on that
// triangulation is
//     phi(u, v) = max(0, 1 - max(|u|, |v|, |u - v|)),   u = s - J, v = t - I,
// whose support is the hexagon with corners (+-1,0), (0,+-1), (1,1), (-1,-1).
// Nested (even) levels give the exact finite-element embedding; odd levels
// are non-nested.  Restriction gathers the same weights, so R = P^T exactly.
template <typename T>
EQ_HD inline T hat(int du_num, int dv_num, T inx, T iny)
{
    const T u = (T)du_num * inx, v = (T)dv_num * iny;
    const T au = u < (T)0 ? -u : u, av = v < (T)0 ? -v : v, d = u - v, ad = d < (T)0 ? -d : d;
    T m = au > av ? au : av;
    m = m > ad ? m : ad;
    const T w = (T)1 - m;
    return w > (T)0 ? w : (T)0;
}


} // namespace eqk
