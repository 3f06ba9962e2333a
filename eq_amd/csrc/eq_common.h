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

constexpr int kTileW = 64;  // one wavefront per tile row
constexpr int kTileH = 4;   // rows per tile (4 waves per 256-thread block)
constexpr int kBlock = kTileW * kTileH;
constexpr int kRowPad = 64; // pitch granularity in elements

// walls in eQ::boundaryCondition::wall order (src/eQ.h:335-340)
enum { kLeft = 0, kRight = 1, kTop = 2, kBottom = 3 };
enum { kNeumann = 0, kDirichlet = 1, kRobin = 2 };

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
    double hx, hy;         // spacing (boundary flux)
    double Ddt;            // dt*D (boundary flux scaling)
};

EQ_HD inline bool is_dirichlet(const LevelProb& P, int i, int j)
{
    return ((P.dmask & 1) && j == 0) || ((P.dmask & 2) && j == P.nx) ||
           ((P.dmask & 4) && i == P.ny) || ((P.dmask & 8) && i == 0);
}

// Dirichlet value of a node; walls are applied left, right, top, bottom as
// fenicsInterface::createHSL pushes them (src/fHSL.cpp:476-537), so a corner
// on two Dirichlet walls keeps the later wall's value.
EQ_HD inline double dirichlet_value(const LevelProb& P, int i, int j)
{
    double v = 0.0;
    if ((P.dmask & 1) && j == 0) v = P.g[kLeft];
    if ((P.dmask & 2) && j == P.nx) v = P.g[kRight];
    if ((P.dmask & 4) && i == P.ny) v = P.g[kTop];
    if ((P.dmask & 8) && i == 0) v = P.g[kBottom];
    return v;
}

template <typename T>
struct Row {
    T mc, mw, me, ms, mn, msw, mne; // mass-like couplings (mass, reaction, Robin)
    T kw, ke, ks, kn;               // stiffness difference weights
};

// Row of A = (1+dt*gamma)M + dt*D*K + Robin at node (i,j) (Dirichlet rows are
// handled by the callers).
template <typename T>
EQ_HD inline Row<T> row_coefs(const LevelProb& P, int i, int j)
{
    const double sw = (i > 0 && j > 0) ? 1.0 : 0.0;
    const double se = (i > 0 && j < P.nx) ? 1.0 : 0.0;
    const double nw = (i < P.ny && j > 0) ? 1.0 : 0.0;
    const double ne = (i < P.ny && j < P.nx) ? 1.0 : 0.0;
    double mc = (2.0 * sw + se + nw + 2.0 * ne) * P.mh;
    double mw = (sw + nw) * P.mo, me = (se + ne) * P.mo;
    double ms = (sw + se) * P.mo, mn = (nw + ne) * P.mo;
    const double msw = 2.0 * sw * P.mo, mne = 2.0 * ne * P.mo;
    // Robin wall edges: dt*r*|e|*[[1/3,1/6],[1/6,1/3]]  (fenics/hslD.h:3284-3441)
    const double dn = (i > 0) ? 1.0 : 0.0, up = (i < P.ny) ? 1.0 : 0.0;
    const double lf = (j > 0) ? 1.0 : 0.0, rt = (j < P.nx) ? 1.0 : 0.0;
    if (j == 0 && P.rob[kLeft] != 0.0) {
        mc += (dn + up) * P.rob[kLeft] * (1.0 / 3.0);
        ms += dn * P.rob[kLeft] * (1.0 / 6.0);
        mn += up * P.rob[kLeft] * (1.0 / 6.0);
    }
    if (j == P.nx && P.rob[kRight] != 0.0) {
        mc += (dn + up) * P.rob[kRight] * (1.0 / 3.0);
        ms += dn * P.rob[kRight] * (1.0 / 6.0);
        mn += up * P.rob[kRight] * (1.0 / 6.0);
    }
    if (i == 0 && P.rob[kBottom] != 0.0) {
        mc += (lf + rt) * P.rob[kBottom] * (1.0 / 3.0);
        mw += lf * P.rob[kBottom] * (1.0 / 6.0);
        me += rt * P.rob[kBottom] * (1.0 / 6.0);
    }
    if (i == P.ny && P.rob[kTop] != 0.0) {
        mc += (lf + rt) * P.rob[kTop] * (1.0 / 3.0);
        mw += lf * P.rob[kTop] * (1.0 / 6.0);
        me += rt * P.rob[kTop] * (1.0 / 6.0);
    }
    Row<T> r;
    r.mc = (T)mc; r.mw = (T)mw; r.me = (T)me; r.ms = (T)ms; r.mn = (T)mn;
    r.msw = (T)msw; r.mne = (T)mne;
    r.kw = (T)((sw + nw) * P.sa);
    r.ke = (T)((se + ne) * P.sa);
    r.ks = (T)((sw + se) * P.sb);
    r.kn = (T)((nw + ne) * P.sb);
    return r;
}

template <typename T>
EQ_HD inline T row_diag(const Row<T>& r)
{
    return r.mc + ((r.kw + r.ke) + (r.ks + r.kn));
}

// y_c = (A v)_c from the seven values around the node
template <typename T>
EQ_HD inline T row_apply(const Row<T>& r, T vc, T vw, T ve, T vs, T vn, T vsw, T vne)
{
    const T stiff = (r.kw * (vc - vw) + r.ke * (vc - ve)) + (r.ks * (vc - vs) + r.kn * (vc - vn));
    const T mass = r.mc * vc + ((r.mw * vw + r.me * ve) + (r.ms * vs + r.mn * vn)) + (r.msw * vsw + r.mne * vne);
    return stiff + mass;
}

// Right-hand side row of level 0: b = M u0 + dt*f*int(phi) + Robin dt*r*s*int(phi)
// (fenics/hslD.h:3466-3529, 3554-3689).
EQ_HD inline double rhs_row(const LevelProb& P, int i, int j, double uc, double uw, double ue, double us,
                            double un, double usw, double une)
{
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

// floor(a / b) for b > 0 (a may be negative)
EQ_HD inline int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

} // namespace eqk
