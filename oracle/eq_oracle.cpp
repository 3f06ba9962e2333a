/*
 * eq_oracle.cpp -- CPU FP64 restatement of eQ's quorum-sensing field step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The
 * product library never links or calls this file.
 *
 * Built with -ffp-contract=off so that every a*b+c rounds twice, as the
 * reference's scalar C++ does on x86-64 without FMA contraction.
 */
#include "eq_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

inline int64_t node_index(const orc_problem* p, int64_t i, int64_t j) { return j + i * (int64_t)(p->nx + 1); }

/* Vertex coordinates of RectangleMesh(p0=(0,0), p1=(W,H), nx, ny) [ext: DOLFIN
 * 2019.1 RectangleMesh::build_tri]: x = a + ix*(b-a)/nx, y = c + iy*(d-c)/ny.
 * createMesh passes p0=(0,0), p1=(trapWidth, trapHeight)  (src/fHSL.cpp:169-172). */
inline double vx(const orc_problem* p, int64_t j) { return 0.0 + ((double)j * (p->width - 0.0)) / (double)p->nx; }
inline double vy(const orc_problem* p, int64_t i) { return 0.0 + ((double)i * (p->height - 0.0)) / (double)p->ny; }

/* "right" diagonal [ext]: square (iy,ix) with v0=(iy,ix), v1=v0+1,
 * v2=v0+(nx+1), v3=v1+(nx+1) is split into (v0,v1,v3) and (v0,v2,v3), i.e.
 * the diagonal runs bottom-left -> top-right. */
struct Tri { int64_t v[3]; };
inline void cell_triangles(const orc_problem* p, int64_t iy, int64_t ix, Tri t[2])
{
    const int64_t v0 = node_index(p, iy, ix), v1 = v0 + 1;
    const int64_t v2 = v0 + (p->nx + 1), v3 = v1 + (p->nx + 1);
    t[0] = Tri{{v0, v1, v3}};
    t[1] = Tri{{v0, v2, v3}};
}

inline void coords_of(const orc_problem* p, int64_t v, double& x, double& y)
{
    const int64_t nw = p->nx + 1;
    x = vx(p, v % nw);
    y = vy(p, v / nw);
}

/* Bilinear cell kernel, restating hsld_cell_integral_0_otherwise::tabulate_tensor
 * (fenics/hslD.h:3123-3259) with the coefficient order D11,D22,D12,D,dt of
 * Form_a (fenics/hslD.h:4675-4690):
 *   A_e = |detJ|*[1/12 diag, 1/24 off]                          (mass, 3237-3245)
 *       + dt * sum_q w_q |detJ| grad(phi_i)^T (D*Dtensor(x_q)) grad(phi_j)
 * with the degree-2 rule w_q = 1/6 at barycentric points (2/3,1/6,1/6) and
 * permutations (3147-3156), Dtensor interpolated in P1.  The reaction term
 * dt*gamma*M (absent from hslD.ufl) is added as gamma-scaled mass. */
void element_matrix(const double xc[3], const double yc[3], const double d11[3], const double d22[3],
                    const double d12[3], double D, double dt, double gamma, double Ae[3][3])
{
    const double J00 = xc[1] - xc[0], J01 = xc[2] - xc[0];
    const double J10 = yc[1] - yc[0], J11 = yc[2] - yc[0];
    const double det = J00 * J11 - J01 * J10;
    const double adet = std::fabs(det);
    /* inverse Jacobian K = J^{-1} */
    const double K00 = J11 / det, K01 = -J01 / det, K10 = -J10 / det, K11 = J00 / det;
    /* reference gradients of phi0=1-X-Y, phi1=X, phi2=Y */
    const double gX[3] = {-1.0, 1.0, 0.0}, gY[3] = {-1.0, 0.0, 1.0};
    double gx[3], gy[3];
    for (int k = 0; k < 3; ++k) {
        gx[k] = K00 * gX[k] + K10 * gY[k];
        gy[k] = K01 * gX[k] + K11 * gY[k];
    }
    static const double bary[3][3] = {{2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                                      {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0},
                                      {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}};
    double S[3][3] = {{0}};
    for (int q = 0; q < 3; ++q) {
        double a11 = 0, a22 = 0, a12 = 0;
        for (int c = 0; c < 3; ++c) {
            a11 += d11[c] * bary[q][c];
            a22 += d22[c] * bary[q][c];
            a12 += d12[c] * bary[q][c];
        }
        a11 *= D; a22 *= D; a12 *= D;
        const double wq = (1.0 / 6.0) * adet * dt;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double fx = a11 * gx[j] + a12 * gy[j];
                const double fy = a12 * gx[j] + a22 * gy[j];
                S[i][j] += wq * (gx[i] * fx + gy[i] * fy);
            }
    }
    const double mscale = 1.0 + dt * gamma;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Ae[i][j] = mscale * adet * (i == j ? 1.0 / 12.0 : 1.0 / 24.0) + S[i][j];
}

/* Galerkin convection element matrix C_e[i][j] = int_T v dphi_j/dx phi_i with v
 * P1 (vertex values vv): int_T v phi_i = |T|/12 (v_i + v_0 + v_1 + v_2) exactly,
 * and dphi_j/dx constant on T */
void convection_matrix(const double xc[3], const double yc[3], const double vv[3], double Ce[3][3])
{
    const double J00 = xc[1] - xc[0], J01 = xc[2] - xc[0];
    const double J10 = yc[1] - yc[0], J11 = yc[2] - yc[0];
    const double det = J00 * J11 - J01 * J10;
    const double adet = std::fabs(det);
    const double K00 = J11 / det, K10 = -J10 / det;
    const double gX[3] = {-1.0, 1.0, 0.0}, gY[3] = {-1.0, 0.0, 1.0};
    const double vs = vv[0] + vv[1] + vv[2];
    for (int i = 0; i < 3; ++i) {
        const double Ii = 0.5 * adet / 12.0 * (vv[i] + vs);
        for (int j = 0; j < 3; ++j) Ce[i][j] = (K00 * gX[j] + K10 * gY[j]) * Ii;
    }
}

/* wall predicates on computed vertex coordinates, DirichletBoundary_TrapEdge::inside
 * (src/fHSL.h:575-590) with near(x, v) = |x - v| < DOLFIN_EPS [ext] */
const double kDolfinEps = 3.0e-16;
inline bool near(double a, double b) { return std::fabs(a - b) < kDolfinEps; }
inline bool on_wall(const orc_problem* p, int w, double x, double y)
{
    switch (w) {
    case ORC_LEFT: return near(x, 0.0);
    case ORC_RIGHT: return near(x, p->width);
    case ORC_TOP: return near(y, p->height);
    default: return near(y, 0.0);
    }
}

/* a vertex on the outer row or column of wall w, by grid index */
inline bool on_wall_index(const orc_problem* p, int w, int64_t v)
{
    const int64_t nw = p->nx + 1, i = v / nw, j = v % nw;
    switch (w) {
    case ORC_LEFT: return j == 0;
    case ORC_RIGHT: return j == p->nx;
    case ORC_TOP: return i == p->ny;
    default: return i == 0;
    }
}

/* exterior facets of the mesh as (v_a, v_b, wall), each with both endpoints on the wall */
struct Facet { int64_t a, b; int wall; };
std::vector<Facet> boundary_facets(const orc_problem* p)
{
    std::vector<Facet> f;
    const int64_t nx = p->nx, ny = p->ny;
    for (int64_t j = 0; j < nx; ++j) f.push_back({node_index(p, 0, j), node_index(p, 0, j + 1), ORC_BOTTOM});
    for (int64_t j = 0; j < nx; ++j) f.push_back({node_index(p, ny, j), node_index(p, ny, j + 1), ORC_TOP});
    for (int64_t i = 0; i < ny; ++i) f.push_back({node_index(p, i, 0), node_index(p, i + 1, 0), ORC_LEFT});
    for (int64_t i = 0; i < ny; ++i) f.push_back({node_index(p, i, nx), node_index(p, i + 1, nx), ORC_RIGHT});
    return f;
}

/* a facet is marked for wall w when the SubDomain holds at both vertices
 * (SubDomain::mark over facet vertices [ext]) */
inline bool facet_on_wall(const orc_problem* p, const Facet& fc, int w)
{
    if (p->grid_walls) return on_wall_index(p, w, fc.a) && on_wall_index(p, w, fc.b);
    double xa, ya, xb, yb;
    coords_of(p, fc.a, xa, ya);
    coords_of(p, fc.b, xb, yb);
    return on_wall(p, w, xa, ya) && on_wall(p, w, xb, yb);
}

inline double facet_length(const orc_problem* p, const Facet& fc)
{
    double xa, ya, xb, yb;
    coords_of(p, fc.a, xa, ya);
    coords_of(p, fc.b, xb, yb);
    const double ex = xb - xa, ey = yb - ya;
    return std::sqrt(ex * ex + ey * ey);
}

} // namespace

extern "C" {

int64_t orc_num_nodes(const orc_problem* p) { return (int64_t)(p->nx + 1) * (p->ny + 1); }

void orc_walls_found(const orc_problem* p, int32_t* found4)
{
    /* every facet of a wall shares the wall's computed coordinate, so the
     * wall is found on all of its facets or on none */
    const std::vector<Facet> facets = boundary_facets(p);
    for (int w = 0; w < 4; ++w) found4[w] = 0;
    for (const Facet& fc : facets)
        if (facet_on_wall(p, fc, fc.wall)) found4[fc.wall] = 1;
}

void orc_mesh_coords(const orc_problem* p, double* xy)
{
    const int64_t n = orc_num_nodes(p);
    for (int64_t v = 0; v < n; ++v) coords_of(p, v, xy[2 * v], xy[2 * v + 1]);
}

int64_t orc_assemble(const orc_problem* p, const double* D11, const double* D22, const double* D12,
                     int64_t* rowptr, int32_t* col, double* val)
{
    const int64_t n = orc_num_nodes(p), nw = p->nx + 1;
    /* fixed 7-point pattern of the "right" mesh: SW,S,W,C,E,N,NE, kept even where
     * the entry is numerically zero (the stiffness diagonal couplings), as
     * DOLFIN's sparsity pattern keeps every mesh edge */
    std::vector<std::vector<std::pair<int32_t, double>>> rows(n);
    auto add = [&](int64_t r, int64_t c, double v) {
        auto& row = rows[r];
        for (auto& e : row)
            if (e.first == c) { e.second += v; return; }
        row.push_back({(int32_t)c, v});
    };
    for (int64_t iy = 0; iy < p->ny; ++iy)
        for (int64_t ix = 0; ix < p->nx; ++ix) {
            Tri t[2];
            cell_triangles(p, iy, ix, t);
            for (int k = 0; k < 2; ++k) {
                double xc[3], yc[3], d11[3], d22[3], d12[3];
                for (int c = 0; c < 3; ++c) {
                    coords_of(p, t[k].v[c], xc[c], yc[c]);
                    /* AnisotropicDiffusionTensor::eval at the vertex (src/Expressions.h:115-121) */
                    d11[c] = D11 ? D11[t[k].v[c]] : 1.0;
                    d22[c] = D22 ? D22[t[k].v[c]] : 1.0;
                    d12[c] = D12 ? D12[t[k].v[c]] : 0.0;
                }
                double Ae[3][3];
                element_matrix(xc, yc, d11, d22, d12, p->D, p->dt, p->gamma, Ae);
                if (p->vx) {
                    double vv[3], Ce[3][3];
                    for (int c = 0; c < 3; ++c) vv[c] = p->vx[t[k].v[c] / (p->nx + 1)];
                    convection_matrix(xc, yc, vv, Ce);
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) Ae[a][b] += p->dt * Ce[a][b];
                }
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) add(t[k].v[a], t[k].v[b], Ae[a][b]);
            }
        }
    /* Robin facets, hsld_exterior_facet_integral_0_{1,2} (fenics/hslD.h:3284-3441):
     * dt*r*|e|*[[1/3,1/6],[1/6,1/3]] */
    for (const Facet& fc : boundary_facets(p)) {
        const orc_bc& bc = p->wall[fc.wall];
        if (bc.kind != ORC_ROBIN || !facet_on_wall(p, fc, fc.wall)) continue;
        const double s = p->dt * bc.r * facet_length(p, fc);
        add(fc.a, fc.a, s * (1.0 / 3.0));
        add(fc.b, fc.b, s * (1.0 / 3.0));
        add(fc.a, fc.b, s * (1.0 / 6.0));
        add(fc.b, fc.a, s * (1.0 / 6.0));
    }
    int64_t nnz = 0;
    rowptr[0] = 0;
    for (int64_t r = 0; r < n; ++r) {
        auto& row = rows[r];
        std::sort(row.begin(), row.end());
        for (auto& e : row) { col[nnz] = e.first; val[nnz] = e.second; ++nnz; }
        rowptr[r + 1] = nnz;
    }
    (void)nw;
    return nnz;
}

/* Linear form, hsld_cell_integral_1_otherwise (fenics/hslD.h:3466-3529):
 *   b_e[i] = sum_q (1/6)|detJ| (dt*f + u0(x_q)) phi_i(x_q)
 * and Robin facets hsld_exterior_facet_integral_1_{1,2} (3554-3689):
 *   b[a] += dt*r*s*|e|/2 for both facet vertices. */
void orc_assemble_rhs(const orc_problem* p, const double* u0, double* b)
{
    const int64_t n = orc_num_nodes(p);
    std::fill(b, b + n, 0.0);
    static const double bary[3][3] = {{2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                                      {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0},
                                      {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}};
    for (int64_t iy = 0; iy < p->ny; ++iy)
        for (int64_t ix = 0; ix < p->nx; ++ix) {
            Tri t[2];
            cell_triangles(p, iy, ix, t);
            for (int k = 0; k < 2; ++k) {
                double xc[3], yc[3];
                for (int c = 0; c < 3; ++c) coords_of(p, t[k].v[c], xc[c], yc[c]);
                const double det = (xc[1] - xc[0]) * (yc[2] - yc[0]) - (xc[2] - xc[0]) * (yc[1] - yc[0]);
                const double adet = std::fabs(det);
                double be[3] = {0, 0, 0};
                for (int q = 0; q < 3; ++q) {
                    double w0 = 0.0;
                    for (int c = 0; c < 3; ++c) w0 += u0[t[k].v[c]] * bary[q][c];
                    const double fw = (p->dt * p->f + w0) * adet * (1.0 / 6.0);
                    for (int i = 0; i < 3; ++i) be[i] += fw * bary[q][i];
                }
                for (int i = 0; i < 3; ++i) b[t[k].v[i]] += be[i];
            }
        }
    for (const Facet& fc : boundary_facets(p)) {
        const orc_bc& bc = p->wall[fc.wall];
        if (bc.kind != ORC_ROBIN || !facet_on_wall(p, fc, fc.wall)) continue;
        const double s = p->dt * bc.r * bc.s * facet_length(p, fc) * 0.5;
        b[fc.a] += s;
        b[fc.b] += s;
    }
}

/* DirichletBC nodes: every vertex of every facet marked on a Dirichlet wall
 * (topological method [ext]); the walls are decoded per src/fHSL.cpp:467-539.
 * Walls are applied in the order left, right, top, bottom, matching the
 * push_back order of shell->dbc, so a corner shared by two Dirichlet walls
 * takes the later wall's value, as the later DirichletBC::apply overwrites. */
int64_t orc_dirichlet_nodes(const orc_problem* p, uint8_t* mask, double* g)
{
    const int64_t n = orc_num_nodes(p);
    std::fill(mask, mask + n, (uint8_t)0);
    if (g) std::fill(g, g + n, 0.0);
    const std::vector<Facet> facets = boundary_facets(p);
    for (int w = 0; w < 4; ++w) {
        if (p->wall[w].kind != ORC_DIRICHLET) continue;
        for (const Facet& fc : facets) {
            if (!facet_on_wall(p, fc, w)) continue;
            mask[fc.a] = mask[fc.b] = 1;
            if (g) {
                const double* gw = p->gwall[w];
                const int64_t nw = p->nx + 1;
                /* position along the wall: column on top/bottom, row on left/right */
                auto pos = [&](int64_t v) { return (w == ORC_TOP || w == ORC_BOTTOM) ? v % nw : v / nw; };
                g[fc.a] = gw ? gw[pos(fc.a)] : p->wall[w].value;
                g[fc.b] = gw ? gw[pos(fc.b)] : p->wall[w].value;
            }
        }
    }
    int64_t cnt = 0;
    for (int64_t i = 0; i < n; ++i) cnt += mask[i];
    return cnt;
}

void orc_apply_dirichlet(const orc_problem* p, const int64_t* rowptr, const int32_t* col, double* val,
                         double* b)
{
    const int64_t n = orc_num_nodes(p);
    std::vector<uint8_t> mask(n);
    std::vector<double> g(n);
    orc_dirichlet_nodes(p, mask.data(), g.data());
    for (int64_t r = 0; r < n; ++r) {
        if (!mask[r]) continue;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) val[k] = (col[k] == r) ? 1.0 : 0.0;
        if (b) b[r] = g[r];
    }
}

void orc_spmv(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val, const double* x,
              double* y)
{
    for (int64_t r = 0; r < n; ++r) {
        double s = 0.0;
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) s += val[k] * x[col[k]];
        y[r] = s;
    }
}

int orc_solve_banded(const orc_problem* p, const int64_t* rowptr, const int32_t* col, const double* val,
                     const double* b, double* x)
{
    const int64_t n = orc_num_nodes(p);
    const int64_t bw = p->nx + 2; /* lower = upper bandwidth of the natural order */
    const int64_t ld = 2 * bw + 1;
    std::vector<double> band((size_t)(n * ld), 0.0);
    auto at = [&](int64_t r, int64_t c) -> double& { return band[(size_t)(r * ld + (c - r + bw))]; };
    for (int64_t r = 0; r < n; ++r)
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            if (std::llabs((long long)(col[k] - r)) > bw) return 2;
            at(r, col[k]) += val[k];
        }
    std::vector<double> y(b, b + n);
    for (int64_t k = 0; k < n; ++k) {
        const double piv = at(k, k);
        if (piv == 0.0) return 1;
        const int64_t rmax = std::min(n - 1, k + bw), cmax = std::min(n - 1, k + bw);
        for (int64_t r = k + 1; r <= rmax; ++r) {
            double& lrk = at(r, k);
            if (lrk == 0.0) continue;
            const double m = lrk / piv;
            lrk = m;
            for (int64_t c = k + 1; c <= cmax; ++c) at(r, c) -= m * at(k, c);
            y[r] -= m * y[k];
        }
    }
    for (int64_t k = n - 1; k >= 0; --k) {
        double s = y[k];
        const int64_t cmax = std::min(n - 1, k + bw);
        for (int64_t c = k + 1; c <= cmax; ++c) s -= at(k, c) * x[c];
        x[k] = s / at(k, k);
    }
    return 0;
}

int orc_fd_step(const orc_problem* p, double* u)
{
    const int64_t nx = p->nx, ny = p->ny, nw = nx + 1, n = orc_num_nodes(p);
    const double hx = p->width / nx, hy = p->height / ny;
    const double Fx = p->D * p->dt / (hx * hx), Fy = p->D * p->dt / (hy * hy);
    /* per wall: Nc, Dc, BV of the reference's Robin form Dc u + Nc du/dn = BV */
    double Nc[4], Dc[4], BV[4];
    for (int w = 0; w < 4; ++w) {
        const orc_bc& bc = p->wall[w];
        if (bc.kind == ORC_DIRICHLET) {
            Nc[w] = 0.0; Dc[w] = 1.0; BV[w] = bc.value;
        } else if (bc.kind == ORC_ROBIN) {
            Nc[w] = 1.0; Dc[w] = bc.r / p->D; BV[w] = bc.r * bc.s / p->D;
        } else {
            Nc[w] = 1.0; Dc[w] = 0.0; BV[w] = 0.0;
        }
    }
    std::vector<int64_t> rowptr(n + 1, 0);
    std::vector<int32_t> col;
    std::vector<double> val, b(n);
    col.reserve(5 * n);
    val.reserve(5 * n);
    for (int64_t i = 0; i <= ny; ++i)
        for (int64_t j = 0; j <= nx; ++j) {
            const int64_t k = j + i * nw;
            const bool L = j == 0, R = j == nx, B = i == 0, T = i == ny;
            const bool dirx = (L && Nc[ORC_LEFT] == 0.0) || (R && Nc[ORC_RIGHT] == 0.0);
            const bool diry = (B && Nc[ORC_BOTTOM] == 0.0) || (T && Nc[ORC_TOP] == 0.0);
            std::vector<std::pair<int32_t, double>> row;
            if (dirx || diry) {
                row.push_back({(int32_t)k, 1.0});
                double g = 0.0;
                if (T && Nc[ORC_TOP] == 0.0) g = BV[ORC_TOP];
                if (B && Nc[ORC_BOTTOM] == 0.0) g = BV[ORC_BOTTOM];
                if (R && Nc[ORC_RIGHT] == 0.0) g = BV[ORC_RIGHT];
                if (L && Nc[ORC_LEFT] == 0.0) g = BV[ORC_LEFT];
                b[k] = g;
            } else {
                double diag = 1.0 + 2.0 * Fx + 2.0 * Fy;
                double rhs = u[k];
                /* x direction: a wall reflects the ghost onto the inward neighbour */
                if (L) {
                    row.push_back({(int32_t)(k + 1), -2.0 * Fx});
                    diag += 2.0 * hx * Fx * Dc[ORC_LEFT] / Nc[ORC_LEFT];
                    rhs += 2.0 * Fx * hx * BV[ORC_LEFT] / Nc[ORC_LEFT];
                } else if (R) {
                    row.push_back({(int32_t)(k - 1), -2.0 * Fx});
                    diag += 2.0 * hx * Fx * Dc[ORC_RIGHT] / Nc[ORC_RIGHT];
                    rhs += 2.0 * Fx * hx * BV[ORC_RIGHT] / Nc[ORC_RIGHT];
                } else {
                    row.push_back({(int32_t)(k - 1), -Fx});
                    row.push_back({(int32_t)(k + 1), -Fx});
                }
                if (B) {
                    row.push_back({(int32_t)(k + nw), -2.0 * Fy});
                    diag += 2.0 * hy * Fy * Dc[ORC_BOTTOM] / Nc[ORC_BOTTOM];
                    rhs += 2.0 * Fy * hy * BV[ORC_BOTTOM] / Nc[ORC_BOTTOM];
                } else if (T) {
                    row.push_back({(int32_t)(k - nw), -2.0 * Fy});
                    diag += 2.0 * hy * Fy * Dc[ORC_TOP] / Nc[ORC_TOP];
                    rhs += 2.0 * Fy * hy * BV[ORC_TOP] / Nc[ORC_TOP];
                } else {
                    row.push_back({(int32_t)(k - nw), -Fy});
                    row.push_back({(int32_t)(k + nw), -Fy});
                }
                row.push_back({(int32_t)k, diag});
                b[k] = rhs;
            }
            std::sort(row.begin(), row.end());
            for (auto& e : row) {
                col.push_back(e.first);
                val.push_back(e.second);
            }
            rowptr[k + 1] = (int64_t)col.size();
        }
    std::vector<double> x(n);
    const int rc = orc_solve_banded(p, rowptr.data(), col.data(), val.data(), b.data(), x.data());
    if (rc) return rc;
    std::copy(x.begin(), x.end(), u);
    return 0;
}

/* one channel: u <- CN step of  (M + k/2 (D K + v C + R)) u = (M - k/2 (D K + v C + R)) u0
 * + k (r1 s1 e_0 + r2 s2 e_N), P1 on a uniform interval mesh (cells of length hc):
 * M_e = hc/6 [[2,1],[1,2]], K_e = 1/hc [[1,-1],[-1,1]], C_e = [[-1/2,1/2],[-1/2,1/2]]
 * (row = test function, the advection term v du/dx w) */
static void channel_cn(int64_t N, double hc, double D, double v, double k, const orc_channel* ch, double* u)
{
    const int64_t n = N + 1;
    std::vector<double> a(n, 0.0), b(n, 0.0), c(n, 0.0), r(n, 0.0);
    /* operator T = D K + v C + R, assembled per element; A = M + k/2 T, B = M - k/2 T */
    std::vector<double> Ml(n, 0.0), Md(n, 0.0), Mu(n, 0.0), Tl(n, 0.0), Td(n, 0.0), Tu(n, 0.0);
    for (int64_t e = 0; e < N; ++e) {
        const int64_t p0 = e, p1 = e + 1;
        Md[p0] += 2.0 * hc / 6.0; Mu[p0] += hc / 6.0;
        Ml[p1] += hc / 6.0;       Md[p1] += 2.0 * hc / 6.0;
        Td[p0] += D / hc;  Tu[p0] += -D / hc;
        Tl[p1] += -D / hc; Td[p1] += D / hc;
        Td[p0] += -0.5 * v; Tu[p0] += 0.5 * v;
        Tl[p1] += -0.5 * v; Td[p1] += 0.5 * v;
    }
    Td[0] += ch->r_left;
    Td[N] += ch->r_right;
    for (int64_t i = 0; i < n; ++i) {
        a[i] = Ml[i] + 0.5 * k * Tl[i];
        b[i] = Md[i] + 0.5 * k * Td[i];
        c[i] = Mu[i] + 0.5 * k * Tu[i];
        double rhs = (Md[i] - 0.5 * k * Td[i]) * u[i];
        if (i > 0) rhs += (Ml[i] - 0.5 * k * Tl[i]) * u[i - 1];
        if (i < N) rhs += (Mu[i] - 0.5 * k * Tu[i]) * u[i + 1];
        r[i] = rhs;
    }
    r[0] += k * ch->r_left * ch->s_left;
    r[N] += k * ch->r_right * ch->s_right;
    /* Thomas algorithm */
    std::vector<double> cp(n), dp(n);
    cp[0] = c[0] / b[0];
    dp[0] = r[0] / b[0];
    for (int64_t i = 1; i < n; ++i) {
        const double m = b[i] - a[i] * cp[i - 1];
        cp[i] = c[i] / m;
        dp[i] = (r[i] - a[i] * dp[i - 1]) / m;
    }
    u[N] = dp[N];
    for (int64_t i = N - 1; i >= 0; --i) u[i] = dp[i] - cp[i] * u[i + 1];
}

void orc_channel_update(const orc_problem* p, const orc_channel* top_ch, const orc_channel* bottom_ch,
                        const double* u, double* top, double* bottom)
{
    const int64_t nx = p->nx, ny = p->ny, nw = nx + 1;
    const double hc = p->width / nx;
    for (int side = 0; side < 2; ++side) {
        const orc_channel* ch = side == 0 ? top_ch : bottom_ch;
        double* cv = side == 0 ? top : bottom;
        if (!ch || !cv) continue;
        const int nsub = ch->substeps > 0 ? ch->substeps : 1;
        std::vector<double> flux(nw);
        for (int64_t j = 0; j < nw; ++j) {
            /* ds * dt * D * gradc / wellScaling with ds = h and gradc = du/h */
            const double du = side == 0 ? u[(ny - 1) * nw + j] - u[ny * nw + j] : u[nw + j] - u[j];
            flux[j] = p->dt * p->D * du / ch->well_scaling;
        }
        const double k = p->dt / nsub;
        for (int s = 0; s < nsub; ++s) {
            for (int64_t j = 0; j < nw; ++j) cv[j] += flux[j] / (double)nsub;
            channel_cn(nx, hc, p->D, ch->v, k, ch, cv);
        }
    }
}

int orc_solve_pcg(const orc_problem* p, const int64_t* rowptr, const int32_t* col, const double* val,
                  const double* b, double* x, double rtol, int32_t maxit, orc_solve_info* info)
{
    /* Symmetric elimination of the Dirichlet rows: x_D = g is fixed, the free
     * rows see b_F - A_FD g.  Its solution equals the row-replaced system's. */
    const int64_t n = orc_num_nodes(p);
    std::vector<uint8_t> mask(n);
    std::vector<double> g(n);
    orc_dirichlet_nodes(p, mask.data(), g.data());
    std::vector<double> bh(n), diag(n, 1.0);
    double bnorm2 = 0.0;
    for (int64_t r = 0; r < n; ++r) bnorm2 += b[r] * b[r];
    for (int64_t r = 0; r < n; ++r) {
        if (mask[r]) { x[r] = g[r]; bh[r] = 0.0; continue; }
        double s = b[r];
        for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) {
            if (mask[col[k]]) s -= val[k] * g[col[k]];
            if (col[k] == r) diag[r] = val[k];
        }
        bh[r] = s;
    }
    auto apply = [&](const std::vector<double>& v, std::vector<double>& out) {
        for (int64_t r = 0; r < n; ++r) {
            if (mask[r]) { out[r] = 0.0; continue; }
            double s = 0.0;
            for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k)
                if (!mask[col[k]]) s += val[k] * v[col[k]];
            out[r] = s;
        }
    };
    std::vector<double> xf(n, 0.0), r(n), z(n), pv(n), q(n);
    for (int64_t i = 0; i < n; ++i) xf[i] = mask[i] ? 0.0 : x[i];
    apply(xf, q);
    double rr = 0.0, rz = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        r[i] = mask[i] ? 0.0 : bh[i] - q[i];
        z[i] = r[i] / diag[i];
        pv[i] = z[i];
        rr += r[i] * r[i];
        rz += r[i] * z[i];
    }
    const double bnorm = std::sqrt(bnorm2);
    const double tol2 = rtol * rtol * bnorm2;
    int32_t it = 0;
    /* the CSR residual has a rounding floor near eps*8*dt*D/h^2 relative;
     * stop on stagnation of the true residual as well */
    double best_true = rr;
    int32_t best_it = 0;
    while (rr > tol2 && it < maxit) {
        apply(pv, q);
        double pq = 0.0;
        for (int64_t i = 0; i < n; ++i) pq += pv[i] * q[i];
        if (!(pq > 0.0) || !(rz > 0.0)) break; /* exact solution reached */
        const double alpha = rz / pq;
        double rz1 = 0.0;
        rr = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            xf[i] += alpha * pv[i];
            r[i] -= alpha * q[i];
            z[i] = r[i] / diag[i];
            rz1 += r[i] * z[i];
            rr += r[i] * r[i];
        }
        const double beta = rz1 / rz;
        rz = rz1;
        for (int64_t i = 0; i < n; ++i) pv[i] = z[i] + beta * pv[i];
        ++it;
        if (it % 50 == 0 || rr <= tol2) {
            /* recompute the true residual to stop on it, not on the recurrence */
            apply(xf, q);
            rr = 0.0;
            for (int64_t i = 0; i < n; ++i) {
                const double ri = mask[i] ? 0.0 : bh[i] - q[i];
                rr += ri * ri;
            }
            if (rr < 0.25 * best_true) { best_true = rr; best_it = it; }
            else if (it - best_it >= 400) break;
        }
    }
    for (int64_t i = 0; i < n; ++i) x[i] = mask[i] ? g[i] : xf[i];
    if (info) {
        info->iterations = it;
        info->bnorm = bnorm;
        info->relres = bnorm > 0 ? std::sqrt(rr) / bnorm : std::sqrt(rr);
    }
    return rr <= tol2 ? 0 : 1;
}

int orc_step(const orc_problem* p, const double* D11, const double* D22, const double* D12, double* u,
             int32_t method, double rtol, int32_t maxit, orc_solve_info* info)
{
    /* stepDiffusion: u0 <- solution_vector (src/fHSL.cpp:104); LVS->solve()
     * reassembles a and L and applies every DirichletBC (105-106); then
     * solution_vector <- u (108). */
    const int64_t n = orc_num_nodes(p);
    std::vector<int64_t> rowptr(n + 1);
    std::vector<int32_t> col(7 * n);
    std::vector<double> val(7 * n), b(n);
    orc_assemble(p, D11, D22, D12, rowptr.data(), col.data(), val.data());
    orc_assemble_rhs(p, u, b.data());
    orc_apply_dirichlet(p, rowptr.data(), col.data(), val.data(), b.data());
    int rc;
    if (method == 0) {
        std::vector<double> x(n);
        rc = orc_solve_banded(p, rowptr.data(), col.data(), val.data(), b.data(), x.data());
        if (info) {
            std::vector<double> ax(n);
            orc_spmv(n, rowptr.data(), col.data(), val.data(), x.data(), ax.data());
            double rr = 0, bb = 0;
            for (int64_t i = 0; i < n; ++i) { rr += (b[i] - ax[i]) * (b[i] - ax[i]); bb += b[i] * b[i]; }
            info->iterations = 0;
            info->bnorm = std::sqrt(bb);
            info->relres = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
        }
        std::copy(x.begin(), x.end(), u);
    } else if (p->vx) {
        rc = 3; /* CG needs a symmetric operator */
    } else {
        rc = orc_solve_pcg(p, rowptr.data(), col.data(), val.data(), b.data(), u, rtol, maxit > 0 ? maxit : 1000000, info);
    }
    return rc;
}

/* ------------------------------------------------------------------ cells */

namespace {

/* eQ::data::ij_from_xy: j = size_t(round(x*n)), i = size_t(round(y*n))  (src/eQ.h:119-124).
 * Negative rounded values (a point left of / below the grid) are clamped to 0
 * here; the reference converts them to size_t, which is undefined. */
inline int64_t to_index(double v, double npm)
{
    const double r = std::round(v * npm);
    return r < 0.0 ? 0 : (int64_t)r;
}

struct Pole { double x, y; };

/* Ecoli::updatePoleCenters (src/abm/Ecoli.cpp:36-63): centre +/-
 * (cos a, sin a)*(r*0.5*L - w/2), clamped to the trap rectangle. */
inline Pole pole_position(const orc_grid* g, const orc_cell* c, int r)
{
    const double s = ((double)r * 0.5) * c->length - c->width / 2.0;
    Pole p{c->x + std::cos(c->angle) * s, c->y + std::sin(c->angle) * s};
    if (p.x < 0.0) p.x = 0.0;
    if (p.y < 0.0) p.y = 0.0;
    if (p.x > g->trap_w) p.x = g->trap_w;
    if (p.y > g->trap_h) p.y = g->trap_h;
    return p;
}

struct Box { double ulx, uly, urx, ury, lrx, lry, llx, lly; double e[4][2]; };

/* Body-A box of a cell with its current length (src/abm/cpmEcoli.cpp:102-124):
 * offset = (length-width)/2, radius = width/2; vertices UL,UR,LR,LL and the
 * edge vectors of computeCellBoudaryEdges (src/abm/cpmEcoli.cpp:71-81). */
inline Box cell_box(const orc_cell* c)
{
    const double shapeLength = c->length - c->width;
    const double radius = c->width * 0.5, offset = shapeLength * 0.5;
    Box b;
    b.ulx = -offset; b.uly = radius;
    b.urx = offset;  b.ury = radius;
    b.lrx = offset;  b.lry = -radius;
    b.llx = -offset; b.lly = -radius;
    b.e[0][0] = b.urx - b.ulx; b.e[0][1] = b.ury - b.uly;
    b.e[1][0] = b.lrx - b.urx; b.e[1][1] = b.lry - b.ury;
    b.e[2][0] = b.llx - b.lrx; b.e[2][1] = b.lly - b.lry;
    b.e[3][0] = b.ulx - b.llx; b.e[3][1] = b.uly - b.lly;
    return b;
}

/* cpmEcoli::pointIsInCell (src/abm/cpmEcoli.cpp:313-328).  The world->local
 * map is Chipmunk 7.0.1's cpBodyWorldToLocal [ext] for a body at position
 * (px,py), angle a and centre of gravity 0:  cpTransformPoint of
 * cpTransformRigidInverse(t) where t = {a:cos, b:sin, c:-sin, d:cos, tx:px, ty:py};
 * cpTransformPoint(t,p) = (t.a*p.x + t.c*p.y + t.tx, t.b*p.x + t.d*p.y + t.ty). */
inline bool point_in_cell(const orc_cell* c, const Box& b, double cs, double sn, double X, double Y)
{
    const double ia = cs, ib = -sn, ic = sn, id = cs;
    const double itx = (-sn) * c->y - cs * c->x;
    const double ity = sn * c->x - cs * c->y;
    const double lx = ia * X + ic * Y + itx;
    const double ly = ib * X + id * Y + ity;
    const double p0x = lx - b.urx, p0y = ly - b.ury;
    const double p1x = lx - b.lrx, p1y = ly - b.lry;
    const double p2x = lx - b.llx, p2y = ly - b.lly;
    const double p3x = lx - b.ulx, p3y = ly - b.uly;
    /* cpvcross(a,b) = a.x*b.y - a.y*b.x, strict < 0 on every edge */
    return (b.e[0][0] * p0y - b.e[0][1] * p0x < 0.0) && (b.e[1][0] * p1y - b.e[1][1] * p1x < 0.0) &&
           (b.e[2][0] * p2y - b.e[2][1] * p2x < 0.0) && (b.e[3][0] * p3y - b.e[3][1] * p3x < 0.0);
}

} // namespace

void orc_cell_window(const orc_grid* g, const orc_cell* c, int32_t* win)
{
    /* findInteriorPoints (src/abm/eQabm.cpp:268-287); nodesToEdge uses the
     * default 1 um width, not the cell's (src/abm/eQabm.cpp:75) */
    const int64_t nodesToEdge = (int64_t)std::round(g->npm * 1.0 / 2.0);
    const Pole A = pole_position(g, c, 1), B = pole_position(g, c, -1);
    const int64_t Ai = to_index(A.y, g->npm), Aj = to_index(A.x, g->npm);
    const int64_t Bi = to_index(B.y, g->npm), Bj = to_index(B.x, g->npm);
    int64_t i1, i2, j1, j2;
    if (Ai > Bi) { i1 = Bi; i2 = Ai; } else { i1 = Ai; i2 = Bi; }
    if (Aj > Bj) { j1 = Bj; j2 = Aj; } else { j1 = Aj; j2 = Bj; }
    const int64_t nh = g->nodes_h, nw = g->nodes_w;
    i1 = (i1 >= nodesToEdge) ? (i1 - nodesToEdge) : 0;
    j1 = (j1 >= nodesToEdge) ? (j1 - nodesToEdge) : 0;
    i2 = ((i2 + nodesToEdge) >= (nh - 1)) ? nh - 1 : (i2 + nodesToEdge);
    j2 = ((j2 + nodesToEdge) >= (nw - 1)) ? nw - 1 : (j2 + nodesToEdge);
    win[0] = (int32_t)i1; win[1] = (int32_t)i2; win[2] = (int32_t)j1; win[3] = (int32_t)j2;
}

int64_t orc_cell_coverage(const orc_grid* g, const orc_cell* c, int32_t* ij, int64_t cap)
{
    int32_t w[4];
    orc_cell_window(g, c, w);
    const Box b = cell_box(c);
    const double cs = std::cos(c->angle), sn = std::sin(c->angle);
    int64_t cnt = 0;
    for (int64_t i = w[0]; i <= w[1]; ++i)
        for (int64_t j = w[2]; j <= w[3]; ++j) {
            /* xy_from_ij: (j/n, i/n)  (src/eQ.h:115-118) */
            const double X = double(j) / g->npm, Y = double(i) / g->npm;
            if (point_in_cell(c, b, cs, sn, X, Y)) {
                if (cnt < cap) { ij[2 * cnt] = (int32_t)i; ij[2 * cnt + 1] = (int32_t)j; }
                ++cnt;
            }
        }
    if (cnt == 0) {
        /* empty set -> centre node (src/abm/eQabm.cpp:299-303) */
        int64_t ci = to_index(c->y, g->npm), cj = to_index(c->x, g->npm);
        ci = std::min<int64_t>(ci, g->nodes_h - 1);
        cj = std::min<int64_t>(cj, g->nodes_w - 1);
        if (cap > 0) { ij[0] = (int32_t)ci; ij[1] = (int32_t)cj; }
        cnt = 1;
    }
    return cnt;
}

void orc_couple(const orc_grid* g, const orc_cell* cells, int64_t nc, int32_t ns, double* const* fields,
                const double* q, double* sampled, int32_t* counts)
{
    /* updateCells (src/abm/eQabm.cpp:234-449): per cell in list order,
     * findInteriorPoints, then readHSL for every species (326-337, 388-393),
     * then writeHSL for every species (338-359, 398-405). */
    const int64_t nw = g->nodes_w;
    std::vector<int32_t> ij;
    for (int64_t k = 0; k < nc; ++k) {
        int32_t w[4];
        orc_cell_window(g, &cells[k], w);
        const int64_t cap = (int64_t)(w[1] - w[0] + 1) * (w[3] - w[2] + 1) + 1;
        ij.resize(2 * cap);
        const int64_t cnt = orc_cell_coverage(g, &cells[k], ij.data(), cap);
        if (counts) counts[k] = (int32_t)cnt;
        for (int32_t s = 0; s < ns; ++s) {
            double hsl = 0.0;
            for (int64_t t = 0; t < cnt; ++t) hsl += fields[s][ij[2 * t + 1] + ij[2 * t] * nw];
            if (sampled) sampled[k * ns + s] = hsl / double(cnt);
        }
        for (int32_t s = 0; s < ns; ++s) {
            const double d = q[k * ns + s] / double(cnt);
            for (int64_t t = 0; t < cnt; ++t) fields[s][ij[2 * t + 1] + ij[2 * t] * nw] += d;
        }
    }
}

void orc_couple_affine(const orc_grid* g, const orc_cell* cells, int64_t nc, int32_t ns, double* const* fields,
                       const double* qa, const double* qb, double* sampled, int32_t* counts)
{
    /* updateCells (src/abm/eQabm.cpp:384-405) with the strain's amount formed
     * from the cell's own reading: readHSL for every species, then
     * q = a + b*e (the membrane term dt*membraneRate*(iHSL - eHSL) of
     * src/Strain.h:312-313 is of this form), then writeHSL. */
    const int64_t nw = g->nodes_w;
    std::vector<int32_t> ij;
    std::vector<double> e(ns);
    for (int64_t k = 0; k < nc; ++k) {
        int32_t w[4];
        orc_cell_window(g, &cells[k], w);
        const int64_t cap = (int64_t)(w[1] - w[0] + 1) * (w[3] - w[2] + 1) + 1;
        ij.resize(2 * cap);
        const int64_t cnt = orc_cell_coverage(g, &cells[k], ij.data(), cap);
        if (counts) counts[k] = (int32_t)cnt;
        for (int32_t s = 0; s < ns; ++s) {
            double hsl = 0.0;
            for (int64_t t = 0; t < cnt; ++t) hsl += fields[s][ij[2 * t + 1] + ij[2 * t] * nw];
            e[s] = hsl / double(cnt);
            if (sampled) sampled[k * ns + s] = e[s];
        }
        for (int32_t s = 0; s < ns; ++s) {
            const double q = qa[k * ns + s] + qb[k * ns + s] * e[s];
            const double d = q / double(cnt);
            for (int64_t t = 0; t < cnt; ++t) fields[s][ij[2 * t + 1] + ij[2 * t] * nw] += d;
        }
    }
}

void orc_cell_tensor(const orc_grid* g, const orc_cell* cells, int64_t nc, double Dx, double Dy, double* D11,
                     double* D22, double* D12)
{
    /* updateCells (src/abm/eQabm.cpp:241-248): the grids start isotropic;
     * setDiffusionTensor (306-325), called per cell in list order after its
     * read and write (407), overwrites the cell's covered nodes. */
    const int64_t nw = g->nodes_w, n = nw * g->nodes_h;
    std::fill(D11, D11 + n, 1.0);
    std::fill(D22, D22 + n, 1.0);
    std::fill(D12, D12 + n, 0.0);
    std::vector<int32_t> ij;
    for (int64_t k = 0; k < nc; ++k) {
        int32_t w[4];
        orc_cell_window(g, &cells[k], w);
        const int64_t cap = (int64_t)(w[1] - w[0] + 1) * (w[3] - w[2] + 1) + 1;
        ij.resize(2 * cap);
        const int64_t cnt = orc_cell_coverage(g, &cells[k], ij.data(), cap);
        const double theta = cells[k].angle;
        const double cos2t = std::cos(theta) * std::cos(theta);
        const double sin2t = std::sin(theta) * std::sin(theta);
        const double sincost = std::sin(theta) * std::cos(theta);
        for (int64_t t = 0; t < cnt; ++t) {
            const int64_t v = ij[2 * t + 1] + ij[2 * t] * nw;
            D11[v] = Dx * cos2t + Dy * sin2t;
            D22[v] = Dx * sin2t + Dy * cos2t;
            D12[v] = (Dx - Dy) * sincost;
        }
    }
}

double orc_boundary_well(double w, double flux, double well_scaling, double decay_rate, double dt,
                         int64_t* underflows)
{
    w += flux / well_scaling;
    w -= dt * decay_rate * w;
    if (w < 0.0) {
        w = 0.0;
        if (underflows) ++*underflows;
    }
    return w;
}

/* Boundary-flux functional, boundary_exterior_facet_integral_0_otherwise
 * (fenics/boundary.h:2652-2742): for every exterior facet, -grad(u).n*|e|
 * with grad(u) the P1 gradient of the triangle owning the facet. */
double orc_boundary_flux(const orc_problem* p, const double* u)
{
    double total = 0.0;
    for (int64_t iy = 0; iy < p->ny; ++iy)
        for (int64_t ix = 0; ix < p->nx; ++ix) {
            Tri t[2];
            cell_triangles(p, iy, ix, t);
            for (int k = 0; k < 2; ++k) {
                double xc[3], yc[3];
                for (int c = 0; c < 3; ++c) coords_of(p, t[k].v[c], xc[c], yc[c]);
                const double J00 = xc[1] - xc[0], J01 = xc[2] - xc[0];
                const double J10 = yc[1] - yc[0], J11 = yc[2] - yc[0];
                const double det = J00 * J11 - J01 * J10;
                const double K00 = J11 / det, K01 = -J01 / det, K10 = -J10 / det, K11 = J00 / det;
                const double dX = u[t[k].v[1]] - u[t[k].v[0]], dY = u[t[k].v[2]] - u[t[k].v[0]];
                const double gx = K00 * dX + K10 * dY, gy = K01 * dX + K11 * dY;
                /* triangle edges (0,1),(1,2),(2,0); an edge is exterior when both
                 * endpoints lie on the same outer row or column of the grid (ds
                 * without a marker covers every exterior facet, whatever the
                 * wall tests find) */
                for (int e = 0; e < 3; ++e) {
                    const int a = e, bq = (e + 1) % 3, o = (e + 2) % 3;
                    for (int w = 0; w < 4; ++w) {
                        if (!(on_wall_index(p, w, t[k].v[a]) && on_wall_index(p, w, t[k].v[bq]))) continue;
                        const double ex = xc[bq] - xc[a], ey = yc[bq] - yc[a];
                        const double len = std::sqrt(ex * ex + ey * ey);
                        double nx = ey / len, ny = -ex / len;
                        /* outward: away from the opposite vertex */
                        if (nx * (xc[o] - xc[a]) + ny * (yc[o] - yc[a]) > 0) { nx = -nx; ny = -ny; }
                        total += -(gx * nx + gy * ny) * len;
                    }
                }
            }
        }
    return total;
}

} // extern "C"
