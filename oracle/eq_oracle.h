/*
 * eq_oracle.h -- CPU FP64 restatement of eQ's quorum-sensing field step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (eq_amd/, libeq_field.so)
 * never links, imports or calls it.
 *
 * What it restates (reference paths are relative to the eQ source tree):
 *   - mesh:        DOLFIN RectangleMesh(p0,p1,nx,ny,"right") as built by
 *                  fenicsInterface::createMesh            src/fHSL.cpp:164-173
 *   - operator:    FFC element tensors of the hslD form    fenics/hslD.ufl:31-49
 *                  bilinear cell kernel                     fenics/hslD.h:3123-3259
 *                  bilinear Robin facets                    fenics/hslD.h:3284-3441
 *                  linear cell kernel                       fenics/hslD.h:3466-3529
 *                  linear Robin facets                      fenics/hslD.h:3554-3689
 *                  coefficient order                        fenics/hslD.h:4675-4730, 4968-5020
 *                  plus the reaction term dt*gamma*M (not in hslD.ufl; gamma=0
 *                  reproduces the reference form exactly)
 *   - boundaries:  per-wall decode of MICROFLUIDIC_TRAP     src/fHSL.cpp:467-539
 *                  DirichletBC::apply row replacement (DOLFIN 2019.1, [ext])
 *   - solve:       LinearVariationalSolver default LU       src/fHSL.h:156-160 ([ext])
 *   - cells:       eQabm::updateCells                       src/abm/eQabm.cpp:234-449
 *                  Ecoli::updatePoleCenters                 src/abm/Ecoli.cpp:36-63
 *                  cpmEcoli::pointIsInCell                  src/abm/cpmEcoli.cpp:313-328
 *                  box vertices / edges                     src/abm/cpmEcoli.cpp:71-81,102-124
 *                  index helpers                            src/eQ.h:115-134
 *
 * Natural node order everywhere: node (i,j), row i = y, column j = x,
 * index j + i*(nx+1)  (src/eQ.h:131-134).
 */
#ifndef EQ_ORACLE_H
#define EQ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* wall order follows eQ::boundaryCondition::wall (src/eQ.h:335-340) */
enum { ORC_LEFT = 0, ORC_RIGHT = 1, ORC_TOP = 2, ORC_BOTTOM = 3 };
/* kinds: Neumann (natural, r = 0), Dirichlet (u = value), Robin (rate r, exterior value s) */
enum { ORC_NEUMANN = 0, ORC_DIRICHLET = 1, ORC_ROBIN = 2 };

typedef struct {
    int32_t kind;
    int32_t pad_;
    double value; /* Dirichlet value */
    double r;     /* Robin rate      */
    double s;     /* Robin exterior concentration */
} orc_bc;

typedef struct {
    int32_t nx, ny;         /* cells in x and y                      */
    double width, height;   /* trap size in microns                  */
    double dt, D, gamma, f; /* step, diffusivity, degradation, source */
    orc_bc wall[4];
    /* per-node Dirichlet values along a wall (NULL: the wall's value): index j
     * (column) on top/bottom, i (row) on left/right -- a DirichletBC bound to a
     * Function instead of a Constant, as the channel-coupled walls are
     * (src/fHSL.cpp:514, 531) */
    const double* gwall[4];
    /* optional advection (the survey's config-4 channel flow; the reference's trap
     * form has none): x-velocity per grid row (ny+1 values), P1 in y, entering
     * as dt * int (v du/dx) phi (Galerkin) -- a nonsymmetric operator; the LU
     * path solves it, PCG refuses it */
    const double* vx;
    /* how walls are found: 0 = the reference's test, near() on the computed
     * vertex coordinates (DirichletBoundary_TrapEdge::inside, src/fHSL.h:575-590,
     * with DOLFIN_EPS [ext]), under which a wall whose computed coordinate
     * misses W or H by more than DOLFIN_EPS carries no condition; 1 = by grid
     * index (the outer rows and columns; the same walls whenever the computed
     * coordinates are exact, e.g. for every integer trap size) */
    int32_t grid_walls;
    int32_t pad_;
} orc_problem;

/* flow channel above or below the trap (fenicsChannel, src/fHSL.h:240-332) */
typedef struct {
    double v;                  /* flow velocity along +x (data.v, src/fHSL.cpp:336-337) */
    double r_left, s_left;     /* Robin rate / exterior value at the channel's left end  */
    double r_right, s_right;   /* ... right end (ds(1), ds(2) of AdvectionDiffusion.ufl) */
    double well_scaling;       /* wellScaling (src/fHSL.cpp:47)                           */
    int32_t substeps;          /* channelSolverNumberIterations (src/main.cpp:414-430)    */
    int32_t pad_;
} orc_channel;

typedef struct {
    double x, y, angle, length, width;
} orc_cell;

typedef struct {
    int32_t nodes_w, nodes_h; /* grid nodes (nx+1, ny+1)                 */
    double npm;               /* nodes per micron used by the ABM maps    */
    double trap_w, trap_h;    /* trap size in microns (pole clamping)     */
} orc_grid;

typedef struct {
    int32_t iterations;
    int32_t pad_;
    double relres;      /* ||b - A x|| / ||b|| on the row-replaced system */
    double bnorm;
} orc_solve_info;

/* ---- mesh ---- */
int64_t orc_num_nodes(const orc_problem* p);
/* found[w] = 1 when wall w carries its condition under p->grid_walls */
void orc_walls_found(const orc_problem* p, int32_t* found4);
void orc_mesh_coords(const orc_problem* p, double* xy);

/* ---- assembly ---- */
/* Assemble A = (1+dt*gamma)*M + dt*D*K(Dtensor) + Robin, no Dirichlet rows yet.
 * D11/D22/D12 are per-node tensor entries in natural order; NULL means the
 * isotropic default (1,1,0) of fenicsClassInit (src/fHSL.cpp:313-323).
 * CSR with sorted columns; rowptr has N+1 entries; returns nnz.  col/val must
 * hold 7*N entries. */
int64_t orc_assemble(const orc_problem* p, const double* D11, const double* D22,
                     const double* D12, int64_t* rowptr, int32_t* col, double* val);
/* b = M*u0 + dt*f*int(phi) + Robin dt*r*s*int(phi) on the marked walls */
void orc_assemble_rhs(const orc_problem* p, const double* u0, double* b);
/* mask[n] = 1 for Dirichlet nodes, g[n] their value */
int64_t orc_dirichlet_nodes(const orc_problem* p, uint8_t* mask, double* g);
/* DirichletBC::apply: zero the row, 1 on the diagonal, b_i = g_i */
void orc_apply_dirichlet(const orc_problem* p, const int64_t* rowptr, const int32_t* col,
                         double* val, double* b);
/* y = A x */
void orc_spmv(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
              const double* x, double* y);

/* ---- solves ---- */
/* Direct banded LU (no pivoting; leading minors are non-singular because
 * Dirichlet rows are unit rows and the free block is SPD) of the
 * row-replaced system.  Returns 0 on success. */
int orc_solve_banded(const orc_problem* p, const int64_t* rowptr, const int32_t* col,
                     const double* val, const double* b, double* x);
/* Jacobi-PCG on the symmetrically eliminated system (same solution).  x holds
 * the initial guess.  Stops at ||b-Ax||/||b|| <= rtol. */
int orc_solve_pcg(const orc_problem* p, const int64_t* rowptr, const int32_t* col,
                  const double* val, const double* b, double* x, double rtol, int32_t maxit,
                  orc_solve_info* info);
/* One reference step: reassemble A and b from u (in/out), apply Dirichlet,
 * solve (method 0 = banded LU, 1 = PCG to rtol or maxit iterations), write u+ into u. */
int orc_step(const orc_problem* p, const double* D11, const double* D22, const double* D12,
             double* u, int32_t method, double rtol, int32_t maxit, orc_solve_info* info);

/* ---- cells ---- */
/* Covered-node set S_c of one cell, in the reference scan order (row-major
 * over the search window).  Writes up to cap (i,j) pairs; returns |S_c|
 * (including the centre-node fallback when the scan finds nothing). */
int64_t orc_cell_coverage(const orc_grid* g, const orc_cell* c, int32_t* ij, int64_t cap);
/* Search window [i1,i2]x[j1,j2] of one cell (eQabm.cpp:274-287). */
void orc_cell_window(const orc_grid* g, const orc_cell* c, int32_t* win4);
/* Sequential read-then-write loop of updateCells over cells in list order for
 * nspecies fields (natural order, length nodes_w*nodes_h each).  q is
 * nc x nspecies (amount per cell, divided over |S_c|), sampled (may be NULL)
 * receives the per-cell mean read before the cell's own write.  counts (may
 * be NULL) receives |S_c|. */
void orc_couple(const orc_grid* g, const orc_cell* cells, int64_t nc, int32_t nspecies,
                double* const* fields, const double* q, double* sampled, int32_t* counts);

/* orc_couple with each cell's amount formed from its own reading: per cell in
 * list order, e_s = mean of species s over S_c, then q_s = a_s + b_s e_s is
 * written as q_s/|S_c| (a strain whose membrane exchange dt*rate*(iHSL - eHSL),
 * src/Strain.h:312-313, is affine in the sample; src/abm/eQabm.cpp:384-405). */
void orc_couple_affine(const orc_grid* g, const orc_cell* cells, int64_t nc, int32_t ns, double* const* fields,
                       const double* qa, const double* qb, double* sampled, int32_t* counts);

/* setDiffusionTensor of eQabm::updateCells (src/abm/eQabm.cpp:241-248,
 * 306-325, 407): D11, D22, D12 (natural order, nodes_w*nodes_h each) reset to
 * (1, 1, 0), then for each cell in list order every node of S_c (the centre
 * fallback included) takes D11 = Dx cos^2 a + Dy sin^2 a, D22 = Dx sin^2 a +
 * Dy cos^2 a, D12 = (Dx - Dy) sin a cos a for the cell's angle a: the last
 * cell covering a node wins.  Dx, Dy: AnisotropicDiffusion_Axial/_Transverse. */
void orc_cell_tensor(const orc_grid* g, const orc_cell* cells, int64_t nc, double Dx, double Dy, double* D11,
                     double* D22, double* D12);

/* ---- finite-difference alternative: diffusionPETSc ----
 * One backward-Euler step of the 5-point scheme of diffusionPETSc
 * (diffuclass.cpp): interior rows (1 + 2Fx + 2Fy) u - Fx (u_W + u_E) -
 * Fy (u_S + u_N) = u_old with Fx = D dt/hx^2, Fy = D dt/hy^2 (MyMatMult,
 * 787-865; the reference has hx = hy = h, 361); wall rows by ghost
 * reflection: the inward coupling doubles and a Robin wall adds
 * 2 h F Dc/Nc to the diagonal and 2 F h BV/Nc to the right-hand side
 * (ApplyBoundaryConditions, 191-339); Dirichlet wall rows are identity with
 * the wall value, the left/right value winning at corners (the right/left
 * block of ApplyBoundaryConditions runs after the top/bottom one).  Wall
 * kinds map as Dirichlet: Nc = 0; Neumann: Nc = 1, Dc = BV = 0; Robin
 * (flux D du/dn = -r (u - s)): Nc = 1, Dc = r/D, BV = r s/D.  A corner where
 * a Neumann/Robin wall meets a Dirichlet wall is a Dirichlet node (the
 * reference's matrix row there reads past the grid).  Solved exactly (banded
 * LU without pivoting: the rows are diagonally dominant); the reference runs
 * KSPFBCGSR without preconditioner at PETSc's default tolerances (389-391).
 * u is updated in place; returns 0 on success. */
int orc_fd_step(const orc_problem* p, double* u);

/* The channel part of fenicsInterface::stepDiffusion (src/fHSL.cpp:110-152)
 * after the trap solve: the flux into each channel from the trap field u
 * (computeBoundaryFlux, src/fHSL.cpp:54-96: dt*D*(u_inner - u_wall)/wellScaling
 * per column), then `substeps` Crank-Nicolson steps of dt/substeps of the P1
 * advection-diffusion form of AdvectionDiffusion.ufl:62-64 on the interval
 * mesh of the trap's columns, each preceded by adding flux/substeps.  top and
 * bottom (nx+1 values, natural order) are updated in place; they are the
 * next trap solve's Dirichlet data on those walls. */
void orc_channel_update(const orc_problem* p, const orc_channel* top_ch, const orc_channel* bottom_ch,
                        const double* u, double* top, double* bottom);

/* Simulation::computeBoundaryWell (src/simulation.cpp:581-604): the well
 * concentration after one step that produced totalBoundaryFlux = flux:
 *   w += flux / well_scaling;  w -= dt * decay_rate * w;
 *   w < 0: w = 0 and *underflows is incremented.
 * With the DIRICHLET_UPDATE boundary the result is every wall's Dirichlet
 * value for the next step (setBoundaryValues, src/fHSL.cpp:601-603). */
double orc_boundary_well(double w, double flux, double well_scaling, double decay_rate, double dt,
                         int64_t* underflows);

/* boundary-flux functional of stepDiffusion (fenics/boundary.ufl:9-12):
 * -int_{boundary} grad(u).n ds over all walls, P1 gradient per boundary triangle */
double orc_boundary_flux(const orc_problem* p, const double* u);

#ifdef __cplusplus
}
#endif
#endif
