/*
 * eq_field.h -- C ABI of the MI355X quorum-sensing field step.
 *
 * One "problem" is one HSL species on one trap: the unit that eQ's
 * fenicsInterface solves on its own MPI rank (src/simulation.cpp:38-41,
 * src/fHSL.cpp:98-161).  A batch holds many problems grouped by trap
 * (problem p = trap * nspecies + species) and advances all of them with one
 * launch per kernel.  A single field (eq_field_*) is a batch of one.
 *
 * Conventions
 *   - Status: every entry point returns 0 (EQ_OK) or a positive eq_status;
 *     nothing throws across the ABI; eq_last_error() describes the last
 *     failure of the calling thread.
 *   - Host pointers are caller-owned and are read or written before return.
 *     *_dev entry points take device pointers and are ordered on the batch's
 *     HIP stream (eq_batch_stream); eq_batch_sync waits for it.
 *   - Node order is natural: node (i,j), row i = y, column j = x, index
 *     j + i*(nx+1) (src/eQ.h:131-134).  eq_*_vertex_map returns the identity
 *     dof_from_vertex, so the host's (i,j)->dof lookup tables stay valid.
 *   - A handle is not thread-safe; distinct handles may be used concurrently.
 */
#ifndef EQ_FIELD_H
#define EQ_FIELD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    EQ_OK = 0,
    EQ_ERR_INVALID = 1,       /* bad argument or descriptor            */
    EQ_ERR_HIP = 2,           /* HIP runtime failure                   */
    EQ_ERR_NOMEM = 3,         /* device or host allocation failed      */
    EQ_ERR_NOT_CONVERGED = 4, /* a solve hit max_iter above rtol       */
    EQ_ERR_NO_DEVICE = 5      /* no HIP device visible                 */
} eq_status;

/* wall order of eQ::boundaryCondition::wall (src/eQ.h:335-340) */
typedef enum { EQ_WALL_LEFT = 0, EQ_WALL_RIGHT = 1, EQ_WALL_TOP = 2, EQ_WALL_BOTTOM = 3 } eq_wall;

/* Per-wall condition, decoded as fenicsInterface::createHSL does
 * (src/fHSL.cpp:467-539): Dirichlet u = value; Neumann adds nothing;
 * Robin adds dt*r*(u - s) on the wall. */
typedef enum { EQ_BC_NEUMANN = 0, EQ_BC_DIRICHLET = 1, EQ_BC_ROBIN = 2 } eq_bc_kind;

typedef struct {
    int32_t kind;
    int32_t pad_;
    double value; /* Dirichlet value                 */
    double r;     /* Robin rate                      */
    double s;     /* Robin exterior concentration    */
} eq_bc;

/* One species on one trap: the trap mesh of fenicsInterface::createMesh
 * (RectangleMesh (0,0)-(width,height), nx x ny cells, "right" diagonal,
 * src/fHSL.cpp:164-173) and the hslD form (fenics/hslD.ufl:31-49) with the
 * reaction term added:  (M + dt*D*K + dt*gamma*M) u+ = M u + dt*f*int(phi). */
typedef enum {
    EQ_SCHEME_P1 = 0, /* P1 finite elements of fenicsInterface (fenics/hslD.ufl)            */
    EQ_SCHEME_FD = 1  /* 5-point finite differences of diffusionPETSc (diffuclass.cpp):
                       * interior (1 + 2Fx + 2Fy) u - Fx (u_W + u_E) - Fy (u_S + u_N) = u_old,
                       * Fx = D dt/hx^2; walls by ghost reflection (MyMatMult,
                       * diffuclass.cpp:787-865; ApplyBoundaryConditions 191-339), Robin
                       * walls as D du/dn = -r (u - s); solved in the equivalent symmetric
                       * form (rows scaled by hx*hy/2 on walls, hx*hy/4 in corners) */
} eq_scheme;

/* How the walls are found on the mesh.  The reference marks a wall's facets
 * with near(coordinate, W or H) on the vertex coordinates DOLFIN computes,
 * x_j = (j*W)/nx and y_i = (i*H)/ny (DirichletBoundary_TrapEdge::inside,
 * src/fHSL.h:575-590; RectangleMesh, src/fHSL.cpp:169-172), with
 * near(a, b) = |a - b| < DOLFIN_EPS = 3e-16.  For an integer trap size these
 * coordinates are exact and both modes find every wall.  For other sizes the
 * right or top coordinate can miss by an ulp (e.g. 218 cells over 21.8 um),
 * and the reference then applies no condition on that wall. */
typedef enum {
    EQ_WALLS_GRID = 0,  /* the outer rows and columns of the grid (default) */
    EQ_WALLS_COORD = 1  /* the reference's coordinate test: a wall it misses
                         * becomes Neumann (no Dirichlet, no Robin term) */
} eq_wall_marking;

typedef struct {
    int32_t nx, ny;
    double width, height; /* microns */
    double dt, D, gamma, f;
    eq_bc wall[4];
    int32_t scheme;       /* eq_scheme (0: the P1 form above)  */
    int32_t marking;      /* eq_wall_marking (0: grid rows and columns) */
} eq_field_desc;

typedef enum {
    EQ_SOLVER_JACOBI_CG = 0,    /* FP64 CG, Jacobi preconditioner                    */
    EQ_SOLVER_GMG_CG = 1,       /* FP64 CG, FP64 geometric-multigrid V-cycle         */
    EQ_SOLVER_MIXED_GMG_CG = 2, /* FP64 CG, FP32 V-cycle (FP64 residual/refinement)  */
    EQ_SOLVER_GMG_BICGSTAB = 3, /* FP64 BiCGStab, FP64 V-cycle (nonsymmetric: advection) */
    EQ_SOLVER_MIXED_GMG_BICGSTAB = 4, /* FP64 BiCGStab, FP32 V-cycle                  */
    EQ_SOLVER_IR_GMG_CG = 5     /* FP64 iterative refinement around FP32 GMG-CG: the
                                   residual b - A x and the iterate x in FP64; each
                                   correction solved by CG in FP32 (FP32 vectors and
                                   stencil, FP32 V-cycle) to ir_inner_rtol          */
} eq_solver_kind;

typedef struct {
    int32_t kind;                /* eq_solver_kind                                  */
    int32_t max_iter;            /* per solve                                       */
    double rtol;                 /* stop at ||b - A u+||_2 <= rtol * ||b||_2 (FP64) */
    int32_t mg_pre, mg_post;     /* damped-Jacobi sweeps per level                  */
    double mg_omega;             /* Jacobi damping (of every sweep, or with
                                    mg_omega2 of every first sweep of a pair)       */
    int32_t mg_coarse_max_nodes; /* coarsen until a level has at most this many    */
    int32_t mg_max_levels;
    int32_t check_every;         /* iterations between host convergence checks      */
    int32_t use_graphs;          /* capture iteration chunks in hipGraphs (0/1)     */
    int32_t mg_fuse_max_nodes;   /* levels with fewer nodes (all problems) run the
                                    fused multigrid kernels; <= 0: 524288            */
    int32_t mg_fuse_pre_max_nodes; /* the same for the two pre-sweeps in one launch;
                                      <= 0: mg_fuse_max_nodes                       */
    int32_t mg_bottom_lds_bytes; /* LDS one workgroup may use for the levels it runs
                                    (the bottom of the V-cycle); <= 0 (default): 40960
                                    for batches of fewer than 64 problems, else
                                    153600                                          */
    int32_t mg_fuse_pre_top;     /* 1: the finest level's first two sweeps run in one
                                    launch (pays off on large fields)               */
    int32_t tile4_min_nodes;     /* coarse levels with at least this many nodes (all
                                    problems) run 4 rows per thread; <= 0: 8 Mi     */
    int32_t tile8_min_nodes;     /* levels with at least this many run 8 rows per
                                    thread (level 0 runs at least 4); <= 0: 32 Mi   */
    int32_t extrapolate;         /* 1: start each solve from 2 u_n - u_(n-1) once two
                                    solutions are kept; 0 (default): from u_n       */
    int32_t ir_max_outer;        /* EQ_SOLVER_IR_GMG_CG: refinement steps per solve  */
    double ir_inner_rtol;        /* EQ_SOLVER_IR_GMG_CG: each inner FP32 solve reduces its
                                    residual by this much (default 1e-4), or by less
                                    when less reaches rtol                          */
    int32_t mg_level_fused;      /* 1 (default): a coarse level smoothed with two pre-
                                    and two post-sweeps runs its way down (sweeps,
                                    residual, restriction) in one launch and its way
                                    up (correction, sweeps) in another; 0: the
                                    separate launches                               */
    double mg_omega2;            /* damping of every second sweep (0: mg_omega):
                                    pre-sweeps alternate mg_omega, mg_omega2, ...;
                                    post-sweeps take the same dampings in reverse,
                                    so the preconditioner stays symmetric.  The
                                    defaults 0.5617, 1.3895 are the degree-2
                                    Chebyshev dampings on [0.5, 2]                  */
    int32_t streams;             /* CG kinds, every level fused: the batch's problems
                                    solved as this many groups (1 to 4, default 2;
                                    at most one per problem), each on its own stream,
                                    each stopping at its own iteration count (same
                                    per-problem results)                            */
} eq_solver_desc;

typedef struct {
    int32_t iterations_max;   /* over problems, last step                        */
    int32_t nconverged;       /* problems at or below rtol                       */
    double relres_max;        /* max over problems of ||b-Au||/||b||             */
    double seconds;           /* host wall time of the call                      */
} eq_step_stats;

/* A rod cell as the ABM sees it: centre, angle, current length and width in
 * microns (Ecoli::getCenter_x/y, getAngle, getLengthMicrons, getWidthMicrons,
 * src/abm/Ecoli.h:39-43).  Array order = the host's list iteration order. */
typedef struct {
    double x, y, angle, length, width;
} eq_cell;

/* coupling flags */
#define EQ_COUPLE_SAMPLE 1u  /* write per-cell means into `sampled`  */
#define EQ_COUPLE_DEPOSIT 2u /* add q_c/|S_c| to every covered node   */
#define EQ_COUPLE_TENSOR 4u  /* set the diffusion tensor of the traps enabled by
                                eq_batch_cell_tensor from the cells' angles   */

typedef struct eq_batch eq_batch;
typedef struct eq_batch eq_field;

const char* eq_last_error(void);
/* found[w] = 1 when wall w (left, right, top, bottom) of this mesh is found by
 * the reference's coordinate test (see eq_wall_marking); left and bottom
 * always are.  A host can call it before eq_batch_create to learn which walls
 * a reference run would have dropped. */
int eq_walls_found(const eq_field_desc* desc, int32_t found[4]);
/* fills the defaults: mixed GMG-CG, rtol 1e-10, V(2,2), omega 0.8, coarse <= 256 nodes,
 * graphs on */
void eq_solver_default(eq_solver_desc* s);

/* ---------------- batch API ---------------- */
/* Replaces one fenicsInterface per species and rank (fenicsInterface::initDiffusion,
 * src/fHSL.cpp:37-53 -> fenicsClassInit 195-328 -> createHSL 367-599).
 * descs has nprob entries, problems t*nspecies .. t*nspecies+nspecies-1 share
 * trap t (same nx, ny, width, height). */
int eq_batch_create(const eq_field_desc* descs, int32_t nprob, int32_t nspecies,
                    const eq_solver_desc* solver, eq_batch** out);
void eq_batch_destroy(eq_batch* b);
int32_t eq_batch_num_problems(const eq_batch* b);
int64_t eq_batch_num_nodes(const eq_batch* b, int32_t prob);
void* eq_batch_stream(const eq_batch* b); /* hipStream_t */
int eq_batch_sync(eq_batch* b);

/* u0->vector()->set_local(solution_vector)  (src/fHSL.cpp:104) */
int eq_batch_upload(eq_batch* b, int32_t prob, const double* u, size_t n);
/* u->vector()->get_local(solution_vector)   (src/fHSL.cpp:108) */
int eq_batch_download(eq_batch* b, int32_t prob, double* u, size_t n);
/* device-pointer variants (natural order, stream ordered) */
int eq_batch_upload_dev(eq_batch* b, int32_t prob, const double* d_u, size_t n);
int eq_batch_download_dev(eq_batch* b, int32_t prob, double* d_u, size_t n);

/* LVS->solve() for every problem  (src/fHSL.cpp:105-106): rebuild b = M u
 * (+ sources, Robin), apply the Dirichlet walls, solve to rtol. */
int eq_batch_step(eq_batch* b, eq_step_stats* stats);
/* per-problem iterations and final relative residual of the last step */
int eq_batch_problem_stats(const eq_batch* b, int32_t* iterations, double* relres);

/* updateCells' coverage/read/write loop (src/abm/eQabm.cpp:234-449) for one
 * trap and all of its species.  npm = nodesPerMicronSignaling.  q is
 * nc x nspecies, sampled (may be NULL) nc x nspecies, counts (may be NULL) nc.
 * Per cell in array order: sample every species, then deposit q/|S_c|. */
int eq_batch_couple(eq_batch* b, int32_t trap, const eq_cell* cells, size_t nc, double npm,
                    const double* q, double* sampled, int32_t* counts, uint32_t flags);
/* Per-node Dirichlet values on one wall of problem `prob` (a DirichletBC bound to a
 * Function rather than a Constant, src/fHSL.cpp:514, 531): values[k] is the value at
 * column k (top/bottom walls, n = nx+1) or row k (left/right, n = ny+1); the wall
 * must be Dirichlet.  NULL reverts to the wall's constant value. */
int eq_batch_set_wall_values(eq_batch* b, int32_t prob, int32_t wall, const double* values, size_t n);

/* Flow channels above (top) and below (bottom) a trap -- fenicsChannel
 * (src/fHSL.h:240-332) with the coupling of fenicsInterface::stepDiffusion
 * (src/fHSL.cpp:110-152): the channel's values are the trap's Dirichlet data on
 * that wall; after every solve the flux dt*D*(u_inner - u_wall)/well_scaling per
 * column (computeBoundaryFlux, src/fHSL.cpp:54-96) is added to the channel in
 * `substeps` equal parts, each followed by a Crank-Nicolson step of dt/substeps of
 * the P1 advection-diffusion form of fenics/AdvectionDiffusion.ufl:62-64 (velocity
 * v along +x, Robin ends r*(u - s) at x = 0 and x = width). */
typedef struct {
    double v;
    double r_left, s_left;
    double r_right, s_right;
    double well_scaling;
    int32_t substeps;
    int32_t pad_;
} eq_channel_desc;
/* top/bottom may be NULL (no channel on that side); the walls with a channel must be
 * Dirichlet.  Channels start at zero (src/fHSL.cpp:423-426). */
int eq_batch_set_channels(eq_batch* b, int32_t prob, const eq_channel_desc* top, const eq_channel_desc* bottom);
/* copy the channel values (nx+1 each, natural order) to the host; NULL skips a side */
int eq_batch_channel_values(eq_batch* b, int32_t prob, double* top, double* bottom);

/* Per-node diffusion tensor of problem `prob`: D11, D22, D12 (n = (nx+1)*(ny+1) each,
 * natural order) are fenicsInterface's D11/D22/D12 vectors (src/fHSL.cpp:313-323,
 * 387-392), P1-interpolated into A = D*[[D11, D12], [D12, D22]] of fenics/hslD.ufl:31
 * (AnisotropicDiffusionTensor::eval, src/Expressions.h:95-124).  All three NULL:
 * back to constant coefficients (the reference's default 1, 1, 0).  Copies the
 * arrays; call between steps. */
int eq_batch_set_tensor(eq_batch* b, int32_t prob, const double* D11, const double* D22, const double* D12,
                        size_t n);

/* setDiffusionTensor of eQabm::updateCells on the device (src/abm/eQabm.cpp:241-248,
 * 306-325, 407).  enable = 1: every species of trap `trap` takes its D11/D22/D12
 * from the coupling calls that carry EQ_COUPLE_TENSOR: the tensor starts from
 * (1, 1, 0) and each cell, in list order, writes on its covered nodes
 *   D11 = Dx cos^2 a + Dy sin^2 a,  D22 = Dx sin^2 a + Dy cos^2 a,
 *   D12 = (Dx - Dy) sin a cos a          (a = the cell's angle)
 * so the last cell covering a node wins.  The coupling call then carries the
 * tensor to the coarse levels (nearest fine node) and rebuilds the coarsest
 * inverses, all on the batch stream with no host round trip.  Dx, Dy are
 * AnisotropicDiffusion_Axial / _Transverse (src/main.cpp:569-570).  enable = 0:
 * constant coefficients again.  Set-up call (synchronises); P1 scheme only, not
 * with a velocity. */
int eq_batch_cell_tensor(eq_batch* b, int32_t trap, double Dx, double Dy, int32_t enable);
/* the finest-level tensor of problem `prob` in natural order ((nx+1)(ny+1)
 * values each; constant coefficients read as 1, 1, 0).  Synchronises. */
int eq_batch_tensor_values(eq_batch* b, int32_t prob, double* D11, double* D22, double* D12, size_t n);

/* Advection (no counterpart in the reference's trap form; the channel-flow case of
 * the config-4 workload): x-velocity per grid row, vx[i] for row i (n = ny+1),
 * P1 in y, adding dt * int (v du/dx) phi (Galerkin) to the operator.  The system
 * is then nonsymmetric: solve it with a BiCGStab solver kind.  NULL removes it. */
int eq_batch_set_velocity(eq_batch* b, int32_t prob, const double* vx, size_t n);

/* Autotuner (no counterpart in the reference, which has one fixed direct solve):
 * times `repeats` steps of each candidate multigrid cycle shape (pre/post sweeps
 * (1,1), (2,1), (2,2), (3,3)) from the batch's current state, restoring the fields,
 * the kept solutions and the Krylov state after every trial, and keeps the fastest
 * shape whose solves all converged.  The state afterwards is the state before.
 * `chosen` (may be NULL) receives the resulting solver settings. */
int eq_batch_autotune(eq_batch* b, int32_t repeats, eq_solver_desc* chosen);

/* device-pointer variant: d_cells/d_q/d_sampled as above, on the batch stream */
int eq_batch_couple_dev(eq_batch* b, int32_t trap, const eq_cell* d_cells, size_t nc, double npm,
                        const double* d_q, double* d_sampled, int32_t* d_counts, uint32_t flags);
/* The same for the traps trap0 .. trap0+ntraps-1 in one pass (one eQabm::updateCells
 * per trap, src/abm/eQabm.cpp:234-449, the traps being independent): their cells are
 * concatenated in d_cells, cell_offsets (HOST array, ntraps+1 entries) gives the first
 * cell of each trap and the total; q/sampled/counts follow the concatenated order. */
int eq_batch_couple_traps_dev(eq_batch* b, int32_t trap0, int32_t ntraps, const int64_t* cell_offsets,
                              const eq_cell* d_cells, double npm, const double* d_q, double* d_sampled,
                              int32_t* d_counts, uint32_t flags);

/* updateCells for strains whose amount depends on the cell's own sample
 * (Strain::computeProteins with membrane exchange, dt*membraneRate*(iHSL - eHSL),
 * src/Strain.h:312-313, is affine in the sampled eHSL): per cell in list order
 * and per species, e = the mean over S_c of the current field (after every
 * earlier cell's deposit), then q = a + b*e is added as q/|S_c| on S_c -- the
 * reference's read -> compute -> write order inside one step
 * (src/abm/eQabm.cpp:384-405).  a, b, sampled: nc x nspecies (device); cells and
 * offsets as eq_batch_couple_traps_dev.  One workgroup per trap walks its cells
 * in order, so a trap's cells are a sequential chain: use the q-upfront entry
 * points when the amounts do not depend on the sample. */
int eq_batch_couple_affine_dev(eq_batch* b, int32_t trap0, int32_t ntraps, const int64_t* cell_offsets,
                               const eq_cell* d_cells, double npm, const double* d_a, const double* d_b,
                               double* d_sampled, int32_t* d_counts);

/* No counterpart in the reference.  eq_batch_couple_traps_dev for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): same arguments, same results, and it
 * calls no hipMalloc, hipFree or allocating library routine.  The coupling's
 * scratch (cell records, cell/tile pairs, sort buffers, per-pair sums) lives in
 * one fixed 64 MiB device buffer per GPU, part of the library's device code and
 * held until the process ends.  A call that needs more returns EQ_ERR_INVALID
 * before it touches any field (couple fewer traps per call: about 300 000 rods
 * of 2-4 um at 10 nodes per micron fit).  Every call on one GPU shares that
 * GPU's buffer: calls that overlap on one GPU (other threads, other streams)
 * race, so a caller serialises them. */
int eq_batch_couple_traps_dev_noalloc(eq_batch* b, int32_t trap0, int32_t ntraps, const int64_t* cell_offsets,
                                      const eq_cell* d_cells, double npm, const double* d_q, double* d_sampled,
                                      int32_t* d_counts, uint32_t flags);

/* totalBoundaryFlux = D*dt * (-int grad(u).n ds)  (src/fHSL.cpp:156-160,
 * fenics/boundary.ufl:9-12) of the current field */
int eq_batch_boundary_flux(eq_batch* b, int32_t prob, double* total);

/* setBoundaryValues(bval) (src/fHSL.cpp:601-603): every Dirichlet wall of
 * problem `prob` takes the value `value` from the next solve on (the one
 * updatingDirchletBoundary that all four DirichletBCs of the DIRICHLET_UPDATE
 * boundary share, src/fHSL.cpp:545-554, src/Expressions.h:13-44).  Walls with
 * per-node values or a flow channel keep those.  Stream ordered: one small
 * kernel on the batch stream, no table upload and no host synchronisation, so
 * a host may call it between steps or while capturing its own graph. */
int eq_batch_set_boundary_value(eq_batch* b, int32_t prob, double value);

/* The boundary well of Simulation::computeBoundaryWell with the DIRICHLET_UPDATE
 * boundary (src/simulation.cpp:581-604, 612-624), run on the device at the end
 * of every eq_batch_step, after the solve and the channels:
 *   flux = totalBoundaryFlux of the new field        (src/fHSL.cpp:156-160)
 *   w   += flux / well_scaling
 *   w   -= dt * decay_rate * w
 *   w < 0: w = 0 and one underflow is counted
 *   every Dirichlet wall takes w for the next solve  (setBoundaryValues)
 * well_scaling is the reference's wellScaling (2*10*(15/lengthScaling)*(W+H)),
 * decay_rate its boundaryDecayRate (simulationFlowRate / trap width), w0 the
 * starting concentration (0 in initBoundaryWell).  The problem's four walls
 * must be Dirichlet, as in the DIRICHLET_UPDATE set-up.  NULL removes the
 * well.  Set-up call: synchronises the batch stream. */
typedef struct {
    double well_scaling;
    double decay_rate;
    double w0;
} eq_well_desc;
int eq_batch_set_well(eq_batch* b, int32_t prob, const eq_well_desc* well);
/* the well's concentration, the last boundary flux it took and the underflow
 * count (src/simulation.cpp:591-595); any pointer may be NULL.  Synchronises. */
int eq_batch_well_state(eq_batch* b, int32_t prob, double* w, double* last_flux, int64_t* underflows);
/* shell->mesh_coords (2N, vertex order) and shell->dof_from_vertex (N)
 * (src/fHSL.h:82-88); vertex order = natural order, dof_from_vertex = identity */
int eq_batch_vertex_map(const eq_batch* b, int32_t prob, double* coords, int32_t* dof_from_vertex);

/* Matrix-free counterparts of the reference's assembled system, for checking:
 *   eq_batch_apply: y = A x with A = (1+dt*gamma)M + dt*D*K + Robin after
 *     DirichletBC::apply (Dirichlet rows are identity rows), i.e. the matrix
 *     LinearVariationalSolver factorises (assemble(A,a) + bc.apply(A),
 *     src/fHSL.h:156-160 [ext]);
 *   eq_batch_rhs: b = M u0 + dt*f*int(phi) + Robin, Dirichlet rows set to the
 *     wall value (assemble(b,L) + bc.apply(b)).
 * Natural order, host pointers, length (nx+1)*(ny+1).  Uses the batch's CG
 * scratch vectors: call between steps only. */
int eq_batch_apply(eq_batch* b, int32_t prob, const double* x, double* y, size_t n);
int eq_batch_rhs(eq_batch* b, int32_t prob, const double* u0, double* rhs, size_t n);

/* ---------------- single-field API (batch of one) ---------------- */
int eq_field_create(const eq_field_desc* desc, const eq_solver_desc* solver, eq_field** out);
void eq_field_destroy(eq_field* f);
int eq_field_upload(eq_field* f, const double* u, size_t n);
int eq_field_download(eq_field* f, double* u, size_t n);
int eq_field_step(eq_field* f, eq_step_stats* stats);
int eq_field_vertex_map(const eq_field* f, double* coords, int32_t* dof_from_vertex);
int eq_field_boundary_flux(eq_field* f, double* total);
int eq_field_set_tensor(eq_field* f, const double* D11, const double* D22, const double* D12, size_t n);
int eq_cells_couple(eq_field* f, const eq_cell* cells, size_t nc, double npm, const double* q,
                    double* sampled, int32_t* counts, uint32_t flags);

/* ---------------- measurement ---------------- */
/* Kernel families for per-launch HIP-event timing on the batch stream.  Every
 * launch of the library belongs to one. */
typedef enum {
    EQ_KERNEL_STENCIL64 = 0,      /* Krylov operator apply q = A p (CG), v, t = A x^ (BiCGStab) */
    EQ_KERNEL_SMOOTH = 1,         /* finest-level Jacobi sweeps of the V-cycle               */
    EQ_KERNEL_CG_UPDATE = 2,      /* x, r update + dot products (CG, BiCGStab)               */
    EQ_KERNEL_COUPLE = 3,         /* cell sample/deposit per tile (k_tile_couple)            */
    EQ_KERNEL_SMOOTH_COARSE = 4,  /* Jacobi sweeps on the coarse levels                      */
    EQ_KERNEL_TRANSFER0 = 5,      /* finest-level residual, restriction from and prolongation
                                     to it (fused forms included)                          */
    EQ_KERNEL_TRANSFER_COARSE = 6,/* the same between coarse levels                          */
    EQ_KERNEL_BOTTOM = 7,         /* the LDS bottom V-cycle (one workgroup per problem)      */
    EQ_KERNEL_FINISH = 8,         /* per-problem reductions of the tile partials             */
    EQ_KERNEL_SOLVE_ONCE = 9,     /* once per solve: right-hand side, initial and true
                                     residuals, first sweeps, copies, the kept solution    */
    EQ_KERNEL_BICG_VEC = 10,      /* BiCGStab direction and s updates                        */
    EQ_KERNEL_COUPLE_SETUP = 11,  /* cell geometry, pair scan, pairs, sort, tile bounds,
                                     per-cell sample sums                                  */
    EQ_KERNEL_BOUNDARY = 12,      /* channels, boundary wells, boundary values               */
    EQ_KERNEL_COUNT_ = 13
} eq_kernel_id;
/* enable (1) / disable (0) event timing around every launch of `kernel`;
 * timing forces eager launches (no graph replay) while enabled */
int eq_batch_profile(eq_batch* b, int32_t kernel, int32_t enable);
/* launches, summed event time (ms) and algorithmic bytes summed over launches */
int eq_batch_profile_read(eq_batch* b, int32_t kernel, int64_t* launches, double* total_ms,
                          double* bytes);
/* Algorithmic HBM bytes (each operand array read once and each output written
 * once per launch, at its element size) and launches issued by the batch since
 * the last reset, graph replays included: the numerator of a whole-step
 * roofline fraction.  reset = 1 zeroes the counters after reading them. */
int eq_batch_traffic(eq_batch* b, double* bytes, int64_t* launches, int32_t reset);

#ifdef __cplusplus
}
#endif
#endif /* EQ_FIELD_H */
