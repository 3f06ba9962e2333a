"""Pins the CPU oracle (FFC-quadrature restatement of eQ's field step) against
independent facts: the textbook-assembly golden fixtures, closed-form stencil
entries, conservation/invariance laws, hand-countable cell coverages and a
brute-force coverage scan.  CPU only."""
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden", "p1_tiny.npz")


def _problem(O, params, walls):
    nx, ny, W, H, dt, D, gamma, f = params
    wd = {w: (int(walls[w][0]), walls[w][1], walls[w][2], walls[w][3]) for w in range(4) if walls[w][0] != 0}
    return O.problem(int(nx), int(ny), W, H, dt, D, gamma, f, walls=wd)


def _cases():
    z = np.load(GOLD)
    return [str(n) for n in z["names"]]


@pytest.mark.parametrize("name", _cases())
def test_oracle_matches_textbook_fixture(oracle, name):
    O = oracle
    z = np.load(GOLD)
    p = _problem(O, z[f"{name}/params"], z[f"{name}/walls"])
    T = z[f"{name}/tensor"]
    n = O.num_nodes(p)
    rp, c, v = O.assemble(p, T[:, 0], T[:, 1], T[:, 2])
    A = O.to_scipy(rp, c, v, n).toarray()
    Aref = z[f"{name}/A"]
    assert np.abs(A - Aref).max() <= 1e-13 * np.abs(Aref).max()
    b = O.assemble_rhs(p, z[f"{name}/u0"])
    v2, b2 = O.apply_dirichlet(p, rp, c, v, b)
    A2 = O.to_scipy(rp, c, v2, n).toarray()
    assert np.abs(A2 - z[f"{name}/A_bc"]).max() <= 1e-13 * np.abs(Aref).max()
    assert np.abs(b2 - z[f"{name}/b_bc"]).max() <= 1e-13 * max(1.0, np.abs(b2).max())
    x = O.solve_banded(p, rp, c, v2, b2)
    xr = z[f"{name}/x"]
    assert np.linalg.norm(x - xr) <= 1e-12 * np.linalg.norm(xr)


def test_interior_stencil_closed_form(oracle):
    O = oracle
    h, dt, D, g = 0.25, 0.1, 400.0, 0.01
    p = O.problem(8, 6, 2.0, 1.5, dt, D, g, walls={O.TOP: (O.DIRICHLET, 0.0, 0, 0)})
    rp, c, v = O.assemble(p)
    A = O.to_scipy(rp, c, v, O.num_nodes(p)).toarray()
    nw = 9
    k = 3 + 2 * nw
    m = 1.0 + dt * g
    assert math.isclose(A[k, k], m * h * h / 2 + 4 * dt * D, rel_tol=1e-14)
    for off in (1, -1, nw, -nw):
        assert math.isclose(A[k, k + off], m * h * h / 12 - dt * D, rel_tol=1e-14)
    for off in (nw + 1, -nw - 1):  # NE, SW: mass only (stiffness coupling is zero)
        assert math.isclose(A[k, k + off], m * h * h / 12, rel_tol=1e-12)
    for off in (nw - 1, -nw + 1):  # NW, SE: not mesh edges of the "right" diagonal
        assert A[k, k + off] == 0.0
    # corner (0,0) belongs to two triangles, edge node (0,3) to three
    assert math.isclose(A[0, 0], m * h * h / 6 + dt * D, rel_tol=1e-13)
    assert math.isclose(A[3, 3], m * h * h / 4 + 2 * dt * D, rel_tol=1e-13)


def test_row_sums_are_mass(oracle):
    """K has zero row sums, so A.1 = (1+dt*gamma) M.1 and sum(M) = W*H."""
    O = oracle
    p = O.problem(10, 7, 3.0, 2.1, 0.1, 250.0, 0.05)
    rp, c, v = O.assemble(p)
    A = O.to_scipy(rp, c, v, O.num_nodes(p))
    ones = np.ones(O.num_nodes(p))
    Aone = A @ ones
    b = O.assemble_rhs(p, ones)  # = M.1
    assert np.allclose(Aone, (1.0 + 0.1 * 0.05) * b, rtol=1e-11, atol=1e-13)
    assert math.isclose(b.sum(), 3.0 * 2.1, rel_tol=1e-13)


def test_constant_field_invariant_and_decay(oracle):
    O = oracle
    p = O.problem(12, 6, 3.0, 1.5, 0.1, 400.0, 0.0)
    u = np.full(O.num_nodes(p), 0.7)
    u1, info = O.step(p, u, "lu")
    assert np.abs(u1 - 0.7).max() < 1e-12
    p = O.problem(12, 6, 3.0, 1.5, 0.1, 400.0, 0.2)
    u1, info = O.step(p, u, "lu")
    assert np.abs(u1 - 0.7 / 1.02).max() < 1e-12


def test_mass_conservation_neumann(oracle):
    O = oracle
    p = O.problem(16, 8, 4.0, 2.0, 0.1, 100.0, 0.0)
    u = np.random.default_rng(3).random(O.num_nodes(p))
    u1, _ = O.step(p, u, "lu")
    ones = np.ones_like(u)
    M1 = O.assemble_rhs(p, ones)
    assert math.isclose(M1 @ u1, M1 @ u, rel_tol=1e-12)


def test_lu_matches_pcg(oracle):
    O = oracle
    p = O.problem(40, 20, 10.0, 5.0, 0.1, 400.0, 0.01, walls={O.TOP: (O.DIRICHLET, 0.0, 0, 0)})
    u = np.random.default_rng(5).random(O.num_nodes(p))
    x1, i1 = O.step(p, u, "lu")
    x2, i2 = O.step(p, u, "pcg", 1e-13)
    assert np.linalg.norm(x1 - x2) <= 1e-10 * np.linalg.norm(x1)


def test_boundary_flux_closed_forms(oracle):
    O = oracle
    nx, ny, W, H = 10, 5, 2.0, 1.0
    p = O.problem(nx, ny, W, H, 0.1, 1.0)
    xy = O.mesh_coords(p)
    lin = 0.3 * xy[:, 0] - 1.7 * xy[:, 1] + 0.2
    assert abs(O.boundary_flux(p, lin)) < 1e-12  # divergence theorem for a linear field
    quad = xy[:, 0] ** 2
    h = W / nx
    assert math.isclose(O.boundary_flux(p, quad), H * (2 * h - 2 * W), rel_tol=1e-12)


# ---------------------------------------------------------------- cells

def test_coverage_axis_aligned_hand_count(oracle):
    O = oracle
    g = O.grid(41, 21, 4.0, 10.0, 5.0)
    # rod of length 3, width 1 at a node: |dx| < 1, |dy| < 0.5 strictly -> 7 x 3 nodes
    cov = O.cell_coverage(g, [5.0, 2.5, 0.0, 3.0, 1.0])
    assert len(cov) == 21
    assert set(map(tuple, cov)) == {(i, j) for i in (9, 10, 11) for j in range(17, 24)}
    # same rod turned by pi/2: |dy| < 1, |dx| < 0.5
    cov = O.cell_coverage(g, [5.0, 2.5, math.pi / 2, 3.0, 1.0])
    assert set(map(tuple, cov)) == {(i, j) for i in range(7, 14) for j in (19, 20, 21)}
    # scan order is row-major over the window
    assert [tuple(r) for r in cov] == sorted(map(tuple, cov))


def test_coverage_fallback_centre_node(oracle):
    O = oracle
    g = O.grid(21, 11, 2.0, 10.0, 5.0)
    # length == width: zero-length box -> nothing strictly inside -> centre node
    cov = O.cell_coverage(g, [3.3, 1.2, 0.4, 1.0, 1.0])
    assert cov.tolist() == [[2, 7]]


def _brute_coverage(cell, nodes_w, nodes_h, npm):
    x, y, a, L, w = cell
    j = np.arange(nodes_w)
    i = np.arange(nodes_h)
    X, Y = np.meshgrid(j / npm, i / npm)
    lx = np.cos(a) * (X - x) + np.sin(a) * (Y - y)
    ly = -np.sin(a) * (X - x) + np.cos(a) * (Y - y)
    inside = (np.abs(lx) < (L - w) / 2) & (np.abs(ly) < w / 2)
    ii, jj = np.nonzero(inside)
    return set(zip(ii.tolist(), jj.tolist())), inside


def test_coverage_matches_brute_force(oracle):
    """Away from rounding ties the window never clips a unit-width rod, so the
    reference scan equals a brute-force scan of the whole grid."""
    O = oracle
    rng = np.random.default_rng(7)
    npm, W, H = 10.0, 20.0, 10.0
    g = O.grid(201, 101, npm, W, H)
    for _ in range(60):
        cell = [rng.uniform(2, 18), rng.uniform(2, 8), rng.uniform(0, math.pi), rng.uniform(2, 4), 1.0]
        ref, _ = _brute_coverage(cell, 201, 101, npm)
        got = set(map(tuple, O.cell_coverage(g, cell).tolist()))
        if not ref:
            continue
        assert got == ref


def test_wide_rod_is_clipped_by_default_width_window(oracle):
    """nodesToEdge uses the default 1 um width (eQabm.cpp:75): a 2 um wide rod
    keeps only the nodes inside the narrower window."""
    O = oracle
    npm = 10.0
    g = O.grid(101, 101, npm, 10.0, 10.0)
    cell = [5.0, 5.0, 0.0, 4.0, 2.0]
    got = set(map(tuple, O.cell_coverage(g, cell).tolist()))
    ref, _ = _brute_coverage(cell, 101, 101, npm)
    i1, i2, j1, j2 = O.cell_window(g, cell)
    assert (i1, i2) == (45, 55)  # round(5*10) -/+ round(10/2)
    assert got == {(i, j) for (i, j) in ref if i1 <= i <= i2 and j1 <= j <= j2}
    assert len(got) < len(ref)


def test_couple_sequential_semantics(oracle):
    O = oracle
    g = O.grid(41, 21, 4.0, 10.0, 5.0)
    cells = np.array([[5.0, 2.5, 0.0, 3.0, 1.0],    # 21 nodes
                      [5.5, 2.5, 0.0, 3.0, 1.0]])   # overlaps the first on 15 nodes
    u = np.zeros(41 * 21)
    q = np.array([[2.1], [4.2]])
    (f,), sampled, counts = O.couple(g, cells, [u], q)
    assert counts.tolist() == [21, 21]
    assert sampled[0, 0] == 0.0
    # the second cell reads the first cell's deposit on the shared nodes
    assert math.isclose(sampled[1, 0], (15 * 2.1 / 21) / 21, rel_tol=1e-14)
    assert math.isclose(f.sum(), 2.1 + 4.2, rel_tol=1e-14)
    # constant field is sampled exactly
    (f2,), s2, _ = O.couple(g, cells, [np.full(41 * 21, 0.3)], np.zeros((2, 1)))
    assert np.allclose(s2, 0.3, rtol=1e-15, atol=0)


def _p1_tensor_assembly(nx, ny, W, H, dt, D, d11, d22, d12):
    """Independent assembly of dt*int grad(v)^T (D*Dtensor) grad(u) on the
    'right' mesh from the P1 gradients of each triangle and the exact
    integral of a P1 tensor (the mean of its vertex values times |T|)."""
    hx, hy = W / nx, H / ny
    n = (nx + 1) * (ny + 1)
    K = np.zeros((n, n))
    idx = lambda i, j: i * (nx + 1) + j
    area = hx * hy / 2.0
    for i in range(ny):
        for j in range(nx):
            bl, br, tl, tr = idx(i, j), idx(i, j + 1), idx(i + 1, j), idx(i + 1, j + 1)
            # lower-right (BL, BR, TR) and upper-left (BL, TL, TR) triangles
            for verts, grads in (((bl, br, tr), ((-1 / hx, 0.0), (1 / hx, -1 / hy), (0.0, 1 / hy))),
                                 ((bl, tl, tr), ((0.0, -1 / hy), (-1 / hx, 1 / hy), (1 / hx, 0.0)))):
                a11 = np.mean([d11[v] for v in verts])
                a22 = np.mean([d22[v] for v in verts])
                a12 = np.mean([d12[v] for v in verts])
                A = D * np.array([[a11, a12], [a12, a22]])
                for a, ga in zip(verts, grads):
                    for b, gb in zip(verts, grads):
                        K[a, b] += dt * area * np.asarray(ga) @ A @ np.asarray(gb)
    return K


def test_tensor_stiffness_matches_gradient_assembly(oracle):
    """The oracle's quadrature of the anisotropic cell kernel against an
    assembly from explicit P1 gradients with the exact triangle mean of the
    tensor; a spatially varying tensor with D12 != 0 (pins the mixed term,
    the coefficient order and the hypotenuse coupling)."""
    nx, ny, W, H, dt, D, gam = 6, 4, 3.0, 2.0, 0.1, 7.0, 0.02
    rng = np.random.default_rng(3)
    n = (nx + 1) * (ny + 1)
    d11 = 1.0 + rng.random(n)
    d22 = 1.0 + rng.random(n)
    d12 = 0.4 * (rng.random(n) - 0.5)
    p = oracle.problem(nx, ny, W, H, dt, D, gam)
    rp, col, val = oracle.assemble(p, d11, d22, d12)
    A = oracle.to_scipy(rp, col, val, n).toarray()
    # the mass part from the isotropic assembly with D = 0
    p0 = oracle.problem(nx, ny, W, H, dt, 0.0, gam)
    rp0, col0, val0 = oracle.assemble(p0)
    M = oracle.to_scipy(rp0, col0, val0, n).toarray()
    K = _p1_tensor_assembly(nx, ny, W, H, dt, D, d11, d22, d12)
    assert np.allclose(A - M, K, rtol=0, atol=1e-13 * np.abs(K).max())
    # a constant tensor diag(a, b) splits into a*Kx + b*Ky of the 5-point stencil
    K2 = _p1_tensor_assembly(nx, ny, W, H, dt, D, np.full(n, 2.0), np.full(n, 0.5), np.zeros(n))
    Kx = _p1_tensor_assembly(nx, ny, W, H, dt, D, np.ones(n), np.zeros(n), np.zeros(n))
    Ky = _p1_tensor_assembly(nx, ny, W, H, dt, D, np.zeros(n), np.ones(n), np.zeros(n))
    assert np.allclose(K2, 2.0 * Kx + 0.5 * Ky, atol=1e-13)
    assert abs(Kx[0, 1] + dt * D * (H / ny) / (2 * W / nx)) < 1e-14  # boundary edge: one triangle


def test_fd_scheme_closed_forms(oracle):
    """The diffusionPETSc scheme: with Neumann walls the discrete cosine
    modes cos(pi k j/nx) are eigenvectors of the ghost-reflection operator,
    so one step divides a mode by exactly 1 + 2F(1 - cos(pi k/nx)); the
    weighted sum (weights 1/2 on walls, 1/4 in corners) is conserved; a
    Dirichlet wall holds its value (left/right win at corners)."""
    nx, ny, W, H, dt, D = 24, 10, 2.4, 1.0, 0.1, 0.3
    p = oracle.problem(nx, ny, W, H, dt, D)
    n = (nx + 1) * (ny + 1)
    j = np.tile(np.arange(nx + 1), ny + 1)
    i = np.repeat(np.arange(ny + 1), nx + 1)
    Fx, Fy = D * dt / (W / nx) ** 2, D * dt / (H / ny) ** 2
    for k in (1, 3):
        u0 = np.cos(np.pi * k * j / nx)
        u1 = oracle.fd_step(p, u0)
        assert np.allclose(u1, u0 / (1.0 + 2.0 * Fx * (1.0 - np.cos(np.pi * k / nx))), atol=1e-14)
        v0 = np.cos(np.pi * k * i / ny)
        v1 = oracle.fd_step(p, v0)
        assert np.allclose(v1, v0 / (1.0 + 2.0 * Fy * (1.0 - np.cos(np.pi * k / ny))), atol=1e-14)
    w = np.where((j == 0) | (j == nx), 0.5, 1.0) * np.where((i == 0) | (i == ny), 0.5, 1.0)
    u0 = np.random.default_rng(2).random(n)
    u1 = oracle.fd_step(p, u0)
    assert abs(w @ u1 - w @ u0) < 1e-12 * (w @ u0)
    # Dirichlet top 0.5, left 2.0: the wall nodes take their values
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.5, 0, 0), oracle.LEFT: (oracle.DIRICHLET, 2.0, 0, 0)}
    p2 = oracle.problem(nx, ny, W, H, dt, D, walls=walls)
    u2 = oracle.fd_step(p2, u0)
    assert np.all(u2[(j == 0)] == 2.0)
    assert np.all(u2[(i == ny) & (j > 0)] == 0.5)


def test_channel_cn_closed_forms(oracle):
    """The flow-channel Crank-Nicolson step (AdvectionDiffusion.ufl): with v = 0
    and no Robin ends, nodal cosines are eigenvectors of M and K (boundary rows
    halved alike), so one step multiplies mode theta by (mu - k D lam/2) /
    (mu + k D lam/2), mu = hc (4 + 2 cos theta)/6, lam = (2 - 2 cos theta)/hc;
    constants are invariant for any v (C 1 = 0); the flux from the trap is
    dt*D*(u_inner - u_wall)/wellScaling per column (src/fHSL.cpp:54-96)."""
    nx, ny, W, H, dt, D = 40, 6, 8.0, 1.2, 0.1, 50.0
    p = oracle.problem(nx, ny, W, H, dt, D)
    hc = W / nx
    u_flat = np.ones((ny + 1) * (nx + 1))  # no flux: equal rows
    ch = oracle.channel(0.0, 0.0, 0.0, 0.0, 0.0, 3.0, 1)
    for kmode in (1, 4):
        th = np.pi * kmode / nx
        c0 = np.cos(th * np.arange(nx + 1))
        t1, b1 = oracle.channel_update(p, ch, ch, u_flat, c0, 2.0 * c0)
        mu, lam = hc * (4 + 2 * np.cos(th)) / 6, (2 - 2 * np.cos(th)) / hc
        g = (mu - dt * D * lam / 2) / (mu + dt * D * lam / 2)
        assert np.allclose(t1, g * c0, atol=1e-13)
        assert np.allclose(b1, 2.0 * g * c0, atol=1e-13)
    chv = oracle.channel(35.0, 0.0, 0.0, 0.0, 0.0, 3.0, 3)
    t1, _ = oracle.channel_update(p, chv, chv, u_flat, np.full(nx + 1, 0.7), np.zeros(nx + 1))
    assert np.allclose(t1, 0.7, atol=1e-14)
    # flux: a trap field whose top row is lower than the next row by d_j
    u = np.zeros((ny + 1, nx + 1))
    d = np.linspace(0.1, 0.5, nx + 1)
    u[ny - 1] = d
    ch1 = oracle.channel(0.0, 0.0, 0.0, 0.0, 0.0, 4.0, 1)
    # with D = 0 in the channel form there is no diffusion: the channel gains exactly the flux
    p0 = oracle.problem(nx, ny, W, H, dt, D)
    t1, b1 = oracle.channel_update(p0, ch1, ch1, u.ravel(), np.zeros(nx + 1), np.zeros(nx + 1))
    tot = (t1 * np.where((np.arange(nx + 1) == 0) | (np.arange(nx + 1) == nx), 0.5, 1.0)).sum()
    ref = (dt * D * d / 4.0 * np.where((np.arange(nx + 1) == 0) | (np.arange(nx + 1) == nx), 0.5, 1.0)).sum()
    assert abs(tot - ref) < 1e-12 * abs(ref)   # diffusion conserves the lumped total M 1
    assert np.all(b1 == 0.0)


def test_convection_matrix_closed_forms(oracle):
    """Advection (config 4): the Galerkin convection part has zero row sums
    (constants are not transported), is skew for a constant velocity away from
    the walls (int v dphi_j/dx phi_i = -int v dphi_i/dx phi_j + boundary terms),
    and its interior row, summed from the six triangles around the node, is
    v*hy/3 * (u_E - u_W) + v*hy/6 * ((u_NE - u_N) + (u_S - u_SW)), which
    gives v*hx*hy = int v du/dx phi for u = x."""
    nx, ny, W, H, dt, D = 8, 6, 4.0, 3.0, 0.1, 0.0
    hx, hy = W / nx, H / ny
    p = oracle.problem(nx, ny, W, H, dt, D)
    n = (nx + 1) * (ny + 1)
    rp0, c0, v0 = oracle.assemble(p)
    M = oracle.to_scipy(rp0, c0, v0, n).toarray()
    v = 2.5
    p.set_velocity(np.full(ny + 1, v))
    rp, c, val = oracle.assemble(p)
    C = (oracle.to_scipy(rp, c, val, n).toarray() - M) / dt
    assert np.allclose(C.sum(axis=1), 0.0, atol=1e-13)
    idx = lambda i, j: i * (nx + 1) + j
    interior = [idx(i, j) for i in range(1, ny) for j in range(1, nx)]
    S = C + C.T
    assert np.allclose(S[np.ix_(interior, interior)], 0.0, atol=1e-13)
    i, j = 3, 4
    u = np.random.default_rng(0).random(n)
    g = lambda a, b: u[idx(a, b)]
    expect = v * hy / 3 * (g(i, j + 1) - g(i, j - 1)) + v * hy / 6 * ((g(i + 1, j + 1) - g(i + 1, j)) + (g(i - 1, j) - g(i - 1, j - 1)))
    assert abs((C @ u)[idx(i, j)] - expect) < 1e-12
    x = np.tile(np.arange(nx + 1) * hx, ny + 1)
    assert abs((C @ x)[idx(i, j)] - v * hx * hy) < 1e-12


def test_boundary_well_closed_forms(oracle):
    """computeBoundaryWell (src/simulation.cpp:581-604) pinned by its closed
    forms: no flux gives geometric decay w0*(1-dt*k)^n; a constant flux F gives
    w_n = w* + (w0 - w*)(1-dt*k)^n with w* = (F/s)(1-dt*k)/(dt*k), because the
    flux is added before the decay (the other order gives a different fixed
    point, F/(s*dt*k)); a large negative flux clamps at zero and counts one
    underflow per clamped step."""
    O = oracle
    w, n = O.boundary_well(1.0, 1.0, 1.0, 0.5, 1.0)   # (1 + 1)(1 - 0.5)
    assert (w, n) == (1.0, 0)
    dt, k, s, F, w0 = 0.1, 0.05, 750.0, 3.0, 0.7
    a = 1.0 - dt * k
    w = w0
    for step in range(1, 400):
        w, n = O.boundary_well(w, 0.0, s, k, dt)
        assert n == 0
        assert math.isclose(w, w0 * a ** step, rel_tol=1e-12)
    ws = (F / s) * a / (dt * k)
    w = w0
    for step in range(1, 400):
        w, _ = O.boundary_well(w, F, s, k, dt)
        assert math.isclose(w, ws + (w0 - ws) * a ** step, rel_tol=1e-12)
    assert not math.isclose(ws, F / (s * dt * k), rel_tol=1e-6)
    w, n = O.boundary_well(0.2, -1e6, s, k, dt)
    assert (w, n) == (0.0, 1)
    w, n = O.boundary_well(0.2, -0.2 * s, s, k, dt)       # lands exactly on zero: no underflow
    assert (w, n) == (0.0, 0)


def test_channel_pure_advection_of_a_linear_profile(oracle):
    """Sign and size of the channel advection term (AdvectionDiffusion.ufl:62-64,
    v c' w not integrated by parts): with D = 0 and no Robin ends, a linear
    profile c = a + b x is transported exactly, c -> c - dt*v*b at every node
    (C c = b * int(phi) and M 1 = int(phi) row by row, so the Crank-Nicolson
    rows hold with c1 = c0 - k v b for each substep).  A flipped advection
    sign would give c0 + dt*v*b."""
    nx, ny, W, H, dt = 50, 4, 10.0, 1.0, 0.1
    p = oracle.problem(nx, ny, W, H, dt, 0.0)   # D = 0: no diffusion, no flux from the trap
    x = np.arange(nx + 1) * (W / nx)
    a, bslope = 0.4, 0.03
    c0 = a + bslope * x
    u = np.random.default_rng(1).random((ny + 1) * (nx + 1))
    for v, subs in ((20.0, 1), (-35.0, 4)):
        ch = oracle.channel(v, 0.0, 0.0, 0.0, 0.0, 2.0, subs)
        t1, b1 = oracle.channel_update(p, ch, ch, u, c0, 2.0 * c0)
        assert np.allclose(t1, c0 - dt * v * bslope, rtol=0, atol=5e-13), v
        assert np.allclose(b1, 2.0 * c0 - dt * v * 2.0 * bslope, rtol=0, atol=5e-13), v
        assert np.abs(t1 - (c0 + dt * v * bslope)).max() > 0.05   # the other sign is far off


def test_fd_robin_wall_holds_its_linear_steady_state(oracle):
    """diffusionPETSc's Robin wall by ghost reflection (ApplyBoundaryConditions,
    diffuclass.cpp:191-339): for D du/dn = -r (u - s) on the left wall and a
    Dirichlet right wall g, u = A + B x with B = r (A - s)/D and A + B W = g is
    a steady state of the discrete scheme (its 5-point Laplacian is zero and
    the ghost u_-1 = u_1 - 2 h r (u_0 - s)/D lies on the same line), so one
    backward-Euler step returns it unchanged.  A wrong factor in the Robin
    row (the 2 h, or r/D) moves the wall column."""
    nx, ny, W, H, dt, D = 30, 12, 3.0, 1.2, 0.1, 0.7
    r, s, g = 1.3, 0.25, 2.0
    walls = {oracle.LEFT: (oracle.ROBIN, 0.0, r, s), oracle.RIGHT: (oracle.DIRICHLET, g, 0, 0)}
    p = oracle.problem(nx, ny, W, H, dt, D, walls=walls)
    A = (g + r * s * W / D) / (1.0 + r * W / D)
    B = r * (A - s) / D
    x = np.tile(np.arange(nx + 1) * (W / nx), ny + 1)
    u0 = A + B * x
    u1 = oracle.fd_step(p, u0)
    assert np.allclose(u1, u0, rtol=0, atol=1e-12)
    # the same line with a Robin rate off by a factor 2 is not steady
    B2 = 2.0 * r * (A - s) / D
    u2 = A + B2 * x
    assert np.abs(oracle.fd_step(p, u2) - u2).max() > 1e-3


def test_pole_window_is_asymmetric(oracle):
    """The search window runs between the poles of Ecoli::updatePoleCenters
    (src/abm/Ecoli.cpp:44-56): centre + (L/2 - w/2) along the axis and
    centre - (L/2 + w/2) against it, then +/- nodesToEdge = round(npm/2)
    (eQabm.cpp:75, 274-290).  Hand-computed for a 4 um x 2 um rod at (5, 5),
    npm 4: angle 0 gives columns [round(2*4) - 2, round(6*4) + 2] = [6, 26]
    (a symmetric window would be [10, 30]); angle pi mirrors it to [14, 34];
    rows are round(5*4) -/+ 2.  The box |dx| < 1, |dy| < 1 (strict) holds
    columns 17..23 and rows 17..23, and the window keeps rows 18..22: 35 nodes.
    Near the left wall the far pole is clamped to x = 0."""
    O = oracle
    g = O.grid(41, 41, 4.0, 10.0, 10.0)
    assert O.cell_window(g, [5.0, 5.0, 0.0, 4.0, 2.0]) == (18, 22, 6, 26)
    assert O.cell_window(g, [5.0, 5.0, math.pi, 4.0, 2.0]) == (18, 22, 14, 34)
    cov = set(map(tuple, O.cell_coverage(g, [5.0, 5.0, 0.0, 4.0, 2.0]).tolist()))
    assert cov == {(i, j) for i in range(18, 23) for j in range(17, 24)}
    # angle pi/2: the poles lie along y; rows [round(2*4) - 2, round(6*4) + 2]
    assert O.cell_window(g, [5.0, 5.0, math.pi / 2, 4.0, 2.0]) == (6, 26, 18, 22)
    # centre 1.5 um from the left wall: the pole at x = 1.5 - 3 = -1.5 is clamped to 0
    assert O.cell_window(g, [1.5, 5.0, 0.0, 4.0, 2.0]) == (18, 22, 0, 12)


def test_cell_tensor_hand_counted(oracle):
    """setDiffusionTensor (src/abm/eQabm.cpp:241-248, 306-325): isotropic
    (1, 1, 0) off the cells; on the 21 nodes of a 3 x 1 um rod at angle 0,
    (Dx, Dy, 0); turned by pi/2, (Dy, Dx, ~0); at pi/4 all three entries from
    cos^2 = sin^2 = sin cos = 1/2 (to rounding); where two rods overlap, the
    later one in list order wins; a rod too short for any node writes its
    centre node (the fallback of findInteriorPoints)."""
    O = oracle
    g = O.grid(41, 21, 4.0, 10.0, 5.0)
    Dx, Dy = 3.0, 0.5
    rod = {(i, j) for i in (9, 10, 11) for j in range(17, 24)}
    D11, D22, D12 = O.cell_tensor(g, np.array([[5.0, 2.5, 0.0, 3.0, 1.0]]), Dx, Dy)
    on = np.zeros(41 * 21, bool)
    for (i, j) in rod:
        on[j + 41 * i] = True
    assert np.all(D11[on] == Dx) and np.all(D22[on] == Dy) and np.all(D12[on] == 0.0)
    assert np.all(D11[~on] == 1.0) and np.all(D22[~on] == 1.0) and np.all(D12[~on] == 0.0)
    D11, D22, D12 = O.cell_tensor(g, np.array([[5.0, 2.5, math.pi / 2, 3.0, 1.0]]), Dx, Dy)
    rod2 = [(i, j) for i in range(7, 14) for j in (19, 20, 21)]
    for (i, j) in rod2:
        v = j + 41 * i
        assert math.isclose(D11[v], Dy, rel_tol=1e-15) and math.isclose(D22[v], Dx, rel_tol=1e-15)
        assert abs(D12[v]) < 1e-15
    D11, D22, D12 = O.cell_tensor(g, np.array([[5.0, 2.5, math.pi / 4, 3.0, 1.0]]), Dx, Dy)
    cov = O.cell_coverage(g, [5.0, 2.5, math.pi / 4, 3.0, 1.0])
    for (i, j) in cov:
        v = j + 41 * i
        assert math.isclose(D11[v], (Dx + Dy) / 2, rel_tol=1e-15) and math.isclose(D22[v], (Dx + Dy) / 2, rel_tol=1e-15)
        assert math.isclose(D12[v], (Dx - Dy) / 2, rel_tol=1e-15)
    # overlap: the second rod (angle pi/2 through the same centre) wins on the shared nodes
    cells = np.array([[5.0, 2.5, 0.0, 3.0, 1.0], [5.0, 2.5, math.pi / 2, 3.0, 1.0]])
    D11, _, _ = O.cell_tensor(g, cells, Dx, Dy)
    shared = rod & set(rod2)
    assert shared and all(math.isclose(D11[j + 41 * i], Dy, rel_tol=1e-15) for (i, j) in shared)
    assert all(D11[j + 41 * i] == Dx for (i, j) in rod - set(rod2))
    D11, _, _ = O.cell_tensor(g, cells[::-1].copy(), Dx, Dy)
    assert all(D11[j + 41 * i] == Dx for (i, j) in rod)
    # fallback: a zero-length rod writes only its centre node
    D11, D22, D12 = O.cell_tensor(g, np.array([[3.3, 1.2, 0.0, 1.0, 1.0]]), Dx, Dy)
    assert np.count_nonzero(D11 != 1.0) == 1 and D11[round(3.3 * 4) + 41 * round(1.2 * 4)] == Dx


def test_couple_affine_semantics(oracle):
    """The read -> compute -> write order of updateCells for a strain whose
    amount depends on its own reading (src/abm/eQabm.cpp:384-405): with b = 0
    it is the plain coupling bit for bit; one cell on a constant field c
    deposits (a + b c)/|S_c| per node; a second cell overlapping the first
    reads the first's deposit before its own amount is formed."""
    O = oracle
    g = O.grid(41, 21, 4.0, 10.0, 5.0)
    rng = np.random.default_rng(3)
    cells = np.array([[5.0, 2.5, 0.0, 3.0, 1.0], [5.5, 2.5, 0.0, 3.0, 1.0], [3.0, 3.0, 0.7, 2.5, 1.0]])
    u = [rng.random(41 * 21), rng.random(41 * 21)]
    a = rng.random((3, 2))
    f1, s1, c1 = O.couple(g, cells, u, a)
    f2, s2, c2 = O.couple_affine(g, cells, u, a, np.zeros((3, 2)))
    assert all(np.array_equal(x, y) for x, y in zip(f1, f2)) and np.array_equal(s1, s2) and np.array_equal(c1, c2)
    # one cell on a constant field 0.4: q = 0.3 - 0.5 * 0.4 = 0.1 over its 21 nodes
    (f,), s, c = O.couple_affine(g, cells[:1], [np.full(41 * 21, 0.4)], np.array([[0.3]]), np.array([[-0.5]]))
    assert c[0] == 21 and math.isclose(s[0, 0], 0.4, rel_tol=1e-15)
    assert math.isclose(f.max(), 0.4 + (0.3 - 0.5 * s[0, 0]) / 21, rel_tol=1e-15) and np.count_nonzero(f != 0.4) == 21
    assert math.isclose(f.max(), 0.4 + 0.1 / 21, rel_tol=1e-14)
    # two overlapping cells (15 shared nodes) on a zero field, q = a + b e
    (f,), s, c = O.couple_affine(g, cells[:2], [np.zeros(41 * 21)], np.array([[2.1], [4.2]]), np.array([[-1.0], [-3.0]]))
    e2 = 15 * (2.1 / 21) / 21
    assert s[0, 0] == 0.0 and math.isclose(s[1, 0], e2, rel_tol=1e-14)
    assert math.isclose(f.sum(), 2.1 + (4.2 - 3.0 * e2), rel_tol=1e-13)
