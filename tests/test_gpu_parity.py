"""GPU parity of the HIP field step against the CPU oracle and the textbook
golden fixtures.  Every call goes through the C ABI (libeq_field.so).

Tolerances:
  operator and right-hand side vs the oracle's assembled CSR: 1e-14 relative
    to the row scale (sum |A_ij x_j|);
  solves vs the oracle's direct LU and vs the fixtures: 1e-9 relative L2 per
    species per step (the solver stops at ||b-Ax||/||b|| <= 1e-10);
  coverage counts: exact; deposited fields: bit-identical (same per-node
    addition order as the sequential loop); sampled means: n*eps relative
    (the per-cell sum runs in another order), 1e-13 at ~30 nodes per cell,
    2e-13 at ~800.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "p1_tiny.npz")
SOLVERS = ["jacobi", "gmg", "mixed", "ir"]


def _kind(field, name):
    return {"jacobi": field.JACOBI_CG, "gmg": field.GMG_CG, "mixed": field.MIXED_GMG_CG,
            "ir": field.IR_GMG_CG}[name]


def _walls_from(arr):
    return {w: (int(arr[w][0]), arr[w][1], arr[w][2], arr[w][3]) for w in range(4) if arr[w][0] != 0}


def _pair(eqf, oracle, nx, ny, W, H, dt, D, gamma=0.0, f=0.0, walls=None):
    return (eqf.field_desc(nx, ny, W, H, dt, D, gamma, f, walls),
            oracle.problem(nx, ny, W, H, dt, D, gamma, f, walls=walls))


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _cases():
    z = np.load(GOLD)
    return [str(n) for n in z["names"]]


def _set_fixture_tensor(fb, z, name):
    """Fixtures with a tensor field other than (1, 1, 0) get it uploaded."""
    t = z[f"{name}/tensor"]
    if np.any(t[:, 0] != 1.0) or np.any(t[:, 1] != 1.0) or np.any(t[:, 2] != 0.0):
        fb.set_tensor(0, t[:, 0], t[:, 1], t[:, 2])


@pytest.mark.parametrize("name", _cases())
def test_operator_entries_match_fixture(eqf, oracle, name):
    """Columns of the matrix-free operator = the textbook-assembled matrix."""
    z = np.load(GOLD)
    nx, ny, W, H, dt, D, gamma, f = z[f"{name}/params"]
    d = eqf.field_desc(int(nx), int(ny), W, H, dt, D, gamma, f, _walls_from(z[f"{name}/walls"]))
    A_ref = z[f"{name}/A_bc"]
    n = A_ref.shape[0]
    with eqf.FieldBatch([d], solver=eqf.solver_desc(eqf.JACOBI_CG)) as fb:
        _set_fixture_tensor(fb, z, name)
        A = np.stack([fb.apply(0, np.eye(n)[k]) for k in range(n)], axis=1)
        assert np.abs(A - A_ref).max() <= 1e-14 * np.abs(A_ref).max()
        b = fb.rhs(0, z[f"{name}/u0"])
        assert np.abs(b - z[f"{name}/b_bc"]).max() <= 1e-14 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("shape", [(1, 1), (3, 2), (64, 4), (130, 67), (160, 80)])
def test_operator_matches_oracle_csr(eqf, oracle, shape):
    nx, ny = shape
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.3, 0, 0), oracle.LEFT: (oracle.ROBIN, 0, 1.7, 0.4)}
    d, p = _pair(eqf, oracle, nx, ny, nx * 0.25, ny * 0.3, 0.1, 400.0, 0.01, 0.2, walls)
    rp, c, v = oracle.assemble(p)
    x = np.random.default_rng(nx * 1000 + ny).random(oracle.num_nodes(p))
    u0 = np.random.default_rng(ny).random(oracle.num_nodes(p))
    b_ref = oracle.assemble_rhs(p, u0)
    v2, b2 = oracle.apply_dirichlet(p, rp, c, v, b_ref)
    y_ref = oracle.spmv(rp, c, v2, x)
    scale = oracle.spmv(rp, c, np.abs(v2), np.abs(x))
    with eqf.FieldBatch([d], solver=eqf.solver_desc(eqf.JACOBI_CG)) as fb:
        y = fb.apply(0, x)
        assert np.all(np.abs(y - y_ref) <= 1e-14 * scale + 1e-300)
        b = fb.rhs(0, u0)
        bscale = oracle.assemble_rhs(p, np.abs(u0)) + 1.0
        assert np.all(np.abs(b - b2) <= 1e-14 * bscale)


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("name", _cases())
def test_solve_matches_fixture(eqf, oracle, solver, name):
    z = np.load(GOLD)
    nx, ny, W, H, dt, D, gamma, f = z[f"{name}/params"]
    d = eqf.field_desc(int(nx), int(ny), W, H, dt, D, gamma, f, _walls_from(z[f"{name}/walls"]))
    with eqf.FieldBatch([d], solver=eqf.solver_desc(_kind(eqf, solver))) as fb:
        _set_fixture_tensor(fb, z, name)
        fb.upload(0, z[f"{name}/u0"])
        st = fb.step()
        x = fb.download(0)
    assert st.relres_max <= 1e-10
    assert _rel(x, z[f"{name}/x"]) <= 1e-9


@pytest.mark.parametrize("solver", SOLVERS)
def test_config0_steps_match_oracle_lu(eqf, oracle, solver):
    """The reference CPU configuration: 160x80 trap, 2 species, 200 rods;
    couple + solve for 4 steps, each step against the oracle's direct LU."""
    from eq_amd import workloads as Wl
    t = Wl.config0()
    walls = Wl.TOP_DIRICHLET
    descs = [eqf.field_desc(t.nx, t.ny, t.width, t.height, 0.1, D, g, 0.0, walls) for D, g in zip(t.D, t.gamma)]
    probs = [oracle.problem(t.nx, t.ny, t.width, t.height, 0.1, D, g, 0.0, walls=walls) for D, g in zip(t.D, t.gamma)]
    grid = oracle.grid(t.nx + 1, t.ny + 1, t.npm, t.width, t.height)
    ref = [u.copy() for u in t.u0]
    q = t.rates * 0.1
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(_kind(eqf, solver))) as fb:
        for s in range(2):
            fb.upload(s, ref[s])
        for step in range(4):
            sampled, counts = fb.couple(0, t.cells, t.npm, q)
            ref, s_ref, c_ref = oracle.couple(grid, t.cells, ref, q)
            assert np.array_equal(counts, c_ref)
            assert np.allclose(sampled, s_ref, rtol=1e-13, atol=1e-300)
            st = fb.step()
            assert st.relres_max <= 1e-10
            for s in range(2):
                ref[s], info = oracle.step(probs[s], ref[s], "lu")
                got = fb.download(s)
                assert _rel(got, ref[s]) <= 1e-9, (step, s, _rel(got, ref[s]))
                fb.upload(s, ref[s])  # continue both from the same state


@pytest.mark.parametrize("solver", SOLVERS)
def test_graph_replays_match_eager_launches(eqf, oracle, solver):
    """The captured iteration graph is replayed many times per step; with
    host work (downloads, uploads, coupling) between the replays every step
    must be bit-identical to eager launches."""
    from eq_amd import workloads as Wl
    t = Wl.config0()
    descs = [eqf.field_desc(t.nx, t.ny, t.width, t.height, 0.1, D, g, 0.0, Wl.TOP_DIRICHLET)
             for D, g in zip(t.D, t.gamma)]
    q = t.rates * 0.1
    out = {}
    for graphs in (0, 1):
        with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(_kind(eqf, solver), use_graphs=graphs)) as fb:
            for s in range(2):
                fb.upload(s, t.u0[s])
            hist = []
            for step in range(5):
                fb.couple(0, t.cells, t.npm, q)
                fb.step()
                hist.append([fb.download(s) for s in range(2)])
                fb.upload(1, hist[-1][1])  # eager host work between replays
            out[graphs] = hist
    for step in range(5):
        for s in range(2):
            assert np.array_equal(out[0][step][s], out[1][step][s]), (step, s)


def test_couple_bit_exact_against_sequential_loop(eqf, oracle):
    from eq_amd import workloads as Wl
    rng = Wl.SplitMix64(11)
    nx, ny, npm, W, H = 400, 200, 20.0, 20.0, 10.0
    cells, rates = Wl.seeded_cells(rng, 700, W, H, 2)
    cells[::7, 4] = 1.6            # some wide rods (window clipping)
    cells[::11, 3] = cells[::11, 4]  # some zero-length boxes (centre fallback)
    u = [np.random.default_rng(1).random((nx + 1) * (ny + 1)), np.zeros((nx + 1) * (ny + 1))]
    d = [eqf.field_desc(nx, ny, W, H, 0.1, 100.0) for _ in range(2)]
    grid = oracle.grid(nx + 1, ny + 1, npm, W, H)
    with eqf.FieldBatch(d, nspecies=2, solver=eqf.solver_desc(eqf.JACOBI_CG)) as fb:
        for s in range(2):
            fb.upload(s, u[s])
        sampled, counts = fb.couple(0, cells, npm, rates)
        got = [fb.download(s) for s in range(2)]
    ref, s_ref, c_ref = oracle.couple(grid, cells, u, rates)
    assert np.array_equal(counts, c_ref)
    for s in range(2):
        assert np.array_equal(got[s], ref[s])
    # sums of up to ~1000 terms in another order: |err| <= n*eps*sum|x|
    assert np.allclose(sampled, s_ref, rtol=2e-13, atol=1e-300)


def test_uneven_batch_matches_oracle(eqf, oracle):
    shapes = [(37, 23, 3.7, 2.3), (64, 33, 6.4, 3.3), (101, 50, 10.1, 5.0)]
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0), oracle.RIGHT: (oracle.ROBIN, 0, 3.0, 0.1)}
    descs, probs, u0 = [], [], []
    rng = np.random.default_rng(9)
    for (nx, ny, W, H) in shapes:
        for D, g in ((300.0, 0.01), (80.0, 0.05)):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, g, 0.0, walls))
            probs.append(oracle.problem(nx, ny, W, H, 0.1, D, g, 0.0, walls=walls))
            u0.append(rng.random((nx + 1) * (ny + 1)))
    for solver in SOLVERS:
        with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(_kind(eqf, solver))) as fb:
            for k, u in enumerate(u0):
                fb.upload(k, u)
            st = fb.step()
            assert st.nconverged == len(descs)
            for k in range(len(descs)):
                ref, _ = oracle.step(probs[k], u0[k], "lu")
                assert _rel(fb.download(k), ref) <= 1e-9, (solver, k)


def test_boundary_flux_and_vertex_map(eqf, oracle):
    d, p = _pair(eqf, oracle, 30, 20, 3.0, 2.0, 0.1, 50.0)
    u = np.random.default_rng(4).random(oracle.num_nodes(p))
    with eqf.FieldBatch([d], solver=eqf.solver_desc(eqf.JACOBI_CG)) as fb:
        fb.upload(0, u)
        flux = fb.boundary_flux(0)
        xy, dof = fb.vertex_map(0)
    assert math.isclose(flux, 0.1 * 50.0 * oracle.boundary_flux(p, u), rel_tol=1e-12)
    assert np.array_equal(xy, oracle.mesh_coords(p))
    assert np.array_equal(dof, np.arange(oracle.num_nodes(p)))


def test_full_size_trap_residual_and_solver_agreement(eqf, oracle):
    """2000x1000 trap (config 1 size): the relative residual of every solver,
    recomputed in FP64 with the oracle's assembled CSR, is <= 1e-10, and the
    three solvers agree to 1e-9."""
    from eq_amd import workloads as Wl
    t = Wl.config1(u0_seed=2)
    walls = Wl.TOP_DIRICHLET
    p = oracle.problem(t.nx, t.ny, t.width, t.height, 0.1, t.D[0], t.gamma[0], 0.0, walls=walls)
    rp, c, v = oracle.assemble(p)
    b = oracle.assemble_rhs(p, t.u0[0])
    v2, b2 = oracle.apply_dirichlet(p, rp, c, v, b)
    sols = {}
    for solver in SOLVERS:
        d = eqf.field_desc(t.nx, t.ny, t.width, t.height, 0.1, t.D[0], t.gamma[0], 0.0, walls)
        with eqf.FieldBatch([d], solver=eqf.solver_desc(_kind(eqf, solver), max_iter=20000)) as fb:
            fb.upload(0, t.u0[0])
            st = fb.step()
            x = fb.download(0)
        res = np.linalg.norm(b2 - oracle.spmv(rp, c, v2, x)) / np.linalg.norm(b2)
        print(f"{solver}: iterations={st.iterations_max} relres(gpu)={st.relres_max:.3e} relres(csr)={res:.3e}",
              flush=True)
        assert st.relres_max <= 1e-10
        assert res <= 1.2e-10  # CSR evaluation adds its own rounding floor
        sols[solver] = x
    assert _rel(sols["gmg"], sols["jacobi"]) <= 1e-9
    assert _rel(sols["mixed"], sols["jacobi"]) <= 1e-9
    assert _rel(sols["ir"], sols["jacobi"]) <= 1e-9


def test_degenerate_inputs(eqf, oracle):
    # 1x1 trap, all walls Dirichlet with distinct values; zero field; pure Neumann
    walls = {w: (oracle.DIRICHLET, 0.1 * (w + 1), 0, 0) for w in range(4)}
    d, p = _pair(eqf, oracle, 1, 1, 1.0, 1.0, 0.1, 1.0, 0.0, 0.0, walls)
    for solver in SOLVERS:
        with eqf.FieldBatch([d], solver=eqf.solver_desc(_kind(eqf, solver))) as fb:
            fb.upload(0, np.zeros(4))
            fb.step()
            ref, _ = oracle.step(p, np.zeros(4), "lu")
            assert np.array_equal(fb.download(0), ref)
    d, p = _pair(eqf, oracle, 17, 9, 1.7, 0.9, 0.1, 10.0)
    with eqf.FieldBatch([d], solver=eqf.solver_desc(eqf.MIXED_GMG_CG)) as fb:
        fb.upload(0, np.zeros(180))
        st = fb.step()
        assert np.array_equal(fb.download(0), np.zeros(180)) and st.nconverged == 1
        fb.upload(0, np.full(180, 0.25))
        fb.step()
        assert np.abs(fb.download(0) - 0.25).max() <= 1e-12


def test_couple_traps_batched_matches_per_trap_and_oracle(eqf, oracle):
    """Several uneven traps coupled in one pass: fields bit-identical to the
    sequential loop per trap, samples to n*eps, and identical to trap-by-trap
    calls."""
    import torch
    from eq_amd import workloads as Wl
    shapes = [(120, 60, 6.0, 3.0, 150), (64, 200, 3.2, 10.0, 400), (300, 90, 15.0, 4.5, 60)]
    npm = 20.0
    rng = Wl.SplitMix64(21)
    descs, cells, rates, u0, grids = [], [], [], [], []
    for (nx, ny, W, H, nc) in shapes:
        c, r = Wl.seeded_cells(rng, nc, W, H, 2)
        cells.append(c)
        rates.append(r)
        grids.append(oracle.grid(nx + 1, ny + 1, npm, W, H))
        g = np.random.default_rng(nx)
        for s in range(2):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, 100.0))
            u0.append(g.random((nx + 1) * (ny + 1)))
    off = np.concatenate([[0], np.cumsum([c.shape[0] for c in cells])]).astype(np.int64)
    dev = torch.device("cuda")
    cells_d = torch.from_numpy(np.concatenate(cells)).to(dev).contiguous()
    q_d = torch.from_numpy(np.concatenate(rates)).to(dev).contiguous()
    samp_d = torch.zeros((int(off[-1]), 2), dtype=torch.float64, device=dev)
    cnt_d = torch.zeros(int(off[-1]), dtype=torch.int32, device=dev)
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(eqf.JACOBI_CG)) as fb:
        for k, u in enumerate(u0):
            fb.upload(k, u)
        fb.couple_traps_dev(0, off, cells_d.data_ptr(), npm, q_d.data_ptr(), samp_d.data_ptr(), cnt_d.data_ptr())
        fb.sync()
        got = [fb.download(k) for k in range(len(descs))]
    sampled, counts = samp_d.cpu().numpy(), cnt_d.cpu().numpy()
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(eqf.JACOBI_CG)) as fb:
        for k, u in enumerate(u0):
            fb.upload(k, u)
        one = [fb.couple(t, cells[t], npm, rates[t]) for t in range(len(shapes))]
        got1 = [fb.download(k) for k in range(len(descs))]
    for t in range(len(shapes)):
        ref, s_ref, c_ref = oracle.couple(grids[t], cells[t], [u0[2 * t], u0[2 * t + 1]], rates[t])
        a, b = off[t], off[t + 1]
        assert np.array_equal(counts[a:b], c_ref)
        assert np.allclose(sampled[a:b], s_ref, rtol=2e-13, atol=1e-300)
        assert np.array_equal(sampled[a:b], one[t][0])
        for s in range(2):
            assert np.array_equal(got[2 * t + s], ref[s]), (t, s)
            assert np.array_equal(got1[2 * t + s], ref[s]), (t, s)


def test_couple_past_the_kept_pair_capacity(eqf, oracle):
    """The tile coupling sizes its pair buffers from earlier calls: a call with
    far more pairs than those (run by the list-order kernel instead) and the
    call after it (buffers grown) both match the sequential loop, fields bit
    for bit."""
    from eq_amd import workloads as Wl
    nx, ny, npm, W, H = 400, 200, 20.0, 20.0, 10.0
    rng = Wl.SplitMix64(31)
    grid = oracle.grid(nx + 1, ny + 1, npm, W, H)
    u = [np.random.default_rng(2).random((nx + 1) * (ny + 1)), np.random.default_rng(3).random((nx + 1) * (ny + 1))]
    d = [eqf.field_desc(nx, ny, W, H, 0.1, 100.0) for _ in range(2)]
    with eqf.FieldBatch(d, nspecies=2, solver=eqf.solver_desc(eqf.JACOBI_CG)) as fb:
        for s in range(2):
            fb.upload(s, u[s])
        for nc in (4, 3000, 2500):  # a few pairs, then ~24 000 (past the kept room), then room for them
            cells, rates = Wl.seeded_cells(rng, nc, W, H, 2)
            sampled, counts = fb.couple(0, cells, npm, rates)
            got = [fb.download(s) for s in range(2)]
            ref, s_ref, c_ref = oracle.couple(grid, cells, u, rates)
            assert np.array_equal(counts, c_ref), nc
            assert np.allclose(sampled, s_ref, rtol=2e-13, atol=1e-300), nc
            for s in range(2):
                assert np.array_equal(got[s], ref[s]), (nc, s)
            u = ref


def test_autotune_restores_state_and_keeps_parity(eqf, oracle):
    """The autotuner's trial steps leave no trace: fields unchanged, and the
    next step (with the chosen cycle shape) still matches the oracle."""
    shapes = [(90, 45, 9.0, 4.5), (40, 70, 4.0, 7.0)]
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0)}
    descs, probs, u0 = [], [], []
    rng = np.random.default_rng(5)
    for (nx, ny, W, H) in shapes:
        for D, g in ((300.0, 0.01), (80.0, 0.05)):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, g, 0.0, walls))
            probs.append(oracle.problem(nx, ny, W, H, 0.1, D, g, 0.0, walls=walls))
            u0.append(rng.random((nx + 1) * (ny + 1)))
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(eqf.MIXED_GMG_CG)) as fb:
        for k, u in enumerate(u0):
            fb.upload(k, u)
        sd = fb.autotune(1)
        assert (sd.mg_pre, sd.mg_post) in ((1, 1), (2, 1), (2, 2), (3, 3))
        for k, u in enumerate(u0):
            assert np.array_equal(fb.download(k), u)
        st = fb.step()
        assert st.nconverged == len(descs)
        for k in range(len(descs)):
            ref, _ = oracle.step(probs[k], u0[k], "lu")
            assert _rel(fb.download(k), ref) <= 1e-9, k


@pytest.mark.parametrize("solver", SOLVERS)
def test_tensor_field_steps_match_oracle(eqf, oracle, solver):
    """A spatially varying SPD tensor (D12 != 0) on uneven traps with Dirichlet
    and Robin walls: every solver against the oracle's LU with the same
    tensor, over two steps; then the tensor is cleared and the constant-
    coefficient step must match again."""
    shapes = [(70, 45, 7.0, 4.5), (33, 60, 3.3, 6.0)]
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0), oracle.RIGHT: (oracle.ROBIN, 0, 2.0, 0.3)}
    descs, probs, u0, tens = [], [], [], []
    rng = np.random.default_rng(17)
    for (nx, ny, W, H) in shapes:
        n = (nx + 1) * (ny + 1)
        for D, g in ((300.0, 0.01), (80.0, 0.05)):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, g, 0.0, walls))
            probs.append(oracle.problem(nx, ny, W, H, 0.1, D, g, 0.0, walls=walls))
            u0.append(rng.random(n))
            d11 = 0.5 + rng.random(n)
            d22 = 0.5 + rng.random(n)
            d12 = 0.8 * np.sqrt(d11 * d22) * (rng.random(n) - 0.5)
            tens.append((d11, d22, d12))
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(_kind(eqf, solver))) as fb:
        for k in range(len(descs)):
            fb.upload(k, u0[k])
            fb.set_tensor(k, *tens[k])
        ref = [u.copy() for u in u0]
        for step in range(2):
            st = fb.step()
            assert st.nconverged == len(descs), step
            for k in range(len(descs)):
                ref[k], _ = oracle.step(probs[k], ref[k], "lu", D11=tens[k][0], D22=tens[k][1], D12=tens[k][2])
                assert _rel(fb.download(k), ref[k]) <= 1e-9, (step, k)
        for k in range(len(descs)):
            fb.set_tensor(k)
        st = fb.step()
        assert st.nconverged == len(descs)
        for k in range(len(descs)):
            r2, _ = oracle.step(probs[k], ref[k], "lu")
            assert _rel(fb.download(k), r2) <= 1e-9, k


@pytest.mark.parametrize("solver", SOLVERS)
def test_fd_scheme_matches_oracle(eqf, oracle, solver):
    """diffusionPETSc's 5-point scheme (ghost-reflection walls, Robin, Dirichlet
    corners) on uneven traps in one batch, three steps against the oracle's
    exact solve of the reference rows."""
    cases = [
        ((48, 30, 4.8, 3.0), {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0)}),
        ((65, 40, 6.5, 2.0), {oracle.LEFT: (oracle.DIRICHLET, 1.0, 0, 0), oracle.TOP: (oracle.DIRICHLET, 0.5, 0, 0),
                              oracle.RIGHT: (oracle.ROBIN, 0, 2.0, 0.3)}),
        ((20, 90, 2.0, 9.0), {}),  # all Neumann
    ]
    descs, probs, u0 = [], [], []
    rng = np.random.default_rng(23)
    for (nx, ny, W, H), walls in cases:
        for D in (300.0, 60.0):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, 0.0, 0.0, walls, scheme=eqf.SCHEME_FD))
            probs.append(oracle.problem(nx, ny, W, H, 0.1, D, 0.0, 0.0, walls=walls))
            u0.append(rng.random((nx + 1) * (ny + 1)))
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(_kind(eqf, solver))) as fb:
        ref = [u.copy() for u in u0]
        for k in range(len(descs)):
            fb.upload(k, u0[k])
        for step in range(3):
            st = fb.step()
            assert st.nconverged == len(descs), step
            for k in range(len(descs)):
                ref[k] = oracle.fd_step(probs[k], ref[k])
                assert _rel(fb.download(k), ref[k]) <= 1e-9, (step, k)


def _channel_rates(v, D, length_left, length_right):
    """Robin rates of the channel ends (fenicsInterface::setRobinBoundaryConditions)."""
    if v > 1e-6:
        return v / (1.0 - np.exp(-length_left * v / D)), v / (np.exp(length_right * v / D) - 1.0)
    return D / length_left, D / length_right


@pytest.mark.parametrize("solver", SOLVERS)
def test_flow_channels_match_oracle(eqf, oracle, solver):
    """Traps whose top and bottom walls are flow channels: each step solves the
    trap with the channels' values as Dirichlet data, then moves the flux into
    the channels and advances them by Crank-Nicolson substeps; three steps of
    fields and channels against the oracle."""
    shapes = [(60, 30, 6.0, 3.0), (41, 50, 4.1, 5.0)]
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0), oracle.BOTTOM: (oracle.DIRICHLET, 0.0, 0, 0),
             oracle.LEFT: (oracle.ROBIN, 0, 1.5, 0.2)}
    descs, probs, u0, chans = [], [], [], []
    rng = np.random.default_rng(31)
    for (nx, ny, W, H) in shapes:
        for D in (300.0, 80.0):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls))
            probs.append(oracle.problem(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls=walls))
            u0.append(rng.random((nx + 1) * (ny + 1)))
            pair = []
            for v, subs in ((20.0, 3), (0.0, 1)):
                rl, rr = _channel_rates(v, D, 5.0, 7.0)
                pair.append((v, rl, 0.0, rr, 0.0, 2.5, subs))
            chans.append(pair)
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(_kind(eqf, solver))) as fb:
        for k in range(len(descs)):
            fb.upload(k, u0[k])
            fb.set_channels(k, eqf.channel_desc(*chans[k][0]), eqf.channel_desc(*chans[k][1]))
        ref = [u.copy() for u in u0]
        tops = [np.zeros(p.nx + 1) for p in probs]
        bots = [np.zeros(p.nx + 1) for p in probs]
        for step in range(3):
            st = fb.step()
            assert st.nconverged == len(descs), step
            for k, p in enumerate(probs):
                p.set_wall_values(oracle.TOP, tops[k])
                p.set_wall_values(oracle.BOTTOM, bots[k])
                ref[k], _ = oracle.step(p, ref[k], "lu")
                tops[k], bots[k] = oracle.channel_update(p, oracle.channel(*chans[k][0]), oracle.channel(*chans[k][1]),
                                                         ref[k], tops[k], bots[k])
                assert _rel(fb.download(k), ref[k]) <= 1e-9, (step, k)
                t, b = fb.channel_values(k)
                assert _rel(t, tops[k]) <= 1e-9 and _rel(b, bots[k]) <= 1e-9, (step, k)
    # per-node wall values on their own
    d, p = _pair(eqf, oracle, 30, 20, 3.0, 2.0, 0.1, 50.0, 0.0, 0.0, {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0)})
    g = np.linspace(0.0, 1.0, 31)
    u = rng.random(31 * 21)
    with eqf.FieldBatch([d], solver=eqf.solver_desc(_kind(eqf, solver))) as fb:
        fb.upload(0, u)
        fb.set_wall_values(0, oracle.TOP, g)
        fb.step()
        p.set_wall_values(oracle.TOP, g)
        r, _ = oracle.step(p, u, "lu")
        assert _rel(fb.download(0), r) <= 1e-9


def _parabolic(ny, H, vmax):
    y = np.arange(ny + 1) * (H / ny)
    return 4.0 * vmax * y * (H - y) / (H * H)


@pytest.mark.parametrize("solver", ["gmg_bicgstab", "mixed_bicgstab"])
def test_advection_bicgstab_matches_oracle(eqf, oracle, solver):
    """Channel-flow advection (parabolic x-velocity, zero at the walls) on uneven
    traps: the nonsymmetric operator is solved by GMG-preconditioned BiCGStab;
    three steps against the oracle's LU of the assembled operator; a trap
    without velocity in the same batch, and the operator itself (apply) vs the
    oracle's CSR; then the flow is reversed and halved for one more step."""
    kind = {"gmg_bicgstab": eqf.GMG_BICGSTAB, "mixed_bicgstab": eqf.MIXED_GMG_BICGSTAB}[solver]
    shapes = [(80, 40, 8.0, 4.0, 50.0), (37, 60, 3.7, 6.0, 20.0), (50, 25, 5.0, 2.5, 0.0)]
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0), oracle.LEFT: (oracle.ROBIN, 0, 1.0, 0.1)}
    descs, probs, u0, vel = [], [], [], []
    rng = np.random.default_rng(41)
    for (nx, ny, W, H, vmax) in shapes:
        for D in (100.0, 25.0):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls))
            p = oracle.problem(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls=walls)
            v = _parabolic(ny, H, vmax) if vmax > 0 else None
            p.set_velocity(v)
            probs.append(p)
            vel.append(v)
            u0.append(rng.random((nx + 1) * (ny + 1)))
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(kind)) as fb:
        for k in range(len(descs)):
            fb.upload(k, u0[k])
            if vel[k] is not None:
                fb.set_velocity(k, vel[k])
        # operator check on problem 0
        rp, c, v = oracle.assemble(probs[0])
        n0 = oracle.num_nodes(probs[0])
        x = rng.random(n0)
        v2, _ = oracle.apply_dirichlet(probs[0], rp, c, v, np.zeros(n0))
        y_ref = oracle.spmv(rp, c, v2, x)
        scale = oracle.spmv(rp, c, np.abs(v2), np.abs(x))
        y = fb.apply(0, x)
        assert np.all(np.abs(y - y_ref) <= 1e-14 * scale + 1e-300)
        ref = [u.copy() for u in u0]
        for step in range(3):
            st = fb.step()
            assert st.nconverged == len(descs), step
            for k in range(len(descs)):
                ref[k], _ = oracle.step(probs[k], ref[k], "lu")
                assert _rel(fb.download(k), ref[k]) <= 1e-9, (step, k)
        # a flow that changes between steps: new row velocities on the same batch
        for k in range(len(descs)):
            if vel[k] is not None:
                v2 = -0.5 * vel[k]
                fb.set_velocity(k, v2)
                probs[k].set_velocity(v2)
        st = fb.step()
        assert st.nconverged == len(descs)
        for k in range(len(descs)):
            ref[k], _ = oracle.step(probs[k], ref[k], "lu")
            assert _rel(fb.download(k), ref[k]) <= 1e-9, ("new flow", k)


@pytest.mark.parametrize("case", ["mixed", "mixed-fused", "gmg-fused", "bicgstab-velocity-fused", "mixed-tiles8",
                                  "gmg-fused-tiles8", "bicgstab-velocity-tiles4"])
def test_interior_tile_paths_match_oracle(eqf, oracle, case):
    """Traps large enough to hold interior tiles (and tiles whose one-node ring
    is interior), where the kernels take their register-window paths: the
    smoother, residual, prolongation, operator, CG update and BiCGStab
    apply, and -- with the finest level's pre-sweeps fused -- the fused
    first two sweeps; with a velocity the rows carry the convection.  Two
    steps against the oracle's LU.  The tiles8 / tiles4 cases lower the
    tile-shape thresholds so these traps run 8 rows per thread on every level,
    or 4 on every coarse level (the shapes large batches take)."""
    kind = {"mixed": eqf.MIXED_GMG_CG, "mixed-fused": eqf.MIXED_GMG_CG, "gmg-fused": eqf.GMG_CG,
            "bicgstab-velocity-fused": eqf.MIXED_GMG_BICGSTAB, "mixed-tiles8": eqf.MIXED_GMG_CG,
            "gmg-fused-tiles8": eqf.GMG_CG, "bicgstab-velocity-tiles4": eqf.MIXED_GMG_BICGSTAB}[case]
    fused = "fused" in case
    vel = case.startswith("bicgstab")
    shapes = [(260, 120, 13.0, 6.0), (200, 90, 10.0, 4.5)]
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0), oracle.LEFT: (oracle.ROBIN, 0, 1.0, 0.2)}
    descs, probs, u0, vels = [], [], [], []
    rng = np.random.default_rng(23)
    for (nx, ny, W, H) in shapes:
        for D in (200.0, 40.0):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls))
            p = oracle.problem(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls=walls)
            v = _parabolic(ny, H, 30.0) if vel else None
            p.set_velocity(v)
            probs.append(p)
            vels.append(v)
            u0.append(rng.random((nx + 1) * (ny + 1)))
    kw = {"mg_fuse_pre_top": 1, "mg_fuse_max_nodes": 2**31 - 1} if fused else {}
    if case.endswith("tiles8"):
        kw["tile8_min_nodes"] = 1
    if case.endswith("tiles4"):
        kw["tile4_min_nodes"] = 1
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(kind, **kw)) as fb:
        for k in range(len(descs)):
            fb.upload(k, u0[k])
            if vels[k] is not None:
                fb.set_velocity(k, vels[k])
        ref = [u.copy() for u in u0]
        for step in range(2):
            print(f"interior-tile case {case} step {step}", flush=True)
            st = fb.step()
            assert st.nconverged == len(descs), step
            for k in range(len(descs)):
                ref[k], _ = oracle.step(probs[k], ref[k], "lu")
                assert _rel(fb.download(k), ref[k]) <= 1e-9, (step, k)


@pytest.mark.parametrize("solver", ["gmg", "mixed", "mixed_bicgstab"])
def test_regridded_hierarchy_matches_oracle(eqf, oracle, solver):
    """Grids whose cell counts stop halving exactly: the hierarchy regrids
    once (about half, rounded up so the levels under it nest again), and the
    regridded transfer runs in separate launches while nested levels fuse.
    One batch mixes shapes that regrid on different levels (260x132 from
    its second coarse level, 65x33, on; 203x97 and 99x33 at once) with one
    that stays nested (128x64), so a level holds nested and regridded
    problems together.  Two steps against the oracle's LU; the BiCGStab case
    adds a velocity.  The spacing is 1/16 um, so every node coordinate is
    exact and the oracle's coordinate test finds every wall (at 203 cells
    over 10.15 um it misses the right wall, which this path marks by index)."""
    kind = {"gmg": eqf.GMG_CG, "mixed": eqf.MIXED_GMG_CG, "mixed_bicgstab": eqf.MIXED_GMG_BICGSTAB}[solver]
    vel = solver.endswith("bicgstab")
    shapes = [(260, 132, 16.25, 8.25), (203, 97, 12.6875, 6.0625), (128, 64, 8.0, 4.0), (99, 33, 6.1875, 2.0625)]
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0), oracle.RIGHT: (oracle.ROBIN, 0, 2.0, 0.1)}
    descs, probs, u0, vels = [], [], [], []
    rng = np.random.default_rng(31)
    for (nx, ny, W, H) in shapes:
        for D in (250.0, 60.0):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, 0.02, 0.0, walls))
            p = oracle.problem(nx, ny, W, H, 0.1, D, 0.02, 0.0, walls=walls)
            v = _parabolic(ny, H, 20.0) if vel else None
            p.set_velocity(v)
            probs.append(p)
            vels.append(v)
            u0.append(rng.random((nx + 1) * (ny + 1)))
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(kind)) as fb:
        for k in range(len(descs)):
            fb.upload(k, u0[k])
            if vels[k] is not None:
                fb.set_velocity(k, vels[k])
        ref = [u.copy() for u in u0]
        for step in range(2):
            print(f"regridded hierarchy {solver} step {step}", flush=True)
            st = fb.step()
            assert st.nconverged == len(descs), step
            for k in range(len(descs)):
                ref[k], _ = oracle.step(probs[k], ref[k], "lu")
                assert _rel(fb.download(k), ref[k]) <= 1e-9, (step, k)


def test_empty_cell_lists_and_rejected_calls(eqf, oracle):
    """No cells: coupling leaves the fields as they are (host and device
    entry points, and a batched call whose middle trap has none); a batched
    call still matches per-trap calls around the empty trap.  Calls the
    library cannot honour come back as EQ_ERR_INVALID with the state intact:
    a velocity under a CG solver, a velocity of the wrong length, a velocity
    on the finite-difference scheme, a problem index out of range."""
    import torch
    from eq_amd import workloads as Wl
    shapes = [(80, 40, 4.0, 2.0, 50), (60, 30, 3.0, 1.5, 0), (70, 50, 3.5, 2.5, 40)]
    npm = 20.0
    rng = Wl.SplitMix64(31)
    descs, cells, rates, u0, grids = [], [], [], [], []
    for (nx, ny, W, H, nc) in shapes:
        c, r = Wl.seeded_cells(rng, nc, W, H, 2)
        cells.append(c)
        rates.append(r)
        grids.append(oracle.grid(nx + 1, ny + 1, npm, W, H))
        g = np.random.default_rng(nx + ny)
        for s in range(2):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, 100.0))
            u0.append(g.random((nx + 1) * (ny + 1)))
    off = np.concatenate([[0], np.cumsum([c.shape[0] for c in cells])]).astype(np.int64)
    dev = torch.device("cuda")
    cells_d = torch.from_numpy(np.concatenate(cells)).to(dev).contiguous()
    q_d = torch.from_numpy(np.concatenate(rates)).to(dev).contiguous()
    samp_d = torch.zeros((int(off[-1]), 2), dtype=torch.float64, device=dev)
    cnt_d = torch.zeros(int(off[-1]), dtype=torch.int32, device=dev)
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(eqf.MIXED_GMG_CG)) as fb:
        for k, u in enumerate(u0):
            fb.upload(k, u)
        # no cells at all, host entry point
        fb.couple(1, np.zeros((0, 5)), npm, np.zeros((0, 2)))
        for k, u in enumerate(u0):
            assert np.array_equal(fb.download(k), u)
        fb.couple_traps_dev(0, off, cells_d.data_ptr(), npm, q_d.data_ptr(), samp_d.data_ptr(), cnt_d.data_ptr())
        fb.sync()
        got = [fb.download(k) for k in range(len(descs))]
        # rejected calls leave everything as it was
        vx = np.linspace(0.0, 1.0, 41)
        with pytest.raises(eqf.EqError) as e:
            fb.set_velocity(0, vx[:-1])
        assert e.value.code == 1
        with pytest.raises(eqf.EqError) as e:
            fb.set_velocity(len(descs), vx)
        assert e.value.code == 1
        fb.set_velocity(0, vx)
        with pytest.raises(eqf.EqError) as e:
            fb.step()
        assert e.value.code == 1 and "BiCGStab" in str(e.value)
        fb.set_velocity(0, None)
        st = fb.step()
        assert st.nconverged == len(descs)
    for t in range(len(shapes)):
        ref, s_ref, c_ref = oracle.couple(grids[t], cells[t], [u0[2 * t], u0[2 * t + 1]], rates[t])
        a, b = off[t], off[t + 1]
        assert np.array_equal(cnt_d.cpu().numpy()[a:b], c_ref)
        for s in range(2):
            assert np.array_equal(got[2 * t + s], ref[s]), (t, s)
    fd = eqf.field_desc(20, 10, 2.0, 1.0, 0.1, 10.0, scheme=eqf.SCHEME_FD)
    with eqf.FieldBatch([fd], solver=eqf.solver_desc(eqf.GMG_BICGSTAB)) as fb:
        with pytest.raises(eqf.EqError) as e:
            fb.set_velocity(0, np.ones(11))
        assert e.value.code == 1


def test_affine_coupling_matches_oracle(eqf, oracle):
    """Amounts that depend on the cell's own sample (q = a + b e, a strain's
    membrane exchange): two traps coupled in one call against the oracle's
    sequential loop; counts exact, samples and fields within rounding of the
    per-cell sums (the workgroup sums the covered nodes in another order, and
    each deposit depends on its sample); then a step against the oracle's LU.
    With b = 0 the fields equal the q-upfront coupling's bit for bit up to
    that summation order."""
    import torch
    from eq_amd import workloads as Wl
    npm = 10.0
    shapes = [(120, 60, 12.0, 6.0, 150), (80, 50, 8.0, 5.0, 90)]
    rng = Wl.SplitMix64(77)
    descs, probs, cells, grids, u, qa, qb = [], [], [], [], [], [], []
    for (nx, ny, W, H, nc) in shapes:
        c, r = Wl.seeded_cells(rng, nc, W, H, 2)
        cells.append(c)
        grids.append(oracle.grid(nx + 1, ny + 1, npm, W, H))
        qa.append(0.1 * r)
        qb.append(-0.05 * np.ones_like(r) * np.array([1.0, 0.5]))
        for D in (200.0, 60.0):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, 0.01, 0.0, Wl.TOP_DIRICHLET))
            probs.append(oracle.problem(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls=Wl.TOP_DIRICHLET))
            u.append(np.random.default_rng(nx + int(D)).random((nx + 1) * (ny + 1)))
    off = np.concatenate([[0], np.cumsum([c.shape[0] for c in cells])]).astype(np.int64)
    dev = torch.device("cuda")
    cells_d = torch.from_numpy(np.concatenate(cells)).to(dev).contiguous()
    a_d = torch.from_numpy(np.concatenate(qa)).to(dev).contiguous()
    b_d = torch.from_numpy(np.concatenate(qb)).to(dev).contiguous()
    samp_d = torch.zeros((int(off[-1]), 2), dtype=torch.float64, device=dev)
    cnt_d = torch.zeros(int(off[-1]), dtype=torch.int32, device=dev)
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(eqf.MIXED_GMG_CG)) as fb:
        for k in range(len(descs)):
            fb.upload(k, u[k])
        fb.couple_affine_dev(0, off, cells_d.data_ptr(), npm, a_d.data_ptr(), b_d.data_ptr(), samp_d.data_ptr(),
                             cnt_d.data_ptr())
        fb.sync()
        samp, cnt = samp_d.cpu().numpy(), cnt_d.cpu().numpy()
        for t in range(len(shapes)):
            a, b = off[t], off[t + 1]
            fields, s_ref, c_ref = oracle.couple_affine(grids[t], cells[t], u[2 * t:2 * t + 2], qa[t], qb[t])
            assert np.array_equal(cnt[a:b], c_ref)
            assert np.allclose(samp[a:b], s_ref, rtol=1e-13, atol=1e-300)
            for s in range(2):
                got = fb.download(2 * t + s)
                assert np.allclose(got, fields[s], rtol=1e-14, atol=1e-15), (t, s)
                u[2 * t + s] = got
        st = fb.step()
        assert st.nconverged == len(descs)
        for k in range(len(descs)):
            ref, _ = oracle.step(probs[k], u[k], "lu")
            assert _rel(fb.download(k), ref) <= 1e-9, k


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["gmg", "mixed"])
@pytest.mark.parametrize("traps,streams", [(1, 2), (3, 2), (3, 3)])
def test_two_problem_groups_match_one(eqf, oracle, solver, traps, streams):
    """The CG kinds solve a batch as problem groups on their own streams
    (solver option streams = 2, the default, or 3) or as one (streams = 1):
    every problem's field bit for bit the same, and its iteration count, over
    three steps of species with different diffusivities (so the groups stop
    at different counts); the oracle's LU within 1e-9."""
    kind = {"gmg": eqf.GMG_CG, "mixed": eqf.MIXED_GMG_CG}[solver]
    shapes = [(300, 150, 18.75, 9.375), (203, 97, 12.6875, 6.0625), (131, 67, 8.1875, 4.1875)][:traps]
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0), oracle.LEFT: (oracle.ROBIN, 0, 1.5, 0.2)}
    descs, probs, u0 = [], [], []
    rng = np.random.default_rng(43)
    for (nx, ny, W, H) in shapes:
        for D in (400.0, 30.0):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls))
            probs.append(oracle.problem(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls=walls))
            u0.append(rng.random((nx + 1) * (ny + 1)))
    out, its = {}, {}
    for ns in (streams, 1):
        with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(kind, streams=ns)) as fb:
            for k in range(len(descs)):
                fb.upload(k, u0[k])
            u, n = [], []
            for step in range(3):
                print(f"problem groups {solver} traps={traps} streams={ns} step {step}", flush=True)
                st = fb.step()
                assert st.nconverged == len(descs), (ns, step)
                u.append([fb.download(k) for k in range(len(descs))])
                n.append(list(fb.problem_stats()[0]))
            out[ns], its[ns] = u, n
    ref = [x.copy() for x in u0]
    for step in range(3):
        assert its[streams][step] == its[1][step], (step, its[streams][step], its[1][step])
        for k in range(len(descs)):
            assert np.array_equal(out[streams][step][k], out[1][step][k]), (step, k)
            ref[k], _ = oracle.step(probs[k], ref[k], "lu")
            assert _rel(out[streams][step][k], ref[k]) <= 1e-9, (step, k)


@pytest.mark.parametrize("solver", ["gmg", "mixed", "ir"])
def test_fused_levels_match_separate_launches(eqf, oracle, solver):
    """The fused level kernels (a level's way down and way up in one launch
    each, level 0 from the Krylov update's first sweep with r.z in its own
    partial slots) against the separate launches (mg_level_fused = 0) and the
    oracle's LU.  One batch holds a trap whose whole V-cycle runs in one
    workgroup (its r.z partials stay in the level's own slots), nested and
    regridded hierarchies, odd sizes, and Dirichlet, Robin and Neumann walls,
    so the fused kernels meet interior tiles, wall tiles, walls with and
    without Dirichlet rows, and both transfer forms.  Two steps: fields
    within 1e-9 of the oracle either way, iteration counts within one."""
    kind = {"gmg": eqf.GMG_CG, "mixed": eqf.MIXED_GMG_CG, "ir": eqf.IR_GMG_CG}[solver]
    shapes = [(40, 20, 2.5, 1.25), (300, 150, 18.75, 9.375), (203, 97, 12.6875, 6.0625), (131, 67, 8.1875, 4.1875)]
    wallsets = [{oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0), oracle.LEFT: (oracle.ROBIN, 0, 1.5, 0.2)},
                {w: (oracle.DIRICHLET, 0.25, 0, 0) for w in range(4)}]
    descs, probs, u0 = [], [], []
    rng = np.random.default_rng(41)
    for t, (nx, ny, W, H) in enumerate(shapes):
        walls = wallsets[t % 2]
        for D in (300.0, 50.0):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls))
            probs.append(oracle.problem(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls=walls))
            u0.append(rng.random((nx + 1) * (ny + 1)))
    out, its = {}, {}
    for fused in (1, 0):
        with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(kind, mg_level_fused=fused)) as fb:
            for k in range(len(descs)):
                fb.upload(k, u0[k])
            u, n = [], []
            for step in range(2):
                print(f"fused levels {solver} fused={fused} step {step}", flush=True)
                st = fb.step()
                assert st.nconverged == len(descs), (fused, step)
                u.append([fb.download(k) for k in range(len(descs))])
                n.append(list(fb.problem_stats()[0]))
            out[fused], its[fused] = u, n
    ref = [x.copy() for x in u0]
    for step in range(2):
        for k in range(len(descs)):
            ref[k], _ = oracle.step(probs[k], ref[k], "lu")
            for fused in (1, 0):
                assert _rel(out[fused][step][k], ref[k]) <= 1e-9, (fused, step, k)
        assert all(abs(a - b) <= 1 for a, b in zip(its[1][step], its[0][step])), (step, its[1][step], its[0][step])
