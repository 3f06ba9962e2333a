"""GPU parity of the diffusion tensor set from the cells' orientations
(setDiffusionTensor of eQabm::updateCells, src/abm/eQabm.cpp:241-248,
306-325, 407), through the C ABI.

Each coupling call that carries EQ_COUPLE_TENSOR resets the enabled traps'
tensor to (1, 1, 0) and writes every covered node from its cells in list
order; the coarse levels and the coarsest inverses follow on the device.
Checked against the oracle's restatement (orc_cell_tensor): the finest-level
tensor within 4 ulp (the device's cos/sin of the angle may differ from the
host libm's by an ulp), the coupled fields bit-identical, and each step
within 1e-9 relative L2 of the oracle's LU step with the oracle's tensor.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("solver", ["jacobi", "gmg", "mixed"])
def test_cell_tensor_steps_match_oracle(eqf, oracle, solver):
    import torch
    from eq_amd import workloads as Wl
    kind = {"jacobi": eqf.JACOBI_CG, "gmg": eqf.GMG_CG, "mixed": eqf.MIXED_GMG_CG}[solver]
    npm, Dx, Dy = 10.0, 3.0, 0.5
    # two traps: the first takes its tensor from its cells, the second stays isotropic
    shapes = [(130, 70, 13.0, 7.0, 90), (90, 50, 9.0, 5.0, 40)]
    walls = Wl.TOP_DIRICHLET
    rng = Wl.SplitMix64(17)
    descs, probs, cells, rates, grids, u = [], [], [], [], [], []
    for (nx, ny, W, H, nc) in shapes:
        c, r = Wl.seeded_cells(rng, nc, W, H, 2)
        cells.append(c)
        rates.append(r * 0.1)
        grids.append(oracle.grid(nx + 1, ny + 1, npm, W, H))
        for D, g in ((300.0, 0.01), (80.0, 0.02)):
            descs.append(eqf.field_desc(nx, ny, W, H, 0.1, D, g, 0.0, walls))
            probs.append(oracle.problem(nx, ny, W, H, 0.1, D, g, 0.0, walls=walls))
            u.append(np.random.default_rng(nx + int(D)).random((nx + 1) * (ny + 1)))
    off = np.concatenate([[0], np.cumsum([c.shape[0] for c in cells])]).astype(np.int64)
    dev = torch.device("cuda")
    q_d = torch.from_numpy(np.concatenate(rates)).to(dev).contiguous()
    samp_d = torch.zeros((int(off[-1]), 2), dtype=torch.float64, device=dev)
    flags = eqf.SAMPLE | eqf.DEPOSIT | eqf.TENSOR
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(kind)) as fb:
        fb.cell_tensor(0, Dx, Dy)
        for step in range(3):
            print(f"cell tensor {solver} step {step}", flush=True)
            cells = [Wl.random_walk_step(c, 17, step, 0.2, shapes[t][2], shapes[t][3]) for t, c in enumerate(cells)]
            cells_d = torch.from_numpy(np.concatenate(cells)).to(dev).contiguous()
            for k in range(len(descs)):
                fb.upload(k, u[k])
            fb.couple_traps_dev(0, off, cells_d.data_ptr(), npm, q_d.data_ptr(), samp_d.data_ptr(), 0, flags)
            tens = []
            for t in range(len(shapes)):
                a, b = off[t], off[t + 1]
                u[2 * t], u[2 * t + 1] = oracle.couple(grids[t], cells[t], u[2 * t:2 * t + 2], rates[t])[0]
                ref_t = oracle.cell_tensor(grids[t], cells[t], Dx, Dy) if t == 0 else None
                tens.append(ref_t)
                for s in range(2):
                    assert np.array_equal(fb.download(2 * t + s), u[2 * t + s]), (step, t, s)
                    got = fb.tensor_values(2 * t + s)
                    if ref_t is None:
                        assert np.all(got[0] == 1.0) and np.all(got[1] == 1.0) and np.all(got[2] == 0.0)
                    else:
                        for c in range(3):
                            assert np.allclose(got[c], ref_t[c], rtol=4 * 2.2e-16, atol=4 * 2.2e-16 * Dx), (step, c)
                        assert np.count_nonzero(got[0] != 1.0) > 1000
            st = fb.step()
            assert st.nconverged == len(descs), step
            for k in range(len(descs)):
                t = k // 2
                kw = {} if tens[t] is None else {"D11": tens[t][0], "D22": tens[t][1], "D12": tens[t][2]}
                u[k], _ = oracle.step(probs[k], u[k], "lu", **kw)
                assert _rel(fb.download(k), u[k]) <= 1e-9, (step, k)
        # back to constant coefficients
        fb.cell_tensor(0, enable=False)
        for k in range(len(descs)):
            fb.upload(k, u[k])
        st = fb.step()
        assert st.nconverged == len(descs)
        for k in range(len(descs)):
            u[k], _ = oracle.step(probs[k], u[k], "lu")
            assert _rel(fb.download(k), u[k]) <= 1e-9, ("isotropic again", k)
            got = fb.tensor_values(k)
            assert np.all(got[0] == 1.0) and np.all(got[2] == 0.0)


def test_cell_tensor_rejects_velocity_and_fd(eqf, oracle):
    d = eqf.field_desc(40, 20, 4.0, 2.0, 0.1, 50.0, 0.01, 0.0, {2: (1, 0.0, 0.0, 0.0)})
    with eqf.FieldBatch([d], solver=eqf.solver_desc(eqf.MIXED_GMG_BICGSTAB)) as fb:
        fb.set_velocity(0, np.linspace(0.0, 1.0, 21))
        with pytest.raises(eqf.EqError) as e:
            fb.cell_tensor(0, 2.0, 1.0)
        assert e.value.code == 1
    fd = eqf.field_desc(20, 10, 2.0, 1.0, 0.1, 10.0, scheme=eqf.SCHEME_FD)
    with eqf.FieldBatch([fd], solver=eqf.solver_desc(eqf.GMG_CG)) as fb:
        with pytest.raises(eqf.EqError) as e:
            fb.cell_tensor(0, 2.0, 1.0)
        assert e.value.code == 1
