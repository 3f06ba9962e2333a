"""GPU parity of the boundary-value updates through the C ABI.

  setBoundaryValues (src/fHSL.cpp:601-603): a stream-ordered update of every
    Dirichlet wall's value between steps, against the oracle's LU with the
    same wall values;
  the boundary well of Simulation::computeBoundaryWell with the
    DIRICHLET_UPDATE boundary (src/simulation.cpp:581-604): updated on the
    device after every step from the new field's boundary flux; each step's
    field within 1e-9 relative L2 of the oracle's LU step whose four walls
    carry the oracle's well value, the well value within 1e-8 relative of the
    oracle's (the flux is a boundary sum of the solution, which carries the
    solve's 1e-10 residual), and the underflow count exact.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


ALL_DIRICHLET = {w: (1, 0.0, 0.0, 0.0) for w in range(4)}


def test_set_boundary_value_between_steps(eqf, oracle):
    nx, ny, W, H = 90, 44, 9.0, 4.4
    walls = {oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0), oracle.LEFT: (oracle.DIRICHLET, 0.0, 0, 0),
             oracle.RIGHT: (oracle.ROBIN, 0, 1.5, 0.2)}
    descs = [eqf.field_desc(nx, ny, W, H, 0.1, D, 0.01, 0.0, walls) for D in (200.0, 50.0)]
    u = [np.random.default_rng(s).random((nx + 1) * (ny + 1)) for s in range(2)]
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(eqf.MIXED_GMG_CG)) as fb:
        for s in range(2):
            fb.upload(s, u[s])
        for step, vals in enumerate([(0.25, 1.5), (0.0, 0.75), (2.0, 2.0)]):
            for s in range(2):
                fb.set_boundary_value(s, vals[s])
            st = fb.step()
            assert st.nconverged == 2
            for s in range(2):
                wl = dict(walls)
                wl[oracle.TOP] = (oracle.DIRICHLET, vals[s], 0, 0)
                wl[oracle.LEFT] = (oracle.DIRICHLET, vals[s], 0, 0)
                p = oracle.problem(nx, ny, W, H, 0.1, descs[s].D, 0.01, 0.0, walls=wl)
                u[s], _ = oracle.step(p, u[s], "lu")
                assert _rel(fb.download(s), u[s]) <= 1e-9, (step, s)
        # a problem without a Dirichlet wall is refused
    with eqf.FieldBatch([eqf.field_desc(8, 4, 2.0, 1.0, 0.1, 1.0)], solver=eqf.solver_desc(eqf.GMG_CG)) as fb:
        with pytest.raises(eqf.EqError) as e:
            fb.set_boundary_value(0, 1.0)
        assert e.value.code == 1


@pytest.mark.parametrize("solver", ["gmg", "mixed", "mixed_bicgstab"])
def test_boundary_well_steps_match_oracle(eqf, oracle, solver):
    """Two species on one trap with DIRICHLET_UPDATE walls; species 0 starts
    from a field above the wall value (outflow raises the well), species 1
    from zero with a high starting well (inflow drains it past zero: a clamp
    and an underflow)."""
    kind = {"gmg": eqf.GMG_CG, "mixed": eqf.MIXED_GMG_CG, "mixed_bicgstab": eqf.MIXED_GMG_BICGSTAB}[solver]
    nx, ny, W, H = 120, 60, 30.0, 15.0
    dt, lengthScaling = 0.1, 5.0
    scaling = 2.0 * 10.0 * (15.0 / lengthScaling) * (W + H)      # initBoundaryWell
    decay = 0.5   # boundaryDecayRate (1/s); dt * decay < 1, as setBoundaryRate requires
    Ds = (400.0, 100.0)
    descs = [eqf.field_desc(nx, ny, W, H, dt, D, 0.01, 0.0, ALL_DIRICHLET) for D in Ds]
    n = (nx + 1) * (ny + 1)
    u = [10.0 * np.random.default_rng(4).random(n), np.zeros(n)]
    w = [0.0, 3.0]
    scal = [scaling, 0.05]
    under = [0, 0]
    with eqf.FieldBatch(descs, nspecies=2, solver=eqf.solver_desc(kind)) as fb:
        for s in range(2):
            fb.upload(s, u[s])
            fb.set_well(s, scal[s], decay, w[s])
        for step in range(5):
            st = fb.step()
            assert st.nconverged == 2
            for s in range(2):
                wl = {k: (oracle.DIRICHLET, w[s], 0, 0) for k in range(4)}
                p = oracle.problem(nx, ny, W, H, dt, Ds[s], 0.01, 0.0, walls=wl)
                u[s], _ = oracle.step(p, u[s], "lu")
                flux = dt * Ds[s] * oracle.boundary_flux(p, u[s])
                w[s], k = oracle.boundary_well(w[s], flux, scal[s], decay, dt)
                under[s] += k
                got = fb.download(s)
                assert _rel(got, u[s]) <= 1e-9, (step, s)
                gw, gflux, gunder = fb.well_state(s)
                assert abs(gw - w[s]) <= 1e-8 * max(abs(w[s]), 1e-12) + 1e-14, (step, s, gw, w[s])
                assert abs(gflux - flux) <= 1e-8 * abs(flux) + 1e-12, (step, s, gflux, flux)
                assert gunder == under[s], (step, s)
        assert under[1] >= 1 and w[0] > 0.0
        # host-side set-up calls after device-side well updates keep the well's wall value
        fb.set_velocity(0, None)
        fb.set_well(1, None)
        st = fb.step()
        for s in range(2):
            wl = {k: (oracle.DIRICHLET, w[s], 0, 0) for k in range(4)}
            p = oracle.problem(nx, ny, W, H, dt, Ds[s], 0.01, 0.0, walls=wl)
            u[s], _ = oracle.step(p, u[s], "lu")
            if s == 0:
                flux = dt * Ds[s] * oracle.boundary_flux(p, u[s])
                w[s], _ = oracle.boundary_well(w[s], flux, scal[s], decay, dt)
            assert _rel(fb.download(s), u[s]) <= 1e-9, ("after removal", s)
        with pytest.raises(eqf.EqError):
            fb.well_state(1)


def test_well_needs_all_walls_dirichlet(eqf, oracle):
    d = eqf.field_desc(20, 10, 2.0, 1.0, 0.1, 10.0, walls={oracle.TOP: (oracle.DIRICHLET, 0.0, 0, 0)})
    with eqf.FieldBatch([d], solver=eqf.solver_desc(eqf.GMG_CG)) as fb:
        with pytest.raises(eqf.EqError) as e:
            fb.set_well(0, 10.0, 0.1, 0.0)
        assert e.value.code == 1
