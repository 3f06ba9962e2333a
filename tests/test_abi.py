"""CPU checks of the C-ABI library: it is built, loads, and exports exactly the
entry points include/eq_field.h declares.  No GPU compute is called here."""
import math
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "eq_field.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(eq_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_binding_list():
    from eq_amd import field
    assert sorted(field.EXPORTED) == declared_functions()


def test_library_exports_every_declared_symbol():
    from eq_amd import field
    assert os.path.exists(field.LIB_PATH), "run eq_amd/build.py (or __graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", field.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (eq_\w+)$", out, flags=re.M))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_and_reports_defaults():
    from eq_amd import field
    L = field.load_library()
    s = field.solver_desc()
    assert s.kind == field.MIXED_GMG_CG and s.rtol == 1e-10
    assert s.mg_pre == 2 and s.mg_post == 2
    # the Chebyshev dampings of two sweeps on [0.5, 2]: 1 / (1.25 -+ 0.75 cos(pi/4))
    w1, w2 = 1.0 / (1.25 + 0.75 * math.cos(math.pi / 4)), 1.0 / (1.25 - 0.75 * math.cos(math.pi / 4))
    assert abs(s.mg_omega - w1) < 1e-15 and abs(s.mg_omega2 - w2) < 1e-15
    assert s.mg_level_fused == 1
    assert s.streams == 2
    assert L.eq_last_error() is not None


def test_invalid_descriptor_is_rejected_before_any_device_work():
    from eq_amd import field
    bad = field.field_desc(0, 4, 1.0, 1.0, 0.1, 1.0)
    with pytest.raises(field.EqError) as e:
        field.FieldBatch([bad])
    assert e.value.code == 1  # EQ_ERR_INVALID


@pytest.mark.parametrize("case", ["nx0", "dt0", "negD", "neg_gamma", "nan_f", "bad_wall", "neg_robin", "bad_scheme",
                                  "huge", "species_grid", "nprob_species", "solver_kind", "rtol0", "mg_pre0"])
def test_bad_inputs_are_rejected_before_any_device_work(case):
    """Every descriptor and solver check runs before the library asks for a
    device, so each bad input comes back as EQ_ERR_INVALID here, without a GPU."""
    from eq_amd import field as F
    good = dict(nx=8, ny=4, width=2.0, height=1.0, dt=0.1, D=1.0, gamma=0.0, f=0.0)
    descs = [F.field_desc(**good)]
    nspecies, solver = 1, F.solver_desc()
    d = descs[0]
    if case == "nx0":
        d.nx = 0
    elif case == "dt0":
        d.dt = 0.0
    elif case == "negD":
        d.D = -1.0
    elif case == "neg_gamma":
        d.gamma = -0.1
    elif case == "nan_f":
        d.f = float("nan")
    elif case == "bad_wall":
        d.wall[2].kind = 7
    elif case == "neg_robin":
        d.wall[0].kind, d.wall[0].r = 2, -1.0
    elif case == "bad_scheme":
        d.scheme = 5
    elif case == "huge":
        d.nx, d.ny = 70000, 4
    elif case == "species_grid":
        e = F.field_desc(**dict(good, nx=9))
        descs, nspecies = [d, e], 2
    elif case == "nprob_species":
        descs, nspecies = [d, F.field_desc(**good), F.field_desc(**good)], 2
    elif case == "solver_kind":
        solver.kind = 9
    elif case == "rtol0":
        solver.rtol = 0.0
    elif case == "mg_pre0":
        solver.mg_pre = 0
    with pytest.raises(F.EqError) as e:
        F.FieldBatch(descs, nspecies=nspecies, solver=solver)
    assert e.value.code == 1, str(e.value)  # EQ_ERR_INVALID


def test_workload_generators_are_deterministic():
    import numpy as np
    from eq_amd import workloads as W
    a, b = W.config0(), W.config0()
    assert np.array_equal(a.cells, b.cells) and np.array_equal(a.rates, b.rates)
    assert a.cells.shape == (200, 5) and a.nodes == 13041
    assert (a.cells[:, 3] >= 2).all() and (a.cells[:, 3] <= 4).all()
    assert (a.rates >= 0.5).all() and (a.rates <= 1.5).all()
    c1 = W.random_walk_step(a.cells, 0, 3, 0.05, 40.0, 20.0)
    c2 = W.random_walk_step(a.cells, 0, 3, 0.05, 40.0, 20.0)
    assert np.array_equal(c1, c2)
    d = (c1[:, :2] - a.cells[:, :2]).ravel()
    assert abs(d.std() - 0.05) < 0.01 and np.abs(d).max() <= 0.3 + 1e-12


def test_walls_found_matches_the_oracle_on_config3_traps():
    """The library's restatement of the reference's coordinate wall test
    agrees with the oracle's on all 512 config-3 traps: 41 lose the top wall
    (the computed y of the top row misses H by more than DOLFIN_EPS), 39 the
    right wall.  Integer trap sizes lose none, because (n*W)/n is then exact.
    Host logic only: no device is asked for."""
    from eq_amd import field as F
    from eq_amd import workloads as W
    from oracle import oracle as O
    lost_top = lost_right = 0
    for t in W.config3_traps(512, 3):
        d = F.field_desc(t.nx, t.ny, t.width, t.height, 0.1, 1.0)
        found = F.walls_found(d)
        assert found == O.walls_found(O.problem(t.nx, t.ny, t.width, t.height, 0.1, 1.0))
        assert found[F.LEFT] and found[F.BOTTOM]
        lost_top += not found[F.TOP]
        lost_right += not found[F.RIGHT]
    assert (lost_top, lost_right) == (41, 39)
    for nx, ny, W_, H in [(160, 80, 40.0, 20.0), (2000, 1000, 100.0, 50.0), (8192, 4096, 204.8, 102.4),
                          (1000, 500, 100.0, 50.0), (37, 23, 37.0, 23.0)]:
        assert all(F.walls_found(F.field_desc(nx, ny, W_, H, 0.1, 1.0)))
    # the grid rule finds every wall in the oracle, whatever the coordinates
    assert all(O.walls_found(O.problem(218, 440, 21.8, 44.0, 0.1, 1.0, grid_walls=True)))


def test_bad_wall_marking_is_rejected():
    from eq_amd import field as F
    d = F.field_desc(8, 4, 2.0, 1.0, 0.1, 1.0)
    d.marking = 3
    with pytest.raises(F.EqError) as e:
        F.walls_found(d)
    assert e.value.code == 1
    with pytest.raises(F.EqError) as e:
        F.FieldBatch([d])
    assert e.value.code == 1
