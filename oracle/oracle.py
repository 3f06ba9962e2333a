"""ctypes view of the CPU oracle (oracle/lib/libeq_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and there only as the checker or the timed CPU
baseline.  The product package ``eq_amd`` never imports this module.

Every function restates a part of eQ's field step; see eq_oracle.h for the
reference file:line each one follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libeq_oracle.so")

NEUMANN, DIRICHLET, ROBIN = 0, 1, 2
LEFT, RIGHT, TOP, BOTTOM = 0, 1, 2, 3


class Bc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("value", C.c_double),
                ("r", C.c_double), ("s", C.c_double)]


class Problem(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("width", C.c_double),
                ("height", C.c_double), ("dt", C.c_double), ("D", C.c_double),
                ("gamma", C.c_double), ("f", C.c_double), ("wall", Bc * 4),
                ("gwall", C.c_void_p * 4), ("vx", C.c_void_p), ("grid_walls", C.c_int32),
                ("pad_", C.c_int32)]

    def set_velocity(self, vx_rows):
        """x-velocity per grid row (ny+1 values), or None."""
        if not hasattr(self, "_keep"):
            self._keep = {}
        if vx_rows is None:
            self.vx = None
            self._keep.pop("vx", None)
            return
        a = np.ascontiguousarray(vx_rows, dtype=np.float64).copy()
        self._keep["vx"] = a
        self.vx = a.ctypes.data

    def set_wall_values(self, wall: int, values):
        """Per-node Dirichlet values along a wall (None: the wall's constant)."""
        if not hasattr(self, "_keep"):
            self._keep = {}
        if values is None:
            self.gwall[wall] = None
            self._keep.pop(wall, None)
            return
        a = np.ascontiguousarray(values, dtype=np.float64).copy()
        self._keep[wall] = a
        self.gwall[wall] = a.ctypes.data


class Channel(C.Structure):
    _fields_ = [("v", C.c_double), ("r_left", C.c_double), ("s_left", C.c_double),
                ("r_right", C.c_double), ("s_right", C.c_double), ("well_scaling", C.c_double),
                ("substeps", C.c_int32), ("pad_", C.c_int32)]


class Cell(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("angle", C.c_double),
                ("length", C.c_double), ("width", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("nodes_w", C.c_int32), ("nodes_h", C.c_int32), ("npm", C.c_double),
                ("trap_w", C.c_double), ("trap_h", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("pad_", C.c_int32), ("relres", C.c_double),
                ("bnorm", C.c_double)]


def build(force: bool = False) -> str:
    """Compile the oracle with make (g++); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        dp, ip64, ip32 = P(C.c_double), P(C.c_int64), P(C.c_int32)
        L.orc_num_nodes.restype = C.c_int64
        L.orc_num_nodes.argtypes = [P(Problem)]
        L.orc_mesh_coords.argtypes = [P(Problem), dp]
        L.orc_walls_found.argtypes = [P(Problem), ip32]
        L.orc_assemble.restype = C.c_int64
        L.orc_assemble.argtypes = [P(Problem), dp, dp, dp, ip64, ip32, dp]
        L.orc_assemble_rhs.argtypes = [P(Problem), dp, dp]
        L.orc_dirichlet_nodes.restype = C.c_int64
        L.orc_dirichlet_nodes.argtypes = [P(Problem), P(C.c_uint8), dp]
        L.orc_apply_dirichlet.argtypes = [P(Problem), ip64, ip32, dp, dp]
        L.orc_spmv.argtypes = [C.c_int64, ip64, ip32, dp, dp, dp]
        L.orc_solve_banded.restype = C.c_int
        L.orc_solve_banded.argtypes = [P(Problem), ip64, ip32, dp, dp, dp]
        L.orc_solve_pcg.restype = C.c_int
        L.orc_solve_pcg.argtypes = [P(Problem), ip64, ip32, dp, dp, dp, C.c_double, C.c_int32,
                                    P(SolveInfo)]
        L.orc_channel_update.restype = None
        L.orc_channel_update.argtypes = [P(Problem), P(Channel), P(Channel), dp, dp, dp]
        L.orc_fd_step.restype = C.c_int
        L.orc_fd_step.argtypes = [P(Problem), dp]
        L.orc_step.restype = C.c_int
        L.orc_step.argtypes = [P(Problem), dp, dp, dp, dp, C.c_int32, C.c_double, C.c_int32, P(SolveInfo)]
        L.orc_cell_coverage.restype = C.c_int64
        L.orc_cell_coverage.argtypes = [P(Grid), P(Cell), ip32, C.c_int64]
        L.orc_cell_window.argtypes = [P(Grid), P(Cell), ip32]
        L.orc_couple.argtypes = [P(Grid), P(Cell), C.c_int64, C.c_int32, P(dp), dp, dp, ip32]
        L.orc_cell_tensor.argtypes = [P(Grid), P(Cell), C.c_int64, C.c_double, C.c_double, dp, dp, dp]
        L.orc_couple_affine.argtypes = [P(Grid), P(Cell), C.c_int64, C.c_int32, P(dp), dp, dp, dp, ip32]
        L.orc_boundary_well.restype = C.c_double
        L.orc_boundary_well.argtypes = [C.c_double] * 5 + [ip64]
        L.orc_boundary_flux.restype = C.c_double
        L.orc_boundary_flux.argtypes = [P(Problem), dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def problem(nx, ny, width, height, dt, D, gamma=0.0, f=0.0, walls=None, grid_walls=False) -> Problem:
    """walls: dict wall-index -> (kind, value, r, s); missing walls are Neumann.
    grid_walls: find the walls by grid index instead of the reference's
    near() test on the computed vertex coordinates."""
    p = Problem()
    p.grid_walls = 1 if grid_walls else 0
    p.nx, p.ny, p.width, p.height = nx, ny, width, height
    p.dt, p.D, p.gamma, p.f = dt, D, gamma, f
    for w in range(4):
        kind, value, r, s = (walls or {}).get(w, (NEUMANN, 0.0, 0.0, 0.0))
        p.wall[w].kind, p.wall[w].value, p.wall[w].r, p.wall[w].s = kind, value, r, s
    return p


def num_nodes(p: Problem) -> int:
    return (p.nx + 1) * (p.ny + 1)


def walls_found(p: Problem) -> list:
    """Per wall (left, right, top, bottom): does the wall carry its condition."""
    f = np.empty(4, np.int32)
    lib().orc_walls_found(C.byref(p), f.ctypes.data_as(C.POINTER(C.c_int32)))
    return [bool(v) for v in f]


def mesh_coords(p: Problem) -> np.ndarray:
    xy = np.empty(2 * num_nodes(p))
    lib().orc_mesh_coords(C.byref(p), _dp(xy))
    return xy.reshape(-1, 2)


def assemble(p: Problem, D11=None, D22=None, D12=None):
    """CSR (rowptr, col, val) of A = (1+dt*gamma)M + dt*D*K + Robin (no Dirichlet)."""
    n = num_nodes(p)
    rowptr = np.empty(n + 1, np.int64)
    col = np.empty(7 * n, np.int32)
    val = np.empty(7 * n)
    t = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (D11, D22, D12)]
    nnz = lib().orc_assemble(C.byref(p), _dp(t[0]), _dp(t[1]), _dp(t[2]),
                             rowptr.ctypes.data_as(C.POINTER(C.c_int64)),
                             col.ctypes.data_as(C.POINTER(C.c_int32)), _dp(val))
    return rowptr, col[:nnz].copy(), val[:nnz].copy()


def to_scipy(rowptr, col, val, n):
    import scipy.sparse as sp
    return sp.csr_matrix((val, col, rowptr), shape=(n, n))


def assemble_rhs(p: Problem, u0) -> np.ndarray:
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    b = np.empty(num_nodes(p))
    lib().orc_assemble_rhs(C.byref(p), _dp(u0), _dp(b))
    return b


def dirichlet_nodes(p: Problem):
    n = num_nodes(p)
    mask = np.empty(n, np.uint8)
    g = np.empty(n)
    lib().orc_dirichlet_nodes(C.byref(p), mask.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(g))
    return mask.astype(bool), g


def apply_dirichlet(p: Problem, rowptr, col, val, b):
    val = val.copy()
    b = None if b is None else b.copy()
    lib().orc_apply_dirichlet(C.byref(p), rowptr.ctypes.data_as(C.POINTER(C.c_int64)),
                              col.ctypes.data_as(C.POINTER(C.c_int32)), _dp(val), _dp(b))
    return val, b


def spmv(rowptr, col, val, x) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(len(rowptr) - 1)
    lib().orc_spmv(len(rowptr) - 1, rowptr.ctypes.data_as(C.POINTER(C.c_int64)),
                   col.ctypes.data_as(C.POINTER(C.c_int32)), _dp(val), _dp(x), _dp(y))
    return y


def solve_banded(p: Problem, rowptr, col, val, b) -> np.ndarray:
    x = np.empty(num_nodes(p))
    rc = lib().orc_solve_banded(C.byref(p), rowptr.ctypes.data_as(C.POINTER(C.c_int64)),
                                col.ctypes.data_as(C.POINTER(C.c_int32)), _dp(val),
                                _dp(np.ascontiguousarray(b)), _dp(x))
    if rc:
        raise RuntimeError(f"banded LU failed rc={rc}")
    return x


def step(p: Problem, u, method="lu", rtol=1e-14, D11=None, D22=None, D12=None, maxit=0):
    """One reference step on a copy of u; returns (u_plus, SolveInfo).
    maxit > 0 caps the PCG iterations (used to time a bounded sample)."""
    u = np.array(u, dtype=np.float64, copy=True)
    info = SolveInfo()
    t = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (D11, D22, D12)]
    rc = lib().orc_step(C.byref(p), _dp(t[0]), _dp(t[1]), _dp(t[2]), _dp(u),
                        0 if method == "lu" else 1, rtol, maxit, C.byref(info))
    if rc and method == "lu":
        raise RuntimeError(f"oracle LU step failed rc={rc}")
    # PCG may stop at the rounding floor of the CSR residual above rtol;
    # callers read info.relres
    return u, info


def channel(v, r_left, s_left, r_right, s_right, well_scaling, substeps) -> Channel:
    ch = Channel()
    ch.v, ch.r_left, ch.s_left, ch.r_right, ch.s_right = v, r_left, s_left, r_right, s_right
    ch.well_scaling, ch.substeps = well_scaling, substeps
    return ch


def channel_update(p: Problem, top_ch, bottom_ch, u, top, bottom):
    """Channel part of the step after the trap solve; returns new (top, bottom)."""
    t = np.array(top, dtype=np.float64, copy=True)
    b = np.array(bottom, dtype=np.float64, copy=True)
    lib().orc_channel_update(C.byref(p), C.byref(top_ch), C.byref(bottom_ch), _dp(np.ascontiguousarray(u, np.float64)),
                             _dp(t), _dp(b))
    return t, b


def fd_step(p: Problem, u):
    """One step of the diffusionPETSc finite-difference scheme (exact solve) on a copy of u."""
    u = np.array(u, dtype=np.float64, copy=True)
    rc = lib().orc_fd_step(C.byref(p), _dp(u))
    if rc:
        raise RuntimeError(f"oracle FD step failed rc={rc}")
    return u


def grid(nodes_w, nodes_h, npm, trap_w, trap_h) -> Grid:
    g = Grid()
    g.nodes_w, g.nodes_h, g.npm, g.trap_w, g.trap_h = nodes_w, nodes_h, npm, trap_w, trap_h
    return g


def cells_array(cells: np.ndarray):
    """cells: (nc, 5) float64 array of x, y, angle, length, width."""
    cells = np.ascontiguousarray(cells, dtype=np.float64)
    assert cells.ndim == 2 and cells.shape[1] == 5
    return cells, cells.ctypes.data_as(C.POINTER(Cell))


def cell_window(g: Grid, cell) -> tuple:
    arr, ptr = cells_array(np.asarray(cell, dtype=np.float64).reshape(1, 5))
    w = np.empty(4, np.int32)
    lib().orc_cell_window(C.byref(g), ptr, w.ctypes.data_as(C.POINTER(C.c_int32)))
    return tuple(int(v) for v in w)


def cell_coverage(g: Grid, cell) -> np.ndarray:
    arr, ptr = cells_array(np.asarray(cell, dtype=np.float64).reshape(1, 5))
    cap = g.nodes_w * g.nodes_h + 1
    ij = np.empty(2 * cap, np.int32)
    n = lib().orc_cell_coverage(C.byref(g), ptr, ij.ctypes.data_as(C.POINTER(C.c_int32)), cap)
    return ij[: 2 * n].reshape(-1, 2).copy()


def couple(g: Grid, cells: np.ndarray, fields, q: np.ndarray):
    """Sequential updateCells loop; fields are modified copies. Returns
    (new_fields, sampled (nc, ns), counts (nc,))."""
    arr, ptr = cells_array(cells)
    nc = arr.shape[0]
    fs = [np.array(f, dtype=np.float64, copy=True) for f in fields]
    ns = len(fs)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(nc, ns)
    sampled = np.empty((nc, ns))
    counts = np.empty(nc, np.int32)
    ptrs = (C.POINTER(C.c_double) * ns)(*[_dp(f) for f in fs])
    lib().orc_couple(C.byref(g), ptr, nc, ns, ptrs, _dp(q), _dp(sampled),
                     counts.ctypes.data_as(C.POINTER(C.c_int32)))
    return fs, sampled, counts


def boundary_flux(p: Problem, u) -> float:
    return lib().orc_boundary_flux(C.byref(p), _dp(np.ascontiguousarray(u, dtype=np.float64)))


def boundary_well(w, flux, well_scaling, decay_rate, dt):
    """One computeBoundaryWell update; returns (w_new, underflowed)."""
    n = C.c_int64(0)
    w2 = lib().orc_boundary_well(w, flux, well_scaling, decay_rate, dt, C.byref(n))
    return w2, n.value


def cell_tensor(g: Grid, cells: np.ndarray, Dx: float, Dy: float):
    """setDiffusionTensor over the cell list: (D11, D22, D12) in natural order."""
    arr, ptr = cells_array(cells)
    n = g.nodes_w * g.nodes_h
    t = [np.empty(n) for _ in range(3)]
    lib().orc_cell_tensor(C.byref(g), ptr, arr.shape[0], Dx, Dy, _dp(t[0]), _dp(t[1]), _dp(t[2]))
    return t


def couple_affine(g: Grid, cells: np.ndarray, fields, qa: np.ndarray, qb: np.ndarray):
    """Sequential updateCells loop with q = a + b * (the cell's own sample).
    Returns (new_fields, sampled (nc, ns), counts (nc,))."""
    arr, ptr = cells_array(cells)
    nc = arr.shape[0]
    fs = [np.array(f, dtype=np.float64, copy=True) for f in fields]
    ns = len(fs)
    qa = np.ascontiguousarray(qa, dtype=np.float64).reshape(nc, ns)
    qb = np.ascontiguousarray(qb, dtype=np.float64).reshape(nc, ns)
    sampled = np.empty((nc, ns))
    counts = np.empty(nc, np.int32)
    ptrs = (C.POINTER(C.c_double) * ns)(*[_dp(f) for f in fs])
    lib().orc_couple_affine(C.byref(g), ptr, nc, ns, ptrs, _dp(qa), _dp(qb), _dp(sampled),
                            counts.ctypes.data_as(C.POINTER(C.c_int32)))
    return fs, sampled, counts
