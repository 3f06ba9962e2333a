"""Python binding of the C ABI in include/eq_field.h (libeq_field.so).

This is the host-side mirror of eQ's ``eQ::diffusionSolver`` /
``fenicsInterface`` surface (src/eQ.h:302-330, src/fHSL.h:333-449) for Python
callers: ``FieldBatch`` holds one problem per (trap, species) and forwards
every call to the HIP library.  There is no CPU fallback: constructing a
batch without the built library or without a HIP device raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libeq_field.so")

SCHEME_P1, SCHEME_FD = 0, 1
WALLS_GRID, WALLS_COORD = 0, 1
NEUMANN, DIRICHLET, ROBIN = 0, 1, 2
LEFT, RIGHT, TOP, BOTTOM = 0, 1, 2, 3
JACOBI_CG, GMG_CG, MIXED_GMG_CG = 0, 1, 2
GMG_BICGSTAB, MIXED_GMG_BICGSTAB = 3, 4
IR_GMG_CG = 5
SAMPLE, DEPOSIT, TENSOR = 1, 2, 4
KERNEL_STENCIL64, KERNEL_SMOOTH, KERNEL_CG_UPDATE, KERNEL_COUPLE = 0, 1, 2, 3
KERNEL_SMOOTH_COARSE, KERNEL_TRANSFER0, KERNEL_TRANSFER_COARSE, KERNEL_BOTTOM = 4, 5, 6, 7
KERNEL_FINISH, KERNEL_SOLVE_ONCE, KERNEL_BICG_VEC, KERNEL_COUPLE_SETUP, KERNEL_BOUNDARY = 8, 9, 10, 11, 12
KERNEL_COUNT = 13
KERNEL_FAMILIES = {
    KERNEL_STENCIL64: "Krylov operator apply (k_ap_dir, k_bicg_apply)",
    KERNEL_SMOOTH: "finest-level way up: correction and sweeps (k_level_up; k_smooth, k_presmooth2, "
                   "k_prolong_smooth on level 0)",
    KERNEL_CG_UPDATE: "Krylov x, r update + dots (k_update, k_bicg_x)",
    KERNEL_COUPLE: "cell sample/deposit per tile (k_tile_couple)",
    KERNEL_SMOOTH_COARSE: "coarse-level way up: correction and sweeps (k_level_up; k_smooth, k_presmooth2, "
                          "k_prolong_smooth)",
    KERNEL_TRANSFER0: "finest-level way down: sweep, residual, restriction (k_level_down; k_residual, k_restrict, "
                      "k_residual_restrict, k_prolong on level 0)",
    KERNEL_TRANSFER_COARSE: "coarse-level way down: sweeps, residual, restriction (k_level_down; k_residual, "
                            "k_restrict, k_residual_restrict)",
    KERNEL_BOTTOM: "LDS bottom V-cycle (k_bottom, k_cast_in, k_cast_out)",
    KERNEL_FINISH: "per-problem reduction finish (k_finish)",
    KERNEL_SOLVE_ONCE: "once per solve (k_rhs, k_init, k_true_residual, k_keep, first sweep, k_copy)",
    KERNEL_BICG_VEC: "BiCGStab vector updates (k_bicg_p, k_bicg_s)",
    KERNEL_COUPLE_SETUP: "cell geometry, pair scan and sort, tile bounds, per-cell sums",
    KERNEL_BOUNDARY: "channels, boundary wells and values, flux",
}

STATUS = {0: "EQ_OK", 1: "EQ_ERR_INVALID", 2: "EQ_ERR_HIP", 3: "EQ_ERR_NOMEM",
          4: "EQ_ERR_NOT_CONVERGED", 5: "EQ_ERR_NO_DEVICE"}

# every symbol include/eq_field.h declares
EXPORTED = [
    "eq_last_error", "eq_solver_default", "eq_walls_found",
    "eq_batch_create", "eq_batch_destroy", "eq_batch_num_problems", "eq_batch_num_nodes",
    "eq_batch_stream", "eq_batch_sync", "eq_batch_upload", "eq_batch_download",
    "eq_batch_upload_dev", "eq_batch_download_dev", "eq_batch_step", "eq_batch_problem_stats",
    "eq_batch_couple", "eq_batch_couple_dev", "eq_batch_couple_traps_dev", "eq_batch_couple_traps_dev_noalloc",
    "eq_batch_couple_affine_dev",
    "eq_batch_autotune",
    "eq_batch_set_tensor", "eq_field_set_tensor", "eq_batch_cell_tensor", "eq_batch_tensor_values", "eq_batch_set_velocity", "eq_batch_set_wall_values", "eq_batch_set_channels",
    "eq_batch_channel_values",
    "eq_batch_boundary_flux", "eq_batch_set_boundary_value", "eq_batch_set_well", "eq_batch_well_state",
    "eq_batch_vertex_map",
    "eq_field_create", "eq_field_destroy", "eq_field_upload", "eq_field_download",
    "eq_field_step", "eq_field_vertex_map", "eq_field_boundary_flux", "eq_cells_couple",
    "eq_batch_profile", "eq_batch_profile_read", "eq_batch_traffic", "eq_batch_apply", "eq_batch_rhs",
]


class Bc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad_", C.c_int32), ("value", C.c_double),
                ("r", C.c_double), ("s", C.c_double)]


class FieldDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("width", C.c_double),
                ("height", C.c_double), ("dt", C.c_double), ("D", C.c_double),
                ("gamma", C.c_double), ("f", C.c_double), ("wall", Bc * 4), ("scheme", C.c_int32),
                ("marking", C.c_int32)]


class ChannelDesc(C.Structure):
    _fields_ = [("v", C.c_double), ("r_left", C.c_double), ("s_left", C.c_double),
                ("r_right", C.c_double), ("s_right", C.c_double), ("well_scaling", C.c_double),
                ("substeps", C.c_int32), ("pad_", C.c_int32)]


def channel_desc(v, r_left, s_left, r_right, s_right, well_scaling, substeps=1) -> ChannelDesc:
    c = ChannelDesc()
    c.v, c.r_left, c.s_left, c.r_right, c.s_right = v, r_left, s_left, r_right, s_right
    c.well_scaling, c.substeps = well_scaling, substeps
    return c


class WellDesc(C.Structure):
    _fields_ = [("well_scaling", C.c_double), ("decay_rate", C.c_double), ("w0", C.c_double)]


class SolverDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("max_iter", C.c_int32), ("rtol", C.c_double),
                ("mg_pre", C.c_int32), ("mg_post", C.c_int32), ("mg_omega", C.c_double),
                ("mg_coarse_max_nodes", C.c_int32), ("mg_max_levels", C.c_int32),
                ("check_every", C.c_int32), ("use_graphs", C.c_int32),
                ("mg_fuse_max_nodes", C.c_int32), ("mg_fuse_pre_max_nodes", C.c_int32),
                ("mg_bottom_lds_bytes", C.c_int32), ("mg_fuse_pre_top", C.c_int32),
                ("tile4_min_nodes", C.c_int32), ("tile8_min_nodes", C.c_int32),
                ("extrapolate", C.c_int32), ("ir_max_outer", C.c_int32), ("ir_inner_rtol", C.c_double),
                ("mg_level_fused", C.c_int32), ("mg_omega2", C.c_double),
                ("streams", C.c_int32)]


class StepStats(C.Structure):
    _fields_ = [("iterations_max", C.c_int32), ("nconverged", C.c_int32),
                ("relres_max", C.c_double), ("seconds", C.c_double)]


class Cell(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("angle", C.c_double),
                ("length", C.c_double), ("width", C.c_double)]


class EqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load_library():
    """Load libeq_field.so; raises if it has not been built.

    A process that also uses torch on the GPU must initialise torch first:
    torch ships its own HIP runtime under the same soname (libamdhip64.so.7)
    and whichever is loaded first serves the process."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m eq_amd.build` (the HIP path "
                           "has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp, dp, ip32, ip64 = C.c_void_p, P(C.c_double), P(C.c_int32), P(C.c_int64)
    L.eq_last_error.restype = C.c_char_p
    L.eq_solver_default.argtypes = [P(SolverDesc)]
    L.eq_walls_found.argtypes = [P(FieldDesc), ip32]
    L.eq_batch_create.argtypes = [P(FieldDesc), C.c_int32, C.c_int32, P(SolverDesc), P(vp)]
    L.eq_batch_destroy.argtypes = [vp]
    L.eq_batch_num_problems.argtypes = [vp]
    L.eq_batch_num_problems.restype = C.c_int32
    L.eq_batch_num_nodes.argtypes = [vp, C.c_int32]
    L.eq_batch_num_nodes.restype = C.c_int64
    L.eq_batch_stream.argtypes = [vp]
    L.eq_batch_stream.restype = vp
    L.eq_batch_sync.argtypes = [vp]
    for f in ("eq_batch_upload", "eq_batch_download", "eq_batch_upload_dev", "eq_batch_download_dev"):
        getattr(L, f).argtypes = [vp, C.c_int32, vp, C.c_size_t]
    L.eq_batch_step.argtypes = [vp, P(StepStats)]
    L.eq_batch_problem_stats.argtypes = [vp, ip32, dp]
    L.eq_batch_couple.argtypes = [vp, C.c_int32, vp, C.c_size_t, C.c_double, vp, vp, vp, C.c_uint32]
    L.eq_batch_autotune.argtypes = [vp, C.c_int32, P(SolverDesc)]
    L.eq_batch_set_tensor.argtypes = [vp, C.c_int32, vp, vp, vp, C.c_size_t]
    L.eq_batch_set_velocity.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    L.eq_batch_cell_tensor.argtypes = [vp, C.c_int32, C.c_double, C.c_double, C.c_int32]
    L.eq_batch_tensor_values.argtypes = [vp, C.c_int32, vp, vp, vp, C.c_size_t]
    L.eq_batch_set_wall_values.argtypes = [vp, C.c_int32, C.c_int32, vp, C.c_size_t]
    L.eq_batch_set_channels.argtypes = [vp, C.c_int32, P(ChannelDesc), P(ChannelDesc)]
    L.eq_batch_channel_values.argtypes = [vp, C.c_int32, vp, vp]
    L.eq_field_set_tensor.argtypes = [vp, vp, vp, vp, C.c_size_t]
    L.eq_batch_couple_traps_dev.argtypes = [vp, C.c_int32, C.c_int32, P(C.c_int64), vp, C.c_double, vp, vp, vp,
                                            C.c_uint32]
    L.eq_batch_couple_dev.argtypes = [vp, C.c_int32, vp, C.c_size_t, C.c_double, vp, vp, vp,
                                      C.c_uint32]
    L.eq_batch_couple_traps_dev_noalloc.argtypes = L.eq_batch_couple_traps_dev.argtypes
    L.eq_batch_couple_affine_dev.argtypes = [vp, C.c_int32, C.c_int32, P(C.c_int64), vp, C.c_double, vp, vp, vp, vp]
    L.eq_batch_boundary_flux.argtypes = [vp, C.c_int32, dp]
    L.eq_batch_set_boundary_value.argtypes = [vp, C.c_int32, C.c_double]
    L.eq_batch_set_well.argtypes = [vp, C.c_int32, P(WellDesc)]
    L.eq_batch_well_state.argtypes = [vp, C.c_int32, dp, dp, ip64]
    L.eq_batch_vertex_map.argtypes = [vp, C.c_int32, dp, ip32]
    L.eq_batch_profile.argtypes = [vp, C.c_int32, C.c_int32]
    L.eq_batch_profile_read.argtypes = [vp, C.c_int32, ip64, dp, dp]
    L.eq_batch_traffic.argtypes = [vp, dp, ip64, C.c_int32]
    L.eq_batch_apply.argtypes = [vp, C.c_int32, vp, vp, C.c_size_t]
    L.eq_batch_rhs.argtypes = [vp, C.c_int32, vp, vp, C.c_size_t]
    L.eq_field_create.argtypes = [P(FieldDesc), P(SolverDesc), P(vp)]
    L.eq_field_destroy.argtypes = [vp]
    for f in ("eq_field_upload", "eq_field_download"):
        getattr(L, f).argtypes = [vp, vp, C.c_size_t]
    L.eq_field_step.argtypes = [vp, P(StepStats)]
    L.eq_field_vertex_map.argtypes = [vp, dp, ip32]
    L.eq_field_boundary_flux.argtypes = [vp, dp]
    L.eq_cells_couple.argtypes = [vp, vp, C.c_size_t, C.c_double, vp, vp, vp, C.c_uint32]
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise EqError(rc, load_library().eq_last_error().decode())


def field_desc(nx, ny, width, height, dt, D, gamma=0.0, f=0.0, walls=None, scheme=0,
               marking=WALLS_GRID) -> FieldDesc:
    """walls: {wall: (kind, value, r, s)}; missing walls are Neumann.
    scheme: SCHEME_P1 (fenicsInterface's P1 form) or SCHEME_FD (diffusionPETSc).
    marking: WALLS_GRID (outer grid rows/columns) or WALLS_COORD (the
    reference's near() test on DOLFIN's computed vertex coordinates)."""
    d = FieldDesc()
    d.nx, d.ny, d.width, d.height = nx, ny, width, height
    d.dt, d.D, d.gamma, d.f = dt, D, gamma, f
    d.scheme = scheme
    d.marking = marking
    for w in range(4):
        kind, value, r, s = (walls or {}).get(w, (NEUMANN, 0.0, 0.0, 0.0))
        d.wall[w].kind, d.wall[w].value, d.wall[w].r, d.wall[w].s = kind, value, r, s
    return d


def walls_found(desc: FieldDesc) -> list:
    """Per wall (left, right, top, bottom): found by the reference's coordinate test."""
    f = np.empty(4, np.int32)
    _check(load_library().eq_walls_found(C.byref(desc), f.ctypes.data_as(C.POINTER(C.c_int32))))
    return [bool(v) for v in f]


def solver_desc(kind=MIXED_GMG_CG, **kw) -> SolverDesc:
    s = SolverDesc()
    load_library().eq_solver_default(C.byref(s))
    s.kind = kind
    for k, v in kw.items():
        setattr(s, k, v)
    return s


@dataclass
class Stats:
    iterations_max: int
    nconverged: int
    relres_max: float
    seconds: float


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


class FieldBatch:
    """Problems (trap x species) advanced together; problem p = trap*nspecies + species."""

    def __init__(self, descs, nspecies: int = 1, solver: SolverDesc | None = None):
        L = load_library()
        arr = (FieldDesc * len(descs))(*descs)
        h = C.c_void_p()
        _check(L.eq_batch_create(arr, len(descs), nspecies, C.byref(solver) if solver else None,
                                 C.byref(h)))
        self._h = h
        self.nprob = len(descs)
        self.nspecies = nspecies
        self.descs = list(descs)

    def close(self):
        if getattr(self, "_h", None):
            load_library().eq_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def num_nodes(self, prob: int) -> int:
        return int(load_library().eq_batch_num_nodes(self._h, prob))

    def stream(self) -> int:
        return int(load_library().eq_batch_stream(self._h) or 0)

    def sync(self):
        _check(load_library().eq_batch_sync(self._h))

    def upload(self, prob: int, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        _check(load_library().eq_batch_upload(self._h, prob, _ptr(u), u.size))

    def download(self, prob: int) -> np.ndarray:
        u = np.empty(self.num_nodes(prob))
        _check(load_library().eq_batch_download(self._h, prob, _ptr(u), u.size))
        return u

    def upload_dev(self, prob: int, dev_ptr: int, n: int):
        _check(load_library().eq_batch_upload_dev(self._h, prob, C.c_void_p(dev_ptr), n))

    def download_dev(self, prob: int, dev_ptr: int, n: int):
        _check(load_library().eq_batch_download_dev(self._h, prob, C.c_void_p(dev_ptr), n))

    def step(self, check: bool = True) -> Stats:
        st = StepStats()
        rc = load_library().eq_batch_step(self._h, C.byref(st))
        if check:
            _check(rc)
        return Stats(st.iterations_max, st.nconverged, st.relres_max, st.seconds)

    def problem_stats(self):
        it = np.empty(self.nprob, np.int32)
        rr = np.empty(self.nprob)
        _check(load_library().eq_batch_problem_stats(
            self._h, it.ctypes.data_as(C.POINTER(C.c_int32)), rr.ctypes.data_as(C.POINTER(C.c_double))))
        return it, rr

    def couple(self, trap: int, cells: np.ndarray, npm: float, q: np.ndarray,
               flags: int = SAMPLE | DEPOSIT):
        """cells: (nc,5) x,y,angle,length,width; q: (nc, nspecies). Returns (sampled, counts)."""
        cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 5)
        nc = cells.shape[0]
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(nc, self.nspecies)
        sampled = np.empty((nc, self.nspecies))
        counts = np.empty(nc, np.int32)
        _check(load_library().eq_batch_couple(self._h, trap, _ptr(cells), nc, npm, _ptr(q),
                                              _ptr(sampled), _ptr(counts), flags))
        return sampled, counts

    def couple_dev(self, trap: int, cells_ptr: int, nc: int, npm: float, q_ptr: int,
                   sampled_ptr: int = 0, counts_ptr: int = 0, flags: int = SAMPLE | DEPOSIT):
        _check(load_library().eq_batch_couple_dev(
            self._h, trap, C.c_void_p(cells_ptr), nc, npm, C.c_void_p(q_ptr),
            C.c_void_p(sampled_ptr) if sampled_ptr else None,
            C.c_void_p(counts_ptr) if counts_ptr else None, flags))

    def couple_traps_dev(self, trap0: int, cell_offsets, cells_ptr: int, npm: float, q_ptr: int,
                         sampled_ptr: int = 0, counts_ptr: int = 0, flags: int = SAMPLE | DEPOSIT):
        """Traps trap0 .. trap0+len(cell_offsets)-2 in one pass; cell_offsets (host) gives
        the first cell of each trap in the concatenated device arrays, and the total."""
        off = np.ascontiguousarray(cell_offsets, dtype=np.int64)
        _check(load_library().eq_batch_couple_traps_dev(
            self._h, trap0, off.size - 1, off.ctypes.data_as(C.POINTER(C.c_int64)), C.c_void_p(cells_ptr), npm,
            C.c_void_p(q_ptr), C.c_void_p(sampled_ptr) if sampled_ptr else None,
            C.c_void_p(counts_ptr) if counts_ptr else None, flags))

    def couple_traps_dev_noalloc(self, trap0: int, cell_offsets, cells_ptr: int, npm: float, q_ptr: int,
                                 sampled_ptr: int = 0, counts_ptr: int = 0, flags: int = SAMPLE | DEPOSIT):
        """couple_traps_dev with the fixed per-device scratch (no allocation)."""
        off = np.ascontiguousarray(cell_offsets, dtype=np.int64)
        _check(load_library().eq_batch_couple_traps_dev_noalloc(
            self._h, trap0, off.size - 1, off.ctypes.data_as(C.POINTER(C.c_int64)), C.c_void_p(cells_ptr), npm,
            C.c_void_p(q_ptr), C.c_void_p(sampled_ptr) if sampled_ptr else None,
            C.c_void_p(counts_ptr) if counts_ptr else None, flags))

    def couple_affine_dev(self, trap0: int, cell_offsets, cells_ptr: int, npm: float, a_ptr: int, b_ptr: int,
                          sampled_ptr: int = 0, counts_ptr: int = 0):
        """Coupling with q = a + b * (the cell's own sample), in list order."""
        off = np.ascontiguousarray(cell_offsets, dtype=np.int64)
        _check(load_library().eq_batch_couple_affine_dev(
            self._h, trap0, off.size - 1, off.ctypes.data_as(C.POINTER(C.c_int64)), C.c_void_p(cells_ptr), npm,
            C.c_void_p(a_ptr), C.c_void_p(b_ptr), C.c_void_p(sampled_ptr) if sampled_ptr else None,
            C.c_void_p(counts_ptr) if counts_ptr else None))

    def set_wall_values(self, prob: int, wall: int, values=None):
        """Per-node Dirichlet values along a Dirichlet wall, or the wall's constant when None."""
        if values is None:
            _check(load_library().eq_batch_set_wall_values(self._h, prob, wall, None, 0))
            return
        a = np.ascontiguousarray(values, dtype=np.float64)
        _check(load_library().eq_batch_set_wall_values(self._h, prob, wall, _ptr(a), a.size))

    def set_channels(self, prob: int, top=None, bottom=None):
        """Flow channels (ChannelDesc or None) above/below the trap of problem prob."""
        _check(load_library().eq_batch_set_channels(self._h, prob, C.byref(top) if top is not None else None,
                                                    C.byref(bottom) if bottom is not None else None))

    def channel_values(self, prob: int):
        nx = self.descs[prob].nx
        t, b = np.empty(nx + 1), np.empty(nx + 1)
        _check(load_library().eq_batch_channel_values(self._h, prob, _ptr(t), _ptr(b)))
        return t, b

    def set_velocity(self, prob: int, vx_rows=None):
        """x-velocity per grid row (ny+1 values) for advection, or None."""
        if vx_rows is None:
            _check(load_library().eq_batch_set_velocity(self._h, prob, None, 0))
            return
        a = np.ascontiguousarray(vx_rows, dtype=np.float64)
        _check(load_library().eq_batch_set_velocity(self._h, prob, _ptr(a), a.size))

    def set_tensor(self, prob: int, D11=None, D22=None, D12=None):
        """Per-node diffusion tensor (natural order), or constant coefficients when None."""
        if D11 is None:
            _check(load_library().eq_batch_set_tensor(self._h, prob, None, None, None, 0))
            return
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (D11, D22, D12)]
        _check(load_library().eq_batch_set_tensor(self._h, prob, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), a[0].size))

    def cell_tensor(self, trap: int, Dx: float = 1.0, Dy: float = 1.0, enable: bool = True):
        """Set the trap's diffusion tensor from the cells' angles in every coupling
        call carrying TENSOR (setDiffusionTensor), or back to constant coefficients."""
        _check(load_library().eq_batch_cell_tensor(self._h, trap, Dx, Dy, 1 if enable else 0))

    def tensor_values(self, prob: int):
        """(D11, D22, D12) of the finest level in natural order."""
        n = self.num_nodes(prob)
        t = [np.empty(n) for _ in range(3)]
        _check(load_library().eq_batch_tensor_values(self._h, prob, _ptr(t[0]), _ptr(t[1]), _ptr(t[2]), n))
        return t

    def autotune(self, repeats: int = 1) -> SolverDesc:
        """Pick the fastest multigrid cycle shape on the current state (which is restored)."""
        sd = SolverDesc()
        _check(load_library().eq_batch_autotune(self._h, repeats, C.byref(sd)))
        return sd

    def boundary_flux(self, prob: int) -> float:
        v = C.c_double()
        _check(load_library().eq_batch_boundary_flux(self._h, prob, C.byref(v)))
        return v.value

    def set_boundary_value(self, prob: int, value: float):
        """setBoundaryValues: every Dirichlet wall of `prob` takes `value` (stream ordered)."""
        _check(load_library().eq_batch_set_boundary_value(self._h, prob, value))

    def set_well(self, prob: int, well_scaling=None, decay_rate=0.0, w0=0.0):
        """Boundary well of the DIRICHLET_UPDATE boundary, updated after every step;
        well_scaling None removes it."""
        if well_scaling is None:
            _check(load_library().eq_batch_set_well(self._h, prob, None))
            return
        w = WellDesc(well_scaling, decay_rate, w0)
        _check(load_library().eq_batch_set_well(self._h, prob, C.byref(w)))

    def well_state(self, prob: int):
        """(w, last flux, underflows) of the problem's boundary well."""
        w, f, n = C.c_double(), C.c_double(), C.c_int64()
        _check(load_library().eq_batch_well_state(self._h, prob, C.byref(w), C.byref(f), C.byref(n)))
        return w.value, f.value, n.value

    def vertex_map(self, prob: int):
        n = self.num_nodes(prob)
        xy = np.empty(2 * n)
        dof = np.empty(n, np.int32)
        _check(load_library().eq_batch_vertex_map(self._h, prob, xy.ctypes.data_as(C.POINTER(C.c_double)),
                                                  dof.ctypes.data_as(C.POINTER(C.c_int32))))
        return xy.reshape(-1, 2), dof

    def apply(self, prob: int, x) -> np.ndarray:
        """y = A x of the Dirichlet-row-replaced operator (natural order)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(load_library().eq_batch_apply(self._h, prob, _ptr(x), _ptr(y), x.size))
        return y

    def rhs(self, prob: int, u0) -> np.ndarray:
        """b = M u0 + sources with Dirichlet rows set to the wall value."""
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        y = np.empty_like(u0)
        _check(load_library().eq_batch_rhs(self._h, prob, _ptr(u0), _ptr(y), u0.size))
        return y

    def profile(self, kernel: int, enable: bool = True):
        _check(load_library().eq_batch_profile(self._h, kernel, 1 if enable else 0))

    def traffic(self, reset: bool = False):
        """(algorithmic bytes, launches) issued since the last reset, graph replays included."""
        by, n = C.c_double(), C.c_int64()
        _check(load_library().eq_batch_traffic(self._h, C.byref(by), C.byref(n), 1 if reset else 0))
        return by.value, n.value

    def profile_read(self, kernel: int):
        n = C.c_int64()
        ms = C.c_double()
        by = C.c_double()
        _check(load_library().eq_batch_profile_read(self._h, kernel, C.byref(n), C.byref(ms), C.byref(by)))
        return n.value, ms.value, by.value
