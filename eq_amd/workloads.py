"""Synthetic, seeded workloads for the quorum-sensing field step.

Every random number comes from SplitMix64 (Steele, Lea & Flood, "Fast
splittable pseudorandom number generators", OOPSLA 2014): state += golden
gamma, then the two xor-shift-multiply rounds of the published finaliser.
Uniform doubles are (z >> 11) * 2^-53.  The reference's own generator
(std::default_random_engine seeded from the clock, src/eQrandom.h:13-44) is
not reproducible and is not reused.

Per-step cell motion is a counter-based random walk: the displacement of cell
c at step k along axis a is sigma * (sum of 12 uniforms - 6) (Irwin-Hall, unit
variance), each uniform being the top 32 bits of SplitMix64 of a counter
derived from (seed, k, c, a, term).  With 32-bit uniforms the sum is exact in
FP64; the new position is x + sigma*g (two roundings, no fused multiply-add)
clamped to the trap.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


class SplitMix64:
    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = np.arange(1, n + 1, dtype=np.uint64)
            states = self.state + idx * GOLDEN
            self.state = self.state + np.uint64(n) * GOLDEN
        return _mix(states)

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
        return lo + (hi - lo) * u


def random_walk_step(cells: np.ndarray, seed: int, step: int, sigma: float, width: float,
                     height: float) -> np.ndarray:
    """Return a copy of cells (nc,5) with x,y advanced by one random-walk step."""
    nc = cells.shape[0]
    out = np.array(cells, dtype=np.float64, copy=True)
    c = np.arange(nc, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _mix(np.array([seed], dtype=np.uint64) ^ (np.uint64(step) * GOLDEN))[0]
    for axis, lim, col in ((0, width, 0), (1, height, 1)):
        acc = np.zeros(nc, dtype=np.float64)
        for term in range(12):
            with np.errstate(over="ignore"):
                ctr = key ^ (c * np.uint64(0xD1B54A32D192ED03)) ^ np.uint64((axis << 8) | term)
            u32 = (_mix(ctr) >> np.uint64(32)).astype(np.float64) * (2.0 ** -32)
            acc += u32  # exact: 12 multiples of 2^-32 below 12
        g = acc - 6.0
        out[:, col] = np.clip(out[:, col] + sigma * g, 0.0, lim)
    return out


@dataclass
class Trap:
    nx: int
    ny: int
    width: float
    height: float
    npm: float
    D: list
    gamma: list
    cells: np.ndarray           # (nc, 5)
    rates: np.ndarray           # (nc, nspecies) production rate per species
    u0: list = field(default_factory=list)  # initial fields (natural order)

    @property
    def nodes(self) -> int:
        return (self.nx + 1) * (self.ny + 1)


def seeded_cells(rng: SplitMix64, nc: int, width: float, height: float, nspecies: int,
                 length=(2.0, 4.0), cell_width: float = 1.0, rate=(0.5, 1.5)):
    """Rod cells: centre uniform in the trap, angle in [0,pi), length in [2,4],
    width 1, production rate per species in [0.5,1.5].  Draw order per cell:
    x, y, angle, length, rate_0 .. rate_{S-1}."""
    draws = rng.uniform(nc * (4 + nspecies)).reshape(nc, 4 + nspecies)
    cells = np.empty((nc, 5))
    cells[:, 0] = draws[:, 0] * width
    cells[:, 1] = draws[:, 1] * height
    cells[:, 2] = draws[:, 2] * math.pi
    cells[:, 3] = length[0] + (length[1] - length[0]) * draws[:, 3]
    cells[:, 4] = cell_width
    rates = rate[0] + (rate[1] - rate[0]) * draws[:, 4:]
    return cells, rates


TOP_DIRICHLET = {2: (1, 0.0, 0.0, 0.0)}  # wall TOP: Dirichlet 0, others Neumann


def config0() -> Trap:
    """40x20 um trap, 160x80 cells (h=0.25), 2 species D=400,100, 200 static rods, seed 0."""
    rng = SplitMix64(0)
    cells, rates = seeded_cells(rng, 200, 40.0, 20.0, 2)
    return Trap(160, 80, 40.0, 20.0, 4.0, [400.0, 100.0], [0.01, 0.01], cells, rates,
                [np.zeros(161 * 81), np.zeros(161 * 81)])


def config1(u0_seed: int | None = None) -> Trap:
    """100x50 um trap, 2000x1000 cells (h=0.05), 2 species, 5000 rods seed 0
    (random walk sigma=0.05 um per step).  u0_seed=2 gives the config-2 start
    fields uniform in [0,1]."""
    rng = SplitMix64(0)
    cells, rates = seeded_cells(rng, 5000, 100.0, 50.0, 2)
    n = 2001 * 1001
    if u0_seed is None:
        u0 = [np.zeros(n), np.zeros(n)]
    else:
        r2 = SplitMix64(u0_seed)
        u0 = [r2.uniform(n), r2.uniform(n)]
    return Trap(2000, 1000, 100.0, 50.0, 20.0, [400.0, 100.0], [0.01, 0.01], cells, rates, u0)


def config3_traps(ntraps: int = 512, seed: int = 3):
    """Uneven batched traps at h=0.1: W~U[20,100], H~U[10,50] rounded to 0.1 um,
    cells U{50..3000}, D1~U[200,600], D2~U[50,150], gamma log-uniform [1e-3,1e-1]."""
    rng = SplitMix64(seed)
    traps = []
    for _ in range(ntraps):
        w, h, nc_u, d1, d2, lg = rng.uniform(6)
        nx = int(round((20.0 + 80.0 * w) * 10.0))
        ny = int(round((10.0 + 40.0 * h) * 10.0))
        width, height = nx / 10.0, ny / 10.0
        nc = 50 + int(math.floor(nc_u * 2951))
        gamma = 10.0 ** (-3.0 + 2.0 * lg)
        cells, rates = seeded_cells(rng, nc, width, height, 2)
        traps.append(Trap(nx, ny, width, height, 10.0, [200.0 + 400.0 * d1, 50.0 + 100.0 * d2],
                          [gamma, gamma], cells, rates,
                          [np.zeros((nx + 1) * (ny + 1)), np.zeros((nx + 1) * (ny + 1))]))
    return traps


def config4() -> Trap:
    """204.8x102.4 um trap, 8192x4096 cells (h=0.025, 33 566 721 nodes/species),
    4 species D=400,250,150,100, gamma 0.01, 100 000 static rods seed 4; the
    fields start at zero (u0 left empty: nothing to upload)."""
    rng = SplitMix64(4)
    cells, rates = seeded_cells(rng, 100000, 204.8, 102.4, 4)
    return Trap(8192, 4096, 204.8, 102.4, 40.0, [400.0, 250.0, 150.0, 100.0], [0.01] * 4, cells, rates, [])


def channel_velocity(ny: int, height: float, vmax: float = 50.0) -> np.ndarray:
    """Parabolic x-velocity per grid row, zero at both walls, vmax mid-height."""
    y = np.arange(ny + 1) * (height / ny)
    return 4.0 * vmax * y * (height - y) / (height * height)


def descs_for(trap: Trap, dt: float = 0.1, walls=None):
    """eq_field_desc-style tuples (nx, ny, W, H, dt, D, gamma, walls) per species."""
    walls = TOP_DIRICHLET if walls is None else walls
    return [(trap.nx, trap.ny, trap.width, trap.height, dt, D, g, walls)
            for D, g in zip(trap.D, trap.gamma)]
