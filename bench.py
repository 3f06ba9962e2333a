"""Benchmark of the quorum-sensing field step on one MI355X.

Workload (the single-trap production configuration): a 100 x 50 um trap at
h = 0.05 um (2000 x 1000 cells, 2 003 001 nodes per species), two HSL species
(D = 400 and 100 um^2/s, gamma = 0.01/s, dt = 0.1 s), Dirichlet 0 on the top
wall and Neumann elsewhere, 5000 rod cells seeded from seed 0 whose centres
take a seeded random walk (sigma = 0.05 um) every step.  One step = couple
(sample + deposit for every cell and species) + implicit solve of every
species to ||b - A u||/||b|| <= 1e-10 in FP64.

Inputs (cell positions for every step, rates, initial fields) are resident in
HBM before the timed region.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# profiles/traffic.json: per workload ("config1", "config3", "config4") and
# kernel family id, the HBM bytes per launch that rocprofv3's FETCH_SIZE and
# WRITE_SIZE counters give for the family's launches


# solver options every batch of the run takes (bench flags)
SOLVER_OPTS = {}


def solver_desc(F, kind, **kw):
    return F.solver_desc(kind, **{**SOLVER_OPTS, **kw})


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_cells(trap, nsteps, seed=0, sigma=0.05):
    from eq_amd import workloads as W
    out = np.empty((nsteps,) + trap.cells.shape)
    cur = trap.cells
    for k in range(nsteps):
        cur = W.random_walk_step(cur, seed, k, sigma, trap.width, trap.height)
        out[k] = cur
    return out


class Run:
    """One batch (trap x species) with device-resident inputs for `nsteps` steps."""

    def __init__(self, F, trap, solver, cells_dev, q_dev, sampled_dev, dt=0.1, **solver_kw):
        import torch
        from eq_amd import workloads as W
        self.F = F
        self.trap = trap
        descs = [F.field_desc(trap.nx, trap.ny, trap.width, trap.height, dt, D, g, 0.0, W.TOP_DIRICHLET)
                 for D, g in zip(trap.D, trap.gamma)]
        self.fb = F.FieldBatch(descs, nspecies=len(descs), solver=solver_desc(F, solver, **solver_kw))
        for s, u in enumerate(trap.u0):
            self.fb.upload(s, u)
        self.cells = cells_dev
        self.q = q_dev
        self.sampled = sampled_dev
        self.nc = trap.cells.shape[0]
        self.iters = []
        self.relres = []

    def step(self, k, check=True):
        if not 0 <= k < self.cells.shape[0]:
            raise IndexError(f"cell positions exist for {self.cells.shape[0]} steps, not step {k}")
        stride = self.nc * 5 * 8
        self.fb.couple_dev(0, self.cells.data_ptr() + k * stride, self.nc, self.trap.npm, self.q.data_ptr(),
                           self.sampled.data_ptr(), 0, self.F.SAMPLE | self.F.DEPOSIT)
        st = self.fb.step(check=check)
        self.iters.append(st.iterations_max)
        self.relres.append(st.relres_max)
        return st


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(trap, gpu_jacobi_iters, budget_iters=(40, 240)):
    """Time the oracle (C++ restatement of the reference step) on a bounded
    sample of the same workload, in the reference's decomposition: the cell
    loop on one core, then one core per species solving concurrently
    (src/simulation.cpp:38-41), and scale it to one full step."""
    import threading
    from oracle import oracle as O
    from eq_amd import workloads as W
    grid = O.grid(trap.nx + 1, trap.ny + 1, trap.npm, trap.width, trap.height)
    cells0 = W.random_walk_step(trap.cells, 0, 0, 0.05, trap.width, trap.height)
    q = trap.rates * 0.1
    t0 = time.perf_counter()
    fields, _, _ = O.couple(grid, cells0, trap.u0, q)
    t_couple = time.perf_counter() - t0
    ns = len(trap.D)
    parts = [None] * ns

    def species(s):
        # the oracle's C calls release the GIL: the species threads run on
        # their own cores, as the reference's species ranks do
        p = O.problem(trap.nx, trap.ny, trap.width, trap.height, 0.1, trap.D[s], trap.gamma[s], 0.0,
                      walls=W.TOP_DIRICHLET)
        ts = []
        for k in budget_iters:
            t1 = time.perf_counter()
            O.step(p, fields[s], "pcg", 1e-10, maxit=k)
            ts.append(time.perf_counter() - t1)
        per_it = (ts[1] - ts[0]) / (budget_iters[1] - budget_iters[0])
        fixed = ts[0] - budget_iters[0] * per_it
        need = int(gpu_jacobi_iters[s])
        parts[s] = (fixed, per_it, need, fixed + need * per_it)

    threads = [threading.Thread(target=species, args=(s,)) for s in range(ns)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    t_solve = max(pt[3] for pt in parts)
    total = t_couple + t_solve
    nodes = trap.nodes * ns
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count()
    sample = (f"oracle reference step of step 0 (reassembly from the FFC element tensors, Dirichlet rows, "
              f"Jacobi-PCG to 1e-10, sequential cell loop over {trap.cells.shape[0]} cells): cell loop "
              f"{t_couple:.2f}s timed in full on 1 core; then {ns} species concurrently, one thread each, each "
              f"timing its assembly and {budget_iters[0]} and {budget_iters[1]} PCG iterations, scaled to the "
              f"{', '.join(str(p[2]) for p in parts)} iterations the same Jacobi-PCG needs (GPU FP64 Jacobi-CG "
              f"count for this solve); step = cell loop + slowest species. Direct LU at 2M DOFs (the reference's "
              f"solver) is out of a 30 s budget")
    return {"value": nodes / total, "unit": "DOF-timesteps/s", "cores": ns, "kind": "port",
            "sample": sample, "seconds_per_step": total,
            "seconds_per_species": [round(pt[3], 3) for pt in parts], "seconds_cell_loop": t_couple,
            "host": {"cpu_model": _cpu_model(), "cpus_in_affinity": affinity, "os_cpu_count": os.cpu_count()}}


def cpu_baseline_config0(nsteps=100):
    """Config 0 in full on the host: the reference's algorithm per step (the
    sequential cell loop over 200 static rods, then per species reassembly from
    the element tensors, Dirichlet rows and a direct banded LU), in the
    reference's decomposition (src/simulation.cpp:38-41): the cell loop on one
    core, then one thread per species on its own core."""
    import threading
    from oracle import oracle as O
    from eq_amd import workloads as W
    t = W.config0()
    grid = O.grid(t.nx + 1, t.ny + 1, t.npm, t.width, t.height)
    probs = [O.problem(t.nx, t.ny, t.width, t.height, 0.1, D, g, 0.0, walls=W.TOP_DIRICHLET)
             for D, g in zip(t.D, t.gamma)]
    q = t.rates * 0.1
    fields = [u.copy() for u in t.u0]
    t_cells = t_solve = 0.0
    out = [None] * len(probs)

    def species(k):
        out[k] = O.step(probs[k], fields[k], "lu")[0]

    for _ in range(nsteps):
        t0 = time.perf_counter()
        fields, _, _ = O.couple(grid, t.cells, fields, q)
        t1 = time.perf_counter()
        th = [threading.Thread(target=species, args=(k,)) for k in range(len(probs))]
        for x in th:
            x.start()
        for x in th:
            x.join()
        fields = list(out)
        t_solve += time.perf_counter() - t1
        t_cells += t1 - t0
    total = t_cells + t_solve
    return {"value": t.nodes * len(probs) * nsteps / total, "unit": "DOF-timesteps/s", "cores": len(probs),
            "kind": "port", "steps": nsteps, "seconds": total, "seconds_cell_loop": t_cells,
            "seconds_solves": t_solve,
            "sample": f"config 0 in full: 160x80 trap, 2 species, 200 static rods, {nsteps} steps; per step the "
                      f"oracle's sequential cell loop on 1 core, then each species on its own thread: reassembly "
                      f"from the element tensors, Dirichlet rows, direct banded LU (the reference's LU solve)"}, fields


def gpu_config0(F, W, nsteps=100):
    """Config 0 on the GPU through the host entry points (cells and q from
    host memory every step, as the reference's controller hands them over),
    mixed GMG-CG to 1e-10; returns throughput and the fields after nsteps."""
    import torch
    t = W.config0()
    descs = [F.field_desc(t.nx, t.ny, t.width, t.height, 0.1, D, g, 0.0, W.TOP_DIRICHLET)
             for D, g in zip(t.D, t.gamma)]
    q = t.rates * 0.1
    with F.FieldBatch(descs, nspecies=2, solver=solver_desc(F, F.MIXED_GMG_CG)) as fb:
        for s_ in range(2):
            fb.upload(s_, t.u0[s_])
        fb.couple(0, t.cells, t.npm, q)
        fb.step()
        for s_ in range(2):
            fb.upload(s_, t.u0[s_])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            fb.couple(0, t.cells, t.npm, q)
            fb.step()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        fields = [fb.download(s_) for s_ in range(2)]
    return {"value": t.nodes * 2 * nsteps / dt, "unit": "DOF-timesteps/s", "steps": nsteps,
            "ms_per_step": dt / nsteps * 1e3}, fields


def family_profile(F, fb, step_fn, nsteps):
    """Every kernel family of the library timed over nsteps eager steps:
    start/stop HIP events attached to each launch on the batch stream
    (hipExtLaunchKernelGGL), so each duration is the kernel's own execution.
    Returns ({family: record}, eager seconds per step)."""
    import torch
    for k in range(F.KERNEL_COUNT):
        fb.profile(k, True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(nsteps):
        step_fn(k)
    torch.cuda.synchronize()
    eager = (time.perf_counter() - t0) / nsteps
    fams = {}
    for k in range(F.KERNEL_COUNT):
        n, ms, by = fb.profile_read(k)
        fb.profile(k, False)
        if n:
            fams[k] = {"family": F.KERNEL_FAMILIES[k], "launches_per_step": n / nsteps, "ms_per_step": ms / nsteps,
                       "bytes_per_step": by / nsteps, "avg_us": ms / n * 1e3,
                       "achieved_gbs": (by / (ms / 1e3) / 1e9) if ms > 0 else 0.0}
    total = sum(f["ms_per_step"] for f in fams.values())
    for f in fams.values():
        f["share"] = f["ms_per_step"] / total if total > 0 else 0.0
    return fams, eager


def roofline_from(fams, step_bytes, step_s, nsteps_profiled, eager_s, workload="config1"):
    """Roofline record of the family with the largest summed kernel time, with
    the whole step's fraction (all launches' algorithmic bytes over the timed
    step) beside it."""
    dom = max(fams, key=lambda k: fams[k]["ms_per_step"])
    f = fams[dom]
    ach = f["achieved_gbs"]
    step_ach = step_bytes / step_s / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": load_traffic(workload, dom),
            "kernel": f["family"], "family_id": dom, "share_of_kernel_time": f["share"],
            "launches_per_step": f["launches_per_step"], "avg_us": f["avg_us"],
            "bytes_per_launch": f["bytes_per_step"] / f["launches_per_step"],
            "step_frac": step_ach / HBM_PEAK_GBS, "step_achieved": step_ach, "step_bytes": step_bytes,
            "timing": f"start/stop HIP events attached to every launch of every family on the batch stream over "
                      f"{nsteps_profiled} eager steps after the timed region ({eager_s * 1e3:.3f} ms/step eager); "
                      f"step_frac = algorithmic bytes of every launch of the timed graph-replayed steps / their "
                      f"wall time / peak"}


def load_traffic(workload, family):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(workload, {}).get(str(family))
    except (OSError, ValueError):
        return None


class OracleLU:
    """The reference step with a direct solve for one species of the
    2000x1000 trap: the oracle assembles A (element tensors, Robin, reaction)
    and replaces the Dirichlet rows; SuperLU factors it once (the operator is
    the same every step), in a background thread, so later steps cost one
    triangular solve each.  The cell loop is the oracle's sequential one."""

    def __init__(self, trap, species):
        import threading
        self.trap, self.s = trap, species
        self.lu = None
        self.thread = threading.Thread(target=self._factor, daemon=True)
        self.thread.start()

    def _factor(self):
        import scipy.sparse.linalg as sla
        from oracle import oracle as O
        from eq_amd import workloads as W
        t = self.trap
        self.O = O
        self.p = O.problem(t.nx, t.ny, t.width, t.height, 0.1, t.D[self.s], t.gamma[self.s], 0.0,
                           walls=W.TOP_DIRICHLET)
        n = O.num_nodes(self.p)
        self.rp, self.c, self.v = O.assemble(self.p)
        v2, _ = O.apply_dirichlet(self.p, self.rp, self.c, self.v, np.zeros(n))
        self.lu = sla.splu(O.to_scipy(self.rp, self.c, v2, n).tocsc(), permc_spec="COLAMD")

    def step(self, u):
        self.thread.join()
        b = self.O.assemble_rhs(self.p, u)
        _, b2 = self.O.apply_dirichlet(self.p, self.rp, self.c, self.v, b)
        return self.lu.solve(b2)


def _errors(a, b):
    return {"rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)), "max_abs": float(np.abs(a - b).max()),
            "max_rel": float(np.abs(a - b).max() / np.abs(b).max())}


def config2_drift(F, W, cells_host, cells_dev, q_dev, nsteps, oracle_lu=None):
    """Config 2: the config-1 trap started from u0 ~ U[0,1] (seed 2), stepped
    independently by the mixed solver, by the refinement solver (FP32 inner
    solves) and by the FP64 GMG-CG solver at rtol 1e-13 (at most 30
    iterations: its rounding floor, ~3e-12, is reached well before), with the
    same cells.  At checkpoints: each solver's fields against the FP64 run
    (the drift of the trajectory), and -- when oracle_lu is given -- each
    solver's step against the oracle's direct solve from the same input
    (per-step error, the input being that solver's own state; computed by
    finalize() once the factorisations are done)."""
    import torch
    trap2 = W.config1(u0_seed=2)
    nc = trap2.cells.shape[0]
    q = trap2.rates * 0.1
    samp = {k: torch.empty((nc, 2), dtype=torch.float64, device="cuda") for k in ("ref", "mixed", "ir")}
    runs = {"ref": Run(F, trap2, F.GMG_CG, cells_dev, q_dev, samp["ref"], rtol=1e-13, max_iter=30),
            "mixed": Run(F, trap2, F.MIXED_GMG_CG, cells_dev, q_dev, samp["mixed"]),
            "ir": Run(F, trap2, F.IR_GMG_CG, cells_dev, q_dev, samp["ir"])}
    checkpoints = sorted({1, 2, 5, 10, 20, 50, 100, 200, 500, 1000, nsteps})
    oracle_points = sorted({1, 10, 100, nsteps})
    hist, snaps = [], []
    t_run = {"mixed": 0.0, "ir": 0.0}
    for k in range(nsteps):
        runs["ref"].step(k, check=False)
        for name in ("mixed", "ir"):
            before = [runs[name].fb.download(s) for s in range(2)] if (oracle_lu and k + 1 in oracle_points) else None
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            runs[name].step(k)
            torch.cuda.synchronize()
            t_run[name] += time.perf_counter() - t0
            if before is not None:
                snaps.append({"solver": name, "step": k + 1, "before": before, "cells": cells_host[k].copy(),
                              "after": [runs[name].fb.download(s) for s in range(2)]})
        if k + 1 in checkpoints:
            rec = {"step": k + 1}
            ref_f = [runs["ref"].fb.download(s) for s in range(2)]
            for name in ("mixed", "ir"):
                rec[name] = [_errors(runs[name].fb.download(s), ref_f[s]) for s in range(2)]
            sr = samp["ref"].cpu().numpy()
            rec["sampled_max_rel"] = {n: float(np.abs(samp[n].cpu().numpy() - sr).max() / max(np.abs(sr).max(), 1e-300))
                                      for n in ("mixed", "ir")}
            hist.append(rec)
            log(f"bench: config 2 step {k + 1}: rel L2 vs FP64 run: mixed " +
                ", ".join(f"{e['rel_l2']:.2e}" for e in rec["mixed"]) + "; refinement " +
                ", ".join(f"{e['rel_l2']:.2e}" for e in rec["ir"]))
    res = {"workload": "config-1 trap, u0 ~ U[0,1] seed 2, 5000 random-walk rods",
           "steps": nsteps,
           "reference": "FP64 GMG-CG, rtol 1e-13 capped at 30 iterations (its rounding floor), stepped independently",
           "mixed_value": trap2.nodes * 2 * nsteps / t_run["mixed"],
           "ir_value": trap2.nodes * 2 * nsteps / t_run["ir"],
           "mixed_iterations_max": int(max(runs["mixed"].iters)), "mixed_relres_max": float(max(runs["mixed"].relres)),
           "ir_inner_iterations_max": int(max(runs["ir"].iters)), "ir_relres_max": float(max(runs["ir"].relres)),
           "reference_relres_max": float(max(runs["ref"].relres)),
           "drift_vs_fp64_run": hist}
    for r in runs.values():
        r.fb.close()

    def finalize():
        """per-step errors against the oracle's direct solve, from each solver's own input state"""
        from oracle import oracle as O
        grid = O.grid(trap2.nx + 1, trap2.ny + 1, trap2.npm, trap2.width, trap2.height)
        out = []
        for sn in snaps:
            fields, _, _ = O.couple(grid, sn["cells"], sn["before"], q)
            ref = [oracle_lu[s].step(fields[s]) for s in range(2)]
            out.append({"solver": sn["solver"], "step": sn["step"],
                        "species": [_errors(sn["after"][s], ref[s]) for s in range(2)]})
        res["per_step_vs_oracle_lu"] = out
        res["per_step_vs_oracle_lu_max_rel_l2"] = max(e["rel_l2"] for o in out for e in o["species"]) if out else None
        log("bench: config 2 per-step error vs the oracle's LU: max rel L2 "
            f"{res['per_step_vs_oracle_lu_max_rel_l2']:.2e}" if out else "bench: config 2: no oracle checkpoints")
        return res

    return res, finalize


def config3_batched(F, W, nsteps, warmup=2, ntraps=512, sigma=0.05, prof=2):
    """Config 3: 512 uneven traps (seed 3) x 2 species in one batch, mixed
    solver, cycle shape picked by the autotuner after the first step; every
    step couples all traps' cells in one pass and solves all 1024 problems
    together.  Cells random-walk (sigma 0.05 um, the config-1 walk) inside
    their own trap."""
    import torch
    traps = W.config3_traps(ntraps, 3)
    descs = [F.field_desc(t.nx, t.ny, t.width, t.height, 0.1, D, g, 0.0, W.TOP_DIRICHLET)
             for t in traps for D, g in zip(t.D, t.gamma)]
    nodes = sum(t.nodes for t in traps) * 2
    counts = [t.cells.shape[0] for t in traps]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    cells0 = np.concatenate([t.cells for t in traps])
    wlim = np.concatenate([np.full(t.cells.shape[0], t.width) for t in traps])
    hlim = np.concatenate([np.full(t.cells.shape[0], t.height) for t in traps])
    q = torch.from_numpy(np.concatenate([t.rates for t in traps]) * 0.1).to("cuda").contiguous()
    total = warmup + nsteps + prof
    walk = np.empty((total,) + cells0.shape)
    cur = cells0
    for k in range(total):
        cur = W.random_walk_step(cur, 3, k, sigma, wlim, hlim)
        walk[k] = cur
    cells_d = torch.from_numpy(walk).to("cuda").contiguous()
    sampled = torch.empty((cells0.shape[0], 2), dtype=torch.float64, device="cuda")
    t_create = time.perf_counter()
    fb = F.FieldBatch(descs, nspecies=2, solver=solver_desc(F, F.MIXED_GMG_CG))
    t_create = time.perf_counter() - t_create
    stride = cells0.shape[0] * 5 * 8

    def step(k):
        fb.couple_traps_dev(0, off, cells_d.data_ptr() + k * stride, 10.0, q.data_ptr(), sampled.data_ptr())
        return fb.step()

    step(0)
    t_tune = time.perf_counter()
    sd = fb.autotune(1)
    t_tune = time.perf_counter() - t_tune
    for k in range(1, warmup):
        step(k)
    torch.cuda.synchronize()
    iters, rel = [], []
    fb.traffic(reset=True)
    t0 = time.perf_counter()
    for k in range(warmup, total - prof):
        st = step(k)
        iters.append(st.iterations_max)
        rel.append(st.relres_max)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    step_bytes = fb.traffic(reset=True)[0] / nsteps
    fams, eager = family_profile(F, fb, lambda k: step(total - prof + k), prof) if prof else ({}, 0.0)
    roof = roofline_from(fams, step_bytes, dt / nsteps, prof, eager, "config3") if fams else None
    res = {"workload": f"{ntraps} uneven traps (seed 3), h=0.1 um, 2 species, {int(off[-1])} rods",
           "value": nodes * nsteps / dt, "unit": "DOF-timesteps/s", "steps": nsteps,
           "ms_per_step": dt / nsteps * 1e3, "dofs": nodes, "problems": len(descs),
           "iterations_max": int(max(iters)), "relres_max": float(max(rel)),
           "autotuned_cycle": {"pre": int(sd.mg_pre), "post": int(sd.mg_post),
                               "fuse_max_nodes": int(sd.mg_fuse_max_nodes), "fuse_pre_finest": int(sd.mg_fuse_pre_top)},
           "autotune_seconds": t_tune, "create_seconds": t_create, "roofline": roof,
           "kernel_families": {f["family"]: f for f in fams.values()} or None}
    log(f"bench: config 3: {res['value']:.4e} DOF-timesteps/s, {res['ms_per_step']:.2f} ms/step, "
        f"V({sd.mg_pre},{sd.mg_post}), iterations max {res['iterations_max']}")
    fb.close()
    return res


def config4_advection(F, W, nsteps, warmup=2, prof=1):
    """Config 4: one 8192x4096 trap (h = 0.025 um), 4 species, parabolic
    channel velocity along x (peak 50 um/s, zero at the walls), mixed
    GMG-BiCGStab to 1e-10, cycle shape and fusion picked by the autotuner
    after the first step; 100 000 static rods (seed 4) coupled every step."""
    import torch
    trap = W.config4()
    vx = W.channel_velocity(trap.ny, trap.height)
    descs = [F.field_desc(trap.nx, trap.ny, trap.width, trap.height, 0.1, D, g, 0.0, W.TOP_DIRICHLET)
             for D, g in zip(trap.D, trap.gamma)]
    nodes = (trap.nx + 1) * (trap.ny + 1) * len(descs)
    nc = trap.cells.shape[0]
    cells_d = torch.from_numpy(trap.cells).to("cuda").contiguous()
    q = torch.from_numpy(trap.rates * 0.1).to("cuda").contiguous()
    sampled = torch.empty((nc, len(descs)), dtype=torch.float64, device="cuda")
    t_create = time.perf_counter()
    fb = F.FieldBatch(descs, nspecies=len(descs), solver=solver_desc(F, F.MIXED_GMG_BICGSTAB))
    for s in range(len(descs)):
        fb.set_velocity(s, vx)
    t_create = time.perf_counter() - t_create

    def step():
        fb.couple_dev(0, cells_d.data_ptr(), nc, trap.npm, q.data_ptr(), sampled.data_ptr(), 0,
                      F.SAMPLE | F.DEPOSIT)
        return fb.step()

    step()
    t_tune = time.perf_counter()
    sd = fb.autotune(1)
    t_tune = time.perf_counter() - t_tune
    for _ in range(warmup - 1):
        step()
    torch.cuda.synchronize()
    iters, rel = [], []
    fb.traffic(reset=True)
    t0 = time.perf_counter()
    for _ in range(nsteps):
        st = step()
        iters.append(st.iterations_max)
        rel.append(st.relres_max)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    step_bytes = fb.traffic(reset=True)[0] / nsteps
    fams, eager = family_profile(F, fb, lambda k: step(), prof) if prof else ({}, 0.0)
    roof = roofline_from(fams, step_bytes, dt / nsteps, prof, eager, "config4") if fams else None
    res = {"workload": "one 204.8x102.4 um trap, 8192x4096 P1 mesh (33 566 721 nodes/species), 4 species, "
                       "parabolic v_x peak 50 um/s, 100000 static rods (seed 4), mixed GMG-BiCGStab",
           "value": nodes * nsteps / dt, "unit": "DOF-timesteps/s", "steps": nsteps, "warmup": warmup,
           "ms_per_step": dt / nsteps * 1e3, "dofs": nodes, "iterations_max": int(max(iters)),
           "relres_max": float(max(rel)), "create_seconds": t_create,
           "autotuned_cycle": {"pre": int(sd.mg_pre), "post": int(sd.mg_post),
                               "fuse_max_nodes": int(sd.mg_fuse_max_nodes), "fuse_pre_finest": int(sd.mg_fuse_pre_top)},
           "autotune_seconds": t_tune, "roofline": roof,
           "kernel_families": {f["family"]: f for f in fams.values()} or None}
    log(f"bench: config 4: {res['value']:.4e} DOF-timesteps/s, {res['ms_per_step']:.2f} ms/step, "
        f"BiCGStab iterations max {res['iterations_max']}, relres {res['relres_max']:.2e}, "
        f"V({sd.mg_pre},{sd.mg_post}) fused-finest-pre {sd.mg_fuse_pre_top}")
    fb.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--solver", default="mixed", choices=["mixed", "gmg", "jacobi", "ir"])
    ap.add_argument("--variant-steps", type=int, default=5)
    ap.add_argument("--no-variants", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--extrapolate", type=int, default=0, choices=[0, 1],
                    help="start each solve from 2 u_n - u_(n-1) (1) or from u_n (0)")
    ap.add_argument("--c3-steps", type=int, default=5,
                    help="timed steps of the config-3 batch (512 uneven traps), part of the variants")
    ap.add_argument("--c2-steps", type=int, default=1000,
                    help="steps of the config-2 drift comparison (mixed vs FP64), part of the variants")
    ap.add_argument("--c2-oracle", type=int, default=1, choices=[0, 1],
                    help="per-step errors of config 2 against the oracle's direct solve (SuperLU factorisation in "
                         "background threads, ~80 s and ~6 GB per species)")
    ap.add_argument("--c4-steps", type=int, default=3,
                    help="timed steps of config 4 (8192x4096 trap, 4 species, advection), part of the variants")
    ap.add_argument("--only", choices=["c3", "c4"], default=None,
                    help="run only the config-3 batch or the config-4 trap (for profiling) and print its line")
    ap.add_argument("--streams", type=int, default=None, choices=[1, 2, 3, 4],
                    help="solver option streams (problem groups of a CG solve; default: the library's, 2); the "
                         "counter runs take 1, so each counted launch covers the whole batch as in the family timing")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 256 MiB read + 256 MiB write elementwise kernel (torch.mul) three times first, "
                         "as the known-size reference for HBM byte counters")
    args = ap.parse_args()
    if args.streams is not None:
        SOLVER_OPTS["streams"] = args.streams
    if args.gpus != 1:
        print(json.dumps({"error": "this field step runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from eq_amd import field as F
    from eq_amd import workloads as W

    torch.cuda.set_device(0)
    if args.copy_probe:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b
    if args.only:
        res = (config3_batched(F, W, args.steps, warmup=max(args.warmup, 2)) if args.only == "c3"
               else config4_advection(F, W, args.steps, warmup=max(args.warmup, 2)))
        line = {"metric": "DOF-timesteps/s at 1e-10 rel. residual (FP64 vs mixed/batched); % HBM peak",
                "value": res["value"], "unit": "DOF-timesteps/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True,
                "scaling": "weak", "vs_baseline": None, "dtype": "f64 (outer) / f32 (multigrid)",
                "data": "synthetic", "config": {"workload": res["workload"]}, "detail": res}
        print(json.dumps(line), flush=True)
        return
    kinds = {"mixed": F.MIXED_GMG_CG, "gmg": F.GMG_CG, "jacobi": F.JACOBI_CG, "ir": F.IR_GMG_CG}
    trap = W.config1()
    nsteps = args.warmup + args.steps
    log(f"bench: building {nsteps} steps of cell positions for {trap.cells.shape[0]} cells")
    need = max(nsteps + min(10, args.steps), args.variant_steps + 3, 0 if args.no_variants else args.c2_steps)
    cells_all = make_cells(trap, need)
    cells_dev = torch.from_numpy(cells_all).to("cuda", dtype=torch.float64).contiguous()
    q_dev = torch.from_numpy(trap.rates * 0.1).to("cuda", dtype=torch.float64).contiguous()
    sampled_dev = torch.empty((trap.cells.shape[0], 2), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    run = Run(F, trap, kinds[args.solver], cells_dev, q_dev, sampled_dev, extrapolate=args.extrapolate)
    for k in range(args.warmup):
        run.step(k)
    # timed region: graph replays, no per-launch events; the traffic counter
    # sums the algorithmic bytes of every launch (graph replays included)
    log(f"bench: timing {args.steps} steps ({args.solver})")
    torch.cuda.synchronize()
    run.fb.traffic(reset=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        run.step(args.warmup + k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    step_bytes, step_launches = run.fb.traffic(reset=True)
    step_bytes /= args.steps
    nodes = trap.nodes * len(trap.D)
    value = nodes * args.steps / elapsed
    iters = run.iters[args.warmup:]
    relres = run.relres[args.warmup:]
    # per-family kernel time over the next prof_steps steps; the roofline is
    # the family with the most kernel time
    roof, fams = None, None
    prof_steps = min(10, args.steps)
    if not args.no_profile and prof_steps > 0:
        fams, eager_s = family_profile(F, run.fb, lambda k: run.step(args.warmup + args.steps + k), prof_steps)
        roof = roofline_from(fams, step_bytes, elapsed / args.steps, prof_steps, eager_s)
        for k, f in sorted(fams.items(), key=lambda kv: -kv[1]["ms_per_step"]):
            log(f"bench: {f['family']}: {f['ms_per_step']:.3f} ms/step ({f['share']:.1%}), "
                f"{f['achieved_gbs']:.0f} GB/s")
    fields_final = [run.fb.download(s) for s in range(2)]
    # each workload below runs on its own batch: this one's streams go now (a
    # CG solve overlaps its two problem groups only while their streams sit on
    # different hardware queues, and the device has few)
    run.fb.close()
    log(f"bench: {value:.4e} DOF-timesteps/s, {elapsed / args.steps * 1e3:.3f} ms/step, "
        f"iterations max {max(iters)}, relres max {max(relres):.2e}")

    out = {
        "metric": "DOF-timesteps/s at 1e-10 rel. residual (FP64 vs mixed/batched); % HBM peak",
        "value": value,
        "unit": "DOF-timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "f64 (Krylov) / f32 (multigrid)", "ir": "f64 (refinement) / f32 (inner Krylov, multigrid)"}
                 .get(args.solver, "f64"),
        "data": "synthetic",
        "config": {"workload": "single 100x50 um trap, 2000x1000 P1 mesh (2 003 001 nodes/species), 2 species, "
                               "5000 random-walk rods, dt 0.1 s, top Dirichlet, rtol 1e-10",
                   "solver": {"mixed": "FP64 CG + FP32 GMG V(2,2) preconditioner",
                              "gmg": "FP64 CG + FP64 GMG V(2,2) preconditioner",
                              "jacobi": "FP64 Jacobi-CG",
                              "ir": "FP64 iterative refinement around FP32 GMG-CG (FP32 vectors and stencil)"}[args.solver],
                   "species": 2, "nodes_per_species": trap.nodes, "cells": int(trap.cells.shape[0]),
                   "parallelism": "single GPU, species batched in one launch per kernel"},
        "iterations": {"max": int(max(iters)), "mean": float(np.mean(iters))},
        "relres_max": float(max(relres)),
        "roofline": roof,
        "kernel_families": {F.KERNEL_FAMILIES[k]: v for k, v in fams.items()} if fams else None,
        "step_traffic": {"bytes_per_step": step_bytes, "launches_per_step": step_launches / args.steps,
                         "achieved_gbs": step_bytes / (elapsed / args.steps) / 1e9,
                         "frac": step_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS},
    }

    if not args.no_variants:
        # the oracle's direct solves for config 2's checkpoints factor in the
        # background while the GPU sections run
        oracle_lu = [OracleLU(W.config1(u0_seed=2), s) for s in range(2)] if args.c2_oracle else None
        variants = {}
        for name in ("gmg", "jacobi", "mixed", "ir"):
            if name == args.solver:
                continue
            r = Run(F, trap, kinds[name], cells_dev, q_dev, sampled_dev, max_iter=20000,
                    check_every=(64 if name == "jacobi" else 4))
            # three untimed steps: the first captures the iteration graphs
            for k in range(3):
                r.step(k)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for k in range(3, 3 + args.variant_steps):
                r.step(k)
            torch.cuda.synchronize()
            dtv = time.perf_counter() - t1
            variants[name] = {"value": nodes * args.variant_steps / dtv, "steps": args.variant_steps, "warmup": 3,
                              "iterations_max": int(max(r.iters[3:])), "relres_max": float(max(r.relres[3:]))}
            if name == "jacobi":
                variants[name]["first_step_iterations"] = [int(v) for v in r.fb.problem_stats()[0]]
                jac0 = Run(F, trap, kinds[name], cells_dev, q_dev, sampled_dev, max_iter=20000, check_every=64)
                jac0.step(0)
                variants[name]["step0_iterations"] = [int(v) for v in jac0.fb.problem_stats()[0]]
                jac0.fb.close()
            r.fb.close()
            log(f"bench: variant {name}: {variants[name]['value']:.4e} DOF-timesteps/s")
        out["variants"] = variants
        # the mixed solver against the FP64 GPU solvers of the same workload
        fp64 = variants.get("gmg", {}).get("value") if args.solver != "gmg" else value
        jac = variants.get("jacobi", {}).get("value")
        mixv = value if args.solver == "mixed" else variants.get("mixed", {}).get("value")
        irv = variants.get("ir", {}).get("value")
        out["speedup_vs_fp64"] = {
            "mixed_over_fp64_gmg_cg": mixv / fp64 if (mixv and fp64) else None,
            "mixed_over_fp64_jacobi_cg": mixv / jac if (mixv and jac) else None,
            "refinement_over_fp64_gmg_cg": irv / fp64 if (irv and fp64) else None,
            "refinement_over_fp64_jacobi_cg": irv / jac if (irv and jac) else None}
        c2, c2_finalize = config2_drift(F, W, cells_all, cells_dev, q_dev, args.c2_steps, oracle_lu)
        out["config2"] = c2
        if args.c3_steps > 0:
            out["config3"] = config3_batched(F, W, args.c3_steps)
        if args.c4_steps > 0:
            out["config4"] = config4_advection(F, W, args.c4_steps)
        if oracle_lu:
            log("bench: config 2 per-step errors against the oracle's direct solve")
            out["config2"] = c2_finalize()
            del oracle_lu

    if not args.no_cpu_baseline:
        jac_iters = out.get("variants", {}).get("jacobi", {}).get("step0_iterations")
        if jac_iters is None:
            jr = Run(F, trap, F.JACOBI_CG, cells_dev, q_dev, sampled_dev, max_iter=20000, check_every=64)
            jr.step(0)
            jac_iters = [int(v) for v in jr.fb.problem_stats()[0]]
            jr.fb.close()
        log("bench: timing the CPU baseline sample")
        out["cpu_baseline"] = cpu_baseline(trap, jac_iters)
        out["cpu_baseline"]["extrapolated"] = True
        log("bench: config 0 in full on the CPU (oracle, banded LU) and on the GPU")
        c0, f_cpu = cpu_baseline_config0()
        g0, f_gpu = gpu_config0(F, W)
        c0["gpu_same_workload"] = g0
        c0["gpu_vs_cpu_rel_l2_after_100_steps"] = [float(np.linalg.norm(a - b) / np.linalg.norm(b))
                                                   for a, b in zip(f_gpu, f_cpu)]
        out["cpu_baseline_config0"] = c0
    del fields_final
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
