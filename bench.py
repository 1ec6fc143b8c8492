#!/usr/bin/env python3
"""SIMPLE hot-path benchmark on one MI355X.

Default workload: the mixed-precision channel case — 200x100x200 box, 15 %
seeded node jitter, 10 % of the hexes split into 6 tetrahedra (6.0 M cells),
laminar channel at Re = 1000, GMRES(50) + Jacobi on every solve.  It runs the
FP64 path and the mixed path (FP32 inner Krylov solves inside FP64 iterative
refinement) on the same case from rest; `value` is the mixed path's SIMPLE
iterations/s and the line also carries the FP64 rate, the speed-up and the
per-field relative L2 difference between the two runs.

One "step" is one SIMPLE iteration (momentum assembly + 3 solves, Rhie-Chow
flux, pressure-correction assembly + solve, corrections, k-epsilon update for
turbulent workloads) with every field, matrix and Krylov vector resident in
HBM.

Other workloads (--workload): `cavity` (32^3 lid-driven cavity, Re 100,
GMRES(30)+Jacobi, 200 iterations, FP64), `step` (backward-facing step,
1.12 M jittered tetrahedra, k-epsilon, Re 5000, FP64 with the reference's
default solvers), `large` (400x200x250 jittered hex channel, 20 M cells,
k-epsilon, mixed GMRES(50)+ILU(0)), `sweep` (64 lid-driven cavities of
edge 16..64 and Re 100..1000 run as one batch, mixed GMRES(30)+Jacobi).

Prints one JSON line: SIMPLE iterations/s, the roofline of the SpMV kernel
(algorithmic bytes per launch / HIP-event-timed launch duration vs 8 TB/s),
and the FP64 CPU oracle timed on the host cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the production SpMV on the channel workload's
# pressure-correction matrix (5,995,990 cells), from rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by a 1 GiB copy (profiles/hbm_traffic_channel_6M.md):
# 276.4 MB read + 25.7 MB written.  Below the algorithmic figure, which prices
# the CSR form (4-byte column ids): the FP32 copy stores 16-bit column offsets.
MEASURED_TRAFFIC = {("k_spmv_sell2<float, false, false, true>", 5_995_990): 302_145_434}
METRIC = "SIMPLE iters/s @4M cells, FP64 vs mixed (<=1e-6 rel err); SpMV % HBM peak"
STAGES = ["total", "build_velocity", "solve_velocity", "build_pressure", "solve_pressure", "other"]


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# --------------------------------------------------------------------------- workloads

def _channel_case(dims, jitter, split, re, turbulent=False):
    import sysu_cfd_parallel_amd as s

    return s.channel(*dims, lx=2.0, ly=1.0, lz=2.0, jitter=jitter, split=split, seed=0, re=re,
                     turbulent=turbulent)


def _gmres_all(p, restart, pc, maxit=20000, rtol=None):
    import sysu_cfd_parallel_amd as s

    # one KSP setting on every solve, as the reference's unprefixed
    # KSPSetFromOptions applies command-line options to all of them; the
    # iteration cap raised as a command-line -ksp_max_it would
    for k in range(3):
        o = p.ksp[k]
        o.type, o.restart, o.pc, o.maxit = s.KSP_GMRES, restart, pc, maxit
        if rtol is not None:
            o.rtol = rtol
    return p


class Channel:
    name = "channel"
    sizes = {"full": (200, 100, 200), "medium": (100, 50, 100), "small": (40, 20, 40)}
    precisions = ["fp64", "mixed"]
    steps, warmup = 5, 1
    turbulent = False
    autotune, graphs = 1, 0
    assembly_fp32 = 0
    # CPU baseline on the GPU case itself from the GPU's post-warm-up state:
    # "steps" times whole SIMPLE steps, "projected" times every phase but the
    # pressure solve in full and the pressure solve per Krylov iteration
    # (a step from this state takes minutes on the host cores); None: a
    # smaller sample from rest (cpu_sample)
    cpu_mode = "projected"

    def __init__(self, size):
        self.dims = self.sizes[size]

    def case(self, dims=None):
        return _channel_case(dims or self.dims, 0.15, 0.10, 1000.0)

    def params(self, prec):
        import sysu_cfd_parallel_amd as s

        p = _gmres_all(s.default_params(), 50, s.PC_JACOBI)
        p.residual_steady = 0.0
        p.precision = prec
        # the FP32 cycle on the rows in the order the generator wrote the
        # cells (lattice order), with 16-bit SELL column offsets
        p.inner_order = 2
        return p

    def describe(self, prec_name):
        d = self.dims
        return (f"laminar channel Re=1000, {d[0]}x{d[1]}x{d[2]} box, 15% jitter, 10% hexes split into 6 tets, "
                f"GMRES(50)+Jacobi, rtol 1e-8 momentum / 1e-6 pressure, iteration caps raised from the reference's "
                f"1000 to 20000 (the pressure solve takes ~3000), "
                + ("FP32 inner solves in FP64 refinement, the FP32 cycle on the rows in the mesh's input cell order"
                   if prec_name == "mixed" else "FP64"))

    def cpu_sample(self):
        # 1/8 of the cells: the full mesh takes minutes per CPU step
        dims = self.sizes["medium"] if self.dims == self.sizes["full"] else self.dims
        return self.case(dims), "channel %dx%dx%d, same jitter/split/Re/solver settings" % dims


class Cavity(Channel):
    name = "cavity"
    sizes = {"full": 32, "medium": 32, "small": 16}
    precisions = ["fp64"]
    steps, warmup = 200, 0
    graphs = 1
    cpu_mode = "steps"

    def case(self, dims=None):
        import sysu_cfd_parallel_amd as s

        return s.cavity(dims or self.dims, re=100.0)

    def params(self, prec):
        import sysu_cfd_parallel_amd as s

        p = _gmres_all(s.default_params(), 30, s.PC_JACOBI, maxit=1000, rtol=1e-8)
        p.ksp[s.KSP_SCALAR].maxit = 500
        p.residual_steady = 0.0
        p.precision = prec
        return p

    def describe(self, prec_name):
        n = self.dims
        return f"lid-driven cavity Re=100, {n}^3 uniform hexes, GMRES(30)+Jacobi rtol 1e-8, FP64"

    def cpu_sample(self):
        return self.case(), "the same cavity"


class Step(Channel):
    name = "step"
    sizes = {"full": 9, "medium": 6, "small": 3}
    precisions = ["fp64"]
    steps, warmup = 5, 1
    turbulent = True
    graphs = 1  # the ILU(0) level sweeps replay from captured graphs
    cpu_mode = "steps"

    def case(self, dims=None):
        import sysu_cfd_parallel_amd as s

        return s.step(dims or self.dims, jitter=0.10, tets=True, turbulent=True, re=5000.0)

    def params(self, prec):
        import sysu_cfd_parallel_amd as s

        # the reference's solver defaults: GMRES(30)+ILU(0) for momentum and
        # k/epsilon, CG+ILU(0) for the symmetric pressure correction
        p = s.default_params()
        p.turb_model = 1
        p.residual_steady = 0.0
        p.precision = prec
        # the pressure correction relaxed by 0.3 (a param.in relaxation line):
        # SIMPLE's default 0.5 is past the outer iteration's stability limit
        # on these tetrahedra, and the run would diverge
        p.urf[3] = 0.3
        return p

    def describe(self, prec_name):
        n = self.dims
        return (f"backward-facing step ER=2, {n} cells per step height, 10%-jittered hex lattice with every hex "
                f"split into 6 tetrahedra (Kuhn split, not a Delaunay fill), k-epsilon with wall functions, Re=5000, "
                f"relaxation 0.6 momentum / 0.3 pressure, GMRES(30)/CG + ILU(0), FP64")

    def cpu_sample(self):
        return self.case(), "the same step case"


class Large(Channel):
    name = "large"
    sizes = {"full": (400, 200, 250), "medium": (200, 100, 125), "small": (40, 20, 25)}
    precisions = ["mixed"]
    steps, warmup = 3, 1
    turbulent = True
    graphs = 1  # the ILU(0) level sweeps replay from captured graphs
    cpu_mode = None  # the oracle's 20 M-cell context needs ~35 GB and minutes to set up

    def case(self, dims=None):
        return _channel_case(dims or self.dims, 0.10, 0.0, 13750.0, turbulent=True)

    def params(self, prec):
        import sysu_cfd_parallel_amd as s

        p = _gmres_all(s.default_params(), 50, s.PC_ILU0)
        p.turb_model = 1
        p.residual_steady = 0.0
        p.precision = prec
        return p

    def describe(self, prec_name):
        d = self.dims
        return (f"turbulent channel (k-epsilon, Re_b=13750 ~ Re_tau 395), {d[0]}x{d[1]}x{d[2]} hexes, 10% jitter, "
                f"GMRES(50)+ILU(0), " + ("FP32 inner solves in FP64 refinement" if prec_name == "mixed" else "FP64"))

    def cpu_sample(self):
        dims = (100, 50, 62) if self.dims == self.sizes["full"] else self.dims
        return self.case(dims), "turbulent channel %dx%dx%d, same jitter/Re/solver settings" % dims


class Sweep(Channel):
    """64 independent lid-driven cavities: edge n ~ U{16..64}, Re ~
    logU[100, 1000] (seeded), run as one batch."""
    name = "sweep"
    sizes = {"full": 64, "medium": 16, "small": 4}
    precisions = ["mixed"]
    steps, warmup = 5, 1
    graphs = 1
    cpu_mode = None

    def __init__(self, size):
        import numpy as np

        self.dims = self.sizes[size]
        rng = np.random.default_rng(2024)
        self.edges = [int(e) for e in rng.integers(16, 65, size=self.dims)]
        self.res = [float(r) for r in np.exp(rng.uniform(np.log(100.0), np.log(1000.0), size=self.dims))]

    def case(self, dims=None):
        import sysu_cfd_parallel_amd as s

        idx = range(self.dims) if dims is None else dims
        return [s.cavity(self.edges[k], re=self.res[k]) for k in idx]

    def params(self, prec):
        import sysu_cfd_parallel_amd as s

        p = _gmres_all(s.default_params(), 30, s.PC_JACOBI, maxit=5000)
        p.residual_steady = 0.0
        p.precision = prec
        # each cavity's FP32 cycle on its rows in the generator's cell order
        p.inner_order = 2
        return p

    def describe(self, prec_name):
        cells = sum(e ** 3 for e in self.edges)
        return (f"{self.dims} lid-driven cavities as one batch, edges U{{16..64}} (sum {cells} cells), "
                f"Re logU[100,1000], seed 2024, GMRES(30)+Jacobi, "
                + ("FP32 inner solves in FP64 refinement, each case's FP32 cycle on its rows in the input cell order"
                   if prec_name == "mixed" else "FP64"))

    def cpu_sample(self):
        # the cases of the batch up to about 300K cells, in batch order
        idx, cells = [], 0
        for k, e in enumerate(self.edges):
            if cells + e ** 3 > 300_000 and idx:
                break
            idx.append(k)
            cells += e ** 3
        return self.case(idx), f"the first {len(idx)} cavities of the batch, run one after the other"


WORKLOADS = {w.name: w for w in (Channel, Cavity, Step, Large, Sweep)}


# --------------------------------------------------------------------------- measurement

def spmv_bytes(nnz: int, nc: int, b_val: int, b_x: int, jacobi: bool = True) -> float:
    # algorithmic bytes of the production y = M^-1 A x: CSR values + column
    # ids, row pointers, x, y and (Jacobi) the inverse diagonal
    return (b_val + 4) * nnz + 4 * (nc + 1) + (3 if jacobi else 2) * b_x * nc


def spmv_kernel_name(shape: dict, elem: str, jacobi: bool = True) -> str:
    tf = lambda v: "true" if v else "false"  # noqa: E731
    if shape["lanes"] == 0:
        two = "2" if shape.get("slices") == 2 else ""  # two slices per wave
        c16 = tf(shape.get("order", 0))  # 16-bit column offsets in the inner order
        return f"k_spmv_sell{two}<{elem}, {tf(shape['nt'])}, {tf(jacobi)}, {c16}>"
    return f"k_spmv_pc<{elem}, {shape['lanes']}, {tf(shape['nt'])}>"


def host_cores() -> dict:
    """The host cores this process may use (its affinity, bounded by
    OMP_NUM_THREADS where that is set) and the CPU model."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(omp)) if omp.isdigit() and int(omp) > 0 else avail
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": avail, "threads": threads, "model": model}


def _timed(fn):
    t = time.perf_counter()
    out = fn()
    return time.perf_counter() - t, out


def cpu_baseline_same_case(wl, case, state: dict, gpu_its: dict, threads: int, budget_s: float = 25.0,
                           p_cap: int = 200) -> dict:
    """FP64 CPU oracle on the GPU's case itself, started from the state the
    GPU's timed steps start from (after its warm-up).  cpu_mode "steps": whole
    SIMPLE steps until the budget is spent.  "projected": one step's phases
    timed in full (momentum assembly and its 3 solves, pressure-correction
    assembly, corrections, k-epsilon), the pressure solve over its first
    p_cap iterations, the step's time = the phases + the GPU FP64 run's mean
    pressure iterations per step x the CPU time per iteration (the same
    GMRES(m) on the same matrix).  Both on `threads` host threads; the
    single-core figure scales the step by the 1-thread / all-thread time of
    a pressure Krylov iteration on the same matrix."""
    from oracle.pyoracle import Oracle

    t_set, o = _timed(lambda: Oracle(case, wl.params(0), threads=threads))
    for n, v in state.items():
        o.set(n, v)
    out = {"setup_s": t_set, "cells": o.sizes.n_cell}
    if wl.cpu_mode == "steps":
        steps, t0 = 0, time.perf_counter()
        while True:
            _, done, rc = o.iterate(1)
            steps += 1
            if rc != 0 or time.perf_counter() - t0 > budget_s or steps >= 200:
                break
        dt = time.perf_counter() - t0
        out.update({"step_s": dt / steps, "steps": steps, "how": f"{steps} whole SIMPLE steps in {dt:.1f}s"})
        p_its = None
    else:
        rtol_p = o.params.ksp[1].rtol
        t_v, (rc, _, vinfo) = _timed(o.calculate_velocity)
        t_bp, _ = _timed(lambda: [o.call(k) for k in ("bc_velocity", "cal_ru_face", "build_pressure_matrix")])
        t_sp, pinfo = _timed(lambda: o.solve_pressure(rtol_p, p_cap))
        k = max(pinfo["iters"], 1)
        t_it = t_sp / k
        t_c, _ = _timed(lambda: (o.correct_fields(), o.call("bc_velocity"), o.call("correct_ru_face")))
        t_ke = _timed(o.update_tur_kepsilon)[0] if wl.turbulent else 0.0
        p_its = gpu_its.get("P", 0.0)
        step = t_v + t_bp + p_its * t_it + t_c + t_ke
        out.update({"step_s": step, "steps": 1, "pressure_s_per_iteration": t_it,
                    "how": (f"momentum phase {t_v:.2f}s ({[i['iters'] for i in vinfo]} its), pressure assembly "
                            f"{t_bp:.2f}s, pressure solve {t_it * 1e3:.2f} ms/iteration over its first {k} "
                            f"iterations x {p_its:.0f} iterations per step (the GPU FP64 run's mean), corrections "
                            f"{t_c:.2f}s" + (f", k-epsilon {t_ke:.2f}s" if wl.turbulent else ""))})
    # single core: a pressure Krylov iteration on 1 thread against all threads
    o.lib.orc_set_threads(1)
    o.call("build_pressure_matrix")
    t1, i1 = _timed(lambda: o.solve_pressure(o.params.ksp[1].rtol, 8))
    o.lib.orc_set_threads(threads)
    tn, iN = _timed(lambda: o.solve_pressure(o.params.ksp[1].rtol, 8))
    r = (t1 / max(i1["iters"], 1)) / max(tn / max(iN["iters"], 1), 1e-12)
    out["single_core_factor"] = r
    o.close()
    return out


def cpu_baseline(wl, threads: int, budget_s: float = 20.0, max_steps: int = 3) -> dict:
    """FP64 CPU oracle on a bounded sample of the same workload, started from
    rest like the GPU run and stepped until the budget is spent.  Where the
    sample is smaller than the GPU case, Krylov iteration counts are lower on
    the sample, so its cells*iters/s overstates the CPU rate on the full mesh:
    the scaled figure is an upper bound for the CPU."""
    from oracle.pyoracle import Oracle

    case, what = wl.cpu_sample()
    cases = case if isinstance(case, list) else [case]
    oracles = [Oracle(c, wl.params(0), threads=threads) for c in cases]
    n = sum(o.sizes.n_cell for o in oracles)
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps:
        for o in oracles:
            o.iterate(1)
        steps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    for o in oracles:
        o.close()
    return {"iters_per_s": steps / dt, "cells": n, "steps": steps, "seconds": dt, "what": what}


def run_path(wl, case, prec: int, steps: int, warmup: int, label: str, snapshot: bool = False,
             shapes: dict | None = None) -> dict:
    """W untimed + K timed SIMPLE steps of one precision path from rest;
    with snapshot the state the timed steps start from is kept; shapes
    (precision index -> lanes, nt, rows, sigma) fixes the GMRES kernel shapes
    instead of the autotuner."""
    import numpy as np

    import sysu_cfd_parallel_amd as s

    t = time.perf_counter()
    p = wl.params(prec)
    p.autotune, p.graphs = wl.autotune, wl.graphs
    if prec == s.MIXED:
        p.assembly_fp32 = wl.assembly_fp32
    if shapes:
        p.autotune = 0
    dev = s.BatchSolver(case, p) if isinstance(case, list) else s.Solver(case, p)
    for k, (lanes, nt, rows, sigma) in (shapes or {}).items():
        dev.set_kernel_shapes(k, lanes, nt=nt, rows=rows, sigma=sigma)
    z = dev.sizes
    log(f"[{label}] device context: {z.n_cell} cells, {z.n_face} faces, nnz {z.nnz} "
        f"({time.perf_counter() - t:.1f}s)")
    for k in range(warmup):
        t = time.perf_counter()
        h, _ = dev.iterate(1)
        log(f"[{label}] warmup step {k + 1}/{warmup}: {time.perf_counter() - t:.2f}s, ResMax {h[-1, 10]:.3e}")
    dev.call("synchronize")
    state = {n: dev.get(n) for n in STATE} if snapshot else None
    dev.stage_times(reset=True)
    kprec = 1 if prec == s.MIXED else 0
    dev.spmv_profile(kprec, 1)  # HIP events around the Arnoldi SpMV launches of the timed steps
    names = ("U", "V", "W", "P", "TE", "ED") if wl.turbulent else ("U", "V", "W", "P")
    rows, its = [], {n: 0 for n in names}
    refinements = 0
    quiet = steps > 20
    t0 = time.perf_counter()
    for k in range(steps):
        t = time.perf_counter()
        h, _ = dev.iterate(1)
        rows.append(h[-1])
        info = dev.last_solve_info()
        for n in its:
            its[n] += info[n]["iters"]
        refinements += sum(info[n]["refinements"] for n in its)
        if not quiet or (k + 1) % 25 == 0 or k + 1 == steps:
            log(f"[{label}] step {k + 1}/{steps}: {time.perf_counter() - t:.3f}s, ResMax {h[-1, 10]:.3e}, "
                f"krylov its U {info['U']['iters']} P {info['P']['iters']}")
    dev.call("synchronize")
    t1 = time.perf_counter()
    stages = dev.stage_times()
    spmv_ms, spmv_n = dev.spmv_profile(kprec, -1)
    shapes = dev.kernel_shapes()
    k = kprec
    tuning = {"shapes": shapes["fp32" if k else "fp64"], "report": dev.tune_report(k)}
    log(f"[{label}] kernel shapes {tuning['shapes']}")
    out = {
        "dev": dev, "sizes": z, "seconds": t1 - t0, "history": np.array(rows), "tuning": tuning, "state": state,
        "spmv_in_solve": {"ms_per_launch": spmv_ms, "launches_timed": spmv_n},
        "fields": {n: dev.get(n) for n in names},
        "krylov_its_per_step": {n: v / steps for n, v in its.items()},
        "refinements_per_step": refinements / steps,
        "stage_ms_per_step": {k: v / steps for k, v in zip(STAGES, stages.tolist())},
    }
    log(f"[{label}] {steps / (t1 - t0):.3f} SIMPLE it/s, {(t1 - t0) * 1e3 / steps:.1f} ms/step")
    return out


# the state one SIMPLE step starts from (matrices and gradients are rebuilt)
STATE = ("R", "U", "V", "W", "P", "T", "TE", "ED", "VISLAM", "VISTUR", "APR", "DELTAP", "BRO", "BU", "BV", "BW",
         "BPRE", "BTE", "BED", "BTEM", "SHEAR", "RUFACE")


def one_step_agreement(dev, names) -> dict:
    """The mixed path's own error, without the drift of two trajectories:
    one SIMPLE step from the same state in mixed precision and in FP64, at the
    workload's solve tolerances and with every solve converged 10^4 times
    further (each path's converged step), on the same context; relative L2
    differences of each field."""
    import sysu_cfd_parallel_amd as s

    state = {n: dev.get(n) for n in STATE}
    p = dev.params
    rtol = [p.ksp[k].rtol for k in range(3)]
    out = {}
    for key, prec, tight in (("mixed", s.MIXED, 1.0), ("fp64", s.FP64, 1.0), ("converged", s.FP64, 1e-4),
                             ("mixed_converged", s.MIXED, 1e-4)):
        for n, v in state.items():
            dev.set(n, v)
        p.precision = prec
        for k in range(3):
            p.ksp[k].rtol = rtol[k] * tight
        dev.set_params(p)
        dev.iterate(1)
        out[key] = {n: dev.get(n) for n in names}
    p.precision = s.MIXED
    for k in range(3):
        p.ksp[k].rtol = rtol[k]
    dev.set_params(p)
    res = {f"mixed_vs_fp64.{n}": l2_rel(out["mixed"][n], out["fp64"][n]) for n in names}
    res.update({f"mixed_vs_converged.{n}": l2_rel(out["mixed"][n], out["converged"][n]) for n in names})
    res.update({f"fp64_vs_converged.{n}": l2_rel(out["fp64"][n], out["converged"][n]) for n in names})
    res.update({f"converged_mixed_vs_fp64.{n}": l2_rel(out["mixed_converged"][n], out["converged"][n])
                for n in names})
    return res


def l2_rel(a, b) -> float:
    import numpy as np

    return float(np.linalg.norm(a - b) / max(float(np.linalg.norm(b)), 1e-300))


def spmv_roofline(dev, z, reps: int = 50) -> dict:
    """Production SpMV launch on the assembled pressure-correction matrix,
    each launch timed alone after the Infinity Cache has been overwritten."""
    out = {}
    for name, prec, b in (("fp64", 0, 8), ("fp32", 1, 4)):
        ms = dev.spmv_time(1, prec, reps)
        shape = dev.kernel_shapes()[name]
        jacobi = not shape.get("order")  # the inner order holds the Jacobi scaling in the matrix
        nb = spmv_bytes(z.nnz, z.n_cell, b, b, jacobi)
        out[name] = {"kernel": spmv_kernel_name(shape, "double" if b == 8 else "float", jacobi),
                     "bytes_per_launch": nb,
                     "ms_per_launch": ms, "achieved": nb / (ms * 1e-3) / 1e9, "timing": "cold, one launch at a time"}
        out[name]["frac"] = out[name]["achieved"] / HBM_PEAK_GBS
    log(f"SpMV fp64 {out['fp64']['ms_per_launch'] * 1e3:.1f} us -> {out['fp64']['achieved']:.0f} GB/s; "
        f"fp32 {out['fp32']['ms_per_launch'] * 1e3:.1f} us -> {out['fp32']['achieved']:.0f} GB/s")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="channel")
    ap.add_argument("--size", choices=["full", "medium", "small"], default="full")
    ap.add_argument("--precision", choices=["default", "both", "fp64", "mixed"], default="default",
                    help="both: FP64 then mixed on the same case, value is the mixed rate")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--shapes", default=None,
                    help="fixed GMRES kernel shapes instead of the autotuner, per precision: "
                         "'fp64=lanes,nt,rows,sigma;fp32=lanes,nt,rows,sigma' (lanes -2/-1/0: SELL-64 with two/one "
                         "slices per wave/by size), e.g. for repeatable counter passes")
    ap.add_argument("--assembly", choices=["fp64", "fp32"], default="fp64",
                    help="the mixed path's coefficient assembly: FP64 (default) or FP32 arithmetic")
    ap.add_argument("--graphs", type=int, choices=[0, 1], default=None,
                    help="replay GMRES cycles from captured graphs (default: per workload)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus above 1 is not supported"}))
        log("only --gpus 1 is supported: every case runs on a single MI355X")
        return 2

    import sysu_cfd_parallel_amd as s

    wl = WORKLOADS[args.workload](args.size)
    if args.graphs is not None:
        wl.graphs = args.graphs
    wl.assembly_fp32 = int(args.assembly == "fp32")
    steps = args.steps if args.steps is not None else wl.steps
    warmup = args.warmup if args.warmup is not None else wl.warmup
    order = {"default": wl.precisions, "both": ["fp64", "mixed"], "fp64": ["fp64"], "mixed": ["mixed"]}[args.precision]

    t = time.perf_counter()
    case = wl.case()
    n_elems = sum(c.n_cell_elems for c in case) if isinstance(case, list) else case.n_cell_elems
    log(f"{wl.name} case built in {time.perf_counter() - t:.1f}s ({n_elems} volume elements)")

    shapes = None
    if args.shapes:
        shapes = {}
        for part in args.shapes.split(";"):
            key, vals = part.split("=")
            shapes[{"fp64": 0, "fp32": 1}[key.strip()]] = tuple(int(v) for v in vals.split(","))
    runs = {}
    want_cpu = not args.no_cpu_baseline and wl.cpu_mode is not None and not isinstance(case, list)
    for k, name in enumerate(order):
        # the CPU baseline starts from the first path's post-warm-up state
        runs[name] = run_path(wl, case, s.MIXED if name == "mixed" else s.FP64, steps, warmup, name,
                              snapshot=want_cpu and k == 0, shapes=shapes)
        if name != order[-1]:
            runs[name]["dev"].close()
    main_name = order[-1]
    r = runs[main_name]
    dev, z = r["dev"], r["sizes"]
    value = steps / r["seconds"]

    agreement = None
    if "fp64" in runs and "mixed" in runs:
        a, b = runs["mixed"], runs["fp64"]
        agreement = {n: l2_rel(a["fields"][n], b["fields"][n]) for n in a["fields"]}
        agreement["max"] = max(agreement.values())
        agreement["residual_history_max_rel"] = float(
            (abs(a["history"][:, :4] - b["history"][:, :4]) / abs(b["history"][:, :4]).clip(1e-300)).max())
        log(f"mixed vs FP64 relative L2 field difference: {agreement}")

    accuracy = None
    if agreement is not None:
        t = time.perf_counter()
        one = one_step_agreement(dev, list(runs["mixed"]["fields"]))
        agreement["one_step_from_the_same_state"] = one
        log(f"mixed vs FP64, one step from the same state: {one} ({time.perf_counter() - t:.1f}s)")
        conv = max(v for k, v in one.items() if k.startswith("converged_mixed_vs_fp64."))
        accuracy = {
            "target_rel_err": 1e-6,
            "converged_step_max": conv, "pass": conv <= 1e-6,
            "converged_step": "one SIMPLE step from the same state, mixed vs FP64, every solve converged 1e4 "
                              "times below the workload's tolerance: the error of the mixed arithmetic itself",
            "trajectory_max": agreement["max"], "trajectory_pass": agreement["max"] <= 1e-6,
            "trajectory": f"after the {warmup} + {steps} steps from rest at the workload's tolerances, where each "
                          "path stops every pressure solve at its own 1e-6 and the two drift apart by that",
        }
        log(f"accuracy vs the 1e-6 target: converged step {conv:.2e} ({'pass' if accuracy['pass'] else 'FAIL'}), "
            f"trajectory {agreement['max']:.2e} ({'pass' if accuracy['trajectory_pass'] else 'fail'})")

    spmv = spmv_roofline(dev, z)
    dom = dict(spmv["fp32" if main_name == "mixed" else "fp64"])
    ins = r["spmv_in_solve"]
    if ins["launches_timed"] > 0:
        # the roofline of the launches inside the timed solves themselves;
        # with ILU(0) the Arnoldi SpMV carries no Jacobi scaling, nor in the
        # inner order, whose FP32 matrix holds it
        shape = dev.kernel_shapes()[main_name if main_name == "fp64" else "fp32"]
        jacobi = dev.params.ksp[s.KSP_PRESSURE].pc == s.PC_JACOBI and not shape.get("order")
        b = 4 if main_name == "mixed" else 8
        dom["kernel"] = spmv_kernel_name(shape, "float" if b == 4 else "double", jacobi)
        dom["bytes_per_launch"] = spmv_bytes(z.nnz, z.n_cell, b, b, jacobi)
        dom["ms_per_launch"] = ins["ms_per_launch"]
        dom["achieved"] = dom["bytes_per_launch"] / (ins["ms_per_launch"] * 1e-3) / 1e9
        dom["frac"] = dom["achieved"] / HBM_PEAK_GBS
        dom["timing"] = f"HIP events around {ins['launches_timed']} Arnoldi SpMV launches of the timed steps"
        log(f"in-solve SpMV {ins['ms_per_launch'] * 1e3:.1f} us over {ins['launches_timed']} launches -> "
            f"{dom['achieved']:.0f} GB/s")
    copy_bytes = 1 << 30
    ms_copy = dev.copy_time(copy_bytes, 20)
    copy_gbs = 2 * copy_bytes / (ms_copy * 1e-3) / 1e9
    log(f"plain device copy of 1 GiB: {ms_copy:.3f} ms -> {copy_gbs:.0f} GB/s read+write")
    dev.close()

    cpu = None
    if not args.no_cpu_baseline:
        host = host_cores()
        threads = args.cpu_threads or host["threads"]
        log(f"CPU oracle baseline on {threads} threads ({host['nproc']} cores seen, {host['affinity']} in this "
            f"process's affinity, {host['model']})")
        if want_cpu:
            first = runs[order[0]]
            its = runs["fp64"]["krylov_its_per_step"] if "fp64" in runs else first["krylov_its_per_step"]
            cb = cpu_baseline_same_case(wl, case, first["state"], its, threads)
            rate = 1.0 / cb["step_s"]
            cpu = {
                "value": rate, "unit": "SIMPLE iterations/s (FP64 oracle)", "cores": threads, "kind": "port",
                "sample": f"the same {cb['cells']}-cell case from the GPU run's state after its {warmup} warm-up "
                          f"steps: {cb['how']}",
                "single_core": {"value": rate / cb["single_core_factor"],
                                "how": f"the step scaled by the 1-thread / {threads}-thread time of a "
                                       f"pressure-correction Krylov iteration ({cb['single_core_factor']:.2f})"},
                "host": host, "setup_s": cb["setup_s"],
            }
        else:
            cb = cpu_baseline(wl, threads)
            cpu = {
                "value": cb["iters_per_s"] * cb["cells"] / z.n_cell,
                "unit": "SIMPLE iterations/s (FP64 oracle; sample cells*iters/s divided by the GPU case's cells)",
                "cores": threads,
                "kind": "port",
                "sample": f"{cb['what']}, {cb['cells']} cells, first {cb['steps']} SIMPLE steps from rest in "
                          f"{cb['seconds']:.1f}s ({cb['iters_per_s']:.4f} it/s on the sample)",
                "host": host,
            }
        log(f"CPU baseline {cpu['value']:.4g} it/s: {cpu['sample']}")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "SIMPLE iterations/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": r["seconds"] * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_name == "mixed" else "f64",
        "data": "synthetic (seeded jittered hex/tet mesh generated in-process)",
        "config": {
            "workload": wl.describe(main_name), "autotune": 0 if shapes else wl.autotune, "graphs": wl.graphs,
            "cells": z.n_cell, "faces": z.n_face, "nnz": z.nnz,
            "cells_iters_per_s": value * z.n_cell,
            "precision": main_name,
            "assembly": "fp32" if main_name == "mixed" and wl.assembly_fp32 else "fp64",
        },
        "roofline": {
            "bound": "hbm", "kernel": dom["kernel"] + " on the pressure-correction matrix",
            "achieved": dom["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": dom["frac"],
            "traffic": MEASURED_TRAFFIC.get((dom["kernel"], z.n_cell)),
            "traffic_over_algorithmic": (MEASURED_TRAFFIC[(dom["kernel"], z.n_cell)] / dom["bytes_per_launch"]
                                         if (dom["kernel"], z.n_cell) in MEASURED_TRAFFIC else None),
            "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/hbm_traffic_channel_6M.md"
            if (dom["kernel"], z.n_cell) in MEASURED_TRAFFIC else None,
            "bytes_per_launch": dom["bytes_per_launch"], "ms_per_launch": dom["ms_per_launch"],
            "timing": dom["timing"],
            "measured_copy_gbs": copy_gbs, "frac_of_measured_copy": dom["achieved"] / copy_gbs,
        },
        "spmv": spmv,
        "paths": {
            n: {"iters_per_s": steps / v["seconds"], "ms_per_step": v["seconds"] * 1e3 / steps,
                "krylov_its_per_step": v["krylov_its_per_step"], "refinements_per_step": v["refinements_per_step"],
                "stage_ms_per_step": v["stage_ms_per_step"], "kernel_tuning": v["tuning"],
                "spmv_in_solve": v["spmv_in_solve"],
                "residual_history_last": v["history"][-1, [0, 1, 2, 3, 10]].tolist()}
            for n, v in runs.items()
        },
        "mixed_vs_fp64": agreement,
        "accuracy": accuracy,
        "cpu_baseline": cpu,
    }
    if "fp64" in runs and "mixed" in runs:
        out["mixed_speedup_vs_fp64"] = runs["fp64"]["seconds"] / runs["mixed"]["seconds"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
