"""Benchmark: 3D smoke-plume steps at 256^3 on one MI355X.

One step is the per-step grid path of mantaflow's plume scenes
(scenes/simpleplume.py, scenes/plume_2d.py) at 256^3:
  density source (cylinder, seeded per-cell density in [0.8, 1.0])
  advectSemiLagrange(density, order=2)      MacCormack, clampMode 2
  advectSemiLagrange(vel, order=2)          MacCormack on the MAC grid
  resetOutflow(real=density)                open top (outflow cells)
  setWallBcs ; addBuoyancy(gravity=(0,-1e-3,0))
  solvePressure(PcMGStatic)                 MG-preconditioned CG to
                                            ||r||_2/||b||_2 <= 1e-4
  setWallBcs
with every grid resident in HBM.  Value = cells * steps / wall time
(Mcell-steps/s) over exactly `steps` timed steps after `warmup` untimed ones,
bracketed by device synchronisation.

Printed as one JSON line with the dominant kernel's roofline figures (HIP
events on its stream during the timed region) and a CPU baseline (the CPU
oracle, a restatement of the reference's plugins, timed on this host's cores
on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s at 256³ (rel. res 1e-4); served solves/s + p50/p99 latency"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per cell of the instrumented kernels (FP32 values, uint8
# flags/matrix code, SoA velocity; each array element counted once per pass)
KERNEL_BYTES = {
    "cg_matvec": 17.0,         # read z 4, p 4, code 1; write p 4, q 4
    "cg_update": 24.0,         # read x, p, r, q; write x, r
    "mg_l0_down": 13.0,        # pre-smooth from 0 + residual: read r 4, code 1; write x 4, r0 4
    "mg_l0_up": 13.5,          # prolong + post-smooth: read r 4, x 4, code 1, coarse x 0.5; write z 4
    "advect_sl_mac": 24.0,     # read vel 12 (gathers cached); write fwd 12
    "advect_maccormack_mac": 37.0,  # read vel 12, fwd 12 (gathers), flags 1; write 12
}


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at n^3 from the committed PMC summary
    (profiles/plume256_traffic.md: FETCH_SIZE and WRITE_SIZE per dispatch,
    each scaled by the same counter's reading for a plain copy of known
    size), or None when no summary covers this kernel and size."""
    path = os.path.join(ROOT, "profiles", "plume256_traffic.md")
    if n != 256 or not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.strip().strip("|").split("|")
        if len(parts) >= 5 and parts[0].strip() == "`%s`" % kernel:
            try:
                return float(parts[4])
            except ValueError:
                return None
    return None


def plume_setup(manta, n, seed=1):
    s = manta.Solver(name="bench", gridSize=manta.vec3(n, n, n), dim=3)
    s.timestep = 1.0
    s.mfxRelTol = 1e-4
    flags = s.create(manta.FlagGrid)
    vel = s.create(manta.MACGrid)
    density = s.create(manta.RealGrid)
    pressure = s.create(manta.RealGrid)
    flags.initDomain(boundaryWidth=1)
    flags.fillGrid()
    manta.setOpenBound(flags, 1, "Y", manta.FlagOutflow | manta.FlagEmpty)
    # source: cylinder of radius 0.14 N at the bottom centre, density U[0.8, 1]
    src = manta.Cylinder(s, center=manta.vec3(0.5 * n, 0.1 * n, 0.5 * n), radius=0.14 * n,
                         z=manta.vec3(0, 0.02 * n, 0))
    mask = s.create(manta.RealGrid)
    src.applyToGrid(grid=mask, value=1.0)
    rng = np.random.default_rng(seed)
    val = rng.uniform(0.8, 1.0, n * n * n).astype(np.float32)
    srcval = s.create(manta.RealGrid)
    srcval.fromNumpy(val)
    srcval.mult(mask)                  # density inside the source
    keep = s.create(manta.RealGrid)
    keep.setConst(1.0)
    keep.sub(mask)                     # 1 outside the source, 0 inside
    return dict(s=s, flags=flags, vel=vel, density=density, pressure=pressure, keep=keep, srcval=srcval)


def plume_step(manta, st):
    f, v, d, p = st["flags"], st["vel"], st["density"], st["pressure"]
    d.mult(st["keep"])
    d.add(st["srcval"])
    manta.advectSemiLagrange(flags=f, vel=v, grid=d, order=2)
    manta.advectSemiLagrange(flags=f, vel=v, grid=v, order=2, strength=1.0)
    manta.resetOutflow(flags=f, real=d)
    manta.setWallBcs(flags=f, vel=v)
    manta.addBuoyancy(density=d, vel=v, gravity=manta.vec3(0, -1e-3, 0), flags=f)
    manta.solvePressure(flags=f, vel=v, pressure=p, cgAccuracy=1e-4, preconditioner=manta.PcMGStatic)
    manta.setWallBcs(flags=f, vel=v)
    return st["s"]._solve_stats


def _oracle_plume_step(o, shape, flags, keep, srcval, dens, vel, tol):
    """One plume step on the oracle (the bench's step order), in place."""
    from oracle import PC_MGSTATIC, TYPE_MAC, TYPE_REAL
    one = o.dtype(1.0)
    dens[:] = dens * keep + srcval
    o.advect(shape, one, flags, vel, dens, TYPE_REAL, 2, 1.0, 2)
    o.advect(shape, one, flags, vel, vel, TYPE_MAC, 2, 1.0, 2)
    o.reset_outflow(shape, flags, None, dens)
    o.set_wall_bcs(shape, flags, vel)
    o.add_buoyancy(shape, one, flags, dens, vel, (0, -1e-3, 0))
    st = o.solve_pressure(shape, vel, np.zeros(dens.shape[0], o.dtype), flags, cgAccuracy=min(1e-4, tol),
                          preconditioner=PC_MGSTATIC, relTol=tol)
    o.set_wall_bcs(shape, flags, vel)
    return st


def plume_step_error(manta, st, threads=16):
    """Error of one GPU plume step against the FP64 oracle's step from the
    same (downloaded) state, with the projection solved to relative residual
    1e-10, next to the CPU FP32 oracle's error for the same step at the
    bench's tolerance (1e-4): max |x - x64| / max |x64| for density and
    velocity."""
    from oracle import Oracle
    s = st["s"]
    n = st["flags"].getSizeX()
    shape = (n, n, n)
    flags = st["flags"].toNumpy()
    keep, srcval = st["keep"].toNumpy(), st["srcval"].toNumpy()
    d0, v0 = st["density"].toNumpy().copy(), st["vel"].toNumpy().copy()
    plume_step(manta, st)
    s.sync()
    gd, gv = st["density"].toNumpy(), st["vel"].toNumpy()
    out = {"grid": [n, n, n], "threads": threads}
    ref = None
    for f64, tol in ((True, 1e-10), (False, 1e-4)):
        o = Oracle(f64=f64)
        o.set_threads(threads)
        dens, vel = d0.astype(o.dtype), v0.astype(o.dtype)
        t0 = time.perf_counter()
        sto = _oracle_plume_step(o, shape, flags, keep.astype(o.dtype), srcval.astype(o.dtype), dens, vel, tol)
        dt = time.perf_counter() - t0
        if f64:
            ref = (dens, vel)
            out["fp64_oracle"] = {"rel_tol": tol, "s": round(dt, 2), "iterations": sto["iterations"]}
        else:
            d32 = dens
            out["cpu_fp32_oracle"] = {"rel_tol": tol, "s": round(dt, 2), "iterations": sto["iterations"],
                                      "density_err_vs_fp64": _rel_err(dens, ref[0]),
                                      "velocity_err_vs_fp64": _rel_err(vel, ref[1]),
                                      "density_err_vs_fp64_detail": _err_detail(dens, ref[0]),
                                      "velocity_err_vs_fp64_detail": _err_detail(vel, ref[1])}
    out["gpu"] = {"rel_tol": 1e-4, "density_err_vs_fp64": _rel_err(gd, ref[0]),
                  "velocity_err_vs_fp64": _rel_err(gv, ref[1]),
                  "density_err_vs_fp64_detail": _err_detail(gd, ref[0]),
                  "velocity_err_vs_fp64_detail": _err_detail(gv, ref[1]),
                  # advection and forces are bit-exact; the density does not
                  # pass through the pressure solve
                  "density_bitexact_vs_cpu_fp32_oracle": bool(np.array_equal(gd, d32))}
    return out


def cpu_baseline(n_sample=256, steps=3, threads=None, warm=2):
    """The CPU oracle (C restatement of the reference's plugins, OpenMP over
    the data-parallel loops) on the same plume step: by default the 256^3
    workload itself, timed over steps 3-5 of the plume (a step takes seconds
    there; --configs times the 128^3 plume over all 100 steps)."""
    from oracle import Oracle, PC_MGSTATIC, TYPE_MAC, TYPE_REAL
    o = Oracle()
    th = threads or min(16, os.cpu_count() or 1)
    o.set_threads(th)
    n = n_sample
    shape = (n, n, n)
    N = n ** 3
    flags = o.init_domain(shape, boundaryWidth=1)
    o.fill_grid(shape, flags)
    o.set_open_bound(shape, flags, 1, "Y")
    mask = np.zeros(N, np.float32)
    o.apply_shape(shape, 3, [0.5 * n, 0.1 * n, 0.5 * n], [0, 0.02 * n, 0], 0.14 * n, mask, TYPE_REAL, 1.0)
    val = np.random.default_rng(1).uniform(0.8, 1.0, N).astype(np.float32) * mask
    keep = (np.float32(1.0) - mask).astype(np.float32)
    vel = np.zeros((N, 3), np.float32)
    dens = np.zeros(N, np.float32)
    p = np.zeros(N, np.float32)

    def step():
        nonlocal dens
        dens = (dens * keep + val).astype(np.float32)
        o.advect(shape, np.float32(1.0), flags, vel, dens, TYPE_REAL, 2, 1.0, 2)
        o.advect(shape, np.float32(1.0), flags, vel, vel, TYPE_MAC, 2, 1.0, 2)
        o.reset_outflow(shape, flags, None, dens)
        o.set_wall_bcs(shape, flags, vel)
        o.add_buoyancy(shape, np.float32(1.0), flags, dens, vel, (0, -1e-3, 0))
        o.solve_pressure(shape, vel, p, flags, cgAccuracy=1e-4, preconditioner=PC_MGSTATIC, relTol=1e-4)
        o.set_wall_bcs(shape, flags, vel)

    for _ in range(warm):  # untimed: let the plume develop like the GPU warm-up
        step()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    dt = time.perf_counter() - t0
    return {"value": N * steps / dt / 1e6, "unit": "Mcell-steps/s", "cores": th, "kind": "port",
            "sample": "%d steps of the same plume step at %d^3 after %d untimed steps, CPU oracle "
                      "(OpenMP data-parallel loops, sequential MG coarsening and coarsest CG as in "
                      "the reference), %.1f s timed" % (steps, n, warm, dt)}


def bench_batched(manta, n2d=32, n3d=32):
    """64 uneven pressure problems: one batched solve vs the problems solved
    one after another on the plain path (same inputs, rel. residual 1e-4).
    The batch is solved ten times untimed (the plane-chunk autotuner's eight
    trials, then the chosen policy's queued launches and the capture of its
    two-iteration graph) and the median of three graph replays is
    reported."""
    from mantaflow_amd import workloads
    probs = workloads.batched_problems(n2d, n3d)
    kw = dict(cgAccuracy=1e-4, preconditioner=manta.PcMGStatic)
    for p in probs:  # build the static multigrid hierarchies
        manta.solvePressure(flags=p.flags, vel=p.vel, pressure=p.pressure, **kw)
    times = []
    for rep in range(13):
        for p in probs:
            p.reset()
            p.s.sync()
        t0 = time.perf_counter()
        manta.solvePressureBatch([p.problem() for p in probs], **kw)
        times.append(time.perf_counter() - t0)
    tfirst = times[0]
    tb = float(np.median(times[-3:]))
    its = [manta.lastSolveStats(p.s)["iterations"] for p in probs]
    rel = max(manta.lastSolveStats(p.s)["relResidual"] for p in probs)
    for p in probs:
        p.reset()
        p.s.sync()
    t0 = time.perf_counter()
    for p in probs:
        manta.solvePressure(flags=p.flags, vel=p.vel, pressure=p.pressure, **kw)
        p.s.sync()
    ts = time.perf_counter() - t0
    cells = sum(p.cells for p in probs)
    return {"problems": len(probs), "cells": cells, "batched_ms": round(tb * 1e3, 3),
            "batched_first_ms": round(tfirst * 1e3, 3),
            "batched_trials_ms": [round(t * 1e3, 3) for t in times[:8]],
            "serial_ms": round(ts * 1e3, 3), "batched_solves_per_s": round(len(probs) / tb, 1),
            "serial_solves_per_s": round(len(probs) / ts, 1), "speedup": round(ts / tb, 2),
            "iterations_max": int(max(its)), "rel_residual_max": rel}


def bench_projection(manta, n=256, frac=0.15, fp64=True, threads=16):
    """n^3 pressure projection: walls plus seeded boxes/spheres covering 15 %
    of the interior, face velocities N(0,1) smoothed by three binomial passes
    per axis, then setWallBcs.  Multigrid-preconditioned CG to relative
    residual 1e-4 and 1e-6, FP32 and mixed precision (BF16 coarse-level
    stencils), FP64 true residual recomputed every 50 iterations.  Reports
    time, iterations, the FP64 true relative residual, the velocity error
    against the FP64 oracle's solution of the same system (MG-PCG to relative
    residual 1e-10; max |u - u64| / max |u64|) next to the CPU FP32 oracle's
    error at the same tolerances, and the projection's roofline fraction at
    115 algorithmic bytes per cell per MG-PCG iteration (CG init counted as
    one iteration)."""
    from mantaflow_amd import workloads
    p = workloads.Problem((n, n, n), frac, seed=3)
    kw = dict(cgAccuracy=1e-4, preconditioner=manta.PcMGStatic)
    s = p.s
    s.mfxTrueResidualEvery = 50

    def solve(prec, tol):
        s.mfxPrecision = prec
        s.mfxRelTol = tol
        p.reset()
        s.sync()
        t0 = time.perf_counter()
        manta.solvePressure(flags=p.flags, vel=p.vel, pressure=p.pressure, **kw)
        s.sync()
        return time.perf_counter() - t0, manta.lastSolveStats(s)

    out = {"grid": [n, n, n], "obstacle_fraction": frac, "true_residual_every": 50, "runs": {}}
    ref = None
    if fp64:
        from oracle import Oracle, PC_MGSTATIC
        flags, vel0 = p.flags.toNumpy(), p.vel0
        o64 = Oracle(f64=True)
        o64.set_threads(threads)
        v64 = vel0.astype(np.float64)
        o64.set_wall_bcs((n, n, n), flags, v64)
        t0 = time.perf_counter()
        st64 = o64.solve_pressure((n, n, n), v64, np.zeros(n ** 3), flags, cgAccuracy=1e-10,
                                  preconditioner=PC_MGSTATIC, relTol=1e-10)
        ref = v64
        out["fp64_oracle"] = {"s": round(time.perf_counter() - t0, 2), "iterations": st64["iterations"],
                              "rel_residual": st64["relResidual"], "threads": threads}
        o32 = Oracle()
        o32.set_threads(threads)
        out["cpu_fp32_oracle"] = {}
        for tol in (1e-4, 1e-6):
            v32 = vel0.copy()
            o32.set_wall_bcs((n, n, n), flags, v32)
            t0 = time.perf_counter()
            # (bounded: in FP32 the reference's recursive residual may stall above 1e-6)
            st32 = o32.solve_pressure((n, n, n), v32, np.zeros(n ** 3, np.float32), flags, cgAccuracy=1e-4,
                                      preconditioner=PC_MGSTATIC, relTol=tol, maxIterOverride=40)
            out["cpu_fp32_oracle"]["%g" % tol] = {
                "s": round(time.perf_counter() - t0, 2), "iterations": st32["iterations"],
                "rel_residual": st32["relResidual"], "velocity_err_vs_fp64": _rel_err(v32, ref),
                "velocity_err_vs_fp64_detail": _err_detail(v32, ref)}
    else:
        s.mfxTrueResidualEvery = 10
        solve(0, 1e-7)
        ref = p.vel.toNumpy().copy()
        s.mfxTrueResidualEvery = 50
    for prec, name in ((0, "fp32"), (1, "mixed")):
        solve(prec, 1e-4)  # builds the hierarchy for this precision
        for tol in (1e-4, 1e-6):
            t, st = solve(prec, tol)
            its = int(st["iterations"])
            gbs = (its + 1) * 115.0 * n ** 3 / t / 1e9
            u = p.vel.toNumpy()
            out["runs"]["%s_%g" % (name, tol)] = {
                "ms": round(t * 1e3, 3), "iterations": its, "true_rel_residual": st["relResidual"],
                "converged": bool(st["converged"]),
                ("velocity_err_vs_fp64" if fp64 else "velocity_err_vs_1e-7"): _rel_err(u, ref),
                ("velocity_err_vs_fp64_detail" if fp64 else "velocity_err_vs_1e-7_detail"): _err_detail(u, ref),
                "achieved_GBs": round(gbs, 1), "roofline_frac": round(gbs / HBM_PEAK_GBS, 4)}
    s.mfxPrecision = 0
    return out


def _err_detail(a, b):
    """RMS error relative to the RMS of b, and the number of cells whose
    largest component error exceeds 1e-3 of max |b| (a max-norm error can
    sit in a handful of cells, e.g. one-cell-wide channels between
    obstacles)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    scale = max(1e-30, float(np.abs(b).max()))
    cell = d.max(axis=1) if d.ndim == 2 else d
    return {"rms": float(np.sqrt((d * d).mean()) / max(1e-30, np.sqrt((b * b).mean()))),
            "cells_over_1e-3": int((cell > 1e-3 * scale).sum())}


def _rel_err(a, b):
    return float(np.abs(np.asarray(a, np.float64) - np.asarray(b, np.float64)).max() /
                 max(1e-30, float(np.abs(np.asarray(b, np.float64)).max())))


def config1_plume2d(manta, steps=200, threads=16):
    """64^2 plume (walls, box source, first-order SL, buoyancy (0,-4e-3),
    plain CG at 1e-3): GPU, CPU oracle (FP32) and FP64 oracle, 200 steps;
    error of the GPU fields against the FP64 oracle after the last step."""
    from oracle import Oracle, PC_NONE, TYPE_MAC, TYPE_REAL
    shape, n = (64, 64, 1), 64 * 64
    s = manta.Solver(name="plume2d", gridSize=manta.vec3(64, 64, 1), dim=2)
    s.timestep = 1.0
    f, v, d, pr = s.create(manta.FlagGrid), s.create(manta.MACGrid), s.create(manta.RealGrid), s.create(manta.RealGrid)
    f.initDomain()
    f.fillGrid()
    src = manta.Box(s, p0=manta.vec3(25.6, 3.2, 0), p1=manta.vec3(38.4, 6.4, 1))
    its = []
    s.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        src.applyToGrid(grid=d, value=1.0)
        manta.advectSemiLagrange(flags=f, vel=v, grid=d, order=1)
        manta.advectSemiLagrange(flags=f, vel=v, grid=v, order=1)
        manta.setWallBcs(flags=f, vel=v)
        manta.addBuoyancy(density=d, vel=v, gravity=manta.vec3(0, -4e-3, 0), flags=f)
        manta.solvePressure(flags=f, vel=v, pressure=pr, cgAccuracy=1e-3, preconditioner=manta.PcNone)
        its.append(manta.lastSolveStats(s)["iterations"])
        s.step()
    s.sync()
    tg = time.perf_counter() - t0
    gd, gv = d.toNumpy(), v.toNumpy()
    out = {"gpu_ms_per_step": round(tg / steps * 1e3, 3), "cg_iterations_mean": float(np.mean(its))}
    for f64 in (False, True):
        o = Oracle(f64=f64)
        o.set_threads(threads)
        flags = o.init_domain(shape)
        o.fill_grid(shape, flags)
        vel, dens, p = np.zeros((n, 3), o.dtype), np.zeros(n, o.dtype), np.zeros(n, o.dtype)
        t0 = time.perf_counter()
        for _ in range(steps):
            o.apply_shape(shape, 1, [25.6, 3.2, 0], [38.4, 6.4, 1], 0, dens, TYPE_REAL, 1.0)
            o.advect(shape, 1.0, flags, vel, dens, TYPE_REAL, 1)
            o.advect(shape, 1.0, flags, vel, vel, TYPE_MAC, 1)
            o.set_wall_bcs(shape, flags, vel)
            o.add_buoyancy(shape, 1.0, flags, dens, vel, (0, -4e-3, 0))
            o.solve_pressure(shape, vel, p, flags, cgAccuracy=1e-3, preconditioner=PC_NONE)
        tc = time.perf_counter() - t0
        key = "fp64_oracle" if f64 else "cpu_oracle_fp32"
        out[key] = {"ms_per_step": round(tc / steps * 1e3, 3), "threads": threads,
                    "gpu_density_err": _rel_err(gd, dens), "gpu_velocity_err": _rel_err(gv, vel)}
    return out


def _plume_integrals(dens, vel, shape):
    """Total density, kinetic energy (half the sum of squared face
    velocities) and the density-weighted mean height (cells) of a plume."""
    d = np.asarray(dens, np.float64)
    v = np.asarray(vel, np.float64)
    tot = float(d.sum())
    j = (np.arange(d.size) // shape[0]) % shape[1]
    return {"density_total": tot, "kinetic_energy": 0.5 * float((v * v).sum()),
            "plume_height": float((d * j).sum() / tot) if tot > 0 else 0.0}


def config2_plume3d(manta, n=128, steps=100, threads=16, record=(1, 10, 25, 50, 75, 100)):
    """128^3 open-top plume, MacCormack, MG-static PCG at rel. res. 1e-4, 100
    steps on the GPU, the CPU oracle (FP32) and the FP64 oracle from the same
    initial state.  Reports the GPU time per step, the oracle times, the
    pairwise field errors (max |a - b| / max |b|) of FP32 oracle vs FP64
    oracle, GPU vs FP64 oracle and GPU vs FP32 oracle at the recorded steps,
    and per-step integrals (total density, kinetic energy, plume height) of
    the three runs with the largest relative spread between them."""
    from oracle import Oracle, PC_MGSTATIC, TYPE_MAC, TYPE_REAL
    st = plume_setup(manta, n)
    s = st["s"]
    its, snaps, integ = [], {}, {"gpu": [], "cpu_oracle_fp32": [], "fp64_oracle": []}
    s.sync()
    tg = 0.0
    for k in range(1, steps + 1):
        t0 = time.perf_counter()
        its.append(plume_step(manta, st)["iterations"])
        s.sync()
        if k > 1:
            tg += time.perf_counter() - t0
        gd, gv = st["density"].toNumpy(), st["vel"].toNumpy()
        integ["gpu"].append(_plume_integrals(gd, gv, (n, n, n)))
        if k in record:
            snaps[("gpu", k)] = (gd.copy(), gv.copy())
    out = {"gpu_ms_per_step": round(tg / (steps - 1) * 1e3, 3),
           "gpu_mcell_steps_per_s": round(n ** 3 * (steps - 1) / tg / 1e6, 1),
           "cg_iterations_mean": float(np.mean(its))}
    shape, N = (n, n, n), n ** 3
    for f64 in (False, True):
        key = "fp64_oracle" if f64 else "cpu_oracle_fp32"
        o = Oracle(f64=f64)
        o.set_threads(threads)
        flags = o.init_domain(shape, boundaryWidth=1)
        o.fill_grid(shape, flags)
        o.set_open_bound(shape, flags, 1, "Y")
        mask = np.zeros(N, o.dtype)
        o.apply_shape(shape, 3, [0.5 * n, 0.1 * n, 0.5 * n], [0, 0.02 * n, 0], 0.14 * n, mask, TYPE_REAL, 1.0)
        val = (np.random.default_rng(1).uniform(0.8, 1.0, N).astype(np.float32) * mask.astype(np.float32)).astype(o.dtype)
        keep = (1.0 - mask).astype(o.dtype)
        vel, dens = np.zeros((N, 3), o.dtype), np.zeros(N, o.dtype)
        t0 = time.perf_counter()
        for k in range(1, steps + 1):
            _oracle_plume_step(o, shape, flags, keep, val, dens, vel, 1e-4)
            integ[key].append(_plume_integrals(dens, vel, shape))
            if k in record:
                snaps[(key, k)] = (dens.copy(), vel.copy())
        tc = time.perf_counter() - t0
        out[key] = {"ms_per_step": round(tc / steps * 1e3, 1), "threads": threads,
                    "mcell_steps_per_s": round(N * steps / tc / 1e6, 2)}
    errs = {}
    for a, b in (("cpu_oracle_fp32", "fp64_oracle"), ("gpu", "fp64_oracle"), ("gpu", "cpu_oracle_fp32")):
        errs["%s_vs_%s" % (a, b)] = {str(k): {"density": _rel_err(snaps[(a, k)][0], snaps[(b, k)][0]),
                                              "velocity": _rel_err(snaps[(a, k)][1], snaps[(b, k)][1])}
                                     for k in record}
    out["field_errors"] = errs
    spread = {}
    for q in ("density_total", "kinetic_energy", "plume_height"):
        worst = 0.0
        for k in range(steps):
            vals = [integ[r][k][q] for r in integ]
            ref = abs(integ["fp64_oracle"][k][q])
            if ref > 0:
                worst = max(worst, (max(vals) - min(vals)) / ref)
        spread[q] = worst
    out["integrals"] = {r: {str(k): integ[r][k - 1] for k in record} for r in integ}
    out["integrals_max_rel_spread"] = spread
    return out


def config3_fp64(manta, n=256, frac=0.15, threads=16):
    """The 256^3 projection problem of the projection leg against the FP64
    oracle (MG-static PCG to relative residual 1e-10): GPU error at 1e-4 and
    1e-6, FP32 and mixed precision; the CPU oracle's FP32 solve time."""
    from oracle import Oracle, PC_MGSTATIC
    from mantaflow_amd import workloads
    shape = (n, n, n)
    rng = np.random.default_rng(3)
    flags = workloads.obstacle_flags_np(shape, frac, rng)
    vel0 = workloads.smooth_field(rng, shape)
    o = Oracle(f64=True)
    o.set_threads(threads)
    vref = vel0.astype(np.float64)
    o.set_wall_bcs(shape, flags, vref)
    pref = np.zeros(n ** 3)
    t0 = time.perf_counter()
    sref = o.solve_pressure(shape, vref, pref, flags, cgAccuracy=1e-10, preconditioner=PC_MGSTATIC, relTol=1e-10)
    tref = time.perf_counter() - t0
    o32 = Oracle()
    o32.set_threads(threads)
    v32 = vel0.copy()
    o32.set_wall_bcs(shape, flags, v32)
    p32 = np.zeros(n ** 3, np.float32)
    t0 = time.perf_counter()
    s32 = o32.solve_pressure(shape, v32, p32, flags, cgAccuracy=1e-4, preconditioner=PC_MGSTATIC, relTol=1e-4)
    t32 = time.perf_counter() - t0
    p = workloads.Problem(shape, frac, seed=3)
    out = {"fp64_oracle": {"s": round(tref, 1), "iterations": sref["iterations"], "rel_residual": sref["relResidual"]},
           "cpu_oracle_fp32_1e-4": {"s": round(t32, 2), "iterations": s32["iterations"], "threads": threads}, "gpu": {}}
    for prec, name in ((0, "fp32"), (1, "mixed")):
        for tol in (1e-4, 1e-6):
            p.s.mfxPrecision, p.s.mfxRelTol, p.s.mfxTrueResidualEvery = prec, tol, 50
            p.reset()
            manta.solvePressure(flags=p.flags, vel=p.vel, pressure=p.pressure, cgAccuracy=1e-4,
                                preconditioner=manta.PcMGStatic)
            out["gpu"]["%s_%g" % (name, tol)] = {"velocity_err": _rel_err(p.vel.toNumpy(), vref),
                                                 "iterations": manta.lastSolveStats(p.s)["iterations"]}
    return out


def config4_cpu(manta, threads=16):
    """The 64 batched-solve problems solved one after another by the CPU
    oracle (FP32, MG-static PCG at 1e-4).  Every problem is a closed domain
    (walls on all sides), a singular system: the reference's GridMg warns
    that it may not converge without a fixed pressure value
    (multigrid.cpp:401-402), so the oracle solves with zeroPressureFixing."""
    from oracle import Oracle, PC_MGSTATIC
    from mantaflow_amd import workloads
    shapes, fracs = workloads.batched_shapes()
    o = Oracle()
    o.set_threads(threads)
    tot, tconv, its, worst, failed = 0.0, 0.0, [], 0.0, []
    for i, (sh, fr) in enumerate(zip(shapes, fracs)):
        rng = np.random.default_rng(1000 + i)
        flags = workloads.obstacle_flags_np(sh, fr, rng)
        vel = workloads.smooth_field(rng, sh)
        o.set_wall_bcs(sh, flags, vel)
        pr = np.zeros(sh[0] * sh[1] * sh[2], np.float32)
        t0 = time.perf_counter()
        st = o.solve_pressure(sh, vel, pr, flags, cgAccuracy=1e-4, preconditioner=PC_MGSTATIC, relTol=1e-4,
                              zeroPressureFixing=True)
        dt = time.perf_counter() - t0
        tot += dt
        if st["relResidual"] <= 1e-4:
            its.append(st["iterations"])
            worst = max(worst, st["relResidual"])
            tconv += dt
        else:
            # a single fixed cell leaves fluid pockets enclosed by obstacles
            # singular: the reference's multigrid stalls there too
            failed.append({"problem": i, "shape": list(sh), "s": round(dt, 2), "rel_residual": st["relResidual"]})
    return {"problems": len(shapes), "cpu_oracle_serial_ms": round(tot * 1e3, 1), "threads": threads,
            "converged": len(its), "converged_serial_ms": round(tconv * 1e3, 1),
            "iterations_max_converged": int(max(its)), "rel_residual_max_converged": worst,
            "not_converged": failed, "zero_pressure_fixing": True}


def bench_configs(manta):
    out = {}
    for name, fn in (("config1_plume2d_64", config1_plume2d), ("config2_plume3d_128", config2_plume3d),
                     ("config3_projection_256_fp64", config3_fp64), ("config4_batched_cpu", config4_cpu)):
        print("bench --configs: %s ..." % name, file=sys.stderr, flush=True)
        try:
            out[name] = fn(manta)
        except Exception as e:
            out[name] = {"error": "%s: %s" % (type(e).__name__, e)}
    return out


def bench_serving(n_sessions=24, n_req=240, rate=200.0):
    """Poisson stream of plume-step / pressure-solve requests on resident
    sessions (64^2 .. 192^3 cells, log-uniform), both through the library's
    serving engine: served (continuous batching, up to four ticks in flight on
    worker threads) vs the plain path one request at a time; in real time at
    `rate` req/s
    (latency) and with every request queued at t=0 (throughput).  In real
    time each request's slowdown is its latency over the isolated latency of
    the same request (same session and kind, plain path, nothing else
    running); p50/p99 of the slowdown are reported beside the latencies."""
    from mantaflow_amd import serve
    sessions = [serve.PlumeScene(sh, seed=20 + i) for i, sh in enumerate(serve.make_sessions(n_sessions))]
    for sess in sessions:  # warm: multigrid hierarchies, pools
        sess.run_single(serve.REQ_STEP)
    trace = serve.make_trace(n_req, n_sessions, rate=rate)
    # isolated latency of every (session, kind): median of three runs of the
    # request alone through the engine's plain path (nothing else running)
    iso = {}
    for i, sess in enumerate(sessions):
        for kind in (serve.REQ_STEP, serve.REQ_SOLVE):
            ts = []
            for _ in range(3):
                lat, _ = serve.serve_serial(sessions, [serve.Request(0, 0.0, i, kind)], realtime=False)
                ts.append(lat[0])
            iso[(i, kind)] = float(np.median(ts))
    single = [iso[(i, serve.REQ_SOLVE)] for i in range(min(8, n_sessions))]
    out = {"sessions": n_sessions, "requests": n_req, "rate_req_per_s": rate,
           "isolated_solve_ms_median": round(float(np.median(single)) * 1e3, 3)}
    by_rid = {r.rid: r for r in trace}

    def slowdown(lat):
        # latency of each request over the isolated latency of the same request
        sl = np.sort([lat[rid] / iso[(by_rid[rid].session, by_rid[rid].kind)] for rid in lat])
        return {"slowdown_p50": round(float(np.percentile(sl, 50)), 3),
                "slowdown_p99": round(float(np.percentile(sl, 99)), 3)}

    for name, realtime in (("realtime", True), ("saturated", False)):
        srv = serve.Server(sessions, workers=4)
        lat, span = srv.run(trace, realtime=realtime)
        out["served_" + name] = {k: round(v, 3) for k, v in serve.summarize(lat, span).items()}
        out["served_" + name]["mean_batch"] = round(float(np.mean(srv.batch_sizes)), 2)
        if realtime:
            out["served_" + name].update(slowdown(lat))
        lat, span = serve.serve_serial(sessions, trace, realtime=realtime)
        out["serial_" + name] = {k: round(v, 3) for k, v in serve.summarize(lat, span).items()}
        if realtime:
            out["serial_" + name].update(slowdown(lat))
    out["throughput_speedup_saturated"] = round(out["served_saturated"]["solves_per_s"] /
                                                out["serial_saturated"]["solves_per_s"], 2)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-res", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-problem batched-solve leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the served request-stream leg")
    ap.add_argument("--configs", action="store_true",
                    help="instead of the bench line: GPU, CPU-oracle and FP64-oracle figures of the 2D plume "
                         "(64^2, 200 steps), the 3D plume (128^3, 100 steps), the 256^3 projection and the "
                         "64 batched problems, as one JSON object (a few minutes)")
    ap.add_argument("--no-projection", action="store_true",
                    help="skip the obstacle-rich projection leg (FP32 and mixed precision, rel. res. 1e-4/1e-6)")
    ap.add_argument("--no-developed", action="store_true",
                    help="skip the developed-plume window (10 steps timed after step 100)")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the FP64-oracle error legs (one plume step, the projection)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run 5 plain 64 MiB device copies first (PMC byte-counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build runs the grid path on a single MI355X "
              "(multi-GPU sharding is a design note only); use --gpus 1", file=sys.stderr)
        return 2

    from mantaflow_amd import manta
    from mantaflow_amd._abi import lib
    import ctypes as C

    if a.configs:
        print(json.dumps(bench_configs(manta)), flush=True)
        return 0

    n = a.res
    st = plume_setup(manta, n)
    s = st["s"]
    if a.calibrate_copy:
        src = s.create(manta.RealGrid)
        dst = s.create(manta.RealGrid)
        src.setConst(1.0)
        for _ in range(5):
            dst.copyFrom(src)  # n^3 floats read + written (64 MiB each at 256^3)
        s.sync()
    for _ in range(a.warmup):
        plume_step(manta, st)
    s.sync()
    lib().mfx_kernel_timing(1)
    lib().mfx_kernel_timing_reset()
    iters = []
    rel = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        stats = plume_step(manta, st)
        iters.append(stats["iterations"])
        rel.append(stats["relResidual"])
    s.sync()
    t1 = time.perf_counter()
    lib().mfx_kernel_timing(0)
    wall = t1 - t0
    cells = n ** 3
    value = cells * a.steps / wall / 1e6

    # roofline: the instrumented kernel with the largest total time
    best = None
    kstats = {}
    for name, bpc in KERNEL_BYTES.items():
        ms, cnt = C.c_double(), C.c_longlong()
        lib().mfx_kernel_timing_get(name.encode(), C.byref(ms), C.byref(cnt))
        if cnt.value == 0:
            continue
        avg = ms.value / cnt.value
        kstats[name] = {"launches": cnt.value, "avg_ms": avg, "total_ms": ms.value}
        if best is None or ms.value > kstats[best]["total_ms"]:
            best = name
    roof = None
    if best:
        avg_s = kstats[best]["avg_ms"] / 1e3
        algo = KERNEL_BYTES[best] * cells
        achieved = algo / avg_s / 1e9
        roof = {"kernel": best, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(best, n),
                "algo_bytes_per_launch": algo, "avg_launch_ms": round(kstats[best]["avg_ms"], 4)}

    # a developed plume: continue untimed to step 100, then time 10 steps
    developed = None
    if not a.no_developed:
        done = a.warmup + a.steps
        for _ in range(max(0, 100 - done)):
            plume_step(manta, st)
        s.sync()
        dits = []
        t0 = time.perf_counter()
        for _ in range(10):
            dits.append(plume_step(manta, st)["iterations"])
        s.sync()
        tdev = time.perf_counter() - t0
        first = max(done, 100) + 1
        developed = {"steps": [first, first + 9], "ms_per_step": round(tdev / 10 * 1e3, 3),
                     "mcell_steps_per_s": round(cells * 10 / tdev / 1e6, 1),
                     "cg_iterations_mean": float(np.mean(dits))}

    batched = served = proj = steperr = None
    if not a.no_errors:
        print("bench: one plume step against the FP64 oracle ...", file=sys.stderr, flush=True)
        try:
            steperr = plume_step_error(manta, st)
        except Exception as e:
            steperr = {"error": str(e)}
    if not a.no_projection:
        print("bench: 256^3 projection ...", file=sys.stderr, flush=True)
        try:
            proj = bench_projection(manta, n, fp64=not a.no_errors)
        except Exception as e:
            proj = {"error": str(e)}
    if not a.no_batched:
        print("bench: batched solves ...", file=sys.stderr, flush=True)
        try:
            batched = bench_batched(manta)
        except Exception as e:
            batched = {"error": str(e)}
    if not a.no_serving:
        print("bench: serving ...", file=sys.stderr, flush=True)
        try:
            served = bench_serving()
        except Exception as e:
            served = {"error": str(e)}

    cpu = None
    if not a.no_cpu_baseline:
        print("bench: CPU baseline ...", file=sys.stderr, flush=True)
        try:
            cpu = cpu_baseline(a.cpu_res, a.cpu_steps)
        except Exception as e:  # the baseline must not hide the GPU number
            cpu = {"error": str(e)}

    mvc, l0c = C.c_int(), C.c_int()
    tq = lib().mfx_testing_zmarch_chunks
    tq.restype, tq.argtypes = None, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    tq(n, n, n, C.byref(mvc), C.byref(l0c))
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(wall / a.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3D smoke plume step %d^3: MacCormack advection (density + MAC velocity), "
                               "open top, wall BCs, buoyancy, MG-preconditioned CG projection to rel. "
                               "residual 1e-4" % n,
                   "grid": [n, n, n], "step_window": [a.warmup + 1, a.warmup + a.steps],
                   "cg_iterations_mean": float(np.mean(iters)),
                   "rel_residual_max": float(np.max(rel)), "parallelism": "single GPU",
                   "zmarch_plane_chunks_autotuned": {"cg_matvec": mvc.value, "mg_level0": l0c.value}},
        "roofline": roof,
        "kernels": kstats,
        "developed_plume": developed,
        "step_error": steperr,
        "projection": proj,
        "batched_solves": batched,
        "serving": served,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
