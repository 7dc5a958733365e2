#!/usr/bin/env python3
"""Benchmark of the batched propagator on one MI355X.

Default workload "ck_grid256": adaptive Cash-Karp (PropagationCK, tolerance
1e-4, minStep 1 kpc, maxStep 1 Mpc) of 2^20 particles in a periodic 256^3 FP32
Kolmogorov Grid3f (spacing 20 kpc, Brms 1 nG, lMin 2 cells, lMax 128 cells,
seed 7), species Z in {1,2,7,14,26} with weights {.5,.2,.15,.1,.05}, E
log-uniform 1-300 EeV, isotropic directions from the box centre, each particle
stopped by MaximumTrajectoryLength at 100 Mpc.

Other workloads (--workload): bp_grid64 (4096 protons, 2000 fixed 0.1 Mpc
Boris steps, 64^3 grid of 50 kpc), bp_planewave (2^20 protons 1e15-1e18 eV,
fixed 0.05 pc Boris steps in PlaneWaveTurbulence with Nm=256, Brms 1 muG,
lMin 1 pc, lMax 100 pc, seed 11, 10000 steps), bp_grid512 / ck_grid512 (2^21
protons, 1000 fixed 20 kpc steps in a 512^3 grid of 10 kpc, seed 3),
ck_jf12 (the reference notebook's Galactic case: 2^16 antiprotons of 1 EeV
from Earth (-8.5, 0, 0) kpc, isotropic, fixed 0.1 pc Cash-Karp steps through
JF12Field with randomStriated(42) and randomTurbulent(7) (256^3 at 4 pc),
20000 steps), sde_grid256 (DiffusionSDE: 2^20 protons 1 TeV-1 PeV from the centre of a
periodic 256^3 Kolmogorov grid of 1 pc (lMin 2 cells, lMax 128 cells, seed 7),
tolerance 1e-4, minStep 0.1 pc, maxStep 10 pc, epsilon 0.1, alpha 1/3,
scale 1, Philox seed 11, MaximumTrajectoryLength 1 kpc), and
adjoint_grid128 (forward + backward of the differentiable Boris model: 65536
protons 1-100 EeV, 4096 steps of 10 kpc through a 128^3 grid of 20 kpc, seed
5, checkpoints every 64 steps; loss = sum |u_final - target|^2).

One step = propagate the whole batch (reset to its initial state) to the end
of the workload.  The metric is accepted particle-steps (process() calls) per
second with inputs resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

pc = 648000 / math.pi * 149597870700.0
kpc = 1e3 * pc
Mpc = 1e6 * pc
eplus = 1.602176487e-19
eV = eplus
EeV = 1e18 * eplus
nG = 1e-13
muG = 1e-10

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BYTES_PER_EVAL = 96  # 8 corners x 3 x FP32 of one trilinear Grid3f evaluation
BYTES_PER_EVAL_FP16 = 52  # 8 corners x 3 x FP16 + one FP32 brick scale
FLOP_PER_MODE = 16   # dot 5, phase 2, xi*A 3, *cos 3, accumulate 3 (cos counted separately)
FLOP_BORIS = 60

WORKLOADS = {
    "ck_grid256": dict(kind="CK", n=2 ** 20, field="grid", grid_n=256, spacing=20 * kpc, lmin_cells=2,
                       lmax_cells=128, seed=7, tolerance=1e-4, min_step=1 * kpc, max_step=1 * Mpc,
                       max_length=100 * Mpc, emin=1 * EeV, emax=300 * EeV, species=(1, 2, 7, 14, 26),
                       weights=(0.5, 0.2, 0.15, 0.1, 0.05), start="centre"),
    "bp_grid64": dict(kind="BP", n=4096, field="grid", grid_n=64, spacing=50 * kpc, lmin_cells=2, lmax_cells=32,
                      seed=42, tolerance=0.42, min_step=0.1 * Mpc, max_step=0.1 * Mpc, max_length=0.0, steps=2000,
                      emin=1 * EeV, emax=100 * EeV, species=(1,), weights=(1.0,), start="centre"),
    "bp_planewave": dict(kind="BP", n=2 ** 20, field="planewave", nm=256, brms=1 * muG, lmin=1 * pc, lmax=100 * pc,
                         seed=11, tolerance=0.42, min_step=0.05 * pc, max_step=0.05 * pc, max_length=0.0,
                         steps=10000, emin=1e15 * eV, emax=1e18 * eV, species=(1,), weights=(1.0,), start="origin"),
    "bp_grid512": dict(kind="BP", n=2 ** 21, field="grid", grid_n=512, spacing=10 * kpc, lmin_cells=2,
                       lmax_cells=256, seed=3, tolerance=0.42, min_step=20 * kpc, max_step=20 * kpc,
                       max_length=0.0, steps=1000, emin=1 * EeV, emax=100 * EeV, species=(1,), weights=(1.0,),
                       start="centre"),
    "sde_grid256": dict(kind="SDE", n=2 ** 20, field="grid", grid_n=256, spacing=1 * pc, lmin_cells=2,
                        lmax_cells=128, seed=7, tolerance=1e-4, min_step=0.1 * pc, max_step=10 * pc,
                        max_length=1 * kpc, epsilon=0.1, alpha=1 / 3., scale=1.0, sde_seed=11,
                        emin=1e12 * eV, emax=1e15 * eV, species=(1,), weights=(1.0,), start="centre"),
    "ck_jf12": dict(kind="CK", n=2 ** 16, field="jf12", striated_seed=42, turb_seed=7, tolerance=1e-4,
                    min_step=0.1 * pc, max_step=0.1 * pc, max_length=0.0, steps=20000, emin=1 * EeV, emax=1 * EeV,
                    species=(-1,), weights=(1.0,), start="earth", seed=13),
    "ck_grid512": dict(kind="CK", n=2 ** 21, field="grid", grid_n=512, spacing=10 * kpc, lmin_cells=2,
                       lmax_cells=256, seed=3, tolerance=1e-4, min_step=20 * kpc, max_step=20 * kpc,
                       max_length=0.0, steps=1000, emin=1 * EeV, emax=100 * EeV, species=(1,), weights=(1.0,),
                       start="centre"),
}


def draw_particles(seed, n, emin, emax, weights=(1.0,)):
    """Species bins, energies and raw directions as a reference steering script draws them: one
    crpropa Random seeded with `seed`; per particle randBin over the cumulative weights (when there
    is more than one species), randPowerLaw(-1, emin, emax), randVector (src/Random.cpp:129-146,
    195-198)."""
    from crpropa3_amd import Random
    r = Random(int(seed))
    cdf = list(np.cumsum(np.asarray(weights, dtype=np.float64)))
    sp = np.zeros(n, dtype=np.int64)
    E = np.empty(n)
    d = np.empty((n, 3))
    for i in range(n):
        if len(cdf) > 1:
            sp[i] = r.randBin(cdf)
        E[i] = r.randPowerLaw(-1.0, emin, emax)
        v = r.randVector()
        d[i] = (v.x, v.y, v.z)
    return sp, E, d


def make_particles(w, n):
    """Start positions, unit directions, energies and charges of a workload's first n particles, drawn
    with the workload's seed (draw_particles); directions normalised as ParticleState::setDirection."""
    sp, E, d = draw_particles(w["seed"], n, w["emin"], w["emax"], w["weights"])
    Z = np.asarray(w["species"], dtype=np.float64)[sp]
    d /= np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])[:, None]
    if w["start"] == "earth":
        pos = np.tile([-8.5 * kpc, 0.0, 0.0], (n, 1))
    elif w["start"] == "centre":
        L = w["grid_n"] * w["spacing"]
        pos = np.tile([L / 2, L / 2, L / 2], (n, 1))
    else:
        pos = np.zeros((n, 3))
    return pos, d, E, Z * eplus


def make_field_data(w):
    """Grid values (numpy) or plane-wave mode table (dict) of the workload."""
    if w["field"] == "jf12":
        from crpropa3_amd import JF12Field
        f = JF12Field()
        f.randomStriated(w["striated_seed"])
        f.randomTurbulent(w["turb_seed"])
        return {"striated": np.array(f.getStriatedGrid().array(), copy=True),
                "turbulent": np.array(f.getTurbulentGrid().array(), copy=True)}
    if w["field"] == "grid":
        from crpropa3_amd.turbulence import simple_grid_turbulence
        return simple_grid_turbulence(w["grid_n"], w["spacing"], 1 * nG, w["lmin_cells"] * w["spacing"],
                                      w["lmax_cells"] * w["spacing"], w["seed"])
    from crpropa3_amd import PlaneWaveTurbulence, SimpleTurbulenceSpectrum
    pw = PlaneWaveTurbulence(SimpleTurbulenceSpectrum(w["brms"], w["lmin"], w["lmax"]), w["nm"], w["seed"])
    return {"k": np.array(pw.getK()), "Ak": np.array(pw.getAk()), "beta": np.array(pw.getBeta()),
            "xi": np.array([[v.x, v.y, v.z] for v in pw.getXi()]),
            "kappa": np.array([[v.x, v.y, v.z] for v in pw.getKappa()])}


def host_threads():
    """OpenMP threads for the CPU baseline: every CPU this process may run on, capped by
    OMP_NUM_THREADS when the environment sets it (the host's share of CPUs for this job)."""
    avail = len(os.sched_getaffinity(0))
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(env, avail) if env > 0 else avail), avail


def cpu_baseline(w, fdata, pos, d, E, q, gpu_final=None, target_s=15.0):
    """The oracle (CPU restatement with glibc libm, OpenMP) on a bounded sample.  With `gpu_final`
    (the GPU pass's final states) the sample is also a parity check: the GPU against the oracle with
    glibc (the reference's arithmetic) and, on its first particles, bit for bit against the oracle
    with the correctly rounded functions the kernels use."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as O
    threads, avail = host_threads()
    if w["field"] == "grid":
        f = O.OracleField.grid(fdata, [0, 0, 0], [w["spacing"]] * 3)
    elif w["field"] == "jf12":
        jp = O.jf12_defaults(False)
        jp.use_striated = jp.use_turbulent = 1
        f = O.OracleField.jf12(jp, fdata["striated"], 0.1 * kpc, fdata["turbulent"], 4 * pc)
    else:
        f = O.OracleField.planewave(fdata)
    kind = O.PROP_CK if w["kind"] == "CK" else O.PROP_BP
    steps = w.get("steps", 10 ** 12)

    def run(m, nsteps):
        st = O.OracleState(pos[:m], d[:m], E[:m], q[:m], normalised=True)
        t0 = time.perf_counter()
        if w["kind"] == "SDE":
            # the reference's threading: one MT19937 per OpenMP thread, glibc libm
            nacc = O.sde_propagate_threads(f, st, min(nsteps, 2 ** 62), w["tolerance"], w["min_step"], w["max_step"],
                                           epsilon=w["epsilon"], alpha=w["alpha"], scale=w["scale"],
                                           max_len=w["max_length"], seed=w["sde_seed"], nthreads=threads)
        else:
            nacc, _ = O.propagate(f, kind, st, nsteps, tolerance=w["tolerance"], min_step=w["min_step"],
                                  max_step=w["max_step"], max_len=w["max_length"], libm=O.LIBM_GLIBC,
                                  nthreads=threads)
        return nacc, time.perf_counter() - t0, st

    m = min(len(E), 16 * threads)
    nsteps = steps
    nacc, dt, st = run(m, nsteps)
    if dt > target_s and steps < 10 ** 12:  # too slow even for the calibration: fewer steps
        nsteps = max(1, int(steps * target_s / dt / 4))
        nacc, dt, st = run(m, nsteps)
    while dt < target_s / 8 and m < len(E):  # grow the sample until every thread has a share of it
        m = min(len(E), 4 * m)
        nacc, dt, st = run(m, nsteps)
    m2 = int(min(len(E), max(m, m * target_s / max(dt, 1e-3))))
    if m2 > m:
        nacc, dt, st = run(m2, nsteps)
        m = m2
    reps = 1
    while dt * reps < 1.0 and reps < 64:  # whole workload too short to time once: repeat it
        reps *= 2
    if reps > 1:
        t = 0.0
        for _ in range(reps):
            nacc, d1, st = run(m, nsteps)
            t += d1
        nacc, dt = nacc * reps, t
    what = "run to completion" if nsteps >= steps else "%d of %d steps" % (nsteps, steps)
    if reps > 1:
        what += ", %d times" % reps
    out = {"value": nacc / dt, "unit": "particle-steps/s", "cores": threads, "host_cpus": avail, "kind": "port",
           "sample": "first %d of the %d particles, same field and propagator, %s (%d accepted steps in %.1f s)"
                     % (m, len(E), what, nacc, dt)}
    if gpu_final is not None and w["kind"] != "SDE":
        if nsteps < steps:
            out["parity"] = "not compared: the sample ran fewer steps than the GPU pass"
        else:
            rel = np.max(np.abs(gpu_final["pos"][:m] - st.pos), axis=1) / np.max(np.abs(st.pos), axis=1)
            k = min(m, 4096)
            cr = O.OracleState(pos[:k], d[:k], E[:k], q[:k], normalised=True)
            O.propagate(f, kind, cr, nsteps, tolerance=w["tolerance"], min_step=w["min_step"],
                        max_step=w["max_step"], max_len=w["max_length"], libm=O.LIBM_CR, nthreads=threads)
            exact = np.ones(k, dtype=bool)
            for key in ("pos", "dir", "trajectory_length", "next_step"):
                a = np.ascontiguousarray(gpu_final[key][:k]).view(np.uint64).reshape(k, -1)
                b = np.ascontiguousarray(getattr(cr, key)).view(np.uint64).reshape(k, -1)
                exact &= np.all(a == b, axis=1)
            out["parity"] = {"vs_oracle_correctly_rounded": {"particles": k, "bit_exact": int(exact.sum())},
                             "vs_oracle_glibc": {"particles": m, "within_1e-9_relative": float(np.mean(rel <= 1e-9)),
                                                 "median_relative": float(np.median(rel))}}
    return out


ADJOINT = dict(n=65536, grid_n=128, spacing=20 * kpc, lmin_cells=2, lmax_cells=128, seed=5, step=10 * kpc,
               steps=4096, K=64, emin=1 * EeV, emax=100 * EeV)
# per step of the backward: recompute (the cell's 128-byte corner line, 48 B of sub-checkpoint per 8
# steps written), reverse (the cell line again, the sub-checkpoint read, the 32-byte record written into
# its tile's range), accumulation (the record read)
ADJ_BYTES_PER_STEP = (128 + 6) + (128 + 6 + 32) + 32


def adjoint_inputs(a, n):
    from crpropa3_amd.turbulence import simple_grid_turbulence
    grid = simple_grid_turbulence(a["grid_n"], a["spacing"], 1 * nG, a["lmin_cells"] * a["spacing"],
                                  a["lmax_cells"] * a["spacing"], a["seed"]).astype(np.float32)
    # particles as draw_particles with the grid's seed, then the target directions from the next
    # seed (randVector, normalised)
    _, E, d = draw_particles(a["seed"], n, a["emin"], a["emax"])
    _, _, tgt = draw_particles(a["seed"] + 1, n, a["emin"], a["emax"])
    d /= np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])[:, None]
    tgt /= np.sqrt((tgt[:, 0] * tgt[:, 0] + tgt[:, 1] * tgt[:, 1]) + tgt[:, 2] * tgt[:, 2])[:, None]
    L = a["grid_n"] * a["spacing"]
    pos = np.tile([L / 2] * 3, (n, 1))
    return grid, pos, d, E, tgt


def adjoint_cpu_baseline(a, grid, pos, d, E, tgt, target_s=15.0):
    """The oracle's FP64 reference gradient (serial, one core) on the first m particles."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as O
    kq = eplus * 2.99792458e8 * a["step"] / (2 * E)
    o, sp = np.zeros(3), np.full(3, a["spacing"])

    def run(m):
        t0 = time.perf_counter()
        xo, uo = O.adjoint_forward(grid, o, sp, 1.0, a["step"], a["steps"], pos[:m], d[:m], kq[:m])
        O.adjoint_backward(grid, o, sp, 1.0, a["step"], a["steps"], pos[:m], d[:m], kq[:m], None,
                           2 * (uo - tgt[:m]))
        return time.perf_counter() - t0

    m = 4
    dt = run(m)
    while dt < target_s / 8 and m < len(E):
        m = min(len(E), 4 * m)
        dt = run(m)
    m2 = int(min(len(E), max(m, m * target_s / max(dt, 1e-3))))
    if m2 > m:
        m = m2
        dt = run(m)
    return {"value": m * a["steps"] / dt, "unit": "particle-steps/s (loss and gradient)", "cores": 1,
            "kind": "port", "sample": "first %d of the %d particles, all %d steps, forward + reverse (%.1f s)"
                                      % (m, len(E), a["steps"], dt)}


def adjoint_measure(args, steps, warmup):
    """Forward (checkpoint every K) + backward (recompute + reverse step, scatter-add) of the
    differentiable Boris model on the C5 workload; HIP-event times of both halves on torch's stream."""
    import torch
    from crpropa3_amd.adjoint import boris_grid_propagate
    a = dict(ADJOINT)
    if args.workload == "adjoint_grid128" and args.propagation_steps:
        a["steps"] = args.propagation_steps
    n = (args.particles if args.workload == "adjoint_grid128" else 0) or a["n"]
    grid_np, pos, d, E, tgt = adjoint_inputs(a, n)
    dev = torch.device("cuda:0")
    grid = torch.tensor(grid_np, device=dev, requires_grad=True)
    x0 = torch.tensor(pos, device=dev)
    u0 = torch.tensor(d, device=dev, requires_grad=True)
    Et = torch.tensor(E, device=dev)
    tg = torch.tensor(tgt, device=dev)
    scale = torch.tensor(1.0, dtype=torch.float64, device=dev, requires_grad=True)
    sp = [a["spacing"]] * 3

    def one():
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ev[0].record()
        _, uo = boris_grid_propagate(grid, x0, u0, Et, eplus, [0, 0, 0], sp, a["step"], a["steps"], a["K"],
                                     scale=scale)
        loss = ((uo - tg) ** 2).sum()
        ev[1].record()
        grid.grad = u0.grad = scale.grad = None
        loss.backward()
        ev[2].record()
        return ev

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    evs = []
    t0 = time.perf_counter()
    for _ in range(steps):
        evs.append(one())
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    fwd = sum(e[0].elapsed_time(e[1]) for e in evs) / steps
    bwd = sum(e[1].elapsed_time(e[2]) for e in evs) / steps
    nsteps = n * a["steps"]
    # the product path's forward on the same particles (PropagationBP, FP64, fixed step, FP32 trilinear
    # grid in the cell layout), for a second denominator
    from crpropa3_amd import capi
    f = capi.Field.grid3f(grid_np, [0, 0, 0], [a["spacing"]] * 3)
    prop = capi.Propagator("BP", f, 0.42, a["step"], a["step"], precision=capi.PRECISION_FP64)
    init = capi.Batch(n)
    init.upload(pos, d, E, np.full(n, eplus))
    work = capi.Batch(n)
    work.copy_from(init)
    prop.propagate(work, max_steps=a["steps"])
    work.copy_from(init)
    bp_ms = prop.propagate(work, max_steps=a["steps"]).kernel_ms
    return {"workload": "adjoint_grid128", "particles": n, "propagation_steps": a["steps"], "step_m": a["step"],
            "checkpoint_every": a["K"], "grid": "%d^3 FP32, spacing %.0f kpc" % (a["grid_n"], a["spacing"] / kpc),
            "forward_ms": fwd, "backward_ms": bwd, "adjoint_over_forward": bwd / fwd,
            "gradient_over_forward": (fwd + bwd) / fwd, "propagate_bp_fp64_ms": bp_ms,
            "adjoint_over_propagate_bp": bwd / bp_ms, "value": nsteps * steps / dt,
            "unit": "particle-steps/s (loss and gradient)", "ms_per_step": dt / steps * 1e3,
            "inputs": (a, grid_np, pos, d, E, tgt)}


def bench_adjoint(args):
    """The adjoint workload alone; value = differentiated particle-steps per second."""
    r = adjoint_measure(args, args.steps, args.warmup)
    a, grid_np, pos, d, E, tgt = r.pop("inputs")
    n = r["particles"]
    bytes_per_launch = n * a["steps"] * ADJ_BYTES_PER_STEP
    achieved = bytes_per_launch / (r["backward_ms"] / 1e3) / 1e9
    config = {k: r[k] for k in ("workload", "particles", "propagation_steps", "step_m", "checkpoint_every", "grid",
                                "forward_ms", "backward_ms", "adjoint_over_forward", "gradient_over_forward",
                                "propagate_bp_fp64_ms", "adjoint_over_propagate_bp")}
    out = {
        "metric": "particle-steps/s (1 MI355X), FP64 vs mixed at stated error; adjoint/fwd time",
        "value": r["value"], "unit": r["unit"],
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: seeded turbulent field, particles and target directions",
        "config": config,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic("adjoint_grid128", "fp64"),
                     "kernel": "backward: k_adj_recompute + k_adj_reverse + k_dep_tiles on three streams",
                     "kernel_ms_avg": r["backward_ms"], "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = adjoint_cpu_baseline(a, grid_np, pos, d, E, tgt)
    print(json.dumps(out), flush=True)


def timed_passes(capi, prop, init, work, max_steps, steps, warmup):
    """`warmup` untimed passes, then `steps` timed passes of the whole batch.  The batch is reset to
    its initial state before each pass, outside the timed region (crp_batch_copy synchronises its
    stream); each pass is bracketed by device synchronisation.  Returns accepted steps, field
    evaluations, kernel milliseconds (HIP events on the kernel's stream), launches and wall seconds."""
    for _ in range(warmup):
        work.copy_from(init)
        prop.propagate(work, max_steps=max_steps)
    capi.check(capi.lib().crp_device_synchronize())
    acc = evals = launches = 0
    kms = dt = 0.0
    for _ in range(steps):
        work.copy_from(init)
        capi.check(capi.lib().crp_device_synchronize())
        t0 = time.perf_counter()
        st = prop.propagate(work, max_steps=max_steps)
        capi.check(capi.lib().crp_device_synchronize())
        dt += time.perf_counter() - t0
        acc += st.accepted_steps
        evals += st.field_evals
        kms += st.kernel_ms
        launches += st.launches
    return acc, evals, kms, launches, dt


def position_error(a, b):
    """Relative final-position error of run a against run b per particle: max-component |dx| / |x|."""
    return np.max(np.abs(a - b), axis=1) / np.max(np.abs(b), axis=1)


def paired_diffusion(pos_a, pos_b, x0, L):
    """Running diffusion coefficients D_perp = <dx^2+dy^2>/(4L), D_par = <dz^2>/(2L) of runs a and b of the
    same particles, their paired difference and its standard error."""
    out = {}
    for name, sel, den in (("perp", (0, 1), 4 * L), ("par", (2,), 2 * L)):
        a = np.sum((pos_a - x0)[:, sel] ** 2, axis=1) / den
        b = np.sum((pos_b - x0)[:, sel] ** 2, axis=1) / den
        diff = a - b
        se = diff.std(ddof=1) / np.sqrt(len(diff))
        out[name] = {"mixed": a.mean(), "fp64": b.mean(), "relative_difference": diff.mean() / b.mean(),
                     "standard_errors": diff.mean() / se if se > 0 else 0.0}
    return out


def mixed_block(capi, w, fdata, field, init, work, pos, final64, value64, args):
    """The mixed path (FP32 field/rotation/direction stages, FP64 positions and step control) on the same
    particles with FP32 cells and with FP16 bricks: rate, speed-up over the FP64 value, and the error of
    its final states against the FP64 pass's."""
    out = {}
    L = w["max_length"] if w["max_length"] > 0 else w.get("steps", 0) * w["max_step"]
    for storage in ("fp32", "fp16"):
        if storage == "fp16":
            field.set_storage("fp16")
        prop = capi.Propagator(w["kind"], field, w["tolerance"], w["min_step"], w["max_step"],
                               max_length=w["max_length"], precision=capi.PRECISION_MIXED)
        acc, evals, kms, launches, dt = timed_passes(capi, prop, init, work, w.get("steps", -1), args.steps,
                                                     min(args.warmup, 2))
        fin = work.download()
        e = position_error(fin["pos"], final64["pos"])
        de = np.max(np.abs(fin["dir"] - final64["dir"]), axis=1)
        kern_s = kms / 1e3 / launches
        bpe = BYTES_PER_EVAL_FP16 if storage == "fp16" else BYTES_PER_EVAL
        achieved = evals * bpe / launches / kern_s / 1e9
        out[storage] = {
            "value": acc / dt, "speedup_vs_fp64": (acc / dt) / value64, "ms_per_step": dt / args.steps * 1e3,
            "kernel_ms_avg": kern_s * 1e3, "accepted_steps_per_pass": acc // args.steps,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "bytes_per_eval": bpe},
            "final_position_error_vs_fp64": {"median": float(np.median(e)), "p90": float(np.percentile(e, 90)),
                                             "p99": float(np.percentile(e, 99)), "max": float(e.max())},
            "final_direction_error_vs_fp64": {"median": float(np.median(de)), "p90": float(np.percentile(de, 90))},
            "diffusion_vs_fp64": paired_diffusion(fin["pos"], final64["pos"], pos, L) if L > 0 else None}
    return out


def read_traffic(workload, precision):
    path = os.path.join(ROOT, "profiles", "traffic_%s%s.json" % (workload, "" if precision == "fp64" else "_mixed"))
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


def read_flops(workload):
    path = os.path.join(ROOT, "profiles", "flops_%s.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("fp64_flops_per_launch")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ck_grid256", choices=sorted(WORKLOADS) + ["adjoint_grid128"])
    ap.add_argument("--particles", type=int, default=0, help="override the particle count")
    ap.add_argument("--propagation-steps", type=int, default=0, help="override the steps of fixed-step workloads")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="fp64: the reference's arithmetic bit for bit; mixed: FP32 field and rotation")
    ap.add_argument("--grid-layout", default="cells", choices=["cells", "points"],
                    help="device storage of the grid (results are identical)")
    ap.add_argument("--grid-storage", default="fp32", choices=["fp32", "fp16"],
                    help="grid values read by the mixed path: FP32, or FP16 with a per-8^3-brick FP32 scale")
    ap.add_argument("--interpolation", default="trilinear", choices=["trilinear", "tricubic", "nearest"],
                    help="Grid::setInterpolationType of a grid workload (tricubic and nearest read the points)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision block of the default workload")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint block of the default workload")
    ap.add_argument("--autotune", type=int, default=0, metavar="PROBE_STEPS",
                    help="run crp_propagator_autotune with this many probe steps before timing, and report its choice")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it (got --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload == "adjoint_grid128":
        return bench_adjoint(args)
    from crpropa3_amd import capi
    w = dict(WORKLOADS[args.workload])
    if args.propagation_steps and "steps" in w:
        w["steps"] = args.propagation_steps
    n = args.particles or w["n"]
    fdata = make_field_data(w)
    pos, d, E, q = make_particles(w, n)
    capi.check(capi.lib().crp_set_device(0))
    if w["field"] == "jf12":
        field = capi.Field.jf12(striated=fdata["striated"], striated_spacing=0.1 * kpc, turbulent=fdata["turbulent"],
                                turbulent_spacing=4 * pc)
    elif w["field"] == "grid":
        field = capi.Field.grid3f(fdata, [0, 0, 0], [w["spacing"]] * 3)
        if args.interpolation != "trilinear":
            if args.grid_storage == "fp16":
                sys.exit("bench.py: --interpolation %s reads the grid's FP32 points (no FP16 storage)"
                         % args.interpolation)
            args.grid_layout = "points"  # the cell table holds trilinear corners
            field.set_interpolation({"tricubic": "TRICUBIC", "nearest": "NEAREST_NEIGHBOUR"}[args.interpolation])
        else:
            field.set_layout(args.grid_layout)
        if args.grid_storage == "fp16":
            if args.precision != "mixed":
                sys.exit("bench.py: --grid-storage fp16 is read by the mixed path only (use --precision mixed)")
            field.set_storage("fp16")
    else:
        field = capi.Field.planewave(fdata["k"], fdata["Ak"], fdata["xi"], fdata["kappa"], fdata["beta"])
    prec = capi.PRECISION_MIXED if args.precision == "mixed" else capi.PRECISION_FP64
    if w["kind"] == "SDE":
        if args.precision != "fp64":
            sys.exit("bench.py: DiffusionSDE runs in FP64 only")
        prop = capi.Propagator("SDE", field, w["tolerance"], w["min_step"], w["max_step"], max_length=w["max_length"],
                               epsilon=w["epsilon"], alpha=w["alpha"], scale=w["scale"], seed=w["sde_seed"])
    else:
        prop = capi.Propagator(w["kind"], field, w["tolerance"], w["min_step"], w["max_step"],
                               max_length=w["max_length"], precision=prec)
    init = capi.Batch(n)
    init.upload(pos, d, E, q)
    work = capi.Batch(n)
    max_steps = w.get("steps", -1)
    tuned = None
    if args.autotune > 0:
        t0 = time.perf_counter()
        tuned = prop.autotune(init, args.autotune)
        tuned["autotune_s"] = time.perf_counter() - t0
        tuned["probe_steps"] = args.autotune

    acc, evals, kms, launches, dt = timed_passes(capi, prop, init, work, max_steps, args.steps, args.warmup)
    final = work.download()
    kern_s = kms / 1e3 / launches
    per_trial = 6 if w["kind"] == "CK" else (1 if w["min_step"] == w["max_step"] else 3)
    config = {"workload": args.workload, "propagator": w["kind"], "precision": args.precision,
              "particles": n, "tolerance": w["tolerance"], "min_step_m": w["min_step"], "max_step_m": w["max_step"],
              "max_length_Mpc": w["max_length"] / Mpc, "propagation_steps": w.get("steps"),
              "accepted_steps_per_pass": acc // args.steps}
    if w["kind"] == "SDE":
        config.update({"propagator": "DiffusionSDE", "epsilon": w["epsilon"], "alpha": w["alpha"],
                       "scale": w["scale"], "field_evals_per_pass": evals // args.steps})
    else:
        config["trials_per_pass"] = (evals // per_trial) // args.steps
    config["launch"] = prop.launch()
    if tuned is not None:
        config["autotuned"] = tuned
    if w["field"] == "grid":
        fp16 = args.grid_storage == "fp16"
        config.update({"grid": "%d^3 %s, spacing %.0f kpc" % (w["grid_n"], "FP16+brick scale" if fp16 else "FP32",
                                                              w["spacing"] / kpc),
                       "grid_layout": "cells16" if fp16 else args.grid_layout, "interpolation": args.interpolation})
        # bytes one evaluation reads: 8 corners (trilinear), 4^3 points (tricubic) or 1 point (nearest), 3 x FP32
        per_eval = {"trilinear": BYTES_PER_EVAL_FP16 if fp16 else BYTES_PER_EVAL, "tricubic": 64 * 12,
                    "nearest": 12}[args.interpolation]
        bytes_per_launch = evals * per_eval / launches
        achieved = bytes_per_launch / kern_s / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS,
                    "traffic": read_traffic(args.workload, args.precision) if args.interpolation == "trilinear" else None,
                    "kernel": ("k_sde<%s>" if w["kind"] == "SDE" else "k_propagate<%s, " + w["kind"] + ">") % (
                        "FieldGridAny" if args.interpolation != "trilinear" else "FieldGridCells16" if fp16 else
                        ("FieldGridCells" if args.grid_layout == "cells" else "FieldGrid")),
                    "kernel_ms_avg": kern_s * 1e3, "algorithmic_bytes_per_launch": bytes_per_launch}
        if args.grid_layout == "cells":
            # every evaluation reads one random line of the cell table; the GPU's rate of such reads,
            # measured on a table of the same size, is the ceiling this kernel can reach
            line = 64 if fp16 else 96
            table = w["grid_n"] ** 3 * (64 if fp16 else 128)
            ceiling = capi.measure_gather(table, line)
            evals_s = evals / launches / kern_s
            roofline["random_line_ceiling"] = {"lines_per_s": ceiling, "bytes_per_line": line, "table_bytes": table,
                                               "evals_per_s": evals_s, "frac": evals_s / ceiling}
    elif w["field"] == "jf12":
        config.update({"field": "JF12Field + randomStriated(%d) + randomTurbulent(%d), 256^3 at 4 pc"
                                % (w["striated_seed"], w["turb_seed"])})
        # FP64 VALU work per launch from the instruction counters of the committed profile (64 lanes per
        # wave instruction, FMA = 2 FLOP): the field's correctly rounded transcendental functions dominate
        fl = read_flops(args.workload)
        achieved = fl / kern_s / 1e12 if fl else None
        roofline = {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / FP64_PEAK_TFLOPS if achieved else None,
                    "traffic": read_traffic(args.workload, args.precision), "kernel": "k_propagate<FieldJF12, CK>",
                    "kernel_ms_avg": kern_s * 1e3, "fp64_flops_per_launch": fl,
                    "field_evals_per_s": evals / launches / kern_s}
    else:
        config.update({"field": "PlaneWaveTurbulence Nm=%d" % w["nm"]})
        flops = evals * (w["nm"] * FLOP_PER_MODE + FLOP_BORIS) / launches
        achieved = flops / kern_s / 1e12
        peak = FP64_PEAK_TFLOPS if args.precision == "fp64" else FP32_PEAK_TFLOPS
        roofline = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                    "traffic": read_traffic(args.workload, args.precision),
                    "kernel": "k_propagate<FieldPlaneWave, %s>" % w["kind"], "kernel_ms_avg": kern_s * 1e3,
                    "algorithmic_flops_per_launch": flops, "cos_per_launch": evals * w["nm"] / launches,
                    "cos_per_s": evals * w["nm"] / launches / kern_s}
    out = {
        "metric": "particle-steps/s (1 MI355X), FP64 vs mixed at stated error; adjoint/fwd time",
        "value": acc / dt,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 field/rotation + f64 position/step control",
        "data": "synthetic: seeded turbulent field and seeded particles",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    headline = (args.workload == "ck_grid256" and args.precision == "fp64" and args.grid_layout == "cells"
                and args.interpolation == "trilinear")
    if headline and not args.no_mixed:
        # the metric's second half: the mixed path on the same particles, error against this FP64 pass
        out["mixed"] = mixed_block(capi, w, fdata, field, init, work, pos, final, acc / dt, args)
    del init, work, prop, field
    if headline and not args.no_adjoint:
        # the metric's third half: adjoint / forward time on the C5 workload
        a = adjoint_measure(args, max(1, min(args.steps, 5)), max(1, min(args.warmup, 2)))
        out["adjoint"] = {k: a[k] for k in ("workload", "particles", "propagation_steps", "checkpoint_every",
                                            "forward_ms", "backward_ms", "adjoint_over_forward",
                                            "gradient_over_forward", "propagate_bp_fp64_ms",
                                            "adjoint_over_propagate_bp", "value", "unit")}
    if not args.no_cpu_baseline and args.interpolation == "trilinear":
        out["cpu_baseline"] = cpu_baseline(w, fdata, pos, d, E, q, gpu_final=final)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
