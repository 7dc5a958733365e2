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
lMin 1 pc, lMax 100 pc, seed 11), bp_grid512 / ck_grid512 (2^21 protons,
1000 fixed 20 kpc steps in a 512^3 grid of 10 kpc, seed 3).

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
                         steps=1000, emin=1e15 * eV, emax=1e18 * eV, species=(1,), weights=(1.0,), start="origin"),
    "bp_grid512": dict(kind="BP", n=2 ** 21, field="grid", grid_n=512, spacing=10 * kpc, lmin_cells=2,
                       lmax_cells=256, seed=3, tolerance=0.42, min_step=20 * kpc, max_step=20 * kpc,
                       max_length=0.0, steps=1000, emin=1 * EeV, emax=100 * EeV, species=(1,), weights=(1.0,),
                       start="centre"),
    "ck_grid512": dict(kind="CK", n=2 ** 21, field="grid", grid_n=512, spacing=10 * kpc, lmin_cells=2,
                       lmax_cells=256, seed=3, tolerance=1e-4, min_step=20 * kpc, max_step=20 * kpc,
                       max_length=0.0, steps=1000, emin=1 * EeV, emax=100 * EeV, species=(1,), weights=(1.0,),
                       start="centre"),
}


def make_particles(w, n):
    rng = np.random.Generator(np.random.PCG64(w["seed"] + 1000))
    sp = rng.choice(len(w["species"]), size=n, p=np.asarray(w["weights"]) / np.sum(w["weights"]))
    Z = np.asarray(w["species"], dtype=np.float64)[sp]
    E = np.exp(rng.uniform(np.log(w["emin"]), np.log(w["emax"]), n))
    z = rng.uniform(-1, 1, n)
    t = rng.uniform(-np.pi, np.pi, n)
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(t), r * np.sin(t), z], axis=1)
    d /= np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])[:, None]
    if w["start"] == "centre":
        L = w["grid_n"] * w["spacing"]
        pos = np.tile([L / 2, L / 2, L / 2], (n, 1))
    else:
        pos = np.zeros((n, 3))
    return pos, d, E, Z * eplus


def make_field_data(w):
    """Grid values (numpy) or plane-wave mode table (dict) of the workload."""
    if w["field"] == "grid":
        from crpropa3_amd.turbulence import simple_grid_turbulence
        return simple_grid_turbulence(w["grid_n"], w["spacing"], 1 * nG, w["lmin_cells"] * w["spacing"],
                                      w["lmax_cells"] * w["spacing"], w["seed"])
    from crpropa3_amd import PlaneWaveTurbulence, SimpleTurbulenceSpectrum
    pw = PlaneWaveTurbulence(SimpleTurbulenceSpectrum(w["brms"], w["lmin"], w["lmax"]), w["nm"], w["seed"])
    return {"k": np.array(pw.getK()), "Ak": np.array(pw.getAk()), "beta": np.array(pw.getBeta()),
            "xi": np.array([[v.x, v.y, v.z] for v in pw.getXi()]),
            "kappa": np.array([[v.x, v.y, v.z] for v in pw.getKappa()])}


def cpu_baseline(w, fdata, pos, d, E, q, target_s=15.0):
    """The oracle (CPU restatement with glibc libm, OpenMP) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    if w["field"] == "grid":
        f = O.OracleField.grid(fdata, [0, 0, 0], [w["spacing"]] * 3)
    else:
        f = O.OracleField.planewave(fdata)
    kind = O.PROP_CK if w["kind"] == "CK" else O.PROP_BP
    steps = w.get("steps", 10 ** 12)

    def run(m, nsteps):
        st = O.OracleState(pos[:m], d[:m], E[:m], q[:m])
        t0 = time.perf_counter()
        nacc, _ = O.propagate(f, kind, st, nsteps, tolerance=w["tolerance"], min_step=w["min_step"],
                              max_step=w["max_step"], max_len=w["max_length"], libm=O.LIBM_GLIBC, nthreads=threads)
        return nacc, time.perf_counter() - t0

    m = min(len(E), 8 * threads)
    nsteps = steps
    nacc, dt = run(m, nsteps)
    if dt > target_s and steps < 10 ** 12:  # too slow even for the calibration: fewer steps
        nsteps = max(1, int(steps * target_s / dt / 4))
        nacc, dt = run(m, nsteps)
    m2 = int(min(len(E), max(m, m * target_s / max(dt, 1e-3))))
    if m2 > m:
        nacc, dt = run(m2, nsteps)
        m = m2
    what = "run to completion" if nsteps >= steps else "%d of %d steps" % (nsteps, steps)
    return {"value": nacc / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "first %d of the %d particles, same field and propagator, %s (%d accepted steps in %.1f s)"
                      % (m, len(E), what, nacc, dt)}


def read_traffic(workload, precision):
    path = os.path.join(ROOT, "profiles", "traffic_%s%s.json" % (workload, "" if precision == "fp64" else "_mixed"))
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ck_grid256", choices=sorted(WORKLOADS))
    ap.add_argument("--particles", type=int, default=0, help="override the particle count")
    ap.add_argument("--propagation-steps", type=int, default=0, help="override the steps of fixed-step workloads")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="fp64: the reference's arithmetic bit for bit; mixed: FP32 field and rotation")
    ap.add_argument("--grid-layout", default="cells", choices=["cells", "points"],
                    help="device storage of the grid (results are identical)")
    ap.add_argument("--grid-storage", default="fp32", choices=["fp32", "fp16"],
                    help="grid values read by the mixed path: FP32, or FP16 with a per-8^3-brick FP32 scale")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it (got --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from crpropa3_amd import capi
    w = dict(WORKLOADS[args.workload])
    if args.propagation_steps and "steps" in w:
        w["steps"] = args.propagation_steps
    n = args.particles or w["n"]
    fdata = make_field_data(w)
    pos, d, E, q = make_particles(w, n)
    capi.check(capi.lib().crp_set_device(0))
    if w["field"] == "grid":
        field = capi.Field.grid3f(fdata, [0, 0, 0], [w["spacing"]] * 3).set_layout(args.grid_layout)
        if args.grid_storage == "fp16":
            if args.precision != "mixed":
                sys.exit("bench.py: --grid-storage fp16 is read by the mixed path only (use --precision mixed)")
            field.set_storage("fp16")
    else:
        field = capi.Field.planewave(fdata["k"], fdata["Ak"], fdata["xi"], fdata["kappa"], fdata["beta"])
    prec = capi.PRECISION_MIXED if args.precision == "mixed" else capi.PRECISION_FP64
    prop = capi.Propagator(w["kind"], field, w["tolerance"], w["min_step"], w["max_step"],
                           max_length=w["max_length"], precision=prec)
    init = capi.Batch(n)
    init.upload(pos, d, E, q)
    work = capi.Batch(n)
    max_steps = w.get("steps", -1)

    for _ in range(args.warmup):
        work.copy_from(init)
        prop.propagate(work, max_steps=max_steps)
    capi.check(capi.lib().crp_device_synchronize())
    acc = evals = launches = 0
    kms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        work.copy_from(init)
        st = prop.propagate(work, max_steps=max_steps)
        acc += st.accepted_steps
        evals += st.field_evals
        kms += st.kernel_ms
        launches += st.launches
    capi.check(capi.lib().crp_device_synchronize())
    dt = time.perf_counter() - t0

    kern_s = kms / 1e3 / launches
    per_trial = 6 if w["kind"] == "CK" else (1 if w["min_step"] == w["max_step"] else 3)
    config = {"workload": args.workload, "propagator": w["kind"], "precision": args.precision,
              "particles": n, "tolerance": w["tolerance"], "min_step_m": w["min_step"], "max_step_m": w["max_step"],
              "max_length_Mpc": w["max_length"] / Mpc, "propagation_steps": w.get("steps"),
              "accepted_steps_per_pass": acc // args.steps,
              "trials_per_pass": (evals // per_trial) // args.steps}
    if w["field"] == "grid":
        fp16 = args.grid_storage == "fp16"
        config.update({"grid": "%d^3 %s, spacing %.0f kpc" % (w["grid_n"], "FP16+brick scale" if fp16 else "FP32",
                                                              w["spacing"] / kpc),
                       "grid_layout": "cells16" if fp16 else args.grid_layout})
        bytes_per_launch = evals * (BYTES_PER_EVAL_FP16 if fp16 else BYTES_PER_EVAL) / launches
        achieved = bytes_per_launch / kern_s / 1e9
        roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(args.workload, args.precision),
                    "kernel": "k_propagate<%s, %s>" % ("FieldGridCells16" if fp16 else (
                        "FieldGridCells" if args.grid_layout == "cells" else "FieldGrid"), w["kind"]),
                    "kernel_ms_avg": kern_s * 1e3, "algorithmic_bytes_per_launch": bytes_per_launch}
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
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, fdata, pos, d, E, q)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
