#!/usr/bin/env python3
"""Benchmark of the batched propagator on one MI355X.

Workload (default, "ck_grid256"): adaptive Cash-Karp (PropagationCK, tolerance
1e-4, minStep 1 kpc, maxStep 1 Mpc) of 2^20 particles in a periodic 256^3 FP32
Kolmogorov Grid3f (spacing 20 kpc, Brms 1 nG, lMin 2 cells, lMax 128 cells,
seed 7), species Z in {1,2,7,14,26} with weights {.5,.2,.15,.1,.05}, E
log-uniform 1-300 EeV, isotropic directions from the box centre, each particle
stopped by MaximumTrajectoryLength at 100 Mpc.

One step = propagate the whole batch (reset to its initial state) until every
particle is inactive.  The metric is accepted particle-steps (process() calls)
per second, inputs resident in HBM.  Prints one JSON line.
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
EeV = 1e18 * eplus
nG = 1e-13

HBM_PEAK_GBS = 8000.0
BYTES_PER_EVAL = 96  # 8 corners x 3 x FP32 of one trilinear Grid3f evaluation

WORKLOADS = {
    "ck_grid256": dict(kind="CK", n=2 ** 20, grid_n=256, spacing=20 * kpc, lmin_cells=2, lmax_cells=128, seed=7,
                       tolerance=1e-4, min_step=1 * kpc, max_step=1 * Mpc, max_length=100 * Mpc,
                       emin=1 * EeV, emax=300 * EeV, species=(1, 2, 7, 14, 26),
                       weights=(0.5, 0.2, 0.15, 0.1, 0.05)),
    "bp_grid64": dict(kind="BP", n=4096, grid_n=64, spacing=50 * kpc, lmin_cells=2, lmax_cells=32, seed=42,
                      tolerance=0.42, min_step=0.1 * Mpc, max_step=0.1 * Mpc, max_length=0.0, steps=2000,
                      emin=1 * EeV, emax=100 * EeV, species=(1,), weights=(1.0,)),
}


def make_workload(w, n=None):
    from crpropa3_amd.turbulence import simple_grid_turbulence
    n = n or w["n"]
    L = w["grid_n"] * w["spacing"]
    grid = simple_grid_turbulence(w["grid_n"], w["spacing"], 1 * nG, w["lmin_cells"] * w["spacing"],
                                  w["lmax_cells"] * w["spacing"], w["seed"])
    rng = np.random.Generator(np.random.PCG64(w["seed"] + 1000))
    sp = rng.choice(len(w["species"]), size=n, p=np.asarray(w["weights"]) / np.sum(w["weights"]))
    Z = np.asarray(w["species"], dtype=np.float64)[sp]
    E = np.exp(rng.uniform(np.log(w["emin"]), np.log(w["emax"]), n))
    z = rng.uniform(-1, 1, n)
    t = rng.uniform(-np.pi, np.pi, n)
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(t), r * np.sin(t), z], axis=1)
    d /= np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])[:, None]
    pos = np.tile([L / 2, L / 2, L / 2], (n, 1))
    return grid, pos, d, E, Z * eplus


def cpu_baseline(w, grid, pos, d, E, q, target_s=15.0):
    """The oracle (CPU restatement with glibc libm, OpenMP) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    f = O.OracleField.grid(grid, [0, 0, 0], [w["spacing"]] * 3)
    kind = O.PROP_CK if w["kind"] == "CK" else O.PROP_BP
    steps = w.get("steps", 10 ** 12)

    def run(m):
        st = O.OracleState(pos[:m], d[:m], E[:m], q[:m])
        t0 = time.perf_counter()
        nacc, _ = O.propagate(f, kind, st, steps, tolerance=w["tolerance"], min_step=w["min_step"],
                              max_step=w["max_step"], max_len=w["max_length"], libm=O.LIBM_GLIBC, nthreads=threads)
        return nacc, time.perf_counter() - t0

    m = min(len(E), 8 * threads)
    nacc, dt = run(m)
    m2 = int(min(len(E), max(m, m * target_s / max(dt, 1e-3))))
    if m2 > m:
        nacc, dt = run(m2)
        m = m2
    return {"value": nacc / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "first %d of the %d particles, same grid and propagator, run to completion "
                      "(%d accepted steps in %.1f s)" % (m, len(E), nacc, dt)}


def read_traffic(workload):
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
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
    ap.add_argument("--particles", type=int, default=0, help="override the particle count (tests only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--grid-layout", default="cells", choices=["cells", "points"],
                    help="device storage of the grid (results are identical)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not part of it (got --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    from crpropa3_amd import capi
    w = WORKLOADS[args.workload]
    grid, pos, d, E, q = make_workload(w, args.particles or None)
    n = len(E)
    capi.check(capi.lib().crp_set_device(0))
    field = capi.Field.grid3f(grid, [0, 0, 0], [w["spacing"]] * 3).set_layout(args.grid_layout)
    prop = capi.Propagator(w["kind"], field, w["tolerance"], w["min_step"], w["max_step"],
                           max_length=w["max_length"])
    init = capi.Batch(n)
    init.upload(pos, d, E, q)
    work = capi.Batch(n)
    max_steps = w.get("steps", -1)

    for _ in range(args.warmup):
        work.copy_from(init)
        prop.propagate(work, max_steps=max_steps)
    capi.check(capi.lib().crp_device_synchronize())
    acc = 0
    evals = 0
    kms = 0.0
    launches = 0
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

    value = acc / dt
    kern_s = kms / 1e3 / launches
    bytes_per_launch = evals * BYTES_PER_EVAL / launches
    achieved = bytes_per_launch / kern_s / 1e9
    out = {
        "metric": "particle-steps/s (1 MI355X), FP64 vs mixed at stated error; adjoint/fwd time",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Kolmogorov Grid3f and seeded particles",
        "config": {"workload": args.workload, "propagator": w["kind"], "precision": "fp64 (bit-exact path)", "grid_layout": args.grid_layout,
                   "particles": n, "grid": "%d^3 FP32" % w["grid_n"], "tolerance": w["tolerance"],
                   "min_step_kpc": w["min_step"] / kpc, "max_step_kpc": w["max_step"] / kpc,
                   "max_length_Mpc": w["max_length"] / Mpc,
                   "accepted_steps_per_pass": acc // args.steps,
                   "trials_per_pass": (evals // (6 if w["kind"] == "CK" else 1)) // args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(args.workload),
                     "kernel": "k_propagate (%s, grid)" % w["kind"], "kernel_ms_avg": kern_s * 1e3,
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, grid, pos, d, E, q)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
