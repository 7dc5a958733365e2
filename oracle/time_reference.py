#!/usr/bin/env python3
"""Times the reference build (oracle/_ref/ref_time) and the oracle port (oracle/liboracle.so, glibc libm)
on the same samples of the C1 and C2 benchmark workloads, with 1 and 8 OpenMP threads, and prints a
markdown summary.  Runs only where the reference is present and `make -C oracle ref` has built
ref_time; the reference does not travel to the GPU machine.

The grids are the oracle's restatement of SimpleGridTurbulence::initTurbulence (the reference's own
generator needs FFTW, absent here): the same MT19937 k-space draws, a double-precision inverse FFT,
scaled to Brms and written in the reference's raw float32 format (src/GridTools.cpp:200-218), which
ref_time reads with the reference's loadGrid.  The particles are the benchmark's draws (one
Random(seed); randBin, randPowerLaw(-1), randVector per particle).

    python3 oracle/time_reference.py [--threads 1,8]
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle_ctypes as O  # noqa: E402

pc = 648000 / math.pi * 149597870700.0
kpc, Mpc = 1e3 * pc, 1e6 * pc
EeV = 1e18 * 1.602176487e-19
nG = 1e-13

# name: (kind, n, spacing, lmin cells, lmax cells, seed, emin, emax, tolerance, minStep, maxStep, maxLength,
#        steps, species weights, particles timed with 1 thread / with more)
WORKLOADS = {
    "C1 bp_grid64": ("BP", 64, 50 * kpc, 2, 32, 42, 1 * EeV, 100 * EeV, 0.42, 0.1 * Mpc, 0.1 * Mpc, 0.0, 2000,
                     ((1, 1.0),), 4096, 4096),
    "C2 ck_grid256": ("CK", 256, 20 * kpc, 2, 128, 7, 1 * EeV, 300 * EeV, 1e-4, 1 * kpc, 1 * Mpc, 100 * Mpc, -1,
                      ((1, 0.5), (2, 0.2), (7, 0.15), (14, 0.1), (26, 0.05)), 2048, 16384),
}


def make_grid(n, spacing, lmin, lmax, seed):
    g, _ = O.turbulence_grid(n, spacing, lmin, lmax, seed=seed)
    rms = math.sqrt(np.mean(np.sum(g * g, axis=-1)))
    return (g * (1 * nG / rms)).astype(np.float32)


def draws(seed, m, emin, emax, weights):
    sp, E, d = O.sample_particles(seed, m, emin, emax, weights=weights)
    return sp, E, d


def time_port(grid, spacing, w, m, threads):
    kind, n = w[0], w[1]
    species = np.array([z for z, _ in w[13]], dtype=np.float64)
    sp, E, d = draws(w[5], m, w[6], w[7], [x for _, x in w[13]])
    q = species[sp] * 1.602176487e-19
    L = n * spacing
    st = O.OracleState(np.tile([L / 2] * 3, (m, 1)), d, E, q)
    f = O.OracleField.grid(grid, [0, 0, 0], [spacing] * 3)
    t0 = time.perf_counter()
    nacc, _ = O.propagate(f, O.PROP_BP if kind == "BP" else O.PROP_CK, st, w[12] if w[12] >= 0 else 10 ** 12,
                          tolerance=w[8], min_step=w[9], max_step=w[10], max_len=w[11], libm=O.LIBM_GLIBC,
                          nthreads=threads)
    dt = time.perf_counter() - t0
    return nacc, dt, float(np.sum(st.pos) / Mpc)


def time_reference(path, w, m, threads):
    kind, n, spacing = w[0], w[1], w[2]
    spec = ",".join("%d:%r" % (z, x) for z, x in w[13])
    cmd = [os.path.join(HERE, "_ref", "ref_time"), kind, path, str(n), repr(spacing), str(m), str(w[5]), repr(w[6]),
           repr(w[7]), repr(w[8]), repr(w[9]), repr(w[10]), repr(w[11]), str(w[12]), spec]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), KISS_LOG_LEVEL="0")
    r = json.loads(subprocess.check_output(cmd, env=env).decode().strip().splitlines()[-1])
    return r["accepted_steps"], r["seconds"], r["position_sum_Mpc"], r["threads"]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--threads", default="1,8")
    args = ap.parse_args()
    if not os.path.exists(os.path.join(HERE, "_ref", "ref_time")):
        sys.exit("time_reference.py: oracle/_ref/ref_time is missing (make -C oracle ref)")
    threads = [int(t) for t in args.threads.split(",")]
    rows = []
    with tempfile.TemporaryDirectory() as tmp:
        for name, w in WORKLOADS.items():
            grid = make_grid(w[1], w[2], w[3] * w[2], w[4] * w[2], w[5])
            path = os.path.join(tmp, "grid.raw")
            grid.tofile(path)  # raw float32, (ix, iy, iz, component), as dumpGrid writes
            for t in threads:
                m = w[14] if t == 1 else w[15]
                ra, rt, rs, rth = time_reference(path, w, m, t)
                pa, pt, ps = time_port(grid, w[2], w, m, t)
                rows.append((name, m, t, rth, ra, rt, pa, pt, rs, ps))
                print("%s: %d particles, %d threads: reference %.3e steps/s, port %.3e steps/s"
                      % (name, m, t, ra / rt, pa / pt), file=sys.stderr, flush=True)
    print("| Workload | particles | OpenMP threads | reference steps | reference s | reference steps/s "
          "| port steps | port s | port steps/s | port / reference |")
    print("|---|---|---|---|---|---|---|---|---|---|")
    for name, m, t, rth, ra, rt, pa, pt, rs, ps in rows:
        print("| %s | %d | %d | %d | %.2f | %.3e | %d | %.2f | %.3e | %.2f |"
              % (name, m, rth, ra, rt, ra / rt, pa, pt, pa / pt, (pa / pt) / (ra / rt)))
    print()
    print("Same accepted steps and final positions (sum of coordinates, Mpc) in both: %s"
          % all(r[4] == r[6] and abs(r[8] - r[9]) <= 1e-9 * max(1.0, abs(r[8])) for r in rows))


if __name__ == "__main__":
    main()
