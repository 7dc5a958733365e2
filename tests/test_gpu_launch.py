"""Launch configuration, hipGraph replay and the autotuner (crp_launch_config).

Every particle's arithmetic is independent of which lane, launch or graph
replay runs it, so results must be bit-identical across block sizes,
workgroups per CU, steps per launch and graph replay, and equal to the
oracle (with the kernels' correctly rounded libm)."""
import numpy as np
import pytest

import oracle_ctypes as O
from crpropa3_amd import capi

pytestmark = pytest.mark.gpu

KEYS = ("pos", "dir", "next_step", "current_step", "trajectory_length", "time", "prev_pos", "prev_dir")

CONFIGS = [dict(block=64, blocks_per_cu=1), dict(block=128), dict(block=256, blocks_per_cu=1),
           dict(block=256, steps_per_launch=7), dict(block=256, use_graph=True),
           dict(block=64, blocks_per_cu=2, steps_per_launch=5, use_graph=True)]


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a, dtype=np.float64).view(np.uint64),
                          np.ascontiguousarray(b, dtype=np.float64).view(np.uint64))


def c2_small(n=3000, seed=5):
    g = O.load_case("grid32")
    gm = g["meta"]["grid"]
    sp, E, d = O.sample_particles(seed, n, 1 * O.EeV, 300 * O.EeV, weights=[0.5, 0.2, 0.15, 0.1, 0.05])
    q = np.array([1, 2, 7, 14, 26])[sp] * O.eplus
    L = gm["n"] * gm["spacing"]
    return g["grid"], gm, np.tile([L / 2] * 3, (n, 1)), d, E, q  # d: raw draws, normalised on upload


def run(prop, pos, d, E, q, chunks):
    b = capi.Batch(len(E))
    b.upload(pos, O.normalise(d), E, q)  # the oracle state normalises d itself (setDirection)
    out, stats = [], []
    for m in chunks:
        st = prop.propagate(b, max_steps=m)
        stats.append((st.accepted_steps, st.rejected_trials, st.field_evals, st.finished, st.launches))
        out.append(b.download())
    return out, stats


@pytest.mark.parametrize("kind", ["CK", "BP"])
def test_results_independent_of_launch_configuration(kind):
    grid, gm, pos, d, E, q = c2_small()
    field = capi.Field.grid3f(grid, gm["origin"], gm["spacing"])
    tol = 1e-4 if kind == "CK" else 1e-3
    chunks = [1, 30, 100, -1]
    base = capi.Propagator(kind, field, tol, 1 * O.kpc, 1 * O.Mpc, max_length=5 * O.Mpc)
    ref, ref_st = run(base, pos, d, E, q, chunks)
    # the default configuration against the oracle
    of = O.OracleField.grid(grid, gm["origin"], gm["spacing"])
    ost = O.OracleState(pos, d, E, q)
    pt = O.PROP_CK if kind == "CK" else O.PROP_BP
    for si, m in enumerate(chunks):
        O.propagate(of, pt, ost, m if m > 0 else 10 ** 9, tolerance=tol, min_step=1 * O.kpc, max_step=1 * O.Mpc,
                    max_len=5 * O.Mpc, libm=O.LIBM_CR)
        assert bits_equal(ref[si]["pos"], ost.pos) and bits_equal(ref[si]["dir"], ost.dir)
    for cfg in CONFIGS:
        p = capi.Propagator(kind, field, tol, 1 * O.kpc, 1 * O.Mpc, max_length=5 * O.Mpc).set_launch(**cfg)
        got, st = run(p, pos, d, E, q, chunks)
        for si in range(len(chunks)):
            for k in KEYS:
                assert bits_equal(got[si][k], ref[si][k]), (cfg, si, k)
            assert np.array_equal(got[si]["active"], ref[si]["active"])
            assert st[si][:4] == ref_st[si][:4], (cfg, si)
        spl = cfg.get("steps_per_launch", 0)
        if spl:
            assert st[2][4] == -(-100 // spl)  # a budget of 100 steps in launches of spl steps
            assert st[3][4] > 1  # open-ended run: several launches
        lc = p.launch()
        assert lc["block"] == cfg["block"] and lc["particles_per_lane"] > 0


def test_graph_replay_every_replay_matches_eager():
    grid, gm, pos, d, E, q = c2_small(2000, seed=9)
    field = capi.Field.grid3f(grid, gm["origin"], gm["spacing"])
    eager = capi.Propagator("CK", field, 1e-4, 1 * O.kpc, 1 * O.Mpc, max_length=4 * O.Mpc)
    graph = capi.Propagator("CK", field, 1e-4, 1 * O.kpc, 1 * O.Mpc, max_length=4 * O.Mpc).set_launch(use_graph=True)
    init = capi.Batch(len(E))
    init.upload(pos, O.normalise(d), E, q)
    w_e, w_g = capi.Batch(len(E)), capi.Batch(len(E))
    for rep in range(4):
        for budget in (50, 50, 120):  # same budget replays the captured graph; a new one re-captures
            w_e.copy_from(init)
            w_g.copy_from(init)
            se = eager.propagate(w_e, max_steps=budget)
            sg = graph.propagate(w_g, max_steps=budget)
            a, b = w_e.download(), w_g.download()
            for k in KEYS:
                assert bits_equal(a[k], b[k]), (rep, budget, k)
            assert (se.accepted_steps, se.rejected_trials, se.finished) == (sg.accepted_steps, sg.rejected_trials,
                                                                            sg.finished)


def test_autotune_leaves_batch_unchanged_and_results_identical():
    grid, gm, pos, d, E, q = c2_small(20000, seed=3)
    field = capi.Field.grid3f(grid, gm["origin"], gm["spacing"])
    p = capi.Propagator("CK", field, 1e-4, 1 * O.kpc, 1 * O.Mpc, max_length=3 * O.Mpc)
    b = capi.Batch(len(E))
    b.upload(pos, O.normalise(d), E, q)
    before = b.download()
    cfg = p.autotune(b, probe_steps=16)
    print("autotune:", cfg)
    assert cfg["block"] in (64, 128, 256)
    assert cfg["blocks_per_cu"] >= 1 and cfg["steps_per_launch"] >= 0 and cfg["particles_per_lane"] > 0
    after = b.download()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    ref = capi.Propagator("CK", field, 1e-4, 1 * O.kpc, 1 * O.Mpc, max_length=3 * O.Mpc)
    b2 = capi.Batch(len(E))
    b2.upload(pos, O.normalise(d), E, q)
    ref.propagate(b2)
    p.propagate(b)
    x, y = b.download(), b2.download()
    for k in KEYS:
        assert bits_equal(x[k], y[k]), k


def test_launch_config_validation():
    p = capi.Propagator("BP", capi.Field.uniform([0, 0, 1e-13]), 0.42, 1e19, 1e19)
    with pytest.raises(capi.CrpError):
        p.set_launch(block=96)
    with pytest.raises(capi.CrpError):
        p.set_launch(blocks_per_cu=-1)
    with pytest.raises(capi.CrpError):
        p.autotune(capi.Batch(4), probe_steps=0)
