"""DiffusionSDE in the oracle (src/module/DiffusionSDE.cpp:48-254), on the reference's own random
stream (one MT19937 seeded once, consumed particle after particle, glibc libm): final positions,
directions, step sizes, lengths and times equal the reference's bit for bit in a uniform field, a
fine turbulent grid (the step search halves) and with advection (neutral particles drift)."""
import numpy as np
import pytest

import oracle_ctypes as O


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("name", ["sde_uniform", "sde_grid32", "sde_advection", "sde_advection_cspherical",
                                  "sde_advection_spherical"])
def test_oracle_sde_bit_exact_vs_reference(name):
    c = O.load_case(name)
    m = c["meta"]
    st = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    O.sde_propagate(O.sde_field(c), st, int(c["nsteps"].max()), rng=O.SDE_MT, **O.sde_args(m))
    assert bits_equal(st.pos, c["final_pos"]) and bits_equal(st.dir, c["final_dir"])
    assert bits_equal(st.trajectory_length, c["final_len"]) and bits_equal(st.next_step, c["final_next"])
    assert bits_equal(st.time, c["final_time"])
    assert np.array_equal(st.active, c["final_active"])


def test_oracle_sde_threads_one_thread_is_the_reference_stream():
    """The CPU baseline's threaded form seeds thread t with seed + t: one thread is the reference."""
    c = O.load_case("sde_uniform")
    st = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    n = O.sde_propagate_threads(O.sde_field(c), st, 10 ** 6, nthreads=1, **O.sde_args(c["meta"]))
    assert n == int(c["nsteps"].sum())
    assert bits_equal(st.pos, c["final_pos"]) and bits_equal(st.dir, c["final_dir"])
