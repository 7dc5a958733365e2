"""The oracle's tricubic and nearest-neighbour grid interpolation (Grid::setInterpolationType and
Grid::interpolate, include/crpropa/Grid.h:255-264, 318-335, 383-536), pinned against the reference:
field samples and trajectories written by the reference's SIMD_EXTENSIONS=avx build (the only build
that evaluates tricubic Grid3f; tests/golden/gen_golden.cpp, gen_interp), the known answers of
test/testCore.cpp, and what the mathematics fixes: Catmull-Rom cubics reproduce polynomials of degree
two along each axis, the nearest point of a grid by brute force, periodicity and mirror symmetry."""
import numpy as np
import pytest

import oracle_ctypes as O

MODES = [("tricubic_periodic", O.GRID_TRICUBIC, False, False),
         ("tricubic_reflective", O.GRID_TRICUBIC, True, False),
         ("tricubic_clip", O.GRID_TRICUBIC, False, True),
         ("nearest_periodic", O.GRID_NEAREST, False, False),
         ("nearest_reflective", O.GRID_NEAREST, True, False)]
TRAJECTORIES = ["ck_grid32_tricubic", "bp_fixed_grid32_tricubic_reflective", "bp_fixed_grid32_nearest_reflective",
                "ck_grid32_modulated_tricubic"]


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def grid1f_eval(mod, morigin, mspacing, pos, minterp, reflective=False, clip=False):
    """Grid1f::interpolate, read through ModulatedMagneticFieldGrid with a one-point Grid3f of value
    (1, 1, 1) taken nearest-neighbour: B = 1 * m exactly."""
    ones = np.ones((1, 1, 1, 3), np.float32)
    f = O.OracleField.modulated(ones, [0, 0, 0], [1, 1, 1], mod, morigin, mspacing, mreflective=reflective,
                                mclip=clip, interp=O.GRID_NEAREST, minterp=minterp)
    B = f.eval(np.atleast_2d(pos))
    assert np.array_equal(B[:, 0], B[:, 1]) and np.array_equal(B[:, 0], B[:, 2])
    return B[:, 0]


def float_eq(a, b, ulps=4):
    """gtest's EXPECT_FLOAT_EQ: within 4 units in the last place of float"""
    a, b = np.float32(a), np.float32(b)
    ia, ib = int(np.array(a).view(np.int32)), int(np.array(b).view(np.int32))
    ia = ia if ia >= 0 else -(ia & 0x7fffffff)
    ib = ib if ib >= 0 else -(ib & 0x7fffffff)
    return abs(ia - ib) <= ulps


# ---------------------------------------------------------------- the reference's own output
@pytest.mark.parametrize("mode,ip,refl,clip", MODES)
def test_grid3f_samples_bit_exact_vs_reference(mode, ip, refl, clip):
    c = O.load_case("interp_grids")
    m = c["meta"]
    f = O.OracleField.grid(c["grid"], m["origin"], m["spacing"], reflective=refl, clip=clip, interp=ip)
    B = f.eval(c["pos"])
    assert bits_equal(B, c["B_" + mode])
    assert np.count_nonzero(B) > (1000 if clip else 10000)


@pytest.mark.parametrize("mode,ip,refl,clip", MODES)
def test_grid1f_samples_bit_exact_vs_reference(mode, ip, refl, clip):
    c = O.load_case("interp_grids")
    m = c["meta"]
    S = grid1f_eval(c["mod"], m["mod_origin"], m["mod_spacing"], c["pos"], ip, refl, clip)
    assert np.array_equal(S, c["m_" + mode].astype(np.float64))


@pytest.mark.parametrize("name", TRAJECTORIES)
def test_trajectories_bit_exact_vs_reference(name):
    """PropagationBP / PropagationCK through grid32 read tricubic or nearest-neighbour (adaptive CK with
    the host's glibc pow, as the reference ran)."""
    c = O.load_case(name)
    m = c["meta"]
    assert m["interpolation"] in ("TRICUBIC", "NEAREST_NEIGHBOUR")
    st = O.OracleState(c["init_pos"], c["init_dir"], c["energy"], c["charge"])
    done = 0
    for si, s in enumerate(c["snap_steps"]):
        O.propagate(O.field_for_case(c), steps=int(s) - done, max_len=m.get("maxTrajectoryLength", 0.0),
                    libm=O.LIBM_GLIBC, state=st, **O.propagator_args(m))
        done = int(s)
        assert bits_equal(st.pos, c["pos"][si]) and bits_equal(st.dir, c["dir"][si]), (name, s)


def test_closest_value_rounds_as_the_reference():
    """crpropa::round (Grid.h:61-63) rounds r + 0.5 before the floor: r = 0.5 - 2^-54 picks the upper
    point, where the C library's round would pick the lower one.  The reference samples hold 19
    positions of this kind; here one is made on purpose."""
    vals = np.array([1.0, 2.0], np.float32).reshape(2, 1, 1)
    r = 0.5 - 2.0 ** -54
    assert r < 0.5 and r + 0.5 == 1.0
    pos = np.array([[0.5 + r, 0.5, 0.5]])  # gridOrigin 0.5, spacing 1: r exactly
    assert grid1f_eval(vals, [0, 0, 0], [1, 1, 1], pos, O.GRID_NEAREST)[0] == 2.0


# ---------------------------------------------------------------- test/testCore.cpp known answers
def test_boundary_indices_known_answers():
    """periodicBoundary / reflectiveBoundary, test/testCore.cpp:458-507; the mirror loop equals
    i mod 2n folded onto [0, n) (the closed form the kernels use) for every index tried."""
    L = O.lib()
    for i, want in [(0, 0), (7, 7), (8, 0), (9, 1)]:
        assert L.orc_periodic_boundary(i, 8) == want
    for i, want in [(8, 7), (9, 6), (0, 0), (-1, 0), (-8, 7), (-9, 7)]:
        assert L.orc_reflective_boundary(i, 8) == want
    for n in (1, 2, 3, 8):
        for i in range(-5 * n - 3, 5 * n + 4):
            m = i % (2 * n)
            assert L.orc_reflective_boundary(i, n) == (m if m < n else 2 * n - 1 - m), (i, n)


def test_grid3f_tricubic_known_answers():
    """TEST(Grid3f, Interpolation), test/testCore.cpp:654-715: one value 1.7 at (0, 0, 1) of a 3^3 grid
    of spacing 2.793; the reference's expected tricubic values, to EXPECT_FLOAT_EQ."""
    sp = 2.793
    g = np.zeros((3, 3, 3, 3), np.float32)
    g[0, 0, 1] = (1.7, 0, 0)
    f = O.OracleField.grid(g, [0, 0, 0], sp, interp=O.GRID_TRICUBIC)
    for p, want in [((0.5, 0.5, 1.5), 1.7), ((0.5, 0.5, 1.4), 1.66005015373), ((0.5, 0.5, 1.6), 1.66005003452),
                    ((0.5, 0.35, 1.6), 1.57507634163), ((0.5, 2.65, 1.6), 0.190802007914)]:
        b = f.eval(np.array([p]) * sp)[0]
        assert float_eq(b[0], want), (p, b[0], want)
    lin = O.OracleField.grid(g, [0, 0, 0], sp)
    assert float_eq(lin.eval(np.array([[0.5, 0.5, 1.4]]) * sp)[0, 0], 1.7 * 0.9)


@pytest.mark.parametrize("ip", [O.GRID_TRICUBIC, O.GRID_NEAREST])
def test_grid1f_value_at_grid_point(ip):
    """TEST(Grid1f, SimpleTest / GridPropertiesConstructor / TestVectorSpacing), test/testCore.cpp:509-604:
    at a grid point every interpolation type returns the point's value; clipVolume gives 0 outside."""
    for spacing, idx, val in [((2.0, 2.0, 2.0), (2, 3, 4), 7.0), ((1.0, 5.0, 3.0), (1, 7, 6), 12.0)]:
        g = np.zeros((5, 8, 10), np.float32)
        g[idx] = val
        origin = np.array([1.0, 2.0, 3.0])
        p = origin + np.array(idx) * np.array(spacing) + np.array(spacing) / 2
        assert float_eq(grid1f_eval(g, origin, spacing, p, ip)[0], val)
    assert grid1f_eval(g, origin, (2.0, 2.0, 2.0), [100, 0, 12], ip, clip=True)[0] == 0.0


def test_grid1f_closest_value_known_answers():
    """TEST(Grid1f, ClosestValue), test/testCore.cpp:606-629"""
    g = np.arange(1, 9, dtype=np.float32).reshape(2, 2, 2)
    for p, want in [((0.1, 0.2, 0.6), 1), ((0.2, 0.1, 1.3), 2), ((0.3, 1.2, 0.2), 3), ((1.7, 1.8, 0.4), 7)]:
        assert grid1f_eval(g, [0, 0, 0], [1, 1, 1], p, O.GRID_NEAREST)[0] == want


@pytest.mark.parametrize("ip", [O.GRID_TRICUBIC, O.GRID_NEAREST])
def test_grid3f_periodicity_and_reflectivity(ip):
    """TEST(Grid3f, Periodicity / Reflectivity), test/testCore.cpp:717-874: B(x + a n s) = B(x)
    periodically, B(x + a) = B(-x - a) reflectively, 0 outside a clipped volume."""
    n, sp = 3, 3.0
    size = n * sp
    ix, iy, iz = np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij")
    g = np.stack([iz + ix, iy * iz, ix - iz * iy], axis=-1).astype(np.float32)
    pos = np.array([1.2, 2.3, 0.7])
    f = O.OracleField.grid(g, [0, 0, 0], sp, interp=ip)
    b = f.eval(pos[None])[0]
    for shift in ([1, 0, 0], [0, 5, 0], [0, 0, -2]):
        b2 = f.eval((pos + np.array(shift) * size)[None])[0]
        assert all(float_eq(x, y) for x, y in zip(b, b2)), shift
    r = O.OracleField.grid(g, [0, 0, 0], sp, reflective=True, interp=ip)
    for shift in ([1, 0, 0], [0, 5, 0], [0, 0, -2]):
        s = np.array(shift) * sp
        b1, b2 = r.eval(np.array([pos + s, -pos - s]))
        assert all(float_eq(x, y) for x, y in zip(b1, b2)), shift
    c = O.OracleField.grid(g, [0, 0, 0], sp, clip=True, interp=ip)
    assert np.all(c.eval((pos + np.array([0, 0, -2]) * size)[None]) == 0)


# ---------------------------------------------------------------- what the mathematics fixes
def quadratic_grid(shape, seed):
    """values of a polynomial of degree 2 in each unit-grid coordinate, and the polynomial"""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (3, 3, 3))  # coefficient of u^i v^j w^k

    def q(u, v, w):
        return sum(a[i, j, k] * u ** i * v ** j * w ** k for i in range(3) for j in range(3) for k in range(3))

    u, v, w = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")
    return q(u, v, w), q


def test_tricubic_reproduces_quadratics():
    """Catmull-Rom interpolation reproduces every polynomial of degree two along an axis (the cubic
    convolution kernel with a = -1/2 has third-order accuracy), so applied along z, y and x it
    returns a tensor-product quadratic exactly up to rounding, away from the continuation.  Grid1f in
    FP64 arithmetic, Grid3f in FP32; trilinear interpolation misses by the curvature term."""
    shape = (9, 8, 10)
    origin, spacing = np.array([-2.0, 1.0, 0.5]), np.array([0.7, 1.3, 0.9])
    rng = np.random.default_rng(3)
    r = np.stack([rng.uniform(1, n - 3, 400) for n in shape], axis=1)  # neighbours -1..+2 inside the grid
    pos = origin + spacing / 2 + r * spacing
    vals = []
    for seed in range(3):
        v, q = quadratic_grid(shape, seed)
        vals.append((v.astype(np.float32), q(r[:, 0], r[:, 1], r[:, 2])))
    for v, want in vals:
        got = grid1f_eval(v, origin, spacing, pos, O.GRID_TRICUBIC)
        scale = np.abs(v).max()
        assert np.abs(got - want).max() <= 2e-6 * scale
        lin = grid1f_eval(v, origin, spacing, pos, O.GRID_TRILINEAR)
        assert np.abs(lin - want).max() > 1e-3 * scale
    g = np.stack([v for v, _ in vals], axis=-1)
    f = O.OracleField.grid(g, origin, spacing, interp=O.GRID_TRICUBIC)
    B = f.eval(pos)
    for a, (v, want) in enumerate(vals):
        assert np.abs(B[:, a] - want).max() <= 2e-5 * np.abs(v).max(), a


def test_nearest_neighbour_brute_force():
    """The nearest grid point by exhaustive search: inside the grid the closest sample position per
    axis, outside the periodic image (index mod n) or the mirror image about the faces."""
    shape = (6, 5, 7)
    origin, spacing = np.array([0.3, -1.0, 2.0]), np.array([1.1, 0.6, 0.8])
    rng = np.random.default_rng(11)
    g = rng.uniform(-1, 1, shape + (3,)).astype(np.float32)
    n = np.array(shape)
    r = np.stack([rng.uniform(-3 * k, 4 * k, 3000) for k in shape], axis=1)
    r = r[np.all(np.abs(r - np.round(r) - 0.5 * np.sign(r - np.round(r))) > 1e-6, axis=1)]  # no ties
    pos = origin + spacing / 2 + r * spacing
    cand = np.arange(-4 * int(n.max()), 5 * int(n.max()))
    near = np.stack([cand[np.argmin(np.abs(r[:, a:a + 1] - cand[None, :]), axis=1)] for a in range(3)], axis=1)
    per = near % n
    got = O.OracleField.grid(g, origin, spacing, interp=O.GRID_NEAREST).eval(pos)
    assert np.array_equal(got, g[per[:, 0], per[:, 1], per[:, 2]].astype(np.float64))
    m = near % (2 * n)
    mir = np.where(m < n, m, 2 * n - 1 - m)
    got = O.OracleField.grid(g, origin, spacing, reflective=True, interp=O.GRID_NEAREST).eval(pos)
    assert np.array_equal(got, g[mir[:, 0], mir[:, 1], mir[:, 2]].astype(np.float64))
