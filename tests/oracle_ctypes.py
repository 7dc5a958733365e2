"""ctypes bindings to the oracle (oracle/liboracle.so): the CPU restatement of
the reference hot path.  Test infrastructure only -- the product never loads it.
"""
import ctypes as C
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")

LIBM_GLIBC = 0
LIBM_CR = 1
FIELD_NONE, FIELD_UNIFORM, FIELD_GRID, FIELD_PLANEWAVE, FIELD_MODULATED, FIELD_JF12 = 0, 1, 2, 3, 4, 5
FIELD_PERIODIC, FIELD_LIST, FIELD_EVOLUTION = 6, 7, 8
PROP_BP, PROP_CK = 0, 1
MOD_BOX, MOD_SPHERE, MOD_MAXLEN, MOD_OBSERVER = 1, 2, 3, 4  # orc_propagator.order

# SI units (include/crpropa/Units.h)
c_light = 2.99792458e8
eplus = 1.602176487e-19
pc = 648000 / np.pi * 149597870700.0
kpc = 1e3 * pc
Mpc = 1e6 * pc
Gpc = 1e9 * pc
eV = eplus * 1.0
EeV = 1e18 * eV
gauss = 1e-4
nG = 1e-9 * gauss
muG = 1e-6 * gauss

_P = C.POINTER
_dp = _P(C.c_double)
_fp = _P(C.c_float)
_ip = _P(C.c_int32)


# Grid::setInterpolationType values (crp_grid_interpolation)
GRID_TRILINEAR, GRID_TRICUBIC, GRID_NEAREST = 0, 1, 2
GRID_INTERP = {"TRILINEAR": GRID_TRILINEAR, "TRICUBIC": GRID_TRICUBIC, "NEAREST_NEIGHBOUR": GRID_NEAREST}


class Field(C.Structure):
    pass


Field._fields_ = [
        ("type", C.c_int),
        ("B0", C.c_double * 3),
        ("grid", _fp),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("origin", C.c_double * 3),
        ("spacing", C.c_double * 3),
        ("reflective", C.c_int), ("clip", C.c_int),
        ("nm", C.c_int),
        ("k", _dp), ("Ak", _dp), ("xi", _dp), ("kappa", _dp), ("beta", _dp),
        ("mod", _fp), ("mnx", C.c_int), ("mny", C.c_int), ("mnz", C.c_int),
        ("morigin", C.c_double * 3), ("mspacing", C.c_double * 3), ("mreflective", C.c_int), ("mclip", C.c_int),
        ("jf12", C.c_void_p),
        ("children", C.POINTER(C.POINTER(Field))), ("nchildren", C.c_int),
        ("extends", C.c_double * 3), ("porigin", C.c_double * 3), ("preflective", C.c_int),
        ("evo_m", C.c_double),
        ("interp", C.c_int), ("minterp", C.c_int),
]


class JF12Params(C.Structure):
    """crp_jf12_params: JF12Field's switches and constants"""
    _fields_ = [(n, C.c_int) for n in ("use_regular", "use_disk", "use_toroidal_halo", "use_x", "use_striated",
                                       "use_turbulent")] + \
               [("r_arms", C.c_double * 8), ("b_disk", C.c_double * 8)] + \
               [(n, C.c_double) for n in ("sin_pitch", "cos_pitch", "tan90_minus_pitch", "b_ring", "h_disk", "w_disk",
                                          "b_north", "b_south", "r_north", "r_south", "w_halo", "z0", "b_x",
                                          "sin_theta_x0", "cos_theta_x0", "tan_theta_x0", "r_xc", "r_x", "sqrt_beta")] + \
               [("b_disk_turb", C.c_double * 8)] + \
               [(n, C.c_double) for n in ("b_disk_turb5", "z_disk_turb", "b_halo_turb", "r_halo_turb", "z_halo_turb")]



class SDE(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("minStep", C.c_double), ("maxStep", C.c_double), ("epsilon", C.c_double),
                ("alpha", C.c_double), ("scale", C.c_double), ("advection", C.c_int), ("adv", C.c_double * 3),
                ("maxTrajectoryLength", C.c_double), ("rng", C.c_int), ("seed", C.c_uint32), ("adv_kind", C.c_int),
                ("adv_origin", C.c_double * 3), ("adv_par", C.c_double * 4), ("adv_n", C.c_int),
                ("adv_list_kind", C.c_int * 8), ("adv_list_par", (C.c_double * 9) * 8)]


SDE_MT, SDE_PHILOX = 0, 1


class Propagator(C.Structure):
    _fields_ = [("type", C.c_int), ("tolerance", C.c_double), ("minStep", C.c_double),
                ("maxStep", C.c_double), ("maxTrajectoryLength", C.c_double),
                ("obsTimes", _dp), ("nObs", C.c_int), ("obsDeactivate", C.c_int),
                ("box", C.c_int), ("boxOrigin", C.c_double * 3), ("boxSize", C.c_double * 3),
                ("sphere", C.c_int), ("sphereLimitStep", C.c_int), ("sphereCenter", C.c_double * 3),
                ("sphereRadius", C.c_double), ("sphereMargin", C.c_double), ("order", C.c_int * 4),
                ("nOrder", C.c_int)]


class State(C.Structure):
    _fields_ = [("pos", _dp), ("dir", _dp), ("energy", _dp), ("charge", _dp), ("redshift", _dp),
                ("current_step", _dp), ("next_step", _dp), ("trajectory_length", _dp), ("time", _dp),
                ("active", _ip), ("prev_pos", _dp), ("prev_dir", _dp),
                ("det_index", _ip), ("snap", _dp), ("n", C.c_size_t),
                ("source_pos", _dp), ("created_pos", _dp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "port"])
        L = C.CDLL(path)
        L.orc_propagate.restype = C.c_longlong
        L.orc_propagate.argtypes = [_P(Field), _P(Propagator), _P(State), C.c_size_t, C.c_longlong, C.c_int,
                                    C.c_int, _P(C.c_longlong)]
        L.orc_field_eval.argtypes = [_P(Field), _dp, C.c_size_t, _dp, C.c_int]
        L.orc_adv_eval.argtypes = [_P(SDE), _dp, C.c_size_t, _dp, C.c_int]
        L.orc_periodic_boundary.argtypes = [C.c_int, C.c_int]
        L.orc_reflective_boundary.argtypes = [C.c_int, C.c_int]
        L.orc_field_eval_z.argtypes = [_P(Field), _dp, _dp, C.c_size_t, _dp, C.c_int]
        L.orc_step_batch.argtypes = [_P(Field), C.c_int, C.c_size_t, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                     _dp, C.c_int]
        L.orc_planewave_init.argtypes = [C.c_double] * 6 + [C.c_int, C.c_int, C.c_int] + [_dp] * 5
        L.orc_cos_batch.argtypes = [_dp, C.c_size_t, _dp, C.c_int]
        L.orc_pow_batch.argtypes = [_dp, C.c_double, C.c_size_t, _dp, C.c_int]
        L.orc_noise_grid.argtypes = [C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_double, _fp]
        L.orc_rng_randint_batch.argtypes = [C.c_uint32, C.c_size_t, _P(C.c_uint32)]
        L.orc_cos_phases.argtypes = [_dp, C.c_size_t, _dp, _P(C.c_int32), _dp]
        L.orc_philox4x32_10.argtypes = [_P(C.c_uint32), _P(C.c_uint32), _P(C.c_uint32)]
        L.orc_sin_log_batch.argtypes = [_dp, C.c_size_t, _dp, _dp]
        L.orc_sde_propagate_threads.restype = C.c_longlong
        L.orc_sde_propagate_threads.argtypes = [_P(Field), _P(SDE), _P(State), C.c_size_t, C.c_longlong, C.c_int]
        L.orc_sde_propagate.restype = C.c_longlong
        L.orc_sde_propagate.argtypes = [_P(Field), _P(SDE), _P(State), C.c_size_t, C.c_longlong, _P(C.c_uint32),
                                        _P(C.c_uint64)]
        L.orc_turbulence_coeffs.restype = C.c_size_t
        L.orc_turbulence_coeffs.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int] + \
            [C.c_double] * 4 + [C.c_uint32, _fp]
        L.orc_sample_particles.argtypes = [C.c_uint32, C.c_size_t, C.c_double, C.c_double, _dp, C.c_size_t,
                                           _ip, _dp, _dp]
        adj = [_fp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_size_t, _dp, _dp,
               _dp]
        L.orc_adjoint_forward.argtypes = adj + [_dp, _dp]
        L.orc_adjoint_backward.argtypes = adj + [_dp] * 6
        _lib = L
    return _lib


def ptr(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


# ---------------------------------------------------------------- fixtures
def load_case(name):
    d = os.path.join(GOLDEN, name)
    out = {}
    for f in os.listdir(d):
        if f.endswith(".npy"):
            out[f[:-4]] = np.load(os.path.join(d, f), allow_pickle=False)
    mp = os.path.join(d, "meta.json")
    if os.path.exists(mp):
        with open(mp) as fh:
            out["meta"] = json.load(fh)
    return out


# ---------------------------------------------------------------- fields
class OracleField:
    """Holds numpy buffers alive for an orc_field."""

    def __init__(self):
        self.s = Field()
        self._keep = []

    @staticmethod
    def uniform(B):
        f = OracleField()
        f.s.type = FIELD_UNIFORM
        for a in range(3):
            f.s.B0[a] = B[a]
        return f

    @staticmethod
    def grid(values, origin, spacing, reflective=False, clip=False, interp=0):
        f = OracleField()
        v = np.ascontiguousarray(values, dtype=np.float32)
        f._keep.append(v)
        f.s.type = FIELD_GRID
        f.s.grid = ptr(v, _fp)
        f.s.nx, f.s.ny, f.s.nz = v.shape[:3]
        sp = np.broadcast_to(np.asarray(spacing, dtype=np.float64), (3,))
        for a in range(3):
            f.s.origin[a] = origin[a]
            f.s.spacing[a] = sp[a]
        f.s.reflective = int(reflective)
        f.s.clip = int(clip)
        f.s.interp = int(interp)
        return f

    @staticmethod
    def modulated(values, origin, spacing, mod, morigin, mspacing, reflective=False, mreflective=True,
                  clip=False, mclip=False, interp=0, minterp=0):
        """ModulatedMagneticFieldGrid: Grid3f values times a Grid1f (one float per point)."""
        f = OracleField.grid(values, origin, spacing, reflective, clip, interp)
        m = np.ascontiguousarray(mod, dtype=np.float32)
        f._keep.append(m)
        f.s.type = FIELD_MODULATED
        f.s.mod = ptr(m, _fp)
        f.s.mnx, f.s.mny, f.s.mnz = m.shape[:3]
        msp = np.broadcast_to(np.asarray(mspacing, dtype=np.float64), (3,))
        for a in range(3):
            f.s.morigin[a] = morigin[a]
            f.s.mspacing[a] = msp[a]
        f.s.mreflective = int(mreflective)
        f.s.mclip = int(mclip)
        f.s.minterp = int(minterp)
        return f

    @staticmethod
    def jf12(params, striated=None, striated_spacing=None, turbulent=None, turbulent_spacing=None):
        """JF12Field with the given JF12Params; striated (Nx,Ny,Nz) and turbulent (Nx,Ny,Nz,3) grids at the
        origin (periodic)"""
        if turbulent is not None:
            f = OracleField.grid(turbulent, [0, 0, 0], turbulent_spacing)
        else:
            f = OracleField()
        f.s.type = FIELD_JF12
        if striated is not None:
            m = np.ascontiguousarray(striated, dtype=np.float32)
            f._keep.append(m)
            f.s.mod = ptr(m, _fp)
            f.s.mnx, f.s.mny, f.s.mnz = m.shape[:3]
            for a in range(3):
                f.s.morigin[a] = 0.0
                f.s.mspacing[a] = striated_spacing
            f.s.mreflective = 0
        f._keep.append(params)
        f.s.jf12 = C.cast(C.pointer(params), C.c_void_p)
        return f

    @staticmethod
    def planewave(tables):
        f = OracleField()
        k, Ak, xi, kappa, beta = [np.ascontiguousarray(tables[n], dtype=np.float64)
                                  for n in ("k", "Ak", "xi", "kappa", "beta")]
        f._keep += [k, Ak, xi, kappa, beta]
        f.s.type = FIELD_PLANEWAVE
        f.s.nm = len(k)
        f.s.k, f.s.Ak, f.s.xi, f.s.kappa, f.s.beta = [ptr(a) for a in (k, Ak, xi, kappa, beta)]
        return f

    @staticmethod
    def _wrap(kind, children):
        f = OracleField()
        f._keep += list(children)
        arr = (C.POINTER(Field) * max(1, len(children)))(*[C.pointer(c.s) for c in children])
        f._keep.append(arr)
        f.s.type = kind
        f.s.children = C.cast(arr, C.POINTER(C.POINTER(Field)))
        f.s.nchildren = len(children)
        return f

    @staticmethod
    def periodic(child, extends, origin=(0, 0, 0), reflective=False):
        """PeriodicMagneticField(field, extends, origin, reflective)"""
        f = OracleField._wrap(FIELD_PERIODIC, [child])
        for a in range(3):
            f.s.extends[a] = extends[a]
            f.s.porigin[a] = origin[a]
        f.s.preflective = int(reflective)
        return f

    @staticmethod
    def field_list(children):
        """MagneticFieldList with addField(children[0]), addField(children[1]), ..."""
        return OracleField._wrap(FIELD_LIST, children)

    @staticmethod
    def evolution(child, m):
        """MagneticFieldEvolution(field, m)"""
        f = OracleField._wrap(FIELD_EVOLUTION, [child])
        f.s.evo_m = m
        return f

    def eval(self, pos, libm=LIBM_GLIBC, z=None):
        """getField(position) at each position, or getField(position, z[i]) when z is given."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        B = np.zeros_like(pos)
        if z is None:
            lib().orc_field_eval(C.byref(self.s), ptr(pos), pos.shape[0], ptr(B), libm)
        else:
            zz = np.ascontiguousarray(np.broadcast_to(np.asarray(z, dtype=np.float64), (pos.shape[0],)))
            lib().orc_field_eval_z(C.byref(self.s), ptr(pos), ptr(zz), pos.shape[0], ptr(B), libm)
        return B


def planewave_tables(meta):
    """Mode table as built by the PlaneWaveTurbulence constructor (host glibc)."""
    nm = int(meta["Nm"])
    k = np.zeros(nm)
    Ak = np.zeros(nm)
    xi = np.zeros((nm, 3))
    kappa = np.zeros((nm, 3))
    beta = np.zeros(nm)
    simple = meta["spectrum"] == "simple"
    lb = 1000 * meta["lMax"] if simple else meta.get("lBendover", 1.0)
    q = 0.0 if simple else meta.get("qIndex", 4.0)
    lib().orc_planewave_init(meta["Brms"], meta["lMin"], meta["lMax"], lb, meta["sIndex"], q, int(simple), nm,
                             int(meta["seed"]), ptr(k), ptr(Ak), ptr(xi), ptr(kappa), ptr(beta))
    return {"k": k, "Ak": Ak, "xi": xi, "kappa": kappa, "beta": beta}


def noise_grid(seed, n, amp):
    g = np.zeros((n, n, n, 3), dtype=np.float32)
    lib().orc_noise_grid(seed, n, n, n, amp, ptr(g, _fp))
    return g


def sample_particles(seed, n, emin, emax, weights=(1.0,)):
    w = np.asarray(weights, dtype=np.float64)
    sp = np.zeros(n, dtype=np.int32)
    E = np.zeros(n)
    d = np.zeros((n, 3))
    lib().orc_sample_particles(seed, n, emin, emax, ptr(w), len(w), ptr(sp, _ip), ptr(E), ptr(d))
    return sp, E, d


# ---------------------------------------------------------------- propagation
def normalise(d):
    """ParticleState::setDirection: dir / dir.getR() (src/ParticleState.cpp:30-32)."""
    d = np.asarray(d, dtype=np.float64)
    n = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    return d / n[:, None]


class OracleState:
    def __init__(self, pos, dir, energy, charge, next_step=None, normalised=False):
        """dir: raw directions, normalised here as ParticleState::setDirection does, or (normalised=True)
        directions that already went through setDirection, taken as they are."""
        n = len(energy)
        self.n = n
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).copy()
        self.dir = np.ascontiguousarray(dir, dtype=np.float64).copy() if normalised else normalise(dir)
        self.energy = np.ascontiguousarray(energy, dtype=np.float64).copy()
        self.charge = np.ascontiguousarray(charge, dtype=np.float64).copy()
        self.redshift = np.zeros(n)
        self.current_step = np.zeros(n)
        self.next_step = np.zeros(n) if next_step is None else np.array(next_step, dtype=np.float64)
        self.trajectory_length = np.zeros(n)
        self.time = np.zeros(n)
        self.active = np.ones(n, dtype=np.int32)
        self.prev_pos = self.pos.copy()
        self.prev_dir = self.dir.copy()
        self.det_index = np.zeros(n, dtype=np.int32)
        self.snap = None  # (K, n, 9) once an observer is used
        self.source_pos = self.pos.copy()
        self.created_pos = self.pos.copy()

    def struct(self):
        s = State()
        for name in ("pos", "dir", "energy", "charge", "redshift", "current_step", "next_step",
                     "trajectory_length", "time", "prev_pos", "prev_dir"):
            setattr(s, name, ptr(getattr(self, name)))
        s.active = ptr(self.active, _ip)
        s.det_index = ptr(self.det_index, _ip)
        s.source_pos = ptr(self.source_pos)
        s.created_pos = ptr(self.created_pos)
        s.snap = ptr(self.snap) if self.snap is not None else None
        s.n = self.n
        return s


def propagate(field, prop_type, state, steps, tolerance=0.42, min_step=0.0, max_step=0.0, max_len=0.0,
              libm=LIBM_GLIBC, nthreads=0, observer_times=None, observer_deactivate=False, box=None,
              sphere=None, order=None):
    """observer_times: detection lengths of an Observer(ObserverTimeEvolution) after the length limit;
    detections land in state.snap[k, i, :] = (pos, dir, length, time, nextStep)."""
    p = Propagator(prop_type, tolerance, min_step, max_step, max_len)
    if box is not None:  # (origin[3], size[3]) of a PeriodicBox right after the propagator
        p.box = 1
        for a in range(3):
            p.boxOrigin[a] = box[0][a]
            p.boxSize[a] = box[1][a]
    if order is not None:  # module names after the propagator in ModuleList order
        ids = {"box": MOD_BOX, "sphere": MOD_SPHERE, "maxlen": MOD_MAXLEN, "observer": MOD_OBSERVER}
        for a, nm in enumerate(order):
            p.order[a] = ids[nm]
        p.nOrder = len(order)
    if sphere is not None:  # (center[3], radius, margin, limit_step) of a SphericalBoundary after the box
        p.sphere = 1
        for a in range(3):
            p.sphereCenter[a] = sphere[0][a]
        p.sphereRadius, p.sphereMargin, p.sphereLimitStep = sphere[1], sphere[2], int(sphere[3])
    keep = None
    if observer_times is not None and len(observer_times):
        keep = np.ascontiguousarray(observer_times, dtype=np.float64)
        p.obsTimes = ptr(keep)
        p.nObs = len(keep)
        p.obsDeactivate = int(bool(observer_deactivate))
        if state.snap is None or state.snap.shape[0] < len(keep):
            state.snap = np.full((len(keep), state.n, 9), np.nan)
    s = state.struct()
    trials = C.c_longlong(0)
    f = C.byref(field.s) if field is not None else None
    n = lib().orc_propagate(f, C.byref(p), C.byref(s), state.n, steps, libm, nthreads, C.byref(trials))
    return n, trials.value


def _adjoint_args(grid, origin, spacing, scale, h, nsteps, x0, u0, kq):
    grid = np.ascontiguousarray(grid, dtype=np.float32)
    nx, ny, nz = grid.shape[:3]
    o = np.ascontiguousarray(origin, dtype=np.float64)
    sp = np.ascontiguousarray(spacing, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    kq = np.ascontiguousarray(kq, dtype=np.float64)
    keep = (grid, o, sp, x0, u0, kq)
    args = [ptr(grid, _fp), nx, ny, nz, ptr(o), ptr(sp), float(scale), float(h), int(nsteps), len(kq), ptr(x0),
            ptr(u0), ptr(kq)]
    return keep, args


def adjoint_forward(grid, origin, spacing, scale, h, nsteps, x0, u0, kq):
    """FP64 fixed-step Boris model through scale * trilinear(grid): (x_final, u_final)."""
    keep, args = _adjoint_args(grid, origin, spacing, scale, h, nsteps, x0, u0, kq)
    n = len(keep[5])
    xo = np.zeros((n, 3))
    uo = np.zeros((n, 3))
    lib().orc_adjoint_forward(*args, ptr(xo), ptr(uo))
    return xo, uo


def adjoint_backward(grid, origin, spacing, scale, h, nsteps, x0, u0, kq, gx_out, gu_out):
    """Reverse-mode gradient of the same model: (grad_grid, grad_x0, grad_u0, grad_scale)."""
    keep, args = _adjoint_args(grid, origin, spacing, scale, h, nsteps, x0, u0, kq)
    n = len(keep[5])
    gxo = None if gx_out is None else np.ascontiguousarray(gx_out, dtype=np.float64)
    guo = np.ascontiguousarray(gu_out, dtype=np.float64)
    gg = np.zeros(keep[0].shape)
    gx0 = np.zeros((n, 3))
    gu0 = np.zeros((n, 3))
    gs = np.zeros(1)
    lib().orc_adjoint_backward(*args, ptr(gxo), ptr(guo), ptr(gg), ptr(gx0), ptr(gu0), ptr(gs))
    return gg, gx0, gu0, gs[0]


def cos_batch(x, libm):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().orc_cos_batch(ptr(x), x.size, ptr(out), libm)
    return out


def pow_batch(x, y, libm):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib().orc_pow_batch(ptr(x), y, x.size, ptr(out), libm)
    return out


def field_for_case(case):
    """Build the oracle field a golden trajectory case ran in."""
    meta = case["meta"]
    f = meta["field"] if "field" in meta else None
    ip = GRID_INTERP[meta.get("interpolation", "TRILINEAR")]
    if f == "grid32":
        g = load_case("grid32")
        gm = g["meta"]["grid"]
        return OracleField.grid(g["grid"], gm["origin"], gm["spacing"], reflective=meta.get("reflective", False),
                                clip=meta.get("clip", False), interp=ip)
    if f == "grid32_modulated":
        g = load_case("grid32")
        gm = g["meta"]["grid"]
        return OracleField.modulated(g["grid"], gm["origin"], gm["spacing"], case["mod"], meta["mod_origin"],
                                     meta["mod_spacing"], mreflective=meta["mod_reflective"], interp=ip, minterp=ip)
    if f in ("list(grid32, uniform [1 nG, 0, 0])", "periodic(grid32) reflective", "evolution(grid32, m=1.5)"):
        g = load_case("grid32")
        gm = g["meta"]["grid"]
        grid = OracleField.grid(g["grid"], gm["origin"], gm["spacing"])
        if f.startswith("list"):
            return OracleField.field_list([grid, OracleField.uniform([1 * nG, 0, 0])])
        if f.startswith("periodic"):
            return OracleField.periodic(grid, meta["extends"], meta["origin"], reflective=True)
        return OracleField.evolution(grid, meta["evolution_m"])
    if isinstance(f, dict) and f["type"] == "planewave":
        return OracleField.planewave(planewave_tables(f))
    if isinstance(f, dict) and f["type"] == "uniform":
        return OracleField.uniform(f["B"])
    if "grid" in meta:  # generated noise grid (C1-shaped case)
        gm = meta["grid"]
        vals = noise_grid(gm["seed"], gm["n"], gm["amp"])
        return OracleField.grid(vals, gm["origin"], gm["spacing"])
    raise ValueError("unknown field in case")


STEP_BP_DY, STEP_BP_TRY, STEP_CK_DYDT, STEP_CK_TRY = 0, 1, 2, 3


def step_batch(field, op, x, u, h, z, q, m, libm=LIBM_CR):
    """The reference's single-step functions on the oracle (orc_step_batch): returns (out_x, out_u, err_x,
    err_u) as (n, 3) arrays (the errors are zero for dY / dYdt)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[0]
    u = np.ascontiguousarray(u, dtype=np.float64)
    arr = [np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,))) for a in (h, z, q, m)]
    outs = [np.zeros((n, 3)) for _ in range(4)]
    lib().orc_step_batch(C.byref(field.s), op, n, ptr(x), ptr(u), *[ptr(a) for a in arr], *[ptr(o) for o in outs],
                         libm)
    return outs


def decorator_fields():
    """The fields of the decorators32 samples (gen_golden.cpp gen_decorators), by fixture name."""
    c = load_case("decorators32")
    m = c["meta"]
    g = load_case("grid32")
    gm = g["meta"]["grid"]
    grid = OracleField.grid(g["grid"], gm["origin"], gm["spacing"])
    uni = OracleField.uniform(m["uniform"])
    pw = OracleField.planewave(planewave_tables(load_case("planewave_simple2")["meta"]))
    inner = OracleField.field_list([grid, uni])
    L = gm["n"] * gm["spacing"]
    fields = {
        "B_periodic": OracleField.periodic(grid, m["extends"], m["origin"]),
        "B_reflective": OracleField.periodic(grid, m["extends"], m["origin"], reflective=True),
        "B_list": OracleField.field_list([grid, uni, pw]),
        "B_evolution": OracleField.evolution(grid, m["evolution_m"]),
        "B_nested": OracleField.field_list([OracleField.periodic(grid, m["extends"], m["origin"], reflective=True),
                                            OracleField.evolution(grid, m["nested_evolution_m"]),
                                            OracleField.periodic(inner, m["nested_extends"])]),
        "B_evolution_list": OracleField.evolution(inner, m["evolution_list_m"]),
    }
    assert L > 0
    return c, fields


def propagator_args(meta):
    if meta["propagator"] == "BP":
        if "fixedStep" in meta:
            return dict(prop_type=PROP_BP, tolerance=0.42, min_step=meta["fixedStep"], max_step=meta["fixedStep"])
        return dict(prop_type=PROP_BP, tolerance=meta["tolerance"], min_step=meta["minStep"],
                    max_step=meta["maxStep"])
    return dict(prop_type=PROP_CK, tolerance=meta["tolerance"], min_step=meta["minStep"], max_step=meta["maxStep"])


def module_args(meta):
    """propagate() keyword arguments for the modules a golden case lists after its propagator."""
    kw = {}
    if "box" in meta:
        kw["box"] = (meta["box"]["origin"], meta["box"]["size"])
    if "sphere" in meta:
        sp = meta["sphere"]
        kw["sphere"] = (sp["center"], sp["radius"], sp["margin"], sp["limitStep"])
    if "order" in meta:
        kw["order"] = list(meta["order"]) + (["maxlen"] if "maxTrajectoryLength" in meta and "maxlen" not in meta["order"]
                                             else [])
    return kw


TURB_SIMPLE, TURB_GENERAL, TURB_GENERAL_POWERLAW = 0, 1, 2


def turbulence_coeffs(n, spacing, lmin, lmax, kind=TURB_SIMPLE, alpha=-11 / 3., lbendover=1.0, sindex=5 / 3.,
                      qindex=4.0, seed=1):
    """k-space field of the reference's turbulence generators (glibc libm, MT19937 seeded with seed):
    complex64 array (3, n, n, n//2 + 1) and the number of modes in range."""
    n2 = n // 2 + 1
    bk = np.zeros((3, n, n, n2, 2), dtype=np.float32)
    modes = lib().orc_turbulence_coeffs(n, spacing, lmin, lmax, kind, alpha, lbendover, sindex, qindex, seed,
                                        ptr(bk, _fp))
    return bk[..., 0] + 1j * bk[..., 1], modes


def turbulence_grid(n, spacing, lmin, lmax, **kw):
    """The unnormalised grid: the coefficients through an unnormalised inverse real FFT in double
    precision (GridTurbulence.cpp:147-185), as an (n, n, n, 3) float64 array."""
    X, modes = turbulence_coeffs(n, spacing, lmin, lmax, **kw)
    out = np.stack([np.fft.irfftn(X[c].astype(np.complex128), s=(n, n, n), axes=(0, 1, 2), norm="forward")
                    for c in range(3)], axis=-1)
    return out, modes


def randint_batch(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    lib().orc_rng_randint_batch(seed, n, out.ctypes.data_as(_P(C.c_uint32)))
    return out


def cos_phases(x):
    """(quick value, decided flag, accurate value) of the correctly rounded cos per argument."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    q, a = np.zeros_like(x), np.zeros_like(x)
    d = np.zeros(x.shape, dtype=np.int32)
    lib().orc_cos_phases(ptr(x), x.size, ptr(q), d.ctypes.data_as(_P(C.c_int32)), ptr(a))
    return q, d.astype(bool), a


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def sin_log(x):
    """(correctly rounded sin, rounded double-double log) of this package's math header."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    s, lg = np.zeros_like(x), np.zeros_like(x)
    lib().orc_sin_log_batch(ptr(x), x.size, ptr(s), ptr(lg))
    return s, lg


def atan2_phases(y, x):
    """(decided flag, quick value, accurate value) of the correctly rounded atan2 per argument pair"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    d = np.zeros(x.size, dtype=np.int32)
    q, a = np.zeros_like(x), np.zeros_like(x)
    L = lib()
    L.orc_atan2_phases.argtypes = [_P(C.c_double), _P(C.c_double), C.c_size_t, _ip, _P(C.c_double), _P(C.c_double)]
    L.orc_atan2_phases(ptr(y), ptr(x), x.size, ptr(d, _ip), ptr(q), ptr(a))
    return d.astype(bool), q, a


def exp_phases(x):
    """(decided flag, quick value, accurate value) of the correctly rounded exp per argument"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    d = np.zeros(x.size, dtype=np.int32)
    q, a = np.zeros_like(x), np.zeros_like(x)
    L = lib()
    L.orc_exp_phases.argtypes = [_P(C.c_double), C.c_size_t, _ip, _P(C.c_double), _P(C.c_double)]
    L.orc_exp_phases(ptr(x), x.size, ptr(d, _ip), ptr(q), ptr(a))
    return d.astype(bool), q, a


def fast_sincos_log(x):
    """crp_sincos_fast and crp_log_fast element-wise: (sin, cos, log)"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    sn, cs, lg = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    L = lib()
    L.orc_fast_batch.argtypes = [_P(C.c_double), C.c_size_t, _P(C.c_double), _P(C.c_double), _P(C.c_double)]
    L.orc_fast_batch(ptr(x), x.size, ptr(sn), ptr(cs), ptr(lg))
    return sn, cs, lg


def exp_atan2(x, y):
    """crp_exp_rn(x) and crp_atan2_rn(y, x) element-wise"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    e, a = np.zeros_like(x), np.zeros_like(x)
    L = lib()
    L.orc_exp_atan2_batch.argtypes = [_P(C.c_double), _P(C.c_double), C.c_size_t, _P(C.c_double), _P(C.c_double)]
    L.orc_exp_atan2_batch(ptr(x), ptr(y), x.size, ptr(e), ptr(a))
    return e, a


def jf12_defaults(planck=False):
    p = JF12Params()
    lib().orc_jf12_defaults(int(planck), C.byref(p))
    return p


def jf12_striated(seed, n=100):
    """JF12Field::randomStriated(seed) (src/magneticField/JF12Field.cpp:81-94): round(rand()) * 2 - 1 per
    point in storage order, rand() = randInt() / (2^32 - 1)"""
    r = randint_batch(seed, n ** 3).astype(np.float64) * (1.0 / 4294967295.0)
    return (np.round(r) * 2 - 1).astype(np.float32).reshape(n, n, n)


def sde_propagate(field, state, steps, tolerance, min_step, max_step, epsilon=0.1, alpha=1 / 3., scale=1.0,
                  advection=None, max_len=0.0, rng=SDE_MT, seed=1, sde_step=None, sde_stream=None):
    """DiffusionSDE steps (src/module/DiffusionSDE.cpp) on the oracle state; rng SDE_MT follows the
    reference's stream (serial, particle by particle), SDE_PHILOX the device's (sde_step: uint32 step
    counters, updated; sde_stream: uint64 stream numbers, default the particle index)."""
    p = SDE(tolerance, min_step, max_step, epsilon, alpha, scale, 0, (C.c_double * 3)(0, 0, 0), max_len, rng, seed)
    set_advection(p, advection)
    s = state.struct()
    cnt = None
    if sde_step is not None:
        assert sde_step.dtype == np.uint32 and sde_step.flags.c_contiguous
        cnt = sde_step.ctypes.data_as(_P(C.c_uint32))
    stream = None
    if sde_stream is not None:
        sde_stream = np.ascontiguousarray(sde_stream, dtype=np.uint64)
        stream = sde_stream.ctypes.data_as(_P(C.c_uint64))
    return lib().orc_sde_propagate(C.byref(field.s), C.byref(p), C.byref(s), state.n, steps, cnt, stream)


# crp_sde_advection_kind
ADV_UNIFORM, ADV_CONSTANT_SPHERICAL, ADV_SPHERICAL, ADV_CARTESIAN_SHOCK, ADV_SPHERICAL_SHOCK_1D, ADV_OBLIQUE_SHOCK, \
    ADV_SPHERICAL_SHOCK = 1, 2, 3, 4, 5, 6, 7


def advection_eval(terms, pos, libm=LIBM_GLIBC):
    """the velocity of an advection-field term list [(kind, [parameters]), ...] at positions (n, 3)"""
    p = SDE()
    set_advection(p, {"terms": terms})
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    v = np.zeros_like(pos)
    lib().orc_adv_eval(C.byref(p), ptr(pos), pos.shape[0], ptr(v), libm)
    return v


def set_advection(p, advection):
    """advection: None, a uniform velocity (3 numbers), a dict {"kind": 2|3, "origin": [..], "par": [..]},
    or {"terms": [(kind, [parameters in the C ABI's order]), ...]} summed as AdvectionFieldList does"""
    if advection is None:
        return
    p.advection = 1
    if isinstance(advection, dict) and "terms" in advection:
        p.adv_n = len(advection["terms"])
        for t, (kind, par) in enumerate(advection["terms"]):
            p.adv_list_kind[t] = kind
            for i, v in enumerate(par):
                p.adv_list_par[t][i] = v
        return
    if isinstance(advection, dict):
        p.adv_kind = advection["kind"]
        for a in range(3):
            p.adv_origin[a] = advection["origin"][a]
        for a, v in enumerate(advection["par"]):
            p.adv_par[a] = v
    else:
        p.adv_kind = 1
        for a in range(3):
            p.adv[a] = advection[a]


def sde_propagate_threads(field, state, steps, tolerance, min_step, max_step, epsilon=0.1, alpha=1 / 3., scale=1.0,
                          advection=None, max_len=0.0, seed=1, nthreads=0):
    """DiffusionSDE with one MT19937 per OpenMP thread (the reference's threading); returns the steps."""
    p = SDE(tolerance, min_step, max_step, epsilon, alpha, scale, 0, (C.c_double * 3)(0, 0, 0), max_len, SDE_MT, seed)
    set_advection(p, advection)
    s = state.struct()
    return lib().orc_sde_propagate_threads(C.byref(field.s), C.byref(p), C.byref(s), state.n, steps, nthreads)


def sde_args(meta):
    kw = dict(tolerance=meta["tolerance"], min_step=meta["minStep"], max_step=meta["maxStep"], epsilon=meta["epsilon"],
              alpha=meta["alpha"], scale=meta["scale"], max_len=meta.get("maxTrajectoryLength", 0.0), seed=meta["seed"])
    if "advection" in meta:
        kw["advection"] = meta["advection"]
    if "advection_kind" in meta:
        kw["advection"] = {"kind": meta["advection_kind"], "origin": meta["advection_origin"],
                           "par": meta["advection_par"]}
    if "advection_terms" in meta:
        kw["advection"] = {"terms": [(k, par) for k, par in meta["advection_terms"]]}
    return kw


def sde_field(case):
    m = case["meta"]
    if "grid_spacing" in m:
        return OracleField.grid(case["grid"], [0, 0, 0], m["grid_spacing"])
    return OracleField.uniform(m["field"]["B"])
