"""ctypes binding of the C ABI in include/crpropa3_amd.h.

This is the thin layer a Python caller (or the reference's steering code)
uses to reach the HIP library.  It loads crpropa3_amd/libcrpropa3_amd.so and
raises if the library is missing: there is no CPU fallback.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcrpropa3_amd.so")

CRP_OK = 0
PRECISION_FP64 = 0
PRECISION_MIXED = 1

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_vp = C.c_void_p


class RunStats(C.Structure):
    _fields_ = [("accepted_steps", C.c_ulonglong), ("rejected_trials", C.c_ulonglong),
                ("field_evals", C.c_ulonglong), ("finished", C.c_ulonglong),
                ("kernel_ms", C.c_double), ("launches", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class HostState(C.Structure):
    _fields_ = [("pos", _dp), ("dir", _dp), ("energy", _dp), ("charge", _dp), ("current_step", _dp),
                ("next_step", _dp), ("trajectory_length", _dp), ("time", _dp), ("active", _ip),
                ("prev_pos", _dp), ("prev_dir", _dp), ("detection_index", _ip), ("source_pos", _dp),
                ("created_pos", _dp), ("sde_step", C.POINTER(C.c_uint32)), ("sde_stream", C.POINTER(C.c_uint64)),
                ("redshift", _dp)]


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


class LaunchConfig(C.Structure):
    """crp_launch_config: block size, workgroups per CU, steps per launch, hipGraph replay."""
    _fields_ = [("block", C.c_int), ("blocks_per_cu", C.c_int), ("steps_per_launch", C.c_longlong),
                ("use_graph", C.c_int), ("particles_per_lane", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class CrpError(RuntimeError):
    """Error reported by the C ABI (the reference raises std::runtime_error)."""


_lib = None


def lib():
    """The loaded HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("crpropa3_amd: %s is missing; build it with __graft_entry__.build() "
                              "or `make -C crpropa3_amd`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.crp_last_error.restype = C.c_char_p
        L.crp_field_create_uniform.argtypes = [_dp, C.POINTER(_vp)]
        L.crp_field_create_grid3f.argtypes = [_fp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                              C.POINTER(_vp)]
        L.crp_field_create_grid3f_device.argtypes = [_vp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                                     C.POINTER(_vp)]
        L.crp_field_create_modulated_grid3f.argtypes = [_fp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                                        _fp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                                        C.POINTER(_vp)]
        L.crp_field_create_planewave.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, C.POINTER(_vp)]
        L.crp_field_destroy.argtypes = [_vp]
        L.crp_jf12_default_params.argtypes = [C.c_int, C.POINTER(JF12Params)]
        L.crp_field_create_jf12.argtypes = [C.POINTER(JF12Params), _fp, C.POINTER(C.c_int), _dp, _dp, C.c_int, _fp,
                                            C.POINTER(C.c_int), _dp, _dp, C.c_int, C.c_int, C.POINTER(_vp)]
        L.crp_field_set_grid_layout.argtypes = [_vp, C.c_int]
        L.crp_field_get_grid_layout.argtypes = [_vp, C.POINTER(C.c_int)]
        L.crp_field_set_grid_storage.argtypes = [_vp, C.c_int]
        L.crp_field_set_grid_interpolation.argtypes = [_vp, C.c_int, C.c_int]
        L.crp_field_get_grid_interpolation.argtypes = [_vp, C.c_int, C.POINTER(C.c_int)]
        L.crp_field_eval_ex.argtypes = [_vp, _dp, C.c_size_t, _dp, C.c_int]
        L.crp_field_eval.argtypes = [_vp, _dp, C.c_size_t, _dp]
        L.crp_field_eval_noalloc.argtypes = [_vp, _dp, C.c_size_t, _dp, C.c_int]
        L.crp_field_create_periodic.argtypes = [_vp, _dp, _dp, C.c_int, C.POINTER(_vp)]
        L.crp_field_create_list.argtypes = [C.POINTER(_vp)]
        L.crp_field_list_add.argtypes = [_vp, _vp]
        L.crp_field_create_evolution.argtypes = [_vp, C.c_double, C.POINTER(_vp)]
        L.crp_field_eval_z.argtypes = [_vp, _dp, _dp, C.c_size_t, _dp]
        L.crp_bp_dy.argtypes = [_vp, C.c_size_t] + [_dp] * 8
        L.crp_bp_try_step.argtypes = [_vp, C.c_size_t] + [_dp] * 10
        L.crp_ck_dydt.argtypes = [_vp, C.c_size_t] + [_dp] * 7
        L.crp_ck_try_step.argtypes = [_vp, C.c_size_t] + [_dp] * 10
        for nm in ("crp_propagator_create_bp", "crp_propagator_create_ck"):
            getattr(L, nm).argtypes = [_vp, C.c_double, C.c_double, C.c_double, C.POINTER(_vp)]
        L.crp_propagator_create_sde.argtypes = [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.POINTER(_vp)]
        L.crp_propagator_set_sde.argtypes = [_vp, C.c_double, C.c_double, C.c_double, C.c_uint32]
        L.crp_propagator_set_sde_advection.argtypes = [_vp, _dp]
        L.crp_propagator_add_sde_advection.argtypes = [_vp, C.c_int, _dp, C.c_int]
        L.crp_propagator_clear_sde_advection.argtypes = [_vp]
        L.crp_propagator_set_sde_constant_spherical_advection.argtypes = [_vp, _dp, C.c_double]
        L.crp_propagator_set_sde_spherical_advection.argtypes = [_vp, _dp, C.c_double, C.c_double, C.c_double,
                                                                 C.c_double]
        L.crp_propagator_set_max_trajectory_length.argtypes = [_vp, C.c_double]
        L.crp_propagator_set_precision.argtypes = [_vp, C.c_int]
        L.crp_propagator_destroy.argtypes = [_vp]
        L.crp_propagator_set_observer.argtypes = [_vp, _dp, C.c_int, C.c_int]
        L.crp_propagator_set_periodic_box.argtypes = [_vp, _dp, _dp]
        L.crp_propagator_set_spherical_boundary.argtypes = [_vp, _dp, C.c_double, C.c_double, C.c_int]
        L.crp_propagator_set_module_order.argtypes = [_vp, C.POINTER(C.c_int), C.c_int]
        L.crp_batch_detections.argtypes = [_vp, C.c_int, _dp]
        L.crp_propagator_set_launch.argtypes = [_vp, C.POINTER(LaunchConfig)]
        L.crp_propagator_get_launch.argtypes = [_vp, C.POINTER(LaunchConfig)]
        L.crp_propagator_autotune.argtypes = [_vp, _vp, C.c_longlong, C.POINTER(LaunchConfig)]
        L.crp_batch_create.argtypes = [C.c_size_t, C.POINTER(_vp)]
        L.crp_batch_destroy.argtypes = [_vp]
        L.crp_batch_size.argtypes = [_vp, C.POINTER(C.c_size_t)]
        L.crp_batch_upload.argtypes = [_vp, C.c_size_t, C.POINTER(HostState)]
        L.crp_batch_download.argtypes = [_vp, C.POINTER(HostState)]
        L.crp_batch_copy.argtypes = [_vp, _vp]
        L.crp_propagate.argtypes = [_vp, _vp, C.c_longlong, C.POINTER(RunStats)]
        L.crp_turbulence_mode_count.argtypes = [C.POINTER(TurbulenceDesc), C.POINTER(C.c_uint64)]
        L.crp_turbulence_generate.argtypes = [C.POINTER(TurbulenceDesc), C.POINTER(C.c_uint32), C.c_uint64, _fp]
        L.crp_measure_gather.argtypes = [C.c_uint64, C.c_int, _dp]
        L.crp_set_device.argtypes = [C.c_int]
        L.crp_get_device_count.argtypes = [C.POINTER(C.c_int)]
        _lib = L
    return _lib


def check(rc):
    if rc != CRP_OK:
        msg = lib().crp_last_error().decode("utf-8", "replace")
        raise CrpError(msg or ("crpropa3_amd error %d" % rc))
    return rc


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Field:
    """A device-side magnetic field (owns its device buffers)."""

    def __init__(self, handle, keep=()):
        self.handle = handle
        self._keep = keep

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.crp_field_destroy(self.handle)
            self.handle = None

    @staticmethod
    def uniform(B):
        B = _f64(B, (3,))
        h = _vp()
        check(lib().crp_field_create_uniform(_p(B), C.byref(h)))
        return Field(h)

    @staticmethod
    def grid3f(values, origin, spacing, reflective=False, clip_volume=False):
        v = np.ascontiguousarray(values, dtype=np.float32)
        if v.ndim != 4 or v.shape[3] != 3:
            raise ValueError("grid values must have shape (Nx, Ny, Nz, 3)")
        o = _f64(origin, (3,))
        s = _f64(np.broadcast_to(np.asarray(spacing, dtype=np.float64), (3,)))
        h = _vp()
        check(lib().crp_field_create_grid3f(_p(v, _fp), v.shape[0], v.shape[1], v.shape[2], _p(o), _p(s),
                                            int(reflective), int(clip_volume), C.byref(h)))
        return Field(h)

    @staticmethod
    def modulated_grid3f(values, origin, spacing, modulation, morigin, mspacing, reflective=False,
                         clip_volume=False, mreflective=True, mclip_volume=False):
        """ModulatedMagneticFieldGrid: Grid3f values (Nx, Ny, Nz, 3) times a Grid1f (Mx, My, Mz)."""
        v = np.ascontiguousarray(values, dtype=np.float32)
        m = np.ascontiguousarray(modulation, dtype=np.float32)
        if v.ndim != 4 or v.shape[3] != 3 or m.ndim != 3:
            raise ValueError("grid values (Nx, Ny, Nz, 3) and modulation (Mx, My, Mz) expected")
        o, mo = _f64(origin, (3,)), _f64(morigin, (3,))
        s = _f64(np.broadcast_to(np.asarray(spacing, dtype=np.float64), (3,)))
        ms = _f64(np.broadcast_to(np.asarray(mspacing, dtype=np.float64), (3,)))
        h = _vp()
        check(lib().crp_field_create_modulated_grid3f(_p(v, _fp), v.shape[0], v.shape[1], v.shape[2], _p(o), _p(s),
                                                      int(reflective), int(clip_volume), _p(m, _fp), m.shape[0],
                                                      m.shape[1], m.shape[2], _p(mo), _p(ms), int(mreflective),
                                                      int(mclip_volume), C.byref(h)))
        return Field(h)

    @staticmethod
    def jf12(params=None, planck=False, striated=None, striated_origin=(0, 0, 0), striated_spacing=None,
             striated_reflective=False, turbulent=None, turbulent_origin=(0, 0, 0), turbulent_spacing=None,
             turbulent_reflective=False, turbulent_clip=False):
        """JF12Field (planck: PlanckJF12bField): params a JF12Params (default: the constructor's values
        with use_striated / use_turbulent set when the grid is given); striated a (Nx, Ny, Nz) float32
        Grid1f, turbulent a (Nx, Ny, Nz, 3) float32 Grid3f."""
        if params is None:
            params = jf12_default_params(planck)
            params.use_striated = int(striated is not None)
            params.use_turbulent = int(turbulent is not None)
        keep = []

        def grid(v, origin, spacing, comps):
            if v is None:
                return None, None, None, None
            v = np.ascontiguousarray(v, dtype=np.float32)
            if v.ndim != (3 if comps == 1 else 4) or (comps == 3 and v.shape[3] != 3):
                raise ValueError("grid of shape (Nx, Ny, Nz%s) expected" % ("" if comps == 1 else ", 3"))
            n = (C.c_int * 3)(*v.shape[:3])
            o = _f64(origin, (3,))
            sp = _f64(np.broadcast_to(np.asarray(spacing, dtype=np.float64), (3,)))
            keep.extend([v, n, o, sp])
            return _p(v, _fp), n, _p(o), _p(sp)

        sv, sn, so, ss = grid(striated, striated_origin, striated_spacing, 1)
        tv, tn, to, ts = grid(turbulent, turbulent_origin, turbulent_spacing, 3)
        h = _vp()
        check(lib().crp_field_create_jf12(C.byref(params), sv, sn, so, ss, int(striated_reflective), tv, tn, to, ts,
                                          int(turbulent_reflective), int(turbulent_clip), C.byref(h)))
        return Field(h)

    @staticmethod
    def grid3f_device(ptr, shape, origin, spacing, reflective=False, clip_volume=False, keep=None):
        """Borrow grid values already in device memory (e.g. a torch tensor's data_ptr())."""
        o = _f64(origin, (3,))
        s = _f64(np.broadcast_to(np.asarray(spacing, dtype=np.float64), (3,)))
        h = _vp()
        check(lib().crp_field_create_grid3f_device(_vp(ptr), shape[0], shape[1], shape[2], _p(o), _p(s),
                                                   int(reflective), int(clip_volume), C.byref(h)))
        return Field(h, keep=(keep,))

    @staticmethod
    def planewave(k, Ak, xi, kappa, beta):
        k, Ak, beta = _f64(k), _f64(Ak), _f64(beta)
        xi, kappa = _f64(xi, (-1, 3)), _f64(kappa, (-1, 3))
        h = _vp()
        check(lib().crp_field_create_planewave(len(k), _p(k), _p(Ak), _p(xi), _p(kappa), _p(beta), C.byref(h)))
        return Field(h)

    @staticmethod
    def periodic(field, extends, origin=(0, 0, 0), reflective=False):
        """PeriodicMagneticField(field, extends, origin, reflective); keeps `field` alive."""
        e, o = _f64(extends, (3,)), _f64(origin, (3,))
        h = _vp()
        check(lib().crp_field_create_periodic(field.handle, _p(e), _p(o), int(reflective), C.byref(h)))
        return Field(h, keep=(field,))

    @staticmethod
    def field_list(fields):
        """MagneticFieldList with addField(f) for each f in order; keeps the fields alive."""
        h = _vp()
        check(lib().crp_field_create_list(C.byref(h)))
        f = Field(h, keep=tuple(fields))
        for c in fields:
            check(lib().crp_field_list_add(h, c.handle))
        return f

    @staticmethod
    def evolution(field, m):
        """MagneticFieldEvolution(field, m); keeps `field` alive."""
        h = _vp()
        check(lib().crp_field_create_evolution(field.handle, float(m), C.byref(h)))
        return Field(h, keep=(field,))

    def eval_z(self, pos, z):
        """getField(position, z) with a redshift per position (FP64)."""
        pos = _f64(pos, (-1, 3))
        zz = _f64(np.broadcast_to(np.asarray(z, dtype=np.float64), (pos.shape[0],)))
        B = np.zeros_like(pos)
        check(lib().crp_field_eval_z(self.handle, _p(pos), _p(zz), pos.shape[0], _p(B)))
        return B

    LAYOUTS = {"points": 0, "cells": 1}

    def set_layout(self, layout):
        """Grid storage on the device: "points" (8 gathers) or "cells" (corner table)."""
        check(lib().crp_field_set_grid_layout(self.handle, self.LAYOUTS[layout]))
        return self

    def layout(self):
        v = C.c_int(-1)
        check(lib().crp_field_get_grid_layout(self.handle, C.byref(v)))
        return {0: "points", 1: "cells"}[v.value]

    STORAGE = {"fp32": 0, "fp16": 1}

    def set_storage(self, storage):
        """Grid storage read by mixed-precision runs: "fp32" or "fp16" (FP16 + per-brick scale)."""
        check(lib().crp_field_set_grid_storage(self.handle, self.STORAGE[storage]))
        return self

    INTERPOLATION = {"TRILINEAR": 0, "TRICUBIC": 1, "NEAREST_NEIGHBOUR": 2}
    GRIDS = {"values": 0, "modulation": 1}

    def set_interpolation(self, interpolation, grid="values"):
        """Grid::setInterpolationType of the Grid3f ("values") or the Grid1f ("modulation") of a grid or
        modulated field: "TRILINEAR", "TRICUBIC" or "NEAREST_NEIGHBOUR"."""
        check(lib().crp_field_set_grid_interpolation(self.handle, self.GRIDS[grid], self.INTERPOLATION[interpolation]))
        return self

    def interpolation(self, grid="values"):
        v = C.c_int(-1)
        check(lib().crp_field_get_grid_interpolation(self.handle, self.GRIDS[grid], C.byref(v)))
        return {0: "TRILINEAR", 1: "TRICUBIC", 2: "NEAREST_NEIGHBOUR"}[v.value]

    def eval(self, pos, precision=PRECISION_FP64):
        pos = _f64(pos, (-1, 3))
        B = np.zeros_like(pos)
        if precision == PRECISION_FP64:
            check(lib().crp_field_eval(self.handle, _p(pos), pos.shape[0], _p(B)))
        else:
            check(lib().crp_field_eval_ex(self.handle, _p(pos), pos.shape[0], _p(B), precision))
        return B


class TurbulenceDesc(C.Structure):
    """crp_turbulence_desc"""
    _fields_ = [("n", C.c_int), ("spacing", C.c_double), ("lmin", C.c_double), ("lmax", C.c_double),
                ("kind", C.c_int), ("alpha", C.c_double), ("lbendover", C.c_double), ("sindex", C.c_double),
                ("qindex", C.c_double)]


TURBULENCE_SIMPLE, TURBULENCE_GENERAL, TURBULENCE_GENERAL_POWERLAW = 0, 1, 2


def turbulence_desc(n, spacing, lmin, lmax, kind=TURBULENCE_SIMPLE, alpha=-11 / 3., lbendover=1.0, sindex=5 / 3.,
                    qindex=4.0):
    return TurbulenceDesc(n, spacing, lmin, lmax, kind, alpha, lbendover, sindex, qindex)


def turbulence_mode_count(desc):
    """Modes inside the turbulence range (the generator takes 4 draws per mode for the simple
    spectrum, 2 otherwise)."""
    m = C.c_uint64()
    check(lib().crp_turbulence_mode_count(C.byref(desc), C.byref(m)))
    return m.value


def turbulence_generate(desc, draws):
    """Unnormalised turbulent grid, (n, n, n, 3) float32, from MT19937 randInt() draws."""
    d = np.ascontiguousarray(draws, dtype=np.uint32)
    out = np.zeros((desc.n, desc.n, desc.n, 3), dtype=np.float32)
    check(lib().crp_turbulence_generate(C.byref(desc), d.ctypes.data_as(C.POINTER(C.c_uint32)), d.size,
                                         out.ctypes.data_as(_fp)))
    return out


MODULES = {"box": 1, "sphere": 2, "maxlen": 3, "observer": 4}  # enum crp_module
# SphericalBoundary's default margin, 0.1 kpc (include/crpropa/Units.h:109-114)
MODULE_SPHERE_DEFAULT_MARGIN = 0.1 * (1e3 * (648000 / math.pi * 149597870700.0))


def jf12_default_params(planck=False):
    """The JF12Field (PlanckJF12bField) constructor's constants (crp_jf12_default_params)."""
    p = JF12Params()
    check(lib().crp_jf12_default_params(int(planck), C.byref(p)))
    return p


class Propagator:
    """PropagationBP / PropagationCK / DiffusionSDE ("BP", "CK", "SDE") with a fused
    MaximumTrajectoryLength.  For "SDE": epsilon, alpha, scale, seed and advection (a uniform
    advection velocity, or None)."""

    def __init__(self, kind, field, tolerance, min_step, max_step, max_length=0.0, precision=PRECISION_FP64,
                 epsilon=0.1, alpha=1 / 3., scale=1.0, seed=0, advection=None):
        h = _vp()
        fh = field.handle if field is not None else None
        if kind == "SDE":
            check(lib().crp_propagator_create_sde(fh, tolerance, min_step, max_step, epsilon, C.byref(h)))
        else:
            fn = lib().crp_propagator_create_bp if kind == "BP" else lib().crp_propagator_create_ck
            check(fn(fh, tolerance, min_step, max_step, C.byref(h)))
        self.handle = h
        self.field = field
        self.kind = kind
        if max_length:
            check(lib().crp_propagator_set_max_trajectory_length(h, max_length))
        if precision != PRECISION_FP64:
            check(lib().crp_propagator_set_precision(h, precision))
        if kind == "SDE":
            self.set_sde(alpha, scale, epsilon, seed)
            if advection is not None:
                self.set_sde_advection(advection)

    def step(self, op, x, u, h, z, q, m):
        """The reference's single-step functions on n phase points: op "dY" / "tryStep" (BP; m = mass,
        h = step) or "dYdt" / "tryStep" (CK; m = energy, h = t = step / c).  Returns (out_x, out_u) or
        (out_x, out_u, err_x, err_u) as (n, 3) arrays."""
        x, u = _f64(x, (-1, 3)), _f64(u, (-1, 3))
        n = x.shape[0]
        col = [_f64(np.broadcast_to(np.asarray(a, dtype=np.float64), (n,))) for a in (h, z, q, m)]
        outs = [np.zeros((n, 3)) for _ in range(4)]
        L = lib()
        if self.kind == "BP" and op == "dY":
            check(L.crp_bp_dy(self.handle, n, _p(x), _p(u), *[_p(a) for a in col], _p(outs[0]), _p(outs[1])))
            return outs[0], outs[1]
        if self.kind == "CK" and op == "dYdt":
            check(L.crp_ck_dydt(self.handle, n, _p(x), _p(u), *[_p(a) for a in col[1:]], _p(outs[0]), _p(outs[1])))
            return outs[0], outs[1]
        fn = L.crp_bp_try_step if self.kind == "BP" else L.crp_ck_try_step
        check(fn(self.handle, n, _p(x), _p(u), *[_p(a) for a in col], *[_p(o) for o in outs]))
        return tuple(outs)

    def set_sde(self, alpha, scale, epsilon, seed):
        check(lib().crp_propagator_set_sde(self.handle, alpha, scale, epsilon, int(seed) & 0xFFFFFFFF))
        return self

    def set_sde_advection(self, v):
        """UniformAdvectionField(v) of the DiffusionSDE; None removes it; a dict {"kind": 2, "origin", "par":
        [vWind]} is a ConstantSphericalAdvectionField, {"kind": 3, "origin", "par": [radius, vMax, tau, alpha]}
        a SphericalAdvectionField; {"terms": [(kind, [parameters]), ...]} any list of the
        crp_sde_advection_kind fields, summed in order (AdvectionFieldList)."""
        if isinstance(v, dict) and "terms" in v:
            check(lib().crp_propagator_clear_sde_advection(self.handle))
            for kind, par in v["terms"]:
                p = _f64(par)
                check(lib().crp_propagator_add_sde_advection(self.handle, int(kind), _p(p), len(p)))
        elif isinstance(v, dict):
            o = _f64(v["origin"], (3,))
            fn = (lib().crp_propagator_set_sde_constant_spherical_advection if v["kind"] == 2
                  else lib().crp_propagator_set_sde_spherical_advection)
            check(fn(self.handle, _p(o), *[float(x) for x in v["par"]]))
        elif v is None:
            check(lib().crp_propagator_set_sde_advection(self.handle, None))
        else:
            a = _f64(v, (3,))
            check(lib().crp_propagator_set_sde_advection(self.handle, _p(a)))
        return self

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.crp_propagator_destroy(self.handle)
            self.handle = None

    def propagate(self, batch, max_steps=-1):
        st = RunStats()
        check(lib().crp_propagate(self.handle, batch.handle, max_steps, C.byref(st)))
        return st

    def set_observer(self, times, deactivate=False):
        """Fused Observer(ObserverTimeEvolution(times)) after the propagator and the length limit;
        times=() removes it.  Detections are read with Batch.detections(len(times))."""
        t = np.ascontiguousarray(times, dtype=np.float64)
        check(lib().crp_propagator_set_observer(self.handle, _p(t) if t.size else None, int(t.size),
                                                1 if deactivate else 0))
        return self

    def set_periodic_box(self, origin, size):
        """Fused PeriodicBox(origin, size) right after the propagator; size (0,0,0) removes it."""
        o = _f64(origin, (3,))
        z = _f64(size, (3,))
        check(lib().crp_propagator_set_periodic_box(self.handle, _p(o), _p(z)))
        return self

    def set_spherical_boundary(self, center, radius, margin=MODULE_SPHERE_DEFAULT_MARGIN, limit_step=True):
        """Fused SphericalBoundary(center, radius) after the box; center None removes it."""
        if center is None:
            check(lib().crp_propagator_set_spherical_boundary(self.handle, None, 0.0, 0.0, 0))
            return self
        c = _f64(center, (3,))
        check(lib().crp_propagator_set_spherical_boundary(self.handle, _p(c), radius, margin, 1 if limit_step else 0))
        return self

    def set_module_order(self, names):
        """Order of the fused modules after the propagator: names from "box", "sphere", "maxlen",
        "observer" (ModuleList order); [] restores box, sphere, maxlen, observer."""
        ids = [MODULES[nm] for nm in names]
        arr = (C.c_int * max(1, len(ids)))(*ids)
        check(lib().crp_propagator_set_module_order(self.handle, arr, len(ids)))
        return self

    def set_launch(self, block=256, blocks_per_cu=0, steps_per_launch=0, use_graph=False):
        cfg = LaunchConfig(block, blocks_per_cu, steps_per_launch, 1 if use_graph else 0, 0.0)
        check(lib().crp_propagator_set_launch(self.handle, C.byref(cfg)))
        return self

    def launch(self):
        cfg = LaunchConfig()
        check(lib().crp_propagator_get_launch(self.handle, C.byref(cfg)))
        return cfg.as_dict()

    def autotune(self, batch, probe_steps=32):
        """Time short runs over launch configurations on a copy of `batch`; keep and return the fastest."""
        cfg = LaunchConfig()
        check(lib().crp_propagator_autotune(self.handle, batch.handle, probe_steps, C.byref(cfg)))
        return cfg.as_dict()


class Batch:
    """Device-resident particle batch (structure of arrays)."""

    def __init__(self, capacity):
        h = _vp()
        check(lib().crp_batch_create(capacity, C.byref(h)))
        self.handle = h
        self.capacity = capacity

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.crp_batch_destroy(self.handle)
            self.handle = None

    def __len__(self):
        n = C.c_size_t(0)
        check(lib().crp_batch_size(self.handle, C.byref(n)))
        return n.value

    def upload(self, pos, dir, energy, charge, current_step=None, next_step=None, trajectory_length=None,
               time=None, active=None, detection_index=None, source_pos=None, created_pos=None, sde_step=None,
               sde_stream=None, redshift=None):
        pos = _f64(pos, (-1, 3))
        n = pos.shape[0]
        arrs = {
            "pos": pos, "dir": _f64(dir, (-1, 3)), "energy": _f64(energy), "charge": _f64(charge),
            "current_step": _f64(current_step if current_step is not None else np.zeros(n)),
            "next_step": _f64(next_step if next_step is not None else np.zeros(n)),
            "trajectory_length": _f64(trajectory_length if trajectory_length is not None else np.zeros(n)),
            "time": _f64(time if time is not None else np.zeros(n)),
        }
        act = np.ascontiguousarray(active if active is not None else np.ones(n), dtype=np.int32)
        s = HostState()
        for k, a in arrs.items():
            setattr(s, k, _p(a))
        s.active = _p(act, _ip)
        det = None
        if detection_index is not None:
            det = np.ascontiguousarray(detection_index, dtype=np.int32)
            s.detection_index = _p(det, _ip)
        extra = {}
        for name, a in (("source_pos", source_pos), ("created_pos", created_pos)):
            if a is not None:
                extra[name] = _f64(a, (-1, 3))
                setattr(s, name, _p(extra[name]))
        if sde_step is not None:
            extra["sde_step"] = np.ascontiguousarray(sde_step, dtype=np.uint32)
            s.sde_step = extra["sde_step"].ctypes.data_as(C.POINTER(C.c_uint32))
        if sde_stream is not None:
            extra["sde_stream"] = np.ascontiguousarray(sde_stream, dtype=np.uint64)
            s.sde_stream = extra["sde_stream"].ctypes.data_as(C.POINTER(C.c_uint64))
        if redshift is not None:
            extra["redshift"] = _f64(np.broadcast_to(np.asarray(redshift, dtype=np.float64), (n,)))
            s.redshift = _p(extra["redshift"])
        for a in extra.values():
            assert a.shape[0] == n
        check(lib().crp_batch_upload(self.handle, n, C.byref(s)))

    def download(self):
        n = len(self)
        out = {"pos": np.zeros((n, 3)), "dir": np.zeros((n, 3)), "energy": np.zeros(n), "charge": np.zeros(n),
               "current_step": np.zeros(n), "next_step": np.zeros(n), "trajectory_length": np.zeros(n),
               "time": np.zeros(n), "prev_pos": np.zeros((n, 3)), "prev_dir": np.zeros((n, 3)),
               "source_pos": np.zeros((n, 3)), "created_pos": np.zeros((n, 3))}
        act = np.zeros(n, dtype=np.int32)
        det = np.zeros(n, dtype=np.int32)
        sst = np.zeros(n, dtype=np.uint32)
        s = HostState()
        for k, a in out.items():
            setattr(s, k, _p(a))
        s.active = _p(act, _ip)
        s.detection_index = _p(det, _ip)
        s.sde_step = sst.ctypes.data_as(C.POINTER(C.c_uint32))
        check(lib().crp_batch_download(self.handle, C.byref(s)))
        out["active"] = act
        out["detection_index"] = det
        out["sde_step"] = sst
        return out

    def detections(self, ntimes):
        """(ntimes, n, 9) array: state at each observer detection (position, direction, trajectory
        length, time, nextStep); NaN for detections that did not happen."""
        n = len(self)
        out = np.full((ntimes, n, 9), np.nan)
        if n and ntimes:
            check(lib().crp_batch_detections(self.handle, ntimes, _p(out)))
        return out

    def copy_from(self, other):
        check(lib().crp_batch_copy(self.handle, other.handle))


def device_count():
    n = C.c_int(0)
    check(lib().crp_get_device_count(C.byref(n)))
    return n.value


def measure_gather(table_bytes, bytes_per_line=96):
    """Random line reads per second of this GPU (the cell-table evaluation's ceiling)."""
    r = C.c_double()
    check(lib().crp_measure_gather(int(table_bytes), int(bytes_per_line), C.byref(r)))
    return r.value
