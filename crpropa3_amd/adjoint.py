"""Differentiable fixed-step Boris propagation through a periodic Grid3f.

`BorisGridFunction` is a torch.autograd.Function whose forward and backward
are the HIP kernels behind crp_adjoint_forward / crp_adjoint_backward
(include/crpropa3_amd.h): the forward stores the state every K steps, the
backward recomputes each K-step segment and runs the hand-written reverse
step, scatter-adding into the grid gradient with FP64 atomics.

Model (per step h; PropagationBP::dY, src/module/PropagationBP.cpp:19-39, and
the two direction normalisations of :121 and src/ParticleState.cpp:30-32):
B = scale * trilinear(grid) in FP64.  Gradients are returned for the grid
values, the initial positions and directions, and the field scale.
Tensors live on the GPU; torch only supplies device memory and the stream.
"""
import ctypes as C

import torch

from . import capi

C_LIGHT = 2.99792458e8


class AdjointConfig(C.Structure):
    _fields_ = [("grid", C.c_void_p), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("origin", C.c_double * 3), ("spacing", C.c_double * 3), ("scale", C.c_double),
                ("step", C.c_double), ("nsteps", C.c_int), ("checkpoint_every", C.c_int)]


def _bind():
    L = capi.lib()
    if getattr(L, "_adjoint_bound", False):
        return L
    P = C.POINTER(AdjointConfig)
    vp = C.c_void_p
    L.crp_adjoint_workspace.argtypes = [P, C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    L.crp_adjoint_forward.argtypes = [P, C.c_size_t, vp, vp, vp, vp, vp, vp, vp]
    L.crp_adjoint_backward.argtypes = [P, C.c_size_t, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L._adjoint_bound = True
    return L


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _config(grid, origin, spacing, scale, step, nsteps, K):
    cfg = AdjointConfig()
    cfg.grid = grid.data_ptr()
    cfg.nx, cfg.ny, cfg.nz = grid.shape[:3]
    for a in range(3):
        cfg.origin[a] = float(origin[a])
        cfg.spacing[a] = float(spacing[a])
    cfg.scale = float(scale)
    cfg.step = float(step)
    cfg.nsteps = int(nsteps)
    cfg.checkpoint_every = int(K)
    return cfg


class BorisGridFunction(torch.autograd.Function):
    """(grid, x0, u0, scale) -> (x_final, u_final); see the module docstring."""

    @staticmethod
    def forward(ctx, grid, x0, u0, scale, kq, origin, spacing, step, nsteps, K):
        if not (grid.is_cuda and x0.is_cuda and u0.is_cuda and kq.is_cuda):
            raise ValueError("BorisGridFunction: tensors must be on the GPU")
        grid = grid.detach().contiguous()
        if grid.dtype != torch.float32 or grid.dim() != 4 or grid.shape[3] != 3:
            raise ValueError("grid must be float32 of shape (nx, ny, nz, 3)")
        x0c = x0.detach().to(torch.float64).contiguous()
        u0c = u0.detach().to(torch.float64).contiguous()
        kqc = kq.detach().to(torch.float64).contiguous()
        n = x0c.shape[0]
        L = _bind()
        cfg = _config(grid, origin, spacing, float(scale), step, nsteps, K)
        cb, sb = C.c_size_t(0), C.c_size_t(0)
        capi.check(L.crp_adjoint_workspace(C.byref(cfg), n, C.byref(cb), C.byref(sb)))
        ckpt = torch.empty(max(cb.value // 8, 1), dtype=torch.float64, device=grid.device)
        xo = torch.empty_like(x0c)
        uo = torch.empty_like(u0c)
        stream = C.c_void_p(torch.cuda.current_stream(grid.device).cuda_stream)
        capi.check(L.crp_adjoint_forward(C.byref(cfg), n, _ptr(x0c), _ptr(u0c), _ptr(kqc), _ptr(ckpt), _ptr(xo),
                                         _ptr(uo), stream))
        ctx.save_for_backward(grid, kqc, ckpt)
        ctx.meta = (origin, spacing, float(scale), step, nsteps, K, sb.value, grid.dtype)
        ctx.scale_is_tensor = torch.is_tensor(scale)
        return xo, uo

    @staticmethod
    def backward(ctx, gx, gu):
        grid, kqc, ckpt = ctx.saved_tensors
        origin, spacing, scale, step, nsteps, K, sbytes, gdtype = ctx.meta
        n = kqc.shape[0]
        dev = grid.device
        L = _bind()
        cfg = _config(grid, origin, spacing, scale, step, nsteps, K)
        gu = torch.zeros((n, 3), dtype=torch.float64, device=dev) if gu is None else gu.to(torch.float64).contiguous()
        gx = None if gx is None else gx.to(torch.float64).contiguous()
        scratch = torch.empty(max(sbytes // 8, 1), dtype=torch.float64, device=dev)
        ggrid = torch.zeros(grid.numel(), dtype=torch.float64, device=dev)
        gscale = torch.zeros(1, dtype=torch.float64, device=dev)
        gx0 = torch.empty((n, 3), dtype=torch.float64, device=dev)
        gu0 = torch.empty((n, 3), dtype=torch.float64, device=dev)
        stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        capi.check(L.crp_adjoint_backward(C.byref(cfg), n, _ptr(kqc), _ptr(ckpt), _ptr(gx), _ptr(gu), _ptr(scratch),
                                          _ptr(ggrid), _ptr(gx0), _ptr(gu0), _ptr(gscale), stream))
        gs = gscale.view(()) if ctx.scale_is_tensor else None
        return (ggrid.view(grid.shape).to(gdtype), gx0, gu0, gs, None, None, None, None, None, None)


def boris_grid_propagate(grid, x0, u0, energy, charge, origin, spacing, step, nsteps, checkpoint_every=64,
                         scale=1.0):
    """Differentiable fixed-step Boris push of particles through `scale * grid`.

    grid: (nx, ny, nz, 3) float32 CUDA tensor (periodic Grid3f values); x0, u0: (n, 3)
    float64 CUDA tensors (positions [m], unit directions); energy [J], charge [C]: (n,)
    tensors or floats.  `scale` may be a float or a 0-d tensor requiring grad.
    Returns (x_final, u_final)."""
    dev = grid.device
    E = torch.as_tensor(energy, dtype=torch.float64, device=dev).expand(x0.shape[0])
    q = torch.as_tensor(charge, dtype=torch.float64, device=dev).expand(x0.shape[0])
    kq = (q * C_LIGHT * step / (2 * E)).contiguous()
    if not torch.is_tensor(scale):
        scale = torch.tensor(float(scale), dtype=torch.float64, device=dev)
    return BorisGridFunction.apply(grid, x0, u0, scale, kq, origin, spacing, step, nsteps, checkpoint_every)
