"""torch.autograd wrapper of the certified contact solve (no counterpart in
the reference, which has no derivative of its solve).

``contact_pressure(gap, p_nominal, ...)`` returns the converged pressure
field as a float64 tensor; its backward runs the library's adjoint (the
transposed KKT system of the contact QP on the converged support, solved by
preconditioned CG on the GPU with the forward solve's own operator): for
dL/dp = w it gives dL/dgap = -mu on the support (0 elsewhere) and
dL/dp_nominal = nu Ns^2, where [S_CC 1; 1' 0][mu; nu] = [w_C; 0].

Tensors may live on the CPU or on a torch device. Device tensors go to the
library as device pointers (the gap, the pressure, dL/dp and dL/dgap stay in
HBM) when torch and the library run on one HIP runtime, i.e. one
libamdhip64 is mapped in the process; otherwise, and for CPU tensors, they
are handed over as host arrays (torch is only the autograd front end here).
The gradient is exact for a strictly complementary solution: points exactly
at the contact threshold have a one-sided derivative.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from ._capi import check, lib
from .api import ContactSolver, solve_nested, solve_nested_device


@dataclass(frozen=True)
class ContactConfig:
    grid_size: int
    backend: str = "fft"              # "fft" | "periodic" | "dense"
    domain_size: float = 1.0
    E_star: float = 1.0
    tol: float = 1e-8
    max_iter: int = 20000
    nested: bool = True
    coarsest: int = 64
    precision: str = "double"         # nested: double | float_then_double | float
    precond: str = "fourier"          # single level: none | fourier | fourier-fft
    adjoint_tol: float = 1e-10
    adjoint_max_iter: int = 5000
    adjoint_precond: str = "auto"


def _solve(cfg: ContactConfig, gap: np.ndarray, p_nominal: float, p_init=None):
    if cfg.nested and cfg.backend != "dense":
        return solve_nested(cfg.grid_size, gap, p_nominal, domain_size=cfg.domain_size,
                            E_star=cfg.E_star, coarsest=min(cfg.coarsest, cfg.grid_size),
                            tol=cfg.tol, max_iter=cfg.max_iter, backend=cfg.backend,
                            precision=cfg.precision, reuse_operators=True)
    s = _solver(cfg)
    return s.solve(gap, p_nominal, tol=cfg.tol, max_iter=cfg.max_iter, precond=cfg.precond,
                   p_init=p_init)


def _solve_device(cfg: ContactConfig, gap_ptr: int, p_nominal: float):
    if cfg.nested and cfg.backend != "dense":
        return solve_nested_device(cfg.grid_size, gap_ptr, p_nominal,
                                   domain_size=cfg.domain_size, E_star=cfg.E_star,
                                   coarsest=min(cfg.coarsest, cfg.grid_size), tol=cfg.tol,
                                   max_iter=cfg.max_iter, backend=cfg.backend,
                                   precision=cfg.precision, reuse_operators=True)
    return _solver(cfg).solve_device(gap_ptr, p_nominal, tol=cfg.tol, max_iter=cfg.max_iter,
                                     precond=cfg.precond)


def _one_hip_runtime() -> bool:
    """True when a single libamdhip64 is mapped in this process: torch and the
    library then share one HIP runtime, and device pointers pass between them."""
    try:
        paths = set()
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    paths.add(line.split()[-1])
        return len(paths) == 1
    except OSError:
        return False


def _on_device(t, n: int) -> bool:
    # the library's device copy moves multiples of 16 bytes
    return t.is_cuda and (8 * n) % 16 == 0 and _one_hip_runtime()


_SOLVERS: dict = {}


def _solver(cfg: ContactConfig) -> ContactSolver:
    """One solver handle per (grid, backend, L, E*), on the cached operator
    shared with the nested driver."""
    key = (cfg.grid_size, cfg.backend, cfg.domain_size, cfg.E_star)
    s = _SOLVERS.get(key)
    if s is None:
        s = ContactSolver(cfg.grid_size, domain_size=cfg.domain_size, E_star=cfg.E_star,
                          backend=cfg.backend, shared_operator=True)
        _SOLVERS[key] = s
    return s


class ContactPressure(torch.autograd.Function):
    """pressure = solve(gap, p_nominal); backward by the GPU adjoint."""

    @staticmethod
    def forward(ctx, gap, p_nominal, cfg: ContactConfig):
        N = cfg.grid_size * cfg.grid_size
        if gap.numel() != N:
            raise ValueError("gap has %d entries, expected %d" % (gap.numel(), N))
        pn = float(p_nominal.detach().cpu()) if torch.is_tensor(p_nominal) else float(p_nominal)
        ctx.device = _on_device(gap, N)
        if ctx.device:
            # the gap stays in HBM: the library reads it and the pressure is
            # copied device to device into the output tensor
            with torch.cuda.device(gap.device):
                g = gap.detach().to(torch.float64).contiguous().reshape(-1)
                torch.cuda.synchronize()
                res = _solve_device(cfg, g.data_ptr(), pn)
        else:
            g = gap.detach().to("cpu", torch.float64).contiguous().numpy().ravel()
            res = _solve(cfg, g, pn)
        if not res.converged:
            raise RuntimeError("contact solve did not converge: %s (%s)"
                               % (res.status, res.status_reason))
        ctx.res = res
        ctx.cfg = cfg
        ctx.shape = gap.shape
        ctx.pn_tensor = torch.is_tensor(p_nominal)
        if ctx.device:
            with torch.cuda.device(gap.device):
                p = torch.empty(N, dtype=torch.float64, device=gap.device)
                check(lib().aspher_copy_device(p.data_ptr(), res.field_device_ptr(0), 8 * N))
            return p.reshape(gap.shape)
        p = torch.from_numpy(res.pressure.reshape(-1).copy()).reshape(gap.shape)
        return p.to(device=gap.device)

    @staticmethod
    def backward(ctx, grad_p):
        cfg = ctx.cfg
        N = cfg.grid_size * cfg.grid_size
        if ctx.device and grad_p.is_cuda:
            with torch.cuda.device(grad_p.device):
                w = grad_p.detach().to(torch.float64).contiguous().reshape(-1)
                g_gap = torch.empty(N, dtype=torch.float64, device=grad_p.device)
                torch.cuda.synchronize()
                gpn, info = _solver(cfg).adjoint_device(
                    ctx.res, w.data_ptr(), g_gap.data_ptr(), tol=cfg.adjoint_tol,
                    max_iter=cfg.adjoint_max_iter, precond=cfg.adjoint_precond)
            g_gap = g_gap.reshape(ctx.shape)
        else:
            w = grad_p.detach().to("cpu", torch.float64).contiguous().numpy().ravel()
            gg, gpn, info = _solver(cfg).adjoint(ctx.res, w, tol=cfg.adjoint_tol,
                                                 max_iter=cfg.adjoint_max_iter,
                                                 precond=cfg.adjoint_precond)
            g_gap = torch.from_numpy(np.ascontiguousarray(gg).reshape(-1)).reshape(ctx.shape)
            g_gap = g_gap.to(device=grad_p.device)
        if not info["converged"]:
            raise RuntimeError("adjoint solve did not converge: %s (residual %g)"
                               % (info["status"], info["rel_residual"]))
        ctx.adjoint_info = info
        g_pn = None
        if ctx.pn_tensor and ctx.needs_input_grad[1]:
            g_pn = torch.tensor(gpn, dtype=torch.float64, device=grad_p.device)
        return g_gap, g_pn, None


def contact_pressure(gap, p_nominal, cfg: ContactConfig = None, **kw):
    """Differentiable certified contact pressure. ``gap``: (Ns, Ns) or (N,)
    float tensor; ``p_nominal``: float or 0-dim tensor (differentiable)."""
    if cfg is None:
        Ns = kw.pop("grid_size", None)
        if Ns is None:
            n = gap.numel()
            Ns = int(round(n ** 0.5))
        cfg = ContactConfig(grid_size=int(Ns), **kw)
    elif kw:
        raise TypeError("pass either cfg or keyword options, not both")
    return ContactPressure.apply(gap, p_nominal, cfg)
