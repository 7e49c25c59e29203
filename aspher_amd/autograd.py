"""torch.autograd wrapper of the certified contact solve (no counterpart in
the reference, which has no derivative of its solve).

``contact_pressure(gap, p_nominal, ...)`` returns the converged pressure
field as a float64 tensor; its backward runs the library's adjoint (the
transposed KKT system of the contact QP on the converged support, solved by
preconditioned CG on the GPU with the forward solve's own operator): for
dL/dp = w it gives dL/dgap = -mu on the support (0 elsewhere) and
dL/dp_nominal = nu Ns^2, where [S_CC 1; 1' 0][mu; nu] = [w_C; 0].

Tensors may live on the CPU or on a torch device; they are handed to the
library as host arrays (torch is only the autograd front end here). The
gradient is exact for a strictly complementary solution: points exactly at
the contact threshold have a one-sided derivative.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .api import ContactSolver, solve_nested


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
        g = gap.detach().to("cpu", torch.float64).contiguous().numpy().ravel()
        pn = float(p_nominal.detach().cpu()) if torch.is_tensor(p_nominal) else float(p_nominal)
        res = _solve(cfg, g, pn)
        if not res.converged:
            raise RuntimeError("contact solve did not converge: %s (%s)"
                               % (res.status, res.status_reason))
        ctx.res = res
        ctx.cfg = cfg
        ctx.shape = gap.shape
        ctx.pn_tensor = torch.is_tensor(p_nominal)
        p = torch.from_numpy(res.pressure.reshape(-1).copy()).reshape(gap.shape)
        return p.to(device=gap.device)

    @staticmethod
    def backward(ctx, grad_p):
        cfg = ctx.cfg
        w = grad_p.detach().to("cpu", torch.float64).contiguous().numpy().ravel()
        gg, gpn, info = _solver(cfg).adjoint(ctx.res, w, tol=cfg.adjoint_tol,
                                             max_iter=cfg.adjoint_max_iter,
                                             precond=cfg.adjoint_precond)
        if not info["converged"]:
            raise RuntimeError("adjoint solve did not converge: %s (residual %g)"
                               % (info["status"], info["rel_residual"]))
        ctx.adjoint_info = info
        g_gap = torch.from_numpy(np.ascontiguousarray(gg).reshape(-1)).reshape(ctx.shape)
        g_gap = g_gap.to(device=grad_p.device)
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
