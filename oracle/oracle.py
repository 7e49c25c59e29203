"""ORACLE — test infrastructure only.

CPU restatement (numpy / scipy.fft, plus the C Love kernel in ``love.c``) of
the reference's spectral contact path. Only ``tests/``, ``__graft_entry__.smoke``
and the ``cpu_baseline`` leg of ``bench.py`` may import this module, and only as
the checker or as the timed CPU baseline; the product path (``aspher_amd``)
never imports, calls or links it.

Every function cites the reference file:line it restates. Transforms use
scipy.fft (pocketfft, the same FFT library the reference vendors in
third_party/pocketfft); the reference's engine is unnormalised
(fwd∘inv = n², src/fft_engine.hpp:18-22), so ``_inv`` multiplies scipy's
normalised inverse back by n².

Pinning (what checks each function, see tests/test_oracle_*.py):
  love_uz / love_table     80-digit fixture tests/golden/love_reference.txt,
                           self-term closed form, parity, homogeneity
  FreeSpaceOperator        brute-force dense convolution (rel <1e-12),
                           K̂ realness, symmetry
  PeriodicOperator         cos-mode eigenvalues 2/(E*|q|), DC -> 0,
                           symmetric/PSD/circulant on odd grids
  StencilPreconditioner    universal-constant tap table of
                           tests/test_stencil.cpp:21-33 (±5e-6), symbol > 0
  FourierPreconditioner    offset invariance and symmetry (<1e-12)
  project_load_simplex     sorting projector, idempotence
  solve_contact            manufactured-KKT exact p*, support enumeration,
                           Westergaard closed form, FW bound identities
  solve_nested             nested == single level, datum invariance
  solve_contact_active     restricted certificate is not global (paraboloid,
                           half-domain candidates: tests/test_active.cpp:513-561),
                           stays zero outside C, equals the full solve on the
                           gap lifted outside C, equals the full solve when C
                           holds the contact set
  dilate_mask              scipy.ndimage.maximum_filter (square, clipped)
  active_level             active nested == standard nested, fallback and
                           expansion regressions, all-levels / finest-only
                           stage names, occupancy gate, third-exit fallback
                           (tests/test_active.cpp:309-511)
  kkt_gradient             central finite differences through the exact
                           support-enumeration solver and solve_contact
  circulant_matrix         equals the operator's matvec on random vectors
  splitmix64               the published known-answer vector (seed 1234567)
  cerruti_uxx / _uxy       Gauss-Legendre quadrature (tests/test_cerruti.cpp:53-78),
                           nu = 0 -> love_uz, parity, closed-form self term
  CerrutiTables            accessor identities, dense block symmetric + PD,
                           continuum symbol eigenvalues
  TangentialOperator       equals the dense block at Ns in {8, 12, 16}
  TangentialPreconditioner symmetry, mask, mean removal
  rough_surface_hash       band limit, mean/rms, Hermitian output, equality
                           with the numpy-phase recipe on the same phases
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np
import scipy.fft as sfft

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None


def build_lib(force: bool = False) -> str:
    """Compile love.c into liboracle.so (gcc, -O2, no fast-math)."""
    src = os.path.join(_HERE, "love.c")
    if force or not os.path.exists(_SO) or os.path.getmtime(_SO) < os.path.getmtime(src):
        tmp = _SO + ".tmp%d" % os.getpid()
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11",
                               "-fno-fast-math", src, "-o", tmp, "-lm"])
        os.replace(tmp, _SO)
    return _SO


def lib():
    global _lib
    if _lib is None:
        path = _SO if os.path.exists(_SO) else build_lib()
        L = ctypes.CDLL(path)
        d = ctypes.c_double
        pd = ctypes.POINTER(ctypes.c_double)
        L.oracle_love_uz.restype = d
        L.oracle_love_uz.argtypes = [d, d, d, d]
        L.oracle_asinh_diff.restype = d
        L.oracle_asinh_diff.argtypes = [d, d, d]
        L.oracle_love_many.restype = None
        L.oracle_love_many.argtypes = [ctypes.c_int64, pd, pd, pd, pd, pd]
        L.oracle_love_table.restype = None
        L.oracle_love_table.argtypes = [ctypes.c_int, d, d, pd]
        L.oracle_cerruti_uxx.restype = d
        L.oracle_cerruti_uxx.argtypes = [d, d, d, d, d]
        L.oracle_cerruti_uxy.restype = d
        L.oracle_cerruti_uxy.argtypes = [d, d, d, d]
        L.oracle_cerruti_many.argtypes = [ctypes.c_int64, pd, pd, pd, pd, d, pd, pd]
        L.oracle_cerruti_tables.argtypes = [ctypes.c_int, d, d, d, pd, pd]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ── L1: Love kernel (src/boussinesq_kernel.cpp:44-106) ──────────────────────

def love_uz(x, y, a, b):
    """Vectorised love_uz (src/boussinesq_kernel.cpp:71-95)."""
    x, y, a, b = np.broadcast_arrays(*(np.asarray(v, dtype=np.float64) for v in (x, y, a, b)))
    x = np.ascontiguousarray(x).ravel(); y = np.ascontiguousarray(y).ravel()
    a = np.ascontiguousarray(a).ravel(); b = np.ascontiguousarray(b).ravel()
    out = np.empty_like(x)
    lib().oracle_love_many(x.size, _ptr(x), _ptr(y), _ptr(a), _ptr(b), _ptr(out))
    return out


def love_table(Ns: int, L: float = 1.0, E_star: float = 1.0) -> np.ndarray:
    """(Ns, Ns) table T[dy, dx] (src/boussinesq_kernel.cpp:97-106)."""
    T = np.empty((Ns, Ns), dtype=np.float64)
    lib().oracle_love_table(int(Ns), float(L), float(E_star), _ptr(T))
    return T


def dense_matrix(Ns: int, L: float = 1.0, E_star: float = 1.0) -> np.ndarray:
    """N x N dense influence matrix (src/boussinesq_kernel.cpp:108-115,
    entry() at include/boussinesq_kernel.hpp:37-41)."""
    T = love_table(Ns, L, E_star)
    ix = np.arange(Ns * Ns) % Ns
    iy = np.arange(Ns * Ns) // Ns
    return T[np.abs(iy[:, None] - iy[None, :]), np.abs(ix[:, None] - ix[None, :])]


# ── transforms (src/fft_engine.hpp: unnormalised r2c / c2r) ─────────────────

# scipy.fft worker threads (1 = the serial restatement; the CPU-baseline leg
# of bench.py raises it to the host's core count)
WORKERS = 1


def _fwd(G):
    return sfft.rfft2(G, workers=WORKERS)


def _inv(C, n):
    return sfft.irfft2(C, s=(n, n), workers=WORKERS) * (n * n)


# ── L3: free-space operator (src/fft_operator.cpp:18-163) ───────────────────

class FreeSpaceOperator:
    """Hockney zero-padded (2Ns)^2 circular convolution with the Love table."""

    def __init__(self, Ns: int, L: float = 1.0, E_star: float = 1.0, table=None):
        self.Ns, self.M = Ns, 2 * Ns
        self.table = love_table(Ns, L, E_star) if table is None else table
        self.build()

    def embed(self) -> np.ndarray:
        """Wrap-around embedding (src/fft_operator.cpp:30-44): offset d at
        index d (d>=0) or M+d (d<0); row/column Ns zero."""
        Ns, M = self.Ns, self.M
        idx = np.arange(M)
        ad = np.where(idx < Ns, idx, M - idx)
        K = self.table[np.minimum(ad, Ns - 1)[:, None], np.minimum(ad, Ns - 1)[None, :]].copy()
        K[Ns, :] = 0.0
        K[:, Ns] = 0.0
        return K

    def build(self):
        """FFT of the embedding, realness check, K̂ = Re/M^2
        (src/fft_operator.cpp:46-68)."""
        M = self.M
        Kc = _fwd(self.embed())                    # (M, Ns+1) [ky, kx]
        max_im = np.abs(Kc.imag).max()
        max_re = np.abs(Kc.real).max()
        if max_im > 1e-10 * max_re:
            raise RuntimeError("FFTOperator: kernel spectrum has a non-negligible "
                               "imaginary part; wrap-around embedding is broken")
        self.Kh = Kc.real * (1.0 / (float(M) * float(M)))
        self.Kh_f = self.Kh.astype(np.float32)     # src/fft_operator.cpp:71-78

    def matvec(self, x, dtype=np.float64):
        """scatter -> fwd -> xK̂ -> inv -> gather (src/fft_operator.cpp:87-138)."""
        Ns, M = self.Ns, self.M
        G = np.zeros((M, M), dtype=dtype)
        G[:Ns, :Ns] = np.asarray(x, dtype=dtype).reshape(Ns, Ns)
        C = _fwd(G)
        C *= (self.Kh if dtype == np.float64 else self.Kh_f)
        return _inv(C, M)[:Ns, :Ns].astype(dtype).ravel()


# ── L3: periodic operator (src/periodic_fft_operator.cpp:11-177) ────────────

class PeriodicOperator:
    """u_hat = scale/|m| p_hat, scale = L/(pi E* Ns^2), zero mode -> 0."""

    def __init__(self, Ns: int, L: float = 1.0, E_star: float = 1.0):
        if Ns < 2:
            raise ValueError("grid_size must be at least 2")
        self.Ns = Ns
        self.scale = L / (math.pi * E_star * float(Ns) * Ns)
        nh = Ns // 2 + 1
        ky = np.arange(Ns)
        my = np.where(ky < (Ns + 1) // 2, ky, ky - Ns).astype(np.float64)
        kx = np.arange(nh).astype(np.float64)
        m2 = kx[None, :] ** 2 + (my ** 2)[:, None]
        with np.errstate(divide="ignore"):
            W = np.where(m2 > 0, self.scale / np.sqrt(m2), 0.0)
        self.W = W                                   # symbol in double, cast per precision
        self.W_f = W.astype(np.float32)

    def matvec(self, x, dtype=np.float64):
        Ns = self.Ns
        C = _fwd(np.asarray(x, dtype=dtype).reshape(Ns, Ns))
        C *= (self.W if dtype == np.float64 else self.W_f)
        return _inv(C, Ns).astype(dtype).ravel()


class DenseOperator:
    """Dense matvec from the Love table (the reference's 'dense' backend)."""

    def __init__(self, Ns: int, L: float = 1.0, E_star: float = 1.0):
        self.Ns = Ns
        self.S = dense_matrix(Ns, L, E_star)

    def matvec(self, x, dtype=np.float64):
        if dtype == np.float64:
            return self.S @ np.asarray(x, dtype=np.float64)
        return (self.S @ np.asarray(x, dtype=np.float64)).astype(dtype)


# ── L4: preconditioners ─────────────────────────────────────────────────────

def stencil_kernel_grid(Nw: int) -> np.ndarray:
    """w = F^-1 |k| (DC 0) normalised by w(0,0) (src/stencil_precond.cpp:19-43)."""
    nh = Nw // 2 + 1
    ky = np.arange(Nw)
    kof = np.where(ky < Nw // 2, ky, ky - Nw).astype(np.float64)
    kx = np.arange(nh).astype(np.float64)
    C = np.hypot(kx[None, :], kof[:, None]).astype(np.complex128)
    C[0, 0] = 0.0
    G = _inv(C, Nw)
    return G / G[0, 0]                               # [iy, ix]


class StencilPreconditioner:
    """13-tap (R=2) real-space |k| stencil (src/stencil_precond.cpp:53-109,199-220)."""

    def __init__(self, Ns: int, radius: int = 2):
        if radius < 1:
            raise ValueError("stencil radius must be >= 1")
        if 2 * radius + 1 > Ns:
            raise ValueError("stencil radius too large for the grid")
        self.Ns, self.R = Ns, radius
        Nw = Ns if Ns < 512 else 512
        w = stencil_kernel_grid(Nw)
        self.dx, self.dy, self.w = [], [], []
        for dy in range(-radius, radius + 1):
            for dx in range(-radius, radius + 1):
                if dx * dx + dy * dy > radius * radius:
                    continue
                self.dx.append(dx)
                self.dy.append(dy)
                self.w.append(float(w[dy % Nw, dx % Nw]))

    def apply(self, g, contact, dtype=None):
        """apply_full (src/stencil_precond.cpp:63-109)."""
        Ns = self.Ns
        dtype = np.dtype(g.dtype if dtype is None else dtype).type
        c = np.asarray(contact).astype(bool).reshape(Ns, Ns)
        gd = g.astype(np.float64).reshape(Ns, Ns)
        ng = int(c.sum())
        gmean = gd[c].sum() / ng if ng else 0.0
        src = np.where(c, gd - gmean, 0.0)
        acc = np.zeros((Ns, Ns))
        for dx, dy, w in zip(self.dx, self.dy, self.w):
            # value at (ix+dx, iy+dy) with periodic wrap
            acc += w * np.roll(np.roll(src, -dy, axis=0), -dx, axis=1)
        z = np.where(c, acc, 0.0).astype(dtype)
        nc = ng
        if nc:
            zsum = acc[c].sum()
            zmean = dtype(zsum / nc)
            z[c] -= zmean
        return z.ravel()


class FourierPreconditioner:
    """Full-grid |k| preconditioner, symbol in float (src/fourier_precond.cpp:12-116)."""

    def __init__(self, Ns: int):
        self.Ns = Ns
        nh = Ns // 2 + 1
        norm = np.float32(1.0) / (np.float32(Ns) * np.float32(Ns))
        ky = np.arange(Ns)
        kof = np.where(ky < Ns // 2, ky, ky - Ns).astype(np.float32)
        kx = np.arange(nh).astype(np.float32)
        wh = norm * np.hypot(kx[None, :], kof[:, None]).astype(np.float32)
        wh[0, 0] = 0.0
        self.wh = wh.astype(np.float32)              # [ky, kx]

    def apply(self, g, contact, dtype=None):
        Ns = self.Ns
        dtype = np.dtype(g.dtype if dtype is None else dtype).type
        c = np.asarray(contact).astype(bool).ravel()
        ng = int(c.sum())
        gsum = g[c].astype(np.float64).sum()
        gmean = dtype(gsum / ng) if ng else dtype(0)
        G = np.where(c, g - gmean, dtype(0)).astype(dtype).reshape(Ns, Ns)
        C = _fwd(G)
        C *= self.wh.astype(dtype)
        Gi = _inv(C, Ns).astype(dtype).ravel()
        z = np.where(c, Gi, dtype(0)).astype(dtype)
        nc = ng
        if nc:
            zsum = Gi[c].astype(np.float64).sum()
            zmean = dtype(zsum / nc)
            z[c] -= zmean
        return z


def stencil_for_level(engine: str, precond: bool, prev_occupancy: float,
                      occupancy_max: float) -> bool:
    """src/nested_solve.cpp:27-39."""
    if not precond:
        return False
    if engine == "stencil":
        return True
    if engine == "fft":
        return False
    return prev_occupancy < 0.0 or prev_occupancy < occupancy_max


# ── L5: load-simplex projection (src/contact_solver.cpp:31-60) ──────────────

def project_load_simplex(y: np.ndarray, P_total: float) -> float:
    """Michelot threshold iteration, in place; returns theta."""
    N = y.size
    if N == 0:
        return 0.0
    yd = y.astype(np.float64)
    ssum = yd.sum()
    nact = N
    theta = (ssum - P_total) / nact
    for _ in range(64):
        m = yd > theta
        n2 = int(m.sum())
        if n2 == 0 or n2 == nact:
            break
        nact = n2
        ssum = yd[m].sum()
        theta = (ssum - P_total) / nact
    th = y.dtype.type(theta)
    y[:] = np.where(y > th, y - th, y.dtype.type(0))
    return theta


# ── L5: certified Polonsky–Keer CG (src/contact_solver.cpp:150-587) ─────────

STATUSES = ("converged", "stagnated", "max_iterations", "nonpositive_curvature",
            "nonfinite", "verification_failed", "resource_limit")


@dataclass
class SolveOptions:
    """include/contact_solver.hpp:151-191 (+ SolveScales :134-147)."""
    tol: float = 1e-8
    requested_tol: float = 0.0
    max_iter: int = 5000
    use_pr: bool = True
    light: bool = False
    record_history: bool = False
    keep_best: bool = True
    load_tol: float = 0.0
    stall_limit: int = 200
    allow_tolerance_relaxation: bool = False
    p_ref: float = 0.0
    g_ref: float = 0.0
    datum_mode: str = "automatic"          # automatic | solver | caller
    datum: float = 0.0


@dataclass
class Result:
    pressure: np.ndarray = None
    displacement: Optional[np.ndarray] = None
    gap: Optional[np.ndarray] = None
    approach: float = 0.0
    objective: float = 0.0
    error: float = 0.0
    iterations: int = 0
    converged: bool = False
    contact_fraction: float = 0.0
    mean_pressure: float = 0.0
    error_history: List[float] = field(default_factory=list)
    status: str = "max_iterations"
    status_reason: str = ""
    pk_error: float = 0.0
    fw_gap: float = 0.0
    fw_error: float = 0.0
    load_error: float = 0.0
    pressure_violation: float = 0.0
    penetration_error: float = 0.0
    requested_tol: float = 0.0
    effective_tol: float = 0.0
    p_ref: float = 0.0
    g_ref: float = 0.0
    matvec_count: int = 0
    verification_matvec_count: int = 0
    active_rounds: int = 0
    active_fallback: bool = False
    precond_count: int = 0
    identification_steps: int = 0
    returned_best: bool = False
    time_total: float = 0.0
    stage_stats: list = field(default_factory=list)


class _Diag:
    __slots__ = ("alpha", "gmin", "pk", "G", "fw", "pen", "load", "psum", "pv", "nc", "finite")


def solve_contact(S: Callable, g0: np.ndarray, p_bar, opt: SolveOptions = None,
                  precond: Optional[Callable] = None, p_init=None,
                  dtype=np.float64) -> Result:
    """Restatement of solve_contact_impl<Real> (src/contact_solver.cpp:150-587).

    S(x) -> S x in `dtype`; precond(g, contact_bool) -> z in `dtype`.
    All reductions accumulate in float64 like the reference."""
    opt = opt or SolveOptions()
    R = np.dtype(dtype).type
    is_double = (np.dtype(dtype) == np.float64)
    g0 = np.ascontiguousarray(g0, dtype=dtype).ravel()
    N = g0.size
    p_bar = R(p_bar)
    if N == 0 or not (p_bar > 0):
        raise ValueError("solve_contact: empty gap or p_bar <= 0")
    if not (opt.tol > 0.0):
        raise ValueError("solve_contact: tol must be positive")
    if opt.max_iter < 0:
        raise ValueError("solve_contact: max_iter must be >= 0")
    if not np.all(np.isfinite(g0)):
        raise ValueError("solve_contact: g0 contains non-finite values")

    P_total = float(p_bar) * N
    gmin_v = float(g0.min())
    gmax_v = float(g0.max())
    g_ref = opt.g_ref
    if not (g_ref > 0.0):
        g_ref = gmax_v - gmin_v
    if not (g_ref > 0.0):
        g_ref = 1.0
    p_ref = opt.p_ref if opt.p_ref > 0.0 else float(p_bar)
    load_tol = opt.load_tol if opt.load_tol > 0.0 else (1e-12 if is_double else 5e-7)
    tol = opt.tol
    req_tol = opt.requested_tol if opt.requested_tol > 0.0 else tol
    if opt.datum_mode == "automatic":
        datum = gmin_v + 0.5 * (gmax_v - gmin_v); subtract = True
    elif opt.datum_mode == "solver":
        datum = opt.datum; subtract = True
    else:
        datum = opt.datum; subtract = False
    gshift = R(datum) if subtract else R(0)
    g0s = (g0 - gshift).astype(dtype)               # (g0(i) - gshift) in Real

    if p_init is not None:
        p = np.array(p_init, dtype=dtype).ravel()
        if p.size != N:
            raise ValueError("solve_contact: p_init size mismatch")
        p = np.maximum(p, R(0))
        s = p.astype(np.float64).sum()
        if s > 0.0:
            p = (p * R(P_total / s)).astype(dtype)
        else:
            p = np.full(N, p_bar, dtype=dtype)
    else:
        p = np.full(N, p_bar, dtype=dtype)
    t = np.zeros(N, dtype=dtype)
    u = np.zeros(N, dtype=dtype)
    g = np.zeros(N, dtype=dtype)
    g_prev = np.zeros(N, dtype=dtype)
    z = np.zeros(N, dtype=dtype)
    r = np.zeros(N, dtype=dtype)
    best_p = None
    res = Result()
    cnt = {"mv": 0, "pc": 0}
    t_start = time.perf_counter()

    def apply(x):
        cnt["mv"] += 1
        return np.asarray(S(x), dtype=dtype).ravel()

    st = {"G_old": 1.0, "delta": R(0), "s_id": 0.0, "u_fresh": False}

    def evaluate():
        nonlocal g
        d = _Diag()
        v = (u + g0s).astype(dtype)
        g = v.copy()
        vd = v.astype(np.float64)
        pos = p > 0
        nc = int(pos.sum())
        vmin = float(vd.min())
        gsum = float(vd[pos].sum())
        d.nc = nc
        d.alpha = gsum / nc if nc else vmin
        d.gmin = vmin - d.alpha
        alpha = R(d.alpha)
        gmin = R(d.gmin)
        pd = p.astype(np.float64)
        pv = float((pd * g.astype(np.float64)).sum())
        g = (g - alpha).astype(dtype)
        gd = g.astype(np.float64)
        gcs = float(gd[pos].sum())
        e = float((pd * np.abs(gd)).sum())
        Gc = float((pd * (g - gmin).astype(dtype).astype(np.float64)).sum())
        psum = float(pd.sum())
        if d.nc > 0:
            delta = gcs / d.nc
            d.alpha += delta
            d.gmin -= delta
        d.pk = e / (P_total * g_ref)
        d.G = Gc
        d.fw = Gc / (P_total * g_ref)
        d.pen = max(0.0, -d.gmin) / g_ref
        d.psum = psum
        d.pv = pv
        d.load = abs(psum - P_total) / P_total
        d.finite = all(math.isfinite(q) for q in (gsum, vmin, psum, e, Gc))
        return d

    def identification_step():
        nonlocal p, u, z, r, t, g_prev
        pd = p.astype(np.float64)
        f_cur = float((pd * (0.5 * u.astype(np.float64) + g0s.astype(np.float64))).sum())
        if not (st["s_id"] > 0.0):
            pp = float((pd * pd).sum())
            pu = float((pd * u.astype(np.float64)).sum())
            st["s_id"] = pp / pu if (pu > 0.0 and pp > 0.0) else 1.0
        for trial in range(12):
            s = R(st["s_id"])
            z = (p - s * g).astype(dtype)
            project_load_simplex(z, P_total)
            r = apply(z)
            zd = z.astype(np.float64)
            f_try = float((zd * (0.5 * r.astype(np.float64) + g0s.astype(np.float64))).sum())
            gd = float((g.astype(np.float64) * (z - p).astype(dtype).astype(np.float64)).sum())
            if math.isfinite(f_try) and gd < 0.0 and f_try <= f_cur + 1e-4 * gd:
                p, z = z, p
                u, r = r, u
                st["u_fresh"] = True
                st["delta"] = R(0)
                t = np.zeros(N, dtype=dtype)
                g_prev = np.zeros(N, dtype=dtype)
                st["G_old"] = 1.0
                res.identification_steps += 1
                if trial == 0:
                    st["s_id"] *= 2.0
                return True
            st["s_id"] *= 0.25
        return False

    status = "max_iterations"
    reason = ""
    best_merit = math.inf
    seen_merit = math.inf
    stall = 0
    forced_id = 0
    stall_limit = opt.stall_limit if opt.stall_limit > 0 else 200
    d = None
    it = 0
    while it < opt.max_iter:
        if not st["u_fresh"]:
            u = apply(p)
        st["u_fresh"] = False
        d = evaluate()
        if not d.finite:
            status = "nonfinite"
            break
        if opt.record_history:
            res.error_history.append(d.pk)
        if d.load <= load_tol and d.fw <= tol and d.pen <= tol:
            status = "converged"
            st["u_fresh"] = True
            break
        merit = max(d.fw, d.pen)
        if opt.keep_best and merit < best_merit:
            best_merit = merit
            best_p = p.copy()
        if merit < seen_merit * (1.0 - 1e-4):
            seen_merit = merit
            stall = 0
        else:
            stall += 1
            if stall >= stall_limit:
                if forced_id < 3 and identification_step():
                    forced_id += 1
                    stall = 0
                    it += 1
                    continue
                status = "stagnated"
                reason = "stall"
                break
        did_cg = False
        if d.nc >= 2:
            contact = p > 0
            if precond is not None:
                z = np.asarray(precond(g, contact), dtype=dtype).ravel()
                cnt["pc"] += 1
            else:
                z = np.where(contact, g, R(0)).astype(dtype)
            gd = g.astype(np.float64)[contact]
            zd = z.astype(np.float64)[contact]
            G = float((zd * gd).sum())
            G_pr = float((zd * (g - g_prev).astype(dtype).astype(np.float64)[contact]).sum())
            gg = float((gd * gd).sum())
            G_old = st["G_old"]
            beta_val = ((max(0.0, G_pr / G_old) if opt.use_pr else G / G_old)
                        if G_old != 0.0 else 0.0)
            beta = st["delta"] * R(beta_val)
            t = np.where(contact, z + beta * t, R(0)).astype(dtype)
            g_prev = g.copy()
            st["G_old"] = G
            r = apply(t)
            rsum = float(r.astype(np.float64)[contact].sum())
            num = float((g.astype(np.float64)[contact] * t.astype(np.float64)[contact]).sum())
            rmean = R(rsum / d.nc if d.nc else 0.0)
            den = float(((r - rmean).astype(dtype).astype(np.float64)[contact]
                         * t.astype(np.float64)[contact]).sum())
            if den > 0.0 and math.isfinite(den) and math.isfinite(num):
                tau = R(num / den)
                tau_ov = R(num * num / (den * gg)) if gg > 0.0 else tau
                was = p > 0
                pi = np.where(was, np.maximum(p - tau * t, R(0)), p).astype(dtype)
                act = (pi == 0) & (g < 0)
                pi = np.where(act, pi - tau_ov * g, pi).astype(dtype)
                changed = bool(np.any((pi > 0) != was))
                p = pi
                st["delta"] = R(0) if changed else R(1)
                total = float(p.astype(np.float64).sum())
                if total > 0.0:
                    p = (p * R(P_total / total)).astype(dtype)
                else:
                    p = np.full(N, p_bar, dtype=dtype)
                did_cg = True
        if not did_cg:
            if not identification_step():
                status = "nonpositive_curvature"
                break
        it += 1

    if status != "converged" and opt.keep_best and best_p is not None:
        merit = max(d.fw, d.pen)
        if not (merit <= best_merit):
            p = best_p
            st["u_fresh"] = False
            res.returned_best = True
    best_p = None
    if not st["u_fresh"]:
        u = apply(p)
    d = evaluate()
    if not d.finite and status != "nonfinite":
        status = "nonfinite"
    if status == "converged" and not (d.load <= load_tol and d.fw <= tol and d.pen <= tol):
        status = "verification_failed"
        reason = "final_check"
    if status == "converged" and req_tol < tol and not opt.allow_tolerance_relaxation:
        if not (d.fw <= req_tol and d.pen <= req_tol):
            status = "stagnated"
            reason = "precision_limit"

    if not opt.light:
        res.displacement = u.astype(np.float64)
        res.gap = g.astype(np.float64)
    res.approach = d.alpha + datum
    obj = float((p.astype(np.float64) * (0.5 * u.astype(np.float64)
                                          + g0s.astype(np.float64))).sum())
    res.objective = obj + datum * P_total
    res.iterations = it
    res.contact_fraction = d.nc / N
    res.mean_pressure = d.psum / N
    res.status = status
    res.status_reason = reason
    res.converged = status == "converged"
    res.error = res.pk_error = d.pk
    res.fw_gap = d.G
    res.fw_error = d.fw
    res.load_error = d.load
    res.penetration_error = d.pen
    res.requested_tol = req_tol
    res.effective_tol = tol
    res.p_ref = p_ref
    res.g_ref = g_ref
    res.matvec_count = cnt["mv"]
    res.precond_count = cnt["pc"]
    res.time_total = time.perf_counter() - t_start
    res.pressure = p.astype(np.float64)
    return res


# ── restricted (active-set) PK on a candidate set ───────────────────────────

def solve_contact_active(S: Callable, g0: np.ndarray, p_bar, idx, opt: SolveOptions = None,
                         precond: Optional[Callable] = None, p_init=None, n_grid: int = 0,
                         dtype=np.float64):
    """Restatement of solve_contact_active_impl<Real> (src/contact_solver.cpp:601-1040).

    The PK iteration of `solve_contact` with every per-iteration O(N) loop
    driven by the candidate index list `idx`: the pressure lives on idx only
    and every reduction, update and the identification step's projection
    range over idx. S and precond act on full-length vectors. Returns
    (Result, certificate) with the certificate of :1020-1030 (pv_sum,
    vmin_local, P_total, g_ref) in the frame of `g0`."""
    opt = opt or SolveOptions()
    R = np.dtype(dtype).type
    is_double = (np.dtype(dtype) == np.float64)
    g0 = np.ascontiguousarray(g0, dtype=dtype).ravel()
    idx = np.asarray(idx, dtype=np.int64).ravel()
    N, Nc = g0.size, idx.size
    p_bar = R(p_bar)
    if N == 0 or Nc == 0 or not (p_bar > 0):
        raise ValueError("solve_contact_active: empty gap/candidate set or p_bar <= 0")
    if not (opt.tol > 0.0):
        raise ValueError("solve_contact_active: tol must be positive")
    if opt.max_iter < 0:
        raise ValueError("solve_contact_active: max_iter must be >= 0")
    if not np.all(np.isfinite(g0)):
        raise ValueError("solve_contact_active: g0 contains non-finite values")
    N_grid = n_grid if n_grid > 0 else N
    P_total = float(p_bar) * N_grid
    gmin_v, gmax_v = float(g0.min()), float(g0.max())
    g_ref = opt.g_ref
    if not (g_ref > 0.0):
        g_ref = gmax_v - gmin_v
    if not (g_ref > 0.0):
        g_ref = 1.0
    if opt.datum_mode == "automatic":
        datum = gmin_v + 0.5 * (gmax_v - gmin_v); subtract = True
    elif opt.datum_mode == "solver":
        datum = opt.datum; subtract = True
    else:
        datum = opt.datum; subtract = False
    gshift = R(datum) if subtract else R(0)
    p_ref = opt.p_ref if opt.p_ref > 0.0 else float(p_bar)
    load_tol = opt.load_tol if opt.load_tol > 0.0 else (1e-12 if is_double else 5e-7)
    tol = opt.tol
    req_tol = opt.requested_tol if opt.requested_tol > 0.0 else tol
    p_unif = R(P_total / Nc)
    g0c = (g0[idx] - gshift).astype(dtype)            # (g0(i) - gshift) on idx

    p = np.zeros(N, dtype=dtype)
    if p_init is not None:
        pi0 = np.asarray(p_init, dtype=dtype).ravel()
        if pi0.size != N:
            raise ValueError("solve_contact_active: p_init size mismatch")
        p[idx] = np.maximum(pi0[idx], R(0))
        s = float(p[idx].astype(np.float64).sum())
        if s > 0.0:
            p[idx] = (p[idx] * R(P_total / s)).astype(dtype)
        else:
            p[idx] = p_unif
    else:
        p[idx] = p_unif
    t = np.zeros(N, dtype=dtype)
    u = np.zeros(N, dtype=dtype)
    g = np.zeros(N, dtype=dtype)
    g_prev = np.zeros(N, dtype=dtype)
    r = np.zeros(N, dtype=dtype)
    contact = np.zeros(N, dtype=bool)
    best_p = None
    res = Result()
    cnt = {"mv": 0, "pc": 0}
    st = {"G_old": 1.0, "delta": R(0), "s_id": 0.0, "u_fresh": False}
    t_start = time.perf_counter()
    f64 = lambda a: a.astype(np.float64)

    def apply(x):
        cnt["mv"] += 1
        return np.asarray(S(x), dtype=dtype).ravel()

    def evaluate():
        d = _Diag()
        v = (u[idx] + g0c).astype(dtype)
        pc_ = p[idx]
        pos = pc_ > 0
        vd = f64(v)
        vmin = float(vd.min())
        nc = int(pos.sum())
        gsum = float(vd[pos].sum())
        d.nc = nc
        d.alpha = gsum / nc if nc else vmin
        d.gmin = vmin - d.alpha
        alpha, gmin = R(d.alpha), R(d.gmin)
        pd = f64(pc_)
        pv = float((pd * vd).sum())
        gc = (v - alpha).astype(dtype)
        g[idx] = gc
        gd = f64(gc)
        gcs = float(gd[pos].sum())
        e = float((pd * np.abs(gd)).sum())
        Gc = float((pd * f64((gc - gmin).astype(dtype))).sum())
        psum = float(pd.sum())
        if nc > 0:
            dl = gcs / nc
            d.alpha += dl
            d.gmin -= dl
        d.pk = e / (P_total * g_ref)
        d.G = Gc
        d.fw = Gc / (P_total * g_ref)
        d.pen = max(0.0, -d.gmin) / g_ref
        d.psum = psum
        d.pv = pv
        d.load = abs(psum - P_total) / P_total
        d.finite = all(math.isfinite(q) for q in (gsum, vmin, psum, e, Gc))
        return d

    def identification_step():
        nonlocal p, u, r, t, g_prev
        pd = f64(p[idx])
        f_cur = float((pd * (0.5 * f64(u[idx]) + f64(g0c))).sum())
        if not (st["s_id"] > 0.0):
            pp = float((pd * pd).sum())
            pu = float((pd * f64(u[idx])).sum())
            st["s_id"] = pp / pu if (pu > 0.0 and pp > 0.0) else 1.0
        for trial in range(12):
            s = R(st["s_id"])
            y = (p[idx] - s * g[idx]).astype(dtype)
            project_load_simplex(y, P_total)
            z = np.zeros(N, dtype=dtype)
            z[idx] = y
            rr = apply(z)
            zd = f64(z[idx])
            f_try = float((zd * (0.5 * f64(rr[idx]) + f64(g0c))).sum())
            gd = float((f64(g[idx]) * f64((z[idx] - p[idx]).astype(dtype))).sum())
            if math.isfinite(f_try) and gd < 0.0 and f_try <= f_cur + 1e-4 * gd:
                p = z
                u, r = rr, u
                st["u_fresh"] = True
                st["delta"] = R(0)
                t = np.zeros(N, dtype=dtype)
                g_prev = np.zeros(N, dtype=dtype)
                st["G_old"] = 1.0
                res.identification_steps += 1
                if trial == 0:
                    st["s_id"] *= 2.0
                return True
            st["s_id"] *= 0.25
        return False

    status, reason = "max_iterations", ""
    best_merit = seen_merit = math.inf
    stall = forced_id = 0
    stall_limit = opt.stall_limit if opt.stall_limit > 0 else 200
    d = None
    it = 0
    while it < opt.max_iter:
        if not st["u_fresh"]:
            u = apply(p)
        st["u_fresh"] = False
        d = evaluate()
        if not d.finite:
            status = "nonfinite"
            break
        if opt.record_history:
            res.error_history.append(d.pk)
        if d.load <= load_tol and d.fw <= tol and d.pen <= tol:
            status = "converged"
            st["u_fresh"] = True
            break
        merit = max(d.fw, d.pen)
        if opt.keep_best and merit < best_merit:
            best_merit = merit
            best_p = p.copy()
        if merit < seen_merit * (1.0 - 1e-4):
            seen_merit = merit
            stall = 0
        else:
            stall += 1
            if stall >= stall_limit:
                if forced_id < 3 and identification_step():
                    forced_id += 1
                    stall = 0
                    it += 1
                    continue
                status, reason = "stagnated", "stall"
                break
        did_cg = False
        if d.nc >= 2:
            contact[idx] = p[idx] > 0
            if precond is not None:
                z = np.asarray(precond(g, contact), dtype=dtype).ravel()
                cnt["pc"] += 1
            else:
                z = np.zeros(N, dtype=dtype)
                z[idx] = np.where(contact[idx], g[idx], R(0))
            ci = idx[contact[idx]]
            gci = f64(g[ci])
            G = float((f64(z[ci]) * gci).sum())
            G_pr = float((f64(z[ci]) * f64((g[ci] - g_prev[ci]).astype(dtype))).sum())
            gg = float((gci * gci).sum())
            G_old = st["G_old"]
            beta_val = ((max(0.0, G_pr / G_old) if opt.use_pr else G / G_old)
                        if G_old != 0.0 else 0.0)
            beta = st["delta"] * R(beta_val)
            t[idx] = np.where(contact[idx], z[idx] + beta * t[idx], R(0)).astype(dtype)
            g_prev[idx] = g[idx]
            st["G_old"] = G
            r = apply(t)
            rsum = float(f64(r[ci]).sum())
            num = float((f64(g[ci]) * f64(t[ci])).sum())
            rmean = R(rsum / d.nc if d.nc else 0.0)
            den = float((f64((r[ci] - rmean).astype(dtype)) * f64(t[ci])).sum())
            if den > 0.0 and math.isfinite(den) and math.isfinite(num):
                tau = R(num / den)
                tau_ov = R(num * num / (den * gg)) if gg > 0.0 else tau
                pc_ = p[idx]
                was = pc_ > 0
                pi = np.where(was, np.maximum(pc_ - tau * t[idx], R(0)), pc_).astype(dtype)
                act = (pi == 0) & (g[idx] < 0)
                pi = np.where(act, pi - tau_ov * g[idx], pi).astype(dtype)
                changed = bool(np.any((pi > 0) != was))
                p[idx] = pi
                st["delta"] = R(0) if changed else R(1)
                total = float(f64(p[idx]).sum())
                if total > 0.0:
                    p[idx] = (p[idx] * R(P_total / total)).astype(dtype)
                else:
                    p[idx] = p_unif
                did_cg = True
        if not did_cg and not identification_step():
            status = "nonpositive_curvature"
            break
        it += 1

    if status != "converged" and opt.keep_best and best_p is not None:
        if not (max(d.fw, d.pen) <= best_merit):
            p = best_p
            st["u_fresh"] = False
            res.returned_best = True
    if not st["u_fresh"]:
        u = apply(p)
    d = evaluate()
    if not d.finite and status != "nonfinite":
        status = "nonfinite"
    if status == "converged" and not (d.load <= load_tol and d.fw <= tol and d.pen <= tol):
        status, reason = "verification_failed", "final_check"
    if status == "converged" and req_tol < tol and not opt.allow_tolerance_relaxation:
        if not (d.fw <= req_tol and d.pen <= req_tol):
            status, reason = "stagnated", "precision_limit"
    obj = float((f64(p[idx]) * (0.5 * f64(u[idx]) + f64(g0c))).sum())
    res.approach = d.alpha + datum
    res.objective = obj + datum * P_total
    res.iterations = it
    res.contact_fraction = d.nc / N_grid
    res.mean_pressure = d.psum / N_grid
    res.status, res.status_reason = status, reason
    res.converged = status == "converged"
    res.error = res.pk_error = d.pk
    res.fw_gap, res.fw_error = d.G, d.fw
    res.load_error, res.penetration_error = d.load, d.pen
    res.requested_tol, res.effective_tol = req_tol, tol
    res.p_ref, res.g_ref = p_ref, g_ref
    res.matvec_count, res.precond_count = cnt["mv"], cnt["pc"]
    res.time_total = time.perf_counter() - t_start
    res.pressure = f64(p)
    cert = dict(pv_sum=d.pv, vmin_local=d.gmin + d.alpha, P_total=P_total, g_ref=g_ref)
    return res, cert


def dilate_mask(m: np.ndarray, Ns: int, r: int) -> np.ndarray:
    """Chebyshev (square) dilation by radius r, window clipped at the grid
    edge, separable: rows then columns (src/nested_solve.cpp:76-101)."""
    M = np.asarray(m, dtype=bool).reshape(Ns, Ns)
    if r <= 0:
        return M.ravel().copy()
    H = np.zeros_like(M)
    for dx in range(-r, r + 1):
        if dx >= 0:
            H[:, :Ns - dx] |= M[:, dx:]
        else:
            H[:, -dx:] |= M[:, :Ns + dx]
    V = np.zeros_like(M)
    for dy in range(-r, r + 1):
        if dy >= 0:
            V[:Ns - dy, :] |= H[dy:, :]
        else:
            V[-dy:, :] |= H[:Ns + dy, :]
    return V.ravel()


def active_level(mv: Callable, precond: Optional[Callable], glvl, datum, p_bar, lvl_tol,
                 max_iter, use_pr, Ns, p_init_d, coarse_gap, light, g_ref, dtype,
                 active_delta=0.05, active_halo=2, active_max_rounds=5,
                 record_history=False) -> Result:
    """One restricted level (src/nested_solve.cpp:117-516) on a full-grid
    operator: candidate set C = dilate(p_init > 0) | {coarse gap < delta},
    restricted solve at tol/2, global verification with a fresh apply of the
    returned pressure (G = G_local + P (vmin_local - vmin_global)), extension
    by dilated violations, full-solve fallback. `glvl` is the level gap in
    double, `mv` applies the operator in `dtype`."""
    glvl = np.ascontiguousarray(glvl, dtype=np.float64).ravel()
    N = Ns * Ns
    R = np.dtype(dtype).type
    gmax, gmin = float(glvl.max()), float(glvl.min())
    g_scale = gmax - gmin if gmax > gmin else 1.0
    delta = active_delta * g_scale
    vthresh = -lvl_tol * g_scale
    cmask = dilate_mask(np.asarray(p_init_d) > 0.0, Ns, active_halo)
    cg = np.asarray(coarse_gap, dtype=np.float64).reshape(Ns // 2, Ns // 2)
    cmask |= (np.repeat(np.repeat(cg, 2, axis=0), 2, axis=1) < delta).ravel()
    g0w = (glvl - datum).astype(dtype)                 # centred working gap
    p0 = np.asarray(p_init_d, dtype=np.float64).astype(dtype)
    it_total = verify_mv = rounds = 0
    certified = False
    hist = []
    res = None
    pr = None
    while rounds < active_max_rounds:
        rounds += 1
        aopt = SolveOptions(tol=0.5 * lvl_tol, requested_tol=lvl_tol, max_iter=max_iter,
                            use_pr=use_pr, record_history=record_history,
                            allow_tolerance_relaxation=True, g_ref=g_ref,
                            datum_mode="caller", datum=datum)
        res, cert = solve_contact_active(mv, g0w, R(p_bar), np.flatnonzero(cmask), aopt,
                                         precond, p0, n_grid=N, dtype=dtype)
        p0 = None
        it_total += res.iterations
        hist += list(res.error_history)
        pr = res.pressure.astype(dtype)
        a = res.approach - datum
        v = np.asarray(mv(pr), dtype=dtype).astype(np.float64) + g0w.astype(np.float64)
        verify_mv += 1
        vmin_glob = float(v.min())
        viol = (~cmask) & (v - a < vthresh)
        nviol = int(viol.sum())
        G_glob = res.fw_gap + cert["P_total"] * (cert["vmin_local"] - vmin_glob)
        fw_glob = G_glob / (cert["P_total"] * cert["g_ref"])
        pen_glob = max(0.0, -(vmin_glob - a)) / cert["g_ref"]
        res.fw_gap, res.fw_error, res.penetration_error = G_glob, fw_glob, pen_glob
        res.effective_tol = lvl_tol
        if res.status == "converged" and nviol == 0 and fw_glob <= lvl_tol \
                and pen_glob <= lvl_tol:
            certified = True
            break
        if nviol == 0:
            break
        cmask |= dilate_mask(viol, Ns, active_halo)
        p0 = pr
    if not certified:
        pf = p0 if p0 is not None else pr
        double = np.dtype(dtype) == np.float64
        fopt = SolveOptions(tol=lvl_tol, max_iter=max_iter, use_pr=use_pr, light=light,
                            record_history=record_history, keep_best=not light,
                            allow_tolerance_relaxation=True, g_ref=g_ref,
                            datum_mode="solver" if double else "caller", datum=datum)
        res = solve_contact(mv, glvl if double else g0w, R(p_bar), fopt, precond, pf,
                            dtype=dtype)
        it_total += res.iterations
        hist += list(res.error_history)
        res.active_fallback = True
    elif not light:
        a = res.approach - datum
        disp = np.asarray(mv(pr), dtype=dtype).astype(np.float64)
        res.displacement = disp
        res.gap = disp + g0w.astype(np.float64) - a
    res.active_rounds = rounds
    res.iterations = it_total
    res.verification_matvec_count = verify_mv
    if not certified and not res.active_fallback and res.status == "converged":
        res.status, res.status_reason = "verification_failed", "global_check"
    res.converged = res.status == "converged"
    if record_history:
        res.error_history = hist
    return res


# ── L6: nested driver, fft / periodic branches (src/nested_solve.cpp) ───────

def restrict_field(f: np.ndarray, Ns: int) -> np.ndarray:
    """2x2 block average (src/nested_solve.cpp:42-54)."""
    F = f.reshape(Ns, Ns)
    return (0.25 * (F[0::2, 0::2] + F[0::2, 1::2] + F[1::2, 0::2] + F[1::2, 1::2])).ravel()


def prolong_field(pc: np.ndarray, Nc: int) -> np.ndarray:
    """Injection (src/nested_solve.cpp:57-71)."""
    P = pc.reshape(Nc, Nc)
    return np.repeat(np.repeat(P, 2, axis=0), 2, axis=1).ravel()


def solve_nested(Ns, g0, p_bar, L=1.0, E_star=1.0, coarsest=64, precond=True,
                 precond_engine="auto", precond_radius=2, precond_occupancy_max=0.4,
                 tol=1e-8, coarse_tol=1e-4, max_iter=20000, use_pr=True,
                 precision="double", float_floor=2e-6,
                 allow_tolerance_relaxation=False, light_result=False,
                 backend="fft", record_error_history=False, leaf_side=8,
                 active_set=False, active_delta=0.05, active_halo=2, active_max_rounds=5,
                 active_all_levels=True, active_occupancy_max=0.4,
                 operators: dict = None) -> Result:
    """solve_contact_nested for backend fft|periodic (src/nested_solve.cpp:518-990).

    With active_set=True (fft backend) every level with a coarser one beneath
    it whose predicted occupancy passes the gate is solved by `active_level`
    (src/nested_solve.cpp:765-870). `operators` may map level size -> prebuilt
    operator (to reuse tables)."""
    g0 = np.ascontiguousarray(g0, dtype=np.float64).ravel()
    if g0.size != Ns * Ns:
        raise ValueError("g0 size != Ns*Ns")
    levels = []
    n = coarsest
    while n <= Ns:
        levels.append(n)
        n *= 2
    if not levels or levels[-1] != Ns:
        raise ValueError("Ns must equal coarsest * 2^k")
    any_float = precision != "double"
    if active_set:
        if backend != "fft":
            raise ValueError("solve_contact_nested: active_set requires backend='fft'")
        if len(levels) < 2:
            raise ValueError("solve_contact_nested: active_set needs a coarse level "
                             "(Ns > coarsest)")
    ncoarse = len(levels) - 1
    gap = [None] * ncoarse
    if ncoarse > 0:
        gap[ncoarse - 1] = restrict_field(g0, Ns)
        for i in range(ncoarse - 2, -1, -1):
            gap[i] = restrict_field(gap[i + 1], levels[i + 1])
    gmn, gmx = float(g0.min()), float(g0.max())
    datum = gmn + 0.5 * (gmx - gmn)
    g_ref = max(gmx - gmn, p_bar * L / E_star)
    p_init = None
    coarse_gap = None
    prev_occ = -1.0
    stages = []
    res = None
    t0 = time.perf_counter()
    for li, n in enumerate(levels):
        finest = li + 1 == len(levels)
        if operators and n in operators:
            op = operators[n]
        elif backend == "fft":
            op = FreeSpaceOperator(n, L, E_star)
        else:
            op = PeriodicOperator(n, L, E_star)
        r_lvl = min(precond_radius, max(1, n // 4), leaf_side)
        engine = "fft" if (backend == "periodic" and precond_engine == "auto") else precond_engine
        use_stencil = stencil_for_level(engine, precond, prev_occ, precond_occupancy_max)
        pc_obj = None
        if use_stencil:
            pc_obj = StencilPreconditioner(n, r_lvl)
        elif precond:
            pc_obj = FourierPreconditioner(n)
        glvl = g0 if finest else gap[li]
        polish_after = finest and precision == "float_then_double"
        lvl_tol = tol if finest else coarse_tol
        if polish_after:
            lvl_tol = float_floor
        elif any_float:
            lvl_tol = max(lvl_tol, float_floor)
        gate_open_next = prev_occ < 0.0 or prev_occ < active_occupancy_max
        keep_gap = (active_set and not finest and gate_open_next
                    and (active_all_levels or li + 2 == len(levels)))
        light = light_result if finest else (not keep_gap)
        use_active = (active_set and p_init is not None and coarse_gap is not None
                      and prev_occ < active_occupancy_max and (finest or active_all_levels))
        lopt = SolveOptions(tol=lvl_tol, requested_tol=(tol if finest else coarse_tol),
                            max_iter=max_iter, use_pr=use_pr, light=light,
                            record_history=finest and record_error_history,
                            keep_best=not light,
                            allow_tolerance_relaxation=(not finest) or polish_after
                            or allow_tolerance_relaxation,
                            g_ref=g_ref, datum=datum,
                            datum_mode=("caller" if any_float else "solver"))
        dt = np.float32 if any_float else np.float64
        pcf = (lambda gg, cc, _o=pc_obj, _d=dt: _o.apply(gg, cc, _d)) if pc_obj else None
        mvf = lambda x, _o=op, _d=dt: _o.matvec(x, _d)
        if use_active:
            res = active_level(mvf, pcf, glvl, datum, p_bar, lvl_tol, max_iter, use_pr, n,
                               p_init, coarse_gap, light, g_ref, dt, active_delta,
                               active_halo, active_max_rounds,
                               record_history=finest and record_error_history)
            res.requested_tol = lopt.requested_tol
            coarse_gap = None
        elif any_float:
            gf = (glvl - datum).astype(np.float32)
            res = solve_contact(mvf, gf, np.float32(p_bar), lopt, pcf,
                                None if p_init is None else p_init.astype(np.float32),
                                dtype=np.float32)
        else:
            res = solve_contact(mvf, glvl, p_bar, lopt, pcf, p_init, dtype=np.float64)
        prev_occ = res.contact_fraction
        stages.append(dict(name=("finest:" if finest else "coarse:") + str(n)
                           + ("(active)" if use_active else ""),
                           precision="float" if any_float else "double",
                           requested_tol=lopt.requested_tol, effective_tol=lopt.tol,
                           iterations=res.iterations, matvec_count=res.matvec_count,
                           status=res.status, fw_error=res.fw_error,
                           penetration_error=res.penetration_error))
        if polish_after:
            popt = SolveOptions(tol=tol, requested_tol=tol, max_iter=max_iter,
                                use_pr=use_pr, light=light_result,
                                record_history=finest and record_error_history,
                                keep_best=not light_result,
                                allow_tolerance_relaxation=allow_tolerance_relaxation,
                                g_ref=g_ref, datum=datum, datum_mode="solver")
            before = res
            pcd = (lambda gg, cc, _o=pc_obj: _o.apply(gg, cc, np.float64)) if pc_obj else None
            res = solve_contact(lambda x, _o=op: _o.matvec(x, np.float64), glvl, p_bar,
                                popt, pcd, before.pressure, dtype=np.float64)
            stages.append(dict(name="polish:" + str(n), precision="double",
                               requested_tol=tol, effective_tol=tol,
                               iterations=res.iterations, matvec_count=res.matvec_count,
                               status=res.status, fw_error=res.fw_error,
                               penetration_error=res.penetration_error))
            res.iterations += before.iterations
            res.matvec_count += before.matvec_count
            res.precond_count += before.precond_count
            if record_error_history:
                res.error_history = list(before.error_history) + list(res.error_history)
        if not finest:
            p_init = prolong_field(res.pressure, n)
            coarse_gap = res.gap if keep_gap else None
    res.stage_stats = stages
    res.time_total = time.perf_counter() - t0
    return res


# ── independent diagnostics (tests/contact_oracle.hpp) ──────────────────────

def diagnose(matvec: Callable, g0, p, P_total, g_ref) -> dict:
    """tests/contact_oracle.hpp:35-61, with an arbitrary FP64 operator."""
    p = np.asarray(p, dtype=np.float64).ravel()
    v = np.asarray(matvec(p), dtype=np.float64) + np.asarray(g0, dtype=np.float64).ravel()
    pos = p > 0
    nc = int(pos.sum())
    approach = v[pos].sum() / nc if nc else v.min()
    min_gap = v.min() - approach
    G = float((p * (v - v.min())).sum())
    return dict(approach=approach, min_gap=min_gap, fw_gap=G,
                fw_error=G / (P_total * g_ref),
                penetration=max(0.0, -min_gap) / g_ref,
                load_error=abs(p.sum() - P_total) / P_total,
                min_pressure=float(p.min()),
                pk_error=float((p * np.abs(v - approach)).sum()) / (P_total * g_ref),
                objective=0.5 * float(p @ (v - np.asarray(g0).ravel())) + float(p @ np.asarray(g0).ravel()),
                n_contact=nc)


def sorting_projection(y, P_total):
    """tests/contact_oracle.hpp:64-80."""
    u = np.sort(np.asarray(y, dtype=np.float64))[::-1]
    css = np.cumsum(u)
    t = (css - P_total) / np.arange(1, u.size + 1)
    ok = u > t
    theta = t[np.nonzero(ok)[0][-1]] if ok.any() else 0.0
    return np.maximum(np.asarray(y, dtype=np.float64) - theta, 0.0)


def enumerate_qp(S, g0, P_total, tol=1e-10):
    """Support enumeration, n <= 12 (tests/contact_oracle.hpp:95-142)."""
    n = g0.size
    best = (math.inf, None, 0.0)
    for m in range(1, 1 << n):
        A = [i for i in range(n) if m & (1 << i)]
        k = len(A)
        K = np.zeros((k + 1, k + 1))
        rhs = np.zeros(k + 1)
        K[:k, :k] = S[np.ix_(A, A)]
        K[:k, k] = -1.0
        K[k, :k] = 1.0
        rhs[:k] = -g0[A]
        rhs[k] = P_total
        try:
            sol = np.linalg.solve(K, rhs)
        except np.linalg.LinAlgError:
            continue
        if np.linalg.norm(K @ sol - rhs) > 1e-8 * (1.0 + np.linalg.norm(rhs)):
            continue
        if np.any(sol[:k] < -tol):
            continue
        p = np.zeros(n)
        p[A] = np.maximum(sol[:k], 0.0)
        alpha = sol[k]
        v = S @ p + g0
        if np.any(v - alpha < -tol * (1.0 + abs(alpha))):
            continue
        obj = 0.5 * p @ (S @ p) + p @ g0
        if obj < best[0]:
            best = (obj, p, alpha)
    return best


def manufacture(S, p_star, g_star, alpha_star):
    """Manufactured KKT fixture (tests/contact_oracle.hpp:156-166)."""
    g0 = alpha_star - S @ p_star + g_star
    return g0, float(p_star.sum())


def spd_kernel(n, decay=1.0):
    """tests/contact_oracle.hpp:169-177."""
    i = np.arange(n)
    d = np.abs(i[:, None] - i[None, :]).astype(np.float64)
    S = np.exp(-decay * d) / (1.0 + 0.25 * d) + 0.1 * np.eye(n)
    return 0.5 * (S + S.T)


# ── adjoint of the contact solve (no counterpart in the reference) ─────────
#
# The reference solves min 1/2 p'Sp + g0'p s.t. p >= 0, sum p = P
# (doc/theory/pcg.tex:380-538; KKT residuals as in tests/contact_oracle.hpp:35-61)
# but has no derivative of its solution. On a strictly complementary solution
# with contact set C the KKT system S_CC p_C + g0_C = alpha 1, 1'p_C = P is
# linear, so for a loss with dL/dp = w:
#     [S_CC 1; 1' 0] [mu; nu] = [w_C; 0],   dL/dg0 = -mu on C (0 off C),
#     dL/dP = nu   (dL/dp_nominal = nu * N).
# Pinned by central finite differences through the exact support-enumeration
# solver and through solve_contact (tests/test_oracle.py).

def kkt_gradient(S, p, grad_p, contact=None):
    """Dense direct solve of the adjoint KKT system. Returns (dL/dg0, dL/dP)."""
    p = np.asarray(p, dtype=np.float64).ravel()
    w = np.asarray(grad_p, dtype=np.float64).ravel()
    c = (p > 0) if contact is None else np.asarray(contact, dtype=bool).ravel()
    idx = np.nonzero(c)[0]
    k = idx.size
    K = np.zeros((k + 1, k + 1))
    K[:k, :k] = S[np.ix_(idx, idx)]
    K[:k, k] = 1.0
    K[k, :k] = 1.0
    rhs = np.zeros(k + 1)
    rhs[:k] = w[idx]
    sol = np.linalg.solve(K, rhs)
    g = np.zeros(p.size)
    g[idx] = -sol[:k]
    return g, float(sol[k])


def circulant_matrix(matvec, Ns):
    """Dense matrix of a translation-invariant periodic operator on the
    Ns x Ns torus from its response to a unit impulse."""
    e = np.zeros(Ns * Ns)
    e[0] = 1.0
    c0 = np.asarray(matvec(e), dtype=np.float64).reshape(Ns, Ns)
    S = np.empty((Ns * Ns, Ns * Ns))
    for jy in range(Ns):
        for jx in range(Ns):
            S[:, jy * Ns + jx] = np.roll(np.roll(c0, jy, axis=0), jx, axis=1).ravel()
    return S


# ── rough-surface generator (the benchmark recipe; rfgen is third-party) ───
#
# The reference's benchmarks draw surfaces from `rfgen` (pyproject.toml:15,
# bench_backend_precision_study.py:37-59), which is not part of the
# reference tree and not installed. Recipe: amplitude |m|^-(1+H) for integer
# wavenumber magnitude in [k_low, k_high] (fftfreq order), uniform random
# phase, real part of ifft2, mean removed, scaled to rms. The device
# generator draws the phase of frequency (kx, ky) as SplitMix64 output number
# ky*Ns + kx (Steele, Lea, Flood, OOPSLA 2014) so that it can be restated here.

_SM_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, idx):
    """Output number idx (0-based) of a SplitMix64 generator seeded with seed."""
    z = np.uint64(seed) + (np.asarray(idx, dtype=np.uint64) + np.uint64(1)) * _SM_GOLDEN
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def hash_uniform(seed, idx):
    """53-bit uniform in [0, 1) from splitmix64."""
    return (splitmix64(seed, idx) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def rough_surface_from_phases(Ns, H, k_low, k_high, rms, phase):
    """phase: (Ns, Ns) in [0, 1), indexed [ky, kx]."""
    k = np.fft.fftfreq(Ns, 1.0 / Ns)
    kk = np.hypot(k[:, None], k[None, :])
    kk[0, 0] = 1.0
    amp = np.where((kk >= k_low) & (kk <= k_high), kk ** (-(1.0 + H)), 0.0)
    amp[0, 0] = 0.0
    h = np.real(np.fft.ifft2(amp * np.exp(2j * np.pi * phase)))
    h -= h.mean()
    h *= rms / h.std()
    return h


def rough_surface_hash(Ns, H=0.8, k_low=4.0, k_high=64.0, rms=0.002, seed=0):
    """The device generator's surface (phases from splitmix64)."""
    idx = np.arange(Ns * Ns, dtype=np.uint64).reshape(Ns, Ns)
    return rough_surface_from_phases(Ns, H, k_low, k_high, rms, hash_uniform(seed, idx))


# ── tangential (Cerruti) operator: src/cerruti_kernel.cpp,
#    src/tangential_operator.cpp:16-157, src/fourier_precond.cpp:234-359 ──────

def cerruti_uxx(x, y, a, b, nu):
    """Vectorised cerruti_uxx bracket (src/cerruti_kernel.cpp:85-95)."""
    x, y, a, b = (np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64),
                                                        np.broadcast(x, y, a, b).shape)).ravel()
                  for v in (x, y, a, b))
    uxx, uxy = np.empty_like(x), np.empty_like(x)
    lib().oracle_cerruti_many(x.size, _ptr(x), _ptr(y), _ptr(a), _ptr(b), float(nu),
                              _ptr(uxx), _ptr(uxy))
    return uxx


def cerruti_uxy(x, y, a, b):
    """Vectorised cerruti_uxy bracket (src/cerruti_kernel.cpp:97-123)."""
    x, y, a, b = (np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64),
                                                        np.broadcast(x, y, a, b).shape)).ravel()
                  for v in (x, y, a, b))
    uxx, uxy = np.empty_like(x), np.empty_like(x)
    lib().oracle_cerruti_many(x.size, _ptr(x), _ptr(y), _ptr(a), _ptr(b), 0.0,
                              _ptr(uxx), _ptr(uxy))
    return uxy


class CerrutiTables:
    """CerrutiKernel (include/cerruti_kernel.hpp:34-72): xx and xy tables and
    the signed-offset accessors."""

    def __init__(self, Ns, L=1.0, E_star=1.0, nu=0.3):
        self.Ns, self.L, self.E, self.nu = Ns, L, E_star, nu
        self.xx = np.empty((Ns, Ns))
        self.xy = np.empty((Ns, Ns))
        lib().oracle_cerruti_tables(int(Ns), float(L), float(E_star), float(nu),
                                    _ptr(self.xx), _ptr(self.xy))

    def xx_off(self, dx, dy):
        dx, dy = np.abs(dx), np.abs(dy)
        ok = (dx < self.Ns) & (dy < self.Ns)
        return np.where(ok, self.xx[np.minimum(dy, self.Ns - 1), np.minimum(dx, self.Ns - 1)], 0.0)

    def yy_off(self, dx, dy):
        return self.xx_off(dy, dx)

    def xy_off(self, dx, dy):
        dx, dy = np.asarray(dx), np.asarray(dy)
        ax, ay = np.abs(dx), np.abs(dy)
        ok = (ax < self.Ns) & (ay < self.Ns) & (dx != 0) & (dy != 0)
        v = self.xy[np.minimum(ay, self.Ns - 1), np.minimum(ax, self.Ns - 1)]
        return np.where(ok, np.where((dx > 0) == (dy > 0), v, -v), 0.0)

    def dense(self):
        """[[XX, XY], [XY, YY]] (src/cerruti_kernel.cpp:129-144)."""
        Ns = self.Ns
        i = np.arange(Ns * Ns)
        dix = (i % Ns)[:, None] - (i % Ns)[None, :]
        diy = (i // Ns)[:, None] - (i // Ns)[None, :]
        xx, xy, yy = self.xx_off(dix, diy), self.xy_off(dix, diy), self.yy_off(dix, diy)
        return np.block([[xx, xy], [xy, yy]])

    def symbol(self, kx, ky):
        """Continuum 2x2 symbol (src/cerruti_kernel.cpp:146-161)."""
        k2 = kx * kx + ky * ky
        c = 2.0 / (self.E * (1.0 - self.nu) * np.sqrt(k2))
        return np.array([[c * (1 - self.nu * kx * kx / k2), -c * self.nu * kx * ky / k2],
                         [-c * self.nu * kx * ky / k2, c * (1 - self.nu * ky * ky / k2)]])


class TangentialOperator:
    """TangentialFFTOperator (src/tangential_operator.cpp:16-157): the three
    tables embedded in wrap-around order on the (2Ns)^2 grid, real spectra
    (checked), u = [Kxx Kxy; Kxy Kyy] q per mode; q, u stacked [x; y]."""

    def __init__(self, Ns, L=1.0, E_star=1.0, nu=0.3):
        self.Ns = Ns
        self.k = CerrutiTables(Ns, L, E_star, nu)
        M = 2 * Ns
        d = np.arange(M)
        d = np.where(d < Ns, d, d - M)
        DX, DY = np.meshgrid(d, d, indexing="xy")      # [iy, ix]
        zero = (np.arange(M)[None, :] == Ns) | (np.arange(M)[:, None] == Ns)
        self.K = {}
        for name, f in (("xx", self.k.xx_off), ("yy", self.k.yy_off), ("xy", self.k.xy_off)):
            E = np.where(zero, 0.0, f(DX, DY))
            C = _fwd(E)
            if np.abs(C.imag).max() > 1e-10 * np.abs(C.real).max():
                raise RuntimeError("tangential spectrum not real")
            self.K[name] = C.real / (M * M)

    def matvec(self, q):
        Ns, M = self.Ns, 2 * self.Ns
        N = Ns * Ns
        q = np.asarray(q, dtype=np.float64)
        G1 = np.zeros((M, M))
        G2 = np.zeros((M, M))
        G1[:Ns, :Ns] = q[:N].reshape(Ns, Ns)
        G2[:Ns, :Ns] = q[N:].reshape(Ns, Ns)
        A, B = _fwd(G1), _fwd(G2)
        U1 = self.K["xx"] * A + self.K["xy"] * B
        U2 = self.K["xy"] * A + self.K["yy"] * B
        u1, u2 = _inv(U1, M), _inv(U2, M)
        return np.concatenate([u1[:Ns, :Ns].ravel(), u2[:Ns, :Ns].ravel()])


class TangentialPreconditioner:
    """TangentialFourierPreconditioner (src/fourier_precond.cpp:234-359):
    masked scatter, per-mode 2x2 multiply by |k|/Ns^2 (I + gamma k k^T/|k|^2),
    gamma = nu/(1-nu), cross term zeroed on self-conjugate Nyquist modes,
    masked gather and optional per-component mask-mean removal."""

    def __init__(self, Ns, nu):
        self.Ns = Ns
        g = nu / (1.0 - nu)
        norm = 1.0 / (Ns * Ns)
        ky = np.arange(Ns)
        fy = np.where(ky < Ns // 2, ky, ky - Ns).astype(np.float64)[:, None]
        fx = np.arange(Ns // 2 + 1, dtype=np.float64)[None, :]
        k2 = fx * fx + fy * fy
        kk = np.sqrt(k2)
        with np.errstate(invalid="ignore", divide="ignore"):
            self.wxx = np.where(k2 > 0, norm * kk * (1 + g * fx * fx / k2), 0.0)
            self.wyy = np.where(k2 > 0, norm * kk * (1 + g * fy * fy / k2), 0.0)
            nyq = (fx == Ns // 2) | (fy == Ns // 2) | (np.arange(Ns)[:, None] == Ns // 2)
            self.wxy = np.where((k2 > 0) & ~nyq, norm * kk * g * fx * fy / k2, 0.0)

    def apply(self, g, mask, remove_mean=False):
        Ns = self.Ns
        N = Ns * Ns
        m = np.asarray(mask).astype(bool).ravel()
        g = np.asarray(g, dtype=np.float64)
        A = _fwd(np.where(m, g[:N], 0.0).reshape(Ns, Ns))
        B = _fwd(np.where(m, g[N:], 0.0).reshape(Ns, Ns))
        Z1 = _inv(self.wxx * A + self.wxy * B, Ns).ravel()
        Z2 = _inv(self.wxy * A + self.wyy * B, Ns).ravel()
        z1 = np.where(m, Z1, 0.0)
        z2 = np.where(m, Z2, 0.0)
        if remove_mean and m.any():
            z1[m] -= Z1[m].sum() / m.sum()
            z2[m] -= Z2[m].sum() / m.sum()
        return np.concatenate([z1, z2])
