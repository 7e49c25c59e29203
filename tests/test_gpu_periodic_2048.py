"""Doubly periodic backend in FP64 at Ns=2048: the self-affine rough surface
(H=0.8, k in [4, 64]/Ns, rms 0.002) for seeds 0-3 at p_nominal 0.001, 0.005
and 0.02, and the single sinusoid (Westergaard) at partial and full contact.

At this size the oracle's own solve takes minutes on the CPU, so parity is
pinned by size-independent properties instead:
- every solve is certified by the oracle's independent FP64 operator
  (FW gap, penetration and load, tests/contact_oracle.hpp:35-61);
- the minimiser is unique (strictly convex QP), so two different GPU paths
  (unpreconditioned single level vs. nested with the full-grid Fourier
  preconditioner, src/nested_solve.cpp:518-990) converged at tol 1e-13 must
  agree in pressure to 1e-10 and in contact set except at threshold points;
- Westergaard's closed form (tests/test_periodic.cpp:94-129).
The matvec itself is pinned against the oracle at this size in
test_gpu_operators.py; the solver logic against the oracle's solve at Ns=256
in test_config_periodic_256_parity_1e10 below.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

NS = 2048


@pytest.fixture(scope="module")
def hc():
    import aspher_amd
    assert aspher_amd.device_count() >= 1, "no GPU visible"
    O.WORKERS = 8
    return aspher_amd


@pytest.fixture(scope="module")
def op2048():
    return O.PeriodicOperator(NS, 1.0, 1.0)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_rough_2048_certified(hc, op2048, seed):
    g0 = hc.surfaces.rough_gap(NS, seed=seed)
    s = hc.ContactSolver(NS, backend="periodic")
    for pbar in (0.001, 0.005, 0.02):
        print(f"seed={seed} pbar={pbar}", flush=True)
        r = s.solve(g0, pbar, tol=1e-8, max_iter=20000, precond="fourier-fft")
        assert r.status == "converged", (seed, pbar, r.status, r.status_reason)
        p = r.pressure.ravel()
        assert p.min() >= 0.0
        assert abs(p.mean() - pbar) <= 1e-12 * pbar
        d = O.diagnose(op2048.matvec, g0, p, pbar * NS * NS, np.ptp(g0))
        assert d["fw_error"] <= 1.01e-8, (seed, pbar, d)
        assert d["penetration"] <= 1.01e-8, (seed, pbar, d)
        assert d["load_error"] <= 1e-12, (seed, pbar, d)
        # the solver's own certificate agrees with the oracle's
        assert abs(r.fw_error - d["fw_error"]) <= 1e-9
        assert 0.0 < r.contact_area < 1.0


@pytest.mark.parametrize("pbar", [0.001, 0.005, 0.02])
def test_rough_2048_two_paths_agree_1e10(hc, op2048, pbar):
    g0 = hc.surfaces.rough_gap(NS, seed=0)
    s = hc.ContactSolver(NS, backend="periodic")
    ra = s.solve(g0, pbar, tol=1e-13, max_iter=50000, precond="none")
    assert ra.status == "converged", (ra.status, ra.status_reason)
    rb = hc.solve_nested(grid_size=NS, gap=g0, p_nominal=pbar, coarsest=64,
                         backend="periodic", tol=1e-13, max_iter=50000)
    assert rb.status == "converged", (rb.status, rb.status_reason)
    pa, pb = ra.pressure.ravel(), rb.pressure.ravel()
    rel = np.linalg.norm(pa - pb) / np.linalg.norm(pb)
    assert rel <= 1e-10, rel
    diff = (pa > 0) != (pb > 0)
    assert np.all(np.maximum(pa, pb)[diff] < 1e-12 * pb.max())
    for p in (pa, pb):
        d = O.diagnose(op2048.matvec, g0, p, pbar * NS * NS, np.ptp(g0))
        assert d["fw_error"] <= 1.01e-13 and d["penetration"] <= 1.01e-13, d


def test_westergaard_2048(hc):
    """Single sinusoid of amplitude 0.01 and wavelength L: full contact at
    p = 1.5 p*, partial at 0.3 p*, p* = pi E* D / L."""
    L, Es, D = 1.0, 1.0, 0.01
    pstar = math.pi * Es * D / L
    g0 = hc.surfaces.westergaard_gap(NS, L=L, amplitude=D)
    s = hc.ContactSolver(NS, domain_size=L, E_star=Es, backend="periodic")
    cosx = np.tile(np.cos(2 * math.pi * np.arange(NS) / NS), (NS, 1)).ravel()
    pbar = 1.5 * pstar
    r = s.solve(g0, pbar, tol=1e-12, precond="fourier-fft")
    assert r.converged
    pex = pbar + pstar * cosx
    assert np.linalg.norm(r.pressure.ravel() - pex) / np.linalg.norm(pex) < 1e-9
    assert r.contact_area == 1.0
    pbar = 0.3 * pstar
    r = s.solve(g0, pbar, tol=1e-10, precond="fourier-fft")
    assert r.converged
    row = r.pressure.ravel()[:NS]
    width = (row > 0).sum() * L / NS
    wex = 2 * (L / math.pi) * math.asin(math.sqrt(pbar / pstar))
    assert abs(width - wex) <= 2 * L / NS
    # every row carries the same pressure profile (the solution is 1-D)
    P = r.pressure.reshape(NS, NS)
    assert np.abs(P - P[0]).max() <= 1e-9 * P.max()


def test_config_periodic_256_parity_1e10(hc):
    """The same surface recipe at Ns=256 against the oracle's own solve
    (src/contact_solver.cpp:150-587 with src/periodic_fft_operator.cpp and
    src/fourier_precond.cpp), both at tol 1e-13."""
    ns = 256
    g0 = hc.surfaces.rough_gap(ns, seed=1)
    s = hc.ContactSolver(ns, backend="periodic")
    r = s.solve(g0, 0.005, tol=1e-13, max_iter=20000, precond="fourier-fft")
    assert r.status == "converged"
    op = O.PeriodicOperator(ns, 1.0, 1.0)
    fp = O.FourierPreconditioner(ns)
    ro = O.solve_contact(op.matvec, g0, 0.005, O.SolveOptions(tol=1e-13, max_iter=20000),
                         fp.apply)
    assert ro.status == "converged"
    p, po = r.pressure.ravel(), ro.pressure
    assert np.linalg.norm(p - po) / np.linalg.norm(po) <= 1e-10
    diff = (p > 0) != (po > 0)
    assert np.all(np.maximum(p, po)[diff] < 1e-12 * po.max())
