"""Adjoint of the certified contact solve on the GPU (no counterpart in the
reference): dL/dgap and dL/dp_nominal of L = 1/2 sum (p - p_target)^2.

Pinned three ways:
- at Ns=64 against the oracle's dense direct solve of the adjoint KKT system
  (oracle.kkt_gradient on the FP64 dense Love matrix, resp. the dense
  circulant of the periodic operator), itself pinned by finite differences
  through exact solutions in tests/test_oracle.py;
- at Ns=1024 by central finite differences of the GPU forward solve (step
  1e-7 rms) at 32 seeded points away from the contact boundary;
- through torch.autograd (the Function's backward is the same adjoint).
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

RMS = 0.002
PBAR = 0.005


@pytest.fixture(scope="module")
def hc():
    import aspher_amd
    assert aspher_amd.device_count() >= 1, "no GPU visible"
    O.WORKERS = 8
    return aspher_amd


def _precond(backend):
    return "fourier-fft" if backend == "periodic" else "fourier"


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("backend", ["fft", "periodic"])
def test_adjoint_matches_dense_reference_64(hc, backend):
    Ns, N = 64, 64 * 64
    g0 = hc.surfaces.rough_gap(Ns, seed=0)
    gt = hc.surfaces.rough_gap(Ns, seed=1)
    s = hc.ContactSolver(Ns, backend=backend)
    pt = s.solve(gt, PBAR, tol=1e-12, max_iter=20000, precond=_precond(backend)).pressure.ravel()
    r = s.solve(g0, PBAR, tol=1e-12, max_iter=20000, precond=_precond(backend))
    assert r.status == "converged"
    p = r.pressure.ravel()
    w = p - pt
    gg, gpn, info = s.adjoint(r, w, tol=1e-12)
    assert info["converged"], info
    assert info["n_contact"] == int((p > 0).sum())
    if backend == "fft":
        S = O.dense_matrix(Ns)
    else:
        S = O.circulant_matrix(O.PeriodicOperator(Ns, 1.0, 1.0).matvec, Ns)
    ref_g, ref_nu = O.kkt_gradient(S, p, w)
    gg = gg.ravel()
    assert _rel(gg, ref_g) <= 1e-8, _rel(gg, ref_g)
    assert abs(gpn - ref_nu * N) <= 1e-8 * abs(ref_nu * N)
    assert np.all(gg[p == 0] == 0.0)
    # the support is the oracle's: an independent FP64 dense solve agrees
    ro = O.solve_contact(lambda v, dtype=np.float64: S @ v, g0, PBAR,
                         O.SolveOptions(tol=1e-13, max_iter=20000))
    diff = (p > 0) != (ro.pressure > 0)
    assert np.all(np.maximum(p, ro.pressure)[diff] < 1e-12 * p.max())


def _fd_points(p, gap, rng, n=32):
    """Seeded points at least two cells inside the contact region or at
    least two cells away from it."""
    Ns = int(round(p.size ** 0.5))
    c = (p > 0).reshape(Ns, Ns)
    interior = c.copy()
    exterior = ~c
    for dy in (-2, -1, 0, 1, 2):
        for dx in (-2, -1, 0, 1, 2):
            sh = np.roll(np.roll(c, dy, 0), dx, 1)
            interior &= sh
            exterior &= ~sh
    ii = np.flatnonzero(interior.ravel())
    ee = np.flatnonzero(exterior.ravel())
    k_in = min(n - 4, ii.size)
    return np.concatenate([rng.choice(ii, k_in, replace=False),
                           rng.choice(ee, n - k_in, replace=False)])


@pytest.mark.parametrize("backend", ["fft", "periodic"])
def test_adjoint_central_differences_1024(hc, backend):
    Ns = 1024
    g0 = hc.surfaces.rough_gap(Ns, seed=0)
    gt = hc.surfaces.rough_gap(Ns, seed=1)
    s = hc.ContactSolver(Ns, backend=backend)
    pc = _precond(backend)
    pt = s.solve(gt, PBAR, tol=1e-10, max_iter=20000, precond=pc).pressure.ravel()
    # the forward solution to FP64 stagnation: tol below what FP64 certifies,
    # the solver returns its best iterate
    deep = dict(tol=1e-30, max_iter=20000, precond=pc)
    r = s.solve(g0, PBAR, **deep)
    p = r.pressure.ravel()
    assert r.fw_error < 1e-12 and r.penetration_error < 1e-12, (r.fw_error, r.status)
    w = p - pt
    gg, gpn, info = s.adjoint(r, w, tol=1e-12)
    assert info["converged"], info
    gg = gg.ravel()
    rng = np.random.default_rng(2024)
    pts = _fd_points(p, g0, rng)
    eps = 1e-7 * RMS
    fd = np.empty(pts.size)
    for k, i in enumerate(pts):
        print(f"{backend} fd point {k} index {i}", flush=True)
        e = np.zeros_like(g0)
        e[i] = eps
        pp = s.solve(g0 + e, PBAR, p_init=p, **deep).pressure.ravel()
        pm = s.solve(g0 - e, PBAR, p_init=p, **deep).pressure.ravel()
        # L(p+) - L(p-) = sum (p+ - p-)((p+ + p-)/2 - pt), without cancellation
        fd[k] = float(((pp - pm) * (0.5 * (pp + pm) - pt)).sum()) / (2 * eps)
    ad = gg[pts]
    scale = np.abs(ad).max()
    assert scale > 0
    err = np.abs(fd - ad)
    assert err.max() <= 1e-3 * scale, (err.max(), scale, list(zip(fd, ad)))
    assert np.linalg.norm(fd - ad) <= 1e-3 * np.linalg.norm(ad)
    # p_nominal: step 1e-7 p_nominal
    ep = 1e-7 * PBAR
    pp = s.solve(g0, PBAR + ep, p_init=p, **deep).pressure.ravel()
    pm = s.solve(g0, PBAR - ep, p_init=p, **deep).pressure.ravel()
    fdp = float(((pp - pm) * (0.5 * (pp + pm) - pt)).sum()) / (2 * ep)
    assert abs(fdp - gpn) <= 1e-3 * abs(gpn), (fdp, gpn)


def test_adjoint_cost_at_most_twice_forward_1024(hc):
    Ns = 1024
    g0 = hc.surfaces.rough_gap(Ns, seed=0)
    gt = hc.surfaces.rough_gap(Ns, seed=1)
    kw = dict(coarsest=64, backend="fft", tol=1e-8, precision="double", reuse_operators=True)
    pt = hc.solve_nested(Ns, gt, PBAR, **kw).pressure.ravel()
    t0 = time.perf_counter()
    r = hc.solve_nested(Ns, g0, PBAR, **kw)
    t_fwd = time.perf_counter() - t0
    assert r.converged
    s = hc.ContactSolver(Ns, backend="fft", shared_operator=True)
    w = r.pressure.ravel() - pt
    s.adjoint(r, w, tol=1e-10)   # warm (preconditioner set-up, first launches)
    t0 = time.perf_counter()
    gg, gpn, info = s.adjoint(r, w, tol=1e-10)
    t_adj = time.perf_counter() - t0
    print(f"forward {t_fwd:.4f} s, adjoint {t_adj:.4f} s ({info['iterations']} iterations, "
          f"{info['precond']})", flush=True)
    assert info["converged"]
    assert t_adj <= 2.0 * t_fwd, (t_adj, t_fwd)


def test_torch_autograd_function(hc):
    torch = pytest.importorskip("torch")
    from aspher_amd.autograd import ContactConfig, contact_pressure
    Ns = 128
    cfg = ContactConfig(grid_size=Ns, backend="periodic", tol=1e-12, coarsest=32)
    g0 = torch.tensor(hc.surfaces.rough_gap(Ns, seed=0).reshape(Ns, Ns), requires_grad=True)
    pn = torch.tensor(PBAR, dtype=torch.float64, requires_grad=True)
    pt = torch.tensor(hc.solve_nested(Ns, hc.surfaces.rough_gap(Ns, seed=1), PBAR, coarsest=32,
                                      backend="periodic", tol=1e-12).pressure)
    p = contact_pressure(g0, pn, cfg)
    assert p.dtype == torch.float64 and p.shape == (Ns, Ns)
    loss = 0.5 * ((p - pt) ** 2).sum()
    dg, dpn = torch.autograd.grad(loss, (g0, pn))
    # the same adjoint called directly
    res = hc.solve_nested(Ns, g0.detach().numpy(), PBAR, coarsest=32, backend="periodic",
                          tol=1e-12, reuse_operators=True, precision="double")
    s = hc.ContactSolver(Ns, backend="periodic", shared_operator=True)
    gg, gpn, _ = s.adjoint(res, (res.pressure - pt.numpy()).ravel())
    assert _rel(dg.numpy().ravel(), gg.ravel()) <= 1e-8
    assert abs(float(dpn) - gpn) <= 1e-8 * abs(gpn)
    # zero off the support, nonzero on it
    pc = p.detach().numpy().ravel() > 0
    assert np.all(dg.numpy().ravel()[~pc] == 0.0) and np.any(dg.numpy().ravel()[pc] != 0.0)


def test_torch_autograd_device_tensors(hc):
    """Device tensors stay in HBM: the forward solve reads the gap through
    its device pointer, the pressure is copied device to device into the
    output tensor, and the backward passes dL/dp and dL/dgap as device
    pointers. Run in a fresh process that imports torch first (torch and the
    library then share torch's HIP runtime): pressure and both gradients
    equal the host-array path's bit for bit."""
    pytest.importorskip("torch")
    import os
    import subprocess
    import sys
    import textwrap
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    script = textwrap.dedent("""
        import sys
        import numpy as np
        import torch
        sys.path.insert(0, %r)
        assert torch.cuda.is_available()
        import aspher_amd as hc
        from aspher_amd import autograd as ag
        assert ag._one_hip_runtime()
        calls = []
        real = ag._solve_device
        def counted(*a, **k):
            calls.append(1)
            return real(*a, **k)
        ag._solve_device = counted
        Ns, pbar = 128, 0.005
        cfg = ag.ContactConfig(grid_size=Ns, backend="fft", tol=1e-12, coarsest=32)
        gap = hc.surfaces.rough_gap(Ns, seed=0).reshape(Ns, Ns)
        pt = hc.solve_nested(Ns, hc.surfaces.rough_gap(Ns, seed=1), pbar, coarsest=32,
                             backend="fft", tol=1e-12).pressure
        out = {}
        for dev in ("cpu", "cuda"):
            g0 = torch.tensor(gap, device=dev, requires_grad=True)
            pn = torch.tensor(pbar, dtype=torch.float64, device=dev, requires_grad=True)
            p = ag.contact_pressure(g0, pn, cfg)
            assert p.device.type == dev and p.dtype == torch.float64
            loss = 0.5 * ((p - torch.tensor(pt, device=dev)) ** 2).sum()
            dg, dpn = torch.autograd.grad(loss, (g0, pn))
            assert dg.device.type == dev
            out[dev] = (p.detach().cpu().numpy(), dg.cpu().numpy(), float(dpn.cpu()))
        assert len(calls) == 1, calls
        assert np.array_equal(out["cpu"][0], out["cuda"][0])
        assert np.array_equal(out["cpu"][1], out["cuda"][1])
        assert out["cpu"][2] == out["cuda"][2]
        assert np.any(out["cuda"][1] != 0.0)
        print("device autograd ok")
    """ % root)
    print("device autograd subprocess", flush=True)
    r = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout[-2000:], r.stderr[-3000:], flush=True)
    assert r.returncode == 0 and "device autograd ok" in r.stdout
