"""Device paths of the load-simplex projection and of the failure statuses,
through the C ABI.

* The device projection against the oracle's sorting projector on the 200-case
  sweep of T21 (tests/test_certification.cpp:443-467) and at 4096^2.
* `nonfinite` from an operator that starts returning NaN (T03,
  tests/test_certification.cpp:270-288) on the single-solve, nested and
  ensemble paths, and `verification_failed` from a final fresh evaluation that
  disagrees with the loop's (src/contact_solver.cpp:529-537). The C ABI's test
  hook (aspher_test_fault) stands in for the reference's wrapped functor.
* The nested precision policy: precision '' or 'double' with
  single_precision=True runs float (src/nested_solve.cpp:584-587)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def hc():
    import aspher_amd
    assert aspher_amd.device_count() >= 1, "no GPU visible"
    return aspher_amd


def _project(y, P, precision=64):
    from aspher_amd._capi import lib, check, dptr
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.empty_like(y)
    th = C.c_double()
    check(lib().aspher_project_load_simplex(y.size, dptr(y), float(P), precision, dptr(out),
                                            C.byref(th)))
    return out, th.value


def test_device_projection_matches_sorting_oracle(hc):
    """T21: 200 cases, n = 1..40, scales 1e-6..1e6, repeated values every
    fifth case; load, sign, oracle agreement and idempotence."""
    rng = np.random.default_rng(7)
    for trial in range(200):
        n = 1 + trial % 40
        scale = 10.0 ** ((trial % 13) - 6)
        y = scale * rng.uniform(-1, 1, n)
        if trial % 5 == 0:
            y[:] = scale
        P = 0.37 * scale * n
        prod, _ = _project(y, P)
        ref = O.sorting_projection(y, P)
        assert prod.min() >= 0, trial
        assert abs(prod.sum() - P) <= 1e-12 * P * max(1, n), trial
        assert np.linalg.norm(prod - ref) / max(np.linalg.norm(ref), 1e-300) < 1e-12, trial
        again, _ = _project(prod, P)
        assert np.linalg.norm(again - prod) <= 1e-12 * max(np.linalg.norm(prod), 1e-300)
        # FP32: the projection of the rounded input, to FP32 accuracy
        y32 = y.astype(np.float32).astype(np.float64)
        p32, _ = _project(y, P, precision=32)
        ref32 = O.sorting_projection(y32, P)
        assert p32.min() >= 0
        assert np.linalg.norm(p32 - ref32) <= 1e-6 * max(np.linalg.norm(ref32), 1e-300) + \
            1e-6 * scale, trial


def test_device_projection_large(hc):
    """4096^2 values shaped like a pressure iterate after a step (many
    negatives, a heavy tail): equals the sorting projector, threshold agrees
    with the oracle's Michelot iteration."""
    n = 4096 * 4096
    rng = np.random.default_rng(3)
    y = rng.standard_normal(n) * 0.004 + 0.001
    y[rng.random(n) < 0.01] += 0.05
    P = 0.005 * n
    prod, th = _project(y, P)
    ref = O.sorting_projection(y, P)
    assert np.linalg.norm(prod - ref) / np.linalg.norm(ref) < 1e-12
    assert abs(prod.sum() - P) <= 1e-12 * P * 64
    yo = y.copy()
    th_o = O.project_load_simplex(yo, P)
    assert abs(th - th_o) <= 1e-12 * abs(th_o)


@pytest.fixture
def fault(hc):
    from aspher_amd._capi import lib, check

    def set_(mode, after=0):
        check(lib().aspher_test_fault(int(mode), int(after)))
    yield set_
    check(lib().aspher_test_fault(0, 0))


def _rough(hc, Ns, seed=0, bump=False):
    g = hc.surfaces.rough_gap(Ns, H=0.8, seed=seed)
    if bump:
        # element 0, where the fault hook writes, far out of contact: a NaN
        # there in the line-search product is then outside the contact sums
        # and surfaces at the next fresh evaluation (0 * NaN in the FW gap)
        g[0] = g.max() + np.ptp(g)
    return g


@pytest.mark.parametrize("backend", ["fft", "periodic"])
def test_nonfinite_operator_single_solve(hc, fault, backend):
    """T03: after two good applies (the first iteration's) every operator
    output carries a NaN; the next fresh evaluation sees it."""
    Ns = 64
    g0 = _rough(hc, Ns, bump=True)
    s = hc.ContactSolver(Ns, backend=backend)
    ok = s.solve(g0, 0.005, tol=1e-10, max_iter=100)
    fault(1, 2)
    r = s.solve(g0, 0.005, tol=1e-10, max_iter=100)
    print(backend, ok.status, r.status, r.iterations, flush=True)
    assert not r.converged
    assert r.status == "nonfinite"
    assert r.iterations <= 2
    fault(0)
    again = s.solve(g0, 0.005, tol=1e-10, max_iter=100)
    assert again.status == ok.status and again.iterations == ok.iterations
    np.testing.assert_array_equal(again.pressure, ok.pressure)


def test_nonfinite_operator_fp32_and_nested(hc, fault):
    Ns = 256
    g0 = _rough(hc, Ns, bump=True)
    fault(1, 0)
    for precision in ("double", "float", "float_then_double"):
        r = hc.solve_nested(Ns, g0, 0.005, backend="fft", tol=1e-8, coarsest=64,
                            precision=precision, max_iter=200)
        assert r.status == "nonfinite" and not r.converged, precision
        assert r.stage_stats[0]["status"] == "nonfinite"
    # the fault armed after the coarse levels: they converge, the finest does not
    fault(0)
    ref = hc.solve_nested(Ns, g0, 0.005, backend="fft", tol=1e-8, coarsest=64,
                          precision="double")
    # matvec_count leaves out the dropped speculative applies (one or two per
    # level), which the device counter sees: a margin of 12
    coarse_mv = sum(st["matvec_count"] for st in ref.stage_stats[:-1])
    assert ref.stage_stats[-1]["matvec_count"] > 30
    fault(1, coarse_mv + 12)
    r = hc.solve_nested(Ns, g0, 0.005, backend="fft", tol=1e-8, coarsest=64,
                        precision="double")
    st = [x["status"] for x in r.stage_stats]
    print(st, flush=True)
    assert st[:-1] == ["converged"] * (len(st) - 1)
    assert st[-1] == "nonfinite" and r.status == "nonfinite" and not r.converged


def test_nonfinite_operator_ensemble(hc, fault):
    """Element 0 of the first instance of each size group turns NaN after a
    few replayed iterations: those instances leave the device loop and end
    nonfinite; the other instances converge as without the fault."""
    spec = [("fft", 64, 0.005), ("fft", 64, 0.02), ("periodic", 128, 0.02),
            ("periodic", 128, 0.005), ("fft", 128, 0.001)]
    probs = [dict(backend=b, grid_size=n, p_nominal=p, gap=_rough(hc, n, seed=40 + k, bump=True))
             for k, (b, n, p) in enumerate(spec)]
    clean, _ = hc.solve_ensemble(probs, tol=1e-10, precision="double")
    fault(1, 20)
    res, st = hc.solve_ensemble(probs, tol=1e-10, precision="double")
    assert st["graph"]
    first = {}
    for k, (b, n, _) in enumerate(spec):
        first.setdefault((b, n), k)
    for k, (r, c) in enumerate(zip(res, clean)):
        print(k, r.status, r.status_reason, flush=True)
        if k in first.values():
            assert r.status == "nonfinite" and not r.converged, k
        else:
            assert r.status == "converged", k
            np.testing.assert_array_equal(r.pressure, c.pressure)


def test_verification_failed_on_disagreeing_final_check(hc, fault):
    """A converged loop whose fresh final evaluation disagrees must not report
    success (src/contact_solver.cpp:529-537)."""
    Ns = 64
    g0 = _rough(hc, Ns)
    s = hc.ContactSolver(Ns, backend="fft")
    fault(2)
    r = s.solve(g0, 0.005, tol=1e-8, precond="fourier")
    assert r.status == "verification_failed" and not r.converged
    assert r.status_reason == "final_check"
    assert r.penetration_error > 1.0
    fault(0)
    assert s.solve(g0, 0.005, tol=1e-8, precond="fourier").status == "converged"


def test_nested_precision_policy_single_precision_flag(hc):
    """precision '' and 'double' defer to single_precision
    (python/bindings.cpp:404-411, src/nested_solve.cpp:584-587)."""
    Ns = 128
    g0 = _rough(hc, Ns)
    for prec in ("", "double"):
        r = hc.solve_nested(Ns, g0, 0.005, backend="fft", tol=1e-5, coarsest=64,
                            precision=prec, single_precision=True)
        assert [x["precision"] for x in r.stage_stats] == ["float", "float"], prec
        r = hc.solve_nested(Ns, g0, 0.005, backend="fft", tol=1e-5, coarsest=64,
                            precision=prec, single_precision=False)
        assert [x["precision"] for x in r.stage_stats] == ["double", "double"], prec
    r = hc.solve_nested(Ns, g0, 0.005, backend="fft", tol=1e-7, coarsest=64,
                        precision="float_then_double", single_precision=True)
    assert [x["name"] for x in r.stage_stats] == ["coarse:64", "finest:128", "polish:128"]
    with pytest.raises(ValueError, match="defer to single_precision"):
        hc.solve_nested(Ns, g0, 0.005, backend="fft", precision="half")
