"""GPU parity of the certified projected CG and the nested driver, through the
C ABI, against the CPU oracle, exact solutions and the reference's own
solver-level tests."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def hc():
    import aspher_amd
    assert aspher_amd.device_count() >= 1, "no GPU visible"
    return aspher_amd


def _parabolic(ns):
    x = (np.arange(ns) + 0.5) / ns - 0.5
    return (x[:, None] ** 2 + x[None, :] ** 2).ravel()


def _rough(ns, seed=7):
    """tests/test_certification_py.py:19-30."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(ns, 1.0 / ns)
    kx, ky = np.meshgrid(k, k, indexing="ij")
    kk = np.hypot(kx, ky)
    kk[0, 0] = 1.0
    amp = np.where((kk >= 2) & (kk <= 16), kk ** (-1.8), 0.0)
    h = np.real(np.fft.ifft2(amp * np.exp(1j * rng.uniform(0, 2 * np.pi, (ns, ns)))))
    h = 0.02 * h / h.std()
    g = (-h).ravel().copy()
    return g - g.min()


def _periodic_rough(ns, seed=3, rms=0.01):
    """tests/test_periodic_py.py:17-25."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(ns, 1.0 / ns)
    kk = np.hypot(k[:, None], k[None, :])
    kk[0, 0] = 1.0
    amp = np.where((kk >= 2) & (kk <= ns // 8), kk ** (-1.8), 0.0)
    h = np.real(np.fft.ifft2(amp * np.exp(2j * np.pi * rng.random((ns, ns)))))
    h *= rms / h.std()
    return np.ascontiguousarray((h.max() - h).ravel())


def test_manufactured_kkt_love_operator(hc):
    """T01 construction (tests/contact_oracle.hpp:156-166) on the Love
    operator: p* and g* complementary, g0 = alpha 1 - S p* + g*; the solver
    must return p* (it is the unique minimiser)."""
    Ns = 8
    S = O.dense_matrix(Ns)
    N = Ns * Ns
    rng = np.random.default_rng(11)
    c = (np.arange(Ns) - 3.5) ** 2
    support = (c[:, None] + c[None, :]).ravel() < 6.0
    p_star = np.where(support, 0.5 + rng.random(N), 0.0)
    g_star = np.where(support, 0.0, 0.01 + 0.05 * rng.random(N))
    g0 = 0.25 - S @ p_star + g_star
    P = p_star.sum()
    for be in ("dense", "fft"):
        s = hc.ContactSolver(Ns, backend=be)
        r = s.solve(g0, P / N, tol=1e-13, max_iter=20000)
        assert r.status == "converged", (be, r.status)
        p = r.pressure.ravel()
        assert np.linalg.norm(p - p_star) / np.linalg.norm(p_star) < 1e-9, be
        d = O.diagnose(lambda v: S @ v, g0, p, P, np.ptp(g0))
        assert d["fw_error"] <= 1e-12 and d["penetration"] <= 1e-12
        assert abs(r.approach - 0.25) < 1e-10


def test_hertz_fft_vs_dense_and_oracle(hc):
    """tests/test_fft.cpp:63-107 (Ns=64): same contact set, pressure to 1e-6
    at tol 1e-8; and the oracle's solve at the same tolerance."""
    Ns, R, pbar = 64, 0.5, 0.05
    h = 1.0 / Ns
    c = (np.arange(Ns) + 0.5) * h - 0.5
    g0 = ((c[:, None] ** 2 + c[None, :] ** 2) / (2 * R)).ravel()
    rd = hc.ContactSolver(Ns, backend="dense").solve(g0, pbar, tol=1e-8, max_iter=2000)
    rf = hc.ContactSolver(Ns, backend="fft").solve(g0, pbar, tol=1e-8, max_iter=2000)
    assert rd.converged and rf.converged
    assert abs(rd.iterations - rf.iterations) <= 1
    assert rd.contact_area == rf.contact_area
    pd, pf = rd.pressure.ravel(), rf.pressure.ravel()
    assert np.linalg.norm(pf - pd) / np.linalg.norm(pd) < 1e-6
    ro = O.solve_contact(O.FreeSpaceOperator(Ns).matvec, g0, pbar,
                         O.SolveOptions(tol=1e-8, max_iter=2000))
    assert abs(ro.iterations - rf.iterations) <= 1
    assert ro.contact_fraction == rf.contact_area
    assert np.linalg.norm(pf - ro.pressure) / np.linalg.norm(ro.pressure) < 1e-6


def test_config1_free_space_256_parity_1e10(hc):
    """Ns=256 self-affine surface (H=0.8, k in [4,64]/Ns, rms 0.002), Fourier
    (stencil) preconditioner: at tol 1e-13 both sides sit on the unique
    solution, pressure rel L2 <= 1e-10 and identical contact sets."""
    Ns = 256
    g0 = hc.surfaces.rough_gap(Ns, seed=0)
    s = hc.ContactSolver(Ns, backend="fft")
    r = s.solve(g0, 0.005, tol=1e-13, max_iter=20000, precond="fourier")
    assert r.status == "converged"
    op = O.FreeSpaceOperator(Ns)
    sp = O.StencilPreconditioner(Ns, 2)
    ro = O.solve_contact(op.matvec, g0, 0.005, O.SolveOptions(tol=1e-13, max_iter=20000),
                         sp.apply)
    assert ro.status == "converged"
    p, po = r.pressure.ravel(), ro.pressure
    assert np.linalg.norm(p - po) / np.linalg.norm(po) <= 1e-10
    # contact sets agree except where p sits at the threshold
    diff = (p > 0) != (po > 0)
    assert np.all(np.maximum(p, po)[diff] < 1e-12 * po.max())
    # independent FP64 certificate with the oracle's operator
    d = O.diagnose(op.matvec, g0, p, 0.005 * Ns * Ns, np.ptp(g0))
    assert d["fw_error"] <= 1e-12 and d["penetration"] <= 1e-12 and d["load_error"] <= 1e-12
    # the tol 1e-8 solve: converged, certified by the oracle operator
    r8 = s.solve(g0, 0.005, tol=1e-8, precond="fourier")
    assert r8.converged
    d8 = O.diagnose(op.matvec, g0, r8.pressure.ravel(), 0.005 * Ns * Ns, np.ptp(g0))
    assert d8["fw_error"] <= 1.01e-8 and d8["penetration"] <= 1.01e-8


def test_westergaard_full_and_partial(hc):
    """tests/test_periodic.cpp:94-129 (n=256)."""
    n, L, Es, D = 256, 2.0, 1.3, 0.01
    pstar = math.pi * Es * D / L
    s = hc.ContactSolver(n, domain_size=L, E_star=Es, backend="periodic")
    cosx = np.tile(np.cos(2 * math.pi * np.arange(n) / n), (n, 1)).ravel()
    g0 = D * (1 - cosx)
    pbar = 1.5 * pstar
    r = s.solve(g0, pbar, tol=1e-12)
    assert r.converged
    pex = pbar + pstar * cosx
    assert np.linalg.norm(r.pressure.ravel() - pex) / np.linalg.norm(pex) < 1e-9
    pbar = 0.3 * pstar
    r = s.solve(g0, pbar, tol=1e-10)
    assert r.converged
    width = (r.pressure.ravel()[:n] > 0).sum() * L / n
    wex = 2 * (L / math.pi) * math.asin(math.sqrt(pbar / pstar))
    assert abs(width - wex) <= 2 * L / n


def test_status_paths_and_input_rejection(hc):
    """T03 (tests/test_certification.cpp:215-323) through the API."""
    Ns = 16
    g0 = _parabolic(Ns)
    s = hc.ContactSolver(Ns, backend="dense")
    for mi in (0, 1):
        r = s.solve(g0, 0.02, tol=1e-10, max_iter=mi)
        assert r.status == "max_iterations" and not r.converged
    bad = g0.copy()
    bad[3] = np.inf
    with pytest.raises(ValueError):
        s.solve(bad, 0.02, tol=1e-10)
    with pytest.raises(ValueError):
        s.solve(g0, -1.0)
    with pytest.raises(ValueError):
        s.solve(g0, 0.02, tol=0.0)
    with pytest.raises(ValueError):
        s.solve(g0, 0.02, max_iter=-3)
    with pytest.raises(ValueError):
        s.solve(g0, 0.02, precond="bogus")
    with pytest.raises(ValueError):
        s.solve(g0[:-1], 0.02)


def test_singleton_warm_start_repaired(hc):
    """tests/test_certification_py.py:47-75 (dense backend, Ns=8)."""
    ns = 8
    g0 = _parabolic(ns)
    op = hc.ContactSolver(grid_size=ns, backend="dense")
    warm = np.zeros(ns * ns)
    warm[0] = 0.01 * ns * ns
    cold = op.solve(g0, 0.01, tol=1e-10)
    got = op.solve(g0, 0.01, tol=1e-10, p_init=warm)
    assert got.status == "converged" and got.converged is True and got.iterations > 0
    S = O.dense_matrix(ns)
    d = O.diagnose(lambda v: S @ v, g0, got.pressure.ravel(), 0.01 * ns * ns, np.ptp(g0))
    assert d["fw_error"] <= 1e-10 and d["penetration"] <= 1e-10 and d["min_pressure"] >= 0
    rel = np.linalg.norm(got.pressure - cold.pressure) / np.linalg.norm(cold.pressure)
    assert rel < 1e-6


def test_status_fields_consistent(hc):
    """tests/test_certification_py.py:78-105."""
    ns = 16
    g0 = _parabolic(ns)
    op = hc.ContactSolver(grid_size=ns, backend="dense")
    r = op.solve(g0, 0.02, tol=1e-10)
    assert r.status == "converged" and r.converged
    assert r.fw_error <= r.effective_tol and r.penetration_error <= r.effective_tol
    assert r.load_error < 1e-12
    assert r.requested_tol == 1e-10 and r.effective_tol == 1e-10
    assert r.matvec_count > 0
    assert r.validation_scope == "solve_operator"
    assert r.operator_error_kind == "unavailable" and r.operator_error is None
    bad = op.solve(g0, 0.02, tol=1e-25, max_iter=300)
    if not bad.converged:
        assert bad.status in ("stagnated", "max_iterations")


def test_nested_periodic_matches_single_level(hc):
    """tests/test_periodic_py.py:53-67."""
    ns = 256
    g = _periodic_rough(ns)
    pbar = 0.01
    s = hc.ContactSolver(ns, domain_size=1.0, E_star=1.0, backend="periodic")
    ref = s.solve(g, pbar, tol=1e-12, precond="fourier-fft").pressure.ravel()
    for prec, tol, bound in (("double", 1e-10, 1e-7), ("float_then_double", 1e-10, 1e-7),
                             ("float", 2e-6, 1e-4)):
        r = hc.solve_nested(grid_size=ns, gap=g, p_nominal=pbar, coarsest=32,
                            backend="periodic", tol=tol, precision=prec)
        assert r.status == "converged", (prec, r.status, r.status_reason)
        p = r.pressure.ravel()
        assert abs(p.mean() - pbar) < 1e-6 * pbar
        assert np.linalg.norm(p - ref) / np.linalg.norm(ref) < bound, prec


def test_nested_auto_engine_is_fft_on_periodic(hc):
    """tests/test_periodic_py.py:70-84."""
    ns = 256
    g = _periodic_rough(ns)
    runs = {e: hc.solve_nested(grid_size=ns, gap=g, p_nominal=0.02, coarsest=32,
                               backend="periodic", tol=1e-10, precond_engine=e)
            for e in ("auto", "fft", "stencil")}
    assert runs["auto"].iterations == runs["fft"].iterations
    assert runs["auto"].matvec_count == runs["fft"].matvec_count
    np.testing.assert_array_equal(runs["auto"].pressure, runs["fft"].pressure)
    assert runs["stencil"].iterations > runs["fft"].iterations


def test_nested_datum_invariance_fft(hc):
    """tests/test_certification_py.py:108-121."""
    ns = 128
    g0 = _rough(ns)
    ref = hc.solve_nested(ns, g0, 0.02, coarsest=32, backend="fft", tol=1e-10)
    assert ref.status == "converged"
    pref = ref.pressure.ravel()
    for shift in (0.0, 1e3, 1e6):
        for single in (False, True):
            r = hc.solve_nested(ns, g0 + shift, 0.02, coarsest=32, backend="fft",
                                single_precision=single, allow_tolerance_relaxation=single,
                                tol=1e-8)
            rel = np.linalg.norm(r.pressure.ravel() - pref) / np.linalg.norm(pref)
            assert r.status == "converged"
            assert abs(r.approach - ref.approach - shift) < 1e-6 * max(1.0, shift)
            assert rel < (1e-4 if single else 1e-5), (shift, single, rel)


def test_nested_precision_stages_fft(hc):
    """T17 (tests/test_contracts.cpp:372-455) on the fft backend."""
    Ns = 64
    x = (np.arange(Ns) + 0.5) / Ns - 0.5
    X, Y = np.meshgrid(x, x, indexing="xy")
    g0 = (0.5 * (X ** 2 + Y ** 2) + 0.002 * np.cos(12 * np.pi * X) * np.cos(10 * np.pi * Y)).ravel()
    tol = 1e-9
    rd = hc.solve_nested(Ns, g0, 0.02, coarsest=32, backend="fft", tol=tol)
    assert rd.converged
    rf = hc.solve_nested(Ns, g0, 0.02, coarsest=32, backend="fft", tol=tol, precision="float")
    assert not rf.converged and rf.status == "stagnated" and rf.status_reason == "precision_limit"
    assert rf.requested_tol < rf.effective_tol
    rfr = hc.solve_nested(Ns, g0, 0.02, coarsest=32, backend="fft", tol=tol,
                          precision="float", allow_tolerance_relaxation=True)
    assert rfr.converged and rfr.effective_tol > rfr.requested_tol
    rp = hc.solve_nested(Ns, g0, 0.02, coarsest=32, backend="fft", tol=tol,
                         precision="float_then_double")
    rel_f = np.linalg.norm(rfr.pressure - rd.pressure) / np.linalg.norm(rd.pressure)
    rel_p = np.linalg.norm(rp.pressure - rd.pressure) / np.linalg.norm(rd.pressure)
    assert rp.converged and rp.fw_error <= tol and rp.penetration_error <= tol
    assert rp.effective_tol == rp.requested_tol and rel_p < rel_f
    st = rp.stage_stats
    assert len(st) == len(rd.stage_stats) + 1
    assert st[-2]["precision"] == "float" and st[-2]["name"].startswith("finest:")
    assert st[-1]["precision"] == "double" and st[-1]["name"].startswith("polish:")
    assert st[-2]["effective_tol"] == 2e-6 and st[-1]["effective_tol"] == tol
    assert rp.iterations == st[-2]["iterations"] + st[-1]["iterations"]


def test_nested_rejects_out_of_contract(hc):
    """tests/test_contracts.cpp:66-67, tests/test_periodic_py.py:98-102."""
    g = _periodic_rough(128)
    with pytest.raises(ValueError):
        hc.solve_nested(grid_size=128, gap=g, p_nominal=0.01, coarsest=32, backend="periodic",
                        active_set=True)
    with pytest.raises(ValueError):
        hc.solve_nested(grid_size=128, gap=g, p_nominal=0.01, coarsest=48, backend="fft")
    with pytest.raises(ValueError):
        hc.solve_nested(grid_size=128, gap=g, p_nominal=0.01, coarsest=32, backend="nope")
    with pytest.raises(ValueError):
        hc.solve_nested(grid_size=128, gap=g, p_nominal=0.0, coarsest=32, backend="fft")


@pytest.mark.parametrize("backend,precision,tol,max_iter", [
    ("fft", "double", 1e-10, 20000),
    ("fft", "float", 1e-6, 20000),
    ("periodic", "double", 1e-25, 300),     # stagnation: stall, forced identification
    ("fft", "double", 1e-10, 7),            # budget ends right after a CG step
])
def test_overlapped_cg_step_matches_synchronous(hc, backend, precision, tol, max_iter):
    """The solver enqueues each CG step before the host reads the iteration's
    diagnostics and drops it on the other branches. An opt-in cost gate that
    never fires (1e9) runs the same iteration with the host synchronised
    before each step: every result must be bit-identical."""
    Ns = 128
    g = hc.surfaces.rough_gap(Ns, H=0.8, seed=3) if backend == "fft" else _periodic_rough(Ns)
    kw = dict(grid_size=Ns, gap=g, p_nominal=0.01, coarsest=32, tol=tol, max_iter=max_iter,
              backend=backend, precision=precision, allow_tolerance_relaxation=True)
    a = hc.solve_nested(**kw)
    b = hc.solve_nested(precond_cost_gate=1e9, **kw)
    assert a.status == b.status and a.status_reason == b.status_reason
    assert a.iterations == b.iterations and a.matvec_count == b.matvec_count
    assert a.precond_count == b.precond_count
    assert a.identification_steps == b.identification_steps
    assert not b.precond_dropped
    np.testing.assert_array_equal(a.pressure, b.pressure)
    np.testing.assert_array_equal(a.gap, b.gap)
    assert [s["iterations"] for s in a.stage_stats] == [s["iterations"] for s in b.stage_stats]


@pytest.mark.parametrize("backend,precision,tol,max_iter,Ns,coarsest", [
    ("fft", "double", 1e-10, 20000, 256, 32),
    ("fft", "float_then_double", 1e-8, 20000, 256, 32),
    ("periodic", "float", 1e-6, 20000, 256, 32),
    ("periodic", "double", 1e-25, 300, 256, 32),   # stagnation: stall, forced identification
    ("fft", "double", 1e-10, 20000, 96, 24),       # generic (non power-of-two) plans, ungated applies
])
def test_graph_replay_matches_launched_kernels(hc, backend, precision, tol, max_iter, Ns,
                                               coarsest):
    """Levels of at most graph_max_points points replay each CG iteration's
    two enqueue sequences from captured hipGraphs; each replay runs between
    the solver's own eager work and host synchronisations (identification
    steps, the final evaluation). With graph_max_points = 0 every kernel is
    launched: every result must be bit-identical, run after run."""
    import ctypes as C
    from aspher_amd import api, _capi
    g = hc.surfaces.rough_gap(Ns, H=0.8, seed=5) if backend == "fft" else _periodic_rough(Ns)
    kw = dict(coarsest=coarsest, q=6, leaf_side=8, precond=True, precond_engine="auto",
              precond_radius=2, precond_occupancy_max=0.4, precond_cost_gate=0.0,
              coarse_tol=1e-4, single_precision=False, precision=precision, float_floor=2e-6,
              allow_tolerance_relaxation=True, light_result=False, backend=backend,
              record_error_history=False, active_set=False, active_delta=0.05, active_halo=2,
              active_max_rounds=5, active_all_levels=True, active_occupancy_max=0.4,
              reuse_operators=False)

    def run(graph_points):
        p = api._nested_params(**kw)
        p.graph_max_points = graph_points
        gg = np.ascontiguousarray(g, dtype=np.float64).ravel()
        out = C.c_void_p()
        _capi.check(_capi.lib().aspher_solve_nested(
            Ns, gg.ctypes.data_as(C.POINTER(C.c_double)), 0.01, 1.0, 1.0, tol, max_iter, 1,
            C.byref(p), C.byref(out)))
        return api.ContactResult(out.value)

    ref = run(0)
    for _ in range(2):
        a = run(Ns * Ns)
        assert a.status == ref.status and a.status_reason == ref.status_reason
        assert a.iterations == ref.iterations and a.matvec_count == ref.matvec_count
        assert a.precond_count == ref.precond_count
        assert a.identification_steps == ref.identification_steps
        np.testing.assert_array_equal(a.pressure, ref.pressure)
        np.testing.assert_array_equal(a.gap, ref.gap)
        assert [s["iterations"] for s in a.stage_stats] == \
            [s["iterations"] for s in ref.stage_stats]
