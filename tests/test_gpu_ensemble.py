"""Ensemble solver (many independent certified solves advanced together,
one HIP graph per iteration): every instance must agree with the single
solve of the same problem (ContactSolver(backend).solve,
python/bindings.cpp:186-234) and carry a certificate that the oracle's
independent FP64 operator confirms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def hc():
    import aspher_amd
    assert aspher_amd.device_count() >= 1, "no GPU visible"
    O.WORKERS = 8
    return aspher_amd


def _problems(hc, spec, seed0=0):
    out = []
    for k, (backend, Ns, pbar) in enumerate(spec):
        g = hc.surfaces.rough_gap(Ns, H=0.6 + 0.1 * (k % 4), seed=seed0 + k)
        out.append(dict(backend=backend, grid_size=Ns, gap=g, p_nominal=pbar))
    return out


def _oracle_op(backend, Ns):
    return O.FreeSpaceOperator(Ns) if backend == "fft" else O.PeriodicOperator(Ns, 1.0, 1.0)


SPEC = [("fft", 64, 0.005), ("periodic", 128, 0.02), ("fft", 128, 0.001),
        ("periodic", 64, 0.01), ("fft", 64, 0.02), ("periodic", 128, 0.005)]


def test_ensemble_fp64_matches_single_solves(hc):
    probs = _problems(hc, SPEC)
    res, st = hc.solve_ensemble(probs, tol=1e-12, precision="double", max_iter=20000)
    assert st["graph"] and st["groups"] == 4
    for pr, r in zip(probs, res):
        Ns, b = pr["grid_size"], pr["backend"]
        print(b, Ns, r.status, r.iterations, flush=True)
        assert r.status == "converged", (b, Ns, r.status, r.status_reason)
        s = hc.ContactSolver(Ns, backend=b)
        ref = s.solve(pr["gap"], pr["p_nominal"], tol=1e-12, max_iter=20000,
                      precond="fourier-fft")
        assert ref.converged
        p, q = r.pressure.ravel(), ref.pressure.ravel()
        assert np.linalg.norm(p - q) / np.linalg.norm(q) < 1e-8
        assert abs(r.approach - ref.approach) < 1e-9 * max(1.0, abs(ref.approach))
        assert abs(r.objective - ref.objective) < 1e-9 * abs(ref.objective)
        d = O.diagnose(_oracle_op(b, Ns).matvec, pr["gap"], p, pr["p_nominal"] * Ns * Ns,
                       np.ptp(pr["gap"]))
        assert d["fw_error"] <= 1.01e-12 and d["penetration"] <= 1.01e-12, d
        assert d["load_error"] <= 1e-12
        assert r.gap.min() >= -1.01e-12 * np.ptp(pr["gap"])
        assert r.stage_stats[0]["name"] == "ensemble:%d" % Ns


@pytest.mark.parametrize("precond", ["auto", "fourier"])
def test_ensemble_graph_replays_equal_eager(hc, precond):
    """The captured iteration replayed must give bit-identical results to the
    same sequence launched eagerly (graph replays with host polls between),
    with the Fourier preconditioner's batched convolutions ("auto") and with
    the stencil's tile work queue ("fourier")."""
    probs = _problems(hc, SPEC, seed0=10)
    ra, sa = hc.solve_ensemble(probs, tol=1e-9, precision="double", use_graph=True,
                               check_every=3, precond=precond)
    rb, sb = hc.solve_ensemble(probs, tol=1e-9, precision="double", use_graph=False,
                               check_every=5, precond=precond)
    assert sa["graph"] and not sb["graph"]
    for a, b in zip(ra, rb):
        assert a.status == b.status == "converged"
        assert a.iterations == b.iterations
        np.testing.assert_array_equal(a.pressure, b.pressure)


def test_ensemble_fp32_certified_by_oracle(hc):
    probs = _problems(hc, SPEC, seed0=20)
    res, st = hc.solve_ensemble(probs, tol=1e-5, precision="float")
    for pr, r in zip(probs, res):
        Ns, b = pr["grid_size"], pr["backend"]
        assert r.status == "converged", (b, Ns, r.status, r.status_reason)
        assert r.effective_tol == 1e-5
        d = O.diagnose(_oracle_op(b, Ns).matvec, pr["gap"], r.pressure.ravel(),
                       pr["p_nominal"] * Ns * Ns, np.ptp(pr["gap"]))
        assert d["fw_error"] <= 2e-5 and d["penetration"] <= 2e-5, (b, Ns, d)
        ref = hc.ContactSolver(Ns, backend=b).solve(pr["gap"], pr["p_nominal"], tol=1e-10,
                                                    precond="fourier")
        assert abs(r.contact_area - ref.contact_area) <= 2e-3


def test_ensemble_host_fallback_on_budget(hc):
    probs = _problems(hc, SPEC[:3], seed0=30)
    res, st = hc.solve_ensemble(probs, tol=1e-12, precision="double", max_iter=4)
    assert st["host_fallbacks"] == 3
    for r in res:
        assert r.status == "max_iterations"
        assert r.status_reason.startswith("ensemble:max_iterations")
        assert r.iterations == 4


def test_ensemble_rejects_bad_input(hc):
    g = np.zeros(48 * 48)
    with pytest.raises(ValueError):
        hc.solve_ensemble([dict(backend="fft", grid_size=48, gap=g, p_nominal=0.01)])
    with pytest.raises(ValueError):
        hc.solve_ensemble([dict(backend="h2", grid_size=64, gap=np.zeros(4096),
                                p_nominal=0.01)])
    with pytest.raises(ValueError):
        hc.solve_ensemble([dict(backend="fft", grid_size=64, gap=np.zeros(4096),
                                p_nominal=0.0)])
    bad = np.zeros(4096)
    bad[5] = np.nan
    with pytest.raises(ValueError):
        hc.solve_ensemble([dict(backend="fft", grid_size=64, gap=bad, p_nominal=0.01)])


def test_ensemble_uneven_sizes_c4_subset(hc):
    """Sizes {256, 512} of the batched workload (Hurst in [0.6, 0.9], p_nominal
    log-uniform in [0.001, 0.02], free/periodic), FP32 with FP64 reductions."""
    rng = np.random.default_rng(4)
    probs = []
    for k in range(12):
        Ns = int(rng.choice([256, 512]))
        H = float(rng.uniform(0.6, 0.9))
        pbar = float(np.exp(rng.uniform(np.log(0.001), np.log(0.02))))
        b = "fft" if rng.random() < 0.5 else "periodic"
        probs.append(dict(backend=b, grid_size=Ns, p_nominal=pbar,
                          gap=hc.surfaces.rough_gap(Ns, H=H, seed=100 + k)))
    res, st = hc.solve_ensemble(probs, tol=1e-5)
    print(st, flush=True)
    for pr, r in zip(probs, res):
        assert r.status == "converged", (pr["backend"], pr["grid_size"], r.status_reason)
        assert r.fw_error <= 1e-5 and r.penetration_error <= 1e-5
        assert abs(r.mean_pressure - pr["p_nominal"]) <= 1e-5 * pr["p_nominal"]


def test_ensemble_c4_workload_certified_by_oracle(hc):
    """The batched workload at its real sizes (64 instances drawn from
    {256, 512, 1024, 2048}^2, Hurst in [0.6, 0.9], p_nominal log-uniform in
    [0.001, 0.02], free/periodic; FP32 with FP64 reductions, tol 1e-5): every
    returned pressure is certified by the oracle's FP64 operator at
    1.01 max(tol, 2e-6), and has the load."""
    import bench
    probs = bench.ensemble_problems(hc)
    tol = 1e-5
    res, st = hc.solve_ensemble(probs, tol=tol)
    print(st, flush=True)
    assert {p["grid_size"] for p in probs} == {256, 512, 1024, 2048}
    assert st["graph"] and st["host_fallbacks"] == 0
    ops = {}
    worst = {}
    for pr, r in zip(probs, res):
        key = (pr["backend"], pr["grid_size"])
        Ns = pr["grid_size"]
        assert r.status == "converged", (key, r.status, r.status_reason)
        if key not in ops:
            ops[key] = _oracle_op(*key)
        d = O.diagnose(ops[key].matvec, pr["gap"], r.pressure.ravel(), pr["p_nominal"] * Ns * Ns,
                       np.ptp(pr["gap"]))
        w = worst.setdefault(key, [0.0, 0.0])
        w[0] = max(w[0], d["fw_error"])
        w[1] = max(w[1], d["penetration"])
        gate = 1.01 * max(tol, 2e-6)
        assert d["fw_error"] <= gate and d["penetration"] <= gate, (key, d)
        assert d["load_error"] <= 1e-6 and d["min_pressure"] >= 0.0
    for k in sorted(worst):
        print("oracle certificate %s %d: fw %.3e pen %.3e" % (k + tuple(worst[k])), flush=True)


def test_ensemble_fp64_1024_group_matches_single_solves(hc):
    """FP64 ensemble of 1024^2 instances of the batched workload (both kinds)
    against the single solves of the same problems and the oracle."""
    import bench
    probs = [p for p in bench.ensemble_problems(hc) if p["grid_size"] == 1024]
    pick = [p for p in probs if p["backend"] == "fft"][:2] + \
        [p for p in probs if p["backend"] == "periodic"][:1]
    assert len(pick) == 3
    tol = 1e-10
    res, st = hc.solve_ensemble(pick, tol=tol, precision="double")
    assert st["graph"]
    for pr, r in zip(pick, res):
        b, Ns = pr["backend"], pr["grid_size"]
        assert r.status == "converged", (b, r.status, r.status_reason)
        ref = hc.ContactSolver(Ns, backend=b).solve(
            pr["gap"], pr["p_nominal"], tol=tol, max_iter=20000, precond="fourier-fft")
        assert ref.converged
        p, q = r.pressure.ravel(), ref.pressure.ravel()
        rel = np.linalg.norm(p - q) / np.linalg.norm(q)
        print(b, Ns, r.iterations, ref.iterations, "rel %.3e" % rel, flush=True)
        assert rel < 1e-6
        assert abs(r.contact_area - ref.contact_area) <= 1e-4
        d = O.diagnose(_oracle_op(b, Ns).matvec, pr["gap"], p, pr["p_nominal"] * Ns * Ns,
                       np.ptp(pr["gap"]))
        assert d["fw_error"] <= 1.01 * tol and d["penetration"] <= 1.01 * tol, d
