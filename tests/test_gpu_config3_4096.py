"""The headline workload at its real size: free-space rough contact at
Ns = 4096 (H = 0.8, rms 0.002, |k| in [4, 64], p_nominal 0.005, tol 1e-6,
nested from 64), mixed precision (float_then_double) and FP64, each with the
active set (the benchmark configuration) and on the full grid, and mixed
with the full-grid Fourier preconditioner on every level
(precond_engine="fft", src/nested_solve.cpp:711-723, where the automatic
choice takes the stencil at this contact fraction). Each returned pressure
is certified by the oracle's independent FP64 operator (scipy.fft of the C
Love table), and every mixed result must match an FP64 one: contact area
within 1e-4, pressure and mean gap close. Two more surfaces (seeds 1, 2) run
the benchmark pair (active set, mixed and FP64), certified the same way."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

NS, PBAR, TOL = 4096, 0.005, 1e-6


@pytest.fixture(scope="module")
def oracle_op():
    O.WORKERS = 16
    return O.FreeSpaceOperator(NS)


def _gap(hc, seed):
    return hc.surfaces.rough_gap(NS, H=0.8, k_low=4, k_high=64, rms=0.002, seed=seed)


def _kw(g):
    return dict(grid_size=NS, gap=g, p_nominal=PBAR, coarsest=64, tol=TOL, backend="fft",
                reuse_operators=True)


def _certify(op, g, r):
    assert r.status == "converged"
    assert r.fw_error <= TOL and r.penetration_error <= TOL
    g_ref = max(np.ptp(g), PBAR)
    d = O.diagnose(op.matvec, g, r.pressure.ravel(), PBAR * NS * NS, g_ref)
    assert d["fw_error"] <= 1.01 * TOL and d["penetration"] <= 1.01 * TOL, d
    assert d["load_error"] <= 1e-12 and d["min_pressure"] >= 0.0
    assert abs(d["n_contact"] / NS ** 2 - r.contact_area) <= 1.0 / NS ** 2


def _close(r, d):
    assert abs(r.contact_area - d.contact_area) <= 1e-4
    p, q = r.pressure.ravel(), d.pressure.ravel()
    assert np.linalg.norm(p - q) / np.linalg.norm(q) <= 1e-4
    assert abs(r.gap.mean() - d.gap.mean()) <= 1e-6 * np.ptp(d.gap)


@pytest.fixture(scope="module")
def c3(oracle_op):
    import aspher_amd as hc
    assert hc.device_count() >= 1, "no GPU visible"
    g = _gap(hc, 0)
    kw = _kw(g)
    out = {
        "mixed": hc.solve_nested(precision="float_then_double", **kw),
        "fp64": hc.solve_nested(precision="double", **kw),
        "active": hc.solve_nested(precision="float_then_double", active_set=True, **kw),
        "active_fp64": hc.solve_nested(precision="double", active_set=True, **kw),
        "mixed_fft": hc.solve_nested(precision="float_then_double", precond_engine="fft",
                                     **kw),
    }
    hc.clear_operator_cache()
    return g, out, oracle_op


@pytest.mark.parametrize("which", ["mixed", "fp64", "active", "active_fp64", "mixed_fft"])
def test_c3_certified_by_oracle(c3, which):
    g, out, op = c3
    r = out[which]
    print(which, r, flush=True)
    _certify(op, g, r)


def test_c3_mixed_matches_fp64(c3):
    _, out, _ = c3
    m, d, a = out["mixed"], out["fp64"], out["active"]
    for r in (m, a, out["mixed_fft"]):
        _close(r, d)
    _close(a, out["active_fp64"])
    _close(out["active_fp64"], d)
    # the float identification stage precedes an FP64 polish at the full tolerance
    st = m.stage_stats
    assert st[-2]["precision"] == "float" and st[-1]["name"] == "polish:4096"
    assert st[-1]["precision"] == "double" and st[-1]["status"] == "converged"


@pytest.mark.parametrize("seed", [1, 2])
def test_c3_active_pair_other_surfaces(oracle_op, seed):
    """The benchmark pair (active set; mixed and FP64) on two more surfaces
    of the same recipe."""
    import aspher_amd as hc
    g = _gap(hc, seed)
    kw = _kw(g)
    m = hc.solve_nested(precision="float_then_double", active_set=True, **kw)
    d = hc.solve_nested(precision="double", active_set=True, **kw)
    hc.clear_operator_cache()
    print(seed, m, d, flush=True)
    _certify(oracle_op, g, m)
    _certify(oracle_op, g, d)
    _close(m, d)
