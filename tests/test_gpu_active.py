"""Active-set nested solve on the free-space backend (GPU): the restricted
level driver must reproduce the standard nested solve with a global
certificate, catch a truncated candidate set (fallback) or expand it, and
follow the level selection rules (tests/test_active.cpp:309-511), and agree
with the oracle's index-driven restatement."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def hc():
    import aspher_amd
    assert aspher_amd.device_count() >= 1, "no GPU visible"
    return aspher_amd


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def base(hc):
    Ns, pbar, tol = 256, 0.005, 1e-8
    g = hc.surfaces.rough_gap(Ns, H=0.8, seed=42)
    kw = dict(grid_size=Ns, gap=g, p_nominal=pbar, coarsest=64, tol=tol, backend="fft",
              max_iter=20000)
    r0 = hc.solve_nested(**kw)
    assert r0.converged
    return kw, r0


def test_active_matches_standard_fp64(hc, base):
    kw, r0 = base
    r1 = hc.solve_nested(active_set=True, **kw)
    print(r1, r1.active_rounds, [s["name"] for s in r1.stage_stats], flush=True)
    assert r1.converged and not r1.active_fallback
    assert 1 <= r1.active_rounds <= 2
    assert r1.verification_matvec_count == r1.active_rounds
    assert _rel(r1.pressure.ravel(), r0.pressure.ravel()) <= 1e-6
    assert abs(r1.contact_area - r0.contact_area) <= 1e-6
    assert np.linalg.norm(r1.gap - r0.gap) <= 1e-6 * (np.linalg.norm(r0.gap) + 1.0)
    assert [s["name"] for s in r1.stage_stats] == ["coarse:64", "coarse:128(active)",
                                                  "finest:256(active)"]
    assert r1.fw_error <= kw["tol"] and r1.penetration_error <= kw["tol"]
    # independent global certificate of the returned pressure
    Ns = kw["grid_size"]
    d = O.diagnose(O.FreeSpaceOperator(Ns).matvec, kw["gap"], r1.pressure.ravel(),
                   kw["p_nominal"] * Ns * Ns, np.ptp(kw["gap"]))
    assert d["fw_error"] <= 1.01 * kw["tol"] and d["penetration"] <= 1.01 * kw["tol"], d
    assert r1.timings["candidate"] >= 0.0 and r1.timings["verification"] > 0.0


def test_active_fp32_and_mixed(hc, base):
    kw, r0 = base
    r2 = hc.solve_nested(active_set=True, single_precision=True, **kw)
    assert r2.converged and not r2.active_fallback
    assert _rel(r2.pressure.ravel(), r0.pressure.ravel()) <= 1e-3
    assert abs(r2.contact_area - r0.contact_area) <= 2e-3
    # float identification restricted, FP64 polish on the full grid
    r3 = hc.solve_nested(active_set=True, precision="float_then_double", **kw)
    assert r3.converged and r3.fw_error <= kw["tol"] and r3.penetration_error <= kw["tol"]
    names = [s["name"] for s in r3.stage_stats]
    assert names[-2] == "finest:256(active)" and names[-1] == "polish:256"
    assert r3.stage_stats[-2]["precision"] == "float"
    assert r3.stage_stats[-1]["precision"] == "double"
    assert _rel(r3.pressure.ravel(), r0.pressure.ravel()) <= 1e-6


def test_active_truncated_candidates(hc, base):
    kw, r0 = base
    r3 = hc.solve_nested(active_set=True, active_delta=0.0, active_halo=0, active_max_rounds=1,
                         **kw)
    assert r3.converged and r3.active_fallback
    assert _rel(r3.pressure.ravel(), r0.pressure.ravel()) <= 1e-4
    assert abs(r3.contact_area - r0.contact_area) <= 1e-4
    r4 = hc.solve_nested(active_set=True, active_delta=0.0, active_halo=0, active_max_rounds=8,
                         **kw)
    assert r4.converged
    assert _rel(r4.pressure.ravel(), r0.pressure.ravel()) <= 1e-4
    assert abs(r4.contact_area - r0.contact_area) <= 1e-4


def test_active_level_selection_gate_and_third_exit(hc, base):
    kw, r0 = base
    r6 = hc.solve_nested(active_set=True, active_all_levels=False, **kw)
    assert [s["name"] for s in r6.stage_stats] == ["coarse:64", "coarse:128",
                                                  "finest:256(active)"]
    assert _rel(r6.pressure.ravel(), r0.pressure.ravel()) <= 1e-6
    r7 = hc.solve_nested(active_set=True, active_occupancy_max=1e-6, **kw)
    assert not any("(active)" in s["name"] for s in r7.stage_stats)
    assert r7.active_rounds == 0
    assert _rel(r7.pressure.ravel(), r0.pressure.ravel()) <= 1e-6
    kw8 = dict(kw, max_iter=2)
    r8 = hc.solve_nested(active_set=True, active_delta=1e9, **kw8)
    assert r8.active_fallback
    p = r8.pressure.ravel()
    assert p.size == kw["grid_size"] ** 2 and np.all(np.isfinite(p)) and np.all(p >= 0)


@pytest.mark.parametrize("precision", ["double", "float"])
def test_active_matches_oracle_restatement(hc, precision):
    """Same level hierarchy through the oracle's index-driven restricted
    solve (src/contact_solver.cpp:601-1040 restated) and the device's lifted
    gap formulation: same rounds and fallback decisions, same pressure."""
    Ns = 128
    g = hc.surfaces.rough_gap(Ns, H=0.8, seed=7)
    ops = {n: O.FreeSpaceOperator(n) for n in (32, 64, 128)}
    ro = O.solve_nested(Ns, g, 0.005, coarsest=32, tol=1e-8, operators=ops, active_set=True,
                        precision=precision)
    rg = hc.solve_nested(grid_size=Ns, gap=g, p_nominal=0.005, coarsest=32, tol=1e-8,
                         backend="fft", active_set=True, precision=precision)
    assert ro.converged and rg.converged
    assert rg.active_rounds == ro.active_rounds
    assert rg.active_fallback == ro.active_fallback
    assert [s["name"] for s in rg.stage_stats] == [s["name"] for s in ro.stage_stats]
    tolp = 1e-6 if precision == "double" else 1e-3
    assert _rel(rg.pressure.ravel(), ro.pressure) <= tolp
    assert abs(rg.contact_area - ro.contact_fraction) <= (1e-6 if precision == "double"
                                                          else 2e-3)
