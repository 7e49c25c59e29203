"""Active-set (restricted) PK and its nested driver in the oracle, pinned by
the reference's own gates (tests/test_active.cpp:309-561): a restricted
certificate is not a global one, the restricted iterate stays zero outside
the candidate set, the active nested solve reproduces the standard one, a
truncated candidate set is caught (fallback) or expanded, and the stage
names / occupancy gate behave as specified. Also pins the device
formulation: the restricted solve equals the full solve on the gap lifted
far above every reachable value outside the candidate set."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O


def _paraboloid(Ns):
    x = (np.arange(Ns) + 0.5) / Ns - 0.5
    X, Y = np.meshgrid(x, x, indexing="xy")
    return (X ** 2 + Y ** 2).ravel()


def test_dilate_mask_is_clipped_square_max_filter():
    rng = np.random.default_rng(0)
    for Ns, r in ((16, 0), (16, 1), (32, 2), (64, 3)):
        m = rng.random(Ns * Ns) < 0.03
        ref = ndimage.maximum_filter(m.reshape(Ns, Ns).astype(np.uint8), size=2 * r + 1,
                                     mode="constant", cval=0).astype(bool).ravel()
        np.testing.assert_array_equal(O.dilate_mask(m, Ns, r), ref)


def test_restricted_certificate_is_not_global():
    """tests/test_active.cpp:513-561: candidates = left half of a paraboloid
    contact; the local subproblem converges, the global check fails."""
    Ns = 32
    N = Ns * Ns
    D = O.DenseOperator(Ns)
    g2 = _paraboloid(Ns)
    pbar = 0.02
    P = pbar * N
    gref = np.ptp(g2)
    idx = [iy * Ns + ix for iy in range(Ns) for ix in range(Ns // 2)]
    o = O.SolveOptions(tol=1e-10, max_iter=20000, g_ref=gref, datum_mode="caller")
    rr, cert = O.solve_contact_active(D.matvec, g2, pbar, idx, o, n_grid=N)
    v = D.matvec(rr.pressure) + g2
    pen_glob = max(0.0, -(v.min() - rr.approach)) / gref
    G_glob = rr.fw_gap + P * (cert["vmin_local"] - v.min())
    assert rr.status == "converged" and rr.fw_error <= 1e-10
    assert pen_glob > 1e-3 and G_glob / (P * gref) > 1e-3
    assert np.abs(rr.pressure.reshape(Ns, Ns)[:, Ns // 2:]).sum() == 0.0
    assert abs(rr.mean_pressure - pbar) < 1e-12


@pytest.mark.parametrize("dtype,tol,ptol", [(np.float64, 1e-10, 1e-10), (np.float32, 1e-5, 1e-6)])
def test_restricted_equals_full_solve_on_lifted_gap(dtype, tol, ptol):
    """The index-driven restricted solve and the full-grid solve whose gap is
    lifted by 1e8 g_ref outside the candidate set (warm-started on the set)
    take the same iterations to the same pressure."""
    Ns = 32
    N = Ns * Ns
    D = O.DenseOperator(Ns)
    rng = np.random.default_rng(3)
    g2 = _paraboloid(Ns) + 0.002 * rng.standard_normal(N)
    gref = np.ptp(g2)
    cand = O.dilate_mask(rng.random(N) < 0.1, Ns, 1) | (g2 < 0.05)
    idx = np.flatnonzero(cand)
    mv = lambda v: D.matvec(v, dtype)
    opt = lambda: O.SolveOptions(tol=tol, max_iter=20000, g_ref=gref, datum_mode="caller")
    ra, _ = O.solve_contact_active(mv, g2.astype(dtype), 0.01, idx, opt(), n_grid=N,
                                   dtype=dtype)
    lifted = np.where(cand, g2, 1e8 * gref).astype(dtype)
    p0 = np.where(cand, 0.01 * N / idx.size, 0.0).astype(dtype)
    rf = O.solve_contact(mv, lifted, 0.01, opt(), p_init=p0, dtype=dtype)
    assert ra.status == rf.status == "converged"
    assert ra.iterations == rf.iterations
    assert np.abs(rf.pressure - ra.pressure).max() <= ptol * ra.pressure.max()
    assert np.all(rf.pressure[~cand] == 0.0)
    assert abs(rf.approach - ra.approach) <= 1e-9 * gref


@pytest.fixture(scope="module")
def nested128():
    from aspher_amd.surfaces import rough_gap
    Ns = 128
    g = rough_gap(Ns, H=0.8, seed=42)
    ops = {n: O.FreeSpaceOperator(n) for n in (32, 64, 128)}
    kw = dict(coarsest=32, tol=1e-8, operators=ops)
    r0 = O.solve_nested(Ns, g, 0.005, **kw)
    assert r0.converged
    return Ns, g, kw, r0


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_active_nested_matches_standard(nested128):
    Ns, g, kw, r0 = nested128
    r1 = O.solve_nested(Ns, g, 0.005, active_set=True, **kw)
    assert r1.converged and not r1.active_fallback
    assert 1 <= r1.active_rounds <= 2
    assert r1.verification_matvec_count == r1.active_rounds
    assert _rel(r1.pressure, r0.pressure) <= 1e-6
    assert abs(r1.contact_fraction - r0.contact_fraction) <= 1e-6
    assert np.linalg.norm(r1.gap - r0.gap) <= 1e-6 * (np.linalg.norm(r0.gap) + 1.0)
    names = [s["name"] for s in r1.stage_stats]
    assert names == ["coarse:32", "coarse:64(active)", "finest:128(active)"]
    r2 = O.solve_nested(Ns, g, 0.005, active_set=True, precision="float", **kw)
    assert r2.converged and not r2.active_fallback
    assert _rel(r2.pressure, r0.pressure) <= 1e-3
    assert abs(r2.contact_fraction - r0.contact_fraction) <= 2e-3


def test_active_truncated_candidates_fall_back_or_expand(nested128):
    Ns, g, kw, r0 = nested128
    r3 = O.solve_nested(Ns, g, 0.005, active_set=True, active_delta=0.0, active_halo=0,
                        active_max_rounds=1, **kw)
    assert r3.converged and r3.active_fallback
    assert _rel(r3.pressure, r0.pressure) <= 1e-4
    r4 = O.solve_nested(Ns, g, 0.005, active_set=True, active_delta=0.0, active_halo=0,
                        active_max_rounds=8, **kw)
    assert r4.converged
    assert _rel(r4.pressure, r0.pressure) <= 1e-4
    assert abs(r4.contact_fraction - r0.contact_fraction) <= 1e-4


def test_active_level_selection_and_gate(nested128):
    Ns, g, kw, r0 = nested128
    r6 = O.solve_nested(Ns, g, 0.005, active_set=True, active_all_levels=False, **kw)
    assert [s["name"] for s in r6.stage_stats] == ["coarse:32", "coarse:64", "finest:128(active)"]
    assert _rel(r6.pressure, r0.pressure) <= 1e-6
    r7 = O.solve_nested(Ns, g, 0.005, active_set=True, active_occupancy_max=1e-6, **kw)
    assert not any("(active)" in s["name"] for s in r7.stage_stats)
    assert r7.active_rounds == 0
    assert _rel(r7.pressure, r0.pressure) <= 1e-6


def test_active_third_exit_fallback(nested128):
    """tests/test_active.cpp:454-511: candidates cover the grid (no
    violations), the restricted CG is starved (max_iter 2): the driver falls
    back and returns a finite non-negative pressure."""
    Ns, g, kw, _ = nested128
    kw = dict(kw, max_iter=2)
    r8 = O.solve_nested(Ns, g, 0.005, active_set=True, active_delta=1e9, **kw)
    assert r8.active_fallback
    assert r8.pressure.size == Ns * Ns
    assert np.all(np.isfinite(r8.pressure)) and np.all(r8.pressure >= 0)


def test_active_rejections():
    g = np.zeros(64 * 64)
    with pytest.raises(ValueError):
        O.solve_nested(64, g, 0.01, coarsest=32, backend="periodic", active_set=True)
    with pytest.raises(ValueError):
        O.solve_nested(64, g, 0.01, coarsest=64, active_set=True)
