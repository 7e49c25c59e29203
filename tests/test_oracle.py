"""Pins the CPU oracle (oracle/) against the reference's own fixtures, closed
forms and mathematical invariants. CPU only — no GPU, no product code.

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails one of them.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _love_fixture():
    rows = []
    with open(os.path.join(GOLD, "love_reference.txt")) as f:
        for line in f:
            if not line.strip() or line.startswith("#"):
                continue
            rows.append([float(v) for v in line.split()])
    return np.array(rows)


# ── a1: love_uz vs the reference's 80-digit fixture ─────────────────────────

def test_love_fixture_1e12():
    """tests/test_kernel_stability.cpp:53-85: 472 cases, rel <= 1e-12."""
    F = _love_fixture()
    assert F.shape[0] > 300
    got = O.love_uz(F[:, 0], F[:, 1], F[:, 2], F[:, 3])
    rel = np.abs(got - F[:, 4]) / np.abs(F[:, 4])
    assert np.all(np.isfinite(got)) and np.all(got > 0)
    assert rel.max() <= 1e-12, rel.max()


def test_love_self_term_parity_homogeneity():
    """tests/test_kernel_stability.cpp:88-100,150-185."""
    self_t = O.love_uz(0.0, 0.0, 0.5, 0.5)[0]
    assert abs(self_t - 4 * math.log(1 + math.sqrt(2))) <= 1e-15 * self_t
    for x in (0.0, 0.3, 1.0, 7.5, 1234.0, 5e5):
        for y in (0.0, 0.2, 2.0, 33.0, 9999.0, 3e5):
            v = O.love_uz(x, y, 0.5, 0.5)[0]
            assert O.love_uz(-x, y, 0.5, 0.5)[0] == v
            assert O.love_uz(x, -y, 0.5, 0.5)[0] == v
            sw = O.love_uz(y, x, 0.5, 0.5)[0]
            assert abs(sw - v) <= 1e-12 * v
    for s in (1e-3, 0.1, 3.0, 1e3, 1e6):
        for x in (0.0, 0.7, 12.0, 4000.0):
            y = 0.31 * x + 0.05
            v = O.love_uz(x, y, 0.5, 0.4)[0]
            vs = O.love_uz(s * x, s * y, s * 0.5, s * 0.4)[0]
            assert abs(vs - s * v) <= 1e-13 * abs(s * v)


def test_love_table_kernel_checks():
    """tests/test_kernel.cpp:15-56 (self term, point-source asymptotics, decay)."""
    Ns, L, E = 64, 1.0, 1.7
    h = L / Ns
    T = O.love_table(Ns, L, E)
    exact = 4 * h * math.log(1 + math.sqrt(2)) / (math.pi * E)
    assert abs(T[0, 0] / exact - 1) < 1e-14

    def rel(dx, dy):
        r = h * math.hypot(dx, dy)
        return abs(T[dy, dx] / (h * h / (math.pi * E * r)) - 1)
    assert 0.03 < rel(1, 0) < 0.05
    assert 0.015 < rel(1, 1) < 0.035
    assert rel(5, 0) < 0.002 and rel(4, 3) < 0.002 and rel(20, 11) < 0.0005
    assert np.all(np.diff(T[0, :]) < 0) and np.all(T > 0)


# ── a3/a5: free-space operator vs brute-force dense ─────────────────────────

@pytest.mark.parametrize("Ns", [8, 24, 32])
def test_free_space_matches_dense(Ns):
    """tests/test_fft.cpp:20-39: double < 1e-12, float < 1e-5."""
    op = O.FreeSpaceOperator(Ns)
    S = O.dense_matrix(Ns)
    x = np.random.default_rng(Ns).uniform(-1, 1, Ns * Ns)
    ref = S @ x
    y = op.matvec(x)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-12
    yf = op.matvec(x.astype(np.float32), np.float32)
    assert np.linalg.norm(yf - ref) / np.linalg.norm(ref) < 1e-5
    # the kernel spectrum is real and even in ky (only (Ns+1)^2 unique values)
    M = 2 * Ns
    assert np.allclose(op.Kh[1:, :], op.Kh[1:, :][::-1, :], rtol=0, atol=1e-13 * np.abs(op.Kh).max())
    assert op.Kh.shape == (M, Ns + 1)


def test_free_space_dense_symmetric():
    S = O.dense_matrix(6)
    assert np.array_equal(S, S.T)
    assert np.linalg.eigvalsh(S).min() > 0


# ── a6: periodic operator closed forms ──────────────────────────────────────

def test_periodic_cos_modes():
    """tests/test_periodic.cpp:16-36 (< 2e-13)."""
    n, L, E = 32, 3.0, 2.5
    op = O.PeriodicOperator(n, L, E)
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    for mx, my in ((1, 0), (0, 3), (2, 5)):
        q = 2 * math.pi * math.hypot(mx, my) / L
        p = np.cos(2 * math.pi * (mx * ix + my * iy) / n).ravel()
        u = op.matvec(p)
        ref = 2.0 / (E * q) * p
        assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 2e-13


@pytest.mark.parametrize("n", [7, 12, 16])
def test_periodic_matrix_properties(n):
    """tests/test_periodic.cpp:56-88: symmetric, PSD, circulant, DC null space."""
    op = O.PeriodicOperator(n, 1.7, 0.8)
    N = n * n
    S = np.stack([op.matvec(np.eye(N)[j]) for j in range(N)], axis=1)
    assert np.linalg.norm(S - S.T) / np.linalg.norm(S) < 1e-14
    ev = np.linalg.eigvalsh(S)
    assert ev[0] > -1e-13 * ev[-1] and ev[1] > 1e-6 * ev[-1]
    assert np.linalg.norm(op.matvec(np.ones(N))) < 1e-13
    sx, sy = 3 % n, 5 % n
    col = S[:, sy * n + sx].reshape(n, n)
    col0 = S[:, 0].reshape(n, n)
    assert np.allclose(col, np.roll(np.roll(col0, sy, 0), sx, 1), rtol=0, atol=1e-14 * S[0, 0])


# ── a7/a8: preconditioners ──────────────────────────────────────────────────

@pytest.mark.parametrize("Ns", [512, 1024])
def test_stencil_taps_match_reference_table(Ns):
    """tests/test_stencil.cpp:36-58: the universal-constant oracle table."""
    fx = json.load(open(os.path.join(GOLD, "stencil_taps.json")))
    sp = O.StencilPreconditioner(Ns, 2)
    assert len(sp.w) == 13
    for dx, dy, w in zip(sp.dx, sp.dy, sp.w):
        a, b = sorted((abs(dx), abs(dy)), reverse=True)
        ref = fx["taps"]["%d,%d" % (a, b)]
        assert abs(w - ref) <= fx["tolerance"], (dx, dy, w, ref)
        if dx == 0 and dy == 0:
            assert w == 1.0


@pytest.mark.parametrize("R", [1, 2, 4])
def test_stencil_symbol_positive(R):
    """tests/test_stencil.cpp:61-84."""
    Ns = 64
    sp = O.StencilPreconditioner(Ns, R)
    k = np.arange(Ns)
    KX, KY = np.meshgrid(k, k, indexing="xy")
    s = np.zeros((Ns, Ns))
    for dx, dy, w in zip(sp.dx, sp.dy, sp.w):
        s += w * np.cos(2 * math.pi * (KX * dx + KY * dy) / Ns)
    s[0, 0] = 1.0
    assert s.min() > 0


def test_preconditioners_offset_invariant_and_symmetric():
    """tests/test_precond.cpp:27-66."""
    Ns = 32
    N = Ns * Ns
    rng = np.random.default_rng(7)
    contact = rng.random(N) < 0.4
    chi = contact.astype(np.float64)
    g, x, y = (rng.uniform(-1, 1, N) for _ in range(3))
    for pc in (O.FourierPreconditioner(Ns), O.StencilPreconditioner(Ns, 2)):
        z0 = pc.apply(g, contact)
        z1 = pc.apply(g + 0.37 * chi, contact)
        assert np.abs(z1 - z0).max() / np.abs(z0).max() < 1e-12
        mx, my = pc.apply(x, contact), pc.apply(y, contact)
        xmy = (x * my)[contact].sum()
        ymx = (y * mx)[contact].sum()
        assert abs(xmy - ymx) / (abs(xmy) + abs(ymx)) < 1e-12
        assert np.all(z0[~contact] == 0)
        assert abs(z0[contact].sum()) < 1e-12 * np.abs(z0).sum()


@pytest.mark.parametrize("Ns", [16, 64, 96])
def test_fourier_preconditioner_cos_mode_eigenvalues(Ns):
    """On a full mask the preconditioner is F^-1 (|k|/Ns^2) F with unnormalised
    transforms, so a cos mode m is an eigenvector with eigenvalue |kof(m)| =
    hypot of the signed wavenumbers (src/fourier_precond.cpp:12-28),
    Nyquist rows and columns included (kof(Ns/2) = -Ns/2, magnitude Ns/2).
    A |k|^2 symbol, a wrong Nyquist kof or a dropped 1/Ns^2 all fail here;
    the symbol is float, so the gate is its rounding (6e-8)."""
    pc = O.FourierPreconditioner(Ns)
    full = np.ones(Ns * Ns, dtype=bool)
    iy, ix = np.meshgrid(np.arange(Ns), np.arange(Ns), indexing="ij")
    h = Ns // 2
    for mx, my in ((1, 0), (0, 1), (3, 5), (h, 0), (0, h), (h, h), (h - 1, h), (2, Ns - 3)):
        g = np.cos(2 * math.pi * (mx * ix + my * iy) / Ns).ravel()
        kx = mx if mx <= h else mx - Ns
        ky = my if my < h else my - Ns
        lam = math.hypot(kx, ky)
        z = pc.apply(g, full)
        assert np.linalg.norm(z - lam * g) <= 1e-7 * lam * np.linalg.norm(g), (mx, my)
    # the DC mode is the null space (mean removal on both sides)
    assert np.abs(pc.apply(np.full(Ns * Ns, 0.7), full)).max() == 0.0


def test_stencil_for_level_table():
    """tests/test_precond.cpp:68-101."""
    f = O.stencil_for_level
    assert not f("stencil", False, -1, 0.4) and not f("auto", False, 0.9, 0.4)
    assert f("stencil", True, 0.9, 0.4) and not f("fft", True, -1, 0.4)
    assert f("auto", True, -1, 0.4) and f("auto", True, 0.39, 0.4)
    assert not f("auto", True, 0.4, 0.4) and not f("auto", True, 0.41, 0.4)
    assert f("auto", True, 0.55, 0.6) and not f("auto", True, 0.65, 0.6)


# ── a9: projection vs sorting oracle ────────────────────────────────────────

def test_projection_matches_sorting():
    """tests/test_certification.cpp:443-467 (T21)."""
    rng = np.random.default_rng(7)
    for trial in range(200):
        n = 1 + trial % 40
        scale = 10.0 ** ((trial % 13) - 6)
        y = scale * rng.uniform(-1, 1, n)
        if trial % 5 == 0:
            y[:] = scale
        P = 0.37 * scale * n
        prod = y.copy()
        O.project_load_simplex(prod, P)
        ref = O.sorting_projection(y, P)
        assert prod.min() >= 0
        assert abs(prod.sum() - P) <= 1e-12 * P * max(1, n)
        assert np.linalg.norm(prod - ref) / max(np.linalg.norm(ref), 1e-300) < 1e-12
        again = prod.copy()
        O.project_load_simplex(again, P)
        assert np.linalg.norm(again - prod) <= 1e-12 * max(np.linalg.norm(prod), 1e-300)


# ── a10–a13: the PK solver against exact solutions ──────────────────────────

def _dense_op(S):
    return lambda v: S @ v


@pytest.mark.parametrize("case", ["singleton", "full", "islands"])
def test_manufactured_kkt(case):
    """T01 (tests/test_certification.cpp:76-152): known p*, enumeration oracle."""
    n = 10
    S = O.spd_kernel(n)
    p = np.zeros(n)
    g = np.zeros(n)
    if case == "singleton":
        p[4] = 1.0
        g = np.array([0.0 if i == 4 else 0.5 + 0.1 * i for i in range(n)])
    elif case == "full":
        p = 0.3 + 0.05 * np.arange(n)
    else:
        p[[0, 1, 7, 8]] = [0.7, 0.4, 0.9, 0.2]
        for i in range(n):
            if p[i] == 0.0:
                g[i] = 0.0 if i == 5 else 0.3 + 0.2 * i
    g0, P = O.manufacture(S, p, g, 0.25)
    gref = g0.max() - g0.min()
    obj_star = 0.5 * p @ S @ p + p @ g0
    en = O.enumerate_qp(S, g0, P)
    assert en[1] is not None and abs(en[0] - obj_star) < 1e-9 * (1 + abs(obj_star))
    for warm in range(3):
        p0 = None
        if warm == 1:
            p0 = p.copy()
        if warm == 2:
            p0 = np.zeros(n)
            p0[-1] = P
        opt = O.SolveOptions(tol=1e-12, max_iter=20000, g_ref=gref)
        r = O.solve_contact(_dense_op(S), g0, P / n, opt, None, p0)
        d = O.diagnose(_dense_op(S), g0, r.pressure, P, gref)
        assert r.status == "converged"
        assert d["fw_error"] <= 1e-12 and d["penetration"] <= 1e-12
        assert d["load_error"] <= 1e-12 and d["min_pressure"] >= 0
        assert np.linalg.norm(r.pressure - p) / np.linalg.norm(p) < 1e-7
        assert abs(r.fw_error - d["fw_error"]) <= 1e-9 * max(1, d["fw_error"])
        assert abs(r.approach - d["approach"]) <= 1e-9 * max(1, abs(d["approach"]))
        if warm == 1:
            assert r.iterations == 0
        if warm == 2:
            assert r.iterations > 0


def test_support_enumeration_random():
    """Brute force on tiny Love problems: PK == enumeration."""
    Ns = 3
    S = O.dense_matrix(Ns)
    rng = np.random.default_rng(3)
    for _ in range(3):
        g0 = rng.uniform(0, 0.05, Ns * Ns)
        P = 0.01 * Ns * Ns
        en = O.enumerate_qp(S, g0, P)
        r = O.solve_contact(_dense_op(S), g0, 0.01, O.SolveOptions(tol=1e-13, max_iter=5000))
        assert r.status == "converged"
        assert np.linalg.norm(r.pressure - en[1]) / np.linalg.norm(en[1]) < 1e-8


def test_status_paths():
    """T03 (tests/test_certification.cpp:215-323)."""
    Ns = 16
    S = O.dense_matrix(Ns)
    x = (np.arange(Ns) + 0.5) / Ns - 0.5
    g0 = (x[:, None] ** 2 + x[None, :] ** 2).ravel()
    for mi in (0, 1):
        r = O.solve_contact(_dense_op(S), g0, 0.02, O.SolveOptions(tol=1e-10, max_iter=mi))
        assert r.status == "max_iterations" and not r.converged
    calls = [0]

    def bad(v):
        calls[0] += 1
        y = S @ v
        if calls[0] > 3:
            y[0] = np.nan
        return y
    r = O.solve_contact(bad, g0, 0.02, O.SolveOptions(tol=1e-10, max_iter=100))
    assert r.status == "nonfinite"
    gb = g0.copy()
    gb[3] = np.inf
    with pytest.raises(ValueError):
        O.solve_contact(_dense_op(S), gb, 0.02)
    for args in ((-1.0, 1e-10, 10), (0.02, 0.0, 10), (0.02, 1e-10, -3)):
        with pytest.raises(ValueError):
            O.solve_contact(_dense_op(S), g0, args[0],
                            O.SolveOptions(tol=args[1], max_iter=args[2]))


def test_fw_bound_identities():
    """T31 (tests/test_certification.cpp:376-440): 0 <= f-f* <= G, |p-p*|_S^2 <= 2G."""
    n = 8
    S = O.spd_kernel(n, 0.7)
    p_star = np.zeros(n)
    p_star[[1, 2, 6]] = [0.5, 1.5, 0.75]
    g_star = np.where(p_star > 0, 0.0, 0.2 + 0.05 * np.arange(n))
    g0, P = O.manufacture(S, p_star, g_star, 0.1)
    gref = g0.max() - g0.min()
    f_star = 0.5 * p_star @ S @ p_star + p_star @ g0
    rng = np.random.default_rng(12345)
    for _ in range(200):
        p = O.sorting_projection(rng.random(n), P)
        d = O.diagnose(_dense_op(S), g0, p, P, gref)
        gap = d["objective"] - f_star
        assert gap >= -1e-12 and gap <= d["fw_gap"] + 1e-12
        e = p - p_star
        assert e @ S @ e <= 2 * d["fw_gap"] + 1e-12


def test_westergaard_full_and_partial():
    """tests/test_periodic.cpp:94-129 (n reduced 256 -> 128 for CPU time)."""
    n, L, Es, D = 128, 2.0, 1.3, 0.01
    pstar = math.pi * Es * D / L
    op = O.PeriodicOperator(n, L, Es)
    cosx = np.tile(np.cos(2 * math.pi * np.arange(n) / n), (n, 1)).ravel()
    g0 = D * (1 - cosx)
    pbar = 1.5 * pstar
    r = O.solve_contact(op.matvec, g0, pbar, O.SolveOptions(tol=1e-12))
    assert r.status == "converged"
    pex = pbar + pstar * cosx
    assert np.linalg.norm(r.pressure - pex) / np.linalg.norm(pex) < 1e-9
    pbar = 0.3 * pstar
    r = O.solve_contact(op.matvec, g0, pbar, O.SolveOptions(tol=1e-10))
    assert r.status == "converged"
    width = (r.pressure[:n] > 0).sum() * L / n
    wex = 2 * (L / math.pi) * math.asin(math.sqrt(pbar / pstar))
    assert abs(width - wex) <= 2 * L / n


def test_hertz_fft_equals_dense():
    """tests/test_fft.cpp:63-107 (Ns 64 -> 32): fft vs dense solve."""
    Ns, R, pbar = 32, 0.5, 0.05
    h = 1.0 / Ns
    c = (np.arange(Ns) + 0.5) * h - 0.5
    g0 = ((c[:, None] ** 2 + c[None, :] ** 2) / (2 * R)).ravel()
    S = O.dense_matrix(Ns)
    op = O.FreeSpaceOperator(Ns)
    rd = O.solve_contact(_dense_op(S), g0, pbar, O.SolveOptions(tol=1e-8, max_iter=2000))
    rf = O.solve_contact(op.matvec, g0, pbar, O.SolveOptions(tol=1e-8, max_iter=2000))
    assert rd.converged and rf.converged
    assert abs(rd.iterations - rf.iterations) <= 1
    assert rd.contact_fraction == rf.contact_fraction
    assert np.linalg.norm(rf.pressure - rd.pressure) / np.linalg.norm(rd.pressure) < 1e-6


def test_nested_matches_single_level_and_datum():
    """tests/test_periodic_py.py:53-67 and tests/test_certification_py.py:96-121."""
    ns = 64
    rng = np.random.default_rng(3)
    k = np.fft.fftfreq(ns, 1.0 / ns)
    kk = np.hypot(k[:, None], k[None, :])
    kk[0, 0] = 1
    amp = np.where((kk >= 2) & (kk <= ns // 8), kk ** -1.8, 0.0)
    h = np.real(np.fft.ifft2(amp * np.exp(2j * np.pi * rng.random((ns, ns)))))
    h *= 0.01 / h.std()
    g = (h.max() - h).ravel()
    op = O.PeriodicOperator(ns)
    fp = O.FourierPreconditioner(ns)
    ref = O.solve_contact(op.matvec, g, 0.01, O.SolveOptions(tol=1e-12), fp.apply)
    for prec, tol, bound in (("double", 1e-10, 1e-7), ("float_then_double", 1e-10, 1e-7),
                             ("float", 2e-6, 1e-4)):
        r = O.solve_nested(ns, g, 0.01, coarsest=16, backend="periodic", tol=tol,
                           precision=prec)
        assert r.status == "converged", (prec, r.status, r.status_reason)
        assert abs(r.pressure.mean() - 0.01) < 1e-6 * 0.01
        assert np.linalg.norm(r.pressure - ref.pressure) / np.linalg.norm(ref.pressure) < bound
    base = O.solve_nested(ns, g, 0.01, coarsest=16, backend="fft", tol=1e-10)
    for shift in (1e3, 1e6):
        r = O.solve_nested(ns, g + shift, 0.01, coarsest=16, backend="fft", tol=1e-10)
        assert r.status == "converged"
        assert abs(r.approach - base.approach - shift) < 1e-6 * shift
        assert np.linalg.norm(r.pressure - base.pressure) / np.linalg.norm(base.pressure) < 1e-5


# ── adjoint: kkt_gradient against central finite differences ───────────────

def _loss(p, pt):
    return 0.5 * float(((p - pt) ** 2).sum())


def test_kkt_gradient_vs_exact_enumeration_fd():
    """Exact solutions (support enumeration) make the finite differences
    accurate to O(eps^2): gradient w.r.t. every g0 entry and P."""
    n = 10
    S = O.spd_kernel(n, 0.7)
    rng = np.random.default_rng(5)
    p_star = np.where(np.arange(n) % 3 != 1, rng.uniform(0.5, 1.5, n), 0.0)
    g_star = np.where(p_star > 0, 0.0, rng.uniform(0.2, 0.6, n))
    g0, P = O.manufacture(S, p_star, g_star, 0.3)
    _, p, _ = O.enumerate_qp(S, g0, P)
    np.testing.assert_allclose(p, p_star, atol=1e-10)
    pt = rng.uniform(0.0, 1.0, n)
    gg, gP = O.kkt_gradient(S, p, p - pt)
    eps = 1e-6
    for i in range(n):
        e = np.zeros(n)
        e[i] = eps
        fd = (_loss(O.enumerate_qp(S, g0 + e, P)[1], pt) -
              _loss(O.enumerate_qp(S, g0 - e, P)[1], pt)) / (2 * eps)
        assert abs(fd - gg[i]) <= 1e-7 * (1 + abs(gg[i])), (i, fd, gg[i])
        if p[i] == 0:
            assert gg[i] == 0.0
    fdP = (_loss(O.enumerate_qp(S, g0, P + eps)[1], pt) -
           _loss(O.enumerate_qp(S, g0, P - eps)[1], pt)) / (2 * eps)
    assert abs(fdP - gP) <= 1e-7 * (1 + abs(gP))
    # a plausible sign slip (dL/dg0 = +mu) fails
    assert np.abs(gg).max() > 1e-3


def test_kkt_gradient_love_operator_fd():
    """Love dense operator at Ns=8 through the oracle's PK solver."""
    Ns = 8
    S = O.dense_matrix(Ns)
    x = (np.arange(Ns) + 0.5) / Ns - 0.5
    g0 = (x[:, None] ** 2 + 0.5 * x[None, :] ** 2).ravel()
    g0 = g0 + 0.01 * np.cos(7 * np.arange(Ns * Ns))
    pbar = 0.05
    opt = O.SolveOptions(tol=1e-15, max_iter=20000)
    mv = lambda v, dtype=np.float64: S @ v  # noqa: E731
    r = O.solve_contact(mv, g0, pbar, opt)
    p = r.pressure
    assert 0 < (p > 0).sum() < Ns * Ns
    pt = np.roll(p, 3)
    gg, gP = O.kkt_gradient(S, p, p - pt)
    eps = 1e-6
    v = S @ p + g0
    gap = v - v[p > 0].mean()
    interior = np.nonzero(p > 0.2 * p.max())[0][:6]
    outside = np.nonzero(gap > 0.2 * gap.max())[0][:3]
    for i in list(interior) + list(outside):
        e = np.zeros(Ns * Ns)
        e[i] = eps
        lp = _loss(O.solve_contact(mv, g0 + e, pbar, opt, p_init=p.copy()).pressure, pt)
        lm = _loss(O.solve_contact(mv, g0 - e, pbar, opt, p_init=p.copy()).pressure, pt)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - gg[i]) <= 1e-5 * max(1.0, np.abs(gg).max()), (i, fd, gg[i])
    N = Ns * Ns
    lp = _loss(O.solve_contact(mv, g0, pbar + eps, opt, p_init=p.copy()).pressure, pt)
    lm = _loss(O.solve_contact(mv, g0, pbar - eps, opt, p_init=p.copy()).pressure, pt)
    fdp = (lp - lm) / (2 * eps)
    assert abs(fdp - gP * N) <= 1e-5 * max(1.0, abs(gP * N))


def test_circulant_matrix_matches_periodic_operator():
    Ns = 8
    op = O.PeriodicOperator(Ns, 1.0, 1.0)
    S = O.circulant_matrix(op.matvec, Ns)
    x = np.random.default_rng(2).random(Ns * Ns)
    np.testing.assert_allclose(S @ x, op.matvec(x), rtol=0, atol=1e-14 * np.abs(S @ x).max())
    np.testing.assert_allclose(S, S.T, atol=1e-15 * np.abs(S).max())


# ── rough-surface generator ─────────────────────────────────────────────────

def test_splitmix64_known_answer():
    """Published SplitMix64 sequence for seed 1234567."""
    out = O.splitmix64(1234567, np.arange(5))
    assert [int(v) for v in out] == [6457827717110365317, 3203168211198807973,
                                      9817491932198370423, 4593380528125082431,
                                      16408922859458223821]
    u = O.hash_uniform(7, np.arange(200000))
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5e-3 and abs(u.var() - 1 / 12) < 2e-3


def test_rough_surface_hash_recipe():
    Ns, H, kl, kh, rms = 64, 0.7, 4, 20, 0.002
    h = O.rough_surface_hash(Ns, H, kl, kh, rms, seed=3)
    assert h.shape == (Ns, Ns) and np.isrealobj(h)
    assert abs(h.mean()) < 1e-18 and abs(h.std() - rms) < 1e-15
    F = np.fft.fft2(h)
    k = np.fft.fftfreq(Ns, 1.0 / Ns)
    kk = np.hypot(k[:, None], k[None, :])
    band = (kk >= kl) & (kk <= kh)
    assert np.abs(F[~band]).max() < 1e-12 * np.abs(F).max()
    # the same code path with numpy's phases is the host recipe
    from aspher_amd.surfaces import rough_surface
    rng = np.random.default_rng(11)
    ph = rng.random((Ns, Ns))
    a = O.rough_surface_from_phases(Ns, H, kl, kh, rms, ph)
    b = rough_surface(Ns, H=H, k_low=kl, k_high=kh, rms=rms, seed=11)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-15)
    # a different seed is a different surface
    assert np.abs(O.rough_surface_hash(Ns, H, kl, kh, rms, seed=4) - h).max() > 0.1 * rms


# ── tangential (Cerruti) operator ───────────────────────────────────────────

def _gauss_legendre(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def test_cerruti_brackets_vs_quadrature():
    """tests/test_cerruti.cpp:53-78: 64-point tensor Gauss-Legendre, 1e-9."""
    h = 1.0 / 32
    a, nu = 0.5 * h, 0.3
    xs, ws = _gauss_legendre(64)
    XI, ETA = np.meshgrid(a * xs, a * xs, indexing="ij")
    W = np.outer(ws, ws) * a * a
    for ox, oy in ((1, 0), (0, 2), (1, 1), (2, 1), (5, 3), (9, 4)):
        x, y = ox * h, oy * h
        s = np.hypot(x - XI, y - ETA)
        qxx = (W * ((1 - nu) / s + nu * (x - XI) ** 2 / s ** 3)).sum()
        qxy = (W * ((x - XI) * (y - ETA) / s ** 3)).sum()
        assert abs(O.cerruti_uxx(x, y, a, a, nu)[0] / qxx - 1) < 1e-9
        if ox and oy:
            assert abs(O.cerruti_uxy(x, y, a, a)[0] / qxy - 1) < 1e-9


def test_cerruti_identities():
    """tests/test_cerruti.cpp:80-101."""
    a, b, x, y, nu = 0.011, 0.017, 0.05, -0.03, 0.31
    assert abs(O.cerruti_uxx(x, y, a, b, 0.0)[0] / O.love_uz(x, y, a, b)[0] - 1) < 1e-13
    assert O.cerruti_uxx(x, y, a, b, nu)[0] == O.cerruti_uxx(-x, y, a, b, nu)[0]
    assert O.cerruti_uxx(x, y, a, b, nu)[0] == O.cerruti_uxx(x, -y, a, b, nu)[0]
    assert O.cerruti_uxy(-x, y, a, b)[0] == -O.cerruti_uxy(x, y, a, b)[0]
    assert O.cerruti_uxy(x, -y, a, b)[0] == -O.cerruti_uxy(x, y, a, b)[0]
    self_t = (2 - nu) * 4 * a * math.log(1 + math.sqrt(2))
    assert abs(O.cerruti_uxx(0.0, 0.0, a, a, nu)[0] / self_t - 1) < 1e-14
    assert O.cerruti_uxy(0.0, 0.0, a, a)[0] == 0.0


def test_cerruti_tables_dense_block():
    """tests/test_cerruti.cpp:103-143: accessors, symmetric PD block."""
    Ns, L, E, nu = 16, 1.0, 1.7, 0.3
    C = O.CerrutiTables(Ns, L, E, nu)
    a = 0.5 * L / Ns
    self_t = (2 - nu) * 4 * a * math.log(1 + math.sqrt(2)) / (math.pi * E * (1 - nu))
    assert abs(C.xx_off(0, 0) / self_t - 1) < 1e-14 and C.xy_off(0, 0) == 0
    assert C.yy_off(3, 1) == C.xx_off(1, 3) and C.xx_off(-3, 1) == C.xx_off(3, 1)
    assert C.xy_off(2, 5) == C.xy_off(-2, -5) and C.xy_off(-2, 5) == -C.xy_off(2, 5)
    assert C.xy_off(0, 4) == 0 and C.xy_off(4, 0) == 0 and C.xx_off(Ns, 0) == 0
    M = O.CerrutiTables(8, L, E, nu).dense()
    assert np.array_equal(M, M.T)
    np.linalg.cholesky(M)


def test_cerruti_symbol():
    """tests/test_cerruti.cpp:145-190: eigenvalues of the continuum symbol and
    the lattice DFT of the tables against it."""
    C = O.CerrutiTables(128, 1.0, 1.7, 0.3)
    kx, ky = 7.3, -2.1
    kk = math.hypot(kx, ky)
    S = C.symbol(kx, ky)
    el, et = np.array([kx, ky]) / kk, np.array([-ky, kx]) / kk
    assert abs(el @ S @ el / (2 / (1.7 * kk)) - 1) < 1e-14
    assert abs(et @ S @ et / (2 / (1.7 * 0.7 * kk)) - 1) < 1e-14
    h = 1.0 / 128
    d = np.arange(-127, 128)
    DX, DY = np.meshgrid(d, d, indexing="xy")
    for mx, my in ((8, 0), (0, 8), (6, 6), (12, 5)):
        qx, qy = 2 * math.pi * mx, 2 * math.pi * my
        c = np.cos(qx * DX * h + qy * DY * h)
        S = C.symbol(qx, qy)
        assert abs((C.xx_off(DX, DY) * c).sum() / S[0, 0] - 1) < 0.18
        assert abs((C.yy_off(DX, DY) * c).sum() / S[1, 1] - 1) < 0.18
        if mx and my:
            # odd table: the sine-sine part carries the cross symbol
            sxy = -(C.xy_off(DX, DY) * np.sin(qx * DX * h) * np.sin(qy * DY * h)).sum()
            assert abs(sxy / S[0, 1] - 1) < 0.10


@pytest.mark.parametrize("Ns", [8, 12, 16])
def test_tangential_fft_equals_dense(Ns):
    """tests/test_tangential.cpp:27-40 (rel < 1e-13)."""
    T = O.TangentialOperator(Ns, 1.0, 1.3, 0.29)
    M = O.CerrutiTables(Ns, 1.0, 1.3, 0.29).dense()
    q = np.random.default_rng(Ns).standard_normal(2 * Ns * Ns)
    u = T.matvec(q)
    ref = M @ q
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-13


def test_tangential_preconditioner_symmetric():
    Ns = 16
    P = O.TangentialPreconditioner(Ns, 0.3)
    rng = np.random.default_rng(1)
    mask = rng.random(Ns * Ns) < 0.6
    a, b = rng.standard_normal(2 * Ns * Ns), rng.standard_normal(2 * Ns * Ns)
    m2 = np.concatenate([mask, mask])
    a[~m2] = 0
    b[~m2] = 0
    za, zb = P.apply(a, mask), P.apply(b, mask)
    assert abs(za @ b - zb @ a) < 1e-12 * np.abs(za @ b)
    assert np.all(za[~m2] == 0)
    zm = P.apply(a, mask, remove_mean=True)
    N = Ns * Ns
    assert abs(zm[:N][mask].sum()) < 1e-12 and abs(zm[N:][mask].sum()) < 1e-12
