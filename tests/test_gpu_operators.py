"""GPU parity of the operators and preconditioners against the CPU oracle, the
reference's golden fixtures and closed forms. Every call goes through the C ABI
(libaspher_hip.so)."""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def hc():
    import aspher_amd
    from aspher_amd import _capi
    _capi.lib()
    assert aspher_amd.device_count() >= 1, "no GPU visible"
    return aspher_amd


def _dev(a, dtype):
    from aspher_amd.device import DeviceArray
    return DeviceArray.from_numpy(np.asarray(a), dtype)


def _love_fixture():
    rows = [[float(v) for v in l.split()] for l in open(os.path.join(GOLD, "love_reference.txt"))
            if l.strip() and not l.startswith("#")]
    return np.array(rows)


def test_device_love_uz_matches_80_digit_fixture(hc):
    """tests/test_kernel_stability.cpp:53-85 on the device kernel."""
    from aspher_amd._capi import lib, check, dptr
    F = np.ascontiguousarray(_love_fixture())
    cols = [np.ascontiguousarray(F[:, k]) for k in range(5)]
    out = np.empty(F.shape[0])
    check(lib().aspher_love_points(F.shape[0], dptr(cols[0]), dptr(cols[1]), dptr(cols[2]),
                                   dptr(cols[3]), dptr(out)))
    rel = np.abs(out - cols[4]) / np.abs(cols[4])
    assert np.all(np.isfinite(out)) and rel.max() <= 1e-12, rel.max()


@pytest.mark.parametrize("Ns,L,E", [(64, 1.0, 1.7), (257, 2.0, 1.0)])
def test_device_love_table_matches_oracle(hc, Ns, L, E):
    from aspher_amd._capi import lib, check, dptr
    out = np.empty(Ns * Ns)
    check(lib().aspher_love_table(Ns, L, E, dptr(out)))
    ref = O.love_table(Ns, L, E).ravel()
    # both sides are within the reference's 1e-12 love_uz accuracy gate
    # (the far-field asinh differences lose ~log10(r/h) digits)
    assert np.max(np.abs(out - ref) / np.abs(ref)) < 1e-12


@pytest.mark.parametrize("Ns", [8, 24, 32, 64])
def test_free_space_matvec_matches_dense(hc, Ns):
    """tests/test_fft.cpp:20-39: FP64 < 1e-12, FP32 < 1e-5 against the dense sum."""
    s = hc.ContactSolver(Ns, backend="fft")
    S = O.dense_matrix(Ns)
    x = np.random.default_rng(Ns).uniform(-1, 1, Ns * Ns)
    ref = S @ x
    y = s.matvec(x)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-12
    xf = _dev(x, np.float32)
    yd = _dev(np.zeros(Ns * Ns), np.float32)
    s.matvec_device(xf.ptr, yd.ptr, 32)
    yf = yd.numpy().astype(np.float64)
    assert np.linalg.norm(yf - ref) / np.linalg.norm(ref) < 1e-5


@pytest.mark.parametrize("backend,Ns", [("fft", 16), ("fft", 128), ("fft", 512), ("fft", 1024),
                                        ("periodic", 32), ("periodic", 256),
                                        ("periodic", 512), ("periodic", 1024)])
def test_register_blocked_sizes_match_oracle(hc, backend, Ns):
    """Every register-blocked transform shape (last stage radix 2, 4, 8 or 16;
    free-space padded lengths 32 .. 2048, periodic 32 .. 1024) in FP64 and
    FP32 against the oracle's scipy.fft operators."""
    s = hc.ContactSolver(Ns, backend=backend)
    op = O.FreeSpaceOperator(Ns) if backend == "fft" else O.PeriodicOperator(Ns, 1.0, 1.0)
    x = np.random.default_rng(Ns + 7).uniform(0, 1, Ns * Ns)
    ref = op.matvec(x)
    y = s.matvec(x)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-12
    xf = _dev(x, np.float32)
    yd = _dev(np.zeros(Ns * Ns), np.float32)
    s.matvec_device(xf.ptr, yd.ptr, 32)
    yf = yd.numpy().astype(np.float64)
    assert np.linalg.norm(yf - ref) / np.linalg.norm(ref) < 1e-5


def test_free_space_matvec_matches_oracle_256(hc):
    Ns = 256
    s = hc.ContactSolver(Ns, domain_size=1.0, E_star=1.0, backend="fft")
    x = np.random.default_rng(0).uniform(0, 1, Ns * Ns)
    ref = O.FreeSpaceOperator(Ns).matvec(x)
    y = s.matvec(x)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13


def test_free_space_repeat_calls_bit_equal(hc):
    """tests/test_fft.cpp:46-59."""
    Ns = 32
    s = hc.ContactSolver(Ns, backend="fft")
    rng = np.random.default_rng(1)
    x, other = rng.random(Ns * Ns), rng.random(Ns * Ns)
    y1 = s.matvec(x)
    s.matvec(other)
    assert np.array_equal(y1, s.matvec(x))


def test_periodic_cos_modes(hc):
    """tests/test_periodic.cpp:16-36 (< 2e-13)."""
    n, L, E = 32, 3.0, 2.5
    s = hc.ContactSolver(n, domain_size=L, E_star=E, backend="periodic")
    iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    for mx, my in ((1, 0), (0, 3), (2, 5)):
        q = 2 * math.pi * math.hypot(mx, my) / L
        p = np.cos(2 * math.pi * (mx * ix + my * iy) / n).ravel()
        u = s.matvec(p)
        ref = 2.0 / (E * q) * p
        assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 2e-13


def test_periodic_dc_repeat_float(hc):
    """tests/test_periodic.cpp:38-51."""
    n = 32
    s = hc.ContactSolver(n, backend="periodic")
    assert np.linalg.norm(s.matvec(np.ones(n * n))) < 1e-13
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, n * n)
    y1 = s.matvec(x)
    s.matvec(rng.random(n * n))
    assert np.array_equal(y1, s.matvec(x))
    xf = _dev(x, np.float32)
    s.matvec_device(xf.ptr, xf.ptr, 32)   # aliased in place
    yf = xf.numpy().astype(np.float64)
    assert np.linalg.norm(yf - y1) / np.linalg.norm(y1) < 2e-6


@pytest.mark.parametrize("n", [7, 12, 16])
def test_periodic_matrix_properties_odd_grids(hc, n):
    """tests/test_periodic.cpp:56-88: symmetric, PSD, circulant; generic-size FFT."""
    s = hc.ContactSolver(n, domain_size=1.7, E_star=0.8, backend="periodic")
    N = n * n
    S = np.stack([s.matvec(np.eye(N)[j]) for j in range(N)], axis=1)
    assert np.linalg.norm(S - S.T) / np.linalg.norm(S) < 1e-14
    ev = np.linalg.eigvalsh(S)
    assert ev[0] > -1e-13 * ev[-1] and ev[1] > 1e-6 * ev[-1]
    ref = O.PeriodicOperator(n, 1.7, 0.8)
    x = np.random.default_rng(n).uniform(-1, 1, N)
    assert np.linalg.norm(s.matvec(x) - ref.matvec(x)) / np.linalg.norm(ref.matvec(x)) < 1e-13


@pytest.mark.parametrize("Ns", [24, 20, 64])
def test_periodic_and_free_space_generic_sizes_match_oracle(hc, Ns):
    x = np.random.default_rng(Ns).uniform(0, 1, Ns * Ns)
    for be, ref in (("periodic", O.PeriodicOperator(Ns)), ("fft", O.FreeSpaceOperator(Ns))):
        y = hc.ContactSolver(Ns, backend=be).matvec(x)
        r = ref.matvec(x)
        assert np.linalg.norm(y - r) / np.linalg.norm(r) < 1e-12, be


def test_periodic_2048_matches_oracle(hc):
    Ns = 2048
    x = np.random.default_rng(0).uniform(0, 1, Ns * Ns)
    y = hc.ContactSolver(Ns, backend="periodic").matvec(x)
    r = O.PeriodicOperator(Ns).matvec(x)
    assert np.linalg.norm(y - r) / np.linalg.norm(r) < 1e-13


def test_dense_backend_matches_oracle(hc):
    Ns = 12
    x = np.random.default_rng(5).uniform(0, 1, Ns * Ns)
    y = hc.ContactSolver(Ns, backend="dense").matvec(x)
    r = O.dense_matrix(Ns) @ x
    assert np.linalg.norm(y - r) / np.linalg.norm(r) < 1e-14


def test_stencil_taps_device_build(hc):
    """tests/test_stencil.cpp:36-58 on the library's own weight build."""
    from aspher_amd._capi import lib, check
    fx = json.load(open(os.path.join(GOLD, "stencil_taps.json")))
    for Ns in (512, 1024, 2048):
        nt = C.c_int()
        dx = (C.c_int * 25)()
        dy = (C.c_int * 25)()
        w = (C.c_double * 25)()
        check(lib().aspher_stencil_taps(Ns, 2, C.byref(nt), dx, dy, w))
        assert nt.value == 13
        ref = O.StencilPreconditioner(Ns, 2)
        for t in range(13):
            a, b = sorted((abs(dx[t]), abs(dy[t])), reverse=True)
            assert abs(w[t] - fx["taps"]["%d,%d" % (a, b)]) <= fx["tolerance"]
            assert (dx[t], dy[t]) == (ref.dx[t], ref.dy[t])
            assert abs(w[t] - ref.w[t]) < 1e-12


def _precond(kind, Ns, g, mask, precision=64, radius=2):
    from aspher_amd._capi import lib, check, dptr
    g = np.ascontiguousarray(g, dtype=np.float64)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    z = np.empty(Ns * Ns)
    check(lib().aspher_precond_apply(Ns, kind.encode(), radius, precision, dptr(g),
                                     m.ctypes.data_as(C.POINTER(C.c_uint8)), dptr(z)))
    return z


@pytest.mark.parametrize("Ns", [32, 256, 512])
def test_preconditioners_match_oracle(hc, Ns):
    """Ns = 512: the FP64 Fourier preconditioner's row workgroups have 352
    threads, so its masked per-block sums reduce over a partial last
    wavefront."""
    rng = np.random.default_rng(Ns)
    mask = rng.random(Ns * Ns) < 0.4
    g = rng.uniform(-1, 1, Ns * Ns)
    for kind, ref in (("fourier", O.StencilPreconditioner(Ns, 2)),
                      ("fourier-fft", O.FourierPreconditioner(Ns))):
        z = _precond(kind, Ns, g, mask)
        r = ref.apply(g, mask)
        assert np.linalg.norm(z - r) / np.linalg.norm(r) < 1e-12, kind
        zf = _precond(kind, Ns, g, mask, precision=32)
        rf = ref.apply(g.astype(np.float32), mask).astype(np.float64)
        assert np.linalg.norm(zf - rf) / np.linalg.norm(rf) < 1e-5, kind


@pytest.mark.parametrize("Ns,radius", [(8, 1), (16, 2), (96, 3), (200, 2), (40, 4), (64, 5), (300, 7)])
def test_stencil_small_and_ragged_grids(hc, Ns, radius):
    """Grids smaller than one 64x16 tile and grids that are not a multiple of
    it: the periodic halo wraps, the tile overhangs the grid."""
    rng = np.random.default_rng(Ns + radius)
    mask = rng.random(Ns * Ns) < 0.5
    g = rng.uniform(-1, 1, Ns * Ns)
    ref = O.StencilPreconditioner(Ns, radius)
    z = _precond("fourier", Ns, g, mask, radius=radius)
    r = ref.apply(g, mask)
    assert np.linalg.norm(z - r) / np.linalg.norm(r) < 1e-12
    zf = _precond("fourier", Ns, g, mask, precision=32, radius=radius)
    rf = ref.apply(g.astype(np.float32), mask).astype(np.float64)
    assert np.linalg.norm(zf - rf) / np.linalg.norm(rf) < 1e-5


def _disk_mask(Ns, rng, n_disks, rmax):
    """Contact patches: n_disks periodic disks (one across the corner), so most
    64x16 tiles of the grid hold no contact point."""
    iy, ix = np.meshgrid(np.arange(Ns), np.arange(Ns), indexing="ij")
    m = np.zeros((Ns, Ns), dtype=bool)
    centres = [(0, 0)] + [tuple(rng.integers(0, Ns, 2)) for _ in range(n_disks - 1)]
    for cy, cx in centres:
        r = rng.uniform(1.0, rmax)
        dy = (iy - cy + Ns // 2) % Ns - Ns // 2
        dx = (ix - cx + Ns // 2) % Ns - Ns // 2
        m |= dy * dy + dx * dx <= r * r
    return m.ravel()


@pytest.mark.parametrize("Ns,radius,n_disks", [(512, 2, 6), (200, 2, 3), (1024, 3, 12),
                                               (300, 7, 4)])
def test_stencil_clustered_masks_skip_empty_tiles(hc, Ns, radius, n_disks):
    """Masks of a few contact patches (ragged grids, a patch across the
    periodic corner, compile-time and run-time radius): tiles without a
    contact point are skipped by the stencil, so the result must still match
    the oracle and be exactly zero off the mask."""
    rng = np.random.default_rng(7 * Ns + radius)
    mask = _disk_mask(Ns, rng, n_disks, max(4.0, Ns / 24))
    tiles = mask.reshape(Ns, Ns)[: Ns // 16 * 16, : Ns // 64 * 64]
    empty = ~tiles.reshape(Ns // 16, 16, Ns // 64, 64).any(axis=(1, 3))
    assert empty.mean() > 0.5
    g = rng.uniform(-1, 1, Ns * Ns)
    ref = O.StencilPreconditioner(Ns, radius)
    r = ref.apply(g, mask)
    for precision, gate in ((64, 1e-12), (32, 1e-5)):
        z = _precond("fourier", Ns, g, mask, precision=precision, radius=radius)
        rr = r if precision == 64 else ref.apply(g.astype(np.float32), mask).astype(np.float64)
        assert np.linalg.norm(z - rr) / np.linalg.norm(rr) < gate, precision
        assert np.all(z[~mask] == 0.0), precision


def test_preconditioners_offset_invariant_symmetric(hc):
    """tests/test_precond.cpp:27-66."""
    Ns = 32
    N = Ns * Ns
    rng = np.random.default_rng(7)
    mask = rng.random(N) < 0.4
    chi = mask.astype(np.float64)
    g, x, y = (rng.uniform(-1, 1, N) for _ in range(3))
    for kind in ("fourier", "fourier-fft"):
        z0 = _precond(kind, Ns, g, mask)
        z1 = _precond(kind, Ns, g + 0.37 * chi, mask)
        assert np.abs(z1 - z0).max() / np.abs(z0).max() < 1e-12
        mx, my = _precond(kind, Ns, x, mask), _precond(kind, Ns, y, mask)
        xmy, ymx = (x * my)[mask].sum(), (y * mx)[mask].sum()
        assert abs(xmy - ymx) / (abs(xmy) + abs(ymx)) < 1e-12
