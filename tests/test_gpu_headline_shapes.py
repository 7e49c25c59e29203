"""GPU parity at the benchmark's own shapes, through the C ABI.

The free-space matvec at Ns = 2048 and 4096 (padded transform lengths 4096 and
8192: the n = 8192 column pass has the split FP64 exchange, the unpaired mirror
chunks of ky = 0 and n/2 and the folded last column round, which exist at no
smaller size), the periodic matvec at 2048 and 4096, and both preconditioners
at 4096, each in FP64 and FP32 against the oracle's scipy.fft restatement,
with the gates of tests/test_fft.cpp:35-36 (rel < 1e-12 FP64, < 1e-5 FP32).
The FP32 results are compared with the FP64 oracle, so the gate bounds the
FP32 pipeline's whole error, not its agreement with another FP32 code."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def hc():
    import aspher_amd
    assert aspher_amd.device_count() >= 1, "no GPU visible"
    O.WORKERS = 16
    return aspher_amd


def _dev(a, dtype):
    from aspher_amd.device import DeviceArray
    return DeviceArray.from_numpy(np.asarray(a), dtype)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _matvec_both(s, x):
    y64 = s.matvec(x)
    xf = _dev(x, np.float32)
    yd = _dev(np.zeros(x.size), np.float32)
    s.matvec_device(xf.ptr, yd.ptr, 32)
    return y64, yd.numpy().astype(np.float64)


@pytest.mark.parametrize("Ns", [2048, 4096])
def test_free_space_matvec_headline_sizes(hc, Ns):
    x = np.random.default_rng(Ns).uniform(0, 1, Ns * Ns)
    ref = O.FreeSpaceOperator(Ns).matvec(x)
    s = hc.ContactSolver(Ns, backend="fft")
    y64, y32 = _matvec_both(s, x)
    e64, e32 = _rel(y64, ref), _rel(y32, ref)
    print("free space Ns=%d: FP64 rel %.3e, FP32 rel %.3e" % (Ns, e64, e32), flush=True)
    assert e64 < 1e-12
    assert e32 < 1e-5
    # a signed input: cancellations the uniform input does not exercise
    x2 = np.random.default_rng(Ns + 1).standard_normal(Ns * Ns)
    ref2 = O.FreeSpaceOperator(Ns).matvec(x2)
    y64, y32 = _matvec_both(s, x2)
    assert _rel(y64, ref2) < 1e-12
    assert _rel(y32, ref2) < 1e-5


@pytest.mark.parametrize("Ns", [2048, 4096])
def test_periodic_matvec_headline_sizes(hc, Ns):
    x = np.random.default_rng(Ns + 2).uniform(-1, 1, Ns * Ns)
    ref = O.PeriodicOperator(Ns).matvec(x)
    y64, y32 = _matvec_both(hc.ContactSolver(Ns, backend="periodic"), x)
    e64, e32 = _rel(y64, ref), _rel(y32, ref)
    print("periodic Ns=%d: FP64 rel %.3e, FP32 rel %.3e" % (Ns, e64, e32), flush=True)
    assert e64 < 1e-12
    assert e32 < 1e-5


def _precond(kind, Ns, g, mask, precision):
    import ctypes as C
    from aspher_amd._capi import lib, check, dptr
    g = np.ascontiguousarray(g, dtype=np.float64)
    m = np.ascontiguousarray(mask, dtype=np.uint8)
    z = np.empty(Ns * Ns)
    check(lib().aspher_precond_apply(Ns, kind.encode(), 2, precision, dptr(g),
                                     m.ctypes.data_as(C.POINTER(C.c_uint8)), dptr(z)))
    return z


@pytest.mark.parametrize("occupancy", [0.05, 0.6])
def test_preconditioners_4096(hc, occupancy):
    """Both preconditioners at 4096^2 on a sparse and a dense contact mask."""
    Ns = 4096
    rng = np.random.default_rng(int(occupancy * 100))
    mask = rng.random(Ns * Ns) < occupancy
    g = rng.uniform(-1, 1, Ns * Ns)
    for kind, ref in (("fourier", O.StencilPreconditioner(Ns, 2)),
                      ("fourier-fft", O.FourierPreconditioner(Ns))):
        r = ref.apply(g, mask)
        z = _precond(kind, Ns, g, mask, 64)
        z32 = _precond(kind, Ns, g, mask, 32)
        e64, e32 = _rel(z, r), _rel(z32, r)
        print("%s occupancy %.2f: FP64 rel %.3e, FP32 rel %.3e" % (kind, occupancy, e64, e32),
              flush=True)
        assert e64 < 1e-12, kind
        assert e32 < 1e-5, kind
        assert np.all(z[~mask] == 0.0) and np.all(z32[~mask] == 0.0)


def test_fourier_preconditioner_cos_modes_4096(hc):
    """On a full mask a cos mode m is an eigenvector of the Fourier
    preconditioner with eigenvalue |kof(m)| (symbol |k|/Ns^2 times the
    unnormalised inverse's Ns^2), including the Nyquist rows and columns,
    where kof(Ns/2) = -Ns/2 (src/fourier_precond.cpp:12-28). The symbol is
    stored in float, so the FP64 gate is the float rounding of |k| (6e-8),
    relative to the eigenvalue.

    In FP32 the forward transform's rounding spreads about eps32 * ||g|| over
    every mode, and the symbol multiplies each by up to max|kof| = Ns/sqrt(2),
    so the FP32 error is bounded relative to the operator norm, not to a
    small eigenvalue: ||z - lam g|| / (max|kof| ||g||) < 2e-6."""
    Ns = 4096
    mask = np.ones(Ns * Ns, dtype=bool)
    iy, ix = np.meshgrid(np.arange(Ns), np.arange(Ns), indexing="ij")
    norm_op = math.hypot(Ns // 2, Ns // 2)
    for mx, my in ((1, 0), (3, 7), (Ns // 2, 0), (0, Ns // 2), (Ns // 2, Ns // 2),
                   (Ns // 2 - 1, 5)):
        g = np.cos(2 * math.pi * (mx * ix + my * iy) / Ns).ravel()
        lam = math.hypot(mx, my)
        z = _precond("fourier-fft", Ns, g, mask, 64)
        e64 = _rel(z, lam * g)
        assert e64 < 1e-7, (mx, my, 64, e64)
        z32 = _precond("fourier-fft", Ns, g, mask, 32)
        e32 = float(np.linalg.norm(z32 - lam * g) / (norm_op * np.linalg.norm(g)))
        assert e32 < 2e-6, (mx, my, 32, e32)
