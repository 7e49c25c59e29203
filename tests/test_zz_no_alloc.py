"""The matvec without allocation (aspher_solver_matvec_device_noalloc), through
the C ABI, against the allocating device matvec (aspher_solver_matvec_device)
on the same handle and input. Both run the same kernels with the same launch
shapes, so the outputs must agree bit for bit; the static scratch keeps the
previous call's contents, so each case calls twice in a row, the second time
on a smaller grid."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hc():
    import aspher_amd
    assert aspher_amd.device_count() >= 1, "no GPU visible"
    return aspher_amd


def _lib():
    from aspher_amd._capi import lib
    return lib()


def _dev(a, dtype):
    from aspher_amd.device import DeviceArray
    return DeviceArray.from_numpy(np.asarray(a), dtype)


def _pair(hc, backend, Ns, prec, seed):
    """(allocating, no-allocation) outputs for one handle and input."""
    from aspher_amd._capi import check
    dt = np.float64 if prec == 64 else np.float32
    s = hc.ContactSolver(Ns, backend=backend)
    x = np.random.default_rng(seed).standard_normal(Ns * Ns)
    xd = _dev(x, dt)
    ya = _dev(np.zeros(Ns * Ns), dt)
    yn = _dev(np.full(Ns * Ns, np.nan), dt)
    check(_lib().aspher_solver_matvec_device(s._h, C.c_void_p(xd.ptr), C.c_void_p(ya.ptr),
                                             prec))
    print("noalloc %s Ns=%d FP%d seed %d" % (backend, Ns, prec, seed), flush=True)
    check(_lib().aspher_solver_matvec_device_noalloc(s._h, C.c_void_p(xd.ptr),
                                                     C.c_void_p(yn.ptr), prec))
    return ya.numpy(), yn.numpy()


@pytest.mark.parametrize("backend,sizes,prec", [
    ("fft", (2048, 256), 64),
    ("fft", (2048, 512), 32),
    ("periodic", (2048, 64), 64),
    ("periodic", (4096, 1024), 32),
])
def test_noalloc_matches_allocating_matvec(hc, backend, sizes, prec):
    for k, Ns in enumerate(sizes):
        ya, yn = _pair(hc, backend, Ns, prec, seed=10 * Ns + k)
        assert np.all(np.isfinite(yn))
        assert np.array_equal(ya, yn), (backend, Ns, prec,
                                        float(np.max(np.abs(ya - yn))))


def test_noalloc_in_place(hc):
    """x and y may alias, as in the allocating entry point."""
    from aspher_amd._capi import check
    Ns = 512
    s = hc.ContactSolver(Ns, backend="fft")
    x = np.random.default_rng(3).uniform(0, 1, Ns * Ns)
    ref = s.matvec(x)
    xd = _dev(x, np.float64)
    check(_lib().aspher_solver_matvec_device_noalloc(s._h, C.c_void_p(xd.ptr),
                                                     C.c_void_p(xd.ptr), 64))
    assert np.array_equal(xd.numpy(), ref)


def test_noalloc_rejections(hc):
    from aspher_amd import _capi
    L = _lib()
    # FP32 not prepared yet on a fresh handle
    s = hc.ContactSolver(256, backend="periodic")
    xd = _dev(np.zeros(256 * 256), np.float32)
    rc = L.aspher_solver_matvec_device_noalloc(s._h, C.c_void_p(xd.ptr), C.c_void_p(xd.ptr), 32)
    assert rc == _capi.ERR_LOGIC
    assert b"not prepared" in L.aspher_last_error()
    # beyond the static scratch: periodic FP64 at Ns = 4096 needs 129 MiB
    s2 = hc.ContactSolver(4096, backend="periodic")
    xd2 = _dev(np.zeros(4096 * 4096), np.float64)
    rc = L.aspher_solver_matvec_device_noalloc(s2._h, C.c_void_p(xd2.ptr), C.c_void_p(xd2.ptr),
                                               64)
    assert rc == _capi.ERR_INVALID
    assert b"72 MiB" in L.aspher_last_error()
    # dense backend, bad precision
    s3 = hc.ContactSolver(16, backend="dense")
    xd3 = _dev(np.zeros(256), np.float64)
    assert L.aspher_solver_matvec_device_noalloc(s3._h, C.c_void_p(xd3.ptr),
                                                 C.c_void_p(xd3.ptr), 64) != 0
    assert L.aspher_solver_matvec_device_noalloc(s._h, C.c_void_p(xd.ptr),
                                                 C.c_void_p(xd.ptr), 16) == _capi.ERR_INVALID
