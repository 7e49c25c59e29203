"""Tangential (Cerruti) operator and preconditioner on the GPU against the
oracle (pinned by quadrature, closed forms, the dense block and symbol
checks in tests/test_oracle.py) and the reference's own gates
(tests/test_cerruti.cpp, tests/test_tangential.cpp)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def tg():
    import aspher_amd
    assert aspher_amd.device_count() >= 1, "no GPU visible"
    from aspher_amd import tangential
    return tangential


def test_device_cerruti_brackets(tg):
    rng = np.random.default_rng(3)
    x = rng.uniform(-0.3, 0.3, 400)
    y = rng.uniform(-0.3, 0.3, 400)
    x[:5] = [0.0, 0.0, 0.01, 2000 * 0.01, 0.0]
    y[:5] = [0.0, 0.02, 0.0, 3.0, 0.0]
    a = np.full(400, 0.01)
    b = np.full(400, 0.013)
    uxx, uxy = tg.cerruti_points(x, y, a, b, 0.3)
    oxx = O.cerruti_uxx(x, y, a, b, 0.3)
    oxy = O.cerruti_uxy(x, y, a, b)
    assert np.abs(uxx / oxx - 1).max() < 1e-12
    nz = oxy != 0
    assert np.abs(uxy[nz] / oxy[nz] - 1).max() < 1e-12 and np.all(uxy[~nz] == 0)
    xx, xy = tg.cerruti_tables(64, 1.0, 1.7, 0.3)
    C = O.CerrutiTables(64, 1.0, 1.7, 0.3)
    assert np.abs(xx / C.xx - 1).max() < 1e-12
    m = C.xy != 0
    assert np.abs(xy[m] / C.xy[m] - 1).max() < 1e-12


@pytest.mark.parametrize("Ns", [16, 32])
def test_tangential_fft_equals_dense(tg, Ns):
    """tests/test_tangential.cpp:27-40: the FFT block matvec reproduces the
    dense CerrutiKernel block to roundoff (< 1e-13)."""
    op = tg.TangentialFFTOperator(Ns, 1.0, 1.7, 0.3)
    D = O.CerrutiTables(Ns, 1.0, 1.7, 0.3).dense()
    q = np.random.default_rng(7 + Ns).standard_normal(2 * Ns * Ns)
    u = op.matvec(q)
    ref = D @ q
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-13


@pytest.mark.parametrize("Ns", [64, 256, 1024])
def test_tangential_matches_oracle_fft(tg, Ns):
    op = tg.TangentialFFTOperator(Ns, 1.3, 2.1, 0.25)
    ref_op = O.TangentialOperator(Ns, 1.3, 2.1, 0.25)
    q = np.random.default_rng(Ns).uniform(-1, 1, 2 * Ns * Ns)
    u = op.matvec(q)
    ref = ref_op.matvec(q)
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-12
    # FP32 path
    from aspher_amd.device import DeviceArray
    x = DeviceArray.from_numpy(q, np.float32)
    y = DeviceArray(2 * Ns * Ns, np.float32)
    op.matvec_device(x.ptr, y.ptr, 32)
    uf = y.numpy().astype(np.float64)
    assert np.linalg.norm(uf - ref) / np.linalg.norm(ref) < 1e-5


def test_tangential_symbol_modes(tg):
    """Longitudinal / transverse cos modes: eigenvalues near the continuum
    symbol 2/(E*|q|) and 2/(E*(1-nu)|q|) (tests/test_cerruti.cpp:145-190)."""
    Ns, E, nu = 128, 1.7, 0.3
    op = tg.TangentialFFTOperator(Ns, 1.0, E, nu)
    x = np.arange(Ns) / Ns
    X, Y = np.meshgrid(x, x, indexing="xy")
    m = 8
    c = np.cos(2 * math.pi * m * X).ravel()
    z = np.zeros_like(c)
    ul = op.matvec(np.concatenate([c, z]))     # longitudinal: q along k (x)
    ut = op.matvec(np.concatenate([z, c]))     # transverse: q normal to k
    kk = 2 * math.pi * m
    N = Ns * Ns
    lam_l = (ul[:N] @ c) / (c @ c)
    lam_t = (ut[N:] @ c) / (c @ c)
    # the finite patch sits a few percent below the continuum (0.979 and
    # 0.970 for this grid in the oracle)
    assert abs(lam_l / (2 / (E * kk)) - 1) < 0.05
    assert abs(lam_t / (2 / (E * (1 - nu) * kk)) - 1) < 0.05


@pytest.mark.parametrize("prec,tol", [(64, 1e-12), (32, 1e-5)])
def test_tangential_preconditioner_matches_oracle(tg, prec, tol):
    Ns, nu = 64, 0.3
    rng = np.random.default_rng(5)
    mask = rng.random(Ns * Ns) < 0.5
    g = rng.standard_normal(2 * Ns * Ns)
    P = O.TangentialPreconditioner(Ns, nu)
    for rm in (False, True):
        z = tg.tangential_precond_apply(Ns, nu, g, mask, remove_mean=rm, precision=prec)
        ref = P.apply(g, mask, remove_mean=rm)
        assert np.linalg.norm(z - ref) / np.linalg.norm(ref) < tol, (prec, rm)


def test_tangential_rejects_bad_input(tg):
    with pytest.raises(ValueError):
        tg.TangentialFFTOperator(24)
    with pytest.raises(ValueError):
        tg.TangentialFFTOperator(64, nu=1.0)
    op = tg.TangentialFFTOperator(16)
    with pytest.raises(ValueError):
        op.matvec(np.zeros(16 * 16))
