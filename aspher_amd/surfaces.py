"""Synthetic inputs of the benchmark workloads (the reference draws its
surfaces from the third-party `rfgen` package, which is not available; these
follow the published recipe instead).

Self-affine Gaussian rough surface: random-phase spectrum with amplitude
|m|^-(1+H) (PSD ~ |q|^-2(1+H)) for integer wavenumbers |m| in [k_low, k_high],
zero elsewhere; Hermitian through the real part of the inverse FFT; mean
removed and scaled to the requested rms. The contact gap is -height.
"""
from __future__ import annotations

import math

import numpy as np


def rough_surface(Ns: int, H: float = 0.8, k_low: float = 4, k_high: float = 64,
                  rms: float = 0.002, seed: int = 0) -> np.ndarray:
    """(Ns, Ns) height field, row-major [iy, ix]."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(Ns, 1.0 / Ns)
    kk = np.hypot(k[:, None], k[None, :])
    kk[0, 0] = 1.0
    amp = np.where((kk >= k_low) & (kk <= k_high), kk ** (-(1.0 + H)), 0.0)
    h = np.real(np.fft.ifft2(amp * np.exp(2j * np.pi * rng.random((Ns, Ns)))))
    h -= h.mean()
    h *= rms / h.std()
    return np.ascontiguousarray(h)


def rough_gap(Ns: int, **kw) -> np.ndarray:
    """Flat gap g0 = -height of `rough_surface`."""
    return np.ascontiguousarray(-rough_surface(Ns, **kw).ravel())


def rough_surface_gpu(Ns: int, H: float = 0.8, k_low: float = 4, k_high: float = 64,
                      rms: float = 0.002, seed: int = 0) -> np.ndarray:
    """The same recipe generated on the GPU (libaspher_hip), with the phase of
    frequency (kx, ky) from a SplitMix64 hash of (seed, ky*Ns + kx) instead
    of numpy's generator. (Ns, Ns) float64 on the host."""
    import ctypes as C
    from ._capi import check, dptr, lib
    out = np.empty(Ns * Ns, dtype=np.float64)
    check(lib().aspher_rough_surface(int(Ns), float(H), float(k_low), float(k_high), float(rms),
                                     C.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF), dptr(out)))
    return out.reshape(Ns, Ns)


def rough_surface_to_device(dev_array, Ns: int, H: float = 0.8, k_low: float = 4,
                            k_high: float = 64, rms: float = 0.002, seed: int = 0) -> None:
    """Generate straight into a device buffer of Ns^2 float64 (a DeviceArray
    or a raw device address): the surface never leaves HBM."""
    import ctypes as C
    from ._capi import check, lib
    ptr = dev_array if isinstance(dev_array, int) else dev_array.ptr
    check(lib().aspher_rough_surface_device(int(Ns), float(H), float(k_low), float(k_high),
                                            float(rms),
                                            C.c_uint64(int(seed) & 0xFFFFFFFFFFFFFFFF),
                                            C.c_void_p(ptr)))


def westergaard_gap(Ns: int, L: float = 1.0, amplitude: float = 0.01) -> np.ndarray:
    """Single sinusoid g0 = D (1 - cos 2 pi x / L) along x (Westergaard 1939)."""
    x = np.arange(Ns) / Ns
    row = amplitude * (1.0 - np.cos(2.0 * math.pi * x))
    return np.ascontiguousarray(np.tile(row, (Ns, 1)).ravel())
