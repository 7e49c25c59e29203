"""Tangential (Cerruti) operator on the GPU.

TangentialFFTOperator (src/tangential_operator.cpp:16-157): the exact 2x2
block convolution u = [Kxx Kxy; Kxy Kyy] q of the CerrutiKernel tables
(src/cerruti_kernel.cpp) on stacked [q_x; q_y] (2 Ns^2 values), and the
tangential Fourier preconditioner (src/fourier_precond.cpp:234-359). The
reference exposes them to Python only inside its friction solver; here they
are plain objects over the C ABI.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._capi import check, dptr, lib


class TangentialFFTOperator:
    def __init__(self, grid_size, domain_size=1.0, E_star=1.0, nu=0.3):
        self._h = C.c_void_p()
        check(lib().aspher_tangential_create(int(grid_size), float(domain_size), float(E_star),
                                             float(nu), C.byref(self._h)))
        self.grid_size = int(grid_size)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().aspher_tangential_destroy(h)
            except Exception:
                pass
            self._h = C.c_void_p(None)

    def matvec(self, q) -> np.ndarray:
        n2 = 2 * self.grid_size * self.grid_size
        x = np.ascontiguousarray(np.asarray(q, dtype=np.float64)).ravel()
        if x.size != n2:
            raise ValueError("q has %d entries, expected %d" % (x.size, n2))
        u = np.empty(n2, dtype=np.float64)
        check(lib().aspher_tangential_matvec(self._h, dptr(x), dptr(u)))
        return u

    def matvec_device(self, q_ptr: int, u_ptr: int, precision: int = 64) -> None:
        check(lib().aspher_tangential_matvec_device(self._h, C.c_void_p(q_ptr),
                                                    C.c_void_p(u_ptr), int(precision)))


def tangential_precond_apply(grid_size, nu, g, mask, remove_mean=False, precision=64):
    n = int(grid_size) ** 2
    gv = np.ascontiguousarray(np.asarray(g, dtype=np.float64)).ravel()
    mv = np.ascontiguousarray(np.asarray(mask).astype(np.uint8)).ravel()
    if gv.size != 2 * n or mv.size != n:
        raise ValueError("g needs 2 Ns^2 values and mask Ns^2")
    z = np.empty(2 * n, dtype=np.float64)
    check(lib().aspher_tangential_precond_apply(
        int(grid_size), float(nu), int(precision), dptr(gv),
        mv.ctypes.data_as(C.POINTER(C.c_uint8)), int(bool(remove_mean)), dptr(z)))
    return z


def cerruti_points(x, y, a, b, nu):
    x, y, a, b = (np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64),
                                                        np.broadcast(x, y, a, b).shape)).ravel()
                  for v in (x, y, a, b))
    uxx, uxy = np.empty_like(x), np.empty_like(x)
    check(lib().aspher_cerruti_points(x.size, dptr(x), dptr(y), dptr(a), dptr(b), float(nu),
                                      dptr(uxx), dptr(uxy)))
    return uxx, uxy


def cerruti_tables(grid_size, domain_size=1.0, E_star=1.0, nu=0.3):
    xx = np.empty((grid_size, grid_size))
    xy = np.empty((grid_size, grid_size))
    check(lib().aspher_cerruti_tables(int(grid_size), float(domain_size), float(E_star),
                                      float(nu), dptr(xx), dptr(xy)))
    return xx, xy
