"""The reference's Python API (module ``aspher``) for the spectral path,
backed by the HIP library through the C ABI.

Signatures, defaults, property names and error classes follow the pybind11
layer of the reference (python/bindings.cpp:637-879): ``ContactSolver``,
``solve_nested``, ``ContactResult`` and ``configure_allocator``. Backends
"fft", "periodic" and "dense" run on the GPU; the reference's "h2" and
"hmatrix" backends are not part of this library and raise ValueError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import check, dptr, lib


def _flat(a, expected: int) -> np.ndarray:
    """numpy (N,) or (Ns, Ns) -> C-contiguous float64 of `expected` entries
    (python/bindings.cpp:33-43)."""
    v = np.ascontiguousarray(np.asarray(a, dtype=np.float64)).ravel()
    if v.size != expected:
        raise ValueError("array has %d entries, expected %d" % (v.size, expected))
    return v


class ContactResult:
    """Read-only view of one solve (python/bindings.cpp:637-802)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        info = _capi.ResultInfo()
        check(lib().aspher_result_get_info(self._h, C.byref(info)))
        self._i = info
        self._Ns = info.grid_size
        self._cache = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().aspher_result_destroy(h)
            except Exception:
                pass
            self._h = C.c_void_p(None)

    def _field(self, which: int, present: bool):
        if not present:
            return None
        if which not in self._cache:
            out = np.empty(self._Ns * self._Ns, dtype=np.float64)
            check(lib().aspher_result_get_field(self._h, which, dptr(out)))
            self._cache[which] = out.reshape(self._Ns, self._Ns)
        return self._cache[which].copy()

    def field_device_ptr(self, which: int) -> int:
        """Device address of pressure (0), displacement (1) or gap (2), FP64,
        valid while this result is alive."""
        p = C.c_void_p()
        check(lib().aspher_result_field_device(self._h, which, C.byref(p)))
        return p.value

    pressure = property(lambda s: s._field(0, True))
    displacement = property(lambda s: s._field(1, bool(s._i.has_displacement)))
    gap = property(lambda s: s._field(2, bool(s._i.has_gap)))
    objective = property(lambda s: s._i.objective)
    contact_area = property(lambda s: s._i.contact_area)
    mean_pressure = property(lambda s: s._i.mean_pressure)
    approach = property(lambda s: s._i.approach)
    iterations = property(lambda s: s._i.iterations)
    error = property(lambda s: s._i.error)
    converged = property(lambda s: bool(s._i.converged))
    active_rounds = property(lambda s: s._i.active_rounds)
    active_fallback = property(lambda s: bool(s._i.active_fallback))
    status = property(lambda s: s._i.status.decode())
    status_reason = property(lambda s: s._i.status_reason.decode())
    pk_error = property(lambda s: s._i.pk_error)
    fw_gap = property(lambda s: s._i.fw_gap)
    fw_error = property(lambda s: s._i.fw_error)
    load_error = property(lambda s: s._i.load_error)
    pressure_violation = property(lambda s: s._i.pressure_violation)
    penetration_error = property(lambda s: s._i.penetration_error)
    requested_tol = property(lambda s: s._i.requested_tol)
    effective_tol = property(lambda s: s._i.effective_tol)
    p_ref = property(lambda s: s._i.p_ref)
    g_ref = property(lambda s: s._i.g_ref)
    validation_scope = property(lambda s: s._i.validation_scope.decode())
    operator_error_kind = property(lambda s: s._i.operator_error_kind.decode())
    operator_error = property(lambda s: None)          # "unavailable" is not zero
    objective_error_bound = property(lambda s: None)
    matvec_count = property(lambda s: s._i.matvec_count)
    verification_matvec_count = property(lambda s: s._i.verification_matvec_count)
    precond_count = property(lambda s: s._i.precond_count)
    identification_steps = property(lambda s: s._i.identification_steps)
    returned_best = property(lambda s: bool(s._i.returned_best))
    precond_dropped = property(lambda s: bool(s._i.precond_dropped))

    @property
    def error_history(self) -> np.ndarray:
        out = np.empty(self._i.n_history, dtype=np.float64)
        if out.size:
            check(lib().aspher_result_get_history(self._h, dptr(out)))
        return out

    @property
    def stage_stats(self) -> list:
        out = []
        for k in range(self._i.n_stages):
            st = _capi.StageInfo()
            check(lib().aspher_result_get_stage(self._h, k, C.byref(st)))
            out.append(dict(name=st.name.decode(), precision=st.precision.decode(), q=st.q,
                            requested_tol=st.requested_tol, effective_tol=st.effective_tol,
                            iterations=st.iterations, matvec_count=st.matvec_count,
                            seconds=st.seconds, status=st.status.decode(),
                            fw_error=st.fw_error, penetration_error=st.penetration_error,
                            precond_dropped=bool(st.precond_dropped)))
        return out

    @property
    def memory(self) -> dict:
        i = self._i
        return dict(cg_state=i.mem_cg_state, best_iterate=i.mem_best_iterate,
                    contact_mask=i.mem_contact_mask, output_arrays=i.mem_output_arrays,
                    peak=i.mem_peak)

    @property
    def timings(self) -> dict:
        """Seconds per phase; ``matvec`` and ``precond`` are GPU time of every
        eighth apply (event-timed), scaled to all applies."""
        i = self._i
        return dict(total=i.t_total, matvec=i.t_matvec, precond=i.t_precond, build=i.t_build,
                    coarse=i.t_coarse, verification=i.t_verification,
                    candidate=i.t_candidate, output=i.t_output)

    def __repr__(self):
        return ("<ContactResult: %s in %d iters, fw_error %g, penetration %g, contact area "
                "%g, mean p %g>" % (self.status, self.iterations, self.fw_error,
                                   self.penetration_error, self.contact_area,
                                   self.mean_pressure))


class ContactSolver:
    """ContactSolver (python/bindings.cpp:115-358, signature :804-834)."""

    def __init__(self, grid_size, domain_size=1.0, E_star=1.0, eta=2.0, aca_tol=1e-6,
                 leaf_size=64, use_hmatrix=True, use_acagp=False, central_fraction=0.3,
                 inline_svd_tol=0.0, backend="", q=4, near_radius=1, h2_leaf_side=8,
                 shared_operator=False):
        """``shared_operator`` (not in the reference) takes the operator from
        the process cache that ``solve_nested(reuse_operators=True)`` fills,
        so an adjoint runs on the forward solve's own kernel spectrum."""
        grid_size = int(grid_size)
        if grid_size <= 0:
            raise ValueError("grid_size must be positive")
        if not (domain_size > 0.0):
            raise ValueError("domain_size must be positive")
        if not (E_star > 0.0):
            raise ValueError("E_star must be positive")
        if not backend:
            backend = "hmatrix" if use_hmatrix else "dense"
        self._h = C.c_void_p()
        create = (lib().aspher_solver_create_shared if shared_operator
                  else lib().aspher_solver_create)
        check(create(grid_size, float(domain_size), float(E_star), backend.encode(),
                     C.byref(self._h)))
        self.grid_size = grid_size
        self.backend = backend
        self.domain_size = float(domain_size)
        self.E_star = float(E_star)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().aspher_solver_destroy(h)
            except Exception:
                pass
            self._h = C.c_void_p(None)

    @property
    def N(self) -> int:
        return self.grid_size * self.grid_size

    def matvec(self, p) -> np.ndarray:
        """u = S p; accepts (N,) or (Ns, Ns); returns (N,) (python/bindings.cpp:172-184)."""
        x = _flat(p, self.N)
        y = np.empty(self.N, dtype=np.float64)
        check(lib().aspher_solver_matvec(self._h, dptr(x), dptr(y)))
        return y

    def matvec_device(self, x_ptr: int, y_ptr: int, precision: int = 64) -> None:
        """u = S p on device buffers of Ns^2 values (FP64 or FP32); x, y may alias."""
        check(lib().aspher_solver_matvec_device(self._h, C.c_void_p(x_ptr),
                                                C.c_void_p(y_ptr), int(precision)))

    def time_matvec(self, x_ptr: int, y_ptr: int, precision: int = 64, reps: int = 20) -> float:
        """Average device milliseconds of `reps` back-to-back matvecs."""
        ms = C.c_double()
        check(lib().aspher_solver_time_matvec(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                              int(precision), int(reps), C.byref(ms)))
        return ms.value

    def time_stages(self, x_ptr: int, y_ptr: int, precision: int = 64, reps: int = 20):
        """Average device milliseconds of the row-forward, column and
        row-inverse launches of one matvec."""
        ms = (C.c_double * 3)()
        check(lib().aspher_solver_time_stages(self._h, C.c_void_p(x_ptr), C.c_void_p(y_ptr),
                                              int(precision), int(reps), ms))
        return [ms[0], ms[1], ms[2]]

    def solve(self, gap, p_nominal, tol=1e-8, max_iter=5000, use_pr=True, precond="none",
              p_init=None) -> ContactResult:
        """python/bindings.cpp:186-234 (defaults :818-828)."""
        g = _flat(gap, self.N)
        p0 = None if p_init is None else _flat(p_init, self.N)
        out = C.c_void_p()
        check(lib().aspher_solver_solve(self._h, dptr(g), float(p_nominal), float(tol),
                                        int(max_iter), int(bool(use_pr)),
                                        str(precond).encode(),
                                        None if p0 is None else dptr(p0), C.byref(out)))
        return ContactResult(out.value)

    def solve_device(self, gap_ptr: int, p_nominal, tol=1e-8, max_iter=5000, use_pr=True,
                     precond="none", p_init_ptr: int = 0) -> ContactResult:
        """solve() with the gap already resident on the device (FP64)."""
        out = C.c_void_p()
        check(lib().aspher_solver_solve_device(self._h, C.c_void_p(gap_ptr), float(p_nominal),
                                               float(tol), int(max_iter), int(bool(use_pr)),
                                               str(precond).encode(),
                                               C.c_void_p(p_init_ptr or None), C.byref(out)))
        return ContactResult(out.value)

    def adjoint(self, result: ContactResult, grad_pressure, tol=1e-10, max_iter=5000,
                precond="auto"):
        """Backward pass of a converged solve on this grid (no reference
        counterpart): for dL/dp = grad_pressure returns (dL/dgap as (Ns, Ns),
        dL/dp_nominal, info dict). Solves [S_CC 1; 1' 0][mu; nu] = [w_C; 0]
        on the support C = {p > 0} by preconditioned CG with this solver's
        operator; dL/dgap = -mu on C, 0 elsewhere; dL/dp_nominal = nu Ns^2."""
        w = _flat(grad_pressure, self.N)
        g = np.empty(self.N, dtype=np.float64)
        prm, info = _adjoint_params(tol, max_iter, precond), _capi.AdjointInfo()
        check(lib().aspher_solver_adjoint(self._h, result._h, dptr(w), C.byref(prm),
                                          dptr(g), C.byref(info)))
        return (g.reshape(self.grid_size, self.grid_size), info.grad_p_nominal,
                _adjoint_info(info))

    def adjoint_device(self, result: ContactResult, grad_pressure_ptr: int,
                       grad_gap_ptr: int, tol=1e-10, max_iter=5000, precond="auto"):
        """adjoint() on device FP64 buffers; returns (dL/dp_nominal, info)."""
        prm, info = _adjoint_params(tol, max_iter, precond), _capi.AdjointInfo()
        check(lib().aspher_solver_adjoint_device(self._h, result._h,
                                                 C.c_void_p(grad_pressure_ptr), C.byref(prm),
                                                 C.c_void_p(grad_gap_ptr), C.byref(info)))
        return info.grad_p_nominal, _adjoint_info(info)

    def hmatrix_info(self) -> dict:
        """Operator statistics (python/bindings.cpp:307-328 for fft/periodic)."""
        info = _capi.OperatorInfo()
        check(lib().aspher_solver_info(self._h, C.byref(info)))
        b = info.backend.decode()
        if b == "fft":
            print("FFTOperator: Ns=%d (N=%d), padded %dx%d, spectrum %.1f MiB, scratch %.1f "
                  "MiB, total %.1f MiB" % (self.grid_size, info.n, info.padded_side,
                                           info.padded_side, info.bytes_spectrum / 1048576.0,
                                           info.bytes_scratch / 1048576.0,
                                           info.bytes / 1048576.0))
            return dict(backend="fft", n=info.n, padded_side=info.padded_side,
                        bytes=info.bytes, bytes_spectrum=info.bytes_spectrum,
                        bytes_scratch=info.bytes_scratch,
                        compression=info.bytes / (8.0 * info.n * info.n))
        if b == "periodic":
            return dict(backend="periodic", n=info.n, periodic_x=True, periodic_y=True,
                        bytes_scratch=info.bytes_scratch)
        print("dense influence matrix,", info.n, "x", info.n)
        return dict(dense=True, bytes=8 * info.n * info.n)

    def block_layout(self) -> np.ndarray:
        """H-matrix block layout: empty for the spectral and dense backends."""
        return np.zeros((0, 6))

    def recompress(self, svd_tol) -> None:
        """H-matrix recompression: nothing to do for the spectral backends."""
        return None


def _adjoint_params(tol, max_iter, precond) -> _capi.AdjointParams:
    p = _capi.AdjointParams()
    lib().aspher_adjoint_params_default(C.byref(p))
    p.tol, p.max_iter = float(tol), int(max_iter)
    p.precond = str(precond).encode()
    return p


def _adjoint_info(i) -> dict:
    return dict(converged=bool(i.converged), iterations=i.iterations, n_contact=i.n_contact,
                matvec_count=i.matvec_count, precond_count=i.precond_count,
                rel_residual=i.rel_residual, multiplier=i.multiplier,
                grad_p_nominal=i.grad_p_nominal, seconds=i.seconds,
                status=i.status.decode(), precond=i.precond.decode())


_PRECOND_ENGINES = ("auto", "stencil", "fft")
_PRECISIONS = ("", "double", "float", "float_then_double")


def _nested_params(coarsest, q, leaf_side, precond, precond_engine, precond_radius,
                   precond_occupancy_max, precond_cost_gate, coarse_tol, single_precision,
                   precision, float_floor, allow_tolerance_relaxation, light_result, backend,
                   record_error_history, active_set, active_delta, active_halo,
                   active_max_rounds, active_all_levels, active_occupancy_max,
                   reuse_operators, graph_max_points=0) -> _capi.NestedParams:
    if precond_engine not in _PRECOND_ENGINES:
        raise ValueError("precond_engine must be 'auto', 'stencil' or 'fft'")
    if precision not in _PRECISIONS:
        raise ValueError("precision must be 'double', 'float', 'float_then_double' or '' "
                         "(defer to single_precision)")
    p = _capi.NestedParams()
    lib().aspher_nested_params_default(C.byref(p))
    p.coarsest, p.q, p.leaf_side = int(coarsest), int(q), int(leaf_side)
    p.precond = int(bool(precond))
    p.precond_engine = precond_engine.encode()
    p.precond_radius = int(precond_radius)
    p.precond_occupancy_max = float(precond_occupancy_max)
    p.precond_cost_gate = float(precond_cost_gate)
    p.coarse_tol = float(coarse_tol)
    p.single_precision = int(bool(single_precision))
    p.precision = precision.encode()
    p.float_floor = float(float_floor)
    p.allow_tolerance_relaxation = int(bool(allow_tolerance_relaxation))
    p.light_result = int(bool(light_result))
    p.backend = str(backend).encode()
    p.record_error_history = int(bool(record_error_history))
    p.active_set = int(bool(active_set))
    p.active_delta = float(active_delta)
    p.active_halo = int(active_halo)
    p.active_max_rounds = int(active_max_rounds)
    p.active_all_levels = int(bool(active_all_levels))
    p.active_occupancy_max = float(active_occupancy_max)
    p.reuse_operators = int(bool(reuse_operators))
    p.graph_max_points = int(graph_max_points)
    return p


def solve_nested(grid_size, gap, p_nominal, domain_size=1.0, E_star=1.0, coarsest=64, q=6,
                 leaf_side=8, precond=True, precond_engine="auto", precond_radius=2,
                 precond_occupancy_max=0.4, precond_cost_gate=0.0, tol=1e-8, coarse_tol=1e-4,
                 max_iter=20000, use_pr=True, single_precision=False, precision="",
                 float_floor=2e-6, allow_tolerance_relaxation=False, light_result=False,
                 backend="h2", record_error_history=False, active_set=False,
                 active_delta=0.05, active_halo=2, active_max_rounds=5,
                 active_all_levels=True, active_occupancy_max=0.4,
                 reuse_operators=False, graph_max_points=0) -> ContactResult:
    """Nested-grid (cascadic) solve (python/bindings.cpp:362-432, defaults :836-853).

    Not in the reference: ``reuse_operators`` keeps each level's built
    operator in a process cache so repeated solves skip the table build;
    levels of at most ``graph_max_points`` points start on the device loop
    (one captured iteration replayed as a hipGraph; 0, the default, launches
    every kernel, with bit-identical results)."""
    grid_size = int(grid_size)
    p = _nested_params(coarsest, q, leaf_side, precond, precond_engine, precond_radius,
                       precond_occupancy_max, precond_cost_gate, coarse_tol,
                       single_precision, precision, float_floor, allow_tolerance_relaxation,
                       light_result, backend, record_error_history, active_set, active_delta,
                       active_halo, active_max_rounds, active_all_levels,
                       active_occupancy_max, reuse_operators, graph_max_points)
    g = _flat(gap, grid_size * grid_size)
    out = C.c_void_p()
    check(lib().aspher_solve_nested(grid_size, dptr(g), float(p_nominal), float(domain_size),
                                    float(E_star), float(tol), int(max_iter),
                                    int(bool(use_pr)), C.byref(p), C.byref(out)))
    return ContactResult(out.value)


def solve_nested_device(grid_size, gap_ptr: int, p_nominal, tol=1e-8, max_iter=20000,
                        use_pr=True, domain_size=1.0, E_star=1.0, **kw) -> ContactResult:
    """solve_nested with the FP64 gap already resident on the device."""
    args = dict(coarsest=64, q=6, leaf_side=8, precond=True, precond_engine="auto",
                precond_radius=2, precond_occupancy_max=0.4, precond_cost_gate=0.0,
                coarse_tol=1e-4, single_precision=False, precision="", float_floor=2e-6,
                allow_tolerance_relaxation=False, light_result=False, backend="h2",
                record_error_history=False, active_set=False, active_delta=0.05,
                active_halo=2, active_max_rounds=5, active_all_levels=True,
                active_occupancy_max=0.4, reuse_operators=False, graph_max_points=0)
    unknown = set(kw) - set(args)
    if unknown:
        raise TypeError("unexpected keyword arguments: %s" % sorted(unknown))
    args.update(kw)
    p = _nested_params(**args)
    out = C.c_void_p()
    check(lib().aspher_solve_nested_device(int(grid_size), C.c_void_p(gap_ptr),
                                           float(p_nominal), float(domain_size), float(E_star),
                                           float(tol), int(max_iter), int(bool(use_pr)),
                                           C.byref(p), C.byref(out)))
    return ContactResult(out.value)


_BACKEND_CODE = {"fft": 0, "periodic": 1}


def _ensemble_call(problems, on_device, tol, max_iter, use_pr, precision, precond,
                   check_every, use_graph, autotune, keep_best, stall_limit, light_result):
    n = len(problems)
    if n == 0:
        raise ValueError("solve_ensemble: no problems")
    kinds, sizes, Ls, Es, pns, keep = [], [], [], [], [], []
    gaps = (C.POINTER(C.c_double) * n)() if not on_device else (C.c_void_p * n)()
    for i, pr in enumerate(problems):
        b = pr.get("backend", "fft")
        if b not in _BACKEND_CODE:
            raise ValueError("ensemble backend must be 'fft' or 'periodic', got %r" % (b,))
        Ns = int(pr["grid_size"])
        kinds.append(_BACKEND_CODE[b])
        sizes.append(Ns)
        Ls.append(float(pr.get("domain_size", 1.0)))
        Es.append(float(pr.get("E_star", 1.0)))
        pns.append(float(pr["p_nominal"]))
        if on_device:
            gaps[i] = C.c_void_p(int(pr["gap_ptr"]))
        else:
            g = _flat(pr["gap"], Ns * Ns)
            keep.append(g)
            gaps[i] = dptr(g)
    prm = _capi.EnsembleParams()
    lib().aspher_ensemble_params_default(C.byref(prm))
    prm.tol, prm.max_iter, prm.use_pr = float(tol), int(max_iter), int(bool(use_pr))
    prm.precision, prm.precond = str(precision).encode(), str(precond).encode()
    prm.check_every, prm.use_graph = int(check_every), int(bool(use_graph))
    prm.autotune, prm.keep_best = int(bool(autotune)), int(bool(keep_best))
    prm.stall_limit, prm.light_result = int(stall_limit), int(bool(light_result))
    ia = lambda v: (C.c_int * n)(*v)          # noqa: E731
    da = lambda v: (C.c_double * n)(*v)       # noqa: E731
    res = (C.c_void_p * n)()
    st = _capi.EnsembleStats()
    secs = (C.c_double * n)()
    fn = lib().aspher_solve_ensemble_device if on_device else lib().aspher_solve_ensemble
    check(fn(n, ia(kinds), ia(sizes), da(Ls), da(Es), da(pns), gaps, C.byref(prm), res,
             C.byref(st), secs))
    results = [ContactResult(res[i]) for i in range(n)]
    stats = {f[0]: getattr(st, f[0]) for f in _capi.EnsembleStats._fields_}
    stats["graph"] = bool(stats["graph"])
    stats["instance_seconds"] = [secs[i] for i in range(n)]
    return results, stats


def solve_ensemble(problems, tol=1e-5, max_iter=5000, use_pr=True, precision="float",
                   precond="auto", check_every=8, use_graph=True, autotune=True,
                   keep_best=True, stall_limit=200, light_result=False):
    """Many independent certified solves advanced together on the GPU (no
    reference counterpart; each instance is ContactSolver(backend).solve).

    ``problems``: list of dicts with ``grid_size``, ``gap`` ((N,) or (Ns, Ns)),
    ``p_nominal`` and optionally ``backend`` ("fft" | "periodic"),
    ``domain_size``, ``E_star``. Returns (list of ContactResult in input
    order, stats dict). FP32 ensembles certify at max(tol, 2e-6).
    ``precond``: "auto" (the full-grid Fourier |q| preconditioner for both
    kinds), "fourier" (13-tap stencil), "fourier-fft" or "none"."""
    return _ensemble_call(problems, False, tol, max_iter, use_pr, precision, precond,
                          check_every, use_graph, autotune, keep_best, stall_limit,
                          light_result)


def solve_ensemble_device(problems, tol=1e-5, max_iter=5000, use_pr=True, precision="float",
                          precond="auto", check_every=8, use_graph=True, autotune=True,
                          keep_best=True, stall_limit=200, light_result=False):
    """solve_ensemble with each gap already resident on the device: problems
    carry ``gap_ptr`` (FP64 device address) instead of ``gap``."""
    return _ensemble_call(problems, True, tol, max_iter, use_pr, precision, precond,
                          check_every, use_graph, autotune, keep_best, stall_limit,
                          light_result)


def configure_allocator(mmap_threshold_bytes=128 * 1024, trim_threshold_bytes=128 * 1024):
    """Host allocator policy (reference: glibc mallopt). Device memory is
    managed by the library; nothing is changed, returns False."""
    applied = C.c_int(0)
    check(lib().aspher_configure_allocator(int(mmap_threshold_bytes),
                                           int(trim_threshold_bytes), C.byref(applied)))
    return bool(applied.value)


def clear_operator_cache() -> None:
    check(lib().aspher_clear_cache())


def device_count() -> int:
    n = C.c_int(0)
    check(lib().aspher_device_count(C.byref(n)))
    return n.value
