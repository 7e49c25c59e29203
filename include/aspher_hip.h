/*
 * C ABI of the MI355X spectral contact library (libaspher_hip.so).
 *
 * Plain pointers, sizes and status codes; no C++ or framework types cross
 * this boundary. Each entry point replaces one call of the reference's
 * pybind11 layer (cofebem/aspher python/bindings.cpp), cited per function,
 * and keeps its argument meaning, defaults and error behaviour:
 *   invalid input (std::invalid_argument -> ValueError)  ASPHER_ERR_INVALID
 *   lifecycle misuse (std::logic_error -> RuntimeError)    ASPHER_ERR_LOGIC
 *   broken construction (std::runtime_error)               ASPHER_ERR_RUNTIME
 *   HIP/device failure                                     ASPHER_ERR_DEVICE
 * Numerical outcomes are never errors: they are the `status` string of the
 * result (converged, stagnated, max_iterations, nonpositive_curvature,
 * nonfinite, verification_failed, resource_limit).
 *
 * Grids are row-major, flat index iy*Ns + ix, FP64 on the host side. Calls
 * are synchronous; a handle must not be used from two threads at once.
 */
#ifndef ASPHER_HIP_H
#define ASPHER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASPHER_OK 0
#define ASPHER_ERR_INVALID 1
#define ASPHER_ERR_LOGIC 2
#define ASPHER_ERR_RUNTIME 3
#define ASPHER_ERR_DEVICE 4

typedef struct aspher_solver aspher_solver;
typedef struct aspher_result aspher_result;

/* Message of the last failed call on this thread ("" if none). */
const char* aspher_last_error(void);

/* Library version string. */
const char* aspher_version(void);

/* Number of visible HIP devices (0 without a GPU; never an error). */
int aspher_device_count(int* count);

/* ContactSolver.__init__ (python/bindings.cpp:117-162, defaults :804-815).
 * backend: "fft" (free-space, FP64 Love table + padded FFT), "periodic",
 * "dense" (direct Love sum on the device). Other reference backends
 * ("h2", "hmatrix", "") are not part of this library: ASPHER_ERR_INVALID. */
int aspher_solver_create(int grid_size, double domain_size, double E_star,
                         const char* backend, aspher_solver** out);
void aspher_solver_destroy(aspher_solver* s);

/* ContactSolver.matvec (python/bindings.cpp:164-184): u = S p, N = Ns^2
 * host doubles in and out. */
int aspher_solver_matvec(aspher_solver* s, const double* p, double* u);

/* Same on device buffers already resident in HBM: x, y hold Ns^2 values of
 * FP64 (precision 64) or FP32 (precision 32); x and y may alias. */
int aspher_solver_matvec_device(aspher_solver* s, const void* x, void* y, int precision);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * u = S x as aspher_solver_matvec_device, on the same kernels, but the call
 * allocates nothing. Its intermediate spectra live in one static 72 MiB
 * device array per GPU, held until the process ends, which bounds the grid:
 * free space up to Ns = 2048, periodic up to Ns = 4096 in FP32 and 2048 in
 * FP64 (power-of-two grids; larger ones are rejected with
 * ASPHER_ERR_INVALID). The handle must be prepared for the precision: FP64 is
 * prepared by aspher_solver_create, FP32 by one aspher_solver_matvec_device
 * call in FP32; otherwise ASPHER_ERR_LOGIC. Every call on a GPU shares that
 * GPU's array, so calls that overlap on one GPU (other threads, other
 * handles) race; calls in sequence are safe. x and y may alias. */
int aspher_solver_matvec_device_noalloc(aspher_solver* s, const void* x, void* y,
                                        int precision);

/* Average device time (ms) of `reps` back-to-back device matvecs, measured
 * with HIP events on the library's own stream (benchmark support). */
int aspher_solver_time_matvec(aspher_solver* s, const void* x, void* y, int precision,
                              int reps, double* ms_per_matvec);

/* Average device time (ms) of each launch of the spectral matvec pipeline
 * (ms[0] row forward transform, ms[1] fused column pass, ms[2] row inverse
 * transform) over `reps` back-to-back matvecs, timed with HIP events on the
 * library's stream. No reference counterpart (profiling support). */
int aspher_solver_time_stages(aspher_solver* s, const void* x, void* y, int precision,
                              int reps, double* ms3);

/* ContactSolver.solve (python/bindings.cpp:186-234, defaults :818-828):
 * precond "none" | "" | "fourier" (13-tap stencil) | "fourier-fft" (full-grid
 * |k|); p_init may be NULL. */
int aspher_solver_solve(aspher_solver* s, const double* gap, double p_nominal, double tol,
                        int max_iter, int use_pr, const char* precond,
                        const double* p_init, aspher_result** out);

/* Same with the gap (and optional p_init) already on the device (FP64). */
int aspher_solver_solve_device(aspher_solver* s, const double* gap_dev, double p_nominal,
                               double tol, int max_iter, int use_pr, const char* precond,
                               const double* p_init_dev, aspher_result** out);

/* ContactSolver.hmatrix_info for fft / periodic / dense
 * (python/bindings.cpp:307-328). */
typedef struct {
    char backend[16];
    int64_t n;
    int padded_side;         /* 2 Ns for fft, Ns otherwise */
    int64_t bytes;           /* spectrum + scratch held */
    int64_t bytes_spectrum;
    int64_t bytes_scratch;
    double build_seconds;
} aspher_operator_info;
int aspher_solver_info(aspher_solver* s, aspher_operator_info* out);

/* solve_nested keyword arguments (python/bindings.cpp:836-853). */
typedef struct {
    int coarsest;                   /* 64 */
    int q;                          /* 6 (accepted, unused: no H2 here) */
    int leaf_side;                  /* 8 */
    int precond;                    /* 1 */
    char precond_engine[16];        /* "auto" | "stencil" | "fft" */
    int precond_radius;             /* 2 */
    double precond_occupancy_max;   /* 0.4 */
    double precond_cost_gate;       /* 0.0 */
    double coarse_tol;              /* 1e-4 */
    int single_precision;           /* 0 */
    char precision[24];             /* "" | "double" | "float" | "float_then_double" */
    double float_floor;             /* 2e-6 */
    int allow_tolerance_relaxation; /* 0 */
    int light_result;               /* 0 */
    char backend[16];               /* "h2" (reference default; not in this build) */
    int record_error_history;       /* 0 */
    int active_set;                 /* 0 */
    double active_delta;            /* 0.05 */
    int active_halo;                /* 2 */
    int active_max_rounds;          /* 5 */
    int active_all_levels;          /* 1 */
    double active_occupancy_max;    /* 0.4 */
    int reuse_operators;            /* 0: rebuild every level like the reference;
                                       1: keep built operators in a process cache */
    int64_t graph_max_points;       /* 0: levels of at most this many points
                                       replay each CG iteration from captured
                                       hipGraphs (0: launch every kernel; same
                                       results either way). No reference
                                       counterpart. */
} aspher_nested_params;
void aspher_nested_params_default(aspher_nested_params* p);

/* solve_nested (python/bindings.cpp:362-432): gap is Ns^2 host doubles. */
int aspher_solve_nested(int grid_size, const double* gap, double p_nominal,
                        double domain_size, double E_star, double tol, int max_iter,
                        int use_pr, const aspher_nested_params* params,
                        aspher_result** out);
int aspher_solve_nested_device(int grid_size, const double* gap_dev, double p_nominal,
                               double domain_size, double E_star, double tol, int max_iter,
                               int use_pr, const aspher_nested_params* params,
                               aspher_result** out);

/* ContactResult scalar properties (python/bindings.cpp:637-802). */
typedef struct {
    int grid_size;
    double objective, contact_area, mean_pressure, approach;
    int iterations;
    double error;
    int converged;
    int active_rounds, active_fallback;
    char status[32];
    char status_reason[32];
    double pk_error, fw_gap, fw_error, load_error, pressure_violation, penetration_error;
    double requested_tol, effective_tol, p_ref, g_ref;
    char validation_scope[32];
    char operator_error_kind[32];
    int64_t matvec_count, verification_matvec_count, precond_count;
    int identification_steps, returned_best, precond_dropped;
    int64_t mem_cg_state, mem_best_iterate, mem_contact_mask, mem_output_arrays, mem_peak;
    /* seconds; t_matvec and t_precond are GPU time from events around every
       eighth operator / preconditioner apply, scaled to all applies */
    double t_total, t_matvec, t_precond, t_build, t_coarse, t_verification, t_candidate,
        t_output;
    int has_displacement, has_gap;
    int n_history, n_stages;
} aspher_result_info;

int aspher_result_get_info(const aspher_result* r, aspher_result_info* out);
/* which: 0 pressure, 1 displacement, 2 gap; out receives Ns^2 doubles.
 * ASPHER_ERR_INVALID when the field is absent (light result). */
int aspher_result_get_field(const aspher_result* r, int which, double* out);
/* device pointer of a field (valid until aspher_result_destroy) */
int aspher_result_field_device(const aspher_result* r, int which, const double** dev);
int aspher_result_get_history(const aspher_result* r, double* out);

typedef struct {
    char name[48];
    char precision[16];
    int q;
    double requested_tol, effective_tol;
    int iterations;
    int64_t matvec_count;
    double seconds;
    char status[32];
    double fw_error, penetration_error;
    int precond_dropped;
} aspher_stage_info;
int aspher_result_get_stage(const aspher_result* r, int index, aspher_stage_info* out);
void aspher_result_destroy(aspher_result* r);

/* configure_allocator (python/bindings.cpp, runtime_policy): host allocator
 * policy; kept for API parity, returns 0 (not applied) through *applied. */
int aspher_configure_allocator(int64_t mmap_threshold_bytes, int64_t trim_threshold_bytes,
                               int* applied);

/* Kernel-level entry points used by the parity tests. */
/* love_uz at n points on the device (src/boussinesq_kernel.cpp:71-95) */
int aspher_love_points(int n, const double* x, const double* y, const double* a,
                       const double* b, double* out);
/* Ns x Ns FP64 influence table (src/boussinesq_kernel.cpp:97-106) */
int aspher_love_table(int grid_size, double domain_size, double E_star, double* out);
/* stencil taps (src/stencil_precond.cpp:199-220); arrays hold >= (2R+1)^2 */
int aspher_stencil_taps(int grid_size, int radius, int* ntaps, int* dx, int* dy, double* w);
/* z = M^-1 g on the contact mask (src/stencil_precond.cpp:63-109,
 * src/fourier_precond.cpp:50-116): kind "fourier" (stencil) | "fourier-fft";
 * precision 64 or 32 (host arrays are FP64 either way). */
int aspher_precond_apply(int grid_size, const char* kind, int radius, int precision,
                         const double* g, const uint8_t* mask, double* z);
/* Drop the process-wide operator cache. */
int aspher_clear_cache(void);

/* project_load_simplex (src/contact_solver.cpp:31-60), the projection of the
 * identification step, run on the device: out = max(y - theta, 0) with
 * sum(out) = P_total, theta by Michelot's threshold iteration (FP64 sums, at
 * most 64 passes). n values; precision 64 or 32 (host arrays are FP64 either
 * way, the FP32 path rounds y first); theta may be NULL. The reference has no
 * binding for it (its test T21 calls the C++ template,
 * tests/test_certification.cpp:443-467); this entry point serves that test. */
int aspher_project_load_simplex(int64_t n, const double* y, double P_total, int precision,
                                double* out, double* theta);

/* Test hook for the failure statuses `nonfinite` and `verification_failed`.
 * The reference's tests inject faults through a wrapped MatVec functor
 * (tests/test_certification.cpp:270-288); the device paths have no such
 * seam, so this hook stands in for it. Process-wide, off by default:
 *   mode 0: off.
 *   mode 1: every operator apply after the first `after` ones, counted on the
 *           device from this call on (graph replays included), writes NaN
 *           into element 0 of its output; in an ensemble, into element 0 of
 *           the first instance of each size group.
 *   mode 2: the final fresh evaluation of every converged single-level
 *           solve sees element 0 of its displacement lowered by 10 g_ref
 *           (a non-deterministic operator), so it must report
 *           verification_failed. */
int aspher_test_fault(int mode, int64_t after);

/* Same as aspher_solver_create, but the operator (Love table + kernel
 * spectrum) comes from the process-wide cache that solve_nested fills when
 * it reuses operators, so a later adjoint runs on the very spectrum of the
 * forward solve. No reference counterpart (the reference rebuilds per call). */
int aspher_solver_create_shared(int grid_size, double domain_size, double E_star,
                                const char* backend, aspher_solver** out);

/* Adjoint of a converged solve: the backward pass of p(gap, p_nominal).
 * No counterpart in the reference (it has no derivative of its solve); it
 * differentiates the KKT system of the QP that solve_contact solves
 * (src/contact_solver.cpp:150-587): for dL/dp = grad_pressure on the support
 * C = {p > 0},
 *     [S_CC 1; 1' 0][mu; nu] = [grad_pressure_C; 0],
 *     grad_gap = -mu on C (0 off C),  dL/dp_nominal = nu Ns^2,
 * solved by preconditioned CG on the support with the solver's operator. */
typedef struct {
    double tol;             /* relative residual of the adjoint CG; default 1e-10 */
    int max_iter;           /* default 5000 */
    char precond[16];       /* "auto" (default), "none", "fourier" (13-tap
                               stencil), "fourier-fft" */
} aspher_adjoint_params;

typedef struct {
    int converged, iterations;
    int64_t n_contact, matvec_count, precond_count;
    double rel_residual;
    double multiplier;      /* nu = dL/dP, P = p_nominal Ns^2 */
    double grad_p_nominal;  /* dL/dp_nominal */
    double seconds;
    char status[32];        /* converged | max_iterations | nonpositive_curvature | nonfinite */
    char precond[16];       /* preconditioner used */
} aspher_adjoint_info;

void aspher_adjoint_params_default(aspher_adjoint_params* p);
/* host arrays of Ns^2 doubles; `res` must come from a solve on the same grid */
int aspher_solver_adjoint(aspher_solver* s, const aspher_result* res,
                          const double* grad_pressure, const aspher_adjoint_params* prm,
                          double* grad_gap, aspher_adjoint_info* info);
/* same on device FP64 buffers */
int aspher_solver_adjoint_device(aspher_solver* s, const aspher_result* res,
                                 const double* grad_pressure_dev,
                                 const aspher_adjoint_params* prm, double* grad_gap_dev,
                                 aspher_adjoint_info* info);

/* Ensemble: many independent certified solves of uneven grid sizes advanced
 * together (no reference counterpart; each instance is the solve of
 * ContactSolver(backend).solve(gap, p_nominal, precond=...),
 * python/bindings.cpp:186-234, single level). Instances are grouped by
 * (backend, grid_size, domain_size, E_star); every O(N) pass of an iteration
 * is one launch over a per-instance work queue, the spectral operators one
 * batched launch per group, and convergence is decided on the device, so the
 * iteration is captured in a HIP graph and replayed; the host polls the
 * instances' flags every `check_every` replays. Tile shapes of the batched
 * transforms are chosen by an on-device autotuner (cached per process). */
typedef struct {
    double tol;              /* 1e-5; FP32 ensembles certify at max(tol, 2e-6) */
    int max_iter;            /* 5000 */
    int use_pr;              /* 1 */
    char precision[16];      /* "float" (FP32 + FP64 reductions) | "double" */
    char precond[16];        /* "auto" (the full-grid Fourier |q| preconditioner
                                for both kinds, as "fourier-fft"), "none",
                                "fourier" (13-tap stencil), "fourier-fft" */
    int check_every;         /* 8 */
    int use_graph;           /* 1 */
    int autotune;            /* 1 */
    int keep_best;           /* 1 */
    int stall_limit;         /* 200 */
    int light_result;        /* 0 */
} aspher_ensemble_params;

typedef struct {
    double seconds, setup_seconds, loop_seconds, finish_seconds, tune_seconds;
    int groups, replays, polls, host_fallbacks, device_iterations, graph;
} aspher_ensemble_stats;

void aspher_ensemble_params_default(aspher_ensemble_params* p);
/* n instances; backend[i]: 0 "fft", 1 "periodic"; gaps[i]: Ns_i^2 doubles
 * (host for aspher_solve_ensemble, device for the _device variant);
 * results[i] receives a result handle; instance_seconds (may be NULL) the
 * time from the start of the loop until instance i left the device. */
int aspher_solve_ensemble(int n, const int* backend, const int* grid_size,
                          const double* domain_size, const double* E_star,
                          const double* p_nominal, const double* const* gaps,
                          const aspher_ensemble_params* params, aspher_result** results,
                          aspher_ensemble_stats* stats, double* instance_seconds);
int aspher_solve_ensemble_device(int n, const int* backend, const int* grid_size,
                                 const double* domain_size, const double* E_star,
                                 const double* p_nominal, const double* const* gaps_dev,
                                 const aspher_ensemble_params* params,
                                 aspher_result** results, aspher_ensemble_stats* stats,
                                 double* instance_seconds);

/* Tangential (Cerruti) 2x2 block operator u = [Kxx Kxy; Kxy Kyy] q on
 * stacked [x; y] components of Ns^2 values each
 * (TangentialFFTOperator, src/tangential_operator.cpp:16-157, with the
 * CerrutiKernel tables of src/cerruti_kernel.cpp:117-127). The reference
 * reaches it only through its friction driver; grid_size a power of two in
 * [16, 2048]. */
typedef struct aspher_tangential aspher_tangential;
int aspher_tangential_create(int grid_size, double domain_size, double E_star, double nu,
                             aspher_tangential** out);
void aspher_tangential_destroy(aspher_tangential* t);
/* host: q, u hold 2 Ns^2 doubles */
int aspher_tangential_matvec(aspher_tangential* t, const double* q, double* u);
/* device buffers of 2 Ns^2 FP64 (precision 64) or FP32 (32) values; may alias */
int aspher_tangential_matvec_device(aspher_tangential* t, const void* q, void* u, int precision);
/* TangentialFourierPreconditioner::apply_into (src/fourier_precond.cpp:291-358):
 * g, z hold 2 Ns^2 doubles, mask Ns^2 bytes; precision 64 or 32. */
int aspher_tangential_precond_apply(int grid_size, double nu, int precision, const double* g,
                                    const uint8_t* mask, int remove_mean, double* z);
/* Cerruti brackets cerruti_uxx / cerruti_uxy (src/cerruti_kernel.cpp:85-123)
 * at n points, and the Ns x Ns scaled tables (tests of the device kernels). */
int aspher_cerruti_points(int n, const double* x, const double* y, const double* a,
                          const double* b, double nu, double* uxx, double* uxy);
int aspher_cerruti_tables(int grid_size, double domain_size, double E_star, double nu,
                          double* xx, double* xy);

/* Self-affine Gaussian rough surface generated on the device. The
 * reference's benchmarks draw their surfaces from the third-party `rfgen`
 * package (pyproject.toml:15), outside the library; this replaces it with
 * the published recipe: amplitude |m|^-(1+hurst) for integer wavenumber
 * magnitudes k_low <= |m| <= k_high, uniform random phase, real part of the
 * inverse 2-D DFT, mean removed, scaled to `rms`. The phase of frequency
 * (kx, ky) is 2 pi times SplitMix64 output number ky*Ns + kx of a generator
 * seeded with `seed` (53-bit uniform). Ns^2 doubles, row-major. */
int aspher_rough_surface(int grid_size, double hurst, double k_low, double k_high, double rms,
                         uint64_t seed, double* out);
int aspher_rough_surface_device(int grid_size, double hurst, double k_low, double k_high,
                                double rms, uint64_t seed, double* out_dev);

/* Device buffers for callers of the *_device entry points that have no
 * allocator of their own (tests, benchmarks). No reference counterpart: the
 * reference is host-only. */
int aspher_device_malloc(int64_t bytes, void** ptr);
int aspher_device_free(void* ptr);
int aspher_memcpy_htod(void* dst, const void* src, int64_t bytes);
int aspher_memcpy_dtoh(void* dst, const void* src, int64_t bytes);
/* device-to-device copy by a plain streaming kernel of this library (reads
 * and writes exactly `bytes`, a multiple of 16): the calibration of the HBM
 * traffic counters. No reference counterpart. */
int aspher_copy_device(void* dst, const void* src, int64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* ASPHER_HIP_H */
