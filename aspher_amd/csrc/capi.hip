// extern "C" boundary (include/aspher_hip.h): exceptions become status codes
// plus a thread-local message; device state stays behind opaque handles.
#include "../../include/aspher_hip.h"
#include "engine.hpp"
#include "ensemble.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

using namespace aspher;

struct aspher_solver {
    std::unique_ptr<Context> ctx;
    std::shared_ptr<Operator> op;
    std::string backend;
    int Ns = 0;
};

struct aspher_result {
    Result r;
};

struct aspher_tangential {
    std::unique_ptr<Context> ctx;
    std::unique_ptr<TangentialOperator> op;
    int Ns = 0;
};

namespace {

thread_local std::string g_err;

template <typename F>
int guard(F&& f) {
    try {
        g_err.clear();
        f();
        return ASPHER_OK;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return ASPHER_ERR_INVALID;
    } catch (const std::logic_error& e) {
        g_err = e.what();
        return ASPHER_ERR_LOGIC;
    } catch (const DeviceError& e) {
        g_err = e.what();
        return ASPHER_ERR_DEVICE;
    } catch (const std::bad_alloc& e) {
        g_err = "host allocation failed";
        return ASPHER_ERR_RUNTIME;
    } catch (const std::exception& e) {
        g_err = e.what();
        return ASPHER_ERR_RUNTIME;
    }
}

void copy_str(char* dst, std::size_t n, const std::string& s) {
    std::strncpy(dst, s.c_str(), n - 1);
    dst[n - 1] = '\0';
}

Precond::Kind precond_kind(const std::string& name) {
    // python/bindings.cpp:193-214
    if (name == "fourier") return Precond::Kind::stencil;
    if (name == "fourier-fft") return Precond::Kind::fourier;
    if (name == "none" || name.empty()) return Precond::Kind::none;
    throw std::invalid_argument("precond must be 'none', 'fourier' or 'fourier-fft'");
}

NestedParams to_params(const aspher_nested_params* p) {
    NestedParams np;
    if (!p) return np;
    np.coarsest = p->coarsest;
    np.q = p->q;
    np.leaf_side = p->leaf_side;
    np.precond = p->precond != 0;
    np.precond_engine = p->precond_engine;
    np.precond_radius = p->precond_radius;
    np.precond_occupancy_max = p->precond_occupancy_max;
    np.precond_cost_gate = p->precond_cost_gate;
    np.coarse_tol = p->coarse_tol;
    np.single_precision = p->single_precision != 0;
    np.precision = p->precision;
    np.float_floor = p->float_floor;
    np.allow_tolerance_relaxation = p->allow_tolerance_relaxation != 0;
    np.light_result = p->light_result != 0;
    np.backend = p->backend;
    np.record_error_history = p->record_error_history != 0;
    np.active_set = p->active_set != 0;
    np.active_delta = p->active_delta;
    np.active_halo = p->active_halo;
    np.active_max_rounds = p->active_max_rounds;
    np.active_all_levels = p->active_all_levels != 0;
    np.active_occupancy_max = p->active_occupancy_max;
    np.graph_max_points = p->graph_max_points;
    return np;
}

void solve_single(aspher_solver* s, const double* gap_dev, double p_nominal, double tol,
                  int max_iter, int use_pr, const char* precond, const double* p_init_dev,
                  aspher_result** out) {
    if (!s || !out) throw std::invalid_argument("null handle");
    const Precond::Kind pk = precond_kind(precond ? precond : "none");
    Precond pc(pk, s->Ns, 2);
    // solve_contact defaults (include/contact_solver.hpp:151-191)
    SolveOptions opt;
    opt.tol = tol;
    opt.max_iter = max_iter;
    opt.use_pr = use_pr != 0;
    auto res = std::make_unique<aspher_result>();
    const std::int64_t N = (std::int64_t)s->Ns * s->Ns;
    DevBuf<double> p0;
    if (p_init_dev) {
        p0.alloc(N);
        VecOps<double>::copy(N, p_init_dev, p0.p, s->ctx->stream);
    }
    solve_impl<double>(*s->ctx, *s->op, gap_dev, p_nominal, opt, &pc,
                       p_init_dev ? p0.p : nullptr, res->r);
    *out = res.release();
}

void create_solver(int grid_size, double domain_size, double E_star, const char* backend,
                   bool shared, aspher_solver** out) {
    if (!out) throw std::invalid_argument("null output handle");
    // validation order of python/bindings.cpp:122-127
    if (grid_size <= 0) throw std::invalid_argument("grid_size must be positive");
    if (!(domain_size > 0.0)) throw std::invalid_argument("domain_size must be positive");
    if (!(E_star > 0.0)) throw std::invalid_argument("E_star must be positive");
    const std::string b = backend ? backend : "";
    Operator::Kind kind;
    if (b == "fft") kind = Operator::Kind::free_space;
    else if (b == "periodic") kind = Operator::Kind::periodic;
    else if (b == "dense") kind = Operator::Kind::dense;
    else if (b == "h2" || b == "hmatrix" || b.empty())
        throw std::invalid_argument("backend '" + (b.empty() ? std::string("hmatrix") : b) +
                                    "' is not part of this build (expected fft, periodic "
                                    "or dense)");
    else
        throw std::invalid_argument("unknown backend: " + b +
                                    " (expected dense, hmatrix, h2, fft, or periodic)");
    auto s = std::make_unique<aspher_solver>();
    s->ctx = std::make_unique<Context>();
    if (shared) {
        s->op = cached_operator(kind, grid_size, domain_size, E_star, s->ctx->stream);
    } else {
        s->op = std::make_shared<Operator>(kind, grid_size, domain_size, E_star);
        s->op->build(s->ctx->stream);
    }
    // FP64 plan and kernel spectrum in the plan's order, made here once so
    // that later FP64 matvecs (the no-allocation one included) find them
    s->op->prepare<double>(s->ctx->stream);
    ASPHER_HIP(hipStreamSynchronize(s->ctx->stream));
    s->backend = b;
    s->Ns = grid_size;
    *out = s.release();
}


AdjointOptions to_adjoint(const aspher_adjoint_params* p) {
    AdjointOptions o;
    if (p) {
        o.tol = p->tol;
        o.max_iter = p->max_iter;
    }
    return o;
}

// "auto": the full-grid Fourier preconditioner on the periodic backend (the
// nested driver's periodic default), otherwise the 13-tap stencil unless the
// support covers more than 40% of the grid (src/nested_solve.cpp:27-40)
Precond::Kind adjoint_precond(const aspher_solver* s, const char* name, long long n_contact) {
    const std::string nm = name ? name : "auto";
    if (nm != "auto") return precond_kind(nm);
    if (s->backend == "periodic") return Precond::Kind::fourier;
    const double occ = (double)n_contact / ((double)s->Ns * s->Ns);
    return occ < 0.4 ? Precond::Kind::stencil : Precond::Kind::fourier;
}

void adjoint_device(aspher_solver* s, const aspher_result* res, const double* w_dev,
                    const aspher_adjoint_params* prm, double* grad_dev,
                    aspher_adjoint_info* info) {
    if (!s || !res || !w_dev || !grad_dev) throw std::invalid_argument("null pointer");
    if (res->r.Ns != s->Ns)
        throw std::invalid_argument("result grid size does not match the solver");
    if (!res->r.pressure.p) throw std::invalid_argument("result holds no pressure field");
    const long long nc = (long long)std::llround(res->r.contact_fraction * s->Ns * (double)s->Ns);
    Precond pc(adjoint_precond(s, prm ? prm->precond : nullptr, nc), s->Ns, 2);
    AdjointInfo ai;
    adjoint_impl(*s->ctx, *s->op, &pc, res->r.pressure.p, w_dev, grad_dev, to_adjoint(prm), ai);
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->converged = ai.converged ? 1 : 0;
        info->iterations = ai.iterations;
        info->n_contact = ai.n_contact;
        info->matvec_count = ai.matvec_count;
        info->precond_count = ai.precond_count;
        info->rel_residual = ai.rel_residual;
        info->multiplier = ai.multiplier;
        info->grad_p_nominal = ai.grad_p_nominal;
        info->seconds = ai.seconds;
        copy_str(info->status, sizeof(info->status), ai.status);
        copy_str(info->precond, sizeof(info->precond),
                 pc.kind() == Precond::Kind::stencil ? "fourier"
                 : pc.kind() == Precond::Kind::fourier ? "fourier-fft" : "none");
    }
}


void ensemble_run(int n, const int* backend, const int* grid_size, const double* L,
                  const double* E, const double* pn, const double* const* gaps, bool on_device,
                  const aspher_ensemble_params* prm, aspher_result** results,
                  aspher_ensemble_stats* stats, double* inst_seconds) {
    if (n <= 0 || !backend || !grid_size || !pn || !gaps || !results)
        throw std::invalid_argument("solve_ensemble: null argument or n <= 0");
    EnsembleOptions o;
    if (prm) {
        o.tol = prm->tol;
        o.max_iter = prm->max_iter;
        o.use_pr = prm->use_pr != 0;
        o.precision = prm->precision;
        o.precond = prm->precond;
        o.check_every = prm->check_every;
        o.use_graph = prm->use_graph != 0;
        o.autotune = prm->autotune != 0;
        o.keep_best = prm->keep_best != 0;
        o.stall_limit = prm->stall_limit;
        o.light = prm->light_result != 0;
    }
    std::vector<EnsembleInstance> inst(n);
    std::vector<DevBuf<double>> staged;
    for (int i = 0; i < n; ++i) {
        auto& e = inst[i];
        e.kind = backend[i];
        e.Ns = grid_size[i];
        e.L = L ? L[i] : 1.0;
        e.E_star = E ? E[i] : 1.0;
        e.p_bar = pn[i];
        if (!gaps[i]) throw std::invalid_argument("solve_ensemble: null gap");
        if (e.Ns <= 0) throw std::invalid_argument("solve_ensemble: grid_size must be positive");
    }
    if (on_device) {
        ASPHER_HIP(hipDeviceSynchronize());
        for (int i = 0; i < n; ++i) inst[i].gap = gaps[i];
    } else {
        staged.resize(n);
        for (int i = 0; i < n; ++i) {
            const std::size_t N = (std::size_t)inst[i].Ns * inst[i].Ns;
            staged[i].alloc(N);
            ASPHER_HIP(hipMemcpy(staged[i].p, gaps[i], sizeof(double) * N,
                                 hipMemcpyHostToDevice));
            inst[i].gap = staged[i].p;
        }
    }
    std::vector<Result> res;
    EnsembleStats st;
    solve_ensemble(inst, o, res, st);
    for (int i = 0; i < n; ++i) {
        auto r = std::make_unique<aspher_result>();
        r->r = std::move(res[i]);
        results[i] = r.release();
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->seconds = st.seconds;
        stats->setup_seconds = st.setup_seconds;
        stats->loop_seconds = st.loop_seconds;
        stats->finish_seconds = st.finish_seconds;
        stats->tune_seconds = st.tune_seconds;
        stats->groups = st.groups;
        stats->replays = st.replays;
        stats->polls = st.polls;
        stats->host_fallbacks = st.host_fallbacks;
        stats->device_iterations = st.device_iterations;
        stats->graph = st.graph ? 1 : 0;
    }
    if (inst_seconds)
        for (int i = 0; i < n; ++i) inst_seconds[i] = st.instance_seconds[i];
}

} // namespace

extern "C" {

const char* aspher_last_error(void) { return g_err.c_str(); }

const char* aspher_version(void) { return "aspher-mi355x 0.1.0 (gfx950)"; }

int aspher_device_count(int* count) {
    return guard([&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        if (count) *count = n;
    });
}

int aspher_solver_create(int grid_size, double domain_size, double E_star, const char* backend,
                         aspher_solver** out) {
    return guard([&] { create_solver(grid_size, domain_size, E_star, backend, false, out); });
}

int aspher_solver_create_shared(int grid_size, double domain_size, double E_star,
                                const char* backend, aspher_solver** out) {
    return guard([&] { create_solver(grid_size, domain_size, E_star, backend, true, out); });
}

void aspher_solver_destroy(aspher_solver* s) {
    if (!s) return;
    try {
        s->op.reset();
        delete s;
    } catch (...) {
    }
}

int aspher_solver_matvec(aspher_solver* s, const double* p, double* u) {
    return guard([&] {
        if (!s || !p || !u) throw std::invalid_argument("null pointer");
        const std::int64_t N = (std::int64_t)s->Ns * s->Ns;
        DevBuf<double> x(N);
        hipStream_t st = s->ctx->stream;
        ASPHER_HIP(hipMemcpyAsync(x.p, p, sizeof(double) * N, hipMemcpyHostToDevice, st));
        s->op->apply(x.p, x.p, st);
        ASPHER_HIP(hipMemcpyAsync(u, x.p, sizeof(double) * N, hipMemcpyDeviceToHost, st));
        ASPHER_HIP(hipStreamSynchronize(st));
    });
}

int aspher_solver_matvec_device(aspher_solver* s, const void* x, void* y, int precision) {
    return guard([&] {
        if (!s || !x || !y) throw std::invalid_argument("null pointer");
        ASPHER_HIP(hipDeviceSynchronize());
        hipStream_t st = s->ctx->stream;
        if (precision == 64) {
            s->op->apply(static_cast<const double*>(x), static_cast<double*>(y), st);
        } else if (precision == 32) {
            s->op->build_single(st);
            s->op->apply(static_cast<const float*>(x), static_cast<float*>(y), st);
        } else {
            throw std::invalid_argument("precision must be 64 or 32");
        }
        ASPHER_HIP(hipStreamSynchronize(st));
    });
}

// ── matvec without allocation ─────────────────────────────────────────────
// The intermediate spectra of aspher_solver_matvec_device_noalloc live in this
// file-scope device array: each GPU the library runs on holds its own copy
// (72 MiB) until the process ends. Host code never uses the name `scratch` as
// an address (on the host it is only a stand-in for the symbol): the device
// address comes from hipGetSymbolAddress, once per device, with that device
// current.
constexpr std::size_t kNoAllocScratchBytes = std::size_t(72) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxScratchDevices] = {};

// device address of `scratch` on the current device
void* scratch_on_current_device() {
    int dev = 0;
    ASPHER_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices)
        throw std::runtime_error("matvec without allocation: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw DeviceError(std::string("hipGetSymbolAddress(scratch): ") +
                              hipGetErrorString(e));
        g_scratch_dev[dev] = ptr;
    }
    return g_scratch_dev[dev];
}
} // namespace

int aspher_solver_matvec_device_noalloc(aspher_solver* s, const void* x, void* y,
                                        int precision) {
    return guard([&] {
        if (!s || !x || !y) throw std::invalid_argument("null pointer");
        if (precision != 64 && precision != 32)
            throw std::invalid_argument("precision must be 64 or 32");
        const std::size_t need = precision == 64 ? s->op->workspace_external<double>()
                                                 : s->op->workspace_external<float>();
        if (need > kNoAllocScratchBytes)
            throw std::invalid_argument(
                "matvec without allocation: this grid needs " + std::to_string(need) +
                " bytes of scratch, more than the static " +
                std::to_string(kNoAllocScratchBytes) + " (72 MiB)");
        void* ws = scratch_on_current_device();
        ASPHER_HIP(hipDeviceSynchronize());
        hipStream_t st = s->ctx->stream;
        if (precision == 64)
            s->op->apply_external(static_cast<const double*>(x), static_cast<double*>(y), ws,
                                  kNoAllocScratchBytes, st);
        else
            s->op->apply_external(static_cast<const float*>(x), static_cast<float*>(y), ws,
                                  kNoAllocScratchBytes, st);
        ASPHER_HIP(hipStreamSynchronize(st));
    });
}

int aspher_solver_time_matvec(aspher_solver* s, const void* x, void* y, int precision, int reps,
                              double* ms) {
    return guard([&] {
        if (!s || !x || !y || !ms || reps <= 0) throw std::invalid_argument("bad arguments");
        ASPHER_HIP(hipDeviceSynchronize());
        hipStream_t st = s->ctx->stream;
        if (precision == 32) s->op->build_single(s->ctx->stream);
        auto run = [&] {
            if (precision == 64)
                s->op->apply(static_cast<const double*>(x), static_cast<double*>(y), st);
            else
                s->op->apply(static_cast<const float*>(x), static_cast<float*>(y), st);
        };
        run();  // warm-up
        hipEvent_t a, b;
        ASPHER_HIP(hipEventCreate(&a));
        ASPHER_HIP(hipEventCreate(&b));
        ASPHER_HIP(hipEventRecord(a, st));
        for (int i = 0; i < reps; ++i) run();
        ASPHER_HIP(hipEventRecord(b, st));
        ASPHER_HIP(hipEventSynchronize(b));
        float t = 0.0f;
        ASPHER_HIP(hipEventElapsedTime(&t, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        *ms = t / reps;
    });
}

int aspher_solver_time_stages(aspher_solver* s, const void* x, void* y, int precision, int reps,
                              double* ms3) {
    return guard([&] {
        if (!s || !x || !y || !ms3 || reps <= 0) throw std::invalid_argument("bad arguments");
        if (precision != 64 && precision != 32)
            throw std::invalid_argument("precision must be 64 or 32");
        ASPHER_HIP(hipDeviceSynchronize());
        hipStream_t st = s->ctx->stream;
        if (precision == 32) s->op->build_single(s->ctx->stream);
        std::vector<hipEvent_t> ev(4 * (std::size_t)reps);
        for (auto& e : ev) ASPHER_HIP(hipEventCreate(&e));
        auto run = [&](hipEvent_t* e) {
            if (precision == 64)
                s->op->apply_timed(static_cast<const double*>(x), static_cast<double*>(y), st, e);
            else
                s->op->apply_timed(static_cast<const float*>(x), static_cast<float*>(y), st, e);
        };
        run(ev.data());  // warm-up
        ASPHER_HIP(hipStreamSynchronize(st));
        for (int r = 0; r < reps; ++r) run(ev.data() + 4 * r);
        ASPHER_HIP(hipStreamSynchronize(st));
        double acc[3] = {0, 0, 0};
        for (int r = 0; r < reps; ++r)
            for (int k = 0; k < 3; ++k) {
                float t = 0.0f;
                ASPHER_HIP(hipEventElapsedTime(&t, ev[4 * r + k], ev[4 * r + k + 1]));
                acc[k] += t;
            }
        for (auto e : ev) (void)hipEventDestroy(e);
        for (int k = 0; k < 3; ++k) ms3[k] = acc[k] / reps;
    });
}

int aspher_solver_solve(aspher_solver* s, const double* gap, double p_nominal, double tol,
                        int max_iter, int use_pr, const char* precond, const double* p_init,
                        aspher_result** out) {
    return guard([&] {
        if (!s || !gap) throw std::invalid_argument("null pointer");
        const std::int64_t N = (std::int64_t)s->Ns * s->Ns;
        hipStream_t st = s->ctx->stream;
        (void)precond_kind(precond ? precond : "none");
        DevBuf<double> g(N), p0;
        ASPHER_HIP(hipMemcpyAsync(g.p, gap, sizeof(double) * N, hipMemcpyHostToDevice, st));
        if (p_init) {
            p0.alloc(N);
            ASPHER_HIP(hipMemcpyAsync(p0.p, p_init, sizeof(double) * N, hipMemcpyHostToDevice, st));
        }
        solve_single(s, g.p, p_nominal, tol, max_iter, use_pr, precond, p_init ? p0.p : nullptr,
                     out);
    });
}

int aspher_solver_solve_device(aspher_solver* s, const double* gap_dev, double p_nominal,
                               double tol, int max_iter, int use_pr, const char* precond,
                               const double* p_init_dev, aspher_result** out) {
    return guard([&] {
        if (!s || !gap_dev) throw std::invalid_argument("null pointer");
        ASPHER_HIP(hipDeviceSynchronize());
        solve_single(s, gap_dev, p_nominal, tol, max_iter, use_pr, precond, p_init_dev, out);
    });
}

int aspher_solver_info(aspher_solver* s, aspher_operator_info* out) {
    return guard([&] {
        if (!s || !out) throw std::invalid_argument("null pointer");
        std::memset(out, 0, sizeof(*out));
        copy_str(out->backend, sizeof(out->backend), s->backend);
        out->n = (int64_t)s->Ns * s->Ns;
        out->padded_side = s->backend == "fft" ? 2 * s->Ns : s->Ns;
        out->bytes_spectrum = (int64_t)s->op->bytes_spectrum();
        out->bytes_scratch = (int64_t)s->op->bytes_scratch();
        out->bytes = out->bytes_spectrum + out->bytes_scratch;
        out->build_seconds = s->op->build_seconds;
    });
}

void aspher_nested_params_default(aspher_nested_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->coarsest = 64;
    p->q = 6;
    p->leaf_side = 8;
    p->precond = 1;
    std::strcpy(p->precond_engine, "auto");
    p->precond_radius = 2;
    p->precond_occupancy_max = 0.4;
    p->precond_cost_gate = 0.0;
    p->coarse_tol = 1e-4;
    p->single_precision = 0;
    p->precision[0] = '\0';
    p->float_floor = 2e-6;
    p->allow_tolerance_relaxation = 0;
    p->light_result = 0;
    std::strcpy(p->backend, "h2");
    p->record_error_history = 0;
    p->active_set = 0;
    p->active_delta = 0.05;
    p->active_halo = 2;
    p->active_max_rounds = 5;
    p->active_all_levels = 1;
    p->active_occupancy_max = 0.4;
    p->reuse_operators = 0;
    p->graph_max_points = 0;
}

int aspher_solve_nested(int grid_size, const double* gap, double p_nominal, double domain_size,
                        double E_star, double tol, int max_iter, int use_pr,
                        const aspher_nested_params* params, aspher_result** out) {
    return guard([&] {
        if (!gap || !out) throw std::invalid_argument("null pointer");
        if (grid_size <= 0)
            throw std::invalid_argument("solve_contact_nested: Ns must be positive");
        const std::int64_t N = (std::int64_t)grid_size * grid_size;
        DevBuf<double> g(N);
        ASPHER_HIP(hipMemcpy(g.p, gap, sizeof(double) * N, hipMemcpyHostToDevice));
        auto res = std::make_unique<aspher_result>();
        const NestedParams np = to_params(params);
        solve_nested(grid_size, domain_size, E_star, g.p, p_nominal, tol, max_iter, use_pr != 0,
                     np, res->r, params && params->reuse_operators != 0);
        *out = res.release();
    });
}

int aspher_solve_nested_device(int grid_size, const double* gap_dev, double p_nominal,
                               double domain_size, double E_star, double tol, int max_iter,
                               int use_pr, const aspher_nested_params* params,
                               aspher_result** out) {
    return guard([&] {
        if (!gap_dev || !out) throw std::invalid_argument("null pointer");
        ASPHER_HIP(hipDeviceSynchronize());
        auto res = std::make_unique<aspher_result>();
        const NestedParams np = to_params(params);
        solve_nested(grid_size, domain_size, E_star, gap_dev, p_nominal, tol, max_iter,
                     use_pr != 0, np, res->r, params && params->reuse_operators != 0);
        *out = res.release();
    });
}

int aspher_result_get_info(const aspher_result* rr, aspher_result_info* o) {
    return guard([&] {
        if (!rr || !o) throw std::invalid_argument("null pointer");
        const Result& r = rr->r;
        std::memset(o, 0, sizeof(*o));
        o->grid_size = r.Ns;
        o->objective = r.objective;
        o->contact_area = r.contact_fraction;
        o->mean_pressure = r.mean_pressure;
        o->approach = r.approach;
        o->iterations = r.iterations;
        o->error = r.error;
        o->converged = r.converged ? 1 : 0;
        o->active_rounds = r.active_rounds;
        o->active_fallback = r.active_fallback ? 1 : 0;
        copy_str(o->status, sizeof(o->status), to_string(r.status));
        copy_str(o->status_reason, sizeof(o->status_reason), r.status_reason);
        o->pk_error = r.pk_error;
        o->fw_gap = r.fw_gap;
        o->fw_error = r.fw_error;
        o->load_error = r.load_error;
        o->pressure_violation = r.pressure_violation;
        o->penetration_error = r.penetration_error;
        o->requested_tol = r.requested_tol;
        o->effective_tol = r.effective_tol;
        o->p_ref = r.p_ref;
        o->g_ref = r.g_ref;
        copy_str(o->validation_scope, sizeof(o->validation_scope), r.validation_scope);
        copy_str(o->operator_error_kind, sizeof(o->operator_error_kind), r.operator_error_kind);
        o->matvec_count = r.matvec_count;
        o->verification_matvec_count = r.verification_matvec_count;
        o->precond_count = r.precond_count;
        o->identification_steps = r.identification_steps;
        o->returned_best = r.returned_best ? 1 : 0;
        o->precond_dropped = r.precond_dropped ? 1 : 0;
        o->mem_cg_state = r.mem_cg_state;
        o->mem_best_iterate = r.mem_best;
        o->mem_contact_mask = r.mem_mask;
        o->mem_output_arrays = r.mem_output;
        o->mem_peak = r.mem_peak;
        o->t_total = r.time_total;
        o->t_matvec = r.time_matvec;
        o->t_precond = r.time_precond;
        o->t_build = r.time_build;
        o->t_coarse = r.time_coarse;
        o->t_verification = r.time_verification;
        o->t_candidate = r.time_candidate;
        o->t_output = r.time_output;
        o->has_displacement = r.displacement.p ? 1 : 0;
        o->has_gap = r.gap.p ? 1 : 0;
        o->n_history = (int)r.error_history.size();
        o->n_stages = (int)r.stage_stats.size();
    });
}

int aspher_result_field_device(const aspher_result* rr, int which, const double** dev) {
    return guard([&] {
        if (!rr || !dev) throw std::invalid_argument("null pointer");
        const Result& r = rr->r;
        const DevBuf<double>* b = which == 0 ? &r.pressure : which == 1 ? &r.displacement
                                  : which == 2 ? &r.gap : nullptr;
        if (!b) throw std::invalid_argument("field index must be 0, 1 or 2");
        if (!b->p) throw std::invalid_argument("field not stored (light result)");
        *dev = b->p;
    });
}

int aspher_result_get_field(const aspher_result* rr, int which, double* out) {
    return guard([&] {
        if (!rr || !out) throw std::invalid_argument("null pointer");
        const Result& r = rr->r;
        const DevBuf<double>* b = which == 0 ? &r.pressure : which == 1 ? &r.displacement
                                  : which == 2 ? &r.gap : nullptr;
        if (!b) throw std::invalid_argument("field index must be 0, 1 or 2");
        if (!b->p) throw std::invalid_argument("field not stored (light result)");
        ASPHER_HIP(hipMemcpy(out, b->p, b->bytes(), hipMemcpyDeviceToHost));
    });
}

int aspher_result_get_history(const aspher_result* rr, double* out) {
    return guard([&] {
        if (!rr || !out) throw std::invalid_argument("null pointer");
        const auto& h = rr->r.error_history;
        if (!h.empty()) std::memcpy(out, h.data(), sizeof(double) * h.size());
    });
}

int aspher_result_get_stage(const aspher_result* rr, int index, aspher_stage_info* o) {
    return guard([&] {
        if (!rr || !o) throw std::invalid_argument("null pointer");
        const auto& st = rr->r.stage_stats;
        if (index < 0 || index >= (int)st.size()) throw std::invalid_argument("stage index");
        const Stage& s = st[index];
        std::memset(o, 0, sizeof(*o));
        copy_str(o->name, sizeof(o->name), s.name);
        copy_str(o->precision, sizeof(o->precision), s.precision);
        o->q = s.q;
        o->requested_tol = s.requested_tol;
        o->effective_tol = s.effective_tol;
        o->iterations = s.iterations;
        o->matvec_count = s.matvec_count;
        o->seconds = s.seconds;
        copy_str(o->status, sizeof(o->status), to_string(s.status));
        o->fw_error = s.fw_error;
        o->penetration_error = s.penetration_error;
        o->precond_dropped = s.precond_dropped ? 1 : 0;
    });
}

void aspher_result_destroy(aspher_result* r) {
    try {
        delete r;
    } catch (...) {
    }
}

int aspher_configure_allocator(int64_t, int64_t, int* applied) {
    if (applied) *applied = 0;
    return ASPHER_OK;
}

int aspher_love_points(int n, const double* x, const double* y, const double* a,
                       const double* b, double* out) {
    return guard([&] {
        if (n <= 0) return;
        DevBuf<double> d(5 * (std::size_t)n);
        ASPHER_HIP(hipMemcpy(d.p, x, sizeof(double) * n, hipMemcpyHostToDevice));
        ASPHER_HIP(hipMemcpy(d.p + n, y, sizeof(double) * n, hipMemcpyHostToDevice));
        ASPHER_HIP(hipMemcpy(d.p + 2 * n, a, sizeof(double) * n, hipMemcpyHostToDevice));
        ASPHER_HIP(hipMemcpy(d.p + 3 * n, b, sizeof(double) * n, hipMemcpyHostToDevice));
        love_points(n, d.p, d.p + n, d.p + 2 * n, d.p + 3 * n, d.p + 4 * n, nullptr);
        ASPHER_HIP(hipMemcpy(out, d.p + 4 * n, sizeof(double) * n, hipMemcpyDeviceToHost));
    });
}

int aspher_love_table(int grid_size, double domain_size, double E_star, double* out) {
    return guard([&] {
        if (grid_size <= 0) throw std::invalid_argument("grid_size must be positive");
        const std::int64_t N = (std::int64_t)grid_size * grid_size;
        DevBuf<double> t(N);
        love_table(grid_size, domain_size, E_star, t.p, nullptr);
        ASPHER_HIP(hipMemcpy(out, t.p, sizeof(double) * N, hipMemcpyDeviceToHost));
    });
}

int aspher_stencil_taps(int grid_size, int radius, int* ntaps, int* dx, int* dy, double* w) {
    return guard([&] {
        if (radius < 1) throw std::invalid_argument("stencil radius must be >= 1");
        if (2 * radius + 1 > grid_size)
            throw std::invalid_argument("stencil radius too large for the grid");
        std::vector<int> vx, vy;
        std::vector<double> vw;
        stencil_weights(grid_size, radius, vx, vy, vw);
        *ntaps = (int)vw.size();
        for (std::size_t t = 0; t < vw.size(); ++t) {
            dx[t] = vx[t];
            dy[t] = vy[t];
            w[t] = vw[t];
        }
    });
}

} // extern "C"

namespace {
template <typename T>
void precond_apply_t(int Ns, Precond::Kind kind, int radius, const double* g,
                     const uint8_t* mask, double* z) {
    Context ctx;
    hipStream_t st = ctx.stream;
    const std::int64_t N = (std::int64_t)Ns * Ns;
    DevBuf<double> gd(N);
    DevBuf<T> gt(N), zt(N);
    DevBuf<uint8_t> m(N);
    ASPHER_HIP(hipMemcpyAsync(gd.p, g, sizeof(double) * N, hipMemcpyHostToDevice, st));
    ASPHER_HIP(hipMemcpyAsync(m.p, mask, N, hipMemcpyHostToDevice, st));
    VecOps<T>::from_double(N, gd.p, 0.0, gt.p, st);
    const int nblk = vec_blocks(N);
    static const int ops[2] = {RSUM, RSUM};
    static const int dst[2] = {S_R0, S_NC};
    VecOps<T>::masked_sum(N, gt.p, m.p, ctx.part.p, nblk, st);
    finalize<T>(ctx.part.p, nblk, 2, ops, dst, ctx.scal.p, STEP_NONE, st);
    ctx.read_scal();
    const double nc = ctx.host_scal[S_NC];
    const int idx[1] = {S_GMEAN};
    const double val[1] = {nc > 0 ? ctx.host_scal[S_R0] / nc : 0.0};
    set_scal(ctx.scal.p, 1, idx, val, st);
    Precond pc(kind, Ns, radius);
    pc.apply<T>(gt.p, m.p, zt.p, ctx.scal.p, ctx.part.p, st);
    VecOps<T>::masked_sub(N, zt.p, m.p, ctx.scal.p, S_ZMEAN, st);
    VecOps<T>::to_double(N, zt.p, gd.p, st);
    ASPHER_HIP(hipMemcpyAsync(z, gd.p, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    ASPHER_HIP(hipStreamSynchronize(st));
}
} // namespace

extern "C" {

int aspher_precond_apply(int grid_size, const char* kind, int radius, int precision,
                         const double* g, const uint8_t* mask, double* z) {
    return guard([&] {
        if (!g || !mask || !z) throw std::invalid_argument("null pointer");
        const Precond::Kind k = precond_kind(kind ? kind : "");
        if (k == Precond::Kind::none) throw std::invalid_argument("no preconditioner named");
        if (precision == 64) precond_apply_t<double>(grid_size, k, radius, g, mask, z);
        else if (precision == 32) precond_apply_t<float>(grid_size, k, radius, g, mask, z);
        else throw std::invalid_argument("precision must be 64 or 32");
    });
}

int aspher_clear_cache(void) {
    return guard([&] { clear_operator_cache(); });
}

int aspher_project_load_simplex(int64_t n, const double* y, double P_total, int precision,
                                double* out, double* theta) {
    return guard([&] {
        if (n <= 0) throw std::invalid_argument("project_load_simplex: n must be positive");
        if (!y || !out) throw std::invalid_argument("null pointer");
        if (!std::isfinite(P_total))
            throw std::invalid_argument("project_load_simplex: P_total must be finite");
        Context ctx;
        hipStream_t st = ctx.stream;
        DevBuf<double> yd((std::size_t)n);
        ASPHER_HIP(hipMemcpyAsync(yd.p, y, sizeof(double) * n, hipMemcpyHostToDevice, st));
        double th = 0.0;
        if (precision == 64) {
            th = project_load_simplex_dev<double>(ctx, n, yd.p, P_total);
        } else if (precision == 32) {
            DevBuf<float> yf((std::size_t)n);
            VecOps<float>::from_double(n, yd.p, 0.0, yf.p, st);
            th = project_load_simplex_dev<float>(ctx, n, yf.p, P_total);
            VecOps<float>::to_double(n, yf.p, yd.p, st);
        } else {
            throw std::invalid_argument("precision must be 64 or 32");
        }
        ASPHER_HIP(hipMemcpyAsync(out, yd.p, sizeof(double) * n, hipMemcpyDeviceToHost, st));
        ASPHER_HIP(hipStreamSynchronize(st));
        if (theta) *theta = th;
    });
}

int aspher_test_fault(int mode, int64_t after) {
    return guard([&] { fault_configure(mode, after); });
}


void aspher_adjoint_params_default(aspher_adjoint_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->tol = 1e-10;
    p->max_iter = 5000;
    std::strcpy(p->precond, "auto");
}

int aspher_solver_adjoint(aspher_solver* s, const aspher_result* res,
                          const double* grad_pressure, const aspher_adjoint_params* prm,
                          double* grad_gap, aspher_adjoint_info* info) {
    return guard([&] {
        if (!s || !grad_pressure || !grad_gap) throw std::invalid_argument("null pointer");
        const std::int64_t N = (std::int64_t)s->Ns * s->Ns;
        DevBuf<double> w(N), g(N);
        hipStream_t st = s->ctx->stream;
        ASPHER_HIP(hipMemcpyAsync(w.p, grad_pressure, sizeof(double) * N,
                                  hipMemcpyHostToDevice, st));
        adjoint_device(s, res, w.p, prm, g.p, info);
        ASPHER_HIP(hipMemcpyAsync(grad_gap, g.p, sizeof(double) * N, hipMemcpyDeviceToHost, st));
        ASPHER_HIP(hipStreamSynchronize(st));
    });
}

int aspher_solver_adjoint_device(aspher_solver* s, const aspher_result* res,
                                 const double* grad_pressure_dev,
                                 const aspher_adjoint_params* prm, double* grad_gap_dev,
                                 aspher_adjoint_info* info) {
    return guard([&] {
        ASPHER_HIP(hipDeviceSynchronize());
        adjoint_device(s, res, grad_pressure_dev, prm, grad_gap_dev, info);
        ASPHER_HIP(hipStreamSynchronize(s->ctx->stream));
    });
}


void aspher_ensemble_params_default(aspher_ensemble_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->tol = 1e-5;
    p->max_iter = 5000;
    p->use_pr = 1;
    std::strcpy(p->precision, "float");
    std::strcpy(p->precond, "auto");
    p->check_every = 8;
    p->use_graph = 1;
    p->autotune = 1;
    p->keep_best = 1;
    p->stall_limit = 200;
    p->light_result = 0;
}

int aspher_solve_ensemble(int n, const int* backend, const int* grid_size,
                          const double* domain_size, const double* E_star,
                          const double* p_nominal, const double* const* gaps,
                          const aspher_ensemble_params* params, aspher_result** results,
                          aspher_ensemble_stats* stats, double* instance_seconds) {
    return guard([&] {
        ensemble_run(n, backend, grid_size, domain_size, E_star, p_nominal, gaps, false,
                     params, results, stats, instance_seconds);
    });
}

int aspher_solve_ensemble_device(int n, const int* backend, const int* grid_size,
                                 const double* domain_size, const double* E_star,
                                 const double* p_nominal, const double* const* gaps_dev,
                                 const aspher_ensemble_params* params,
                                 aspher_result** results, aspher_ensemble_stats* stats,
                                 double* instance_seconds) {
    return guard([&] {
        ensemble_run(n, backend, grid_size, domain_size, E_star, p_nominal, gaps_dev, true,
                     params, results, stats, instance_seconds);
    });
}


int aspher_tangential_create(int grid_size, double domain_size, double E_star, double nu,
                             aspher_tangential** out) {
    return guard([&] {
        if (!out) throw std::invalid_argument("null output handle");
        auto t = std::make_unique<aspher_tangential>();
        t->ctx = std::make_unique<Context>();
        t->op = std::make_unique<TangentialOperator>(grid_size, domain_size, E_star, nu);
        t->op->build(t->ctx->stream);
        t->Ns = grid_size;
        *out = t.release();
    });
}

void aspher_tangential_destroy(aspher_tangential* t) {
    if (!t) return;
    try {
        t->op.reset();
        delete t;
    } catch (...) {
    }
}

int aspher_tangential_matvec(aspher_tangential* t, const double* q, double* u) {
    return guard([&] {
        if (!t || !q || !u) throw std::invalid_argument("null pointer");
        const std::size_t n2 = (std::size_t)2 * t->Ns * t->Ns;
        DevBuf<double> x(n2);
        hipStream_t st = t->ctx->stream;
        ASPHER_HIP(hipMemcpyAsync(x.p, q, sizeof(double) * n2, hipMemcpyHostToDevice, st));
        t->op->apply(x.p, x.p, st);
        ASPHER_HIP(hipMemcpyAsync(u, x.p, sizeof(double) * n2, hipMemcpyDeviceToHost, st));
        ASPHER_HIP(hipStreamSynchronize(st));
    });
}

int aspher_tangential_matvec_device(aspher_tangential* t, const void* q, void* u, int precision) {
    return guard([&] {
        if (!t || !q || !u) throw std::invalid_argument("null pointer");
        ASPHER_HIP(hipDeviceSynchronize());
        hipStream_t st = t->ctx->stream;
        if (precision == 64)
            t->op->apply(static_cast<const double*>(q), static_cast<double*>(u), st);
        else if (precision == 32)
            t->op->apply(static_cast<const float*>(q), static_cast<float*>(u), st);
        else
            throw std::invalid_argument("precision must be 64 or 32");
        ASPHER_HIP(hipStreamSynchronize(st));
    });
}

int aspher_tangential_precond_apply(int grid_size, double nu, int precision, const double* g,
                                    const uint8_t* mask, int remove_mean, double* z) {
    return guard([&] {
        if (!g || !mask || !z) throw std::invalid_argument("null pointer");
        if (precision != 64 && precision != 32)
            throw std::invalid_argument("precision must be 64 or 32");
        TangentialPrecond pc(grid_size, nu);
        Context ctx;
        hipStream_t st = ctx.stream;
        const std::size_t N = (std::size_t)grid_size * grid_size;
        DevBuf<std::uint8_t> m(N);
        DevBuf<double> gd(2 * N), zd(2 * N);
        ASPHER_HIP(hipMemcpyAsync(m.p, mask, N, hipMemcpyHostToDevice, st));
        ASPHER_HIP(hipMemcpyAsync(gd.p, g, sizeof(double) * 2 * N, hipMemcpyHostToDevice, st));
        if (precision == 64) {
            pc.apply<double>(gd.p, m.p, remove_mean != 0, zd.p, st);
        } else {
            DevBuf<float> gf(2 * N), zf(2 * N);
            VecOps<float>::from_double((std::int64_t)(2 * N), gd.p, 0.0, gf.p, st);
            pc.apply<float>(gf.p, m.p, remove_mean != 0, zf.p, st);
            VecOps<float>::to_double((std::int64_t)(2 * N), zf.p, zd.p, st);
            ASPHER_HIP(hipStreamSynchronize(st));
        }
        ASPHER_HIP(hipMemcpyAsync(z, zd.p, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, st));
        ASPHER_HIP(hipStreamSynchronize(st));
    });
}

int aspher_cerruti_points(int n, const double* x, const double* y, const double* a,
                          const double* b, double nu, double* uxx, double* uxy) {
    return guard([&] {
        if (n < 0 || !x || !y || !a || !b || !uxx || !uxy)
            throw std::invalid_argument("bad arguments");
        if (n == 0) return;
        Context ctx;
        hipStream_t st = ctx.stream;
        DevBuf<double> d(6 * (std::size_t)n);
        const double* src[4] = {x, y, a, b};
        for (int k = 0; k < 4; ++k)
            ASPHER_HIP(hipMemcpyAsync(d.p + (std::size_t)k * n, src[k], sizeof(double) * n,
                                      hipMemcpyHostToDevice, st));
        cerruti_points(n, d.p, d.p + n, d.p + 2 * (std::size_t)n, d.p + 3 * (std::size_t)n, nu,
                       d.p + 4 * (std::size_t)n, d.p + 5 * (std::size_t)n, st);
        ASPHER_HIP(hipMemcpyAsync(uxx, d.p + 4 * (std::size_t)n, sizeof(double) * n,
                                  hipMemcpyDeviceToHost, st));
        ASPHER_HIP(hipMemcpyAsync(uxy, d.p + 5 * (std::size_t)n, sizeof(double) * n,
                                  hipMemcpyDeviceToHost, st));
        ASPHER_HIP(hipStreamSynchronize(st));
    });
}

int aspher_cerruti_tables(int grid_size, double domain_size, double E_star, double nu,
                          double* xx, double* xy) {
    return guard([&] {
        if (grid_size <= 0 || !xx || !xy) throw std::invalid_argument("bad arguments");
        if (!(nu >= 0.0 && nu < 1.0)) throw std::invalid_argument("nu must be in [0, 1)");
        Context ctx;
        hipStream_t st = ctx.stream;
        const std::size_t N = (std::size_t)grid_size * grid_size;
        DevBuf<double> a(N), b(N);
        cerruti_tables(grid_size, domain_size, E_star, nu, a.p, b.p, st);
        ASPHER_HIP(hipMemcpyAsync(xx, a.p, sizeof(double) * N, hipMemcpyDeviceToHost, st));
        ASPHER_HIP(hipMemcpyAsync(xy, b.p, sizeof(double) * N, hipMemcpyDeviceToHost, st));
        ASPHER_HIP(hipStreamSynchronize(st));
    });
}

int aspher_rough_surface(int grid_size, double hurst, double k_low, double k_high, double rms,
                         uint64_t seed, double* out) {
    return guard([&] {
        if (!out) throw std::invalid_argument("null output");
        if (grid_size <= 0) throw std::invalid_argument("grid_size must be positive");
        const std::size_t N = (std::size_t)grid_size * grid_size;
        Context ctx;
        DevBuf<double> d(N);
        rough_surface_device(grid_size, hurst, k_low, k_high, rms, seed, d.p, ctx.stream);
        ASPHER_HIP(hipMemcpyAsync(out, d.p, sizeof(double) * N, hipMemcpyDeviceToHost,
                                  ctx.stream));
        ASPHER_HIP(hipStreamSynchronize(ctx.stream));
    });
}

int aspher_rough_surface_device(int grid_size, double hurst, double k_low, double k_high,
                                double rms, uint64_t seed, double* out_dev) {
    return guard([&] {
        if (!out_dev) throw std::invalid_argument("null output");
        ASPHER_HIP(hipDeviceSynchronize());
        Context ctx;
        rough_surface_device(grid_size, hurst, k_low, k_high, rms, seed, out_dev, ctx.stream);
        ASPHER_HIP(hipStreamSynchronize(ctx.stream));
    });
}

int aspher_device_malloc(int64_t bytes, void** ptr) {
    return guard([&] {
        if (!ptr || bytes <= 0) throw std::invalid_argument("bad allocation request");
        ASPHER_HIP(hipMalloc(ptr, (std::size_t)bytes));
    });
}

int aspher_device_free(void* ptr) {
    return guard([&] {
        if (ptr) ASPHER_HIP(hipFree(ptr));
    });
}

int aspher_memcpy_htod(void* dst, const void* src, int64_t bytes) {
    return guard([&] {
        if (!dst || !src || bytes < 0) throw std::invalid_argument("bad copy");
        ASPHER_HIP(hipMemcpy(dst, src, (std::size_t)bytes, hipMemcpyHostToDevice));
    });
}

int aspher_copy_device(void* dst, const void* src, int64_t bytes) {
    return guard([&] {
        if (!dst || !src || bytes < 0 || bytes % 16)
            throw std::invalid_argument("bad copy (bytes must be a multiple of 16)");
        copy16(bytes, src, dst, nullptr);
        ASPHER_HIP(hipDeviceSynchronize());
    });
}

int aspher_memcpy_dtoh(void* dst, const void* src, int64_t bytes) {
    return guard([&] {
        if (!dst || !src || bytes < 0) throw std::invalid_argument("bad copy");
        ASPHER_HIP(hipMemcpy(dst, src, (std::size_t)bytes, hipMemcpyDeviceToHost));
    });
}

} // extern "C"
