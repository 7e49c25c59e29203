// C++17 host side of the spectral contact path: operators, preconditioners,
// the certified Polonsky–Keer state machine and the nested-grid driver. All
// vectors live on the device for the whole solve; the host reads one block
// of FP64 diagnostics per iteration.
#pragma once

#include "common.hpp"
#include "spectral.hpp"

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace aspher {

// RAII device allocation.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    std::size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(std::size_t count) { alloc(count); }
    ~DevBuf() { reset(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    void alloc(std::size_t count) {
        reset();
        if (count) p = static_cast<T*>(dev_alloc(sizeof(T) * count));
        n = count;
    }
    void reset() {
        if (p) dev_free(p, sizeof(T) * n);
        p = nullptr;
        n = 0;
    }
    std::size_t bytes() const { return sizeof(T) * n; }
};

// Termination state (include/contact_solver.hpp:13-21).
enum class Status : int {
    converged = 0, stagnated, max_iterations, nonpositive_curvature, nonfinite,
    verification_failed, resource_limit,
};
const char* to_string(Status s);

// Solver policy (include/contact_solver.hpp:134-191).
struct SolveOptions {
    double tol = 1e-8;
    double requested_tol = 0.0;
    int max_iter = 5000;
    bool use_pr = true;
    bool light = false;
    bool record_history = false;
    bool keep_best = true;
    double load_tol = 0.0;
    int stall_limit = 200;
    double precond_cost_gate = 0.0;
    bool allow_tolerance_relaxation = false;
    double p_ref = 0.0;
    double g_ref = 0.0;
    enum class Datum { automatic, solver, caller } datum_mode = Datum::automatic;
    double datum = 0.0;
    // grids of at most this many points start on the device-controlled loop
    // (one captured iteration replayed as a hipGraph) instead of launching
    // every kernel (0: never, the default). Measured on MI355X it loses at
    // the iteration counts of the nested levels and of single solves up to
    // ~100 iterations: capture and instantiation cost ~0.45 ms per solve,
    // and a replayed iteration of ~25 dependent small kernels saves only
    // ~20 us at 64^2 and nothing from 256^2 up.
    long long graph_max_points = 0;
};

struct Stage {
    std::string name, precision;
    int q = 0;
    double requested_tol = 0.0, effective_tol = 0.0;
    int iterations = 0;
    long long matvec_count = 0;
    double seconds = 0.0;
    Status status = Status::max_iterations;
    double fw_error = 0.0, penetration_error = 0.0;
    bool precond_dropped = false;
};

// ContactResult (include/contact_solver.hpp:25-148). Fields stay on the
// device (FP64) until the caller copies them out.
struct Result {
    int Ns = 0;
    DevBuf<double> pressure, displacement, gap;
    double approach = 0.0, objective = 0.0, error = 0.0;
    int iterations = 0;
    bool converged = false;
    double contact_fraction = 0.0, mean_pressure = 0.0;
    std::vector<double> error_history;
    int active_rounds = 0;
    bool active_fallback = false;
    Status status = Status::max_iterations;
    std::string status_reason;
    double pk_error = 0.0, fw_gap = 0.0, fw_error = 0.0, load_error = 0.0;
    double pressure_violation = 0.0, penetration_error = 0.0;
    // min over the solve's points of u + g0 in the solver's frame (the
    // restricted certificate's vmin_local, src/contact_solver.cpp:1020-1030)
    double vmin_local = 0.0;
    double requested_tol = 0.0, effective_tol = 0.0, p_ref = 0.0, g_ref = 0.0;
    std::string validation_scope = "solve_operator";
    std::string operator_error_kind = "unavailable";
    long long mem_cg_state = 0, mem_best = 0, mem_mask = 0, mem_output = 0, mem_peak = 0;
    double time_total = 0.0, time_matvec = 0.0, time_precond = 0.0, time_build = 0.0;
    double time_coarse = 0.0, time_verification = 0.0, time_candidate = 0.0,
           time_output = 0.0;
    std::vector<Stage> stage_stats;
    long long matvec_count = 0, verification_matvec_count = 0, precond_count = 0;
    int identification_steps = 0;
    bool returned_best = false, precond_dropped = false;
};

// Influence operator on the device: free-space (padded Love convolution),
// doubly periodic (2/(E*|q|)) or dense (direct Love sum, small grids).
class Operator {
public:
    enum class Kind { free_space, periodic, dense };
    Operator(Kind kind, int Ns, double L, double E_star);
    ~Operator();
    Operator(const Operator&) = delete;
    Operator& operator=(const Operator&) = delete;

    // free space: FP64 Love table + FP64 padded r2c of its embedding,
    // realness check, (Ns+1)^2 real spectrum (src/fft_operator.cpp:18-69)
    void build(hipStream_t s);
    void build_single(hipStream_t s);    // FP32 spectrum copy + FP32 plan
    void release_single();
    bool built() const { return built_; }

    // y = S x on device buffers (x, y may alias for the spectral kinds)
    void apply(const double* x, double* y, hipStream_t s);
    void apply(const float* x, float* y, hipStream_t s);
    // same, recording 4 events around the three spectral launches
    void apply_timed(const double* x, double* y, hipStream_t s, hipEvent_t* ev);
    void apply_timed(const float* x, float* y, hipStream_t s, hipEvent_t* ev);
    // same, skipped on the device while the device scalar gscal[gate] is 0
    // (register-blocked spectral plans; other plans and the dense kind apply
    // unconditionally)
    template <typename T>
    void apply_gated(const T* x, T* y, hipStream_t s, const double* gscal, int gate);

    // y = S x (spectral kinds) with the intermediate spectra in the caller's
    // workspace `ws` (ws_bytes bytes) instead of the plan's own. Allocates
    // nothing: the plan and, for free space, the kernel spectrum of this
    // precision must already exist (`prepared`), else std::logic_error;
    // a workspace smaller than workspace_external<T>() is invalid_argument.
    template <typename T>
    void apply_external(const T* x, T* y, void* ws, std::size_t ws_bytes, hipStream_t s);
    template <typename T> bool prepared() const;
    // set-up for apply_external in precision T (allocates: the plan and, for
    // free space, the kernel spectrum in the plan's order)
    template <typename T> void prepare(hipStream_t s);
    // caller workspace bytes of apply_external (0 if the plan of this
    // precision does not exist yet)
    template <typename T> std::size_t workspace_external() const;

    Kind kind() const { return kind_; }
    int Ns() const { return Ns_; }
    double L() const { return L_; }
    double E() const { return E_; }
    std::size_t bytes_spectrum() const;
    std::size_t bytes_scratch() const;
    double build_seconds = 0.0;
    // device Love table (dense kind; free space keeps it only if asked)
    const double* table() const { return table_.p; }
    // free-space kernel spectrum in the Symbol::table layout that plan `pl`
    // reads (made on demand per precision); pitch receives its row pitch
    template <typename T>
    const T* table_for(const SpectralPlan<T>& pl, hipStream_t s, int& pitch);

private:
    template <typename T>
    void apply_t(const T* x, T* y, hipStream_t s, hipEvent_t* ev,
                 const SpecBatch* gate = nullptr);
    template <typename T> SpectralPlan<T>& plan(hipStream_t s);

    Kind kind_;
    int Ns_;
    double L_, E_;
    bool built_ = false;
    DevBuf<double> table_;
    DevBuf<double> kh64_;      // (Ns+1) x (Ns+1), [kx][ky]
    DevBuf<float> kh32_;
    DevBuf<double> khp64_;     // (Ns+1) x 2Ns in the register-blocked column order
    DevBuf<float> khp32_;
    std::unique_ptr<SpectralPlan<double>> plan64_;
    std::unique_ptr<SpectralPlan<float>> plan32_;
    DevBuf<double> tmp64_;     // dense aliasing scratch
    DevBuf<float> tmp32_;
};

// Tangential (Cerruti) 2x2 block operator on stacked [x; y] components
// (src/tangential_operator.cpp:16-157): FP64 tables and spectra, FP32 copies
// on demand. Power-of-two grids 16 .. 2048.
class TangentialOperator {
public:
    TangentialOperator(int Ns, double L, double E_star, double nu);
    ~TangentialOperator();
    TangentialOperator(const TangentialOperator&) = delete;
    TangentialOperator& operator=(const TangentialOperator&) = delete;
    void build(hipStream_t s);
    void build_single(hipStream_t s);
    // u = C q, 2 Ns^2 values each; q and u may alias
    void apply(const double* q, double* u, hipStream_t s);
    void apply(const float* q, float* u, hipStream_t s);
    int Ns() const { return Ns_; }
    double build_seconds = 0.0;

private:
    template <typename T> void apply_t(const T* q, T* u, hipStream_t s);
    int Ns_;
    double L_, E_, nu_;
    bool built_ = false;
    DevBuf<double> k64_[3];      // unique (Ns+1)^2 spectra: xx, yy, xy
    DevBuf<double> perm64_[3];   // register-order tables of the FP64 plan
    DevBuf<float> perm32_[3];
    std::unique_ptr<SpectralPlan<double>> plan64_;
    std::unique_ptr<SpectralPlan<float>> plan32_;
};

// Tangential Fourier preconditioner (src/fourier_precond.cpp:234-359): masked
// scatter of both components, per-mode |k|/Ns^2 (I + gamma k k^T/|k|^2),
// gamma = nu/(1-nu), masked gather, optional per-component mask-mean removal.
class TangentialPrecond {
public:
    TangentialPrecond(int Ns, double nu);
    template <typename T>
    void apply(const T* g, const std::uint8_t* mask, bool remove_mean, T* z, hipStream_t s);

private:
    int Ns_;
    double nu_;
    DevBuf<double> scal_, part_;
    std::unique_ptr<SpectralPlan<double>> plan64_;
    std::unique_ptr<SpectralPlan<float>> plan32_;
};

// Preconditioner: none, 13-tap stencil (src/stencil_precond.cpp) or full-grid
// Fourier |k| (src/fourier_precond.cpp).
class Precond {
public:
    enum class Kind { none, stencil, fourier };
    Precond(Kind kind, int Ns, int radius = 2);
    Kind kind() const { return kind_; }
    // z = M^-1 g on the mask (mean-zeroed); uses scal[S_GMEAN] and leaves
    // scal[S_ZMEAN] for the consumer. Returns false for `none`.
    // gate >= 0: skipped on the device while scal[gate] is 0; zero_off =
    // false: the stencil leaves z unwritten off the mask (callers that read
    // z only on the mask)
    template <typename T>
    void apply(const T* g, const std::uint8_t* mask, T* z, double* scal, double* part,
               hipStream_t s, int gate = -1, bool zero_off = true);
    // builds what apply<T> would build on its first call (the Fourier plan)
    template <typename T> void prepare();
    const std::vector<int>& tap_dx() const { return dx_; }
    const std::vector<int>& tap_dy() const { return dy_; }
    const std::vector<double>& tap_w() const { return w_; }
    // device tap table (dx, dy interleaved) and weights of the stencil kind
    const int* taps_dev() const { return tapd_.p; }
    const double* weights_dev() const { return tapw_.p; }
    int radius() const { return R_; }
    float fourier_norm() const { return norm_; }

private:
    Kind kind_;
    int Ns_;
    int R_ = 0;
    std::vector<int> dx_, dy_;
    std::vector<double> w_;
    DevBuf<int> tapd_;
    DevBuf<double> tapw_;
    std::unique_ptr<SpectralPlan<double>> plan64_;
    std::unique_ptr<SpectralPlan<float>> plan32_;
    float norm_ = 0.0f;
};

// Host-side stencil weights: w = F^-1|k| on an Nw x Nw grid (DC zeroed),
// normalised by w(0,0), sampled on the l2 disc of radius R.
void stencil_weights(int Ns, int R, std::vector<int>& dx, std::vector<int>& dy,
                     std::vector<double>& w);

// Per-call device context: stream + reusable scalar/partial buffers.
struct Context {
    hipStream_t stream = nullptr;
    DevBuf<double> scal, part;
    double* host_scal = nullptr;   // pinned mirror
    Context();
    ~Context();
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    void read_scal();              // D2H of the scalar block + stream sync
    // D2H of the scalar block without waiting: the stream runs on, and
    // wait_scal() blocks only until that copy has landed
    void copy_scal_async();
    void wait_scal();
    hipEvent_t scal_ready = nullptr;
};

// Load-simplex projection (src/contact_solver.cpp:31-60) of N device values
// in place: y <- max(y - theta, 0) with sum = P_total, theta by Michelot's
// threshold iteration (FP64 sums, at most 64 counting passes, one host read
// each). Returns theta.
template <typename Real>
double project_load_simplex_dev(Context& ctx, std::int64_t N, Real* y, double P_total);

// One certified PK solve (src/contact_solver.cpp:150-587) on device inputs:
// g0 (N values of Real), optional p_init (consumed, may be modified).
template <typename Real>
void solve_impl(Context& ctx, Operator& op, const Real* g0, Real p_bar,
                const SolveOptions& opt, Precond* pc, Real* p_init, Result& res);

// Nested-grid driver parameters (include/nested_solve.hpp:11-119; the fft
// and periodic branches).
struct NestedParams {
    int coarsest = 64;
    int q = 6;
    int leaf_side = 8;
    bool precond = true;
    std::string precond_engine = "auto";   // auto | stencil | fft
    int precond_radius = 2;
    double precond_occupancy_max = 0.4;
    double precond_cost_gate = 0.0;
    double coarse_tol = 1e-4;
    std::string precision = "";            // double | float | float_then_double | ""
    bool single_precision = false;
    double float_floor = 2e-6;
    bool allow_tolerance_relaxation = false;
    bool light_result = false;
    std::string backend = "h2";
    bool record_error_history = false;
    bool active_set = false;
    double active_delta = 0.05;
    int active_halo = 2;
    int active_max_rounds = 5;
    bool active_all_levels = true;
    double active_occupancy_max = 0.4;
    long long graph_max_points = 0;   // SolveOptions::graph_max_points
};

bool stencil_for_level(const std::string& engine, bool precond, double prev_occupancy,
                       double occupancy_max);

// One restricted (active-set) level of the nested driver
// (src/nested_solve.cpp:117-516) on the free-space operator. glvl: the
// level's FP64 gap (device), datum: the hierarchy's datum; p_init: the
// prolonged pressure (Ns^2, consumed); coarse_gap: the gap field of the level
// below ((Ns/2)^2). The result carries the global certificate.
template <typename Real>
void active_level(Context& ctx, Operator& op, Precond* pc, const double* glvl, double datum,
                  double p_bar, double lvl_tol, int max_iter, bool use_pr,
                  const NestedParams& np, DevBuf<double>& p_init,
                  const DevBuf<double>& coarse_gap, bool record_history, bool light,
                  double g_ref, Result& res);

// gap: device FP64 Ns*Ns. Operators are built per level (or taken from the
// process-wide cache when `reuse_operators` is on).
void solve_nested(int Ns, double L, double E_star, const double* gap_dev, double p_bar,
                  double tol, int max_iter, bool use_pr, const NestedParams& np,
                  Result& res, bool reuse_operators = false);

// Adjoint of a converged solve (no counterpart in the reference). For a
// loss with dL/dp = w it returns dL/dg0 (Ns^2, zero off the contact set) and
// dL/dp_nominal, from the transposed KKT system of the contact QP on the
// support C = {p > 0}:
//   [S_CC 1; 1' 0][mu; nu] = [w_C; 0],  dL/dg0 = -mu on C,
//   dL/dP = nu, dL/dp_nominal = nu Ns^2,
// solved by preconditioned CG on the zero-mean subspace of C with the
// forward operator (same spectrum) and preconditioner; iterates are
// recomputed, nothing of the forward CG is stored.
struct AdjointOptions {
    double tol = 1e-10;        // relative residual ||r|| / ||r0||
    int max_iter = 5000;
};

struct AdjointInfo {
    bool converged = false;
    int iterations = 0;
    long long n_contact = 0;
    double rel_residual = 0.0;
    double multiplier = 0.0;        // nu = dL/dP
    double grad_p_nominal = 0.0;    // nu * Ns^2
    long long matvec_count = 0, precond_count = 0;
    double seconds = 0.0;
    std::string status;             // converged | max_iterations | nonpositive_curvature
};

// p: converged pressure (device FP64), w: dL/dp (device FP64), grad: output
// dL/dg0 (device FP64); all Ns^2 values.
void adjoint_impl(Context& ctx, Operator& op, Precond* pc, const double* p, const double* w,
                  double* grad, const AdjointOptions& opt, AdjointInfo& info);

// Self-affine Gaussian rough surface on the device (the benchmark recipe:
// random phase, amplitude |m|^-(1+H) for k_low <= |m| <= k_high, real part
// of the inverse 2-D DFT, mean removed, scaled to rms), with the phase of
// frequency (kx, ky) drawn as SplitMix64 output number ky*Ns + kx of a
// generator seeded with `seed`. out: Ns^2 FP64 (device), row-major.
void rough_surface_device(int Ns, double hurst, double k_low, double k_high, double rms,
                          unsigned long long seed, double* out, hipStream_t s);

// Process-wide operator cache (kind, Ns, L, E) -> built operator.
std::shared_ptr<Operator> cached_operator(Operator::Kind kind, int Ns, double L, double E,
                                          hipStream_t s);
void clear_operator_cache();

} // namespace aspher
