// Nested-grid (cascadic) continuation for the fft and periodic backends: the
// behaviour of solve_contact_nested (src/nested_solve.cpp:518-990) on
// device-resident fields.
//
// The driver has three parts.
//  1. validate_request(): ordered rule tables, checked before anything is
//     allocated; the first failing rule raises with the reference's message.
//  2. plan_levels(): the hierarchy coarsest*2^k .. Ns is laid out up front as
//     one LevelPlan row per level: side, working precision, the tolerance the
//     level certifies and the one it was asked for, whether relaxation is a
//     stage policy there, the stencil radius, and whether an FP64 polish stage
//     follows. Nothing in a row depends on a solve result.
//  3. the level loop executes the table. A Carry record hands the prolonged
//     pressure, the gap field kept for a restricted successor and the contact
//     fraction measured below from one level to the next. Only the choices
//     that need that measured fraction are made at run time: stencil or
//     Fourier preconditioner, restricted (active-set) level, keeping the gap.
// The gap pyramid is built on the device by 2x2 restriction of the caller's
// FP64 gap; the finest level reads the caller's buffer in place.
#include "engine.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <initializer_list>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>

namespace aspher {

namespace {

using Clock = std::chrono::steady_clock;
double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

enum class Policy { fp64, fp32, fp32_then_fp64 };

struct Rule {
    bool ok;
    const char* msg;
};

void enforce(std::initializer_list<Rule> rules) {
    for (const Rule& r : rules)
        if (!r.ok) throw std::invalid_argument(r.msg);
}

// '' and 'double' both defer to the legacy single_precision flag
// (python/bindings.cpp:404-411 with src/nested_solve.cpp:584-587)
Policy resolve_policy(const NestedParams& np) {
    if (np.precision == "float") return Policy::fp32;
    if (np.precision == "float_then_double") return Policy::fp32_then_fp64;
    if (np.precision.empty() || np.precision == "double")
        return np.single_precision ? Policy::fp32 : Policy::fp64;
    throw std::invalid_argument(
        "precision must be 'double', 'float', 'float_then_double' or '' "
        "(defer to single_precision)");
}

Policy validate_request(int Ns, double L, double E_star, double p_bar, double tol,
                        int max_iter, const NestedParams& np) {
    const std::int64_t N64 = (std::int64_t)Ns * Ns;
    const bool leaf_pow2 = np.leaf_side > 0 && (np.leaf_side & (np.leaf_side - 1)) == 0;
    enforce({
        {Ns > 0, "solve_contact_nested: Ns must be positive"},
        {L > 0.0 && std::isfinite(L), "solve_contact_nested: L must be positive"},
        {E_star > 0.0 && std::isfinite(E_star), "solve_contact_nested: E_star must be positive"},
        {N64 <= (std::int64_t)std::numeric_limits<int>::max(),
         "solve_contact_nested: Ns*Ns exceeds the int index range"},
        {p_bar > 0.0 && std::isfinite(p_bar), "solve_contact_nested: p_nominal must be positive"},
        {tol > 0.0 && std::isfinite(tol), "solve_contact_nested: tol must be positive"},
        {max_iter >= 0, "solve_contact_nested: max_iter must be >= 0"},
        {np.coarsest > 0, "solve_contact_nested: coarsest must be positive"},
        {np.coarsest <= Ns, "solve_contact_nested: coarsest > Ns"},
        {np.q >= 2, "solve_contact_nested: q must be >= 2"},
        {leaf_pow2, "solve_contact_nested: leaf_side must be a positive power of two"},
        {np.coarse_tol > 0.0, "solve_contact_nested: coarse_tol must be positive"},
    });
    if (np.active_set)
        enforce({
            {np.active_halo >= 0, "solve_contact_nested: active_halo must be >= 0"},
            {np.active_max_rounds >= 1, "solve_contact_nested: active_max_rounds must be >= 1"},
            {np.active_delta >= 0.0, "solve_contact_nested: active_delta must be >= 0"},
            {np.active_occupancy_max > 0.0,
             "solve_contact_nested: active_occupancy_max must be > 0"},
        });
    enforce({{np.precond_occupancy_max > 0.0,
              "solve_contact_nested: precond_occupancy_max must be > 0"}});
    const Policy pol = resolve_policy(np);
    // float_then_double with active_set: rejected by the reference (its h2
    // polish would need a restricted double stage); with fft the float stage
    // runs restricted and the polish on the full grid
    enforce({
        {np.float_floor > 0.0, "solve_contact_nested: float_floor must be positive"},
        {!(pol == Policy::fp32_then_fp64 && np.active_set && np.backend != "fft"),
         "solve_contact_nested: float_then_double does not yet route its polish through "
         "the active-set path; use float_only or double_only with active_set=true"},
    });
    return pol;
}

// Backend checks that need the level count (src/nested_solve.cpp:597-617).
void validate_backend(const NestedParams& np, std::size_t nlevels) {
    enforce({{np.backend == "h2" || np.backend == "fft" || np.backend == "periodic",
              "solve_contact_nested: backend must be 'h2', 'fft', or 'periodic'"}});
    if (np.active_set)
        enforce({
            // the restricted solve runs on the free-space operator (the
            // reference's masked-matvec backend 'h2' is not part of this build)
            {np.backend == "fft",
             "solve_contact_nested: active_set requires backend='fft' (or the reference's "
             "'h2', not part of this build)"},
            {nlevels >= 2, "solve_contact_nested: active_set needs a coarse level (Ns > coarsest)"},
        });
    enforce({
        {np.backend != "h2",
         "solve_contact_nested: backend 'h2' is not part of this build; use backend='fft' "
         "(free space) or backend='periodic'"},
        {np.precond_engine == "auto" || np.precond_engine == "stencil" ||
             np.precond_engine == "fft",
         "precond_engine must be 'auto', 'stencil' or 'fft'"},
    });
}

struct LevelPlan {
    int side = 0;
    bool finest = false;
    bool below_finest = false;  // the level right under the finest
    bool fp32 = false;          // working precision of the level's first stage
    bool polish = false;        // an FP64 stage to the requested tol follows
    double stage_tol = 0.0;     // tolerance the first stage certifies
    double asked_tol = 0.0;     // tolerance asked of this level
    bool may_relax = false;     // a looser certificate is stage policy here
    int radius = 2;             // stencil radius (never wider than the level)
    bool history = false;
};

// Sides coarsest * 2^k up to Ns, with the per-level policy of
// src/nested_solve.cpp:703-790: coarse levels are solved to coarse_tol, a
// float stage never certifies below float_floor, and under
// float_then_double the finest float stage only identifies the contact at
// the floor before the FP64 polish takes the requested tolerance.
std::vector<LevelPlan> plan_levels(int Ns, double tol, Policy pol, const NestedParams& np) {
    std::vector<LevelPlan> table;
    for (long long side = np.coarsest; side <= Ns; side *= 2) {
        LevelPlan row;
        row.side = (int)side;
        table.push_back(row);
    }
    if (table.empty() || table.back().side != Ns)
        throw std::invalid_argument("solve_contact_nested: Ns must equal coarsest * 2^k");
    const std::size_t last = table.size() - 1;
    for (std::size_t i = 0; i <= last; ++i) {
        LevelPlan& row = table[i];
        row.finest = i == last;
        row.below_finest = i + 1 == last;
        row.fp32 = pol != Policy::fp64;
        row.polish = row.finest && pol == Policy::fp32_then_fp64;
        row.asked_tol = row.finest ? tol : np.coarse_tol;
        if (row.polish) row.stage_tol = np.float_floor;
        else row.stage_tol = row.fp32 ? std::max(row.asked_tol, np.float_floor) : row.asked_tol;
        row.may_relax = !row.finest || row.polish || np.allow_tolerance_relaxation;
        row.radius = std::min({np.precond_radius, std::max(1, row.side / 4), np.leaf_side});
        row.history = row.finest && np.record_error_history;
    }
    return table;
}

// What one level hands to the next.
struct Carry {
    DevBuf<double> pressure;    // prolonged warm start (FP64, next level's side)
    DevBuf<double> gap_field;   // this level's gap, kept for a restricted successor
    double occupancy = -1.0;    // contact fraction measured below; -1: none yet
};

Stage stage_row(const std::string& name, const char* precision, int q, double asked,
                double certified, const Result& r) {
    Stage st;
    st.name = name;
    st.precision = precision;
    st.q = q;
    st.requested_tol = asked;
    st.effective_tol = certified;
    st.iterations = r.iterations;
    st.matvec_count = r.matvec_count;
    st.seconds = r.time_total;
    st.status = r.status;
    st.fw_error = r.fw_error;
    st.penetration_error = r.penetration_error;
    st.precond_dropped = r.precond_dropped;
    return st;
}

// The finest level's reported work is the sum of its stages.
void fold_first_stage(Result& total, const Result& first) {
    total.iterations += first.iterations;
    total.matvec_count += first.matvec_count;
    total.precond_count += first.precond_count;
    total.time_matvec += first.time_matvec;
    total.time_precond += first.time_precond;
    total.mem_peak = std::max(total.mem_peak, first.mem_peak);
    if (!first.error_history.empty()) {
        std::vector<double> h = first.error_history;
        h.insert(h.end(), total.error_history.begin(), total.error_history.end());
        total.error_history = std::move(h);
    }
}

std::mutex g_cache_mu;
std::map<std::tuple<int, int, double, double>, std::shared_ptr<Operator>> g_cache;

} // namespace

bool stencil_for_level(const std::string& engine, bool precond, double prev_occupancy,
                       double occupancy_max) {
    // src/nested_solve.cpp:27-39
    if (!precond) return false;
    if (engine == "stencil") return true;
    if (engine == "fft") return false;
    return prev_occupancy < 0.0 || prev_occupancy < occupancy_max;
}

std::shared_ptr<Operator> cached_operator(Operator::Kind kind, int Ns, double L, double E,
                                          hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    auto key = std::make_tuple((int)kind, Ns, L, E);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) return it->second;
    auto op = std::make_shared<Operator>(kind, Ns, L, E);
    op->build(s);
    g_cache[key] = op;
    return op;
}

void clear_operator_cache() {
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        g_cache.clear();
    }
    // the freed operators' memory, and every other cached block, goes back
    // to the runtime
    (void)hipDeviceSynchronize();
    dev_release_cached();
}

void solve_nested(int Ns, double L, double E_star, const double* gap_dev, double p_bar,
                  double tol, int max_iter, bool use_pr, const NestedParams& np, Result& res,
                  bool reuse_operators) {
    const Policy pol = validate_request(Ns, L, E_star, p_bar, tol, max_iter, np);
    const std::vector<LevelPlan> levels = plan_levels(Ns, tol, pol, np);
    validate_backend(np, levels.size());

    Context ctx;
    hipStream_t s = ctx.stream;
    const auto t_begin = Clock::now();
    const std::int64_t N = (std::int64_t)Ns * Ns;

    // one datum and one gap scale for the whole hierarchy, from the finest
    // gap: the range midpoint and max(range, p_bar L / E*)
    double g_lo = 0.0, g_hi = 0.0;
    {
        static const int ops[3] = {RMIN, RMAX, RSUM};
        static const int dst[3] = {S_R0, S_R1, S_R2};
        const int nb = vec_blocks(N);
        VecOps<double>::stats(N, gap_dev, ctx.part.p, nb, s);
        finalize<double>(ctx.part.p, nb, 3, ops, dst, ctx.scal.p, STEP_NONE, s);
        ctx.read_scal();
        if (ctx.host_scal[S_R2] != 0.0)
            throw std::invalid_argument("solve_contact_nested: gap contains non-finite values");
        g_lo = ctx.host_scal[S_R0];
        g_hi = ctx.host_scal[S_R1];
    }
    const double datum = g_lo + 0.5 * (g_hi - g_lo);
    const double g_ref = std::max(g_hi - g_lo, p_bar * L / E_star);

    // gap pyramid: pyramid[i] is level i's restricted gap (the finest level
    // reads the caller's buffer)
    const std::size_t nlev = levels.size();
    std::vector<DevBuf<double>> pyramid(nlev);
    for (std::size_t i = nlev - 1; i-- > 0;) {
        const double* finer = i + 1 == nlev - 1 ? gap_dev : pyramid[i + 1].p;
        pyramid[i].alloc((std::size_t)levels[i].side * levels[i].side);
        restrict2(levels[i + 1].side, finer, pyramid[i].p, s);
    }

    const Operator::Kind kind =
        np.backend == "fft" ? Operator::Kind::free_space : Operator::Kind::periodic;
    const std::string engine =
        (np.backend == "periodic" && np.precond_engine == "auto") ? std::string("fft")
                                                                  : np.precond_engine;
    auto options_for = [&](const LevelPlan& row, bool light) {
        SolveOptions o;
        o.tol = row.stage_tol;
        o.requested_tol = row.asked_tol;
        o.max_iter = max_iter;
        o.use_pr = use_pr;
        o.light = light;
        o.record_history = row.history;
        o.keep_best = !light;
        o.allow_tolerance_relaxation = row.may_relax;
        o.g_ref = g_ref;
        o.datum = datum;
        // FP32 stages get the datum removed in the cast, before any rounding;
        // FP64 stages subtract it on the fly
        o.datum_mode = row.fp32 ? SolveOptions::Datum::caller : SolveOptions::Datum::solver;
        o.precond_cost_gate = np.precond_cost_gate;
        o.graph_max_points = np.graph_max_points;
        return o;
    };

    Carry carry;
    std::vector<Stage> stages;
    double t_build = 0.0, t_coarse = 0.0, t_candidate = 0.0;

    for (std::size_t i = 0; i < nlev; ++i) {
        const LevelPlan& row = levels[i];
        const int n = row.side;
        const std::int64_t nn = (std::int64_t)n * n;
        const double* gl = row.finest ? gap_dev : pyramid[i].p;

        const auto t_b = Clock::now();
        std::shared_ptr<Operator> op;
        if (reuse_operators) {
            op = cached_operator(kind, n, L, E_star, s);
        } else {
            op = std::make_shared<Operator>(kind, n, L, E_star);
            op->build(s);
        }
        const Precond::Kind pk =
            !np.precond ? Precond::Kind::none
            : stencil_for_level(engine, true, carry.occupancy, np.precond_occupancy_max)
                ? Precond::Kind::stencil
                : Precond::Kind::fourier;
        Precond pc(pk, n, row.radius);
        t_build += seconds_since(t_b);

        // run-time choices from the contact fraction measured below: a
        // level is restricted once a coarser one has predicted its contact
        // and that fraction is under the gate; a coarse level keeps its gap
        // field when its successor is predicted to be restricted
        const bool gate_open = carry.occupancy < np.active_occupancy_max;
        const bool restricted = np.active_set && carry.pressure.p && carry.gap_field.p &&
                                gate_open && (row.finest || np.active_all_levels);
        const bool keep_gap = np.active_set && !row.finest && gate_open &&
                              (np.active_all_levels || row.below_finest);
        const bool light = row.finest ? np.light_result : !keep_gap;
        const SolveOptions opt = options_for(row, light);

        Result r;
        if (restricted) {
            if (row.fp32) {
                op->build_single(s);
                active_level<float>(ctx, *op, &pc, gl, datum, p_bar, row.stage_tol, max_iter,
                                    use_pr, np, carry.pressure, carry.gap_field, row.history,
                                    light, g_ref, r);
            } else {
                active_level<double>(ctx, *op, &pc, gl, datum, p_bar, row.stage_tol, max_iter,
                                     use_pr, np, carry.pressure, carry.gap_field, row.history,
                                     light, g_ref, r);
            }
            r.requested_tol = row.asked_tol;
            t_candidate += r.time_candidate;
        } else if (row.fp32) {
            op->build_single(s);
            DevBuf<float> g32(nn), p32;
            VecOps<float>::from_double(nn, gl, datum, g32.p, s);
            if (carry.pressure.p) {
                p32.alloc(nn);
                VecOps<float>::from_double(nn, carry.pressure.p, 0.0, p32.p, s);
            }
            carry.pressure.reset();
            solve_impl<float>(ctx, *op, g32.p, static_cast<float>(p_bar), opt, &pc, p32.p, r);
        } else {
            solve_impl<double>(ctx, *op, gl, p_bar, opt, &pc, carry.pressure.p, r);
        }
        carry.gap_field.reset();
        carry.pressure.reset();
        carry.occupancy = r.contact_fraction;
        stages.push_back(stage_row((row.finest ? "finest:" : "coarse:") + std::to_string(n) +
                                       (restricted ? "(active)" : ""),
                                   row.fp32 ? "float" : "double", np.q, row.asked_tol,
                                   row.stage_tol, r));

        if (row.polish) {
            // fresh FP64 solve with the same operator, warm-started from the
            // float pressure, to the requested tolerance; no conjugacy state
            // crosses the precision switch
            if (!reuse_operators) op->release_single();
            LevelPlan fp64_row = row;
            fp64_row.fp32 = false;
            fp64_row.stage_tol = tol;
            fp64_row.asked_tol = tol;
            fp64_row.may_relax = np.allow_tolerance_relaxation;
            const SolveOptions popt = options_for(fp64_row, np.light_result);
            DevBuf<double> warm = std::move(r.pressure);
            Result fin;
            solve_impl<double>(ctx, *op, gl, p_bar, popt, &pc, warm.p, fin);
            stages.push_back(stage_row("polish:" + std::to_string(n), "double", np.q, tol, tol,
                                       fin));
            fold_first_stage(fin, r);
            r = std::move(fin);
        }

        if (row.finest) {
            res = std::move(r);
        } else {
            t_coarse += r.time_total;
            pyramid[i].reset();
            carry.pressure.alloc((std::size_t)4 * nn);
            prolong2(n, r.pressure.p, carry.pressure.p, s);
            if (keep_gap && r.gap.p) carry.gap_field = std::move(r.gap);
        }
    }
    ASPHER_HIP(hipStreamSynchronize(s));
    res.stage_stats = std::move(stages);
    res.time_build = t_build;
    res.time_coarse = t_coarse;
    res.time_candidate = t_candidate;
    res.time_total = seconds_since(t_begin);
}

} // namespace aspher
