// Certified Polonsky–Keer projected CG (src/contact_solver.cpp:150-587) on
// device-resident state, and the load-simplex projection of its
// identification step (src/contact_solver.cpp:31-60).
#include "engine.hpp"
#include "kernels.hpp"
#include "reduce.hpp"
#include "timer.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>

namespace aspher {

namespace {

using Clock = std::chrono::steady_clock;
inline double since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

const int kS1[1] = {RSUM};
const int kS2[2] = {RSUM, RSUM};
const int kS3[3] = {RSUM, RSUM, RSUM};
const int kS5[5] = {RSUM, RSUM, RSUM, RSUM, RSUM};

struct Diag {
    double alpha = 0, gmin = 0, pk = 0, G = 0, fw = 0, pen = 0, load = 0, psum = 0, pv = 0;
    long long nc = 0;
    bool finite = true;
};

Diag read_diag(const double* h) {
    Diag d;
    d.alpha = h[S_ALPHA_REF];
    d.gmin = h[S_GMIN_REF];
    d.pk = h[S_PK];
    d.G = h[S_GC];
    d.fw = h[S_FW];
    d.pen = h[S_PEN];
    d.load = h[S_LOAD];
    d.psum = h[S_PSUM];
    d.pv = h[S_PV];
    d.nc = (long long)h[S_NC];
    d.finite = h[S_FINITE] != 0.0;
    return d;
}


} // namespace

template <typename Real>
double project_load_simplex_dev(Context& ctx, std::int64_t N, Real* y, double P_total) {
    if (N <= 0) return 0.0;
    hipStream_t s = ctx.stream;
    const int nblk = vec_blocks(N);
    const double* hs = ctx.host_scal;
    static const int d1[1] = {S_R0};
    static const int d2[2] = {S_R0, S_R1};
    VecOps<Real>::sum(N, y, ctx.part.p, nblk, s);
    finalize<Real>(ctx.part.p, nblk, 1, kS1, d1, ctx.scal.p, STEP_NONE, s);
    ctx.read_scal();
    // Michelot: theta from the mean of the retained entries; each counting
    // pass drops the entries at or below it (at most 64 passes)
    double retained = (double)N;
    double theta = (hs[S_R0] - P_total) / retained;
    for (int pass = 0; pass < 64; ++pass) {
        VecOps<Real>::count_above(N, y, theta, ctx.part.p, nblk, s);
        finalize<Real>(ctx.part.p, nblk, 2, kS2, d2, ctx.scal.p, STEP_NONE, s);
        ctx.read_scal();
        const double above = hs[S_R1];
        if (above == 0.0 || above == retained) break;
        retained = above;
        theta = (hs[S_R0] - P_total) / retained;
    }
    VecOps<Real>::project(N, y, theta, s);
    return theta;
}

template double project_load_simplex_dev<double>(Context&, std::int64_t, double*, double);
template double project_load_simplex_dev<float>(Context&, std::int64_t, float*, double);

template <typename Real>
void solve_impl(Context& ctx, Operator& op, const Real* g0, Real p_bar,
                const SolveOptions& opt, Precond* pc, Real* p_init, Result& res) {
    using V = VecOps<Real>;
    constexpr bool is_double = std::is_same_v<Real, double>;
    hipStream_t s = ctx.stream;
    const int Ns = op.Ns();
    const std::int64_t N = (std::int64_t)Ns * Ns;
    if (N == 0 || !(p_bar > Real(0)))
        throw std::invalid_argument("solve_contact: empty gap or p_bar <= 0");
    if (!(opt.tol > 0.0)) throw std::invalid_argument("solve_contact: tol must be positive");
    if (opt.max_iter < 0) throw std::invalid_argument("solve_contact: max_iter must be >= 0");
    const auto t_start = Clock::now();
    const int nblk = vec_blocks(N);
    double* scal = ctx.scal.p;
    double* part = ctx.part.p;
    const double* hs = ctx.host_scal;

    // non-finite input is rejected before any operator call
    {
        static const int ops[3] = {RMIN, RMAX, RSUM};
        static const int dst[3] = {S_R0, S_R1, S_R2};
        V::stats(N, g0, part, nblk, s);
        finalize<Real>(part, nblk, 3, ops, dst, scal, STEP_NONE, s);
        ctx.read_scal();
    }
    if (hs[S_R2] != 0.0)
        throw std::invalid_argument("solve_contact: g0 contains non-finite values");
    const double gmin_v = hs[S_R0], gmax_v = hs[S_R1];

    const double P_total = static_cast<double>(p_bar) * N;
    double g_ref = opt.g_ref;
    if (!(g_ref > 0.0)) g_ref = gmax_v - gmin_v;
    if (!(g_ref > 0.0)) g_ref = 1.0;
    const double p_ref = opt.p_ref > 0.0 ? opt.p_ref : static_cast<double>(p_bar);
    const double load_tol = opt.load_tol > 0.0 ? opt.load_tol : (is_double ? 1e-12 : 5e-7);
    const double tol = opt.tol;
    const double req_tol = opt.requested_tol > 0.0 ? opt.requested_tol : tol;
    double datum = 0.0;
    bool subtract = false;
    switch (opt.datum_mode) {
        case SolveOptions::Datum::automatic:
            datum = gmin_v + 0.5 * (gmax_v - gmin_v);
            subtract = true;
            break;
        case SolveOptions::Datum::solver:
            datum = opt.datum;
            subtract = true;
            break;
        case SolveOptions::Datum::caller:
            datum = opt.datum;
            break;
    }
    const Real gshift = subtract ? static_cast<Real>(datum) : Real(0);

    // device state: p, t, u, g, g_prev, z, r (+ mask, centred gap). The
    // pressure rotates through two buffers (three with keep_best): the
    // projected update writes out of place, so marking the best iterate is
    // an index, not a copy of p.
    DevBuf<Real> bg0s(N), bp(N), bp2(N), bt(N), bu(N), bg(N), bgp(N), bz(N), br(N), bbest;
    DevBuf<std::uint8_t> bmask(N);
    Real *P = bp.p, *T_ = bt.p, *U = bu.p, *G = bg.p, *GP = bgp.p, *Z = bz.p, *R = br.p;
    Real* g0s = bg0s.p;
    std::uint8_t* mask = bmask.p;
    V::shift(N, g0, static_cast<double>(gshift), g0s, s);

    if (p_init) {
        V::copy(N, p_init, P, s);
        V::clamp_sum(N, P, part, nblk, s);
        static const int d1[1] = {S_R0};
        finalize<Real>(part, nblk, 1, kS1, d1, scal, STEP_NONE, s);
        ctx.read_scal();
        const double ssum = hs[S_R0];
        if (ssum > 0.0) V::scale(N, P, static_cast<double>(static_cast<Real>(P_total / ssum)), s);
        else V::fill(N, P, static_cast<double>(p_bar), s);
    } else {
        V::fill(N, P, static_cast<double>(p_bar), s);
    }
    V::fill(N, T_, 0.0, s);
    V::fill(N, GP, 0.0, s);
    V::fill(N, G, 0.0, s);
    {
        const int idx[6] = {S_PTOT, S_GREF, S_USEPR, S_PBAR, S_GOLD, S_DELTA};
        const double val[6] = {P_total, g_ref, opt.use_pr ? 1.0 : 0.0,
                               static_cast<double>(p_bar), 1.0, 0.0};
        set_scal(scal, 6, idx, val, s);
    }
    if (opt.keep_best) bbest.alloc(N);
    Real* pb[3] = {bp.p, bp2.p, bbest.p};
    const int nbuf = opt.keep_best ? 3 : 2;
    int cur = 0, best_idx = -1;
    auto pick_out = [&]() -> int {
        for (int i = 0; i < nbuf; ++i)
            if (i != cur && i != best_idx) return i;
        throw std::logic_error("solve_contact: no free pressure buffer");
    };

    long long mv = 0, pcount = 0;
    // every eighth apply is timed: the applies of a level are alike, and each
    // event record idles the device for a few microseconds
    EventTimer tm_mv(8), tm_pc(8);
    auto sync_read = [&]() {
        ctx.read_scal();
        tm_mv.drain();
        tm_pc.drain();
    };
    auto apply = [&](const Real* x, Real* y) {
        tm_mv.start(s);
        op.apply(x, y, s);
        tm_mv.stop(s);
        ++mv;
    };
    auto evaluate = [&]() {
        static const int o1[3] = {RSUM, RSUM, RMIN};
        static const int d1[3] = {S_GSUM, S_NC, S_VMIN};
        static const int d2[5] = {S_PV, S_GCS, S_E, S_GC, S_PSUM};
        V::eval1(N, U, g0s, P, G, part, nblk, s);
        finalize<Real>(part, nblk, 3, o1, d1, scal, STEP_EVAL1, s);
        V::eval2(N, P, G, mask, scal, part, nblk, s);
        finalize<Real>(part, nblk, 5, kS5, d2, scal, STEP_EVAL2, s);
    };
    auto set_delta_restart = [&]() {
        const int idx[2] = {S_DELTA, S_GOLD};
        const double val[2] = {0.0, 1.0};
        set_scal(scal, 2, idx, val, s);
    };

    // feasible identification step (src/contact_solver.cpp:314-364)
    double s_id = 0.0;
    bool u_fresh = false;
    auto identification_step = [&]() -> bool {
        static const int d3[3] = {S_R0, S_R1, S_R2};
        static const int d2[2] = {S_R0, S_R1};
        V::objective(N, P, U, g0s, part, nblk, s);
        finalize<Real>(part, nblk, 3, kS3, d3, scal, STEP_NONE, s);
        sync_read();
        const double f_cur = hs[S_R0], pp = hs[S_R1], pu = hs[S_R2];
        if (!(s_id > 0.0)) s_id = (pu > 0.0 && pp > 0.0) ? pp / pu : 1.0;
        for (int trial = 0; trial < 12; ++trial) {
            V::axpy_id(N, P, G, Z, s_id, s);
            project_load_simplex_dev<Real>(ctx, N, Z, P_total);
            apply(Z, R);
            V::idcheck(N, Z, R, g0s, G, P, part, nblk, s);
            finalize<Real>(part, nblk, 2, kS2, d2, scal, STEP_NONE, s);
            sync_read();
            const double f_try = hs[S_R0], gd = hs[S_R1];
            if (std::isfinite(f_try) && gd < 0.0 && f_try <= f_cur + 1e-4 * gd) {
                {
                    const int o = pick_out();   // Z is reused; keep p in the rotation
                    V::copy(N, Z, pb[o], s);
                    cur = o;
                    P = pb[o];
                }
                std::swap(U, R);
                u_fresh = true;
                set_delta_restart();
                V::fill(N, T_, 0.0, s);
                V::fill(N, GP, 0.0, s);
                ++res.identification_steps;
                if (trial == 0) s_id *= 2.0;
                return true;
            }
            s_id *= 0.25;
        }
        return false;
    };

    bool precond_on = pc && pc->kind() != Precond::Kind::none;
    bool precond_dropped = false;

    // one CG step: direction, second operator apply, line search, projected
    // update and load rescale; the update is predicated on the device and
    // writes out of place into the returned pressure buffer (`commit` makes
    // it the iterate)
    auto enqueue_cg = [&]() -> int {
        static const int dg[3] = {S_G, S_GPR, S_GG};
        static const int dl1[2] = {S_RSUM, S_NUM};
        static const int dl2[1] = {S_DEN};
        static const int ou[2] = {RMAX, RSUM};
        static const int du[2] = {S_CHANGED, S_TOTAL};
        const bool has_pc = precond_on;
        if (has_pc) {
            tm_pc.start(s);
            pc->apply<Real>(G, mask, Z, scal, part, s, -1, false);
            tm_pc.stop(s);
            ++pcount;
            if (opt.precond_cost_gate > 0.0 && pcount >= 1 && mv >= 1) {
                // opt-in cost gate (src/contact_solver.cpp:132-146)
                sync_read();
                const double per_pc = tm_pc.seconds() / pcount;
                const double per_mv = tm_mv.seconds() / mv;
                if (per_mv > 0.0 && per_pc > opt.precond_cost_gate * per_mv) {
                    precond_on = false;
                    precond_dropped = true;
                    const int idx[1] = {S_DELTA};
                    const double val[1] = {0.0};
                    set_scal(scal, 1, idx, val, s);
                }
            }
        }
        V::cgdot(N, G, GP, Z, mask, scal, has_pc, part, nblk, s);
        finalize<Real>(part, nblk, 3, kS3, dg, scal, STEP_CGDOT, s);
        V::tupdate(N, G, nullptr, Z, mask, T_, scal, has_pc, s);
        apply(T_, R);
        V::line1(N, R, G, T_, mask, part, nblk, s);
        finalize<Real>(part, nblk, 2, kS2, dl1, scal, STEP_LINE1, s);
        V::line2(N, R, T_, mask, scal, part, nblk, s);
        finalize<Real>(part, nblk, 1, kS1, dl2, scal, STEP_LINE2, s);
        const int o = pick_out();
        V::update(N, P, pb[o], T_, G, scal, part, nblk, s);
        finalize<Real>(part, nblk, 2, ou, du, scal, STEP_UPDATE, s);
        V::rescale(N, pb[o], scal, s);
        return o;
    };
    auto commit = [&](int o) {
        cur = o;
        P = pb[o];
    };

    // The CG step of an iteration is enqueued before the host has read that
    // iteration's diagnostics, so the device runs it while the host decides.
    // It is kept when the host takes the CG branch (the common case). On any
    // other branch it is dropped: the loop then ends (the step wrote only the
    // CG state and a pressure buffer that is not the iterate or the best
    // one), or an identification step follows, which resets that CG state
    // (t, g_prev, delta, G_old) before use. The opt-in cost gate synchronises
    // inside the step, so it runs without this overlap.
    const bool speculate = !(opt.precond_cost_gate > 0.0);
    int spec_o = -1;
    long long spec_mv0 = 0, spec_pc0 = 0;
    auto discard_spec = [&]() {
        if (spec_o < 0) return;
        mv = spec_mv0;   // a dropped step's operator and preconditioner
        pcount = spec_pc0;   // applies are not counted
        spec_o = -1;
    };

    Status status = Status::max_iterations;
    std::string reason;
    double best_merit = std::numeric_limits<double>::infinity();
    double seen_merit = std::numeric_limits<double>::infinity();
    int stall = 0, forced_id = 0;
    const int stall_limit = opt.stall_limit > 0 ? opt.stall_limit : 200;
    bool have_best = false;
    bool pending = false;
    Diag d;

    // ── device-controlled iterations (small grids, opt-in) ──────────────────
    // Where the kernels are short, the host round trip of every iteration
    // (read the diagnostics, decide, enqueue) sets the pace. On grids of at
    // most opt.graph_max_points points (0 by default: the capture and
    // instantiation outweigh the replay savings at the iteration counts of
    // nested levels) the loop starts on the device: one iteration -- apply, evaluation, device
    // control (single_control), best-iterate copy, CG step -- is captured as
    // a hipGraph and replayed, the host reading the scalars only every
    // kDevPoll replays. Every kernel of the replay is gated on S_ACTIVE, so
    // replays after the control has stopped are empty. Buffers are fixed
    // here: the update is in place, g_prev is a copy and the best iterate a
    // copy into its own buffer; the arithmetic is that of the host loop.
    // When the control stops on anything but convergence, the host loop
    // takes over at that iteration (its evaluation done) with the device's
    // bookkeeping, and decides it exactly as it would have.
    int it0 = 0;
    bool resumed = false, device_converged = false;
    const bool device_loop = opt.graph_max_points > 0 && N <= opt.graph_max_points &&
                             speculate && !opt.record_history && opt.max_iter > 1 &&
                             fault_mode() == FAULT_NONE;
    if (device_loop) {
        constexpr int kDevPoll = 8;
        op.prepare<Real>(s);   // nothing may be allocated while capturing
        if (pc) pc->prepare<Real>();
        {
            const int idx[8] = {S_ACTIVE, S_STATUS, S_HOST, S_PENDING, S_ITER, S_MAXIT, S_TOL,
                                S_LOADTOL};
            const double val[8] = {1.0, 0.0, 0.0, 0.0, 0.0, (double)opt.max_iter, tol, load_tol};
            set_scal(scal, 8, idx, val, s);
            const int idx2[6] = {S_BEST, S_SAVEBEST, S_SEEN, S_STALL, S_STALL_LIMIT, S_KEEPBEST};
            const double val2[6] = {1e300, 0.0, 1e300, 0.0, (double)stall_limit,
                                    opt.keep_best ? 1.0 : 0.0};
            set_scal(scal, 6, idx2, val2, s);
            const int idx3[1] = {S_OK};
            const double val3[1] = {1.0};
            set_scal(scal, 1, idx3, val3, s);
        }
        Ens eg;   // single solve, every launch gated on S_ACTIVE
        eg.gscal = scal;
        eg.gate = S_ACTIVE;
        Real* PB = opt.keep_best ? bbest.p : nullptr;
        const bool has_pc = precond_on;
        auto iteration = [&](bool timed) {
            static const int o1[3] = {RSUM, RSUM, RMIN};
            static const int d1[3] = {S_GSUM, S_NC, S_VMIN};
            static const int d2[5] = {S_PV, S_GCS, S_E, S_GC, S_PSUM};
            static const int dg[3] = {S_G, S_GPR, S_GG};
            static const int dl1[2] = {S_RSUM, S_NUM};
            static const int dl2[1] = {S_DEN};
            static const int ou[2] = {RMAX, RSUM};
            static const int du[2] = {S_CHANGED, S_TOTAL};
            if (timed) tm_mv.start(s);
            op.apply_gated(P, U, s, scal, S_ACTIVE);
            if (timed) tm_mv.stop(s);
            V::eval1(N, U, g0s, P, G, part, nblk, s, eg);
            finalize<Real>(part, nblk, 3, o1, d1, scal, STEP_EVAL1, s, eg);
            V::eval2(N, P, G, mask, scal, part, nblk, s, eg);
            finalize<Real>(part, nblk, 5, kS5, d2, scal, STEP_S_EVAL2, s, eg);
            if (has_pc) {
                if (timed) tm_pc.start(s);
                pc->apply<Real>(G, mask, Z, scal, part, s, S_ACTIVE, false);
                if (timed) tm_pc.stop(s);
            }
            V::cgdot(N, G, GP, Z, mask, scal, has_pc, part, nblk, s, eg);
            finalize<Real>(part, nblk, 3, kS3, dg, scal, STEP_CGDOT, s, eg);
            V::tupdate(N, G, GP, Z, mask, T_, scal, has_pc, s, eg);
            if (timed) tm_mv.start(s);
            op.apply_gated(T_, R, s, scal, S_ACTIVE);
            if (timed) tm_mv.stop(s);
            V::line1(N, R, G, T_, mask, part, nblk, s, eg);
            finalize<Real>(part, nblk, 2, kS2, dl1, scal, STEP_LINE1, s, eg);
            V::line2(N, R, T_, mask, scal, part, nblk, s, eg);
            finalize<Real>(part, nblk, 1, kS1, dl2, scal, STEP_LINE2, s, eg);
            // the best-iterate copy rides on the update (P is unchanged
            // until then)
            V::update(N, P, P, T_, G, scal, part, nblk, s, eg, PB, S_SAVEBEST);
            finalize<Real>(part, nblk, 2, ou, du, scal, STEP_UPDATE, s, eg);
            V::rescale(N, P, scal, s, eg);
        };
        hipGraph_t graph = nullptr;
        hipGraphExec_t gexec = nullptr;
        auto release = [&]() {
            if (gexec) (void)hipGraphExecDestroy(gexec);
            if (graph) (void)hipGraphDestroy(graph);
            gexec = nullptr;
            graph = nullptr;
        };
        try {
            // the first iteration eagerly: its applies are the timers' samples
            iteration(true);
            sync_read();
            if (hs[S_ACTIVE] != 0.0) {
                ASPHER_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try {
                    iteration(false);
                } catch (...) {
                    hipGraph_t g2 = nullptr;
                    (void)hipStreamEndCapture(s, &g2);
                    if (g2) (void)hipGraphDestroy(g2);
                    throw;
                }
                ASPHER_HIP(hipStreamEndCapture(s, &graph));
                ASPHER_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
                // each replay either runs an iteration or (stopped) nothing;
                // the control stops by the iteration budget at the latest
                for (long long r = 0; r <= opt.max_iter && hs[S_ACTIVE] != 0.0;) {
                    for (int k = 0; k < kDevPoll; ++k, ++r) ASPHER_HIP(hipGraphLaunch(gexec, s));
                    sync_read();
                }
                if (hs[S_ACTIVE] != 0.0)
                    throw std::logic_error("solve_contact: device iterations overran the budget");
            }
        } catch (...) {
            release();
            throw;
        }
        release();
        // iterations 0 .. it0-1 took the CG branch on the device; iteration
        // it0 was evaluated (one apply) and stopped there
        it0 = (int)hs[S_ITER];
        mv += 2LL * it0 + 1;
        if (has_pc) pcount += it0;
        for (long long k = 0; k < 2LL * it0 + 1 - 2; ++k) tm_mv.skip();
        if (has_pc)
            for (long long k = 0; k < (long long)it0 - 1; ++k) tm_pc.skip();
        d = read_diag(hs);
        if (hs[S_STATUS] == 1.0) {
            status = Status::converged;
            u_fresh = true;
            device_converged = true;   // no host iterations
        } else {
            resumed = true;
            pending = hs[S_PENDING] != 0.0;
            if (opt.keep_best && hs[S_BEST] < 1e300) {
                // the best iterate is the copy in its own buffer
                best_merit = hs[S_BEST];
                best_idx = 2;
                have_best = true;
            }
            if (hs[S_SEEN] < 1e300) seen_merit = hs[S_SEEN];
            stall = (int)hs[S_STALL];
        }
    }

    // g_prev by rotation: the gradient buffer of the previous iteration
    // becomes g_prev and the next evaluation writes the other one. It equals
    // the reference's copy (g_prev = g in the CG step) whenever g_prev is
    // used: beta carries the factor delta, which is 1 only after a CG step
    // without support change, i.e. when the previous iteration took the CG
    // branch; after a restart (delta = 0) its value does not enter. Both
    // buffers start at zero, so it is always finite.
    int it = it0;
    for (; !device_converged && it < opt.max_iter; ++it) {
        if (resumed) {
            // the device loop stopped here: this iteration is evaluated and
            // its diagnostics are on the host
            resumed = false;
        } else {
            if (!u_fresh) apply(P, U);
            u_fresh = false;
            std::swap(G, GP);
            evaluate();
            if (speculate) {
                ctx.copy_scal_async();
                spec_mv0 = mv;
                spec_pc0 = pcount;
                spec_o = enqueue_cg();
                ctx.wait_scal();
                tm_mv.drain();
                tm_pc.drain();
            } else {
                sync_read();
            }
            d = read_diag(hs);
        }
        if (pending) {
            pending = false;
            if (hs[S_OK] == 0.0) {
                // the previous step had no usable curvature: run its
                // identification step now, on the identical state. Its
                // update left the pressure unchanged, so this iteration's
                // apply re-applied the same pressure; the reference runs the
                // identification step within the failed iteration and never
                // makes that apply, so it is not counted.
                discard_spec();
                --mv;
                if (!identification_step()) {
                    status = Status::nonpositive_curvature;
                    --it;
                    break;
                }
                u_fresh = false;
                evaluate();
                sync_read();
                d = read_diag(hs);
            }
        }
        if (!d.finite) {
            discard_spec();
            status = Status::nonfinite;
            break;
        }
        if (opt.record_history) res.error_history.push_back(d.pk);
        if (d.load <= load_tol && d.fw <= tol && d.pen <= tol) {
            discard_spec();
            status = Status::converged;
            u_fresh = true;
            break;
        }
        const double merit = std::max(d.fw, d.pen);
        if (opt.keep_best && merit < best_merit) {
            best_merit = merit;
            best_idx = cur;
            have_best = true;
        }
        if (merit < seen_merit * (1.0 - 1e-4)) {
            seen_merit = merit;
            stall = 0;
        } else if (++stall >= stall_limit) {
            discard_spec();
            if (forced_id < 3 && identification_step()) {
                ++forced_id;
                stall = 0;
                continue;
            }
            status = Status::stagnated;
            reason = "stall";
            break;
        }
        if (d.nc >= 2) {
            if (spec_o >= 0) {
                commit(spec_o);
                spec_o = -1;
            } else {
                commit(enqueue_cg());
            }
            pending = true;
        } else {
            discard_spec();
            if (!identification_step()) {
                status = Status::nonpositive_curvature;
                break;
            }
        }
    }
    discard_spec();
    if (pending) {
        // the loop ended on the iteration budget right after a CG step
        sync_read();
        if (hs[S_OK] == 0.0 && !identification_step()) {
            status = Status::nonpositive_curvature;
            it = std::max(0, it - 1);
        }
    }

    // final, freshly evaluated diagnostics on the returned pressure
    if (status != Status::converged && opt.keep_best && have_best) {
        const double merit = std::max(d.fw, d.pen);
        if (!(merit <= best_merit)) {
            cur = best_idx;
            P = pb[cur];
            u_fresh = false;
            res.returned_best = true;
        }
    }
    if (!u_fresh) apply(P, U);
    if (status == Status::converged && fault_mode() == FAULT_FINAL_CHECK)
        fault_before_final_check<Real>(U, 10.0 * g_ref, s);
    evaluate();
    sync_read();
    d = read_diag(hs);
    if (!d.finite && status != Status::nonfinite) status = Status::nonfinite;
    if (status == Status::converged && !(d.load <= load_tol && d.fw <= tol && d.pen <= tol)) {
        status = Status::verification_failed;
        reason = "final_check";
    }
    if (status == Status::converged && req_tol < tol && !opt.allow_tolerance_relaxation) {
        if (!(d.fw <= req_tol && d.pen <= req_tol)) {
            status = Status::stagnated;
            reason = "precision_limit";
        }
    }
    {
        static const int d3[3] = {S_R0, S_R1, S_R2};
        V::objective(N, P, U, g0s, part, nblk, s);
        finalize<Real>(part, nblk, 3, kS3, d3, scal, STEP_NONE, s);
    }
    res.pressure.alloc(N);
    V::to_double(N, P, res.pressure.p, s);
    if (!opt.light) {
        res.displacement.alloc(N);
        res.gap.alloc(N);
        V::to_double(N, U, res.displacement.p, s);
        V::to_double(N, G, res.gap.p, s);
    } else {
        res.displacement.reset();
        res.gap.reset();
    }
    sync_read();
    res.Ns = Ns;
    res.objective = hs[S_R0] + datum * P_total;
    res.approach = d.alpha + datum;
    res.iterations = it;
    res.contact_fraction = double(d.nc) / N;
    res.mean_pressure = d.psum / N;
    res.status = status;
    res.status_reason = reason;
    res.converged = status == Status::converged;
    res.error = res.pk_error = d.pk;
    res.fw_gap = d.G;
    res.fw_error = d.fw;
    res.load_error = d.load;
    res.pressure_violation = 0.0;
    res.penetration_error = d.pen;
    res.vmin_local = d.gmin + d.alpha;
    res.requested_tol = req_tol;
    res.effective_tol = tol;
    res.p_ref = p_ref;
    res.g_ref = g_ref;
    res.matvec_count = mv;
    res.precond_count = pcount;
    res.precond_dropped = precond_dropped;
    const long long rb = sizeof(Real);
    res.mem_cg_state = 8 * N * rb;
    res.mem_best = opt.keep_best ? N * rb : 0;
    res.mem_mask = N;
    res.mem_output = (rb == 8 ? 0 : N * 8) + (opt.light ? 0 : 2 * N * 8);
    res.mem_peak = res.mem_cg_state + res.mem_best + res.mem_mask + res.mem_output;
    res.time_matvec = tm_mv.seconds();
    res.time_precond = tm_pc.seconds();
    res.time_total = since(t_start);
}

template void solve_impl<double>(Context&, Operator&, const double*, double,
                                 const SolveOptions&, Precond*, double*, Result&);
template void solve_impl<float>(Context&, Operator&, const float*, float,
                                const SolveOptions&, Precond*, float*, Result&);

} // namespace aspher
