// Ensemble solver: many independent certified Polonsky–Keer solves
// (src/contact_solver.cpp:150-587, one per instance) advanced together.
//
// Layout in HBM: the instances are ordered by group (kind, Ns, L, E*) and
// their vectors concatenated, instance i at element offset off[i]; scalar
// blocks (S_COUNT doubles) and partial blocks (pstride doubles) per
// instance. One iteration is a fixed launch sequence:
//   per group : batched matvec u = S p          (3 launches, instances in y)
//   all       : evaluate pass 1 + finalize      (one launch each)
//   all       : evaluate pass 2 + finalize with the device-side control
//   all       : best-iterate copy (predicated per instance)
//   per group : preconditioner (stencil work queue / batched Fourier)
//   all       : zsum finalize, CG direction dots + finalize, direction update
//   per group : batched matvec r = S t
//   all       : line search (2 passes + finalizes), projected update +
//               finalize, load rescale
// Every "all" launch walks a work queue of (instance, chunk) items, so grids
// of 256^2 and 2048^2 share a launch without idle blocks; instances whose
// flag S_ACTIVE is 0 drop out at the top of every kernel. The sequence has
// no host dependency and is captured in a HIP graph once, then replayed.
#include "ensemble.hpp"
#include "kernels.hpp"
#include "reduce.hpp"
#include "timer.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
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

constexpr double kFloatFloor = 2e-6;   // nested float_floor default (src/nested_solve.cpp)

int stencil_tiles(int Ns) { return ((Ns + 63) / 64) * ((Ns + 15) / 16); }

// ── autotuner ─────────────────────────────────────────────────────────────

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int, int, int, int>, SpectralTuning> g_tune;

template <typename T>
double time_convolve(SpectralPlan<T>& pl, const T* x, T* y, const T* ktab, int kp,
                     Symbol sym, int batch, std::int64_t xs, std::int64_t ys, hipStream_t s,
                     double* stage_ms) {
    hipEvent_t ev[4];
    for (auto& e : ev) ASPHER_HIP(hipEventCreate(&e));
    SpecBatch sb;
    sb.count = batch;
    sb.x_stride = xs;
    sb.y_stride = ys;
    const double sc = 1e-3;
    double acc[3] = {0, 0, 0};
    const int reps = 3;
    try {
        pl.convolve(x, pl.cols_in, y, pl.cols_out, sym, ktab, kp, sc, s, RowIn::plain,
                    RowOut::plain, nullptr, nullptr, 0, nullptr, nullptr, &sb);
        for (int r = 0; r < reps; ++r) {
            pl.convolve(x, pl.cols_in, y, pl.cols_out, sym, ktab, kp, sc, s, RowIn::plain,
                        RowOut::plain, nullptr, nullptr, 0, nullptr, ev, &sb);
            ASPHER_HIP(hipEventSynchronize(ev[3]));
            for (int k = 0; k < 3; ++k) {
                float ms = 0.0f;
                ASPHER_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
                acc[k] += ms;
            }
        }
    } catch (...) {
        for (auto& e : ev) (void)hipEventDestroy(e);
        throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int k = 0; k < 3; ++k) stage_ms[k] = acc[k] / reps;
    return (acc[0] + acc[1] + acc[2]) / reps;
}

} // namespace

template <typename T>
SpectralTuning autotune_plan(int n, int rows_in, int cols_in, int rows_out, int cols_out,
                             int batch, Symbol sym, hipStream_t s) {
    const auto key = std::make_tuple(n, rows_in, cols_in, rows_out, cols_out, batch,
                                     (int)sizeof(T), (int)sym);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) return it->second;
    }
    const int L = n / 2;
    const std::size_t cs = sizeof(cx<T>);
    const std::size_t lds_max = 160 * 1024;
    std::vector<int> Bs, Ws;
    for (int B = 1; B <= 64; B *= 2)
        if (B * (L / 16) <= 512 && B * cs * (std::size_t)(L + L / 16 + 1) <= lds_max &&
            B <= std::max(rows_in, rows_out))
            Bs.push_back(B);
    for (int W = 1; W <= 64; W *= 2)
        if (W * (n / 16) <= 512 && W * cs * (std::size_t)(n + n / 16 + 1) <= lds_max &&
            W <= n / 2 + 1)
            Ws.push_back(W);
    if (Bs.empty() || Ws.empty()) return SpectralTuning{};
    const std::int64_t xs = (std::int64_t)rows_in * cols_in;
    const std::int64_t ys = (std::int64_t)rows_out * cols_out;
    DevBuf<T> x((std::size_t)xs * batch), y((std::size_t)ys * batch);
    VecOps<T>::fill(xs * batch, x.p, 0.5, s);
    DevBuf<T> kt;
    int kp = 0;
    if (sym == Symbol::table) {
        // the register-blocked layout: n/2+1 rows of n values
        kp = n;
        kt.alloc((std::size_t)(n / 2 + 1) * n);
        VecOps<T>::fill((std::int64_t)(n / 2 + 1) * n, kt.p, 1e-3, s);
    }
    SpectralTuning best{};
    double t_best_rows = 1e30, t_best_cols = 1e30;
    int B_best = Bs[0], W_best = Ws[0];
    const int W0 = Ws[Ws.size() / 2];
    for (int B : Bs) {
        SpectralPlan<T> pl(n, rows_in, cols_in, rows_out, cols_out, {B, W0, true}, batch);
        double st[3];
        time_convolve(pl, x.p, y.p, kt.p, kp, sym, batch, xs, ys, s, st);
        if (st[0] + st[2] < t_best_rows) {
            t_best_rows = st[0] + st[2];
            B_best = B;
        }
    }
    for (int W : Ws) {
        SpectralPlan<T> pl(n, rows_in, cols_in, rows_out, cols_out, {B_best, W, true}, batch);
        double st[3];
        time_convolve(pl, x.p, y.p, kt.p, kp, sym, batch, xs, ys, s, st);
        if (st[1] < t_best_cols) {
            t_best_cols = st[1];
            W_best = W;
        }
    }
    double t_on, t_off;
    {
        SpectralPlan<T> pl(n, rows_in, cols_in, rows_out, cols_out, {B_best, W_best, true},
                           batch);
        double st[3];
        t_on = time_convolve(pl, x.p, y.p, kt.p, kp, sym, batch, xs, ys, s, st);
    }
    {
        SpectralPlan<T> pl(n, rows_in, cols_in, rows_out, cols_out, {B_best, W_best, false},
                           batch);
        double st[3];
        t_off = time_convolve(pl, x.p, y.p, kt.p, kp, sym, batch, xs, ys, s, st);
    }
    best = SpectralTuning{B_best, W_best, t_on <= t_off};
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = best;
    return best;
}

template SpectralTuning autotune_plan<float>(int, int, int, int, int, int, Symbol, hipStream_t);
template SpectralTuning autotune_plan<double>(int, int, int, int, int, int, Symbol,
                                              hipStream_t);

namespace {

// One group of equal-size instances: batched operator + preconditioner.
template <typename Real>
struct Group {
    int kind = 0, Ns = 0, first = 0, count = 0;
    double L = 1.0, E = 1.0;
    std::shared_ptr<Operator> op;
    std::unique_ptr<SpectralPlan<Real>> plan;      // matvec
    std::unique_ptr<Precond> pc;
    std::unique_ptr<SpectralPlan<Real>> pc_plan;   // Fourier preconditioner
    DevBuf<WorkItem> tiles;                        // stencil work queue
    int ntiles = 0;
    const Real* ktab = nullptr;
    int kpitch = 0;
    double sym_scale = 0.0;
};

template <typename Real>
void ensemble_impl(const std::vector<EnsembleInstance>& inst, const EnsembleOptions& opt,
                   std::vector<Result>& out, EnsembleStats& st) {
    using V = VecOps<Real>;
    constexpr bool is_double = std::is_same_v<Real, double>;
    const auto t_start = Clock::now();
    const int ni = (int)inst.size();
    if (ni == 0) throw std::invalid_argument("solve_ensemble: no instances");
    if (!(opt.tol > 0.0)) throw std::invalid_argument("solve_ensemble: tol must be positive");
    if (opt.max_iter < 0) throw std::invalid_argument("solve_ensemble: max_iter must be >= 0");
    if (opt.check_every < 1) throw std::invalid_argument("solve_ensemble: check_every >= 1");
    for (const auto& e : inst) {
        if (e.kind != 0 && e.kind != 1)
            throw std::invalid_argument("solve_ensemble: kind must be fft (0) or periodic (1)");
        const int nmax = e.kind == 0 ? 4096 : 8192;
        if (e.Ns < 32 || e.Ns > nmax || (e.Ns & (e.Ns - 1)))
            throw std::invalid_argument("solve_ensemble: Ns must be a power of two in [32, " +
                                        std::to_string(nmax) + "]");
        if (!(e.p_bar > 0.0)) throw std::invalid_argument("solve_ensemble: p_bar must be > 0");
        if (!(e.L > 0.0) || !(e.E_star > 0.0))
            throw std::invalid_argument("solve_ensemble: L and E* must be positive");
        if (!e.gap) throw std::invalid_argument("solve_ensemble: null gap");
    }
    // "auto": the full-grid Fourier |q| preconditioner on both kinds. On free
    // space it costs a batched convolution per iteration against the
    // stencil's O(N) pass, but cuts the iterations: on the 64-instance
    // benchmark workload (tol 1e-5, FP32) free-space instances took 152 /
    // 204 iterations on average at 1024^2 / 2048^2 with the stencil (worst
    // 470) and 85 / 82 with it (worst 164); the ensemble ran 0.48 s -> 0.28 s.
    Precond::Kind pk_free, pk_per;
    if (opt.precond == "auto") {
        pk_free = Precond::Kind::fourier;
        pk_per = Precond::Kind::fourier;
    } else if (opt.precond == "none") {
        pk_free = pk_per = Precond::Kind::none;
    } else if (opt.precond == "fourier") {
        pk_free = pk_per = Precond::Kind::stencil;
    } else if (opt.precond == "fourier-fft") {
        pk_free = pk_per = Precond::Kind::fourier;
    } else {
        throw std::invalid_argument("precond must be 'auto', 'none', 'fourier' or 'fourier-fft'");
    }
    const bool has_pc = pk_free != Precond::Kind::none;

    Context ctx;
    hipStream_t s = ctx.stream;

    // ── order by group, offsets ──
    std::vector<int> order(ni);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const auto& A = inst[a];
        const auto& B = inst[b];
        return std::make_tuple(A.kind, A.Ns, A.L, A.E_star) <
               std::make_tuple(B.kind, B.Ns, B.L, B.E_star);
    });
    std::vector<std::int64_t> off(ni), len(ni);
    std::int64_t Ntot = 0;
    for (int k = 0; k < ni; ++k) {
        const auto& e = inst[order[k]];
        off[k] = Ntot;
        len[k] = (std::int64_t)e.Ns * e.Ns;
        Ntot += len[k];
    }
    std::vector<Group<Real>> groups;
    for (int k = 0; k < ni; ++k) {
        const auto& e = inst[order[k]];
        if (groups.empty() || groups.back().kind != e.kind || groups.back().Ns != e.Ns ||
            groups.back().L != e.L || groups.back().E != e.E_star) {
            groups.emplace_back();
            auto& g = groups.back();
            g.kind = e.kind;
            g.Ns = e.Ns;
            g.L = e.L;
            g.E = e.E_star;
            g.first = k;
        }
        ++groups.back().count;
    }
    st = EnsembleStats{};
    st.groups = (int)groups.size();

    // ── device state (sorted order) ──
    DevBuf<Real> bg0s(Ntot), bp(Ntot), bt(Ntot), bu(Ntot), bg(Ntot), bgp(Ntot), bz(Ntot),
        br(Ntot), bbest;
    DevBuf<std::uint8_t> bmask(Ntot);
    if (opt.keep_best) bbest.alloc(Ntot);
    Real *P = bp.p, *T_ = bt.p, *U = bu.p, *G = bg.p, *GP = bgp.p, *Z = bz.p, *R = br.p;

    // per-instance partial rows: vector chunks, stencil tiles, Fourier row blocks
    std::vector<int> vrows(ni), prow(ni, 1);
    std::int64_t pst_rows = 1;
    for (int k = 0; k < ni; ++k) {
        vrows[k] = (int)((len[k] + kChunk - 1) / kChunk);
        pst_rows = std::max<std::int64_t>(pst_rows, vrows[k]);
    }

    // ── setup per instance: scales, datum, shifted gap, initial pressure ──
    std::vector<double> datum(ni), gref(ni), Ptot(ni);
    {
        DevBuf<double> scal1(S_COUNT), part1((std::size_t)kMaxPartRows * kPartVals);
        for (int k = 0; k < ni; ++k) {
            const auto& e = inst[order[k]];
            const int nb = vec_blocks(len[k]);
            static const int ops[3] = {RMIN, RMAX, RSUM};
            static const int dst[3] = {S_R0, S_R1, S_R2};
            VecOps<double>::stats(len[k], e.gap, part1.p, nb, s);
            finalize<double>(part1.p, nb, 3, ops, dst, scal1.p, STEP_NONE, s);
            double h[S_COUNT];
            ASPHER_HIP(hipMemcpyAsync(h, scal1.p, sizeof(h), hipMemcpyDeviceToHost, s));
            ASPHER_HIP(hipStreamSynchronize(s));
            if (h[S_R2] != 0.0)
                throw std::invalid_argument("solve_ensemble: gap of instance " +
                                            std::to_string(order[k]) +
                                            " contains non-finite values");
            const double gmin = h[S_R0], gmax = h[S_R1];
            datum[k] = gmin + 0.5 * (gmax - gmin);
            gref[k] = gmax - gmin > 0.0 ? gmax - gmin : 1.0;
            Ptot[k] = e.p_bar * (double)len[k];
            V::from_double(len[k], e.gap, datum[k], bg0s.p + off[k], s);
            V::fill(len[k], P + off[k], e.p_bar, s);
        }
    }
    V::fill(Ntot, T_, 0.0, s);
    V::fill(Ntot, GP, 0.0, s);
    const double eff_tol = (!is_double && opt.tol < kFloatFloor) ? kFloatFloor : opt.tol;
    const double load_tol = is_double ? 1e-12 : 5e-7;
    std::vector<double> hs0((std::size_t)ni * S_COUNT, 0.0);
    for (int k = 0; k < ni; ++k) {
        double* h = hs0.data() + (std::size_t)k * S_COUNT;
        const auto& e = inst[order[k]];
        h[S_PTOT] = Ptot[k];
        h[S_GREF] = gref[k];
        h[S_USEPR] = opt.use_pr ? 1.0 : 0.0;
        h[S_PBAR] = e.p_bar;
        h[S_GOLD] = 1.0;
        h[S_DELTA] = 0.0;
        h[S_ACTIVE] = 1.0;
        h[S_MAXIT] = opt.max_iter;
        h[S_TOL] = eff_tol;
        h[S_LOADTOL] = load_tol;
        h[S_BEST] = 1e300;
        h[S_SEEN] = 1e300;
        h[S_STALL_LIMIT] = opt.stall_limit > 0 ? opt.stall_limit : 200;
        h[S_KEEPBEST] = opt.keep_best ? 1.0 : 0.0;
        h[S_OK] = 1.0;
    }

    // ── groups: operators, batched plans (autotuned), preconditioners ──
    const auto t_tune = Clock::now();
    for (auto& g : groups) {
        const Operator::Kind kind = g.kind == 0 ? Operator::Kind::free_space
                                                : Operator::Kind::periodic;
        g.op = cached_operator(kind, g.Ns, g.L, g.E, s);
        const int n = g.kind == 0 ? 2 * g.Ns : g.Ns;
        const Symbol sym = g.kind == 0 ? Symbol::table : Symbol::periodic;
        SpectralTuning tn{};
        if (opt.autotune) tn = autotune_plan<Real>(n, g.Ns, g.Ns, g.Ns, g.Ns, g.count, sym, s);
        g.plan = std::make_unique<SpectralPlan<Real>>(n, g.Ns, g.Ns, g.Ns, g.Ns, tn, g.count);
        if (g.kind == 0) {
            g.ktab = g.op->template table_for<Real>(*g.plan, s, g.kpitch);
        } else {
            const double pi = std::acos(-1.0);
            g.sym_scale = g.L / (pi * g.E * (double)g.Ns * g.Ns);
        }
        const Precond::Kind pk = g.kind == 0 ? pk_free : pk_per;
        const int R = std::min({2, std::max(1, g.Ns / 4), 8});
        g.pc = std::make_unique<Precond>(pk, g.Ns, R);
        if (pk == Precond::Kind::stencil) {
            g.ntiles = stencil_tiles(g.Ns);
            std::vector<WorkItem> w((std::size_t)g.ntiles * g.count);
            for (int l = 0; l < g.count; ++l)
                for (int t = 0; t < g.ntiles; ++t)
                    w[(std::size_t)l * g.ntiles + t] = WorkItem{t, t + 1, g.first + l, t};
            g.tiles.alloc(w.size());
            ASPHER_HIP(hipMemcpy(g.tiles.p, w.data(), sizeof(WorkItem) * w.size(),
                                 hipMemcpyHostToDevice));
            for (int l = 0; l < g.count; ++l) prow[g.first + l] = g.ntiles;
            pst_rows = std::max<std::int64_t>(pst_rows, g.ntiles);
        } else if (pk == Precond::Kind::fourier) {
            SpectralTuning tp{};
            if (opt.autotune)
                tp = autotune_plan<Real>(g.Ns, g.Ns, g.Ns, g.Ns, g.Ns, g.count,
                                         Symbol::precond, s);
            g.pc_plan = std::make_unique<SpectralPlan<Real>>(g.Ns, g.Ns, g.Ns, g.Ns, g.Ns, tp,
                                                             g.count);
            for (int l = 0; l < g.count; ++l) prow[g.first + l] = g.pc_plan->row_out_blocks();
            pst_rows = std::max<std::int64_t>(pst_rows, g.pc_plan->row_out_blocks());
        }
    }
    st.tune_seconds = since(t_tune);
    const std::int64_t pstride = pst_rows * kPartVals;

    // ── work queues and tables ──
    std::vector<WorkItem> vw;
    for (int k = 0; k < ni; ++k)
        for (int r = 0; r < vrows[k]; ++r)
            vw.push_back(WorkItem{(std::int64_t)r * kChunk,
                                  std::min<std::int64_t>(len[k], (std::int64_t)(r + 1) * kChunk),
                                  k, r});
    std::vector<int> nsv(ni);
    for (int k = 0; k < ni; ++k) nsv[k] = inst[order[k]].Ns;
    DevBuf<WorkItem> dvw(vw.size());
    DevBuf<std::int64_t> doff(ni);
    DevBuf<int> dns(ni), dvrows(ni), dprow(ni);
    DevBuf<double> scal((std::size_t)ni * S_COUNT), part((std::size_t)ni * pstride);
    ASPHER_HIP(hipMemcpy(dvw.p, vw.data(), sizeof(WorkItem) * vw.size(), hipMemcpyHostToDevice));
    ASPHER_HIP(hipMemcpy(doff.p, off.data(), sizeof(std::int64_t) * ni, hipMemcpyHostToDevice));
    ASPHER_HIP(hipMemcpy(dns.p, nsv.data(), sizeof(int) * ni, hipMemcpyHostToDevice));
    ASPHER_HIP(hipMemcpy(dvrows.p, vrows.data(), sizeof(int) * ni, hipMemcpyHostToDevice));
    ASPHER_HIP(hipMemcpy(dprow.p, prow.data(), sizeof(int) * ni, hipMemcpyHostToDevice));
    ASPHER_HIP(hipMemcpy(scal.p, hs0.data(), sizeof(double) * hs0.size(),
                         hipMemcpyHostToDevice));

    Ens ev;   // vector passes over every instance
    ev.work = dvw.p;
    ev.nwork = (int)vw.size();
    ev.off = doff.p;
    ev.sstride = S_COUNT;
    ev.pstride = pstride;
    ev.gscal = scal.p;
    ev.gate = S_ACTIVE;
    ev.ns = dns.p;
    ev.ninst = ni;
    ev.rows = dvrows.p;
    // instance lengths Ns^2 (Ns >= 32) and chunks of kChunk elements:
    // every boundary is a multiple of 4
    ev.vec4 = kChunk % 4 == 0 &&
              std::all_of(len.begin(), len.end(), [](std::int64_t l) { return l % 4 == 0; });
    Ens epc = ev;   // preconditioner partials
    epc.rows = dprow.p;

    auto matvec = [&](const Real* x, Real* y) {
        for (auto& g : groups) {
            SpecBatch sb;
            sb.count = g.count;
            sb.x_stride = (std::int64_t)g.Ns * g.Ns;
            sb.y_stride = sb.x_stride;
            sb.first = g.first;
            sb.sstride = S_COUNT;
            sb.pstride = pstride;
            sb.gscal = scal.p;
            sb.gate = S_ACTIVE;
            const Symbol sym = g.kind == 0 ? Symbol::table : Symbol::periodic;
            g.plan->convolve(x + off[g.first], g.Ns, y + off[g.first], g.Ns, sym, g.ktab,
                             g.kpitch, g.sym_scale, s, RowIn::plain, RowOut::plain, nullptr,
                             nullptr, 0, nullptr, nullptr, &sb);
            fault_after_apply<Real>(y + off[g.first], s);
        }
    };
    auto precondition = [&]() {
        for (auto& g : groups) {
            if (g.pc->kind() == Precond::Kind::stencil) {
                Ens et = ev;
                et.work = g.tiles.p;
                et.nwork = g.ntiles * g.count;
                V::stencil(g.Ns, G, bmask.p, scal.p, g.pc->taps_dev(), g.pc->weights_dev(),
                           (int)g.pc->tap_w().size(), g.pc->radius(), Z, part.p, 0, s, et);
            } else if (g.pc->kind() == Precond::Kind::fourier) {
                SpecBatch sb;
                sb.count = g.count;
                sb.x_stride = (std::int64_t)g.Ns * g.Ns;
                sb.y_stride = sb.x_stride;
                sb.first = g.first;
                sb.sstride = S_COUNT;
                sb.pstride = pstride;
                sb.gscal = scal.p;
                sb.gate = S_ACTIVE;
                const std::int64_t o = off[g.first];
                g.pc_plan->convolve(G + o, g.Ns, Z + o, g.Ns, Symbol::precond, nullptr, 0,
                                    (double)g.pc->fourier_norm(), s, RowIn::masked,
                                    RowOut::masked, bmask.p + o, scal.p, S_GMEAN, part.p,
                                    nullptr, &sb);
            }
        }
        static const int dz[1] = {S_ZSUM};
        finalize<Real>(part.p, 0, 1, kS1, dz, scal.p, STEP_ZMEAN, s, epc);
    };
    auto iterate = [&]() {
        static const int o1[3] = {RSUM, RSUM, RMIN};
        static const int d1[3] = {S_GSUM, S_NC, S_VMIN};
        static const int d2[5] = {S_PV, S_GCS, S_E, S_GC, S_PSUM};
        static const int dg[3] = {S_G, S_GPR, S_GG};
        static const int dl1[2] = {S_RSUM, S_NUM};
        static const int dl2[1] = {S_DEN};
        static const int ou[2] = {RMAX, RSUM};
        static const int du[2] = {S_CHANGED, S_TOTAL};
        matvec(P, U);
        V::eval1(Ntot, U, bg0s.p, P, G, part.p, 0, s, ev);
        finalize<Real>(part.p, 0, 3, o1, d1, scal.p, STEP_EVAL1, s, ev);
        V::eval2(Ntot, P, G, bmask.p, scal.p, part.p, 0, s, ev);
        finalize<Real>(part.p, 0, 5, kS5, d2, scal.p, STEP_B_EVAL2, s, ev);
        if (has_pc) precondition();
        V::cgdot(Ntot, G, GP, Z, bmask.p, scal.p, has_pc, part.p, 0, s, ev);
        finalize<Real>(part.p, 0, 3, kS3, dg, scal.p, STEP_CGDOT, s, ev);
        V::tupdate(Ntot, G, GP, Z, bmask.p, T_, scal.p, has_pc, s, ev);
        matvec(T_, R);
        V::line1(Ntot, R, G, T_, bmask.p, part.p, 0, s, ev);
        finalize<Real>(part.p, 0, 2, kS2, dl1, scal.p, STEP_LINE1, s, ev);
        V::line2(Ntot, R, T_, bmask.p, scal.p, part.p, 0, s, ev);
        finalize<Real>(part.p, 0, 1, kS1, dl2, scal.p, STEP_LINE2, s, ev);
        // the best-iterate copy rides on the update (P is unchanged until then)
        V::update(Ntot, P, P, T_, G, scal.p, part.p, 0, s, ev,
                  opt.keep_best ? bbest.p : nullptr, S_SAVEBEST);
        finalize<Real>(part.p, 0, 2, ou, du, scal.p, STEP_UPDATE, s, ev);
        V::rescale(Ntot, P, scal.p, s, ev);
    };

    ASPHER_HIP(hipStreamSynchronize(s));
    st.setup_seconds = since(t_start);

    // ── the loop: eager first iteration, then graph replays ──
    const auto t_loop = Clock::now();
    std::vector<double> hs((std::size_t)ni * S_COUNT);
    std::vector<double> left(ni, -1.0);
    auto poll = [&]() -> int {
        ASPHER_HIP(hipMemcpyAsync(hs.data(), scal.p, sizeof(double) * hs.size(),
                                  hipMemcpyDeviceToHost, s));
        ASPHER_HIP(hipStreamSynchronize(s));
        ++st.polls;
        int active = 0;
        const double t = since(t_loop);
        for (int k = 0; k < ni; ++k) {
            if (hs[(std::size_t)k * S_COUNT + S_ACTIVE] != 0.0) ++active;
            else if (left[k] < 0.0) left[k] = t;
        }
        return active;
    };
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    auto cleanup = [&]() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    };
    try {
        iterate();
        ++st.replays;
        int active = poll();
        if (active > 0 && opt.use_graph) {
            ASPHER_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                iterate();
            } catch (...) {
                hipGraph_t g2 = nullptr;
                (void)hipStreamEndCapture(s, &g2);
                if (g2) (void)hipGraphDestroy(g2);
                throw;
            }
            ASPHER_HIP(hipStreamEndCapture(s, &graph));
            ASPHER_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            st.graph = true;
        }
        // every instance leaves by convergence, a host reason or its budget
        // (S_ITER >= S_MAXIT): the replay count is bounded
        const long long max_replays = (long long)opt.max_iter + 2;
        while (active > 0 && st.replays < max_replays) {
            for (int r = 0; r < opt.check_every && st.replays < max_replays; ++r) {
                if (gexec) ASPHER_HIP(hipGraphLaunch(gexec, s));
                else iterate();
                ++st.replays;
            }
            active = poll();
        }
        if (active > 0) throw std::runtime_error("solve_ensemble: iteration budget overrun");
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
    st.loop_seconds = since(t_loop);
    st.instance_seconds.assign(ni, 0.0);

    // ── results ──
    const auto t_fin = Clock::now();
    out.clear();
    out.resize(ni);
    std::unique_ptr<Context> fctx;
    for (int k = 0; k < ni; ++k) {
        const int oi = order[k];
        const auto& e = inst[oi];
        const double* h = hs.data() + (std::size_t)k * S_COUNT;
        Result& res = out[oi];
        st.instance_seconds[oi] = left[k];
        const int iters = (int)h[S_ITER];
        st.device_iterations = std::max(st.device_iterations, iters);
        const std::int64_t N = len[k];
        Real* Pk = P + off[k];
        if (h[S_STATUS] == 1.0) {
            // converged on the device: U and G are the fresh evaluation
            DevBuf<double> part1((std::size_t)vec_blocks(N) * kPartVals);
            static const int d3[3] = {S_R0, S_R1, S_R2};
            const int nb = vec_blocks(N);
            V::objective(N, Pk, U + off[k], bg0s.p + off[k], part1.p, nb, s);
            finalize<Real>(part1.p, nb, 3, kS3, d3, scal.p + (std::size_t)k * S_COUNT,
                           STEP_NONE, s);
            res.pressure.alloc(N);
            V::to_double(N, Pk, res.pressure.p, s);
            if (!opt.light) {
                res.displacement.alloc(N);
                res.gap.alloc(N);
                V::to_double(N, U + off[k], res.displacement.p, s);
                V::to_double(N, G + off[k], res.gap.p, s);
            }
            double hf[S_COUNT];
            ASPHER_HIP(hipMemcpyAsync(hf, scal.p + (std::size_t)k * S_COUNT, sizeof(hf),
                                      hipMemcpyDeviceToHost, s));
            ASPHER_HIP(hipStreamSynchronize(s));
            res.Ns = e.Ns;
            res.objective = hf[S_R0] + datum[k] * Ptot[k];
            res.approach = h[S_ALPHA_REF] + datum[k];
            res.iterations = iters;
            res.contact_fraction = h[S_NC] / (double)N;
            res.mean_pressure = h[S_PSUM] / (double)N;
            res.status = Status::converged;
            res.converged = true;
            res.error = res.pk_error = h[S_PK];
            res.fw_gap = h[S_GC];
            res.fw_error = h[S_FW];
            res.load_error = h[S_LOAD];
            res.penetration_error = h[S_PEN];
            res.matvec_count = 2LL * iters + 1;
            res.precond_count = has_pc ? iters : 0;
        } else {
            // host decision needed: finish on the single-solve path from the
            // device state (best iterate when it is better)
            ++st.host_fallbacks;
            if (!fctx) fctx = std::make_unique<Context>();
            ASPHER_HIP(hipStreamSynchronize(s));
            auto op = cached_operator(e.kind == 0 ? Operator::Kind::free_space
                                                  : Operator::Kind::periodic,
                                      e.Ns, e.L, e.E_star, fctx->stream);
            const Precond::Kind pk = e.kind == 0 ? pk_free : pk_per;
            Precond pc1(pk, e.Ns, std::min({2, std::max(1, e.Ns / 4), 8}));
            SolveOptions o;
            o.tol = eff_tol;
            o.requested_tol = opt.tol;
            o.allow_tolerance_relaxation = true;
            o.max_iter = std::max(0, opt.max_iter - iters);
            o.use_pr = opt.use_pr;
            o.keep_best = opt.keep_best;
            o.stall_limit = opt.stall_limit;
            o.light = opt.light;
            o.g_ref = gref[k];
            o.p_ref = e.p_bar;
            o.datum_mode = SolveOptions::Datum::caller;
            o.datum = datum[k];
            const bool use_best = opt.keep_best && h[S_BEST] < 1e300 &&
                                  h[S_BEST] < std::max(h[S_FW], h[S_PEN]);
            Real* pinit = use_best ? bbest.p + off[k] : Pk;
            solve_impl<Real>(*fctx, *op, bg0s.p + off[k], static_cast<Real>(e.p_bar), o,
                             &pc1, pinit, res);
            res.iterations += iters;
            res.matvec_count += 2LL * iters + 1;
            if (has_pc) res.precond_count += iters;
            res.status_reason = (h[S_HOST] == HOST_IDENT   ? "ensemble:identification"
                                 : h[S_HOST] == HOST_STALL ? "ensemble:stall"
                                 : h[S_HOST] == HOST_MAXIT ? "ensemble:max_iterations"
                                                           : "ensemble:nonfinite") +
                                (res.status_reason.empty() ? std::string()
                                                           : "," + res.status_reason);
        }
        res.requested_tol = opt.tol;
        res.effective_tol = eff_tol;
        res.p_ref = e.p_bar;
        res.g_ref = gref[k];
        res.validation_scope = "solve_operator";
        Stage sg;
        sg.name = "ensemble:" + std::to_string(e.Ns);
        sg.precision = is_double ? "double" : "float";
        sg.requested_tol = opt.tol;
        sg.effective_tol = eff_tol;
        sg.iterations = res.iterations;
        sg.matvec_count = res.matvec_count;
        sg.seconds = left[k];
        sg.status = res.status;
        sg.fw_error = res.fw_error;
        sg.penetration_error = res.penetration_error;
        res.stage_stats.push_back(sg);
        res.time_total = left[k];
    }
    ASPHER_HIP(hipStreamSynchronize(s));
    st.finish_seconds = since(t_fin);
    st.seconds = since(t_start);
}

} // namespace

void solve_ensemble(const std::vector<EnsembleInstance>& inst, const EnsembleOptions& opt,
                    std::vector<Result>& out, EnsembleStats& stats) {
    if (opt.precision == "float") ensemble_impl<float>(inst, opt, out, stats);
    else if (opt.precision == "double") ensemble_impl<double>(inst, opt, out, stats);
    else throw std::invalid_argument("solve_ensemble: precision must be 'float' or 'double'");
}

} // namespace aspher
