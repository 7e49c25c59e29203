// Active-set (restricted) level of the nested driver for the free-space
// backend: src/nested_solve.cpp:117-516 with src/contact_solver.cpp:601-1040
// as the restricted solve.
//
// The reference compresses the state to the candidate set C because its
// masked H2 matvec gets cheaper with |C|. The FFT operator costs the same on
// any support, so here the state stays on the full grid and the restriction
// is carried by the gap: outside C the working gap is lifted to
// 1e8 * g_ref, far above any value a candidate can take. Those points can
// then never hold pressure (they start at zero and are never in contact, the
// overlap step needs g < 0), never set the minimum gap, and never survive the
// threshold of the load projection, so the standard device PK iteration
// performs exactly the restricted iteration of the reference. The global
// certificate comes from a fresh apply of the returned pressure over the
// whole grid: G = G_local + P (vmin_local - vmin_global).
#include "engine.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <type_traits>

namespace aspher {

namespace {

using Clock = std::chrono::steady_clock;
inline double since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

#define GRID_LOOP(i, N)                                                          \
    for (std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;   \
         i < (N); i += (std::int64_t)gridDim.x * blockDim.x)

template <int OP>
__device__ __forceinline__ void put(double v, double* red, double* part, int idx) {
    v = block_reduce<OP>(v, red);
    if (threadIdx.x == 0) part[(std::size_t)blockIdx.x * kPartVals + idx] = v;
}

__global__ void __launch_bounds__(256)
k_cand_pos(std::int64_t N, const double* __restrict__ p, std::uint8_t* __restrict__ m) {
    GRID_LOOP(i, N) m[i] = p[i] > 0.0 ? 1 : 0;
}

// square (Chebyshev) dilation, window clipped at the grid edge
// (src/nested_solve.cpp:76-101): rows, then columns
__global__ void __launch_bounds__(256)
k_dilate_rows(int Ns, int r, const std::uint8_t* __restrict__ in, std::uint8_t* __restrict__ out) {
    const std::int64_t N = (std::int64_t)Ns * Ns;
    GRID_LOOP(i, N) {
        const int iy = (int)(i / Ns), ix = (int)(i - (std::int64_t)iy * Ns);
        const std::uint8_t* row = in + (std::size_t)iy * Ns;
        const int x1 = min(Ns - 1, ix + r);
        std::uint8_t v = 0;
        for (int x = max(0, ix - r); x <= x1 && !v; ++x) v = row[x];
        out[i] = v;
    }
}

__global__ void __launch_bounds__(256)
k_dilate_cols(int Ns, int r, const std::uint8_t* __restrict__ in, std::uint8_t* __restrict__ out) {
    const std::int64_t N = (std::int64_t)Ns * Ns;
    GRID_LOOP(i, N) {
        const int iy = (int)(i / Ns), ix = (int)(i - (std::int64_t)iy * Ns);
        const int y1 = min(Ns - 1, iy + r);
        std::uint8_t v = 0;
        for (int y = max(0, iy - r); y <= y1 && !v; ++y) v = in[(std::size_t)y * Ns + ix];
        out[i] = v;
    }
}

// m |= coarse gap (sampled by injection from the (Ns/2)^2 level below) < delta
__global__ void __launch_bounds__(256)
k_cand_coarse(int Ns, const double* __restrict__ cg, double delta, std::uint8_t* __restrict__ m) {
    const std::int64_t N = (std::int64_t)Ns * Ns;
    const int nc = Ns / 2;
    GRID_LOOP(i, N) {
        const int iy = (int)(i / Ns), ix = (int)(i - (std::int64_t)iy * Ns);
        if (cg[(std::size_t)(iy >> 1) * nc + (ix >> 1)] < delta) m[i] = 1;
    }
}

__global__ void __launch_bounds__(256)
k_or_mask(std::int64_t N, const std::uint8_t* __restrict__ a, std::uint8_t* __restrict__ m) {
    GRID_LOOP(i, N) if (a[i]) m[i] = 1;
}

// working gap: T(g - datum) on C, `big` elsewhere
template <typename T>
__global__ void __launch_bounds__(256)
k_lift_gap(std::int64_t N, const double* __restrict__ g, double datum,
           const std::uint8_t* __restrict__ m, T big, T* __restrict__ out) {
    GRID_LOOP(i, N) out[i] = m[i] ? static_cast<T>(g[i] - datum) : big;
}

// warm start restricted to C: out = m ? T(p) : 0; partials {sum_C max(T(p),0), |C|}
template <typename T>
__global__ void __launch_bounds__(256)
k_cand_warm(std::int64_t N, const double* __restrict__ p, const std::uint8_t* __restrict__ m,
            T* __restrict__ out, double* __restrict__ part) {
    __shared__ double red[8];
    double s = 0.0, c = 0.0;
    GRID_LOOP(i, N) {
        T v = T(0);
        if (m[i]) {
            v = static_cast<T>(p[i]);
            const T cl = v < T(0) ? T(0) : v;
            s += (double)cl;
            c += 1.0;
        }
        out[i] = v;
    }
    put<RSUM>(s, red, part, 0);
    put<RSUM>(c, red, part, 1);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_fill_masked(std::int64_t N, const std::uint8_t* __restrict__ m, T v, T* __restrict__ out) {
    GRID_LOOP(i, N) out[i] = m[i] ? v : T(0);
}

// global verification over every grid point (src/nested_solve.cpp:255-310):
// v = u + (g0 - datum) in the level's working frame; partials
// {min v, #violations}; viol = !C && v - a < vthresh
template <typename T>
__global__ void __launch_bounds__(256)
k_active_verify(std::int64_t N, const T* __restrict__ u, const double* __restrict__ g,
                double datum, const std::uint8_t* __restrict__ m, double a, double vthresh,
                std::uint8_t* __restrict__ viol, double* __restrict__ part) {
    __shared__ double red[8];
    double vmin = __builtin_huge_val(), nv = 0.0;
    GRID_LOOP(i, N) {
        const double v = (double)u[i] + (double)static_cast<T>(g[i] - datum);
        if (v < vmin) vmin = v;
        const bool bad = !m[i] && (v - a < vthresh);
        viol[i] = bad ? 1 : 0;
        if (bad) nv += 1.0;
    }
    put<RMIN>(vmin, red, part, 0);
    put<RSUM>(nv, red, part, 1);
}

// result fields of a certified restricted level: displacement and gap
template <typename T>
__global__ void __launch_bounds__(256)
k_active_fields(std::int64_t N, const T* __restrict__ u, const double* __restrict__ g,
                double datum, double a, double* __restrict__ disp, double* __restrict__ gap) {
    GRID_LOOP(i, N) {
        const double d = (double)u[i];
        disp[i] = d;
        gap[i] = d + (double)static_cast<T>(g[i] - datum) - a;
    }
}

#define LV(k, ...) k<<<nblk, kVecThreads, 0, s>>>(__VA_ARGS__); ASPHER_LAUNCH_CHECK()

void dilate(int Ns, int r, std::uint8_t* m, std::uint8_t* tmp, hipStream_t s) {
    if (r <= 0) return;
    const int nblk = vec_blocks((std::int64_t)Ns * Ns);
    LV(k_dilate_rows, Ns, r, m, tmp);
    LV(k_dilate_cols, Ns, r, tmp, m);
}

} // namespace

template <typename Real>
void active_level(Context& ctx, Operator& op, Precond* pc, const double* glvl, double datum,
                  double p_bar, double lvl_tol, int max_iter, bool use_pr,
                  const NestedParams& np, DevBuf<double>& p_init,
                  const DevBuf<double>& coarse_gap, bool record_history, bool light,
                  double g_ref, Result& res) {
    constexpr bool is_double = std::is_same_v<Real, double>;
    hipStream_t s = ctx.stream;
    const int Ns = op.Ns();
    const std::int64_t N = (std::int64_t)Ns * Ns;
    const int nblk = vec_blocks(N);
    double* scal = ctx.scal.p;
    double* part = ctx.part.p;
    const double* hs = ctx.host_scal;
    if (p_init.n != (std::size_t)N || coarse_gap.n != (std::size_t)N / 4)
        throw std::logic_error("active_level: warm start / coarse gap size mismatch");

    // the level's gap scale (src/nested_solve.cpp:137-143)
    {
        static const int ops[3] = {RMIN, RMAX, RSUM};
        static const int dst[3] = {S_R0, S_R1, S_R2};
        VecOps<double>::stats(N, glvl, part, nblk, s);
        finalize<double>(part, nblk, 3, ops, dst, scal, STEP_NONE, s);
        ctx.read_scal();
    }
    const double gmin = hs[S_R0], gmax = hs[S_R1];
    const double g_scale = gmax > gmin ? gmax - gmin : 1.0;
    const double delta = np.active_delta * g_scale;
    const double vthresh = -lvl_tol * g_scale;
    const double P_total = p_bar * (double)N;
    const Real big = static_cast<Real>(1e8 * g_ref);

    // candidate set C = dilate(p_init > 0) | {coarse gap < delta}
    double t_candidate = 0.0;
    auto t_c0 = Clock::now();
    DevBuf<std::uint8_t> cmask(N), tmp(N), viol(N);
    DevBuf<Real> g0r(N), p0(N), u(N), pw(N);
    LV(k_cand_pos, N, p_init.p, cmask.p);
    dilate(Ns, np.active_halo, cmask.p, tmp.p, s);
    LV(k_cand_coarse, Ns, coarse_gap.p, delta, cmask.p);
    LV(k_lift_gap<Real>, N, glvl, datum, cmask.p, big, g0r.p);
    {
        // warm start on C (src/contact_solver.cpp:654-677): uniform P/|C|
        // when the prolonged pressure carries no load there
        static const int d2[2] = {S_R0, S_R1};
        static const int o2[2] = {RSUM, RSUM};
        LV(k_cand_warm<Real>, N, p_init.p, cmask.p, p0.p, part);
        finalize<Real>(part, nblk, 2, o2, d2, scal, STEP_NONE, s);
        ctx.read_scal();
        const double ssum = hs[S_R0], nc = hs[S_R1];
        if (!(nc > 0.0))
            throw std::invalid_argument(
                "solve_contact_active: empty gap/candidate set or p_bar <= 0");
        if (!(ssum > 0.0))
            LV(k_fill_masked<Real>, N, cmask.p, static_cast<Real>(P_total / nc), p0.p);
    }
    p_init.reset();
    t_candidate += since(t_c0);

    Result r;
    std::vector<double> hist;
    int it_total = 0, rounds = 0;
    long long verify_mv = 0;
    double t_verify = 0.0, t_output = 0.0;
    bool certified = false, p0_valid = true;
    double a = 0.0;

    while (rounds < np.active_max_rounds) {
        ++rounds;
        // restricted stage at half the level budget: with every outside gap
        // >= -tol/2 g_ref the global certificate gains at most tol/2
        SolveOptions aopt;
        aopt.tol = 0.5 * lvl_tol;
        aopt.requested_tol = lvl_tol;
        aopt.max_iter = max_iter;
        aopt.use_pr = use_pr;
        aopt.record_history = record_history;
        aopt.allow_tolerance_relaxation = true;
        aopt.light = true;
        aopt.keep_best = true;
        aopt.g_ref = g_ref;
        aopt.datum_mode = SolveOptions::Datum::caller;
        aopt.datum = datum;
        aopt.precond_cost_gate = np.precond_cost_gate;
        aopt.graph_max_points = np.graph_max_points;
        r = Result();
        solve_impl<Real>(ctx, op, g0r.p, static_cast<Real>(p_bar), aopt, pc, p0.p, r);
        p0_valid = false;
        it_total += r.iterations;
        if (record_history)
            hist.insert(hist.end(), r.error_history.begin(), r.error_history.end());

        // verification: fresh apply of the returned pressure in working
        // precision, global minimum and violations outside C
        const auto t_v0 = Clock::now();
        VecOps<Real>::from_double(N, r.pressure.p, 0.0, pw.p, s);
        op.apply(pw.p, u.p, s);
        ++verify_mv;
        a = r.approach - datum;
        {
            static const int o2[2] = {RMIN, RSUM};
            static const int d2[2] = {S_R0, S_R1};
            LV(k_active_verify<Real>, N, u.p, glvl, datum, cmask.p, a, vthresh, viol.p, part);
            finalize<Real>(part, nblk, 2, o2, d2, scal, STEP_NONE, s);
            ctx.read_scal();
        }
        t_verify += since(t_v0);
        const double vmin_glob = hs[S_R0];
        const long long nviol = (long long)hs[S_R1];
        const double G_glob = r.fw_gap + P_total * (r.vmin_local - vmin_glob);
        const double fw_glob = G_glob / (P_total * g_ref);
        const double pen_glob = std::max(0.0, -(vmin_glob - a)) / g_ref;
        r.fw_gap = G_glob;
        r.fw_error = fw_glob;
        r.penetration_error = pen_glob;
        r.effective_tol = lvl_tol;
        if (r.status == Status::converged && nviol == 0 && fw_glob <= lvl_tol &&
            pen_glob <= lvl_tol) {
            certified = true;
            break;
        }
        if (nviol == 0) break;   // nothing to extend into: fall back
        // extend C with the dilated violations and resume from the returned
        // pressure (zero on the new candidates)
        t_c0 = Clock::now();
        dilate(Ns, np.active_halo, viol.p, tmp.p, s);
        LV(k_or_mask, N, viol.p, cmask.p);
        LV(k_lift_gap<Real>, N, glvl, datum, cmask.p, big, g0r.p);
        VecOps<Real>::copy(N, pw.p, p0.p, s);
        p0_valid = true;
        t_candidate += since(t_c0);
    }

    if (!certified) {
        // safety net: full-grid solve warm-started from the last restricted
        // pressure (src/nested_solve.cpp:367-452)
        if (!p0_valid) VecOps<Real>::from_double(N, r.pressure.p, 0.0, p0.p, s);
        SolveOptions fopt;
        fopt.tol = lvl_tol;
        fopt.max_iter = max_iter;
        fopt.use_pr = use_pr;
        fopt.light = light;
        fopt.record_history = record_history;
        fopt.keep_best = !light;
        fopt.allow_tolerance_relaxation = true;
        fopt.g_ref = g_ref;
        fopt.datum = datum;
        fopt.precond_cost_gate = np.precond_cost_gate;
        fopt.graph_max_points = np.graph_max_points;
        Result f;
        if constexpr (is_double) {
            fopt.datum_mode = SolveOptions::Datum::solver;
            solve_impl<double>(ctx, op, glvl, p_bar, fopt, pc, p0.p, f);
        } else {
            fopt.datum_mode = SolveOptions::Datum::caller;
            VecOps<Real>::from_double(N, glvl, datum, g0r.p, s);
            solve_impl<Real>(ctx, op, g0r.p, static_cast<Real>(p_bar), fopt, pc, p0.p, f);
        }
        it_total += f.iterations;
        if (record_history)
            hist.insert(hist.end(), f.error_history.begin(), f.error_history.end());
        r = std::move(f);
        r.active_fallback = true;
    } else if (!light) {
        const auto t_o0 = Clock::now();
        r.displacement.alloc(N);
        r.gap.alloc(N);
        LV(k_active_fields<Real>, N, u.p, glvl, datum, a, r.displacement.p, r.gap.p);
        ASPHER_HIP(hipStreamSynchronize(s));
        t_output = since(t_o0);
    }
    r.time_verification = t_verify;
    r.time_output = t_output;
    r.time_candidate = t_candidate;
    r.active_rounds = rounds;
    r.iterations = it_total;
    r.verification_matvec_count = verify_mv;
    if (!certified && !r.active_fallback && r.status == Status::converged) {
        r.status = Status::verification_failed;
        r.status_reason = "global_check";
    }
    r.converged = r.status == Status::converged;
    if (record_history) r.error_history = std::move(hist);
    res = std::move(r);
}

template void active_level<double>(Context&, Operator&, Precond*, const double*, double, double,
                                   double, int, bool, const NestedParams&, DevBuf<double>&,
                                   const DevBuf<double>&, bool, bool, double, Result&);
template void active_level<float>(Context&, Operator&, Precond*, const double*, double, double,
                                  double, int, bool, const NestedParams&, DevBuf<double>&,
                                  const DevBuf<double>&, bool, bool, double, Result&);

} // namespace aspher
