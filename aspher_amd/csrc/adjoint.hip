// Adjoint of the certified contact solve: the backward pass of p(g0, P).
//
// No counterpart in the reference. The reference solves
//   min 1/2 p'Sp + g0'p  s.t.  p >= 0, sum p = P
// (KKT residuals as in tests/contact_oracle.hpp:35-61). On a strictly
// complementary solution with support C the optimality system is linear:
//   S_CC p_C - alpha 1 = -g0_C,  1'p_C = P,
// so a loss L with dL/dp = w has dL/dg0 = -mu on C (zero off C) and
// dL/dP = nu, where [S_CC 1; 1' 0][mu; nu] = [w_C; 0]. mu is the solution of
// S_CC mu = w_C - nu 1 with 1'mu = 0: a symmetric positive definite system on
// the zero-mean subspace of C, solved here by preconditioned CG with the
// forward operator (the same cached kernel spectrum) and the forward
// preconditioners, which already centre over the support. Iterates are
// recomputed, nothing of the forward CG is kept. After convergence
// nu = mean_C (w - S mu).
#include "engine.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

#include <chrono>
#include <cmath>

namespace aspher {

void adjoint_impl(Context& ctx, Operator& op, Precond* pc, const double* p, const double* w,
                  double* grad, const AdjointOptions& opt, AdjointInfo& info) {
    if (!(opt.tol > 0.0) || opt.max_iter < 1)
        throw std::invalid_argument("adjoint: tol must be > 0 and max_iter >= 1");
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = ctx.stream;
    const int Ns = op.Ns();
    const std::int64_t N = (std::int64_t)Ns * Ns;
    const int nblk = vec_blocks(N);
    double* scal = ctx.scal.p;
    double* part = ctx.part.p;
    const bool has_pc = pc && pc->kind() != Precond::Kind::none;
    using A = AdjOps<double>;

    DevBuf<double> x(N), r(N), t(N), q(N), z(has_pc ? N : 0);
    DevBuf<std::uint8_t> mask(N);
    static const int s2[2] = {RSUM, RSUM};
    static const int s3[3] = {RSUM, RSUM, RSUM};
    static const int d_init[2] = {S_A_WSUM, S_NC};
    static const int d_dot[2] = {S_A_RZ, S_A_RR};
    static const int d_curv[3] = {S_A_QSUM, S_A_QT, S_A_TSUM};
    static const int d_rsum[1] = {S_A_RSUM};
    static const int d_fin[1] = {S_A_WQSUM};

    info = AdjointInfo{};
    A::init(N, p, w, mask.p, r.p, x.p, t.p, part, nblk, s);
    finalize<double>(part, nblk, 2, s2, d_init, scal, STEP_ADJ_INIT, s);
    // centre the right-hand side over C, then record its mean (~0) for the
    // preconditioner's centring
    VecOps<double>::masked_sub(N, r.p, mask.p, scal, S_GMEAN, s);
    VecOps<double>::masked_sum(N, r.p, mask.p, part, nblk, s);
    finalize<double>(part, nblk, 1, s2, d_rsum, scal, STEP_ADJ_RMEAN, s);

    double rr0 = 0.0;
    info.status = "max_iterations";
    for (int it = 0;; ++it) {
        if (has_pc) {
            pc->apply<double>(r.p, mask.p, z.p, scal, part, s);
            ++info.precond_count;
        }
        A::dot(N, r.p, has_pc ? z.p : nullptr, mask.p, scal, part, nblk, s);
        finalize<double>(part, nblk, 2, s2, d_dot, scal, STEP_ADJ_DOT, s);
        ctx.read_scal();
        const double* h = ctx.host_scal;
        if (it == 0) info.n_contact = (long long)h[S_NC];
        if (it > 0 && h[S_A_OK] == 0.0) {
            info.status = "nonpositive_curvature";
            break;
        }
        const double rr = h[S_A_RR];
        if (!std::isfinite(rr) || !std::isfinite(h[S_A_RZ])) {
            info.status = "nonfinite";
            break;
        }
        if (it == 0) rr0 = rr;
        info.rel_residual = rr0 > 0.0 ? std::sqrt(rr / rr0) : 0.0;
        if (rr0 == 0.0 || info.rel_residual <= opt.tol) {
            info.converged = true;
            info.status = "converged";
            break;
        }
        if (it >= opt.max_iter) break;
        A::dir(N, r.p, has_pc ? z.p : nullptr, mask.p, scal, t.p, s);
        op.apply(t.p, q.p, s);
        ++info.matvec_count;
        A::curv(N, q.p, t.p, mask.p, part, nblk, s);
        finalize<double>(part, nblk, 3, s3, d_curv, scal, STEP_ADJ_CURV, s);
        A::step(N, x.p, r.p, t.p, q.p, mask.p, scal, part, nblk, s);
        finalize<double>(part, nblk, 1, s2, d_rsum, scal, STEP_ADJ_RMEAN, s);
        info.iterations = it + 1;
    }
    // nu = mean_C (w - S mu), gradient = -mu on C
    op.apply(x.p, q.p, s);
    ++info.matvec_count;
    A::finish(N, x.p, w, q.p, mask.p, grad, part, nblk, s);
    finalize<double>(part, nblk, 1, s2, d_fin, scal, STEP_ADJ_FINISH, s);
    ctx.read_scal();
    info.multiplier = ctx.host_scal[S_A_NU];
    info.grad_p_nominal = info.multiplier * (double)N;
    info.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

} // namespace aspher
