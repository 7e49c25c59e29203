// Tangential (Cerruti) 2x2 FFT operator and its Fourier preconditioner:
// src/tangential_operator.cpp:16-157 and src/fourier_precond.cpp:234-359.
//
// u = [Kxx Kxy; Kxy Kyy] q on stacked [x; y] components. The three offset
// tables are embedded in wrap-around order on the (2Ns)^2 grid and
// transformed once in FP64; xx and yy are even, xy odd in each axis, so all
// three spectra are real (checked like the normal operator's). A matvec is
// the two-component pipeline of the register-blocked plan: one batched row
// pass over both components, one column pass that transforms both
// components of a column, mixes them per mode and transforms back, one
// batched inverse row pass.
#include "engine.hpp"
#include "kernels.hpp"
#include "reduce.hpp"
#include "spectral_rb.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <type_traits>

namespace aspher {

TangentialOperator::TangentialOperator(int Ns, double L, double E_star, double nu)
    : Ns_(Ns), L_(L), E_(E_star), nu_(nu) {
    if (Ns < 16 || Ns > 2048 || (Ns & (Ns - 1)))
        throw std::invalid_argument("tangential operator: grid_size must be a power of two in [16, 2048]");
    if (!(L > 0.0)) throw std::invalid_argument("domain_size must be positive");
    if (!(E_star > 0.0)) throw std::invalid_argument("E_star must be positive");
    if (!(nu >= 0.0 && nu < 1.0)) throw std::invalid_argument("nu must be in [0, 1)");
}

TangentialOperator::~TangentialOperator() = default;

void TangentialOperator::build(hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    const std::int64_t N = (std::int64_t)Ns_ * Ns_;
    const int M = 2 * Ns_;
    const std::size_t nu2 = (std::size_t)(Ns_ + 1) * (Ns_ + 1);
    DevBuf<double> xx(N), xy(N);
    cerruti_tables(Ns_, L_, E_, nu_, xx.p, xy.p, s);
    {
        SpectralPlan<double> bplan(M, M, M, 0, 0);
        DevBuf<double> part((std::size_t)2 * bplan.col_blocks());
        const char* names[3] = {"xx", "yy", "xy"};
        const double* src[3] = {xx.p, xx.p, xy.p};
        const int embed[3] = {2, 3, 4};
        for (int k = 0; k < 3; ++k) {
            k64_[k].alloc(nu2);
            bplan.build_spectrum(src[k], Ns_, k64_[k].p, part.p, s, embed[k]);
            std::vector<double> h(part.n);
            ASPHER_HIP(hipMemcpyAsync(h.data(), part.p, part.bytes(), hipMemcpyDeviceToHost, s));
            ASPHER_HIP(hipStreamSynchronize(s));
            double mi = 0.0, mr = 0.0;
            for (std::size_t b = 0; b < h.size() / 2; ++b) {
                mi = std::max(mi, h[2 * b]);
                mr = std::max(mr, h[2 * b + 1]);
            }
            if (mi > 1e-10 * mr)
                throw std::runtime_error(std::string("TangentialFFTOperator: ") + names[k] +
                                         " spectrum parity broken (wrap-around embedding)");
        }
    }
    plan64_ = std::make_unique<SpectralPlan<double>>(M, Ns_, Ns_, Ns_, Ns_, SpectralTuning{}, 2);
    for (int k = 0; k < 3; ++k) {
        perm64_[k].alloc(plan64_->table_elems());
        plan64_->permute_table(k64_[k].p, Ns_ + 1, perm64_[k].p, s, k == 2);
    }
    ASPHER_HIP(hipStreamSynchronize(s));
    built_ = true;
    build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void TangentialOperator::build_single(hipStream_t s) {
    if (!built_) throw std::logic_error("TangentialFFTOperator: build() not called");
    if (plan32_) return;
    const int M = 2 * Ns_;
    const std::size_t nu2 = (std::size_t)(Ns_ + 1) * (Ns_ + 1);
    plan32_ = std::make_unique<SpectralPlan<float>>(M, Ns_, Ns_, Ns_, Ns_, SpectralTuning{}, 2);
    DevBuf<float> tmp(nu2);
    for (int k = 0; k < 3; ++k) {
        cast_d2f((std::int64_t)nu2, k64_[k].p, tmp.p, s);
        perm32_[k].alloc(plan32_->table_elems());
        plan32_->permute_table(tmp.p, Ns_ + 1, perm32_[k].p, s, k == 2);
    }
    ASPHER_HIP(hipStreamSynchronize(s));
}

template <typename T>
void TangentialOperator::apply_t(const T* q, T* u, hipStream_t s) {
    if (!built_) throw std::logic_error("TangentialFFTOperator: build() not called");
    const std::int64_t N = (std::int64_t)Ns_ * Ns_;
    if constexpr (std::is_same_v<T, double>) {
        plan64_->convolve2(q, Ns_, N, u, Ns_, N, 5, perm64_[0].p, perm64_[1].p, perm64_[2].p,
                           2 * Ns_, 0.0, 0.0, s);
    } else {
        build_single(s);
        plan32_->convolve2(q, Ns_, N, u, Ns_, N, 5, perm32_[0].p, perm32_[1].p, perm32_[2].p,
                           2 * Ns_, 0.0, 0.0, s);
    }
}

void TangentialOperator::apply(const double* q, double* u, hipStream_t s) { apply_t(q, u, s); }
void TangentialOperator::apply(const float* q, float* u, hipStream_t s) { apply_t(q, u, s); }

// ── preconditioner ─────────────────────────────────────────────────────────

TangentialPrecond::TangentialPrecond(int Ns, double nu) : Ns_(Ns), nu_(nu) {
    if (Ns < 32 || Ns > 4096 || (Ns & (Ns - 1)))
        throw std::invalid_argument("tangential preconditioner: grid_size must be a power of two in [32, 4096]");
    if (!(nu >= 0.0 && nu < 1.0)) throw std::invalid_argument("nu must be in [0, 1)");
    scal_.alloc(S_COUNT);
}

template <typename T>
void TangentialPrecond::apply(const T* g, const std::uint8_t* mask, bool remove_mean, T* z,
                              hipStream_t s) {
    const std::int64_t N = (std::int64_t)Ns_ * Ns_;
    SpectralPlan<T>* pl;
    if constexpr (std::is_same_v<T, double>) {
        if (!plan64_) plan64_ = std::make_unique<SpectralPlan<double>>(Ns_, Ns_, Ns_, Ns_, Ns_,
                                                                       SpectralTuning{}, 2);
        pl = plan64_.get();
    } else {
        if (!plan32_) plan32_ = std::make_unique<SpectralPlan<float>>(Ns_, Ns_, Ns_, Ns_, Ns_,
                                                                      SpectralTuning{}, 2);
        pl = plan32_.get();
    }
    const int rows = pl->row_out_blocks();
    const std::size_t need =
        (std::size_t)std::max(2 * rows, vec_blocks(N)) * kPartVals;
    if (part_.n < need) part_.alloc(need);
    // the masked scatter subtracts scal[S_GMEAN]: zero for this operator
    VecOps<double>::fill(S_COUNT, scal_.p, 0.0, s);
    const double gamma = nu_ / (1.0 - nu_);
    const double norm = 1.0 / ((double)Ns_ * Ns_);
    pl->convolve2(g, Ns_, N, z, Ns_, N, 6, nullptr, nullptr, nullptr, 0, gamma, norm, s,
                  RowIn::masked, RowOut::masked, mask, scal_.p, S_GMEAN, part_.p,
                  (std::int64_t)rows * kPartVals);
    if (remove_mean) {
        // per-component mean over the mask (src/fourier_precond.cpp:341-357)
        static const int ops[2] = {RSUM, RSUM};
        static const int dst[2] = {S_ZSUM, S_NC};
        const int nb = vec_blocks(N);
        for (int c = 0; c < 2; ++c) {
            VecOps<T>::masked_sum(N, z + c * N, mask, part_.p, nb, s);
            finalize<T>(part_.p, nb, 2, ops, dst, scal_.p, STEP_ZMEAN, s);
            VecOps<T>::masked_sub(N, z + c * N, mask, scal_.p, S_ZMEAN, s);
        }
    }
}

template void TangentialPrecond::apply<double>(const double*, const std::uint8_t*, bool, double*,
                                               hipStream_t);
template void TangentialPrecond::apply<float>(const float*, const std::uint8_t*, bool, float*,
                                              hipStream_t);

} // namespace aspher
