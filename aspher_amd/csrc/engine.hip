// Host engine of the spectral contact path (see engine.hpp).
//
// The projected-CG state machine follows src/contact_solver.cpp:150-587 step
// by step; every O(N) loop is a device pass (kernels.hip), every operator or
// preconditioner application a spectral pipeline (spectral.hip). The host
// reads the FP64 diagnostic block once per iteration. The CG update of an
// iteration is launched without waiting for its line-search verdict: the
// verdict (positive curvature or not) is read together with the next
// iteration's diagnostics. When the verdict is negative the update kernels
// were predicated off on the device, the pressure is unchanged, and the
// speculatively re-evaluated state is bit-identical to the state the
// reference hands to its identification step, so the control flow is the
// reference's exactly.
#include "engine.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

namespace aspher {

const char* to_string(Status s) {
    switch (s) {
        case Status::converged: return "converged";
        case Status::stagnated: return "stagnated";
        case Status::max_iterations: return "max_iterations";
        case Status::nonpositive_curvature: return "nonpositive_curvature";
        case Status::nonfinite: return "nonfinite";
        case Status::verification_failed: return "verification_failed";
        case Status::resource_limit: return "resource_limit";
    }
    return "unknown";
}

namespace {

using Clock = std::chrono::steady_clock;
inline double since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

const int kSum1[1] = {RSUM};
const int kSum2[2] = {RSUM, RSUM};

} // namespace

// ── Context ────────────────────────────────────────────────────────────────

// ── device memory with reuse ───────────────────────────────────────────────

namespace {

struct DevPool {
    std::multimap<std::size_t, void*> blocks;   // rounded size -> cached block
    std::size_t cached = 0, cap = 0;
};

std::mutex& pool_mutex() {
    static std::mutex* m = new std::mutex;   // never destroyed: no exit-order issues
    return *m;
}
std::map<int, DevPool>& pools() {
    static auto* p = new std::map<int, DevPool>;
    return *p;
}
std::size_t pool_round(std::size_t bytes) {
    return (bytes + 4095) & ~static_cast<std::size_t>(4095);
}

void evict_all(DevPool& P) {
    for (auto& kv : P.blocks) (void)hipFree(kv.second);
    P.blocks.clear();
    P.cached = 0;
}

} // namespace

void* dev_alloc(std::size_t bytes) {
    if (bytes == 0) return nullptr;
    const std::size_t sz = pool_round(bytes);
    int dev = 0;
    ASPHER_HIP(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> lk(pool_mutex());
        DevPool& P = pools()[dev];
        auto it = P.blocks.find(sz);
        if (it != P.blocks.end()) {
            void* q = it->second;
            P.blocks.erase(it);
            P.cached -= sz;
            return q;
        }
    }
    void* q = nullptr;
    if (hipMalloc(&q, sz) != hipSuccess) {
        // out of memory: give the cached blocks back and try once more
        (void)hipGetLastError();
        (void)hipDeviceSynchronize();
        {
            std::lock_guard<std::mutex> lk(pool_mutex());
            evict_all(pools()[dev]);
        }
        ASPHER_HIP(hipMalloc(&q, sz));
    }
    return q;
}

void dev_free(void* q, std::size_t bytes) noexcept {
    if (!q) return;
    const std::size_t sz = pool_round(bytes);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(q);
        return;
    }
    std::lock_guard<std::mutex> lk(pool_mutex());
    DevPool& P = pools()[dev];
    if (P.cap == 0) {
        std::size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            tot = std::size_t(16) << 30;
        }
        P.cap = std::min<std::size_t>(tot / 8, std::size_t(24) << 30);
    }
    if (sz > P.cap) {
        (void)hipFree(q);
        return;
    }
    P.blocks.emplace(sz, q);
    P.cached += sz;
    while (P.cached > P.cap && !P.blocks.empty()) {
        auto it = std::prev(P.blocks.end());   // largest first
        (void)hipFree(it->second);
        P.cached -= it->first;
        P.blocks.erase(it);
    }
}

void dev_release_cached() noexcept {
    std::lock_guard<std::mutex> lk(pool_mutex());
    for (auto& kv : pools()) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        (void)hipSetDevice(kv.first);
        evict_all(kv.second);
        (void)hipSetDevice(cur);
    }
}

Context::Context() {
    ASPHER_HIP(hipStreamCreate(&stream));
    scal.alloc(S_COUNT);
    part.alloc((std::size_t)kMaxPartRows * kPartVals);
    ASPHER_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_scal), sizeof(double) * S_COUNT));
    ASPHER_HIP(hipMemsetAsync(scal.p, 0, sizeof(double) * S_COUNT, stream));
    ASPHER_HIP(hipEventCreateWithFlags(&scal_ready, hipEventDisableTiming));
}

Context::~Context() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (scal_ready) (void)hipEventDestroy(scal_ready);
    if (host_scal) (void)hipHostFree(host_scal);
    scal.reset();
    part.reset();
    if (stream) (void)hipStreamDestroy(stream);
}

void Context::read_scal() {
    ASPHER_HIP(hipMemcpyAsync(host_scal, scal.p, sizeof(double) * S_COUNT,
                              hipMemcpyDeviceToHost, stream));
    ASPHER_HIP(hipStreamSynchronize(stream));
}

void Context::copy_scal_async() {
    ASPHER_HIP(hipMemcpyAsync(host_scal, scal.p, sizeof(double) * S_COUNT,
                              hipMemcpyDeviceToHost, stream));
    ASPHER_HIP(hipEventRecord(scal_ready, stream));
}

void Context::wait_scal() { ASPHER_HIP(hipEventSynchronize(scal_ready)); }

// ── Operator ───────────────────────────────────────────────────────────────

Operator::Operator(Kind kind, int Ns, double L, double E_star)
    : kind_(kind), Ns_(Ns), L_(L), E_(E_star) {
    if (Ns <= 0) throw std::invalid_argument("grid_size must be positive");
    if (!(L > 0.0)) throw std::invalid_argument("domain_size must be positive");
    if (!(E_star > 0.0)) throw std::invalid_argument("E_star must be positive");
    if (kind == Kind::periodic && Ns < 2)
        throw std::invalid_argument("grid_size must be at least 2");
}

Operator::~Operator() = default;

void Operator::build(hipStream_t s) {
    const auto t0 = Clock::now();
    const std::int64_t N = (std::int64_t)Ns_ * Ns_;
    if (kind_ == Kind::free_space) {
        // FP64 Love table, then the forward transform of its (2Ns)^2
        // wrap-around embedding; K is real and even, so only the (Ns+1)^2
        // unique values Re/M^2 are kept (src/fft_operator.cpp:18-69).
        const int M = 2 * Ns_;
        DevBuf<double> table(N);
        love_table(Ns_, L_, E_, table.p, s);
        kh64_.alloc((std::size_t)(Ns_ + 1) * (Ns_ + 1));
        {
            SpectralPlan<double> bplan(M, M, M, 0, 0);
            DevBuf<double> part((std::size_t)2 * bplan.col_blocks());
            bplan.build_spectrum(table.p, Ns_, kh64_.p, part.p, s);
            std::vector<double> h(part.n);
            ASPHER_HIP(hipMemcpyAsync(h.data(), part.p, part.bytes(), hipMemcpyDeviceToHost,
                                      s));
            ASPHER_HIP(hipStreamSynchronize(s));
            double mi = 0.0, mr = 0.0;
            for (std::size_t b = 0; b < h.size() / 2; ++b) {
                mi = std::max(mi, h[2 * b]);
                mr = std::max(mr, h[2 * b + 1]);
            }
            if (mi > 1e-10 * mr)
                throw std::runtime_error(
                    "FFTOperator: kernel spectrum has a non-negligible imaginary part; "
                    "wrap-around embedding is broken");
        }
        plan64_ = std::make_unique<SpectralPlan<double>>(M, Ns_, Ns_, Ns_, Ns_);
    } else if (kind_ == Kind::periodic) {
        plan64_ = std::make_unique<SpectralPlan<double>>(Ns_, Ns_, Ns_, Ns_, Ns_);
    } else {
        table_.alloc(N);
        love_table(Ns_, L_, E_, table_.p, s);
    }
    ASPHER_HIP(hipStreamSynchronize(s));
    built_ = true;
    build_seconds = since(t0);
}

void Operator::build_single(hipStream_t s) {
    if (!built_) throw std::logic_error("FFTOperator::build_single_caches: build() not called");
    if (kind_ == Kind::free_space && !kh32_.p) {
        kh32_.alloc(kh64_.n);
        cast_d2f((std::int64_t)kh64_.n, kh64_.p, kh32_.p, s);
    }
    if (kind_ != Kind::dense && !plan32_) {
        const int n = kind_ == Kind::free_space ? 2 * Ns_ : Ns_;
        plan32_ = std::make_unique<SpectralPlan<float>>(n, Ns_, Ns_, Ns_, Ns_);
    }
}

void Operator::release_single() {
    plan32_.reset();
    kh32_.reset();
    khp32_.reset();
    tmp32_.reset();
}

template <>
SpectralPlan<double>& Operator::plan<double>(hipStream_t) {
    return *plan64_;
}
template <>
SpectralPlan<float>& Operator::plan<float>(hipStream_t s) {
    if (!plan32_) build_single(s);
    return *plan32_;
}
template <typename T>
void Operator::apply_t(const T* x, T* y, hipStream_t s, hipEvent_t* ev, const SpecBatch* gate) {
    if (!built_) throw std::logic_error("FFTOperator::matvec: build() not called");
    if (kind_ == Kind::free_space) {
        SpectralPlan<T>& pl = plan<T>(s);
        int kp = 0;
        const T* kh = table_for<T>(pl, s, kp);
        pl.convolve(x, Ns_, y, Ns_, Symbol::table, kh, kp, 0.0, s, RowIn::plain, RowOut::plain,
                    nullptr, nullptr, 0, nullptr, ev, pl.register_blocked() ? gate : nullptr);
    } else if (kind_ == Kind::periodic) {
        // q = 2 pi k/L, u_hat = 2 p_hat/(E*|q|), unnormalised pair -> 1/Ns^2
        // (src/periodic_fft_operator.cpp:11-24)
        const double pi = std::acos(-1.0);
        const double sc = L_ / (pi * E_ * static_cast<double>(Ns_) * Ns_);
        SpectralPlan<T>& pl = plan<T>(s);
        pl.convolve(x, Ns_, y, Ns_, Symbol::periodic, nullptr, 0, sc, s, RowIn::plain,
                    RowOut::plain, nullptr, nullptr, 0, nullptr, ev,
                    pl.register_blocked() ? gate : nullptr);
    } else {
        if (ev) throw std::invalid_argument("stage timing needs a spectral backend");
        const std::int64_t N = (std::int64_t)Ns_ * Ns_;
        if (x == y) {
            DevBuf<T>& tmp = [&]() -> DevBuf<T>& {
                if constexpr (std::is_same_v<T, double>) return tmp64_;
                else return tmp32_;
            }();
            if (tmp.n != (std::size_t)N) tmp.alloc(N);
            VecOps<T>::copy(N, x, tmp.p, s);
            VecOps<T>::dense(Ns_, table_.p, tmp.p, y, s);
        } else {
            VecOps<T>::dense(Ns_, table_.p, x, y, s);
        }
    }
    fault_after_apply<T>(y, s);
}

template <typename T>
bool Operator::prepared() const {
    if (!built_ || kind_ == Kind::dense) return false;
    const SpectralPlan<T>* pl;
    const DevBuf<T>* perm;
    if constexpr (std::is_same_v<T, double>) {
        pl = plan64_.get();
        perm = &khp64_;
    } else {
        pl = plan32_.get();
        perm = &khp32_;
    }
    if (!pl || !pl->register_blocked()) return false;
    return kind_ == Kind::periodic || perm->p != nullptr;
}

template <typename T>
void Operator::prepare(hipStream_t s) {
    if (!built_) throw std::logic_error("FFTOperator::prepare: build() not called");
    if (kind_ == Kind::dense) return;
    SpectralPlan<T>& pl = plan<T>(s);
    if (kind_ == Kind::free_space) {
        int kp = 0;
        (void)table_for<T>(pl, s, kp);
    }
}
template void Operator::prepare<double>(hipStream_t);
template void Operator::prepare<float>(hipStream_t);

template <typename T>
std::size_t Operator::workspace_external() const {
    const SpectralPlan<T>* pl;
    if constexpr (std::is_same_v<T, double>) pl = plan64_.get();
    else pl = plan32_.get();
    return pl ? pl->external_ws_bytes() : 0;
}

template <typename T>
void Operator::apply_external(const T* x, T* y, void* ws, std::size_t ws_bytes,
                              hipStream_t s) {
    if (kind_ == Kind::dense)
        throw std::invalid_argument("matvec without allocation: spectral backends only "
                                    "(fft, periodic)");
    if (!prepared<T>())
        throw std::logic_error(std::string("matvec without allocation: the operator is not "
                                           "prepared in ") +
                               (std::is_same_v<T, double> ? "FP64" : "FP32") +
                               " (power-of-two grids 16..4096 free space, 32..8192 periodic; "
                               "run one allocating matvec in this precision first)");
    SpectralPlan<T>& pl = plan<T>(s);
    const std::size_t need = pl.external_ws_bytes();
    if (ws_bytes < need)
        throw std::invalid_argument("matvec without allocation: needs " + std::to_string(need) +
                                    " bytes of workspace, " + std::to_string(ws_bytes) +
                                    " given");
    if (kind_ == Kind::free_space) {
        const T* kh;
        if constexpr (std::is_same_v<T, double>) kh = khp64_.p;
        else kh = khp32_.p;
        pl.convolve(x, Ns_, y, Ns_, Symbol::table, kh, pl.table_pitch_half(), 0.0, s,
                    RowIn::plain, RowOut::plain, nullptr, nullptr, 0, nullptr, nullptr, nullptr,
                    ws);
    } else {
        const double pi = std::acos(-1.0);
        const double sc = L_ / (pi * E_ * static_cast<double>(Ns_) * Ns_);
        pl.convolve(x, Ns_, y, Ns_, Symbol::periodic, nullptr, 0, sc, s, RowIn::plain,
                    RowOut::plain, nullptr, nullptr, 0, nullptr, nullptr, nullptr, ws);
    }
    fault_after_apply<T>(y, s);
}
template void Operator::apply_external<double>(const double*, double*, void*, std::size_t,
                                               hipStream_t);
template void Operator::apply_external<float>(const float*, float*, void*, std::size_t,
                                              hipStream_t);
template bool Operator::prepared<double>() const;
template bool Operator::prepared<float>() const;
template std::size_t Operator::workspace_external<double>() const;
template std::size_t Operator::workspace_external<float>() const;

template <typename T>
const T* Operator::table_for(const SpectralPlan<T>& pl, hipStream_t s, int& pitch) {
    if (kind_ != Kind::free_space || !built_)
        throw std::logic_error("table_for: free-space operator not built");
    const T* kh;
    DevBuf<T>* perm;
    if constexpr (std::is_same_v<T, double>) {
        kh = kh64_.p;
        perm = &khp64_;
    } else {
        if (!kh32_.p) build_single(s);
        kh = kh32_.p;
        perm = &khp32_;
    }
    if (!pl.table_permuted()) {
        pitch = Ns_ + 1;
        return kh;
    }
    if (pl.n != 2 * Ns_ || pl.nh != Ns_ + 1)
        throw std::invalid_argument("table_for: plan does not match the operator");
    if (!perm->p) {
        perm->alloc(pl.table_elems_half());
        pl.permute_table_half(kh, Ns_ + 1, perm->p, s);
    }
    pitch = pl.table_pitch_half();
    return perm->p;
}
template const double* Operator::table_for<double>(const SpectralPlan<double>&, hipStream_t,
                                                   int&);
template const float* Operator::table_for<float>(const SpectralPlan<float>&, hipStream_t, int&);

template <typename T>
void Operator::apply_gated(const T* x, T* y, hipStream_t s, const double* gscal, int gate) {
    SpecBatch sb;
    sb.count = 1;
    sb.gscal = gscal;
    sb.gate = gate;
    apply_t<T>(x, y, s, nullptr, &sb);
}
template void Operator::apply_gated<double>(const double*, double*, hipStream_t, const double*,
                                            int);
template void Operator::apply_gated<float>(const float*, float*, hipStream_t, const double*, int);

void Operator::apply(const double* x, double* y, hipStream_t s) { apply_t<double>(x, y, s, nullptr); }
void Operator::apply(const float* x, float* y, hipStream_t s) { apply_t<float>(x, y, s, nullptr); }
void Operator::apply_timed(const double* x, double* y, hipStream_t s, hipEvent_t* ev) {
    apply_t<double>(x, y, s, ev);
}
void Operator::apply_timed(const float* x, float* y, hipStream_t s, hipEvent_t* ev) {
    apply_t<float>(x, y, s, ev);
}

std::size_t Operator::bytes_spectrum() const {
    return kh64_.bytes() + kh32_.bytes() + khp64_.bytes() + khp32_.bytes();
}
std::size_t Operator::bytes_scratch() const {
    std::size_t b = 0;
    if (plan64_) b += plan64_->workspace_bytes();
    if (plan32_) b += plan32_->workspace_bytes();
    return b;
}

// ── Rough-surface generator ────────────────────────────────────────────────

void rough_surface_device(int Ns, double hurst, double k_low, double k_high, double rms,
                          unsigned long long seed, double* out, hipStream_t s) {
    if (Ns < 32 || Ns > 8192 || (Ns & (Ns - 1)))
        throw std::invalid_argument("rough_surface: Ns must be a power of two in [32, 8192]");
    if (!(rms > 0.0) || !std::isfinite(hurst) || !(k_high >= k_low))
        throw std::invalid_argument("rough_surface: need rms > 0 and k_low <= k_high");
    const std::int64_t N = (std::int64_t)Ns * Ns;
    SpectralPlan<double> pl(Ns, Ns, Ns, Ns, Ns);
    pl.generate_surface(hurst, k_low, k_high, 1.0 / (double)N, seed, out, s);
    // mean removal, then population std after centring (numpy h.std())
    const int nb = vec_blocks(N);
    DevBuf<double> scal(S_COUNT), part((std::size_t)nb * kPartVals);
    static const int d2[2] = {S_R0, S_R1};
    double h[S_COUNT];
    auto moments = [&]() {
        VecOps<double>::moments(N, out, part.p, nb, s);
        finalize<double>(part.p, nb, 2, kSum2, d2, scal.p, STEP_NONE, s);
        ASPHER_HIP(hipMemcpyAsync(h, scal.p, sizeof(h), hipMemcpyDeviceToHost, s));
        ASPHER_HIP(hipStreamSynchronize(s));
    };
    moments();
    VecOps<double>::shift(N, out, h[S_R0] / (double)N, out, s);
    moments();
    const double sd = std::sqrt(h[S_R1] / (double)N);
    if (!(sd > 0.0))
        throw std::invalid_argument("rough_surface: empty band (no wavenumber in [k_low, k_high])");
    VecOps<double>::scale(N, out, rms / sd, s);
    ASPHER_HIP(hipStreamSynchronize(s));
}

// ── Preconditioners ────────────────────────────────────────────────────────

void stencil_weights(int Ns, int R, std::vector<int>& dx, std::vector<int>& dy,
                     std::vector<double>& w) {
    // w = F^-1 |k| on Nw = min(Ns, 512) (src/stencil_precond.cpp:19-43,
    // 199-220): the half spectrum C(kx, ky) = hypot(kx, kof(ky)), DC zeroed,
    // inverse c2c along ky then c2r along kx (Im of the DC and Nyquist
    // columns ignored), evaluated directly at the (2R+1)^2 needed offsets.
    static std::mutex mu;
    static std::map<std::pair<int, int>, std::tuple<std::vector<int>, std::vector<int>,
                                                    std::vector<double>>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(Ns, R);
    auto it = cache.find(key);
    if (it != cache.end()) {
        std::tie(dx, dy, w) = it->second;
        return;
    }
    const int Nw = Ns < 512 ? Ns : 512;
    const int nh = Nw / 2 + 1;
    const long double two_pi = 6.283185307179586476925286766559005768L;
    std::vector<long double> cs(Nw), sn(Nw);
    for (int m = 0; m < Nw; ++m) {
        cs[m] = cosl(two_pi * m / Nw);
        sn[m] = sinl(two_pi * m / Nw);
    }
    auto kof = [Nw](int i) { return (i < Nw / 2) ? i : i - Nw; };
    // the offsets actually needed, as grid indices
    std::vector<int> need;
    for (int d = -R; d <= R; ++d) need.push_back(((d % Nw) + Nw) % Nw);
    std::sort(need.begin(), need.end());
    need.erase(std::unique(need.begin(), need.end()), need.end());
    auto G = [&](int ix, int iy) -> long double {
        long double acc = 0.0L;
        for (int kx = 0; kx < nh; ++kx) {
            // B(kx, iy) = sum_ky C(kx, ky) exp(+2 pi i ky iy / Nw)
            long double br = 0.0L, bi = 0.0L;
            for (int ky = 0; ky < Nw; ++ky) {
                if (kx == 0 && ky == 0) continue;
                const long double c = std::hypot((double)kx, (double)kof(ky));
                const int m = (int)(((long long)ky * iy) % Nw);
                br += c * cs[m];
                bi += c * sn[m];
            }
            const bool edge = (kx == 0) || (Nw % 2 == 0 && kx == Nw / 2);
            if (edge) {
                const int m = (int)(((long long)kx * ix) % Nw);
                acc += br * cs[m];
            } else {
                const int m = (int)(((long long)kx * ix) % Nw);
                acc += 2.0L * (br * cs[m] - bi * sn[m]);
            }
        }
        return acc;
    };
    std::map<std::pair<int, int>, long double> val;
    for (int iy : need)
        for (int ix : need) val[{ix, iy}] = G(ix, iy);
    const long double c0 = val[{0, 0}];
    dx.clear();
    dy.clear();
    w.clear();
    for (int ddy = -R; ddy <= R; ++ddy)
        for (int ddx = -R; ddx <= R; ++ddx) {
            if (ddx * ddx + ddy * ddy > R * R) continue;
            const int wx = ((ddx % Nw) + Nw) % Nw, wy = ((ddy % Nw) + Nw) % Nw;
            dx.push_back(ddx);
            dy.push_back(ddy);
            w.push_back((ddx == 0 && ddy == 0) ? 1.0 : (double)(val[{wx, wy}] / c0));
        }
    cache[key] = std::make_tuple(dx, dy, w);
}

Precond::Precond(Kind kind, int Ns, int radius) : kind_(kind), Ns_(Ns), R_(radius) {
    if (kind_ == Kind::stencil) {
        if (radius < 1) throw std::invalid_argument("stencil radius must be >= 1");
        if (2 * radius + 1 > Ns)
            throw std::invalid_argument("stencil radius too large for the grid");
        // the device stencil stages a (64 + 2R) x (16 + 2R) FP64 halo in LDS
        if (radius > 26)
            throw std::invalid_argument("stencil radius above 26 (the LDS halo of the device "
                                        "stencil) is not supported");
        stencil_weights(Ns, radius, dx_, dy_, w_);
        std::vector<int> d(2 * dx_.size());
        for (std::size_t t = 0; t < dx_.size(); ++t) {
            d[2 * t] = dx_[t];
            d[2 * t + 1] = dy_[t];
        }
        tapd_.alloc(d.size());
        tapw_.alloc(w_.size());
        ASPHER_HIP(hipMemcpy(tapd_.p, d.data(), sizeof(int) * d.size(), hipMemcpyHostToDevice));
        ASPHER_HIP(hipMemcpy(tapw_.p, w_.data(), sizeof(double) * w_.size(),
                             hipMemcpyHostToDevice));
    } else if (kind_ == Kind::fourier) {
        // |k|/Ns^2 in float (src/fourier_precond.cpp:12-28)
        norm_ = 1.0f / (static_cast<float>(Ns) * static_cast<float>(Ns));
    }
}

template <typename T>
void Precond::apply(const T* g, const std::uint8_t* mask, T* z, double* scal, double* part,
                    hipStream_t s, int gate, bool zero_off) {
    static const int dz[1] = {S_ZSUM};
    // gate >= 0: every launch is skipped on the device while scal[gate] is 0
    Ens eg;
    eg.gscal = scal;
    eg.gate = gate;
    if (kind_ == Kind::stencil) {
        const int nblk = vec_blocks((std::int64_t)Ns_ * Ns_);
        VecOps<T>::stencil(Ns_, g, mask, scal, tapd_.p, tapw_.p, (int)w_.size(), R_, z, part,
                           nblk, s, eg, zero_off);
        finalize<T>(part, nblk, 1, kSum1, dz, scal, STEP_ZMEAN, s, eg);
    } else if (kind_ == Kind::fourier) {
        prepare<T>();
        SpectralPlan<T>* pl;
        if constexpr (std::is_same_v<T, double>) pl = plan64_.get();
        else pl = plan32_.get();
        if (pl->row_out_blocks() > kMaxPartRows)
            throw std::invalid_argument("grid too large for the preconditioner reduction");
        SpecBatch sb;
        sb.gscal = scal;
        sb.gate = gate;
        pl->convolve(g, Ns_, z, Ns_, Symbol::precond, nullptr, 0, (double)norm_, s,
                     RowIn::masked, RowOut::masked, mask, scal, S_GMEAN, part, nullptr,
                     gate >= 0 && pl->register_blocked() ? &sb : nullptr);
        finalize<T>(part, pl->row_out_blocks(), 1, kSum1, dz, scal, STEP_ZMEAN, s, eg);
    }
}

template <typename T>
void Precond::prepare() {
    if (kind_ != Kind::fourier) return;
    if constexpr (std::is_same_v<T, double>) {
        if (!plan64_) plan64_ = std::make_unique<SpectralPlan<double>>(Ns_, Ns_, Ns_, Ns_, Ns_);
    } else {
        if (!plan32_) plan32_ = std::make_unique<SpectralPlan<float>>(Ns_, Ns_, Ns_, Ns_, Ns_);
    }
}
template void Precond::prepare<double>();
template void Precond::prepare<float>();

template void Precond::apply<double>(const double*, const std::uint8_t*, double*, double*,
                                     double*, hipStream_t, int, bool);
template void Precond::apply<float>(const float*, const std::uint8_t*, float*, double*,
                                    double*, hipStream_t, int, bool);

} // namespace aspher
