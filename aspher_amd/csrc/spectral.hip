// Spectral convolution kernels for gfx950 (see spectral.hpp for the
// pipeline). Row kernels: one workgroup transforms `batch` rows in LDS; a real
// row of even length n is packed as n/2 complex values, transformed, and
// split into the n/2+1 half-spectrum values. Column kernel: one workgroup
// owns W adjacent spectrum columns, holds each complete column in LDS and
// runs forward transform, symbol multiply and inverse transform back to back
// (the reference fuses the same three steps per 16-column block in cache,
// src/periodic_fft_operator.cpp:56-101). Workgroup-to-tile mapping is
// XCD-aware: gfx950 dispatches consecutive workgroups round-robin over its 8
// XCDs, so the tiles are renumbered to give each XCD a contiguous range of
// columns whose strided row segments then share that XCD's L2 lines.
#include "fft_core.hpp"
#include "reduce.hpp"
#include "spectral.hpp"
#include "spectral_rb.hpp"

#include <type_traits>

#include <cmath>
#include <map>
#include <mutex>

namespace aspher {

using fft::Radices;

namespace {

__device__ __forceinline__ int xcd_tile(int b, int G) {
    const int q = G >> 3, r = G & 7;
    const int x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

template <typename T, bool POW2, int IN>
__global__ void __launch_bounds__(1024)
k_rows_fwd(const T* __restrict__ x, int x_pitch, int ncols, int nrows,
           cx<T>* __restrict__ A, int pitchA, int n, int logL, int batch,
           Radices rad, const cx<T>* __restrict__ tw,
           const std::uint8_t* __restrict__ mask, const double* __restrict__ scal,
           int gmean_idx, const double* __restrict__ table, int Ns_embed) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const bool packed = !(n & 1);
    const int L = packed ? (n >> 1) : n;
    const int bstride = L;
    const int row0 = blockIdx.x * batch;
    T gmean = T(0);
    if constexpr (IN == 1) gmean = static_cast<T>(scal[gmean_idx]);

    auto load = [&](int row, int col) -> T {
        if (col >= ncols) return T(0);
        if constexpr (IN == 2) {
            const int Ns = Ns_embed;
            if (row == Ns || col == Ns) return T(0);
            const int ady = row < Ns ? row : n - row;
            const int adx = col < Ns ? col : n - col;
            return static_cast<T>(table[(size_t)ady * Ns + adx]);
        } else {
            const size_t i = (size_t)row * x_pitch + col;
            if constexpr (IN == 1) return mask[i] ? T(x[i] - gmean) : T(0);
            else return x[i];
        }
    };

    for (int e = threadIdx.x; e < batch * L; e += blockDim.x) {
        const int b = e / L, j = e - b * L;
        const int row = row0 + b;
        cx<T> z = {T(0), T(0)};
        if (row < nrows) {
            if (packed) {
                z.re = load(row, 2 * j);
                z.im = load(row, 2 * j + 1);
            } else {
                z.re = load(row, j);
            }
        }
        buf[b * bstride + j] = z;
    }
    __syncthreads();
    cx<T>* res = buf;
    const int twmul = packed ? 2 : 1;
    if constexpr (POW2)
        fft::fft_pow2<T, false>(buf, logL, batch, bstride, tw, twmul);
    else
        res = fft::fft_generic<T, false>(buf, buf + batch * bstride, L, batch,
                                         bstride, rad, tw, twmul);
    const int nh = (n >> 1) + 1;
    for (int e = threadIdx.x; e < batch * nh; e += blockDim.x) {
        const int b = e / nh, k = e - b * nh;
        const int row = row0 + b;
        if (row >= nrows) continue;
        cx<T> X;
        if (packed) {
            const cx<T> Zk = res[b * bstride + (k == L ? 0 : k)];
            const cx<T> Zc = conjg(res[b * bstride + (k == 0 ? 0 : L - k)]);
            const cx<T> E = scale(Zk + Zc, T(0.5));
            const cx<T> D = Zk - Zc;
            const cx<T> O = {D.im * T(0.5), -D.re * T(0.5)};
            X = E + tw[k] * O;
        } else {
            X = res[b * bstride + k];
        }
        A[(size_t)row * pitchA + k] = X;
    }
}

template <typename T, int SYM>
__device__ __forceinline__ T symbol(int kx, int ky, int n, const T* __restrict__ ktab,
                                    int kp, double sc) {
    if constexpr (SYM == 0) {
        const int kyp = (ky <= (n >> 1)) ? ky : n - ky;
        return ktab[(size_t)kx * kp + kyp];
    } else if constexpr (SYM == 1) {
        const int my = (ky < ((n + 1) >> 1)) ? ky : ky - n;
        const double m2 = (double)kx * kx + (double)my * my;
        return m2 > 0.0 ? static_cast<T>(sc / sqrt(m2)) : T(0);
    } else {
        const int kof = (ky < (n >> 1)) ? ky : ky - n;
        const float h = (float)sqrt((double)kx * kx + (double)kof * kof);
        const float w = (kx == 0 && ky == 0) ? 0.0f : (float)sc * h;
        return static_cast<T>(w);
    }
}

template <typename T, bool POW2, int SYM>
__global__ void __launch_bounds__(1024)
k_cols(cx<T>* __restrict__ A, int pitchA, int ncols, int rows_in, int rows_out,
       int n, int logn, int W, int bstride, Radices rad,
       const cx<T>* __restrict__ tw, const T* __restrict__ ktab, int kp,
       double sc, T* __restrict__ kout, double* __restrict__ part, int remap) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[32];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const int tile = remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const int c0 = tile * W;
    for (int e = threadIdx.x; e < W * n; e += blockDim.x) {
        const int iy = e / W, c = e - iy * W;
        const int kx = c0 + c;
        cx<T> v = {T(0), T(0)};
        if (iy < rows_in && kx < ncols) v = A[(size_t)iy * pitchA + kx];
        buf[c * bstride + iy] = v;
    }
    __syncthreads();
    cx<T>* res = buf;
    cx<T>* other = buf + W * bstride;
    if constexpr (POW2) {
        fft::fft_pow2<T, false>(buf, logn, W, bstride, tw, 1);
    } else {
        res = fft::fft_generic<T, false>(buf, other, n, W, bstride, rad, tw, 1);
        other = (res == buf) ? buf + W * bstride : buf;
    }
    if constexpr (SYM == 3) {
        // kernel-spectrum build (src/fft_operator.cpp:46-68): realness check
        // over every mode, store Re/n^2 for ky in [0, n/2] (even in ky)
        const double norm = 1.0 / ((double)n * (double)n);
        double mi = 0.0, mr = 0.0;
        for (int e = threadIdx.x; e < W * n; e += blockDim.x) {
            const int c = e / n, ky = e - c * n;
            const int kx = c0 + c;
            if (kx >= ncols) continue;
            const cx<T> v = res[c * bstride + ky];
            mi = fmax(mi, fabs((double)v.im));
            mr = fmax(mr, fabs((double)v.re));
            if (ky <= (n >> 1)) kout[(size_t)kx * kp + ky] = static_cast<T>(norm * (double)v.re);
        }
        mi = block_reduce<RMAX>(mi, red);
        if (threadIdx.x == 0) part[2 * blockIdx.x] = mi;
        mr = block_reduce<RMAX>(mr, red);
        if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = mr;
        return;
    } else {
        for (int e = threadIdx.x; e < W * n; e += blockDim.x) {
            const int c = e / n, ky = e - c * n;
            const int kx = c0 + c;
            const T s = (kx < ncols) ? symbol<T, SYM>(kx, ky, n, ktab, kp, sc) : T(0);
            res[c * bstride + ky] = scale(res[c * bstride + ky], s);
        }
        __syncthreads();
        cx<T>* out = res;
        if constexpr (POW2)
            fft::fft_pow2<T, true>(res, logn, W, bstride, tw, 1);
        else
            out = fft::fft_generic<T, true>(res, other, n, W, bstride, rad, tw, 1);
        for (int e = threadIdx.x; e < W * rows_out; e += blockDim.x) {
            const int iy = e / W, c = e - iy * W;
            const int kx = c0 + c;
            if (kx < ncols) A[(size_t)iy * pitchA + kx] = out[c * bstride + iy];
        }
    }
}

template <typename T, bool POW2, int OUT>
__global__ void __launch_bounds__(1024)
k_rows_inv(const cx<T>* __restrict__ A, int pitchA, int nrows, int n, int logL,
           int batch, Radices rad, const cx<T>* __restrict__ tw,
           T* __restrict__ y, int y_pitch, int ncols,
           const std::uint8_t* __restrict__ mask, double* __restrict__ part) {
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[32];
    cx<T>* buf = reinterpret_cast<cx<T>*>(smem);
    const bool packed = !(n & 1);
    const int L = packed ? (n >> 1) : n;
    const int bstride = L;
    const int row0 = blockIdx.x * batch;
    const int nh = (n >> 1) + 1;
    if (packed) {
        // Z_k = (X_k + conj X_{L-k}) + i (X_k - conj X_{L-k}) exp(+2 pi i k/n);
        // Im of the DC and Nyquist modes is ignored (c2r convention)
        for (int e = threadIdx.x; e < batch * L; e += blockDim.x) {
            const int b = e / L, k = e - b * L;
            const int row = row0 + b;
            cx<T> z = {T(0), T(0)};
            if (row < nrows) {
                const cx<T>* a = A + (size_t)row * pitchA;
                cx<T> Xk = a[k], Xc = a[L - k];
                if (k == 0) {
                    Xk.im = T(0);
                    Xc.im = T(0);
                }
                Xc = conjg(Xc);
                const cx<T> S = Xk + Xc;
                const cx<T> D = (Xk - Xc) * conjg(tw[k]);
                z = {S.re - D.im, S.im + D.re};
            }
            buf[b * bstride + k] = z;
        }
    } else {
        for (int e = threadIdx.x; e < batch * n; e += blockDim.x) {
            const int b = e / n, m = e - b * n;
            const int row = row0 + b;
            cx<T> z = {T(0), T(0)};
            if (row < nrows) {
                const cx<T>* a = A + (size_t)row * pitchA;
                z = (m < nh) ? a[m] : conjg(a[n - m]);
                if (m == 0) z.im = T(0);
            }
            buf[b * bstride + m] = z;
        }
    }
    __syncthreads();
    cx<T>* res = buf;
    const int twmul = packed ? 2 : 1;
    if constexpr (POW2)
        fft::fft_pow2<T, true>(buf, logL, batch, bstride, tw, twmul);
    else
        res = fft::fft_generic<T, true>(buf, buf + batch * bstride, L, batch,
                                        bstride, rad, tw, twmul);
    double acc = 0.0;
    for (int e = threadIdx.x; e < batch * ncols; e += blockDim.x) {
        const int b = e / ncols, col = e - b * ncols;
        const int row = row0 + b;
        if (row >= nrows) continue;
        T v;
        if (packed) {
            const cx<T> z = res[b * bstride + (col >> 1)];
            v = (col & 1) ? z.im : z.re;
        } else {
            v = res[b * bstride + col].re;
        }
        const size_t i = (size_t)row * y_pitch + col;
        if constexpr (OUT == 1) {
            if (mask[i]) acc += (double)v;
            else v = T(0);
        }
        y[i] = v;
    }
    if constexpr (OUT == 1) {
        acc = block_reduce<RSUM>(acc, red);
        if (threadIdx.x == 0) part[(std::size_t)blockIdx.x * kPartVals] = acc;
    }
}

// hipFuncSetAttribute once per kernel for LDS above the 64 KiB default.
void allow_lds(const void* fn, std::size_t bytes) {
    static std::mutex mu;
    static std::map<const void*, std::size_t> done;
    if (bytes <= 65536) return;
    std::lock_guard<std::mutex> lk(mu);
    auto it = done.find(fn);
    if (it != done.end() && it->second >= bytes) return;
    ASPHER_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(bytes)));
    done[fn] = bytes;
}

Radices to_radices(const std::vector<int>& v) {
    Radices r{};
    if (v.size() > 24) throw std::invalid_argument("transform length has too many factors");
    r.n = static_cast<int>(v.size());
    for (std::size_t i = 0; i < v.size(); ++i) r.r[i] = v[i];
    return r;
}

int ilog2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

constexpr std::size_t kMaxLds = 160 * 1024;

} // namespace

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

std::vector<int> factorize(int L) {
    std::vector<int> r;
    int m = L;
    for (int f : {8, 4, 2}) {
        while (m % f == 0 && (f != 8 || m >= 8)) {
            r.push_back(f);
            m /= f;
        }
    }
    for (int p = 3; m > 1 && p * p <= m; p += 2)
        while (m % p == 0) {
            r.push_back(p);
            m /= p;
        }
    if (m > 1) r.push_back(m);
    return r;
}

// exp(-2 pi i t/n), t < n, in precision T: evaluated once per (n, T) in
// long double (n long-double sin/cos cost ~1.5 ms at n = 8192, and every
// operator build creates several plans) and kept for the process
template <typename T>
const std::vector<cx<T>>& twiddle_host(int n) {
    static std::mutex mu;
    static std::map<int, std::vector<cx<T>>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(n);
    if (it != cache.end()) return it->second;
    std::vector<cx<T>> h(n);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < n; ++t) {
        long long a = t, N = n;
        // angle = 2 pi a / N ; use exact symmetries where they apply
        long double ang = two_pi * (long double)a / (long double)N;
        long double c = cosl(ang), s = sinl(ang);
        if (4 * a == N) { c = 0; s = 1; }
        if (2 * a == N) { c = -1; s = 0; }
        if (4 * a == 3 * N) { c = 0; s = -1; }
        if (a == 0) { c = 1; s = 0; }
        h[t] = {static_cast<T>(c), static_cast<T>(-s)};
    }
    return cache.emplace(n, std::move(h)).first->second;
}

template <typename T>
Twiddles<T>::Twiddles(int n_) : n(n_) {
    const std::vector<cx<T>>& h = twiddle_host<T>(n);
    dev = static_cast<cx<T>*>(dev_alloc(sizeof(cx<T>) * n));
    ASPHER_HIP(hipMemcpy(dev, h.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
}

template <typename T>
Twiddles<T>::~Twiddles() {
    dev_free(dev, sizeof(cx<T>) * n);
}

// Mirror map of the register-blocked column transform of length n = 2^logn
// (host restatement of rb::Shape): for each column thread t, the chunk of the
// half-size symbol table holding its 16 values and whether it reads them
// reversed. Thread t' is t's mirror when it holds the frequencies n - ky of
// t's registers i at registers 15 - i; a pair shares one chunk (the lower
// thread reads it forward), every other thread gets a chunk of its own.
// Evaluated once per size for the process.
const std::vector<int>& mirror_map(int logn, int& nchunks) {
    static std::mutex mu;
    static std::map<int, std::pair<std::vector<int>, int>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(logn);
    if (it == cache.end()) {
        const int n = 1 << logn, A = logn / 4, RL = logn % 4, m = A + (RL ? 1 : 0);
        int N[8], S[8], P[8];
        for (int j = 0; j < m; ++j) N[j] = j < A ? 16 : (1 << RL);
        for (int j = 0; j < m; ++j) {
            S[j] = 1;
            for (int i = j + 1; i < m; ++i) S[j] *= N[i];
            P[j] = 1;
            for (int i = 0; i < j; ++i) P[j] *= N[i];
        }
        auto pos = [&](int st, int g, int v) {
            return (g / S[st]) * N[st] * S[st] + v * S[st] + g % S[st];
        };
        auto kidx = [&](int p) {
            int k = 0;
            for (int j = 0; j < m; ++j) k += ((p / S[j]) % N[j]) * P[j];
            return k;
        };
        const int Tn = n / 16, Nl = N[m - 1], Gl = 16 / Nl;
        std::vector<int> ky((std::size_t)Tn * 16), own_t(n), own_i(n);
        for (int t = 0; t < Tn; ++t)
            for (int q = 0; q < Gl; ++q)
                for (int v = 0; v < Nl; ++v) {
                    const int i = q * Nl + v, k = kidx(pos(m - 1, t * Gl + q, v));
                    ky[(std::size_t)t * 16 + i] = k;
                    own_t[k] = t;
                    own_i[k] = i;
                }
        std::vector<int> km(Tn, -1);
        int nc = 0;
        for (int t = 0; t < Tn; ++t) {
            if (km[t] >= 0) continue;
            const int tp = own_t[(n - ky[(std::size_t)t * 16]) % n];
            bool pair = tp > t && km[tp] < 0;
            for (int i = 0; i < 16 && pair; ++i) {
                const int k2 = (n - ky[(std::size_t)t * 16 + i]) % n;
                pair = own_t[k2] == tp && own_i[k2] == 15 - i;
            }
            km[t] = nc << 1;
            if (pair) km[tp] = (nc << 1) | 1;
            ++nc;
        }
        it = cache.emplace(logn, std::make_pair(std::move(km), nc)).first;
    }
    nchunks = it->second.second;
    return it->second.first;
}

template <typename T>
SpectralPlan<T>::SpectralPlan(int n_, int rin, int cin, int rout, int cout,
                              SpectralTuning tune, int batch)
    : n(n_), rows_in(rin), cols_in(cin), rows_out(rout), cols_out(cout), batch_(batch) {
    if (n < 2) throw std::invalid_argument("transform size must be >= 2");
    if (batch < 1) throw std::invalid_argument("batch must be >= 1");
    nh = n / 2 + 1;
    pitch = (nh + 15) / 16 * 16;
    const bool packed = (n % 2 == 0);
    const int Lr = packed ? n / 2 : n;
    pow2_ = is_pow2(n) && Lr >= 8;
    logn_ = ilog2(n);
    logL_row_ = ilog2(Lr);
    xcd_remap_ = tune.xcd_remap;
    const std::size_t cs = sizeof(cx<T>);
    rb_ = is_pow2(n) && n >= 32 && n <= 8192;
    if (rb_) {
        // threads per row/column are L/16 resp. n/16; the inverse row pass
        // aims for 256-thread workgroups (more resident ones: 6% faster than
        // 512 at 4096^2 FP32), the column pass for 512, while the LDS of one
        // workgroup stays <= 96 KiB
        const std::size_t rowlds = cs * (std::size_t)(Lr + Lr / 16 + 1);
        const std::size_t collds = cs * (std::size_t)(n + n / 16 + 1);
        int B = tune.row_batch > 0 ? tune.row_batch : std::max(1, 256 / (Lr / 16));
        while (B > 1 && (B * (Lr / 16) > 512 || B * rowlds > 98304)) --B;
        int W = tune.col_width > 0 ? tune.col_width : std::max(1, 512 / (n / 16));
        while (W > 1 && (W * (n / 16) > 512 || W * collds > 98304)) --W;
        if (B * (Lr / 16) > 512 || W * (n / 16) > 512 || B * rowlds > kMaxLds ||
            W * collds > kMaxLds)
            throw std::invalid_argument("spectral tiling exceeds the workgroup limits");
        row_batch_ = B;
        col_width_ = W;
        // the forward row pass writes the transposed half spectrum in runs of
        // B rows: it takes as many rows as 1024 threads and 144 KiB of LDS
        // allow (one resident workgroup), measured faster than two smaller
        // ones; the inverse pass keeps B (more resident workgroups)
        int Bf = tune.row_batch > 0 ? B : std::max(1, 1024 / (Lr / 16));
        while (tune.row_batch <= 0 && Bf > B && Bf * rowlds > 147456) --Bf;
        row_batch_fwd_ = std::max(Bf, B);
        // small transforms: the batches above would put a whole pass in one
        // or two workgroups (64 rows of a 128-point transform in one
        // 1024-thread group), so they are halved until a pass has 256
        // workgroups (one per CU) or a single row/column each. Measured on
        // the nested 4096^2 solves (levels 64^2 .. 1024^2): mixed with the
        // Fourier engine 0.0904 -> 0.0875 s, with the stencil 0.3455 ->
        // 0.3436 s; targets of 16 and 64 groups gained less.
        {
            const int nr = std::max(rows_in, rows_out);
            auto spread = [](int v, int count) {
                while (v > 1 && (count + v - 1) / v < 256) v = (v + 1) / 2;
                return v;
            };
            if (tune.row_batch <= 0) {
                row_batch_ = spread(row_batch_, nr);
                row_batch_fwd_ = std::max(spread(row_batch_fwd_, nr), row_batch_);
            }
            if (tune.col_width <= 0) col_width_ = spread(col_width_, nh);
        }
        pitchT_ =(std::max(rows_in, rows_out) + 31) / 32 * 32 + 32;
        // The column pass writes its result back into AT, transposed: every
        // store is a contiguous run, where row-major stores of single
        // elements into BR were written back ~2.15x over (measured at
        // 4096^2 FP32). Measured at 4096^2: FP32 matvec 0.326 -> 0.292 ms,
        // FP64 0.541 -> 0.535 ms; 2048^2 FP64 0.137 -> 0.118 ms.
        // The inverse row pass then reads runs of row_batch_inv_ = 2 rows per
        // frequency (16-byte runs in FP32, 32-byte in FP64; at least the
        // row-major pass's batch, within the forward pass's LDS budget): two
        // 512-thread workgroups are resident per CU, so one's loads overlap
        // the other's transform. Measured at 4096^2 FP32, inverse rows: 4
        // rows (one 1024-thread workgroup per CU) 84 us, 2 rows 78 us, 1
        // row (8-byte runs) 124 us; mixed solve 0.357 -> 0.352 s.
        tcols_ = tune.col_out_t;
        row_batch_inv_ = row_batch_;
        if (tcols_) row_batch_inv_ = std::min(row_batch_fwd_, std::max(row_batch_, 2));
        tw64_ = std::make_unique<Twiddles<double>>(n);
        a_rows_ = static_cast<std::size_t>(std::max(rows_out, 1));
        br_stride_ = (std::int64_t)a_rows_ * pitch;
        at_stride_ = (std::int64_t)nh * pitchT_;
        a_bytes_ = cs * (std::size_t)br_stride_ * batch_;
        at_bytes_ = cs * (std::size_t)at_stride_ * batch_;
        A_ = static_cast<cx<T>*>(dev_alloc(a_bytes_));
        AT_ = static_cast<cx<T>*>(dev_alloc(at_bytes_));
        const std::vector<int>& km = mirror_map(logn_, kmap_chunks_);
        kmap_ = static_cast<int*>(dev_alloc(sizeof(int) * km.size()));
        ASPHER_HIP(hipMemcpy(kmap_, km.data(), sizeof(int) * km.size(), hipMemcpyHostToDevice));
        return;
    }
    if (batch_ != 1)
        throw std::invalid_argument("batched transforms need a power-of-two length in [32, 8192]");
    if (pow2_) {
        row_batch_ = tune.row_batch > 0 ? tune.row_batch : std::max(1, 2048 / Lr);
        while (Lr * row_batch_ > 8192) row_batch_ /= 2;
        int w = tune.col_width > 0 ? tune.col_width
                                   : static_cast<int>(std::max<std::size_t>(
                                         1, std::min<std::size_t>(16, 65536 / (n * cs))));
        while (w > 1 && (w * n > 8192 || w * (std::size_t)(n + 1) * cs > kMaxLds)) w /= 2;
        col_width_ = w;
        if ((std::size_t)n * cs > kMaxLds || n > 8192 || Lr * row_batch_ > 8192)
            throw std::invalid_argument("transform length exceeds the LDS capacity");
    } else {
        rad_row_ = factorize(Lr);
        rad_col_ = factorize(n);
        row_batch_ = tune.row_batch > 0 ? tune.row_batch : std::max(1, std::min(64, 1024 / Lr));
        int w = tune.col_width > 0 ? tune.col_width : std::max(1, std::min(16, 1024 / n));
        while (w > 1 && 2 * w * (std::size_t)(n + 1) * cs > kMaxLds) w /= 2;
        col_width_ = w;
        if (2 * (std::size_t)(n + 1) * cs > kMaxLds || 2 * (std::size_t)Lr * cs * row_batch_ > kMaxLds)
            throw std::invalid_argument("non power-of-two transform length too large for LDS");
    }
    tw_ = std::make_unique<Twiddles<T>>(n);
    a_rows_ = static_cast<std::size_t>(std::max(rows_in, rows_out));
    a_bytes_ = cs * a_rows_ * pitch;
    A_ = static_cast<cx<T>*>(dev_alloc(a_bytes_));
}

template <typename T>
SpectralPlan<T>::~SpectralPlan() {
    dev_free(A_, a_bytes_);
    dev_free(AT_, at_bytes_);
    if (kmap_) dev_free(kmap_, sizeof(int) * (std::size_t)(n / 16));
}

template <typename T>
std::size_t SpectralPlan<T>::ext_br_bytes() const {
    // with the transposed column output the convolution never touches BR
    if (tcols_) return 0;
    return (sizeof(cx<T>) * (std::size_t)br_stride_ + 255) / 256 * 256;
}

template <typename T>
std::size_t SpectralPlan<T>::external_ws_bytes() const {
    if (!rb_) return 0;
    return ext_br_bytes() + sizeof(cx<T>) * (std::size_t)at_stride_;
}

template <typename T>
std::size_t SpectralPlan<T>::workspace_bytes() const {
    return sizeof(cx<T>) * a_rows_ * pitch * batch_ +
           (rb_ ? sizeof(cx<double>) * n + sizeof(cx<T>) * (std::size_t)nh * pitchT_ * batch_
                : sizeof(cx<T>) * n);
}

template <typename T>
void SpectralPlan<T>::generate_surface(double hurst, double k_low, double k_high, double scale,
                                       unsigned long long seed, T* out, hipStream_t s) {
    if constexpr (!std::is_same_v<T, double>) {
        throw std::logic_error("generate_surface: FP64 plans only");
    } else {
        if (!rb_ || rows_in != n || rows_out != n || cols_in != n || cols_out != n)
            throw std::invalid_argument(
                "generate_surface: needs a power-of-two n x n plan, 32 <= n <= 8192");
        RbGen g;
        g.BR = A_;
        g.pitch = pitch;
        g.hurst = hurst;
        g.k_low = k_low;
        g.k_high = k_high;
        g.scale = scale;
        g.seed = seed;
        g.tw = tw64_->dev;
        g.W = col_width_;
        g.remap = xcd_remap_ ? 1 : 0;
        rb_gen_cols(logn_, g, s);
        RbRowsInv<double> r;
        r.BR = A_;
        r.pitch = pitch;
        r.nrows = n;
        r.B = row_batch_;
        r.tw = tw64_->dev;
        r.y = out;
        r.y_pitch = n;
        r.ncols = n;
        r.mode = 0;
        r.remap = g.remap;
        rb_rows_inv<double>(ilog2(n / 2), r, s);
    }
}

template <typename T>
void SpectralPlan<T>::permute_table(const T* kunique, int kp, T* kout, hipStream_t s,
                                    bool odd) const {
    if (!rb_) throw std::logic_error("permute_table: generic plans read the unique table");
    rb_permute_table<T>(logn_, kunique, kp, nh, kout, s, odd);
}

template <typename T>
void SpectralPlan<T>::permute_table_half(const T* kunique, int kp, T* kout,
                                         hipStream_t s) const {
    if (!rb_) throw std::logic_error("permute_table_half: register-blocked plans only");
    rb_permute_table_half<T>(logn_, kunique, kp, nh, kmap_, table_pitch_half(), kout, s);
}

template <typename T>
int SpectralPlan<T>::row_out_blocks() const {
    const int b = rb_ ? row_batch_inv_ : row_batch_;
    return (rows_out + b - 1) / b;
}

template <typename T>
int SpectralPlan<T>::col_blocks() const {
    return (nh + col_width_ - 1) / col_width_;
}

#define ROWS_FWD_ARGS                                                                  \
    x, x_pitch, ncols, nrows, A_, pitch, n, logL_row_, row_batch_, rad, tw_->dev, mask, \
        scal, gmean_idx, table, Ns_embed

template <typename T>
void SpectralPlan<T>::rows_fwd(const T* x, int x_pitch, RowIn mode,
                               const std::uint8_t* mask, const double* scal,
                               int gmean_idx, const double* table, int Ns_embed,
                               int nrows, int ncols, hipStream_t s) {
    const bool packed = (n % 2 == 0);
    const int Lr = packed ? n / 2 : n;
    const Radices rad = to_radices(rad_row_);
    const int grid = (nrows + row_batch_ - 1) / row_batch_;
    const int block = pow2_ ? (Lr * row_batch_) / 8 : 256;
    const std::size_t lds = sizeof(cx<T>) * (std::size_t)Lr * row_batch_ * (pow2_ ? 1 : 2);
    const void* fn = nullptr;
    if (pow2_) {
        if (mode == RowIn::plain) fn = (const void*)k_rows_fwd<T, true, 0>;
        else if (mode == RowIn::masked) fn = (const void*)k_rows_fwd<T, true, 1>;
        else fn = (const void*)k_rows_fwd<T, true, 2>;
    } else {
        if (mode == RowIn::plain) fn = (const void*)k_rows_fwd<T, false, 0>;
        else if (mode == RowIn::masked) fn = (const void*)k_rows_fwd<T, false, 1>;
        else fn = (const void*)k_rows_fwd<T, false, 2>;
    }
    allow_lds(fn, lds);
    if (pow2_) {
        if (mode == RowIn::plain) k_rows_fwd<T, true, 0><<<grid, block, lds, s>>>(ROWS_FWD_ARGS);
        else if (mode == RowIn::masked) k_rows_fwd<T, true, 1><<<grid, block, lds, s>>>(ROWS_FWD_ARGS);
        else k_rows_fwd<T, true, 2><<<grid, block, lds, s>>>(ROWS_FWD_ARGS);
    } else {
        if (mode == RowIn::plain) k_rows_fwd<T, false, 0><<<grid, block, lds, s>>>(ROWS_FWD_ARGS);
        else if (mode == RowIn::masked) k_rows_fwd<T, false, 1><<<grid, block, lds, s>>>(ROWS_FWD_ARGS);
        else k_rows_fwd<T, false, 2><<<grid, block, lds, s>>>(ROWS_FWD_ARGS);
    }
    ASPHER_LAUNCH_CHECK();
}

#define COLS_ARGS                                                                    \
    A_, pitch, nh, rin, rout, n, logn_, col_width_, bstride, rad, tw_->dev, ktab, kp, \
        sc, kout, part, xcd_remap_ ? 1 : 0

template <typename T>
void SpectralPlan<T>::cols(Symbol sym, const T* ktab, int kp, double sc, T* kout,
                           double* part, int rin, int rout, hipStream_t s) {
    const Radices rad = to_radices(rad_col_);
    const int bstride = n + (col_width_ > 1 ? 1 : 0);
    const int grid = (nh + col_width_ - 1) / col_width_;
    const int block = pow2_ ? (n * col_width_) / 8 : 256;
    const std::size_t lds = sizeof(cx<T>) * (std::size_t)bstride * col_width_ * (pow2_ ? 1 : 2);
    const void* fn = nullptr;
#define COLS_FN(P, S) fn = (const void*)k_cols<T, P, S>
    switch (sym) {
        case Symbol::table: if (pow2_) COLS_FN(true, 0); else COLS_FN(false, 0); break;
        case Symbol::periodic: if (pow2_) COLS_FN(true, 1); else COLS_FN(false, 1); break;
        case Symbol::precond: if (pow2_) COLS_FN(true, 2); else COLS_FN(false, 2); break;
        case Symbol::build: if (pow2_) COLS_FN(true, 3); else COLS_FN(false, 3); break;
    }
#undef COLS_FN
    allow_lds(fn, lds);
#define COLS_LAUNCH(P, S) k_cols<T, P, S><<<grid, block, lds, s>>>(COLS_ARGS)
    switch (sym) {
        case Symbol::table: if (pow2_) COLS_LAUNCH(true, 0); else COLS_LAUNCH(false, 0); break;
        case Symbol::periodic: if (pow2_) COLS_LAUNCH(true, 1); else COLS_LAUNCH(false, 1); break;
        case Symbol::precond: if (pow2_) COLS_LAUNCH(true, 2); else COLS_LAUNCH(false, 2); break;
        case Symbol::build: if (pow2_) COLS_LAUNCH(true, 3); else COLS_LAUNCH(false, 3); break;
    }
#undef COLS_LAUNCH
    ASPHER_LAUNCH_CHECK();
}

#define ROWS_INV_ARGS                                                                   \
    A_, pitch, rows_out, n, logL_row_, row_batch_, rad, tw_->dev, y, y_pitch, cols_out, \
        mask, part

template <typename T>
void SpectralPlan<T>::rows_inv(T* y, int y_pitch, RowOut mode, const std::uint8_t* mask,
                               double* part, hipStream_t s) {
    const bool packed = (n % 2 == 0);
    const int Lr = packed ? n / 2 : n;
    const Radices rad = to_radices(rad_row_);
    const int grid = row_out_blocks();
    const int block = pow2_ ? (Lr * row_batch_) / 8 : 256;
    const std::size_t lds = sizeof(cx<T>) * (std::size_t)Lr * row_batch_ * (pow2_ ? 1 : 2);
    const void* fn = nullptr;
    if (pow2_) fn = (mode == RowOut::plain) ? (const void*)k_rows_inv<T, true, 0>
                                            : (const void*)k_rows_inv<T, true, 1>;
    else fn = (mode == RowOut::plain) ? (const void*)k_rows_inv<T, false, 0>
                                      : (const void*)k_rows_inv<T, false, 1>;
    allow_lds(fn, lds);
    if (pow2_) {
        if (mode == RowOut::plain) k_rows_inv<T, true, 0><<<grid, block, lds, s>>>(ROWS_INV_ARGS);
        else k_rows_inv<T, true, 1><<<grid, block, lds, s>>>(ROWS_INV_ARGS);
    } else {
        if (mode == RowOut::plain) k_rows_inv<T, false, 0><<<grid, block, lds, s>>>(ROWS_INV_ARGS);
        else k_rows_inv<T, false, 1><<<grid, block, lds, s>>>(ROWS_INV_ARGS);
    }
    ASPHER_LAUNCH_CHECK();
}

template <typename T>
void SpectralPlan<T>::convolve(const T* x, int x_pitch, T* y, int y_pitch, Symbol sym,
                               const T* ktab, int ktab_pitch, double sym_scale,
                               hipStream_t s, RowIn in_mode, RowOut out_mode,
                               const std::uint8_t* mask, const double* scal,
                               int gmean_idx, double* part, hipEvent_t* ev,
                               const SpecBatch* sb, void* ext_ws) {
    if (sym == Symbol::build) throw std::invalid_argument("convolve: build symbol");
    if (sb && (!rb_ || sb->count > batch_ || sb->count < 1))
        throw std::invalid_argument("convolve: batch exceeds the plan");
    if (ext_ws && (!rb_ || sb))
        throw std::invalid_argument("convolve: caller workspace needs a single-instance "
                                    "register-blocked plan");
    if (rb_) {
        // intermediate spectra: the plan's own, or the caller's workspace
        // (BR at its start, AT after it at a 256-byte boundary)
        cx<T>* ws_br = A_;
        cx<T>* ws_at = AT_;
        if (ext_ws) {
            ws_br = static_cast<cx<T>*>(ext_ws);
            ws_at = reinterpret_cast<cx<T>*>(static_cast<char*>(ext_ws) + ext_br_bytes());
        }
        RbBatch bt;
        if (sb) {
            bt.count = sb->count;
            bt.x_stride = sb->x_stride;
            bt.y_stride = sb->y_stride;
            bt.at_stride = at_stride_;
            bt.br_stride = br_stride_;
            bt.first = sb->first;
            bt.sstride = sb->sstride;
            bt.pstride = sb->pstride;
            bt.gscal = sb->gscal;
            bt.gate = sb->gate;
        }
        const int Lr = n / 2;
        RbRowsFwd<T> f;
        f.x = x;
        f.x_pitch = x_pitch;
        f.ncols = cols_in;
        f.nrows = rows_in;
        f.AT = ws_at;
        f.pitchT = pitchT_;
        f.B = row_batch_fwd_;
        f.tw = tw64_->dev;
        f.mode = in_mode == RowIn::masked ? 1 : 0;
        f.mask = mask;
        f.scal = scal;
        f.gmean_idx = gmean_idx;
        f.remap = xcd_remap_ ? 1 : 0;
        f.bt = bt;
        RbCols<T> c;
        c.bt = bt;
        c.AT = ws_at;
        c.pitchT = pitchT_;
        c.BR = ws_br;
        c.pitch = pitch;
        c.ncols = nh;
        c.rows_in = rows_in;
        c.rows_out = rows_out;
        c.W = col_width_;
        c.tw = f.tw;
        c.sym = static_cast<int>(sym);
        c.ktab = ktab;
        c.kp = ktab_pitch;
        c.kmap = kmap_;
        c.sc = sym_scale;
        c.remap = f.remap;
        if (tcols_) {
            // the column pass writes back into AT; the inverse rows read it
            // transposed in runs of row_batch_inv_ rows
            c.out_t = 1;
            c.BR = ws_at;
        }
        RbRowsInv<T> r;
        r.BR = ws_br;
        r.pitch = pitch;
        r.nrows = rows_out;
        if (tcols_) {
            r.in_t = 1;
            r.BR = ws_at;
            r.pitch = pitchT_;
        }
        r.B = row_batch_inv_;
        r.tw = f.tw;
        r.y = y;
        r.y_pitch = y_pitch;
        r.ncols = cols_out;
        r.mode = out_mode == RowOut::masked ? 1 : 0;
        r.mask = mask;
        r.part = part;
        r.remap = f.remap;
        r.bt = bt;
        if (ev) ASPHER_HIP(hipEventRecord(ev[0], s));
        rb_rows_fwd<T>(ilog2(Lr), f, s);
        if (ev) ASPHER_HIP(hipEventRecord(ev[1], s));
        rb_cols<T>(logn_, c, s);
        if (ev) ASPHER_HIP(hipEventRecord(ev[2], s));
        rb_rows_inv<T>(ilog2(Lr), r, s);
        if (ev) ASPHER_HIP(hipEventRecord(ev[3], s));
        return;
    }
    if (ev) ASPHER_HIP(hipEventRecord(ev[0], s));
    rows_fwd(x, x_pitch, in_mode, mask, scal, gmean_idx, nullptr, 0, rows_in, cols_in, s);
    if (ev) ASPHER_HIP(hipEventRecord(ev[1], s));
    cols(sym, ktab, ktab_pitch, sym_scale, nullptr, nullptr, rows_in, rows_out, s);
    if (ev) ASPHER_HIP(hipEventRecord(ev[2], s));
    rows_inv(y, y_pitch, out_mode, mask, part, s);
    if (ev) ASPHER_HIP(hipEventRecord(ev[3], s));
}

template <typename T>
void SpectralPlan<T>::build_spectrum(const double* table, int Ns, T* kout, double* part,
                                     hipStream_t s, int embed) {
    if (rows_in != n) throw std::invalid_argument("build_spectrum: plan must read all rows");
    if (rb_) {
        RbRowsFwd<T> f;
        f.ncols = n;
        f.nrows = n;
        f.AT = AT_;
        f.pitchT = pitchT_;
        f.B = row_batch_fwd_;
        f.tw = tw64_->dev;
        f.mode = embed;
        f.table = table;
        f.Ns_embed = Ns;
        f.remap = xcd_remap_ ? 1 : 0;
        RbCols<T> c;
        c.AT = AT_;
        c.pitchT = pitchT_;
        c.ncols = nh;
        c.rows_in = n;
        c.rows_out = 0;
        c.W = col_width_;
        c.tw = f.tw;
        c.sym = 3;
        c.kp = nh;
        c.kout = kout;
        c.part = part;
        c.remap = f.remap;
        rb_rows_fwd<T>(ilog2(n / 2), f, s);
        rb_cols<T>(logn_, c, s);
        return;
    }
    if (embed != 2) throw std::invalid_argument("build_spectrum: table kinds need a power-of-two plan");
    rows_fwd(nullptr, 0, RowIn::embed, nullptr, nullptr, 0, table, Ns, n, n, s);
    cols(Symbol::build, nullptr, nh, 0.0, kout, part, n, 0, s);
}

template <typename T>
void SpectralPlan<T>::convolve2(const T* x, int x_pitch, std::int64_t x_stride, T* y,
                                int y_pitch, std::int64_t y_stride, int sym, const T* kxx,
                                const T* kyy, const T* kxy, int kp, double gamma, double norm,
                                hipStream_t s, RowIn in_mode, RowOut out_mode,
                                const std::uint8_t* mask, const double* scal, int gmean_idx,
                                double* part, std::int64_t pstride) {
    if (!rb_ || batch_ < 2 || n > 4096)
        throw std::invalid_argument("convolve2: needs a power-of-two plan of batch >= 2, n <= 4096");
    RbBatch bt;
    bt.count = 2;
    bt.x_stride = x_stride;
    bt.y_stride = y_stride;
    bt.m_stride = 0;
    bt.at_stride = at_stride_;
    bt.br_stride = br_stride_;
    bt.pstride = pstride;
    RbRowsFwd<T> f;
    f.x = x;
    f.x_pitch = x_pitch;
    f.ncols = cols_in;
    f.nrows = rows_in;
    f.AT = AT_;
    f.pitchT = pitchT_;
    f.B = row_batch_fwd_;
    f.tw = tw64_->dev;
    f.mode = in_mode == RowIn::masked ? 1 : 0;
    f.mask = mask;
    f.scal = scal;
    f.gmean_idx = gmean_idx;
    f.remap = xcd_remap_ ? 1 : 0;
    f.bt = bt;
    RbCols2<T> c;
    c.AT = AT_;
    c.pitchT = pitchT_;
    c.at_stride = at_stride_;
    c.BR = A_;
    c.pitch = pitch;
    c.br_stride = br_stride_;
    c.ncols = nh;
    c.rows_in = rows_in;
    c.rows_out = rows_out;
    c.tw = f.tw;
    c.sym = sym;
    c.kxx = kxx;
    c.kyy = kyy;
    c.kxy = kxy;
    c.kp = kp;
    c.gamma = gamma;
    c.norm = norm;
    c.remap = f.remap;
    RbRowsInv<T> r;
    r.BR = A_;
    r.pitch = pitch;
    r.nrows = rows_out;
    r.B = row_batch_inv_;
    r.tw = f.tw;
    r.y = y;
    r.y_pitch = y_pitch;
    r.ncols = cols_out;
    r.mode = out_mode == RowOut::masked ? 1 : 0;
    r.mask = mask;
    r.part = part;
    r.remap = f.remap;
    r.bt = bt;
    rb_rows_fwd<T>(ilog2(n / 2), f, s);
    rb_cols2<T>(logn_, c, s);
    rb_rows_inv<T>(ilog2(n / 2), r, s);
}

template class Twiddles<float>;
template class Twiddles<double>;
template class SpectralPlan<float>;
template class SpectralPlan<double>;

} // namespace aspher
