// Power-of-two spectral convolution kernels built on the register-blocked
// FFT (fft_rb.hpp). Data flow of y = inv(s . fwd(x)) on an n x n grid:
//
//   rows_fwd : x rows (contiguous reads) -> half spectrum written
//              TRANSPOSED, AT[kx][iy] (strided writes, merged in L2)
//   cols     : AT columns (contiguous reads) -> forward, symbol, inverse ->
//              BR[iy][kx] row-major (strided writes)
//   rows_inv : BR rows (contiguous reads) -> y rows (contiguous writes)
//
// Every strided access is a store: stores retire without stalling the wave
// and the partial lines of neighbouring rows/columns, launched on the same
// XCD by the XCD-aware block mapping, merge in that XCD's L2 before they
// reach HBM; every load is a full, coalesced line.
#include "fft_rb.hpp"
#include "reduce.hpp"
#include "spectral_rb.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

namespace aspher {

namespace {

// two consecutive reals as one 8- (FP32) or 16-byte (FP64) access
template <typename T>
struct alignas(2 * sizeof(T)) Pair {
    T a, b;
};

__host__ __device__ __forceinline__ bool pair_aligned(const void* p, std::size_t elt) {
    return (reinterpret_cast<std::uintptr_t>(p) % (2 * elt)) == 0;
}

// Slot of natural-order element k in a row pass's LDS row (the
// digit-reversed write and packed-spectrum split of the forward pass, the
// transposed load and gather of the inverse). The B rows of a workgroup are
// read and written together there, lane t on row t % B at element t / B + c:
// rows take Skew32x2 and start P / B slots apart mod P (row_stride; P = 32
// for 8-byte elements, 16 for 16-byte ones), so the lanes the LDS serves
// together fall in distinct banks.
template <typename T>
__device__ __forceinline__ int row_ix(int k) {
    return rb::Skew32x2::ix(k);
}
template <typename T>
__host__ __device__ inline int row_stride(int L, int B) {
    const int base = rb::lds_len(L);
    constexpr int P = sizeof(cx<T>) >= 16 ? 16 : 32;
    if (B < 2 || B > P) return base;
    return base + ((P / B - base % P) % P + P) % P;
}

__device__ __forceinline__ int xcd_block(int b, int G) {
    const int q = G >> 3, r = G & 7;
    const int x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// ── rows: real -> half spectrum (transposed store) ─────────────────────────
// Z: the row's nonzero reals are x[0 .. L) of n = 2L (zero-padded free-space
// rows), so the upper half of the packed input is never loaded and the first
// radix-16 stage is pruned.
template <typename T, int LOGL, int IN, bool Z>
__global__ void __launch_bounds__(1024)
k_rb_rows_fwd(const T* __restrict__ x, int x_pitch, int ncols, int nrows,
              cx<T>* __restrict__ AT, int pitchT, int B, const cx<double>* __restrict__ tw,
              const std::uint8_t* __restrict__ mask, const double* __restrict__ scal,
              int gmean_idx, const double* __restrict__ table, int Ns_embed, int remap,
              RbBatch bt) {
    using Sh = rb::Shape<LOGL>;
    constexpr int L = Sh::n, n = 2 * L, TPR = Sh::threads;
    const int LL = row_stride<T>(L, B);
    extern __shared__ __align__(16) unsigned char smem[];
    {
        const int li = blockIdx.y, gi = bt.first + li;
        if (bt.gate >= 0 && bt.gscal[(std::size_t)gi * bt.sstride + bt.gate] == 0.0) return;
        x += li * bt.x_stride;
        if (mask) mask += li * (bt.m_stride >= 0 ? bt.m_stride : bt.x_stride);
        AT += li * bt.at_stride;
        if (scal) scal += (std::size_t)gi * bt.sstride;
    }
    cx<T>* lds_all = reinterpret_cast<cx<T>*>(smem);
    const int blk = remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int b = threadIdx.x / TPR, th = threadIdx.x - b * TPR;
    const int row = blk * B + b;
    cx<T>* lds = lds_all + b * LL;
    T gmean = T(0);
    if constexpr (IN == 1) gmean = static_cast<T>(scal[gmean_idx]);

    auto load = [&](int col) -> T {
        if (row >= nrows || col >= ncols) return T(0);
        if constexpr (IN >= 2 && IN <= 4) {
            // wrap-around embedding of an offset table (index d or n + d;
            // row/column Ns zero): 2 even, 3 even transposed, 4 odd-odd
            const int Ns = Ns_embed;
            if (row == Ns || col == Ns) return T(0);
            const int dy = row < Ns ? row : row - n;
            const int dx = col < Ns ? col : col - n;
            const int ady = dy < 0 ? -dy : dy, adx = dx < 0 ? -dx : dx;
            if constexpr (IN == 2) {
                return static_cast<T>(table[(size_t)ady * Ns + adx]);
            } else if constexpr (IN == 3) {
                return static_cast<T>(table[(size_t)adx * Ns + ady]);
            } else {
                if (dx == 0 || dy == 0) return T(0);
                const T v = static_cast<T>(table[(size_t)ady * Ns + adx]);
                return ((dx > 0) == (dy > 0)) ? v : T(-v);
            }
        } else {
            const size_t i = (size_t)row * x_pitch + col;
            if constexpr (IN == 1) return mask[i] ? T(x[i] - gmean) : T(0);
            else return x[i];
        }
    };

    cx<T> r[16];
    // plain or masked input on an even pitch and width from a pair-aligned
    // base: both reals of a packed element in one access (and their two
    // mask bytes in one)
    const bool pairs = (IN == 0 || IN == 1) && ((x_pitch | ncols) & 1) == 0 &&
                       pair_aligned(x, sizeof(T)) &&
                       (IN == 0 || (reinterpret_cast<std::uintptr_t>(mask) & 1u) == 0);
    if (pairs) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int p = Sh::pos(0, th, v);
            r[v] = {T(0), T(0)};
            if ((Z && v >= 8) || row >= nrows || 2 * p >= ncols) continue;
            const std::size_t i = (size_t)row * x_pitch + 2 * p;
            const Pair<T> q = *reinterpret_cast<const Pair<T>*>(x + i);
            if constexpr (IN == 1) {
                const unsigned mm = *reinterpret_cast<const unsigned short*>(mask + i);
                r[v] = {(mm & 0xffu) ? T(q.a - gmean) : T(0), (mm >> 8) ? T(q.b - gmean) : T(0)};
            } else {
                r[v] = {q.a, q.b};
            }
        }
    } else {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int p = Sh::pos(0, th, v);   // packed index j: reals 2j, 2j+1
            if (Z && v >= 8) r[v] = {T(0), T(0)};
            else r[v] = {load(2 * p), load(2 * p + 1)};
        }
    }
    rb::forward<T, LOGL, Z>(r, lds, th, tw, 2);
    // digit-reversed registers -> natural order in LDS, then split the
    // packed spectrum: X[k] = (Z_k + conj Z_{L-k})/2 - i W^k (Z_k - conj Z_{L-k})/2
    __syncthreads();
    constexpr int Gl = Sh::G(Sh::m - 1), Nl = Sh::N(Sh::m - 1);
#pragma unroll
    for (int q = 0; q < Gl; ++q)
#pragma unroll
        for (int v = 0; v < Nl; ++v)
            lds[row_ix<T>(Sh::kidx(Sh::pos(Sh::m - 1, th * Gl + q, v)))] = r[q * Nl + v];
    __syncthreads();
    // store: rows of the workgroup fastest, so B consecutive rows of one kx
    // form one contiguous run of the transposed layout. Element e =
    // threadIdx.x + i*blockDim.x of the B*(L+1) outputs is row e % B,
    // frequency e / B; with blockDim.x = B*TPR that is row threadIdx.x % B and
    // frequency threadIdx.x / B + i*TPR, so no division per element
    const int nb = (nrows - blk * B < B) ? (nrows - blk * B) : B;
    const int bb = threadIdx.x % B, k0 = threadIdx.x / B;
    if (bb < nb) {
        const cx<T>* lb = lds_all + bb * LL;
        cx<T>* out = AT + blk * B + bb;
        constexpr int KI = L / TPR + 1;
#pragma unroll
        for (int i = 0; i < KI; ++i) {
            const int k = k0 + i * TPR;
            if (k > L) break;
            const cx<T> Zk = lb[row_ix<T>(k == L ? 0 : k)];
            const cx<T> Zc = conjg(lb[row_ix<T>(k == 0 ? 0 : L - k)]);
            const cx<T> E = scale(Zk + Zc, T(0.5));
            const cx<T> D = Zk - Zc;
            const cx<T> O = {D.im * T(0.5), -D.re * T(0.5)};
            // W^k = W^(k - k%64) W^(k%64): both factors from a small,
            // cache-hot subset of the table
            const cx<double> wd = tw[k & ~63] * tw[k & 63];
            const cx<T> X = E + cx<T>{static_cast<T>(wd.re), static_cast<T>(wd.im)} * O;
            out[(size_t)k * pitchT] = X;
        }
    }
}

// 16 contiguous values from a 64-byte (FP32) / 128-byte (FP64) aligned
// address as 16-byte vector loads
__device__ __forceinline__ void load16(const float* __restrict__ p, float* v) {
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float4 t = q[j];
        v[4 * j] = t.x;
        v[4 * j + 1] = t.y;
        v[4 * j + 2] = t.z;
        v[4 * j + 3] = t.w;
    }
}
__device__ __forceinline__ void load16(const double* __restrict__ p, double* v) {
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const double2 t = q[j];
        v[2 * j] = t.x;
        v[2 * j + 1] = t.y;
    }
}

// ── columns: forward . symbol . inverse (transposed in, row-major out) ─────
template <typename T, int SYM, int LOGN>
__device__ __forceinline__ T symbol_rb(int kx, int ky, int n, const T* __restrict__ ktab,
                                       int kp, double sc) {
    if constexpr (SYM == 1) {
        const int my = (ky < ((n + 1) >> 1)) ? ky : ky - n;
        const double m2 = (double)kx * kx + (double)my * my;
        return m2 > 0.0 ? static_cast<T>(sc / sqrt(m2)) : T(0);
    } else {
        const int kof = (ky < (n >> 1)) ? ky : ky - n;
        // |k| rounded once to float. Up to n = 4096, kx^2 + kof^2 <= n^2/2 is
        // an exact float and the correctly rounded FP32 root of it is that
        // rounding (checked for every integer below 2^24); beyond, FP64.
        float h;
        if constexpr (LOGN <= 12) h = sqrtf((float)(kx * kx + kof * kof));
        else h = (float)sqrt((double)kx * kx + (double)kof * kof);
        const float w = (kx == 0 && ky == 0) ? 0.0f : (float)sc * h;
        return static_cast<T>(w);
    }
}

// Free-space column passes that exchange through a real (T) LDS buffer:
// FP64 at n = 8192, where the halved LDS lets a second 512-thread workgroup
// be resident (measured 3.5% faster at 4096^2; at n = 4096 FP64 it was 15%
// slower and is not used). FP32 columns exchange complex values: their
// 8-byte exchanges fit two workgroups in LDS anyway, and with bank-conflict-
// free layouts the halved barrier count wins (FP32 column pass at 4096^2:
// 128 -> 122 us against the split exchanges).
template <typename T, int LOGN, int SYM>
constexpr bool kColsSplit = SYM == 0 && sizeof(T) == 8 && LOGN >= 13;

// Free-space column passes held to 128 VGPRs (4 waves per SIMD, two resident
// 512-thread workgroups), FP32 and FP64 at n = 8192; forcing 80 VGPRs for a
// third workgroup spilled and was ~12% slower.
template <typename T, int LOGN, int SYM>
constexpr bool kColsBound4 = SYM == 0 && (sizeof(T) == 4 || LOGN >= 13);

template <typename T, int LOGN, int SYM>
__global__ void __launch_bounds__(512, (kColsBound4<T, LOGN, SYM> ? 4 : 1))
k_rb_cols(const cx<T>* AT, int pitchT, cx<T>* BR, int pitch,
          int ncols, int rows_in, int rows_out, int W, const cx<double>* __restrict__ tw,
          const T* __restrict__ ktab, int kp, double sc, T* __restrict__ kout,
          double* __restrict__ part, int remap, RbBatch bt, const int* __restrict__ kmap,
          int out_t) {
    using Sh = rb::Shape<LOGN>;
    {
        const int li = blockIdx.y, gi = bt.first + li;
        if (bt.gate >= 0 && bt.gscal[(std::size_t)gi * bt.sstride + bt.gate] == 0.0) return;
        AT += li * bt.at_stride;
        BR += li * (out_t ? bt.at_stride : bt.br_stride);
    }
    constexpr int n = Sh::n;
    constexpr int LL = rb::lds_len(n);
    constexpr int Gl = Sh::G(Sh::m - 1), Nl = Sh::N(Sh::m - 1);
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[16];
    const int tile0 = remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    // tiles beyond the grid (the host folds a small last round, e.g. the
    // Nyquist column of an (Ns+1)-column half spectrum, into the first
    // blocks) are taken in a second pass; the pass condition is uniform
    for (int tile = tile0; tile * W < ncols; tile += gridDim.x) {
    // the lane's indices are made opaque per pass, so the address arithmetic
    // derived from them is recomputed in the pass rather than kept live
    // across it (which would cost occupancy)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    // columns interleaved across lanes: the W columns' outputs of one row
    // are adjacent within each store instruction (W-fold longer runs of
    // the strided row-major stores; -14% at n = 4096, FP32)
    const int c = tid % W, th = tid / W;
    // FP64 table symbol at n = 8192: the exchanges pass real and imaginary
    // parts separately through half the LDS (more resident workgroups)
    using LT = std::conditional_t<kColsSplit<T, LOGN, SYM>, T, cx<T>>;
    LT* lds = reinterpret_cast<LT*>(smem) + c * LL;
    const int kx = tile * W + c;
    const bool live = kx < ncols;

    // free-space symbol: rows_in, rows_out <= n/2, so the zero-padded upper
    // half is neither loaded nor transformed forward, nor produced inverse
    constexpr bool ZP = (SYM == 0);
    cx<T> r[16];
    const cx<T>* col = AT + (size_t)(live ? kx : 0) * pitchT;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int p = Sh::pos(0, th, v);
        if (ZP && v >= 8) r[v] = {T(0), T(0)};
        else r[v] = (live && p < rows_in) ? col[p] : cx<T>{T(0), T(0)};
    }
    // the kernel table in the plan's digit-reversed layout: the 16 values of
    // this thread are contiguous at kx*kp + 16*th (Gl*Nl = 16); loaded
    // inside the last forward exchange so the HBM latency is hidden
    T kv[16];
    // half-size table: mirror-paired threads share a chunk, the upper one
    // reads it reversed (kmap); without a map, one chunk per thread
    int km = th << 1;
    // FP32 table symbol: the inverse reuses the forward's stage-KS twiddle
    // powers (15 values more per thread, within the bounded columns' 128
    // VGPRs; the computed-symbol passes have no such bound and would lose a
    // resident workgroup)
    constexpr int KS = kColsBound4<T, LOGN, SYM> && std::is_same_v<T, float>
                           ? rb::keep_stage<LOGN>() : -1;
    rb::TwKeep keep;
    if constexpr (SYM == 0) {
        if (kmap) km = kmap[th];
        const T* kr = ktab + (size_t)(live ? kx : 0) * kp + 16 * (km >> 1);
        auto hook = [&]() { load16(kr, kv); };
        rb::forward<T, LOGN, ZP, decltype(hook), LT, KS>(r, lds, th, tw, 1, hook, &keep);
    } else {
        rb::forward<T, LOGN, ZP, rb::NoHook, LT, KS>(r, lds, th, tw, 1, rb::NoHook{}, &keep);
    }

    if constexpr (SYM == 3) {
        // kernel-spectrum build (src/fft_operator.cpp:46-68): realness check
        // over every mode, Re/n^2 stored for ky in [0, n/2] (even in ky)
        const double norm = 1.0 / ((double)n * (double)n);
        double mi = 0.0, mr = 0.0;
#pragma unroll
        for (int q = 0; q < Gl; ++q)
#pragma unroll
            for (int v = 0; v < Nl; ++v) {
                const int ky = Sh::kidx(Sh::pos(Sh::m - 1, th * Gl + q, v));
                const cx<T> z = r[q * Nl + v];
                if (live) {
                    mi = fmax(mi, fabs((double)z.im));
                    mr = fmax(mr, fabs((double)z.re));
                    if (ky <= (n >> 1)) kout[(size_t)kx * kp + ky] = static_cast<T>(norm * (double)z.re);
                }
            }
        mi = block_reduce<RMAX>(mi, red);
        if (threadIdx.x == 0) part[2 * blockIdx.x] = mi;
        mr = block_reduce<RMAX>(mr, red);
        if (threadIdx.x == 0) part[2 * blockIdx.x + 1] = mr;
        return;
    } else {
        if constexpr (SYM == 0) {
            const bool rev = (km & 1) != 0;
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const T k = rev ? kv[15 - i] : kv[i];
                r[i] = scale(r[i], live ? k : T(0));
            }
        } else {
#pragma unroll
            for (int q = 0; q < Gl; ++q)
#pragma unroll
                for (int v = 0; v < Nl; ++v) {
                    const int ky = Sh::kidx(Sh::pos(Sh::m - 1, th * Gl + q, v));
                    const T s = live ? symbol_rb<T, SYM, LOGN>(kx, ky, n, ktab, kp, sc) : T(0);
                    r[q * Nl + v] = scale(r[q * Nl + v], s);
                }
        }
        rb::inverse<T, LOGN, ZP, LT, KS>(r, lds, th, tw, 1, &keep);
        if (live && out_t) {
            // transposed, in place: the column's rows [0, rows_out) back into
            // its own AT column (every thread's loads of it were consumed
            // before the first exchange barrier); consecutive lanes hold
            // consecutive rows, so each store is one contiguous run
            cx<T>* oc = BR + (size_t)kx * pitchT;
#pragma unroll
            for (int v = 0; v < (ZP ? 8 : 16); ++v) {
                const int p = Sh::pos(0, th, v);
                if (p < rows_out) oc[p] = r[v];
            }
        } else if (live) {
#pragma unroll
            for (int v = 0; v < (ZP ? 8 : 16); ++v) {
                const int p = Sh::pos(0, th, v);
                if (p < rows_out) BR[(size_t)p * pitch + kx] = r[v];
            }
        }
    }
    }
}

// ── rows: half spectrum (row-major or transposed) -> real ──────────────────
// H: only the reals y[0 .. L) of n = 2L are wanted (free-space rows), so the
// last radix-16 stage computes half of its outputs.
template <typename T, int LOGL, int OUT, bool H>
__global__ void __launch_bounds__(1024)
k_rb_rows_inv(const cx<T>* __restrict__ BR, int pitch, int nrows, int B,
              const cx<double>* __restrict__ tw, T* __restrict__ y, int y_pitch, int ncols,
              const std::uint8_t* __restrict__ mask, double* __restrict__ part, int remap,
              RbBatch bt, int in_t) {
    using Sh = rb::Shape<LOGL>;
    {
        const int li = blockIdx.y, gi = bt.first + li;
        if (bt.gate >= 0 && bt.gscal[(std::size_t)gi * bt.sstride + bt.gate] == 0.0) return;
        BR += li * (in_t ? bt.at_stride : bt.br_stride);
        y += li * bt.y_stride;
        if (mask) mask += li * (bt.m_stride >= 0 ? bt.m_stride : bt.y_stride);
        if (part) part += (std::size_t)gi * bt.pstride;
    }
    constexpr int L = Sh::n, TPR = Sh::threads;
    const int LL = row_stride<T>(L, B);
    constexpr int Gl = Sh::G(Sh::m - 1), Nl = Sh::N(Sh::m - 1);
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ double red[16];
    cx<T>* lds_all = reinterpret_cast<cx<T>*>(smem);
    const int blk = remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int b = threadIdx.x / TPR, th = threadIdx.x - b * TPR;
    const int row = blk * B + b;
    cx<T>* lds = lds_all + b * LL;
    cx<T> r[16];
    // coalesced load of the L+1 half-spectrum values of each row
    // (e = threadIdx.x + j*blockDim.x, blockDim.x = B L/16): every load
    // is issued before the first is waited for
    constexpr int J = 17;
    const int lim = B * (L + 1);
    cx<T> xv[J];
    if (in_t) {
        // transposed input BR[k*pitch + row] (pitch = the column pitch): the
        // workgroup's B rows fastest, so each k is one run of B values.
        // Element e is row e % B, frequency e / B; with blockDim.x = B*TPR
        // that is row threadIdx.x % B, frequency threadIdx.x / B + j*TPR
        static_assert(J == L / TPR + 1, "one load per frequency step");
        const int bb = threadIdx.x % B, k0 = threadIdx.x / B;
        const int rr = blk * B + bb;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int k = k0 + j * TPR;
            xv[j] = (k <= L && rr < nrows) ? BR[(size_t)k * pitch + rr] : cx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int k = k0 + j * TPR;
            if (k <= L) lds_all[bb * LL + row_ix<T>(k)] = xv[j];
        }
    } else {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int e = threadIdx.x + j * (int)blockDim.x;
            const int bb = e / (L + 1), k = e - bb * (L + 1);
            const int rr = blk * B + bb;
            xv[j] = (e < lim && rr < nrows) ? BR[(size_t)rr * pitch + k] : cx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int e = threadIdx.x + j * (int)blockDim.x;
            const int bb = e / (L + 1), k = e - bb * (L + 1);
            if (e < lim) lds_all[bb * LL + row_ix<T>(k)] = xv[j];
        }
    }
    __syncthreads();
    // Z_k = (X_k + conj X_{L-k}) + i (X_k - conj X_{L-k}) exp(+2 pi i k/n),
    // gathered straight into the digit-reversed grouping of the inverse.
    // The last digit v of a group's positions is the top digit of k:
    // k = kg + v L/Nl, so W^k = W^kg W_32^(16 v/Nl) (one table read per group).
#pragma unroll
    for (int q = 0; q < Gl; ++q) {
        const int kg = Sh::kidx(Sh::pos(Sh::m - 1, th * Gl + q, 0));
        const cx<double> wg = tw[kg];
#pragma unroll
        for (int v = 0; v < Nl; ++v) {
            const int k = kg + v * (L / Nl);
            cx<T> Xk = lds[row_ix<T>(k)], Xc = lds[row_ix<T>(L - k)];
            if (k == 0) {
                Xk.im = T(0);
                Xc.im = T(0);
            }
            Xc = conjg(Xc);
            const cx<T> S = Xk + Xc;
            constexpr int js = 16 / Nl;
            const cx<double> wd =
                v == 0 ? wg : wg * cx<double>{rb::kW32re[(v * js) & 15], rb::kW32im[(v * js) & 15]};
            const cx<T> D = (Xk - Xc) * cx<T>{static_cast<T>(wd.re), static_cast<T>(-wd.im)};
            r[q * Nl + v] = {S.re - D.im, S.im + D.re};
        }
    }
    // every wave has gathered from the whole row before the transform's
    // first exchange writes (it may be ordered within each wave only)
    __syncthreads();
    rb::inverse<T, LOGL, H>(r, lds, th, tw, 2);
    double acc = 0.0;
    const bool pairs = (OUT == 0 || OUT == 1) && ((y_pitch | ncols) & 1) == 0 &&
                       pair_aligned(y, sizeof(T)) &&
                       (OUT == 0 || (reinterpret_cast<std::uintptr_t>(mask) & 1u) == 0);
    if (row < nrows && pairs) {
        // plain or masked output on an even pitch and width: pair stores
        // (and the two mask bytes in one load)
#pragma unroll
        for (int v = 0; v < (H ? 8 : 16); ++v) {
            const int p = Sh::pos(0, th, v);
            if (2 * p >= ncols) continue;
            const std::size_t i = (size_t)row * y_pitch + 2 * p;
            T a = r[v].re, b = r[v].im;
            if constexpr (OUT == 1) {
                const unsigned mm = *reinterpret_cast<const unsigned short*>(mask + i);
                if (mm & 0xffu) acc += (double)a;
                else a = T(0);
                if (mm >> 8) acc += (double)b;
                else b = T(0);
            }
            *reinterpret_cast<Pair<T>*>(y + i) = Pair<T>{a, b};
        }
    } else if (row < nrows) {
#pragma unroll
        for (int v = 0; v < (H ? 8 : 16); ++v) {
            const int p = Sh::pos(0, th, v);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int col = 2 * p + h;
                if (col >= ncols) continue;
                T val = h ? r[v].im : r[v].re;
                const size_t i = (size_t)row * y_pitch + col;
                if constexpr (OUT == 1) {
                    if (mask[i]) acc += (double)val;
                    else val = T(0);
                }
                y[i] = val;
            }
        }
    }
    if constexpr (OUT == 1) {
        acc = block_reduce<RSUM>(acc, red);
        if (threadIdx.x == 0) part[(std::size_t)blockIdx.x * kPartVals] = acc;
    }
}

void allow_lds_rb(const void* fn, std::size_t bytes) {
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

template <typename T, int LOGL>
void launch_rows_fwd(const RbRowsFwd<T>& a, hipStream_t s) {
    constexpr int L = 1 << LOGL;
    // every operand the mode reads must be present
    const bool ok = a.mode == 0   ? a.x != nullptr
                    : a.mode == 1 ? (a.x && a.mask && a.scal)
                    : a.mode <= 4 ? a.table != nullptr
                                  : false;
    if (!ok) throw std::invalid_argument("row transform: missing operand for its input mode");
    const int block = a.B * (L / 16);
    const dim3 grid((a.nrows + a.B - 1) / a.B, a.bt.count);
    const std::size_t lds = sizeof(cx<T>) * (std::size_t)a.B * row_stride<T>(L, a.B);
    const void* fn;
#define RF_ARGS a.x, a.x_pitch, a.ncols, a.nrows, a.AT, a.pitchT, a.B, a.tw, a.mask, a.scal, \
        a.gmean_idx, a.table, a.Ns_embed, a.remap, a.bt
    if (a.mode == 0 && a.ncols <= L) {
        fn = (const void*)k_rb_rows_fwd<T, LOGL, 0, true>;
        allow_lds_rb(fn, lds);
        k_rb_rows_fwd<T, LOGL, 0, true><<<grid, block, lds, s>>>(RF_ARGS);
    } else if (a.mode == 0) {
        fn = (const void*)k_rb_rows_fwd<T, LOGL, 0, false>;
        allow_lds_rb(fn, lds);
        k_rb_rows_fwd<T, LOGL, 0, false><<<grid, block, lds, s>>>(RF_ARGS);
    } else if (a.mode == 1) {
        fn = (const void*)k_rb_rows_fwd<T, LOGL, 1, false>;
        allow_lds_rb(fn, lds);
        k_rb_rows_fwd<T, LOGL, 1, false><<<grid, block, lds, s>>>(RF_ARGS);
    } else if (a.mode == 2) {
        fn = (const void*)k_rb_rows_fwd<T, LOGL, 2, false>;
        allow_lds_rb(fn, lds);
        k_rb_rows_fwd<T, LOGL, 2, false><<<grid, block, lds, s>>>(RF_ARGS);
    } else if (a.mode == 3) {
        fn = (const void*)k_rb_rows_fwd<T, LOGL, 3, false>;
        allow_lds_rb(fn, lds);
        k_rb_rows_fwd<T, LOGL, 3, false><<<grid, block, lds, s>>>(RF_ARGS);
    } else {
        fn = (const void*)k_rb_rows_fwd<T, LOGL, 4, false>;
        allow_lds_rb(fn, lds);
        k_rb_rows_fwd<T, LOGL, 4, false><<<grid, block, lds, s>>>(RF_ARGS);
    }
#undef RF_ARGS
    ASPHER_LAUNCH_CHECK();
}

// Concurrent workgroups of a kernel on the device (occupancy x CUs), cached.
int resident_blocks(const void* fn, int block, std::size_t lds) {
    static std::mutex mu;
    static std::map<std::tuple<int, const void*, int, std::size_t>, int> done;
    int dev = 0;
    ASPHER_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_tuple(dev, fn, block, lds);   // per device: CU counts differ
    auto it = done.find(key);
    if (it != done.end()) return it->second;
    int per_cu = 0, cus = 0;
    ASPHER_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds));
    ASPHER_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int r = std::max(1, per_cu) * std::max(1, cus);
    done[key] = r;
    return r;
}

// Grid of the column pass: when the tiles exceed one round of resident
// workgroups by a small remainder (at most 1/16 of a round), the remainder
// is folded into the first blocks as a second pass instead of a last round
// that would run almost alone.
int cols_grid(int tiles, int resident) {
    if (tiles <= resident) return tiles;
    const int rem = tiles % resident;
    if (rem == 0 || rem > resident / 16) return tiles;
    return tiles - rem;
}

template <typename T, int LOGN>
void launch_cols(const RbCols<T>& a, hipStream_t s) {
    constexpr int n = 1 << LOGN;
    const int block = a.W * (n / 16);
    const int tiles = (a.ncols + a.W - 1) / a.W;
    const std::size_t elt = (a.sym == 0 && kColsSplit<T, LOGN, 0>) ? sizeof(T) : sizeof(cx<T>);
    const std::size_t lds = elt * (std::size_t)a.W * rb::lds_len(n);
    int gx = tiles;
    if (a.sym != 3 && a.bt.count == 1) {
        const void* f = a.sym == 0 ? (const void*)k_rb_cols<T, LOGN, 0>
                        : a.sym == 1 ? (const void*)k_rb_cols<T, LOGN, 1>
                                     : (const void*)k_rb_cols<T, LOGN, 2>;
        allow_lds_rb(f, lds);
        gx = cols_grid(tiles, resident_blocks(f, block, lds));
    }
    const dim3 grid(gx, a.bt.count);
#define RC_ARGS a.AT, a.pitchT, a.BR, a.pitch, a.ncols, a.rows_in, a.rows_out, a.W, a.tw, a.ktab, \
        a.kp, a.sc, a.kout, a.part, a.remap, a.bt, a.kmap, a.out_t
#define RC_CASE(S)                                                     \
    case S: {                                                          \
        const void* fn = (const void*)k_rb_cols<T, LOGN, S>;           \
        allow_lds_rb(fn, lds);                                         \
        k_rb_cols<T, LOGN, S><<<grid, block, lds, s>>>(RC_ARGS);       \
        break;                                                         \
    }
    if (a.sym == 0 && (a.rows_in > n / 2 || a.rows_out > n / 2))
        throw std::invalid_argument("table symbol needs rows_in, rows_out <= n/2");
    if (a.out_t && (a.BR != a.AT || a.sym == 3 || a.rows_out > a.pitchT))
        throw std::invalid_argument("transposed column output is in place (BR == AT)");
    switch (a.sym) {
        RC_CASE(0)
        RC_CASE(1)
        RC_CASE(2)
        RC_CASE(3)
        default: throw std::invalid_argument("unknown symbol");
    }
#undef RC_CASE
#undef RC_ARGS
    ASPHER_LAUNCH_CHECK();
}

template <typename T, int LOGL>
void launch_rows_inv(const RbRowsInv<T>& a, hipStream_t s) {
    constexpr int L = 1 << LOGL;
    const bool ok = a.y && (a.mode == 0 ? true : a.mode == 1 ? (a.mask && a.part) : false);
    if (!ok) throw std::invalid_argument("row transform: missing operand for its output mode");
    const int block = a.B * (L / 16);
    const dim3 grid((a.nrows + a.B - 1) / a.B, a.bt.count);
    const std::size_t lds = sizeof(cx<T>) * (std::size_t)a.B * row_stride<T>(L, a.B);
#define RI_ARGS a.BR, a.pitch, a.nrows, a.B, a.tw, a.y, a.y_pitch, a.ncols, a.mask, a.part, a.remap, \
        a.bt, a.in_t
    if (a.mode == 0 && a.ncols <= L) {
        const void* fn = (const void*)k_rb_rows_inv<T, LOGL, 0, true>;
        allow_lds_rb(fn, lds);
        k_rb_rows_inv<T, LOGL, 0, true><<<grid, block, lds, s>>>(RI_ARGS);
    } else if (a.mode == 0) {
        const void* fn = (const void*)k_rb_rows_inv<T, LOGL, 0, false>;
        allow_lds_rb(fn, lds);
        k_rb_rows_inv<T, LOGL, 0, false><<<grid, block, lds, s>>>(RI_ARGS);
    } else {
        const void* fn = (const void*)k_rb_rows_inv<T, LOGL, 1, false>;
        allow_lds_rb(fn, lds);
        k_rb_rows_inv<T, LOGL, 1, false><<<grid, block, lds, s>>>(RI_ARGS);
    }
#undef RI_ARGS
    ASPHER_LAUNCH_CHECK();
}


// ── two-component (tangential) column pass ───────────────────────────────
// One workgroup per column kx: threads [0, TPC) transform the x component
// (plan instance 0), threads [TPC, 2 TPC) the y component (instance 1). After
// the forward transforms both halves swap their spectra through LDS (the two
// components hold the same digit-reversed positions), mix them by the real
// symmetric 2x2 symbol and run their inverse transforms.
template <typename T, int LOGN, int SYM>
__global__ void __launch_bounds__(512)
k_rb_cols2(const cx<T>* __restrict__ AT, int pitchT, std::int64_t at_stride,
           cx<T>* __restrict__ BR, int pitch, std::int64_t br_stride, int ncols, int rows_in,
           int rows_out, const cx<double>* __restrict__ tw, const T* __restrict__ kxx,
           const T* __restrict__ kyy, const T* __restrict__ kxy, int kp, double gamma,
           double norm, int remap) {
    using Sh = rb::Shape<LOGN>;
    constexpr int n = Sh::n, TPC = Sh::threads;
    constexpr int LL = rb::lds_len(n);
    constexpr int Gl = Sh::G(Sh::m - 1), Nl = Sh::N(Sh::m - 1);
    constexpr bool ZP = (SYM == 5);
    extern __shared__ __align__(16) unsigned char smem[];
    const int kx = remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int c = threadIdx.x / TPC, th = threadIdx.x - c * TPC;
    const bool live = kx < ncols;
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem) + c * LL;
    const cx<T>* ldo = reinterpret_cast<cx<T>*>(smem) + (1 - c) * LL;

    cx<T> r[16];
    const cx<T>* col = AT + c * at_stride + (size_t)(live ? kx : 0) * pitchT;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int p = Sh::pos(0, th, v);
        if (ZP && v >= 8) r[v] = {T(0), T(0)};
        else r[v] = (live && p < rows_in) ? col[p] : cx<T>{T(0), T(0)};
    }
    rb::forward<T, LOGN, ZP>(r, lds, th, tw, 1);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; ++i) lds[rb::lds_ix(16 * th + i)] = r[i];
    __syncthreads();
    cx<T> o[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) o[i] = ldo[rb::lds_ix(16 * th + i)];
    if constexpr (SYM == 5) {
        const std::size_t at = (size_t)(live ? kx : 0) * kp + 16 * th;
        T kd[16], kc[16];
        load16((c == 0 ? kxx : kyy) + at, kd);
        load16(kxy + at, kc);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const T a = live ? kd[i] : T(0), b = live ? kc[i] : T(0);
            r[i] = {a * r[i].re + b * o[i].re, a * r[i].im + b * o[i].im};
        }
    } else {
#pragma unroll
        for (int q = 0; q < Gl; ++q)
#pragma unroll
            for (int v = 0; v < Nl; ++v) {
                const int i = q * Nl + v;
                const int ky = Sh::kidx(Sh::pos(Sh::m - 1, th * Gl + q, v));
                const double fx = (double)kx;
                const double fy = (double)(ky < (n >> 1) ? ky : ky - n);
                const double k2 = fx * fx + fy * fy;
                double wd = 0.0, wc = 0.0;
                if (live && k2 > 0.0) {
                    const double kk = sqrt(k2);
                    const double f = c == 0 ? fx : fy;
                    wd = norm * kk * (1.0 + gamma * f * f / k2);
                    wc = (kx == (n >> 1) || ky == (n >> 1)) ? 0.0
                                                            : norm * kk * gamma * fx * fy / k2;
                }
                const T a = static_cast<T>(wd), b = static_cast<T>(wc);
                r[i] = {a * r[i].re + b * o[i].re, a * r[i].im + b * o[i].im};
            }
    }
    // the other component has read this one's spectrum before the inverse's
    // first exchange overwrites it (that exchange may be ordered within
    // each wave only)
    __syncthreads();
    rb::inverse<T, LOGN, ZP>(r, lds, th, tw, 1);
    if (live) {
        cx<T>* out = BR + c * br_stride;
#pragma unroll
        for (int v = 0; v < (ZP ? 8 : 16); ++v) {
            const int p = Sh::pos(0, th, v);
            if (p < rows_out) out[(size_t)p * pitch + kx] = r[v];
        }
    }
}

template <typename T, int LOGN>
void launch_cols2(const RbCols2<T>& a, hipStream_t s) {
    constexpr int n = 1 << LOGN;
    static_assert(2 * (n / 16) <= 512, "two-component columns up to n = 4096");
    const int block = 2 * (n / 16);
    const int grid = a.ncols;
    const std::size_t lds = sizeof(cx<T>) * 2 * (std::size_t)rb::lds_len(n);
    if (a.sym == 5) {
        if (a.rows_in > n / 2 || a.rows_out > n / 2)
            throw std::invalid_argument("tangential columns need rows <= n/2");
        const void* fn = (const void*)k_rb_cols2<T, LOGN, 5>;
        allow_lds_rb(fn, lds);
        k_rb_cols2<T, LOGN, 5><<<grid, block, lds, s>>>(
            a.AT, a.pitchT, a.at_stride, a.BR, a.pitch, a.br_stride, a.ncols, a.rows_in,
            a.rows_out, a.tw, a.kxx, a.kyy, a.kxy, a.kp, a.gamma, a.norm, a.remap);
    } else {
        const void* fn = (const void*)k_rb_cols2<T, LOGN, 6>;
        allow_lds_rb(fn, lds);
        k_rb_cols2<T, LOGN, 6><<<grid, block, lds, s>>>(
            a.AT, a.pitchT, a.at_stride, a.BR, a.pitch, a.br_stride, a.ncols, a.rows_in,
            a.rows_out, a.tw, a.kxx, a.kyy, a.kxy, a.kp, a.gamma, a.norm, a.remap);
    }
    ASPHER_LAUNCH_CHECK();
}

// ── free-space kernel spectrum in the column kernel's register order ──────
// kout[kx*n + pos] = K(kx, ky) with ky = kidx(pos) the frequency that the
// forward column transform leaves at digit-reversed position pos, folded by
// evenness to the (n/2+1)-wide unique table kin[kx*kp + min(ky, n-ky)].
template <typename T, int LOGN>
__global__ void __launch_bounds__(256)
k_rb_permute_table(const T* __restrict__ kin, int kp, int nh, T* __restrict__ kout, bool odd) {
    using Sh = rb::Shape<LOGN>;
    constexpr int n = Sh::n;
    const std::int64_t total = (std::int64_t)nh * n;
    for (std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(i / n), pos = (int)(i - (std::int64_t)kx * n);
        const int ky = Sh::kidx(pos);
        const int kyp = (ky <= (n >> 1)) ? ky : n - ky;
        const T v = kin[(std::size_t)kx * kp + kyp];
        kout[i] = (odd && ky > (n >> 1)) ? T(-v) : v;
    }
}

template <typename T, int LOGN>
void launch_permute(const T* kin, int kp, int nh, T* kout, hipStream_t s, bool odd) {
    const std::int64_t total = (std::int64_t)nh * (1 << LOGN);
    const int grid = (int)std::min<std::int64_t>((total + 255) / 256, 65536);
    k_rb_permute_table<T, LOGN><<<grid, 256, 0, s>>>(kin, kp, nh, kout, odd);
    ASPHER_LAUNCH_CHECK();
}

template <typename T, int LOGN>
__global__ void __launch_bounds__(256)
k_rb_permute_table_half(const T* __restrict__ kin, int kp, int nh, const int* __restrict__ kmap,
                        int kph, T* __restrict__ kout) {
    using Sh = rb::Shape<LOGN>;
    constexpr int n = Sh::n, Tn = Sh::threads;
    constexpr int Gl = Sh::G(Sh::m - 1), Nl = Sh::N(Sh::m - 1);
    const std::int64_t total = (std::int64_t)nh * Tn * 16;
    for (std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (std::int64_t)gridDim.x * blockDim.x) {
        const int kx = (int)(i / (Tn * 16)), e = (int)(i - (std::int64_t)kx * (Tn * 16));
        const int th = e >> 4, j = e & 15;
        const int km = kmap[th];
        if (km & 1) continue;   // this chunk is written by its forward reader
        const int q = j / Nl, v = j - q * Nl;
        const int ky = Sh::kidx(Sh::pos(Sh::m - 1, th * Gl + q, v));
        const int kyp = (ky <= (n >> 1)) ? ky : n - ky;
        kout[(std::size_t)kx * kph + 16 * (km >> 1) + j] = kin[(std::size_t)kx * kp + kyp];
    }
}

template <typename T, int LOGN>
void launch_permute_half(const T* kin, int kp, int nh, const int* kmap, int kph, T* kout,
                         hipStream_t s) {
    const std::int64_t total = (std::int64_t)nh * (1 << LOGN);
    const int grid = (int)std::min<std::int64_t>((total + 255) / 256, 65536);
    k_rb_permute_table_half<T, LOGN><<<grid, 256, 0, s>>>(kin, kp, nh, kmap, kph, kout);
    ASPHER_LAUNCH_CHECK();
}

// ── rough-surface generator: spectrum synthesis + inverse column pass ─────
__device__ __forceinline__ double hash_uniform(unsigned long long seed, unsigned long long idx) {
    // SplitMix64 output number idx of a generator seeded with `seed`
    unsigned long long z = seed + (idx + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * 0x1.0p-53;
}

template <int LOGN>
__global__ void __launch_bounds__(512)
k_rb_gen_cols(cx<double>* __restrict__ BR, int pitch, double hurst, double k_low,
              double k_high, double scale, unsigned long long seed,
              const cx<double>* __restrict__ tw, int W, int remap) {
    using Sh = rb::Shape<LOGN>;
    constexpr int n = Sh::n, TPC = Sh::threads, nh = n / 2 + 1;
    constexpr int LL = rb::lds_len(n);
    constexpr int Gl = Sh::G(Sh::m - 1), Nl = Sh::N(Sh::m - 1);
    extern __shared__ __align__(16) unsigned char smem[];
    const int tile = remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int c = threadIdx.x / TPC, th = threadIdx.x - c * TPC;
    const int kx = tile * W + c;
    const bool live = kx < nh;
    cx<double>* lds = reinterpret_cast<cx<double>*>(smem) + c * LL;
    const double two_pi = 6.283185307179586476925286766559;
    cx<double> r[16];
#pragma unroll
    for (int q = 0; q < Gl; ++q)
#pragma unroll
        for (int v = 0; v < Nl; ++v) {
            const int ky = Sh::kidx(Sh::pos(Sh::m - 1, th * Gl + q, v));
            cx<double> h = {0.0, 0.0};
            if (live) {
                const int mx = kx < n / 2 ? kx : kx - n;   // fftfreq order
                const int my = ky < n / 2 ? ky : ky - n;
                const double mag = sqrt((double)mx * mx + (double)my * my);
                if (mag >= k_low && mag <= k_high && mag > 0.0) {
                    const double A = 0.5 * scale * pow(mag, -(1.0 + hurst));
                    const int kxn = (n - kx) % n, kyn = (n - ky) % n;
                    const double p1 = two_pi * hash_uniform(seed, (unsigned long long)ky * n + kx);
                    const double p2 = two_pi * hash_uniform(seed, (unsigned long long)kyn * n + kxn);
                    double s1, c1, s2, c2;
                    sincos(p1, &s1, &c1);
                    sincos(p2, &s2, &c2);
                    h = {A * (c1 + c2), A * (s1 - s2)};
                }
            }
            r[q * Nl + v] = h;
        }
    rb::inverse<double, LOGN>(r, lds, th, tw, 1);
    if (live) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int p = Sh::pos(0, th, v);
            BR[(size_t)p * pitch + kx] = r[v];
        }
    }
}

template <int LOGN>
void launch_gen(const RbGen& a, hipStream_t s) {
    constexpr int n = 1 << LOGN;
    const int block = a.W * (n / 16);
    const int grid = (n / 2 + 1 + a.W - 1) / a.W;
    const std::size_t lds = sizeof(cx<double>) * (std::size_t)a.W * rb::lds_len(n);
    const void* fn = (const void*)k_rb_gen_cols<LOGN>;
    allow_lds_rb(fn, lds);
    k_rb_gen_cols<LOGN><<<grid, block, lds, s>>>(a.BR, a.pitch, a.hurst, a.k_low, a.k_high,
                                                 a.scale, a.seed, a.tw, a.W, a.remap);
    ASPHER_LAUNCH_CHECK();
}
} // namespace

template <typename T>
void rb_rows_fwd(int logL, const RbRowsFwd<T>& a, hipStream_t s) {
    switch (logL) {
        case 4: launch_rows_fwd<T, 4>(a, s); break;
        case 5: launch_rows_fwd<T, 5>(a, s); break;
        case 6: launch_rows_fwd<T, 6>(a, s); break;
        case 7: launch_rows_fwd<T, 7>(a, s); break;
        case 8: launch_rows_fwd<T, 8>(a, s); break;
        case 9: launch_rows_fwd<T, 9>(a, s); break;
        case 10: launch_rows_fwd<T, 10>(a, s); break;
        case 11: launch_rows_fwd<T, 11>(a, s); break;
        case 12: launch_rows_fwd<T, 12>(a, s); break;
        case 13: launch_rows_fwd<T, 13>(a, s); break;
        default: throw std::invalid_argument("row transform length out of range");
    }
}

template <typename T>
void rb_cols(int logN, const RbCols<T>& a, hipStream_t s) {
    switch (logN) {
        case 5: launch_cols<T, 5>(a, s); break;
        case 6: launch_cols<T, 6>(a, s); break;
        case 7: launch_cols<T, 7>(a, s); break;
        case 8: launch_cols<T, 8>(a, s); break;
        case 9: launch_cols<T, 9>(a, s); break;
        case 10: launch_cols<T, 10>(a, s); break;
        case 11: launch_cols<T, 11>(a, s); break;
        case 12: launch_cols<T, 12>(a, s); break;
        case 13: launch_cols<T, 13>(a, s); break;
        default: throw std::invalid_argument("column transform length out of range");
    }
}

template <typename T>
void rb_rows_inv(int logL, const RbRowsInv<T>& a, hipStream_t s) {
    switch (logL) {
        case 4: launch_rows_inv<T, 4>(a, s); break;
        case 5: launch_rows_inv<T, 5>(a, s); break;
        case 6: launch_rows_inv<T, 6>(a, s); break;
        case 7: launch_rows_inv<T, 7>(a, s); break;
        case 8: launch_rows_inv<T, 8>(a, s); break;
        case 9: launch_rows_inv<T, 9>(a, s); break;
        case 10: launch_rows_inv<T, 10>(a, s); break;
        case 11: launch_rows_inv<T, 11>(a, s); break;
        case 12: launch_rows_inv<T, 12>(a, s); break;
        case 13: launch_rows_inv<T, 13>(a, s); break;
        default: throw std::invalid_argument("row transform length out of range");
    }
}

template <typename T>
void rb_cols2(int logN, const RbCols2<T>& a, hipStream_t s) {
    switch (logN) {
        case 5: launch_cols2<T, 5>(a, s); break;
        case 6: launch_cols2<T, 6>(a, s); break;
        case 7: launch_cols2<T, 7>(a, s); break;
        case 8: launch_cols2<T, 8>(a, s); break;
        case 9: launch_cols2<T, 9>(a, s); break;
        case 10: launch_cols2<T, 10>(a, s); break;
        case 11: launch_cols2<T, 11>(a, s); break;
        case 12: launch_cols2<T, 12>(a, s); break;
        default: throw std::invalid_argument("two-component transform length out of range");
    }
}
template void rb_cols2<float>(int, const RbCols2<float>&, hipStream_t);
template void rb_cols2<double>(int, const RbCols2<double>&, hipStream_t);

void rb_gen_cols(int logN, const RbGen& a, hipStream_t s) {
    switch (logN) {
        case 5: launch_gen<5>(a, s); break;
        case 6: launch_gen<6>(a, s); break;
        case 7: launch_gen<7>(a, s); break;
        case 8: launch_gen<8>(a, s); break;
        case 9: launch_gen<9>(a, s); break;
        case 10: launch_gen<10>(a, s); break;
        case 11: launch_gen<11>(a, s); break;
        case 12: launch_gen<12>(a, s); break;
        case 13: launch_gen<13>(a, s); break;
        default: throw std::invalid_argument("surface size out of range");
    }
}

template <typename T>
void rb_permute_table(int logN, const T* kin, int kp, int nh, T* kout, hipStream_t s,
                      bool odd) {
    switch (logN) {
        case 5: launch_permute<T, 5>(kin, kp, nh, kout, s, odd); break;
        case 6: launch_permute<T, 6>(kin, kp, nh, kout, s, odd); break;
        case 7: launch_permute<T, 7>(kin, kp, nh, kout, s, odd); break;
        case 8: launch_permute<T, 8>(kin, kp, nh, kout, s, odd); break;
        case 9: launch_permute<T, 9>(kin, kp, nh, kout, s, odd); break;
        case 10: launch_permute<T, 10>(kin, kp, nh, kout, s, odd); break;
        case 11: launch_permute<T, 11>(kin, kp, nh, kout, s, odd); break;
        case 12: launch_permute<T, 12>(kin, kp, nh, kout, s, odd); break;
        case 13: launch_permute<T, 13>(kin, kp, nh, kout, s, odd); break;
        default: throw std::invalid_argument("column transform length out of range");
    }
}
template <typename T>
void rb_permute_table_half(int logN, const T* kin, int kp, int nh, const int* kmap, int kph,
                           T* kout, hipStream_t s) {
    if (!kmap) throw std::invalid_argument("permute_table_half: no mirror map");
    switch (logN) {
        case 5: launch_permute_half<T, 5>(kin, kp, nh, kmap, kph, kout, s); break;
        case 6: launch_permute_half<T, 6>(kin, kp, nh, kmap, kph, kout, s); break;
        case 7: launch_permute_half<T, 7>(kin, kp, nh, kmap, kph, kout, s); break;
        case 8: launch_permute_half<T, 8>(kin, kp, nh, kmap, kph, kout, s); break;
        case 9: launch_permute_half<T, 9>(kin, kp, nh, kmap, kph, kout, s); break;
        case 10: launch_permute_half<T, 10>(kin, kp, nh, kmap, kph, kout, s); break;
        case 11: launch_permute_half<T, 11>(kin, kp, nh, kmap, kph, kout, s); break;
        case 12: launch_permute_half<T, 12>(kin, kp, nh, kmap, kph, kout, s); break;
        case 13: launch_permute_half<T, 13>(kin, kp, nh, kmap, kph, kout, s); break;
        default: throw std::invalid_argument("column transform length out of range");
    }
}
template void rb_permute_table_half<float>(int, const float*, int, int, const int*, int, float*,
                                           hipStream_t);
template void rb_permute_table_half<double>(int, const double*, int, int, const int*, int,
                                            double*, hipStream_t);
template void rb_permute_table<float>(int, const float*, int, int, float*, hipStream_t, bool);
template void rb_permute_table<double>(int, const double*, int, int, double*, hipStream_t,
                                       bool);

template void rb_rows_fwd<float>(int, const RbRowsFwd<float>&, hipStream_t);
template void rb_rows_fwd<double>(int, const RbRowsFwd<double>&, hipStream_t);
template void rb_cols<float>(int, const RbCols<float>&, hipStream_t);
template void rb_cols<double>(int, const RbCols<double>&, hipStream_t);
template void rb_rows_inv<float>(int, const RbRowsInv<float>&, hipStream_t);
template void rb_rows_inv<double>(int, const RbRowsInv<double>&, hipStream_t);

} // namespace aspher
