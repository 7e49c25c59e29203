// Register-blocked power-of-two FFT for gfx950.
//
// A length n = 2^LOGN transform is factored n = N_0 N_1 ... N_{m-1} with
// N_s = 16 for the leading stages and one trailing 2/4/8 stage. Each thread
// owns 16 elements; a stage is a batch of in-register DFTs of size N_s, and
// the only LDS traffic is one exchange between consecutive stages.
//
// Position of an element: pos = sum_j d_j S_j with S_j = prod_{i>j} N_i. The
// group of stage s (the N_s elements one small DFT combines) is every digit
// but digit s: pos = high*N_s*S_s + v*S_s + low, g = high*S_s + low; thread
// th handles groups g = th*G_s + q, G_s = 16/N_s.
//
// Forward (decimation in frequency): stage s DFT over digit s, then for
// s < m-1 multiply by W_{N_s S_s}^{v*low}. Natural-order input, output in
// digit-reversed order: the value at pos is X[k], k = sum_j d_j P_j with
// P_j = prod_{i<j} N_i. Inverse (decimation in time): stages m-1 .. 0 with
// the conjugate twiddle applied before each stage s < m-1; it consumes the
// digit-reversed order and returns natural order. Forward stage m-1 and
// inverse stage m-1 share their grouping, so a spectral multiply between
// them needs no exchange.
#pragma once

#include "common.hpp"

#include <type_traits>

namespace aspher {
namespace rb {

// cos(pi j / 16), j = 0..8
__device__ constexpr double kCos16[9] = {
    1.0,
    0.98078528040323044912618223613423903697,
    0.92387953251128675612818318939678828683,
    0.83146961230254523707878837761790575673,
    0.70710678118654752440084436210484903928,
    0.55557023301960222474283081394853287438,
    0.38268343236508977172845998403039886676,
    0.19509032201612826784828486847702224093,
    0.0};

__device__ __forceinline__ constexpr double cospi16(int j) {
    j = ((j % 32) + 32) % 32;
    if (j > 16) j = 32 - j;
    return (j <= 8) ? kCos16[j] : -kCos16[16 - j];
}
__device__ __forceinline__ constexpr double sinpi16(int j) { return cospi16(j - 8); }

template <typename T, bool INV>
__device__ __forceinline__ cx<T> rot90(cx<T> a) {
    return INV ? cx<T>{-a.im, a.re} : cx<T>{a.im, -a.re};
}

// a * W_32^j (forward: exp(-2 pi i j/32); inverse: conjugate). j must fold
// to a constant after unrolling.
template <typename T, bool INV>
__device__ __forceinline__ cx<T> w32(cx<T> a, int j) {
    j = ((j % 32) + 32) % 32;
    if (j == 0) return a;
    if (j == 8) return rot90<T, INV>(a);
    if (j == 16) return cx<T>{-a.re, -a.im};
    if (j == 24) return rot90<T, !INV>(a);
    const T c = static_cast<T>(cospi16(j)), s = static_cast<T>(sinpi16(j));
    return INV ? cx<T>{a.re * c - a.im * s, a.re * s + a.im * c}
               : cx<T>{a.re * c + a.im * s, a.im * c - a.re * s};
}

template <typename T, bool INV>
__device__ __forceinline__ void dft2(cx<T>* a) {
    const cx<T> t = a[0];
    a[0] = t + a[1];
    a[1] = t - a[1];
}

template <typename T, bool INV>
__device__ __forceinline__ void dft4(cx<T>* a) {
    const cx<T> s0 = a[0] + a[2], d0 = a[0] - a[2];
    const cx<T> s1 = a[1] + a[3], d1 = rot90<T, INV>(a[1] - a[3]);
    a[0] = s0 + s1;
    a[2] = s0 - s1;
    a[1] = d0 + d1;
    a[3] = d0 - d1;
}

template <typename T, bool INV>
__device__ __forceinline__ void dft8(cx<T>* a) {
    cx<T> e[4] = {a[0], a[2], a[4], a[6]};
    cx<T> o[4] = {a[1], a[3], a[5], a[7]};
    dft4<T, INV>(e);
    dft4<T, INV>(o);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const cx<T> t = w32<T, INV>(o[k], 4 * k);
        a[k] = e[k] + t;
        a[k + 4] = e[k] - t;
    }
}

// 16-point DFT as 4 x 4 (decimation in frequency), natural order in and out.
template <typename T, bool INV>
__device__ __forceinline__ void dft16(cx<T>* a) {
    cx<T> b[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        cx<T> t[4] = {a[n2], a[4 + n2], a[8 + n2], a[12 + n2]};
        dft4<T, INV>(t);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) b[4 * k1 + n2] = w32<T, INV>(t[k1], 2 * k1 * n2);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cx<T> t[4] = {b[4 * k1], b[4 * k1 + 1], b[4 * k1 + 2], b[4 * k1 + 3]};
        dft4<T, INV>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = t[k2];
    }
}

// dft16 with a[8..15] == 0 (the zero-padded half of a free-space
// transform): the first 4-point pass sees two zero inputs.
template <typename T, bool INV>
__device__ __forceinline__ void dft16_zin(cx<T>* a) {
    cx<T> b[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        const cx<T> x0 = a[n2], x1 = a[4 + n2];
        const cx<T> r1 = rot90<T, INV>(x1);
        const cx<T> t[4] = {x0 + x1, x0 + r1, x0 - x1, x0 - r1};
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) b[4 * k1 + n2] = w32<T, INV>(t[k1], 2 * k1 * n2);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cx<T> t[4] = {b[4 * k1], b[4 * k1 + 1], b[4 * k1 + 2], b[4 * k1 + 3]};
        dft4<T, INV>(t);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = t[k2];
    }
}

// dft16 computing only the outputs a[0..7] (the half of a free-space
// transform that is kept); a[8..15] are left unspecified.
template <typename T, bool INV>
__device__ __forceinline__ void dft16_hout(cx<T>* a) {
    cx<T> b[16];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        cx<T> t[4] = {a[n2], a[4 + n2], a[8 + n2], a[12 + n2]};
        dft4<T, INV>(t);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) b[4 * k1 + n2] = w32<T, INV>(t[k1], 2 * k1 * n2);
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        const cx<T> s0 = b[4 * k1] + b[4 * k1 + 2], d0 = b[4 * k1] - b[4 * k1 + 2];
        const cx<T> s1 = b[4 * k1 + 1] + b[4 * k1 + 3];
        const cx<T> d1 = rot90<T, INV>(b[4 * k1 + 1] - b[4 * k1 + 3]);
        a[k1] = s0 + s1;
        a[k1 + 4] = d0 + d1;
    }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dft(cx<T>* a) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) dft2<T, INV>(a);
    else if constexpr (R == 4) dft4<T, INV>(a);
    else if constexpr (R == 8) dft8<T, INV>(a);
    else dft16<T, INV>(a);
}

template <int LOGN>
struct Shape {
    static_assert(LOGN >= 4 && LOGN <= 13, "register-blocked sizes 16 .. 8192");
    static constexpr int n = 1 << LOGN;
    static constexpr int A = LOGN / 4;
    static constexpr int RL = LOGN % 4;
    static constexpr int m = A + (RL ? 1 : 0);
    static constexpr int threads = n / 16;
    static constexpr int N(int s) { return s < A ? 16 : (1 << RL); }
    static constexpr int S(int s) {
        int p = 1;
        for (int j = s + 1; j < m; ++j) p *= N(j);
        return p;
    }
    static constexpr int P(int s) {
        int p = 1;
        for (int j = 0; j < s; ++j) p *= N(j);
        return p;
    }
    static constexpr int G(int s) { return 16 / N(s); }
    // position of element v of group g in stage s
    static __device__ __forceinline__ int pos(int s, int g, int v) {
        const int Ss = S(s), Ns = N(s);
        return (g / Ss) * Ns * Ss + v * Ss + (g % Ss);
    }
    // frequency index of a digit-reversed position
    static __device__ __forceinline__ int kidx(int p) {
        int k = 0;
#pragma unroll
        for (int j = 0; j < m; ++j) k += ((p / S(j)) % N(j)) * P(j);
        return k;
    }
};

// LDS index with a skew of one element every 16. The exchanges read and
// write with lane strides of 1, 16 and 32 elements; with this skew those
// strides map 16 lanes of 8-byte elements (and 8 lanes of 16-byte elements)
// to distinct banks (a skew of one per 32 left two-way conflicts on the
// stride-32 exchanges).
__device__ __forceinline__ int lds_ix(int p) { return p + (p >> 4); }
__host__ __device__ constexpr int lds_len(int n) { return n + (n >> 4) + 1; }

// Element -> slot layouts of LDS buffers. The LDS serves a wave's 4- and
// 8-byte accesses in two halves of 32 lanes, whose slots are conflict-free
// when they differ mod 32, and its 16-byte accesses in four quarters of 16
// lanes, conflict-free when they differ mod 16 (measured: a 2-way conflict
// costs two cycles per access in the first case, four in the second).
// Skew16 (one slot of skew per 16 elements) keeps stride-16 and
// stride-32-pair patterns apart but folds the two ends of a 32-element run
// onto one bank; Skew32x2 (two slots per 32) keeps runs of 32 and the
// stride-32 pairs apart. Both fit in lds_len(n) slots.
struct Skew16 {
    static __host__ __device__ constexpr int ix(int p) { return p + (p >> 4); }
};
struct Skew32x2 {
    static __host__ __device__ constexpr int ix(int p) { return p + 2 * (p >> 5); }
};

// twiddles W^{v*idx}, v = 0..N-1, applied to a group from one FP64 table
// read w = tw[idx] (tw[t] = exp(-2 pi i t/len)).
// FP64 data: powers as a running FP64 product (at most 15 roundings,
// ~1e-15 relative). FP32 data: w, w^2, w^4, w^8 by FP64 squaring, rounded
// once to FP32; the other powers are FP32 products of at most three of
// those (<= 4 roundings, ~2.5e-7 relative, the accuracy of an FP32 twiddle
// table). One table read per group measured faster than reading every power
// (the table reads, not the arithmetic, bound the stage).
// FP32 twiddle powers of one radix-16 group (f[v] = W^{v idx}, f[0]
// unused), formed by a forward stage and reused, conjugated, by the same
// stage of the inverse, which would otherwise form the powers of the
// conjugated table value again (the same values in exact arithmetic; the
// compiled sequences differ in FMA contraction, so results move at the
// FP32 rounding level: 1.1e-7 relative on a 4096^2 free-space matvec).
struct TwKeep {
    cx<float> f[16];
};

// KEEP: 0 = form the powers; 1 = form them and keep them in *keep
// (forward); 2 = apply the conjugates of *keep (inverse, same group)
template <typename T, int NS, bool INV, int KEEP = 0>
__device__ __forceinline__ void twiddle_group(cx<T>* a, const cx<double>* __restrict__ tw,
                                              int idx, int /*tmask*/, TwKeep* keep = nullptr) {
    if (idx == 0) return;
    if constexpr (KEEP == 2) {
        static_assert(std::is_same_v<T, float> && NS == 16, "kept powers: FP32 radix 16");
#pragma unroll
        for (int v = 1; v < 16; ++v) a[v] = a[v] * cx<float>{keep->f[v].re, -keep->f[v].im};
    } else if constexpr (std::is_same_v<T, float> && NS == 16) {
        cx<double> w = tw[idx];
        if (INV) w.im = -w.im;
        const cx<double> w2 = w * w, w4 = w2 * w2, w8 = w4 * w4;
        auto f32 = [](cx<double> z) { return cx<float>{(float)z.re, (float)z.im}; };
        cx<float> f[16];
        f[1] = f32(w);
        f[2] = f32(w2);
        f[4] = f32(w4);
        f[8] = f32(w8);
        f[3] = f[2] * f[1];
        f[5] = f[4] * f[1];
        f[6] = f[4] * f[2];
        f[7] = f[4] * f[3];
#pragma unroll
        for (int v = 9; v < 16; ++v) f[v] = f[8] * f[v - 8];
        if constexpr (KEEP == 1) {
#pragma unroll
            for (int v = 1; v < 16; ++v) keep->f[v] = f[v];
        }
#pragma unroll
        for (int v = 1; v < 16; ++v) a[v] = a[v] * f[v];
    } else {
        static_assert(KEEP == 0, "kept powers: FP32 radix 16");
        cx<double> w = tw[idx];
        if (INV) w.im = -w.im;
        cx<double> wv = w;
#pragma unroll
        for (int v = 1; v < NS; ++v) {
            if (v > 1) wv = wv * w;
            a[v] = a[v] * cx<T>{static_cast<T>(wv.re), static_cast<T>(wv.im)};
        }
    }
}

// exp(-2 pi i v/32), v = 0..15: the twiddles of a radix-16 stage of stride 2
// (followed by a radix-2 stage), whose groups take either these or ones.
constexpr double kW32re[16] = {
    1.0, 0.98078528040323044913, 0.92387953251128675613, 0.83146961230254523708,
    0.7071067811865475244, 0.55557023301960222474, 0.38268343236508977173,
    0.19509032201612826785, 0.0, -0.19509032201612826785, -0.38268343236508977173,
    -0.55557023301960222474, -0.7071067811865475244, -0.83146961230254523708,
    -0.92387953251128675613, -0.98078528040323044913};
constexpr double kW32im[16] = {
    0.0, -0.19509032201612826785, -0.38268343236508977173, -0.55557023301960222474,
    -0.7071067811865475244, -0.83146961230254523708, -0.92387953251128675613,
    -0.98078528040323044913, -1.0, -0.98078528040323044913, -0.92387953251128675613,
    -0.83146961230254523708, -0.7071067811865475244, -0.55557023301960222474,
    -0.38268343236508977173, -0.19509032201612826785};

template <typename T, bool INV>
__device__ __forceinline__ void twiddle_w32(cx<T>* a, int low) {
    if (low == 0) return;
#pragma unroll
    for (int v = 1; v < 16; ++v) {
        const cx<T> w = {static_cast<T>(kW32re[v]),
                         static_cast<T>(INV ? -kW32im[v] : kW32im[v])};
        a[v] = a[v] * w;
    }
}

// Stage s of the forward (INV=false, DFT then twiddle) or inverse (INV=true,
// twiddle then DFT) transform on the registers of thread th. PR (stage 0
// only, where every thread holds one group of 16 at positions v*S_0 + th):
// 1 = inputs v >= 8 are zero, 2 = only outputs v < 8 are needed.
// KS: the stage whose FP32 twiddle powers the forward keeps in *keep and
// the inverse reuses (-1: none; see keep_stage)
template <typename T, int LOGN, int s, bool INV, int PR = 0, int KS = -1>
__device__ __forceinline__ void stage(cx<T>* r, int th, const cx<double>* __restrict__ tw,
                                      int twmul, TwKeep* keep = nullptr) {
    using Sh = Shape<LOGN>;
    constexpr int Ns = Sh::N(s), Ss = Sh::S(s), Gs = Sh::G(s);
    constexpr int tstep = Sh::n / (Ns * Ss);     // table stride of W_{Ns Ss}
    constexpr int KP = (s == KS) ? (INV ? 2 : 1) : 0;
    static_assert(PR == 0 || (s == 0 && Ns == 16 && Gs == 1), "pruned stage 0 only");
    static_assert(KP == 0 || (Ns == 16 && Gs == 1 && Ss > 2 && s < Sh::m - 1),
                  "kept powers: a twiddled radix-16 stage with one group per thread");
#pragma unroll
    for (int q = 0; q < Gs; ++q) {
        cx<T>* a = r + q * Ns;
        const int low = (th * Gs + q) % Ss;
        // stride 2: the group's twiddles are W32^v (low = 1) or ones,
        // constants rather than table reads and products
        constexpr bool W32 = (Ns == 16 && Ss == 2);
        if constexpr (INV) {
            if constexpr (s < Sh::m - 1) {
                if constexpr (W32) twiddle_w32<T, true>(a, low);
                else twiddle_group<T, Ns, true, KP>(a, tw, low * tstep * twmul,
                                                    Sh::n * twmul - 1, keep);
            }
            if constexpr (PR == 2) dft16_hout<T, true>(a);
            else dft<T, Ns, true>(a);
        } else {
            if constexpr (PR == 1) dft16_zin<T, false>(a);
            else dft<T, Ns, false>(a);
            if constexpr (s < Sh::m - 1) {
                if constexpr (W32) twiddle_w32<T, false>(a, low);
                else twiddle_group<T, Ns, false, KP>(a, tw, low * tstep * twmul,
                                                     Sh::n * twmul - 1, keep);
            }
        }
    }
}

// The stage whose twiddle powers a forward-then-inverse kernel can keep:
// stage 1 where it is a generic twiddled radix-16 stage (n >= 1024).
template <int LOGN>
__host__ __device__ constexpr int keep_stage() {
    using Sh = Shape<LOGN>;
    return (Sh::m >= 3 && Sh::N(1) == 16 && Sh::G(1) == 1 && Sh::S(1) > 2) ? 1 : -1;
}

// No-op hook of the exchanges.
struct NoHook {
    __device__ __forceinline__ void operator()() const {}
};

// Positions of the registers of thread th in stage s's grouping, as
// base(th) + off(q, v) = Shape::pos(s, th*G + q, v): a radix-16 stage holds
// one group (off = v*S), the last stage of a radix below 16 holds G groups
// of stride 1 (base = 16 th, off = q*N + v). The skew does not carry
// between the two parts (for Skew16, base mod 16 plus off mod 16 stays below
// 16 in both cases; for Skew32x2, the same mod 32 on the exchanges it is
// used for), so Lay::ix(base + off) = Lay::ix(base) + Lay::ix(off): one
// address per thread and compile-time offsets.
template <int LOGN, int s, typename Lay = Skew16>
struct Grouping {
    using Sh = Shape<LOGN>;
    static constexpr int N = Sh::N(s), S = Sh::S(s), G = Sh::G(s);
    static_assert(G == 1 || S == 1, "only the last stage holds several groups");
    static __device__ __forceinline__ int base(int th) {
        if constexpr (G == 1) return (th / S) * N * S + th % S;
        else return 16 * th;
    }
    static constexpr int off(int q, int v) {
        const int o = G == 1 ? v * S : q * N + v;
        return Lay::ix(o);
    }
};

// Layout of the exchange between stages e and e + 1 (both directions). For
// transforms with whole waves per transform (n >= 1024) every exchange but
// the last takes Skew32x2: its groupings are runs of 32 (or two runs of 16)
// on one side and runs or stride-32 pairs on the other, which Skew16
// conflicts on and Skew32x2 does not, for 4-, 8- and 16-byte elements
// alike; the last exchange's stride-16 side needs Skew16. Both stay linear in the same way as above:
// ix(base + off) = ix(base) + ix(off) for every base and offset used, so
// addressing stays one address per thread plus compile-time offsets.
template <typename T, int LOGN, int e>
using ExLayout = std::conditional_t<(LOGN >= 10 && e + 2 < Shape<LOGN>::m), Skew32x2, Skew16>;

// Whether the exchange between groupings sa and sb stays inside groups of
// `lanes` consecutive threads: each group's positions are the same set in
// both groupings, so no element crosses to another group. `lanes` is the
// fewest threads of one transform a wave can hold: 64 / W with W up to
// 512 / threads transforms per workgroup (a column workgroup interleaves
// its W columns across threads), 64 when one transform fills the wave.
template <int LOGN, int sa, int sb>
constexpr bool wave_local() {
    using Sh = Shape<LOGN>;
    constexpr int T = Sh::threads;
    constexpr int Wmax = T >= 512 ? 1 : 512 / T;
    constexpr int lanes = 64 / Wmax;
    if (T < 64 || lanes < 1) return false;
    auto pos = [](int s, int g, int v) {
        return (g / Sh::S(s)) * Sh::N(s) * Sh::S(s) + v * Sh::S(s) + (g % Sh::S(s));
    };
    for (int g0 = 0; g0 < T; g0 += lanes) {
        int lo[2] = {1 << 30, 1 << 30}, hi[2] = {-1, -1};
        const int st[2] = {sa, sb};
        for (int k = 0; k < 2; ++k)
            for (int th = g0; th < g0 + lanes; ++th)
                for (int q = 0; q < Sh::G(st[k]); ++q)
                    for (int v = 0; v < Sh::N(st[k]); ++v) {
                        const int p = pos(st[k], th * Sh::G(st[k]) + q, v);
                        lo[k] = p < lo[k] ? p : lo[k];
                        hi[k] = p > hi[k] ? p : hi[k];
                    }
        // 16 distinct positions per thread: equal bounds spanning exactly
        // 16 * lanes positions mean equal sets
        if (lo[0] != lo[1] || hi[0] != hi[1] || hi[0] - lo[0] + 1 != 16 * lanes) return false;
    }
    return true;
}

// Order a wave's LDS writes before its later LDS reads: the LDS executes one
// wave's accesses in issue order, so this only keeps the compiler from
// moving the reads above the writes (no instruction is emitted).
__device__ __forceinline__ void wave_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exchanges of FP32 data that stay inside every wave are ordered by
// wave_order instead of workgroup barriers.
template <typename T, int LOGN, int sa, int sb>
__device__ __forceinline__ void exchange_sync() {
    if constexpr (wave_local<LOGN, sa, sb>()) wave_order();
    else __syncthreads();
}

// Regroup registers from stage sa's grouping to stage sb's through LDS.
// `hook` runs between the LDS writes and the barrier, where the registers
// of the data are free: a caller issues global loads there (the kernel
// spectrum of the column pass) so that their latency overlaps the barrier,
// the LDS reads and the next stage.
template <typename T, int LOGN, int sa, int sb, typename Hook = NoHook>
__device__ __forceinline__ void exchange(cx<T>* r, cx<T>* lds, int th,
                                         const Hook& hook = Hook{}) {
    using Lay = ExLayout<T, LOGN, (sa < sb ? sa : sb)>;
    using A = Grouping<LOGN, sa, Lay>;
    using B = Grouping<LOGN, sb, Lay>;
    exchange_sync<T, LOGN, sa, sb>();
    cx<T>* wa = lds + Lay::ix(A::base(th));
#pragma unroll
    for (int q = 0; q < A::G; ++q)
#pragma unroll
        for (int v = 0; v < A::N; ++v) wa[A::off(q, v)] = r[q * A::N + v];
    hook();
    exchange_sync<T, LOGN, sa, sb>();
    const cx<T>* rbp = lds + Lay::ix(B::base(th));
#pragma unroll
    for (int q = 0; q < B::G; ++q)
#pragma unroll
        for (int v = 0; v < B::N; ++v) r[q * B::N + v] = rbp[B::off(q, v)];
}

// The same regrouping through a real (T) LDS buffer of half the size: real
// parts first, then imaginary parts (four barriers instead of two). The
// second write reads only the imaginary parts, which the first read left
// untouched. Half the LDS per transform lets more workgroups be resident.
template <typename T, int LOGN, int sa, int sb, typename Hook = NoHook>
__device__ __forceinline__ void exchange(cx<T>* r, T* lds, int th,
                                         const Hook& hook = Hook{}) {
    using Lay = ExLayout<T, LOGN, (sa < sb ? sa : sb)>;
    using A = Grouping<LOGN, sa, Lay>;
    using B = Grouping<LOGN, sb, Lay>;
    T* wa = lds + Lay::ix(A::base(th));
    const T* rbp = lds + Lay::ix(B::base(th));
    exchange_sync<T, LOGN, sa, sb>();
#pragma unroll
    for (int q = 0; q < A::G; ++q)
#pragma unroll
        for (int v = 0; v < A::N; ++v) wa[A::off(q, v)] = r[q * A::N + v].re;
    hook();
    exchange_sync<T, LOGN, sa, sb>();
#pragma unroll
    for (int q = 0; q < B::G; ++q)
#pragma unroll
        for (int v = 0; v < B::N; ++v) r[q * B::N + v].re = rbp[B::off(q, v)];
    exchange_sync<T, LOGN, sa, sb>();
#pragma unroll
    for (int q = 0; q < A::G; ++q)
#pragma unroll
        for (int v = 0; v < A::N; ++v) wa[A::off(q, v)] = r[q * A::N + v].im;
    exchange_sync<T, LOGN, sa, sb>();
#pragma unroll
    for (int q = 0; q < B::G; ++q)
#pragma unroll
        for (int v = 0; v < B::N; ++v) r[q * B::N + v].im = rbp[B::off(q, v)];
}

// LDS element type L: cx<T> (one pass per exchange) or T (split passes)
template <typename T, int LOGN, int s, bool ZIN, typename Hook, typename L, int KS>
__device__ __forceinline__ void forward_from(cx<T>* r, L* lds, int th,
                                             const cx<double>* __restrict__ tw, int twmul,
                                             const Hook& hook, TwKeep* keep) {
    stage<T, LOGN, s, false, (s == 0 && ZIN) ? 1 : 0, KS>(r, th, tw, twmul, keep);
    if constexpr (s + 1 < Shape<LOGN>::m) {
        if constexpr (s + 2 == Shape<LOGN>::m) exchange<T, LOGN, s, s + 1>(r, lds, th, hook);
        else exchange<T, LOGN, s, s + 1>(r, lds, th);
        forward_from<T, LOGN, s + 1, ZIN, Hook, L, KS>(r, lds, th, tw, twmul, hook, keep);
    } else if constexpr (s == 0) {
        hook();   // single-stage transform: no exchange to hide the loads in
    }
}

template <typename T, int LOGN, int s, bool HOUT, typename L, int KS>
__device__ __forceinline__ void inverse_from(cx<T>* r, L* lds, int th,
                                             const cx<double>* __restrict__ tw, int twmul,
                                             TwKeep* keep) {
    stage<T, LOGN, s, true, (s == 0 && HOUT) ? 2 : 0, KS>(r, th, tw, twmul, keep);
    if constexpr (s > 0) {
        exchange<T, LOGN, s, s - 1>(r, lds, th);
        inverse_from<T, LOGN, s - 1, HOUT, L, KS>(r, lds, th, tw, twmul, keep);
    }
}

// forward: registers in stage-0 grouping (natural order) -> stage m-1
// grouping (digit-reversed); inverse: the reverse. ZIN: the upper half of
// the input (positions >= n/2, i.e. v >= 8 of stage 0) is zero; HOUT: only
// the lower half of the output is used.
// `hook` runs once, inside the last exchange (see `exchange`). KS >= 0: the
// forward keeps stage KS's twiddle powers in *keep, and an inverse of the
// same thread with the same KS reuses them.
template <typename T, int LOGN, bool ZIN = false, typename Hook = NoHook, typename L = cx<T>,
          int KS = -1>
__device__ __forceinline__ void forward(cx<T>* r, L* lds, int th,
                                        const cx<double>* __restrict__ tw, int twmul,
                                        const Hook& hook = Hook{}, TwKeep* keep = nullptr) {
    forward_from<T, LOGN, 0, ZIN, Hook, L, KS>(r, lds, th, tw, twmul, hook, keep);
}
template <typename T, int LOGN, bool HOUT = false, typename L = cx<T>, int KS = -1>
__device__ __forceinline__ void inverse(cx<T>* r, L* lds, int th,
                                        const cx<double>* __restrict__ tw, int twmul,
                                        TwKeep* keep = nullptr) {
    inverse_from<T, LOGN, Shape<LOGN>::m - 1, HOUT, L, KS>(r, lds, th, tw, twmul, keep);
}

} // namespace rb
} // namespace aspher
