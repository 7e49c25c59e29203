// Device-side FFT building blocks for gfx950: batched complex transforms of
// length L held in LDS, Stockham autosort passes (natural-order output, no
// bit reversal), radix 8/4/2 butterflies in registers for powers of two and a
// direct-DFT mixed-radix path for every other length.
//
// Conventions (match the reference engine, src/fft_engine.hpp:18-22):
// forward = exp(-2 pi i jk/L), inverse = exp(+2 pi i jk/L), both
// unnormalised. Twiddles come from a host-built table tw[t] = exp(-2 pi i t/n)
// accurate to an ulp; a length-L transform reads tw[t * twmul] with
// twmul = n / L.
//
// Stockham pass with radix R and span p (product of earlier radices), for
// butterfly i in [0, L/R) and k = i mod p:
//   u_j = x[i + j L/R] * w^(j k),  w = exp(-+2 pi i / (p R))
//   y[(i - k) R + k + j p] = DFT_R(u)_j
#pragma once

#include "common.hpp"

namespace aspher {
namespace fft {

template <typename T, bool INV>
__device__ __forceinline__ cx<T> rot90(cx<T> a) {
    // forward: multiply by -i; inverse: by +i
    return INV ? cx<T>{-a.im, a.re} : cx<T>{a.im, -a.re};
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
    const T r = T(0.70710678118654752440084436210484903928);
    const cx<T> o1 = o[1], o3 = o[3];
    if (INV) {
        o[1] = {r * (o1.re - o1.im), r * (o1.re + o1.im)};
        o[3] = {-r * (o3.re + o3.im), r * (o3.re - o3.im)};
    } else {
        o[1] = {r * (o1.re + o1.im), r * (o1.im - o1.re)};
        o[3] = {r * (o3.im - o3.re), -r * (o3.re + o3.im)};
    }
    o[2] = rot90<T, INV>(o[2]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = e[k] + o[k];
        a[k + 4] = e[k] - o[k];
    }
}

template <typename T, int R, bool INV>
__device__ __forceinline__ void dftR(cx<T>* a) {
    if constexpr (R == 2) dft2<T, INV>(a);
    else if constexpr (R == 4) dft4<T, INV>(a);
    else dft8<T, INV>(a);
}

template <typename T, bool INV>
__device__ __forceinline__ cx<T> twiddle(const cx<T>* __restrict__ tw, int idx) {
    cx<T> w = tw[idx];
    if (INV) w.im = -w.im;
    return w;
}

// One in-place Stockham pass over `batch` transforms of length L = 2^logL
// stored at buf + b*bstride. Every thread owns exactly 8 elements:
// blockDim.x * 8 == L * batch (enforced by the launchers). All threads read
// their inputs, barrier, then write their outputs, so one LDS buffer suffices.
template <typename T, int R, bool INV>
__device__ __forceinline__ void pass_pow2(cx<T>* buf, int logL, int batch_unused,
                                          int bstride, int logp,
                                          const cx<T>* __restrict__ tw, int twmul) {
    constexpr int K = 8 / R;
    constexpr int logR = (R == 2) ? 1 : (R == 4) ? 2 : 3;
    const int lognb = logL - logR;          // butterflies per transform = L/R
    const int nbm = (1 << lognb) - 1;
    const int pm = (1 << logp) - 1;
    const int twstep = (1 << (logL - logp - logR)) * twmul;  // L/(pR) in table units
    cx<T> v[K][R];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int bf = threadIdx.x + k * blockDim.x;
        const int b = bf >> lognb, i = bf & nbm;
        const cx<T>* src = buf + b * bstride + i;
#pragma unroll
        for (int j = 0; j < R; ++j) v[k][j] = src[j << lognb];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int bf = threadIdx.x + k * blockDim.x;
        const int b = bf >> lognb, i = bf & nbm;
        const int kk = i & pm;
        if (logp > 0) {
#pragma unroll
            for (int j = 1; j < R; ++j)
                v[k][j] = v[k][j] * twiddle<T, INV>(tw, j * kk * twstep);
        }
        dftR<T, R, INV>(v[k]);
        cx<T>* dst = buf + b * bstride + ((i - kk) << logR) + kk;
#pragma unroll
        for (int j = 0; j < R; ++j) dst[j << logp] = v[k][j];
    }
    __syncthreads();
}

// Complete power-of-two transform (L >= 8). Precondition: data in LDS and a
// barrier has been passed. Postcondition: natural-order result in place,
// barrier passed.
template <typename T, bool INV>
__device__ void fft_pow2(cx<T>* buf, int logL, int batch, int bstride,
                         const cx<T>* __restrict__ tw, int twmul) {
    int logp = 0;
    int rem = logL;
    if (rem % 3 == 1) {
        pass_pow2<T, 2, INV>(buf, logL, batch, bstride, logp, tw, twmul);
        logp += 1;
        rem -= 1;
    } else if (rem % 3 == 2) {
        pass_pow2<T, 4, INV>(buf, logL, batch, bstride, logp, tw, twmul);
        logp += 2;
        rem -= 2;
    }
    while (rem > 0) {
        pass_pow2<T, 8, INV>(buf, logL, batch, bstride, logp, tw, twmul);
        logp += 3;
        rem -= 3;
    }
}

// Mixed-radix factorisation passed by value to the generic kernels.
struct Radices {
    int n;          // number of passes
    int r[24];      // radix of each pass (product == L)
};

// Generic-length transform: out-of-place ping-pong between two LDS buffers,
// each butterfly a direct DFT of its radix using the shared table. Any radix
// works (a prime L is one O(L^2) pass); meant for the non-power-of-two grids
// the API allows, not for the benchmark sizes. Returns the buffer holding
// the natural-order result; ends with a barrier.
template <typename T, bool INV>
__device__ cx<T>* fft_generic(cx<T>* buf0, cx<T>* buf1, int L, int batch,
                              int bstride, const Radices& rad,
                              const cx<T>* __restrict__ tw, int twmul) {
    cx<T>* in = buf0;
    cx<T>* out = buf1;
    int p = 1;
    for (int ps = 0; ps < rad.n; ++ps) {
        const int R = rad.r[ps];
        const int nb = L / R;
        const int unit = L / (p * R);
        for (int bf = threadIdx.x; bf < nb * batch; bf += blockDim.x) {
            const int b = bf / nb, i = bf - b * nb;
            const int k = i % p;
            const cx<T>* src = in + b * bstride + i;
            cx<T>* dst = out + b * bstride + (i - k) * R + k;
            for (int m = 0; m < R; ++m) {
                cx<T> acc = {T(0), T(0)};
                for (int j = 0; j < R; ++j) {
                    const long e = (long)j * (k + m * p) * unit;
                    acc = acc + src[j * nb] * twiddle<T, INV>(tw, (int)(e % L) * twmul);
                }
                dst[m * p] = acc;
            }
        }
        __syncthreads();
        cx<T>* t = in;
        in = out;
        out = t;
        p *= R;
    }
    return in;
}

} // namespace fft
} // namespace aspher
