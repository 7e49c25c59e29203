// Launch interface of the register-blocked power-of-two spectral kernels.
#pragma once

#include "common.hpp"

#include <cstdint>

namespace aspher {

// Instances of one launch (blockIdx.y): equal-size problems of an ensemble
// group. Local instance l uses x + l*x_stride (and its mask), y + l*y_stride,
// its own slice of the plan workspace, and the scalar/partial blocks of
// global instance first + l; instances whose gate flag is 0 are skipped.
struct RbBatch {
    int count = 1;
    std::int64_t x_stride = 0, y_stride = 0, at_stride = 0, br_stride = 0;
    std::int64_t m_stride = -1;   // mask stride (-1: that of x, resp. y)
    int first = 0;
    int sstride = 0;
    std::int64_t pstride = 0;
    const double* gscal = nullptr;
    int gate = -1;
};

template <typename T>
struct RbRowsFwd {
    const T* x = nullptr;
    int x_pitch = 0, ncols = 0, nrows = 0;
    cx<T>* AT = nullptr;        // transposed half spectrum, AT[kx*pitchT + iy]
    int pitchT = 0;
    int B = 1;                  // rows per workgroup
    const cx<double>* tw = nullptr;   // exp(-2 pi i t/n), n = 2L
    int mode = 0;               // 0 plain, 1 masked (mask ? x - gmean : 0),
                                // 2/3/4 embedding of an even / transposed even /
                                // odd-odd (xy) influence table
    const std::uint8_t* mask = nullptr;
    const double* scal = nullptr;
    int gmean_idx = 0;
    const double* table = nullptr;
    int Ns_embed = 0;
    int remap = 1;
    RbBatch bt;
};

template <typename T>
struct RbCols {
    const cx<T>* AT = nullptr;
    int pitchT = 0;
    cx<T>* BR = nullptr;        // row-major output BR[iy*pitch + kx]
    int pitch = 0;
    int ncols = 0, rows_in = 0, rows_out = 0;
    int W = 1;                  // columns per workgroup
    const cx<double>* tw = nullptr;   // exp(-2 pi i t/n)
    int sym = 0;                // 0 table, 1 periodic, 2 precond, 3 build
    const T* ktab = nullptr;
    int kp = 0;
    double sc = 0.0;
    T* kout = nullptr;
    double* part = nullptr;
    int remap = 1;
    // table symbol: per column thread (chunk << 1) | reversed into the
    // half-size table (SpectralPlan::permute_table_half)
    const int* kmap = nullptr;
    // 1: the result goes back into the AT column it came from (BR == AT),
    // transposed like the input, instead of row-major into BR
    int out_t = 0;
    RbBatch bt;
};

template <typename T>
struct RbRowsInv {
    const cx<T>* BR = nullptr;
    int pitch = 0, nrows = 0;
    int B = 1;
    const cx<double>* tw = nullptr;   // exp(-2 pi i t/n), n = 2L
    T* y = nullptr;
    int y_pitch = 0, ncols = 0;
    int mode = 0;               // 0 plain, 1 masked with per-block sum
    const std::uint8_t* mask = nullptr;
    double* part = nullptr;
    int remap = 1;
    // 1: BR is the transposed half spectrum (value of row iy, frequency k
    // at BR[k*pitch + iy], pitch the column pitch; instance stride at_stride)
    int in_t = 0;
    RbBatch bt;
};

// Self-affine random-phase spectrum, column pass of the surface generator:
// BR[ky][kx] (ky < n, kx <= n/2) = inverse column DFT of
//   H(kx, ky) = scale/2 * A(m) * (exp(i phi(k)) + exp(-i phi(-k))),
// A(m) = |m|^-(1+hurst) for k_low <= |m| <= k_high (else 0), m the signed
// frequency of k, phi(k) = 2 pi u(seed, ky*n + kx) from a SplitMix64 hash.
struct RbGen {
    cx<double>* BR = nullptr;
    int pitch = 0;
    double hurst = 0.8, k_low = 4.0, k_high = 64.0, scale = 1.0;
    unsigned long long seed = 0;
    const cx<double>* tw = nullptr;
    int W = 1;
    int remap = 1;
};
void rb_gen_cols(int logN, const RbGen& a, hipStream_t s);

template <typename T> void rb_rows_fwd(int logL, const RbRowsFwd<T>& a, hipStream_t s);
template <typename T> void rb_cols(int logN, const RbCols<T>& a, hipStream_t s);
template <typename T> void rb_rows_inv(int logL, const RbRowsInv<T>& a, hipStream_t s);
// Two-component (tangential) column pass: instance 0 and 1 of the plan are
// the x and y components; after the forward transform each mode is mixed by
// a symmetric real 2x2 symbol
//   sym 5: tables Kxx, Kyy, Kxy in the register order (free space, pruned)
//   sym 6: |k|/Ns^2 (I + gamma k k^T/|k|^2), cross term 0 on Nyquist modes
template <typename T>
struct RbCols2 {
    const cx<T>* AT = nullptr;
    int pitchT = 0;
    std::int64_t at_stride = 0;
    cx<T>* BR = nullptr;
    int pitch = 0;
    std::int64_t br_stride = 0;
    int ncols = 0, rows_in = 0, rows_out = 0;
    const cx<double>* tw = nullptr;
    int sym = 5;
    const T *kxx = nullptr, *kyy = nullptr, *kxy = nullptr;
    int kp = 0;
    double gamma = 0.0, norm = 0.0;
    int remap = 1;
};
template <typename T> void rb_cols2(int logN, const RbCols2<T>& a, hipStream_t s);

// unique (n/2+1)^2 free-space spectrum -> nh x n table in the column
// kernel's register order (see k_rb_permute_table)
// (odd: the table is odd in ky, values for ky > n/2 take the opposite sign)
template <typename T>
void rb_permute_table(int logN, const T* kin, int kp, int nh, T* kout, hipStream_t s,
                      bool odd = false);
// the half-size layout: kout[kx*kph + 16*chunk + i] = K(kx, ky) of register i
// of the thread that reads chunk forward (kmap: per thread (chunk << 1) |
// reversed)
template <typename T>
void rb_permute_table_half(int logN, const T* kin, int kp, int nh, const int* kmap, int kph,
                           T* kout, hipStream_t s);

} // namespace aspher
