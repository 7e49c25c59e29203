// Host-side interface of the spectral (FFT-convolution) kernels.
//
// A 2-D spectral convolution y = F^-1 (s . F x) on an n x n grid is three
// launches, each one HBM round trip:
//   rows_fwd : real rows -> half spectrum A[iy][kx], kx in [0, n/2]
//   cols     : per column kx: forward c2c along iy, multiply by the symbol,
//              inverse c2c, in place in A (pruned: only rows_in rows read,
//              only rows_out rows written)
//   rows_inv : A rows -> real rows (only ncols_out values written)
// The free-space operator runs it on the (2Ns)^2 zero-padded grid with
// rows_in = rows_out = Ns (src/fft_operator.cpp:87-138); the periodic
// operator and the Fourier preconditioner on the Ns^2 torus.
#pragma once

#include "common.hpp"

#include <cstdint>
#include <memory>
#include <vector>

namespace aspher {

enum class Symbol : int {
    table = 0,     // real even kernel spectrum K[kx][min(ky, n-ky)] (free space)
    periodic = 1,  // scale / |m| evaluated in double (periodic compliance)
    precond = 2,   // |k| / Ns^2 evaluated in float (Fourier preconditioner)
    build = 3,     // forward only: write Re/n^2 into the table, max |Im|, |Re|
};

enum class RowIn : int {
    plain = 0,     // x[row*pitch + col]
    masked = 1,    // mask ? x - gmean : 0 (gmean read from a device scalar)
    embed = 2,     // wrap-around embedding of the Love table (K build)
};

enum class RowOut : int {
    plain = 0,     // y = value
    masked = 1,    // y = mask ? value : 0, per-block sum over the mask
};

// Launch configuration of one transform size (tuned by the autotuner).
struct SpectralTuning {
    int row_batch = 0;   // rows per workgroup in the row kernels (0 = auto)
    int col_width = 0;   // columns per workgroup in the column kernel (0 = auto)
    bool xcd_remap = true;
    // register-blocked plans: the column pass writes its result back into
    // the transposed half spectrum and the inverse row pass reads it there
    bool col_out_t = true;
};

// Equal-size instances of one convolve launch (ensemble groups, register-
// blocked sizes only): instance l reads x + l*x_stride, writes y + l*y_stride
// and uses the scalar/partial blocks of global instance first + l; instances
// whose gate flag is 0 are skipped.
struct SpecBatch {
    int count = 1;
    std::int64_t x_stride = 0, y_stride = 0;
    int first = 0, sstride = 0;
    std::int64_t pstride = 0;
    const double* gscal = nullptr;
    int gate = -1;
};

// Twiddle table exp(-2 pi i t/n), t in [0, n), built on the host in long
// double and uploaded once per (n, precision).
template <typename T>
struct Twiddles {
    int n = 0;
    cx<T>* dev = nullptr;
    explicit Twiddles(int n_);
    ~Twiddles();
    Twiddles(const Twiddles&) = delete;
    Twiddles& operator=(const Twiddles&) = delete;
};

// A planned n x n spectral convolution in precision T. Owns the twiddles and
// the half-spectrum workspace A (rows_alloc x pitch complex).
template <typename T>
class SpectralPlan {
public:
    SpectralPlan(int n, int rows_in, int cols_in, int rows_out, int cols_out,
                 SpectralTuning tune = {}, int batch = 1);
    ~SpectralPlan();
    SpectralPlan(const SpectralPlan&) = delete;
    SpectralPlan& operator=(const SpectralPlan&) = delete;

    // Full convolution: y = inv(sym . fwd(x)). x and y may alias.
    // ktab: symbol table (Symbol::table) with row pitch ktab_pitch.
    // mask/scal/part: for RowIn::masked / RowOut::masked (gmean at
    // scal[gmean_idx]; per-block partial sums written to part[block]).
    void convolve(const T* x, int x_pitch, T* y, int y_pitch, Symbol sym,
                  const T* ktab, int ktab_pitch, double sym_scale,
                  hipStream_t s, RowIn in_mode = RowIn::plain,
                  RowOut out_mode = RowOut::plain,
                  const std::uint8_t* mask = nullptr,
                  const double* scal = nullptr, int gmean_idx = 0,
                  double* part = nullptr, hipEvent_t* stage_events = nullptr,
                  const SpecBatch* batch = nullptr, void* ext_ws = nullptr);
    // Bytes of caller workspace that convolve(..., ext_ws) uses in place of
    // the plan's own intermediate spectra (single instance, register-blocked
    // plans; 0 for the generic path, which takes no caller workspace).
    std::size_t external_ws_bytes() const;
    bool register_blocked() const { return rb_; }

    // Forward-only pass used to build the free-space kernel spectrum from
    // the Love table (T = double): writes Re/n^2 into kout[(kx)(n/2+1)+ky]
    // for ky in [0, n/2] and per-block max |Im|, max |Re| into part.
    // embed: 2 even table, 3 even transposed (yy from xx), 4 odd-odd (xy)
    void build_spectrum(const double* table, int Ns, T* kout, double* part,
                        hipStream_t s, int embed = 2);

    // Two-component (tangential) convolution on a plan of batch >= 2: the x
    // and y components at x and x + x_stride (outputs y, y + y_stride) are
    // transformed, mixed per mode by a real symmetric 2x2 symbol
    // (sym 5: tables in the register order, sym 6: the tangential
    // preconditioner symbol with `gamma`, `norm`) and transformed back.
    // Masked modes share one mask for both components; with RowOut::masked
    // the per-block sums of component c go to part + c*pstride.
    void convolve2(const T* x, int x_pitch, std::int64_t x_stride, T* y, int y_pitch,
                   std::int64_t y_stride, int sym, const T* kxx, const T* kyy, const T* kxy,
                   int kp, double gamma, double norm, hipStream_t s,
                   RowIn in_mode = RowIn::plain, RowOut out_mode = RowOut::plain,
                   const std::uint8_t* mask = nullptr, const double* scal = nullptr,
                   int gmean_idx = 0, double* part = nullptr, std::int64_t pstride = 0);

    // Rough-surface synthesis on an n x n plan (rows/cols in and out = n):
    // out (n^2, row-major) = c2r of the random-phase self-affine spectrum,
    // scaled by `scale` (see RbGen). Register-blocked sizes, T = double.
    void generate_surface(double hurst, double k_low, double k_high, double scale,
                          unsigned long long seed, T* out, hipStream_t s);

    int row_out_blocks() const;   // blocks of rows_inv (partial count)
    int batch_capacity() const { return batch_; }
    // Symbol::table layout the kernels of this plan read: the register-
    // blocked path wants the spectrum in its column register order (nh rows
    // of n values, pitch n); the generic path the unique (n/2+1)-wide table.
    bool table_permuted() const { return rb_; }
    std::size_t table_elems() const { return (std::size_t)nh * n; }
    // The layout the table-symbol column pass reads (register-blocked plans):
    // the spectrum is even in ky and the threads' frequency sets come in
    // mirror pairs (thread t' holds n - ky of thread t, in reversed register
    // order), so each pair shares one 16-value chunk; chunks per column and
    // the table size follow.
    int table_chunks() const { return kmap_chunks_; }
    int table_pitch_half() const { return 16 * kmap_chunks_; }
    std::size_t table_elems_half() const { return (std::size_t)nh * 16 * kmap_chunks_; }
    void permute_table_half(const T* kunique, int kp, T* kout, hipStream_t s) const;
    void permute_table(const T* kunique, int kp, T* kout, hipStream_t s,
                       bool odd = false) const;
    SpectralTuning tuning() const { return {row_batch_, col_width_, xcd_remap_, tcols_}; }
    int col_blocks() const;       // blocks of cols (partial count, build)
    std::size_t workspace_bytes() const;

    int n, rows_in, cols_in, rows_out, cols_out, nh, pitch;

private:
    std::size_t ext_br_bytes() const;
    void rows_fwd(const T* x, int x_pitch, RowIn mode, const std::uint8_t* mask,
                  const double* scal, int gmean_idx, const double* table,
                  int Ns_embed, int nrows, int ncols, hipStream_t s);
    void cols(Symbol sym, const T* ktab, int ktab_pitch, double sym_scale,
              T* kout, double* part, int rin, int rout, hipStream_t s);
    void rows_inv(T* y, int y_pitch, RowOut mode, const std::uint8_t* mask,
                  double* part, hipStream_t s);

    std::unique_ptr<Twiddles<T>> tw_;
    std::size_t a_bytes_ = 0, at_bytes_ = 0;
    cx<T>* A_ = nullptr;          // row-major half spectrum (BR of the rb path)
    std::size_t a_rows_ = 0;
    bool pow2_ = false;
    // register-blocked path (powers of two, 32 <= n <= 8192): transposed
    // half spectrum AT (nh columns of pitchT) + FP64 twiddles
    bool rb_ = false;
    cx<T>* AT_ = nullptr;
    int pitchT_ = 0;
    std::unique_ptr<Twiddles<double>> tw64_;
    int logL_row_ = 0, logn_ = 0;
    int batch_ = 1;
    std::int64_t at_stride_ = 0, br_stride_ = 0;
    int row_batch_ = 1, col_width_ = 1;
    int row_batch_fwd_ = 1;       // rows per workgroup of the forward row pass
    bool tcols_ = false;          // transposed column output (SpectralTuning::col_out_t)
    int row_batch_inv_ = 1;       // rows per workgroup of the register-blocked inverse pass
    bool xcd_remap_ = true;
    // per column thread: (chunk << 1) | reversed, device copy
    int* kmap_ = nullptr;
    int kmap_chunks_ = 0;
    std::vector<int> rad_row_, rad_col_;
};

// Power-of-two test and the mixed-radix factorisation used by the generic
// (non-power-of-two) kernels.
bool is_pow2(int n);
std::vector<int> factorize(int L);

} // namespace aspher
