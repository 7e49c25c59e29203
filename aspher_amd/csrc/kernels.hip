// O(N) kernels of the certified projected CG (restating the loops of
// src/contact_solver.cpp:150-587 as fused device passes with FP64
// accumulation), the FP64 Love table (src/boussinesq_kernel.cpp:44-106), the
// dense operator and the nested-grid transfers (src/nested_solve.cpp:42-71).
//
// Every reduction writes one partial row per block (grid-stride loop in a
// fixed order, fixed shuffle tree) and is finished by `finalize`, a single
// block reducing the partials in index order: deterministic, bit-reproducible
// results. The finalize steps compute the derived scalars (approach, beta,
// step lengths, ...) on the device so the host only reads the diagnostics.
#include "kernels.hpp"
#include "reduce.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <stdexcept>

namespace aspher {

namespace {

#define GRID_LOOP(i, N)                                                          \
    for (std::int64_t i = (std::int64_t)blockIdx.x * blockDim.x + threadIdx.x;   \
         i < (N); i += (std::int64_t)gridDim.x * blockDim.x)

template <int OP>
__device__ __forceinline__ void put(double v, double* red, double* part, int nv, int idx) {
    v = block_reduce<OP>(v, red);
    if (threadIdx.x == 0) part[(std::size_t)blockIdx.x * kPartVals + idx] = v;
    (void)nv;
}

// Ensemble prologue: the element range of this block (a work item of one
// instance, or the grid-stride range of a single solve), the instance's
// vector offset, and the gate on its activity flag.
#define ENS_SETUP(N) ENS_SETUP_V(N, 1)
// VW consecutive elements per thread and step: the range boundaries (the
// vector length, instance offsets, work items) must be multiples of VW,
// which the launchers check before choosing VW > 1.
#define ENS_SETUP_V(N, VW)                                                             \
    std::int64_t i0_, i1_, is_;                                                        \
    int inst_, row_;                                                                   \
    if (ens.work) {                                                                    \
        const WorkItem w_ = ens.work[blockIdx.x];                                      \
        i0_ = w_.begin + (std::int64_t)threadIdx.x * (VW);                             \
        i1_ = w_.end;                                                                  \
        is_ = (std::int64_t)blockDim.x * (VW);                                         \
        inst_ = w_.inst;                                                               \
        row_ = w_.row;                                                                 \
    } else {                                                                           \
        i0_ = ((std::int64_t)blockIdx.x * blockDim.x + threadIdx.x) * (VW);            \
        i1_ = (N);                                                                     \
        is_ = (std::int64_t)gridDim.x * blockDim.x * (VW);                             \
        inst_ = 0;                                                                     \
        row_ = blockIdx.x;                                                             \
    }                                                                                  \
    if (ens.gate >= 0 && ens.gscal[(std::size_t)inst_ * ens.sstride + ens.gate] == 0.0) \
        return;                                                                        \
    const std::int64_t o_ = ens.off ? ens.off[inst_] : 0;                             \
    (void)o_;                                                                          \
    (void)row_
#define ENS_LOOP(i) for (std::int64_t i = i0_; i < i1_; i += is_)
#define ENS_SCAL(p) ((p) + (std::size_t)inst_ * ens.sstride)
#define ENS_PART(p) ((p) + (std::size_t)inst_ * ens.pstride + (std::size_t)row_ * kPartVals)

// VW consecutive values from an aligned address: one 16-byte access per 4
// FP32 values or 2 FP64 values, one 4-byte access per 4 mask bytes
template <int VW>
__device__ __forceinline__ void ldv(const float* p, float* v) {
    if constexpr (VW == 1) {
        v[0] = p[0];
    } else {
        static_assert(VW == 4, "vector width 1 or 4");
        const float4 q = *reinterpret_cast<const float4*>(p);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    }
}
template <int VW>
__device__ __forceinline__ void ldv(const double* p, double* v) {
    if constexpr (VW == 1) {
        v[0] = p[0];
    } else {
        static_assert(VW == 4, "vector width 1 or 4");
        const double2 a = *reinterpret_cast<const double2*>(p);
        const double2 b = *reinterpret_cast<const double2*>(p + 2);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
}
template <int VW>
__device__ __forceinline__ void stv(float* p, const float* v) {
    if constexpr (VW == 1) p[0] = v[0];
    else *reinterpret_cast<float4*>(p) = float4{v[0], v[1], v[2], v[3]};
}
template <int VW>
__device__ __forceinline__ void stv(double* p, const double* v) {
    if constexpr (VW == 1) {
        p[0] = v[0];
    } else {
        *reinterpret_cast<double2*>(p) = double2{v[0], v[1]};
        *reinterpret_cast<double2*>(p + 2) = double2{v[2], v[3]};
    }
}
// mask bytes as one word (byte k = element k)
template <int VW>
__device__ __forceinline__ unsigned ldm(const std::uint8_t* p) {
    if constexpr (VW == 1) return p[0];
    else return *reinterpret_cast<const unsigned*>(p);
}
template <int VW>
__device__ __forceinline__ void stm(std::uint8_t* p, unsigned m) {
    if constexpr (VW == 1) p[0] = (std::uint8_t)m;
    else *reinterpret_cast<unsigned*>(p) = m;
}
__device__ __forceinline__ bool mbit(unsigned m, int k) { return ((m >> (8 * k)) & 0xffu) != 0; }

template <int OP>
__device__ __forceinline__ void putr(double v, double* red, double* prow, int idx) {
    v = block_reduce<OP>(v, red);
    if (threadIdx.x == 0) prow[idx] = v;
}

// ── evaluate (src/contact_solver.cpp:260-303) ──────────────────────────────

template <typename T, int VW>
__global__ void __launch_bounds__(256)
k_eval1(std::int64_t N, const T* __restrict__ u, const T* __restrict__ g0s,
        const T* __restrict__ p, T* __restrict__ g, double* __restrict__ part, Ens ens) {
    __shared__ double red[8];
    ENS_SETUP_V(N, VW);
    u += o_; g0s += o_; p += o_; g += o_;
    double* pr = ENS_PART(part);
    double gsum = 0.0, nc = 0.0, vmin = __builtin_huge_val();
    ENS_LOOP(i) {
        T uv[VW], gv[VW], pv[VW];
        ldv<VW>(u + i, uv);
        ldv<VW>(g0s + i, gv);
        ldv<VW>(p + i, pv);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            const T v = uv[k] + gv[k];
            gv[k] = v;
            const double vd = (double)v;
            if (vd < vmin) vmin = vd;
            if (pv[k] > T(0)) {
                gsum += vd;
                nc += 1.0;
            }
        }
        stv<VW>(g + i, gv);
    }
    putr<RSUM>(gsum, red, pr, 0);
    putr<RSUM>(nc, red, pr, 1);
    putr<RMIN>(vmin, red, pr, 2);
}

template <typename T, int VW>
__global__ void __launch_bounds__(256)
k_eval2(std::int64_t N, const T* __restrict__ p, T* __restrict__ g,
        std::uint8_t* __restrict__ mask, const double* __restrict__ scal,
        double* __restrict__ part, Ens ens) {
    __shared__ double red[8];
    ENS_SETUP_V(N, VW);
    p += o_; g += o_; mask += o_;
    scal = ENS_SCAL(scal);
    double* pr = ENS_PART(part);
    const T alpha = static_cast<T>(scal[S_ALPHA]);
    const T gmin = static_cast<T>(scal[S_GMIN]);
    double pv = 0.0, gcs = 0.0, e = 0.0, Gc = 0.0, psum = 0.0;
    ENS_LOOP(i) {
        T pa[VW], ga[VW];
        ldv<VW>(p + i, pa);
        ldv<VW>(g + i, ga);
        unsigned m = 0u;
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            const T pi = pa[k];
            const double pd = (double)pi;
            T gi = ga[k];
            pv += pd * (double)gi;
            gi = gi - alpha;
            ga[k] = gi;
            const double gd = (double)gi;
            const bool c = pi > T(0);
            if (c) gcs += gd;
            e += pd * fabs(gd);
            Gc += pd * (double)(T)(gi - gmin);
            psum += pd;
            m |= (c ? 1u : 0u) << (8 * k);
        }
        stv<VW>(g + i, ga);
        stm<VW>(mask + i, m);
    }
    putr<RSUM>(pv, red, pr, 0);
    putr<RSUM>(gcs, red, pr, 1);
    putr<RSUM>(e, red, pr, 2);
    putr<RSUM>(Gc, red, pr, 3);
    putr<RSUM>(psum, red, pr, 4);
}

// ── stencil preconditioner apply_full (src/stencil_precond.cpp:63-109) ────
// Tiles of 64 x 16 outputs: the centred, masked residual of the tile plus an
// R-cell periodic halo is staged once in LDS (FP64), then every contact
// point sums its taps from LDS in the reference's tap order. A single solve
// walks the tiles grid-stride; an ensemble takes one tile per work item
// (begin = tile index, row = partial row).
constexpr int kStX = 64, kStY = 16;

// RC > 0: the radius as a compile-time constant (halo shape and its index
// arithmetic fold into constants, the staging loop unrolls and issues all
// its loads together); RC = 0: radius R at run time.
template <typename T, int RC>
__global__ void __launch_bounds__(256)
k_stencil(int Ns, const T* __restrict__ g, const std::uint8_t* __restrict__ mask,
          const double* __restrict__ scal, const int* __restrict__ tapd,
          const double* __restrict__ tapw, int ntaps, int Rrt, T* __restrict__ z,
          double* __restrict__ part, Ens ens, int zero_off) {
    __shared__ double red[8];
    extern __shared__ __align__(16) unsigned char smem[];
    double* tile = reinterpret_cast<double*>(smem);
    const int R = RC > 0 ? RC : Rrt;
    int t0, t1, ts, inst = 0, row = blockIdx.x;
    if (ens.work) {
        const WorkItem w = ens.work[blockIdx.x];
        t0 = (int)w.begin;
        t1 = (int)w.end;
        ts = 1;
        inst = w.inst;
        row = w.row;
        Ns = ens.ns[inst];
    } else {
        t0 = blockIdx.x;
        t1 = ((Ns + kStX - 1) / kStX) * ((Ns + kStY - 1) / kStY);
        ts = gridDim.x;
    }
    if (ens.gate >= 0 && ens.gscal[(std::size_t)inst * ens.sstride + ens.gate] == 0.0) return;
    const std::int64_t o = ens.off ? ens.off[inst] : 0;
    g += o; mask += o; z += o;
    scal += (std::size_t)inst * ens.sstride;
    double* pr = part + (std::size_t)inst * ens.pstride + (std::size_t)row * kPartVals;
    const int HX = kStX + 2 * R, HY = kStY + 2 * R, HN = HX * HY;
    const float inv_hx = 1.0f / (float)HX;
    const double gmean = scal[S_GMEAN];
    const int ntx = (Ns + kStX - 1) / kStX;
    const int tx = threadIdx.x & 63, ty0 = threadIdx.x >> 6;
    double zsum = 0.0;
    // halo element e of tile t: the grid index it reads, or -1 where no grid
    // point of the tile reads it (a tile may overhang a grid smaller than
    // itself). Within the read part the periodic wrap is one add or
    // subtract, since 2R + 1 <= Ns.
    auto halo_index = [&](int t, int e, int hy) -> std::int64_t {
        const int x0 = (t % ntx) * kStX, y0 = (t / ntx) * kStY;
        const int hxv = min(kStX, Ns - x0) + 2 * R, hyv = min(kStY, Ns - y0) + 2 * R;
        const int hx = e - hy * HX;
        if (hx >= hxv || hy >= hyv) return -1;
        int jx = x0 - R + hx, jy = y0 - R + hy;
        jx += jx < 0 ? Ns : 0;
        jx -= jx >= Ns ? Ns : 0;
        jy += jy < 0 ? Ns : 0;
        jy -= jy >= Ns ? Ns : 0;
        return (std::int64_t)jy * Ns + jx;
    };
    // taps: for a compile-time radius the l2 disc is enumerated here in the
    // host's order (dy outer, dx inner, both ascending; the host checks the
    // tap count), so the LDS offsets are constants of the unrolled loop; the
    // weights are uniform (scalar registers for R <= 2, LDS broadcast
    // reads otherwise). The summation order is the tap order either way.
    constexpr int NT = RC == 1 ? 5 : RC == 2 ? 13 : RC == 3 ? 29 : RC == 4 ? 49 : 1;
    __shared__ double s_w[RC >= 3 ? NT : 1];
    // the tile's own mask bytes, staged with its halo
    __shared__ std::uint8_t s_cm[kStY * kStX];
    double tw_r[RC >= 1 && RC <= 2 ? NT : 1];
    if constexpr (RC >= 1 && RC <= 2) {
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const double wq = tapw[q];
            const unsigned long long b = __double_as_longlong(wq);
            const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(b & 0xffffffffu));
            const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
            tw_r[q] = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
        }
    } else if constexpr (RC >= 3) {
        if ((int)threadIdx.x < NT) s_w[threadIdx.x] = tapw[threadIdx.x];
        // the first tile's barrier orders these writes before any read
    }
    auto compute_tile = [&](int t) {
        const int x0 = (t % ntx) * kStX, y0 = (t / ntx) * kStY;
        const int ix = x0 + tx;
#pragma unroll
        for (int k = 0; k < kStY / 4; ++k) {
            const int ty = ty0 + 4 * k, iy = y0 + ty;
            if (ix >= Ns || iy >= Ns) continue;
            const std::size_t i = (std::size_t)iy * Ns + ix;
            if (!s_cm[ty * kStX + tx]) {
                // z off the mask is zero; a caller that reads z only on the
                // mask skips these stores
                if (zero_off) z[i] = T(0);
                continue;
            }
            double acc = 0.0;
            const double* c = tile + (ty + R) * HX + (tx + R);
            if constexpr (RC > 0) {
                constexpr int HXc = kStX + 2 * RC;
                int q = 0;
#pragma unroll
                for (int dy = -RC; dy <= RC; ++dy)
#pragma unroll
                    for (int dx = -RC; dx <= RC; ++dx) {
                        if (dx * dx + dy * dy > RC * RC) continue;
                        const double wq = RC <= 2 ? tw_r[RC <= 2 ? q : 0] : s_w[RC >= 3 ? q : 0];
                        acc += wq * c[dy * HXc + dx];
                        ++q;
                    }
            } else {
                for (int q = 0; q < ntaps; ++q)
                    acc += tapw[q] * c[tapd[2 * q + 1] * HX + tapd[2 * q]];
            }
            z[i] = static_cast<T>(acc);
            zsum += acc;
        }
    };
    // Tiles without a contact point have no output on the mask and add
    // nothing to zsum; at a few percent contact most tiles are such (83% of
    // them at 6.9% contact on a 4096^2 rough surface), so each tile
    // is first tested on its 64 x 16 mask bytes, and only tiles with a
    // contact point fetch their halo. Each thread tests 4 consecutive bytes
    // of one tile row: one 4-byte load where Ns is a multiple of 4 (a
    // block-uniform choice, so the loads of several tiles go out together).
    const bool words = (Ns & 3) == 0 && (reinterpret_cast<std::uintptr_t>(mask) & 3u) == 0;
    // this thread's 4 mask bytes of tile t (t < t1): grid offset, or -1
    auto tile_at = [&](int t) -> std::int64_t {
        const int x0 = (t % ntx) * kStX, y0 = (t / ntx) * kStY;
        const int iy = y0 + (int)(threadIdx.x >> 4), cx = x0 + 4 * (int)(threadIdx.x & 15);
        return (iy < Ns && cx < Ns) ? (std::int64_t)iy * Ns + cx : -1;
    };
    // Ns a multiple of 4: the 4 bytes are one aligned word, loaded without a
    // branch (offset 0 stands in for tiles past t1 and points off the grid)
    auto tile_word = [&](int t) -> bool {
        const std::int64_t at = t < t1 ? tile_at(t) : -1;
        const unsigned w = *reinterpret_cast<const unsigned*>(mask + (at >= 0 ? at : 0));
        return at >= 0 && w != 0u;
    };
    auto tile_bytes = [&](int t) -> bool {
        const std::int64_t at = t < t1 ? tile_at(t) : -1;
        if (at < 0) return false;
        const int cx = (int)(at % Ns);
        bool a = false;
        for (int j = 0; j < 4 && cx + j < Ns; ++j) a = a || mask[at + j] != 0;
        return a;
    };
    // an empty tile's outputs, for callers that read z everywhere
    auto zero_tile = [&](int t) {
        const int x0 = (t % ntx) * kStX, y0 = (t / ntx) * kStY;
        const int ix = x0 + tx;
#pragma unroll
        for (int k = 0; k < kStY / 4; ++k) {
            const int iy = y0 + ty0 + 4 * k;
            if (ix < Ns && iy < Ns) z[(std::size_t)iy * Ns + ix] = T(0);
        }
    };
    // the masked, centred residual of a tile and its halo, and the tile's
    // mask bytes, into LDS: RC > 0 from values fetched into registers
    // beforehand (so the next tile's loads overlap this tile's compute)
    constexpr int HXc = kStX + 2 * (RC > 0 ? RC : 0), HNc = HXc * (kStY + 2 * (RC > 0 ? RC : 0));
    constexpr int K = (HNc + 255) / 256;
    auto fetch = [&](int t, T* gv, std::uint8_t* mv) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int e = threadIdx.x + 256 * k;
            const std::int64_t j = e < HNc ? halo_index(t, e, e / HXc) : -1;
            gv[k] = j >= 0 ? g[j] : T(0);
            mv[k] = j >= 0 ? mask[j] : std::uint8_t(0);
        }
    };
    auto put = [&](const T* gv, const std::uint8_t* mv) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int e = threadIdx.x + 256 * k;
            if (e >= HNc) continue;
            tile[e] = mv[k] ? (double)gv[k] - gmean : 0.0;
            const int hy = e / HXc, hx = e - hy * HXc;
            if (hy >= RC && hy < RC + kStY && hx >= RC && hx < RC + kStX)
                s_cm[(hy - RC) * kStX + (hx - RC)] = mv[k];
        }
    };
    auto stage_generic = [&](int t) {
        // e / HX by the rounded reciprocal: exact here, since (e + 0.5)/HX
        // is at least 0.5/HX away from an integer, far above float rounding
        for (int e = threadIdx.x; e < HN; e += 256) {
            const int hy = (int)(((float)e + 0.5f) * inv_hx), hx = e - hy * HX;
            const std::int64_t j = halo_index(t, e, hy);
            if (j < 0) continue;
            const std::uint8_t mj = mask[j];
            tile[e] = mj ? (double)g[j] - gmean : 0.0;
            if (hy >= R && hy < R + kStY && hx >= R && hx < R + kStX)
                s_cm[(hy - R) * kStX + (hx - R)] = mj;
        }
    };
    constexpr int kAhead = 8;   // tiles tested per batch of mask loads
    for (int tb = t0; tb < t1; tb += kAhead * ts) {
        bool has[kAhead];
        if (words) {
#pragma unroll
            for (int q = 0; q < kAhead; ++q) has[q] = tile_word(tb + q * ts);
        } else {
#pragma unroll
            for (int q = 0; q < kAhead; ++q) has[q] = tile_bytes(tb + q * ts);
        }
        // the batch's tiles with a contact point (block-uniform); these
        // barriers also order the previous batch's LDS reads before this
        // batch's writes
        unsigned live = 0u;
#pragma unroll
        for (int q = 0; q < kAhead; ++q)
            if (__syncthreads_or(has[q])) live |= 1u << q;
        if (zero_off) {
#pragma unroll
            for (int q = 0; q < kAhead; ++q) {
                const int t = tb + q * ts;
                if (t < t1 && !((live >> q) & 1u)) zero_tile(t);
            }
        }
        int q = live ? __builtin_ctz(live) : kAhead;
        if constexpr (RC > 0) {
            T gv[K];
            std::uint8_t mv[K];
            if (q < kAhead) fetch(tb + q * ts, gv, mv);
            while (q < kAhead) {
                const unsigned rest = live & ~((2u << q) - 1u);
                const int qn = rest ? __builtin_ctz(rest) : kAhead;
                __syncthreads();
                put(gv, mv);
                if (qn < kAhead) fetch(tb + qn * ts, gv, mv);
                __syncthreads();
                compute_tile(tb + q * ts);
                q = qn;
            }
        } else {
            while (q < kAhead) {
                const unsigned rest = live & ~((2u << q) - 1u);
                __syncthreads();
                stage_generic(tb + q * ts);
                __syncthreads();
                compute_tile(tb + q * ts);
                q = rest ? __builtin_ctz(rest) : kAhead;
            }
        }
    }
    putr<RSUM>(zsum, red, pr, 0);
}

// ── CG direction (src/contact_solver.cpp:411-440) ──────────────────────────

template <typename T, bool PC, int VW>
__global__ void __launch_bounds__(256)
k_cgdot(std::int64_t N, const T* __restrict__ g, const T* __restrict__ gprev,
        const T* __restrict__ z, const std::uint8_t* __restrict__ mask,
        const double* __restrict__ scal, double* __restrict__ part, Ens ens) {
    __shared__ double red[8];
    ENS_SETUP_V(N, VW);
    g += o_; gprev += o_; mask += o_;
    if (PC) z += o_;
    scal = ENS_SCAL(scal);
    double* pr = ENS_PART(part);
    const T zmean = PC ? static_cast<T>(scal[S_ZMEAN]) : T(0);
    double G = 0.0, Gpr = 0.0, gg = 0.0;
    ENS_LOOP(i) {
        const unsigned m = ldm<VW>(mask + i);
        if (m == 0u) continue;
        T ga[VW], gpa[VW], za[VW];
        ldv<VW>(g + i, ga);
        ldv<VW>(gprev + i, gpa);
        if (PC) ldv<VW>(z + i, za);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            if (!mbit(m, k)) continue;
            const T gi = ga[k];
            const T zi = PC ? T(za[k] - zmean) : gi;
            const double gd = (double)gi, zd = (double)zi;
            G += zd * gd;
            Gpr += zd * (double)(T)(gi - gpa[k]);
            gg += gd * gd;
        }
    }
    putr<RSUM>(G, red, pr, 0);
    putr<RSUM>(Gpr, red, pr, 1);
    putr<RSUM>(gg, red, pr, 2);
}

template <typename T, bool PC, int VW>
__global__ void __launch_bounds__(256)
k_tupdate(std::int64_t N, const T* __restrict__ g, T* __restrict__ gprev,
          const T* __restrict__ z, const std::uint8_t* __restrict__ mask,
          T* __restrict__ t, const double* __restrict__ scal, Ens ens) {
    ENS_SETUP_V(N, VW);
    g += o_; mask += o_; t += o_;
    if (gprev) gprev += o_;
    if (PC) z += o_;
    scal = ENS_SCAL(scal);
    const T zmean = PC ? static_cast<T>(scal[S_ZMEAN]) : T(0);
    const T beta = static_cast<T>(scal[S_BETA]);
    // gprev == nullptr: the caller keeps g_prev by rotating buffers (no copy)
    const bool copy = gprev != nullptr;
    ENS_LOOP(i) {
        const unsigned m = ldm<VW>(mask + i);
        T ta[VW];
        if (m == 0u) {
#pragma unroll
            for (int k = 0; k < VW; ++k) ta[k] = T(0);
        } else {
            T za[VW];
            ldv<VW>(t + i, ta);
            if (PC) ldv<VW>(z + i, za);
            else ldv<VW>(g + i, za);
#pragma unroll
            for (int k = 0; k < VW; ++k) {
                const T zi = PC ? T(za[k] - zmean) : za[k];
                ta[k] = mbit(m, k) ? T(zi + beta * ta[k]) : T(0);
            }
        }
        stv<VW>(t + i, ta);
        if (copy) {
            T ga[VW];
            ldv<VW>(g + i, ga);
            stv<VW>(gprev + i, ga);
        }
    }
}

// ── exact line search (src/contact_solver.cpp:446-459) ─────────────────────

template <typename T, int VW>
__global__ void __launch_bounds__(256)
k_line1(std::int64_t N, const T* __restrict__ r, const T* __restrict__ g,
        const T* __restrict__ t, const std::uint8_t* __restrict__ mask,
        double* __restrict__ part, Ens ens) {
    __shared__ double red[8];
    ENS_SETUP_V(N, VW);
    r += o_; g += o_; t += o_; mask += o_;
    double* pr = ENS_PART(part);
    double rsum = 0.0, num = 0.0;
    ENS_LOOP(i) {
        const unsigned m = ldm<VW>(mask + i);
        if (m == 0u) continue;
        T ra[VW], ga[VW], ta[VW];
        ldv<VW>(r + i, ra);
        ldv<VW>(g + i, ga);
        ldv<VW>(t + i, ta);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            if (!mbit(m, k)) continue;
            rsum += (double)ra[k];
            num += (double)ga[k] * (double)ta[k];
        }
    }
    putr<RSUM>(rsum, red, pr, 0);
    putr<RSUM>(num, red, pr, 1);
}

template <typename T, int VW>
__global__ void __launch_bounds__(256)
k_line2(std::int64_t N, const T* __restrict__ r, const T* __restrict__ t,
        const std::uint8_t* __restrict__ mask, const double* __restrict__ scal,
        double* __restrict__ part, Ens ens) {
    __shared__ double red[8];
    ENS_SETUP_V(N, VW);
    r += o_; t += o_; mask += o_;
    scal = ENS_SCAL(scal);
    double* pr = ENS_PART(part);
    const T rmean = static_cast<T>(scal[S_RMEAN]);
    double den = 0.0;
    ENS_LOOP(i) {
        const unsigned m = ldm<VW>(mask + i);
        if (m == 0u) continue;
        T ra[VW], ta[VW];
        ldv<VW>(r + i, ra);
        ldv<VW>(t + i, ta);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            if (!mbit(m, k)) continue;
            den += (double)(T)(ra[k] - rmean) * (double)ta[k];
        }
    }
    putr<RSUM>(den, red, pr, 0);
}

// ── update, overlap step, load rescale (src/contact_solver.cpp:477-501) ────

// best != nullptr: where the instance's scalar best_flag is set, the
// incoming pressure is also stored to best (the best iterate, marked by the
// evaluation; the pressure is not written between that and this update).
template <typename T, int VW>
__global__ void __launch_bounds__(256)
k_update(std::int64_t N, const T* pin, T* pout, const T* __restrict__ t,
         const T* __restrict__ g, const double* __restrict__ scal,
         double* __restrict__ part, Ens ens, T* __restrict__ best, int best_flag) {
    __shared__ double red[8];
    ENS_SETUP_V(N, VW);
    pin += o_; pout += o_; t += o_; g += o_;
    scal = ENS_SCAL(scal);
    double* pr = ENS_PART(part);
    const bool save = best != nullptr && scal[best_flag] != 0.0;
    if (save) best += o_;
    if (scal[S_OK] == 0.0) {
        // no usable curvature: the pressure is unchanged
        if (pout != pin || save) {
            ENS_LOOP(i) {
                T pa[VW];
                ldv<VW>(pin + i, pa);
                if (pout != pin) stv<VW>(pout + i, pa);
                if (save) stv<VW>(best + i, pa);
            }
        }
        if (threadIdx.x == 0) {
            pr[0] = 0.0;
            pr[1] = 0.0;
        }
        return;
    }
    const T tau = static_cast<T>(scal[S_TAU]);
    const T tau_ov = static_cast<T>(scal[S_TAUOV]);
    double changed = 0.0, total = 0.0;
    ENS_LOOP(i) {
        T pa[VW], ga[VW], ta[VW];
        ldv<VW>(pin + i, pa);
        ldv<VW>(g + i, ga);
        if (save) stv<VW>(best + i, pa);
        bool any = false;
#pragma unroll
        for (int k = 0; k < VW; ++k) any = any || pa[k] > T(0);
        // t is zero off the contact set: read only where some p > 0
        if (any) ldv<VW>(t + i, ta);
#pragma unroll
        for (int k = 0; k < VW; ++k) {
            T pi = pa[k];
            const bool was = pi > T(0);
            if (was) {
                const T q = pi - tau * ta[k];
                pi = (q < T(0)) ? T(0) : q;  // std::max(q, 0): NaN propagates
            }
            const T gi = ga[k];
            if (pi == T(0) && gi < T(0)) pi = pi - tau_ov * gi;
            if ((pi > T(0)) != was) changed = 1.0;
            pa[k] = pi;
            total += (double)pi;
        }
        stv<VW>(pout + i, pa);
    }
    putr<RMAX>(changed, red, pr, 0);
    putr<RSUM>(total, red, pr, 1);
}

template <typename T, int VW>
__global__ void __launch_bounds__(256)
k_rescale(std::int64_t N, T* __restrict__ p, const double* __restrict__ scal, Ens ens) {
    ENS_SETUP_V(N, VW);
    p += o_;
    scal = ENS_SCAL(scal);
    if (scal[S_OK] == 0.0) return;
    const bool fill = scal[S_PFILL] != 0.0;
    const T f = static_cast<T>(scal[S_F]);
    const T pbar = static_cast<T>(scal[S_PBAR]);
    ENS_LOOP(i) {
        T pa[VW];
        if (fill) {
#pragma unroll
            for (int k = 0; k < VW; ++k) pa[k] = pbar;
        } else {
            ldv<VW>(p + i, pa);
#pragma unroll
            for (int k = 0; k < VW; ++k) pa[k] = T(pa[k] * f);
        }
        stv<VW>(p + i, pa);
    }
}

// ── identification step and simplex projection (src/contact_solver.cpp:31-60,314-364)

template <typename T>
__global__ void __launch_bounds__(256)
k_objective(std::int64_t N, const T* __restrict__ p, const T* __restrict__ u,
            const T* __restrict__ g0s, double* __restrict__ part) {
    __shared__ double red[8];
    double f = 0.0, pp = 0.0, pu = 0.0;
    GRID_LOOP(i, N) {
        const double pd = (double)p[i], ud = (double)u[i];
        f += pd * (0.5 * ud + (double)g0s[i]);
        pp += pd * pd;
        pu += pd * ud;
    }
    put<RSUM>(f, red, part, 3, 0);
    put<RSUM>(pp, red, part, 3, 1);
    put<RSUM>(pu, red, part, 3, 2);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_axpy_id(std::int64_t N, const T* __restrict__ p, const T* __restrict__ g,
          T* __restrict__ z, T s) {
    GRID_LOOP(i, N) z[i] = p[i] - s * g[i];
}

template <typename T>
__global__ void __launch_bounds__(256)
k_sum(std::int64_t N, const T* __restrict__ y, double* __restrict__ part) {
    __shared__ double red[8];
    double s = 0.0;
    GRID_LOOP(i, N) s += (double)y[i];
    put<RSUM>(s, red, part, 1, 0);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_moments(std::int64_t N, const T* __restrict__ y, double* __restrict__ part) {
    __shared__ double red[8];
    double s = 0.0, q = 0.0;
    GRID_LOOP(i, N) {
        const double v = (double)y[i];
        s += v;
        q += v * v;
    }
    put<RSUM>(s, red, part, 2, 0);
    put<RSUM>(q, red, part, 2, 1);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_count_above(std::int64_t N, const T* __restrict__ y, double theta,
              double* __restrict__ part) {
    __shared__ double red[8];
    double s = 0.0, c = 0.0;
    GRID_LOOP(i, N) {
        const double v = (double)y[i];
        if (v > theta) {
            s += v;
            c += 1.0;
        }
    }
    put<RSUM>(s, red, part, 2, 0);
    put<RSUM>(c, red, part, 2, 1);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_project(std::int64_t N, T* __restrict__ y, T th) {
    GRID_LOOP(i, N) {
        const T v = y[i];
        y[i] = (v > th) ? T(v - th) : T(0);
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_idcheck(std::int64_t N, const T* __restrict__ z, const T* __restrict__ r,
          const T* __restrict__ g0s, const T* __restrict__ g, const T* __restrict__ p,
          double* __restrict__ part) {
    __shared__ double red[8];
    double f = 0.0, gd = 0.0;
    GRID_LOOP(i, N) {
        const double zd = (double)z[i];
        f += zd * (0.5 * (double)r[i] + (double)g0s[i]);
        gd += (double)g[i] * (double)(T)(z[i] - p[i]);
    }
    put<RSUM>(f, red, part, 2, 0);
    put<RSUM>(gd, red, part, 2, 1);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_stats(std::int64_t N, const T* __restrict__ y, double* __restrict__ part) {
    __shared__ double red[8];
    double mn = __builtin_huge_val(), mx = -__builtin_huge_val(), bad = 0.0;
    GRID_LOOP(i, N) {
        const double v = (double)y[i];
        if (!isfinite(v)) bad += 1.0;
        if (v < mn) mn = v;
        if (v > mx) mx = v;
    }
    put<RMIN>(mn, red, part, 3, 0);
    put<RMAX>(mx, red, part, 3, 1);
    put<RSUM>(bad, red, part, 3, 2);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_clamp_sum(std::int64_t N, T* __restrict__ p, double* __restrict__ part) {
    __shared__ double red[8];
    double s = 0.0;
    GRID_LOOP(i, N) {
        const T v = p[i];
        const T c = (v < T(0)) ? T(0) : v;   // cwiseMax(0)
        p[i] = c;
        s += (double)c;
    }
    put<RSUM>(s, red, part, 1, 0);
}

template <typename T>
__global__ void __launch_bounds__(256) k_scale(std::int64_t N, T* __restrict__ p, T f) {
    GRID_LOOP(i, N) p[i] = p[i] * f;
}

template <typename T>
__global__ void __launch_bounds__(256) k_fill(std::int64_t N, T* __restrict__ p, T v) {
    GRID_LOOP(i, N) p[i] = v;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_from_double(std::int64_t N, const double* __restrict__ x, double shift, T* __restrict__ y) {
    GRID_LOOP(i, N) y[i] = static_cast<T>(x[i] - shift);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_masked_sum(std::int64_t N, const T* __restrict__ y, const std::uint8_t* __restrict__ mask,
             double* __restrict__ part) {
    __shared__ double red[8];
    double s = 0.0, c = 0.0;
    GRID_LOOP(i, N) {
        if (!mask[i]) continue;
        s += (double)y[i];
        c += 1.0;
    }
    put<RSUM>(s, red, part, 2, 0);
    put<RSUM>(c, red, part, 2, 1);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_masked_sub(std::int64_t N, T* __restrict__ y, const std::uint8_t* __restrict__ mask,
             const double* __restrict__ scal, int idx) {
    const T c = static_cast<T>(scal[idx]);
    GRID_LOOP(i, N) if (mask[i]) y[i] = y[i] - c;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_shift(std::int64_t N, const T* __restrict__ x, T c, T* __restrict__ y) {
    GRID_LOOP(i, N) y[i] = x[i] - c;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_to_double(std::int64_t N, const T* __restrict__ x, double* __restrict__ y) {
    GRID_LOOP(i, N) y[i] = (double)x[i];
}

template <typename T>
__global__ void __launch_bounds__(256)
k_dense(int Ns, const double* __restrict__ table, const T* __restrict__ x, T* __restrict__ y) {
    const std::int64_t N = (std::int64_t)Ns * Ns;
    GRID_LOOP(i, N) {
        const int iy = (int)(i / Ns), ix = (int)(i - (std::int64_t)iy * Ns);
        double acc = 0.0;
        for (int jy = 0; jy < Ns; ++jy) {
            const double* trow = table + (std::size_t)abs(iy - jy) * Ns;
            const T* xr = x + (std::size_t)jy * Ns;
            for (int jx = 0; jx < Ns; ++jx) acc += trow[abs(ix - jx)] * (double)xr[jx];
        }
        y[i] = static_cast<T>(acc);
    }
}

// ── finalize: partials -> scalars, then the derived-scalar step ────────────


// Ensemble iteration control after a fresh evaluation: the loop body of
// src/contact_solver.cpp:366-407 (convergence, best iterate, stall) decided
// on the device, so an instance leaves the replayed iteration by itself.
// Branches that need the host (identification step after a failed line
// search or with fewer than two contact points, stall, iteration budget)
// clear S_ACTIVE and record the reason in S_HOST.
__device__ void ens_control(double* s) {
    const bool pend = s[S_PENDING] != 0.0;
    s[S_PENDING] = 0.0;
    s[S_SAVEBEST] = 0.0;
    if (pend && s[S_OK] == 0.0) {
        s[S_ACTIVE] = 0.0;
        s[S_HOST] = HOST_IDENT;
        return;
    }
    if (s[S_FINITE] == 0.0) {
        s[S_ACTIVE] = 0.0;
        s[S_HOST] = HOST_NONFINITE;
        return;
    }
    const double fw = s[S_FW], pen = s[S_PEN];
    if (s[S_LOAD] <= s[S_LOADTOL] && fw <= s[S_TOL] && pen <= s[S_TOL]) {
        s[S_ACTIVE] = 0.0;
        s[S_STATUS] = 1.0;   // converged
        return;
    }
    const double merit = fw > pen ? fw : pen;
    if (s[S_KEEPBEST] != 0.0 && merit < s[S_BEST]) {
        s[S_BEST] = merit;
        s[S_SAVEBEST] = 1.0;
    }
    if (merit < s[S_SEEN] * (1.0 - 1e-4)) {
        s[S_SEEN] = merit;
        s[S_STALL] = 0.0;
    } else {
        s[S_STALL] += 1.0;
        if (s[S_STALL] >= s[S_STALL_LIMIT]) {
            s[S_ACTIVE] = 0.0;
            s[S_HOST] = HOST_STALL;
            return;
        }
    }
    if (s[S_ITER] >= s[S_MAXIT]) {
        s[S_ACTIVE] = 0.0;
        s[S_HOST] = HOST_MAXIT;
        return;
    }
    if (s[S_NC] < 2.0) {
        s[S_ACTIVE] = 0.0;
        s[S_HOST] = HOST_IDENT;
        return;
    }
    s[S_ITER] += 1.0;
    s[S_PENDING] = 1.0;
}

// Single-solve iteration control after a fresh evaluation (the device loop
// of solve_impl): the iteration continues on the device only where the host
// loop (src/contact_solver.cpp:366-407 and the CG branch) would take the plain
// CG branch, and then does that branch's bookkeeping (best iterate, stall
// counter, iteration count). Every other branch -- identification after a
// failed line search, non-finite values, convergence, the stall limit, the
// last iteration of the budget, fewer than two contact points -- clears
// S_ACTIVE before any bookkeeping, so the host loop redoes the iteration's
// decisions from the state of the previous one.
__device__ void single_control(double* s) {
    const double fw = s[S_FW], pen = s[S_PEN];
    const double merit = fw > pen ? fw : pen;
    const bool improves = merit < s[S_SEEN] * (1.0 - 1e-4);
    s[S_SAVEBEST] = 0.0;
    bool go = false;
    if (s[S_PENDING] != 0.0 && s[S_OK] == 0.0) {
        s[S_HOST] = HOST_IDENT;
    } else if (s[S_FINITE] == 0.0) {
        s[S_HOST] = HOST_NONFINITE;
    } else if (s[S_LOAD] <= s[S_LOADTOL] && fw <= s[S_TOL] && pen <= s[S_TOL]) {
        s[S_STATUS] = 1.0;   // converged
    } else if (!improves && s[S_STALL] + 1.0 >= s[S_STALL_LIMIT]) {
        s[S_HOST] = HOST_STALL;
    } else if (s[S_ITER] + 1.0 >= s[S_MAXIT]) {
        s[S_HOST] = HOST_MAXIT;
    } else if (s[S_NC] < 2.0) {
        s[S_HOST] = HOST_IDENT;
    } else {
        go = true;
    }
    if (!go) {
        s[S_ACTIVE] = 0.0;
        return;
    }
    if (s[S_KEEPBEST] != 0.0 && merit < s[S_BEST]) {
        s[S_BEST] = merit;
        s[S_SAVEBEST] = 1.0;
    }
    if (improves) {
        s[S_SEEN] = merit;
        s[S_STALL] = 0.0;
    } else {
        s[S_STALL] += 1.0;
    }
    s[S_ITER] += 1.0;
    s[S_PENDING] = 1.0;
}

template <typename T>
__device__ void run_step(int step, double* s) {
    const double nc = s[S_NC];
    const double P = s[S_PTOT], gref = s[S_GREF];
    switch (step) {
        case STEP_EVAL1: {
            const double alpha = nc > 0.0 ? s[S_GSUM] / nc : s[S_VMIN];
            s[S_ALPHA] = alpha;
            s[S_GMIN] = s[S_VMIN] - alpha;
            break;
        }
        case STEP_EVAL2:
        case STEP_B_EVAL2:
        case STEP_S_EVAL2: {
            double a = s[S_ALPHA], gm = s[S_GMIN];
            if (nc > 0.0) {  // refine_contact_mean (src/contact_solver.cpp:94-99)
                const double delta = s[S_GCS] / nc;
                a += delta;
                gm -= delta;
            }
            s[S_ALPHA_REF] = a;
            s[S_GMIN_REF] = gm;
            s[S_PK] = s[S_E] / (P * gref);
            s[S_FW] = s[S_GC] / (P * gref);
            s[S_PEN] = ((-gm > 0.0) ? -gm : 0.0) / gref;
            s[S_LOAD] = fabs(s[S_PSUM] - P) / P;
            s[S_FINITE] = (isfinite(s[S_GSUM]) && isfinite(s[S_VMIN]) &&
                           isfinite(s[S_PSUM]) && isfinite(s[S_E]) && isfinite(s[S_GC]))
                              ? 1.0 : 0.0;
            s[S_GMEAN] = nc > 0.0 ? s[S_GCS] / nc : 0.0;
            if (step == STEP_B_EVAL2) ens_control(s);
            if (step == STEP_S_EVAL2) single_control(s);
            break;
        }
        case STEP_ZMEAN:
            s[S_ZMEAN] = nc > 0.0 ? s[S_ZSUM] / nc : 0.0;
            break;
        case STEP_CGDOT: {
            const double Gold = s[S_GOLD];
            double bv = 0.0;
            if (Gold != 0.0) {
                if (s[S_USEPR] != 0.0) {
                    const double q = s[S_GPR] / Gold;
                    bv = (0.0 < q) ? q : 0.0;   // std::max(0.0, q)
                } else {
                    bv = s[S_G] / Gold;
                }
            }
            s[S_BETA] = (double)(static_cast<T>(s[S_DELTA]) * static_cast<T>(bv));
            s[S_GOLD] = s[S_G];
            break;
        }
        case STEP_LINE1:
            s[S_RMEAN] = (double)static_cast<T>(nc > 0.0 ? s[S_RSUM] / nc : 0.0);
            break;
        case STEP_LINE2: {
            const double den = s[S_DEN], num = s[S_NUM];
            const bool ok = den > 0.0 && isfinite(den) && isfinite(num);
            s[S_OK] = ok ? 1.0 : 0.0;
            if (ok) {
                const T tau = static_cast<T>(num / den);
                const double gg = s[S_GG];
                s[S_TAU] = (double)tau;
                s[S_TAUOV] = gg > 0.0 ? (double)static_cast<T>(num * num / (den * gg))
                                      : (double)tau;
            }
            break;
        }
        case STEP_UPDATE:
            if (s[S_OK] != 0.0) {
                s[S_DELTA] = s[S_CHANGED] != 0.0 ? 0.0 : 1.0;
                const double total = s[S_TOTAL];
                if (total > 0.0) {
                    s[S_F] = (double)static_cast<T>(P / total);
                    s[S_PFILL] = 0.0;
                } else {
                    s[S_PFILL] = 1.0;
                }
            }
            break;
        case STEP_ADJ_INIT:
            s[S_GMEAN] = nc > 0.0 ? s[S_A_WSUM] / nc : 0.0;
            s[S_A_RZOLD] = 0.0;
            s[S_A_BETA] = 0.0;
            s[S_A_OK] = 1.0;
            break;
        case STEP_ADJ_DOT: {
            const double rz = s[S_A_RZ], old = s[S_A_RZOLD];
            s[S_A_BETA] = old != 0.0 ? rz / old : 0.0;
            s[S_A_RZOLD] = rz;
            break;
        }
        case STEP_ADJ_CURV: {
            const double qmean = nc > 0.0 ? s[S_A_QSUM] / nc : 0.0;
            const double den = s[S_A_QT] - qmean * s[S_A_TSUM];
            const double rz = s[S_A_RZOLD];
            const bool ok = den > 0.0 && isfinite(den) && isfinite(rz);
            s[S_A_QMEAN] = qmean;
            s[S_A_OK] = ok ? 1.0 : 0.0;
            s[S_A_ALPHA] = ok ? rz / den : 0.0;
            break;
        }
        case STEP_ADJ_RMEAN:
            s[S_GMEAN] = nc > 0.0 ? s[S_A_RSUM] / nc : 0.0;
            break;
        case STEP_ADJ_FINISH:
            s[S_A_NU] = nc > 0.0 ? s[S_A_WQSUM] / nc : 0.0;
            break;
        default:
            break;
    }
}

struct FinSpec {
    int nv;
    int ops[kPartVals];
    int dst[kPartVals];
};

// NV: the number of values (spec.nv) as a compile-time constant, so every
// load of a batch is unconditional and the whole batch is in flight at once
// (a runtime count left one memory round trip per value). SUM: every value
// is a sum, so the per-value choice of operation is not compiled in.
template <typename T, int NV, bool SUM>
__global__ void __launch_bounds__(256)
k_finalize(const double* __restrict__ part, int nblk, FinSpec spec, double* __restrict__ scal,
           int step, Ens ens) {
    __shared__ double red[8];
    // one block per instance (a single solve: one block)
    const int inst = blockIdx.x;
    if (ens.gate >= 0 && ens.gscal[(std::size_t)inst * ens.sstride + ens.gate] == 0.0) return;
    part += (std::size_t)inst * ens.pstride;
    scal += (std::size_t)inst * ens.sstride;
    if (ens.rows) nblk = ens.rows[inst];
    // one pass over the partial rows for all values, then one combination
    // of the waves: per value the same order as a reduction of that value
    // alone (thread-strided rows, shuffle tree, waves in index order)
    constexpr int nv = NV;
    double acc[kPartVals];
#pragma unroll
    for (int v = 0; v < kPartVals; ++v)
        acc[v] = (SUM || spec.ops[v] == RSUM) ? 0.0
                 : (spec.ops[v] == RMIN ? __builtin_huge_val() : -__builtin_huge_val());
    // rows b, b + 256, ... in that order; eight rows are loaded before any is
    // accumulated, so a batch costs one memory latency, not eight
    constexpr int kBatch = 8;
    for (int b0 = threadIdx.x; b0 < nblk; b0 += kBatch * blockDim.x) {
        double x[kBatch][kPartVals];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int b = b0 + u * (int)blockDim.x;
            const double* row = part + (std::size_t)(b < nblk ? b : 0) * kPartVals;
#pragma unroll
            for (int v = 0; v < kPartVals; ++v) x[u][v] = v < nv ? row[v] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            if (b0 + u * (int)blockDim.x >= nblk) break;
#pragma unroll
            for (int v = 0; v < kPartVals; ++v) {
                if (v >= nv) break;
                const double xv = x[u][v];
                const int op = SUM ? RSUM : spec.ops[v];
                acc[v] = op == RSUM ? acc[v] + xv
                         : (op == RMIN ? ((xv < acc[v]) ? xv : acc[v])
                                       : ((xv > acc[v]) ? xv : acc[v]));
            }
        }
    }
    __shared__ double wred[8 * kPartVals];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int v = 0; v < kPartVals; ++v) {
        if (v >= nv) break;
        const int op = SUM ? RSUM : spec.ops[v];
        acc[v] = op == RSUM ? wave_reduce<RSUM>(acc[v])
                 : (op == RMIN ? wave_reduce<RMIN>(acc[v]) : wave_reduce<RMAX>(acc[v]));
        if (lane == 0) wred[wave * kPartVals + v] = acc[v];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int v = 0; v < nv; ++v) {
            const int op = SUM ? RSUM : spec.ops[v];
            double r = op == RSUM ? 0.0 : (op == RMIN ? __builtin_huge_val() : -__builtin_huge_val());
            for (int w = 0; w < nw; ++w) {
                const double x = wred[w * kPartVals + v];
                r = op == RSUM ? r + x : (op == RMIN ? ((x < r) ? x : r) : ((x > r) ? x : r));
            }
            scal[spec.dst[v]] = r;
        }
        run_step<T>(step, scal);
    }
    (void)red;
}

struct SetSpec {
    int n;
    int idx[8];
    double val[8];
};

__global__ void k_set_scal(double* scal, SetSpec s) {
    if ((int)threadIdx.x < s.n) scal[s.idx[threadIdx.x]] = s.val[threadIdx.x];
}

// ── Love kernel (src/boussinesq_kernel.cpp:44-95) ──────────────────────────

__device__ double d_asinh_diff(double u, double v, double du) {
    const double huge_arg = 1e150;
    if (fabs(u) > huge_arg || fabs(v) > huge_arg) return asinh(u) - asinh(v);
    const double su = sqrt(1.0 + u * u), sv = sqrt(1.0 + v * v);
    const double den = u * sv + v * su;
    const double w = (u * v > 0.0 && den != 0.0) ? du * (u + v) / den : u * sv - v * su;
    return asinh(w);
}

__device__ double d_term(double t, double s1, double s2, double ds) {
    if (t == 0.0) return 0.0;
    const double c = fabs(t);
    return t * d_asinh_diff(s1 / c, s2 / c, ds / c);
}

__device__ double d_love_uz(double x, double y, double a, double b) {
    x = fabs(x);
    y = fabs(y);
    const double m = (a > b) ? a : b;
    const double r2 = x * x + y * y;
    const double sw = 1500.0 * m;
    if (r2 > sw * sw) {
        const double r = sqrt(r2);
        const double A = 4.0 * a * b;
        const double quad = (a * a * (3.0 * x * x - r2) + b * b * (3.0 * y * y - r2)) /
                            (6.0 * r2 * r2);
        return A * (1.0 + quad) / r;
    }
    const double xp = x + a, xm = x - a, yp = y + b, ym = y - b;
    const double da = 2.0 * a, db = 2.0 * b;
    return d_term(xp, yp, ym, db) + d_term(yp, xp, xm, da) + d_term(xm, ym, yp, -db) +
           d_term(ym, xm, xp, -da);
}

// ── Cerruti (tangential) cell integrals (src/cerruti_kernel.cpp:1-127) ─────
// uxx = (1-nu) love_uz + nu ylog (ylog = integral of (x-xi)^2/rho^3, same
// far-field switch); uxy = integral of (x-xi)(y-eta)/rho^3 in the
// rationalised, cancellation-free product form, odd in x and in y.
__device__ double d_cerruti_ylog(double xa, double ya, double a, double b) {
    const double mmax = (a > b) ? a : b;
    const double r2 = xa * xa + ya * ya;
    const double sw = 1500.0 * mmax;
    if (r2 > sw * sw) {
        const double r = sqrt(r2);
        const double r3 = r2 * r, r5 = r3 * r2, r7 = r5 * r2;
        const double A = 4.0 * a * b;
        const double x2 = xa * xa, y2 = ya * ya;
        return A * (x2 / r3 + (a * a / 6.0) * (2.0 / r3 - 15.0 * x2 / r5 + 15.0 * x2 * x2 / r7) +
                    (b * b / 6.0) * (-3.0 * x2 / r5 + 15.0 * x2 * y2 / r7));
    }
    const double k = xa + a, l = xa - a, m = ya + b, n = ya - b;
    const double da = 2.0 * a;
    return d_term(m, k, l, da) + d_term(n, l, k, -da);
}

__device__ double d_cerruti_uxx(double x, double y, double a, double b, double nu) {
    const double xa = fabs(x), ya = fabs(y);
    return (1.0 - nu) * d_love_uz(xa, ya, a, b) + nu * d_cerruti_ylog(xa, ya, a, b);
}

__device__ double d_cerruti_uxy(double x, double y, double a, double b) {
    const double xa = fabs(x), ya = fabs(y);
    const double k = xa + a, l = xa - a, m = ya + b, n = ya - b;
    const double Rkm = hypot(k, m), Rkn = hypot(k, n), Rlm = hypot(l, m), Rln = hypot(l, n);
    const double num = 16.0 * a * b * xa * ya * (1.0 / (Rkn + Rln) + 1.0 / (Rkm + Rlm));
    const double den = (Rlm + Rln) * (Rkn + Rkm);
    const double v = den > 0.0 ? num / den : 0.0;
    return ((x >= 0.0) ? 1.0 : -1.0) * ((y >= 0.0) ? 1.0 : -1.0) * v;
}

__global__ void __launch_bounds__(256)
k_cerruti_tables(int Ns, double h, double nu, double pref, double prefnu,
                 double* __restrict__ xx, double* __restrict__ xy) {
    const std::int64_t N = (std::int64_t)Ns * Ns;
    const double a = 0.5 * h;
    GRID_LOOP(i, N) {
        const int dy = (int)(i / Ns), dx = (int)(i - (std::int64_t)dy * Ns);
        xx[i] = pref * d_cerruti_uxx(dx * h, dy * h, a, a, nu);
        xy[i] = prefnu * d_cerruti_uxy(dx * h, dy * h, a, a);
    }
}

__global__ void __launch_bounds__(256)
k_cerruti_points(int n, const double* x, const double* y, const double* a, const double* b,
                 double nu, double* uxx, double* uxy) {
    GRID_LOOP(i, n) {
        uxx[i] = d_cerruti_uxx(x[i], y[i], a[i], b[i], nu);
        uxy[i] = d_cerruti_uxy(x[i], y[i], a[i], b[i]);
    }
}

__global__ void __launch_bounds__(256)
k_love_table(int Ns, double h, double scale, double* __restrict__ table) {
    const std::int64_t N = (std::int64_t)Ns * Ns;
    const double a = 0.5 * h;
    GRID_LOOP(i, N) {
        const int dy = (int)(i / Ns), dx = (int)(i - (std::int64_t)dy * Ns);
        table[i] = scale * d_love_uz(dx * h, dy * h, a, a);
    }
}

__global__ void __launch_bounds__(256)
k_love_points(int n, const double* x, const double* y, const double* a, const double* b,
              double* out) {
    GRID_LOOP(i, n) out[i] = d_love_uz(x[i], y[i], a[i], b[i]);
}

__global__ void __launch_bounds__(256)
k_restrict2(int Ns, const double* __restrict__ f, double* __restrict__ out) {
    const int c = Ns / 2;
    const std::int64_t N = (std::int64_t)c * c;
    GRID_LOOP(i, N) {
        const int iy = (int)(i / c), ix = (int)(i - (std::int64_t)iy * c);
        const std::size_t x0 = 2 * ix, y0 = 2 * iy;
        out[i] = 0.25 * (f[y0 * Ns + x0] + f[y0 * Ns + x0 + 1] + f[(y0 + 1) * Ns + x0] +
                         f[(y0 + 1) * Ns + x0 + 1]);
    }
}

__global__ void __launch_bounds__(256)
k_prolong2(int Nc, const double* __restrict__ pc, double* __restrict__ out) {
    const int f = 2 * Nc;
    const std::int64_t N = (std::int64_t)f * f;
    GRID_LOOP(i, N) {
        const int iy = (int)(i / f), ix = (int)(i - (std::int64_t)iy * f);
        out[i] = pc[(std::size_t)(iy >> 1) * Nc + (ix >> 1)];
    }
}

__global__ void __launch_bounds__(256)
k_cast_d2f(std::int64_t N, const double* __restrict__ x, float* __restrict__ y) {
    GRID_LOOP(i, N) y[i] = (float)x[i];
}


// ── adjoint: linear CG on a fixed contact set ──────────────────────────────
// Solves [S_CC 1; 1' 0][mu; nu] = [w_C; 0] (the transposed KKT system of the
// contact QP on its converged support) as CG on the zero-mean subspace of C.

template <typename T>
__global__ void __launch_bounds__(256)
k_adj_init(std::int64_t N, const double* __restrict__ p, const T* __restrict__ w,
           std::uint8_t* __restrict__ mask, T* __restrict__ r, T* __restrict__ x,
           T* __restrict__ t, double* __restrict__ part) {
    __shared__ double red[8];
    double ws = 0.0, nc = 0.0;
    GRID_LOOP(i, N) {
        const bool c = p[i] > 0.0;
        mask[i] = c ? 1 : 0;
        const T wi = c ? w[i] : T(0);
        r[i] = wi;
        x[i] = T(0);
        t[i] = T(0);
        if (c) {
            ws += (double)wi;
            nc += 1.0;
        }
    }
    put<RSUM>(ws, red, part, 2, 0);
    put<RSUM>(nc, red, part, 2, 1);
}

template <typename T, bool PC>
__global__ void __launch_bounds__(256)
k_adj_dot(std::int64_t N, const T* __restrict__ r, const T* __restrict__ z,
          const std::uint8_t* __restrict__ mask, const double* __restrict__ scal,
          double* __restrict__ part) {
    __shared__ double red[8];
    const T zmean = PC ? static_cast<T>(scal[S_ZMEAN]) : T(0);
    double rz = 0.0, rr = 0.0;
    GRID_LOOP(i, N) {
        if (!mask[i]) continue;
        const double ri = (double)r[i];
        const double zi = PC ? (double)(T)(z[i] - zmean) : ri;
        rz += ri * zi;
        rr += ri * ri;
    }
    put<RSUM>(rz, red, part, 2, 0);
    put<RSUM>(rr, red, part, 2, 1);
}

template <typename T, bool PC>
__global__ void __launch_bounds__(256)
k_adj_dir(std::int64_t N, const T* __restrict__ r, const T* __restrict__ z,
          const std::uint8_t* __restrict__ mask, const double* __restrict__ scal,
          T* __restrict__ t) {
    const T zmean = PC ? static_cast<T>(scal[S_ZMEAN]) : T(0);
    const T beta = static_cast<T>(scal[S_A_BETA]);
    GRID_LOOP(i, N) {
        const T zi = PC ? T(z[i] - zmean) : r[i];
        t[i] = mask[i] ? T(zi + beta * t[i]) : T(0);
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_adj_curv(std::int64_t N, const T* __restrict__ q, const T* __restrict__ t,
           const std::uint8_t* __restrict__ mask, double* __restrict__ part) {
    __shared__ double red[8];
    double qs = 0.0, qt = 0.0, ts = 0.0;
    GRID_LOOP(i, N) {
        if (!mask[i]) continue;
        const double qi = (double)q[i], ti = (double)t[i];
        qs += qi;
        qt += qi * ti;
        ts += ti;
    }
    put<RSUM>(qs, red, part, 3, 0);
    put<RSUM>(qt, red, part, 3, 1);
    put<RSUM>(ts, red, part, 3, 2);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_adj_step(std::int64_t N, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ t,
           const T* __restrict__ q, const std::uint8_t* __restrict__ mask,
           const double* __restrict__ scal, double* __restrict__ part) {
    __shared__ double red[8];
    const bool ok = scal[S_A_OK] != 0.0;
    const T a = static_cast<T>(scal[S_A_ALPHA]), qm = static_cast<T>(scal[S_A_QMEAN]);
    double rs = 0.0;
    GRID_LOOP(i, N) {
        if (!mask[i]) continue;
        T ri = r[i];
        if (ok) {
            x[i] = x[i] + a * t[i];
            ri = ri - a * (q[i] - qm);
            r[i] = ri;
        }
        rs += (double)ri;
    }
    put<RSUM>(rs, red, part, 1, 0);
}

template <typename T>
__global__ void __launch_bounds__(256)
k_adj_finish(std::int64_t N, const T* __restrict__ x, const T* __restrict__ w,
             const T* __restrict__ q, const std::uint8_t* __restrict__ mask,
             double* __restrict__ grad, double* __restrict__ part) {
    __shared__ double red[8];
    double s = 0.0;
    GRID_LOOP(i, N) {
        const bool c = mask[i] != 0;
        grad[i] = c ? -(double)x[i] : 0.0;
        if (c) s += (double)w[i] - (double)q[i];
    }
    put<RSUM>(s, red, part, 1, 0);
}


// Plain streaming copy in 16-byte words (calibration of the HBM counters:
// reads and writes exactly `n16` * 16 bytes).
__global__ void __launch_bounds__(256)
k_copy16(std::int64_t n16, const float4* __restrict__ x, float4* __restrict__ y) {
    GRID_LOOP(i, n16) y[i] = x[i];
}

} // namespace

#define L1(k, ...) k<<<nblk, kVecThreads, 0, s>>>(__VA_ARGS__); ASPHER_LAUNCH_CHECK()
// 4 consecutive elements per access when every range boundary is a multiple
// of 4 (single vector: N; ensemble: flagged by its builder) and every
// pointer is 16-byte aligned (null pointers are not read)
inline bool a16(const void* p) { return (reinterpret_cast<std::uintptr_t>(p) & 15u) == 0; }
template <typename... P>
bool vec4_ok(std::int64_t N, const Ens& ens, P... ptrs) {
    return N % 4 == 0 && (!ens.work || ens.vec4) && (a16(ptrs) && ...);
}
#define L1S(k, ...) k<<<vec_blocks(N), kVecThreads, 0, s>>>(__VA_ARGS__); ASPHER_LAUNCH_CHECK()
// ensemble-aware launches: one block per work item, else the single-solve grid
#define LE(k, ...) k<<<ens.work ? ens.nwork : nblk, kVecThreads, 0, s>>>(__VA_ARGS__); \
    ASPHER_LAUNCH_CHECK()
#define LES(k, ...) k<<<ens.work ? ens.nwork : vec_blocks(N), kVecThreads, 0, s>>>(__VA_ARGS__); \
    ASPHER_LAUNCH_CHECK()

template <typename T>
void VecOps<T>::eval1(std::int64_t N, const T* u, const T* g0s, const T* p, T* g,
                      double* part, int nblk, hipStream_t s, const Ens& ens) {
    if (vec4_ok(N, ens, u, g0s, p, g)) { LE((k_eval1<T, 4>), N, u, g0s, p, g, part, ens); }
    else { LE((k_eval1<T, 1>), N, u, g0s, p, g, part, ens); }
}
template <typename T>
void VecOps<T>::eval2(std::int64_t N, const T* p, T* g, std::uint8_t* mask,
                      const double* scal, double* part, int nblk, hipStream_t s,
                      const Ens& ens) {
    if (vec4_ok(N, ens, p, g, mask)) { LE((k_eval2<T, 4>), N, p, g, mask, scal, part, ens); }
    else { LE((k_eval2<T, 1>), N, p, g, mask, scal, part, ens); }
}
template <typename T>
void VecOps<T>::stencil(int Ns, const T* g, const std::uint8_t* mask, const double* scal,
                        const int* tapd, const double* tapw, int ntaps, int R, T* z,
                        double* part, int nblk, hipStream_t s, const Ens& ens,
                        bool zero_off) {
    const std::size_t lds = sizeof(double) * (kStX + 2 * R) * (kStY + 2 * R);
    const int grid = ens.work ? ens.nwork : nblk;
    const int zo = zero_off ? 1 : 0;
    // the unrolled kernels assume the l2 disc's tap count of their radius
    static const int disc_taps[5] = {0, 5, 13, 29, 49};
    switch (R >= 1 && R <= 4 && ntaps == disc_taps[R] ? R : 0) {
#define ST_CASE(RR) \
        case RR: k_stencil<T, RR><<<grid, kVecThreads, lds, s>>>(Ns, g, mask, scal, tapd, tapw, \
                                                               ntaps, R, z, part, ens, zo); break;
        ST_CASE(1) ST_CASE(2) ST_CASE(3) ST_CASE(4)
#undef ST_CASE
        default:
            k_stencil<T, 0><<<grid, kVecThreads, lds, s>>>(Ns, g, mask, scal, tapd, tapw, ntaps, R,
                                                          z, part, ens, zo);
    }
    ASPHER_LAUNCH_CHECK();
}
template <typename T>
void VecOps<T>::cgdot(std::int64_t N, const T* g, const T* gprev, const T* z,
                      const std::uint8_t* mask, const double* scal, bool has_pc,
                      double* part, int nblk, hipStream_t s, const Ens& ens) {
    const bool v4 = vec4_ok(N, ens, g, gprev, has_pc ? z : nullptr, mask);
    if (has_pc && v4) { LE((k_cgdot<T, true, 4>), N, g, gprev, z, mask, scal, part, ens); }
    else if (has_pc) { LE((k_cgdot<T, true, 1>), N, g, gprev, z, mask, scal, part, ens); }
    else if (v4) { LE((k_cgdot<T, false, 4>), N, g, gprev, z, mask, scal, part, ens); }
    else { LE((k_cgdot<T, false, 1>), N, g, gprev, z, mask, scal, part, ens); }
}
template <typename T>
void VecOps<T>::tupdate(std::int64_t N, const T* g, T* gprev, const T* z,
                        const std::uint8_t* mask, T* t, const double* scal, bool has_pc,
                        hipStream_t s, const Ens& ens) {
    const bool v4 = vec4_ok(N, ens, g, gprev, has_pc ? z : nullptr, mask, t);
    if (has_pc && v4) { LES((k_tupdate<T, true, 4>), N, g, gprev, z, mask, t, scal, ens); }
    else if (has_pc) { LES((k_tupdate<T, true, 1>), N, g, gprev, z, mask, t, scal, ens); }
    else if (v4) { LES((k_tupdate<T, false, 4>), N, g, gprev, z, mask, t, scal, ens); }
    else { LES((k_tupdate<T, false, 1>), N, g, gprev, z, mask, t, scal, ens); }
}
template <typename T>
void VecOps<T>::line1(std::int64_t N, const T* r, const T* g, const T* t,
                      const std::uint8_t* mask, double* part, int nblk, hipStream_t s,
                      const Ens& ens) {
    if (vec4_ok(N, ens, r, g, t, mask)) { LE((k_line1<T, 4>), N, r, g, t, mask, part, ens); }
    else { LE((k_line1<T, 1>), N, r, g, t, mask, part, ens); }
}
template <typename T>
void VecOps<T>::line2(std::int64_t N, const T* r, const T* t, const std::uint8_t* mask,
                      const double* scal, double* part, int nblk, hipStream_t s,
                      const Ens& ens) {
    if (vec4_ok(N, ens, r, t, mask)) { LE((k_line2<T, 4>), N, r, t, mask, scal, part, ens); }
    else { LE((k_line2<T, 1>), N, r, t, mask, scal, part, ens); }
}
template <typename T>
void VecOps<T>::update(std::int64_t N, const T* pin, T* pout, const T* t, const T* g,
                       const double* scal, double* part, int nblk, hipStream_t s,
                       const Ens& ens, T* best, int best_flag) {
    if (vec4_ok(N, ens, pin, pout, t, g, best)) {
        LE((k_update<T, 4>), N, pin, pout, t, g, scal, part, ens, best, best_flag);
    } else {
        LE((k_update<T, 1>), N, pin, pout, t, g, scal, part, ens, best, best_flag);
    }
}
template <typename T>
void VecOps<T>::rescale(std::int64_t N, T* p, const double* scal, hipStream_t s,
                        const Ens& ens) {
    if (vec4_ok(N, ens, p)) { LES((k_rescale<T, 4>), N, p, scal, ens); }
    else { LES((k_rescale<T, 1>), N, p, scal, ens); }
}
template <typename T>
void VecOps<T>::objective(std::int64_t N, const T* p, const T* u, const T* g0s,
                          double* part, int nblk, hipStream_t s) {
    L1(k_objective<T>, N, p, u, g0s, part);
}
template <typename T>
void VecOps<T>::axpy_id(std::int64_t N, const T* p, const T* g, T* z, double s_id,
                        hipStream_t s) {
    L1S(k_axpy_id<T>, N, p, g, z, static_cast<T>(s_id));
}
template <typename T>
void VecOps<T>::sum(std::int64_t N, const T* y, double* part, int nblk, hipStream_t s) {
    L1(k_sum<T>, N, y, part);
}
template <typename T>
void VecOps<T>::moments(std::int64_t N, const T* y, double* part, int nblk, hipStream_t s) {
    L1(k_moments<T>, N, y, part);
}
template <typename T>
void VecOps<T>::count_above(std::int64_t N, const T* y, double theta, double* part,
                            int nblk, hipStream_t s) {
    L1(k_count_above<T>, N, y, theta, part);
}
template <typename T>
void VecOps<T>::project(std::int64_t N, T* y, double theta, hipStream_t s) {
    L1S(k_project<T>, N, y, static_cast<T>(theta));
}
template <typename T>
void VecOps<T>::idcheck(std::int64_t N, const T* z, const T* r, const T* g0s, const T* g,
                        const T* p, double* part, int nblk, hipStream_t s) {
    L1(k_idcheck<T>, N, z, r, g0s, g, p, part);
}
template <typename T>
void VecOps<T>::stats(std::int64_t N, const T* y, double* part, int nblk, hipStream_t s) {
    L1(k_stats<T>, N, y, part);
}
template <typename T>
void VecOps<T>::clamp_sum(std::int64_t N, T* p, double* part, int nblk, hipStream_t s) {
    L1(k_clamp_sum<T>, N, p, part);
}
template <typename T>
void VecOps<T>::scale(std::int64_t N, T* p, double f, hipStream_t s) {
    L1S(k_scale<T>, N, p, static_cast<T>(f));
}
template <typename T>
void VecOps<T>::fill(std::int64_t N, T* p, double v, hipStream_t s) {
    L1S(k_fill<T>, N, p, static_cast<T>(v));
}
template <typename T>
void VecOps<T>::from_double(std::int64_t N, const double* x, double shift, T* y,
                            hipStream_t s) {
    L1S(k_from_double<T>, N, x, shift, y);
}
template <typename T>
void VecOps<T>::to_double(std::int64_t N, const T* x, double* y, hipStream_t s) {
    L1S(k_to_double<T>, N, x, y);
}
template <typename T>
void VecOps<T>::shift(std::int64_t N, const T* x, double c, T* y, hipStream_t s) {
    L1S(k_shift<T>, N, x, static_cast<T>(c), y);
}
template <typename T>
void VecOps<T>::masked_sum(std::int64_t N, const T* y, const std::uint8_t* mask, double* part,
                           int nblk, hipStream_t s) {
    L1(k_masked_sum<T>, N, y, mask, part);
}
template <typename T>
void VecOps<T>::masked_sub(std::int64_t N, T* y, const std::uint8_t* mask, const double* scal,
                           int idx, hipStream_t s) {
    L1S(k_masked_sub<T>, N, y, mask, scal, idx);
}
template <typename T>
void VecOps<T>::copy(std::int64_t N, const T* x, T* y, hipStream_t s) {
    ASPHER_HIP(hipMemcpyAsync(y, x, sizeof(T) * N, hipMemcpyDeviceToDevice, s));
}
template <typename T>
void VecOps<T>::dense(int Ns, const double* table, const T* x, T* y, hipStream_t s) {
    const std::int64_t N = (std::int64_t)Ns * Ns;
    k_dense<T><<<(int)((N + 255) / 256), 256, 0, s>>>(Ns, table, x, y);
    ASPHER_LAUNCH_CHECK();
}

template struct VecOps<float>;
template struct VecOps<double>;

template <typename T>
void AdjOps<T>::init(std::int64_t N, const double* p, const T* w, std::uint8_t* mask, T* r,
                     T* x, T* t, double* part, int nblk, hipStream_t s) {
    L1(k_adj_init<T>, N, p, w, mask, r, x, t, part);
}
template <typename T>
void AdjOps<T>::dot(std::int64_t N, const T* r, const T* z, const std::uint8_t* mask,
                    const double* scal, double* part, int nblk, hipStream_t s) {
    if (z) { L1((k_adj_dot<T, true>), N, r, z, mask, scal, part); }
    else { L1((k_adj_dot<T, false>), N, r, z, mask, scal, part); }
}
template <typename T>
void AdjOps<T>::dir(std::int64_t N, const T* r, const T* z, const std::uint8_t* mask,
                    const double* scal, T* t, hipStream_t s) {
    if (z) { L1S((k_adj_dir<T, true>), N, r, z, mask, scal, t); }
    else { L1S((k_adj_dir<T, false>), N, r, z, mask, scal, t); }
}
template <typename T>
void AdjOps<T>::curv(std::int64_t N, const T* q, const T* t, const std::uint8_t* mask,
                     double* part, int nblk, hipStream_t s) {
    L1(k_adj_curv<T>, N, q, t, mask, part);
}
template <typename T>
void AdjOps<T>::step(std::int64_t N, T* x, T* r, const T* t, const T* q,
                     const std::uint8_t* mask, const double* scal, double* part, int nblk,
                     hipStream_t s) {
    L1(k_adj_step<T>, N, x, r, t, q, mask, scal, part);
}
template <typename T>
void AdjOps<T>::finish(std::int64_t N, const T* x, const T* w, const T* q,
                       const std::uint8_t* mask, double* grad, double* part, int nblk,
                       hipStream_t s) {
    L1(k_adj_finish<T>, N, x, w, q, mask, grad, part);
}

template struct AdjOps<double>;

void copy16(std::int64_t bytes, const void* x, void* y, hipStream_t s) {
    const std::int64_t n16 = bytes / 16;
    const int nb = (int)std::min<std::int64_t>((n16 + 255) / 256, 8192);
    k_copy16<<<std::max(nb, 1), 256, 0, s>>>(n16, static_cast<const float4*>(x),
                                             static_cast<float4*>(y));
    ASPHER_LAUNCH_CHECK();
}

template <typename T>
void finalize(const double* part, int nblk, int nv, const int* ops, const int* dst,
              double* scal, Step step, hipStream_t s, const Ens& ens) {
    FinSpec f{};
    f.nv = nv;
    for (int v = 0; v < nv; ++v) {
        f.ops[v] = ops[v];
        f.dst[v] = dst[v];
    }
    bool all_sum = true;
    for (int v = 0; v < nv; ++v) all_sum = all_sum && ops[v] == RSUM;
    const dim3 grid(std::max(ens.ninst, 1));
    const int st = static_cast<int>(step);
    switch (nv) {
#define FIN_CASE(V) \
        case V: \
            if (all_sum) k_finalize<T, V, true><<<grid, 256, 0, s>>>(part, nblk, f, scal, st, ens); \
            else k_finalize<T, V, false><<<grid, 256, 0, s>>>(part, nblk, f, scal, st, ens); \
            break;
        FIN_CASE(1) FIN_CASE(2) FIN_CASE(3) FIN_CASE(4)
        FIN_CASE(5) FIN_CASE(6) FIN_CASE(7) FIN_CASE(8)
#undef FIN_CASE
        default: throw std::invalid_argument("finalize: 1 to 8 values");
    }
    ASPHER_LAUNCH_CHECK();
}
template void finalize<float>(const double*, int, int, const int*, const int*, double*,
                              Step, hipStream_t, const Ens&);
template void finalize<double>(const double*, int, int, const int*, const int*, double*,
                               Step, hipStream_t, const Ens&);

void set_scal(double* scal, int n, const int* idx, const double* val, hipStream_t s) {
    SetSpec sp{};
    if (n > 8) throw std::invalid_argument("set_scal: at most 8 values");
    sp.n = n;
    for (int i = 0; i < n; ++i) {
        sp.idx[i] = idx[i];
        sp.val[i] = val[i];
    }
    k_set_scal<<<1, 64, 0, s>>>(scal, sp);
    ASPHER_LAUNCH_CHECK();
}

void cerruti_tables(int Ns, double L, double E_star, double nu, double* xx, double* xy,
                    hipStream_t s) {
    const double pi = 3.14159265358979323846;
    const double pref = 1.0 / (pi * E_star * (1.0 - nu));
    const std::int64_t N = (std::int64_t)Ns * Ns;
    k_cerruti_tables<<<vec_blocks(N), kVecThreads, 0, s>>>(Ns, L / Ns, nu, pref, nu * pref, xx,
                                                           xy);
    ASPHER_LAUNCH_CHECK();
}

void cerruti_points(int n, const double* x, const double* y, const double* a, const double* b,
                    double nu, double* uxx, double* uxy, hipStream_t s) {
    k_cerruti_points<<<vec_blocks(n), kVecThreads, 0, s>>>(n, x, y, a, b, nu, uxx, uxy);
    ASPHER_LAUNCH_CHECK();
}

void love_table(int Ns, double L, double E_star, double* table, hipStream_t s) {
    const double h = L / Ns;
    const double scale = 1.0 / (M_PI * E_star);
    const std::int64_t N = (std::int64_t)Ns * Ns;
    k_love_table<<<vec_blocks(N), kVecThreads, 0, s>>>(Ns, h, scale, table);
    ASPHER_LAUNCH_CHECK();
}

void love_points(int n, const double* x, const double* y, const double* a, const double* b,
                 double* out, hipStream_t s) {
    k_love_points<<<vec_blocks(n), kVecThreads, 0, s>>>(n, x, y, a, b, out);
    ASPHER_LAUNCH_CHECK();
}

void restrict2(int Ns, const double* f, double* out, hipStream_t s) {
    const std::int64_t N = (std::int64_t)(Ns / 2) * (Ns / 2);
    k_restrict2<<<vec_blocks(N), kVecThreads, 0, s>>>(Ns, f, out);
    ASPHER_LAUNCH_CHECK();
}

void prolong2(int Nc, const double* pc, double* out, hipStream_t s) {
    const std::int64_t N = 4LL * Nc * Nc;
    k_prolong2<<<vec_blocks(N), kVecThreads, 0, s>>>(Nc, pc, out);
    ASPHER_LAUNCH_CHECK();
}

void cast_d2f(std::int64_t N, const double* x, float* y, hipStream_t s) {
    k_cast_d2f<<<vec_blocks(N), kVecThreads, 0, s>>>(N, x, y);
    ASPHER_LAUNCH_CHECK();
}

// ── test hooks for the failure statuses ────────────────────────────────────

namespace {

std::atomic<int> g_fault_mode{FAULT_NONE};
std::atomic<long long> g_fault_after{0};
// applies counted on the device: a captured graph replays the counting
// kernel, so replays count like eager launches
unsigned long long* g_fault_count = nullptr;

__global__ void k_fault_reset(unsigned long long* cnt) { *cnt = 0ULL; }

template <typename T>
__global__ void k_fault_nan(T* y, unsigned long long* cnt, long long after) {
    const unsigned long long seen = atomicAdd(cnt, 1ULL);
    if ((long long)seen >= after) y[0] = static_cast<T>(__builtin_nan(""));
}

template <typename T>
__global__ void k_fault_drop(T* u, double drop) {
    u[0] = static_cast<T>((double)u[0] - drop);
}

} // namespace

void fault_configure(int mode, long long after) {
    if (mode != FAULT_NONE && mode != FAULT_NONFINITE && mode != FAULT_FINAL_CHECK)
        throw std::invalid_argument("fault mode must be 0 (off), 1 (nonfinite) or 2 (final_check)");
    if (after < 0) throw std::invalid_argument("fault: after must be >= 0");
    ASPHER_HIP(hipDeviceSynchronize());
    if (mode == FAULT_NONFINITE) {
        if (!g_fault_count) g_fault_count = static_cast<unsigned long long*>(dev_alloc(8));
        k_fault_reset<<<1, 1>>>(g_fault_count);
        ASPHER_LAUNCH_CHECK();
        ASPHER_HIP(hipDeviceSynchronize());
    }
    g_fault_after.store(after);
    g_fault_mode.store(mode);
}

int fault_mode() { return g_fault_mode.load(std::memory_order_relaxed); }

template <typename T>
void fault_after_apply(T* y, hipStream_t s) {
    if (fault_mode() != FAULT_NONFINITE || !g_fault_count) return;
    k_fault_nan<T><<<1, 1, 0, s>>>(y, g_fault_count, g_fault_after.load());
    ASPHER_LAUNCH_CHECK();
}
template void fault_after_apply<double>(double*, hipStream_t);
template void fault_after_apply<float>(float*, hipStream_t);

template <typename T>
void fault_before_final_check(T* u, double drop, hipStream_t s) {
    k_fault_drop<T><<<1, 1, 0, s>>>(u, drop);
    ASPHER_LAUNCH_CHECK();
}
template void fault_before_final_check<double>(double*, double, hipStream_t);
template void fault_before_final_check<float>(float*, double, hipStream_t);

} // namespace aspher
