// Deterministic block-level reductions (64-wide wavefronts).
//
// Every reduction in this library is two-level: each block reduces its
// grid-stride share in a fixed order (per-thread loop, then a fixed shuffle
// tree, then waves in index order) and writes one partial per value; a single
// block later reduces the partials in index order. Results are therefore
// bit-reproducible run to run, unlike an atomic or OpenMP reduction.
#pragma once

#include "common.hpp"

namespace aspher {

enum RedOp : int { RSUM = 0, RMIN = 1, RMAX = 2 };

template <int OP>
__device__ __forceinline__ double red_apply(double a, double b) {
    if constexpr (OP == RSUM) return a + b;
    // NaN-ignoring like the reference's `if (v < vmin) vmin = v`
    else if constexpr (OP == RMIN) return (b < a) ? b : a;
    else return (b > a) ? b : a;
}

template <int OP>
__device__ __forceinline__ double red_identity() {
    if constexpr (OP == RSUM) return 0.0;
    else if constexpr (OP == RMIN) return __builtin_huge_val();
    else return -__builtin_huge_val();
}

// Result valid in lane 0. nlanes: the wave's launched lanes (a block whose
// size is not a multiple of 64 ends in a partial wave); a shuffle from a lane
// at or beyond nlanes reads a register no thread wrote and is skipped. For a
// full wave the tree and its order are the plain one.
template <int OP>
__device__ __forceinline__ double wave_reduce(double v, int nlanes = 64) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(v, off, 64);
        if (lane + off < nlanes) v = red_apply<OP>(v, o);
    }
    return v;
}

// Reduces one value per thread over the block; the result is valid in
// thread 0. `scratch` must hold (blockDim.x + 63)/64 doubles. Contains
// barriers: every thread of the block must call it.
template <int OP>
__device__ __forceinline__ double block_reduce(double v, double* scratch) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    const int rest = (int)blockDim.x - 64 * wave;
    v = wave_reduce<OP>(v, rest < 64 ? rest : 64);
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double r = red_identity<OP>();
    if (threadIdx.x == 0)
        for (int w = 0; w < nw; ++w) r = red_apply<OP>(r, scratch[w]);
    return r;
}

} // namespace aspher
