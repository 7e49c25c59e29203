// Shared host/device definitions for the MI355X spectral contact path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace aspher {

// HIP status -> exception. The C-ABI layer catches and maps it to a code.
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define ASPHER_HIP(call)                                                        \
    do {                                                                        \
        hipError_t err__ = (call);                                              \
        if (err__ != hipSuccess)                                                \
            throw ::aspher::DeviceError(std::string(#call) + ": " +             \
                                        hipGetErrorString(err__) + " (" +       \
                                        __FILE__ + ":" +                        \
                                        std::to_string(__LINE__) + ")");        \
    } while (0)

#define ASPHER_LAUNCH_CHECK() ASPHER_HIP(hipGetLastError())

// Device memory of the library's own buffers, with reuse: a freed block is
// kept (per device, up to a capacity) and handed out again for a request of
// the same rounded size, so repeated solves do not pay hipMalloc/hipFree per
// level. dev_free synchronises the device first, as hipFree does, so a
// reused block is never still in use by queued work.
void* dev_alloc(std::size_t bytes);
void dev_free(void* p, std::size_t bytes) noexcept;
// Returns every cached block to the runtime (the device must be idle).
void dev_release_cached() noexcept;

// Interleaved complex value; alignment lets the compiler emit 8/16-byte
// vector loads and stores for one element.
template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) {
    return {a.re + b.re, a.im + b.im};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) {
    return {a.re - b.re, a.im - b.im};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> operator*(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> scale(cx<T> a, T s) {
    return {a.re * s, a.im * s};
}
template <typename T>
__host__ __device__ __forceinline__ cx<T> conjg(cx<T> a) {
    return {a.re, -a.im};
}

// Number of 256-thread blocks for grid-stride O(N) passes with block-level
// reductions: enough to fill 256 CUs several times over, bounded so the
// partial-sum buffer stays small and the final reduction is one block.
constexpr int kVecThreads = 256;
constexpr int kMaxVecBlocks = 2048;
// partial-sum buffer: one row of up to kPartVals values per block
constexpr int kPartVals = 8;
// rows of the partial buffer of one solve context (>= any reduction's block count)
constexpr int kMaxPartRows = 16384;

inline int vec_blocks(std::int64_t n) {
    std::int64_t b = (n + 4LL * kVecThreads - 1) / (4LL * kVecThreads);
    if (b < 1) b = 1;
    if (b > kMaxVecBlocks) b = kMaxVecBlocks;
    return static_cast<int>(b);
}

// ── ensembles: many independent solves in one launch ────────────────────
// Vectors of all instances are concatenated (instance i at element offset
// off[i]); scalar blocks sit at scal + i*sstride, partial blocks at
// part + i*pstride. An ensemble launch takes one work item per block: a
// chunk [begin, end) of one instance's elements, whose block reduction goes
// to partial row `row` of that instance. Instances whose activity flag
// (scalar `gate`) is 0 are skipped. The default Ens is a single solve:
// grid-stride loops over one vector, exactly as without an ensemble.
constexpr int kChunk = 4096;   // elements per work item (16 per thread)

struct WorkItem {
    std::int64_t begin, end;
    int inst, row;
};

struct Ens {
    const WorkItem* work = nullptr;   // null: single solve
    int nwork = 0;
    const std::int64_t* off = nullptr;
    int sstride = 0;
    std::int64_t pstride = 0;
    const double* gscal = nullptr;
    int gate = -1;
    const int* ns = nullptr;          // per-instance grid side (tile work items)
    int ninst = 1;                    // instances (finalize grid)
    const int* rows = nullptr;        // per-instance partial rows (finalize)
    // every instance offset and work-item boundary is a multiple of 4, so
    // the element loops may take 4 consecutive elements per access
    bool vec4 = false;
};

} // namespace aspher
