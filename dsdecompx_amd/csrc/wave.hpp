// wave.hpp -- wave-level building blocks shared by the gfx950 kernel files
// (halo_kernels.hip, grid_kernels.hip): lane arithmetic, the XCD-aware block
// order, the periodic wrap, chunk loads, the LDS-staged column scan, and the
// host-side launch helpers (shape dispatch, timed launches).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "device.hpp"
#include "timing.hpp"

namespace dsdx {
namespace kern {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// A value that the pass reads once and only reads: the owner records a ghost
// copy or refresh gathers, the positions the classification reads.  STREAM
// (the particle arrays are larger than the 256 MB memory-side cache,
// Layout::streaming): the load is nontemporal, so these lines do not displace
// what the caches hold.  Otherwise the arrays may be cache-resident and the
// load is ordinary.  A compile-time choice: the compiler merges a run-time
// select of the two loads into one ordinary load.
template <bool STREAM, class T>
__device__ __forceinline__ T read_once(const T* p)
{
    if constexpr (STREAM) return __builtin_nontemporal_load(p);
    return *p;
}

// Calls f(std::integral_constant<bool, stream>) for the run-time flag.
template <class F>
hipError_t dispatch_stream(uint32_t stream, F&& f)
{
    if (stream) return f(std::integral_constant<bool, true>{});
    return f(std::integral_constant<bool, false>{});
}

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, kWave);
        if (lane >= static_cast<uint32_t>(d)) v += o;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) v |= __shfl_xor(v, d, kWave);
    return v;
}

// MI355X places block b of a launch on XCD b % 8, and each XCD has its own
// L2.  Blocks are renumbered so that each XCD works through runs of `run`
// consecutive blocks (the XCDs take the runs in turn): neighbouring chunks
// share cache lines (compacted entries, owner maps, ghost runs, count
// columns), and a line touched by one XCD only is fetched and written back
// once.  The last partial round of runs keeps the hardware order; run 0
// keeps it everywhere.  A bijection on [0, gridDim.x).
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t xcd_block(uint32_t run)
{
    const uint32_t b = blockIdx.x;
    if (run == 0) return b;
    const uint32_t group = kXcds * run;
    if (b >= gridDim.x - gridDim.x % group) return b;
    const uint32_t l = b % group;
    return b - l + (l % kXcds) * run + l / kXcds;
}

// The chunk (or tile) of this wave among `count` (wave-uniform); a
// descending launch walks them last to first.  Indices >= count (the padding
// waves of the last block) are returned unchanged.
template <int WAVES>
__device__ __forceinline__ uint32_t wave_chunk(const Layout& L, uint32_t count)
{
    const uint32_t c = uniform(xcd_block(L.xcd_run) * WAVES + threadIdx.x / kWave);
    return (L.descending && c < count) ? count - 1u - c : c;
}

// Replica of a chunk.  The descriptor table and the chunk -> replica map
// reach every kernel as __restrict__ parameters: nothing the kernel stores
// aliases them, so their wave-uniform reads compile to scalar loads.
__device__ __forceinline__ uint32_t replica_of(const uint32_t* __restrict__ chunk_replica,
                                               uint32_t chunk)
{
    return chunk_replica ? uniform(chunk_replica[chunk]) : 0u;
}

// Periodic wrap of migrate at P = 1 (src/migration.cpp:106-113):
// v = x - lower; v -= floor(v / extent) * extent; v + lower, each operation
// rounded on its own (explicit _rn intrinsics, never fused).
__device__ __forceinline__ double wrap(double value, double lower, double extent)
{
    const double v = __dsub_rn(value, lower);
    // 0 < v < extent: the rounded quotient v / extent lies in (0, 1) (a
    // quotient just below 1 rounds to at most 1 - 2^-53), so floor gives +0,
    // 0 * extent is +0 and v - 0 == v exactly.  The result is then v + lower,
    // bit for bit, without the division.
    if (v > 0.0 && v < extent) return __dadd_rn(v, lower);
    const double q = floor(__ddiv_rn(v, extent));
    return __dadd_rn(__dsub_rn(v, __dmul_rn(q, extent)), lower);
}

// Wraps the periodic axes of one lane's value and stores it back when it
// changed (used by the fused migrate + classify pass).
__device__ __forceinline__ void wrap_store(double& v, double* dst, bool valid, int periodic,
                                           double lower, double extent)
{
    if (!periodic) return;
    const double w = wrap(v, lower, extent);
    if (valid && __double_as_longlong(w) != __double_as_longlong(v)) *dst = w;
    v = w;
}

// Loads v[k] = src[first + 64 k + lane] for the chunk's items (0 past
// `owned`); a chunk entirely below `owned` takes the unpredicated loads.
// STREAM: nontemporal loads (read_once).
template <int ITEMS, bool STREAM = false>
__device__ __forceinline__ void load_items(const double* __restrict__ src, uint32_t first,
                                           uint32_t owned, bool full, double (&v)[ITEMS])
{
    const uint32_t lane = lane_id();
    auto ld = [&](uint32_t i) { return read_once<STREAM>(src + i); };
    if (full) {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) v[k] = ld(first + k * kWave + lane);
    } else {
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t i = first + k * kWave + lane;
            v[k] = i < owned ? ld(i) : 0.0;
        }
    }
}

constexpr int kMaxScanThreads = 1024;
constexpr int kScanPer = 8;                               // values per thread per tile
constexpr int kScanTile = kMaxScanThreads * kScanPer;

// LDS index of tile element p: one pad word per 32 keeps the thread-
// contiguous reads (stride kScanPer) off a single bank.
__device__ __forceinline__ uint32_t scan_pad(uint32_t p) { return p + (p >> 5); }

__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* s_wave, uint32_t waves,
                                                    uint32_t* total)
{
    const uint32_t t = threadIdx.x;
    const uint32_t wave = t / kWave;
    const uint32_t incl = wave_inclusive_sum(v);
    if (lane_id() == kWave - 1) s_wave[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        const uint32_t w = lane_id() < waves ? s_wave[lane_id()] : 0u;
        const uint32_t wi = wave_inclusive_sum(w);
        if (lane_id() < waves) s_wave[lane_id()] = wi - w;
        if (lane_id() == waves - 1) s_wave[kMaxScanThreads / kWave] = wi;
    }
    __syncthreads();
    const uint32_t out = s_wave[wave] + incl - v;
    *total = s_wave[kMaxScanThreads / kWave];
    __syncthreads();
    return out;
}


// Exclusive scan of src[0, n) into dst by the whole block (blockDim.x
// threads); returns the total.  Tiles of T * kScanPer values go through LDS:
// coalesced loads, a scan of kScanPer consecutive values per thread,
// coalesced stores.
__device__ __forceinline__ uint32_t scan_column(const uint32_t* src, uint32_t* dst, uint32_t n,
                                                uint32_t* s_tile, uint32_t* s_wave)
{
    const uint32_t T = blockDim.x;
    const uint32_t t = threadIdx.x;
    const uint32_t waves = T / kWave;
    uint32_t carry = 0;
    for (uint32_t tile = 0; tile < n; tile += T * kScanPer) {
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            const uint32_t idx = k * T + t;
            s_tile[scan_pad(idx)] = tile + idx < n ? src[tile + idx] : 0u;
        }
        __syncthreads();
        uint32_t v[kScanPer];
        uint32_t sum = 0;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            v[j] = s_tile[scan_pad(t * kScanPer + j)];
            sum += v[j];
        }
        uint32_t total = 0;
        uint32_t run = carry + block_exclusive(sum, s_wave, waves, &total);
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            s_tile[scan_pad(t * kScanPer + j)] = run;
            run += v[j];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            const uint32_t idx = k * T + t;
            if (tile + idx < n) dst[tile + idx] = s_tile[scan_pad(idx)];
        }
        carry += total;
        __syncthreads();
    }
    return carry;
}

// ---------------------------------------------------------------------------
// Host side.

// Shape dispatch: items per lane 4 / 8 / 16, waves per block 2 / 4 / 8.
template <class F>
hipError_t dispatch(const Shape& s, F&& f)
{
    auto by_waves = [&](auto items) -> hipError_t {
        switch (s.waves) {
        case 2: return f(items, std::integral_constant<int, 2>{});
        case 4: return f(items, std::integral_constant<int, 4>{});
        case 8: return f(items, std::integral_constant<int, 8>{});
        default: return hipErrorInvalidValue;
        }
    };
    switch (s.items) {
    case 4: return by_waves(std::integral_constant<int, 4>{});
    case 8: return by_waves(std::integral_constant<int, 8>{});
    case 16: return by_waves(std::integral_constant<int, 16>{});
    default: return hipErrorInvalidValue;
    }
}

template <class F>
hipError_t dispatch_waves(const Shape& s, F&& f)
{
    switch (s.waves) {
    case 2: return f(std::integral_constant<int, 2>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 8: return f(std::integral_constant<int, 8>{});
    default: return hipErrorInvalidValue;
    }
}

inline uint32_t wave_blocks(uint32_t chunks, int waves)
{
    return std::max<uint32_t>(1, (chunks + waves - 1) / waves);
}

inline dim3 chunk_grid(const Layout& L, int waves) { return dim3(wave_blocks(L.chunks, waves)); }
inline dim3 tile_grid(const Layout& L, int waves)
{
    return dim3(wave_blocks(L.tile_waves ? L.tile_waves : L.tiles, waves));
}

// The layout of one launch of a pass (its walking order from Shape::descending).
inline Layout ordered(const Layout& L, const Shape& s, int bit)
{
    Layout out = L;
    out.descending = (s.descending & bit) ? 1u : 0u;
    return out;
}

// Every kernel goes through launch(): when a timing region is open on this
// thread the dispatch itself carries the region's start / stop events.
template <typename... Params, typename... Args>
hipError_t launch(void (*kernel)(Params...), dim3 grid, dim3 block, hipStream_t stream, Args... args)
{
    hipEvent_t start = nullptr, stop = nullptr;
    timing::claim(&start, &stop);
    hipExtLaunchKernelGGL(kernel, grid, block, 0, stream, start, stop, 0, args...);
    return hipGetLastError();
}

}  // namespace kern
}  // namespace dsdx
