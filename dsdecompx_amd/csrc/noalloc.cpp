// noalloc.cpp -- dsdx_halo_build_noalloc: the single-box halo build with its
// plan state in a fixed device array instead of plan buffers, for callers
// that cannot allocate (code under graph capture, serving loops).  The same
// kernels as dsdx_halo_build run on the same inputs, so the ghosts are the
// same bit for bit.
//
// The array lives in this file as a static __device__ symbol.  On the host
// its name is only a stand-in: its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and kept, and every
// kernel gets that address.  Each GPU holds its own copy until the process
// ends; every call on a GPU shares it, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "abi_internal.hpp"
#include "dsdx_hip.h"
#include "kernels.hpp"

namespace dsdx::abi {
namespace {

// Fixed launch shape: chunks of 4 items per lane (256 particles), 4 waves per
// block, scan blocks of 1024 threads.
constexpr int kItems = 4;
constexpr uint64_t kChunk = static_cast<uint64_t>(kItems) * kWave;
constexpr uint64_t kMaxChunks = DSDX_NOALLOC_MAX_OWNED / kChunk;
constexpr uint64_t kSlotsMax = kMaxChunks * kChunk;
static_assert(DSDX_NOALLOC_MAX_OWNED % kChunk == 0, "whole chunks");

constexpr uint64_t up256(uint64_t b) { return (b + 255) & ~uint64_t{255}; }

// Regions of the array, each 256-byte aligned.
constexpr uint64_t kOffDesc = 0;
constexpr uint64_t kOffTotals = kOffDesc + up256(sizeof(ReplicaDesc));
constexpr uint64_t kOffCounts = kOffTotals + up256(kTotals * 4);
constexpr uint64_t kOffPrefix = kOffCounts + up256(kCols * kMaxChunks * 4);
constexpr uint64_t kOffEntries = kOffPrefix + up256(kCols * kMaxChunks * 4);
constexpr uint64_t kOffBpos = kOffEntries + up256(kSlotsMax * 4);
constexpr uint64_t kOffOwnerIndex = kOffBpos + up256(3 * kSlotsMax * 8);
constexpr uint64_t kOffOwnerCode = kOffOwnerIndex + up256(kSlotsMax * 4);
constexpr uint64_t kOffImageStart = kOffOwnerCode + up256(kSlotsMax);
constexpr uint64_t kOffOwnerEntry = kOffImageStart + up256(kSlotsMax * 4);
constexpr uint64_t kOffGhostList = kOffOwnerEntry + up256(kSlotsMax * 4);
constexpr uint64_t kScratchBytes = kOffGhostList + up256(DSDX_NOALLOC_MAX_GHOSTS * 4);

}  // namespace
}  // namespace dsdx::abi

static __device__ __align__(256) unsigned char scratch[dsdx::abi::kScratchBytes];

namespace dsdx::abi {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's copy of `scratch`, looked up on the
// first call with the device current.
int scratch_address(hipStream_t stream, unsigned char** out)
{
    int device = -1;
    DSDX_HIP(hipGetDevice(&device), "hipGetDevice");
    int stream_device = -1;
    DSDX_HIP(hipStreamGetDevice(stream, &stream_device), "hipStreamGetDevice");
    if (stream_device != device)
        return fail(DSDX_EINVAL, "dsdx_halo_build_noalloc: the stream must belong to the current device");
    if (device < 0 || device >= kMaxDevices) return fail(DSDX_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        DSDX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
        g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[device];
    return DSDX_OK;
}

}  // namespace
}  // namespace dsdx::abi

using namespace dsdx;
using namespace dsdx::abi;

extern "C" int dsdx_halo_build_noalloc(const dsdx_box* box, const dsdx_particles* particles,
                                       const dsdx_field* fields, uint32_t field_count, uint64_t owned,
                                       uint64_t capacity, int32_t* ghost_owners, void* stream,
                                       uint64_t* ghosts)
{
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (owned > DSDX_NOALLOC_MAX_OWNED)
        return fail(DSDX_EINVAL, "dsdx_halo_build_noalloc takes at most 262144 owned particles");
    if (capacity < owned) return fail(DSDX_EINVAL, "particle capacity below the owned count");
    BoxParams bp;
    if (int rc = make_box_params(box, &bp)) return rc;
    if (int rc = check_particles(particles, owned)) return rc;
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    unsigned char* base = nullptr;
    if (int rc = scratch_address(s, &base)) return rc;

    const uint32_t chunks = static_cast<uint32_t>((owned + kChunk - 1) / kChunk);
    ReplicaDesc d;
    std::memset(&d, 0, sizeof(d));
    d.bp = bp;
    d.capacity = capacity;
    d.owned = static_cast<uint32_t>(owned);
    d.chunks = chunks;
    d.ghost_list_capacity = std::min<uint64_t>(capacity - owned, DSDX_NOALLOC_MAX_GHOSTS);
    route_to_self(&d, 0);
    // owner tiles: a build that fits has at most min(owned, ghosts) boundary owners
    const uint64_t bound = std::min<uint64_t>(owned, d.ghost_list_capacity);
    const uint32_t tiles = static_cast<uint32_t>((bound + kOwnerTile - 1) / kOwnerTile);

    Shape shape;
    shape.items = kItems;
    shape.waves = 4;
    shape.scan_threads = 1024;
    Layout L{};
    L.desc = reinterpret_cast<const ReplicaDesc*>(base + kOffDesc);
    L.replicas = 1;
    L.chunks = chunks;
    L.max_replica_chunks = chunks;
    L.chunk_size = static_cast<uint32_t>(kChunk);
    L.tiles = tiles;
    L.tile_waves = tiles;
    L.units = (tiles + kRunTiles - 1) / kRunTiles;
    L.ghost_runs = shape.ghost_runs > 0 ? 1u : 0u;
    L.xcd_run = static_cast<uint32_t>(shape.xcd_run);
    PlanBuffers B{};
    B.entries = reinterpret_cast<uint32_t*>(base + kOffEntries);
    B.bpos = reinterpret_cast<double*>(base + kOffBpos);
    B.bpos_stride = static_cast<uint64_t>(std::max<uint32_t>(1, chunks)) * kChunk;
    B.counts = reinterpret_cast<uint32_t*>(base + kOffCounts);
    B.prefix = reinterpret_cast<uint32_t*>(base + kOffPrefix);
    B.totals = reinterpret_cast<uint32_t*>(base + kOffTotals);
    B.owner_index = reinterpret_cast<uint32_t*>(base + kOffOwnerIndex);
    B.owner_code = base + kOffOwnerCode;
    B.image_start = reinterpret_cast<uint32_t*>(base + kOffImageStart);
    B.ghost_list = reinterpret_cast<uint32_t*>(base + kOffGhostList);
    B.owner_entry = reinterpret_cast<uint32_t*>(base + kOffOwnerEntry);

    // Every word a pass reads is written earlier in the same call (the
    // descriptor, then counts, prefix and totals, then the entries and owner
    // maps of the boundary particles), so nothing is cleared.
    DSDX_HIP(launch_store_desc(d, reinterpret_cast<ReplicaDesc*>(base + kOffDesc), s), "store replica table");
    DSDX_HIP(launch_classify(particles->x, particles->y, particles->z, false, L, B, shape, s),
             "launch halo classify");
    DSDX_HIP(launch_scan(L, B, shape, s), "launch halo scan");
    DSDX_HIP(launch_route(L, B, shape, s), "launch halo route");
    DSDX_HIP(launch_copy(*particles, fp, L, B, shape, s), "launch halo copy");
    if (ghost_owners)
        DSDX_HIP(launch_ghost_owners(ghost_owners, L, B, shape, s), "launch ghost owners");
    if (!ghosts) return DSDX_OK;
    uint32_t totals[kTotals] = {};
    DSDX_HIP(hipMemcpyAsync(totals, B.totals, sizeof(totals), hipMemcpyDeviceToHost, s),
             "copy halo counts to the host");
    DSDX_HIP(hipStreamSynchronize(s), "wait for halo counts");
    *ghosts = totals[kColImages];
    if (totals[kStatusWord] & kOverflowStorage)
        return fail(DSDX_ECAPACITY, "device halo output capacity is too small");
    if (totals[kStatusWord] & kOverflowGhostList)
        return fail(DSDX_ECAPACITY, "dsdx_halo_build_noalloc holds at most 524288 ghosts");
    return DSDX_OK;
}
