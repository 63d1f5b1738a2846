// abi_internal.hpp -- host-side plan state shared by the C ABI's single-box
// and batched entry points (abi.cpp), process grids (grid.cpp) and
// communication records (wire.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "device.hpp"
#include "dsdx_hip.h"

namespace dsdx::abi {

int fail(int code, const std::string& message);
int hip_fail(hipError_t status, const char* operation);

#define DSDX_HIP(call, what)                                                  \
    do {                                                                      \
        const hipError_t status_ = (call);                                    \
        if (status_ != hipSuccess) return ::dsdx::abi::hip_fail(status_, what); \
    } while (0)

// Device allocation that grows by 1.5x (at least 4 KiB) and keeps nothing on
// growth (src/cuda/cuda_buffer.cpp:10-54 keeps the contents; no caller here
// needs them).  Waits for the device before freeing the old block, which may
// still be read by queued kernels.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    hipError_t reserve(size_t need);
    template <class T> T* as() const { return static_cast<T*>(ptr); }
    ~DeviceBuffer();
};

// A device halo plan over R replicas (R = 1 for the single-box entry points).
struct Plan {
    Shape shape;
    std::vector<ReplicaDesc> desc;        // host copy of the table on the device
    std::vector<uint32_t> chunk_replica;  // host copy (R > 1 only)
    std::vector<uint32_t> tile_replica;   // host copy (R > 1 only)
    std::vector<uint32_t> unit_replica;   // host copy (R > kWave only)
    DeviceBuffer d_desc, d_chunk_replica, d_tile_replica, d_unit_replica;
    bool desc_on_device = false;
    uint32_t chunks = 0;
    uint32_t max_replica_chunks = 0;
    uint32_t tiles = 0;
    uint32_t units = 0;
    // Up to 64 replicas: live owner tiles of the last build whose counts the
    // host read (0 before the first).  Owner-major launches size their grid
    // from it; their waves walk however many tiles the device counts hold.
    uint32_t live_tiles = 0;
    // Single box: the chunk size follows the particle count (auto_items) until
    // a caller or DSDX_ITEMS fixes it.
    bool items_auto = true;

    void note_live_tiles()
    {
        if (desc.size() > static_cast<size_t>(kWave) || !host_totals) return;
        counts_seen = true;
        uint32_t n = 0;
        for (size_t r = 0; r < desc.size(); ++r) {
            const uint32_t* t = host_totals + r * kTotals;
            if (!t[kStatusWord]) n += (t[kColOwners] + kOwnerTile - 1) / kOwnerTile;
        }
        live_tiles = n;
    }

    // Set by set_single (the single-box entry points): such a plan's copy and
    // reverse may take the chunk units (Layout::unit_chunks).
    bool single_box = false;
    // The owner maps and ghost list of the last build are written.  A build
    // over chunk units leaves them out; ensure_routed writes them for the
    // calls that walk them (refresh, field exchanges, ghost owners).
    mutable bool routed = true;
    // The last build's copy took the chunk units and stored the chunk rows
    // (PlanBuffers::chunk_rows) that reverse_halo over chunk units reads.
    bool rows_built = false;
    // host_totals hold the counts of some earlier build (the chunk units are
    // sized from them).
    bool counts_seen = false;

    bool classified = false;   // classify + scan of the current table are on the device
    bool built = false;        // ghosts and owner map written: reverse/forward usable
    bool totals_cached = false;

    DeviceBuffer entries, bpos, counts, prefix, totals, owner_index, owner_code, image_start,
        ghost_list, owner_entry;
    DeviceBuffer recv;   // process grids only: receive-block bases and the grid status
    DeviceBuffer unit_rows;   // first ghost-list value of each slot per ghost-run unit
    DeviceBuffer chunk_rows;  // one box over chunk units: each chunk's prefix row
    bool grid = false;   // replicas send to each other (dsdx_grid), not to themselves
    uint64_t ghost_list_capacity = 0;   // elements
    uint32_t* host_totals = nullptr;    // pinned copy of totals (R x kTotals)
    uint32_t* host_totals_dev = nullptr;  // the same memory, as the device addresses it
    size_t host_totals_words = 0;
    // Single box: the scan's stamped column totals (PlanBuffers::host_counts),
    // the stamp of the last classify and the stream it ran on.
    uint64_t* host_counts = nullptr;
    uint64_t* host_counts_dev = nullptr;
    uint32_t stamp = 0;
    hipStream_t count_stream = nullptr;
    bool counts_captured = false;   // the last classify was captured into a graph

    Plan() = default;
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;
    ~Plan()
    {
        if (host_totals) (void)hipHostFree(host_totals);
        if (host_counts) (void)hipHostFree(host_counts);
    }

    uint32_t chunk_size() const { return static_cast<uint32_t>(shape.items) * kWave; }

    Layout layout() const
    {
        Layout L;
        L.desc = d_desc.as<ReplicaDesc>();
        L.chunk_replica = desc.size() > 1 ? d_chunk_replica.as<uint32_t>() : nullptr;
        L.tile_replica = desc.size() > 1 ? d_tile_replica.as<uint32_t>() : nullptr;
        L.replicas = static_cast<uint32_t>(desc.size());
        L.chunks = chunks;
        L.max_replica_chunks = max_replica_chunks;
        L.chunk_size = chunk_size();
        L.tiles = tiles;
        L.tile_waves = tiles;
        L.unit_replica = desc.size() > static_cast<size_t>(kWave) ? d_unit_replica.as<uint32_t>() : nullptr;
        L.units = units;
        if (desc.size() <= static_cast<size_t>(kWave) && live_tiles)
            L.tile_waves = std::min(tiles, live_tiles + live_tiles / 16 + 8);
        // Reads that a pass makes once stream past the caches when the owned
        // records no longer fit the 256 MB memory-side cache (16 M particles:
        // copy 168 -> 146 us, classify 81 -> 66 us); below that the lines may
        // be cache hits, and streaming them measured slower (1 M particles:
        // copy 27.5 -> 33 us).
        uint64_t owned_total = 0;
        for (const ReplicaDesc& d : desc) owned_total += d.owned;
        L.streaming = owned_total * kRecordBytes > kMallBytes ? 1u : 0u;
        L.ghost_runs = shape.ghost_runs < 0 ? L.streaming : static_cast<uint32_t>(shape.ghost_runs != 0);
        L.xcd_run = static_cast<uint32_t>(shape.xcd_run);
        L.descending = 0;
        L.unit_chunks = L.reverse_unit_chunks = 0;
        if (single_box && !grid && desc.size() == 1 && shape.unit_owners > 0 && L.ghost_runs && unit_rows.ptr &&
            chunk_rows.ptr && chunks) {
            const uint32_t target = static_cast<uint32_t>(shape.unit_owners);
            // The copy's wave w walks the unit's chunks w, w + kRunWaves, ...
            // and the block waits for its slowest wave, so from kRunWaves
            // chunks on a unit holds a multiple of kRunWaves (16 M particles:
            // units of 9 chunks 131.6 us, of 8 chunks 128.7 us).
            const uint32_t g = unit_chunks(target);
            L.unit_chunks = g >= static_cast<uint32_t>(kRunWaves) ? g - g % kRunWaves : g;
            L.reverse_unit_chunks = unit_chunks(std::min(target, kReverseUnitOwners));
        }
        return L;
    }

    // Chunks per unit: about `target` owners (and twice as many images) by
    // the mean of the last counts the host read, 1 to 64 chunks; 4 before any
    // count is known.  Any value gives the same results.
    uint32_t unit_chunks(uint32_t target_owners) const
    {
        if (!counts_seen) return 4;
        const uint64_t owners = host_totals[kColOwners], images = host_totals[kColImages];
        const uint64_t target = target_owners;
        uint64_t g = 64;
        if (owners) g = std::min<uint64_t>(g, target * chunks / owners);
        if (images) g = std::min<uint64_t>(g, 2 * target * chunks / images);
        return static_cast<uint32_t>(std::max<uint64_t>(1, g));
    }

    PlanBuffers buffers() const
    {
        return {entries.as<uint32_t>(), bpos.as<double>(),
                static_cast<uint64_t>(std::max<uint32_t>(1, chunks)) * chunk_size(),
                counts.as<uint32_t>(), prefix.as<uint32_t>(),
                totals.as<uint32_t>(), owner_index.as<uint32_t>(), owner_code.as<uint8_t>(),
                image_start.as<uint32_t>(), ghost_list.as<uint32_t>(),
                owner_entry.as<uint32_t>(), desc.size() == 1 && !grid ? host_counts_dev : nullptr,
                grid ? recv.as<uint32_t>() : nullptr, unit_rows.as<uint32_t>(), stamp,
                chunk_rows.as<uint32_t>()};
    }
};

int make_box_params(const dsdx_box* box, BoxParams* bp);
void route_to_self(ReplicaDesc* d, uint32_t replica);
int make_field_params(const dsdx_field* fields, uint32_t count, FieldParams* fp);
int make_typed_fields(const dsdx_field* fields, uint32_t count, TypedFields* tf);
int check_particles(const dsdx_particles* p, uint64_t owned);
int reserve_buffers(Plan* plan, uint64_t ghost_list_elems);
void number_chunks(Plan* plan);
void number_tiles(Plan* plan);
int upload_table(Plan* plan, hipStream_t stream);
int set_single(Plan* plan, const BoxParams& bp, uint64_t owned, uint64_t capacity,
               uint64_t ghost_list_elems, hipStream_t stream);
int read_totals(Plan* plan, hipStream_t stream);
int wait_counts(Plan* plan);
int run_classify(Plan* plan, const dsdx_particles* p, bool wrap_first, hipStream_t s);
int run_scatter(Plan* plan, const dsdx_particles* p, const FieldParams& fp, hipStream_t s);
int ensure_routed(const Plan* plan, hipStream_t s);
int shape_from_env(Shape* s);

}  // namespace dsdx::abi

// The opaque handles of include/dsdx_hip.h: a single-box plan and a batch are
// plans over one and over R replicas.
struct dsdx_plan : dsdx::abi::Plan {};
struct dsdx_batch : dsdx::abi::Plan {};
