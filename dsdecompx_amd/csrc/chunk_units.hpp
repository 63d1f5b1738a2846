// chunk_units.hpp -- the ghost copy and reverse_halo of one box in units of
// whole chunks, read straight from the classification's entries and the
// scan's prefixes, so that a rebuild needs no route pass
// (halo_kernels.hip: copy_chunks_kernel, reverse_chunks_kernel).
//
// A unit is Layout::unit_chunks consecutive chunks (reverse_unit_chunks for
// reverse_halo), one workgroup of kRunWaves waves.  A chunk's boundary
// owners are its compacted entries, in owned order, and the
// ghost of an owner through send slot s has ghost-list value
//   slot_first[s] + prefix[s][chunk] + (its rank among the chunk's owners with s),
// slot_first[s] being the first ghost of the receive block slot s feeds, the
// numbering the route pass writes into the ghost list.  So the unit's images
// through slot s are one run of consecutive ghosts, from slot_first[s] +
// prefix[s][c0] to slot_first[s] + prefix[s][c1].  In the copy wave w walks
// the unit's chunks w, w + kRunWaves, ..., ranks each chunk's owners per slot
// with ballots, stages their records in LDS and maps each image to its run
// index; the unit's runs are then written with consecutive lanes (as the
// staged runs of runs.hpp are).  reverse_halo reads the unit's ghost forces
// as dense runs into LDS, copies the unit's entries into LDS in owner order,
// and wave t adds the images of owners 64 t .. 64 t + 63 in ascending send
// slot.  A unit with more owners or images than LDS holds is copied, or
// reduced, one lane per owner.
#pragma once

#include "device.hpp"
#include "owners.hpp"
#include "runs.hpp"
#include "wave.hpp"

namespace dsdx::kern {

// Owners and images of one unit that fit the copy's LDS stage; the host picks
// unit_chunks so that a unit holds about Shape::unit_owners owners.
constexpr uint32_t kChunkOwners = 640;
constexpr uint32_t kChunkMap = 2 * kChunkOwners;

// The unit's layout, the same in every wave of the block.
struct ChunkUnit {
    uint32_t c0, c1;       // the unit's chunks [c0, c1)
    uint32_t owner0;       // boundary owners of the chunks before c0
    uint32_t owners;       // the unit's boundary owners
    uint32_t pre0;         // lane l < kCols: prefix[l][c0]
    uint32_t slot_first;   // lane s < kSlots: ghost-list value of the first ghost of receive block 25 - s
    uint32_t overflow;     // the build's overflow bits (the route pass's test)
    UnitRuns runs;         // lane s < kSlots: the unit's run of slot s
};

__device__ __forceinline__ uint32_t chunk_units_of(const Layout& L)
{
    return (L.chunks + L.unit_chunks - 1) / L.unit_chunks;
}

// The unit of chunks [c0, c1) from lane l's prefix of column l at c0 and at
// c1 (the column total when c1 is the last chunk) and the replica's totals
// row.  Every lane of the wave must be active.
__device__ __forceinline__ ChunkUnit chunk_unit_of(const ReplicaDesc* d, uint32_t c0, uint32_t c1, uint32_t pre0,
                                                   uint32_t pre1, uint32_t total_row)
{
    const uint32_t lane = lane_id();
    ChunkUnit u;
    u.c0 = c0;
    u.c1 = c1;
    u.pre0 = pre0;
    const uint32_t received = static_cast<uint32_t>(__builtin_amdgcn_readlane(total_row, kColReceived));
    const uint32_t images = static_cast<uint32_t>(__builtin_amdgcn_readlane(total_row, kColImages));
    u.overflow = 0;
    if (static_cast<uint64_t>(d->owned) + received > d->capacity) u.overflow |= kOverflowStorage;
    if (images > d->ghost_list_capacity) u.overflow |= kOverflowGhostList;
    // receive block m holds send slot 25 - m: block m starts after the
    // totals of slots 25 .. 26 - m
    const uint32_t from = lane < kSlots ? kSlots - 1 - lane : 0u;
    const uint32_t block_size = static_cast<uint32_t>(__shfl(static_cast<int>(total_row), static_cast<int>(from), kWave));
    const uint32_t recv_base = wave_inclusive_sum(block_size) - block_size;
    u.slot_first = lane < kSlots ? static_cast<uint32_t>(__shfl(static_cast<int>(recv_base), static_cast<int>(from), kWave))
                                 : 0u;
    u.owner0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(u.pre0, kColOwners));
    u.owners = static_cast<uint32_t>(__builtin_amdgcn_readlane(pre1, kColOwners)) - u.owner0;
    u.runs = unit_runs(u.slot_first + u.pre0, u.slot_first + pre1);
    return u;
}

// Unit `unit` of the layout.  rows: the chunks' prefix rows
// (PlanBuffers::chunk_rows, once the copy has stored them), else the prefix
// columns are read.  Every lane of the wave must be active.
__device__ __forceinline__ ChunkUnit chunk_unit(const ReplicaDesc* d, uint32_t unit, const Layout& L,
                                                const PlanBuffers& B, const uint32_t* rows = nullptr)
{
    const uint32_t lane = lane_id();
    const uint32_t c0 = unit * L.unit_chunks;
    const uint32_t c1 = min(L.chunks, c0 + L.unit_chunks);
    const uint64_t col = lane < kCols ? lane : 0u;
    const uint32_t total_row = B.totals[lane & (kTotals - 1)];
    // past the last chunk the prefix is the column total
    uint32_t pre0, pre1;
    if (rows) {
        pre0 = rows[static_cast<uint64_t>(c0) * kTotals + col];
        pre1 = c1 < L.chunks ? rows[static_cast<uint64_t>(c1) * kTotals + col] : total_row;
    } else {
        pre0 = B.prefix[col * L.chunks + c0];
        pre1 = c1 < L.chunks ? B.prefix[col * L.chunks + c1] : total_row;
    }
    return chunk_unit_of(d, c0, c1, pre0, pre1, total_row);
}


// The copy's wave takes its chunk's owners two rounds of 64 at a time, every
// load of both rounds issued before the first is used: at 16 M particles a
// chunk holds about 65 boundary owners, and a second round whose loads
// waited for the first's cost 8 us in a wave-per-chunk reverse_halo.
constexpr uint32_t kChunkRounds = 2;

// One chunk of the unit as its wave sees it: lane l < kCols holds
// prefix[l][c]; the first kChunkRounds * 64 entries are loaded with it (the
// entry region of a chunk always holds chunk_size >= 256 entries).
struct ChunkHead {
    uint32_t pc;
    uint32_t entry[kChunkRounds];
    uint32_t n;            // boundary owners of the chunk
};

__device__ __forceinline__ ChunkHead chunk_head(uint32_t c, const Layout& L, const PlanBuffers& B)
{
    const uint32_t lane = lane_id();
    const uint64_t col = lane < kCols ? lane : 0u;
    ChunkHead h;
    h.pc = B.prefix[col * L.chunks + c];
#pragma unroll
    for (uint32_t q = 0; q < kChunkRounds; ++q)
        h.entry[q] = B.entries[static_cast<uint64_t>(c) * L.chunk_size + q * kWave + lane];
    h.n = uniform(B.counts[static_cast<uint64_t>(kColOwners) * L.chunks + c]);
    return h;
}

// The part of a chunk's head the tile-lane reverse needs: the owners before
// the chunk (uniform, in pc, from its prefix row), its owner count and its
// first entries.
__device__ __forceinline__ ChunkHead chunk_owner_head(uint32_t c, const Layout& L, const PlanBuffers& B)
{
    const uint32_t lane = lane_id();
    ChunkHead h;
    h.pc = uniform(B.chunk_rows[static_cast<uint64_t>(c) * kTotals + kColOwners]);
#pragma unroll
    for (uint32_t q = 0; q < kChunkRounds; ++q)
        h.entry[q] = B.entries[static_cast<uint64_t>(c) * L.chunk_size + q * kWave + lane];
    h.n = uniform(B.counts[static_cast<uint64_t>(kColOwners) * L.chunks + c]);
    return h;
}

// Entry e = e0 + 64 q + lane of the chunk (0 past its owners): from the head
// for the first rounds, else loaded.
__device__ __forceinline__ uint32_t chunk_entry(const ChunkHead& h, uint32_t e0, uint32_t q, uint64_t ebase,
                                                const PlanBuffers& B)
{
    const uint32_t e = e0 + q * kWave + lane_id();
    if (e >= h.n) return 0u;
    return e0 == 0 ? h.entry[q] : B.entries[ebase + e];
}

// Calls image(s, k, mine) for each send slot s present in this round of 64
// owners (ascending), k being the owner's rank among the chunk's owners with
// slot s; `run` (lane s: owners with s in earlier rounds) advances.  Every
// lane of the wave must be active.
template <class Image>
__device__ __forceinline__ void rank_round(uint32_t slots, uint32_t& run, Image&& image)
{
    const uint32_t lane = lane_id();
    for (uint32_t present = uniform(wave_or(slots)); present; present &= present - 1u) {
        const uint32_t s = uniform(static_cast<uint32_t>(__builtin_ctz(present)));
        const bool mine = (slots >> s) & 1u;
        const uint64_t b = __ballot(mine);
        const uint32_t k = static_cast<uint32_t>(__builtin_amdgcn_readlane(run, s)) + lanes_below(b);
        image(s, k, mine);
        if (lane == s) run += static_cast<uint32_t>(__popcll(b));
    }
}

}  // namespace dsdx::kern
