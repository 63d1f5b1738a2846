// owners.hpp -- device helpers of the owner-major kernels: one wave per tile
// of 64 boundary owners of a replica (dense owner slots, owned order), one
// lane per owner, and the slots of an owner's ghosts.
#pragma once

#include "device.hpp"
#include "wave.hpp"

namespace dsdx::kern {

// Particle slot of ghost `rel` sent by replica d through slot s: ghosts
// follow the receiving replica's owned particles.
__device__ __forceinline__ uint64_t ghost_slot(const ReplicaDesc* __restrict__ desc,
                                               const ReplicaDesc* d, uint32_t s, uint32_t rel)
{
    if (d->self_routed) return d->offset + d->owned + rel;
    const ReplicaDesc* to = desc + d->dest[s];
    return to->offset + to->owned + rel;
}

// Per-wave table of the slots, in LDS: the particle slot of the first ghost
// of the receive block each send slot feeds, and the slot's shift.  Lanes
// read entries of their own slots from it, so the ghost loop of an owner-major
// kernel issues no vector memory load for them.
struct SlotTable {
    uint64_t ghost0[kSlots];
    double shift[kSlots][3];
};

// Fills the wave's table; every lane of the wave must be active.
__device__ __forceinline__ const SlotTable& fill_slot_table(SlotTable& t,
                                                           const ReplicaDesc* __restrict__ desc,
                                                           const ReplicaDesc* d)
{
    const uint32_t lane = lane_id();
    if (lane < kSlots) {
        t.ghost0[lane] = (d->self_routed || d->dest[lane] != kNoPeer) ? ghost_slot(desc, d, lane, 0u) : 0u;
        t.shift[lane][0] = d->bp.shift[lane][0];
        t.shift[lane][1] = d->bp.shift[lane][1];
        t.shift[lane][2] = d->bp.shift[lane][2];
    }
    // LDS operations of one wave complete in order: the table is visible to
    // every lane of the wave after this point without a block barrier.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    return t;
}

// Ghost-list entries an owner-major kernel loads before its ghost loop.  On
// gfx9 the vector memory counter counts stores as well as loads, so a load
// issued inside a loop of stores makes the wave wait for the stores too.
constexpr uint32_t kPreloaded = 2;

// Owner-major kernels: one wave per owner tile, one lane per boundary owner
// (the owner slots of a replica are dense, in owned order).  A replica whose
// last build overflowed has no owner map and is skipped.
struct OwnerLane {
    const ReplicaDesc* d;
    uint64_t u;      // owner slot
    bool live;       // this lane holds an owner
};

// Calls body(lane) for every owner tile this wave walks, all lanes active
// (a lane without an owner has live == false).  The launch has L.tile_waves
// waves; wave w walks tiles w, w + L.tile_waves, ...
//
// Up to 64 replicas: the tiles walked are exactly the live ones, ceil(owners_r
// / kOwnerTile) of each replica r in replica order, counted from the device
// totals (lane r reads replica r's), so the launch may have fewer waves than
// tiles or more, and no wave visits an empty tile.  A replica whose build
// overflowed has none.  More replicas: the tiles are those numbered on the
// host (bound_r per replica), one per wave, and a tile past its replica's
// owners is skipped.
template <int WAVES, class Body>
__device__ __forceinline__ void for_owner_tiles(const ReplicaDesc* __restrict__ desc,
                                                const uint32_t* __restrict__ tile_replica,
                                                const Layout& L, const PlanBuffers& B, Body&& body)
{
    const uint32_t lane = lane_id();
    const uint32_t first = uniform(xcd_block(L.xcd_run) * WAVES + threadIdx.x / kWave);
    const uint32_t stride = gridDim.x * WAVES;
    if (L.replicas <= kWave) {
        uint32_t owners_r = 0;
        if (lane < L.replicas) {
            const uint32_t* t = B.totals + static_cast<uint64_t>(lane) * kTotals;
            owners_r = t[kStatusWord] ? 0u : t[kColOwners];
        }
        const uint32_t live_r = (owners_r + kOwnerTile - 1) / kOwnerTile;
        const uint32_t end_r = wave_inclusive_sum(live_r);
        const uint32_t begin_r = end_r - live_r;
        const uint32_t total = uniform(__shfl(end_r, kWave - 1, kWave));
        for (uint32_t c = first; c < total; c += stride) {
            const uint32_t tile = L.descending ? total - 1u - c : c;
            const uint64_t hit = __ballot(begin_r <= tile && tile < end_r);   // one lane
            const uint32_t r = static_cast<uint32_t>(__builtin_ctzll(hit));
            const uint32_t e0 = (tile - __builtin_amdgcn_readlane(begin_r, r)) * kOwnerTile;
            const uint32_t owners = __builtin_amdgcn_readlane(owners_r, r);
            const ReplicaDesc* d = desc + r;
            const uint32_t e = e0 + lane;
            body(OwnerLane{d, static_cast<uint64_t>(d->first_chunk) * L.chunk_size + e, e < owners});
        }
        return;
    }
    for (uint32_t c = first; c < L.tiles; c += stride) {
        const uint32_t tile = L.descending ? L.tiles - 1u - c : c;
        const uint32_t r = uniform(tile_replica[tile]);
        const ReplicaDesc* d = desc + r;
        const uint32_t* totals = B.totals + static_cast<uint64_t>(r) * kTotals;
        if (uniform(totals[kStatusWord])) continue;
        const uint32_t owners = uniform(totals[kColOwners]);
        const uint32_t e0 = (tile - d->first_tile) * kOwnerTile;
        if (e0 >= owners) continue;
        const uint32_t e = e0 + lane;
        body(OwnerLane{d, static_cast<uint64_t>(d->first_chunk) * L.chunk_size + e, e < owners});
    }
}

// Ghost-run units: kRunTiles consecutive owner tiles of one replica, one
// workgroup each.  Calls body(d, e0, owners) for every unit this workgroup
// walks (units g, g + grid, ...): d the replica, e0 its first owner
// (relative to the replica's owner slots), owners the replica's owner count.
// Every wave of the block computes the same units, so the calls are
// block-uniform and the body may hold block barriers.  Up to 64 replicas the
// units walked are exactly the live ones, counted from the device totals as
// for_owner_tiles counts tiles; more replicas: the units numbered on the
// host (ceil(tiles_r / kRunTiles) per replica), one past its replica's
// owners skipped.  A replica whose build overflowed has none.
template <class Body>
__device__ __forceinline__ void for_run_units(const ReplicaDesc* __restrict__ desc,
                                              const uint32_t* __restrict__ unit_replica,
                                              const Layout& L, const PlanBuffers& B, Body&& body)
{
    const uint32_t lane = lane_id();
    const uint32_t first = xcd_block(L.xcd_run);
    if (L.replicas <= kWave) {
        uint32_t owners_r = 0;
        if (lane < L.replicas) {
            const uint32_t* t = B.totals + static_cast<uint64_t>(lane) * kTotals;
            owners_r = t[kStatusWord] ? 0u : t[kColOwners];
        }
        const uint32_t units_r = (owners_r + kRunOwners - 1) / kRunOwners;
        const uint32_t end_r = wave_inclusive_sum(units_r);
        const uint32_t begin_r = end_r - units_r;
        const uint32_t total = uniform(__shfl(end_r, kWave - 1, kWave));
        for (uint32_t c = first; c < total; c += gridDim.x) {
            const uint32_t unit = L.descending ? total - 1u - c : c;
            const uint64_t hit = __ballot(begin_r <= unit && unit < end_r);   // one lane
            const uint32_t r = static_cast<uint32_t>(__builtin_ctzll(hit));
            const uint32_t e0 = (unit - __builtin_amdgcn_readlane(begin_r, r)) * kRunOwners;
            body(desc + r, e0, static_cast<uint32_t>(__builtin_amdgcn_readlane(owners_r, r)));
        }
        return;
    }
    for (uint32_t c = first; c < L.units; c += gridDim.x) {
        const uint32_t unit = L.descending ? L.units - 1u - c : c;
        const uint32_t r = uniform(unit_replica[unit]);
        const ReplicaDesc* d = desc + r;
        const uint32_t* totals = B.totals + static_cast<uint64_t>(r) * kTotals;
        if (uniform(totals[kStatusWord])) continue;
        const uint32_t owners = uniform(totals[kColOwners]);
        const uint32_t e0 = (unit - d->first_unit) * kRunOwners;
        if (e0 >= owners) continue;
        body(d, e0, owners);
    }
}

// The ghosts of one owner: its slot set and its ghost-list entries (one per
// set bit, ascending slot).  at(k, s) is the particle slot of the k-th ghost,
// sent through slot s.
struct OwnerGhosts {
    const ReplicaDesc* desc;
    const ReplicaDesc* d;
    const uint32_t* list;
    uint32_t slots;
    __device__ uint64_t at(uint32_t k, uint32_t s) const { return ghost_slot(desc, d, s, list[k]); }
};

__device__ __forceinline__ OwnerGhosts ghosts_of(const ReplicaDesc* __restrict__ desc,
                                                 const OwnerLane& o, const PlanBuffers& B)
{
    return {desc, o.d, B.ghost_list + o.d->ghost_list_offset + B.image_start[o.u],
            o.d->bp.lut[B.owner_code[o.u]]};
}


// The first kPreloaded images of an owner, read before any of its stores or
// sums: image k's send slot and ghost-list entry.  Past the owner's last
// image (k >= n, n >= 1) they repeat the last image, so every lane loads
// without a branch; image k is used only when k < n.
struct FirstImages {
    uint32_t n;
    uint32_t slot[kPreloaded];
    uint32_t rel[kPreloaded];
    uint32_t rest;   // slot set of the images after the first kPreloaded
};

__device__ __forceinline__ FirstImages first_images(const OwnerGhosts& g)
{
    FirstImages f;
    f.n = static_cast<uint32_t>(__popc(g.slots));
    uint32_t rest = g.slots, last = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPreloaded; ++k) {
        f.rel[k] = g.list[k < f.n ? k : f.n - 1];
        if (rest) last = static_cast<uint32_t>(__builtin_ctz(rest));
        f.slot[k] = last;
        rest &= rest - 1u;
    }
    f.rest = rest;
    return f;
}

// One step of reduce_field_values (src/halo.cpp:1351-1372).
template <class T>
__device__ __forceinline__ T reduce_value(T dst, T v, int32_t op)
{
    switch (op) {
    case DSDX_SUM: return static_cast<T>(dst + v);
    case DSDX_MIN: return (v < dst) ? v : dst;   // std::min(dst, v)
    case DSDX_MAX: return (dst < v) ? v : dst;   // std::max(dst, v)
    default: return v;                            // Replace: last slot wins
    }
}

}  // namespace dsdx::kern
