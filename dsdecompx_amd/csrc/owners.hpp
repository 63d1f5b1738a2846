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

// Owner-major kernels: one wave per owner tile, one lane per boundary owner
// (the owner slots of a replica are dense, in owned order).  A replica whose
// last build overflowed has no owner map and is skipped.
struct OwnerLane {
    const ReplicaDesc* d;
    uint64_t u;      // owner slot
    bool live;       // this lane holds an owner
};

template <int WAVES>
__device__ __forceinline__ OwnerLane owner_of_lane(const ReplicaDesc* __restrict__ desc,
                                                   const uint32_t* __restrict__ tile_replica,
                                                   const Layout& L, const PlanBuffers& B)
{
    OwnerLane o{nullptr, 0, false};
    const uint32_t tile = wave_chunk<WAVES>(L, L.tiles);
    if (tile >= L.tiles) return o;
    const uint32_t r = tile_replica ? uniform(tile_replica[tile]) : 0u;
    o.d = desc + r;
    const uint32_t* totals = B.totals + static_cast<uint64_t>(r) * kTotals;
    if (uniform(totals[kStatusWord])) return o;
    const uint32_t owners = uniform(totals[kColOwners]);
    const uint32_t e = (tile - o.d->first_tile) * kOwnerTile + lane_id();
    if (e >= owners) return o;
    o.u = static_cast<uint64_t>(o.d->first_chunk) * L.chunk_size + e;
    o.live = true;
    return o;
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
