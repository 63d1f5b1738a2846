// runs.hpp -- ghost runs staged through LDS, shared by the ghost copy, the
// reused-plan refresh and reverse_halo (halo_kernels.hip).
//
// A workgroup of kRunWaves waves takes a unit of kRunTiles owner tiles of one
// replica (kRunOwners consecutive boundary owners, for_run_units).  The
// unit's ghosts through send slot s are consecutive ghosts of one receive
// block (their ranks among the slot's owners are consecutive), so the unit
// has one run per slot.  The route pass writes, per unit, where each slot's
// run starts (PlanBuffers::unit_rows); the copy then
//   1. gathers each owner's record into LDS (one lane per owner) and counts
//      the unit's images per tile and slot;
//   2. turns the counts into, for every image of the unit, its index in the
//      unit's runs laid end to end, and map[index] = its owner;
//   3. has lane k of a round write run index k, so a store instruction covers
//      whole cache lines of at most a few runs.
// The per-owner kernels write each tile's short runs lane by lane: with 13
// arrays and 26 receive blocks open at once, the partly written lines leave
// the L2 before their neighbours arrive.  A unit whose images exceed the map
// is written by one lane per owner from the same LDS records.
#pragma once

#include "device.hpp"
#include "owners.hpp"
#include "wave.hpp"

namespace dsdx::kern {

constexpr uint32_t kRunMap = 2 * kRunOwners;
static_assert(kRunTiles == kRunWaves, "one owner tile per wave");

// The runs of one unit from its rows of PlanBuffers::unit_rows: lane s <
// kSlots holds slot s's run (the ghost-list value of its first image, its
// first index in the unit's runs laid end to end, its length), every lane the
// unit's image count.
struct UnitRuns {
    uint32_t first, start, count, images;
};

__device__ __forceinline__ const uint32_t* unit_row(const ReplicaDesc* __restrict__ desc, const ReplicaDesc* d,
                                                    uint32_t e0, const PlanBuffers& B)
{
    return B.unit_rows +
           static_cast<uint64_t>(d->first_unit + static_cast<uint32_t>(d - desc) + e0 / kRunOwners) * kSlots;
}

// first_raw, next_raw: lane s's words of the unit's row and of the next row
// (lanes past the slots load the last slot's and ignore it).
__device__ __forceinline__ UnitRuns unit_runs(uint32_t first_raw, uint32_t next_raw)
{
    const uint32_t lane = lane_id();
    UnitRuns u;
    u.first = lane < kSlots ? first_raw : 0u;
    u.count = lane < kSlots ? next_raw - first_raw : 0u;
    const uint32_t end = wave_inclusive_sum(u.count);
    u.start = end - u.count;
    u.images = uniform(static_cast<uint32_t>(__shfl(static_cast<int>(end), kWave - 1, kWave)));
    return u;
}

// The slot whose run holds index r < images: the last slot whose run starts
// at or before r (an empty run starts where the next one does).
__device__ __forceinline__ uint32_t run_slot(uint32_t start, uint32_t r)
{
    uint32_t s = 0;
#pragma unroll
    for (int k = 1; k < kSlots; ++k) s += static_cast<uint32_t>(__builtin_amdgcn_readlane(start, k)) <= r ? 1u : 0u;
    return s;
}

// Lane s's v; every lane of the wave must be active (a shuffle reads 0 from
// an inactive lane).
__device__ __forceinline__ uint32_t lane_value(uint32_t v, uint32_t s)
{
    return static_cast<uint32_t>(__shfl(static_cast<int>(v), static_cast<int>(s), kWave));
}

struct RunStage {
    double px[kRunOwners], py[kRunOwners], pz[kRunOwners];
    double vx[kRunOwners], vy[kRunOwners], vz[kRunOwners], m[kRunOwners], q[kRunOwners];
    int32_t ty[kRunOwners], id[kRunOwners];
    uint16_t map[kRunMap];      // run index -> owner (index in the unit)
    uint32_t tile_count[kRunTiles][kSlots];   // owners of each tile with each slot
    SlotTable table;
};

// Lane s: how many owners of this wave have slot s (every lane active).
__device__ __forceinline__ uint32_t tile_slot_count(uint32_t slots, uint32_t present)
{
    const uint32_t lane = lane_id();
    uint32_t mine = 0;
    for (uint32_t rest = present; rest; rest &= rest - 1u) {
        const uint32_t s = static_cast<uint32_t>(__builtin_ctz(rest));
        const uint64_t b = __ballot((slots >> s) & 1u);
        if (lane == s) mine = static_cast<uint32_t>(__popcll(b));
    }
    return mine;
}

// Lane s: owners with slot s in the unit's tiles before tile t (after the
// block barrier that follows the counts).
__device__ __forceinline__ uint32_t earlier_tiles(const uint32_t (&count)[kRunTiles][kSlots], uint32_t t)
{
    const uint32_t lane = lane_id();
    uint32_t before = 0;
    if (lane < kSlots)
        for (uint32_t q = 0; q < t; ++q) before += count[q][lane];
    return before;
}

// One unit (e0: its first owner, owners: the replica's owner count), all
// waves of the block.  Each owner's position comes from the classification
// pass (bpos, FROM_ENTRIES) or from the particle arrays, and the rest of its
// record is staged when RECORDS.  Both are compile-time choices: a run-time
// select between an ordinary load of bpos and a nontemporal load of x, y, z
// compiles to one ordinary load.  fields(i, s, rel) runs on the owner's lane
// for each image when RECORDS (i: the owner's particle slot, rel: the image's
// ghost-list value), write(ol, s, rel) once per image from the LDS records.
template <bool STREAM, bool FROM_ENTRIES, bool RECORDS, class Fields, class Write>
__device__ __forceinline__ void ghost_runs_unit(RunStage& sh, const dsdx_particles& p,
                                                const ReplicaDesc* __restrict__ desc, const ReplicaDesc* d,
                                                uint32_t e0, uint32_t owners, const Layout& L,
                                                const PlanBuffers& B, bool with_fields, Fields&& fields,
                                                Write&& write)
{
    const uint32_t lane = lane_id();
    const uint32_t t = uniform(threadIdx.x / kWave);   // this wave's tile of the unit
    const uint64_t slot0 = static_cast<uint64_t>(d->first_chunk) * L.chunk_size;
    const uint32_t* glist = B.ghost_list + d->ghost_list_offset;
    if (t == 0) fill_slot_table(sh.table, desc, d);
    const uint32_t ol = t * kWave + lane;
    const uint32_t e = e0 + ol;
    const bool live = e < owners;
    const uint64_t u = slot0 + (live ? e : e0);
    const uint32_t sl = lane < kSlots ? lane : kSlots - 1u;
    const uint32_t* row = unit_row(desc, d, e0, B);
    const uint32_t first_raw = row[sl];
    const uint32_t next_raw = row[kSlots + sl];
    const uint32_t oi = B.owner_index[u];
    const uint32_t code_raw = B.owner_code[u];
    const uint32_t lut_l = d->bp.lut[lane];
    const uint32_t slots = lane_value(lut_l, live ? code_raw : 0u);
    // 1. records into LDS; per tile and slot the owner count
    if (live) {
        const uint64_t i = d->offset + oi;
        if constexpr (FROM_ENTRIES) {
            const uint32_t E = B.owner_entry[u];
            sh.px[ol] = B.bpos[E];
            sh.py[ol] = B.bpos[B.bpos_stride + E];
            sh.pz[ol] = B.bpos[2 * B.bpos_stride + E];
        } else {
            sh.px[ol] = read_once<STREAM>(p.x + i);
            sh.py[ol] = read_once<STREAM>(p.y + i);
            sh.pz[ol] = read_once<STREAM>(p.z + i);
        }
        if constexpr (RECORDS) {
            sh.vx[ol] = read_once<STREAM>(p.vx + i);
            sh.vy[ol] = read_once<STREAM>(p.vy + i);
            sh.vz[ol] = read_once<STREAM>(p.vz + i);
            sh.m[ol] = read_once<STREAM>(p.mass + i);
            sh.q[ol] = read_once<STREAM>(p.charge + i);
            sh.ty[ol] = read_once<STREAM>(p.type + i);
            sh.id[ol] = read_once<STREAM>(p.global_id + i);
            if (with_fields) {
                const uint32_t* list = glist + B.image_start[u];
                uint32_t k = 0;
                for (uint32_t rest = slots; rest; rest &= rest - 1u, ++k)
                    fields(i, static_cast<uint32_t>(__builtin_ctz(rest)), list[k]);
            }
        }
    }
    const uint32_t present = uniform(wave_or(slots));
    const uint32_t mine = tile_slot_count(slots, present);
    if (lane < kSlots) sh.tile_count[t][lane] = mine;
    const UnitRuns runs = unit_runs(first_raw, next_raw);
    __syncthreads();
    if (runs.images <= kRunMap) {
        // 2. run index of every image -> its owner and slot
        const uint32_t before = earlier_tiles(sh.tile_count, t);
        for (uint32_t rest = present; rest; rest &= rest - 1u) {
            const uint32_t s = static_cast<uint32_t>(__builtin_ctz(rest));
            const uint64_t b = __ballot((slots >> s) & 1u);
            if ((slots >> s) & 1u) {
                const uint32_t r = static_cast<uint32_t>(__builtin_amdgcn_readlane(runs.start, s)) +
                                   static_cast<uint32_t>(__builtin_amdgcn_readlane(before, s)) + lanes_below(b);
                sh.map[r] = static_cast<uint16_t>(ol);
            }
        }
        __syncthreads();
        // 3. lane k of a round writes run index k (every lane takes part in
        // the shuffles of each round)
        for (uint32_t r0 = 0; r0 < runs.images; r0 += kRunWaves * kWave) {
            const uint32_t r = r0 + threadIdx.x;
            const uint32_t s = run_slot(runs.start, r);
            const uint32_t rel = lane_value(runs.first, s) + (r - lane_value(runs.start, s));
            if (r < runs.images) write(static_cast<uint32_t>(sh.map[r]), s, rel);
        }
    } else if (live) {
        // more images than the map holds: one lane per owner
        const uint32_t* list = glist + B.image_start[u];
        uint32_t k = 0;
        for (uint32_t rest = slots; rest; rest &= rest - 1u, ++k)
            write(ol, static_cast<uint32_t>(__builtin_ctz(rest)), list[k]);
    }
    __syncthreads();
}

// reverse_halo over a unit with the ghost forces read as dense runs: one
// wave per owner tile holds its owners' forces; the unit's ghost forces come
// into LDS in run order, then each owner adds its images in ascending send
// slot (the per-owner kernel's order) from there.  The unit's runs come from
// its rows, loaded with the owner maps, so the ghost loads need no barrier;
// they are issued before the owners' forces, whose loads then stay in flight
// across the barrier (a wave waits for its loads in issue order).

struct ReverseStage {
    double gx[kRunMap], gy[kRunMap], gz[kRunMap];
    uint32_t tile_count[kRunTiles][kSlots];   // owners of each tile with each slot
};

// ghost(s, rel, gx, gy, gz) loads the returned force of the image with
// ghost-list value rel sent through slot s (from the ghost particle, or from
// its record).
template <class Ghost>
__device__ __forceinline__ void reverse_runs_unit(ReverseStage& sh, double* __restrict__ fx,
                                                  double* __restrict__ fy, double* __restrict__ fz,
                                                  const ReplicaDesc* __restrict__ desc, const ReplicaDesc* d,
                                                  uint32_t e0, uint32_t owners, const Layout& L,
                                                  const PlanBuffers& B, Ghost&& ghost)
{
    const uint32_t lane = lane_id();
    const uint32_t t = uniform(threadIdx.x / kWave);
    const uint64_t slot0 = static_cast<uint64_t>(d->first_chunk) * L.chunk_size;
    const uint32_t e = e0 + t * kWave + lane;
    const bool live = e < owners;
    const uint64_t u = slot0 + (live ? e : e0);
    const uint32_t sl = lane < kSlots ? lane : kSlots - 1u;
    const uint32_t* row = unit_row(desc, d, e0, B);
    // every load the unit's layout needs, issued together
    const uint32_t first_raw = row[sl];
    const uint32_t next_raw = row[kSlots + sl];
    const uint32_t oi_raw = B.owner_index[u];
    const uint32_t code_raw = B.owner_code[u];
    const uint32_t lut_l = d->bp.lut[lane];
    const UnitRuns runs = unit_runs(first_raw, next_raw);
    const uint32_t slots = lane_value(lut_l, live ? code_raw : 0u);
    const bool staged = runs.images <= kRunMap;
    if (staged) {
        // every lane takes part in each round: a shuffle reads 0 from an inactive lane
        for (uint32_t r0 = 0; r0 < runs.images; r0 += kRunWaves * kWave) {
            const uint32_t r = r0 + threadIdx.x;
            const uint32_t s = run_slot(runs.start, r);
            const uint32_t rel = lane_value(runs.first, s) + (r - lane_value(runs.start, s));
            if (r < runs.images) ghost(s, rel, sh.gx[r], sh.gy[r], sh.gz[r]);
        }
    }
    const uint64_t i = d->offset + (live ? oi_raw : 0u);
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (live) {
        ax = fx[i];
        ay = fy[i];
        az = fz[i];
    }
    if (staged) {
        const uint32_t present = uniform(wave_or(slots));
        const uint32_t mine = tile_slot_count(slots, present);
        if (lane < kSlots) sh.tile_count[t][lane] = mine;
        __syncthreads();
        const uint32_t before = earlier_tiles(sh.tile_count, t);
        for (uint32_t rest = present; rest; rest &= rest - 1u) {
            const uint32_t s = static_cast<uint32_t>(__builtin_ctz(rest));
            const uint64_t b = __ballot((slots >> s) & 1u);
            if ((slots >> s) & 1u) {
                const uint32_t r = static_cast<uint32_t>(__builtin_amdgcn_readlane(runs.start, s)) +
                                   static_cast<uint32_t>(__builtin_amdgcn_readlane(before, s)) + lanes_below(b);
                ax = ax + sh.gx[r];
                ay = ay + sh.gy[r];
                az = az + sh.gz[r];
            }
        }
    } else if (live) {
        // more images than LDS holds: each owner reads its own images
        const uint32_t* list = B.ghost_list + d->ghost_list_offset + B.image_start[u];
        uint32_t k = 0;
        for (uint32_t rest = slots; rest; rest &= rest - 1u, ++k) {
            double gx, gy, gz;
            ghost(static_cast<uint32_t>(__builtin_ctz(rest)), list[k], gx, gy, gz);
            ax = ax + gx;
            ay = ay + gy;
            az = az + gz;
        }
    }
    if (live) {
        fx[i] = ax;
        fy[i] = ay;
        fz[i] = az;
    }
    __syncthreads();
}

// Owned records larger than the memory-side cache take the runs by default
// (below that size the per-owner kernels measured faster: the arrays may be
// cache hits); DSDX_GHOST_RUNS forces either kernel.  The runs need the unit
// rows, which a plan without them (the allocation-free build) does not keep.
inline bool ghost_runs(const Layout& L, const PlanBuffers& B) { return L.ghost_runs != 0 && B.unit_rows; }

inline dim3 run_grid(const Layout& L)
{
    const uint32_t units = L.replicas <= static_cast<uint32_t>(kWave)
                               ? ((L.tile_waves ? L.tile_waves : L.tiles) + kRunTiles - 1) / kRunTiles
                               : L.units;
    return dim3(std::max<uint32_t>(1, units));
}

}  // namespace dsdx::kern
