// device.hpp -- shared host/device definitions of the gfx950 halo kernels.
//
// Every launch covers a batch of R independent boxes ("replicas"); a single
// box is a batch of one.  Replica r owns the particle slots
// [offset_r, offset_r + capacity_r) of the SoA arrays (owned first, then its
// ghosts) and the chunks [first_chunk_r, first_chunk_r + chunks_r) of the
// global chunk numbering; a chunk is kItems * 64 consecutive owned particles
// of one replica, processed by one wave.
//
// Plan buffers in HBM (u32 unless noted), C = total chunks, K = chunk size:
//   entries[C * K]          per chunk, its boundary particles compacted in
//                           owned order: (index in chunk << 8) | slab code
//   bpos[3][C * K] (f64)    the positions of those entries (written by the
//                           classify pass, read densely by the scatter)
//   counts[kCols][C]        per chunk: send count of each of the 26 slots,
//                           boundary owners, images (column-major)
//   prefix[kCols][C]        exclusive scan of counts over each replica's chunks
//   totals[R][kTotals]      per replica: column totals, the status word, then
//                           the ghosts it receives (kColReceived)
//   owner_index[C * K]      owner slot first_chunk_r * K + u -> owned index
//   owner_code[C * K] (u8)  ... -> slab code
//   image_start[C * K]      ... -> first entry of its ghosts in ghost_list
//   ghost_list[...]         per replica at ghost_list_offset_r: per owner its
//                           ghosts (index relative to the replica's owned
//                           count) in ascending send-slot order
//   owner_entry[C * K]      ... -> its entry index (chunk * K + rank)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "dsdx_hip.h"

namespace dsdx {

constexpr int kWave = 64;
constexpr int kSlots = 26;
constexpr int kCols = 28;          // 26 slots + owners + images
constexpr int kColOwners = 26;
constexpr int kColImages = 27;
constexpr int kStatusWord = kCols; // totals[r][kStatusWord]: overflow bits
constexpr int kColReceived = 29;   // totals[r][kColReceived]: ghosts the replica receives
constexpr int kTotals = 32;        // words per replica in totals
constexpr uint32_t kNoPeer = 0xFFFFFFFFu;
constexpr uint32_t kOverflowStorage = 1u;
constexpr uint32_t kOverflowGhostList = 2u;

// Slab code: bit 2a = low slab of axis a (x < lower+width), bit 2a+1 = high
// slab (x > upper-width).  lut[code] is the set of active send slots whose
// direction is covered by the code (src/halo.cpp:212-221).
struct BoxParams {
    double lo_thr[3];
    double hi_thr[3];
    double shift[kSlots][3];
    uint32_t lut[64];
    double lower[3];
    double extent[3];
    int32_t periodic[3];
    int32_t pad;
};

struct ReplicaDesc {
    BoxParams bp;
    uint64_t offset;               // first particle slot of the replica
    uint64_t capacity;             // particle slots (owned + ghosts)
    uint64_t ghost_list_offset;
    uint64_t ghost_list_capacity;
    uint32_t owned;
    uint32_t first_chunk;
    uint32_t chunks;
    uint32_t first_tile;           // first owner tile (kOwnerTile owner slots each)
    // Replica that receives send slot s (into its receive block 25 - s):
    // the replica itself on a self-loopback box, the neighbouring subdomain
    // on a process grid, kNoPeer where the slot has no neighbour.
    uint32_t dest[kSlots];
    uint32_t self_routed;          // every dest[s] is the replica itself
    uint32_t first_unit;           // first ghost-run unit (kRunTiles owner tiles each)
};

// The geometry migrate's wrap needs for one box, passed by value.
struct WrapBox {
    double lower[3];
    double extent[3];
    int32_t periodic[3];
    uint32_t owned;
};

// Bits of Shape::descending.
constexpr int kOrderMigrate = 1;
constexpr int kOrderClassify = 2;
constexpr int kOrderRoute = 4;
constexpr int kOrderCopy = 8;
constexpr int kOrderOwners = 16;
constexpr int kOrderMigrateScatter = 32;   // process grids: the migration's record moves
constexpr int kOrderFused = 64;            // the fused migrate + classify pass

// Launch shape of the classify / scatter / owner-major kernels: items per
// lane (the chunk is kWave * items particles) and waves per block.
struct Shape {
    int items = 16;
    int waves = 4;
    int scan_threads = 1024;
    int xcd_run = 16;   // consecutive blocks per XCD turn (0: hardware order)
    // Bit mask of passes that walk their chunks / tiles last to first, each
    // choice measured at 16 M particles.  migrate, the first pass over x, y,
    // z after the application wrote them, walks backwards: the most recently
    // written lines are the ones still held by the 256 MB memory-side cache
    // (70.8 against 97 us forwards).  The classify pass that follows walks
    // backwards too (64.4 against 65.9 us).  The owner-major passes after the
    // rebuild (reverse, refresh, field exchanges) walk backwards for the
    // reason migrate does: the application has just written the forces (or
    // positions) in index order, and the highest owners' lines are the cached
    // ones (reverse 95 against 111 us).  The route and copy passes walking
    // backwards as well measured 0.7 us faster per step (reverse_halo after
    // them).  The fused migrate + classify pass, which takes its chunk in
    // rounds of 8 items, walks forwards (71.0 against 74.9 us).
    int descending = kOrderMigrate | kOrderClassify | kOrderRoute | kOrderCopy | kOrderOwners;
    // Ghost writes staged through LDS (copy_runs_kernel): -1 when the owned
    // records exceed the memory-side cache, 0 never, 1 always.
    int ghost_runs = -1;
    // reverse_halo with the ghost forces read as dense runs (1), per owner
    // (0), or as the ghost writes go (-1: runs wherever the copy stages them).
    int reverse_runs = -1;
    // One box whose ghost writes take the staged runs: the copy walks units
    // of whole chunks holding about this many boundary owners (reverse_halo
    // at most kReverseUnitOwners), and the build runs no route pass (0:
    // units of owner tiles after a route pass, as batches and grids do).
    int unit_owners = 544;
};

// Owner tiles: the owner slots [0, bound_r) of replica r (relative to
// first_chunk_r * K), bound_r = min(owned_r, capacity_r - owned_r,
// ghost_list_capacity_r) >= its boundary owners, cut into tiles of
// kOwnerTile; one wave of the copy pass per tile.
constexpr uint32_t kOwnerTile = kWave;

// Ghost-run units of the staged copy: kRunTiles consecutive owner tiles of one
// replica, one workgroup of kRunWaves waves each.
constexpr uint32_t kRunTiles = 8;
constexpr uint32_t kRunWaves = 8;
constexpr uint32_t kRunOwners = kRunTiles * kOwnerTile;

struct Layout {
    const ReplicaDesc* desc;        // device, one per replica
    const uint32_t* chunk_replica;  // device, replica of each chunk; null for one replica
    const uint32_t* tile_replica;   // device, replica of each owner tile; null for one replica
    uint32_t replicas;
    uint32_t chunks;                // total over all replicas
    uint32_t max_replica_chunks;    // largest chunks_r (scan work per block)
    uint32_t chunk_size;            // kWave * items (a power of two)
    uint32_t tiles;                 // owner tiles over all replicas
    uint32_t tile_waves;            // waves of an owner-major launch (each walks tiles w, w + tile_waves, ...)
    const uint32_t* unit_replica;   // device, replica of each ghost-run unit; null up to kWave replicas
    uint32_t units;                 // ghost-run units over all replicas (ceil(tiles_r / kRunTiles) each)
    uint32_t streaming;             // the owned records exceed the memory-side cache: passes that
                                    // read an array once load it nontemporally (read_once, load_items)
    uint32_t ghost_runs;            // ghost writes through copy_runs_kernel (Shape::ghost_runs)
    uint32_t xcd_run;               // block renumbering (see xcd_block)
    uint32_t descending;            // this launch walks its chunks / tiles last to first
    uint32_t unit_chunks;           // one box: chunks per unit of the route-free copy
                                    // (copy_chunks_kernel); 0: owner-tile units after a route pass
    uint32_t reverse_unit_chunks;   // the same for reverse_halo (reverse_chunks_kernel), whose
                                    // units hold at most kRunOwners owners
};

// The owners a unit of reverse_chunks_kernel is sized for: its stage holds
// kRunOwners, and the owners of a unit of whole chunks vary about the mean.
constexpr uint32_t kReverseUnitOwners = 7 * kRunOwners / 8;

struct PlanBuffers {
    uint32_t* entries;
    double* bpos;          // x, y, z of each entry (3 planes of bpos_stride doubles)
    uint64_t bpos_stride;  // C * K
    uint32_t* counts;
    uint32_t* prefix;
    uint32_t* totals;
    uint32_t* owner_index;
    uint8_t* owner_code;
    uint32_t* image_start;
    uint32_t* ghost_list;
    uint32_t* owner_entry;   // owner slot -> its entry index (chunk * K + rank)
    // Single box (null otherwise): pinned host memory into which the scan also
    // stores each column total tagged with the build's stamp,
    // host_counts[col] = stamp << 32 | total, one 8-byte store per column.
    // The host takes a column once its stamp matches, so no event on the
    // stream marks the counts (an event between the scan and the route held
    // the GPU idle for 5.5 us).
    uint64_t* host_counts;
    // Process grids only (null when every replica sends to itself):
    // recv[r * kTotals + m] = first ghost of receive block m of replica r,
    // recv[R * kTotals] = overflow bits of the whole grid.
    uint32_t* recv;
    // Written by the route pass for the ghost-run kernels (null: not kept):
    // unit_rows[(first_unit_r + r + k) * kSlots + s] = the ghost-list value of
    // the first image through send slot s at or after owner k * kRunOwners of
    // replica r, and row k = the replica's unit count holds one past each
    // slot's last value, so row k + 1 minus row k counts unit k's images.
    uint32_t* unit_rows;
    uint32_t stamp;   // this build's tag in host_counts
    // One box over chunk units (null: not kept): the copy stores each
    // chunk's prefix of all columns as one row, chunk_rows[chunk * kTotals +
    // col] = prefix[col][chunk], for reverse_halo, which then reads a unit's
    // bounds from two rows instead of two columns of 28 lines each.
    uint32_t* chunk_rows;
};

// Migration between the subdomains of a process grid (grid_kernels.hip).
constexpr int kMaxRanks = 64;   // subdomains of one grid (one per lane in the rank tables)

struct MigrateParams {
    double cuts[3][kMaxRanks + 1];  // subdomain boundaries per axis (dims + 1 each)
    int32_t dims[3];
    uint32_t ranks;
    uint8_t* dest;        // per chunk slot: destination rank of the owned particle
    uint32_t* counts;     // [ranks][C]: per chunk, particles going to each rank
    uint32_t* prefix;     // [ranks][C]: exclusive scan over each source replica's chunks
    uint32_t* totals;     // [R][kMaxRanks]: particles replica r sends to each rank
    uint32_t* base;       // [ranks][kMaxRanks]: first new slot in rank q of those from r
    uint32_t* new_owned;  // [ranks] device, and the same in pinned host memory
    uint32_t* new_owned_host;
};

struct FieldParams {
    uint32_t count;
    uint32_t wide;                     // bit f: field f moves as 8-byte units (even words, 8-byte aligned)
    uint32_t words[DSDX_MAX_FIELDS];   // 32-bit words per particle
    void* data[DSDX_MAX_FIELDS];
};

// Registered fields inside communication records: the particle record
// layout (from byte kRecordBytes, back to back) or the exchange record layout
// (from byte 0, each field aligned to its element size); `stride` bytes per
// record, a multiple of 8.
struct RecordFields {
    uint32_t count;
    uint32_t stride;
    int32_t type[DSDX_MAX_FIELDS];
    int32_t op[DSDX_MAX_FIELDS];
    uint32_t components[DSDX_MAX_FIELDS];
    uint32_t words[DSDX_MAX_FIELDS];    // 32-bit words per particle
    uint32_t offset[DSDX_MAX_FIELDS];   // byte offset inside the record
    void* data[DSDX_MAX_FIELDS];
};

constexpr uint32_t kRecordBytes = 96;   // ParticleRecord: 11 doubles, 2 int32
constexpr uint64_t kMallBytes = 256ull << 20;   // MI355X memory-side (Infinity) cache

// Runs of one loopback exchange: block b copies words[b] 8-byte words.
struct BlockMap {
    uint64_t src[kSlots];
    uint64_t dst[kSlots];
    uint64_t words[kSlots];
};

struct TypedFields {
    uint32_t count;
    int32_t type[DSDX_MAX_FIELDS];
    int32_t op[DSDX_MAX_FIELDS];
    uint32_t components[DSDX_MAX_FIELDS];
    void* data[DSDX_MAX_FIELDS];
};

}  // namespace dsdx
