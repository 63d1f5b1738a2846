// device.hpp -- shared host/device definitions of the gfx950 halo kernels.
//
// Layout of a halo plan in HBM (all u32 unless noted), for a box of `owned`
// particles split into `chunks` chunks of kChunk consecutive particles:
//   entries[chunks * kChunk]   boundary particles of each chunk, compacted in
//                              owned order: (index in chunk << 8) | slab code
//   counts[kCols][chunks]      per chunk: send count of each of the 26 slots,
//                              boundary owners, images (column-major so the
//                              scan reads each column contiguously)
//   prefix[kCols][chunks]      exclusive scan of counts over chunks
//   totals[kCols + 2]          column totals, then the overflow status word
//   owner_index[owned]         owner u -> owned index (first U entries used)
//   owner_code[owned] (u8)     owner u -> slab code
//   image_start[owned]         owner u -> first entry in ghost_list
//   ghost_list[capacity]       per owner, its ghosts (index relative to
//                              `owned`) in ascending send-slot order
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
constexpr int kStatusWord = kCols; // totals[kCols]: overflow bits
constexpr uint32_t kOverflowStorage = 1u;
constexpr uint32_t kOverflowGhostList = 2u;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
// Items per lane of the classify and scatter passes; a chunk is one wave's
// contiguous share of kWave * kItems particles.
constexpr int kItems = 8;
constexpr int kChunk = kWave * kItems;

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

struct FieldParams {
    uint32_t count;
    uint32_t words[DSDX_MAX_FIELDS];   // 32-bit words per particle
    void* data[DSDX_MAX_FIELDS];
};

struct TypedFields {
    uint32_t count;
    int32_t type[DSDX_MAX_FIELDS];
    int32_t op[DSDX_MAX_FIELDS];
    uint32_t components[DSDX_MAX_FIELDS];
    void* data[DSDX_MAX_FIELDS];
};

struct PlanView {
    const uint32_t* totals;
    const uint32_t* owner_index;
    const uint8_t* owner_code;
    const uint32_t* image_start;
    const uint32_t* ghost_list;
};

}  // namespace dsdx
