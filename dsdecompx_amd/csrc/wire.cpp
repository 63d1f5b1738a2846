// wire.cpp -- C ABI of the communication records (include/dsdx_hip.h, "Wire
// format"): layouts, the pack / unpack / reduce steps of the reference's
// device operator layer (src/cuda/kernels.hpp:40-176) and the self messages
// of a loopback box.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "abi_internal.hpp"
#include "dsdx_hip.h"
#include "kernels.hpp"
#include "timing.hpp"

using namespace dsdx;
using namespace dsdx::abi;

namespace {

hipStream_t as_stream(void* stream) { return static_cast<hipStream_t>(stream); }

constexpr uint64_t kMaxRecordOwned = (1ull << 31) - 1;

uint32_t element_bytes(int32_t type)
{
    switch (type) {
    case DSDX_INT32: case DSDX_UINT32: case DSDX_FLOAT32: return 4;
    case DSDX_INT64: case DSDX_UINT64: case DSDX_FLOAT64: return 8;
    default: return 0;
    }
}

// make_device_wire_layout (src/device_fields.cpp:35-51): fields back to back
// after the 96-byte ParticleRecord; the exchange records of
// src/halo.cpp:1237-1271 and :1422-1457: from byte 0, each field aligned to
// its element size.  Both strides round up to 8.
int make_record_fields(const dsdx_field* fields, uint32_t count, int32_t kind, bool typed,
                       RecordFields* rf)
{
    std::memset(rf, 0, sizeof(*rf));
    if (kind != DSDX_WIRE_PARTICLE && kind != DSDX_WIRE_EXCHANGE)
        return fail(DSDX_EINVAL, "unknown record layout");
    if (count > DSDX_MAX_FIELDS) return fail(DSDX_EINVAL, "at most 16 fields travel in device records");
    if (count && !fields) return fail(DSDX_EINVAL, "null field table");
    uint64_t offset = kind == DSDX_WIRE_PARTICLE ? kRecordBytes : 0;
    for (uint32_t f = 0; f < count; ++f) {
        const uint32_t element = element_bytes(fields[f].type);
        if (!fields[f].data || fields[f].components <= 0 || element == 0)
            return fail(DSDX_EINVAL, "invalid custom particle field");
        if (typed && (fields[f].op < DSDX_SUM || fields[f].op > DSDX_REPLACE))
            return fail(DSDX_EINVAL, "invalid reverse exchange operation");
        const uint64_t bytes = static_cast<uint64_t>(element) * static_cast<uint64_t>(fields[f].components);
        if (bytes % 4 != 0) return fail(DSDX_EINVAL, "field size must be a multiple of 4 bytes");
        if (kind == DSDX_WIRE_EXCHANGE) offset = (offset + element - 1) / element * element;
        rf->type[f] = fields[f].type;
        rf->op[f] = fields[f].op;
        rf->components[f] = static_cast<uint32_t>(fields[f].components);
        rf->words[f] = static_cast<uint32_t>(bytes / 4);
        rf->offset[f] = static_cast<uint32_t>(offset);
        rf->data[f] = fields[f].data;
        offset += bytes;
    }
    const uint64_t stride = (offset + 7) / 8 * 8;
    if (stride > UINT32_MAX) return fail(DSDX_EINVAL, "record too large");
    rf->stride = static_cast<uint32_t>(stride);
    rf->count = count;
    return DSDX_OK;
}

// A single-box plan with a halo built for `owned` particles.
int check_plan(const dsdx_plan* plan, uint64_t owned)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!plan->built) return fail(DSDX_ESTATE, "no device halo plan");
    if (plan->desc.size() != 1) return fail(DSDX_EINVAL, "records are packed for single-box plans");
    if (owned != plan->desc[0].owned)
        return fail(DSDX_EINVAL, "owned count differs from the last device halo");
    return DSDX_OK;
}

}  // namespace

extern "C" {

// launch_halo_classify + count readback + launch_halo_scatter
// (src/halo.cpp:464-531): the halo of `owned` particles classified, routed
// and packed as particle records into `send` (send-buffer layout), with no
// ghost written to the particle arrays.  The records' positions come from
// the classification pass.  Waits for the counts (*records = ghosts); a send
// buffer of fewer than that many records gives DSDX_ECAPACITY with the plan
// built, after which dsdx_halo_pack fills a larger buffer.
int dsdx_halo_build_records(dsdx_plan* plan, const dsdx_box* box, const dsdx_particles* particles,
                            const dsdx_field* fields, uint32_t field_count, uint64_t owned, void* send,
                            uint64_t send_capacity, void* stream, uint64_t* records)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!stream || !records) return fail(DSDX_EINVAL, "record build needs a stream and a count output");
    if (owned > kMaxRecordOwned) return fail(DSDX_EINVAL, "too many particles for one device plan");
    BoxParams bp;
    if (int rc = make_box_params(box, &bp)) return rc;
    if (int rc = check_particles(particles, owned)) return rc;
    RecordFields rf;
    if (int rc = make_record_fields(fields, field_count, DSDX_WIRE_PARTICLE, false, &rf)) return rc;
    const hipStream_t s = as_stream(stream);
    plan->classified = plan->built = plan->totals_cached = false;
    // no ghost goes to particle storage: only the ghost list bounds the route
    const uint64_t unbounded = owned + (1ull << 40);
    const uint64_t guess = std::max<uint64_t>(plan->ghost_list_capacity, owned + 1024);
    if (int rc = set_single(plan, bp, owned, unbounded, guess, s)) return rc;
    if (int rc = run_classify(plan, particles, false, s)) return rc;
    if (int rc = wait_counts(plan)) return rc;
    const uint64_t g = plan->host_totals[kColImages];
    if (plan->host_totals[kStatusWord] & kOverflowGhostList)
        if (int rc = set_single(plan, bp, owned, unbounded, g + g / 4 + 1024, s)) return rc;
    {
        timing::Scope timed(timing::kScatter, s);
        DSDX_HIP(launch_route(plan->layout(), plan->buffers(), plan->shape, s), "launch halo route");
    }
    plan->routed = true;
    plan->rows_built = false;
    plan->host_totals[kStatusWord] = 0;
    plan->built = true;
    *records = g;
    if (g > send_capacity) return fail(DSDX_ECAPACITY, "send buffer too small for the halo records");
    if (g && !send) return fail(DSDX_EINVAL, "null send buffer");
    timing::Scope timed(timing::kPack, s);
    DSDX_HIP(launch_pack_halo(*particles, rf, 0, 1, static_cast<unsigned char*>(send), plan->layout(),
                              plan->buffers(), plan->shape, s),
             "launch halo pack");
    return DSDX_OK;
}

int dsdx_wire_layout(const dsdx_field* fields, uint32_t count, int32_t kind, uint32_t* offsets,
                     uint32_t* stride)
{
    RecordFields rf;
    if (int rc = make_record_fields(fields, count, kind, false, &rf)) return rc;
    for (uint32_t f = 0; f < count; ++f)
        if (offsets) offsets[f] = rf.offset[f];
    if (stride) *stride = rf.stride;
    return DSDX_OK;
}

int dsdx_halo_pack(const dsdx_plan* plan, const dsdx_particles* particles, const dsdx_field* fields,
                   uint32_t field_count, uint64_t owned, int32_t positions_only, void* send, void* stream)
{
    if (int rc = check_plan(plan, owned)) return rc;
    if (!stream || !send) return fail(DSDX_EINVAL, "packing needs an output buffer and a stream");
    if (int rc = check_particles(particles, owned)) return rc;
    RecordFields rf;
    if (int rc = make_record_fields(positions_only ? nullptr : fields, positions_only ? 0 : field_count,
                                    DSDX_WIRE_PARTICLE, false, &rf))
        return rc;
    const hipStream_t s = as_stream(stream);
    if (int rc = ensure_routed(plan, s)) return rc;
    timing::Scope timed(timing::kPack, s);
    DSDX_HIP(launch_pack_halo(*particles, rf, positions_only ? 1 : 0, 0, static_cast<unsigned char*>(send),
                              plan->layout(), plan->buffers(), plan->shape, s),
             "launch halo pack");
    return DSDX_OK;
}

int dsdx_unpack_records(const void* records, uint64_t count, const dsdx_particles* particles,
                        uint64_t dest_offset, const dsdx_field* fields, uint32_t field_count, void* stream)
{
    if (!stream || !particles || (count && !records)) return fail(DSDX_EINVAL, "invalid record unpack");
    RecordFields rf;
    if (int rc = make_record_fields(fields, field_count, DSDX_WIRE_PARTICLE, false, &rf)) return rc;
    timing::Scope timed(timing::kUnpack, as_stream(stream));
    DSDX_HIP(launch_unpack_records(static_cast<const unsigned char*>(records), count, *particles, dest_offset,
                                   rf, as_stream(stream)),
             "launch record unpack");
    return DSDX_OK;
}

int dsdx_unpack_positions(const double* positions, uint64_t count, const dsdx_particles* particles,
                          uint64_t dest_offset, void* stream)
{
    if (!stream || !particles || (count && !positions)) return fail(DSDX_EINVAL, "invalid position unpack");
    DSDX_HIP(launch_unpack_positions(positions, count, *particles, dest_offset, as_stream(stream)),
             "launch position unpack");
    return DSDX_OK;
}

int dsdx_pack_reverse_records(const double* fx, const double* fy, const double* fz, uint64_t ghost_offset,
                              uint64_t count, void* out, void* stream)
{
    if (!stream || (count && (!fx || !fy || !fz || !out))) return fail(DSDX_EINVAL, "invalid reverse pack");
    timing::Scope timed(timing::kPackReverse, as_stream(stream));
    DSDX_HIP(launch_pack_reverse(fx, fy, fz, ghost_offset, count, static_cast<unsigned char*>(out),
                                 as_stream(stream)),
             "launch reverse pack");
    return DSDX_OK;
}

int dsdx_accumulate_reverse_records(const dsdx_plan* plan, const void* records, double* fx, double* fy,
                                    double* fz, uint64_t owned, void* stream)
{
    if (int rc = check_plan(plan, owned)) return rc;
    if (!stream || !records || !fx || !fy || !fz) return fail(DSDX_EINVAL, "invalid reverse accumulate");
    if (int rc = ensure_routed(plan, as_stream(stream))) return rc;
    timing::Scope timed(timing::kAccumulate, as_stream(stream));
    DSDX_HIP(launch_accumulate_reverse(static_cast<const unsigned char*>(records), fx, fy, fz, plan->layout(),
                                       plan->buffers(), plan->shape, as_stream(stream)),
             "launch reverse accumulate");
    return DSDX_OK;
}

int dsdx_pack_field_records(const dsdx_field* fields, uint32_t field_count, uint64_t ghost_offset,
                            uint64_t count, void* out, void* stream)
{
    if (!stream || (count && field_count && !out)) return fail(DSDX_EINVAL, "invalid field record pack");
    RecordFields rf;
    if (int rc = make_record_fields(fields, field_count, DSDX_WIRE_EXCHANGE, false, &rf)) return rc;
    DSDX_HIP(launch_pack_field_records(rf, ghost_offset, count, static_cast<unsigned char*>(out),
                                       as_stream(stream)),
             "launch field record pack");
    return DSDX_OK;
}

int dsdx_scatter_field_records(const dsdx_field* fields, uint32_t field_count, const void* records,
                               uint64_t count, uint64_t ghost_offset, void* stream)
{
    if (!stream || (count && field_count && !records)) return fail(DSDX_EINVAL, "invalid field record scatter");
    RecordFields rf;
    if (int rc = make_record_fields(fields, field_count, DSDX_WIRE_EXCHANGE, false, &rf)) return rc;
    DSDX_HIP(launch_scatter_field_records(rf, static_cast<const unsigned char*>(records), count, ghost_offset,
                                          as_stream(stream)),
             "launch field record scatter");
    return DSDX_OK;
}

int dsdx_gather_field_records(const dsdx_plan* plan, const dsdx_field* fields, uint32_t field_count,
                              uint64_t owned, void* out, void* stream)
{
    if (int rc = check_plan(plan, owned)) return rc;
    if (!stream || (field_count && !out)) return fail(DSDX_EINVAL, "invalid field record gather");
    RecordFields rf;
    if (int rc = make_record_fields(fields, field_count, DSDX_WIRE_EXCHANGE, false, &rf)) return rc;
    if (int rc = ensure_routed(plan, as_stream(stream))) return rc;
    DSDX_HIP(launch_gather_field_records(rf, static_cast<unsigned char*>(out), plan->layout(), plan->buffers(),
                                         plan->shape, as_stream(stream)),
             "launch field record gather");
    return DSDX_OK;
}

int dsdx_reduce_field_records(const dsdx_plan* plan, const dsdx_field* fields, uint32_t field_count,
                              const void* records, uint64_t owned, void* stream)
{
    if (int rc = check_plan(plan, owned)) return rc;
    if (!stream || (field_count && !records)) return fail(DSDX_EINVAL, "invalid field record reduce");
    RecordFields rf;
    if (int rc = make_record_fields(fields, field_count, DSDX_WIRE_EXCHANGE, true, &rf)) return rc;
    if (int rc = ensure_routed(plan, as_stream(stream))) return rc;
    DSDX_HIP(launch_reduce_field_records(rf, static_cast<const unsigned char*>(records), plan->layout(),
                                         plan->buffers(), plan->shape, as_stream(stream)),
             "launch field record reduce");
    return DSDX_OK;
}

// The self messages of a loopback box (src/backends/mpi_backend.cpp:366-375:
// the message of slot s arrives in receive block 25 - s).  Forward: `send`
// holds one block per send slot (slot order), `recv` receives the receive
// blocks in order, i.e. the ghosts in ghost order.  Reverse: `send` holds
// records in ghost order, `recv` gets back, per send slot, the records of the
// ghosts that slot produced.  Waits for the plan's counts if they are not on
// the host yet.
int dsdx_loopback_exchange(dsdx_plan* plan, const void* send, void* recv, uint32_t record_bytes,
                           int32_t reverse, void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!plan->built) return fail(DSDX_ESTATE, "no device halo plan");
    if (plan->desc.size() != 1) return fail(DSDX_EINVAL, "loopback exchange needs a single-box plan");
    if (!stream || record_bytes == 0 || record_bytes % 8 != 0)
        return fail(DSDX_EINVAL, "record size must be a positive multiple of 8 bytes");
    const hipStream_t s = as_stream(stream);
    if (!plan->totals_cached)
        if (int rc = read_totals(plan, s)) return rc;
    const uint32_t* t = plan->host_totals;
    uint64_t send_start[kSlots], recv_start[kSlots];
    uint64_t a = 0, b = 0;
    for (int k = 0; k < kSlots; ++k) {
        send_start[k] = a;
        recv_start[k] = b;
        a += t[k];
        b += t[kSlots - 1 - k];
    }
    if (a && (!send || !recv)) return fail(DSDX_EINVAL, "null exchange buffer");
    const uint64_t words = record_bytes / 8;
    BlockMap map;
    for (int m = 0; m < kSlots; ++m) {
        const int from = kSlots - 1 - m;    // receive block m holds send slot 25 - m
        map.words[m] = t[from] * words;
        map.src[m] = (reverse ? recv_start[m] : send_start[from]) * words;
        map.dst[m] = (reverse ? send_start[from] : recv_start[m]) * words;
    }
    timing::Scope timed(timing::kLoopback, s);
    DSDX_HIP(launch_block_copy(send, recv, map, s), "launch loopback exchange");
    return DSDX_OK;
}

}  // extern "C"
