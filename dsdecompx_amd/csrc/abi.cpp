// abi.cpp -- C ABI of the gfx950 halo kernels (include/dsdx_hip.h): argument
// validation, the device halo plan and its growable buffers, launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>

#include "dsdx_hip.h"
#include "kernels.hpp"
#include "timing.hpp"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& message)
{
    g_error = "dsDecompX: " + message;
    return code;
}

int hip_fail(hipError_t status, const char* operation)
{
    g_error = std::string("dsDecompX HIP: ") + operation + ": " + hipGetErrorString(status);
    return DSDX_EDEVICE;
}

#define DSDX_HIP(call, what)                         \
    do {                                             \
        const hipError_t status_ = (call);           \
        if (status_ != hipSuccess) return hip_fail(status_, what); \
    } while (0)

// Device allocation that grows by 1.5x (at least 4 KiB) and keeps nothing on
// growth (src/cuda/cuda_buffer.cpp:10-54 keeps the contents; no caller here
// needs them).  Waits for the device before freeing the old block, which may
// still be read by queued kernels.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;

    hipError_t reserve(size_t need)
    {
        if (need <= bytes) return hipSuccess;
        const size_t grown = std::max<size_t>({need, bytes + bytes / 2, 4096});
        if (ptr) {
            const hipError_t s = hipDeviceSynchronize();
            if (s != hipSuccess) return s;
            (void)hipFree(ptr);
            ptr = nullptr;
            bytes = 0;
        }
        const hipError_t s = hipMalloc(&ptr, grown);
        if (s == hipSuccess) bytes = grown;
        return s;
    }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
    ~DeviceBuffer() { if (ptr) (void)hipFree(ptr); }
};

}  // namespace

struct dsdx_plan {
    dsdx::BoxParams bp{};
    bool classified = false;   // classify + scan of `owned` particles are on the device
    bool built = false;        // ghosts and owner map written: reverse/forward usable
    uint64_t owned = 0;
    uint32_t chunks = 0;
    DeviceBuffer entries, counts, prefix, totals, owner_index, owner_code, image_start, ghost_list;
    uint64_t ghost_list_capacity = 0;   // elements
    uint32_t* host_totals = nullptr;    // pinned copy of totals
    bool totals_cached = false;         // host_totals hold the current plan's counts

    dsdx::PlanView view() const
    {
        return {totals.as<uint32_t>(), owner_index.as<uint32_t>(), owner_code.as<uint8_t>(),
                image_start.as<uint32_t>(), ghost_list.as<uint32_t>()};
    }
};

namespace {

using dsdx::BoxParams;

int direction_slot(int dx, int dy, int dz)
{
    const int raw = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
    return raw < 13 ? raw : raw - 1;
}

// Neighbour table of the 1x1x1 grid (src/topology.cpp:85-166 with every
// direction a self-neighbour): active send slots, their periodic shifts, and
// the slab-code -> slot-set table of the Host membership rule
// (src/halo.cpp:212-221).
int make_box_params(const dsdx_box* box, BoxParams* bp)
{
    if (!box) return fail(DSDX_EINVAL, "null box");
    for (int a = 0; a < 3; ++a) {
        if (!(box->upper[a] > box->lower[a]))
            return fail(DSDX_EINVAL, "global-domain extent must be positive");
    }
    if (box->width < 0.0) return fail(DSDX_EINVAL, "halo width cannot be negative");
    std::memset(bp, 0, sizeof(*bp));
    double extent[3];
    for (int a = 0; a < 3; ++a) {
        extent[a] = box->upper[a] - box->lower[a];
        bp->lo_thr[a] = box->lower[a] + box->width;
        bp->hi_thr[a] = box->upper[a] - box->width;
        bp->lower[a] = box->lower[a];
        bp->extent[a] = extent[a];
        bp->periodic[a] = box->periodic[a] ? 1 : 0;
    }
    uint32_t active = 0;
    int dir[dsdx::kSlots][3] = {};
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                if (!box->edges_and_corners && std::abs(dx) + std::abs(dy) + std::abs(dz) != 1)
                    continue;
                const int d[3] = {dx, dy, dz};
                const int s = direction_slot(dx, dy, dz);
                bool inside = true;
                for (int a = 0; a < 3; ++a) {
                    dir[s][a] = d[a];
                    if (d[a] != 0 && !box->periodic[a]) inside = false;
                    bp->shift[s][a] = d[a] == 0 ? 0.0 : (d[a] < 0 ? extent[a] : -extent[a]);
                }
                if (inside) active |= 1u << s;
            }
    for (uint32_t code = 0; code < 64; ++code) {
        uint32_t mask = 0;
        for (int s = 0; s < dsdx::kSlots; ++s) {
            if (!((active >> s) & 1u)) continue;
            bool covered = true;
            for (int a = 0; a < 3; ++a) {
                if (dir[s][a] < 0 && !((code >> (2 * a)) & 1u)) covered = false;
                if (dir[s][a] > 0 && !((code >> (2 * a + 1)) & 1u)) covered = false;
            }
            if (covered) mask |= 1u << s;
        }
        bp->lut[code] = mask;
    }
    return DSDX_OK;
}

size_t field_type_size(int32_t type)
{
    switch (type) {
    case DSDX_INT32: case DSDX_UINT32: case DSDX_FLOAT32: return 4;
    case DSDX_INT64: case DSDX_UINT64: case DSDX_FLOAT64: return 8;
    default: return 0;
    }
}

int make_field_params(const dsdx_field* fields, uint32_t count, dsdx::FieldParams* fp)
{
    std::memset(fp, 0, sizeof(*fp));
    if (count > DSDX_MAX_FIELDS)
        return fail(DSDX_EINVAL, "at most 16 fields can be exchanged on device memory");
    if (count && !fields) return fail(DSDX_EINVAL, "null field table");
    for (uint32_t f = 0; f < count; ++f) {
        const size_t elem = field_type_size(fields[f].type);
        if (!fields[f].data || fields[f].components <= 0 || elem == 0)
            return fail(DSDX_EINVAL, "invalid custom particle field");
        const size_t bytes = elem * static_cast<size_t>(fields[f].components);
        if (bytes % 4 != 0) return fail(DSDX_EINVAL, "field size must be a multiple of 4 bytes");
        fp->words[f] = static_cast<uint32_t>(bytes / 4);
        fp->data[f] = fields[f].data;
    }
    fp->count = count;
    return DSDX_OK;
}

int check_particles(const dsdx_particles* p, uint64_t owned)
{
    if (!p) return fail(DSDX_EINVAL, "null particle view");
    if (owned && (!p->x || !p->y || !p->z || !p->vx || !p->vy || !p->vz || !p->mass ||
                  !p->charge || !p->type || !p->global_id))
        return fail(DSDX_EINVAL, "incomplete device particle view");
    return DSDX_OK;
}

hipStream_t as_stream(void* stream) { return static_cast<hipStream_t>(stream); }

constexpr uint64_t kMaxOwned = (1ull << 31) - 1;

int reserve_plan(dsdx_plan* plan, uint64_t owned, uint64_t ghost_list_elems)
{
    const uint64_t chunks = (owned + dsdx::kChunk - 1) / dsdx::kChunk;
    DSDX_HIP(plan->entries.reserve(std::max<uint64_t>(1, chunks) * dsdx::kChunk * 4), "hipMalloc");
    DSDX_HIP(plan->counts.reserve(std::max<uint64_t>(1, chunks) * dsdx::kCols * 4), "hipMalloc");
    DSDX_HIP(plan->prefix.reserve(std::max<uint64_t>(1, chunks) * dsdx::kCols * 4), "hipMalloc");
    DSDX_HIP(plan->totals.reserve((dsdx::kCols + 2) * 4), "hipMalloc");
    DSDX_HIP(plan->owner_index.reserve(std::max<uint64_t>(1, owned) * 4), "hipMalloc");
    DSDX_HIP(plan->owner_code.reserve(std::max<uint64_t>(1, owned)), "hipMalloc");
    DSDX_HIP(plan->image_start.reserve(std::max<uint64_t>(1, owned) * 4), "hipMalloc");
    if (ghost_list_elems > plan->ghost_list_capacity) {
        DSDX_HIP(plan->ghost_list.reserve(ghost_list_elems * 4), "hipMalloc");
        plan->ghost_list_capacity = plan->ghost_list.bytes / 4;
    }
    if (!plan->host_totals) {
        void* host = nullptr;
        DSDX_HIP(hipHostMalloc(&host, (dsdx::kCols + 2) * 4, hipHostMallocDefault), "hipHostMalloc");
        plan->host_totals = static_cast<uint32_t*>(host);
    }
    return DSDX_OK;
}

int read_totals(dsdx_plan* plan, hipStream_t stream)
{
    DSDX_HIP(hipMemcpyAsync(plan->host_totals, plan->totals.ptr, (dsdx::kCols + 1) * 4,
                            hipMemcpyDeviceToHost, stream),
             "copy halo counts to the host");
    DSDX_HIP(hipStreamSynchronize(stream), "wait for halo counts");
    plan->totals_cached = true;
    return DSDX_OK;
}

int launch_scatter(dsdx_plan* plan, const dsdx_particles* p, const dsdx::FieldParams& fp,
                   uint64_t capacity, hipStream_t stream)
{
    dsdx::timing::Scope timed(dsdx::timing::kScatter, stream);
    DSDX_HIP(dsdx::launch_scatter(*p, static_cast<uint32_t>(plan->owned), plan->bp, fp,
                                  plan->entries.as<uint32_t>(), plan->counts.as<uint32_t>(),
                                  plan->prefix.as<uint32_t>(), plan->totals.as<uint32_t>(),
                                  plan->owner_index.as<uint32_t>(), plan->owner_code.as<uint8_t>(),
                                  plan->image_start.as<uint32_t>(), plan->ghost_list.as<uint32_t>(),
                                  plan->chunks, capacity, plan->ghost_list_capacity, stream),
             "launch halo scatter");
    return DSDX_OK;
}

// Scatter again after an overflow: clear the status word, grow the owner ->
// ghost map to the known ghost count, launch, and wait for the status.
int rescatter(dsdx_plan* plan, const dsdx_particles* p, const dsdx::FieldParams& fp,
              uint64_t capacity, hipStream_t stream, uint64_t ghosts)
{
    if (plan->owned + ghosts > capacity)
        return fail(DSDX_ECAPACITY, "device halo output capacity is too small");
    if (int rc = reserve_plan(plan, plan->owned, ghosts + ghosts / 4 + 1024)) return rc;
    DSDX_HIP(hipMemsetAsync(plan->totals.as<uint32_t>() + dsdx::kStatusWord, 0, 4, stream),
             "clear halo status");
    if (int rc = launch_scatter(plan, p, fp, capacity, stream)) return rc;
    if (int rc = read_totals(plan, stream)) return rc;
    if (plan->host_totals[dsdx::kStatusWord] != 0)
        return fail(DSDX_EDEVICE, "halo scatter overflow after growth");
    plan->built = true;
    return DSDX_OK;
}

}  // namespace

extern "C" {

const char* dsdx_last_error(void) { return g_error.c_str(); }

int dsdx_plan_create(dsdx_plan** plan)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan pointer");
    *plan = new (std::nothrow) dsdx_plan();
    return *plan ? DSDX_OK : fail(DSDX_EDEVICE, "out of host memory");
}

int dsdx_plan_destroy(dsdx_plan* plan)
{
    if (!plan) return DSDX_OK;
    if (plan->host_totals) (void)hipHostFree(plan->host_totals);
    delete plan;
    return DSDX_OK;
}

int dsdx_plan_reset(dsdx_plan* plan)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    plan->classified = false;
    plan->built = false;
    plan->totals_cached = false;
    plan->owned = 0;
    plan->chunks = 0;
    return DSDX_OK;
}

int dsdx_plan_reserve(dsdx_plan* plan, uint64_t owned, uint64_t ghosts)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (owned > kMaxOwned) return fail(DSDX_EINVAL, "too many particles for one device plan");
    return reserve_plan(plan, owned, std::max<uint64_t>(ghosts, 1));
}

int dsdx_migrate_wrap(const dsdx_box* box, double* x, double* y, double* z, uint64_t owned,
                      void* stream)
{
    BoxParams bp;
    if (int rc = make_box_params(box, &bp)) return rc;
    if (!stream) return fail(DSDX_EINVAL, "device migration requires an explicit stream");
    if (owned && (!x || !y || !z)) return fail(DSDX_EINVAL, "incomplete device particle view");
    dsdx::timing::Scope timed(dsdx::timing::kMigrate, as_stream(stream));
    DSDX_HIP(dsdx::launch_migrate_wrap(x, y, z, owned, bp, as_stream(stream)), "launch migrate wrap");
    return DSDX_OK;
}

int dsdx_halo_build(dsdx_plan* plan, const dsdx_box* box, const dsdx_particles* particles,
                    const dsdx_field* fields, uint32_t field_count, uint64_t owned,
                    uint64_t capacity, void* stream, uint64_t* ghosts)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (owned > kMaxOwned) return fail(DSDX_EINVAL, "too many particles for one device plan");
    if (capacity < owned) return fail(DSDX_EINVAL, "particle capacity below the owned count");
    BoxParams bp;
    if (int rc = make_box_params(box, &bp)) return rc;
    if (int rc = check_particles(particles, owned)) return rc;
    dsdx::FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    const hipStream_t s = as_stream(stream);

    plan->classified = plan->built = plan->totals_cached = false;
    const uint64_t guess = std::max<uint64_t>(plan->ghost_list_capacity, owned + 1024);
    if (int rc = reserve_plan(plan, owned, guess)) return rc;
    plan->bp = bp;
    plan->owned = owned;
    plan->chunks = static_cast<uint32_t>((owned + dsdx::kChunk - 1) / dsdx::kChunk);

    {
    dsdx::timing::Scope timed(dsdx::timing::kClassify, s);
    DSDX_HIP(dsdx::launch_classify(particles->x, particles->y, particles->z,
                                   static_cast<uint32_t>(owned), bp, plan->entries.as<uint32_t>(),
                                   plan->counts.as<uint32_t>(), plan->chunks, s),
             "launch halo classify");
    }
    {
    dsdx::timing::Scope timed(dsdx::timing::kScan, s);
    DSDX_HIP(dsdx::launch_scan(plan->counts.as<uint32_t>(), plan->prefix.as<uint32_t>(),
                               plan->totals.as<uint32_t>(), plan->chunks, s),
             "launch halo scan");
    }
    plan->classified = true;
    if (int rc = launch_scatter(plan, particles, fp, capacity, s)) return rc;
    if (!ghosts) {
        plan->built = true;
        return DSDX_OK;
    }
    if (int rc = read_totals(plan, s)) return rc;
    const uint64_t g = plan->host_totals[dsdx::kColImages];
    const uint32_t status = plan->host_totals[dsdx::kStatusWord];
    *ghosts = g;
    if (status == 0) {
        plan->built = true;
        return DSDX_OK;
    }
    if (status & dsdx::kOverflowStorage)
        return fail(DSDX_ECAPACITY, "device halo output capacity is too small");
    return rescatter(plan, particles, fp, capacity, s, g);
}

int dsdx_halo_scatter(dsdx_plan* plan, const dsdx_particles* particles, const dsdx_field* fields,
                      uint32_t field_count, uint64_t owned, uint64_t capacity, void* stream)
{
    if (!plan || !plan->classified) return fail(DSDX_ESTATE, "no classified halo to scatter");
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (owned != plan->owned) return fail(DSDX_EINVAL, "owned count differs from the classified halo");
    if (int rc = check_particles(particles, owned)) return rc;
    dsdx::FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    const hipStream_t s = as_stream(stream);
    if (int rc = read_totals(plan, s)) return rc;
    return rescatter(plan, particles, fp, capacity, s, plan->host_totals[dsdx::kColImages]);
}

int dsdx_halo_refresh(dsdx_plan* plan, const dsdx_box* box, const dsdx_particles* particles,
                      const dsdx_field* fields, uint32_t field_count, uint64_t owned,
                      int32_t positions_only, void* stream)
{
    if (!plan || !plan->built) return fail(DSDX_ESTATE, "no device halo plan to refresh");
    if (box) {
        BoxParams bp;
        if (int rc = make_box_params(box, &bp)) return rc;
        std::memcpy(plan->bp.shift, bp.shift, sizeof(bp.shift));
    }
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (owned != plan->owned) return fail(DSDX_EINVAL, "owned count differs from the halo plan");
    if (int rc = check_particles(particles, owned)) return rc;
    dsdx::FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    dsdx::timing::Scope timed(dsdx::timing::kRefresh, as_stream(stream));
    DSDX_HIP(dsdx::launch_refresh(*particles, fp, static_cast<uint32_t>(owned), positions_only ? 1 : 0,
                                  plan->view(), plan->bp, as_stream(stream)),
             "launch halo refresh");
    return DSDX_OK;
}

int dsdx_plan_query(const dsdx_plan* plan, uint32_t* send_counts, uint32_t* receive_counts,
                    uint64_t* ghosts, uint64_t* owners, int32_t* overflow, void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    uint32_t local[dsdx::kCols + 1] = {};
    if (plan->classified) {
        // A null stream returns the counts of the last waited build without
        // touching the device.
        if (stream) {
            if (int rc = read_totals(const_cast<dsdx_plan*>(plan), as_stream(stream))) return rc;
        }
        if (!stream && !plan->totals_cached)
            return fail(DSDX_EINVAL, "plan query without a stream needs a waited build");
        std::memcpy(local, plan->host_totals, sizeof(local));
    }
    for (int s = 0; s < dsdx::kSlots; ++s) {
        if (send_counts) send_counts[s] = local[s];
        if (receive_counts) receive_counts[s] = local[dsdx::kSlots - 1 - s];
    }
    if (ghosts) *ghosts = local[dsdx::kColImages];
    if (owners) *owners = local[dsdx::kColOwners];
    if (overflow) *overflow = static_cast<int32_t>(local[dsdx::kStatusWord]);
    return DSDX_OK;
}

int dsdx_plan_ghost_owners(const dsdx_plan* plan, int32_t* owners, void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!plan->built) return DSDX_OK;
    if (!stream || !owners) return fail(DSDX_EINVAL, "ghost owner query needs output and stream");
    dsdx::timing::Scope timed(dsdx::timing::kGhostOwners, as_stream(stream));
    DSDX_HIP(dsdx::launch_ghost_owners(owners, static_cast<uint32_t>(plan->owned), plan->view(),
                                       plan->bp, as_stream(stream)),
             "launch ghost owners");
    return DSDX_OK;
}

int dsdx_reverse_halo(const dsdx_plan* plan, double* fx, double* fy, double* fz, uint64_t owned,
                      void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!stream) return fail(DSDX_EINVAL, "device reverse halo requires an explicit stream");
    if (!plan->built) return DSDX_OK;   // no device halo: nothing to return
    if (owned != plan->owned)
        return fail(DSDX_EINVAL, "reverse exchange owned count differs from the last device halo");
    if (!fx || !fy || !fz) return fail(DSDX_EINVAL, "invalid device reverse-halo view");
    dsdx::timing::Scope timed(dsdx::timing::kReverse, as_stream(stream));
    DSDX_HIP(dsdx::launch_reverse_forces(fx, fy, fz, static_cast<uint32_t>(owned), plan->view(),
                                         plan->bp, as_stream(stream)),
             "launch reverse halo");
    return DSDX_OK;
}

int dsdx_reverse_exchange(const dsdx_plan* plan, const dsdx_field* fields, uint32_t field_count,
                          uint64_t owned, void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!stream) return fail(DSDX_EINVAL, "device reverse exchange requires an explicit stream");
    if (field_count > DSDX_MAX_FIELDS)
        return fail(DSDX_EINVAL, "at most 16 fields can be reverse-exchanged on device memory");
    dsdx::TypedFields tf;
    std::memset(&tf, 0, sizeof(tf));
    for (uint32_t f = 0; f < field_count; ++f) {
        if (!fields[f].data || fields[f].components <= 0 || field_type_size(fields[f].type) == 0 ||
            fields[f].op < DSDX_SUM || fields[f].op > DSDX_REPLACE)
            return fail(DSDX_EINVAL, "invalid reverse exchange field");
        tf.type[f] = fields[f].type;
        tf.op[f] = fields[f].op;
        tf.components[f] = static_cast<uint32_t>(fields[f].components);
        tf.data[f] = fields[f].data;
    }
    tf.count = field_count;
    if (!plan->built || field_count == 0) return DSDX_OK;
    if (owned != plan->owned)
        return fail(DSDX_EINVAL, "reverse exchange owned count differs from the last device halo");
    dsdx::timing::Scope timed(dsdx::timing::kReverseFields, as_stream(stream));
    DSDX_HIP(dsdx::launch_reverse_fields(tf, static_cast<uint32_t>(owned), plan->view(), plan->bp,
                                         as_stream(stream)),
             "launch reverse exchange");
    return DSDX_OK;
}

int dsdx_forward_exchange(const dsdx_plan* plan, const dsdx_field* fields, uint32_t field_count,
                          uint64_t owned, void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!stream) return fail(DSDX_EINVAL, "device forward exchange requires an explicit stream");
    dsdx::FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    if (!plan->built || field_count == 0) return DSDX_OK;
    if (owned != plan->owned)
        return fail(DSDX_EINVAL, "forward exchange owned count differs from the last device halo");
    dsdx::timing::Scope timed(dsdx::timing::kForward, as_stream(stream));
    DSDX_HIP(dsdx::launch_forward_fields(fp, static_cast<uint32_t>(owned), plan->view(), plan->bp,
                                         as_stream(stream)),
             "launch forward exchange");
    return DSDX_OK;
}

}  // extern "C"
