// abi.cpp -- C ABI of the gfx950 halo kernels (include/dsdx_hip.h): argument
// validation, the device halo plan (a replica table plus growable buffers),
// launches, and the batched form.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "abi_internal.hpp"
#include "dsdx_hip.h"
#include "kernels.hpp"
#include "timing.hpp"

namespace dsdx::abi {

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

hipError_t DeviceBuffer::reserve(size_t need)
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

DeviceBuffer::~DeviceBuffer()
{
    if (ptr) (void)hipFree(ptr);
}

namespace {

int direction_slot(int dx, int dy, int dz)
{
    const int raw = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
    return raw < 13 ? raw : raw - 1;
}

size_t field_type_size(int32_t type)
{
    switch (type) {
    case DSDX_INT32: case DSDX_UINT32: case DSDX_FLOAT32: return 4;
    case DSDX_INT64: case DSDX_UINT64: case DSDX_FLOAT64: return 8;
    default: return 0;
    }
}

}  // namespace

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
    if (!(box->width >= 0.0)) return fail(DSDX_EINVAL, "halo width cannot be negative");
    if (box->local_given) {
        for (int a = 0; a < 3; ++a)
            if (!(box->local_upper[a] > box->local_lower[a]))
                return fail(DSDX_EINVAL, "local-domain extent must be positive");
    }
    std::memset(bp, 0, sizeof(*bp));
    double extent[3];
    for (int a = 0; a < 3; ++a) {
        extent[a] = box->upper[a] - box->lower[a];
        // the local domain of the 1x1x1 grid: the cuts lower + extent * i / 1
        // (src/topology.cpp:66-72) unless the caller gives them
        const double cut_lo = box->local_given ? box->local_lower[a] : box->lower[a] + extent[a] * 0.0 / 1.0;
        const double cut_hi = box->local_given ? box->local_upper[a] : box->lower[a] + extent[a] * 1.0 / 1.0;
        bp->lo_thr[a] = cut_lo + box->width;
        bp->hi_thr[a] = cut_hi - box->width;
        bp->lower[a] = box->lower[a];
        bp->extent[a] = extent[a];
        bp->periodic[a] = box->periodic[a] ? 1 : 0;
    }
    uint32_t active = 0;
    int dir[kSlots][3] = {};
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
        for (int s = 0; s < kSlots; ++s) {
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

// Self-loopback routing: every send slot of replica r lands in r itself.
void route_to_self(ReplicaDesc* d, uint32_t replica)
{
    for (int s = 0; s < kSlots; ++s) d->dest[s] = replica;
    d->self_routed = 1;
}

int make_field_params(const dsdx_field* fields, uint32_t count, FieldParams* fp)
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
        if (bytes % 8 == 0 && reinterpret_cast<uintptr_t>(fields[f].data) % 8 == 0) fp->wide |= 1u << f;
    }
    fp->count = count;
    return DSDX_OK;
}

int make_typed_fields(const dsdx_field* fields, uint32_t count, TypedFields* tf)
{
    std::memset(tf, 0, sizeof(*tf));
    if (count > DSDX_MAX_FIELDS)
        return fail(DSDX_EINVAL, "at most 16 fields can be reverse-exchanged on device memory");
    if (count && !fields) return fail(DSDX_EINVAL, "null field table");
    for (uint32_t f = 0; f < count; ++f) {
        if (!fields[f].data || fields[f].components <= 0 || field_type_size(fields[f].type) == 0 ||
            fields[f].op < DSDX_SUM || fields[f].op > DSDX_REPLACE)
            return fail(DSDX_EINVAL, "invalid reverse exchange field");
        tf->type[f] = fields[f].type;
        tf->op[f] = fields[f].op;
        tf->components[f] = static_cast<uint32_t>(fields[f].components);
        tf->data[f] = fields[f].data;
    }
    tf->count = count;
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

// Sizes every buffer for the current table and shape (ghost list: at least
// `ghost_list_elems`).
int reserve_buffers(Plan* plan, uint64_t ghost_list_elems)
{
    const uint64_t c = std::max<uint64_t>(1, plan->chunks);
    const uint64_t slots = c * plan->chunk_size();
    // entry indices (chunk * K + rank) are 32-bit
    if (slots > UINT32_MAX) return fail(DSDX_EINVAL, "too many particles for one device plan");
    const uint64_t replicas = std::max<size_t>(1, plan->desc.size());
    DSDX_HIP(plan->entries.reserve(slots * 4), "hipMalloc");
    DSDX_HIP(plan->bpos.reserve(slots * 3 * sizeof(double)), "hipMalloc");
    DSDX_HIP(plan->counts.reserve(c * kCols * 4), "hipMalloc");
    DSDX_HIP(plan->prefix.reserve(c * kCols * 4), "hipMalloc");
    DSDX_HIP(plan->totals.reserve(replicas * kTotals * 4), "hipMalloc");
    DSDX_HIP(plan->owner_index.reserve(slots * 4), "hipMalloc");
    DSDX_HIP(plan->owner_code.reserve(slots), "hipMalloc");
    DSDX_HIP(plan->image_start.reserve(slots * 4), "hipMalloc");
    DSDX_HIP(plan->owner_entry.reserve(slots * 4), "hipMalloc");
    // a replica of owned_r particles has at most ceil(owned_r / kRunOwners)
    // ghost-run units, plus its end row
    DSDX_HIP(plan->unit_rows.reserve((slots / kRunOwners + 2 * replicas) * kSlots * 4), "hipMalloc");
    DSDX_HIP(plan->chunk_rows.reserve(c * kTotals * 4), "hipMalloc");
    if (ghost_list_elems > plan->ghost_list_capacity) {
        DSDX_HIP(plan->ghost_list.reserve(ghost_list_elems * 4), "hipMalloc");
        plan->ghost_list_capacity = plan->ghost_list.bytes / 4;
    }
    const size_t words = replicas * kTotals;
    if (words > plan->host_totals_words) {
        if (plan->host_totals) (void)hipHostFree(plan->host_totals);
        plan->host_totals = nullptr;
        plan->host_totals_dev = nullptr;
        plan->host_totals_words = 0;
        void* host = nullptr;
        // mapped and coherent: the scan kernel stores the counts straight into it
        DSDX_HIP(hipHostMalloc(&host, words * 4, hipHostMallocMapped | hipHostMallocCoherent),
                 "hipHostMalloc");
        plan->host_totals = static_cast<uint32_t*>(host);
        plan->host_totals_words = words;
        void* dev = nullptr;
        DSDX_HIP(hipHostGetDevicePointer(&dev, host, 0), "hipHostGetDevicePointer");
        plan->host_totals_dev = static_cast<uint32_t*>(dev);
    }
    if (!plan->host_counts) {
        void* host = nullptr;
        // mapped and coherent: the scan stores the stamped column totals into it
        DSDX_HIP(hipHostMalloc(&host, kCols * sizeof(uint64_t), hipHostMallocMapped | hipHostMallocCoherent),
                 "hipHostMalloc");
        plan->host_counts = static_cast<uint64_t*>(host);
        std::memset(host, 0, kCols * sizeof(uint64_t));   // stamp 0 is never a build's
        void* dev = nullptr;
        DSDX_HIP(hipHostGetDevicePointer(&dev, host, 0), "hipHostGetDevicePointer");
        plan->host_counts_dev = static_cast<uint64_t*>(dev);
    }
    return DSDX_OK;
}

// Chunk numbering of the table for the current shape.
void number_chunks(Plan* plan)
{
    const uint32_t K = plan->chunk_size();
    uint32_t next = 0;
    plan->max_replica_chunks = 0;
    plan->chunk_replica.clear();
    for (size_t r = 0; r < plan->desc.size(); ++r) {
        ReplicaDesc& d = plan->desc[r];
        d.first_chunk = next;
        d.chunks = (d.owned + K - 1) / K;
        next += d.chunks;
        plan->max_replica_chunks = std::max(plan->max_replica_chunks, d.chunks);
        if (plan->desc.size() > 1)
            plan->chunk_replica.insert(plan->chunk_replica.end(), d.chunks, static_cast<uint32_t>(r));
    }
    plan->chunks = next;
}

// Owner tiles of every replica: ceil(bound_r / kOwnerTile).  A replica that
// builds without overflow has at most min(owned, ghosts) boundary owners and
// at most min(capacity - owned, ghost list capacity) ghosts.
void number_tiles(Plan* plan)
{
    uint32_t next = 0, next_unit = 0;
    plan->tile_replica.clear();
    plan->unit_replica.clear();
    for (size_t r = 0; r < plan->desc.size(); ++r) {
        ReplicaDesc& d = plan->desc[r];
        // a self-routed replica's ghosts are its own sends, so its storage
        // bounds them too; on a process grid only the ghost list does
        const uint64_t bound =
            plan->grid ? std::min<uint64_t>(d.owned, d.ghost_list_capacity)
                       : std::min<uint64_t>({static_cast<uint64_t>(d.owned), d.capacity - d.owned,
                                             d.ghost_list_capacity});
        const uint32_t tiles = static_cast<uint32_t>((bound + kOwnerTile - 1) / kOwnerTile);
        d.first_tile = next;
        next += tiles;
        if (plan->desc.size() > 1)
            plan->tile_replica.insert(plan->tile_replica.end(), tiles, static_cast<uint32_t>(r));
        const uint32_t units = (tiles + kRunTiles - 1) / kRunTiles;
        d.first_unit = next_unit;
        next_unit += units;
        if (plan->desc.size() > static_cast<size_t>(kWave))
            plan->unit_replica.insert(plan->unit_replica.end(), units, static_cast<uint32_t>(r));
    }
    plan->tiles = next;
    plan->units = next_unit;
}

// Copies the table to the device.  A single box's descriptor goes down as the
// argument of a store kernel (stream-ordered, no host wait, capturable); a
// batch's tables are copied and waited for (the host copy may change after).
int upload_table(Plan* plan, hipStream_t stream)
{
    DSDX_HIP(plan->d_desc.reserve(plan->desc.size() * sizeof(ReplicaDesc)), "hipMalloc");
    if (plan->desc.size() == 1) {
        DSDX_HIP(launch_store_desc(plan->desc[0], plan->d_desc.as<ReplicaDesc>(), stream),
                 "store replica table");
        plan->desc_on_device = true;
        return DSDX_OK;
    }
    DSDX_HIP(hipMemcpyAsync(plan->d_desc.ptr, plan->desc.data(), plan->desc.size() * sizeof(ReplicaDesc),
                            hipMemcpyHostToDevice, stream),
             "upload replica table");
    if (plan->desc.size() > 1) {
        DSDX_HIP(plan->d_chunk_replica.reserve(std::max<size_t>(1, plan->chunk_replica.size()) * 4),
                 "hipMalloc");
        if (!plan->chunk_replica.empty())
            DSDX_HIP(hipMemcpyAsync(plan->d_chunk_replica.ptr, plan->chunk_replica.data(),
                                    plan->chunk_replica.size() * 4, hipMemcpyHostToDevice, stream),
                     "upload chunk table");
        DSDX_HIP(plan->d_tile_replica.reserve(std::max<size_t>(1, plan->tile_replica.size()) * 4),
                 "hipMalloc");
        if (!plan->tile_replica.empty())
            DSDX_HIP(hipMemcpyAsync(plan->d_tile_replica.ptr, plan->tile_replica.data(),
                                    plan->tile_replica.size() * 4, hipMemcpyHostToDevice, stream),
                     "upload tile table");
        DSDX_HIP(plan->d_unit_replica.reserve(std::max<size_t>(1, plan->unit_replica.size()) * 4),
                 "hipMalloc");
        if (!plan->unit_replica.empty())
            DSDX_HIP(hipMemcpyAsync(plan->d_unit_replica.ptr, plan->unit_replica.data(),
                                    plan->unit_replica.size() * 4, hipMemcpyHostToDevice, stream),
                     "upload unit table");
    }
    DSDX_HIP(hipStreamSynchronize(stream), "wait for the replica table");
    plan->desc_on_device = true;
    return DSDX_OK;
}

// Single-box table: one replica at offset 0; uploaded only when it changed.
// With ghost_list_elems == 0 only the table is kept (no plan buffers).
// Lane items per chunk of a single box of `owned` particles: 16 (chunks of
// 1024) from 4096 such chunks up, where a wave per chunk fills the GPU many
// times over; 8 from 4096 chunks of 512; below that 4, so that a small box
// still has thousands of waves.  Measured per step: 16 M particles, 16 items
// fastest (0.476 against 0.499 ms at 8); 1 M particles (the clustered
// fields workload), 4 items fastest (0.093 against 0.096 at 8 and 0.102 ms
// at 16).
int auto_items(uint64_t owned)
{
    if (owned >= 4096ull * 1024) return 16;
    if (owned >= 4096ull * 512) return 8;
    return 4;
}

// The chunk size of a single-box plan for `owned` particles, unless fixed.
void apply_auto_items(Plan* plan, uint64_t owned)
{
    if (!plan->items_auto || plan->shape.items == auto_items(owned)) return;
    plan->shape.items = auto_items(owned);
    plan->classified = plan->built = plan->totals_cached = false;
    plan->desc_on_device = false;
    plan->live_tiles = 0;
}

int set_single(Plan* plan, const BoxParams& bp, uint64_t owned, uint64_t capacity,
               uint64_t ghost_list_elems, hipStream_t stream)
{
    apply_auto_items(plan, owned);
    plan->single_box = true;
    ReplicaDesc d;
    std::memset(&d, 0, sizeof(d));
    d.bp = bp;
    d.capacity = capacity;
    d.owned = static_cast<uint32_t>(owned);
    route_to_self(&d, 0);
    ReplicaDesc before;
    std::memset(&before, 0, sizeof(before));
    if (plan->desc.size() == 1) before = plan->desc[0];
    plan->desc.assign(1, d);
    number_chunks(plan);
    if (ghost_list_elems) {
        if (int rc = reserve_buffers(plan, ghost_list_elems)) return rc;
        plan->desc[0].ghost_list_capacity = plan->ghost_list_capacity;
        number_tiles(plan);
    }
    if (!plan->desc_on_device || std::memcmp(&before, &plan->desc[0], sizeof(ReplicaDesc)) != 0)
        return upload_table(plan, stream);
    return DSDX_OK;
}

namespace {

// Column c of the last single-box classify has arrived when its stamp does.
uint32_t arrived_columns(const Plan* plan)
{
    uint32_t c = 0;
    while (c < static_cast<uint32_t>(kCols) &&
           static_cast<uint32_t>(__atomic_load_n(plan->host_counts + c, __ATOMIC_ACQUIRE) >> 32) == plan->stamp)
        ++c;
    return c;
}

}  // namespace

// Single-box plans: the scan also stores the column totals, each tagged with
// the build's stamp, into pinned host memory, and the host waits for the
// scan only, never for the ghost records (the reference's rebuild likewise
// waits for the classification counts alone, src/halo.cpp:482-491).  The
// host spins on the stamps.  Only after 50 ms without them does it ask the
// stream, every 10 ms, whether it failed or drained without them (then they
// can no longer arrive): a stream query queues work of its own on the
// device, which held the GPU idle for 5 us after the ghost copy.  The
// overflow status is derived here exactly as the route pass derives it on
// the device.
int wait_counts(Plan* plan)
{
    if (!plan->host_counts) return fail(DSDX_ESTATE, "no halo counts to wait for");
    if (plan->counts_captured)
        return fail(DSDX_ESTATE, "the counts of a build captured into a graph cannot be waited for");
    using clock = std::chrono::steady_clock;
    const clock::time_point start = clock::now();
    clock::time_point next_query = start + std::chrono::milliseconds(50);
    for (uint32_t spin = 1; arrived_columns(plan) < static_cast<uint32_t>(kCols); ++spin) {
#if defined(__x86_64__)
        __builtin_ia32_pause();
#endif
        if (spin % 256 || clock::now() < next_query) continue;
        next_query = clock::now() + std::chrono::milliseconds(10);
        const hipError_t q = hipStreamQuery(plan->count_stream);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return hip_fail(q, "wait for halo counts");
        if (arrived_columns(plan) < static_cast<uint32_t>(kCols))
            return fail(DSDX_EDEVICE, "the halo counts of the last classify did not arrive");
    }
    uint32_t* t = plan->host_totals;
    for (int c = 0; c < kCols; ++c) t[c] = static_cast<uint32_t>(plan->host_counts[c]);
    t[kColReceived] = t[kColImages];
    const ReplicaDesc& d = plan->desc[0];
    const uint64_t ghosts = t[kColImages];
    uint32_t status = 0;
    if (d.owned + ghosts > d.capacity) status |= kOverflowStorage;
    if (ghosts > d.ghost_list_capacity) status |= kOverflowGhostList;
    t[kStatusWord] = status;
    plan->totals_cached = true;
    plan->note_live_tiles();
    return DSDX_OK;
}

int read_totals(Plan* plan, hipStream_t stream)
{
    const size_t words = plan->desc.size() * kTotals;
    DSDX_HIP(hipMemcpyAsync(plan->host_totals, plan->totals.ptr, words * 4, hipMemcpyDeviceToHost, stream),
             "copy halo counts to the host");
    DSDX_HIP(hipStreamSynchronize(stream), "wait for halo counts");
    plan->totals_cached = true;
    plan->note_live_tiles();
    return DSDX_OK;
}

int run_classify(Plan* plan, const dsdx_particles* p, bool wrap_first, hipStream_t s)
{
    // A build the host may wait for takes a fresh stamp for the scan's host
    // copy of the counts (0 is never used).  A captured build is never waited
    // for: its scan leaves the host copy alone, so a replay cannot overwrite
    // the counts of the last build the host waited for.
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    DSDX_HIP(hipStreamIsCapturing(s, &capture), "query stream capture");
    const bool waitable = capture == hipStreamCaptureStatusNone;
    if (waitable) {
        plan->stamp = plan->stamp + 1 ? plan->stamp + 1 : 1;
        plan->count_stream = s;
    }
    plan->counts_captured = !waitable;
    const Layout L = plan->layout();
    PlanBuffers B = plan->buffers();
    if (!waitable) B.host_counts = nullptr;
    {
        timing::Scope timed(wrap_first ? timing::kMigrateClassify : timing::kClassify, s);
        DSDX_HIP(launch_classify(p->x, p->y, p->z, wrap_first, L, B, plan->shape, s),
                 "launch halo classify");
    }
    {
        timing::Scope timed(timing::kScan, s);
        DSDX_HIP(launch_scan(L, B, plan->shape, s), "launch halo scan");
    }
    plan->classified = true;
    return DSDX_OK;
}

int run_scatter(Plan* plan, const dsdx_particles* p, const FieldParams& fp, hipStream_t s)
{
    const Layout L = plan->layout();
    const PlanBuffers B = plan->buffers();
    // The copy over chunk units places every ghost from the entries and
    // prefixes: the route waits until a call needs the owner maps.  A build
    // captured into a graph routes at once: each replay rebuilds without the
    // host seeing it, so only a route in the same graph keeps the owner maps
    // those of the last replayed build.
    plan->routed = L.unit_chunks == 0 || plan->counts_captured;
    plan->rows_built = L.unit_chunks != 0;
    if (plan->routed) {
        timing::Scope timed(timing::kScatter, s);
        DSDX_HIP(launch_route(L, B, plan->shape, s), "launch halo route");
    }
    {
        timing::Scope timed(timing::kCopy, s);
        DSDX_HIP(launch_copy(*p, fp, L, B, plan->shape, s), "launch halo copy");
    }
    return DSDX_OK;
}

// The owner maps and ghost list of the last build, for the calls that walk
// them; a no-op once written.  The route derives everything from the build's
// entries, prefixes and totals, which stay on the device until the next
// build, so it may run any time after the build on the same stream order.
int ensure_routed(const Plan* plan, hipStream_t s)
{
    if (plan->routed || !plan->classified) return DSDX_OK;
    Layout L = plan->layout();
    L.unit_chunks = 0;
    {
        timing::Scope timed(timing::kScatter, s);
        DSDX_HIP(launch_route(L, plan->buffers(), plan->shape, s), "launch halo route");
    }
    plan->routed = true;
    return DSDX_OK;
}

int shape_from_env(Shape* s)
{
    if (const char* v = std::getenv("DSDX_ITEMS")) s->items = std::atoi(v);
    if (const char* v = std::getenv("DSDX_WAVES")) s->waves = std::atoi(v);
    if (const char* v = std::getenv("DSDX_SCAN_THREADS")) s->scan_threads = std::atoi(v);
    if (const char* v = std::getenv("DSDX_XCD_RUN")) s->xcd_run = std::atoi(v);
    if (const char* v = std::getenv("DSDX_DESCENDING")) s->descending = std::atoi(v);
    if (const char* v = std::getenv("DSDX_GHOST_RUNS")) s->ghost_runs = std::atoi(v);
    if (const char* v = std::getenv("DSDX_REVERSE_RUNS")) s->reverse_runs = std::atoi(v);
    if (const char* v = std::getenv("DSDX_UNIT_OWNERS")) s->unit_owners = std::atoi(v);
    return shape_supported(*s) ? DSDX_OK : fail(DSDX_EINVAL, "unsupported launch shape");
}

}  // namespace dsdx::abi

using namespace dsdx;
using namespace dsdx::abi;


namespace {

hipStream_t as_stream(void* stream) { return static_cast<hipStream_t>(stream); }

constexpr uint64_t kMaxOwned = (1ull << 31) - 1;

// Single box: scatter again after an overflow with the capacity and ghost
// list grown to the known ghost count, then wait for the status.
int rescatter(Plan* plan, const dsdx_particles* p, const FieldParams& fp, uint64_t capacity,
              hipStream_t stream, uint64_t ghosts)
{
    const uint64_t owned = plan->desc[0].owned;
    if (owned + ghosts > capacity) return fail(DSDX_ECAPACITY, "device halo output capacity is too small");
    const BoxParams bp = plan->desc[0].bp;
    if (int rc = set_single(plan, bp, owned, capacity, ghosts + ghosts / 4 + 1024, stream)) return rc;
    DSDX_HIP(hipMemsetAsync(plan->totals.as<uint32_t>() + kStatusWord, 0, 4, stream), "clear halo status");
    if (int rc = run_scatter(plan, p, fp, stream)) return rc;
    if (int rc = wait_counts(plan)) return rc;
    if (plan->host_totals[kStatusWord] != 0) return fail(DSDX_EDEVICE, "halo scatter overflow after growth");
    plan->built = true;
    return DSDX_OK;
}

bool single_matches(const Plan* plan, uint64_t owned)
{
    return plan->desc.size() == 1 && owned == plan->desc[0].owned;
}

}  // namespace

static int halo_build(dsdx_plan* plan, const dsdx_box* box, const dsdx_particles* particles,
                      const dsdx_field* fields, uint32_t field_count, uint64_t owned,
                      uint64_t capacity, void* stream, uint64_t* ghosts, bool wrap_first);

extern "C" {

const char* dsdx_last_error(void) { return g_error.c_str(); }

int dsdx_plan_create(dsdx_plan** plan)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan pointer");
    *plan = new (std::nothrow) dsdx_plan();
    if (!*plan) return fail(DSDX_EDEVICE, "out of host memory");
    if (int rc = shape_from_env(&(*plan)->shape)) {
        delete *plan;
        *plan = nullptr;
        return rc;
    }
    (*plan)->items_auto = std::getenv("DSDX_ITEMS") == nullptr;
    return DSDX_OK;
}

int dsdx_plan_destroy(dsdx_plan* plan)
{
    delete plan;
    return DSDX_OK;
}

int dsdx_plan_reset(dsdx_plan* plan)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    plan->classified = false;
    plan->built = false;
    plan->totals_cached = false;
    return DSDX_OK;
}

int dsdx_plan_reserve(dsdx_plan* plan, uint64_t owned, uint64_t ghosts)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (owned > kMaxOwned) return fail(DSDX_EINVAL, "too many particles for one device plan");
    apply_auto_items(plan, owned);   // the chunk size a build of `owned` will use
    const uint32_t K = plan->chunk_size();
    const uint32_t saved = plan->chunks;
    plan->chunks = static_cast<uint32_t>((owned + K - 1) / K);
    const int rc = reserve_buffers(plan, std::max<uint64_t>(ghosts, 1));
    plan->chunks = saved;
    if (rc) return rc;
    // the single-box descriptor the builds store
    DSDX_HIP(plan->d_desc.reserve(sizeof(ReplicaDesc)), "hipMalloc");
    return DSDX_OK;
}

int dsdx_plan_set_shape(dsdx_plan* plan, int32_t items, int32_t waves, int32_t scan_threads)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    Shape s = plan->shape;   // the other choices (XCD run, order, staged kernels) stay
    s.items = items;
    s.waves = waves;
    s.scan_threads = scan_threads;
    if (!shape_supported(s)) return fail(DSDX_EINVAL, "unsupported launch shape");
    if (s.items != plan->shape.items) {
        plan->classified = plan->built = plan->totals_cached = false;
        plan->desc_on_device = false;
    }
    plan->shape = s;
    plan->items_auto = false;
    return DSDX_OK;
}

int dsdx_plan_get_shape(const dsdx_plan* plan, int32_t* items, int32_t* waves, int32_t* scan_threads)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (items) *items = plan->shape.items;
    if (waves) *waves = plan->shape.waves;
    if (scan_threads) *scan_threads = plan->shape.scan_threads;
    return DSDX_OK;
}

int dsdx_migrate_wrap(const dsdx_box* box, double* x, double* y, double* z, uint64_t owned, void* stream)
{
    BoxParams bp;
    if (int rc = make_box_params(box, &bp)) return rc;
    if (!stream) return fail(DSDX_EINVAL, "device migration requires an explicit stream");
    if (owned && (!x || !y || !z)) return fail(DSDX_EINVAL, "incomplete device particle view");
    if (owned > kMaxOwned) return fail(DSDX_EINVAL, "too many particles for one device plan");
    if (owned == 0) return DSDX_OK;
    // The box goes to the kernel as an argument: no table, no device memory,
    // no host wait.  Launch shape: the library default (or DSDX_*), chunk
    // size by the particle count as for a plan.
    Shape shape;
    if (int rc = shape_from_env(&shape)) return rc;
    if (!std::getenv("DSDX_ITEMS")) shape.items = auto_items(owned);
    WrapBox wb;
    for (int a = 0; a < 3; ++a) {
        wb.lower[a] = bp.lower[a];
        wb.extent[a] = bp.extent[a];
        wb.periodic[a] = bp.periodic[a];
    }
    wb.owned = static_cast<uint32_t>(owned);
    const hipStream_t s = as_stream(stream);
    timing::Scope timed(timing::kMigrate, s);
    DSDX_HIP(launch_wrap_box(x, y, z, wb, shape, s), "launch migrate wrap");
    return DSDX_OK;
}

int dsdx_halo_build(dsdx_plan* plan, const dsdx_box* box, const dsdx_particles* particles,
                    const dsdx_field* fields, uint32_t field_count, uint64_t owned, uint64_t capacity,
                    void* stream, uint64_t* ghosts)
{
    return halo_build(plan, box, particles, fields, field_count, owned, capacity, stream, ghosts,
                      false);
}

int dsdx_migrate_halo_build(dsdx_plan* plan, const dsdx_box* box, const dsdx_particles* particles,
                            const dsdx_field* fields, uint32_t field_count, uint64_t owned,
                            uint64_t capacity, void* stream, uint64_t* ghosts)
{
    return halo_build(plan, box, particles, fields, field_count, owned, capacity, stream, ghosts,
                      true);
}

}  // extern "C"

static int halo_build(dsdx_plan* plan, const dsdx_box* box, const dsdx_particles* particles,
                      const dsdx_field* fields, uint32_t field_count, uint64_t owned,
                      uint64_t capacity, void* stream, uint64_t* ghosts, bool wrap_first)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (owned > kMaxOwned) return fail(DSDX_EINVAL, "too many particles for one device plan");
    if (capacity < owned) return fail(DSDX_EINVAL, "particle capacity below the owned count");
    BoxParams bp;
    if (int rc = make_box_params(box, &bp)) return rc;
    if (int rc = check_particles(particles, owned)) return rc;
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    const hipStream_t s = as_stream(stream);

    plan->classified = plan->built = plan->totals_cached = false;
    // The ghost list keeps its size (dsdx_plan_reserve, or the last growth);
    // a build whose ghosts do not fit it raises the overflow status and the
    // list grows to the known count in rescatter.  A fresh plan starts from
    // the ghosts the storage could hold.
    const uint64_t guess =
        plan->ghost_list_capacity
            ? plan->ghost_list_capacity
            : std::max<uint64_t>(1024, std::min<uint64_t>(capacity - owned, owned + 1024));
    if (int rc = set_single(plan, bp, owned, capacity, guess, s)) return rc;
    if (int rc = run_classify(plan, particles, wrap_first, s)) return rc;
    if (int rc = run_scatter(plan, particles, fp, s)) return rc;
    if (!ghosts) {
        plan->built = true;
        return DSDX_OK;
    }
    if (int rc = wait_counts(plan)) return rc;
    const uint64_t g = plan->host_totals[kColImages];
    const uint32_t status = plan->host_totals[kStatusWord];
    *ghosts = g;
    if (status == 0) {
        plan->built = true;
        return DSDX_OK;
    }
    if (status & kOverflowStorage) return fail(DSDX_ECAPACITY, "device halo output capacity is too small");
    return rescatter(plan, particles, fp, capacity, s, g);
}

extern "C" {

int dsdx_halo_scatter(dsdx_plan* plan, const dsdx_particles* particles, const dsdx_field* fields,
                      uint32_t field_count, uint64_t owned, uint64_t capacity, void* stream)
{
    if (!plan || !plan->classified || plan->desc.size() != 1)
        return fail(DSDX_ESTATE, "no classified halo to scatter");
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (owned != plan->desc[0].owned) return fail(DSDX_EINVAL, "owned count differs from the classified halo");
    if (int rc = check_particles(particles, owned)) return rc;
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    const hipStream_t s = as_stream(stream);
    if (int rc = wait_counts(plan)) return rc;
    return rescatter(plan, particles, fp, capacity, s, plan->host_totals[kColImages]);
}

int dsdx_halo_refresh(dsdx_plan* plan, const dsdx_box* box, const dsdx_particles* particles,
                      const dsdx_field* fields, uint32_t field_count, uint64_t owned,
                      int32_t positions_only, void* stream)
{
    if (!plan || !plan->built) return fail(DSDX_ESTATE, "no device halo plan to refresh");
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (!single_matches(plan, owned)) return fail(DSDX_EINVAL, "owned count differs from the halo plan");
    if (int rc = check_particles(particles, owned)) return rc;
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    const hipStream_t s = as_stream(stream);
    if (box) {
        BoxParams bp;
        if (int rc = make_box_params(box, &bp)) return rc;
        if (std::memcmp(plan->desc[0].bp.shift, bp.shift, sizeof(bp.shift)) != 0) {
            std::memcpy(plan->desc[0].bp.shift, bp.shift, sizeof(bp.shift));
            if (int rc = upload_table(plan, s)) return rc;
        }
    }
    if (int rc = ensure_routed(plan, s)) return rc;
    timing::Scope timed(timing::kRefresh, s);
    DSDX_HIP(launch_refresh(*particles, fp, positions_only ? 1 : 0, plan->layout(), plan->buffers(),
                            plan->shape, s),
             "launch halo refresh");
    return DSDX_OK;
}

int dsdx_plan_query(const dsdx_plan* plan, uint32_t* send_counts, uint32_t* receive_counts,
                    uint64_t* ghosts, uint64_t* owners, int32_t* overflow, void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    uint32_t local[kTotals] = {};
    if (plan->classified) {
        // A null stream returns the counts of the last waited build without
        // touching the device.
        if (stream) {
            if (int rc = wait_counts(const_cast<dsdx_plan*>(plan))) return rc;
        } else if (!plan->totals_cached) {
            return fail(DSDX_EINVAL, "plan query without a stream needs a waited build");
        }
        std::memcpy(local, plan->host_totals, sizeof(local));
    }
    for (int s = 0; s < kSlots; ++s) {
        if (send_counts) send_counts[s] = local[s];
        if (receive_counts) receive_counts[s] = local[kSlots - 1 - s];
    }
    if (ghosts) *ghosts = local[kColImages];
    if (owners) *owners = local[kColOwners];
    if (overflow) *overflow = static_cast<int32_t>(local[kStatusWord]);
    return DSDX_OK;
}

int dsdx_plan_ghost_owners(const dsdx_plan* plan, int32_t* owners, void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!plan->built) return DSDX_OK;
    if (!stream || !owners) return fail(DSDX_EINVAL, "ghost owner query needs output and stream");
    const hipStream_t s = as_stream(stream);
    if (int rc = ensure_routed(plan, s)) return rc;
    timing::Scope timed(timing::kGhostOwners, s);
    DSDX_HIP(launch_ghost_owners(owners, plan->layout(), plan->buffers(), plan->shape, s),
             "launch ghost owners");
    return DSDX_OK;
}

int dsdx_reverse_halo(const dsdx_plan* plan, double* fx, double* fy, double* fz, uint64_t owned,
                      void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!stream) return fail(DSDX_EINVAL, "device reverse halo requires an explicit stream");
    if (!plan->built) return DSDX_OK;   // no device halo: nothing to return
    if (!single_matches(plan, owned))
        return fail(DSDX_EINVAL, "reverse exchange owned count differs from the last device halo");
    if (!fx || !fy || !fz) return fail(DSDX_EINVAL, "invalid device reverse-halo view");
    const hipStream_t s = as_stream(stream);
    Layout L = plan->layout();
    if (!plan->rows_built) L.reverse_unit_chunks = 0;
    if (!chunk_reverse(L, plan->shape))
        if (int rc = ensure_routed(plan, s)) return rc;
    timing::Scope timed(timing::kReverse, s);
    DSDX_HIP(launch_reverse_forces(fx, fy, fz, L, plan->buffers(), plan->shape, s), "launch reverse halo");
    return DSDX_OK;
}

int dsdx_reverse_exchange(const dsdx_plan* plan, const dsdx_field* fields, uint32_t field_count,
                          uint64_t owned, void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!stream) return fail(DSDX_EINVAL, "device reverse exchange requires an explicit stream");
    TypedFields tf;
    if (int rc = make_typed_fields(fields, field_count, &tf)) return rc;
    if (!plan->built || field_count == 0) return DSDX_OK;
    if (!single_matches(plan, owned))
        return fail(DSDX_EINVAL, "reverse exchange owned count differs from the last device halo");
    const hipStream_t s = as_stream(stream);
    if (int rc = ensure_routed(plan, s)) return rc;
    timing::Scope timed(timing::kReverseFields, s);
    DSDX_HIP(launch_reverse_fields(tf, plan->layout(), plan->buffers(), plan->shape, s),
             "launch reverse exchange");
    return DSDX_OK;
}

int dsdx_forward_exchange(const dsdx_plan* plan, const dsdx_field* fields, uint32_t field_count,
                          uint64_t owned, void* stream)
{
    if (!plan) return fail(DSDX_EINVAL, "null plan");
    if (!stream) return fail(DSDX_EINVAL, "device forward exchange requires an explicit stream");
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    if (!plan->built || field_count == 0) return DSDX_OK;
    if (!single_matches(plan, owned))
        return fail(DSDX_EINVAL, "forward exchange owned count differs from the last device halo");
    const hipStream_t s = as_stream(stream);
    if (int rc = ensure_routed(plan, s)) return rc;
    timing::Scope timed(timing::kForward, s);
    DSDX_HIP(launch_forward_fields(fp, plan->layout(), plan->buffers(), plan->shape, s),
             "launch forward exchange");
    return DSDX_OK;
}

// ---------------------------------------------------------------------------
// Batched form: R replicas in one set of SoA arrays, one launch per phase,
// nothing waits on the host (graph-capturable).

int dsdx_batch_create(const dsdx_replica* replicas, uint32_t count, void* stream, dsdx_batch** out)
{
    if (!out) return fail(DSDX_EINVAL, "null batch pointer");
    *out = nullptr;
    if (!replicas || count == 0) return fail(DSDX_EINVAL, "a batch needs at least one replica");
    if (!stream) return fail(DSDX_EINVAL, "batch setup requires an explicit stream");
    dsdx_batch* b = new (std::nothrow) dsdx_batch();
    if (!b) return fail(DSDX_EDEVICE, "out of host memory");
    // Over a batch of small boxes the fused migrate + classify pass walks
    // last to first (512 boxes of 2,048-262,144 particles: graph step 2.300
    // against 2.331 ms first to last); DSDX_DESCENDING overrides.
    b->shape.descending |= kOrderFused;
    int rc = shape_from_env(&b->shape);
    b->desc.resize(count);
    uint64_t ghost_list = 0;
    for (uint32_t r = 0; r < count && !rc; ++r) {
        const dsdx_replica& in = replicas[r];
        ReplicaDesc& d = b->desc[r];
        std::memset(&d, 0, sizeof(d));
        rc = make_box_params(&in.box, &d.bp);
        if (!rc && (in.owned > kMaxOwned || in.capacity < in.owned))
            rc = fail(DSDX_EINVAL, "replica capacity below its owned count");
        if (!rc && r > 0 && in.offset < replicas[r - 1].offset + replicas[r - 1].capacity)
            rc = fail(DSDX_EINVAL, "replica storage regions must be ascending and disjoint");
        d.offset = in.offset;
        d.capacity = in.capacity;
        d.owned = static_cast<uint32_t>(in.owned);
        route_to_self(&d, static_cast<uint32_t>(r));
        d.ghost_list_offset = ghost_list;
        d.ghost_list_capacity = in.capacity - in.owned;
        ghost_list += d.ghost_list_capacity;
    }
    if (!rc) {
        number_chunks(b);
        number_tiles(b);
        rc = reserve_buffers(b, std::max<uint64_t>(1, ghost_list));
    }
    if (!rc) rc = upload_table(b, as_stream(stream));
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return DSDX_OK;
}

int dsdx_batch_destroy(dsdx_batch* batch)
{
    delete batch;
    return DSDX_OK;
}

int dsdx_batch_set_shape(dsdx_batch* batch, int32_t items, int32_t waves, int32_t scan_threads,
                         void* stream)
{
    if (!batch) return fail(DSDX_EINVAL, "null batch");
    if (!stream) return fail(DSDX_EINVAL, "batch setup requires an explicit stream");
    Shape s = batch->shape;
    s.items = items;
    s.waves = waves;
    s.scan_threads = scan_threads;
    if (!shape_supported(s)) return fail(DSDX_EINVAL, "unsupported launch shape");
    const bool relayout = s.items != batch->shape.items;
    batch->shape = s;
    if (!relayout) return DSDX_OK;
    batch->classified = batch->built = batch->totals_cached = false;
    number_chunks(batch);
    if (int rc = reserve_buffers(batch, batch->ghost_list_capacity)) return rc;
    return upload_table(batch, as_stream(stream));
}

int dsdx_batch_get_shape(const dsdx_batch* batch, int32_t* items, int32_t* waves, int32_t* scan_threads)
{
    if (!batch) return fail(DSDX_EINVAL, "null batch");
    if (items) *items = batch->shape.items;
    if (waves) *waves = batch->shape.waves;
    if (scan_threads) *scan_threads = batch->shape.scan_threads;
    return DSDX_OK;
}

int dsdx_batch_migrate_wrap(const dsdx_batch* batch, double* x, double* y, double* z, void* stream)
{
    if (!batch) return fail(DSDX_EINVAL, "null batch");
    if (!stream) return fail(DSDX_EINVAL, "device migration requires an explicit stream");
    if (!x || !y || !z) return fail(DSDX_EINVAL, "incomplete device particle view");
    const hipStream_t s = as_stream(stream);
    timing::Scope timed(timing::kMigrate, s);
    DSDX_HIP(launch_migrate_wrap(x, y, z, batch->layout(), batch->shape, s), "launch migrate wrap");
    return DSDX_OK;
}

static int batch_build(dsdx_batch* batch, const dsdx_particles* particles, const dsdx_field* fields,
                       uint32_t field_count, void* stream, bool wrap_first)
{
    if (!batch) return fail(DSDX_EINVAL, "null batch");
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (int rc = check_particles(particles, 1)) return rc;
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    const hipStream_t s = as_stream(stream);
    batch->totals_cached = false;
    if (int rc = run_classify(batch, particles, wrap_first, s)) return rc;
    if (int rc = run_scatter(batch, particles, fp, s)) return rc;
    batch->built = true;
    return DSDX_OK;
}

int dsdx_batch_halo_build(dsdx_batch* batch, const dsdx_particles* particles, const dsdx_field* fields,
                          uint32_t field_count, void* stream)
{
    return batch_build(batch, particles, fields, field_count, stream, false);
}

int dsdx_batch_migrate_halo_build(dsdx_batch* batch, const dsdx_particles* particles,
                                  const dsdx_field* fields, uint32_t field_count, void* stream)
{
    return batch_build(batch, particles, fields, field_count, stream, true);
}

int dsdx_batch_refresh(dsdx_batch* batch, const dsdx_particles* particles, const dsdx_field* fields,
                       uint32_t field_count, int32_t positions_only, void* stream)
{
    if (!batch || !batch->built) return fail(DSDX_ESTATE, "no device halo plan to refresh");
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (int rc = check_particles(particles, 1)) return rc;
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    const hipStream_t s = as_stream(stream);
    timing::Scope timed(timing::kRefresh, s);
    DSDX_HIP(launch_refresh(*particles, fp, positions_only ? 1 : 0, batch->layout(), batch->buffers(),
                            batch->shape, s),
             "launch halo refresh");
    return DSDX_OK;
}

int dsdx_batch_reverse_halo(const dsdx_batch* batch, double* fx, double* fy, double* fz, void* stream)
{
    if (!batch) return fail(DSDX_EINVAL, "null batch");
    if (!stream) return fail(DSDX_EINVAL, "device reverse halo requires an explicit stream");
    if (!fx || !fy || !fz) return fail(DSDX_EINVAL, "invalid device reverse-halo view");
    if (!batch->built) return DSDX_OK;
    const hipStream_t s = as_stream(stream);
    timing::Scope timed(timing::kReverse, s);
    DSDX_HIP(launch_reverse_forces(fx, fy, fz, batch->layout(), batch->buffers(), batch->shape, s),
             "launch reverse halo");
    return DSDX_OK;
}

int dsdx_batch_reverse_exchange(const dsdx_batch* batch, const dsdx_field* fields, uint32_t field_count,
                                void* stream)
{
    if (!batch) return fail(DSDX_EINVAL, "null batch");
    if (!stream) return fail(DSDX_EINVAL, "device reverse exchange requires an explicit stream");
    TypedFields tf;
    if (int rc = make_typed_fields(fields, field_count, &tf)) return rc;
    if (!batch->built || field_count == 0) return DSDX_OK;
    const hipStream_t s = as_stream(stream);
    timing::Scope timed(timing::kReverseFields, s);
    DSDX_HIP(launch_reverse_fields(tf, batch->layout(), batch->buffers(), batch->shape, s),
             "launch reverse exchange");
    return DSDX_OK;
}

int dsdx_batch_forward_exchange(const dsdx_batch* batch, const dsdx_field* fields, uint32_t field_count,
                                void* stream)
{
    if (!batch) return fail(DSDX_EINVAL, "null batch");
    if (!stream) return fail(DSDX_EINVAL, "device forward exchange requires an explicit stream");
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    if (!batch->built || field_count == 0) return DSDX_OK;
    const hipStream_t s = as_stream(stream);
    timing::Scope timed(timing::kForward, s);
    DSDX_HIP(launch_forward_fields(fp, batch->layout(), batch->buffers(), batch->shape, s),
             "launch forward exchange");
    return DSDX_OK;
}

int dsdx_batch_ghost_owners(const dsdx_batch* batch, int32_t* owners, void* stream)
{
    if (!batch) return fail(DSDX_EINVAL, "null batch");
    if (!stream || !owners) return fail(DSDX_EINVAL, "ghost owner query needs output and stream");
    if (!batch->built) return DSDX_OK;
    const hipStream_t s = as_stream(stream);
    DSDX_HIP(launch_ghost_owners(owners, batch->layout(), batch->buffers(), batch->shape, s),
             "launch ghost owners");
    return DSDX_OK;
}

int dsdx_batch_counts(const dsdx_batch* batch, uint32_t* ghosts, uint32_t* owners, uint32_t* status,
                      void* stream)
{
    if (!batch) return fail(DSDX_EINVAL, "null batch");
    if (!stream) return fail(DSDX_EINVAL, "batch counts require an explicit stream");
    Plan* plan = const_cast<dsdx_batch*>(batch);
    if (!plan->classified) return fail(DSDX_ESTATE, "no halo has been built");
    if (int rc = read_totals(plan, as_stream(stream))) return rc;
    for (size_t r = 0; r < plan->desc.size(); ++r) {
        const uint32_t* t = plan->host_totals + r * kTotals;
        if (ghosts) ghosts[r] = t[kColImages];
        if (owners) owners[r] = t[kColOwners];
        if (status) status[r] = t[kStatusWord];
    }
    return DSDX_OK;
}

const uint32_t* dsdx_batch_device_totals(const dsdx_batch* batch)
{
    return batch ? batch->totals.as<uint32_t>() : nullptr;
}

// A step captured on a stream (thread-local capture mode: only this thread's
// allocations and waits are barred while it records) and instantiated once.
struct dsdx_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int dsdx_graph_begin(void* stream)
{
    if (!stream) return fail(DSDX_EINVAL, "graph capture needs a stream other than the null stream");
    DSDX_HIP(hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal),
             "begin graph capture");
    return DSDX_OK;
}

int dsdx_graph_end(void* stream, dsdx_graph** graph)
{
    if (!stream || !graph) return fail(DSDX_EINVAL, "null stream or graph pointer");
    *graph = nullptr;
    hipGraph_t g = nullptr;
    DSDX_HIP(hipStreamEndCapture(static_cast<hipStream_t>(stream), &g), "end graph capture");
    hipGraphExec_t exec = nullptr;
    const hipError_t st = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    if (st != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(st, "instantiate captured graph");
    }
    auto* out = new (std::nothrow) dsdx_graph();
    if (!out) {
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(g);
        return fail(DSDX_EDEVICE, "out of host memory");
    }
    out->graph = g;
    out->exec = exec;
    *graph = out;
    return DSDX_OK;
}

int dsdx_graph_launch(const dsdx_graph* graph, void* stream)
{
    if (!graph) return fail(DSDX_EINVAL, "null graph");
    DSDX_HIP(hipGraphLaunch(graph->exec, static_cast<hipStream_t>(stream)), "launch captured graph");
    return DSDX_OK;
}

int dsdx_graph_destroy(dsdx_graph* graph)
{
    if (!graph) return DSDX_OK;
    const hipError_t a = hipGraphExecDestroy(graph->exec);
    const hipError_t b = hipGraphDestroy(graph->graph);
    delete graph;
    if (a != hipSuccess) return hip_fail(a, "destroy graph executable");
    if (b != hipSuccess) return hip_fail(b, "destroy graph");
    return DSDX_OK;
}

}  // extern "C"
