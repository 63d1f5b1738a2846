// grid.cpp -- C ABI of process grids (dsdx_grid_*): the subdomains of a
// dims[0] x dims[1] x dims[2] Cartesian decomposition of one box, all on one
// GPU, with the semantics of the reference's Host path on that many MPI ranks
// (src/topology.cpp, src/halo.cpp, src/migration.cpp).  Each subdomain is a
// replica of the halo kernels whose send slots land in the neighbouring
// replica; migration moves particles between replicas.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "abi_internal.hpp"
#include "dsdx_hip.h"
#include "kernels.hpp"
#include "timing.hpp"

using namespace dsdx;
using namespace dsdx::abi;

struct dsdx_grid : Plan {
    dsdx_box box{};
    int32_t dims[3] = {1, 1, 1};
    uint32_t ranks = 1;
    uint64_t rank_capacity = 0;
    MigrateParams mp{};
    DeviceBuffer mig_dest, mig_counts, mig_prefix, mig_totals, mig_base, mig_new_owned;
    uint32_t* new_owned_host = nullptr;
    uint32_t* new_owned_host_dev = nullptr;
    hipEvent_t migrated = nullptr;   // recorded after the migration counts
    hipEvent_t counted = nullptr;    // recorded after the ghost-count exchange
    // pinned staging of the replica, chunk and tile tables: the upload does
    // not wait; the next one waits for `staged` before it rewrites the stage
    unsigned char* stage = nullptr;
    size_t stage_bytes = 0;
    hipEvent_t staged = nullptr;

    ~dsdx_grid()
    {
        if (staged) (void)hipEventSynchronize(staged);
        if (new_owned_host) (void)hipHostFree(new_owned_host);
        if (stage) (void)hipHostFree(stage);
        if (migrated) (void)hipEventDestroy(migrated);
        if (counted) (void)hipEventDestroy(counted);
        if (staged) (void)hipEventDestroy(staged);
    }
};

namespace {

hipStream_t as_stream(void* stream) { return static_cast<hipStream_t>(stream); }

int slot_of(int dx, int dy, int dz)
{
    const int raw = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
    return raw < 13 ? raw : raw - 1;
}

// Replica r's descriptor: its subdomain's slab thresholds (src/halo.cpp:199-205
// against the local domain), and per send slot the neighbour rank and periodic
// shift of src/topology.cpp:57-176 -- directions along an axis of dimension 1
// skipped (:118-127), a step off a non-periodic axis without neighbour.  The
// wrap parameters stay those of the global box (src/migration.cpp:102-128).
void rank_desc(dsdx_grid* g, uint32_t r, ReplicaDesc* d)
{
    std::memset(d, 0, sizeof(*d));
    BoxParams& bp = d->bp;
    const int32_t c[3] = {static_cast<int32_t>(r) / (g->dims[1] * g->dims[2]),
                          static_cast<int32_t>(r) / g->dims[2] % g->dims[1],
                          static_cast<int32_t>(r) % g->dims[2]};
    for (int a = 0; a < 3; ++a) {
        const double lower = g->mp.cuts[a][c[a]];
        const double upper = g->mp.cuts[a][c[a] + 1];
        bp.lo_thr[a] = lower + g->box.width;
        bp.hi_thr[a] = upper - g->box.width;
        bp.lower[a] = g->box.lower[a];
        bp.extent[a] = g->box.upper[a] - g->box.lower[a];
        bp.periodic[a] = g->box.periodic[a] ? 1 : 0;
    }
    uint32_t active = 0;
    int dir[kSlots][3] = {};
    for (int s = 0; s < kSlots; ++s) d->dest[s] = kNoPeer;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                if (!g->box.edges_and_corners && std::abs(dx) + std::abs(dy) + std::abs(dz) != 1)
                    continue;
                const int dd[3] = {dx, dy, dz};
                const int s = slot_of(dx, dy, dz);
                bool skip = false, outside = false;
                int32_t nc[3];
                double shift[3] = {0.0, 0.0, 0.0};
                for (int a = 0; a < 3; ++a) {
                    dir[s][a] = dd[a];
                    if (dd[a] != 0 && g->dims[a] == 1) skip = true;
                }
                if (skip) continue;
                for (int a = 0; a < 3; ++a) {
                    nc[a] = c[a] + dd[a];
                    if (nc[a] < 0 || nc[a] >= g->dims[a]) {
                        if (!g->box.periodic[a]) {
                            outside = true;
                            break;
                        }
                        const double extent = g->box.upper[a] - g->box.lower[a];
                        if (nc[a] < 0) {
                            nc[a] += g->dims[a];
                            shift[a] = extent;
                        } else {
                            nc[a] -= g->dims[a];
                            shift[a] = -extent;
                        }
                    }
                }
                if (outside) continue;
                for (int a = 0; a < 3; ++a) bp.shift[s][a] = shift[a];
                d->dest[s] = static_cast<uint32_t>((nc[0] * g->dims[1] + nc[1]) * g->dims[2] + nc[2]);
                active |= 1u << s;
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
        bp.lut[code] = mask;
    }
    d->self_routed = 0;
    d->offset = static_cast<uint64_t>(r) * g->rank_capacity;
    d->capacity = g->rank_capacity;
    d->ghost_list_offset = static_cast<uint64_t>(r) * g->rank_capacity;
    d->ghost_list_capacity = g->rank_capacity;
}

// The replica, chunk and tile tables to the device through the pinned stage,
// in stream order, without waiting for the copies.
int upload_tables(dsdx_grid* g, hipStream_t s)
{
    const size_t desc_bytes = g->desc.size() * sizeof(ReplicaDesc);
    const size_t chunk_bytes = g->chunk_replica.size() * 4;
    const size_t tile_bytes = g->tile_replica.size() * 4;
    const size_t need = desc_bytes + chunk_bytes + tile_bytes;
    if (!g->staged) DSDX_HIP(hipEventCreateWithFlags(&g->staged, hipEventDisableTiming), "hipEventCreate");
    else DSDX_HIP(hipEventSynchronize(g->staged), "wait for the previous table upload");
    if (need > g->stage_bytes) {
        const size_t grown = std::max(need, g->stage_bytes + g->stage_bytes / 2);
        if (g->stage) (void)hipHostFree(g->stage);
        g->stage = nullptr;
        g->stage_bytes = 0;
        void* host = nullptr;
        DSDX_HIP(hipHostMalloc(&host, grown, hipHostMallocDefault), "hipHostMalloc");
        g->stage = static_cast<unsigned char*>(host);
        g->stage_bytes = grown;
    }
    DSDX_HIP(g->d_desc.reserve(std::max<size_t>(1, desc_bytes)), "hipMalloc");
    DSDX_HIP(g->d_chunk_replica.reserve(std::max<size_t>(4, chunk_bytes)), "hipMalloc");
    DSDX_HIP(g->d_tile_replica.reserve(std::max<size_t>(4, tile_bytes)), "hipMalloc");
    std::memcpy(g->stage, g->desc.data(), desc_bytes);
    std::memcpy(g->stage + desc_bytes, g->chunk_replica.data(), chunk_bytes);
    std::memcpy(g->stage + desc_bytes + chunk_bytes, g->tile_replica.data(), tile_bytes);
    DSDX_HIP(hipMemcpyAsync(g->d_desc.ptr, g->stage, desc_bytes, hipMemcpyHostToDevice, s),
             "upload replica table");
    if (chunk_bytes)
        DSDX_HIP(hipMemcpyAsync(g->d_chunk_replica.ptr, g->stage + desc_bytes, chunk_bytes,
                                hipMemcpyHostToDevice, s),
                 "upload chunk table");
    if (tile_bytes)
        DSDX_HIP(hipMemcpyAsync(g->d_tile_replica.ptr, g->stage + desc_bytes + chunk_bytes, tile_bytes,
                                hipMemcpyHostToDevice, s),
                 "upload tile table");
    DSDX_HIP(hipEventRecord(g->staged, s), "record table upload");
    g->desc_on_device = true;
    return DSDX_OK;
}

// Owned counts changed: renumber chunks and tiles, size the buffers, upload.
int relayout(dsdx_grid* g, hipStream_t s)
{
    g->classified = g->built = g->totals_cached = false;
    number_chunks(g);
    number_tiles(g);
    if (int rc = reserve_buffers(g, static_cast<uint64_t>(g->ranks) * g->rank_capacity)) return rc;
    DSDX_HIP(g->recv.reserve((static_cast<size_t>(g->ranks) + 1) * kTotals * 4), "hipMalloc");
    return upload_tables(g, s);
}

int check_full(const dsdx_particles* p)
{
    if (!p || !p->x || !p->y || !p->z || !p->vx || !p->vy || !p->vz || !p->fx || !p->fy ||
        !p->fz || !p->mass || !p->charge || !p->type || !p->global_id)
        return fail(DSDX_EINVAL, "a process grid needs every particle array");
    return DSDX_OK;
}

}  // namespace

extern "C" {

int dsdx_grid_create(const dsdx_box* box, const int32_t* dims, uint64_t capacity_per_rank,
                     void* stream, dsdx_grid** out)
{
    if (!out) return fail(DSDX_EINVAL, "null output handle");
    *out = nullptr;
    if (!stream) return fail(DSDX_EINVAL, "grid setup requires an explicit stream");
    if (!dims) return fail(DSDX_EINVAL, "null grid dimensions");
    BoxParams check;
    if (int rc = make_box_params(box, &check)) return rc;
    int64_t ranks = 1;
    for (int a = 0; a < 3; ++a) {
        if (dims[a] < 1) return fail(DSDX_EINVAL, "grid dimensions must be positive");
        ranks *= dims[a];
    }
    if (ranks > kMaxRanks) return fail(DSDX_EINVAL, "at most 64 subdomains per grid");
    if (capacity_per_rank == 0 || capacity_per_rank > (1ull << 31) ||
        static_cast<uint64_t>(ranks) * capacity_per_rank > (1ull << 32))
        return fail(DSDX_EINVAL, "grid capacity out of range");
    auto* g = new dsdx_grid();
    g->grid = true;
    g->box = *box;
    g->ranks = static_cast<uint32_t>(ranks);
    g->rank_capacity = capacity_per_rank;
    g->mp.ranks = g->ranks;
    for (int a = 0; a < 3; ++a) {
        g->dims[a] = dims[a];
        g->mp.dims[a] = dims[a];
        // src/topology.cpp:66-72: lower + extent * i / dims
        const double extent = box->upper[a] - box->lower[a];
        for (int i = 0; i <= dims[a]; ++i)
            g->mp.cuts[a][i] = box->lower[a] + extent * static_cast<double>(i) / static_cast<double>(dims[a]);
    }
    // Chunks of 8 items per lane: the migration's count and scatter passes
    // hold every item of a chunk in registers, and at 16 items they run at
    // 4-5 waves per SIMD against 7-8 (2x2x2 grid of 16 M particles: count
    // 103 -> 82 us, scatter 683 -> 651 us per step; the halo passes are no
    // slower).
    g->shape.items = 8;
    int rc = shape_from_env(&g->shape);
    g->desc.resize(g->ranks);
    for (uint32_t r = 0; r < g->ranks; ++r) rank_desc(g, r, &g->desc[r]);
    if (!rc) rc = relayout(g, as_stream(stream));
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    return DSDX_OK;
}

int dsdx_grid_destroy(dsdx_grid* grid)
{
    delete grid;
    return DSDX_OK;
}

int dsdx_grid_set_owned(dsdx_grid* grid, const uint64_t* owned, void* stream)
{
    if (!grid || !owned) return fail(DSDX_EINVAL, "null grid or counts");
    if (!stream) return fail(DSDX_EINVAL, "grid setup requires an explicit stream");
    for (uint32_t r = 0; r < grid->ranks; ++r)
        if (owned[r] > grid->rank_capacity) return fail(DSDX_EINVAL, "owned count above the subdomain capacity");
    for (uint32_t r = 0; r < grid->ranks; ++r) grid->desc[r].owned = static_cast<uint32_t>(owned[r]);
    return relayout(grid, as_stream(stream));
}

int dsdx_grid_counts(const dsdx_grid* grid, uint64_t* owned, uint64_t* ghosts, uint64_t* owners)
{
    if (!grid) return fail(DSDX_EINVAL, "null grid");
    for (uint32_t r = 0; r < grid->ranks; ++r) {
        if (owned) owned[r] = grid->desc[r].owned;
        if (ghosts) ghosts[r] = grid->built ? grid->host_totals[r * kTotals + kColReceived] : 0;
        if (owners) owners[r] = grid->built ? grid->host_totals[r * kTotals + kColOwners] : 0;
    }
    return DSDX_OK;
}

// src/migration.cpp:391-591 on every subdomain: `from` holds the current
// particles (subdomain r in slots [r * capacity, r * capacity + owned_r)),
// `to` receives the migrated ones in the same layout, and each registered
// field travels from fields_from[f] to fields_to[f] with its particle.  Waits
// for the new counts; on DSDX_ECAPACITY nothing is written to `to`.
int dsdx_grid_migrate(dsdx_grid* grid, const dsdx_particles* from, const dsdx_particles* to,
                      const dsdx_field* fields_from, const dsdx_field* fields_to,
                      uint32_t field_count, void* stream)
{
    if (!grid) return fail(DSDX_EINVAL, "null grid");
    if (!stream) return fail(DSDX_EINVAL, "device migration requires an explicit stream");
    if (int rc = check_full(from)) return rc;
    if (int rc = check_full(to)) return rc;
    if (from->x == to->x) return fail(DSDX_EINVAL, "migration needs separate source and target storage");
    FieldParams ff, ft;
    if (int rc = make_field_params(fields_from, field_count, &ff)) return rc;
    if (int rc = make_field_params(fields_to, field_count, &ft)) return rc;
    // 8-byte moves only where both sides allow them
    ff.wide &= ft.wide;
    ft.wide = ff.wide;
    for (uint32_t f = 0; f < field_count; ++f) {
        if (ff.words[f] != ft.words[f] || fields_from[f].type != fields_to[f].type)
            return fail(DSDX_EINVAL, "source and target fields differ in type or size");
        if (ff.data[f] == ft.data[f])
            return fail(DSDX_EINVAL, "migration needs separate source and target field storage");
    }
    const hipStream_t s = as_stream(stream);
    const Layout L = grid->layout();
    const uint64_t slots = static_cast<uint64_t>(std::max<uint32_t>(1, grid->chunks)) * grid->chunk_size();
    DSDX_HIP(grid->mig_dest.reserve(slots), "hipMalloc");
    DSDX_HIP(grid->mig_counts.reserve(static_cast<size_t>(grid->ranks) * std::max<uint32_t>(1, grid->chunks) * 4),
             "hipMalloc");
    DSDX_HIP(grid->mig_prefix.reserve(static_cast<size_t>(grid->ranks) * std::max<uint32_t>(1, grid->chunks) * 4),
             "hipMalloc");
    DSDX_HIP(grid->mig_totals.reserve(static_cast<size_t>(grid->ranks) * kMaxRanks * 4), "hipMalloc");
    DSDX_HIP(grid->mig_base.reserve(static_cast<size_t>(grid->ranks) * kMaxRanks * 4), "hipMalloc");
    DSDX_HIP(grid->mig_new_owned.reserve(static_cast<size_t>(kMaxRanks) * 4), "hipMalloc");
    if (!grid->new_owned_host) {
        void* host = nullptr;
        DSDX_HIP(hipHostMalloc(&host, kMaxRanks * 4, hipHostMallocMapped | hipHostMallocCoherent),
                 "hipHostMalloc");
        grid->new_owned_host = static_cast<uint32_t*>(host);
        void* dev = nullptr;
        DSDX_HIP(hipHostGetDevicePointer(&dev, host, 0), "hipHostGetDevicePointer");
        grid->new_owned_host_dev = static_cast<uint32_t*>(dev);
    }
    MigrateParams M = grid->mp;
    M.dest = grid->mig_dest.as<uint8_t>();
    M.counts = grid->mig_counts.as<uint32_t>();
    M.prefix = grid->mig_prefix.as<uint32_t>();
    M.totals = grid->mig_totals.as<uint32_t>();
    M.base = grid->mig_base.as<uint32_t>();
    M.new_owned = grid->mig_new_owned.as<uint32_t>();
    M.new_owned_host = grid->new_owned_host_dev;
    {
        timing::Scope timed(timing::kGridMigrate, s);
        DSDX_HIP(launch_migrate_count(*from, L, M, grid->shape, s), "launch migrate count");
    }
    DSDX_HIP(launch_migrate_scan(L, M, grid->shape, s), "launch migrate scan");
    DSDX_HIP(launch_migrate_bases(M, s), "launch migrate bases");
    if (!grid->migrated)
        DSDX_HIP(hipEventCreateWithFlags(&grid->migrated, hipEventDisableTiming), "hipEventCreate");
    DSDX_HIP(hipEventRecord(grid->migrated, s), "record migration counts");
    // the scatter skips itself when a subdomain would overflow; the host
    // reads the counts while it runs
    {
        timing::Scope timed(timing::kGridScatter, s);
        DSDX_HIP(launch_migrate_scatter(*from, *to, ff, ft, L, M, grid->shape, s), "launch migrate scatter");
    }
    DSDX_HIP(hipEventSynchronize(grid->migrated), "wait for the migration counts");
    std::vector<uint32_t> owned(grid->ranks);
    for (uint32_t q = 0; q < grid->ranks; ++q) {
        owned[q] = grid->new_owned_host[q];
        if (owned[q] > grid->rank_capacity)
            return fail(DSDX_ECAPACITY, "a subdomain cannot hold its particles after migration");
    }
    for (uint32_t q = 0; q < grid->ranks; ++q) grid->desc[q].owned = owned[q];
    return relayout(grid, s);
}

// src/halo.cpp:164-238 on every subdomain and the exchange of
// src/backends/mpi_backend.cpp:344-433: subdomain q's ghosts follow its owned
// particles, receive block m holding what its neighbour in direction m sent
// through the opposite slot; registered fields are copied into the ghosts
// with their records.  Waits for the ghost counts (ghosts[r]).
int dsdx_grid_halo_build(dsdx_grid* grid, const dsdx_particles* particles, const dsdx_field* fields,
                         uint32_t field_count, void* stream, uint64_t* ghosts)
{
    if (!grid) return fail(DSDX_EINVAL, "null grid");
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (int rc = check_full(particles)) return rc;
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    const hipStream_t s = as_stream(stream);
    grid->classified = grid->built = grid->totals_cached = false;
    if (int rc = run_classify(grid, particles, false, s)) return rc;
    {
        timing::Scope timed(timing::kGridExchange, s);
        DSDX_HIP(launch_exchange_counts(grid->layout(), grid->buffers(), grid->host_totals_dev, s),
                 "launch exchange counts");
    }
    if (!grid->counted) DSDX_HIP(hipEventCreateWithFlags(&grid->counted, hipEventDisableTiming), "hipEventCreate");
    DSDX_HIP(hipEventRecord(grid->counted, s), "record ghost counts");
    // route and copy skip themselves when a subdomain overflows; the host
    // reads the counts while they run
    if (int rc = run_scatter(grid, particles, fp, s)) return rc;
    DSDX_HIP(hipEventSynchronize(grid->counted), "wait for the ghost counts");
    grid->totals_cached = true;
    for (uint32_t r = 0; r < grid->ranks; ++r)
        if (grid->host_totals[r * kTotals + kStatusWord])
            return fail(DSDX_ECAPACITY, "a subdomain cannot hold its ghosts");
    for (uint32_t r = 0; r < grid->ranks; ++r)
        if (ghosts) ghosts[r] = grid->host_totals[r * kTotals + kColReceived];
    grid->note_live_tiles();
    grid->built = true;
    return DSDX_OK;
}

// src/halo.cpp:1288-1347 on every subdomain: each owner adds the forces of its
// ghosts (held by the neighbouring subdomains) in ascending send slot.
int dsdx_grid_reverse_halo(dsdx_grid* grid, double* fx, double* fy, double* fz, void* stream)
{
    if (!grid) return fail(DSDX_EINVAL, "null grid");
    if (!grid->built) return fail(DSDX_ESTATE, "no halo plan on the grid");
    if (!stream) return fail(DSDX_EINVAL, "device reverse halo requires an explicit stream");
    if (!fx || !fy || !fz) return fail(DSDX_EINVAL, "reverse halo needs fx, fy, fz");
    const hipStream_t s = as_stream(stream);
    timing::Scope timed(timing::kReverse, s);
    DSDX_HIP(launch_reverse_forces(fx, fy, fz, grid->layout(), grid->buffers(), grid->shape, s),
             "launch reverse halo");
    return DSDX_OK;
}

// src/halo.cpp:1396-1532 on every subdomain: Sum / Min / Max / Replace of the
// ghosts' values into their owners, in ascending send slot.
int dsdx_grid_reverse_exchange(dsdx_grid* grid, const dsdx_field* fields, uint32_t field_count,
                               void* stream)
{
    if (!grid) return fail(DSDX_EINVAL, "null grid");
    if (!stream) return fail(DSDX_EINVAL, "device reverse exchange requires an explicit stream");
    TypedFields tf;
    if (int rc = make_typed_fields(fields, field_count, &tf)) return rc;
    if (!grid->built) return fail(DSDX_ESTATE, "no halo plan on the grid");
    if (field_count == 0) return DSDX_OK;
    const hipStream_t s = as_stream(stream);
    timing::Scope timed(timing::kReverseFields, s);
    DSDX_HIP(launch_reverse_fields(tf, grid->layout(), grid->buffers(), grid->shape, s),
             "launch reverse exchange");
    return DSDX_OK;
}

// src/halo.cpp:1084-1207 on every subdomain: owner values into every ghost
// that copies the owner, in the neighbouring subdomains.
int dsdx_grid_forward_exchange(dsdx_grid* grid, const dsdx_field* fields, uint32_t field_count,
                               void* stream)
{
    if (!grid) return fail(DSDX_EINVAL, "null grid");
    if (!stream) return fail(DSDX_EINVAL, "device forward exchange requires an explicit stream");
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    if (!grid->built) return fail(DSDX_ESTATE, "no halo plan on the grid");
    if (field_count == 0) return DSDX_OK;
    const hipStream_t s = as_stream(stream);
    timing::Scope timed(timing::kForward, s);
    DSDX_HIP(launch_forward_fields(fp, grid->layout(), grid->buffers(), grid->shape, s),
             "launch forward exchange");
    return DSDX_OK;
}

// Reused-plan refresh on every subdomain (launch_halo_gather /
// launch_halo_gather_positions + launch_unpack_positions,
// src/cuda/kernels.hpp:103-128): every ghost re-reads its owner along the
// last build's routes with the build's periodic shifts; positions_only
// rewrites x, y, z and zeroes the forces, otherwise the whole record and the
// registered fields.
int dsdx_grid_refresh(dsdx_grid* grid, const dsdx_particles* particles, const dsdx_field* fields,
                      uint32_t field_count, int32_t positions_only, void* stream)
{
    if (!grid) return fail(DSDX_EINVAL, "null grid");
    if (!stream) return fail(DSDX_EINVAL, "device halo requires an explicit stream");
    if (int rc = check_full(particles)) return rc;
    FieldParams fp;
    if (int rc = make_field_params(fields, field_count, &fp)) return rc;
    if (!grid->built) return fail(DSDX_ESTATE, "no halo plan on the grid");
    const hipStream_t s = as_stream(stream);
    timing::Scope timed(timing::kRefresh, s);
    DSDX_HIP(launch_refresh(*particles, fp, positions_only ? 1 : 0, grid->layout(), grid->buffers(),
                            grid->shape, s),
             "launch halo refresh");
    return DSDX_OK;
}

}  // extern "C"
