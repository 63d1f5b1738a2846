// grid_kernels.hip -- gfx950 kernels of process grids: several subdomains
// ("ranks") of one periodic box held on one GPU, each a replica of the halo
// kernels (halo_kernels.hip) whose send slots land in the neighbouring
// replica.  This file adds what crosses subdomains: the receive-block table of
// every replica and migration between replicas (src/migration.cpp:391-591).
#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"
#include "wave.hpp"

namespace dsdx {
using namespace kern;
namespace {

// One wave per replica q: receive block m of q holds what its neighbour in
// direction m sends through the opposite slot 25 - m
// (src/backends/mpi_backend.cpp:391-433), blocks in slot order
// (src/runtime.cpp:946-957).  Writes the block bases, the replica's ghost
// count, and ORs the overflow bits of every replica into the grid status;
// the ghost count, the distinct boundary owners and the replica's overflow
// bits also go to `host` (mapped host memory, kTotals words per replica).
__global__ __launch_bounds__(kWave) void exchange_counts_kernel(const ReplicaDesc* __restrict__ desc,
                                                                Layout L, PlanBuffers B,
                                                                uint32_t* __restrict__ host)
{
    const uint32_t q = blockIdx.x;
    if (q >= L.replicas) return;
    const ReplicaDesc* d = desc + q;
    const uint32_t lane = lane_id();
    const uint32_t peer = lane < kSlots ? d->dest[lane] : kNoPeer;
    const uint32_t count =
        peer != kNoPeer ? B.totals[static_cast<uint64_t>(peer) * kTotals + kSlots - 1 - lane] : 0u;
    const uint32_t incl = wave_inclusive_sum(count);
    if (lane < kSlots) B.recv[static_cast<uint64_t>(q) * kTotals + lane] = incl - count;
    const uint32_t received = __builtin_amdgcn_readlane(incl, kWave - 1);
    if (lane == 0) {
        uint32_t* tot = B.totals + static_cast<uint64_t>(q) * kTotals;
        tot[kColReceived] = received;
        uint32_t overflow = 0;
        if (static_cast<uint64_t>(d->owned) + received > d->capacity) overflow |= kOverflowStorage;
        if (tot[kColImages] > d->ghost_list_capacity) overflow |= kOverflowGhostList;
        if (overflow) atomicOr(B.recv + static_cast<uint64_t>(L.replicas) * kTotals, overflow);
        host[static_cast<uint64_t>(q) * kTotals + kColReceived] = received;
        host[static_cast<uint64_t>(q) * kTotals + kColOwners] = tot[kColOwners];
        host[static_cast<uint64_t>(q) * kTotals + kStatusWord] = overflow;
    }
}

// owner_rank (src/migration.cpp:74-98) of a position whose periodic
// coordinates owner_rank has wrapped itself: the cut interval by
// upper_bound - 1, clamped; ranks row-major.
__device__ __forceinline__ uint32_t owner_rank(const double (&v)[3], const MigrateParams& M)
{
    int32_t c[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int32_t k = 0;
        for (int32_t i = 0; i <= M.dims[a]; ++i) k += (v[a] < M.cuts[a][i]) ? 0 : 1;
        c[a] = min(max(k - 1, 0), M.dims[a] - 1);
    }
    return static_cast<uint32_t>((c[0] * M.dims[1] + c[1]) * M.dims[2] + c[2]);
}

// The wrapped position of item k (wrapped_record, src/migration.cpp:102-128:
// periodic axes only).
template <int ITEMS>
__device__ __forceinline__ void wrapped(const ReplicaDesc* d, const double (&x)[ITEMS],
                                        const double (&y)[ITEMS], const double (&z)[ITEMS], int k,
                                        double (&v)[3])
{
    v[0] = d->bp.periodic[0] ? wrap(x[k], d->bp.lower[0], d->bp.extent[0]) : x[k];
    v[1] = d->bp.periodic[1] ? wrap(y[k], d->bp.lower[1], d->bp.extent[1]) : y[k];
    v[2] = d->bp.periodic[2] ? wrap(z[k], d->bp.lower[2], d->bp.extent[2]) : z[k];
}

// Pass 1 of migration: one wave per chunk of a source replica; destination
// rank of every owned particle and, per chunk, the count for each rank
// (LDS atomics in the wave's own row).
template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void migrate_count_kernel(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const ReplicaDesc* __restrict__ desc, const uint32_t* __restrict__ chunk_replica, Layout L,
    MigrateParams M)
{
    constexpr uint32_t K = ITEMS * kWave;
    __shared__ uint32_t s_cnt[WAVES][kMaxRanks];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t chunk = wave_chunk<WAVES>(L, L.chunks);
    if (chunk >= L.chunks) return;
    const ReplicaDesc* d = desc + replica_of(chunk_replica, chunk);
    s_cnt[wave][lane] = 0u;
    const uint32_t owned = d->owned;
    const uint64_t off = d->offset;
    const uint32_t first = (chunk - d->first_chunk) * K;
    const bool full = first + K <= owned;
    double px[ITEMS], py[ITEMS], pz[ITEMS];
    load_items<ITEMS>(x + off, first, owned, full, px);
    load_items<ITEMS>(y + off, first, owned, full, py);
    load_items<ITEMS>(z + off, first, owned, full, pz);
    const uint64_t slot0 = static_cast<uint64_t>(chunk) * K;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t i = first + k * kWave + lane;
        if (i >= owned) continue;
        // migration asks owner_rank for the wrapped record's position, and
        // owner_rank wraps it again (src/migration.cpp:83-87, :440): a record
        // that wrapped to exactly the upper face belongs to the lower one
        double v[3];
        wrapped<ITEMS>(d, px, py, pz, k, v);
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (d->bp.periodic[a]) v[a] = wrap(v[a], d->bp.lower[a], d->bp.extent[a]);
        const uint32_t q = owner_rank(v, M);
        M.dest[slot0 + k * kWave + lane] = static_cast<uint8_t>(q);
        atomicAdd(&s_cnt[wave][q], 1u);
    }
    // LDS operations of one wave complete in order: the atomics above are
    // visible to the reads below without a barrier.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < M.ranks) M.counts[static_cast<uint64_t>(lane) * L.chunks + chunk] = s_cnt[wave][lane];
}

// Pass 2: exclusive scan of column q (particles going to rank q) over source
// replica r's chunks; grid (ranks, replicas).
__global__ __launch_bounds__(kMaxScanThreads) void migrate_scan_kernel(
    const ReplicaDesc* __restrict__ desc, Layout L, MigrateParams M)
{
    __shared__ uint32_t s_tile[kScanTile + kScanTile / 32];
    __shared__ uint32_t s_wave[kMaxScanThreads / kWave + 1];
    const uint32_t q = blockIdx.x;
    const uint32_t r = blockIdx.y;
    const ReplicaDesc* d = desc + r;
    const uint64_t base = static_cast<uint64_t>(q) * L.chunks + d->first_chunk;
    const uint32_t total = scan_column(M.counts + base, M.prefix + base, d->chunks, s_tile, s_wave);
    if (threadIdx.x == 0) M.totals[static_cast<uint64_t>(r) * kMaxRanks + q] = total;
}

// Pass 3: one wave per destination rank q.  Rank q's new owned array is its
// retained particles, then those of ranks 0, 1, ..., P-1 in rank order
// (src/migration.cpp:456-518): base[q][r] = particles of the sources before r
// in that order.  Writes q's new owned count (device and host memory).
__global__ __launch_bounds__(kWave) void migrate_bases_kernel(MigrateParams M)
{
    const uint32_t q = blockIdx.x;
    if (q >= M.ranks) return;
    const uint32_t lane = lane_id();
    const uint32_t count = lane < M.ranks ? M.totals[static_cast<uint64_t>(lane) * kMaxRanks + q] : 0u;
    auto key = [q](uint32_t r) { return r == q ? 0u : (r < q ? r + 1u : r); };
    uint32_t base = 0;
    for (uint32_t r = 0; r < M.ranks; ++r) {
        const uint32_t c = __builtin_amdgcn_readlane(count, r);
        if (key(r) < key(lane)) base += c;
    }
    if (lane < M.ranks) M.base[static_cast<uint64_t>(q) * kMaxRanks + lane] = base;
    const uint32_t incl = wave_inclusive_sum(count);
    if (lane == kWave - 1) {
        M.new_owned[q] = incl;
        M.new_owned_host[q] = incl;
    }
}

// One array of a chunk moved to its new slots: the chunk's ITEMS loads are
// issued together (coalesced, owned order), then the ITEMS stores.  `op`
// transforms each value on the way (the periodic wrap of a coordinate).
template <int ITEMS, typename T, typename Op>
__device__ __forceinline__ void move_column(const T* __restrict__ src, T* __restrict__ dst,
                                            uint32_t first, uint32_t owned, bool full,
                                            const uint32_t (&to)[ITEMS], Op op)
{
    const uint32_t lane = lane_id();
    T v[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t i = first + k * kWave + lane;
        v[k] = (full || i < owned) ? src[i] : T();
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t i = first + k * kWave + lane;
        if (full || i < owned) dst[to[k]] = op(v[k]);
    }
}

// One registered field of a chunk moved to its new slots, `units` values of
// T per particle.
template <int ITEMS, typename T>
__device__ __forceinline__ void move_field_units(const T* __restrict__ from, T* __restrict__ into,
                                                 uint32_t units, uint32_t first, uint32_t owned,
                                                 const uint32_t (&to)[ITEMS])
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const uint32_t i = first + k * kWave + lane;
        if (i >= owned) continue;
        const uint64_t a = static_cast<uint64_t>(i) * units;
        const uint64_t b = static_cast<uint64_t>(to[k]) * units;
        for (uint32_t c = 0; c < units; ++c) into[b + c] = from[a + c];
    }
}

struct Same {
    template <typename T>
    __device__ __forceinline__ T operator()(T v) const { return v; }
};

struct WrapAxis {
    int periodic;
    double lower, extent;
    __device__ __forceinline__ double operator()(double v) const
    {
        return periodic ? wrap(v, lower, extent) : v;
    }
};

// Pass 4: one wave per chunk of a source replica r.  Each owned particle goes
// to slot base[q][r] + prefix[q][chunk] + its rank among the chunk's
// particles bound for q (owned order) of rank q's new arrays, as its wrapped
// record (wrapped_record: position wrapped, every other member copied) with
// the words of every registered field (from `ff` into `ft`).  The new slots
// of the chunk's particles are found first (one round per distinct
// destination of an item row, usually one); then every array moves as a
// column.
template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void migrate_scatter_kernel(
    dsdx_particles src, dsdx_particles dst, FieldParams ff, FieldParams ft,
    const ReplicaDesc* __restrict__ desc, const uint32_t* __restrict__ chunk_replica, Layout L,
    MigrateParams M)
{
    constexpr uint32_t K = ITEMS * kWave;
    __shared__ uint32_t s_run[WAVES][kMaxRanks];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t chunk = wave_chunk<WAVES>(L, L.chunks);
    if (chunk >= L.chunks) return;
    // nothing moves when any subdomain would overflow
    if (__ballot(lane < M.ranks && M.new_owned[lane] > desc[lane].capacity)) return;
    const uint32_t r = replica_of(chunk_replica, chunk);
    const ReplicaDesc* d = desc + r;
    s_run[wave][lane] = 0u;
    const uint32_t owned = d->owned;
    const uint64_t off = d->offset;
    const uint32_t first = (chunk - d->first_chunk) * K;
    const bool full = first + K <= owned;
    const uint64_t slot0 = static_cast<uint64_t>(chunk) * K;
    uint32_t to[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        bool pending = first + k * kWave + lane < owned;
        const uint32_t q = pending ? M.dest[slot0 + k * kWave + lane] : 0u;
        to[k] = 0u;
        for (uint64_t left = __ballot(pending); left; left = __ballot(pending)) {
            const uint32_t q0 = __builtin_amdgcn_readlane(q, static_cast<int>(__builtin_ctzll(left)));
            const bool mine = pending && q == q0;
            const uint64_t group = __ballot(mine);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const uint32_t run = uniform(s_run[wave][q0]);
            if (mine) {
                // ranks * capacity <= 2^32: a slot fits 32 bits
                to[k] = static_cast<uint32_t>(desc[q0].offset) +
                        M.base[static_cast<uint64_t>(q0) * kMaxRanks + r] +
                        M.prefix[static_cast<uint64_t>(q0) * L.chunks + chunk] + run + lanes_below(group);
                pending = false;
            }
            if (lane == 0) s_run[wave][q0] = run + static_cast<uint32_t>(__popcll(group));
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    }
    const BoxParams& bp = d->bp;
    move_column<ITEMS>(src.x + off, dst.x, first, owned, full, to,
                       WrapAxis{bp.periodic[0], bp.lower[0], bp.extent[0]});
    move_column<ITEMS>(src.y + off, dst.y, first, owned, full, to,
                       WrapAxis{bp.periodic[1], bp.lower[1], bp.extent[1]});
    move_column<ITEMS>(src.z + off, dst.z, first, owned, full, to,
                       WrapAxis{bp.periodic[2], bp.lower[2], bp.extent[2]});
    move_column<ITEMS>(src.vx + off, dst.vx, first, owned, full, to, Same{});
    move_column<ITEMS>(src.vy + off, dst.vy, first, owned, full, to, Same{});
    move_column<ITEMS>(src.vz + off, dst.vz, first, owned, full, to, Same{});
    move_column<ITEMS>(src.fx + off, dst.fx, first, owned, full, to, Same{});
    move_column<ITEMS>(src.fy + off, dst.fy, first, owned, full, to, Same{});
    move_column<ITEMS>(src.fz + off, dst.fz, first, owned, full, to, Same{});
    move_column<ITEMS>(src.mass + off, dst.mass, first, owned, full, to, Same{});
    move_column<ITEMS>(src.charge + off, dst.charge, first, owned, full, to, Same{});
    move_column<ITEMS>(src.type + off, dst.type, first, owned, full, to, Same{});
    move_column<ITEMS>(src.global_id + off, dst.global_id, first, owned, full, to, Same{});
    for (uint32_t f = 0; f < ff.count; ++f) {
        if ((ff.wide >> f) & 1u)
            move_field_units<ITEMS>(static_cast<const uint64_t*>(ff.data[f]) + off * (ff.words[f] / 2),
                                    static_cast<uint64_t*>(ft.data[f]), ff.words[f] / 2, first, owned, to);
        else
            move_field_units<ITEMS>(static_cast<const uint32_t*>(ff.data[f]) + off * ff.words[f],
                                    static_cast<uint32_t*>(ft.data[f]), ff.words[f], first, owned, to);
    }
}

}  // namespace

hipError_t launch_exchange_counts(const Layout& L, const PlanBuffers& B, uint32_t* host,
                                  hipStream_t stream)
{
    if (L.replicas == 0) return hipSuccess;
    return launch(exchange_counts_kernel, dim3(L.replicas), dim3(kWave), stream, L.desc, L, B, host);
}

hipError_t launch_migrate_count(const dsdx_particles& p, const Layout& L, const MigrateParams& M,
                                const Shape& s, hipStream_t stream)
{
    if (L.chunks == 0) return hipSuccess;
    const Layout O = ordered(L, s, kOrderMigrate);
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        return launch(migrate_count_kernel<I, W>, chunk_grid(L, W), dim3(W * kWave), stream,
                      static_cast<const double*>(p.x), static_cast<const double*>(p.y),
                      static_cast<const double*>(p.z), L.desc, L.chunk_replica, O, M);
    });
}

hipError_t launch_migrate_scan(const Layout& L, const MigrateParams& M, const Shape& s,
                               hipStream_t stream)
{
    if (L.replicas == 0) return hipSuccess;
    return launch(migrate_scan_kernel, dim3(M.ranks, L.replicas), dim3(s.scan_threads), stream,
                  L.desc, L, M);
}

hipError_t launch_migrate_bases(const MigrateParams& M, hipStream_t stream)
{
    return launch(migrate_bases_kernel, dim3(M.ranks), dim3(kWave), stream, M);
}

hipError_t launch_migrate_scatter(const dsdx_particles& from, const dsdx_particles& to,
                                  const FieldParams& fields_from, const FieldParams& fields_to,
                                  const Layout& L, const MigrateParams& M, const Shape& s,
                                  hipStream_t stream)
{
    if (L.chunks == 0) return hipSuccess;
    const Layout O = ordered(L, s, kOrderMigrateScatter);
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        return launch(migrate_scatter_kernel<I, W>, chunk_grid(L, W), dim3(W * kWave), stream, from,
                      to, fields_from, fields_to, L.desc, L.chunk_replica, O, M);
    });
}

}  // namespace dsdx
