// halo_kernels.hip -- gfx950 kernels of the loopback particle-communication
// path.  Bandwidth-bound stream compaction and gathers; no MFMA.  Built with
// -ffp-contract=off: every FP64 operation is separately rounded, as in the
// reference Host build, so positions and sums are bit-identical to it.
//
// Every kernel walks a batch of replicas (device/halo layout in device.hpp);
// one wave owns one chunk of one replica, so replica parameters are
// wave-uniform and come in through scalar loads.
#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"
#include "chunk_units.hpp"
#include "owners.hpp"
#include "runs.hpp"
#include "wave.hpp"

namespace dsdx {
using namespace kern;
namespace {
__device__ __forceinline__ uint32_t slab_code(double px, double py, double pz, const BoxParams& bp)
{
    // Strict compares of src/halo.cpp:200-205 against thresholds computed on
    // the host exactly as lower+width and upper-width.
    return static_cast<uint32_t>(px < bp.lo_thr[0]) |
           (static_cast<uint32_t>(px > bp.hi_thr[0]) << 1) |
           (static_cast<uint32_t>(py < bp.lo_thr[1]) << 2) |
           (static_cast<uint32_t>(py > bp.hi_thr[1]) << 3) |
           (static_cast<uint32_t>(pz < bp.lo_thr[2]) << 4) |
           (static_cast<uint32_t>(pz > bp.hi_thr[2]) << 5);
}

// ---------------------------------------------------------------------------
// Pass 1: one wave per chunk.  Slab-classifies every owned particle and
// compacts the boundary ones (non-empty slot set) into the chunk's entry
// region in owned order, with their positions; counts the sends of each slot
// with one LDS atomic per (boundary particle, slot).  Every LDS array is
// per wave, so no block barrier is needed.  The slot-set table lut[64] sits one
// entry per lane and is read with a cross-lane permute.  Reads 24 B per
// particle, writes 28 B per boundary particle and 28 counters per chunk.
// With WRAP the pass first applies migrate's periodic wrap to the positions
// it loaded (storing the changed ones) and classifies the wrapped values:
// migrate followed by a rebuild with one read of x, y, z.
template <int ITEMS, int WAVES, bool WRAP, bool STREAM, int SUB>
__global__ __launch_bounds__(WAVES * kWave) void classify_kernel(
    double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    const ReplicaDesc* __restrict__ desc, const uint32_t* __restrict__ chunk_replica, Layout L,
    PlanBuffers B)
{
    constexpr uint32_t K = ITEMS * kWave;
    constexpr uint32_t kRing = 2 * kWave;
    static_assert(ITEMS % SUB == 0, "the chunk's items split into whole rounds");
    __shared__ uint32_t s_count[WAVES][kWave];
    __shared__ uint32_t s_ring_e[WAVES][kRing];
    __shared__ double s_ring_p[WAVES][3][kRing];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t chunk = wave_chunk<WAVES>(L, L.chunks);
    if (chunk >= L.chunks) return;
    const ReplicaDesc* d = desc + replica_of(chunk_replica, chunk);
    const uint32_t lut = d->bp.lut[lane];
    s_count[wave][lane] = 0u;
    const uint32_t owned = d->owned;
    const uint64_t off = d->offset;
    const uint32_t first = (chunk - d->first_chunk) * K;
    const bool full = first + K <= owned;
    // Boundary particles go through a per-wave LDS ring of 2 * 64 slots and
    // leave it 64 at a time, so the entry and position stores are full-wave
    // and contiguous.  The ring never holds more than 127 entries: it is
    // drained whenever 64 are waiting, and one item adds at most 64.
    const uint64_t ebase = static_cast<uint64_t>(chunk) * K;
    auto drain = [&](uint32_t from, bool active) {
        // LDS operations of one wave complete in order: the ring writes of
        // the other lanes are visible here without a barrier.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (active) {
            const uint32_t q = (from + lane) & (kRing - 1);
            const uint64_t e = ebase + from + lane;
            B.entries[e] = s_ring_e[wave][q];
            // the scatter reads these densely instead of gathering x, y, z
            B.bpos[e] = s_ring_p[wave][0][q];
            B.bpos[B.bpos_stride + e] = s_ring_p[wave][1][q];
            B.bpos[2 * B.bpos_stride + e] = s_ring_p[wave][2][q];
        }
        __builtin_amdgcn_wave_barrier();
    };
    uint32_t running = 0;
    uint32_t drained = 0;
    uint32_t images = 0;
    // The chunk's items in rounds of SUB: the loads of one round are in flight
    // together; fewer items per round hold fewer registers, so more waves run.
#pragma unroll
    for (uint32_t k0 = 0; k0 < ITEMS; k0 += SUB) {
        const uint32_t round_first = first + k0 * kWave;
        double px[SUB], py[SUB], pz[SUB];
        load_items<SUB, STREAM>(x + off, round_first, owned, full, px);
        load_items<SUB, STREAM>(y + off, round_first, owned, full, py);
        load_items<SUB, STREAM>(z + off, round_first, owned, full, pz);
        if (WRAP) {
#pragma unroll
            for (int k = 0; k < SUB; ++k) {
                const uint32_t i = round_first + k * kWave + lane;
                const bool valid = i < owned;
                wrap_store(px[k], x + off + i, valid, d->bp.periodic[0], d->bp.lower[0], d->bp.extent[0]);
                wrap_store(py[k], y + off + i, valid, d->bp.periodic[1], d->bp.lower[1], d->bp.extent[1]);
                wrap_store(pz[k], z + off + i, valid, d->bp.periodic[2], d->bp.lower[2], d->bp.extent[2]);
            }
        }
#pragma unroll
        for (int k = 0; k < SUB; ++k) {
            const uint32_t item = k0 + static_cast<uint32_t>(k);
            const uint32_t i = first + item * kWave + lane;
            const uint32_t code = i < owned ? slab_code(px[k], py[k], pz[k], d->bp) : 0u;
            const uint32_t slots = static_cast<uint32_t>(__shfl(static_cast<int>(lut), static_cast<int>(code), kWave));
            const bool boundary = slots != 0u;
            const uint64_t ballot = __ballot(boundary);
            if (boundary) {
                const uint32_t q = (running + lanes_below(ballot)) & (kRing - 1);
                s_ring_e[wave][q] = ((item * kWave + lane) << 8) | code;
                s_ring_p[wave][0][q] = px[k];
                s_ring_p[wave][1][q] = py[k];
                s_ring_p[wave][2][q] = pz[k];
                images += static_cast<uint32_t>(__popc(slots));
                for (uint32_t m = slots; m; m &= m - 1u) atomicAdd(&s_count[wave][__builtin_ctz(m)], 1u);
            }
            running += static_cast<uint32_t>(__popcll(ballot));
            if (running - drained >= kWave) {
                drain(drained, true);
                drained += kWave;
            }
        }
    }
    if (running > drained) drain(drained, lane < running - drained);
    const uint32_t image_total = __builtin_amdgcn_readlane(wave_inclusive_sum(images), kWave - 1);
    // LDS operations of one wave complete in order: the atomics above are
    // visible to the reads below without a barrier.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane < kCols) {
        // lane s < 26: send count of slot s; lane 26: owners; lane 27: images
        const uint32_t acc = lane < kSlots ? s_count[wave][lane]
                             : lane == kColOwners ? running : image_total;
        B.counts[static_cast<uint64_t>(lane) * L.chunks + chunk] = acc;
    }
}

// ---------------------------------------------------------------------------
// Pass 2: exclusive scan of one count column over one replica's chunks.
// totals[r][col] receives the column total (and the scan clears the status
// word, sets a loopback replica's received count, copies both to the mapped
// host totals of a single box, and clears a process grid's status).
__device__ __forceinline__ void column_done(const PlanBuffers& B, const Layout& L, uint32_t col, uint32_t r,
                                            uint32_t carry)
{
    uint32_t* tot = B.totals + static_cast<uint64_t>(r) * kTotals;
    tot[col] = carry;
    if (col == 0) tot[kStatusWord] = 0u;
    // a replica that sends to itself receives what it sends; a process
    // grid's exchange pass overwrites this with the neighbours' sends
    if (col == kColImages) tot[kColReceived] = carry;
    // one 8-byte store carries the total and the stamp, so the host sees
    // both or neither
    if (B.host_counts)
        __hip_atomic_store(B.host_counts + col, (static_cast<uint64_t>(B.stamp) << 32) | carry, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    if (B.recv && col == 0 && r == 0) B.recv[static_cast<uint64_t>(L.replicas) * kTotals] = 0u;
}

// One block per (count column, replica): grid (kCols, replicas).
__global__ __launch_bounds__(kMaxScanThreads) void scan_kernel(const ReplicaDesc* __restrict__ desc,
                                                               Layout L, PlanBuffers B)
{
    __shared__ uint32_t s_tile[kScanTile + kScanTile / 32];
    __shared__ uint32_t s_wave[kMaxScanThreads / kWave + 1];
    const uint32_t col = blockIdx.x;
    const uint32_t r = blockIdx.y;
    const ReplicaDesc* d = desc + r;
    const uint64_t base = static_cast<uint64_t>(col) * L.chunks + d->first_chunk;
    const uint32_t carry = scan_column(B.counts + base, B.prefix + base, d->chunks, s_tile, s_wave);
    if (threadIdx.x == 0) column_done(B, L, col, r, carry);
}

// Long columns cut into parts of one LDS tile (blockDim.x * kScanPer values):
// grid (kCols * parts, replicas).  The block of part p loads its tile and,
// in the same round of loads, every value before the tile, whose sum it
// reduces itself; the parts of a column therefore need nothing from each
// other, and every block waits for one round of loads instead of one per
// tile.  The part that ends the column writes its total.  A part reduces up
// to kMaxScanParts - 1 tiles before its own; longer columns take scan_kernel.
constexpr uint32_t kMaxScanParts = 8;

__global__ __launch_bounds__(kMaxScanThreads) void scan_parts_kernel(const ReplicaDesc* __restrict__ desc,
                                                                     Layout L, PlanBuffers B)
{
    __shared__ uint32_t s_tile[kScanTile + kScanTile / 32];
    __shared__ uint32_t s_wave[kMaxScanThreads / kWave + 1];
    const uint32_t T = blockDim.x;
    const uint32_t t = threadIdx.x;
    const uint32_t col = blockIdx.x % kCols;
    const uint32_t part = blockIdx.x / kCols;
    const uint32_t r = blockIdx.y;
    const ReplicaDesc* d = desc + r;
    const uint32_t n = d->chunks;
    const uint32_t span = T * kScanPer;
    const uint32_t lo = part * span;
    if (part > 0 && lo >= n) return;   // block-uniform: this replica has fewer parts
    const uint64_t base = static_cast<uint64_t>(col) * L.chunks + d->first_chunk;
    const uint32_t* src = B.counts + base;
    uint32_t* dst = B.prefix + base;
    const uint32_t hi = min(n, lo + span);
    uint32_t v[kScanPer];
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const uint32_t i = lo + k * T + t;
        v[k] = i < hi ? src[i] : 0u;
    }
    uint32_t before = 0;
    for (uint32_t i0 = 0; i0 < lo; i0 += kScanPer * T) {
        uint32_t w[kScanPer];
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            const uint32_t i = i0 + k * T + t;
            w[k] = i < lo ? src[i] : 0u;
        }
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) before += w[k];
    }
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) s_tile[scan_pad(k * T + t)] = v[k];
    uint32_t carry = 0;
    (void)block_exclusive(before, s_wave, T / kWave, &carry);   // its barriers also publish s_tile
    uint32_t mine[kScanPer];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        mine[j] = s_tile[scan_pad(t * kScanPer + j)];
        sum += mine[j];
    }
    uint32_t total = 0;
    uint32_t run = carry + block_exclusive(sum, s_wave, T / kWave, &total);
#pragma unroll
    for (int j = 0; j < kScanPer; ++j) {
        s_tile[scan_pad(t * kScanPer + j)] = run;
        run += mine[j];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        const uint32_t i = lo + k * T + t;
        if (i < hi) dst[i] = s_tile[scan_pad(k * T + t)];
    }
    if (hi == n && t == 0) column_done(B, L, col, r, carry + total);
}

// Short columns (every replica at most kShortScanChunks chunks, as in a batch
// of small boxes): one wave per (column, replica), kShortScanPer consecutive
// values per lane and a wave scan, no block barrier; a block holds
// kShortScanWaves such columns.
constexpr uint32_t kShortScanPer = 8;
constexpr uint32_t kShortScanWaves = 8;
constexpr uint32_t kShortScanChunks = 2 * kWave * kShortScanPer;

__global__ __launch_bounds__(kShortScanWaves * kWave) void scan_short_kernel(const ReplicaDesc* __restrict__ desc,
                                                                            Layout L, PlanBuffers B)
{
    const uint32_t lane = lane_id();
    const uint32_t pair = blockIdx.x * kShortScanWaves + uniform(threadIdx.x / kWave);
    if (pair >= static_cast<uint32_t>(kCols) * L.replicas) return;
    const uint32_t col = pair % kCols;
    const uint32_t r = pair / kCols;
    const ReplicaDesc* d = desc + r;
    const uint64_t base = static_cast<uint64_t>(col) * L.chunks + d->first_chunk;
    const uint32_t* src = B.counts + base;
    uint32_t* dst = B.prefix + base;
    const uint32_t n = d->chunks;
    uint32_t carry = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += kWave * kShortScanPer) {
        const uint32_t at = k0 + lane * kShortScanPer;
        uint32_t v[kShortScanPer];
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < kShortScanPer; ++j) {
            v[j] = at + j < n ? src[at + j] : 0u;
            sum += v[j];
        }
        const uint32_t incl = wave_inclusive_sum(sum);
        uint32_t run = carry + incl - sum;
#pragma unroll
        for (uint32_t j = 0; j < kShortScanPer; ++j) {
            if (at + j < n) dst[at + j] = run;
            run += v[j];
        }
        carry += uniform(__shfl(incl, kWave - 1, kWave));
    }
    if (lane == 0) column_done(B, L, col, r, carry);
}

// ---------------------------------------------------------------------------
// Copies of registered fields, in units of T.
template <class T>
__device__ __forceinline__ void copy_units(T* dd, uint32_t units, uint64_t from, uint64_t to)
{
    for (uint32_t k = 0; k < units; ++k) dd[to * units + k] = dd[from * units + k];
}

// Every registered field of particle `from` into particle `to`, in 8-byte
// units where the field allows it.
__device__ __forceinline__ void copy_field_words(const FieldParams& fp, uint64_t from, uint64_t to)
{
    for (uint32_t f = 0; f < fp.count; ++f) {
        if ((fp.wide >> f) & 1u)
            copy_units(static_cast<uint64_t*>(fp.data[f]), fp.words[f] / 2, from, to);
        else
            copy_units(static_cast<uint32_t*>(fp.data[f]), fp.words[f], from, to);
    }
}

// Pass 3 (route): one wave per chunk.  Walks the chunk's boundary particles
// in owned order; for each of its send slots s (ascending) a particle gets
// the ghost at receive block 25-s, position prefix[s][chunk] + rank in the
// chunk.  Writes the owner -> ghost map (owner_index, owner_code,
// image_start, owner_entry, ghost_list), from which the owner-major passes
// work.  The capacity test is uniform over a replica: an overflowing replica
// gets its status word set and no route.
template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void route_kernel(
    const ReplicaDesc* __restrict__ desc, const uint32_t* __restrict__ chunk_replica, Layout L,
    PlanBuffers B)
{
    constexpr uint32_t K = ITEMS * kWave;
    const uint32_t lane = lane_id();
    const uint32_t chunk = wave_chunk<WAVES>(L, L.chunks);
    if (chunk >= L.chunks) return;
    const uint32_t ebase = chunk * K;
    const uint32_t r = replica_of(chunk_replica, chunk);
    const ReplicaDesc* d = desc + r;
    uint32_t* totals = B.totals + static_cast<uint64_t>(r) * kTotals;
    // Every load the chunk needs before its walk goes out together, before
    // anything waits: the chunk's prefix of every column (lane c: column c),
    // the replica's totals row (lane w: word w), the chunk's boundary count
    // and its first 64 entries (the entry region always holds K); 18.6 ->
    // 17.1 us at 16 M particles with the prefix and entries.
    const uint32_t col_prefix = lane < kCols ? B.prefix[static_cast<uint64_t>(lane) * L.chunks + chunk] : 0u;
    const uint32_t total_row = totals[lane & (kTotals - 1)];
    const uint32_t entry_first = B.entries[ebase + lane];
    const uint32_t n_boundary = uniform(B.counts[static_cast<uint64_t>(kColOwners) * L.chunks + chunk]);
    const uint32_t lut = d->bp.lut[lane];

    const uint32_t owned = d->owned;
    const uint32_t received = static_cast<uint32_t>(__builtin_amdgcn_readlane(total_row, kColReceived));
    const uint32_t images = static_cast<uint32_t>(__builtin_amdgcn_readlane(total_row, kColImages));
    uint32_t overflow = 0;
    if (static_cast<uint64_t>(owned) + received > d->capacity) overflow |= kOverflowStorage;
    if (images > d->ghost_list_capacity) overflow |= kOverflowGhostList;
    // a process grid routes nothing when any of its subdomains overflowed
    if (B.recv) overflow |= uniform(B.recv[static_cast<uint64_t>(L.replicas) * kTotals]);
    if (overflow) {
        if (chunk == d->first_chunk && lane == 0) totals[kStatusWord] = overflow;
        return;
    }

    // Lane s ends with slot_base = first ghost, in the receiving replica, of
    // this chunk's sends through slot s.  On a self-loopback box receive
    // block m holds send slot 25-m: recv_base[m] = sum of totals of slots
    // 25..26-m.  On a process grid the exchange pass tabulated the receive
    // blocks of every replica.
    uint32_t slot_first = 0;
    if (B.recv) {
        const uint32_t peer = lane < kSlots ? d->dest[lane] : kNoPeer;
        slot_first = peer != kNoPeer ? B.recv[static_cast<uint64_t>(peer) * kTotals + kSlots - 1 - lane] : 0u;
    } else {
        const uint32_t block_size =
            static_cast<uint32_t>(__shfl(static_cast<int>(total_row), lane < kSlots ? kSlots - 1 - lane : 0, kWave));
        const uint32_t recv_base = wave_inclusive_sum(block_size) - block_size;
        slot_first = __shfl(recv_base, lane < kSlots ? kSlots - 1 - lane : 0, kWave);
    }
    const uint32_t slot_base =
        lane < kSlots ? slot_first + col_prefix : 0u;
    uint32_t slot_run = 0;   // lane s: ghosts of slot s already routed from this chunk

    const uint32_t owner_base = static_cast<uint32_t>(__builtin_amdgcn_readlane(col_prefix, kColOwners));
    uint32_t image_run = static_cast<uint32_t>(__builtin_amdgcn_readlane(col_prefix, kColImages));
    const uint32_t first = (chunk - d->first_chunk) * K;
    const uint64_t slot0 = static_cast<uint64_t>(d->first_chunk) * K + owner_base;
    uint32_t* glist = B.ghost_list + d->ghost_list_offset;
    uint32_t* rows = B.unit_rows ? B.unit_rows + static_cast<uint64_t>(d->first_unit + r) * kSlots : nullptr;
    if (rows && chunk == d->first_chunk && lane < kSlots) {
        // the end row: one past the replica's last ghost of each slot
        const uint32_t owners = static_cast<uint32_t>(__builtin_amdgcn_readlane(total_row, kColOwners));
        rows[static_cast<uint64_t>((owners + kRunOwners - 1) / kRunOwners) * kSlots + lane] =
            slot_first + total_row;
    }

    for (uint32_t e0 = 0; e0 < n_boundary; e0 += kWave) {
        const uint32_t e = e0 + lane;
        const bool valid = e < n_boundary;
        const uint32_t entry = valid ? (e0 == 0 ? entry_first : B.entries[ebase + e]) : 0u;
        const uint32_t code = entry & 63u;
        const uint32_t slots = static_cast<uint32_t>(__shfl(static_cast<int>(lut), static_cast<int>(code), kWave));
        const uint32_t images = static_cast<uint32_t>(__popc(slots));
        const uint32_t incl = wave_inclusive_sum(images);
        const uint32_t first_image = image_run + incl - images;
        image_run += uniform(__shfl(incl, kWave - 1, kWave));
        if (valid) {
            const uint64_t u = slot0 + e;
            B.owner_index[u] = first + (entry >> 8);
            B.owner_code[u] = static_cast<uint8_t>(code);
            B.image_start[u] = first_image;
            B.owner_entry[u] = ebase + e;
        }
        // These 64 owners hold at most one first owner of a ghost-run unit:
        // the one in lane `at`, if any.
        const uint32_t at = (kRunOwners - (owner_base + e0) % kRunOwners) % kRunOwners;
        const bool unit_start = rows && at < kWave && e0 + at < n_boundary;
        const uint64_t before_at = at < kWave ? (1ull << at) - 1u : ~0ull;
        uint32_t row_value = slot_base + slot_run;   // lane s: slot s's next ghost-list value
        uint32_t present = uniform(wave_or(slots));
        uint32_t written = 0;
        while (present) {
            const uint32_t s = uniform(static_cast<uint32_t>(__builtin_ctz(present)));
            present &= present - 1u;
            const bool mine = (slots >> s) & 1u;
            const uint64_t ballot = __ballot(mine);
            const uint32_t base = __builtin_amdgcn_readlane(slot_base, s) +
                                  __builtin_amdgcn_readlane(slot_run, s);
            if (mine) {
                const uint32_t rel = base + lanes_below(ballot);
                glist[first_image + written] = rel;
                ++written;
            }
            if (lane == s) {
                slot_run += static_cast<uint32_t>(__popcll(ballot));
                row_value += static_cast<uint32_t>(__popcll(ballot & before_at));
            }
        }
        if (unit_start && lane < kSlots)
            rows[static_cast<uint64_t>((owner_base + e0 + at) / kRunOwners) * kSlots + lane] = row_value;
    }
}

// The particle slots of an owner's images: the first kPreloaded from the
// entries loaded with the owner, the rest from its ghost list.
struct Images {
    const OwnerGhosts& g;
    const FirstImages& f;
    const SlotTable& t;
    __device__ uint64_t first(uint32_t k) const { return t.ghost0[f.slot[k]] + f.rel[k]; }
    __device__ uint64_t later(uint32_t k, uint32_t s) const { return t.ghost0[s] + g.list[k]; }
};

// Calls put(send slot, particle slot) for every image of an owner, in
// ascending send slot.
template <class Put>
__device__ __forceinline__ void each_image(const Images& im, Put&& put)
{
#pragma unroll
    for (uint32_t k = 0; k < kPreloaded; ++k) {
        if (k >= im.f.n) break;
        put(im.f.slot[k], im.first(k));
    }
    uint32_t k = kPreloaded;
    for (uint32_t rest = im.f.rest; rest; rest &= rest - 1u, ++k) {
        const uint32_t slot = static_cast<uint32_t>(__builtin_ctz(rest));
        put(slot, im.later(k, slot));
    }
}

// Pass 4 (copy): one wave per owner tile, one lane per boundary owner, in
// owned order (one ascending sweep over the owner arrays).  Each owner's
// ghosts get position + shift[slot] (the owner's position comes from the
// classify pass, densely), velocity, mass, charge, type, id and the
// registered fields of the owner, forces zero (src/halo.cpp:1232-1262).
template <int WAVES, bool STREAM>
__global__ __launch_bounds__(WAVES * kWave) void copy_kernel(
    dsdx_particles p, FieldParams fp, const ReplicaDesc* __restrict__ desc,
    const uint32_t* __restrict__ tile_replica, Layout L, PlanBuffers B)
{
    __shared__ SlotTable s_table[WAVES];
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        const SlotTable& t = fill_slot_table(s_table[threadIdx.x / kWave], desc, o.d);
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const uint32_t E = B.owner_entry[o.u];
        const OwnerGhosts g = ghosts_of(desc, o, B);
        const double px = B.bpos[E];
        const double py = B.bpos[B.bpos_stride + E];
        const double pz = B.bpos[2 * B.bpos_stride + E];
        const double vx = read_once<STREAM>(p.vx + i), vy = read_once<STREAM>(p.vy + i), vz = read_once<STREAM>(p.vz + i);
        const double m = read_once<STREAM>(p.mass + i), q = read_once<STREAM>(p.charge + i);
        const int32_t ty = read_once<STREAM>(p.type + i), id = read_once<STREAM>(p.global_id + i);
        const FirstImages f = first_images(g);
        each_image(Images{g, f, t}, [&](uint32_t s, uint64_t j) {
            p.x[j] = px + t.shift[s][0];
            p.y[j] = py + t.shift[s][1];
            p.z[j] = pz + t.shift[s][2];
            p.vx[j] = vx; p.vy[j] = vy; p.vz[j] = vz;
            if (p.fx) p.fx[j] = 0.0;
            if (p.fy) p.fy[j] = 0.0;
            if (p.fz) p.fz[j] = 0.0;
            p.mass[j] = m; p.charge[j] = q;
            p.type[j] = ty; p.global_id[j] = id;
            if (fp.count) copy_field_words(fp, i, j);
        });
    });
}

// The ghost copy and the reused-plan refresh through the staged runs of
// runs.hpp.  MODE kRunsBuild: the rebuild's copy, positions from the
// classify pass (bpos).  kRunsRefresh: the refresh of the whole record, the
// owner's current position.  kRunsPositions: the positions-only refresh
// (x, y, z and zeroed forces).  The receive block of slot s is the
// replica's own (a loopback box or batch) or the neighbouring subdomain's (a
// process grid), through the slot table.  Registered fields are copied per
// owner while the records are staged.
constexpr int kRunsBuild = 0;
constexpr int kRunsRefresh = 1;
constexpr int kRunsPositions = 2;

template <int MODE, class Stage>
__device__ __forceinline__ void write_run_ghost(const dsdx_particles& p, const Stage& sh, uint32_t ol,
                                                uint32_t s, uint64_t j)
{
    p.x[j] = sh.px[ol] + sh.table.shift[s][0];
    p.y[j] = sh.py[ol] + sh.table.shift[s][1];
    p.z[j] = sh.pz[ol] + sh.table.shift[s][2];
    if (p.fx) p.fx[j] = 0.0;
    if (p.fy) p.fy[j] = 0.0;
    if (p.fz) p.fz[j] = 0.0;
    if (MODE != kRunsPositions) {
        p.vx[j] = sh.vx[ol]; p.vy[j] = sh.vy[ol]; p.vz[j] = sh.vz[ol];
        p.mass[j] = sh.m[ol]; p.charge[j] = sh.q[ol];
        p.type[j] = sh.ty[ol]; p.global_id[j] = sh.id[ol];
    }
}

template <int MODE, bool STREAM>
__global__ __launch_bounds__(kRunWaves * kWave) void copy_runs_kernel(dsdx_particles p, FieldParams fp,
                                                                      const ReplicaDesc* __restrict__ desc,
                                                                      const uint32_t* __restrict__ unit_replica,
                                                                      Layout L, PlanBuffers B)
{
    __shared__ RunStage sh;
    for_run_units(desc, unit_replica, L, B, [&](const ReplicaDesc* d, uint32_t e0, uint32_t owners) {
        ghost_runs_unit<STREAM, MODE == kRunsBuild, MODE != kRunsPositions>(
            sh, p, desc, d, e0, owners, L, B, fp.count != 0,
            [&](uint64_t i, uint32_t s, uint32_t rel) { copy_field_words(fp, i, ghost_slot(desc, d, s, rel)); },
            [&](uint32_t ol, uint32_t s, uint32_t rel) {
                write_run_ghost<MODE>(p, sh, ol, s, sh.table.ghost0[s] + rel);
            });
    });
}

// reverse_halo: f[i] = ((f[i] + g_s1) + g_s2) + ... over the owner's send
// slots s1 < s2 < ..., the order of the Host reduction (src/halo.cpp:1357-1370
// for n ascending).  No atomics: one lane sums one owner.  The forces of the
// first kPreloaded images are loaded together, then added in slot order.
template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void reverse_forces_kernel(
    double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
    const ReplicaDesc* __restrict__ desc, const uint32_t* __restrict__ tile_replica, Layout L,
    PlanBuffers B)
{
    __shared__ SlotTable s_table[WAVES];
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        const SlotTable& t = fill_slot_table(s_table[threadIdx.x / kWave], desc, o.d);
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const OwnerGhosts g = ghosts_of(desc, o, B);
        double ax = fx[i], ay = fy[i], az = fz[i];
        const FirstImages f = first_images(g);
        const Images im{g, f, t};
        double gx[kPreloaded], gy[kPreloaded], gz[kPreloaded];
#pragma unroll
        for (uint32_t k = 0; k < kPreloaded; ++k) {
            const uint64_t j = im.first(k);
            gx[k] = fx[j]; gy[k] = fy[j]; gz[k] = fz[j];
        }
#pragma unroll
        for (uint32_t k = 0; k < kPreloaded; ++k) {
            if (k < f.n) {
                ax = ax + gx[k];
                ay = ay + gy[k];
                az = az + gz[k];
            }
        }
        uint32_t k = kPreloaded;
        for (uint32_t rest = f.rest; rest; rest &= rest - 1u, ++k) {
            const uint64_t j = im.later(k, static_cast<uint32_t>(__builtin_ctz(rest)));
            ax = ax + fx[j];
            ay = ay + fy[j];
            az = az + fz[j];
        }
        fx[i] = ax; fy[i] = ay; fz[i] = az;
    });
}

// reverse_halo through the staged runs (Shape::reverse_runs): a unit's ghost
// forces read as dense runs into LDS, each owner's images added from there in
// ascending send slot (runs.hpp).  Against the per-owner kernel: 512 replica
// boxes of 2,048-262,144 particles 434 -> 398 us; one box of 16 M particles
// 102.7 -> 94.0 us.
__global__ __launch_bounds__(kRunWaves * kWave) void reverse_runs_kernel(double* __restrict__ fx,
                                                                         double* __restrict__ fy,
                                                                         double* __restrict__ fz,
                                                                         const ReplicaDesc* __restrict__ desc,
                                                                         const uint32_t* __restrict__ unit_replica,
                                                                         Layout L, PlanBuffers B)
{
    __shared__ ReverseStage sh;
    for_run_units(desc, unit_replica, L, B, [&](const ReplicaDesc* d, uint32_t e0, uint32_t owners) {
        reverse_runs_unit(sh, fx, fy, fz, desc, d, e0, owners, L, B,
                          [&](uint32_t s, uint32_t rel, double& gx, double& gy, double& gz) {
                              const uint64_t j = ghost_slot(desc, d, s, rel);
                              gx = fx[j];
                              gy = fy[j];
                              gz = fz[j];
                          });
    });
}

// ---------------------------------------------------------------------------
// The rebuild's copy and reverse_halo of one box over chunk units
// (chunk_units.hpp): every ghost position comes from the entries and the
// prefixes, so the build skips the route pass; the owner maps and the ghost
// list are routed only when a later call needs them.

struct ChunkStage {
    double px[kChunkOwners], py[kChunkOwners], pz[kChunkOwners];
    double vx[kChunkOwners], vy[kChunkOwners], vz[kChunkOwners], m[kChunkOwners], q[kChunkOwners];
    int32_t ty[kChunkOwners], id[kChunkOwners];
    uint16_t map[kChunkMap];   // run index -> owner (index in the unit)
    SlotTable table;
};

// One workgroup per unit.  Wave w ranks the owners of the unit's chunks w, w
// + kRunWaves, ... per slot, gathers their records (positions from the
// classification pass) into LDS and maps each image's run index to its
// owner; after one barrier lane k of a round writes run index k.  A unit
// that does not fit the stage writes each owner's images from its lane.
template <bool STREAM>
__global__ __launch_bounds__(kRunWaves * kWave) void copy_chunks_kernel(dsdx_particles p, FieldParams fp,
                                                                        const ReplicaDesc* __restrict__ desc,
                                                                        Layout L, PlanBuffers B)
{
    __shared__ ChunkStage sh;
    const uint32_t units = chunk_units_of(L);
    const uint32_t b = xcd_block(L.xcd_run);
    if (b >= units) return;
    const uint32_t unit = L.descending ? units - 1u - b : b;
    const ReplicaDesc* d = desc;
    const uint32_t lane = lane_id();
    const uint32_t w = uniform(threadIdx.x / kWave);
    if (w == 0) fill_slot_table(sh.table, desc, d);
    const ChunkUnit u = chunk_unit(d, unit, L, B);
    // a single box sends every slot to itself: its ghosts follow its owned
    // particles
    const uint64_t ghost0 = d->offset + d->owned;
    if (u.overflow) {
        // the status word the route pass sets; nothing is copied
        if (unit == 0 && threadIdx.x == 0) B.totals[kStatusWord] = u.overflow;
        return;
    }
    const bool staged = u.owners <= kChunkOwners && u.runs.images <= kChunkMap;
    const uint32_t lut = d->bp.lut[lane];
    const uint32_t K = L.chunk_size;
    for (uint32_t c = u.c0 + w; c < u.c1; c += kRunWaves) {
        const ChunkHead h = chunk_head(c, L, B);
        if (lane < kCols) B.chunk_rows[static_cast<uint64_t>(c) * kTotals + lane] = h.pc;
        const uint64_t ebase = static_cast<uint64_t>(c) * K;
        const uint64_t first = d->offset + static_cast<uint64_t>(c - d->first_chunk) * K;
        const uint32_t ol0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(h.pc, kColOwners)) - u.owner0;
        const uint32_t rbase = u.runs.start + (h.pc - u.pre0);   // lane s: run index of the chunk's first image through s
        const uint32_t gbase = u.slot_first + h.pc;              // lane s: its ghost-list value
        uint32_t run = 0;
        for (uint32_t e0 = 0; e0 < h.n; e0 += kChunkRounds * kWave) {
            uint32_t entry[kChunkRounds];
#pragma unroll
            for (uint32_t k = 0; k < kChunkRounds; ++k) entry[k] = chunk_entry(h, e0, k, ebase, B);
            if (staged) {
                // every record of the rounds loaded first, straight into LDS,
                // then each round's ranks
#pragma unroll
                for (uint32_t k = 0; k < kChunkRounds; ++k) {
                    const uint32_t e = e0 + k * kWave + lane;
                    if (e < h.n) {
                        const uint64_t i = first + (entry[k] >> 8);
                        const uint32_t ol = ol0 + e;
                        sh.px[ol] = B.bpos[ebase + e];
                        sh.py[ol] = B.bpos[B.bpos_stride + ebase + e];
                        sh.pz[ol] = B.bpos[2 * B.bpos_stride + ebase + e];
                        sh.vx[ol] = read_once<STREAM>(p.vx + i);
                        sh.vy[ol] = read_once<STREAM>(p.vy + i);
                        sh.vz[ol] = read_once<STREAM>(p.vz + i);
                        sh.m[ol] = read_once<STREAM>(p.mass + i);
                        sh.q[ol] = read_once<STREAM>(p.charge + i);
                        sh.ty[ol] = read_once<STREAM>(p.type + i);
                        sh.id[ol] = read_once<STREAM>(p.global_id + i);
                    }
                }
#pragma unroll
                for (uint32_t k = 0; k < kChunkRounds; ++k) {
                    if (e0 + k * kWave >= h.n) break;   // wave-uniform
                    const uint32_t e = e0 + k * kWave + lane;
                    const uint32_t code_slots = lane_value(lut, entry[k] & 63u);
                    const uint32_t slots = e < h.n ? code_slots : 0u;
                    const uint64_t i = first + (entry[k] >> 8);
                    const uint32_t ol = ol0 + e;
                    rank_round(slots, run, [&](uint32_t s, uint32_t r, bool mine) {
                        const uint32_t rb = static_cast<uint32_t>(__builtin_amdgcn_readlane(rbase, s));
                        const uint32_t g = static_cast<uint32_t>(__builtin_amdgcn_readlane(gbase, s));
                        if (!mine) return;
                        sh.map[rb + r] = static_cast<uint16_t>(ol);
                        if (fp.count) copy_field_words(fp, i, ghost0 + (g + r));
                    });
                }
            } else {
                // a unit larger than the stage: one round at a time, each
                // owner writing its images from registers
#pragma unroll
                for (uint32_t k = 0; k < kChunkRounds; ++k) {
                    if (e0 + k * kWave >= h.n) break;   // wave-uniform
                    const uint32_t e = e0 + k * kWave + lane;
                    const bool valid = e < h.n;
                    const uint32_t code_slots = lane_value(lut, entry[k] & 63u);
                    const uint32_t slots = valid ? code_slots : 0u;
                    const uint64_t i = first + (entry[k] >> 8);
                    double x = 0.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0, m = 0.0, q = 0.0;
                    int32_t ty = 0, id = 0;
                    if (valid) {
                        x = B.bpos[ebase + e];
                        y = B.bpos[B.bpos_stride + ebase + e];
                        z = B.bpos[2 * B.bpos_stride + ebase + e];
                        vx = read_once<STREAM>(p.vx + i);
                        vy = read_once<STREAM>(p.vy + i);
                        vz = read_once<STREAM>(p.vz + i);
                        m = read_once<STREAM>(p.mass + i);
                        q = read_once<STREAM>(p.charge + i);
                        ty = read_once<STREAM>(p.type + i);
                        id = read_once<STREAM>(p.global_id + i);
                    }
                    rank_round(slots, run, [&](uint32_t s, uint32_t r, bool mine) {
                        const uint32_t g = static_cast<uint32_t>(__builtin_amdgcn_readlane(gbase, s));
                        if (!mine) return;
                        const uint64_t j = ghost0 + (g + r);
                        p.x[j] = x + d->bp.shift[s][0];
                        p.y[j] = y + d->bp.shift[s][1];
                        p.z[j] = z + d->bp.shift[s][2];
                        if (p.fx) p.fx[j] = 0.0;
                        if (p.fy) p.fy[j] = 0.0;
                        if (p.fz) p.fz[j] = 0.0;
                        p.vx[j] = vx; p.vy[j] = vy; p.vz[j] = vz;
                        p.mass[j] = m; p.charge[j] = q;
                        p.type[j] = ty; p.global_id[j] = id;
                        if (fp.count) copy_field_words(fp, i, j);
                    });
                }
            }
        }
    }
    if (!staged) return;
    __syncthreads();
    // lane k of a round writes run index k (every lane takes part in the
    // shuffles of each round)
    for (uint32_t r0 = 0; r0 < u.runs.images; r0 += kRunWaves * kWave) {
        const uint32_t r = r0 + threadIdx.x;
        const uint32_t s = run_slot(u.runs.start, r);
        const uint32_t rel = lane_value(u.runs.first, s) + (r - lane_value(u.runs.start, s));
        if (r < u.runs.images) write_run_ghost<kRunsBuild>(p, sh, sh.map[r], s, sh.table.ghost0[s] + rel);
    }
}

// reverse_halo over chunk units, one lane per owner reading each image's
// force from its ghost: the units whose owners or images exceed the LDS
// stage of reverse_chunks_kernel.  Wave w walks the unit's chunks w, w +
// kRunWaves, ... and ranks each round of 64 owners per slot with ballots.
__device__ __forceinline__ void reverse_chunk_walk(double* __restrict__ fx, double* __restrict__ fy,
                                                   double* __restrict__ fz, const ReplicaDesc* d,
                                                   const ChunkUnit& u, uint64_t ghost0, uint32_t lut,
                                                   const Layout& L, const PlanBuffers& B)
{
    const uint32_t lane = lane_id();
    const uint32_t w = uniform(threadIdx.x / kWave);
    const uint32_t K = L.chunk_size;
    for (uint32_t c = u.c0 + w; c < u.c1; c += kRunWaves) {
        const ChunkHead h = chunk_head(c, L, B);
        const uint64_t ebase = static_cast<uint64_t>(c) * K;
        const uint64_t first = d->offset + static_cast<uint64_t>(c - d->first_chunk) * K;
        const uint32_t gbase = u.slot_first + h.pc;   // lane s: ghost-list value of the chunk's first image through s
        uint32_t run = 0;
        for (uint32_t e0 = 0; e0 < h.n; e0 += kWave) {
            const uint32_t e = e0 + lane;
            const bool valid = e < h.n;
            const uint32_t entry = valid ? (e0 == 0 ? h.entry[0] : B.entries[ebase + e]) : 0u;
            const uint32_t code_slots = lane_value(lut, entry & 63u);
            const uint32_t slots = valid ? code_slots : 0u;
            const uint64_t i = first + (entry >> 8);
            double ax = 0.0, ay = 0.0, az = 0.0;
            if (valid) {
                ax = fx[i];
                ay = fy[i];
                az = fz[i];
            }
            rank_round(slots, run, [&](uint32_t s, uint32_t r, bool mine) {
                const uint32_t g = static_cast<uint32_t>(__builtin_amdgcn_readlane(gbase, s));
                if (!mine) return;
                const uint64_t j = ghost0 + g + r;
                ax = ax + fx[j];
                ay = ay + fy[j];
                az = az + fz[j];
            });
            if (valid) {
                fx[i] = ax;
                fy[i] = ay;
                fz[i] = az;
            }
        }
    }
}

// The stage of reverse_chunks_kernel: a unit of at most kRunOwners owners
// (kRunTiles tiles of 64) and kReverseStage images.
constexpr uint32_t kReverseRounds = 2;
constexpr uint32_t kReverseStage = kReverseRounds * kRunWaves * kWave;

struct ChunkReverseStage {
    double gx[kReverseStage], gy[kReverseStage], gz[kReverseStage];
    uint32_t entry[kRunOwners];   // the unit's entries in owner order
    uint32_t at[kRunOwners];      // their entry indices (chunk * chunk_size + rank)
    uint32_t tile_count[kRunTiles][kSlots];   // owners of each tile with each slot
};

// reverse_halo over chunk units: the unit's ghost forces come into LDS as
// dense runs; each wave copies its chunk's entries into LDS in owner order,
// and after a barrier wave t takes owners 64 t .. 64 t + 63 of the unit (one
// full tile instead of a chunk and the few owners past its first 64), counts
// them per slot with ballots, and after a second barrier adds each owner's
// images in ascending send slot, the order of the per-owner kernel.  The
// loads go out in this order: the chunk's entries, the ghost forces, then,
// after the first barrier, the owner forces, which stay in flight while the
// ghost forces are written to LDS (a wave waits for its loads in issue
// order).  Units that exceed the stage: reverse_chunk_walk.
__global__ __launch_bounds__(kRunWaves * kWave) void reverse_chunks_kernel(double* __restrict__ fx,
                                                                           double* __restrict__ fy,
                                                                           double* __restrict__ fz,
                                                                           const ReplicaDesc* __restrict__ desc,
                                                                           Layout L, PlanBuffers B)
{
    __shared__ ChunkReverseStage sh;
    const uint32_t units = chunk_units_of(L);
    const uint32_t b = xcd_block(L.xcd_run);
    if (b >= units) return;
    const uint32_t unit = L.descending ? units - 1u - b : b;
    const ReplicaDesc* d = desc;
    const uint32_t lane = lane_id();
    const uint32_t w = uniform(threadIdx.x / kWave);
    const ChunkUnit u = chunk_unit(d, unit, L, B, B.chunk_rows);
    const uint64_t ghost0 = d->offset + d->owned;   // every slot sends to the box itself
    if (u.overflow) return;
    const uint32_t lut = d->bp.lut[lane];
    if (u.owners > kRunOwners || u.runs.images > kReverseStage) {
        reverse_chunk_walk(fx, fy, fz, d, u, ghost0, lut, L, B);
        return;
    }
    const uint32_t K = L.chunk_size;
    // the wave's chunk (a wave past the unit's chunks reads the last one)
    const uint32_t c = u.c0 + w;
    ChunkHead h = chunk_owner_head(min(c, u.c1 - 1u), L, B);
    if (c >= u.c1) h.n = 0;
    // the ghost forces, unconditionally (an idle lane reads the box's first
    // particle), so that the compiler counts the loads exactly
    double gx[kReverseRounds], gy[kReverseRounds], gz[kReverseRounds];
#pragma unroll
    for (uint32_t k = 0; k < kReverseRounds; ++k) {
        const uint32_t r = k * kRunWaves * kWave + threadIdx.x;
        const uint32_t s = run_slot(u.runs.start, r);
        const uint32_t rel = lane_value(u.runs.first, s) + (r - lane_value(u.runs.start, s));
        const uint64_t j = r < u.runs.images ? ghost0 + rel : d->offset;
        gx[k] = fx[j];
        gy[k] = fy[j];
        gz[k] = fz[j];
    }
    __builtin_amdgcn_sched_barrier(0);
    // the unit's entries into LDS in owner order (a unit of more chunks than
    // waves, or a chunk of more than kChunkRounds * 64 owners, loads the rest)
    for (uint32_t cc = c; cc < u.c1; cc += kRunWaves) {
        if (cc != c) h = chunk_owner_head(cc, L, B);
        const uint32_t o0 = h.pc - u.owner0;
        const uint64_t ebase = static_cast<uint64_t>(cc) * K;
        for (uint32_t e0 = 0; e0 < h.n; e0 += kWave) {
            const uint32_t e = e0 + lane;
            if (e < h.n) {
                sh.entry[o0 + e] = e0 == 0 ? h.entry[0] : e0 == kWave ? h.entry[1] : B.entries[ebase + e];
                sh.at[o0 + e] = static_cast<uint32_t>(ebase + e);
            }
        }
    }
    __syncthreads();
    // tile w of the unit
    const uint32_t o = w * kWave + lane;
    const bool live = o < u.owners;
    const uint32_t entry = live ? sh.entry[o] : 0u;
    const uint32_t at = live ? sh.at[o] : u.c0 * K;
    const uint64_t i = d->offset + (at & ~(K - 1u)) - static_cast<uint64_t>(d->first_chunk) * K + (entry >> 8);
    double ax = fx[i], ay = fy[i], az = fz[i];
    const uint32_t code_slots = lane_value(lut, entry & 63u);
    const uint32_t slots = live ? code_slots : 0u;
    const uint32_t present = uniform(wave_or(slots));
    const uint32_t mine = tile_slot_count(slots, present);
    if (lane < kSlots) sh.tile_count[w][lane] = mine;
    __builtin_amdgcn_sched_barrier(0);
    // every stage slot is written, used or not: a write only where used would
    // let the compiler sink the ghost loads to it
#pragma unroll
    for (uint32_t k = 0; k < kReverseRounds; ++k) {
        const uint32_t r = k * kRunWaves * kWave + threadIdx.x;
        sh.gx[r] = gx[k];
        sh.gy[r] = gy[k];
        sh.gz[r] = gz[k];
    }
    __syncthreads();
    const uint32_t before = earlier_tiles(sh.tile_count, w);
    for (uint32_t rest = present; rest; rest &= rest - 1u) {
        const uint32_t s = static_cast<uint32_t>(__builtin_ctz(rest));
        const uint64_t bl = __ballot((slots >> s) & 1u);
        if ((slots >> s) & 1u) {
            const uint32_t r = static_cast<uint32_t>(__builtin_amdgcn_readlane(u.runs.start, s)) +
                               static_cast<uint32_t>(__builtin_amdgcn_readlane(before, s)) + lanes_below(bl);
            ax = ax + sh.gx[r];
            ay = ay + sh.gy[r];
            az = az + sh.gz[r];
        }
    }
    if (live) {
        fx[i] = ax;
        fy[i] = ay;
        fz[i] = az;
    }
}

// Up to C components at once: one pass over the images, each image's
// components read together, those of the first kPreloaded images all before
// the first reduction; every component still accumulates its images in
// ascending send slot.
template <class T, int C>
__device__ __forceinline__ void reduce_components(T* data, int32_t op, uint64_t i, const Images& im)
{
    T acc[C];
#pragma unroll
    for (int e = 0; e < C; ++e) acc[e] = data[i * C + e];
    T v[kPreloaded][C];
#pragma unroll
    for (uint32_t k = 0; k < kPreloaded; ++k) {
        const T* src = data + im.first(k) * C;
#pragma unroll
        for (int e = 0; e < C; ++e) v[k][e] = src[e];
    }
#pragma unroll
    for (uint32_t k = 0; k < kPreloaded; ++k) {
        if (k < im.f.n) {
#pragma unroll
            for (int e = 0; e < C; ++e) acc[e] = reduce_value<T>(acc[e], v[k][e], op);
        }
    }
    uint32_t k = kPreloaded;
    for (uint32_t rest = im.f.rest; rest; rest &= rest - 1u, ++k) {
        const T* src = data + im.later(k, static_cast<uint32_t>(__builtin_ctz(rest))) * C;
#pragma unroll
        for (int e = 0; e < C; ++e) acc[e] = reduce_value<T>(acc[e], src[e], op);
    }
#pragma unroll
    for (int e = 0; e < C; ++e) data[i * C + e] = acc[e];
}

template <class T>
__device__ __forceinline__ void reduce_field(T* data, uint32_t c, int32_t op, uint64_t i, const Images& im)
{
    switch (c) {
    case 1: reduce_components<T, 1>(data, op, i, im); return;
    case 2: reduce_components<T, 2>(data, op, i, im); return;
    case 3: reduce_components<T, 3>(data, op, i, im); return;
    case 4: reduce_components<T, 4>(data, op, i, im); return;
    default: break;
    }
    for (uint32_t e = 0; e < c; ++e) {
        T acc = data[i * c + e];
        uint32_t k = 0;
        for (uint32_t rest = im.g.slots; rest; rest &= rest - 1u, ++k)
            acc = reduce_value<T>(acc, data[im.later(k, static_cast<uint32_t>(__builtin_ctz(rest))) * c + e], op);
        data[i * c + e] = acc;
    }
}

// Integer sums wrap as unsigned (two's complement), like the Host's
// destination += value on int32/int64.
template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void reverse_fields_kernel(
    TypedFields tf, const ReplicaDesc* __restrict__ desc, const uint32_t* __restrict__ tile_replica,
    Layout L, PlanBuffers B)
{
    __shared__ SlotTable s_table[WAVES];
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        const SlotTable& t = fill_slot_table(s_table[threadIdx.x / kWave], desc, o.d);
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const OwnerGhosts g = ghosts_of(desc, o, B);
        const FirstImages f = first_images(g);
        const Images ghosts{g, f, t};
        for (uint32_t fi = 0; fi < tf.count; ++fi) {
            const uint32_t cc = tf.components[fi];
            const int32_t op = tf.op[fi];
            switch (tf.type[fi]) {
            case DSDX_INT32:
                if (op == DSDX_SUM)
                    reduce_field<uint32_t>(static_cast<uint32_t*>(tf.data[fi]), cc, op, i, ghosts);
                else
                    reduce_field<int32_t>(static_cast<int32_t*>(tf.data[fi]), cc, op, i, ghosts);
                break;
            case DSDX_INT64:
                if (op == DSDX_SUM)
                    reduce_field<uint64_t>(static_cast<uint64_t*>(tf.data[fi]), cc, op, i, ghosts);
                else
                    reduce_field<int64_t>(static_cast<int64_t*>(tf.data[fi]), cc, op, i, ghosts);
                break;
            case DSDX_UINT32:
                reduce_field<uint32_t>(static_cast<uint32_t*>(tf.data[fi]), cc, op, i, ghosts);
                break;
            case DSDX_UINT64:
                reduce_field<uint64_t>(static_cast<uint64_t*>(tf.data[fi]), cc, op, i, ghosts);
                break;
            case DSDX_FLOAT32:
                reduce_field<float>(static_cast<float*>(tf.data[fi]), cc, op, i, ghosts);
                break;
            default:
                reduce_field<double>(static_cast<double*>(tf.data[fi]), cc, op, i, ghosts);
                break;
            }
        }
    });
}


// A field of U units: the owner's units loaded once, then stored to every image.
template <class T, int U>
__device__ __forceinline__ void forward_record(T* dd, uint64_t i, const Images& im)
{
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = dd[i * U + u];
    each_image(im, [&](uint32_t, uint64_t j) {
#pragma unroll
        for (int u = 0; u < U; ++u) dd[j * U + u] = v[u];
    });
}

template <class T>
__device__ __forceinline__ void forward_units(T* dd, uint32_t units, uint64_t i, const Images& im)
{
    switch (units) {
    case 1: forward_record<T, 1>(dd, i, im); return;
    case 2: forward_record<T, 2>(dd, i, im); return;
    case 3: forward_record<T, 3>(dd, i, im); return;
    case 4: forward_record<T, 4>(dd, i, im); return;
    default: break;
    }
    for (uint32_t c = 0; c < units; ++c) {
        const T value = dd[i * units + c];
        each_image(im, [&](uint32_t, uint64_t j) { dd[j * units + c] = value; });
    }
}

// forward_exchange: owner values copied into each of its ghost slots.
template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void forward_fields_kernel(
    FieldParams fp, const ReplicaDesc* __restrict__ desc, const uint32_t* __restrict__ tile_replica,
    Layout L, PlanBuffers B)
{
    __shared__ SlotTable s_table[WAVES];
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        const SlotTable& t = fill_slot_table(s_table[threadIdx.x / kWave], desc, o.d);
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const OwnerGhosts g = ghosts_of(desc, o, B);
        const FirstImages f = first_images(g);
        const Images ghosts{g, f, t};
        for (uint32_t fi = 0; fi < fp.count; ++fi) {
            if ((fp.wide >> fi) & 1u)
                forward_units(static_cast<uint64_t*>(fp.data[fi]), fp.words[fi] / 2, i, ghosts);
            else
                forward_units(static_cast<uint32_t*>(fp.data[fi]), fp.words[fi], i, ghosts);
        }
    });
}

// Reused-plan refresh: each ghost re-reads its owner; the slot of the k-th
// ghost of an owner is the k-th set bit of its slot set.  The owner's values
// and its first kPreloaded ghost-list entries are loaded before its stores.
template <int WAVES, bool STREAM>
__global__ __launch_bounds__(WAVES * kWave) void refresh_kernel(
    dsdx_particles p, FieldParams fp, int32_t positions_only, const ReplicaDesc* __restrict__ desc,
    const uint32_t* __restrict__ tile_replica, Layout L, PlanBuffers B)
{
    __shared__ SlotTable s_table[WAVES];
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        const SlotTable& t = fill_slot_table(s_table[threadIdx.x / kWave], desc, o.d);
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const OwnerGhosts g = ghosts_of(desc, o, B);
        const double px = read_once<STREAM>(p.x + i), py = read_once<STREAM>(p.y + i), pz = read_once<STREAM>(p.z + i);
        double vx = 0.0, vy = 0.0, vz = 0.0, m = 0.0, q = 0.0;
        int32_t ty = 0, id = 0;
        if (!positions_only) {
            vx = read_once<STREAM>(p.vx + i); vy = read_once<STREAM>(p.vy + i); vz = read_once<STREAM>(p.vz + i);
            m = read_once<STREAM>(p.mass + i); q = read_once<STREAM>(p.charge + i);
            ty = read_once<STREAM>(p.type + i); id = read_once<STREAM>(p.global_id + i);
        }
        const FirstImages f = first_images(g);
        each_image(Images{g, f, t}, [&](uint32_t s, uint64_t j) {
            p.x[j] = px + t.shift[s][0];
            p.y[j] = py + t.shift[s][1];
            p.z[j] = pz + t.shift[s][2];
            if (p.fx) p.fx[j] = 0.0;
            if (p.fy) p.fy[j] = 0.0;
            if (p.fz) p.fz[j] = 0.0;
            if (!positions_only) {
                p.vx[j] = vx; p.vy[j] = vy; p.vz[j] = vz;
                p.mass[j] = m; p.charge[j] = q;
                p.type[j] = ty; p.global_id[j] = id;
                if (fp.count) copy_field_words(fp, i, j);
            }
        });
    });
}

// Owner of every ghost: out[offset_q + rel] = owned index within the
// sending replica, for ghost rel of receiving replica q.
template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void ghost_owners_kernel(
    int32_t* __restrict__ out, const ReplicaDesc* __restrict__ desc,
    const uint32_t* __restrict__ tile_replica, Layout L, PlanBuffers B)
{
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        if (!o.live) return;
        const uint32_t i = B.owner_index[o.u];
        const OwnerGhosts ghosts = ghosts_of(desc, o, B);
        uint32_t k = 0;
        for (uint32_t rest = ghosts.slots; rest; rest &= rest - 1u, ++k) {
            // ghost slot relative to the receiving replica's first owned slot
            const uint32_t s = static_cast<uint32_t>(__builtin_ctz(rest));
            const ReplicaDesc* to = desc + o.d->dest[s];
            out[ghosts.at(k, s) - to->owned] = static_cast<int32_t>(i);
        }
    });
}

// ---------------------------------------------------------------------------
// migrate at P = 1 (the wrap itself is defined above classify).  One wave per
// chunk, loads before compute; only a changed coordinate is stored.

// The wrap of one chunk: particles [first, first + 64 * ITEMS) of a box with
// `owned` particles whose arrays start at x, y, z.
template <int ITEMS>
__device__ __forceinline__ void wrap_chunk(double* x, double* y, double* z, uint32_t owned, uint32_t first,
                                           const int32_t (&periodic)[3], const double (&lower)[3],
                                           const double (&extent)[3])
{
    constexpr uint32_t K = ITEMS * kWave;
    const bool full = first + K <= owned;
    double* axis[3] = {x, y, z};
    double val[3][ITEMS];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (periodic[a]) load_items<ITEMS>(axis[a], first, owned, full, val[a]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!periodic[a]) continue;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t i = first + k * kWave + lane_id();
            const double w = wrap(val[a][k], lower[a], extent[a]);
            if (i < owned && __double_as_longlong(w) != __double_as_longlong(val[a][k])) axis[a][i] = w;
        }
    }
}

template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void migrate_wrap_kernel(
    double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    const ReplicaDesc* __restrict__ desc, const uint32_t* __restrict__ chunk_replica, Layout L)
{
    constexpr uint32_t K = ITEMS * kWave;
    const uint32_t chunk = wave_chunk<WAVES>(L, L.chunks);
    if (chunk >= L.chunks) return;
    const ReplicaDesc* d = desc + replica_of(chunk_replica, chunk);
    const uint64_t off = d->offset;
    wrap_chunk<ITEMS>(x + off, y + off, z + off, d->owned, (chunk - d->first_chunk) * K, d->bp.periodic,
                      d->bp.lower, d->bp.extent);
}

// The same wrap for one box whose geometry comes in as a kernel argument:
// no replica table, so the call needs no device memory of its own.
template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void wrap_box_kernel(
    double* __restrict__ x, double* __restrict__ y, double* __restrict__ z, WrapBox b, Layout L)
{
    constexpr uint32_t K = ITEMS * kWave;
    const uint32_t chunk = wave_chunk<WAVES>(L, L.chunks);
    if (chunk >= L.chunks) return;
    wrap_chunk<ITEMS>(x, y, z, b.owned, chunk * K, b.periodic, b.lower, b.extent);
}

// One replica descriptor into device memory, by one wave in stream order:
// the launch carries the descriptor, so the host copy may change as soon as
// the launch returns, no host wait is needed, and the store can be captured
// into a graph.
__global__ __launch_bounds__(kWave) void store_desc_kernel(ReplicaDesc d, ReplicaDesc* __restrict__ out)
{
    static_assert(sizeof(ReplicaDesc) % 8 == 0, "descriptor copied in 8-byte words");
    constexpr uint32_t kWords = sizeof(ReplicaDesc) / 8;
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&d);
    uint64_t* dst = reinterpret_cast<uint64_t*>(out);
    for (uint32_t w = lane_id(); w < kWords; w += kWave) dst[w] = src[w];
}

}  // namespace

bool shape_supported(const Shape& s)
{
    return (s.items == 4 || s.items == 8 || s.items == 16) &&
           (s.waves == 2 || s.waves == 4 || s.waves == 8) &&
           (s.scan_threads == 256 || s.scan_threads == 512 || s.scan_threads == 1024) &&
           s.unit_owners >= 0 && s.unit_owners <= 1024;
}

// ---------------------------------------------------------------------------
// Launchers (called by the C ABI in abi.cpp).

hipError_t launch_classify(double* x, double* y, double* z, bool wrap_first, const Layout& L,
                           const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    const Layout O = ordered(L, s, wrap_first ? kOrderFused : kOrderClassify);
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        return dispatch_stream(L.streaming, [&](auto streaming) {
            constexpr int I = decltype(items)::value, W = decltype(waves)::value;
            constexpr bool S = decltype(streaming)::value;
            // The fused pass takes a chunk in rounds of at most 8 items: with
            // the wrap's stores it holds fewer registers and runs more waves
            // (16 M particles: 80 -> 74.7 us; rounds of 4 measured the same).
            // The plain pass measured the same either way and keeps one round.
            constexpr int kWrapRound = I < 8 ? I : 8;
            if (wrap_first)
                return launch(classify_kernel<I, W, true, S, kWrapRound>, chunk_grid(L, W), dim3(W * kWave),
                              stream, x, y, z, L.desc, L.chunk_replica, O, B);
            return launch(classify_kernel<I, W, false, S, I>, chunk_grid(L, W), dim3(W * kWave),
                          stream, x, y, z, L.desc, L.chunk_replica, O, B);
        });
    });
}

hipError_t launch_scan(const Layout& L, const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    if (L.replicas == 0) return hipSuccess;
    if (L.max_replica_chunks <= kShortScanChunks) {
        const uint32_t pairs = static_cast<uint32_t>(kCols) * L.replicas;
        return launch(scan_short_kernel, dim3((pairs + kShortScanWaves - 1) / kShortScanWaves),
                      dim3(kShortScanWaves * kWave), stream, L.desc, L, B);
    }
    // Up to kMaxScanParts tiles per column: one block per tile, each reducing
    // the values before its tile itself (16 M particles: 9.5 -> 5.0 us).  Longer
    // columns: one block per column walking its tiles.
    const uint32_t span = static_cast<uint32_t>(s.scan_threads) * kScanPer;
    const uint32_t parts = std::max<uint32_t>(1, (L.max_replica_chunks + span - 1) / span);
    if (parts <= kMaxScanParts)
        return launch(scan_parts_kernel, dim3(kCols * parts, L.replicas), dim3(s.scan_threads), stream, L.desc,
                      L, B);
    return launch(scan_kernel, dim3(kCols, L.replicas), dim3(s.scan_threads), stream, L.desc, L, B);
}

hipError_t launch_route(const Layout& L, const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    const Layout O = ordered(L, s, kOrderRoute);
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        return launch(route_kernel<I, W>, chunk_grid(L, W), dim3(W * kWave), stream, L.desc,
                      L.chunk_replica, O, B);
    });
}

hipError_t launch_copy(const dsdx_particles& p, const FieldParams& fp, const Layout& L,
                       const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    const Layout O = ordered(L, s, kOrderCopy);
    if (L.tiles == 0) return hipSuccess;
    if (L.unit_chunks)
        return dispatch_stream(L.streaming, [&](auto streaming) {
            constexpr bool S = decltype(streaming)::value;
            return launch(copy_chunks_kernel<S>, dim3((L.chunks + L.unit_chunks - 1) / L.unit_chunks),
                          dim3(kRunWaves * kWave), stream, p, fp, L.desc, O, B);
        });
    // Each workgroup walks units g, g + grid, ... up to the device's owner count.
    if (ghost_runs(L, B))
        return dispatch_stream(L.streaming, [&](auto streaming) {
            constexpr bool S = decltype(streaming)::value;
            return launch(copy_runs_kernel<kRunsBuild, S>, run_grid(L), dim3(kRunWaves * kWave), stream, p, fp,
                          L.desc, L.unit_replica, O, B);
        });
    return dispatch_waves(s, [&](auto waves) {
        return dispatch_stream(L.streaming, [&](auto streaming) {
            constexpr int W = decltype(waves)::value;
            constexpr bool S = decltype(streaming)::value;
            return launch(copy_kernel<W, S>, tile_grid(L, W), dim3(W * kWave), stream, p, fp, L.desc,
                          L.tile_replica, O, B);
        });
    });
}

hipError_t launch_reverse_forces(double* fx, double* fy, double* fz, const Layout& L,
                                 const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    const Layout O = ordered(L, s, kOrderOwners);
    if (L.tiles == 0) return hipSuccess;
    if (chunk_reverse(L, s)) {
        Layout R = O;
        R.unit_chunks = L.reverse_unit_chunks;
        return launch(reverse_chunks_kernel, dim3((R.chunks + R.unit_chunks - 1) / R.unit_chunks),
                      dim3(kRunWaves * kWave), stream, fx, fy, fz, L.desc, R, B);
    }
    if (B.unit_rows && (s.reverse_runs < 0 ? ghost_runs(L, B) : s.reverse_runs > 0))
        return launch(reverse_runs_kernel, run_grid(L), dim3(kRunWaves * kWave), stream, fx, fy, fz, L.desc,
                      L.unit_replica, O, B);
    return dispatch_waves(s, [&](auto waves) {
        constexpr int W = decltype(waves)::value;
        return launch(reverse_forces_kernel<W>, tile_grid(L, W), dim3(W * kWave), stream, fx, fy, fz,
                      L.desc, L.tile_replica, O, B);
    });
}

hipError_t launch_reverse_fields(const TypedFields& tf, const Layout& L, const PlanBuffers& B,
                                 const Shape& s, hipStream_t stream)
{
    const Layout O = ordered(L, s, kOrderOwners);
    if (L.tiles == 0) return hipSuccess;
    return dispatch_waves(s, [&](auto waves) {
        constexpr int W = decltype(waves)::value;
        return launch(reverse_fields_kernel<W>, tile_grid(L, W), dim3(W * kWave), stream, tf,
                      L.desc, L.tile_replica, O, B);
    });
}

hipError_t launch_forward_fields(const FieldParams& fp, const Layout& L, const PlanBuffers& B,
                                 const Shape& s, hipStream_t stream)
{
    const Layout O = ordered(L, s, kOrderOwners);
    if (L.tiles == 0) return hipSuccess;
    return dispatch_waves(s, [&](auto waves) {
        constexpr int W = decltype(waves)::value;
        return launch(forward_fields_kernel<W>, tile_grid(L, W), dim3(W * kWave), stream, fp,
                      L.desc, L.tile_replica, O, B);
    });
}

hipError_t launch_refresh(const dsdx_particles& p, const FieldParams& fp, int32_t positions_only,
                          const Layout& L, const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    const Layout O = ordered(L, s, kOrderOwners);
    if (L.tiles == 0) return hipSuccess;
    if (ghost_runs(L, B))
        return dispatch_stream(L.streaming, [&](auto streaming) {
            constexpr bool S = decltype(streaming)::value;
            if (positions_only)
                return launch(copy_runs_kernel<kRunsPositions, S>, run_grid(L), dim3(kRunWaves * kWave), stream,
                              p, fp, L.desc, L.unit_replica, O, B);
            return launch(copy_runs_kernel<kRunsRefresh, S>, run_grid(L), dim3(kRunWaves * kWave), stream, p, fp,
                          L.desc, L.unit_replica, O, B);
        });
    return dispatch_waves(s, [&](auto waves) {
        return dispatch_stream(L.streaming, [&](auto streaming) {
            constexpr int W = decltype(waves)::value;
            constexpr bool S = decltype(streaming)::value;
            return launch(refresh_kernel<W, S>, tile_grid(L, W), dim3(W * kWave), stream, p, fp,
                          positions_only, L.desc, L.tile_replica, O, B);
        });
    });
}

hipError_t launch_ghost_owners(int32_t* out, const Layout& L, const PlanBuffers& B, const Shape& s,
                               hipStream_t stream)
{
    const Layout O = ordered(L, s, kOrderOwners);
    if (L.tiles == 0) return hipSuccess;
    return dispatch_waves(s, [&](auto waves) {
        constexpr int W = decltype(waves)::value;
        return launch(ghost_owners_kernel<W>, tile_grid(L, W), dim3(W * kWave), stream, out, L.desc,
                      L.tile_replica, O, B);
    });
}

hipError_t launch_migrate_wrap(double* x, double* y, double* z, const Layout& L, const Shape& s,
                               hipStream_t stream)
{
    const Layout O = ordered(L, s, kOrderMigrate);
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        return launch(migrate_wrap_kernel<I, W>, chunk_grid(L, W), dim3(W * kWave), stream,
                      x, y, z, L.desc, L.chunk_replica, O);
    });
}

hipError_t launch_wrap_box(double* x, double* y, double* z, const WrapBox& b, const Shape& s,
                           hipStream_t stream)
{
    const uint32_t K = static_cast<uint32_t>(s.items) * kWave;
    Layout L{};
    L.chunks = (b.owned + K - 1) / K;
    L.chunk_size = K;
    L.xcd_run = static_cast<uint32_t>(s.xcd_run);
    const Layout O = ordered(L, s, kOrderMigrate);
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        return launch(wrap_box_kernel<I, W>, chunk_grid(L, W), dim3(W * kWave), stream, x, y, z, b, O);
    });
}

hipError_t launch_store_desc(const ReplicaDesc& d, ReplicaDesc* out, hipStream_t stream)
{
    return launch(store_desc_kernel, dim3(1), dim3(kWave), stream, d, out);
}

}  // namespace dsdx
