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

namespace dsdx {
namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, kWave);
        if (lane >= static_cast<uint32_t>(d)) v += o;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v)
{
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) v |= __shfl_xor(v, d, kWave);
    return v;
}

// The chunk of this wave and its replica (both wave-uniform).
template <int WAVES>
__device__ __forceinline__ uint32_t wave_chunk()
{
    return uniform(blockIdx.x * WAVES + threadIdx.x / kWave);
}

__device__ __forceinline__ uint32_t replica_of(const Layout& L, uint32_t chunk)
{
    return L.chunk_replica ? uniform(L.chunk_replica[chunk]) : 0u;
}

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

// Periodic wrap of migrate at P = 1 (src/migration.cpp:106-113):
// v = x - lower; v -= floor(v / extent) * extent; v + lower, each operation
// rounded on its own (explicit _rn intrinsics, never fused).
__device__ __forceinline__ double wrap(double value, double lower, double extent)
{
    const double v = __dsub_rn(value, lower);
    const double q = floor(__ddiv_rn(v, extent));
    return __dadd_rn(__dsub_rn(v, __dmul_rn(q, extent)), lower);
}

// Wraps the periodic axes of one lane's value and stores it back when it
// changed (used by the fused migrate + classify pass).
__device__ __forceinline__ void wrap_store(double& v, double* dst, bool valid, int periodic,
                                           double lower, double extent)
{
    if (!periodic) return;
    const double w = wrap(v, lower, extent);
    if (valid && __double_as_longlong(w) != __double_as_longlong(v)) *dst = w;
    v = w;
}

// ---------------------------------------------------------------------------
// Pass 1: one wave per chunk.  Slab-classifies every owned particle,
// compacts the boundary ones (non-empty slot set) into the chunk's entry
// region in owned order, and counts per send slot through an LDS histogram of
// slab codes (one LDS atomic per boundary particle).  Reads 24 B per
// particle, writes 4 B per boundary particle and 28 counters per chunk.
// With WRAP the pass first applies migrate's periodic wrap to the positions
// it loaded (storing the changed ones) and classifies the wrapped values:
// migrate followed by a rebuild with one read of x, y, z.
template <int ITEMS, int WAVES, bool WRAP>
__global__ __launch_bounds__(WAVES * kWave) void classify_kernel(
    double* __restrict__ x, double* __restrict__ y, double* __restrict__ z, Layout L,
    PlanBuffers B)
{
    constexpr uint32_t K = ITEMS * kWave;
    __shared__ uint32_t s_lut[WAVES][64];
    __shared__ uint32_t s_hist[WAVES][64];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t chunk = wave_chunk<WAVES>();
    const bool live = chunk < L.chunks;
    const ReplicaDesc* d = nullptr;
    if (live) {
        d = L.desc + replica_of(L, chunk);
        s_lut[wave][lane] = d->bp.lut[lane];
    }
    s_hist[wave][lane] = 0u;
    __syncthreads();

    uint32_t running = 0;
    if (live) {
        const uint32_t owned = d->owned;
        const uint64_t off = d->offset;
        const uint32_t first = (chunk - d->first_chunk) * K;
        double px[ITEMS], py[ITEMS], pz[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t i = first + k * kWave + lane;
            const bool valid = i < owned;
            px[k] = valid ? x[off + i] : 0.0;
            py[k] = valid ? y[off + i] : 0.0;
            pz[k] = valid ? z[off + i] : 0.0;
        }
        if (WRAP) {
#pragma unroll
            for (int k = 0; k < ITEMS; ++k) {
                const uint32_t i = first + k * kWave + lane;
                const bool valid = i < owned;
                wrap_store(px[k], x + off + i, valid, d->bp.periodic[0], d->bp.lower[0], d->bp.extent[0]);
                wrap_store(py[k], y + off + i, valid, d->bp.periodic[1], d->bp.lower[1], d->bp.extent[1]);
                wrap_store(pz[k], z + off + i, valid, d->bp.periodic[2], d->bp.lower[2], d->bp.extent[2]);
            }
        }
        const uint64_t ebase = static_cast<uint64_t>(chunk) * K;
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t i = first + k * kWave + lane;
            const uint32_t code = i < owned ? slab_code(px[k], py[k], pz[k], d->bp) : 0u;
            const bool boundary = s_lut[wave][code] != 0u;
            const uint64_t ballot = __ballot(boundary);
            if (boundary) {
                B.entries[ebase + running + lanes_below(ballot)] =
                    ((static_cast<uint32_t>(k) * kWave + lane) << 8) | code;
                atomicAdd(&s_hist[wave][code], 1u);
            }
            running += static_cast<uint32_t>(__popcll(ballot));
        }
    }
    __syncthreads();
    if (live) {
        // lane s < 26: send count of slot s; lane 26: owners; lane 27: images
        uint32_t acc = 0;
        for (int c = 1; c < 64; ++c) {
            const uint32_t h = s_hist[wave][c];
            const uint32_t m = s_lut[wave][c];
            if (lane < kSlots) acc += ((m >> lane) & 1u) ? h : 0u;
            else if (lane == kColImages) acc += h * static_cast<uint32_t>(__popc(m));
        }
        if (lane == kColOwners) acc = running;
        if (lane < kCols) B.counts[static_cast<uint64_t>(lane) * L.chunks + chunk] = acc;
    }
}

// ---------------------------------------------------------------------------
// Pass 2: exclusive scan of one count column over one replica's chunks; grid
// (kCols, replicas).  totals[r][col] receives the column total.
constexpr int kMaxScanThreads = 1024;

__global__ __launch_bounds__(kMaxScanThreads) void scan_kernel(Layout L, PlanBuffers B)
{
    __shared__ uint32_t s_wave[kMaxScanThreads / kWave];
    const uint32_t col = blockIdx.x;
    const uint32_t r = blockIdx.y;
    const ReplicaDesc* d = L.desc + r;
    const uint32_t n = d->chunks;
    const uint32_t threads = blockDim.x;
    const uint32_t waves = threads / kWave;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + threads - 1) / threads;
    const uint32_t begin = min(t * per, n);
    const uint32_t end = min(begin + per, n);
    const uint64_t base = static_cast<uint64_t>(col) * L.chunks + d->first_chunk;
    const uint32_t* src = B.counts + base;
    uint32_t* dst = B.prefix + base;
    uint32_t sum = 0;
    for (uint32_t k = begin; k < end; ++k) sum += src[k];
    const uint32_t incl = wave_inclusive_sum(sum);
    const uint32_t wave = t / kWave;
    if (lane_id() == kWave - 1) s_wave[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        const uint32_t w = lane_id() < waves ? s_wave[lane_id()] : 0u;
        const uint32_t wi = wave_inclusive_sum(w);
        if (lane_id() < waves) s_wave[lane_id()] = wi - w;
        if (lane_id() == waves - 1) B.totals[static_cast<uint64_t>(r) * kTotals + col] = wi;
    }
    __syncthreads();
    uint32_t run = s_wave[wave] + incl - sum;
    for (uint32_t k = begin; k < end; ++k) {
        const uint32_t v = src[k];
        dst[k] = run;
        run += v;
    }
    if (col == 0 && t == 0) B.totals[static_cast<uint64_t>(r) * kTotals + kStatusWord] = 0u;
}

// ---------------------------------------------------------------------------
// Pass 3: one wave per chunk.  Walks the chunk's boundary particles in owned
// order; for each of its send slots s (ascending) a particle gets the ghost
// at receive block 25-s, position prefix[s][chunk] + rank in the chunk.
// Writes the 13 ghost arrays, the registered fields and the owner -> ghost
// map.  The capacity test is uniform over a replica, so an overflowing
// replica leaves every one of its ghost slots untouched.
__device__ __forceinline__ void copy_field_words(const FieldParams& fp, uint64_t from, uint64_t to)
{
    for (uint32_t f = 0; f < fp.count; ++f) {
        const uint32_t w = fp.words[f];
        uint32_t* dd = static_cast<uint32_t*>(fp.data[f]);
        for (uint32_t k = 0; k < w; ++k) dd[to * w + k] = dd[from * w + k];
    }
}

template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void scatter_kernel(dsdx_particles p, FieldParams fp,
                                                                Layout L, PlanBuffers B)
{
    constexpr uint32_t K = ITEMS * kWave;
    __shared__ uint32_t s_lut[WAVES][64];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t chunk = wave_chunk<WAVES>();
    if (chunk >= L.chunks) return;
    const uint32_t r = replica_of(L, chunk);
    const ReplicaDesc* d = L.desc + r;
    uint32_t* totals = B.totals + static_cast<uint64_t>(r) * kTotals;
    s_lut[wave][lane] = d->bp.lut[lane];

    const uint32_t owned = d->owned;
    const uint32_t ghosts = uniform(totals[kColImages]);
    uint32_t overflow = 0;
    if (static_cast<uint64_t>(owned) + ghosts > d->capacity) overflow |= kOverflowStorage;
    if (ghosts > d->ghost_list_capacity) overflow |= kOverflowGhostList;
    if (overflow) {
        if (chunk == d->first_chunk && lane == 0) totals[kStatusWord] = overflow;
        return;
    }
    __builtin_amdgcn_wave_barrier();

    // Receive block m holds send slot 25-m: recv_base[m] = sum of totals of
    // slots 25..26-m.  Lane s ends with slot_base = recv_base[25-s] + prefix.
    const uint32_t block_size = lane < kSlots ? totals[kSlots - 1 - lane] : 0u;
    const uint32_t recv_base = wave_inclusive_sum(block_size) - block_size;
    const uint32_t mirrored = __shfl(recv_base, lane < kSlots ? kSlots - 1 - lane : 0, kWave);
    const uint32_t slot_base =
        lane < kSlots ? mirrored + B.prefix[static_cast<uint64_t>(lane) * L.chunks + chunk] : 0u;
    uint32_t slot_run = 0;   // lane s: ghosts of slot s already written from this chunk

    const uint32_t n_boundary = uniform(B.counts[static_cast<uint64_t>(kColOwners) * L.chunks + chunk]);
    const uint32_t owner_base = uniform(B.prefix[static_cast<uint64_t>(kColOwners) * L.chunks + chunk]);
    uint32_t image_run = uniform(B.prefix[static_cast<uint64_t>(kColImages) * L.chunks + chunk]);
    const uint64_t off = d->offset;
    const uint32_t first = (chunk - d->first_chunk) * K;
    const uint64_t ebase = static_cast<uint64_t>(chunk) * K;
    const uint64_t slot0 = static_cast<uint64_t>(d->first_chunk) * K + owner_base;
    uint32_t* glist = B.ghost_list + d->ghost_list_offset;

    for (uint32_t e0 = 0; e0 < n_boundary; e0 += kWave) {
        const uint32_t e = e0 + lane;
        const bool valid = e < n_boundary;
        const uint32_t entry = valid ? B.entries[ebase + e] : 0u;
        const uint32_t code = entry & 63u;
        const uint32_t il = first + (entry >> 8);
        const uint64_t i = off + il;
        const uint32_t slots = valid ? s_lut[wave][code] : 0u;
        const uint32_t images = static_cast<uint32_t>(__popc(slots));
        const uint32_t incl = wave_inclusive_sum(images);
        const uint32_t first_image = image_run + incl - images;
        image_run += uniform(__shfl(incl, kWave - 1, kWave));

        double px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0, m = 0, q = 0;
        int32_t ty = 0, id = 0;
        if (valid) {
            const uint64_t u = slot0 + e;
            B.owner_index[u] = il;
            B.owner_code[u] = static_cast<uint8_t>(code);
            B.image_start[u] = first_image;
            px = p.x[i]; py = p.y[i]; pz = p.z[i];
            vx = p.vx[i]; vy = p.vy[i]; vz = p.vz[i];
            m = p.mass[i]; q = p.charge[i];
            ty = p.type[i]; id = p.global_id[i];
        }
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
                const uint64_t j = off + owned + rel;
                p.x[j] = px + d->bp.shift[s][0];
                p.y[j] = py + d->bp.shift[s][1];
                p.z[j] = pz + d->bp.shift[s][2];
                p.vx[j] = vx; p.vy[j] = vy; p.vz[j] = vz;
                if (p.fx) p.fx[j] = 0.0;
                if (p.fy) p.fy[j] = 0.0;
                if (p.fz) p.fz[j] = 0.0;
                p.mass[j] = m; p.charge[j] = q;
                p.type[j] = ty; p.global_id[j] = id;
                if (fp.count) copy_field_words(fp, i, j);
                glist[first_image + written] = rel;
                ++written;
            }
            if (lane == s) slot_run += static_cast<uint32_t>(__popcll(ballot));
        }
    }
}

// ---------------------------------------------------------------------------
// Owner-major kernels: one wave per chunk, one lane per boundary owner found
// in that chunk.  A replica whose last build overflowed has no owner map and
// is skipped.
struct OwnerCursor {
    const ReplicaDesc* d;
    uint64_t slot0;      // owner slot of the chunk's first owner
    uint32_t count;      // owners found in the chunk
    bool live;
};

template <int ITEMS, int WAVES>
__device__ __forceinline__ OwnerCursor owners_of_wave(const Layout& L, const PlanBuffers& B)
{
    constexpr uint32_t K = ITEMS * kWave;
    OwnerCursor c{nullptr, 0, 0, false};
    const uint32_t chunk = wave_chunk<WAVES>();
    if (chunk >= L.chunks) return c;
    const uint32_t r = replica_of(L, chunk);
    c.d = L.desc + r;
    if (uniform(B.totals[static_cast<uint64_t>(r) * kTotals + kStatusWord])) return c;
    c.count = uniform(B.counts[static_cast<uint64_t>(kColOwners) * L.chunks + chunk]);
    c.slot0 = static_cast<uint64_t>(c.d->first_chunk) * K +
              uniform(B.prefix[static_cast<uint64_t>(kColOwners) * L.chunks + chunk]);
    c.live = true;
    return c;
}

// reverse_halo: f[i] = ((f[i] + g_s1) + g_s2) + ... over the owner's send
// slots s1 < s2 < ..., the order of the Host reduction (src/halo.cpp:1357-1370
// for n ascending).  No atomics: one lane sums one owner.
template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void reverse_forces_kernel(
    double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz, Layout L,
    PlanBuffers B)
{
    const OwnerCursor c = owners_of_wave<ITEMS, WAVES>(L, B);
    if (!c.live) return;
    const uint64_t off = c.d->offset;
    const uint64_t goff = off + c.d->owned;
    const uint32_t* glist = B.ghost_list + c.d->ghost_list_offset;
    for (uint32_t e = lane_id(); e < c.count; e += kWave) {
        const uint64_t u = c.slot0 + e;
        const uint64_t i = off + B.owner_index[u];
        const uint32_t first = B.image_start[u];
        const uint32_t n_img = static_cast<uint32_t>(__popc(c.d->bp.lut[B.owner_code[u]]));
        double ax = fx[i], ay = fy[i], az = fz[i];
        for (uint32_t k = 0; k < n_img; ++k) {
            const uint64_t j = goff + glist[first + k];
            ax = ax + fx[j];
            ay = ay + fy[j];
            az = az + fz[j];
        }
        fx[i] = ax; fy[i] = ay; fz[i] = az;
    }
}

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

template <class T>
__device__ void reduce_field(T* data, uint32_t c, int32_t op, uint64_t i, uint64_t goff,
                             const uint32_t* ghosts, uint32_t n_img)
{
    for (uint32_t e = 0; e < c; ++e) {
        T acc = data[i * c + e];
        for (uint32_t k = 0; k < n_img; ++k) acc = reduce_value<T>(acc, data[(goff + ghosts[k]) * c + e], op);
        data[i * c + e] = acc;
    }
}

// Integer sums wrap as unsigned (two's complement), like the Host's
// destination += value on int32/int64.
template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void reverse_fields_kernel(TypedFields tf, Layout L,
                                                                       PlanBuffers B)
{
    const OwnerCursor c = owners_of_wave<ITEMS, WAVES>(L, B);
    if (!c.live) return;
    const uint64_t off = c.d->offset;
    const uint64_t goff = off + c.d->owned;
    const uint32_t* glist = B.ghost_list + c.d->ghost_list_offset;
    for (uint32_t e = lane_id(); e < c.count; e += kWave) {
        const uint64_t u = c.slot0 + e;
        const uint64_t i = off + B.owner_index[u];
        const uint32_t* ghosts = glist + B.image_start[u];
        const uint32_t n_img = static_cast<uint32_t>(__popc(c.d->bp.lut[B.owner_code[u]]));
        for (uint32_t f = 0; f < tf.count; ++f) {
            const uint32_t cc = tf.components[f];
            const int32_t op = tf.op[f];
            switch (tf.type[f]) {
            case DSDX_INT32:
                if (op == DSDX_SUM)
                    reduce_field<uint32_t>(static_cast<uint32_t*>(tf.data[f]), cc, op, i, goff, ghosts, n_img);
                else
                    reduce_field<int32_t>(static_cast<int32_t*>(tf.data[f]), cc, op, i, goff, ghosts, n_img);
                break;
            case DSDX_INT64:
                if (op == DSDX_SUM)
                    reduce_field<uint64_t>(static_cast<uint64_t*>(tf.data[f]), cc, op, i, goff, ghosts, n_img);
                else
                    reduce_field<int64_t>(static_cast<int64_t*>(tf.data[f]), cc, op, i, goff, ghosts, n_img);
                break;
            case DSDX_UINT32:
                reduce_field<uint32_t>(static_cast<uint32_t*>(tf.data[f]), cc, op, i, goff, ghosts, n_img);
                break;
            case DSDX_UINT64:
                reduce_field<uint64_t>(static_cast<uint64_t*>(tf.data[f]), cc, op, i, goff, ghosts, n_img);
                break;
            case DSDX_FLOAT32:
                reduce_field<float>(static_cast<float*>(tf.data[f]), cc, op, i, goff, ghosts, n_img);
                break;
            default:
                reduce_field<double>(static_cast<double*>(tf.data[f]), cc, op, i, goff, ghosts, n_img);
                break;
            }
        }
    }
}

// forward_exchange: owner values copied into each of its ghost slots.
template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void forward_fields_kernel(FieldParams fp, Layout L,
                                                                       PlanBuffers B)
{
    const OwnerCursor c = owners_of_wave<ITEMS, WAVES>(L, B);
    if (!c.live) return;
    const uint64_t off = c.d->offset;
    const uint64_t goff = off + c.d->owned;
    const uint32_t* glist = B.ghost_list + c.d->ghost_list_offset;
    for (uint32_t e = lane_id(); e < c.count; e += kWave) {
        const uint64_t u = c.slot0 + e;
        const uint64_t i = off + B.owner_index[u];
        const uint32_t* ghosts = glist + B.image_start[u];
        const uint32_t n_img = static_cast<uint32_t>(__popc(c.d->bp.lut[B.owner_code[u]]));
        for (uint32_t f = 0; f < fp.count; ++f) {
            const uint32_t w = fp.words[f];
            uint32_t* dd = static_cast<uint32_t*>(fp.data[f]);
            for (uint32_t k = 0; k < w; ++k) {
                const uint32_t value = dd[i * w + k];
                for (uint32_t g = 0; g < n_img; ++g) dd[(goff + ghosts[g]) * w + k] = value;
            }
        }
    }
}

// Reused-plan refresh: each ghost re-reads its owner; the slot of the k-th
// ghost of an owner is the k-th set bit of its slot set.
template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void refresh_kernel(dsdx_particles p, FieldParams fp,
                                                                int32_t positions_only, Layout L,
                                                                PlanBuffers B)
{
    const OwnerCursor c = owners_of_wave<ITEMS, WAVES>(L, B);
    if (!c.live) return;
    const uint64_t off = c.d->offset;
    const uint64_t goff = off + c.d->owned;
    const uint32_t* glist = B.ghost_list + c.d->ghost_list_offset;
    for (uint32_t e = lane_id(); e < c.count; e += kWave) {
        const uint64_t u = c.slot0 + e;
        const uint64_t i = off + B.owner_index[u];
        const uint32_t* ghosts = glist + B.image_start[u];
        uint32_t slots = c.d->bp.lut[B.owner_code[u]];
        const double px = p.x[i], py = p.y[i], pz = p.z[i];
        for (uint32_t k = 0; slots; ++k) {
            const uint32_t s = static_cast<uint32_t>(__builtin_ctz(slots));
            slots &= slots - 1u;
            const uint64_t j = goff + ghosts[k];
            p.x[j] = px + c.d->bp.shift[s][0];
            p.y[j] = py + c.d->bp.shift[s][1];
            p.z[j] = pz + c.d->bp.shift[s][2];
            if (p.fx) p.fx[j] = 0.0;
            if (p.fy) p.fy[j] = 0.0;
            if (p.fz) p.fz[j] = 0.0;
            if (!positions_only) {
                p.vx[j] = p.vx[i]; p.vy[j] = p.vy[i]; p.vz[j] = p.vz[i];
                p.mass[j] = p.mass[i]; p.charge[j] = p.charge[i];
                p.type[j] = p.type[i]; p.global_id[j] = p.global_id[i];
                if (fp.count) copy_field_words(fp, i, j);
            }
        }
    }
}

// Owner of every ghost: out[offset_r + rel] = owned index (replica-local).
template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void ghost_owners_kernel(int32_t* __restrict__ out,
                                                                     Layout L, PlanBuffers B)
{
    const OwnerCursor c = owners_of_wave<ITEMS, WAVES>(L, B);
    if (!c.live) return;
    const uint32_t* glist = B.ghost_list + c.d->ghost_list_offset;
    for (uint32_t e = lane_id(); e < c.count; e += kWave) {
        const uint64_t u = c.slot0 + e;
        const uint32_t i = B.owner_index[u];
        const uint32_t* ghosts = glist + B.image_start[u];
        const uint32_t n_img = static_cast<uint32_t>(__popc(c.d->bp.lut[B.owner_code[u]]));
        for (uint32_t k = 0; k < n_img; ++k) out[c.d->offset + ghosts[k]] = static_cast<int32_t>(i);
    }
}

// ---------------------------------------------------------------------------
// migrate at P = 1 (the wrap itself is defined above classify).  One wave per
// chunk, loads before compute; only a changed coordinate is stored.

template <int ITEMS, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void migrate_wrap_kernel(
    double* __restrict__ x, double* __restrict__ y, double* __restrict__ z, Layout L)
{
    constexpr uint32_t K = ITEMS * kWave;
    const uint32_t chunk = wave_chunk<WAVES>();
    if (chunk >= L.chunks) return;
    const ReplicaDesc* d = L.desc + replica_of(L, chunk);
    const uint32_t owned = d->owned;
    const uint64_t off = d->offset;
    const uint32_t first = (chunk - d->first_chunk) * K;
    double* axis[3] = {x, y, z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!d->bp.periodic[a]) continue;
        double* v = axis[a] + off;
        const double lower = d->bp.lower[a];
        const double extent = d->bp.extent[a];
        double val[ITEMS];
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t i = first + k * kWave + lane_id();
            val[k] = i < owned ? v[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < ITEMS; ++k) {
            const uint32_t i = first + k * kWave + lane_id();
            const double w = wrap(val[k], lower, extent);
            if (i < owned && __double_as_longlong(w) != __double_as_longlong(val[k])) v[i] = w;
        }
    }
}

// ---------------------------------------------------------------------------
// Shape dispatch: items per lane 4 / 8 / 16, waves per block 2 / 4 / 8.
template <class F>
hipError_t dispatch(const Shape& s, F&& f)
{
    auto by_waves = [&](auto items) -> hipError_t {
        switch (s.waves) {
        case 2: return f(items, std::integral_constant<int, 2>{});
        case 4: return f(items, std::integral_constant<int, 4>{});
        case 8: return f(items, std::integral_constant<int, 8>{});
        default: return hipErrorInvalidValue;
        }
    };
    switch (s.items) {
    case 4: return by_waves(std::integral_constant<int, 4>{});
    case 8: return by_waves(std::integral_constant<int, 8>{});
    case 16: return by_waves(std::integral_constant<int, 16>{});
    default: return hipErrorInvalidValue;
    }
}

uint32_t wave_blocks(uint32_t chunks, int waves)
{
    return std::max<uint32_t>(1, (chunks + waves - 1) / waves);
}

}  // namespace

bool shape_supported(const Shape& s)
{
    return (s.items == 4 || s.items == 8 || s.items == 16) &&
           (s.waves == 2 || s.waves == 4 || s.waves == 8) &&
           (s.scan_threads == 256 || s.scan_threads == 512 || s.scan_threads == 1024);
}

// ---------------------------------------------------------------------------
// Launchers (called by the C ABI in abi.cpp).
hipError_t launch_classify(double* x, double* y, double* z, bool wrap_first, const Layout& L,
                           const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        if (wrap_first)
            hipLaunchKernelGGL((classify_kernel<I, W, true>), dim3(wave_blocks(L.chunks, W)),
                               dim3(W * kWave), 0, stream, x, y, z, L, B);
        else
            hipLaunchKernelGGL((classify_kernel<I, W, false>), dim3(wave_blocks(L.chunks, W)),
                               dim3(W * kWave), 0, stream, x, y, z, L, B);
        return hipGetLastError();
    });
}

hipError_t launch_scan(const Layout& L, const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    if (L.replicas == 0) return hipSuccess;
    hipLaunchKernelGGL(scan_kernel, dim3(kCols, L.replicas), dim3(s.scan_threads), 0, stream, L, B);
    return hipGetLastError();
}

hipError_t launch_scatter(const dsdx_particles& p, const FieldParams& fp, const Layout& L,
                          const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        hipLaunchKernelGGL((scatter_kernel<I, W>), dim3(wave_blocks(L.chunks, W)), dim3(W * kWave), 0,
                           stream, p, fp, L, B);
        return hipGetLastError();
    });
}

hipError_t launch_reverse_forces(double* fx, double* fy, double* fz, const Layout& L,
                                 const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        hipLaunchKernelGGL((reverse_forces_kernel<I, W>), dim3(wave_blocks(L.chunks, W)),
                           dim3(W * kWave), 0, stream, fx, fy, fz, L, B);
        return hipGetLastError();
    });
}

hipError_t launch_reverse_fields(const TypedFields& tf, const Layout& L, const PlanBuffers& B,
                                 const Shape& s, hipStream_t stream)
{
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        hipLaunchKernelGGL((reverse_fields_kernel<I, W>), dim3(wave_blocks(L.chunks, W)),
                           dim3(W * kWave), 0, stream, tf, L, B);
        return hipGetLastError();
    });
}

hipError_t launch_forward_fields(const FieldParams& fp, const Layout& L, const PlanBuffers& B,
                                 const Shape& s, hipStream_t stream)
{
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        hipLaunchKernelGGL((forward_fields_kernel<I, W>), dim3(wave_blocks(L.chunks, W)),
                           dim3(W * kWave), 0, stream, fp, L, B);
        return hipGetLastError();
    });
}

hipError_t launch_refresh(const dsdx_particles& p, const FieldParams& fp, int32_t positions_only,
                          const Layout& L, const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        hipLaunchKernelGGL((refresh_kernel<I, W>), dim3(wave_blocks(L.chunks, W)), dim3(W * kWave), 0,
                           stream, p, fp, positions_only, L, B);
        return hipGetLastError();
    });
}

hipError_t launch_ghost_owners(int32_t* out, const Layout& L, const PlanBuffers& B, const Shape& s,
                               hipStream_t stream)
{
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        hipLaunchKernelGGL((ghost_owners_kernel<I, W>), dim3(wave_blocks(L.chunks, W)),
                           dim3(W * kWave), 0, stream, out, L, B);
        return hipGetLastError();
    });
}

hipError_t launch_migrate_wrap(double* x, double* y, double* z, const Layout& L, const Shape& s,
                               hipStream_t stream)
{
    if (L.chunks == 0) return hipSuccess;
    return dispatch(s, [&](auto items, auto waves) {
        constexpr int I = decltype(items)::value, W = decltype(waves)::value;
        hipLaunchKernelGGL((migrate_wrap_kernel<I, W>), dim3(wave_blocks(L.chunks, W)),
                           dim3(W * kWave), 0, stream, x, y, z, L);
        return hipGetLastError();
    });
}

}  // namespace dsdx
