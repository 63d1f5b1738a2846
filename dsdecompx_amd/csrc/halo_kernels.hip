// halo_kernels.hip -- gfx950 kernels of the loopback particle-communication
// path.  Bandwidth-bound stream compaction; no MFMA.  Built with
// -ffp-contract=off: every FP64 operation is separately rounded, as in the
// reference Host build, so positions and sums are bit-identical to it.
#include "device.hpp"
#include "kernels.hpp"

namespace dsdx {
namespace {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

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

__device__ __forceinline__ uint32_t uniform(uint32_t v)
{
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint32_t slab_code(double px, double py, double pz,
                                              const BoxParams& bp)
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
// Pass 1: one wave per chunk.  Slab-classifies every owned particle,
// compacts the boundary ones (non-empty slot set) into the chunk's entry
// region in owned order, and counts per send slot through an LDS histogram of
// slab codes.  Reads 24 B per particle, writes 4 B per boundary particle.
__global__ __launch_bounds__(kBlock) void classify_kernel(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, uint32_t owned, BoxParams bp,
    uint32_t* __restrict__ entries, uint32_t* __restrict__ counts, uint32_t chunks)
{
    __shared__ uint32_t s_lut[64];
    __shared__ uint32_t s_hist[kWavesPerBlock][64];
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t chunk = blockIdx.x * kWavesPerBlock + wave;
    if (threadIdx.x < 64) s_lut[threadIdx.x] = bp.lut[threadIdx.x];
    s_hist[wave][lane] = 0u;
    __syncthreads();

    uint32_t running = 0;
    if (chunk < chunks) {
        const uint64_t base = static_cast<uint64_t>(chunk) * kChunk;
        double px[kItems], py[kItems], pz[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t i = base + static_cast<uint64_t>(k) * kWave + lane;
            const bool valid = i < owned;
            px[k] = valid ? x[i] : 0.0;
            py[k] = valid ? y[i] : 0.0;
            pz[k] = valid ? z[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t i = base + static_cast<uint64_t>(k) * kWave + lane;
            const uint32_t code = i < owned ? slab_code(px[k], py[k], pz[k], bp) : 0u;
            const bool boundary = s_lut[code] != 0u;
            const uint64_t ballot = __ballot(boundary);
            if (boundary) {
                entries[base + running + lanes_below(ballot)] =
                    ((static_cast<uint32_t>(k) * kWave + lane) << 8) | code;
                atomicAdd(&s_hist[wave][code], 1u);
            }
            running += static_cast<uint32_t>(__popcll(ballot));
        }
    }
    __syncthreads();
    if (chunk < chunks) {
        // lane s < 26: send count of slot s; lane 26: owners; lane 27: images
        uint32_t acc = 0;
        for (int c = 1; c < 64; ++c) {
            const uint32_t h = s_hist[wave][c];
            const uint32_t m = s_lut[c];
            if (lane < kSlots) acc += ((m >> lane) & 1u) ? h : 0u;
            else if (lane == kColImages) acc += h * static_cast<uint32_t>(__popc(m));
        }
        if (lane == kColOwners) acc = running;
        if (lane < kCols) counts[static_cast<uint64_t>(lane) * chunks + chunk] = acc;
    }
}

// ---------------------------------------------------------------------------
// Pass 2: exclusive scan of every count column over the chunks; one block of
// 1024 threads per column.  totals[col] receives the column total.
constexpr int kScanThreads = 1024;

__global__ __launch_bounds__(kScanThreads) void scan_kernel(
    const uint32_t* __restrict__ counts, uint32_t* __restrict__ prefix,
    uint32_t* __restrict__ totals, uint32_t chunks)
{
    __shared__ uint32_t s_wave[kScanThreads / kWave];
    const uint32_t col = blockIdx.x;
    const uint32_t t = threadIdx.x;
    const uint32_t per = (chunks + kScanThreads - 1) / kScanThreads;
    const uint32_t begin = min(t * per, chunks);
    const uint32_t end = min(begin + per, chunks);
    const uint32_t* src = counts + static_cast<uint64_t>(col) * chunks;
    uint32_t* dst = prefix + static_cast<uint64_t>(col) * chunks;
    uint32_t sum = 0;
    for (uint32_t k = begin; k < end; ++k) sum += src[k];
    const uint32_t incl = wave_inclusive_sum(sum);
    const uint32_t wave = t / kWave;
    if (lane_id() == kWave - 1) s_wave[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        const uint32_t w = lane_id() < kScanThreads / kWave ? s_wave[lane_id()] : 0u;
        const uint32_t wi = wave_inclusive_sum(w);
        if (lane_id() < kScanThreads / kWave) s_wave[lane_id()] = wi - w;
        if (lane_id() == kScanThreads / kWave - 1) totals[col] = wi;
    }
    __syncthreads();
    uint32_t run = s_wave[wave] + incl - sum;
    for (uint32_t k = begin; k < end; ++k) {
        const uint32_t v = src[k];
        dst[k] = run;
        run += v;
    }
    if (col == 0 && t == 0) totals[kStatusWord] = 0u;
}

// ---------------------------------------------------------------------------
// Pass 3: one wave per chunk.  Walks the chunk's boundary particles in owned
// order; for each of its send slots s (ascending) a particle gets the ghost
// at receive block 25-s, position prefix[s][chunk] + rank in the chunk.
// Writes the 13 ghost arrays, the registered fields and the owner -> ghost
// map.  The capacity test is uniform over the grid, so an overflow leaves
// every ghost slot untouched.
__device__ __forceinline__ void copy_field_words(const FieldParams& fp, uint64_t from,
                                                 uint64_t to)
{
    for (uint32_t f = 0; f < fp.count; ++f) {
        const uint32_t w = fp.words[f];
        uint32_t* d = static_cast<uint32_t*>(fp.data[f]);
        for (uint32_t k = 0; k < w; ++k) d[to * w + k] = d[from * w + k];
    }
}

__global__ __launch_bounds__(kBlock) void scatter_kernel(
    dsdx_particles p, uint32_t owned, BoxParams bp, FieldParams fp,
    const uint32_t* __restrict__ entries, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ prefix, uint32_t* __restrict__ totals,
    uint32_t* __restrict__ owner_index, uint8_t* __restrict__ owner_code,
    uint32_t* __restrict__ image_start, uint32_t* __restrict__ ghost_list,
    uint32_t chunks, uint64_t capacity, uint64_t ghost_list_capacity)
{
    __shared__ uint32_t s_lut[64];
    if (threadIdx.x < 64) s_lut[threadIdx.x] = bp.lut[threadIdx.x];
    __syncthreads();

    const uint32_t ghosts = totals[kColImages];
    uint32_t overflow = 0;
    if (static_cast<uint64_t>(owned) + ghosts > capacity) overflow |= kOverflowStorage;
    if (ghosts > ghost_list_capacity) overflow |= kOverflowGhostList;
    if (overflow) {
        if (blockIdx.x == 0 && threadIdx.x == 0) totals[kStatusWord] = overflow;
        return;
    }
    const uint32_t lane = lane_id();
    const uint32_t chunk = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    if (chunk >= chunks) return;

    // Receive block m holds send slot 25-m: recv_base[m] = sum of totals of
    // slots 25..26-m.  Lane s ends with slot_base = recv_base[25-s] + prefix.
    const uint32_t block_size = lane < kSlots ? totals[kSlots - 1 - lane] : 0u;
    const uint32_t recv_base = wave_inclusive_sum(block_size) - block_size;
    const uint32_t mirrored = __shfl(recv_base, lane < kSlots ? kSlots - 1 - lane : 0, kWave);
    const uint32_t slot_base =
        lane < kSlots ? mirrored + prefix[static_cast<uint64_t>(lane) * chunks + chunk] : 0u;
    uint32_t slot_run = 0;   // lane s: ghosts of slot s already written in this chunk

    const uint32_t n_boundary = counts[static_cast<uint64_t>(kColOwners) * chunks + chunk];
    const uint32_t owner_base = prefix[static_cast<uint64_t>(kColOwners) * chunks + chunk];
    uint32_t image_run = prefix[static_cast<uint64_t>(kColImages) * chunks + chunk];
    const uint64_t chunk_first = static_cast<uint64_t>(chunk) * kChunk;

    for (uint32_t e0 = 0; e0 < n_boundary; e0 += kWave) {
        const uint32_t e = e0 + lane;
        const bool valid = e < n_boundary;
        const uint32_t entry = valid ? entries[chunk_first + e] : 0u;
        const uint32_t code = entry & 63u;
        const uint64_t i = chunk_first + (entry >> 8);
        const uint32_t slots = valid ? s_lut[code] : 0u;
        const uint32_t images = static_cast<uint32_t>(__popc(slots));
        const uint32_t incl = wave_inclusive_sum(images);
        const uint32_t first_image = image_run + incl - images;
        image_run += uniform(__shfl(incl, kWave - 1, kWave));

        double px = 0, py = 0, pz = 0, vx = 0, vy = 0, vz = 0, m = 0, q = 0;
        int32_t ty = 0, id = 0;
        if (valid) {
            const uint32_t u = owner_base + e;
            owner_index[u] = static_cast<uint32_t>(i);
            owner_code[u] = static_cast<uint8_t>(code);
            image_start[u] = first_image;
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
                const uint64_t j = static_cast<uint64_t>(owned) + rel;
                p.x[j] = px + bp.shift[s][0];
                p.y[j] = py + bp.shift[s][1];
                p.z[j] = pz + bp.shift[s][2];
                p.vx[j] = vx; p.vy[j] = vy; p.vz[j] = vz;
                if (p.fx) p.fx[j] = 0.0;
                if (p.fy) p.fy[j] = 0.0;
                if (p.fz) p.fz[j] = 0.0;
                p.mass[j] = m; p.charge[j] = q;
                p.type[j] = ty; p.global_id[j] = id;
                if (fp.count) copy_field_words(fp, i, j);
                ghost_list[first_image + written] = rel;
                ++written;
            }
            if (lane == s) slot_run += static_cast<uint32_t>(__popcll(ballot));
        }
    }
}

// ---------------------------------------------------------------------------
// Owner-major kernels over the plan: thread per boundary owner, grid-stride.
// Owners of the last build; zero when that build overflowed (its scatter
// wrote no owner map), so the owner-major kernels then touch nothing.
__device__ __forceinline__ uint32_t owner_count(const PlanView& v)
{
    return v.totals[kStatusWord] ? 0u : v.totals[kColOwners];
}

// reverse_halo: f[i] = ((f[i] + g_s1) + g_s2) + ... over the owner's send
// slots s1 < s2 < ..., the order of the Host reduction (src/halo.cpp:1357-1370
// for n ascending).  No atomics: each owner is summed by one thread.
__global__ __launch_bounds__(kBlock) void reverse_forces_kernel(
    double* __restrict__ fx, double* __restrict__ fy, double* __restrict__ fz,
    uint32_t owned, PlanView v, BoxParams bp)
{
    __shared__ uint32_t s_lut[64];
    if (threadIdx.x < 64) s_lut[threadIdx.x] = bp.lut[threadIdx.x];
    __syncthreads();
    const uint32_t n_owners = owner_count(v);
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_owners;
         u += gridDim.x * blockDim.x) {
        const uint32_t i = v.owner_index[u];
        const uint32_t first = v.image_start[u];
        const uint32_t n_img = static_cast<uint32_t>(__popc(s_lut[v.owner_code[u]]));
        double ax = fx[i], ay = fy[i], az = fz[i];
        for (uint32_t k = 0; k < n_img; ++k) {
            const uint64_t j = static_cast<uint64_t>(owned) + v.ghost_list[first + k];
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
__device__ void reduce_field(T* data, uint32_t c, int32_t op, uint32_t i, uint32_t owned,
                             const uint32_t* ghosts, uint32_t n_img)
{
    for (uint32_t e = 0; e < c; ++e) {
        T acc = data[static_cast<uint64_t>(i) * c + e];
        for (uint32_t k = 0; k < n_img; ++k) {
            const uint64_t j = static_cast<uint64_t>(owned) + ghosts[k];
            acc = reduce_value<T>(acc, data[j * c + e], op);
        }
        data[static_cast<uint64_t>(i) * c + e] = acc;
    }
}

// Integer sums wrap as unsigned (two's complement), like the Host's
// destination += value on int32/int64.
__global__ __launch_bounds__(kBlock) void reverse_fields_kernel(
    TypedFields tf, uint32_t owned, PlanView v, BoxParams bp)
{
    __shared__ uint32_t s_lut[64];
    if (threadIdx.x < 64) s_lut[threadIdx.x] = bp.lut[threadIdx.x];
    __syncthreads();
    const uint32_t n_owners = owner_count(v);
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_owners;
         u += gridDim.x * blockDim.x) {
        const uint32_t i = v.owner_index[u];
        const uint32_t* ghosts = v.ghost_list + v.image_start[u];
        const uint32_t n_img = static_cast<uint32_t>(__popc(s_lut[v.owner_code[u]]));
        for (uint32_t f = 0; f < tf.count; ++f) {
            const uint32_t c = tf.components[f];
            const int32_t op = tf.op[f];
            switch (tf.type[f]) {
            case DSDX_INT32:
                if (op == DSDX_SUM)
                    reduce_field<uint32_t>(static_cast<uint32_t*>(tf.data[f]), c, op, i, owned, ghosts, n_img);
                else
                    reduce_field<int32_t>(static_cast<int32_t*>(tf.data[f]), c, op, i, owned, ghosts, n_img);
                break;
            case DSDX_INT64:
                if (op == DSDX_SUM)
                    reduce_field<uint64_t>(static_cast<uint64_t*>(tf.data[f]), c, op, i, owned, ghosts, n_img);
                else
                    reduce_field<int64_t>(static_cast<int64_t*>(tf.data[f]), c, op, i, owned, ghosts, n_img);
                break;
            case DSDX_UINT32:
                reduce_field<uint32_t>(static_cast<uint32_t*>(tf.data[f]), c, op, i, owned, ghosts, n_img);
                break;
            case DSDX_UINT64:
                reduce_field<uint64_t>(static_cast<uint64_t*>(tf.data[f]), c, op, i, owned, ghosts, n_img);
                break;
            case DSDX_FLOAT32:
                reduce_field<float>(static_cast<float*>(tf.data[f]), c, op, i, owned, ghosts, n_img);
                break;
            default:
                reduce_field<double>(static_cast<double*>(tf.data[f]), c, op, i, owned, ghosts, n_img);
                break;
            }
        }
    }
}

// forward_exchange: owner values copied into each of its ghost slots.
__global__ __launch_bounds__(kBlock) void forward_fields_kernel(
    FieldParams fp, uint32_t owned, PlanView v, BoxParams bp)
{
    __shared__ uint32_t s_lut[64];
    if (threadIdx.x < 64) s_lut[threadIdx.x] = bp.lut[threadIdx.x];
    __syncthreads();
    const uint32_t n_owners = owner_count(v);
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_owners;
         u += gridDim.x * blockDim.x) {
        const uint32_t i = v.owner_index[u];
        const uint32_t* ghosts = v.ghost_list + v.image_start[u];
        const uint32_t n_img = static_cast<uint32_t>(__popc(s_lut[v.owner_code[u]]));
        for (uint32_t f = 0; f < fp.count; ++f) {
            const uint32_t w = fp.words[f];
            uint32_t* d = static_cast<uint32_t*>(fp.data[f]);
            for (uint32_t k = 0; k < w; ++k) {
                const uint32_t value = d[static_cast<uint64_t>(i) * w + k];
                for (uint32_t g = 0; g < n_img; ++g)
                    d[(static_cast<uint64_t>(owned) + ghosts[g]) * w + k] = value;
            }
        }
    }
}

// Reused-plan refresh: each ghost re-reads its owner; the slot of the k-th
// ghost of an owner is the k-th set bit of its slot set.
__global__ __launch_bounds__(kBlock) void refresh_kernel(
    dsdx_particles p, FieldParams fp, uint32_t owned, int32_t positions_only,
    PlanView v, BoxParams bp)
{
    __shared__ uint32_t s_lut[64];
    if (threadIdx.x < 64) s_lut[threadIdx.x] = bp.lut[threadIdx.x];
    __syncthreads();
    const uint32_t n_owners = owner_count(v);
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_owners;
         u += gridDim.x * blockDim.x) {
        const uint32_t i = v.owner_index[u];
        const uint32_t* ghosts = v.ghost_list + v.image_start[u];
        uint32_t slots = s_lut[v.owner_code[u]];
        const double px = p.x[i], py = p.y[i], pz = p.z[i];
        for (uint32_t k = 0; slots; ++k) {
            const uint32_t s = static_cast<uint32_t>(__builtin_ctz(slots));
            slots &= slots - 1u;
            const uint64_t j = static_cast<uint64_t>(owned) + ghosts[k];
            p.x[j] = px + bp.shift[s][0];
            p.y[j] = py + bp.shift[s][1];
            p.z[j] = pz + bp.shift[s][2];
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

__global__ __launch_bounds__(kBlock) void ghost_owners_kernel(
    int32_t* __restrict__ out, PlanView v, BoxParams bp)
{
    __shared__ uint32_t s_lut[64];
    if (threadIdx.x < 64) s_lut[threadIdx.x] = bp.lut[threadIdx.x];
    __syncthreads();
    const uint32_t n_owners = owner_count(v);
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_owners;
         u += gridDim.x * blockDim.x) {
        const uint32_t i = v.owner_index[u];
        const uint32_t* ghosts = v.ghost_list + v.image_start[u];
        const uint32_t n_img = static_cast<uint32_t>(__popc(s_lut[v.owner_code[u]]));
        for (uint32_t k = 0; k < n_img; ++k) out[ghosts[k]] = static_cast<int32_t>(i);
    }
}

// ---------------------------------------------------------------------------
// migrate at P = 1: periodic wrap in place (src/migration.cpp:106-113):
// v = x - lower; v -= floor(v / extent) * extent; v + lower, each operation
// rounded on its own (explicit _rn intrinsics, never fused).  Only a changed
// coordinate is stored.
__device__ __forceinline__ double wrap(double value, double lower, double extent)
{
    const double v = __dsub_rn(value, lower);
    const double q = floor(__ddiv_rn(v, extent));
    return __dadd_rn(__dsub_rn(v, __dmul_rn(q, extent)), lower);
}

constexpr int kWrapItems = 4;

__global__ __launch_bounds__(kBlock) void migrate_wrap_kernel(
    double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    uint64_t owned, BoxParams bp)
{
    double* axis[3] = {x, y, z};
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i0 < owned; i0 += stride * kWrapItems) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (!bp.periodic[a]) continue;
            double* d = axis[a];
            double val[kWrapItems];
#pragma unroll
            for (int k = 0; k < kWrapItems; ++k) {
                const uint64_t i = i0 + k * stride;
                val[k] = i < owned ? d[i] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < kWrapItems; ++k) {
                const uint64_t i = i0 + k * stride;
                const double w = wrap(val[k], bp.lower[a], bp.extent[a]);
                if (i < owned && __double_as_longlong(w) != __double_as_longlong(val[k]))
                    d[i] = w;
            }
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers (called by the C ABI in abi.cpp).
hipError_t launch_classify(const double* x, const double* y, const double* z, uint32_t owned,
                           const BoxParams& bp, uint32_t* entries, uint32_t* counts,
                           uint32_t chunks, hipStream_t stream)
{
    if (chunks == 0) return hipSuccess;
    const uint32_t blocks = (chunks + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL(classify_kernel, dim3(blocks), dim3(kBlock), 0, stream, x, y, z, owned,
                       bp, entries, counts, chunks);
    return hipGetLastError();
}

hipError_t launch_scan(const uint32_t* counts, uint32_t* prefix, uint32_t* totals,
                       uint32_t chunks, hipStream_t stream)
{
    hipLaunchKernelGGL(scan_kernel, dim3(kCols), dim3(kScanThreads), 0, stream, counts, prefix,
                       totals, chunks);
    return hipGetLastError();
}

hipError_t launch_scatter(const dsdx_particles& p, uint32_t owned, const BoxParams& bp,
                          const FieldParams& fp, const uint32_t* entries, const uint32_t* counts,
                          const uint32_t* prefix, uint32_t* totals, uint32_t* owner_index,
                          uint8_t* owner_code, uint32_t* image_start, uint32_t* ghost_list,
                          uint32_t chunks, uint64_t capacity, uint64_t ghost_list_capacity,
                          hipStream_t stream)
{
    const uint32_t blocks = std::max<uint32_t>(1, (chunks + kWavesPerBlock - 1) / kWavesPerBlock);
    hipLaunchKernelGGL(scatter_kernel, dim3(blocks), dim3(kBlock), 0, stream, p, owned, bp, fp,
                       entries, counts, prefix, totals, owner_index, owner_code, image_start,
                       ghost_list, chunks, capacity, ghost_list_capacity);
    return hipGetLastError();
}

static uint32_t owner_grid(uint64_t owned)
{
    const uint64_t want = (owned + kBlock - 1) / kBlock;
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(want, 2048)));
}

hipError_t launch_reverse_forces(double* fx, double* fy, double* fz, uint32_t owned,
                                 const PlanView& v, const BoxParams& bp, hipStream_t stream)
{
    hipLaunchKernelGGL(reverse_forces_kernel, dim3(owner_grid(owned)), dim3(kBlock), 0, stream,
                       fx, fy, fz, owned, v, bp);
    return hipGetLastError();
}

hipError_t launch_reverse_fields(const TypedFields& tf, uint32_t owned, const PlanView& v,
                                 const BoxParams& bp, hipStream_t stream)
{
    hipLaunchKernelGGL(reverse_fields_kernel, dim3(owner_grid(owned)), dim3(kBlock), 0, stream,
                       tf, owned, v, bp);
    return hipGetLastError();
}

hipError_t launch_forward_fields(const FieldParams& fp, uint32_t owned, const PlanView& v,
                                 const BoxParams& bp, hipStream_t stream)
{
    hipLaunchKernelGGL(forward_fields_kernel, dim3(owner_grid(owned)), dim3(kBlock), 0, stream,
                       fp, owned, v, bp);
    return hipGetLastError();
}

hipError_t launch_refresh(const dsdx_particles& p, const FieldParams& fp, uint32_t owned,
                          int32_t positions_only, const PlanView& v, const BoxParams& bp,
                          hipStream_t stream)
{
    hipLaunchKernelGGL(refresh_kernel, dim3(owner_grid(owned)), dim3(kBlock), 0, stream, p, fp,
                       owned, positions_only, v, bp);
    return hipGetLastError();
}

hipError_t launch_ghost_owners(int32_t* out, uint32_t owned, const PlanView& v,
                               const BoxParams& bp, hipStream_t stream)
{
    hipLaunchKernelGGL(ghost_owners_kernel, dim3(owner_grid(owned)), dim3(kBlock), 0, stream, out,
                       v, bp);
    return hipGetLastError();
}

hipError_t launch_migrate_wrap(double* x, double* y, double* z, uint64_t owned,
                               const BoxParams& bp, hipStream_t stream)
{
    if (owned == 0) return hipSuccess;
    const uint64_t per_block = static_cast<uint64_t>(kBlock) * kWrapItems;
    const uint64_t blocks = std::min<uint64_t>((owned + per_block - 1) / per_block, 1u << 20);
    hipLaunchKernelGGL(migrate_wrap_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0,
                       stream, x, y, z, owned, bp);
    return hipGetLastError();
}

}  // namespace dsdx
