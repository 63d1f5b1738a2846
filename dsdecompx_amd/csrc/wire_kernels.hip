// wire_kernels.hip -- gfx950 kernels of the reference's communication
// records, for callers that carry the records over a transport of their own:
// particle records (ParticleRecord, include/field.hpp:150-157, 96 bytes, then
// the registered fields), position records (3 doubles), returned force
// records and exchange records of registered fields.
//
// The plan-aware kernels walk the owner tiles of a single-box plan (one lane
// per boundary owner, its images in ascending send slot) and address each
// image's record in a send buffer of one block per send slot, blocks in slot
// order, each in owned order -- the layout of the reference's HaloPackSpec
// offsets (src/cuda/kernels.hpp:30-37, src/cuda/pack.cu:146-180).
#include "device.hpp"
#include "kernels.hpp"
#include "owners.hpp"
#include "wave.hpp"

namespace dsdx {
using namespace kern;
namespace {

// Lane s < 26 of the calling wave (all lanes active): the offset from a ghost
// of send slot s (relative to `owned`) to its record in the send buffer.  On
// a loopback box receive block m holds slot 25 - m's sends in the same order,
// so a ghost at recv_start[25 - s] + j is record send_start[s] + j.
__device__ __forceinline__ uint32_t send_delta(const uint32_t* __restrict__ totals)
{
    const uint32_t lane = lane_id();
    const uint32_t sent = lane < kSlots ? totals[lane] : 0u;
    const uint32_t send_start = wave_inclusive_sum(sent) - sent;
    const uint32_t got = lane < kSlots ? totals[kSlots - 1 - lane] : 0u;   // receive block `lane`
    const uint32_t recv_start = wave_inclusive_sum(got) - got;
    const uint32_t mirror = static_cast<uint32_t>(
        __shfl(static_cast<int>(recv_start), lane < kSlots ? kSlots - 1 - lane : 0, kWave));
    return send_start - mirror;   // modulo 2^32; ghost + delta is the record index
}

// Each wave keeps its slot deltas in its own LDS row (written and read by the
// same wave: LDS operations of one wave complete in order).
template <int WAVES>
__device__ __forceinline__ const uint32_t* wave_deltas(uint32_t (&s_delta)[WAVES][kWave],
                                                       const PlanBuffers& B)
{
    const uint32_t wave = threadIdx.x / kWave;
    s_delta[wave][lane_id()] = send_delta(B.totals);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    return s_delta[wave];
}

__device__ __forceinline__ void copy_words(const uint32_t* __restrict__ from, uint32_t* __restrict__ to,
                                           uint32_t words)
{
    for (uint32_t w = 0; w < words; ++w) to[w] = from[w];
}

// Records are staged in LDS 13 words apart (12 used): 64-bit reads of lane l
// at word 13 l spread over the banks.
constexpr uint32_t kRecordWords = kRecordBytes / 8;
constexpr uint32_t kStageWords = kRecordWords + 1;

__device__ __forceinline__ uint64_t bits_of(double v) { return static_cast<uint64_t>(__double_as_longlong(v)); }
__device__ __forceinline__ double double_of(uint64_t v) { return __longlong_as_double(static_cast<long long>(v)); }

// store_halo_record (src/cuda/pack.cu:129-142) of every image of every
// boundary owner: owner + periodic shift of the send slot, velocities, zero
// forces, mass, charge, type, id, then the field words; or, positions_only,
// the three shifted coordinates (halo_gather_positions_kernel, :204-216).
// Reads the owners' current values: after a build it is launch_halo_scatter's
// output, on a reused plan launch_halo_gather's.  `fresh`: the positions have
// not changed since the classification pass, which stored them densely per
// boundary entry (bpos); read them there instead of gathering x, y, z.  The
// owner gathers bound it (one lane per owner, sparse in owned order);
// staging the records through LDS for full-wave stores measured no faster.
template <int WAVES, bool STREAM>
__global__ __launch_bounds__(WAVES * kWave) void pack_halo_kernel(
    dsdx_particles p, RecordFields rf, int32_t positions_only, int32_t fresh, unsigned char* __restrict__ out,
    const ReplicaDesc* __restrict__ desc, const uint32_t* __restrict__ tile_replica, Layout L,
    PlanBuffers B)
{
    __shared__ uint32_t s_delta[WAVES][kWave];
    const uint32_t* delta = wave_deltas<WAVES>(s_delta, B);
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        if (!o.live) return;
        const BoxParams& bp = o.d->bp;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const OwnerGhosts g = ghosts_of(desc, o, B);
        double x, y, z;
        if (fresh) {
            const uint32_t E = B.owner_entry[o.u];
            x = B.bpos[E];
            y = B.bpos[B.bpos_stride + E];
            z = B.bpos[2 * B.bpos_stride + E];
        } else {
            x = read_once<STREAM>(p.x + i);
            y = read_once<STREAM>(p.y + i);
            z = read_once<STREAM>(p.z + i);
        }
        // every image's record, first the preloaded ones
        auto each_image = [&](auto&& put) {
            const FirstImages f = first_images(g);
#pragma unroll
            for (uint32_t k = 0; k < kPreloaded; ++k) {
                if (k >= f.n) break;
                put(f.slot[k], f.rel[k] + delta[f.slot[k]]);
            }
            uint32_t k = kPreloaded;
            for (uint32_t rest = f.rest; rest; rest &= rest - 1u, ++k) {
                const uint32_t s = static_cast<uint32_t>(__builtin_ctz(rest));
                put(s, g.list[k] + delta[s]);
            }
        };
        if (positions_only) {
            each_image([&](uint32_t s, uint32_t at) {
                double* r = reinterpret_cast<double*>(out) + 3ull * at;
                r[0] = x + bp.shift[s][0];
                r[1] = y + bp.shift[s][1];
                r[2] = z + bp.shift[s][2];
            });
            return;
        }
        const double vx = read_once<STREAM>(p.vx + i), vy = read_once<STREAM>(p.vy + i), vz = read_once<STREAM>(p.vz + i);
        const double m = read_once<STREAM>(p.mass + i), q = read_once<STREAM>(p.charge + i);
        const uint64_t ids = static_cast<uint32_t>(read_once<STREAM>(p.type + i)) |
                             (static_cast<uint64_t>(static_cast<uint32_t>(read_once<STREAM>(p.global_id + i))) << 32);
        each_image([&](uint32_t s, uint32_t at) {
            unsigned char* rec = out + static_cast<uint64_t>(at) * rf.stride;
            double* r = reinterpret_cast<double*>(rec);
            r[0] = x + bp.shift[s][0];
            r[1] = y + bp.shift[s][1];
            r[2] = z + bp.shift[s][2];
            r[3] = vx; r[4] = vy; r[5] = vz;
            r[6] = 0.0; r[7] = 0.0; r[8] = 0.0;
            r[9] = m; r[10] = q;
            reinterpret_cast<uint64_t*>(rec)[11] = ids;
            for (uint32_t f = 0; f < rf.count; ++f)
                copy_words(static_cast<const uint32_t*>(rf.data[f]) + i * rf.words[f],
                           reinterpret_cast<uint32_t*>(rec + rf.offset[f]), rf.words[f]);
        });
    });
}

// unpack_kernel (src/cuda/unpack.cu:16-31): record k into particle dest + k,
// every member as received (null arrays skipped), then the field words.  A
// wave takes 64 records: their header words come in through consecutive
// lanes into LDS, then each lane writes one particle.
__global__ __launch_bounds__(256) void unpack_records_kernel(const unsigned char* __restrict__ in,
                                                             uint64_t count, dsdx_particles p,
                                                             uint64_t dest, RecordFields rf)
{
    __shared__ uint64_t s_rec[4][kWave * kStageWords];
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    const uint64_t base = (blockIdx.x * 4ull + wave) * kWave;
    if (base >= count) return;
    const uint32_t n = static_cast<uint32_t>(min<uint64_t>(kWave, count - base));
    const uint32_t stride_w = rf.stride / 8;
    const uint64_t* in64 = reinterpret_cast<const uint64_t*>(in);
    uint64_t* stage = s_rec[wave];
    for (uint32_t w = lane; w < n * kRecordWords; w += kWave) {
        const uint32_t j = w / kRecordWords, c = w - j * kRecordWords;
        stage[j * kStageWords + c] = in64[(base + j) * stride_w + c];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane >= n) return;
    const uint64_t* r = stage + lane * kStageWords;
    const uint64_t k = base + lane;
    const uint64_t i = dest + k;
    if (p.x) p.x[i] = double_of(r[0]);
    if (p.y) p.y[i] = double_of(r[1]);
    if (p.z) p.z[i] = double_of(r[2]);
    if (p.vx) p.vx[i] = double_of(r[3]);
    if (p.vy) p.vy[i] = double_of(r[4]);
    if (p.vz) p.vz[i] = double_of(r[5]);
    if (p.fx) p.fx[i] = double_of(r[6]);
    if (p.fy) p.fy[i] = double_of(r[7]);
    if (p.fz) p.fz[i] = double_of(r[8]);
    if (p.mass) p.mass[i] = double_of(r[9]);
    if (p.charge) p.charge[i] = double_of(r[10]);
    if (p.type) p.type[i] = static_cast<int32_t>(static_cast<uint32_t>(r[11]));
    if (p.global_id) p.global_id[i] = static_cast<int32_t>(static_cast<uint32_t>(r[11] >> 32));
    const unsigned char* rec = in + k * rf.stride;
    for (uint32_t f = 0; f < rf.count; ++f)
        copy_words(reinterpret_cast<const uint32_t*>(rec + rf.offset[f]),
                   static_cast<uint32_t*>(rf.data[f]) + i * rf.words[f], rf.words[f]);
}

// unpack_positions_kernel (src/cuda/unpack.cu:33-42).
__global__ __launch_bounds__(256) void unpack_positions_kernel(const double* __restrict__ in,
                                                               uint64_t count, dsdx_particles p,
                                                               uint64_t dest)
{
    const uint64_t k = blockIdx.x * 256ull + threadIdx.x;
    if (k >= count) return;
    const uint64_t i = dest + k;
    if (p.x) p.x[i] = in[3 * k];
    if (p.y) p.y[i] = in[3 * k + 1];
    if (p.z) p.z[i] = in[3 * k + 2];
    if (p.fx) p.fx[i] = 0.0;
    if (p.fy) p.fy[i] = 0.0;
    if (p.fz) p.fz[i] = 0.0;
}

// pack_reverse_kernel (src/cuda/pack.cu:219-227): a default record carrying
// ghost first + k's force.  A wave takes 64 records: the forces go through
// LDS and the 768 words leave through consecutive lanes.
__global__ __launch_bounds__(256) void pack_reverse_kernel(const double* __restrict__ fx,
                                                           const double* __restrict__ fy,
                                                           const double* __restrict__ fz,
                                                           uint64_t first, uint64_t count,
                                                           unsigned char* __restrict__ out)
{
    __shared__ double s_f[4][3][kWave];
    const uint32_t wave = threadIdx.x / kWave;
    const uint32_t lane = lane_id();
    const uint64_t base = (blockIdx.x * 4ull + wave) * kWave;
    if (base >= count) return;
    const uint32_t n = static_cast<uint32_t>(min<uint64_t>(kWave, count - base));
    if (lane < n) {
        s_f[wave][0][lane] = fx[first + base + lane];
        s_f[wave][1][lane] = fy[first + base + lane];
        s_f[wave][2][lane] = fz[first + base + lane];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    uint64_t* o = reinterpret_cast<uint64_t*>(out) + base * kRecordWords;
    for (uint32_t w = lane; w < n * kRecordWords; w += kWave) {
        const uint32_t j = w / kRecordWords, c = w - j * kRecordWords;
        o[w] = (c >= 6 && c <= 8) ? bits_of(s_f[wave][c - 6][j]) : 0u;
    }
}

// The returned force records (send-buffer layout) added to their owners, one
// lane per owner over its images in ascending send slot: the order of the
// reference's one accumulate launch per neighbour block
// (src/halo.cpp:1069-1079, src/cuda/unpack.cu:46-57), without atomics.
template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void accumulate_reverse_kernel(
    const unsigned char* __restrict__ in, double* __restrict__ fx, double* __restrict__ fy,
    double* __restrict__ fz, const ReplicaDesc* __restrict__ desc,
    const uint32_t* __restrict__ tile_replica, Layout L, PlanBuffers B)
{
    __shared__ uint32_t s_delta[WAVES][kWave];
    const uint32_t* delta = wave_deltas<WAVES>(s_delta, B);
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const OwnerGhosts g = ghosts_of(desc, o, B);
        double ax = fx[i], ay = fy[i], az = fz[i];
        auto forces = [&](uint32_t at) {
            return reinterpret_cast<const double*>(in + static_cast<uint64_t>(at) * kRecordBytes) + 6;
        };
        // the forces of the first kPreloaded images loaded together, then
        // added in slot order
        const FirstImages f = first_images(g);
        double gx[kPreloaded], gy[kPreloaded], gz[kPreloaded];
#pragma unroll
        for (uint32_t k = 0; k < kPreloaded; ++k) {
            const double* r = forces(f.rel[k] + delta[f.slot[k]]);
            gx[k] = r[0]; gy[k] = r[1]; gz[k] = r[2];
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
            const double* r = forces(g.list[k] + delta[__builtin_ctz(rest)]);
            ax = ax + r[0];
            ay = ay + r[1];
            az = az + r[2];
        }
        fx[i] = ax; fy[i] = ay; fz[i] = az;
    });
}

// pack_reverse_fields_kernel / scatter_fields_kernel
// (src/cuda/reverse_fields.cu:128-144, :196-210): particle first + k to and
// from exchange record k.
__global__ __launch_bounds__(256) void pack_field_records_kernel(RecordFields rf, uint64_t first,
                                                                 uint64_t count,
                                                                 unsigned char* __restrict__ out)
{
    const uint64_t k = blockIdx.x * 256ull + threadIdx.x;
    if (k >= count) return;
    for (uint32_t f = 0; f < rf.count; ++f)
        copy_words(static_cast<const uint32_t*>(rf.data[f]) + (first + k) * rf.words[f],
                   reinterpret_cast<uint32_t*>(out + k * rf.stride + rf.offset[f]), rf.words[f]);
}

__global__ __launch_bounds__(256) void scatter_field_records_kernel(RecordFields rf,
                                                                    const unsigned char* __restrict__ in,
                                                                    uint64_t count, uint64_t first)
{
    const uint64_t k = blockIdx.x * 256ull + threadIdx.x;
    if (k >= count) return;
    for (uint32_t f = 0; f < rf.count; ++f)
        copy_words(reinterpret_cast<const uint32_t*>(in + k * rf.stride + rf.offset[f]),
                   static_cast<uint32_t*>(rf.data[f]) + (first + k) * rf.words[f], rf.words[f]);
}

// gather_fields_kernel (src/cuda/reverse_fields.cu:182-194) over every send
// block: each image's exchange record holds its owner's fields.
template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void gather_field_records_kernel(
    RecordFields rf, unsigned char* __restrict__ out, const ReplicaDesc* __restrict__ desc,
    const uint32_t* __restrict__ tile_replica, Layout L, PlanBuffers B)
{
    __shared__ uint32_t s_delta[WAVES][kWave];
    const uint32_t* delta = wave_deltas<WAVES>(s_delta, B);
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const OwnerGhosts g = ghosts_of(desc, o, B);
        uint32_t k = 0;
        for (uint32_t rest = g.slots; rest; rest &= rest - 1u, ++k) {
            const uint32_t s = static_cast<uint32_t>(__builtin_ctz(rest));
            unsigned char* rec = out + static_cast<uint64_t>(g.list[k] + delta[s]) * rf.stride;
            for (uint32_t f = 0; f < rf.count; ++f)
                copy_words(static_cast<const uint32_t*>(rf.data[f]) + i * rf.words[f],
                           reinterpret_cast<uint32_t*>(rec + rf.offset[f]), rf.words[f]);
        }
    });
}

template <class T>
__device__ void reduce_records(T* data, uint32_t c, int32_t op, uint64_t i, const OwnerGhosts& g,
                               const uint32_t* delta, const unsigned char* in, uint32_t stride,
                               uint32_t offset)
{
    for (uint32_t e = 0; e < c; ++e) {
        T acc = data[i * c + e];
        uint32_t k = 0;
        for (uint32_t rest = g.slots; rest; rest &= rest - 1u, ++k) {
            const uint32_t s = static_cast<uint32_t>(__builtin_ctz(rest));
            const T* v = reinterpret_cast<const T*>(in + static_cast<uint64_t>(g.list[k] + delta[s]) * stride +
                                                    offset);
            acc = reduce_value<T>(acc, v[e], op);
        }
        data[i * c + e] = acc;
    }
}

// reduce_reverse_fields_kernel (src/cuda/reverse_fields.cu:146-180) over the
// returned blocks in slot order, one lane per owner, ascending send slot --
// the order of the reference's one launch per neighbour block; integer sums
// wrap as unsigned.
template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void reduce_field_records_kernel(
    RecordFields rf, const unsigned char* __restrict__ in, const ReplicaDesc* __restrict__ desc,
    const uint32_t* __restrict__ tile_replica, Layout L, PlanBuffers B)
{
    __shared__ uint32_t s_delta[WAVES][kWave];
    const uint32_t* delta = wave_deltas<WAVES>(s_delta, B);
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const OwnerGhosts g = ghosts_of(desc, o, B);
        for (uint32_t f = 0; f < rf.count; ++f) {
            const uint32_t c = rf.components[f];
            const int32_t op = rf.op[f];
            const uint32_t off = rf.offset[f];
            switch (rf.type[f]) {
            case DSDX_INT32:
                if (op == DSDX_SUM)
                    reduce_records<uint32_t>(static_cast<uint32_t*>(rf.data[f]), c, op, i, g, delta, in, rf.stride, off);
                else
                    reduce_records<int32_t>(static_cast<int32_t*>(rf.data[f]), c, op, i, g, delta, in, rf.stride, off);
                break;
            case DSDX_INT64:
                if (op == DSDX_SUM)
                    reduce_records<uint64_t>(static_cast<uint64_t*>(rf.data[f]), c, op, i, g, delta, in, rf.stride, off);
                else
                    reduce_records<int64_t>(static_cast<int64_t*>(rf.data[f]), c, op, i, g, delta, in, rf.stride, off);
                break;
            case DSDX_UINT32:
                reduce_records<uint32_t>(static_cast<uint32_t*>(rf.data[f]), c, op, i, g, delta, in, rf.stride, off);
                break;
            case DSDX_UINT64:
                reduce_records<uint64_t>(static_cast<uint64_t*>(rf.data[f]), c, op, i, g, delta, in, rf.stride, off);
                break;
            case DSDX_FLOAT32:
                reduce_records<float>(static_cast<float*>(rf.data[f]), c, op, i, g, delta, in, rf.stride, off);
                break;
            default:
                reduce_records<double>(static_cast<double*>(rf.data[f]), c, op, i, g, delta, in, rf.stride, off);
                break;
            }
        }
    });
}

// The self messages of a loopback box (src/backends/mpi_backend.cpp:366-375):
// block b moves words[b] 8-byte words; grid (x, 26), blockIdx.y = b.
__global__ __launch_bounds__(256) void block_copy_kernel(const uint64_t* __restrict__ src,
                                                         uint64_t* __restrict__ dst, BlockMap map)
{
    const uint32_t b = blockIdx.y;
    const uint64_t words = map.words[b];
    const uint64_t from = map.src[b], to = map.dst[b];
    for (uint64_t w = blockIdx.x * 256ull + threadIdx.x; w < words; w += gridDim.x * 256ull)
        dst[to + w] = src[from + w];
}

unsigned blocks_for(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

hipError_t launch_pack_halo(const dsdx_particles& p, const RecordFields& rf, int32_t positions_only,
                            int32_t fresh, unsigned char* out, const Layout& L, const PlanBuffers& B,
                            const Shape& s, hipStream_t stream)
{
    if (L.tiles == 0) return hipSuccess;
    const Layout O = ordered(L, s, kOrderOwners);
    return dispatch_waves(s, [&](auto waves) {
        return dispatch_stream(L.streaming, [&](auto streaming) {
            constexpr int W = decltype(waves)::value;
            constexpr bool S = decltype(streaming)::value;
            return launch(pack_halo_kernel<W, S>, tile_grid(L, W), dim3(W * kWave), stream, p, rf,
                          positions_only, fresh, out, L.desc, L.tile_replica, O, B);
        });
    });
}

hipError_t launch_unpack_records(const unsigned char* in, uint64_t count, const dsdx_particles& p,
                                 uint64_t dest, const RecordFields& rf, hipStream_t stream)
{
    if (count == 0) return hipSuccess;
    return launch(unpack_records_kernel, dim3(blocks_for(count)), dim3(256), stream, in, count, p, dest, rf);
}

hipError_t launch_unpack_positions(const double* in, uint64_t count, const dsdx_particles& p, uint64_t dest,
                                   hipStream_t stream)
{
    if (count == 0) return hipSuccess;
    return launch(unpack_positions_kernel, dim3(blocks_for(count)), dim3(256), stream, in, count, p, dest);
}

hipError_t launch_pack_reverse(const double* fx, const double* fy, const double* fz, uint64_t first,
                               uint64_t count, unsigned char* out, hipStream_t stream)
{
    if (count == 0) return hipSuccess;
    return launch(pack_reverse_kernel, dim3(blocks_for(count)), dim3(256), stream, fx, fy, fz, first, count,
                  out);
}

hipError_t launch_accumulate_reverse(const unsigned char* in, double* fx, double* fy, double* fz,
                                     const Layout& L, const PlanBuffers& B, const Shape& s,
                                     hipStream_t stream)
{
    if (L.tiles == 0) return hipSuccess;
    const Layout O = ordered(L, s, kOrderOwners);
    return dispatch_waves(s, [&](auto waves) {
        constexpr int W = decltype(waves)::value;
        return launch(accumulate_reverse_kernel<W>, tile_grid(L, W), dim3(W * kWave), stream, in, fx, fy, fz,
                      L.desc, L.tile_replica, O, B);
    });
}

hipError_t launch_pack_field_records(const RecordFields& rf, uint64_t first, uint64_t count,
                                     unsigned char* out, hipStream_t stream)
{
    if (count == 0 || rf.count == 0) return hipSuccess;
    return launch(pack_field_records_kernel, dim3(blocks_for(count)), dim3(256), stream, rf, first, count, out);
}

hipError_t launch_scatter_field_records(const RecordFields& rf, const unsigned char* in, uint64_t count,
                                        uint64_t first, hipStream_t stream)
{
    if (count == 0 || rf.count == 0) return hipSuccess;
    return launch(scatter_field_records_kernel, dim3(blocks_for(count)), dim3(256), stream, rf, in, count,
                  first);
}

hipError_t launch_gather_field_records(const RecordFields& rf, unsigned char* out, const Layout& L,
                                       const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    if (L.tiles == 0 || rf.count == 0) return hipSuccess;
    const Layout O = ordered(L, s, kOrderOwners);
    return dispatch_waves(s, [&](auto waves) {
        constexpr int W = decltype(waves)::value;
        return launch(gather_field_records_kernel<W>, tile_grid(L, W), dim3(W * kWave), stream, rf, out, L.desc,
                      L.tile_replica, O, B);
    });
}

hipError_t launch_reduce_field_records(const RecordFields& rf, const unsigned char* in, const Layout& L,
                                       const PlanBuffers& B, const Shape& s, hipStream_t stream)
{
    if (L.tiles == 0 || rf.count == 0) return hipSuccess;
    const Layout O = ordered(L, s, kOrderOwners);
    return dispatch_waves(s, [&](auto waves) {
        constexpr int W = decltype(waves)::value;
        return launch(reduce_field_records_kernel<W>, tile_grid(L, W), dim3(W * kWave), stream, rf, in, L.desc,
                      L.tile_replica, O, B);
    });
}

hipError_t launch_block_copy(const void* src, void* dst, const BlockMap& map, hipStream_t stream)
{
    uint64_t most = 0;
    for (int b = 0; b < kSlots; ++b) most = std::max(most, map.words[b]);
    if (most == 0) return hipSuccess;
    const unsigned x = std::min<unsigned>(blocks_for(most), 2048u);
    return launch(block_copy_kernel, dim3(x, kSlots), dim3(256), stream, static_cast<const uint64_t*>(src),
                  static_cast<uint64_t*>(dst), map);
}

}  // namespace dsdx
