// floor_probe.cpp -- what the memory access patterns of the step cost on their
// own, on one MI355X, with the library's own plan of a uniform periodic box
// (the benchmark workload: particles uniform in [0, L)^3, width 2.8).  Each
// probe is one kernel of this file that does only the named accesses, timed
// with HIP events over `iters` launches after one warm-up; the figures bound
// the library's kernels from below (their other work comes on top):
//
//   stream_read_xyz   read x, y, z of every owned particle (migrate and the
//                     classification read exactly this)
//   owner_read_f      read fx, fy, fz of every boundary owner (one lane per
//                     owner over the library's owner tiles, owned order)
//   owner_rmw_f       read, add, write fx, fy, fz of every boundary owner
//                     (what reverse_halo must do before any ghost read)
//   owner_gather7     read vx, vy, vz, mass, charge (nontemporal), type, id
//                     and the owner's position from the classification pass,
//                     as the ghost copy's gathers
//   dense_write13     write the 13 ghost arrays of G ghosts contiguously (the
//                     ghost copy's writes, were they one run)
//   gather_write_dense  owner_gather7, then the 13 arrays of one record per
//                     owner written at consecutive slots after the owned
//                     particles: the copy's gathers and writes in one pass,
//                     with no staging and every write in one dense run
//   owner_rmw_first   owner_rmw_f, adding the force of the owner's first image
//                     read from where the copy put it (reverse_halo's reads
//                     and writes for owners with one image)
//   library_copy,     the library's own ghost copy and reverse_halo on the
//   library_reverse   same plan, timed the same way
//
// Each probe writes a small dense output (one value per owner or per wave) so
// that its loads are not removed.  The first three are timed twice: back to
// back (clean caches), and right after a dense rewrite of the three arrays
// they read, as the application's position update precedes migrate and its
// force fill precedes reverse_halo (the write-backs then share the bus).
// Prints one JSON object on stdout.
//
// usage: floor_probe [particles [L [iters]]]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "abi_internal.hpp"
#include "dsdx_hip.h"
#include "kernels.hpp"
#include "owners.hpp"
#include "wave.hpp"

using namespace dsdx;
using namespace dsdx::kern;

namespace {

void ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "HIP error %s: %s\n", what, hipGetErrorString(e));
        std::exit(3);
    }
}

void dsdx_ok(int rc, const char* what)
{
    if (rc != DSDX_OK) {
        std::fprintf(stderr, "dsdx error %s: %d %s\n", what, rc, dsdx_last_error());
        std::exit(4);
    }
}

__device__ __forceinline__ uint64_t mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double unit(uint64_t stream, uint64_t i)
{
    return static_cast<double>(mix(0x9E3779B97F4A7C15ull * (stream + 1) + i) >> 11) * 0x1.0p-53;
}

__global__ void place(dsdx_particles p, uint64_t n, double L)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    p.x[i] = L * unit(0, i); p.y[i] = L * unit(1, i); p.z[i] = L * unit(2, i);
    p.vx[i] = unit(3, i); p.vy[i] = unit(4, i); p.vz[i] = unit(5, i);
    p.fx[i] = unit(6, i); p.fy[i] = unit(7, i); p.fz[i] = unit(8, i);
    p.mass[i] = 1.0; p.charge[i] = unit(9, i);
    p.type[i] = static_cast<int32_t>(i & 3); p.global_id[i] = static_cast<int32_t>(i);
}

// Every owned x, y, z; one sum per wave.
__global__ __launch_bounds__(256) void stream_read_xyz(const double* __restrict__ x, const double* __restrict__ y,
                                                       const double* __restrict__ z, uint64_t n, double* out)
{
    double s = 0.0;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull)
        s += __builtin_nontemporal_load(x + i) + __builtin_nontemporal_load(y + i) +
             __builtin_nontemporal_load(z + i);
    if (s == -1.0) out[blockIdx.x] = s;   // never true for these values; keeps the loads
}

template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void owner_read_f(const double* __restrict__ fx,
                                                              const double* __restrict__ fy,
                                                              const double* __restrict__ fz, double* __restrict__ out,
                                                              const ReplicaDesc* __restrict__ desc,
                                                              const uint32_t* __restrict__ tile_replica, Layout L,
                                                              PlanBuffers B)
{
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        out[o.u] = fx[i] + fy[i] + fz[i];
    });
}

template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void owner_rmw_f(double* __restrict__ fx, double* __restrict__ fy,
                                                             double* __restrict__ fz,
                                                             const ReplicaDesc* __restrict__ desc,
                                                             const uint32_t* __restrict__ tile_replica, Layout L,
                                                             PlanBuffers B)
{
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        fx[i] = fx[i] + 1.0;
        fy[i] = fy[i] + 1.0;
        fz[i] = fz[i] + 1.0;
    });
}

template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void owner_gather7(dsdx_particles p, double* __restrict__ out,
                                                               const ReplicaDesc* __restrict__ desc,
                                                               const uint32_t* __restrict__ tile_replica, Layout L,
                                                               PlanBuffers B)
{
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const uint32_t E = B.owner_entry[o.u];
        const double pos = B.bpos[E] + B.bpos[B.bpos_stride + E] + B.bpos[2 * B.bpos_stride + E];
        const double rec = read_once<true>(p.vx + i) + read_once<true>(p.vy + i) + read_once<true>(p.vz + i) +
                           read_once<true>(p.mass + i) + read_once<true>(p.charge + i);
        const int32_t ids = read_once<true>(p.type + i) + read_once<true>(p.global_id + i);
        out[o.u] = pos + rec + static_cast<double>(ids);
    });
}

// The copy's gathers, then one record per owner at slot first + u.
template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void gather_write_dense(dsdx_particles p, uint64_t first,
                                                                    const ReplicaDesc* __restrict__ desc,
                                                                    const uint32_t* __restrict__ tile_replica,
                                                                    Layout L, PlanBuffers B)
{
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const uint32_t E = B.owner_entry[o.u];
        const uint64_t j = first + o.u;
        const double px = B.bpos[E], py = B.bpos[B.bpos_stride + E], pz = B.bpos[2 * B.bpos_stride + E];
        const double vx = read_once<true>(p.vx + i), vy = read_once<true>(p.vy + i), vz = read_once<true>(p.vz + i);
        const double m = read_once<true>(p.mass + i), q = read_once<true>(p.charge + i);
        const int32_t ty = read_once<true>(p.type + i), id = read_once<true>(p.global_id + i);
        p.x[j] = px; p.y[j] = py; p.z[j] = pz;
        p.vx[j] = vx; p.vy[j] = vy; p.vz[j] = vz;
        p.fx[j] = 0.0; p.fy[j] = 0.0; p.fz[j] = 0.0;
        p.mass[j] = m; p.charge[j] = q;
        p.type[j] = ty; p.global_id[j] = id;
    });
}

// owner_rmw_f plus the force of the owner's first image.
template <int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void owner_rmw_first(double* __restrict__ fx, double* __restrict__ fy,
                                                                 double* __restrict__ fz,
                                                                 const ReplicaDesc* __restrict__ desc,
                                                                 const uint32_t* __restrict__ tile_replica, Layout L,
                                                                 PlanBuffers B)
{
    for_owner_tiles<WAVES>(desc, tile_replica, L, B, [&](const OwnerLane& o) {
        if (!o.live) return;
        const uint64_t i = o.d->offset + B.owner_index[o.u];
        const uint64_t j = o.d->offset + o.d->owned + B.ghost_list[o.d->ghost_list_offset + B.image_start[o.u]];
        const double gx = fx[j], gy = fy[j], gz = fz[j];
        fx[i] = fx[i] + gx;
        fy[i] = fy[i] + gy;
        fz[i] = fz[i] + gz;
    });
}

__global__ __launch_bounds__(256) void dense_write13(dsdx_particles p, uint64_t first, uint64_t count)
{
    const uint64_t k = blockIdx.x * 256ull + threadIdx.x;
    if (k >= count) return;
    const uint64_t j = first + k;
    const double v = static_cast<double>(k);
    p.x[j] = v; p.y[j] = v; p.z[j] = v;
    p.vx[j] = v; p.vy[j] = v; p.vz[j] = v;
    p.fx[j] = 0.0; p.fy[j] = 0.0; p.fz[j] = 0.0;
    p.mass[j] = v; p.charge[j] = v;
    p.type[j] = static_cast<int32_t>(k); p.global_id[j] = static_cast<int32_t>(k);
}

// Rewrites a[i] = scale * unit(stream + k, i) for the three arrays (the
// values place() wrote, so positions stay in the box).
__global__ __launch_bounds__(256) void rewrite3(double* a0, double* a1, double* a2, uint64_t n, uint64_t stream,
                                                double scale)
{
    const uint64_t i = blockIdx.x * 256ull + threadIdx.x;
    if (i >= n) return;
    a0[i] = scale * unit(stream, i);
    a1[i] = scale * unit(stream + 1, i);
    a2[i] = scale * unit(stream + 2, i);
}

unsigned blocks(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

int main(int argc, char** argv)
{
    const uint64_t n = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 16777216ull;
    const double L = argc > 2 ? std::atof(argv[2]) : 256.0;
    const int iters = argc > 3 ? std::atoi(argv[3]) : 20;
    if (n == 0 || iters < 1) {
        std::fprintf(stderr, "usage: %s [particles [L [iters]]]\n", argv[0]);
        return 2;
    }
    const uint64_t cap = n + n / 2 + 4096;
    hipStream_t s = nullptr;
    ok(hipStreamCreate(&s), "stream");
    double* a[11];
    int32_t* ia[2];
    for (auto& q : a) ok(hipMalloc(&q, cap * sizeof(double)), "malloc");
    for (auto& q : ia) ok(hipMalloc(&q, cap * sizeof(int32_t)), "malloc");
    const dsdx_particles p{a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], a[9], a[10], ia[0], ia[1]};
    hipLaunchKernelGGL(place, dim3(blocks(n)), dim3(256), 0, s, p, n, L);
    ok(hipGetLastError(), "place");

    dsdx_box box{};
    for (int k = 0; k < 3; ++k) {
        box.upper[k] = L;
        box.periodic[k] = 1;
    }
    box.width = 2.8;
    box.edges_and_corners = 1;
    dsdx_plan* plan = nullptr;
    dsdx_ok(dsdx_plan_create(&plan), "plan");
    uint64_t ghosts = 0, owners = 0;
    dsdx_ok(dsdx_halo_build(plan, &box, &p, nullptr, 0, n, cap, s, &ghosts), "halo build");
    dsdx_ok(dsdx_plan_query(plan, nullptr, nullptr, nullptr, &owners, nullptr, nullptr), "query");
    const abi::Plan* P = plan;
    // the probes walk the owner maps, which a build over chunk units leaves
    // to the route pass
    dsdx_ok(abi::ensure_routed(P, s), "route");
    const Layout Lay = P->layout();
    const PlanBuffers B = P->buffers();
    Layout LD = Lay;
    LD.descending = 1;   // reverse_halo's walk order
    double* out = nullptr;
    ok(hipMalloc(&out, std::max<uint64_t>(owners, 4096) * sizeof(double)), "malloc");
    constexpr int W = 4;
    const dim3 tiles = tile_grid(Lay, W);

    struct Probe {
        const char* name;
        double bytes;   // the probe's own useful bytes per launch
        int rewrite;    // the arrays rewritten before the second timing: 0 x, y, z; 6 fx, fy, fz; -1 none
        double ms, after_write_ms;
    };
    std::vector<Probe> probes = {
        {"stream_read_xyz", 24.0 * n, 0, 0.0, 0.0},
        {"owner_read_f", 24.0 * owners, 6, 0.0, 0.0},
        {"owner_rmw_f", 48.0 * owners, 6, 0.0, 0.0},
        {"owner_gather7", 72.0 * owners, -1, 0.0, 0.0},
        {"dense_write13", 96.0 * ghosts, -1, 0.0, 0.0},
        {"gather_write_dense", 168.0 * owners, -1, 0.0, 0.0},
        {"owner_rmw_first", 72.0 * owners, 6, 0.0, 0.0},
        {"library_copy", 168.0 * ghosts, -1, 0.0, 0.0},
        {"library_reverse", 24.0 * ghosts + 48.0 * owners, 6, 0.0, 0.0},
    };
    auto run = [&](int k) {
        switch (k) {
        case 0: hipLaunchKernelGGL(stream_read_xyz, dim3(4096), dim3(256), 0, s, a[0], a[1], a[2], n, out); break;
        case 1: hipLaunchKernelGGL((owner_read_f<W>), tiles, dim3(W * kWave), 0, s, a[6], a[7], a[8], out, Lay.desc,
                                   Lay.tile_replica, LD, B); break;
        case 2: hipLaunchKernelGGL((owner_rmw_f<W>), tiles, dim3(W * kWave), 0, s, a[6], a[7], a[8], Lay.desc,
                                   Lay.tile_replica, LD, B); break;
        case 3: hipLaunchKernelGGL((owner_gather7<W>), tiles, dim3(W * kWave), 0, s, p, out, Lay.desc,
                                   Lay.tile_replica, Lay, B); break;
        case 4: hipLaunchKernelGGL(dense_write13, dim3(blocks(ghosts)), dim3(256), 0, s, p, n, ghosts); break;
        case 5: hipLaunchKernelGGL((gather_write_dense<W>), tiles, dim3(W * kWave), 0, s, p, n, Lay.desc,
                                   Lay.tile_replica, Lay, B); break;
        case 6: hipLaunchKernelGGL((owner_rmw_first<W>), tiles, dim3(W * kWave), 0, s, a[6], a[7], a[8], Lay.desc,
                                   Lay.tile_replica, LD, B); break;
        case 7: ok(launch_copy(p, FieldParams{}, Lay, B, P->shape, s), "library copy"); break;
        case 8: ok(launch_reverse_forces(a[6], a[7], a[8], Lay, B, P->shape, s), "library reverse"); break;
        }
        ok(hipGetLastError(), probes[k].name);
    };
    hipEvent_t e0, e1;
    ok(hipEventCreate(&e0), "event");
    ok(hipEventCreate(&e1), "event");
    auto timed = [&](int k, bool rewrite) {
        float total = 0.0f;
        for (int it = 0; it < iters; ++it) {
            if (rewrite) {
                const int r = probes[k].rewrite;
                hipLaunchKernelGGL(rewrite3, dim3(blocks(n)), dim3(256), 0, s, a[r], a[r + 1], a[r + 2], n,
                                   static_cast<uint64_t>(r), r == 0 ? L : 1.0);
                ok(hipGetLastError(), "rewrite3");
            }
            ok(hipEventRecord(e0, s), "record");
            run(k);
            ok(hipEventRecord(e1, s), "record");
            ok(hipEventSynchronize(e1), "sync");
            float ms = 0.0f;
            ok(hipEventElapsedTime(&ms, e0, e1), "elapsed");
            total += ms;
        }
        return total / iters;
    };
    for (size_t k = 0; k < probes.size(); ++k) {
        run(static_cast<int>(k));   // warm-up
        probes[k].ms = timed(static_cast<int>(k), false);
        if (probes[k].rewrite >= 0) probes[k].after_write_ms = timed(static_cast<int>(k), true);
    }
    std::printf("{\"particles\": %llu, \"box\": %.4g, \"ghosts\": %llu, \"owners\": %llu, \"iters\": %d, \"probes\": {",
                static_cast<unsigned long long>(n), L, static_cast<unsigned long long>(ghosts),
                static_cast<unsigned long long>(owners), iters);
    for (size_t k = 0; k < probes.size(); ++k) {
        std::printf("%s\"%s\": {\"us\": %.2f, \"bytes\": %.0f, \"GBps\": %.1f", k ? ", " : "", probes[k].name,
                    1e3 * probes[k].ms, probes[k].bytes, probes[k].bytes / (probes[k].ms * 1e-3) / 1e9);
        if (probes[k].rewrite >= 0) std::printf(", \"after_write_us\": %.2f", 1e3 * probes[k].after_write_ms);
        std::printf("}");
    }
    std::printf("}}\n");
    dsdx_ok(dsdx_plan_destroy(plan), "plan destroy");
    ok(hipStreamDestroy(s), "stream");
    return 0;
}
