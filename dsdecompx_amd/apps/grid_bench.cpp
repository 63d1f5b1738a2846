// grid_bench.cpp -- timing driver of process grids (dsdx_grid_*) on one
// MI355X: the particles of one periodic box split over a dims[0] x dims[1] x
// dims[2] Cartesian decomposition, every subdomain on this GPU.  Each step
// moves every owned particle by N(0, sigma) per axis, then runs migrate ->
// halo build -> [forces U[-1,1] on owned + ghosts] -> reverse halo, the step
// of the reference's examples/communication_test/main.cu:222-236 on P ranks.
//
// HIP events on the stream bracket migrate + halo build and the reverse
// halo; their sum over `steps` steps is the hot-path time.  A second pass of
// `kernel_steps` steps runs with the library's per-kernel timers on.  Prints
// one JSON object on stdout.
//
// usage: grid_bench particles L width sigma dx dy dz steps warmup seed [kernel_steps]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dsdx_hip.h"

namespace {

void ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "HIP error %s: %s\n", what, hipGetErrorString(e));
        std::exit(3);
    }
}

void check(int rc, const char* what)
{
    if (rc != DSDX_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, dsdx_last_error());
        std::exit(3);
    }
}

__device__ __forceinline__ uint64_t mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double unit(uint64_t seed, uint64_t stream, uint64_t i)
{
    const uint64_t z = mix(seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull + i);
    return static_cast<double>(z >> 11) * 0x1.0p-53;
}

// Particle i of the whole box, uniform in [0, L)^3, written to subdomain
// i / per (an arbitrary first owner: the first migration sorts them out).
__global__ void place(dsdx_particles p, uint64_t n, uint64_t per, uint64_t cap, double L, uint64_t seed)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const uint64_t j = (i / per) * cap + i % per;
    p.x[j] = L * unit(seed, 0, i);
    p.y[j] = L * unit(seed, 1, i);
    p.z[j] = L * unit(seed, 2, i);
    p.vx[j] = unit(seed, 3, i) - 0.5;
    p.vy[j] = unit(seed, 4, i) - 0.5;
    p.vz[j] = unit(seed, 5, i) - 0.5;
    p.mass[j] = 1.0;
    p.charge[j] = 0.0;
    p.type[j] = 0;
    p.global_id[j] = static_cast<int32_t>(i);
}

// N(0, sigma) per axis (Box-Muller) on the owned particles of every
// subdomain: owned[r] of them from slot r * cap.
__global__ void move(dsdx_particles p, const uint64_t* owned, uint64_t cap, uint32_t ranks,
                     double sigma, uint64_t seed, uint64_t step)
{
    const uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    const uint64_t r = j / cap;
    if (r >= ranks || j % cap >= owned[r]) return;
    double* axis[3] = {p.x, p.y, p.z};
    for (int a = 0; a < 3; ++a) {
        const double u1 = 1.0 - unit(seed, 16 + 2 * a + 8 * step, j);
        const double u2 = unit(seed, 17 + 2 * a + 8 * step, j);
        axis[a][j] += sigma * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

__global__ void fill_forces(dsdx_particles p, const uint64_t* count, uint64_t cap, uint32_t ranks,
                            uint64_t seed, uint64_t step)
{
    const uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    const uint64_t r = j / cap;
    if (r >= ranks || j % cap >= count[r]) return;
    p.fx[j] = 2.0 * unit(seed, 1000 + 3 * step, j) - 1.0;
    p.fy[j] = 2.0 * unit(seed, 1001 + 3 * step, j) - 1.0;
    p.fz[j] = 2.0 * unit(seed, 1002 + 3 * step, j) - 1.0;
}

unsigned grid(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

template <typename T>
T* zeros(uint64_t count, std::vector<void*>* keep)
{
    T* a = nullptr;
    ok(hipMalloc(&a, count * sizeof(T)), "malloc");
    ok(hipMemset(a, 0, count * sizeof(T)), "memset");
    keep->push_back(a);
    return a;
}

struct Storage {
    dsdx_particles p{};
    std::vector<void*> arrays;

    explicit Storage(uint64_t slots)
    {
        for (double** a : {&p.x, &p.y, &p.z, &p.vx, &p.vy, &p.vz, &p.fx, &p.fy, &p.fz, &p.mass, &p.charge})
            *a = zeros<double>(slots, &arrays);
        p.type = zeros<int32_t>(slots, &arrays);
        p.global_id = zeros<int32_t>(slots, &arrays);
    }
    ~Storage()
    {
        for (void* a : arrays) (void)hipFree(a);
    }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 11) {
        std::fprintf(stderr, "usage: %s particles L width sigma dx dy dz steps warmup seed [kernel_steps]\n",
                     argv[0]);
        return 2;
    }
    const uint64_t n = std::strtoull(argv[1], nullptr, 10);
    const double L = std::atof(argv[2]);
    const double w = std::atof(argv[3]);
    const double sigma = std::atof(argv[4]);
    const int32_t dims[3] = {std::atoi(argv[5]), std::atoi(argv[6]), std::atoi(argv[7])};
    const int steps = std::atoi(argv[8]);
    const int warmup = std::atoi(argv[9]);
    const uint64_t seed = std::strtoull(argv[10], nullptr, 10);
    const int kernel_steps = argc > 11 ? std::atoi(argv[11]) : steps;
    const uint32_t ranks = static_cast<uint32_t>(dims[0] * dims[1] * dims[2]);
    const uint64_t per = (n + ranks - 1) / ranks;
    // owned share plus the halo (boundary fraction of a subdomain) and slack
    double boundary = 1.0;
    for (int a = 0; a < 3; ++a) boundary *= 1.0 - 2.0 * w * dims[a] / L;
    const uint64_t cap = static_cast<uint64_t>(per * (1.15 + 2.0 * (1.0 - boundary))) + 4096;

    hipStream_t s = nullptr;
    ok(hipStreamCreate(&s), "stream");
    dsdx_box box{};
    for (int a = 0; a < 3; ++a) {
        box.upper[a] = L;
        box.periodic[a] = 1;
    }
    box.width = w;
    box.edges_and_corners = 1;
    dsdx_grid* g = nullptr;
    check(dsdx_grid_create(&box, dims, cap, s, &g), "grid create");
    Storage a(ranks * cap), b(ranks * cap);
    Storage* cur = &a;
    Storage* spare = &b;
    hipLaunchKernelGGL(place, dim3(grid(n)), dim3(256), 0, s, cur->p, n, per, cap, L, seed);
    std::vector<uint64_t> owned(ranks), ghosts(ranks), owners(ranks), count(ranks);
    for (uint32_t r = 0; r < ranks; ++r) owned[r] = std::min<uint64_t>(per, n - std::min<uint64_t>(n, r * per));
    check(dsdx_grid_set_owned(g, owned.data(), s), "set owned");
    uint64_t* d_counts = nullptr;
    ok(hipMalloc(&d_counts, 2 * ranks * sizeof(uint64_t)), "malloc");
    ok(hipStreamSynchronize(s), "place");

    // Timing-only events: without the system-scope fence (cache write-back
    // and invalidation) that an event otherwise performs, whose cost would
    // sit inside the measured interval (4 us per step at 16 M particles).
    hipEvent_t ev[4];
    for (auto& e : ev) ok(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "event");
    uint64_t step = 0;
    double ghosts_sum = 0.0, owners_sum = 0.0;
    auto pass = [&](int k_steps, bool timed, bool count_ghosts) {
        double hot = 0.0;
        for (int k = 0; k < k_steps; ++k, ++step) {
            check(dsdx_grid_counts(g, owned.data(), nullptr, nullptr), "counts");
            ok(hipMemcpyAsync(d_counts, owned.data(), ranks * 8, hipMemcpyHostToDevice, s), "copy");
            hipLaunchKernelGGL(move, dim3(grid(ranks * cap)), dim3(256), 0, s, cur->p, d_counts, cap, ranks,
                               sigma, seed, step);
            if (timed) ok(hipEventRecord(ev[0], s), "record");
            check(dsdx_grid_migrate(g, &cur->p, &spare->p, nullptr, nullptr, 0, s), "migrate");
            std::swap(cur, spare);
            check(dsdx_grid_halo_build(g, &cur->p, nullptr, 0, s, ghosts.data()), "halo build");
            if (timed) ok(hipEventRecord(ev[1], s), "record");
            check(dsdx_grid_counts(g, owned.data(), nullptr, owners.data()), "counts");
            for (uint32_t r = 0; r < ranks; ++r) count[r] = owned[r] + ghosts[r];
            ok(hipMemcpyAsync(d_counts + ranks, count.data(), ranks * 8, hipMemcpyHostToDevice, s), "copy");
            hipLaunchKernelGGL(fill_forces, dim3(grid(ranks * cap)), dim3(256), 0, s, cur->p, d_counts + ranks,
                               cap, ranks, seed, step);
            if (timed) ok(hipEventRecord(ev[2], s), "record");
            check(dsdx_grid_reverse_halo(g, cur->p.fx, cur->p.fy, cur->p.fz, s), "reverse halo");
            if (timed) {
                ok(hipEventRecord(ev[3], s), "record");
                ok(hipEventSynchronize(ev[3]), "sync");
                float t0 = 0.0f, t1 = 0.0f;
                ok(hipEventElapsedTime(&t0, ev[0], ev[1]), "elapsed");
                ok(hipEventElapsedTime(&t1, ev[2], ev[3]), "elapsed");
                hot += t0 + t1;
            }
            if (count_ghosts)
                for (uint32_t r = 0; r < ranks; ++r) {
                    ghosts_sum += static_cast<double>(ghosts[r]);
                    owners_sum += static_cast<double>(owners[r]);
                }
        }
        return hot;
    };

    pass(warmup, false, false);
    ok(hipStreamSynchronize(s), "sync");
    const double hot_ms = pass(steps, true, true);
    double kernel_hot_ms = 0.0;
    if (kernel_steps > 0) {
        dsdx_timing_reset();
        dsdx_timing_enable(1);
        kernel_hot_ms = pass(kernel_steps, true, false);
        ok(hipStreamSynchronize(s), "sync");
        dsdx_timing_enable(0);
    }
    uint64_t owned_min = ~0ull, owned_max = 0;
    for (uint32_t r = 0; r < ranks; ++r) {
        owned_min = std::min(owned_min, owned[r]);
        owned_max = std::max(owned_max, owned[r]);
    }
    const char* kernels[] = {"grid_migrate_count", "grid_migrate_scatter", "halo_classify", "halo_scan",
                             "grid_exchange_counts", "halo_scatter", "halo_copy", "reverse_halo"};
    std::printf("{\"particles\": %llu, \"dims\": [%d, %d, %d], \"capacity_per_rank\": %llu, \"steps\": %d, "
                "\"warmup\": %d, \"hot_ms_per_step\": %.6f, \"ghosts_per_step\": %.1f, \"owners_per_step\": %.1f, "
                "\"owned_min\": %llu, \"owned_max\": %llu, \"kernel_steps\": %d, "
                "\"kernel_pass_hot_ms_per_step\": %.6f, \"kernels\": {",
                static_cast<unsigned long long>(n), dims[0], dims[1], dims[2],
                static_cast<unsigned long long>(cap), steps, warmup, hot_ms / steps, ghosts_sum / steps,
                owners_sum / steps,
                static_cast<unsigned long long>(owned_min), static_cast<unsigned long long>(owned_max),
                kernel_steps, kernel_steps > 0 ? kernel_hot_ms / kernel_steps : 0.0);
    for (int k = 0; k < 8; ++k) {
        double ms = 0.0;
        uint64_t launches = 0;
        dsdx_timing_read(kernels[k], &ms, &launches);
        std::printf("%s\"%s\": {\"total_ms\": %.6f, \"launches\": %llu}", k ? ", " : "", kernels[k], ms,
                    static_cast<unsigned long long>(launches));
    }
    std::printf("}}\n");
    for (auto& e : ev) (void)hipEventDestroy(e);
    (void)hipFree(d_counts);
    check(dsdx_grid_destroy(g), "grid destroy");
    (void)hipStreamDestroy(s);
    return 0;
}
