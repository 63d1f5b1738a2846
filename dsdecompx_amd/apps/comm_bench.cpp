// comm_bench.cpp -- timing driver of the drop-in C++ API on one MI355X, the
// counterpart of the reference's examples/communication_test/main.cu:
// every step moves all particles by N(0, sigma) per axis, then runs
// migrate -> exchange_halo (rebuild) -> [forces U[-1,1] on owned + ghosts]
// -> reverse_halo through dsdecompx::Decomp on a HIP stream.
//
// HIP events on that stream bracket migrate + exchange_halo and reverse_halo;
// their sum over `steps` steps is the hot-path time.  A second pass of
// `kernel_steps` steps runs with the library's per-kernel timers on
// (dsdx_timing_*: each dispatch carries its own start / stop events) and
// gives each kernel's average duration.  An optional plain copy of
// `calibration_bytes` at the end is the reference dispatch for rocprofv3's
// FETCH_SIZE / WRITE_SIZE counters.  Prints one JSON object on stdout.
//
// usage: comm_bench particles L width sigma z_periodic steps warmup seed
//                   [calibration_bytes [kernel_steps [mode]]]
// mode "split" (default): migrate + exchange_halo(rebuild) + reverse_halo.
// mode "fused": Decomp::migrate_and_exchange_halo instead of migrate +
// exchange_halo (same results, one pass over the positions).
// modes "refresh" / "refresh_pos": the reference benchmark's non-rebuild step
// (examples/communication_test/main.cu:232-236): the plan built in the first
// warm-up step is reused, exchange_halo(rebuild = false), full records or
// positions only, then reverse_halo; no migrate.
// mode "fields": positions in 64 Gaussian clusters (sigma 4, centres uniform,
// wrapped into the box) and two registered fields that travel with the
// particles and into the ghosts -- "density" (f64, uniform [0, 2]) and "vec"
// (3 x f64, uniform [-1, 1]); the step is the split one plus
// reverse_exchange(Sum) of both fields after reverse_halo.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "dsdecompx.hpp"
#include "dsdx_hip.h"

using namespace dsdecompx;

namespace {

void ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "HIP error %s: %s\n", what, hipGetErrorString(e));
        std::exit(3);
    }
}

__device__ __forceinline__ uint64_t mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [0, 1) from (seed, stream, index)
__device__ __forceinline__ double unit(uint64_t seed, uint64_t stream, uint64_t i)
{
    const uint64_t z = mix(seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull + i);
    return static_cast<double>(z >> 11) * 0x1.0p-53;
}

__global__ void place(double* x, double* y, double* z, double* vx, double* vy, double* vz,
                      double* m, double* q, int* t, int* id, uint64_t n, double L, uint64_t seed,
                      int clusters)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    if (clusters) {
        // a Gaussian (sigma 4) around one of `clusters` uniform centres, wrapped
        const uint64_t c = static_cast<uint64_t>(unit(seed, 6, i) * clusters);
        double* axis[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) {
            const double centre = L * unit(seed, 7 + a, c);
            const double u1 = 1.0 - unit(seed, 10 + 2 * a, i);
            const double u2 = unit(seed, 11 + 2 * a, i);
            double v = centre + 4.0 * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
            v -= floor(v / L) * L;
            axis[a][i] = v < L ? v : 0.0;
        }
    } else {
        x[i] = L * unit(seed, 0, i);
        y[i] = L * unit(seed, 1, i);
        z[i] = L * unit(seed, 2, i);
    }
    vx[i] = unit(seed, 3, i) - 0.5;
    vy[i] = unit(seed, 4, i) - 0.5;
    vz[i] = unit(seed, 5, i) - 0.5;
    m[i] = 1.0;
    q[i] = 0.0;
    t[i] = 0;
    id[i] = static_cast<int>(i);
}

// Box-Muller displacement N(0, sigma) per axis; with z non-periodic the
// application reflects z at the walls 0 and L (the library only wraps
// periodic axes).
__global__ void move(double* x, double* y, double* z, uint64_t n, double sigma, uint64_t seed,
                     uint64_t step, int zper, double L)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    double* axis[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        const double u1 = 1.0 - unit(seed, 16 + 2 * a + 8 * step, i);
        const double u2 = unit(seed, 17 + 2 * a + 8 * step, i);
        double v = axis[a][i] + sigma * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        if (a == 2 && !zper) {
            if (v < 0.0) v = -v;
            if (v > L) v = 2.0 * L - v;
        }
        axis[a][i] = v;
    }
}

__global__ void fill_fields(double* density, double* vec, uint64_t n, uint64_t seed)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    density[i] = 2.0 * unit(seed, 30, i);
    for (int c = 0; c < 3; ++c) vec[3 * i + c] = 2.0 * unit(seed, 31 + c, i) - 1.0;
}

__global__ void fill_forces(double* fx, double* fy, double* fz, uint64_t count, uint64_t seed,
                            uint64_t step)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= count) return;
    fx[i] = 2.0 * unit(seed, 1000 + 3 * step, i) - 1.0;
    fy[i] = 2.0 * unit(seed, 1001 + 3 * step, i) - 1.0;
    fz[i] = 2.0 * unit(seed, 1002 + 3 * step, i) - 1.0;
}

// Particles within `w` of a face that has images: the distinct owners U of
// the halo (bench bookkeeping for the reverse-halo byte count).
__global__ void count_boundary(const double* x, const double* y, const double* z, uint64_t n,
                               double lo, double hi, int zper, unsigned long long* out)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    bool b = false;
    if (i < n) {
        b = x[i] < lo || x[i] > hi || y[i] < lo || y[i] > hi;
        if (zper) b = b || z[i] < lo || z[i] > hi;
    }
    const unsigned long long ballot = __ballot(b);
    if ((threadIdx.x & 63) == 0 && ballot) atomicAdd(out, static_cast<unsigned long long>(__popcll(ballot)));
}

// Coordinates that migrate's wrap rewrites: those outside [0, L) on a
// periodic axis (the box starts at 0).
__global__ void count_outside(const double* x, const double* y, const double* z, uint64_t n, double L,
                              int zper, unsigned long long* out)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    unsigned c = 0;
    if (i < n) {
        c = (x[i] < 0.0 || x[i] >= L) + (y[i] < 0.0 || y[i] >= L);
        if (zper) c += (z[i] < 0.0 || z[i] >= L);
    }
    for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, static_cast<unsigned long long>(c));
}

// A plain copy of known traffic (n doubles read, n written): rocprofv3's
// FETCH_SIZE / WRITE_SIZE of this dispatch scale the counters of the halo
// kernels in the same run to bytes.
__global__ void calibration_copy(const double* __restrict__ src, double* __restrict__ dst, uint64_t n)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

unsigned grid(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 9) {
        std::fprintf(stderr,
                     "usage: %s particles L width sigma z_periodic steps warmup seed [calibration_bytes "
                     "[kernel_steps [split|fused|refresh|refresh_pos|fields]]]\n",
                     argv[0]);
        return 2;
    }
    const uint64_t n = std::strtoull(argv[1], nullptr, 10);
    const double L = std::atof(argv[2]);
    const double w = std::atof(argv[3]);
    const double sigma = std::atof(argv[4]);
    const bool zper = std::atoi(argv[5]) != 0;
    const int steps = std::atoi(argv[6]);
    const int warmup = std::atoi(argv[7]);
    const uint64_t seed = std::strtoull(argv[8], nullptr, 10);
    const uint64_t calibration = argc > 9 ? std::strtoull(argv[9], nullptr, 10) / 8 : 0;
    const int kernel_steps = argc > 10 ? std::atoi(argv[10]) : steps;
    const std::string mode = argc > 11 ? argv[11] : "split";
    if (mode != "split" && mode != "fused" && mode != "refresh" && mode != "refresh_pos" && mode != "fields") {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 2;
    }
    const bool refresh = mode == "refresh" || mode == "refresh_pos";
    const bool fields = mode == "fields";

    Config config;
    config.domain.upper = {L, L, L};
    config.domain.periodic = {true, true, zper};
    config.halo.width = w;
    Decomp decomp(MPI_COMM_WORLD, config);
    hipStream_t s = nullptr;
    ok(hipStreamCreate(&s), "stream");

    ParticleInfo p = ParticleInfo::allocate(MemorySpace::CudaDevice, n);
    hipLaunchKernelGGL(place, dim3(grid(n)), dim3(256), 0, s, p.x_data(), p.y_data(), p.z_data(),
                       p.vx_data(), p.vy_data(), p.vz_data(), p.mass_data(), p.charge_data(),
                       p.type_data(), p.global_id_data(), n, L, seed, fields ? 64 : 0);
    // registered fields (mode "fields"): application-owned device arrays with
    // room for the ghosts
    double* density = nullptr;
    double* vec = nullptr;
    const uint64_t field_capacity = 2 * n + 4096;
    if (fields) {
        ok(hipMalloc(&density, field_capacity * sizeof(double)), "malloc");
        ok(hipMalloc(&vec, 3 * field_capacity * sizeof(double)), "malloc");
        hipLaunchKernelGGL(fill_fields, dim3(grid(n)), dim3(256), 0, s, density, vec, n, seed);
        p.add_field<double>("density", density, field_capacity);
        p.add_field<double>("vec", vec, field_capacity, 3);
    }
    const std::vector<ReverseExchangeField> reduced = {{"density", ReduceOp::Sum}, {"vec", ReduceOp::Sum}};
    ok(hipStreamSynchronize(s), "place");

    // Timing-only events: without the system-scope fence (cache write-back
    // and invalidation) that an event otherwise performs, whose cost would
    // sit inside the measured interval (4 us per step at 16 M particles).
    hipEvent_t ev[4];
    for (auto& e : ev) ok(hipEventCreateWithFlags(&e, hipEventDisableSystemFence), "event");
    uint64_t step = 0;
    uint64_t last_ghosts = 0;
    // One pass of `count` steps; returns the hot-path milliseconds summed over
    // the steps and adds each step's ghost count to *ghosts.
    // `wraps`: counts the coordinates each step's migrate rewrites (an extra
    // kernel, so only in a pass of its own).
    unsigned long long* d_wraps = nullptr;
    ok(hipMalloc(&d_wraps, sizeof(*d_wraps)), "malloc");
    ok(hipMemsetAsync(d_wraps, 0, sizeof(*d_wraps), s), "memset");
    auto pass = [&](int count, bool timed, double* ghosts, bool wraps = false) {
        double hot = 0.0;
        for (int k = 0; k < count; ++k, ++step) {
            hipLaunchKernelGGL(move, dim3(grid(n)), dim3(256), 0, s, p.x_data(), p.y_data(), p.z_data(),
                               n, sigma, seed, step, zper ? 1 : 0, L);
            if (wraps && !(refresh && step > 0))
                hipLaunchKernelGGL(count_outside, dim3(grid(n)), dim3(256), 0, s, p.x_data(), p.y_data(),
                                   p.z_data(), n, L, zper ? 1 : 0, d_wraps);
            if (timed) ok(hipEventRecord(ev[0], s), "record");
            HaloResult r;
            if (mode == "fused") {
                r = decomp.migrate_and_exchange_halo(p, {true, s});
            } else if (refresh && step > 0) {
                r = decomp.exchange_halo(p, {false, s, mode == "refresh_pos"});
            } else {
                decomp.migrate(p, s);
                r = decomp.exchange_halo(p, {true, s});
            }
            if (timed) ok(hipEventRecord(ev[1], s), "record");
            const uint64_t total = n + r.ghost_count;
            hipLaunchKernelGGL(fill_forces, dim3(grid(total)), dim3(256), 0, s, p.fx_data(), p.fy_data(),
                               p.fz_data(), total, seed, step);
            if (timed) ok(hipEventRecord(ev[2], s), "record");
            decomp.reverse_halo(p, s);
            if (fields) decomp.reverse_exchange(p, reduced, s);
            if (timed) {
                ok(hipEventRecord(ev[3], s), "record");
                ok(hipEventSynchronize(ev[3]), "sync");
                float a = 0.0f, b = 0.0f;
                ok(hipEventElapsedTime(&a, ev[0], ev[1]), "elapsed");
                ok(hipEventElapsedTime(&b, ev[2], ev[3]), "elapsed");
                hot += a + b;
            }
            if (ghosts) *ghosts += static_cast<double>(r.ghost_count);
            last_ghosts = r.ghost_count;
        }
        return hot;
    };

    pass(warmup, false, nullptr);
    ok(hipStreamSynchronize(s), "sync");
    // Timed pass: the library's per-kernel timers stay off (a timed dispatch
    // is preceded by several microseconds of idle time).
    double ghosts_sum = 0.0;
    const auto wall0 = std::chrono::steady_clock::now();
    const double hot_ms = pass(steps, true, &ghosts_sum);
    ok(hipStreamSynchronize(s), "sync");
    const double wall_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    // Kernel pass: the same steps again with every kernel's dispatch timed.
    double kernel_pass_hot_ms = 0.0;
    if (kernel_steps > 0) {
        dsdx_timing_reset();
        dsdx_timing_enable(1);
        kernel_pass_hot_ms = pass(kernel_steps, true, nullptr);
        ok(hipStreamSynchronize(s), "sync");
        dsdx_timing_enable(0);
    }

    // Count pass: coordinates rewritten by migrate per step.
    const int count_steps = std::min(steps, 20);
    pass(count_steps, false, nullptr, true);
    unsigned long long wraps = 0;
    ok(hipMemcpyAsync(&wraps, d_wraps, sizeof(wraps), hipMemcpyDeviceToHost, s), "copy");
    ok(hipStreamSynchronize(s), "sync");
    ok(hipFree(d_wraps), "free");

    unsigned long long* d_u = nullptr;
    ok(hipMalloc(&d_u, sizeof(*d_u)), "malloc");
    ok(hipMemsetAsync(d_u, 0, sizeof(*d_u), s), "memset");
    hipLaunchKernelGGL(count_boundary, dim3(grid(n)), dim3(256), 0, s, p.x_data(), p.y_data(),
                       p.z_data(), n, w, L - w, zper ? 1 : 0, d_u);
    unsigned long long owners = 0;
    ok(hipMemcpyAsync(&owners, d_u, sizeof(owners), hipMemcpyDeviceToHost, s), "copy");
    ok(hipStreamSynchronize(s), "sync");
    ok(hipFree(d_u), "free");
    if (calibration) {
        double *src = nullptr, *dst = nullptr;
        ok(hipMalloc(&src, calibration * 8), "malloc");
        ok(hipMalloc(&dst, calibration * 8), "malloc");
        hipLaunchKernelGGL(calibration_copy, dim3(grid(calibration)), dim3(256), 0, s, src, dst, calibration);
        ok(hipStreamSynchronize(s), "calibration copy");
        ok(hipFree(src), "free");
        ok(hipFree(dst), "free");
    }

    const char* kernels[] = {"migrate_wrap", "halo_classify", "halo_scan", "halo_scatter", "halo_copy",
                             "reverse_halo", "migrate_classify", "halo_refresh", "reverse_exchange"};
    std::printf("{\"particles\": %llu, \"steps\": %d, \"warmup\": %d, \"hot_ms_per_step\": %.6f, "
                "\"wall_ms_per_step\": %.6f, \"ghosts_per_step\": %.1f, \"last_ghosts\": %llu, "
                "\"owners\": %llu, \"kernel_steps\": %d, \"kernel_pass_hot_ms_per_step\": %.6f, "
                "\"wrapped_coords_per_step\": %.1f, \"kernels\": {",
                static_cast<unsigned long long>(n), steps, warmup, hot_ms / steps, wall_ms / steps,
                ghosts_sum / steps, static_cast<unsigned long long>(last_ghosts), owners, kernel_steps,
                kernel_steps > 0 ? kernel_pass_hot_ms / kernel_steps : 0.0,
                count_steps > 0 ? static_cast<double>(wraps) / count_steps : 0.0);
    for (int k = 0; k < 9; ++k) {
        double ms = 0.0;
        uint64_t launches = 0;
        dsdx_timing_read(kernels[k], &ms, &launches);
        std::printf("%s\"%s\": {\"total_ms\": %.6f, \"launches\": %llu}", k ? ", " : "", kernels[k], ms,
                    static_cast<unsigned long long>(launches));
    }
    std::printf("}}\n");
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (density) (void)hipFree(density);
    if (vec) (void)hipFree(vec);
    (void)hipStreamDestroy(s);
    return 0;
}
