// batch_bench.cpp -- the batched replica workload through the C ABI, timed
// and checked in C++ with HIP's own graph API (no framework in between).
//
// Replicas: `replicas` periodic boxes, particle counts log-uniform in
// [2048, 262144], density U[0.6, 1.2] (side (n / density)^(1/3)), halo width
// 2.8, displacement N(0, 0.1) per axis per step; every replica's storage
// holds its owned particles and room for 1.35 times the expected shell of
// ghosts.  One step, identical in every form: the application moves the
// owned particles (one launch over all replicas), then migrate ->
// exchange_halo (rebuild) -> the application fills the forces (one launch)
// -> reverse_halo; a one-thread kernel advances the step counter the
// displacement reads, so a captured step replays with fresh displacements.
//
// Forms (each unfused, migrate then exchange_halo, and fused,
// migrate_and_exchange_halo):
//   loop    the plain per-box loop: one single-box plan per replica, one
//           C-ABI call per replica and phase (a rebuild waits for its ghost
//           count, as the single-box call does)
//   batch   one launch per phase over every replica (dsdx_batch_*), eager
//   graph   the batch step captured once with hipStreamBeginCapture and
//           replayed with hipGraphLaunch
// usage: batch_bench replicas steps warmup seed [bench|check|graph]
//   bench: the batch's launch shape tuned first (items per lane, waves per
//          block, scan threads: the fastest eager fused step), then
//          replica-steps per second of every form, one JSON line.
//   check: every form runs `steps` steps from the same start; all 13 arrays
//          of every form must equal the unfused loop's bit for bit.  Then one
//          box of 100,000 particles: dsdx_plan_reserve, capture of
//          migrate -> dsdx_halo_build (no count wait) -> forces ->
//          reverse_halo in global capture mode (any allocation or host wait
//          inside would fail the capture), and `steps` replays against the
//          same steps run eagerly on a second plan.  Exit status 1 on any
//          mismatch.
//   graph: the captured batch step only, unfused then fused, at the launch
//          shape the environment gives (DSDX_ITEMS, DSDX_WAVES,
//          DSDX_SCAN_THREADS; no tuning), so that a kernel trace of the run
//          holds the batch's kernels alone.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
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

void dsdx(int rc, const char* what)
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

__device__ __forceinline__ double unit(uint64_t seed, uint64_t stream, uint64_t i)
{
    const uint64_t z = mix(seed * 0x9E3779B97F4A7C15ull + stream * 0xD1B54A32D192ED03ull + i);
    return static_cast<double>(z >> 11) * 0x1.0p-53;
}

// Owned particles of one replica at [off, off + n): uniform in [0, L)^3.
__global__ void place(double* x, double* y, double* z, double* vx, double* vy, double* vz, double* m,
                      double* q, int* t, int* id, uint64_t off, uint64_t n, double L, uint64_t seed)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const uint64_t j = off + i;
    x[j] = L * unit(seed, 0, j);
    y[j] = L * unit(seed, 1, j);
    z[j] = L * unit(seed, 2, j);
    vx[j] = unit(seed, 3, j) - 0.5;
    vy[j] = unit(seed, 4, j) - 0.5;
    vz[j] = unit(seed, 5, j) - 0.5;
    m[j] = 1.0;
    q[j] = unit(seed, 6, j);
    t[j] = static_cast<int>(j & 3);
    id[j] = static_cast<int>(i);
}

// The application's displacement, N(0, sigma) per axis, of the owned slots.
__global__ void move(double* x, double* y, double* z, const uint8_t* owned, uint64_t total, double sigma,
                     uint64_t seed, const uint64_t* step)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= total || !owned[i]) return;
    const uint64_t s = *step;
    double* axis[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        const double u1 = 1.0 - unit(seed, 16 + 2 * a + 8 * s, i);
        const double u2 = unit(seed, 17 + 2 * a + 8 * s, i);
        axis[a][i] = axis[a][i] + sigma * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
}

// The application's forces, U[-1, 1] on every slot.
__global__ void fill_forces(double* fx, double* fy, double* fz, uint64_t total, uint64_t seed,
                            const uint64_t* step)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i >= total) return;
    const uint64_t s = *step;
    fx[i] = 2.0 * unit(seed, 1000 + 3 * s, i) - 1.0;
    fy[i] = 2.0 * unit(seed, 1001 + 3 * s, i) - 1.0;
    fz[i] = 2.0 * unit(seed, 1002 + 3 * s, i) - 1.0;
}

__global__ void advance(uint64_t* step) { *step += 1; }

__global__ void count_diff(const uint64_t* a, const uint64_t* b, uint64_t words, unsigned long long* out)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < words && a[i] != b[i]) atomicAdd(out, 1ull);
}

unsigned grid(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// 13 SoA arrays over `total` slots, allocated as one block.
struct Storage {
    uint64_t total = 0;
    void* block = nullptr;
    double* f64[11] = {};
    int* i32[2] = {};
    explicit Storage(uint64_t n) : total(n)
    {
        ok(hipMalloc(&block, bytes()), "malloc storage");
        auto* at = static_cast<unsigned char*>(block);
        for (auto& p : f64) {
            p = reinterpret_cast<double*>(at);
            at += n * 8;
        }
        for (auto& p : i32) {
            p = reinterpret_cast<int*>(at);
            at += n * 4;
        }
    }
    ~Storage() { (void)hipFree(block); }
    uint64_t bytes() const { return total * (11 * 8 + 2 * 4); }
    dsdx_particles view(uint64_t off = 0) const
    {
        return {f64[0] + off, f64[1] + off, f64[2] + off, f64[3] + off, f64[4] + off,
                f64[5] + off, f64[6] + off, f64[7] + off, f64[8] + off, f64[9] + off,
                f64[10] + off, i32[0] + off, i32[1] + off};
    }
};

struct Replicas {
    std::vector<dsdx_replica> table;
    uint64_t total = 0;
    uint64_t particles = 0;
};

Replicas make_replicas(int count, uint64_t seed)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    Replicas r;
    uint64_t off = 0;
    for (int k = 0; k < count; ++k) {
        const double n = std::exp(std::log(2048.0) + u(rng) * (std::log(262144.0) - std::log(2048.0)));
        const uint64_t owned = static_cast<uint64_t>(n);
        const double rho = 0.6 + 0.6 * u(rng);
        const double L = std::cbrt(static_cast<double>(owned) / rho);
        const double shell = (std::pow(L + 5.6, 3) - std::pow(L, 3)) / std::pow(L, 3);
        dsdx_replica d;
        std::memset(&d, 0, sizeof(d));
        for (int a = 0; a < 3; ++a) {
            d.box.upper[a] = L;
            d.box.periodic[a] = 1;
        }
        d.box.width = 2.8;
        d.box.edges_and_corners = 1;
        d.offset = off;
        d.owned = owned;
        d.capacity = owned + static_cast<uint64_t>(1.35 * shell * static_cast<double>(owned)) + 256;
        off += d.capacity;
        r.particles += owned;
        r.table.push_back(d);
    }
    r.total = off;
    return r;
}

struct Workload {
    const Replicas& R;
    Storage start;
    uint8_t* owned_mask = nullptr;
    uint64_t* step = nullptr;
    uint64_t seed;
    Workload(const Replicas& replicas, uint64_t s, hipStream_t stream) : R(replicas), start(replicas.total), seed(s)
    {
        ok(hipMemsetAsync(start.block, 0, start.bytes(), stream), "clear");
        std::vector<uint8_t> mask(R.total, 0);
        for (const dsdx_replica& d : R.table) {
            hipLaunchKernelGGL(place, dim3(grid(d.owned)), dim3(256), 0, stream, start.f64[0], start.f64[1],
                               start.f64[2], start.f64[3], start.f64[4], start.f64[5], start.f64[9],
                               start.f64[10], start.i32[0], start.i32[1], d.offset, d.owned, d.box.upper[0], seed);
            std::memset(mask.data() + d.offset, 1, d.owned);
        }
        ok(hipMalloc(&owned_mask, R.total), "malloc");
        ok(hipMemcpy(owned_mask, mask.data(), R.total, hipMemcpyHostToDevice), "upload");
        ok(hipMalloc(&step, 8), "malloc");
        ok(hipStreamSynchronize(stream), "place");
    }
    ~Workload()
    {
        (void)hipFree(owned_mask);
        (void)hipFree(step);
    }
    void reset(Storage& s, hipStream_t stream) const
    {
        ok(hipMemcpyAsync(s.block, start.block, s.bytes(), hipMemcpyDeviceToDevice, stream), "restore");
        ok(hipMemsetAsync(step, 0, 8, stream), "step");
    }
    void app_move(const Storage& s, hipStream_t stream) const
    {
        hipLaunchKernelGGL(move, dim3(grid(R.total)), dim3(256), 0, stream, s.f64[0], s.f64[1], s.f64[2],
                           owned_mask, R.total, 0.1, seed, step);
    }
    void app_forces(const Storage& s, hipStream_t stream) const
    {
        hipLaunchKernelGGL(fill_forces, dim3(grid(R.total)), dim3(256), 0, stream, s.f64[6], s.f64[7],
                           s.f64[8], R.total, seed, step);
    }
    void app_advance(hipStream_t stream) const
    {
        hipLaunchKernelGGL(advance, dim3(1), dim3(1), 0, stream, step);
    }
};

struct Loop {
    const Replicas& R;
    std::vector<dsdx_plan*> plans;
    explicit Loop(const Replicas& r) : R(r), plans(r.table.size(), nullptr)
    {
        for (auto& p : plans) dsdx(dsdx_plan_create(&p), "plan create");
    }
    ~Loop()
    {
        for (auto* p : plans) dsdx_plan_destroy(p);
    }
    void step(const Workload& w, const Storage& s, bool fused, hipStream_t stream)
    {
        w.app_move(s, stream);
        for (size_t k = 0; k < plans.size(); ++k) {
            const dsdx_replica& d = R.table[k];
            const dsdx_particles p = s.view(d.offset);
            uint64_t g = 0;
            if (fused) {
                dsdx(dsdx_migrate_halo_build(plans[k], &d.box, &p, nullptr, 0, d.owned, d.capacity, stream, &g),
                     "migrate+halo");
            } else {
                dsdx(dsdx_migrate_wrap(&d.box, p.x, p.y, p.z, d.owned, stream), "migrate");
                dsdx(dsdx_halo_build(plans[k], &d.box, &p, nullptr, 0, d.owned, d.capacity, stream, &g), "halo");
            }
        }
        w.app_forces(s, stream);
        for (size_t k = 0; k < plans.size(); ++k) {
            const dsdx_replica& d = R.table[k];
            const dsdx_particles p = s.view(d.offset);
            dsdx(dsdx_reverse_halo(plans[k], p.fx, p.fy, p.fz, d.owned, stream), "reverse");
        }
        w.app_advance(stream);
    }
};

struct Batch {
    dsdx_batch* b = nullptr;
    Batch(const Replicas& r, hipStream_t stream)
    {
        dsdx(dsdx_batch_create(r.table.data(), static_cast<uint32_t>(r.table.size()), stream, &b), "batch");
    }
    ~Batch() { dsdx_batch_destroy(b); }
    void step(const Workload& w, const Storage& s, bool fused, hipStream_t stream)
    {
        const dsdx_particles p = s.view();
        w.app_move(s, stream);
        if (fused) {
            dsdx(dsdx_batch_migrate_halo_build(b, &p, nullptr, 0, stream), "batch migrate+halo");
        } else {
            dsdx(dsdx_batch_migrate_wrap(b, p.x, p.y, p.z, stream), "batch migrate");
            dsdx(dsdx_batch_halo_build(b, &p, nullptr, 0, stream), "batch halo");
        }
        w.app_forces(s, stream);
        dsdx(dsdx_batch_reverse_halo(b, p.fx, p.fy, p.fz, stream), "batch reverse");
        w.app_advance(stream);
    }
};

// One step of `body` captured on `stream` and instantiated.
template <class Body>
hipGraphExec_t capture(hipStream_t stream, hipStreamCaptureMode mode, Body&& body)
{
    hipGraph_t graph = nullptr;
    ok(hipStreamBeginCapture(stream, mode), "begin capture");
    body();
    ok(hipStreamEndCapture(stream, &graph), "end capture");
    hipGraphExec_t exec = nullptr;
    ok(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate");
    ok(hipGraphDestroy(graph), "graph destroy");
    return exec;
}

template <class F>
double wall_seconds(hipStream_t stream, F&& f)
{
    ok(hipStreamSynchronize(stream), "sync");
    const auto t0 = std::chrono::steady_clock::now();
    f();
    ok(hipStreamSynchronize(stream), "sync");
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

uint64_t differences(const Storage& a, const Storage& b, hipStream_t stream)
{
    unsigned long long* d = nullptr;
    ok(hipMalloc(&d, 8), "malloc");
    ok(hipMemsetAsync(d, 0, 8, stream), "memset");
    const uint64_t words = a.bytes() / 8;
    hipLaunchKernelGGL(count_diff, dim3(grid(words)), dim3(256), 0, stream, static_cast<const uint64_t*>(a.block),
                       static_cast<const uint64_t*>(b.block), words, d);
    unsigned long long n = 0;
    ok(hipMemcpyAsync(&n, d, 8, hipMemcpyDeviceToHost, stream), "copy");
    ok(hipStreamSynchronize(stream), "sync");
    ok(hipFree(d), "free");
    return n;
}

// One box: a captured step right after dsdx_plan_reserve against the same
// steps run eagerly; returns the number of differing 8-byte words.
uint64_t single_box_capture_check(int steps, uint64_t seed, hipStream_t stream)
{
    Replicas one;
    dsdx_replica d;
    std::memset(&d, 0, sizeof(d));
    const uint64_t n = 100000;
    const double L = std::cbrt(static_cast<double>(n));
    for (int a = 0; a < 3; ++a) {
        d.box.upper[a] = L;
        d.box.periodic[a] = 1;
    }
    d.box.width = 2.8;
    d.box.edges_and_corners = 1;
    d.owned = n;
    d.capacity = 2 * n;
    one.table.push_back(d);
    one.total = d.capacity;
    one.particles = n;
    Workload w(one, seed, stream);
    Storage eager(one.total), captured(one.total);
    dsdx_plan* pe = nullptr;
    dsdx_plan* pc = nullptr;
    dsdx(dsdx_plan_create(&pe), "plan");
    dsdx(dsdx_plan_create(&pc), "plan");
    // reserve: builds of n particles with up to n ghosts allocate nothing
    dsdx(dsdx_plan_reserve(pc, n, n), "reserve");
    auto step = [&](dsdx_plan* plan, const Storage& s) {
        const dsdx_particles p = s.view();
        w.app_move(s, stream);
        dsdx(dsdx_migrate_wrap(&d.box, p.x, p.y, p.z, n, stream), "migrate");
        dsdx(dsdx_halo_build(plan, &d.box, &p, nullptr, 0, n, d.capacity, stream, nullptr), "halo");
        w.app_forces(s, stream);
        dsdx(dsdx_reverse_halo(plan, p.fx, p.fy, p.fz, n, stream), "reverse");
        w.app_advance(stream);
    };
    w.reset(eager, stream);
    for (int k = 0; k < steps; ++k) step(pe, eager);
    ok(hipStreamSynchronize(stream), "sync");
    w.reset(captured, stream);
    ok(hipStreamSynchronize(stream), "sync");
    hipGraphExec_t exec = capture(stream, hipStreamCaptureModeGlobal, [&] { step(pc, captured); });
    for (int k = 0; k < steps; ++k) ok(hipGraphLaunch(exec, stream), "graph launch");
    ok(hipStreamSynchronize(stream), "sync");
    const uint64_t diff = differences(eager, captured, stream);
    ok(hipGraphExecDestroy(exec), "exec destroy");
    dsdx_plan_destroy(pe);
    dsdx_plan_destroy(pc);
    return diff;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s replicas steps warmup seed [bench|check|graph]\n", argv[0]);
        return 2;
    }
    const int count = std::atoi(argv[1]);
    const int steps = std::atoi(argv[2]);
    const int warmup = std::atoi(argv[3]);
    const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
    const std::string mode = argc > 5 ? argv[5] : "bench";
    if (count < 1 || steps < 1 || (mode != "bench" && mode != "check" && mode != "graph")) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    hipStream_t stream = nullptr;
    ok(hipStreamCreate(&stream), "stream");
    const Replicas R = make_replicas(count, seed);
    Workload w(R, seed, stream);
    Storage s(R.total);
    Loop loop(R);
    Batch batch(R, stream);

    if (mode == "check") {
        Storage ref(R.total);
        w.reset(ref, stream);
        for (int k = 0; k < steps; ++k) loop.step(w, ref, false, stream);
        ok(hipStreamSynchronize(stream), "sync");
        struct Form {
            const char* name;
            uint64_t diff;
        };
        std::vector<Form> forms;
        w.reset(s, stream);
        for (int k = 0; k < steps; ++k) loop.step(w, s, true, stream);
        forms.push_back({"loop_fused", differences(ref, s, stream)});
        for (int fused = 0; fused < 2; ++fused) {
            w.reset(s, stream);
            for (int k = 0; k < steps; ++k) batch.step(w, s, fused != 0, stream);
            forms.push_back({fused ? "batch_fused" : "batch", differences(ref, s, stream)});
            w.reset(s, stream);
            ok(hipStreamSynchronize(stream), "sync");
            hipGraphExec_t exec = capture(stream, hipStreamCaptureModeGlobal,
                                          [&] { batch.step(w, s, fused != 0, stream); });
            for (int k = 0; k < steps; ++k) ok(hipGraphLaunch(exec, stream), "graph launch");
            forms.push_back({fused ? "graph_fused" : "graph", differences(ref, s, stream)});
            ok(hipGraphExecDestroy(exec), "exec destroy");
        }
        const uint64_t single = single_box_capture_check(steps, seed + 1, stream);
        bool all = single == 0;
        std::printf("{\"replicas\": %d, \"particles\": %llu, \"steps\": %d, \"differing_words\": {", count,
                    static_cast<unsigned long long>(R.particles), steps);
        for (size_t k = 0; k < forms.size(); ++k) {
            std::printf("%s\"%s\": %llu", k ? ", " : "", forms[k].name,
                        static_cast<unsigned long long>(forms[k].diff));
            all = all && forms[k].diff == 0;
        }
        std::printf(", \"single_box_graph_after_reserve\": %llu}, \"bit_exact\": %s}\n",
                    static_cast<unsigned long long>(single), all ? "true" : "false");
        ok(hipStreamDestroy(stream), "stream");
        return all ? 0 : 1;
    }

    if (mode == "graph") {
        int shape[3] = {0, 0, 0};
        dsdx(dsdx_batch_get_shape(batch.b, &shape[0], &shape[1], &shape[2]), "shape");
        double ms[2] = {0.0, 0.0};
        for (int fused = 0; fused < 2; ++fused) {
            w.reset(s, stream);
            ok(hipStreamSynchronize(stream), "sync");
            hipGraphExec_t exec = capture(stream, hipStreamCaptureModeGlobal,
                                          [&] { batch.step(w, s, fused != 0, stream); });
            for (int k = 0; k < warmup; ++k) ok(hipGraphLaunch(exec, stream), "graph launch");
            ms[fused] = 1e3 * wall_seconds(stream, [&] {
                for (int k = 0; k < steps; ++k) ok(hipGraphLaunch(exec, stream), "graph launch");
            }) / steps;
            ok(hipGraphExecDestroy(exec), "exec destroy");
        }
        std::printf("{\"replicas\": %d, \"particles\": %llu, \"steps\": %d, \"batch_shape\": [%d, %d, %d], "
                    "\"ms_per_step\": {\"graph\": %.4f, \"graph_fused\": %.4f}}\n",
                    count, static_cast<unsigned long long>(R.particles), steps, shape[0], shape[1], shape[2], ms[0],
                    ms[1]);
        ok(hipStreamDestroy(stream), "stream");
        return 0;
    }

    // Launch shape of the batch: the fastest eager step over items per lane
    // x waves per block x scan threads (each from the same start).
    int best[3] = {0, 0, 0};
    double best_ms = 1e30;
    for (int items : {8, 16})
        for (int waves : {4, 8})
            for (int scan : {256, 1024}) {
                dsdx(dsdx_batch_set_shape(batch.b, items, waves, scan, stream), "shape");
                w.reset(s, stream);
                for (int k = 0; k < 2; ++k) batch.step(w, s, true, stream);
                const double t = wall_seconds(stream, [&] {
                    for (int k = 0; k < 3; ++k) batch.step(w, s, true, stream);
                });
                if (t < best_ms) {
                    best_ms = t;
                    best[0] = items; best[1] = waves; best[2] = scan;
                }
            }
    dsdx(dsdx_batch_set_shape(batch.b, best[0], best[1], best[2], stream), "shape");

    // bench: every form from the same start, `warmup` untimed steps first
    struct Timed {
        const char* name;
        double seconds;
    };
    std::vector<Timed> timed;
    for (int fused = 0; fused < 2; ++fused) {
        w.reset(s, stream);
        for (int k = 0; k < warmup; ++k) loop.step(w, s, fused != 0, stream);
        timed.push_back({fused ? "loop_fused" : "loop", wall_seconds(stream, [&] {
                             for (int k = 0; k < steps; ++k) loop.step(w, s, fused != 0, stream);
                         })});
        w.reset(s, stream);
        for (int k = 0; k < warmup; ++k) batch.step(w, s, fused != 0, stream);
        timed.push_back({fused ? "batch_fused" : "batch", wall_seconds(stream, [&] {
                             for (int k = 0; k < steps; ++k) batch.step(w, s, fused != 0, stream);
                         })});
        w.reset(s, stream);
        ok(hipStreamSynchronize(stream), "sync");
        hipGraphExec_t exec = capture(stream, hipStreamCaptureModeGlobal,
                                      [&] { batch.step(w, s, fused != 0, stream); });
        for (int k = 0; k < warmup; ++k) ok(hipGraphLaunch(exec, stream), "graph launch");
        timed.push_back({fused ? "graph_fused" : "graph", wall_seconds(stream, [&] {
                             for (int k = 0; k < steps; ++k) ok(hipGraphLaunch(exec, stream), "graph launch");
                         })});
        ok(hipGraphExecDestroy(exec), "exec destroy");
    }
    std::printf("{\"replicas\": %d, \"particles\": %llu, \"slots\": %llu, \"steps\": %d, \"warmup\": %d, "
                "\"ms_per_step\": {",
                count, static_cast<unsigned long long>(R.particles), static_cast<unsigned long long>(R.total),
                steps, warmup);
    for (size_t k = 0; k < timed.size(); ++k)
        std::printf("%s\"%s\": %.4f", k ? ", " : "", timed[k].name, 1e3 * timed[k].seconds / steps);
    std::printf("}, \"batch_shape\": [%d, %d, %d], \"replica_steps_per_s\": {", best[0], best[1], best[2]);
    for (size_t k = 0; k < timed.size(); ++k)
        std::printf("%s\"%s\": %.1f", k ? ", " : "", timed[k].name, count * steps / timed[k].seconds);
    std::printf("}}\n");
    ok(hipStreamDestroy(stream), "stream");
    return 0;
}
