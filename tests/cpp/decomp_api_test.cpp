// decomp_api_test.cpp -- the public C++ API (Decomp / ParticleInfo on HIP
// device memory) checked like the reference's own programs
// (examples/getting_started, examples/registered_field,
// examples/sanity_check/mpi/regression.cpp, cuda/halo_reuse.cu,
// cuda/forward_exchange.cu) and, step by step, bit-exactly against the CPU
// oracle (linked as the checker only).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "dsdecompx.hpp"
#include "dsdx_oracle.h"

using namespace dsdecompx;

static int g_failures = 0;

#define EXPECT(cond, what)                                                            \
    do {                                                                              \
        if (!(cond)) {                                                                \
            std::printf("FAILED: %s (%s:%d)\n", what, __FILE__, __LINE__);            \
            ++g_failures;                                                             \
        }                                                                             \
    } while (0)

static void hip_ok(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::printf("HIP error %s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

template <class T> static std::vector<T> download(const T* d, std::size_t n)
{
    std::vector<T> h(n);
    if (n) hip_ok(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost), "download");
    return h;
}

template <class T> static void upload(T* d, const std::vector<T>& h)
{
    if (!h.empty()) hip_ok(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
}

static bool same_bits(const double* a, const double* b, std::size_t n)
{
    return std::memcmp(a, b, n * sizeof(double)) == 0;
}

static uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double uniform(uint64_t& s) { return static_cast<double>(splitmix(s) >> 11) * 0x1.0p-53; }

static Config make_config(double L, double w, bool zperiodic = true)
{
    Config c;
    c.domain.upper = {L, L, L};
    c.domain.periodic = {true, true, zperiodic};
    c.halo.width = w;
    return c;
}

// examples/getting_started + registered_field on one GPU: a corner particle
// has 7 self-images and returns fx = 7.
static void test_getting_started(hipStream_t s)
{
    Decomp decomp(MPI_COMM_WORLD, make_config(8.0, 0.25));
    EXPECT(decomp.topology().neighbors().size() == 26, "26 loopback neighbours");
    EXPECT(decomp.backend_kind() == BackendKind::Loopback, "loopback backend");
    ParticleInfo p = ParticleInfo::allocate(MemorySpace::CudaDevice, 1);
    upload(p.x_data(), std::vector<double>{0.05});
    upload(p.y_data(), std::vector<double>{0.05});
    upload(p.z_data(), std::vector<double>{0.05});
    upload(p.global_id_data(), std::vector<int>{0});
    upload(p.mass_data(), std::vector<double>{1.0});
    decomp.migrate(p, s);
    const HaloResult r = decomp.exchange_halo(p, {true, s});
    EXPECT(r.ghost_count == 7 && p.ghost_count() == 7, "corner particle has 7 ghosts");
    EXPECT(r.layout_changed && !r.plan_reused, "a rebuild changes the layout");
    std::vector<double> fx(8, 1.0);
    fx[0] = 0.0;
    upload(p.fx_data(), fx);
    hip_ok(hipMemsetAsync(p.fy_data(), 0, 8 * sizeof(double), s), "memset");
    hip_ok(hipMemsetAsync(p.fz_data(), 0, 8 * sizeof(double), s), "memset");
    decomp.reverse_halo(p, s);
    hip_ok(hipStreamSynchronize(s), "sync");
    const double returned = download(p.fx_data(), 1)[0];
    EXPECT(returned == 7.0, "returned_fx=7");
    // registered_field: contribution 1.0 per ghost reduces to the same value
    double* contribution = nullptr;
    hip_ok(hipMalloc(&contribution, 64 * sizeof(double)), "malloc");
    p.add_field<double>("contribution", contribution, 64);
    decomp.exchange_halo(p, {true, s});
    std::vector<double> c(8, 1.0);
    c[0] = 0.0;
    upload(contribution, c);
    decomp.reverse_exchange(p, {{"contribution", ReduceOp::Sum}}, s);
    hip_ok(hipStreamSynchronize(s), "sync");
    EXPECT(download(contribution, 1)[0] == returned, "registered_field contribution == fx");
    hip_ok(hipFree(contribution), "free");
}

// Errors the reference raises (halo_reuse.cu:170-198, forward_exchange.cu:173-179).
static void test_errors(hipStream_t s)
{
    Decomp decomp(MPI_COMM_WORLD, make_config(8.0, 1.0));
    ParticleInfo p = ParticleInfo::allocate(MemorySpace::CudaDevice, 16);
    bool threw = false;
    try { decomp.exchange_halo(p, {true, nullptr}); } catch (const std::invalid_argument&) { threw = true; }
    EXPECT(threw, "null stream is invalid_argument");
    threw = false;
    try { decomp.end_halo_exchange(p); } catch (const std::logic_error&) { threw = true; }
    EXPECT(threw, "end without begin is logic_error");
    decomp.begin_halo_exchange(p, {true, s});
    threw = false;
    try { decomp.begin_halo_exchange(p, {true, s}); } catch (const std::logic_error&) { threw = true; }
    EXPECT(threw, "double begin is logic_error");
    decomp.end_halo_exchange(p);
    threw = false;
    try { decomp.forward_exchange(p, {"no_such_field"}, s); } catch (const std::exception&) { threw = true; }
    EXPECT(threw, "unknown field is an error");
    threw = false;
    try { decomp.set_halo_width(-1.0); } catch (const std::invalid_argument&) { threw = true; }
    EXPECT(threw, "negative width is invalid_argument");
    threw = false;
    try { Decomp bad(MPI_COMM_NULL, make_config(8.0, 1.0)); } catch (const std::invalid_argument&) { threw = true; }
    EXPECT(threw, "null communicator is invalid_argument");
    ParticleInfo host(4);
    threw = false;
    try { decomp.exchange_halo(host, {true, s}); } catch (const std::invalid_argument&) { threw = true; }
    EXPECT(threw, "host particle memory is rejected");
}

struct HostState {
    std::vector<double> f64[11];
    std::vector<int32_t> i32[2];
};

// 32,768-particle loop through the C++ API, every array compared with the oracle
// after every step; the halo is rebuilt each step after migrate; every third
// step also refreshes the reused plan (positions only, then full).  With
// `fused` the step calls migrate_and_exchange_halo instead of migrate +
// exchange_halo; the oracle side is the same two operations.
static void test_steps_against_oracle(hipStream_t s, std::size_t n, double L, double w,
                                      bool zperiodic, int steps, bool fused = false)
{
    Decomp decomp(MPI_COMM_WORLD, make_config(L, w, zperiodic));
    ParticleInfo p = ParticleInfo::allocate(MemorySpace::CudaDevice, n);
    const std::size_t cap = n * 4 + 64;
    std::vector<double> density(cap, 0.0), vec3(cap * 3, 0.0);
    double* d_density = nullptr;
    double* d_vec = nullptr;
    hip_ok(hipMalloc(&d_density, cap * sizeof(double)), "malloc");
    hip_ok(hipMalloc(&d_vec, cap * 3 * sizeof(double)), "malloc");
    p.add_field<double>("density", d_density, cap).add_field<double>("vec", d_vec, cap, 3);

    dsdxo_box box{};
    for (int a = 0; a < 3; ++a) {
        box.lower[a] = 0.0;
        box.upper[a] = L;
        box.periodic[a] = 1;
    }
    box.periodic[2] = zperiodic ? 1 : 0;
    box.width = w;
    box.edges_and_corners = 1;
    dsdxo_neighbors nb;
    dsdxo_neighbors_build(&box, &nb);
    dsdxo_plan plan;
    std::memset(&plan, 0, sizeof(plan));

    uint64_t seed = 12345;
    std::vector<double> x(cap), y(cap), z(cap), vx(cap), vy(cap), vz(cap), fx(cap), fy(cap), fz(cap),
        m(cap, 1.0), q(cap);
    std::vector<int32_t> t(cap), id(cap);
    for (std::size_t i = 0; i < n; ++i) {
        x[i] = L * uniform(seed); y[i] = L * uniform(seed); z[i] = L * uniform(seed);
        vx[i] = uniform(seed) - 0.5; vy[i] = uniform(seed) - 0.5; vz[i] = uniform(seed) - 0.5;
        q[i] = uniform(seed); t[i] = static_cast<int32_t>(splitmix(seed) % 4); id[i] = static_cast<int32_t>(i);
        density[i] = 2.0 * uniform(seed);
        for (int c = 0; c < 3; ++c) vec3[3 * i + c] = 2.0 * uniform(seed) - 1.0;
    }
    auto push = [&](std::size_t count) {
        hip_ok(hipMemcpy(p.x_data(), x.data(), count * 8, hipMemcpyHostToDevice), "up");
        hip_ok(hipMemcpy(p.y_data(), y.data(), count * 8, hipMemcpyHostToDevice), "up");
        hip_ok(hipMemcpy(p.z_data(), z.data(), count * 8, hipMemcpyHostToDevice), "up");
    };
    upload(p.x_data(), std::vector<double>(x.begin(), x.begin() + n));
    upload(p.y_data(), std::vector<double>(y.begin(), y.begin() + n));
    upload(p.z_data(), std::vector<double>(z.begin(), z.begin() + n));
    upload(p.vx_data(), std::vector<double>(vx.begin(), vx.begin() + n));
    upload(p.vy_data(), std::vector<double>(vy.begin(), vy.begin() + n));
    upload(p.vz_data(), std::vector<double>(vz.begin(), vz.begin() + n));
    upload(p.mass_data(), std::vector<double>(m.begin(), m.begin() + n));
    upload(p.charge_data(), std::vector<double>(q.begin(), q.begin() + n));
    upload(p.type_data(), std::vector<int32_t>(t.begin(), t.begin() + n));
    upload(p.global_id_data(), std::vector<int32_t>(id.begin(), id.begin() + n));
    upload(d_density, density);
    upload(d_vec, vec3);
    hip_ok(hipMemset(p.fx_data(), 0, n * 8), "memset");
    hip_ok(hipMemset(p.fy_data(), 0, n * 8), "memset");
    hip_ok(hipMemset(p.fz_data(), 0, n * 8), "memset");

    dsdxo_particles op{x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(), fx.data(),
                       fy.data(), fz.data(), m.data(), q.data(), t.data(), id.data()};
    dsdxo_field of[2] = {{density.data(), DSDXO_FLOAT64, 1, DSDXO_SUM},
                         {vec3.data(), DSDXO_FLOAT64, 3, DSDXO_SUM}};

    for (int step = 0; step < steps; ++step) {
        for (std::size_t i = 0; i < n; ++i) {
            x[i] += 0.1 * (uniform(seed) - 0.5);
            y[i] += 0.1 * (uniform(seed) - 0.5);
            z[i] += 0.1 * (uniform(seed) - 0.5);
        }
        push(n);
        HaloResult hr;
        if (fused) {
            hr = decomp.migrate_and_exchange_halo(p, {false, s});
            EXPECT(p.owned_count() == n, "fused migrate keeps every particle");
        } else {
            const MigrationResult mr = decomp.migrate(p, s);
            EXPECT(mr.new_owned_count == n && !mr.ownership_changed, "migrate keeps every particle");
            hr = decomp.exchange_halo(p, {false, s});
        }
        dsdxo_migrate_wrap(&box, static_cast<int64_t>(n), x.data(), y.data(), z.data());
        EXPECT(!hr.plan_reused, "migrate invalidates the plan");
        dsdxo_halo_build(&box, &nb, static_cast<int64_t>(n), x.data(), y.data(), z.data(), &plan);
        dsdxo_halo_write_ghosts(&nb, &plan, static_cast<int64_t>(n), &op, of, 2);
        const std::size_t g = static_cast<std::size_t>(plan.ghost_total);
        EXPECT(hr.ghost_count == g, "ghost count equals the oracle");
        if (step % 3 == 2) {
            // reused plan: positions-only (no fields allowed) is not taken with
            // fields registered, so this is a full refresh (same ghosts).
            const HaloResult rr = decomp.exchange_halo(p, {false, s, true});
            EXPECT(rr.plan_reused && !rr.layout_changed && rr.ghost_count == g, "reused plan");
        }
        for (std::size_t i = 0; i < n + g; ++i) {
            fx[i] = 2.0 * uniform(seed) - 1.0;
            fy[i] = 2.0 * uniform(seed) - 1.0;
            fz[i] = 2.0 * uniform(seed) - 1.0;
        }
        hip_ok(hipMemcpy(p.fx_data(), fx.data(), (n + g) * 8, hipMemcpyHostToDevice), "up");
        hip_ok(hipMemcpy(p.fy_data(), fy.data(), (n + g) * 8, hipMemcpyHostToDevice), "up");
        hip_ok(hipMemcpy(p.fz_data(), fz.data(), (n + g) * 8, hipMemcpyHostToDevice), "up");
        decomp.reverse_halo(p, s);
        dsdxo_reverse_halo(&nb, &plan, static_cast<int64_t>(n), fx.data(), fy.data(), fz.data());
        // density: ghost contributions summed back (the clustered workload's reverse_exchange)
        for (std::size_t i = n; i < n + g; ++i) density[i] = uniform(seed);
        hip_ok(hipMemcpy(d_density, density.data(), (n + g) * 8, hipMemcpyHostToDevice), "up");
        decomp.reverse_exchange(p, {{"density", ReduceOp::Sum}}, s);
        dsdxo_field dsum = {density.data(), DSDXO_FLOAT64, 1, DSDXO_SUM};
        dsdxo_reverse_exchange(&nb, &plan, static_cast<int64_t>(n), &dsum, 1);
        hip_ok(hipStreamSynchronize(s), "sync");

        const std::size_t total = n + g;
        bool ok = same_bits(download(p.x_data(), total).data(), x.data(), total) &&
                  same_bits(download(p.y_data(), total).data(), y.data(), total) &&
                  same_bits(download(p.z_data(), total).data(), z.data(), total) &&
                  same_bits(download(p.vx_data(), total).data(), vx.data(), total) &&
                  same_bits(download(p.fx_data(), total).data(), fx.data(), total) &&
                  same_bits(download(p.fy_data(), total).data(), fy.data(), total) &&
                  same_bits(download(p.fz_data(), total).data(), fz.data(), total) &&
                  same_bits(download(p.charge_data(), total).data(), q.data(), total) &&
                  same_bits(download(d_density, total).data(), density.data(), total) &&
                  same_bits(download(d_vec, total * 3).data(), vec3.data(), total * 3) &&
                  download(p.global_id_data(), total) == std::vector<int32_t>(id.begin(), id.begin() + total) &&
                  download(p.type_data(), total) == std::vector<int32_t>(t.begin(), t.begin() + total);
        char what[128];
        std::snprintf(what, sizeof(what), "step %d (n=%zu, ghosts=%zu) bit-exact with the oracle", step, n, g);
        EXPECT(ok, what);
    }
    dsdxo_plan_free(&plan);
    hip_ok(hipFree(d_density), "free");
    hip_ok(hipFree(d_vec), "free");
}

// forward_exchange.cu:103-167: exact copies and adjointness with reverse(Sum).
static void test_forward_adjoint(hipStream_t s)
{
    const std::size_t n = 4000;
    const double L = 12.0;
    Decomp decomp(MPI_COMM_WORLD, make_config(L, 2.0));
    ParticleInfo p = ParticleInfo::allocate(MemorySpace::CudaDevice, n);
    uint64_t seed = 99;
    std::vector<double> h(n);
    for (auto* arr : {p.x_data(), p.y_data(), p.z_data()}) {
        for (auto& v : h) v = L * uniform(seed);
        upload(arr, h);
    }
    const std::size_t cap = 4 * n;
    const int wide = 7;
    double* feat = nullptr;
    hip_ok(hipMalloc(&feat, cap * wide * sizeof(double)), "malloc");
    p.add_field<double>("feat", feat, cap, wide);
    const HaloResult r = decomp.exchange_halo(p, {true, s});
    const std::size_t g = r.ghost_count;
    std::vector<double> xv(n * wide), yv(g * wide), a(cap * wide, 0.0);
    for (auto& v : xv) v = uniform(seed) - 0.5;
    for (auto& v : yv) v = uniform(seed) - 0.5;
    std::copy(xv.begin(), xv.end(), a.begin());
    upload(feat, a);
    decomp.forward_exchange(p, {"feat"}, s);
    hip_ok(hipStreamSynchronize(s), "sync");
    std::vector<double> fwd = download(feat, (n + g) * wide);
    double s1 = 0.0;
    for (std::size_t k = 0; k < g * wide; ++k) s1 += fwd[n * wide + k] * yv[k];
    std::fill(a.begin(), a.end(), 0.0);
    std::copy(yv.begin(), yv.end(), a.begin() + n * wide);
    upload(feat, a);
    decomp.reverse_exchange(p, {{"feat", ReduceOp::Sum}}, s);
    hip_ok(hipStreamSynchronize(s), "sync");
    std::vector<double> rev = download(feat, n * wide);
    double s2 = 0.0;
    for (std::size_t k = 0; k < n * wide; ++k) s2 += xv[k] * rev[k];
    EXPECT(std::abs(s1 - s2) <= 1e-10 * (1.0 + std::abs(s1)), "forward and reverse(Sum) are adjoint");
    hip_ok(hipFree(feat), "free");
}

int main()
{
    hipStream_t s = nullptr;
    hip_ok(hipStreamCreate(&s), "stream");
    try {
        test_getting_started(s);
        test_errors(s);
        test_forward_adjoint(s);
        test_steps_against_oracle(s, 32768, 32.0, 2.8, true, 6);
        test_steps_against_oracle(s, 20000, 27.0, 2.8, false, 4);
        test_steps_against_oracle(s, 2048, 11.95, 2.8, true, 4);
        test_steps_against_oracle(s, 32768, 32.0, 2.8, true, 4, true);
        test_steps_against_oracle(s, 20000, 27.0, 2.8, false, 3, true);
    } catch (const std::exception& e) {
        std::printf("FAILED: exception %s\n", e.what());
        ++g_failures;
    }
    hip_ok(hipStreamDestroy(s), "stream");
    std::printf(g_failures ? "decomp_api_test: %d FAILED\n" : "decomp_api_test: all checks passed\n",
                g_failures);
    return g_failures ? 1 : 0;
}
