// decomp_api_test.cpp -- the public C++ API (Decomp / ParticleInfo on HIP
// device memory) checked like the reference's own programs
// (examples/getting_started, examples/registered_field,
// examples/sanity_check/mpi/regression.cpp, cuda/halo_reuse.cu,
// cuda/forward_exchange.cu) and, step by step, bit-exactly against the CPU
// oracle (linked as the checker only).
#include <hip/hip_runtime.h>

#include <array>
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

// Device copies of 13 SoA arrays the test owns (borrowed by ParticleInfo).
struct DeviceArrays {
    double* f64[11] = {};
    int* i32[2] = {};
    std::size_t capacity = 0;
    explicit DeviceArrays(std::size_t cap) : capacity(cap)
    {
        for (auto& p : f64) hip_ok(hipMalloc(&p, cap * sizeof(double)), "malloc");
        for (auto& p : i32) hip_ok(hipMalloc(&p, cap * sizeof(int)), "malloc");
        for (auto& p : f64) hip_ok(hipMemset(p, 0, cap * sizeof(double)), "memset");
        for (auto& p : i32) hip_ok(hipMemset(p, 0, cap * sizeof(int)), "memset");
    }
    ~DeviceArrays()
    {
        for (auto& p : f64) (void)hipFree(p);
        for (auto& p : i32) (void)hipFree(p);
    }
    ParticleInfo borrow(std::size_t owned)
    {
        ParticleInfo p = ParticleInfo::external(MemorySpace::CudaDevice, capacity, owned);
        p.positions(f64[0], f64[1], f64[2]).velocities(f64[3], f64[4], f64[5]);
        p.forces(f64[6], f64[7], f64[8]).properties(f64[9], f64[10], i32[0], i32[1]);
        return p;
    }
};

// The same particles on the host, for the oracle.
struct HostArrays {
    std::vector<double> f64[11];
    std::vector<int32_t> i32[2];
    explicit HostArrays(std::size_t cap)
    {
        for (auto& v : f64) v.assign(cap, 0.0);
        for (auto& v : i32) v.assign(cap, 0);
    }
    dsdxo_particles view()
    {
        return {f64[0].data(), f64[1].data(), f64[2].data(), f64[3].data(), f64[4].data(), f64[5].data(),
                f64[6].data(), f64[7].data(), f64[8].data(), f64[9].data(), f64[10].data(), i32[0].data(),
                i32[1].data()};
    }
    void push(DeviceArrays& d, std::size_t count) const
    {
        for (int k = 0; k < 11; ++k)
            hip_ok(hipMemcpy(d.f64[k], f64[k].data(), count * 8, hipMemcpyHostToDevice), "up");
        for (int k = 0; k < 2; ++k)
            hip_ok(hipMemcpy(d.i32[k], i32[k].data(), count * 4, hipMemcpyHostToDevice), "up");
    }
    // every array of the device copy equals this one over [0, count)
    bool same(DeviceArrays& d, std::size_t count) const
    {
        bool ok = true;
        for (int k = 0; k < 11; ++k) ok = ok && same_bits(download(d.f64[k], count).data(), f64[k].data(), count);
        for (int k = 0; k < 2; ++k)
            ok = ok && download(d.i32[k], count) == std::vector<int32_t>(i32[k].begin(), i32[k].begin() + count);
        return ok;
    }
};

__global__ void owned_only_work(double* vx, double* mass, std::size_t owned)
{
    const std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
    if (i < owned) {
        vx[i] = vx[i] * 0.5 + 1.0;
        mass[i] = mass[i] + 2.0;
    }
}

// examples/sanity_check/cuda/halo_reuse.cu:84-198 on the loopback grid, every
// step bit-exact against the oracle: refreshes on a reused plan move the
// ghosts with their owners (a particle that jumps out of its slab keeps its
// images until the next rebuild); a split-phase exchange with owned-only work
// between begin and end equals exchange_halo followed by that work; an
// invalidated plan is rebuilt and drops the jumped particle's images.
static void test_halo_reuse(hipStream_t s)
{
    const double L = 8.0, w = 0.5, drift = 0.01, jump = 1.0;
    const std::size_t owned = 8, cap = 256;
    Decomp decomp(MPI_COMM_WORLD, make_config(L, w));
    Decomp twin_decomp(MPI_COMM_WORLD, make_config(L, w));
    DeviceArrays dev(cap), twin(cap);
    HostArrays host(cap);
    for (std::size_t k = 0; k < owned; ++k) {
        const double offset = k < 4 ? 0.05 + 0.01 * static_cast<double>(k) : 1.0 + 0.1 * static_cast<double>(k);
        for (int a = 0; a < 3; ++a) host.f64[a][k] = offset;
        host.f64[3][k] = 0.25 * static_cast<double>(k);
        host.f64[9][k] = 1.0;
        host.i32[1][k] = static_cast<int32_t>(k);
    }
    host.push(dev, owned);
    host.push(twin, owned);
    ParticleInfo p = dev.borrow(owned);
    ParticleInfo q = twin.borrow(owned);
    EXPECT(p.uses_external_memory() && p.capacity() == cap, "borrowed arrays");

    dsdxo_box box{};
    for (int a = 0; a < 3; ++a) {
        box.upper[a] = L;
        box.periodic[a] = 1;
    }
    box.width = w;
    box.edges_and_corners = 1;
    dsdxo_neighbors nb;
    dsdxo_neighbors_build(&box, &nb);
    dsdxo_plan plan;
    std::memset(&plan, 0, sizeof(plan));
    dsdxo_particles hv = host.view();

    const HaloResult built = decomp.exchange_halo(p, {true, s});
    twin_decomp.exchange_halo(q, {true, s});
    dsdxo_halo_build(&box, &nb, static_cast<int64_t>(owned), hv.x, hv.y, hv.z, &plan);
    dsdxo_halo_write_ghosts(&nb, &plan, static_cast<int64_t>(owned), &hv, nullptr, 0);
    const std::size_t ghosts = static_cast<std::size_t>(plan.ghost_total);
    hip_ok(hipStreamSynchronize(s), "sync");
    EXPECT(built.ghost_count == ghosts && ghosts > 0 && built.layout_changed, "reuse: first build");
    EXPECT(host.same(dev, owned + ghosts), "reuse: first build bit-exact with the oracle");

    auto move_owners = [&](double by, bool jump_first) {
        for (std::size_t k = 0; k < owned; ++k)
            for (int a = 0; a < 3; ++a) host.f64[a][k] += by + (jump_first && k == 0 ? jump : 0.0);
        for (int a = 0; a < 3; ++a) {
            hip_ok(hipMemcpy(dev.f64[a], host.f64[a].data(), owned * 8, hipMemcpyHostToDevice), "up");
            hip_ok(hipMemcpy(twin.f64[a], host.f64[a].data(), owned * 8, hipMemcpyHostToDevice), "up");
        }
    };
    for (int step = 1; step <= 3; ++step) {
        move_owners(drift, step == 2);
        const HaloResult r = decomp.exchange_halo(p, {false, s});
        twin_decomp.exchange_halo(q, {false, s});
        dsdxo_halo_write_ghosts(&nb, &plan, static_cast<int64_t>(owned), &hv, nullptr, 0);
        hip_ok(hipStreamSynchronize(s), "sync");
        EXPECT(r.ghost_count == ghosts && r.plan_reused && !r.layout_changed, "reuse: plan kept");
        EXPECT(p.ghost_count() == ghosts, "reuse: ghost count committed");
        EXPECT(host.same(dev, owned + ghosts), "reuse: refreshed ghosts bit-exact with the oracle");
    }
    // the jumped particle 0 still has its images after the refreshes
    std::vector<int32_t> owner_of(ghosts);
    dsdxo_ghost_owners(&nb, &plan, owner_of.data());
    std::size_t images_of_0 = 0;
    for (int32_t o : owner_of) images_of_0 += o == 0;
    EXPECT(images_of_0 == 7, "reuse: the jumped corner particle keeps its 7 images");

    // split phase with owned-only work between begin and end, against
    // exchange_halo followed by the same work
    EXPECT(decomp.can_refresh_halo(p), "split: refreshable");
    move_owners(drift, false);
    decomp.begin_halo_exchange(p, {false, s});
    bool second_begin = false;
    try { decomp.begin_halo_exchange(p, {false, s}); } catch (const std::logic_error&) { second_begin = true; }
    EXPECT(second_begin, "split: a second begin is a logic_error");
    hipLaunchKernelGGL(owned_only_work, dim3(1), dim3(64), 0, s, dev.f64[3], dev.f64[9], owned);
    const HaloResult split = decomp.end_halo_exchange(p);
    twin_decomp.exchange_halo(q, {false, s});
    hipLaunchKernelGGL(owned_only_work, dim3(1), dim3(64), 0, s, twin.f64[3], twin.f64[9], owned);
    dsdxo_halo_write_ghosts(&nb, &plan, static_cast<int64_t>(owned), &hv, nullptr, 0);
    for (std::size_t k = 0; k < owned; ++k) {
        host.f64[3][k] = host.f64[3][k] * 0.5 + 1.0;
        host.f64[9][k] = host.f64[9][k] + 2.0;
    }
    hip_ok(hipStreamSynchronize(s), "sync");
    EXPECT(split.ghost_count == ghosts && split.plan_reused && !split.layout_changed, "split: plan kept");
    bool twin_same = true;
    for (int k = 0; k < 11; ++k)
        twin_same = twin_same && same_bits(download(dev.f64[k], owned + ghosts).data(),
                                           download(twin.f64[k], owned + ghosts).data(), owned + ghosts);
    for (int k = 0; k < 2; ++k)
        twin_same = twin_same && download(dev.i32[k], owned + ghosts) == download(twin.i32[k], owned + ghosts);
    EXPECT(twin_same, "split: begin / owned work / end equals exchange_halo then the work, bit for bit");
    EXPECT(host.same(dev, owned + ghosts), "split: bit-exact with the oracle");
    bool end_again = false;
    try { decomp.end_halo_exchange(p); } catch (const std::logic_error&) { end_again = true; }
    EXPECT(end_again, "split: end without begin is a logic_error");

    // an invalidated plan is rebuilt: the jumped particle left its slabs
    decomp.invalidate_halo();
    EXPECT(!decomp.can_refresh_halo(p), "invalidate: not refreshable");
    const HaloResult replanned = decomp.exchange_halo(p, {false, s});
    dsdxo_halo_build(&box, &nb, static_cast<int64_t>(owned), hv.x, hv.y, hv.z, &plan);
    dsdxo_halo_write_ghosts(&nb, &plan, static_cast<int64_t>(owned), &hv, nullptr, 0);
    hip_ok(hipStreamSynchronize(s), "sync");
    const std::size_t g2 = static_cast<std::size_t>(plan.ghost_total);
    EXPECT(replanned.ghost_count == g2 && g2 + 7 == ghosts && replanned.layout_changed && !replanned.plan_reused,
           "invalidate: the rebuild drops the 7 images of the jumped particle");
    EXPECT(host.same(dev, owned + g2), "invalidate: rebuild bit-exact with the oracle");
    dsdxo_plan_free(&plan);
}

// rescale_domain (src/decomp.cpp:109-144): the plan survives, a refresh gives
// the ghosts the new extent's shifts, and the next rebuild classifies against
// the rescaled cuts (the upper cut of [9.8, 53.6) is one ulp below 53.6 and
// moves with the box), all bit-exact against the oracle.
static void test_rescale_domain(hipStream_t s)
{
    const double lo = 9.8, up = 53.6, w = 1.1, up2 = 60.1;
    const std::size_t n = 3000, cap = 3 * n;
    Config c = make_config(0.0, w);
    c.domain.lower = {lo, lo, lo};
    c.domain.upper = {up, up, up};
    Decomp decomp(MPI_COMM_WORLD, c);
    DeviceArrays dev(cap);
    HostArrays host(cap);
    uint64_t seed = 777;
    for (std::size_t i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) host.f64[a][i] = lo + (up - lo) * uniform(seed);
        host.f64[3][i] = uniform(seed) - 0.5;
        host.f64[9][i] = 1.0;
        host.i32[1][i] = static_cast<int32_t>(i);
    }
    // ties on the upper slab threshold of the fresh cut
    host.f64[0][0] = 52.5;
    host.f64[1][1] = 52.5;
    host.push(dev, n);
    ParticleInfo p = dev.borrow(n);

    dsdxo_box box{};
    for (int a = 0; a < 3; ++a) {
        box.lower[a] = lo;
        box.upper[a] = up;
        box.periodic[a] = 1;
    }
    box.width = w;
    box.edges_and_corners = 1;
    dsdxo_neighbors nb;
    dsdxo_neighbors_build(&box, &nb);
    dsdxo_plan plan;
    std::memset(&plan, 0, sizeof(plan));
    dsdxo_particles hv = host.view();
    const HaloResult built = decomp.exchange_halo(p, {true, s});
    dsdxo_halo_build(&box, &nb, static_cast<int64_t>(n), hv.x, hv.y, hv.z, &plan);
    dsdxo_halo_write_ghosts(&nb, &plan, static_cast<int64_t>(n), &hv, nullptr, 0);
    hip_ok(hipStreamSynchronize(s), "sync");
    std::size_t g = static_cast<std::size_t>(plan.ghost_total);
    EXPECT(built.ghost_count == g, "rescale: first build ghost count");
    EXPECT(host.same(dev, n + g), "rescale: first build bit-exact (upper cut one ulp below upper)");

    // rescale: same plan, new shifts on refresh
    double cut_lo[3], cut_hi[3];
    dsdxo_local_domain(&box, cut_lo, cut_hi);
    dsdxo_box box2 = box;
    for (int a = 0; a < 3; ++a) box2.upper[a] = up2;
    dsdxo_rescale_cuts(box.lower, box.upper, box2.lower, box2.upper, cut_lo, cut_hi);
    GlobalDomain d2 = c.domain;
    d2.upper = {up2, up2, up2};
    decomp.rescale_domain(d2);
    const LocalDomain& local = decomp.topology().local_domain();
    EXPECT(local.upper[0] == cut_hi[0] && local.lower[0] == cut_lo[0], "rescale: cuts moved as the oracle's");
    dsdxo_neighbors nb2;
    dsdxo_neighbors_build(&box2, &nb2);
    const HaloResult refreshed = decomp.exchange_halo(p, {false, s});
    dsdxo_halo_write_ghosts(&nb2, &plan, static_cast<int64_t>(n), &hv, nullptr, 0);
    hip_ok(hipStreamSynchronize(s), "sync");
    EXPECT(refreshed.plan_reused && refreshed.ghost_count == g, "rescale: the plan is kept");
    EXPECT(host.same(dev, n + g), "rescale: refreshed ghosts carry the new shifts, bit-exact");

    // rebuild on the rescaled cuts, with particles on the new thresholds
    for (int a = 0; a < 3; ++a) {
        box2.local_lower[a] = cut_lo[a];
        box2.local_upper[a] = cut_hi[a];
    }
    box2.local_given = 1;
    host.f64[0][2] = cut_hi[0] - w;
    host.f64[0][3] = up2 - w;
    host.f64[2][4] = cut_lo[2] + w;
    host.push(dev, n);
    const HaloResult rebuilt = decomp.exchange_halo(p, {true, s});
    dsdxo_halo_build(&box2, &nb2, static_cast<int64_t>(n), hv.x, hv.y, hv.z, &plan);
    dsdxo_halo_write_ghosts(&nb2, &plan, static_cast<int64_t>(n), &hv, nullptr, 0);
    hip_ok(hipStreamSynchronize(s), "sync");
    g = static_cast<std::size_t>(plan.ghost_total);
    EXPECT(rebuilt.ghost_count == g, "rescale: rebuild ghost count on the rescaled cuts");
    EXPECT(host.same(dev, n + g), "rescale: rebuild bit-exact on the rescaled cuts");
    dsdxo_plan_free(&plan);
}

// A Provider that hands out device arrays and grows them on request.
class GrowingProvider : public ParticleInfo::Provider {
public:
    ~GrowingProvider() override { release(); }
    ParticleView particle_info_view(std::size_t minimum) override { return make(minimum, 0); }
    StorageResizeResult resize_particle_info(std::size_t required, std::size_t preserve) override
    {
        ++grown;
        return {make(required + 100, preserve), true};
    }
    int grown = 0;

private:
    ParticleView make(std::size_t count, std::size_t preserve)
    {
        ParticleView v;
        v.memory_space = MemorySpace::CudaDevice;
        v.count = count;
        double* d[11];
        int* t[2];
        for (auto& p : d) hip_ok(hipMalloc(&p, count * 8), "malloc");
        for (auto& p : t) hip_ok(hipMalloc(&p, count * 4), "malloc");
        double* old_d[11] = {view_.x, view_.y, view_.z, view_.vx, view_.vy, view_.vz,
                             view_.fx, view_.fy, view_.fz, view_.mass, view_.charge};
        int* old_t[2] = {view_.type, view_.global_id};
        for (int k = 0; k < 11 && preserve; ++k)
            hip_ok(hipMemcpy(d[k], old_d[k], preserve * 8, hipMemcpyDeviceToDevice), "copy");
        for (int k = 0; k < 2 && preserve; ++k)
            hip_ok(hipMemcpy(t[k], old_t[k], preserve * 4, hipMemcpyDeviceToDevice), "copy");
        release();
        v.x = d[0]; v.y = d[1]; v.z = d[2]; v.vx = d[3]; v.vy = d[4]; v.vz = d[5];
        v.fx = d[6]; v.fy = d[7]; v.fz = d[8]; v.mass = d[9]; v.charge = d[10];
        v.type = t[0]; v.global_id = t[1];
        view_ = v;
        return v;
    }
    void release()
    {
        for (void* p : {static_cast<void*>(view_.x), static_cast<void*>(view_.y), static_cast<void*>(view_.z),
                        static_cast<void*>(view_.vx), static_cast<void*>(view_.vy), static_cast<void*>(view_.vz),
                        static_cast<void*>(view_.fx), static_cast<void*>(view_.fy), static_cast<void*>(view_.fz),
                        static_cast<void*>(view_.mass), static_cast<void*>(view_.charge),
                        static_cast<void*>(view_.type), static_cast<void*>(view_.global_id)})
            if (p) (void)hipFree(p);
        view_ = ParticleView{};
    }
    ParticleView view_;
};

// ParticleInfo's storage paths: library-owned device storage that grows and
// keeps its particles, borrowed arrays that cannot be resized, a Provider
// that grows borrowed arrays during an exchange, field registration, moves,
// and the error types of the reference (src/runtime.cpp:451-755).
static void test_particle_storage(hipStream_t s)
{
    ParticleInfo a = ParticleInfo::allocate(MemorySpace::CudaDevice, 10);
    EXPECT(a.owned_count() == 10 && a.capacity() >= 10 && !a.uses_external_memory() &&
               a.memory_space() == MemorySpace::CudaDevice,
           "storage: allocate");
    std::vector<int> ids(10);
    for (int k = 0; k < 10; ++k) ids[k] = 100 + k;
    upload(a.global_id_data(), ids);
    upload(a.mass_data(), std::vector<double>(10, 3.5));
    a.resize_owned(50000);
    EXPECT(a.owned_count() == 50000 && a.capacity() >= 50000, "storage: grown");
    EXPECT(download(a.global_id_data(), 10) == ids, "storage: growth keeps the particles (ids)");
    EXPECT(download(a.mass_data(), 10) == std::vector<double>(10, 3.5), "storage: growth keeps the particles (mass)");
    ParticleInfo moved(std::move(a));
    EXPECT(moved.owned_count() == 50000 && a.owned_count() == 0 && a.capacity() == 0, "storage: move");
    EXPECT(download(moved.global_id_data(), 10) == ids, "storage: moved arrays");
    bool threw = false;
    try { moved.reset_counts(moved.capacity(), 1); } catch (const std::invalid_argument&) { threw = true; }
    EXPECT(threw, "storage: reset_counts beyond capacity is invalid_argument");
    double* field = nullptr;
    hip_ok(hipMalloc(&field, 60000 * sizeof(double)), "malloc");
    moved.add_field<double>("f", field, 60000);
    threw = false;
    try { moved.add_field<double>("f", field, 60000); } catch (const std::invalid_argument&) { threw = true; }
    EXPECT(threw, "storage: duplicate field is invalid_argument");
    threw = false;
    try { moved.add_field<double>("g", field, 10); } catch (const std::invalid_argument&) { threw = true; }
    EXPECT(threw, "storage: undersized field is invalid_argument");
    threw = false;
    try { moved.resize_owned(70000); } catch (const std::length_error&) { threw = true; }
    EXPECT(threw, "storage: growth past a field's capacity is length_error");
    hip_ok(hipFree(field), "free");

    DeviceArrays arrays(64);
    ParticleInfo b = arrays.borrow(8);
    EXPECT(b.uses_external_memory() && b.x_data() == arrays.f64[0] && b.capacity() == 64, "storage: borrowed");
    threw = false;
    try { b.resize_owned(10); } catch (const std::logic_error&) { threw = true; }
    EXPECT(threw, "storage: resizing borrowed arrays is logic_error");
    threw = false;
    try { ParticleInfo bad(70, 0, ParticleView{MemorySpace::CudaDevice, 64}); } catch (const std::invalid_argument&) { threw = true; }
    EXPECT(threw, "storage: a view smaller than its counts is invalid_argument");
    threw = false;
    try { b.reset_view(ParticleView{MemorySpace::CudaDevice, 4}); } catch (const std::invalid_argument&) { threw = true; }
    EXPECT(threw, "storage: reset_view smaller than the particles is invalid_argument");

    ParticleInfo h(4);
    EXPECT(h.capacity() == 4 && h.memory_space() == MemorySpace::Host && h.x_data() == h.x.data() &&
               h.num_particles == 4,
           "storage: host vectors");

    // a Provider grows borrowed storage when the ghosts arrive; the fixed
    // borrowed arrays cannot and report length_error
    GrowingProvider provider;
    {
        const std::size_t n = 2000;
        const double L = 10.0, w = 2.0;
        ParticleInfo pp(n, 0, provider);
        uint64_t seed = 5;
        std::vector<double> pos(n);
        for (double* arr : {pp.x_data(), pp.y_data(), pp.z_data()}) {
            for (auto& v : pos) v = L * uniform(seed);
            upload(arr, pos);
        }
        std::vector<int> id(n);
        for (std::size_t i = 0; i < n; ++i) id[i] = static_cast<int>(i);
        upload(pp.global_id_data(), id);
        Decomp decomp(MPI_COMM_WORLD, make_config(L, w));
        const HaloResult r = decomp.exchange_halo(pp, {true, s});
        hip_ok(hipStreamSynchronize(s), "sync");
        EXPECT(provider.grown >= 1 && r.storage_reallocated && pp.capacity() >= n + r.ghost_count,
               "storage: the provider grew the arrays");
        EXPECT(download(pp.global_id_data(), n) == id, "storage: provider growth kept the owned particles");
        DeviceArrays fixed(n + 10);
        ParticleInfo fp = fixed.borrow(n);
        for (int k = 0; k < 3; ++k)
            hip_ok(hipMemcpy(fixed.f64[k], std::array<double*, 3>{pp.x_data(), pp.y_data(), pp.z_data()}[k],
                             n * 8, hipMemcpyDeviceToDevice), "copy");
        threw = false;
        try { decomp.exchange_halo(fp, {true, s}); } catch (const std::length_error&) { threw = true; }
        EXPECT(threw, "storage: fixed borrowed arrays too small for the ghosts are length_error");
    }
}

int main()
{
    hipStream_t s = nullptr;
    hip_ok(hipStreamCreate(&s), "stream");
    try {
        test_getting_started(s);
        test_errors(s);
        test_particle_storage(s);
        test_halo_reuse(s);
        test_rescale_domain(s);
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
