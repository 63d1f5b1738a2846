// ref_grid_fixture.cpp -- TEST INFRASTRUCTURE: a program of this project that
// drives the reference library's public API (dsDecompX, unmodified, Host
// execution) on 8 MPI ranks, a 2x2x2 periodic process grid, and writes what
// every rank holds after each operation.  tests/golden/make_ref_fixtures.py
// builds it against the reference sources (oracle/ref.mk), runs it and turns
// the output into the fixture tests/golden/ref_grid_2x2x2.npz, which the
// oracle's P-rank restatement must reproduce bit for bit.  Nothing of this
// program or of the reference is part of the product or travels to the GPU.
//
// usage: mpiexec -n 8 ref_grid_fixture fixture OUTDIR
//        mpiexec -n 8 ref_grid_fixture time N L WIDTH SIGMA STEPS
// fixture: 3000 particles, global id k starts on rank k % 8 (so the first
// migrate moves most of them), positions spread 10 % past every face of the
// box [-3.7, 12.1) x [1.3, 20.9) x [0.25, 9.85) (so migrate wraps them),
// halo width 1.1, one registered field "aux" (3 doubles).  Two steps of:
// migrate; (first step) six owned particles of every rank moved exactly onto
// its slab thresholds or one ulp past them; exchange_halo(rebuild); forces
// set on owned + ghosts;
// reverse_halo; ghost values of "aux" set, reverse_exchange(aux, Sum); owned
// positions displaced.  Each rank writes OUTDIR/rank<r>.txt: one line per
// array, "<step> <name> <count> <values>", doubles in %a notation.
// time: N particles uniform in a periodic cube of side L, the reference's
// communication_test step (migrate, exchange_halo(rebuild), forces,
// reverse_halo) timed with MPI_Wtime; rank 0 prints the slowest rank's mean
// milliseconds per step.
#include <dsdecompx.hpp>
#include <mpi.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

namespace {

uint64_t mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [0, 1) from a key
double unit(uint64_t key) { return static_cast<double>(mix(key) >> 11) * 0x1.0p-53; }

uint64_t key(uint64_t stream, uint64_t a, uint64_t b = 0) { return mix(stream * 0x100000001B3ull ^ mix(a) ^ (b << 1)); }

struct Writer {
    FILE* f;
    int step = 0;
    void doubles(const char* name, const double* v, std::size_t n)
    {
        std::fprintf(f, "%d %s %zu", step, name, n);
        for (std::size_t i = 0; i < n; ++i) std::fprintf(f, " %a", v[i]);
        std::fprintf(f, "\n");
    }
    void ints(const char* name, const int* v, std::size_t n)
    {
        std::fprintf(f, "%d %s %zu", step, name, n);
        for (std::size_t i = 0; i < n; ++i) std::fprintf(f, " %d", v[i]);
        std::fprintf(f, "\n");
    }
};

dsdecompx::Config grid_config(const dsdecompx::Vec3d& lower, const dsdecompx::Vec3d& upper, double width)
{
    dsdecompx::Config c;
    c.execution = dsdecompx::ExecutionMode::Host;
    c.domain.lower = lower;
    c.domain.upper = upper;
    c.domain.periodic = {true, true, true};
    c.halo.width = width;
    c.process_grid.dimensions = {2, 2, 2};
    c.dynamic.interval = 0;   // no automatic rebalance: the cuts stay where the topology put them
    return c;
}

int fixture(const std::string& out_dir)
{
    const dsdecompx::Vec3d lower{-3.7, 1.3, 0.25};
    const dsdecompx::Vec3d upper{12.1, 20.9, 9.85};
    const double width = 1.1;
    dsdecompx::Decomp decomp(MPI_COMM_WORLD, grid_config(lower, upper, width));
    const int rank = decomp.topology().rank();
    const int ranks = decomp.topology().size();
    const std::size_t total = 3000;
    std::vector<std::size_t> mine;
    for (std::size_t k = 0; k < total; ++k)
        if (static_cast<int>(k % static_cast<std::size_t>(ranks)) == rank) mine.push_back(k);
    dsdecompx::ParticleInfo p(mine.size());
    const std::size_t field_capacity = 4 * total;
    std::vector<double> aux(3 * field_capacity, 0.0);
    for (std::size_t i = 0; i < mine.size(); ++i) {
        const uint64_t k = mine[i];
        double* pos[3] = {&p.x[i], &p.y[i], &p.z[i]};
        for (int a = 0; a < 3; ++a) {
            const double extent = upper[a] - lower[a];
            *pos[a] = lower[a] - 0.1 * extent + 1.2 * extent * unit(key(1, k, a));
        }
        p.vx[i] = unit(key(2, k)) - 0.5;
        p.vy[i] = unit(key(3, k)) - 0.5;
        p.vz[i] = unit(key(4, k)) - 0.5;
        p.mass[i] = 1.0 + 0.001 * static_cast<double>(k);
        p.charge[i] = unit(key(5, k));
        p.type[i] = static_cast<int>(k % 5);
        p.global_id[i] = static_cast<int>(k);
        for (int c = 0; c < 3; ++c) aux[3 * i + c] = unit(key(6, k, c)) - 0.5;
    }
    p.add_field<double>("aux", aux.data(), field_capacity, 3);

    const std::string path = out_dir + "/rank" + std::to_string(rank) + ".txt";
    Writer out{std::fopen(path.c_str(), "w")};
    if (!out.f) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        return 1;
    }
    const std::size_t m = mine.size();
    out.ints("start_global_id", p.global_id.data(), m);
    out.ints("start_type", p.type.data(), m);
    out.doubles("start_x", p.x.data(), m);
    out.doubles("start_y", p.y.data(), m);
    out.doubles("start_z", p.z.data(), m);
    out.doubles("start_vx", p.vx.data(), m);
    out.doubles("start_vy", p.vy.data(), m);
    out.doubles("start_vz", p.vz.data(), m);
    out.doubles("start_mass", p.mass.data(), m);
    out.doubles("start_charge", p.charge.data(), m);
    out.doubles("start_aux", aux.data(), 3 * m);
    for (int step = 0; step < 2; ++step) {
        out.step = step;
        decomp.migrate(p);
        std::size_t n = p.owned_count();
        out.ints("migrate_global_id", p.global_id.data(), n);
        out.doubles("migrate_x", p.x.data(), n);
        out.doubles("migrate_y", p.y.data(), n);
        out.doubles("migrate_z", p.z.data(), n);
        out.doubles("migrate_vx", p.vx.data(), n);
        out.doubles("migrate_aux", aux.data(), 3 * n);
        if (step == 0 && n >= 6) {
            // six owned particles exactly on this rank's slab thresholds (the
            // reference's own local domain) or one ulp beyond them
            const auto& local = decomp.topology().local_domain();
            p.x[0] = local.lower[0] + width;
            p.x[1] = std::nextafter(local.lower[0] + width, -1e300);
            p.x[2] = local.upper[0] - width;
            p.x[3] = std::nextafter(local.upper[0] - width, 1e300);
            p.y[4] = local.upper[1] - width;
            p.z[4] = local.lower[2] + width;
            p.z[5] = std::nextafter(local.upper[2] - width, 1e300);
            p.y[5] = std::nextafter(local.lower[1] + width, -1e300);
        }
        out.doubles("ties_x", p.x.data(), n);
        out.doubles("ties_y", p.y.data(), n);
        out.doubles("ties_z", p.z.data(), n);
        decomp.exchange_halo(p, {true, nullptr});
        const std::size_t size = p.size();
        out.ints("halo_global_id", p.global_id.data() + n, size - n);
        out.doubles("halo_x", p.x.data() + n, size - n);
        out.doubles("halo_y", p.y.data() + n, size - n);
        out.doubles("halo_z", p.z.data() + n, size - n);
        out.doubles("halo_vz", p.vz.data() + n, size - n);
        out.doubles("halo_mass", p.mass.data() + n, size - n);
        out.ints("halo_type", p.type.data() + n, size - n);
        out.doubles("halo_aux", aux.data() + 3 * n, 3 * (size - n));
        for (std::size_t i = 0; i < size; ++i) {
            const uint64_t id = static_cast<uint64_t>(p.global_id[i]);
            p.fx[i] = 2.0 * unit(key(10 + 3 * step, id, i)) - 1.0;
            p.fy[i] = 2.0 * unit(key(11 + 3 * step, id, i)) - 1.0;
            p.fz[i] = 2.0 * unit(key(12 + 3 * step, id, i)) - 1.0;
        }
        out.doubles("forces_fx", p.fx.data(), size);
        out.doubles("forces_fy", p.fy.data(), size);
        out.doubles("forces_fz", p.fz.data(), size);
        decomp.reverse_halo(p);
        out.doubles("reverse_fx", p.fx.data(), n);
        out.doubles("reverse_fy", p.fy.data(), n);
        out.doubles("reverse_fz", p.fz.data(), n);
        for (std::size_t i = n; i < size; ++i)
            for (int c = 0; c < 3; ++c) aux[3 * i + c] = unit(key(20 + step, static_cast<uint64_t>(p.global_id[i]), 3 * i + c)) - 0.5;
        out.doubles("aux_ghosts", aux.data() + 3 * n, 3 * (size - n));
        decomp.reverse_exchange(p, {{"aux", dsdecompx::ReduceOp::Sum}});
        out.doubles("aux_reduced", aux.data(), 3 * n);
        for (std::size_t i = 0; i < n; ++i) {
            const uint64_t id = static_cast<uint64_t>(p.global_id[i]);
            p.x[i] += 1.5 * (unit(key(30 + step, id, 0)) - 0.5);
            p.y[i] += 1.5 * (unit(key(30 + step, id, 1)) - 0.5);
            p.z[i] += 1.5 * (unit(key(30 + step, id, 2)) - 0.5);
        }
        p.reset_counts(n, 0);
        out.doubles("displaced_x", p.x.data(), n);
        out.doubles("displaced_y", p.y.data(), n);
        out.doubles("displaced_z", p.z.data(), n);
    }
    std::fclose(out.f);
    return 0;
}

// A normal deviate from two uniforms (Box-Muller).
double normal(uint64_t a, uint64_t b)
{
    const double u1 = 1.0 - unit(a), u2 = unit(b);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

int timing(std::size_t total, double L, double width, double sigma, int steps)
{
    dsdecompx::Decomp decomp(MPI_COMM_WORLD, grid_config({0.0, 0.0, 0.0}, {L, L, L}, width));
    const int rank = decomp.topology().rank();
    const int ranks = decomp.topology().size();
    const std::size_t first = total * static_cast<std::size_t>(rank) / static_cast<std::size_t>(ranks);
    const std::size_t last = total * static_cast<std::size_t>(rank + 1) / static_cast<std::size_t>(ranks);
    const auto& local = decomp.topology().local_domain();
    dsdecompx::ParticleInfo p(last - first);
    for (std::size_t i = 0; i < last - first; ++i) {
        const uint64_t k = first + i;
        p.x[i] = local.lower[0] + (local.upper[0] - local.lower[0]) * unit(key(1, k, 0));
        p.y[i] = local.lower[1] + (local.upper[1] - local.lower[1]) * unit(key(1, k, 1));
        p.z[i] = local.lower[2] + (local.upper[2] - local.lower[2]) * unit(key(1, k, 2));
        p.mass[i] = 1.0;
        p.global_id[i] = static_cast<int>(k);
    }
    double busy = 0.0;
    for (int step = -1; step < steps; ++step) {
        const std::size_t n = p.owned_count();
        for (std::size_t i = 0; i < n; ++i) {
            const uint64_t id = static_cast<uint64_t>(p.global_id[i]);
            p.x[i] += sigma * normal(key(40, id, 2 * step + 100), key(41, id, 2 * step + 100));
            p.y[i] += sigma * normal(key(42, id, 2 * step + 100), key(43, id, 2 * step + 100));
            p.z[i] += sigma * normal(key(44, id, 2 * step + 100), key(45, id, 2 * step + 100));
        }
        MPI_Barrier(MPI_COMM_WORLD);
        const double t0 = MPI_Wtime();
        decomp.migrate(p);
        decomp.exchange_halo(p, {true, nullptr});
        const double t1 = MPI_Wtime();
        for (std::size_t i = 0; i < p.size(); ++i) {
            p.fx[i] = 2.0 * unit(key(50, i, step + 1)) - 1.0;
            p.fy[i] = 2.0 * unit(key(51, i, step + 1)) - 1.0;
            p.fz[i] = 2.0 * unit(key(52, i, step + 1)) - 1.0;
        }
        const double t2 = MPI_Wtime();
        decomp.reverse_halo(p);
        const double t3 = MPI_Wtime();
        if (step >= 0) busy += (t1 - t0) + (t3 - t2);
    }
    double slowest = 0.0;
    MPI_Reduce(&busy, &slowest, 1, MPI_DOUBLE, MPI_MAX, 0, MPI_COMM_WORLD);
    if (rank == 0)
        std::printf("{\"particles\": %zu, \"ranks\": %d, \"steps\": %d, \"ms_per_step\": %.3f}\n", total, ranks,
                    steps, 1e3 * slowest / steps);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    int provided = MPI_THREAD_SINGLE;
    MPI_Init_thread(&argc, &argv, MPI_THREAD_FUNNELED, &provided);
    int rc = 2;
    {
        const std::string mode = argc > 1 ? argv[1] : "";
        if (mode == "fixture" && argc > 2) {
            rc = fixture(argv[2]);
        } else if (mode == "time" && argc > 6) {
            rc = timing(std::strtoull(argv[2], nullptr, 10), std::atof(argv[3]), std::atof(argv[4]),
                        std::atof(argv[5]), std::atoi(argv[6]));
        } else {
            std::fprintf(stderr, "usage: %s fixture OUTDIR | time N L WIDTH SIGMA STEPS\n", argv[0]);
        }
    }
    MPI_Finalize();
    return rc;
}
