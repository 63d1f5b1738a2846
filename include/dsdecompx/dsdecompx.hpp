// dsdecompx.hpp -- the public API of the MI355X build: Decomp and
// ParticleInfo with the reference's signatures (include/dsdecompx.hpp:18-248),
// executed by gfx950 HIP kernels on a 1x1x1 self-loopback grid.
//
// Include with -I<repo>/include/dsdecompx (as the reference's
// `#include <dsdecompx.hpp>`) and link libdsdecompx_amd.so.
#pragma once

#include <mpi.h>

#include <cstddef>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "config.hpp"
#include "field.hpp"
#include "perf.hpp"
#include "topology.hpp"
#include "version.hpp"

namespace dsdecompx {

struct HaloResult {
    std::size_t ghost_count = 0;
    bool layout_changed = false;
    bool storage_reallocated = false;
    // Device memory only: the previous plan was refreshed in place.
    bool plan_reused = false;
};

struct HaloExchangeOptions {
    bool rebuild = false;
    ExecutionStream stream = nullptr;
    // Reused plan, no application fields: only x, y, z are refreshed and the
    // ghost forces zeroed; every other ghost member keeps its last value.
    bool positions_only = false;
};
#define DSDECOMPX_HALO_POSITIONS_ONLY 1
#define DSDECOMPX_HOST_HALO_REFRESHES_VELOCITY 1
#define DSDECOMPX_RESCALE_DOMAIN 1
inline const char* header_path() noexcept { return __FILE__; }

struct StorageResizeResult {
    ParticleView view;
    bool reallocated = false;
};

struct RebalanceOptions {
    double local_work = -1.0;
    bool force = false;
};

struct RebalanceResult {
    bool attempted = false;
    bool geometry_changed = false;
    double imbalance_before = 1.0;
    MigrationResult migration;
};

// Particle storage of every Decomp operation.  allocate(CudaDevice, n) gives
// library-owned device arrays that grow as ghosts arrive; external(...) plus
// the setters borrow caller arrays; a Provider grows borrowed storage.  The
// public members and functions are the reference's (include/dsdecompx.hpp:
// 70-168); where the arrays live is the library's own business
// (dsdecompx_amd/csrc/host/particle_info.cpp).
class ParticleInfo {
public:
    class Provider {
    public:
        virtual ~Provider() = default;
        virtual ParticleView particle_info_view(std::size_t minimum) = 0;
        virtual StorageResizeResult resize_particle_info(std::size_t required,
                                                         std::size_t preserve) = 0;
    };

    ParticleInfo();
    explicit ParticleInfo(std::size_t owned);
    ~ParticleInfo();
    ParticleInfo(ParticleInfo&& other) noexcept;
    ParticleInfo& operator=(ParticleInfo&& other) noexcept;
    ParticleInfo(const ParticleInfo&) = delete;
    ParticleInfo& operator=(const ParticleInfo&) = delete;

    ParticleInfo(std::size_t owned, std::size_t ghosts, ParticleView storage);
    ParticleInfo(std::size_t owned, std::size_t ghosts, Provider& provider);

    static ParticleInfo allocate(MemorySpace memory_space, std::size_t owned);
    static ParticleInfo external(MemorySpace memory_space, std::size_t capacity, std::size_t owned,
                                 std::size_t ghosts = 0);
    ParticleInfo& positions(double* x, double* y, double* z) noexcept;
    ParticleInfo& velocities(double* vx, double* vy, double* vz) noexcept;
    ParticleInfo& forces(double* fx, double* fy, double* fz) noexcept;
    ParticleInfo& properties(double* mass, double* charge, int* type, int* global_id) noexcept;

    template <class T>
    ParticleInfo& add_field(std::string name, T* data, std::size_t capacity = 0,
                            std::size_t components = 1)
    {
        return register_field(ParticleFieldView{std::move(name), field_type_v<T>, components,
                                                capacity ? capacity : this->capacity(), data});
    }

    std::vector<double> x, y, z;
    std::vector<double> vx, vy, vz;
    std::vector<double> fx, fy, fz;
    std::vector<double> mass, charge;
    std::vector<int> type, global_id;
    std::size_t num_particles = 0;

    void resize(std::size_t count);

    std::size_t owned_count() const noexcept;
    std::size_t ghost_count() const noexcept;
    std::size_t size() const noexcept;
    std::size_t capacity() const noexcept;
    MemorySpace memory_space() const noexcept;
    bool uses_external_memory() const noexcept;

    double* x_data() noexcept;
    double* y_data() noexcept;
    double* z_data() noexcept;
    double* vx_data() noexcept;
    double* vy_data() noexcept;
    double* vz_data() noexcept;
    double* fx_data() noexcept;
    double* fy_data() noexcept;
    double* fz_data() noexcept;
    double* mass_data() noexcept;
    double* charge_data() noexcept;
    int* type_data() noexcept;
    int* global_id_data() noexcept;

    void resize_owned(std::size_t count);
    void reset_counts(std::size_t owned, std::size_t ghosts = 0);
    void reset_view(ParticleView storage);

private:
    friend class Decomp;
    // The arrays with room for at least `need` particles, fields attached.
    ConstParticleView const_view(std::size_t need) const;
    // Room for `required` particles, keeping the first `preserve` of every
    // array when the storage moves.
    StorageResizeResult ensure_capacity(std::size_t required, std::size_t preserve);
    void set_counts(std::size_t owned, std::size_t ghosts) noexcept;
    ParticleInfo& register_field(ParticleFieldView field);

    // Null after a move: the object then reads as empty host storage.
    struct State;
    std::unique_ptr<State> state_;
    State& state();
};

// The user entry point.  Every operation runs on the caller's HIP stream; a
// stream is required for device memory, as in the reference's CUDA mode.
class Decomp {
public:
    Decomp(MPI_Comm communicator, Config config = {});
    ~Decomp();

    Decomp(Decomp&&) noexcept;
    Decomp& operator=(Decomp&&) noexcept;
    Decomp(const Decomp&) = delete;
    Decomp& operator=(const Decomp&) = delete;

    const Topology& topology() const noexcept;
    BackendKind backend_kind() const noexcept;
    bool mpi_takes_device_pointers() const noexcept;
    int owner_rank(const Vec3d& position) const;
    bool any_rank_true(bool local_flag) const;
    bool has_out_of_domain_particles(const ParticleInfo& particles) const;

    void update_domain(const GlobalDomain& domain);
    void rescale_domain(const GlobalDomain& domain);
    void set_halo_width(double width);
    double halo_width() const noexcept;
    void invalidate_halo() noexcept;

    HaloResult exchange_halo(ParticleInfo& particles, HaloExchangeOptions options = {});
    void begin_halo_exchange(ParticleInfo& particles, HaloExchangeOptions options = {});
    HaloResult end_halo_exchange(ParticleInfo& particles);
    bool can_refresh_halo(const ParticleInfo& particles) const;
    MigrationResult migrate(ParticleInfo& particles, ExecutionStream stream = nullptr);
    // Extension with no counterpart in the reference: migrate(particles,
    // options.stream) followed by exchange_halo(particles, {rebuild, ...}),
    // with bit-identical results, but migrate's periodic wrap runs inside the
    // halo classification pass (one read of the positions instead of two).
    // For loops in which nothing touches the positions between the two calls.
    HaloResult migrate_and_exchange_halo(ParticleInfo& particles, HaloExchangeOptions options = {});
    RebalanceResult rebalance(ParticleInfo& particles, RebalanceOptions options = {},
                              ExecutionStream stream = nullptr);
    void reverse_halo(ParticleInfo& particles, ExecutionStream stream = nullptr);
    void reverse_exchange(ParticleInfo& particles, const std::vector<ReverseExchangeField>& fields,
                          ExecutionStream stream = nullptr);
    void forward_exchange(ParticleInfo& particles, const std::vector<std::string>& fields,
                          ExecutionStream stream = nullptr);

    void set_performance_counters_enabled(bool enabled, bool neighbor_detail = false) noexcept;
    PerfStats get_stats() const;
    PerfStats get_global_stats() const;
    void reset_stats() noexcept;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace dsdecompx
