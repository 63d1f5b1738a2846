// decomp.cpp -- Decomp on a 1x1x1 self-loopback grid: the reference's public
// operations (src/runtime.cpp:757-1083, src/decomp.cpp, src/halo.cpp,
// src/migration.cpp device branches) orchestrating the gfx950 kernels
// through the C ABI of include/dsdx_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <unordered_set>

#include "dsdecompx.hpp"
#include "dsdx_hip.h"

namespace dsdecompx {

namespace detail {
struct TopologyAccess {
    static int& rank(Topology& t) { return t.rank_; }
    static int& size(Topology& t) { return t.size_; }
    static Vec3i& dimensions(Topology& t) { return t.dimensions_; }
    static Vec3i& coordinates(Topology& t) { return t.coordinates_; }
    static LocalDomain& local_domain(Topology& t) { return t.local_domain_; }
    static std::vector<Neighbor>& neighbors(Topology& t) { return t.neighbors_; }
    static std::array<std::vector<double>, 3>& axis_cuts(Topology& t) { return t.axis_cuts_; }
};
}  // namespace detail

namespace {

// Status code of the C ABI -> the exception the reference throws for it.
void check(int rc)
{
    if (rc == DSDX_OK) return;
    const std::string message = dsdx_last_error();
    switch (rc) {
    case DSDX_EINVAL: throw std::invalid_argument(message);
    case DSDX_ESTATE: throw std::logic_error(message);
    case DSDX_ECAPACITY: throw std::length_error(message);
    default: throw std::runtime_error(message);
    }
}

int field_code(FieldType t) { return static_cast<int>(t); }

// Name/type/components of every field, in order: equal signatures mean the
// same ghost layout (src/halo.cpp:92-116 same_wire_layout).
std::string field_signature(const ConstParticleView& v)
{
    std::string s;
    for (std::size_t f = 0; f < v.field_count; ++f) {
        s += v.fields[f].name;
        s += '\x1f';
        s += std::to_string(field_code(v.fields[f].type));
        s += ':';
        s += std::to_string(v.fields[f].components);
        s += '\x1e';
    }
    return s;
}

void require_device(MemorySpace space, const char* what)
{
    if (space != MemorySpace::CudaDevice && space != MemorySpace::CudaManaged)
        throw std::invalid_argument(std::string("dsDecompX: ") + what +
                                    " needs device particle memory in the MI355X build");
}

dsdx_particles to_abi(const ParticleView& v)
{
    dsdx_particles p;
    p.x = v.x; p.y = v.y; p.z = v.z;
    p.vx = v.vx; p.vy = v.vy; p.vz = v.vz;
    p.fx = v.fx; p.fy = v.fy; p.fz = v.fz;
    p.mass = v.mass; p.charge = v.charge;
    p.type = v.type; p.global_id = v.global_id;
    return p;
}

dsdx_particles to_abi(const ConstParticleView& v)
{
    dsdx_particles p;
    p.x = const_cast<double*>(v.x); p.y = const_cast<double*>(v.y);
    p.z = const_cast<double*>(v.z); p.vx = const_cast<double*>(v.vx);
    p.vy = const_cast<double*>(v.vy); p.vz = const_cast<double*>(v.vz);
    p.fx = const_cast<double*>(v.fx); p.fy = const_cast<double*>(v.fy);
    p.fz = const_cast<double*>(v.fz); p.mass = const_cast<double*>(v.mass);
    p.charge = const_cast<double*>(v.charge); p.type = const_cast<int*>(v.type);
    p.global_id = const_cast<int*>(v.global_id);
    return p;
}

// All registered fields of a view as C-ABI field descriptors, validated as
// the reference does (src/halo.cpp:57-71; device wire fields are 32-bit word
// multiples and at most 16).
std::vector<dsdx_field> all_fields(const ConstParticleView& v, std::size_t min_capacity)
{
    if (v.field_count > DSDX_MAX_FIELDS)
        throw std::invalid_argument("dsDecompX: at most 16 custom fields travel on device memory");
    std::unordered_set<std::string> names;
    std::vector<dsdx_field> out;
    for (std::size_t f = 0; f < v.field_count; ++f) {
        const ParticleFieldView& fld = v.fields[f];
        if (fld.name.empty() || !fld.data || fld.components == 0 || fld.capacity < min_capacity)
            throw std::invalid_argument("dsDecompX: invalid or undersized custom particle field");
        if (!names.insert(fld.name).second)
            throw std::invalid_argument("dsDecompX: duplicate custom particle field '" + fld.name + "'");
        if (fld.record_bytes() % 4 != 0)
            throw std::invalid_argument("dsDecompX: custom field '" + fld.name +
                                        "' must be a multiple of 4 bytes per particle");
        out.push_back({fld.data, field_code(fld.type), static_cast<int32_t>(fld.components),
                       DSDX_SUM});
    }
    return out;
}

std::size_t field_capacity(const ConstParticleView& v, std::size_t particle_capacity)
{
    std::size_t cap = particle_capacity;
    for (std::size_t f = 0; f < v.field_count; ++f) cap = std::min(cap, v.fields[f].capacity);
    return cap;
}

// Looks up named fields for reverse/forward exchanges (src/halo.cpp:1237-1271,
// :1422-1455): unique names, present, large enough for owned + ghosts.
const ParticleFieldView& find_field(const ParticleView& v, const std::string& name,
                                    std::size_t required, const char* what)
{
    for (std::size_t f = 0; f < v.field_count; ++f) {
        const ParticleFieldView& fld = v.fields[f];
        if (fld.name != name) continue;
        if (!fld.data || fld.components == 0 || fld.capacity < required) break;
        return fld;
    }
    throw std::invalid_argument(std::string("dsDecompX: missing or undersized ") + what +
                                " field '" + name + "'");
}

}  // namespace

const char* backend_selection_name(BackendSelection s) noexcept
{
    switch (s) {
    case BackendSelection::Auto: return "auto";
    case BackendSelection::Mpi: return "mpi";
    case BackendSelection::Nccl: return "nccl";
    case BackendSelection::Nvshmem: return "nvshmem";
    }
    return "unknown";
}

BackendSelection parse_backend_selection(std::string_view value)
{
    std::string v(value);
    for (char& c : v) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    if (v.empty() || v == "auto") return BackendSelection::Auto;
    if (v == "mpi") return BackendSelection::Mpi;
    if (v == "nccl") return BackendSelection::Nccl;
    if (v == "nvshmem") return BackendSelection::Nvshmem;
    throw std::invalid_argument("dsDecompX: unknown backend '" + std::string(value) + "'");
}

// Only the single-GPU loopback transport is compiled into this build; Auto
// selects it.
bool backend_compiled(BackendSelection selection) noexcept
{
    return selection == BackendSelection::Auto;
}

GridShape choose_process_grid(int nranks, const DomainShape&, const TopologyInfo&, GridPolicy,
                              GridShape manual)
{
    if (nranks != 1)
        throw std::invalid_argument("dsDecompX: the MI355X loopback build runs one rank");
    (void)manual;
    return GridShape{1, 1, 1};
}

struct Decomp::Impl {
    Config config;
    Topology topology;
    dsdx_box box{};
    dsdx_plan* plan = nullptr;

    // device halo plan state (src/internal/decomp_impl.hpp:58-106)
    bool device_halo = false;      // the last exchange left a usable ghost -> owner map
    bool plan_valid = false;       // ... and it may be refreshed in place
    bool plan_reused = false;
    std::size_t plan_owned = 0;
    std::size_t plan_ghosts = 0;
    std::string plan_schema;
    bool shifts_changed = false;   // rescale_domain since the last build
    bool wrap_in_build = false;    // the next rebuild also applies migrate's wrap

    struct Pending {
        bool active = false;
        HaloExchangeOptions options;
        std::size_t owned = 0;
        bool reuse = false;
        bool compact = false;
        bool reallocated = false;
        std::string schema;
    } pending;
    bool operation_active = false;

    bool perf_enabled = false;
    bool neighbor_detail = false;
    PerfStats stats;

    ~Impl()
    {
        if (plan) dsdx_plan_destroy(plan);
    }

    // The two cut planes of each axis of the 1x1x1 grid, as a fresh topology
    // places them: lower + extent * i / 1 (src/topology.cpp:66-72).  The
    // upper cut can lie one ulp away from the domain's upper bound.
    void place_cuts()
    {
        auto& cuts = detail::TopologyAccess::axis_cuts(topology);
        for (int a = 0; a < 3; ++a) {
            const double extent = config.domain.upper[a] - config.domain.lower[a];
            cuts[a] = {config.domain.lower[a] + extent * 0.0 / 1.0,
                       config.domain.lower[a] + extent * 1.0 / 1.0};
        }
    }

    // Neighbours, shifts and the local domain from the domain and the cuts
    // (the cuts themselves are placed once, then only rescaled).
    void refresh_geometry()
    {
        auto& dims = detail::TopologyAccess::dimensions(topology);
        auto& coords = detail::TopologyAccess::coordinates(topology);
        auto& local = detail::TopologyAccess::local_domain(topology);
        auto& nbs = detail::TopologyAccess::neighbors(topology);
        const auto& cuts = detail::TopologyAccess::axis_cuts(topology);
        dims = {1, 1, 1};
        coords = {0, 0, 0};
        detail::TopologyAccess::rank(topology) = 0;
        detail::TopologyAccess::size(topology) = 1;
        for (int a = 0; a < 3; ++a) {
            local.lower[a] = cuts[a][0];
            local.upper[a] = cuts[a][1];
            box.lower[a] = config.domain.lower[a];
            box.upper[a] = config.domain.upper[a];
            box.periodic[a] = config.domain.periodic[a] ? 1 : 0;
            box.local_lower[a] = cuts[a][0];
            box.local_upper[a] = cuts[a][1];
        }
        box.local_given = 1;
        box.width = config.halo.width;
        box.edges_and_corners = config.halo.include_edges_and_corners ? 1 : 0;
        // src/topology.cpp:85-166 with every direction a self-neighbour
        nbs.clear();
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    if (dx == 0 && dy == 0 && dz == 0) continue;
                    if (!config.halo.include_edges_and_corners &&
                        std::abs(dx) + std::abs(dy) + std::abs(dz) != 1)
                        continue;
                    Neighbor n;
                    n.direction = {dx, dy, dz};
                    bool outside = false;
                    for (int a = 0; a < 3; ++a) {
                        if (n.direction[a] == 0) continue;
                        if (!config.domain.periodic[a]) {
                            outside = true;
                            break;
                        }
                        const double extent = config.domain.upper[a] - config.domain.lower[a];
                        n.periodic_shift[a] = n.direction[a] < 0 ? extent : -extent;
                    }
                    const int raw = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
                    n.direction_slot = raw < 13 ? raw : raw - 1;
                    n.rank = outside ? -1 : 0;
                    n.is_self = n.rank == 0;
                    nbs.push_back(n);
                }
    }

    void drop_device_halo()
    {
        if (plan) check(dsdx_plan_reset(plan));
        device_halo = plan_valid = plan_reused = false;
        plan_owned = plan_ghosts = 0;
    }
};

Decomp::Decomp(MPI_Comm communicator, Config config) : impl_(std::make_unique<Impl>())
{
    if (communicator == MPI_COMM_NULL) throw std::invalid_argument("dsDecompX: MPI communicator is null");
    if (config.execution != ExecutionMode::Cuda)
        throw std::invalid_argument(
            "dsDecompX: the MI355X build executes on the GPU (ExecutionMode::Cuda)");
    for (int a = 0; a < 3; ++a) {
        if (!(config.domain.upper[a] > config.domain.lower[a]))
            throw std::invalid_argument("dsDecompX: global-domain extent must be positive");
        const int d = config.process_grid.dimensions[a];
        if (d != 0 && d != 1)
            throw std::invalid_argument("dsDecompX: the MI355X loopback build runs a 1x1x1 grid");
    }
    if (config.halo.width < 0.0) throw std::invalid_argument("dsDecompX: halo width cannot be negative");
    if (config.process_grid.allow_rank_reorder)
        throw std::invalid_argument(
            "dsDecompX CUDA execution currently requires allow_rank_reorder=false");
    if (!(config.dynamic.imbalance_threshold >= 1.0) ||
        !(config.dynamic.relaxation > 0.0 && config.dynamic.relaxation <= 1.0) ||
        !(config.dynamic.minimum_cell_fraction > 0.0 && config.dynamic.minimum_cell_fraction <= 1.0))
        throw std::invalid_argument("dsDecompX: invalid dynamic load-balance configuration");
    if (config.advanced.selection != BackendSelection::Auto)
        throw std::invalid_argument(std::string("dsDecompX: backend '") +
                                    backend_selection_name(config.advanced.selection) +
                                    "' is not compiled; this build has the single-GPU loopback");
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices <= 0)
        throw std::runtime_error("dsDecompX: could not select a GPU (no visible device)");
    int device = -1;
    if (config.advanced.cuda_device >= 0) {
        if (config.advanced.cuda_device >= devices)
            throw std::runtime_error("dsDecompX: Config::advanced.cuda_device out of range");
        device = config.advanced.cuda_device;
    } else if (config.advanced.cuda_device == kCudaDeviceByLocalRank) {
        device = 0;
    }
    if (device >= 0 && hipSetDevice(device) != hipSuccess)
        throw std::runtime_error("dsDecompX: hipSetDevice failed");
    config.process_grid.dimensions = {1, 1, 1};
    impl_->config = std::move(config);
    impl_->place_cuts();
    impl_->refresh_geometry();
    check(dsdx_plan_create(&impl_->plan));
}

Decomp::~Decomp() = default;
Decomp::Decomp(Decomp&&) noexcept = default;
Decomp& Decomp::operator=(Decomp&&) noexcept = default;

const Topology& Decomp::topology() const noexcept { return impl_->topology; }
BackendKind Decomp::backend_kind() const noexcept { return BackendKind::Loopback; }
bool Decomp::mpi_takes_device_pointers() const noexcept { return false; }

// One rank owns every position: the wrapped coordinate is clamped into the
// single cell of each axis (src/migration.cpp:74-98).
int Decomp::owner_rank(const Vec3d&) const { return 0; }
bool Decomp::any_rank_true(bool local_flag) const { return local_flag; }
bool Decomp::has_out_of_domain_particles(const ParticleInfo&) const { return false; }

void Decomp::rescale_domain(const GlobalDomain& domain)
{
    if (domain.periodic != impl_->config.domain.periodic)
        throw std::invalid_argument(
            "dsDecompX: changing periodicity after topology creation requires reconstructing Decomp");
    for (int a = 0; a < 3; ++a)
        if (!(domain.upper[a] > domain.lower[a]))
            throw std::invalid_argument("dsDecompX: global-domain extent must be positive");
    // The cuts move with the domain: each keeps its fraction of the extent
    // (src/decomp.cpp:117-124); the plan keeps its ghost lists and takes the
    // new extent's shifts at the next refresh.
    auto& cuts = detail::TopologyAccess::axis_cuts(impl_->topology);
    const GlobalDomain previous = impl_->config.domain;
    for (int a = 0; a < 3; ++a) {
        const double old_extent = previous.upper[a] - previous.lower[a];
        const double new_extent = domain.upper[a] - domain.lower[a];
        for (double& cut : cuts[a]) cut = domain.lower[a] + (cut - previous.lower[a]) / old_extent * new_extent;
    }
    impl_->config.domain = domain;
    impl_->refresh_geometry();
    impl_->shifts_changed = true;
}

void Decomp::update_domain(const GlobalDomain& domain)
{
    rescale_domain(domain);
    invalidate_halo();
}

void Decomp::set_halo_width(double width)
{
    if (width < 0.0) throw std::invalid_argument("dsDecompX: halo width cannot be negative");
    if (impl_->config.halo.width != width) {
        impl_->config.halo.width = width;
        impl_->box.width = width;
        invalidate_halo();
    }
}

double Decomp::halo_width() const noexcept { return impl_->config.halo.width; }

void Decomp::invalidate_halo() noexcept { impl_->plan_valid = false; }

HaloResult Decomp::exchange_halo(ParticleInfo& particles, HaloExchangeOptions options)
{
    begin_halo_exchange(particles, options);
    return end_halo_exchange(particles);
}

bool Decomp::can_refresh_halo(const ParticleInfo& particles) const
{
    const std::size_t owned = particles.owned_count();
    ConstParticleView input = particles.const_view(owned);
    input.count = owned;
    if (input.memory_space == MemorySpace::Host) return false;
    return impl_->plan_valid && impl_->plan_owned == owned &&
           impl_->plan_schema == field_signature(input);
}

// begin: rebuilds are launched speculatively with the current storage (the
// classification, scan and ghost scatter queue on the stream without a host
// wait); a reused plan refreshes in place.  end: reads the ghost count back
// (the one host wait of the reference, src/halo.cpp:482-491), grows storage
// and scatters again if the ghosts did not fit.
void Decomp::begin_halo_exchange(ParticleInfo& particles, HaloExchangeOptions options)
{
    Impl& m = *impl_;
    if (m.pending.active)
        throw std::logic_error("dsDecompX: a split-phase halo exchange is already in flight");
    if (m.operation_active) throw std::logic_error("dsDecompX: an operation is already in flight");
    if (options.rebuild) invalidate_halo();
    const std::size_t owned = particles.owned_count();
    ConstParticleView input = particles.const_view(owned);
    input.count = owned;
    require_device(input.memory_space, "exchange_halo");
    if (!options.stream)
        throw std::invalid_argument("dsDecompX CUDA halo requires an explicit communication stream");
    if (owned && (!input.x || !input.y || !input.z || !input.vx || !input.vy || !input.vz ||
                  !input.mass || !input.charge || !input.type || !input.global_id))
        throw std::invalid_argument("dsDecompX: incomplete CUDA particle view");
    std::vector<dsdx_field> fields = all_fields(input, owned);
    const std::string schema = field_signature(input);
    const bool reuse = m.plan_valid && m.plan_owned == owned && m.plan_schema == schema;
    const bool compact = options.positions_only && reuse && fields.empty();

    Impl::Pending pending;
    pending.options = options;
    pending.owned = owned;
    pending.reuse = reuse;
    pending.compact = compact;
    pending.schema = schema;
    if (reuse) {
        StorageResizeResult storage = particles.ensure_capacity(owned + m.plan_ghosts, owned);
        if (storage.reallocated && compact)
            throw std::logic_error(
                "dsDecompX: a positions-only halo refresh needed a larger particle storage; "
                "the ghosts' other fields are gone (use a full exchange)");
        pending.reallocated = storage.reallocated;
        const dsdx_particles p = to_abi(storage.view);
        check(dsdx_halo_refresh(m.plan, m.shifts_changed ? &m.box : nullptr, &p, fields.data(),
                                static_cast<uint32_t>(fields.size()), owned, compact ? 1 : 0,
                                options.stream));
        m.shifts_changed = false;
    } else {
        m.plan_valid = false;
        m.device_halo = false;
        const dsdx_particles p = to_abi(input);
        const std::size_t capacity = field_capacity(input, particles.capacity());
        const bool wrap = m.wrap_in_build;
        m.wrap_in_build = false;
        check((wrap ? dsdx_migrate_halo_build : dsdx_halo_build)(
            m.plan, &m.box, &p, fields.data(), static_cast<uint32_t>(fields.size()), owned, capacity,
            options.stream, nullptr));
        m.shifts_changed = false;
    }
    pending.active = true;
    m.pending = std::move(pending);
    m.operation_active = true;
}

HaloResult Decomp::end_halo_exchange(ParticleInfo& particles)
{
    Impl& m = *impl_;
    if (!m.pending.active)
        throw std::logic_error("dsDecompX: end_halo_exchange without a matching begin_halo_exchange");
    Impl::Pending pending = std::move(m.pending);
    m.pending = Impl::Pending{};
    m.operation_active = false;
    const std::size_t owned = pending.owned;
    std::size_t ghosts = m.plan_ghosts;
    bool reallocated = pending.reallocated;
    if (!pending.reuse) {
        uint64_t g = 0;
        int32_t overflow = 0;
        check(dsdx_plan_query(m.plan, nullptr, nullptr, &g, nullptr, &overflow, pending.options.stream));
        ghosts = static_cast<std::size_t>(g);
        if (overflow) {
            StorageResizeResult storage = particles.ensure_capacity(owned + ghosts, owned);
            reallocated = storage.reallocated;
            ConstParticleView view = as_const_view(storage.view);
            std::vector<dsdx_field> fields = all_fields(view, owned + ghosts);
            const dsdx_particles p = to_abi(storage.view);
            check(dsdx_halo_scatter(m.plan, &p, fields.data(), static_cast<uint32_t>(fields.size()),
                                    owned, field_capacity(view, storage.view.count),
                                    pending.options.stream));
        }
        m.plan_valid = true;
        m.plan_owned = owned;
        m.plan_ghosts = ghosts;
        m.plan_schema = pending.schema;
    }
    m.device_halo = true;
    m.plan_reused = pending.reuse;
    const bool layout_changed = pending.options.rebuild || particles.ghost_count() != ghosts ||
                                (ghosts != 0 && !pending.reuse);
    particles.set_counts(owned, ghosts);
    if (m.perf_enabled) {
        ++m.stats.halo_calls;
        std::size_t record = pending.compact ? 3 * sizeof(double) : sizeof(ParticleRecord);
        const ConstParticleView v = particles.const_view(owned + ghosts);
        if (!pending.compact)
            for (std::size_t f = 0; f < v.field_count; ++f) record += v.fields[f].record_bytes();
        m.stats.bytes_sent += ghosts * record;
        m.stats.bytes_received += ghosts * record;
    }
    return HaloResult{ghosts, layout_changed, reallocated, pending.reuse};
}

namespace {

void check_migration_view(const ConstParticleView& input, std::size_t owned, ExecutionStream stream)
{
    if (!stream)
        throw std::invalid_argument(
            "dsDecompX CUDA migration requires an explicit communication stream");
    if (owned && (!input.x || !input.y || !input.z || !input.vx || !input.vy || !input.vz ||
                  !input.fx || !input.fy || !input.fz || !input.mass || !input.charge ||
                  !input.type || !input.global_id))
        throw std::invalid_argument("dsDecompX: incomplete CUDA migration view");
}

}  // namespace

HaloResult Decomp::migrate_and_exchange_halo(ParticleInfo& particles, HaloExchangeOptions options)
{
    Impl& m = *impl_;
    if (m.operation_active)
        throw std::logic_error("dsDecompX: migration cannot overlap an active halo operation");
    const std::size_t owned = particles.owned_count();
    ConstParticleView input = particles.const_view(owned);
    input.count = owned;
    require_device(input.memory_space, "migrate");
    check_migration_view(input, owned, options.stream);
    all_fields(input, owned);
    // migrate()'s bookkeeping, then a rebuild whose classification pass wraps
    // the positions first (dsdx_migrate_halo_build).
    m.drop_device_halo();
    particles.ensure_capacity(owned, 0);
    particles.set_counts(owned, 0);
    if (m.perf_enabled) ++m.stats.migration_calls;
    options.rebuild = true;
    m.wrap_in_build = true;
    try {
        begin_halo_exchange(particles, options);
    } catch (...) {
        m.wrap_in_build = false;
        throw;
    }
    return end_halo_exchange(particles);
}

MigrationResult Decomp::migrate(ParticleInfo& particles, ExecutionStream stream)
{
    Impl& m = *impl_;
    if (m.operation_active)
        throw std::logic_error("dsDecompX: migration cannot overlap an active halo operation");
    const std::size_t owned = particles.owned_count();
    ConstParticleView input = particles.const_view(owned);
    input.count = owned;
    require_device(input.memory_space, "migrate");
    check_migration_view(input, owned, stream);
    all_fields(input, owned);
    // At 1x1x1 every particle stays, in order, with its periodic coordinates
    // wrapped; ghosts are discarded and the device halo dropped
    // (src/migration.cpp:355-359, :437-479).
    check(dsdx_migrate_wrap(&m.box, const_cast<double*>(input.x), const_cast<double*>(input.y),
                            const_cast<double*>(input.z), owned, stream));
    m.drop_device_halo();
    particles.ensure_capacity(owned, 0);
    particles.set_counts(owned, 0);
    if (m.perf_enabled) ++m.stats.migration_calls;
    MigrationResult result;
    result.previous_owned_count = owned;
    result.retained_count = owned;
    result.new_owned_count = owned;
    return result;
}

RebalanceResult Decomp::rebalance(ParticleInfo& particles, RebalanceOptions, ExecutionStream stream)
{
    // One rank: there is no cut plane to move (src/runtime.cpp:345-447 needs
    // more than one rank), so rebalance is a migration.
    RebalanceResult result;
    result.migration = migrate(particles, stream);
    return result;
}

void Decomp::reverse_halo(ParticleInfo& particles, ExecutionStream stream)
{
    Impl& m = *impl_;
    if (m.operation_active)
        throw std::logic_error("dsDecompX: reverse halo cannot overlap an active operation");
    StorageResizeResult storage = particles.ensure_capacity(particles.size(), particles.size());
    require_device(storage.view.memory_space, "reverse_halo");
    if (!stream) throw std::invalid_argument("dsDecompX: device reverse halo requires an explicit stream");
    if (!m.device_halo) return;
    const std::size_t owned = particles.owned_count();
    if (storage.view.count < owned + m.plan_ghosts || !storage.view.fx || !storage.view.fy ||
        !storage.view.fz)
        throw std::invalid_argument("dsDecompX: invalid device reverse-halo view");
    check(dsdx_reverse_halo(m.plan, storage.view.fx, storage.view.fy, storage.view.fz, owned, stream));
}

void Decomp::reverse_exchange(ParticleInfo& particles,
                              const std::vector<ReverseExchangeField>& fields,
                              ExecutionStream stream)
{
    Impl& m = *impl_;
    if (m.operation_active)
        throw std::logic_error("dsDecompX: reverse exchange cannot overlap an active operation");
    StorageResizeResult storage = particles.ensure_capacity(particles.size(), particles.size());
    require_device(storage.view.memory_space, "reverse_exchange");
    if (!stream)
        throw std::invalid_argument("dsDecompX: device reverse exchange requires an explicit stream");
    if (fields.size() > DSDX_MAX_FIELDS)
        throw std::invalid_argument(
            "dsDecompX: at most 16 fields can be reverse-exchanged on CUDA particle memory");
    const std::size_t owned = particles.owned_count();
    const std::size_t required = owned + (m.device_halo ? m.plan_ghosts : 0);
    std::unordered_set<std::string> names;
    std::vector<dsdx_field> out;
    for (const auto& spec : fields) {
        if (!names.insert(spec.name).second)
            throw std::invalid_argument("dsDecompX: duplicate reverse exchange field '" + spec.name + "'");
        const ParticleFieldView& f = find_field(storage.view, spec.name, required, "reverse exchange");
        out.push_back({f.data, field_code(f.type), static_cast<int32_t>(f.components),
                       static_cast<int32_t>(spec.operation)});
    }
    if (!m.device_halo || out.empty()) return;
    check(dsdx_reverse_exchange(m.plan, out.data(), static_cast<uint32_t>(out.size()), owned, stream));
}

void Decomp::forward_exchange(ParticleInfo& particles, const std::vector<std::string>& fields,
                              ExecutionStream stream)
{
    Impl& m = *impl_;
    if (m.operation_active)
        throw std::logic_error("dsDecompX: forward exchange cannot overlap an active operation");
    StorageResizeResult storage = particles.ensure_capacity(particles.size(), particles.size());
    require_device(storage.view.memory_space, "forward_exchange");
    if (!stream)
        throw std::invalid_argument("dsDecompX: device forward exchange requires an explicit stream");
    if (fields.size() > DSDX_MAX_FIELDS)
        throw std::invalid_argument(
            "dsDecompX: at most 16 fields can be forward-exchanged on CUDA particle memory");
    const std::size_t owned = particles.owned_count();
    const std::size_t required = owned + (m.device_halo ? m.plan_ghosts : 0);
    std::unordered_set<std::string> names;
    std::vector<dsdx_field> out;
    for (const auto& name : fields) {
        if (!names.insert(name).second)
            throw std::invalid_argument("dsDecompX: duplicate forward exchange field '" + name + "'");
        const ParticleFieldView& f = find_field(storage.view, name, required, "forward exchange");
        if (f.record_bytes() % 4 != 0)
            throw std::invalid_argument("dsDecompX: forward exchange field '" + name +
                                        "' must be a multiple of 4 bytes per particle");
        out.push_back({f.data, field_code(f.type), static_cast<int32_t>(f.components), DSDX_REPLACE});
    }
    if (!m.device_halo || out.empty()) return;
    check(dsdx_forward_exchange(m.plan, out.data(), static_cast<uint32_t>(out.size()), owned, stream));
}

void Decomp::set_performance_counters_enabled(bool enabled, bool neighbor_detail) noexcept
{
    impl_->perf_enabled = enabled;
    impl_->neighbor_detail = neighbor_detail;
}

PerfStats Decomp::get_stats() const { return impl_->stats; }
PerfStats Decomp::get_global_stats() const
{
    PerfStats s = impl_->stats;
    s.neighbor_bytes.clear();
    return s;
}
void Decomp::reset_stats() noexcept { impl_->stats = PerfStats{}; }

}  // namespace dsdecompx
