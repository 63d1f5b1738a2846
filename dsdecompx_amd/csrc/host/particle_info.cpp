// particle_info.cpp -- ParticleInfo storage (host vectors, library-owned HIP
// device/managed arrays, borrowed views and providers).  Semantics of the
// reference's src/runtime.cpp:451-755: capacity, growth by at least 1.5x
// (minimum 64) preserving the first `preserve` particles, borrowed storage
// that is too small is a length_error unless a Provider grows it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "dsdecompx.hpp"

namespace dsdecompx {
namespace {

void hip_require(hipError_t status, const char* what)
{
    if (status != hipSuccess)
        throw std::runtime_error(std::string("dsDecompX HIP: ") + what + ": " +
                                 hipGetErrorString(status));
}

template <class T> void device_alloc(MemorySpace space, T** p, std::size_t count)
{
    const std::size_t bytes = count * sizeof(T);
    if (space == MemorySpace::CudaManaged)
        hip_require(hipMallocManaged(reinterpret_cast<void**>(p), bytes), "hipMallocManaged");
    else
        hip_require(hipMalloc(reinterpret_cast<void**>(p), bytes), "hipMalloc");
}

template <class T> void device_free(T*& p) noexcept
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_device_view(ParticleView& v) noexcept
{
    device_free(v.x); device_free(v.y); device_free(v.z);
    device_free(v.vx); device_free(v.vy); device_free(v.vz);
    device_free(v.fx); device_free(v.fy); device_free(v.fz);
    device_free(v.mass); device_free(v.charge);
    device_free(v.type); device_free(v.global_id);
    v = ParticleView{};
}

ParticleView allocate_device_view(MemorySpace space, std::size_t capacity)
{
    ParticleView v;
    v.memory_space = space;
    v.count = capacity;
    if (capacity == 0) return v;
    try {
        device_alloc(space, &v.x, capacity); device_alloc(space, &v.y, capacity);
        device_alloc(space, &v.z, capacity); device_alloc(space, &v.vx, capacity);
        device_alloc(space, &v.vy, capacity); device_alloc(space, &v.vz, capacity);
        device_alloc(space, &v.fx, capacity); device_alloc(space, &v.fy, capacity);
        device_alloc(space, &v.fz, capacity); device_alloc(space, &v.mass, capacity);
        device_alloc(space, &v.charge, capacity); device_alloc(space, &v.type, capacity);
        device_alloc(space, &v.global_id, capacity);
    } catch (...) {
        free_device_view(v);
        throw;
    }
    return v;
}

void copy_device_view(const ParticleView& from, ParticleView& to, std::size_t count)
{
    if (!count) return;
    auto copy = [count](auto* dst, const auto* src) {
        using T = std::remove_pointer_t<decltype(dst)>;
        hip_require(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyDefault),
                    "hipMemcpy while growing ParticleInfo");
    };
    copy(to.x, from.x); copy(to.y, from.y); copy(to.z, from.z);
    copy(to.vx, from.vx); copy(to.vy, from.vy); copy(to.vz, from.vz);
    copy(to.fx, from.fx); copy(to.fy, from.fy); copy(to.fz, from.fz);
    copy(to.mass, from.mass); copy(to.charge, from.charge);
    copy(to.type, from.type); copy(to.global_id, from.global_id);
}

void attach_fields(ParticleView& view, std::vector<ParticleFieldView>& fields)
{
    view.fields = fields.empty() ? nullptr : fields.data();
    view.field_count = fields.size();
}

}  // namespace

ParticleInfo::~ParticleInfo() { release_device_memory(); }

ParticleInfo::ParticleInfo(ParticleInfo&& other) noexcept { *this = std::move(other); }

ParticleInfo& ParticleInfo::operator=(ParticleInfo&& other) noexcept
{
    if (this == &other) return *this;
    release_device_memory();
    x = std::move(other.x); y = std::move(other.y); z = std::move(other.z);
    vx = std::move(other.vx); vy = std::move(other.vy); vz = std::move(other.vz);
    fx = std::move(other.fx); fy = std::move(other.fy); fz = std::move(other.fz);
    mass = std::move(other.mass); charge = std::move(other.charge);
    type = std::move(other.type); global_id = std::move(other.global_id);
    num_particles = other.num_particles;
    external_ = other.external_;
    owns_device_memory_ = other.owns_device_memory_;
    owned_count_ = other.owned_count_;
    ghost_count_ = other.ghost_count_;
    fields_ = std::move(other.fields_);
    storage_ = other.storage_;
    provider_ = other.provider_;
    attach_fields(storage_, fields_);
    other.num_particles = 0;
    other.external_ = other.owns_device_memory_ = false;
    other.owned_count_ = other.ghost_count_ = 0;
    other.fields_.clear();
    other.storage_ = ParticleView{};
    other.provider_ = nullptr;
    return *this;
}

void ParticleInfo::release_device_memory() noexcept
{
    if (!owns_device_memory_) return;
    free_device_view(storage_);
    owns_device_memory_ = false;
}

ParticleInfo ParticleInfo::allocate(MemorySpace memory_space, std::size_t owned)
{
    if (memory_space == MemorySpace::Host) return ParticleInfo(owned);
    ParticleInfo result;
    result.external_ = true;
    result.owns_device_memory_ = true;
    result.owned_count_ = owned;
    result.storage_ = allocate_device_view(memory_space, owned);
    return result;
}

ParticleInfo::ParticleInfo(std::size_t owned, std::size_t ghosts, ParticleView storage)
    : external_(true), owned_count_(owned), ghost_count_(ghosts), storage_(storage)
{
    if (storage_.count < owned + ghosts)
        throw std::invalid_argument(
            "dsDecompX: external particle view is smaller than its particle counts");
    if (storage.fields && storage.field_count)
        fields_.assign(storage.fields, storage.fields + storage.field_count);
    attach_fields(storage_, fields_);
}

ParticleInfo::ParticleInfo(std::size_t owned, std::size_t ghosts, Provider& provider)
    : ParticleInfo(owned, ghosts, provider.particle_info_view(owned + ghosts))
{
    provider_ = &provider;
}

ParticleInfo ParticleInfo::external(MemorySpace memory_space, std::size_t capacity,
                                    std::size_t owned, std::size_t ghosts)
{
    ParticleView view;
    view.memory_space = memory_space;
    view.count = capacity;
    return ParticleInfo(owned, ghosts, view);
}

ParticleInfo& ParticleInfo::positions(double* px, double* py, double* pz) noexcept
{
    storage_.x = px; storage_.y = py; storage_.z = pz;
    return *this;
}

ParticleInfo& ParticleInfo::velocities(double* pvx, double* pvy, double* pvz) noexcept
{
    storage_.vx = pvx; storage_.vy = pvy; storage_.vz = pvz;
    return *this;
}

ParticleInfo& ParticleInfo::forces(double* pfx, double* pfy, double* pfz) noexcept
{
    storage_.fx = pfx; storage_.fy = pfy; storage_.fz = pfz;
    return *this;
}

ParticleInfo& ParticleInfo::properties(double* pm, double* pq, int* pt, int* pid) noexcept
{
    storage_.mass = pm; storage_.charge = pq; storage_.type = pt; storage_.global_id = pid;
    return *this;
}

std::size_t ParticleInfo::capacity() const noexcept { return external_ ? storage_.count : x.size(); }

MemorySpace ParticleInfo::memory_space() const noexcept
{
    return external_ ? storage_.memory_space : MemorySpace::Host;
}

double* ParticleInfo::x_data() noexcept { return external_ ? storage_.x : x.data(); }
double* ParticleInfo::y_data() noexcept { return external_ ? storage_.y : y.data(); }
double* ParticleInfo::z_data() noexcept { return external_ ? storage_.z : z.data(); }
double* ParticleInfo::vx_data() noexcept { return external_ ? storage_.vx : vx.data(); }
double* ParticleInfo::vy_data() noexcept { return external_ ? storage_.vy : vy.data(); }
double* ParticleInfo::vz_data() noexcept { return external_ ? storage_.vz : vz.data(); }
double* ParticleInfo::fx_data() noexcept { return external_ ? storage_.fx : fx.data(); }
double* ParticleInfo::fy_data() noexcept { return external_ ? storage_.fy : fy.data(); }
double* ParticleInfo::fz_data() noexcept { return external_ ? storage_.fz : fz.data(); }
double* ParticleInfo::mass_data() noexcept { return external_ ? storage_.mass : mass.data(); }
double* ParticleInfo::charge_data() noexcept { return external_ ? storage_.charge : charge.data(); }
int* ParticleInfo::type_data() noexcept { return external_ ? storage_.type : type.data(); }
int* ParticleInfo::global_id_data() noexcept
{
    return external_ ? storage_.global_id : global_id.data();
}

void ParticleInfo::commit_counts(std::size_t owned, std::size_t ghosts) noexcept
{
    owned_count_ = owned;
    ghost_count_ = ghosts;
}

void ParticleInfo::reset_counts(std::size_t owned, std::size_t ghosts)
{
    const std::size_t total = owned + ghosts;
    if (read_view(total).count < total)
        throw std::invalid_argument("dsDecompX: particle storage is smaller than the supplied counts");
    commit_counts(owned, ghosts);
}

void ParticleInfo::resize(std::size_t count)
{
    x.resize(count); y.resize(count); z.resize(count);
    vx.resize(count); vy.resize(count); vz.resize(count);
    fx.resize(count); fy.resize(count); fz.resize(count);
    mass.resize(count); charge.resize(count);
    type.resize(count); global_id.resize(count);
    num_particles = count;
}

ConstParticleView ParticleInfo::host_const_view(std::size_t count) const
{
    ConstParticleView v;
    v.memory_space = MemorySpace::Host;
    v.count = count;
    v.x = x.data(); v.y = y.data(); v.z = z.data();
    v.vx = vx.data(); v.vy = vy.data(); v.vz = vz.data();
    v.fx = fx.data(); v.fy = fy.data(); v.fz = fz.data();
    v.mass = mass.data(); v.charge = charge.data();
    v.type = type.data(); v.global_id = global_id.data();
    v.fields = fields_.empty() ? nullptr : fields_.data();
    v.field_count = fields_.size();
    return v;
}

ParticleView ParticleInfo::host_view(std::size_t count)
{
    ParticleView v;
    v.memory_space = MemorySpace::Host;
    v.count = count;
    v.x = x.data(); v.y = y.data(); v.z = z.data();
    v.vx = vx.data(); v.vy = vy.data(); v.vz = vz.data();
    v.fx = fx.data(); v.fy = fy.data(); v.fz = fz.data();
    v.mass = mass.data(); v.charge = charge.data();
    v.type = type.data(); v.global_id = global_id.data();
    attach_fields(v, fields_);
    return v;
}

void ParticleInfo::resize_owned(std::size_t count)
{
    if (external_) {
        if (!owns_device_memory_)
            throw std::logic_error("dsDecompX: resize_owned cannot resize borrowed particle memory");
        resize_storage(count, std::min(count, owned_count_));
        commit_counts(count, 0);
        return;
    }
    for (const auto& f : fields_)
        if (f.capacity < count)
            throw std::length_error("dsDecompX: custom particle field '" + f.name +
                                    "' has insufficient capacity");
    resize(count);
    commit_counts(count, 0);
}

ConstParticleView ParticleInfo::read_view(std::size_t count) const
{
    if (external_) {
        if (storage_.count < count)
            throw std::invalid_argument(
                "dsDecompX: external particle memory has insufficient capacity");
        return as_const_view(storage_);
    }
    return host_const_view(count);
}

ParticleInfo& ParticleInfo::add_field_view(ParticleFieldView field)
{
    if (field.name.empty() || !field.data || field.components == 0 || field.capacity < size())
        throw std::invalid_argument("dsDecompX: invalid or undersized custom particle field");
    for (const auto& f : fields_)
        if (f.name == field.name)
            throw std::invalid_argument("dsDecompX: duplicate custom particle field '" +
                                        field.name + "'");
    fields_.push_back(std::move(field));
    attach_fields(storage_, fields_);
    return *this;
}

StorageResizeResult ParticleInfo::resize_storage(std::size_t required, std::size_t preserve)
{
    for (const auto& f : fields_)
        if (f.capacity < required)
            throw std::length_error("dsDecompX: custom particle field '" + f.name +
                                    "' has insufficient capacity");
    if (!external_) {
        const double* before = x.data();
        resize(required);
        return {host_view(required), before != x.data()};
    }
    if (required <= storage_.count) return {storage_, false};
    if (owns_device_memory_) {
        const std::size_t grown =
            std::max(required, std::max<std::size_t>(64, storage_.count + storage_.count / 2));
        // Stream-ordered work may still read the old arrays.
        hip_require(hipDeviceSynchronize(), "synchronize before growing ParticleInfo");
        ParticleView replacement = allocate_device_view(storage_.memory_space, grown);
        try {
            copy_device_view(storage_, replacement, std::min(preserve, storage_.count));
        } catch (...) {
            free_device_view(replacement);
            throw;
        }
        free_device_view(storage_);
        storage_ = replacement;
        attach_fields(storage_, fields_);
        return {storage_, true};
    }
    if (!provider_) throw std::length_error("dsDecompX: fixed external particle storage is too small");
    StorageResizeResult result = provider_->resize_particle_info(required, preserve);
    storage_ = result.view;
    attach_fields(storage_, fields_);
    result.view = storage_;
    return result;
}

void ParticleInfo::reset_view(ParticleView storage)
{
    if (storage.count < size())
        throw std::invalid_argument(
            "dsDecompX: replacement particle view is smaller than current particles");
    release_device_memory();
    external_ = true;
    owns_device_memory_ = false;
    provider_ = nullptr;
    storage_ = storage;
    fields_.clear();
    if (storage.fields && storage.field_count)
        fields_.assign(storage.fields, storage.fields + storage.field_count);
    attach_fields(storage_, fields_);
}

}  // namespace dsdecompx
