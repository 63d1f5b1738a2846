// particle_info.cpp -- where a ParticleInfo keeps its 13 particle arrays.
//
// Three kinds of backing:
//   host vectors  the public std::vector members (host storage; the GPU
//                 operations reject it);
//   borrowed      arrays the caller owns: a view, external() plus the pointer
//                 setters, or a Provider that can hand out larger ones;
//   block         library-owned device (or managed) memory, one allocation per
//                 capacity holding all 13 arrays back to back, each starting
//                 on a 256-byte boundary.
// A block grows to the required particle count plus a quarter, rounded up to
// whole pages of 1024 particles, and keeps the first `preserve` particles of
// every array.  Growth waits for the device first: work queued on any stream
// may still read the old arrays.  Registered fields are borrowed in every
// case; views handed out carry them.  The exception types are the
// reference's (invalid_argument for bad sizes and views, length_error for
// storage that cannot hold the particles, logic_error for resizing borrowed
// arrays, runtime_error for device failures).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <stdexcept>
#include <string>
#include <utility>

#include "dsdecompx.hpp"

namespace dsdecompx {

namespace {

enum class Backing { HostVectors, Borrowed, Block };

constexpr std::size_t kPage = 1024;   // particles
constexpr std::size_t kAlign = 256;   // bytes

void hip_check(hipError_t status, const char* what)
{
    if (status != hipSuccess)
        throw std::runtime_error(std::string("dsDecompX HIP: ") + what + ": " + hipGetErrorString(status));
}

// The arrays of a view in a fixed order: 11 of doubles, then 2 of ints.
constexpr double* ParticleView::*kDoubles[11] = {
    &ParticleView::x, &ParticleView::y, &ParticleView::z, &ParticleView::vx, &ParticleView::vy,
    &ParticleView::vz, &ParticleView::fx, &ParticleView::fy, &ParticleView::fz, &ParticleView::mass,
    &ParticleView::charge};
constexpr int* ParticleView::*kInts[2] = {&ParticleView::type, &ParticleView::global_id};

std::size_t round_up(std::size_t v, std::size_t to) { return (v + to - 1) / to * to; }

std::size_t block_bytes(std::size_t capacity)
{
    return 11 * round_up(capacity * sizeof(double), kAlign) + 2 * round_up(capacity * sizeof(int), kAlign);
}

// The arrays of a block of `capacity` particles at `base`.
ParticleView carve(void* base, std::size_t capacity, MemorySpace space)
{
    ParticleView v;
    v.memory_space = space;
    v.count = capacity;
    auto* at = static_cast<unsigned char*>(base);
    for (auto member : kDoubles) {
        v.*member = reinterpret_cast<double*>(at);
        at += round_up(capacity * sizeof(double), kAlign);
    }
    for (auto member : kInts) {
        v.*member = reinterpret_cast<int*>(at);
        at += round_up(capacity * sizeof(int), kAlign);
    }
    return v;
}

void* allocate_block(MemorySpace space, std::size_t capacity)
{
    void* p = nullptr;
    if (space == MemorySpace::CudaManaged)
        hip_check(hipMallocManaged(&p, block_bytes(capacity)), "allocate managed particle storage");
    else
        hip_check(hipMalloc(&p, block_bytes(capacity)), "allocate device particle storage");
    return p;
}

// Capacity of a block that has to hold `required` particles.
std::size_t block_capacity(std::size_t required)
{
    return std::max(kPage, round_up(required + required / 4, kPage));
}

}  // namespace

struct ParticleInfo::State {
    Backing backing = Backing::HostVectors;
    ParticleView arrays;              // borrowed or block: pointers, capacity, space (no fields)
    void* block = nullptr;            // block: the allocation
    Provider* provider = nullptr;     // borrowed: can replace the arrays with larger ones
    std::vector<ParticleFieldView> fields;
    std::size_t owned = 0;
    std::size_t ghosts = 0;

    State() = default;
    State(const State&) = delete;
    State& operator=(const State&) = delete;
    ~State() { drop_block(); }

    void drop_block() noexcept
    {
        if (block) (void)hipFree(block);
        block = nullptr;
    }

    ParticleView view() const
    {
        ParticleView v = arrays;
        v.fields = fields.empty() ? nullptr : const_cast<ParticleFieldView*>(fields.data());
        v.field_count = fields.size();
        return v;
    }

    void adopt(const ParticleView& v)
    {
        arrays = v;
        arrays.fields = nullptr;
        arrays.field_count = 0;
        fields.clear();
        if (v.fields && v.field_count) fields.assign(v.fields, v.fields + v.field_count);
    }

    void require_field_capacity(std::size_t required) const
    {
        for (const ParticleFieldView& f : fields)
            if (f.capacity < required)
                throw std::length_error("dsDecompX: registered field '" + f.name + "' holds fewer than " +
                                        std::to_string(required) + " particles");
    }

    // A new block of room for `required` particles with the first `preserve`
    // of every array copied over.
    void grow_block(std::size_t required, std::size_t preserve)
    {
        const std::size_t capacity = block_capacity(required);
        const MemorySpace space = arrays.memory_space;
        hip_check(hipDeviceSynchronize(), "wait before moving particle storage");
        void* fresh = allocate_block(space, capacity);
        const ParticleView next = carve(fresh, capacity, space);
        const std::size_t keep = block ? std::min(preserve, arrays.count) : 0;
        if (keep) {
            for (auto member : kDoubles) {
                const hipError_t e = hipMemcpy(next.*member, arrays.*member, keep * sizeof(double),
                                               hipMemcpyDeviceToDevice);
                if (e != hipSuccess) {
                    (void)hipFree(fresh);
                    hip_check(e, "copy particle storage");
                }
            }
            for (auto member : kInts) {
                const hipError_t e = hipMemcpy(next.*member, arrays.*member, keep * sizeof(int),
                                               hipMemcpyDeviceToDevice);
                if (e != hipSuccess) {
                    (void)hipFree(fresh);
                    hip_check(e, "copy particle storage");
                }
            }
        }
        drop_block();
        block = fresh;
        arrays = next;
    }
};

ParticleInfo::State& ParticleInfo::state()
{
    if (!state_) state_ = std::make_unique<State>();
    return *state_;
}

ParticleInfo::ParticleInfo() : state_(std::make_unique<State>()) {}

ParticleInfo::ParticleInfo(std::size_t owned) : state_(std::make_unique<State>()) { resize_owned(owned); }

ParticleInfo::~ParticleInfo() = default;

ParticleInfo::ParticleInfo(ParticleInfo&& other) noexcept
    : x(std::move(other.x)), y(std::move(other.y)), z(std::move(other.z)), vx(std::move(other.vx)),
      vy(std::move(other.vy)), vz(std::move(other.vz)), fx(std::move(other.fx)), fy(std::move(other.fy)),
      fz(std::move(other.fz)), mass(std::move(other.mass)), charge(std::move(other.charge)),
      type(std::move(other.type)), global_id(std::move(other.global_id)),
      num_particles(std::exchange(other.num_particles, 0)), state_(std::move(other.state_))
{
}

ParticleInfo& ParticleInfo::operator=(ParticleInfo&& other) noexcept
{
    if (this != &other) {
        x = std::move(other.x); y = std::move(other.y); z = std::move(other.z);
        vx = std::move(other.vx); vy = std::move(other.vy); vz = std::move(other.vz);
        fx = std::move(other.fx); fy = std::move(other.fy); fz = std::move(other.fz);
        mass = std::move(other.mass); charge = std::move(other.charge);
        type = std::move(other.type); global_id = std::move(other.global_id);
        num_particles = std::exchange(other.num_particles, 0);
        state_ = std::move(other.state_);   // frees this object's block, if any
    }
    return *this;
}

ParticleInfo::ParticleInfo(std::size_t owned, std::size_t ghosts, ParticleView storage)
    : state_(std::make_unique<State>())
{
    if (storage.count < owned + ghosts)
        throw std::invalid_argument("dsDecompX: a particle view of " + std::to_string(storage.count) +
                                    " slots cannot hold " + std::to_string(owned + ghosts) + " particles");
    state_->backing = Backing::Borrowed;
    state_->adopt(storage);
    state_->owned = owned;
    state_->ghosts = ghosts;
}

ParticleInfo::ParticleInfo(std::size_t owned, std::size_t ghosts, Provider& provider)
    : ParticleInfo(owned, ghosts, provider.particle_info_view(owned + ghosts))
{
    state_->provider = &provider;
}

ParticleInfo ParticleInfo::allocate(MemorySpace memory_space, std::size_t owned)
{
    if (memory_space == MemorySpace::Host) return ParticleInfo(owned);
    ParticleInfo p;
    State& s = *p.state_;
    s.backing = Backing::Block;
    s.arrays.memory_space = memory_space;
    s.grow_block(owned, 0);
    s.owned = owned;
    return p;
}

ParticleInfo ParticleInfo::external(MemorySpace memory_space, std::size_t capacity, std::size_t owned,
                                    std::size_t ghosts)
{
    ParticleView v;
    v.memory_space = memory_space;
    v.count = capacity;
    return ParticleInfo(owned, ghosts, v);
}

ParticleInfo& ParticleInfo::positions(double* px, double* py, double* pz) noexcept
{
    if (!state_) state_.reset(new (std::nothrow) State());
    if (state_) {
        state_->arrays.x = px; state_->arrays.y = py; state_->arrays.z = pz;
    }
    return *this;
}

ParticleInfo& ParticleInfo::velocities(double* pvx, double* pvy, double* pvz) noexcept
{
    if (!state_) state_.reset(new (std::nothrow) State());
    if (state_) {
        state_->arrays.vx = pvx; state_->arrays.vy = pvy; state_->arrays.vz = pvz;
    }
    return *this;
}

ParticleInfo& ParticleInfo::forces(double* pfx, double* pfy, double* pfz) noexcept
{
    if (!state_) state_.reset(new (std::nothrow) State());
    if (state_) {
        state_->arrays.fx = pfx; state_->arrays.fy = pfy; state_->arrays.fz = pfz;
    }
    return *this;
}

ParticleInfo& ParticleInfo::properties(double* pm, double* pq, int* pt, int* pid) noexcept
{
    if (!state_) state_.reset(new (std::nothrow) State());
    if (state_) {
        state_->arrays.mass = pm; state_->arrays.charge = pq;
        state_->arrays.type = pt; state_->arrays.global_id = pid;
    }
    return *this;
}

std::size_t ParticleInfo::owned_count() const noexcept { return state_ ? state_->owned : 0; }
std::size_t ParticleInfo::ghost_count() const noexcept { return state_ ? state_->ghosts : 0; }
std::size_t ParticleInfo::size() const noexcept { return owned_count() + ghost_count(); }

std::size_t ParticleInfo::capacity() const noexcept
{
    return state_ && state_->backing != Backing::HostVectors ? state_->arrays.count : x.size();
}

MemorySpace ParticleInfo::memory_space() const noexcept
{
    return state_ && state_->backing != Backing::HostVectors ? state_->arrays.memory_space
                                                              : MemorySpace::Host;
}

bool ParticleInfo::uses_external_memory() const noexcept
{
    return state_ && state_->backing == Backing::Borrowed;
}

// Host storage hands out the vectors; every other backing its arrays.
#define DSDX_ARRAY_ACCESSOR(T, name)                                                     \
    T* ParticleInfo::name##_data() noexcept                                             \
    {                                                                                    \
        return state_ && state_->backing != Backing::HostVectors ? state_->arrays.name : name.data(); \
    }
DSDX_ARRAY_ACCESSOR(double, x)
DSDX_ARRAY_ACCESSOR(double, y)
DSDX_ARRAY_ACCESSOR(double, z)
DSDX_ARRAY_ACCESSOR(double, vx)
DSDX_ARRAY_ACCESSOR(double, vy)
DSDX_ARRAY_ACCESSOR(double, vz)
DSDX_ARRAY_ACCESSOR(double, fx)
DSDX_ARRAY_ACCESSOR(double, fy)
DSDX_ARRAY_ACCESSOR(double, fz)
DSDX_ARRAY_ACCESSOR(double, mass)
DSDX_ARRAY_ACCESSOR(double, charge)
DSDX_ARRAY_ACCESSOR(int, type)
DSDX_ARRAY_ACCESSOR(int, global_id)
#undef DSDX_ARRAY_ACCESSOR

void ParticleInfo::resize(std::size_t count)
{
    for (std::vector<double>* v : {&x, &y, &z, &vx, &vy, &vz, &fx, &fy, &fz, &mass, &charge})
        v->resize(count);
    type.resize(count);
    global_id.resize(count);
    num_particles = count;
}

void ParticleInfo::set_counts(std::size_t owned, std::size_t ghosts) noexcept
{
    if (!state_) state_.reset(new (std::nothrow) State());
    if (state_) {
        state_->owned = owned;
        state_->ghosts = ghosts;
    }
}

void ParticleInfo::reset_counts(std::size_t owned, std::size_t ghosts)
{
    if (capacity() < owned + ghosts)
        throw std::invalid_argument("dsDecompX: " + std::to_string(owned + ghosts) +
                                    " particles do not fit storage of " + std::to_string(capacity()));
    State& s = state();
    s.owned = owned;
    s.ghosts = ghosts;
}

void ParticleInfo::resize_owned(std::size_t count)
{
    State& s = state();
    switch (s.backing) {
    case Backing::Borrowed:
        throw std::logic_error("dsDecompX: resize_owned cannot resize arrays the caller owns");
    case Backing::Block:
        ensure_capacity(count, std::min(count, s.owned));
        break;
    case Backing::HostVectors:
        s.require_field_capacity(count);
        resize(count);
        break;
    }
    s.owned = count;
    s.ghosts = 0;
}

ConstParticleView ParticleInfo::const_view(std::size_t need) const
{
    if (capacity() < need)
        throw std::invalid_argument("dsDecompX: particle storage of " + std::to_string(capacity()) +
                                    " slots is smaller than " + std::to_string(need));
    if (state_ && state_->backing != Backing::HostVectors) return as_const_view(state_->view());
    ConstParticleView v;
    v.memory_space = MemorySpace::Host;
    v.count = need;
    v.x = x.data(); v.y = y.data(); v.z = z.data();
    v.vx = vx.data(); v.vy = vy.data(); v.vz = vz.data();
    v.fx = fx.data(); v.fy = fy.data(); v.fz = fz.data();
    v.mass = mass.data(); v.charge = charge.data();
    v.type = type.data(); v.global_id = global_id.data();
    if (state_ && !state_->fields.empty()) {
        v.fields = state_->fields.data();
        v.field_count = state_->fields.size();
    }
    return v;
}

StorageResizeResult ParticleInfo::ensure_capacity(std::size_t required, std::size_t preserve)
{
    State& s = state();
    s.require_field_capacity(required);
    switch (s.backing) {
    case Backing::HostVectors: {
        const double* before = x.data();
        resize(required);
        ParticleView v;
        v.memory_space = MemorySpace::Host;
        v.count = required;
        v.x = x.data(); v.y = y.data(); v.z = z.data();
        v.vx = vx.data(); v.vy = vy.data(); v.vz = vz.data();
        v.fx = fx.data(); v.fy = fy.data(); v.fz = fz.data();
        v.mass = mass.data(); v.charge = charge.data();
        v.type = type.data(); v.global_id = global_id.data();
        v.fields = s.fields.empty() ? nullptr : s.fields.data();
        v.field_count = s.fields.size();
        return {v, before != x.data()};
    }
    case Backing::Block:
        if (required <= s.arrays.count) return {s.view(), false};
        s.grow_block(required, preserve);
        return {s.view(), true};
    case Backing::Borrowed:
        break;
    }
    if (required <= s.arrays.count) return {s.view(), false};
    if (!s.provider)
        throw std::length_error("dsDecompX: borrowed particle arrays of " + std::to_string(s.arrays.count) +
                                " slots cannot hold " + std::to_string(required) + " particles");
    const StorageResizeResult grown = s.provider->resize_particle_info(required, preserve);
    s.arrays = grown.view;   // the registered fields stay the ones registered here
    s.arrays.fields = nullptr;
    s.arrays.field_count = 0;
    return {s.view(), grown.reallocated};
}

ParticleInfo& ParticleInfo::register_field(ParticleFieldView field)
{
    State& s = state();
    if (field.name.empty() || !field.data || field.components == 0 || field.capacity < size())
        throw std::invalid_argument("dsDecompX: registered field '" + field.name +
                                    "' needs a name, data, components and room for every particle");
    for (const ParticleFieldView& f : s.fields)
        if (f.name == field.name)
            throw std::invalid_argument("dsDecompX: a field named '" + field.name + "' is already registered");
    s.fields.push_back(std::move(field));
    return *this;
}

void ParticleInfo::reset_view(ParticleView storage)
{
    if (storage.count < size())
        throw std::invalid_argument("dsDecompX: the new particle view has " + std::to_string(storage.count) +
                                    " slots for " + std::to_string(size()) + " particles");
    State& s = state();
    s.drop_block();
    s.backing = Backing::Borrowed;
    s.provider = nullptr;
    s.adopt(storage);
}

}  // namespace dsdecompx
