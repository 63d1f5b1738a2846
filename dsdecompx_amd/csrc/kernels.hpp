// kernels.hpp -- launchers of the gfx950 halo kernels (halo_kernels.hip).
#pragma once

#include <algorithm>

#include "device.hpp"

namespace dsdx {

hipError_t launch_classify(const double* x, const double* y, const double* z, uint32_t owned,
                           const BoxParams& bp, uint32_t* entries, uint32_t* counts,
                           uint32_t chunks, hipStream_t stream);
hipError_t launch_scan(const uint32_t* counts, uint32_t* prefix, uint32_t* totals,
                       uint32_t chunks, hipStream_t stream);
hipError_t launch_scatter(const dsdx_particles& p, uint32_t owned, const BoxParams& bp,
                          const FieldParams& fp, const uint32_t* entries, const uint32_t* counts,
                          const uint32_t* prefix, uint32_t* totals, uint32_t* owner_index,
                          uint8_t* owner_code, uint32_t* image_start, uint32_t* ghost_list,
                          uint32_t chunks, uint64_t capacity, uint64_t ghost_list_capacity,
                          hipStream_t stream);
hipError_t launch_reverse_forces(double* fx, double* fy, double* fz, uint32_t owned,
                                 const PlanView& v, const BoxParams& bp, hipStream_t stream);
hipError_t launch_reverse_fields(const TypedFields& tf, uint32_t owned, const PlanView& v,
                                 const BoxParams& bp, hipStream_t stream);
hipError_t launch_forward_fields(const FieldParams& fp, uint32_t owned, const PlanView& v,
                                 const BoxParams& bp, hipStream_t stream);
hipError_t launch_refresh(const dsdx_particles& p, const FieldParams& fp, uint32_t owned,
                          int32_t positions_only, const PlanView& v, const BoxParams& bp,
                          hipStream_t stream);
hipError_t launch_ghost_owners(int32_t* out, uint32_t owned, const PlanView& v,
                               const BoxParams& bp, hipStream_t stream);
hipError_t launch_migrate_wrap(double* x, double* y, double* z, uint64_t owned,
                               const BoxParams& bp, hipStream_t stream);

}  // namespace dsdx
