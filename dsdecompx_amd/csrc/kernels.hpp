// kernels.hpp -- launchers of the gfx950 halo kernels (halo_kernels.hip).
#pragma once

#include <algorithm>

#include "device.hpp"

namespace dsdx {

bool shape_supported(const Shape& s);

// wrap_first: apply migrate's periodic wrap in the same pass (fused step).
hipError_t launch_classify(double* x, double* y, double* z, bool wrap_first, const Layout& L,
                           const PlanBuffers& B, const Shape& s, hipStream_t stream);
hipError_t launch_scan(const Layout& L, const PlanBuffers& B, const Shape& s, hipStream_t stream);
hipError_t launch_route(const Layout& L, const PlanBuffers& B, const Shape& s, hipStream_t stream);
// The build's copy over chunk units leaves the owner maps and the ghost list
// unwritten; reverse_halo then needs no route pass when it takes the same
// units.
inline bool chunk_reverse(const Layout& L, const Shape& s) { return L.reverse_unit_chunks != 0 && s.reverse_runs != 0; }
hipError_t launch_copy(const dsdx_particles& p, const FieldParams& fp, const Layout& L,
                       const PlanBuffers& B, const Shape& s, hipStream_t stream);
hipError_t launch_reverse_forces(double* fx, double* fy, double* fz, const Layout& L,
                                 const PlanBuffers& B, const Shape& s, hipStream_t stream);
hipError_t launch_reverse_fields(const TypedFields& tf, const Layout& L, const PlanBuffers& B,
                                 const Shape& s, hipStream_t stream);
hipError_t launch_forward_fields(const FieldParams& fp, const Layout& L, const PlanBuffers& B,
                                 const Shape& s, hipStream_t stream);
hipError_t launch_refresh(const dsdx_particles& p, const FieldParams& fp, int32_t positions_only,
                          const Layout& L, const PlanBuffers& B, const Shape& s, hipStream_t stream);
hipError_t launch_ghost_owners(int32_t* out, const Layout& L, const PlanBuffers& B, const Shape& s,
                               hipStream_t stream);
hipError_t launch_migrate_wrap(double* x, double* y, double* z, const Layout& L, const Shape& s,
                               hipStream_t stream);
// migrate's wrap of one box given by value (no replica table).
hipError_t launch_wrap_box(double* x, double* y, double* z, const WrapBox& b, const Shape& s,
                           hipStream_t stream);
// One replica descriptor stored into device memory by a kernel (stream-ordered, no host wait).
hipError_t launch_store_desc(const ReplicaDesc& d, ReplicaDesc* out, hipStream_t stream);

// Process grids (grid_kernels.hip).
hipError_t launch_exchange_counts(const Layout& L, const PlanBuffers& B, uint32_t* host,
                                  hipStream_t stream);
hipError_t launch_migrate_count(const dsdx_particles& p, const Layout& L, const MigrateParams& M,
                                const Shape& s, hipStream_t stream);
hipError_t launch_migrate_scan(const Layout& L, const MigrateParams& M, const Shape& s,
                               hipStream_t stream);
hipError_t launch_migrate_bases(const MigrateParams& M, hipStream_t stream);
hipError_t launch_migrate_scatter(const dsdx_particles& from, const dsdx_particles& to,
                                  const FieldParams& fields_from, const FieldParams& fields_to,
                                  const Layout& L, const MigrateParams& M, const Shape& s,
                                  hipStream_t stream);

// Communication records (wire_kernels.hip).  The plan-aware launchers need a
// single-box plan; their buffers hold one block per send slot, slot order.
hipError_t launch_pack_halo(const dsdx_particles& p, const RecordFields& rf, int32_t positions_only,
                            int32_t fresh, unsigned char* out, const Layout& L, const PlanBuffers& B,
                            const Shape& s, hipStream_t stream);
hipError_t launch_unpack_records(const unsigned char* in, uint64_t count, const dsdx_particles& p,
                                 uint64_t dest, const RecordFields& rf, hipStream_t stream);
hipError_t launch_unpack_positions(const double* in, uint64_t count, const dsdx_particles& p, uint64_t dest,
                                   hipStream_t stream);
hipError_t launch_pack_reverse(const double* fx, const double* fy, const double* fz, uint64_t first,
                               uint64_t count, unsigned char* out, hipStream_t stream);
hipError_t launch_accumulate_reverse(const unsigned char* in, double* fx, double* fy, double* fz,
                                     const Layout& L, const PlanBuffers& B, const Shape& s,
                                     hipStream_t stream);
hipError_t launch_pack_field_records(const RecordFields& rf, uint64_t first, uint64_t count,
                                     unsigned char* out, hipStream_t stream);
hipError_t launch_scatter_field_records(const RecordFields& rf, const unsigned char* in, uint64_t count,
                                        uint64_t first, hipStream_t stream);
hipError_t launch_gather_field_records(const RecordFields& rf, unsigned char* out, const Layout& L,
                                       const PlanBuffers& B, const Shape& s, hipStream_t stream);
hipError_t launch_reduce_field_records(const RecordFields& rf, const unsigned char* in, const Layout& L,
                                       const PlanBuffers& B, const Shape& s, hipStream_t stream);
hipError_t launch_block_copy(const void* src, void* dst, const BlockMap& map, hipStream_t stream);

}  // namespace dsdx
