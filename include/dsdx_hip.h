/*
 * dsdx_hip.h -- C ABI of the MI355X (gfx950) particle-communication kernels.
 *
 * These entry points replace the reference's device operator layer
 * (src/cuda/kernels.hpp:40-176, namespace dsdecompx::cuda) on a 1x1x1
 * self-loopback grid: each of the 26 neighbour slots is this GPU, and the
 * ghost block of receive slot m is filled directly from the owners in the
 * send list of the opposite slot 25-m (the reference's self-neighbour rule,
 * src/backends/mpi_backend.cpp:366-375), so the direct entry points involve
 * no wire records and no transport.  The record entry points ("Communication
 * records") produce and consume the reference's wire format for callers that
 * keep a transport of their own, and dsdx_grid_* holds several subdomains.
 *
 * Conventions
 *   - Plain pointers and sizes only.  Particle arrays are device (or managed)
 *     SoA arrays of length >= capacity; ghosts are written after the owned
 *     particles, at [owned, owned + ghosts).
 *   - `stream` is a hipStream_t passed as void* (the reference's
 *     ExecutionStream, include/field.hpp:17).  All work is ordered on it.
 *   - Every function returns DSDX_OK or a negative DSDX_E* code;
 *     dsdx_last_error() returns the calling thread's message for the last
 *     failure.  Nothing throws across this boundary.
 *   - Results are bit-identical to the reference Host path
 *     (src/halo.cpp, src/migration.cpp): strict FP64 slab compares
 *     (x < lower+width, x > upper-width), the periodic wrap without fold-back,
 *     ghost order = receive slot, then owned order, and reverse sums added
 *     per owner in ascending send-slot order (no atomics).
 */
#ifndef DSDX_HIP_H
#define DSDX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSDX_MAX_NEIGHBORS 26
#define DSDX_MAX_FIELDS 16

enum {
    DSDX_OK = 0,
    DSDX_EINVAL = -1,    /* bad argument (std::invalid_argument) */
    DSDX_ECAPACITY = -2, /* particle storage too small for owned + ghosts */
    DSDX_EDEVICE = -3,   /* HIP runtime / launch failure (std::runtime_error) */
    DSDX_ESTATE = -4     /* no plan, or plan does not match (std::logic_error) */
};

/* FieldType and ReduceOp, same enumerator values as include/field.hpp:22-36. */
enum { DSDX_INT32 = 0, DSDX_INT64, DSDX_UINT32, DSDX_UINT64, DSDX_FLOAT32,
       DSDX_FLOAT64 };
enum { DSDX_SUM = 0, DSDX_MIN, DSDX_MAX, DSDX_REPLACE };

/* Device SoA particle storage (ParticleView, include/field.hpp:113-131). */
typedef struct dsdx_particles {
    double* x; double* y; double* z;
    double* vx; double* vy; double* vz;
    double* fx; double* fy; double* fz;
    double* mass; double* charge;
    int32_t* type; int32_t* global_id;
} dsdx_particles;

/* One registered application field (ParticleFieldView,
 * include/field.hpp:74-84): AoS, `components` elements of `type` per
 * particle.  `op` is only read by dsdx_reverse_exchange. */
typedef struct dsdx_field {
    void* data;
    int32_t type;
    int32_t components;
    int32_t op;
} dsdx_field;

/* Global domain and halo settings (GlobalDomain + HaloConfig,
 * include/domain.hpp:11-30).  The migrate wrap and the periodic shifts use
 * the global domain.  The slab tests use the local domain of the 1x1x1 grid
 * (Topology::local_domain, src/halo.cpp:200-205): with local_given == 0 it
 * is the fresh topology's cut planes lower + extent * 0 / 1 and
 * lower + extent * 1 / 1, extent = upper - lower (src/topology.cpp:66-72;
 * the upper cut can differ from `upper` by one ulp); with local_given != 0
 * it is local_lower / local_upper, e.g. the cuts that
 * Decomp::rescale_domain rescaled in place (src/decomp.cpp:117-124). */
typedef struct dsdx_box {
    double lower[3];
    double upper[3];
    int32_t periodic[3];
    double width;
    int32_t edges_and_corners;
    double local_lower[3];
    double local_upper[3];
    int32_t local_given;
} dsdx_box;

/* Device-resident halo plan: per-chunk slab classification, per-slot send
 * counts, the owner -> ghost map used by reverse/forward exchanges.
 * Replaces the device plan state of src/internal/decomp_impl.hpp:58-106 and
 * the source_indices buffer of src/cuda/execution.hpp:14-52. */
typedef struct dsdx_plan dsdx_plan;

const char* dsdx_last_error(void);

int dsdx_plan_create(dsdx_plan** plan);
int dsdx_plan_destroy(dsdx_plan* plan);

/* Drops the plan (the device halo state cleared by a migration,
 * src/migration.cpp:355-359): reverse and forward exchanges become no-ops. */
int dsdx_plan_reset(dsdx_plan* plan);

/* Pre-sizes the plan's buffers for builds of `owned` particles with at most
 * `ghosts` ghosts.  A following dsdx_halo_build / dsdx_migrate_halo_build of
 * that many particles with a null `ghosts` pointer then allocates no memory
 * and waits for nothing on the host (the box descriptor goes down as a kernel
 * argument), so it can be captured into a hipGraph together with
 * dsdx_migrate_wrap and dsdx_reverse_halo.  A build whose ghosts exceed the
 * reserved count writes none of them and records the overflow
 * (dsdx_plan_query); the reverse and forward exchanges then skip it. */
int dsdx_plan_reserve(dsdx_plan* plan, uint64_t owned, uint64_t ghosts);

/* Periodic wrap of the owned positions in place; replaces
 * launch_classify_destinations + launch_build_migration_layout + the unpack
 * (src/cuda/kernels.hpp:152-176) at P = 1, where every particle is retained
 * in order (src/migration.cpp:102-116, :437-479).  Non-periodic axes are left
 * unchanged.  One kernel launch with the box as its argument: no device
 * memory of its own, no host wait (capturable). */
int dsdx_migrate_wrap(const dsdx_box* box, double* x, double* y, double* z,
                      uint64_t owned, void* stream);

/* Halo rebuild: slab classification of the owned particles, per-slot stable
 * compaction and the ghost write, fused for the loopback grid.  Replaces
 * launch_halo_classify + launch_halo_scatter + launch_unpack_records
 * (src/cuda/kernels.hpp:78-96, :132-137) and the host wire path
 * (src/halo.cpp:164-238).  Ghosts receive the owner's fields with the
 * periodic shift added to x, y, z and zero forces; every registered field is
 * copied.  All 13 particle arrays must be non-null (src/halo.cpp:369-376)
 * except fx, fy, fz, which are zeroed only when present.
 *
 * If `ghosts` is non-null the call waits for the ghost count and stores it
 * there (the reference's count readback, src/halo.cpp:482-491); when
 * owned + ghosts > capacity it returns DSDX_ECAPACITY without writing any
 * ghost, and dsdx_halo_scatter may be called with larger storage.  If
 * `ghosts` is null nothing waits; an overflow is recorded in the plan and
 * reported by dsdx_plan_query.  A build captured into a graph must pass a
 * null `ghosts`: waiting for the counts of a captured build (here, or in
 * dsdx_plan_query with a stream) returns DSDX_ESTATE.
 *
 * A box whose particle records exceed the device's 256 MB memory-side cache
 * is built without writing the owner map and ghost list: the ghosts and
 * dsdx_reverse_halo need neither.  The first later call that walks them
 * (dsdx_halo_refresh, the field exchanges, dsdx_plan_ghost_owners, the record
 * calls) writes them first, with one more kernel on its stream.  A build
 * captured into a graph writes them in the graph, so that every replay
 * leaves the maps of the build it replayed. */
int dsdx_halo_build(dsdx_plan* plan, const dsdx_box* box,
                    const dsdx_particles* particles, const dsdx_field* fields,
                    uint32_t field_count, uint64_t owned, uint64_t capacity,
                    void* stream, uint64_t* ghosts);

/* migrate (dsdx_migrate_wrap) followed by dsdx_halo_build in one pass over
 * the positions: the classification pass wraps the periodic coordinates it
 * loaded, stores the changed ones and classifies the wrapped values.  Same
 * results as the two calls; reads x, y, z once instead of twice.  For
 * callers that rebuild right after migrating (the reference's
 * communication_test step, examples/communication_test/main.cu:222-236). */
int dsdx_migrate_halo_build(dsdx_plan* plan, const dsdx_box* box,
                            const dsdx_particles* particles,
                            const dsdx_field* fields, uint32_t field_count,
                            uint64_t owned, uint64_t capacity, void* stream,
                            uint64_t* ghosts);

/* dsdx_halo_build without a plan and without allocating: no counterpart in
 * the reference.  For callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same kernels write the same ghosts, bit for bit; their plan state (per
 * chunk entries and counts, the owner map, the ghost list) lives in one fixed
 * device array of the library, about 13 MB, which each GPU the library runs
 * on holds until the process ends.  Limits: at most DSDX_NOALLOC_MAX_OWNED
 * owned particles (DSDX_EINVAL beyond) and DSDX_NOALLOC_MAX_GHOSTS ghosts
 * (DSDX_ECAPACITY beyond, nothing written).  Every call on a GPU shares that
 * array and keeps the last call's plan in it, so calls that overlap on one
 * GPU race: order them on one stream.  The stream must belong to the
 * current device.  `ghost_owners`, if non-null, is a device array of at
 * least capacity - owned int32 that receives the owned index of every ghost
 * in ghost order (as dsdx_plan_ghost_owners), enough for a caller's own
 * reverse sum.  `ghosts` as for dsdx_halo_build: non-null waits for the count
 * (and reports an overflow as DSDX_ECAPACITY), null waits for nothing. */
#define DSDX_NOALLOC_MAX_OWNED 262144
#define DSDX_NOALLOC_MAX_GHOSTS 524288
int dsdx_halo_build_noalloc(const dsdx_box* box, const dsdx_particles* particles,
                            const dsdx_field* fields, uint32_t field_count,
                            uint64_t owned, uint64_t capacity,
                            int32_t* ghost_owners, void* stream, uint64_t* ghosts);

/* Writes the ghosts of the plan's last classification (after a
 * DSDX_ECAPACITY, with storage of at least owned + ghosts). */
int dsdx_halo_scatter(dsdx_plan* plan, const dsdx_particles* particles,
                      const dsdx_field* fields, uint32_t field_count,
                      uint64_t owned, uint64_t capacity, void* stream);

/* Reused-plan refresh: every ghost re-reads its owner along the existing
 * plan (launch_halo_gather / launch_halo_gather_positions +
 * launch_unpack_positions, src/cuda/kernels.hpp:103-128).  With
 * positions_only only x, y, z are rewritten and forces zeroed; otherwise the
 * whole record and the registered fields are rewritten.  A non-null `box`
 * replaces the periodic shifts of the plan (Decomp::rescale_domain keeps the
 * ghost list but moves the images, src/decomp.cpp:109-144); null keeps the
 * shifts of the build. */
int dsdx_halo_refresh(dsdx_plan* plan, const dsdx_box* box,
                      const dsdx_particles* particles,
                      const dsdx_field* fields, uint32_t field_count,
                      uint64_t owned, int32_t positions_only, void* stream);

/* Waits for the plan's counts: per-slot send counts (send_counts[s] for
 * slot s), receive counts (receive slot m receives send_counts[25-m]), total
 * ghosts and distinct owners, and the overflow flag of an unwaited build.
 * Any output pointer may be null.  With a null stream nothing waits: the
 * counts read back by the last waited dsdx_halo_build are returned. */
int dsdx_plan_query(const dsdx_plan* plan, uint32_t* send_counts,
                    uint32_t* receive_counts, uint64_t* ghosts,
                    uint64_t* owners, int32_t* overflow, void* stream);

/* Owner (owned index) of every ghost, in ghost order: the reference's
 * halo_source_indices permuted into receive order (src/cuda/pack.cu:172-177).
 * `owners` is a device array of at least `ghosts` int32. */
int dsdx_plan_ghost_owners(const dsdx_plan* plan, int32_t* owners,
                           void* stream);

/* reverse_halo: fx, fy, fz of every ghost added to its owner, per owner in
 * ascending send-slot order, deterministically (Host order of
 * src/halo.cpp:1351-1372 via :1336-1346).  Replaces
 * launch_pack_reverse_records + launch_accumulate_reverse_records
 * (src/cuda/kernels.hpp:139-150), which reach the same order through one
 * atomic-add launch per neighbour block (src/halo.cpp:1069-1079); here one
 * pass, one lane per owner, no atomics and no records. */
int dsdx_reverse_halo(const dsdx_plan* plan, double* fx, double* fy,
                      double* fz, uint64_t owned, void* stream);

/* reverse_exchange of registered fields with Sum / Min / Max / Replace, in
 * the same deterministic order (src/halo.cpp:1396-1532).  Replaces
 * launch_pack_reverse_fields + launch_reduce_reverse_fields
 * (src/cuda/kernels.hpp:55-64). */
int dsdx_reverse_exchange(const dsdx_plan* plan, const dsdx_field* fields,
                          uint32_t field_count, uint64_t owned, void* stream);

/* forward_exchange: owner values copied into every ghost slot that holds a
 * copy of the owner.  Replaces launch_gather_fields + launch_scatter_fields
 * (src/cuda/kernels.hpp:69-72, src/halo.cpp:1084-1207). */
int dsdx_forward_exchange(const dsdx_plan* plan, const dsdx_field* fields,
                          uint32_t field_count, uint64_t owned, void* stream);

/* Launch shape of the plan's kernels: lane items per chunk (4, 8, 16; a
 * chunk is 64 x items particles, one wave each), waves per block (2, 4, 8)
 * and scan block threads (256, 512, 1024).  Defaults: items by the owned
 * count of each build (16 from 4 M particles, 8 from 2 M, else 4) / 4 /
 * 1024, or the environment variables DSDX_ITEMS, DSDX_WAVES,
 * DSDX_SCAN_THREADS; the autotuner (dsdecompx_amd.autotune) picks them by
 * measurement.  Setting a shape fixes the items; changing items drops the
 * plan. */
int dsdx_plan_set_shape(dsdx_plan* plan, int32_t items, int32_t waves,
                        int32_t scan_threads);
int dsdx_plan_get_shape(const dsdx_plan* plan, int32_t* items, int32_t* waves,
                        int32_t* scan_threads);

/* ---- Batched form: many independent boxes ("replicas") in one set of SoA
 * arrays, each phase one launch over all of them, nothing waiting on the
 * host, so a whole step can be captured in a hipGraph.  Replica r owns the
 * particle slots [offset, offset + capacity) in every particle array (and in
 * every registered field): its owned particles first, its ghosts after them.
 * Each replica's result is bit-identical to the single-box entry points on
 * its own slots.  No counterpart in the reference (which has one box per
 * rank); for ensembles of replica systems. */
typedef struct dsdx_replica {
    dsdx_box box;
    uint64_t offset;
    uint64_t capacity;
    uint64_t owned;
} dsdx_replica;

typedef struct dsdx_batch dsdx_batch;

/* Copies the replica table to the device and sizes every buffer (the ghost
 * list to the sum of capacity - owned), so later calls allocate nothing.
 * Regions must be ascending and disjoint. */
int dsdx_batch_create(const dsdx_replica* replicas, uint32_t count,
                      void* stream, dsdx_batch** batch);
int dsdx_batch_destroy(dsdx_batch* batch);
int dsdx_batch_set_shape(dsdx_batch* batch, int32_t items, int32_t waves,
                         int32_t scan_threads, void* stream);
int dsdx_batch_get_shape(const dsdx_batch* batch, int32_t* items,
                         int32_t* waves, int32_t* scan_threads);
int dsdx_batch_migrate_wrap(const dsdx_batch* batch, double* x, double* y,
                            double* z, void* stream);
/* A replica whose ghosts do not fit its capacity writes none of them and
 * sets its status word (read by dsdx_batch_counts); the owner-major calls
 * then skip it. */
int dsdx_batch_halo_build(dsdx_batch* batch, const dsdx_particles* particles,
                          const dsdx_field* fields, uint32_t field_count,
                          void* stream);
/* dsdx_batch_migrate_wrap + dsdx_batch_halo_build in one pass over x, y, z. */
int dsdx_batch_migrate_halo_build(dsdx_batch* batch,
                                  const dsdx_particles* particles,
                                  const dsdx_field* fields,
                                  uint32_t field_count, void* stream);
int dsdx_batch_refresh(dsdx_batch* batch, const dsdx_particles* particles,
                       const dsdx_field* fields, uint32_t field_count,
                       int32_t positions_only, void* stream);
int dsdx_batch_reverse_halo(const dsdx_batch* batch, double* fx, double* fy,
                            double* fz, void* stream);
int dsdx_batch_reverse_exchange(const dsdx_batch* batch,
                                const dsdx_field* fields, uint32_t field_count,
                                void* stream);
int dsdx_batch_forward_exchange(const dsdx_batch* batch,
                                const dsdx_field* fields, uint32_t field_count,
                                void* stream);
/* owners[offset_r + g] = owned index (replica-local) of ghost g of replica r */
int dsdx_batch_ghost_owners(const dsdx_batch* batch, int32_t* owners,
                            void* stream);
/* Waits; per replica: ghosts, distinct owners, status (0 or overflow bits). */
int dsdx_batch_counts(const dsdx_batch* batch, uint32_t* ghosts,
                      uint32_t* owners, uint32_t* status, void* stream);
/* Device pointer to the per-replica counters, 32 u32 words per replica:
 * [0..25] send count per slot, [26] owners, [27] ghosts, [28] status.  For
 * kernels that must know a replica's ghost count without a host wait. */
const uint32_t* dsdx_batch_device_totals(const dsdx_batch* batch);

/* A step captured into a HIP graph and replayed: dsdx_graph_begin starts a
 * thread-local capture on `stream` (not the null stream); the calls that
 * follow on that stream (dsdx_batch_*, and the application's own kernels)
 * are recorded, not run; dsdx_graph_end ends the capture and instantiates
 * the graph; dsdx_graph_launch replays it on any stream.  A call that
 * allocates or waits on the host inside the capture ends it with an error
 * (the batch calls do neither).  No counterpart in the reference (its steps
 * are host-driven); for callers that replay a fixed step. */
typedef struct dsdx_graph dsdx_graph;
int dsdx_graph_begin(void* stream);
int dsdx_graph_end(void* stream, dsdx_graph** graph);
int dsdx_graph_launch(const dsdx_graph* graph, void* stream);
int dsdx_graph_destroy(dsdx_graph* graph);

/* ---- Communication records ------------------------------------------------
 * The reference's wire formats, for callers that carry records over a
 * transport of their own (MPI, RCCL) between the pack and unpack steps:
 *   particle record  ParticleRecord (include/field.hpp:150-157): x, y, z, vx,
 *                    vy, vz, fx, fy, fz, mass, charge (double), type,
 *                    global_id (int32) -- DSDX_RECORD_BYTES -- then the
 *                    registered fields back to back; stride rounded up to 8
 *                    (make_device_wire_layout, src/device_fields.cpp:16-53)
 *   exchange record  the fields of a forward / reverse exchange from byte 0,
 *                    each aligned to its element size; stride rounded up to 8
 *                    (src/halo.cpp:1237-1271, :1422-1457)
 *   position record  x, y, z (24 bytes)
 * A send buffer holds one block per send slot in slot order (the reference's
 * neighbour order), each in owned order, block s starting at the sum of the
 * send counts of the slots before it (dsdx_plan_query); a receive buffer
 * holds the receive blocks in slot order, i.e. the ghosts in ghost order.
 * The plan-aware entry points need a single-box plan with a built halo. */
#define DSDX_RECORD_BYTES 96
enum { DSDX_WIRE_PARTICLE = 0, DSDX_WIRE_EXCHANGE = 1 };

/* launch_halo_classify + count readback + launch_halo_scatter
 * (src/halo.cpp:464-531, src/cuda/kernels.hpp:78-101): the halo of `owned`
 * particles classified, routed and packed as particle records into `send`
 * (send-buffer layout); no ghost is written to the particle arrays, and the
 * records' positions come from the classification pass.  Waits for the
 * counts (*records = number of records).  With send_capacity (in records)
 * below that: DSDX_ECAPACITY, the plan built, and dsdx_halo_pack fills a
 * larger buffer. */
int dsdx_halo_build_records(dsdx_plan* plan, const dsdx_box* box,
                            const dsdx_particles* particles, const dsdx_field* fields,
                            uint32_t field_count, uint64_t owned, void* send,
                            uint64_t send_capacity, void* stream, uint64_t* records);
/* Byte offset of every field inside a record of `kind`, and the stride. */
int dsdx_wire_layout(const dsdx_field* fields, uint32_t count, int32_t kind,
                     uint32_t* offsets, uint32_t* stride);
/* Particle records of every ghost of the plan into its send block: owner +
 * periodic shift, zero forces, the fields (positions_only: position records,
 * no fields).  After a build: launch_halo_scatter's records; on a reused
 * plan: launch_halo_gather / launch_halo_gather_positions
 * (src/cuda/kernels.hpp:89-121). */
int dsdx_halo_pack(const dsdx_plan* plan, const dsdx_particles* particles,
                   const dsdx_field* fields, uint32_t field_count, uint64_t owned,
                   int32_t positions_only, void* send, void* stream);
/* launch_unpack_records / launch_unpack_positions (src/cuda/kernels.hpp:
 * 124-137): record k into particle dest_offset + k (null arrays skipped);
 * positions also zero the forces. */
int dsdx_unpack_records(const void* records, uint64_t count,
                        const dsdx_particles* particles, uint64_t dest_offset,
                        const dsdx_field* fields, uint32_t field_count, void* stream);
int dsdx_unpack_positions(const double* positions, uint64_t count,
                          const dsdx_particles* particles, uint64_t dest_offset,
                          void* stream);
/* launch_pack_reverse_records (src/cuda/kernels.hpp:139-143): particle
 * records of ghosts [ghost_offset, ghost_offset + count) carrying only their
 * forces. */
int dsdx_pack_reverse_records(const double* fx, const double* fy, const double* fz,
                              uint64_t ghost_offset, uint64_t count, void* out,
                              void* stream);
/* The returned force records (send-buffer layout) added to their owners in
 * ascending send slot: the reference's launch_accumulate_reverse_records once
 * (src/cuda/kernels.hpp:146-150) once per neighbour block
 * (src/halo.cpp:1069-1079), as one ordered pass. */
int dsdx_accumulate_reverse_records(const dsdx_plan* plan, const void* records,
                                    double* fx, double* fy, double* fz, uint64_t owned,
                                    void* stream);
/* launch_pack_reverse_fields / launch_scatter_fields (src/cuda/kernels.hpp:
 * 56-60, :71-72): exchange records of particles
 * [ghost_offset, ghost_offset + count). */
int dsdx_pack_field_records(const dsdx_field* fields, uint32_t field_count,
                            uint64_t ghost_offset, uint64_t count, void* out, void* stream);
int dsdx_scatter_field_records(const dsdx_field* fields, uint32_t field_count,
                               const void* records, uint64_t count, uint64_t ghost_offset,
                               void* stream);
/* launch_gather_fields over every send block (exchange records of the
 * owners, send-buffer layout), and launch_reduce_reverse_fields over the
 * returned blocks with each field's op, in ascending send slot
 * (src/cuda/kernels.hpp:62-70). */
int dsdx_gather_field_records(const dsdx_plan* plan, const dsdx_field* fields,
                              uint32_t field_count, uint64_t owned, void* out, void* stream);
int dsdx_reduce_field_records(const dsdx_plan* plan, const dsdx_field* fields,
                              uint32_t field_count, const void* records, uint64_t owned,
                              void* stream);
/* The self messages of the loopback box (src/backends/mpi_backend.cpp:
 * 366-375): forward, send blocks -> receive blocks (block m takes slot
 * 25 - m); reverse (non-zero `reverse`), records in ghost order -> back to
 * the send slots that produced them.  record_bytes: a multiple of 8.  Waits
 * for the plan's counts if they are not on the host yet. */
int dsdx_loopback_exchange(dsdx_plan* plan, const void* send, void* recv,
                           uint32_t record_bytes, int32_t reverse, void* stream);

/* ---- Process grids -------------------------------------------------------
 * The dims[0] x dims[1] x dims[2] Cartesian decomposition of one box, every
 * subdomain ("rank", row-major as MPI_Cart_rank without reordering) held on
 * this GPU, with the semantics of the reference's Host path on that many MPI
 * ranks: cuts lower + extent * i / dims, neighbour lists with the
 * unsplit-axis skip (src/topology.cpp:57-176), halo slab tests against each
 * subdomain (src/halo.cpp:164-238), receive blocks filled by the neighbour in
 * each direction (src/backends/mpi_backend.cpp:344-433), reverse sums in
 * ascending send slot (src/halo.cpp:1288-1532) and migration to owner_rank,
 * retained particles first, then those of ranks 0..P-1
 * (src/migration.cpp:74-128, :391-591).  No reference counterpart as one
 * object: it stands for P processes of the reference.
 *
 * Storage is the caller's: every particle array holds subdomain r in slots
 * [r * capacity_per_rank, (r + 1) * capacity_per_rank), owned first, then its
 * ghosts; a registered field's array (dsdx_field, AoS of `components`
 * values per slot) has the same slots.  All 13 particle arrays are required.
 * At most 64 subdomains.  A grid's kernels take chunks of 8 items per lane
 * (DSDX_ITEMS and the other shape variables apply as for a plan). */
typedef struct dsdx_grid dsdx_grid;

int dsdx_grid_create(const dsdx_box* box, const int32_t* dims, uint64_t capacity_per_rank,
                     void* stream, dsdx_grid** grid);
int dsdx_grid_destroy(dsdx_grid* grid);
/* Owned counts of every subdomain (after the caller filled the arrays). */
int dsdx_grid_set_owned(dsdx_grid* grid, const uint64_t* owned, void* stream);
/* Host copies: owned counts, and the ghost counts and distinct boundary
 * owners (particles with at least one image) of the last halo build.  Any
 * output may be null. */
int dsdx_grid_counts(const dsdx_grid* grid, uint64_t* owned, uint64_t* ghosts, uint64_t* owners);
/* Migrates from `from` into the separate storage `to` (same layout), each
 * registered field from fields_from[f] into fields_to[f] with its particle
 * (Decomp::migrate with the fields of add_field, src/migration.cpp:391-591);
 * waits for the new counts.  DSDX_ECAPACITY when a subdomain would overflow
 * (nothing is written then). */
int dsdx_grid_migrate(dsdx_grid* grid, const dsdx_particles* from, const dsdx_particles* to,
                      const dsdx_field* fields_from, const dsdx_field* fields_to,
                      uint32_t field_count, void* stream);
/* Rebuilds every subdomain's halo, the registered fields copied into the
 * ghosts with their records; waits for the ghost counts.  DSDX_ECAPACITY
 * when a subdomain cannot hold its ghosts (no ghost is written then). */
int dsdx_grid_halo_build(dsdx_grid* grid, const dsdx_particles* particles,
                         const dsdx_field* fields, uint32_t field_count, void* stream,
                         uint64_t* ghosts);
/* Ghost forces of every subdomain added to their owners. */
int dsdx_grid_reverse_halo(dsdx_grid* grid, double* fx, double* fy, double* fz, void* stream);
/* reverse_exchange / forward_exchange of registered fields on every
 * subdomain (src/halo.cpp:1396-1532, :1084-1207), as dsdx_reverse_exchange
 * and dsdx_forward_exchange. */
int dsdx_grid_reverse_exchange(dsdx_grid* grid, const dsdx_field* fields, uint32_t field_count,
                               void* stream);
int dsdx_grid_forward_exchange(dsdx_grid* grid, const dsdx_field* fields, uint32_t field_count,
                               void* stream);
/* Reused-plan refresh of every subdomain's ghosts, as dsdx_halo_refresh with
 * the shifts of the build. */
int dsdx_grid_refresh(dsdx_grid* grid, const dsdx_particles* particles, const dsdx_field* fields,
                      uint32_t field_count, int32_t positions_only, void* stream);

/* Per-kernel device timing, process-wide: when enabled, every kernel
 * dispatch carries its own start / stop HIP events (a timed dispatch starts a
 * few microseconds later than an untimed one).  dsdx_timing_read waits for
 * the recorded events and returns the accumulated milliseconds and launch
 * count of one kernel: "migrate_wrap", "halo_classify", "halo_scan",
 * "halo_scatter" (the route pass: owner map and ghost routes), "halo_copy"
 * (the ghost records), "reverse_halo", "reverse_exchange",
 * "forward_exchange", "halo_refresh", "ghost_owners", "migrate_classify"
 * (the fused pass), "grid_migrate_count", "grid_migrate_scatter",
 * "grid_exchange_counts", "halo_pack", "unpack_records", "pack_reverse",
 * "accumulate_reverse" or "loopback_exchange".  The GPU-event counterpart of the reference's host-clock
 * timings (include/decomposition.hpp:30-38, src/halo.cpp:316-323). */
int dsdx_timing_enable(int32_t enable);
int dsdx_timing_reset(void);
int dsdx_timing_read(const char* kernel, double* total_ms, uint64_t* launches);

#ifdef __cplusplus
}
#endif

#endif
