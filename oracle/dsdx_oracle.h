/*
 * dsdx_oracle.h -- CPU restatement of dsDecompX's Host communication path on a
 * 1x1x1 periodic (self-loopback) process grid.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (dsdecompx_amd) never
 * links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * plain IEEE double with FP contraction disabled (the Makefile passes
 * -ffp-contract=off), matching the reference Host build (g++ -O3 on x86-64,
 * no -march, hence no FMA).
 *
 * Parity pinning (see tests/test_oracle_*.py):
 *   dsdxo_neighbors_build   closed-form slot/shift table, opp(opp(m)) == m,
 *                           independent itertools enumeration
 *   dsdxo_halo_build        exact-tie known answers (x == lo+w, x == hi-w),
 *                           exact-rational geometric brute force of the ghost
 *                           set, hand-derived ghost order, the reference
 *                           examples' ghost counts (corner particle -> 7)
 *   dsdxo_halo_write_ghosts owner-copy invariants, +0.0 shift, zero forces
 *   dsdxo_reverse_exchange  order-sensitive known answers (catastrophic
 *                           cancellation), exact integer-valued conservation
 *                           against an independent bincount, Min/Max/Replace
 *                           known answers, NaN/-0.0 std::min/max semantics
 *   dsdxo_forward_exchange  exact copies, adjointness with reverse(Sum)
 *   dsdxo_migrate_wrap      x=-1e-17 -> L (no fold-back), exact-rational
 *                           congruence mod L, non-periodic axes untouched
 */
#ifndef DSDX_ORACLE_H
#define DSDX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSDXO_MAX_NEIGHBORS 26
#define DSDXO_MAX_FIELDS 16

/* Global box, halo width and neighbour policy (include/domain.hpp:11-30),
 * and the local domain whose faces the slab tests use: with local_given == 0
 * the cuts of a fresh 1x1x1 topology, lower + extent * 0 / 1 and
 * lower + extent * 1 / 1 (src/topology.cpp:66-72), else local_lower /
 * local_upper (a subdomain of a process grid, or cuts rescaled by
 * rescale_domain, src/decomp.cpp:117-124). */
typedef struct {
    double lower[3];
    double upper[3];
    int32_t periodic[3];
    double width;
    int32_t edges_and_corners;
    double local_lower[3];
    double local_upper[3];
    int32_t local_given;
} dsdxo_box;

/* Neighbour table of the 1x1x1 grid (include/topology.hpp:14-20). */
typedef struct {
    int32_t count;
    int32_t direction[DSDXO_MAX_NEIGHBORS][3];
    double shift[DSDXO_MAX_NEIGHBORS][3];
    int32_t slot[DSDXO_MAX_NEIGHBORS];
    int32_t active[DSDXO_MAX_NEIGHBORS]; /* rank >= 0 */
} dsdxo_neighbors;

/* Host halo plan (include/decomposition.hpp HaloPlanEntry): per neighbour
 * index n the owned indices it sends, and per neighbour index the receive
 * count.  recv_from[n] is the neighbour index whose send list fills receive
 * block n (opposite slot, src/backends/mpi_backend.cpp:366-375). */
typedef struct {
    int64_t send_count[DSDXO_MAX_NEIGHBORS];
    int32_t* source[DSDXO_MAX_NEIGHBORS];
    int64_t send_capacity[DSDXO_MAX_NEIGHBORS];
    int64_t recv_count[DSDXO_MAX_NEIGHBORS];
    int32_t recv_from[DSDXO_MAX_NEIGHBORS];
    int64_t ghost_total;
} dsdxo_plan;

/* SoA particle arrays; any of vx..type may be NULL (src/halo.cpp:44-53). */
typedef struct {
    double* x; double* y; double* z;
    double* vx; double* vy; double* vz;
    double* fx; double* fy; double* fz;
    double* mass; double* charge;
    int32_t* type; int32_t* global_id;
} dsdxo_particles;

/* FieldType / ReduceOp enumerators of include/field.hpp:22-36. */
enum { DSDXO_INT32 = 0, DSDXO_INT64, DSDXO_UINT32, DSDXO_UINT64,
       DSDXO_FLOAT32, DSDXO_FLOAT64 };
enum { DSDXO_SUM = 0, DSDXO_MIN, DSDXO_MAX, DSDXO_REPLACE };

/* One registered application field (include/field.hpp:74-84): AoS,
 * `components` elements of `type` per particle. */
typedef struct {
    void* data;
    int32_t type;
    int32_t components;
    int32_t op;
} dsdxo_field;

int dsdxo_direction_slot(int dx, int dy, int dz);
int dsdxo_opposite_slot(int dx, int dy, int dz);
void dsdxo_neighbors_build(const dsdxo_box* box, dsdxo_neighbors* out);

int dsdxo_halo_build(const dsdxo_box* box, const dsdxo_neighbors* nb,
                     int64_t owned, const double* x, const double* y,
                     const double* z, dsdxo_plan* plan);
void dsdxo_plan_free(dsdxo_plan* plan);

/* Owner index of every ghost, in ghost-array order (length ghost_total). */
void dsdxo_ghost_owners(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                        int32_t* owners);

int dsdxo_halo_write_ghosts(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                            int64_t owned, const dsdxo_particles* p,
                            const dsdxo_field* fields, int32_t field_count);

int dsdxo_reverse_exchange(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                           int64_t owned, const dsdxo_field* fields,
                           int32_t field_count);

int dsdxo_reverse_halo(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                       int64_t owned, double* fx, double* fy, double* fz);

int dsdxo_forward_exchange(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                           int64_t owned, const dsdxo_field* fields,
                           int32_t field_count);

void dsdxo_migrate_wrap(const dsdxo_box* box, int64_t owned, double* x,
                        double* y, double* z);

/* The local domain the slab tests use (src/topology.cpp:66-72): the given
 * one, or the fresh 1x1x1 cuts lower + extent * i / 1. */
void dsdxo_local_domain(const dsdxo_box* box, double lower[3], double upper[3]);

/* Decomposition::rescale_domain (src/decomp.cpp:117-124): every cut keeps its
 * fraction of the extent, fraction = (cut - old_lower) / old_extent,
 * cut = new_lower + fraction * new_extent, in place on cut_lo / cut_hi. */
void dsdxo_rescale_cuts(const double old_lower[3], const double old_upper[3],
                        const double new_lower[3], const double new_upper[3],
                        double cut_lo[3], double cut_hi[3]);

/* ---- Wire format (dsdx_oracle_wire.c) ------------------------------------
 * ParticleRecord (include/field.hpp:150-157), the particle and exchange
 * record layouts and the device path's pack / unpack / reduce steps on the
 * loopback plan.  Parity pinning (tests/test_oracle_wire.py):
 *   dsdxo_record / dsdxo_wire_layout  byte offsets and sizes of the record
 *                            members, stride known answers for every field
 *                            type mix (both layouts)
 *   dsdxo_halo_pack + dsdxo_loopback_exchange + dsdxo_unpack_records
 *                            round trip equal to dsdxo_halo_write_ghosts
 *                            (records and fields), every record decoded by
 *                            numpy against its owner + shift
 *   dsdxo_pack_reverse_records + dsdxo_accumulate_reverse
 *                            equal to dsdxo_reverse_halo bit for bit
 *   field records            pack / scatter / gather / reduce round trips
 *                            equal to dsdxo_forward_exchange and
 *                            dsdxo_reverse_exchange for every type and op */
typedef struct {
    double x, y, z, vx, vy, vz, fx, fy, fz, mass, charge;
    int32_t type, global_id;
} dsdxo_record;

enum { DSDXO_WIRE_PARTICLE = 0, DSDXO_WIRE_EXCHANGE = 1 };

void dsdxo_reduce_elements(int32_t type, int32_t op, void* dst, const void* src,
                           int32_t components);
int dsdxo_wire_layout(const dsdxo_field* fields, int32_t count, int32_t kind, int64_t* offsets,
                      int64_t* stride);
int64_t dsdxo_halo_pack(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                        const dsdxo_particles* p, const dsdxo_field* fields, int32_t field_count,
                        int32_t positions_only, unsigned char* out);
void dsdxo_unpack_records(const unsigned char* records, int64_t count, const dsdxo_particles* p,
                          int64_t dest, const dsdxo_field* fields, int32_t field_count);
void dsdxo_unpack_positions(const double* positions, int64_t count, const dsdxo_particles* p,
                            int64_t dest);
void dsdxo_pack_reverse_records(const double* fx, const double* fy, const double* fz,
                                int64_t ghost_offset, int64_t count, unsigned char* out);
void dsdxo_accumulate_reverse(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                              const unsigned char* records, double* fx, double* fy, double* fz);
void dsdxo_pack_field_records(const dsdxo_field* fields, int32_t field_count, int64_t first,
                              int64_t count, unsigned char* out);
void dsdxo_scatter_field_records(const dsdxo_field* fields, int32_t field_count,
                                 const unsigned char* records, int64_t count, int64_t first);
int64_t dsdxo_gather_field_records(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                                   const dsdxo_field* fields, int32_t field_count,
                                   unsigned char* out);
void dsdxo_reduce_field_records(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                                const dsdxo_field* fields, int32_t field_count,
                                const unsigned char* records);
void dsdxo_loopback_exchange(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                             const unsigned char* send, unsigned char* recv, int64_t record_bytes,
                             int32_t reverse);

/* ---- P-rank process grids (dsdx_oracle_grid.c) ---------------------------
 * The unmodified Host path on a dims[0] x dims[1] x dims[2] Cartesian grid,
 * every rank restated in one process.  Parity pinning (tests/test_oracle_grid.py):
 *   dsdxo_grid_build         closed-form cuts, row-major ranks, neighbour
 *                            ranks and shifts against an independent
 *                            enumeration; unsplit-axis skip; the reference
 *                            examples' neighbour counts
 *   dsdxo_grid_owner_rank    brute force over the subdomain boxes
 *   dsdxo_grid_migrate       conservation of the id multiset, every particle
 *                            inside its owner's subdomain, retained-first
 *                            order, the reference example particle_migration
 *   dsdxo_grid_halo          the reference examples' ghost counts and
 *                            returned forces (getting_started: 1/3/7 ghosts on
 *                            2/4/8 ranks), exact-rational geometric brute force
 *                            of every rank's ghost set, the 1x1x1 loopback
 *                            oracle when the unsplit-axis skip is lifted
 *   dsdxo_grid_reverse_exchange / dsdxo_grid_forward_exchange
 *                            exact conservation, adjointness, the loopback
 *                            oracle at 1x1x1 */
#define DSDXO_MAX_RANKS 64

typedef struct {
    dsdxo_box global;
    int32_t dims[3];
    int32_t ranks;
    double cuts[3][DSDXO_MAX_RANKS + 1];
    int32_t coord[DSDXO_MAX_RANKS][3];
    dsdxo_box local[DSDXO_MAX_RANKS];          /* subdomain, global width/periodicity */
    dsdxo_neighbors nb[DSDXO_MAX_RANKS];       /* neighbour list of every rank */
    int32_t nb_rank[DSDXO_MAX_RANKS][DSDXO_MAX_NEIGHBORS]; /* -1: outside a wall */
} dsdxo_grid;

int dsdxo_grid_build(const dsdxo_box* box, const int32_t dims[3], int32_t lift_unsplit,
                     dsdxo_grid* g);
int32_t dsdxo_grid_owner_rank(const dsdxo_grid* g, double x, double y, double z);
int dsdxo_grid_migrate(const dsdxo_grid* g, dsdxo_particles* p, const int64_t* capacity,
                       int64_t* owned, const dsdxo_field* fields, int32_t field_count,
                       int64_t* sent, int64_t* received);
int dsdxo_grid_halo(const dsdxo_grid* g, dsdxo_particles* p, const int64_t* capacity,
                    const int64_t* owned, dsdxo_plan* plans, int32_t* recv_rank,
                    int64_t* ghosts, const dsdxo_field* fields, int32_t field_count);
int dsdxo_grid_reverse_exchange(const dsdxo_grid* g, const dsdxo_plan* plans,
                                const int64_t* owned, const dsdxo_field* fields,
                                int32_t field_count);
int dsdxo_grid_forward_exchange(const dsdxo_grid* g, const dsdxo_plan* plans,
                                const int32_t* recv_rank, const int64_t* owned,
                                const dsdxo_field* fields, int32_t field_count);

#ifdef __cplusplus
}
#endif

#endif
