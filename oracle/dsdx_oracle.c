/*
 * dsdx_oracle.c -- CPU restatement of dsDecompX's Host communication path on a
 * 1x1x1 periodic self-loopback grid.  TEST INFRASTRUCTURE ONLY (see the
 * header): the product library never links or loads this file.
 *
 * Loopback semantics: the reference's own self-neighbour rule (a neighbour
 * slot m whose peer is this rank receives what this rank sent to the opposite
 * slot) applied to the 26 self-neighbours of a 1x1x1 grid, i.e. the Host
 * path with the single-rank short-circuits (src/halo.cpp:328-350) and the
 * unsplit-axis skip (src/topology.cpp:118-127) lifted.
 */
#include "dsdx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* src/topology.cpp:11-14 */
int dsdxo_direction_slot(int dx, int dy, int dz)
{
    const int raw = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
    return raw < 13 ? raw : raw - 1;
}

/* src/halo.cpp:30-34, src/backends/mpi_backend.cpp:29-35 */
int dsdxo_opposite_slot(int dx, int dy, int dz)
{
    return dsdxo_direction_slot(-dx, -dy, -dz);
}

/* src/topology.cpp:57-176 for dimensions {1,1,1}, coordinates {0,0,0}, with
 * the unsplit-axis skip (:118-127) lifted so that every direction is a
 * self-neighbour.  A step along a non-periodic axis leaves the box, so that
 * neighbour has rank -1 (:140-143, :160-162). */
void dsdxo_neighbors_build(const dsdxo_box* box, dsdxo_neighbors* out)
{
    memset(out, 0, sizeof(*out));
    int n = 0;
    for (int dx = -1; dx <= 1; ++dx) {
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dz = -1; dz <= 1; ++dz) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                if (!box->edges_and_corners &&
                    abs(dx) + abs(dy) + abs(dz) != 1)
                    continue;
                const int d[3] = {dx, dy, dz};
                double shift[3] = {0.0, 0.0, 0.0};
                int outside = 0;
                for (int a = 0; a < 3; ++a) {
                    /* neighbour coordinate 0 + d[a] is outside [0, 1) iff d != 0 */
                    if (d[a] != 0) {
                        if (!box->periodic[a]) {
                            outside = 1;
                            break;
                        }
                        const double extent = box->upper[a] - box->lower[a];
                        shift[a] = d[a] < 0 ? extent : -extent;
                    }
                }
                for (int a = 0; a < 3; ++a) {
                    out->direction[n][a] = d[a];
                    out->shift[n][a] = shift[a];
                }
                out->slot[n] = dsdxo_direction_slot(dx, dy, dz);
                out->active[n] = outside ? 0 : 1;
                ++n;
            }
        }
    }
    out->count = n;
}

static int neighbor_with_slot(const dsdxo_neighbors* nb, int slot)
{
    for (int k = 0; k < nb->count; ++k)
        if (nb->active[k] && nb->slot[k] == slot) return k;
    return -1;
}

static int push_index(dsdxo_plan* plan, int n, int32_t value)
{
    if (plan->send_count[n] == plan->send_capacity[n]) {
        const int64_t grown = plan->send_capacity[n] ? plan->send_capacity[n] * 2 : 64;
        int32_t* p = (int32_t*)realloc(plan->source[n], (size_t)grown * sizeof(int32_t));
        if (!p) return -1;
        plan->source[n] = p;
        plan->send_capacity[n] = grown;
    }
    plan->source[n][plan->send_count[n]++] = value;
    return 0;
}

/* src/halo.cpp:164-238 (rebuild_halo_plan): six slab tests per owned
 * particle, strict comparisons against the local domain's lower+width and
 * upper-width, then every active neighbour whose direction is covered.  Send
 * lists keep owned order.  Receive counts: src/backends/mpi_backend.cpp:
 * 366-375 (receive block n is the send list of the neighbour in the opposite
 * slot). */
int dsdxo_halo_build(const dsdxo_box* box, const dsdxo_neighbors* nb,
                     int64_t owned, const double* x, const double* y,
                     const double* z, dsdxo_plan* plan)
{
    for (int n = 0; n < DSDXO_MAX_NEIGHBORS; ++n) plan->send_count[n] = 0;
    const double width = box->width;
    double lo[3], hi[3];
    dsdxo_local_domain(box, lo, hi);
    for (int64_t i = 0; i < owned; ++i) {
        const double px = x[i], py = y[i], pz = z[i];
        const int in_xlo = px < lo[0] + width;
        const int in_xhi = px > hi[0] - width;
        const int in_ylo = py < lo[1] + width;
        const int in_yhi = py > hi[1] - width;
        const int in_zlo = pz < lo[2] + width;
        const int in_zhi = pz > hi[2] - width;
        if (!in_xlo && !in_xhi && !in_ylo && !in_yhi && !in_zlo && !in_zhi)
            continue;
        for (int n = 0; n < nb->count; ++n) {
            if (!nb->active[n]) continue;
            const int32_t* d = nb->direction[n];
            if (d[0] < 0 && !in_xlo) continue;
            if (d[0] > 0 && !in_xhi) continue;
            if (d[1] < 0 && !in_ylo) continue;
            if (d[1] > 0 && !in_yhi) continue;
            if (d[2] < 0 && !in_zlo) continue;
            if (d[2] > 0 && !in_zhi) continue;
            if (push_index(plan, n, (int32_t)i)) return -1;
        }
    }
    plan->ghost_total = 0;
    for (int n = 0; n < nb->count; ++n) {
        plan->recv_from[n] = -1;
        plan->recv_count[n] = 0;
        if (!nb->active[n]) continue;
        const int32_t* d = nb->direction[n];
        const int source = neighbor_with_slot(nb, dsdxo_opposite_slot(d[0], d[1], d[2]));
        if (source < 0) continue;
        plan->recv_from[n] = source;
        plan->recv_count[n] = plan->send_count[source];
        plan->ghost_total += plan->recv_count[n];
    }
    return 0;
}

void dsdxo_plan_free(dsdxo_plan* plan)
{
    for (int n = 0; n < DSDXO_MAX_NEIGHBORS; ++n) {
        free(plan->source[n]);
        plan->source[n] = NULL;
        plan->send_capacity[n] = 0;
        plan->send_count[n] = 0;
    }
}

/* Ghost g of receive block n at position k copies owned particle
 * source[recv_from[n]][k]; blocks follow in neighbour order
 * (src/runtime.cpp:946-957). */
void dsdxo_ghost_owners(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                        int32_t* owners)
{
    int64_t g = 0;
    for (int n = 0; n < nb->count; ++n) {
        const int s = plan->recv_from[n];
        if (s < 0) continue;
        for (int64_t k = 0; k < plan->recv_count[n]; ++k) owners[g++] = plan->source[s][k];
    }
}

static size_t type_size(int32_t type)
{
    switch (type) {
    case DSDXO_INT32: case DSDXO_UINT32: case DSDXO_FLOAT32: return 4;
    default: return 8;
    }
}

/* Ghost records (src/halo.cpp:37-55, make_record) placed after the owned
 * particles in receive order (src/runtime.cpp:946-957 via copy_record
 * :83-98), and the application fields copied byte-for-byte in the same order
 * (src/halo.cpp:120-147 pack, :874-890 unpack).  A NULL source array leaves
 * the record member at its default 0; a NULL destination array is skipped. */
int dsdxo_halo_write_ghosts(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                            int64_t owned, const dsdxo_particles* p,
                            const dsdxo_field* fields, int32_t field_count)
{
    int64_t g = owned;
    for (int n = 0; n < nb->count; ++n) {
        const int s = plan->recv_from[n];
        if (s < 0) continue;
        const double* shift = nb->shift[s];
        for (int64_t k = 0; k < plan->recv_count[n]; ++k, ++g) {
            const int64_t i = plan->source[s][k];
            const double rx = p->x[i] + shift[0];
            const double ry = p->y[i] + shift[1];
            const double rz = p->z[i] + shift[2];
            const double rvx = p->vx ? p->vx[i] : 0.0;
            const double rvy = p->vy ? p->vy[i] : 0.0;
            const double rvz = p->vz ? p->vz[i] : 0.0;
            const double rm = p->mass ? p->mass[i] : 0.0;
            const double rq = p->charge ? p->charge[i] : 0.0;
            const int32_t rt = p->type ? p->type[i] : 0;
            const int32_t rid = p->global_id[i];
            p->x[g] = rx; p->y[g] = ry; p->z[g] = rz;
            if (p->vx) p->vx[g] = rvx;
            if (p->vy) p->vy[g] = rvy;
            if (p->vz) p->vz[g] = rvz;
            if (p->fx) p->fx[g] = 0.0;
            if (p->fy) p->fy[g] = 0.0;
            if (p->fz) p->fz[g] = 0.0;
            if (p->mass) p->mass[g] = rm;
            if (p->charge) p->charge[g] = rq;
            if (p->type) p->type[g] = rt;
            p->global_id[g] = rid;
            for (int32_t f = 0; f < field_count; ++f) {
                const size_t bytes = type_size(fields[f].type) * (size_t)fields[f].components;
                unsigned char* base = (unsigned char*)fields[f].data;
                memmove(base + (size_t)g * bytes, base + (size_t)i * bytes, bytes);
            }
        }
    }
    return 0;
}

/* One element of reduce_field_values (src/halo.cpp:1351-1372).  std::min(a,b)
 * is (b < a) ? b : a and std::max(a,b) is (a < b) ? b : a; integer sums wrap
 * (computed unsigned here to keep C well defined). */
#define DSDXO_REDUCE(T, UT)                                                     \
    static void reduce_##T(T* dst, T v, int op)                                 \
    {                                                                           \
        switch (op) {                                                           \
        case DSDXO_SUM: *dst = (T)((UT)*dst + (UT)v); break;                    \
        case DSDXO_MIN: *dst = (v < *dst) ? v : *dst; break;                    \
        case DSDXO_MAX: *dst = (*dst < v) ? v : *dst; break;                    \
        default: *dst = v; break;                                               \
        }                                                                       \
    }
DSDXO_REDUCE(int32_t, uint32_t)
DSDXO_REDUCE(int64_t, uint64_t)
DSDXO_REDUCE(uint32_t, uint32_t)
DSDXO_REDUCE(uint64_t, uint64_t)

static void reduce_float(float* dst, float v, int op)
{
    switch (op) {
    case DSDXO_SUM: *dst = *dst + v; break;
    case DSDXO_MIN: *dst = (v < *dst) ? v : *dst; break;
    case DSDXO_MAX: *dst = (*dst < v) ? v : *dst; break;
    default: *dst = v; break;
    }
}

static void reduce_double(double* dst, double v, int op)
{
    switch (op) {
    case DSDXO_SUM: *dst = *dst + v; break;
    case DSDXO_MIN: *dst = (v < *dst) ? v : *dst; break;
    case DSDXO_MAX: *dst = (*dst < v) ? v : *dst; break;
    default: *dst = v; break;
    }
}

/* `components` elements of `type` at src reduced into dst with `op`, element
 * by element (reduce_field_values, src/halo.cpp:1351-1372). */
void dsdxo_reduce_elements(int32_t type, int32_t op, void* dst, const void* src,
                           int32_t components)
{
    for (int32_t e = 0; e < components; ++e) {
        switch (type) {
        case DSDXO_INT32: reduce_int32_t((int32_t*)dst + e, ((const int32_t*)src)[e], op); break;
        case DSDXO_INT64: reduce_int64_t((int64_t*)dst + e, ((const int64_t*)src)[e], op); break;
        case DSDXO_UINT32: reduce_uint32_t((uint32_t*)dst + e, ((const uint32_t*)src)[e], op); break;
        case DSDXO_UINT64: reduce_uint64_t((uint64_t*)dst + e, ((const uint64_t*)src)[e], op); break;
        case DSDXO_FLOAT32: reduce_float((float*)dst + e, ((const float*)src)[e], op); break;
        default: reduce_double((double*)dst + e, ((const double*)src)[e], op); break;
        }
    }
}

/* src/halo.cpp:1396-1532 (Host reverse_exchange).  Per field, send block n
 * holds the ghosts of receive block n; the exchange goes through self
 * messages matched by opposite-slot tags (direct_self_copy=false,
 * src/backends/mpi_backend.cpp:510-595), so the values reduced into send
 * list n are those of the ghost block whose neighbour sits in the opposite
 * slot.  Reduction order: neighbour n ascending, then k ascending. */
int dsdxo_reverse_exchange(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                           int64_t owned, const dsdxo_field* fields,
                           int32_t field_count)
{
    int64_t block_start[DSDXO_MAX_NEIGHBORS];
    int64_t offset = owned;
    for (int n = 0; n < nb->count; ++n) {
        block_start[n] = offset;
        offset += plan->recv_count[n];
    }
    for (int32_t f = 0; f < field_count; ++f) {
        const int32_t c = fields[f].components;
        const int op = fields[f].op;
        for (int n = 0; n < nb->count; ++n) {
            if (!nb->active[n]) continue;
            const int32_t* d = nb->direction[n];
            const int from = neighbor_with_slot(nb, dsdxo_opposite_slot(d[0], d[1], d[2]));
            if (from < 0) continue;
            if (plan->recv_count[from] != plan->send_count[n]) return -1;
            const int64_t base = block_start[from];
            for (int64_t k = 0; k < plan->send_count[n]; ++k) {
                const int64_t i = plan->source[n][k];
                for (int32_t e = 0; e < c; ++e) {
                    const int64_t dst = i * c + e, src = (base + k) * c + e;
                    switch (fields[f].type) {
                    case DSDXO_INT32: { int32_t* a = (int32_t*)fields[f].data; reduce_int32_t(a + dst, a[src], op); break; }
                    case DSDXO_INT64: { int64_t* a = (int64_t*)fields[f].data; reduce_int64_t(a + dst, a[src], op); break; }
                    case DSDXO_UINT32: { uint32_t* a = (uint32_t*)fields[f].data; reduce_uint32_t(a + dst, a[src], op); break; }
                    case DSDXO_UINT64: { uint64_t* a = (uint64_t*)fields[f].data; reduce_uint64_t(a + dst, a[src], op); break; }
                    case DSDXO_FLOAT32: { float* a = (float*)fields[f].data; reduce_float(a + dst, a[src], op); break; }
                    default: { double* a = (double*)fields[f].data; reduce_double(a + dst, a[src], op); break; }
                    }
                }
            }
        }
    }
    return 0;
}

/* src/halo.cpp:1288-1347: Host reverse_halo is reverse_exchange(Sum) over
 * the three force components wrapped as Float64 fields. */
int dsdxo_reverse_halo(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                       int64_t owned, double* fx, double* fy, double* fz)
{
    dsdxo_field f[3] = {{fx, DSDXO_FLOAT64, 1, DSDXO_SUM},
                        {fy, DSDXO_FLOAT64, 1, DSDXO_SUM},
                        {fz, DSDXO_FLOAT64, 1, DSDXO_SUM}};
    return dsdxo_reverse_exchange(nb, plan, owned, f, 3);
}

/* src/halo.cpp:1084-1207, :1210-1286 (forward_exchange): owner values of send
 * list s copied into the ghost slots of the receive block that s feeds, ghost
 * blocks in neighbour order. */
int dsdxo_forward_exchange(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                           int64_t owned, const dsdxo_field* fields,
                           int32_t field_count)
{
    for (int32_t f = 0; f < field_count; ++f) {
        const size_t bytes = type_size(fields[f].type) * (size_t)fields[f].components;
        unsigned char* base = (unsigned char*)fields[f].data;
        int64_t g = owned;
        for (int n = 0; n < nb->count; ++n) {
            const int s = plan->recv_from[n];
            if (s < 0) continue;
            for (int64_t k = 0; k < plan->recv_count[n]; ++k, ++g)
                memmove(base + (size_t)g * bytes,
                        base + (size_t)plan->source[s][k] * bytes, bytes);
        }
    }
    return 0;
}

/* src/migration.cpp:102-116 (wrapped_record): on a periodic axis
 * v = x - lower; v -= floor(v / extent) * extent; v + lower, every operation
 * separately rounded and without folding v == extent back to 0.  At 1x1x1
 * every particle is retained in its original order (:437-479), so migrate is
 * this wrap applied in place; non-periodic axes are unchanged. */
void dsdxo_migrate_wrap(const dsdxo_box* box, int64_t owned, double* x,
                        double* y, double* z)
{
    double* axes[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
        if (!box->periodic[a]) continue;
        const double lower = box->lower[a];
        const double extent = box->upper[a] - lower;
        double* v = axes[a];
        for (int64_t i = 0; i < owned; ++i) {
            double value = v[i] - lower;
            const double q = floor(value / extent);
            const double m = q * extent;
            value = value - m;
            v[i] = value + lower;
        }
    }
}

/* src/topology.cpp:66-72: the cut planes of a fresh 1x1x1 topology,
 * global_lower + extent * i / dims for i = 0, 1, unless given. */
void dsdxo_local_domain(const dsdxo_box* box, double lower[3], double upper[3])
{
    for (int a = 0; a < 3; ++a) {
        const double extent = box->upper[a] - box->lower[a];
        lower[a] = box->local_given ? box->local_lower[a] : box->lower[a] + extent * 0.0 / 1.0;
        upper[a] = box->local_given ? box->local_upper[a] : box->lower[a] + extent * 1.0 / 1.0;
    }
}

/* src/decomp.cpp:117-124 */
void dsdxo_rescale_cuts(const double old_lower[3], const double old_upper[3],
                        const double new_lower[3], const double new_upper[3],
                        double cut_lo[3], double cut_hi[3])
{
    for (int a = 0; a < 3; ++a) {
        const double old_extent = old_upper[a] - old_lower[a];
        const double new_extent = new_upper[a] - new_lower[a];
        double* cuts[2] = {&cut_lo[a], &cut_hi[a]};
        for (int c = 0; c < 2; ++c) {
            const double fraction = (*cuts[c] - old_lower[a]) / old_extent;
            *cuts[c] = new_lower[a] + fraction * new_extent;
        }
    }
}
