/*
 * dsdx_oracle_grid.c -- CPU restatement of dsDecompX's Host communication
 * path on a P-rank Cartesian process grid, every rank held in one process.
 * TEST INFRASTRUCTURE ONLY (see dsdx_oracle.h): the product library never
 * links or loads this file.
 *
 * The per-rank arithmetic is the single-box restatement of dsdx_oracle.c
 * (slab tests, records, reductions) applied to each rank's subdomain and
 * neighbour list; what this file adds is the topology (cuts, row-major ranks,
 * neighbour ranks and shifts, the unsplit-axis skip), the routing of each
 * send list to the peer that receives it (tag matching by direction slot),
 * and migration between ranks.
 */
#include "dsdx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* MPI_Cart_rank of a Cartesian communicator created without reordering:
 * row-major, the last dimension fastest (src/backends/mpi_backend.cpp:92-144). */
static int32_t cart_rank(const int32_t dims[3], const int32_t c[3])
{
    return (c[0] * dims[1] + c[1]) * dims[2] + c[2];
}

/* src/topology.cpp:57-176 (refresh_domain_geometry) for every rank:
 * cuts lower + extent * i / dims, local domain [cut[c], cut[c + 1]), and the
 * neighbour list in (dx, dy, dz) nested order.  A direction that moves along
 * an axis of dimension 1 is skipped (:118-127) unless lift_unsplit is set
 * (the 1x1x1 self-loopback); a step off a non-periodic axis has rank -1 and
 * no shift; a step across a periodic boundary wraps with shift +extent
 * (coordinate < 0) or -extent (coordinate >= dims). */
int dsdxo_grid_build(const dsdxo_box* box, const int32_t dims[3], int32_t lift_unsplit,
                     dsdxo_grid* g)
{
    memset(g, 0, sizeof(*g));
    for (int a = 0; a < 3; ++a)
        if (dims[a] < 1) return -1;
    const int64_t ranks = (int64_t)dims[0] * dims[1] * dims[2];
    if (ranks > DSDXO_MAX_RANKS) return -1;
    g->global = *box;
    g->ranks = (int32_t)ranks;
    for (int a = 0; a < 3; ++a) {
        g->dims[a] = dims[a];
        const double extent = box->upper[a] - box->lower[a];
        for (int i = 0; i <= dims[a]; ++i)
            g->cuts[a][i] = box->lower[a] + extent * (double)i / (double)dims[a];
    }
    for (int32_t c0 = 0; c0 < dims[0]; ++c0)
        for (int32_t c1 = 0; c1 < dims[1]; ++c1)
            for (int32_t c2 = 0; c2 < dims[2]; ++c2) {
                const int32_t c[3] = {c0, c1, c2};
                const int32_t r = cart_rank(dims, c);
                dsdxo_box* local = &g->local[r];
                *local = *box;
                for (int a = 0; a < 3; ++a) {
                    g->coord[r][a] = c[a];
                    local->lower[a] = g->cuts[a][c[a]];
                    local->upper[a] = g->cuts[a][c[a] + 1];
                    local->local_lower[a] = local->lower[a];
                    local->local_upper[a] = local->upper[a];
                }
                local->local_given = 1;
                dsdxo_neighbors* nb = &g->nb[r];
                int n = 0;
                for (int dx = -1; dx <= 1; ++dx)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dz = -1; dz <= 1; ++dz) {
                            if (dx == 0 && dy == 0 && dz == 0) continue;
                            if (!box->edges_and_corners && abs(dx) + abs(dy) + abs(dz) != 1)
                                continue;
                            const int d[3] = {dx, dy, dz};
                            int unsplit = 0;
                            for (int a = 0; a < 3; ++a)
                                if (d[a] != 0 && dims[a] == 1) unsplit = 1;
                            if (unsplit && !lift_unsplit) continue;
                            int32_t nc[3];
                            double shift[3] = {0.0, 0.0, 0.0};
                            int outside = 0;
                            for (int a = 0; a < 3; ++a) {
                                nc[a] = c[a] + d[a];
                                if (nc[a] < 0 || nc[a] >= dims[a]) {
                                    if (!box->periodic[a]) {
                                        outside = 1;
                                        break;
                                    }
                                    const double extent = box->upper[a] - box->lower[a];
                                    if (nc[a] < 0) {
                                        nc[a] += dims[a];
                                        shift[a] = extent;
                                    } else {
                                        nc[a] -= dims[a];
                                        shift[a] = -extent;
                                    }
                                }
                            }
                            for (int a = 0; a < 3; ++a) {
                                nb->direction[n][a] = d[a];
                                nb->shift[n][a] = outside ? 0.0 : shift[a];
                            }
                            nb->slot[n] = dsdxo_direction_slot(dx, dy, dz);
                            nb->active[n] = outside ? 0 : 1;
                            g->nb_rank[r][n] = outside ? -1 : cart_rank(dims, nc);
                            ++n;
                        }
                nb->count = n;
            }
    return 0;
}

/* src/migration.cpp:74-98 (owner_rank): wrap each periodic coordinate, then
 * the cut interval by upper_bound - 1, clamped to [0, dims - 1]. */
int32_t dsdxo_grid_owner_rank(const dsdxo_grid* g, double x, double y, double z)
{
    const double pos[3] = {x, y, z};
    int32_t c[3];
    for (int a = 0; a < 3; ++a) {
        double v = pos[a];
        if (g->global.periodic[a]) {
            const double lower = g->global.lower[a];
            const double extent = g->global.upper[a] - lower;
            v = v - lower;
            v = v - floor(v / extent) * extent;
            v = v + lower;
        }
        /* upper_bound over cuts[0..dims]: first index with cut > v */
        int k = 0;
        while (k <= g->dims[a] && !(v < g->cuts[a][k])) ++k;
        int coordinate = k - 1;
        if (coordinate < 0) coordinate = 0;
        if (coordinate > g->dims[a] - 1) coordinate = g->dims[a] - 1;
        c[a] = coordinate;
    }
    return cart_rank(g->dims, c);
}

static size_t grid_type_size(int32_t type)
{
    switch (type) {
    case DSDXO_INT32: case DSDXO_UINT32: case DSDXO_FLOAT32: return 4;
    default: return 8;
    }
}

/* One owned record in flight (ParticleRecord, include/field.hpp:150-157),
 * with the bytes of every registered field. */
typedef struct {
    double x, y, z, vx, vy, vz, fx, fy, fz, mass, charge;
    int32_t type, id;
    unsigned char* field_bytes;
} record_t;

static double get_or_zero(const double* a, int64_t i) { return a ? a[i] : 0.0; }

/* src/migration.cpp:391-591 (Host migrate): every owned particle becomes its
 * wrapped record (wrapped_record, :102-128; non-periodic axes unchanged) and
 * goes to owner_rank of the wrapped position.  Rank q's new owned array is
 * its retained records in their original order, followed by the records
 * received from ranks 0, 1, ..., P-1 in rank order, each in its sender's
 * order (:511-518); fields travel the same way.  In place on each rank's
 * arrays; owned[] is updated.  Returns 0, or -1 when a rank would exceed its
 * capacity (nothing is changed then) or memory runs out. */
int dsdxo_grid_migrate(const dsdxo_grid* g, dsdxo_particles* p, const int64_t* capacity,
                       int64_t* owned, const dsdxo_field* fields, int32_t field_count,
                       int64_t* sent, int64_t* received)
{
    const int32_t P = g->ranks;
    size_t field_bytes = 0;
    for (int32_t f = 0; f < field_count; ++f)
        field_bytes += grid_type_size(fields[f].type) * (size_t)fields[f].components;
    int64_t total = 0;
    for (int32_t r = 0; r < P; ++r) total += owned[r];
    record_t* rec = (record_t*)calloc((size_t)(total ? total : 1), sizeof(record_t));
    int32_t* from = (int32_t*)calloc((size_t)(total ? total : 1), sizeof(int32_t));
    int32_t* dest = (int32_t*)calloc((size_t)(total ? total : 1), sizeof(int32_t));
    unsigned char* bytes = (unsigned char*)calloc((size_t)(total ? total : 1), field_bytes ? field_bytes : 1);
    int64_t* count = (int64_t*)calloc((size_t)P, sizeof(int64_t));
    if (!rec || !from || !dest || !bytes || !count) {
        free(rec); free(from); free(dest); free(bytes); free(count);
        return -1;
    }
    int64_t k = 0;
    for (int32_t r = 0; r < P; ++r) {
        const dsdxo_particles* q = &p[r];
        for (int64_t i = 0; i < owned[r]; ++i, ++k) {
            record_t* t = &rec[k];
            double pos[3] = {q->x[i], q->y[i], q->z[i]};
            for (int a = 0; a < 3; ++a) {
                if (!g->global.periodic[a]) continue;
                const double lower = g->global.lower[a];
                const double extent = g->global.upper[a] - lower;
                double v = pos[a] - lower;
                const double qq = floor(v / extent);
                const double m = qq * extent;
                v = v - m;
                pos[a] = v + lower;
            }
            t->x = pos[0]; t->y = pos[1]; t->z = pos[2];
            t->vx = get_or_zero(q->vx, i); t->vy = get_or_zero(q->vy, i); t->vz = get_or_zero(q->vz, i);
            t->fx = get_or_zero(q->fx, i); t->fy = get_or_zero(q->fy, i); t->fz = get_or_zero(q->fz, i);
            t->mass = get_or_zero(q->mass, i); t->charge = get_or_zero(q->charge, i);
            t->type = q->type ? q->type[i] : 0;
            t->id = q->global_id[i];
            t->field_bytes = bytes + (size_t)k * field_bytes;
            size_t off = 0;
            for (int32_t f = 0; f < field_count; ++f) {
                const dsdxo_field* fr = &fields[(size_t)r * (size_t)field_count + (size_t)f];
                const size_t b = grid_type_size(fr->type) * (size_t)fr->components;
                memcpy(t->field_bytes + off, (const unsigned char*)fr->data + (size_t)i * b, b);
                off += b;
            }
            from[k] = r;
            dest[k] = dsdxo_grid_owner_rank(g, t->x, t->y, t->z);
            count[dest[k]]++;
        }
    }
    for (int32_t r = 0; r < P; ++r) {
        if (count[r] > capacity[r]) {
            free(rec); free(from); free(dest); free(bytes); free(count);
            return -1;
        }
    }
    for (int32_t r = 0; r < P; ++r) {
        if (sent) sent[r] = 0;
        if (received) received[r] = 0;
    }
    /* start of every sender's records */
    int64_t* start = (int64_t*)calloc((size_t)P + 1, sizeof(int64_t));
    if (!start) {
        free(rec); free(from); free(dest); free(bytes); free(count);
        return -1;
    }
    for (int32_t r = 0; r < P; ++r) start[r + 1] = start[r] + owned[r];
    for (int32_t q = 0; q < P; ++q) {
        dsdxo_particles* o = &p[q];
        int64_t w = 0;
        /* retained first, then senders 0..P-1 (the own rank already placed) */
        for (int32_t pass = -1; pass < P; ++pass) {
            const int32_t s = pass < 0 ? q : pass;
            if (pass >= 0 && s == q) continue;
            for (int64_t j = start[s]; j < start[s + 1]; ++j) {
                if (dest[j] != q) continue;
                const record_t* t = &rec[j];
                o->x[w] = t->x; o->y[w] = t->y; o->z[w] = t->z;
                if (o->vx) o->vx[w] = t->vx;
                if (o->vy) o->vy[w] = t->vy;
                if (o->vz) o->vz[w] = t->vz;
                if (o->fx) o->fx[w] = t->fx;
                if (o->fy) o->fy[w] = t->fy;
                if (o->fz) o->fz[w] = t->fz;
                if (o->mass) o->mass[w] = t->mass;
                if (o->charge) o->charge[w] = t->charge;
                if (o->type) o->type[w] = t->type;
                o->global_id[w] = t->id;
                size_t off = 0;
                for (int32_t f = 0; f < field_count; ++f) {
                    const dsdxo_field* fr = &fields[(size_t)q * (size_t)field_count + (size_t)f];
                    const size_t b = grid_type_size(fr->type) * (size_t)fr->components;
                    memcpy((unsigned char*)fr->data + (size_t)w * b, t->field_bytes + off, b);
                    off += b;
                }
                if (s != q) {
                    if (sent) sent[s]++;
                    if (received) received[q]++;
                }
                ++w;
            }
        }
        owned[q] = w;
    }
    free(start);
    free(rec); free(from); free(dest); free(bytes); free(count);
    return 0;
}

/* Entry of rank s's list that receives what rank r sends through its entry n:
 * the entry of s with the opposite direction (tags: the sender's direction
 * slot is matched by the receiver's opposite slot,
 * src/backends/mpi_backend.cpp:391-433). */
static int peer_entry(const dsdxo_grid* g, int32_t s, const int32_t d[3])
{
    const dsdxo_neighbors* nb = &g->nb[s];
    for (int k = 0; k < nb->count; ++k) {
        const int32_t* e = nb->direction[k];
        if (e[0] == -d[0] && e[1] == -d[1] && e[2] == -d[2]) return k;
    }
    return -1;
}

/* src/halo.cpp:164-238 on every rank (its subdomain and neighbour list),
 * then src/backends/mpi_backend.cpp:344-433: rank q's receive block n holds
 * the send list of rank nb_rank[q][n] through that rank's entry of the
 * opposite direction, shifted by the sender's periodic shift; blocks follow
 * in q's neighbour order after its owned particles (src/runtime.cpp:946-957).
 * plans[r].recv_from[n] = sender's entry, recv_rank[r * 26 + n] = sender's
 * rank (-1 when none).  Ghost records as make_record (src/halo.cpp:37-55):
 * forces zero, fields byte-for-byte.  Returns 0, or -1 when a rank's
 * owned + ghosts exceed its capacity (its plan is then built but nothing is
 * written) or memory runs out. */
int dsdxo_grid_halo(const dsdxo_grid* g, dsdxo_particles* p, const int64_t* capacity,
                    const int64_t* owned, dsdxo_plan* plans, int32_t* recv_rank,
                    int64_t* ghosts, const dsdxo_field* fields, int32_t field_count)
{
    const int32_t P = g->ranks;
    for (int32_t r = 0; r < P; ++r)
        if (dsdxo_halo_build(&g->local[r], &g->nb[r], owned[r], p[r].x, p[r].y, p[r].z, &plans[r]))
            return -1;
    int rc = 0;
    for (int32_t q = 0; q < P; ++q) {
        const dsdxo_neighbors* nb = &g->nb[q];
        int64_t total = 0;
        for (int n = 0; n < DSDXO_MAX_NEIGHBORS; ++n) {
            plans[q].recv_from[n] = -1;
            plans[q].recv_count[n] = 0;
            recv_rank[q * DSDXO_MAX_NEIGHBORS + n] = -1;
        }
        for (int n = 0; n < nb->count; ++n) {
            const int32_t s = g->nb_rank[q][n];
            if (s < 0) continue;
            const int e = peer_entry(g, s, nb->direction[n]);
            if (e < 0 || g->nb_rank[s][e] != q) continue;
            plans[q].recv_from[n] = e;
            plans[q].recv_count[n] = plans[s].send_count[e];
            recv_rank[q * DSDXO_MAX_NEIGHBORS + n] = s;
            total += plans[q].recv_count[n];
        }
        plans[q].ghost_total = total;
        ghosts[q] = total;
        if (owned[q] + total > capacity[q]) rc = -1;
    }
    if (rc) return rc;
    for (int32_t q = 0; q < P; ++q) {
        const dsdxo_neighbors* nb = &g->nb[q];
        int64_t w = owned[q];
        dsdxo_particles* o = &p[q];
        for (int n = 0; n < nb->count; ++n) {
            const int32_t s = recv_rank[q * DSDXO_MAX_NEIGHBORS + n];
            if (s < 0) continue;
            const int e = plans[q].recv_from[n];
            const double* shift = g->nb[s].shift[e];
            const dsdxo_particles* src = &p[s];
            for (int64_t k = 0; k < plans[s].send_count[e]; ++k, ++w) {
                const int64_t i = plans[s].source[e][k];
                o->x[w] = src->x[i] + shift[0];
                o->y[w] = src->y[i] + shift[1];
                o->z[w] = src->z[i] + shift[2];
                if (o->vx) o->vx[w] = get_or_zero(src->vx, i);
                if (o->vy) o->vy[w] = get_or_zero(src->vy, i);
                if (o->vz) o->vz[w] = get_or_zero(src->vz, i);
                if (o->fx) o->fx[w] = 0.0;
                if (o->fy) o->fy[w] = 0.0;
                if (o->fz) o->fz[w] = 0.0;
                if (o->mass) o->mass[w] = get_or_zero(src->mass, i);
                if (o->charge) o->charge[w] = get_or_zero(src->charge, i);
                if (o->type) o->type[w] = src->type ? src->type[i] : 0;
                o->global_id[w] = src->global_id[i];
                for (int32_t f = 0; f < field_count; ++f) {
                    const dsdxo_field* fo = &fields[(size_t)q * (size_t)field_count + (size_t)f];
                    const dsdxo_field* fs = &fields[(size_t)s * (size_t)field_count + (size_t)f];
                    const size_t b = grid_type_size(fo->type) * (size_t)fo->components;
                    memcpy((unsigned char*)fo->data + (size_t)w * b,
                           (const unsigned char*)fs->data + (size_t)i * b, b);
                }
            }
        }
    }
    return 0;
}

#define GRID_REDUCE(T, UT)                                                      \
    static void grid_reduce_##T(T* dst, T v, int op)                            \
    {                                                                           \
        switch (op) {                                                           \
        case DSDXO_SUM: *dst = (T)((UT)*dst + (UT)v); break;                    \
        case DSDXO_MIN: *dst = (v < *dst) ? v : *dst; break;                    \
        case DSDXO_MAX: *dst = (*dst < v) ? v : *dst; break;                    \
        default: *dst = v; break;                                               \
        }                                                                       \
    }
GRID_REDUCE(int32_t, uint32_t)
GRID_REDUCE(int64_t, uint64_t)
GRID_REDUCE(uint32_t, uint32_t)
GRID_REDUCE(uint64_t, uint64_t)

static void grid_reduce_float(float* dst, float v, int op)
{
    switch (op) {
    case DSDXO_SUM: *dst = *dst + v; break;
    case DSDXO_MIN: *dst = (v < *dst) ? v : *dst; break;
    case DSDXO_MAX: *dst = (*dst < v) ? v : *dst; break;
    default: *dst = v; break;
    }
}

static void grid_reduce_double(double* dst, double v, int op)
{
    switch (op) {
    case DSDXO_SUM: *dst = *dst + v; break;
    case DSDXO_MIN: *dst = (v < *dst) ? v : *dst; break;
    case DSDXO_MAX: *dst = (*dst < v) ? v : *dst; break;
    default: *dst = v; break;
    }
}

/* Start of receive block n of rank q's ghosts. */
static int64_t block_start(const dsdxo_grid* g, const dsdxo_plan* plans, const int64_t* owned,
                           int32_t q, int n)
{
    int64_t s = owned[q];
    for (int k = 0; k < n; ++k) s += plans[q].recv_count[k];
    (void)g;
    return s;
}

/* src/halo.cpp:1396-1532 (Host reverse_exchange) on every rank: per field,
 * for the rank's neighbour entries n ascending, the ghost values that the
 * peer received through send list n (its entry of the opposite direction)
 * come back and are reduced into the owners source[n][k], k ascending.
 * fields holds field_count entries per rank, rank-major. */
int dsdxo_grid_reverse_exchange(const dsdxo_grid* g, const dsdxo_plan* plans,
                                const int64_t* owned, const dsdxo_field* fields,
                                int32_t field_count)
{
    for (int32_t r = 0; r < g->ranks; ++r) {
        const dsdxo_neighbors* nb = &g->nb[r];
        for (int32_t f = 0; f < field_count; ++f) {
            const dsdxo_field* fr = &fields[(size_t)r * (size_t)field_count + (size_t)f];
            const int32_t c = fr->components;
            const int op = fr->op;
            for (int n = 0; n < nb->count; ++n) {
                const int32_t s = g->nb_rank[r][n];
                if (s < 0 || plans[r].send_count[n] == 0) continue;
                const int e = peer_entry(g, s, nb->direction[n]);
                if (e < 0 || g->nb_rank[s][e] != r) continue;
                if (plans[s].recv_count[e] != plans[r].send_count[n]) return -1;
                const dsdxo_field* fs = &fields[(size_t)s * (size_t)field_count + (size_t)f];
                const int64_t base = block_start(g, plans, owned, s, e);
                for (int64_t k = 0; k < plans[r].send_count[n]; ++k) {
                    const int64_t i = plans[r].source[n][k];
                    for (int32_t el = 0; el < c; ++el) {
                        const int64_t dst = i * c + el, src = (base + k) * c + el;
                        switch (fr->type) {
                        case DSDXO_INT32: grid_reduce_int32_t((int32_t*)fr->data + dst, ((int32_t*)fs->data)[src], op); break;
                        case DSDXO_INT64: grid_reduce_int64_t((int64_t*)fr->data + dst, ((int64_t*)fs->data)[src], op); break;
                        case DSDXO_UINT32: grid_reduce_uint32_t((uint32_t*)fr->data + dst, ((uint32_t*)fs->data)[src], op); break;
                        case DSDXO_UINT64: grid_reduce_uint64_t((uint64_t*)fr->data + dst, ((uint64_t*)fs->data)[src], op); break;
                        case DSDXO_FLOAT32: grid_reduce_float((float*)fr->data + dst, ((float*)fs->data)[src], op); break;
                        default: grid_reduce_double((double*)fr->data + dst, ((double*)fs->data)[src], op); break;
                        }
                    }
                }
            }
        }
    }
    return 0;
}

/* src/halo.cpp:1210-1286 (forward_exchange) on every rank: each ghost slot
 * receives the current value of its owner on the sending rank. */
int dsdxo_grid_forward_exchange(const dsdxo_grid* g, const dsdxo_plan* plans,
                                const int32_t* recv_rank, const int64_t* owned,
                                const dsdxo_field* fields, int32_t field_count)
{
    for (int32_t q = 0; q < g->ranks; ++q) {
        const dsdxo_neighbors* nb = &g->nb[q];
        for (int32_t f = 0; f < field_count; ++f) {
            const dsdxo_field* fq = &fields[(size_t)q * (size_t)field_count + (size_t)f];
            const size_t b = grid_type_size(fq->type) * (size_t)fq->components;
            int64_t w = owned[q];
            for (int n = 0; n < nb->count; ++n) {
                const int32_t s = recv_rank[q * DSDXO_MAX_NEIGHBORS + n];
                if (s < 0) continue;
                const int e = plans[q].recv_from[n];
                const dsdxo_field* fs = &fields[(size_t)s * (size_t)field_count + (size_t)f];
                for (int64_t k = 0; k < plans[s].send_count[e]; ++k, ++w)
                    memcpy((unsigned char*)fq->data + (size_t)w * b,
                           (const unsigned char*)fs->data + (size_t)plans[s].source[e][k] * b, b);
            }
        }
    }
    return 0;
}
