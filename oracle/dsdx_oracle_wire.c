/*
 * dsdx_oracle_wire.c -- CPU restatement of dsDecompX's wire format and the
 * packing steps of its device path, on the 1x1x1 loopback plan of
 * dsdx_oracle.c.
 *
 * TEST INFRASTRUCTURE ONLY (see dsdx_oracle.h).
 *
 * Formats:
 *   particle record  ParticleRecord (include/field.hpp:150-157): 11 doubles
 *                    x..charge, then int type, int global_id -- 96 bytes;
 *                    registered fields follow back to back from byte 96,
 *                    record stride rounded up to 8 (make_device_wire_layout,
 *                    src/device_fields.cpp:16-53)
 *   exchange record  the fields of a forward / reverse exchange back to back
 *                    from byte 0, each aligned to its element size, stride
 *                    rounded up to 8 (src/halo.cpp:1237-1271, :1422-1457)
 *   position record  x, y, z (src/cuda/pack.cu:204-216), 24 bytes
 * Buffers: a send buffer holds one block per neighbour n in neighbour order,
 * each in send-list order (HaloPackSpec offsets, src/cuda/kernels.hpp:30-37;
 * src/cuda/pack.cu:146-180); a receive buffer holds the received blocks in
 * neighbour order (src/halo.cpp:1017-1024 for the returned records).
 */
#include <stdlib.h>
#include <string.h>

#include "dsdx_oracle.h"

static int64_t type_bytes(int32_t type)
{
    return (type == DSDXO_INT32 || type == DSDXO_UINT32 || type == DSDXO_FLOAT32) ? 4 : 8;
}

static int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

int dsdxo_wire_layout(const dsdxo_field* fields, int32_t count, int32_t kind, int64_t* offsets,
                      int64_t* stride)
{
    int64_t offset = kind == DSDXO_WIRE_PARTICLE ? (int64_t)sizeof(dsdxo_record) : 0;
    for (int32_t f = 0; f < count; ++f) {
        const int64_t element = type_bytes(fields[f].type);
        if (kind == DSDXO_WIRE_EXCHANGE) offset = align_up(offset, element);
        offsets[f] = offset;
        offset += element * fields[f].components;
    }
    *stride = align_up(offset, 8);
    return 0;
}

/* store_halo_record (src/cuda/pack.cu:129-142): owner + periodic shift,
 * velocities, zero forces, mass, charge, type, id, then the field bytes. */
static void store_record(const dsdxo_particles* p, int64_t i, const double* shift,
                         const dsdxo_field* fields, int32_t field_count, const int64_t* offsets,
                         unsigned char* out)
{
    dsdxo_record r;
    memset(&r, 0, sizeof(r));
    r.x = p->x[i] + shift[0];
    r.y = p->y[i] + shift[1];
    r.z = p->z[i] + shift[2];
    r.vx = p->vx ? p->vx[i] : 0.0;
    r.vy = p->vy ? p->vy[i] : 0.0;
    r.vz = p->vz ? p->vz[i] : 0.0;
    r.mass = p->mass ? p->mass[i] : 0.0;
    r.charge = p->charge ? p->charge[i] : 0.0;
    r.type = p->type ? p->type[i] : 0;
    r.global_id = p->global_id[i];
    memcpy(out, &r, sizeof(r));
    for (int32_t f = 0; f < field_count; ++f) {
        const int64_t bytes = type_bytes(fields[f].type) * fields[f].components;
        memcpy(out + offsets[f], (const unsigned char*)fields[f].data + i * bytes, (size_t)bytes);
    }
}

int64_t dsdxo_halo_pack(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                        const dsdxo_particles* p, const dsdxo_field* fields, int32_t field_count,
                        int32_t positions_only, unsigned char* out)
{
    int64_t offsets[DSDXO_MAX_FIELDS], stride = 24;
    if (!positions_only) dsdxo_wire_layout(fields, field_count, DSDXO_WIRE_PARTICLE, offsets, &stride);
    int64_t slot = 0;
    for (int n = 0; n < nb->count; ++n) {
        const double* shift = nb->shift[n];
        for (int64_t k = 0; k < plan->send_count[n]; ++k, ++slot) {
            const int64_t i = plan->source[n][k];
            unsigned char* rec = out + slot * stride;
            if (positions_only) {
                /* halo_gather_positions_kernel (src/cuda/pack.cu:204-216) */
                const double xyz[3] = {p->x[i] + shift[0], p->y[i] + shift[1], p->z[i] + shift[2]};
                memcpy(rec, xyz, sizeof(xyz));
            } else {
                store_record(p, i, shift, fields, field_count, offsets, rec);
            }
        }
    }
    return slot;
}

/* unpack_kernel (src/cuda/unpack.cu:16-31): record k into particle dest + k,
 * every member as sent, fields via load_wire_fields. */
void dsdxo_unpack_records(const unsigned char* records, int64_t count, const dsdxo_particles* p,
                          int64_t dest, const dsdxo_field* fields, int32_t field_count)
{
    int64_t offsets[DSDXO_MAX_FIELDS], stride;
    dsdxo_wire_layout(fields, field_count, DSDXO_WIRE_PARTICLE, offsets, &stride);
    for (int64_t k = 0; k < count; ++k) {
        const unsigned char* rec = records + k * stride;
        dsdxo_record r;
        memcpy(&r, rec, sizeof(r));
        const int64_t i = dest + k;
        p->x[i] = r.x; p->y[i] = r.y; p->z[i] = r.z;
        if (p->vx) p->vx[i] = r.vx;
        if (p->vy) p->vy[i] = r.vy;
        if (p->vz) p->vz[i] = r.vz;
        if (p->fx) p->fx[i] = r.fx;
        if (p->fy) p->fy[i] = r.fy;
        if (p->fz) p->fz[i] = r.fz;
        if (p->mass) p->mass[i] = r.mass;
        if (p->charge) p->charge[i] = r.charge;
        if (p->type) p->type[i] = r.type;
        p->global_id[i] = r.global_id;
        for (int32_t f = 0; f < field_count; ++f) {
            const int64_t bytes = type_bytes(fields[f].type) * fields[f].components;
            memcpy((unsigned char*)fields[f].data + i * bytes, rec + offsets[f], (size_t)bytes);
        }
    }
}

/* unpack_positions_kernel (src/cuda/unpack.cu:33-42): positions, zero forces. */
void dsdxo_unpack_positions(const double* positions, int64_t count, const dsdxo_particles* p,
                            int64_t dest)
{
    for (int64_t k = 0; k < count; ++k) {
        const int64_t i = dest + k;
        p->x[i] = positions[3 * k];
        p->y[i] = positions[3 * k + 1];
        p->z[i] = positions[3 * k + 2];
        if (p->fx) p->fx[i] = 0.0;
        if (p->fy) p->fy[i] = 0.0;
        if (p->fz) p->fz[i] = 0.0;
    }
}

/* pack_reverse_kernel (src/cuda/pack.cu:219-227): a default record with the
 * ghost's forces. */
void dsdxo_pack_reverse_records(const double* fx, const double* fy, const double* fz,
                                int64_t ghost_offset, int64_t count, unsigned char* out)
{
    for (int64_t k = 0; k < count; ++k) {
        dsdxo_record r;
        memset(&r, 0, sizeof(r));
        r.fx = fx[ghost_offset + k];
        r.fy = fy[ghost_offset + k];
        r.fz = fz[ghost_offset + k];
        memcpy(out + k * (int64_t)sizeof(r), &r, sizeof(r));
    }
}

/* The returned records, neighbour block n at the sum of the send counts
 * before it, added to their owners one block after another
 * (src/halo.cpp:1069-1079; accumulate_reverse_kernel,
 * src/cuda/unpack.cu:46-57): an owner gets its contributions in ascending n. */
void dsdxo_accumulate_reverse(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                              const unsigned char* records, double* fx, double* fy, double* fz)
{
    int64_t slot = 0;
    for (int n = 0; n < nb->count; ++n) {
        for (int64_t k = 0; k < plan->send_count[n]; ++k, ++slot) {
            dsdxo_record r;
            memcpy(&r, records + slot * (int64_t)sizeof(r), sizeof(r));
            const int64_t i = plan->source[n][k];
            fx[i] = fx[i] + r.fx;
            fy[i] = fy[i] + r.fy;
            fz[i] = fz[i] + r.fz;
        }
    }
}

/* pack_reverse_fields_kernel / scatter_fields_kernel
 * (src/cuda/reverse_fields.cu:128-144, :196-210):
 * particle first + k <-> exchange record k. */
void dsdxo_pack_field_records(const dsdxo_field* fields, int32_t field_count, int64_t first,
                              int64_t count, unsigned char* out)
{
    int64_t offsets[DSDXO_MAX_FIELDS], stride;
    dsdxo_wire_layout(fields, field_count, DSDXO_WIRE_EXCHANGE, offsets, &stride);
    for (int64_t k = 0; k < count; ++k)
        for (int32_t f = 0; f < field_count; ++f) {
            const int64_t bytes = type_bytes(fields[f].type) * fields[f].components;
            memcpy(out + k * stride + offsets[f],
                   (const unsigned char*)fields[f].data + (first + k) * bytes, (size_t)bytes);
        }
}

void dsdxo_scatter_field_records(const dsdxo_field* fields, int32_t field_count,
                                 const unsigned char* records, int64_t count, int64_t first)
{
    int64_t offsets[DSDXO_MAX_FIELDS], stride;
    dsdxo_wire_layout(fields, field_count, DSDXO_WIRE_EXCHANGE, offsets, &stride);
    for (int64_t k = 0; k < count; ++k)
        for (int32_t f = 0; f < field_count; ++f) {
            const int64_t bytes = type_bytes(fields[f].type) * fields[f].components;
            memcpy((unsigned char*)fields[f].data + (first + k) * bytes,
                   records + k * stride + offsets[f], (size_t)bytes);
        }
}

/* gather_fields_kernel (src/cuda/reverse_fields.cu:182-194, via forward_device_fields,
 * src/halo.cpp:1084-1207): send block n, record k = owner source[n][k]. */
int64_t dsdxo_gather_field_records(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                                   const dsdxo_field* fields, int32_t field_count,
                                   unsigned char* out)
{
    int64_t offsets[DSDXO_MAX_FIELDS], stride;
    dsdxo_wire_layout(fields, field_count, DSDXO_WIRE_EXCHANGE, offsets, &stride);
    int64_t slot = 0;
    for (int n = 0; n < nb->count; ++n)
        for (int64_t k = 0; k < plan->send_count[n]; ++k, ++slot)
            for (int32_t f = 0; f < field_count; ++f) {
                const int64_t bytes = type_bytes(fields[f].type) * fields[f].components;
                memcpy(out + slot * stride + offsets[f],
                       (const unsigned char*)fields[f].data + plan->source[n][k] * bytes,
                       (size_t)bytes);
            }
    return slot;
}

/* reduce_reverse_fields_kernel (src/cuda/reverse_fields.cu:146-180) over the
 * returned blocks in neighbour order (src/halo.cpp:1069-1079): per owner,
 * ascending n. */
void dsdxo_reduce_field_records(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                                const dsdxo_field* fields, int32_t field_count,
                                const unsigned char* records)
{
    int64_t offsets[DSDXO_MAX_FIELDS], stride;
    dsdxo_wire_layout(fields, field_count, DSDXO_WIRE_EXCHANGE, offsets, &stride);
    int64_t slot = 0;
    for (int n = 0; n < nb->count; ++n)
        for (int64_t k = 0; k < plan->send_count[n]; ++k, ++slot)
            for (int32_t f = 0; f < field_count; ++f) {
                const int64_t bytes = type_bytes(fields[f].type) * fields[f].components;
                dsdxo_reduce_elements(fields[f].type, fields[f].op,
                                      (unsigned char*)fields[f].data + plan->source[n][k] * bytes,
                                      records + slot * stride + offsets[f], fields[f].components);
            }
}

/* The self messages of the loopback grid (src/backends/mpi_backend.cpp:
 * 366-375, tags matched by opposite slot): forward, receive block n takes the
 * send block of list recv_from[n]; reverse, the returned block of list s is
 * the receive block fed by s.  `record_bytes` per record. */
void dsdxo_loopback_exchange(const dsdxo_neighbors* nb, const dsdxo_plan* plan,
                             const unsigned char* send, unsigned char* recv, int64_t record_bytes,
                             int32_t reverse)
{
    int64_t send_start[DSDXO_MAX_NEIGHBORS], recv_start[DSDXO_MAX_NEIGHBORS];
    int64_t a = 0, b = 0;
    for (int n = 0; n < nb->count; ++n) {
        send_start[n] = a;
        recv_start[n] = b;
        a += plan->send_count[n];
        b += plan->recv_count[n];
    }
    for (int n = 0; n < nb->count; ++n) {
        const int s = plan->recv_from[n];
        if (s < 0 || plan->recv_count[n] == 0) continue;
        const size_t bytes = (size_t)(plan->recv_count[n] * record_bytes);
        if (!reverse)
            memcpy(recv + recv_start[n] * record_bytes, send + send_start[s] * record_bytes, bytes);
        else
            memcpy(recv + send_start[s] * record_bytes, send + recv_start[n] * record_bytes, bytes);
    }
}
