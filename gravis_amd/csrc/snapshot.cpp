// snapshot.cpp -- sim-core's binary snapshot format (host code).
//
// Reference: crates/sim-core/src/snapshot.rs:5-157. Layout, little endian:
//   magic "NBODY00v" (v = 1, 2, 3), time f64, step u64, softening f64, dt f64,
//   count u64, then x, y, z, vx, vy, vz, mass (count f64 each);
//   v2+: particle_type (count bytes); v3: internal_energy, smoothing_length,
//   density (count f64 each).
// Writing always produces v3 (Snapshot::write_to). Reading accepts v1-v3;
// accelerations and the fields a version does not carry come back zero
// (Snapshot::read_from), so a restart must recompute accelerations first
// (Snapshot::initialize).
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/gravis.h"

namespace {

bool little_endian() {
    const uint16_t one = 1;
    uint8_t b;
    std::memcpy(&b, &one, 1);
    return b == 1;
}

bool put(FILE* f, const void* p, size_t bytes) { return std::fwrite(p, 1, bytes, f) == bytes; }
bool get(FILE* f, void* p, size_t bytes) { return std::fread(p, 1, bytes, f) == bytes; }

struct Header {
    int version = 0;
    double time = 0, softening = 0, dt = 0;
    uint64_t step = 0, count = 0;
};

gravis_status read_header(FILE* f, Header* h) {
    char magic[8];
    if (!get(f, magic, 8)) return GRAVIS_INVALID_ARG;
    if (std::memcmp(magic, "NBODY003", 8) == 0) h->version = 3;
    else if (std::memcmp(magic, "NBODY002", 8) == 0) h->version = 2;
    else if (std::memcmp(magic, "NBODY001", 8) == 0) h->version = 1;
    else return GRAVIS_INVALID_ARG;  // "Invalid snapshot magic bytes"
    if (!get(f, &h->time, 8) || !get(f, &h->step, 8) || !get(f, &h->softening, 8) ||
        !get(f, &h->dt, 8) || !get(f, &h->count, 8))
        return GRAVIS_INVALID_ARG;
    return GRAVIS_OK;
}

}  // namespace

extern "C" gravis_status gravis_snapshot_write(const char* path, double time, uint64_t step,
                                               double softening, double dt,
                                               const gravis_sph_state* s) {
    if (!path || !s || !little_endian()) return GRAVIS_INVALID_ARG;
    const size_t n = s->n;
    if (n > 0 && (!s->x || !s->y || !s->z || !s->vx || !s->vy || !s->vz || !s->m || !s->type ||
                  !s->u || !s->h || !s->rho))
        return GRAVIS_INVALID_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return GRAVIS_INVALID_ARG;
    const uint64_t count = n;
    bool ok = put(f, "NBODY003", 8) && put(f, &time, 8) && put(f, &step, 8) &&
              put(f, &softening, 8) && put(f, &dt, 8) && put(f, &count, 8);
    const double* arrays[7] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->m};
    for (const double* a : arrays) ok = ok && put(f, a, n * 8);
    ok = ok && put(f, s->type, n);
    const double* gas[3] = {s->u, s->h, s->rho};
    for (const double* a : gas) ok = ok && put(f, a, n * 8);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? GRAVIS_OK : GRAVIS_INVALID_ARG;
}

extern "C" gravis_status gravis_snapshot_info(const char* path, int* version, double* time,
                                              uint64_t* step, double* softening, double* dt,
                                              uint64_t* count) {
    if (!path) return GRAVIS_INVALID_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return GRAVIS_INVALID_ARG;
    Header h;
    const gravis_status st = read_header(f, &h);
    std::fclose(f);
    if (st != GRAVIS_OK) return st;
    if (version) *version = h.version;
    if (time) *time = h.time;
    if (step) *step = h.step;
    if (softening) *softening = h.softening;
    if (dt) *dt = h.dt;
    if (count) *count = h.count;
    return GRAVIS_OK;
}

extern "C" gravis_status gravis_snapshot_read(const char* path, gravis_sph_state* s) {
    if (!path || !s || !little_endian()) return GRAVIS_INVALID_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return GRAVIS_INVALID_ARG;
    Header h;
    gravis_status st = read_header(f, &h);
    if (st != GRAVIS_OK || h.count != s->n) {
        std::fclose(f);
        return GRAVIS_INVALID_ARG;
    }
    const size_t n = s->n;
    double* all[17] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->ax, s->ay, s->az,
                       s->m, s->rho, s->p, s->u, s->dudt, s->h, s->c, s->alpha};
    for (double* a : all)
        if (n > 0 && !a) {
            std::fclose(f);
            return GRAVIS_INVALID_ARG;
        }
    if (n > 0 && !s->type) {
        std::fclose(f);
        return GRAVIS_INVALID_ARG;
    }
    for (double* a : all)
        for (size_t i = 0; i < n; ++i) a[i] = 0.0;
    double* arrays[7] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->m};
    bool ok = true;
    for (double* a : arrays) ok = ok && get(f, a, n * 8);
    if (h.version >= 2) ok = ok && get(f, s->type, n);
    else std::memset(s->type, 0, n);
    if (h.version >= 3) {
        double* gas[3] = {s->u, s->h, s->rho};
        for (double* a : gas) ok = ok && get(f, a, n * 8);
    }
    std::fclose(f);
    return ok ? GRAVIS_OK : GRAVIS_INVALID_ARG;
}
