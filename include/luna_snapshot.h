/* Snapshots in LUNA's HDF5 layout, as luna-cpp's hydro_sedov writes them (src/hydro/snapshot.hpp:68-135):
 * <dir>/snapdir_NNN/snapshot_NNN.0.hdf5 with a group `cells` (attributes time, step, gamma, spatial_dimension, order,
 * Lx, Ly, Lz, mesh = "voronoi", N_cells and the build's provenance; datasets x, y, z, volume, density, velocity_x/y/z,
 * internal_energy = P / (gamma - 1), pressure, U [5][N] = Q / V) and, when units are set, a group `units`.
 * Host code (libluna_snapshot.so, linked with the image's HDF5). */
#ifndef LUNA_SNAPSHOT_H
#define LUNA_SNAPSHOT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int index;                   /* snapshot number NNN, also the `step` attribute */
    double time;
    double sides[3];             /* Lx, Ly, Lz */
    int order;
    double gamma, smallr, smallc; /* the gas (primitive() with its floors converts U) */
    double length_cm, mass_g, velocity_cm_s; /* all 1: no units group */
    long long cells;
    const double* positions;     /* N x 3 generators */
    const double* volume;        /* N */
    const double* q;             /* N x 5 extensive conserved state (FiniteVolume::state()); U = Q / V */
    const char* commit;          /* provenance: the library's commit (or "") */
    int dirty;
    /* the mesh producer's provenance, written as voronoi_cpp_commit, voronoi_cpp_dirty and voronoi_cpp_origin (the
     * reference's build_info of its mesh library); NULL strings are written empty */
    const char* mesh_commit;
    int mesh_dirty;
    const char* mesh_origin;
} luna_snapshot;

/* Writes the snapshot; name (capacity cap) receives the file's path. 0 on success, -1 with luna_snapshot_error(). */
int luna_snapshot_write(const char* dir, const luna_snapshot* s, char* name, size_t cap);

/* Reads back one dataset of group `cells` (count doubles; U as 5 x N) and one scalar attribute (as double; integer
 * attributes converted); for round trips and analysis. 0 on success, -1 with luna_snapshot_error(). */
int luna_snapshot_read_dataset(const char* path, const char* name, double* out, size_t count);
int luna_snapshot_read_attribute(const char* path, const char* group, const char* name, double* out);
int luna_snapshot_read_string(const char* path, const char* group, const char* name, char* out, size_t cap);

const char* luna_snapshot_error(void);

/* FNV-1a (64-bit) over bytes, continuing from h (start: 1469598103934665603): the reference drivers' checksum of a state's
 * values, little-endian bytes of each double in row-major order (hydro_sedov.cpp:159-174). */
unsigned long long luna_fnv1a(const void* data, size_t bytes, unsigned long long h);

#ifdef __cplusplus
}
#endif

#endif
