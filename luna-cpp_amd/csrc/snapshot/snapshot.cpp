// LUNA-layout HDF5 snapshots (include/luna_snapshot.h), with the HDF5 C library of the image. The primitives of each cell
// are computed from U = Q / V with the floors of primitive() (euler.hpp:104-113), as the reference's writer does.
#include "../../../include/luna_snapshot.h"

#include <hdf5.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(const std::string& msg) {
    g_error = msg;
    return -1;
}

// Closes an HDF5 handle when it leaves scope.
struct Handle {
    hid_t id;
    herr_t (*close)(hid_t);
    Handle(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    ~Handle() {
        if (id >= 0) close(id);
    }
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
};

bool attribute(hid_t group, const char* name, hid_t type, const void* value) {
    Handle space(H5Screate(H5S_SCALAR), H5Sclose);
    Handle attr(H5Acreate2(group, name, type, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
    return attr.id >= 0 && H5Awrite(attr.id, type, value) >= 0;
}
bool attribute(hid_t group, const char* name, double v) { return attribute(group, name, H5T_NATIVE_DOUBLE, &v); }
bool attribute(hid_t group, const char* name, long long v) { return attribute(group, name, H5T_NATIVE_LLONG, &v); }
bool attribute(hid_t group, const char* name, const std::string& v) {
    Handle type(H5Tcopy(H5T_C_S1), H5Tclose);
    H5Tset_size(type.id, std::max<size_t>(1, v.size()));
    return attribute(group, name, type.id, v.empty() ? "" : v.c_str());
}

// `values` as `rows` equal rows (rows == 1: a 1D dataset).
bool dataset(hid_t group, const char* name, const std::vector<double>& values, hsize_t rows = 1) {
    const hsize_t dims[2] = {rows, values.size() / rows};
    Handle space(H5Screate_simple(rows == 1 ? 1 : 2, rows == 1 ? dims + 1 : dims, nullptr), H5Sclose);
    Handle set(H5Dcreate2(group, name, H5T_NATIVE_DOUBLE, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
    return set.id >= 0 && H5Dwrite(set.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, values.data()) >= 0;
}

// euler.hpp:104-113 primitive() for D = 3.
void primitive3(const double* U, double gamma, double smallr, double smallc, double* W) {
    const double rho = std::fmax(U[0], smallr);
    const double inv = 1.0 / rho;
    const double v[3] = {U[1] * inv, U[2] * inv, U[3] * inv};
    const double smalle = smallc * smallc / gamma / (gamma - 1.0);
    const double kinetic = 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double e = std::fmax(U[4] - kinetic, smalle * rho);
    W[0] = rho;
    W[1] = v[0];
    W[2] = v[1];
    W[3] = v[2];
    W[4] = (gamma - 1.0) * e;
}

}  // namespace

extern "C" {

const char* luna_snapshot_error(void) { return g_error.c_str(); }

unsigned long long luna_fnv1a(const void* data, size_t bytes, unsigned long long h) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t k = 0; k < bytes; ++k) {
        h ^= p[k];
        h *= 1099511628211ULL;
    }
    return h;
}

int luna_snapshot_write(const char* dir, const luna_snapshot* s, char* name, size_t cap) {
    if (!dir || !s || !s->positions || !s->volume || !s->q || s->cells < 0) return fail("luna_snapshot_write: null argument");
    char number[16];
    std::snprintf(number, sizeof number, "%03d", s->index);
    const std::string folder = std::string(dir) + "/snapdir_" + number;
    std::error_code ec;
    std::filesystem::create_directories(folder, ec);
    if (ec) return fail("luna_snapshot_write: cannot create " + folder + ": " + ec.message());
    const std::string path = folder + "/snapshot_" + number + ".0.hdf5";
    const size_t n = static_cast<size_t>(s->cells);
    std::vector<double> x(n), y(n), z(n), vol(n), rho(n), vx(n), vy(n), vz(n), e_int(n), P(n), U(5 * n);
    for (size_t i = 0; i < n; ++i) {
        x[i] = s->positions[3 * i];
        y[i] = s->positions[3 * i + 1];
        z[i] = s->positions[3 * i + 2];
        vol[i] = s->volume[i];
        double u[5], W[5];
        for (int v = 0; v < 5; ++v) u[v] = s->q[5 * i + v] / s->volume[i];
        primitive3(u, s->gamma, s->smallr, s->smallc, W);
        rho[i] = W[0];
        vx[i] = W[1];
        vy[i] = W[2];
        vz[i] = W[3];
        P[i] = W[4];
        e_int[i] = W[4] / (s->gamma - 1.0);
        for (int v = 0; v < 5; ++v) U[v * n + i] = u[v];
    }
    Handle file(H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail("luna_snapshot_write: cannot create " + path);
    {
        Handle cells(H5Gcreate2(file.id, "cells", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        bool ok = cells.id >= 0;
        ok = ok && attribute(cells.id, "time", s->time) && attribute(cells.id, "step", static_cast<long long>(s->index)) &&
             attribute(cells.id, "gamma", s->gamma) && attribute(cells.id, "spatial_dimension", 3LL) &&
             attribute(cells.id, "order", static_cast<long long>(s->order)) && attribute(cells.id, "Lx", s->sides[0]) &&
             attribute(cells.id, "Ly", s->sides[1]) && attribute(cells.id, "Lz", s->sides[2]) &&
             attribute(cells.id, "mesh", std::string("voronoi")) && attribute(cells.id, "N_cells", static_cast<long long>(n)) &&
             attribute(cells.id, "luna_cpp_commit", std::string(s->commit ? s->commit : "")) &&
             attribute(cells.id, "luna_cpp_dirty", static_cast<long long>(s->dirty ? 1 : 0)) &&
             attribute(cells.id, "voronoi_cpp_commit", std::string(s->mesh_commit ? s->mesh_commit : "")) &&
             attribute(cells.id, "voronoi_cpp_dirty", static_cast<long long>(s->mesh_dirty ? 1 : 0)) &&
             attribute(cells.id, "voronoi_cpp_origin", std::string(s->mesh_origin ? s->mesh_origin : ""));
        ok = ok && dataset(cells.id, "x", x) && dataset(cells.id, "y", y) && dataset(cells.id, "z", z) && dataset(cells.id, "volume", vol) &&
             dataset(cells.id, "density", rho) && dataset(cells.id, "velocity_x", vx) && dataset(cells.id, "velocity_y", vy) &&
             dataset(cells.id, "velocity_z", vz) && dataset(cells.id, "internal_energy", e_int) && dataset(cells.id, "pressure", P) &&
             dataset(cells.id, "U", U, 5);
        if (!ok) return fail("luna_snapshot_write: HDF5 write failed in " + path);
    }
    const bool units = !(s->length_cm == 1 && s->mass_g == 1 && s->velocity_cm_s == 1);
    if (units) {
        Handle g(H5Gcreate2(file.id, "units", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
        const double time_s = s->length_cm / s->velocity_cm_s;
        const double density = s->mass_g / (s->length_cm * s->length_cm * s->length_cm);
        const bool ok = g.id >= 0 && attribute(g.id, "UnitLength_in_cm", s->length_cm) && attribute(g.id, "UnitMass_in_g", s->mass_g) &&
                        attribute(g.id, "UnitVelocity_in_cm_per_s", s->velocity_cm_s) && attribute(g.id, "UnitTime_in_s", time_s) &&
                        attribute(g.id, "UnitDensity_in_cgs", density) &&
                        attribute(g.id, "UnitPressure_in_cgs", density * s->velocity_cm_s * s->velocity_cm_s) &&
                        attribute(g.id, "UnitEnergy_in_cgs", s->mass_g * s->velocity_cm_s * s->velocity_cm_s);
        if (!ok) return fail("luna_snapshot_write: HDF5 units failed in " + path);
    }
    if (name && cap) {
        std::strncpy(name, path.c_str(), cap - 1);
        name[cap - 1] = 0;
    }
    return 0;
}

int luna_snapshot_read_dataset(const char* path, const char* name, double* out, size_t count) {
    if (!path || !name || !out) return fail("luna_snapshot_read_dataset: null argument");
    Handle file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(std::string("cannot open ") + path);
    Handle set(H5Dopen2(file.id, (std::string("cells/") + name).c_str(), H5P_DEFAULT), H5Dclose);
    if (set.id < 0) return fail(std::string("no dataset cells/") + name);
    Handle space(H5Dget_space(set.id), H5Sclose);
    if (static_cast<size_t>(H5Sget_simple_extent_npoints(space.id)) != count) return fail(std::string("size of cells/") + name);
    return H5Dread(set.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) >= 0 ? 0 : fail("read failed");
}

int luna_snapshot_read_attribute(const char* path, const char* group, const char* name, double* out) {
    if (!path || !group || !name || !out) return fail("luna_snapshot_read_attribute: null argument");
    Handle file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(std::string("cannot open ") + path);
    Handle g(H5Gopen2(file.id, group, H5P_DEFAULT), H5Gclose);
    if (g.id < 0) return fail(std::string("no group ") + group);
    Handle attr(H5Aopen(g.id, name, H5P_DEFAULT), H5Aclose);
    if (attr.id < 0) return fail(std::string("no attribute ") + name);
    return H5Aread(attr.id, H5T_NATIVE_DOUBLE, out) >= 0 ? 0 : fail("attribute read failed");
}

int luna_snapshot_read_string(const char* path, const char* group, const char* name, char* out, size_t cap) {
    if (!path || !group || !name || !out || !cap) return fail("luna_snapshot_read_string: null argument");
    Handle file(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
    if (file.id < 0) return fail(std::string("cannot open ") + path);
    Handle g(H5Gopen2(file.id, group, H5P_DEFAULT), H5Gclose);
    Handle attr(H5Aopen(g.id, name, H5P_DEFAULT), H5Aclose);
    if (attr.id < 0) return fail(std::string("no attribute ") + name);
    Handle type(H5Aget_type(attr.id), H5Tclose);
    const size_t size = H5Tget_size(type.id);
    std::vector<char> buf(size + 1, 0);
    if (H5Aread(attr.id, type.id, buf.data()) < 0) return fail("attribute read failed");
    std::strncpy(out, buf.data(), cap - 1);
    out[cap - 1] = 0;
    return 0;
}

}  // extern "C"
