// The C-ABI of libluna_hydro.so (include/luna_hydro.h): solver handles, device memory, the launch sequence of a step,
// HIP-event kernel timing and error reporting. One HIP stream per handle (default priority).
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <algorithm>
#include <cstdio>
#include <thread>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/luna_hydro.h"
#include "kernels.hpp"

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& what) {
    g_error = what;
    return code;
}

#define LUNA_HIP(call)                                                                                       \
    do {                                                                                                     \
        const hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return fail(LUNA_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// A device array that grows by reallocation (moving meshes change their face count every step).
template <class T>
struct Buf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n) {
        if (n <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        const size_t want = n + n / 8;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (want > 0 ? want : 1));
        if (e == hipSuccess) cap = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Kernel timing: events around each launch while on; read at the next synchronisation.
struct Profiler {
    bool on = false;
    struct Pending {
        int name;
        hipEvent_t start, stop;
    };
    std::vector<std::string> names;
    std::vector<long long> launches;
    std::vector<double> total_ms;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;

    int id(const char* name) {
        for (size_t k = 0; k < names.size(); ++k)
            if (names[k] == name) return static_cast<int>(k);
        names.emplace_back(name);
        launches.push_back(0);
        total_ms.push_back(0);
        return static_cast<int>(names.size() - 1);
    }
    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    hipError_t collect() {
        for (const Pending& p : pending) {
            hipError_t e = hipEventSynchronize(p.stop);
            if (e != hipSuccess) return e;
            float ms = 0;
            e = hipEventElapsedTime(&ms, p.start, p.stop);
            if (e != hipSuccess) return e;
            launches[p.name] += 1;
            total_ms[p.name] += ms;
            pool.push_back(p.start);
            pool.push_back(p.stop);
        }
        pending.clear();
        return hipSuccess;
    }
    void clear() {
        (void)collect();
        for (auto& t : total_ms) t = 0;
        for (auto& l : launches) l = 0;
    }
    void destroy() {
        (void)collect();
        for (hipEvent_t e : pool) (void)hipEventDestroy(e);
        pool.clear();
    }
};

// Runs launch() between two events when profiling.
template <class F>
hipError_t timed(Profiler& prof, hipStream_t s, const char* name, F&& launch) {
    if (!prof.on) return launch();
    Profiler::Pending p{prof.id(name), prof.take(), prof.take()};
    hipError_t e = hipEventRecord(p.start, s);
    if (e != hipSuccess) return e;
    e = launch();
    if (e != hipSuccess) return e;
    e = hipEventRecord(p.stop, s);
    if (e != hipSuccess) return e;
    prof.pending.push_back(p);
    return hipSuccess;
}

// IEEE binary16 bits -> float (exact): subnormals m 2^-24, normals (1024 + m) 2^(e - 25).
float half_value(unsigned short h) {
    const int e = (h >> 10) & 0x1f, m = h & 0x3ff;
    const float sign = (h & 0x8000) ? -1.0f : 1.0f;
    if (e == 0x1f) return m ? NAN : sign * INFINITY;
    return sign * (e == 0 ? std::ldexp(static_cast<float>(m), -24) : std::ldexp(static_cast<float>(1024 + m), e - 25));
}

double bits_to_double(unsigned long long b) {
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

}  // namespace

struct luna_fv {
    luna_fv_options opt{};
    int D = 3, V = 5, N = 0, F = 0, E = 0, slots = 0;
    hipStream_t stream = nullptr;
    Buf<int> count, slice, list, other, steps;
    Buf<double> cell, volume, facerec, slotgeo;
    Buf<double> q, grad, w, x, old_centroid, flux;
    Buf<unsigned int> exponents;  // grad16: the per-cell group exponents (int8 x 3)
    Buf<unsigned char> hll;       // the HLL fallback's per-cell compression flags
    Buf<unsigned long long> counters;  // [0, 1] dt slots, [2, 3] fallbacks of the step's two flux passes
    // the derivatives' rows (first use): luna::Derivs
    Buf<double> d_dq, d_dW, d_dgrad, d_dflux, d_sidebar, d_bbar, d_boundbar, d_out;
    Buf<int> d_extreme;
    // the moving step's reverse (first use)
    Buf<double> m_row, m_Ab, m_Xfb, m_Wbar, m_gbar, m_whbar, m_c0bar, m_cbar, m_xbar, m_beta, m_boundbar, m_entry, m_angle, m_Vbar, m_out, m_qbar;
    Buf<int> m_extreme, m_fmax;
    // the moving step's tangent (first use)
    Buf<double> t_dx, t_dx1, t_dV, t_dc, t_dc0, t_dA, t_dxf, t_dW, t_dcs, t_dw, t_dg, t_dG, t_dq;
    // the learned limiter: its features and troubled flags (recorded by each sweep) and the network copy
    Buf<unsigned short> features;
    Buf<unsigned char> troubled, net;
    bool has_net = false, features_ready = false;
    unsigned long long* host = nullptr;  // pinned: read-backs
    hipEvent_t sized = nullptr;          // the geometry pass's slot count and checks are on the host
    double side[3] = {0, 0, 0};
    luna::Grid grid{};
    int parity = 0;
    long long updates = 0;  // mesh updates of moving steps (MeshStats::updates, finite_volume.hpp:1237)
    bool prims_ready = false, swept = false, begun = false, new_mesh = false;
    bool tangent_swept = false;  // time_step_tangent wrote the fit's tangent rows for the next step_tangent (same dq)
    Profiler prof;
    // a host mesh's arrays in device memory (the geometry pass reads them as it reads a device mesh); the centroids,
    // positions and volumes share src_cells
    Buf<int> src_owner, src_neighbour, src_start, src_list;
    Buf<double> src_area, src_fc, src_cells;
    // the geometry pass's total slot count and its checks' flags
    Buf<long long> geo_info;
    Buf<unsigned int> geo_flags;
    bool mesh_ok = false;  // the last mesh given was accepted
    // The accepted mesh's arrays: the step reads its face arrays in place (a device mesh's, or the copies of a host mesh in
    // the src_ buffers); `records` once the derivatives' face records are built for it.
    luna::MeshIn mesh{};
    bool records = false;

    luna::Geo geo() const {
        luna::Geo g{};
        g.N = N;
        g.F = F;
        g.count = count.p;
        g.slice = slice.p;
        g.list = list.p;
        g.other = other.p;
        g.cell = cell.p;
        g.volume = volume.p;
        g.face = facerec.p;
        g.owner = mesh.owner;
        g.neighbour = mesh.neighbour;
        g.area = mesh.area;
        g.fcen = mesh.face_centroid;
        g.slot = slotgeo.p;
        g.slots = slots;
        for (int a = 0; a < 3; ++a) {
            g.side[a] = side[a];
            g.inverse[a] = side[a] > 0 ? 1.0 / side[a] : 0.0;
        }
        return g;
    }
    int grad_row() const { return D == 3 ? luna::grad_row<3>() : luna::grad_row<2>(); }
    luna::Rows rows() const {
        return luna::Rows{q.p, grad.p, w.p, x.p, old_centroid.p, steps.p, flux.p, features.p, troubled.p, opt.precision, exponents.p,
                          opt.hll_compression > 0 ? hll.p : nullptr, opt.hll_compression, opt.sweep_lanes};
    }
    // the kernels' Riemann selector: 2 = HLLC with the HLL fallback at flagged cells
    int riemann() const {
        return opt.hll_compression > 0 && opt.riemann == LUNA_RIEMANN_HLLC ? 2 : opt.riemann;
    }
    luna::Gas gas() const { return luna::Gas{opt.gamma, opt.smallr, opt.smallc, 1.0 / (opt.gamma - 1.0)}; }
    luna::Counters counters_now() const {
        return luna::Counters{counters.p + parity, counters.p + (parity ^ 1), counters.p + 2};
    }
};

namespace {

int check_geometry(const luna_geometry* m, int D, int N) {
    if (!m) return fail(LUNA_ERR_ARGUMENT, "luna: null mesh");
    if (m->dim != D) return fail(LUNA_ERR_ARGUMENT, "luna: mesh dimension differs from the solver's");
    if (m->cells != N) return fail(LUNA_ERR_ARGUMENT, "luna: mesh cell count differs from the solver's");
    if (m->faces < 0 || !m->owner || !m->neighbour || !m->start || !m->list || !m->area || !m->face_centroid || !m->centroid ||
        !m->volume || !m->positions)
        return fail(LUNA_ERR_ARGUMENT, "luna: mesh arrays missing");
    for (int a = 0; a < D; ++a)
        if (!(m->side[a] > 0)) return fail(LUNA_ERR_ARGUMENT, "luna: box sides must be positive");
    return LUNA_OK;
}

// The mesh a handle steps on, from the mesh producer's arrays (device pointers in `in`) to the rows the kernels read: the
// geometry pass of geometry_kernels.hip on the handle's stream. One read-back sizes the list slots after the scan, one
// at the end collects the pass's checks. The face arrays (owner, neighbour, area, face centroid) are read in place by
// the steps that follow, until the next mesh; the cell arrays and lists are not read after this returns. A malformed
// mesh (CSR offsets, cell or face indices out of range, more than 2^31 slots) is refused, and the handle takes no step
// until a valid mesh is set.
int records(luna_fv* fv);

int derive(luna_fv* fv, const luna_geometry* m, luna::MeshIn in, bool eager_records) {
    const int D = fv->D, N = fv->N;
    fv->mesh_ok = false;
    if (fv->opt.motion == LUNA_MOTION_MOVING) {
        if (!(m->step > 0)) return fail(LUNA_ERR_ARGUMENT, "luna: a moving mesh needs its snapping grid step");
        for (int a = 0; a < D; ++a)
            if (m->period[a] <= 0) return fail(LUNA_ERR_ARGUMENT, "luna: a moving mesh needs the grid's periods");
    }
    const int F = m->faces;
    const int S = (N + luna::SLICE - 1) / luna::SLICE;
    for (int a = 0; a < 3; ++a) {
        in.side[a] = a < D ? m->side[a] : 0.0;
        in.inverse[a] = in.side[a] > 0 ? 1.0 / in.side[a] : 0.0;
    }
    hipStream_t s = fv->stream;
    LUNA_HIP(fv->slice.reserve(S));
    LUNA_HIP(fv->facerec.reserve(static_cast<size_t>(F) * luna::FACE_ROW));
    LUNA_HIP(fv->flux.reserve(static_cast<size_t>(F) * fv->V));
    LUNA_HIP(hipMemsetAsync(fv->geo_flags.p, 0, sizeof(unsigned int), s));
    LUNA_HIP(timed(fv->prof, s, "geometry_cells", [&] {
        return D == 3 ? luna::launch_geometry_cells<3>(s, in, fv->cell.p, fv->volume.p, fv->count.p, fv->slice.p, fv->geo_flags.p)
                      : luna::launch_geometry_cells<2>(s, in, fv->cell.p, fv->volume.p, fv->count.p, fv->slice.p, fv->geo_flags.p);
    }));
    LUNA_HIP(timed(fv->prof, s, "geometry_scan", [&] { return luna::launch_geometry_scan(s, fv->slice.p, S, fv->geo_info.p, fv->geo_flags.p); }));
    unsigned long long* back = fv->host + 4;
    LUNA_HIP(hipMemcpyAsync(back, fv->geo_info.p, sizeof(long long), hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipMemcpyAsync(back + 1, fv->geo_flags.p, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipEventRecord(fv->sized, s));
    // The faces' cell indices need neither the lists nor their size: the GPU checks them while the host reads the count.
    LUNA_HIP(timed(fv->prof, s, "geometry_pairs", [&] { return luna::launch_geometry_pairs(s, in, fv->geo_flags.p); }));
    LUNA_HIP(hipEventSynchronize(fv->sized));
    unsigned int flags = static_cast<unsigned int>(back[1]);
    if (flags & luna::GEO_BAD_START) return fail(LUNA_ERR_ARGUMENT, "luna: CSR start must run from 0 to 2 x faces without decreasing");
    if (flags & luna::GEO_TOO_MANY_SLOTS) return fail(LUNA_ERR_ARGUMENT, "luna: mesh lists exceed 2^31 slots");
    const long long slots = static_cast<long long>(back[0]);
    LUNA_HIP(fv->list.reserve(slots > 0 ? slots : 1));
    LUNA_HIP(fv->other.reserve(slots > 0 ? slots : 1));
    LUNA_HIP(fv->slotgeo.reserve(static_cast<size_t>(D + 1) * (slots > 0 ? slots : 1)));
    LUNA_HIP(timed(fv->prof, s, "geometry_slots", [&] {
        return D == 3 ? luna::launch_geometry_slots<3>(s, in, fv->count.p, fv->slice.p, fv->geo_info.p, fv->list.p, fv->other.p, fv->slotgeo.p,
                                                       fv->geo_flags.p)
                      : luna::launch_geometry_slots<2>(s, in, fv->count.p, fv->slice.p, fv->geo_info.p, fv->list.p, fv->other.p, fv->slotgeo.p,
                                                       fv->geo_flags.p);
    }));
    LUNA_HIP(hipMemcpyAsync(back + 1, fv->geo_flags.p, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipStreamSynchronize(s));
    flags = static_cast<unsigned int>(back[1]);
    if (flags & luna::GEO_BAD_CELL) return fail(LUNA_ERR_ARGUMENT, "luna: face cell index out of range");
    if (flags & luna::GEO_BAD_FACE) return fail(LUNA_ERR_ARGUMENT, "luna: list entry out of range or not a face of its cell");
    fv->F = F;
    fv->E = 2 * F;
    fv->slots = static_cast<int>(slots);
    fv->mesh = in;  // the face arrays only are read from here on
    fv->mesh.start = fv->mesh.list = nullptr;
    fv->mesh.centroid = fv->mesh.volume = fv->mesh.positions = nullptr;
    fv->records = false;
    fv->mesh_ok = true;
    if (eager_records)
        if (int rc = records(fv)) return rc;
    for (int a = 0; a < 3; ++a) fv->side[a] = a < D ? m->side[a] : 0.0;
    fv->grid.step = m->step;
    fv->grid.per_step = m->step > 0 ? 1.0 / m->step : 0.0;
    for (int a = 0; a < 3; ++a) fv->grid.period[a] = a < D ? m->period[a] : 1;
    return LUNA_OK;
}

luna::MeshIn mesh_in(const luna_geometry* m) {
    luna::MeshIn in{};
    in.N = m->cells;
    in.F = m->faces;
    in.owner = m->owner;
    in.neighbour = m->neighbour;
    in.start = m->start;
    in.list = m->list;
    in.area = m->area;
    in.face_centroid = m->face_centroid;
    in.centroid = m->centroid;
    in.volume = m->volume;
    in.positions = m->positions;
    return in;
}

// A mesh in device memory (the arrays the producer left on the GPU), read in place.
int upload_device(luna_fv* fv, const luna_geometry* m) {
    if (int rc = check_geometry(m, fv->D, fv->N)) return rc;
    return derive(fv, m, mesh_in(m), false);
}

// A mesh in host memory: its arrays are copied to the handle's device buffers, then derived as a device mesh. Host
// checks first, which the copies' sizes rely on: the CSR offsets' range and the face count.
int upload(luna_fv* fv, const luna_geometry* m) {
    const int D = fv->D, N = fv->N;
    if (int rc = check_geometry(m, D, N)) return rc;
    const size_t F = m->faces, E = static_cast<size_t>(m->start[N]);
    if (E != 2 * F) return fail(LUNA_ERR_ARGUMENT, "luna: every face must appear in two lists");
    hipStream_t s = fv->stream;
    LUNA_HIP(fv->src_owner.reserve(F > 0 ? F : 1));
    LUNA_HIP(fv->src_neighbour.reserve(F > 0 ? F : 1));
    LUNA_HIP(fv->src_start.reserve(static_cast<size_t>(N) + 1));
    LUNA_HIP(fv->src_list.reserve(E > 0 ? E : 1));
    LUNA_HIP(fv->src_area.reserve(F * D > 0 ? F * D : 1));
    LUNA_HIP(fv->src_fc.reserve(F * D > 0 ? F * D : 1));
    LUNA_HIP(fv->src_cells.reserve(static_cast<size_t>(N) * (2 * D + 1)));
    double* centroid = fv->src_cells.p;
    double* positions = centroid + static_cast<size_t>(N) * D;
    double* volume = positions + static_cast<size_t>(N) * D;
    LUNA_HIP(hipMemcpyAsync(fv->src_owner.p, m->owner, sizeof(int) * F, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->src_neighbour.p, m->neighbour, sizeof(int) * F, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->src_start.p, m->start, sizeof(int) * (N + 1), hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->src_list.p, m->list, sizeof(int) * E, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->src_area.p, m->area, sizeof(double) * F * D, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->src_fc.p, m->face_centroid, sizeof(double) * F * D, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(centroid, m->centroid, sizeof(double) * N * D, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(positions, m->positions, sizeof(double) * N * D, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(volume, m->volume, sizeof(double) * N, hipMemcpyHostToDevice, s));
    luna::MeshIn in = mesh_in(m);
    in.owner = fv->src_owner.p;
    in.neighbour = fv->src_neighbour.p;
    in.start = fv->src_start.p;
    in.list = fv->src_list.p;
    in.area = fv->src_area.p;
    in.face_centroid = fv->src_fc.p;
    in.centroid = centroid;
    in.volume = volume;
    in.positions = positions;
    return derive(fv, m, in, true);
}

// The derivatives read each face as one 64 B record (k_geo_faces, from the face arrays the step reads in place; the pass
// has checked their cell indices). Built on their first use of a mesh, stream-ordered; a host mesh's at once.
int records(luna_fv* fv) {
    if (fv->records || !fv->mesh_ok) return LUNA_OK;
    hipStream_t s = fv->stream;
    LUNA_HIP(timed(fv->prof, s, "geometry_faces", [&] {
        return fv->D == 3 ? luna::launch_geometry_faces<3>(s, fv->mesh, fv->facerec.p, fv->geo_flags.p)
                          : luna::launch_geometry_faces<2>(s, fv->mesh, fv->facerec.p, fv->geo_flags.p);
    }));
    fv->records = true;
    return LUNA_OK;
}

int primitives(luna_fv* fv) {
    const luna::Geo g = fv->geo();
    const luna::Rows r = fv->rows();
    const luna::Gas gas = fv->gas();
    hipStream_t s = fv->stream;
    LUNA_HIP(timed(fv->prof, s, "primitives", [&] {
        return fv->D == 3 ? luna::launch_primitives<3>(s, g, r, gas) : luna::launch_primitives<2>(s, g, r, gas);
    }));
    return LUNA_OK;
}

// The network copy the kernels read, carved from one device buffer: W1, W2 (BF16), then b1, b2, w3, b3 (FP32).
constexpr size_t NET_W2 = sizeof(unsigned short) * luna::HIDDEN * luna::FEATURES;
constexpr size_t NET_REST = NET_W2 + sizeof(unsigned short) * luna::HIDDEN * luna::HIDDEN;
constexpr size_t NET_BYTES = NET_REST + sizeof(float) * (3 * luna::HIDDEN + 1);

luna::Network network_at(unsigned char* p) {
    const float* rest = reinterpret_cast<const float*>(p + NET_REST);
    return luna::Network{reinterpret_cast<const unsigned short*>(p), reinterpret_cast<const unsigned short*>(p + NET_W2), rest,
                         rest + luna::HIDDEN, rest + 2 * luna::HIDDEN, rest + 3 * luna::HIDDEN};
}

// The network copy of FP32 parameters (luna::P_* order), on the host.
std::vector<unsigned char> network_bytes(const float* params) {
    std::vector<unsigned char> out(NET_BYTES);
    unsigned short* w1 = reinterpret_cast<unsigned short*>(out.data());
    unsigned short* w2 = reinterpret_cast<unsigned short*>(out.data() + NET_W2);
    float* rest = reinterpret_cast<float*>(out.data() + NET_REST);
    for (int k = 0; k < luna::HIDDEN * luna::FEATURES; ++k) w1[k] = luna::bf16_bits(params[luna::P_W1 + k]);
    for (int k = 0; k < luna::HIDDEN * luna::HIDDEN; ++k) w2[k] = luna::bf16_bits(params[luna::P_W2 + k]);
    for (int k = 0; k < luna::HIDDEN; ++k) {
        rest[k] = params[luna::P_B1 + k];
        rest[luna::HIDDEN + k] = params[luna::P_B2 + k];
        rest[2 * luna::HIDDEN + k] = params[luna::P_W3 + k];
    }
    rest[3 * luna::HIDDEN] = params[luna::P_B3];
    return out;
}

int sweep(luna_fv* fv, double courant, bool dt_only = false) {
    const luna::Geo g = fv->geo();
    const luna::Rows r = fv->rows();
    const luna::Gas gas = fv->gas();
    const luna::Counters c = fv->counters_now();
    const bool moving = fv->opt.motion == LUNA_MOTION_MOVING, grad = fv->opt.order == 1 && !dt_only;
    const bool learned = fv->opt.limiter == LUNA_LIMITER_LEARNED;
    const int limit = learned ? 2 : fv->opt.limiter == LUNA_LIMITER_FACE_CENTROID ? 1 : 0;
    hipStream_t s = fv->stream;
    LUNA_HIP(timed(fv->prof, s, "sweep", [&] {
        return fv->D == 3 ? luna::launch_sweep<3>(s, g, r, gas, courant, moving, grad, limit, fv->opt.step_rule, c)
                          : luna::launch_sweep<2>(s, g, r, gas, courant, moving, grad, limit, fv->opt.step_rule, c);
    }));
    fv->parity ^= 1;
    if (learned) {
        fv->features_ready = true;
        if (fv->has_net) {
            const luna::Network net = network_at(fv->net.p);
            LUNA_HIP(timed(fv->prof, s, "limiter", [&] {
                return luna::launch_limiter_apply(s, net, fv->features.p, fv->troubled.p, fv->N * fv->V, fv->V, fv->D, fv->grad_row(),
                                                  fv->grad.p, nullptr);
            }));
        }
    }
    return LUNA_OK;
}

int ensure_swept(luna_fv* fv) {
    if (!fv->mesh_ok) return fail(LUNA_ERR_LOGIC, "luna: the last mesh given was refused; set a valid mesh first");
    if (fv->swept) return LUNA_OK;
    if (!fv->prims_ready)
        if (int rc = primitives(fv)) return rc;
    fv->prims_ready = false;
    return sweep(fv, 1.0);
}

// One flux pass and its update; returns through *prims whether the update wrote the next primitives (the per-face loop
// with next_prims), else the next time_step computes them.
int flux_and_update(luna_fv* fv, int pass, double flux_dt, double apply_dt, double unit, bool next_prims, bool* prims = nullptr) {
    const luna::Grid* drift = pass == 0 ? &fv->grid : nullptr;
    const luna::Geo g = fv->geo();
    const luna::Rows r = fv->rows();
    const luna::Gas gas = fv->gas();
    hipStream_t s = fv->stream;
    unsigned long long* count = fv->counters.p + (pass == 1 ? 3 : 2);
    const int riemann = fv->riemann(), order = fv->opt.order;
    if (prims) *prims = next_prims && fv->opt.update != LUNA_UPDATE_PER_CELL;
    if (fv->opt.update == LUNA_UPDATE_PER_CELL) {
        if (int rc = records(fv)) return rc;  // the per-cell loop reads each face's record
        LUNA_HIP(timed(fv->prof, s, pass < 0 || pass == 2 ? "fluxes_and_update" : pass == 0 ? "fluxes_and_update_pass1" : "fluxes_and_update_pass2", [&] {
            return fv->D == 3 ? luna::launch_fluxes_and_update<3>(s, g, r, gas, riemann, order, pass, flux_dt, apply_dt, unit, drift, count)
                              : luna::launch_fluxes_and_update<2>(s, g, r, gas, riemann, order, pass, flux_dt, apply_dt, unit, drift, count);
        }));
        return LUNA_OK;
    }
    LUNA_HIP(timed(fv->prof, s, pass < 0 || pass == 2 ? "flux" : pass == 0 ? "flux_pass1" : "flux_pass2", [&] {
        return fv->D == 3 ? luna::launch_flux<3>(s, g, r, gas, riemann, order, pass, flux_dt, unit, count)
                          : luna::launch_flux<2>(s, g, r, gas, riemann, order, pass, flux_dt, unit, count);
    }));
    LUNA_HIP(timed(fv->prof, s, pass < 0 || pass == 2 ? "update" : pass == 0 ? "update_pass1" : "update_pass2", [&] {
        return fv->D == 3 ? luna::launch_update<3>(s, g, r, gas, apply_dt, next_prims, drift)
                          : luna::launch_update<2>(s, g, r, gas, apply_dt, next_prims, drift);
    }));
    return LUNA_OK;
}

}  // namespace

extern "C" {

const char* luna_last_error(void) { return g_error.c_str(); }

}  // extern "C"

namespace {

int create(const luna_fv_options* opt, const luna_geometry* mesh, luna_fv** out, bool device) {
    if (!opt || !mesh || !out) return fail(LUNA_ERR_ARGUMENT, "luna_fv_create: null argument");
    *out = nullptr;
    if (opt->dim != 2 && opt->dim != 3) return fail(LUNA_ERR_ARGUMENT, "hydro::FiniteVolume: D = 2 or 3 (1D: luna_fv1_create)");
    if (opt->order != 0 && opt->order != 1) return fail(LUNA_ERR_ARGUMENT, "hydro::FiniteVolume: order 0 or 1");
    if (opt->limiter != LUNA_LIMITER_FACE_CENTROID && opt->limiter != LUNA_LIMITER_NONE && opt->limiter != LUNA_LIMITER_LEARNED)
        return fail(LUNA_ERR_ARGUMENT, "hydro::FiniteVolume: limiter face_centroid, none or learned in 2D/3D");
    if (opt->limiter == LUNA_LIMITER_LEARNED && opt->order != 1)
        return fail(LUNA_ERR_ARGUMENT, "luna: the learned limiter needs order 1");
    if (opt->riemann != LUNA_RIEMANN_HLLC && opt->riemann != LUNA_RIEMANN_HLL) return fail(LUNA_ERR_ARGUMENT, "luna: unknown Riemann solver");
    if (opt->motion != LUNA_MOTION_STATIC && opt->motion != LUNA_MOTION_MOVING) return fail(LUNA_ERR_ARGUMENT, "luna: unknown motion");
    if (opt->step_rule == LUNA_STEP_LUNA && opt->dim != 2) return fail(LUNA_ERR_ARGUMENT, "StepRule::luna: 2D");
    if (opt->precision < LUNA_PRECISION_DOUBLE || opt->precision > LUNA_PRECISION_GRAD16)
        return fail(LUNA_ERR_ARGUMENT, "hydro::Precision: double, grad32 or grad16");
    if (opt->precision != LUNA_PRECISION_DOUBLE && opt->limiter == LUNA_LIMITER_LEARNED)
        return fail(LUNA_ERR_ARGUMENT, "luna: the learned limiter's features need the double policy");
    if (!(opt->hll_compression >= 0) || std::isinf(opt->hll_compression))
        return fail(LUNA_ERR_ARGUMENT, "luna: hll_compression must be finite and not negative (0: off)");
    if (opt->hll_compression > 0 &&
        (opt->order != 1 || opt->precision != LUNA_PRECISION_DOUBLE || opt->limiter == LUNA_LIMITER_LEARNED))
        return fail(LUNA_ERR_ARGUMENT, "luna: the HLL fallback needs order 1, the double policy and the face-centroid limiter or none");
    if (!(opt->gamma > 1)) return fail(LUNA_ERR_ARGUMENT, "luna: gamma must exceed 1");
    if (opt->update != LUNA_UPDATE_PER_FACE && opt->update != LUNA_UPDATE_PER_CELL)
        return fail(LUNA_ERR_ARGUMENT, "hydro::Update: per_face or per_cell");
    if (opt->update == LUNA_UPDATE_PER_CELL && opt->order != 1)
        return fail(LUNA_ERR_ARGUMENT, "luna: the per-cell update here steps order 1 (the order-0 fluxes read Q, which it updates in place)");
    if (opt->sweep_lanes != 0 && opt->sweep_lanes != 1 && opt->sweep_lanes != 2 && opt->sweep_lanes != 4 && opt->sweep_lanes != 8)
        return fail(LUNA_ERR_ARGUMENT, "LUNA_SWEEP_LANES: 1 (the one-lane sweep), 2, 4 or 8");
    if (opt->sweep_lanes > 1 && (opt->order != 1 || opt->precision != LUNA_PRECISION_DOUBLE || opt->limiter == LUNA_LIMITER_LEARNED))
        return fail(LUNA_ERR_ARGUMENT, "LaneSweep: order 1, the double policy and the face-centroid limiter or none");
    if (mesh->cells <= 0) return fail(LUNA_ERR_ARGUMENT, "luna: no cells");
    luna_fv* fv = new luna_fv;
    fv->opt = *opt;
    fv->D = opt->dim;
    fv->V = opt->dim + 2;
    fv->N = mesh->cells;
    auto cleanup = [&](int rc) {
        luna_fv_destroy(fv);
        return rc;
    };
    if (hipStreamCreate(&fv->stream) != hipSuccess) return cleanup(fail(LUNA_ERR_DEVICE, "luna: hipStreamCreate failed"));
    if (hipEventCreateWithFlags(&fv->sized, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(LUNA_ERR_DEVICE, "luna: hipEventCreate failed"));
    const size_t N = fv->N, D = fv->D, V = fv->V;
    hipError_t e = hipSuccess;
    if ((e = fv->q.reserve(N * V)) != hipSuccess || (e = fv->count.reserve(N)) != hipSuccess ||
        (e = fv->grad.reserve(N * fv->grad_row())) != hipSuccess || (e = fv->w.reserve(N * D)) != hipSuccess ||
        (e = fv->cell.reserve(N * luna::CELL_ROW)) != hipSuccess || (e = fv->volume.reserve(N)) != hipSuccess ||
        (e = fv->counters.reserve(4)) != hipSuccess || (e = fv->geo_info.reserve(1)) != hipSuccess ||
        (e = fv->geo_flags.reserve(1)) != hipSuccess)
        return cleanup(fail(LUNA_ERR_MEMORY, std::string("luna: device allocation: ") + hipGetErrorString(e)));
    if (opt->limiter == LUNA_LIMITER_LEARNED) {
        if ((e = fv->features.reserve(N * V * luna::FEATURES)) != hipSuccess || (e = fv->troubled.reserve(N)) != hipSuccess ||
            (e = fv->net.reserve(NET_BYTES)) != hipSuccess)
            return cleanup(fail(LUNA_ERR_MEMORY, std::string("luna: device allocation: ") + hipGetErrorString(e)));
    }
    if (opt->precision == LUNA_PRECISION_GRAD16 && (e = fv->exponents.reserve(N)) != hipSuccess)
        return cleanup(fail(LUNA_ERR_MEMORY, std::string("luna: device allocation: ") + hipGetErrorString(e)));
    if (opt->hll_compression > 0 && (e = fv->hll.reserve(N)) != hipSuccess)
        return cleanup(fail(LUNA_ERR_MEMORY, std::string("luna: device allocation: ") + hipGetErrorString(e)));
    if (opt->motion == LUNA_MOTION_MOVING) {
        if ((e = fv->x.reserve(N * D)) != hipSuccess || (e = fv->old_centroid.reserve(N * D)) != hipSuccess ||
            (e = fv->steps.reserve(N * D)) != hipSuccess)
            return cleanup(fail(LUNA_ERR_MEMORY, std::string("luna: device allocation: ") + hipGetErrorString(e)));
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&fv->host), 8 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(LUNA_ERR_MEMORY, "luna: pinned host allocation failed"));
    const unsigned long long init[4] = {luna::INF_BITS, luna::INF_BITS, 0, 0};
    if (hipMemcpy(fv->counters.p, init, sizeof init, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(LUNA_ERR_DEVICE, "luna: counter initialisation failed"));
    std::vector<double> zero(N * V, 0.0);
    if (hipMemcpy(fv->q.p, zero.data(), sizeof(double) * N * V, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(LUNA_ERR_DEVICE, "luna: state initialisation failed"));
    if (int rc = device ? upload_device(fv, mesh) : upload(fv, mesh)) return cleanup(rc);
    if (opt->motion == LUNA_MOTION_MOVING) {  // copy_positions (finite_volume.hpp:342-349)
        if (hipMemcpy(fv->x.p, mesh->positions, sizeof(double) * N * D, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(LUNA_ERR_DEVICE, "luna: position upload failed"));
        std::vector<double> zw(N * D, 0.0);
        if (hipMemcpy(fv->w.p, zw.data(), sizeof(double) * N * D, hipMemcpyHostToDevice) != hipSuccess)
            return cleanup(fail(LUNA_ERR_DEVICE, "luna: velocity initialisation failed"));
    }
    *out = fv;
    return LUNA_OK;
}

}  // namespace

extern "C" {

int luna_fv_create(const luna_fv_options* opt, const luna_geometry* mesh, luna_fv** out) { return create(opt, mesh, out, false); }

int luna_fv_create_device(const luna_fv_options* opt, const luna_geometry* mesh, luna_fv** out) { return create(opt, mesh, out, true); }

int luna_fv_destroy(luna_fv* fv) {
    if (!fv) return LUNA_OK;
    if (fv->stream) (void)hipStreamSynchronize(fv->stream);
    fv->prof.destroy();
    for (auto* b : {&fv->cell, &fv->volume, &fv->facerec, &fv->slotgeo, &fv->q, &fv->grad, &fv->w, &fv->x, &fv->old_centroid, &fv->flux}) b->release();
    for (auto* b : {&fv->count, &fv->slice, &fv->list, &fv->other, &fv->steps, &fv->d_extreme}) b->release();
    fv->exponents.release();
    fv->hll.release();
    for (auto* b : {&fv->d_dq, &fv->d_dW, &fv->d_dgrad, &fv->d_dflux, &fv->d_sidebar, &fv->d_bbar, &fv->d_boundbar, &fv->d_out}) b->release();
    for (auto* b : {&fv->m_row, &fv->m_Ab, &fv->m_Xfb, &fv->m_Wbar, &fv->m_gbar, &fv->m_whbar, &fv->m_c0bar, &fv->m_cbar, &fv->m_xbar,
                    &fv->m_beta, &fv->m_boundbar, &fv->m_entry, &fv->m_angle, &fv->m_Vbar, &fv->m_out, &fv->m_qbar})
        b->release();
    fv->m_extreme.release();
    fv->m_fmax.release();
    for (auto* b : {&fv->t_dx, &fv->t_dx1, &fv->t_dV, &fv->t_dc, &fv->t_dc0, &fv->t_dA, &fv->t_dxf, &fv->t_dW, &fv->t_dcs, &fv->t_dw, &fv->t_dg,
                    &fv->t_dG, &fv->t_dq})
        b->release();
    for (auto* b : {&fv->src_owner, &fv->src_neighbour, &fv->src_start, &fv->src_list}) b->release();
    for (auto* b : {&fv->src_area, &fv->src_fc, &fv->src_cells}) b->release();
    fv->geo_info.release();
    fv->geo_flags.release();
    fv->counters.release();
    fv->features.release();
    fv->troubled.release();
    fv->net.release();
    if (fv->host) (void)hipHostFree(fv->host);
    if (fv->sized) (void)hipEventDestroy(fv->sized);
    if (fv->stream) (void)hipStreamDestroy(fv->stream);
    delete fv;
    return LUNA_OK;
}

int luna_fv_set_state(luna_fv* fv, const double* q) {
    if (!fv || !q) return fail(LUNA_ERR_ARGUMENT, "luna_fv_set_state: null argument");
    LUNA_HIP(hipMemcpyAsync(fv->q.p, q, sizeof(double) * fv->N * fv->V, hipMemcpyHostToDevice, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    fv->prims_ready = false;
    fv->swept = false;
    return LUNA_OK;
}

int luna_fv_get_state(luna_fv* fv, double* q) {
    if (!fv || !q) return fail(LUNA_ERR_ARGUMENT, "luna_fv_get_state: null argument");
    LUNA_HIP(hipMemcpyAsync(q, fv->q.p, sizeof(double) * fv->N * fv->V, hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

int luna_fv_state_device(luna_fv* fv, double** q) {
    if (!fv || !q) return fail(LUNA_ERR_ARGUMENT, "luna_fv_state_device: null argument");
    *q = fv->q.p;
    fv->prims_ready = false;
    fv->swept = false;
    return LUNA_OK;
}

int luna_fv_time_step(luna_fv* fv, double courant, double* dt) {
    if (!fv || !dt) return fail(LUNA_ERR_ARGUMENT, "luna_fv_time_step: null argument");
    if (fv->begun) return fail(LUNA_ERR_LOGIC, "luna_fv_time_step: a moving step is between begin and finish");
    if (!fv->mesh_ok) return fail(LUNA_ERR_LOGIC, "luna: the last mesh given was refused; set a valid mesh first");
    if (!fv->prims_ready)
        if (int rc = primitives(fv)) return rc;
    fv->prims_ready = false;
    const int slot = fv->parity;
    if (int rc = sweep(fv, courant)) return rc;
    LUNA_HIP(hipMemcpyAsync(fv->host, fv->counters.p + slot, sizeof(unsigned long long), hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    LUNA_HIP(fv->prof.collect());
    *dt = bits_to_double(fv->host[0]);
    fv->swept = true;
    fv->tangent_swept = false;
    return LUNA_OK;
}

int luna_fv_step(luna_fv* fv, double dt) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv_step: null argument");
    if (fv->opt.motion != LUNA_MOTION_STATIC)
        return fail(LUNA_ERR_LOGIC, "luna_fv_step: a moving mesh steps with luna_fv_moving_begin / luna_fv_moving_finish");
    if (int rc = ensure_swept(fv)) return rc;
    fv->swept = false;
    bool prims = false;
    if (int rc = flux_and_update(fv, -1, dt, dt, 0.0, true, &prims)) return rc;
    fv->prims_ready = prims;
    return LUNA_OK;
}

namespace {

int derivatives_ready(luna_fv* fv, const char* what) {
    if (fv->opt.motion != LUNA_MOTION_STATIC || fv->opt.order != 1)
        return fail(LUNA_ERR_LOGIC, std::string(what) + ": derivatives of a static order-1 step");
    if (fv->opt.limiter == LUNA_LIMITER_LEARNED)
        return fail(LUNA_ERR_LOGIC, std::string(what) + ": derivatives of the face-centroid limiter or none (not the learned one)");
    if (fv->opt.precision != LUNA_PRECISION_DOUBLE) return fail(LUNA_ERR_LOGIC, std::string(what) + ": derivatives of the double policy");
    if (fv->opt.hll_compression > 0) return fail(LUNA_ERR_LOGIC, std::string(what) + ": derivatives without the HLL fallback");
    if (fv->opt.sweep_lanes > 1) return fail(LUNA_ERR_LOGIC, std::string(what) + ": derivatives of the one-lane sweep");
    if (fv->begun) return fail(LUNA_ERR_LOGIC, std::string(what) + ": a moving step is between begin and finish");
    if (!fv->mesh_ok) return fail(LUNA_ERR_LOGIC, "luna: the last mesh given was refused; set a valid mesh first");
    if (int rc = records(fv)) return rc;
    const size_t N = fv->N, V = fv->V, F = fv->F;
    LUNA_HIP(fv->d_dq.reserve(N * V));
    LUNA_HIP(fv->d_dW.reserve(N * V));
    LUNA_HIP(fv->d_dgrad.reserve(N * fv->grad_row()));
    LUNA_HIP(fv->d_dflux.reserve(2 * F * V > 0 ? 2 * F * V : 1));  // each face's flux and its tangent
    LUNA_HIP(fv->d_sidebar.reserve(2 * F * V > 0 ? 2 * F * V : 1));
    LUNA_HIP(fv->d_bbar.reserve(N * V * (fv->D + 1)));
    LUNA_HIP(fv->d_boundbar.reserve(N * V));
    LUNA_HIP(fv->d_extreme.reserve(N * V));
    LUNA_HIP(fv->d_out.reserve(N * V));
    return LUNA_OK;
}

luna::Derivs derivs_of(luna_fv* fv) {
    return luna::Derivs{fv->d_dq.p, fv->d_dW.p, fv->d_dgrad.p, fv->d_dflux.p, fv->d_sidebar.p, fv->d_bbar.p, fv->d_boundbar.p,
                        fv->d_extreme.p, fv->d_out.p};
}

}  // namespace

int luna_fv_step_tangent(luna_fv* fv, double dt, double* dq, int device) {
    if (!fv || !dq) return fail(LUNA_ERR_ARGUMENT, "luna_fv_step_tangent: null argument");
    if (int rc = derivatives_ready(fv, "luna_fv_step_tangent")) return rc;
    hipStream_t s = fv->stream;
    const size_t bytes = sizeof(double) * fv->N * fv->V;
    LUNA_HIP(hipMemcpyAsync(fv->d_dq.p, dq, bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    const luna::Geo g = fv->geo();
    const luna::Rows r = fv->rows();
    const luna::Gas gas = fv->gas();
    const luna::Derivs w = derivs_of(fv);
    const bool limit = fv->opt.limiter == LUNA_LIMITER_FACE_CENTROID;
    const bool fitted = fv->swept && fv->tangent_swept;  // time_step_tangent's rows (the reference's contract: the same dq)
    LUNA_HIP(timed(fv->prof, s, "step_tangent", [&] {
        return fv->D == 3 ? luna::launch_step_tangent<3>(s, g, r, gas, fv->opt.riemann, dt, limit, w, fitted)
                          : luna::launch_step_tangent<2>(s, g, r, gas, fv->opt.riemann, dt, limit, w, fitted);
    }));
    fv->tangent_swept = false;
    LUNA_HIP(hipMemcpyAsync(dq, fv->d_dq.p, bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipStreamSynchronize(s));
    fv->prims_ready = false;
    fv->swept = false;
    return LUNA_OK;
}

int luna_fv_time_step_tangent(luna_fv* fv, double courant, const double* dq, int device, double* dt) {
    if (!fv || !dq || !dt) return fail(LUNA_ERR_ARGUMENT, "luna_fv_time_step_tangent: null argument");
    if (int rc = derivatives_ready(fv, "luna_fv_time_step_tangent")) return rc;
    // As the reference's (its sweep with the tangent hook): the primitives and their tangents, the limited fit and its
    // tangent (k_fit_tangent writes the gradients the step uses and their tangents), and dt from a sweep of the time
    // step's terms alone (the same terms, so the same dt as time_step's). The next step_tangent with the same dq reuses
    // the rows.
    hipStream_t s = fv->stream;
    LUNA_HIP(hipMemcpyAsync(fv->d_dq.p, dq, sizeof(double) * fv->N * fv->V, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    if (!fv->prims_ready)
        if (int rc = primitives(fv)) return rc;
    fv->prims_ready = false;
    const int slot = fv->parity;
    if (int rc = sweep(fv, courant, true)) return rc;
    const luna::Geo g = fv->geo();
    const luna::Rows r = fv->rows();
    const luna::Gas gas = fv->gas();
    const luna::Derivs w = derivs_of(fv);
    const bool limit = fv->opt.limiter == LUNA_LIMITER_FACE_CENTROID;
    LUNA_HIP(timed(fv->prof, s, "sweep_tangent", [&] {
        return fv->D == 3 ? luna::launch_sweep_tangent<3>(s, g, r, gas, limit, w) : luna::launch_sweep_tangent<2>(s, g, r, gas, limit, w);
    }));
    LUNA_HIP(hipMemcpyAsync(fv->host, fv->counters.p + slot, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipStreamSynchronize(s));
    LUNA_HIP(fv->prof.collect());
    *dt = bits_to_double(fv->host[0]);
    fv->swept = true;
    fv->tangent_swept = true;
    return LUNA_OK;
}

int luna_fv_step_adjoint(luna_fv* fv, double dt, double* qbar, int device) {
    if (!fv || !qbar) return fail(LUNA_ERR_ARGUMENT, "luna_fv_step_adjoint: null argument");
    if (int rc = derivatives_ready(fv, "luna_fv_step_adjoint")) return rc;
    hipStream_t s = fv->stream;
    const size_t bytes = sizeof(double) * fv->N * fv->V;
    LUNA_HIP(hipMemcpyAsync(fv->d_dq.p, qbar, bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    if (int rc = primitives(fv)) return rc;  // the forward's rows at Q: primitives and the limited gradients
    if (int rc = sweep(fv, 1.0)) return rc;
    const luna::Geo g = fv->geo();
    const luna::Rows r = fv->rows();
    const luna::Gas gas = fv->gas();
    const luna::Derivs w = derivs_of(fv);
    const bool limit = fv->opt.limiter == LUNA_LIMITER_FACE_CENTROID;
    LUNA_HIP(timed(fv->prof, s, "step_adjoint", [&] {
        return fv->D == 3 ? luna::launch_step_adjoint<3>(s, g, r, gas, fv->opt.riemann, dt, limit, w)
                          : luna::launch_step_adjoint<2>(s, g, r, gas, fv->opt.riemann, dt, limit, w);
    }));
    LUNA_HIP(hipMemcpyAsync(qbar, fv->d_out.p, bytes, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipStreamSynchronize(s));
    fv->prims_ready = false;
    fv->swept = false;
    return LUNA_OK;
}

}  // extern "C"

namespace {

// The work rows of luna_fv_step_adjoint_noalloc: a file-scope device array (each GPU holds its own copy from the library's
// load until the process ends), carved per call into the adjoint's rows. Its device address comes from
// hipGetSymbolAddress once per device (the array's name on the host is only a stand-in, never a device address).
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_lock;
void* g_scratch[MAX_DEVICES] = {};

// The scratch array's address on the current device (looked up on the first call there).
int scratch_address(void** out) {
    int device = 0;
    LUNA_HIP(hipGetDevice(&device));
    if (device < 0 || device >= MAX_DEVICES) return fail(LUNA_ERR_DEVICE, "luna: device index beyond the scratch table");
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(LUNA_ERR_DEVICE, std::string("luna: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return LUNA_OK;
}

// The adjoint's rows carved from [base, base + SCRATCH_BYTES), each 256-byte aligned; false if they do not fit.
bool carve_adjoint(const luna_fv* fv, unsigned char* base, luna::Derivs& w, size_t& need) {
    const size_t N = fv->N, V = fv->V, F = fv->F, U = fv->D + 1;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t here = at;
        at += (bytes + 255) / 256 * 256;
        return here;
    };
    const size_t o_sidebar = take(sizeof(double) * 2 * F * V), o_dW = take(sizeof(double) * N * V), o_bbar = take(sizeof(double) * N * V * U),
                 o_bound = take(sizeof(double) * N * V), o_ext = take(sizeof(int) * N * V), o_out = take(sizeof(double) * N * V);
    need = at;
    if (need > SCRATCH_BYTES) return false;
    w = luna::Derivs{nullptr, reinterpret_cast<double*>(base + o_dW), nullptr, nullptr, reinterpret_cast<double*>(base + o_sidebar),
                     reinterpret_cast<double*>(base + o_bbar), reinterpret_cast<double*>(base + o_bound), reinterpret_cast<int*>(base + o_ext),
                     reinterpret_cast<double*>(base + o_out)};
    return true;
}

}  // namespace

extern "C" {

int luna_fv_step_adjoint_noalloc(luna_fv* fv, double dt, double* qbar) {
    if (!fv || !qbar) return fail(LUNA_ERR_ARGUMENT, "luna_fv_step_adjoint_noalloc: null argument");
    if (fv->opt.motion != LUNA_MOTION_STATIC || fv->opt.order != 1)
        return fail(LUNA_ERR_LOGIC, "luna_fv_step_adjoint_noalloc: derivatives of a static order-1 step");
    if (fv->opt.limiter == LUNA_LIMITER_LEARNED || fv->opt.precision != LUNA_PRECISION_DOUBLE || fv->opt.hll_compression > 0 ||
        fv->opt.sweep_lanes > 1)
        return fail(LUNA_ERR_LOGIC,
                    "luna_fv_step_adjoint_noalloc: the double policy, the face-centroid limiter or none, no HLL fallback, the one-lane sweep");
    if (fv->begun) return fail(LUNA_ERR_LOGIC, "luna_fv_step_adjoint_noalloc: a moving step is between begin and finish");
    if (!fv->mesh_ok) return fail(LUNA_ERR_LOGIC, "luna: the last mesh given was refused; set a valid mesh first");
    void* base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    luna::Derivs w{};
    size_t need = 0;
    if (!carve_adjoint(fv, static_cast<unsigned char*>(base), w, need))
        return fail(LUNA_ERR_ARGUMENT, "luna_fv_step_adjoint_noalloc: this mesh needs " + std::to_string(need) + " bytes of work rows, beyond the " +
                                           std::to_string(SCRATCH_BYTES) + " of the library's scratch (use luna_fv_step_adjoint)");
    w.dq = qbar;  // read only: the cotangent after the step
    hipStream_t s = fv->stream;
    if (int rc = records(fv)) return rc;  // launches only: the records' memory is the handle's, reserved with the mesh
    if (int rc = primitives(fv)) return rc;  // the forward's rows at Q: primitives and the limited gradients
    if (int rc = sweep(fv, 1.0)) return rc;
    const luna::Geo g = fv->geo();
    const luna::Rows r = fv->rows();
    const luna::Gas gas = fv->gas();
    const bool limit = fv->opt.limiter == LUNA_LIMITER_FACE_CENTROID;
    LUNA_HIP(timed(fv->prof, s, "step_adjoint", [&] {
        return fv->D == 3 ? luna::launch_step_adjoint<3>(s, g, r, gas, fv->opt.riemann, dt, limit, w)
                          : luna::launch_step_adjoint<2>(s, g, r, gas, fv->opt.riemann, dt, limit, w);
    }));
    LUNA_HIP(hipMemcpyAsync(qbar, w.out, sizeof(double) * fv->N * fv->V, hipMemcpyDeviceToDevice, s));
    LUNA_HIP(hipStreamSynchronize(s));
    fv->prims_ready = false;
    fv->swept = false;
    return LUNA_OK;
}

int luna_fv_moving_adjoint(luna_fv* fv, double dt, const luna_geometry* mesh0, const luna_geometry* mesh1, luna_geometry_adjoint_fn adjoint0,
                           const void* ctx0, luna_geometry_adjoint_fn adjoint1, const void* ctx1, int steps_fixed, double* qbar, double* xbar) {
    if (!fv || !mesh0 || !mesh1 || !adjoint0 || !adjoint1 || !qbar || !xbar) return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_adjoint: null argument");
    if (fv->opt.motion != LUNA_MOTION_MOVING) return fail(LUNA_ERR_LOGIC, "hydro::FiniteVolume::step_moving_adjoint: a moving mesh");
    if (fv->opt.limiter == LUNA_LIMITER_LEARNED)
        return fail(LUNA_ERR_LOGIC, "luna_fv_moving_adjoint: derivatives of the face-centroid limiter or none (not the learned one)");
    if (fv->opt.precision != LUNA_PRECISION_DOUBLE) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_adjoint: derivatives of the double policy");
    if (fv->opt.hll_compression > 0) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_adjoint: derivatives without the HLL fallback");
    if (fv->opt.sweep_lanes > 1) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_adjoint: derivatives of the one-lane sweep");
    if (fv->begun) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_adjoint: a moving step is between begin and finish");
    if (!(dt > 0)) return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_adjoint: dt must be positive");
    if (mesh0->cells != fv->N || mesh1->cells != fv->N) return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_adjoint: both meshes must have the handle's cells");
    const int D = fv->D, V = fv->V, N = fv->N, U = D + 1;
    const size_t ND = static_cast<size_t>(N) * D, NV = static_cast<size_t>(N) * V;
    hipStream_t s = fv->stream;
    const bool limit = fv->opt.limiter == LUNA_LIMITER_FACE_CENTROID;
    const int row_width = D == 3 ? 2 * (5 + 15 + 3) + 9 : 2 * (4 + 8 + 2) + 6;
    // t^n: its mesh and positions, W, the limited gradients and point velocities (the sweep keeps c^n), the grid steps.
    if (int rc = luna_fv_set_geometry(fv, mesh0)) return rc;  // and x^n
    if (int rc = primitives(fv)) return rc;
    if (int rc = sweep(fv, 1.0)) return rc;
    if (fv->opt.order == 1) {
        const luna::Geo g = fv->geo();
        const luna::Rows r = fv->rows();
        LUNA_HIP(D == 3 ? luna::launch_keep_centroids<3>(s, g, r, fv->grid, dt) : luna::launch_keep_centroids<2>(s, g, r, fv->grid, dt));
    }
    const double unit = fv->grid.step / dt;
    // work arrays (zeroed where summed into)
    const size_t Fmax = static_cast<size_t>(std::max(std::max(mesh0->faces, mesh1->faces), 1));
    LUNA_HIP(fv->m_row.reserve(Fmax * row_width));
    LUNA_HIP(fv->m_Ab.reserve(Fmax * D));
    LUNA_HIP(fv->m_Xfb.reserve(Fmax * D));
    for (Buf<double>* b : {&fv->m_whbar, &fv->m_c0bar, &fv->m_cbar, &fv->m_xbar, &fv->m_angle}) LUNA_HIP(b->reserve(2 * ND));
    LUNA_HIP(fv->m_Wbar.reserve(NV));
    LUNA_HIP(fv->m_gbar.reserve(NV * D));
    LUNA_HIP(fv->m_beta.reserve(NV * U));
    LUNA_HIP(fv->m_boundbar.reserve(NV));
    LUNA_HIP(fv->m_Vbar.reserve(N));
    LUNA_HIP(fv->m_out.reserve(NV));
    LUNA_HIP(fv->m_qbar.reserve(NV));
    LUNA_HIP(fv->m_extreme.reserve(NV));
    LUNA_HIP(fv->m_fmax.reserve(N));
    LUNA_HIP(hipMemsetAsync(fv->m_Wbar.p, 0, sizeof(double) * NV, s));
    LUNA_HIP(hipMemsetAsync(fv->m_gbar.p, 0, sizeof(double) * NV * D, s));
    LUNA_HIP(hipMemsetAsync(fv->m_whbar.p, 0, sizeof(double) * ND, s));
    LUNA_HIP(hipMemsetAsync(fv->m_c0bar.p, 0, sizeof(double) * ND, s));
    LUNA_HIP(hipMemsetAsync(fv->m_cbar.p, 0, sizeof(double) * ND, s));
    LUNA_HIP(hipMemcpyAsync(fv->m_qbar.p, qbar, sizeof(double) * NV, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->m_xbar.p, xbar, sizeof(double) * ND, hipMemcpyHostToDevice, s));
    auto bars = [&]() {
        return luna::MovingBars{fv->m_row.p, fv->m_Ab.p, fv->m_Xfb.p, fv->m_Wbar.p, fv->m_gbar.p, fv->m_whbar.p, fv->m_c0bar.p, fv->m_cbar.p,
                                fv->m_xbar.p, fv->m_beta.p, fv->m_boundbar.p, nullptr, fv->m_angle.p, fv->m_Vbar.p, fv->m_out.p,
                                fv->m_extreme.p, fv->m_fmax.p};
    };
    const luna::Gas gas = fv->gas();
    if (fv->opt.order == 0) {
        // step_moving_adjoint at order 0 (finite_volume.hpp:2490-2598): the one pass on t^n (sides W, the point velocities),
        // x^n+1 = x^n + dt w (w's bar takes dt of x^n+1's, x^n's is x^n+1's), then the point velocity's and the sound
        // speed's adjoints and W = primitive(Q / V) per cell (zero gradient bars: the fit adds nothing).
        LUNA_HIP(fv->m_entry.reserve(static_cast<size_t>(fv->slots > 0 ? fv->slots : 1) * (2 * D + 1)));
        const luna::Geo g = fv->geo();
        const luna::Rows r = fv->rows();
        luna::MovingBars b = bars();
        b.entry = fv->m_entry.p;
        // w's bar from x^n+1's first (m_xbar holds the caller's bar), then the pass adds the separations' bars to x^n's
        LUNA_HIP(luna::launch_drift_adjoint(s, static_cast<long long>(ND), dt, fv->m_xbar.p, fv->m_whbar.p));
        LUNA_HIP(D == 3 ? luna::launch_moving_pass_adjoint<3>(s, g, r, gas, fv->opt.riemann, dt, unit, false, fv->m_qbar.p, b, true)
                        : luna::launch_moving_pass_adjoint<2>(s, g, r, gas, fv->opt.riemann, dt, unit, false, fv->m_qbar.p, b, true));
        LUNA_HIP(D == 3 ? luna::launch_moving_cell_adjoint<3>(s, g, r, gas, false, false, fv->m_qbar.p, b)
                        : luna::launch_moving_cell_adjoint<2>(s, g, r, gas, false, false, fv->m_qbar.p, b));
        const size_t F0 = static_cast<size_t>(mesh0->faces);
        std::vector<double> hV(N, 0.0), hc(ND), hA(Fmax * D), hxf(Fmax * D), hx(ND);
        LUNA_HIP(hipMemcpyAsync(hV.data(), fv->m_Vbar.p, sizeof(double) * N, hipMemcpyDeviceToHost, s));
        LUNA_HIP(hipMemcpyAsync(hc.data(), fv->m_cbar.p, sizeof(double) * ND, hipMemcpyDeviceToHost, s));
        LUNA_HIP(hipMemcpyAsync(hA.data(), fv->m_Ab.p, sizeof(double) * F0 * D, hipMemcpyDeviceToHost, s));
        LUNA_HIP(hipMemcpyAsync(hxf.data(), fv->m_Xfb.p, sizeof(double) * F0 * D, hipMemcpyDeviceToHost, s));
        LUNA_HIP(hipMemcpyAsync(hx.data(), fv->m_xbar.p, sizeof(double) * ND, hipMemcpyDeviceToHost, s));
        LUNA_HIP(hipMemcpyAsync(qbar, fv->m_out.p, sizeof(double) * NV, hipMemcpyDeviceToHost, s));
        LUNA_HIP(hipStreamSynchronize(s));
        if (adjoint0(ctx0, hV.data(), hc.data(), hA.data(), hxf.data(), hx.data()) != 0)
            return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_adjoint: the geometry adjoint of the mesh of t^n failed");
        std::memcpy(xbar, hx.data(), sizeof(double) * ND);
        fv->prims_ready = false;
        fv->swept = false;
        return LUNA_OK;
    }
    // Pass 2 on t^n+1 (the rows W^n, g^n, c^n and the steps stay), then x^n+1's bar with its geometry's adjoint.
    fv->begun = true;  // keeps W, g across the upload
    int rc = upload(fv, mesh1);
    fv->begun = false;
    if (rc != LUNA_OK) return rc;
    {
        const luna::Geo g = fv->geo();
        const luna::Rows r = fv->rows();
        const luna::MovingBars b = bars();
        LUNA_HIP(D == 3 ? luna::launch_moving_pass_adjoint<3>(s, g, r, gas, fv->opt.riemann, dt, unit, true, fv->m_qbar.p, b)
                        : luna::launch_moving_pass_adjoint<2>(s, g, r, gas, fv->opt.riemann, dt, unit, true, fv->m_qbar.p, b));
    }
    const size_t F1 = static_cast<size_t>(mesh1->faces), F0 = static_cast<size_t>(mesh0->faces);
    std::vector<double> hV(N, 0.0), hc(ND), hA(Fmax * D), hxf(Fmax * D), hx(ND);
    LUNA_HIP(hipMemcpyAsync(hc.data(), fv->m_cbar.p, sizeof(double) * ND, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipMemcpyAsync(hA.data(), fv->m_Ab.p, sizeof(double) * F1 * D, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipMemcpyAsync(hxf.data(), fv->m_Xfb.p, sizeof(double) * F1 * D, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipMemcpyAsync(hx.data(), fv->m_xbar.p, sizeof(double) * ND, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipStreamSynchronize(s));
    if (adjoint1(ctx1, hV.data(), hc.data(), hA.data(), hxf.data(), hx.data()) != 0)
        return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_adjoint: the geometry adjoint of the mesh of t^n+1 failed");
    // Back to t^n: x^n's bar starts as x^n+1's (x^n+1 = x^n + dt w-hat), w-hat's gets dt of it unless the steps are held.
    fv->begun = true;
    rc = upload(fv, mesh0);
    fv->begun = false;
    if (rc != LUNA_OK) return rc;
    LUNA_HIP(hipMemcpyAsync(fv->m_xbar.p, hx.data(), sizeof(double) * ND, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->m_cbar.p, fv->m_c0bar.p, sizeof(double) * ND, hipMemcpyDeviceToDevice, s));
    if (!steps_fixed) LUNA_HIP(luna::launch_drift_adjoint(s, static_cast<long long>(ND), dt, fv->m_xbar.p, fv->m_whbar.p));
    LUNA_HIP(fv->m_entry.reserve(static_cast<size_t>(fv->slots > 0 ? fv->slots : 1) * (2 * D + 1)));
    {
        const luna::Geo g = fv->geo();
        const luna::Rows r = fv->rows();
        luna::MovingBars b = bars();
        b.entry = fv->m_entry.p;
        LUNA_HIP(D == 3 ? luna::launch_moving_pass_adjoint<3>(s, g, r, gas, fv->opt.riemann, dt, unit, false, fv->m_qbar.p, b)
                        : luna::launch_moving_pass_adjoint<2>(s, g, r, gas, fv->opt.riemann, dt, unit, false, fv->m_qbar.p, b));
        LUNA_HIP(D == 3 ? luna::launch_moving_cell_adjoint<3>(s, g, r, gas, limit, steps_fixed != 0, fv->m_qbar.p, b)
                        : luna::launch_moving_cell_adjoint<2>(s, g, r, gas, limit, steps_fixed != 0, fv->m_qbar.p, b));
    }
    LUNA_HIP(hipMemcpyAsync(hV.data(), fv->m_Vbar.p, sizeof(double) * N, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipMemcpyAsync(hc.data(), fv->m_cbar.p, sizeof(double) * ND, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipMemcpyAsync(hA.data(), fv->m_Ab.p, sizeof(double) * F0 * D, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipMemcpyAsync(hxf.data(), fv->m_Xfb.p, sizeof(double) * F0 * D, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipMemcpyAsync(hx.data(), fv->m_xbar.p, sizeof(double) * ND, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipMemcpyAsync(qbar, fv->m_out.p, sizeof(double) * NV, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipStreamSynchronize(s));
    if (adjoint0(ctx0, hV.data(), hc.data(), hA.data(), hxf.data(), hx.data()) != 0)
        return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_adjoint: the geometry adjoint of the mesh of t^n failed");
    std::memcpy(xbar, hx.data(), sizeof(double) * ND);
    fv->prims_ready = false;
    fv->swept = false;
    return LUNA_OK;
}

int luna_fv_moving_tangent(luna_fv* fv, double dt, const luna_geometry* mesh0, const luna_geometry* mesh1, luna_geometry_tangent_fn tangent0,
                           const void* ctx0, luna_geometry_tangent_fn tangent1, const void* ctx1, int steps_fixed, double* dq, double* dx) {
    if (!fv || !mesh0 || !mesh1 || !tangent0 || !tangent1 || !dq || !dx) return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_tangent: null argument");
    if (fv->opt.motion != LUNA_MOTION_MOVING) return fail(LUNA_ERR_LOGIC, "hydro::FiniteVolume::step_moving_tangent: a moving mesh");
    if (fv->opt.limiter == LUNA_LIMITER_LEARNED)
        return fail(LUNA_ERR_LOGIC, "luna_fv_moving_tangent: derivatives of the face-centroid limiter or none (not the learned one)");
    if (fv->opt.precision != LUNA_PRECISION_DOUBLE) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_tangent: derivatives of the double policy");
    if (fv->opt.hll_compression > 0) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_tangent: derivatives without the HLL fallback");
    if (fv->opt.sweep_lanes > 1) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_tangent: derivatives of the one-lane sweep");
    if (fv->begun) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_tangent: a moving step is between begin and finish");
    if (!(dt > 0)) return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_tangent: dt must be positive");
    if (mesh0->cells != fv->N || mesh1->cells != fv->N) return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_tangent: both meshes must have the handle's cells");
    const int D = fv->D, V = fv->V, N = fv->N;
    const size_t ND = static_cast<size_t>(N) * D, NV = static_cast<size_t>(N) * V;
    hipStream_t s = fv->stream;
    const bool limit = fv->opt.limiter == LUNA_LIMITER_FACE_CENTROID, order0 = fv->opt.order == 0;
    // t^n: its mesh and positions, W, the limited gradients and point velocities, c^n and the grid steps
    if (int rc = luna_fv_set_geometry(fv, mesh0)) return rc;
    if (int rc = primitives(fv)) return rc;
    if (int rc = sweep(fv, 1.0)) return rc;
    if (!order0) {
        const luna::Geo g = fv->geo();
        const luna::Rows r = fv->rows();
        LUNA_HIP(D == 3 ? luna::launch_keep_centroids<3>(s, g, r, fv->grid, dt) : luna::launch_keep_centroids<2>(s, g, r, fv->grid, dt));
    }
    const double unit = fv->grid.step / dt;
    const size_t Fmax = static_cast<size_t>(std::max(std::max(mesh0->faces, mesh1->faces), 1));
    for (Buf<double>* b : {&fv->t_dx, &fv->t_dx1, &fv->t_dc, &fv->t_dc0, &fv->t_dw}) LUNA_HIP(b->reserve(ND));
    LUNA_HIP(fv->t_dV.reserve(N));
    LUNA_HIP(fv->t_dcs.reserve(N));
    LUNA_HIP(fv->t_dA.reserve(Fmax * D));
    LUNA_HIP(fv->t_dxf.reserve(Fmax * D));
    LUNA_HIP(fv->t_dW.reserve(NV));
    LUNA_HIP(fv->t_dg.reserve(NV * D));
    LUNA_HIP(fv->t_dG.reserve(Fmax * V));
    LUNA_HIP(fv->t_dq.reserve(NV));
    // the geometry's tangent of the mesh of t^n (the mesh library's, on the host)
    const size_t F0 = static_cast<size_t>(mesh0->faces), F1 = static_cast<size_t>(mesh1->faces);
    std::vector<double> hV(N), hc(ND), hA(Fmax * D), hxf(Fmax * D), hx1(ND);
    if (tangent0(ctx0, dx, hV.data(), hc.data(), hA.data(), hxf.data()) != 0)
        return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_tangent: the geometry tangent of the mesh of t^n failed");
    LUNA_HIP(hipMemcpyAsync(fv->t_dq.p, dq, sizeof(double) * NV, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->t_dx.p, dx, sizeof(double) * ND, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->t_dV.p, hV.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->t_dc.p, hc.data(), sizeof(double) * ND, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->t_dA.p, hA.data(), sizeof(double) * F0 * D, hipMemcpyHostToDevice, s));
    LUNA_HIP(hipMemcpyAsync(fv->t_dxf.p, hxf.data(), sizeof(double) * F0 * D, hipMemcpyHostToDevice, s));
    luna::MovingDots t{fv->t_dx.p, fv->t_dV.p, fv->t_dc.p, fv->t_dc0.p, fv->t_dA.p, fv->t_dxf.p, fv->t_dW.p, fv->t_dcs.p, fv->t_dw.p, fv->t_dg.p,
                       fv->t_dG.p};
    const luna::Gas gas = fv->gas();
    {
        const luna::Geo g = fv->geo();
        const luna::Rows r = fv->rows();
        LUNA_HIP(D == 3 ? luna::launch_moving_tangent_cells<3>(s, g, r, gas, limit, !order0, fv->t_dq.p, t)
                        : luna::launch_moving_tangent_cells<2>(s, g, r, gas, limit, !order0, fv->t_dq.p, t));
        if (!order0 && steps_fixed) LUNA_HIP(hipMemsetAsync(fv->t_dw.p, 0, sizeof(double) * ND, s));  // no derivative through w-hat
        // pass 1 on t^n (order 0: the one pass over dt with the sweep's w)
        LUNA_HIP(D == 3 ? luna::launch_moving_pass_tangent<3>(s, g, r, gas, fv->opt.riemann, dt, unit, false, order0, t, fv->t_dq.p)
                        : luna::launch_moving_pass_tangent<2>(s, g, r, gas, fv->opt.riemann, dt, unit, false, order0, t, fv->t_dq.p));
    }
    // x^n+1 = x^n + dt w-hat (order 0: + dt w, off the grid)
    LUNA_HIP(luna::launch_axpy(s, static_cast<long long>(ND), dt, fv->t_dx.p, fv->t_dw.p, fv->t_dx1.p));
    LUNA_HIP(hipMemcpyAsync(hx1.data(), fv->t_dx1.p, sizeof(double) * ND, hipMemcpyDeviceToHost, s));
    if (!order0) {
        LUNA_HIP(hipMemcpyAsync(fv->t_dc0.p, fv->t_dc.p, sizeof(double) * ND, hipMemcpyDeviceToDevice, s));
        LUNA_HIP(hipStreamSynchronize(s));
        // pass 2 on t^n+1 with its geometry's tangent along dx^n+1 (W^n, g^n, c^n and the steps stay)
        if (tangent1(ctx1, hx1.data(), hV.data(), hc.data(), hA.data(), hxf.data()) != 0)
            return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_tangent: the geometry tangent of the mesh of t^n+1 failed");
        LUNA_HIP(hipMemcpyAsync(fv->t_dV.p, hV.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
        LUNA_HIP(hipMemcpyAsync(fv->t_dc.p, hc.data(), sizeof(double) * ND, hipMemcpyHostToDevice, s));
        LUNA_HIP(hipMemcpyAsync(fv->t_dA.p, hA.data(), sizeof(double) * F1 * D, hipMemcpyHostToDevice, s));
        LUNA_HIP(hipMemcpyAsync(fv->t_dxf.p, hxf.data(), sizeof(double) * F1 * D, hipMemcpyHostToDevice, s));
        fv->begun = true;  // keeps W, g across the upload
        int rc = upload(fv, mesh1);
        fv->begun = false;
        if (rc != LUNA_OK) return rc;
        t.dx = fv->t_dx1.p;
        const luna::Geo g = fv->geo();
        const luna::Rows r = fv->rows();
        LUNA_HIP(D == 3 ? luna::launch_moving_pass_tangent<3>(s, g, r, gas, fv->opt.riemann, dt, unit, true, false, t, fv->t_dq.p)
                        : luna::launch_moving_pass_tangent<2>(s, g, r, gas, fv->opt.riemann, dt, unit, true, false, t, fv->t_dq.p));
    }
    LUNA_HIP(hipMemcpyAsync(dq, fv->t_dq.p, sizeof(double) * NV, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipStreamSynchronize(s));
    std::memcpy(dx, hx1.data(), sizeof(double) * ND);
    if (!order0) {  // back on the mesh of t^n
        fv->begun = true;
        int rc = upload(fv, mesh0);
        fv->begun = false;
        if (rc != LUNA_OK) return rc;
    }
    fv->prims_ready = false;
    fv->swept = false;
    return LUNA_OK;
}

int luna_fv_moving_begin(luna_fv* fv, double dt, double* positions) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_begin: null argument");
    if (fv->opt.motion != LUNA_MOTION_MOVING) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_begin: a static mesh steps with luna_fv_step");
    if (fv->begun) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_begin: the last step was not finished");
    if (!(dt > 0)) return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_begin: dt must be positive");
    if (int rc = ensure_swept(fv)) return rc;
    fv->swept = false;
    const luna::Geo g = fv->geo();
    const luna::Rows r = fv->rows();
    hipStream_t s = fv->stream;
    const luna::Grid grid = fv->grid;
    if (fv->opt.order == 0) {  // finite_volume.hpp:216-225: advance<0, true>() on the mesh of t^n, then move()
        if (int rc = flux_and_update(fv, 2, dt, dt, 0.0, false)) return rc;
        LUNA_HIP(timed(fv->prof, s, "drift", [&] {
            return fv->D == 3 ? luna::launch_drift<3>(s, g, r, dt) : luna::launch_drift<2>(s, g, r, dt);
        }));
        if (positions) LUNA_HIP(hipMemcpyAsync(positions, fv->x.p, sizeof(double) * fv->N * fv->D, hipMemcpyDeviceToHost, s));
        LUNA_HIP(hipStreamSynchronize(s));
        fv->begun = true;
        fv->new_mesh = false;
        return LUNA_OK;
    }
    LUNA_HIP(timed(fv->prof, s, "keep_centroids", [&] {
        return fv->D == 3 ? luna::launch_keep_centroids<3>(s, g, r, grid, dt) : luna::launch_keep_centroids<2>(s, g, r, grid, dt);
    }));
    const double unit = fv->grid.step / dt;
    if (int rc = flux_and_update(fv, 0, 0.0, 0.5 * dt, unit, false)) return rc;  // the update also drifts x
    if (positions) LUNA_HIP(hipMemcpyAsync(positions, fv->x.p, sizeof(double) * fv->N * fv->D, hipMemcpyDeviceToHost, s));
    LUNA_HIP(hipStreamSynchronize(s));
    fv->begun = true;
    fv->new_mesh = false;
    return LUNA_OK;
}

}  // extern "C"

namespace {

int set_geometry(luna_fv* fv, const luna_geometry* mesh, bool device) {
    if (!fv || !mesh) return fail(LUNA_ERR_ARGUMENT, "luna_fv_set_geometry: null argument");
    if (fv->begun) fv->new_mesh = false;
    if (int rc = device ? upload_device(fv, mesh) : upload(fv, mesh)) return rc;
    if (fv->begun) {
        fv->new_mesh = true;
        ++fv->updates;
    } else {  // mesh_moved() (finite_volume.hpp:152-156): the cached rows go; a moving solver takes the mesh's positions
        fv->prims_ready = false;
        fv->swept = false;
        if (fv->opt.motion == LUNA_MOTION_MOVING) {
            LUNA_HIP(hipMemcpyAsync(fv->x.p, mesh->positions, sizeof(double) * fv->N * fv->D,
                                    device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, fv->stream));
            LUNA_HIP(hipStreamSynchronize(fv->stream));
        }
    }
    return LUNA_OK;
}

}  // namespace

extern "C" {

int luna_fv_set_geometry(luna_fv* fv, const luna_geometry* mesh) { return set_geometry(fv, mesh, false); }

int luna_fv_set_geometry_device(luna_fv* fv, const luna_geometry* mesh) { return set_geometry(fv, mesh, true); }

int luna_fv_moving_finish(luna_fv* fv, const luna_geometry* mesh, double dt) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv_moving_finish: null argument");
    if (!fv->begun) return fail(LUNA_ERR_LOGIC, "luna_fv_moving_finish: no step was begun");
    if (mesh) {
        fv->new_mesh = false;
        if (int rc = upload(fv, mesh)) return rc;
        ++fv->updates;
    } else if (!fv->new_mesh) {
        return fail(LUNA_ERR_LOGIC, "luna_fv_moving_finish: no mesh of the new positions (pass one or call luna_fv_set_geometry)");
    }
    if (fv->opt.order == 0) {  // Q stays; the new mesh's primitives
        const luna::Geo g = fv->geo();
        const luna::Rows r = fv->rows();
        const luna::Gas gas = fv->gas();
        hipStream_t s = fv->stream;
        LUNA_HIP(timed(fv->prof, s, "primitives", [&] {
            return fv->D == 3 ? luna::launch_primitives<3>(s, g, r, gas) : luna::launch_primitives<2>(s, g, r, gas);
        }));
        fv->prims_ready = true;
        fv->begun = false;
        fv->new_mesh = false;
        return LUNA_OK;
    }
    const double unit = fv->grid.step / dt;
    bool prims = false;
    if (int rc = flux_and_update(fv, 1, dt, 0.5 * dt, unit, true, &prims)) return rc;
    fv->prims_ready = prims;
    fv->begun = false;
    fv->new_mesh = false;
    return LUNA_OK;
}

int luna_synchronize(void) {
    LUNA_HIP(hipDeviceSynchronize());
    return LUNA_OK;
}

int luna_fv_fallbacks(luna_fv* fv, long long* count) {
    if (!fv || !count) return fail(LUNA_ERR_ARGUMENT, "luna_fv_fallbacks: null argument");
    LUNA_HIP(hipMemcpyAsync(fv->host + 2, fv->counters.p + 2, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    LUNA_HIP(fv->prof.collect());
    *count = static_cast<long long>(fv->host[2] + fv->host[3]);
    return LUNA_OK;
}

int luna_fv_positions(luna_fv* fv, double* x) {
    if (!fv || !x) return fail(LUNA_ERR_ARGUMENT, "luna_fv_positions: null argument");
    if (fv->opt.motion == LUNA_MOTION_MOVING)
        LUNA_HIP(hipMemcpyAsync(x, fv->x.p, sizeof(double) * fv->N * fv->D, hipMemcpyDeviceToHost, fv->stream));
    else  // the mesh's positions, from the cell rows
        LUNA_HIP(hipMemcpy2DAsync(x, sizeof(double) * fv->D, fv->cell.p, sizeof(double) * luna::CELL_ROW, sizeof(double) * fv->D, fv->N,
                                  hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

int luna_fv_point_velocities(luna_fv* fv, double* w) {
    if (!fv || !w) return fail(LUNA_ERR_ARGUMENT, "luna_fv_point_velocities: null argument");
    if (fv->opt.motion != LUNA_MOTION_MOVING) return fail(LUNA_ERR_LOGIC, "luna_fv_point_velocities: static mesh");
    LUNA_HIP(hipMemcpyAsync(w, fv->w.p, sizeof(double) * fv->N * fv->D, hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

int luna_fv_gradients(luna_fv* fv, double* g) {
    if (!fv || !g) return fail(LUNA_ERR_ARGUMENT, "luna_fv_gradients: null argument");
    if (fv->opt.order != 1) return fail(LUNA_ERR_LOGIC, "luna_fv_gradients: order 0 keeps no gradients");
    const size_t VD = static_cast<size_t>(fv->V) * fv->D, N = fv->N;
    if (fv->opt.precision == LUNA_PRECISION_DOUBLE) {
        const size_t width = sizeof(double) * VD;
        LUNA_HIP(hipMemcpy2DAsync(g, width, fv->grad.p, sizeof(double) * fv->grad_row(), width, fv->N, hipMemcpyDeviceToHost, fv->stream));
        LUNA_HIP(hipStreamSynchronize(fv->stream));
        return LUNA_OK;
    }
    // reduced precision: the stored rows, decoded on the host
    const size_t row = fv->opt.precision == LUNA_PRECISION_GRAD32 ? sizeof(float) * (fv->D == 3 ? luna::lp_row<3>() : luna::lp_row<2>())
                                                                  : sizeof(unsigned short) * luna::HALF_ROW;
    std::vector<unsigned char> raw(N * row);
    std::vector<unsigned int> k(fv->opt.precision == LUNA_PRECISION_GRAD16 ? N : 0);
    LUNA_HIP(hipMemcpyAsync(raw.data(), fv->grad.p, raw.size(), hipMemcpyDeviceToHost, fv->stream));
    if (!k.empty()) LUNA_HIP(hipMemcpyAsync(k.data(), fv->exponents.p, sizeof(unsigned int) * N, hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    for (size_t i = 0; i < N; ++i)
        for (size_t c = 0; c < VD; ++c) {
            if (k.empty()) {
                float f;
                std::memcpy(&f, raw.data() + i * row + c * sizeof(float), sizeof f);
                g[i * VD + c] = f;
            } else {
                unsigned short h;
                std::memcpy(&h, raw.data() + i * row + c * sizeof h, sizeof h);
                const int group = c < static_cast<size_t>(fv->D) ? 0 : c < static_cast<size_t>(fv->D + fv->D * fv->D) ? 1 : 2;
                const int e = static_cast<signed char>(static_cast<unsigned char>(k[i] >> (8 * group)));
                g[i * VD + c] = static_cast<double>(half_value(h) * std::ldexp(1.0f, -e));
            }
        }
    return LUNA_OK;
}

int luna_fv_primitives(luna_fv* fv, double* rows) {
    if (!fv || !rows) return fail(LUNA_ERR_ARGUMENT, "luna_fv_primitives: null argument");
    const size_t pitch = sizeof(double) * (fv->V + 2), row = sizeof(double) * luna::CELL_ROW;
    // W and c_s from the primitive half of the cell rows, r from the geometry half
    LUNA_HIP(hipMemcpy2DAsync(rows, pitch, fv->cell.p + luna::GEO_ROW, row, sizeof(double) * (fv->V + 1), fv->N, hipMemcpyDeviceToHost,
                              fv->stream));
    LUNA_HIP(hipMemcpy2DAsync(rows + fv->V + 1, pitch, fv->cell.p + 2 * fv->D + 1, row, sizeof(double), fv->N, hipMemcpyDeviceToHost,
                              fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

int luna_fv_faces(luna_fv* fv, int* faces) {
    if (!fv || !faces) return fail(LUNA_ERR_ARGUMENT, "luna_fv_faces: null argument");
    *faces = fv->F;
    return LUNA_OK;
}

int luna_fv_set_step_rule(luna_fv* fv, int rule) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv_set_step_rule: null argument");
    if (rule != LUNA_STEP_AREPO && rule != LUNA_STEP_LUNA) return fail(LUNA_ERR_ARGUMENT, "set_step_rule: AREPO or LUNA");
    if (rule == LUNA_STEP_LUNA && fv->D != 2) return fail(LUNA_ERR_ARGUMENT, "StepRule::luna: 2D");
    fv->opt.step_rule = rule;
    fv->swept = false;  // the next step sweeps again with the rule's dt terms
    return LUNA_OK;
}

int luna_fv_release_mesh(luna_fv* fv) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv_release_mesh: null argument");
    if (fv->opt.motion == LUNA_MOTION_MOVING) return fail(LUNA_ERR_LOGIC, "hydro::FiniteVolume::release_mesh: a moving mesh updates the mesh");
    // A device mesh's face arrays are read in place: the handle copies them into its own memory, so the caller may free
    // the mesh (a host mesh's are the handle's copies already).
    if (fv->mesh_ok && fv->mesh.owner != fv->src_owner.p) {
        const size_t F = fv->F, FD = F * fv->D;
        hipStream_t s = fv->stream;
        LUNA_HIP(fv->src_owner.reserve(F > 0 ? F : 1));
        LUNA_HIP(fv->src_neighbour.reserve(F > 0 ? F : 1));
        LUNA_HIP(fv->src_area.reserve(FD > 0 ? FD : 1));
        LUNA_HIP(fv->src_fc.reserve(FD > 0 ? FD : 1));
        LUNA_HIP(hipMemcpyAsync(fv->src_owner.p, fv->mesh.owner, sizeof(int) * F, hipMemcpyDeviceToDevice, s));
        LUNA_HIP(hipMemcpyAsync(fv->src_neighbour.p, fv->mesh.neighbour, sizeof(int) * F, hipMemcpyDeviceToDevice, s));
        LUNA_HIP(hipMemcpyAsync(fv->src_area.p, fv->mesh.area, sizeof(double) * FD, hipMemcpyDeviceToDevice, s));
        LUNA_HIP(hipMemcpyAsync(fv->src_fc.p, fv->mesh.face_centroid, sizeof(double) * FD, hipMemcpyDeviceToDevice, s));
        LUNA_HIP(hipStreamSynchronize(s));
        fv->mesh.owner = fv->src_owner.p;
        fv->mesh.neighbour = fv->src_neighbour.p;
        fv->mesh.area = fv->src_area.p;
        fv->mesh.face_centroid = fv->src_fc.p;
    }
    return LUNA_OK;
}

int luna_fv_mesh_stats(luna_fv* fv, long long* updates) {
    if (!fv || !updates) return fail(LUNA_ERR_ARGUMENT, "luna_fv_mesh_stats: null argument");
    *updates = fv->updates;
    return LUNA_OK;
}

int luna_fv_profile(luna_fv* fv, int on) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv_profile: null argument");
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    fv->prof.clear();
    fv->prof.on = on != 0;
    return LUNA_OK;
}

int luna_fv_profile_read(luna_fv* fv, luna_kernel_time* out, int max, int* count) {
    if (!fv || !count) return fail(LUNA_ERR_ARGUMENT, "luna_fv_profile_read: null argument");
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    LUNA_HIP(fv->prof.collect());
    const int n = static_cast<int>(fv->prof.names.size());
    *count = n;
    for (int k = 0; k < n && k < max && out; ++k) {
        std::snprintf(out[k].name, sizeof out[k].name, "%s", fv->prof.names[k].c_str());
        out[k].launches = fv->prof.launches[k];
        out[k].total_ms = fv->prof.total_ms[k];
    }
    return LUNA_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------------------- D = 1

struct luna_fv1 {
    luna_fv1_options opt{};
    int n = 0, periodic = 0, wall_left = 1, wall_right = 1;
    long long period = 0;
    double step = 0;
    hipStream_t stream = nullptr;
    Buf<long long> keys;
    Buf<double> q, next, dq, dnext;  // dq, dnext: the derivatives' rows (first use)
    Buf<unsigned long long> counters;
    unsigned long long* host = nullptr;
    int parity = 0;

    // The moving line. Q, next, x, x_next by cell, `capacity` cells; the line's slots at [lo, lo + n) of line_key[cur]
    // (its keys), line_key[cur ^ 1] (the keys after a step) and line_id, `slots` long, so that a cell from the left wall
    // takes slot lo - 1 without moving the others.
    bool moving = false;
    int regularise = 1;
    double side = 0, inflow[2][3] = {};
    int capacity = 0, slots = 0, lo = 0, cur = 0, dparity = 0;
    Buf<long long> line_key[2];
    Buf<int> line_id;
    Buf<double> x, x_next, ends, inflow_dev;
    Buf<unsigned long long> descents, packed;
    Buf<unsigned char> sort_temp;
    size_t sort_bytes = 0;
    double* host_ends = nullptr;
    unsigned long long* host_descents = nullptr;
    // The cell supply (finite_volume_1d.hpp:991-996): which ends supply, whether gas enters (watching), whether this
    // step's time_step() looked at the ends, what they asked for, steps to the next look; counters.
    bool supply_left = false, supply_right = false, watching = false, looked = false;
    double spacing = 0;
    int wants = 0, until_look = 0;
    long long added = 0, updates = 0, sorts = 0;

    luna::LineArgs args() const {
        luna::LineArgs a{};
        a.n = n;
        a.periodic = periodic;
        a.wall_left = wall_left;
        a.wall_right = wall_right;
        a.order = opt.order;
        a.limiter = opt.limiter;
        a.riemann = opt.riemann;
        a.gas = luna::Gas{opt.gamma, opt.smallr, opt.smallc, 1.0 / (opt.gamma - 1.0)};
        a.key = keys.p;
        a.period = period;
        a.step = step;
        a.q = q.p;
        a.next = next.p;
        if (moving) {
            a.moving = 1;
            a.regularise = regularise;
            a.key = line_key[cur].p + lo;
            a.key_next = line_key[cur ^ 1].p + lo;
            a.id = line_id.p + lo;
            a.side = side;
            a.inflow = inflow_dev.p;
            a.x = x.p;
            a.x_next = x_next.p;
        }
        return a;
    }
    luna::Counters counters_now() const { return luna::Counters{counters.p + parity, counters.p + (parity ^ 1), counters.p + 2}; }
};

namespace {

// A device array grown to `count` elements keeping `keep` of them (moved from offset `from` to offset `to`).
template <class T>
hipError_t grow(Buf<T>& b, size_t count, size_t keep, size_t from, size_t to, hipStream_t s) {
    Buf<T> nb;
    hipError_t e = nb.reserve(count);
    if (e != hipSuccess) return e;
    if (keep) e = hipMemcpyAsync(nb.p + to, b.p + from, sizeof(T) * keep, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        nb.release();
        return e;
    }
    b.release();
    b = nb;
    return hipSuccess;
}

// finite_volume_1d.hpp:965-981 reserve_cells(): room for `count` cells, a quarter more when full.
hipError_t reserve_cells(luna_fv1* fv, int count) {
    if (count <= fv->capacity) return hipSuccess;
    const int cap = std::max(count, fv->capacity + fv->capacity / 4);
    hipError_t e;
    if ((e = grow(fv->q, 3 * static_cast<size_t>(cap), 3 * static_cast<size_t>(fv->n), 0, 0, fv->stream)) != hipSuccess) return e;
    if ((e = grow(fv->next, 3 * static_cast<size_t>(cap), 0, 0, 0, fv->stream)) != hipSuccess) return e;
    if ((e = grow(fv->x, cap, fv->n, 0, 0, fv->stream)) != hipSuccess) return e;
    if ((e = grow(fv->x_next, cap, 0, 0, 0, fv->stream)) != hipSuccess) return e;
    fv->capacity = cap;
    return hipSuccess;
}

// Room for a slot before the line (left) and after it (right): the slot arrays re-made with a margin of a quarter of
// the line at each end when one is short.
hipError_t reserve_slots(luna_fv1* fv, bool left, bool right) {
    if ((!left || fv->lo > 0) && (!right || fv->lo + fv->n < fv->slots)) return hipSuccess;
    const int margin = std::max(64, fv->n / 4);
    const int slots = fv->n + 2 * margin + 2, lo = margin + 1;
    hipError_t e;
    if ((e = grow(fv->line_key[fv->cur], slots, fv->n, fv->lo, lo, fv->stream)) != hipSuccess) return e;
    if ((e = grow(fv->line_key[fv->cur ^ 1], slots, 0, 0, 0, fv->stream)) != hipSuccess) return e;
    if ((e = grow(fv->line_id, slots, fv->n, fv->lo, lo, fv->stream)) != hipSuccess) return e;
    fv->slots = slots;
    fv->lo = lo;
    return hipSuccess;
}

double host_snap(const luna_fv1* fv, double x) {
    long long k = static_cast<long long>(std::round(x * (1.0 / fv->step)));
    if (fv->periodic) {
        k %= fv->period;
        if (k < 0) k += fv->period;
    }
    return static_cast<double>(k);
}

// The ends of the line (luna::LINE_ENDS values) into fv->host_ends; blocks.
int read_ends(luna_fv1* fv) {
    LUNA_HIP(luna::launch_line_ends(fv->stream, fv->args(), fv->spacing, fv->supply_left, fv->supply_right, fv->ends.p));
    LUNA_HIP(hipMemcpyAsync(fv->host_ends, fv->ends.p, sizeof(double) * luna::LINE_ENDS, hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

// finite_volume_1d.hpp:419-465 supply_cells(): a wanted cell's generator goes `spacing` inside the end generator's
// distance from its wall, in a new end slot; it takes from the old end cell the fraction it overlaps (its whole length,
// from the wall to halfway to the old end generator) of mass, momentum and energy. New cells are numbered after the
// last, the left wall's first. Returns whether cells were added into *added_any.
int supply_cells(luna_fv1* fv, bool* added_any) {
    *added_any = false;
    const int rc = read_ends(fv);
    if (rc != LUNA_OK) return rc;
    const double* e = fv->host_ends;
    fv->watching = fv->watching || e[4] > 0;
    const bool left = e[1] > 0, right = e[3] > 0;
    if (!left && !right) return LUNA_OK;
    const int requests = int(left) + int(right);
    LUNA_HIP(reserve_cells(fv, fv->n + requests));
    LUNA_HIP(reserve_slots(fv, left, right));
    luna::LineSplit split[2];
    int r = 0;
    const long long key0 = static_cast<long long>(e[7]), keyl = static_cast<long long>(e[8]);
    int n = fv->n;
    if (left) {
        const double wanted = e[0] - fv->spacing;
        const long long k = static_cast<long long>(host_snap(fv, wanted));
        const double overlap = (static_cast<double>(2 * k) + static_cast<double>(key0 - k)) * (0.5 * fv->step);
        split[r] = luna::LineSplit{n, static_cast<int>(e[9]), overlap / e[5], wanted, k, fv->line_key[fv->cur].p + fv->lo - 1,
                                   fv->line_id.p + fv->lo - 1};
        fv->lo -= 1;
        ++n;
        ++r;
    }
    if (right) {
        const double wanted = fv->side - (e[2] - fv->spacing);
        const long long k = static_cast<long long>(host_snap(fv, wanted));
        const double overlap = (static_cast<double>(k - keyl) + static_cast<double>(2 * (fv->period - k))) * (0.5 * fv->step);
        split[r] = luna::LineSplit{n, static_cast<int>(e[10]), overlap / e[6], wanted, k, fv->line_key[fv->cur].p + fv->lo + n,
                                   fv->line_id.p + fv->lo + n};
        ++n;
        ++r;
    }
    LUNA_HIP(luna::launch_line_split(fv->stream, fv->q.p, fv->x.p, split, r));
    fv->n = n;
    fv->added += r;
    *added_any = true;
    return LUNA_OK;
}

// finite_volume_1d.hpp:954-961 due(): look at the ends after this step: as the kernels asked while gas enters, else
// every 64 steps.
bool due(luna_fv1* fv) {
    if (!fv->supply_left && !fv->supply_right) return false;
    if (fv->watching) return !fv->looked || fv->wants > 0;
    if (--fv->until_look > 0) return false;
    fv->until_look = 64;
    return true;
}

// Mesh<1>::order(): the slots in key order again when the step's new keys left two adjacent slots out of order (the
// points crossed, or one crossed a periodic line's seam: the line is kept starting at its smallest key).
int order_line(luna_fv1* fv, unsigned long long descents, bool* sorted) {
    fv->updates += 1;
    *sorted = descents != 0;
    if (!descents) return LUNA_OK;
    const int n = fv->n;
    LUNA_HIP(fv->packed.reserve(2 * static_cast<size_t>(n)));
    size_t bytes = 0;
    LUNA_HIP(luna::launch_line_sort(fv->stream, nullptr, nullptr, n, fv->packed.p, nullptr, &bytes));
    LUNA_HIP(fv->sort_temp.reserve(bytes > 0 ? bytes : 1));
    bytes = fv->sort_temp.cap;
    LUNA_HIP(luna::launch_line_sort(fv->stream, fv->line_key[fv->cur].p + fv->lo, fv->line_id.p + fv->lo, n, fv->packed.p, fv->sort_temp.p,
                                    &bytes));
    fv->sorts += 1;
    return LUNA_OK;
}

// The moving kernel, then Q, the positions and the keys swap; returns the descents it counted (blocks).
int moving_kernel(luna_fv1* fv, double dt, double courant, bool fused, unsigned long long* descents) {
    const int d = fv->dparity;
    LUNA_HIP(luna::launch_line_step_moving(fv->stream, fv->args(), dt, courant, fused, fv->counters_now(), fv->descents.p + d,
                                           fv->descents.p + (d ^ 1)));
    fv->dparity ^= 1;
    std::swap(fv->q, fv->next);
    std::swap(fv->x, fv->x_next);
    fv->cur ^= 1;
    LUNA_HIP(hipMemcpyAsync(fv->host_descents, fv->descents.p + d, sizeof(unsigned long long), hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    *descents = fv->host_descents[0];
    return LUNA_OK;
}

}  // namespace

extern "C" {

int luna_fv1_create(const luna_fv1_options* opt, const luna_line* line, luna_fv1** out) {
    if (!opt || !line || !out || !line->keys) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create: null argument");
    *out = nullptr;
    if (opt->order != 0 && opt->order != 1) return fail(LUNA_ERR_ARGUMENT, "hydro::FiniteVolume<1>: order 0 or 1");
    if (opt->limiter < 0 || opt->limiter > 3) return fail(LUNA_ERR_ARGUMENT, "luna: unknown limiter");
    if (opt->riemann != LUNA_RIEMANN_HLLC && opt->riemann != LUNA_RIEMANN_HLL) return fail(LUNA_ERR_ARGUMENT, "luna: unknown Riemann solver");
    if (line->cells < 3) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create: at least 3 cells");
    if (line->period <= 0 || !(line->step > 0)) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create: bad grid");
    for (int s = 0; s < line->cells; ++s) {
        if (line->keys[s] < 0 || line->keys[s] >= line->period) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create: key outside the line");
        if (s > 0 && line->keys[s] <= line->keys[s - 1]) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create: keys must ascend");
    }
    for (int wall : {line->wall_left, line->wall_right})
        if (!line->periodic && wall != LUNA_WALL_REFLECTIVE && wall != LUNA_WALL_OUTFLOW)
            return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create: walls reflective or outflow (inflow supplies cells on a moving line only)");
    luna_fv1* fv = new luna_fv1;
    fv->opt = *opt;
    fv->n = line->cells;
    fv->periodic = line->periodic;
    fv->wall_left = line->wall_left;
    fv->wall_right = line->wall_right;
    fv->period = line->period;
    fv->step = line->step;
    auto cleanup = [&](int rc) {
        luna_fv1_destroy(fv);
        return rc;
    };
    const size_t n = fv->n;
    if (hipStreamCreate(&fv->stream) != hipSuccess) return cleanup(fail(LUNA_ERR_DEVICE, "luna: hipStreamCreate failed"));
    if (fv->keys.reserve(n) != hipSuccess || fv->q.reserve(3 * n) != hipSuccess || fv->next.reserve(3 * n) != hipSuccess ||
        fv->counters.reserve(4) != hipSuccess)
        return cleanup(fail(LUNA_ERR_MEMORY, "luna: device allocation failed"));
    if (hipHostMalloc(reinterpret_cast<void**>(&fv->host), 8 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(LUNA_ERR_MEMORY, "luna: pinned host allocation failed"));
    const unsigned long long init[4] = {luna::INF_BITS, luna::INF_BITS, 0, 0};
    std::vector<double> zero(3 * n, 0.0);
    if (hipMemcpy(fv->counters.p, init, sizeof init, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(fv->keys.p, line->keys, sizeof(long long) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(fv->q.p, zero.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(LUNA_ERR_DEVICE, "luna: initial upload failed"));
    *out = fv;
    return LUNA_OK;
}

int luna_fv1_create_moving(const luna_fv1_options* opt, const luna_moving_line* line, luna_fv1** out) {
    if (!opt || !line || !out || !line->positions) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create_moving: null argument");
    *out = nullptr;
    if (opt->order != 0 && opt->order != 1) return fail(LUNA_ERR_ARGUMENT, "hydro::FiniteVolume<1>: order 0 or 1");
    if (opt->limiter < 0 || opt->limiter > 3) return fail(LUNA_ERR_ARGUMENT, "luna: unknown limiter");
    if (opt->riemann != LUNA_RIEMANN_HLLC && opt->riemann != LUNA_RIEMANN_HLL) return fail(LUNA_ERR_ARGUMENT, "luna: unknown Riemann solver");
    if (line->cells < 3) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create_moving: at least 3 cells");
    if (line->period <= 0 || line->period > (1ll << 30) || !(line->side > 0) || !(line->spacing >= 0))
        return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create_moving: bad grid (period in 1 .. 2^30, side > 0, spacing >= 0)");
    for (int wall : {line->wall_left, line->wall_right})
        if (!line->periodic && wall != LUNA_WALL_REFLECTIVE && wall != LUNA_WALL_OUTFLOW && wall != LUNA_WALL_INFLOW)
            return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create_moving: unknown wall");
    for (int s = 0; s < line->cells; ++s) {
        const double p = line->positions[s];
        if (!(p >= 0) || !(line->periodic ? p < line->side : p <= line->side))
            return fail(LUNA_ERR_ARGUMENT, "luna_fv1_create_moving: a position outside the line");
    }
    luna_fv1* fv = new luna_fv1;
    fv->opt = *opt;
    fv->moving = true;
    fv->n = line->cells;
    fv->periodic = line->periodic;
    fv->wall_left = line->periodic ? LUNA_WALL_OUTFLOW : line->wall_left;
    fv->wall_right = line->periodic ? LUNA_WALL_OUTFLOW : line->wall_right;
    fv->period = line->period;
    fv->side = line->side;
    fv->step = line->side / static_cast<double>(line->period);
    fv->regularise = line->regularise != 0;
    for (int v = 0; v < 3; ++v) {
        fv->inflow[0][v] = line->inflow_left[v];
        fv->inflow[1][v] = line->inflow_right[v];
    }
    auto cleanup = [&](int rc) {
        luna_fv1_destroy(fv);
        return rc;
    };
    const int n = fv->n;
    // The slots in (key, cell) order of the snapped positions (Mesh<1>::update).
    std::vector<std::pair<long long, int>> order(n);
    for (int i = 0; i < n; ++i) order[i] = {static_cast<long long>(host_snap(fv, line->positions[i])), i};
    std::sort(order.begin(), order.end());
    const int margin = std::max(64, n / 4);
    fv->slots = n + 2 * margin + 2;
    fv->lo = margin + 1;
    fv->capacity = n;
    std::vector<long long> key(fv->slots, 0);
    std::vector<int> id(fv->slots, 0);
    for (int s = 0; s < n; ++s) {
        key[fv->lo + s] = order[s].first;
        id[fv->lo + s] = order[s].second;
    }
    if (hipStreamCreate(&fv->stream) != hipSuccess) return cleanup(fail(LUNA_ERR_DEVICE, "luna: hipStreamCreate failed"));
    if (fv->q.reserve(3 * static_cast<size_t>(n)) != hipSuccess || fv->next.reserve(3 * static_cast<size_t>(n)) != hipSuccess ||
        fv->x.reserve(n) != hipSuccess || fv->x_next.reserve(n) != hipSuccess || fv->line_key[0].reserve(fv->slots) != hipSuccess ||
        fv->line_key[1].reserve(fv->slots) != hipSuccess || fv->line_id.reserve(fv->slots) != hipSuccess ||
        fv->counters.reserve(4) != hipSuccess || fv->descents.reserve(2) != hipSuccess || fv->ends.reserve(luna::LINE_ENDS) != hipSuccess ||
        fv->inflow_dev.reserve(6) != hipSuccess)
        return cleanup(fail(LUNA_ERR_MEMORY, "luna: device allocation failed"));
    if (hipHostMalloc(reinterpret_cast<void**>(&fv->host), 4 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&fv->host_ends), luna::LINE_ENDS * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&fv->host_descents), sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(LUNA_ERR_MEMORY, "luna: pinned host allocation failed"));
    const unsigned long long init[4] = {luna::INF_BITS, luna::INF_BITS, 0, 0}, zeros[2] = {0, 0};
    std::vector<double> zero(3 * static_cast<size_t>(n), 0.0);
    if (hipMemcpy(fv->counters.p, init, sizeof init, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(fv->descents.p, zeros, sizeof zeros, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(fv->line_key[0].p, key.data(), sizeof(long long) * fv->slots, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(fv->line_id.p, id.data(), sizeof(int) * fv->slots, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(fv->x.p, line->positions, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(fv->inflow_dev.p, fv->inflow, sizeof fv->inflow, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(fv->q.p, zero.data(), sizeof(double) * 3 * n, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail(LUNA_ERR_DEVICE, "luna: initial upload failed"));
    fv->spacing = line->spacing;
    (void)luna_fv1_set_inflow_spacing(fv, line->spacing);
    *out = fv;
    return LUNA_OK;
}

int luna_fv1_set_inflow_spacing(luna_fv1* fv, double spacing) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_set_inflow_spacing: null handle");
    if (!(spacing >= 0)) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_set_inflow_spacing: spacing >= 0");
    fv->spacing = spacing;
    fv->supply_left = fv->moving && !fv->periodic && fv->wall_left != LUNA_WALL_REFLECTIVE && spacing > 0;
    fv->supply_right = fv->moving && !fv->periodic && fv->wall_right != LUNA_WALL_REFLECTIVE && spacing > 0;
    return LUNA_OK;
}

int luna_fv1_cells(luna_fv1* fv, int* cells, long long* added) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_cells: null handle");
    if (cells) *cells = fv->n;
    if (added) *added = fv->added;
    return LUNA_OK;
}

int luna_fv1_positions(luna_fv1* fv, double* x) {
    if (!fv || !x) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_positions: null argument");
    if (!fv->moving) return fail(LUNA_ERR_LOGIC, "luna_fv1_positions: a static line holds no positions");
    LUNA_HIP(hipMemcpyAsync(x, fv->x.p, sizeof(double) * fv->n, hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

int luna_fv1_line(luna_fv1* fv, long long* keys, int* cells) {
    if (!fv || !keys || !cells) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_line: null argument");
    if (fv->moving) {
        LUNA_HIP(hipMemcpyAsync(keys, fv->line_key[fv->cur].p + fv->lo, sizeof(long long) * fv->n, hipMemcpyDeviceToHost, fv->stream));
        LUNA_HIP(hipMemcpyAsync(cells, fv->line_id.p + fv->lo, sizeof(int) * fv->n, hipMemcpyDeviceToHost, fv->stream));
    } else {
        LUNA_HIP(hipMemcpyAsync(keys, fv->keys.p, sizeof(long long) * fv->n, hipMemcpyDeviceToHost, fv->stream));
        for (int s = 0; s < fv->n; ++s) cells[s] = s;
    }
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

int luna_fv1_mesh_stats(luna_fv1* fv, long long* updates, long long* sorts) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_mesh_stats: null handle");
    if (updates) *updates = fv->updates;
    if (sorts) *sorts = fv->sorts;
    return LUNA_OK;
}

int luna_fv1_destroy(luna_fv1* fv) {
    if (!fv) return LUNA_OK;
    if (fv->stream) (void)hipStreamSynchronize(fv->stream);
    fv->line_key[0].release();
    fv->line_key[1].release();
    fv->line_id.release();
    fv->x.release();
    fv->x_next.release();
    fv->ends.release();
    fv->inflow_dev.release();
    fv->descents.release();
    fv->packed.release();
    fv->sort_temp.release();
    if (fv->host_ends) (void)hipHostFree(fv->host_ends);
    if (fv->host_descents) (void)hipHostFree(fv->host_descents);
    fv->keys.release();
    fv->q.release();
    fv->next.release();
    fv->dq.release();
    fv->dnext.release();
    fv->counters.release();
    if (fv->host) (void)hipHostFree(fv->host);
    if (fv->stream) (void)hipStreamDestroy(fv->stream);
    delete fv;
    return LUNA_OK;
}

int luna_fv1_set_state(luna_fv1* fv, const double* q) {
    if (!fv || !q) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_set_state: null argument");
    LUNA_HIP(hipMemcpyAsync(fv->q.p, q, sizeof(double) * 3 * fv->n, hipMemcpyHostToDevice, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

int luna_fv1_get_state(luna_fv1* fv, double* q) {
    if (!fv || !q) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_get_state: null argument");
    LUNA_HIP(hipMemcpyAsync(q, fv->q.p, sizeof(double) * 3 * fv->n, hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

int luna_fv1_time_step(luna_fv1* fv, double courant, int rule, double* dt) {
    if (!fv || !dt) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_time_step: null argument");
    const int slot = fv->parity;
    LUNA_HIP(luna::launch_line_time_step(fv->stream, fv->args(), courant, rule == LUNA_STEP_LUNA, fv->counters_now()));
    fv->parity ^= 1;
    LUNA_HIP(hipMemcpyAsync(fv->host, fv->counters.p + slot, sizeof(unsigned long long), hipMemcpyDeviceToHost, fv->stream));
    if (fv->moving && fv->watching) {
        // finite_volume_1d.hpp:372-381: while gas enters, what the end slots ask their walls for
        const int rc = read_ends(fv);
        if (rc != LUNA_OK) return rc;
        fv->wants = static_cast<int>(fv->host_ends[1] + fv->host_ends[3]);
        fv->looked = true;
    }
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    *dt = bits_to_double(fv->host[0]);
    return LUNA_OK;
}

int luna_fv1_step(luna_fv1* fv, double dt) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_step: null argument");
    if (!fv->moving) {
        LUNA_HIP(luna::launch_line_step(fv->stream, fv->args(), dt, 0.0, false, fv->counters_now()));
        std::swap(fv->q, fv->next);
        return LUNA_OK;
    }
    // finite_volume_1d.hpp:389-407 (moving): the kernel, the order, and the walls' cells when due.
    unsigned long long descents = 0;
    int rc = moving_kernel(fv, dt, 0.0, false, &descents);
    if (rc != LUNA_OK) return rc;
    bool sorted = false, added = false;
    if ((rc = order_line(fv, descents, &sorted)) != LUNA_OK) return rc;
    if (due(fv) && (rc = supply_cells(fv, &added)) != LUNA_OK) return rc;
    fv->looked = false;
    return LUNA_OK;
}

int luna_fv1_step_tangent(luna_fv1* fv, double dt, double* dq) {
    if (!fv || !dq) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_step_tangent: null argument");
    if (fv->moving) return fail(LUNA_ERR_LOGIC, "hydro::FiniteVolume<1>::step_tangent: a static line");
    const size_t bytes = sizeof(double) * 3 * fv->n;
    LUNA_HIP(fv->dq.reserve(3 * static_cast<size_t>(fv->n)));
    LUNA_HIP(fv->dnext.reserve(3 * static_cast<size_t>(fv->n)));
    LUNA_HIP(hipMemcpyAsync(fv->dq.p, dq, bytes, hipMemcpyHostToDevice, fv->stream));
    LUNA_HIP(luna::launch_line_tangent(fv->stream, fv->args(), dt, fv->dq.p, fv->dnext.p));
    std::swap(fv->q, fv->next);
    LUNA_HIP(hipMemcpyAsync(dq, fv->dnext.p, bytes, hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

int luna_fv1_step_adjoint(luna_fv1* fv, double dt, double* qbar) {
    if (!fv || !qbar) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_step_adjoint: null argument");
    if (fv->moving) return fail(LUNA_ERR_LOGIC, "hydro::FiniteVolume<1>::step_adjoint: a static line");
    const size_t bytes = sizeof(double) * 3 * fv->n;
    LUNA_HIP(fv->dq.reserve(3 * static_cast<size_t>(fv->n)));
    LUNA_HIP(fv->dnext.reserve(3 * static_cast<size_t>(fv->n)));
    LUNA_HIP(hipMemcpyAsync(fv->dq.p, qbar, bytes, hipMemcpyHostToDevice, fv->stream));
    LUNA_HIP(luna::launch_line_adjoint(fv->stream, fv->args(), dt, fv->dq.p, fv->dnext.p));
    LUNA_HIP(hipMemcpyAsync(qbar, fv->dnext.p, bytes, hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

int luna_fv1_step_and_time_step(luna_fv1* fv, double dt, double courant, double* next) {
    if (!fv || !next) return fail(LUNA_ERR_ARGUMENT, "luna_fv1_step_and_time_step: null argument");
    const int slot = fv->parity;
    if (fv->moving) {
        // finite_volume_1d.hpp:679-708 (moving): the ends' wishes on the line before the step, the fused kernel, the
        // order; after a sort or new cells the next dt is time_step(courant, luna) of the new line.
        int wants = 0, rc;
        if (fv->watching) {
            if ((rc = read_ends(fv)) != LUNA_OK) return rc;
            wants = static_cast<int>(fv->host_ends[1] + fv->host_ends[3]);
        }
        unsigned long long descents = 0;
        if ((rc = moving_kernel(fv, dt, courant, true, &descents)) != LUNA_OK) return rc;
        fv->parity ^= 1;
        LUNA_HIP(hipMemcpyAsync(fv->host, fv->counters.p + slot, sizeof(unsigned long long), hipMemcpyDeviceToHost, fv->stream));
        LUNA_HIP(hipStreamSynchronize(fv->stream));
        *next = bits_to_double(fv->host[0]);
        bool again = false, added = false;
        if ((rc = order_line(fv, descents, &again)) != LUNA_OK) return rc;
        fv->wants = wants;
        fv->looked = fv->watching;
        if (due(fv)) {
            if ((rc = supply_cells(fv, &added)) != LUNA_OK) return rc;
            again = again || added;
        }
        fv->looked = false;
        if (again) return luna_fv1_time_step(fv, courant, LUNA_STEP_LUNA, next);
        return LUNA_OK;
    }
    LUNA_HIP(luna::launch_line_step(fv->stream, fv->args(), dt, courant, true, fv->counters_now()));
    fv->parity ^= 1;
    std::swap(fv->q, fv->next);
    LUNA_HIP(hipMemcpyAsync(fv->host, fv->counters.p + slot, sizeof(unsigned long long), hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    *next = bits_to_double(fv->host[0]);
    return LUNA_OK;
}

int luna_fv_set_limiter_network(luna_fv* fv, const float* params, int count) {
    if (!fv) return fail(LUNA_ERR_ARGUMENT, "luna_fv_set_limiter_network: null handle");
    if (fv->opt.limiter != LUNA_LIMITER_LEARNED) return fail(LUNA_ERR_LOGIC, "luna_fv_set_limiter_network: the solver's limiter is not the learned one");
    if (!params) {
        fv->has_net = false;
        return LUNA_OK;
    }
    if (count != luna::PARAMS) return fail(LUNA_ERR_ARGUMENT, "luna_fv_set_limiter_network: count must be LUNA_LIMITER_PARAMS");
    for (int k = 0; k < count; ++k)
        if (!std::isfinite(params[k])) return fail(LUNA_ERR_ARGUMENT, "luna_fv_set_limiter_network: parameters must be finite");
    const std::vector<unsigned char> bytes = network_bytes(params);
    LUNA_HIP(hipMemcpyAsync(fv->net.p, bytes.data(), bytes.size(), hipMemcpyHostToDevice, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    fv->has_net = true;
    fv->swept = false;  // the next step sweeps (and limits) again
    return LUNA_OK;
}

int luna_fv_limiter_features(luna_fv* fv, unsigned short* features, unsigned char* troubled) {
    if (!fv || !features || !troubled) return fail(LUNA_ERR_ARGUMENT, "luna_fv_limiter_features: null argument");
    if (fv->opt.limiter != LUNA_LIMITER_LEARNED) return fail(LUNA_ERR_LOGIC, "luna_fv_limiter_features: the solver's limiter is not the learned one");
    if (!fv->features_ready) return fail(LUNA_ERR_LOGIC, "luna_fv_limiter_features: no sweep yet (call luna_fv_time_step)");
    const size_t rows = static_cast<size_t>(fv->N) * fv->V;
    LUNA_HIP(hipMemcpyAsync(features, fv->features.p, sizeof(unsigned short) * rows * luna::FEATURES, hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipMemcpyAsync(troubled, fv->troubled.p, fv->N, hipMemcpyDeviceToHost, fv->stream));
    LUNA_HIP(hipStreamSynchronize(fv->stream));
    return LUNA_OK;
}

}  // extern "C"

// The network's trainer: FP32 parameters with Adam's moments, the BF16 network copy the kernels read, the per-wavefront
// partial sums.
struct luna_mlp {
    hipStream_t stream = nullptr;
    Buf<float> master, m, v, partials, out;
    Buf<unsigned char> net;
    int t = 0;
};

namespace {

int mlp_batch(luna_mlp* mlp, const unsigned short* x, const float* target, const float* weight, int rows, double weight_sum,
              luna::Batch& b, int& waves, const char* what) {
    if (!mlp || !x || !target || !weight) return fail(LUNA_ERR_ARGUMENT, std::string(what) + ": null argument");
    if (rows <= 0) return fail(LUNA_ERR_ARGUMENT, std::string(what) + ": rows must be positive");
    if (!(weight_sum > 0) || !std::isfinite(weight_sum)) return fail(LUNA_ERR_ARGUMENT, std::string(what) + ": weight_sum must be positive");
    b = luna::Batch{x, target, weight, rows, static_cast<float>(1.0 / weight_sum)};
    waves = luna::mlp_waves(rows);
    LUNA_HIP(mlp->partials.reserve(static_cast<size_t>(waves) * luna::PARTIAL_ROW));
    return LUNA_OK;
}

}  // namespace

extern "C" {

int luna_mlp_create(luna_mlp** out) {
    if (!out) return fail(LUNA_ERR_ARGUMENT, "luna_mlp_create: null argument");
    *out = nullptr;
    luna_mlp* mlp = new luna_mlp;
    hipError_t e = hipStreamCreate(&mlp->stream);
    if (e == hipSuccess) e = mlp->master.reserve(luna::PARAMS);
    if (e == hipSuccess) e = mlp->m.reserve(luna::PARAMS);
    if (e == hipSuccess) e = mlp->v.reserve(luna::PARAMS);
    if (e == hipSuccess) e = mlp->out.reserve(luna::PARAMS + 1);
    if (e == hipSuccess) e = mlp->net.reserve(NET_BYTES);
    if (e != hipSuccess) {
        luna_mlp_destroy(mlp);
        return fail(LUNA_ERR_MEMORY, std::string("luna_mlp_create: ") + hipGetErrorString(e));
    }
    *out = mlp;
    return LUNA_OK;
}

int luna_mlp_destroy(luna_mlp* mlp) {
    if (!mlp) return LUNA_OK;
    if (mlp->stream) (void)hipStreamSynchronize(mlp->stream);
    for (auto* b : {&mlp->master, &mlp->m, &mlp->v, &mlp->partials, &mlp->out}) b->release();
    mlp->net.release();
    if (mlp->stream) (void)hipStreamDestroy(mlp->stream);
    delete mlp;
    return LUNA_OK;
}

int luna_mlp_set_params(luna_mlp* mlp, const float* params, int count) {
    if (!mlp || !params) return fail(LUNA_ERR_ARGUMENT, "luna_mlp_set_params: null argument");
    if (count != luna::PARAMS) return fail(LUNA_ERR_ARGUMENT, "luna_mlp_set_params: count must be LUNA_LIMITER_PARAMS");
    const std::vector<unsigned char> bytes = network_bytes(params);
    const std::vector<float> zero(luna::PARAMS, 0.0f);
    LUNA_HIP(hipMemcpyAsync(mlp->master.p, params, sizeof(float) * luna::PARAMS, hipMemcpyHostToDevice, mlp->stream));
    LUNA_HIP(hipMemcpyAsync(mlp->m.p, zero.data(), sizeof(float) * luna::PARAMS, hipMemcpyHostToDevice, mlp->stream));
    LUNA_HIP(hipMemcpyAsync(mlp->v.p, zero.data(), sizeof(float) * luna::PARAMS, hipMemcpyHostToDevice, mlp->stream));
    LUNA_HIP(hipMemcpyAsync(mlp->net.p, bytes.data(), bytes.size(), hipMemcpyHostToDevice, mlp->stream));
    LUNA_HIP(hipStreamSynchronize(mlp->stream));
    mlp->t = 0;
    return LUNA_OK;
}

int luna_mlp_get_params(luna_mlp* mlp, float* params, int count) {
    if (!mlp || !params) return fail(LUNA_ERR_ARGUMENT, "luna_mlp_get_params: null argument");
    if (count != luna::PARAMS) return fail(LUNA_ERR_ARGUMENT, "luna_mlp_get_params: count must be LUNA_LIMITER_PARAMS");
    LUNA_HIP(hipMemcpyAsync(params, mlp->master.p, sizeof(float) * luna::PARAMS, hipMemcpyDeviceToHost, mlp->stream));
    LUNA_HIP(hipStreamSynchronize(mlp->stream));
    return LUNA_OK;
}

int luna_mlp_forward(luna_mlp* mlp, const unsigned short* x, int rows, float* fraction) {
    if (!mlp || !x || !fraction) return fail(LUNA_ERR_ARGUMENT, "luna_mlp_forward: null argument");
    if (rows < 0) return fail(LUNA_ERR_ARGUMENT, "luna_mlp_forward: rows must not be negative");
    LUNA_HIP(luna::launch_limiter_apply(mlp->stream, network_at(mlp->net.p), x, nullptr, rows, 1, 1, 0, nullptr, fraction));
    LUNA_HIP(hipStreamSynchronize(mlp->stream));
    return LUNA_OK;
}

int luna_mlp_gradient(luna_mlp* mlp, const unsigned short* x, const float* target, const float* weight, int rows, double weight_sum,
                      float* grad, float* loss) {
    if (!grad || !loss) return fail(LUNA_ERR_ARGUMENT, "luna_mlp_gradient: null argument");
    luna::Batch b{};
    int waves = 0;
    if (int rc = mlp_batch(mlp, x, target, weight, rows, weight_sum, b, waves, "luna_mlp_gradient")) return rc;
    LUNA_HIP(luna::launch_mlp_partials(mlp->stream, network_at(mlp->net.p), b, mlp->partials.p, waves));
    LUNA_HIP(luna::launch_mlp_reduce(mlp->stream, mlp->partials.p, waves, mlp->out.p, nullptr));
    LUNA_HIP(hipMemcpyAsync(grad, mlp->out.p, sizeof(float) * luna::PARAMS, hipMemcpyDeviceToHost, mlp->stream));
    LUNA_HIP(hipMemcpyAsync(loss, mlp->out.p + luna::PARAMS, sizeof(float), hipMemcpyDeviceToHost, mlp->stream));
    LUNA_HIP(hipStreamSynchronize(mlp->stream));
    return LUNA_OK;
}

int luna_mlp_train_step(luna_mlp* mlp, const unsigned short* x, const float* target, const float* weight, int rows, double weight_sum,
                        float lr, float* loss) {
    luna::Batch b{};
    int waves = 0;
    if (int rc = mlp_batch(mlp, x, target, weight, rows, weight_sum, b, waves, "luna_mlp_train_step")) return rc;
    if (!(lr > 0) || !std::isfinite(lr)) return fail(LUNA_ERR_ARGUMENT, "luna_mlp_train_step: lr must be positive");
    mlp->t += 1;
    const double beta1 = 0.9, beta2 = 0.999;
    luna::Adam a{};
    a.master = mlp->master.p;
    a.m = mlp->m.p;
    a.v = mlp->v.p;
    a.lr = lr;
    a.beta1 = static_cast<float>(beta1);
    a.beta2 = static_cast<float>(beta2);
    a.eps = 1e-8f;
    a.c1 = static_cast<float>(1.0 / (1.0 - std::pow(beta1, mlp->t)));
    a.c2 = static_cast<float>(1.0 / (1.0 - std::pow(beta2, mlp->t)));
    a.w1 = reinterpret_cast<unsigned short*>(mlp->net.p);
    a.w2 = reinterpret_cast<unsigned short*>(mlp->net.p + NET_W2);
    a.rest = reinterpret_cast<float*>(mlp->net.p + NET_REST);
    LUNA_HIP(luna::launch_mlp_partials(mlp->stream, network_at(mlp->net.p), b, mlp->partials.p, waves));
    LUNA_HIP(luna::launch_mlp_reduce(mlp->stream, mlp->partials.p, waves, mlp->out.p, &a));
    if (loss) {
        LUNA_HIP(hipMemcpyAsync(loss, mlp->out.p, sizeof(float), hipMemcpyDeviceToHost, mlp->stream));
        LUNA_HIP(hipStreamSynchronize(mlp->stream));
    }
    return LUNA_OK;
}

int luna_mlp_synchronize(luna_mlp* mlp) {
    if (!mlp) return fail(LUNA_ERR_ARGUMENT, "luna_mlp_synchronize: null handle");
    LUNA_HIP(hipStreamSynchronize(mlp->stream));
    return LUNA_OK;
}

}  // extern "C"
