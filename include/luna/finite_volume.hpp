// C++ facade of libluna_hydro.so: the member functions of luna-cpp's hydro::FiniteVolume<D> (src/hydro/finite_volume.hpp)
// and FiniteVolume<1> (src/hydro/finite_volume_1d.hpp), with the same names and argument meanings, over the C-ABI of
// luna_hydro.h. Errors throw what the reference throws: std::invalid_argument for LUNA_ERR_ARGUMENT, std::logic_error for
// LUNA_ERR_LOGIC, std::runtime_error otherwise. Views are host std::vectors here (the reference returns Kokkos views);
// state_device() gives the device array. Header-only; link with -lluna_hydro.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../luna_hydro.h"

namespace luna {

inline void check(int rc, const char* what) {
    if (rc == LUNA_OK) return;
    const std::string msg = std::string(what) + ": " + luna_last_error();
    if (rc == LUNA_ERR_ARGUMENT) throw std::invalid_argument(msg);
    if (rc == LUNA_ERR_LOGIC) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}

// hydro::Gas (euler.hpp:82-101).
struct Gas {
    double gamma = 5.0 / 3.0, smallr = 1e-10, smallc = 1e-10;
};

// learned: this library's learned limiter (no counterpart in the reference; 2D/3D, order 1; luna_hydro.h).
enum class Limiter { face_centroid = LUNA_LIMITER_FACE_CENTROID, none = LUNA_LIMITER_NONE, minmod = LUNA_LIMITER_MINMOD,
                     van_leer = LUNA_LIMITER_VAN_LEER, learned = LUNA_LIMITER_LEARNED };
enum class Motion { static_mesh = LUNA_MOTION_STATIC, moving = LUNA_MOTION_MOVING };
enum class Riemann { hllc = LUNA_RIEMANN_HLLC, hll = LUNA_RIEMANN_HLL };
enum class StepRule { arepo = LUNA_STEP_AREPO, luna = LUNA_STEP_LUNA };
// The gradient slot of hydro::Precision (precision.hpp:41-54): Precision<>, Grad32, Grad16. A run-time choice here.
enum class Precision { double_ = LUNA_PRECISION_DOUBLE, grad32 = LUNA_PRECISION_GRAD32, grad16 = LUNA_PRECISION_GRAD16 };
// hydro::Update (finite_volume.hpp:25): the per-face loop or the per-cell one (fluxes_and_update).
enum class Update { per_face = LUNA_UPDATE_PER_FACE, per_cell = LUNA_UPDATE_PER_CELL };

// hydro::FiniteVolume<D> for D = 2, 3 (Update::per_face; the precision policy a constructor argument, double by default).
template <int D>
class FiniteVolume {
public:
    static_assert(D == 2 || D == 3, "luna::FiniteVolume: D = 1 is luna::Line");
    static constexpr int V = D + 2;

    // FiniteVolume(exec, mesh, gas, order, update, limiter, motion) (finite_volume.hpp:79-107); mesh: the voronoi-cpp
    // output contract as luna_geometry.
    FiniteVolume(const luna_geometry& mesh, const Gas& gas, int order, Limiter limiter = Limiter::face_centroid,
                 Motion motion = Motion::static_mesh, Riemann riemann = Riemann::hllc, Precision precision = Precision::double_,
                 double hll_compression = 0,  // the HLL fallback at extreme compression (luna_hydro.h; none in the reference)
                 Update update = Update::per_face,
                 int sweep_lanes = 1)  // LUNA_SWEEP_LANES (LaneSweep, :880-1037): 1, 2, 4 or 8
        : n_(mesh.cells) {
        luna_fv_options o{};
        o.dim = D;
        o.order = order;
        o.limiter = static_cast<int>(limiter);
        o.riemann = static_cast<int>(riemann);
        o.motion = static_cast<int>(motion);
        o.step_rule = LUNA_STEP_AREPO;
        o.gamma = gas.gamma;
        o.smallr = gas.smallr;
        o.smallc = gas.smallc;
        o.precision = static_cast<int>(precision);
        o.hll_compression = hll_compression;
        o.update = static_cast<int>(update);
        o.sweep_lanes = sweep_lanes;
        options_ = o;
        check(luna_fv_create(&o, &mesh, &fv_), "hydro::FiniteVolume");
    }
    FiniteVolume(const FiniteVolume&) = delete;
    FiniteVolume& operator=(const FiniteVolume&) = delete;
    ~FiniteVolume() { luna_fv_destroy(fv_); }

    int cells() const { return n_; }
    int faces() const {
        int f = 0;
        check(luna_fv_faces(fv_, &f), "faces");
        return f;
    }
    // state() (:145-148): here a host copy, and set_state() to write it.
    std::vector<double> state() const {
        std::vector<double> q(static_cast<size_t>(n_) * V);
        check(luna_fv_get_state(fv_, q.data()), "state");
        return q;
    }
    void set_state(const std::vector<double>& q) { check(luna_fv_set_state(fv_, q.data()), "set_state"); }
    // The learned limiter's network (LUNA_LIMITER_PARAMS FP32 values; empty: none, the face-centroid limiter's gradients).
    void set_limiter_network(const std::vector<float>& params) {
        check(luna_fv_set_limiter_network(fv_, params.empty() ? nullptr : params.data(), static_cast<int>(params.size())),
              "set_limiter_network");
    }
    double* state_device() {
        double* q = nullptr;
        check(luna_fv_state_device(fv_, &q), "state_device");
        return q;
    }
    // time_step(courant) (:190-201).
    double time_step(double courant) {
        double dt = 0;
        check(luna_fv_time_step(fv_, courant, &dt), "time_step");
        return dt;
    }
    // set_step_rule (:181), release_mesh (:134-139: the handle copies a device mesh's face arrays, which it reads in place,
    // and lets the mesh go; refused on a moving mesh), and of
    // mesh_stats (:160-163) the mesh updates taken.
    void set_step_rule(StepRule rule) { check(luna_fv_set_step_rule(fv_, static_cast<int>(rule)), "set_step_rule"); }
    void release_mesh() { check(luna_fv_release_mesh(fv_), "release_mesh"); }
    long long mesh_updates() const {
        long long n = 0;
        check(luna_fv_mesh_stats(fv_, &n), "mesh_stats");
        return n;
    }
    // Derivatives of one static order-1 step at fixed dt (:1941-2231): step_tangent advances Q and dq (dq becomes J dq);
    // step_adjoint maps qbar after the step to J^T qbar before it (Q stays); time_step_tangent is time_step() with the
    // fit's tangent for dq, which the next step_tangent with the same dq reuses.
    void step_tangent(double dt, std::vector<double>& dq) { check(luna_fv_step_tangent(fv_, dt, dq.data(), 0), "step_tangent"); }
    void step_adjoint(double dt, std::vector<double>& qbar) { check(luna_fv_step_adjoint(fv_, dt, qbar.data(), 0), "step_adjoint"); }
    double time_step_tangent(double courant, const std::vector<double>& dq) {
        double dt = 0;
        check(luna_fv_time_step_tangent(fv_, courant, dq.data(), 0, &dt), "time_step_tangent");
        return dt;
    }
    // step(dt) (:205-226) on a static mesh. On a moving mesh the reference's step() calls the mesh library itself; here the
    // caller passes a producer: std::vector<double> positions -> luna_geometry (kept alive by the caller until return).
    void step(double dt) { check(luna_fv_step(fv_, dt), "step"); }
    template <class MeshOf>
    void step(double dt, MeshOf&& mesh_of) {
        std::vector<double> x(static_cast<size_t>(n_) * D);
        check(luna_fv_moving_begin(fv_, dt, x.data()), "step (moving: begin)");
        const luna_geometry& next = mesh_of(x);
        check(luna_fv_moving_finish(fv_, &next, dt), "step (moving: finish)");
    }
    long long fallbacks() const {
        long long c = 0;
        check(luna_fv_fallbacks(fv_, &c), "fallbacks");
        return c;
    }
    std::vector<double> positions() const {
        std::vector<double> x(static_cast<size_t>(n_) * D);
        check(luna_fv_positions(fv_, x.data()), "positions");
        return x;
    }
    std::vector<double> point_velocities() const {
        std::vector<double> w(static_cast<size_t>(n_) * D);
        check(luna_fv_point_velocities(fv_, w.data()), "point_velocities");
        return w;
    }
    // density_gradients() (:165-180): the first D entries of each cell's gradient row.
    std::vector<double> density_gradients() const {
        std::vector<double> g(static_cast<size_t>(n_) * V * D), out(static_cast<size_t>(n_) * D);
        check(luna_fv_gradients(fv_, g.data()), "density_gradients");
        for (int i = 0; i < n_; ++i)
            for (int a = 0; a < D; ++a) out[static_cast<size_t>(i) * D + a] = g[static_cast<size_t>(i) * V * D + a];
        return out;
    }
    // mesh_moved() (:152-156): the caller changed the mesh itself.
    void mesh_moved(const luna_geometry& mesh) { check(luna_fv_set_geometry(fv_, &mesh), "mesh_moved"); }
    // A moving step with a mesh producer that leaves its arrays on the GPU (voronoi-cpp's device views, read in place as
    // :257-327 reads them): mesh_of(positions, host copy) -> luna_geometry of device pointers. The new mesh's face arrays
    // are read by the steps that follow, so the producer keeps them until it hands over the next mesh.
    template <class DeviceMeshOf>
    void step_device(double dt, DeviceMeshOf&& mesh_of) {
        std::vector<double> x(static_cast<size_t>(n_) * D);
        check(luna_fv_moving_begin(fv_, dt, x.data()), "step (moving: begin)");
        const luna_geometry& next = mesh_of(x);
        check(luna_fv_set_geometry_device(fv_, &next), "step (moving: the new mesh)");
        check(luna_fv_moving_finish(fv_, nullptr, dt), "step (moving: finish)");
    }
    void mesh_moved_device(const luna_geometry& mesh) { check(luna_fv_set_geometry_device(fv_, &mesh), "mesh_moved"); }
    luna_fv* handle() { return fv_; }

private:
    luna_fv* fv_ = nullptr;
    luna_fv_options options_{};
    int n_ = 0;
};

// FiniteVolume<1> (finite_volume_1d.hpp:41-53) on a sorted line: static (the line's keys) or moving (the generators'
// positions; the walls supply cells, so cells() grows; the accessors of :57-71).
class Line {
public:
    Line(const luna_line& line, const Gas& gas, int order, Limiter limiter = Limiter::face_centroid, Riemann riemann = Riemann::hllc) {
        const luna_fv1_options o = options(gas, order, limiter, riemann);
        check(luna_fv1_create(&o, &line, &fv_), "hydro::FiniteVolume<1>");
    }
    Line(const luna_moving_line& line, const Gas& gas, int order, Limiter limiter = Limiter::face_centroid, Riemann riemann = Riemann::hllc) {
        const luna_fv1_options o = options(gas, order, limiter, riemann);
        check(luna_fv1_create_moving(&o, &line, &fv_), "hydro::FiniteVolume<1> (moving)");
    }
    Line(const Line&) = delete;
    Line& operator=(const Line&) = delete;
    ~Line() { luna_fv1_destroy(fv_); }
    int cells() const {
        int n = 0;
        check(luna_fv1_cells(fv_, &n, nullptr), "cells");
        return n;
    }
    long long cells_added() const {
        long long added = 0;
        check(luna_fv1_cells(fv_, nullptr, &added), "cells_added");
        return added;
    }
    std::vector<double> state() const {
        std::vector<double> q(static_cast<size_t>(cells()) * 3);
        check(luna_fv1_get_state(fv_, q.data()), "state");
        return q;
    }
    void set_state(const std::vector<double>& q) { check(luna_fv1_set_state(fv_, q.data()), "set_state"); }
    std::vector<double> positions() const {
        std::vector<double> x(static_cast<size_t>(cells()));
        check(luna_fv1_positions(fv_, x.data()), "positions");
        return x;
    }
    // The line's slots in key order: their keys and cells (Mesh<1>::line()).
    void line(std::vector<long long>& keys, std::vector<int>& cells_of_slots) const {
        keys.resize(static_cast<size_t>(cells()));
        cells_of_slots.resize(keys.size());
        check(luna_fv1_line(fv_, keys.data(), cells_of_slots.data()), "line");
    }
    void set_inflow_spacing(double spacing) { check(luna_fv1_set_inflow_spacing(fv_, spacing), "set_inflow_spacing"); }
    double time_step(double courant, StepRule rule = StepRule::arepo) {
        double dt = 0;
        check(luna_fv1_time_step(fv_, courant, static_cast<int>(rule), &dt), "time_step");
        return dt;
    }
    void step(double dt) { check(luna_fv1_step(fv_, dt), "step"); }
    double step_and_time_step(double dt, double courant) {
        double next = 0;
        check(luna_fv1_step_and_time_step(fv_, dt, courant, &next), "step_and_time_step");
        return next;
    }

private:
    static luna_fv1_options options(const Gas& gas, int order, Limiter limiter, Riemann riemann) {
        luna_fv1_options o{};
        o.order = order;
        o.limiter = static_cast<int>(limiter);
        o.riemann = static_cast<int>(riemann);
        o.gamma = gas.gamma;
        o.smallr = gas.smallr;
        o.smallc = gas.smallc;
        return o;
    }
    luna_fv1* fv_ = nullptr;
};

}  // namespace luna
