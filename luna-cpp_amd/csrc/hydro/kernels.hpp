// Device kernels of the hydro step on a Voronoi mesh (D = 2, 3) and on a sorted line (D = 1), their argument records and
// the row layouts in HBM. Launched by api.hip on the handle's stream. FP64 throughout, -ffp-contract=off (physics.hpp).
#pragma once

#include "learned.hpp"
#include "physics.hpp"

namespace luna {

constexpr unsigned long long INF_BITS = 0x7FF0000000000000ull;  // +inf as the bits a min-reduction starts from

// Row layouts (doubles per row), each row aligned to its size:
//   cell row     [x (D), c (D), V, r | W (D + 2), c_s], 16 (128 B, one L2 line). c: centroid relative to the generator;
//                r: cbrt(3V/4pi) / sqrt(V/pi). The geometry half (first 8) comes with the mesh; the primitive half is
//                written by the primitives and by the update. A neighbour's position, centroid and state cost one line.
//   face record  [A (D), fc (D), |A|, (owner, neighbour) as two int32 in the last slot], 8 (64 B)
//   gradient row [g (D + 2) x D], padded to 16 (3D) or 8 (2D)
//   flux row     [F (D + 2)] per face, in face order
constexpr int CELL_ROW = 16, GEO_ROW = 8, FACE_ROW = 8;
template <int D>
constexpr int grad_row() {
    return D == 3 ? 16 : 8;
}
template <int D>
struct CellAt {
    static constexpr int X = 0, C = D, VOL = 2 * D, RAD = 2 * D + 1, W = GEO_ROW, CS = GEO_ROW + D + 2;
};
template <int D>
struct FaceAt {
    static constexpr int A = 0, FC = D, MAG = 2 * D, PAIR = 7;
};

// Face lists in slices of 64 cells (one wavefront): entry t of cell i sits in slot slice[i / 64] + 64 t + i % 64, so the
// t-th entries of a wavefront's cells are one contiguous run; a slice is as long as its longest list (the rest is padding
// that no kernel reads). Within a cell the entries keep the mesh's list order.
constexpr int SLICE = 64;

// The mesh as the kernels read it (device pointers): finite_volume.hpp:257-327 Geometry, packed.
struct Geo {
    int N, F;
    const int* count;   // N: list length of each cell
    const int* slice;   // ceil(N / 64): first slot of each slice
    const int* list;    // slots: f if the cell owns face f, ~f otherwise
    const int* other;   // slots: the cell on the other side of the entry's face
    double* cell;       // N x CELL_ROW (the kernels write only the primitive half)
    const double* volume;  // N: the cell rows' volumes again, densely, for the kernels that read no other geometry
    const double* face; // F x FACE_ROW records, for the derivatives (built on their first use of a mesh)
    // The mesh's face arrays, read in place as the reference reads the mesh's views (finite_volume.hpp:257-327): owner
    // and neighbour (F), area vectors and face centroids relative to the owner's generator (F x D)
    const int* owner;
    const int* neighbour;
    const double* area;
    const double* fcen;
    // per slot, in the slices' layout, D + 1 planes of `slots` doubles: the face's |A|, then the face centroid relative
    // to this cell's centroid (Geometry::face_offset, finite_volume.hpp:322-326), one plane per axis
    const double* slot;
    long long slots;
    double side[3], inverse[3];  // the periodic box and 1 / side
};

// The mesh producer's arrays as the reference reads them (device pointers; the voronoi-cpp output contract,
// finite_volume.hpp:257-327): the input of the geometry pass that derives Geo's rows (geometry_kernels.hip).
struct MeshIn {
    int N, F;
    const int* owner;
    const int* neighbour;
    const int* start;          // N + 1
    const int* list;           // 2 F entries, f or ~f
    const double* area;        // F x D
    const double* face_centroid;  // F x D, relative to the owner's generator
    const double* centroid;    // N x D, relative to the cell's generator
    const double* volume;      // N
    const double* positions;   // N x D
    double side[3], inverse[3];
};
// Flags the geometry pass raises on a malformed mesh (it then reads nothing out of range and the host refuses the mesh).
constexpr unsigned int GEO_BAD_START = 1, GEO_BAD_CELL = 2, GEO_BAD_FACE = 4, GEO_TOO_MANY_SLOTS = 8;

// Per-cell rows of the step (device pointers).
struct Rows {
    double* q;             // N x V, extensive
    double* grad;          // N x grad_row<D>()
    double* w;             // N x D point velocities (moving)
    double* x;             // N x D solver positions (moving)
    double* old_centroid;  // N x D centroids of t^n (moving; written by the sweep)
    int* steps;            // N x D whole grid steps over the step (moving)
    double* flux;          // F x V, face order
    unsigned short* features;  // N V x 16 BF16 (the learned limiter's, recorded by the sweep with limit 2)
    unsigned char* troubled;   // N (the same)
    // The precision policy of the gradient rows (precision.hpp:26-54): 0 double (rows of grad_row<D>() doubles); 1 grad32
    // (fit, limiter and face-state change in float, rows of lp_row<D>() floats); 2 grad16 (the same arithmetic, rows of
    // 16 FP16 values with one power-of-two exponent per group rho, v, P in `exponents`, int8 each). The rows share the
    // double rows' allocation.
    int precision;
    unsigned int* exponents;  // N (grad16)
    // The HLL fallback at extreme compression (no counterpart in the reference): the sweep flags cell i when
    // r_i (div v)_i / c_s,i < -hll_threshold, div v from the limited gradient; faces with a flagged side take HLL
    // instead of HLLC. Null when off.
    unsigned char* hll;
    double hll_threshold;
    // LaneSweep (finite_volume.hpp:880-1037): 0 or 1 the one-lane sweep; 2, 4 or 8 lanes per cell (double policy, order 1,
    // the face-centroid limiter or none).
    int sweep_lanes;
};
template <int D>
constexpr int lp_row() {  // row_width<float>(V D): whole 32-byte sectors
    return D == 3 ? 16 : 8;
}
constexpr int HALF_ROW = 16;  // row_width<uint16_t>(V D) in 2D and 3D: 32 bytes

struct Grid {
    double step, per_step;  // grid step and its inverse
    long long period[3];
};

struct Counters {
    unsigned long long* dt;         // the min-reduction this sweep writes
    unsigned long long* dt_next;    // the other slot, reset for the next sweep
    unsigned long long* fallbacks;  // [2]: face sides that took the cell's state, per flux pass of the step
};

// Kernel launchers (api.hip calls these; each returns the HIP launch status).
template <int D>
hipError_t launch_primitives(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas);
// limit: 0 none, 1 the face-centroid limiter, 2 the same recording the learned limiter's features (r.features, r.troubled).
template <int D>
hipError_t launch_sweep(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double courant, bool moving, bool grad, int limit,
                        int rule, const Counters& c);
// pass: -1 static, 0 a moving step's first pass (mesh of t^n), 1 its second (mesh of t^n+1).
template <int D>
hipError_t launch_flux(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, int pass, double dt, double unit,
                       unsigned long long* fallbacks);
// drift: a moving step's first pass, which also drifts the positions on the grid (finite_volume.hpp:1219-1231).
template <int D>
hipError_t launch_update(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double dt, bool next_prims, const Grid* drift);
// Update::per_cell (finite_volume.hpp:1544-1566): the fluxes and the update in one per-cell kernel (each face's flux
// computed by both its cells, the same bits as the per-face loop); pass, order and flux_dt as launch_flux, dt the update's,
// drift as launch_update. It leaves the next primitives to the next time_step.
template <int D>
hipError_t launch_fluxes_and_update(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, int pass, double flux_dt,
                                    double dt, double unit, const Grid* drift, unsigned long long* fallbacks);
template <int D>
hipError_t launch_keep_centroids(hipStream_t s, const Geo& g, const Rows& r, const Grid& grid, double dt);
// move() off the grid (order 0 on a moving mesh): x += dt w, wrapped into the box.
template <int D>
hipError_t launch_drift(hipStream_t s, const Geo& g, const Rows& r, double dt);

// The geometry pass (geometry_kernels.hip), in this order on one stream: cells (the geometry half of the cell rows, the
// list lengths, each slice's longest list into slice), scan (slice's first slots, the total into info[0]), pairs (the
// faces' owner and neighbour checked); the host then checks the flags and sizes the slot arrays; slots (lists, other
// cells, |A| and face offsets). faces (the F x FACE_ROW records) runs on a derivative's first use of a mesh.
template <int D>
hipError_t launch_geometry_cells(hipStream_t s, const MeshIn& m, double* cell, double* volume, int* count, int* slice, unsigned int* flags);
hipError_t launch_geometry_scan(hipStream_t s, int* slice, int S, long long* info, unsigned int* flags);
template <int D>
hipError_t launch_geometry_faces(hipStream_t s, const MeshIn& m, double* face, unsigned int* flags);
hipError_t launch_geometry_pairs(hipStream_t s, const MeshIn& m, unsigned int* flags);
template <int D>
hipError_t launch_geometry_slots(hipStream_t s, const MeshIn& m, const int* count, const int* slice, const long long* info, int* list,
                                 int* other, double* slot, unsigned int* flags);

// Derivatives of the static order-1 step (HLLC or HLL) at fixed dt (adjoint_kernels.hip). Device rows:
//   dq       N x V   in: the tangent of Q (tangent) or Q bar after the step (adjoint); the tangent advances it to J dq
//   dW       N x V   W tangents (tangent) or W bars (adjoint)
//   dgrad    N x grad_row<D>()   gradient tangents
//   dflux    F x 2V  each face's flux row and its tangent, side by side
//   sidebar  F x 2V  the bars of each face's two sides
//   bbar     N x V (D + 1), boundbar N x V, extreme N x V: each cell's fit adjoint for its neighbours
//   out      N x V   J^T Q bar (adjoint)
struct Derivs {
    double* dq;
    double* dW;
    double* dgrad;
    double* dflux;
    double* sidebar;
    double* bbar;
    double* boundbar;
    int* extreme;
    double* out;
};
// The tangent writes Q' into r.q, each face's flux and its tangent into dflux and the limited gradients into r.grad
// (fitted: time_step_tangent has written the gradients, dW and dgrad already); the adjoint needs the primitive rows and the
// limited gradients of Q (the sweep) and leaves Q as it is.
template <int D>
hipError_t launch_step_tangent(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, double dt, bool limit, const Derivs& w,
                               bool fitted = false);
// The primitives' and the limited fit's tangents (and the forward gradients), as time_step_tangent takes them.
template <int D>
hipError_t launch_sweep_tangent(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, bool limit, const Derivs& w);
template <int D>
hipError_t launch_step_adjoint(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, double dt, bool limit, const Derivs& w);

// The moving step's reverse (order 1, fixed topology): device work arrays. Per face of the pass's mesh: row
// (pass_row), Ab, Xfb (A's and x_f's bars); per cell: W, g, w-hat, c^n bars (summed over both passes), the mesh's c and x
// bars, the fit's beta, bound bar and extreme face, V's bar and the Q bar out; per slot the entry rows; per cell the
// angle face and its bars.
struct MovingBars {
    double *row, *Ab, *Xfb, *Wbar, *gbar, *whbar, *c0bar, *cbar, *xbar, *beta, *boundbar, *entry, *angle, *Vbar, *out;
    int *extreme, *fmax_of;
};
// One pass on the mesh the Geo holds (second: t^n+1): the faces' rows and A, x_f bars (written), then the cells' gathers
// (added to W, g, w-hat, c^n, c and x bars). order0: the order-0 moving step's single pass (sides W, the sweep's w for
// w-hat, Q moved by the whole dt; no gradient bars).
template <int D>
hipError_t launch_moving_pass_adjoint(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, double dt, double unit, bool second,
                                      const double* qbar, const MovingBars& b, bool order0 = false);
// w-hat bar += dt x bar (n entries).
hipError_t launch_drift_adjoint(hipStream_t s, long long n, double dt, const double* xbar, double* whbar);
// On t^n: the point velocity's and the fit's adjoints per cell, the neighbours' gather with the primitive adjoint (out, V
// bar), the faces' A and x_f bars (added).
template <int D>
hipError_t launch_moving_cell_adjoint(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, bool limit, bool steps_fixed, const double* qbar,
                                      const MovingBars& b);

// The moving step's tangent (order 0 and 1, fixed topology): device rows. Per cell: dx (positions), dV, dc (the pass's mesh's
// geometry tangents), dc0 (c^n's: the first mesh's dc), dW (N x V), dcs (N), dw (point velocities, zero with the steps held),
// dg (N x V D limited gradients); per face of the pass's mesh: dA, dxf, and the dG rows.
struct MovingDots {
    double *dx, *dV, *dc, *dc0, *dA, *dxf, *dW, *dcs, *dw, *dg, *dG;
};
// On t^n: W and c_s tangents from dq and dV, the point velocities' and (fit) the limited gradients' tangents.
template <int D>
hipError_t launch_moving_tangent_cells(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, bool limit, bool fit, const double* dq,
                                       const MovingDots& t);
// One pass on the mesh the Geo holds (second: t^n+1, with t.dc0): the faces' dG, then dq -= dt/2 sum +-dG per cell in list
// order (order0: the order-0 step's one pass with the sweep's w and no gradients, dq -= dt sum).
template <int D>
hipError_t launch_moving_pass_tangent(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, double dt, double unit, bool second,
                                      bool order0, const MovingDots& t, double* dq);
// out = x + dt y (n entries).
hipError_t launch_axpy(hipStream_t s, long long n, double dt, const double* x, const double* y, double* out);

// D = 1: the sorted line. Slots in key order; Q (and on a moving line the positions) by cell. Walls: 0 reflective,
// 1 outflow, 2 inflow (a moving line's open wall holding the state inflow[side]).
struct LineArgs {
    int n, periodic, wall_left, wall_right, order, limiter, riemann;
    Gas gas;
    const long long* key;  // by slot
    long long period;
    double step;
    const double* q;  // by cell
    double* next;
    // moving line
    int moving, regularise;
    const int* id;  // slot -> cell; nullptr: slot s holds cell s
    double side;    // period * step
    const double* inflow;  // device: the left, then the right inflow wall's primitives (6)
    const double* x;  // positions by cell
    double* x_next;
    long long* key_next;  // by slot: the key of the slot's point after the step
};
hipError_t launch_line_time_step(hipStream_t s, const LineArgs& a, double courant, int luna_rule, const Counters& c);
hipError_t launch_line_step(hipStream_t s, const LineArgs& a, double dt, double courant, bool fused, const Counters& c);
// The moving step: Q, positions and keys after the step into a.next, a.x_next, a.key_next; the count of adjacent slots
// out of order after the move into *descents (*descents_next reset for the next step); fused, the next dt by LUNA's rule.
hipError_t launch_line_step_moving(hipStream_t s, const LineArgs& a, double dt, double courant, bool fused, const Counters& c,
                                   unsigned long long* descents, unsigned long long* descents_next);
// The ends of a moving line with walls, as the cell supply reads them (11 doubles): the left end generator's distance
// from its wall, the cells that end asks for, the same at the right, whether gas enters at an open end, the end cells'
// volumes, the end keys and the end cells.
constexpr int LINE_ENDS = 11;
hipError_t launch_line_ends(hipStream_t s, const LineArgs& a, double spacing, int supply_left, int supply_right, double* out);
// A cell supplied by a wall: cell `cell` takes `fraction` of Q of cell `donor` and sits at `x`; the line gets the slot
// (*key_at, *id_at) = (key, cell).
struct LineSplit {
    int cell, donor;
    double fraction, x;
    long long key;
    long long* key_at;
    int* id_at;
};
hipError_t launch_line_split(hipStream_t s, double* q, double* x, const LineSplit* splits, int count);
// The slots in (key, cell) order again: device radix sort of key * 2^32 + cell. temp: scratch of *temp_bytes (a null
// temp only sets *temp_bytes); packed: 2 n 64-bit words.
hipError_t launch_line_sort(hipStream_t s, long long* key, int* id, int n, unsigned long long* packed, void* temp, size_t* temp_bytes);
// Derivatives of the static step at fixed dt: the tangent writes the step into a.next and J dq into dnext; the
// adjoint writes J^T qbar (q: the state before the step) into out.
hipError_t launch_line_tangent(hipStream_t s, const LineArgs& a, double dt, const double* dq, double* dnext);
hipError_t launch_line_adjoint(hipStream_t s, const LineArgs& a, double dt, const double* qbar, double* out);

}  // namespace luna
