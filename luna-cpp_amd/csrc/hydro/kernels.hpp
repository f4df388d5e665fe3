// Device kernels of the hydro step on a Voronoi mesh (D = 2, 3) and on a sorted line (D = 1), and their argument records.
// Launched by api.hip on the handle's stream. FP64 throughout, compiled with -ffp-contract=off (physics.hpp).
#pragma once

#include "physics.hpp"

namespace luna {

constexpr unsigned long long INF_BITS = 0x7FF0000000000000ull;  // +inf as the bits a min-reduction starts from

// The mesh arrays as the kernels read them (device pointers), finite_volume.hpp:257-327 Geometry.
struct Geo {
    int N, F;
    const int* owner;
    const int* neighbour;
    const int* start;
    const int* list;
    const double* area;          // F x D
    const double* fc;            // F x D, relative to the owner's generator
    const double* magnitude;     // F: |A|, computed once per mesh
    double* magnitude_out;       // the same array, written by a moving step's second flux pass
    const double* centroid;      // N x D, relative to the generator
    const double* volume;        // N
    const double* radius;        // N: the radius of a sphere (circle) of the cell's volume, from the host's cbrt (sqrt)
    const double* positions;     // N x D
    double side[3];
};

// Per-cell rows of the step (device pointers).
struct Rows {
    double* q;             // N x V, extensive
    double* prims;         // N x (V + 2): W, sound speed, radius
    double* grad;          // N x (V D)
    double* w;             // N x D point velocities (moving)
    double* x;             // N x D solver positions (moving)
    double* old_centroid;  // N x D centroids of t^n (moving)
    int* steps;            // N x D whole grid steps over the step (moving)
    double* flux;          // F x V
};

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
template <int D>
hipError_t launch_magnitudes(hipStream_t s, const Geo& g, double* magnitude);
template <int D>
hipError_t launch_sweep(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double courant, bool moving, bool grad, bool limit,
                        int rule, const Counters& c);
// pass: -1 static, 0 a moving step's first pass (mesh of t^n), 1 its second (mesh of t^n+1).
template <int D>
hipError_t launch_flux(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, int pass, double dt, double unit,
                       unsigned long long* fallbacks);
template <int D>
hipError_t launch_update(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double dt, bool next_prims);
template <int D>
hipError_t launch_keep_centroids(hipStream_t s, const Geo& g, const Rows& r, const Grid& grid, double dt);
template <int D>
hipError_t launch_drift(hipStream_t s, int N, const Rows& r, const Grid& grid);

// D = 1: the static sorted line.
struct LineArgs {
    int n, periodic, wall_left, wall_right, order, limiter, riemann;
    Gas gas;
    const long long* key;
    long long period;
    double step;
    const double* q;
    double* next;
};
hipError_t launch_line_time_step(hipStream_t s, const LineArgs& a, double courant, int luna_rule, const Counters& c);
hipError_t launch_line_step(hipStream_t s, const LineArgs& a, double dt, double courant, bool fused, const Counters& c);

}  // namespace luna
