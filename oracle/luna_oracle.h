/* CPU oracle of luna-cpp's finite-volume hydro step (TEST INFRASTRUCTURE ONLY).
 *
 * A plain C99 + OpenMP restatement of the reference's FP64 arithmetic, in the reference's operation order, compiled with
 * -ffp-contract=off. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only as the checker
 * or as the timed CPU baseline; the product (libluna_hydro.so) never calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to the luna-cpp checkout). Parity pins:
 *   orc_hll / orc_hllc ............ LUNA golden fluxes (tests/golden/riemann.json), closed forms, symmetry, rotation
 *   orc_primitive / orc_conserved . round trips, floors (closed forms)
 *   orc_line_* .................... LUNA's N=16 final states (MC) and 5-step forward states (minmod, MC, van Leer),
 *                                   exact Riemann L1 bounds
 *   orc_time_step / orc_step_static (D = 2, 3): closed forms (linear field reproduced by the unlimited fit, uniform state
 *                                   stays uniform, conservation), brute-force lstsq of the fit, limiter range invariant,
 *                                   dt closed form on a lattice
 *   orc_half_step (2D, 3D) ........ the primitive-variable Euler Jacobians, the Euler equations (complex-step derivatives
 *                                   of the exact flux), rotation covariance
 *   orc_face_velocity ............. the normal speed of the moving bisector plane at the face centroid (central
 *                                   differences of the geometry), both sides of the 0.9 |d| limit
 *   orc_regularised_point_velocity  each branch of the face-angle ramp (closed forms), and in the sweep on a mesh where
 *                                   all three fire
 *   the face-centroid limiter ..... maximality: limited = min(1, min over faces of the needed factor) x unlimited
 *   orc_face_sides ................ static W - dt/2 A g + g r, moving pass 1 W + g (x_f^n - c^n), pass 2
 *                                   W - dt A g + g (x_f^n+1 - c^n), against the Jacobians and the meshes' lab positions
 *   orc_moving_begin / orc_moving_finish: uniform bulk flow stays uniform, conservation, Galilean invariance of a blast
 *                                   (hydro_test.cpp:474-510, the reference's bounds); the whole moving step against the
 *                                   reference's own build is "parity unpinned" (voronoi-cpp and Kokkos are absent).
 */
#ifndef LUNA_ORACLE_H
#define LUNA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* euler.hpp:82-101 */
typedef struct {
    double gamma, smallr, smallc, inverse_gm1;
} orc_gas;

void orc_gas_init(orc_gas* gas, double gamma);
void orc_set_threads(int n);
int orc_max_threads(void);
void orc_primitive(int D, const orc_gas* gas, const double* U, double* W);
void orc_conserved(int D, const orc_gas* gas, const double* W, double* U);
double orc_sound_speed(int D, const orc_gas* gas, const double* W);
void orc_half_step(int D, const orc_gas* gas, const double* W, const double* g, double dt, double* out);
double orc_cbrt_rn(double v); /* the cube root rounded to nearest (the cell radius) */
void orc_hll(int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w, double* F);
void orc_hllc(int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w, double* F);

/* The mesh arrays the step reads (the voronoi-cpp output contract, finite_volume.hpp:257-327).
 * list entries: f for a face the cell owns, ~f (negative) for one it does not. */
typedef struct {
    int D, N, F;
    const int* owner;
    const int* neighbour;
    const int* start;
    const int* list;
    const double* area;          /* F x D, |A| n, owner to neighbour */
    const double* face_centroid; /* F x D, relative to the owner's generator */
    const double* centroid;      /* N x D, relative to the cell's generator */
    const double* volume;        /* N */
    const double* positions;     /* N x D, in [0, side) */
    double side[3];
    double step;                 /* snapping grid step */
    long long period[3];         /* grid steps per side */
} orc_geometry;

#define ORC_FEATURES 16 /* features per cell and variable of the learned limiter */
unsigned short orc_bf16_bits(double x);

/* Solver state (arrays owned by the caller). limiter: 0 face centroid, 1 none, 2 face centroid recording the learned
 * limiter's features and troubled flags (the gradients are those of 0). riemann: 0 HLLC, 1 HLL.
 * step_rule: 0 AREPO, 1 LUNA (2D only). */
typedef struct {
    int D, N, order, limiter, riemann, moving, step_rule;
    orc_gas gas;
    double* q;            /* N x V, extensive */
    double* prims;        /* N x (V + 2): W, sound speed, radius */
    double* grad;         /* N x (V D) */
    double* w;            /* N x D point velocities (moving) */
    double* x;            /* N x D positions (moving) */
    double* old_centroid; /* N x D (moving) */
    int* steps;           /* N x D whole grid steps (moving) */
    double* flux;         /* F x V scratch (at least the largest mesh's F) */
    long long fallbacks[2];
    int prims_ready;
    unsigned short* features; /* limiter 2 (learned): N V x 16 BF16 bit patterns, written by each sweep */
    unsigned char* troubled;  /* limiter 2: N */
    int hold_steps;           /* moving: orc_moving_begin keeps the given steps (the adjoint's central differences) */
    int precision;            /* the gradient policy (precision.hpp:26-54): 0 double, 1 grad32, 2 grad16; grad keeps the
                                 stored rows decoded to double (floats, exactly) */
    double hll_compression;   /* > 0: the HLL fallback at extreme compression (no counterpart in the reference): the sweep
                                 sets hll[i] when r_i (div v)_i / c_s,i < -hll_compression (div v of the limited gradient);
                                 HLLC takes HLL at faces with a flagged side */
    unsigned char* hll;       /* N, written by the sweep when hll_compression > 0 */
    int sweep_lanes;          /* 0 or 1: the sweep (finite_volume.hpp:463-569); 2, 4, 8: LaneSweep (:880-1037), order 1,
                                 double, face-centroid limiter or none */
} orc_fv;

void orc_primitives(orc_fv* fv, const orc_geometry* geo);
double orc_time_step(orc_fv* fv, const orc_geometry* geo, double courant);
void orc_step_static(orc_fv* fv, const orc_geometry* geo, double dt);
void orc_moving_begin(orc_fv* fv, const orc_geometry* geo, double dt);
void orc_moving_finish(orc_fv* fv, const orc_geometry* geo, double dt);
/* Pieces of the step the tests pin on their own: the face velocity (finite_volume.hpp:1380-1397; d owner -> neighbour,
 * xf relative to the owner's generator), the regularised point velocity (:526-537) and a face's two states as the flux
 * solves them (pass -1 static, 0 / 1 a moving step's first / second pass, 2 order 0 moving; returns |A|). */
void orc_face_velocity(int D, const double* d, const double* xf, const double* wL, const double* wR, double* wf);
void orc_regularised_point_velocity(int D, const double* Wi, const double* centroid, double max_angle, double* w);
double orc_face_sides(const orc_fv* fv, const orc_geometry* geo, int f, double dt, int pass, double unit, double* L, double* R, double* n,
                      double* wf, int* fallbacks);
/* Derivatives of the static order-1 step (HLLC or HLL) at fixed dt (finite_volume.hpp:1670-2231): the tangent advances q and dq (J dq)
 * in place; the adjoint maps qbar after the step to J^T qbar before it (q, the state before the step, stays). Return -1
 * for a moving mesh, order 0 or D = 1 (see orc_line_*). */
int orc_step_tangent(orc_fv* fv, const orc_geometry* geo, double dt, double* dq);
int orc_step_adjoint(orc_fv* fv, const orc_geometry* geo, double dt, double* qbar);
/* The moving step's reverse at order 1 and fixed topology (finite_volume.hpp:2984-3155 step_moving_adjoint_1): fv holds the
 * state at t^n (q, x on the grid; its prims and sweep are recomputed on g0), g1 is the mesh of the positions the forward
 * step reached. qbar, xbar (N x V, N x D: the bars of Q, x after the step) become those before it. The geometries'
 * adjoints are the caller's (the mesh library's derivative): geometry_adjoint(ctx, Vbar, cbar, Abar, xfbar, xbar) adds to
 * xbar, called for g1 (ctx1), then g0 (ctx0). steps_fixed: the whole grid steps the forward took are held (derivative 0);
 * otherwise the quantisation is the identity (x^n+1 = x^n + dt w). The forward with held steps: orc_moving_begin with
 * fv->hold_steps set and fv->steps given. */
typedef void (*orc_geometry_adjoint)(void* ctx, const double* Vbar, const double* cbar, const double* Abar, const double* xfbar, double* xbar);
int orc_moving_adjoint(orc_fv* fv, const orc_geometry* g0, const orc_geometry* g1, double dt, int steps_fixed, double* qbar, double* xbar,
                       orc_geometry_adjoint adjoint0, void* ctx0, orc_geometry_adjoint adjoint1, void* ctx1);
/* The moving step's tangent (finite_volume.hpp:2428-2486 order 0, :2764-2870 order 1), fixed topology: from the state held
 * (Q, x at t^n on g0; g1 the mesh the forward step reaches), dq (N x V) and dx (N x D) at t^n become their tangents at
 * t^n+1, in place; Q and x stay. tangent0 / tangent1: the two meshes' geometry tangents (dx -> dV, dc, dA, dx_f, the
 * mesh library's GeometryDerivative::tangent). steps_fixed as orc_moving_adjoint. -1 where the adjoint refuses too. */
typedef void (*orc_geometry_tangent)(void* ctx, const double* dx, double* dV, double* dc, double* dA, double* dxf);
int orc_moving_tangent(orc_fv* fv, const orc_geometry* g0, const orc_geometry* g1, double dt, int steps_fixed, double* dq, double* dx,
                       orc_geometry_tangent tangent0, void* ctx0, orc_geometry_tangent tangent1, void* ctx1);
/* Its pieces, pinned on their own: the face velocity's tangent (:2322-2330) and the moving face's (euler.hpp:861-891). */
void orc_face_velocity_tangent(int D, const double* d, const double* xf, const double* wL, const double* wR, const double* dd, const double* dxf,
                               const double* dwL, const double* dwR, double* dwf);
void orc_moving_face_tangent(int kind, int D, const orc_gas* gas, const double* L, const double* R, const double* A, const double* w,
                             const double* dL, const double* dR, const double* dA, const double* dw, double* dG);

long long orc_snap(double x, double step, long long period);
/* precision.hpp:55-130, half::from_float / to_float (IEEE binary16, round to nearest even), group_exponent,
 * power_of_two. */
unsigned short orc_half_from_float(float f);
float orc_half_to_float(unsigned short h);
int orc_group_exponent(float m);
float orc_power_of_two(int k);
int orc_delta(long long k0, long long k1, long long period);

/* 1D sorted line (finite_volume_1d.hpp). Static: slots are cells (id NULL). Moving: see orc_mline_*, which owns one.
 * wall kinds: 0 reflective, 1 outflow, 2 inflow (moving only). limiter: 0 face centroid (MC on a uniform line), 1 none,
 * 2 minmod, 3 van Leer. */
typedef struct {
    int n, periodic, wall_left, wall_right, order, limiter, riemann;
    orc_gas gas;
    const long long* key; /* ascending grid keys, by slot */
    long long period;
    double step;
    double* q;    /* n x 3, by cell */
    double* next; /* n x 3, by cell */
    /* moving line (zero / NULL on a static one) */
    int moving, regularise;
    const int* id;       /* slot -> cell; NULL: slot s holds cell s */
    double side;         /* period * step */
    double inflow[2][3]; /* primitives of an inflow wall, left and right */
} orc_line;

double orc_line_time_step(orc_line* line, double courant, int luna_rule);
void orc_line_step(orc_line* line, double dt);
double orc_line_step_fused(orc_line* line, double dt, double courant);
double orc_line_volume(const orc_line* line, int s);
/* Derivatives of a static step, dt fixed (finite_volume_1d.hpp:710-935, euler.hpp:115-832): the tangent advances q and dq
 * (J dq) in place; the adjoint maps qbar after the step to J^T qbar before it, with q the state before the step. */
void orc_line_step_tangent(orc_line* line, double dt, double* dq);
void orc_line_step_adjoint(orc_line* line, double dt, double* qbar);
void orc_primitive_tangent(int D, const orc_gas* gas, const double* U, const double* dU, double* dW);
void orc_primitive_adjoint(int D, const orc_gas* gas, const double* U, const double* Wbar, double* Ubar);
void orc_half_step_tangent(int D, const orc_gas* gas, const double* W, const double* dW, const double* g, const double* dg, double dt,
                           double* out);
void orc_half_step_adjoint(int D, const orc_gas* gas, const double* W, const double* g, double dt, const double* hbar, double* Wbar,
                           double* gbar);
void orc_hllc_tangent(int D, const orc_gas* gas, const double* L, const double* R, const double* dL, const double* dR, const double* n,
                      const double* w, double* F, double* dF);
void orc_hllc_adjoint(int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w, const double* Fbar,
                      double* Lbar, double* Rbar);
void orc_hll_tangent(int D, const orc_gas* gas, const double* L, const double* R, const double* dL, const double* dR, const double* n,
                     const double* w, double* F, double* dF);
void orc_hll_adjoint(int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w, const double* Fbar,
                     double* Lbar, double* Rbar);
double orc_line_centroid(const orc_line* line, int s);

/* 1D moving line (finite_volume_1d.hpp with Motion::moving): the generators drift with the gas plus the regularisation,
 * the line keeps its slots in key order, and an open wall (outflow or inflow) supplies cells split from the end cell.
 * Positions x in [0, side] (walls) or [0, side) (periodic); the grid has `period` keys per side. Cells are numbered in
 * the order they were made; q and x are by cell, the line by slot. */
typedef struct orc_mline orc_mline;
orc_mline* orc_mline_create(int n, const double* x, double side, long long period, int periodic, int wall_left, int wall_right,
                            const double* inflow_left, const double* inflow_right, int order, int limiter, int riemann, double gamma,
                            double spacing, int regularise);
void orc_mline_destroy(orc_mline* m);
int orc_mline_cells(const orc_mline* m);
void orc_mline_set_state(orc_mline* m, const double* q);
void orc_mline_get_state(const orc_mline* m, double* q);
void orc_mline_get_positions(const orc_mline* m, double* x);
void orc_mline_get_line(const orc_mline* m, long long* keys, int* ids);
/* what: 0 cells added, 1 order updates, 2 sorts */
long long orc_mline_stat(const orc_mline* m, int what);
double orc_mline_time_step(orc_mline* m, double courant, int luna_rule);
void orc_mline_step(orc_mline* m, double dt);
double orc_mline_step_fused(orc_mline* m, double dt, double courant);

#ifdef __cplusplus
}
#endif
#endif
