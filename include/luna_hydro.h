/* luna_hydro: the finite-volume hydro step of luna-cpp on MI355X (gfx950), as a C-ABI shared library (libluna_hydro.so).
 *
 * This is the drop-in boundary of luna-cpp's luna_hydro_library for the hydro step: each entry point replaces a member of
 * hydro::FiniteVolume<D> (src/hydro/finite_volume.hpp) or FiniteVolume<1> (src/hydro/finite_volume_1d.hpp), cited
 * beside it. The reference is a C++ template class on Kokkos; its C++ facade over this C-ABI is include/luna/finite_volume.hpp.
 * Every kernel is hand-written HIP for CDNA4, FP64 throughout (the reference's Precision<> policy). One HIP stream per
 * handle; a handle is not thread-safe (as the reference). Errors: every entry point returns LUNA_OK (0) or an error code,
 * with the message in luna_last_error() (the reference throws std::invalid_argument / std::logic_error instead).
 */
#ifndef LUNA_HYDRO_H
#define LUNA_HYDRO_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LUNA_OK = 0,
    LUNA_ERR_ARGUMENT = 1, /* std::invalid_argument in the reference (order not 0/1, bad sizes) */
    LUNA_ERR_LOGIC = 2,    /* std::logic_error in the reference (call not valid for this motion / state) */
    LUNA_ERR_DEVICE = 3,   /* a HIP runtime error */
    LUNA_ERR_MEMORY = 4    /* device allocation failed */
};

/* finite_volume.hpp:25-34; LUNA_LIMITER_LEARNED (2D/3D, order 1) has no counterpart in the reference: see below. */
enum { LUNA_LIMITER_FACE_CENTROID = 0, LUNA_LIMITER_NONE = 1, LUNA_LIMITER_MINMOD = 2, LUNA_LIMITER_VAN_LEER = 3, LUNA_LIMITER_LEARNED = 4 };
enum { LUNA_RIEMANN_HLLC = 0, LUNA_RIEMANN_HLL = 1 };
enum { LUNA_MOTION_STATIC = 0, LUNA_MOTION_MOVING = 1 };
enum { LUNA_STEP_AREPO = 0, LUNA_STEP_LUNA = 1 };
enum { LUNA_WALL_REFLECTIVE = 0, LUNA_WALL_OUTFLOW = 1, LUNA_WALL_INFLOW = 2 }; /* hydro::Wall, finite_volume_1d.hpp:9 */
/* The gradient slot of hydro::Precision (precision.hpp:26-54): double (Precision<>), grad32 (Grad32: the fit, the limiter
 * and the face-state change in float, float rows), grad16 (Grad16: the same arithmetic, rows of FP16 with one
 * power-of-two exponent per group rho, v, P). Q, W, volumes, centroids, the Riemann solve and the flux sums stay double
 * in every policy. */
enum { LUNA_PRECISION_DOUBLE = 0, LUNA_PRECISION_GRAD32 = 1, LUNA_PRECISION_GRAD16 = 2 };
/* hydro::Update (finite_volume.hpp:25, apply() :1539-1631): the per-face loop (flux rows, then a per-cell gather in list
 * order) or the per-cell one (fluxes_and_update, :1544-1566: each cell computes its faces' fluxes itself, each face twice;
 * the same bits, no flux rows). */
enum { LUNA_UPDATE_PER_FACE = 0, LUNA_UPDATE_PER_CELL = 1 };

/* hydro::Gas (euler.hpp:82-101) and the constructor's options (finite_volume.hpp:79-80). The update loop is the
 * reference's per-face one (Update::per_face: flux rows, then a deterministic per-cell gather in list order) unless
 * `update` asks for the per-cell one. */
typedef struct {
    int dim;       /* 2 or 3 */
    int order;     /* 0 or 1 */
    int limiter;   /* LUNA_LIMITER_FACE_CENTROID, LUNA_LIMITER_NONE or LUNA_LIMITER_LEARNED */
    int riemann;   /* LUNA_RIEMANN_HLLC or LUNA_RIEMANN_HLL */
    int motion;    /* LUNA_MOTION_STATIC or LUNA_MOTION_MOVING */
    int step_rule; /* LUNA_STEP_AREPO, or LUNA_STEP_LUNA in 2D (set_step_rule, finite_volume.hpp:181) */
    double gamma, smallr, smallc;
    int precision; /* LUNA_PRECISION_*; a zero-initialised struct asks for double. The reduced policies step forward
                      (static and moving, order 1) with the face-centroid limiter or none; the learned limiter and the
                      derivatives need double. The reference's default (DefaultPrecision<D>, precision.hpp:49-54) is
                      grad16 in 3D and grad32 in 2D; here the default is double. */
    double hll_compression; /* > 0: the HLL fallback at extreme compression, which has no counterpart in the reference.
                               Each time_step's sweep flags cell i when r_i (div v)_i / c_s,i < -hll_compression,
                               div v = sum_a dv_a/dx_a of the limited gradient and r the cell radius. The HLLC solver then
                               takes HLL at every face with a flagged side. 0 (a zero-initialised struct): off. Needs
                               order 1, the double policy and the face-centroid limiter or none. The derivatives refuse it. */
    int update;             /* LUNA_UPDATE_PER_FACE (a zero-initialised struct) or LUNA_UPDATE_PER_CELL (order 1; the
                               derivatives and the adjoint's checkpoints use the per-face rows either way) */
    int sweep_lanes;        /* the reference's LUNA_SWEEP_LANES (finite_volume.hpp:880-1037 LaneSweep): 0 or 1 (a
                               zero-initialised struct) the one-lane sweep; 2, 4 or 8 lanes per cell, each taking every
                               L-th list entry, their sums combined by a fixed tree and the limiter by the smallest factor
                               over the faces: the same bits run to run, other bits than the one-lane sweep's. Needs order
                               1, the double policy and the face-centroid limiter or none; the derivatives refuse it. */
} luna_fv_options;

/* The mesh arrays the step reads, host pointers: voronoi-cpp's output contract as finite_volume.hpp:257-327 and :353-364
 * read it (luna_mesh_record of include/luna_mesh.h starts with the same fields). Row-major [rows][dim] FP64, int32
 * indices; list entries f (the cell owns face f) or ~f (it does not) — or plain face ids: the owner bit is set from
 * owner[] as the reference's owner_layout does (finite_volume.hpp:78-87), so a list with or without it steps the same;
 * area A = |A| n points owner -> neighbour; face centroids relative to the owner's generator, cell centroids relative to
 * their own. */
typedef struct {
    int dim, cells, faces;
    const int* owner;
    const int* neighbour;
    const int* start;
    const int* list;
    const double* area;
    const double* face_centroid;
    const double* centroid;
    const double* volume;
    const double* positions;
    double side[3];
    double step;         /* the mesh's snapping grid step (exact_predicates::BoxN::step) */
    long long period[3]; /* grid steps per box side (BoxN::period) */
} luna_geometry;

typedef struct luna_fv luna_fv;

/* FiniteVolume(exec, mesh, gas, order, Update::per_face, limiter, motion) (finite_volume.hpp:79-107): allocates the
 * solver on the current HIP device, uploads the mesh; Q starts at zero. */
int luna_fv_create(const luna_fv_options* options, const luna_geometry* mesh, luna_fv** out);
int luna_fv_destroy(luna_fv* fv);

/* state() (finite_volume.hpp:145-148): Q, N x (dim + 2) extensive (mass, momentum, energy), row-major FP64.
 * set/get copy through the host; state_device returns the device array (and, as state(), drops the cached primitives). */
int luna_fv_set_state(luna_fv* fv, const double* q);
int luna_fv_get_state(luna_fv* fv, double* q);
int luna_fv_state_device(luna_fv* fv, double** q);

/* time_step(courant) (finite_volume.hpp:190-201): primitives, the sweep (limited gradients; moving: point velocities),
 * and the global dt, read back to the host (blocking). */
int luna_fv_time_step(luna_fv* fv, double courant, double* dt);

/* step(dt) on a static mesh (finite_volume.hpp:205-226, per-face loop). */
int luna_fv_step(luna_fv* fv, double dt);

/* step(dt) on a moving mesh at order 1 (step_moving, finite_volume.hpp:1497-1519), split where the reference calls the
 * mesh library (move(), :1212-1250). begin keeps the centroids of t^n, puts the points on the grid, applies the fluxes of
 * the mesh of t^n over dt/2 and drifts the points by their whole grid steps; positions (N x dim, host) receives the new
 * positions unless NULL (then luna_fv_positions reads them). The caller builds the mesh of those positions and passes it
 * to finish, or to luna_fv_set_geometry first and NULL to finish; finish applies the fluxes of the new mesh over dt/2 and
 * writes the next primitives. */
int luna_fv_moving_begin(luna_fv* fv, double dt, double* positions);
int luna_fv_moving_finish(luna_fv* fv, const luna_geometry* mesh, double dt);
/* At order 0 (advance<0, true>() + move(), finite_volume.hpp:216-225, 1211-1250): begin applies one flux pass over dt on
 * the mesh of t^n, with face velocities from the sweep's point velocities, and moves the points off the grid
 * (x + dt w wrapped into the box); finish takes the new mesh and writes its primitives (Q is unchanged). */

/* Takes a mesh. Between moving_begin and moving_finish: the mesh of the new positions. Otherwise: a caller moved the
 * mesh itself (mesh_moved(), finite_volume.hpp:152-156), and the cached rows are dropped; a moving solver's positions
 * become the mesh's (a restart from a checkpoint of Q and x). The handle derives what its kernels read (the cell rows,
 * lists in 64-cell slices with |A| and the face offsets x_f - c_i, the cell radii) in one pass on the GPU; a host mesh's
 * arrays are copied into the handle first, so the caller's may change once the call returns. A malformed mesh (CSR offsets not
 * running from 0 to 2 faces, a cell or face index out of range) gives LUNA_ERR_ARGUMENT, and the handle then refuses to
 * step (LUNA_ERR_LOGIC) until a valid mesh is set. */
int luna_fv_set_geometry(luna_fv* fv, const luna_geometry* mesh);

/* set_geometry and the constructor with the mesh arrays in device memory (pointers from hipMalloc on the handle's
 * device, or torch tensors' data): voronoi-cpp's device views, which the reference reads in place
 * (finite_volume.hpp:257-327, 353-364). The side, step and period fields are values as for a host mesh. The arrays must
 * be ready when the call is made (the handle's stream does not wait on another stream). The cell arrays and the lists
 * are read only during the call; the face arrays (owner, neighbour, area, face_centroid) are read in place by the steps
 * that follow, as the reference reads the mesh it holds, so they must stay valid and unchanged until the next mesh is
 * set, luna_fv_release_mesh (static) copies them, or the handle is destroyed. */
int luna_fv_set_geometry_device(luna_fv* fv, const luna_geometry* mesh);
int luna_fv_create_device(const luna_fv_options* options, const luna_geometry* mesh, luna_fv** out);

/* fallbacks() (finite_volume.hpp:230-238): face sides of the last step that took the cell's state. Fences. */
int luna_fv_fallbacks(luna_fv* fv, long long* count);

/* positions(), point_velocities() (finite_volume.hpp:150, 157), host copies N x dim. */
int luna_fv_positions(luna_fv* fv, double* x);
int luna_fv_point_velocities(luna_fv* fv, double* w);

/* Host copy of the limited gradients of the last time_step(), N x (dim + 2) x dim (density_gradients(),
 * finite_volume.hpp:165-180, gives the first dim of each row); under grad32/grad16 the stored rows, decoded (FP16 times
 * 2^-k per group, FaceFlux::gradient(), finite_volume.hpp:1288-1300) and widened to double. */
int luna_fv_gradients(luna_fv* fv, double* g);

/* Host copy of the primitive rows the last sweep read (or the next one will), N x (dim + 4): W, sound speed, radius
 * (finite_volume.hpp:67-68 Prims, written by primitives() :403-415). */
int luna_fv_primitives(luna_fv* fv, double* rows);

/* faces() (finite_volume.hpp:141). */
int luna_fv_faces(luna_fv* fv, int* faces);

/* set_step_rule(StepRule) (finite_volume.hpp:181): the rule of the next time steps; LUNA_STEP_LUNA in 2D only
 * (LUNA_ERR_ARGUMENT otherwise). */
int luna_fv_set_step_rule(luna_fv* fv, int rule);
/* release_mesh() (finite_volume.hpp:134-139): the reference stops reading the caller's mesh on a static mesh and refuses on a
 * moving one (LUNA_ERR_LOGIC). On a static mesh the handle copies the face arrays of a device mesh, which it reads in
 * place, into its own memory (a host mesh's are its copies already); the caller may then free the mesh. */
int luna_fv_release_mesh(luna_fv* fv);
/* mesh_stats() (finite_volume.hpp:160-163): of the reference's counters, `updates`, the mesh updates of moving steps
 * (:1237; here the meshes a moving step took between begin and finish); the rest are the mesh library's own
 * (voronoi-cpp's update report). A mesh set outside a step (mesh_moved()) is not an update. */
int luna_fv_mesh_stats(luna_fv* fv, long long* updates);

/* Kernel timing with HIP events on the handle's stream (on: 1, off: 0; turning it on clears the sums). read: up to max
 * entries of (name, launches, total milliseconds), for the launches since it was turned on. */
typedef struct {
    char name[32];
    long long launches;
    double total_ms;
} luna_kernel_time;
int luna_fv_profile(luna_fv* fv, int on);
int luna_fv_profile_read(luna_fv* fv, luna_kernel_time* out, int max, int* count);

/* ---------------------------------------------------------------- D = 1: FiniteVolume<1> on a sorted line */

/* The line as finite_volume_1d.hpp reads voronoi-cpp's sorted_mesh::Line: cells in ascending order of their integer grid
 * keys, grid step `step`, `period` keys per side; walls at 0 and period * step, or periodic. */
typedef struct {
    int cells;
    const long long* keys;
    long long period;
    double step;
    int periodic;
    int wall_left, wall_right; /* LUNA_WALL_* (finite_volume_1d.hpp:9-14) */
} luna_line;

typedef struct {
    int order;   /* 0 or 1 */
    int limiter; /* any LUNA_LIMITER_* */
    int riemann;
    double gamma, smallr, smallc;
} luna_fv1_options;

typedef struct luna_fv1 luna_fv1;

/* FiniteVolume<1>(exec, mesh, gas, order, limiter, static_mesh, walls) (finite_volume_1d.hpp:41-53). */
int luna_fv1_create(const luna_fv1_options* options, const luna_line* line, luna_fv1** out);
int luna_fv1_destroy(luna_fv1* fv);
int luna_fv1_set_state(luna_fv1* fv, const double* q);
int luna_fv1_get_state(luna_fv1* fv, double* q);
/* time_step(courant, rule) (finite_volume_1d.hpp:332-386); rule LUNA_STEP_AREPO or LUNA_STEP_LUNA. */
int luna_fv1_time_step(luna_fv1* fv, double courant, int rule, double* dt);
/* step(dt) (finite_volume_1d.hpp:389-407). */
int luna_fv1_step(luna_fv1* fv, double dt);
/* step_and_time_step(dt, courant) (finite_volume_1d.hpp:412-417): the step and the next dt by LUNA's rule, one launch. */
int luna_fv1_step_and_time_step(luna_fv1* fv, double dt, double courant, double* next);

/* The moving line: the generators as voronoi-cpp's Mesh<1> holds them, with the walls and the cell supply of
 * FiniteVolume<1>(exec, mesh, gas, order, limiter, Motion::moving, walls, regularise) (finite_volume_1d.hpp:41-53) and
 * set_inflow_spacing() (:62-67). The line's keys are the positions snapped to the grid (step = side / period). */
typedef struct {
    int cells;
    const double* positions; /* cells generators in [0, side] (walls) or [0, side) (periodic) */
    double side;
    long long period; /* grid keys per side */
    int periodic;
    int wall_left, wall_right;              /* LUNA_WALL_*; an inflow wall holds inflow_left / inflow_right (Walls, :11-14) */
    double inflow_left[3], inflow_right[3]; /* primitives rho, v, P */
    double spacing;                         /* inflow_spacing(): cells from open walls this far apart; 0: none */
    int regularise;                         /* the point velocity's regularisation (drift(), :301-314) */
} luna_moving_line;

/* FiniteVolume<1> with Motion::moving. Then time_step (AREPO's or LUNA's rule on |w - v|), step (the points drift, the
 * order is updated, open walls supply cells when due, :389-407, :419-465) and step_and_time_step (:679-708) as on a static
 * line; set_state / get_state take cells x 3 rows by cell, for the current number of cells (luna_fv1_cells), which grows
 * as walls supply cells (new cells numbered after the last). */
int luna_fv1_create_moving(const luna_fv1_options* options, const luna_moving_line* line, luna_fv1** out);
/* cells() (:58) and cells_added() (:59). */
int luna_fv1_cells(luna_fv1* fv, int* cells, long long* added);
/* positions() (:57): cells positions, by cell. */
int luna_fv1_positions(luna_fv1* fv, double* x);
/* The line (Mesh<1>::line()): per slot in key order, its key and its cell; cells entries each. */
int luna_fv1_line(luna_fv1* fv, long long* keys, int* cells);
/* set_inflow_spacing(spacing) (:62-67). */
int luna_fv1_set_inflow_spacing(luna_fv1* fv, double spacing);
/* mesh_stats() (:71): order updates and the sorts among them. */
int luna_fv1_mesh_stats(luna_fv1* fv, long long* updates, long long* sorts);

/* FiniteVolume<1>::step_tangent(dt, dq) (finite_volume_1d.hpp:850-880): one step of Q and of its tangent; dq (host, cells x 3,
 * a tangent of the extensive Q as state()) becomes J dq in place. Static line, dt fixed. */
int luna_fv1_step_tangent(luna_fv1* fv, double dt, double* dq);

/* FiniteVolume<1>::step_adjoint(dt, qbar) (finite_volume_1d.hpp:882-933): qbar (host, cells x 3) after the step becomes J^T qbar
 * before it, in place; Q is the state at the start of the step and stays. Static line, dt fixed. */
int luna_fv1_step_adjoint(luna_fv1* fv, double dt, double* qbar);

/* Derivatives of one static step at fixed dt (finite_volume.hpp:1670-2231; static mesh, order 1, HLLC or HLL, either limiter), on
 * N x (dim + 2) rows of the extensive state; device = 1: dq / qbar are device pointers (as luna_fv_state_device), else host.
 * FiniteVolume::step_tangent(dt, dq) (:2039-2074): Q and dq advance by one step, dq becoming J dq.
 * FiniteVolume::step_adjoint(dt, qbar) (:2201-2231): qbar after the step becomes J^T qbar before it; Q (the state at the start
 * of the step) stays. Both block until done. LUNA_ERR_LOGIC on a moving mesh or order 0. */
int luna_fv_step_tangent(luna_fv* fv, double dt, double* dq, int device);
int luna_fv_step_adjoint(luna_fv* fv, double dt, double* qbar, int device);
/* No counterpart in the reference: luna_fv_step_adjoint (device = 1) for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It allocates nothing: the adjoint's work
 * rows (the faces' side bars, the cells' W bars, fit bars, bound bars and extremes, the output) are carved from the
 * library's own scratch, a file-scope device array of 64 MiB per GPU (held from the library's load until the process
 * ends), which bounds the mesh: about 70,000 cells in 3D and 160,000 in 2D; a larger mesh gives LUNA_ERR_ARGUMENT
 * (use luna_fv_step_adjoint). qbar: a device array (cells x (dim + 2)), read and then overwritten with J^T qbar. Every
 * call on a GPU uses that GPU's one scratch array, so calls that overlap on one GPU (two threads, or two streams in
 * flight) race; each call writes every row it reads before reading it. Blocks, as luna_fv_step_adjoint. */
int luna_fv_step_adjoint_noalloc(luna_fv* fv, double dt, double* qbar);
/* FiniteVolume::time_step_tangent(courant, dq) (:1941-1964): dt, as luna_fv_time_step (the same terms, the same bits),
 * with the primitives' and the limited fit's tangents for dq (N x V, host or device as `device` says) kept for the next
 * luna_fv_step_tangent, which must be called with the same dq (the reference's contract); any other call in between
 * (set_state, time_step, step, ...) drops them and step_tangent takes them itself. The derivatives' conditions apply
 * (LUNA_ERR_LOGIC otherwise). */
int luna_fv_time_step_tangent(luna_fv* fv, double courant, const double* dq, int device, double* dt);

/* The geometry's derivative the moving step's reverse reads (hydro::GeometryDerivative::adjoint, mesh_derivative.hpp:28-29): from
 * the bars of V, c (cells) and A, x_f (faces), add to xbar (cells x dim); host arrays; 0 on success. The mesh library's
 * luna_mesh_derivative_adjoint (include/luna_mesh.h) has this shape, with its handle as ctx. */
typedef int (*luna_geometry_adjoint_fn)(const void* ctx, const double* Vbar, const double* cbar, const double* Abar, const double* xfbar,
                                         double* xbar);

/* FiniteVolume::step_moving_adjoint(dt, qbar, xbar, geometry, steps_fixed) at order 1 (finite_volume.hpp:2488-2495, 2984-3155), at
 * fixed topology: the handle holds Q at t^n (as set); mesh0 is the mesh of t^n (its positions are x^n), mesh1 the mesh the forward
 * step reached from it. qbar (cells x (dim + 2)) and xbar (cells x dim), host, the bars of Q and x after the step, become those
 * before it. adjoint0 / adjoint1: the two meshes' geometry adjoints (called on the host, mesh1's first). steps_fixed: the whole
 * grid steps the forward took are held (derivative 0, as central differences small enough to change no step see it);
 * otherwise the quantisation is the identity (x^n+1 = x^n + dt w). Leaves the handle on mesh0 with Q unchanged; blocks.
 * At order 0 (step_moving_adjoint, :2490-2598: one pass on mesh0 with the point velocities, x^n+1 = x^n + dt w) mesh1,
 * adjoint1 and steps_fixed are not used (pass mesh0 and adjoint0 for them). LUNA_ERR_LOGIC on a static mesh, the learned
 * limiter, a reduced precision policy or the HLL fallback. */
int luna_fv_moving_adjoint(luna_fv* fv, double dt, const luna_geometry* mesh0, const luna_geometry* mesh1, luna_geometry_adjoint_fn adjoint0,
                           const void* ctx0, luna_geometry_adjoint_fn adjoint1, const void* ctx1, int steps_fixed, double* qbar, double* xbar);

/* The geometry's tangent the moving step's tangent reads (hydro::GeometryDerivative::tangent, mesh_derivative.hpp:26-27): from dx
 * (cells x dim), write dV (cells), dc (cells x dim), dA and dxf (faces x dim); host arrays; 0 on success. The mesh library's
 * luna_mesh_derivative_tangent (include/luna_mesh.h) has this shape, with its handle as ctx. */
typedef int (*luna_geometry_tangent_fn)(const void* ctx, const double* dx, double* dV, double* dc, double* dA, double* dxf);

/* FiniteVolume::step_moving_tangent(dt, dq, dx, geometry, steps_fixed) (finite_volume.hpp:2428-2486 order 0, 2764-2870 order 1),
 * at fixed topology: the forward mode of the moving step at the state the handle holds (Q at t^n, as set; mesh0 the mesh of t^n,
 * its positions x^n; mesh1 the mesh the forward step reached from it, not read at order 0). dq (cells x (dim + 2)) and dx
 * (cells x dim), host, the tangents of Q and x at t^n, become those at t^n+1, in place. tangent0 / tangent1: the two meshes'
 * geometry tangents (called on the host). steps_fixed as luna_fv_moving_adjoint. The reference's member also takes the
 * forward step; here the caller's forward (moving_begin / finish) takes it, and the handle is left on mesh0 with Q unchanged,
 * as luna_fv_moving_adjoint leaves it. Blocks. LUNA_ERR_LOGIC where luna_fv_moving_adjoint refuses. */
int luna_fv_moving_tangent(luna_fv* fv, double dt, const luna_geometry* mesh0, const luna_geometry* mesh1, luna_geometry_tangent_fn tangent0,
                           const void* ctx0, luna_geometry_tangent_fn tangent1, const void* ctx1, int steps_fixed, double* dq, double* dx);

/* ------------------------------------------------------------ the learned slope limiter (no counterpart in luna-cpp)
 *
 * With LUNA_LIMITER_LEARNED each sweep runs the face-centroid limiter (finite_volume.hpp:538-565) and records, per cell and
 * variable, 16 features (BF16) and a troubled-cell flag; with a network set, an MLP 16 -> 64 -> 64 -> 1 (ReLU; BF16
 * operands, FP32 accumulation on the matrix cores) gives each (cell, variable) a fraction sigmoid(z) in [0, 1] that
 * scales its limited gradient, except in troubled cells (a compressive pressure jump, or a neighbour at under a tenth of
 * the cell's density or pressure), which keep the face-centroid gradient. The scaled gradient is never steeper than the
 * face-centroid one, so every face-centroid value stays within the neighbours' range. Without a network the learned mode
 * gives the face-centroid limiter's results bit for bit. The parameters are FP32 in the order W1 [64][16], b1 [64],
 * W2 [64][64], b2 [64], w3 [64], b3: h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), z = w3 . h2 + b3. */
#define LUNA_LIMITER_FEATURES 16
#define LUNA_LIMITER_PARAMS 5313

/* Sets (params host, count LUNA_LIMITER_PARAMS) or, with params NULL, removes the network of a learned-limiter solver. */
int luna_fv_set_limiter_network(luna_fv* fv, const float* params, int count);
/* Host copies of the last sweep's features (N x (dim + 2) rows of LUNA_LIMITER_FEATURES BF16 bit patterns, row i (dim + 2) + v)
 * and troubled flags (N bytes). Feature k of variable v: 0 the face-centroid factor phi_v (limited = phi_v x unlimited),
 * 1 (max - W)/S, 2 (W - min)/S, 3 the largest unlimited face-centroid increment / S, 4 minus the smallest / S,
 * 5 the least-squares intercept / S, 6 clipped faces / list length, 7 min over variables of phi, 8 density range /
 * density, 9 pressure range / pressure, 10 div(v) r / c, 11 |curl v| r / c, 12 |v| / c, 13 [v = density],
 * 14 [v = pressure], 15 S / scale; min, max over the neighbours (tiny faces excluded as the limiter does),
 * S = max - min + 1e-12 scale + 1e-300, scale = c for a velocity and |W| otherwise, each feature clamped to [-8, 8]. */
int luna_fv_limiter_features(luna_fv* fv, unsigned short* features, unsigned char* troubled);

/* The network's trainer, on the GPU: FP32 parameters with Adam (beta 0.9, 0.999, eps 1e-8) and the BF16 copy the kernels
 * read. A batch is rows of features x (rows x 16 BF16), targets t and weights w (rows FP32), all device pointers that are
 * ready when the call is made; the loss is sum_r w_r (sigmoid(z_r) - t_r)^2 / weight_sum. Sums run in a fixed order, so a
 * batch gives the same bits every time. */
typedef struct luna_mlp luna_mlp;
int luna_mlp_create(luna_mlp** out);
int luna_mlp_destroy(luna_mlp* mlp);
/* Sets the parameters (host, LUNA_LIMITER_PARAMS) and resets Adam. */
int luna_mlp_set_params(luna_mlp* mlp, const float* params, int count);
int luna_mlp_get_params(luna_mlp* mlp, float* params, int count);
/* fraction (device, rows FP32) = sigmoid(z) of each row. Blocks. */
int luna_mlp_forward(luna_mlp* mlp, const unsigned short* x, int rows, float* fraction);
/* The loss and its gradient in the parameters (host outputs). Blocks. */
int luna_mlp_gradient(luna_mlp* mlp, const unsigned short* x, const float* target, const float* weight, int rows, double weight_sum,
                      float* grad, float* loss);
/* One Adam step at learning rate lr; loss (host, the loss before the step) if not NULL, which blocks; else asynchronous. */
int luna_mlp_train_step(luna_mlp* mlp, const unsigned short* x, const float* target, const float* weight, int rows, double weight_sum,
                        float lr, float* loss);
int luna_mlp_synchronize(luna_mlp* mlp);

/* The last error message of this thread. */
const char* luna_last_error(void);

/* Waits for all work on the current device (hipDeviceSynchronize). */
int luna_synchronize(void);

#ifdef __cplusplus
}
#endif
#endif
