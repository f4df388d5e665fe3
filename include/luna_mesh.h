/* Periodic Voronoi geometry producer (host, C++/OpenMP): the mesh record the hydro step takes as input.
 *
 * It stands in for voronoi-cpp's voronoi::Mesh<D> output contract (the arrays finite_volume.hpp:257-327 and :353-364 read):
 * faces as owner/neighbour pairs, area vectors A = |A| n pointing owner -> neighbour, face centroids relative to the owner's
 * generator, cell volumes, cell centroids relative to their generators, generator positions snapped to a grid of 2^-30 of
 * the longest side, and per-cell face lists in CSR form whose entries carry the owner bit (f if the cell owns face f, ~f
 * otherwise). The owner of a face is its lower-numbered cell; faces are numbered by owner, then neighbour; each cell's
 * list is in ascending face order. Build time is reported, not optimised.
 */
#ifndef LUNA_MESH_H
#define LUNA_MESH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int dim, cells, faces;
    int* owner;            /* faces */
    int* neighbour;        /* faces */
    int* start;            /* cells + 1 */
    int* list;             /* start[cells] entries */
    double* area;          /* faces x dim */
    double* face_centroid; /* faces x dim, relative to the owner's generator */
    double* centroid;      /* cells x dim, relative to the cell's generator */
    double* volume;        /* cells */
    double* positions;     /* cells x dim, snapped, in [0, side) */
    double side[3];
    double step;           /* the snapping grid step */
    long long period[3];   /* grid steps per side */
    double build_seconds;
} luna_mesh_record;

/* Builds the periodic Voronoi mesh of `cells` generators (cells x dim, any real coordinates; wrapped into the box and
 * snapped). dim 2 or 3; side: dim box lengths. Returns 0, or a nonzero code with luna_mesh_last_error() set.
 * The record's arrays belong to the library until luna_mesh_release(). */
int luna_mesh_build(int dim, int cells, const double* positions, const double* side, luna_mesh_record* out);
void luna_mesh_release(luna_mesh_record* record);
const char* luna_mesh_last_error(void);

/* The derivative of a mesh's geometry in its generator positions at fixed topology (the same faces and lists), as the
 * moving step's derivatives read it: the role of voronoi-cpp's GeometryDerivative behind the reference's
 * hydro::GeometryDerivative (mesh_derivative.hpp:21-46). Each vertex is the meeting point of its cell's bisectors; the
 * volumes, centroids, area vectors and face centroids are differentiated through the producer's own formulas.
 * create: from a built record (its positions, lists and faces); the topology is that mesh's. */
typedef struct luna_mesh_derivative luna_mesh_derivative;
int luna_mesh_derivative_create(const luna_mesh_record* mesh, luna_mesh_derivative** out);
void luna_mesh_derivative_destroy(luna_mesh_derivative* derivative);
/* GeometryDerivative::tangent (mesh_derivative.hpp:26-27): dx (cells x dim) -> dV (cells), dc (cells x dim), dA, dxf
 * (faces x dim), written. */
int luna_mesh_derivative_tangent(const luna_mesh_derivative* derivative, const double* dx, double* dV, double* dc, double* dA,
                                 double* dxf);
/* GeometryDerivative::adjoint (:28-29): from the bars of V, c (cells) and A, x_f (faces), adds to xbar (cells x dim). */
int luna_mesh_derivative_adjoint(const luna_mesh_derivative* derivative, const double* Vbar, const double* cbar, const double* Abar,
                                 const double* xfbar, double* xbar);

#ifdef __cplusplus
}
#endif
#endif
