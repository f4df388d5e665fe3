// Derivatives of the static order-1 step (HLLC or HLL) on a Voronoi mesh (D = 2, 3) at fixed dt, for gfx950: the tangent
// (Q, dQ) -> (Q', J dQ) and the adjoint Qbar' -> J^T Qbar' (finite_volume.hpp:1670-2231). Every sum over a cell's faces
// runs in list order, one work-item per cell, and every face's share is computed by one work-item per face, so the
// results are deterministic (no atomics) and follow the CPU restatement's operation order. The forward pieces (fit,
// limiter, face sides) are recomputed with the forward's own arithmetic.
#include "derivatives.hpp"
#include "kernels.hpp"

namespace luna {
namespace {

constexpr int BLOCK = 64;  // one wavefront per work-group: the per-cell walks differ in length

inline dim3 grid_of(long long n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

// The heaviest per-cell kernels (the fit's tangent, the cell adjoint) need every register a lane can hold; at 2 waves per
// SIMD they spill a few bytes and run faster than at 1 (the compiler's choice without a floor).
#ifndef LUNA_DERIV_WAVES
#define LUNA_DERIV_WAVES 2
#endif

// The logical block of this work-group, as the forward kernels map them: work-groups go to the 8 XCDs in turn (block b
// to XCD b mod 8), so XCD x takes the contiguous run [x q + min(x, r), ...) of q + (x < r) blocks (q = n / 8, r = n mod
// 8), and the neighbours a cell's walk gathers sit in that XCD's L2.
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, n = gridDim.x;
    const int q = n >> 3, rem = n & 7, x = b & 7, k = b >> 3;
    return x * q + min(x, rem) + k;
}

__device__ __forceinline__ int first_slot(const Geo& g, int i) { return g.slice[i >> 6] + (i & (SLICE - 1)); }

template <int D>
__device__ __forceinline__ void load_vec(const double* p, double* v) {
#pragma unroll
    for (int a = 0; a < D; ++a) v[a] = p[a];
}

template <int D>
__device__ __forceinline__ void separation(const Geo& g, const double* xi, const double* xj, double* d) {
#pragma unroll
    for (int a = 0; a < D; ++a) d[a] = wrap(xj[a] - xi[a], g.side[a], g.inverse[a]);
}

// Geometry::face_offset (finite_volume.hpp:322-326) of face f from cell i (fc is stored relative to the owner).
template <int D>
__device__ __forceinline__ void face_offset(const Geo& g, const double* face, bool owned, const double* xi, const double* ci, int other,
                                            double* r) {
    using C = CellAt<D>;
    double rel[D];
    load_vec<D>(face + FaceAt<D>::FC, rel);
    if (!owned) {
        double xo[D], s[D];
        load_vec<D>(g.cell + (size_t)other * CELL_ROW + C::X, xo);
        separation<D>(g, xi, xo, s);
#pragma unroll
        for (int a = 0; a < D; ++a) rel[a] = rel[a] + s[a];
    }
#pragma unroll
    for (int a = 0; a < D; ++a) r[a] = rel[a] - ci[a];
}

// What the face-centroid limiter did per variable (finite_volume.hpp:709-799): the last clipping face (list position, -1
// none), its offset, the bound and the face whose neighbour holds it (the first entry in list order with the extreme).
template <int D>
struct Pattern {
    int last[D + 2], extreme[D + 2];
    bool upper[D + 2], zero[D + 2];
    double offset[D + 2][D], bound[D + 2], dbound[D + 2];
};

// The cell's fit and limiter as the sweep computes them (the same bits for g), with the unlimited gradient g0, its tangent
// dg0 (Tangent), M's Cholesky factor L and first row M0, and the limiter's pattern. dW: N x V tangents of W.
template <int D, bool Tangent>
__device__ void fit_cell(const Geo& g, int i, const double* dW, bool limit, double* gl, double* g0, double* dg0, double (*L)[D + 1],
                         double* M0, Pattern<D>& pat, double* intercept = nullptr) {
    constexpr int V = D + 2, U = D + 1;
    using C = CellAt<D>;
    using FA = FaceAt<D>;
    const double* mine = g.cell + (size_t)i * CELL_ROW;
    double Wi[V], xi[D], ci[D];
#pragma unroll
    for (int v = 0; v < V; ++v) Wi[v] = mine[C::W + v];
    load_vec<D>(mine + C::X, xi);
    load_vec<D>(mine + C::C, ci);
    double surface = 0, smallest = INFINITY;
    double lo[V], hi[V], dlo[V], dhi[V];
    int lo_f[V], hi_f[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        lo[v] = 1.7976931348623157e308;
        hi[v] = -1.7976931348623157e308;
        dlo[v] = dhi[v] = 0;
        lo_f[v] = hi_f[v] = -1;
    }
    double M[U][U], b[V][U], db[V][U];
#pragma unroll
    for (int p = 0; p < U; ++p)
#pragma unroll
        for (int s = 0; s < U; ++s) M[p][s] = 0;
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
        for (int p = 0; p < U; ++p) b[v][p] = db[v][p] = 0;
    const int count = g.count[i], base = first_slot(g, i);
    for (int k = 0; k < count; ++k) {
        const int slot = base + k * SLICE;
        const int e = g.list[slot], f = e >= 0 ? e : ~e, j = g.other[slot];
        const double* row = g.cell + (size_t)j * CELL_ROW;
        double xj[D], d[D], X[U], Wj[V];
        load_vec<D>(row + C::X, xj);
        separation<D>(g, xi, xj, d);
#pragma unroll
        for (int v = 0; v < V; ++v) Wj[v] = row[C::W + v];
        const double w = g.face[(size_t)f * FACE_ROW + FA::MAG];
        surface += w;
        smallest = fmin(smallest, w);
        X[0] = 1;
#pragma unroll
        for (int p = 1; p < U; ++p) X[p] = d[p - 1] + row[C::C + p - 1] - ci[p - 1];
#pragma unroll
        for (int p = 0; p < U; ++p)
#pragma unroll
            for (int s = p; s < U; ++s) M[p][s] += w * X[p] * X[s];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double dWv = Wj[v] - Wi[v];
#pragma unroll
            for (int p = 0; p < U; ++p) b[v][p] += w * X[p] * dWv;
            double dWj = 0;
            if constexpr (Tangent) {
                dWj = dW[(size_t)j * V + v];
                const double ddW = w * (dWj - dW[(size_t)i * V + v]);
#pragma unroll
                for (int p = 0; p < U; ++p) db[v][p] += X[p] * ddW;
            }
            if (Wj[v] < lo[v]) lo[v] = Wj[v], lo_f[v] = k, dlo[v] = dWj;
            if (Wj[v] > hi[v]) hi[v] = Wj[v], hi_f[v] = k, dhi[v] = dWj;
        }
    }
    const double tiny = 1e-6 * surface;
    const bool any_tiny = smallest < tiny;
    if (any_tiny) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            lo[v] = 1.7976931348623157e308;
            hi[v] = -1.7976931348623157e308;
            lo_f[v] = hi_f[v] = -1;
        }
        for (int k = 0; k < count; ++k) {
            const int slot = base + k * SLICE;
            const int e = g.list[slot], f = e >= 0 ? e : ~e, j = g.other[slot];
            if (g.face[(size_t)f * FACE_ROW + FA::MAG] < tiny) continue;
            const double* row = g.cell + (size_t)j * CELL_ROW + C::W;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const double dWj = Tangent ? dW[(size_t)j * V + v] : 0.0;
                if (row[v] < lo[v]) lo[v] = row[v], lo_f[v] = k, dlo[v] = dWj;
                if (row[v] > hi[v]) hi[v] = row[v], hi_f[v] = k, dhi[v] = dWj;
            }
        }
    }
    // Cholesky of M (upper triangle) and the solves, in the sweep's operation order
#pragma unroll
    for (int p = 0; p < U; ++p)
#pragma unroll
        for (int q = 0; q < U; ++q) L[p][q] = 0;
#pragma unroll
    for (int p = 0; p < U; ++p)
#pragma unroll
        for (int q = 0; q <= p; ++q) {
            double sum = M[q][p];
#pragma unroll
            for (int m = 0; m < q; ++m) sum -= L[p][m] * L[q][m];
            L[p][q] = p == q ? sqrt(sum) : sum / L[q][q];
        }
#pragma unroll
    for (int p = 0; p < U; ++p) M0[p] = M[0][p];
#pragma unroll
    for (int pass = 0; pass < (Tangent ? 2 : 1); ++pass) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            double z[U], x[U];
#pragma unroll
            for (int p = 0; p < U; ++p) {
                double sum = pass == 0 ? b[v][p] : db[v][p];
#pragma unroll
                for (int m = 0; m < p; ++m) sum -= L[p][m] * z[m];
                z[p] = sum / L[p][p];
            }
#pragma unroll
            for (int p = U - 1; p >= 0; --p) {
                double sum = z[p];
#pragma unroll
                for (int m = p + 1; m < U; ++m) sum -= L[m][p] * x[m];
                x[p] = sum / L[p][p];
            }
#pragma unroll
            for (int a = 0; a < D; ++a) {
                if (pass == 0)
                    g0[v * D + a] = gl[v * D + a] = x[1 + a];
                else
                    dg0[v * D + a] = x[1 + a];
            }
            if (pass == 0 && intercept) intercept[v] = x[0];
        }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
        pat.last[v] = pat.extreme[v] = -1;
        pat.upper[v] = pat.zero[v] = false;
    }
    if (limit)
        for (int k = 0; k < count; ++k) {
            const int slot = base + k * SLICE;
            const int e = g.list[slot], f = e >= 0 ? e : ~e;
            const double* face = g.face + (size_t)f * FACE_ROW;
            if (any_tiny && face[FA::MAG] < tiny) continue;
            double offset[D];
            face_offset<D>(g, face, e >= 0, xi, ci, g.other[slot], offset);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                double* gv = gl + v * D;
                const double dp = dot<D>(gv, offset);
                if (dp > 0 && Wi[v] + dp > hi[v]) {
                    const double s = hi[v] > Wi[v] ? (hi[v] - Wi[v]) / dp : 0.0;
#pragma unroll
                    for (int a = 0; a < D; ++a) gv[a] = gv[a] * s;
                    pat.last[v] = k, pat.upper[v] = true, pat.zero[v] = !(hi[v] > Wi[v]);
#pragma unroll
                    for (int a = 0; a < D; ++a) pat.offset[v][a] = offset[a];
                } else if (dp < 0 && Wi[v] + dp < lo[v]) {
                    const double s = lo[v] < Wi[v] ? (lo[v] - Wi[v]) / dp : 0.0;
#pragma unroll
                    for (int a = 0; a < D; ++a) gv[a] = gv[a] * s;
                    pat.last[v] = k, pat.upper[v] = false, pat.zero[v] = !(lo[v] < Wi[v]);
#pragma unroll
                    for (int a = 0; a < D; ++a) pat.offset[v][a] = offset[a];
                }
            }
        }
#pragma unroll
    for (int v = 0; v < V; ++v) {
        pat.extreme[v] = pat.last[v] < 0 ? -1 : pat.upper[v] ? hi_f[v] : lo_f[v];
        pat.bound[v] = pat.upper[v] ? hi[v] : lo[v];
        pat.dbound[v] = pat.upper[v] ? dhi[v] : dlo[v];
    }
}

// A face from the mesh's arrays, read in place: A, the face centroid, owner and neighbour, and |A| as k_geo_faces takes
// it. The per-face kernels stream these (the rows' gathers then wait on the 8 B of the cell pair, not on a 64 B record).
template <int D>
struct FaceIn {
    double A[D], fc[D], mag;
    int2 pair;
};
template <int D>
__device__ __forceinline__ void face_in(const Geo& g, int f, FaceIn<D>& x) {
    x.pair = make_int2(g.owner[f], g.neighbour[f]);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        x.A[k] = g.area[(size_t)f * D + k];
        x.fc[k] = g.fcen[(size_t)f * D + k];
    }
    x.mag = sqrt(dot<D>(x.A, x.A));
}

// One face's two sides as the static flux computes them (half step + g.r, positivity fallback), with their fallback
// flags, offsets, the unit normal and |A|.
template <int D>
__device__ __forceinline__ void face_sides(const Geo& g, const Rows& r, const Gas& gas, const FaceIn<D>& face, double dt, double* Ls, double* Rs,
                                           bool* fallback, double* ro, double* rj, double* n, double& magnitude) {
    constexpr int V = D + 2;
    using C = CellAt<D>;
    const int2 pair = face.pair;
    magnitude = face.mag;
    const double inv = 1.0 / magnitude;
#pragma unroll
    for (int k = 0; k < D; ++k) n[k] = face.A[k] * inv;
    const double* ro_row = g.cell + (size_t)pair.x * CELL_ROW;
    const double* rj_row = g.cell + (size_t)pair.y * CELL_ROW;
    double xo[D], xj[D], d[D];
    load_vec<D>(ro_row + C::X, xo);
    load_vec<D>(rj_row + C::X, xj);
    separation<D>(g, xj, xo, d);  // the owner seen from the neighbour
    const double half = 0.5 * dt, wf = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        ro[k] = face.fc[k] - ro_row[C::C + k];
        rj[k] = face.fc[k] + d[k] - rj_row[C::C + k];
        ro[k] = ro[k] + wf * half;
        rj[k] = rj[k] + wf * half;
    }
    const int cells[2] = {pair.x, pair.y};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        double* out = s == 0 ? Ls : Rs;
        const double* rs = s == 0 ? ro : rj;
        const double* W = g.cell + (size_t)cells[s] * CELL_ROW + C::W;
        const double* gr = r.grad + (size_t)cells[s] * grad_row<D>();
        double Wc[V], gc[V * D];
#pragma unroll
        for (int v = 0; v < V; ++v) Wc[v] = W[v];
#pragma unroll
        for (int k = 0; k < V * D; ++k) gc[k] = gr[k];
        half_step<D>(gas, Wc, gc, dt, out);
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] += dot<D>(gc + v * D, rs);
        fallback[s] = !(out[0] > 0 && out[D + 1] > 0);
        if (fallback[s])
#pragma unroll
            for (int v = 0; v < V; ++v) out[v] = Wc[v];
    }
}

// ---- tangent ----

// The primitive half of the cell rows (as k_primitives) and dW = J dU (dU = dq / V).
template <int D>
__global__ __launch_bounds__(BLOCK) void k_prim_tangent(Geo g, Rows r, Gas gas, const double* dq, double* dW) {
    constexpr int V = D + 2;
    using C = CellAt<D>;
    const int i = xcd_block() * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    double* row = g.cell + (size_t)i * CELL_ROW;
    const double inv = 1.0 / row[C::VOL];
    double U[V], dU[V], W[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        U[v] = r.q[(size_t)i * V + v] * inv;
        dU[v] = dq[(size_t)i * V + v] * inv;
    }
    primitive<D>(gas, U, W);
#pragma unroll
    for (int v = 0; v < V; ++v) row[C::W + v] = W[v];
    row[C::CS] = sound_speed<D>(gas, W);
    primitive_tangent<D>(gas, U, dU, dW + (size_t)i * V);
}

// The limited gradient (into r.grad) and its tangent: after the last clip g = g0 (bound - W_i) / (g0 . o), whatever clips
// came before (finite_volume.hpp:809-831).
template <int D>
__global__ __launch_bounds__(BLOCK, LUNA_DERIV_WAVES) void k_fit_tangent(Geo g, Rows r, int limit, const double* dW, double* dgrad) {
    constexpr int V = D + 2, U = D + 1, GR = grad_row<D>();
    using C = CellAt<D>;
    const int i = xcd_block() * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    double gl[V * D], g0[V * D], dg0[V * D], L[U][U], M0[U];
    Pattern<D> pat;
    fit_cell<D, true>(g, i, dW, limit != 0, gl, g0, dg0, L, M0, pat);
    const double* Wi = g.cell + (size_t)i * CELL_ROW + C::W;
    const double* dWi = dW + (size_t)i * V;
    double* out = r.grad + (size_t)i * GR;
    double* dout = dgrad + (size_t)i * GR;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const double* gv = g0 + v * D;
        const double* dgv = dg0 + v * D;
#pragma unroll
        for (int a = 0; a < D; ++a) out[v * D + a] = gl[v * D + a];
        if (pat.last[v] < 0) {
#pragma unroll
            for (int a = 0; a < D; ++a) dout[v * D + a] = dgv[a];
        } else if (pat.zero[v]) {
#pragma unroll
            for (int a = 0; a < D; ++a) dout[v * D + a] = 0.0;
        } else {
            const double* o = pat.offset[v];
            const double den = dot<D>(gv, o);
            const double s = (pat.bound[v] - Wi[v]) / den, ds = (pat.dbound[v] - dWi[v] - s * dot<D>(dgv, o)) / den;
#pragma unroll
            for (int a = 0; a < D; ++a) dout[v * D + a] = dgv[a] * s + gv[a] * ds;
        }
    }
}

// One face's flux row and its tangent (x |A|).
template <int D>
__global__ __launch_bounds__(BLOCK) void k_flux_tangent(Geo g, Rows r, Gas gas, int riemann, double dt, const double* dW, const double* dgrad,
                                                        double* dflux) {
    constexpr int V = D + 2, GR = grad_row<D>();
    using C = CellAt<D>;
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= g.F) return;
    // the face's flux and its tangent side by side (2V doubles a face), so that the update gathers one row per entry
    double* Fv = dflux + (size_t)f * 2 * V;
    double* dF = Fv + V;
    FaceIn<D> face;
    face_in<D>(g, f, face);
    if (face.mag == 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) Fv[v] = dF[v] = 0;
        return;
    }
    double Ls[V], Rs[V], ro[D], rj[D], n[D], magnitude, dLs[V], dRs[V];
    bool fallback[2];
    face_sides<D>(g, r, gas, face, dt, Ls, Rs, fallback, ro, rj, n, magnitude);
    const int2 pair = face.pair;
    const int cells[2] = {pair.x, pair.y};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        double* out = s == 0 ? dLs : dRs;
        const int i = cells[s];
        if (fallback[s]) {
#pragma unroll
            for (int v = 0; v < V; ++v) out[v] = dW[(size_t)i * V + v];
            continue;
        }
        double W[V], dWi[V], gi[V * D], dgi[V * D];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            W[v] = g.cell[(size_t)i * CELL_ROW + C::W + v];
            dWi[v] = dW[(size_t)i * V + v];
        }
#pragma unroll
        for (int k = 0; k < V * D; ++k) {
            gi[k] = r.grad[(size_t)i * GR + k];
            dgi[k] = dgrad[(size_t)i * GR + k];
        }
        half_step_tangent<D>(gas, W, dWi, gi, dgi, dt, out);
        const double* rs = s == 0 ? ro : rj;
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] += dot<D>(dgi + v * D, rs);
    }
    const double w0[D] = {};
    double Fl[V], dFl[V];
    riemann_tangent<D>(riemann, gas, Ls, Rs, dLs, dRs, n, w0, Fl, dFl);
#pragma unroll
    for (int v = 0; v < V; ++v) {
        Fv[v] = Fl[v] * magnitude;
        dF[v] = dFl[v] * magnitude;
    }
}

// Q' = Q - dt sum_k sign_k F and dQ' = dQ - dt sum_k sign_k dF, in list order.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_update_tangent(Geo g, Rows r, double dt, const double* dflux, double* dq) {
    constexpr int V = D + 2;
    const int i = xcd_block() * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    double sum[V], dsum[V];
#pragma unroll
    for (int v = 0; v < V; ++v) sum[v] = dsum[v] = 0;
    const int count = g.count[i], base = first_slot(g, i);
    constexpr int UN = 4;  // the rows of UN entries in flight, summed in list order
    int k = 0;
    for (; k + UN - 1 < count; k += UN) {
        int e[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) e[u] = g.list[base + (k + u) * SLICE];
        double F[UN][V], dF[UN][V];
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const double* row = dflux + (size_t)(e[u] >= 0 ? e[u] : ~e[u]) * 2 * V;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                F[u][v] = row[v];
                dF[u][v] = row[V + v];
            }
        }
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const double sign = e[u] >= 0 ? 1.0 : -1.0;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                sum[v] += sign * F[u][v];
                dsum[v] += sign * dF[u][v];
            }
        }
    }
    for (; k < count; ++k) {
        const int e = g.list[base + k * SLICE], f = e >= 0 ? e : ~e;
        const double sign = e >= 0 ? 1.0 : -1.0;
        const double* row = dflux + (size_t)f * 2 * V;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            sum[v] += sign * row[v];
            dsum[v] += sign * row[V + v];
        }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) {
        r.q[(size_t)i * V + v] = r.q[(size_t)i * V + v] - dt * sum[v];
        dq[(size_t)i * V + v] = dq[(size_t)i * V + v] - dt * dsum[v];
    }
}

// ---- adjoint ----

// Per face: the bars of its two sides, (L bar, R bar) = J^T Fbar of the Riemann solve, Fbar = (dt qbar_j - dt qbar_o) |A|.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_face_adjoint(Geo g, Rows r, Gas gas, int riemann, double dt, const double* qbar, double* sidebar) {
    constexpr int V = D + 2;
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= g.F) return;
    double Lbar[V], Rbar[V];
#pragma unroll
    for (int v = 0; v < V; ++v) Lbar[v] = Rbar[v] = 0;
    FaceIn<D> face;
    face_in<D>(g, f, face);
    if (face.mag != 0) {
        double Ls[V], Rs[V], ro[D], rj[D], n[D], magnitude, Fbar[V];
        bool fallback[2];
        face_sides<D>(g, r, gas, face, dt, Ls, Rs, fallback, ro, rj, n, magnitude);
        const int2 pair = face.pair;
#pragma unroll
        for (int v = 0; v < V; ++v) Fbar[v] = (dt * qbar[(size_t)pair.y * V + v] - dt * qbar[(size_t)pair.x * V + v]) * magnitude;
        const double w0[D] = {};
        riemann_adjoint<D>(riemann, gas, Ls, Rs, n, w0, Fbar, Lbar, Rbar);
    }
    double* out = sidebar + (size_t)f * 2 * V;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        out[v] = Lbar[v];
        out[V + v] = Rbar[v];
    }
}

// Per cell, in list order: its sides' adjoints (half step and g.r) into W bar and g bar; then the limited fit's adjoint
// (finite_volume.hpp:832-864): the bound's bar (for the neighbour holding it), b bar = M^-1 x bar (for the neighbours),
// and the cell's own share.
template <int D>
__global__ __launch_bounds__(BLOCK, LUNA_DERIV_WAVES) void k_cell_adjoint(Geo g, Rows r, Gas gas, double dt, int limit, const double* sidebar, double* Wbar,
                                                        double* bbar, double* boundbar, int* extreme) {
    constexpr int V = D + 2, U = D + 1, GR = grad_row<D>();
    using C = CellAt<D>;
    const int i = xcd_block() * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    double Wi[V], gi[V * D];
#pragma unroll
    for (int v = 0; v < V; ++v) Wi[v] = g.cell[(size_t)i * CELL_ROW + C::W + v];
#pragma unroll
    for (int k = 0; k < V * D; ++k) gi[k] = r.grad[(size_t)i * GR + k];
    double Wb[V], gb[V * D];
#pragma unroll
    for (int v = 0; v < V; ++v) Wb[v] = 0;
#pragma unroll
    for (int k = 0; k < V * D; ++k) gb[k] = 0;
    // This cell's side of each of its faces, as face_sides computes it for the face kernel: the half step is the cell's
    // own (the same for all its faces), plus g . r with r the face centroid in this cell's frame minus its centroid; only
    // the fallback flag and r are needed here, so the other side's rows are not read.
    double hs[V], ci[D], xi[D];
    half_step<D>(gas, Wi, gi, dt, hs);
    load_vec<D>(g.cell + (size_t)i * CELL_ROW + C::C, ci);
    load_vec<D>(g.cell + (size_t)i * CELL_ROW + C::X, xi);
    const double half = 0.5 * dt, wf = 0.0;
    const int count = g.count[i], base = first_slot(g, i);
    for (int k = 0; k < count; ++k) {
        const int e = g.list[base + k * SLICE], f = e >= 0 ? e : ~e;
        const double* face = g.face + (size_t)f * FACE_ROW;
        if (face[FaceAt<D>::MAG] == 0) continue;
        const int s = e >= 0 ? 0 : 1;
        double rr[D];
        if (s == 0) {
#pragma unroll
            for (int c = 0; c < D; ++c) {
                rr[c] = face[FaceAt<D>::FC + c] - ci[c];
                rr[c] = rr[c] + wf * half;
            }
        } else {
            const int2 pair = *reinterpret_cast<const int2*>(face + FaceAt<D>::PAIR);
            double xo[D], d[D];
            load_vec<D>(g.cell + (size_t)pair.x * CELL_ROW + C::X, xo);
            separation<D>(g, xi, xo, d);  // the owner seen from this cell, the neighbour
#pragma unroll
            for (int c = 0; c < D; ++c) {
                rr[c] = face[FaceAt<D>::FC + c] + d[c] - ci[c];
                rr[c] = rr[c] + wf * half;
            }
        }
        double Wf[V];
#pragma unroll
        for (int v = 0; v < V; ++v) Wf[v] = hs[v] + dot<D>(gi + v * D, rr);
        const bool fallback = !(Wf[0] > 0 && Wf[D + 1] > 0);
        const double* side = sidebar + (size_t)f * 2 * V + s * V;
        double sb[V];
#pragma unroll
        for (int v = 0; v < V; ++v) sb[v] = side[v];
        if (fallback) {
#pragma unroll
            for (int v = 0; v < V; ++v) Wb[v] += sb[v];
            continue;
        }
        double hb[V * D];
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int c = 0; c < D; ++c) hb[v * D + c] = sb[v] * rr[c];
        half_step_adjoint<D>(gas, Wi, gi, dt, sb, Wb, hb);
#pragma unroll
        for (int q = 0; q < V * D; ++q) gb[q] += hb[q];
    }
    double gl[V * D], g0[V * D], dg0[V * D], Lc[U][U], M0[U];
    Pattern<D> pat;
    fit_cell<D, false>(g, i, nullptr, limit != 0, gl, g0, dg0, Lc, M0, pat);
    double xbar[V][U];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        double g0bar[D];
#pragma unroll
        for (int c = 0; c < D; ++c) g0bar[c] = gb[v * D + c];
        double t = 0;
        int ext = -1;
        if (pat.last[v] >= 0 && pat.zero[v]) {
#pragma unroll
            for (int c = 0; c < D; ++c) g0bar[c] = 0;
        } else if (pat.last[v] >= 0) {
            const double* o = pat.offset[v];
            const double den = dot<D>(g0 + v * D, o);
            const double sc = (pat.bound[v] - Wi[v]) / den;
            t = dot<D>(gb + v * D, g0 + v * D) / den;
#pragma unroll
            for (int c = 0; c < D; ++c) g0bar[c] = gb[v * D + c] * sc - o[c] * (sc * t);
            const int ee = g.list[base + pat.extreme[v] * SLICE];
            ext = ee >= 0 ? ee : ~ee;
            Wb[v] -= t;
        }
        boundbar[(size_t)i * V + v] = t;
        extreme[(size_t)i * V + v] = ext;
        xbar[v][0] = 0;
#pragma unroll
        for (int c = 0; c < D; ++c) xbar[v][1 + c] = g0bar[c];
    }
    // b bar = M^-1 x bar (M symmetric), with the factor the fit used
#pragma unroll
    for (int v = 0; v < V; ++v) {
        double z[U], x[U];
#pragma unroll
        for (int p = 0; p < U; ++p) {
            double sum = xbar[v][p];
#pragma unroll
            for (int m = 0; m < p; ++m) sum -= Lc[p][m] * z[m];
            z[p] = sum / Lc[p][p];
        }
#pragma unroll
        for (int p = U - 1; p >= 0; --p) {
            double sum = z[p];
#pragma unroll
            for (int m = p + 1; m < U; ++m) sum -= Lc[m][p] * x[m];
            x[p] = sum / Lc[p][p];
        }
#pragma unroll
        for (int p = 0; p < U; ++p) {
            bbar[((size_t)i * V + v) * U + p] = x[p];
            Wb[v] -= M0[p] * x[p];
        }
        Wbar[(size_t)i * V + v] = Wb[v];
    }
}

// Per cell j, in list order: the shares of W_j in its neighbours' fits (|A| X . b bar, and the bound's bar where the
// neighbour's bound came from W_j), then the primitive adjoint: out = qbar + Ubar / V.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_gather_adjoint(Geo g, Rows r, Gas gas, const double* qbar, const double* Wbar, const double* bbar,
                                                          const double* boundbar, const int* extreme, double* out) {
    constexpr int V = D + 2, U = D + 1;
    using C = CellAt<D>;
    const int j = xcd_block() * BLOCK + threadIdx.x;
    if (j >= g.N) return;
    const double* mine = g.cell + (size_t)j * CELL_ROW;
    double xj[D], cj[D], Wb[V];
    load_vec<D>(mine + C::X, xj);
    load_vec<D>(mine + C::C, cj);
#pragma unroll
    for (int v = 0; v < V; ++v) Wb[v] = Wbar[(size_t)j * V + v];
    const int count = g.count[j], base = first_slot(g, j);
    for (int k = 0; k < count; ++k) {
        const int slot = base + k * SLICE;
        const int e = g.list[slot], f = e >= 0 ? e : ~e, i = g.other[slot];
        const double* row = g.cell + (size_t)i * CELL_ROW;
        double xi[D], d[D], X[U];
        load_vec<D>(row + C::X, xi);
        separation<D>(g, xi, xj, d);  // j seen from i, as i's fit saw it
        const double w = g.slot[slot];  // |A|, the face record's value, read coalesced from the entry's slot
        X[0] = 1;
#pragma unroll
        for (int p = 1; p < U; ++p) X[p] = d[p - 1] + cj[p - 1] - row[C::C + p - 1];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            double share = 0;
#pragma unroll
            for (int p = 0; p < U; ++p) share += w * X[p] * bbar[((size_t)i * V + v) * U + p];
            if (extreme[(size_t)i * V + v] == f) share += boundbar[(size_t)i * V + v];
            Wb[v] += share;
        }
    }
    const double inv = 1.0 / mine[C::VOL];
    double Uj[V], Ub[V];
#pragma unroll
    for (int v = 0; v < V; ++v) Uj[v] = r.q[(size_t)j * V + v] * inv;
    primitive_adjoint<D>(gas, Uj, Wb, Ub);
#pragma unroll
    for (int v = 0; v < V; ++v) out[(size_t)j * V + v] = qbar[(size_t)j * V + v] + Ub[v] * inv;
}

// ---- the moving step's reverse, order 1, fixed topology (finite_volume.hpp:2233-3155) ----
// Two passes (mesh t^n+1, then t^n), each one work-item per face writing its sides' bars into a face row and its own
// face's bars, then one per cell gathering its entries in list order; then per cell the point velocity's and the fit's
// adjoints (entry rows for the neighbours), a gather per cell and one per face. No atomics.

// Face row of a pass: [W bar, g bar, r bar] of the owner's side, the same of the neighbour's, w-hat bars of both cells,
// the bar of the separation owner -> neighbour.
template <int D>
constexpr int pass_row() {
    return 2 * ((D + 2) + (D + 2) * D + D) + 3 * D;
}
// Entry row of the fit's adjoint (per list slot): the bar of the entry's between vector, of its weight |A|, of its face
// offset (the limiter's last clip).
template <int D>
constexpr int entry_row() {
    return 2 * D + 1;
}

// euler.hpp:848-851 turn_bar(): the n bar of <a, Omega b>: out += sign (a (n.b) - b (n.a)).
template <int D>
__device__ __forceinline__ void turn_bar_add(const double* n, const double* a, const double* b, double sign, double* out) {
    const double nb = dot<D>(n, b), na = dot<D>(n, a);
#pragma unroll
    for (int k = 0; k < D; ++k) out[k] += sign * (a[k] * nb - b[k] * na);
}

// euler.hpp:893-916 MovingFace::adjoint(): G = |A| F(L, R, A / |A|, w); adds J^T Gbar to Lbar, Rbar, Abar, wbar (the normal's
// and the face velocity's parts from the flux's rotation and Galilean invariance, euler.hpp:835-841).
template <int D>
__device__ void moving_face_adjoint(int riemann_kind, const Gas& gas, const double* L, const double* R, const double* A, const double* w,
                                    const double* Gbar, double* Lbar, double* Rbar, double* Abar, double* wbar) {
    constexpr int V = D + 2;
    const double s = sqrt(dot<D>(A, A));
    double n[D], F[V], Fr[V], X[V], mw[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        n[k] = A[k] * (1.0 / s);
        mw[k] = -w[k];
    }
    riemann<D>(riemann_kind, gas, L, R, n, w, F);
#pragma unroll
    for (int v = 0; v < V; ++v) Fr[v] = F[v];
    to_lab<D>(Fr, mw);  // the flux in the face's frame
    double sbar = 0;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        sbar += Gbar[v] * F[v];
        X[v] = s * Gbar[v];
    }
    double nbar[D], shiftbar[D];
#pragma unroll
    for (int k = 0; k < D; ++k) nbar[k] = 0;
    turn_bar_add<D>(n, X + 1, F + 1, 1.0, nbar);
#pragma unroll
    for (int k = 0; k < D; ++k) shiftbar[k] = X[1 + k] * Fr[0] + X[D + 1] * (Fr[1 + k] + w[k] * Fr[0]);
    double Lr[V], Rr[V];
#pragma unroll
    for (int v = 0; v < V; ++v) Lr[v] = Rr[v] = 0;
    riemann_adjoint<D>(riemann_kind, gas, L, R, n, w, X, Lr, Rr);
#pragma unroll
    for (int v = 0; v < V; ++v) {
        Lbar[v] += Lr[v];
        Rbar[v] += Rr[v];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) shiftbar[k] -= Lr[1 + k] + Rr[1 + k];
    turn_bar_add<D>(n, Lr + 1, L + 1, -1.0, nbar);
    turn_bar_add<D>(n, Rr + 1, R + 1, -1.0, nbar);
#pragma unroll
    for (int k = 0; k < D; ++k) wbar[k] += shiftbar[k];
    turn_bar_add<D>(n, shiftbar, w, -1.0, nbar);
    const double nn = dot<D>(n, nbar);
#pragma unroll
    for (int k = 0; k < D; ++k) Abar[k] += (nbar[k] - n[k] * nn) * (1.0 / s) + n[k] * sbar;
}

// FaceFlux::face_velocity() (finite_volume.hpp:1380-1397) and its adjoint (:2331-2369), at the forward's branch.
template <int D>
__device__ __forceinline__ void face_velocity(const double* d, const double* xf, const double* wL, const double* wR, double* wf) {
    double c[D], dw[D];
    const double dist = sqrt(dot<D>(d, d)), inv_dist = 1.0 / dist;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        c[a] = xf[a] - d[a] * 0.5;
        dw[a] = wL[a] - wR[a];
    }
    double facv = dot<D>(c, dw) * inv_dist;
    const double limit = 0.9 * dist, c2 = dot<D>(c, c);
    if (c2 > limit * limit) facv *= limit / sqrt(c2);
#pragma unroll
    for (int a = 0; a < D; ++a) wf[a] = (wL[a] + wR[a]) * 0.5 + d[a] * (facv * inv_dist);
}
template <int D>
__device__ void face_velocity_adjoint(const double* d, const double* xf, const double* wL, const double* wR, const double* wfbar, double* dbar,
                                      double* xfbar, double* wLbar, double* wRbar) {
    double c[D], u[D], cbar[D];
    const double dist = sqrt(dot<D>(d, d));
#pragma unroll
    for (int a = 0; a < D; ++a) {
        c[a] = xf[a] - d[a] * 0.5;
        u[a] = wL[a] - wR[a];
        cbar[a] = 0;
    }
    const double f0 = dot<D>(c, u) / dist, limit = 0.9 * dist, c2 = dot<D>(c, c), cc = sqrt(c2);
    const bool clipped = c2 > limit * limit;
    const double sc = clipped ? limit / cc : 1.0, facv = f0 * sc;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        wLbar[a] += wfbar[a] * 0.5;
        wRbar[a] += wfbar[a] * 0.5;
        dbar[a] += wfbar[a] * (facv / dist);
    }
    const double qb = dot<D>(wfbar, d), facvbar = qb / dist;
    double distbar = -qb * facv / (dist * dist);
    const double f0bar = facvbar * sc;
    if (clipped) {
        const double scbar = facvbar * f0;
        distbar += scbar * 0.9 / cc;
#pragma unroll
        for (int a = 0; a < D; ++a) cbar[a] += c[a] * (-scbar * sc / (cc * cc));
    }
    distbar -= f0bar * f0 / dist;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        cbar[a] += u[a] * (f0bar / dist);
        const double ub = c[a] * (f0bar / dist);
        wLbar[a] += ub;
        wRbar[a] -= ub;
        xfbar[a] += cbar[a];
        dbar[a] += -cbar[a] * 0.5 + d[a] * (distbar / dist);
    }
}

// One moving pass's face (order0: the order-0 step's pass, sides W, w-hat = w, Gbar = -dt (qbar_o - qbar_j), no gradient
// bars): its two sides (first pass W + g.r; second half_step(W, g, 2 dt) + g.(r + c - c^n + dt w-hat)),
// G = |A| F at the face velocity, Gbar = -dt/2 (qbar_o - qbar_j); the sides' bars (W, g, offset), the w-hat bars, the
// separation's bar into the row; A's and x_f's bars (x_f also from both offsets) into this face's rows of Ab, Xfb.
template <int D, bool Second>
__global__ __launch_bounds__(BLOCK) void k_mpass_face(Geo g, Rows r, Gas gas, int riemann_kind, double dt, double unit, const double* qbar,
                                                      double* row, double* Ab, double* Xfb, int order0) {
    constexpr int V = D + 2, RW = pass_row<D>(), GR = grad_row<D>(), SIDE = V + V * D + D;
    using C = CellAt<D>;
    using FA = FaceAt<D>;
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= g.F) return;
    double* out = row + (size_t)f * RW;
#pragma unroll
    for (int k = 0; k < RW; ++k) out[k] = 0;
    double Abar[D], Xfbar[D];
#pragma unroll
    for (int a = 0; a < D; ++a) Abar[a] = Xfbar[a] = 0;
    const double* face = g.face + (size_t)f * FACE_ROW;
    if (face[FA::MAG] != 0) {
        const int2 pair = *reinterpret_cast<const int2*>(face + FA::PAIR);
        const int cells[2] = {pair.x, pair.y};
        double A[D], fc[D], xs[2][D], d[D], back[D], wh[2][D], wf[D];
        load_vec<D>(face + FA::A, A);
        load_vec<D>(face + FA::FC, fc);
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
            load_vec<D>(g.cell + (size_t)cells[s2] * CELL_ROW + C::X, xs[s2]);
#pragma unroll
            for (int a = 0; a < D; ++a)  // order 0: the sweep's point velocity itself
                wh[s2][a] = order0 ? r.w[(size_t)cells[s2] * D + a] : (double)r.steps[(size_t)cells[s2] * D + a] * unit;
        }
        separation<D>(g, xs[0], xs[1], d);     // owner -> neighbour
        separation<D>(g, xs[1], xs[0], back);  // neighbour -> owner
        face_velocity<D>(d, fc, wh[0], wh[1], wf);
        double W[2][V], gr[2][V * D], rr[2][D], sides[2][V];
        bool fallback[2];
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
            const int i = cells[s2];
            const double* cr = g.cell + (size_t)i * CELL_ROW;
#pragma unroll
            for (int v = 0; v < V; ++v) W[s2][v] = cr[C::W + v];
#pragma unroll
            for (int k = 0; k < V * D; ++k) gr[s2][k] = order0 ? 0.0 : r.grad[(size_t)i * GR + k];  // order 0: no gradient
#pragma unroll
            for (int a = 0; a < D; ++a) {
                rr[s2][a] = (s2 == 0 ? fc[a] : fc[a] + back[a]) - cr[C::C + a];
                if constexpr (Second) rr[s2][a] = rr[s2][a] + cr[C::C + a] - r.old_centroid[(size_t)i * D + a] + wh[s2][a] * (0.5 * (2 * dt));
            }
            if constexpr (Second) {
                half_step<D>(gas, W[s2], gr[s2], 2 * dt, sides[s2]);
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) sides[s2][v] = W[s2][v];
            }
#pragma unroll
            for (int v = 0; v < V; ++v) sides[s2][v] += dot<D>(gr[s2] + v * D, rr[s2]);
            fallback[s2] = !(sides[s2][0] > 0 && sides[s2][D + 1] > 0);
            if (fallback[s2])
#pragma unroll
                for (int v = 0; v < V; ++v) sides[s2][v] = W[s2][v];
        }
        double Gbar[V], Lbar[V], Rbar[V], wfbar[D], dbar[D];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            Gbar[v] = (order0 ? -dt : -0.5 * dt) * (qbar[(size_t)pair.x * V + v] - qbar[(size_t)pair.y * V + v]);
            Lbar[v] = Rbar[v] = 0;
        }
#pragma unroll
        for (int a = 0; a < D; ++a) wfbar[a] = dbar[a] = 0;
        moving_face_adjoint<D>(riemann_kind, gas, sides[0], sides[1], A, wf, Gbar, Lbar, Rbar, Abar, wfbar);
        double* whb_o = out + 2 * SIDE;
        double* whb_j = whb_o + D;
        face_velocity_adjoint<D>(d, fc, wh[0], wh[1], wfbar, dbar, Xfbar, whb_o, whb_j);
#pragma unroll
        for (int a = 0; a < D; ++a) out[2 * SIDE + 2 * D + a] = dbar[a];
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
            const double* sb = s2 == 0 ? Lbar : Rbar;
            double* o = out + s2 * SIDE;  // [W bar V, g bar V D, r bar D]
            if (fallback[s2]) {
#pragma unroll
                for (int v = 0; v < V; ++v) o[v] = sb[v];
                continue;
            }
            double Wb[V], gb[V * D];
#pragma unroll
            for (int v = 0; v < V; ++v) Wb[v] = 0;
#pragma unroll
            for (int k = 0; k < V * D; ++k) gb[k] = 0;
            if constexpr (Second) {
                half_step_adjoint<D>(gas, W[s2], gr[s2], 2 * dt, sb, Wb, gb);
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) Wb[v] += sb[v];
            }
            double rb[D];
#pragma unroll
            for (int a = 0; a < D; ++a) rb[a] = 0;
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    if (!order0) gb[v * D + a] += sb[v] * rr[s2][a];
                    rb[a] += sb[v] * gr[s2][v * D + a];
                }
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = Wb[v];
#pragma unroll
            for (int k = 0; k < V * D; ++k) o[V + k] = gb[k];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                o[V + V * D + a] = rb[a];
                Xfbar[a] += rb[a];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) {
        Ab[(size_t)f * D + a] = Abar[a];
        Xfb[(size_t)f * D + a] = Xfbar[a];
    }
}

// Per cell, in list order: its sides' rows into W bar, g bar, w-hat bar; the offsets' bars into c bar (and on the second
// pass into c^n's bar and w-hat's); the separations' bars into x bar. Adds to every array.
template <int D, bool Second>
__global__ __launch_bounds__(BLOCK) void k_mpass_gather(Geo g, double dt, const double* row, double* Wbar, double* gbar, double* whbar,
                                                        double* c0bar, double* cbar, double* xbar) {
    constexpr int V = D + 2, RW = pass_row<D>(), SIDE = V + V * D + D;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    double Wb[V], gb[V * D], whb[D], c0b[D], cb[D], xb[D];
#pragma unroll
    for (int v = 0; v < V; ++v) Wb[v] = Wbar[(size_t)i * V + v];
#pragma unroll
    for (int k = 0; k < V * D; ++k) gb[k] = gbar[(size_t)i * V * D + k];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        whb[a] = whbar[(size_t)i * D + a];
        c0b[a] = c0bar[(size_t)i * D + a];
        cb[a] = cbar[(size_t)i * D + a];
        xb[a] = xbar[(size_t)i * D + a];
    }
    const int count = g.count[i], base = first_slot(g, i);
    for (int k = 0; k < count; ++k) {
        const int e = g.list[base + k * SLICE], f = e >= 0 ? e : ~e;
        const double* rw = row + (size_t)f * RW;
        const int s2 = e >= 0 ? 0 : 1;
        const double* o = rw + s2 * SIDE;
#pragma unroll
        for (int v = 0; v < V; ++v) Wb[v] += o[v];
#pragma unroll
        for (int q = 0; q < V * D; ++q) gb[q] += o[V + q];
        const double* rb = o + V + V * D;
        const double* wb = rw + 2 * SIDE + s2 * D;
        const double* db = rw + 2 * SIDE + 2 * D;
        const double* rb_j = rw + SIDE + V + V * D;  // the neighbour side's offset bar: its separation neighbour -> owner
#pragma unroll
        for (int a = 0; a < D; ++a) {
            whb[a] += wb[a];
            cb[a] -= rb[a];
            if constexpr (Second) {
                cb[a] += rb[a];
                c0b[a] -= rb[a];
                whb[a] += dt * rb[a];
            }
            if (s2 == 0)
                xb[a] += rb_j[a] - db[a];
            else
                xb[a] += db[a] - rb_j[a];
        }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) Wbar[(size_t)i * V + v] = Wb[v];
#pragma unroll
    for (int k = 0; k < V * D; ++k) gbar[(size_t)i * V * D + k] = gb[k];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        whbar[(size_t)i * D + a] = whb[a];
        c0bar[(size_t)i * D + a] = c0b[a];
        cbar[(size_t)i * D + a] = cb[a];
        xbar[(size_t)i * D + a] = xb[a];
    }
}

// x^n+1 = x^n + dt w-hat: w-hat's bar from x^n+1's (the quantisation the identity).
__global__ __launch_bounds__(BLOCK) void k_drift_adjoint(long long n, double dt, const double* xbar, double* whbar) {
    const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (k < n) whbar[k] += dt * xbar[k];
}

// finite_volume.hpp:2237-2243 face_angle_slope()
template <int D>
__device__ __forceinline__ double face_angle_slope(double area) {
    if constexpr (D == 3) return area > 0 ? 0.5 / (M_PI * sqrt(area / M_PI)) : 0.0;
    return 0.5;
}

// Per cell on t^n: the point velocity's adjoint (finite_volume.hpp:2245-2306 Regularised, unless the steps are held), the
// sound speed's, and the limited fit's with its geometry (:2629-2760): W_i's own bar, beta = M^-1 xbar, the bound's bar;
// per list slot the bars of its between vector, weight and face offset; the angle face and the bars of its A and
// separation; the cell's own c and x bars.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_mcell_adjoint(Geo g, Rows r, Gas gas, int limit, int steps_fixed, double* Wbar, const double* gbar,
                                                         const double* whbar, double* cbar, double* xbar, double* beta, double* boundbar,
                                                         int* extreme, double* entry, int* fmax_of, double* angle) {
    constexpr int V = D + 2, U = D + 1, ER = entry_row<D>();
    using C = CellAt<D>;
    using FA = FaceAt<D>;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    const double* mine = g.cell + (size_t)i * CELL_ROW;
    double Wi[V], xi[D], ci[D], Wb[V], cb[D], xb[D];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        Wi[v] = mine[C::W + v];
        Wb[v] = Wbar[(size_t)i * V + v];
    }
    load_vec<D>(mine + C::X, xi);
    load_vec<D>(mine + C::C, ci);
#pragma unroll
    for (int a = 0; a < D; ++a) {
        cb[a] = cbar[(size_t)i * D + a];
        xb[a] = xbar[(size_t)i * D + a];
    }
    const int count = g.count[i], base = first_slot(g, i);
    // the point velocity w = v + c (fraction c_s / |c|), fraction from the largest face angle
    double csbar = 0, angle_A[D], angle_d[D];
    int fmax = -1;
#pragma unroll
    for (int a = 0; a < D; ++a) angle_A[a] = angle_d[a] = 0;
    if (!steps_fixed) {
        double amax = 0, dmax[D];
        int other = -1;
#pragma unroll
        for (int a = 0; a < D; ++a) dmax[a] = 0;
        for (int k = 0; k < count; ++k) {
            const int slot = base + k * SLICE;
            const int e = g.list[slot], f = e >= 0 ? e : ~e, j = g.other[slot];
            double xj[D], d[D];
            load_vec<D>(g.cell + (size_t)j * CELL_ROW + C::X, xj);
            separation<D>(g, xi, xj, d);
            const double ang = face_angle<D>(g.face[(size_t)f * FACE_ROW + FA::MAG]) / (0.5 * sqrt(dot<D>(d, d)));
            if (ang > amax) {
                amax = ang;
                fmax = f;
                other = j;
#pragma unroll
                for (int a = 0; a < D; ++a) dmax[a] = d[a];
            }
        }
        const double Fc = 2.25, shaping = 0.5;
        double fraction = 0, slope = 0;
        if (amax > 0.75 * Fc) {
            fraction = amax > Fc ? shaping : shaping * (amax - 0.75 * Fc) / (0.25 * Fc);
            slope = amax > Fc ? 0.0 : shaping / (0.25 * Fc);
        }
        double wbar[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            wbar[a] = whbar[(size_t)i * D + a];
            Wb[1 + a] += wbar[a];
        }
        const double cd = sqrt(dot<D>(ci, ci)), cs = mine[C::CS];
        if (cd > 0 && fraction > 0) {
            const double k = fraction * cs / cd, kbar = dot<D>(wbar, ci);
            const double fractionbar = kbar * cs / cd, cdbar = -kbar * k / cd;
            csbar += kbar * fraction / cd;
#pragma unroll
            for (int a = 0; a < D; ++a) cb[a] += wbar[a] * k + ci[a] * (cdbar / cd);
            if (slope != 0) {
                const double abar = slope * fractionbar;
                const double* A = g.face + (size_t)fmax * FACE_ROW + FA::A;
                const double mag = sqrt(dot<D>(A, A)), dist = sqrt(dot<D>(dmax, dmax));
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    angle_A[a] = A[a] * (abar * face_angle_slope<D>(mag) / mag / (0.5 * dist));
                    angle_d[a] = -dmax[a] * (abar * amax / (dist * dist));
                    xb[a] -= angle_d[a];  // d = x_other - x_i
                }
            }
        }
        (void)other;
    }
    fmax_of[i] = fmax;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        angle[(size_t)i * 2 * D + a] = angle_A[a];
        angle[(size_t)i * 2 * D + D + a] = angle_d[a];
    }
    if (mine[C::CS] > 0) {  // c_s = sqrt(gamma P / rho)
        const double c2bar = csbar / (2 * mine[C::CS]);
        Wb[D + 1] += c2bar * gas.gamma / Wi[0];
        Wb[0] -= c2bar * gas.gamma * Wi[D + 1] / (Wi[0] * Wi[0]);
    }
    // the limited fit
    double gl[V * D], g0[V * D], dg0[V * D], Lc[U][U], M0[U], icpt[V], gb[V * D];
    Pattern<D> pat;
    fit_cell<D, false>(g, i, nullptr, limit != 0, gl, g0, dg0, Lc, M0, pat, icpt);
#pragma unroll
    for (int k = 0; k < V * D; ++k) gb[k] = gbar[(size_t)i * V * D + k];
    for (int k = 0; k < count; ++k) {
        double* er = entry + (size_t)(base + k * SLICE) * ER;
#pragma unroll
        for (int q = 0; q < ER; ++q) er[q] = 0;
    }
    double xbv[V][U];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        double g0bar[D];
#pragma unroll
        for (int a = 0; a < D; ++a) g0bar[a] = gb[v * D + a];
        double t = 0;
        int ext = -1;
        if (pat.last[v] >= 0 && pat.zero[v]) {
#pragma unroll
            for (int a = 0; a < D; ++a) g0bar[a] = 0;
        } else if (pat.last[v] >= 0) {
            const double* o = pat.offset[v];
            const double den = dot<D>(g0 + v * D, o);
            const double sc = (pat.bound[v] - Wi[v]) / den;
            t = dot<D>(gb + v * D, g0 + v * D) / den;
            double* ob = entry + (size_t)(base + pat.last[v] * SLICE) * ER + D + 1;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                g0bar[a] = gb[v * D + a] * sc - o[a] * (sc * t);
                ob[a] += -sc * t * g0[v * D + a];
            }
            const int ee = g.list[base + pat.extreme[v] * SLICE];
            ext = ee >= 0 ? ee : ~ee;
            Wb[v] -= t;
        }
        boundbar[(size_t)i * V + v] = t;
        extreme[(size_t)i * V + v] = ext;
        xbv[v][0] = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) xbv[v][1 + a] = g0bar[a];
    }
    double bb[V][U], x[V][U];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        double z[U];
#pragma unroll
        for (int p = 0; p < U; ++p) {
            double sum = xbv[v][p];
#pragma unroll
            for (int m = 0; m < p; ++m) sum -= Lc[p][m] * z[m];
            z[p] = sum / Lc[p][p];
        }
#pragma unroll
        for (int p = U - 1; p >= 0; --p) {
            double sum = z[p];
#pragma unroll
            for (int m = p + 1; m < U; ++m) sum -= Lc[m][p] * bb[v][m];
            bb[v][p] = sum / Lc[p][p];
        }
#pragma unroll
        for (int p = 0; p < U; ++p) {
            beta[((size_t)i * V + v) * U + p] = bb[v][p];
            Wb[v] -= M0[p] * bb[v][p];
        }
        x[v][0] = icpt[v];
#pragma unroll
        for (int a = 0; a < D; ++a) x[v][1 + a] = g0[v * D + a];
    }
    double Mbar[U][U];
#pragma unroll
    for (int p = 0; p < U; ++p)
#pragma unroll
        for (int q = 0; q < U; ++q) {
            double m = 0;
#pragma unroll
            for (int v = 0; v < V; ++v) m -= bb[v][p] * x[v][q];
            Mbar[p][q] = m;
        }
    for (int k = 0; k < count; ++k) {
        const int slot = base + k * SLICE;
        const int e = g.list[slot], f = e >= 0 ? e : ~e, j = g.other[slot];
        const double* row = g.cell + (size_t)j * CELL_ROW;
        double xj[D], d[D], X[U], Xbar[U];
        load_vec<D>(row + C::X, xj);
        separation<D>(g, xi, xj, d);
        const double w = g.face[(size_t)f * FACE_ROW + FA::MAG];
        X[0] = 1;
#pragma unroll
        for (int p = 1; p < U; ++p) X[p] = d[p - 1] + row[C::C + p - 1] - ci[p - 1];
#pragma unroll
        for (int p = 0; p < U; ++p) Xbar[p] = 0;
        double wbar = 0;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double dW = row[C::W + v] - Wi[v];
            double xb2 = 0;
#pragma unroll
            for (int p = 0; p < U; ++p) xb2 += X[p] * bb[v][p];
            wbar += xb2 * dW;
#pragma unroll
            for (int p = 0; p < U; ++p) Xbar[p] += w * dW * bb[v][p];
        }
#pragma unroll
        for (int p = 0; p < U; ++p)
#pragma unroll
            for (int q = 0; q < U; ++q) {
                wbar += X[p] * Mbar[p][q] * X[q];
                Xbar[p] += w * (Mbar[p][q] + Mbar[q][p]) * X[q];
            }
        double* er = entry + (size_t)slot * ER;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            er[a] = Xbar[1 + a];
            // between = d + c_j - c_i (own parts: - c_i, - x_i); offset = x_f (+ x_o - x_i if not owned) - c_i
            cb[a] -= Xbar[1 + a] + er[D + 1 + a];
            xb[a] -= Xbar[1 + a];
            if (e < 0) xb[a] -= er[D + 1 + a];
        }
        er[D] = wbar;
    }
#pragma unroll
    for (int v = 0; v < V; ++v) Wbar[(size_t)i * V + v] = Wb[v];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        cbar[(size_t)i * D + a] = cb[a];
        xbar[(size_t)i * D + a] = xb[a];
    }
}

// The slot of face f in cell i's list.
__device__ __forceinline__ int slot_of(const Geo& g, int i, int f) {
    const int count = g.count[i], base = first_slot(g, i);
    for (int k = 0; k < count; ++k) {
        const int e = g.list[base + k * SLICE];
        if ((e >= 0 ? e : ~e) == f) return base + k * SLICE;
    }
    return base;
}

// Per cell j on t^n, in list order: W_j's shares in its neighbours' fits and bounds; the neighbours' entry rows' bars of
// c_j and x_j (their between vectors, their offsets through x_owner, their angle face's separation); then the primitive
// adjoint: out = qbar + Ubar / V, and V's bar.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_mgather_adjoint(Geo g, Rows r, Gas gas, const double* qbar, const double* Wbar, const double* beta,
                                                           const double* boundbar, const int* extreme, const double* entry, const int* fmax_of,
                                                           const double* angle, double* cbar, double* xbar, double* Vbar, double* out) {
    constexpr int V = D + 2, U = D + 1, ER = entry_row<D>();
    using C = CellAt<D>;
    const int j = blockIdx.x * BLOCK + threadIdx.x;
    if (j >= g.N) return;
    const double* mine = g.cell + (size_t)j * CELL_ROW;
    double xj[D], cj[D], Wb[V], cb[D], xb[D];
    load_vec<D>(mine + C::X, xj);
    load_vec<D>(mine + C::C, cj);
#pragma unroll
    for (int v = 0; v < V; ++v) Wb[v] = Wbar[(size_t)j * V + v];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        cb[a] = cbar[(size_t)j * D + a];
        xb[a] = xbar[(size_t)j * D + a];
    }
    const int count = g.count[j], base = first_slot(g, j);
    for (int k = 0; k < count; ++k) {
        const int slot = base + k * SLICE;
        const int e = g.list[slot], f = e >= 0 ? e : ~e, i = g.other[slot];
        const double* row = g.cell + (size_t)i * CELL_ROW;
        double xi[D], d[D], X[U];
        load_vec<D>(row + C::X, xi);
        separation<D>(g, xi, xj, d);
        const double w = g.face[(size_t)f * FACE_ROW + FaceAt<D>::MAG];
        X[0] = 1;
#pragma unroll
        for (int p = 1; p < U; ++p) X[p] = d[p - 1] + cj[p - 1] - row[C::C + p - 1];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            double share = 0;
#pragma unroll
            for (int p = 0; p < U; ++p) share += w * X[p] * beta[((size_t)i * V + v) * U + p];
            if (extreme[(size_t)i * V + v] == f) share += boundbar[(size_t)i * V + v];
            Wb[v] += share;
        }
        const double* er = entry + (size_t)slot_of(g, i, f) * ER;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            cb[a] += er[a];  // i's between = x_j - x_i + c_j - c_i
            xb[a] += er[a];
            if (e >= 0) xb[a] += er[D + 1 + a];  // j owns f: i's offset carries x_j - x_i
            if (fmax_of[i] == f) xb[a] += angle[(size_t)i * 2 * D + D + a];
        }
    }
    const double inv = 1.0 / mine[C::VOL];
    double Uj[V], Ub[V];
#pragma unroll
    for (int v = 0; v < V; ++v) Uj[v] = r.q[(size_t)j * V + v] * inv;
    primitive_adjoint<D>(gas, Uj, Wb, Ub);
    double vb = 0;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        out[(size_t)j * V + v] = qbar[(size_t)j * V + v] + Ub[v] * inv;
        vb -= Ub[v] * Uj[v] * inv;
    }
    Vbar[j] = vb;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        cbar[(size_t)j * D + a] = cb[a];
        xbar[(size_t)j * D + a] = xb[a];
    }
}

// Per face on t^n: A's bar from the two entries' weight bars and the angle faces; x_f's from the two entries' offsets.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_mfaces_bar(Geo g, const double* entry, const int* fmax_of, const double* angle, double* Ab, double* Xfb) {
    constexpr int ER = entry_row<D>();
    using FA = FaceAt<D>;
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= g.F) return;
    const double* face = g.face + (size_t)f * FACE_ROW;
    const int2 pair = *reinterpret_cast<const int2*>(face + FA::PAIR);
    const double* eo = entry + (size_t)slot_of(g, pair.x, f) * ER;
    const double* ej = entry + (size_t)slot_of(g, pair.y, f) * ER;
    const double mag = face[FA::MAG];
    const double wb = mag > 0 ? (eo[D] + ej[D]) / mag : 0.0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        Ab[(size_t)f * D + a] += face[FA::A + a] * wb + (fmax_of[pair.x] == f ? angle[(size_t)pair.x * 2 * D + a] : 0.0) +
                                 (fmax_of[pair.y] == f ? angle[(size_t)pair.y * 2 * D + a] : 0.0);
        Xfb[(size_t)f * D + a] += eo[D + 1 + a] + ej[D + 1 + a];
    }
}


// ---- the moving step's tangent (forward mode, fixed topology; finite_volume.hpp:2428-2486, 2764-2870) ----
// Device rows of the tangent: the geometry's tangents of the pass's mesh (dV, dc per cell; dA, dxf per face), the positions'
// (dx), c^n's (dc0, the first mesh's dc), the W and c_s tangents (dW: N x V, dcs: N), the point velocities' (dw) and the
// limited gradients' (dg: N x V D); the faces' dG rows.

// FaceVelocity::tangent (finite_volume.hpp:2322-2330), at the forward's branch of the 0.9 |d| limit.
template <int D>
__device__ __forceinline__ void face_velocity_tangent(const double* d, const double* xf, const double* wL, const double* wR, const double* dd,
                                                      const double* dxf, const double* dwL, const double* dwR, double* dwf) {
    double c[D], u[D], dc[D], du[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        c[a] = xf[a] - d[a] * 0.5;
        u[a] = wL[a] - wR[a];
        dc[a] = dxf[a] - dd[a] * 0.5;
        du[a] = dwL[a] - dwR[a];
    }
    const double dist = sqrt(dot<D>(d, d)), ddist = dot<D>(d, dd) / dist;
    const double f0 = dot<D>(c, u) / dist, df0 = (dot<D>(dc, u) + dot<D>(c, du)) / dist - f0 * ddist / dist;
    const double limit = 0.9 * dist, c2 = dot<D>(c, c);
    double facv = f0, dfacv = df0;
    if (c2 > limit * limit) {
        const double cc = sqrt(c2), sc = limit / cc, dsc = 0.9 * ddist / cc - sc * dot<D>(c, dc) / c2;
        facv = f0 * sc;
        dfacv = df0 * sc + f0 * dsc;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) dwf[a] = (dwL[a] + dwR[a]) * 0.5 + dd[a] * (facv / dist) + d[a] * (dfacv / dist - facv * ddist / (dist * dist));
}

// MovingFace::tangent (euler.hpp:861-891): dG of G = |A| F(L, R, A / |A|, w). F = to_lab(Phi(L - w, R - w, n), w); Phi's change
// with n from its rotation covariance (Omega = dn n^T - n dn^T): one Riemann tangent at (L, R) with dL - dw - Omega L_w and
// dR - dw - Omega R_w, plus Omega Phi and to_lab's own change with w.
template <int D>
__device__ void moving_face_tangent(int riemann_kind, const Gas& gas, const double* L, const double* R, const double* A, const double* w,
                                    const double* dL, const double* dR, const double* dA, const double* dw, double* dG) {
    constexpr int V = D + 2, P = D + 1;
    const double s = sqrt(dot<D>(A, A));
    double n[D], dn[D], mw[D], lw[D], rw[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        n[k] = A[k] * (1.0 / s);
        mw[k] = -w[k];
        lw[k] = L[1 + k] - w[k];
        rw[k] = R[1 + k] - w[k];
    }
    const double ds = dot<D>(n, dA);
#pragma unroll
    for (int k = 0; k < D; ++k) dn[k] = (dA[k] - n[k] * ds) / s;
    double eL[V], eR[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        eL[v] = dL[v];
        eR[v] = dR[v];
    }
    const double nl = dot<D>(n, lw), dnl = dot<D>(dn, lw), nr = dot<D>(n, rw), dnr = dot<D>(dn, rw);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        eL[1 + k] -= dw[k] + (dn[k] * nl - n[k] * dnl);
        eR[1 + k] -= dw[k] + (dn[k] * nr - n[k] * dnr);
    }
    double F[V], dF[V], Phi[V], X[V];
    riemann_tangent<D>(riemann_kind, gas, L, R, eL, eR, n, w, F, dF);
#pragma unroll
    for (int v = 0; v < V; ++v) Phi[v] = F[v];
    to_lab<D>(Phi, mw);
    const double nP = dot<D>(n, Phi + 1), dnP = dot<D>(dn, Phi + 1);
    X[0] = 0;
    X[P] = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) X[1 + k] = dn[k] * nP - n[k] * dnP;
    to_lab<D>(X, w);
#pragma unroll
    for (int v = 0; v < V; ++v) dF[v] += X[v];
    dF[P] += dot<D>(dw, Phi + 1) + dot<D>(w, dw) * Phi[0];
#pragma unroll
    for (int k = 0; k < D; ++k) dF[1 + k] += dw[k] * Phi[0];
#pragma unroll
    for (int v = 0; v < V; ++v) dG[v] = ds * F[v] + s * dF[v];
}

// The tangent of face_offset(f, i): dx_f - dc_i, plus the owner's separation's when i does not own f.
template <int D>
__device__ __forceinline__ void offset_dot(const double* dxf, bool owned, const double* dxi, const double* dxo, const double* dci, double* dr) {
#pragma unroll
    for (int a = 0; a < D; ++a) dr[a] = dxf[a] - dci[a] + (owned ? 0.0 : dxo[a] - dxi[a]);
}

// W, c_s tangents per cell (finite_volume.hpp:2408-2426 moving_primitives): dU = (dq - U dV) / V.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_mprim_tangent(Geo g, Rows r, Gas gas, const double* dq, const double* dV, double* dW, double* dcs) {
    constexpr int V = D + 2;
    using C = CellAt<D>;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    const double* row = g.cell + (size_t)i * CELL_ROW;
    const double inv = 1.0 / row[C::VOL];
    double U[V], dU[V], Wt[V];
#pragma unroll
    for (int v = 0; v < V; ++v) U[v] = r.q[(size_t)i * V + v] * inv;
#pragma unroll
    for (int v = 0; v < V; ++v) dU[v] = (dq[(size_t)i * V + v] - U[v] * dV[i]) * inv;
    primitive_tangent<D>(gas, U, dU, Wt);
    const double rho = row[C::W], P = row[C::W + D + 1], cs = row[C::CS];
    double* out = dW + (size_t)i * V;
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = Wt[v];
    dcs[i] = cs > 0 ? 0.5 * gas.gamma * (Wt[D + 1] / rho - P * Wt[0] / (rho * rho)) / cs : 0.0;
}

// Per cell on t^n: the point velocity's tangent (finite_volume.hpp:2251-2306 Regularised::tangent) and, at order 1, the limited
// fit's with its geometry (:2687-2762 moving_gradient_tangent: x_v = M^-1 b_v moves by M^-1 (db_v - dM x_v); after the last
// clip g = g0 (bound - W_i) / (g0 . o), o that face's offset, which moves too).
template <int D>
__global__ __launch_bounds__(BLOCK) void k_mcell_tangent(Geo g, Rows r, int limit, int fit, const double* dW, const double* dcs, const double* dx,
                                                         const double* dc, const double* dA, const double* dxf, double* dw, double* dg) {
    constexpr int V = D + 2, U = D + 1;
    using C = CellAt<D>;
    using FA = FaceAt<D>;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    const double* mine = g.cell + (size_t)i * CELL_ROW;
    double Wi[V], xi[D], ci[D], dxi[D], dci[D], dWi[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        Wi[v] = mine[C::W + v];
        dWi[v] = dW[(size_t)i * V + v];
    }
    load_vec<D>(mine + C::X, xi);
    load_vec<D>(mine + C::C, ci);
    load_vec<D>(dx + (size_t)i * D, dxi);
    load_vec<D>(dc + (size_t)i * D, dci);
    const int count = g.count[i], base = first_slot(g, i);
    // the point velocity w = v + c (fraction c_s / |c|), fraction from the largest face angle
    {
        double amax = 0, dmax[D];
        int fmax = -1, other = -1;
#pragma unroll
        for (int a = 0; a < D; ++a) dmax[a] = 0;
        for (int k = 0; k < count; ++k) {
            const int slot = base + k * SLICE;
            const int e = g.list[slot], f = e >= 0 ? e : ~e, j = g.other[slot];
            double xj[D], d[D];
            load_vec<D>(g.cell + (size_t)j * CELL_ROW + C::X, xj);
            separation<D>(g, xi, xj, d);
            const double ang = face_angle<D>(g.face[(size_t)f * FACE_ROW + FA::MAG]) / (0.5 * sqrt(dot<D>(d, d)));
            if (ang > amax) {
                amax = ang;
                fmax = f;
                other = j;
#pragma unroll
                for (int a = 0; a < D; ++a) dmax[a] = d[a];
            }
        }
        const double Fc = 2.25, shaping = 0.5;
        double fraction = 0, slope = 0;
        if (amax > 0.75 * Fc) {
            fraction = amax > Fc ? shaping : shaping * (amax - 0.75 * Fc) / (0.25 * Fc);
            slope = amax > Fc ? 0.0 : shaping / (0.25 * Fc);
        }
        double t[D];
#pragma unroll
        for (int a = 0; a < D; ++a) t[a] = dWi[1 + a];
        const double cd = sqrt(dot<D>(ci, ci)), cs = mine[C::CS];
        if (cd > 0 && fraction > 0) {
            const double kk = fraction * cs / cd;
            double dfraction = 0;
            if (slope != 0) {
                const double* A = g.face + (size_t)fmax * FACE_ROW + FA::A;
                double dd[D];
#pragma unroll
                for (int a = 0; a < D; ++a) dd[a] = dx[(size_t)other * D + a] - dxi[a];
                const double mag = sqrt(dot<D>(A, A)), dist = sqrt(dot<D>(dmax, dmax));
                const double dmag = dot<D>(A, dA + (size_t)fmax * D) / mag, ddist = dot<D>(dmax, dd) / dist;
                dfraction = slope * (face_angle_slope<D>(mag) * dmag / (0.5 * dist) - amax * ddist / dist);
            }
            const double dcd = dot<D>(ci, dci) / cd;
            const double dk = (dfraction * cs + fraction * dcs[i]) / cd - kk * dcd / cd;
#pragma unroll
            for (int a = 0; a < D; ++a) t[a] += dci[a] * kk + ci[a] * dk;
        }
#pragma unroll
        for (int a = 0; a < D; ++a) dw[(size_t)i * D + a] = t[a];
    }
    if (!fit) return;
    // the limited fit: the forward's values and pattern (the bound's tangent from dW), then the geometry's parts
    double gl[V * D], g0[V * D], dg0[V * D], Lc[U][U], M0[U], icpt[V];
    Pattern<D> pat;
    fit_cell<D, true>(g, i, dW, limit != 0, gl, g0, dg0, Lc, M0, pat, icpt);
    double dM[U][U], db[V][U];
#pragma unroll
    for (int p = 0; p < U; ++p)
#pragma unroll
        for (int q = 0; q < U; ++q) dM[p][q] = 0;
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
        for (int p = 0; p < U; ++p) db[v][p] = 0;
    for (int k = 0; k < count; ++k) {
        const int slot = base + k * SLICE;
        const int e = g.list[slot], f = e >= 0 ? e : ~e, j = g.other[slot];
        const double* row = g.cell + (size_t)j * CELL_ROW;
        double xj[D], d[D], X[U], dX[U];
        load_vec<D>(row + C::X, xj);
        separation<D>(g, xi, xj, d);
        const double* face = g.face + (size_t)f * FACE_ROW;
        const double w = face[FA::MAG];
        double Af[D], dAf[D];
        load_vec<D>(face + FA::A, Af);
        load_vec<D>(dA + (size_t)f * D, dAf);
        const double dwt = w > 0 ? dot<D>(Af, dAf) / w : 0.0;
        X[0] = 1;
        dX[0] = 0;
#pragma unroll
        for (int p = 1; p < U; ++p) {
            X[p] = d[p - 1] + row[C::C + p - 1] - ci[p - 1];
            dX[p] = (dx[(size_t)j * D + p - 1] - dxi[p - 1]) + dc[(size_t)j * D + p - 1] - dci[p - 1];
        }
#pragma unroll
        for (int p = 0; p < U; ++p)
#pragma unroll
            for (int q = 0; q < U; ++q) dM[p][q] += dwt * X[p] * X[q] + w * (dX[p] * X[q] + X[p] * dX[q]);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double delta = row[C::W + v] - Wi[v], ddelta = dW[(size_t)j * V + v] - dWi[v];
#pragma unroll
            for (int p = 0; p < U; ++p) db[v][p] += dwt * X[p] * delta + w * dX[p] * delta + w * X[p] * ddelta;
        }
    }
    double* out = dg + (size_t)i * V * D;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        double z[U], xs[U], rhs[U];
#pragma unroll
        for (int p = 0; p < U; ++p) {
            double mx = dM[p][0] * icpt[v];
#pragma unroll
            for (int a = 0; a < D; ++a) mx += dM[p][1 + a] * g0[v * D + a];
            rhs[p] = db[v][p] - mx;
        }
#pragma unroll
        for (int p = 0; p < U; ++p) {
            double sum = rhs[p];
#pragma unroll
            for (int m = 0; m < p; ++m) sum -= Lc[p][m] * z[m];
            z[p] = sum / Lc[p][p];
        }
#pragma unroll
        for (int p = U - 1; p >= 0; --p) {
            double sum = z[p];
#pragma unroll
            for (int m = p + 1; m < U; ++m) sum -= Lc[m][p] * xs[m];
            xs[p] = sum / Lc[p][p];
        }
        const double* gv = g0 + v * D;
        if (pat.last[v] < 0) {
#pragma unroll
            for (int a = 0; a < D; ++a) out[v * D + a] = xs[1 + a];
        } else if (pat.zero[v]) {
#pragma unroll
            for (int a = 0; a < D; ++a) out[v * D + a] = 0.0;
        } else {
            const int slot = base + pat.last[v] * SLICE;
            const int e = g.list[slot], f = e >= 0 ? e : ~e;
            double dor[D];
            offset_dot<D>(dxf + (size_t)f * D, e >= 0, dxi, dx + (size_t)g.other[slot] * D, dci, dor);
            const double* o = pat.offset[v];
            double dg0v[D];
#pragma unroll
            for (int a = 0; a < D; ++a) dg0v[a] = xs[1 + a];
            const double den = dot<D>(gv, o);
            const double sc = (pat.bound[v] - Wi[v]) / den;
            const double dsc = (pat.dbound[v] - dWi[v] - sc * (dot<D>(dg0v, o) + dot<D>(gv, dor))) / den;
#pragma unroll
            for (int a = 0; a < D; ++a) out[v * D + a] = dg0v[a] * sc + gv[a] * dsc;
        }
    }
}

// One moving pass's face tangent (finite_volume.hpp:2829-2870 moving_pass_tangent): dG of G = |A| F(side_o, side_j, A / |A|,
// face_velocity(d, x_f, w-hat_o, w-hat_j)); first pass (order0: the order-0 step's one pass, w-hat = the sweep's w, no
// gradient) sides W + g.r, second half_step(W, g, 2 dt) + g.(r + c - c^n + dt w-hat).
template <int D, bool Second>
__global__ __launch_bounds__(BLOCK) void k_mpass_face_tangent(Geo g, Rows r, Gas gas, int riemann_kind, double dt, double unit, int order0,
                                                              const double* dW, const double* dg, const double* dw, const double* dx,
                                                              const double* dc, const double* dA, const double* dxf, const double* dc0,
                                                              double* dG) {
    constexpr int V = D + 2, GR = grad_row<D>();
    using C = CellAt<D>;
    using FA = FaceAt<D>;
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= g.F) return;
    double* out = dG + (size_t)f * V;
    const double* face = g.face + (size_t)f * FACE_ROW;
    if (face[FA::MAG] == 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = 0;
        return;
    }
    const int2 pair = *reinterpret_cast<const int2*>(face + FA::PAIR);
    const int cells[2] = {pair.x, pair.y};
    double A[D], fc[D], xs[2][D], d[D], back[D], wh[2][D], dwh[2][D], dxs[2][D], wf[D], dwf[D], dd[D];
    load_vec<D>(face + FA::A, A);
    load_vec<D>(face + FA::FC, fc);
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
        const int i = cells[s2];
        load_vec<D>(g.cell + (size_t)i * CELL_ROW + C::X, xs[s2]);
        load_vec<D>(dx + (size_t)i * D, dxs[s2]);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            wh[s2][a] = order0 ? r.w[(size_t)i * D + a] : (double)r.steps[(size_t)i * D + a] * unit;
            dwh[s2][a] = dw[(size_t)i * D + a];
        }
    }
    separation<D>(g, xs[0], xs[1], d);     // owner -> neighbour
    separation<D>(g, xs[1], xs[0], back);  // neighbour -> owner
#pragma unroll
    for (int a = 0; a < D; ++a) dd[a] = dxs[1][a] - dxs[0][a];
    face_velocity<D>(d, fc, wh[0], wh[1], wf);
    face_velocity_tangent<D>(d, fc, wh[0], wh[1], dd, dxf + (size_t)f * D, dwh[0], dwh[1], dwf);
    double sides[2][V], dsides[2][V];
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2) {
        const int i = cells[s2];
        const double* cr = g.cell + (size_t)i * CELL_ROW;
        double W[V], dWi[V], gr[V * D], dgr[V * D], rr[D], dr[D];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            W[v] = cr[C::W + v];
            dWi[v] = dW[(size_t)i * V + v];
        }
#pragma unroll
        for (int k = 0; k < V * D; ++k) {
            gr[k] = order0 ? 0.0 : r.grad[(size_t)i * GR + k];
            dgr[k] = order0 ? 0.0 : dg[(size_t)i * V * D + k];
        }
        const double* dci = dc + (size_t)i * D;
        offset_dot<D>(dxf + (size_t)f * D, s2 == 0, dxs[s2], dxs[0], dci, dr);
#pragma unroll
        for (int a = 0; a < D; ++a) {
            rr[a] = (s2 == 0 ? fc[a] : fc[a] + back[a]) - cr[C::C + a];
            if constexpr (Second) {
                rr[a] = rr[a] + cr[C::C + a] - r.old_centroid[(size_t)i * D + a] + wh[s2][a] * (0.5 * (2 * dt));
                dr[a] = dr[a] + dci[a] - dc0[(size_t)i * D + a] + dwh[s2][a] * dt;
            }
        }
        if constexpr (Second) {
            half_step<D>(gas, W, gr, 2 * dt, sides[s2]);
            half_step_tangent<D>(gas, W, dWi, gr, dgr, 2 * dt, dsides[s2]);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                sides[s2][v] = W[v];
                dsides[s2][v] = dWi[v];
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) sides[s2][v] += dot<D>(gr + v * D, rr);
        if (sides[s2][0] > 0 && sides[s2][D + 1] > 0) {
#pragma unroll
            for (int v = 0; v < V; ++v) dsides[s2][v] += dot<D>(dgr + v * D, rr) + dot<D>(gr + v * D, dr);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                sides[s2][v] = W[v];
                dsides[s2][v] = dWi[v];
            }
        }
    }
    moving_face_tangent<D>(riemann_kind, gas, sides[0], sides[1], A, wf, dsides[0], dsides[1], dA + (size_t)f * D, dwf, out);
}

// dq -= scale sum_k sign_k dG_k per cell in list order (scale: dt / 2 per pass, dt at order 0).
template <int D>
__global__ __launch_bounds__(BLOCK) void k_mpass_gather_tangent(Geo g, double scale, const double* dG, double* dq) {
    constexpr int V = D + 2;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    double sum[V];
#pragma unroll
    for (int v = 0; v < V; ++v) sum[v] = 0;
    const int count = g.count[i], base = first_slot(g, i);
    for (int k = 0; k < count; ++k) {
        const int e = g.list[base + k * SLICE], f = e >= 0 ? e : ~e;
        const double sign = e >= 0 ? 1.0 : -1.0;
#pragma unroll
        for (int v = 0; v < V; ++v) sum[v] += sign * dG[(size_t)f * V + v];
    }
#pragma unroll
    for (int v = 0; v < V; ++v) dq[(size_t)i * V + v] -= scale * sum[v];
}

// out = x + dt y (n entries).
__global__ __launch_bounds__(BLOCK) void k_axpy(long long n, double dt, const double* x, const double* y, double* out) {
    const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
    if (k < n) out[k] = x[k] + dt * y[k];
}
}  // namespace

template <int D>
hipError_t launch_sweep_tangent(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, bool limit, const Derivs& w) {
    if (g.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_prim_tangent<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, gas, w.dq, w.dW);
    hipLaunchKernelGGL(k_fit_tangent<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, limit ? 1 : 0, w.dW, w.dgrad);
    return hipGetLastError();
}

// LDS the update tangent's launch reserves and leaves unused: 32 KiB holds a CU to 5 of its work-groups (20 wavefronts),
// so a face's [F | dF] row, gathered by its owner and by its neighbour, is found in L2 the second time more often (128^3:
// tangent 5.83 -> 5.70 ms; 40 KiB the same, 54 and 64 KiB slower). The adjoint's gather measured no better this way.
constexpr size_t UPDATE_TANGENT_LDS_RESERVE = 32768;

template <int D>
hipError_t launch_step_tangent(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, double dt, bool limit, const Derivs& w,
                               bool fitted) {
    if (g.N == 0) return hipSuccess;
    if (!fitted)
        if (hipError_t e = launch_sweep_tangent<D>(s, g, r, gas, limit, w)) return e;
    if (g.F > 0) hipLaunchKernelGGL(k_flux_tangent<D>, grid_of(g.F), dim3(BLOCK), 0, s, g, r, gas, riemann, dt, w.dW, w.dgrad, w.dflux);
    hipLaunchKernelGGL(k_update_tangent<D>, grid_of(g.N), dim3(BLOCK), UPDATE_TANGENT_LDS_RESERVE, s, g, r, dt, w.dflux, w.dq);
    return hipGetLastError();
}

template <int D>
hipError_t launch_step_adjoint(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, double dt, bool limit, const Derivs& w) {
    if (g.N == 0) return hipSuccess;
    if (g.F > 0) hipLaunchKernelGGL(k_face_adjoint<D>, grid_of(g.F), dim3(BLOCK), 0, s, g, r, gas, riemann, dt, w.dq, w.sidebar);
    hipLaunchKernelGGL(k_cell_adjoint<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, gas, dt, limit ? 1 : 0, w.sidebar, w.dW, w.bbar, w.boundbar,
                       w.extreme);
    hipLaunchKernelGGL(k_gather_adjoint<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, gas, w.dq, w.dW, w.bbar, w.boundbar, w.extreme, w.out);
    return hipGetLastError();
}

template <int D>
hipError_t launch_moving_pass_adjoint(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, double dt, double unit, bool second,
                                      const double* qbar, const MovingBars& b, bool order0) {
    if (g.N == 0) return hipSuccess;
    if (g.F > 0) {
        if (second)
            hipLaunchKernelGGL((k_mpass_face<D, true>), grid_of(g.F), dim3(BLOCK), 0, s, g, r, gas, riemann, dt, unit, qbar, b.row, b.Ab, b.Xfb, 0);
        else
            hipLaunchKernelGGL((k_mpass_face<D, false>), grid_of(g.F), dim3(BLOCK), 0, s, g, r, gas, riemann, dt, unit, qbar, b.row, b.Ab, b.Xfb,
                               order0 ? 1 : 0);
    }
    if (second)
        hipLaunchKernelGGL((k_mpass_gather<D, true>), grid_of(g.N), dim3(BLOCK), 0, s, g, dt, b.row, b.Wbar, b.gbar, b.whbar, b.c0bar, b.cbar, b.xbar);
    else
        hipLaunchKernelGGL((k_mpass_gather<D, false>), grid_of(g.N), dim3(BLOCK), 0, s, g, dt, b.row, b.Wbar, b.gbar, b.whbar, b.c0bar, b.cbar,
                           b.xbar);
    return hipGetLastError();
}

hipError_t launch_drift_adjoint(hipStream_t s, long long n, double dt, const double* xbar, double* whbar) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_drift_adjoint, grid_of(n), dim3(BLOCK), 0, s, n, dt, xbar, whbar);
    return hipGetLastError();
}

template <int D>
hipError_t launch_moving_cell_adjoint(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, bool limit, bool steps_fixed, const double* qbar,
                                      const MovingBars& b) {
    if (g.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mcell_adjoint<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, gas, limit ? 1 : 0, steps_fixed ? 1 : 0, b.Wbar, b.gbar, b.whbar,
                       b.cbar, b.xbar, b.beta, b.boundbar, b.extreme, b.entry, b.fmax_of, b.angle);
    hipLaunchKernelGGL(k_mgather_adjoint<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, gas, qbar, b.Wbar, b.beta, b.boundbar, b.extreme, b.entry,
                       b.fmax_of, b.angle, b.cbar, b.xbar, b.Vbar, b.out);
    if (g.F > 0) hipLaunchKernelGGL(k_mfaces_bar<D>, grid_of(g.F), dim3(BLOCK), 0, s, g, b.entry, b.fmax_of, b.angle, b.Ab, b.Xfb);
    return hipGetLastError();
}

template <int D>
hipError_t launch_moving_tangent_cells(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, bool limit, bool fit, const double* dq,
                                       const MovingDots& t) {
    hipLaunchKernelGGL(k_mprim_tangent<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, gas, dq, t.dV, t.dW, t.dcs);
    hipLaunchKernelGGL(k_mcell_tangent<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, limit ? 1 : 0, fit ? 1 : 0, t.dW, t.dcs, t.dx, t.dc, t.dA, t.dxf,
                       t.dw, t.dg);
    return hipGetLastError();
}

template <int D>
hipError_t launch_moving_pass_tangent(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, double dt, double unit, bool second,
                                      bool order0, const MovingDots& t, double* dq) {
    if (g.F > 0) {
        if (second)
            hipLaunchKernelGGL((k_mpass_face_tangent<D, true>), grid_of(g.F), dim3(BLOCK), 0, s, g, r, gas, riemann, dt, unit, 0, t.dW, t.dg, t.dw,
                               t.dx, t.dc, t.dA, t.dxf, t.dc0, t.dG);
        else
            hipLaunchKernelGGL((k_mpass_face_tangent<D, false>), grid_of(g.F), dim3(BLOCK), 0, s, g, r, gas, riemann, dt, unit, order0 ? 1 : 0, t.dW,
                               t.dg, t.dw, t.dx, t.dc, t.dA, t.dxf, t.dc0, t.dG);
    }
    hipLaunchKernelGGL(k_mpass_gather_tangent<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, order0 ? dt : 0.5 * dt, t.dG, dq);
    return hipGetLastError();
}

hipError_t launch_axpy(hipStream_t s, long long n, double dt, const double* x, const double* y, double* out) {
    if (n > 0) hipLaunchKernelGGL(k_axpy, grid_of(n), dim3(BLOCK), 0, s, n, dt, x, y, out);
    return hipGetLastError();
}

template hipError_t launch_moving_tangent_cells<2>(hipStream_t, const Geo&, const Rows&, const Gas&, bool, bool, const double*, const MovingDots&);
template hipError_t launch_moving_tangent_cells<3>(hipStream_t, const Geo&, const Rows&, const Gas&, bool, bool, const double*, const MovingDots&);
template hipError_t launch_moving_pass_tangent<2>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, double, bool, bool, const MovingDots&,
                                                  double*);
template hipError_t launch_moving_pass_tangent<3>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, double, bool, bool, const MovingDots&,
                                                  double*);
template hipError_t launch_moving_pass_adjoint<2>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, double, bool, const double*,
                                                  const MovingBars&, bool);
template hipError_t launch_moving_pass_adjoint<3>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, double, bool, const double*,
                                                  const MovingBars&, bool);
template hipError_t launch_moving_cell_adjoint<2>(hipStream_t, const Geo&, const Rows&, const Gas&, bool, bool, const double*, const MovingBars&);
template hipError_t launch_moving_cell_adjoint<3>(hipStream_t, const Geo&, const Rows&, const Gas&, bool, bool, const double*, const MovingBars&);
template hipError_t launch_sweep_tangent<2>(hipStream_t, const Geo&, const Rows&, const Gas&, bool, const Derivs&);
template hipError_t launch_sweep_tangent<3>(hipStream_t, const Geo&, const Rows&, const Gas&, bool, const Derivs&);
template hipError_t launch_step_tangent<2>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, bool, const Derivs&, bool);
template hipError_t launch_step_tangent<3>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, bool, const Derivs&, bool);
template hipError_t launch_step_adjoint<2>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, bool, const Derivs&);
template hipError_t launch_step_adjoint<3>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, bool, const Derivs&);

}  // namespace luna
