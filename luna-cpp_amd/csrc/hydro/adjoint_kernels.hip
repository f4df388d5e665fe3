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
                         double* M0, Pattern<D>& pat) {
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

// One face's two sides as the static flux computes them (half step + g.r, positivity fallback), with their fallback
// flags, offsets, the unit normal and |A|.
template <int D>
__device__ __forceinline__ void face_sides(const Geo& g, const Rows& r, const Gas& gas, int f, double dt, double* Ls, double* Rs,
                                           bool* fallback, double* ro, double* rj, double* n, double& magnitude) {
    constexpr int V = D + 2;
    using C = CellAt<D>;
    using FA = FaceAt<D>;
    const double* face = g.face + (size_t)f * FACE_ROW;
    const int2 pair = *reinterpret_cast<const int2*>(face + FA::PAIR);
    magnitude = face[FA::MAG];
    const double inv = 1.0 / magnitude;
#pragma unroll
    for (int k = 0; k < D; ++k) n[k] = face[FA::A + k] * inv;
    const double* ro_row = g.cell + (size_t)pair.x * CELL_ROW;
    const double* rj_row = g.cell + (size_t)pair.y * CELL_ROW;
    double xo[D], xj[D], d[D];
    load_vec<D>(ro_row + C::X, xo);
    load_vec<D>(rj_row + C::X, xj);
    separation<D>(g, xj, xo, d);  // the owner seen from the neighbour
    const double half = 0.5 * dt, wf = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        ro[k] = face[FA::FC + k] - ro_row[C::C + k];
        rj[k] = face[FA::FC + k] + d[k] - rj_row[C::C + k];
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
    const int i = blockIdx.x * BLOCK + threadIdx.x;
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
__global__ __launch_bounds__(BLOCK) void k_fit_tangent(Geo g, Rows r, int limit, const double* dW, double* dgrad) {
    constexpr int V = D + 2, U = D + 1, GR = grad_row<D>();
    using C = CellAt<D>;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
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
    double* Fv = r.flux + (size_t)f * V;
    double* dF = dflux + (size_t)f * V;
    const double* face = g.face + (size_t)f * FACE_ROW;
    if (face[FaceAt<D>::MAG] == 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) Fv[v] = dF[v] = 0;
        return;
    }
    double Ls[V], Rs[V], ro[D], rj[D], n[D], magnitude, dLs[V], dRs[V];
    bool fallback[2];
    face_sides<D>(g, r, gas, f, dt, Ls, Rs, fallback, ro, rj, n, magnitude);
    const int2 pair = *reinterpret_cast<const int2*>(face + FaceAt<D>::PAIR);
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
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    double sum[V], dsum[V];
#pragma unroll
    for (int v = 0; v < V; ++v) sum[v] = dsum[v] = 0;
    const int count = g.count[i], base = first_slot(g, i);
    for (int k = 0; k < count; ++k) {
        const int e = g.list[base + k * SLICE], f = e >= 0 ? e : ~e;
        const double sign = e >= 0 ? 1.0 : -1.0;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            sum[v] += sign * r.flux[(size_t)f * V + v];
            dsum[v] += sign * dflux[(size_t)f * V + v];
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
    const double* face = g.face + (size_t)f * FACE_ROW;
    if (face[FaceAt<D>::MAG] != 0) {
        double Ls[V], Rs[V], ro[D], rj[D], n[D], magnitude, Fbar[V];
        bool fallback[2];
        face_sides<D>(g, r, gas, f, dt, Ls, Rs, fallback, ro, rj, n, magnitude);
        const int2 pair = *reinterpret_cast<const int2*>(face + FaceAt<D>::PAIR);
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
__global__ __launch_bounds__(BLOCK) void k_cell_adjoint(Geo g, Rows r, Gas gas, double dt, int limit, const double* sidebar, double* Wbar,
                                                        double* bbar, double* boundbar, int* extreme) {
    constexpr int V = D + 2, U = D + 1, GR = grad_row<D>();
    using C = CellAt<D>;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
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
    const int count = g.count[i], base = first_slot(g, i);
    for (int k = 0; k < count; ++k) {
        const int e = g.list[base + k * SLICE], f = e >= 0 ? e : ~e;
        if (g.face[(size_t)f * FACE_ROW + FaceAt<D>::MAG] == 0) continue;
        double Ls[V], Rs[V], ro[D], rj[D], n[D], magnitude;
        bool fallback[2];
        face_sides<D>(g, r, gas, f, dt, Ls, Rs, fallback, ro, rj, n, magnitude);
        const int s = e >= 0 ? 0 : 1;
        const double* side = sidebar + (size_t)f * 2 * V + s * V;
        double sb[V];
#pragma unroll
        for (int v = 0; v < V; ++v) sb[v] = side[v];
        if (fallback[s]) {
#pragma unroll
            for (int v = 0; v < V; ++v) Wb[v] += sb[v];
            continue;
        }
        const double* rr = s == 0 ? ro : rj;
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
    const int j = blockIdx.x * BLOCK + threadIdx.x;
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
        const double w = g.face[(size_t)f * FACE_ROW + FaceAt<D>::MAG];
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

}  // namespace

template <int D>
hipError_t launch_step_tangent(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, double dt, bool limit, const Derivs& w) {
    if (g.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_prim_tangent<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, gas, w.dq, w.dW);
    hipLaunchKernelGGL(k_fit_tangent<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, limit ? 1 : 0, w.dW, w.dgrad);
    if (g.F > 0) hipLaunchKernelGGL(k_flux_tangent<D>, grid_of(g.F), dim3(BLOCK), 0, s, g, r, gas, riemann, dt, w.dW, w.dgrad, w.dflux);
    hipLaunchKernelGGL(k_update_tangent<D>, grid_of(g.N), dim3(BLOCK), 0, s, g, r, dt, w.dflux, w.dq);
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

template hipError_t launch_step_tangent<2>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, bool, const Derivs&);
template hipError_t launch_step_tangent<3>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, bool, const Derivs&);
template hipError_t launch_step_adjoint<2>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, bool, const Derivs&);
template hipError_t launch_step_adjoint<3>(hipStream_t, const Geo&, const Rows&, const Gas&, int, double, bool, const Derivs&);

}  // namespace luna
