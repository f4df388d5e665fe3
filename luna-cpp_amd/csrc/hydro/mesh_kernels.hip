// The hydro step on a Voronoi mesh (D = 2, 3), hand-written for CDNA4 (gfx950): one work-item per cell for the sweep
// and the update, one per face for the fluxes. FP64, -ffp-contract=off; each cell sums over its face list in list order,
// so every run gives the same bits, and those of the CPU restatement of the reference's arithmetic.
#include "kernels.hpp"

namespace luna {
namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// The block's smallest value into *out (as the bits of a non-negative double: they order as unsigned integers; NaN
// terms never win, as Kokkos::Min's `<` and fmin drop them).
__device__ __forceinline__ void block_min(double v, unsigned long long* out) {
    __shared__ double part[BLOCK / 64];
    v = wave_min(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = part[0];
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) m = fmin(m, part[k]);
        if (m < INFINITY) atomicMin(out, (unsigned long long)__double_as_longlong(m));
    }
}
__device__ __forceinline__ void block_count(int v, unsigned long long* out) {
    __shared__ int part[BLOCK / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += part[k];
        if (s) atomicAdd(out, (unsigned long long)s);
    }
}

// finite_volume.hpp:279-282 Geometry::separation(i, j): generator j seen from generator i, nearest image.
template <int D>
__device__ __forceinline__ void separation(const Geo& g, int i, int j, double* d) {
#pragma unroll
    for (int a = 0; a < D; ++a) d[a] = wrap(g.positions[(size_t)j * D + a] - g.positions[(size_t)i * D + a], g.side[a]);
}

// finite_volume.hpp:322-326 Geometry::face_offset(f, i); owned: whether i is the owner of f (the list's owner bit).
template <int D>
__device__ __forceinline__ void face_offset(const Geo& g, int f, int i, bool owned, double* r) {
    double rel[D];
    if (owned) {
#pragma unroll
        for (int a = 0; a < D; ++a) rel[a] = g.fc[(size_t)f * D + a];
    } else {
        double s[D];
        separation<D>(g, i, g.owner[f], s);
#pragma unroll
        for (int a = 0; a < D; ++a) rel[a] = g.fc[(size_t)f * D + a] + s[a];
    }
#pragma unroll
    for (int a = 0; a < D; ++a) r[a] = rel[a] - g.centroid[(size_t)i * D + a];
}

// finite_volume.hpp:403-415 primitives()
template <int D>
__global__ __launch_bounds__(BLOCK) void k_primitives(Geo g, Rows r, Gas gas) {
    constexpr int V = D + 2, PW = V + 2;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    const double inv = 1.0 / g.volume[i];
    double U[V], W[V];
#pragma unroll
    for (int v = 0; v < V; ++v) U[v] = r.q[(size_t)i * V + v] * inv;
    primitive<D>(gas, U, W);
    double* row = r.prims + (size_t)i * PW;
#pragma unroll
    for (int v = 0; v < V; ++v) row[v] = W[v];
    row[V] = sound_speed<D>(gas, W);
    row[V + 1] = g.radius[i];
}

// |A| of every face (the magnitudes the reference caches, finite_volume.hpp:1582-1585).
template <int D>
__global__ __launch_bounds__(BLOCK) void k_magnitudes(Geo g, double* magnitude) {
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    if (f >= g.F) return;
    const double* a = g.area + (size_t)f * D;
    magnitude[f] = sqrt(dot<D>(a, a));
}

// The sweep (finite_volume.hpp:463-569 Sweep::operator(), double policy) with the time step's terms (:1040-1104) fused:
// each cell takes its own term and, for every face of its list, the face term with its own radius. The pair term vsig and
// |d| have the same bits from either side, so the minimum is the reference's (the face kernel's terms are the same set).
// The first block also resets the other dt slot and the step's fallback counters.
template <int D, bool Moving, bool Grad, bool Limit, int Rule>
__global__ __launch_bounds__(BLOCK) void k_sweep(Geo g, Rows r, Gas gas, double courant, Counters c) {
    constexpr int V = D + 2, U = D + 1, PW = V + 2;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *c.dt_next = INF_BITS;
        c.fallbacks[0] = 0;
        c.fallbacks[1] = 0;
    }
    double smallest = INFINITY;
    if (i < g.N) {
        const double* Wrow = r.prims + (size_t)i * PW;
        double Wi[V];
#pragma unroll
        for (int v = 0; v < V; ++v) Wi[v] = Wrow[v];
        const double ci = Wrow[V], ri = Wrow[V + 1];
        double centroid[D];
#pragma unroll
        for (int a = 0; a < D; ++a) centroid[a] = g.centroid[(size_t)i * D + a];
        double surface = 0, smallest_face = INFINITY, max_angle = 0;
        double lo[V], hi[V], M[U][U], b[V][U];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            lo[v] = 1.7976931348623157e308;
            hi[v] = -1.7976931348623157e308;
#pragma unroll
            for (int p = 0; p < U; ++p) b[v][p] = 0;
        }
#pragma unroll
        for (int p = 0; p < U; ++p)
#pragma unroll
            for (int s = 0; s < U; ++s) M[p][s] = 0;
        const int first = g.start[i], end = g.start[i + 1];
        for (int k = first; k < end; ++k) {
            const int e = g.list[k], f = e >= 0 ? e : ~e;
            const int j = e >= 0 ? g.neighbour[f] : g.owner[f];
            const double* Wrj = r.prims + (size_t)j * PW;
            double Wj[V];
#pragma unroll
            for (int v = 0; v < V; ++v) Wj[v] = Wrj[v];
            double d[D];
            separation<D>(g, i, j, d);
            const double magnitude = g.magnitude[f];
            surface += magnitude;
            smallest_face = fmin(smallest_face, magnitude);
            const double dist = sqrt(dot<D>(d, d));
            if constexpr (Moving) max_angle = fmax(max_angle, face_angle<D>(magnitude) / (0.5 * dist));
            if constexpr (Grad) {
                double X[U];
                X[0] = 1;
#pragma unroll
                for (int p = 1; p < U; ++p) X[p] = d[p - 1] + g.centroid[(size_t)j * D + p - 1] - centroid[p - 1];
                const double w = magnitude;
#pragma unroll
                for (int p = 0; p < U; ++p)
#pragma unroll
                    for (int s = p; s < U; ++s) M[p][s] += w * X[p] * X[s];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const double dW = Wj[v] - Wi[v];
#pragma unroll
                    for (int p = 0; p < U; ++p) b[v][p] += w * X[p] * dW;
                    lo[v] = fmin(lo[v], Wj[v]);
                    hi[v] = fmax(hi[v], Wj[v]);
                }
            }
            if constexpr (Rule == 0) {  // FaceStep (:1079-1089) from this side
                double dv[D];
#pragma unroll
                for (int a = 0; a < D; ++a) dv[a] = Wj[1 + a] - Wi[1 + a];
                const double vsig = ci + Wrj[V] - dot<D>(dv, d) / dist;
                if (vsig > 0) smallest = fmin(smallest, courant * ((dist + ri) / vsig));
            }
        }
        const double tiny = 1e-6 * surface;
        const bool any_tiny = smallest_face < tiny;
        if constexpr (Grad) {
            if (any_tiny) {  // neighbour_range (:572-588)
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    lo[v] = 1.7976931348623157e308;
                    hi[v] = -1.7976931348623157e308;
                }
                for (int k = first; k < end; ++k) {
                    const int e = g.list[k], f = e >= 0 ? e : ~e;
                    if (g.magnitude[f] < tiny) continue;
                    const int j = e >= 0 ? g.neighbour[f] : g.owner[f];
                    const double* Wrj = r.prims + (size_t)j * PW;
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        lo[v] = fmin(lo[v], Wrj[v]);
                        hi[v] = fmax(hi[v], Wrj[v]);
                    }
                }
            }
        }
        double grad[V][D];
        if constexpr (Grad) {  // Sweep::solve (:590-615): Cholesky of M, then M x = b per variable
            double L[U][U];
#pragma unroll
            for (int p = 0; p < U; ++p)
#pragma unroll
                for (int s = 0; s < U; ++s) L[p][s] = 0;
#pragma unroll
            for (int p = 0; p < U; ++p)
#pragma unroll
                for (int s = 0; s <= p; ++s) {
                    double sum = M[s][p];
#pragma unroll
                    for (int m = 0; m < s; ++m) sum -= L[p][m] * L[s][m];
                    L[p][s] = p == s ? sqrt(sum) : sum / L[s][s];
                }
#pragma unroll
            for (int v = 0; v < V; ++v) {
                double z[U], x[U];
#pragma unroll
                for (int p = 0; p < U; ++p) {
                    double sum = b[v][p];
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
                for (int a = 0; a < D; ++a) grad[v][a] = x[1 + a];
            }
        }
        double wi[D];
#pragma unroll
        for (int a = 0; a < D; ++a) wi[a] = Wi[1 + a];
        if constexpr (Moving) {  // the regularised point velocity (:526-537; constants :74)
            const double shaping = 0.5, factor = 2.25;
            double fraction = 0;
            if (max_angle > 0.75 * factor)
                fraction = max_angle > factor ? shaping : shaping * (max_angle - 0.75 * factor) / (0.25 * factor);
            const double cd = sqrt(dot<D>(centroid, centroid));
            if (cd > 0 && fraction > 0) {
                const double s = fraction * ci / cd;
#pragma unroll
                for (int a = 0; a < D; ++a) wi[a] = wi[a] + centroid[a] * s;
            }
#pragma unroll
            for (int a = 0; a < D; ++a) r.w[(size_t)i * D + a] = wi[a];
        }
        if constexpr (Rule == 0) {  // CellStep (:1046-1051)
            double rel[D];
#pragma unroll
            for (int a = 0; a < D; ++a) rel[a] = Moving ? Wi[1 + a] - wi[a] : Wi[1 + a];
            const double step = ri / (ci + sqrt(dot<D>(rel, rel)));
            smallest = fmin(smallest, courant * step);
        } else {  // LunaStep (:1061-1070), 2D
            const double dx = ri * sqrt(M_PI);
            double rate = 0;
#pragma unroll
            for (int a = 0; a < D; ++a) {
                double relative = Wi[1 + a];
                if constexpr (Moving) relative -= wi[a];
                rate += (fabs(relative) + ci) / dx;
            }
            smallest = fmin(smallest, courant / rate);
        }
        if constexpr (Grad) {
            if constexpr (Limit) {  // the face-centroid limiter, face by face in list order (:538-565)
                for (int k = first; k < end; ++k) {
                    const int e = g.list[k], f = e >= 0 ? e : ~e;
                    if (any_tiny && g.magnitude[f] < tiny) continue;
                    double offset[D];
                    face_offset<D>(g, f, i, e >= 0, offset);
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        const double dp = dot<D>(grad[v], offset);
                        if (dp > 0 && Wi[v] + dp > hi[v]) {
                            const double s = hi[v] > Wi[v] ? (hi[v] - Wi[v]) / dp : 0.0;
#pragma unroll
                            for (int a = 0; a < D; ++a) grad[v][a] = grad[v][a] * s;
                        } else if (dp < 0 && Wi[v] + dp < lo[v]) {
                            const double s = lo[v] < Wi[v] ? (lo[v] - Wi[v]) / dp : 0.0;
#pragma unroll
                            for (int a = 0; a < D; ++a) grad[v][a] = grad[v][a] * s;
                        }
                    }
                }
            }
            double* out = r.grad + (size_t)i * V * D;
#pragma unroll
            for (int v = 0; v < V; ++v)
#pragma unroll
                for (int a = 0; a < D; ++a) out[v * D + a] = grad[v][a];
        }
    }
    block_min(smallest, c.dt);
}

// finite_volume.hpp:1306-1329 FaceFlux::side() at order 1: half step (over step_dt) + g.r, positivity fallback.
template <int D>
__device__ __forceinline__ void face_side(const Rows& r, const Gas& gas, int i, const double* off, double step_dt, int order,
                                          double* Wf, int& fallbacks) {
    constexpr int V = D + 2, PW = V + 2;
    const double* W = r.prims + (size_t)i * PW;
    double Wc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) Wc[v] = W[v];
    if (order == 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) Wf[v] = Wc[v];
        return;
    }
    double gr[V * D];
#pragma unroll
    for (int k = 0; k < V * D; ++k) gr[k] = r.grad[(size_t)i * V * D + k];
    half_step<D>(gas, Wc, gr, step_dt, Wf);
#pragma unroll
    for (int v = 0; v < V; ++v) Wf[v] += dot<D>(gr + v * D, off);
    if (Wf[0] > 0 && Wf[D + 1] > 0) return;
    ++fallbacks;
#pragma unroll
    for (int v = 0; v < V; ++v) Wf[v] = Wc[v];
}

// finite_volume.hpp:1451-1465 MovingFlux<false>::pass_side(): W + g.r on the mesh of t^n.
template <int D>
__device__ __forceinline__ void pass_side_first(const Rows& r, int i, const double* off, double* Wf, int& fallbacks) {
    constexpr int V = D + 2, PW = V + 2;
    const double* W = r.prims + (size_t)i * PW;
    double Wc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) Wc[v] = W[v];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        double gv[D];
#pragma unroll
        for (int a = 0; a < D; ++a) gv[a] = r.grad[(size_t)i * V * D + v * D + a];
        Wf[v] = Wc[v] + dot<D>(gv, off);
    }
    if (Wf[0] > 0 && Wf[D + 1] > 0) return;
    ++fallbacks;
#pragma unroll
    for (int v = 0; v < V; ++v) Wf[v] = Wc[v];
}

// finite_volume.hpp:1380-1397 FaceFlux::face_velocity()
template <int D>
__device__ __forceinline__ void face_velocity(const Geo& g, int f, int o, int j, const double* wL, const double* wR, double* wf) {
    double d[D], c[D], dw[D];
    separation<D>(g, o, j, d);
    const double dist = sqrt(dot<D>(d, d));
    const double inv_dist = 1.0 / dist;
#pragma unroll
    for (int a = 0; a < D; ++a) c[a] = g.fc[(size_t)f * D + a] - d[a] * 0.5;
#pragma unroll
    for (int a = 0; a < D; ++a) dw[a] = wL[a] - wR[a];
    double facv = dot<D>(c, dw) * inv_dist;
    const double limit = 0.9 * dist, c2 = dot<D>(c, c);
    if (c2 > limit * limit) facv *= limit / sqrt(c2);
#pragma unroll
    for (int a = 0; a < D; ++a) wf[a] = (wL[a] + wR[a]) * 0.5 + d[a] * (facv * inv_dist);
}

// One face's flux row: FaceFlux::operator() (:1415-1434) statically, MovingFlux::operator() (:1466-1490) on a moving step.
template <int D, int Pass, int Riemann>
__global__ __launch_bounds__(BLOCK) void k_flux(Geo g, Rows r, Gas gas, int order, double dt, double unit, unsigned long long* fallbacks) {
    constexpr int V = D + 2;
    const int f = blockIdx.x * BLOCK + threadIdx.x;
    int count = 0;
    if (f < g.F) {
        const int o = g.owner[f], j = g.neighbour[f];
        double a[D];
#pragma unroll
        for (int k = 0; k < D; ++k) a[k] = g.area[(size_t)f * D + k];
        const double magnitude = sqrt(dot<D>(a, a));
        if constexpr (Pass == 1) g.magnitude_out[f] = magnitude;  // for the next sweep on this mesh
        double F[V];
#pragma unroll
        for (int v = 0; v < V; ++v) F[v] = 0;
        if (magnitude != 0) {
            double L[V], R[V], n[D], wf[D], ro[D], rj[D];
            const double inv = 1.0 / magnitude;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                n[k] = a[k] * inv;
                wf[k] = 0;
            }
            face_offset<D>(g, f, o, true, ro);
            face_offset<D>(g, f, j, false, rj);
            if constexpr (Pass < 0) {
                const double half = 0.5 * dt;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    ro[k] = ro[k] + wf[k] * half;
                    rj[k] = rj[k] + wf[k] * half;
                }
                face_side<D>(r, gas, o, ro, dt, order, L, count);
                face_side<D>(r, gas, j, rj, dt, order, R, count);
            } else {
                double wo[D], wj[D];
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    wo[k] = (double)r.steps[(size_t)o * D + k] * unit;
                    wj[k] = (double)r.steps[(size_t)j * D + k] * unit;
                }
                face_velocity<D>(g, f, o, j, wo, wj, wf);
                if constexpr (Pass == 0) {
                    pass_side_first<D>(r, o, ro, L, count);
                    pass_side_first<D>(r, j, rj, R, count);
                } else {
                    // side(i, r + c^n+1 - c^n + w dt) with the FaceFlux dt = 2 dt (:1452-1454)
                    const double step_dt = 2 * dt;
                    double r2[D];
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        const double moved = wo[k] * (0.5 * step_dt);
                        r2[k] = ro[k] + g.centroid[(size_t)o * D + k] - r.old_centroid[(size_t)o * D + k] + moved;
                    }
                    face_side<D>(r, gas, o, r2, step_dt, 1, L, count);
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        const double moved = wj[k] * (0.5 * step_dt);
                        r2[k] = rj[k] + g.centroid[(size_t)j * D + k] - r.old_centroid[(size_t)j * D + k] + moved;
                    }
                    face_side<D>(r, gas, j, r2, step_dt, 1, R, count);
                }
            }
            if constexpr (Riemann == 1)
                hll<D>(gas, L, R, n, wf, F);
            else
                hllc<D>(gas, L, R, n, wf, F);
#pragma unroll
            for (int v = 0; v < V; ++v) F[v] *= magnitude;
        }
#pragma unroll
        for (int v = 0; v < V; ++v) r.flux[(size_t)f * V + v] = F[v];
    }
    block_count(count, fallbacks);
}

// finite_volume.hpp:1587-1603 update_body: Q -= dt sum_k sign_k F in list order; NextPrims: the next primitive rows
// (a moving step's second pass, and here every static step, which the next time_step() would compute the same way).
template <int D, bool NextPrims>
__global__ __launch_bounds__(BLOCK) void k_update(Geo g, Rows r, Gas gas, double dt) {
    constexpr int V = D + 2, PW = V + 2;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    double sum[V];
#pragma unroll
    for (int v = 0; v < V; ++v) sum[v] = 0;
    const int first = g.start[i], end = g.start[i + 1];
    for (int k = first; k < end; ++k) {
        const int e = g.list[k], f = e >= 0 ? e : ~e;
        const double sign = e >= 0 ? 1.0 : -1.0;
#pragma unroll
        for (int v = 0; v < V; ++v) sum[v] += sign * r.flux[(size_t)f * V + v];
    }
    double next[V];
#pragma unroll
    for (int v = 0; v < V; ++v) r.q[(size_t)i * V + v] = next[v] = r.q[(size_t)i * V + v] - dt * sum[v];
    if constexpr (NextPrims) {
        const double inv = 1.0 / g.volume[i];
        double U[V], W[V];
#pragma unroll
        for (int v = 0; v < V; ++v) U[v] = next[v] * inv;
        primitive<D>(gas, U, W);
        double* row = r.prims + (size_t)i * PW;
#pragma unroll
        for (int v = 0; v < V; ++v) row[v] = W[v];
        row[V] = sound_speed<D>(gas, W);
        row[V + 1] = g.radius[i];
    }
}

__device__ __forceinline__ long long snap_key(double x, const Grid& grid, int a) {
    long long k = (long long)round(x * grid.per_step) % grid.period[a];
    if (k < 0) k += grid.period[a];
    return k;
}

// finite_volume.hpp:1506-1514 keep_centroids: c^n kept; x to its grid point; the move to x + dt w in whole grid steps.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_keep_centroids(Geo g, Rows r, Grid grid, double dt) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= g.N) return;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const size_t at = (size_t)i * D + a;
        r.old_centroid[at] = g.centroid[at];
        const double x = r.x[at];
        const long long k0 = snap_key(x, grid, a), k1 = snap_key(x + dt * r.w[at], grid, a);
        r.x[at] = (double)k0 * grid.step;
        long long d = k1 - k0;
        const long long p = grid.period[a];
        if (d > p / 2) d -= p;
        if (d < -(p / 2)) d += p;
        r.steps[at] = (int)d;
    }
}

// finite_volume.hpp:1219-1231 drift on the grid: x to its grid point plus its whole steps, wrapped.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_drift(int N, Rows r, Grid grid) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= N) return;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const size_t at = (size_t)i * D + a;
        const long long k = (long long)round(r.x[at] * grid.per_step) + r.steps[at];
        const long long p = grid.period[a];
        r.x[at] = (double)(((k % p) + p) % p) * grid.step;
    }
}

inline dim3 blocks(long long n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

}  // namespace

template <int D>
hipError_t launch_primitives(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas) {
    if (g.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_primitives<D>, blocks(g.N), dim3(BLOCK), 0, s, g, r, gas);
    return hipGetLastError();
}

template <int D>
hipError_t launch_magnitudes(hipStream_t s, const Geo& g, double* magnitude) {
    if (g.F == 0) return hipSuccess;
    hipLaunchKernelGGL(k_magnitudes<D>, blocks(g.F), dim3(BLOCK), 0, s, g, magnitude);
    return hipGetLastError();
}

template <int D, bool Moving, bool Grad, bool Limit>
static hipError_t sweep_rule(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double courant, int rule, const Counters& c) {
    const dim3 grid = blocks(g.N > 0 ? g.N : 1);
    if (rule == 1)
        hipLaunchKernelGGL((k_sweep<D, Moving, Grad, Limit, 1>), grid, dim3(BLOCK), 0, s, g, r, gas, courant, c);
    else
        hipLaunchKernelGGL((k_sweep<D, Moving, Grad, Limit, 0>), grid, dim3(BLOCK), 0, s, g, r, gas, courant, c);
    return hipGetLastError();
}

template <int D>
hipError_t launch_sweep(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double courant, bool moving, bool grad, bool limit,
                        int rule, const Counters& c) {
    if (moving) {
        if (grad) return limit ? sweep_rule<D, true, true, true>(s, g, r, gas, courant, rule, c)
                               : sweep_rule<D, true, true, false>(s, g, r, gas, courant, rule, c);
        return sweep_rule<D, true, false, false>(s, g, r, gas, courant, rule, c);
    }
    if (grad) return limit ? sweep_rule<D, false, true, true>(s, g, r, gas, courant, rule, c)
                           : sweep_rule<D, false, true, false>(s, g, r, gas, courant, rule, c);
    return sweep_rule<D, false, false, false>(s, g, r, gas, courant, rule, c);
}

template <int D, int Pass>
static hipError_t flux_pass(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, double dt, double unit,
                            unsigned long long* fallbacks) {
    const dim3 grid = blocks(g.F > 0 ? g.F : 1);
    if (riemann == 1)
        hipLaunchKernelGGL((k_flux<D, Pass, 1>), grid, dim3(BLOCK), 0, s, g, r, gas, order, dt, unit, fallbacks);
    else
        hipLaunchKernelGGL((k_flux<D, Pass, 0>), grid, dim3(BLOCK), 0, s, g, r, gas, order, dt, unit, fallbacks);
    return hipGetLastError();
}

template <int D>
hipError_t launch_flux(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, int pass, double dt, double unit,
                       unsigned long long* fallbacks) {
    if (pass < 0) return flux_pass<D, -1>(s, g, r, gas, riemann, order, dt, unit, fallbacks);
    if (pass == 0) return flux_pass<D, 0>(s, g, r, gas, riemann, order, dt, unit, fallbacks);
    return flux_pass<D, 1>(s, g, r, gas, riemann, order, dt, unit, fallbacks);
}

template <int D>
hipError_t launch_update(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double dt, bool next_prims) {
    if (g.N == 0) return hipSuccess;
    if (next_prims)
        hipLaunchKernelGGL((k_update<D, true>), blocks(g.N), dim3(BLOCK), 0, s, g, r, gas, dt);
    else
        hipLaunchKernelGGL((k_update<D, false>), blocks(g.N), dim3(BLOCK), 0, s, g, r, gas, dt);
    return hipGetLastError();
}

template <int D>
hipError_t launch_keep_centroids(hipStream_t s, const Geo& g, const Rows& r, const Grid& grid, double dt) {
    if (g.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_keep_centroids<D>, blocks(g.N), dim3(BLOCK), 0, s, g, r, grid, dt);
    return hipGetLastError();
}

template <int D>
hipError_t launch_drift(hipStream_t s, int N, const Rows& r, const Grid& grid) {
    if (N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_drift<D>, blocks(N), dim3(BLOCK), 0, s, N, r, grid);
    return hipGetLastError();
}

#define LUNA_INSTANTIATE(D)                                                                                                          \
    template hipError_t launch_primitives<D>(hipStream_t, const Geo&, const Rows&, const Gas&);                                     \
    template hipError_t launch_magnitudes<D>(hipStream_t, const Geo&, double*);                                                     \
    template hipError_t launch_sweep<D>(hipStream_t, const Geo&, const Rows&, const Gas&, double, bool, bool, bool, int,             \
                                        const Counters&);                                                                           \
    template hipError_t launch_flux<D>(hipStream_t, const Geo&, const Rows&, const Gas&, int, int, int, double, double,              \
                                       unsigned long long*);                                                                        \
    template hipError_t launch_update<D>(hipStream_t, const Geo&, const Rows&, const Gas&, double, bool);                           \
    template hipError_t launch_keep_centroids<D>(hipStream_t, const Geo&, const Rows&, const Grid&, double);                         \
    template hipError_t launch_drift<D>(hipStream_t, int, const Rows&, const Grid&);

LUNA_INSTANTIATE(2)
LUNA_INSTANTIATE(3)

}  // namespace luna
