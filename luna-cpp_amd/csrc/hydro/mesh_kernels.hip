// The hydro step on a Voronoi mesh (D = 2, 3), hand-written for CDNA4 (gfx950): one work-item per cell for the sweep
// and the update, one per face for the fluxes. FP64, -ffp-contract=off; each cell sums over its face list in list order,
// so every run gives the same bits, and those of the CPU restatement of the reference's arithmetic.
//
// Memory (kernels.hpp): a neighbour is one 128 B cell row (position, centroid, state) and a face one 64 B record; the
// lists are stored in 64-cell slices, so a wavefront reads the t-th entries of its cells as one contiguous run and the
// entries of the next group are in flight while the current group is summed; work-groups are mapped so that each XCD
// takes a contiguous run of cells (neighbours in Morton order share its L2).
#include <hip/hip_fp16.h>

#include <type_traits>

#include "kernels.hpp"

// Tuning knobs (compile-time; the defaults are the measured best): entries of the sweep's first walk and of its limiter
// walk loaded together, flux rows of the update loaded together, and the occupancy floor (waves per SIMD) of the sweep,
// flux and LaneSweep kernels.
#ifndef LUNA_SWEEP_UNROLL
#define LUNA_SWEEP_UNROLL 1
#endif
#ifndef LUNA_LIMIT_UNROLL
#define LUNA_LIMIT_UNROLL 4
#endif
#ifndef LUNA_UPDATE_UNROLL
#define LUNA_UPDATE_UNROLL 4
#endif
#ifndef LUNA_SWEEP_WAVES
#define LUNA_SWEEP_WAVES 3
#endif
#ifndef LUNA_FLUX_WAVES
#define LUNA_FLUX_WAVES 1
#endif
#ifndef LUNA_LANE_WAVES  // LaneSweep's occupancy floor: 2 and 8 lanes measured best at 2 waves (no spill), 4 lanes at 3
#define LUNA_LANE_WAVES(L) ((L) == 4 ? 3 : 2)
#endif
// 1: every sweep takes the neighbours' range (min, max of W) in a loop of its own after the fit, re-reading the
// neighbours' W (cache-resident), instead of carrying it through the first walk (20 registers fewer there); 0: only the
// moving FP64 sweep does (the static one fits 3 waves per SIMD without it; the moving one spills with it carried). The
// reduced-precision sweeps load two entries together and carry the range (their float sums are half the registers).
#ifndef LUNA_RANGE_PASS
#define LUNA_RANGE_PASS 0
#endif

namespace luna {
namespace {

// Work-group sizes: the per-cell list walks (sweep, update) run one wavefront per work-group, so a wavefront that ends
// early frees its slot without waiting for the others of its group; the rest run 256.
constexpr int BLOCK = 256, CELL_BLOCK = 64;

// The logical block of this work-group. Work-groups go to the 8 XCDs in turn (block b to XCD b mod 8), so XCD x gets the
// contiguous logical run [x q + min(x, r), ...) of q + (x < r) blocks (q = n / 8, r = n mod 8): a bijection.
__device__ __forceinline__ int xcd_block() {
    const int b = blockIdx.x, n = gridDim.x;
    const int q = n >> 3, rem = n & 7, x = b & 7, k = b >> 3;
    return x * q + min(x, rem) + k;
}

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
template <int B>
__device__ __forceinline__ void block_min(double v, unsigned long long* out) {
    __shared__ double part[B / 64];
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
template <int B>
__device__ __forceinline__ void block_count(int v, unsigned long long* out) {
    __shared__ int part[B / 64];
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
__device__ __forceinline__ void separation(const Geo& g, const double* xi, const double* xj, double* d) {
#pragma unroll
    for (int a = 0; a < D; ++a) d[a] = wrap(xj[a] - xi[a], g.side[a], g.inverse[a]);
}

template <int D>
__device__ __forceinline__ void load_vec(const double* p, double* v) {
#pragma unroll
    for (int a = 0; a < D; ++a) v[a] = p[a];
}

// The slot of cell i's first list entry; entry t is 64 t further.
__device__ __forceinline__ int first_slot(const Geo& g, int i) { return g.slice[i >> 6] + (i & (SLICE - 1)); }

// finite_volume.hpp:403-415 primitives(): the primitive half of the cell row.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_primitives(Geo g, Rows r, Gas gas) {
    constexpr int V = D + 2;
    using C = CellAt<D>;
    const int i = xcd_block() * BLOCK + threadIdx.x;
    if (i >= g.N) return;
    double* row = g.cell + (size_t)i * CELL_ROW;
    const double inv = 1.0 / g.volume[i];
    double U[V], W[V];
#pragma unroll
    for (int v = 0; v < V; ++v) U[v] = r.q[(size_t)i * V + v] * inv;
    primitive<D>(gas, U, W);
#pragma unroll
    for (int v = 0; v < V; ++v) row[C::W + v] = W[v];
    row[C::CS] = sound_speed<D>(gas, W);
}

// One list entry as the sweep's first walk reads it: the face's |A| (from the entry's slot) and the other cell's row.
template <int D>
struct Near {
    double mag, x[D], c[D], W[D + 2], cs;
};
template <int D>
__device__ __forceinline__ void load_near(const Geo& g, int slot, int j, Near<D>& n) {
    using C = CellAt<D>;
    n.mag = g.slot[slot];
    const double* row = g.cell + (size_t)j * CELL_ROW;
    load_vec<D>(row + C::X, n.x);
    load_vec<D>(row + C::C, n.c);
#pragma unroll
    for (int v = 0; v < D + 2; ++v) n.W[v] = row[C::W + v];
    n.cs = row[C::CS];
}

// The gradient rows of the precision policies (precision.hpp:26-54, 55-130). FP16 (grad16): each group (rho, v, P) of the
// cell's gradient is scaled by 2^k so that its largest component is in [2^14, 2^15) (k = 0 for an all-zero group, within
// [-114, 126]); the three k are kept as int8 in one 32-bit word per cell. The conversions are the hardware's
// round-to-nearest-even ones.
template <int D>
__device__ __forceinline__ int grad_group(int c) {  // half::group(): 0 rho, 1 velocity, 2 pressure
    return c < D ? 0 : c < D + D * D ? 1 : 2;
}
__device__ __forceinline__ float power_of_two(int k) { return __uint_as_float((unsigned)(k + 127) << 23); }
__device__ __forceinline__ int group_exponent(float m) {
    if (!(m > 0)) return 0;
    const int e = (int)((__float_as_uint(m) >> 23) & 0xff) - 127;
    return min(max(14 - e, -114), 126);
}
template <int D, int Prec, class G>
__device__ __forceinline__ void store_gradient(const Rows& r, int i, const G (&grad)[D + 2][D]) {
    constexpr int V = D + 2;
    if constexpr (Prec == 0) {
        double* out = r.grad + (size_t)i * grad_row<D>();
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int a = 0; a < D; ++a) out[v * D + a] = grad[v][a];
    } else if constexpr (Prec == 1) {
        float* out = reinterpret_cast<float*>(r.grad) + (size_t)i * lp_row<D>();
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int a = 0; a < D; ++a) out[v * D + a] = grad[v][a];
    } else {
        float largest[3] = {0, 0, 0};
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const int q = grad_group<D>(D * v + a);
                largest[q] = fmaxf(largest[q], fabsf(grad[v][a]));
            }
        int k[3];
        float scale[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            k[q] = group_exponent(largest[q]);
            scale[q] = power_of_two(k[q]);
        }
        unsigned short* out = reinterpret_cast<unsigned short*>(r.grad) + (size_t)i * HALF_ROW;
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int a = 0; a < D; ++a) out[v * D + a] = __half_as_ushort(__float2half_rn(grad[v][a] * scale[grad_group<D>(D * v + a)]));
        r.exponents[i] = (unsigned)(unsigned char)(signed char)k[0] | (unsigned)(unsigned char)(signed char)k[1] << 8 |
                         (unsigned)(unsigned char)(signed char)k[2] << 16;
    }
}
// The gradient row of cell i in float (reduced-precision policies), decoded from FP16 if need be (FaceFlux::gradient,
// :1288-1300).
template <int D, int Prec>
__device__ __forceinline__ void load_gradient_lp(const Rows& r, int i, float* gr) {
    constexpr int VD = (D + 2) * D;
    if constexpr (Prec == 1) {
        const float* p = reinterpret_cast<const float*>(r.grad) + (size_t)i * lp_row<D>();
#pragma unroll
        for (int c = 0; c < VD; ++c) gr[c] = p[c];
    } else {
        const unsigned short* p = reinterpret_cast<const unsigned short*>(r.grad) + (size_t)i * HALF_ROW;
        const unsigned k = r.exponents[i];
        float scale[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) scale[q] = power_of_two(-(int)(signed char)(unsigned char)(k >> (8 * q)));
#pragma unroll
        for (int c = 0; c < VD; ++c) gr[c] = __half2float(__ushort_as_half(p[c])) * scale[grad_group<D>(c)];
    }
}

// The largest finite value of the fit's type (finite_max_v).
template <class G>
__device__ __forceinline__ constexpr G finite_max() {
    return sizeof(G) == 8 ? (G)1.7976931348623157e308 : (G)3.4028234663852886e38f;
}

// The sweep's per-cell sums (finite_volume.hpp:486-515), in list order. G: the fit's type (double, or float for the
// reduced-precision policies, whose range is of W_j - W_i, rounded once, :466-469, 510-512).
template <int D, bool Moving, bool Grad, int Rule, class G = double, bool RangePass = false>
struct Sums {
    static constexpr int V = D + 2, U = D + 1;
    static constexpr bool exact = sizeof(G) == 8;
    double surface = 0, smallest_face = INFINITY, max_angle = 0, smallest = INFINITY;
    G lo[V], hi[V], M[U][U], b[V][U];

    __device__ __forceinline__ void init() {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            lo[v] = finite_max<G>();
            hi[v] = -finite_max<G>();
#pragma unroll
            for (int p = 0; p < U; ++p) b[v][p] = 0;
        }
#pragma unroll
        for (int p = 0; p < U; ++p)
#pragma unroll
            for (int s = 0; s < U; ++s) M[p][s] = 0;
    }
    __device__ __forceinline__ void add(const Geo& g, const Near<D>& n, const double* xi, const double* centroid, const double* Wi,
                                        double csi, double ri, double courant) {
        double d[D];
        separation<D>(g, xi, n.x, d);
        const double magnitude = n.mag;
        surface += magnitude;
        smallest_face = fmin(smallest_face, magnitude);
        const double dist = sqrt(dot<D>(d, d));
        if constexpr (Moving) max_angle = fmax(max_angle, face_angle<D>(magnitude) / (0.5 * dist));
        if constexpr (Grad) {
            G X[U];
            X[0] = 1;
#pragma unroll
            for (int p = 1; p < U; ++p) X[p] = (G)(d[p - 1] + n.c[p - 1] - centroid[p - 1]);
            const G w = (G)magnitude;
#pragma unroll
            for (int p = 0; p < U; ++p)
#pragma unroll
                for (int s = p; s < U; ++s) M[p][s] += w * X[p] * X[s];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const double dW = n.W[v] - Wi[v];
#pragma unroll
                for (int p = 0; p < U; ++p) b[v][p] += w * X[p] * (G)dW;
                if constexpr (!RangePass) {
                    const G value = exact ? (G)n.W[v] : (G)dW;
                    lo[v] = fmin(lo[v], value);
                    hi[v] = fmax(hi[v], value);
                }
            }
        }
        if constexpr (Rule == 0) {  // FaceStep (:1079-1089) from this side
            double dv[D];
#pragma unroll
            for (int a = 0; a < D; ++a) dv[a] = n.W[1 + a] - Wi[1 + a];
            const double vsig = csi + n.cs - dot<D>(dv, d) / dist;
            if (vsig > 0) smallest = fmin(smallest, courant * ((dist + ri) / vsig));
        }
    }
};

// The sweep (finite_volume.hpp:463-569 Sweep::operator(), double policy) with the time step's terms (:1040-1104) fused:
// each cell takes its own term and, for every face of its list, the face term with its own radius. The pair term vsig and
// |d| have the same bits from either side, so the minimum is the reference's (the face kernel's terms are the same set).
// The first block also resets the other dt slot and the step's fallback counters. Limit: 0 none, 1 the face-centroid
// limiter, 2 the same limiter recording the learned limiter's features and troubled-cell flag (the gradients are those of 1).
template <int D, bool Moving, bool Grad, int Limit, int Rule, int Prec = 0>
__global__ __launch_bounds__(CELL_BLOCK, Limit == 2 ? 1 : LUNA_SWEEP_WAVES) void k_sweep(Geo g, Rows r, Gas gas, double courant, Counters c) {
    constexpr int V = D + 2, U = D + 1;
    using G = typename std::conditional<Prec == 0, double, float>::type;
    constexpr bool exact = Prec == 0;
    static_assert(exact || Limit < 2, "the learned limiter's features need the double policy");
    using C = CellAt<D>;
    const int i = xcd_block() * CELL_BLOCK + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *c.dt_next = INF_BITS;
        c.fallbacks[0] = 0;
        c.fallbacks[1] = 0;
    }
    double smallest = INFINITY;
    if (i < g.N) {
        const double* mine = g.cell + (size_t)i * CELL_ROW;
        double Wi[V], xi[D], centroid[D];
#pragma unroll
        for (int v = 0; v < V; ++v) Wi[v] = mine[C::W + v];
        const double ci = mine[C::CS], ri = mine[C::RAD];
        load_vec<D>(mine + C::X, xi);
        load_vec<D>(mine + C::C, centroid);
        constexpr bool range_pass = LUNA_RANGE_PASS || (Moving && exact);  // the float sums fit 3 waves without it
        Sums<D, Moving, Grad, Rule, G, range_pass> s;
        s.init();
        const int count = g.count[i], base = first_slot(g, i);
        // Groups of UN entries: their face magnitudes and rows in flight together, summed in list order; the next group's
        // other-cell indices are read one group ahead, so each group waits for one level of loads.
        constexpr int UN = exact ? LUNA_SWEEP_UNROLL : 2;
        int nj[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u)
            if (u < count) nj[u] = g.other[base + u * SLICE];
        int t = 0;
        for (; t + UN - 1 < count; t += UN) {
            int cj[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                cj[u] = nj[u];
                if (t + UN + u < count) nj[u] = g.other[base + (t + UN + u) * SLICE];
            }
            Near<D> n[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u) load_near<D>(g, base + (t + u) * SLICE, cj[u], n[u]);
#pragma unroll
            for (int u = 0; u < UN; ++u) s.add(g, n[u], xi, centroid, Wi, ci, ri, courant);
        }
#pragma unroll
        for (int u = 0; u < UN - 1; ++u)
            if (t + u < count) {
                Near<D> a;
                load_near<D>(g, base + (t + u) * SLICE, nj[u], a);
                s.add(g, a, xi, centroid, Wi, ci, ri, courant);
            }
        smallest = s.smallest;
        const double tiny = 1e-6 * s.surface;
        const bool any_tiny = s.smallest_face < tiny;
        // neighbour_range (:572-588): the range of the neighbours' W, faces under 1e-6 of the surface left out if any
        auto neighbour_range = [&]() {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                s.lo[v] = finite_max<G>();
                s.hi[v] = -finite_max<G>();
            }
            for (int k = 0; k < count; ++k) {
                if (any_tiny && g.slot[base + k * SLICE] < tiny) continue;
                const double* row = g.cell + (size_t)g.other[base + k * SLICE] * CELL_ROW + C::W;
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const G value = exact ? (G)row[v] : (G)(row[v] - Wi[v]);
                    s.lo[v] = fmin(s.lo[v], value);
                    s.hi[v] = fmax(s.hi[v], value);
                }
            }
        };
        if constexpr (Grad && !range_pass) {
            if (any_tiny) neighbour_range();
        }
        G grad[V][D], intercept[V];
        if constexpr (Grad) {  // Sweep::solve (:590-615): Cholesky of M, then M x = b per variable
            G L[U][U];
#pragma unroll
            for (int p = 0; p < U; ++p)
#pragma unroll
                for (int q = 0; q < U; ++q) L[p][q] = 0;
#pragma unroll
            for (int p = 0; p < U; ++p)
#pragma unroll
                for (int q = 0; q <= p; ++q) {
                    G sum = s.M[q][p];
#pragma unroll
                    for (int m = 0; m < q; ++m) sum -= L[p][m] * L[q][m];
                    L[p][q] = p == q ? sqrt(sum) : sum / L[q][q];
                }
#pragma unroll
            for (int v = 0; v < V; ++v) {
                G z[U], x[U];
#pragma unroll
                for (int p = 0; p < U; ++p) {
                    G sum = s.b[v][p];
#pragma unroll
                    for (int m = 0; m < p; ++m) sum -= L[p][m] * z[m];
                    z[p] = sum / L[p][p];
                }
#pragma unroll
                for (int p = U - 1; p >= 0; --p) {
                    G sum = z[p];
#pragma unroll
                    for (int m = p + 1; m < U; ++m) sum -= L[m][p] * x[m];
                    x[p] = sum / L[p][p];
                }
#pragma unroll
                for (int a = 0; a < D; ++a) grad[v][a] = x[1 + a];
                intercept[v] = x[0];
            }
        }
        double wi[D];
#pragma unroll
        for (int a = 0; a < D; ++a) wi[a] = Wi[1 + a];
        if constexpr (Moving) {  // the regularised point velocity (:526-537; constants :74)
            const double shaping = 0.5, factor = 2.25;
            double fraction = 0;
            if (s.max_angle > 0.75 * factor)
                fraction = s.max_angle > factor ? shaping : shaping * (s.max_angle - 0.75 * factor) / (0.25 * factor);
            const double cd = sqrt(dot<D>(centroid, centroid));
            if (cd > 0 && fraction > 0) {
                const double sc = fraction * ci / cd;
#pragma unroll
                for (int a = 0; a < D; ++a) wi[a] = wi[a] + centroid[a] * sc;
            }
#pragma unroll
            for (int a = 0; a < D; ++a) {
                r.w[(size_t)i * D + a] = wi[a];
                r.old_centroid[(size_t)i * D + a] = centroid[a];  // keep_centroids (:1506-1514): c^n for the second pass
            }
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
            if constexpr (Limit > 0 && range_pass) neighbour_range();
            if constexpr (Limit > 0) {  // the face-centroid limiter, face by face in list order (:538-565)
                // Geometry::face_offset (:322-326), kept per slot, of UL entries loaded together, then applied in order.
                constexpr int UL = LUNA_LIMIT_UNROLL;
                Record<D, Limit == 2> rec;
                if constexpr (exact) rec.init(grad);
                for (int t2 = 0; t2 < count; t2 += UL) {
                    double offset[UL][D];
                    bool use[UL];
#pragma unroll
                    for (int u = 0; u < UL; ++u) {
                        use[u] = false;
                        if (t2 + u < count) {
                            const long long slot = base + (t2 + u) * SLICE;
                            use[u] = !(any_tiny && g.slot[slot] < tiny);
#pragma unroll
                            for (int a = 0; a < D; ++a) offset[u][a] = g.slot[(a + 1) * g.slots + slot];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < UL; ++u) {
                        if (!use[u]) continue;
                        if constexpr (exact) {
                            rec.face(offset[u]);
#pragma unroll
                            for (int v = 0; v < V; ++v) {
                                const double dp = dot<D>(grad[v], offset[u]);
                                if (dp > 0 && Wi[v] + dp > s.hi[v]) {
                                    const double sc = s.hi[v] > Wi[v] ? (s.hi[v] - Wi[v]) / dp : 0.0;
#pragma unroll
                                    for (int a = 0; a < D; ++a) grad[v][a] = grad[v][a] * sc;
                                    rec.clip(v, sc);
                                } else if (dp < 0 && Wi[v] + dp < s.lo[v]) {
                                    const double sc = s.lo[v] < Wi[v] ? (s.lo[v] - Wi[v]) / dp : 0.0;
#pragma unroll
                                    for (int a = 0; a < D; ++a) grad[v][a] = grad[v][a] * sc;
                                    rec.clip(v, sc);
                                }
                            }
                        } else {  // reduced precision (:558-562): the offset, the face's change and the range relative to W_i in G
                            G rg[D];
#pragma unroll
                            for (int a = 0; a < D; ++a) rg[a] = (G)offset[u][a];
#pragma unroll
                            for (int v = 0; v < V; ++v) {
                                const G dp = dot<D>(grad[v], rg);
                                if (dp > 0 && dp > s.hi[v]) {
                                    const G sc = s.hi[v] > 0 ? s.hi[v] / dp : G(0);
#pragma unroll
                                    for (int a = 0; a < D; ++a) grad[v][a] = grad[v][a] * sc;
                                } else if (dp < 0 && dp < s.lo[v]) {
                                    const G sc = s.lo[v] < 0 ? s.lo[v] / dp : G(0);
#pragma unroll
                                    for (int a = 0; a < D; ++a) grad[v][a] = grad[v][a] * sc;
                                }
                            }
                        }
                    }
                }
                if constexpr (Limit == 2 && exact) {
                    Cell<D> cell;
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        cell.W[v] = Wi[v];
                        cell.lo[v] = s.lo[v];
                        cell.hi[v] = s.hi[v];
                        cell.intercept[v] = intercept[v];
                    }
                    cell.cs = ci;
                    cell.r = ri;
                    cell.count = count;
                    r.troubled[i] = limiter_features<D>(cell, rec, r.features + (size_t)i * V * FEATURES);
                }
            }
            store_gradient<D, Prec>(r, i, grad);
            if constexpr (exact && Limit < 2) {
                if (r.hll) {  // the compression indicator of the HLL fallback
                    double div = 0;
#pragma unroll
                    for (int a = 0; a < D; ++a) div += grad[1 + a][a];
                    r.hll[i] = div * ri / ci < -r.hll_threshold ? 1 : 0;
                }
            }
        }
    }
    block_min<CELL_BLOCK>(smallest, c.dt);
}

// finite_volume.hpp:880-1037 LaneSweep: L lanes per cell (L = 2, 4 or 8; 64 / L cells per wavefront). Lane l takes the
// list entries l, l + L, ... in list order with the one-lane sweep's arithmetic per entry (Sums::add, the time step's
// terms included); the lanes' sums are combined by a butterfly of shuffles (stage o = L/2, ..., 1: each lane adds lane
// l^o's value; addition commutes, so every lane of the cell holds the same bits after each stage), the minima and maxima
// likewise. The limiter takes, per variable, the smallest factor over the faces, each from the unlimited gradient, and
// scales the gradient once. Bits: the same run to run, other bits than the one-lane sweep's (its sums in list order, its
// clips one after the other). Double policy, order 1, the face-centroid limiter (Limit 1) or none (Limit 0).
template <int L>
__device__ __forceinline__ double lanes_sum(double v) {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v = v + __shfl_xor(v, o);
    return v;
}
template <int L>
__device__ __forceinline__ double lanes_min(double v) {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
template <int L>
__device__ __forceinline__ double lanes_max(double v) {
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

template <int D, bool Moving, int Limit, int Rule, int L>
__global__ __launch_bounds__(CELL_BLOCK, LUNA_LANE_WAVES(L)) void k_sweep_lanes(Geo g, Rows r, Gas gas, double courant, Counters c) {
    constexpr int V = D + 2, U = D + 1, CELLS = CELL_BLOCK / L;
    using C = CellAt<D>;
    const int lane = threadIdx.x & (L - 1);
    const int i = xcd_block() * CELLS + threadIdx.x / L;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *c.dt_next = INF_BITS;
        c.fallbacks[0] = 0;
        c.fallbacks[1] = 0;
    }
    double smallest = INFINITY;
    if (i < g.N) {  // the same for the L lanes of a cell, so their shuffles stay among active lanes
        const double* mine = g.cell + (size_t)i * CELL_ROW;
        double Wi[V], xi[D], centroid[D];
#pragma unroll
        for (int v = 0; v < V; ++v) Wi[v] = mine[C::W + v];
        const double ci = mine[C::CS], ri = mine[C::RAD];
        load_vec<D>(mine + C::X, xi);
        load_vec<D>(mine + C::C, centroid);
        Sums<D, Moving, true, Rule, double, false> s;
        s.init();
        const int count = g.count[i], base = first_slot(g, i);
        for (int t = lane; t < count; t += L) {
            Near<D> n;
            load_near<D>(g, base + t * SLICE, g.other[base + t * SLICE], n);
            s.add(g, n, xi, centroid, Wi, ci, ri, courant);
        }
        s.surface = lanes_sum<L>(s.surface);
        s.smallest_face = lanes_min<L>(s.smallest_face);
        if constexpr (Moving) s.max_angle = lanes_max<L>(s.max_angle);
        smallest = lanes_min<L>(s.smallest);
#pragma unroll
        for (int p = 0; p < U; ++p)
#pragma unroll
            for (int q = p; q < U; ++q) s.M[p][q] = lanes_sum<L>(s.M[p][q]);
#pragma unroll
        for (int v = 0; v < V; ++v) {
#pragma unroll
            for (int p = 0; p < U; ++p) s.b[v][p] = lanes_sum<L>(s.b[v][p]);
            s.lo[v] = lanes_min<L>(s.lo[v]);
            s.hi[v] = lanes_max<L>(s.hi[v]);
        }
        const double tiny = 1e-6 * s.surface;
        const bool any_tiny = s.smallest_face < tiny;
        if (any_tiny) {  // rare: the range again without the tiny faces, each lane walking the whole list
#pragma unroll
            for (int v = 0; v < V; ++v) {
                s.lo[v] = finite_max<double>();
                s.hi[v] = -finite_max<double>();
            }
            for (int k = 0; k < count; ++k) {
                if (g.slot[base + k * SLICE] < tiny) continue;
                const double* row = g.cell + (size_t)g.other[base + k * SLICE] * CELL_ROW + C::W;
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    s.lo[v] = fmin(s.lo[v], row[v]);
                    s.hi[v] = fmax(s.hi[v], row[v]);
                }
            }
        }
        double grad[V][D];
        {  // Sweep::solve (:590-615), as the one-lane sweep
            double Lc[U][U];
#pragma unroll
            for (int p = 0; p < U; ++p)
#pragma unroll
                for (int q = 0; q < U; ++q) Lc[p][q] = 0;
#pragma unroll
            for (int p = 0; p < U; ++p)
#pragma unroll
                for (int q = 0; q <= p; ++q) {
                    double sum = s.M[q][p];
#pragma unroll
                    for (int m = 0; m < q; ++m) sum -= Lc[p][m] * Lc[q][m];
                    Lc[p][q] = p == q ? sqrt(sum) : sum / Lc[q][q];
                }
#pragma unroll
            for (int v = 0; v < V; ++v) {
                double z[U], x[U];
#pragma unroll
                for (int p = 0; p < U; ++p) {
                    double sum = s.b[v][p];
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
                for (int a = 0; a < D; ++a) grad[v][a] = x[1 + a];
            }
        }
        double wi[D];
#pragma unroll
        for (int a = 0; a < D; ++a) wi[a] = Wi[1 + a];
        if constexpr (Moving) {  // the regularised point velocity (:526-537), as the one-lane sweep
            const double shaping = 0.5, factor = 2.25;
            double fraction = 0;
            if (s.max_angle > 0.75 * factor)
                fraction = s.max_angle > factor ? shaping : shaping * (s.max_angle - 0.75 * factor) / (0.25 * factor);
            const double cd = sqrt(dot<D>(centroid, centroid));
            if (cd > 0 && fraction > 0) {
                const double sc = fraction * ci / cd;
#pragma unroll
                for (int a = 0; a < D; ++a) wi[a] = wi[a] + centroid[a] * sc;
            }
            if (lane == 0) {
#pragma unroll
                for (int a = 0; a < D; ++a) {
                    r.w[(size_t)i * D + a] = wi[a];
                    r.old_centroid[(size_t)i * D + a] = centroid[a];
                }
            }
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
        if constexpr (Limit > 0) {  // LaneSweep's limiter (:1006-1034): the smallest factor per variable
            double f[V];
#pragma unroll
            for (int v = 0; v < V; ++v) f[v] = 1.0;
            for (int t = lane; t < count; t += L) {
                const long long slot = base + t * SLICE;
                if (any_tiny && g.slot[slot] < tiny) continue;
                double offset[D];
#pragma unroll
                for (int a = 0; a < D; ++a) offset[a] = g.slot[(a + 1) * g.slots + slot];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const double dp = dot<D>(grad[v], offset);
                    double factor = 1.0;
                    if (dp > 0 && Wi[v] + dp > s.hi[v])
                        factor = s.hi[v] > Wi[v] ? (s.hi[v] - Wi[v]) / dp : 0.0;
                    else if (dp < 0 && Wi[v] + dp < s.lo[v])
                        factor = s.lo[v] < Wi[v] ? (s.lo[v] - Wi[v]) / dp : 0.0;
                    f[v] = fmin(f[v], factor);
                }
            }
#pragma unroll
            for (int v = 0; v < V; ++v) {
                f[v] = lanes_min<L>(f[v]);
                if (f[v] < 1.0)
#pragma unroll
                    for (int a = 0; a < D; ++a) grad[v][a] = grad[v][a] * f[v];
            }
        }
        if (lane == 0) {
            store_gradient<D, 0>(r, i, grad);
            if (r.hll) {  // the compression indicator of the HLL fallback
                double div = 0;
#pragma unroll
                for (int a = 0; a < D; ++a) div += grad[1 + a][a];
                r.hll[i] = div * ri / ci < -r.hll_threshold ? 1 : 0;
            }
        }
    }
    block_min<CELL_BLOCK>(smallest, c.dt);
}

// finite_volume.hpp:1306-1329 FaceFlux::side() at order 1: half step (over step_dt) + g.r, positivity fallback. prow,
// grow: the cell's primitive half and gradient row, read here so that only one side's rows are live at a time.
template <int D>
__device__ __forceinline__ void face_side(const Gas& gas, const double* prow, const double* grow, const double* off, double step_dt,
                                          int order, double* Wf, int& fallbacks) {
    constexpr int V = D + 2;
    double W[V];
#pragma unroll
    for (int v = 0; v < V; ++v) W[v] = prow[v];
    if (order == 0) {
#pragma unroll
        for (int v = 0; v < V; ++v) Wf[v] = W[v];
        return;
    }
    double gr[V * D];
#pragma unroll
    for (int k = 0; k < V * D; ++k) gr[k] = grow[k];
    half_step<D>(gas, W, gr, step_dt, Wf);
#pragma unroll
    for (int v = 0; v < V; ++v) Wf[v] += dot<D>(gr + v * D, off);
    if (Wf[0] > 0 && Wf[D + 1] > 0) return;
    ++fallbacks;
#pragma unroll
    for (int v = 0; v < V; ++v) Wf[v] = W[v];
}

// finite_volume.hpp:1451-1465 MovingFlux<false>::pass_side(): W + g.r on the mesh of t^n.
template <int D>
__device__ __forceinline__ void pass_side_first(const double* prow, const double* grow, const double* off, double* Wf, int& fallbacks) {
    constexpr int V = D + 2;
    double W[V];
#pragma unroll
    for (int v = 0; v < V; ++v) W[v] = prow[v];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        double gv[D];
#pragma unroll
        for (int a = 0; a < D; ++a) gv[a] = grow[v * D + a];
        Wf[v] = W[v] + dot<D>(gv, off);
    }
    if (Wf[0] > 0 && Wf[D + 1] > 0) return;
    ++fallbacks;
#pragma unroll
    for (int v = 0; v < V; ++v) Wf[v] = W[v];
}

// The sides of the reduced-precision policies (FaceFlux::side(), :1306-1329, float branch; MovingFlux<false>::pass_side(),
// :1451-1465): W in double; the gradient, the offset, the half step's rate and dt / 2 in float; W + (double)(g.r - dt/2
// rate) or W + (double)(g.r).
template <int D, int Prec>
__device__ __forceinline__ void face_side_p(const Gas& gas, const Rows& r, int i, const double* prow, const double* grow, const double* off,
                                            double step_dt, int order, double* Wf, int& fallbacks) {
    constexpr int V = D + 2;
    if constexpr (Prec == 0) {
        face_side<D>(gas, prow, grow, off, step_dt, order, Wf, fallbacks);
    } else {
        double W[V];
#pragma unroll
        for (int v = 0; v < V; ++v) W[v] = prow[v];
        if (order == 0) {
#pragma unroll
            for (int v = 0; v < V; ++v) Wf[v] = W[v];
            return;
        }
        float gr[V * D], Wg[V], A[V], rg[D];
        load_gradient_lp<D, Prec>(r, i, gr);
#pragma unroll
        for (int v = 0; v < V; ++v) Wg[v] = (float)W[v];
        half_step_rate<D, float>(gas, Wg, gr, A);
#pragma unroll
        for (int a = 0; a < D; ++a) rg[a] = (float)off[a];
        const float h = 0.5f * (float)step_dt;
#pragma unroll
        for (int v = 0; v < V; ++v) Wf[v] = W[v] + (double)(dot<D>(gr + v * D, rg) - h * A[v]);
        if (Wf[0] > 0 && Wf[D + 1] > 0) return;
        ++fallbacks;
#pragma unroll
        for (int v = 0; v < V; ++v) Wf[v] = W[v];
    }
}
template <int D, int Prec>
__device__ __forceinline__ void pass_side_first_p(const Rows& r, int i, const double* prow, const double* grow, const double* off, double* Wf,
                                                  int& fallbacks) {
    constexpr int V = D + 2;
    if constexpr (Prec == 0) {
        pass_side_first<D>(prow, grow, off, Wf, fallbacks);
    } else {
        double W[V];
#pragma unroll
        for (int v = 0; v < V; ++v) W[v] = prow[v];
        float gr[V * D], rg[D];
        load_gradient_lp<D, Prec>(r, i, gr);
#pragma unroll
        for (int a = 0; a < D; ++a) rg[a] = (float)off[a];
#pragma unroll
        for (int v = 0; v < V; ++v) Wf[v] = W[v] + (double)dot<D>(gr + v * D, rg);
        if (Wf[0] > 0 && Wf[D + 1] > 0) return;
        ++fallbacks;
#pragma unroll
        for (int v = 0; v < V; ++v) Wf[v] = W[v];
    }
}

// The position and centroid of one side of a face, and where its primitive half and gradient row are.
template <int D>
struct SideRows {
    double x[D], c[D];
    const double* p;
    const double* g;
};
template <int D>
__device__ __forceinline__ void load_side(const Geo& g, const Rows& r, int i, SideRows<D>& s) {
    using C = CellAt<D>;
    const double* row = g.cell + (size_t)i * CELL_ROW;
    s.p = row + C::W;
    s.g = r.grad + (size_t)i * grad_row<D>();
    load_vec<D>(row + C::X, s.x);
    load_vec<D>(row + C::C, s.c);
}

__device__ __forceinline__ long long wrap_key(long long k, long long p);

// One face's flux times |A|: FaceFlux::operator() (:1415-1434) statically, MovingFlux::operator() (:1466-1490) on a moving
// step; the positivity fallbacks of the owner's and the neighbour's side added to count_o and count_j. Record: read the
// face from its 64 B record (the per-cell loop, which gathers faces through the lists: one line each), else from the
// mesh's face arrays in place (the per-face loop, which streams them).
template <int D, int Pass, int Riemann, int Prec, bool Record = false>
__device__ __forceinline__ void flux_of_face(const Geo& g, const Rows& r, const Gas& gas, int order, double dt, double unit, int f, double* F,
                                             int& count_o, int& count_j) {
    constexpr int V = D + 2;
    using FA = FaceAt<D>;
    {
        // A, the face centroid (relative to the owner's generator), owner, neighbour, |A|
        double rec[FACE_ROW - 1];
        int o, j;
        double magnitude;
        if constexpr (Record) {
            const double* face = g.face + (size_t)f * FACE_ROW;
#pragma unroll
            for (int k = 0; k < FACE_ROW - 1; ++k) rec[k] = face[k];
            const int2 pair = *reinterpret_cast<const int2*>(face + FA::PAIR);
            o = pair.x;
            j = pair.y;
            magnitude = rec[FA::MAG];
        } else {
#pragma unroll
            for (int k = 0; k < D; ++k) {
                rec[FA::A + k] = g.area[(size_t)f * D + k];
                rec[FA::FC + k] = g.fcen[(size_t)f * D + k];
            }
            o = g.owner[f];
            j = g.neighbour[f];
            magnitude = sqrt(dot<D>(rec + FA::A, rec + FA::A));  // |A| as k_geo_faces writes it
        }
#pragma unroll
        for (int v = 0; v < V; ++v) F[v] = 0;
        if (magnitude != 0) {
            SideRows<D> so, sj;
            load_side<D>(g, r, o, so);
            load_side<D>(g, r, j, sj);
            double L[V], R[V], n[D], wf[D], ro[D], rj[D], d[D];
            const double inv = 1.0 / magnitude;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                n[k] = rec[FA::A + k] * inv;
                wf[k] = 0;
            }
            // face_offset (:322-326) of both sides
            separation<D>(g, sj.x, so.x, d);  // the owner seen from the neighbour
#pragma unroll
            for (int k = 0; k < D; ++k) {
                ro[k] = rec[FA::FC + k] - so.c[k];
                rj[k] = rec[FA::FC + k] + d[k] - sj.c[k];
            }
            if constexpr (Pass < 0) {
                const double half = 0.5 * dt;
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    ro[k] = ro[k] + wf[k] * half;
                    rj[k] = rj[k] + wf[k] * half;
                }
                face_side_p<D, Prec>(gas, r, o, so.p, so.g, ro, dt, order, L, count_o);
                face_side_p<D, Prec>(gas, r, j, sj.p, sj.g, rj, dt, order, R, count_j);
            } else {
                double wo[D], wj[D];
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    if constexpr (Pass == 2) {  // order 0 on a moving mesh: the sweep's point velocities (FaceFlux<0, true>)
                        wo[k] = r.w[(size_t)o * D + k];
                        wj[k] = r.w[(size_t)j * D + k];
                    } else {
                        wo[k] = (double)r.steps[(size_t)o * D + k] * unit;
                        wj[k] = (double)r.steps[(size_t)j * D + k] * unit;
                    }
                }
                // face_velocity (:1380-1397)
                double sep[D], cf[D], dw[D];
                separation<D>(g, so.x, sj.x, sep);
                const double dist = sqrt(dot<D>(sep, sep));
                const double inv_dist = 1.0 / dist;
#pragma unroll
                for (int k = 0; k < D; ++k) cf[k] = rec[FA::FC + k] - sep[k] * 0.5;
#pragma unroll
                for (int k = 0; k < D; ++k) dw[k] = wo[k] - wj[k];
                double facv = dot<D>(cf, dw) * inv_dist;
                const double limit = 0.9 * dist, c2 = dot<D>(cf, cf);
                if (c2 > limit * limit) facv *= limit / sqrt(c2);
#pragma unroll
                for (int k = 0; k < D; ++k) wf[k] = (wo[k] + wj[k]) * 0.5 + sep[k] * (facv * inv_dist);
                if constexpr (Pass == 2) {
                    face_side_p<D, Prec>(gas, r, o, so.p, so.g, ro, dt, 0, L, count_o);
                    face_side_p<D, Prec>(gas, r, j, sj.p, sj.g, rj, dt, 0, R, count_j);
                } else if constexpr (Pass == 0) {
                    pass_side_first_p<D, Prec>(r, o, so.p, so.g, ro, L, count_o);
                    pass_side_first_p<D, Prec>(r, j, sj.p, sj.g, rj, R, count_j);
                } else {
                    // side(i, r + c^n+1 - c^n + w dt) with the FaceFlux dt = 2 dt (:1452-1454)
                    const double step_dt = 2 * dt;
                    double r2[D];
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        const double moved = wo[k] * (0.5 * step_dt);
                        r2[k] = ro[k] + so.c[k] - r.old_centroid[(size_t)o * D + k] + moved;
                    }
                    face_side_p<D, Prec>(gas, r, o, so.p, so.g, r2, step_dt, 1, L, count_o);
#pragma unroll
                    for (int k = 0; k < D; ++k) {
                        const double moved = wj[k] * (0.5 * step_dt);
                        r2[k] = rj[k] + sj.c[k] - r.old_centroid[(size_t)j * D + k] + moved;
                    }
                    face_side_p<D, Prec>(gas, r, j, sj.p, sj.g, r2, step_dt, 1, R, count_j);
                }
            }
            if constexpr (Riemann == 1)
                hll<D>(gas, L, R, n, wf, F);
            else if constexpr (Riemann == 2) {  // HLLC, HLL where a side is flagged as extremely compressed
                if (r.hll[o] | r.hll[j])
                    hll<D>(gas, L, R, n, wf, F);
                else
                    hllc<D>(gas, L, R, n, wf, F);
            } else
                hllc<D>(gas, L, R, n, wf, F);
#pragma unroll
            for (int v = 0; v < V; ++v) F[v] *= magnitude;
        }
    }
}

// One face's flux row per work-item (the per-face loop, Update::per_face).
template <int D, int Pass, int Riemann, int Prec = 0>
__global__ __launch_bounds__(BLOCK, LUNA_FLUX_WAVES) void k_flux(Geo g, Rows r, Gas gas, int order, double dt, double unit, unsigned long long* fallbacks) {
    constexpr int V = D + 2;
    const int f = xcd_block() * BLOCK + threadIdx.x;
    int count = 0;
    if (f < g.F) {
        double F[V];
        int co = 0, cj = 0;
        flux_of_face<D, Pass, Riemann, Prec>(g, r, gas, order, dt, unit, f, F, co, cj);
        count = co + cj;
#pragma unroll
        for (int v = 0; v < V; ++v) r.flux[(size_t)f * V + v] = F[v];
    }
    block_count<BLOCK>(count, fallbacks);
}

// finite_volume.hpp:1544-1566 fluxes_and_update (Update::per_cell): each cell computes the flux of every face in its list
// (each face twice, once from each side, with the same arithmetic, so the same bits) and Q -= dt sum_k sign_k F in list order,
// in place (the fluxes read only the primitive and gradient rows); each side's fallbacks are counted by its own cell. No
// flux rows; the next primitives are left to the next time_step (the fluxes of the other cells still read this cell's).
template <int D, int Pass, int Riemann, int Prec, bool Drift>
__global__ __launch_bounds__(CELL_BLOCK) void k_fluxes_and_update(Geo g, Rows r, Gas gas, int order, double flux_dt, double dt, double unit,
                                                                  Grid grid, unsigned long long* fallbacks) {
    constexpr int V = D + 2;
    const int i = xcd_block() * CELL_BLOCK + threadIdx.x;
    int count = 0;
    if (i < g.N) {
        double sum[V];
#pragma unroll
        for (int v = 0; v < V; ++v) sum[v] = 0;
        const int n = g.count[i], base = first_slot(g, i);
        for (int t = 0; t < n; ++t) {
            const int e = g.list[base + t * SLICE], f = e >= 0 ? e : ~e;
            double F[V];
            int co = 0, cj = 0;
            flux_of_face<D, Pass, Riemann, Prec, true>(g, r, gas, order, flux_dt, unit, f, F, co, cj);
            count += e >= 0 ? co : cj;
            const double sign = e >= 0 ? 1.0 : -1.0;
#pragma unroll
            for (int v = 0; v < V; ++v) sum[v] += sign * F[v];
        }
#pragma unroll
        for (int v = 0; v < V; ++v) r.q[(size_t)i * V + v] = r.q[(size_t)i * V + v] - dt * sum[v];
        if constexpr (Drift) {  // finite_volume.hpp:1219-1231, as k_update
#pragma unroll
            for (int a = 0; a < D; ++a) {
                const size_t at = (size_t)i * D + a;
                const long long k = (long long)round(r.x[at] * grid.per_step) + r.steps[at];
                r.x[at] = (double)wrap_key(k, grid.period[a]) * grid.step;
            }
        }
    }
    block_count<CELL_BLOCK>(count, fallbacks);
}

// k mod p in [0, p): one conditional step for a key within a period of the box (every position and move here), the
// 64-bit remainder (a long instruction sequence on the GPU) only beyond that.
__device__ __forceinline__ long long wrap_key(long long k, long long p) {
    long long m = k >= p ? k - p : k < 0 ? k + p : k;
    if (m < 0 || m >= p) m = ((k % p) + p) % p;
    return m;
}

// finite_volume.hpp:1587-1603 update_body: Q -= dt sum_k sign_k F in list order; NextPrims: the primitive half of the
// cell row for the next step (a moving step's second pass, and every static step, which the next time_step() would
// compute the same way).
template <int D, bool NextPrims, bool Drift>
__global__ __launch_bounds__(CELL_BLOCK) void k_update(Geo g, Rows r, Gas gas, double dt, Grid grid) {
    constexpr int V = D + 2;
    using C = CellAt<D>;
    const int i = xcd_block() * CELL_BLOCK + threadIdx.x;
    if (i >= g.N) return;
    double sum[V];
#pragma unroll
    for (int v = 0; v < V; ++v) sum[v] = 0;
    const int count = g.count[i], base = first_slot(g, i);
    constexpr int UN = LUNA_UPDATE_UNROLL;
    int t = 0;
    for (; t + UN - 1 < count; t += UN) {  // UN rows in flight, summed in list order
        int e[UN];
#pragma unroll
        for (int u = 0; u < UN; ++u) e[u] = g.list[base + (t + u) * SLICE];
        double F[UN][V];
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const double* row = r.flux + (size_t)(e[u] >= 0 ? e[u] : ~e[u]) * V;
#pragma unroll
            for (int v = 0; v < V; ++v) F[u][v] = row[v];
        }
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const double sign = e[u] >= 0 ? 1.0 : -1.0;
#pragma unroll
            for (int v = 0; v < V; ++v) sum[v] += sign * F[u][v];
        }
    }
    for (; t < count; ++t) {
        const int e = g.list[base + t * SLICE], f = e >= 0 ? e : ~e;
        const double sign = e >= 0 ? 1.0 : -1.0;
#pragma unroll
        for (int v = 0; v < V; ++v) sum[v] += sign * r.flux[(size_t)f * V + v];
    }
    double next[V];
#pragma unroll
    for (int v = 0; v < V; ++v) r.q[(size_t)i * V + v] = next[v] = r.q[(size_t)i * V + v] - dt * sum[v];
    if constexpr (NextPrims) {
        double* row = g.cell + (size_t)i * CELL_ROW;
        const double inv = 1.0 / g.volume[i];  // the dense copy: this kernel then only writes the row
        double U[V], W[V];
#pragma unroll
        for (int v = 0; v < V; ++v) U[v] = next[v] * inv;
        primitive<D>(gas, U, W);
#pragma unroll
        for (int v = 0; v < V; ++v) row[C::W + v] = W[v];
        row[C::CS] = sound_speed<D>(gas, W);
    }
    if constexpr (Drift) {  // finite_volume.hpp:1219-1231: x to its grid point plus its whole steps, wrapped
#pragma unroll
        for (int a = 0; a < D; ++a) {
            const size_t at = (size_t)i * D + a;
            const long long k = (long long)round(r.x[at] * grid.per_step) + r.steps[at];
            r.x[at] = (double)wrap_key(k, grid.period[a]) * grid.step;
        }
    }
}

__device__ __forceinline__ long long snap_key(double x, const Grid& grid, int a) {
    return wrap_key((long long)round(x * grid.per_step), grid.period[a]);
}

// finite_volume.hpp:1506-1514 keep_centroids: x to its grid point; the move to x + dt w in whole grid steps (c^n is kept
// by the sweep, which reads it).
template <int D>
__global__ __launch_bounds__(BLOCK) void k_keep_centroids(Geo g, Rows r, Grid grid, double dt) {
    const int i = xcd_block() * BLOCK + threadIdx.x;
    if (i >= g.N) return;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const size_t at = (size_t)i * D + a;
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

// finite_volume.hpp:1227-1228 move() off the grid (order 0 on a moving mesh): x + dt w, wrapped into [0, side).
template <int D>
__global__ __launch_bounds__(BLOCK) void k_drift(Geo g, Rows r, double dt) {
    const int i = xcd_block() * BLOCK + threadIdx.x;
    if (i >= g.N) return;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        const size_t at = (size_t)i * D + a;
        const double moved = r.x[at] + dt * r.w[at];
        r.x[at] = moved - g.side[a] * floor(moved / g.side[a]);
    }
}

inline dim3 blocks(long long n, int b = BLOCK) { return dim3((unsigned)((n + b - 1) / b)); }

}  // namespace

template <int D>
hipError_t launch_drift(hipStream_t s, const Geo& g, const Rows& r, double dt) {
    if (g.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_drift<D>, blocks(g.N), dim3(BLOCK), 0, s, g, r, dt);
    return hipGetLastError();
}

template <int D>
hipError_t launch_primitives(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas) {
    if (g.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_primitives<D>, blocks(g.N), dim3(BLOCK), 0, s, g, r, gas);
    return hipGetLastError();
}

template <int D, bool Moving, bool Grad, int Limit, int Prec>
static hipError_t sweep_prec(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double courant, int rule, const Counters& c) {
    const dim3 grid = blocks(g.N > 0 ? g.N : 1, CELL_BLOCK);
    if (rule == 1)
        hipLaunchKernelGGL((k_sweep<D, Moving, Grad, Limit, 1, Prec>), grid, dim3(CELL_BLOCK), 0, s, g, r, gas, courant, c);
    else
        hipLaunchKernelGGL((k_sweep<D, Moving, Grad, Limit, 0, Prec>), grid, dim3(CELL_BLOCK), 0, s, g, r, gas, courant, c);
    return hipGetLastError();
}
template <int D, bool Moving, int Limit, int L>
static hipError_t sweep_lanes(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double courant, int rule, const Counters& c) {
    const dim3 grid = blocks(g.N > 0 ? (long long)g.N * L : 1, CELL_BLOCK);
    if (rule == 1)
        hipLaunchKernelGGL((k_sweep_lanes<D, Moving, Limit, 1, L>), grid, dim3(CELL_BLOCK), 0, s, g, r, gas, courant, c);
    else
        hipLaunchKernelGGL((k_sweep_lanes<D, Moving, Limit, 0, L>), grid, dim3(CELL_BLOCK), 0, s, g, r, gas, courant, c);
    return hipGetLastError();
}
template <int D, bool Moving, bool Grad, int Limit>
static hipError_t sweep_rule(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double courant, int rule, const Counters& c) {
    if constexpr (Grad && Limit < 2) {
        if (r.precision == 0 && r.sweep_lanes == 2) return sweep_lanes<D, Moving, Limit, 2>(s, g, r, gas, courant, rule, c);
        if (r.precision == 0 && r.sweep_lanes == 4) return sweep_lanes<D, Moving, Limit, 4>(s, g, r, gas, courant, rule, c);
        if (r.precision == 0 && r.sweep_lanes == 8) return sweep_lanes<D, Moving, Limit, 8>(s, g, r, gas, courant, rule, c);
        if (r.precision == 1) return sweep_prec<D, Moving, Grad, Limit, 1>(s, g, r, gas, courant, rule, c);
        if (r.precision == 2) return sweep_prec<D, Moving, Grad, Limit, 2>(s, g, r, gas, courant, rule, c);
    }
    return sweep_prec<D, Moving, Grad, Limit, 0>(s, g, r, gas, courant, rule, c);
}

template <int D>
hipError_t launch_sweep(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double courant, bool moving, bool grad, int limit,
                        int rule, const Counters& c) {
    if (moving) {
        if (!grad) return sweep_rule<D, true, false, 0>(s, g, r, gas, courant, rule, c);
        if (limit == 2) return sweep_rule<D, true, true, 2>(s, g, r, gas, courant, rule, c);
        return limit ? sweep_rule<D, true, true, 1>(s, g, r, gas, courant, rule, c) : sweep_rule<D, true, true, 0>(s, g, r, gas, courant, rule, c);
    }
    if (!grad) return sweep_rule<D, false, false, 0>(s, g, r, gas, courant, rule, c);
    if (limit == 2) return sweep_rule<D, false, true, 2>(s, g, r, gas, courant, rule, c);
    return limit ? sweep_rule<D, false, true, 1>(s, g, r, gas, courant, rule, c) : sweep_rule<D, false, true, 0>(s, g, r, gas, courant, rule, c);
}

template <int D, int Pass, int Prec>
static hipError_t flux_prec(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, double dt, double unit,
                            unsigned long long* fallbacks) {
    const dim3 grid = blocks(g.F > 0 ? g.F : 1);
    if (riemann == 1)
        hipLaunchKernelGGL((k_flux<D, Pass, 1, Prec>), grid, dim3(BLOCK), 0, s, g, r, gas, order, dt, unit, fallbacks);
    else if (Prec == 0 && riemann == 2)
        hipLaunchKernelGGL((k_flux<D, Pass, 2, 0>), grid, dim3(BLOCK), 0, s, g, r, gas, order, dt, unit, fallbacks);
    else
        hipLaunchKernelGGL((k_flux<D, Pass, 0, Prec>), grid, dim3(BLOCK), 0, s, g, r, gas, order, dt, unit, fallbacks);
    return hipGetLastError();
}
template <int D, int Pass>
static hipError_t flux_pass(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, double dt, double unit,
                            unsigned long long* fallbacks) {
    if (order == 1 && r.precision == 1) return flux_prec<D, Pass, 1>(s, g, r, gas, riemann, order, dt, unit, fallbacks);
    if (order == 1 && r.precision == 2) return flux_prec<D, Pass, 2>(s, g, r, gas, riemann, order, dt, unit, fallbacks);
    return flux_prec<D, Pass, 0>(s, g, r, gas, riemann, order, dt, unit, fallbacks);
}

template <int D>
hipError_t launch_flux(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, int pass, double dt, double unit,
                       unsigned long long* fallbacks) {
    if (pass < 0) return flux_pass<D, -1>(s, g, r, gas, riemann, order, dt, unit, fallbacks);
    if (pass == 0) return flux_pass<D, 0>(s, g, r, gas, riemann, order, dt, unit, fallbacks);
    if (pass == 2) return flux_pass<D, 2>(s, g, r, gas, riemann, 0, dt, unit, fallbacks);
    return flux_pass<D, 1>(s, g, r, gas, riemann, order, dt, unit, fallbacks);
}

template <int D, int Pass, int Riemann, int Prec>
static hipError_t per_cell_drift(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int order, double flux_dt, double dt, double unit,
                                 const Grid* drift, unsigned long long* fallbacks) {
    const dim3 grid = blocks(g.N > 0 ? g.N : 1, CELL_BLOCK);
    const Grid gr = drift ? *drift : Grid{};
    if (drift)
        hipLaunchKernelGGL((k_fluxes_and_update<D, Pass, Riemann, Prec, true>), grid, dim3(CELL_BLOCK), 0, s, g, r, gas, order, flux_dt, dt, unit,
                           gr, fallbacks);
    else
        hipLaunchKernelGGL((k_fluxes_and_update<D, Pass, Riemann, Prec, false>), grid, dim3(CELL_BLOCK), 0, s, g, r, gas, order, flux_dt, dt, unit,
                           gr, fallbacks);
    return hipGetLastError();
}
template <int D, int Pass, int Prec>
static hipError_t per_cell_riemann(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, double flux_dt, double dt,
                                   double unit, const Grid* drift, unsigned long long* fallbacks) {
    if (riemann == 1) return per_cell_drift<D, Pass, 1, Prec>(s, g, r, gas, order, flux_dt, dt, unit, drift, fallbacks);
    if (Prec == 0 && riemann == 2) return per_cell_drift<D, Pass, 2, 0>(s, g, r, gas, order, flux_dt, dt, unit, drift, fallbacks);
    return per_cell_drift<D, Pass, 0, Prec>(s, g, r, gas, order, flux_dt, dt, unit, drift, fallbacks);
}
template <int D, int Pass>
static hipError_t per_cell_prec(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, double flux_dt, double dt,
                                double unit, const Grid* drift, unsigned long long* fallbacks) {
    if (order == 1 && r.precision == 1) return per_cell_riemann<D, Pass, 1>(s, g, r, gas, riemann, order, flux_dt, dt, unit, drift, fallbacks);
    if (order == 1 && r.precision == 2) return per_cell_riemann<D, Pass, 2>(s, g, r, gas, riemann, order, flux_dt, dt, unit, drift, fallbacks);
    return per_cell_riemann<D, Pass, 0>(s, g, r, gas, riemann, order, flux_dt, dt, unit, drift, fallbacks);
}

template <int D>
hipError_t launch_fluxes_and_update(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, int riemann, int order, int pass, double flux_dt,
                                    double dt, double unit, const Grid* drift, unsigned long long* fallbacks) {
    if (g.N == 0) return hipSuccess;
    if (pass < 0) return per_cell_prec<D, -1>(s, g, r, gas, riemann, order, flux_dt, dt, unit, nullptr, fallbacks);
    if (pass == 0) return per_cell_prec<D, 0>(s, g, r, gas, riemann, order, flux_dt, dt, unit, drift, fallbacks);
    if (pass == 2) return per_cell_prec<D, 2>(s, g, r, gas, riemann, 0, flux_dt, dt, unit, nullptr, fallbacks);
    return per_cell_prec<D, 1>(s, g, r, gas, riemann, order, flux_dt, dt, unit, nullptr, fallbacks);
}

// LDS the update's launches reserve and leave unused: 16 KiB a work-group holds a CU to 10 of them (160 KiB of LDS). Each
// flux row is read twice, by its owner and by its neighbour, and the update gathers faster than L2 keeps lines: with
// fewer wavefronts in flight the second read finds the row in L2 more often (128^3: 0.451 -> 0.424 ms a launch; 8 KiB
// and 32 KiB measured slower).
constexpr size_t UPDATE_LDS_RESERVE = 16384;

template <int D>
hipError_t launch_update(hipStream_t s, const Geo& g, const Rows& r, const Gas& gas, double dt, bool next_prims, const Grid* drift) {
    if (g.N == 0) return hipSuccess;
    const Grid grid = drift ? *drift : Grid{};
    const size_t lds = UPDATE_LDS_RESERVE;
    if (drift)
        hipLaunchKernelGGL((k_update<D, false, true>), blocks(g.N, CELL_BLOCK), dim3(CELL_BLOCK), lds, s, g, r, gas, dt, grid);
    else if (next_prims)
        hipLaunchKernelGGL((k_update<D, true, false>), blocks(g.N, CELL_BLOCK), dim3(CELL_BLOCK), lds, s, g, r, gas, dt, grid);
    else
        hipLaunchKernelGGL((k_update<D, false, false>), blocks(g.N, CELL_BLOCK), dim3(CELL_BLOCK), lds, s, g, r, gas, dt, grid);
    return hipGetLastError();
}

template <int D>
hipError_t launch_keep_centroids(hipStream_t s, const Geo& g, const Rows& r, const Grid& grid, double dt) {
    if (g.N == 0) return hipSuccess;
    hipLaunchKernelGGL(k_keep_centroids<D>, blocks(g.N), dim3(BLOCK), 0, s, g, r, grid, dt);
    return hipGetLastError();
}

#define LUNA_INSTANTIATE(D)                                                                                                          \
    template hipError_t launch_primitives<D>(hipStream_t, const Geo&, const Rows&, const Gas&);                                     \
    template hipError_t launch_sweep<D>(hipStream_t, const Geo&, const Rows&, const Gas&, double, bool, bool, int, int,              \
                                        const Counters&);                                                                           \
    template hipError_t launch_flux<D>(hipStream_t, const Geo&, const Rows&, const Gas&, int, int, int, double, double,              \
                                       unsigned long long*);                                                                        \
    template hipError_t launch_update<D>(hipStream_t, const Geo&, const Rows&, const Gas&, double, bool, const Grid*);              \
    template hipError_t launch_fluxes_and_update<D>(hipStream_t, const Geo&, const Rows&, const Gas&, int, int, int, double, double, double, \
                                                    const Grid*, unsigned long long*);                                            \
    template hipError_t launch_keep_centroids<D>(hipStream_t, const Geo&, const Rows&, const Grid&, double);                      \
    template hipError_t launch_drift<D>(hipStream_t, const Geo&, const Rows&, double);

LUNA_INSTANTIATE(2)
LUNA_INSTANTIATE(3)

}  // namespace luna
