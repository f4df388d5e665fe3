// The 1D step on a sorted line (finite_volume_1d.hpp), one work-item per slot: the five-cell stencil in registers, three
// limited gradients, two face states, two Riemann fluxes and the new state in one kernel; fused, the next time step by
// LUNA's rule from the new state as a block min-reduction. On a moving line the same kernel also gives each point its
// velocity, new position and key; the slots' order is checked in it and restored by a device radix sort when points
// crossed, and the walls' new cells are split from the end cells. FP64, -ffp-contract=off.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include "derivatives.hpp"
#include "kernels.hpp"

namespace luna {
namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
    return v;
}
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

// sorted_mesh::Line as finite_volume_1d.hpp reads it: the gap (in grid steps) after slot s; a wall's gap is twice the
// distance to it (the mirror ghost, :152-155, :636-653).
__device__ __forceinline__ long long gap(const LineArgs& l, int s) {
    const int n = l.n;
    if (s >= 0 && s < n - 1) return l.key[s + 1] - l.key[s];
    if (l.periodic) return l.key[0] + l.period - l.key[n - 1];
    if (s < 0) return 2 * l.key[0];
    return 2 * (l.period - l.key[n - 1]);
}
__device__ __forceinline__ double length(const LineArgs& l, int s) { return (double)gap(l, s) * l.step; }
// finite_volume_1d.hpp:173, 176: V and the centroid relative to the generator from the gaps either side.
__device__ __forceinline__ double volume(const LineArgs& l, int s) { return ((double)gap(l, s - 1) + (double)gap(l, s)) * (0.5 * l.step); }
__device__ __forceinline__ double centroid(const LineArgs& l, int s) { return ((double)gap(l, s) - (double)gap(l, s - 1)) * (0.25 * l.step); }
// line.order(s): the cell in slot s.
__device__ __forceinline__ int cell_of(const LineArgs& l, int s) { return l.id ? l.id[s] : s; }
__device__ __forceinline__ int slot(const LineArgs& l, int s, int k) {
    const int t = s + k;
    if (l.periodic) return ((t % l.n) + l.n) % l.n;
    return t >= 0 && t < l.n ? t : -1;
}

// finite_volume_1d.hpp:107-111 primitive_of()
__device__ __forceinline__ void primitive_of(const LineArgs& l, int s, double* W) {
    const double inv = 1.0 / volume(l, s);
    const int i = cell_of(l, s);
    const double U[3] = {l.q[i * 3] * inv, l.q[i * 3 + 1] * inv, l.q[i * 3 + 2] * inv};
    primitive<1>(l.gas, U, W);
}

// finite_volume_1d.hpp:117-128 ghost(), ghost_face(): reflective mirrors the velocity, inflow is the wall's state,
// outflow copies.
__device__ __forceinline__ void ghost(const LineArgs& l, const double* W, bool left, double* out) {
    const int wall = left ? l.wall_left : l.wall_right;
    double a0 = W[0], a1 = wall == 0 ? -W[1] : W[1], a2 = W[2];
    if (wall == 2) {
        // read as global memory: a generic pointer here keeps the stencil's private arrays out of registers (scratch)
        typedef const __attribute__((address_space(1))) double global_double;
        global_double* in = (global_double*)(l.inflow) + (left ? 0 : 3);
        a0 = in[0];
        a1 = in[1];
        a2 = in[2];
    }
    out[0] = a0;
    out[1] = a1;
    out[2] = a2;
}
__device__ __forceinline__ void ghost_face(const LineArgs& l, const double* Wf, const double* W, bool left, double* out) {
    const int wall = left ? l.wall_left : l.wall_right;
    if (wall == 0) {
        out[0] = Wf[0];
        out[1] = -Wf[1];
        out[2] = Wf[2];
        return;
    }
    ghost(l, W, left, out);
}

struct Cell {
    double W[3], c, V;
    bool real;
};

// finite_volume_1d.hpp:132-156 load<2>()
__device__ __forceinline__ void load(const LineArgs& l, int s, Cell* cell, double* g) {
    int sl[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        sl[k] = slot(l, s, k - 2);
        cell[k].W[0] = cell[k].W[1] = cell[k].W[2] = 0;
        cell[k].c = cell[k].V = 0;
        cell[k].real = true;
    }
#pragma unroll
    for (int k = 0; k < 5; ++k)
        if (sl[k] >= 0) {
            primitive_of(l, sl[k], cell[k].W);
            cell[k].c = centroid(l, sl[k]);
            cell[k].V = volume(l, sl[k]);
        }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (sl[k] >= 0) continue;
        cell[k].real = false;
        const int end = k < 2 ? k + 1 : k - 1;
        if (sl[end] < 0) continue;
        ghost(l, cell[end].W, k < 2, cell[k].W);
        cell[k].c = -cell[end].c;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int a = sl[k], b = sl[k + 1];
        g[k] = a >= 0 ? length(l, a) : b >= 0 ? length(l, -1) : 0.0;
    }
}

// finite_volume_1d.hpp:183-205 Stencil::gradient(): limiter 0 face centroid (MC on a uniform line), 1 none, 2 minmod,
// 3 van Leer.
__device__ __forceinline__ double gradient(const LineArgs& l, const Cell* cell, double gapL, double gapR, int v) {
    const Cell& me = cell[0];
    const double dL = -gapL + cell[-1].c - me.c, dR = gapR + cell[1].c - me.c;
    const double Wi = me.W[v], left = cell[-1].W[v], right = cell[1].W[v];
    if (l.limiter == 2 || l.limiter == 3) {
        const double sL = (left - Wi) / dL, sR = (right - Wi) / dR;
        if (!(sL * sR > 0)) return 0.0;
        if (l.limiter == 3) return 2 * sL * sR / (sL + sR);
        return fabs(sL) < fabs(sR) ? sL : sR;
    }
    double g = ((right - Wi) - (left - Wi)) / (dR - dL);
    if (l.limiter == 1) return g;
    const double r[2] = {-0.5 * gapL - me.c, 0.5 * gapR - me.c};
    const double lo = fmin(left, right), hi = fmax(left, right);
#pragma unroll
    for (int f = 0; f < 2; ++f) {
        const double dp = g * r[f];
        if (dp > 0 && Wi + dp > hi)
            g = g * (hi > Wi ? (hi - Wi) / dp : 0.0);
        else if (dp < 0 && Wi + dp < lo)
            g = g * (lo < Wi ? (lo - Wi) / dp : 0.0);
    }
    return g;
}

// finite_volume_1d.hpp:492-505 Kernel::face_state() (static)
__device__ __forceinline__ void face_state(const LineArgs& l, const double* W, const double* g, double r, double dt, double* out) {
    if (l.order == 0) {
        out[0] = W[0];
        out[1] = W[1];
        out[2] = W[2];
        return;
    }
    half_step<1>(l.gas, W, g, dt, out);
#pragma unroll
    for (int v = 0; v < 3; ++v) out[v] += g[v] * r;
    if (out[0] > 0 && out[2] > 0) return;
    out[0] = W[0];
    out[1] = W[1];
    out[2] = W[2];
}

// ---- the moving line (finite_volume_1d.hpp with Motion::moving)

// finite_volume_1d.hpp:301-314 drift(): the point's velocity relative to its cell's from AREPO-ng's regularisation, for
// primitives W, density gradient g, centroid c (relative to the generator) and volume.
__device__ __forceinline__ double drift(const LineArgs& l, const double* W, double g, double c, double vol) {
    constexpr double shaping = 0.5;  // cell_shaping_speed, finite_volume_1d.hpp:26
    const double r = 0.5 * vol;
    if (!l.regularise || fabs(c) <= 0.5 * r) return 0.0;
    double offset = c;
    const double dgrad = fabs(g);
    if (dgrad > 0) {
        const double tmp = 3 * r + W[0] / dgrad;
        const double x = (tmp - sqrt(tmp * tmp - 8 * r * r)) / 4;
        if (x < 0.25 * r) offset += x / dgrad * g;
    }
    const double d = fabs(offset);
    if (d <= 0.75 * r) return 0.0;
    const double fraction = d > r ? shaping : shaping * (d - 0.75 * r) / (0.25 * r);
    return offset * (fraction * sound_speed<1>(l.gas, W) / d);
}

// finite_volume_1d.hpp:319-328 point_velocity(): drift(), plus, for a generator beyond 0.75 r of its centroid, up to the
// fluid's velocity at the generator, linear between this cell's centroid and that of its neighbour on that side.
__device__ __forceinline__ double point_velocity(const LineArgs& l, const Cell* cell, double gapL, double gapR, double g_rho, double vol) {
    const Cell& me = cell[0];
    const double r = 0.5 * vol, d = fabs(me.c), weight = fmin(1.0, (d - 0.75 * r) / (0.25 * r));
    const double drifting = drift(l, me.W, g_rho, me.c, vol);
    if (!(weight > 0)) return drifting;
    const bool right = me.c < 0;  // the neighbour on the generator's side
    const double between = right ? gapR + cell[1].c - me.c : gapL + me.c - cell[-1].c;
    const double v_side = right ? cell[1].W[1] : cell[-1].W[1];
    return drifting + weight * (v_side - me.W[1]) * (d / between);
}

// finite_volume_1d.hpp:546-550 moved(): where a point at x0 is after dt (wrapped on a periodic line, held between walls).
__device__ __forceinline__ double moved(const LineArgs& l, double x0, double velocity, double dt) {
    const double y = fma(dt, velocity, x0);
    return l.periodic ? y - l.side * floor(y / l.side) : fmin(fmax(y, 0.0), l.side);
}

// The grid key of a position: the nearest grid point, wrapped into [0, period) on a periodic line (sorted_mesh::snap).
__device__ __forceinline__ long long snap(const LineArgs& l, double x) {
    long long k = (long long)round(x * (1.0 / l.step));
    if (l.periodic) {
        k %= l.period;
        if (k < 0) k += l.period;
    }
    return k;
}

// The line's order: by key, then by cell (sorted_mesh::before).
__device__ __forceinline__ bool before(long long ka, int a, long long kb, int b) { return ka < kb || (ka == kb && a < b); }

// finite_volume_1d.hpp:509-524 extrapolated(): W + g r, a whole step on (MUSCL-Hancock) when `second`; W itself if rho
// or P would not stay positive.
__device__ __forceinline__ void extrapolated(const LineArgs& l, const double* W, const double* g, double r, int second, double dt,
                                             double* out) {
    if (l.order == 0) {
        out[0] = W[0];
        out[1] = W[1];
        out[2] = W[2];
        return;
    }
    if (second) {
        half_step<1>(l.gas, W, g, 2 * dt, out);
    } else {
        out[0] = W[0];
        out[1] = W[1];
        out[2] = W[2];
    }
#pragma unroll
    for (int v = 0; v < 3; ++v) out[v] += g[v] * r;
    if (out[0] > 0 && out[2] > 0) return;
    out[0] = W[0];
    out[1] = W[1];
    out[2] = W[2];
}

__device__ __forceinline__ void block_sum(unsigned long long v, unsigned long long* out) {
    __shared__ unsigned long long part[BLOCK / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m += part[k];
        if (m) atomicAdd(out, m);
    }
}

// finite_volume_1d.hpp:555-661 Kernel::body() (moving): gradients and point velocities of slots s - 1, s, s + 1, their new
// keys and moves in grid steps; each face's flux the mean of those on the meshes of t^n and t^n+1 (the face moving by the
// mean of its generators' moves; a wall face static); the new state, position and key of the slot's point; whether slot
// s and s + 1 are out of order after the move; fused, the next dt by LUNA's rule from the new keys and |w - v| of t^n.
__global__ __launch_bounds__(BLOCK) void k_line_step_moving(LineArgs l, double dt, double courant, int fused, unsigned long long* dt_out,
                                                            unsigned long long* dt_next, unsigned long long* descents,
                                                            unsigned long long* descents_next) {
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *descents_next = 0;
        if (fused) *dt_next = INF_BITS;
    }
    double smallest = INFINITY;
    unsigned long long descent = 0;
    if (s < l.n) {
        Cell cell[5];
        double g[4];
        load(l, s, cell, g);
        double grad[5][3], point[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 5; ++k) grad[k][0] = grad[k][1] = grad[k][2] = 0;
#pragma unroll
        for (int k = 1; k <= 3; ++k) {
            if (!cell[k].real) continue;
            if (l.order == 1)
#pragma unroll
                for (int v = 0; v < 3; ++v) grad[k][v] = gradient(l, cell + k, g[k - 1], g[k], v);
            point[k] = cell[k].W[1] + point_velocity(l, cell + k, g[k - 1], g[k], grad[k][0], 0.5 * (g[k - 1] + g[k]));
        }
        long long key[5] = {0, 0, 0, 0, 0};
        double moves[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 1; k <= 3; ++k) {
            if (!cell[k].real) continue;
            const int t = slot(l, s, k - 2);
            key[k] = snap(l, moved(l, l.x[cell_of(l, t)], point[k], dt));
            long long d = key[k] - l.key[t];
            if (l.periodic) d = d > l.period / 2 ? d - l.period : d < -l.period / 2 ? d + l.period : d;
            moves[k] = (double)d;
        }
        double F[2][3];
        const double one[1] = {1.0};
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            const int a = 1 + f, b = 2 + f;
            const bool inside = cell[a].real && cell[b].real;
            const double travel = inside ? 0.5 * (moves[a] + moves[b]) * l.step : 0.0;
            const double w[1] = {travel / dt};
            const double ra = 0.5 * g[a] - cell[a].c, rb = -0.5 * g[a] - cell[b].c;
            double on[2][3];
            const int meshes = l.order == 0 ? 1 : 2;
#pragma unroll
            for (int second = 0; second < 2; ++second) {
                if (second >= meshes) continue;
                double L[3] = {0, 0, 0}, R[3] = {0, 0, 0};
                if (cell[a].real) extrapolated(l, cell[a].W, grad[a], ra + second * travel, second, dt, L);
                if (cell[b].real) extrapolated(l, cell[b].W, grad[b], rb + second * travel, second, dt, R);
                if (!cell[a].real) ghost_face(l, R, cell[b].W, true, L);
                if (!cell[b].real) ghost_face(l, L, cell[a].W, false, R);
                if (l.riemann == 1)
                    hll<1>(l.gas, L, R, one, w, on[second]);
                else
                    hllc<1>(l.gas, L, R, one, w, on[second]);
            }
#pragma unroll
            for (int v = 0; v < 3; ++v) F[f][v] = meshes == 1 ? on[0][v] : 0.5 * (on[0][v] + on[1][v]);
        }
        const int i = cell_of(l, s);
        double Unew[3];
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            Unew[v] = l.q[i * 3 + v] - dt * (F[1][v] - F[0][v]);
            l.next[i * 3 + v] = Unew[v];
        }
        l.x_next[i] = moved(l, l.x[i], point[2], dt);
        l.key_next[s] = key[2];
        if (cell[3].real && s < l.n - 1) descent = !before(key[2], i, key[3], cell_of(l, s + 1));
        if (fused) {
            double left, right;
            if (cell[1].real) {
                const long long d = key[2] - key[1];
                left = (double)(d < 0 && l.periodic ? d + l.period : d);
            } else {
                left = (double)(2 * key[2]);
            }
            if (cell[3].real) {
                const long long d = key[3] - key[2];
                right = (double)(d < 0 && l.periodic ? d + l.period : d);
            } else {
                right = (double)(2 * (l.period - key[2]));
            }
            const double vol = (left + right) * (0.5 * l.step);
            const double relative = fabs(point[2] - cell[2].W[1]);
            const double inv = 1.0 / vol;
            const double U[3] = {Unew[0] * inv, Unew[1] * inv, Unew[2] * inv};
            double Wn[3];
            primitive<1>(l.gas, U, Wn);
            smallest = courant * vol / (relative + sound_speed<1>(l.gas, Wn));
        }
    }
    block_sum(descent, descents);
    if (fused) block_min(smallest, dt_out);
}

// finite_volume_1d.hpp:86-94 room_at(), wants_cell(): the cells end slot s asks its wall for.
__device__ __forceinline__ int wants_cell(const LineArgs& l, int s, double spacing, int supply_left, int supply_right) {
    int wants = 0;
    if (s == 0 && supply_left) {
        const double vol = volume(l, 0);
        wants += vol > 2 * spacing && (double)l.key[0] * l.step >= 1.5 * spacing;
    }
    if (s == l.n - 1 && supply_right) {
        const double vol = volume(l, s);
        wants += vol > 2 * spacing && (double)(l.period - l.key[s]) * l.step >= 1.5 * spacing;
    }
    return wants;
}

// finite_volume_1d.hpp:422-432 (the "hydro1::ends" kernel of supply_cells()), with the end cells' volumes, keys and cells
// the split needs.
__global__ void k_line_ends(LineArgs l, double spacing, int supply_left, int supply_right, double* out) {
    if (threadIdx.x != 0) return;
    const int last = l.n - 1;
    double W0[3], Wl[3];
    primitive_of(l, 0, W0);
    primitive_of(l, last, Wl);
    out[0] = (double)l.key[0] * l.step;
    out[1] = wants_cell(l, 0, spacing, supply_left, supply_right);
    out[2] = (double)(l.period - l.key[last]) * l.step;
    out[3] = wants_cell(l, last, spacing, supply_left, supply_right);
    out[4] = (supply_left && W0[1] > 0) || (supply_right && Wl[1] < 0);
    out[5] = volume(l, 0);
    out[6] = volume(l, last);
    out[7] = (double)l.key[0];
    out[8] = (double)l.key[last];
    out[9] = cell_of(l, 0);
    out[10] = cell_of(l, last);
}

// finite_volume_1d.hpp:450-462 (the "hydro1::split" kernel): each new cell takes its fraction of the donor's Q.
__global__ void k_line_split(double* q, double* x, LineSplit a, LineSplit b, int count) {
    const int r = threadIdx.x;
    if (r >= count) return;
    const LineSplit& p = r == 0 ? a : b;
    const int i = p.cell, old = p.donor;
#pragma unroll
    for (int v = 0; v < 3; ++v) q[i * 3 + v] = 0;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const double part = p.fraction * q[old * 3 + v];
        q[old * 3 + v] -= part;
        q[i * 3 + v] += part;
    }
    x[i] = p.x;
    *p.key_at = p.key;
    *p.id_at = i;
}

__global__ void k_line_pack(const long long* key, const int* id, int n, unsigned long long* packed) {
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s < n) packed[s] = ((unsigned long long)key[s] << 32) | (unsigned int)id[s];
}
__global__ void k_line_unpack(const unsigned long long* packed, int n, long long* key, int* id) {
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s < n) {
        key[s] = (long long)(packed[s] >> 32);
        id[s] = (int)(packed[s] & 0xffffffffull);
    }
}

// finite_volume_1d.hpp:555-661 Kernel::body() (static; fused: the next dt by LUNA's rule).
__global__ __launch_bounds__(BLOCK) void k_line_step(LineArgs l, double dt, double courant, int fused, unsigned long long* dt_out,
                                                     unsigned long long* dt_next) {
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0 && fused) *dt_next = INF_BITS;
    double smallest = INFINITY;
    if (s < l.n) {
        Cell cell[5];
        double g[4];
        load(l, s, cell, g);
        double grad[5][3];
#pragma unroll
        for (int k = 0; k < 5; ++k) grad[k][0] = grad[k][1] = grad[k][2] = 0;
#pragma unroll
        for (int k = 1; k <= 3; ++k) {
            if (!cell[k].real) continue;
            if (l.order == 1)
#pragma unroll
                for (int v = 0; v < 3; ++v) grad[k][v] = gradient(l, cell + k, g[k - 1], g[k], v);
        }
        double F[2][3];
        const double one[1] = {1.0}, zero[1] = {0.0};
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            const int a = 1 + f, b = 2 + f;
            const double w = 0, shift = 0.5 * dt * w;
            double L[3] = {0, 0, 0}, R[3] = {0, 0, 0};
            if (cell[a].real) face_state(l, cell[a].W, grad[a], 0.5 * g[a] - cell[a].c + shift, dt, L);
            if (cell[b].real) face_state(l, cell[b].W, grad[b], -0.5 * g[a] - cell[b].c + shift, dt, R);
            if (!cell[a].real) ghost_face(l, R, cell[b].W, true, L);
            if (!cell[b].real) ghost_face(l, L, cell[a].W, false, R);
            if (l.riemann == 1)
                hll<1>(l.gas, L, R, one, zero, F[f]);
            else
                hllc<1>(l.gas, L, R, one, zero, F[f]);
        }
        double Unew[3];
        const int i = cell_of(l, s);
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            Unew[v] = l.q[i * 3 + v] - dt * (F[1][v] - F[0][v]);
            l.next[i * 3 + v] = Unew[v];
        }
        if (fused) {
            const double vol = cell[2].V;
            const double inv = 1.0 / vol;
            const double U[3] = {Unew[0] * inv, Unew[1] * inv, Unew[2] * inv};
            double Wn[3];
            primitive<1>(l.gas, U, Wn);
            const double relative = fabs(Wn[1]);
            smallest = courant * vol / (relative + sound_speed<1>(l.gas, Wn));
        }
    }
    if (fused) block_min(smallest, dt_out);
}

// finite_volume_1d.hpp:332-386 time_step() (static): LUNA's rule or AREPO's with the wall ghosts' pair terms.
__global__ __launch_bounds__(BLOCK) void k_line_time_step(LineArgs l, double courant, int luna_rule, unsigned long long* dt_out,
                                                          unsigned long long* dt_next) {
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) *dt_next = INF_BITS;
    double smallest = INFINITY;
    if (s < l.n) {
        double Wi[3];
        primitive_of(l, s, Wi);
        const double ci = sound_speed<1>(l.gas, Wi);
        const double vol = volume(l, s);
        double vi = fabs(Wi[1]);
        if (l.moving) {
            // |w - v| of the point velocity, from the cell's density gradient (:340-356)
            Cell cell[5];
            double g[4];
            load(l, s, cell, g);
            const double grho = l.order == 1 ? gradient(l, cell + 2, g[1], g[2], 0) : 0.0;
            vi = fabs(point_velocity(l, cell + 2, g[1], g[2], grho, vol));
        }
        if (luna_rule) {
            smallest = courant * vol / (vi + ci);
        } else {
            const double r = 0.5 * vol;
            double step = r / (ci + vi);
            for (int side = -1; side <= 1; side += 2) {
                const int t = slot(l, s, side);
                double Wj[3];
                if (t >= 0)
                    primitive_of(l, t, Wj);
                else
                    ghost(l, Wi, side < 0, Wj);
                const double dist = length(l, side < 0 ? s - 1 : s);
                const double vsig = ci + sound_speed<1>(l.gas, Wj) - (Wj[1] - Wi[1]) * side;
                if (vsig > 0) step = fmin(step, (dist + r) / vsig);
            }
            smallest = courant * step;
        }
    }
    block_min(smallest, dt_out);
}


// ---- derivatives of the static step (finite_volume_1d.hpp:710-935), dt fixed: the face-by-face chain rule of k_line_step.

// finite_volume_1d.hpp:234-253 clips(): the last face that clipped variable v under the face-centroid limiter (-1: none),
// to which bound, and whether the clip zeroed the slope.
__device__ __forceinline__ void clips(const Cell* cell, double gapL, double gapR, int v, int& last, bool& upper, bool& zero) {
    const Cell& me = cell[0];
    const double dL = -gapL + cell[-1].c - me.c, dR = gapR + cell[1].c - me.c;
    const double Wi = me.W[v], left = cell[-1].W[v], right = cell[1].W[v];
    double g = ((right - Wi) - (left - Wi)) / (dR - dL);
    const double r[2] = {-0.5 * gapL - me.c, 0.5 * gapR - me.c};
    const double lo = fmin(left, right), hi = fmax(left, right);
    last = -1;
    upper = zero = false;
#pragma unroll
    for (int f = 0; f < 2; ++f) {
        const double dp = g * r[f];
        if (dp > 0 && Wi + dp > hi) {
            g = g * (hi > Wi ? (hi - Wi) / dp : 0.0);
            last = f, upper = true, zero = !(hi > Wi);
        } else if (dp < 0 && Wi + dp < lo) {
            g = g * (lo < Wi ? (lo - Wi) / dp : 0.0);
            last = f, upper = false, zero = !(lo < Wi);
        }
    }
}

// finite_volume_1d.hpp:209-233: d g for the tangents of W[v] of the left neighbour, the cell and the right neighbour; a
// clipped slope is (bound - W_i) / r of its last clip, the bound from the neighbour fmax / fmin returned.
__device__ __forceinline__ double gradient_tangent(const LineArgs& l, const Cell* cell, double gapL, double gapR, int v, double dl, double di,
                                                   double dr) {
    const Cell& me = cell[0];
    const double dL = -gapL + cell[-1].c - me.c, dR = gapR + cell[1].c - me.c;
    const double Wi = me.W[v], left = cell[-1].W[v], right = cell[1].W[v];
    if (l.limiter == 2 || l.limiter == 3) {
        const double sL = (left - Wi) / dL, sR = (right - Wi) / dR, dsL = (dl - di) / dL, dsR = (dr - di) / dR;
        if (!(sL * sR > 0)) return 0.0;
        if (l.limiter == 3) {
            const double sum = sL + sR, g = 2 * sL * sR / sum;
            return (2 * (dsL * sR + sL * dsR) - g * (dsL + dsR)) / sum;
        }
        return fabs(sL) < fabs(sR) ? dsL : dsR;
    }
    const double dg = (dr - dl) / (dR - dL);
    if (l.limiter == 1) return dg;
    int last;
    bool upper, zero;
    clips(cell, gapL, gapR, v, last, upper, zero);
    if (last < 0) return dg;
    if (zero) return 0.0;
    const double r = last == 0 ? -0.5 * gapL - me.c : 0.5 * gapR - me.c;
    const bool from_left = upper ? !(right > left) : !(right < left);
    return ((from_left ? dl : dr) - di) / r;
}

// finite_volume_1d.hpp:255-290: adds J^T gbar of gradient() to the bars of W[v] of the left neighbour, the cell and the
// right neighbour.
__device__ __forceinline__ void gradient_adjoint(const LineArgs& l, const Cell* cell, double gapL, double gapR, int v, double gbar, double& bl,
                                                 double& bi, double& br) {
    const Cell& me = cell[0];
    const double dL = -gapL + cell[-1].c - me.c, dR = gapR + cell[1].c - me.c;
    const double Wi = me.W[v], left = cell[-1].W[v], right = cell[1].W[v];
    if (l.limiter == 2 || l.limiter == 3) {
        const double sL = (left - Wi) / dL, sR = (right - Wi) / dR;
        if (!(sL * sR > 0)) return;
        double aL, aR;  // d g = aL dsL + aR dsR
        if (l.limiter == 3) {
            const double sum = sL + sR, g = 2 * sL * sR / sum;
            aL = (2 * sR - g) / sum, aR = (2 * sL - g) / sum;
        } else {
            const bool pick_left = fabs(sL) < fabs(sR);
            aL = pick_left ? 1.0 : 0.0, aR = pick_left ? 0.0 : 1.0;
        }
        bl += gbar * aL / dL, bi -= gbar * aL / dL;
        br += gbar * aR / dR, bi -= gbar * aR / dR;
        return;
    }
    int last = -1;
    bool upper = false, zero = false;
    if (l.limiter != 1) clips(cell, gapL, gapR, v, last, upper, zero);
    if (last < 0) {
        br += gbar / (dR - dL), bl -= gbar / (dR - dL);
        return;
    }
    if (zero) return;
    const double r = last == 0 ? -0.5 * gapL - me.c : 0.5 * gapR - me.c;
    const bool from_left = upper ? !(right > left) : !(right < left);
    if (from_left)
        bl += gbar / r;
    else
        br += gbar / r;
    bi -= gbar / r;
}

// ghost() as a map of tangents (its own adjoint): outflow copies the end cell, reflective mirrors its velocity.
__device__ __forceinline__ void ghost_tangent(const LineArgs& l, const double* dW, bool left, double* out) {
    const int wall = left ? l.wall_left : l.wall_right;
    out[0] = dW[0];
    out[1] = wall == 0 ? -dW[1] : dW[1];
    out[2] = dW[2];
}
__device__ __forceinline__ void mirror(const double* a, double* out) {
    out[0] = a[0];
    out[1] = -a[1];
    out[2] = a[2];
}

__device__ __forceinline__ void slope(const LineArgs& l, const Cell* cell, const double* gap, int k, double* g) {
    g[0] = g[1] = g[2] = 0;
    if (l.order == 1 && cell[k].real)
#pragma unroll
        for (int v = 0; v < 3; ++v) g[v] = gradient(l, cell + k, gap[k - 1], gap[k], v);
}
__device__ __forceinline__ void slope_tangent(const LineArgs& l, const Cell* cell, const double* gap, const double (*dW)[3], int k, double* dg) {
    dg[0] = dg[1] = dg[2] = 0;
    if (l.order == 1 && cell[k].real)
#pragma unroll
        for (int v = 0; v < 3; ++v) dg[v] = gradient_tangent(l, cell + k, gap[k - 1], gap[k], v, dW[k - 1][v], dW[k][v], dW[k + 1][v]);
}

// finite_volume_1d.hpp:730-743 face_value(): face_state() with its fallback flag.
__device__ __forceinline__ void face_value(const LineArgs& l, const double* W, const double* g, double r, double dt, double* out, bool& fallback) {
    fallback = false;
    if (l.order == 0) {
        out[0] = W[0];
        out[1] = W[1];
        out[2] = W[2];
        return;
    }
    half_step<1>(l.gas, W, g, dt, out);
#pragma unroll
    for (int v = 0; v < 3; ++v) out[v] += g[v] * r;
    fallback = !(out[0] > 0 && out[2] > 0);
    if (fallback) {
        out[0] = W[0];
        out[1] = W[1];
        out[2] = W[2];
    }
}
__device__ __forceinline__ void face_value_tangent(const LineArgs& l, const double* W, const double* dW, const double* g, const double* dg,
                                                   double r, double dt, bool fallback, double* out) {
    if (l.order == 0 || fallback) {
        out[0] = dW[0];
        out[1] = dW[1];
        out[2] = dW[2];
        return;
    }
    half_step_tangent<1>(l.gas, W, dW, g, dg, dt, out);
#pragma unroll
    for (int v = 0; v < 3; ++v) out[v] += dg[v] * r;
}
__device__ __forceinline__ void face_value_adjoint(const LineArgs& l, const double* W, const double* g, double r, double dt, bool fallback,
                                                   const double* Wfbar, double* Wbar, double* gbar) {
    if (l.order == 0 || fallback) {
#pragma unroll
        for (int v = 0; v < 3; ++v) Wbar[v] += Wfbar[v];
        return;
    }
    double gb[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) gb[v] = Wfbar[v] * r;
    half_step_adjoint<1>(l.gas, W, g, dt, Wfbar, Wbar, gb);
#pragma unroll
    for (int v = 0; v < 3; ++v) gbar[v] += gb[v];
}

// finite_volume_1d.hpp:766-792 face_tangent(): the flux of the face between stencil positions a and a + 1 and its tangent.
__device__ __forceinline__ void face_tangent(const LineArgs& l, const Cell* cell, const double* gap, const double (*dW)[3], int a, double dt,
                                             double* F, double* dF) {
    const int b = a + 1;
    double ga[3], gb[3], dga[3], dgb[3];
    slope(l, cell, gap, a, ga);
    slope(l, cell, gap, b, gb);
    slope_tangent(l, cell, gap, dW, a, dga);
    slope_tangent(l, cell, gap, dW, b, dgb);
    const double ra = 0.5 * gap[a] - cell[a].c, rb = -0.5 * gap[a] - cell[b].c;
    double L[3] = {0, 0, 0}, R[3] = {0, 0, 0}, dLs[3] = {0, 0, 0}, dRs[3] = {0, 0, 0};
    bool fallback;
    if (cell[a].real) {
        face_value(l, cell[a].W, ga, ra, dt, L, fallback);
        face_value_tangent(l, cell[a].W, dW[a], ga, dga, ra, dt, fallback, dLs);
    }
    if (cell[b].real) {
        face_value(l, cell[b].W, gb, rb, dt, R, fallback);
        face_value_tangent(l, cell[b].W, dW[b], gb, dgb, rb, dt, fallback, dRs);
    }
    if (!cell[a].real) {
        ghost_face(l, R, cell[b].W, true, L);
        if (l.wall_left == 0)
            mirror(dRs, dLs);
        else
            ghost_tangent(l, dW[b], true, dLs);
    }
    if (!cell[b].real) {
        ghost_face(l, L, cell[a].W, false, R);
        if (l.wall_right == 0)
            mirror(dLs, dRs);
        else
            ghost_tangent(l, dW[a], false, dRs);
    }
    const double one[1] = {1.0}, zero[1] = {0.0};
    riemann_tangent<1>(l.riemann, l.gas, L, R, dLs, dRs, one, zero, F, dF);
}

// finite_volume_1d.hpp:794-830 face_adjoint(): adds J^T Fbar of the face between positions a and a + 1 to the bars of the
// stencil's W (a ghost position's bar is mapped to its wall's end cell by the caller).
__device__ __forceinline__ void face_adjoint(const LineArgs& l, const Cell* cell, const double* gap, int a, double dt, const double* Fbar,
                                             double (*Wbar)[3]) {
    const int b = a + 1;
    double ga[3], gb[3];
    slope(l, cell, gap, a, ga);
    slope(l, cell, gap, b, gb);
    const double ra = 0.5 * gap[a] - cell[a].c, rb = -0.5 * gap[a] - cell[b].c;
    double L[3] = {0, 0, 0}, R[3] = {0, 0, 0};
    bool fallback_a = false, fallback_b = false;
    if (cell[a].real) face_value(l, cell[a].W, ga, ra, dt, L, fallback_a);
    if (cell[b].real) face_value(l, cell[b].W, gb, rb, dt, R, fallback_b);
    if (!cell[a].real) ghost_face(l, R, cell[b].W, true, L);
    if (!cell[b].real) ghost_face(l, L, cell[a].W, false, R);
    double Lbar[3] = {0, 0, 0}, Rbar[3] = {0, 0, 0};
    const double one[1] = {1.0}, zero[1] = {0.0};
    riemann_adjoint<1>(l.riemann, l.gas, L, R, one, zero, Fbar, Lbar, Rbar);
    // the ghost states in reverse order of the forward: R from L, then L from R
    if (!cell[b].real) {
        double back[3];
        if (l.wall_right == 0) {
            mirror(Rbar, back);
#pragma unroll
            for (int v = 0; v < 3; ++v) Lbar[v] += back[v];
        } else {
            ghost_tangent(l, Rbar, false, back);
#pragma unroll
            for (int v = 0; v < 3; ++v) Wbar[a][v] += back[v];
        }
    }
    if (!cell[a].real) {
        double back[3];
        if (l.wall_left == 0) {
            mirror(Lbar, back);
#pragma unroll
            for (int v = 0; v < 3; ++v) Rbar[v] += back[v];
        } else {
            ghost_tangent(l, Lbar, true, back);
#pragma unroll
            for (int v = 0; v < 3; ++v) Wbar[b][v] += back[v];
        }
    }
    double gabar[3] = {0, 0, 0}, gbbar[3] = {0, 0, 0};
    if (cell[a].real) face_value_adjoint(l, cell[a].W, ga, ra, dt, fallback_a, Lbar, Wbar[a], gabar);
    if (cell[b].real) face_value_adjoint(l, cell[b].W, gb, rb, dt, fallback_b, Rbar, Wbar[b], gbbar);
    if (l.order == 1) {
        if (cell[a].real)
#pragma unroll
            for (int v = 0; v < 3; ++v) gradient_adjoint(l, cell + a, gap[a - 1], gap[a], v, gabar[v], Wbar[a - 1][v], Wbar[a][v], Wbar[a + 1][v]);
        if (cell[b].real)
#pragma unroll
            for (int v = 0; v < 3; ++v) gradient_adjoint(l, cell + b, gap[b - 1], gap[b], v, gbbar[v], Wbar[b - 1][v], Wbar[b][v], Wbar[b + 1][v]);
    }
}

// finite_volume_1d.hpp:831-848 load_tangent(): the W tangents of slot s's stencil from dq (ghosts through the ghost map).
__device__ __forceinline__ void load_tangent(const LineArgs& l, const double* dq, int s, double (*dW)[3]) {
    int sl[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        sl[k] = slot(l, s, k - 2);
        dW[k][0] = dW[k][1] = dW[k][2] = 0;
        if (sl[k] < 0) continue;
        const int i = sl[k];
        const double inv = 1.0 / volume(l, i);
        const double U[3] = {l.q[i * 3] * inv, l.q[i * 3 + 1] * inv, l.q[i * 3 + 2] * inv};
        const double dU[3] = {dq[i * 3] * inv, dq[i * 3 + 1] * inv, dq[i * 3 + 2] * inv};
        primitive_tangent<1>(l.gas, U, dU, dW[k]);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (sl[k] >= 0) continue;
        const int end = k < 2 ? k + 1 : k - 1;
        if (sl[end] >= 0) ghost_tangent(l, dW[end], k < 2, dW[k]);
    }
}

// finite_volume_1d.hpp:850-880 step_tangent(): next = the step of q; dnext = J dq (dq: a tangent of the extensive state).
__global__ __launch_bounds__(BLOCK) void k_line_tangent(LineArgs l, double dt, const double* dq, double* dnext) {
    const int s = blockIdx.x * BLOCK + threadIdx.x;
    if (s >= l.n) return;
    Cell cell[5];
    double gap[4], dW[5][3], F0[3], F1[3], dF0[3], dF1[3];
    load(l, s, cell, gap);
    load_tangent(l, dq, s, dW);
    face_tangent(l, cell, gap, dW, 1, dt, F0, dF0);
    face_tangent(l, cell, gap, dW, 2, dt, F1, dF1);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        l.next[s * 3 + v] = l.q[s * 3 + v] - dt * (F1[v] - F0[v]);
        dnext[s * 3 + v] = dq[s * 3 + v] - dt * (dF1[v] - dF0[v]);
    }
}

// finite_volume_1d.hpp:882-933 step_adjoint(): out = J^T qbar with q the state before the step. Per slot j, the adjoints of
// the faces whose fluxes read W_j (the left faces of slots j - 1 .. j + 2, and the right wall's face), each from its own
// stencil, then the primitive adjoint: no atomics.
__global__ __launch_bounds__(BLOCK) void k_line_adjoint(LineArgs l, double dt, const double* qbar, double* out) {
    const int j = blockIdx.x * BLOCK + threadIdx.x;
    const int n = l.n;
    if (j >= n) return;
    double Wj_bar[3] = {0, 0, 0};
    for (int d = -1; d <= 3; ++d) {
        int s, a;
        if (d <= 2) {
            s = slot(l, j, d);
            a = 1;
        } else {
            if (l.periodic || j < n - 2) continue;
            s = n - 1;
            a = 2;  // the right wall's face
        }
        if (s < 0) continue;
        Cell cell[5];
        double gap[4];
        load(l, s, cell, gap);
        const int left = slot(l, s, a - 2), right = slot(l, s, a - 1);
        double Fbar[3];
#pragma unroll
        for (int v = 0; v < 3; ++v) Fbar[v] = (right >= 0 ? dt * qbar[right * 3 + v] : 0.0) - (left >= 0 ? dt * qbar[left * 3 + v] : 0.0);
        double Wbar[5][3];
#pragma unroll
        for (int k = 0; k < 5; ++k) Wbar[k][0] = Wbar[k][1] = Wbar[k][2] = 0;
        face_adjoint(l, cell, gap, a, dt, Fbar, Wbar);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const int t = slot(l, s, k - 2);
            if (t == j) {
#pragma unroll
                for (int v = 0; v < 3; ++v) Wj_bar[v] += Wbar[k][v];
            } else if (t < 0) {
                const int end = k < 2 ? k + 1 : k - 1;
                if (slot(l, s, end - 2) == j) {
                    double back[3];
                    ghost_tangent(l, Wbar[k], k < 2, back);
#pragma unroll
                    for (int v = 0; v < 3; ++v) Wj_bar[v] += back[v];
                }
            }
        }
    }
    const double inv = 1.0 / volume(l, j);
    const double U[3] = {l.q[j * 3] * inv, l.q[j * 3 + 1] * inv, l.q[j * 3 + 2] * inv};
    double Ubar[3];
    primitive_adjoint<1>(l.gas, U, Wj_bar, Ubar);
#pragma unroll
    for (int v = 0; v < 3; ++v) out[j * 3 + v] = qbar[j * 3 + v] + Ubar[v] * inv;
}

}  // namespace

hipError_t launch_line_time_step(hipStream_t s, const LineArgs& a, double courant, int luna_rule, const Counters& c) {
    hipLaunchKernelGGL(k_line_time_step, dim3((a.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a, courant, luna_rule, c.dt, c.dt_next);
    return hipGetLastError();
}

hipError_t launch_line_step(hipStream_t s, const LineArgs& a, double dt, double courant, bool fused, const Counters& c) {
    hipLaunchKernelGGL(k_line_step, dim3((a.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a, dt, courant, fused ? 1 : 0, c.dt, c.dt_next);
    return hipGetLastError();
}

hipError_t launch_line_step_moving(hipStream_t s, const LineArgs& a, double dt, double courant, bool fused, const Counters& c,
                                   unsigned long long* descents, unsigned long long* descents_next) {
    hipLaunchKernelGGL(k_line_step_moving, dim3((a.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a, dt, courant, fused ? 1 : 0, c.dt,
                       c.dt_next, descents, descents_next);
    return hipGetLastError();
}

hipError_t launch_line_ends(hipStream_t s, const LineArgs& a, double spacing, int supply_left, int supply_right, double* out) {
    hipLaunchKernelGGL(k_line_ends, dim3(1), dim3(64), 0, s, a, spacing, supply_left, supply_right, out);
    return hipGetLastError();
}

hipError_t launch_line_split(hipStream_t s, double* q, double* x, const LineSplit* splits, int count) {
    if (count < 1 || count > 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_line_split, dim3(1), dim3(64), 0, s, q, x, splits[0], splits[count - 1], count);
    return hipGetLastError();
}

hipError_t launch_line_sort(hipStream_t s, long long* key, int* id, int n, unsigned long long* packed, void* temp, size_t* temp_bytes) {
    if (!temp) return hipcub::DeviceRadixSort::SortKeys(nullptr, *temp_bytes, packed, packed + n, n, 0, 63, s);
    const dim3 grid((n + BLOCK - 1) / BLOCK);
    hipLaunchKernelGGL(k_line_pack, grid, dim3(BLOCK), 0, s, key, id, n, packed);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipcub::DeviceRadixSort::SortKeys(temp, *temp_bytes, packed, packed + n, n, 0, 63, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_line_unpack, grid, dim3(BLOCK), 0, s, packed + n, n, key, id);
    return hipGetLastError();
}

hipError_t launch_line_tangent(hipStream_t s, const LineArgs& a, double dt, const double* dq, double* dnext) {
    hipLaunchKernelGGL(k_line_tangent, dim3((a.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a, dt, dq, dnext);
    return hipGetLastError();
}

hipError_t launch_line_adjoint(hipStream_t s, const LineArgs& a, double dt, const double* qbar, double* out) {
    hipLaunchKernelGGL(k_line_adjoint, dim3((a.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a, dt, qbar, out);
    return hipGetLastError();
}

}  // namespace luna
