// The static 1D step on a sorted line (finite_volume_1d.hpp), one work-item per slot: the five-cell stencil in
// registers, three limited gradients, two face states, two Riemann fluxes and the new state in one kernel; fused, the
// next time step by LUNA's rule from the new state as a block min-reduction. FP64, -ffp-contract=off.
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
__device__ __forceinline__ int slot(const LineArgs& l, int s, int k) {
    const int t = s + k;
    if (l.periodic) return ((t % l.n) + l.n) % l.n;
    return t >= 0 && t < l.n ? t : -1;
}

// finite_volume_1d.hpp:107-111 primitive_of()
__device__ __forceinline__ void primitive_of(const LineArgs& l, int s, double* W) {
    const double inv = 1.0 / volume(l, s);
    const double U[3] = {l.q[s * 3] * inv, l.q[s * 3 + 1] * inv, l.q[s * 3 + 2] * inv};
    primitive<1>(l.gas, U, W);
}

// finite_volume_1d.hpp:117-128 ghost(), ghost_face(): reflective mirrors the velocity, outflow copies.
__device__ __forceinline__ void ghost(const LineArgs& l, const double* W, bool left, double* out) {
    const int wall = left ? l.wall_left : l.wall_right;
    out[0] = W[0];
    out[1] = wall == 0 ? -W[1] : W[1];
    out[2] = W[2];
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
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            Unew[v] = l.q[s * 3 + v] - dt * (F[1][v] - F[0][v]);
            l.next[s * 3 + v] = Unew[v];
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
        const double vi = fabs(Wi[1]);
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

}  // namespace

hipError_t launch_line_time_step(hipStream_t s, const LineArgs& a, double courant, int luna_rule, const Counters& c) {
    hipLaunchKernelGGL(k_line_time_step, dim3((a.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a, courant, luna_rule, c.dt, c.dt_next);
    return hipGetLastError();
}

hipError_t launch_line_step(hipStream_t s, const LineArgs& a, double dt, double courant, bool fused, const Counters& c) {
    hipLaunchKernelGGL(k_line_step, dim3((a.n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, a, dt, courant, fused ? 1 : 0, c.dt, c.dt_next);
    return hipGetLastError();
}

}  // namespace luna
