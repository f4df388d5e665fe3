// Pointwise physics of the hydro step for the device (HIP, gfx950): the FP64 arithmetic of luna-cpp's src/hydro/euler.hpp
// and finite_volume.hpp, in their operation order. Every file that includes this is compiled with -ffp-contract=off, so no
// product is fused into a sum and the bits follow the reference's own host build. D = 1, 2, 3; V = D + 2 variables;
// W = (rho, v, P), U = (rho, rho v, E) per volume.
#pragma once

#include <hip/hip_runtime.h>

namespace luna {

struct Gas {
    double gamma, smallr, smallc, inverse_gm1;
};

// euler.hpp:69-73: left to right, written out per D.
template <int D>
__device__ __forceinline__ double dot(const double* a, const double* b) {
    if constexpr (D == 1) return a[0] * b[0];
    if constexpr (D == 2) return a[0] * b[0] + a[1] * b[1];
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

template <int D>
__device__ __forceinline__ float dot(const float* a, const float* b) {
    if constexpr (D == 1) return a[0] * b[0];
    if constexpr (D == 2) return a[0] * b[0] + a[1] * b[1];
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// euler.hpp:104-113 primitive(): U -> W with the RAMSES floors.
template <int D>
__device__ __forceinline__ void primitive(const Gas& gas, const double* U, double* W) {
    const double rho = fmax(U[0], gas.smallr);
    const double inv_rho = 1.0 / rho;
    double v[D];
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = U[1 + k] * inv_rho;
    const double smalle = gas.smallc * gas.smallc / gas.gamma / (gas.gamma - 1.0);
    const double kinetic = 0.5 * rho * dot<D>(v, v);
    const double e = fmax(U[D + 1] - kinetic, smalle * rho);
    W[0] = rho;
#pragma unroll
    for (int k = 0; k < D; ++k) W[1 + k] = v[k];
    W[D + 1] = (gas.gamma - 1.0) * e;
}

// euler.hpp:175-178
template <int D>
__device__ __forceinline__ double sound_speed(const Gas& gas, const double* W) {
    return sqrt(gas.gamma * W[D + 1] / W[0]);
}

// finite_volume.hpp:248-253 face_angle()
template <int D>
__device__ __forceinline__ double face_angle(double area) {
    if constexpr (D == 3)
        return sqrt(area / M_PI);
    else
        return 0.5 * area;
}

// The cube root of v > 0 rounded to the nearest double. The library's cbrt is within an ulp of it; the result is then
// corrected by the sign of (y +- half an ulp)^3 - v: y^2 and y^3 as double-double products (fma gives a product's
// rounding error exactly), y^3 - v exact by Sterbenz, the half-ulp terms 3 y^2 h + 3 y h^2 + h^3 added after. A tie
// cannot occur (a 54-bit midpoint cubed is no double); only |(y + h)^3 - v| below about 2^-100 v could be misjudged.
// The same function, written the same way in C, gives the oracle's radius, so both sides take the same r.
__device__ __forceinline__ double cube_excess(double y, double h, double v) {
    const double p2 = y * y, e2 = fma(y, y, -p2);
    const double p3 = p2 * y, e3 = fma(p2, y, -p3);
    const double r0 = p3 - v;
    const double big = r0 + 3.0 * (p2 * h);
    return big + (e3 + e2 * y) + 3.0 * (e2 * h) + 3.0 * (y * h * h) + h * h * h;
}
__device__ __forceinline__ double cbrt_rn(double v) {
    double y = cbrt(v);
    if (!(v > 0) || !(v < INFINITY)) return y;
    for (int k = 0; k < 4; ++k) {
        const double up = nextafter(y, INFINITY), down = nextafter(y, 0.0);
        if (cube_excess(y, 0.5 * (up - y), v) < 0) {
            y = up;
        } else if (cube_excess(y, -0.5 * (y - down), v) > 0) {
            y = down;
        } else {
            break;
        }
    }
    return y;
}

// finite_volume.hpp:241-246 radius(): cbrt(3V / 4 pi) in 3D, sqrt(V / pi) in 2D.
template <int D>
__device__ __forceinline__ double cell_radius(double v) {
    if constexpr (D == 3)
        return cbrt_rn(v * (3.0 / (4.0 * M_PI)));
    else
        return sqrt(v * (1.0 / M_PI));
}

// euler.hpp:186-214 half_step_rate() and half_step(): W - dt/2 sum_d A_d(W) dW/dx_d. g: [V][D].
template <int D>
__device__ __forceinline__ void half_step(const Gas& gas, const double* W, const double* g, double dt, double* out) {
    constexpr int P = D + 1;
    const double rho = W[0], inv_rho = 1.0 / W[0], rho_c2 = gas.gamma * W[P];
    double A[D + 2];
#pragma unroll
    for (int v = 0; v < D + 2; ++v) A[v] = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double vn = W[1 + d];
        A[0] += vn * g[0 * D + d] + rho * g[(1 + d) * D + d];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double row = vn * g[(1 + k) * D + d];
            if (k == d) row = row + inv_rho * g[P * D + d];
            A[1 + k] += row;
        }
        A[P] += rho_c2 * g[(1 + d) * D + d] + vn * g[P * D + d];
    }
#pragma unroll
    for (int v = 0; v < D + 2; ++v) out[v] = W[v] - 0.5 * dt * A[v];
}

// euler.hpp:186-204 half_step_rate() in the type T of a reduced-precision policy (float): sum_d A_d(W) dW/dx_d, with W
// and g already in T. Used by the reduced-precision face sides.
template <int D, class T>
__device__ __forceinline__ void half_step_rate(const Gas& gas, const T* W, const T* g, T* A) {
    constexpr int P = D + 1;
    const T rho = W[0], inv_rho = T(1) / W[0], rho_c2 = T(gas.gamma) * W[P];
#pragma unroll
    for (int v = 0; v < D + 2; ++v) A[v] = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const T vn = W[1 + d];
        A[0] += vn * g[0 * D + d] + rho * g[(1 + d) * D + d];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            T row = vn * g[(1 + k) * D + d];
            if (k == d) row = row + inv_rho * g[P * D + d];
            A[1 + k] += row;
        }
        A[P] += rho_c2 * g[(1 + d) * D + d] + vn * g[P * D + d];
    }
}

// euler.hpp:316-336 Side and side_with(): one side in the face frame.
template <int D>
struct Side {
    double v[D];
    double vn, c, E;
    double F[D + 2], U[D + 2];
};

template <int D>
__device__ __forceinline__ void side_with(const Gas& gas, const double* W, const double* n, const double* w, double vn, Side<D>& s) {
    constexpr int P = D + 1;
#pragma unroll
    for (int k = 0; k < D; ++k) s.v[k] = W[1 + k] - w[k];
    s.vn = vn;
    s.c = 0;
    s.E = W[P] * gas.inverse_gm1 + W[0] * (0.5 * dot<D>(s.v, s.v));
    const double m = W[0] * s.vn;
    s.F[0] = m;
#pragma unroll
    for (int k = 0; k < D; ++k) s.F[1 + k] = m * s.v[k] + W[P] * n[k];
    s.F[P] = (s.E + W[P]) * s.vn;
    s.U[0] = W[0];
#pragma unroll
    for (int k = 0; k < D; ++k) s.U[1 + k] = W[0] * s.v[k];
    s.U[P] = s.E;
}

// euler.hpp:340-346 side(): with the sound speed.
template <int D>
__device__ __forceinline__ void side_full(const Gas& gas, const double* W, const double* n, const double* w, Side<D>& s) {
    constexpr int P = D + 1;
    double rel[D];
#pragma unroll
    for (int k = 0; k < D; ++k) rel[k] = W[1 + k] - w[k];
    side_with<D>(gas, W, n, w, dot<D>(rel, n), s);
    const double c2 = gas.gamma * W[P] / W[0];
    s.c = c2 > 0 ? sqrt(c2) : 0.0;
}

// euler.hpp:350-356 to_lab(): face frame -> lab frame.
template <int D>
__device__ __forceinline__ void to_lab(double* F, const double* w) {
    if constexpr (D == 1) F[2] += w[0] * F[1] + 0.5 * dot<1>(w, w) * F[0];
    if constexpr (D == 2) F[3] += w[0] * F[1] + w[1] * F[2] + 0.5 * dot<2>(w, w) * F[0];
    if constexpr (D == 3) F[4] += w[0] * F[1] + w[1] * F[2] + w[2] * F[3] + 0.5 * dot<3>(w, w) * F[0];
#pragma unroll
    for (int k = 0; k < D; ++k) F[1 + k] += w[k] * F[0];
}

// euler.hpp:444-461 HLL::flux()
template <int D>
__device__ __forceinline__ void hll(const Gas& gas, const double* L, const double* R, const double* n, const double* w, double* F) {
    Side<D> l, r;
    side_full<D>(gas, L, n, w, l);
    side_full<D>(gas, R, n, w, r);
    const double SL = fmin(l.vn - l.c, r.vn - r.c);
    const double SR = fmax(l.vn + l.c, r.vn + r.c);
    if (SL >= 0.0) {
#pragma unroll
        for (int v = 0; v < D + 2; ++v) F[v] = l.F[v];
    } else if (SR <= 0.0) {
#pragma unroll
        for (int v = 0; v < D + 2; ++v) F[v] = r.F[v];
    } else {
        const double inv_dS = 1.0 / (SR - SL);
#pragma unroll
        for (int v = 0; v < D + 2; ++v) F[v] = (SR * l.F[v] - SL * r.F[v] + SL * SR * (r.U[v] - l.U[v])) * inv_dS;
    }
    to_lab<D>(F, w);
}

// euler.hpp:547-550 HLLC::nonzero()
__device__ __forceinline__ double nonzero(double d) {
    const double eps = 1e-20;
    return fabs(d) < eps ? (d < 0 ? -eps : eps) : d;
}

// euler.hpp:609-642 HLLC::flux(): Davis wave speeds, LUNA's branch order F_L, F*_L, F*_R, F_R and guards.
template <int D>
__device__ __forceinline__ void hllc(const Gas& gas, const double* L, const double* R, const double* n, const double* w, double* F) {
    constexpr int P = D + 1;
    double rl[D], rr[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        rl[k] = L[1 + k] - w[k];
        rr[k] = R[1 + k] - w[k];
    }
    const double vnL = dot<D>(rl, n), vnR = dot<D>(rr, n);
    const double inv_rhoL = 1.0 / L[0], inv_rhoR = 1.0 / R[0];
    const double c2L = gas.gamma * L[P] * inv_rhoL, c2R = gas.gamma * R[P] * inv_rhoR;
    const bool ok = c2L >= 0 && c2R >= 0;
    const double cL = c2L > 0 ? sqrt(c2L) : 0.0, cR = c2R > 0 ? sqrt(c2R) : 0.0;
    const double SL = fmin(vnL - cL, vnR - cR);
    const double SR = fmax(vnL + cL, vnR + cR);
    const double denomL = L[0] * (SL - vnL), denomR = R[0] * (SR - vnR);
    const double SM = (R[P] - L[P] + denomL * vnL - denomR * vnR) / nonzero(denomL - denomR);
    const bool left = ok && (SL >= 0 || SM >= 0);
    const bool star = ok && SL < 0 && (SM >= 0 || SR > 0);
    double W[D + 2];
#pragma unroll
    for (int v = 0; v < D + 2; ++v) W[v] = left ? L[v] : R[v];
    Side<D> s;
    side_with<D>(gas, W, n, w, left ? vnL : vnR, s);
    if (!star) {
#pragma unroll
        for (int v = 0; v < D + 2; ++v) F[v] = s.F[v];
        to_lab<D>(F, w);
        return;
    }
    const double S = left ? SL : SR, denom = left ? denomL : denomR;
    const double rho_star = denom / nonzero(S - SM);
    const double E_star = s.E * (left ? inv_rhoL : inv_rhoR) + (SM - s.vn) * (SM + W[P] * (1.0 / nonzero(denom)));
    double U_star[D + 2];
    U_star[0] = rho_star;
#pragma unroll
    for (int k = 0; k < D; ++k) U_star[1 + k] = rho_star * (s.v[k] + n[k] * (SM - s.vn));
    U_star[P] = rho_star * E_star;
#pragma unroll
    for (int v = 0; v < D + 2; ++v) F[v] = s.F[v] + S * (U_star[v] - s.U[v]);
    to_lab<D>(F, w);
}

template <int D>
__device__ __forceinline__ void riemann(int kind, const Gas& gas, const double* L, const double* R, const double* n, const double* w, double* F) {
    if (kind == 1)
        hll<D>(gas, L, R, n, w, F);
    else
        hllc<D>(gas, L, R, n, w, F);
}

// The nearest periodic image of a separation, given the box side and its inverse (the role of tet_mesh::wrap(d, side,
// inverse) in Geometry::separation, finite_volume.hpp:279-282).
__device__ __forceinline__ double wrap(double d, double side, double inverse) { return d - side * round(d * inverse); }

}  // namespace luna
