// Derivatives of the pointwise physics for the device: tangents (J dx) and adjoints (J^T ybar, added into the bars) of
// primitive(), half_step(), side_with() / side_full(), to_lab() and the HLLC and HLL fluxes, at fixed dt, face normal and face velocity. At every
// switch of the forward (a floor, a wave case, a 1e-20 guard) the derivative is that of the branch the forward took, as
// luna-cpp's own derivatives are defined (src/hydro/euler.hpp:115-117); written from the forward arithmetic of physics.hpp
// by the chain rule. -ffp-contract=off like the rest.
#pragma once

#include "physics.hpp"

namespace luna {

// euler.hpp:119-126: which floors primitive() applied (bit 0 density, bit 1 internal energy).
template <int D>
__device__ __forceinline__ int primitive_floors(const Gas& gas, const double* U) {
    const double rho = fmax(U[0], gas.smallr);
    const double inv_rho = 1.0 / rho;
    double v[D];
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = U[1 + k] * inv_rho;
    const double smalle = gas.smallc * gas.smallc / gas.gamma / (gas.gamma - 1.0);
    return (U[0] >= gas.smallr ? 0 : 1) | (U[D + 1] - 0.5 * rho * dot<D>(v, v) >= smalle * rho ? 0 : 2);
}

// euler.hpp:128-140: a floored density or energy has derivative zero in what the floor replaced.
template <int D>
__device__ __forceinline__ void primitive_tangent(const Gas& gas, const double* U, const double* dU, double* dW) {
    const int floors = primitive_floors<D>(gas, U);
    const double rho = fmax(U[0], gas.smallr), inv_rho = 1.0 / rho;
    const double drho = (floors & 1) ? 0.0 : dU[0];
    double v[D], dv[D];
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = U[1 + k] * inv_rho;
#pragma unroll
    for (int k = 0; k < D; ++k) dv[k] = (dU[1 + k] - v[k] * drho) * inv_rho;
    const double smalle = gas.smallc * gas.smallc / gas.gamma / (gas.gamma - 1.0);
    const double dkinetic = 0.5 * drho * dot<D>(v, v) + rho * dot<D>(v, dv);
    const double de = (floors & 2) ? smalle * drho : dU[D + 1] - dkinetic;
    dW[0] = drho;
#pragma unroll
    for (int k = 0; k < D; ++k) dW[1 + k] = dv[k];
    dW[D + 1] = (gas.gamma - 1.0) * de;
}

// euler.hpp:142-162: Ubar = J^T Wbar.
template <int D>
__device__ __forceinline__ void primitive_adjoint(const Gas& gas, const double* U, const double* Wbar, double* Ubar) {
    const int floors = primitive_floors<D>(gas, U);
    const double rho = fmax(U[0], gas.smallr), inv_rho = 1.0 / rho;
    double v[D], vbar[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        v[k] = U[1 + k] * inv_rho;
        vbar[k] = Wbar[1 + k];
    }
    const double smalle = gas.smallc * gas.smallc / gas.gamma / (gas.gamma - 1.0);
    const double ebar = (gas.gamma - 1.0) * Wbar[D + 1];
    double rhobar = Wbar[0], energybar = 0.0;
    if (floors & 2) {
        rhobar += smalle * ebar;  // e = smalle rho
    } else {                      // e = E - rho v.v / 2
        energybar = ebar;
        rhobar -= 0.5 * dot<D>(v, v) * ebar;
#pragma unroll
        for (int k = 0; k < D; ++k) vbar[k] = vbar[k] - v[k] * (rho * ebar);
    }
    rhobar -= dot<D>(v, vbar) * inv_rho;  // v = m / rho
    Ubar[0] = (floors & 1) ? 0.0 : rhobar;
#pragma unroll
    for (int k = 0; k < D; ++k) Ubar[1 + k] = vbar[k] * inv_rho;
    Ubar[D + 1] = energybar;
}

// euler.hpp:216-237: d(half_step) for dW, dg. g, dg: [V][D].
template <int D>
__device__ __forceinline__ void half_step_tangent(const Gas& gas, const double* W, const double* dW, const double* g, const double* dg,
                                                  double dt, double* out) {
    constexpr int P = D + 1;
    const double rho = W[0], inv_rho = 1.0 / W[0];
    const double rho_c2 = gas.gamma * W[P];
    const double drho = dW[0], dinv_rho = -inv_rho * inv_rho * dW[0], drho_c2 = gas.gamma * dW[P];
    double dA[D + 2];
#pragma unroll
    for (int v = 0; v < D + 2; ++v) dA[v] = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double vn = W[1 + d], dvn = dW[1 + d];
        dA[0] += dvn * g[0 * D + d] + vn * dg[0 * D + d] + drho * g[(1 + d) * D + d] + rho * dg[(1 + d) * D + d];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            double row = dvn * g[(1 + k) * D + d] + vn * dg[(1 + k) * D + d];
            if (k == d) row += dinv_rho * g[P * D + d] + inv_rho * dg[P * D + d];
            dA[1 + k] += row;
        }
        dA[P] += drho_c2 * g[(1 + d) * D + d] + rho_c2 * dg[(1 + d) * D + d] + dvn * g[P * D + d] + vn * dg[P * D + d];
    }
#pragma unroll
    for (int v = 0; v < D + 2; ++v) out[v] = dW[v] - 0.5 * dt * dA[v];
}

// euler.hpp:239-272: adds J^T hbar of half_step() to Wbar and gbar.
template <int D>
__device__ __forceinline__ void half_step_adjoint(const Gas& gas, const double* W, const double* g, double dt, const double* hbar,
                                                  double* Wbar, double* gbar) {
    constexpr int P = D + 1;
    const double rho = W[0], inv_rho = 1.0 / W[0], rho_c2 = gas.gamma * W[P];
    double Abar[D + 2];
#pragma unroll
    for (int v = 0; v < D + 2; ++v) {
        Wbar[v] += hbar[v];
        Abar[v] = -0.5 * dt * hbar[v];
    }
    double inv_rhobar = 0.0, rho_c2bar = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double vn = W[1 + d];
        Wbar[1 + d] += Abar[0] * g[0 * D + d];
        gbar[0 * D + d] += Abar[0] * vn;
        Wbar[0] += Abar[0] * g[(1 + d) * D + d];
        gbar[(1 + d) * D + d] += Abar[0] * rho;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            Wbar[1 + d] += Abar[1 + k] * g[(1 + k) * D + d];
            gbar[(1 + k) * D + d] += Abar[1 + k] * vn;
            if (k == d) {
                inv_rhobar += Abar[1 + k] * g[P * D + d];
                gbar[P * D + d] += Abar[1 + k] * inv_rho;
            }
        }
        rho_c2bar += Abar[P] * g[(1 + d) * D + d];
        gbar[(1 + d) * D + d] += Abar[P] * rho_c2;
        Wbar[1 + d] += Abar[P] * g[P * D + d];
        gbar[P * D + d] += Abar[P] * vn;
    }
    Wbar[0] -= inv_rho * inv_rho * inv_rhobar;
    Wbar[P] += gas.gamma * rho_c2bar;
}

// euler.hpp:358-375: the tangent of side_with() with vn's tangent given (w fixed).
template <int D>
__device__ __forceinline__ void side_with_tangent(const Gas& gas, const double* W, const double* dW, const double* n, const Side<D>& s,
                                                  double dvn, Side<D>& d) {
    constexpr int P = D + 1;
#pragma unroll
    for (int k = 0; k < D; ++k) d.v[k] = dW[1 + k];
    d.vn = dvn;
    d.c = 0;
    d.E = dW[P] * gas.inverse_gm1 + dW[0] * (0.5 * dot<D>(s.v, s.v)) + W[0] * dot<D>(s.v, d.v);
    const double m = W[0] * s.vn, dm = dW[0] * s.vn + W[0] * dvn;
    d.F[0] = dm;
#pragma unroll
    for (int k = 0; k < D; ++k) d.F[1 + k] = dm * s.v[k] + m * d.v[k] + dW[P] * n[k];
    d.F[P] = (d.E + dW[P]) * s.vn + (s.E + W[P]) * dvn;
    d.U[0] = dW[0];
#pragma unroll
    for (int k = 0; k < D; ++k) d.U[1 + k] = dW[0] * s.v[k] + W[0] * d.v[k];
    d.U[P] = d.E;
}

// euler.hpp:387-411: the adjoint of side_with() from the bars of F, U, E and v; adds J^T to Wbar and returns vn's bar.
template <int D>
__device__ __forceinline__ double side_with_adjoint(const Gas& gas, const double* W, const double* n, const Side<D>& s, const double* Fbar,
                                                    const double* Ubar, double Ebar, double* vbar, double* Wbar) {
    constexpr int P = D + 1;
    const double m = W[0] * s.vn;
    Ebar += Fbar[P] * s.vn + Ubar[P];
    double Pbar = Fbar[P] * s.vn, vnbar = Fbar[P] * (s.E + W[P]), mbar = Fbar[0], rhobar = Ubar[0];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        Pbar += Fbar[1 + k] * n[k];
        mbar += Fbar[1 + k] * s.v[k];
        vbar[k] += Fbar[1 + k] * m + Ubar[1 + k] * W[0];
        rhobar += Ubar[1 + k] * s.v[k];
    }
    rhobar += mbar * s.vn;
    vnbar += mbar * W[0];
    Pbar += Ebar * gas.inverse_gm1;
    rhobar += Ebar * (0.5 * dot<D>(s.v, s.v));
#pragma unroll
    for (int k = 0; k < D; ++k) vbar[k] = vbar[k] + s.v[k] * (Ebar * W[0]);
    Wbar[0] += rhobar;
#pragma unroll
    for (int k = 0; k < D; ++k) Wbar[1 + k] += vbar[k];
    Wbar[P] += Pbar;
    return vnbar;
}

// euler.hpp:428-437: the adjoint of to_lab() (linear in F, w fixed).
template <int D>
__device__ __forceinline__ void to_lab_adjoint(const double* Fbar, const double* w, double* G) {
    constexpr int P = D + 1;
#pragma unroll
    for (int v = 0; v < D + 2; ++v) G[v] = Fbar[v];
    G[0] += 0.5 * dot<D>(w, w) * Fbar[P];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        G[0] += w[k] * Fbar[1 + k];
        G[1 + k] += w[k] * Fbar[P];
    }
}

// euler.hpp:649-695: HLLC's wave speeds as hllc() computes them, and the branch pattern: bit 0 S_L from L, bit 1 S_R from L,
// bit 2 left, bit 3 star, bit 4 ok, bits 5-6 c^2 > 0 on L, R, bits 7-9 the 1e-20 guard on S_M's, rho*'s and E*'s divisor.
struct Waves {
    double vnL, vnR, inv_rhoL, inv_rhoR, c2L, c2R, cL, cR, SL, SR, denomL, denomR, den, SM;
    int pattern;
};
template <int D>
__device__ __forceinline__ void waves(const Gas& gas, const double* L, const double* R, const double* n, const double* w, Waves& a) {
    constexpr int P = D + 1;
    const double eps = 1e-20;
    double rl[D], rr[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        rl[k] = L[1 + k] - w[k];
        rr[k] = R[1 + k] - w[k];
    }
    a.vnL = dot<D>(rl, n);
    a.vnR = dot<D>(rr, n);
    a.inv_rhoL = 1.0 / L[0];
    a.inv_rhoR = 1.0 / R[0];
    a.c2L = gas.gamma * L[P] * a.inv_rhoL;
    a.c2R = gas.gamma * R[P] * a.inv_rhoR;
    const int ok = a.c2L >= 0 && a.c2R >= 0;
    a.cL = a.c2L > 0 ? sqrt(a.c2L) : 0.0;
    a.cR = a.c2R > 0 ? sqrt(a.c2R) : 0.0;
    const double sl = a.vnL - a.cL, sr = a.vnR - a.cR, tl = a.vnL + a.cL, tr = a.vnR + a.cR;
    a.SL = fmin(sl, sr);
    a.SR = fmax(tl, tr);
    a.denomL = L[0] * (a.SL - a.vnL);
    a.denomR = R[0] * (a.SR - a.vnR);
    a.den = a.denomL - a.denomR;
    a.SM = (R[P] - L[P] + a.denomL * a.vnL - a.denomR * a.vnR) / nonzero(a.den);
    const int left = ok && (a.SL >= 0 || a.SM >= 0);
    const int star = ok && a.SL < 0 && (a.SM >= 0 || a.SR > 0);
    const double S = left ? a.SL : a.SR, denom = left ? a.denomL : a.denomR;
    a.pattern = (sr < sl ? 0 : 1) | (tr > tl ? 0 : 2) | left << 2 | star << 3 | ok << 4 | (a.c2L > 0) << 5 | (a.c2R > 0) << 6 |
                (fabs(a.den) < eps) << 7 | (star && fabs(S - a.SM) < eps) << 8 | (star && fabs(denom) < eps) << 9;
}

// euler.hpp:697-745: dF = J (dL, dR) of hllc() at fixed n, w (a guarded divisor has derivative zero); F: the flux from the
// same intermediates.
template <int D>
__device__ __forceinline__ void hllc_tangent(const Gas& gas, const double* L, const double* R, const double* dL, const double* dR,
                                             const double* n, const double* w, double* F, double* dF) {
    constexpr int P = D + 1;
    Waves a;
    waves<D>(gas, L, R, n, w, a);
    const bool left = (a.pattern >> 2) & 1, star = (a.pattern >> 3) & 1;
    double dvl[D], dvr[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        dvl[k] = dL[1 + k];
        dvr[k] = dR[1 + k];
    }
    const double dvnL = dot<D>(dvl, n), dvnR = dot<D>(dvr, n);
    const double dinvL = -a.inv_rhoL * a.inv_rhoL * dL[0], dinvR = -a.inv_rhoR * a.inv_rhoR * dR[0];
    const double dcL = a.c2L > 0 ? gas.gamma * (dL[P] * a.inv_rhoL + L[P] * dinvL) / (2 * a.cL) : 0.0;
    const double dcR = a.c2R > 0 ? gas.gamma * (dR[P] * a.inv_rhoR + R[P] * dinvR) / (2 * a.cR) : 0.0;
    const double dSL = (a.pattern & 1) ? dvnL - dcL : dvnR - dcR, dSR = (a.pattern & 2) ? dvnL + dcL : dvnR + dcR;
    const double ddenomL = dL[0] * (a.SL - a.vnL) + L[0] * (dSL - dvnL), ddenomR = dR[0] * (a.SR - a.vnR) + R[0] * (dSR - dvnR);
    const double dnum = dR[P] - dL[P] + ddenomL * a.vnL + a.denomL * dvnL - ddenomR * a.vnR - a.denomR * dvnR;
    const double dden = (a.pattern & 128) ? 0.0 : ddenomL - ddenomR, den = nonzero(a.den);
    const double dSM = (dnum - a.SM * dden) / den;
    double W[D + 2], dW[D + 2];
#pragma unroll
    for (int v = 0; v < D + 2; ++v) {
        W[v] = left ? L[v] : R[v];
        dW[v] = left ? dL[v] : dR[v];
    }
    const double vn = left ? a.vnL : a.vnR, dvn = left ? dvnL : dvnR;
    Side<D> s, ds;
    side_with<D>(gas, W, n, w, vn, s);
    side_with_tangent<D>(gas, W, dW, n, s, dvn, ds);
    if (!star) {
#pragma unroll
        for (int v = 0; v < D + 2; ++v) {
            F[v] = s.F[v];
            dF[v] = ds.F[v];
        }
        to_lab<D>(F, w);
        to_lab<D>(dF, w);
        return;
    }
    const double S = left ? a.SL : a.SR, dS = left ? dSL : dSR;
    const double denom = left ? a.denomL : a.denomR, ddenom = left ? ddenomL : ddenomR;
    const double inv_side = left ? a.inv_rhoL : a.inv_rhoR, dinv_side = left ? dinvL : dinvR;
    const double q = nonzero(S - a.SM), dq = (a.pattern & 256) ? 0.0 : dS - dSM;
    const double rho_star = denom / q, drho_star = (ddenom - rho_star * dq) / q;
    const double inv_denom = 1.0 / nonzero(denom), dinv_denom = (a.pattern & 512) ? 0.0 : -inv_denom * inv_denom * ddenom;
    const double a2 = a.SM - s.vn, da2 = dSM - dvn;
    const double a3 = a.SM + W[P] * inv_denom, da3 = dSM + dW[P] * inv_denom + W[P] * dinv_denom;
    const double E_star = s.E * inv_side + a2 * a3;
    const double dE_star = ds.E * inv_side + s.E * dinv_side + da2 * a3 + a2 * da3;
#pragma unroll
    for (int v = 0; v < D + 2; ++v) {
        double U_star, dU_star;
        if (v == 0) {
            U_star = rho_star;
            dU_star = drho_star;
        } else if (v == P) {
            U_star = rho_star * E_star;
            dU_star = drho_star * E_star + rho_star * dE_star;
        } else {
            const double vs = s.v[v - 1] + n[v - 1] * a2, dvs = ds.v[v - 1] + n[v - 1] * da2;
            U_star = rho_star * vs;
            dU_star = drho_star * vs + rho_star * dvs;
        }
        F[v] = s.F[v] + S * (U_star - s.U[v]);
        dF[v] = ds.F[v] + dS * (U_star - s.U[v]) + S * (dU_star - ds.U[v]);
    }
    to_lab<D>(F, w);
    to_lab<D>(dF, w);
}

// euler.hpp:746-832: adds J^T Fbar of hllc() to Lbar and Rbar; only the case the forward took is differentiated.
template <int D>
__device__ __forceinline__ void hllc_adjoint(const Gas& gas, const double* L, const double* R, const double* n, const double* w,
                                             const double* Fbar_lab, double* Lbar, double* Rbar) {
    constexpr int P = D + 1;
    Waves a;
    waves<D>(gas, L, R, n, w, a);
    const bool left = (a.pattern >> 2) & 1, star = (a.pattern >> 3) & 1;
    double Fbar[D + 2];
    to_lab_adjoint<D>(Fbar_lab, w, Fbar);
    double W[D + 2], Wbar[D + 2];
#pragma unroll
    for (int v = 0; v < D + 2; ++v) {
        W[v] = left ? L[v] : R[v];
        Wbar[v] = 0;
    }
    const double vn = left ? a.vnL : a.vnR;
    Side<D> s;
    side_with<D>(gas, W, n, w, vn, s);
    double vnLbar = 0, vnRbar = 0;
    if (!star) {
        double Ubar0[D + 2], vbar[D];
#pragma unroll
        for (int v = 0; v < D + 2; ++v) Ubar0[v] = 0;
#pragma unroll
        for (int k = 0; k < D; ++k) vbar[k] = 0;
        const double vnbar = side_with_adjoint<D>(gas, W, n, s, Fbar, Ubar0, 0.0, vbar, Wbar);
        if (left)
            vnLbar += vnbar;
        else
            vnRbar += vnbar;
    } else {
        const double S = left ? a.SL : a.SR, denom = left ? a.denomL : a.denomR, inv_side = left ? a.inv_rhoL : a.inv_rhoR;
        const double q = nonzero(S - a.SM), rho_star = denom / q;
        const double inv_denom = 1.0 / nonzero(denom);
        const double a2 = a.SM - s.vn, a3 = a.SM + W[P] * inv_denom;
        const double E_star = s.E * inv_side + a2 * a3;
        double U_star[D + 2], vs[D];
        U_star[0] = rho_star;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            vs[k] = s.v[k] + n[k] * a2;
            U_star[1 + k] = rho_star * vs[k];
        }
        U_star[P] = rho_star * E_star;
        // F = s.F + S (U* - s.U)
        double sUbar[D + 2], Ustarbar[D + 2], Sbar = 0;
#pragma unroll
        for (int v = 0; v < D + 2; ++v) {
            Sbar += Fbar[v] * (U_star[v] - s.U[v]);
            Ustarbar[v] = S * Fbar[v];
            sUbar[v] = -S * Fbar[v];
        }
        // U* = (rho*, rho* v*, rho* E*)
        double rho_starbar = Ustarbar[0] + Ustarbar[P] * E_star, svbar[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            rho_starbar += Ustarbar[1 + k] * vs[k];
            svbar[k] = Ustarbar[1 + k] * rho_star;
        }
        const double E_starbar = Ustarbar[P] * rho_star;
        // v* = s.v + n a2; E* = s.E inv_side + a2 a3
        double a2bar = E_starbar * a3;
#pragma unroll
        for (int k = 0; k < D; ++k) a2bar += svbar[k] * n[k];
        const double sEbar = E_starbar * inv_side, inv_sidebar = E_starbar * s.E, a3bar = E_starbar * a2;
        // a3 = SM + W_P inv_denom; a2 = SM - s.vn
        double SMbar = a3bar + a2bar;
        Wbar[P] += a3bar * inv_denom;
        const double inv_denombar = a3bar * W[P];
        // inv_denom = 1 / nonzero(denom); rho* = denom / q; q = nonzero(S - SM)
        double denombar = (a.pattern & 512) ? 0.0 : -inv_denom * inv_denom * inv_denombar;
        denombar += rho_starbar / q;
        const double qbar = -rho_starbar * rho_star / q;
        if (!(a.pattern & 256)) {
            Sbar += qbar;
            SMbar -= qbar;
        }
        const double vnbar = side_with_adjoint<D>(gas, W, n, s, Fbar, sUbar, sEbar, svbar, Wbar) - a2bar;
        if (left)
            vnLbar += vnbar;
        else
            vnRbar += vnbar;
        double inv_rhoLbar = left ? inv_sidebar : 0.0, inv_rhoRbar = left ? 0.0 : inv_sidebar;
        double SLbar = left ? Sbar : 0.0, SRbar = left ? 0.0 : Sbar;
        double denomLbar = left ? denombar : 0.0, denomRbar = left ? 0.0 : denombar;
        // SM = num / nonzero(den), num = R_P - L_P + denomL vnL - denomR vnR, den = denomL - denomR
        const double den = nonzero(a.den);
        const double numbar = SMbar / den;
        const double denbar = (a.pattern & 128) ? 0.0 : -SMbar * a.SM / den;
        Rbar[P] += numbar;
        Lbar[P] -= numbar;
        denomLbar += numbar * a.vnL + denbar;
        vnLbar += numbar * a.denomL;
        denomRbar += -numbar * a.vnR - denbar;
        vnRbar -= numbar * a.denomR;
        // denomL = L0 (SL - vnL), denomR = R0 (SR - vnR)
        Lbar[0] += denomLbar * (a.SL - a.vnL);
        SLbar += denomLbar * L[0];
        vnLbar -= denomLbar * L[0];
        Rbar[0] += denomRbar * (a.SR - a.vnR);
        SRbar += denomRbar * R[0];
        vnRbar -= denomRbar * R[0];
        // S_L = min(vnL - cL, vnR - cR), S_R = max(vnL + cL, vnR + cR): the side each took
        double cLbar = 0, cRbar = 0;
        if (a.pattern & 1) {
            vnLbar += SLbar;
            cLbar -= SLbar;
        } else {
            vnRbar += SLbar;
            cRbar -= SLbar;
        }
        if (a.pattern & 2) {
            vnLbar += SRbar;
            cLbar += SRbar;
        } else {
            vnRbar += SRbar;
            cRbar += SRbar;
        }
        // c = sqrt(gamma P inv_rho) where c^2 > 0
        if (a.c2L > 0) {
            const double c2bar = cLbar / (2 * a.cL);
            Lbar[P] += c2bar * gas.gamma * a.inv_rhoL;
            inv_rhoLbar += c2bar * gas.gamma * L[P];
        }
        if (a.c2R > 0) {
            const double c2bar = cRbar / (2 * a.cR);
            Rbar[P] += c2bar * gas.gamma * a.inv_rhoR;
            inv_rhoRbar += c2bar * gas.gamma * R[P];
        }
        Lbar[0] -= inv_rhoLbar * a.inv_rhoL * a.inv_rhoL;
        Rbar[0] -= inv_rhoRbar * a.inv_rhoR * a.inv_rhoR;
    }
#pragma unroll
    for (int v = 0; v < D + 2; ++v) {
        if (left)
            Lbar[v] += Wbar[v];
        else
            Rbar[v] += Wbar[v];
    }
    // vn = (v - w) . n
#pragma unroll
    for (int k = 0; k < D; ++k) {
        Lbar[1 + k] += vnLbar * n[k];
        Rbar[1 + k] += vnRbar * n[k];
    }
}

// ---- HLL (euler.hpp:444-540) ----

// euler.hpp:377-385: the tangent of side_full(), the sound speed's included (zero where c^2 <= 0, as the forward).
template <int D>
__device__ __forceinline__ void side_full_tangent(const Gas& gas, const double* W, const double* dW, const double* n, const Side<D>& s, Side<D>& d) {
    constexpr int P = D + 1;
    double dv[D];
#pragma unroll
    for (int k = 0; k < D; ++k) dv[k] = dW[1 + k];
    side_with_tangent<D>(gas, W, dW, n, s, dot<D>(dv, n), d);
    const double c2 = gas.gamma * W[P] / W[0];
    d.c = c2 > 0 ? (gas.gamma * dW[P] - c2 * dW[0]) / W[0] / (2 * s.c) : 0.0;
}

// euler.hpp:413-426: the adjoint of side_full() from the bars of F, U, vn and c; adds J^T to Wbar.
template <int D>
__device__ __forceinline__ void side_full_adjoint(const Gas& gas, const double* W, const double* n, const Side<D>& s, const double* Fbar,
                                                  const double* Ubar, double vnbar, double cbar, double* Wbar) {
    constexpr int P = D + 1;
    double vbar[D];
#pragma unroll
    for (int k = 0; k < D; ++k) vbar[k] = 0;
    vnbar += side_with_adjoint<D>(gas, W, n, s, Fbar, Ubar, 0.0, vbar, Wbar);
#pragma unroll
    for (int k = 0; k < D; ++k) Wbar[1 + k] += vnbar * n[k];
    const double c2 = gas.gamma * W[P] / W[0];
    if (c2 > 0) {
        const double c2bar = cbar / (2 * s.c);
        Wbar[P] += c2bar * gas.gamma / W[0];
        Wbar[0] -= c2bar * c2 / W[0];
    }
}

// HLL's wave speeds and case as hll() computes them: bit 0 S_L from L, bit 1 S_R from L; case 1 F_L, 2 F_R, 0 the HLL state.
template <int D>
__device__ __forceinline__ int hll_case(const Side<D>& l, const Side<D>& r, double& SL, double& SR, int& pattern) {
    const double sl = l.vn - l.c, sr = r.vn - r.c, tl = l.vn + l.c, tr = r.vn + r.c;
    SL = fmin(sl, sr);
    SR = fmax(tl, tr);
    pattern = (sr < sl ? 0 : 1) | (tr > tl ? 0 : 2);
    return SL >= 0.0 ? 1 : SR <= 0.0 ? 2 : 0;
}

// euler.hpp:466-505: F from the same intermediates and dF = J (dL, dR), n and w fixed.
template <int D>
__device__ __forceinline__ void hll_tangent(const Gas& gas, const double* L, const double* R, const double* dL, const double* dR, const double* n,
                                            const double* w, double* F, double* dF) {
    Side<D> l, r, dl, dr;
    side_full<D>(gas, L, n, w, l);
    side_full<D>(gas, R, n, w, r);
    side_full_tangent<D>(gas, L, dL, n, l, dl);
    side_full_tangent<D>(gas, R, dR, n, r, dr);
    double SL, SR;
    int pattern;
    const int which = hll_case<D>(l, r, SL, SR, pattern);
    const double dSL = (pattern & 1) ? dl.vn - dl.c : dr.vn - dr.c, dSR = (pattern & 2) ? dl.vn + dl.c : dr.vn + dr.c;
    if (which == 1) {
#pragma unroll
        for (int v = 0; v < D + 2; ++v) F[v] = l.F[v], dF[v] = dl.F[v];
    } else if (which == 2) {
#pragma unroll
        for (int v = 0; v < D + 2; ++v) F[v] = r.F[v], dF[v] = dr.F[v];
    } else {
        const double inv_dS = 1.0 / (SR - SL), dinv = -inv_dS * inv_dS * (dSR - dSL);
#pragma unroll
        for (int v = 0; v < D + 2; ++v) {
            const double N = SR * l.F[v] - SL * r.F[v] + SL * SR * (r.U[v] - l.U[v]);
            const double dN = dSR * l.F[v] + SR * dl.F[v] - dSL * r.F[v] - SL * dr.F[v] + (dSL * SR + SL * dSR) * (r.U[v] - l.U[v]) +
                              SL * SR * (dr.U[v] - dl.U[v]);
            F[v] = N * inv_dS;
            dF[v] = dN * inv_dS + N * dinv;
        }
    }
    to_lab<D>(F, w);
    to_lab<D>(dF, w);
}

// euler.hpp:507-540: adds J^T Fbar of hll() to Lbar and Rbar.
template <int D>
__device__ __forceinline__ void hll_adjoint(const Gas& gas, const double* L, const double* R, const double* n, const double* w, const double* Fbar_lab,
                                            double* Lbar, double* Rbar) {
    Side<D> l, r;
    side_full<D>(gas, L, n, w, l);
    side_full<D>(gas, R, n, w, r);
    double SL, SR, Fbar[D + 2], zero[D + 2];
    int pattern;
    const int which = hll_case<D>(l, r, SL, SR, pattern);
    to_lab_adjoint<D>(Fbar_lab, w, Fbar);
#pragma unroll
    for (int v = 0; v < D + 2; ++v) zero[v] = 0;
    if (which == 1) {
        side_full_adjoint<D>(gas, L, n, l, Fbar, zero, 0.0, 0.0, Lbar);
        return;
    }
    if (which == 2) {
        side_full_adjoint<D>(gas, R, n, r, Fbar, zero, 0.0, 0.0, Rbar);
        return;
    }
    const double inv_dS = 1.0 / (SR - SL);
    double lF[D + 2], rF[D + 2], lU[D + 2], rU[D + 2], SLbar = 0, SRbar = 0, inv_bar = 0;
#pragma unroll
    for (int v = 0; v < D + 2; ++v) {
        const double dU = r.U[v] - l.U[v];
        const double N = SR * l.F[v] - SL * r.F[v] + SL * SR * dU;
        const double Nbar = Fbar[v] * inv_dS;
        inv_bar += Fbar[v] * N;
        SRbar += Nbar * (l.F[v] + SL * dU);
        SLbar += Nbar * (SR * dU - r.F[v]);
        lF[v] = Nbar * SR;
        rF[v] = -Nbar * SL;
        rU[v] = Nbar * (SL * SR);
        lU[v] = -Nbar * (SL * SR);
    }
    const double gapbar = -inv_dS * inv_dS * inv_bar;  // inv_dS = 1 / (SR - SL)
    SRbar += gapbar;
    SLbar -= gapbar;
    double lvn = 0, lc = 0, rvn = 0, rc = 0;
    if (pattern & 1)
        lvn += SLbar, lc -= SLbar;
    else
        rvn += SLbar, rc -= SLbar;
    if (pattern & 2)
        lvn += SRbar, lc += SRbar;
    else
        rvn += SRbar, rc += SRbar;
    side_full_adjoint<D>(gas, L, n, l, lF, lU, lvn, lc, Lbar);
    side_full_adjoint<D>(gas, R, n, r, rF, rU, rvn, rc, Rbar);
}

// The flux derivatives of the solver in use (0 HLLC, 1 HLL).
template <int D>
__device__ __forceinline__ void riemann_tangent(int kind, const Gas& gas, const double* L, const double* R, const double* dL, const double* dR,
                                                const double* n, const double* w, double* F, double* dF) {
    if (kind == 1)
        hll_tangent<D>(gas, L, R, dL, dR, n, w, F, dF);
    else
        hllc_tangent<D>(gas, L, R, dL, dR, n, w, F, dF);
}
template <int D>
__device__ __forceinline__ void riemann_adjoint(int kind, const Gas& gas, const double* L, const double* R, const double* n, const double* w,
                                                const double* Fbar, double* Lbar, double* Rbar) {
    if (kind == 1)
        hll_adjoint<D>(gas, L, R, n, w, Fbar, Lbar, Rbar);
    else
        hllc_adjoint<D>(gas, L, R, n, w, Fbar, Lbar, Rbar);
}

}  // namespace luna
