/* CPU oracle of luna-cpp's finite-volume hydro step (TEST INFRASTRUCTURE ONLY; see luna_oracle.h).
 *
 * Build: gcc -O2 -std=c99 -fopenmp -ffp-contract=off -fPIC -shared (oracle/Makefile). No contraction: every product and
 * sum rounds on its own, as the reference's host build does on x86-64 without FMA.
 * Reference paths are relative to the luna-cpp checkout. */
#include "luna_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXV 5
#define MAXD 3

/* euler.hpp:69-73: left to right, written out per D. */
static double dot(int D, const double* a, const double* b) {
    if (D == 1) return a[0] * b[0];
    if (D == 2) return a[0] * b[0] + a[1] * b[1];
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* euler.hpp:91-93 */
/* The OpenMP threads the oracle's loops use (the CPU baseline's single-thread round). */
void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_max_threads(void) { return omp_get_max_threads(); }

void orc_gas_init(orc_gas* gas, double gamma) {
    gas->gamma = gamma;
    gas->smallr = 1e-10;
    gas->smallc = 1e-10;
    gas->inverse_gm1 = 1.0 / (gamma - 1.0);
}

/* euler.hpp:104-113 primitive(): U -> W with the RAMSES floors. */
void orc_primitive(int D, const orc_gas* gas, const double* U, double* W) {
    const double rho = fmax(U[0], gas->smallr);
    const double inv_rho = 1.0 / rho;
    double v[MAXD];
    for (int k = 0; k < D; ++k) v[k] = U[1 + k] * inv_rho;
    const double smalle = gas->smallc * gas->smallc / gas->gamma / (gas->gamma - 1.0);
    const double kinetic = 0.5 * rho * dot(D, v, v);
    const double e = fmax(U[D + 1] - kinetic, smalle * rho);
    W[0] = rho;
    for (int k = 0; k < D; ++k) W[1 + k] = v[k];
    W[D + 1] = (gas->gamma - 1.0) * e;
}

/* euler.hpp:165-172 conserved(): W -> U without floors. */
void orc_conserved(int D, const orc_gas* gas, const double* W, double* U) {
    const double rho = W[0];
    const double kinetic = 0.5 * rho * dot(D, W + 1, W + 1);
    U[0] = rho;
    for (int k = 0; k < D; ++k) U[1 + k] = rho * W[1 + k];
    U[D + 1] = W[D + 1] / (gas->gamma - 1.0) + kinetic;
}

/* euler.hpp:175-178 */
double orc_sound_speed(int D, const orc_gas* gas, const double* W) { return sqrt(gas->gamma * W[D + 1] / W[0]); }

/* euler.hpp:186-204 half_step_rate(): sum_d A_d(W) dW/dx_d, axes in order. g: [V][D]. */
static void half_step_rate(int D, const orc_gas* gas, const double* W, const double* g, double* A) {
    const int P = D + 1;
    const double rho = W[0], inv_rho = 1.0 / W[0], rho_c2 = gas->gamma * W[P];
    for (int v = 0; v < D + 2; ++v) A[v] = 0;
    for (int d = 0; d < D; ++d) {
        const double vn = W[1 + d];
        A[0] += vn * g[0 * D + d] + rho * g[(1 + d) * D + d];
        for (int k = 0; k < D; ++k) {
            double row = vn * g[(1 + k) * D + d];
            if (k == d) row = row + inv_rho * g[P * D + d];
            A[1 + k] += row;
        }
        A[P] += rho_c2 * g[(1 + d) * D + d] + vn * g[P * D + d];
    }
}

/* euler.hpp:186-204 half_step_rate() in float (the reduced-precision face sides): W and g rounded to float first. */
static void half_step_rate_f(int D, const orc_gas* gas, const float* W, const float* g, float* A) {
    const int P = D + 1;
    const float rho = W[0], inv_rho = 1.0f / W[0], rho_c2 = (float)gas->gamma * W[P];
    for (int v = 0; v < D + 2; ++v) A[v] = 0;
    for (int d = 0; d < D; ++d) {
        const float vn = W[1 + d];
        A[0] += vn * g[0 * D + d] + rho * g[(1 + d) * D + d];
        for (int k = 0; k < D; ++k) {
            float row = vn * g[(1 + k) * D + d];
            if (k == d) row = row + inv_rho * g[P * D + d];
            A[1 + k] += row;
        }
        A[P] += rho_c2 * g[(1 + d) * D + d] + vn * g[P * D + d];
    }
}

/* euler.hpp:207-214 half_step(). */
void orc_half_step(int D, const orc_gas* gas, const double* W, const double* g, double dt, double* out) {
    double A[MAXV];
    half_step_rate(D, gas, W, g, A);
    for (int v = 0; v < D + 2; ++v) out[v] = W[v] - 0.5 * dt * A[v];
}

/* euler.hpp:316-320 */
typedef struct {
    double v[MAXD];
    double vn, c, E;
    double F[MAXV], U[MAXV];
} side_t;

/* euler.hpp:324-336 side_with() */
static void side_with(int D, const orc_gas* gas, const double* W, const double* n, const double* w, double vn, side_t* s) {
    const int P = D + 1;
    for (int k = 0; k < D; ++k) s->v[k] = W[1 + k] - w[k];
    s->vn = vn;
    s->c = 0;
    s->E = W[P] * gas->inverse_gm1 + W[0] * (0.5 * dot(D, s->v, s->v));
    const double m = W[0] * s->vn;
    s->F[0] = m;
    for (int k = 0; k < D; ++k) s->F[1 + k] = m * s->v[k] + W[P] * n[k];
    s->F[P] = (s->E + W[P]) * s->vn;
    s->U[0] = W[0];
    for (int k = 0; k < D; ++k) s->U[1 + k] = W[0] * s->v[k];
    s->U[P] = s->E;
}

/* euler.hpp:340-346 side() */
static void side_full(int D, const orc_gas* gas, const double* W, const double* n, const double* w, side_t* s) {
    const int P = D + 1;
    double rel[MAXD];
    for (int k = 0; k < D; ++k) rel[k] = W[1 + k] - w[k];
    side_with(D, gas, W, n, w, dot(D, rel, n), s);
    const double c2 = gas->gamma * W[P] / W[0];
    s->c = c2 > 0 ? sqrt(c2) : 0.0;
}

/* euler.hpp:350-356 to_lab() */
static void to_lab(int D, double* F, const double* w) {
    if (D == 1) F[2] += w[0] * F[1] + 0.5 * dot(1, w, w) * F[0];
    if (D == 2) F[3] += w[0] * F[1] + w[1] * F[2] + 0.5 * dot(2, w, w) * F[0];
    if (D == 3) F[4] += w[0] * F[1] + w[1] * F[2] + w[2] * F[3] + 0.5 * dot(3, w, w) * F[0];
    for (int k = 0; k < D; ++k) F[1 + k] += w[k] * F[0];
}

/* euler.hpp:444-461 HLL::flux() */
void orc_hll(int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w, double* F) {
    side_t l, r;
    side_full(D, gas, L, n, w, &l);
    side_full(D, gas, R, n, w, &r);
    const double SL = fmin(l.vn - l.c, r.vn - r.c);
    const double SR = fmax(l.vn + l.c, r.vn + r.c);
    if (SL >= 0.0) {
        for (int v = 0; v < D + 2; ++v) F[v] = l.F[v];
    } else if (SR <= 0.0) {
        for (int v = 0; v < D + 2; ++v) F[v] = r.F[v];
    } else {
        const double inv_dS = 1.0 / (SR - SL);
        for (int v = 0; v < D + 2; ++v) F[v] = (SR * l.F[v] - SL * r.F[v] + SL * SR * (r.U[v] - l.U[v])) * inv_dS;
    }
    to_lab(D, F, w);
}

/* euler.hpp:547-550 HLLC::nonzero() */
static double nonzero(double d) {
    const double eps = 1e-20;
    return fabs(d) < eps ? (d < 0 ? -eps : eps) : d;
}

/* euler.hpp:609-642 HLLC::flux() (LUNA_FLUX_LEVER 0) */
void orc_hllc(int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w, double* F) {
    const int P = D + 1;
    double rl[MAXD], rr[MAXD];
    for (int k = 0; k < D; ++k) {
        rl[k] = L[1 + k] - w[k];
        rr[k] = R[1 + k] - w[k];
    }
    const double vnL = dot(D, rl, n), vnR = dot(D, rr, n);
    const double inv_rhoL = 1.0 / L[0], inv_rhoR = 1.0 / R[0];
    const double c2L = gas->gamma * L[P] * inv_rhoL, c2R = gas->gamma * R[P] * inv_rhoR;
    const int ok = c2L >= 0 && c2R >= 0;
    const double cL = c2L > 0 ? sqrt(c2L) : 0.0, cR = c2R > 0 ? sqrt(c2R) : 0.0;
    const double SL = fmin(vnL - cL, vnR - cR);
    const double SR = fmax(vnL + cL, vnR + cR);
    const double denomL = L[0] * (SL - vnL), denomR = R[0] * (SR - vnR);
    const double SM = (R[P] - L[P] + denomL * vnL - denomR * vnR) / nonzero(denomL - denomR);
    const int left = ok && (SL >= 0 || SM >= 0);
    const int star = ok && SL < 0 && (SM >= 0 || SR > 0);
    double W[MAXV];
    for (int v = 0; v < D + 2; ++v) W[v] = left ? L[v] : R[v];
    side_t s;
    side_with(D, gas, W, n, w, left ? vnL : vnR, &s);
    if (!star) {
        for (int v = 0; v < D + 2; ++v) F[v] = s.F[v];
        to_lab(D, F, w);
        return;
    }
    const double S = left ? SL : SR, denom = left ? denomL : denomR;
    const double rho_star = denom / nonzero(S - SM);
    const double E_star = s.E * (left ? inv_rhoL : inv_rhoR) + (SM - s.vn) * (SM + W[P] * (1.0 / nonzero(denom)));
    double U_star[MAXV];
    U_star[0] = rho_star;
    for (int k = 0; k < D; ++k) U_star[1 + k] = rho_star * (s.v[k] + n[k] * (SM - s.vn));
    U_star[P] = rho_star * E_star;
    for (int v = 0; v < D + 2; ++v) F[v] = s.F[v] + S * (U_star[v] - s.U[v]);
    to_lab(D, F, w);
}

static void riemann(int kind, int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w, double* F) {
    if (kind == 1)
        orc_hll(D, gas, L, R, n, w, F);
    else
        orc_hllc(D, gas, L, R, n, w, F);
}

/* ------------------------------------------------------------------ D = 2, 3 on a Voronoi mesh ----------------------------- */

/* The nearest periodic image of a separation, with the side's inverse as tet_mesh::wrap(d, side, inverse) takes it
 * (finite_volume.hpp:279-282). */
static double wrap(double d, double side) { return d - side * round(d * (1.0 / side)); }

/* finite_volume.hpp:279-282 Geometry::separation(i, j): generator j seen from generator i. */
static void separation(const orc_geometry* g, int i, int j, double* d) {
    for (int a = 0; a < g->D; ++a) d[a] = wrap(g->positions[j * g->D + a] - g->positions[i * g->D + a], g->side[a]);
}

/* finite_volume.hpp:322-326 Geometry::face_offset(f, i): face centroid relative to cell i's centroid. */
static void face_offset(const orc_geometry* g, int f, int i, double* r) {
    const int D = g->D, o = g->owner[f];
    double rel[MAXD];
    if (o == i) {
        for (int a = 0; a < D; ++a) rel[a] = g->face_centroid[f * D + a];
    } else {
        double s[MAXD];
        separation(g, i, o, s);
        for (int a = 0; a < D; ++a) rel[a] = g->face_centroid[f * D + a] + s[a];
    }
    for (int a = 0; a < D; ++a) r[a] = rel[a] - g->centroid[i * D + a];
}

static int list_face(int e) { return e >= 0 ? e : ~e; }

static double magnitude_of(const orc_geometry* g, int f) { return sqrt(dot(g->D, g->area + f * g->D, g->area + f * g->D)); }

/* The cube root of v > 0 rounded to the nearest double (the reference calls its backend's cbrt, whose last bit differs
 * between glibc and the GPUs' libraries; the correctly rounded one is the same on both sides of the parity tests):
 * the library's cbrt, within an ulp, then corrected by the sign of (y +- half an ulp)^3 - v from double-double products
 * (fma gives each product's rounding error exactly; y^3 - v is exact by Sterbenz). No tie is possible. */
static double cube_excess(double y, double h, double v) {
    const double p2 = y * y, e2 = fma(y, y, -p2);
    const double p3 = p2 * y, e3 = fma(p2, y, -p3);
    const double r0 = p3 - v;
    const double big = r0 + 3.0 * (p2 * h);
    return big + (e3 + e2 * y) + 3.0 * (e2 * h) + 3.0 * (y * h * h) + h * h * h;
}
double orc_cbrt_rn(double v) {
    double y = cbrt(v);
    if (!(v > 0) || !(v < INFINITY)) return y;
    for (int k = 0; k < 4; ++k) {
        const double up = nextafter(y, INFINITY), down = nextafter(y, 0.0);
        if (cube_excess(y, 0.5 * (up - y), v) < 0)
            y = up;
        else if (cube_excess(y, -0.5 * (y - down), v) > 0)
            y = down;
        else
            break;
    }
    return y;
}

/* finite_volume.hpp:241-246 radius() */
static double radius(int D, double v) {
    if (D == 3) return orc_cbrt_rn(v * (3.0 / (4.0 * M_PI)));
    return sqrt(v * (1.0 / M_PI));
}
/* finite_volume.hpp:248-253 face_angle() */
static double face_angle(int D, double area) {
    if (D == 3) return sqrt(area / M_PI);
    return 0.5 * area;
}

/* finite_volume.hpp:403-415 primitives(): prims row = (W, c_s, r) from Q / V. */
void orc_primitives(orc_fv* fv, const orc_geometry* geo) {
    const int D = fv->D, V = D + 2, PW = V + 2;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < fv->N; ++i) {
        const double inv = 1.0 / geo->volume[i];
        double U[MAXV], W[MAXV];
        for (int v = 0; v < V; ++v) U[v] = fv->q[i * V + v] * inv;
        orc_primitive(D, &fv->gas, U, W);
        double* row = fv->prims + (size_t)i * PW;
        for (int v = 0; v < V; ++v) row[v] = W[v];
        row[V] = orc_sound_speed(D, &fv->gas, W);
        row[V + 1] = radius(D, geo->volume[i]);
    }
}

/* finite_volume.hpp:590-615 Sweep::solve(): Cholesky of M (upper triangle), then M x_v = b_v. */
static void solve_fit(int D, double M[MAXD + 1][MAXD + 1], double b[MAXV][MAXD + 1], double* grad, double* intercept) {
    const int U = D + 1, V = D + 2;
    double L[MAXD + 1][MAXD + 1];
    memset(L, 0, sizeof L);
    for (int p = 0; p < U; ++p)
        for (int s = 0; s <= p; ++s) {
            double sum = M[s][p];
            for (int m = 0; m < s; ++m) sum -= L[p][m] * L[s][m];
            L[p][s] = p == s ? sqrt(sum) : sum / L[s][s];
        }
    for (int v = 0; v < V; ++v) {
        double z[MAXD + 1], x[MAXD + 1];
        for (int p = 0; p < U; ++p) {
            double sum = b[v][p];
            for (int m = 0; m < p; ++m) sum -= L[p][m] * z[m];
            z[p] = sum / L[p][p];
        }
        for (int p = U - 1; p >= 0; --p) {
            double sum = z[p];
            for (int m = p + 1; m < U; ++m) sum -= L[m][p] * x[m];
            x[p] = sum / L[p][p];
        }
        for (int a = 0; a < D; ++a) grad[v * D + a] = x[1 + a];
        if (intercept) intercept[v] = x[0];
    }
}

/* The learned limiter's record of the limiter walk (no counterpart in luna-cpp): the unlimited gradient, the product of
 * the clip factors, the extreme unlimited face-centroid increments (0 included) and the clip count of each variable. */
typedef struct {
    double g0[MAXV * MAXD], phi[MAXV], up[MAXV], down[MAXV];
    int clips[MAXV];
} limiter_record;

static void record_init(limiter_record* r, int D, const double* grad) {
    for (int v = 0; v < D + 2; ++v) {
        for (int a = 0; a < D; ++a) r->g0[v * D + a] = grad[v * D + a];
        r->phi[v] = 1.0;
        r->up[v] = 0.0;
        r->down[v] = 0.0;
        r->clips[v] = 0;
    }
}
static void record_face(limiter_record* r, int D, const double* offset) {
    for (int v = 0; v < D + 2; ++v) {
        const double dp = dot(D, r->g0 + v * D, offset);
        r->up[v] = fmax(r->up[v], dp);
        r->down[v] = fmin(r->down[v], dp);
    }
}
static void record_clip(limiter_record* r, int v, double s) {
    r->phi[v] = r->phi[v] * s;
    r->clips[v] += 1;
}

/* Round to BF16 bits: double to float (nearest even), then float to BF16 (nearest even); NaN stays a quiet NaN. */
unsigned short orc_bf16_bits(double x) {
    const float f = (float)x;
    unsigned u;
    memcpy(&u, &f, sizeof u);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

static double clamp8(double x) { return fmin(fmax(x, -8.0), 8.0); }

/* The 16 features of each variable and the troubled flag (the definitions are those of include/luna_hydro.h,
 * luna_fv_limiter_features). */
static unsigned char limiter_features(int D, const double* W, const double* lo, const double* hi, const double* intercept, double cs,
                                      double r, int count, const limiter_record* rec, unsigned short* out) {
    const int V = D + 2;
    const double inv_c = 1.0 / cs;
    double speed2 = 0, div = 0;
    for (int a = 0; a < D; ++a) {
        speed2 += W[1 + a] * W[1 + a];
        div += rec->g0[(1 + a) * D + a];
    }
    double curl;
    if (D == 2) {
        curl = fabs(rec->g0[2 * D + 0] - rec->g0[1 * D + 1]);
    } else {
        const double cx = rec->g0[3 * D + 1] - rec->g0[2 * D + 2], cy = rec->g0[1 * D + 2] - rec->g0[3 * D + 0],
                     cz = rec->g0[2 * D + 0] - rec->g0[1 * D + 1];
        curl = sqrt(cx * cx + cy * cy + cz * cz);
    }
    double phimin = rec->phi[0];
    for (int v = 1; v < V; ++v) phimin = fmin(phimin, rec->phi[v]);
    const double rho_jump = (hi[0] - lo[0]) / W[0], p_jump = (hi[V - 1] - lo[V - 1]) / W[V - 1];
    const double compression = div * r * inv_c;
    for (int v = 0; v < V; ++v) {
        const double scale = (v > 0 && v < V - 1) ? cs : fabs(W[v]);
        const double S = (hi[v] - lo[v]) + 1e-12 * scale + 1e-300;
        const double inv = 1.0 / S;
        double f[ORC_FEATURES];
        f[0] = rec->phi[v];
        f[1] = clamp8((hi[v] - W[v]) * inv);
        f[2] = clamp8((W[v] - lo[v]) * inv);
        f[3] = clamp8(rec->up[v] * inv);
        f[4] = clamp8(-rec->down[v] * inv);
        f[5] = clamp8(intercept[v] * inv);
        f[6] = (double)rec->clips[v] / (double)count;
        f[7] = phimin;
        f[8] = clamp8(rho_jump);
        f[9] = clamp8(p_jump);
        f[10] = clamp8(compression);
        f[11] = clamp8(curl * r * inv_c);
        f[12] = clamp8(sqrt(speed2) * inv_c);
        f[13] = v == 0 ? 1.0 : 0.0;
        f[14] = v == V - 1 ? 1.0 : 0.0;
        f[15] = clamp8(S / (scale + 1e-300));
        /* + 0.0: a zero is always +0 (fmin / fmax may return either zero of a +0 / -0 tie) */
        for (int k = 0; k < ORC_FEATURES; ++k) out[v * ORC_FEATURES + k] = orc_bf16_bits(f[k] + 0.0);
    }
    const int shock = compression < -0.05 && p_jump > 0.5;
    const int thin = lo[0] < 0.1 * W[0] || lo[V - 1] < 0.1 * W[V - 1];
    return (unsigned char)((shock || thin) ? 1 : 0);
}

/* precision.hpp:55-130. binary16: sign, 5 exponent bits (bias 15), 10 fraction bits; below 2^-14 the values are
 * multiples of 2^-24. The encoder rounds to nearest, ties to even; 65520 and above become infinity. */
unsigned short orc_half_from_float(float f) {
    uint32_t x;
    memcpy(&x, &f, sizeof x);
    const unsigned short sign = (unsigned short)((x >> 16) & 0x8000u);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (unsigned short)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
    if (ax >= 0x477ff000u) return (unsigned short)(sign | 0x7c00u); /* |f| >= 65520 */
    if (ax < 0x38800000u) {                                          /* |f| < 2^-14: a multiple of 2^-24 */
        float a;
        memcpy(&a, &ax, sizeof a);
        return (unsigned short)(sign | (unsigned short)rintf(a * 16777216.0f)); /* the product is exact; rintf: ties even */
    }
    const uint32_t r = ax + 0xfffu + ((ax >> 13) & 1u); /* 23 -> 10 fraction bits, ties to even; a carry moves the exponent */
    return (unsigned short)(sign | ((r - (112u << 23)) >> 13));
}
float orc_half_to_float(unsigned short h) {
    const int e = (h >> 10) & 0x1f, m = h & 0x3ff;
    const float sign = (h & 0x8000) ? -1.0f : 1.0f;
    if (e == 0x1f) return m ? NAN : sign * INFINITY;
    return sign * (e == 0 ? ldexpf((float)m, -24) : ldexpf((float)(1024 + m), e - 25));
}
float orc_power_of_two(int k) { return ldexpf(1.0f, k); }
/* The exponent k that puts a group's largest magnitude m at m 2^k in [2^14, 2^15), from m's biased exponent bits
 * (a subnormal m counts as 2^-127), 0 for m = 0, within [-114, 126]. */
int orc_group_exponent(float m) {
    if (!(m > 0)) return 0;
    uint32_t bits;
    memcpy(&bits, &m, sizeof bits);
    const int e = (int)((bits >> 23) & 0xffu) - 127, k = 14 - e;
    return k < -114 ? -114 : k > 126 ? 126 : k;
}
static int half_group(int D, int c) { return c < D ? 0 : c < D + D * D ? 1 : 2; }

static float dotf(int D, const float* a, const float* b) {
    if (D == 1) return a[0] * b[0];
    if (D == 2) return a[0] * b[0] + a[1] * b[1];
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* finite_volume.hpp:526-537 (constants :74): the regularised point velocity of a cell with primitive row Wi (W, c_s),
 * centroid relative to its generator and largest face angle max_angle. */
void orc_regularised_point_velocity(int D, const double* Wi, const double* centroid, double max_angle, double* w) {
    const int V = D + 2;
    const double shaping = 0.5, factor = 2.25;
    double fraction = 0;
    if (max_angle > 0.75 * factor) fraction = max_angle > factor ? shaping : shaping * (max_angle - 0.75 * factor) / (0.25 * factor);
    double wi[MAXD];
    for (int a = 0; a < D; ++a) wi[a] = Wi[1 + a];
    const double cd = sqrt(dot(D, centroid, centroid));
    if (cd > 0 && fraction > 0) {
        const double s = fraction * Wi[V] / cd;
        for (int a = 0; a < D; ++a) wi[a] = wi[a] + centroid[a] * s;
    }
    for (int a = 0; a < D; ++a) w[a] = wi[a];
}

static void regularised_velocity(orc_fv* fv, int i, const double* Wi, const double* centroid, double max_angle) {
    orc_regularised_point_velocity(fv->D, Wi, centroid, max_angle, fv->w + (size_t)i * fv->D);
}

/* finite_volume.hpp:463-569 Sweep::operator() under Grad32 / Grad16 (G = float): the fit's sums in float from the double
 * separations, the neighbours' range of W_j - W_i rounded once (:469, 510-512), the float Cholesky solve (:590-615), the
 * limiter on float offsets against that range (:558-562); the row as stored (:618-641): float, or FP16 of g 2^k per group
 * with k from the group's largest |g|, kept here as FaceFlux::gradient() (:1288-1300) decodes it. */
static void sweep_cell_lp(orc_fv* fv, const orc_geometry* geo, int i) {
    const int D = fv->D, V = D + 2, U = D + 1, PW = V + 2;
    const double* Wi = fv->prims + (size_t)i * PW;
    double surface = 0, smallest = INFINITY, max_angle = 0;
    float lo[MAXV], hi[MAXV], M[MAXD + 1][MAXD + 1], b[MAXV][MAXD + 1];
    for (int v = 0; v < V; ++v) {
        lo[v] = FLT_MAX;
        hi[v] = -FLT_MAX;
        for (int p = 0; p < U; ++p) b[v][p] = 0;
    }
    for (int p = 0; p < U; ++p)
        for (int s = 0; s < U; ++s) M[p][s] = 0;
    const double* centroid = geo->centroid + (size_t)i * D;
    const int first = geo->start[i], end = geo->start[i + 1];
    for (int k = first; k < end; ++k) {
        const int e = geo->list[k], f = list_face(e);
        const int j = e >= 0 ? geo->neighbour[f] : geo->owner[f];
        const double* Wj = fv->prims + (size_t)j * PW;
        double d[MAXD];
        separation(geo, i, j, d);
        const double magnitude = magnitude_of(geo, f);
        surface += magnitude;
        smallest = fmin(smallest, magnitude);
        if (fv->moving) max_angle = fmax(max_angle, face_angle(D, magnitude) / (0.5 * sqrt(dot(D, d, d))));
        float X[MAXD + 1];
        X[0] = 1;
        for (int p = 1; p < U; ++p) X[p] = (float)(d[p - 1] + geo->centroid[(size_t)j * D + p - 1] - centroid[p - 1]);
        const float w = (float)magnitude;
        for (int p = 0; p < U; ++p)
            for (int s = p; s < U; ++s) M[p][s] += w * X[p] * X[s];
        for (int v = 0; v < V; ++v) {
            const double dW = Wj[v] - Wi[v];
            for (int p = 0; p < U; ++p) b[v][p] += w * X[p] * (float)dW;
            lo[v] = fminf(lo[v], (float)dW);
            hi[v] = fmaxf(hi[v], (float)dW);
        }
    }
    const double tiny = 1e-6 * surface;
    const int any_tiny = smallest < tiny;
    if (any_tiny) {
        for (int v = 0; v < V; ++v) {
            lo[v] = FLT_MAX;
            hi[v] = -FLT_MAX;
        }
        for (int k = first; k < end; ++k) {
            const int e = geo->list[k], f = list_face(e);
            const int j = e >= 0 ? geo->neighbour[f] : geo->owner[f];
            if (magnitude_of(geo, f) < tiny) continue;
            const double* Wj = fv->prims + (size_t)j * PW;
            for (int v = 0; v < V; ++v) {
                lo[v] = fminf(lo[v], (float)(Wj[v] - Wi[v]));
                hi[v] = fmaxf(hi[v], (float)(Wj[v] - Wi[v]));
            }
        }
    }
    float L[MAXD + 1][MAXD + 1], grad[MAXV * MAXD];
    for (int p = 0; p < U; ++p)
        for (int s = 0; s < U; ++s) L[p][s] = 0;
    for (int p = 0; p < U; ++p)
        for (int s = 0; s <= p; ++s) {
            float sum = M[s][p];
            for (int m = 0; m < s; ++m) sum -= L[p][m] * L[s][m];
            L[p][s] = p == s ? sqrtf(sum) : sum / L[s][s];
        }
    for (int v = 0; v < V; ++v) {
        float z[MAXD + 1], x[MAXD + 1];
        for (int p = 0; p < U; ++p) {
            float sum = b[v][p];
            for (int m = 0; m < p; ++m) sum -= L[p][m] * z[m];
            z[p] = sum / L[p][p];
        }
        for (int p = U - 1; p >= 0; --p) {
            float sum = z[p];
            for (int m = p + 1; m < U; ++m) sum -= L[m][p] * x[m];
            x[p] = sum / L[p][p];
        }
        for (int a = 0; a < D; ++a) grad[v * D + a] = x[1 + a];
    }
    if (fv->moving) regularised_velocity(fv, i, Wi, centroid, max_angle);
    if (fv->limiter == 0) {
        for (int k = first; k < end; ++k) {
            const int f = list_face(geo->list[k]);
            if (any_tiny && magnitude_of(geo, f) < tiny) continue;
            double offset[MAXD];
            face_offset(geo, f, i, offset);
            float rg[MAXD];
            for (int a = 0; a < D; ++a) rg[a] = (float)offset[a];
            for (int v = 0; v < V; ++v) {
                float* g = grad + v * D;
                const float dp = dotf(D, g, rg);
                if (dp > 0 && dp > hi[v]) {
                    const float s = hi[v] > 0 ? hi[v] / dp : 0.0f;
                    for (int a = 0; a < D; ++a) g[a] = g[a] * s;
                } else if (dp < 0 && dp < lo[v]) {
                    const float s = lo[v] < 0 ? lo[v] / dp : 0.0f;
                    for (int a = 0; a < D; ++a) g[a] = g[a] * s;
                }
            }
        }
    }
    double* out = fv->grad + (size_t)i * V * D;
    if (fv->precision == 1) {
        for (int c = 0; c < V * D; ++c) out[c] = grad[c];
        return;
    }
    float largest[3] = {0, 0, 0};
    for (int c = 0; c < V * D; ++c) largest[half_group(D, c)] = fmaxf(largest[half_group(D, c)], fabsf(grad[c]));
    int kq[3];
    for (int q = 0; q < 3; ++q) kq[q] = orc_group_exponent(largest[q]);
    for (int c = 0; c < V * D; ++c) {
        const int q = half_group(D, c);
        const unsigned short h = orc_half_from_float(grad[c] * orc_power_of_two(kq[q]));
        out[c] = orc_half_to_float(h) * orc_power_of_two(-kq[q]);
    }
}

/* finite_volume.hpp:463-569 Sweep::operator() (double policy, the face-centroid limiter or none), with the moving mesh's
 * point velocity (:526-537, constants :74) and neighbour_range (:572-588). */
/* finite_volume.hpp:880-1037 LaneSweep<Moving> (order 1, double, face-centroid limiter or none): L lanes per cell, lane l
 * taking the list entries l, l + L, ... in list order with the one-lane sweep's arithmetic per entry. The lanes' sums
 * combine by a butterfly (stage o = L/2, ..., 1: lane l adds lane l^o's value; addition commutes, so every lane holds the
 * same bits after each stage, and the total is lane 0's), the minima and maxima likewise. The reference combines them with
 * Kokkos's vector reduction, a fixed tree of its own backend; this restates the tree the GPU kernel uses. The limiter
 * takes, per variable, the smallest factor over the faces, each from the unlimited gradient, and scales once. */
static void lanes_combine(int L, double* a, int op) { /* op 0 sum, 1 min, 2 max; a[0] gets the result */
    double t[8];
    for (int o = L / 2; o > 0; o /= 2) {
        for (int l = 0; l < L; ++l) t[l] = op == 0 ? a[l] + a[l ^ o] : op == 1 ? fmin(a[l], a[l ^ o]) : fmax(a[l], a[l ^ o]);
        for (int l = 0; l < L; ++l) a[l] = t[l];
    }
}

static void sweep_cell_lanes(orc_fv* fv, const orc_geometry* geo, int i, int L) {
    const int D = fv->D, V = D + 2, U = D + 1, PW = V + 2;
    const double* Wi = fv->prims + (size_t)i * PW;
    const double* centroid = geo->centroid + (size_t)i * D;
    const int first = geo->start[i], end = geo->start[i + 1];
    double surface[8], smallest[8], angle[8], lo[MAXV][8], hi[MAXV][8], M[MAXD + 1][MAXD + 1][8], b[MAXV][MAXD + 1][8];
    for (int l = 0; l < L; ++l) {
        surface[l] = 0;
        smallest[l] = INFINITY;
        angle[l] = 0;
        for (int v = 0; v < V; ++v) {
            lo[v][l] = DBL_MAX;
            hi[v][l] = -DBL_MAX;
            for (int p = 0; p < U; ++p) b[v][p][l] = 0;
        }
        for (int p = 0; p < U; ++p)
            for (int s = 0; s < U; ++s) M[p][s][l] = 0;
    }
    for (int k = first; k < end; ++k) {
        const int l = (k - first) % L;
        const int e = geo->list[k], f = list_face(e);
        const int j = e >= 0 ? geo->neighbour[f] : geo->owner[f];
        const double* Wj = fv->prims + (size_t)j * PW;
        double d[MAXD];
        separation(geo, i, j, d);
        const double magnitude = magnitude_of(geo, f);
        surface[l] += magnitude;
        smallest[l] = fmin(smallest[l], magnitude);
        if (fv->moving) angle[l] = fmax(angle[l], face_angle(D, magnitude) / (0.5 * sqrt(dot(D, d, d))));
        double X[MAXD + 1];
        X[0] = 1;
        for (int p = 1; p < U; ++p) X[p] = d[p - 1] + geo->centroid[(size_t)j * D + p - 1] - centroid[p - 1];
        const double w = magnitude;
        for (int p = 0; p < U; ++p)
            for (int s = p; s < U; ++s) M[p][s][l] += w * X[p] * X[s];
        for (int v = 0; v < V; ++v) {
            const double dW = Wj[v] - Wi[v];
            for (int p = 0; p < U; ++p) b[v][p][l] += w * X[p] * dW;
            lo[v][l] = fmin(lo[v][l], Wj[v]);
            hi[v][l] = fmax(hi[v][l], Wj[v]);
        }
    }
    lanes_combine(L, surface, 0);
    lanes_combine(L, smallest, 1);
    lanes_combine(L, angle, 2);
    double Mt[MAXD + 1][MAXD + 1], bt[MAXV][MAXD + 1], lot[MAXV], hit[MAXV];
    memset(Mt, 0, sizeof Mt);
    for (int p = 0; p < U; ++p)
        for (int s = p; s < U; ++s) {
            lanes_combine(L, M[p][s], 0);
            Mt[p][s] = M[p][s][0];
        }
    for (int v = 0; v < V; ++v) {
        for (int p = 0; p < U; ++p) {
            lanes_combine(L, b[v][p], 0);
            bt[v][p] = b[v][p][0];
        }
        lanes_combine(L, lo[v], 1);
        lanes_combine(L, hi[v], 2);
        lot[v] = lo[v][0];
        hit[v] = hi[v][0];
    }
    /* Rare: faces below 1e-6 of the surface leave the range and the limiter (the range taken again, one walk). */
    const double tiny = 1e-6 * surface[0];
    const int any_tiny = smallest[0] < tiny;
    if (any_tiny) {
        for (int v = 0; v < V; ++v) {
            lot[v] = DBL_MAX;
            hit[v] = -DBL_MAX;
        }
        for (int k = first; k < end; ++k) {
            const int e = geo->list[k], f = list_face(e);
            const int j = e >= 0 ? geo->neighbour[f] : geo->owner[f];
            if (magnitude_of(geo, f) < tiny) continue;
            const double* Wj = fv->prims + (size_t)j * PW;
            for (int v = 0; v < V; ++v) {
                lot[v] = fmin(lot[v], Wj[v]);
                hit[v] = fmax(hit[v], Wj[v]);
            }
        }
    }
    double grad[MAXV * MAXD], intercept[MAXV];
    solve_fit(D, Mt, bt, grad, intercept);
    if (fv->moving) regularised_velocity(fv, i, Wi, centroid, angle[0]);
    if (fv->limiter == 0) {
        double factor[MAXV][8];
        for (int v = 0; v < V; ++v)
            for (int l = 0; l < L; ++l) factor[v][l] = 1.0;
        for (int k = first; k < end; ++k) {
            const int l = (k - first) % L, f = list_face(geo->list[k]);
            if (any_tiny && magnitude_of(geo, f) < tiny) continue;
            double offset[MAXD];
            face_offset(geo, f, i, offset);
            for (int v = 0; v < V; ++v) {
                const double dp = dot(D, grad + v * D, offset);
                double s = 1.0;
                if (dp > 0 && Wi[v] + dp > hit[v])
                    s = hit[v] > Wi[v] ? (hit[v] - Wi[v]) / dp : 0.0;
                else if (dp < 0 && Wi[v] + dp < lot[v])
                    s = lot[v] < Wi[v] ? (lot[v] - Wi[v]) / dp : 0.0;
                factor[v][l] = fmin(factor[v][l], s);
            }
        }
        for (int v = 0; v < V; ++v) {
            lanes_combine(L, factor[v], 1);
            if (factor[v][0] < 1.0)
                for (int a = 0; a < D; ++a) grad[v * D + a] = grad[v * D + a] * factor[v][0];
        }
    }
    memcpy(fv->grad + (size_t)i * V * D, grad, sizeof(double) * V * D);
    if (fv->hll_compression > 0) {
        double div = 0;
        for (int a = 0; a < D; ++a) div += grad[(1 + a) * D + a];
        fv->hll[i] = div * Wi[V + 1] / Wi[V] < -fv->hll_compression ? 1 : 0;
    }
}

static void sweep_cell(orc_fv* fv, const orc_geometry* geo, int i) {
    if (fv->precision != 0 && fv->order == 1) {
        sweep_cell_lp(fv, geo, i);
        return;
    }
    if (fv->sweep_lanes > 1 && fv->order == 1) {
        sweep_cell_lanes(fv, geo, i, fv->sweep_lanes);
        return;
    }
    const int D = fv->D, V = D + 2, U = D + 1, PW = V + 2;
    const double* Wi = fv->prims + (size_t)i * PW;
    double surface = 0, smallest = INFINITY, max_angle = 0;
    double lo[MAXV], hi[MAXV];
    for (int v = 0; v < V; ++v) {
        lo[v] = DBL_MAX;
        hi[v] = -DBL_MAX;
    }
    double M[MAXD + 1][MAXD + 1], b[MAXV][MAXD + 1];
    memset(M, 0, sizeof M);
    memset(b, 0, sizeof b);
    const double* centroid = geo->centroid + (size_t)i * D;
    const int first = geo->start[i], end = geo->start[i + 1];
    const int grad_on = fv->order == 1;
    for (int k = first; k < end; ++k) {
        const int e = geo->list[k], f = list_face(e);
        const int j = e >= 0 ? geo->neighbour[f] : geo->owner[f];
        const double* Wj = fv->prims + (size_t)j * PW;
        double d[MAXD];
        separation(geo, i, j, d);
        const double magnitude = magnitude_of(geo, f);
        surface += magnitude;
        smallest = fmin(smallest, magnitude);
        if (fv->moving) max_angle = fmax(max_angle, face_angle(D, magnitude) / (0.5 * sqrt(dot(D, d, d))));
        if (grad_on) {
            double X[MAXD + 1];
            X[0] = 1;
            for (int p = 1; p < U; ++p) X[p] = d[p - 1] + geo->centroid[(size_t)j * D + p - 1] - centroid[p - 1];
            const double w = magnitude;
            for (int p = 0; p < U; ++p)
                for (int s = p; s < U; ++s) M[p][s] += w * X[p] * X[s];
            for (int v = 0; v < V; ++v) {
                const double dW = Wj[v] - Wi[v];
                for (int p = 0; p < U; ++p) b[v][p] += w * X[p] * dW;
                lo[v] = fmin(lo[v], Wj[v]);
                hi[v] = fmax(hi[v], Wj[v]);
            }
        }
    }
    const double tiny = 1e-6 * surface;
    const int any_tiny = smallest < tiny;
    if (grad_on && any_tiny) {
        for (int v = 0; v < V; ++v) {
            lo[v] = DBL_MAX;
            hi[v] = -DBL_MAX;
        }
        for (int k = first; k < end; ++k) {
            const int e = geo->list[k], f = list_face(e);
            const int j = e >= 0 ? geo->neighbour[f] : geo->owner[f];
            if (magnitude_of(geo, f) < tiny) continue;
            const double* Wj = fv->prims + (size_t)j * PW;
            for (int v = 0; v < V; ++v) {
                lo[v] = fmin(lo[v], Wj[v]);
                hi[v] = fmax(hi[v], Wj[v]);
            }
        }
    }
    double grad[MAXV * MAXD], intercept[MAXV];
    if (grad_on) solve_fit(D, M, b, grad, intercept);
    if (fv->moving) regularised_velocity(fv, i, Wi, centroid, max_angle);
    if (!grad_on) return;
    if (fv->limiter == 0 || fv->limiter == 2) {
        limiter_record rec;
        record_init(&rec, D, grad);
        for (int k = first; k < end; ++k) {
            const int f = list_face(geo->list[k]);
            if (any_tiny && magnitude_of(geo, f) < tiny) continue;
            double offset[MAXD];
            face_offset(geo, f, i, offset);
            record_face(&rec, D, offset);
            for (int v = 0; v < V; ++v) {
                double* g = grad + v * D;
                const double dp = dot(D, g, offset);
                if (dp > 0 && Wi[v] + dp > hi[v]) {
                    const double s = hi[v] > Wi[v] ? (hi[v] - Wi[v]) / dp : 0.0;
                    for (int a = 0; a < D; ++a) g[a] = g[a] * s;
                    record_clip(&rec, v, s);
                } else if (dp < 0 && Wi[v] + dp < lo[v]) {
                    const double s = lo[v] < Wi[v] ? (lo[v] - Wi[v]) / dp : 0.0;
                    for (int a = 0; a < D; ++a) g[a] = g[a] * s;
                    record_clip(&rec, v, s);
                }
            }
        }
        if (fv->limiter == 2)
            fv->troubled[i] = limiter_features(D, Wi, lo, hi, intercept, Wi[V], Wi[V + 1], end - first, &rec,
                                               fv->features + (size_t)i * V * ORC_FEATURES);
    }
    memcpy(fv->grad + (size_t)i * V * D, grad, sizeof(double) * V * D);
    if (fv->hll_compression > 0 && fv->limiter != 2) { /* the HLL fallback's compression indicator */
        double div = 0;
        for (int a = 0; a < D; ++a) div += grad[(1 + a) * D + a];
        fv->hll[i] = div * Wi[V + 1] / Wi[V] < -fv->hll_compression ? 1 : 0;
    }
}

/* finite_volume.hpp:1040-1052 CellStep, :1054-1071 LunaStep, :1074-1090 FaceStep, :1107-1149 sweep(). */
static double cell_term(const orc_fv* fv, int i, double courant) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    const double* row = fv->prims + (size_t)i * PW;
    double rel[MAXD];
    for (int a = 0; a < D; ++a) rel[a] = row[1 + a];
    if (fv->moving)
        for (int a = 0; a < D; ++a) rel[a] = rel[a] - fv->w[(size_t)i * D + a];
    const double step = row[V + 1] / (row[V] + sqrt(dot(D, rel, rel)));
    return courant * step;
}

static double luna_term(const orc_fv* fv, int i, double courant) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    const double* row = fv->prims + (size_t)i * PW;
    const double c = row[V], dx = row[V + 1] * sqrt(M_PI);
    double rate = 0;
    for (int a = 0; a < D; ++a) {
        double relative = row[1 + a];
        if (fv->moving) relative -= fv->w[(size_t)i * D + a];
        rate += (fabs(relative) + c) / dx;
    }
    return courant / rate;
}

static double face_term(const orc_fv* fv, const orc_geometry* geo, int f, double courant) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    const int o = geo->owner[f], j = geo->neighbour[f];
    double d[MAXD], dv[MAXD];
    separation(geo, o, j, d);
    const double dist = sqrt(dot(D, d, d));
    const double* po = fv->prims + (size_t)o * PW;
    const double* pj = fv->prims + (size_t)j * PW;
    for (int a = 0; a < D; ++a) dv[a] = pj[1 + a] - po[1 + a];
    const double vsig = po[V] + pj[V] - dot(D, dv, d) / dist;
    double smallest = INFINITY;
    if (vsig > 0) {
        smallest = fmin(smallest, courant * ((dist + po[V + 1]) / vsig));
        smallest = fmin(smallest, courant * ((dist + pj[V + 1]) / vsig));
    }
    return smallest;
}

/* finite_volume.hpp:190-201 time_step(): primitives (unless the moving step's second pass wrote them), the sweep, dt. */
double orc_time_step(orc_fv* fv, const orc_geometry* geo, double courant) {
    if (!fv->prims_ready) orc_primitives(fv, geo);
    fv->prims_ready = 0;
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < fv->N; ++i) sweep_cell(fv, geo, i);
    double smallest = INFINITY;
    if (fv->D == 2 && fv->step_rule == 1) {
#pragma omp parallel for reduction(min : smallest)
        for (int i = 0; i < fv->N; ++i) smallest = fmin(smallest, luna_term(fv, i, courant));
        return smallest;
    }
    const int total = geo->F > fv->N ? geo->F : fv->N;
#pragma omp parallel for reduction(min : smallest)
    for (int k = 0; k < total; ++k) {
        if (k < geo->F) smallest = fmin(smallest, face_term(fv, geo, k, courant));
        if (k < fv->N) smallest = fmin(smallest, cell_term(fv, k, courant));
    }
    return smallest;
}

/* finite_volume.hpp:1306-1329 FaceFlux::side() at order 1 (double): half step + g.r, positivity fallback.
 * step_dt: the FaceFlux's dt (dt of the step; 2 dt on a moving step's second pass). */
static void face_side(const orc_fv* fv, int i, const double* r, double step_dt, double* Wf, int* fallbacks) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    const double* W = fv->prims + (size_t)i * PW;
    const double* g = fv->grad + (size_t)i * V * D;
    if (fv->order == 0) {
        for (int v = 0; v < V; ++v) Wf[v] = W[v];
        return;
    }
    if (fv->precision != 0) { /* :1319-1324: W in double, the rest in float */
        float gf[MAXV * MAXD], Wg[MAXV], A[MAXV], rg[MAXD];
        for (int c = 0; c < V * D; ++c) gf[c] = (float)g[c];
        for (int v = 0; v < V; ++v) Wg[v] = (float)W[v];
        half_step_rate_f(D, &fv->gas, Wg, gf, A);
        for (int a = 0; a < D; ++a) rg[a] = (float)r[a];
        const float h = 0.5f * (float)step_dt;
        for (int v = 0; v < V; ++v) Wf[v] = W[v] + (double)(dotf(D, gf + v * D, rg) - h * A[v]);
    } else {
        orc_half_step(D, &fv->gas, W, g, step_dt, Wf);
        for (int v = 0; v < V; ++v) Wf[v] += dot(D, g + v * D, r);
    }
    if (Wf[0] > 0 && Wf[D + 1] > 0) return;
    ++*fallbacks;
    for (int v = 0; v < V; ++v) Wf[v] = W[v];
}

/* finite_volume.hpp:1451-1465 MovingFlux<false>::pass_side(): W + g.r on the mesh of t^n, no half step. */
static void pass_side_first(const orc_fv* fv, int i, const double* r, double* Wf, int* fallbacks) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    const double* W = fv->prims + (size_t)i * PW;
    const double* g = fv->grad + (size_t)i * V * D;
    if (fv->precision != 0) {
        float gf[MAXV * MAXD], rg[MAXD];
        for (int c = 0; c < V * D; ++c) gf[c] = (float)g[c];
        for (int a = 0; a < D; ++a) rg[a] = (float)r[a];
        for (int v = 0; v < V; ++v) Wf[v] = W[v] + (double)dotf(D, gf + v * D, rg);
    } else {
        for (int v = 0; v < V; ++v) Wf[v] = W[v] + dot(D, g + v * D, r);
    }
    if (Wf[0] > 0 && Wf[D + 1] > 0) return;
    ++*fallbacks;
    for (int v = 0; v < V; ++v) Wf[v] = W[v];
}

/* finite_volume.hpp:1380-1397 FaceFlux::face_velocity(): d the separation owner -> neighbour, xf the face centroid relative
 * to the owner's generator, wL / wR the two point velocities. */
void orc_face_velocity(int D, const double* d, const double* xf, const double* wL, const double* wR, double* wf) {
    double c[MAXD], dw[MAXD];
    const double dist = sqrt(dot(D, d, d));
    const double inv_dist = 1.0 / dist;
    for (int a = 0; a < D; ++a) c[a] = xf[a] - d[a] * 0.5;
    for (int a = 0; a < D; ++a) dw[a] = wL[a] - wR[a];
    double facv = dot(D, c, dw) * inv_dist;
    const double limit = 0.9 * dist, c2 = dot(D, c, c);
    if (c2 > limit * limit) facv *= limit / sqrt(c2);
    for (int a = 0; a < D; ++a) wf[a] = (wL[a] + wR[a]) * 0.5 + d[a] * (facv * inv_dist);
}

static void face_velocity(const orc_geometry* geo, int f, int o, int j, const double* wL, const double* wR, double* wf) {
    double d[MAXD];
    separation(geo, o, j, d);
    orc_face_velocity(geo->D, d, geo->face_centroid + (size_t)f * geo->D, wL, wR, wf);
}

/* The moving step's point velocity from its grid steps (finite_volume.hpp:1446-1448). */
static void grid_velocity(const orc_fv* fv, int i, double unit, double* w) {
    for (int a = 0; a < fv->D; ++a) w[a] = (double)fv->steps[(size_t)i * fv->D + a] * unit;
}

/* One face's flux times |A| (finite_volume.hpp:1415-1434 FaceFlux::operator(), :1466-1490 MovingFlux::operator()).
 * pass: -1 static, 0 a moving step's first pass (mesh t^n), 1 its second (mesh t^n+1). */
static double face_states(const orc_fv* fv, const orc_geometry* geo, int f, double dt, int pass, double unit, double* L, double* R,
                          double* n, double* wf, int* fallbacks) {
    const int D = fv->D;
    const int o = geo->owner[f], j = geo->neighbour[f];
    const double* a = geo->area + (size_t)f * D;
    const double magnitude = sqrt(dot(D, a, a));
    for (int k = 0; k < D; ++k) wf[k] = 0;
    if (magnitude == 0) return 0;
    double ro[MAXD], rj[MAXD];
    if (pass < 0) {
        const double inv = 1.0 / magnitude;
        for (int k = 0; k < D; ++k) n[k] = a[k] * inv;
        face_offset(geo, f, o, ro);
        face_offset(geo, f, j, rj);
        const double half = 0.5 * dt;
        for (int k = 0; k < D; ++k) {
            ro[k] = ro[k] + wf[k] * half;
            rj[k] = rj[k] + wf[k] * half;
        }
        face_side(fv, o, ro, dt, L, fallbacks);
        face_side(fv, j, rj, dt, R, fallbacks);
    } else {
        const double inv = 1.0 / magnitude;
        for (int k = 0; k < D; ++k) n[k] = a[k] * inv;
        double wo[MAXD], wj[MAXD];
        if (pass == 2) { /* order 0 on a moving mesh: the point velocities of the sweep (FaceFlux<0, true>) */
            for (int k = 0; k < D; ++k) {
                wo[k] = fv->w[(size_t)o * D + k];
                wj[k] = fv->w[(size_t)j * D + k];
            }
        } else {
            grid_velocity(fv, o, unit, wo);
            grid_velocity(fv, j, unit, wj);
        }
        face_velocity(geo, f, o, j, wo, wj, wf);
        face_offset(geo, f, o, ro);
        face_offset(geo, f, j, rj);
        if (pass == 2) {
            face_side(fv, o, ro, dt, L, fallbacks);
            face_side(fv, j, rj, dt, R, fallbacks);
        } else if (pass == 0) {
            pass_side_first(fv, o, ro, L, fallbacks);
            pass_side_first(fv, j, rj, R, fallbacks);
        } else {
            /* Second pass: side(i, r + c^n+1 - c^n + w dt) with the FaceFlux dt = 2 dt (its half step spans dt). */
            const double step_dt = 2 * dt;
            double r2[MAXD];
            const int cells[2] = {o, j};
            const double* rs[2] = {ro, rj};
            double* out[2] = {L, R};
            for (int s = 0; s < 2; ++s) {
                const int i = cells[s];
                double wi[MAXD];
                grid_velocity(fv, i, unit, wi);
                for (int k = 0; k < D; ++k) {
                    const double moved = wi[k] * (0.5 * step_dt);
                    r2[k] = rs[s][k] + geo->centroid[(size_t)i * D + k] - fv->old_centroid[(size_t)i * D + k] + moved;
                }
                face_side(fv, i, r2, step_dt, out[s], fallbacks);
            }
        }
    }
    return magnitude;
}

static void face_flux(const orc_fv* fv, const orc_geometry* geo, int f, double dt, int pass, double unit, double* F, int* fallbacks) {
    const int D = fv->D, V = D + 2;
    const int o = geo->owner[f], j = geo->neighbour[f];
    double L[MAXV], R[MAXV], n[MAXD], wf[MAXD];
    for (int v = 0; v < V; ++v) F[v] = 0;
    const double magnitude = face_states(fv, geo, f, dt, pass, unit, L, R, n, wf, fallbacks);
    if (magnitude == 0) return;
    /* the HLL fallback: HLLC takes HLL where a side is flagged as extremely compressed */
    const int solver = fv->riemann == 0 && fv->hll_compression > 0 && (fv->hll[o] || fv->hll[j]) ? 1 : fv->riemann;
    riemann(solver, D, &fv->gas, L, R, n, wf, F);
    for (int v = 0; v < V; ++v) F[v] *= magnitude;
}

/* The two face states of face f as face_flux() hands them to the Riemann solver (for the tests that pin them): L, R [V],
 * the unit normal n and the face velocity wf [D]; returns |A| (0: a zero-area face, nothing written but wf = 0) and the
 * positivity fallbacks taken in *fallbacks. pass and unit as face_flux(): -1 static (dt the step's), 0 / 1 a moving
 * step's first / second pass (dt as apply() passes it: 0 / the step's dt; unit = grid step / dt), 2 order 0 moving. */
double orc_face_sides(const orc_fv* fv, const orc_geometry* geo, int f, double dt, int pass, double unit, double* L, double* R, double* n,
                      double* wf, int* fallbacks) {
    *fallbacks = 0;
    return face_states(fv, geo, f, dt, pass, unit, L, R, n, wf, fallbacks);
}

/* finite_volume.hpp:1539-1631 apply(), per-face loop: the flux rows, then Q -= dt sum_k sign_k F in list order.
 * pass 1 (a moving step's second) also writes the next primitive rows (:1596-1602). */
static void apply(orc_fv* fv, const orc_geometry* geo, double flux_dt, double apply_dt, int pass, double unit) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    long long count = 0;
#pragma omp parallel for reduction(+ : count) schedule(static)
    for (int f = 0; f < geo->F; ++f) {
        int fallbacks = 0;
        face_flux(fv, geo, f, flux_dt, pass, unit, fv->flux + (size_t)f * V, &fallbacks);
        count += fallbacks;
    }
    fv->fallbacks[pass == 1 ? 1 : 0] = count;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < fv->N; ++i) {
        double sum[MAXV] = {0, 0, 0, 0, 0};
        for (int k = geo->start[i]; k < geo->start[i + 1]; ++k) {
            const int e = geo->list[k], f = list_face(e);
            const double sign = e >= 0 ? 1.0 : -1.0;
            for (int v = 0; v < V; ++v) sum[v] += sign * fv->flux[(size_t)f * V + v];
        }
        double next[MAXV];
        for (int v = 0; v < V; ++v) fv->q[(size_t)i * V + v] = next[v] = fv->q[(size_t)i * V + v] - apply_dt * sum[v];
        if (pass == 1) {
            const double inv = 1.0 / geo->volume[i];
            double U[MAXV], W[MAXV];
            for (int v = 0; v < V; ++v) U[v] = next[v] * inv;
            orc_primitive(D, &fv->gas, U, W);
            double* row = fv->prims + (size_t)i * PW;
            for (int v = 0; v < V; ++v) row[v] = W[v];
            row[V] = orc_sound_speed(D, &fv->gas, W);
            row[V + 1] = radius(D, geo->volume[i]);
        }
    }
}

/* finite_volume.hpp:205-226 step() on a static mesh, order 0 or 1, per-face loop (advance -> apply). */
void orc_step_static(orc_fv* fv, const orc_geometry* geo, double dt) {
    fv->fallbacks[1] = 0;
    apply(fv, geo, dt, dt, -1, 0.0);
}

/* The integer grid key of x (the role of exact_predicates::snap): the nearest grid point, wrapped into [0, period). */
long long orc_snap(double x, double step, long long period) {
    long long k = (long long)round(x * (1.0 / step));
    k %= period;
    if (k < 0) k += period;
    return k;
}
/* The move k0 -> k1 in grid steps to the nearest periodic image (the role of exact_predicates::delta). */
int orc_delta(long long k0, long long k1, long long period) {
    long long d = k1 - k0;
    if (d > period / 2) d -= period;
    if (d < -(period / 2)) d += period;
    return (int)d;
}

/* finite_volume.hpp:1497-1516 step_moving() up to the mesh update: keep c^n, snap x, whole grid steps; pass 1 (half dt);
 * drift (:1219-1231) onto the grid point the steps reach. The caller then builds the mesh of the new positions. */
void orc_moving_begin(orc_fv* fv, const orc_geometry* geo, double dt) {
    const int D = fv->D;
    if (fv->order == 0) { /* finite_volume.hpp:216-225: advance<0, true>() on the mesh of t^n, then move(): x += dt w */
        fv->fallbacks[1] = 0;
        apply(fv, geo, dt, dt, 2, 0.0);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < fv->N; ++i)
            for (int a = 0; a < D; ++a) { /* :1227-1228, off the grid */
                const size_t at = (size_t)i * D + a;
                const double moved = fv->x[at] + dt * fv->w[at];
                fv->x[at] = moved - geo->side[a] * floor(moved / geo->side[a]);
            }
        return;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < fv->N; ++i)
        for (int a = 0; a < D; ++a) {
            const size_t at = (size_t)i * D + a;
            fv->old_centroid[at] = geo->centroid[at];
            const long long k0 = orc_snap(fv->x[at], geo->step, geo->period[a]);
            const long long k1 = orc_snap(fv->x[at] + dt * fv->w[at], geo->step, geo->period[a]);
            fv->x[at] = (double)k0 * geo->step;
            if (!fv->hold_steps) fv->steps[at] = orc_delta(k0, k1, geo->period[a]);
        }
    fv->fallbacks[1] = 0;
    apply(fv, geo, 0.0, 0.5 * dt, 0, geo->step / dt);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < fv->N; ++i)
        for (int a = 0; a < D; ++a) {
            const size_t at = (size_t)i * D + a;
            const long long k = (long long)round(fv->x[at] * (1.0 / geo->step)) + fv->steps[at];
            const long long p = geo->period[a];
            fv->x[at] = (double)(((k % p) + p) % p) * geo->step;
        }
}

/* finite_volume.hpp:1517-1518: the second pass on the mesh of t^n+1 (half dt); its update writes the next primitives. */
void orc_moving_finish(orc_fv* fv, const orc_geometry* geo, double dt) {
    if (fv->order == 0) { /* the new mesh's primitives */
        orc_primitives(fv, geo);
        fv->prims_ready = 1;
        return;
    }
    apply(fv, geo, dt, 0.5 * dt, 1, geo->step / dt);
    fv->prims_ready = 1;
}

/* ------------------------------------------------------------------ D = 1 on a sorted line ---------------------------------- */

/* A line this short runs on one thread: its loops are microseconds, below OpenMP's fork cost (the results do not depend
 * on the thread count: each slot is computed alone, and the reductions are a minimum and an integer count). */
#define LINE_PARALLEL 16384

/* sorted_mesh::Line as finite_volume_1d.hpp reads it: gaps in grid steps between neighbouring generators, a wall's gap
 * twice the distance to it (the mirror ghost, :152-155, :636-653). */
static long long line_gap(const orc_line* l, int s) {
    const int n = l->n;
    if (s >= 0 && s < n - 1) return l->key[s + 1] - l->key[s];
    if (l->periodic) return l->key[0] + l->period - l->key[n - 1];
    if (s < 0) return 2 * l->key[0];
    return 2 * (l->period - l->key[n - 1]);
}
static double line_length(const orc_line* l, int s) { return (double)line_gap(l, s) * l->step; }
/* finite_volume_1d.hpp:173, 176 (load_inner): V and the centroid relative to the generator from the two gaps. */
double orc_line_volume(const orc_line* l, int s) { return ((double)line_gap(l, s - 1) + (double)line_gap(l, s)) * (0.5 * l->step); }
double orc_line_centroid(const orc_line* l, int s) { return ((double)line_gap(l, s) - (double)line_gap(l, s - 1)) * (0.25 * l->step); }
/* line.order(s): the cell in slot s. */
static int line_id(const orc_line* l, int s) { return l->id ? l->id[s] : s; }
static int line_slot(const orc_line* l, int s, int k) {
    const int t = s + k;
    if (l->periodic) return ((t % l->n) + l->n) % l->n;
    return t >= 0 && t < l->n ? t : -1;
}

/* finite_volume_1d.hpp:107-111 primitive_of() */
static void line_primitive(const orc_line* l, int s, double* W) {
    const double inv = 1.0 / orc_line_volume(l, s);
    const int i = line_id(l, s);
    const double U[3] = {l->q[i * 3] * inv, l->q[i * 3 + 1] * inv, l->q[i * 3 + 2] * inv};
    orc_primitive(1, &l->gas, U, W);
}

/* finite_volume_1d.hpp:117-128 ghost(), ghost_face(): reflective mirrors the velocity, inflow is the wall's state,
 * outflow copies. */
static void line_ghost(const orc_line* l, const double* W, int left, double* out) {
    const int wall = left ? l->wall_left : l->wall_right;
    for (int v = 0; v < 3; ++v) out[v] = wall == 2 ? l->inflow[left ? 0 : 1][v] : W[v];
    if (wall == 0) out[1] = -out[1];
}
static void line_ghost_face(const orc_line* l, const double* Wf, const double* W, int left, double* out) {
    const int wall = left ? l->wall_left : l->wall_right;
    if (wall == 0) {
        for (int v = 0; v < 3; ++v) out[v] = Wf[v];
        out[1] = -out[1];
        return;
    }
    line_ghost(l, W, left, out);
}

typedef struct {
    double W[3], c, V;
    int real;
} line_cell;

/* finite_volume_1d.hpp:132-156 load<R>() with R = 2. */
static void line_load(const orc_line* l, int s, line_cell cell[5], double gap[4]) {
    int slot[5];
    for (int k = 0; k < 5; ++k) {
        slot[k] = line_slot(l, s, k - 2);
        cell[k].W[0] = cell[k].W[1] = cell[k].W[2] = 0;
        cell[k].c = cell[k].V = 0;
        cell[k].real = 1;
    }
    for (int k = 0; k < 5; ++k)
        if (slot[k] >= 0) {
            line_primitive(l, slot[k], cell[k].W);
            cell[k].c = orc_line_centroid(l, slot[k]);
            cell[k].V = orc_line_volume(l, slot[k]);
        }
    for (int k = 0; k < 5; ++k) {
        if (slot[k] >= 0) continue;
        cell[k].real = 0;
        const int end = k < 2 ? k + 1 : k - 1;
        if (slot[end] < 0) continue;
        line_ghost(l, cell[end].W, k < 2, cell[k].W);
        cell[k].c = -cell[end].c;
    }
    for (int k = 0; k < 4; ++k) {
        const int a = slot[k], b = slot[k + 1];
        gap[k] = a >= 0 ? line_length(l, a) : b >= 0 ? line_length(l, -1) : 0.0;
    }
}

/* finite_volume_1d.hpp:183-205 Stencil::gradient(): the limited gradient of variable v (MC face limiter, minmod, van Leer, none). */
static double line_gradient(const orc_line* l, const line_cell* cell, double gapL, double gapR, int v) {
    const line_cell* me = cell;
    const double dL = -gapL + cell[-1].c - me->c, dR = gapR + cell[1].c - me->c;
    const double Wi = me->W[v], left = cell[-1].W[v], right = cell[1].W[v];
    if (l->limiter == 2 || l->limiter == 3) {
        const double sL = (left - Wi) / dL, sR = (right - Wi) / dR;
        if (!(sL * sR > 0)) return 0.0;
        if (l->limiter == 3) return 2 * sL * sR / (sL + sR);
        return fabs(sL) < fabs(sR) ? sL : sR;
    }
    double g = ((right - Wi) - (left - Wi)) / (dR - dL);
    if (l->limiter == 1) return g;
    const double r[2] = {-0.5 * gapL - me->c, 0.5 * gapR - me->c};
    const double lo = fmin(left, right), hi = fmax(left, right);
    for (int f = 0; f < 2; ++f) {
        const double dp = g * r[f];
        if (dp > 0 && Wi + dp > hi)
            g = g * (hi > Wi ? (hi - Wi) / dp : 0.0);
        else if (dp < 0 && Wi + dp < lo)
            g = g * (lo < Wi ? (lo - Wi) / dp : 0.0);
    }
    return g;
}

/* finite_volume_1d.hpp:492-505 Kernel::face_state() (static). */
static void line_face_state(const orc_line* l, const double* W, const double* g, double r, double dt, double* out) {
    if (l->order == 0) {
        for (int v = 0; v < 3; ++v) out[v] = W[v];
        return;
    }
    orc_half_step(1, &l->gas, W, g, dt, out);
    for (int v = 0; v < 3; ++v) out[v] += g[v] * r;
    if (out[0] > 0 && out[2] > 0) return;
    for (int v = 0; v < 3; ++v) out[v] = W[v];
}

/* finite_volume_1d.hpp:555-661 Kernel::body() (static): the slot's two face fluxes and its new state; fused: the next dt by
 * LUNA's rule from the new state. */
static double line_body(const orc_line* l, int s, double dt, double courant, int fused) {
    line_cell cell[5];
    double gap[4];
    line_load(l, s, cell, gap);
    double grad[5][3];
    memset(grad, 0, sizeof grad);
    for (int k = 1; k <= 3; ++k) {
        if (!cell[k].real) continue;
        if (l->order == 1)
            for (int v = 0; v < 3; ++v) grad[k][v] = line_gradient(l, cell + k, gap[k - 1], gap[k], v);
    }
    double F[2][3];
    const double one = 1.0, zero = 0.0;
    for (int f = 0; f < 2; ++f) {
        const int a = 1 + f, b = 2 + f;
        const double w = 0, shift = 0.5 * dt * w;
        double L[3] = {0, 0, 0}, R[3] = {0, 0, 0};
        if (cell[a].real) line_face_state(l, cell[a].W, grad[a], 0.5 * gap[a] - cell[a].c + shift, dt, L);
        if (cell[b].real) line_face_state(l, cell[b].W, grad[b], -0.5 * gap[a] - cell[b].c + shift, dt, R);
        if (!cell[a].real) line_ghost_face(l, R, cell[b].W, 1, L);
        if (!cell[b].real) line_ghost_face(l, L, cell[a].W, 0, R);
        riemann(l->riemann, 1, &l->gas, L, R, &one, &zero, F[f]);
    }
    double Unew[3];
    const int i = line_id(l, s);
    for (int v = 0; v < 3; ++v) {
        Unew[v] = l->q[i * 3 + v] - dt * (F[1][v] - F[0][v]);
        l->next[i * 3 + v] = Unew[v];
    }
    if (!fused) return INFINITY;
    const double volume = cell[2].V;
    const double inv = 1.0 / volume;
    const double U[3] = {Unew[0] * inv, Unew[1] * inv, Unew[2] * inv};
    double Wnew[3];
    orc_primitive(1, &l->gas, U, Wnew);
    const double relative = fabs(Wnew[1]);
    return courant * volume / (relative + orc_sound_speed(1, &l->gas, Wnew));
}

/* finite_volume_1d.hpp:301-314 drift(): the point's velocity relative to its cell's from AREPO-ng's regularisation, for
 * primitives W, density gradient g, centroid c (relative to the generator) and volume. */
#define LINE_SHAPING_SPEED 0.5 /* cell_shaping_speed, finite_volume_1d.hpp:26 */
static double line_drift(const orc_line* l, const double* W, double g, double c, double volume) {
    const double r = 0.5 * volume;
    if (!l->regularise || fabs(c) <= 0.5 * r) return 0.0;
    double offset = c;
    const double dgrad = fabs(g);
    if (dgrad > 0) {
        const double tmp = 3 * r + W[0] / dgrad;
        const double x = (tmp - sqrt(tmp * tmp - 8 * r * r)) / 4;
        if (x < 0.25 * r) offset += x / dgrad * g;
    }
    const double d = fabs(offset);
    if (d <= 0.75 * r) return 0.0;
    const double fraction = d > r ? LINE_SHAPING_SPEED : LINE_SHAPING_SPEED * (d - 0.75 * r) / (0.25 * r);
    return offset * (fraction * orc_sound_speed(1, &l->gas, W) / d);
}

/* finite_volume_1d.hpp:319-328 point_velocity(): drift(), plus, for a generator beyond 0.75 r of its centroid, up to the
 * fluid's velocity at the generator, linear between this cell's centroid and that of the neighbour on that side. */
static double line_point_velocity(const orc_line* l, const line_cell* cell, double gapL, double gapR, double g_rho, double volume) {
    const line_cell* me = cell;
    const double r = 0.5 * volume, d = fabs(me->c), weight = fmin(1.0, (d - 0.75 * r) / (0.25 * r));
    const double drifting = line_drift(l, me->W, g_rho, me->c, volume);
    if (!(weight > 0)) return drifting;
    const int side = me->c < 0 ? 1 : -1;
    const double between = side > 0 ? gapR + cell[1].c - me->c : gapL + me->c - cell[-1].c;
    return drifting + weight * (cell[side].W[1] - me->W[1]) * (d / between);
}

/* finite_volume_1d.hpp:546-550 moved(): where a point at x0 is after dt (wrapped on a periodic line, held between the
 * walls otherwise). */
static double line_moved(const orc_line* l, double x0, double velocity, double dt) {
    const double side = l->side;
    const double y = fma(dt, velocity, x0);
    return l->periodic ? y - side * floor(y / side) : fmin(fmax(y, 0.0), side);
}

/* sorted_mesh::snap (voronoi-cpp, absent): the nearest grid key of x, wrapped into [0, period) on a periodic line. */
static long long line_snap(const orc_line* l, double x) {
    long long k = (long long)round(x * (1.0 / l->step));
    if (l->periodic) {
        k %= l->period;
        if (k < 0) k += l->period;
    }
    return k;
}

/* sorted_mesh::before (voronoi-cpp, absent): the line's order, by key, then by cell. */
static int line_before(long long ka, int a, long long kb, int b) { return ka < kb || (ka == kb && a < b); }

/* finite_volume_1d.hpp:509-524 extrapolated(): W + g r, a whole step on (MUSCL-Hancock) when `second`; W itself if rho or
 * P would not stay positive. */
static void line_extrapolated(const orc_line* l, const double* W, const double* g, double r, int second, double dt, double* out) {
    if (l->order == 0) {
        for (int v = 0; v < 3; ++v) out[v] = W[v];
        return;
    }
    if (second)
        orc_half_step(1, &l->gas, W, g, 2 * dt, out);
    else
        for (int v = 0; v < 3; ++v) out[v] = W[v];
    for (int v = 0; v < 3; ++v) out[v] += g[v] * r;
    if (out[0] > 0 && out[2] > 0) return;
    for (int v = 0; v < 3; ++v) out[v] = W[v];
}

/* finite_volume_1d.hpp:555-661 Kernel::body() (moving): the gradients and point velocities of slots s - 1, s, s + 1, their
 * new keys and moves in grid steps; each face's flux the mean of those on the meshes of t^n and t^n+1, the face moving by
 * the mean of its generators' moves; the new state and position of the slot's cell, its new key into key_next. Returns the
 * next dt by LUNA's rule (fused) and whether slot s and the next slot are out of order after the move (*descent). */
static double line_body_moving(const orc_line* l, int s, double dt, double courant, int fused, const double* x, double* x_next,
                               long long* key_next, int* descent) {
    line_cell cell[5];
    double gap[4];
    line_load(l, s, cell, gap);
    double grad[5][3], point[5] = {0, 0, 0, 0, 0};
    memset(grad, 0, sizeof grad);
    for (int k = 1; k <= 3; ++k) {
        if (!cell[k].real) continue;
        if (l->order == 1)
            for (int v = 0; v < 3; ++v) grad[k][v] = line_gradient(l, cell + k, gap[k - 1], gap[k], v);
        point[k] = cell[k].W[1] + line_point_velocity(l, cell + k, gap[k - 1], gap[k], grad[k][0], 0.5 * (gap[k - 1] + gap[k]));
    }
    long long key[5] = {0, 0, 0, 0, 0};
    double moves[5] = {0, 0, 0, 0, 0};
    for (int k = 1; k <= 3; ++k) {
        if (!cell[k].real) continue;
        const int t = line_slot(l, s, k - 2);
        key[k] = line_snap(l, line_moved(l, x[line_id(l, t)], point[k], dt));
        long long d = key[k] - l->key[t];
        if (l->periodic) d = d > l->period / 2 ? d - l->period : d < -l->period / 2 ? d + l->period : d;
        moves[k] = (double)d;
    }
    double F[2][3];
    for (int f = 0; f < 2; ++f) {
        const int a = 1 + f, b = 2 + f;
        const int inside = cell[a].real && cell[b].real;
        const double travel = inside ? 0.5 * (moves[a] + moves[b]) * l->step : 0.0;
        const double w = travel / dt;
        const double ra = 0.5 * gap[a] - cell[a].c, rb = -0.5 * gap[a] - cell[b].c;
        const double one = 1.0;
        double on[2][3];
        const int meshes = l->order == 0 ? 1 : 2;
        for (int second = 0; second < meshes; ++second) {
            double L[3] = {0, 0, 0}, R[3] = {0, 0, 0};
            if (cell[a].real) line_extrapolated(l, cell[a].W, grad[a], ra + second * travel, second, dt, L);
            if (cell[b].real) line_extrapolated(l, cell[b].W, grad[b], rb + second * travel, second, dt, R);
            if (!cell[a].real) line_ghost_face(l, R, cell[b].W, 1, L);
            if (!cell[b].real) line_ghost_face(l, L, cell[a].W, 0, R);
            riemann(l->riemann, 1, &l->gas, L, R, &one, &w, on[second]);
        }
        for (int v = 0; v < 3; ++v) F[f][v] = meshes == 1 ? on[0][v] : 0.5 * (on[0][v] + on[1][v]);
    }
    const int i = line_id(l, s);
    double Unew[3];
    for (int v = 0; v < 3; ++v) {
        Unew[v] = l->q[i * 3 + v] - dt * (F[1][v] - F[0][v]);
        l->next[i * 3 + v] = Unew[v];
    }
    x_next[i] = line_moved(l, x[i], point[2], dt);
    key_next[s] = key[2];
    *descent = 0;
    if (cell[3].real && s < l->n - 1) *descent = !line_before(key[2], i, key[3], line_id(l, s + 1));
    if (!fused) return INFINITY;
    /* :636-659: V from the new keys (a wall's gap twice the distance to it), |w - v| of t^n. */
    double left, right;
    if (cell[1].real) {
        const long long d = key[2] - key[1];
        left = (double)(d < 0 && l->periodic ? d + l->period : d);
    } else {
        left = (double)(2 * key[2]);
    }
    if (cell[3].real) {
        const long long d = key[3] - key[2];
        right = (double)(d < 0 && l->periodic ? d + l->period : d);
    } else {
        right = (double)(2 * (l->period - key[2]));
    }
    const double volume = (left + right) * (0.5 * l->step);
    const double relative = fabs(point[2] - cell[2].W[1]);
    const double inv = 1.0 / volume;
    const double U[3] = {Unew[0] * inv, Unew[1] * inv, Unew[2] * inv};
    double Wnew[3];
    orc_primitive(1, &l->gas, U, Wnew);
    return courant * volume / (relative + orc_sound_speed(1, &l->gas, Wnew));
}

/* finite_volume_1d.hpp:332-386 time_step(): LUNA's rule or AREPO's (with the wall ghosts' pair terms); on a moving line |v|
 * is |w - v| of the point velocity, from the cell's density gradient. */
double orc_line_time_step(orc_line* l, double courant, int luna_rule) {
    double smallest = INFINITY;
#pragma omp parallel for reduction(min : smallest) if (l->n >= LINE_PARALLEL)
    for (int s = 0; s < l->n; ++s) {
        double Wi[3];
        line_primitive(l, s, Wi);
        const double ci = orc_sound_speed(1, &l->gas, Wi);
        const double volume = orc_line_volume(l, s);
        double vi = fabs(Wi[1]);
        if (l->moving) {
            line_cell cell[5];
            double gap[4];
            line_load(l, s, cell, gap);
            const double g = l->order == 1 ? line_gradient(l, cell + 2, gap[1], gap[2], 0) : 0.0;
            vi = fabs(line_point_velocity(l, cell + 2, gap[1], gap[2], g, volume));
        }
        if (luna_rule) {
            smallest = fmin(smallest, courant * volume / (vi + ci));
            continue;
        }
        const double r = 0.5 * volume;
        double step = r / (ci + vi);
        for (int side = -1; side <= 1; side += 2) {
            const int t = line_slot(l, s, side);
            double Wj[3];
            if (t >= 0)
                line_primitive(l, t, Wj);
            else
                line_ghost(l, Wi, side < 0, Wj);
            const double dist = line_length(l, side < 0 ? s - 1 : s);
            const double vsig = ci + orc_sound_speed(1, &l->gas, Wj) - (Wj[1] - Wi[1]) * side;
            if (vsig > 0) step = fmin(step, (dist + r) / vsig);
        }
        smallest = fmin(smallest, courant * step);
    }
    return smallest;
}

static void line_swap(orc_line* l) {
    double* t = l->q;
    l->q = l->next;
    l->next = t;
}

/* finite_volume_1d.hpp:389-407 step() (static): one kernel over the slots, then Q and next swap. */
void orc_line_step(orc_line* l, double dt) {
#pragma omp parallel for schedule(static) if (l->n >= LINE_PARALLEL)
    for (int s = 0; s < l->n; ++s) line_body(l, s, dt, 0.0, 0);
    line_swap(l);
}

/* finite_volume_1d.hpp:412-417, 679-708 step_and_time_step() (static): the step and the next dt by LUNA's rule. */
double orc_line_step_fused(orc_line* l, double dt, double courant) {
    double smallest = INFINITY;
#pragma omp parallel for reduction(min : smallest) schedule(static) if (l->n >= LINE_PARALLEL)
    for (int s = 0; s < l->n; ++s) smallest = fmin(smallest, line_body(l, s, dt, courant, 1));
    line_swap(l);
    return smallest;
}

/* ------------------------------------------------------------------ D = 1 on a moving line ----------------------------------- */

struct orc_mline {
    orc_line l;               /* l.key, l.id, l.q, l.next point into the arrays below */
    long long *keys, *keys_next;
    int* ids;
    double *q, *next, *x, *x_next;
    int capacity;
    int supply_left, supply_right, watching, looked, wants, until_look;
    double spacing;
    long long added, updates, sorts;
};

static void mline_bind(orc_mline* m) {
    m->l.key = m->keys;
    m->l.id = m->ids;
    m->l.q = m->q;
    m->l.next = m->next;
}

/* finite_volume_1d.hpp:86-88 room_at(): an end cell of this volume, its generator d grid steps from the wall. */
static int mline_room_at(const orc_mline* m, double volume, long long d) {
    return volume > 2 * m->spacing && (double)d * m->l.step >= 1.5 * m->spacing;
}
/* finite_volume_1d.hpp:89-94 wants_cell(): 0..2 cells end slot s asks its wall for. */
static int mline_wants_cell(const orc_mline* m, int s) {
    const orc_line* l = &m->l;
    int wants = 0;
    if (s == 0 && m->supply_left) wants += mline_room_at(m, orc_line_volume(l, 0), l->key[0]);
    if (s == l->n - 1 && m->supply_right) wants += mline_room_at(m, orc_line_volume(l, s), l->period - l->key[s]);
    return wants;
}

/* finite_volume_1d.hpp:62-67 set_inflow_spacing(): an end supplies cells on a moving line with walls, unless reflective. */
static void mline_set_spacing(orc_mline* m, double spacing) {
    m->spacing = spacing;
    m->supply_left = !m->l.periodic && m->l.wall_left != 0 && spacing > 0;
    m->supply_right = !m->l.periodic && m->l.wall_right != 0 && spacing > 0;
}

orc_mline* orc_mline_create(int n, const double* x, double side, long long period, int periodic, int wall_left, int wall_right,
                            const double* inflow_left, const double* inflow_right, int order, int limiter, int riemann, double gamma,
                            double spacing, int regularise) {
    orc_mline* m = (orc_mline*)calloc(1, sizeof(orc_mline));
    m->capacity = n;
    m->keys = (long long*)malloc(sizeof(long long) * n);
    m->keys_next = (long long*)malloc(sizeof(long long) * n);
    m->ids = (int*)malloc(sizeof(int) * n);
    m->q = (double*)calloc(3 * (size_t)n, sizeof(double));
    m->next = (double*)calloc(3 * (size_t)n, sizeof(double));
    m->x = (double*)malloc(sizeof(double) * n);
    m->x_next = (double*)malloc(sizeof(double) * n);
    orc_line* l = &m->l;
    l->n = n;
    l->periodic = periodic;
    l->wall_left = wall_left;
    l->wall_right = wall_right;
    l->order = order;
    l->limiter = limiter;
    l->riemann = riemann;
    orc_gas_init(&l->gas, gamma);
    l->period = period;
    l->step = side / (double)period;
    l->side = side;
    l->moving = 1;
    l->regularise = regularise;
    for (int v = 0; v < 3; ++v) {
        l->inflow[0][v] = inflow_left ? inflow_left[v] : 0.0;
        l->inflow[1][v] = inflow_right ? inflow_right[v] : 0.0;
    }
    for (int i = 0; i < n; ++i) {
        m->x[i] = x[i];
        m->ids[i] = i;
    }
    mline_bind(m);
    for (int i = 0; i < n; ++i) m->keys[i] = line_snap(l, x[i]);
    mline_set_spacing(m, spacing);
    return m;
}

void orc_mline_destroy(orc_mline* m) {
    if (!m) return;
    free(m->keys);
    free(m->keys_next);
    free(m->ids);
    free(m->q);
    free(m->next);
    free(m->x);
    free(m->x_next);
    free(m);
}

int orc_mline_cells(const orc_mline* m) { return m->l.n; }
void orc_mline_set_state(orc_mline* m, const double* q) { memcpy(m->q, q, sizeof(double) * 3 * m->l.n); }
void orc_mline_get_state(const orc_mline* m, double* q) { memcpy(q, m->q, sizeof(double) * 3 * m->l.n); }
void orc_mline_get_positions(const orc_mline* m, double* x) { memcpy(x, m->x, sizeof(double) * m->l.n); }
void orc_mline_get_line(const orc_mline* m, long long* keys, int* ids) {
    memcpy(keys, m->keys, sizeof(long long) * m->l.n);
    memcpy(ids, m->ids, sizeof(int) * m->l.n);
}
long long orc_mline_stat(const orc_mline* m, int what) { return what == 0 ? m->added : what == 1 ? m->updates : m->sorts; }

/* finite_volume_1d.hpp:332-386, 372-381: time_step(); while gas enters at an end, also what the end slots ask their walls for. */
double orc_mline_time_step(orc_mline* m, double courant, int luna_rule) {
    const double dt = orc_line_time_step(&m->l, courant, luna_rule);
    if (m->watching) {
        m->wants = mline_wants_cell(m, 0) + mline_wants_cell(m, m->l.n - 1);
        m->looked = 1;
    }
    return dt;
}

typedef struct {
    long long key;
    int id;
} mline_entry;
static int mline_compare(const void* a, const void* b) {
    const mline_entry* x = (const mline_entry*)a;
    const mline_entry* y = (const mline_entry*)b;
    if (line_before(x->key, x->id, y->key, y->id)) return -1;
    return line_before(y->key, y->id, x->key, x->id) ? 1 : 0;
}

/* Mesh<1>::order(points) (voronoi-cpp, absent): the slots in key order again after the points moved. The keys are those
 * the step computed for the new positions; a sort runs only when two slots are out of order. A periodic line is kept
 * starting at its smallest key, so a point crossing the seam also re-sorts (each slot's neighbours are the same either way). */
static void mline_order(orc_mline* m, int descents) {
    m->updates += 1;
    if (descents == 0) return;
    const int n = m->l.n;
    mline_entry* e = (mline_entry*)malloc(sizeof(mline_entry) * n);
    for (int s = 0; s < n; ++s) {
        e[s].key = m->keys[s];
        e[s].id = m->ids[s];
    }
    qsort(e, n, sizeof(mline_entry), mline_compare);
    for (int s = 0; s < n; ++s) {
        m->keys[s] = e[s].key;
        m->ids[s] = e[s].id;
    }
    free(e);
    m->sorts += 1;
}

/* finite_volume_1d.hpp:965-981 reserve_cells(): room for `count` cells, a quarter more when full. */
static void mline_reserve(orc_mline* m, int count) {
    if (count <= m->capacity) return;
    int cap = m->capacity + m->capacity / 4;
    if (cap < count) cap = count;
    m->keys = (long long*)realloc(m->keys, sizeof(long long) * cap);
    m->keys_next = (long long*)realloc(m->keys_next, sizeof(long long) * cap);
    m->ids = (int*)realloc(m->ids, sizeof(int) * cap);
    m->q = (double*)realloc(m->q, sizeof(double) * 3 * cap);
    m->next = (double*)realloc(m->next, sizeof(double) * 3 * cap);
    m->x = (double*)realloc(m->x, sizeof(double) * cap);
    m->x_next = (double*)realloc(m->x_next, sizeof(double) * cap);
    m->capacity = cap;
    mline_bind(m);
}

/* finite_volume_1d.hpp:419-465 supply_cells(): the ends' distances from their walls and wishes; while gas enters at an open
 * end the line is watched every step. A wanted cell's generator goes `spacing` inside the end generator's distance from the
 * wall; Mesh<1>::insert (voronoi-cpp, absent) puts it in a new end slot, and the new cell takes from the old end cell the
 * fraction of it that it overlaps (its whole length, from the wall to halfway to the old end generator) of mass, momentum
 * and energy. New cells are numbered after the last, the left wall's first. */
static int mline_supply(orc_mline* m) {
    orc_line* l = &m->l;
    const int last = l->n - 1;
    double ends[5];
    ends[0] = (double)l->key[0] * l->step;
    ends[1] = mline_wants_cell(m, 0);
    ends[2] = (double)(l->period - l->key[last]) * l->step;
    ends[3] = mline_wants_cell(m, last);
    double W0[3], Wl[3];
    line_primitive(l, 0, W0);
    line_primitive(l, last, Wl);
    ends[4] = (m->supply_left && W0[1] > 0) || (m->supply_right && Wl[1] < 0);
    m->watching = m->watching || ends[4] > 0;
    double wanted[2];
    int left[2], requests = 0;
    if (ends[1] > 0) {
        left[requests] = 1;
        wanted[requests++] = ends[0] - m->spacing;
    }
    if (ends[3] > 0) {
        left[requests] = 0;
        wanted[requests++] = l->side - (ends[2] - m->spacing);
    }
    if (requests == 0) return 0;
    mline_reserve(m, l->n + requests);
    l = &m->l;
    for (int r = 0; r < requests; ++r) {
        const int n = l->n, i = n;
        const long long k = line_snap(l, wanted[r]);
        double overlap, volume;
        int old;
        if (left[r]) {
            old = m->ids[0];
            overlap = ((double)(2 * k) + (double)(m->keys[0] - k)) * (0.5 * l->step);
            volume = orc_line_volume(l, 0);
            memmove(m->keys + 1, m->keys, sizeof(long long) * n);
            memmove(m->ids + 1, m->ids, sizeof(int) * n);
            m->keys[0] = k;
            m->ids[0] = i;
        } else {
            old = m->ids[n - 1];
            overlap = ((double)(k - m->keys[n - 1]) + (double)(2 * (l->period - k))) * (0.5 * l->step);
            volume = orc_line_volume(l, n - 1);
            m->keys[n] = k;
            m->ids[n] = i;
        }
        for (int v = 0; v < 3; ++v) m->q[i * 3 + v] = 0;
        const double fraction = overlap / volume;
        for (int v = 0; v < 3; ++v) {
            const double part = fraction * m->q[old * 3 + v];
            m->q[old * 3 + v] -= part;
            m->q[i * 3 + v] += part;
        }
        m->x[i] = wanted[r];
        l->n = n + 1;
        m->added += 1;
    }
    return 1;
}

/* finite_volume_1d.hpp:954-961 due(): look at the ends after this step: as the kernels asked while gas enters, else every
 * 64 steps. */
static int mline_due(orc_mline* m) {
    if (!m->supply_left && !m->supply_right) return 0;
    if (m->watching) return !m->looked || m->wants > 0;
    if (--m->until_look > 0) return 0;
    m->until_look = 64;
    return 1;
}

static int mline_kernel(orc_mline* m, double dt, double courant, int fused, double* smallest) {
    orc_line* l = &m->l;
    int descents = 0;
    double least = INFINITY;
#pragma omp parallel for reduction(min : least) reduction(+ : descents) schedule(static) if (l->n >= LINE_PARALLEL)
    for (int s = 0; s < l->n; ++s) {
        int d = 0;
        least = fmin(least, line_body_moving(l, s, dt, courant, fused, m->x, m->x_next, m->keys_next, &d));
        descents += d;
    }
    *smallest = least;
    double* t = m->q;
    m->q = m->next;
    m->next = t;
    t = m->x;
    m->x = m->x_next;
    m->x_next = t;
    long long* k = m->keys;
    m->keys = m->keys_next;
    m->keys_next = k;
    mline_bind(m);
    return descents;
}

/* finite_volume_1d.hpp:389-407 step() (moving): the kernel, Q and the positions swap, the order is updated, and the ends
 * supply cells when due. */
void orc_mline_step(orc_mline* m, double dt) {
    double unused;
    const int descents = mline_kernel(m, dt, 0.0, 0, &unused);
    mline_order(m, descents);
    if (mline_due(m)) mline_supply(m);
    m->looked = 0;
}

/* finite_volume_1d.hpp:412-417, 679-708 step_and_time_step() (moving): the step and the next dt by LUNA's rule from the new
 * keys; after a sort or new cells the next dt is time_step(courant, luna) of the new line. */
double orc_mline_step_fused(orc_mline* m, double dt, double courant) {
    int wants = 0;
    if (m->watching) wants = mline_wants_cell(m, 0) + mline_wants_cell(m, m->l.n - 1);
    double smallest;
    const int descents = mline_kernel(m, dt, courant, 1, &smallest);
    int again = 0;
    mline_order(m, descents);
    if (descents != 0) again = 1;
    m->wants = wants;
    m->looked = m->watching;
    if (mline_due(m)) again = mline_supply(m) || again;
    m->looked = 0;
    if (again) return orc_mline_time_step(m, courant, 1);
    return smallest;
}

/* ------------------------------------------------------------------ derivatives of a static step --------------------------- */
static void riemann_tangent(int kind, int D, const orc_gas* gas, const double* L, const double* R, const double* dL, const double* dR,
                            const double* n, const double* w, double* F, double* dF);
static void riemann_adjoint(int kind, int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w,
                            const double* Fbar, double* Lbar, double* Rbar);
/* Tangent: J dx; adjoint: J^T ybar (added into the bars). dt and the geometry are fixed. At every switch the forward takes
 * (a floor, a wave case, a guard, a limiter clip, a positivity fallback) the derivative is that of the branch it took,
 * as euler.hpp:115-117 states for the reference's own derivatives. Written from the forward arithmetic above by the
 * chain rule. */

/* euler.hpp:119-126: which floors primitive() applied (bit 0 density, bit 1 internal energy). */
static int primitive_floors(int D, const orc_gas* gas, const double* U) {
    const double rho = fmax(U[0], gas->smallr);
    const double inv_rho = 1.0 / rho;
    double v[MAXD];
    for (int k = 0; k < D; ++k) v[k] = U[1 + k] * inv_rho;
    const double smalle = gas->smallc * gas->smallc / gas->gamma / (gas->gamma - 1.0);
    return (U[0] >= gas->smallr ? 0 : 1) | (U[D + 1] - 0.5 * rho * dot(D, v, v) >= smalle * rho ? 0 : 2);
}

/* euler.hpp:128-140 primitive_tangent(): a floored density or energy has derivative zero in what the floor replaced. */
void orc_primitive_tangent(int D, const orc_gas* gas, const double* U, const double* dU, double* dW) {
    const int floors = primitive_floors(D, gas, U);
    const double rho = fmax(U[0], gas->smallr), inv_rho = 1.0 / rho;
    const double drho = (floors & 1) ? 0.0 : dU[0];
    double v[MAXD], dv[MAXD];
    for (int k = 0; k < D; ++k) v[k] = U[1 + k] * inv_rho;
    for (int k = 0; k < D; ++k) dv[k] = (dU[1 + k] - v[k] * drho) * inv_rho;
    const double smalle = gas->smallc * gas->smallc / gas->gamma / (gas->gamma - 1.0);
    const double dkinetic = 0.5 * drho * dot(D, v, v) + rho * dot(D, v, dv);
    const double de = (floors & 2) ? smalle * drho : dU[D + 1] - dkinetic;
    dW[0] = drho;
    for (int k = 0; k < D; ++k) dW[1 + k] = dv[k];
    dW[D + 1] = (gas->gamma - 1.0) * de;
}

/* euler.hpp:142-162 primitive_adjoint(): Ubar = J^T Wbar. */
void orc_primitive_adjoint(int D, const orc_gas* gas, const double* U, const double* Wbar, double* Ubar) {
    const int floors = primitive_floors(D, gas, U);
    const double rho = fmax(U[0], gas->smallr), inv_rho = 1.0 / rho;
    double v[MAXD], vbar[MAXD];
    for (int k = 0; k < D; ++k) {
        v[k] = U[1 + k] * inv_rho;
        vbar[k] = Wbar[1 + k];
    }
    const double smalle = gas->smallc * gas->smallc / gas->gamma / (gas->gamma - 1.0);
    const double ebar = (gas->gamma - 1.0) * Wbar[D + 1];
    double rhobar = Wbar[0], energybar = 0.0;
    if (floors & 2) {
        rhobar += smalle * ebar; /* e = smalle rho */
    } else {                     /* e = E - rho v.v / 2 */
        energybar = ebar;
        rhobar -= 0.5 * dot(D, v, v) * ebar;
        for (int k = 0; k < D; ++k) vbar[k] = vbar[k] - v[k] * (rho * ebar);
    }
    rhobar -= dot(D, v, vbar) * inv_rho; /* v = m / rho */
    Ubar[0] = (floors & 1) ? 0.0 : rhobar;
    for (int k = 0; k < D; ++k) Ubar[1 + k] = vbar[k] * inv_rho;
    Ubar[D + 1] = energybar;
}

/* euler.hpp:216-237 half_step_tangent() (dt fixed). g, dg: [V][D]. */
void orc_half_step_tangent(int D, const orc_gas* gas, const double* W, const double* dW, const double* g, const double* dg, double dt,
                           double* out) {
    const int P = D + 1;
    const double rho = W[0], inv_rho = 1.0 / W[0];
    const double rho_c2 = gas->gamma * W[P];
    const double drho = dW[0], dinv_rho = -inv_rho * inv_rho * dW[0], drho_c2 = gas->gamma * dW[P];
    double dA[MAXV] = {0, 0, 0, 0, 0};
    for (int d = 0; d < D; ++d) {
        const double vn = W[1 + d], dvn = dW[1 + d];
        dA[0] += dvn * g[0 * D + d] + vn * dg[0 * D + d] + drho * g[(1 + d) * D + d] + rho * dg[(1 + d) * D + d];
        for (int k = 0; k < D; ++k) {
            double row = dvn * g[(1 + k) * D + d] + vn * dg[(1 + k) * D + d];
            if (k == d) row += dinv_rho * g[P * D + d] + inv_rho * dg[P * D + d];
            dA[1 + k] += row;
        }
        dA[P] += drho_c2 * g[(1 + d) * D + d] + rho_c2 * dg[(1 + d) * D + d] + dvn * g[P * D + d] + vn * dg[P * D + d];
    }
    for (int v = 0; v < D + 2; ++v) out[v] = dW[v] - 0.5 * dt * dA[v];
}

/* euler.hpp:239-272 half_step_adjoint(): adds J^T hbar to Wbar and gbar. */
void orc_half_step_adjoint(int D, const orc_gas* gas, const double* W, const double* g, double dt, const double* hbar, double* Wbar,
                           double* gbar) {
    const int P = D + 1;
    const double rho = W[0], inv_rho = 1.0 / W[0], rho_c2 = gas->gamma * W[P];
    double Abar[MAXV];
    for (int v = 0; v < D + 2; ++v) {
        Wbar[v] += hbar[v];
        Abar[v] = -0.5 * dt * hbar[v];
    }
    double inv_rhobar = 0.0, rho_c2bar = 0.0;
    for (int d = 0; d < D; ++d) {
        const double vn = W[1 + d];
        /* A[0] += vn g0_d + rho g(1+d)_d */
        Wbar[1 + d] += Abar[0] * g[0 * D + d];
        gbar[0 * D + d] += Abar[0] * vn;
        Wbar[0] += Abar[0] * g[(1 + d) * D + d];
        gbar[(1 + d) * D + d] += Abar[0] * rho;
        /* A[1+k] += vn g(1+k)_d (+ inv_rho gP_d if k == d) */
        for (int k = 0; k < D; ++k) {
            Wbar[1 + d] += Abar[1 + k] * g[(1 + k) * D + d];
            gbar[(1 + k) * D + d] += Abar[1 + k] * vn;
            if (k == d) {
                inv_rhobar += Abar[1 + k] * g[P * D + d];
                gbar[P * D + d] += Abar[1 + k] * inv_rho;
            }
        }
        /* A[P] += rho_c2 g(1+d)_d + vn gP_d */
        rho_c2bar += Abar[P] * g[(1 + d) * D + d];
        gbar[(1 + d) * D + d] += Abar[P] * rho_c2;
        Wbar[1 + d] += Abar[P] * g[P * D + d];
        gbar[P * D + d] += Abar[P] * vn;
    }
    Wbar[0] -= inv_rho * inv_rho * inv_rhobar;
    Wbar[P] += gas->gamma * rho_c2bar;
}

/* euler.hpp:358-375: the tangent of side_with() with vn's tangent given (w fixed). */
static void side_with_tangent(int D, const orc_gas* gas, const double* W, const double* dW, const double* n, const side_t* s,
                              double dvn, side_t* d) {
    const int P = D + 1;
    for (int k = 0; k < D; ++k) d->v[k] = dW[1 + k];
    d->vn = dvn;
    d->c = 0;
    d->E = dW[P] * gas->inverse_gm1 + dW[0] * (0.5 * dot(D, s->v, s->v)) + W[0] * dot(D, s->v, d->v);
    const double m = W[0] * s->vn, dm = dW[0] * s->vn + W[0] * dvn;
    d->F[0] = dm;
    for (int k = 0; k < D; ++k) d->F[1 + k] = dm * s->v[k] + m * d->v[k] + dW[P] * n[k];
    d->F[P] = (d->E + dW[P]) * s->vn + (s->E + W[P]) * dvn;
    d->U[0] = dW[0];
    for (int k = 0; k < D; ++k) d->U[1 + k] = dW[0] * s->v[k] + W[0] * d->v[k];
    d->U[P] = d->E;
}

/* euler.hpp:387-411: the adjoint of side_with(): from the bars of F, U, E and v, adds J^T to Wbar; returns vn's bar. */
static double side_with_adjoint(int D, const orc_gas* gas, const double* W, const double* n, const side_t* s, const double* Fbar,
                                const double* Ubar, double Ebar, double* vbar, double* Wbar) {
    const int P = D + 1;
    const double m = W[0] * s->vn;
    Ebar += Fbar[P] * s->vn + Ubar[P];
    double Pbar = Fbar[P] * s->vn, vnbar = Fbar[P] * (s->E + W[P]), mbar = Fbar[0], rhobar = Ubar[0];
    for (int k = 0; k < D; ++k) {
        Pbar += Fbar[1 + k] * n[k];
        mbar += Fbar[1 + k] * s->v[k];
        vbar[k] += Fbar[1 + k] * m + Ubar[1 + k] * W[0];
        rhobar += Ubar[1 + k] * s->v[k];
    }
    rhobar += mbar * s->vn;
    vnbar += mbar * W[0];
    Pbar += Ebar * gas->inverse_gm1;
    rhobar += Ebar * (0.5 * dot(D, s->v, s->v));
    for (int k = 0; k < D; ++k) vbar[k] = vbar[k] + s->v[k] * (Ebar * W[0]);
    Wbar[0] += rhobar;
    for (int k = 0; k < D; ++k) Wbar[1 + k] += vbar[k];
    Wbar[P] += Pbar;
    return vnbar;
}

/* euler.hpp:428-437: the adjoint of to_lab() (linear in F, w fixed). */
static void to_lab_adjoint(int D, const double* Fbar, const double* w, double* G) {
    const int P = D + 1;
    for (int v = 0; v < D + 2; ++v) G[v] = Fbar[v];
    G[0] += 0.5 * dot(D, w, w) * Fbar[P];
    for (int k = 0; k < D; ++k) {
        G[0] += w[k] * Fbar[1 + k];
        G[1 + k] += w[k] * Fbar[P];
    }
}

/* euler.hpp:649-695 HLLC::waves(): the wave speeds as flux() computes them and the branch pattern: bit 0 S_L from L, bit 1
 * S_R from L, bit 2 left, bit 3 star, bit 4 ok, bits 5-6 c^2 > 0 on L, R, bits 7-9 the 1e-20 guard on S_M's, rho*'s and
 * E*'s divisor. */
typedef struct {
    double vnL, vnR, inv_rhoL, inv_rhoR, c2L, c2R, cL, cR, SL, SR, denomL, denomR, den, SM;
    int pattern;
} hllc_waves;

static void hllc_waves_of(int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w, hllc_waves* a) {
    const int P = D + 1;
    const double eps = 1e-20;
    double rl[MAXD], rr[MAXD];
    for (int k = 0; k < D; ++k) {
        rl[k] = L[1 + k] - w[k];
        rr[k] = R[1 + k] - w[k];
    }
    a->vnL = dot(D, rl, n);
    a->vnR = dot(D, rr, n);
    a->inv_rhoL = 1.0 / L[0];
    a->inv_rhoR = 1.0 / R[0];
    a->c2L = gas->gamma * L[P] * a->inv_rhoL;
    a->c2R = gas->gamma * R[P] * a->inv_rhoR;
    const int ok = a->c2L >= 0 && a->c2R >= 0;
    a->cL = a->c2L > 0 ? sqrt(a->c2L) : 0.0;
    a->cR = a->c2R > 0 ? sqrt(a->c2R) : 0.0;
    const double sl = a->vnL - a->cL, sr = a->vnR - a->cR, tl = a->vnL + a->cL, tr = a->vnR + a->cR;
    a->SL = fmin(sl, sr);
    a->SR = fmax(tl, tr);
    a->denomL = L[0] * (a->SL - a->vnL);
    a->denomR = R[0] * (a->SR - a->vnR);
    a->den = a->denomL - a->denomR;
    a->SM = (R[P] - L[P] + a->denomL * a->vnL - a->denomR * a->vnR) / nonzero(a->den);
    const int left = ok && (a->SL >= 0 || a->SM >= 0);
    const int star = ok && a->SL < 0 && (a->SM >= 0 || a->SR > 0);
    const double S = left ? a->SL : a->SR, denom = left ? a->denomL : a->denomR;
    a->pattern = (sr < sl ? 0 : 1) | (tr > tl ? 0 : 2) | left << 2 | star << 3 | ok << 4 | (a->c2L > 0) << 5 | (a->c2R > 0) << 6 |
                 (fabs(a->den) < eps) << 7 | (star && fabs(S - a->SM) < eps) << 8 | (star && fabs(denom) < eps) << 9;
}

/* euler.hpp:697-745 HLLC::flux_tangent(): dF = J (dL, dR) at fixed n, w; a guarded divisor has derivative zero. F: the flux
 * from the same intermediates (the forward's operations). */
void orc_hllc_tangent(int D, const orc_gas* gas, const double* L, const double* R, const double* dL, const double* dR, const double* n,
                      const double* w, double* F, double* dF) {
    const int P = D + 1;
    hllc_waves a;
    hllc_waves_of(D, gas, L, R, n, w, &a);
    const int left = (a.pattern >> 2) & 1, star = (a.pattern >> 3) & 1;
    double dvl[MAXD], dvr[MAXD];
    for (int k = 0; k < D; ++k) {
        dvl[k] = dL[1 + k];
        dvr[k] = dR[1 + k];
    }
    const double dvnL = dot(D, dvl, n), dvnR = dot(D, dvr, n);
    const double dinvL = -a.inv_rhoL * a.inv_rhoL * dL[0], dinvR = -a.inv_rhoR * a.inv_rhoR * dR[0];
    const double dcL = a.c2L > 0 ? gas->gamma * (dL[P] * a.inv_rhoL + L[P] * dinvL) / (2 * a.cL) : 0.0;
    const double dcR = a.c2R > 0 ? gas->gamma * (dR[P] * a.inv_rhoR + R[P] * dinvR) / (2 * a.cR) : 0.0;
    const double dSL = (a.pattern & 1) ? dvnL - dcL : dvnR - dcR, dSR = (a.pattern & 2) ? dvnL + dcL : dvnR + dcR;
    const double ddenomL = dL[0] * (a.SL - a.vnL) + L[0] * (dSL - dvnL), ddenomR = dR[0] * (a.SR - a.vnR) + R[0] * (dSR - dvnR);
    const double dnum = dR[P] - dL[P] + ddenomL * a.vnL + a.denomL * dvnL - ddenomR * a.vnR - a.denomR * dvnR;
    const double dden = (a.pattern & 128) ? 0.0 : ddenomL - ddenomR, den = nonzero(a.den);
    const double dSM = (dnum - a.SM * dden) / den;
    double W[MAXV], dW[MAXV];
    for (int v = 0; v < D + 2; ++v) {
        W[v] = left ? L[v] : R[v];
        dW[v] = left ? dL[v] : dR[v];
    }
    const double vn = left ? a.vnL : a.vnR, dvn = left ? dvnL : dvnR;
    side_t s, ds;
    side_with(D, gas, W, n, w, vn, &s);
    side_with_tangent(D, gas, W, dW, n, &s, dvn, &ds);
    if (!star) {
        for (int v = 0; v < D + 2; ++v) {
            F[v] = s.F[v];
            dF[v] = ds.F[v];
        }
        to_lab(D, F, w);
        to_lab(D, dF, w);
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
    to_lab(D, F, w);
    to_lab(D, dF, w);
}

/* euler.hpp:746-832 HLLC::flux_adjoint(): adds J^T Fbar to Lbar and Rbar; only the case the forward took is differentiated. */
void orc_hllc_adjoint(int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w, const double* Fbar_lab,
                      double* Lbar, double* Rbar) {
    const int P = D + 1;
    hllc_waves a;
    hllc_waves_of(D, gas, L, R, n, w, &a);
    const int left = (a.pattern >> 2) & 1, star = (a.pattern >> 3) & 1;
    double Fbar[MAXV];
    to_lab_adjoint(D, Fbar_lab, w, Fbar);
    double W[MAXV], Wbar[MAXV] = {0, 0, 0, 0, 0}; /* the upwind side's bar, added to its side's at the end */
    for (int v = 0; v < D + 2; ++v) W[v] = left ? L[v] : R[v];
    const double vn = left ? a.vnL : a.vnR;
    side_t s;
    side_with(D, gas, W, n, w, vn, &s);
    double vnLbar = 0, vnRbar = 0;
    if (!star) {
        double Ubar0[MAXV] = {0, 0, 0, 0, 0}, vbar[MAXD] = {0, 0, 0};
        const double vnbar = side_with_adjoint(D, gas, W, n, &s, Fbar, Ubar0, 0.0, vbar, Wbar);
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
        double U_star[MAXV], vs[MAXD];
        U_star[0] = rho_star;
        for (int k = 0; k < D; ++k) {
            vs[k] = s.v[k] + n[k] * a2;
            U_star[1 + k] = rho_star * vs[k];
        }
        U_star[P] = rho_star * E_star;
        /* F = s.F + S (U* - s.U) */
        double sFbar[MAXV], sUbar[MAXV], Ustarbar[MAXV], Sbar = 0;
        for (int v = 0; v < D + 2; ++v) {
            sFbar[v] = Fbar[v];
            Sbar += Fbar[v] * (U_star[v] - s.U[v]);
            Ustarbar[v] = S * Fbar[v];
            sUbar[v] = -S * Fbar[v];
        }
        /* U* = (rho*, rho* v*, rho* E*) */
        double rho_starbar = Ustarbar[0] + Ustarbar[P] * E_star, vsbar[MAXD];
        for (int k = 0; k < D; ++k) {
            rho_starbar += Ustarbar[1 + k] * vs[k];
            vsbar[k] = Ustarbar[1 + k] * rho_star;
        }
        const double E_starbar = Ustarbar[P] * rho_star;
        /* v* = s.v + n a2; E* = s.E inv_side + a2 a3 */
        double svbar[MAXD], a2bar = E_starbar * a3;
        for (int k = 0; k < D; ++k) {
            svbar[k] = vsbar[k];
            a2bar += vsbar[k] * n[k];
        }
        const double sEbar = E_starbar * inv_side, inv_sidebar = E_starbar * s.E, a3bar = E_starbar * a2;
        /* a3 = SM + W_P inv_denom; a2 = SM - s.vn */
        double SMbar = a3bar + a2bar;
        Wbar[P] += a3bar * inv_denom;
        const double inv_denombar = a3bar * W[P];
        const double svnbar = -a2bar;
        /* inv_denom = 1 / nonzero(denom); rho* = denom / q; q = nonzero(S - SM) */
        double denombar = (a.pattern & 512) ? 0.0 : -inv_denom * inv_denom * inv_denombar;
        denombar += rho_starbar / q;
        const double qbar = -rho_starbar * rho_star / q;
        if (!(a.pattern & 256)) {
            Sbar += qbar;
            SMbar -= qbar;
        }
        const double vnbar = side_with_adjoint(D, gas, W, n, &s, sFbar, sUbar, sEbar, svbar, Wbar) + svnbar;
        if (left) {
            vnLbar += vnbar;
        } else {
            vnRbar += vnbar;
        }
        /* the side's 1 / rho, wave speed and denominator */
        double inv_rhoLbar = left ? inv_sidebar : 0.0, inv_rhoRbar = left ? 0.0 : inv_sidebar;
        double SLbar = left ? Sbar : 0.0, SRbar = left ? 0.0 : Sbar;
        double denomLbar = left ? denombar : 0.0, denomRbar = left ? 0.0 : denombar;
        /* SM = num / nonzero(den), num = R_P - L_P + denomL vnL - denomR vnR, den = denomL - denomR */
        const double den = nonzero(a.den);
        const double numbar = SMbar / den;
        const double denbar = (a.pattern & 128) ? 0.0 : -SMbar * a.SM / den;
        Rbar[P] += numbar;
        Lbar[P] -= numbar;
        denomLbar += numbar * a.vnL + denbar;
        vnLbar += numbar * a.denomL;
        denomRbar += -numbar * a.vnR - denbar;
        vnRbar -= numbar * a.denomR;
        /* denomL = L0 (SL - vnL), denomR = R0 (SR - vnR) */
        Lbar[0] += denomLbar * (a.SL - a.vnL);
        SLbar += denomLbar * L[0];
        vnLbar -= denomLbar * L[0];
        Rbar[0] += denomRbar * (a.SR - a.vnR);
        SRbar += denomRbar * R[0];
        vnRbar -= denomRbar * R[0];
        /* SL = min(vnL - cL, vnR - cR), SR = max(vnL + cL, vnR + cR): the side each took */
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
        /* c = sqrt(gamma P inv_rho) where c^2 > 0 */
        if (a.c2L > 0) {
            const double c2bar = cLbar / (2 * a.cL);
            Lbar[P] += c2bar * gas->gamma * a.inv_rhoL;
            inv_rhoLbar += c2bar * gas->gamma * L[P];
        }
        if (a.c2R > 0) {
            const double c2bar = cRbar / (2 * a.cR);
            Rbar[P] += c2bar * gas->gamma * a.inv_rhoR;
            inv_rhoRbar += c2bar * gas->gamma * R[P];
        }
        Lbar[0] -= inv_rhoLbar * a.inv_rhoL * a.inv_rhoL;
        Rbar[0] -= inv_rhoRbar * a.inv_rhoR * a.inv_rhoR;
    }
    for (int v = 0; v < D + 2; ++v) {
        if (left)
            Lbar[v] += Wbar[v];
        else
            Rbar[v] += Wbar[v];
    }
    /* vn = (v - w) . n */
    for (int k = 0; k < D; ++k) {
        Lbar[1 + k] += vnLbar * n[k];
        Rbar[1 + k] += vnRbar * n[k];
    }
}

/* ---- D = 1 (finite_volume_1d.hpp:710-935) ---- */

/* finite_volume_1d.hpp:234-253 clips(): for the face-centroid limiter, the last face that clipped variable v (-1: none),
 * to which bound, and whether the clip zeroed the slope. */
static void line_clips(const line_cell* cell, double gapL, double gapR, int v, int* last, int* upper, int* zero) {
    const line_cell* me = cell;
    const double dL = -gapL + cell[-1].c - me->c, dR = gapR + cell[1].c - me->c;
    const double Wi = me->W[v], left = cell[-1].W[v], right = cell[1].W[v];
    double g = ((right - Wi) - (left - Wi)) / (dR - dL);
    const double r[2] = {-0.5 * gapL - me->c, 0.5 * gapR - me->c};
    const double lo = fmin(left, right), hi = fmax(left, right);
    *last = -1;
    *upper = *zero = 0;
    for (int f = 0; f < 2; ++f) {
        const double dp = g * r[f];
        if (dp > 0 && Wi + dp > hi) {
            g = g * (hi > Wi ? (hi - Wi) / dp : 0.0);
            *last = f, *upper = 1, *zero = !(hi > Wi);
        } else if (dp < 0 && Wi + dp < lo) {
            g = g * (lo < Wi ? (lo - Wi) / dp : 0.0);
            *last = f, *upper = 0, *zero = !(lo < Wi);
        }
    }
}

/* finite_volume_1d.hpp:209-233 gradient_tangent(): d g of line_gradient() for the tangents of W[v] of the left neighbour, the
 * cell and the right neighbour. A clipped slope is (bound - W_i) / r of its last clip; the bound is the neighbour fmax /
 * fmin returned. */
static double line_gradient_tangent(const orc_line* l, const line_cell* cell, double gapL, double gapR, int v, double dl, double di,
                                    double dr) {
    const line_cell* me = cell;
    const double dL = -gapL + cell[-1].c - me->c, dR = gapR + cell[1].c - me->c;
    const double Wi = me->W[v], left = cell[-1].W[v], right = cell[1].W[v];
    if (l->limiter == 2 || l->limiter == 3) {
        const double sL = (left - Wi) / dL, sR = (right - Wi) / dR, dsL = (dl - di) / dL, dsR = (dr - di) / dR;
        if (!(sL * sR > 0)) return 0.0;
        if (l->limiter == 3) {
            const double sum = sL + sR, g = 2 * sL * sR / sum;
            return (2 * (dsL * sR + sL * dsR) - g * (dsL + dsR)) / sum;
        }
        return fabs(sL) < fabs(sR) ? dsL : dsR;
    }
    const double dg = (dr - dl) / (dR - dL);
    if (l->limiter == 1) return dg;
    int last, upper, zero;
    line_clips(cell, gapL, gapR, v, &last, &upper, &zero);
    if (last < 0) return dg;
    if (zero) return 0.0;
    const double r[2] = {-0.5 * gapL - me->c, 0.5 * gapR - me->c};
    const int from_left = upper ? !(right > left) : !(right < left);
    return ((from_left ? dl : dr) - di) / r[last];
}

/* finite_volume_1d.hpp:255-290 gradient_adjoint(): adds J^T gbar to the bars of W[v] of the left neighbour, the cell and
 * the right neighbour. */
static void line_gradient_adjoint(const orc_line* l, const line_cell* cell, double gapL, double gapR, int v, double gbar, double* bl,
                                  double* bi, double* br) {
    const line_cell* me = cell;
    const double dL = -gapL + cell[-1].c - me->c, dR = gapR + cell[1].c - me->c;
    const double Wi = me->W[v], left = cell[-1].W[v], right = cell[1].W[v];
    if (l->limiter == 2 || l->limiter == 3) {
        const double sL = (left - Wi) / dL, sR = (right - Wi) / dR;
        if (!(sL * sR > 0)) return;
        double aL, aR; /* d g = aL dsL + aR dsR */
        if (l->limiter == 3) {
            const double sum = sL + sR, g = 2 * sL * sR / sum;
            aL = (2 * sR - g) / sum, aR = (2 * sL - g) / sum;
        } else {
            const int pick_left = fabs(sL) < fabs(sR);
            aL = pick_left ? 1.0 : 0.0, aR = pick_left ? 0.0 : 1.0;
        }
        *bl += gbar * aL / dL, *bi -= gbar * aL / dL;
        *br += gbar * aR / dR, *bi -= gbar * aR / dR;
        return;
    }
    int last = -1, upper = 0, zero = 0;
    if (l->limiter != 1) line_clips(cell, gapL, gapR, v, &last, &upper, &zero);
    if (last < 0) {
        *br += gbar / (dR - dL), *bl -= gbar / (dR - dL);
        return;
    }
    if (zero) return;
    const double r[2] = {-0.5 * gapL - me->c, 0.5 * gapR - me->c};
    const int from_left = upper ? !(right > left) : !(right < left);
    if (from_left)
        *bl += gbar / r[last];
    else
        *br += gbar / r[last];
    *bi -= gbar / r[last];
}

/* finite_volume_1d.hpp:117-128 as a map of tangents (and its own adjoint): an outflow ghost copies the end cell, a
 * reflective one mirrors its velocity. */
static void line_ghost_tangent(const orc_line* l, const double* dW, int left, double* out) {
    const int wall = left ? l->wall_left : l->wall_right;
    for (int v = 0; v < 3; ++v) out[v] = dW[v];
    if (wall == 0) out[1] = -out[1];
}
static void mirror3(const double* a, double* out) {
    out[0] = a[0];
    out[1] = -a[1];
    out[2] = a[2];
}

static void line_slope(const orc_line* l, const line_cell cell[5], const double gap[4], int k, double* g) {
    g[0] = g[1] = g[2] = 0;
    if (l->order == 1 && cell[k].real)
        for (int v = 0; v < 3; ++v) g[v] = line_gradient(l, cell + k, gap[k - 1], gap[k], v);
}
static void line_slope_tangent(const orc_line* l, const line_cell cell[5], const double gap[4], double dW[5][3], int k, double* dg) {
    dg[0] = dg[1] = dg[2] = 0;
    if (l->order == 1 && cell[k].real)
        for (int v = 0; v < 3; ++v)
            dg[v] = line_gradient_tangent(l, cell + k, gap[k - 1], gap[k], v, dW[k - 1][v], dW[k][v], dW[k + 1][v]);
}

/* finite_volume_1d.hpp:730-743 face_value(): line_face_state() with its fallback flag. */
static void line_face_value(const orc_line* l, const double* W, const double* g, double r, double dt, double* out, int* fallback) {
    *fallback = 0;
    if (l->order == 0) {
        for (int v = 0; v < 3; ++v) out[v] = W[v];
        return;
    }
    orc_half_step(1, &l->gas, W, g, dt, out);
    for (int v = 0; v < 3; ++v) out[v] += g[v] * r;
    *fallback = !(out[0] > 0 && out[2] > 0);
    if (*fallback)
        for (int v = 0; v < 3; ++v) out[v] = W[v];
}
static void line_face_value_tangent(const orc_line* l, const double* W, const double* dW, const double* g, const double* dg, double r,
                                    double dt, int fallback, double* out) {
    if (l->order == 0 || fallback) {
        for (int v = 0; v < 3; ++v) out[v] = dW[v];
        return;
    }
    orc_half_step_tangent(1, &l->gas, W, dW, g, dg, dt, out);
    for (int v = 0; v < 3; ++v) out[v] += dg[v] * r;
}
static void line_face_value_adjoint(const orc_line* l, const double* W, const double* g, double r, double dt, int fallback,
                                    const double* Wfbar, double* Wbar, double* gbar) {
    if (l->order == 0 || fallback) {
        for (int v = 0; v < 3; ++v) Wbar[v] += Wfbar[v];
        return;
    }
    double gb[3];
    for (int v = 0; v < 3; ++v) gb[v] = Wfbar[v] * r;
    orc_half_step_adjoint(1, &l->gas, W, g, dt, Wfbar, Wbar, gb);
    for (int v = 0; v < 3; ++v) gbar[v] += gb[v];
}

/* finite_volume_1d.hpp:766-792 face_tangent(): the flux of the face between stencil positions a and a + 1, and its tangent. */
static void line_face_tangent(const orc_line* l, const line_cell cell[5], const double gap[4], double dW[5][3], int a, double dt, double* F,
                              double* dF) {
    const int b = a + 1;
    double ga[3], gb[3], dga[3], dgb[3];
    line_slope(l, cell, gap, a, ga);
    line_slope(l, cell, gap, b, gb);
    line_slope_tangent(l, cell, gap, dW, a, dga);
    line_slope_tangent(l, cell, gap, dW, b, dgb);
    const double ra = 0.5 * gap[a] - cell[a].c, rb = -0.5 * gap[a] - cell[b].c;
    double L[3] = {0, 0, 0}, R[3] = {0, 0, 0}, dLs[3] = {0, 0, 0}, dRs[3] = {0, 0, 0};
    int fallback;
    if (cell[a].real) {
        line_face_value(l, cell[a].W, ga, ra, dt, L, &fallback);
        line_face_value_tangent(l, cell[a].W, dW[a], ga, dga, ra, dt, fallback, dLs);
    }
    if (cell[b].real) {
        line_face_value(l, cell[b].W, gb, rb, dt, R, &fallback);
        line_face_value_tangent(l, cell[b].W, dW[b], gb, dgb, rb, dt, fallback, dRs);
    }
    if (!cell[a].real) {
        line_ghost_face(l, R, cell[b].W, 1, L);
        if (l->wall_left == 0)
            mirror3(dRs, dLs);
        else
            line_ghost_tangent(l, dW[b], 1, dLs);
    }
    if (!cell[b].real) {
        line_ghost_face(l, L, cell[a].W, 0, R);
        if (l->wall_right == 0)
            mirror3(dLs, dRs);
        else
            line_ghost_tangent(l, dW[a], 0, dRs);
    }
    const double one = 1.0, zero = 0.0;
    riemann_tangent(l->riemann, 1, &l->gas, L, R, dLs, dRs, &one, &zero, F, dF);
}

/* finite_volume_1d.hpp:794-830 face_adjoint(): adds J^T Fbar of the face between positions a and a + 1 to the bars of the
 * stencil's W (a ghost position's bar is mapped to its wall's end cell by the caller). */
static void line_face_adjoint(const orc_line* l, const line_cell cell[5], const double gap[4], int a, double dt, const double* Fbar,
                              double Wbar[5][3]) {
    const int b = a + 1;
    double ga[3], gb[3];
    line_slope(l, cell, gap, a, ga);
    line_slope(l, cell, gap, b, gb);
    const double ra = 0.5 * gap[a] - cell[a].c, rb = -0.5 * gap[a] - cell[b].c;
    double L[3] = {0, 0, 0}, R[3] = {0, 0, 0};
    int fallback_a = 0, fallback_b = 0;
    if (cell[a].real) line_face_value(l, cell[a].W, ga, ra, dt, L, &fallback_a);
    if (cell[b].real) line_face_value(l, cell[b].W, gb, rb, dt, R, &fallback_b);
    if (!cell[a].real) line_ghost_face(l, R, cell[b].W, 1, L);
    if (!cell[b].real) line_ghost_face(l, L, cell[a].W, 0, R);
    double Lbar[3] = {0, 0, 0}, Rbar[3] = {0, 0, 0};
    const double one = 1.0, zero = 0.0;
    riemann_adjoint(l->riemann, 1, &l->gas, L, R, &one, &zero, Fbar, Lbar, Rbar);
    /* the ghost states in reverse order of the forward: R from L, then L from R */
    if (!cell[b].real) {
        double back[3];
        if (l->wall_right == 0) {
            mirror3(Rbar, back);
            for (int v = 0; v < 3; ++v) Lbar[v] += back[v];
        } else {
            line_ghost_tangent(l, Rbar, 0, back);
            for (int v = 0; v < 3; ++v) Wbar[a][v] += back[v];
        }
    }
    if (!cell[a].real) {
        double back[3];
        if (l->wall_left == 0) {
            mirror3(Lbar, back);
            for (int v = 0; v < 3; ++v) Rbar[v] += back[v];
        } else {
            line_ghost_tangent(l, Lbar, 1, back);
            for (int v = 0; v < 3; ++v) Wbar[b][v] += back[v];
        }
    }
    double gabar[3] = {0, 0, 0}, gbbar[3] = {0, 0, 0};
    if (cell[a].real) line_face_value_adjoint(l, cell[a].W, ga, ra, dt, fallback_a, Lbar, Wbar[a], gabar);
    if (cell[b].real) line_face_value_adjoint(l, cell[b].W, gb, rb, dt, fallback_b, Rbar, Wbar[b], gbbar);
    if (l->order == 1) {
        if (cell[a].real)
            for (int v = 0; v < 3; ++v)
                line_gradient_adjoint(l, cell + a, gap[a - 1], gap[a], v, gabar[v], &Wbar[a - 1][v], &Wbar[a][v], &Wbar[a + 1][v]);
        if (cell[b].real)
            for (int v = 0; v < 3; ++v)
                line_gradient_adjoint(l, cell + b, gap[b - 1], gap[b], v, gbbar[v], &Wbar[b - 1][v], &Wbar[b][v], &Wbar[b + 1][v]);
    }
}

/* finite_volume_1d.hpp:831-848 load_tangent(): the W tangents of slot s's stencil from dq (ghosts through the ghost map). */
static void line_load_tangent(const orc_line* l, const double* dq, int s, double dW[5][3]) {
    int slot[5];
    for (int k = 0; k < 5; ++k) {
        slot[k] = line_slot(l, s, k - 2);
        dW[k][0] = dW[k][1] = dW[k][2] = 0;
        if (slot[k] < 0) continue;
        const int i = slot[k];
        const double inv = 1.0 / orc_line_volume(l, i);
        const double U[3] = {l->q[i * 3] * inv, l->q[i * 3 + 1] * inv, l->q[i * 3 + 2] * inv};
        const double dU[3] = {dq[i * 3] * inv, dq[i * 3 + 1] * inv, dq[i * 3 + 2] * inv};
        orc_primitive_tangent(1, &l->gas, U, dU, dW[k]);
    }
    for (int k = 0; k < 5; ++k) {
        if (slot[k] >= 0) continue;
        const int end = k < 2 ? k + 1 : k - 1;
        if (slot[end] >= 0) line_ghost_tangent(l, dW[end], k < 2, dW[k]);
    }
}

/* finite_volume_1d.hpp:850-880 step_tangent(): the step (order 0 or 1, its walls) and dq -> J dq, dt fixed; q and dq are
 * advanced in place (dq: a tangent of the extensive state). */
void orc_line_step_tangent(orc_line* l, double dt, double* dq) {
    double* dnext = (double*)malloc(sizeof(double) * 3 * (size_t)l->n);
#pragma omp parallel for schedule(static) if (l->n >= LINE_PARALLEL)
    for (int s = 0; s < l->n; ++s) {
        line_cell cell[5];
        double gap[4], dW[5][3], F0[3], F1[3], dF0[3], dF1[3];
        line_load(l, s, cell, gap);
        line_load_tangent(l, dq, s, dW);
        line_face_tangent(l, cell, gap, dW, 1, dt, F0, dF0);
        line_face_tangent(l, cell, gap, dW, 2, dt, F1, dF1);
        for (int v = 0; v < 3; ++v) {
            l->next[s * 3 + v] = l->q[s * 3 + v] - dt * (F1[v] - F0[v]);
            dnext[s * 3 + v] = dq[s * 3 + v] - dt * (dF1[v] - dF0[v]);
        }
    }
    memcpy(dq, dnext, sizeof(double) * 3 * (size_t)l->n);
    free(dnext);
    line_swap(l);
}

/* finite_volume_1d.hpp:882-933 step_adjoint(): qbar (after the step) -> J^T qbar (before it), in place; q is the state at the
 * start of the step and stays. Per slot j: the adjoints of the faces whose fluxes read W_j (the left faces of slots
 * j - 1 .. j + 2, and the right wall's face), then the primitive adjoint. */
void orc_line_step_adjoint(orc_line* l, double dt, double* qbar) {
    const int n = l->n;
    double* out = (double*)malloc(sizeof(double) * 3 * (size_t)n);
#pragma omp parallel for schedule(static) if (n >= LINE_PARALLEL)
    for (int j = 0; j < n; ++j) {
        double Wj_bar[3] = {0, 0, 0};
        for (int d = -1; d <= 3; ++d) {
            int s, a;
            if (d <= 2) {
                s = line_slot(l, j, d);
                a = 1;
            } else {
                if (l->periodic || j < n - 2) continue;
                s = n - 1;
                a = 2; /* the right wall's face */
            }
            if (s < 0) continue;
            line_cell cell[5];
            double gap[4];
            line_load(l, s, cell, gap);
            const int left = line_slot(l, s, a - 2), right = line_slot(l, s, a - 1);
            double Fbar[3];
            for (int v = 0; v < 3; ++v)
                Fbar[v] = (right >= 0 ? dt * qbar[right * 3 + v] : 0.0) - (left >= 0 ? dt * qbar[left * 3 + v] : 0.0);
            double Wbar[5][3];
            memset(Wbar, 0, sizeof Wbar);
            line_face_adjoint(l, cell, gap, a, dt, Fbar, Wbar);
            for (int k = 0; k < 5; ++k) {
                const int t = line_slot(l, s, k - 2);
                if (t == j) {
                    for (int v = 0; v < 3; ++v) Wj_bar[v] += Wbar[k][v];
                } else if (t < 0) {
                    const int end = k < 2 ? k + 1 : k - 1;
                    if (line_slot(l, s, end - 2) == j) {
                        double back[3];
                        line_ghost_tangent(l, Wbar[k], k < 2, back);
                        for (int v = 0; v < 3; ++v) Wj_bar[v] += back[v];
                    }
                }
            }
        }
        const double inv = 1.0 / orc_line_volume(l, j);
        const double U[3] = {l->q[j * 3] * inv, l->q[j * 3 + 1] * inv, l->q[j * 3 + 2] * inv};
        double Ubar[3];
        orc_primitive_adjoint(1, &l->gas, U, Wj_bar, Ubar);
        for (int v = 0; v < 3; ++v) out[j * 3 + v] = qbar[j * 3 + v] + Ubar[v] * inv;
    }
    memcpy(qbar, out, sizeof(double) * 3 * (size_t)n);
    free(out);
}

/* ---- D = 2, 3: the static order-1 step (finite_volume.hpp:1670-2231) ---- */

/* Cholesky of M (upper triangle) and the solve with its factor, in Sweep::solve's operation order (:590-615). */
static void cholesky(int D, double M[MAXD + 1][MAXD + 1], double L[MAXD + 1][MAXD + 1]) {
    const int U = D + 1;
    memset(L, 0, sizeof(double) * (MAXD + 1) * (MAXD + 1));
    for (int p = 0; p < U; ++p)
        for (int s = 0; s <= p; ++s) {
            double sum = M[s][p];
            for (int m = 0; m < s; ++m) sum -= L[p][m] * L[s][m];
            L[p][s] = p == s ? sqrt(sum) : sum / L[s][s];
        }
}
static void solve_factored(int D, double L[MAXD + 1][MAXD + 1], double b[MAXV][MAXD + 1], double x[MAXV][MAXD + 1]) {
    const int U = D + 1, V = D + 2;
    for (int v = 0; v < V; ++v) {
        double z[MAXD + 1];
        for (int p = 0; p < U; ++p) {
            double sum = b[v][p];
            for (int m = 0; m < p; ++m) sum -= L[p][m] * z[m];
            z[p] = sum / L[p][p];
        }
        for (int p = U - 1; p >= 0; --p) {
            double sum = z[p];
            for (int m = p + 1; m < U; ++m) sum -= L[m][p] * x[v][m];
            x[v][p] = sum / L[p][p];
        }
    }
}

/* finite_volume.hpp:709-799 limited_core(): what the limiter did per variable. last: the list position (from the cell's
 * first entry) of the last clipping face (-1: none); extreme: the face whose neighbour holds the bound (the first entry in
 * list order with the extreme value); offset: the last clipping face's offset. */
typedef struct {
    int last[MAXV], upper[MAXV], zero[MAXV], extreme[MAXV];
    double offset[MAXV][MAXD], bound[MAXV], dbound[MAXV];
} limit_pattern;

/* The cell's fit and limiter as sweep_cell() computes them (the same bits for g), with what the derivatives need: the
 * unlimited gradient g0 and, with dW (N x V tangents of the prims), its tangent dg0; M's Cholesky factor and first row M0
 * (M0[p] = sum_k |A_k| X_k[p]); the limiter's pattern. */
static void fit_cell_x(const orc_fv* fv, const orc_geometry* geo, int i, const double* dW, double* g, double* g0, double* dg0,
                       double L[MAXD + 1][MAXD + 1], double* M0, limit_pattern* pat, double* intercept);
static void fit_cell(const orc_fv* fv, const orc_geometry* geo, int i, const double* dW, double* g, double* g0, double* dg0,
                     double L[MAXD + 1][MAXD + 1], double* M0, limit_pattern* pat) {
    fit_cell_x(fv, geo, i, dW, g, g0, dg0, L, M0, pat, NULL);
}
/* fit_cell() with the fit's intercepts (intercept: V, or NULL). */
static void fit_cell_x(const orc_fv* fv, const orc_geometry* geo, int i, const double* dW, double* g, double* g0, double* dg0,
                       double L[MAXD + 1][MAXD + 1], double* M0, limit_pattern* pat, double* intercept) {
    const int D = fv->D, V = D + 2, U = D + 1, PW = V + 2;
    const double* Wi = fv->prims + (size_t)i * PW;
    double surface = 0, smallest = INFINITY;
    double lo[MAXV], hi[MAXV], dlo[MAXV], dhi[MAXV];
    int lo_f[MAXV], hi_f[MAXV];
    for (int v = 0; v < V; ++v) {
        lo[v] = DBL_MAX;
        hi[v] = -DBL_MAX;
        dlo[v] = dhi[v] = 0;
        lo_f[v] = hi_f[v] = -1;
    }
    double M[MAXD + 1][MAXD + 1], b[MAXV][MAXD + 1], db[MAXV][MAXD + 1];
    memset(M, 0, sizeof M);
    memset(b, 0, sizeof b);
    memset(db, 0, sizeof db);
    const double* centroid = geo->centroid + (size_t)i * D;
    const int first = geo->start[i], end = geo->start[i + 1];
    for (int k = first; k < end; ++k) {
        const int e = geo->list[k], f = list_face(e);
        const int j = e >= 0 ? geo->neighbour[f] : geo->owner[f];
        const double* Wj = fv->prims + (size_t)j * PW;
        double d[MAXD], X[MAXD + 1];
        separation(geo, i, j, d);
        const double w = magnitude_of(geo, f);
        surface += w;
        smallest = fmin(smallest, w);
        X[0] = 1;
        for (int p = 1; p < U; ++p) X[p] = d[p - 1] + geo->centroid[(size_t)j * D + p - 1] - centroid[p - 1];
        for (int p = 0; p < U; ++p)
            for (int s = p; s < U; ++s) M[p][s] += w * X[p] * X[s];
        for (int v = 0; v < V; ++v) {
            const double dWv = Wj[v] - Wi[v];
            for (int p = 0; p < U; ++p) b[v][p] += w * X[p] * dWv;
            if (dW) {
                const double ddW = w * (dW[(size_t)j * V + v] - dW[(size_t)i * V + v]);
                for (int p = 0; p < U; ++p) db[v][p] += X[p] * ddW;
            }
            if (Wj[v] < lo[v]) lo[v] = Wj[v], lo_f[v] = k - first, dlo[v] = dW ? dW[(size_t)j * V + v] : 0.0;
            if (Wj[v] > hi[v]) hi[v] = Wj[v], hi_f[v] = k - first, dhi[v] = dW ? dW[(size_t)j * V + v] : 0.0;
        }
    }
    const double tiny = 1e-6 * surface;
    const int any_tiny = smallest < tiny;
    if (any_tiny) {
        for (int v = 0; v < V; ++v) {
            lo[v] = DBL_MAX;
            hi[v] = -DBL_MAX;
            lo_f[v] = hi_f[v] = -1;
        }
        for (int k = first; k < end; ++k) {
            const int e = geo->list[k], f = list_face(e);
            const int j = e >= 0 ? geo->neighbour[f] : geo->owner[f];
            if (magnitude_of(geo, f) < tiny) continue;
            const double* Wj = fv->prims + (size_t)j * PW;
            for (int v = 0; v < V; ++v) {
                if (Wj[v] < lo[v]) lo[v] = Wj[v], lo_f[v] = k - first, dlo[v] = dW ? dW[(size_t)j * V + v] : 0.0;
                if (Wj[v] > hi[v]) hi[v] = Wj[v], hi_f[v] = k - first, dhi[v] = dW ? dW[(size_t)j * V + v] : 0.0;
            }
        }
    }
    cholesky(D, M, L);
    for (int p = 0; p < U; ++p) M0[p] = M[0][p];
    double xs[MAXV][MAXD + 1];
    solve_factored(D, L, b, xs);
    for (int v = 0; v < V; ++v)
        for (int a = 0; a < D; ++a) g0[v * D + a] = g[v * D + a] = xs[v][1 + a];
    if (intercept)
        for (int v = 0; v < V; ++v) intercept[v] = xs[v][0];
    if (dW) {
        solve_factored(D, L, db, xs);
        for (int v = 0; v < V; ++v)
            for (int a = 0; a < D; ++a) dg0[v * D + a] = xs[v][1 + a];
    }
    for (int v = 0; v < V; ++v) {
        pat->last[v] = pat->extreme[v] = -1;
        pat->upper[v] = pat->zero[v] = 0;
    }
    if (fv->limiter == 0)
        for (int k = first; k < end; ++k) {
            const int f = list_face(geo->list[k]);
            if (any_tiny && magnitude_of(geo, f) < tiny) continue;
            double offset[MAXD];
            face_offset(geo, f, i, offset);
            for (int v = 0; v < V; ++v) {
                double* gv = g + v * D;
                const double dp = dot(D, gv, offset);
                if (dp > 0 && Wi[v] + dp > hi[v]) {
                    const double s = hi[v] > Wi[v] ? (hi[v] - Wi[v]) / dp : 0.0;
                    for (int a = 0; a < D; ++a) gv[a] = gv[a] * s;
                    pat->last[v] = k - first, pat->upper[v] = 1, pat->zero[v] = !(hi[v] > Wi[v]);
                    memcpy(pat->offset[v], offset, sizeof(double) * D);
                } else if (dp < 0 && Wi[v] + dp < lo[v]) {
                    const double s = lo[v] < Wi[v] ? (lo[v] - Wi[v]) / dp : 0.0;
                    for (int a = 0; a < D; ++a) gv[a] = gv[a] * s;
                    pat->last[v] = k - first, pat->upper[v] = 0, pat->zero[v] = !(lo[v] < Wi[v]);
                    memcpy(pat->offset[v], offset, sizeof(double) * D);
                }
            }
        }
    for (int v = 0; v < V; ++v) {
        pat->extreme[v] = pat->last[v] < 0 ? -1 : pat->upper[v] ? hi_f[v] : lo_f[v];
        pat->bound[v] = pat->upper[v] ? hi[v] : lo[v];
        pat->dbound[v] = pat->upper[v] ? dhi[v] : dlo[v];
    }
}

/* finite_volume.hpp:809-831 limited_gradient_tangent(): after the last clip g = g0 (bound - W_i) / (g0 . o), whatever
 * clips came before (each only rescales g). */
static void limited_tangent(int D, const double* Wi, const double* dWi, const double* g0, const double* dg0, const limit_pattern* pat,
                            double* dg) {
    const int V = D + 2;
    for (int v = 0; v < V; ++v) {
        const double* gv = g0 + v * D;
        const double* dgv = dg0 + v * D;
        if (pat->last[v] < 0) {
            for (int a = 0; a < D; ++a) dg[v * D + a] = dgv[a];
        } else if (pat->zero[v]) {
            for (int a = 0; a < D; ++a) dg[v * D + a] = 0.0;
        } else {
            const double* o = pat->offset[v];
            const double den = dot(D, gv, o);
            const double s = (pat->bound[v] - Wi[v]) / den, ds = (pat->dbound[v] - dWi[v] - s * dot(D, dgv, o)) / den;
            for (int a = 0; a < D; ++a) dg[v * D + a] = dgv[a] * s + gv[a] * ds;
        }
    }
}

/* One face's two sides as the static flux computes them (face_flux, pass -1), with their fallback flags and offsets. */
static void face_sides(const orc_fv* fv, const orc_geometry* geo, int f, double dt, double* L, double* R, int* fallback, double* ro,
                       double* rj, double* n, double* magnitude) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    const int o = geo->owner[f], j = geo->neighbour[f];
    const double* a = geo->area + (size_t)f * D;
    *magnitude = sqrt(dot(D, a, a));
    const double inv = 1.0 / *magnitude;
    for (int k = 0; k < D; ++k) n[k] = a[k] * inv;
    face_offset(geo, f, o, ro);
    face_offset(geo, f, j, rj);
    const double half = 0.5 * dt, wf = 0.0;
    for (int k = 0; k < D; ++k) {
        ro[k] = ro[k] + wf * half;
        rj[k] = rj[k] + wf * half;
    }
    const int cells[2] = {o, j};
    const double* rs[2] = {ro, rj};
    double* out[2] = {L, R};
    for (int s = 0; s < 2; ++s) {
        const double* W = fv->prims + (size_t)cells[s] * PW;
        const double* g = fv->grad + (size_t)cells[s] * V * D;
        orc_half_step(D, &fv->gas, W, g, dt, out[s]);
        for (int v = 0; v < V; ++v) out[s][v] += dot(D, g + v * D, rs[s]);
        fallback[s] = !(out[s][0] > 0 && out[s][D + 1] > 0);
        if (fallback[s])
            for (int v = 0; v < V; ++v) out[s][v] = W[v];
    }
}

static int static_derivatives_ok(const orc_fv* fv) {
    return !fv->moving && fv->order == 1 && fv->D >= 2 && fv->precision == 0 && !(fv->hll_compression > 0);
}

/* finite_volume.hpp:2039-2074 step_tangent(): the static step of q and dq -> J dq (dq: N x V, extensive), dt fixed.
 * Returns 0, or -1 outside a static order-1 HLLC mesh step. */
int orc_step_tangent(orc_fv* fv, const orc_geometry* geo, double dt, double* dq) {
    if (!static_derivatives_ok(fv)) return -1;
    const int D = fv->D, V = D + 2, N = fv->N, F = geo->F, PW = V + 2;
    orc_primitives(fv, geo);
    double* dW = (double*)malloc(sizeof(double) * (size_t)N * V);
    double* dg = (double*)malloc(sizeof(double) * (size_t)N * V * D);
    double* dflux = (double*)malloc(sizeof(double) * (size_t)F * V);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        const double inv = 1.0 / geo->volume[i];
        double U[MAXV], dU[MAXV];
        for (int v = 0; v < V; ++v) {
            U[v] = fv->q[(size_t)i * V + v] * inv;
            dU[v] = dq[(size_t)i * V + v] * inv;
        }
        orc_primitive_tangent(D, &fv->gas, U, dU, dW + (size_t)i * V);
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        double g0[MAXV * MAXD], dg0[MAXV * MAXD], L[MAXD + 1][MAXD + 1], M0[MAXD + 1];
        limit_pattern pat;
        fit_cell(fv, geo, i, dW, fv->grad + (size_t)i * V * D, g0, dg0, L, M0, &pat);
        limited_tangent(D, fv->prims + (size_t)i * PW, dW + (size_t)i * V, g0, dg0, &pat, dg + (size_t)i * V * D);
    }
#pragma omp parallel for schedule(static)
    for (int f = 0; f < F; ++f) {
        double L[MAXV], R[MAXV], ro[MAXD], rj[MAXD], n[MAXD], magnitude, dLs[MAXV], dRs[MAXV];
        int fallback[2];
        double* Fv = fv->flux + (size_t)f * V;
        double* dF = dflux + (size_t)f * V;
        for (int v = 0; v < V; ++v) Fv[v] = dF[v] = 0;
        const double* a = geo->area + (size_t)f * D;
        if (sqrt(dot(D, a, a)) == 0) continue;
        face_sides(fv, geo, f, dt, L, R, fallback, ro, rj, n, &magnitude);
        const int cells[2] = {geo->owner[f], geo->neighbour[f]};
        const double* rs[2] = {ro, rj};
        double* outs[2] = {dLs, dRs};
        for (int s = 0; s < 2; ++s) {
            const int i = cells[s];
            if (fallback[s]) {
                for (int v = 0; v < V; ++v) outs[s][v] = dW[(size_t)i * V + v];
                continue;
            }
            const double* gi = fv->grad + (size_t)i * V * D;
            const double* dgi = dg + (size_t)i * V * D;
            orc_half_step_tangent(D, &fv->gas, fv->prims + (size_t)i * PW, dW + (size_t)i * V, gi, dgi, dt, outs[s]);
            for (int v = 0; v < V; ++v) outs[s][v] += dot(D, dgi + v * D, rs[s]);
        }
        const double w0[MAXD] = {0, 0, 0};
        riemann_tangent(fv->riemann, D, &fv->gas, L, R, dLs, dRs, n, w0, Fv, dF);
        for (int v = 0; v < V; ++v) {
            Fv[v] *= magnitude;
            dF[v] *= magnitude;
        }
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        double sum[MAXV] = {0, 0, 0, 0, 0}, dsum[MAXV] = {0, 0, 0, 0, 0};
        for (int k = geo->start[i]; k < geo->start[i + 1]; ++k) {
            const int e = geo->list[k], f = list_face(e);
            const double sign = e >= 0 ? 1.0 : -1.0;
            for (int v = 0; v < V; ++v) {
                sum[v] += sign * fv->flux[(size_t)f * V + v];
                dsum[v] += sign * dflux[(size_t)f * V + v];
            }
        }
        for (int v = 0; v < V; ++v) {
            fv->q[(size_t)i * V + v] = fv->q[(size_t)i * V + v] - dt * sum[v];
            dq[(size_t)i * V + v] = dq[(size_t)i * V + v] - dt * dsum[v];
        }
    }
    free(dW);
    free(dg);
    free(dflux);
    fv->prims_ready = 0;
    return 0;
}

/* finite_volume.hpp:2201-2231 step_adjoint() (AdjointFaces :2078-2102, AdjointCell :2107-2169, AdjointGather :2171-2195):
 * qbar after the static step -> J^T qbar before it, in place; q (the state before the step) stays. Per face: the
 * Riemann adjoint of the flux's bar; per cell, in list order: its sides' adjoints (half step and g.r) into W bar and g bar,
 * then the limited fit's adjoint (the bound's bar and b bar, which the neighbours collect); per cell, in list order: the
 * neighbours' shares, then the primitive adjoint. Deterministic, no atomics. Returns 0, or -1 as orc_step_tangent. */
int orc_step_adjoint(orc_fv* fv, const orc_geometry* geo, double dt, double* qbar) {
    if (!static_derivatives_ok(fv)) return -1;
    const int D = fv->D, V = D + 2, U = D + 1, N = fv->N, F = geo->F, PW = V + 2;
    orc_primitives(fv, geo);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) sweep_cell(fv, geo, i);
    double* sidebar = (double*)malloc(sizeof(double) * (size_t)F * 2 * V); /* L bar, R bar per face */
    double* Wbar = (double*)malloc(sizeof(double) * (size_t)N * V);
    double* bbar = (double*)malloc(sizeof(double) * (size_t)N * V * U);
    double* boundbar = (double*)malloc(sizeof(double) * (size_t)N * V);
    int* extreme = (int*)malloc(sizeof(int) * (size_t)N * V); /* face id of the bound's entry, -1 none */
#pragma omp parallel for schedule(static)
    for (int f = 0; f < F; ++f) {
        double* Lbar = sidebar + (size_t)f * 2 * V;
        double* Rbar = Lbar + V;
        for (int v = 0; v < 2 * V; ++v) Lbar[v] = 0;
        const double* a = geo->area + (size_t)f * D;
        if (sqrt(dot(D, a, a)) == 0) continue;
        double L[MAXV], R[MAXV], ro[MAXD], rj[MAXD], n[MAXD], magnitude, Fbar[MAXV];
        int fallback[2];
        face_sides(fv, geo, f, dt, L, R, fallback, ro, rj, n, &magnitude);
        const int o = geo->owner[f], j = geo->neighbour[f];
        for (int v = 0; v < V; ++v) Fbar[v] = (dt * qbar[(size_t)j * V + v] - dt * qbar[(size_t)o * V + v]) * magnitude;
        const double w0[MAXD] = {0, 0, 0};
        riemann_adjoint(fv->riemann, D, &fv->gas, L, R, n, w0, Fbar, Lbar, Rbar);
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        const double* Wi = fv->prims + (size_t)i * PW;
        const double* gi = fv->grad + (size_t)i * V * D;
        double Wb[MAXV] = {0, 0, 0, 0, 0}, gb[MAXV * MAXD];
        memset(gb, 0, sizeof gb);
        for (int k = geo->start[i]; k < geo->start[i + 1]; ++k) {
            const int e = geo->list[k], f = list_face(e);
            const double* a = geo->area + (size_t)f * D;
            if (sqrt(dot(D, a, a)) == 0) continue;
            double L[MAXV], R[MAXV], ro[MAXD], rj[MAXD], n[MAXD], magnitude;
            int fallback[2];
            face_sides(fv, geo, f, dt, L, R, fallback, ro, rj, n, &magnitude);
            const int s = e >= 0 ? 0 : 1;
            const double* side = sidebar + (size_t)f * 2 * V + s * V;
            const double* r = s == 0 ? ro : rj;
            if (fallback[s]) {
                for (int v = 0; v < V; ++v) Wb[v] += side[v];
                continue;
            }
            double hb[MAXV * MAXD];
            for (int v = 0; v < V; ++v)
                for (int c = 0; c < D; ++c) hb[v * D + c] = side[v] * r[c];
            orc_half_step_adjoint(D, &fv->gas, Wi, gi, dt, side, Wb, hb);
            for (int v = 0; v < V * D; ++v) gb[v] += hb[v];
        }
        /* the limited fit (finite_volume.hpp:832-864 limited_gradient_adjoint) */
        double g[MAXV * MAXD], g0[MAXV * MAXD], Lc[MAXD + 1][MAXD + 1], M0[MAXD + 1];
        limit_pattern pat;
        fit_cell(fv, geo, i, NULL, g, g0, NULL, Lc, M0, &pat);
        double xbar[MAXV][MAXD + 1], bb[MAXV][MAXD + 1];
        for (int v = 0; v < V; ++v) {
            double g0bar[MAXD];
            for (int c = 0; c < D; ++c) g0bar[c] = gb[v * D + c];
            boundbar[(size_t)i * V + v] = 0;
            extreme[(size_t)i * V + v] = -1;
            if (pat.last[v] >= 0 && pat.zero[v]) {
                for (int c = 0; c < D; ++c) g0bar[c] = 0;
            } else if (pat.last[v] >= 0) {
                const double* o = pat.offset[v];
                const double den = dot(D, g0 + v * D, o);
                const double sc = (pat.bound[v] - Wi[v]) / den, t = dot(D, gb + v * D, g0 + v * D) / den;
                for (int c = 0; c < D; ++c) g0bar[c] = gb[v * D + c] * sc - o[c] * (sc * t);
                boundbar[(size_t)i * V + v] = t;
                extreme[(size_t)i * V + v] = list_face(geo->list[geo->start[i] + pat.extreme[v]]);
                Wb[v] -= t;
            }
            xbar[v][0] = 0;
            for (int c = 0; c < D; ++c) xbar[v][1 + c] = g0bar[c];
        }
        solve_factored(D, Lc, xbar, bb); /* M is symmetric: b bar = M^-1 x bar */
        for (int v = 0; v < V; ++v) {
            for (int p = 0; p < U; ++p) {
                bbar[((size_t)i * V + v) * U + p] = bb[v][p];
                Wb[v] -= M0[p] * bb[v][p];
            }
            Wbar[(size_t)i * V + v] = Wb[v];
        }
    }
    double* out = (double*)malloc(sizeof(double) * (size_t)N * V);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j) {
        double Wb[MAXV];
        for (int v = 0; v < V; ++v) Wb[v] = Wbar[(size_t)j * V + v];
        for (int k = geo->start[j]; k < geo->start[j + 1]; ++k) {
            const int e = geo->list[k], f = list_face(e);
            const int i = e >= 0 ? geo->neighbour[f] : geo->owner[f]; /* the neighbour whose fit read W_j */
            double d[MAXD], X[MAXD + 1];
            separation(geo, i, j, d);
            const double w = magnitude_of(geo, f);
            X[0] = 1;
            for (int p = 1; p < U; ++p) X[p] = d[p - 1] + geo->centroid[(size_t)j * D + p - 1] - geo->centroid[(size_t)i * D + p - 1];
            for (int v = 0; v < V; ++v) {
                double share = 0;
                for (int p = 0; p < U; ++p) share += w * X[p] * bbar[((size_t)i * V + v) * U + p];
                if (extreme[(size_t)i * V + v] == f) share += boundbar[(size_t)i * V + v];
                Wb[v] += share;
            }
        }
        const double inv = 1.0 / geo->volume[j];
        double Uj[MAXV], Ub[MAXV];
        for (int v = 0; v < V; ++v) Uj[v] = fv->q[(size_t)j * V + v] * inv;
        orc_primitive_adjoint(D, &fv->gas, Uj, Wb, Ub);
        for (int v = 0; v < V; ++v) out[(size_t)j * V + v] = qbar[(size_t)j * V + v] + Ub[v] * inv;
    }
    memcpy(qbar, out, sizeof(double) * (size_t)N * V);
    free(out);
    free(sidebar);
    free(Wbar);
    free(bbar);
    free(boundbar);
    free(extreme);
    return 0;
}

/* ---- The moving step's reverse, order 1, fixed topology (finite_volume.hpp:2233-3155) ---- */

/* euler.hpp:848-851 turn_bar(): the n bar of <a, Omega b>, Omega = dn n^T - n dn^T: out += sign (a (n.b) - b (n.a)). */
static void turn_bar_add(int D, const double* n, const double* a, const double* b, double sign, double* out) {
    const double nb = dot(D, n, b), na = dot(D, n, a);
    for (int k = 0; k < D; ++k) out[k] += sign * (a[k] * nb - b[k] * na);
}

/* euler.hpp:893-916 MovingFace::adjoint(): G = |A| F(L, R, A / |A|, w); adds J^T Gbar to Lbar, Rbar, Abar, wbar. The normal's
 * and the face velocity's parts come from the flux's rotation and Galilean invariance (euler.hpp:835-841): only the
 * Riemann solver's (L, R) adjoint is needed. */
static void moving_face_adjoint(int kind, int D, const orc_gas* gas, const double* L, const double* R, const double* A, const double* w,
                                const double* Gbar, double* Lbar, double* Rbar, double* Abar, double* wbar) {
    const int V = D + 2;
    const double s = sqrt(dot(D, A, A));
    double n[MAXD], F[MAXV], Fr[MAXV], X[MAXV], mw[MAXD];
    for (int k = 0; k < D; ++k) {
        n[k] = A[k] * (1.0 / s);
        mw[k] = -w[k];
    }
    riemann(kind, D, gas, L, R, n, w, F);
    memcpy(Fr, F, sizeof(double) * V);
    to_lab(D, Fr, mw); /* the flux in the face's frame */
    double sbar = 0;
    for (int v = 0; v < V; ++v) {
        sbar += Gbar[v] * F[v];
        X[v] = s * Gbar[v];
    }
    double nbar[MAXD] = {0, 0, 0}, shiftbar[MAXD];
    turn_bar_add(D, n, X + 1, F + 1, 1.0, nbar);
    for (int k = 0; k < D; ++k) shiftbar[k] = X[1 + k] * Fr[0] + X[D + 1] * (Fr[1 + k] + w[k] * Fr[0]);
    double Lr[MAXV] = {0, 0, 0, 0, 0}, Rr[MAXV] = {0, 0, 0, 0, 0};
    riemann_adjoint(kind, D, gas, L, R, n, w, X, Lr, Rr);
    for (int v = 0; v < V; ++v) {
        Lbar[v] += Lr[v];
        Rbar[v] += Rr[v];
    }
    for (int k = 0; k < D; ++k) shiftbar[k] -= Lr[1 + k] + Rr[1 + k];
    turn_bar_add(D, n, Lr + 1, L + 1, -1.0, nbar);
    turn_bar_add(D, n, Rr + 1, R + 1, -1.0, nbar);
    for (int k = 0; k < D; ++k) wbar[k] += shiftbar[k];
    turn_bar_add(D, n, shiftbar, w, -1.0, nbar);
    const double nn = dot(D, n, nbar);
    for (int k = 0; k < D; ++k) Abar[k] += (nbar[k] - n[k] * nn) * (1.0 / s) + n[k] * sbar;
}

/* finite_volume.hpp:2331-2369 FaceVelocity::adjoint(): of face_velocity(d, x_f, w_L, w_R) (d owner to neighbour, x_f relative to
 * the owner), at the forward's branch. */
static void face_velocity_adjoint(int D, const double* d, const double* xf, const double* wL, const double* wR, const double* wfbar,
                                  double* dbar, double* xfbar, double* wLbar, double* wRbar) {
    double c[MAXD], u[MAXD], cbar[MAXD] = {0, 0, 0}, ubar[MAXD];
    const double dist = sqrt(dot(D, d, d));
    for (int a = 0; a < D; ++a) {
        c[a] = xf[a] - d[a] * 0.5;
        u[a] = wL[a] - wR[a];
    }
    const double f0 = dot(D, c, u) / dist, limit = 0.9 * dist, c2 = dot(D, c, c), cc = sqrt(c2);
    const int clipped = c2 > limit * limit;
    const double sc = clipped ? limit / cc : 1.0, facv = f0 * sc;
    for (int a = 0; a < D; ++a) {
        wLbar[a] += wfbar[a] * 0.5;
        wRbar[a] += wfbar[a] * 0.5;
        dbar[a] += wfbar[a] * (facv / dist);
    }
    const double qb = dot(D, wfbar, d), facvbar = qb / dist;
    double distbar = -qb * facv / (dist * dist);
    const double f0bar = facvbar * sc;
    if (clipped) {
        const double scbar = facvbar * f0;
        distbar += scbar * 0.9 / cc;
        for (int a = 0; a < D; ++a) cbar[a] += c[a] * (-scbar * sc / (cc * cc));
    }
    for (int a = 0; a < D; ++a) {
        cbar[a] += u[a] * (f0bar / dist);
        ubar[a] = c[a] * (f0bar / dist);
    }
    distbar -= f0bar * f0 / dist;
    for (int a = 0; a < D; ++a) {
        wLbar[a] += ubar[a];
        wRbar[a] -= ubar[a];
        xfbar[a] += cbar[a];
        dbar[a] += -cbar[a] * 0.5 + d[a] * (distbar / dist);
    }
}

/* finite_volume.hpp:2237-2243 face_angle_slope(): d face_angle / d |A|. */
static double face_angle_slope(int D, double area) {
    if (D == 3) return area > 0 ? 0.5 / (M_PI * sqrt(area / M_PI)) : 0.0;
    return 0.5;
}

/* Bars of the geometry a moving step reads: per cell c (centroid), V, x (through the separations it reads); per face A,
 * x_f. */
typedef struct {
    double *V, *c, *x, *A, *xf;
} geo_bars;

static void add_sep_bar(int D, geo_bars* b, int from, int to, const double* bar) { /* bar of separation(from, to) = x_to - x_from */
    for (int a = 0; a < D; ++a) {
        b->x[(size_t)to * D + a] += bar[a];
        b->x[(size_t)from * D + a] -= bar[a];
    }
}

/* The bar of face_offset(f, i) (x_f relative to c_i; through the owner's separation when i does not own f). */
static void face_offset_bar(const orc_geometry* g, geo_bars* b, int f, int i, const double* rbar) {
    const int D = g->D, o = g->owner[f];
    for (int a = 0; a < D; ++a) {
        b->xf[(size_t)f * D + a] += rbar[a];
        b->c[(size_t)i * D + a] -= rbar[a];
    }
    if (o != i) add_sep_bar(D, b, i, o, rbar);
}

/* One moving pass's adjoint (finite_volume.hpp:2876-2983 moving_pass_adjoint / moving_pass_gather): Q -= dt/2 sum_f +-G_f
 * with G = |A| F(side_o, side_j, A/|A|, face_velocity(d, x_f, w-hat_o, w-hat_j)). First pass (mesh t^n): side W + g.r;
 * second (mesh t^n+1): half_step(W, g, 2 dt) + g.(r + c^n+1 - c^n + dt w-hat). Adds to Wbar, gbar, whbar, c0bar (per cell) and
 * to the mesh's bars. */
static void moving_pass_adjoint(const orc_fv* fv, const orc_geometry* g, int second, double dt, const double* wh, const double* qbar,
                                double* Wbar, double* gbar, double* whbar, double* c0bar, geo_bars* b) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    for (int f = 0; f < g->F; ++f) {
        const int o = g->owner[f], j = g->neighbour[f];
        const double* A = g->area + (size_t)f * D;
        if (sqrt(dot(D, A, A)) == 0) continue;
        double d[MAXD], wf[MAXD], sides[2][MAXV], r[2][MAXD];
        int fallback[2];
        separation(g, o, j, d);
        face_velocity(g, f, o, j, wh + (size_t)o * D, wh + (size_t)j * D, wf);
        const int cells[2] = {o, j};
        for (int s2 = 0; s2 < 2; ++s2) {
            const int i = cells[s2];
            const double* W = fv->prims + (size_t)i * PW;
            const double* gi = fv->grad + (size_t)i * V * D;
            face_offset(g, f, i, r[s2]);
            if (second) {
                for (int a = 0; a < D; ++a)
                    r[s2][a] = r[s2][a] + g->centroid[(size_t)i * D + a] - fv->old_centroid[(size_t)i * D + a] + wh[(size_t)i * D + a] * (0.5 * (2 * dt));
                orc_half_step(D, &fv->gas, W, gi, 2 * dt, sides[s2]);
            } else {
                for (int v = 0; v < V; ++v) sides[s2][v] = W[v];
            }
            for (int v = 0; v < V; ++v) sides[s2][v] += dot(D, gi + v * D, r[s2]);
            fallback[s2] = !(sides[s2][0] > 0 && sides[s2][D + 1] > 0);
            if (fallback[s2])
                for (int v = 0; v < V; ++v) sides[s2][v] = W[v];
        }
        double Gbar[MAXV], Lbar[MAXV] = {0, 0, 0, 0, 0}, Rbar[MAXV] = {0, 0, 0, 0, 0}, wfbar[MAXD] = {0, 0, 0};
        for (int v = 0; v < V; ++v) Gbar[v] = -0.5 * dt * (qbar[(size_t)o * V + v] - qbar[(size_t)j * V + v]);
        moving_face_adjoint(fv->riemann, D, &fv->gas, sides[0], sides[1], A, wf, Gbar, Lbar, Rbar, b->A + (size_t)f * D, wfbar);
        double dbar[MAXD] = {0, 0, 0};
        face_velocity_adjoint(D, d, g->face_centroid + (size_t)f * D, wh + (size_t)o * D, wh + (size_t)j * D, wfbar, dbar,
                              b->xf + (size_t)f * D, whbar + (size_t)o * D, whbar + (size_t)j * D);
        add_sep_bar(D, b, o, j, dbar);
        const double* sidebar[2] = {Lbar, Rbar};
        for (int s2 = 0; s2 < 2; ++s2) {
            const int i = cells[s2];
            const double* W = fv->prims + (size_t)i * PW;
            const double* gi = fv->grad + (size_t)i * V * D;
            const double* sb = sidebar[s2];
            double* Wb = Wbar + (size_t)i * V;
            double* gb = gbar + (size_t)i * V * D;
            if (fallback[s2]) {
                for (int v = 0; v < V; ++v) Wb[v] += sb[v];
                continue;
            }
            if (second) {
                orc_half_step_adjoint(D, &fv->gas, W, gi, 2 * dt, sb, Wb, gb);
            } else {
                for (int v = 0; v < V; ++v) Wb[v] += sb[v];
            }
            double rbar[MAXD] = {0, 0, 0};
            for (int v = 0; v < V; ++v)
                for (int a = 0; a < D; ++a) {
                    gb[v * D + a] += sb[v] * r[s2][a];
                    rbar[a] += sb[v] * gi[v * D + a];
                }
            face_offset_bar(g, b, f, i, rbar);
            if (second)
                for (int a = 0; a < D; ++a) {
                    b->c[(size_t)i * D + a] += rbar[a];
                    c0bar[(size_t)i * D + a] -= rbar[a];
                    whbar[(size_t)i * D + a] += dt * rbar[a];
                }
        }
    }
}

/* finite_volume.hpp:2245-2306 Regularised::adjoint() for sweep_cell()'s point velocity w = v + c (fraction c_s / |c|),
 * fraction from the largest face angle: adds to Wbar (velocity), *csbar, and the bars of c_i, of the angle face's A and of its
 * separation. */
static void regularised_adjoint(const orc_fv* fv, const orc_geometry* g, int i, const double* wbar, double* Wbar, double* csbar, geo_bars* b) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    const double* Wi = fv->prims + (size_t)i * PW;
    double amax = 0;
    int fmax = -1, other = -1;
    for (int k = g->start[i]; k < g->start[i + 1]; ++k) {
        const int e = g->list[k], f = list_face(e), j = e >= 0 ? g->neighbour[f] : g->owner[f];
        double d[MAXD];
        separation(g, i, j, d);
        const double angle = face_angle(D, magnitude_of(g, f)) / (0.5 * sqrt(dot(D, d, d)));
        if (angle > amax) {
            amax = angle;
            fmax = f;
            other = j;
        }
    }
    const double F = 2.25, shaping = 0.5;
    double fraction = 0, slope = 0;
    if (amax > 0.75 * F) {
        fraction = amax > F ? shaping : shaping * (amax - 0.75 * F) / (0.25 * F);
        slope = amax > F ? 0.0 : shaping / (0.25 * F);
    }
    for (int a = 0; a < D; ++a) Wbar[1 + a] += wbar[a];
    const double* c = g->centroid + (size_t)i * D;
    const double cd = sqrt(dot(D, c, c)), cs = Wi[V];
    if (!(cd > 0 && fraction > 0)) return;
    const double k = fraction * cs / cd, kbar = dot(D, wbar, c);
    const double fractionbar = kbar * cs / cd, cdbar = -kbar * k / cd;
    *csbar += kbar * fraction / cd;
    for (int a = 0; a < D; ++a) b->c[(size_t)i * D + a] += wbar[a] * k + c[a] * (cdbar / cd);
    if (slope == 0) return;
    const double abar = slope * fractionbar;
    const double* A = g->area + (size_t)fmax * D;
    double d[MAXD], Ab[MAXD], db[MAXD];
    separation(g, i, other, d);
    const double mag = sqrt(dot(D, A, A)), dist = sqrt(dot(D, d, d));
    for (int a = 0; a < D; ++a) {
        Ab[a] = A[a] * (abar * face_angle_slope(D, mag) / mag / (0.5 * dist));
        db[a] = -d[a] * (abar * amax / (dist * dist));
        b->A[(size_t)fmax * D + a] += Ab[a];
    }
    add_sep_bar(D, b, i, other, db);
}

/* finite_volume.hpp:2629-2760 moving_gradient_adjoint(): the limited fit's adjoint in W and in the geometry it reads. The fit
 * x_v = M^-1 b_v (X_k = [1, d_k + c_j - c_i], weights |A_k|) moves by M^-1 (db - dM x); the last clip's scale
 * s = (bound - W_i) / (g0 . o) also by o's bar. Writes beta (V x (D + 1): M^-1 x_v bar), boundbar, the extreme face and
 * adds W_i's own bar; the neighbours' W bars are gathered from beta by each neighbour. */
static void moving_fit_adjoint(const orc_fv* fv, const orc_geometry* g, int i, const double* gb, double* Wb_i, double* beta, double* boundbar,
                               int* extreme, geo_bars* b) {
    const int D = fv->D, V = D + 2, U = D + 1, PW = V + 2;
    const double* Wi = fv->prims + (size_t)i * PW;
    double gl[MAXV * MAXD], g0[MAXV * MAXD], Lc[MAXD + 1][MAXD + 1], M0[MAXD + 1], icpt[MAXV];
    limit_pattern pat;
    fit_cell_x(fv, g, i, NULL, gl, g0, NULL, Lc, M0, &pat, icpt);
    const int first = g->start[i];
    double xbar[MAXV][MAXD + 1], bb[MAXV][MAXD + 1];
    for (int v = 0; v < V; ++v) {
        double g0bar[MAXD];
        for (int a = 0; a < D; ++a) g0bar[a] = gb[v * D + a];
        boundbar[v] = 0;
        extreme[v] = -1;
        if (pat.last[v] >= 0 && pat.zero[v]) {
            for (int a = 0; a < D; ++a) g0bar[a] = 0;
        } else if (pat.last[v] >= 0) {
            const double* o = pat.offset[v];
            const double den = dot(D, g0 + v * D, o);
            const double sc = (pat.bound[v] - Wi[v]) / den, t = dot(D, gb + v * D, g0 + v * D) / den;
            double obar[MAXD];
            for (int a = 0; a < D; ++a) {
                g0bar[a] = gb[v * D + a] * sc - o[a] * (sc * t);
                obar[a] = -sc * t * g0[v * D + a];
            }
            boundbar[v] = t;
            extreme[v] = list_face(g->list[first + pat.extreme[v]]);
            Wb_i[v] -= t;
            face_offset_bar(g, b, list_face(g->list[first + pat.last[v]]), i, obar);
        }
        xbar[v][0] = 0;
        for (int a = 0; a < D; ++a) xbar[v][1 + a] = g0bar[a];
    }
    solve_factored(D, Lc, xbar, bb); /* beta_v = M^-1 x_v bar (M symmetric) */
    double x[MAXV][MAXD + 1], Mbar[MAXD + 1][MAXD + 1];
    for (int v = 0; v < V; ++v) {
        x[v][0] = icpt[v];
        for (int a = 0; a < D; ++a) x[v][1 + a] = g0[v * D + a];
    }
    for (int p = 0; p < U; ++p)
        for (int q = 0; q < U; ++q) {
            Mbar[p][q] = 0;
            for (int v = 0; v < V; ++v) Mbar[p][q] -= bb[v][p] * x[v][q];
        }
    for (int v = 0; v < V; ++v)
        for (int p = 0; p < U; ++p) {
            beta[v * U + p] = bb[v][p];
            Wb_i[v] -= M0[p] * bb[v][p];
        }
    const double* ci = g->centroid + (size_t)i * D;
    for (int k = first; k < g->start[i + 1]; ++k) {
        const int e = g->list[k], f = list_face(e), j = e >= 0 ? g->neighbour[f] : g->owner[f];
        const double* Wj = fv->prims + (size_t)j * PW;
        double d[MAXD], X[MAXD + 1], Xbar[MAXD + 1] = {0, 0, 0, 0};
        separation(g, i, j, d);
        const double w = magnitude_of(g, f);
        X[0] = 1;
        for (int p = 1; p < U; ++p) X[p] = d[p - 1] + g->centroid[(size_t)j * D + p - 1] - ci[p - 1];
        double wbar = 0;
        for (int v = 0; v < V; ++v) {
            const double dW = Wj[v] - Wi[v];
            double xb = 0;
            for (int p = 0; p < U; ++p) xb += X[p] * bb[v][p];
            wbar += xb * dW;
            for (int p = 0; p < U; ++p) Xbar[p] += w * dW * bb[v][p];
        }
        for (int p = 0; p < U; ++p)
            for (int q = 0; q < U; ++q) {
                wbar += X[p] * Mbar[p][q] * X[q];
                Xbar[p] += w * (Mbar[p][q] + Mbar[q][p]) * X[q];
            }
        if (w > 0) {
            const double* A = g->area + (size_t)f * D;
            for (int a = 0; a < D; ++a) b->A[(size_t)f * D + a] += wbar * A[a] / w;
        }
        /* between = d + c_j - c_i */
        add_sep_bar(D, b, i, j, Xbar + 1);
        for (int a = 0; a < D; ++a) {
            b->c[(size_t)j * D + a] += Xbar[1 + a];
            b->c[(size_t)i * D + a] -= Xbar[1 + a];
        }
    }
}

/* finite_volume.hpp:2490-2598 step_moving_adjoint (order 0): Q^n+1 = Q^n - dt sum F(W_o, W_j, A, w_f) on the mesh of t^n,
 * x^n+1 = x^n + dt w (wrapped: the identity's derivative), w the regularised point velocity. The pass is the first pass's
 * adjoint with zero gradients (sides W), w in place of w-hat and dt for dt/2; then w's bar takes dt of x^n+1's, the point
 * velocity's and the sound speed's adjoints follow, and W = primitive(Q / V) gives Q's and V's bars. */
static int moving_adjoint_order0(orc_fv* fv, const orc_geometry* g0, double dt, double* qbar, double* xbar, orc_geometry_adjoint adjoint0,
                                 void* ctx0) {
    const int D = fv->D, V = D + 2, N = fv->N, PW = V + 2;
    orc_primitives(fv, g0);
    for (int i = 0; i < N; ++i) sweep_cell(fv, g0, i);
    memset(fv->grad, 0, sizeof(double) * (size_t)N * V * D);
    double* Wbar = (double*)calloc((size_t)N * V, sizeof(double));
    double* gbar = (double*)calloc((size_t)N * V * D, sizeof(double));
    double* wbar = (double*)calloc((size_t)N * D, sizeof(double));
    double* c0bar = (double*)calloc((size_t)N * D, sizeof(double));
    const int F0 = g0->F;
    geo_bars b0 = {(double*)calloc(N, sizeof(double)), (double*)calloc((size_t)N * D, sizeof(double)), (double*)calloc((size_t)N * D, sizeof(double)),
                   (double*)calloc((size_t)(F0 > 0 ? F0 : 1) * D, sizeof(double)), (double*)calloc((size_t)(F0 > 0 ? F0 : 1) * D, sizeof(double))};
    moving_pass_adjoint(fv, g0, 0, 2 * dt, fv->w, qbar, Wbar, gbar, wbar, c0bar, &b0);
    for (size_t k = 0; k < (size_t)N * D; ++k) {
        b0.x[k] += xbar[k];
        wbar[k] += dt * xbar[k];
    }
    for (int i = 0; i < N; ++i) {
        const double* Wi = fv->prims + (size_t)i * PW;
        double* Wb = Wbar + (size_t)i * V;
        double csbar = 0;
        regularised_adjoint(fv, g0, i, wbar + (size_t)i * D, Wb, &csbar, &b0);
        if (Wi[V] > 0) {
            const double c2bar = csbar / (2 * Wi[V]);
            Wb[D + 1] += c2bar * fv->gas.gamma / Wi[0];
            Wb[0] -= c2bar * fv->gas.gamma * Wi[D + 1] / (Wi[0] * Wi[0]);
        }
    }
    for (int j = 0; j < N; ++j) {
        const double inv = 1.0 / g0->volume[j];
        double Uj[MAXV], Ub[MAXV];
        for (int v = 0; v < V; ++v) Uj[v] = fv->q[(size_t)j * V + v] * inv;
        orc_primitive_adjoint(D, &fv->gas, Uj, Wbar + (size_t)j * V, Ub);
        double vb = 0;
        for (int v = 0; v < V; ++v) {
            qbar[(size_t)j * V + v] += Ub[v] * inv;
            vb -= Ub[v] * Uj[v] * inv;
        }
        b0.V[j] += vb;
    }
    adjoint0(ctx0, b0.V, b0.c, b0.A, b0.xf, b0.x);
    memcpy(xbar, b0.x, sizeof(double) * (size_t)N * D);
    free(Wbar);
    free(gbar);
    free(wbar);
    free(c0bar);
    free(b0.V);
    free(b0.c);
    free(b0.x);
    free(b0.A);
    free(b0.xf);
    return 0;
}

int orc_moving_adjoint(orc_fv* fv, const orc_geometry* g0, const orc_geometry* g1, double dt, int steps_fixed, double* qbar, double* xbar,
                       orc_geometry_adjoint adjoint0, void* ctx0, orc_geometry_adjoint adjoint1, void* ctx1) {
    if (!fv->moving || fv->D < 2 || fv->precision != 0 || fv->hll_compression > 0) return -1;
    if (fv->order == 0) return moving_adjoint_order0(fv, g0, dt, qbar, xbar, adjoint0, ctx0);
    if (fv->order != 1) return -1;
    const int D = fv->D, V = D + 2, U = D + 1, N = fv->N, PW = V + 2;
    /* t^n: W, the limited gradients and the point velocities (the sweep), c^n, the grid steps and w-hat. */
    orc_primitives(fv, g0);
    for (int i = 0; i < N; ++i) sweep_cell(fv, g0, i);
    const double unit = g0->step / dt;
    double* wh = (double*)calloc((size_t)N * D, sizeof(double));
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < D; ++a) {
            const size_t at = (size_t)i * D + a;
            fv->old_centroid[at] = g0->centroid[at];
            if (!fv->hold_steps) {
                const long long k0 = orc_snap(fv->x[at], g0->step, g0->period[a]);
                const long long k1 = orc_snap(fv->x[at] + dt * fv->w[at], g0->step, g0->period[a]);
                fv->steps[at] = orc_delta(k0, k1, g0->period[a]);
            }
            wh[at] = (double)fv->steps[at] * unit;
        }
    double* Wbar = (double*)calloc((size_t)N * V, sizeof(double));
    double* gbar = (double*)calloc((size_t)N * V * D, sizeof(double));
    double* whbar = (double*)calloc((size_t)N * D, sizeof(double));
    double* c0bar = (double*)calloc((size_t)N * D, sizeof(double));
    const int F1 = g1->F, F0 = g0->F;
    geo_bars b1 = {(double*)calloc(N, sizeof(double)), (double*)calloc((size_t)N * D, sizeof(double)), (double*)calloc((size_t)N * D, sizeof(double)),
                   (double*)calloc((size_t)(F1 > 0 ? F1 : 1) * D, sizeof(double)), (double*)calloc((size_t)(F1 > 0 ? F1 : 1) * D, sizeof(double))};
    geo_bars b0 = {(double*)calloc(N, sizeof(double)), (double*)calloc((size_t)N * D, sizeof(double)), (double*)calloc((size_t)N * D, sizeof(double)),
                   (double*)calloc((size_t)(F0 > 0 ? F0 : 1) * D, sizeof(double)), (double*)calloc((size_t)(F0 > 0 ? F0 : 1) * D, sizeof(double))};
    /* Pass 2 on t^n+1, then x^n+1's bar: the caller's, the separations', the geometry's. */
    moving_pass_adjoint(fv, g1, 1, dt, wh, qbar, Wbar, gbar, whbar, c0bar, &b1);
    for (size_t k = 0; k < (size_t)N * D; ++k) b1.x[k] += xbar[k];
    adjoint1(ctx1, b1.V, b1.c, b1.A, b1.xf, b1.x);
    /* x^n+1 = x^n + dt w-hat (the quantisation the identity; held steps: w-hat fixed). */
    for (size_t k = 0; k < (size_t)N * D; ++k) {
        b0.x[k] = b1.x[k];
        if (!steps_fixed) whbar[k] += dt * b1.x[k];
    }
    /* Pass 1 on t^n. */
    moving_pass_adjoint(fv, g0, 0, dt, wh, qbar, Wbar, gbar, whbar, c0bar, &b0);
    for (size_t k = 0; k < (size_t)N * D; ++k) b0.c[k] += c0bar[k];
    /* Per cell: the point velocity (w-hat = w), the sound speed, the limited fit. */
    double* beta = (double*)calloc((size_t)N * V * U, sizeof(double));
    double* boundbar = (double*)calloc((size_t)N * V, sizeof(double));
    int* extreme = (int*)malloc(sizeof(int) * (size_t)N * V);
    for (int i = 0; i < N; ++i) {
        const double* Wi = fv->prims + (size_t)i * PW;
        double* Wb = Wbar + (size_t)i * V;
        double csbar = 0;
        if (!steps_fixed) regularised_adjoint(fv, g0, i, whbar + (size_t)i * D, Wb, &csbar, &b0);
        if (Wi[V] > 0) { /* c_s = sqrt(gamma P / rho) */
            const double c2bar = csbar / (2 * Wi[V]);
            Wb[D + 1] += c2bar * fv->gas.gamma / Wi[0];
            Wb[0] -= c2bar * fv->gas.gamma * Wi[D + 1] / (Wi[0] * Wi[0]);
        }
        moving_fit_adjoint(fv, g0, i, gbar + (size_t)i * V * D, Wb, beta + (size_t)i * V * U, boundbar + (size_t)i * V, extreme + (size_t)i * V, &b0);
    }
    /* Per cell: the neighbours' fits' bars of its W, then W = primitive(Q / V): Q's and V's bars. */
    for (int j = 0; j < N; ++j) {
        double Wb[MAXV];
        for (int v = 0; v < V; ++v) Wb[v] = Wbar[(size_t)j * V + v];
        for (int k = g0->start[j]; k < g0->start[j + 1]; ++k) {
            const int e = g0->list[k], f = list_face(e);
            const int i = e >= 0 ? g0->neighbour[f] : g0->owner[f];
            double d[MAXD], X[MAXD + 1];
            separation(g0, i, j, d);
            const double w = magnitude_of(g0, f);
            X[0] = 1;
            for (int p = 1; p < U; ++p) X[p] = d[p - 1] + g0->centroid[(size_t)j * D + p - 1] - g0->centroid[(size_t)i * D + p - 1];
            for (int v = 0; v < V; ++v) {
                double share = 0;
                for (int p = 0; p < U; ++p) share += w * X[p] * beta[((size_t)i * V + v) * U + p];
                if (extreme[(size_t)i * V + v] == f) share += boundbar[(size_t)i * V + v];
                Wb[v] += share;
            }
        }
        const double inv = 1.0 / g0->volume[j];
        double Uj[MAXV], Ub[MAXV];
        for (int v = 0; v < V; ++v) Uj[v] = fv->q[(size_t)j * V + v] * inv;
        orc_primitive_adjoint(D, &fv->gas, Uj, Wb, Ub);
        double vb = 0;
        for (int v = 0; v < V; ++v) {
            qbar[(size_t)j * V + v] += Ub[v] * inv;
            vb -= Ub[v] * Uj[v] * inv;
        }
        b0.V[j] += vb;
    }
    adjoint0(ctx0, b0.V, b0.c, b0.A, b0.xf, b0.x);
    memcpy(xbar, b0.x, sizeof(double) * (size_t)N * D);
    free(wh);
    free(Wbar);
    free(gbar);
    free(whbar);
    free(c0bar);
    free(beta);
    free(boundbar);
    free(extreme);
    geo_bars* all[2] = {&b0, &b1};
    for (int k = 0; k < 2; ++k) {
        free(all[k]->V);
        free(all[k]->c);
        free(all[k]->x);
        free(all[k]->A);
        free(all[k]->xf);
    }
    return 0;
}

/* ---- HLL (euler.hpp:444-540): tangent and adjoint, of the case the forward took ---- */

/* euler.hpp:377-385: the tangent of side(), the sound speed's included (zero where c^2 <= 0, as the forward). */
static void side_full_tangent(int D, const orc_gas* gas, const double* W, const double* dW, const double* n, const side_t* s, side_t* d) {
    const int P = D + 1;
    double dv[MAXD];
    for (int k = 0; k < D; ++k) dv[k] = dW[1 + k];
    side_with_tangent(D, gas, W, dW, n, s, dot(D, dv, n), d);
    const double c2 = gas->gamma * W[P] / W[0];
    d->c = c2 > 0 ? (gas->gamma * dW[P] - c2 * dW[0]) / W[0] / (2 * s->c) : 0.0;
}

/* euler.hpp:413-426: the adjoint of side() from the bars of F, U, vn and c; adds J^T to Wbar. */
static void side_full_adjoint(int D, const orc_gas* gas, const double* W, const double* n, const side_t* s, const double* Fbar,
                              const double* Ubar, double vnbar, double cbar, double* Wbar) {
    const int P = D + 1;
    double vbar[MAXD] = {0, 0, 0};
    vnbar += side_with_adjoint(D, gas, W, n, s, Fbar, Ubar, 0.0, vbar, Wbar);
    for (int k = 0; k < D; ++k) Wbar[1 + k] += vnbar * n[k];
    const double c2 = gas->gamma * W[P] / W[0];
    if (c2 > 0) {
        const double c2bar = cbar / (2 * s->c);
        Wbar[P] += c2bar * gas->gamma / W[0];
        Wbar[0] -= c2bar * c2 / W[0];
    }
}

/* HLL's wave speeds and case as orc_hll computes them: bit 0 S_L from L, bit 1 S_R from L; case 1 F_L, 2 F_R, 0 the HLL state. */
static int hll_case(const side_t* l, const side_t* r, double* SL, double* SR, int* pattern) {
    const double sl = l->vn - l->c, sr = r->vn - r->c, tl = l->vn + l->c, tr = r->vn + r->c;
    *SL = fmin(sl, sr);
    *SR = fmax(tl, tr);
    *pattern = (sr < sl ? 0 : 1) | (tr > tl ? 0 : 2);
    return *SL >= 0.0 ? 1 : *SR <= 0.0 ? 2 : 0;
}

/* euler.hpp:466-505 HLL::flux_tangent(): F from the same intermediates and dF = J (dL, dR), n and w fixed. */
void orc_hll_tangent(int D, const orc_gas* gas, const double* L, const double* R, const double* dL, const double* dR, const double* n,
                     const double* w, double* F, double* dF) {
    side_t l, r, dl, dr;
    side_full(D, gas, L, n, w, &l);
    side_full(D, gas, R, n, w, &r);
    side_full_tangent(D, gas, L, dL, n, &l, &dl);
    side_full_tangent(D, gas, R, dR, n, &r, &dr);
    double SL, SR;
    int pattern;
    const int which = hll_case(&l, &r, &SL, &SR, &pattern);
    const double dSL = (pattern & 1) ? dl.vn - dl.c : dr.vn - dr.c, dSR = (pattern & 2) ? dl.vn + dl.c : dr.vn + dr.c;
    if (which == 1) {
        for (int v = 0; v < D + 2; ++v) F[v] = l.F[v], dF[v] = dl.F[v];
    } else if (which == 2) {
        for (int v = 0; v < D + 2; ++v) F[v] = r.F[v], dF[v] = dr.F[v];
    } else {
        const double inv_dS = 1.0 / (SR - SL), dinv = -inv_dS * inv_dS * (dSR - dSL);
        for (int v = 0; v < D + 2; ++v) {
            const double N = SR * l.F[v] - SL * r.F[v] + SL * SR * (r.U[v] - l.U[v]);
            const double dN = dSR * l.F[v] + SR * dl.F[v] - dSL * r.F[v] - SL * dr.F[v] + (dSL * SR + SL * dSR) * (r.U[v] - l.U[v]) +
                              SL * SR * (dr.U[v] - dl.U[v]);
            F[v] = N * inv_dS;
            dF[v] = dN * inv_dS + N * dinv;
        }
    }
    to_lab(D, F, w);
    to_lab(D, dF, w);
}

/* euler.hpp:507-540 HLL::flux_adjoint(): adds J^T Fbar to Lbar and Rbar. */
void orc_hll_adjoint(int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w, const double* Fbar_lab,
                     double* Lbar, double* Rbar) {
    side_t l, r;
    side_full(D, gas, L, n, w, &l);
    side_full(D, gas, R, n, w, &r);
    double SL, SR, Fbar[MAXV];
    int pattern;
    const int which = hll_case(&l, &r, &SL, &SR, &pattern);
    to_lab_adjoint(D, Fbar_lab, w, Fbar);
    const double zero[MAXV] = {0, 0, 0, 0, 0};
    if (which == 1) {
        side_full_adjoint(D, gas, L, n, &l, Fbar, zero, 0.0, 0.0, Lbar);
        return;
    }
    if (which == 2) {
        side_full_adjoint(D, gas, R, n, &r, Fbar, zero, 0.0, 0.0, Rbar);
        return;
    }
    const double inv_dS = 1.0 / (SR - SL);
    double lF[MAXV], rF[MAXV], lU[MAXV], rU[MAXV], SLbar = 0, SRbar = 0, inv_bar = 0;
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
    const double gapbar = -inv_dS * inv_dS * inv_bar; /* inv_dS = 1 / (SR - SL) */
    SRbar += gapbar;
    SLbar -= gapbar;
    double lvn = 0, lc = 0, rvn = 0, rc = 0;
    if (pattern & 1) lvn += SLbar, lc -= SLbar;
    else rvn += SLbar, rc -= SLbar;
    if (pattern & 2) lvn += SRbar, lc += SRbar;
    else rvn += SRbar, rc += SRbar;
    side_full_adjoint(D, gas, L, n, &l, lF, lU, lvn, lc, Lbar);
    side_full_adjoint(D, gas, R, n, &r, rF, rU, rvn, rc, Rbar);
}

/* The flux derivative of the solver in use (0 HLLC, 1 HLL). */
static void riemann_tangent(int kind, int D, const orc_gas* gas, const double* L, const double* R, const double* dL, const double* dR,
                            const double* n, const double* w, double* F, double* dF) {
    if (kind == 1)
        orc_hll_tangent(D, gas, L, R, dL, dR, n, w, F, dF);
    else
        orc_hllc_tangent(D, gas, L, R, dL, dR, n, w, F, dF);
}
static void riemann_adjoint(int kind, int D, const orc_gas* gas, const double* L, const double* R, const double* n, const double* w,
                            const double* Fbar, double* Lbar, double* Rbar) {
    if (kind == 1)
        orc_hll_adjoint(D, gas, L, R, n, w, Fbar, Lbar, Rbar);
    else
        orc_hllc_adjoint(D, gas, L, R, n, w, Fbar, Lbar, Rbar);
}

/* ---- The moving step's tangent (forward mode), fixed topology (finite_volume.hpp:2428-2486, 2764-2870) ----
 * The forward-mode counterpart of orc_moving_adjoint: from the state held (Q, x at t^n on g0) and the mesh g1 the forward step
 * reaches, (dq, dx) at t^n -> at t^n+1, in place; Q, x stay. The geometry's tangents come from the mesh library's derivative
 * (orc_geometry_tangent: dx -> dV, dc per cell, dA, dx_f per face). */

/* The tangents of the positions and of one mesh's geometry. */
typedef struct {
    const double *x, *V, *c, *A, *xf;
} geo_dots;

/* The tangent of face_offset(f, i): dx_f - dc_i, plus the owner's separation's when i does not own f. */
static void offset_dot(const orc_geometry* g, const geo_dots* t, int f, int i, double* dr) {
    const int D = g->D, o = g->owner[f];
    for (int a = 0; a < D; ++a) dr[a] = t->xf[(size_t)f * D + a] - t->c[(size_t)i * D + a];
    if (o != i)
        for (int a = 0; a < D; ++a) dr[a] += t->x[(size_t)o * D + a] - t->x[(size_t)i * D + a];
}

/* finite_volume.hpp:2322-2330 FaceVelocity::tangent(): of face_velocity(d, x_f, w_L, w_R) at the forward's branch of the
 * 0.9 |d| limit. */
void orc_face_velocity_tangent(int D, const double* d, const double* xf, const double* wL, const double* wR, const double* dd, const double* dxf,
                               const double* dwL, const double* dwR, double* dwf) {
    double c[MAXD], u[MAXD], dc[MAXD], du[MAXD];
    for (int a = 0; a < D; ++a) {
        c[a] = xf[a] - d[a] * 0.5;
        u[a] = wL[a] - wR[a];
        dc[a] = dxf[a] - dd[a] * 0.5;
        du[a] = dwL[a] - dwR[a];
    }
    const double dist = sqrt(dot(D, d, d)), ddist = dot(D, d, dd) / dist;
    const double f0 = dot(D, c, u) / dist, df0 = (dot(D, dc, u) + dot(D, c, du)) / dist - f0 * ddist / dist;
    const double limit = 0.9 * dist, c2 = dot(D, c, c);
    double facv = f0, dfacv = df0;
    if (c2 > limit * limit) {
        const double cc = sqrt(c2), sc = limit / cc, dsc = 0.9 * ddist / cc - sc * dot(D, c, dc) / c2;
        facv = f0 * sc;
        dfacv = df0 * sc + f0 * dsc;
    }
    for (int a = 0; a < D; ++a) dwf[a] = (dwL[a] + dwR[a]) * 0.5 + dd[a] * (facv / dist) + d[a] * (dfacv / dist - facv * ddist / (dist * dist));
}

/* euler.hpp:861-891 MovingFace::tangent(): dG of G = |A| F(L, R, A / |A|, w). The flux is the face frame's Phi(L - w, R - w, n)
 * taken to the lab (to_lab, linear in Phi); Phi's change with n follows from its rotation covariance: for the rotation
 * Omega = dn n^T - n dn^T (Omega n = dn), Phi_n dn = Omega Phi - Phi_L (Omega L_w) - Phi_R (Omega R_w). So one tangent of the
 * Riemann solver at (L, R) with dL - dw - Omega L_w, dR - dw - Omega R_w, plus Omega Phi and to_lab's own change with w. */
void orc_moving_face_tangent(int kind, int D, const orc_gas* gas, const double* L, const double* R, const double* A, const double* w,
                             const double* dL, const double* dR, const double* dA, const double* dw, double* dG) {
    const int V = D + 2, P = D + 1;
    const double s = sqrt(dot(D, A, A));
    double n[MAXD], dn[MAXD], mw[MAXD], lw[MAXD], rw[MAXD];
    for (int k = 0; k < D; ++k) {
        n[k] = A[k] * (1.0 / s);
        mw[k] = -w[k];
        lw[k] = L[1 + k] - w[k];
        rw[k] = R[1 + k] - w[k];
    }
    const double ds = dot(D, n, dA);
    for (int k = 0; k < D; ++k) dn[k] = (dA[k] - n[k] * ds) / s;
    double eL[MAXV], eR[MAXV];
    memcpy(eL, dL, sizeof(double) * V);
    memcpy(eR, dR, sizeof(double) * V);
    const double nl = dot(D, n, lw), dnl = dot(D, dn, lw), nr = dot(D, n, rw), dnr = dot(D, dn, rw);
    for (int k = 0; k < D; ++k) {
        eL[1 + k] -= dw[k] + (dn[k] * nl - n[k] * dnl);
        eR[1 + k] -= dw[k] + (dn[k] * nr - n[k] * dnr);
    }
    double F[MAXV], dF[MAXV], Phi[MAXV], X[MAXV];
    riemann_tangent(kind, D, gas, L, R, eL, eR, n, w, F, dF);
    memcpy(Phi, F, sizeof(double) * V);
    to_lab(D, Phi, mw);
    const double nP = dot(D, n, Phi + 1), dnP = dot(D, dn, Phi + 1);
    X[0] = 0;
    X[P] = 0;
    for (int k = 0; k < D; ++k) X[1 + k] = dn[k] * nP - n[k] * dnP;
    to_lab(D, X, w);
    for (int v = 0; v < V; ++v) dF[v] += X[v];
    dF[P] += dot(D, dw, Phi + 1) + dot(D, w, dw) * Phi[0];
    for (int k = 0; k < D; ++k) dF[1 + k] += dw[k] * Phi[0];
    for (int v = 0; v < V; ++v) dG[v] = ds * F[v] + s * dF[v];
}

/* W, c_s tangents of cell i (finite_volume.hpp:2408-2426 moving_primitives): dU = (dq - U dV) / V. */
static void moving_prim_tangent(const orc_fv* fv, const orc_geometry* g, const geo_dots* t, const double* dq, int i, double* dW, double* dcs) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    const double* Wi = fv->prims + (size_t)i * PW;
    const double inv = 1.0 / g->volume[i];
    double U[MAXV], dU[MAXV];
    for (int v = 0; v < V; ++v) U[v] = fv->q[(size_t)i * V + v] * inv;
    for (int v = 0; v < V; ++v) dU[v] = (dq[(size_t)i * V + v] - U[v] * t->V[i]) * inv;
    orc_primitive_tangent(D, &fv->gas, U, dU, dW);
    const double cs = Wi[V];
    *dcs = cs > 0 ? 0.5 * fv->gas.gamma * (dW[D + 1] / Wi[0] - Wi[D + 1] * dW[0] / (Wi[0] * Wi[0])) / cs : 0.0;
}

/* finite_volume.hpp:2251-2306 Regularised::tangent(): of the point velocity w = v + c (fraction c_s / |c|), at the forward's
 * branches (fraction from the largest face angle, its slope zero outside the ramp). */
static void regularised_tangent(const orc_fv* fv, const orc_geometry* g, const geo_dots* t, int i, const double* dWi, double dcs, double* dw) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    const double* Wi = fv->prims + (size_t)i * PW;
    double amax = 0;
    int fmax = -1, other = -1;
    for (int k = g->start[i]; k < g->start[i + 1]; ++k) {
        const int e = g->list[k], f = list_face(e), j = e >= 0 ? g->neighbour[f] : g->owner[f];
        double d[MAXD];
        separation(g, i, j, d);
        const double angle = face_angle(D, magnitude_of(g, f)) / (0.5 * sqrt(dot(D, d, d)));
        if (angle > amax) {
            amax = angle;
            fmax = f;
            other = j;
        }
    }
    const double F = 2.25, shaping = 0.5;
    double fraction = 0, slope = 0;
    if (amax > 0.75 * F) {
        fraction = amax > F ? shaping : shaping * (amax - 0.75 * F) / (0.25 * F);
        slope = amax > F ? 0.0 : shaping / (0.25 * F);
    }
    for (int a = 0; a < D; ++a) dw[a] = dWi[1 + a];
    const double* c = g->centroid + (size_t)i * D;
    const double* dc = t->c + (size_t)i * D;
    const double cd = sqrt(dot(D, c, c)), cs = Wi[V];
    if (!(cd > 0 && fraction > 0)) return;
    const double k = fraction * cs / cd;
    double dfraction = 0;
    if (slope != 0) {
        const double* A = g->area + (size_t)fmax * D;
        double d[MAXD], dd[MAXD];
        separation(g, i, other, d);
        for (int a = 0; a < D; ++a) dd[a] = t->x[(size_t)other * D + a] - t->x[(size_t)i * D + a];
        const double mag = sqrt(dot(D, A, A)), dist = sqrt(dot(D, d, d));
        const double dmag = dot(D, A, t->A + (size_t)fmax * D) / mag, ddist = dot(D, d, dd) / dist;
        dfraction = slope * (face_angle_slope(D, mag) * dmag / (0.5 * dist) - amax * ddist / dist);
    }
    const double dcd = dot(D, c, dc) / cd;
    const double dk = (dfraction * cs + fraction * dcs) / cd - k * dcd / cd;
    for (int a = 0; a < D; ++a) dw[a] += dc[a] * k + c[a] * dk;
}

/* finite_volume.hpp:2687-2762 moving_gradient_tangent(): the limited fit's tangent in W and in the geometry it reads. The fit
 * x_v = M^-1 b_v (X_k = [1, d_k + c_j - c_i], weights |A_k|) moves by M^-1 (db_v - dM x_v); after the last clip
 * g = g0 (bound - W_i) / (g0 . o), o that face's offset (which moves with the geometry too). dW: all cells' W tangents. */
static void moving_fit_tangent(const orc_fv* fv, const orc_geometry* g, const geo_dots* t, int i, const double* dW, double* dg) {
    const int D = fv->D, V = D + 2, U = D + 1, PW = V + 2;
    const double* Wi = fv->prims + (size_t)i * PW;
    const double* dWi = dW + (size_t)i * V;
    double gl[MAXV * MAXD], g0[MAXV * MAXD], dg_unused[MAXV * MAXD], Lc[MAXD + 1][MAXD + 1], M0[MAXD + 1], icpt[MAXV];
    limit_pattern pat;
    fit_cell_x(fv, g, i, dW, gl, g0, dg_unused, Lc, M0, &pat, icpt);
    double dM[MAXD + 1][MAXD + 1], db[MAXV][MAXD + 1];
    memset(dM, 0, sizeof dM);
    memset(db, 0, sizeof db);
    const double* ci = g->centroid + (size_t)i * D;
    const int first = g->start[i];
    for (int k = first; k < g->start[i + 1]; ++k) {
        const int e = g->list[k], f = list_face(e), j = e >= 0 ? g->neighbour[f] : g->owner[f];
        const double* Wj = fv->prims + (size_t)j * PW;
        double d[MAXD], X[MAXD + 1], dX[MAXD + 1];
        separation(g, i, j, d);
        const double w = magnitude_of(g, f);
        const double dwt = w > 0 ? dot(D, g->area + (size_t)f * D, t->A + (size_t)f * D) / w : 0.0;
        X[0] = 1;
        dX[0] = 0;
        for (int p = 1; p < U; ++p) {
            X[p] = d[p - 1] + g->centroid[(size_t)j * D + p - 1] - ci[p - 1];
            dX[p] = (t->x[(size_t)j * D + p - 1] - t->x[(size_t)i * D + p - 1]) + t->c[(size_t)j * D + p - 1] - t->c[(size_t)i * D + p - 1];
        }
        for (int p = 0; p < U; ++p)
            for (int q = 0; q < U; ++q) dM[p][q] += dwt * X[p] * X[q] + w * (dX[p] * X[q] + X[p] * dX[q]);
        for (int v = 0; v < V; ++v) {
            const double delta = Wj[v] - Wi[v], ddelta = dW[(size_t)j * V + v] - dWi[v];
            for (int p = 0; p < U; ++p) db[v][p] += dwt * X[p] * delta + w * dX[p] * delta + w * X[p] * ddelta;
        }
    }
    double rhs[MAXV][MAXD + 1], dx[MAXV][MAXD + 1];
    for (int v = 0; v < V; ++v)
        for (int p = 0; p < U; ++p) {
            double mx = dM[p][0] * icpt[v];
            for (int a = 0; a < D; ++a) mx += dM[p][1 + a] * g0[v * D + a];
            rhs[v][p] = db[v][p] - mx;
        }
    solve_factored(D, Lc, rhs, dx);
    for (int v = 0; v < V; ++v) {
        const double* gv = g0 + v * D;
        double dg0[MAXD];
        for (int a = 0; a < D; ++a) dg0[a] = dx[v][1 + a];
        if (pat.last[v] < 0) {
            for (int a = 0; a < D; ++a) dg[v * D + a] = dg0[a];
        } else if (pat.zero[v]) {
            for (int a = 0; a < D; ++a) dg[v * D + a] = 0.0;
        } else {
            const double* o = pat.offset[v];
            double dor[MAXD];
            offset_dot(g, t, list_face(g->list[first + pat.last[v]]), i, dor);
            const double den = dot(D, gv, o);
            const double s = (pat.bound[v] - Wi[v]) / den;
            const double ds = (pat.dbound[v] - dWi[v] - s * (dot(D, dg0, o) + dot(D, gv, dor))) / den;
            for (int a = 0; a < D; ++a) dg[v * D + a] = dg0[a] * s + gv[a] * ds;
        }
    }
}

/* One moving pass's tangent (finite_volume.hpp:2829-2870 moving_pass_tangent): dq -= dt/2 sum_k sign_k dG_k per cell in list
 * order, G = |A| F(side_o, side_j, A / |A|, face_velocity(d, x_f, w-hat_o, w-hat_j)). First pass (mesh t^n): sides W + g.r;
 * second (mesh t^n+1): half_step(W, g, 2 dt) + g.(r + c^n+1 - c^n + dt w-hat), with c^n's tangent dc0. */
static void moving_pass_tangent(const orc_fv* fv, const orc_geometry* g, const geo_dots* t, int second, double dt, const double* wh, const double* dwh,
                                const double* dW, const double* dg, const double* dc0, double* dq) {
    const int D = fv->D, V = D + 2, PW = V + 2;
    double* dG = (double*)calloc((size_t)(g->F > 0 ? g->F : 1) * V, sizeof(double));
#pragma omp parallel for schedule(static)
    for (int f = 0; f < g->F; ++f) {
        const int o = g->owner[f], j = g->neighbour[f];
        const double* A = g->area + (size_t)f * D;
        if (sqrt(dot(D, A, A)) == 0) continue;
        double d[MAXD], dd[MAXD], wf[MAXD], dwf[MAXD], sides[2][MAXV], dsides[2][MAXV];
        separation(g, o, j, d);
        for (int a = 0; a < D; ++a) dd[a] = t->x[(size_t)j * D + a] - t->x[(size_t)o * D + a];
        face_velocity(g, f, o, j, wh + (size_t)o * D, wh + (size_t)j * D, wf);
        orc_face_velocity_tangent(D, d, g->face_centroid + (size_t)f * D, wh + (size_t)o * D, wh + (size_t)j * D, dd, t->xf + (size_t)f * D,
                                  dwh + (size_t)o * D, dwh + (size_t)j * D, dwf);
        const int cells[2] = {o, j};
        for (int s2 = 0; s2 < 2; ++s2) {
            const int i = cells[s2];
            const double* W = fv->prims + (size_t)i * PW;
            const double* gi = fv->grad + (size_t)i * V * D;
            const double* dWi = dW + (size_t)i * V;
            const double* dgi = dg + (size_t)i * V * D;
            double r[MAXD], dr[MAXD];
            face_offset(g, f, i, r);
            offset_dot(g, t, f, i, dr);
            if (second) {
                for (int a = 0; a < D; ++a) {
                    r[a] = r[a] + g->centroid[(size_t)i * D + a] - fv->old_centroid[(size_t)i * D + a] + wh[(size_t)i * D + a] * (0.5 * (2 * dt));
                    dr[a] = dr[a] + t->c[(size_t)i * D + a] - dc0[(size_t)i * D + a] + dwh[(size_t)i * D + a] * dt;
                }
                orc_half_step(D, &fv->gas, W, gi, 2 * dt, sides[s2]);
                orc_half_step_tangent(D, &fv->gas, W, dWi, gi, dgi, 2 * dt, dsides[s2]);
            } else {
                for (int v = 0; v < V; ++v) {
                    sides[s2][v] = W[v];
                    dsides[s2][v] = dWi[v];
                }
            }
            for (int v = 0; v < V; ++v) sides[s2][v] += dot(D, gi + v * D, r);
            if (sides[s2][0] > 0 && sides[s2][D + 1] > 0) {
                for (int v = 0; v < V; ++v) dsides[s2][v] += dot(D, dgi + v * D, r) + dot(D, gi + v * D, dr);
            } else {
                for (int v = 0; v < V; ++v) {
                    sides[s2][v] = W[v];
                    dsides[s2][v] = dWi[v];
                }
            }
        }
        orc_moving_face_tangent(fv->riemann, D, &fv->gas, sides[0], sides[1], A, wf, dsides[0], dsides[1], t->A + (size_t)f * D, dwf,
                                dG + (size_t)f * V);
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < fv->N; ++i) {
        double sum[MAXV] = {0, 0, 0, 0, 0};
        for (int k = g->start[i]; k < g->start[i + 1]; ++k) {
            const int e = g->list[k], f = list_face(e);
            const double sign = e >= 0 ? 1.0 : -1.0;
            for (int v = 0; v < V; ++v) sum[v] += sign * dG[(size_t)f * V + v];
        }
        for (int v = 0; v < V; ++v) dq[(size_t)i * V + v] -= 0.5 * dt * sum[v];
    }
    free(dG);
}

static double* dots_alloc(size_t n) { return (double*)calloc(n > 0 ? n : 1, sizeof(double)); }

int orc_moving_tangent(orc_fv* fv, const orc_geometry* g0, const orc_geometry* g1, double dt, int steps_fixed, double* dq, double* dx,
                       orc_geometry_tangent tangent0, void* ctx0, orc_geometry_tangent tangent1, void* ctx1) {
    if (!fv->moving || fv->D < 2 || fv->precision != 0 || fv->hll_compression > 0 || fv->limiter == 2) return -1;
    if (fv->order != 0 && fv->order != 1) return -1;
    const int D = fv->D, V = D + 2, N = fv->N;
    const size_t ND = (size_t)N * D;
    /* t^n: W, the limited gradients and the point velocities (the sweep), c^n, the grid steps and w-hat */
    orc_primitives(fv, g0);
    for (int i = 0; i < N; ++i) sweep_cell(fv, g0, i);
    if (fv->order == 0) memset(fv->grad, 0, sizeof(double) * ND * V);
    double* wh = dots_alloc(ND);
    if (fv->order == 0) {
        memcpy(wh, fv->w, sizeof(double) * ND);
    } else {
        const double unit = g0->step / dt;
        for (int i = 0; i < N; ++i)
            for (int a = 0; a < D; ++a) {
                const size_t at = (size_t)i * D + a;
                fv->old_centroid[at] = g0->centroid[at];
                if (!fv->hold_steps) {
                    const long long k0 = orc_snap(fv->x[at], g0->step, g0->period[a]);
                    const long long k1 = orc_snap(fv->x[at] + dt * fv->w[at], g0->step, g0->period[a]);
                    fv->steps[at] = orc_delta(k0, k1, g0->period[a]);
                }
                wh[at] = (double)fv->steps[at] * unit;
            }
    }
    const int F0 = g0->F, F1 = g1->F;
    double *dV0 = dots_alloc(N), *dc0 = dots_alloc(ND), *dA0 = dots_alloc((size_t)F0 * D), *dxf0 = dots_alloc((size_t)F0 * D);
    tangent0(ctx0, dx, dV0, dc0, dA0, dxf0);
    const geo_dots t0 = {dx, dV0, dc0, dA0, dxf0};
    double *dW = dots_alloc((size_t)N * V), *dcs = dots_alloc(N), *dw = dots_alloc(ND), *dg = dots_alloc((size_t)N * V * D);
    for (int i = 0; i < N; ++i) moving_prim_tangent(fv, g0, &t0, dq, i, dW + (size_t)i * V, dcs + i);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        regularised_tangent(fv, g0, &t0, i, dW + (size_t)i * V, dcs[i], dw + (size_t)i * D);
        if (fv->order == 1) moving_fit_tangent(fv, g0, &t0, i, dW, dg + (size_t)i * V * D);
    }
    if (fv->order == 1 && steps_fixed) memset(dw, 0, sizeof(double) * ND); /* the grid steps held: no derivative through w-hat */
    double* dx1 = dots_alloc(ND);
    if (fv->order == 0) {
        /* finite_volume.hpp:2428-2486: one pass over dt on t^n with the point velocities (sides W), then x += dt w */
        moving_pass_tangent(fv, g0, &t0, 0, 2 * dt, wh, dw, dW, dg, NULL, dq);
        for (size_t k = 0; k < ND; ++k) dx[k] = dx[k] + dt * dw[k];
    } else {
        moving_pass_tangent(fv, g0, &t0, 0, dt, wh, dw, dW, dg, NULL, dq);
        for (size_t k = 0; k < ND; ++k) dx1[k] = dx[k] + dt * dw[k];
        double *dV1 = dots_alloc(N), *dc1 = dots_alloc(ND), *dA1 = dots_alloc((size_t)F1 * D), *dxf1 = dots_alloc((size_t)F1 * D);
        tangent1(ctx1, dx1, dV1, dc1, dA1, dxf1);
        const geo_dots t1 = {dx1, dV1, dc1, dA1, dxf1};
        moving_pass_tangent(fv, g1, &t1, 1, dt, wh, dw, dW, dg, dc0, dq);
        memcpy(dx, dx1, sizeof(double) * ND);
        free(dV1);
        free(dc1);
        free(dA1);
        free(dxf1);
    }
    free(dx1);
    free(wh);
    free(dV0);
    free(dc0);
    free(dA0);
    free(dxf0);
    free(dW);
    free(dcs);
    free(dw);
    free(dg);
    return 0;
}
