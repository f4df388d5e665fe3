// The learned slope limiter (no counterpart in luna-cpp; it acts on the gradients of the face-centroid limiter,
// finite_volume.hpp:538-565): per cell and variable, 16 features recorded by the sweep, an MLP 16 -> 64 -> 64 -> 1
// (ReLU, BF16 operands with FP32 accumulation on MFMA), and the gradient scaled by sigmoid(z) in [0, 1], so the scaled
// gradient is never steeper than the face-centroid limiter's and every face-centroid value stays within the neighbours'
// range. A troubled cell (a compressive pressure jump, or a neighbour at under a tenth of the cell's density or
// pressure) keeps the face-centroid limiter's gradient.
#pragma once

#include "physics.hpp"

namespace luna {

constexpr int FEATURES = 16, HIDDEN = 64;

// The network's parameters, FP32, in this order (the C-ABI's luna_limiter_params layout):
//   W1 [64][16], b1 [64], W2 [64][64], b2 [64], w3 [64], b3
// h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), z = w3 . h2 + b3, fraction = 1 / (1 + exp(-z)).
constexpr int P_W1 = 0, P_B1 = P_W1 + HIDDEN * FEATURES, P_W2 = P_B1 + HIDDEN, P_B2 = P_W2 + HIDDEN * HIDDEN, P_W3 = P_B2 + HIDDEN,
              P_B3 = P_W3 + HIDDEN, PARAMS = P_B3 + 1;

// The device copy the kernels read: W1 and W2 as BF16 bits, the rest FP32.
struct Network {
    const unsigned short* w1;  // [64][16]
    const unsigned short* w2;  // [64][64]
    const float* b1;           // [64]
    const float* b2;           // [64]
    const float* w3;           // [64]
    const float* b3;           // [1]
};

// Round to BF16 bits: double to float (nearest even), then float to BF16 (nearest even); NaN stays a quiet NaN.
__host__ __device__ inline unsigned short bf16_bits(float f) {
    unsigned u = __builtin_bit_cast(unsigned, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__host__ __device__ inline unsigned short bf16_bits(double x) { return bf16_bits((float)x); }

// What the limiter walk records for the features (On = false: nothing).
template <int D, bool On>
struct Record {
    __device__ __forceinline__ void init(const double (&)[D + 2][D]) {}
    __device__ __forceinline__ void face(const double*) {}
    __device__ __forceinline__ void clip(int, double) {}
};
template <int D>
struct Record<D, true> {
    static constexpr int V = D + 2;
    double g0[V][D];  // the unlimited gradient
    double phi[V];    // the product of the clip factors: the limited gradient is phi g0
    double up[V];     // the largest g0 . (x_f - c) over the faces the limiter visits, at least 0
    double down[V];   // the smallest, at most 0
    int clips[V];
    __device__ __forceinline__ void init(const double (&g)[V][D]) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
#pragma unroll
            for (int a = 0; a < D; ++a) g0[v][a] = g[v][a];
            phi[v] = 1.0;
            up[v] = 0.0;
            down[v] = 0.0;
            clips[v] = 0;
        }
    }
    __device__ __forceinline__ void face(const double* offset) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double dp = dot<D>(g0[v], offset);
            up[v] = fmax(up[v], dp);
            down[v] = fmin(down[v], dp);
        }
    }
    __device__ __forceinline__ void clip(int v, double sc) {
        phi[v] = phi[v] * sc;
        ++clips[v];
    }
};

// The cell's own values the features use: W, the neighbours' range (after the tiny-face exclusion), the fit's intercept
// (the fitted value at the centroid minus W: zero for linear data), sound speed, radius, list length.
template <int D>
struct Cell {
    double W[D + 2], lo[D + 2], hi[D + 2], intercept[D + 2], cs, r;
    int count;
};

__host__ __device__ inline double clamp8(double x) { return fmin(fmax(x, -8.0), 8.0); }

// The 16 features of each variable v (row v of out, BF16 bits), all scale-free:
//   0 phi_v   1 (hi - W)/S   2 (W - lo)/S   3 up/S   4 -down/S   5 intercept/S   6 clips/count   7 min over v of phi
//   8 (hi_rho - lo_rho)/rho   9 (hi_P - lo_P)/P   10 div(v) r/c   11 |curl v| r/c   12 |v|/c   13 [v is density]
//   14 [v is pressure]   15 S/scale
// with S = hi - lo + 1e-12 scale + 1e-300 and scale = c for a velocity, |W| otherwise; each clamped to [-8, 8], and a
// zero is always +0.
// Returns the troubled flag: div(v) r/c < -0.05 with a pressure jump above 0.5, or lo_rho < rho/10, or lo_P < P/10.
template <int D>
__device__ inline unsigned char limiter_features(const Cell<D>& c, const Record<D, true>& rec, unsigned short* out) {
    constexpr int V = D + 2;
    const double inv_c = 1.0 / c.cs;
    double speed2 = 0, div = 0;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        speed2 += c.W[1 + a] * c.W[1 + a];
        div += rec.g0[1 + a][a];
    }
    double curl;
    if constexpr (D == 2) {
        curl = fabs(rec.g0[2][0] - rec.g0[1][1]);
    } else {
        const double cx = rec.g0[3][1] - rec.g0[2][2], cy = rec.g0[1][2] - rec.g0[3][0], cz = rec.g0[2][0] - rec.g0[1][1];
        curl = sqrt(cx * cx + cy * cy + cz * cz);
    }
    double phimin = rec.phi[0];
#pragma unroll
    for (int v = 1; v < V; ++v) phimin = fmin(phimin, rec.phi[v]);
    const double rho_jump = (c.hi[0] - c.lo[0]) / c.W[0], p_jump = (c.hi[V - 1] - c.lo[V - 1]) / c.W[V - 1];
    const double compression = div * c.r * inv_c;
    const double f8 = clamp8(rho_jump), f9 = clamp8(p_jump), f10 = clamp8(compression), f11 = clamp8(curl * c.r * inv_c),
                 f12 = clamp8(sqrt(speed2) * inv_c);
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const double scale = (v > 0 && v < V - 1) ? c.cs : fabs(c.W[v]);
        const double S = (c.hi[v] - c.lo[v]) + 1e-12 * scale + 1e-300;
        const double inv = 1.0 / S;
        double f[FEATURES];
        f[0] = rec.phi[v];
        f[1] = clamp8((c.hi[v] - c.W[v]) * inv);
        f[2] = clamp8((c.W[v] - c.lo[v]) * inv);
        f[3] = clamp8(rec.up[v] * inv);
        f[4] = clamp8(-rec.down[v] * inv);
        f[5] = clamp8(c.intercept[v] * inv);
        f[6] = (double)rec.clips[v] / (double)c.count;
        f[7] = phimin;
        f[8] = f8;
        f[9] = f9;
        f[10] = f10;
        f[11] = f11;
        f[12] = f12;
        f[13] = v == 0 ? 1.0 : 0.0;
        f[14] = v == V - 1 ? 1.0 : 0.0;
        f[15] = clamp8(S / (scale + 1e-300));
        // + 0.0 turns -0 into +0: min and max may return either zero of a tie between +0 and -0 (IEEE leaves it
        // open, and the host's library and the GPU choose differently), so a feature's bits never depend on it
        unsigned packed[FEATURES / 2];
#pragma unroll
        for (int k = 0; k < FEATURES / 2; ++k)
            packed[k] = (unsigned)bf16_bits(f[2 * k] + 0.0) | ((unsigned)bf16_bits(f[2 * k + 1] + 0.0) << 16);
        uint4* row = reinterpret_cast<uint4*>(out + v * FEATURES);
        row[0] = make_uint4(packed[0], packed[1], packed[2], packed[3]);
        row[1] = make_uint4(packed[4], packed[5], packed[6], packed[7]);
    }
    const bool shock = compression < -0.05 && p_jump > 0.5;
    const bool thin = c.lo[0] < 0.1 * c.W[0] || c.lo[V - 1] < 0.1 * c.W[V - 1];
    return (shock || thin) ? 1 : 0;
}

// Learned-limiter kernels (limiter_kernels.hip). rows = N V feature rows (row i V + v), 16 BF16 each, 32-B aligned.
// apply: fraction = sigmoid(MLP(x)) per row; unless the row's cell is troubled, grad[i * grad_row + v * D + a] *= fraction.
// grad may be null (then only fraction, if not null, is written).
hipError_t launch_limiter_apply(hipStream_t s, const Network& net, const unsigned short* features, const unsigned char* troubled, int rows,
                                int V, int D, int grad_row, double* grad, float* fraction);
// Training: the weighted squared error sum_r w_r (sigmoid(z_r) - t_r)^2 * inv_weight and its gradient in the parameters.
// partials: one row of PARTIAL_ROW floats per wave (gradient in the parameter order, the loss at PARAMS); waves: the
// grid's wave count (fixed, so the sums' order is fixed). reduce: sums the waves in order into out[PARAMS + 1] or, with
// adam, takes one Adam step of master (FP32) and writes the network copy.
constexpr int PARTIAL_ROW = 5376;
struct Batch {
    const unsigned short* x;
    const float* target;
    const float* weight;
    int rows;
    float inv_weight;
};
struct Adam {
    float* master;  // PARAMS
    float* m;
    float* v;
    float lr, beta1, beta2, eps;
    float c1, c2;   // the bias corrections 1 / (1 - beta^t) of step t
    unsigned short* w1;
    unsigned short* w2;
    float* rest;    // b1, b2, w3, b3 of the network copy, contiguous
};
int mlp_waves(int rows);
hipError_t launch_mlp_partials(hipStream_t s, const Network& net, const Batch& b, float* partials, int waves);
hipError_t launch_mlp_reduce(hipStream_t s, const float* partials, int waves, float* out, const Adam* adam);

}  // namespace luna
