// The learned limiter's network on CDNA4 matrix cores: inference fused with the gradient scaling, and the training
// kernels (forward, backward, deterministic reduction, Adam), BF16 operands with FP32 accumulation
// (v_mfma_f32_16x16x16_bf16, and gfx950's double-rate v_mfma_f32_16x16x32_bf16 for the 64 x 64 layer).
//
// Every product is computed transposed, rows in the MFMA's N position: a tile is 16 feature rows, and lane l (group
// g = l / 16, row n = l % 16) holds row n's hidden units 16 b + 4 g + j (j < 4) of each 16-unit block b. That is the
// MFMA's output layout and also its B-operand layout for the next layer's 16-wide slice of the contraction, so the
// activations go from one layer to the next in registers. The weight gradients contract over rows, so their operands
// go through LDS once per tile.
#include <algorithm>

#include "kernels.hpp"

namespace luna {
namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef short short8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// D = A B + C for 16 x 16 tiles: lane l gives A[l % 16][4 (l / 16) + j] and B[4 (l / 16) + j][l % 16] and gets
// D[4 (l / 16) + j][l % 16].
__device__ __forceinline__ float4_t mfma16(short4_t a, short4_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// D = A B + C over a contraction of 32: lane l gives 8 values of A's row l % 16 and 8 of B's column l % 16, at the same 8
// contraction indices (for A and B alike) chosen by l / 16 and the element. A sum over all of them does not depend on
// which indices those are, so the contraction slots can carry any fixed assignment of hidden units (see Weights::w2).
__device__ __forceinline__ float4_t mfma32(short8_t a, short8_t b, float4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
}

// FP32 to BF16 bits, nearest even (one v_cvt_pk_bf16_f32 on gfx950; the values here are finite).
__device__ __forceinline__ short bf16s(float f) { return __builtin_bit_cast(short, (__bf16)f); }

__device__ __forceinline__ short8_t join(short4_t a, short4_t b) {  // as 32-bit words: no byte shuffles
    const uint2 x = __builtin_bit_cast(uint2, a), y = __builtin_bit_cast(uint2, b);
    return __builtin_bit_cast(short8_t, make_uint4(x.x, x.y, y.x, y.y));
}

__device__ __forceinline__ short4_t load4(const unsigned short* p) {  // 4 BF16, 8-byte aligned
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    short4_t r;
    r[0] = (short)(u.x & 0xffffu);
    r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xffffu);
    r[3] = (short)(u.y >> 16);
    return r;
}
__device__ __forceinline__ float4_t loadf4(const float* p) {  // 16-byte aligned
    const float4 u = *reinterpret_cast<const float4*>(p);
    float4_t r;
    r[0] = u.x;
    r[1] = u.y;
    r[2] = u.z;
    r[3] = u.w;
    return r;
}

// This lane's operands of the forward pass. Layer 2 contracts over hidden-1 in two chunks of 32: the lane's 8 slots of
// chunk c carry units 32 c + 4 g + j (its layer-1 block 2 c) and 32 c + 16 + 4 g + j (block 2 c + 1), j < 4, which is
// what the lane holds after layer 1, in A and B alike.
struct Weights {
    short4_t w1[4];     // A of layer 1, block b: W1[16 b + n][4 g + j]
    short8_t w2[4][2];  // A of layer 2, [out block o][chunk c]: W2[16 o + n][32 c + 4 g + j], W2[16 o + n][32 c + 16 + 4 g + j]
    float4_t b1[4], b2[4], w3[4];  // b1[16 b + 4 g + j], ...
    float b3;
};
__device__ __forceinline__ void load_weights(const Network& net, int g, int n, Weights& W) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        W.w1[b] = load4(net.w1 + (16 * b + n) * FEATURES + 4 * g);
#pragma unroll
        for (int c = 0; c < 2; ++c)
            W.w2[b][c] = join(load4(net.w2 + (16 * b + n) * HIDDEN + 32 * c + 4 * g), load4(net.w2 + (16 * b + n) * HIDDEN + 32 * c + 16 + 4 * g));
        W.b1[b] = loadf4(net.b1 + 16 * b + 4 * g);
        W.b2[b] = loadf4(net.b2 + 16 * b + 4 * g);
        W.w3[b] = loadf4(net.w3 + 16 * b + 4 * g);
    }
    W.b3 = net.b3[0];
}

// One tile's forward pass. x: this lane's 4 features 4 g .. 4 g + 3 of row n.
struct Tile {
    float4_t pre1[4];  // layer-1 pre-activations (hidden 16 b + 4 g + j of row n)
    short4_t h1[4];    // relu, BF16
    float4_t pre2[4];
    float z;           // the output of row n (all four groups hold it)
};
__device__ __forceinline__ void forward(const Weights& W, short4_t x, Tile& t) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        t.pre1[b] = mfma16(W.w1[b], x, W.b1[b]);
#pragma unroll
        for (int j = 0; j < 4; ++j) t.h1[b][j] = bf16s(fmaxf(t.pre1[b][j], 0.0f));
    }
    const short8_t h1c[2] = {join(t.h1[0], t.h1[1]), join(t.h1[2], t.h1[3])};
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        float4_t acc = W.b2[o];
#pragma unroll
        for (int c = 0; c < 2; ++c) acc = mfma32(W.w2[o][c], h1c[c], acc);
        t.pre2[o] = acc;
    }
    float part = 0.0f;
#pragma unroll
    for (int o = 0; o < 4; ++o)
#pragma unroll
        for (int j = 0; j < 4; ++j) part += W.w3[o][j] * fmaxf(t.pre2[o][j], 0.0f);
    part += __shfl_xor(part, 16);
    part += __shfl_xor(part, 32);
    t.z = part + W.b3;
}

__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + expf(-z)); }

// A persistent grid: each wavefront loads the weights once and takes tiles wave, wave + waves, ...
constexpr int APPLY_BLOCK = 256, APPLY_WAVES = 256 * 16;  // 256 CUs, 4 wavefronts per SIMD (126 registers)

__global__ __launch_bounds__(APPLY_BLOCK) void k_limiter_apply(Network net, const unsigned short* x, const unsigned char* troubled, int rows,
                                                               int V, int D, int grad_row, double* grad, float* fraction) {
    const int lane = threadIdx.x & 63, g = lane >> 4, n = lane & 15;
    const int wave = blockIdx.x * (APPLY_BLOCK / 64) + (threadIdx.x >> 6), waves = gridDim.x * (APPLY_BLOCK / 64);
    const int tiles = (rows + 15) >> 4;
    if (wave >= tiles) return;  // the whole wavefront
    Weights W;
    load_weights(net, g, n, W);
    for (int t = wave; t < tiles; t += waves) {  // uniform over the wavefront: the MFMAs run with every lane
        const int row = t * 16 + n;
        short4_t xv = {0, 0, 0, 0};
        if (row < rows) xv = load4(x + (size_t)row * FEATURES + 4 * g);
        Tile tl;
        forward(W, xv, tl);
        if (row < rows) {
            const float frac = sigmoid(tl.z);  // every group holds z: group g scales component g of the row's gradient
            if (fraction && g == 0) fraction[row] = frac;
            if (grad && g < D) {
                const int i = row / V, v = row - i * V;
                if (!troubled[i]) {
                    double* p = grad + (size_t)i * grad_row + v * D + g;
                    *p = *p * (double)frac;
                }
            }
        }
    }
}

// Per wavefront: the gradient of sum_r w_r (sigmoid(z_r) - t_r)^2 inv_weight over tiles wave, wave + waves, ..., and
// the loss, into its row of partials.
__global__ __launch_bounds__(64) void k_mlp_partials(Network net, Batch b, float* partials) {
    __shared__ short h1s[16][HIDDEN], dh2s[16][HIDDEN], dh1s[16][HIDDEN], xs[16][FEATURES];
    __shared__ short w2s[HIDDEN][HIDDEN];  // W2, read transposed as the A operand of dh1
    const int lane = threadIdx.x, g = lane >> 4, n = lane & 15;
    const int wave = blockIdx.x, waves = gridDim.x;
    const int tiles = (b.rows + 15) >> 4;
    Weights W;
    load_weights(net, g, n, W);
    for (int k = lane; k < HIDDEN * HIDDEN; k += 64) w2s[k / HIDDEN][k % HIDDEN] = (short)net.w2[k];
    __syncthreads();
    const float4_t zero = {0.0f, 0.0f, 0.0f, 0.0f};
    float4_t gW1[4], gW2[4][4], gb1[4], gb2[4], gw3[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        gW1[k] = zero;
        gb1[k] = zero;
        gb2[k] = zero;
        gw3[k] = zero;
#pragma unroll
        for (int m = 0; m < 4; ++m) gW2[k][m] = zero;
    }
    float gb3 = 0.0f, loss = 0.0f;
    for (int t = wave; t < tiles; t += waves) {
        const int row = t * 16 + n;
        const bool valid = row < b.rows;
        short4_t xv = {0, 0, 0, 0};
        if (valid) xv = load4(b.x + (size_t)row * FEATURES + 4 * g);
        Tile tl;
        forward(W, xv, tl);
        const float sig = sigmoid(tl.z);
        const float target = valid ? b.target[row] : 0.0f, weight = valid ? b.weight[row] : 0.0f;
        const float diff = sig - target;
        const float dz = 2.0f * weight * diff * sig * (1.0f - sig) * b.inv_weight;
        if (g == 0) {
            loss += weight * diff * diff * b.inv_weight;
            gb3 += dz;
        }
        short4_t dh2b[4];
#pragma unroll
        for (int o = 0; o < 4; ++o)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float pre = tl.pre2[o][j];
                const float d = pre > 0.0f ? dz * W.w3[o][j] : 0.0f;
                gw3[o][j] += dz * fmaxf(pre, 0.0f);
                gb2[o][j] += d;
                dh2b[o][j] = bf16s(d);
            }
        short4_t dh1b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float4_t acc = zero;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                short4_t a;  // W2[16 m + 4 g + j][16 k + n]
#pragma unroll
                for (int j = 0; j < 4; ++j) a[j] = w2s[16 * m + 4 * g + j][16 * k + n];
                acc = mfma16(a, dh2b[m], acc);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float d = tl.pre1[k][j] > 0.0f ? acc[j] : 0.0f;
                gb1[k][j] += d;
                dh1b[k][j] = bf16s(d);
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                h1s[n][16 * k + 4 * g + j] = tl.h1[k][j];
                dh2s[n][16 * k + 4 * g + j] = dh2b[k][j];
                dh1s[n][16 * k + 4 * g + j] = dh1b[k][j];
            }
#pragma unroll
        for (int j = 0; j < 4; ++j) xs[n][4 * g + j] = xv[j];
        __syncthreads();
        // dW2[16 m + 4 g + j][16 k + n] += sum over the tile's rows r of dh2[r][16 m + .] h1[r][16 k + .]
        short4_t xr;
#pragma unroll
        for (int j = 0; j < 4; ++j) xr[j] = xs[4 * g + j][n];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            short4_t a, a1;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                a[j] = dh2s[4 * g + j][16 * m + n];
                a1[j] = dh1s[4 * g + j][16 * m + n];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                short4_t bb;
#pragma unroll
                for (int j = 0; j < 4; ++j) bb[j] = h1s[4 * g + j][16 * k + n];
                gW2[m][k] = mfma16(a, bb, gW2[m][k]);
            }
            gW1[m] = mfma16(a1, xr, gW1[m]);  // dW1[16 m + 4 g + j][n]
        }
        __syncthreads();
    }
    float* out = partials + (size_t)wave * PARTIAL_ROW;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int h = 16 * m + 4 * g + j;
            out[P_W1 + h * FEATURES + n] = gW1[m][j];
#pragma unroll
            for (int k = 0; k < 4; ++k) out[P_W2 + h * HIDDEN + 16 * k + n] = gW2[m][k][j];
        }
    // The bias and w3 sums over the 16 rows of each group, the scalars over the wavefront.
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float s1 = gb1[k][j], s2 = gb2[k][j], s3 = gw3[k][j];
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                s1 += __shfl_xor(s1, o);
                s2 += __shfl_xor(s2, o);
                s3 += __shfl_xor(s3, o);
            }
            if (n == 0) {
                out[P_B1 + 16 * k + 4 * g + j] = s1;
                out[P_B2 + 16 * k + 4 * g + j] = s2;
                out[P_W3 + 16 * k + 4 * g + j] = s3;
            }
        }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        gb3 += __shfl_xor(gb3, o);
        loss += __shfl_xor(loss, o);
    }
    if (lane == 0) {
        out[P_B3] = gb3;
        out[PARAMS] = loss;
    }
}

// Sums the partial rows in a fixed order (so a batch always gives the same bits), then either writes the sums (gradient
// and loss) or takes one Adam step of the FP32 parameters and refreshes the network copy (loss into out[0] if out).
// A work-group takes 64 parameters; its 16 wavefronts each sum one sixteenth of the rows in row order, and the first
// wavefront adds the 16 sums in wavefront order (one long chain of dependent loads per parameter took 175 us).
constexpr int REDUCE_SPLIT = 16;
__global__ __launch_bounds__(64 * REDUCE_SPLIT) void k_mlp_reduce(const float* partials, int waves, float* out, Adam a, int adam) {
    __shared__ float part[REDUCE_SPLIT][64];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    const int per = (waves + REDUCE_SPLIT - 1) / REDUCE_SPLIT, w0 = q * per, w1 = min(waves, w0 + per);
    float s = 0.0f;
    if (p <= PARAMS) {
#pragma unroll 8
        for (int w = w0; w < w1; ++w) s += partials[(size_t)w * PARTIAL_ROW + p];
    }
    part[q][lane] = s;
    __syncthreads();
    if (q != 0 || p > PARAMS) return;
    s = part[0][lane];
    for (int k = 1; k < REDUCE_SPLIT; ++k) s += part[k][lane];
    if (!adam) {
        out[p] = s;
        return;
    }
    if (p == PARAMS) {
        if (out) out[0] = s;
        return;
    }
    const float m = a.beta1 * a.m[p] + (1.0f - a.beta1) * s;
    const float v = a.beta2 * a.v[p] + (1.0f - a.beta2) * s * s;
    a.m[p] = m;
    a.v[p] = v;
    const float value = a.master[p] - a.lr * (m * a.c1) / (sqrtf(v * a.c2) + a.eps);
    a.master[p] = value;
    if (p < P_B1)
        a.w1[p - P_W1] = bf16_bits(value);
    else if (p < P_W2)
        a.rest[p - P_B1] = value;
    else if (p < P_B2)
        a.w2[p - P_W2] = bf16_bits(value);
    else
        a.rest[HIDDEN + (p - P_B2)] = value;
}

}  // namespace

hipError_t launch_limiter_apply(hipStream_t s, const Network& net, const unsigned short* features, const unsigned char* troubled, int rows,
                                int V, int D, int grad_row, double* grad, float* fraction) {
    if (rows <= 0) return hipSuccess;
    if (D > 4) return hipErrorInvalidValue;  // one gradient component per lane group
    const long long tiles = (rows + 15) / 16, waves = std::min<long long>(tiles, APPLY_WAVES);
    const unsigned blocks = (unsigned)((waves + APPLY_BLOCK / 64 - 1) / (APPLY_BLOCK / 64));
    hipLaunchKernelGGL(k_limiter_apply, dim3(blocks), dim3(APPLY_BLOCK), 0, s, net, features, troubled, rows, V, D, grad_row, grad, fraction);
    return hipGetLastError();
}

int mlp_waves(int rows) {
    const int tiles = (rows + 15) / 16;
    return std::max(1, std::min(1024, (tiles + 3) / 4));
}

hipError_t launch_mlp_partials(hipStream_t s, const Network& net, const Batch& b, float* partials, int waves) {
    hipLaunchKernelGGL(k_mlp_partials, dim3(waves), dim3(64), 0, s, net, b, partials);
    return hipGetLastError();
}

hipError_t launch_mlp_reduce(hipStream_t s, const float* partials, int waves, float* out, const Adam* adam) {
    const Adam a = adam ? *adam : Adam{};
    hipLaunchKernelGGL(k_mlp_reduce, dim3((PARAMS + 1 + 63) / 64), dim3(64 * REDUCE_SPLIT), 0, s, partials, waves, out, a, adam ? 1 : 0);
    return hipGetLastError();
}

}  // namespace luna
