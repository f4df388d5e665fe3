"""CPU restatement of the learned limiter's network (test infrastructure: the checker of the device kernels in
luna-cpp_amd/csrc/hydro/limiter_kernels.hip, never called by the product).

The network has no counterpart in luna-cpp. Parameters are FP32 in the order W1 [64][16], b1 [64], W2 [64][64], b2 [64],
w3 [64], b3; h1 = relu(W1 x + b1), h2 = relu(W2 h1 + b2), z = w3 . h2 + b3, fraction = sigmoid(z). The device rounds to
BF16 at the matrix-core operands: the features x, W1, W2, h1 (after relu), and, in the backward pass, dh2 (for dh1 and
dW2) and dh1 (for dW1); everything else is FP32. This restatement rounds at the same points and accumulates in float64,
so it differs from the device only by FP32 accumulation order (and the rare BF16 rounding that order tips).
The loss of a batch is sum_r w_r (sigmoid(z_r) - t_r)^2 / weight_sum.
"""
import numpy as np

FEATURES, HIDDEN = 16, 64
P_W1 = 0
P_B1 = P_W1 + HIDDEN * FEATURES
P_W2 = P_B1 + HIDDEN
P_B2 = P_W2 + HIDDEN * HIDDEN
P_W3 = P_B2 + HIDDEN
P_B3 = P_W3 + HIDDEN
PARAMS = P_B3 + 1


def bf16_bits(x):
    """float32 (after a float64 -> float32 cast, nearest even) to BF16 bits, nearest even; NaN stays quiet."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return r.astype(np.uint16)


def bf16_value(bits):
    """The float value of BF16 bits."""
    b = np.ascontiguousarray(np.asarray(bits, dtype=np.uint16)).astype(np.uint32) << 16
    return b.view(np.float32).astype(np.float64)


def bf16(x):
    return bf16_value(bf16_bits(x))


def unpack(params, rounding=True):
    """The parameter blocks, as FP32 values (rounding) or as given in float64."""
    p = np.asarray(params, dtype=np.float32).astype(np.float64) if rounding else np.asarray(params, dtype=np.float64)
    if p.shape != (PARAMS,):
        raise ValueError("expected %d parameters" % PARAMS)
    W1 = p[P_W1:P_B1].reshape(HIDDEN, FEATURES)
    W2 = p[P_W2:P_B2].reshape(HIDDEN, HIDDEN)
    return W1, p[P_B1:P_W2], W2, p[P_B2:P_W3], p[P_W3:P_B3], p[P_B3]


def _same(x):
    return np.asarray(x, dtype=np.float64)


def forward(params, x_bits, rounding=True):
    """z, fraction and the intermediates of each row (x_bits: rows x 16 BF16 bits). rounding=False keeps every value in
    float64 (the plain MLP, against which the derivative formulas are checked by finite differences)."""
    r = bf16 if rounding else _same
    W1, b1, W2, b2, w3, b3 = unpack(params, rounding)
    x = bf16_value(x_bits).reshape(-1, FEATURES)
    pre1 = x @ r(W1).T + b1
    h1 = r(np.maximum(pre1, 0.0))
    pre2 = h1 @ r(W2).T + b2
    h2 = np.maximum(pre2, 0.0)
    z = h2 @ w3 + b3
    frac = 1.0 / (1.0 + np.exp(-z))
    return {"x": x, "pre1": pre1, "h1": h1, "pre2": pre2, "h2": h2, "z": z, "fraction": frac}


def init_params(seed=0):
    """He-normal weights, zero biases, b3 = 3 (fraction ~0.95 at the start: close to the face-centroid limiter)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, dtype=np.float32)
    p[P_W1:P_B1] = rng.normal(0, np.sqrt(2.0 / FEATURES), HIDDEN * FEATURES)
    p[P_W2:P_B2] = rng.normal(0, np.sqrt(2.0 / HIDDEN), HIDDEN * HIDDEN)
    p[P_W3:P_B3] = rng.normal(0, np.sqrt(1.0 / HIDDEN), HIDDEN)
    p[P_B3] = 3.0
    return p


def gradient(params, x_bits, target, weight, weight_sum, rounding=True):
    """(loss, gradient in the parameter order) of one batch, rounding as the device does (or not: rounding=False)."""
    r = bf16 if rounding else _same
    W1, b1, W2, b2, w3, b3 = unpack(params, rounding)
    f = forward(params, x_bits, rounding)
    t = np.asarray(target, dtype=np.float32).astype(np.float64)
    w = np.asarray(weight, dtype=np.float32).astype(np.float64)
    inv = float(np.float32(1.0 / weight_sum)) if rounding else 1.0 / weight_sum
    sig = f["fraction"]
    diff = sig - t
    loss = np.sum(w * diff * diff) * inv
    dz = 2.0 * w * diff * sig * (1.0 - sig) * inv
    dh2 = np.where(f["pre2"] > 0, dz[:, None] * w3[None, :], 0.0)
    dh2b = r(dh2)
    dh1 = np.where(f["pre1"] > 0, dh2b @ r(W2), 0.0)
    dh1b = r(dh1)
    g = np.zeros(PARAMS)
    g[P_W1:P_B1] = (dh1b.T @ f["x"]).ravel()
    g[P_B1:P_W2] = dh1.sum(axis=0)
    g[P_W2:P_B2] = (dh2b.T @ f["h1"]).ravel()
    g[P_B2:P_W3] = dh2.sum(axis=0)
    g[P_W3:P_B3] = dz @ f["h2"]
    g[P_B3] = dz.sum()
    return loss, g
