"""CPU reference of the warm-start MLP (8 -> 128 -> 128 -> 7, GELU), FP64
numpy -- TEST INFRASTRUCTURE ONLY: the checker of the HIP forward/backward
kernels, never shipped or called by the product.

The network has no counterpart in the reference (its warm start is LTE,
sim_lines.tpp:10-16); it follows the arithmetic the HIP kernels define:
  * weights and activations enter every matrix product rounded to bfloat16
    (round to nearest even), products accumulate in FP32 (here FP64 of the
    bf16-rounded operands, exact to FP32 rounding);
  * GELU is the tanh form 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)));
  * the loss is the mean over rows and the 7 outputs of the squared error of
    the normalised targets;
  * Adam (beta1 0.9, beta2 0.999, eps 1e-8) on FP32 master weights.
Shapes: W1 [8,128], W2 [128,128], W3 [128,7] as [in, out]; biases [out].
"""
import numpy as np

SQRT_2_OVER_PI = 0.7978845608028654
GELU_C = 0.044715


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    t = np.tanh(SQRT_2_OVER_PI * (z + GELU_C * z ** 3))
    return 0.5 * z * (1.0 + t)


def gelu_grad(z):
    u = SQRT_2_OVER_PI * (z + GELU_C * z ** 3)
    t = np.tanh(u)
    du = SQRT_2_OVER_PI * (1.0 + 3.0 * GELU_C * z * z)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * du


def forward(p, X):
    """p: dict W1 b1 W2 b2 W3 b3 (FP32 master values).  Returns Y and the cache."""
    Xb = bf16(X)
    Z1 = Xb @ bf16(p["W1"]) + p["b1"]
    H1 = bf16(gelu(Z1))
    Z2 = H1 @ bf16(p["W2"]) + p["b2"]
    H2 = bf16(gelu(Z2))
    Y = H2 @ bf16(p["W3"]) + p["b3"]
    return Y, dict(Xb=Xb, Z1=Z1, H1=H1, Z2=Z2, H2=H2)


def loss_and_grads(p, X, T):
    """Mean squared error over rows and outputs, and its gradients with the
    same bf16 rounding of the operands of every product as the kernels."""
    Y, c = forward(p, X)
    n = X.shape[0]
    D = Y - T
    loss = float(np.mean(D * D))
    dY = bf16(2.0 * D / (n * T.shape[1]))
    g = {}
    g["W3"] = c["H2"].T @ dY
    g["b3"] = dY.sum(axis=0)
    dZ2 = bf16((dY @ bf16(p["W3"]).T) * gelu_grad(c["Z2"]))
    g["W2"] = c["H1"].T @ dZ2
    g["b2"] = dZ2.sum(axis=0)
    dZ1 = bf16((dZ2 @ bf16(p["W2"]).T) * gelu_grad(c["Z1"]))
    g["W1"] = c["Xb"].T @ dZ1
    g["b1"] = dZ1.sum(axis=0)
    return loss, g


def loss_exact(p, X, T):
    """The loss without any bf16 rounding (FP64 throughout), for finite
    differences of the unrounded function."""
    Z1 = X @ p["W1"] + p["b1"]
    H1 = gelu(Z1)
    Z2 = H1 @ p["W2"] + p["b2"]
    H2 = gelu(Z2)
    Y = H2 @ p["W3"] + p["b3"]
    return float(np.mean((Y - T) ** 2))


def grads_exact(p, X, T):
    """Analytic gradients of loss_exact (FP64, no rounding)."""
    Z1 = X @ p["W1"] + p["b1"]
    H1 = gelu(Z1)
    Z2 = H1 @ p["W2"] + p["b2"]
    H2 = gelu(Z2)
    Y = H2 @ p["W3"] + p["b3"]
    dY = 2.0 * (Y - T) / Y.size
    g = {"W3": H2.T @ dY, "b3": dY.sum(0)}
    dZ2 = (dY @ p["W3"].T) * gelu_grad(Z2)
    g["W2"] = H1.T @ dZ2
    g["b2"] = dZ2.sum(0)
    dZ1 = (dZ2 @ p["W2"].T) * gelu_grad(Z1)
    g["W1"] = X.T @ dZ1
    g["b1"] = dZ1.sum(0)
    return g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    for k in p:
        m[k] = b1 * m[k] + (1 - b1) * g[k]
        v[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
        mh = m[k] / (1 - b1 ** t)
        vh = v[k] / (1 - b2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
