"""CPU oracle of the microphysics-surrogate MLP -- TEST INFRASTRUCTURE ONLY.

The surrogate has no reference counterpart (the reference has no learned part); this file
restates the network the GPU kernels implement (outbreaklab_amd/csrc/k_mlp.hip) in float64
numpy, so that the GPU forward pass and the GPU gradients can be checked against it.

Network: 16 features (14 inputs + 2 zero pads) -> 128 -> 128 -> 128 -> 16 outputs (9 used),
GELU (tanh form) between layers. Weights W1 [16][128], W2 [128][128], W3 [128][128],
W4 [128][16]; biases b1..b4. Loss: mean over rows and the 9 used outputs of (y - t)^2.

Features (x = raw inputs: T, p0, rho0, qv, qc, qr, qi, qs, qg, Nr, Ni, w, dz, dt):
    f0 = (T - 250) / 30,  f1 = p0 / 1e5 - 0.6,  f2 = rho0 - 0.7,
    f3..f8 = (ln(q + 1e-8) + 12) / 4 for qv, qc, qr, qi, qs, qg,
    f9 = (ln(Nr + 1) - 5) / 5,  f10 = (ln(Ni + 1) - 5) / 5,  f11 = w / 10,
    f12 = dz / 500 - 1,  f13 = dt / 5 - 1,  f14 = f15 = 0.
Targets: the 9 tendencies divided by per-output scales (the RMS of each in the training set).
"""
import numpy as np

IN, H, OUT, NOUT = 16, 128, 16, 9
GELU_C = np.sqrt(2.0 / np.pi)


def features(x):
    x = np.asarray(x, np.float64)
    f = np.zeros(x.shape[:-1] + (IN,))
    f[..., 0] = (x[..., 0] - 250.0) / 30.0
    f[..., 1] = x[..., 1] / 1e5 - 0.6
    f[..., 2] = x[..., 2] - 0.7
    for j in range(6):
        f[..., 3 + j] = (np.log(np.maximum(x[..., 3 + j], 0.0) + 1e-8) + 12.0) / 4.0
    f[..., 9] = (np.log(np.maximum(x[..., 9], 0.0) + 1.0) - 5.0) / 5.0
    f[..., 10] = (np.log(np.maximum(x[..., 10], 0.0) + 1.0) - 5.0) / 5.0
    f[..., 11] = x[..., 11] / 10.0
    f[..., 12] = x[..., 12] / 500.0 - 1.0
    f[..., 13] = x[..., 13] / 5.0 - 1.0
    return f


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    u = GELU_C * (z + 0.044715 * z ** 3)
    t = np.tanh(u)
    du = GELU_C * (1.0 + 3 * 0.044715 * z * z)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * du


def forward(p, f):
    z1 = f @ p["W1"] + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"] + p["b2"]
    h2 = gelu(z2)
    z3 = h2 @ p["W3"] + p["b3"]
    h3 = gelu(z3)
    y = h3 @ p["W4"] + p["b4"]
    return y, (f, z1, h1, z2, h2, z3, h3)


def loss_and_grads(p, f, t):
    """Mean squared error over rows and the NOUT used outputs, and its gradients."""
    y, (f, z1, h1, z2, h2, z3, h3) = forward(p, f)
    n = f.shape[0]
    d = np.zeros_like(y)
    d[:, :NOUT] = y[:, :NOUT] - t[:, :NOUT]
    loss = float(np.sum(d * d) / (n * NOUT))
    dy = 2.0 * d / (n * NOUT)
    g = {"W4": h3.T @ dy, "b4": dy.sum(0)}
    dz3 = (dy @ p["W4"].T) * gelu_grad(z3)
    g["W3"], g["b3"] = h2.T @ dz3, dz3.sum(0)
    dz2 = (dz3 @ p["W3"].T) * gelu_grad(z2)
    g["W2"], g["b2"] = h1.T @ dz2, dz2.sum(0)
    dz1 = (dz2 @ p["W2"].T) * gelu_grad(z1)
    g["W1"], g["b1"] = f.T @ dz1, dz1.sum(0)
    return loss, g


def bf16(a):
    """Round to the nearest BF16 value, ties to even (finite inputs), returned as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, f):
    """The quantised forward pass (training with quant on, and the BF16 inference kernels):
    features, weights and hidden activations rounded to BF16, biases and sums in high precision."""
    h = bf16(f)
    for k in ("1", "2", "3"):
        h = bf16(gelu(h @ bf16(p["W" + k]) + p["b" + k]))
    return h @ bf16(p["W4"]) + p["b4"]


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, (a, b) in {"1": (IN, H), "2": (H, H), "3": (H, H), "4": (H, OUT)}.items():
        p["W" + name] = rng.standard_normal((a, b)) * np.sqrt(2.0 / a)
        p["b" + name] = np.zeros(b)
    return p
