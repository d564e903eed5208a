"""CPU restatement of the chemistry surrogate's network -- TEST INFRASTRUCTURE.

Only tests/ (and the benchmark's CPU baseline) use this module, as the checker
of the HIP network in eu3d-gpu_amd/csrc/eu3d_mlp.hip; nothing in the product
imports it.  The reference has no surrogate (its DNN mode, src/Euler.cpp:
1473-1480, runs no chemistry), so this restatement is pinned by mathematics:
tests/test_oracle_pins.py checks its gradients against central differences
of its own float64 loss, and the BF16 emulation against exact rounding cases.

Network: x (n_in raw features) -> normalize -> pad to 32 -> [W0,b0] GELU ->
[W1,b1] GELU -> [W2,b2] GELU -> [W3,b3] -> first n_out outputs (normalized
units).  Loss: mean over the batch and the n_out real outputs of the squared
error against normalized targets.  Parameters are a flat float32 vector in
the layout of eu3d_mlp_get_params: W0 [32 x H], b0 [H], W1 [H x H], b1,
W2 [H x H], b2, W3 [H x 32], b3 [32].

emulate=True applies the device's rounding points: BF16 (round to nearest
even) of the weights, of the normalized input, of each GELU output and of
every back-propagated dZ; float64 elsewhere (the device accumulates in
float32, so the two agree to about 1e-3 relative, not bitwise).
"""
import numpy as np
from scipy.special import erf

K_IN = 32
K_OUT = 32


def bf16(a):
    """Round to BF16 (nearest even) and return float64 values."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    finite = (u & 0x7F800000) != 0x7F800000
    quiet = np.where((u & 0xFFFF) != 0, np.uint64(0x40), np.uint64(0))
    r = np.where(finite, (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16),
                 (u >> np.uint64(16)) | quiet)
    out = (r.astype(np.uint32) << 16).view(np.float32)
    return out.astype(np.float64).reshape(np.shape(a))


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def shapes(H):
    return [(K_IN, H), (H, H), (H, H), (H, K_OUT)]


def param_count(H):
    return sum(k * n + n for k, n in shapes(H))


def unpack(params, H):
    p = np.asarray(params, dtype=np.float64)
    out, o = [], 0
    for k, n in shapes(H):
        W = p[o:o + k * n].reshape(k, n)
        o += k * n
        b = p[o:o + n]
        o += n
        out.append((W, b))
    assert o == p.size
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def _norm_inputs(x, norm, n_in):
    in_mean, in_std = norm[0], norm[1]
    X = np.zeros((x.shape[0], K_IN))
    X[:, :n_in] = (np.asarray(x, np.float64) - in_mean[:n_in]) / in_std[:n_in]
    return X


def forward(params, H, x, norm, n_in, emulate=True):
    """Normalized outputs [B x 32] and the cache for backward.
    norm = (in_mean, in_std, out_mean, out_std) arrays."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, H)
    X = r(_norm_inputs(x, norm, n_in))
    A, cache = X, []
    for l, (W, b) in enumerate(layers):
        Z = A @ r(W) + b
        cache.append((A, Z))
        if l < len(layers) - 1:
            A = r(gelu(Z))
        else:
            A = Z
    return A, cache


def predict(params, H, x, norm, n_in, n_out, emulate=True):
    """De-normalized outputs [B x n_out] (eu3d_mlp_predict)."""
    Y, _ = forward(params, H, x, norm, n_in, emulate)
    return Y[:, :n_out] * norm[3][:n_out] + norm[2][:n_out]


def loss_and_grads(params, H, x, t, norm, n_in, n_out, emulate=True):
    """(loss, flat gradient) of one batch (eu3d_mlp_gradients)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    layers = unpack(params, H)
    Y, cache = forward(params, H, x, norm, n_in, emulate)
    B = Y.shape[0]
    T0 = np.zeros_like(Y)
    T0[:, :n_out] = (np.asarray(t, np.float64) - norm[2][:n_out]) / norm[3][:n_out]
    d = np.zeros_like(Y)
    d[:, :n_out] = Y[:, :n_out] - T0[:, :n_out]
    loss = float(np.sum(d * d) / (B * n_out))
    dZ = r(2.0 * d / (B * n_out))
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        A, _ = cache[l]
        W, _ = layers[l]
        grads[l] = (A.T @ dZ, dZ.sum(axis=0))
        if l > 0:
            _, Zprev = cache[l - 1]
            dZ = r((dZ @ r(W).T) * gelu_grad(Zprev))
    return loss, pack(grads)


def init_params(H, n_in, n_out, seed):
    """Glorot-uniform weights on the real block (padded rows/columns 0), zero
    biases -- the distribution of eu3d_mlp_create, not its exact values."""
    rng = np.random.default_rng(seed)
    layers = []
    for l, (k, n) in enumerate(shapes(H)):
        kin = n_in if l == 0 else k
        nout = n_out if l == 3 else n
        a = np.sqrt(6.0 / (kin + nout))
        W = np.zeros((k, n))
        W[:kin, :nout] = rng.uniform(-a, a, size=(kin, nout))
        layers.append((W, np.zeros(n)))
    return pack(layers).astype(np.float32)
