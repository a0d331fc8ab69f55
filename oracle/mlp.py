"""CPU oracle of the learned QED models (test infrastructure).

The networks have no counterpart in the reference: they are defined by the
mathematics they implement, NIN -> 64 -> 64 -> 64 -> 1 with SiLU (NIN = 2:
the photon-energy sampler (log10 chi, r) -> log delta; NIN = 3: the
emission rate (log10 chi, r1, sigma) -> log f), mean squared error on the
normalised output.  This module restates its forward pass and its gradient in numpy
(float64), optionally rounding the operands of the two 64 x 64 products to
BF16 exactly where the GPU feeds them to the matrix cores, so that the GPU's
forward and backward kernels can be compared with it.  The gradient itself is
pinned by central finite differences of the float64 loss (tests).

Parameter layout (flat, float32 on the device): W1[64][NIN] b1[64] W2[64][64]
b2[64] W3[64][64] b3[64] W4[64] b4.
"""
import numpy as np

def _slices(nin):
    out, o = {}, 0
    for name, shape in (("W1", (64, nin)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)),
                        ("W3", (64, 64)), ("b3", (64,)), ("W4", (64,)), ("b4", (1,))):
        n = int(np.prod(shape))
        out[name] = (o, o + n, shape)
        o += n
    return out, o


SLICES, NPARAM = _slices(2)
SLICES3, NPARAM3 = _slices(3)
assert NPARAM == 8577 and NPARAM3 == 8641


def _sl(n):
    return SLICES if n == NPARAM else SLICES3


def unpack(flat):
    flat = np.asarray(flat, np.float64)
    return {k: flat[a:b].reshape(s) for k, (a, b, s) in _sl(flat.size).items()}


def pack(P):
    nin = np.asarray(P["W1"]).shape[1]
    sl = SLICES if nin == 2 else SLICES3
    out = np.zeros(NPARAM if nin == 2 else NPARAM3)
    for k, (a, b, s) in sl.items():
        out[a:b] = np.asarray(P[k], np.float64).reshape(-1)
    return out


def bf16(a):
    """Round to the nearest BF16 value, ties to even (finite inputs)."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def normalise(x, norm):
    """raw (log10 chi, r[, sigma]) -> network inputs; norm = (lc_min, lc_max,
    mu, sd); sigma passes unscaled"""
    lc_min, lc_max = norm[0], norm[1]
    x = np.asarray(x, np.float64)
    x0 = (x[:, 0] - 0.5 * (lc_min + lc_max)) * (2.0 / (lc_max - lc_min))
    cols = [x0, 2.0 * x[:, 1] - 1.0] + ([x[:, 2]] if x.shape[1] == 3 else [])
    return np.stack(cols, 1)


def forward(flat, X, emulate=True):
    """Network output and cache for normalised inputs X [n, 2]."""
    P = unpack(flat)
    q = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    z1 = X @ P["W1"].T + P["b1"]
    z2 = q(silu(z1)) @ q(P["W2"]).T + P["b2"]
    z3 = q(silu(z2)) @ q(P["W3"]).T + P["b3"]
    o = silu(z3) @ P["W4"] + P["b4"][0]
    return o, (X, z1, z2, z3)


def loss_grad(flat, X, t, emulate=True):
    """Mean squared error against normalised targets t, and its gradient."""
    P = unpack(flat)
    q = bf16 if emulate else (lambda v: np.asarray(v, np.float64))
    o, (X, z1, z2, z3) = forward(flat, X, emulate)
    n = X.shape[0]
    e = o - t
    loss = float(np.mean(e * e))
    dO = 2.0 * e / n
    G = {}
    h3 = silu(z3)
    G["W4"] = dO @ h3
    G["b4"] = np.array([dO.sum()])
    dz3 = np.outer(dO, P["W4"]) * dsilu(z3)
    G["W3"] = q(dz3).T @ q(silu(z2))
    G["b3"] = dz3.sum(0)
    dz2 = (q(dz3) @ q(P["W3"])) * dsilu(z2)
    G["W2"] = q(dz2).T @ q(silu(z1))
    G["b2"] = dz2.sum(0)
    dz1 = (q(dz2) @ q(P["W2"])) * dsilu(z1)
    G["W1"] = dz1.T @ X
    G["b1"] = dz1.sum(0)
    return loss, pack(G)


def init_params(seed=0, nin=2):
    """He-style initial weights, zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    P = {"W1": rng.normal(0, 1 / np.sqrt(nin), (64, nin)), "b1": np.zeros(64),
         "W2": rng.normal(0, 1 / np.sqrt(64), (64, 64)), "b2": np.zeros(64),
         "W3": rng.normal(0, 1 / np.sqrt(64), (64, 64)), "b3": np.zeros(64),
         "W4": rng.normal(0, 1 / np.sqrt(64), 64), "b4": np.zeros(1)}
    return pack(P)
