"""The learned photon-energy sampler (no counterpart in the reference).

An MLP (log10 chi, r) -> log delta, 2 -> 64 -> 64 -> 64 -> 1 with SiLU, BF16
weights on the matrix cores and FP32 accumulation, trained on the GPU by the
library's own forward / backward / Adam kernels on samples of the
optical-depth photon-emission inverse CDF (qed/optical_depth_tables.py:
59-101).  It replaces that table lookup in the mixed-precision step
(``Simulation(..., precision="mixed", sampler=...)``).  Bins of log10 chi
whose KS distance to the table exceeds a threshold fall back to the table.

Weights are saved as ``.npz`` (flat FP32 parameters, normalisation, fallback
mask) and loaded with ``numpy.load`` (no pickle).
"""
import ctypes as C
import os

import numpy as np

from .. import _lib as L

NPARAM = 64 * 2 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1
NBINS = 64
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables",
                            "photon_sampler_mlp.npz")


def init_params(seed=0):
    """He-style initial weights, zero biases (flat layout W1 b1 W2 b2 W3 b3 W4 b4)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = [rng.normal(0, 1 / np.sqrt(2), 128), np.zeros(64),
             rng.normal(0, 1 / np.sqrt(64), 4096), np.zeros(64),
             rng.normal(0, 1 / np.sqrt(64), 4096), np.zeros(64),
             rng.normal(0, 1 / np.sqrt(64), 64), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


class PhotonSampler:
    """Handle of a network on the GPU."""

    def __init__(self, seed=0, lc_range=(-3.0, 2.0)):
        h = C.c_void_p()
        L.call("sfp_mlp_create", C.byref(h))
        self._h = h.value
        self.fallback = 0
        self.set_params(init_params(seed), (lc_range[0], lc_range[1], 0.0, 1.0))

    def __del__(self):
        try:
            if getattr(self, "_h", None) and L._lib is not None:
                L._lib.sfp_mlp_destroy(self._h)
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------ params --
    def set_params(self, flat, norm):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == NPARAM
        nm = np.ascontiguousarray(norm, np.float32)
        L.call("sfp_mlp_set_params", self._h, flat.ctypes.data, nm.ctypes.data)

    def params(self):
        flat = np.empty(NPARAM, np.float32)
        nm = np.empty(4, np.float32)
        L.call("sfp_mlp_get_params", self._h, flat.ctypes.data, nm.ctypes.data)
        return flat, nm

    def set_fallback(self, mask: int):
        self.fallback = int(mask)
        L.call("sfp_mlp_set_fallback", self._h, C.c_uint64(self.fallback))

    def save(self, path=DEFAULT_PATH):
        flat, nm = self.params()
        np.savez(path, params=flat, norm=nm, fallback=np.array([self.fallback], np.uint64))

    @classmethod
    def load(cls, path=DEFAULT_PATH):
        d = np.load(path, allow_pickle=False)
        s = cls()
        s.set_params(d["params"], d["norm"])
        s.set_fallback(int(d["fallback"][0]))
        return s

    # ---------------------------------------------------------- training --
    def make_data(self, n, seed=0, tables=None):
        """n training samples on the device: x [n, 2] raw (log10 chi, r) and
        y [n] log delta from the table inverse CDF."""
        from . import device_tables
        t = tables or device_tables()
        x = L.DeviceArray(8 * n)
        y = L.DeviceArray(4 * n)
        L.call("sfp_mlp_make_data", self._h, t, int(n), C.c_uint64(seed), x.ptr, y.ptr)
        return x, y

    def fit(self, n_data=10_000_000, steps=4000, batch=4096, lr=3e-3, seed=0):
        """Train on the GPU; returns the per-step mean squared error."""
        x, y = self.make_data(n_data, seed)
        ys = np.empty(min(n_data, 1 << 20), np.float32)
        L.call("sfp_memcpy_dtoh", ys.ctypes.data, y.ptr, ys.nbytes)
        flat, nm = self.params()
        self.set_params(flat, (nm[0], nm[1], float(ys.mean()), float(ys.std())))
        curve = np.empty(steps, np.float32)
        L.call("sfp_mlp_train", self._h, x.ptr, y.ptr, int(n_data), int(steps), int(batch),
               float(lr), curve.ctypes.data)
        return curve

    def grad(self, x, y):
        """Gradient of the mean squared error over the samples (GPU kernels)."""
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        dx, dy = L.DeviceArray.of(x), L.DeviceArray.of(y)
        g = np.empty(NPARAM, np.float32)
        loss = C.c_float()
        L.call("sfp_mlp_grad", self._h, dx.ptr, dy.ptr, int(len(y)), g.ctypes.data,
               C.byref(loss))
        return g, float(loss.value)

    # --------------------------------------------------------- inference --
    def forward(self, x):
        """Normalised network output for raw inputs x [n, 2] (log10 chi, r)."""
        x = np.ascontiguousarray(x, np.float32)
        dx = L.DeviceArray.of(x)
        out = L.DeviceArray(4 * len(x))
        L.call("sfp_mlp_forward", self._h, dx.ptr, out.ptr, int(len(x)))
        o = np.empty(len(x), np.float32)
        out.download(o)
        return o

    def sample(self, lc, r, method="mlp", tables=None):
        """delta for (log10 chi, r) pairs by the network or by the table;
        returns (delta, kernel milliseconds)."""
        from . import device_tables
        x = np.ascontiguousarray(np.stack([lc, r], 1), np.float32)
        dx = L.DeviceArray.of(x)
        out = L.DeviceArray(8 * len(x))
        ms = C.c_float()
        L.call("sfp_mlp_sample", self._h, tables or device_tables(), dx.ptr, int(len(x)),
               0 if method == "mlp" else 1, out.ptr, C.byref(ms))
        d = np.empty(len(x), np.float64)
        out.download(d)
        return d, float(ms.value)

    def calibrate_fallback(self, threshold=None, per_bin=20000, seed=7):
        """Per-bin two-sample KS distance between network and table samples
        of delta; bins above `threshold` (default: the p = 0.01 critical
        value for the sample size) use the table.  Returns the distances."""
        from scipy import stats
        flat, nm = self.params()
        rng = np.random.Generator(np.random.PCG64(seed))
        edges = np.linspace(nm[0], nm[1], NBINS + 1)
        lc = np.concatenate([rng.uniform(edges[i], edges[i + 1], per_bin) for i in range(NBINS)])
        d_net, _ = self.sample(lc, rng.random(lc.size), "mlp")
        d_tab, _ = self.sample(lc, rng.random(lc.size), "table")
        if threshold is None:
            threshold = 1.63 * np.sqrt(2.0 / per_bin)
        ks = np.array([stats.ks_2samp(d_net[i * per_bin:(i + 1) * per_bin],
                                      d_tab[i * per_bin:(i + 1) * per_bin]).statistic
                       for i in range(NBINS)])
        mask = 0
        for i in range(NBINS):
            if ks[i] > threshold:
                mask |= 1 << i
        self.set_fallback(mask)
        return ks, threshold


_default = None


def default_sampler():
    """The trained network shipped with the package (trained on the GPU by
    ``python -m sfparticles_amd.qed.mlp``)."""
    global _default
    if _default is None:
        if not os.path.exists(DEFAULT_PATH):
            raise FileNotFoundError(
                f"{DEFAULT_PATH} is missing: train it with python -m sfparticles_amd.qed.mlp")
        _default = PhotonSampler.load(DEFAULT_PATH)
    return _default


def main():
    """Train the default network on the GPU and save it."""
    import argparse
    import time
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--samples", type=int, default=10_000_000)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_PATH)
    a = ap.parse_args()
    s = PhotonSampler(seed=a.seed)
    t0 = time.time()
    curve = s.fit(a.samples, a.steps, a.batch, a.lr, a.seed)
    t1 = time.time()
    ks, thr = s.calibrate_fallback(per_bin=50000)
    print(f"trained {a.steps} steps on {a.samples} samples in {t1 - t0:.1f} s; loss "
          f"{curve[:10].mean():.4g} -> {curve[-100:].mean():.4g}; KS per bin max {ks.max():.4f} "
          f"median {np.median(ks):.4f} (threshold {thr:.4f}); fallback bins "
          f"{bin(s.fallback).count('1')}", flush=True)
    s.save(a.out)


if __name__ == "__main__":
    main()
