"""The learned QED models (no counterpart in the reference).

PhotonSampler: an MLP (log10 chi, r) -> log delta, 2 -> 64 -> 64 -> 64 -> 1
with SiLU, BF16 weights on the matrix cores and FP32 accumulation, trained on
the GPU by the library's own forward / backward / Adam kernels on samples of
the optical-depth photon-emission inverse CDF (qed/optical_depth_tables.py:
59-101).  It replaces that table lookup in the mixed-precision step
(``Simulation(..., precision="mixed", sampler=...)``).  Bins of log10 chi
whose KS distance to the table exceeds a threshold fall back to the table.

EmissionRateNet: the same network with three inputs, (log10 chi, r1, sigma =
s.zeta) -> log(f + 1e-6), f = 3 r1^2 w(chi, r1^3, sigma) / (alpha m c^2 / hbar) the
acceptance density of rejection sampling (qed/rejection_sampling.py:68-85)
with the spin-resolved rate, trained on that rate's table; it replaces the
table lookup for spin-carrying leptons (``Simulation(..., precision="mixed",
spin_emission=True, rate_net=...)``).  Bins of log10 chi whose integrated
emission probability differs from the table's by more than a threshold fall
back to the table.

Weights are saved as ``.npz`` (flat FP32 parameters, normalisation, fallback
mask) and loaded with ``numpy.load`` (no pickle).
"""
import ctypes as C
import os

import numpy as np

from .. import _lib as L

NPARAM = 64 * 2 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1
NPARAM_RATE = NPARAM + 64
NBINS = 64
_TABLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tables")
DEFAULT_PATH = os.path.join(_TABLES, "photon_sampler_mlp.npz")
DEFAULT_RATE_PATH = os.path.join(_TABLES, "emission_rate_mlp.npz")


def init_params(seed=0, nin=2):
    """He-style initial weights, zero biases (flat layout W1 b1 W2 b2 W3 b3 W4 b4)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = [rng.normal(0, 1 / np.sqrt(nin), 64 * nin), np.zeros(64),
             rng.normal(0, 1 / np.sqrt(64), 4096), np.zeros(64),
             rng.normal(0, 1 / np.sqrt(64), 4096), np.zeros(64),
             rng.normal(0, 1 / np.sqrt(64), 64), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


class _Network:
    """Handle of a network on the GPU (NIN inputs)."""
    NIN = 2
    _create = "sfp_mlp_create"

    def __init__(self, seed=0, lc_range=(-3.0, 2.0)):
        h = C.c_void_p()
        L.call(self._create, C.byref(h))
        self._h = h.value
        self.fallback = 0
        self.nparam = NPARAM if self.NIN == 2 else NPARAM_RATE
        self.set_params(init_params(seed, self.NIN), (lc_range[0], lc_range[1], 0.0, 1.0))

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
        assert flat.size == self.nparam
        nm = np.ascontiguousarray(norm, np.float32)
        L.call("sfp_mlp_set_params", self._h, flat.ctypes.data, nm.ctypes.data)

    def params(self):
        flat = np.empty(self.nparam, np.float32)
        nm = np.empty(4, np.float32)
        L.call("sfp_mlp_get_params", self._h, flat.ctypes.data, nm.ctypes.data)
        return flat, nm

    def set_fallback(self, mask: int):
        self.fallback = int(mask)
        L.call("sfp_mlp_set_fallback", self._h, C.c_uint64(self.fallback))

    def save(self, path=None):
        flat, nm = self.params()
        np.savez(path or self.DEFAULT, params=flat, norm=nm,
                 fallback=np.array([self.fallback], np.uint64))

    @classmethod
    def load(cls, path=None):
        d = np.load(path or cls.DEFAULT, allow_pickle=False)
        s = cls()
        s.set_params(d["params"], d["norm"])
        s.set_fallback(int(d["fallback"][0]))
        return s

    # ---------------------------------------------------------- training --
    def make_data(self, n, seed=0, tables=None):
        """n training samples on the device: x [n, NIN] raw inputs and y [n]
        targets from the table."""
        from . import device_tables
        t = tables or device_tables()
        x = L.DeviceArray(4 * self.NIN * n)
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
        g = np.empty(self.nparam, np.float32)
        loss = C.c_float()
        L.call("sfp_mlp_grad", self._h, dx.ptr, dy.ptr, int(len(y)), g.ctypes.data,
               C.byref(loss))
        return g, float(loss.value)

    # --------------------------------------------------------- inference --
    def forward(self, x):
        """Normalised network output for raw inputs x [n, NIN]."""
        x = np.ascontiguousarray(x, np.float32)
        dx = L.DeviceArray.of(x)
        out = L.DeviceArray(4 * len(x))
        L.call("sfp_mlp_forward", self._h, dx.ptr, out.ptr, int(len(x)))
        o = np.empty(len(x), np.float32)
        out.download(o)
        return o


class PhotonSampler(_Network):
    """The photon-energy sampler (log10 chi, r) -> log delta."""
    NIN = 2
    _create = "sfp_mlp_create"
    DEFAULT = DEFAULT_PATH

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


class EmissionRateNet(_Network):
    """The spin-resolved emission rate (log10 chi, r1, sigma) -> log(f + 1e-6)."""
    NIN = 3
    _create = "sfp_mlp_create_rate"
    DEFAULT = DEFAULT_RATE_PATH

    def rate(self, lc, r1, sigma, method="mlp", tables=None):
        """f = 3 r1^2 w(chi, r1^3, sigma) / (alpha m c^2 / hbar) by the network
        or the table; returns (f, kernel milliseconds)."""
        from . import device_tables
        x = np.ascontiguousarray(np.stack([lc, r1, sigma], 1), np.float32)
        dx = L.DeviceArray.of(x)
        out = L.DeviceArray(8 * len(x))
        ms = C.c_float()
        L.call("sfp_mlp_rate", self._h, tables or device_tables(), dx.ptr, int(len(x)),
               0 if method == "mlp" else 1, out.ptr, C.byref(ms))
        f = np.empty(len(x), np.float64)
        out.download(f)
        return f, float(ms.value)

    def bin_errors(self, per_bin=8192, seed=7):
        """Per bin of log10 chi: relative error of the network's emission
        probability (the mean of f over r1 ~ U(0,1), at sigma = -1, 0, 1 on the
        same points) against the table's; the largest of the three."""
        flat, nm = self.params()
        rng = np.random.Generator(np.random.PCG64(seed))
        edges = np.linspace(nm[0], nm[1], NBINS + 1)
        lc = np.concatenate([rng.uniform(edges[i], edges[i + 1], per_bin) for i in range(NBINS)])
        r1 = rng.random(lc.size)
        err = np.zeros(NBINS)
        for sig in (-1.0, 0.0, 1.0):
            s = np.full(lc.size, sig)
            fn, _ = self.rate(lc, r1, s, "mlp")
            ft, _ = self.rate(lc, r1, s, "table")
            mn = fn.reshape(NBINS, per_bin).mean(1)
            mt = ft.reshape(NBINS, per_bin).mean(1)
            err = np.maximum(err, np.abs(mn - mt) / np.maximum(mt, 1e-300))
        return err

    def calibrate_fallback(self, threshold=0.01, per_bin=8192, seed=7):
        """Bins whose emission probability errs by more than `threshold`
        (relative) use the table.  Returns (errors, threshold)."""
        err = self.bin_errors(per_bin, seed)
        mask = 0
        for i in range(NBINS):
            if err[i] > threshold:
                mask |= 1 << i
        self.set_fallback(mask)
        return err, threshold


_default = None
_default_rate = None


def default_rate_net():
    """The trained emission-rate network shipped with the package (trained on
    the GPU by ``python -m sfparticles_amd.qed.mlp --rate``)."""
    global _default_rate
    if _default_rate is None:
        if not os.path.exists(DEFAULT_RATE_PATH):
            raise FileNotFoundError(
                f"{DEFAULT_RATE_PATH} is missing: train it with python -m sfparticles_amd.qed.mlp --rate")
        _default_rate = EmissionRateNet.load(DEFAULT_RATE_PATH)
    return _default_rate


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
    """Train a default network on the GPU and save it."""
    import argparse
    import time
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--rate", action="store_true", help="the emission-rate network")
    ap.add_argument("--samples", type=int, default=10_000_000)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.rate:
        s = EmissionRateNet(seed=a.seed)
        t0 = time.time()
        curve = s.fit(a.samples, a.steps, a.batch, a.lr, a.seed)
        t1 = time.time()
        err, thr = s.calibrate_fallback()
        print(f"trained {a.steps} steps on {a.samples} samples in {t1 - t0:.1f} s; loss "
              f"{curve[:10].mean():.4g} -> {curve[-100:].mean():.4g}; emission probability "
              f"error per bin max {err.max():.4f} median {np.median(err):.4f} (threshold "
              f"{thr}); fallback bins {bin(s.fallback).count('1')}", flush=True)
        s.save(a.out or DEFAULT_RATE_PATH)
        return
    s = PhotonSampler(seed=a.seed)
    t0 = time.time()
    curve = s.fit(a.samples, a.steps, a.batch, a.lr, a.seed)
    t1 = time.time()
    ks, thr = s.calibrate_fallback(per_bin=50000)
    print(f"trained {a.steps} steps on {a.samples} samples in {t1 - t0:.1f} s; loss "
          f"{curve[:10].mean():.4g} -> {curve[-100:].mean():.4g}; KS per bin max {ks.max():.4f} "
          f"median {np.median(ks):.4f} (threshold {thr:.4f}); fallback bins "
          f"{bin(s.fallback).count('1')}", flush=True)
    s.save(a.out or DEFAULT_PATH)


if __name__ == "__main__":
    main()
