"""Ensembles of independent laser shots in one simulation (no counterpart in
the reference, which runs one field per ``Simulation``).

Every particle carries a ``shot`` index; the kernels evaluate the field of
that shot, and particles a shot creates inherit its index, so hundreds of
shots of uneven particle counts advance in the same kernel launches.  QED
draws stay keyed by particle id, so a shot's result is the same as running it
alone with the same ids.

``shot_histogram`` reduces per-shot counts and energy spectra on the device.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L
from .fields import Fields

C_LIGHT = 299792458.0
LAMBDA = 0.8e-6
DT = LAMBDA / C_LIGHT / 100  # T/100
CTAU_30FS = C_LIGHT * 30e-15 / math.sqrt(2 * math.log(2))


def head_on_shots(nshots, nsteps, seed=0, a0=(20.0, 500.0), count=(1e3, 1e6),
                  gamma=(500.0, 5000.0), spread=0.05, w0=2e-6):
    """A head-on ensemble: shot k has a0_k ~ U(a0), N_k = round(10^U(log10
    count)) electrons with gamma ~ N(gamma_k, spread * gamma_k), gamma_k ~
    U(gamma), moving along -x, x ~ N(d, 3 um), y, z ~ N(0, 1 um), against a
    linearly polarised paraxial Gaussian pulse (lambda 0.8 um, w0 2 um,
    tau_FWHM 30 fs) starting at x0 = -d, d = c dt nsteps / 2.

    Draw order (numpy PCG64(seed)): a0, log10 N, gamma_k for all shots; then
    per shot gamma, x, y, z.  Returns (fields per shot, props, shot index,
    per-shot table)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a0s = rng.uniform(a0[0], a0[1], nshots)
    ns = np.rint(10.0 ** rng.uniform(math.log10(count[0]), math.log10(count[1]), nshots))
    ns = ns.astype(np.int64)
    gs = rng.uniform(gamma[0], gamma[1], nshots)
    d = C_LIGHT * DT * nsteps / 2
    parts = []
    for k in range(nshots):
        n = int(ns[k])
        g = rng.normal(gs[k], spread * gs[k], n)
        x = rng.normal(d, 3e-6, n)
        y = rng.normal(0.0, 1e-6, n)
        z = rng.normal(0.0, 1e-6, n)
        parts.append((x, y, z, -np.sqrt(g ** 2 - 1)))
    cat = [np.concatenate([p[j] for p in parts]) if parts else np.zeros(0) for j in range(4)]
    ntot = cat[0].size
    props = (cat[0], cat[1], cat[2], cat[3], np.zeros(ntot), np.zeros(ntot))
    shot = np.repeat(np.arange(nshots, dtype=np.uint32), ns)
    fields = [Fields(terms=[(2, (float(a0s[k]), w0, CTAU_30FS, 1, -d, LAMBDA, 0.0, 0.0))])
              for k in range(nshots)]
    table = dict(a0=a0s, n=ns, gamma=gs)
    return fields, props, shot, table


def shot_histogram(particles, nshots, nbins=0, log10_range=(0.0, 4.0)):
    """Per-shot counts (and energy spectra) of a species' live particles,
    reduced on the device: (counts [nshots], hist [nshots, nbins])."""
    particles._ensure_device()
    out = np.zeros((nshots, nbins + 1), np.int64)
    L.call("sfp_store_histogram", particles._store, int(particles.m == 0), int(nshots),
           int(nbins), float(log10_range[0]), float(log10_range[1]), out.ctypes.data)
    return out[:, 0].copy(), out[:, 1:].copy()
