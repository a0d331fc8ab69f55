"""The learned photon-energy sampler.

CPU: the numpy oracle of the network (oracle/mlp.py) is pinned -- its
gradient against central finite differences of its own float64 loss, its BF16
rounding against exact values.
GPU: the library's forward pass (which also fixes the matrix-core operand
layout) and its backward pass against the oracle with the same BF16
rounding; training on the GPU lowers the loss and the trained network draws
delta distributed like the table's (two-sample KS per bin of log10 chi, the
table taking over in bins that fail); and what the network gains over the
table it replaces.
"""
import numpy as np
import pytest

import oracle.mlp as OM


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.Generator(np.random.PCG64(3))
    flat = OM.init_params(1) + rng.normal(0, 0.05, OM.NPARAM)
    X = rng.uniform(-1, 1, (64, 2))
    t = rng.normal(0, 1, 64)
    loss, g = OM.loss_grad(flat, X, t, emulate=False)
    idx = np.concatenate([rng.choice(OM.NPARAM, 60, replace=False), [OM.NPARAM - 1, 0, 200]])
    h = 1e-6
    for i in idx:
        p = flat.copy()
        p[i] += h
        lp, _ = OM.loss_grad(p, X, t, emulate=False)
        p[i] -= 2 * h
        lm, _ = OM.loss_grad(p, X, t, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_oracle_bf16_rounding():
    """BF16 keeps 7 fraction bits: spacing 2^-7 at 1; ties go to the even
    neighbour.  Also against torch's bfloat16 conversion on random values."""
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, 1.0 + 3 * 2 ** -8, -2.5,
                  3.0e-3, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0030059814453125, 0.0])
    got = OM.bf16(v)
    assert np.array_equal(got, want), got
    import torch
    rng = np.random.Generator(np.random.PCG64(1))
    r = (rng.normal(0, 1, 100000) * 10.0 ** rng.uniform(-20, 20, 100000)).astype(np.float32)
    ref = torch.from_numpy(r).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(OM.bf16(r), ref)


@pytest.mark.gpu
def test_forward_matches_oracle():
    from sfparticles_amd.qed.mlp import PhotonSampler
    rng = np.random.Generator(np.random.PCG64(5))
    flat = (OM.init_params(2) + rng.normal(0, 0.05, OM.NPARAM)).astype(np.float32)
    norm = (-3.0, 2.0, -5.0, 3.0)
    s = PhotonSampler()
    s.set_params(flat, norm)
    x = np.stack([rng.uniform(-3, 2, 5000), rng.random(5000)], 1).astype(np.float32)
    got = s.forward(x)
    X = OM.normalise(x, norm)
    want, _ = OM.forward(flat.astype(np.float64), X, emulate=True)
    exact, _ = OM.forward(flat.astype(np.float64), X, emulate=False)
    err = np.max(np.abs(got - want))
    print("forward: max |gpu - oracle(bf16)|", err, " max |gpu - oracle(fp64)|",
          np.max(np.abs(got - exact)))
    assert err < 2e-3
    assert np.max(np.abs(got - exact)) < 0.1


@pytest.mark.gpu
def test_gradient_matches_oracle():
    from sfparticles_amd.qed.mlp import PhotonSampler
    rng = np.random.Generator(np.random.PCG64(6))
    flat = (OM.init_params(3) + rng.normal(0, 0.05, OM.NPARAM)).astype(np.float32)
    norm = (-3.0, 2.0, -5.0, 3.0)
    s = PhotonSampler()
    s.set_params(flat, norm)
    n = 1024
    x = np.stack([rng.uniform(-3, 2, n), rng.random(n)], 1).astype(np.float32)
    y = rng.normal(-5, 3, n).astype(np.float32)
    g, loss = s.grad(x, y)
    X = OM.normalise(x, norm)
    t = (y.astype(np.float64) - norm[2]) / norm[3]
    l_ref, g_ref = OM.loss_grad(flat.astype(np.float64), X, t, emulate=True)
    l_ex, g_ex = OM.loss_grad(flat.astype(np.float64), X, t, emulate=False)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    cos = float(g @ g_ex / (np.linalg.norm(g) * np.linalg.norm(g_ex)))
    print("gradient: |gpu - oracle(bf16)| / |oracle|", rel, " cosine to fp64", cos,
          " loss", loss, l_ref)
    assert abs(loss - l_ref) <= 1e-3 * l_ref
    assert rel < 1e-2
    assert cos > 0.999
    # every parameter block is checked, not only the largest
    for k, (a, b, _) in OM.SLICES.items():
        rb = np.linalg.norm(g[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-30)
        assert rb < 3e-2, (k, rb)


@pytest.fixture(scope="module")
def trained():
    from sfparticles_amd.qed.mlp import PhotonSampler
    s = PhotonSampler(seed=0)
    curve = s.fit(n_data=2_000_000, steps=3000, batch=4096, lr=3e-3, seed=0)
    ks, thr = s.calibrate_fallback(per_bin=20000)
    return s, curve, ks, thr


@pytest.mark.gpu
def test_training_lowers_the_loss(trained):
    s, curve, ks, thr = trained
    print("loss: first", curve[:5].mean(), "last", curve[-50:].mean())
    assert curve[-50:].mean() < 0.02 * curve[:5].mean()
    assert curve[-50:].mean() < 2e-3


@pytest.mark.gpu
def test_trained_network_matches_table_spectrum(trained):
    """Per bin of log10 chi, the KS distance between network and table
    samples of delta; bins above the p = 0.01 critical value fall back to the
    table, and the sampler as used (network or fallback) matches everywhere."""
    from scipy import stats
    s, curve, ks, thr = trained
    nfb = bin(s.fallback).count("1")
    print("KS per bin: max", ks.max(), "median", np.median(ks), "threshold", thr,
          "fallback bins", nfb)
    assert nfb <= 16  # the network serves most of the range
    # whole range, sampler as used: fresh draws
    rng = np.random.Generator(np.random.PCG64(11))
    lc = rng.uniform(-3, 2, 400000)
    d_net, _ = s.sample(lc, rng.random(lc.size), "mlp")
    d_tab, _ = s.sample(lc, rng.random(lc.size), "table")
    r = stats.ks_2samp(np.log(d_net), np.log(d_tab))
    print("all chi: KS distance", r.statistic, "p", r.pvalue)
    # over the whole range the largest gap between the two CDFs of delta
    # stays under 0.5 % (measured 0.3 %, KS p 0.06 with 4e5 draws each; the
    # p = 0.01 critical distance at this size is 0.36 %)
    assert r.statistic < 5e-3


@pytest.mark.gpu
def test_network_gain_over_table(trained):
    """What the network gains over the table it replaces: the bytes it needs
    (17 KB of weights that live in LDS against the 2 MB table the bisection
    walks in HBM/L2), at batched sampling throughput on the matrix cores on
    par with the table's (measured 9.6 against 9.5 G samples/s; asserted
    within 25 %, as both are timed once)."""
    s, curve, ks, thr = trained
    from sfparticles_amd.qed.tables import load_od_tables
    table_bytes = load_od_tables().photon_cdf.nbytes
    net_bytes = 2 * 2 * 4096 + 4 * (128 + 64 * 4 + 1)
    rng = np.random.Generator(np.random.PCG64(12))
    n = 4_000_000
    lc = rng.uniform(-3, 2, n)
    r = rng.random(n)
    s.sample(lc[:4096], r[:4096], "mlp")
    s.sample(lc[:4096], r[:4096], "table")
    _, ms_net = s.sample(lc, r, "mlp")
    _, ms_tab = s.sample(lc, r, "table")
    print(f"bytes: network {net_bytes}, table {table_bytes} ({table_bytes / net_bytes:.0f}x); "
          f"{n} samples: network {ms_net:.3f} ms ({n / ms_net / 1e6:.2f} G/s), "
          f"table {ms_tab:.3f} ms ({n / ms_tab / 1e6:.2f} G/s)")
    assert net_bytes * 100 < table_bytes
    assert ms_net < 1.25 * ms_tab
