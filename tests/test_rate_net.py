"""The learned spin-resolved emission rate (no counterpart in the reference).

CPU: the numpy oracle's gradient for the three-input network against central
finite differences of its float64 loss.
GPU: the library's forward and backward passes of the three-input network
against the oracle with the same BF16 rounding; training on the GPU lowers
the loss; per bin of log10 chi, the emission probability the trained network
gives (the mean acceptance density over r1, at sigma = -1, 0, 1) against the
table's, the table taking over in bins that err by more than 1 %; in a static
magnetic field, the photon yields of electrons with spin along +zeta and
-zeta with the network in the step against the expectation from the table's
rate (3 sigma); and what the network costs next to the table it replaces.
"""
import numpy as np
import pytest

import oracle.mlp as OM


def test_oracle_gradient_three_inputs():
    rng = np.random.Generator(np.random.PCG64(8))
    flat = OM.init_params(2, nin=3) + rng.normal(0, 0.05, OM.NPARAM3)
    X = rng.uniform(-1, 1, (48, 3))
    t = rng.normal(0, 1, 48)
    loss, g = OM.loss_grad(flat, X, t, emulate=False)
    h = 1e-6
    a, b, _ = OM.SLICES3["W1"]
    for i in list(range(a, b, 7)) + list(rng.choice(OM.NPARAM3, 30, replace=False)):
        p = flat.copy()
        p[i] += h
        lp, _ = OM.loss_grad(p, X, t, emulate=False)
        p[i] -= 2 * h
        lm, _ = OM.loss_grad(p, X, t, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


@pytest.mark.gpu
def test_rate_net_forward_and_gradient_match_oracle():
    from sfparticles_amd.qed.mlp import EmissionRateNet
    rng = np.random.Generator(np.random.PCG64(9))
    flat = (OM.init_params(4, nin=3) + rng.normal(0, 0.05, OM.NPARAM3)).astype(np.float32)
    norm = (-3.0, 2.0, -8.0, 6.0)
    s = EmissionRateNet()
    s.set_params(flat, norm)
    n = 2048
    x = np.stack([rng.uniform(-3, 2, n), rng.random(n), rng.uniform(-1, 1, n)], 1).astype(np.float32)
    X = OM.normalise(x, norm)
    got = s.forward(x)
    want, _ = OM.forward(flat.astype(np.float64), X, emulate=True)
    print("forward: max |gpu - oracle(bf16)|", np.max(np.abs(got - want)))
    assert np.max(np.abs(got - want)) < 2e-3
    y = rng.normal(-8, 6, n).astype(np.float32)
    g, loss = s.grad(x, y)
    t = (y.astype(np.float64) - norm[2]) / norm[3]
    l_ref, g_ref = OM.loss_grad(flat.astype(np.float64), X, t, emulate=True)
    assert abs(loss - l_ref) <= 1e-3 * l_ref
    for k, (a, b, _) in OM.SLICES3.items():
        rb = np.linalg.norm(g[a:b] - g_ref[a:b]) / max(np.linalg.norm(g_ref[a:b]), 1e-30)
        print(k, rb)
        assert rb < 3e-2, (k, rb)


@pytest.fixture(scope="module")
def trained_rate():
    from sfparticles_amd.qed.mlp import EmissionRateNet
    s = EmissionRateNet(seed=0)
    curve = s.fit(n_data=8_000_000, steps=40000, batch=16384, lr=2e-3, seed=1)
    err, thr = s.calibrate_fallback(0.01)
    return s, curve, err


@pytest.mark.gpu
def test_rate_net_training_and_bins(trained_rate):
    s, curve, err = trained_rate
    nfb = bin(s.fallback).count("1")
    print("loss: first", curve[:5].mean(), "last", curve[-50:].mean(), "; bin errors max",
          err.max(), "median", np.median(err), "fallback bins", nfb)
    assert curve[-50:].mean() < 0.02 * curve[:5].mean()
    assert nfb <= 8  # the network serves most of the chi range
    assert np.median(err) < 0.005


@pytest.mark.gpu
def test_rate_net_in_the_step_static_field(trained_rate):
    """Spin asymmetry in a static field with the learned rate in the step:
    each yield within 3 sigma of the table's expectation."""
    from tests.test_spin_emission import _static_field_run, _expected_probability
    s, _, _ = trained_rate
    out = _static_field_run("mixed", rate_net=s)
    (k_up, k_dn, m, chi, gamma, dt) = out
    p_up = _expected_probability(chi, gamma, dt, 1.0)
    p_dn = _expected_probability(chi, gamma, dt, -1.0)
    print(f"network in the step: +zeta {k_up} (expected {m * p_up:.0f}), "
          f"-zeta {k_dn} (expected {m * p_dn:.0f})")
    for k, p in ((k_up, p_up), (k_dn, p_dn)):
        assert abs(k - m * p) <= 3 * np.sqrt(m * p * (1 - p)) + 1


@pytest.mark.gpu
def test_rate_net_cost_against_table(trained_rate):
    """What the network costs next to the table lookup it replaces (batched
    evaluation throughput and the bytes each needs); recorded, not asserted
    beyond both running."""
    s, _, _ = trained_rate
    rng = np.random.Generator(np.random.PCG64(12))
    n = 4_000_000
    lc, r1, sg = rng.uniform(-3, 2, n), rng.random(n), rng.uniform(-1, 1, n)
    s.rate(lc[:4096], r1[:4096], sg[:4096], "mlp")
    s.rate(lc[:4096], r1[:4096], sg[:4096], "table")
    _, ms_net = s.rate(lc, r1, sg, "mlp")
    _, ms_tab = s.rate(lc, r1, sg, "table")
    net_bytes = 2 * 2 * 4096 + 4 * (192 + 64 * 4 + 1)
    tab_bytes = 1023 * (32 + 16)
    print(f"{n} rates: network {ms_net:.3f} ms ({n / ms_net / 1e6:.2f} G/s), FP64 table "
          f"{ms_tab:.3f} ms ({n / ms_tab / 1e6:.2f} G/s); bytes network {net_bytes}, table "
          f"{tab_bytes}")
    assert ms_net > 0 and ms_tab > 0
