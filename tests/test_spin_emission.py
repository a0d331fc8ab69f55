"""Spin-resolved photon emission (no counterpart in the reference, whose
emission ignores spin).

CPU pins of the oracle's restatement:
  * the spin term against an independent evaluation of the LCFA spectrum in
    Bessel form with scipy (kv and quad): the ratio (rate at sigma = +-1) /
    (unpolarised rate) = 1 -+ u K_1/3(y) / [(1 - u + 1/(1-u)) K_2/3(y) -
    int_y K_1/3], y = 2u / (3 chi (1-u)) -- a dropped factor, a wrong power of
    z or a sign error fails it;
  * sigma = 0 gives the reference's rate exactly, the +-1 mean gives it to
    rounding, and the rate stays non-negative for |sigma| <= 1;
  * the Ai table against the reference-pinned int Ai table: Simpson's rule
    on Ai over two intervals gives the difference of int_z^inf Ai;
  * zeta against beta x dbeta/dt computed from the Lorentz force itself
    (including the -beta (beta.E) term the closed form drops).
GPU: the fused FP64 step with spin-resolved emission bit for bit against the
oracle (counts, ids, compaction order, states); in a static magnetic field the
photon yields of electrons with spin along +zeta and -zeta each match the
expectation integrated from the oracle's rate (3 sigma), in FP64 and in mixed
precision.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests.cases import head_on_case, run_gpu_case, run_oracle_case, compare_final

E_CHARGE = 1.602176634e-19
C_LIGHT = 299792458.0


def _lib():
    from sfparticles_amd.qed.tables import load_rs_tables
    L = oracle.lib(False)
    return L, oracle.rs_struct(load_rs_tables())


def test_spin_term_matches_bessel_form():
    from scipy.integrate import quad
    from scipy.special import kv
    L, T = _lib()
    worst = 0.0
    for chi in (0.05, 0.3, 1.0, 4.0, 20.0):
        for u in (1e-3, 0.02, 0.1, 0.3, 0.6, 0.9):
            y = 2 * u / (3 * chi * (1 - u))
            if y > 60:
                continue
            ik = quad(lambda t: kv(1 / 3, t), y, np.inf, limit=200)[0]
            ratio = u * kv(1 / 3, y) / ((1 - u + 1 / (1 - u)) * kv(2 / 3, y) - ik)
            w0 = L.orc_photon_prob_rate(chi, u, C.byref(T))
            for sig in (1.0, -1.0):
                w = L.orc_photon_prob_rate_spin(chi, u, sig, C.byref(T))
                err = abs(w / w0 - (1 - sig * ratio))
                worst = max(worst, err)
                # table interpolation error of the two brackets
                assert err < 2e-3, (chi, u, sig, w / w0, 1 - sig * ratio)
    print("max |ratio - Bessel form|", worst)


def test_spin_rate_reduces_to_reference_rate_and_stays_positive():
    L, T = _lib()
    rng = np.random.Generator(np.random.PCG64(2))
    for chi, u in zip(10 ** rng.uniform(-3, 2, 3000), rng.uniform(1e-6, 1 - 1e-6, 3000)):
        w0 = L.orc_photon_prob_rate(chi, u, C.byref(T))
        assert L.orc_photon_prob_rate_spin(chi, u, 0.0, C.byref(T)) == w0
        up = L.orc_photon_prob_rate_spin(chi, u, 1.0, C.byref(T))
        dn = L.orc_photon_prob_rate_spin(chi, u, -1.0, C.byref(T))
        assert up >= 0.0 and dn >= 0.0 and up <= w0 <= dn
        assert abs(0.5 * (up + dn) - w0) <= 1e-12 * w0


def test_ai_table_is_the_derivative_of_int_ai():
    from sfparticles_amd.qed.tables import load_rs_tables
    T = load_rs_tables()
    h = T.z_delta
    # int Ai(z_i) - int Ai(z_i+2) = integral of Ai over [z_i, z_i+2], Simpson
    lhs = T.int_Ai[:-2] - T.int_Ai[2:]
    rhs = h / 3 * (T.Ai[:-2] + 4 * T.Ai[1:-1] + T.Ai[2:])
    err = np.max(np.abs(lhs - rhs))
    print("max |Simpson(Ai) - difference of int Ai|", err)
    assert err < 1e-6 * np.max(lhs)
    assert T.Ai[0] == pytest.approx(0.355028053887817, rel=1e-12)  # Ai(0)


def test_zeta_is_beta_cross_dbeta_dt():
    L, _ = _lib()
    rng = np.random.Generator(np.random.PCG64(4))
    for _ in range(300):
        q = rng.choice([-1.0, 1.0]) * E_CHARGE
        u = rng.normal(0, 300, 3)
        E = rng.normal(0, 1e12, 3)
        B = rng.normal(0, 3e3, 3)
        s = rng.normal(size=3)
        s /= np.linalg.norm(s)
        g = np.sqrt(1 + u @ u)
        beta = u / g
        # dbeta/dt from the Lorentz force, m = 1 (only the direction matters)
        dbeta = q / g * (E + C_LIGHT * np.cross(beta, B) - beta * (beta @ E))
        zeta = np.cross(beta, dbeta)
        want = s @ zeta / np.linalg.norm(zeta)
        got = L.orc_spin_sigma(q, *u, 1 / g, *E, *B, *s)
        assert abs(got - want) < 1e-9, (got, want)
    # the textbook case: electron along +x in B along +z -> zeta = +z
    assert L.orc_spin_sigma(-E_CHARGE, 1000.0, 0, 0, 1e-3, 0, 0, 0, 0, 0, 1e6, 0, 0, 1) == 1.0
    assert L.orc_spin_sigma(E_CHARGE, 1000.0, 0, 0, 1e-3, 0, 0, 0, 0, 0, 1e6, 0, 0, 1) == -1.0


@pytest.mark.gpu
def test_spin_emission_matches_oracle():
    case = head_on_case(2000, 600, seed=7, a0=150.0, spin=True)
    got, _ = run_gpu_case(case, spin_emission=True)
    ref = run_oracle_case(case, spin_emission=True)
    n = {k: v["n"] for k, v in ref.items()}
    print("counts", n)
    assert n["photon"] > 100
    compare_final(got, ref)


def _expected_probability(chi, gamma, dt, sigma, threshold=2.0, m=100000):
    """Per-step emission probability of a hard photon for a lepton at chi,
    gamma: the mean over r1 of min(1, 3 r1^2 w(chi, r1^3, sigma) dt/gamma),
    delta = r1^3 >= threshold / gamma, from the oracle's rate."""
    L, T = _lib()
    r = (np.arange(m) + 0.5) / m
    w = np.array([L.orc_photon_prob_rate_spin(chi, x ** 3, sigma, C.byref(T)) for x in r])
    p = np.minimum(1.0, 3 * r ** 2 * w * dt / gamma)
    return float(np.mean(np.where(r ** 3 * gamma >= threshold, p, 0.0)))


def _static_field_run(precision, rate_net=None, n=400000, gamma=1000.0, chi=1.0):
    """One step of n electrons along +x in a static B along +z (zeta = +z),
    half with spin +z, half -z; returns the counts of emitters of each half
    (electrons that recoiled: the field only rotates u, a hard photon takes
    at least 2/gamma of |u|) and the parameters."""
    import sfparticles_amd as sfa
    from sfparticles_amd import qed as Q
    from sfparticles_amd.fields import Fields
    Q.set_model("rs")
    # chi = gamma |beta x B| c / E_s for u along x, B along z
    E_s = 1.32328547e18
    B = chi * E_s / C_LIGHT / np.sqrt(gamma ** 2 - 1)
    dt = 0.8e-6 / C_LIGHT / 100
    ux = np.full(n, np.sqrt(gamma ** 2 - 1))
    z0 = np.zeros(n)
    sz = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    e = sfa.SpinParticles("electron", -1, 1, N=n,
                          props=(z0, z0.copy(), z0.copy(), ux, z0.copy(), z0.copy(), z0.copy(),
                                 z0.copy(), sz))
    g = sfa.Particles("photon", 0, 0)
    e.set_photon(g)
    sim = sfa.Simulation(e, g, dt=dt, fields=Fields(terms=[(0, (0, 0, 0, 0, 0, B))]),
                         print_every=None, seed=3, precision=precision, spin_emission=True,
                         rate_net=rate_net)
    sim.start(1)
    u_after = np.sqrt(np.array(e.ux)[:n] ** 2 + np.array(e.uy)[:n] ** 2 + np.array(e.uz)[:n] ** 2)
    emitted = u_after < ux * (1 - 1e-4)
    k_up, k_dn = int(emitted[sz > 0].sum()), int(emitted[sz < 0].sum())
    assert k_up + k_dn == g.N_buffered
    return k_up, k_dn, n // 2, chi, gamma, dt


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_spin_asymmetry_in_static_field(precision):
    k_up, k_dn, m, chi, gamma, dt = _static_field_run(precision)
    p_up = _expected_probability(chi, gamma, dt, 1.0)
    p_dn = _expected_probability(chi, gamma, dt, -1.0)
    print(f"{precision}: emitters spin +zeta {k_up} (expected {m * p_up:.0f}), "
          f"-zeta {k_dn} (expected {m * p_dn:.0f})")
    for k, p in ((k_up, p_up), (k_dn, p_dn)):
        assert abs(k - m * p) <= 3 * np.sqrt(m * p * (1 - p)) + 1
    assert k_dn - k_up > 5 * np.sqrt(k_dn + k_up)  # the asymmetry itself is resolved
