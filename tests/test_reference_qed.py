"""The reference's statistical QED tests on the GPU, as the reference writes
them: tests/test_qed.py (TestPhotonNumber, TestPairNumber) and
tests/test_qed_optical_depth.py (the same classes under the optical-depth
model).  Same particle counts (1e7 electrons, 1e6 photons), the same
chi values, fields and intervals, the same expected counts (the quad
integral of the closed-form rate, qed/optical_depth_tables.py:126-140) and
the same 3-sigma bounds, with nothing added.  Each test runs the
per-operation GPU kernels through the reference's Particles methods.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GAMMA = 10000
UX = np.sqrt(GAMMA ** 2 - 1)


def _bfield(chi):
    from sfparticles_amd.fields import static_field
    from sfparticles_amd.constants import m_e, c, e, hbar
    return static_field(Bz=chi * m_e ** 2 * c ** 2 / e / hbar / UX)


def _beam(name, q, m, N, **kw):
    import sfparticles_amd as sfa
    z = np.zeros(N)
    return sfa.Particles(name, q, m, N=N, props=(z, z.copy(), z.copy(), np.full(N, UX),
                                                 z.copy(), z.copy()), **kw)


@pytest.fixture(params=["rs", "od"])
def model(request):
    from sfparticles_amd import qed as Q
    Q.set_model(request.param)
    yield request.param
    Q.set_model("rs")


def test_hard_photon(model):
    """test_qed.py:54-68: after _pick_hard_photon(2.0) every created photon
    has gamma > 2 (chi = 1, dt = 1e-17 s, 1e7 electrons)."""
    import sfparticles_amd as sfa
    N = 10_000_000
    ele = _beam("ele", -1, 1, N, RR=sfa.RadiationReactionType.NONE)
    pho = sfa.Particles("pho", 0, 0)
    ele.set_photon(pho)
    ele._eval_field(_bfield(1.0), 0)
    ele._calculate_chi()
    ev, delta = ele._photon_event(1e-17)
    n_events = int(np.count_nonzero(np.array(ev)[:N]))
    ele._pick_hard_photon(2.0)
    n_hard = int(np.count_nonzero(np.array(ele.event)[:N]))
    ele._create_photon(pho)
    g = np.array(pho.gamma)[:pho.N_buffered]
    print(model, "events", n_events, "hard", n_hard, "photons", pho.N_buffered)
    assert pho.N_buffered == n_hard > 0
    assert n_hard < n_events  # the cut removed the soft photons
    assert (g > 2.0).all()


def test_photon_number(model):
    """test_qed.py:70-107: photon number within 3 sigma of N W(chi) tau for
    chi = 0.1, 0.5, 1, 2, 5, and the counts in the delta bins [0.01, 0.011]
    and [0.1, 0.11] within 3 sigma of the trapezoid of the closed-form rate;
    1e7 electrons reused across the chi values as the reference does."""
    import sfparticles_amd as sfa
    from sfparticles_amd.qed.tables import photon_rate_for_delta, photon_total_rate
    N = 10_000_000
    ele = _beam("ele", -1, 1, N, RR=sfa.RadiationReactionType.NONE)
    failures = []
    for chi_e in (0.1, 0.5, 1.0, 2.0, 5.0):
        interval = 1e-17 / chi_e
        tau = interval / GAMMA
        pho = sfa.Particles("pho", 0, 0)
        ele.set_photon(pho)
        ele._eval_field(_bfield(chi_e), 0)
        ele._calculate_chi()
        ele._photon_event(interval)
        ele._create_photon(pho)
        n_photon = pho.Npart
        p = photon_total_rate(chi_e) * tau
        expect = p * N
        tor = 3 * np.sqrt(expect * (1 - p))
        print(f"{model} chi {chi_e}: photons {n_photon}, expected {expect:.1f} +- {tor / 3:.1f}")
        if not abs(n_photon - expect) < tor:
            failures.append((chi_e, n_photon, expect, tor))
        P = photon_rate_for_delta(chi_e)
        u = np.array(pho.ux)[:pho.N_buffered] / GAMMA
        for delta in (0.01, 0.1):
            nb = np.histogram(u, bins=[delta, delta * 1.1])[0][0]
            pb = (P(delta) + P(delta * 1.1)) * (delta * 0.1) / 2 * tau
            eb = pb * N
            tb = 3 * np.sqrt(eb * (1 - pb))
            print(f"    bin [{delta}, {delta * 1.1:.3g}]: {nb}, expected {eb:.1f} +- {tb / 3:.1f}")
            if not abs(nb - eb) < tb:
                failures.append((chi_e, delta, nb, eb, tb))
    assert not failures, failures


def test_pair_number(model):
    """test_qed.py:176-200: pair number within 3 sigma of N W_pair(chi) tau for
    chi_gamma = 1, 2, 5 (1e6 photons, interval 1e-17 s); a fresh photon beam
    per chi, so that no photon is already converted."""
    import sfparticles_amd as sfa
    from sfparticles_amd.qed.tables import pair_total_rate
    N = 1_000_000
    interval = 1e-17
    tau = interval / GAMMA
    failures = []
    for chi_gamma in (1.0, 2.0, 5.0):
        pho = _beam("pho", 0, 0, N)
        ele = sfa.Particles("ele", -1, 1)
        pos = sfa.Particles("pos", 1, 1)
        pho.set_pair(ele, pos)
        pho._eval_field(_bfield(chi_gamma), 0)
        pho._calculate_chi()
        pho._pair_event(interval)
        pho._create_pair(ele, pos)
        n_pair = pos.Npart
        p = pair_total_rate(chi_gamma) * tau
        expect = p * N
        tor = 3 * np.sqrt(expect * (1 - p))
        print(f"{model} chi_gamma {chi_gamma}: pairs {n_pair}, expected {expect:.1f} +- {tor / 3:.1f}")
        assert ele.Npart == n_pair and N - pho.Npart == n_pair
        if not abs(n_pair - expect) < tor:
            failures.append((chi_gamma, n_pair, expect, tor))
    assert not failures, failures
