"""Device diagnostics of the exact QED evaluations (Simulation.qed_diag,
sfp_sim_qed_diag) and the dense lepton step (Simulation.stats
"dense_steps"), against counts computed on the CPU from the same Philox
draws and the oracle's rate.

In a static magnetic field with radiation reaction off, an electron's chi
is the same every step (the Boris rotation keeps gamma), so the
rejection-sampling probability of every (particle, step) is
p = 3 r1^2 w(chi, r1^3) dt / gamma with r1 from its own Philox words
(qed/rejection_sampling.py:68-85): the number of p > 1 evaluations, the
largest p (as a float) and the number of draws with z beyond the table
(delta / ((1 - delta) chi) > z_max^1.5, qed/rejection_sampling_tables.py:
40-51) follow exactly.  Under optical depth with chi above the table's top
(qed/optical_depth_tables.py:39-40) every evaluation is a clamp.
"""
import ctypes as C

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GAMMA = 1e4


def _sim(chi, dt, n, model="rs", steps=3):
    import sfparticles_amd as sfa
    from sfparticles_amd import qed as Q
    from sfparticles_amd.fields import Fields
    from sfparticles_amd.constants import m_e, c, e, hbar
    Q.set_model(model)
    ux = np.sqrt(GAMMA ** 2 - 1)
    z = np.zeros(n)
    ele = sfa.Particles("ele", -1, 1, N=n, props=(z, z.copy(), z.copy(), np.full(n, ux),
                                                z.copy(), z.copy()),
                        RR=sfa.RadiationReactionType.NONE)
    pho = sfa.Particles("pho", 0, 0)
    ele.set_photon(pho)
    Bz = chi * m_e ** 2 * c ** 2 / e / hbar / ux
    sim = sfa.Simulation(ele, pho, dt=dt, fields=Fields(terms=[(0, (0, 0, 0, 0, 0, Bz))]),
                         print_every=None, photon_threshold=0.0, seed=0)
    sim.run_steps(steps)
    d = sim.qed_diag(0)
    Q.set_model("rs")
    return d, sim


def _expected_rs(chi_target, dt, n, steps):
    """The counts from the electrons' own draws: chi and 1/gamma as the step
    computes them (the oracle's update_chi on the same field and momentum;
    the Boris rotation in a pure magnetic field changes neither beyond
    rounding), p = 3 r1^2 w(chi, r1^3) dt / gamma."""
    from sfparticles_amd.qed.tables import load_rs_tables
    from sfparticles_amd.constants import m_e, c, e, hbar
    L = oracle.lib()
    T = oracle.rs_struct(load_rs_tables())
    ux = np.sqrt(GAMMA ** 2 - 1)
    ig = 1.0 / np.sqrt(1.0 + ux * ux)
    Bz = chi_target * m_e ** 2 * c ** 2 / e / hbar / ux
    z, chi = np.zeros(1), np.zeros(1)
    args = (z, z, z, z, z, np.array([Bz]), np.array([ux]), z, z, np.array([ig]), chi)
    no = np.zeros(1, np.uint8)
    L.orc_update_chi(*[oracle.ptr(a) for a in args], 1, oracle.ptr(no))
    chi = float(chi[0])
    p1 = zc = 0
    pmax = 0.0
    zmax15 = 100.0 ** 1.5
    for step in range(steps):
        for i in range(n):
            d = oracle.draw4(0, i, step, 0, 0)
            r1 = oracle.u01(d[0], d[1])
            r13 = r1 * r1 * r1
            w = L.orc_photon_prob_rate(chi, r13, C.byref(T))
            p = 3 * (r1 * r1) * w * (dt * ig)
            p1 += p > 1
            zc += r13 > zmax15 * chi * (1.0 - r13)
            pmax = max(pmax, float(np.float32(p)))
    return p1, zc, pmax


def test_probability_above_one_and_its_maximum():
    """chi = 3 and a step long enough that the soft end of the spectrum has
    p > 1 (the reference samples one event per step regardless)."""
    n, steps, chi, dt = 3000, 3, 3.0, 8e-16
    d, _ = _sim(chi, dt, n, steps=steps)
    p1, zc, pmax = _expected_rs(chi, dt, n, steps)
    print("device", d, "expected", p1, zc, pmax)
    assert 0 < p1 < n * steps
    assert d["p_over1"] == p1 and d["z_clamp"] == zc
    assert d["p_max"] == pmax and d["chi_nan"] == 0 and d["chi_clamp"] == 0


def test_photon_z_beyond_the_table():
    """chi = 1e-3: a fifth of the delta = r1^3 draws lie beyond z_max."""
    n, steps, chi, dt = 3000, 2, 1e-3, 1e-16
    d, _ = _sim(chi, dt, n, steps=steps)
    p1, zc, pmax = _expected_rs(chi, dt, n, steps)
    print("device", d, "expected", p1, zc, pmax)
    assert zc > 0.15 * n * steps
    assert d["z_clamp"] == zc and d["p_over1"] == p1 == 0 and d["p_max"] == pmax


def test_optical_depth_chi_above_the_table():
    n, steps = 2000, 3
    d, _ = _sim(200.0, 1e-19, n, model="od", steps=steps)
    print("device", d)
    assert d["chi_clamp"] == n * steps and d["p_over1"] == 0


def test_dense_steps_in_the_collision():
    """The head-on collision lists every electron at its centre: the QED
    kernel then visits the store by index (dense steps) and the result is
    still the oracle's, bit for bit."""
    from tests.cases import head_on_case, run_gpu_case, run_oracle_case, compare_final
    case = head_on_case(n=3000, nsteps=240, seed=11)
    gpu, sim = run_gpu_case(case)
    st = sim.stats(0)
    print("electron stats", st)
    assert st["dense_steps"] > 50
    compare_final(gpu, run_oracle_case(case))


@pytest.mark.parametrize("graph", [False, True])
def test_lookback_flags_cleared_between_batches(graph):
    """The look-back flags and epochs cleared before every batch (the
    library does it every 2^28 steps, before the 29-bit epoch could wrap):
    ranks, ids and states stay the oracle's, eagerly and in graph replays."""
    from tests.cases import head_on_case, build_gpu_case, run_oracle_case, compare_final, _collect
    case = head_on_case(n=3000, nsteps=240, seed=11)
    sim, sp = build_gpu_case(case, graph=graph)
    sim.set_flag_clear_every(1)
    for m in (7, 64, 64, 41):  # uneven run calls; full batches replay as graphs
        sim.run_steps(m)
    sim.start(64)  # the last call through start(), which prunes as the oracle's does
    if graph:
        assert sim.stats(0)["graph_batches"] >= 3
    compare_final(_collect(sp), run_oracle_case(case))
