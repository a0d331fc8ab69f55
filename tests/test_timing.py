"""Sampled kernel timing (Simulation.run_steps(timing=True, timing_every=k)).

The benchmark times the step kernels of every k-th step with HIP events, so
that the events' own cost stays out of most steps.  Timing must not change
results, and exactly the launches of the sampled steps are counted.
"""
import numpy as np
import pytest

from tests.cases import head_on_case

KEYS = ("x", "y", "z", "ux", "uy", "uz")


def _make(case):
    import sfparticles_amd as sfa
    from sfparticles_amd import qed as Q
    from sfparticles_amd.fields import Fields
    Q.set_model(case["qed"])
    n = case["props"][0].shape[0]
    e = sfa.Particles("electron", -1, 1, N=n, props=case["props"],
                      RR=sfa.RadiationReactionType.PHOTON)
    p = sfa.Particles("positron", 1, 1)
    g = sfa.Particles("photon", 0, 0)
    e.set_photon(g)
    p.set_photon(g)
    g.set_pair(e, p)
    sim = sfa.Simulation(e, p, g, dt=case["dt"], fields=Fields(terms=case["field"]),
                         print_every=None, photon_threshold=case["threshold"], seed=case["seed"])
    # room for every particle the run creates: no capacity halt, whose replay
    # would launch (and time) the halted step's kernels again
    for s, cap in zip((e, p, g), (20000, 20000, 400000)):
        s._ensure_device()
        s._extend(cap - s.N_buffered)
    return sim, (e, p, g)


@pytest.mark.gpu
def test_sampled_timing_counts_the_sampled_steps_and_changes_nothing():
    case = head_on_case(3000, 80, seed=21, a0=150.0)
    ref, rs = _make(case)
    ref.run_steps(40)
    ref.run_steps(40)
    sim, ss = _make(case)
    sim.run_steps(40, timing=True, timing_every=3)
    sim.run_steps(40, timing=True, timing_every=3)
    # one launch per kernel and step without halts; steps 0, 3, ..., 78 are timed
    ms, n = sim.kernel_time(0, 0)
    ms1, n1 = sim.kernel_time(0, 1)
    print("electron sweep", ms, "ms over", n, "launches; QED", ms1, n1, "; photons",
          ss[2].N_buffered, "; replays", sim.stats(0)["replays"])
    assert sim.stats(0)["replays"] == 0
    assert n == n1 == len(range(0, 80, 3)) and ms > 0.0 and ms1 > 0.0
    assert ss[2].N_buffered > 0
    for a, b in zip(rs, ss):
        assert a.N_buffered == b.N_buffered
        for k in KEYS:
            assert np.array_equal(np.asarray(getattr(a, k)), np.asarray(getattr(b, k)))


@pytest.mark.gpu
def test_timing_kernel_is_timed_at_every_step():
    """The kernel named by timing_kernel (the electron QED kernel, as the
    benchmark names it) is timed at every step, the others on the sampled
    steps only; results unchanged."""
    case = head_on_case(3000, 80, seed=22, a0=150.0)
    ref, rs = _make(case)
    ref.run_steps(80)
    sim, ss = _make(case)
    sim.run_steps(80, timing=True, timing_every=5, timing_kernel=(0, 1))
    ms0, n0 = sim.kernel_time(0, 0)
    ms1, n1 = sim.kernel_time(0, 1)
    ms2, n2 = sim.kernel_time(2, 1)
    print("electron sweep", n0, "launches; electron QED", n1, "; photon QED", n2)
    assert sim.stats(0)["replays"] == 0
    assert n1 == 80 and ms1 > 0.0
    assert n0 == n2 == len(range(0, 80, 5))
    for a, b in zip(rs, ss):
        assert a.N_buffered == b.N_buffered
        for k in KEYS:
            assert np.array_equal(np.asarray(getattr(a, k)), np.asarray(getattr(b, k)))


def test_timing_kernel_rejects_a_missing_kernel():
    """Checked on the host before anything else (no device needed)."""
    import os
    from sfparticles_amd import _lib as L
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("library not built")
    rc = L.load().sfp_sim_set_timing_kernel(None, 0, 1)
    with pytest.raises(L.SFPError, match="null simulation"):
        L.check(rc)


def test_timing_every_rejects_zero():
    """Checked on the host before anything else (no device needed)."""
    import os
    from sfparticles_amd import _lib as L
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("library not built")
    rc = L.load().sfp_sim_set_timing_every(None, 0)
    with pytest.raises(L.SFPError, match="timing_every"):
        L.check(rc)
