"""Device-side trajectory recorder (Simulation.record / records).

The reference records trajectories with a per-step callback that copies the
electron arrays to the host (example/trajectory.py:62-70), called after the
step's push and before its clock advances (simulation.py:140-144).  The recorder
stores the same six arrays on the device from a kernel at the end of each
recorded step, also inside graph replays and across capacity halts.  The GPU
tests hold it to the callback pattern bit for bit: every recorded row equals
what a host read after that step returns in a separate run with the same
inputs (the step itself is parity-tested against the oracle elsewhere), slots
past the species' count read NaN, and the recorded times equal the
simulation clock.
"""
import numpy as np
import pytest

from tests.cases import head_on_case

KEYS = ("x", "y", "z", "ux", "uy", "uz")


def _make(case, **kw):
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
                         print_every=None, photon_threshold=case["threshold"], seed=case["seed"],
                         **kw)
    return sim, e, g


def _callback_rows(case, nsteps, n, every, **kw):
    """Rows read on the host after every `every`-th step (the reference's
    callback pattern), NaN-padded to n slots, with the clock."""
    sim, e, _ = _make(case, **kw)
    rows, ts = [], []

    def cb(istep):
        if (istep + 1) % every:
            return
        m = min(n, e.N_buffered)
        row = np.full((n, 6), np.nan)
        for j, k in enumerate(KEYS):
            row[:m, j] = np.array(getattr(e, k))[:m]
        rows.append(row)
        ts.append(sim.t + sim.dt)  # the callback sees the step's start time

    sim.start(nsteps, call_back=cb)
    return np.array(rows), np.array(ts)


def _recorded(case, nsteps, n, every, **kw):
    sim, e, g = _make(case, **kw)
    r0 = sim.records()
    assert r0["t"].size == 0
    sim.record(e, nsteps, n=n, every=every)
    sim.start(nsteps)
    r = sim.records()
    rows = np.stack([r[k] for k in KEYS], axis=-1)
    return rows, r["t"], sim, e, g


@pytest.mark.gpu
def test_recorder_matches_per_step_callback():
    case = head_on_case(3000, 200, seed=11, a0=150.0)
    want, t_want = _callback_rows(case, 200, 1000, 1)
    got, t_got, sim, _, g = _recorded(case, 200, 1000, 1)
    print("rows", got.shape, "photons", g.N_buffered)
    assert got.shape == (200, 1000, 6) and g.N_buffered > 0
    assert np.array_equal(got, want)
    assert np.array_equal(t_got, t_want)


@pytest.mark.gpu
def test_recorder_in_graph_replays_with_growth_and_padding():
    # stores start at the seed count and grow (capacity halts inside graph
    # replays); n exceeds the electron count, which grows through pairs
    case = head_on_case(6000, 256, seed=12, a0=250.0)
    n = 6400
    want, t_want = _callback_rows(case, 256, n, 4)
    got, t_got, sim, e, g = _recorded(case, 256, n, 4, graph=True)
    st = sim.stats(0)
    print("rows", got.shape, "graph batches", st["graph_batches"], "replays", st["replays"],
          "electrons", e.N_buffered, "photons", g.N_buffered)
    assert got.shape == (64, n, 6) and st["graph_batches"] >= 3
    assert np.isnan(got[0, 6000:]).all()
    assert np.array_equal(got, want, equal_nan=True)
    assert np.array_equal(t_got, t_want)


@pytest.mark.gpu
def test_recorder_rejects_photons_and_can_be_removed():
    from sfparticles_amd._lib import SFPError
    case = head_on_case(500, 64, seed=13, a0=50.0)
    sim, e, g = _make(case)
    with pytest.raises(SFPError, match="lepton"):
        sim.record(g, 10)
    sim.record(e, 10, n=100)
    sim.start(4)
    assert sim.records()["x"].shape == (4, 100)
    sim.record(e, 10, n=0)
    sim.start(4)
    assert sim.records()["t"].size == 0


@pytest.mark.gpu
def test_callback_every_k_steps():
    """start(call_back=..., callback_every=k) calls back after the steps whose
    index + 1 is a multiple of k, with the state a per-step callback sees."""
    case = head_on_case(2000, 64, seed=14, a0=150.0)
    every1, every5 = {}, {}

    def grab(store, sim, e):
        def cb(istep):
            store[istep] = (np.array(e.x)[:100].copy(), sim.step)
        return cb

    sim, e, _ = _make(case)
    sim.start(40, call_back=grab(every1, sim, e))
    sim5, e5, _ = _make(case)
    sim5.start(40, call_back=grab(every5, sim5, e5), callback_every=5)
    assert sorted(every5) == [4, 9, 14, 19, 24, 29, 34, 39]
    for istep, (x, step) in every5.items():
        assert step == istep  # the reference's counter at its callback
        assert np.array_equal(x, every1[istep][0])


def _oracle_rows(case, nsteps, n, every):
    """The oracle's start(call_back=...) (simulation.py:140-141) recording the
    first n electrons every `every` steps (NaN past the count), with the
    time after the step and the clock and counter the callback sees."""
    from tests.cases import build_oracle_case
    sim, sp = build_oracle_case(case)
    e = sp[0]
    rows, ts, seen = [], [], []

    def cb(istep):
        seen.append((istep, sim.step, sim.t))
        if (istep + 1) % every:
            return
        m = min(n, e.N_buffered)
        row = np.full((n, 6), np.nan)
        for j, k in enumerate(KEYS):
            row[:m, j] = getattr(e, k)[:m]
        rows.append(row)
        ts.append(sim.t + sim.dt)

    sim.start(nsteps, call_back=cb)
    return np.array(rows), np.array(ts), seen


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_recorder_matches_oracle_callback(graph):
    """Device recorder rows (eager and in graph replays, with store growth)
    equal the oracle's per-step callback rows bit for bit, and the recorded
    times the oracle clock after each recorded step."""
    case = head_on_case(3000, 192, seed=15, a0=150.0)
    want, t_want, _ = _oracle_rows(case, 192, 3200, 3)
    got, t_got, sim, e, g = _recorded(case, 192, 3200, 3, graph=graph)
    print("rows", got.shape, "photons", g.N_buffered, "graph batches", sim.stats(0)["graph_batches"])
    assert got.shape == (64, 3200, 6) and g.N_buffered > 0
    assert np.array_equal(got, want, equal_nan=True)
    assert np.array_equal(t_got, t_want)


@pytest.mark.gpu
def test_callback_every_matches_oracle_callback():
    """start(call_back=..., callback_every=k): at each call the electrons'
    state, the clock and the step counter are those the oracle's callback
    sees after the same step."""
    case = head_on_case(2000, 60, seed=16, a0=150.0)
    _, _, seen = _oracle_rows(case, 60, 1, 1)
    ref = {}
    from tests.cases import build_oracle_case
    osim, osp = build_oracle_case(case)
    osim.start(60, call_back=lambda i: ref.__setitem__(i, np.array(osp[0].x[:200])))
    got = {}
    sim, e, _ = _make(case)
    sim.start(60, call_back=lambda i: got.__setitem__(i, (np.array(e.x)[:200], sim.step, sim.t)),
              callback_every=6)
    assert sorted(got) == list(range(5, 60, 6))
    clock = {i: (step, t) for i, step, t in seen}
    for istep, (x, step, t) in got.items():
        assert np.array_equal(x, ref[istep])
        assert (step, t) == clock[istep]
