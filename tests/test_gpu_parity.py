"""GPU parity: the HIP implementation against the CPU oracle on the same
seeded inputs (Philox draws replayed on the CPU).

Bar: bit-exact particle counts, ids (= compaction order) and FP64 state for
the fused step and the per-operation kernels; QED-off trajectories of the
oracle's libm build within 1e-10 relative (the reference's own criterion for
deterministic trajectories).
"""
import ctypes as C

import numpy as np
import pytest

from tests.cases import (head_on_case, cascade_case, run_gpu_case, run_oracle_case,
                         compare_final)

pytestmark = pytest.mark.gpu


def _names(out):
    return {k: v["n"] for k, v in out.items()}


def test_device_math_matches_host():
    """sfp_math.h on the device gives the same bits as on the host."""
    import oracle
    from sfparticles_amd import _lib as L
    rng = np.random.default_rng(1)
    n = 200000
    cases = {
        "exp": (rng.uniform(-740, 705, n), None),
        "log": (np.exp(rng.uniform(-700, 700, n)), None),
        "sin": (rng.uniform(-3000, 3000, n), None),
        "cos": (rng.uniform(-3000, 3000, n), None),
        "atan": (np.sinh(rng.uniform(-30, 30, n)), None),
        "pow": (np.exp(rng.uniform(-50, 50, n)), rng.uniform(-3, 3, n)),
        "exp10": (rng.uniform(-300, 300, n), None),
        "log10": (np.exp(rng.uniform(-700, 700, n)), None),
    }
    names = ["exp", "log", "sin", "cos", "atan", "pow", "exp10", "log10"]
    for fn, (x, y) in cases.items():
        yv = np.zeros_like(x) if y is None else y
        dx, dy, do = L.DeviceArray.of(x), L.DeviceArray.of(yv), L.DeviceArray(x.nbytes)
        L.call("sfp_math_eval", names.index(fn), dx.ptr, dy.ptr, do.ptr, n, None)
        got = np.empty_like(x)
        do.download(got)
        want = oracle.math(fn, x, yv)
        assert np.array_equal(got, want), f"{fn}: {np.sum(got != want)} values differ"
    # IEEE sqrt and division are correctly rounded on both sides
    x = np.exp(rng.uniform(-300, 300, n))
    y = np.exp(rng.uniform(-300, 300, n))
    for code, want in ((8, np.sqrt(x)), (9, x / y)):
        dx, dy, do = L.DeviceArray.of(x), L.DeviceArray.of(y), L.DeviceArray(x.nbytes)
        L.call("sfp_math_eval", code, dx.ptr, dy.ptr, do.ptr, n, None)
        got = np.empty_like(x)
        do.download(got)
        assert np.array_equal(got, want), f"op {code}: {np.sum(got != want)} differ"


def test_head_on_rs_bit_exact():
    case = head_on_case(n=3000, nsteps=240, seed=11)
    gpu, sim = run_gpu_case(case)
    ref = run_oracle_case(case)
    print("counts", _names(gpu), "stats", [sim.stats(i) for i in range(3)])
    assert ref["photon"]["n"] > 1000  # the collision happened
    compare_final(gpu, ref)
    st = [sim.stats(i) for i in range(3)]
    # stores started with capacity 1: appends halted and replayed on the device
    assert st[0]["replays"] > 0
    # the envelope bound culled part of the photons' QED work (photons
    # leaving the pulse) and everything still matches
    assert st[2]["hot"] < st[2]["processed"]
    # photons behind the pulse were pushed several steps at once, bit-exactly
    assert st[2]["slept"] > 0 and st[2]["sleeps"] > 0


def test_head_on_rs_chunked_equals_single():
    """start() called in several chunks (prune in between) replays the same draws."""
    case = head_on_case(n=1500, nsteps=160, seed=5)
    g1, _ = run_gpu_case(case, chunks=1)
    g4, _ = run_gpu_case(case, chunks=4)
    r4 = run_oracle_case(case, chunks=4)
    compare_final(g4, r4)
    # pruning between chunks does not change what survives
    for name in g1:
        assert g1[name]["n"] == g4[name]["n"]
        assert np.array_equal(g1[name]["id"], g4[name]["id"])
        assert np.array_equal(g1[name]["x"], g4[name]["x"])


@pytest.mark.parametrize("qed", ["rs", "od"])
def test_head_on_window_cache_bit_exact(qed):
    """Stores reserved up front, as in the benchmark: the photon visit list
    fits the window cache (state kept by list position between rebuilds and
    written back at each window's last step) -- bit for bit the oracle."""
    case = head_on_case(n=3000, nsteps=240, seed=13, qed=qed)
    gpu, sim = run_gpu_case(case, reserve=(60000, 60000, 400000))
    ref = run_oracle_case(case)
    st = sim.stats(2)
    print("counts", _names(gpu), "replays", st["replays"], "photon capacity", st["capacity"])
    assert st["capacity"] == 400000 and ref["photon"]["n"] > 1000
    compare_final(gpu, ref)


def test_head_on_od_bit_exact():
    case = head_on_case(n=3000, nsteps=240, seed=12, qed="od")
    gpu, _ = run_gpu_case(case)
    ref = run_oracle_case(case)
    print("counts", _names(gpu))
    assert ref["photon"]["n"] > 1000
    compare_final(gpu, ref)


def test_cascade_od_bit_exact():
    # a0 = 500 and pulses starting 1 ctau from the node keep the cascade at
    # ~1e4 particles (64 seeds -> ~230 e-, ~170 e+, ~7000 photons by step 1600)
    case = cascade_case(n=64, nsteps=1600, seed=2, a0=500.0, x0_ctau=1.0)
    gpu, sim = run_gpu_case(case)
    ref = run_oracle_case(case)
    st = sim.stats(2)
    print("counts", _names(gpu), "photon stats", st)
    assert ref["positron"]["n"] > 0  # pairs were created
    # the pair stores started full: conversions that found no room were
    # deferred, made after the stores grew, at the positions and ids they
    # would have had
    assert st["deferrals"] > 0
    compare_final(gpu, ref)


def test_qed_off_spin_ll_trajectories():
    """QED off (Boris + LL + T-BMT): bit-exact against the oracle, and within
    1e-10 relative of the oracle built on the platform libm."""
    case = head_on_case(n=2000, nsteps=300, seed=13, species=("e",), rr="LL", spin=True)
    gpu, _ = run_gpu_case(case)
    ref = run_oracle_case(case)
    compare_final(gpu, ref)
    ref_libm = run_oracle_case(case, libm=True)
    compare_final(gpu, ref_libm, rtol=1e-10)


@pytest.mark.parametrize("qed", ["rs", "od"])
def test_head_on_focused_pulse_bit_exact(qed):
    """QED with a focused pulse: the culling bound and the photons' sleep
    along their paths use the focused envelope; every event still matches."""
    case = head_on_case(n=3000, nsteps=240, seed=14, qed=qed)
    c = case["field"][0][1]
    case["field"] = [(3, (100.0, 2e-6, c[2], 1, 0.0, c[4], 0.8e-6, 0.3, 0.1))]
    gpu, st = run_gpu_case(case)
    ref = run_oracle_case(case)
    print("counts", _names(gpu))
    assert ref["photon"]["n"] > 1000
    compare_final(gpu, ref)


def test_focused_and_simple_fields_push():
    """Focused Gaussian and a sum of plane pulses through the fused push."""
    case = head_on_case(n=1000, nsteps=120, seed=21, species=("e",), rr="CLL")
    c = head_on_case(n=1000, nsteps=120, seed=21)["field"][0][1]
    case["field"] = [(3, (100.0, 2e-6, c[2], 1, 0.0, c[4], 0.8e-6, 0.3, 0.1)),
                     (1, (30.0, 3e-6, c[2], 2, -1, 1e-6, 0.8e-6, 0.7, 0.2))]
    gpu, _ = run_gpu_case(case)
    ref = run_oracle_case(case)
    compare_final(gpu, ref)


def test_per_op_methods_match_oracle():
    """Particles._eval_field/_calculate_chi/_photon_event/_pick_hard_photon/
    _create_photon/_push_* on the GPU against the oracle's passes."""
    import sfparticles_amd as sfa
    from sfparticles_amd import qed as Q
    from sfparticles_amd.fields import Fields
    import oracle
    from oracle.sim import OParticles, OSimulation
    from sfparticles_amd.qed.tables import load_rs_tables, load_od_tables

    Q.set_model("rs")
    case = head_on_case(n=4000, nsteps=2, seed=7)
    case["props"] = (np.zeros(4000) + 1e-7,) + tuple(case["props"][1:])
    e = sfa.Particles("e", -1, 1, N=4000, props=case["props"])
    g = sfa.Particles("g", 0, 0)
    e.set_photon(g)
    F = Fields(terms=case["field"])
    oe = OParticles("e", -1, 1, N=4000, props=[a.copy() for a in case["props"]])
    og = OParticles("g", 0, 0)
    oe.set_photon(og)
    osim = OSimulation(oe, og, dt=case["dt"], fields=case["field"], rs_tables=load_rs_tables(),
                       od_tables=load_od_tables())
    t = 40e-15
    for it in range(3):
        e._eval_field(F, t)
        e._calculate_chi()
        e._photon_event(case["dt"] * 20)
        e._pick_hard_photon(2.0)
        e._create_photon(g)
        e._push_position(case["dt"])
        osim.eval_field(oe, t)
        osim.calculate_chi(oe)
        osim.qed_event(oe, case["dt"] * 20, it, slot_base=Q.PER_OP_SLOT)
        osim.pick_hard_photon(oe, 2.0)
        osim.create_photon(oe, og)
        osim.push_position(oe, case["dt"])
        assert g.N_buffered == og.N_buffered
        for a in ("x", "ux", "chi", "Ey", "Bz", "inv_gamma"):
            assert np.array_equal(getattr(e, a)[:e.N_buffered], getattr(oe, a)[:oe.N_buffered]), a
        for a in ("x", "ux", "uy", "id"):
            assert np.array_equal(getattr(g, a)[:g.N_buffered], getattr(og, a)[:og.N_buffered]), a
    assert g.N_buffered > 0


def test_reference_pusher_tests():
    """tests/test_pusher.py of the reference, on the GPU backend functions."""
    from sfparticles_amd.particles import boris, LL_push
    from sfparticles_amd.constants import e, m_e, c, pi
    r = 1e-6
    ux0, uz0 = 100, 10
    u0 = np.sqrt(ux0 ** 2 + uz0 ** 2)
    v = ux0 / np.sqrt(1 + ux0 ** 2) * c
    T = 2 * pi * r / v
    bz = m_e * c / e * ux0 / r
    ux = np.asarray([ux0], dtype=float)
    uy = np.asarray([0], dtype=float)
    uz = np.asarray([uz0], dtype=float)
    inv_gamma = 1 / np.sqrt(1 + ux ** 2 + uy ** 2 + uz ** 2)
    z = np.zeros(1)
    Bz = np.asarray([bz], dtype=float)
    pr = np.array([False])
    for _ in range(20):
        boris(ux, uy, uz, inv_gamma, z, z, z, z, z, Bz, -e, 1, pr, T / 20)
    assert abs(np.sqrt(ux ** 2 + uy ** 2 + uz ** 2) - u0) / u0 < 1e-10

    ux = np.asarray([1000.0])
    uy = np.zeros(1)
    uz = np.zeros(1)
    inv_gamma = 1 / np.sqrt(1 + ux ** 2)
    LL_push(ux, uy, uz, inv_gamma, np.asarray([1.0]), 1, pr, 2.67e-17)
    assert ux[0] < 1000.0
    assert round((1000.0 - ux[0]) / 1000.0 - 0.1, 2) == 0


def test_pair_bookkeeping_and_momentum():
    """tests/test_qed.py TestPairNumber: bookkeeping and momentum conservation."""
    import sfparticles_amd as sfa
    from sfparticles_amd import qed as Q
    from sfparticles_amd.fields import static_field
    from sfparticles_amd.constants import m_e, c, e, hbar
    for model in ("rs", "od"):
        Q.set_model(model)
        gamma = 10000
        ux = np.sqrt(gamma ** 2 - 1)
        N = 200_000
        pho = sfa.Particles("pho", 0, 0, N=N, props=(np.zeros(N), np.zeros(N), np.zeros(N),
                                                    np.full(N, ux), np.zeros(N), np.zeros(N)))
        ele = sfa.Particles("ele", -1, 1)
        pos = sfa.Particles("pos", 1, 1)
        pho.set_pair(ele, pos)
        pho._eval_field(static_field(Bz=m_e ** 2 * c ** 2 / e / hbar / ux), 0)
        pho._calculate_chi()
        ev, _ = pho._pair_event(1e-17)
        ev = np.array(ev[:N])
        ux_pho = np.array(pho.ux[:N])[ev]
        pho._create_pair(ele, pos)
        assert pos.N_buffered == pos.Npart == ele.Npart == ev.sum() > 0
        assert N - pho.Npart == pos.Npart
        assert (np.abs(ele.ux[:ele.N_buffered] + pos.ux[:pos.N_buffered] - ux_pho) < 1e-10).all()
        pho._pair_event(1e-17)
        pho._create_pair(ele, pos)
        assert pos.N_buffered == pos.Npart
        assert N - pho.Npart == pos.Npart


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["gaussian", "standing_wave", "focused"])
def test_envelope_bounds_are_upper_bounds(which):
    """The FP32 envelope bounds the step culls with exceed |E| + c|B| of the
    exact FP64 field (oracle): the point bound at the point, the path bound at
    every sampled point (x + v s, t + s), s in [0, T], of a photon's straight
    path -- and neither is vacuous."""
    import oracle
    from sfparticles_amd import _lib as L
    from sfparticles_amd.fields import terms_to_c
    C_LIGHT = 299792458.0
    rng = np.random.Generator(np.random.PCG64(7))
    if which in ("gaussian", "focused"):
        case = head_on_case(n=1, nsteps=3000)
        if which == "focused":  # at its focus at t0, polarisation and CEP not 0
            c = case["field"][0][1]
            case["field"] = [(3, (100.0, 2e-6, c[2], 1, 0.0, c[4], 0.8e-6, 0.3, 0.1))]
    else:
        case = cascade_case(n=1, nsteps=100, a0=500.0, x0_ctau=1.0)
    terms = case["field"]
    tc = terms_to_c(terms)
    m = 4000
    t0 = 1500 * case["dt"] if which != "standing_wave" else 0.0
    x = rng.uniform(-20e-6, 20e-6, m)
    y = rng.normal(0, 2e-6, m)
    z = rng.normal(0, 2e-6, m)
    d = rng.normal(size=(3, m))
    d[0] += np.where(rng.random(m) < 0.5, -3.0, 3.0)  # mostly along the axis, both ways
    d /= np.linalg.norm(d, axis=0)
    v = C_LIGHT * d

    def bound(T, path):
        dev = [L.DeviceArray.of(np.ascontiguousarray(a)) for a in (x, y, z, v[0], v[1], v[2])]
        out = L.DeviceArray(4 * m)
        vel = [a.ptr for a in dev[3:]] if path else [None, None, None]
        L.call("sfp_field_bound_eval", tc, len(terms), dev[0].ptr, dev[1].ptr, dev[2].ptr,
               float(t0), *vel, float(T), m, out.ptr, None)
        b = np.empty(m, np.float32)
        out.download(b)
        return b.astype(np.float64)

    def mag(s):
        F = oracle.eval_field(terms, x + v[0] * s, y + v[1] * s, z + v[2] * s, t0 + s)
        return (np.sqrt(F[0] ** 2 + F[1] ** 2 + F[2] ** 2)
                + C_LIGHT * np.sqrt(F[3] ** 2 + F[4] ** 2 + F[5] ** 2))

    b0 = bound(0.0, False)
    m0 = mag(0.0)
    assert np.all(b0 >= m0), np.max(m0 - b0)
    print(which, "point: max field/bound", np.max(m0 / b0))
    assert np.max(m0 / b0) > 0.05
    for nsteps in (8, 64, 512, 3000):
        T = nsteps * case["dt"]
        b = bound(T, True)
        assert np.all(b >= b0 * 0.999999)
        worst = 0.0
        for s in np.concatenate([[0.0, T], rng.random(10) * T]):
            mm = mag(s)
            assert np.all(b >= mm), (nsteps, s, np.max(mm - b))
            worst = max(worst, float(np.max(mm / b)))
        print(which, nsteps, "path: max field/bound", worst)
        assert worst > 0.05


@pytest.mark.parametrize("qed,graph", [("rs", False), ("od", False), ("rs", True)])
def test_photon_sleep_carries_across_run_calls(qed, graph):
    """Run calls that continue each other keep QED photons asleep across the
    call boundary (no restart of the sleep clock, no visit of every photon
    at the first step of a call) and keep the photon visit lists and window
    caches; positions read between calls are those of that step, and the end
    state still equals the oracle's bit for bit (the optical depth of
    sleeping photons included; also with graph replays)."""
    import sfparticles_amd as sfa
    from sfparticles_amd import qed as Q
    from sfparticles_amd.fields import Fields
    case = head_on_case(n=3000, nsteps=240, seed=11, qed=qed)
    Q.set_model(qed)

    def make():
        n = case["props"][0].shape[0]
        e = sfa.Particles("electron", -1, 1, N=n, props=[a.copy() for a in case["props"]],
                          RR=sfa.RadiationReactionType.PHOTON)
        p = sfa.Particles("positron", 1, 1)
        g = sfa.Particles("photon", 0, 0)
        e.set_photon(g)
        p.set_photon(g)
        g.set_pair(e, p)
        sim = sfa.Simulation(e, p, g, dt=case["dt"], fields=Fields(terms=case["field"]),
                             print_every=None, photon_threshold=case["threshold"],
                             seed=case["seed"], graph=graph)
        # the photon store reserved up front: a window cache sized once (one
        # grown at the start of a call rebuilds the lists there)
        g._ensure_device()
        g._extend(1 << 20)
        return sim, (e, p, g)

    # calls of uneven length, none a multiple of the 16-step list window; the
    # last one through start(), which prunes the converted photons at its end
    sim, sp = make()

    def photons(g):
        n = g.N_buffered
        live = ~np.array(g._to_be_pruned)[:n]  # converted photons are holes until pruned
        return {a: np.array(getattr(g, a)[:n])[live] for a in ("x", "y", "z", "id")}

    for m in (37, 64, 83) if not graph else (37, 128, 19):  # graph replays need 64-step batches
        sim.run_steps(m)
    mid = photons(sp[2])
    # calls of one step, and one that ends the run 5 steps before the clock's
    # next rebuild, keep the visit lists and window caches of the call before
    # (the first of them across the host read above, whose catch-up leaves
    # them be); the longer calls before rebuild theirs at their first step
    for _ in range(3):
        sim.run_steps(1)
    sim.start(53)
    st = sim.stats(2)
    print("clock restarts", st["clock_restarts"], "slept", st["slept"], "sleeps", st["sleeps"],
          "visit lists kept", st["list_keeps"])
    assert st["clock_restarts"] == 1 and st["slept"] > 0
    assert st["list_keeps"] == 4
    # positions between calls: a single call stopped at the same step, read
    # after its simulation is gone (which brings sleeping photons up to date)
    sim1, sp1 = make()
    sim1.run_steps(184)
    del sim1
    ref_mid = photons(sp1[2])
    for a in ("x", "y", "z", "id"):
        assert np.array_equal(mid[a], ref_mid[a]), a
    got = {}
    for s_ in sp:
        d = {"n": int(s_.N_buffered)}
        for a in ("x", "y", "z", "ux", "uy", "uz", "id") + (("tau",) if "tau" in s_.attrs else ()):
            d[a] = np.array(getattr(s_, a)[: s_.N_buffered])
        got[s_.name] = d
    if graph:
        assert sim.stats(0)["graph_batches"] >= 2
    compare_final(got, run_oracle_case(case))
    # a change of a store from the host (here the prune) restarts the clock
    # and rebuilds the lists
    sim.run_steps(8)
    assert sim.stats(2)["clock_restarts"] == 2 and sim.stats(2)["list_keeps"] == 4
