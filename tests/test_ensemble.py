"""Ensembles of shots, graph-captured step batches and the launch-shape tuner.

None of these changes what a step computes, so the GPU tests hold them to
bit-exact agreement with plain runs:
  * an ensemble's seed electrons end bit for bit where the same electrons end
    when their shot runs alone (QED on: draws are keyed by particle id, and
    an electron's own draws and field fix its trajectory);
  * per-shot photon and pair yields of the ensemble agree with the separate
    runs within their statistical spread (created particles get other ids in
    the ensemble, hence other draws);
  * every species of a run replayed as captured graphs, with the online
    tuner (grid and tile rows), or with pinned small grids and tiles of few
    rows is bit-identical to the eager run, down to the order and ids of the
    created particles;
  * the on-device per-shot histogram against numpy on the downloaded arrays.
CPU: the ensemble generator's draw order and the host checks.
"""
import numpy as np
import pytest

from tests.cases import head_on_case, run_gpu_case, compare_final


def test_head_on_shots_draw_order():
    from sfparticles_amd.ensemble import head_on_shots
    fields, props, shot, table = head_on_shots(5, 1000, seed=3, count=(10, 100))
    rng = np.random.Generator(np.random.PCG64(3))
    a0 = rng.uniform(20, 500, 5)
    ns = np.rint(10.0 ** rng.uniform(1, 2, 5)).astype(np.int64)
    gs = rng.uniform(500, 5000, 5)
    assert np.array_equal(table["a0"], a0) and np.array_equal(table["n"], ns)
    g0 = rng.normal(gs[0], 0.05 * gs[0], ns[0])
    assert np.array_equal(props[3][:ns[0]], -np.sqrt(g0 ** 2 - 1))
    assert np.array_equal(np.bincount(shot, minlength=5), ns)
    assert len(fields) == 5 and fields[2].terms[0][1][0] == a0[2]


def test_shot_index_without_field_is_rejected():
    import sfparticles_amd as sfa
    from sfparticles_amd.fields import Fields
    e = sfa.Particles("electron", -1, 1, N=4, props=tuple(np.ones(4) for _ in range(6)))
    e.shot = np.array([0, 1, 2, 1], np.uint32)
    f = Fields(terms=[(0, (0, 0, 0, 0, 0, 1.0))])
    with pytest.raises(ValueError, match="shot index 2"):
        sfa.Simulation(e, dt=1e-17, fields=None, shots=[f, f], print_every=None)


def _ensemble(fields, props, shot, nsteps, qed=True, precision="fp64", ids=None, **kw):
    import sfparticles_amd as sfa
    from sfparticles_amd import qed as Q
    Q.set_model("rs")
    n = props[0].size
    rr = sfa.RadiationReactionType.PHOTON if qed else sfa.RadiationReactionType.LL
    e = sfa.Particles("electron", -1, 1, N=n, props=props, RR=rr)
    e.shot = shot
    if ids is not None:
        e.id = ids
        e.next_id = int(ids.max()) + 1 if ids.size else 0
    sp = [e]
    if qed:
        p = sfa.Particles("positron", 1, 1)
        g = sfa.Particles("photon", 0, 0)
        e.set_photon(g)
        p.set_photon(g)
        g.set_pair(e, p)
        sp += [p, g]
    sim = sfa.Simulation(*sp, dt=_dt(), fields=None, shots=fields, print_every=None,
                         photon_threshold=2.0, seed=0, precision=precision, **kw)
    sim.start(nsteps)
    return sim, sp


def _dt():
    from sfparticles_amd.ensemble import DT
    return DT


@pytest.mark.gpu
@pytest.mark.parametrize("qed", [False, True])
def test_ensemble_seed_electrons_match_separate_runs(qed):
    from sfparticles_amd.ensemble import head_on_shots, shot_histogram
    nsteps = 600
    fields, props, shot, table = head_on_shots(4, nsteps, seed=1, count=(500, 4000),
                                               a0=(50, 300), gamma=(1000, 3000))
    ids = np.arange(shot.size, dtype=np.uint64)
    sim, sp = _ensemble(fields, props, shot, nsteps, qed=qed, ids=ids)
    e = sp[0]
    off = np.concatenate([[0], np.cumsum(table["n"])])
    got = {a: np.array(getattr(e, a)) for a in ("x", "y", "z", "ux", "uy", "uz", "id")}
    counts = [shot_histogram(s, 4)[0] for s in sp]
    for k in range(4):
        sl = slice(off[k], off[k + 1])
        sub = tuple(np.ascontiguousarray(a[sl]) for a in props)
        s1, sp1 = _ensemble([fields[k]], sub, np.zeros(sub[0].size, np.uint32), nsteps, qed=qed,
                            ids=ids[sl].copy())
        e1 = sp1[0]
        # the seed electrons keep their slots (appends go after them)
        for a in ("id", "x", "y", "z", "ux", "uy", "uz"):
            want = np.array(getattr(e1, a))[: table["n"][k]]
            assert np.array_equal(got[a][sl], want), (k, a)
        if qed:
            names = ("electron", "positron", "photon")
            for j, s in enumerate(sp1):
                alone = s.N_buffered - (int(table["n"][k]) if j == 0 else 0)
                ens = int(counts[j][k]) - (int(table["n"][k]) if j == 0 else 0)
                tol = 4.0 * np.sqrt(2.0 * max(alone, 1)) + 3
                print(f"shot {k} a0 {table['a0'][k]:.0f} {names[j]}: created alone {alone}, "
                      f"in the ensemble {ens}")
                assert abs(ens - alone) <= tol, (k, names[j], ens, alone)
        del s1, sp1


@pytest.mark.gpu
def test_graph_autotune_and_grids_are_bit_identical():
    case = head_on_case(20000, 640, seed=4, a0=150.0)
    ref, _ = run_gpu_case(case)
    for kw in (dict(graph=True), dict(autotune=True), dict(graph=True, autotune=True),
               dict(grids={(0, 0): 256, (0, 1): 256, (2, 0): 256, (2, 1): 512}),
               dict(grids={(0, 0): (512, 4), (0, 1): (256, 3), (1, 0): (768, 1),
                           (1, 1): (256, 1), (2, 1): (512, 2)})):
        got, sim = run_gpu_case(case, **kw)
        st = sim.stats(0)
        print(kw, "graph batches", st["graph_batches"], "captures", st["captures"],
              "tuning batches", st["tune_batches"],
              "shapes", [sim.launch_shape(s, ph) for s in range(3) for ph in (0, 1)])
        if kw.get("graph"):
            assert st["graph_batches"] >= 5
        if kw.get("autotune"):
            assert st["tune_batches"] >= 1
            # the tuner's candidates: grids times 3 tile-row caps for the
            # electron sweep and QED kernels
            for ph, rmax in ((0, 32), (1, 16)):
                g, t, r, n = sim.launch_shape(0, ph)
                assert n >= 3 and n % 3 == 0 and r in (rmax, rmax // 2, rmax // 4)
        if "grids" in kw and isinstance(kw["grids"][(0, 0)], tuple):
            assert sim.launch_shape(0, 0)[:3] == (512, 256, 4)
            assert sim.launch_shape(0, 1)[:3] == (256, 256, 3)
        compare_final(got, ref)
        del sim


@pytest.mark.gpu
def test_launch_shape_pins_are_checked():
    from sfparticles_amd._lib import SFPError
    case = head_on_case(2000, 8, seed=4, a0=150.0)
    _, sim = run_gpu_case(case)
    with pytest.raises(SFPError):
        sim.set_shape(0, 0, 256, 33)
    with pytest.raises(SFPError):
        sim.set_shape(0, 1, 0, 4)
    sim.set_shape(0, 1, 256, 32)  # clamped to the QED kernel's 16 rows
    assert sim.launch_shape(0, 1)[:4] == (256, 256, 16, 1)
    sim.set_shape(0, 1, 0)
    assert sim.launch_shape(0, 1)[3] > 1


@pytest.mark.gpu
def test_graph_replays_across_a_capacity_halt():
    """Stores that start too small halt inside a graph replay; the replay
    from the halted kernel (eager) and the later graph replays still give the
    eager result."""
    case = head_on_case(20000, 640, seed=5, a0=200.0)
    ref, _ = run_gpu_case(case)
    got, sim = run_gpu_case(case, graph=True)
    st = sim.stats(0)
    print("replays", st["replays"], "graph batches", st["graph_batches"])
    assert st["replays"] >= 1 and st["graph_batches"] >= 5
    compare_final(got, ref)


@pytest.mark.gpu
def test_shot_histogram_matches_numpy():
    from sfparticles_amd.ensemble import head_on_shots, shot_histogram
    fields, props, shot, table = head_on_shots(6, 600, seed=2, count=(300, 3000))
    sim, sp = _ensemble(fields, props, shot, 600, qed=True)
    for s in sp:
        n = s.N_buffered
        live = ~np.array(s._to_be_pruned)[:n]
        sh = np.array(s.shot)[:n][live]
        u2 = (np.array(s.ux)[:n] ** 2 + np.array(s.uy)[:n] ** 2 + np.array(s.uz)[:n] ** 2)[live]
        en = np.sqrt(u2) if s.m == 0 else np.sqrt(1 + u2)
        cnt, hist = shot_histogram(s, 6, nbins=40, log10_range=(0.0, 4.0))
        assert np.array_equal(cnt, np.bincount(sh, minlength=6)), s.name
        b = (np.log10(en) - 0.0) * (40 / 4.0)
        ok = (b >= 0) & (b < 40)
        want = np.zeros((6, 40), np.int64)
        np.add.at(want, (sh[ok], b[ok].astype(np.int64)), 1)
        assert np.array_equal(hist, want), s.name
        # large shot counts take the global-atomics path
        cnt2, _ = shot_histogram(s, 5000)
        assert np.array_equal(cnt2[:6], cnt) and not cnt2[6:].any()
