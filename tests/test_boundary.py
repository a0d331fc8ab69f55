"""The drop-in boundary, callable by callable: every function of the
reference's backend seam (sfparticles/cpu.py:5-232, the 12 callables
particles.py:14-25 imports) and of its QED seam
(qed/rejection_sampling.py:68-101, qed/optical_depth.py:15-110), called
through ``sfparticles_amd.backend`` with numpy arrays exactly as the
reference's ``Particles`` calls them, against the oracle's pass on the same
arrays -- bit for bit.  Inputs: random momenta, fields, chi, events and
deltas with a tenth of the particles pruned."""
import ctypes as C

import numpy as np
import pytest

import oracle
from oracle import ptr

pytestmark = pytest.mark.gpu

N = 20000
DT = 2.6685e-17
Q_E = -1.602176634e-19


def _inputs(seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    d = {}
    for a in ("x", "y", "z"):
        d[a] = rng.normal(0, 2e-6, N)
    u = rng.normal(0, 1, (3, N)) * np.exp(rng.uniform(0, 9, N))
    d["ux"], d["uy"], d["uz"] = u[0].copy(), u[1].copy(), u[2].copy()
    d["inv_gamma"] = 1 / np.sqrt(1 + d["ux"] ** 2 + d["uy"] ** 2 + d["uz"] ** 2)
    s = rng.normal(size=(3, N))
    s /= np.linalg.norm(s, axis=0)
    d["sx"], d["sy"], d["sz"] = s[0].copy(), s[1].copy(), s[2].copy()
    for a in ("Ex", "Ey", "Ez"):
        d[a] = rng.normal(0, 4e13, N)
    for a in ("Bx", "By", "Bz"):
        d[a] = rng.normal(0, 1.3e5, N)
    d["pruned"] = rng.random(N) < 0.1
    d["chi"] = np.exp(rng.uniform(np.log(1e-4), np.log(50.0), N))
    d["event"] = rng.random(N) < 0.3
    d["delta"] = rng.random(N)
    d["tau"] = np.where(rng.random(N) < 0.5, 0.0, rng.exponential(1.0, N) * 1e-3)
    d["ids"] = rng.permutation(N).astype(np.uint64) + np.uint64(1 << 33)
    return d


def _copy(d):
    return {k: v.copy() for k, v in d.items()}


def _u8(a):
    return np.ascontiguousarray(a.astype(np.uint8))


def _same(got, want, names):
    for k in names:
        g, w = np.asarray(got[k]), np.asarray(want[k])
        if g.dtype == np.bool_ or w.dtype == np.bool_:
            g, w = g.astype(np.uint8), w.astype(np.uint8)
        bad = ~((g == w) | (np.isnan(g) & np.isnan(w)) if g.dtype.kind == "f" else (g == w))
        assert not bad.any(), f"{k}: {bad.sum()} of {bad.size} differ, e.g. {g[bad][:3]} vs {w[bad][:3]}"


def test_pushers_and_chi():
    """push_position, boris, boris_tbmt, LL_push, CLL_push, update_chi."""
    from sfparticles_amd import backend as B
    L = oracle.lib()
    d0 = _inputs()
    g, o = _copy(d0), _copy(d0)
    pr = _u8(d0["pruned"])

    B.push_position(g["x"], g["y"], g["z"], g["ux"], g["uy"], g["uz"], g["inv_gamma"], N,
                    g["pruned"], DT)
    L.orc_push_position(*[ptr(o[a]) for a in ("x", "y", "z", "ux", "uy", "uz", "inv_gamma")], N,
                        ptr(pr), DT)
    _same(g, o, ("x", "y", "z"))
    assert not np.array_equal(g["x"], d0["x"])

    F = ("Ex", "Ey", "Ez", "Bx", "By", "Bz")
    B.boris(g["ux"], g["uy"], g["uz"], g["inv_gamma"], *[g[a] for a in F], Q_E, N, g["pruned"], DT)
    L.orc_boris(*[ptr(o[a]) for a in ("ux", "uy", "uz", "inv_gamma") + F], Q_E, N, ptr(pr), DT)
    _same(g, o, ("ux", "uy", "uz", "inv_gamma"))

    B.boris_tbmt(*[g[a] for a in ("ux", "uy", "uz", "inv_gamma", "sx", "sy", "sz") + F], Q_E,
                 1.14e-3, N, g["pruned"], DT)
    L.orc_boris_tbmt(*[ptr(o[a]) for a in ("ux", "uy", "uz", "inv_gamma", "sx", "sy", "sz") + F],
                     Q_E, 1.14e-3, N, ptr(pr), DT)
    _same(g, o, ("ux", "uy", "uz", "inv_gamma", "sx", "sy", "sz"))

    B.update_chi(*[g[a] for a in F + ("ux", "uy", "uz", "inv_gamma", "chi")], N, g["pruned"])
    L.orc_update_chi(*[ptr(o[a]) for a in F + ("ux", "uy", "uz", "inv_gamma", "chi")], N, ptr(pr))
    _same(g, o, ("chi",))
    assert np.all(g["chi"][~d0["pruned"]] != d0["chi"][~d0["pruned"]])

    for fn, orc in ((B.LL_push, L.orc_LL_push), (B.CLL_push, L.orc_CLL_push)):
        fn(g["ux"], g["uy"], g["uz"], g["inv_gamma"], g["chi"], N, g["pruned"], DT)
        orc(*[ptr(o[a]) for a in ("ux", "uy", "uz", "inv_gamma", "chi")], N, ptr(pr), DT)
        _same(g, o, ("ux", "uy", "uz", "inv_gamma"))
    # pruned particles are untouched everywhere
    for a in ("x", "ux", "sx", "chi"):
        assert np.array_equal(g[a][d0["pruned"]], d0[a][d0["pruned"]])


def test_events_recoil_and_counts():
    """pick_hard_photon, photon_recoil, bool_sum, find_event_index."""
    from sfparticles_amd import backend as B
    L = oracle.lib()
    d0 = _inputs(4)
    g, o = _copy(d0), _copy(d0)
    o_event = _u8(d0["event"])
    B.pick_hard_photon(g["event"], g["delta"], g["inv_gamma"], 2.0, N)
    L.orc_pick_hard_photon(ptr(o_event), ptr(o["delta"]), ptr(o["inv_gamma"]), 2.0, N)
    assert np.array_equal(g["event"].astype(np.uint8), o_event)
    assert 0 < g["event"].sum() < d0["event"].sum()

    B.photon_recoil(g["ux"], g["uy"], g["uz"], g["inv_gamma"], g["event"], g["delta"], N,
                    g["pruned"])
    o_pruned = _u8(d0["pruned"])  # held: the oracle reads it through a raw pointer
    L.orc_photon_recoil(ptr(o["ux"]), ptr(o["uy"]), ptr(o["uz"]), ptr(o["inv_gamma"]),
                        ptr(o_event), ptr(o["delta"]), N, ptr(o_pruned))
    _same(g, o, ("ux", "uy", "uz", "inv_gamma"))

    ev8 = _u8(g["event"])
    for n in (0, 1, 63, 64, 65, 4097, N):
        want = int(L.orc_bool_sum(ptr(ev8), n)) if n else 0
        assert B.bool_sum(g["event"], n) == want
        idx = np.zeros(max(want, 1), np.int64)
        cnt = int(L.orc_find_event_index(ptr(ev8), want, n, ptr(idx))) if n else 0
        got = B.find_event_index(g["event"], want, n)
        assert cnt == want and np.array_equal(got, idx[:want])


def test_create_photon_and_pair():
    """create_photon and create_pair (both halves: delta for the electron,
    1 - delta for the positron) into buffers that already hold particles."""
    from sfparticles_amd import backend as B
    L = oracle.lib()
    d0 = _inputs(5)
    ev = d0["event"] & ~d0["pruned"]
    idx = np.flatnonzero(ev).astype(np.int64)
    n_ev = idx.size
    nbuf = 1000
    cap = nbuf + n_ev + 7

    def dst(seed):
        r = np.random.Generator(np.random.PCG64(seed))
        return {a: r.normal(size=cap) for a in ("x", "y", "z", "ux", "uy", "uz", "inv_gamma")} | {
            "pruned": r.random(cap) < 0.5}

    src = ("x", "y", "z", "ux", "uy", "uz")
    outs = ("x", "y", "z", "ux", "uy", "uz", "inv_gamma")
    gd, od = dst(1), dst(1)
    B.create_photon(*[d0[a] for a in src], *[gd[a] for a in outs], gd["pruned"], idx, d0["delta"],
                    nbuf, n_ev)
    opr = _u8(od["pruned"])
    ids = np.zeros(cap, np.uint64)
    L.orc_create_photon(*[ptr(d0[a]) for a in src], *[ptr(od[a]) for a in outs], ptr(opr),
                        ptr(idx), ptr(d0["delta"]), nbuf, n_ev, ptr(ids), 0, None, None)
    od["pruned"] = opr
    _same(gd, od, outs + ("pruned",))

    for inverse in (0, 1):
        gs, os_ = d0["pruned"].copy(), _u8(d0["pruned"])
        gd, od = dst(2), dst(2)
        B.create_pair(*[d0[a] for a in src], gs, *[gd[a] for a in outs], gd["pruned"], idx,
                      d0["delta"], nbuf, n_ev, inverse)
        opr = _u8(od["pruned"])
        L.orc_create_pair(*[ptr(d0[a]) for a in src], ptr(os_), *[ptr(od[a]) for a in outs],
                          ptr(opr), ptr(idx), ptr(d0["delta"]), nbuf, n_ev, inverse,
                          ptr(ids), 0, None, None, None, None, None)
        od["pruned"] = opr
        _same(gd, od, outs + ("pruned",))
        assert np.array_equal(gs.astype(np.uint8), os_) and gs[idx].all()


@pytest.mark.parametrize("species", ["photon", "pair"])
def test_qed_seam(species):
    """photon_from_rejection_sampling / pair_from_rejection_sampling and
    update_tau_e / update_tau_gamma, with the same Philox stream (ids, seed,
    step, tag, slot) on both sides; tau carried over two calls."""
    from sfparticles_amd import backend as B
    from sfparticles_amd.qed.tables import load_rs_tables, load_od_tables
    from sfparticles_amd import qed as Q
    L = oracle.lib()
    rs, od = oracle.rs_struct(load_rs_tables()), oracle.od_struct(load_od_tables())
    d0 = _inputs(6)
    if species == "photon":
        ig = d0["inv_gamma"]
        rs_g, rs_o = B.photon_from_rejection_sampling, L.orc_photon_from_rejection_sampling
        od_g, od_o = B.update_tau_e, L.orc_update_tau_e
    else:  # photons: 1/|u|
        ig = 1 / np.sqrt(d0["ux"] ** 2 + d0["uy"] ** 2 + d0["uz"] ** 2)
        rs_g, rs_o = B.pair_from_rejection_sampling, L.orc_pair_from_rejection_sampling
        od_g, od_o = B.update_tau_gamma, L.orc_update_tau_gamma
    dt = 40 * DT  # events in a good fraction of the particles
    pr = _u8(d0["pruned"])
    for step in (0, 7):
        ge, gd = np.zeros(N, bool), np.zeros(N)
        oe, odl = np.zeros(N, np.uint8), np.zeros(N)
        rs_g(ig, d0["chi"], dt, N, d0["pruned"], ge, gd, ids=d0["ids"], seed=5, step=step, tag=2)
        rs_o(ptr(ig), ptr(d0["chi"]), dt, N, ptr(pr), ptr(oe), ptr(odl), ptr(d0["ids"]), 5, step,
             2, Q.PER_OP_SLOT, C.byref(rs))
        assert np.array_equal(ge.astype(np.uint8), oe) and np.array_equal(gd, odl)
        assert 0 < ge.sum() < N
    gt, ot = d0["tau"].copy(), d0["tau"].copy()
    for step in (0, 1):
        ge, gd = np.zeros(N, bool), np.zeros(N)
        oe, odl = np.zeros(N, np.uint8), np.zeros(N)
        od_g(gt, ig, d0["chi"], dt, N, d0["pruned"], ge, gd, ids=d0["ids"], seed=5, step=step,
             tag=2)
        od_o(ptr(ot), ptr(ig), ptr(d0["chi"]), dt, N, ptr(pr), ptr(oe), ptr(odl), ptr(d0["ids"]), 5,
             step, 2, Q.PER_OP_SLOT, C.byref(od))
        assert np.array_equal(ge.astype(np.uint8), oe) and np.array_equal(gd, odl)
        assert np.array_equal(gt, ot)
        assert 0 < ge.sum() < N
