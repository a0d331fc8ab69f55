"""Mixed precision against FP64, per observable.

The mixed step evaluates fields, pushers and QED sampling in FP32 and keeps
positions, time and chi in FP64.  Its uniforms are the leading 24 bits of the
same Philox words as the FP64 step, so the two runs make the same decisions
except near a threshold, and observables can be compared run against run.

Bars (written in each test): QED off, the electron momentum within 1e-4
relative, the position within 1e-3 of the distance travelled in one step and
the spin within 1e-3 of the FP64 oracle after 300 steps; QED on, photon and
pair yields within 1% (or 3 sigma of the FP64 count, whichever is larger),
the photon energy spectra indistinguishable by a two-sample KS test
(p > 0.01) and the mean electron energy within 1e-3; cascades, mean counts
over six seeds within three standard errors.  With the learned photon-energy
sampler the spectrum bar is the spectrum error, the KS distance, within 1%.
"""
import os

import numpy as np
import pytest
from scipy import stats

from tests.cases import DT, head_on_case, cascade_case, run_gpu_case, run_oracle_case

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def _by_id(d):
    o = np.argsort(d["id"])
    return {k: (v[o] if isinstance(v, np.ndarray) else v) for k, v in d.items()}


def test_mixed_qed_off_trajectories():
    case = head_on_case(n=2000, nsteps=300, seed=13, species=("e",), rr="LL", spin=True)
    mix, _ = run_gpu_case(case, precision="mixed")
    ref = run_oracle_case(case)
    m, r = _by_id(mix["electron"]), _by_id(ref["electron"])
    assert np.array_equal(m["id"], r["id"])
    um = np.stack([m["ux"], m["uy"], m["uz"]])
    ur = np.stack([r["ux"], r["uy"], r["uz"]])
    du = np.max(np.linalg.norm(um - ur, axis=0) / np.linalg.norm(ur, axis=0))
    dx = np.max(np.linalg.norm(np.stack([m["x"] - r["x"], m["y"] - r["y"], m["z"] - r["z"]]), axis=0))
    ds = np.max(np.linalg.norm(np.stack([m["sx"] - r["sx"], m["sy"] - r["sy"], m["sz"] - r["sz"]]),
                               axis=0))
    step = C_LIGHT * DT
    print(f"QED off, 300 steps: max |du|/|u| {du:.3e}, max |dx| {dx:.3e} m "
          f"({dx / step:.3e} of c dt), max |ds| {ds:.3e}")
    assert du < 1e-4
    assert dx < 1e-3 * step
    assert ds < 1e-3


def _compare_qed(case, what):
    f64, _ = run_gpu_case(case)
    mix, _ = run_gpu_case(case, precision="mixed")
    rows = []
    for name in ("photon", "positron"):
        n64, nm = f64[name]["n"], mix[name]["n"]
        tol = max(0.01 * n64, 3 * np.sqrt(max(n64, 1)))
        rows.append((f"{name} yield", n64, nm, (nm - n64) / max(n64, 1)))
        assert abs(nm - n64) <= tol, (what, name, n64, nm)
    e64 = np.sqrt(1 + f64["photon"]["ux"] ** 2 + f64["photon"]["uy"] ** 2 + f64["photon"]["uz"] ** 2)
    em = np.sqrt(1 + mix["photon"]["ux"] ** 2 + mix["photon"]["uy"] ** 2 + mix["photon"]["uz"] ** 2)
    ks = stats.ks_2samp(e64, em)
    rows.append(("photon spectrum KS", len(e64), len(em), ks.pvalue))
    assert ks.pvalue > 0.01, (what, ks)
    g64 = np.mean(np.sqrt(1 + f64["electron"]["ux"] ** 2 + f64["electron"]["uy"] ** 2
                          + f64["electron"]["uz"] ** 2))
    gm = np.mean(np.sqrt(1 + mix["electron"]["ux"] ** 2 + mix["electron"]["uy"] ** 2
                         + mix["electron"]["uz"] ** 2))
    rows.append(("mean electron gamma", g64, gm, (gm - g64) / g64))
    assert abs(gm - g64) < 1e-3 * g64
    for r in rows:
        print(what, r)


def _sampler():
    """The shipped network if present, else one trained here on the GPU."""
    from sfparticles_amd.qed import mlp
    if os.path.exists(mlp.DEFAULT_PATH):
        return mlp.default_sampler()
    s = mlp.PhotonSampler(seed=0)
    s.fit(n_data=2_000_000, steps=3000, batch=4096, lr=3e-3, seed=0)
    s.calibrate_fallback(per_bin=20000)
    return s


def _compare_net(case, what):
    """mixed precision with the learned photon-energy sampler against FP64:
    photon and pair yields within 1% (or 3 sigma), photon spectrum error (KS
    distance, the largest difference of the cumulative spectra) within 1%."""
    f64, _ = run_gpu_case(case)
    net, _ = run_gpu_case(case, precision="mixed", sampler=_sampler())
    for name in ("photon", "positron"):
        n64, nm = f64[name]["n"], net[name]["n"]
        tol = max(0.01 * n64, 3 * np.sqrt(max(n64, 1)))
        print(what, name, "yield", n64, nm, (nm - n64) / max(n64, 1))
        assert abs(nm - n64) <= tol, (what, name, n64, nm)
    e64 = np.sqrt(f64["photon"]["ux"] ** 2 + f64["photon"]["uy"] ** 2 + f64["photon"]["uz"] ** 2)
    em = np.sqrt(net["photon"]["ux"] ** 2 + net["photon"]["uy"] ** 2 + net["photon"]["uz"] ** 2)
    ks = stats.ks_2samp(e64, em)
    print(what, "photon spectrum KS distance", ks.statistic, "p", ks.pvalue)
    # spectrum error: largest difference of the two cumulative spectra <= 1%
    assert ks.statistic <= 0.01, (what, ks)


def test_mixed_mlp_head_on_rs():
    _compare_net(head_on_case(n=40000, nsteps=240, seed=11, a0=250.0), "head-on RS + network")


def test_mixed_mlp_head_on_od():
    _compare_net(head_on_case(n=40000, nsteps=240, seed=12, a0=250.0, qed="od"),
                 "head-on OD + network")


def test_mixed_head_on_rs_yields_and_spectrum():
    _compare_qed(head_on_case(n=40000, nsteps=240, seed=11, a0=250.0), "head-on RS")


def test_mixed_head_on_od_yields_and_spectrum():
    _compare_qed(head_on_case(n=40000, nsteps=240, seed=12, a0=250.0, qed="od"), "head-on OD")


def test_mixed_cascade_od_counts():
    """A cascade is a branching process: one decision that flips early
    changes a whole subtree, so single runs are not comparable.  Over six
    Philox seeds, the mean counts of the two precisions agree within three
    standard errors of the seed-to-seed spread."""
    counts = {"fp64": [], "mixed": []}
    for seed in range(6):
        case = cascade_case(n=64, nsteps=1600, seed=2, a0=500.0, x0_ctau=1.0, philox_seed=seed)
        for prec in counts:
            out, _ = run_gpu_case(case, precision=prec)
            counts[prec].append([out[k]["n"] for k in ("electron", "positron", "photon")])
    c64 = np.array(counts["fp64"], float)
    cm = np.array(counts["mixed"], float)
    for j, name in enumerate(("electron", "positron", "photon")):
        se = np.sqrt(c64[:, j].var(ddof=1) / 6 + cm[:, j].var(ddof=1) / 6)
        d = cm[:, j].mean() - c64[:, j].mean()
        print("cascade", name, c64[:, j].mean(), cm[:, j].mean(), "diff", d, "se", se)
        assert abs(d) <= 3 * se + 1.0, (name, d, se)
