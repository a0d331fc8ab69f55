"""The benchmark's kernel-trace analysis (bench.py --summarize-profile): steps
and timed windows found in a dispatch sequence, and the species tags of the
per-species kernels, for the three launch layouts a step can have -- shared
lepton launches (one sweep launch, k_sweep_pair, and one QED launch for both
species) with the count snapshot in the sweep (the default), one launch of
each kernel per lepton species, and a step opened by the prologue launch."""
import bench

SW = "k_sweep<false, true, 1, true, false>"
SWP = "k_sweep_pair<1, true, false>"
HOT = "k_hot<0, 1, true, 1, false>"
EMIT = "k_emit<0, 1, false>"
PH = ["k_awake", "k_photon<0, 1, true, false>"]


def _seq(step_kernels, nsteps, settles=True):
    seq, t = [], 0

    def add(name):
        nonlocal t
        seq.append((name, t, t + 1))
        t += 2
    if settles:
        add("k_catchup")
        add("k_catchup_slow")
    for _ in range(nsteps):
        for k in step_kernels:
            add(k)
    if settles:
        add("k_catchup")
        add("k_catchup_slow")
    return seq


def test_shared_launches_layout():
    seq = bench._by_species(_seq([SWP, HOT, EMIT] + PH, 3))
    names = [r[0] for r in seq]
    assert names[2:7] == [f"{SWP} [e-,e+]", f"{HOT} [e-,e+]", f"{EMIT} [e-,e+]",
                          "k_awake [photon]", "k_photon<0, 1, true, false> [photon]"]
    assert bench._step_opens(seq) == [2, 7, 12]
    wins = bench._windows(seq, 3)
    assert len(wins) == 1 and len(wins[0]) == 3 * 5 + 2
    assert bench._windows(seq, 4) == []


def test_separate_launches_layout():
    seq = bench._by_species(_seq([SW, HOT, EMIT, SW, HOT, EMIT] + PH, 2))
    names = [r[0] for r in seq]
    assert names[2:8] == [f"{SW} [e-]", f"{HOT} [e-]", f"{EMIT} [e-]",
                          f"{SW} [e+]", f"{HOT} [e+]", f"{EMIT} [e+]"]
    # the positrons' sweep follows the electrons' QED kernels: no step opens there
    assert bench._step_opens(seq) == [2, 10]
    assert len(bench._windows(seq, 2)) == 1


def test_prologue_layout():
    seq = bench._by_species(_seq(["k_step_prologue", SW, SW, HOT, EMIT] + PH, 2))
    assert bench._step_opens(seq) == [2, 9]
    assert [r[0] for r in seq][4] == f"{SW} [e+]"
    assert len(bench._windows(seq, 2)) == 1
