"""Shared launches: the electrons and positrons of a cascade in one launch
of the sweep kernel (k_sweep_pair), of the lepton QED kernel and of the
emission kernel with the learned sampler give bit for bit what one launch of
each per species gives -- the
positrons' photons appended after all of the electrons', with the same ids
-- on stores that start at the seed count and grow through capacity halts.
(The parity tests against the oracle run the shared launches, the default.)
"""
import pytest

from tests.cases import cascade_case, run_gpu_case, compare_final


@pytest.mark.gpu
@pytest.mark.parametrize("qed,precision,sampler", [
    ("od", "fp64", None), ("rs", "fp64", None), ("od", "mixed", None), ("rs", "mixed", "mlp")])
def test_shared_and_separate_launches_are_bit_identical(qed, precision, sampler):
    # a0 = 500 and pulses from 1 ctau: a cascade of ~1e4 particles whose
    # positrons emit photons in most steps of its second half
    case = cascade_case(n=64, nsteps=1600, seed=3, a0=500.0, x0_ctau=1.0, qed=qed)
    shared, s1 = run_gpu_case(case, precision=precision, sampler=sampler)
    separate, s2 = run_gpu_case(case, precision=precision, sampler=sampler, shared=False)
    # the positrons' kernels have no launch shape of their own when shared
    for ph in (0, 1):
        assert s1.launch_shape(1, ph)[3] == 0 and s2.launch_shape(1, ph)[3] > 0
    e, p = s1.stats(0), s1.stats(1)
    print(qed, precision, sampler, "counts", {k: v["n"] for k, v in shared.items()},
          "photons appended by e-", e["spawned"], "by e+", p["spawned"],
          "capacity replays", e["replays"])
    assert p["spawned"] > 100 and e["spawned"] > 100
    assert e["replays"] > 0  # the stores grew through capacity halts and replays
    compare_final(shared, separate)
