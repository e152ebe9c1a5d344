"""The launch-shape autotuner: valid choices, cached by signature, and a
tuned context computes the same trajectory as an untuned one (to FP64
summation order: positions within 1e-12 relative after a few steps)."""
import numpy as np
import pytest


def _ensemble(dml):
    from dml import workloads as W
    cfgs, states = [], []
    for k, n in enumerate((2500, 4000, 3000)):
        band = k != 1
        sysd = W.system_replica(n, 900 + k, reservoir_band=band)
        res = (dml.PISTON, dml.NO_RESERVOIR, dml.GCMC)[k]
        cfgs.append(dml.dana_config(sysd["box"], sysd["skin"], n + 1024, h=sysd["h"],
                                    z0=sysd["z0"], zmax=sysd["zmax"],
                                    integrator=dml.ERMAK if k % 2 else dml.BROWNIAN,
                                    reservoir=res, precision=dml.FP64,
                                    overlap_mode=dml.OVERLAP_ROUNDS, seed=31 + k,
                                    act=0.01 if res == dml.GCMC else 0.0,
                                    nadj=10 if res == dml.GCMC else 0))
        states.append((sysd["pos"], sysd["species"]))
    return cfgs, states


def test_signature_is_stable_and_distinguishes(dml):
    from dml import autotune
    cfgs, states = _ensemble(dml)
    sizes = [len(p) for p, _ in states]
    assert autotune.signature(cfgs, sizes) == autotune.signature(cfgs, sizes)
    assert autotune.signature(cfgs, sizes) != autotune.signature(cfgs[:2], sizes[:2])


@pytest.mark.gpu
def test_autotune_choices_and_same_trajectory(dml, tmp_path):
    from dml import autotune
    cfgs, states = _ensemble(dml)
    cache = str(tmp_path / "tune.json")
    choice = autotune.tune(cfgs, states, cache_path=cache)
    print("choice", choice, flush=True)
    assert choice["lanes_per_atom"] in autotune.FORCE_LANES
    assert choice["list_kernel"] in autotune.LIST_KERNELS
    assert choice["overlap_lanes"] in autotune.OVERLAP_LANES
    assert choice["overlap_rounds"] in autotune.OVERLAP_ROUNDS
    assert autotune.tune(cfgs, states, cache_path=cache)["cached"]
    out = []
    for tuned in (False, True):
        c = dml.Context(cfgs)
        if tuned:
            autotune.apply(c, choice)
        else:  # the round count is part of the schedule, not a launch shape
            c.set_tuning(overlap_rounds=choice["overlap_rounds"])
        for k, (p, s) in enumerate(states):
            c.set_state(p, species=s, rep=k)
        c.run(6, graph=True)
        out.append([c.state(k) for k in range(len(cfgs))])
        c.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a["species"], b["species"])
        d = np.abs(a["pos"] - b["pos"]).max() / np.abs(b["pos"]).max()
        assert d <= 1e-12, d
