"""Batched GCMC (DML_GCMC_BATCHED): the trials of one gcmc_run call
(dana.F90:590-713) evaluated in parallel and committed in trial order give
the sequential result bit for bit.  Checked against the oracle on injected
uniforms, against the sequential device kernel on calls of thousands of
trials that interact (atoms inserted and deleted in the same call, hard-core
neighbours deleted before a later insertion, earlier points blocking later
ones), and over whole Philox steps with the batch inside the captured step."""
import numpy as np
import pytest

from test_gpu_parity import _assert_same_groups, _step_pair, electrode, rel_err, rsa, say

pytestmark = pytest.mark.gpu


def _gcmc_system(rng, box, nx, n_ion, dmin=3.3):
    el = electrode(3.51, nx, nx, 2, 1.755)
    ions = rsa(rng, n_ion, np.array([0, 0, 4.0]), np.array([box[0], box[1], box[2] - 1.0]), dmin,
               box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    return pos, z, ions


def _alternate(u):
    """force the kinds: even trials insert, odd ones delete"""
    u = u.copy()
    u[0::2, 0] = 0.25
    u[1::2, 0] = 0.75
    return u


def test_batched_gcmc_matches_oracle(orc, dml):
    """the oracle's sequential gcmc on the same injected uniforms and
    normals, four calls with Brownian steps between; every call ran batched"""
    rng = np.random.default_rng(61)
    box = np.array([40.0, 40.0, 60.0])
    pos, z, ions = _gcmc_system(rng, box, 11, 220)
    z0, zmax = 20.0, box[2] - 0.5
    say("gcmc batch oracle", len(pos))
    s = orc.System(orc.dana_params(box, 2.0, h=5e-5, z0=z0, zmax=zmax, ermak=False, piston=False),
                   pos, z)
    cfg = dml.dana_config(box, 2.0, len(pos) + 1024, h=5e-5, z0=z0, zmax=zmax,
                          integrator=dml.BROWNIAN, reservoir=dml.GCMC,
                          overlap_mode=dml.OVERLAP_EXACT, gcmc_mode=dml.GCMC_BATCHED)
    c = dml.Context(cfg)
    c.set_state(pos, species=z)
    s.test_update(); c.test_update()
    s.gcmc_init()
    vol = box[0] * box[1] * (zmax - z0)
    n_in = int(((ions[:, 2] >= z0) & (ions[:, 2] <= zmax)).sum())
    act = n_in / vol
    nadj = 120
    for it in range(4):
        u = rng.random((nadj, 5))
        g = rng.standard_normal((nadj, 3))
        assert s.gcmc(act, nadj, u, g) == 0
        c.gcmc(act, nadj, u, g)
        out = c.gcmc_outcomes(nadj)
        assert out is not None, "call %d ran trial by trial" % it
        _assert_same_groups(s, c, "gcmc call %d" % it)
        assert (s.c.ins, s.c.del_) == (c.counters()["ins"], c.counters()["del_"])
        _step_pair(orc, dml, s, c, rng, "exact", 2, gcmc=True)
        _assert_same_groups(s, c, "steps after call %d" % it)
    assert s.c.ins > 20 and s.c.del_ > 20, (s.c.ins, s.c.del_)
    c.close()


def _state_equal(a, b, tag):
    sa, sb = a.state(), b.state()
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg="%s %s" % (tag, k))
    np.testing.assert_array_equal(a.ref_order(), b.ref_order(), err_msg=tag)
    np.testing.assert_array_equal(a.gcmc_order(), b.gcmc_order(), err_msg=tag)
    for x, y in zip(a.slots(), b.slots()):
        np.testing.assert_array_equal(x, y, err_msg=tag)
    na, _, ia = a.lists()
    nb, _, ib = b.lists()
    np.testing.assert_array_equal(na, nb, err_msg=tag)
    np.testing.assert_array_equal(ia, ib, err_msg=tag)
    ca, cb = a.cells(), b.cells()
    np.testing.assert_array_equal(ca[3], cb[3], err_msg=tag)


def test_batched_gcmc_equals_sequential_on_large_calls(dml):
    """a dense control volume, calls of 2,048 trials with the kinds
    alternating: the batch against the sequential kernel on the same
    uniforms, whole state after each call and after steps; the calls reach
    the cases where trials of one call interact"""
    rng = np.random.default_rng(62)
    box = np.array([80.0, 80.0, 70.0])
    pos, z, ions = _gcmc_system(rng, box, 22, 4000, dmin=3.25)
    z0, zmax = 30.0, box[2] - 0.5
    say("gcmc batch vs sequential", len(pos))
    vol = box[0] * box[1] * (zmax - z0)
    n_in = int(((ions[:, 2] >= z0) & (ions[:, 2] <= zmax)).sum())
    ctx = []
    for mode in (0, dml.GCMC_BATCHED):
        cfg = dml.dana_config(box, 2.0, len(pos) + 8192, h=5e-5, z0=z0, zmax=zmax,
                              integrator=dml.BROWNIAN, reservoir=dml.GCMC,
                              overlap_mode=dml.OVERLAP_ROUNDS, gcmc_mode=mode, seed=9)
        c = dml.Context(cfg)
        c.set_state(pos, species=z)
        c.test_update()
        ctx.append(c)
    seq, bat = ctx
    nadj = 2048
    reinserted = 0
    for it in range(3):
        act = (1.0 + 0.5 * it) * n_in / vol
        u = _alternate(rng.random((nadj, 5)))
        g = rng.standard_normal((nadj, 3))
        for c in ctx:
            c.gcmc(act, nadj, u, g)
        out = bat.gcmc_outcomes(nadj)
        assert out is not None, "call %d ran trial by trial" % it
        _state_equal(seq, bat, "call %d" % it)
        # the call's own insertions deleted again later in it
        ins = set(out["slot"][(out["kind"] == 0) & (out["accepted"] == 1)].tolist())
        dels = out["slot"][(out["kind"] == 1) & (out["accepted"] == 1)]
        reinserted += sum(1 for x in dels if x in ins)
        say("call", it, "accepted", int(out["accepted"].sum()), "deleted own", reinserted)
        for c in ctx:
            c.run(2, graph=True)
        _state_equal(seq, bat, "steps after call %d" % it)
    assert reinserted > 0
    for c in ctx:
        c.close()


def test_batched_gcmc_philox_steps_equal_sequential(dml):
    """whole Philox steps in hipGraphs, 2,048 trials a step with alternating
    kinds (the C5 reservoir form): batch and sequential give the same bits"""
    rng = np.random.default_rng(63)
    box = np.array([70.0, 70.0, 80.0])
    pos, z, ions = _gcmc_system(rng, box, 19, 3500, dmin=3.25)
    z0, zmax = 40.0, box[2] - 0.5
    ctx = []
    for mode in (dml.GCMC_ALTERNATE, dml.GCMC_ALTERNATE | dml.GCMC_BATCHED):
        cfg = dml.dana_config(box, 10.0, len(pos) + 16384, h=1e-2, z0=z0, zmax=zmax,
                              integrator=dml.ERMAK, reservoir=dml.GCMC,
                              precision=dml.FP32_FORCE, overlap_mode=dml.OVERLAP_ROUNDS,
                              act=0.01, nadj=2048, gcmc_mode=mode, seed=17)
        c = dml.Context(cfg)
        c.set_state(pos, species=z)
        ctx.append(c)
    for k in range(4):
        for c in ctx:
            c.run(5, graph=True)
        assert ctx[1].gcmc_outcomes(2048) is not None
        _state_equal(ctx[0], ctx[1], "after %d steps" % (5 * (k + 1)))
        cnt = [c.counters() for c in ctx]
        assert (cnt[0]["ins"], cnt[0]["del_"]) == (cnt[1]["ins"], cnt[1]["del_"])
    assert cnt[1]["ins"] > 100 and cnt[1]["del_"] > 100, cnt[1]
    for c in ctx:
        c.close()
