"""Serving: requests admitted into the slots of one running context give the
same result as each request run alone (bit for bit: a replica's kernels read
only its own state, its Philox streams are keyed by its own seed and step),
and a request that arrives while others run is admitted when a slot is free.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(dml, W):
    from dml import serve
    sizes = [1500, 2600, 1800, 3900, 2200, 1200]
    steps = [24, 40, 16, 32, 20, 28]
    reqs = []
    for k, (n, st) in enumerate(zip(sizes, steps)):
        res = [dml.PISTON, dml.NO_RESERVOIR, dml.GCMC][k % 3]
        sysd = W.system_replica(n, 5000 + k, reservoir_band=res != dml.NO_RESERVOIR)
        kw = dict(z0=sysd["z0"], zmax=sysd["zmax"])
        need = len(sysd["pos"]) + (512 if res == dml.GCMC else 0)
        cfg = dml.dana_config(sysd["box"], sysd["skin"], need, h=sysd["h"],
                              integrator=dml.ERMAK if k % 2 else dml.BROWNIAN, reservoir=res,
                              precision=dml.FP32_FORCE, overlap_mode=dml.OVERLAP_ROUNDS,
                              seed=777 + k, act=0.01 if res == dml.GCMC else 0.0,
                              nadj=10 if res == dml.GCMC else 0, **kw)
        reqs.append(serve.Request(k, 0.0, cfg, sysd["pos"], sysd["species"], st, need=need))
    return reqs


def test_requests_alone_equal_in_mix(dml):
    from dml import serve, workloads as W
    reqs = _requests(dml, W)
    big = W.system_replica(4096, 1)["box"]
    # 3 slots for 6 requests: later ones wait for a slot to free
    pool = serve.SlotPool([(3072, 2, big), (4608, 1, big)])
    done, wall = serve.serve(pool, reqs, segment=8)
    pool.close()
    assert len(done) == len(reqs)
    assert any(q.t_admit > 0 for q in done)  # admitted while others ran
    for q in done:
        print("request", q.rid, "n", q.n, "steps", q.steps, "slot", q.rep, flush=True)
        assert q.status == 0
        c = dml.Context(q.cfg)
        c.set_state(q.pos, species=q.species)
        c.run(q.steps, graph=True)
        alone = c.state()
        c.close()
        for k in ("pos", "vel", "species", "flags"):
            np.testing.assert_array_equal(q.result[k], alone[k], err_msg="request %d %s" % (q.rid, k))
