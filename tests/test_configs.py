"""The benchmark configurations themselves, on the GPU, through the C ABI.

* C2: the 100k system, one step with injected noise against the oracle
  (cells/rows of src/Neighbor.F90:465-548 entry for entry; fuerza,
  src/dana.F90:1055-1139, and the one-step positions within 1e-12 in FP64;
  FP32 pair forces within 1e-5), the Ermak and the Brownian variants.
* C3: the first six replicas of the 256-replica ensemble (all three
  reservoirs under both integrators, FP32 forces) in ONE ensemble context,
  two whole steps of per-routine calls with injected noise, each replica
  against its own oracle; and the full 256-replica ensemble over 100
  device-resident graph steps through its invariants.
* C4: requests drawn from the serving distribution (lognormal sizes, median
  10k, sigma 1, clipped to 1k-200k, one of them above 100k), each the same
  alone as in the mix.

Bars as in test_gpu_parity.py.
"""
import os
import sys

import numpy as np
import pytest

from test_gpu_parity import _noise, assert_lists_equal, rel_err, say

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def force_err(f32, f64):
    """max_i |F32_i - F64_i| / max(|F64_i|, 1e-3 rms |F64|) over the atoms with a force"""
    na = np.linalg.norm(f32 - f64, axis=1)
    nb = np.linalg.norm(f64, axis=1)
    hot = nb > 0
    if not hot.any():  # no pair within its cutoff: both must be zero
        return float(np.max(na)) if len(na) else 0.0
    floor = 1e-3 * np.sqrt(np.mean(nb[hot] ** 2))
    return float(np.max(na[hot] / np.maximum(nb[hot], floor)))


@pytest.fixture(scope="module")
def c2():
    from dml import workloads as W
    return W.system_c2()


def _c2_pair(orc, dml, c2, ermak, precision=None):
    box, pos, z = c2["box"], c2["pos"], c2["species"]
    s = orc.System(orc.dana_params(box, c2["skin"], h=c2["h"], z0=c2["z0"], zmax=c2["zmax"],
                                   ermak=ermak, piston=False), pos, z)
    cfg = dml.dana_config(box, c2["skin"], len(pos), h=c2["h"], z0=c2["z0"], zmax=c2["zmax"],
                          integrator=dml.ERMAK if ermak else dml.BROWNIAN,
                          reservoir=dml.NO_RESERVOIR,
                          precision=dml.FP64 if precision is None else precision,
                          overlap_mode=dml.OVERLAP_ROUNDS)
    c = dml.Context(cfg)
    c.set_state(pos, species=z)
    return s, c


def test_c2_ermak_one_step_matches_oracle(orc, dml, c2):
    n = len(c2["pos"])
    say("C2 ermak", n)
    s, c = _c2_pair(orc, dml, c2, ermak=True)
    s.test_update(); c.test_update()
    assert_lists_equal(c, s)
    s.fuerza(); c.forces()
    so, sd = s.state(), c.state()
    assert rel_err(sd["force"], so["force"]) <= 1e-12
    assert rel_err(sd["epot"][:, None], so["epot"][:, None]) <= 1e-12
    say("C2 FP32 forces")
    _, c32 = _c2_pair(orc, dml, c2, ermak=True, precision=dml.FP32_FORCE)
    c32.test_update(); c32.forces()
    e32 = force_err(c32.state()["force"], so["force"])
    say("C2 FP32 force error", e32)
    assert e32 <= 1e-5
    c32.close()
    rng = np.random.default_rng(202)
    nref = s.nref
    nz = _noise(rng, nref, 6)
    ud = rng.random(nref)
    assert s.ermak_a(nz, ud) == 0
    c.ermak_a(nz, ud)
    so, sd = s.state(), c.state()
    assert rel_err(sd["pos"], so["pos"]) <= 1e-12
    s.fuerza(); c.forces()
    s.ermak_b(); c.ermak_b()
    so, sd = s.state(), c.state()
    assert rel_err(sd["force"], so["force"]) <= 1e-12
    assert rel_err(sd["vel"], so["vel"]) <= 1e-12
    st, r1 = s.test_update()
    assert st == 0 and r1 == c.test_update()[0]
    u = rng.random((16, c.total))
    assert s.overlap_rounds(u[:, :s.n].reshape(-1), 16) == 0
    c.overlap_rounds(u, 16)
    st, r2 = s.test_update()
    assert st == 0 and r2 == c.test_update()[0]
    so, sd = s.state(), c.state()
    assert rel_err(sd["pos"], so["pos"]) <= 1e-12
    np.testing.assert_array_equal(sd["species"], so["species"])
    np.testing.assert_array_equal(sd["flags"], so["flags"])
    assert_lists_equal(c, s)
    c.close()


def test_c2_brownian_one_step_matches_oracle(orc, dml, c2):
    """the hard-sphere Brownian variant (no forces in cbrownian_hs)"""
    say("C2 brownian", len(c2["pos"]))
    s, c = _c2_pair(orc, dml, c2, ermak=False)
    s.test_update(); c.test_update()
    rng = np.random.default_rng(203)
    nref = s.nref
    nz = _noise(rng, nref, 3)
    ud = rng.random(nref)
    assert s.brownian(nz, ud) == 0
    c.brownian(nz, ud)
    so, sd = s.state(), c.state()
    assert rel_err(sd["pos"], so["pos"]) <= 1e-12
    np.testing.assert_array_equal(sd["species"], so["species"])
    st, r1 = s.test_update()
    assert st == 0 and r1 == c.test_update()[0]
    u = rng.random((32, c.total))
    assert s.overlap_rounds(u[:, :s.n].reshape(-1), 32) == 0
    c.overlap_rounds(u, 32)
    st, r2 = s.test_update()
    assert st == 0 and r2 == c.test_update()[0]
    s.f_to_cg(); c.f_to_cg()
    so, sd = s.state(), c.state()
    assert rel_err(sd["pos"], so["pos"]) <= 1e-12
    np.testing.assert_array_equal(sd["species"], so["species"])
    np.testing.assert_array_equal(sd["flags"], so["flags"])
    np.testing.assert_array_equal(c.ref_order(), s.ref_order())
    assert_lists_equal(c, s)
    c.close()


def _oracle_replica(orc, dml, r, cfg):
    """the oracle system of one C3 replica (same parameters as its config)"""
    box = np.array(cfg.box[:])
    res = cfg.reservoir
    s = orc.System(orc.dana_params(box, cfg.skin, h=cfg.h, prob=cfg.prob, z0=cfg.z0, zmax=cfg.zmax,
                                   ermak=cfg.integrator == dml.ERMAK,
                                   piston=res == dml.PISTON), r["pos"], r["species"])
    if res == dml.CHUNKS:
        s.set_z1(r["z1"])
        s.chunk_init(r["chunk"], r["dist"])
    return s


def test_c3_six_replicas_in_one_ensemble_match_oracle(orc, dml):
    """replicas 0-5 of the C3 ensemble (piston, chunks, GCMC under the
    Brownian and the Ermak integrator) in one context; two steps of the main
    loop (dana.F90:173-263) as per-routine calls with injected noise; every
    replica against an oracle system of its own: Brownian replicas to 1e-12,
    the Ermak replicas' FP32 forces to 1e-5 and their positions to the error
    those forces carry into the second step; species, flags, ref and gcmc
    order, reservoir heights exact"""
    import bench
    reps = bench.build_replicas(6)
    cfgs = bench.c3_configs(dml, reps)
    say("C3 six replicas", [len(r["pos"]) for r in reps])
    ens = dml.Context(cfgs)
    orcs = []
    for k, (r, cfg) in enumerate(zip(reps, cfgs)):
        ens.set_state(r["pos"], species=r["species"], rep=k)
        if "chunk" in r:
            ens.set_chunk(r["chunk"], r["dist"], r["z1"], rep=k)
        orcs.append(_oracle_replica(orc, dml, r, cfg))
    off = [ens.ref_offset(k) for k in range(6)]
    for k, s in enumerate(orcs):
        s.test_update()
        if cfgs[k].reservoir == dml.GCMC:
            s.gcmc_init()
    r0 = ens.test_update()
    rho0 = {k: s.calc_rho() for k, s in enumerate(orcs) if cfgs[k].reservoir == dml.PISTON}
    ermak = [cfgs[k].integrator == dml.ERMAK for k in range(6)]
    rng = np.random.default_rng(303)
    T = ens.total
    for step in range(2):
        nz3 = rng.standard_normal(3 * T); nz6 = rng.standard_normal(6 * T); ud = rng.random(T)
        for k, s in enumerate(orcs):
            o, nref = off[k], s.nref
            if ermak[k]:
                assert s.ermak_a(nz6[6 * o:6 * (o + nref)], ud[o:o + nref]) == 0
            else:
                assert s.brownian(nz3[3 * o:3 * (o + nref)], ud[o:o + nref]) == 0
        ens.brownian(nz3, ud)
        ens.ermak_a(nz6, ud)
        ens.forces()
        for k, s in enumerate(orcs):
            if ermak[k]:
                s.fuerza()
                f_d = ens.state(k)["force"]
                assert force_err(f_d, s.state()["force"]) <= 1e-5, (step, k)
                s.ermak_b()
        ens.ermak_b()
        reb = ens.test_update()
        for k, s in enumerate(orcs):
            st, r1 = s.test_update()
            assert st == 0 and r1 == reb[k], (step, k)
        u = rng.random((16, T))
        ens.overlap_rounds(u, 16)
        for k, s in enumerate(orcs):
            assert s.overlap_rounds(u[:, off[k]:off[k] + s.n].reshape(-1), 16) == 0
        reb = ens.test_update()
        for k, s in enumerate(orcs):
            st, r2 = s.test_update()
            assert st == 0 and r2 == reb[k], (step, k)
            s.f_to_cg(cfgs[k].reservoir == dml.GCMC)
        ens.f_to_cg()
        for k, s in enumerate(orcs):
            if cfgs[k].reservoir == dml.GCMC:
                ug = rng.random((10, 5)); gg = rng.standard_normal((10, 3))
                assert s.gcmc(cfgs[k].act, 10, ug, gg) == 0
                ens.gcmc(cfgs[k].act, 10, ug, gg, rep=k)
        rho_d = ens.calc_rho()
        for k, s in enumerate(orcs):
            res = cfgs[k].reservoir
            rho_o = s.calc_rho(res == dml.CHUNKS)
            if res != dml.GCMC:
                assert rho_d[k] == rho_o, (step, k)
            if res == dml.CHUNKS:
                assert s.bloques(rho_o)[0] == 0
            elif res == dml.PISTON:
                s.maxz(rho_o, rho0[k])
        ens.bloques()
        ens.maxz()
        ens.end_step()
        for k, s in enumerate(orcs):
            so, sd = s.state(), ens.state(k)
            tag = "step %d replica %d" % (step, k)
            assert ens.nslots(k) == s.n, tag
            np.testing.assert_array_equal(sd["species"], so["species"], err_msg=tag)
            np.testing.assert_array_equal(sd["flags"], so["flags"], err_msg=tag)
            np.testing.assert_array_equal(ens.ref_order(k), s.ref_order(), err_msg=tag)
            tol = 1e-12 if not ermak[k] or step == 0 else 1e-6
            assert rel_err(sd["pos"], so["pos"]) <= tol, tag
            if cfgs[k].reservoir == dml.GCMC:
                np.testing.assert_array_equal(ens.gcmc_order(k), s.gcmc_order(), err_msg=tag)
            if cfgs[k].reservoir != dml.GCMC:
                assert ens.zmax(k) == s.reservoir()[2], tag
    ens.close()


def test_c3_ensemble_100_graph_steps_invariants(dml):
    """the whole 256-replica C3 ensemble, 100 device-resident graph steps:
    no replica halts, overlap calls that end with atoms still overlapping
    (the round schedule's bound of 16 rounds per step) stay below 1 in 1000
    replica-steps, atoms stay in their boxes, CG only grows, Li+ count
    changes only by GCMC moves and chunk blocks, and the particle-step
    counter adds the atoms alive each step"""
    import bench
    reps = bench.build_replicas(256)
    cfgs = bench.c3_configs(dml, reps)
    ens = dml.Context(cfgs)
    for k, r in enumerate(reps):
        ens.set_state(r["pos"], species=r["species"], rep=k)
        if "chunk" in r:
            ens.set_chunk(r["chunk"], r["dist"], r["z1"], rep=k)
    say("C3 ensemble", sum(len(r["pos"]) for r in reps))
    c0 = [ens.counters(k) for k in range(256)]
    ens.run(100, graph=True)
    unfinished = 0
    for k, r in enumerate(reps):
        tag = "replica %d" % k
        assert ens.replica_status(k) == 0, tag
        cnt = ens.counters(k)
        unfinished += cnt["ov_unfinished"] - c0[k]["ov_unfinished"]
        st = ens.state(k)
        alive = (st["flags"] & 1) != 0
        sp = st["species"][alive]
        p = st["pos"][alive]
        box = np.array(cfgs[k].box[:])
        zmax = ens.zmax(k)
        assert np.all((p[:, 0] >= 0) & (p[:, 0] < box[0]) & (p[:, 1] >= 0) & (p[:, 1] < box[1])), tag
        ions = sp != 2
        assert np.all(p[ions, 2] > 0) and np.all(p[ions, 2] <= zmax + 1e-9), tag
        assert (sp == 3).sum() == 0, tag  # F became CG at step end
        n_cg0 = int((r["species"] == 2).sum())
        assert (sp == 2).sum() >= n_cg0, tag
        res = bench.RESERVOIR_CYCLE[k % 3]
        n_li = int(alive.sum())
        n0 = len(r["pos"])
        ps = cnt["particle_steps"] - c0[k]["particle_steps"]
        if res == "GCMC":
            assert n_li == n0 + cnt["ins"] - cnt["del_"], tag
            assert 100 * (n0 - cnt["del_"]) <= ps <= 100 * (n0 + cnt["ins"]), tag
        elif res == "CHUNKS":
            assert (n_li - n0) % len(r["chunk"]) == 0, tag
            assert 100 * n0 <= ps <= 100 * n_li, tag
        else:
            assert n_li == n0, tag
            assert ps == 100 * n0, tag
    say("overlap calls unfinished", unfinished, "of", 256 * 100)
    assert unfinished <= 256 * 100 // 1000
    ens.close()


def test_c4_requests_alone_equal_in_mix(dml):
    """requests from the serving distribution, one above 100k atoms, through
    the continuous-batching loop with fewer slots than requests; each result
    equals the same request run alone, bit for bit"""
    import math
    from dml import serve, workloads as W
    rng = np.random.default_rng(404)
    m = 7
    sizes = np.clip(np.exp(rng.normal(math.log(10000.0), 1.0, m)), 1000, 200000).astype(int)
    sizes[int(np.argmax(sizes))] = max(int(sizes.max()), 120_000)
    steps = np.maximum(1, np.round(rng.integers(1000, 100001, m) * 5e-4)).astype(int)
    cycle = (dml.PISTON, dml.NO_RESERVOIR, dml.GCMC)
    reqs = []
    for k, (n, st) in enumerate(zip(sizes, steps)):
        res = cycle[k % 3]
        sysd = W.system_replica(int(n), 6000 + k, reservoir_band=res != dml.NO_RESERVOIR)
        nat = len(sysd["pos"])
        need = nat + (min(max(1024, nat // 4), 10 * int(st) + 64) if res == dml.GCMC else 0)
        cfg = dml.dana_config(sysd["box"], sysd["skin"], need, h=sysd["h"],
                              integrator=dml.ERMAK if k % 2 else dml.BROWNIAN, reservoir=res,
                              precision=dml.FP32_FORCE, overlap_mode=dml.OVERLAP_ROUNDS,
                              seed=888 + k, act=0.01 if res == dml.GCMC else 0.0,
                              nadj=10 if res == dml.GCMC else 0, z0=sysd["z0"], zmax=sysd["zmax"])
        reqs.append(serve.Request(k, 0.02 * k, cfg, sysd["pos"], sysd["species"], st, need=need))
    say("C4 requests", list(zip(sizes.tolist(), steps.tolist())))
    big = max(q.need for q in reqs)
    pool = serve.SlotPool([(32768, 2, W.system_replica_box(32768)),
                           (big, 1, W.system_replica_box(big))])
    done, wall = serve.serve(pool, reqs, segment=4)
    pool.close()
    assert len(done) == m
    assert any(q.t_admit > q.t_arrival for q in done)  # some request waited for a slot
    for q in done:
        say("request", q.rid, "n", q.n, "steps", q.steps, "status", q.status)
        c = dml.Context(q.cfg)
        c.set_state(q.pos, species=q.species)
        c.set_step_limit(q.steps, 0)
        c.run(q.steps, graph=True)
        alone = c.state()
        assert c.replica_status(0) == q.status
        c.close()
        for key in ("pos", "vel", "species", "flags"):
            np.testing.assert_array_equal(q.result[key], alone[key],
                                          err_msg="request %d %s" % (q.rid, key))
