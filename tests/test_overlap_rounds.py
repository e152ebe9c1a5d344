"""The round schedule of the overlap pass (the device's fast path, oracle
orc_overlap_rounds) pinned against the reference's overlap_moveback
(src/dana.F90:849-943, oracle orc_overlap, which the three golden frames pin).

The checks run the C oracle and a pure-Python restatement of both algorithms
(oracle/ovl.py) on small dense clusters whose ions moved by about a Brownian
step at the reference's defaults (sqrt(2Dh) = 2.2 A per component at D = 250,
h = 1e-2), with electrode (CG) atoms among them:

* the restatement equals the C oracle, both algorithms, bit for bit;
* at prob = 1 (the golden runs and every benchmark configuration), the
  schedule gives the reference pass's result bit for bit unless a round
  moves an atom to within rcut of an atom the same round scans later in ref
  order (in-round interference, ovl.round_interference);
* at prob = 1 its result is always one the reference pass produces for some
  ordering of the ref list (all orders enumerated on 6-ion clusters);
* both end in the state the reference's recursion ends in: no pair of listed
  non-CG ref atoms within rcut unless one of them deposited;
* every contact trial draws exactly one uniform.
"""
import itertools

import numpy as np
import pytest

import ovl


def _cluster(orc, seed, n_ion=7, n_el=2, prob=1.0, piston=False, box=(14.0, 14.0, 20.0), disp=1.8):
    """ions (and a few CG atoms) by random sequential addition at 3.25 A,
    then one Brownian step with injected noise of `disp` A per component; the
    oracle system after the step's first test_update, and its model state"""
    rng = np.random.default_rng(seed)
    box = np.array(box)
    pts = [[rng.random() * box[0], rng.random() * box[1], 0.5] for _ in range(n_el)]
    while len(pts) < n_el + n_ion:
        p = rng.random(3) * [box[0], box[1], 6.0] + [0, 0, 2.0]
        d = np.array(pts) - p
        d[:, :2] -= box[:2] * np.round(d[:, :2] / box[:2])
        if np.min((d * d).sum(1)) >= 3.25 ** 2:
            pts.append(list(p))
    z = np.array([2] * n_el + [1] * n_ion, np.int8)
    s = orc.System(orc.dana_params(box, 2.0, prob=prob, piston=piston, ermak=False, zmax=box[2],
                                   z0=100.0), np.array(pts), z)
    s.test_update()
    noise = rng.standard_normal(3 * n_ion) * disp / np.sqrt(2 * 250 * 1e-2)
    s.brownian(np.concatenate([noise, np.zeros(3)]), np.ones(n_ion + 1))
    s.test_update()
    st = s.state()
    nn, rp, idx = s.lists()
    n = len(st["species"])
    m = dict(pos=st["pos"].copy(), old=st["old_cg"].copy(), vel=st["vel"].copy(),
             acel=st["acel"].copy(), zs=st["species"].astype(int), skip=(st["flags"] & 8) != 0,
             rows=[list(idx[rp[i]:rp[i + 1]]) for i in range(n)], order=list(s.ref_order()),
             limbo=(st["flags"] & 2) != 0, box=box, rcut=3.2, prob=prob, z0=100.0, piston=piston)
    return s, m


def _same(a, b):
    return (np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["zs"], b["zs"])
            and np.array_equal(a["skip"], b["skip"]) and np.array_equal(a["vel"], b["vel"]))


def _oracle_state(s, m):
    st = s.state()
    return dict(pos=st["pos"], zs=st["species"].astype(int), skip=(st["flags"] & 8) != 0, vel=st["vel"])


@pytest.mark.parametrize("prob,piston", [(1.0, False), (0.5, False), (0.5, True)])
def test_restatement_equals_oracle(orc, prob, piston):
    """ovl.seq_overlap = orc_overlap and ovl.round_overlap = orc_overlap_rounds,
    positions, species, skip flags, velocities and counters"""
    for seed in range(40):
        rng = np.random.default_rng(100 + seed)
        s, m = _cluster(orc, seed, n_ion=10, n_el=3, prob=prob, piston=piston, box=(20.0, 20.0, 24.0),
                        disp=1.6)
        uc = rng.random(4096)
        it = iter(uc)
        py, pc, pst = ovl.seq_overlap(m, lambda i, k, p: next(it))
        t0, c0, c3 = s.c.try_, s.c.choques, s.c.choques3
        cst, used = s.overlap(uc)
        assert (pst == 0) == (cst == 0), seed
        assert _same(py, _oracle_state(s, m)), seed
        assert (pc["try_"], pc["choques"], pc["choques3"]) == (s.c.try_ - t0, s.c.choques - c0,
                                                              s.c.choques3 - c3), seed
        assert used == pc["try_"]  # one uniform per contact trial

        s2, m2 = _cluster(orc, seed, n_ion=10, n_el=3, prob=prob, piston=piston,
                          box=(20.0, 20.0, 24.0), disp=1.6)
        ur = rng.random((64, len(m2["zs"])))
        py2, pc2, pst2 = ovl.round_overlap(m2, lambda i, k, r: ur[r, i], 64)
        c0 = s2.c.choques
        cst2 = s2.overlap_rounds(ur, 64)
        assert (pst2 == 0) == (cst2 == 0), seed
        assert _same(py2, _oracle_state(s2, m2)), seed
        assert pc2["choques"] == s2.c.choques - c0, seed


def test_rounds_equal_reference_pass_without_in_round_interference(orc):
    """prob = 1: bit-exact whenever no round moves an atom within rcut of an
    atom it scans later; the interference is common in these dense clusters
    but changes the outcome rarely"""
    cases = differ = interfered = 0
    for seed in range(800):
        s, m = _cluster(orc, 5000 + seed)
        if sum(1 for i in m["order"] if not m["skip"][i]) < 2:
            continue
        U = np.random.default_rng(seed).random((len(m["zs"]), 64))
        u = lambda i, k, r: U[i, k]
        sq, _, sst = ovl.seq_overlap(m, u, 1000)
        rr, _, rst = ovl.round_overlap(m, u, 64)
        if sst != 0 or rst != 0:
            continue
        cases += 1
        inter = ovl.round_interference(m, u)
        interfered += inter
        if not (np.array_equal(sq["pos"], rr["pos"]) and np.array_equal(sq["zs"], rr["zs"])):
            differ += 1
            assert inter, "rounds differ from the reference pass without interference (seed %d)" % seed
    assert cases > 700 and interfered > 20, (cases, interfered)
    assert differ <= cases // 100, (differ, cases)


def test_rounds_result_reachable_by_some_ref_order(orc):
    """prob = 1: the schedule's result is one the reference pass gives for
    some ordering of the ref list (every ordering of 6-ion clusters)"""
    checked = 0
    for seed in range(60):
        s, m = _cluster(orc, 7000 + seed, n_ion=6, n_el=2, box=(13.0, 13.0, 20.0))
        if sum(1 for i in m["order"] if not m["skip"][i]) < 3:
            continue
        u = lambda i, k, r: 0.0
        rr, _, rst = ovl.round_overlap(m, u, 64)
        assert rst == 0
        key = lambda st: (st["pos"].tobytes(), tuple(st["zs"]))
        target = key(rr)
        found = False
        for perm in itertools.permutations(m["order"]):
            sq, _, sst = ovl.seq_overlap(dict(m, order=list(perm)), u, 1000)
            if sst == 0 and key(sq) == target:
                found = True
                break
        assert found, seed
        checked += 1
    assert checked > 30, checked


@pytest.mark.parametrize("algo", ["reference", "rounds"])
def test_terminal_state(orc, algo):
    """prob = 1, no piston: when the recursion ends, no listed pair of
    non-CG ref atoms lies within rcut unless one of them deposited (F)"""
    for seed in range(300):
        s, m = _cluster(orc, 9000 + seed)
        u = lambda i, k, r: 0.0
        st, _, status = (ovl.seq_overlap(m, u, 1000) if algo == "reference"
                         else ovl.round_overlap(m, u, 64))
        if status != 0:
            continue
        ref = m["order"]
        for a, b in itertools.combinations(ref, 2):
            if st["zs"][a] in (2, 3) or st["zs"][b] in (2, 3):
                continue
            if b not in m["rows"][a] and a not in m["rows"][b]:
                continue
            v = ovl.vdist(st, st["pos"][a], st["pos"][b])
            assert v @ v > 3.2 ** 2, (seed, a, b)
