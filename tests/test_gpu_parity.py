"""Parity of the HIP path (through the C ABI) with the oracle and with the
reference-built cell/list fixtures, on identical inputs with injected noise.

Bars:
* cells and neighbour lists: bit-exact, entry for entry (list order too);
* FP64 forces, energies and one-step positions: relative error <= 1e-12
  (norm with a floor of 1e-3 * rms so near-cutoff zeros do not divide by 0);
* FP32-force mode: force error <= 1e-5 with the same norm;
* integer/flag state (species, skip, ref order): exact.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dml():
    import dml as _dml
    return _dml


def rel_err(a, b):
    """max_i |a_i - b_i| / max(|b_i|, 1e-3 * rms |b|) over rows"""
    a = np.atleast_2d(a); b = np.atleast_2d(b)
    na = np.linalg.norm(a - b, axis=-1)
    nb = np.linalg.norm(b, axis=-1)
    rms = np.sqrt(np.mean(nb ** 2)) if nb.size else 0.0
    floor = max(1e-3 * rms, 1e-300)
    return float(np.max(na / np.maximum(nb, floor))) if na.size else 0.0


def rsa(rng, n, lo, hi, dmin, box):
    pts = []
    tries = 0
    while len(pts) < n and tries < 100000:
        tries += 1
        p = lo + rng.random(3) * (hi - lo)
        if pts:
            d = np.array(pts) - p
            d[:, :2] -= box[:2] * np.round(d[:, :2] / box[:2])
            if np.min((d * d).sum(1)) < dmin * dmin:
                continue
        pts.append(p)
    return np.array(pts)


def electrode(a, nx, ny, nl, ztop):
    pts = []
    for l in range(nl):
        off = 0.5 * (l % 2)
        for i in range(nx):
            for j in range(ny):
                pts.append([(i + off) * a, (j + off) * a, ztop - l * a / 2])
    return np.array(pts)


def pair(orc, dml, box, skin, pos, z, **kw):
    """an oracle system and a device context holding the same atoms"""
    ermak = kw.get("ermak", True)
    piston = kw.get("piston", False)
    okw = dict(h=kw.get("h", 1e-2), prob=kw.get("prob", 1.0), z0=kw.get("z0", 100.0),
               zmax=kw.get("zmax", box[2]), dif_sc=kw.get("dif_sc", 250.0),
               dif_sei=kw.get("dif_sei", 250.0), ermak=ermak, piston=piston)
    s = orc.System(orc.dana_params(box, skin, **okw), pos, z)
    cfg = dml.dana_config(box, skin, len(pos) + 64, h=okw["h"], prob=okw["prob"], z0=okw["z0"],
                          zmax=okw["zmax"], dif_sc=okw["dif_sc"], dif_sei=okw["dif_sei"],
                          integrator=dml.ERMAK if ermak else dml.BROWNIAN,
                          reservoir=dml.PISTON if piston else dml.NO_RESERVOIR,
                          precision=kw.get("precision", dml.FP64),
                          overlap_mode=kw.get("overlap_mode", dml.OVERLAP_EXACT),
                          lanes_per_atom=kw.get("lpa", 0))
    c = dml.Context(cfg)
    c.set_state(pos, species=z)
    return s, c


def c1_system(seed=3, n_ion=600, skin=10.0):
    rng = np.random.default_rng(seed)
    box = np.array([50.0, 50.0, 100.0])
    el = electrode(3.51, 14, 14, 2, 1.755)
    ions = rsa(rng, n_ion, np.array([0, 0, 4.0]), np.array([50, 50, 99.0]), 2.6, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    return box, skin, pos, z


def assert_lists_equal(c, s):
    """rows of the ref atoms, entry for entry (the only rows any routine
    reads; an atom that left ref keeps a stale row in the reference)"""
    nn_d, rp_d, idx_d = c.lists()
    nn_o, rp_o, idx_o = s.lists()
    ref = (s.state()["flags"] & 4) != 0
    nn_o = np.where(ref, nn_o, 0)
    idx_o = np.concatenate([idx_o[rp_o[i]:rp_o[i + 1]] for i in np.nonzero(ref)[0]] + [idx_o[:0]])
    np.testing.assert_array_equal(nn_d, nn_o)
    np.testing.assert_array_equal(idx_d, idx_o)


NB = np.load(os.path.join(G, "nb_cases.npz"))


@pytest.mark.parametrize("lb_mode", ["0", "1", "3", "4", "5"])
@pytest.mark.parametrize("name", list(NB["names"]))
def test_cells_and_lists_equal_reference_modules(dml, name, lb_mode, monkeypatch):
    """lb_mode 0: cell-tile list build through LDS; 1: one wave per atom;
    3: one wave per cell; 4, 5: one wave per cell, candidates in registers
    (3 or 2 chunks of 64 a pass)"""
    monkeypatch.setenv("DML_LISTBUILD", lb_mode)
    d = {k.split("/", 1)[1]: NB[k] for k in NB.files if k.startswith(name + "/")}
    say("case", name, len(d["z"]), "lb_mode", lb_mode)
    rc, skin = d["rc_skin"]
    cfg = dml.dana_config(d["box"], skin, len(d["z"]) + 64, rcut=rc, integrator=dml.BROWNIAN,
                          reservoir=dml.NO_RESERVOIR)
    c = dml.Context(cfg)
    c.set_state(d["pos"], species=d["z"])
    reb = c.test_update()
    assert reb[0] == 1
    ncells, cell, use_cells, ci = c.cells()
    assert use_cells == bool(d["tess"])
    np.testing.assert_array_equal(ncells, d["ncells"])
    if use_cells:
        np.testing.assert_array_equal(cell, d["cell"])
        np.testing.assert_array_equal(ci, d["ci"])
    nn, rp, idx = c.lists()
    for r, row in enumerate(d["rows"]):
        np.testing.assert_array_equal(idx[rp[row]:rp[row + 1]], d["idx"][d["rowptr"][r]:d["rowptr"][r + 1]])
    assert int(nn.sum()) == len(d["idx"])
    c.close()


@pytest.mark.parametrize("lpa", [1, 8, 64])
def test_forces_fp64_match_oracle(orc, dml, lpa):
    box, skin, pos, z = c1_system()
    # a few F atoms (deposited this step) interact with Li and F
    z = z.copy()
    ions = np.nonzero(z == 1)[0]
    z[ions[::37]] = 3
    say("forces fp64 lpa", lpa, len(pos))
    s, c = pair(orc, dml, box, skin, pos, z, lpa=lpa)
    s.test_update(); c.test_update()
    assert_lists_equal(c, s)
    s.fuerza(); c.forces()
    so, sd = s.state(), c.state()
    ref = so["species"] != 2
    assert rel_err(sd["force"][ref], so["force"][ref]) <= 1e-12
    assert rel_err(sd["epot"][ref, None], so["epot"][ref, None]) <= 1e-12
    assert np.abs(so["force"][ref]).max() > 1.0  # the system has real forces
    c.close()


def test_forces_ghost_layer_rows(orc, dml):
    """atoms above box(3): their rows are found by nobody; the reaction on
    their neighbours comes from the ghost scatter."""
    rng = np.random.default_rng(9)
    box = np.array([60.0, 60.0, 60.0])
    pos = rsa(rng, 500, np.zeros(3), np.array([60, 60, 70.0]), 2.6, box)
    z = np.ones(len(pos), np.int8)
    say("ghost forces", len(pos))
    s, c = pair(orc, dml, box, 10.0, pos, z, zmax=80.0)
    s.test_update(); c.test_update()
    assert_lists_equal(c, s)
    _, _, _, ci = s.cells()
    assert (ci[:, 2] == 5).any() or (ci[:, 2] > 4).any()
    s.fuerza(); c.forces()
    so, sd = s.state(), c.state()
    assert rel_err(sd["force"], so["force"]) <= 1e-12
    assert rel_err(sd["epot"][:, None], so["epot"][:, None]) <= 1e-12
    c.close()


def test_forces_fp32_error_bound(orc, dml):
    box, skin, pos, z = c1_system(seed=4)
    say("forces fp32", len(pos))
    s, c = pair(orc, dml, box, skin, pos, z, precision=dml.FP32_FORCE)
    s.test_update(); c.test_update()
    s.fuerza(); c.forces()
    so, sd = s.state(), c.state()
    ref = so["species"] != 2
    err = rel_err(sd["force"][ref], so["force"][ref])
    say("fp32 force rel err", err)
    assert err <= 1e-5
    assert rel_err(sd["epot"][ref, None], so["epot"][ref, None]) <= 1e-5
    c.close()


def _noise(rng, nref, k):
    return rng.standard_normal(nref * k)


def test_brownian_step_matches_oracle(orc, dml):
    rng = np.random.default_rng(11)
    box = np.array([50.0, 50.0, 100.0])
    low = rsa(rng, 60, np.array([0, 0, 0.2]), np.array([50, 50, 1.0]), 3.2, box)
    pos = np.vstack([low, rsa(rng, 340, np.array([0, 0, 4.5]), np.array([50, 50, 99.5]), 3.2, box)])
    z = np.ones(len(pos), np.int8)
    say("brownian", len(pos))
    s, c = pair(orc, dml, box, 2.0, pos, z, ermak=False, zmax=100.0, dif_sei=100.0, prob=0.5)
    nz = _noise(rng, len(pos), 3)
    ud = rng.random(len(pos))
    assert s.brownian(nz, ud) == 0
    c.brownian(nz, ud)
    so, sd = s.state(), c.state()
    assert rel_err(sd["pos"], so["pos"]) <= 1e-12
    assert rel_err(sd["vel"], so["vel"]) <= 1e-12
    np.testing.assert_array_equal(sd["old_cg"], so["old_cg"])
    np.testing.assert_array_equal(sd["species"], so["species"])
    np.testing.assert_array_equal(sd["flags"], so["flags"])
    assert (so["species"] == 3).any()  # some electrode deposits happened
    c.close()


def test_ermak_step_matches_oracle(orc, dml):
    box, skin, pos, z = c1_system(seed=5)
    rng = np.random.default_rng(12)
    say("ermak a/b", len(pos))
    s, c = pair(orc, dml, box, skin, pos, z, zmax=100.0, prob=0.7)
    vel = rng.standard_normal((len(pos), 3)) * 5
    acel = rng.standard_normal((len(pos), 3)) * 50
    s.set_state(vel=vel, acel=acel); c.set_state(pos, species=z, vel=vel, acel=acel)
    s.test_update(); c.test_update()
    s.fuerza(); c.forces()
    nref = s.nref
    nz = _noise(rng, nref, 6)
    ud = rng.random(nref)
    assert s.ermak_a(nz, ud) == 0
    c.ermak_a(nz, ud)
    so, sd = s.state(), c.state()
    assert rel_err(sd["pos"], so["pos"]) <= 1e-12
    np.testing.assert_array_equal(sd["species"], so["species"])
    s.fuerza(); c.forces()
    s.ermak_b(); c.ermak_b()
    so, sd = s.state(), c.state()
    ref = so["species"] != 2
    assert rel_err(sd["vel"][ref], so["vel"][ref]) <= 1e-12
    assert rel_err(sd["acel"][ref], so["acel"][ref]) <= 1e-12
    # salida's energia and kion temperature on these states
    ob_d, ob_o = c.observables()[0], s.observables()
    assert ob_o[0] > 0
    np.testing.assert_allclose(ob_d[:2], ob_o[:2], rtol=1e-12)
    assert ob_d[2] == ob_o[2] and ob_d[3] == ob_o[3]
    c.close()


def _step_pair(orc, dml, s, c, rng, mode, nsteps, brown=True, gcmc=False, r1s=None):
    """integrator + test_update + overlap + test_update + F->CG, on both"""
    for k in range(nsteps):
        nref = s.nref
        if brown:
            nz = _noise(rng, nref, 3)
            ud = rng.random(nref)
            assert s.brownian(nz, ud) == 0
            c.brownian(nz, ud)
        else:
            nz = _noise(rng, nref, 6)
            ud = rng.random(nref)
            assert s.ermak_a(nz, ud) == 0
            c.ermak_a(nz, ud)
            s.fuerza(); c.forces()
            s.ermak_b(); c.ermak_b()
        st, r1 = s.test_update()
        assert st == 0
        r1d = c.test_update()[0]
        assert r1 == r1d, k
        if r1s is not None:
            r1s.append(r1)
        if mode == "exact":
            u = rng.random(4096)
            st, used = s.overlap(u)
            assert st == 0
            c.overlap(u)
        else:
            u = rng.random((16, c.total))
            assert s.overlap_rounds(u[:, :s.n].reshape(-1), 16) == 0
            c.overlap_rounds(u, 16)
        st, r2 = s.test_update()
        assert st == 0
        assert r2 == c.test_update()[0], k
        s.f_to_cg(gcmc); c.f_to_cg()
        so, sd = s.state(), c.state()
        np.testing.assert_array_equal(sd["species"], so["species"], err_msg="step %d" % k)
        np.testing.assert_array_equal(sd["flags"], so["flags"], err_msg="step %d" % k)
        np.testing.assert_array_equal(c.ref_order(), s.ref_order())
        assert rel_err(sd["pos"], so["pos"]) <= 1e-12, k
    return so, sd


@pytest.mark.parametrize("mode", ["exact", "rounds", "rounds_lpa1", "rounds_tail"])
def test_brownian_steps_with_overlap_match_oracle(orc, dml, mode, monkeypatch):
    """several whole steps: the exact resolver follows the reference order
    bit for bit; the round schedule follows its own oracle, also with every
    round after the first in the one-workgroup tail kernel (rounds_tail)."""
    # h and the skin are chosen like the reference's own runs (skin 10 with
    # steps of ~1 A): moved-back atoms then stay inside the list radius
    rng = np.random.default_rng(21)
    box = np.array([40.0, 40.0, 60.0])
    el = electrode(3.51, 11, 11, 2, 3.51)
    ions = rsa(rng, 700, np.array([0, 0, 6.0]), np.array([40, 40, 59.0]), 3.2, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    say("steps", mode, len(pos))
    if mode == "rounds_lpa1":
        monkeypatch.setenv("DML_OVERLAP_LPA", "1")
    if mode == "rounds_tail":
        monkeypatch.setenv("DML_OV_TAIL", "1")
    s, c = pair(orc, dml, box, 10.0, pos, z, ermak=False, zmax=60.0, z0=30.0, h=1e-3,
                overlap_mode=dml.OVERLAP_EXACT if mode == "exact" else dml.OVERLAP_ROUNDS)
    s.test_update(); c.test_update()
    so, sd = _step_pair(orc, dml, s, c, rng, "exact" if mode == "exact" else "rounds", 12)
    cnt_o, cnt_d = s.c, c.counters()
    assert cnt_o.choques == cnt_d["choques"]
    assert cnt_o.depo == cnt_d["depo"]
    assert cnt_o.choques > 0
    c.close()


def test_ermak_piston_steps_match_oracle(orc, dml):
    box, skin, pos, z = c1_system(seed=7, n_ion=300)
    rng = np.random.default_rng(22)
    say("ermak piston steps", len(pos))
    s, c = pair(orc, dml, box, skin, pos, z, zmax=100.0, z0=50.0, piston=True, prob=0.8)
    s.test_update(); c.test_update()
    s.fuerza(); c.forces()
    rho0 = s.calc_rho()
    for k in range(6):
        so, sd = _step_pair(orc, dml, s, c, rng, "exact", 1, brown=False)
        # the piston reservoir of the step (dana.F90:244, 261): calc_rho
        # (dana.F90:521-549) counts the same atoms, maxz (dana.F90:776-794)
        # moves the atoms above z0 and zmax by the same factor
        rho_o = s.calc_rho()
        rho_d = c.calc_rho()[0]
        assert rho_d == rho_o, (k, rho_d, rho_o)
        s.maxz(rho_o, rho0)
        c.maxz()
        assert c.zmax() == s.reservoir()[2], k
        so, sd = s.state(), c.state()
        assert rel_err(sd["pos"], so["pos"]) <= 1e-12, k
        np.testing.assert_array_equal(sd["pos"][:, 2] > 50.0, so["pos"][:, 2] > 50.0)
    assert rel_err(sd["vel"][so["species"] != 2], so["vel"][so["species"] != 2]) <= 1e-12
    c.close()


def test_ermak_steps_with_round_overlap_match_oracle(orc, dml):
    """Ermak steps with the round schedule: fuerza runs between the move and
    the overlap pass, so on steps whose first skin test does not rebuild,
    round 0 takes the atoms whose force row held nothing within rcut as
    resolved without scanning them; the oracle scans every atom."""
    rng = np.random.default_rng(31)
    box = np.array([40.0, 40.0, 60.0])
    el = electrode(3.51, 11, 11, 2, 3.51)
    ions = rsa(rng, 700, np.array([0, 0, 6.0]), np.array([40, 40, 59.0]), 3.2, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    say("ermak round steps", len(pos))
    s, c = pair(orc, dml, box, 10.0, pos, z, zmax=60.0, z0=30.0, h=1e-2,
                overlap_mode=dml.OVERLAP_ROUNDS)
    s.test_update(); c.test_update()
    s.fuerza(); c.forces()
    r1s = []
    so, sd = _step_pair(orc, dml, s, c, rng, "rounds", 10, brown=False, r1s=r1s)
    assert rel_err(sd["vel"][so["species"] != 2], so["vel"][so["species"] != 2]) <= 1e-12
    assert 0 in r1s  # some step kept its rows, so the shortcut was open
    cnt_o, cnt_d = s.c, c.counters()
    assert cnt_o.choques == cnt_d["choques"]
    assert cnt_o.choques > 0
    c.close()


@pytest.mark.parametrize("lb_mode,wide,lpa", [("3", False, "8"), ("1", False, "8"),
                                              ("0", False, "8"), ("3", True, "8"),
                                              ("1", True, "8"), ("0", True, "8"),
                                              ("3", True, "1"), ("0", True, "2"),
                                              ("4", False, "8"), ("4", True, "8")])
def test_brownian_round_overlap_after_rebuilds_matches_oracle(orc, dml, lb_mode, wide, lpa,
                                                              monkeypatch):
    """Brownian steps (h 3e-3: some first skin tests rebuild, some do not)
    with the round schedule, on rows from the wave-per-cell (lb_mode 3), the
    wave-per-atom (lb_mode 1) and the cell-tile (lb_mode 0) row builds.  The
    cell-tile build, and the wave-per-cell build in the wide box (4 cells of
    13.2 A across x and y), list each atom's entries within rcut, and round 0
    after a rebuild visits only those."""
    monkeypatch.setenv("DML_LISTBUILD", lb_mode)
    monkeypatch.setenv("DML_OVERLAP_LPA", lpa)
    rng = np.random.default_rng(41 if wide else 31)
    L, nel, nion = (60.0, 17, 1600) if wide else (40.0, 11, 700)
    box = np.array([L, L, 60.0])
    assert (int(L // 13.2) >= 4) == wide
    el = electrode(3.51, nel, nel, 2, 3.51)
    ions = rsa(rng, nion, np.array([0, 0, 6.0]), np.array([L, L, 59.0]), 3.2, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    say("brownian round steps", len(pos), "lb_mode", lb_mode, "wide", wide, "lpa", lpa)
    s, c = pair(orc, dml, box, 10.0, pos, z, ermak=False, zmax=60.0, z0=30.0, h=3e-3,
                overlap_mode=dml.OVERLAP_ROUNDS)
    s.test_update(); c.test_update()
    r1s = []
    _step_pair(orc, dml, s, c, rng, "rounds", 10, r1s=r1s)
    assert 0 in r1s and 1 in r1s
    cnt_o, cnt_d = s.c, c.counters()
    assert cnt_o.choques == cnt_d["choques"]
    assert cnt_o.depo == cnt_d["depo"]
    assert cnt_o.choques > 0
    c.close()


def test_set_order_after_forces_matches_recorded_overlap_pass(dml):
    """dml_set_order after fuerza drops the force kernel's entries-within-rcut
    records, so round 0 of the next overlap pass scans every ref atom; with the
    same order handed back, that pass must end exactly where the pass that
    used the records ends."""
    rng = np.random.default_rng(51)
    box = np.array([60.0, 60.0, 60.0])
    el = electrode(3.51, 17, 17, 2, 3.51)
    # ions 3.0 A apart at least: shallow overlaps (r0 3.2 A) that rounds resolve
    ions = rsa(rng, 1600, np.array([0, 0, 6.0]), np.array([60, 60, 59.0]), 3.0, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    say("set_order after forces", len(pos))
    cfg = dml.dana_config(box, 10.0, len(pos) + 64, h=1e-2, z0=30.0, zmax=60.0,
                          integrator=dml.ERMAK, overlap_mode=dml.OVERLAP_ROUNDS)
    a, b = dml.Context(cfg), dml.Context(cfg)
    for c in (a, b):
        c.set_state(pos, species=z)
        c.test_update()
        c.forces()
    a.set_order(ref_order=a.ref_order())
    u = rng.random((32, a.total))
    for c in (a, b):
        c.overlap_rounds(u, 32)
    sa, sb = a.state(), b.state()
    for k in ("pos", "species", "flags"):
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    np.testing.assert_array_equal(a.ref_order(), b.ref_order())
    assert a.counters()["choques"] == b.counters()["choques"] > 0
    a.close()
    b.close()


def test_gcmc_trial_energy_matches_oracle(orc, dml):
    rng = np.random.default_rng(31)
    box = np.array([60.0, 60.0, 60.0])
    pos = rsa(rng, 900, np.zeros(3), box, 2.8, box)
    z = np.where(rng.random(len(pos)) < 0.2, 2, 1).astype(np.int8)
    say("gcmc energy", len(pos))
    s = orc.System(orc.dana_params(box, 2.0, ermak=False, piston=False), pos, z)
    cfg = dml.dana_config(box, 2.0, len(pos) + 64, integrator=dml.BROWNIAN, reservoir=dml.GCMC)
    c = dml.Context(cfg)
    flags = (1 | np.where(z != 2, 4, 0) | np.where(z == 1, 16, 0)).astype(np.uint8)
    c.set_state(pos, species=z, flags=flags)
    s.test_update(); c.test_update()
    # the oracle's gcmc membership: set by flags
    fo = s.state()["flags"]
    s.set_state(flags=flags)
    trials = rng.random((2000, 3)) * box
    du_d, hard_d = c.gcmc_trial_energy(trials)
    # oracle over all atoms with gcmc membership from the flags
    import ctypes
    so = s.state()
    d = trials[:, None, :] - pos[None, :, :]
    d[..., :2] -= box[:2] * np.round(d[..., :2] / box[:2])
    r2 = (d * d).sum(-1)
    hard_o = ((r2 < 3.2 * 3.2) & (z == 1)[None, :]).any(1)
    np.testing.assert_array_equal(hard_d.astype(bool), hard_o)
    r0 = np.where(z == 1, 3.2, 3.5); eps = np.where(z == 1, 2313.6, 0.0)
    inside = r2 <= r0 ** 2
    rr = np.sqrt(np.where(inside, r2, 1.0))
    b = (r0 / rr) ** 6
    e = np.where(inside, eps * b * (b - 2) + eps, 0.0).sum(1)
    np.testing.assert_allclose(du_d, e, rtol=1e-10, atol=1e-9)
    c.close()


def test_run_graph_equals_eager_and_conserves(dml):
    """Philox steps: graph replays give the same bits as eager launches; atom
    count and species bookkeeping hold; every step advances."""
    rng = np.random.default_rng(41)
    box = np.array([60.0, 60.0, 80.0])
    el = electrode(3.51, 17, 17, 2, 3.51)
    ions = rsa(rng, 1500, np.array([0, 0, 6.0]), np.array([60, 60, 79.0]), 3.2, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    outs = []
    for graph in (False, True):
        for ermak in (False, True):
            say("run graph", graph, "ermak", ermak)
            cfg = dml.dana_config(box, 10.0, len(pos) + 64, integrator=dml.ERMAK if ermak else dml.BROWNIAN,
                                  reservoir=dml.NO_RESERVOIR, overlap_mode=dml.OVERLAP_ROUNDS,
                                  zmax=80.0, z0=40.0, seed=77, h=1e-3)
            c = dml.Context(cfg)
            c.set_state(pos, species=z)
            c.run(20, graph=graph)
            st = c.state()
            outs.append(st)
            assert np.isfinite(st["pos"]).all()
            assert (st["pos"][:, :2] >= 0).all() and (st["pos"][:, :2] < box[:2] + 1e-9).all()
            assert (st["pos"][:, 2] > 0).all()
            assert set(np.unique(st["species"])) <= {1, 2}
            nm, reb, steps = c.list_stats()
            assert steps == 20 and reb > 0
            c.close()
    for a, b in ((outs[0], outs[2]), (outs[1], outs[3])):
        np.testing.assert_array_equal(a["pos"], b["pos"])
        np.testing.assert_array_equal(a["species"], b["species"])


def test_ensemble_replicas_match_single_runs(dml):
    """replicas of different size and integrator in one context give the same
    bits as each replica run alone (noise is keyed per replica seed)."""
    rng = np.random.default_rng(51)
    systems = []
    for k, (n, L) in enumerate(((300, 40.0), (900, 60.0), (500, 50.0))):
        box = np.array([L, L, 60.0])
        pos = rsa(rng, n, np.array([0, 0, 3.0]), np.array([L, L, 59.0]), 3.2, box)
        systems.append((box, pos))
    cfgs = [dml.dana_config(b, 10.0, len(p) + 64, integrator=dml.ERMAK if k % 2 else dml.BROWNIAN,
                            reservoir=dml.NO_RESERVOIR, overlap_mode=dml.OVERLAP_ROUNDS,
                            zmax=60.0, z0=30.0, seed=1000 + k, h=1e-3,
                            precision=dml.FP32_FORCE if k == 1 else dml.FP64)
            for k, (b, p) in enumerate(systems)]
    say("ensemble of", len(cfgs))
    ens = dml.Context(cfgs)
    for k, (b, p) in enumerate(systems):
        ens.set_state(p, rep=k)
    ens.run(15, graph=True)
    ob_ens = ens.observables()
    for k, (b, p) in enumerate(systems):
        one = dml.Context(cfgs[k])
        one.set_state(p)
        one.run(15, graph=True)
        a, o = ens.state(rep=k), one.state()
        np.testing.assert_array_equal(a["pos"], o["pos"])
        np.testing.assert_array_equal(a["species"], o["species"])
        np.testing.assert_array_equal(ob_ens[k], one.observables()[0])
        one.close()
    ens.close()


def test_fatal_condition_halts_only_its_replica(dml):
    """deposit above z0 (dana.F90:905-908) stops the reference; in an ensemble
    it halts that replica only.  Replica 1 gets z0 below every ion, so its
    first CG contact is fatal; replica 0 keeps stepping."""
    rng = np.random.default_rng(61)
    box = np.array([40.0, 40.0, 60.0])
    el = electrode(3.51, 11, 11, 2, 3.51)
    ions = rsa(rng, 500, np.array([0, 0, 6.5]), np.array([40, 40, 59.0]), 3.2, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    cfgs = [dml.dana_config(box, 10.0, len(pos) + 64, integrator=dml.BROWNIAN, h=1e-2,
                            reservoir=dml.NO_RESERVOIR, overlap_mode=dml.OVERLAP_ROUNDS,
                            zmax=60.0, z0=z0, seed=5 + k) for k, z0 in enumerate((60.0, 0.5))]
    say("halt test")
    ens = dml.Context(cfgs)
    ens.set_state(pos, species=z, rep=0)
    ens.set_state(pos, species=z, rep=1)
    ens.run(30, graph=True)
    assert ens.replica_status(0) == 0
    assert ens.replica_status(1) == -6  # DML_E_DEPOSIT_Z0
    s0 = ens.list_stats(0)[2]
    s1 = ens.list_stats(1)[2]
    assert s0 == 30 and s1 < 30
    frozen = ens.state(rep=1)["pos"]
    ens.run(5, graph=True)
    assert ens.list_stats(0)[2] == 35
    np.testing.assert_array_equal(ens.state(rep=1)["pos"], frozen)
    # set_state clears the halt
    ens.set_state(pos, species=z, rep=1)
    assert ens.replica_status(1) == 0
    ens.close()


def _assert_same_groups(s, c, tag):
    so, sd = s.state(), c.state()
    assert c.nslots() == s.n, tag
    np.testing.assert_array_equal(sd["flags"], so["flags"], err_msg=tag)
    np.testing.assert_array_equal(sd["species"], so["species"], err_msg=tag)
    for k in ("pos", "pos_old", "vel", "old_cg"):
        assert rel_err(sd[k], so[k]) <= 1e-12, (tag, k)
    np.testing.assert_array_equal(c.ref_order(), s.ref_order(), err_msg=tag)
    np.testing.assert_array_equal(c.gcmc_order(), s.gcmc_order(), err_msg=tag)
    alive = (so["flags"] & 1) != 0
    np.testing.assert_array_equal(c.slots()[0][alive], s.b_slots()[alive], err_msg=tag)
    assert_lists_equal(c, s)


@pytest.mark.parametrize("geom", ["cells", "verlet"])
def test_gcmc_trials_match_oracle(orc, dml, geom):
    """gcmc_run with injected uniforms/normals: accept/reject decisions, slot
    and hs%b reuse, limbo, push-front into the cell chains, incremental rows
    and the gcmc list order, then Brownian steps over the changed system."""
    rng = np.random.default_rng(41 if geom == "cells" else 42)
    if geom == "cells":
        box = np.array([40.0, 40.0, 60.0]); nx, n_ion, nadj = 11, 160, 80   # 7 x 7 x 11 cells
    else:
        box = np.array([15.0, 15.0, 15.0]); nx, n_ion, nadj = 4, 12, 30     # 2 x 2 x 2: Verlet
    el = electrode(3.51, nx, nx, 2, 1.755)
    ions = rsa(rng, n_ion, np.array([0, 0, 4.0]), np.array([box[0], box[1], box[2] - 1.0]), 3.3, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    # z0 above every possible deposit (electrode top + rcut): no fatal stop
    z0, zmax = max(box[2] / 3.0, 7.0), box[2] - 0.5
    say("gcmc", geom, len(pos))
    # h = 5e-5 with skin 2, as in the reference's gcmc run
    s = orc.System(orc.dana_params(box, 2.0, h=5e-5, z0=z0, zmax=zmax, ermak=False, piston=False),
                   pos, z)
    cfg = dml.dana_config(box, 2.0, len(pos) + 512, h=5e-5, z0=z0, zmax=zmax,
                          integrator=dml.BROWNIAN, reservoir=dml.GCMC,
                          overlap_mode=dml.OVERLAP_EXACT)
    c = dml.Context(cfg)
    c.set_state(pos, species=z)
    s.test_update(); c.test_update()
    s.gcmc_init()
    assert c.cells()[2] == (geom == "cells")
    _assert_same_groups(s, c, "initial")
    vol = box[0] * box[1] * (zmax - z0)
    n_in = int(((ions[:, 2] >= z0) & (ions[:, 2] <= zmax)).sum())
    act = max(n_in, 1) / vol  # accept about half of each kind
    for it in range(4):
        u = rng.random((nadj, 5))
        g = rng.standard_normal((nadj, 3))
        assert s.gcmc(act, nadj, u, g) == 0
        c.gcmc(act, nadj, u, g)
        _assert_same_groups(s, c, "gcmc call %d" % it)
        assert (s.c.ins, s.c.del_) == (c.counters()["ins"], c.counters()["del_"])
        _step_pair(orc, dml, s, c, rng, "exact", 2, gcmc=True)
        _assert_same_groups(s, c, "steps after call %d" % it)
    assert s.c.ins > 5 and s.c.del_ > 5, (s.c.ins, s.c.del_)
    c.close()


def test_gcmc_run_alone_equals_in_ensemble_and_graph(dml):
    """whole steps with the GCMC reservoir on Philox draws: the replica's
    result does not depend on its neighbours or on graph replay."""
    rng = np.random.default_rng(43)
    box = np.array([40.0, 40.0, 60.0])
    el = electrode(3.51, 11, 11, 2, 1.755)
    ions = rsa(rng, 160, np.array([0, 0, 4.0]), np.array([40, 40, 59.0]), 3.3, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    z0, zmax = 20.0, 59.5
    act = 100.0 / (box[0] * box[1] * (zmax - z0))

    def cfg(seed, reservoir=dml.GCMC):
        return dml.dana_config(box, 2.0, len(pos) + 1024, h=5e-5, z0=z0, zmax=zmax,
                               integrator=dml.BROWNIAN, reservoir=reservoir, seed=seed,
                               overlap_mode=dml.OVERLAP_ROUNDS, act=act, nadj=16)
    say("gcmc run", len(pos))
    a = dml.Context(cfg(7))
    a.set_state(pos, species=z)
    a.run(25, graph=False)
    assert a.replica_status(0) == 0
    sa = a.state()
    ca = a.counters()
    assert ca["ins"] > 0 and ca["del_"] > 0, ca
    b = dml.Context([cfg(99, dml.NO_RESERVOIR), cfg(7)])
    b.set_state(pos, species=z, rep=0)
    b.set_state(pos, species=z, rep=1)
    b.run(25, graph=True)
    assert b.replica_status(1) == 0
    sb = b.state(1)
    assert b.nslots(1) == a.nslots(0)
    for k in ("pos", "vel", "species", "flags"):
        np.testing.assert_array_equal(sb[k], sa[k], err_msg=k)
    np.testing.assert_array_equal(b.gcmc_order(1), a.gcmc_order(0))
    assert b.counters(1) == ca
    a.close(); b.close()


def test_row_overflow_grows_and_rebuilds(orc, dml):
    """rows longer than max_neighbors: a per-routine rebuild halts the
    replica in the kernel, then the host grows the rows and replays it"""
    box, skin, pos, z = c1_system(seed=9)
    s = orc.System(orc.dana_params(box, skin, ermak=False, piston=False), pos, z)
    cfg = dml.dana_config(box, skin, len(pos) + 64, integrator=dml.BROWNIAN,
                          reservoir=dml.NO_RESERVOIR, max_neighbors=16)
    say("overflow", len(pos))
    c = dml.Context(cfg)
    c.set_state(pos, species=z)
    s.test_update(); c.test_update()
    assert c.replica_status(0) == 0
    assert_lists_equal(c, s)
    nn, _, _ = c.lists()
    assert nn.max() > 16
    s.fuerza(); c.forces()  # the grown rows serve the force loop
    assert rel_err(c.state()["force"], s.state()["force"]) <= 1e-12
    c.close()


@pytest.mark.parametrize("band", ["fires", "holds"])
def test_chunks_reservoir_matches_oracle(orc, dml, band):
    """calc_rho up to the sensor z1 and bloques: the density test, slot and
    hs%b reuse of the new block, the reservoir and box moving up by dist,
    the regrown tessellation and the rebuilt rows, over whole steps."""
    rng = np.random.default_rng(51 if band == "fires" else 52)
    box0 = np.array([40.0, 40.0, 0.0])
    z0, dist = 20.0, 10.0 + 3.2  # dist + rcut (config_inic)
    z1 = z0 + dist
    zmax = z1 + dist
    box = np.array([40.0, 40.0, zmax])
    el = electrode(3.51, 11, 11, 2, 1.755)
    if band == "fires":   # sensor density far above rhomedia
        ions = rsa(rng, 200, np.array([0, 0, 4.0]), np.array([40, 40, zmax - 1.0]), 3.3, box)
    else:                 # 12 ions in (z0, z1): rho within 18.6 % of rhomedia
        low = rsa(rng, 60, np.array([0, 0, 4.0]), np.array([40, 40, z0 - 1.0]), 3.3, box)
        mid = rsa(rng, 12, np.array([0, 0, z0 + 2.0]), np.array([40, 40, z1 - 2.0]), 3.3, box)
        ions = np.vstack([low, mid])
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    chunk = rsa(rng, 30, np.zeros(3), np.array([40, 40, dist - 0.5]), 3.3, box)
    say("chunks", band, len(pos), len(chunk))
    s = orc.System(orc.dana_params(box, 2.0, h=5e-5, z0=z0, zmax=zmax, ermak=False, piston=False),
                   pos, z)
    s.set_z1(z1)
    s.chunk_init(chunk, dist)
    cfg = dml.dana_config(box, 2.0, len(pos) + 8 * len(chunk) + 64, h=5e-5, z0=z0, zmax=zmax,
                          integrator=dml.BROWNIAN, reservoir=dml.CHUNKS,
                          overlap_mode=dml.OVERLAP_EXACT)
    c = dml.Context(cfg)
    c.set_state(pos, species=z)
    c.set_chunk(chunk, dist, z1)
    s.test_update(); c.test_update()
    fired = 0
    for it in range(3):
        _step_pair(orc, dml, s, c, rng, "exact", 1)
        rho_o = s.calc_rho(True)
        rho_d = c.calc_rho()[0]
        assert rho_d == rho_o, (it, rho_d, rho_o)
        st, f = s.bloques(rho_o)
        assert st == 0
        c.bloques()
        fired += f
        res = s.reservoir()
        assert c.zmax() == res[2], it
        nc_o, cell_o, uc_o, ci_o = s.cells()
        nc_d, cell_d, uc_d, ci_d = c.cells()
        np.testing.assert_array_equal(nc_d, nc_o)
        np.testing.assert_array_equal(cell_d, cell_o)
        if f:  # the device sorts its buckets when it rebuilds (the reference at every test_update)
            np.testing.assert_array_equal(ci_d, ci_o)
        _assert_same_groups(s, c, "bloques %d" % it)
    assert fired == (3 if band == "fires" else 0)
    c.close()


def test_round_overlap_with_subset_ref_order_matches_restatement(dml):
    """dml_set_order with a shuffled ref list that leaves some ions out: the
    flags follow the list (REF on exactly the listed atoms), an ion outside
    the list is reverted like any non-CG neighbour (dana.F90:928-934) and
    waits for nobody; the device's round schedule equals the restatement of
    oracle/ovl.py (itself equal to orc_overlap_rounds) on the device's own
    state before the call, with the same uniforms.  A repeated entry is
    refused."""
    import ovl
    rng = np.random.default_rng(61)
    box = np.array([40.0, 40.0, 30.0])
    el = electrode(3.51, 11, 11, 2, 1.755)
    ions = rsa(rng, 220, np.array([0, 0, 4.0]), np.array([40, 40, 29.0]), 3.3, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    say("subset ref order", len(pos))
    c = dml.Context(dml.dana_config(box, 2.0, len(pos) + 64, h=1e-2, zmax=30.0, prob=0.6,
                                    integrator=dml.BROWNIAN, reservoir=dml.NO_RESERVOIR,
                                    overlap_mode=dml.OVERLAP_ROUNDS))
    c.set_state(pos, species=z)
    c.test_update()
    n = len(pos)
    ion_idx = np.arange(len(el), n)
    c.brownian(rng.standard_normal(3 * n), np.ones(n))
    c.test_update()
    keep = ion_idx[rng.random(len(ion_idx)) > 0.2]
    order = rng.permutation(keep).astype(np.int32)
    with pytest.raises(dml.DmlError):
        c.set_order(ref_order=np.concatenate([order[:5], order[:1]]))
    c.set_order(ref_order=order)
    st = c.state()
    np.testing.assert_array_equal((st["flags"] & 4) != 0, np.isin(np.arange(n), order))
    nn, rp, idx = c.lists()
    m = dict(pos=st["pos"].copy(), old=st["old_cg"].copy(), vel=st["vel"].copy(),
             acel=st["acel"].copy(), zs=st["species"].astype(int), skip=(st["flags"] & 8) != 0,
             rows=[list(idx[rp[i]:rp[i + 1]]) for i in range(n)], order=list(order),
             limbo=(st["flags"] & 2) != 0, box=box, rcut=3.2, prob=0.6, z0=100.0, piston=False)
    R = 32
    ur = rng.random((R, c.total))
    want, wc, wst = ovl.round_overlap(m, lambda i, k, r: ur[r, i], R)
    assert wst == 0
    c0 = c.counters()
    c.overlap_rounds(ur, R)
    got = c.state()
    outside = np.setdiff1d(ion_idx, order)
    moved_out = [i for i in outside if not np.array_equal(want["pos"][i], m["pos"][i])]
    assert moved_out, "no ion outside the ref list was reverted: the case does not test it"
    np.testing.assert_array_equal(got["pos"], want["pos"])
    np.testing.assert_array_equal(got["species"].astype(int), want["zs"])
    np.testing.assert_array_equal((got["flags"] & 8) != 0, want["skip"])
    np.testing.assert_array_equal(got["vel"], want["vel"])
    c1 = c.counters()
    assert (c1["try_"] - c0["try_"], c1["choques"] - c0["choques"]) == (wc["try_"], wc["choques"])
    c.close()


def test_device_brownian_noise_is_the_philox_stream(dml):
    """dml_brownian without injected noise draws, for the atom of particle id
    p at step s, gauss2(philox(seed, p, s, RS_BROWN, 0)) and the first of
    gauss2(... draw 1) (cbrownian_hs's three gasdev, dana.F90:823-833): the
    normals recovered from one step of well-separated atoms, D small enough
    that nothing touches, equal the host restatement of tests/philox_host.py
    (pinned by the SC11 vectors) within 1e-12."""
    import philox_host as ph
    rng = np.random.default_rng(71)
    box = np.array([60.0, 60.0, 60.0])
    pos = rsa(rng, 300, np.array([5.0, 5.0, 10.0]), np.array([55.0, 55.0, 50.0]), 4.0, box)
    z = np.ones(len(pos), np.int8)
    seed = 0x1234ABCD5678
    say("philox stream", len(pos))
    cfg = dml.dana_config(box, 2.0, len(pos) + 64, h=1e-2, zmax=60.0, dif_sc=1.0, dif_sei=1.0,
                          integrator=dml.BROWNIAN, reservoir=dml.NO_RESERVOIR, seed=seed,
                          overlap_mode=dml.OVERLAP_ROUNDS)
    c = dml.Context(cfg)
    c.set_state(pos, species=z)
    c.test_update()
    for step in range(3):
        before = c.state()["pos"]
        c.brownian()
        after = c.state()["pos"]
        xi = (after - before) / np.sqrt(2.0 * 1.0 * 1e-2)
        want = np.zeros_like(xi)
        for i in range(len(pos)):
            g0, g1 = ph.gauss2(ph.draw(seed, i, step, ph.RS_BROWN, 0))
            g2, _ = ph.gauss2(ph.draw(seed, i, step, ph.RS_BROWN, 1))
            want[i] = (g0, g1, g2)
        assert np.max(np.abs(xi - want)) <= 1e-12, step
        c.end_step()
    c.close()
