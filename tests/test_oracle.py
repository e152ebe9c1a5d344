"""Pinning the CPU oracle before it is trusted as the checker.

* whole dana runs reproduce the reference's three golden frames bit for bit
  (tests/{ermak,brown,gcmc}/ref.xyz, stored parsed in tests/golden);
* cells and neighbour lists equal, entry for entry, what the reference's own
  Cells/Neighbor modules build (fixtures written by oracle/_ref/nbdump);
* neighbour sets equal a brute-force O(N^2) search;
* the WCA pair force and the Ermak constants equal their closed forms.
"""
import json
import os

import numpy as np
import pytest

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _run(orc, name, nst=None):
    cfg = json.load(open(os.path.join(G, "ref_runs.json")))[name]
    fr = np.load(os.path.join(G, "ref_frames.npz"))
    chunk = fr[name + "_chunk"] if name + "_chunk" in fr.files else None
    out = orc.dana_run(cfg["idum"], cfg["prob"], cfg["h"], nst or cfg["nst"], cfg["nwr"], cfg["xi"],
                       cfg["yi"], cfg["dist"], cfg["z0"], cfg["zmax"], cfg["dif_sc"], cfg["dif_sei"],
                       cfg["nb_dcut"], cfg["ermak"], cfg["reservoir"], cfg.get("act", 0.0),
                       cfg.get("nadj", 0), chunk=chunk)
    return cfg, fr, out


@pytest.mark.parametrize("name", ["gcmc", "brown", "ermak"])
def test_golden_frame_bit_exact(orc, name):
    cfg, fr, out = _run(orc, name)
    assert out["nat"] == cfg["golden_nat"]
    assert out["zmax"] == cfg["golden_zmax"]
    np.testing.assert_array_equal(out["z"], fr[name + "_z"])
    # bit-exact doubles (the golden text is the shortest round-trip form)
    assert np.array_equal(out["pos"], fr[name + "_pos"])


def test_golden_frame_detects_planted_error(orc):
    """A one-step-short run (a plausible off-by-one) must not match."""
    cfg, fr, out = _run(orc, "gcmc", nst=4999)
    assert not (out["nat"] == cfg["golden_nat"] and np.array_equal(out["pos"], fr["gcmc_pos"]))


def _nb_case(nb, name):
    return {k.split("/", 1)[1]: nb[k] for k in nb.files if k.startswith(name + "/")}


def _oracle_lists(orc, d):
    rc, skin = d["rc_skin"]
    p = orc.dana_params(d["box"], skin, rcut=rc)
    s = orc.System(p, d["pos"], d["z"])
    st, reb = s.test_update()
    assert st == 0 and reb == 1
    return s


NB = np.load(os.path.join(G, "nb_cases.npz"))


@pytest.mark.parametrize("name", list(NB["names"]))
def test_lists_equal_reference_modules(orc, name):
    d = _nb_case(NB, name)
    s = _oracle_lists(orc, d)
    ncells, cell, use_cells, ci = s.cells()
    assert use_cells == bool(d["tess"])
    np.testing.assert_array_equal(ncells, d["ncells"])
    if use_cells:
        np.testing.assert_array_equal(cell, d["cell"])
        np.testing.assert_array_equal(ci, d["ci"])
    nn, rp, idx = s.lists()
    # rows of the reference are ref atoms in ref order; entries in list order
    for r, row in enumerate(d["rows"]):
        ref_entries = d["idx"][d["rowptr"][r]:d["rowptr"][r + 1]]
        mine = idx[rp[row]:rp[row + 1]]
        np.testing.assert_array_equal(mine, ref_entries)
    assert int(nn.sum()) == len(d["idx"])


def _brute(box, rl, pos, z):
    n = len(pos)
    ref = [i for i in range(n) if z[i] != 2]
    out = {}
    for i in ref:
        d = pos - pos[i]
        d[:, :2] -= box[:2] * np.round(d[:, :2] / box[:2])
        r2 = (d * d).sum(1)
        s = set(np.nonzero(r2 < rl * rl)[0].tolist()) - {i}
        out[i] = s
    return out


@pytest.mark.parametrize("name", ["gcmc_like", "c1_like", "dense_slab"])
def test_lists_equal_brute_force_sets(orc, name):
    d = _nb_case(NB, name)
    s = _oracle_lists(orc, d)
    nn, rp, idx = s.lists()
    rc, skin = d["rc_skin"]
    bf = _brute(d["box"], rc + skin, d["pos"], d["z"])
    for i, want in bf.items():
        got = idx[rp[i]:rp[i + 1]].tolist()
        assert len(got) == len(set(got))
        assert set(got) == want, i


def test_ghost_layer_lists_are_asymmetric(orc):
    """Atoms above box(3) sit in the top ghost layer: they find others but
    are found by nobody (Cells.F90:289 + cell_pbc wrap)."""
    d = _nb_case(NB, "ghost_top")
    s = _oracle_lists(orc, d)
    nn, rp, idx = s.lists()
    ghosts = np.nonzero(d["ci"][:, 2] == d["ncells"][2] + 1)[0]
    assert len(ghosts) > 0
    found = set(idx.tolist())
    for g in ghosts:
        assert g not in found
    assert max(nn[g] for g in ghosts) > 0


def test_wca_pair_closed_form(orc):
    """Two Li at distance r < r0: F = 12 eps r0^6/r^7 ((r0/r)^6 - 1) along the
    pair axis, counted twice (full list + reaction), E_i = u(r) (two halves)."""
    box = np.array([50.0, 50.0, 50.0])
    for r in (2.0, 2.9, 3.1999):
        pos = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]])
        p = orc.dana_params(box, 2.0)
        s = orc.System(p, pos, [1, 1])
        s.test_update()
        s.fuerza()
        st = s.state()
        eps, r0 = 2313.6, 3.2
        f = 12 * eps * r0 ** 6 / r ** 7 * ((r0 / r) ** 6 - 1)
        u = eps * (r0 / r) ** 6 * ((r0 / r) ** 6 - 2) + eps
        # (b - 1) cancels near r0: allow the rounding it amplifies, 1e-16/(b-1)
        tol = max(1e-13, 2e-15 / abs((r0 / r) ** 6 - 1))
        np.testing.assert_allclose(st["force"][0], [-2 * f, 0, 0], rtol=tol)
        np.testing.assert_allclose(st["force"][1], [2 * f, 0, 0], rtol=tol)
        np.testing.assert_allclose(st["epot"], [u, u], rtol=1e-12)
    # beyond r0: no force, no energy
    pos = np.array([[10.0, 10.0, 10.0], [13.3, 10.0, 10.0]])
    s = orc.System(orc.dana_params(box, 2.0), pos, [1, 1])
    s.test_update(); s.fuerza()
    assert np.all(s.state()["force"] == 0)


def test_li_cg_pair_has_no_force(orc):
    box = np.array([50.0, 50.0, 50.0])
    pos = np.array([[10.0, 10.0, 10.0], [12.0, 10.0, 10.0]])
    s = orc.System(orc.dana_params(box, 2.0), pos, [1, 2])
    s.test_update(); s.fuerza()
    assert np.all(s.state()["force"][0] == 0)


def test_ermak_constants_closed_form(orc):
    """Allen & Tildesley BD constants for h=1e-2, gamma=1, T=300."""
    box = np.array([50.0, 50.0, 50.0])
    s = orc.System(orc.dana_params(box, 2.0, h=1e-2), np.array([[1.0, 1, 1]]), [1])
    cc0, cc1, cc2, sdr, sdv, crv1, crv2, skt = s.ermak_consts()
    h, g = 1e-2, 1.0
    e = np.exp(-h * g)
    np.testing.assert_allclose(cc0, e, rtol=1e-15)
    np.testing.assert_allclose(cc1, (1 - e) / g, rtol=1e-13)
    np.testing.assert_allclose(cc2, (1 - (1 - e) / (g * h)) / g, rtol=1e-9)  # 1 - cc1/h cancels
    var_r = h / g * (2 - (3 - 4 * e + e * e) / (h * g))
    np.testing.assert_allclose(sdr, np.sqrt(var_r), rtol=1e-9)
    np.testing.assert_allclose(sdv, np.sqrt(1 - e * e), rtol=1e-13)
    np.testing.assert_allclose(crv1, (1 - e) ** 2 / (g * sdr * sdv), rtol=1e-9)
    # series limits for small h: sdr^2 ~ 2h^3/3, crv1 ~ sqrt(3)/2
    np.testing.assert_allclose(sdr ** 2, 2 * h ** 3 / 3, rtol=2e-2)
    np.testing.assert_allclose(crv1, np.sqrt(3) / 2, rtol=2e-2)
    kB = 8.617330350e-5 * 96.485 * 100
    np.testing.assert_allclose(skt, np.sqrt(kB * 300), rtol=1e-15)


def test_brownian_step_closed_form(orc):
    """pos' = pos + sqrt(2 D h) xi with injected xi; D switches at z = 80."""
    box = np.array([100.0, 100.0, 200.0])
    rng = np.random.default_rng(5)
    pos = np.column_stack([rng.random(50) * 100, rng.random(50) * 100, 5 + rng.random(50) * 150])
    p = orc.dana_params(box, 2.0, h=1e-2, dif_sc=250.0, dif_sei=100.0, ermak=False, piston=False, zmax=200.0)
    s = orc.System(p, pos, np.ones(50))
    noise = rng.standard_normal(150) * 0.3
    s.brownian(noise, np.full(50, 0.5))
    got = s.state()["pos"]
    D = np.where(pos[:, 2] > 80, 250.0, 100.0)
    want = pos + np.sqrt(2 * D * 1e-2)[:, None] * noise.reshape(50, 3)
    want[:, :2] = np.where(want[:, :2] > 100, want[:, :2] - 100, want[:, :2])
    want[:, :2] = np.where(want[:, :2] < 0, want[:, :2] + 100, want[:, :2])
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_nr_ran_is_uniform_and_reseeds(orc):
    u = orc.nr_stream(-104012, 20000)
    assert u.min() > 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01
    assert np.array_equal(u[:50], orc.nr_stream(-104012, 50))
    g = orc.nr_stream(-7, 20000, "gauss")
    assert abs(g.mean()) < 0.03 and abs(g.std() - 1) < 0.03


def test_gcmc_trial_energy_brute_force(orc):
    """orc_gcmc_trial_energy against a direct numpy sum of fuerza's pair law
    (Li at the trial point against every atom; r0, eps of the (Li, m) pair),
    for insertion points and for removals (the atom itself left out)."""
    rng = np.random.default_rng(61)
    box = np.array([40.0, 40.0, 40.0])
    pos = rng.random((500, 3)) * box
    z = np.where(rng.random(500) < 0.3, 2, 1).astype(np.int8)
    z[::17] = 3
    s = orc.System(orc.dana_params(box, 2.0, ermak=False, piston=False), pos, z)
    s.gcmc_init()
    pts = np.vstack([rng.random((40, 3)) * box, pos[:40]])
    ex = np.concatenate([-np.ones(40, np.int32), np.arange(40, dtype=np.int32)])
    du, hard = s.gcmc_trial_energy(pts, exclude=ex)
    d = pts[:, None, :] - pos[None, :, :]
    d[..., :2] -= box[:2] * np.round(d[..., :2] / box[:2])
    r2 = (d * d).sum(-1)
    keep = np.ones_like(r2, bool)
    keep[np.arange(40, 80), np.arange(40)] = False
    r0 = np.array([3.2, 3.5, 1.564])[z - 1][None, :]
    eps = np.array([2313.6, 0.0, 529.1])[z - 1][None, :]
    inside = (r2 <= r0 ** 2) & keep
    rr = np.sqrt(np.where(inside, r2, 1.0))
    b = (r0 / rr) ** 6
    e = np.where(inside, eps * b * (b - 2) + eps, 0.0).sum(1)
    np.testing.assert_allclose(du, e, rtol=1e-12, atol=1e-9)
    hard_o = ((r2 < 3.2 ** 2) & keep & (z == 1)[None, :]).any(1)
    np.testing.assert_array_equal(hard.astype(bool), hard_o)


def test_observables_closed_forms(orc):
    """salida's energia and kion (dana.F90:1141-1183, 1342-1376) in the
    oracle: energia after fuerza equals twice the WCA pair energy summed over
    unordered ion pairs (the full list visits each ion pair twice, adding half
    of u to both atoms each time; pairs with CG have eps 0); the temperature
    is T exactly when every ion has m |v|^2 = 3 kB_ui T (equipartition), and
    CG velocities do not enter it."""
    rng = np.random.default_rng(71)
    box = np.array([30.0, 30.0, 30.0])
    pos = rng.random((300, 3)) * box
    z = np.where(rng.random(300) < 0.25, 2, 1).astype(np.int8)
    z[::13] = 3
    s = orc.System(orc.dana_params(box, 2.0, ermak=True, piston=False), pos, z)
    s.test_update()
    s.fuerza()
    ions = z != 2
    d = pos[:, None, :] - pos[None, :, :]
    d[..., :2] -= box[:2] * np.round(d[..., :2] / box[:2])
    r2 = (d * d).sum(-1)
    eps_t = np.array([[2313.6, 0, 529.1], [0, 0, 0], [529.1, 0, 121.0]])
    r0_t = np.array([[3.2, 3.5, 1.564], [3.5, 0, 0], [1.564, 0, 3.61]])
    eps = eps_t[z[:, None] - 1, z[None, :] - 1]
    r0 = r0_t[z[:, None] - 1, z[None, :] - 1]
    iu = np.triu_indices(len(pos), 1)
    inside = (r2[iu] <= r0[iu] ** 2) & ions[iu[0]] & ions[iu[1]]
    rr = np.sqrt(np.where(inside, r2[iu], 1.0))
    b = (r0[iu] / np.where(inside, rr, 1.0)) ** 6
    e_pairs = np.where(inside, eps[iu] * b * (b - 2) + eps[iu], 0.0).sum()
    assert inside.sum() > 5
    ob = s.observables()
    np.testing.assert_allclose(ob[0], 2.0 * e_pairs, rtol=1e-12)
    # equipartition
    kB_ui = 8.617330350e-5 * 96.485 * 100.0
    m = 6.94
    u = rng.normal(size=(len(pos), 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    vel = u * np.sqrt(3.0 * kB_ui * 300.0 / m)
    vel[~ions] *= 50.0  # CG atoms are skipped by kion
    s.set_state(vel=vel)
    ob = s.observables()
    assert ob[2] == ions.sum() and ob[3] == len(pos)
    np.testing.assert_allclose(ob[1], 300.0, rtol=1e-12)
