"""The dendrite-regime configuration, scaled down: a 12-layer electrode with
hemispherical protrusions and ions at 0.01 A^-3 (the generator of the 2M
system with a smaller base), whole GCMC steps on the device, then batched
trial energies of insertions and removals against the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_c5_scaled_trial_energies_match_oracle(orc, dml):
    from dml import workloads as W
    sysd = W.system_c5(n_total=60_000, seed=77, L=120.0, n_prot=4)
    box, pos, z = sysd["box"], sysd["pos"], sysd["species"]
    z0 = 0.75 * box[2]
    print("c5 scaled", len(pos), box, flush=True)
    cfg = dml.dana_config(box, sysd["skin"], len(pos) + 4096, h=sysd["h"], z0=z0, zmax=box[2],
                          integrator=dml.BROWNIAN, reservoir=dml.GCMC, precision=dml.FP64,
                          overlap_mode=dml.OVERLAP_ROUNDS, act=0.01, nadj=10, seed=5)
    c = dml.Context(cfg)
    c.set_state(pos, species=z)
    c.run(20, graph=True)
    assert c.replica_status(0) == 0
    st = c.state()
    alive = np.nonzero(st["flags"] & 1)[0]
    rng = np.random.default_rng(5)
    ins = np.column_stack([rng.random(512) * box[0], rng.random(512) * box[1],
                           z0 + rng.random(512) * (box[2] - z0)])
    gc = c.gcmc_order()
    rem = rng.choice(gc, 512, replace=False)
    pts = np.vstack([ins, st["pos"][rem]])
    ex = np.concatenate([-np.ones(512, np.int32), rem.astype(np.int32)])
    du_d, hard_d = c.gcmc_trial_energy(pts, exclude=ex)
    # the oracle on the same atoms (slot order kept: dead slots sit far away)
    p = st["pos"].copy()
    p[st["flags"] & 1 == 0] = [1e7, 1e7, 1e7]
    s = orc.System(orc.dana_params(box, sysd["skin"], ermak=False, piston=False), p, st["species"])
    s.set_state(flags=st["flags"])
    s.gcmc_init()
    du_o, hard_o = s.gcmc_trial_energy(pts, exclude=ex)
    np.testing.assert_array_equal(hard_d, hard_o)
    np.testing.assert_allclose(du_d, du_o, rtol=1e-12, atol=1e-9)
    # insertions near an ion cost energy; a removal's energy is zero here,
    # since after the overlap resolution no Li pair is closer than r0 = rcut
    # and pairs with CG have eps 0
    assert (hard_o[:512] == 1).any() and (du_o[:512] > 0).any()
    assert (hard_o[512:] == 0).all() and (du_o[512:] == 0).all()
    c.close()


def test_c5_full_system_trial_energies_on_subvolume(orc, dml):
    """The 2M-particle system itself: 3 GCMC steps on the device, then 512
    insertion and 512 removal trial energies inside a 170 x 170 A column;
    the FP64 oracle holds only the atoms of the 200 x 200 A column around it
    (every atom within the pair cutoff of a trial point is there)."""
    from dml import workloads as W
    sysd = W.system_c5()
    box, pos, z = sysd["box"], sysd["pos"], sysd["species"]
    z0 = 0.75 * box[2]
    print("c5 full", len(pos), box, flush=True)
    cfg = dml.dana_config(box, sysd["skin"], len(pos) + 8192, h=sysd["h"], z0=z0, zmax=box[2],
                          integrator=dml.ERMAK, reservoir=dml.GCMC, precision=dml.FP32_FORCE,
                          overlap_mode=dml.OVERLAP_ROUNDS, act=0.01, nadj=10, seed=11)
    c = dml.Context(cfg)
    c.set_state(pos, species=z)
    c.run(3, graph=True)
    assert c.replica_status(0) == 0
    st = c.state()
    alive = (st["flags"] & 1) != 0
    p = st["pos"]
    lo, hi = 200.0, 400.0
    inner = (p[:, 0] > lo + 15) & (p[:, 0] < hi - 15) & (p[:, 1] > lo + 15) & (p[:, 1] < hi - 15)
    rng = np.random.default_rng(11)
    ins = np.column_stack([rng.uniform(lo + 15, hi - 15, 512), rng.uniform(lo + 15, hi - 15, 512),
                           z0 + rng.random(512) * (box[2] - z0)])
    gc = c.gcmc_order()
    cand = gc[inner[gc]]
    rem = rng.choice(cand, 512, replace=False).astype(np.int32)
    du_i, hard_i = c.gcmc_trial_energy(ins)
    du_r, hard_r = c.gcmc_trial_energy(None, exclude=rem)
    c.close()
    # the oracle on the column's atoms
    col = alive & (p[:, 0] >= lo) & (p[:, 0] < hi) & (p[:, 1] >= lo) & (p[:, 1] < hi)
    idx = np.nonzero(col)[0]
    back = -np.ones(len(p), np.int64)
    back[idx] = np.arange(len(idx))
    s = orc.System(orc.dana_params(box, sysd["skin"], ermak=False, piston=False), p[idx],
                   st["species"][idx])
    s.set_state(flags=st["flags"][idx])
    s.gcmc_init()
    du_o, hard_o = s.gcmc_trial_energy(np.vstack([ins, p[rem]]),
                                       exclude=np.concatenate([-np.ones(512, np.int32),
                                                               back[rem].astype(np.int32)]))
    np.testing.assert_array_equal(np.concatenate([hard_i, hard_r]), hard_o)
    np.testing.assert_allclose(np.concatenate([du_i, du_r]), du_o, rtol=1e-12, atol=1e-9)
    assert (hard_o[:512] == 1).any() and (hard_o[:512] == 0).any()
