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
