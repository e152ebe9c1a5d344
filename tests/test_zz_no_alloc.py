"""dml_gcmc_trial_energy_noalloc (the trial-energy entry point that stages
through the library's fixed device array instead of allocating) against
dml_gcmc_trial_energy on the same state: bit-exact, on two calls in a row
(the second smaller, so stale entries of the first would show), for points
with exclusions and for removals at the atoms' own positions, from two
contexts on one GPU; more trials than the array holds is DML_E_NOMEM."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _context(dml, W, seed):
    sysd = W.system_c5(n_total=20_000, seed=seed, L=80.0, n_prot=2)
    box, pos, z = sysd["box"], sysd["pos"], sysd["species"]
    z0 = 0.75 * box[2]
    cfg = dml.dana_config(box, sysd["skin"], len(pos) + 2048, h=sysd["h"], z0=z0, zmax=box[2],
                          integrator=dml.BROWNIAN, reservoir=dml.GCMC, precision=dml.FP64,
                          overlap_mode=dml.OVERLAP_ROUNDS, act=0.01, nadj=10, seed=seed)
    c = dml.Context(cfg)
    c.set_state(pos, species=z)
    c.run(5, graph=True)
    assert c.replica_status(0) == 0
    return c, box, z0


def _trials(c, box, z0, n_ins, n_rem, rng):
    st = c.state()
    ins = np.column_stack([rng.random(n_ins) * box[0], rng.random(n_ins) * box[1],
                           z0 + rng.random(n_ins) * (box[2] - z0)])
    gc = c.gcmc_order()
    assert len(gc) >= 100
    rem = rng.choice(gc, min(n_rem, len(gc)), replace=False).astype(np.int32)
    pts = np.vstack([ins, st["pos"][rem]])
    ex = np.concatenate([-np.ones(n_ins, np.int32), rem])
    return pts, ex, rem


def _same(a, b):
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


def test_noalloc_trial_energy_equals_allocating_entry_point(dml):
    from dml import workloads as W
    rng = np.random.default_rng(21)
    c1, box1, z01 = _context(dml, W, 31)
    c2, box2, z02 = _context(dml, W, 32)
    try:
        for n_ins, n_rem in ((5000, 3000), (400, 250)):
            for c, box, z0 in ((c1, box1, z01), (c2, box2, z02)):
                print("noalloc", n_ins, n_rem, flush=True)
                pts, ex, rem = _trials(c, box, z0, n_ins, n_rem, rng)
                ref = c.gcmc_trial_energy(pts, exclude=ex)
                got = c.gcmc_trial_energy(pts, exclude=ex, noalloc=True)
                _same(ref, got)
                assert len(got[0]) == n_ins + len(rem)
                assert (ref[1][:n_ins] == 1).any() and (ref[1][:n_ins] == 0).any()
                # points only, and removals at the atoms' own positions
                _same(c.gcmc_trial_energy(pts[:n_ins]),
                      c.gcmc_trial_energy(pts[:n_ins], noalloc=True))
                _same(c.gcmc_trial_energy(None, exclude=rem),
                      c.gcmc_trial_energy(None, exclude=rem, noalloc=True))
        # at the maximum, and one above it
        n = dml.NOALLOC_MAX_TRIAL
        pts = np.column_stack([rng.random(n) * box1[0], rng.random(n) * box1[1],
                               z01 + rng.random(n) * (box1[2] - z01)])
        _same(c1.gcmc_trial_energy(pts), c1.gcmc_trial_energy(pts, noalloc=True))
        with pytest.raises(dml.DmlError) as err:
            c1.gcmc_trial_energy(np.vstack([pts, pts[:1]]), noalloc=True)
        assert err.value.status == -3
        # the state is untouched by either entry point
        _same(c1.gcmc_trial_energy(pts[:100]), c1.gcmc_trial_energy(pts[:100], noalloc=True))
    finally:
        c1.close()
        c2.close()
