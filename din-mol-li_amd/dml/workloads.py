"""Synthetic, seeded systems of the benchmark configurations.

The reference builds its initial state with an O(N^2) rejection loop
(pos_inic, dana.F90:330-396) and has no explicit electrode; these generators
place ions by random sequential addition on a hash grid (O(N)) with the same
3.2 A hard-core rule, and add bcc(100) Li0 electrode layers (species CG).

Decisions recorded where the configuration text is open:
* C1: 2 bcc layers of 14 x 14 sites (a = 3.51) plus 8 adatoms -> 400 CG; no
  reservoir (the configuration names none), z0 = zmax = 100 (z0 then only
  feeds the deposit-above-z0 stop); with a piston band the deposit front
  reaches any z0 up to 90 A before 10^4 steps and stops the run
  (dana.F90:905-908).
* C2: 4 bcc layers of 71 x 71 sites with 20% vacancies; N = 100,000 kept by
  filling ions at 0.01 A^-3 in a slab above the electrode (0.01 x the whole
  box would be 250k).
* C3: replica sizes log-uniform in [2k, 50k]; each replica is a square box
  L x L x 2L, L = (N / 0.01)^(1/3), 2 electrode layers with 20% vacancies,
  ions at 0.01 A^-3 above (piston and GCMC replicas: over the whole height,
  band z0 = 1.5 L); integrators alternate Brownian / Ermak.
"""
import math

import numpy as np

A_BCC = 3.51
RHO_ION = 0.01
DMIN = 3.2


def bcc_layers(lx, ly, nl, z_bottom=0.0, vac=0.0, rng=None, a=A_BCC):
    """nl bcc(100) layers, bottom layer at z_bottom, spacing a/2."""
    nx, ny = int(lx // a), int(ly // a)
    pts = []
    for l in range(nl):
        off = 0.5 * (l % 2)
        ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        p = np.column_stack([((ii + off) * a).ravel(), ((jj + off) * a).ravel(),
                             np.full(nx * ny, z_bottom + l * a / 2)])
        if vac > 0:
            p = p[rng.random(len(p)) >= vac]
        pts.append(p)
    return np.vstack(pts)


def pool_map(fn, jobs, workers, chunksize=1):
    """map over a forked process pool whose workers are let to finish and
    exit on their own (close + join, no terminate signal: a profiler's
    signal handler in a forked child can otherwise stall the parent)"""
    import multiprocessing as mp
    pool = mp.get_context("fork").Pool(workers)
    try:
        out = pool.map(fn, jobs, chunksize=chunksize)
        pool.close()
        pool.join()
    except BaseException:
        pool.terminate()
        raise
    return out


class _Near:
    """points of `existing` with zlo <= z < zhi, queried for neighbours
    closer than dmin under the x, y periodic minimum image"""

    def __init__(self, existing, box, zlo, zhi, dmin):
        self.tree = None
        self.dmin = dmin
        if existing is None or not len(existing):
            return
        e = existing[(existing[:, 2] >= zlo) & (existing[:, 2] < zhi)]
        if not len(e):
            return
        from scipy.spatial import cKDTree
        self.box = np.asarray(box[:2], float)
        self.zoff = zlo - dmin
        bz = 3.0 * (zhi + dmin - self.zoff)  # no z image within dmin
        self.tree = cKDTree(self._map(e), boxsize=[self.box[0], self.box[1], bz])

    def _map(self, p):
        xy = np.mod(p[:, :2], self.box)
        xy = np.where(xy >= self.box, xy - self.box, xy)
        return np.column_stack([xy, p[:, 2] - self.zoff])

    def blocked(self, cand):
        if self.tree is None:
            return np.zeros(len(cand), bool)
        d, _ = self.tree.query(self._map(cand), k=1, distance_upper_bound=self.dmin)
        return d < self.dmin

    def free_fraction(self, rng, lo, hi, samples=8192):
        """fraction of [lo, hi) farther than dmin from every point"""
        if self.tree is None:
            return 1.0
        p = np.asarray(lo) + rng.random((samples, 3)) * (np.asarray(hi) - np.asarray(lo))
        return float(1.0 - self.blocked(p).mean())


def rsa_fill(rng, n, lo, hi, box, existing=None, dmin=DMIN, batch=None):
    """n points uniform in [lo, hi) (x, y periodic in box) at least dmin from
    each other and from `existing`, by batched random sequential addition."""
    lo = np.asarray(lo, float); hi = np.asarray(hi, float)
    g = dmin / math.sqrt(3.0) * 0.999
    nx, ny = int(math.ceil(box[0] / g)), int(math.ceil(box[1] / g))
    gx, gy = box[0] / nx, box[1] / ny
    zlo = lo[2] - 2 * dmin
    nz = int(math.ceil((hi[2] + 2 * dmin - zlo) / g)) + 1
    grid = -np.ones((nx, ny, nz), np.int64)  # the new points, at most one per cell
    pts = np.zeros((0, 3))
    # existing points (an electrode lattice may put two in one grid cell) are
    # checked with a k-d tree, periodic in x and y
    near = _Near(existing, box, zlo, hi[2] + 2 * dmin, dmin)
    n0 = 0
    offs = np.array([(a, b, c) for a in range(-2, 3) for b in range(-2, 3) for c in range(-2, 3)])
    batch = batch or max(256, min(20000, n // 4 + 256))
    stall = 0
    while len(pts) - n0 < n and stall < 200:
        cand = lo + rng.random((batch, 3)) * (hi - lo)
        cx = np.floor(cand[:, 0] / gx).astype(int) % nx
        cy = np.floor(cand[:, 1] / gy).astype(int) % ny
        cz = np.floor((cand[:, 2] - zlo) / g).astype(int)
        bad = near.blocked(cand)
        # earliest candidate of each cell (a cell holds at most one point)
        tmp = np.full((nx, ny, nz), batch, np.int64)
        np.minimum.at(tmp, (cx, cy, cz), np.arange(batch))
        bad |= tmp[cx, cy, cz] != np.arange(batch)
        for dx, dy, dz in offs:
            qx, qy, qz = (cx + dx) % nx, (cy + dy) % ny, cz + dz
            ok = (qz >= 0) & (qz < nz)
            qz = np.clip(qz, 0, nz - 1)
            j = np.where(ok, grid[qx, qy, qz], -1)
            m = j >= 0
            if m.any():
                d = cand[m] - pts[j[m]]
                d[:, :2] -= box[:2] * np.round(d[:, :2] / box[:2])
                bad[np.nonzero(m)[0][(d * d).sum(1) < dmin * dmin]] = True
            k = np.where(ok, tmp[qx, qy, qz], batch)
            m = (k < np.arange(batch)) & ((dx, dy, dz) != (0, 0, 0))
            if m.any():
                d = cand[m] - cand[k[m]]
                d[:, :2] -= box[:2] * np.round(d[:, :2] / box[:2])
                bad[np.nonzero(m)[0][(d * d).sum(1) < dmin * dmin]] = True
        good = np.nonzero(~bad)[0][: n - (len(pts) - n0)]
        if len(good) == 0:
            stall += 1
            continue
        grid[cx[good], cy[good], cz[good]] = len(pts) + np.arange(len(good))
        pts = np.vstack([pts, cand[good]])
    if len(pts) - n0 < n:
        raise RuntimeError("placement stalled at %d of %d points" % (len(pts) - n0, n))
    return pts[n0:]


def system_c1(seed=104012):
    """1,000 particles: 400 CG electrode atoms + 600 Li+, 50 x 50 x 100 A."""
    rng = np.random.default_rng(seed)
    box = np.array([50.0, 50.0, 100.0])
    el = bcc_layers(50.0, 50.0, 2, 0.0)
    # 8 adatoms on hollow sites above the top layer (the next layer's sites)
    sites = np.array([[i * A_BCC, j * A_BCC, A_BCC] for i in range(14) for j in range(14)])
    el = np.vstack([el, sites[rng.choice(len(sites), 8, replace=False)]])
    ions = rsa_fill(rng, 600, [0, 0, A_BCC + DMIN], [50, 50, 100.0], box, existing=el)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    return dict(name="C1", box=box, pos=pos, species=z, skin=10.0, h=1e-2, z0=100.0, zmax=100.0)


def system_c2(seed=104012 + 2, n_total=100_000):
    """100k particles: 4-layer bcc electrode with 20% vacancies + ions at 0.01."""
    rng = np.random.default_rng(seed)
    box = np.array([250.0, 250.0, 400.0])
    el = bcc_layers(250.0, 250.0, 4, 0.0, vac=0.2, rng=rng)
    ztop = el[:, 2].max()
    n_ion = n_total - len(el)
    h = n_ion / (RHO_ION * box[0] * box[1])
    ions = rsa_fill(rng, n_ion, [0, 0, ztop + DMIN], [250, 250, ztop + DMIN + h], box, existing=el)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    return dict(name="C2", box=box, pos=pos, species=z, skin=10.0, h=1e-2, z0=200.0, zmax=400.0)


def replica_sizes(nrep=256, lo=2000, hi=50000, seed=104012 + 3):
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), nrep)).astype(int)


def system_replica(n_total, seed, reservoir_band=False):
    """One replica: box L x L x 2L with L = (N / 0.01)^(1/3) (at least 4 cells
    of 13.3 A), 2 bcc layers with 20 % vacancies, ions by RSA.  Without a
    reservoir band the ions fill a slab above the electrode at 0.01 A^-3 and
    z0 = zmax = 2L (z0 is then read only by the deposit-above-z0 stop).  With
    reservoir_band (piston and GCMC runs) the same ions fill the whole height
    and the band is the top quarter, z0 = 1.5 L: at 0.01 A^-3 and prob = 1 the
    deposit front otherwise reaches a band inside the slab within a few
    hundred steps and stops the run (dana.F90:905-908)."""
    rng = np.random.default_rng(seed)
    L = (n_total / RHO_ION) ** (1.0 / 3.0)
    L = max(L, 4 * 13.3)
    box = np.array([L, L, 2 * L])
    el = bcc_layers(L, L, 2, 0.0, vac=0.2, rng=rng)
    ztop = el[:, 2].max()
    n_ion = max(n_total - len(el), 10)
    hgt = n_ion / (RHO_ION * L * L)
    top = 2 * L - 1 if reservoir_band else min(ztop + DMIN + hgt, 2 * L - 1)
    ions = rsa_fill(rng, n_ion, [0, 0, ztop + DMIN], [L, L, top], box, existing=el)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    z0 = 1.5 * L if reservoir_band else 2 * L
    return dict(box=box, pos=pos, species=z, skin=10.0, h=1e-2, z0=z0, zmax=2 * L)


def system_replica_box(n_total):
    """the box system_replica(n_total, ...) uses"""
    L = max((n_total / RHO_ION) ** (1.0 / 3.0), 4 * 13.3)
    return np.array([L, L, 2 * L])


RHO_CHUNK = 6.0e-4   # 1 M, the concentration the reference's chunks sensor expects


def system_chunks(n_total, seed, rho=RHO_CHUNK):
    """A chunks-reservoir replica in the regime of the reference's brown run
    (tests/brown: ions at 1 M, the sensor band of bloques around rhomedia):
    ions by RSA in an L x L x L column (no electrode: deposition at z <= 0),
    z0 = L, the step dist = L/2 + rcut, z1 = z0 + dist, zmax = box(3) =
    z1 + dist, and a block template of the same density, L x L x L/2."""
    rng = np.random.default_rng(seed)
    L = max((n_total / rho) ** (1.0 / 3.0), 4 * 13.3)
    dist = 0.5 * L + 3.2
    z0 = L
    z1 = z0 + dist
    zmax = z1 + dist
    box = np.array([L, L, zmax])
    ions = rsa_fill(rng, int(n_total), [0, 0, 0.0], [L, L, L], box)
    nch = max(1, int(round(rho * L * L * (0.5 * L))))
    chunk = rsa_fill(rng, nch, [0, 0, 0.0], [L, L, 0.5 * L], box)
    z = np.ones(len(ions), np.int8)
    return dict(box=box, pos=ions, species=z, skin=10.0, h=1e-2, z0=z0, zmax=zmax, z1=z1,
                dist=dist, chunk=chunk)


def _slab_job(args):
    seed, k, n, lo, hi, box, existing = args
    return rsa_fill(np.random.default_rng([seed, k]), n, lo, hi, box, existing=existing,
                    batch=min(100000, max(2048, n // 2)))


def rsa_fill_slabs(seed, n, lo, hi, box, existing=None, workers=16, slabs=32):
    """rsa_fill in z slabs: the even slabs in parallel first, then the odd
    ones against their neighbours (same hard-core rule, points spread over
    the slabs by the volume `existing` leaves free)."""
    lo = np.asarray(lo, float); hi = np.asarray(hi, float)
    slabs = max(2, min(slabs, int((hi[2] - lo[2]) // (4 * DMIN))))
    edges = np.linspace(lo[2], hi[2], slabs + 1)
    ex = existing if existing is not None else np.zeros((0, 3))
    near = _Near(ex, box, lo[2] - 2 * DMIN, hi[2] + 2 * DMIN, DMIN)
    frng = np.random.default_rng([seed, 1 << 20])
    free = np.array([near.free_fraction(frng, [lo[0], lo[1], edges[k]], [hi[0], hi[1], edges[k + 1]])
                     for k in range(slabs)])
    cum = np.concatenate([[0.0], np.cumsum(free)])
    counts = np.diff(np.round(n * cum / cum[-1]).astype(int))
    out = [None] * slabs

    def jobs(parity):
        js = []
        for k in range(parity, slabs, 2):
            near = [ex]
            for q in (k - 1, k + 1):
                if 0 <= q < slabs and out[q] is not None:
                    near.append(out[q])
            js.append((seed, k, int(counts[k]), [lo[0], lo[1], edges[k]],
                       [hi[0], hi[1], edges[k + 1]], box, np.vstack(near)))
        return js

    for parity in (0, 1):
        js = jobs(parity)
        if workers > 1:
            res = pool_map(_slab_job, js, min(workers, len(js)))
        else:
            res = [_slab_job(j) for j in js]
        for j, r in zip(js, res):
            out[j[1]] = r
    return np.vstack(out)


def system_c5(n_total=2_000_000, seed=104012 + 5, L=600.0, n_prot=30, workers=16):
    """The dendrite-regime system: a 12-layer bcc(100) electrode (a = 3.51,
    species CG) on L x L, seeded hemispherical bcc protrusions of radius
    10-40 A on its top surface, and Li+ ions by RSA at 0.01 A^-3 above it,
    n_total particles in all.  Box L x L x (electrode + ion slab + 20 A)."""
    rng = np.random.default_rng(seed)
    a = A_BCC
    slab = bcc_layers(L, L, 12, 0.0, rng=rng)
    ztop = slab[:, 2].max()
    nx = int(L // a)
    prot = []
    for _ in range(n_prot):
        r = rng.uniform(10.0, 40.0)
        cx, cy = rng.uniform(0, L, 2)
        # lattice sites of the two bcc sublattices above the slab inside the sphere
        for off in (0.0, 0.5):
            i0 = np.arange(int((cx - r) // a) - 1, int((cx + r) // a) + 2)
            j0 = np.arange(int((cy - r) // a) - 1, int((cy + r) // a) + 2)
            k0 = np.arange(1, int(r // (a / 2)) + 2)
            ii, jj, kk = np.meshgrid(i0, j0, k0, indexing="ij")
            x = (ii + off) * a
            y = (jj + off) * a
            z = ztop + kk * a  # whole cells above the top layer
            if off:
                z = z - a / 2
            p = np.column_stack([x.ravel(), y.ravel(), z.ravel()])
            d = p - np.array([cx, cy, ztop])
            d[:, :2] -= L * np.round(d[:, :2] / L)
            p = p[(d * d).sum(1) <= r * r]
            p[:, :2] %= L
            prot.append(p)
    prot = np.vstack(prot)
    # one atom per lattice site
    key = np.round(prot / (a / 2)).astype(np.int64)
    _, first = np.unique(key[:, 0] * 10**12 + key[:, 1] * 10**6 + key[:, 2], return_index=True)
    prot = prot[np.sort(first)]
    prot[:, :2] = np.where(prot[:, :2] >= L - 1e-9, 0.0, prot[:, :2])
    el = np.vstack([slab, prot])
    etop = el[:, 2].max()
    n_ion = n_total - len(el)
    # the ion slab starts DMIN above the flat top; the protrusions take part
    # of its volume, so it is made taller by the volume they exclude
    z_lo = ztop + DMIN
    excl = 1.0 - _Near(el, [L, L], z_lo - 2 * DMIN, etop + 2 * DMIN, DMIN).free_fraction(
        np.random.default_rng([seed, 7]), [0, 0, z_lo], [L, L, etop + DMIN], samples=200_000)
    hgt = n_ion / (RHO_ION * L * L) + excl * (etop + DMIN - z_lo)
    box = np.array([L, L, ztop + DMIN + hgt + 20.0])
    ions = rsa_fill_slabs(seed, n_ion, [0, 0, ztop + DMIN], [L, L, ztop + DMIN + hgt], box,
                          existing=el, workers=workers)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    return dict(name="C5", box=box, pos=pos, species=z, skin=10.0, h=1e-2, z0=box[2],
                zmax=box[2], electrode_top=etop, n_electrode=len(el))
