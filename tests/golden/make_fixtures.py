"""Write the golden fixtures under tests/golden/ (run in the build container).

1. ref_runs.json / ref_frames.npz: the reference's own test inputs
   (tests/{ermak,brown,gcmc}/entrada.ini, movedor.ini, chunk.xyz) and final
   golden frames (ref.xyz) parsed to numbers.
2. nb_cases.npz: cell assignments and neighbour lists that the reference's
   own Cells/Neighbor modules build (oracle/_ref/nbdump, compiled from the
   reference sources by oracle/build_ref.sh) for a set of configurations that
   exercise the cell quirks: ghost layers, the all-dims-<4 Verlet fallback,
   ncells < 3 duplicates, boundary rounding, electrode CG atoms.

Usage: python tests/golden/make_fixtures.py
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
NBDUMP = os.path.join(ROOT, "oracle", "_ref", "nbdump")


def parse_entrada(path):
    vals = []
    for line in open(path):
        tok = line.split()
        if tok:
            vals.append(float(tok[0]))
    keys = ["idum", "prob", "h", "nst", "nwr", "xi", "yi", "dist", "z0", "zmax", "dif_sc", "dif_sei", "nb_dcut"]
    d = dict(zip(keys, vals))
    for k in ("idum", "nst", "nwr"):
        d[k] = int(d[k])
    return d


def parse_movedor(path):
    lines = [l.split() for l in open(path) if l.strip()]
    d = {"ermak": lines[0][0].lower().startswith(".t"), "reservoir": lines[1][0]}
    if d["reservoir"] == "gcmc":
        d["act"] = float(lines[2][0])
        d["nadj"] = int(lines[2][1])
    return d


def parse_xyz_frame(path):
    L = open(path).read().split("\n")
    n = int(L[0].split()[0])
    zmax = float(L[1].split()[1])
    sym, pos, z = [], [], []
    for i in range(n):
        t = L[2 + i].split()
        sym.append(t[0])
        pos.append([float(x) for x in t[1:4]])
        z.append(int(t[4]))
    return n, zmax, np.array(pos), np.array(z, np.int8)


def write_ref_runs():
    runs, frames = {}, {}
    for name in ("ermak", "brown", "gcmc"):
        d = os.path.join(REF, "tests", name)
        cfg = parse_entrada(os.path.join(d, "entrada.ini"))
        cfg.update(parse_movedor(os.path.join(d, "movedor.ini")))
        if cfg["reservoir"] == "chunks":
            L = open(os.path.join(d, "chunk.xyz")).read().split("\n")
            nch = int(L[0])
            ch = np.array([[float(x) for x in L[2 + j].split()[1:4]] for j in range(nch)])
            frames[name + "_chunk"] = ch
        n, zmax, pos, z = parse_xyz_frame(os.path.join(d, "ref.xyz"))
        cfg["golden_nat"] = n
        cfg["golden_zmax"] = zmax
        frames[name + "_pos"] = pos
        frames[name + "_z"] = z
        runs[name] = cfg
    json.dump(runs, open(os.path.join(HERE, "ref_runs.json"), "w"), indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "ref_frames.npz"), **frames)


def nbdump(box, rc, skin, z, pos):
    lines = ["%.17g %.17g %.17g" % tuple(box), "%.17g %.17g" % (rc, skin), "%d" % len(z)]
    for zi, p in zip(z, pos):
        lines.append("%d %.17g %.17g %.17g" % (zi, p[0], p[1], p[2]))
    r = subprocess.run([NBDUMP], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    out = [l for l in r.stdout.split("\n") if l.strip() and not l.startswith("#")]
    t = out[0].split()
    assert t[0] == "cells"
    tess = t[1] == "T"
    ncells = np.array([int(x) for x in t[2:5]], np.int32)
    cell = np.array([float(x.replace("D", "E")) for x in t[5:8]])
    n = len(z)
    ci = np.array([[int(x) for x in out[1 + i].split()] for i in range(n)], np.int32)
    k = 1 + n
    rows, rowptr, idx = [], [0], []
    while k < len(out):
        t = out[k].split()
        assert t[0] == "row"
        rid, nn = int(t[1]), int(t[2])
        ids = [int(out[k + 1 + m]) for m in range(nn)]
        rows.append(rid - 1)
        idx.extend([x - 1 for x in ids])
        rowptr.append(len(idx))
        k += 1 + nn
    return dict(tess=tess, ncells=ncells, cell=cell, ci=ci, rows=np.array(rows, np.int32),
                rowptr=np.array(rowptr, np.int64), idx=np.array(idx, np.int32))


def rsa(rng, n, lo, hi, dmin, box):
    """random sequential placement with a minimum distance (x,y periodic)"""
    pts = []
    tries = 0
    while len(pts) < n and tries < 200000:
        tries += 1
        p = lo + rng.random(3) * (hi - lo)
        ok = True
        for q in pts:
            d = p - q
            d[:2] -= box[:2] * np.round(d[:2] / box[:2])
            if d @ d < dmin * dmin:
                ok = False
                break
        if ok:
            pts.append(p)
    return np.array(pts)


def bcc_layers(a, nx, ny, nl, z_top=0.0, vac=0.0, rng=None):
    pts = []
    for l in range(nl):
        zc = z_top - l * a / 2.0
        off = 0.5 * (l % 2)
        for i in range(nx):
            for j in range(ny):
                if rng is not None and vac > 0 and rng.random() < vac:
                    continue
                pts.append([(i + off) * a, (j + off) * a, zc])
    return np.array(pts)


def cases():
    rng = np.random.default_rng(20240611)
    out = []
    # 1. gcmc-test geometry at step 0: 602 Li, box 100^3, skin 2
    box = np.array([100.0, 100.0, 100.0])
    pos = rsa(rng, 602, np.zeros(3), box, 3.2, box)
    out.append(("gcmc_like", box, 3.2, 2.0, np.ones(len(pos), np.int8), pos))
    # 2. config-1 geometry: 2 bcc(100) layers of CG at z<=0 (shifted to 0..a/2)
    #    + 600 Li; box 50x50x100, skin 10 -> ncells 3,3,7
    box = np.array([50.0, 50.0, 100.0])
    el = bcc_layers(3.51, 14, 14, 2, z_top=1.755)
    ions = rsa(rng, 600, np.array([0, 0, 5.0]), np.array([50, 50, 99.9]), 3.2, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    out.append(("c1_like", box, 3.2, 10.0, z, pos))
    # 3. ghost layer: box(3)=100 while ions reach z<114 (as in the chunks test)
    box = np.array([100.0, 100.0, 100.0])
    pos = rsa(rng, 700, np.zeros(3), np.array([100, 100, 113.9]), 3.2, box)
    out.append(("ghost_top", box, 3.2, 10.0, np.ones(len(pos), np.int8), pos))
    # 4. Verlet fallback: all ncells < 4
    box = np.array([30.0, 30.0, 40.0])
    pos = rsa(rng, 120, np.zeros(3), box, 3.2, box)
    z = np.where(rng.random(len(pos)) < 0.2, 2, 1).astype(np.int8)
    out.append(("verlet", box, 3.2, 10.0, z, pos))
    # 5. ncells = 1 and 2 in x/y: stencil wrap revisits cells (duplicates)
    box = np.array([20.0, 30.0, 80.0])
    pos = rsa(rng, 150, np.zeros(3), box, 3.2, box)
    out.append(("dup_cells", box, 3.2, 10.0, np.ones(len(pos), np.int8), pos))
    # 6. boundary rounding: atoms at 0, at box-ulp, at cell boundaries
    box = np.array([66.0, 66.0, 66.0])
    rl = 3.2 + 2.0
    nc = int(66.0 / rl)
    cell = 66.0 / nc
    special = []
    for v in (0.0, np.nextafter(66.0, 0), cell, np.nextafter(cell, 0), 2 * cell, 33.0):
        special.append([v, 10.0 + len(special) * 4, 20.0])
        special.append([12.0 + len(special) * 4, v, 30.0])
        special.append([40.0, 40.0 + len(special) % 5, max(v, 0.5)])
    pos = np.vstack([np.array(special), rsa(rng, 300, np.zeros(3), box, 3.2, box)])
    z = np.where(rng.random(len(pos)) < 0.1, 2, 1).astype(np.int8)
    out.append(("boundary", box, 3.2, 2.0, z, pos))
    # 7. dense slab at rho = 0.01 (config-2 density), skin 2
    box = np.array([60.0, 60.0, 60.0])
    el = bcc_layers(3.51, 17, 17, 4, z_top=1.755 + 3 * 1.755, vac=0.2, rng=rng)
    ions = rsa(rng, 1500, np.array([0, 0, 10.0]), np.array([60, 60, 59.0]), 3.2, box)
    pos = np.vstack([el, ions])
    z = np.array([2] * len(el) + [1] * len(ions), np.int8)
    out.append(("dense_slab", box, 3.2, 2.0, z, pos))
    return out


def write_nb_cases():
    arrays = {}
    names = []
    for name, box, rc, skin, z, pos in cases():
        d = nbdump(box, rc, skin, z, pos)
        names.append(name)
        arrays[name + "/box"] = box
        arrays[name + "/rc_skin"] = np.array([rc, skin])
        arrays[name + "/z"] = z
        arrays[name + "/pos"] = pos
        for k, v in d.items():
            arrays[name + "/" + k] = np.asarray(v)
        print(name, len(z), "tess", d["tess"], "ncells", d["ncells"], "entries", len(d["idx"]))
    arrays["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "nb_cases.npz"), **arrays)


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("the reference tree is needed to regenerate fixtures")
    subprocess.check_call([os.path.join(ROOT, "oracle", "build_ref.sh")])
    write_ref_runs()
    write_nb_cases()
