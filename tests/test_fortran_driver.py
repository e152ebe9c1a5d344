"""The Fortran side of the boundary: dana_dml (din-mol-li_amd/fortran) reads
the reference's input files, calls libdml through the ISO_C_BINDING module
dml_f in the main-loop order of dana.F90:172-262 and writes Li.xyz frames and
the E/T/rho/try/depo series like salida (dana.F90:1141-1183).  Its output must
equal, bit for bit, the same call sequence made through the C ABI from
Python: the interface passes every argument, layout and enum the way the
library reads them.

Inputs: the final frames of the reference's own ermak, gcmc and brown runs
(reference tests/*/ref.xyz, kept in tests/golden/ref_frames.npz, with the
brown run's chunk.xyz) as pos_inic.xyz, with the run parameters of their
entrada.ini / movedor.ini.
"""
import json
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "din-mol-li_amd", "dml", "dana_dml")
G = os.path.join(ROOT, "tests", "golden")
SYM = {1: "Li", 2: "CG", 3: "F"}


def write_inputs(d, cfg, pos, z, nst, nwr, chunk=None):
    with open(os.path.join(d, "entrada.ini"), "w") as f:
        for v in (cfg["idum"], cfg["prob"], cfg["h"], nst, nwr, cfg["xi"], cfg["yi"], cfg["dist"],
                  cfg["z0"], cfg["zmax"], cfg["dif_sc"], cfg["dif_sei"], cfg["nb_dcut"]):
            f.write("%r\n" % v)
    with open(os.path.join(d, "movedor.ini"), "w") as f:
        f.write(".true.\n" if cfg["ermak"] else ".false.\n")
        f.write(cfg["reservoir"] + "\n")
        if cfg["reservoir"] == "gcmc":
            f.write("%r %d\n" % (cfg["act"], cfg["nadj"]))
    with open(os.path.join(d, "pos_inic.xyz"), "w") as f:
        f.write("%d\n\n" % len(pos))
        for p, s in zip(pos, z):
            f.write("%s %.17g %.17g %.17g 6.94\n" % (SYM[int(s)], p[0], p[1], p[2]))
    if chunk is not None:
        with open(os.path.join(d, "chunk.xyz"), "w") as f:
            f.write("%d\n\n" % len(chunk))
            for p in chunk:
                f.write("Li %.17g %.17g %.17g\n" % (p[0], p[1], p[2]))


def read_series(path):
    with open(path) as f:
        return np.array([[float(v) for v in line.split()] for line in f if line.strip()])


def read_frames(path):
    frames = []
    with open(path) as f:
        lines = f.read().splitlines()
    q = 0
    while q < len(lines):
        n = int(lines[q].split()[0])
        zmax = float(lines[q + 1].split()[1])
        rows = [lines[q + 2 + k].split() for k in range(n)]
        pos = np.array([[float(v) for v in r[1:4]] for r in rows])
        z = np.array([int(r[4]) for r in rows])
        frames.append((zmax, pos, z))
        q += 2 + n
    return frames


def python_mirror(dml, cfg, pos, z, nst, nwr, chunk=None):
    """the driver's call sequence through the ctypes binding"""
    res = {"piston": dml.PISTON, "gcmc": dml.GCMC, "chunks": dml.CHUNKS}[cfg["reservoir"]]
    nadj = cfg.get("nadj", 0)
    cap = len(pos) + (max(1024, 2 * nadj * nst) if res == dml.GCMC else 0)
    z0, zmax = cfg["z0"], cfg["zmax"]
    if res == dml.CHUNKS:  # config_inic (dana.F90:437-442)
        dist = cfg["dist"] + 3.2
        z1 = z0 + dist
        zmax = z1 + dist
        cap = len(pos) + 64 * len(chunk)
    c = dml.Context(dml.dana_config([cfg["xi"], cfg["yi"], zmax], cfg["nb_dcut"], cap,
                                    h=cfg["h"], prob=cfg["prob"], z0=z0, zmax=zmax,
                                    dif_sc=cfg["dif_sc"], dif_sei=cfg["dif_sei"],
                                    integrator=dml.ERMAK if cfg["ermak"] else dml.BROWNIAN,
                                    reservoir=res, precision=dml.FP64,
                                    overlap_mode=dml.OVERLAP_EXACT, seed=abs(cfg["idum"]),
                                    act=cfg.get("act", 0.0), nadj=nadj))
    c.set_state(pos, species=z)
    frames, series = [], []
    last = {"try_": 0, "depo": 0}
    t = 0.0
    rho = None

    def salida():
        st = c.state()
        _, _, seq = c.slots()
        alive = np.nonzero(st["flags"] & 1)[0]
        order = alive[np.argsort(seq[alive], kind="stable")]
        frames.append((c.zmax(), st["pos"][order], st["species"][order].astype(int)))
        ob = c.observables()[0]
        cnt = c.counters()
        series.append([t, ob[0], ob[1], rho, cnt["try_"] - last["try_"], cnt["depo"] - last["depo"]])
        last["try_"], last["depo"] = cnt["try_"], cnt["depo"]

    c.test_update()
    if cfg["ermak"]:
        c.forces()
    rho = c.calc_rho()[0]
    if res == dml.CHUNKS:
        c.set_chunk(chunk, dist, z1)
    salida()
    for i in range(1, nst + 1):
        if cfg["ermak"]:
            c.ermak_a(); c.forces(); c.ermak_b()
        else:
            c.brownian()
        c.test_update()
        c.overlap(max_passes=100000)
        c.test_update()
        c.f_to_cg()
        if res == dml.GCMC:
            c.gcmc(cfg["act"], nadj)
        rho = c.calc_rho()[0]
        if res == dml.CHUNKS:
            c.bloques()
        if i % nwr == 0:
            salida()
        if res == dml.PISTON:
            c.maxz()
        c.end_step()
        t = t + cfg["h"]
    cnt = c.counters()
    c.close()
    return frames, np.array(series), cnt


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["ermak", "gcmc", "brown"])
def test_fortran_driver_equals_c_abi_calls(dml, tmp_path, name):
    cfg = json.load(open(os.path.join(G, "ref_runs.json")))[name]
    fr = np.load(os.path.join(G, "ref_frames.npz"))
    pos, z = fr[name + "_pos"], fr[name + "_z"]
    chunk = fr["brown_chunk"] if name == "brown" else None
    cfg = dict(cfg)
    cfg["zmax"] = cfg["golden_zmax"]
    if name == "brown":  # restart the chunks run where its golden frame ends
        cfg["z0"] = cfg["golden_zmax"] - 2 * (cfg["dist"] + 3.2)
    nst, nwr = 20, 10
    write_inputs(str(tmp_path), cfg, pos, z, nst, nwr, chunk)
    print("driver", name, len(pos), flush=True)
    out = subprocess.run([DRIVER], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = read_frames(os.path.join(str(tmp_path), "Li.xyz"))
    want, series, cnt = python_mirror(dml, cfg, pos, z, nst, nwr, chunk)
    assert len(got) == len(want) == 1 + nst // nwr
    for (zg, pg, sg), (zw, pw, sw) in zip(got, want):
        assert zg == zw
        np.testing.assert_array_equal(sg, sw)
        np.testing.assert_array_equal(pg, pw)
    for k, name_dat in enumerate(("E", "T", "rho", "try", "depo")):
        dat = read_series(os.path.join(str(tmp_path), name_dat + ".dat"))
        np.testing.assert_array_equal(dat[:, 0], series[:, 0])
        np.testing.assert_array_equal(dat[:, 1], series[:, 1 + k], err_msg=name_dat)
    line = out.stdout.split()
    assert int(line[line.index("nupd") + 1]) == cnt["nupd"]
    assert int(line[line.index("ins") + 1]) == cnt["ins"]
    assert int(line[line.index("del") + 1]) == cnt["del_"]
    if name == "gcmc":
        assert cnt["ins"] + cnt["del_"] > 0
