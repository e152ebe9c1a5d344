"""Whole runs of the reference's three test configurations (tests/golden:
ermak with the piston, brown with chunks, gcmc) with the reference's
overlap_moveback and with the round schedule in its place (16 rounds per
call, the device's default), over many seeds: deposits, collisions and final
state.  Prints mean +- standard error per configuration and mode.

  python oracle/overlap_stats.py [seeds]

Uses only the oracle (liborc.so); no GPU."""
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import orc  # noqa: E402

GOLD = os.path.join(os.path.dirname(HERE), "tests", "golden")


def _run(args):
    name, idum, rounds, seed = args
    cfgs = json.load(open(os.path.join(GOLD, "ref_runs.json")))
    fr = np.load(os.path.join(GOLD, "ref_frames.npz"))
    c = cfgs[name]
    chunk = fr[name + "_chunk"] if name + "_chunk" in fr.files else None
    try:
        out = orc.dana_run(idum, c["prob"], c["h"], c["nst"], c["nwr"], c["xi"], c["yi"], c["dist"],
                           c["z0"], c["zmax"], c["dif_sc"], c["dif_sei"], c["nb_dcut"], c["ermak"],
                           c["reservoir"], act=c.get("act", 0.0), nadj=c.get("nadj", 0), chunk=chunk,
                           overlap_rounds=rounds, rounds_seed=seed)
    except RuntimeError:
        return name, rounds, None  # a fatal stop of the reference (e.g. deposit above z0)
    cc = out["counters"]
    return name, rounds, dict(depo=cc["depo"], choques=cc["choques"], choques3=cc["choques3"],
                              nat=out["nat"], zmax=out["zmax"], ins=cc["ins"],
                              unfinished=cc["ov_unfinished"])


def main(nseed=24, workers=8):
    orc.build()
    jobs = [(name, -104012 - 7 * k, rounds, 1000 + k) for name in ("ermak", "brown", "gcmc")
            for k in range(nseed) for rounds in (0, 16)]
    with Pool(workers) as p:
        res = p.map(_run, jobs, chunksize=1)
    keys = ("depo", "choques", "choques3", "nat", "zmax", "ins", "unfinished")
    for name in ("ermak", "brown", "gcmc"):
        for rounds, mode in ((0, "overlap_moveback"), (16, "round schedule")):
            rs = [r for n, q, r in res if n == name and q == rounds and r]
            fails = sum(1 for n, q, r in res if n == name and q == rounds and not r)
            print("%-6s %-17s runs %2d (stopped %d) " % (name, mode, len(rs), fails) +
                  "  ".join("%s %.1f +- %.1f" % (k, np.mean([r[k] for r in rs]),
                                                  np.std([r[k] for r in rs]) / np.sqrt(len(rs)))
                            for k in keys))


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 24)
