"""Launch-shape autotuner.

Picks, for one ensemble (a list of replica configs and their initial
states), the launch shapes that make a step fastest:

* lanes per ref atom of the force kernel (2..32),
* the list kernel: cell tiles staged through LDS (1), one wave per atom (2),
  each replica by its atoms per cell (3) or one wave per cell (4),
* lanes per ref atom of the overlap round scan (4, 8, 16),
* overlap rounds per step: the fewest of 8, 16, 32 with which no overlap
  call of the tuning steps ends with atoms still overlapping.

Everything runs on a scratch context holding the same replicas, so the
context being tuned keeps its exact trajectory.  The first two are timed per
launch with HIP events on the scratch context's stream (`dml_time_kernel`,
which rebuilds at the current positions first); the overlap lanes by short
whole-step runs, alternating the candidates twice.  Results depend on the
shapes only through the FP64 summation order of the force sums.  Choices are
cached in JSON by workload signature.
"""
import hashlib
import json
import os

import numpy as np

from . import K_FORCE, K_LISTBUILD, Context

FORCE_LANES = (2, 4, 8, 16, 32)
LIST_KERNELS = (1, 2, 3, 4, 5, 6)
OVERLAP_LANES = (1, 2, 4, 8, 16)
OVERLAP_ROUNDS = (8, 16, 32)


def signature(configs, sizes):
    """replica count, sizes and the per-replica settings that change the work"""
    h = hashlib.sha1()
    for c, n in zip(configs, sizes):
        h.update(np.array([n, c.integrator, c.reservoir, c.precision, c.overlap_mode],
                          np.int64).tobytes())
        h.update(np.array([c.skin, c.rcut, c.box[0], c.box[1], c.box[2]], np.float64).tobytes())
    return "%d-%d-%s" % (len(configs), int(sum(sizes)), h.hexdigest()[:16])


def _cache_path(path):
    return path or os.environ.get("DML_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "dml_tuning.json")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, cache):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path + ".tmp", "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(path + ".tmp", path)
    except OSError:
        pass


def tune(configs, states, steps=4, reps=10, cache_path=None, use_cache=True, log=None):
    """configs: list of dml Config; states: list of (pos, species) or
    (pos, species, chunk) per replica, chunk = (template, dist, z1) of a
    chunks replica.  Returns dict(lanes_per_atom, list_kernel, overlap_lanes,
    timings)."""
    path = _cache_path(cache_path)
    sig = signature(configs, [len(st[0]) for st in states])
    cache = _load(path) if use_cache else {}
    if sig in cache:
        return dict(cache[sig], cached=True)
    ctx = Context(list(configs))
    try:
        for k, st in enumerate(states):
            ctx.set_state(st[0], species=st[1], rep=k)
            if len(st) > 2 and st[2] is not None:
                ctx.set_chunk(*st[2], rep=k)
        ctx.run(2, graph=False)  # lists, forces, a first overlap pass
        t_force = {}
        for lanes in FORCE_LANES:
            ctx.set_tuning(lanes_per_atom=lanes)
            t_force[lanes] = ctx.time_kernel(K_FORCE, reps)
        best_lanes = min(t_force, key=t_force.get)
        ctx.set_tuning(lanes_per_atom=best_lanes)
        t_list = {}
        for kern in LIST_KERNELS:
            ctx.set_tuning(list_kernel=kern)
            t_list[kern] = ctx.time_kernel(K_LISTBUILD, reps)
        best_list = min(t_list, key=t_list.get)
        ctx.set_tuning(list_kernel=best_list)
        t_ov = {lanes: [] for lanes in OVERLAP_LANES}
        for _ in range(2):
            for lanes in OVERLAP_LANES:
                ctx.set_tuning(overlap_lanes=lanes)
                ctx.run(1, graph=True)  # capture with this shape
                t_ov[lanes].append(ctx.run(steps, graph=True) / steps)
        best_ov = min(t_ov, key=lambda x: min(t_ov[x]))
        ctx.set_tuning(overlap_lanes=best_ov)
        nrep = len(configs)
        # from the most rounds down, stopping at the first count that leaves
        # an overlap unresolved (its leftovers would burden the next trial)
        best_rounds = OVERLAP_ROUNDS[-1]
        for rounds in sorted(OVERLAP_ROUNDS, reverse=True):
            ctx.set_tuning(overlap_rounds=rounds)
            before = sum(ctx.counters(k)["ov_unfinished"] for k in range(nrep))
            ctx.run(steps, graph=True)
            if sum(ctx.counters(k)["ov_unfinished"] for k in range(nrep)) != before:
                break
            best_rounds = rounds
    finally:
        ctx.close()
    out = dict(lanes_per_atom=int(best_lanes), list_kernel=int(best_list),
               overlap_lanes=int(best_ov), overlap_rounds=int(best_rounds),
               timings=dict(force_ms={str(k): v for k, v in t_force.items()},
                            list_ms={str(k): v for k, v in t_list.items()},
                            step_ms_by_overlap_lanes={str(k): min(v) for k, v in t_ov.items()}))
    if log:
        log("autotune %s: %s" % (sig, json.dumps(out)))
    cache[sig] = out
    _save(path, cache)
    return dict(out, cached=False)


def apply(ctx, choice):
    ctx.set_tuning(lanes_per_atom=choice["lanes_per_atom"], list_kernel=choice["list_kernel"],
                   overlap_lanes=choice["overlap_lanes"],
                   overlap_rounds=choice.get("overlap_rounds", 16))
