#!/usr/bin/env python
"""Throughput of the din-mol-Li hot path on one MI355X.

Legs (synthetic, seeded systems; see din-mol-li_amd/dml/workloads.py):
  plain    C2: one 100,000-particle system, FP64 forces, Ermak integrator,
           kernels launched one by one (no graph).
  batched  C3: 256 replicas, sizes log-uniform in 2k..50k, FP32 pair forces
           with FP64 sums and positions, Brownian and Ermak alternating,
           every phase one launch over all replicas, one hipGraph per step.
value = batched particle-steps/s (particles of all replicas x steps / time),
timed with HIP events around exactly --steps steps after --warmup steps.
  c5       the 2M-particle dendrite-regime system with GCMC trial energies.
Also reported: the dominant kernel against the HBM roofline and the CPU
oracle on the reference's Ermak test (1,204 particles, 10^4 steps) on one
host core.

  serving  C4: Poisson arrivals at 20/s, continuous batching over a slot pool;
           p50/p99 completion latency.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "din-mol-li_amd"))

HBM_PEAK_GBS = 8000.0
# reservoir of replica k is RESERVOIR_CYCLE[k % 3]
RESERVOIR_CYCLE = ("PISTON", "CHUNKS", "GCMC")
# serving requests cycle without chunks: a chunks run grows its box by one
# block step per firing, beyond the cell memory of a slot class
SERVING_CYCLE = ("PISTON", "NO_RESERVOIR", "GCMC")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _replica(args):
    n, seed, res = args
    from dml import workloads as W
    if res == "CHUNKS":
        return W.system_chunks(int(n), int(seed))
    return W.system_replica(int(n), int(seed), reservoir_band=res != "NO_RESERVOIR")


def build_replicas(nrep, seed0=104012):
    from dml import workloads as W
    sizes = W.replica_sizes(nrep)
    jobs = [(int(n), seed0 + 1000 + k, RESERVOIR_CYCLE[k % 3]) for k, n in enumerate(sizes)]
    workers = min(16, os.cpu_count() or 1)
    if workers > 1 and nrep > 4:
        return W.pool_map(_replica, jobs, workers, chunksize=2)
    return [_replica(j) for j in jobs]


def force_rel_err(f32, f64):
    import numpy as np
    na = np.linalg.norm(f32 - f64, axis=1)
    nb = np.linalg.norm(f64, axis=1)
    floor = 1e-3 * np.sqrt(np.mean(nb ** 2))
    return float(np.max(na / np.maximum(nb, floor)))


def energy_drift(dml, mk, snap, steps, every=100):
    """<sum epot> of FP32-force vs FP64 Ermak runs from the same state with
    the same Philox streams: |<E>_32 - <E>_64| / |<E>_64| over the run, and
    each run's raw slope dE/dt (least squares over samples every `every`
    steps; energia of salida, computed on the device)."""
    import numpy as np
    out = {}
    # the third run is FP64 from positions moved by 1e-9 A: two FP64
    # trajectories that part like the FP32 one does (the dynamics is
    # chaotic), the scale against which the FP32 difference is read
    rng = np.random.default_rng(5)
    kick = 1e-9 * rng.standard_normal(snap["pos"].shape)
    for key, prec, dp in (("fp64", dml.FP64, 0.0), ("fp32", dml.FP32_FORCE, 0.0),
                          ("fp64_kicked", dml.FP64, kick)):
        c = dml.Context(mk(prec))
        c.set_state(snap["pos"] + dp, species=snap["species"], vel=snap["vel"], acel=snap["acel"])
        e = []
        for _ in range(steps // every):
            c.run(every, graph=True)
            e.append(c.observables()[0][0])
        out[key] = np.array(e)
        c.close()
    h = mk(dml.FP64).h
    t = (np.arange(len(out["fp64"])) + 1) * every * h
    m = {k: float(v.mean()) for k, v in out.items()}
    return dict(steps=steps, samples=len(t), mean_epot_fp64=m["fp64"], mean_epot_fp32=m["fp32"],
                rel=abs(m["fp32"] - m["fp64"]) / abs(m["fp64"]),
                rel_fp64_kicked=abs(m["fp64_kicked"] - m["fp64"]) / abs(m["fp64"]),
                slope_fp64_per_ps=float(np.polyfit(t, out["fp64"], 1)[0]),
                slope_fp32_per_ps=float(np.polyfit(t, out["fp32"], 1)[0]),
                slope_fp64_kicked_per_ps=float(np.polyfit(t, out["fp64_kicked"], 1)[0]))


def plain_leg(dml, W, steps, warmup, drift_steps=10000):
    import numpy as np
    c2 = W.system_c2()
    n = len(c2["pos"])
    mk = lambda prec: dml.dana_config(c2["box"], c2["skin"], n, h=c2["h"], z0=c2["z0"],
                                      zmax=c2["zmax"], integrator=dml.ERMAK,
                                      reservoir=dml.NO_RESERVOIR, precision=prec,
                                      overlap_mode=dml.OVERLAP_ROUNDS, seed=104012)
    from dml import autotune
    choice = autotune.tune([mk(dml.FP64)], [(c2["pos"], c2["species"])], log=log)
    ctx = dml.Context(mk(dml.FP64))
    autotune.apply(ctx, choice)
    ctx.set_state(c2["pos"], species=c2["species"])
    ctx.run(warmup, graph=True)
    ctx.sync()
    t0 = time.perf_counter()
    ms = ctx.run(steps, graph=True)
    wall = time.perf_counter() - t0
    mean_nn, reb, st = ctx.list_stats()
    snap = ctx.state()
    ctx.close()
    # the configuration's own integrator, Brownian (cbrownian_hs computes no
    # forces, so the FP32-force comparison below needs the Ermak run)
    mkb = lambda: dml.dana_config(c2["box"], c2["skin"], n, h=c2["h"], z0=c2["z0"],
                                  zmax=c2["zmax"], integrator=dml.BROWNIAN,
                                  reservoir=dml.NO_RESERVOIR, precision=dml.FP64,
                                  overlap_mode=dml.OVERLAP_ROUNDS, seed=104012)
    choice_b = autotune.tune([mkb()], [(c2["pos"], c2["species"])], log=log)
    cb = dml.Context(mkb())
    autotune.apply(cb, choice_b)
    cb.set_state(c2["pos"], species=c2["species"])
    cb.run(warmup, graph=True)
    cb.sync()
    ms_b = cb.run(steps, graph=True)
    cb.close()
    # FP32 vs FP64 forces where fuerza runs in the step: right after ermak_a
    # (dana.F90:178-184), before the overlap resolution removes the pairs
    # inside r0.  Both contexts draw the same Philox noise.
    out = {}
    for prec in (dml.FP64, dml.FP32_FORCE):
        c = dml.Context(mk(prec))
        c.set_state(snap["pos"], species=snap["species"], vel=snap["vel"], acel=snap["acel"])
        c.test_update()
        c.ermak_a()
        c.forces()
        out[prec] = c.state()["force"]
        c.close()
    ions = snap["species"] != 2
    f64 = out[dml.FP64][ions]
    hot = np.linalg.norm(f64, axis=1) > 0
    err = force_rel_err(out[dml.FP32_FORCE][ions][hot], f64[hot]) if hot.any() else None
    drift = energy_drift(dml, mk, snap, drift_steps) if drift_steps else None
    return dict(n=n, ms=ms, wall=wall, rate=n * steps / (ms * 1e-3), mean_nn=mean_nn,
                rebuilds=reb, steps_done=st, force_err=err, interacting=int(hot.sum()),
                energy_drift=drift, brownian_rate=n * steps / (ms_b * 1e-3))


def c3_configs(dml, reps, precision=None):
    """the configs of the C3 ensemble: replica k has reservoir
    RESERVOIR_CYCLE[k % 3], the Brownian (k even) or Ermak (k odd)
    integrator, FP32 pair forces and the round schedule"""
    cfgs = []
    for k, r in enumerate(reps):
        n = len(r["pos"])
        res = RESERVOIR_CYCLE[k % 3]
        # GCMC: control volume from 3/4 of the ion slab to the box top at the
        # ion density, 10 trials per step as in the reference's gcmc run
        kw = dict(reservoir=getattr(dml, res))
        kw.update(z0=r["z0"], zmax=r["zmax"])
        if res == "GCMC":
            kw.update(act=0.01, nadj=10)
            n += max(1024, n // 4)
        elif res == "CHUNKS":
            n += 4 * len(r["chunk"])  # room for four blocks (the sensor holds after ~3)
        cfgs.append(dml.dana_config(r["box"], r["skin"], n, h=r["h"],
                                    integrator=dml.ERMAK if k % 2 else dml.BROWNIAN,
                                    precision=dml.FP32_FORCE if precision is None else precision,
                                    overlap_mode=dml.OVERLAP_ROUNDS, seed=104012 + 7919 * k, **kw))
    return cfgs


def batched_leg(dml, nrep, steps, warmup, full_steps=0, list_kernel=0):
    t0 = time.perf_counter()
    reps = build_replicas(nrep)
    log("built %d replicas in %.1f s" % (nrep, time.perf_counter() - t0))
    cfgs = c3_configs(dml, reps)
    from dml import autotune
    t0 = time.perf_counter()
    choice = autotune.tune(cfgs, [(r["pos"], r["species"],
                                   (r["chunk"], r["dist"], r["z1"]) if "chunk" in r else None)
                                  for r in reps], log=log)
    log("autotune %.1f s%s" % (time.perf_counter() - t0, " (cached)" if choice.get("cached") else ""))
    if list_kernel:
        # a fixed row-build kernel (profiles of one variant against another)
        choice = dict(choice, list_kernel=list_kernel)
    ens = dml.Context(cfgs)
    autotune.apply(ens, choice)
    ntot = 0
    for k, r in enumerate(reps):
        ens.set_state(r["pos"], species=r["species"], rep=k)
        if "chunk" in r:
            ens.set_chunk(r["chunk"], r["dist"], r["z1"], rep=k)
        ntot += len(r["pos"])
    ens.run(warmup, graph=True)
    ens.sync()
    before = [ens.counters(k) for k in range(nrep)]
    t0 = time.perf_counter()
    ms = ens.run(steps, graph=True)
    wall = time.perf_counter() - t0
    nn = [ens.list_stats(k) for k in range(nrep)]
    mean_nn = sum(a for a, b, c in nn) / nrep
    reb = sum(b for a, b, c in nn)
    stepsum = sum(c for a, b, c in nn)
    # particle-steps actually taken (atoms alive in each completed step): a
    # replica halted by a fatal condition of the reference (e.g. a deposit
    # above z0) stops advancing, a GCMC replica grows and shrinks
    after = [ens.counters(k) for k in range(nrep)]
    psteps = sum(a["particle_steps"] - b["particle_steps"] for a, b in zip(after, before))
    trials = sum(a["ins"] + a["del_"] - b["ins"] - b["del_"] for a, b in zip(after, before))
    # overlap calls of the timed steps that ran out of rounds with atoms
    # still overlapping (the state carries into the next step)
    ov_left = sum(a["ov_unfinished"] - b["ov_unfinished"] for a, b in zip(after, before))
    halted = sum(1 for k in range(nrep) if ens.replica_status(k) != 0)
    # dominant kernel: the neighbour-row build (see profiles/), timed on the
    # context's own stream over every replica; its rows are those of the ref
    # atoms now (GCMC and chunks replicas have grown)
    c0 = [ens.counters(k) for k in range(nrep)]
    kl_ms = ens.time_kernel(dml.K_LISTBUILD, 10)
    c1 = [ens.counters(k) for k in range(nrep)]
    # the timing rebuilds once before its 10 launches: 11 builds counted
    rows_now = sum(b["rows"] - a["rows"] for a, b in zip(c0, c1)) / 11.0
    entries_now = sum(b["row_entries"] - a["row_entries"] for a, b in zip(c0, c1)) / 11.0
    mean_nn_w = entries_now / max(rows_now, 1.0)
    kf_ms = ens.time_kernel(dml.K_FORCE, 10)
    full = None
    if full_steps > warmup + steps:
        # the configuration's own length: the same ensemble run on to
        # full_steps steps in all (deposit fronts, chunk blocks and GCMC
        # moves change the work as it goes), timed in segments
        done_steps, f_ms = warmup + steps, 0.0
        p0 = sum(ens.counters(k)["particle_steps"] for k in range(nrep))
        while done_steps < full_steps:
            seg = min(500, full_steps - done_steps)
            f_ms += ens.run(seg, graph=True)
            done_steps += seg
            log("c3 full run: %d / %d steps, %.1f s" % (done_steps, full_steps, f_ms * 1e-3))
        p1 = sum(ens.counters(k)["particle_steps"] for k in range(nrep))
        full = dict(steps=full_steps, steps_timed=full_steps - warmup - steps, seconds=f_ms * 1e-3,
                    rate=(p1 - p0) / (f_ms * 1e-3),
                    mean_particles=(p1 - p0) / (full_steps - warmup - steps),
                    halted=sum(1 for k in range(nrep) if ens.replica_status(k) != 0),
                    overlap_unfinished=sum(ens.counters(k)["ov_unfinished"] for k in range(nrep)))
    ens.close()
    return dict(n=ntot, ms=ms, wall=wall, rate=psteps / (ms * 1e-3), mean_nn=mean_nn,
                particles_stepped=psteps / steps, full_run=full,
                rows_at_timing=rows_now, mean_nn_rows=mean_nn_w,
                rebuilds_per_rep_step=reb / max(stepsum, 1), list_ms=kl_ms, force_ms=kf_ms,
                halted=halted, gcmc_accepted=trials, overlap_unfinished=ov_left,
                tuning={k: choice[k] for k in ("lanes_per_atom", "list_kernel", "overlap_lanes",
                                               "overlap_rounds")},
                reservoirs={x: sum(1 for k in range(nrep) if RESERVOIR_CYCLE[k % 3] == x)
                            for x in RESERVOIR_CYCLE})


def _request_system(args):
    n, seed, res = args
    from dml import workloads as W
    return W.system_replica(int(n), int(seed), reservoir_band=res != "NO_RESERVOIR")


SERVE_POOL = "8192:48,32768:16,131072:4,212992:2"


def serving_leg(dml, duration, rate=20.0, step_scale=0.01, seed=104012 + 4, pool_spec=SERVE_POOL):
    """C4 on one GPU: Poisson arrivals at `rate` per second for `duration`
    seconds; sizes lognormal (median 10k, sigma 1) clipped to 1k..200k; step
    counts uniform in 1e3..1e5 times `step_scale`.  Requests are admitted into
    a slot pool reserved at start and served by continuous batching."""
    import math
    import numpy as np
    from dml import serve, workloads as W
    rng = np.random.default_rng(seed)
    t, arr = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        arr.append(t)
    m = len(arr)
    sizes = np.clip(np.exp(rng.normal(math.log(10000.0), 1.0, m)), 1000, 200000).astype(int)
    steps = np.maximum(1, np.round(rng.integers(1000, 100001, m) * step_scale)).astype(int)
    t0 = time.perf_counter()
    jobs = [(int(n), seed + 17 * k, SERVING_CYCLE[k % 3]) for k, n in enumerate(sizes)]
    workers = min(16, os.cpu_count() or 1)
    systems = W.pool_map(_request_system, jobs, workers, chunksize=4)
    log("built %d requests in %.1f s" % (m, time.perf_counter() - t0))
    reqs = []
    for k, (sysd, st) in enumerate(zip(systems, steps)):
        res = SERVING_CYCLE[k % 3]
        n = len(sysd["pos"])
        kw = dict(reservoir=getattr(dml, res), z0=sysd["z0"], zmax=sysd["zmax"])
        need = n
        if res == "GCMC":
            kw.update(act=0.01, nadj=10)
            need = n + min(max(1024, n // 4), 10 * int(st) + 64)
        cfg = dml.dana_config(sysd["box"], sysd["skin"], need, h=sysd["h"],
                              integrator=dml.ERMAK if k % 2 else dml.BROWNIAN,
                              precision=dml.FP32_FORCE, overlap_mode=dml.OVERLAP_ROUNDS,
                              seed=104012 + 7919 * k, **kw)
        reqs.append(serve.Request(k, arr[k], cfg, sysd["pos"], sysd["species"], st, need=need))
    # slot classes: the HBM reserved for serving (state + rows)
    classes = []
    for item in pool_spec.split(","):
        cap, count = (int(v) for v in item.split(":"))
        classes.append((cap, count, W.system_replica_box(cap)))
    pool = serve.SlotPool(classes)
    log("serving pool: %d slots, %.1f GB reserved" % (sum(c[1] for c in classes),
                                                     pool.hbm_bytes() / 1e9))
    done, wall = serve.serve(pool, reqs, segment=8)
    pool.close()
    out = serve.summarize(done, wall)
    offered = sum(q.n * q.steps for q in reqs) / duration
    out.update(requests=m, arrival_rate_per_s=rate, duration_s=duration, step_scale=step_scale,
               offered_particle_steps_per_s=offered, slots=sum(c[1] for c in classes),
               reserved_gb=pool.hbm_bytes() / 1e9)
    return out


def c5_leg(dml, steps, warmup, trials=4096):
    """C5: the 2M-particle dendrite-regime system (12-layer electrode with 30
    hemispherical protrusions, ions at 0.01 A^-3), Ermak with FP32 pair
    forces, the GCMC reservoir in the top quarter with 4,096 insertion and
    4,096 deletion trials every step (alternating kinds): evaluated in the
    batch kernels (points, hard-core neighbours, FP64 trial energies) and
    committed in trial order, with the sequential gcmc_run's result, inside
    the captured step."""
    from dml import workloads as W
    t0 = time.perf_counter()
    sysd = W.system_c5()
    log("built C5 (%d particles) in %.1f s" % (len(sysd["pos"]), time.perf_counter() - t0))
    box = sysd["box"]
    z0 = 0.75 * box[2]
    n = len(sysd["pos"])
    nadj = 2 * trials
    cfg = dml.dana_config(box, sysd["skin"], n + 65536, h=sysd["h"], z0=z0, zmax=box[2],
                          integrator=dml.ERMAK, reservoir=dml.GCMC, precision=dml.FP32_FORCE,
                          overlap_mode=dml.OVERLAP_ROUNDS, act=0.01, nadj=nadj, seed=104017,
                          gcmc_mode=dml.GCMC_BATCHED | dml.GCMC_ALTERNATE)
    from dml import autotune
    t0 = time.perf_counter()
    choice = autotune.tune([cfg], [(sysd["pos"], sysd["species"])], log=log)
    log("c5 autotune %.1f s" % (time.perf_counter() - t0))
    c = dml.Context(cfg)
    autotune.apply(c, choice)
    c.set_state(sysd["pos"], species=sysd["species"])
    c.run(warmup, graph=True)
    c.sync()
    cnt0 = c.counters()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        seg = min(500, steps - done)
        c.run(seg, graph=True)
        done += seg
        log("c5: %d / %d steps, %.1f s" % (done, steps, time.perf_counter() - t0))
    c.sync()
    wall = time.perf_counter() - t0
    cnt = c.counters()
    out = c.gcmc_outcomes(nadj)
    st = c.replica_status(0)
    ps = cnt["particle_steps"] - cnt0["particle_steps"]
    moves = (cnt["ins"] - cnt0["ins"]) + (cnt["del_"] - cnt0["del_"])
    last = None
    if out is not None:
        ins = (out["kind"] == 0)
        acc = out["accepted"] == 1
        last = dict(insertions_accepted=int((ins & acc).sum()),
                    deletions_accepted=int((~ins & acc).sum()),
                    mean_du_accepted_insertion=float(out["du"][ins & acc].mean()) if (ins & acc).any() else None,
                    hard_core_or_rejected_insertions=int((ins & ~acc).sum()))
    c.close()
    return dict(n=n, steps=steps, wall=wall, rate=ps / wall if wall > 0 else None,
                trials_per_step=nadj, trials_per_s=nadj * steps / wall,
                moves_committed_per_s=moves / wall, batched=out is not None, last_call=last,
                status=st, tuning={k: choice[k] for k in ("lanes_per_atom", "list_kernel",
                                                          "overlap_lanes", "overlap_rounds")})


def cpu_baseline():
    """The CPU oracle (a bit-exact C restatement of the reference's routines,
    one host core) on C1: 1,000 particles (400 CG electrode atoms, 600 Li+)
    in 50 x 50 x 100 A, 10^4 Ermak steps, no reservoir, skin 10; also on the reference's own ermak test (1,204 particles, 10^4
    steps), on which the reference Fortran built at -O2 takes 10.3 s on one
    core of the build container's host (1.17e6 particle-steps/s)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    from dml import workloads as W
    c1 = W.system_c1()
    nst = 10000
    out = orc.dana_run(-104012, 1.0, c1["h"], nst, nst, c1["box"][0], c1["box"][1], 50.0,
                       c1["z0"], c1["zmax"], 250.0, 250.0, c1["skin"], 1, "none",
                       init_pos=c1["pos"], init_z=c1["species"])
    n = len(c1["pos"])
    rate = n * nst / out["loop_seconds"]
    cfg = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_runs.json")))["ermak"]
    e = orc.dana_run(cfg["idum"], cfg["prob"], cfg["h"], cfg["nst"], cfg["nwr"], cfg["xi"],
                     cfg["yi"], cfg["dist"], cfg["z0"], cfg["zmax"], cfg["dif_sc"],
                     cfg["dif_sei"], cfg["nb_dcut"], cfg["ermak"], cfg["reservoir"])
    return dict(value=rate, unit="particle-steps/s", cores=1, kind="port",
                sample="C1: %d particles (400 CG) x %d Ermak steps, no reservoir, skin 10, the C "
                       "restatement of the reference's routines on one host core, %.2f s; "
                       "on the reference's ermak test (%d particles x %d steps) it does %.3g "
                       "particle-steps/s, where the reference Fortran at -O2 does 1.17e6 on "
                       "one core of the build host" % (n, nst, out["loop_seconds"], e["nat"],
                                                       cfg["nst"], e["nat"] * cfg["nst"] /
                                                       e["loop_seconds"]))


def copy_ceiling(nbytes=1 << 30, reps=10):
    """HBM bandwidth a plain device-to-device copy reaches (bytes read +
    written per second, HIP events around `reps` copies of nbytes)"""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def calibrate_copy(nbytes=1 << 30):
    """One plain device copy of known size: under rocprofv3 --pmc the
    FETCH_SIZE / WRITE_SIZE this kernel reads give the counter's scale."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    log("copy calibration: %d bytes read, %d bytes written" % (nbytes, nbytes))


ROW_KERNELS = {1: "k_listbuild_cells", 2: "k_listbuild<4>", 4: "k_listbuild_wc",
               5: "k_listbuild_wr<3, 7>", 6: "k_listbuild_wr<2, 8>"}


def kernel_file_stem(name):
    """a kernel's name as a file name: k_listbuild_wr<3, 7> -> k_listbuild_wr_3_7"""
    import re
    return re.sub(r"_+$", "", re.sub(r"[^A-Za-z0-9]+", "_", name))


def pmc_summary(fetch_csv, write_csv, kernel="k_listbuild_wr<3, 7>", last=10, copy_bytes=1 << 30):
    """HBM bytes per launch of `kernel` from two rocprofv3 --pmc runs of
    `bench.py --legs batched --steps 1 --warmup 0 --no-cpu --calibrate-copy`
    (FETCH_SIZE, then WRITE_SIZE; counter_collection.csv of each): the mean
    over its last `last` dispatches (the timed launches of the row build over
    every replica), each counter scaled by what it reads for the 1 GiB device
    copy of the calibration (the largest copyBuffer dispatch)."""
    import csv

    def read(path, counter):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        mine = [float(r["Counter_Value"]) for r in rows
                if r["Kernel_Name"].replace("void ", "", 1).startswith(kernel + "(")]
        copy = max(float(r["Counter_Value"]) for r in rows
                   if r["Kernel_Name"].startswith("__amd_rocclr_copyBuffer"))
        return sum(mine[-last:]) / len(mine[-last:]), copy, len(mine)

    f, fc, nf = read(fetch_csv, "FETCH_SIZE")
    w, wc, nw = read(write_csv, "WRITE_SIZE")
    fs, ws = copy_bytes / 1024.0 / fc, copy_bytes / 1024.0 / wc
    rd, wr = f * 1024.0 * fs, w * 1024.0 * ws
    return {"kernel": kernel, "dispatches_seen": [nf, nw], "averaged_last": last,
            "fetch_size_kib_per_launch": f, "write_size_kib_per_launch": w,
            "copy_calibration": {"bytes": copy_bytes, "fetch_size_kib": fc, "write_size_kib": wc},
            "fetch_scale": fs, "write_scale": ws, "read_bytes": rd, "write_bytes": wr,
            "hbm_bytes_per_launch": rd + wr}


def flops_summary(csv_path, last=10):
    """FP32 / FP64 FLOP rates per kernel from one rocprofv3 --pmc run of the
    eight SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_{F32,F64} counters: 64 x (ADD +
    MUL + TRANS + 2 FMA) per dispatch (every lane counted) over that
    dispatch's own duration, averaged over each kernel's last `last`
    dispatches; peaks FP32 157.3, FP64 78.6 TFLOP/s (vector)."""
    import csv
    per = {}
    for r in csv.DictReader(open(csv_path)):
        name = r["Kernel_Name"].replace("void ", "", 1).split("(")[0]
        d = per.setdefault(name, {}).setdefault(r["Dispatch_Id"], {"ns": 0.0})
        d["ns"] = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    out = {}
    for name, ds in per.items():
        keys = sorted(ds, key=int)[-last:]
        f32 = f64 = ns = 0.0
        for k in keys:
            d = ds[k]
            g = lambda c: d.get(c, 0.0)
            f32 += 64 * (g("SQ_INSTS_VALU_ADD_F32") + g("SQ_INSTS_VALU_MUL_F32") +
                         g("SQ_INSTS_VALU_TRANS_F32") + 2 * g("SQ_INSTS_VALU_FMA_F32"))
            f64 += 64 * (g("SQ_INSTS_VALU_ADD_F64") + g("SQ_INSTS_VALU_MUL_F64") +
                         g("SQ_INSTS_VALU_TRANS_F64") + 2 * g("SQ_INSTS_VALU_FMA_F64"))
            ns += d["ns"]
        if ns <= 0:
            continue
        t32, t64 = f32 / ns * 1e-3, f64 / ns * 1e-3
        out[name] = {"dispatches": len(keys), "avg_ms": round(ns / len(keys) * 1e-6, 4),
                     "fp32_tflops": round(t32, 3), "fp64_tflops": round(t64, 3),
                     "fp32_frac_of_157_3": round(t32 / 157.3, 4),
                     "fp64_frac_of_78_6": round(t64 / 78.6, 4)}
    return dict(sorted(out.items(), key=lambda kv: -kv[1]["avg_ms"] * kv[1]["dispatches"]))


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--legs", default="plain,batched,serving,c5")
    ap.add_argument("--drift-steps", type=int, default=10000,
                    help="Ermak steps of each FP32/FP64 energy-drift run on C2 (0 = skip)")
    ap.add_argument("--c5-steps", type=int, default=2000,
                    help="steps of the C5 leg (its configuration: 2e3)")
    ap.add_argument("--c3-full-steps", type=int, default=5000,
                    help="run the C3 ensemble on to this many steps (its configuration: 5e3; "
                         "0 = only the timed steps)")
    ap.add_argument("--serve-seconds", type=float, default=30.0,
                    help="arrival window of the serving leg")
    ap.add_argument("--serve-pool", default=SERVE_POOL,
                    help="serving slot classes, capacity:count,... (the HBM reserved for serving)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--list-kernel", type=int, default=0,
                    help="row-build kernel of the batched leg (dml_set_tuning's list_kernel; "
                         "0 = the autotuner's choice)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print the per-launch HBM traffic of the row build from two "
                         "rocprofv3 --pmc counter_collection.csv files and exit")
    ap.add_argument("--flops-summary", metavar="CSV",
                    help="print FP32/FP64 FLOP rates per kernel from a rocprofv3 --pmc run of "
                         "the eight SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_{F32,F64} counters and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 1 GiB device copy (scale of the rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: only one GPU is measured here (replicas and requests shard across "
              "GPUs with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import dml
    from dml import workloads as W
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None
    legs = args.legs.split(",")
    if args.flops_summary:
        print(json.dumps(flops_summary(args.flops_summary), indent=1))
        return
    if args.pmc_summary:
        print(json.dumps(pmc_summary(*args.pmc_summary,
                                     kernel=ROW_KERNELS[args.list_kernel or 5]), indent=1))
        return
    plain = batched = None
    if "plain" in legs:
        log("plain leg: C2 100k FP64")
        sync()
        plain = plain_leg(dml, W, args.steps, args.warmup, args.drift_steps)
        sync()
        log("plain", json.dumps(plain))
    if "batched" in legs:
        log("batched leg: C3 %d replicas FP32 forces" % args.replicas)
        sync()
        batched = batched_leg(dml, args.replicas, args.steps, args.warmup, args.c3_full_steps,
                              args.list_kernel)
        sync()
        log("batched", json.dumps(batched))
    serving = None
    if "serving" in legs:
        log("serving leg: Poisson 20/s for %.0f s" % args.serve_seconds)
        sync()
        serving = serving_leg(dml, args.serve_seconds, pool_spec=args.serve_pool)
        sync()
        log("serving", json.dumps(serving))
    c5 = None
    if "c5" in legs:
        log("c5 leg: 2M dendrite-regime system, %d steps + 8192 trial energies per step" %
            args.c5_steps)
        sync()
        c5 = c5_leg(dml, args.c5_steps, 5)
        sync()
        log("c5", json.dumps(c5))
    head = batched or plain
    if head is None:  # serving / c5 only
        print(json.dumps({"serving": serving, "c5": c5}), flush=True)
        return
    res = {
        "metric": "particle-steps/s batched vs plain at force err<=1e-5 vs FP64; p99 run latency",
        "value": head["rate"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp64 (fp32 pair forces)" if batched else "fp64",
        "data": "synthetic (seeded bcc electrode + random-sequential Li+ at 0.01/A^3)",
        "config": {"workload": "C3 batched ensemble, %d replicas (2k-50k log-uniform), FP32 forces, "
                               "hipGraph; plain = C2 single 100k system FP64, Ermak (Brownian "
                               "in plain_brownian_value)" % args.replicas,
                   "particles": head["n"], "parallelism": "replicas on one GPU"},
    }
    if plain:
        res["plain_value"] = plain["rate"]
        res["plain_brownian_value"] = plain["brownian_rate"]
        res["force_rel_err_fp32_vs_fp64"] = plain["force_err"]
        res["energy_drift_fp32_vs_fp64"] = plain["energy_drift"]
    if plain and batched:
        res["batched_over_plain"] = batched["rate"] / plain["rate"]
        res["batched_over_plain_brownian"] = batched["rate"] / plain["brownian_rate"]
    res["p99_run_latency_s"] = serving["p99_latency_s"] if serving else None
    if c5:
        res["c5"] = c5
    if serving:
        res["serving"] = serving
    if batched:
        # neighbour-row build: compulsory bytes per ref atom = own position
        # (24) + row count (4) + 4 per list entry, plus per b atom its bucket
        # slot and bucket-ordered position (4 + 24)
        per_atom = 24 + 4 + 4 * batched["mean_nn_rows"] + 28
        traffic = per_atom * batched["rows_at_timing"]
        ach = traffic / (batched["list_ms"] * 1e-3) / 1e9
        # HBM bytes per launch from the two rocprofv3 PMC passes (FETCH_SIZE,
        # WRITE_SIZE, scaled by a 1 GiB copy) of the same 256-replica launch
        measured = None
        kname = ROW_KERNELS.get(batched["tuning"]["list_kernel"], "row build")
        tf = os.path.join(ROOT, "profiles", "%s_c3_traffic.json" % kernel_file_stem(kname))
        if args.replicas == 256 and os.path.exists(tf):
            measured = json.load(open(tf))["hbm_bytes_per_launch"]
        # priced against HBM (its compulsory bytes); what holds it back is
        # latency: its waves wait on memory about half their cycles and
        # issue VALU a fifth of them (instruction counters under profiles/)
        res["roofline"] = {"kernel": kname, "bound": "hbm", "achieved": ach,
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                           "traffic": measured, "avg_launch_ms": batched["list_ms"],
                           "algorithmic_bytes_per_launch": traffic,
                           "limiter": "memory latency (waits ~50% of wave cycles, VALU issue ~20%)"}
        res["kernel_ms"] = {kname: batched["list_ms"], "k_force": batched["force_ms"]}
        res["particles_stepped_mean"] = batched["particles_stepped"]
        res["c3_full_run"] = batched["full_run"]
        res["mean_list_length"] = batched["mean_nn"]
        res["rebuilds_per_step"] = batched["rebuilds_per_rep_step"]
        res["replicas_halted"] = batched["halted"]
        res["reservoir_mix"] = batched["reservoirs"]
        res["tuning"] = batched["tuning"]
        res["gcmc_moves_accepted"] = batched["gcmc_accepted"]
        res["overlap_calls_unfinished"] = batched["overlap_unfinished"]
    try:
        res["hbm_copy_ceiling_gbs"] = copy_ceiling()
    except Exception as e:  # torch without a device: the line still prints
        log("copy ceiling not measured: %s" % e)
    fl = os.path.join(ROOT, "profiles", "flops_c3_256rep.json")
    if batched and args.replicas == 256 and os.path.exists(fl):
        res["flop_rates"] = json.load(open(fl))["rates"]
    if not args.no_cpu:
        log("cpu baseline: oracle on C1")
        res["cpu_baseline"] = cpu_baseline()
    if args.calibrate_copy:
        calibrate_copy()
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
