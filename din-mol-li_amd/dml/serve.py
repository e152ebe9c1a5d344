"""Continuous-batching server for deposition runs on one GPU.

A request is one system (atoms, box, integrator, reservoir) and a number of
main-loop steps.  All requests share one libdml context whose replicas are
slots of a few capacity classes, reserved at start (that reservation is the
HBM budget: `SlotPool.hbm_bytes`).  A request that arrives is admitted into
the smallest free slot that holds it (`dml_reconfigure` + `dml_set_state` +
`dml_set_step_limit`); requests that do not fit anywhere yet wait, without
blocking smaller requests behind them.  Every `segment` steps of `dml_run`
(one hipGraph launch per step over every slot, paused slots skipped) the loop
collects the requests that reached their step count, reads their final state
and frees their slots.  A replica's results do not depend on the others, so a
request's result is the same alone and in the mix.
"""
import collections
import time

import numpy as np

from . import Context, dana_config, lib

# bytes per slot held on the device: FP64 state (22 x 8), narrow per-slot
# arrays (~30), the FP32 mirror (16), plus the neighbour row (4 x maxnn)
STATE_BYTES_PER_SLOT = 22 * 8 + 30 + 16


class Request:
    __slots__ = ("rid", "t_arrival", "cfg", "pos", "species", "steps", "need", "t_admit",
                 "t_done", "rep", "result", "status", "n", "steps_done")

    def __init__(self, rid, t_arrival, cfg, pos, species, steps, need=None):
        self.rid = rid
        self.t_arrival = t_arrival
        self.cfg = cfg
        self.pos = pos
        self.species = species
        self.steps = int(steps)
        self.n = len(pos)
        self.need = int(need or len(pos))
        self.t_admit = self.t_done = None
        self.rep = None
        self.result = None
        self.status = None
        self.steps_done = 0


class SlotPool:
    """classes: list of (capacity, count, box) — box is the largest system box
    a slot of that class will hold (it sizes the slot's cell memory)."""

    def __init__(self, classes, skin=10.0):
        cfgs, self.cap = [], []
        for cap, count, box in classes:
            for _ in range(count):
                cfgs.append(dana_config(box, skin, cap))
                self.cap.append(int(cap))
        self.ctx = Context(cfgs)
        self.free = set(range(len(cfgs)))
        for r in range(len(cfgs)):
            self.ctx.set_active(False, r)
        self.order = sorted(range(len(cfgs)), key=lambda r: self.cap[r])

    def hbm_bytes(self, row_entries=1024):
        """device bytes the slots reserve (state plus neighbour rows)"""
        total = sum(self.cap)
        return total * (STATE_BYTES_PER_SLOT + 4 * row_entries)

    def slot_for(self, need):
        for r in self.order:
            if r in self.free and self.cap[r] >= need:
                return r
        return None

    def close(self):
        self.ctx.close()


def serve(pool, requests, segment=8, clock=time.perf_counter, on_done=None):
    """Run the requests (sorted by t_arrival, seconds from the start) through
    the pool; returns the completed requests and the wall time."""
    ctx = pool.ctx
    pending = collections.deque(sorted(requests, key=lambda q: q.t_arrival))
    queue = []
    running = {}
    done = []
    t0 = clock()
    while pending or queue or running:
        now = clock() - t0
        while pending and pending[0].t_arrival <= now:
            queue.append(pending.popleft())
        still = []
        for q in queue:
            r = pool.slot_for(q.need)
            if r is None:
                still.append(q)
                continue
            cfg = q.cfg
            cfg.capacity = min(pool.cap[r], max(cfg.capacity, q.need))
            ctx.reconfigure(cfg, r)
            ctx.set_state(q.pos, species=q.species, rep=r)
            ctx.set_step_limit(q.steps, r)
            pool.free.discard(r)
            q.rep, q.t_admit = r, clock() - t0
            running[r] = q
        queue = still
        if not running:
            if pending:
                time.sleep(max(0.0, pending[0].t_arrival - (clock() - t0)))
            continue
        ctx.run(segment, graph=True)
        steps, status, live = ctx.progress()
        for r, q in list(running.items()):
            if live[r]:
                continue
            q.status = int(status[r])
            q.steps_done = int(steps[r])
            q.result = ctx.state(r)
            q.t_done = clock() - t0
            ctx.set_active(False, r)
            del running[r]
            pool.free.add(r)
            done.append(q)
            if on_done:
                on_done(q)
    return done, clock() - t0


def summarize(done, wall):
    lat = np.array([q.t_done - q.t_arrival for q in done])
    wait = np.array([q.t_admit - q.t_arrival for q in done])
    # particle-steps executed (a run stopped by a fatal condition of the
    # reference counts the steps it made)
    ps = sum(q.n * q.steps_done for q in done)
    bad = collections.Counter(q.status for q in done if q.status != 0)
    return dict(completed=len(done), failed=sum(bad.values()),
                failed_by_status={str(k): v for k, v in sorted(bad.items())},
                p50_latency_s=float(np.percentile(lat, 50)) if len(lat) else None,
                p99_latency_s=float(np.percentile(lat, 99)) if len(lat) else None,
                mean_queue_s=float(wait.mean()) if len(wait) else None,
                p99_queue_s=float(np.percentile(wait, 99)) if len(wait) else None,
                particle_steps_per_s=ps / wall if wall > 0 else None, wall_s=wall)


__all__ = ["Request", "SlotPool", "serve", "summarize", "lib"]
