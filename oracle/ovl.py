"""Pure-Python restatement of the overlap pass, for small cases.

Test infrastructure only (like the rest of oracle/): it is the checker, never
the thing measured or shipped.  Two functions over the same small state:

* `seq_overlap` -- overlap_moveback as the reference runs it
  (src/dana.F90:849-943): passes over the ref list in order while a pass
  reverted a neighbour; an atom with skip = .false. sets skip, walks its row
  in order and, for each entry within rcut, either makes one contact trial
  with a CG atom (deposit as F with probability prob, else move itself back
  to old_cg and clear skip) and stops, or moves the neighbour back to old_cg
  (zeroing its velocity and acceleration, clearing its skip), except, for
  the piston, when both sit at their old_cg (choques3, dana.F90:920-927).
  Every revert is visible at once to the atoms visited after it.
* `round_overlap` -- the device's round schedule (orc.c orc_overlap_rounds,
  dml_kernels.hip k_ov_scan_l / k_ov_apply): positions frozen per round; an
  active atom (skip = 0) with an earlier active atom within rcut waits
  (dominated); the other active atoms walk their rows as above; the writes
  land after the round.

The state is a dict of numpy arrays: pos, old (old_cg), vel, acel (n x 3),
zs (species 1 Li, 2 CG, 3 F), skip (bool), and rows (list of row lists),
order (the ref list), limbo (bool), with box, rcut, prob, z0, piston.
Contact uniforms come from a callable u(i, k, rnd): atom i's k-th trial of
this call, in round rnd (the sequential pass passes its pass number).
"""
import copy

import numpy as np

CG, F = 2, 3


def _nint(x):
    """Fortran idnint: round half away from zero"""
    return np.copysign(np.floor(np.abs(x) + 0.5), x)


def vdist(st, a, b):
    """vdistance with mic on x, y (src/Groups.F90:995-1016)"""
    v = np.array(a, float) - np.array(b, float)
    box = st["box"]
    for k in (0, 1):
        v[k] = v[k] - box[k] * _nint(v[k] / box[k])
    return v


def _d2(st, i, j):
    v = vdist(st, st["pos"][i], st["pos"][j])
    return (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]


def _same(st, i):
    return bool(np.all(st["pos"][i] == st["old"][i]))


def seq_overlap(st0, u, max_passes=100000):
    """the reference's pass (src/dana.F90:849-943) on a copy of st0; returns
    (state, counters, status) with status 0, or 'loop' past max_passes, or
    'z0' for a deposit above z0 (dana.F90:905-908)"""
    st = copy.deepcopy(st0)
    cnt = dict(try_=0, depo=0, choques=0, choques3=0, passes=0)
    ntrial = {}
    again, npass = True, 0
    while again:
        if npass >= max_passes:
            return st, cnt, "loop"
        again = False
        cnt["passes"] += 1
        for i in st["order"]:
            if st["skip"][i]:
                continue
            st["skip"][i] = True
            for j in st["rows"][i]:
                if st["limbo"][j]:
                    continue
                if _d2(st, i, j) > st["rcut"] ** 2:
                    continue
                if st["zs"][j] == CG:
                    cnt["try_"] += 1
                    k = ntrial.get(i, 0)
                    ntrial[i] = k + 1
                    if u(i, k, npass) < st["prob"]:
                        cnt["depo"] += 1
                        st["zs"][i] = F
                        if st["pos"][i][2] > st["z0"]:
                            return st, cnt, "z0"
                    else:
                        st["pos"][i] = st["old"][i].copy()
                        st["skip"][i] = False
                    break
                if st["piston"] and _same(st, j) and _same(st, i):
                    cnt["choques3"] += 1
                    continue
                st["pos"][j] = st["old"][j].copy()
                st["acel"][j] = 0.0
                st["vel"][j] = 0.0
                st["skip"][j] = False
                cnt["choques"] += 1
                again = True
        npass += 1
    return st, cnt, 0


def round_overlap(st0, u, max_rounds=64):
    """the round schedule (orc.c orc_overlap_rounds) on a copy of st0; returns
    (state, counters, status) like seq_overlap ('loop' when rounds ran out)"""
    st = copy.deepcopy(st0)
    cnt = dict(try_=0, depo=0, choques=0, choques3=0, passes=0)
    pos_in = {i: k for k, i in enumerate(st["order"])}
    ntrial = {}
    rc2 = st["rcut"] ** 2
    for rnd in range(max_rounds):
        act = {i for i in st["order"] if not st["skip"][i]}
        if not act:
            return st, cnt, 0
        cnt["passes"] += 1
        dom = set()
        for i in act:
            for j in st["rows"][i]:
                if st["limbo"][j] or j not in act or pos_in.get(j, 1 << 30) >= pos_in[i]:
                    continue
                if _d2(st, i, j) <= rc2:
                    dom.add(i)
                    break
        rev, selfrev, dep = set(), set(), set()
        for i in st["order"]:
            if i not in act or i in dom:
                continue
            for j in st["rows"][i]:
                if st["limbo"][j] or _d2(st, i, j) > rc2:
                    continue
                if st["zs"][j] == CG:
                    cnt["try_"] += 1
                    k = ntrial.get(i, 0)
                    ntrial[i] = k + 1
                    (dep if u(i, k, rnd) < st["prob"] else selfrev).add(i)
                    break
                if st["piston"] and _same(st, j) and _same(st, i):
                    cnt["choques3"] += 1
                    continue
                rev.add(j)
                cnt["choques"] += 1
        for i in act - dom:
            st["skip"][i] = True
        for i in sorted(dep):
            cnt["depo"] += 1
            st["zs"][i] = F
            if st["pos"][i][2] > st["z0"]:
                return st, cnt, "z0"
        for i in sorted(selfrev):
            st["pos"][i] = st["old"][i].copy()
            st["skip"][i] = False
        for i in sorted(rev):
            st["pos"][i] = st["old"][i].copy()
            st["acel"][i] = 0.0
            st["vel"][i] = 0.0
            st["skip"][i] = False
        if not rev and not dom:
            return st, cnt, 0
    return st, cnt, "loop"


def round_interference(st0, u, max_rounds=64):
    """Replays the round schedule and returns True when some round makes a
    move (a revert, or a failed contact trial's move back) that lies, before
    or after it, within rcut of an atom the same round scans later in ref
    order and whose row lists the moved atom.  The reference pass would have
    shown that atom the move; the schedule does not.  Without such a move
    the two give the same result at prob = 1 (tests/test_overlap_rounds.py)."""
    st = copy.deepcopy(st0)
    pos_in = {i: k for k, i in enumerate(st["order"])}
    rc2 = st["rcut"] ** 2
    ntrial = {}
    for rnd in range(max_rounds):
        act = {i for i in st["order"] if not st["skip"][i]}
        if not act:
            return False
        dom = set()
        for i in act:
            for j in st["rows"][i]:
                if (not st["limbo"][j] and j in act and pos_in.get(j, 1 << 30) < pos_in[i]
                        and _d2(st, i, j) <= rc2):
                    dom.add(i)
                    break
        scanned = [i for i in st["order"] if i in act and i not in dom]
        moves, rev, selfrev, dep = [], set(), set(), set()
        for i in scanned:
            for j in st["rows"][i]:
                if st["limbo"][j] or _d2(st, i, j) > rc2:
                    continue
                if st["zs"][j] == CG:
                    k = ntrial.get(i, 0)
                    ntrial[i] = k + 1
                    if u(i, k, rnd) < st["prob"]:
                        dep.add(i)
                    else:
                        selfrev.add(i)
                        moves.append((pos_in[i], i))
                    break
                if st["piston"] and _same(st, j) and _same(st, i):
                    continue
                rev.add(j)
                moves.append((pos_in[i], j))
        for pm, w in moves:
            if np.array_equal(st["pos"][w], st["old"][w]):
                continue
            for i in scanned:
                if pos_in[i] <= pm or w not in st["rows"][i] and i != w:
                    continue
                if i == w:
                    return True
                a = vdist(st, st["pos"][i], st["pos"][w])
                b = vdist(st, st["pos"][i], st["old"][w])
                if min(a @ a, b @ b) <= rc2:
                    return True
        for i in act - dom:
            st["skip"][i] = True
        for i in dep:
            st["zs"][i] = F
        for i in selfrev:
            st["pos"][i] = st["old"][i].copy()
            st["skip"][i] = False
        for i in rev:
            st["pos"][i] = st["old"][i].copy()
            st["acel"][i] = 0.0
            st["vel"][i] = 0.0
            st["skip"][i] = False
        if not rev and not dom:
            return False
    return False
