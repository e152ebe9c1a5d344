"""Python binding of libdml (the HIP implementation of the din-mol-Li hot path).

This module is a thin ctypes layer over the C ABI of include/dml.h, mirroring
the reference's operator sequence (test_update, fuerza, cbrownian_hs,
ermak_a/ermak_b, overlap_moveback, gcmc_run).  It loads the in-tree
libdml.so and raises if that library is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdml.so")

BROWNIAN, ERMAK = 0, 1
PISTON, CHUNKS, GCMC, NO_RESERVOIR = 0, 1, 2, 3
FP64, FP32_FORCE = 0, 1
OVERLAP_EXACT, OVERLAP_ROUNDS = 0, 1
GCMC_BATCHED, GCMC_ALTERNATE = 1, 2
GCMC_BATCH_MAX = 8192  # DML_GCMC_BATCH_MAX
FL_ALIVE, FL_LIMBO, FL_REF, FL_SKIP, FL_GCMC = 1, 2, 4, 8, 16

STATUS = {0: "ok", -1: "bad argument", -2: "HIP runtime error", -3: "out of device memory",
          -4: "particle out of tessellation", -5: "neighbour row capacity exceeded",
          -6: "deposit above z0", -7: "chosen particle does not exist",
          -8: "overlap resolution did not finish", -9: "bad state", -10: "unsupported"}

# every symbol include/dml.h declares
EXPORTS = ["dml_create", "dml_create_ensemble", "dml_destroy", "dml_sync", "dml_strerror",
           "dml_num_replicas", "dml_total_slots", "dml_set_state", "dml_set_order",
           "dml_get_state", "dml_get_ref_order", "dml_set_zmax", "dml_get_zmax", "dml_set_box",
           "dml_test_update", "dml_update", "dml_cells", "dml_neighbors", "dml_forces",
           "dml_brownian", "dml_ermak_a", "dml_ermak_b", "dml_ref_offset",
           "dml_overlap_moveback", "dml_overlap_rounds", "dml_f_to_cg", "dml_gcmc",
           "dml_gcmc_trial_energy", "dml_gcmc_trial_energy_noalloc", "dml_run", "dml_get_counters", "dml_list_stats",
           "dml_last_run_ms", "dml_time_kernel", "dml_set_tuning",
           "dml_replica_status", "dml_get_gcmc_order", "dml_gcmc_outcomes", "dml_get_slots",
           "dml_calc_rho", "dml_observables", "dml_maxz", "dml_end_step",
           "dml_get_tuning", "dml_reconfigure", "dml_set_step_limit", "dml_set_active",
           "dml_progress", "dml_set_chunk", "dml_bloques"]
K_LISTBUILD, K_FORCE, K_CELLS, K_DISPMAX = 0, 1, 2, 3
NOALLOC_MAX_TRIAL = 8192  # DML_NOALLOC_MAX_TRIAL


class DmlError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        super().__init__("%s: %s (%d)" % (where, STATUS.get(status, "?"), status))


class Config(C.Structure):
    _fields_ = [("box", C.c_double * 3), ("rcut", C.c_double), ("skin", C.c_double),
                ("eps", C.c_double * 9), ("r0", C.c_double * 9), ("mass", C.c_double * 3),
                ("h", C.c_double), ("gamma", C.c_double), ("temp", C.c_double),
                ("dif_sc", C.c_double), ("dif_sei", C.c_double), ("z_switch", C.c_double),
                ("prob", C.c_double), ("z0", C.c_double), ("zmax", C.c_double),
                ("integrator", C.c_int32), ("reservoir", C.c_int32), ("precision", C.c_int32),
                ("overlap_mode", C.c_int32), ("capacity", C.c_int32),
                ("max_neighbors", C.c_int32), ("seed", C.c_uint64), ("act", C.c_double),
                ("nadj", C.c_int32), ("lanes_per_atom", C.c_int32),
                ("gcmc_mode", C.c_int32), ("reserved", C.c_int32)]


class Counters(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("try_", "depo", "choques", "choques3", "nupd",
                                         "ins", "del_", "passes", "particle_steps",
                                         "row_entries", "rows", "ov_unfinished")]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libdml.so not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.dml_create_ensemble.argtypes = [C.c_int32, C.POINTER(Config), C.POINTER(vp)]
        L.dml_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
        L.dml_destroy.argtypes = [vp]
        L.dml_strerror.restype = C.c_char_p
        L.dml_total_slots.restype = C.c_int64
        L.dml_total_slots.argtypes = [vp]
        L.dml_ref_offset.restype = C.c_int64
        L.dml_ref_offset.argtypes = [vp, C.c_int32]
        L.dml_set_state.argtypes = [vp, C.c_int32, C.c_int32] + [vp] * 7
        L.dml_set_order.argtypes = [vp, C.c_int32, C.c_int32, vp, vp]
        L.dml_get_state.argtypes = [vp, C.c_int32, C.c_int32] + [vp] * 9
        L.dml_get_ref_order.argtypes = [vp, C.c_int32, vp, vp]
        L.dml_set_zmax.argtypes = [vp, C.c_int32, C.c_double, C.c_double]
        L.dml_get_zmax.argtypes = [vp, C.c_int32, vp]
        L.dml_set_box.argtypes = [vp, C.c_int32, vp]
        L.dml_test_update.argtypes = [vp, vp]
        L.dml_update.argtypes = [vp]
        L.dml_cells.argtypes = [vp, C.c_int32, vp, vp, vp, vp]
        L.dml_neighbors.argtypes = [vp, C.c_int32, vp, vp, vp, C.c_int64, vp]
        L.dml_forces.argtypes = [vp]
        L.dml_brownian.argtypes = [vp, vp, vp]
        L.dml_ermak_a.argtypes = [vp, vp, vp]
        L.dml_ermak_b.argtypes = [vp]
        L.dml_overlap_moveback.argtypes = [vp, vp, C.c_int64, C.c_int32]
        L.dml_overlap_rounds.argtypes = [vp, vp, C.c_int32]
        L.dml_f_to_cg.argtypes = [vp]
        L.dml_gcmc.argtypes = [vp, C.c_int32, C.c_double, C.c_int32, vp, vp]
        L.dml_gcmc_trial_energy.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp]
        L.dml_gcmc_outcomes.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp, vp]
        L.dml_gcmc_trial_energy_noalloc.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, vp]
        L.dml_run.argtypes = [vp, C.c_int32, C.c_int32]
        L.dml_get_counters.argtypes = [vp, C.c_int32, C.POINTER(Counters)]
        L.dml_list_stats.argtypes = [vp, C.c_int32, vp, vp, vp]
        L.dml_last_run_ms.argtypes = [vp, vp]
        L.dml_time_kernel.argtypes = [vp, C.c_int32, C.c_int32, vp]
        L.dml_set_tuning.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.dml_get_tuning.argtypes = [vp, vp, vp, vp, vp]
        L.dml_reconfigure.argtypes = [vp, C.c_int32, C.POINTER(Config)]
        L.dml_set_step_limit.argtypes = [vp, C.c_int32, C.c_int64]
        L.dml_set_active.argtypes = [vp, C.c_int32, C.c_int32]
        L.dml_progress.argtypes = [vp, vp, vp, vp]
        L.dml_set_chunk.argtypes = [vp, C.c_int32, C.c_int32, vp, C.c_double, C.c_double]
        L.dml_bloques.argtypes = [vp]
        L.dml_replica_status.argtypes = [vp, C.c_int32, vp]
        L.dml_sync.argtypes = [vp]
        L.dml_get_gcmc_order.argtypes = [vp, C.c_int32, vp, vp]
        L.dml_get_slots.argtypes = [vp, C.c_int32, vp, vp, vp, vp]
        L.dml_calc_rho.argtypes = [vp, vp]
        L.dml_maxz.argtypes = [vp]
        L.dml_end_step.argtypes = [vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _chk(st, where):
    if st != 0:
        raise DmlError(st, where)


def dana_config(box, skin, capacity, h=1e-2, prob=1.0, z0=100.0, zmax=200.0, dif_sc=250.0,
                dif_sei=250.0, integrator=ERMAK, reservoir=PISTON, precision=FP64,
                overlap_mode=OVERLAP_EXACT, seed=104012, rcut=3.2, max_neighbors=0, act=0.0,
                nadj=0, lanes_per_atom=0, gcmc_mode=0):
    """A Config with dana's constants (dana.F90:12-17, 86-100, 113)."""
    c = Config()
    c.box[:] = list(box)
    c.rcut, c.skin = rcut, skin
    eps = np.zeros((3, 3)); r0 = np.zeros((3, 3))
    r0[1, 0] = r0[0, 1] = 3.5
    eps[0, 0] = 2313.6; r0[0, 0] = 3.2
    eps[2, 2] = 121.0; r0[2, 2] = 3.61
    eps[0, 2] = eps[2, 0] = 529.1
    r0[0, 2] = r0[2, 0] = 1.564
    c.eps[:] = list(eps.ravel()); c.r0[:] = list(r0.ravel())
    c.mass[:] = [6.94, 6.94, 6.94]
    c.h, c.gamma, c.temp = h, 1.0, 300.0
    c.dif_sc, c.dif_sei, c.z_switch = dif_sc, dif_sei, 80.0
    c.prob, c.z0, c.zmax = prob, z0, zmax
    c.integrator, c.reservoir, c.precision, c.overlap_mode = integrator, reservoir, precision, overlap_mode
    c.capacity = capacity
    c.max_neighbors = max_neighbors
    c.seed = seed
    c.act, c.nadj, c.lanes_per_atom = act, nadj, lanes_per_atom
    c.gcmc_mode = gcmc_mode
    return c


class Context:
    """One libdml context (one or more replicas on the current GPU)."""

    def __init__(self, configs):
        if isinstance(configs, Config):
            configs = [configs]
        self.L = lib()
        arr = (Config * len(configs))(*configs)
        h = C.c_void_p()
        _chk(self.L.dml_create_ensemble(len(configs), arr, C.byref(h)), "dml_create_ensemble")
        self.h = h
        self.nrep = len(configs)
        self.n = [0] * self.nrep
        self.total = self.L.dml_total_slots(self.h)

    def close(self):
        if self.h:
            self.L.dml_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ref_offset(self, rep=0):
        return self.L.dml_ref_offset(self.h, rep)

    def set_state(self, pos, species=None, vel=None, acel=None, pos_old=None, old_cg=None,
                  flags=None, rep=0):
        pos = np.asarray(pos, np.float64)
        n = len(pos)

        def soa(a):
            return None if a is None else np.ascontiguousarray(np.asarray(a, np.float64).T.reshape(-1))
        arrs = [soa(pos), soa(vel), soa(acel), soa(pos_old), soa(old_cg),
                None if species is None else np.ascontiguousarray(species, np.int8),
                None if flags is None else np.ascontiguousarray(flags, np.uint8)]
        _chk(self.L.dml_set_state(self.h, rep, n, *[_p(a) for a in arrs]), "dml_set_state")
        self.n[rep] = n

    def set_order(self, ref_order=None, bkey=None, rep=0):
        ro = None if ref_order is None else np.ascontiguousarray(ref_order, np.int32)
        bk = None if bkey is None else np.ascontiguousarray(bkey, np.int32)
        _chk(self.L.dml_set_order(self.h, rep, 0 if ro is None else len(ro), _p(ro), _p(bk)),
             "dml_set_order")

    def nslots(self, rep=0):
        """Slots in use (hs%amax): grows when GCMC appends atoms."""
        n = C.c_int32()
        _chk(self.L.dml_get_slots(self.h, rep, C.byref(n), None, None, None), "dml_get_slots")
        self.n[rep] = n.value
        return n.value

    def slots(self, rep=0):
        """(bkey, pid, seq) per slot: the hs%b index, the particle id and the
        creation number (sys list order)."""
        n = self.nslots(rep)
        bk = np.zeros(max(n, 1), np.int32)
        pid = np.zeros(max(n, 1), np.uint32)
        seq = np.zeros(max(n, 1), np.int32)
        nn = C.c_int32()
        _chk(self.L.dml_get_slots(self.h, rep, C.byref(nn), _p(bk), _p(pid), _p(seq)),
             "dml_get_slots")
        return bk[:n], pid[:n], seq[:n]

    def gcmc_order(self, rep=0):
        ng = C.c_int32()
        _chk(self.L.dml_get_gcmc_order(self.h, rep, C.byref(ng), None), "dml_get_gcmc_order")
        out = np.zeros(max(ng.value, 1), np.int32)
        _chk(self.L.dml_get_gcmc_order(self.h, rep, C.byref(ng), _p(out)), "dml_get_gcmc_order")
        return out[:ng.value]

    def gcmc(self, act, nadj, u=None, g=None, rep=0):
        """nadj GCMC trials (gcmc_run, dana.F90:590-713) on replica rep; u is
        (nadj, 5) uniforms, g (nadj, 3) normals, or both None for Philox."""
        uu = None if u is None else np.ascontiguousarray(u, np.float64).reshape(-1)
        gg = None if g is None else np.ascontiguousarray(g, np.float64).reshape(-1)
        if uu is not None and (len(uu) < 5 * nadj or gg is None or len(gg) < 3 * nadj):
            raise ValueError("u needs 5 and g 3 values per trial")
        _chk(self.L.dml_gcmc(self.h, rep, float(act), int(nadj), _p(uu), _p(gg)), "dml_gcmc")
        self.nslots(rep)

    def state(self, rep=0):
        n = self.nslots(rep)
        d = {k: np.zeros(3 * n) for k in ("pos", "vel", "acel", "force", "pos_old", "old_cg")}
        d["epot"] = np.zeros(n)
        d["species"] = np.zeros(n, np.int8)
        d["flags"] = np.zeros(n, np.uint8)
        _chk(self.L.dml_get_state(self.h, rep, n, _p(d["pos"]), _p(d["vel"]), _p(d["acel"]),
                                  _p(d["force"]), _p(d["epot"]), _p(d["pos_old"]),
                                  _p(d["old_cg"]), _p(d["species"]), _p(d["flags"])),
             "dml_get_state")
        for k in ("pos", "vel", "acel", "force", "pos_old", "old_cg"):
            d[k] = d[k].reshape(3, n).T.copy()
        return d

    def ref_order(self, rep=0):
        nref = C.c_int32()
        _chk(self.L.dml_get_ref_order(self.h, rep, C.byref(nref), None), "dml_get_ref_order")
        out = np.zeros(max(nref.value, 1), np.int32)
        _chk(self.L.dml_get_ref_order(self.h, rep, C.byref(nref), _p(out)), "dml_get_ref_order")
        return out[:nref.value]

    def set_zmax(self, zmax, z0, rep=0):
        _chk(self.L.dml_set_zmax(self.h, rep, zmax, z0), "dml_set_zmax")

    def zmax(self, rep=0):
        z = C.c_double()
        _chk(self.L.dml_get_zmax(self.h, rep, C.byref(z)), "dml_get_zmax")
        return z.value

    def set_box(self, box, rep=0):
        b = np.ascontiguousarray(box, np.float64)
        _chk(self.L.dml_set_box(self.h, rep, _p(b)), "dml_set_box")

    def test_update(self):
        reb = np.zeros(self.nrep, np.int32)
        _chk(self.L.dml_test_update(self.h, _p(reb)), "dml_test_update")
        return reb

    def update(self):
        _chk(self.L.dml_update(self.h), "dml_update")

    def cells(self, rep=0):
        nc = np.zeros(3, np.int32); cell = np.zeros(3); uc = C.c_int32()
        n = self.nslots(rep)
        ci = np.zeros(3 * max(n, 1), np.int32)
        _chk(self.L.dml_cells(self.h, rep, _p(nc), _p(cell), C.byref(uc), _p(ci)), "dml_cells")
        return nc, cell, bool(uc.value), ci[:3 * n].reshape(3, n).T.copy()

    def lists(self, rep=0):
        n = self.nslots(rep)
        tot = C.c_int64()
        nn = np.zeros(max(n, 1), np.int32)
        _chk(self.L.dml_neighbors(self.h, rep, _p(nn), None, None, 0, C.byref(tot)), "dml_neighbors")
        rp = np.zeros(n + 1, np.int64)
        idx = np.zeros(max(tot.value, 1), np.int32)
        _chk(self.L.dml_neighbors(self.h, rep, _p(nn), _p(rp), _p(idx), tot.value, C.byref(tot)),
             "dml_neighbors")
        return nn[:n], rp, idx[:tot.value]

    def forces(self):
        _chk(self.L.dml_forces(self.h), "dml_forces")

    def _padded(self, a, per):
        if a is None:
            return None
        a = np.asarray(a, np.float64).reshape(-1)
        out = np.zeros(per * self.total)
        out[:len(a)] = a
        return out

    def brownian(self, noise=None, u_dep=None):
        """noise/u_dep are indexed by global ref position (replica 0 starts at 0)."""
        nz, ud = self._padded(noise, 3), self._padded(u_dep, 1)
        _chk(self.L.dml_brownian(self.h, _p(nz), _p(ud)), "dml_brownian")

    def ermak_a(self, noise=None, u_dep=None):
        nz, ud = self._padded(noise, 6), self._padded(u_dep, 1)
        _chk(self.L.dml_ermak_a(self.h, _p(nz), _p(ud)), "dml_ermak_a")

    def ermak_b(self):
        _chk(self.L.dml_ermak_b(self.h), "dml_ermak_b")

    def overlap(self, u_contact=None, n_u=0, max_passes=100000):
        """u_contact: array (nrep, n_u) of uniforms consumed in contact order."""
        u = None
        if u_contact is not None:
            u = np.ascontiguousarray(np.asarray(u_contact, np.float64).reshape(self.nrep, -1))
            n_u = u.shape[1]
        _chk(self.L.dml_overlap_moveback(self.h, _p(u), n_u, max_passes), "dml_overlap_moveback")

    def overlap_rounds(self, u_round=None, max_rounds=8):
        """u_round: array (max_rounds, total_slots) or None."""
        u = None if u_round is None else np.ascontiguousarray(u_round, np.float64)
        _chk(self.L.dml_overlap_rounds(self.h, _p(u), max_rounds), "dml_overlap_rounds")

    def f_to_cg(self):
        _chk(self.L.dml_f_to_cg(self.h), "dml_f_to_cg")

    def calc_rho(self):
        """calc_rho (dana.F90:521-549) of every replica"""
        rho = np.zeros(self.nrep)
        _chk(self.L.dml_calc_rho(self.h, _p(rho)), "dml_calc_rho")
        return rho

    def observables(self):
        """salida's per-frame observables of every replica, shape (nrep, 4):
        energia (sum of epot over sys), the kion temperature, the count of
        non-CG atoms, the atom count (dana.F90:1141-1183, 1342-1376)"""
        out = np.zeros((self.nrep, 4))
        _chk(self.L.dml_observables(self.h, _p(out)), "dml_observables")
        return out

    def maxz(self):
        """maxz (dana.F90:776-794) of the piston replicas"""
        _chk(self.L.dml_maxz(self.h), "dml_maxz")

    def set_chunk(self, pos, dist, z1, rep=0):
        """config_chunk: the block template (chunk.xyz positions, (n, 3)),
        the reservoir step dist (+ rcut) and the sensor top z1"""
        p = np.ascontiguousarray(np.asarray(pos, np.float64).T.reshape(-1))
        _chk(self.L.dml_set_chunk(self.h, rep, len(pos), _p(p), float(dist), float(z1)),
             "dml_set_chunk")

    def bloques(self):
        """bloques (dana.F90:716-773) of every chunks replica"""
        _chk(self.L.dml_bloques(self.h), "dml_bloques")

    def end_step(self):
        _chk(self.L.dml_end_step(self.h), "dml_end_step")

    def gcmc_trial_energy(self, pos3, rep=0, exclude=None, noalloc=False):
        """trial energies at pos3 (t, 3); exclude (t,) slots left out (-1 none);
        pos3 None: removals of the exclude atoms at their current positions.
        noalloc: dml_gcmc_trial_energy_noalloc (at most DML_NOALLOC_MAX_TRIAL
        trials, staged in the library's fixed device array)"""
        p = None if pos3 is None else np.ascontiguousarray(pos3, np.float64)
        t = len(p) if p is not None else len(exclude)
        ex = None if exclude is None else np.ascontiguousarray(exclude, np.int32)
        du = np.zeros(max(t, 1)); hard = np.zeros(max(t, 1), np.int32)
        fn = "dml_gcmc_trial_energy_noalloc" if noalloc else "dml_gcmc_trial_energy"
        _chk(getattr(self.L, fn)(self.h, rep, t, _p(p), _p(ex), _p(du), _p(hard)), fn)
        return du[:t], hard[:t]

    def gcmc_outcomes(self, ntrial, rep=0):
        """the trials of the last batched gcmc call: dict of kind, accepted,
        slot, du, or None when that call ran trial by trial"""
        kind = np.zeros(max(ntrial, 1), np.int32); acc = np.zeros_like(kind)
        slot = np.zeros_like(kind); du = np.zeros(max(ntrial, 1))
        b = C.c_int32()
        _chk(self.L.dml_gcmc_outcomes(self.h, rep, ntrial, _p(kind), _p(acc), _p(slot), _p(du),
                                      C.byref(b)), "dml_gcmc_outcomes")
        if not b.value:
            return None
        return dict(kind=kind[:ntrial], accepted=acc[:ntrial], slot=slot[:ntrial], du=du[:ntrial])

    def run(self, nsteps, graph=True):
        _chk(self.L.dml_run(self.h, nsteps, int(graph)), "dml_run")
        ms = C.c_double()
        self.L.dml_last_run_ms(self.h, C.byref(ms))
        return ms.value

    def counters(self, rep=0):
        c = Counters()
        _chk(self.L.dml_get_counters(self.h, rep, C.byref(c)), "dml_get_counters")
        return {k: getattr(c, k) for k, _ in Counters._fields_}

    def list_stats(self, rep=0):
        m = C.c_double(); rb = C.c_int64(); st = C.c_int64()
        _chk(self.L.dml_list_stats(self.h, rep, C.byref(m), C.byref(rb), C.byref(st)),
             "dml_list_stats")
        return m.value, rb.value, st.value

    def replica_status(self, rep=0):
        st = C.c_int32()
        _chk(self.L.dml_replica_status(self.h, rep, C.byref(st)), "dml_replica_status")
        return st.value

    def time_kernel(self, which, reps=10):
        ms = C.c_double()
        _chk(self.L.dml_time_kernel(self.h, which, reps, C.byref(ms)), "dml_time_kernel")
        return ms.value

    def set_tuning(self, lanes_per_atom=0, overlap_rounds=0, overlap_lanes=0, list_kernel=0):
        """launch shapes (0 keeps a value); list_kernel 1 = cell tiles, 2 = wave per atom"""
        _chk(self.L.dml_set_tuning(self.h, lanes_per_atom, overlap_rounds, overlap_lanes,
                                   list_kernel), "dml_set_tuning")

    def tuning(self):
        v = [C.c_int32() for _ in range(4)]
        _chk(self.L.dml_get_tuning(self.h, *[C.byref(x) for x in v]), "dml_get_tuning")
        return dict(lanes_per_atom=v[0].value, overlap_rounds=v[1].value,
                    overlap_lanes=v[2].value, list_kernel=v[3].value)

    def reconfigure(self, cfg, rep):
        """a new system description for replica rep (paused until set_state)"""
        _chk(self.L.dml_reconfigure(self.h, rep, C.byref(cfg)), "dml_reconfigure")
        self.n[rep] = 0

    def set_step_limit(self, limit, rep):
        _chk(self.L.dml_set_step_limit(self.h, rep, int(limit)), "dml_set_step_limit")

    def progress(self):
        """(steps, status, running) arrays over the replicas"""
        st = np.zeros(self.nrep, np.int64)
        ss = np.zeros(self.nrep, np.int32)
        ru = np.zeros(self.nrep, np.int32)
        _chk(self.L.dml_progress(self.h, _p(st), _p(ss), _p(ru)), "dml_progress")
        return st, ss, ru.astype(bool)

    def set_active(self, active, rep):
        _chk(self.L.dml_set_active(self.h, rep, int(bool(active))), "dml_set_active")

    def sync(self):
        _chk(self.L.dml_sync(self.h), "dml_sync")
