/*
 * dml.h -- C ABI of the MI355X (gfx950) implementation of the din-mol-Li
 * per-step hot path.
 *
 * The reference is a Fortran program whose hot routines are internal
 * procedures of `program dana` or type-bound procedures of the gems modules;
 * it has no C ABI.  Each entry point below replaces one of those routines
 * (cited per function) so that an ISO_C_BINDING interface module
 * (din-mol-li_amd/fortran/dml_f.F90) can call it in place of the CPU routine.
 *
 * Conventions
 *  - Every function returns 0 (DML_OK) or a negative dml_status; nothing
 *    calls exit()/stop.  dml_strerror() names the status.
 *  - A context owns its device buffers and one HIP stream.  Calls on one
 *    context are asynchronous with respect to the host except those that
 *    return host data (they synchronise the stream).  Distinct contexts may be
 *    used from distinct host threads; one context is single-threaded.
 *  - A context holds one or more independent replicas (systems).  Host arrays
 *    are per replica, structure-of-arrays, column-major n x 3 like a Fortran
 *    `real(c_double) :: pos(n,3)`: element (i,k) at [k*n + i].
 *  - Particle index = slot, the index the reference's ngroup `hs` gives an
 *    atom (0-based here); neighbour lists store slots.
 *  - Species: 1 = Li (ion), 2 = CG (deposited metal), 3 = F (deposited this
 *    step), as set_z in dana.F90:81-84.
 *  - Injected random numbers (noise, uniforms) are optional; NULL means the
 *    device Philox4x32-10 stream keyed by (replica seed, particle id, step).
 */
#ifndef DML_H
#define DML_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dml_ctx dml_ctx;

typedef enum dml_status {
  DML_OK = 0,
  DML_E_ARG = -1,           /* bad argument */
  DML_E_HIP = -2,           /* HIP runtime error */
  DML_E_NOMEM = -3,         /* device allocation failed / capacity exceeded */
  DML_E_OUT_OF_TESS = -4,   /* 'Particle out of tessellation', Cells.F90:293-294 */
  DML_E_LIST_FULL = -5,     /* a neighbour row exceeded max_neighbors */
  DML_E_DEPOSIT_Z0 = -6,    /* 'supero z0', dana.F90:905-908 */
  DML_E_NO_PARTICLE = -7,   /* 'Chosen particle does not exists', dana.F90:701 */
  DML_E_OVERLAP = -8,       /* overlap resolution did not finish in its bound */
  DML_E_STATE = -9,         /* call not valid in the current state */
  DML_E_UNSUPPORTED = -10
} dml_status;

enum { DML_BROWNIAN = 0, DML_ERMAK = 1 };
enum { DML_PISTON = 0, DML_CHUNKS = 1, DML_GCMC = 2, DML_NO_RESERVOIR = 3 };
enum { DML_FP64 = 0, DML_FP32_FORCE = 1 };
enum { DML_OVERLAP_EXACT = 0, DML_OVERLAP_ROUNDS = 1 };
/* gcmc_mode bits.  BATCHED: the nadj trials of a call are evaluated in
 * parallel (points, hard-core neighbours, FP64 trial energies) and then
 * committed in trial order; the result is the sequential gcmc_run's, bit for
 * bit (calls the batch cannot take fall back to the sequential kernel).
 * ALTERNATE: trial t inserts when t is even and deletes when t is odd
 * instead of the ran < 0.5 coin (dana.F90:626), so a call of nadj trials is
 * nadj/2 insertions plus nadj/2 deletions. */
enum { DML_GCMC_BATCHED = 1, DML_GCMC_ALTERNATE = 2 };
/* largest nadj a batched call takes (larger calls run sequentially) */
#define DML_GCMC_BATCH_MAX 8192

/* flags bits per slot */
#define DML_FL_ALIVE 1u
#define DML_FL_LIMBO 2u /* detached while listed (Neighbor.F90:262-267) */
#define DML_FL_REF 4u   /* member of hs%ref (moves, owns a list) */
#define DML_FL_SKIP 8u  /* atom%skip (Groups.F90:296) */
#define DML_FL_GCMC 16u /* member of the GCMC group (dana.F90:130-137) */

typedef struct dml_config {
  double box[3];     /* box(1:3) of gems_program_types (box(3) = tbox(3,3)) */
  double rcut;       /* hs%rcut, 3.2 in dana (dana.F90:113) */
  double skin;       /* nb_dcut (entrada.ini line 13) */
  double eps[9];     /* eps(k,m) at [(k-1)*3 + (m-1)] (dana.F90:86-100) */
  double r0[9];
  double mass[3];
  double h, gamma, temp;            /* h, gama = 1, Tsist = 300 */
  double dif_sc, dif_sei, z_switch; /* diffusion above / below z_switch = 80 */
  double prob, z0, zmax;            /* deposition probability, reservoir */
  int32_t integrator;   /* DML_BROWNIAN (cbrownian_hs) or DML_ERMAK */
  int32_t reservoir;    /* DML_PISTON / DML_CHUNKS / DML_GCMC / none */
  int32_t precision;    /* DML_FP64 or DML_FP32_FORCE */
  int32_t overlap_mode; /* DML_OVERLAP_EXACT (reference order) or ROUNDS */
  int32_t capacity;     /* max slots of this replica */
  int32_t max_neighbors;/* neighbour-row capacity (0 = 512) */
  uint64_t seed;        /* Philox key of this replica */
  double act;           /* GCMC activity (movedor.ini line 3) */
  int32_t nadj;         /* GCMC trials per step */
  int32_t lanes_per_atom; /* force/overlap row lanes (0 = tuned default) */
  int32_t gcmc_mode;    /* DML_GCMC_* bits (0: gcmc_run as in the reference) */
  int32_t reserved;     /* 0 */
} dml_config;

typedef struct dml_counters {
  int64_t try_, depo, choques, choques3, nupd, ins, del, passes;
  int64_t particle_steps; /* sum over completed steps of the atoms alive */
  int64_t row_entries, rows; /* neighbour-row entries and rows built so far */
  int64_t ov_unfinished;     /* round-schedule overlap calls that ended with
                                atoms still overlapping (rounds exhausted) */
} dml_counters;

/* --- lifetime ---------------------------------------------------------- */
int dml_create(const dml_config *cfg, dml_ctx **out);
/* nrep independent replicas in one context; every device phase covers all of
 * them in one launch. */
int dml_create_ensemble(int32_t nrep, const dml_config *cfgs, dml_ctx **out);
int dml_destroy(dml_ctx *ctx);
int dml_sync(dml_ctx *ctx);
const char *dml_strerror(int status);
int32_t dml_num_replicas(const dml_ctx *ctx);
/* total slots of all replicas (sizes of the injected-noise arrays) */
int64_t dml_total_slots(const dml_ctx *ctx);

/* --- state (config_inic, dana.F90:430-517; atom type Groups.F90:251-328) --
 * n slots; NULL arrays take the defaults of atom_allocate / config_inic:
 * vel = acel = 0, pos_old = pos, old_cg = 1e8, flags = ALIVE|REF for
 * non-CG species.  The ref order (hs%ref list order) defaults to ascending
 * slots of REF atoms and the cell-chain order (hs%b index) to the slot. */
int dml_set_state(dml_ctx *ctx, int32_t rep, int32_t n, const double *pos,
                  const double *vel, const double *acel, const double *pos_old,
                  const double *old_cg, const int8_t *species,
                  const uint8_t *flags);
int dml_set_order(dml_ctx *ctx, int32_t rep, int32_t nref,
                  const int32_t *ref_order, const int32_t *bkey);
int dml_get_state(dml_ctx *ctx, int32_t rep, int32_t n, double *pos,
                  double *vel, double *acel, double *force, double *epot,
                  double *pos_old, double *old_cg, int8_t *species,
                  uint8_t *flags);
int dml_get_ref_order(dml_ctx *ctx, int32_t rep, int32_t *nref,
                      int32_t *ref_order);
int dml_set_zmax(dml_ctx *ctx, int32_t rep, double zmax, double z0);
int dml_get_zmax(dml_ctx *ctx, int32_t rep, double *zmax);
int dml_set_box(dml_ctx *ctx, int32_t rep, const double box[3]);

/* --- neighbours -------------------------------------------------------- */
/* test_update (Neighbor.F90:668-713): do_pbc, dispmax skin test, rebuild of
 * cells (Cells.F90:180-302) and lists (Neighbor.F90:465-548) when needed.
 * rebuilt[rep] (may be NULL) receives 1 where the list was rebuilt. */
int dml_test_update(dml_ctx *ctx, int32_t *rebuilt);
/* update (Neighbor.F90:608-633): unconditional rebuild of every replica */
int dml_update(dml_ctx *ctx);
/* cells of the last tessellation; cell_of (n x 3, 1-based with ghost layers,
 * -1 for non-b slots) may be NULL */
int dml_cells(dml_ctx *ctx, int32_t rep, int32_t ncells[3], double cell[3],
              int32_t *use_cells, int32_t *cell_of);
/* neighbour rows in list order: nn[slot], rowptr[slot..slot+1], idx[] slots */
int dml_neighbors(dml_ctx *ctx, int32_t rep, int32_t *nn, int64_t *rowptr,
                  int32_t *idx, int64_t cap, int64_t *total);

/* --- forces and integrators ------------------------------------------- */
/* fuerza (dana.F90:1055-1139) for every replica, in its precision mode */
int dml_forces(dml_ctx *ctx);
/* cbrownian_hs + atom_pbc (dana.F90:798-846, 1187-1250).  noise: 3 normals
 * per ref position [t*3+k]; u_dep: one uniform per ref position (used only on
 * electrode contact).  Ref positions of replica r start at
 * dml_ref_offset(ctx, r).  NULL = device Philox. */
int dml_brownian(dml_ctx *ctx, const double *noise, const double *u_dep);
/* ermak_a + atom_pbc (dana.F90:974-1028): noise [t*6 + 2*k + {0:r1,1:r2}] */
int dml_ermak_a(dml_ctx *ctx, const double *noise, const double *u_dep);
/* ermak_b (dana.F90:1031-1052) */
int dml_ermak_b(dml_ctx *ctx);
int64_t dml_ref_offset(const dml_ctx *ctx, int32_t rep);

/* --- overlap and deposition ------------------------------------------ */
/* overlap_moveback (dana.F90:849-943) in reference order for every replica;
 * u_contact: n_u uniforms per replica (replica r at [r*n_u, (r+1)*n_u)),
 * consumed in order of CG contacts; NULL = Philox. */
int dml_overlap_moveback(dml_ctx *ctx, const double *u_contact, int64_t n_u,
                         int32_t max_passes);
/* deterministic parallel rounds (this build's schedule): round r of slot i
 * draws u_round[r*dml_total_slots + global slot]; NULL = Philox */
int dml_overlap_rounds(dml_ctx *ctx, const double *u_round, int32_t max_rounds);
/* F -> CG and removal from the ref list (dana.F90:228-236) */
int dml_f_to_cg(dml_ctx *ctx);

/* --- piston reservoir and the step counter -----------------------------
 * calc_rho (dana.F90:521-549): count of atoms with z0 < z < zmax (z1 for a
 * chunks replica) per replica, rho[rep] = count / (Lx Ly (top - z0)) (may be
 * NULL); the count stays for dml_maxz and dml_bloques. */
int dml_calc_rho(dml_ctx *ctx, double *rho);

/* The per-frame observables of salida (dana.F90:1141-1183) for every
 * replica, out[4*rep + k]: k = 0 energia, the sum of epot over sys
 * (dana.F90:1150-1158); 1 the kinetic temperature of kion
 * (dana.F90:1342-1376): sum of m |v|^2 over the atoms that are not CG,
 * divided by (j * 3 * kB_ui), j their count (NaN when j = 0, as in the
 * reference); 2 that count j; 3 the atom count of sys.  FP64 sums in a fixed
 * order (per particle block, then per replica); synchronous. */
int dml_observables(dml_ctx *ctx, double *out);
/* maxz (dana.F90:776-794) of the piston replicas, with the count of the last
 * dml_calc_rho */
int dml_maxz(dml_ctx *ctx);
/* end of one main-loop step (dana.F90:262): advances the step number that
 * keys the Philox streams; dml_run does this itself */
int dml_end_step(dml_ctx *ctx);

/* --- chunks reservoir ---------------------------------------------------
 * config_chunk (dana.F90:552-587) for replica rep: n template atoms at the
 * chunk.xyz positions pos (n x 3 SoA; placed at zmax + z as the reference
 * does), dist = the reservoir step (dist + rcut, dana.F90:437-439) and the
 * sensor top z1 (z0 + dist).  Reserves cell memory for the box growth the
 * replica's capacity allows. */
int dml_set_chunk(dml_ctx *ctx, int32_t rep, int32_t n, const double *pos,
                  double dist, double z1);
/* bloques (dana.F90:716-773) of every chunks replica with the count of the
 * last dml_calc_rho: adds a block when the sensor density leaves its band,
 * moves the reservoir up, grows the box and rebuilds that replica's lists */
int dml_bloques(dml_ctx *ctx);

/* --- GCMC reservoir (dana.F90:590-713) ----------------------------------
 * nadj trials of replica rep with uniforms u[t*5 + {0 coin,1 accept,
 * 2..4 coords | 2 pick}] and velocity normals g[t*3 + k]; NULL = Philox. */
int dml_gcmc(dml_ctx *ctx, int32_t rep, double act, int32_t nadj,
             const double *u, const double *g);
/* the gcmc group in list order (slots of rep; the template is order[0], the
 * velocity of an insertion lands on order[ngc-1]: dana.F90:660-668) */
int dml_get_gcmc_order(dml_ctx *ctx, int32_t rep, int32_t *ngc, int32_t *order);
/* the trials of rep's last gcmc call when it ran batched (*batched = 1;
 * 0 when it ran trial by trial, and nothing else is written): kind (0
 * insertion, 1 deletion), accepted (0/1), slot inserted or deleted (-1 when
 * rejected) and the FP64 trial energy (insertion: the WCA sum at the point
 * over the atoms of the call's start, as dml_gcmc_trial_energy; deletion:
 * the removed atom's, at the call's start).  Arrays may be NULL.  No
 * counterpart in the reference (gcmc_run keeps no record). */
int dml_gcmc_outcomes(dml_ctx *ctx, int32_t rep, int32_t ntrial, int32_t *kind,
                      int32_t *accepted, int32_t *slot, double *du, int32_t *batched);
/* per slot: hs%b index (cell-chain order key), particle id and creation
 * number (the order of sys%alist, which salida writes); *nslots is the
 * hs%amax of igroup_attach_atom (Groups.F90:1083-1093).  Arrays may be NULL. */
int dml_get_slots(dml_ctx *ctx, int32_t rep, int32_t *nslots, int32_t *bkey,
                  uint32_t *pid, int32_t *seq);
/* batched trial energies: the WCA pair sum (FP64, fuerza's pair law,
 * dana.F90:1110-1131) of a Li at pos3[t] against every atom within its pair
 * cutoff, and the hard-core indicator of dana.F90:638-653 (a gcmc atom
 * closer than rcut).  exclude[t] (may be NULL; -1 = none) leaves one slot
 * out: the removal energy of atom exclude[t] at its own position.  With
 * pos3 NULL the points are the current positions of the exclude[t] atoms
 * (hard[t] = -1, du[t] = 0 for a slot that holds no atom). */
int dml_gcmc_trial_energy(dml_ctx *ctx, int32_t rep, int32_t ntrial,
                          const double *pos3, const int32_t *exclude, double *du,
                          int32_t *hard);

/* Largest ntrial of dml_gcmc_trial_energy_noalloc. */
#define DML_NOALLOC_MAX_TRIAL 8192
/* dml_gcmc_trial_energy without any device allocation or free: has no
 * counterpart in the reference.  It is for callers that cannot allocate
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernel and results; inputs and outputs are
 * staged in one fixed 320 KiB device array per GPU that the library holds
 * until the process ends.  ntrial > DML_NOALLOC_MAX_TRIAL: DML_E_NOMEM;
 * cell memory still to be grown after dml_set_chunk: DML_E_STATE (any
 * dml_run or dml_gcmc_trial_energy call grows it).  All contexts on one GPU
 * share that array: calls that overlap in time on one GPU race. */
int dml_gcmc_trial_energy_noalloc(dml_ctx *ctx, int32_t rep, int32_t ntrial,
                                  const double *pos3, const int32_t *exclude,
                                  double *du, int32_t *hard);

/* --- whole steps -------------------------------------------------------
 * nsteps device-resident steps of every replica (the order of the main loop,
 * dana.F90:173-263) with Philox noise; use_graph != 0 captures one step in a
 * hipGraph and replays it. */
int dml_run(dml_ctx *ctx, int32_t nsteps, int32_t use_graph);
int dml_get_counters(dml_ctx *ctx, int32_t rep, dml_counters *out);
/* a fatal condition of the reference (werr / stop: out of tessellation,
 * deposit above z0, overlap not resolved) halts only its replica: dml_run
 * continues with the others and this reports the halted replica's status
 * (DML_OK while it runs).  dml_set_state clears it. */
int dml_replica_status(dml_ctx *ctx, int32_t rep, int32_t *status);
/* mean list length and rebuild count of a replica since creation */
int dml_list_stats(dml_ctx *ctx, int32_t rep, double *mean_nn,
                   int64_t *rebuilds, int64_t *steps);

/* --- timing (HIP events on the context's stream) ---------------------- */
int dml_last_run_ms(dml_ctx *ctx, double *ms);
/* average duration of one launch of a phase, reps back-to-back launches;
 * DML_K_LISTBUILD and DML_K_CELLS first rebuild at the current positions. */
enum { DML_K_LISTBUILD = 0, DML_K_FORCE = 1, DML_K_CELLS = 2, DML_K_DISPMAX = 3 };
int dml_time_kernel(dml_ctx *ctx, int32_t which, int32_t reps, double *avg_ms);
/* launch shapes: lanes per ref atom of the force kernel (power of two <= 64)
 * and of the overlap scan's first round (1, 2, 4, 8 or 16; the second and third
 * rounds use at least 2, later ones at least 4),
 * overlap rounds per step of dml_run (default 16, at most 32),
 * list kernel (1 = cell tiles through LDS, 2 = one wave per atom, 3 = 1 or 2
 * per replica by atoms per cell, 4 = one wave per cell, 5 = one wave per
 * cell with its candidates in registers, 3 chunks of 64 a pass, the
 * default, 6 = 5 with 2 chunks a pass at more waves per SIMD); 0
 * keeps the current value.  Results do not depend on them beyond FP64 summation
 * order.  No counterpart in the reference: used by the autotuner. */
int dml_set_tuning(dml_ctx *ctx, int32_t lanes_per_atom, int32_t overlap_rounds,
                   int32_t overlap_lanes, int32_t list_kernel);
int dml_get_tuning(dml_ctx *ctx, int32_t *lanes_per_atom, int32_t *overlap_rounds,
                   int32_t *overlap_lanes, int32_t *list_kernel);

/* --- serving (continuous batching over the replicas of one context) ----
 * No counterpart in the reference.  A replica is a slot that runs one
 * request at a time: dml_reconfigure gives it a new system description
 * (within the capacity and cell memory it was created with; DML_E_NOMEM
 * otherwise) and pauses it; dml_set_state loads the atoms and starts it;
 * dml_set_step_limit pauses it after that many steps of dml_run;
 * dml_set_active pauses / resumes it.  A paused replica costs no kernel
 * work, and a replica's results do not depend on the others. */
int dml_reconfigure(dml_ctx *ctx, int32_t rep, const dml_config *cfg);
int dml_set_step_limit(dml_ctx *ctx, int32_t rep, int64_t limit);
int dml_set_active(dml_ctx *ctx, int32_t rep, int32_t active);
/* per replica, in one call: steps since dml_set_state, status (as
 * dml_replica_status) and whether it is running (arrays of nrep; NULL skips) */
int dml_progress(dml_ctx *ctx, int64_t *steps, int32_t *status, int32_t *running);

#ifdef __cplusplus
}
#endif
#endif
