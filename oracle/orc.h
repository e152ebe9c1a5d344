/*
 * orc.h -- CPU restatement of the din-mol-Li per-step hot path.
 *
 * THIS IS TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg may link or call it, and only
 * as the checker (or the timed CPU baseline).  The product path lives in
 * din-mol-li_amd/csrc and never calls into this library.
 *
 * The restatement follows the reference Fortran routine by routine (file:line
 * citations beside each function in orc.c).  Every floating-point expression
 * keeps the reference's evaluation order, its integer powers use the
 * square-and-multiply chains the -O0 reference build evaluates
 * (x^6 = x^2*x^4, x^7 = (x*x^2)*x^4), and the Numerical-Recipes RNG and the
 * single-precision polar Gaussian are reproduced bit for bit, so that a whole
 * run reproduces tests/{ermak,brown,gcmc}/ref.xyz exactly.
 *
 * Pinning (what checks each part):
 *   whole trajectory (RNG, pos_inic, integrators, fuerza, overlap, GCMC,
 *     piston, chunks, cells, lists)   -> the three reference golden frames
 *   cells / neighbour sets            -> the reference modules themselves
 *     (oracle/_ref/nbdump, built from /root/reference/src), brute force
 *   fuerza                            -> ermak golden, closed-form WCA pair
 *   Ermak constants                   -> closed form (Allen & Tildesley)
 *   parallel overlap variant          -> parity unpinned against the
 *     reference (it is this build's own deterministic schedule, not a
 *     reference algorithm); pinned by invariants in the tests.
 *
 * Index space: every per-particle array is indexed by the "hs slot" (the
 * index the reference's ngroup `hs` gives an atom, 0-based here); neighbour
 * lists store hs slots, exactly as Neighbor.F90 stores `aj%gid(g)`.
 */
#ifndef DML_ORC_H
#define DML_ORC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LI = 1, ORC_CG = 2, ORC_F = 3 };
enum { ORC_NULL = 0, ORC_ALIVE = 1, ORC_LIMBO = 2 };

/* error codes (mirror the werr/stop sites of the reference) */
enum {
  ORC_OK = 0,
  ORC_E_OUT_OF_TESS = -1,   /* Cells.F90:293-294 */
  ORC_E_DEPOSIT_Z0 = -2,    /* dana.F90:905-908 */
  ORC_E_NO_PARTICLE = -3,   /* dana.F90:701 */
  ORC_E_PLACEMENT = -4,     /* dana.F90:386-389 */
  ORC_E_RESERVOIR = -5,     /* dana.F90:423 */
  ORC_E_OVERLAP_LOOP = -6,  /* overlap recursion did not end (no ref check) */
  ORC_E_ALLOC = -7,
  ORC_E_INDEX = -8          /* Groups.F90:1088 index inconsistency */
};

/* where a random draw is consumed; used by array-backed draw sources */
enum {
  ORC_SITE_BROWN = 0,   /* i = ref position, k = component 0..2        */
  ORC_SITE_ERMAK = 1,   /* i = ref position, k = 2*comp + {0:r1, 1:r2}  */
  ORC_SITE_DEPO = 2,    /* i = ref position (electrode contact uniform) */
  ORC_SITE_CONTACT = 3, /* i = contact event counter (overlap CG hit)   */
  ORC_SITE_GCMC_U = 4,  /* i = trial, k = 0 coin, 1 accept, 2..4 coords / 2 pick */
  ORC_SITE_GCMC_G = 5   /* i = trial, k = 0..2 velocity gaussians       */
};

typedef struct orc_rng orc_rng;
struct orc_rng {
  double (*unif)(orc_rng *, int site, int64_t i, int k);
  double (*gauss)(orc_rng *, int site, int64_t i, int k);
  void *user;
  /* Numerical-Recipes stream state (dana.F90:1379-1428) */
  int32_t idum, ix, iy;
  int gaus_stored;
  double gsave;
  int64_t ncalls;
  /* array-backed draws: arrays indexed by site */
  const double *arr[6];
  int64_t arr_n[6];
  int arr_stride[6];
  int64_t contact_next; /* running counter for ORC_SITE_CONTACT */
  int err;              /* set when an array draw runs out of range */
};

void orc_rng_init_nr(orc_rng *r, int32_t idum);
double orc_nr_ran(orc_rng *r);
double orc_nr_gasdev(orc_rng *r);
void orc_rng_init_arrays(orc_rng *r);

typedef struct orc_sys orc_sys;

typedef struct orc_params {
  double box[3];
  double rcut;      /* hs%rcut (3.2 in dana) */
  double skin;      /* nb_dcut */
  double eps[9];    /* eps(k,m) stored [(k-1)*3 + (m-1)] */
  double r0[9];
  double mass[3];
  double h, gama, temp;
  double dif_sc, dif_sei, z_switch;
  double prob, z0, zmax;
  int ermak;        /* integrador: 1 Ermak, 0 Brownian */
  int piston;       /* s_piston (enables the unsolvable-overlap check) */
} orc_params;

/* counters of one run / routine */
typedef struct orc_counters {
  int64_t try_, depo, choques, choques2, choques3, nupd, ins, del;
  int64_t passes;
  int64_t ov_unfinished; /* round-schedule calls that ran out of rounds */
} orc_counters;

/* ---- system construction and access (hs-slot indexed) ---- */
orc_sys *orc_sys_new(const orc_params *p, int cap);
void orc_sys_free(orc_sys *s);
/* set the ermak constants of set_ermak (dana.F90:947-971) from p */
void orc_set_ermak(orc_sys *s);
void orc_get_ermak(const orc_sys *s, double out[8]);
/* attach an atom the way config_inic does (dana.F90:457-489); returns hs slot */
int orc_attach_initial(orc_sys *s, int z, const double pos[3]);
int orc_nslots(const orc_sys *s);
int orc_nref(const orc_sys *s);
/* state export/import, SoA column-major (n x 3), n = orc_nslots */
void orc_get_state(const orc_sys *s, double *pos, double *vel, double *acel,
                   double *force, double *epot, double *pos_old, double *old_cg,
                   int8_t *species, uint8_t *flags);
void orc_set_state(orc_sys *s, const double *pos, const double *vel,
                   const double *acel, const double *pos_old,
                   const double *old_cg, const int8_t *species,
                   const uint8_t *flags);
/* ref order as hs slots */
void orc_get_ref_order(const orc_sys *s, int32_t *out);
void orc_set_zmax(orc_sys *s, double zmax);
void orc_set_box(orc_sys *s, const double box[3]);
int orc_get_b_slot(const orc_sys *s, int hs);

/* flags bits (shared with the device ABI) */
#define ORC_FL_ALIVE 1u
#define ORC_FL_LIMBO 2u
#define ORC_FL_REF 4u
#define ORC_FL_SKIP 8u
#define ORC_FL_GCMC 16u

/* ---- per-routine entry points ---- */
/* do_pbc + tessellate + sort + dispmax test + (update) : Neighbor.F90:668-713 */
int orc_test_update(orc_sys *s, int *rebuilt, orc_counters *c);
/* forced rebuild (update(), Neighbor.F90:608-633) */
int orc_update(orc_sys *s, orc_counters *c);
/* cells of the last tessellation */
void orc_cells(const orc_sys *s, int32_t ncells[3], double cell[3],
               int *use_cells);
/* cell index (1-based, with ghost layers) of each hs slot; -1 for non-b */
void orc_cell_of(const orc_sys *s, int32_t *ci3);
/* neighbour list export: nn per hs slot, entries (hs slots) row by row */
int64_t orc_list_total(const orc_sys *s);
void orc_list_export(const orc_sys *s, int32_t *nn, int64_t *rowptr,
                     int32_t *idx);
/* fuerza (dana.F90:1055-1139) */
void orc_fuerza(orc_sys *s);
/* cbrownian_hs (dana.F90:798-846) incl. atom_pbc */
int orc_brownian(orc_sys *s, orc_rng *r, orc_counters *c);
/* ermak_a / ermak_b (dana.F90:974-1052); ranv is held inside s */
int orc_ermak_a(orc_sys *s, orc_rng *r, orc_counters *c);
void orc_ermak_b(orc_sys *s);
/* overlap_moveback (dana.F90:849-943), exact sequential semantics */
int orc_overlap(orc_sys *s, orc_rng *r, orc_counters *c, int max_passes);
/* this build's deterministic parallel-round overlap schedule (not a
 * reference routine; see orc.c) with counter-keyed uniforms u_key */
int orc_overlap_rounds(orc_sys *s, const double *u_round, int max_rounds,
                       orc_counters *c);
/* F -> CG conversion at step end (dana.F90:228-236) */
void orc_f_to_cg(orc_sys *s, int gcmc);
/* gcmc_run (dana.F90:590-713) */
int orc_gcmc(orc_sys *s, orc_rng *r, double act, int nadj, orc_counters *c);
/* chunks reservoir: the block template of config_chunk and bloques */
typedef struct orc_chunk {
  int n;
  double dist;            /* dist + rcut (config_inic, dana.F90:437-441) */
  double (*pos)[3], (*old)[3];
} orc_chunk;
/* config_chunk (dana.F90:552-587) from chunk.xyz positions pos3[3*j+k] */
int orc_chunk_init(const orc_sys *s, orc_chunk *ch, int n, const double *pos3, double dist);
void orc_chunk_free(orc_chunk *ch);
/* bloques (dana.F90:716-773) with the rho of calc_rho; *fired = 1 when the
 * block was added (then the box grew and test_update ran) */
int orc_bloques(orc_sys *s, orc_chunk *ch, double rho, int *fired, orc_counters *c);
/* calc_rho (dana.F90:521-549): chunks = 1 counts up to z1 */
double orc_calc_rho(orc_sys *s, int chunks);
/* maxz (dana.F90:776-794) with this step's rho and the initial rho0 */
void orc_maxz(orc_sys *s, double rho, double rho0);
/* salida's observables (dana.F90:1141-1183) in sys order: out = energia,
 * kion temperature (dana.F90:1342-1376), j (non-CG count), sys%nat */
void orc_observables(const orc_sys *s, double out[4]);
void orc_set_z1(orc_sys *s, double z1);
/* z0, z1, zmax, box(3) */
void orc_get_reservoir(const orc_sys *s, double out[4]);

/* gcmc group of config_inic: Li atoms in sys order (dana.F90:505-513) */
void orc_gcmc_init(orc_sys *s);
/* gcmc group in list order; returns its size */
int orc_gcmc_order(const orc_sys *s, int32_t *out);
/* GCMC insertion trial energy: WCA sum over neighbours within 27 cells and
 * the hard-core indicator of dana.F90:638-653 (over the gcmc group). */
void orc_gcmc_trial_energy(const orc_sys *s, int ntrial, const double *pos3,
                           const int32_t *exclude, double *du, int32_t *hard);
/* ranv export (ermak) */
void orc_get_ranv(const orc_sys *s, double *ranv);

/* ---- whole dana run (dana.F90:1-305) ---- */
typedef struct orc_run_cfg {
  int32_t idum;
  double prob, h;
  int32_t nst, nwr;
  double xi, yi, dist, z0, zmax, dif_sc, dif_sei, nb_dcut;
  int32_t ermak;      /* movedor.ini line 1 */
  int32_t reservoir;  /* 0 piston, 1 chunks, 2 gcmc, 3 none (z0 = zmax) */
  double act;
  int32_t nadj;
  int32_t nchunk;
  const double *chunk_pos; /* nchunk x 3, row-major as read from chunk.xyz */
  /* optional: start from these positions instead of pos_inic */
  int32_t ninit;
  const double *init_pos;  /* ninit x 3 row-major */
  const int8_t *init_z;    /* species, NULL = all Li */
  /* 0: overlap_moveback as the reference (dana.F90:849-943); n > 0: the
     round schedule (orc_overlap_rounds) with at most n rounds per call and
     its uniforms from a stream of its own, a call that runs out of rounds
     counted and the run continued (statistics of the schedule only) */
  int32_t overlap_rounds;
  uint64_t rounds_seed;
} orc_run_cfg;

typedef struct orc_run_out {
  int32_t nat;       /* final sys%nat */
  double zmax;       /* final zmax */
  double *pos;       /* caller-provided, cap x 3 row-major, sys order */
  int8_t *z;         /* caller-provided, cap */
  int32_t cap;
  orc_counters c;
  double loop_seconds;
  int64_t particle_steps;
} orc_run_out;

int orc_dana_run(const orc_run_cfg *cfg, orc_run_out *out);

#ifdef __cplusplus
}
#endif
#endif
