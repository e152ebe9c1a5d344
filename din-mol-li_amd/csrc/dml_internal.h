// dml_internal.h -- device data layout and shared helpers of libdml.
//
// Layout (HBM):  every replica r owns a slot range [poff, poff + pcap) of the
// global structure-of-arrays particle buffers, pcap a multiple of the
// 256-thread block, so one block never straddles two replicas and a
// block -> replica table (blk2rep) resolves the replica with one scalar load.
// Cells are buckets: replica r owns cells [coff, coff + ccap) (ghost layers
// included) and every cell a fixed run of `bcap` bucket entries holding the
// slots of its atoms in cell-chain order.  Neighbour rows are
// ELL-style, one fixed-capacity row of `maxnn` int32 slots per particle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dml.h"

#define DML_BLOCK 256
#define OVA_PER 8  // items per thread of one k_ov_apply block pass
#define NEAR_MAX 4  // row positions within rcut listed per atom by the row build
#define DML_GCMC_BLOCK 256  // one workgroup per GCMC replica

// species codes (dana.F90:81-84)
#define SP_LI 1
#define SP_CG 2
#define SP_F 3

// error bits collected on the device
#define ERR_TESS 1
#define ERR_LIST 2
#define ERR_DEPZ0 4
#define ERR_NOPART 8
#define ERR_OVERLAP 16
#define ERR_BUCKET 32
#define ERR_UNSUP 64

// largest box edge: the FP32 position mirror used as a distance prefilter is
// given a margin that holds for coordinates up to this size
#define DML_MAX_BOX 4096.0

// RepParams::halted bits
// batched GCMC state of a call (RepParams::gb_state)
#define GB_OFF 0       // sequential trials (k_gcmc_run)
#define GB_ON 1        // the batch kernels commit this call
#define GB_FALLBACK 2  // a batched replica's call the batch cannot take: k_gcmc_run
#define HALT_FATAL 1   // a fatal condition of the reference (err_bits says which)
#define HALT_PAUSED 2  // not running: free serving slot, or step_limit reached
#define HALT_MASK 4    // held out of one phase (dml_run's first rebuild)

struct RepParams {
  double box[3], one_box[3];
  double rcut, rcut2, rl2, skin;
  float pf_margin;   // band (A^2) around a squared cutoff decided in FP64
  double zmax, z0, prob, h, z_switch, dif_sc, dif_sei;
  double cc0, cc1, cc2, sdr, sdv, crv1, crv2, skt;
  double eps[9], r0[9], mass[3];
  double cell[3];
  double rho0, lohi;
  double act;
  int32_t ncells[3];
  int32_t use_cells;
  int32_t poff, pcap;
  int32_t coff, ccap;
  int64_t boff;   // first bucket entry of this replica
  int32_t bcap;   // bucket entries per cell
  int32_t nslots; // slots in use (amax)
  int32_t nref;
  int32_t listed;
  int32_t integrator, reservoir, piston, precision, overlap_mode;
  int32_t step;
  int32_t nadj;
  int32_t rebuild;      // set by the skin test of the current test_update
  int32_t any_ghost;    // some ref row was built from a ghost-layer cell
  int32_t ov_pending;   // overlap rounds not finished
  int32_t ov_round;     // rounds done in the current overlap call
  int32_t any_f;        // F atoms to convert / ref list to compact
  int32_t dispmax_done; // blocks finished in the dispmax reduction
  int32_t rho_done;     // blocks finished in calc_rho
  int32_t rho_count;
  int32_t max_nn_seen;
  int32_t nref_new;
  int32_t halted;    // HALT_* bits: any set bit keeps every kernel off this replica
  int32_t step_limit; // > 0: pause (HALT_PAUSED) once step reaches it
  int32_t err_bits;  // which one (ERR_*)
  // group bookkeeping of the GCMC reservoir (Groups.F90 igroup semantics)
  int32_t nat, nlimbo;     // hs members alive / in limbo
  int32_t b_amax, b_nat;   // hs%b index range and members
  int32_t ngc;             // members of the gcmc group (list order in gc_order)
  // one gcmc_run call: lengths of the working gcmc list (wl[1]), of the
  // in-volume list (ref_tmp) and of the ref list with deleted entries, and
  // the stage the grid-wide list kernels reached (1 staged, 2 trials done,
  // 3 counted; reset at the next call)
  int32_t gc_len, gc_inv, ref_len, gc_pending;
  // step + 1 of the fuerza call or row build whose record of the entries
  // within rcut (ov_near counts, nearl lists) describes the current
  // positions and rows (0: none; cleared by every kernel that moves or adds
  // atoms, and by a rebuild whose row kernel keeps no record)
  int32_t ovn_stamp;
  // search hints: every slot below free_lo is alive or in limbo, every hs%b
  // index below bfree_lo is taken (lowered wherever one is freed)
  int32_t free_lo, bfree_lo;
  uint32_t pid_next;       // next particle id (Philox counter)
  int32_t seq_next;        // next creation number (sys list order)
  int32_t gcmc_calls;
  uint64_t seed;
  double d1, d2;        // top-2 squared displacements of the last skin test
  // chunks reservoir (dana.F90:437-441, 552-587, 716-773)
  double z1, dist;         // sensor top, block step (dist + rcut)
  int64_t chunk_off;       // first template atom in the chunk pool
  int32_t nchunk;          // template atoms
  int32_t chunk_fired;     // bloques added a block in this step
  int64_t creg, breg;      // reserved cell counts / bucket entries (grow room)
  int32_t lb_kind;         // list kernel of this replica: 0 cell tiles, 1 wave per atom
  // batched GCMC (DML_GCMC_BATCHED): trial pool range, this call's state
  // (GB_OFF / GB_ON / GB_FALLBACK) and the group sizes at its start
  int32_t gcmc_mode;
  int32_t gb_off, gb_cap;
  int32_t gb_state, gb_t;
  int32_t gb_ilen, gb_glen, gb_rlen, gb_nslots, gb_bamax;
  int32_t gb_nF, gb_nFB;   // free slots below nslots / free hs%b indices below b_amax
  int32_t gb_nins, gb_ndel, gb_nalist;
  uint32_t gb_stamp;       // gcmc_calls + 1 of this call (stamped per-cell / per-slot marks)
  double gb_act;
};

// cgroup_tessellate (Cells.F90:180-265) on the replica's box: the fast path
// keeps ncells when the box barely changed; all(ncells < 4) selects the
// O(N^2) Verlet scan.  Returns 1 for the reference's stale-cell quirk (the
// fast path declines while cells were in use and the new grid is all < 4).
__host__ __device__ inline int tessellate_rep(RepParams &p, int first) {
  double brcut = p.rcut + p.skin;
  if (!first && p.use_cells) {
    int ok = 1, ok2 = 1;
    for (int k = 0; k < 3; k++)
      if (!((double)(p.ncells[k] + 1) >= p.box[k] / brcut)) ok = 0;
    for (int k = 0; k < 3; k++)
      if (!(brcut < p.box[k] / (double)p.ncells[k])) ok2 = 0;
    if (ok && ok2) {
      for (int k = 0; k < 3; k++) p.cell[k] = p.box[k] / (double)p.ncells[k];
      return 0;
    }
  }
  int nc[3];
  for (int k = 0; k < 3; k++) nc[k] = (int)(p.box[k] / brcut);
  if (nc[0] < 4 && nc[1] < 4 && nc[2] < 4) {
    if (!first && p.use_cells) return 1;
    for (int k = 0; k < 3; k++) {
      p.ncells[k] = nc[k];
      p.cell[k] = p.box[k];
    }
    p.use_cells = 0;
    return 0;
  }
  for (int k = 0; k < 3; k++) {
    p.ncells[k] = nc[k];
    p.cell[k] = p.box[k] / (double)nc[k];
  }
  p.use_cells = 1;
  return 0;
}

struct DevCounters {
  unsigned long long try_, depo, choques, choques3, nupd, ins, del, passes;
  unsigned long long nn_sum, nn_rows, steps, ov_unfinished;
  unsigned long long particle_steps;  // sum over steps of the atoms alive
};

// ------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
// easy as 1, 2, 3", SC'11), counter-based.
// ------------------------------------------------------------------------
struct u32x4 {
  uint32_t x, y, z, w;
};

__host__ __device__ inline void philox_mulhilo(uint32_t a, uint32_t b,
                                               uint32_t *hi, uint32_t *lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0,
                                               uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t hi0, lo0, hi1, lo1;
    philox_mulhilo(M0, c.x, &hi0, &lo0);
    philox_mulhilo(M1, c.z, &hi1, &lo1);
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// uniform double in (0,1) from two words: 53 bits, centred in their ulp
__host__ __device__ inline double u01_from(uint32_t a, uint32_t b) {
  uint64_t m = ((((uint64_t)a) << 32) | b) >> 11;
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// counter layout: (particle id, step, site << 24 | draw, replica seed hi)
enum { RS_BROWN = 1, RS_ERMAK = 2, RS_DEPO = 3, RS_CONTACT = 4, RS_GCMC = 5 };

__host__ __device__ inline u32x4 dml_philox(uint64_t seed, uint32_t pid,
                                            uint32_t step, uint32_t site,
                                            uint32_t draw) {
  u32x4 c;
  c.x = pid;
  c.y = step;
  c.z = (site << 24) | (draw & 0xFFFFFFu);
  c.w = 0x6d6c6921u;  // fixed word: keeps streams of this library distinct
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Box-Muller pair from four Philox words
__device__ inline void gauss2(u32x4 w, double *g0, double *g1) {
  double u1 = u01_from(w.x, w.y);
  double u2 = u01_from(w.z, w.w);
  double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  *g0 = r * c;
  *g1 = r * s;
}

// ------------------------------------------------------------------------
// geometry exactly as the reference evaluates it
// ------------------------------------------------------------------------
__host__ __device__ inline double dot3(double a, double b, double c) {
  return (a * a + b * b) + c * c;
}

// vdistance with mic, pbc = (T,T,F) (Groups.F90:995-1016): round half away
// from zero like idnint
__device__ inline void vdist_mic(const RepParams &p, double ax, double ay,
                                 double az, double bx, double by, double bz,
                                 double *dx, double *dy, double *dz) {
  double vx = ax - bx, vy = ay - by, vz = az - bz;
  vx = vx - p.box[0] * round(vx * p.one_box[0]);
  vy = vy - p.box[1] * round(vy * p.one_box[1]);
  *dx = vx;
  *dy = vy;
  *dz = vz;
}

// the 27-cell stencil in the reference's order (Cells.F90:28-36)
__constant__ static const int8_t kMap[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {1, 1, 0},   {0, 1, 0},  {-1, 1, 0},
    {1, 0, -1},  {1, 1, -1},  {0, 1, -1},  {-1, 1, -1}, {1, 0, 1},
    {1, 1, 1},   {0, 1, 1},   {-1, 1, 1},  {0, 0, 1},  {-1, 0, 0},
    {-1, -1, 0}, {0, -1, 0},  {1, -1, 0},  {-1, 0, 1}, {-1, -1, 1},
    {0, -1, 1},  {1, -1, 1},  {-1, 0, -1}, {-1, -1, -1}, {0, -1, -1},
    {1, -1, -1}, {0, 0, -1}};

// cell_pbc with mic (Cells.F90:387-391), C '%' == Fortran mod here
__device__ inline int cell_wrap(int r, int n) { return ((r - 1 + n) % n) + 1; }
// the same for r in [-1, n + 2] (a stencil step from any reserved cell,
// ghost layer included) without an integer division
__device__ inline int cell_wrap1(int r, int n) {
  const int w = r + (r < 1 ? n : 0);
  return n == 1 ? 1 : w - (w > n ? n : 0);
}

__device__ inline int cell_lin(const RepParams &p, int i, int j, int k) {
  return i + (p.ncells[0] + 2) * (j + (p.ncells[1] + 2) * k);
}

struct DevState {
  // particle SoA
  double *px, *py, *pz, *vx, *vy, *vz, *ax, *ay, *az, *fx, *fy, *fz, *epot;
  double *ox, *oy, *oz;   // pos_old
  double *cx, *cy, *cz;   // old_cg
  double *rvx, *rvy, *rvz; // ranv of ermak_a
  int8_t *species;
  uint8_t *flags;
  uint8_t *rowflag;       // bit0: row built from a ghost-layer cell
  int32_t *bkey;          // hs%b slot: cell-chain order key
  uint32_t *pid;          // stable particle id (Philox counter)
  int32_t *ref_order, *ref_tmp, *ref_pos;
  int32_t *nn, *nbr;
  int32_t maxnn;
  // cells
  int32_t *cell_cnt;      // per cell
  int32_t *bucket;        // slots, bucket-ordered
  double *bx, *by, *bz;   // positions, bucket-ordered
  // overlap rounds
  int32_t *ov_rev;        // round stamp of a revert
  uint8_t *ov_st;         // bit0 active bit1 dominated bit2 deposit bit3 self-revert
  uint8_t *ov_near;       // row entries within rcut (FP32, with margin) listed in
                          // nearl by fuerza or the row build (255: not listed)
  int32_t *nearl;         // NEAR_MAX row positions per atom
  int32_t *wl[2];         // overlap worklists (global ref positions), by round parity
  // overlap rounds after the first visit only their work lists: ovw[q % 2]
  // the ref positions still unresolved (skip = 0) for round q, ovr[q % 2]
  // the atoms reverted in round q; wl_cnt = [|ovw0|, |ovw1|, |ovr0|, |ovr1|]
  int32_t *wl_cnt;
  int32_t *ovw[2], *ovr[2];
  int32_t *ov_more;       // [3][nrep]: the replica reverted or deferred in a round
  int32_t *gc_order;      // gcmc group in list order (global slots)
  int32_t *b2slot;        // hs%b index -> global slot, -1 when free
  int32_t *seq;           // creation number: the order of sys%alist
  double *chx, *chy, *chz;   // chunk templates (pool over the replicas)
  double *chox, *choy, *choz;
  // flat work lists (rebuilt by the host when the layout changes)
  int32_t *cm_rep, *cm_cell; // every reserved cell of every replica
  int32_t *tm_rep, *tm_cell; // cells of the replicas built by cell tiles
  int32_t *wm_blk;           // particle blocks of the replicas built wave per atom
  int64_t n_cm, n_tm, n_wm;
  // FP32 mirror of the positions, w = species bits: conservative prefilters
  float4 *pf4;            // per slot
  float4 *bf4;            // bucket-ordered
  // reductions
  double *part_d;         // 2 per particle block
  double *obs;            // 4 per particle block, then 4 per replica
  int32_t *part_i;        // 2 per particle block
  // tables
  int32_t *blk2rep;
  RepParams *rp;
  DevCounters *cnt;
  int32_t *err;
  // batched GCMC.  Per trial (pool of every batched replica, gb_off + t):
  // point, (u1, u2), energy, kind (0 insertion, 1 deletion, 2 point outside
  // the cells), cell, hard-core neighbours (in-volume list positions, -1 in
  // nblk: a neighbour that cannot leave), earlier insertion points within
  // rcut, decision (insertion rank / in-volume list position, -1 rejected)
  // and the slot inserted or deleted.
  double *gb_pt, *gb_u, *gb_du;
  double *gb_g;           // an insertion's three velocity normals (3 per trial)
  int32_t *gb_kind, *gb_cell, *gb_nblk, *gb_blk, *gb_ncf, *gb_cf, *gb_res, *gb_slot;
  // per accepted insertion (rank k): trial, next in its cell's list, trial
  // of its deletion (INT_MAX: none), template
  int32_t *gi_trial, *gi_next, *gi_del, *gi_tmpl;
  // per accepted deletion (index d): trial, in-volume position, slot, and
  // for atoms present at the call's start: hs%b index and its rank, ref
  // list position, flags, cell
  int32_t *gd_trial, *gd_vpos, *gd_slot, *gd_bso, *gd_brank, *gd_rpos, *gd_oflags, *gd_cell;
  int32_t *gb_free, *gb_freeb, *gb_frank;  // first free slots / hs%b indices, ranks of the latter
  int32_t *gb_U;          // 3 per trial: the free hs%b indices in ascending order
  int32_t *gb_part;       // 2 per particle block
  int32_t *gb_vpos, *gb_delt, *gb_alist;  // per slot
  unsigned long long *gb_app;             // per slot: stamped counts
  unsigned long long *gb_chead, *gb_ccnt; // per cell: stamped list heads and counts
  int32_t *gb_claim;                      // per cell: stamp of the call that rebuilt it
  int32_t nblocks;        // particle blocks
  int32_t nrep;
};
