// dml_kernels.hip -- HIP kernels (gfx950, wave64) of the din-mol-Li hot path.
//
// Each kernel cites the reference routine whose semantics it carries.  All
// kernels cover every replica of a context in one launch: particle kernels
// run one thread per slot (or per ref position) and read the replica from
// the block -> replica table; cell kernels run a 2-D grid (cells x replica).
#include "dml_kernels.h"

#define WAVE 64

__device__ inline int same_pos(const DevState &s, int i) {
  return s.px[i] == s.cx[i] && s.py[i] == s.cy[i] && s.pz[i] == s.cz[i];
}

// FP32 mirror of a slot's position and species (prefilters only; every
// decision near a threshold is re-taken in FP64)
__device__ inline void put_pf4(const DevState &s, int i, double x, double y, double z, int sp) {
  s.pf4[i] = make_float4((float)x, (float)y, (float)z, __int_as_float(sp));
}
__device__ inline void put_pf4_sp(const DevState &s, int i, int sp) {
  reinterpret_cast<float *>(&s.pf4[i])[3] = __int_as_float(sp);
}
// the bucket-ordered FP32 mirror of slot i: position, species (bits 0-7 of
// w) and the slot's flags at the time of writing (bits 8-15), so the row
// build reads whether a bucket entry is a ref atom with its position
__device__ inline float4 bucket_f4(const DevState &s, int i) {
  float4 f = s.pf4[i];
  f.w = __int_as_float((__float_as_int(f.w) & 0xff) | (int)s.flags[i] << 8);
  return f;
}
// squared minimum-image distance in FP32.  A pair is decided in FP32 only
// outside the band cutoff^2 +- RepParams::pf_margin (set by the host from
// the largest box and the list radius, see pf_margin_for); inside it the FP64
// test of the reference decides.  The image chosen near half a box may
// differ from idnint's, which changes d2 only by O(box * 2^-24).
__device__ inline float d2_f32(const RepParams &p, float4 a, float4 b, float bx, float by,
                               float ibx, float iby) {
  float vx = a.x - b.x, vy = a.y - b.y, vz = a.z - b.z;
  vx = vx - bx * rintf(vx * ibx);
  vy = vy - by * rintf(vy * iby);
  return (vx * vx + vy * vy) + vz * vz;
}

// a fatal condition of the reference (werr/stop) ends this replica only: its
// kernels skip it from now on and dml_replica_status() reports why
__device__ inline void rep_fail(const DevState &s, int r, int bit) {
  atomicOr(&s.rp[r].err_bits, bit);
  atomicOr(&s.rp[r].halted, HALT_FATAL);
}

__device__ inline const RepParams &rep_of_block(const DevState &s, int blk,
                                                int *r) {
  *r = s.blk2rep[blk];
  return s.rp[*r];
}

__device__ inline double atomic_load_d(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void top2_merge(double &a1, double &a2, double b1, double b2) {
  double n1 = fmax(a1, b1);
  double n2 = fmax(fmin(a1, b1), fmax(a2, b2));
  a1 = n1;
  a2 = n2;
}

__device__ inline void block_top2(double &a1, double &a2, double *sh) {
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    double b1 = __shfl_xor(a1, off), b2 = __shfl_xor(a2, off);
    top2_merge(a1, a2, b1, b2);
  }
  int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  if (lane == 0) {
    sh[2 * w] = a1;
    sh[2 * w + 1] = a2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x / WAVE); q++) top2_merge(a1, a2, sh[2 * q], sh[2 * q + 1]);
  }
  __syncthreads();
}

// ------------------------------------------------------------------------
// test_update part 1: do_pbc (Groups.F90:1440-1467) and inq_dispmax
// (Neighbor.F90:635-666) with the rebuild decision of Neighbor.F90:696-710.
// The last block of a replica merges the per-block top-2 values.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(DML_BLOCK) k_pbc_dispmax(DevState s) {
  __shared__ double sh[2 * DML_BLOCK / WAVE];
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  int slot = blockIdx.x * DML_BLOCK + threadIdx.x;
  int local = slot - P.poff;
  double d = -1.0;
  if (local < P.nslots && (s.flags[slot] & DML_FL_ALIVE)) {
    double x = s.px[slot], y = s.py[slot];
    double bx = P.box[0], by = P.box[1];
    int moved = 0;
    if (x >= bx) {
      x -= bx; s.px[slot] = x; s.ox[slot] -= bx; moved = 1;
    } else if (x < 0.0) {
      x += bx; s.px[slot] = x; s.ox[slot] += bx; moved = 1;
    }
    if (y >= by) {
      y -= by; s.py[slot] = y; s.oy[slot] -= by; moved = 1;
    } else if (y < 0.0) {
      y += by; s.py[slot] = y; s.oy[slot] += by; moved = 1;
    }
    double z = s.pz[slot];
    if (moved) put_pf4(s, slot, x, y, z, s.species[slot]);
    d = dot3(x - s.ox[slot], y - s.oy[slot], z - s.oz[slot]);
  }
  double a1 = d, a2 = -1.0;
  block_top2(a1, a2, sh);
  if (threadIdx.x == 0) {
    s.part_d[2 * blockIdx.x] = a1;
    s.part_d[2 * blockIdx.x + 1] = a2;
  }
}

// one block per replica merges the block partials (a kernel boundary orders
// them; no device-scope fences, which on this part write back per-XCD L2)
__global__ void __launch_bounds__(DML_BLOCK) k_dispmax_final(DevState s) {
  __shared__ double sh[2 * DML_BLOCK / WAVE];
  const int r = blockIdx.x;
  RepParams *Pw = s.rp + r;
  if (Pw->halted) return;
  const int b0 = Pw->poff / DML_BLOCK, nb = Pw->pcap / DML_BLOCK;
  double a1 = -1.0, a2 = -1.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x)
    top2_merge(a1, a2, s.part_d[2 * (b0 + q)], s.part_d[2 * (b0 + q) + 1]);
  block_top2(a1, a2, sh);
  if (threadIdx.x == 0) {
    top2_merge(a1, a2, 1e-16, 1e-16);  // initial values of inq_dispmax
    Pw->d1 = a1;
    Pw->d2 = a2;
    int reb = (!Pw->listed) || (sqrt(a1) + sqrt(a2) > Pw->skin);
    Pw->rebuild = reb;
    if (reb) {
      Pw->listed = 1;
      Pw->any_ghost = 0;
      Pw->nlimbo = 0;  // the rebuild cleans the limbo (Groups.F90:1036-1053)
      // new rows: the cell-tile and wave-per-cell row builds list their
      // entries within rcut
      Pw->ovn_stamp = Pw->lb_kind != 1 ? Pw->step + 1 : 0;
      atomicAdd(&s.cnt[r].nupd, 1ull);
    }
  }
}

__global__ void k_force_rebuild(DevState s) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= s.nrep) return;
  RepParams *Pw = s.rp + r;
  if (Pw->halted) return;
  Pw->rebuild = 1;
  Pw->listed = 1;
  Pw->any_ghost = 0;
  Pw->nlimbo = 0;
  Pw->ovn_stamp = Pw->lb_kind != 1 ? Pw->step + 1 : 0;
  atomicAdd(&s.cnt[r].nupd, 1ull);
}

// ------------------------------------------------------------------------
// cells: cgroup_tessellate/sort (Cells.F90:180-302) as buckets
// ------------------------------------------------------------------------
__device__ inline int ncell_total(const RepParams &P) {
  return P.use_cells ? (P.ncells[0] + 2) * (P.ncells[1] + 2) * (P.ncells[2] + 2) : 1;
}

// resort: 1 = cells of this replica are rebuilt in this test_update: on a
// list rebuild, and for a GCMC replica whose trials come before the next
// test_update (gcs = 1; its trials read the cells at the current positions)
__device__ inline int want_cells(const RepParams &P, int gcs) {
  return P.rebuild || (gcs && P.reservoir == DML_GCMC);
}

__global__ void k_zero_cells(DevState s, int gcs) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < s.n_cm;
       g += (int64_t)gridDim.x * blockDim.x) {
    const RepParams &P = s.rp[s.cm_rep[g]];
    const int c = s.cm_cell[g];
    if (P.halted || !want_cells(P, gcs) || c >= ncell_total(P)) continue;
    s.cell_cnt[P.coff + c] = 0;
  }
}

__global__ void __launch_bounds__(DML_BLOCK) k_cell_key(DevState s, int gcs) {
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  if (!want_cells(P, gcs)) return;
  int slot = blockIdx.x * DML_BLOCK + threadIdx.x;
  int local = slot - P.poff;
  if (local >= P.nslots) return;
  uint8_t f = s.flags[slot];
  if (P.rebuild && (f & DML_FL_LIMBO)) {  // igroup_clean (Groups.F90:1036-1053)
    s.flags[slot] = f & (uint8_t)~DML_FL_LIMBO;
    atomicMin(&s.rp[r].free_lo, local);
    return;
  }
  if (!(f & DML_FL_ALIVE)) return;
  double x = s.px[slot], y = s.py[slot], z = s.pz[slot];
  if (P.rebuild) {  // update(): pos_old = pos (Neighbor.F90:620-624)
    s.ox[slot] = x;
    s.oy[slot] = y;
    s.oz[slot] = z;
  }
  int cell = 0;
  if (P.use_cells) {
    int c0 = (int)(x / P.cell[0]) + 1;
    int c1 = (int)(y / P.cell[1]) + 1;
    int c2 = (int)(z / P.cell[2]) + 1;
    if (c0 < 0 || c1 < 0 || c2 < 0 || c0 > P.ncells[0] + 1 || c1 > P.ncells[1] + 1 ||
        c2 > P.ncells[2] + 1) {
      rep_fail(s, r, ERR_TESS);
      return;
    }
    cell = cell_lin(P, c0, c1, c2);
  }
  int rank = atomicAdd(&s.cell_cnt[P.coff + cell], 1);
  if (rank >= P.bcap) {
    atomicOr(s.err, ERR_BUCKET);
    rep_fail(s, r, ERR_BUCKET);
    return;
  }
  s.bucket[P.boff + (int64_t)cell * P.bcap + rank] = slot;
}

// order each bucket like the head/next chain: descending hs%b slot for
// cells (push-front, Cells.F90:299-300), ascending for the Verlet scan
// (Neighbor.F90:397); then copy positions into bucket order.
__global__ void k_cell_sort(DevState s, int gcs) {
  int r = blockIdx.y;
  const RepParams &P = s.rp[r];
  if (P.halted) return;
  if (!want_cells(P, gcs)) return;
  int nc = ncell_total(P);
  int desc = P.use_cells;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
    int cnt = s.cell_cnt[P.coff + c];
    if (cnt > P.bcap) cnt = P.bcap;
    if (cnt == 0) continue;
    int64_t base = P.boff + (int64_t)c * P.bcap;
    int32_t *b = s.bucket + base;
    for (int q = 1; q < cnt; q++) {
      int v = b[q], kv = s.bkey[v];
      int w = q - 1;
      while (w >= 0) {
        int kw = s.bkey[b[w]];
        if (desc ? (kw >= kv) : (kw <= kv)) break;
        b[w + 1] = b[w];
        w--;
      }
      b[w + 1] = v;
    }
    for (int q = 0; q < cnt; q++) {
      int v = b[q];
      s.bx[base + q] = s.px[v];
      s.by[base + q] = s.py[v];
      s.bz[base + q] = s.pz[v];
      s.bf4[base + q] = bucket_f4(s, v);
    }
  }
}

// ------------------------------------------------------------------------
// neighbour rows: ngroup_cells (Neighbor.F90:465-548) -- one wave per ref
// atom; candidates of the 27 stencil cells in stencil order, then chain
// order; strict '<' (':530'); compaction by ballot keeps that order.
// ngroup_verlet (358-424) when the box has fewer than 4 cells per side.
// ------------------------------------------------------------------------
template <int WPB>
__global__ void __launch_bounds__(WPB *WAVE) k_listbuild(DevState s) {
  int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  // launch block b covers WPB atoms of particle block wm_blk[b / (256 / WPB)]
  const int per = DML_BLOCK / WPB;
  int t = s.wm_blk[blockIdx.x / per] * DML_BLOCK + (blockIdx.x % per) * WPB + w;
  int r = s.blk2rep[t / DML_BLOCK];
  const RepParams &P = s.rp[r];
  if (P.halted) return;
  if (!P.rebuild || P.lb_kind != 1) return;
  int local = t - P.poff;
  if (local >= P.nref) return;
  int i = s.ref_order[t];
  double xi = s.px[i], yi = s.py[i], zi = s.pz[i];
  int ghost = 0;
  int cnt = 0;
  int64_t base = 0;
  if (P.use_cells) {
    int c0 = (int)(xi / P.cell[0]) + 1;
    int c1 = (int)(yi / P.cell[1]) + 1;
    int c2 = (int)(zi / P.cell[2]) + 1;
    if (c0 < 0 || c1 < 0 || c2 < 0 || c0 > P.ncells[0] + 1 || c1 > P.ncells[1] + 1 ||
        c2 > P.ncells[2] + 1) {
      // out of tessellation: k_cell_key has flagged it; leave the row empty
      if (lane == 0) s.nn[i] = 0;
      return;
    }
    ghost = (c0 < 1 || c1 < 1 || c2 < 1 || c0 > P.ncells[0] || c1 > P.ncells[1] ||
             c2 > P.ncells[2]);
    if (lane < 27) {
      int n0 = cell_wrap(c0 + kMap[lane][0], P.ncells[0]);
      int n1 = cell_wrap(c1 + kMap[lane][1], P.ncells[1]);
      int n2 = cell_wrap(c2 + kMap[lane][2], P.ncells[2]);
      int cell = cell_lin(P, n0, n1, n2);
      cnt = s.cell_cnt[P.coff + cell];
      if (cnt > P.bcap) cnt = P.bcap;
      base = P.boff + (int64_t)cell * P.bcap;
    }
  } else if (lane == 0) {
    cnt = s.cell_cnt[P.coff];
    if (cnt > P.bcap) cnt = P.bcap;
    base = P.boff;
  }
  // inclusive scan of the 27 counts held by lanes 0..26
  int incl = cnt;
  for (int off = 1; off < 32; off <<= 1) {
    int v = __shfl_up(incl, off);
    if (lane >= off) incl += v;
  }
  const int pre = incl - cnt;  // exclusive prefix of stencil cell `lane`
  const int total = __shfl(incl, 26);
  const double rl2 = P.rl2;
  const int strict = P.use_cells;
  const int maxnn = s.maxnn;
  int32_t *row = s.nbr + (int64_t)i * maxnn;
  int nn = 0;
  for (int c0 = 0; c0 < total; c0 += WAVE) {
    int c = c0 + lane;
    int cc = c < total ? c : total - 1;
    int lo = 0, hi = 26;  // largest k with pre(k) <= cc
    for (int it = 0; it < 5; it++) {
      int mid = (lo + hi + 1) >> 1;
      int pm = __shfl(pre, mid);
      if (lo < hi) {
        if (pm <= cc) lo = mid; else hi = mid - 1;
      }
    }
    int64_t bk = __shfl(base, lo);
    int pk = __shfl(pre, lo);
    int hit = 0, j = -1;
    if (c < total) {
      int64_t pos = bk + (cc - pk);
      j = s.bucket[pos];
      double dx, dy, dz;
      vdist_mic(P, s.bx[pos], s.by[pos], s.bz[pos], xi, yi, zi, &dx, &dy, &dz);
      double rd = dot3(dx, dy, dz);
      hit = (strict ? (rd < rl2) : !(rd > rl2)) && (j != i);
    }
    unsigned long long m = __ballot(hit);
    int pr = __popcll(m & ((1ull << lane) - 1ull));
    if (hit && nn + pr < maxnn) row[nn + pr] = j;
    nn += __popcll(m);
  }
  if (lane == 0) {
    if (nn > maxnn) {
      atomicOr(s.err, ERR_LIST);
      atomicMax(&s.rp[r].max_nn_seen, nn);
      rep_fail(s, r, ERR_LIST);
      nn = maxnn;
    }
    s.nn[i] = nn;
    s.rowflag[i] = (uint8_t)ghost;
    if (ghost) s.rp[r].any_ghost = 1;
    atomicAdd(&s.cnt[r].nn_sum, (unsigned long long)nn);
    atomicAdd(&s.cnt[r].nn_rows, 1ull);
  }
}

template __global__ void k_listbuild<4>(DevState s);

// ------------------------------------------------------------------------
// neighbour rows, cell-tile form: one workgroup per cell.  The candidates of
// the 27 stencil cells (stencil order, chain order inside each cell) are
// staged through LDS in tiles and shared by every ref atom of the centre
// cell; each wave owns some centre atoms and appends to their rows by
// ballot compaction, so every row keeps the order of ngroup_cells
// (Neighbor.F90:509-535).  Same rows as k_listbuild, ~cell-occupancy times
// fewer L2/HBM reads.  Like k_listbuild_wc it lists each ref atom's entries
// within rcut (nearl, count in ov_near) for the overlap pass.
// ------------------------------------------------------------------------
#define LB_TILE 1024
#define LB_CMAX 256
__global__ void __launch_bounds__(DML_BLOCK) k_listbuild_cells(DevState s) {
  __shared__ int32_t t_slot[LB_TILE];
  __shared__ float4 t_f[LB_TILE];
  __shared__ int32_t c_slot[LB_CMAX];
  __shared__ double c_x[LB_CMAX], c_y[LB_CMAX], c_z[LB_CMAX];
  __shared__ int32_t c_nn[LB_CMAX], c_nc[LB_CMAX];
  __shared__ int64_t st_base[27];
  __shared__ int32_t st_pre[28];
  __shared__ int32_t sh_A, sh_nrefc;
  __shared__ unsigned long long sh_nnsum, sh_rows;
  const int tid = threadIdx.x, w = tid / WAVE, lane = tid % WAVE;
  const int nwav = DML_BLOCK / WAVE;
  const int maxnn = s.maxnn;
  if (tid == 0) { sh_nnsum = 0; sh_rows = 0; }
  // one cell of the flat list of cell-tile replicas per iteration
  for (int64_t g = blockIdx.x; g < s.n_tm; g += gridDim.x) {
    const int r = s.tm_rep[g];
    const int c = s.tm_cell[g];
    const RepParams &P = s.rp[r];
    if (P.halted || !P.rebuild || P.lb_kind != 0 || c >= ncell_total(P)) continue;
    const double rl2 = P.rl2;
    const int strict = P.use_cells;
    const float rl2lo = (float)rl2 - P.pf_margin, rl2hi = (float)rl2 + P.pf_margin;
    const float rcf = (float)P.rcut2 + P.pf_margin;
    const float bxf = (float)P.box[0], byf = (float)P.box[1];
    const float ibx = (float)P.one_box[0], iby = (float)P.one_box[1];
    if (tid == 0) {
      int A = s.cell_cnt[P.coff + c];
      sh_A = A > P.bcap ? P.bcap : A;
    }
    __syncthreads();
    const int Atot = sh_A;
    const int64_t cbase = P.boff + (int64_t)c * P.bcap;
    int ghost = 0;
    int ci0 = 0, ci1 = 0, ci2 = 0;
    if (P.use_cells) {
      ci0 = c % (P.ncells[0] + 2);
      ci1 = (c / (P.ncells[0] + 2)) % (P.ncells[1] + 2);
      ci2 = c / ((P.ncells[0] + 2) * (P.ncells[1] + 2));
      ghost = (ci0 < 1 || ci1 < 1 || ci2 < 1 || ci0 > P.ncells[0] || ci1 > P.ncells[1] ||
               ci2 > P.ncells[2]);
    }
    for (int a0 = 0; a0 < Atot; a0 += LB_CMAX) {
      const int A = (Atot - a0) < LB_CMAX ? (Atot - a0) : LB_CMAX;
      if (tid == 0) sh_nrefc = 0;
      __syncthreads();
      for (int q = tid; q < A; q += DML_BLOCK) {
        int v = s.bucket[cbase + a0 + q];
        int isref = (s.flags[v] & DML_FL_REF) != 0;
        c_slot[q] = isref ? v : -1;
        c_x[q] = s.bx[cbase + a0 + q];
        c_y[q] = s.by[cbase + a0 + q];
        c_z[q] = s.bz[cbase + a0 + q];
        c_nn[q] = 0;
        c_nc[q] = 0;
        if (isref) atomicAdd(&sh_nrefc, 1);
      }
      if (tid < 27) {
        int cnt = 0;
        int64_t base = 0;
        if (P.use_cells) {
          int n0 = cell_wrap(ci0 + kMap[tid][0], P.ncells[0]);
          int n1 = cell_wrap(ci1 + kMap[tid][1], P.ncells[1]);
          int n2 = cell_wrap(ci2 + kMap[tid][2], P.ncells[2]);
          int cell = cell_lin(P, n0, n1, n2);
          cnt = s.cell_cnt[P.coff + cell];
          if (cnt > P.bcap) cnt = P.bcap;
          base = P.boff + (int64_t)cell * P.bcap;
        } else if (tid == 0) {
          cnt = Atot;
          base = P.boff;
        }
        st_base[tid] = base;
        st_pre[tid + 1] = cnt;  // counts, scanned below
      }
      __syncthreads();
      if (tid == 0) {
        st_pre[0] = 0;
        for (int k = 1; k <= 27; k++) st_pre[k] += st_pre[k - 1];
      }
      __syncthreads();
      const int total = st_pre[27];
      if (sh_nrefc == 0 || total == 0) {
        __syncthreads();
        continue;
      }
      for (int t0 = 0; t0 < total; t0 += LB_TILE) {
        const int tn = (total - t0) < LB_TILE ? (total - t0) : LB_TILE;
        for (int q = tid; q < tn; q += DML_BLOCK) {
          int cidx = t0 + q;
          int lo = 0, hi = 26;
          while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (st_pre[mid] <= cidx) lo = mid; else hi = mid - 1;
          }
          int64_t pos = st_base[lo] + (cidx - st_pre[lo]);
          t_slot[q] = s.bucket[pos];
          t_f[q] = s.bf4[pos];
        }
        __syncthreads();
        for (int a = w; a < A; a += nwav) {
          const int i = c_slot[a];
          if (i < 0) continue;
          const double xi = c_x[a], yi = c_y[a], zi = c_z[a];
          const float4 fi = make_float4((float)xi, (float)yi, (float)zi, 0.f);
          int nn = c_nn[a], nc = c_nc[a];
          int32_t *row = s.nbr + (int64_t)i * maxnn;
          for (int q0 = 0; q0 < tn; q0 += WAVE) {
            int q = q0 + lane;
            int hit = 0, j = -1;
            float df = 3.0e38f;
            if (q < tn) {
              j = t_slot[q];
              df = d2_f32(P, t_f[q], fi, bxf, byf, ibx, iby);
              if (df < rl2lo) {
                hit = 1;  // inside by more than the FP32 error
              } else if (df <= rl2hi) {
                double dx, dy, dz;  // near the list radius: decide in FP64
                vdist_mic(P, s.px[j], s.py[j], s.pz[j], xi, yi, zi, &dx, &dy, &dz);
                double rd = dot3(dx, dy, dz);
                hit = strict ? (rd < rl2) : !(rd > rl2);
              }
              hit = hit && (j != i);
            }
            unsigned long long m = __ballot(hit);
            int pr = __popcll(m & ((1ull << lane) - 1ull));
            if (hit && nn + pr < maxnn) row[nn + pr] = j;
            // the hits within rcut: their row positions
            const int nr = hit && df <= rcf;
            const unsigned long long nm = __ballot(nr);
            if (nm) {
              const int qn = nc + __popcll(nm & ((1ull << lane) - 1ull));
              if (nr && qn < NEAR_MAX) s.nearl[(int64_t)i * NEAR_MAX + qn] = nn + pr;
              nc += __popcll(nm);
            }
            nn += __popcll(m);
          }
          if (lane == 0) {
            c_nn[a] = nn;
            c_nc[a] = nc;
          }
        }
        __syncthreads();
      }
      for (int q = tid; q < A; q += DML_BLOCK) {
        int i = c_slot[q];
        if (i < 0) continue;
        int nn = c_nn[q];
        if (nn > maxnn) {
          atomicOr(s.err, ERR_LIST);
          atomicMax(&s.rp[r].max_nn_seen, nn);
          rep_fail(s, r, ERR_LIST);
          nn = maxnn;
        }
        s.nn[i] = nn;
        s.rowflag[i] = (uint8_t)ghost;
        s.ov_near[i] = (uint8_t)(c_nc[q] <= NEAR_MAX ? c_nc[q] : 255);
        if (ghost) s.rp[r].any_ghost = 1;
        atomicAdd(&sh_nnsum, (unsigned long long)nn);
        atomicAdd(&sh_rows, 1ull);
      }
      __syncthreads();
    }
    __syncthreads();  // sh_A and the LDS tiles are rewritten by the next cell
    if (tid == 0 && sh_rows) {
      atomicAdd(&s.cnt[r].nn_sum, sh_nnsum);
      atomicAdd(&s.cnt[r].nn_rows, sh_rows);
      sh_nnsum = 0;
      sh_rows = 0;
    }
  }
}

// ------------------------------------------------------------------------
// neighbour rows, one wave per cell: the wave holds up to 64 centre atoms of
// its cell in lanes (position, row length) and streams the candidates of the
// 27 stencil cells (stencil order, chain order inside each cell), 64 at a
// time, one per lane; every ref centre atom in turn tests the whole chunk
// (FP32 prefilter, FP64 inside the band) and appends its hits by ballot
// compaction, so rows keep the order of ngroup_cells (Neighbor.F90:509-535).
// With the image shift the same tests also list, per ref atom, the row
// positions of its entries within rcut (FP32 with margin; at most NEAR_MAX,
// count in ov_near, 255 when not listed) for the overlap pass at these
// positions.
// No LDS and no block barriers: candidate loads are shared by all atoms of
// the cell through registers, and dense and sparse cells cost in proportion
// to their work.  With at least 4 cells in x and y, a candidate's x, y are
// moved once per chunk to their image nearest the centre cell's midpoint,
// which is the minimum image for every centre atom whenever the pair can lie
// within the list radius (a pair whose images disagree is more than 1.5
// cells apart either way), so the pair test needs no image arithmetic.
// ------------------------------------------------------------------------
// a wave-uniform value read into scalar registers
__device__ inline int rfl_i(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline float rfl_f(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ inline int64_t rfl_l(int64_t v) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(v & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  return (int64_t)(uint32_t)lo | ((int64_t)hi << 32);
}
__device__ inline double rfl_d(double v) { return __longlong_as_double(rfl_l(__double_as_longlong(v))); }

// inclusive prefix sum over the wave by DPP moves: row_shr 1, 2, 4, 8
// inside each row of 16 lanes, then row_bcast 15 and 31 carry the row
// totals (rows 1, 3 and rows 2, 3); a lane with no source adds 0.  No lane
// index registers (a shuffle's addresses, hoisted by the compiler, would
// hold VGPRs through the kernel)
__device__ inline int wave_scan_incl(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);
  return v;
}
// the same within each row of 16 lanes, and within each half of 32
__device__ inline int row_scan16(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);
  return v;
}
__device__ inline int half_scan32(int v) {
  v = row_scan16(v);
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);
  return v;
}

__device__ inline float rdlane_f(float v, int a) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), a));
}
__device__ inline double rdlane_d(double v, int a) {
  int64_t b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), a);
  int hi = __builtin_amdgcn_readlane((int)(b >> 32), a);
  return __longlong_as_double(((int64_t)(uint32_t)lo) | ((int64_t)hi << 32));
}

// 5 waves per SIMD: the pair tests are a serial chain per centre atom, so
// the issue rate follows the waves resident; the hint brings the kernel to
// 95 VGPRs (5 waves) without spills, 3.8 -> 3.3 ms on a 256-replica ensemble
__global__ void __launch_bounds__(DML_BLOCK) __attribute__((amdgpu_waves_per_eu(5)))
k_listbuild_wc(DevState s) {
  const int lane = threadIdx.x % WAVE;
  const int maxnn = s.maxnn;
  // this lane's stencil offset (lanes 0-26), three fields of 4 bits, biased by 1
  const int kofs = lane < 27 ? (kMap[lane][0] + 1) | (kMap[lane][1] + 1) << 4 |
                                   (kMap[lane][2] + 1) << 8
                             : 0x111;
  // XCD-aware: workgroups are dealt round-robin to the 8 XCDs, so
  // workgroup b works on slice (b % 8) of the cell list and, inside it,
  // every wave on one contiguous run of cells; the waves of one XCD then
  // sweep neighbouring cells together and share their candidates in its L2
  const int nb = gridDim.x;  // a multiple of 8
  const int64_t vb = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  const int64_t wid =
      vb * (DML_BLOCK / WAVE) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const int64_t nwv = (int64_t)nb * (DML_BLOCK / WAVE);
  const int64_t run = (s.n_cm + nwv - 1) / nwv;
  const int64_t g_end = min(s.n_cm, (wid + 1) * run);
  int cur_r = -1;
  unsigned long long acc_nn = 0, acc_rows = 0;
  for (int64_t g0 = wid * run; g0 < g_end; g0 += WAVE) {
    // 64 cells of the run at once, one per lane: most cells of sparse
    // replicas are empty, and their checks need no serial chain of loads
    int lr = 0, lc = 0, lcnt = 0;
    if (g0 + lane < g_end) {
      lr = s.cm_rep[g0 + lane];
      lc = s.cm_cell[g0 + lane];
      const RepParams &Q = s.rp[lr];
      if (!Q.halted && Q.rebuild && Q.lb_kind == 2 && lc < ncell_total(Q))
        lcnt = min(s.cell_cnt[Q.coff + lc], Q.bcap);
    }
    unsigned long long work = __ballot(lcnt > 0);
    while (work) {
    const int q = __builtin_ctzll(work);
    work &= work - 1;
    const int r = __builtin_amdgcn_readlane(lr, q);
    const int c = __builtin_amdgcn_readlane(lc, q);
    const int ccnt = __builtin_amdgcn_readlane(lcnt, q);
    const RepParams &P = s.rp[r];
    if (r != cur_r) {
      if (cur_r >= 0 && lane == 0 && acc_rows) {
        atomicAdd(&s.cnt[cur_r].nn_sum, acc_nn);
        atomicAdd(&s.cnt[cur_r].nn_rows, acc_rows);
      }
      cur_r = r;
      acc_nn = acc_rows = 0;
    }
    const int64_t cbase = P.boff + (int64_t)c * P.bcap;
    int ghost = 0, ci0 = 0, ci1 = 0, ci2 = 0;
    if (P.use_cells) {
      ci0 = c % (P.ncells[0] + 2);
      ci1 = (c / (P.ncells[0] + 2)) % (P.ncells[1] + 2);
      ci2 = c / ((P.ncells[0] + 2) * (P.ncells[1] + 2));
      ghost = (ci0 < 1 || ci1 < 1 || ci2 < 1 || ci0 > P.ncells[0] || ci1 > P.ncells[1] ||
               ci2 > P.ncells[2]);
    }
    // the 27 stencil cells: lane k holds count and first entry of cell k
    int kcnt = 0;
    int64_t kbase = 0;
    if (P.use_cells) {
      if (lane < 27) {
        int n0 = cell_wrap(ci0 + (kofs & 15) - 1, P.ncells[0]);
        int n1 = cell_wrap(ci1 + (kofs >> 4 & 15) - 1, P.ncells[1]);
        int n2 = cell_wrap(ci2 + (kofs >> 8 & 15) - 1, P.ncells[2]);
        int cell = cell_lin(P, n0, n1, n2);
        kcnt = s.cell_cnt[P.coff + cell];
        if (kcnt > P.bcap) kcnt = P.bcap;
        kbase = P.boff + (int64_t)cell * P.bcap;
      }
    } else if (lane == 0) {
      kcnt = ccnt;
      kbase = P.boff;
    }
    int incl = kcnt;
    for (int off = 1; off < 32; off <<= 1) {
      int v = __shfl_up(incl, off);
      if (lane >= off) incl += v;
    }
    const int kpre = incl - kcnt;
    const int total = __shfl(incl, 26);
    const double rl2 = P.rl2;
    const int strict = P.use_cells;
    const float rl2lo = (float)rl2 - P.pf_margin, rl2hi = (float)rl2 + P.pf_margin;
    const float rcf = (float)P.rcut2 + P.pf_margin;
    const float bxf = (float)P.box[0], byf = (float)P.box[1];
    const float ibx = (float)P.one_box[0], iby = (float)P.one_box[1];
    const int img = __builtin_amdgcn_readfirstlane(
        (P.use_cells && P.ncells[0] >= 4 && P.ncells[1] >= 4) ? 1 : 0);
    const float mxf = (float)((ci0 - 0.5) * P.cell[0]), myf = (float)((ci1 - 0.5) * P.cell[1]);
    for (int a0 = 0; a0 < ccnt; a0 += WAVE) {
      // centre atoms a0 .. a0+63 of this cell, one per lane
      const int na = (ccnt - a0) < WAVE ? (ccnt - a0) : WAVE;
      // (a centre atom's FP64 position is read again, at a wave-uniform
      // address, on the rare FP64 path: no registers held for it)
      int my_i = -1, my_ref = 0;
      float4 my_f = make_float4(0.f, 0.f, 0.f, 0.f);
      const int64_t e0 = cbase + a0;
      if (lane < na) {
        my_i = s.bucket[e0 + lane];
        my_ref = (s.flags[my_i] & DML_FL_REF) != 0;
        my_f = make_float4((float)s.bx[e0 + lane], (float)s.by[e0 + lane], (float)s.bz[e0 + lane], 0.f);
      }
      const unsigned long long refm = __ballot(my_ref);
      if (!refm) continue;
      // row length (bits 0-23) and entries within rcut (bits 24-31,
      // saturating) in one register
      int my_nn = 0;
      // candidate cidx of the stream: its bucket entry (the stencil cell
      // holding it by a binary search over the cell starts)
      auto fetch = [&](int cidx, int *cj, float4 *cf) {
        const int cc = cidx < total ? cidx : total - 1;
        int lo = 0, hi = 26;
        for (int it = 0; it < 5; it++) {
          int mid = (lo + hi + 1) >> 1;
          int pm = __shfl(kpre, mid);
          if (lo < hi) {
            if (pm <= cc) lo = mid; else hi = mid - 1;
          }
        }
        const int64_t pos = __shfl(kbase, lo) + (cc - __shfl(kpre, lo));
        *cj = s.bucket[pos];
        *cf = s.bf4[pos];
      };
      // the next chunk's loads are issued before this chunk's tests
      int nj;
      float4 nf;
      fetch(lane, &nj, &nf);
      for (int q0 = 0; q0 < total; q0 += WAVE) {
        const int cidx = q0 + lane;
        const int cj = nj;
        const float4 cf = nf;
        if (q0 + WAVE < total) fetch(q0 + WAVE + lane, &nj, &nf);
        const int valid = cidx < total;
        if (img) {
          // x, y at the candidate's image nearest the centre cell's midpoint
          // (a lane past the stream's end sits infinitely far away)
          const float cx = valid ? cf.x - bxf * rintf((cf.x - mxf) * ibx) : 3.0e38f;
          const float cy = cf.y - byf * rintf((cf.y - myf) * iby);
          unsigned long long m = refm;
          while (m) {
            const int a = __builtin_ctzll(m);
            m &= m - 1;
            const int i = __builtin_amdgcn_readlane(my_i, a);
            const float dx = cx - rdlane_f(my_f.x, a), dy = cy - rdlane_f(my_f.y, a);
            const float dz = cf.z - rdlane_f(my_f.z, a);
            const float df = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            // hits append in candidate order; the masks stay in scalar
            // registers (no per-lane 0/1 round trip through a ballot)
            auto append = [&](bool hit, unsigned long long hm) {
              if (!hm) return;
              const int pk = __builtin_amdgcn_readlane(my_nn, a);
              const int nn0 = pk & 0xffffff;
              const int nh = __popcll(hm);
              const int pr = __builtin_amdgcn_mbcnt_hi((uint32_t)(hm >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)hm, 0u));
              int32_t *row = s.nbr + (int64_t)i * maxnn + nn0;
              if (nn0 + nh <= maxnn) {
                if (hit) row[pr] = cj;
              } else if (hit && nn0 + pr < maxnn) {
                row[pr] = cj;
              }
              int npk = pk + nh;
              // the hits within rcut: their row positions
              const unsigned long long nm = __builtin_amdgcn_ballot_w64(df <= rcf) & hm;
              if (nm) {
                const int nc0 = (int)((uint32_t)pk >> 24);
                const int q = nc0 + (int)__builtin_amdgcn_mbcnt_hi(
                                        (uint32_t)(nm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)nm, 0u));
                if (hit && df <= rcf && q < NEAR_MAX) s.nearl[(int64_t)i * NEAR_MAX + q] = nn0 + pr;
                const int add = __popcll(nm);
                npk += (add < 255 - nc0 ? add : 255 - nc0) << 24;
              }
              my_nn = lane == a ? npk : my_nn;
            };
            const unsigned long long gem = __builtin_amdgcn_ballot_w64(df >= rl2lo);
            if (gem & __builtin_amdgcn_ballot_w64(df <= rl2hi)) {
              // near the list radius: decide in FP64 (rare)
              bool hit = df < rl2lo;
              if (df >= rl2lo && df <= rl2hi) {
                double ex, ey, ez;
                const uint32_t uj = (uint32_t)cj;
                vdist_mic(P, s.px[uj], s.py[uj], s.pz[uj], s.bx[e0 + a], s.by[e0 + a], s.bz[e0 + a],
                          &ex, &ey, &ez);
                const double rd = dot3(ex, ey, ez);
                hit = strict ? (rd < rl2) : !(rd > rl2);
              }
              const bool h = hit && cj != i;
              append(h, __builtin_amdgcn_ballot_w64(h));
            } else {
              // a ballot of a plain compare is that compare's own mask; the
              // hits are the lanes below rl2lo (not in gem) that are not i
              const bool other = cj != i;
              append(!(df >= rl2lo) && other, ~gem & __builtin_amdgcn_ballot_w64(other));
            }
          }
          continue;
        }
        unsigned long long m = refm;
        while (m) {
          const int a = __builtin_ctzll(m);
          m &= m - 1;
          const int i = __builtin_amdgcn_readlane(my_i, a);
          const float4 fi = make_float4(rdlane_f(my_f.x, a), rdlane_f(my_f.y, a),
                                        rdlane_f(my_f.z, a), 0.f);
          int hit = 0;
          if (valid) {
            const float df = d2_f32(P, cf, fi, bxf, byf, ibx, iby);
            if (df < rl2lo) {
              hit = 1;
            } else if (df <= rl2hi) {
              double dx, dy, dz;  // near the list radius: decide in FP64
              const uint32_t uj = (uint32_t)cj;
              vdist_mic(P, s.px[uj], s.py[uj], s.pz[uj], s.bx[e0 + a], s.by[e0 + a], s.bz[e0 + a],
                        &dx, &dy, &dz);
              const double rd = dot3(dx, dy, dz);
              hit = strict ? (rd < rl2) : !(rd > rl2);
            }
            hit = hit && (cj != i);
          }
          const unsigned long long hm = __ballot(hit);
          if (hm) {
            const int pk = __builtin_amdgcn_readlane(my_nn, a);
            const int nn0 = pk & 0xffffff;
            const int pr = __popcll(hm & ((1ull << lane) - 1ull));
            if (hit && nn0 + pr < maxnn) s.nbr[(int64_t)i * maxnn + nn0 + pr] = cj;
            if (lane == a) my_nn = pk + __popcll(hm);
          }
        }
      }
      const int my_nc = (int)((uint32_t)my_nn >> 24);
      my_nn &= 0xffffff;
      if (my_ref) {
        int nn = my_nn;
        if (nn > maxnn) {
          atomicOr(s.err, ERR_LIST);
          atomicMax(&s.rp[r].max_nn_seen, nn);
          rep_fail(s, r, ERR_LIST);
          nn = maxnn;
        }
        s.nn[my_i] = nn;
        s.rowflag[my_i] = (uint8_t)ghost;
        s.ov_near[my_i] = (uint8_t)(img && my_nc <= NEAR_MAX ? my_nc : 255);
        if (ghost) s.rp[r].any_ghost = 1;
      }
      // row statistics of this group (wave-uniform: kept in scalar registers)
      int v = my_ref ? my_nn : 0;
      for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
      acc_nn += (unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane(v);
      acc_rows += (unsigned long long)__popcll(refm);
    }
    }
  }
  if (cur_r >= 0 && lane == 0 && acc_rows) {
    atomicAdd(&s.cnt[cur_r].nn_sum, acc_nn);
    atomicAdd(&s.cnt[cur_r].nn_rows, acc_rows);
  }
}

// ------------------------------------------------------------------------
// Rows of one centre cell with its candidates held in registers: a wave
// loads up to NCH chunks of 64 candidates of the 27-cell stream at a time
// (WPE waves per SIMD: the walk is latency-bound, occupancy pays),
// then walks the cell's ref atoms one at a time; each atom's count, row
// address and position stay in scalar registers across the chunks, and a
// chunk's hits are one compare, one ballot and one store of the chunk's
// lanes in stream order (stencil order, then chain order: ngroup_cells,
// Neighbor.F90:509-535, strict '<'; ngroup_verlet, ':358-424', rejects on
// '>').  FP32 prefilter with margin, FP64 in the band.  Entries within
// rcut (FP32 with margin) are listed like k_listbuild_wc lists them.
// ------------------------------------------------------------------------
template <int NCH, int WPE>
__global__ void __launch_bounds__(DML_BLOCK) __attribute__((amdgpu_waves_per_eu(WPE)))
k_listbuild_wr(DevState s) {
  // one wave's LDS slice, at a wave-uniform base (one lane offset serves
  // every array)
  struct Slice {
    int32_t pos[NCH * WAVE];  // bucket entries of a pass's stream positions
    // the 64 cells of a batch (replica, cell, count, x, y|z|ghost, image
    // reference x, y), kept here rather than in registers
    int32_t bat[7][WAVE];
    int32_t nn0[WAVE];  // a lane's row length at the start of a pass
    int4 ctr[WAVE];     // the cell's current centre atoms: slot, FP32 position
  };
  __shared__ Slice lbr[DML_BLOCK / WAVE];
  Slice &L = lbr[__builtin_amdgcn_readfirstlane(threadIdx.x / WAVE)];
  int32_t *wpos = L.pos;
  int32_t(*bat)[WAVE] = L.bat;
  int32_t *nn0s = L.nn0;
  int4 *ctr = L.ctr;
  const int lane = threadIdx.x % WAVE;
  const int maxnn = s.maxnn;
  const int kofs = lane < 27 ? (kMap[lane][0] + 1) | (kMap[lane][1] + 1) << 4 |
                                   (kMap[lane][2] + 1) << 8
                             : 0x111;
  // XCD-aware runs of cells, as in k_listbuild_wc
  const int nb = gridDim.x;  // a multiple of 8
  const int64_t vb = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
  const int64_t wid =
      vb * (DML_BLOCK / WAVE) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  // (32-bit division: the cell list holds fewer than 2^31 cells)
  const uint32_t nwv = (uint32_t)nb * (DML_BLOCK / WAVE);
  const int64_t run = (int64_t)(((uint32_t)s.n_cm + nwv - 1u) / nwv);
  const int64_t g_end = min(s.n_cm, (wid + 1) * run);
  int cur_r = -1;
  unsigned long long acc_nn = 0, acc_rows = 0;
  // the current replica's parameters, read once per replica (the parameter
  // block is written by this kernel's rare error paths, so its fields are
  // vector loads; read in the cell loop they would be a chain of waits)
  int64_t q_boff = 0;
  int q_cells = 0, q_n0 = 0, q_n1 = 0, q_n2 = 0, q_coff = 0, q_bcap = 0, q_img = 0;
  double q_rl2 = 0.0;
  float q_lo = 0.f, q_hi = 0.f, q_rcf = 0.f, q_bx = 0.f, q_by = 0.f, q_ibx = 0.f, q_iby = 0.f;
  float q_bandw = 0.f;
  for (int64_t g0 = wid * run; g0 < g_end; g0 += WAVE) {
    // 64 cells of the run, one per lane: count, coordinates (y, z packed,
    // the ghost-layer flag in bit 31) and the image reference point, so
    // the cell loop below needs no division and no parameter reads
    int lr = 0, lc = 0, lcnt = 0, lx = 0, lyz = 0;
    float lmx = 0.f, lmy = 0.f;
    if (g0 + lane < g_end) {
      lr = s.cm_rep[g0 + lane];
      lc = s.cm_cell[g0 + lane];
      const RepParams &Q = s.rp[lr];
      if (!Q.halted && Q.rebuild && Q.lb_kind == 3 && lc < ncell_total(Q)) {
        lcnt = min(s.cell_cnt[Q.coff + lc], Q.bcap);
        if (Q.use_cells) {
          const int w0 = Q.ncells[0] + 2, w1 = Q.ncells[1] + 2;
          const int x = lc % w0, y = lc / w0 % w1, z = lc / (w0 * w1);
          const int gh = x < 1 || y < 1 || z < 1 || x > Q.ncells[0] || y > Q.ncells[1] ||
                         z > Q.ncells[2];
          lx = x;
          lyz = y | z << 16 | gh << 31;
          lmx = (float)((x - 0.5) * Q.cell[0]);
          lmy = (float)((y - 0.5) * Q.cell[1]);
        }
      }
    }
    unsigned long long work = __ballot(lcnt > 0);
    __builtin_amdgcn_wave_barrier();
    bat[0][lane] = lr;
    bat[1][lane] = lc;
    bat[2][lane] = lcnt;
    bat[3][lane] = lx;
    bat[4][lane] = lyz;
    bat[5][lane] = __float_as_int(lmx);
    bat[6][lane] = __float_as_int(lmy);
    __builtin_amdgcn_wave_barrier();
    while (work) {
      const int q = __builtin_ctzll(work);
      work &= work - 1;
      const int r = rfl_i(bat[0][q]);
      const int c = rfl_i(bat[1][q]);
      const int ccnt = rfl_i(bat[2][q]);
      const int cx0 = rfl_i(bat[3][q]);
      const int cyz = rfl_i(bat[4][q]);
      const float mxf = __int_as_float(rfl_i(bat[5][q])), myf = __int_as_float(rfl_i(bat[6][q]));
      const RepParams &P = s.rp[r];
      if (r != cur_r) {
        if (cur_r >= 0 && lane == 0 && acc_rows) {
          atomicAdd(&s.cnt[cur_r].nn_sum, acc_nn);
          atomicAdd(&s.cnt[cur_r].nn_rows, acc_rows);
        }
        cur_r = r;
        acc_nn = acc_rows = 0;
        q_boff = rfl_l(P.boff);
        q_cells = rfl_i(P.use_cells);
        q_n0 = rfl_i(P.ncells[0]);
        q_n1 = rfl_i(P.ncells[1]);
        q_n2 = rfl_i(P.ncells[2]);
        q_coff = rfl_i(P.coff);
        q_bcap = rfl_i(P.bcap);
        q_img = q_cells && q_n0 >= 4 && q_n1 >= 4;
        q_rl2 = rfl_d(P.rl2);
        const float pfm = rfl_f(P.pf_margin);
        q_lo = (float)q_rl2 - pfm;
        q_hi = (float)q_rl2 + pfm;
        q_rcf = (float)rfl_d(P.rcut2) + pfm;
        q_bx = (float)rfl_d(P.box[0]);
        q_by = (float)rfl_d(P.box[1]);
        q_ibx = (float)rfl_d(P.one_box[0]);
        q_iby = (float)rfl_d(P.one_box[1]);
        // a pair is near the list radius when |d^2 - rl^2| <= bandw in
        // FP32: a superset of [rl2lo, rl2hi] (the difference of two floats
        // this close is exact; rl2lo, rl2hi are within half an ulp of
        // rl^2 -+ margin)
        q_bandw = pfm * 1.0001f + 1.0e-6f * (float)q_rl2;
      }
      const int64_t cbase = q_boff + (int64_t)c * q_bcap;
      int ghost = 0, ci0 = 0, ci1 = 0, ci2 = 0;
      if (q_cells) {
        ci0 = cx0;
        ci1 = cyz & 0xffff;
        ci2 = cyz >> 16 & 0x7fff;
        ghost = (uint32_t)cyz >> 31;
      }
      // the first 64 centre atoms' entries, loaded with the stencil counts
      // (one wait for both)
      int4 my_c = make_int4(-1, 0, 0, 0);
      int my_ref = 0;
      if (lane < ccnt) {
        my_c.x = s.bucket[cbase + lane];
        const float4 f = s.bf4[cbase + lane];
        my_c.y = __float_as_int(f.x);
        my_c.z = __float_as_int(f.y);
        my_c.w = __float_as_int(f.z);
        my_ref = (__float_as_int(f.w) >> 8 & DML_FL_REF) != 0;
      }
      // the 27 stencil cells: lane k holds count and first entry of cell k
      // (the first entry relative to the replica's bucket region, which
      // holds fewer than 2^31 entries)
      int kcnt = 0, kb32 = 0;
      // (the offsets are unpacked here, each cell: kept unpacked across the
      // loop they would hold three registers)
      int ko = kofs;
      asm volatile("" : "+v"(ko));
      if (q_cells) {
        if (lane < 27) {
          const int n0 = cell_wrap1(ci0 + (ko & 15) - 1, q_n0);
          const int n1 = cell_wrap1(ci1 + (ko >> 4 & 15) - 1, q_n1);
          const int n2 = cell_wrap1(ci2 + (ko >> 8 & 15) - 1, q_n2);
          const int cell = n0 + (q_n0 + 2) * (n1 + (q_n1 + 2) * n2);
          kcnt = min(s.cell_cnt[q_coff + cell], q_bcap);
          kb32 = cell * q_bcap;
        }
      } else if (lane == 0) {
        kcnt = ccnt;
      }
      const int incl = wave_scan_incl(kcnt);
      const int kpre = incl - kcnt;
      // (read from lane 26 into a scalar register: the compiler then keeps
      // the chunk loops' bounds and masks uniform)
      const int total = __builtin_amdgcn_readlane(incl, 26);
      const double rl2 = q_rl2;
      const int strict = q_cells;
      const float rl2lo = q_lo, rl2hi = q_hi, rcf = q_rcf;
      const float bxf = q_bx, byf = q_by, ibx = q_ibx, iby = q_iby;
      const int img = q_img;
      for (int a0 = 0; a0 < ccnt; a0 += WAVE) {
        // centre atoms a0 .. a0+63 of this cell, one per lane
        const int na = (ccnt - a0) < WAVE ? (ccnt - a0) : WAVE;
        const int64_t e0 = cbase + a0;
        if (a0 > 0) {
          // the next 64 centre atoms
          my_c = make_int4(-1, 0, 0, 0);
          my_ref = 0;
          if (lane < na) {
            my_c.x = s.bucket[e0 + lane];
            const float4 f = s.bf4[e0 + lane];
            my_c.y = __float_as_int(f.x);
            my_c.z = __float_as_int(f.y);
            my_c.w = __float_as_int(f.z);
            my_ref = (__float_as_int(f.w) >> 8 & DML_FL_REF) != 0;
          }
        }
        const unsigned long long refm = __ballot(my_ref);
        if (!refm) continue;
        // (read back at wave-uniform addresses by the walks below)
        __builtin_amdgcn_wave_barrier();
        ctr[lane] = my_c;
        __builtin_amdgcn_wave_barrier();
        // row length (bits 0-23) and entries within rcut (24-31, saturating)
        int my_nn = 0;
        for (int s0 = 0; s0 < total; s0 += NCH * WAVE) {
          // (the previous super-chunk's reads of the slice are done)
          __builtin_amdgcn_wave_barrier();
          const int nch = min(NCH, (total - s0 + WAVE - 1) / WAVE);
          const int last = min(total - s0, NCH * WAVE) - 1;
          // the bucket entries of stream positions s0 .. s0+last, in stream
          // order: each stencil lane writes those of its cell (one wave's
          // LDS slice, written and read by that wave only, in order)
          {
            const int lo = max(0, s0 - kpre), hi = min(kcnt, s0 + last + 1 - kpre);
            for (int e = lo; e < hi; e++) wpos[kpre + e - s0] = kb32 + e;
          }
          __builtin_amdgcn_wave_barrier();
          // (the slots' reads first, then the loads: no wait between chunks)
          int slot[NCH];
#pragma unroll
          for (int h = 0; h < NCH; h++) slot[h] = wpos[min(h * WAVE + lane, last)];
          int cj[NCH];
          float cx[NCH], cy[NCH], cz[NCH];
#pragma unroll
          for (int h = 0; h < NCH; h++) {
            if (h < nch) {
              const int64_t pos = q_boff + slot[h];
              cj[h] = s.bucket[pos];
              const float4 f = s.bf4[pos];
              cx[h] = f.x;
              cy[h] = f.y;
              cz[h] = f.z;
            }
          }
          if (img) {
            // each candidate at its image nearest the centre cell's midpoint
#pragma unroll
            for (int h = 0; h < NCH; h++) {
              if (h < nch) {
                cx[h] = cx[h] - bxf * rintf((cx[h] - mxf) * ibx);
                cy[h] = cy[h] - byf * rintf((cy[h] - myf) * iby);
              }
            }
          }
          // lanes of the last chunk that hold a stream entry
          const int nl = last + 1 - (nch - 1) * WAVE;
          const unsigned long long vlast = nl >= WAVE ? ~0ull : ((1ull << nl) - 1ull);
          // the candidates are in registers before the centre loop starts
          // (its row stores would otherwise be waited for at each chunk)
          __builtin_amdgcn_s_waitcnt(0);
          nn0s[lane] = my_nn;
          // centres with a pair near the list radius: their rows of this
          // super-chunk are written again below, the band decided in FP64
          unsigned long long redo = 0;
          const float rl2f = (float)rl2, bandw = q_bandw;
          auto d2f = [&](auto img_c, float x, float y, float z, float xa, float ya,
                         float za) -> float {
            if (decltype(img_c)::value) {
              const float dx = x - xa, dy = y - ya, dz = z - za;
              return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
            }
            return d2_f32(P, make_float4(x, y, z, 0.f), make_float4(xa, ya, za, 0.f), bxf, byf,
                          ibx, iby);
          };
          // the centre loop, one form per pair-distance arithmetic
          // one centre atom's walk over the chunks: its row count, near
          // count and row address in scalar registers
          struct Ctr {
            int a, i, nn, nc;
            float xa, ya, za, bmin;
            int32_t *row;
          };
          auto start = [&](Ctr &c, int a) {
            c.a = a;
            const int4 v = ctr[a];
            c.i = rfl_i(v.x);
            c.xa = __int_as_float(rfl_i(v.y));
            c.ya = __int_as_float(rfl_i(v.z));
            c.za = __int_as_float(rfl_i(v.w));
            const int pk = __builtin_amdgcn_readlane(my_nn, a);
            c.nn = pk & 0xffffff;
            c.nc = (int)((uint32_t)pk >> 24);
            c.row = s.nbr + (int64_t)c.i * maxnn;
            // the least |d^2 - rl^2| this lane meets (lanes past the
            // stream's end repeat its last entry)
            c.bmin = 3.0e38f;
          };
          auto step = [&](auto img_c, Ctr &c, int h) {
            const unsigned long long vm = h == nch - 1 ? vlast : ~0ull;
            const float df = d2f(img_c, cx[h], cy[h], cz[h], c.xa, c.ya, c.za);
            c.bmin = fminf(c.bmin, fabsf(df - rl2f));
            const unsigned long long lt = __builtin_amdgcn_ballot_w64(df < rl2lo);
            const unsigned long long hm = lt & vm & __builtin_amdgcn_ballot_w64(cj[h] != c.i);
            const int nh = __popcll(hm);
            const uint32_t pr = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(hm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)hm, 0u));
            const unsigned long long ok =
                c.nn + nh <= maxnn ? hm : hm & __builtin_amdgcn_ballot_w64(c.nn + (int)pr < maxnn);
            if (__builtin_amdgcn_inverse_ballot_w64(ok)) c.row[(uint32_t)c.nn + pr] = cj[h];
            const unsigned long long nm = __builtin_amdgcn_ballot_w64(df <= rcf) & hm;
            if (nm) {
              const int qn = c.nc + (int)__builtin_amdgcn_mbcnt_hi(
                                        (uint32_t)(nm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)nm, 0u));
              if (__builtin_amdgcn_inverse_ballot_w64(nm) && qn < NEAR_MAX)
                s.nearl[(int64_t)c.i * NEAR_MAX + qn] = c.nn + (int)pr;
              const int nc2 = c.nc + __popcll(nm);
              c.nc = nc2 < 255 ? nc2 : 255;
            }
            c.nn += nh;
          };
          auto finish = [&](Ctr &c) {
            if (__builtin_amdgcn_ballot_w64(c.bmin <= bandw)) redo |= 1ull << c.a;
            my_nn = lane == c.a ? (c.nn | c.nc << 24) : my_nn;
          };
          // the centre loop, one form per pair-distance arithmetic
          auto centres = [&](auto img_c) {
            unsigned long long m = refm;
            while (m) {
              Ctr A;
              start(A, __builtin_ctzll(m));
              m &= m - 1;
#pragma unroll
              for (int h = 0; h < NCH; h++)
                if (h < nch) step(img_c, A, h);
              finish(A);
            }
          };
          if (img)
            centres(std::true_type());
          else
            centres(std::false_type());
          // the band, in FP64, for the centres that met it: their rows of
          // this super-chunk from its start, candidates read again
          while (redo) {
            const int a = __builtin_ctzll(redo);
            redo &= redo - 1;
            const int4 v = ctr[a];
            const int i = rfl_i(v.x);
            const float xa = __int_as_float(rfl_i(v.y)), ya = __int_as_float(rfl_i(v.z));
            const float za = __int_as_float(rfl_i(v.w));
            const int pk = rfl_i(nn0s[a]);
            int nn = pk & 0xffffff, nc = (int)((uint32_t)pk >> 24);
            int32_t *row = s.nbr + (int64_t)i * maxnn;
            for (int h = 0; h < nch; h++) {
              const int cidx = s0 + h * WAVE + lane;
              const int64_t pos = q_boff + wpos[min(h * WAVE + lane, last)];
              const int j = s.bucket[pos];
              const float4 f = s.bf4[pos];
              float fx = f.x, fy = f.y;
              if (img) {
                fx = f.x - bxf * rintf((f.x - mxf) * ibx);
                fy = f.y - byf * rintf((f.y - myf) * iby);
              }
              const float df = img ? d2f(std::true_type(), fx, fy, f.z, xa, ya, za)
                                   : d2f(std::false_type(), fx, fy, f.z, xa, ya, za);
              const bool valid = cidx < total;
              bool hit = valid && df < rl2lo;
              if (valid && !hit && df <= rl2hi) {
                double ex, ey, ez;
                const uint32_t uj = (uint32_t)j;
                vdist_mic(P, s.px[uj], s.py[uj], s.pz[uj], s.bx[e0 + a], s.by[e0 + a],
                          s.bz[e0 + a], &ex, &ey, &ez);
                const double rd = dot3(ex, ey, ez);
                hit = strict ? (rd < rl2) : !(rd > rl2);
              }
              hit = hit && j != i;
              const unsigned long long hm = __builtin_amdgcn_ballot_w64(hit);
              if (hm) {
                const int nh = __popcll(hm);
                const uint32_t pr = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(hm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)hm, 0u));
                if (hit && nn + (int)pr < maxnn) row[(uint32_t)nn + pr] = j;
                const bool near = hit && df <= rcf;
                const unsigned long long nm = __builtin_amdgcn_ballot_w64(near);
                if (nm) {
                  const int qn = nc + (int)__builtin_amdgcn_mbcnt_hi(
                                          (uint32_t)(nm >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)nm, 0u));
                  if (near && qn < NEAR_MAX) s.nearl[(int64_t)i * NEAR_MAX + qn] = nn + (int)pr;
                  const int nc2 = nc + __popcll(nm);
                  nc = nc2 < 255 ? nc2 : 255;
                }
                nn += nh;
              }
            }
            my_nn = lane == a ? (nn | nc << 24) : my_nn;
          }
        }
        const int my_nc = (int)((uint32_t)my_nn >> 24);
        my_nn &= 0xffffff;
        const bool mine = (refm >> lane) & 1ull;
        if (mine) {
          const int my_i = ctr[lane].x;
          int nn = my_nn;
          if (nn > maxnn) {
            atomicOr(s.err, ERR_LIST);
            atomicMax(&s.rp[r].max_nn_seen, nn);
            rep_fail(s, r, ERR_LIST);
            nn = maxnn;
          }
          s.nn[my_i] = nn;
          s.rowflag[my_i] = (uint8_t)ghost;
          s.ov_near[my_i] = (uint8_t)(img && my_nc <= NEAR_MAX ? my_nc : 255);
          if (ghost) s.rp[r].any_ghost = 1;
        }
        const int v = wave_scan_incl(mine ? my_nn : 0);
        acc_nn += (unsigned long long)(uint32_t)__builtin_amdgcn_readlane(v, WAVE - 1);
        acc_rows += (unsigned long long)__popcll(refm);
      }
    }
  }
  if (cur_r >= 0 && lane == 0 && acc_rows) {
    atomicAdd(&s.cnt[cur_r].nn_sum, acc_nn);
    atomicAdd(&s.cnt[cur_r].nn_rows, acc_rows);
  }
}

template __global__ void k_listbuild_wr<3, 7>(DevState s);
template __global__ void k_listbuild_wr<2, 8>(DevState s);


// bucket order by a rank sort inside one wave (cells of up to 256 atoms,
// four per lane: electrode cells hold ~100; larger cells fall back to an
// insertion sort by lane 0)
#define SORT_W 4
__global__ void __launch_bounds__(DML_BLOCK) k_cell_sort_w(DevState s, int gcs) {
  const int lane = threadIdx.x % WAVE;
  // each wave one contiguous run of the cell list, its metadata read 64
  // cells at a time (one per lane), then the non-empty cells one by one
  const int64_t wid = (int64_t)blockIdx.x * (DML_BLOCK / WAVE) + threadIdx.x / WAVE;
  const int64_t nwv = (int64_t)gridDim.x * (DML_BLOCK / WAVE);
  const int64_t run = (s.n_cm + nwv - 1) / nwv;
  const int64_t g_end = min(s.n_cm, (wid + 1) * run);
  for (int64_t g0 = wid * run; g0 < g_end; g0 += WAVE) {
    int lr = 0, lc = 0, lcnt = 0;
    if (g0 + lane < g_end) {
      lr = s.cm_rep[g0 + lane];
      lc = s.cm_cell[g0 + lane];
      const RepParams &Q = s.rp[lr];
      if (!Q.halted && want_cells(Q, gcs) && lc < ncell_total(Q))
        lcnt = min(s.cell_cnt[Q.coff + lc], Q.bcap);
    }
    unsigned long long work = __ballot(lcnt > 0);
    while (work) {
    const int q = __builtin_ctzll(work);
    work &= work - 1;
    const int cnt = __builtin_amdgcn_readlane(lcnt, q);
    if (cnt <= WAVE / 2 && work &&
        __builtin_amdgcn_readlane(lcnt, __builtin_ctzll(work)) <= WAVE / 2) {
      // two cells of up to 32 entries at once, one per half-wave: their
      // loads are in flight together (a cell costs a round trip to memory)
      const int q2 = __builtin_ctzll(work);
      work &= work - 1;
      const int half = lane >> 5, sub = lane & 31;
      const int cnt2 = __builtin_amdgcn_readlane(lcnt, q2);
      const int rq = half ? __builtin_amdgcn_readlane(lr, q2) : __builtin_amdgcn_readlane(lr, q);
      const int cq = half ? __builtin_amdgcn_readlane(lc, q2) : __builtin_amdgcn_readlane(lc, q);
      const int mycnt = half ? cnt2 : cnt;
      const RepParams &H = s.rp[rq];
      const int hdesc = H.use_cells;
      const int64_t hbase = H.boff + (int64_t)cq * H.bcap;
      const bool in = sub < mycnt;
      const int v = in ? s.bucket[hbase + sub] : 0;
      int k = 0;
      double x = 0.0, y = 0.0, z = 0.0;
      float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
      if (in) {
        k = s.bkey[v];
        x = s.px[v];
        y = s.py[v];
        z = s.pz[v];
        f = bucket_f4(s, v);
      }
      int rank = 0;
      const int cmax = cnt > cnt2 ? cnt : cnt2;
      for (int j = 0; j < cmax; j++) {  // j is wave-uniform
        const int ka = __builtin_amdgcn_readlane(k, j);
        const int kb = __builtin_amdgcn_readlane(k, j + WAVE / 2);
        const int kq = half ? kb : ka;
        rank += (j < mycnt) & (hdesc ? (kq > k) : (kq < k));
      }
      if (in) {
        s.bucket[hbase + rank] = v;
        s.bx[hbase + rank] = x;
        s.by[hbase + rank] = y;
        s.bz[hbase + rank] = z;
        s.bf4[hbase + rank] = f;
      }
      continue;
    }
    const RepParams &P = s.rp[__builtin_amdgcn_readlane(lr, q)];
    const int c = __builtin_amdgcn_readlane(lc, q);
    const int desc = P.use_cells;
    const int64_t base = P.boff + (int64_t)c * P.bcap;
    int32_t *b = s.bucket + base;
    if (cnt <= WAVE) {
      // one entry per lane; its key and position loads are issued together,
      // before the ranks are known
      const bool in = lane < cnt;
      const int v = in ? b[lane] : 0;
      int k = 0;
      double x = 0.0, y = 0.0, z = 0.0;
      float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
      if (in) {
        k = s.bkey[v];
        x = s.px[v];
        y = s.py[v];
        z = s.pz[v];
        f = bucket_f4(s, v);
      }
      int rank = 0;
      for (int q = 0; q < cnt; q++) {
        const int kq = __builtin_amdgcn_readlane(k, q);  // q is wave-uniform
        rank += desc ? (kq > k) : (kq < k);
      }
      if (in) {
        b[rank] = v;
        s.bx[base + rank] = x;
        s.by[base + rank] = y;
        s.bz[base + rank] = z;
        s.bf4[base + rank] = f;
      }
    } else if (cnt <= SORT_W * WAVE) {
      // up to SORT_W entries per lane; an entry's place is the number of
      // keys before it (the hs%b keys of one bucket are distinct)
      int v[SORT_W], k[SORT_W], rank[SORT_W];
#pragma unroll
      for (int j = 0; j < SORT_W; j++) {
        const int q = j * WAVE + lane;
        v[j] = q < cnt ? b[q] : 0;
        k[j] = q < cnt ? s.bkey[v[j]] : 0;
        rank[j] = 0;
      }
#pragma unroll
      for (int j2 = 0; j2 < SORT_W; j2++) {
        const int lim = min(WAVE, cnt - j2 * WAVE);
        for (int q = 0; q < lim; q++) {
          const int kq = __builtin_amdgcn_readlane(k[j2], q);  // q is wave-uniform
#pragma unroll
          for (int j = 0; j < SORT_W; j++) rank[j] += desc ? (kq > k[j]) : (kq < k[j]);
        }
      }
#pragma unroll
      for (int j = 0; j < SORT_W; j++) {
        if (j * WAVE + lane < cnt) {
          const int vj = v[j], rj = rank[j];
          b[rj] = vj;
          s.bx[base + rj] = s.px[vj];
          s.by[base + rj] = s.py[vj];
          s.bz[base + rj] = s.pz[vj];
          s.bf4[base + rj] = bucket_f4(s, vj);
        }
      }
    } else {
      if (lane == 0) {
        for (int q = 1; q < cnt; q++) {
          int v = b[q], kv = s.bkey[v];
          int j = q - 1;
          while (j >= 0) {
            int kw = s.bkey[b[j]];
            if (desc ? (kw >= kv) : (kw <= kv)) break;
            b[j + 1] = b[j];
            j--;
          }
          b[j + 1] = v;
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      for (int q = lane; q < cnt; q += WAVE) {
        int v = __hip_atomic_load(&b[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        s.bx[base + q] = s.px[v];
        s.by[base + q] = s.py[v];
        s.bz[base + q] = s.pz[v];
        s.bf4[base + q] = bucket_f4(s, v);
      }
    }
    }
  }
}



// ------------------------------------------------------------------------
// fuerza (dana.F90:1055-1139) as a gather: the full list plus reaction of
// the reference counts every ref-ref pair twice, so entry j of row i
// weighs 2 when j is a ref atom (its row holds i) and 1 otherwise.  Rows
// built from a ghost-layer cell are found by nobody; their reactions are
// added by k_ghost_scatter.  Optionally fused with ermak_b (1031-1052).
// ------------------------------------------------------------------------
__device__ inline double pow6d(double x) {
  double x2 = x * x, x4 = x2 * x2;
  return x2 * x4;
}
__device__ inline double pow7d(double x) {
  double x2 = x * x, x4 = x2 * x2;
  return (x * x2) * x4;
}
__device__ inline float pow6f(float x) {
  float x2 = x * x, x4 = x2 * x2;
  return x2 * x4;
}

__device__ inline void ermak_b_one(const DevState &s, const RepParams &P, int i,
                                   double fx, double fy, double fz) {
  if (s.species[i] == SP_CG) return;
  double m = P.mass[s.species[i] - 1];
  double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
  double ax = s.ax[i], ay = s.ay[i], az = s.az[i];
  s.vx[i] = P.cc0 * vx + (P.cc1 - P.cc2) * ax + P.cc2 * fx / m + s.rvx[i];
  s.vy[i] = P.cc0 * vy + (P.cc1 - P.cc2) * ay + P.cc2 * fy / m + s.rvy[i];
  s.vz[i] = P.cc0 * vz + (P.cc1 - P.cc2) * az + P.cc2 * fz / m + s.rvz[i];
  s.ax[i] = fx / m;
  s.ay[i] = fy / m;
  s.az[i] = fz / m;
}

// v of the lane `off` away within groups of LPA lanes (off < LPA): DPP
// quad permutes for 1 and 2, a shuffle beyond (a shuffle's lane addresses,
// hoisted by the compiler, would hold registers through the kernel)
template <int LPA>
__device__ inline int lpa_xor(int v, int off) {
  if (off == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xf, 0xf, false);  // [1,0,3,2]
  if (off == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xf, 0xf, false);  // [2,3,0,1]
  return __shfl_xor(v, off, LPA);
}
template <int LPA>
__device__ inline double lpa_xor_d(double v, int off) {
  const int64_t b = __double_as_longlong(v);
  const int lo = lpa_xor<LPA>((int)(b & 0xffffffff), off), hi = lpa_xor<LPA>((int)(b >> 32), off);
  return __longlong_as_double((int64_t)(uint32_t)lo | ((int64_t)hi << 32));
}

template <int LPA, int FUSEB>
__global__ void __launch_bounds__(DML_BLOCK) k_force(DevState s, int only_ermak) {
  int gid = blockIdx.x * DML_BLOCK + threadIdx.x;
  int t = gid / LPA, sub = gid % LPA;
  int r = s.blk2rep[t / DML_BLOCK];
  const RepParams &P = s.rp[r];
  if (P.halted) return;
  if (only_ermak && P.integrator != DML_ERMAK) return;
  const int FP32 = P.precision == DML_FP32_FORCE;
  int local = t - P.poff;
  int active = local < P.nref;
  int i = active ? s.ref_order[t] : 0;
  double fx = 0.0, fy = 0.0, fz = 0.0, ea = 0.0;
  // entries within rcut (FP32 with margin), the ones overlap_moveback looks
  // at: their number on this lane and the row position of the last one
  int ncnt = 0, ne = 0;
  if (local == 0 && sub == 0) s.rp[r].ovn_stamp = P.step + 1;
  if (active) {
    int k = s.species[i];
    const float rcf = (float)P.rcut2 + P.pf_margin;
    double xi = s.px[i], yi = s.py[i], zi = s.pz[i];
    double hbx = P.box[0] * .5, hby = P.box[1] * .5;
    int ghost = s.rowflag[i] & 1;
    int nn = s.nn[i];
    const int32_t *row = s.nbr + (int64_t)i * s.maxnn;
    const float4 fi = s.pf4[i];
    const float bxf = (float)P.box[0], byf = (float)P.box[1];
    const float ibx = (float)P.one_box[0], iby = (float)P.one_box[1];
    // this atom's FP32 prefilter bound r0^2 + margin against each of the
    // three species (held in registers: no table load per entry)
    const float th1 = (float)P.r0[(k - 1) * 3 + 0] * (float)P.r0[(k - 1) * 3 + 0] + P.pf_margin;
    const float th2 = (float)P.r0[(k - 1) * 3 + 1] * (float)P.r0[(k - 1) * 3 + 1] + P.pf_margin;
    const float th3 = (float)P.r0[(k - 1) * 3 + 2] * (float)P.r0[(k - 1) * 3 + 2] + P.pf_margin;
    for (int e = sub; e < nn; e += LPA) {
      int j = row[e];
      const float4 fj4 = s.pf4[j];
      {  // surely beyond the pair's r0: no force, no energy
        const int sj = __float_as_int(fj4.w);
        const float th = sj == 1 ? th1 : (sj == 2 ? th2 : th3);
        const float d2f = d2_f32(P, fi, fj4, bxf, byf, ibx, iby);
        if (d2f <= rcf) {
          ncnt++;
          ne = e;
        }
        if (d2f > th) continue;
      }
      uint8_t fj = s.flags[j];
      if (fj & DML_FL_LIMBO) continue;
      int m = s.species[j];
      double vx = xi - s.px[j], vy = yi - s.py[j], vz = zi - s.pz[j];
      if (vx > hbx) vx = vx - P.box[0];
      else if (vx < -hbx) vx = vx + P.box[0];
      if (vy > hby) vy = vy - P.box[1];
      else if (vy < -hby) vy = vy + P.box[1];
      if (m == SP_CG && k == SP_CG) continue;
      double dr2 = dot3(vx, vy, vz);
      int km = (k - 1) * 3 + (m - 1);
      double r0 = P.r0[km];
      if (dr2 > r0 * r0) continue;
      double eps = P.eps[km];
      double w = ((fj & DML_FL_REF) && !ghost) ? 2.0 : 1.0;
      if (FP32) {
        // FP32 pair arithmetic without cancellation: r0^2 - r^2 comes from
        // the FP64 cutoff test, then b - 1 = (r0^6 - r^6)/r^6 and
        // u = eps (b - 1)^2 are formed in FP32; sums stay FP64.
        float q = (float)(r0 * r0 - dr2);
        float s2 = (float)dr2, a2 = (float)(r0 * r0);
        float s6 = s2 * s2 * s2;
        float bm1 = q * (a2 * a2 + a2 * s2 + s2 * s2) / s6;
        float b = 1.0f + bm1;
        float epsf = (float)eps;
        float aux = 12.0f * epsf * b * bm1 / s2;  // |F| / r
        fx += w * (double)(aux * (float)vx);
        fy += w * (double)(aux * (float)vy);
        fz += w * (double)(aux * (float)vz);
        ea += w * (double)(0.5f * epsf * bm1 * bm1);
      } else {
        double dr = sqrt(dr2);
        double b = pow6d(r0);
        double c = eps * 12.0 * b;
        c = c / pow7d(dr);
        b = b / pow6d(dr);
        double aux = c * (b - 1);
        fx += w * (aux * vx / dr);
        fy += w * (aux * vy / dr);
        fz += w * (aux * vz / dr);
        double u = eps * b * (b - 2);
        u = u + eps;
        ea += w * (u * .5);
      }
    }
  }
  // the atom's list: one row position per lane that met exactly one entry
  // (any order; none listed when a lane met more, or more than NEAR_MAX)
  const unsigned long long one = __ballot(ncnt == 1);
  int tot = ncnt, ovf = ncnt > 1;
  for (int off = LPA / 2; off > 0; off >>= 1) {
    fx += lpa_xor_d<LPA>(fx, off);
    fy += lpa_xor_d<LPA>(fy, off);
    fz += lpa_xor_d<LPA>(fz, off);
    ea += lpa_xor_d<LPA>(ea, off);
    tot += lpa_xor<LPA>(tot, off);
    ovf |= lpa_xor<LPA>(ovf, off);
  }
  const int listed = !ovf && tot <= NEAR_MAX;
  if (active && listed && ncnt == 1) {
    const int g0 = (threadIdx.x % WAVE) & ~(LPA - 1);  // first lane of this atom
    const unsigned long long gm = LPA == WAVE ? one : (one >> g0) & ((1ull << (LPA & 63)) - 1ull);
    s.nearl[(int64_t)i * NEAR_MAX + __popcll(gm & ((1ull << sub) - 1ull))] = ne;
  }
  if (active && sub == 0) {
    s.ov_near[i] = (uint8_t)(listed ? tot : 255);
    s.fx[i] = fx;
    s.fy[i] = fy;
    s.fz[i] = fz;
    s.epot[i] = ea;
    if (FUSEB && !P.any_ghost) ermak_b_one(s, P, i, fx, fy, fz);
  }
}

#define INST_FORCE(L)                                              \
  template __global__ void k_force<L, 0>(DevState s, int only_ermak); \
  template __global__ void k_force<L, 1>(DevState s, int only_ermak);
INST_FORCE(1)
INST_FORCE(2)
INST_FORCE(4)
INST_FORCE(8)
INST_FORCE(16)
INST_FORCE(32)
INST_FORCE(64)

// reactions of rows built from ghost-layer cells (nobody lists those atoms)
__global__ void __launch_bounds__(DML_BLOCK) k_ghost_scatter(DevState s, int only_ermak) {
  int t = blockIdx.x * DML_BLOCK + threadIdx.x;
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  if (!P.any_ghost) return;
  if (only_ermak && P.integrator != DML_ERMAK) return;
  if (t - P.poff >= P.nref) return;
  int i = s.ref_order[t];
  if (!(s.rowflag[i] & 1)) return;
  int k = s.species[i];
  double xi = s.px[i], yi = s.py[i], zi = s.pz[i];
  double hbx = P.box[0] * .5, hby = P.box[1] * .5;
  const int32_t *row = s.nbr + (int64_t)i * s.maxnn;
  for (int e = 0; e < s.nn[i]; e++) {
    int j = row[e];
    uint8_t fj = s.flags[j];
    if ((fj & DML_FL_LIMBO) || !(fj & DML_FL_REF)) continue;
    int m = s.species[j];
    double vx = xi - s.px[j], vy = yi - s.py[j], vz = zi - s.pz[j];
    if (vx > hbx) vx = vx - P.box[0];
    else if (vx < -hbx) vx = vx + P.box[0];
    if (vy > hby) vy = vy - P.box[1];
    else if (vy < -hby) vy = vy + P.box[1];
    if (m == SP_CG && k == SP_CG) continue;
    double dr2 = dot3(vx, vy, vz);
    int km = (k - 1) * 3 + (m - 1);
    double r0 = P.r0[km];
    if (dr2 > r0 * r0) continue;
    double eps = P.eps[km];
    double dr = sqrt(dr2);
    double b = pow6d(r0);
    double c = eps * 12.0 * b;
    c = c / pow7d(dr);
    b = b / pow6d(dr);
    double aux = c * (b - 1);
    atomicAdd(&s.fx[j], -(aux * vx / dr));
    atomicAdd(&s.fy[j], -(aux * vy / dr));
    atomicAdd(&s.fz[j], -(aux * vz / dr));
    double u = eps * b * (b - 2);
    u = u + eps;
    atomicAdd(&s.epot[j], u * .5);
  }
}

// ermak_b after the ghost reactions (only when some row is a ghost row or
// when called on its own)
__global__ void __launch_bounds__(DML_BLOCK) k_ermak_b(DevState s, int only_if_ghost) {
  int t = blockIdx.x * DML_BLOCK + threadIdx.x;
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  if (P.integrator != DML_ERMAK) return;
  if (only_if_ghost && !P.any_ghost) return;
  if (t - P.poff >= P.nref) return;
  int i = s.ref_order[t];
  ermak_b_one(s, P, i, s.fx[i], s.fy[i], s.fz[i]);
}

// ------------------------------------------------------------------------
// integrators: cbrownian_hs (dana.F90:798-846), ermak_a (974-1028), both
// with atom_pbc (1187-1250)
// ------------------------------------------------------------------------
__device__ inline int atom_pbc(const DevState &s, const RepParams &P, int r,
                               int i, int64_t t, double &x, double &y, double &z,
                               const double *udep) {
  int depos = 0;
  double bx = P.box[0], by = P.box[1];
  if (x > bx) { x = x - bx; s.ox[i] = s.ox[i] - bx; }
  if (x < 0.0) { x = x + bx; s.ox[i] = s.ox[i] + bx; }
  if (y > by) { y = y - by; s.oy[i] = s.oy[i] - by; }
  if (y < 0.0) { y = y + by; s.oy[i] = s.oy[i] + by; }
  if (z > P.zmax) {
    if (P.integrator == DML_ERMAK) {
      z = z - 2.0 * (z - P.zmax);
      s.vz[i] = -s.vz[i];
    } else {
      x = s.cx[i]; y = s.cy[i]; z = s.cz[i];
    }
  }
  if (z <= 0.0) {
    atomicAdd(&s.cnt[r].try_, 1ull);
    double ne;
    if (udep) {
      ne = udep[t];
    } else {
      u32x4 w = dml_philox(P.seed, s.pid[i], (uint32_t)P.step, RS_DEPO, 0);
      ne = u01_from(w.x, w.y);
    }
    if (ne < P.prob) {
      atomicAdd(&s.cnt[r].depo, 1ull);
      s.species[i] = SP_F;
      s.rp[r].any_f = 1;
      depos = 1;
    }
    x = s.cx[i]; y = s.cy[i]; z = s.cz[i];
  }
  return depos;
}

__global__ void __launch_bounds__(DML_BLOCK) k_brownian(DevState s, const double *noise,
                                                         const double *udep) {
  int t = blockIdx.x * DML_BLOCK + threadIdx.x;
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  if (P.integrator != DML_BROWNIAN) return;
  if (t == P.poff) s.rp[r].ovn_stamp = 0;  // positions move
  if (t - P.poff >= P.nref) return;
  int i = s.ref_order[t];
  double x = s.px[i], y = s.py[i], z = s.pz[i];
  s.cx[i] = x; s.cy[i] = y; s.cz[i] = z;
  double dif = z > P.z_switch ? P.dif_sc : P.dif_sei;
  double fac1 = sqrt(2.0 * dif * P.h);
  double g0, g1, g2;
  if (noise) {
    g0 = noise[3 * t]; g1 = noise[3 * t + 1]; g2 = noise[3 * t + 2];
  } else {
    double g3;
    gauss2(dml_philox(P.seed, s.pid[i], (uint32_t)P.step, RS_BROWN, 0), &g0, &g1);
    gauss2(dml_philox(P.seed, s.pid[i], (uint32_t)P.step, RS_BROWN, 1), &g2, &g3);
  }
  double nx = x + g0 * fac1, ny = y + g1 * fac1, nz = z + g2 * fac1;
  s.vx[i] = (nx - x) / P.h;
  s.vy[i] = (ny - y) / P.h;
  s.vz[i] = (nz - z) / P.h;
  int depos = atom_pbc(s, P, r, i, t, nx, ny, nz, udep);
  s.px[i] = nx; s.py[i] = ny; s.pz[i] = nz;
  put_pf4(s, i, nx, ny, nz, s.species[i]);
  if (!depos) s.flags[i] &= (uint8_t)~DML_FL_SKIP;
}

__global__ void __launch_bounds__(DML_BLOCK) k_ermak_a(DevState s, const double *noise,
                                                        const double *udep) {
  int t = blockIdx.x * DML_BLOCK + threadIdx.x;
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  if (P.integrator != DML_ERMAK) return;
  if (t == P.poff) s.rp[r].ovn_stamp = 0;  // positions move
  if (t - P.poff >= P.nref) return;
  int i = s.ref_order[t];
  double x = s.px[i], y = s.py[i], z = s.pz[i];
  s.cx[i] = x; s.cy[i] = y; s.cz[i] = z;
  double g[6];
  if (noise) {
    for (int q = 0; q < 6; q++) g[q] = noise[6 * t + q];
  } else {
    for (int q = 0; q < 3; q++)
      gauss2(dml_philox(P.seed, s.pid[i], (uint32_t)P.step, RS_ERMAK, q), &g[2 * q], &g[2 * q + 1]);
  }
  double m = P.mass[s.species[i] - 1];
  double sq = P.skt / sqrt(m);
  double np[3] = {x, y, z};
  double rv[3];
  const double vel[3] = {s.vx[i], s.vy[i], s.vz[i]};
  const double acc[3] = {s.ax[i], s.ay[i], s.az[i]};
  for (int k = 0; k < 3; k++) {
    double r1 = g[2 * k], r2 = g[2 * k + 1];
    double ranr = sq * P.sdr * r1;
    np[k] = np[k] + P.cc1 * vel[k] + P.cc2 * P.h * acc[k] + ranr;
    rv[k] = sq * P.sdv * (P.crv1 * r1 + P.crv2 * r2);
  }
  s.rvx[i] = rv[0]; s.rvy[i] = rv[1]; s.rvz[i] = rv[2];
  int depos = atom_pbc(s, P, r, i, t, np[0], np[1], np[2], udep);
  s.px[i] = np[0]; s.py[i] = np[1]; s.pz[i] = np[2];
  put_pf4(s, i, np[0], np[1], np[2], s.species[i]);
  if (!depos) s.flags[i] &= (uint8_t)~DML_FL_SKIP;
}

// ------------------------------------------------------------------------
// overlap_moveback (dana.F90:849-943)
// ------------------------------------------------------------------------

// exact reference order: one wave per replica walks the ref list in order;
// rows are scanned 64 entries at a time, hits acted on in list order and
// re-checked against the current positions (an earlier revert may have
// moved a repeated entry).
__global__ void __launch_bounds__(WAVE) k_ov_exact(DevState s, const double *u,
                                                    int64_t n_u, int max_passes) {
  int r = blockIdx.x;
  if (r >= s.nrep) return;
  const RepParams &P = s.rp[r];
  if (P.halted) return;
  if (P.overlap_mode != DML_OVERLAP_EXACT) return;
  int lane = threadIdx.x;
  const double *ur = u ? u + (int64_t)r * n_u : nullptr;
  int64_t ucount = 0;
  unsigned long long n_try = 0, n_depo = 0, n_ch = 0, n_ch3 = 0, n_pass = 0;
  int nref = P.nref;
  int again = 1;
  for (int pass = 0; again; pass++) {
    if (pass >= max_passes) {
      if (lane == 0) rep_fail(s, r, ERR_OVERLAP);
      break;
    }
    again = 0;
    n_pass++;
    for (int t0 = 0; t0 < nref; t0 += WAVE) {
      int tl = t0 + lane;
      int myi = tl < nref ? s.ref_order[P.poff + tl] : -1;
      unsigned long long mask =
          __ballot(myi >= 0 && !(s.flags[myi >= 0 ? myi : 0] & DML_FL_SKIP));
      while (mask) {
        int q = __ffsll((long long)mask) - 1;
        int ii = __shfl(myi, q);
        int reverted = 0;
        s.flags[ii] |= DML_FL_SKIP;
        double xi = s.px[ii], yi = s.py[ii], zi = s.pz[ii];
        int nn = s.nn[ii];
        const int32_t *row = s.nbr + (int64_t)ii * s.maxnn;
        int done = 0;
        for (int e0 = 0; e0 < nn && !done; e0 += WAVE) {
          int e = e0 + lane;
          int j = e < nn ? row[e] : -1;
          int hit = 0;
          if (j >= 0 && !(s.flags[j] & DML_FL_LIMBO)) {
            double dx, dy, dz;
            vdist_mic(P, xi, yi, zi, s.px[j], s.py[j], s.pz[j], &dx, &dy, &dz);
            hit = !(dot3(dx, dy, dz) > P.rcut2);
          }
          unsigned long long hm = __ballot(hit);
          while (hm) {
            int q2 = __ffsll((long long)hm) - 1;
            hm &= hm - 1ull;
            int jj = __shfl(j, q2);
            double dx, dy, dz;
            vdist_mic(P, xi, yi, zi, s.px[jj], s.py[jj], s.pz[jj], &dx, &dy, &dz);
            // every lane takes the same decisions from the same reads, then
            // writes the same values: the reads come before any write
            const bool near = !(dot3(dx, dy, dz) > P.rcut2);
            const bool cg = near && s.species[jj] == SP_CG;
            const bool stay = near && !cg && P.piston && same_pos(s, jj) && same_pos(s, ii);
            __builtin_amdgcn_wave_barrier();
            if (!near) continue;
            if (cg) {
              n_try++;
              double ne;
              if (ur) {
                ne = ucount < n_u ? ur[ucount] : 0.5;
                if (ucount >= n_u && lane == 0) rep_fail(s, r, ERR_OVERLAP);
              } else {
                u32x4 w = dml_philox(P.seed, s.pid[ii], (uint32_t)P.step, RS_CONTACT,
                                     (uint32_t)ucount);
                ne = u01_from(w.x, w.y);
              }
              ucount++;
              if (ne < P.prob) {
                n_depo++;
                s.species[ii] = SP_F;
                put_pf4_sp(s, ii, SP_F);
                if (lane == 0) s.rp[r].any_f = 1;
                if (zi > P.z0 && lane == 0) rep_fail(s, r, ERR_DEPZ0);
              } else {
                s.px[ii] = s.cx[ii]; s.py[ii] = s.cy[ii]; s.pz[ii] = s.cz[ii];
                put_pf4(s, ii, s.cx[ii], s.cy[ii], s.cz[ii], s.species[ii]);
                s.flags[ii] &= (uint8_t)~DML_FL_SKIP;
                reverted = 1;
              }
              done = 1;
              break;
            }
            if (stay) {
              n_ch3++;
              continue;
            }
            s.px[jj] = s.cx[jj]; s.py[jj] = s.cy[jj]; s.pz[jj] = s.cz[jj];
            put_pf4(s, jj, s.cx[jj], s.cy[jj], s.cz[jj], s.species[jj]);
            s.vx[jj] = 0.0; s.vy[jj] = 0.0; s.vz[jj] = 0.0;
            s.ax[jj] = 0.0; s.ay[jj] = 0.0; s.az[jj] = 0.0;
            s.flags[jj] &= (uint8_t)~DML_FL_SKIP;
            n_ch++;
            again = 1;
            reverted = 1;
          }
        }
        unsigned long long above = (q == 63) ? 0ull : (~0ull << (q + 1));
        if (reverted) {
          int fl = myi >= 0 ? s.flags[myi] : DML_FL_SKIP;
          mask = __ballot(myi >= 0 && !(fl & DML_FL_SKIP)) & above;
        } else {
          mask &= above;
        }
      }
    }
  }
  if (lane == 0) {
    atomicAdd(&s.cnt[r].try_, n_try);
    atomicAdd(&s.cnt[r].depo, n_depo);
    atomicAdd(&s.cnt[r].choques, n_ch);
    atomicAdd(&s.cnt[r].choques3, n_ch3);
    atomicAdd(&s.cnt[r].passes, n_pass);
  }
}

// deterministic parallel rounds (this build's schedule; orc.c
// orc_overlap_rounds is its specification).  Round 0 covers every ref
// position; round r > 0 covers the worklist of atoms left with skip = 0 by
// round r-1, plus the atoms reverted in round r.  Worklist entries are global
// ref positions, deduplicated by a per-slot round stamp.
__device__ inline int ov_stamp(const RepParams &P, int round) { return P.ov_round + round + 1; }

__global__ void k_ov_begin(DevState s) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = q; k < 3 * s.nrep; k += gridDim.x * blockDim.x) s.ov_more[k] = 0;
  if (q == 0) s.wl_cnt[2] = 0;  // round 0's list of reverted atoms outside ref
}

// one round's scan as workgroup bid of nbk (k_ov_scan_l: the whole grid;
// k_ov_tail: one workgroup)
template <int LPA>
__device__ __forceinline__ void ov_scan_body(const DevState &s, const double *u_round,
                                             int64_t u_stride, int round, int64_t ptotal,
                                             int bid, int nbk) {
  // progress flags rotate over three buffers: this round writes round % 3,
  // is gated by (round - 1) % 3, and clears (round + 1) % 3 for the next one
  const int cur = round % 3, prv = (round + 2) % 3, nxt = (round + 1) % 3;
  if (bid == 0) {
    for (int q = threadIdx.x; q < s.nrep; q += DML_BLOCK) s.ov_more[nxt * s.nrep + q] = 0;
    if (threadIdx.x == 0) s.wl_cnt[(round + 1) % 2] = 0;  // the next round's list
  }
  // round 0: every ref position; later rounds: the positions left unresolved
  // (lists hold at most ptotal entries, each a position below ptotal)
  const int64_t nitems =
      round == 0 ? ptotal : min((int64_t)s.wl_cnt[round % 2], ptotal);
  const int32_t *wl = s.ovw[round % 2];
  const int gtid = bid * DML_BLOCK + threadIdx.x;
  const int sub = gtid % LPA;
  const int64_t stride = (int64_t)nbk * (DML_BLOCK / LPA);
  // one ref position: its verdict, and what it adds to its replica's
  // counters (k_r = the replica, -1 when the position is not scanned)
  auto scan_one = [&](int64_t item, int &k_try, int &k_ch, int &k_ch3, int &k_more, int &k_r) {
    const int t = round == 0 ? (int)item : wl[item];
    if ((uint32_t)t >= (uint64_t)ptotal) return;
    const int r = s.blk2rep[t / DML_BLOCK];
    const RepParams &P = s.rp[r];
    if (P.halted || P.overlap_mode != DML_OVERLAP_ROUNDS) return;
    // the sequential pass recurses only after a revert or a deferred atom
    if (round > 0 && !s.ov_more[prv * s.nrep + r]) return;
    const int local = t - P.poff;
    if (local >= P.nref) return;
    const int i = s.ref_order[t];
    if (s.flags[i] & DML_FL_SKIP) {
      if (sub == 0) s.ov_st[i] = 0;
      return;
    }
    // fuerza or the row build of this step, at these same positions and on
    // this same row, listed the row positions of the entries within rcut
    // (FP32 with margin): every other entry lies beyond rcut.  None listed:
    // no contact, no neighbour to revert and no earlier atom to wait for
    int nlist = round == 0 && P.ovn_stamp == P.step + 1 ? s.ov_near[i] : 255;
    if (nlist == 0) {
      if (sub == 0) s.ov_st[i] = 1;
      return;
    }
    const double xi = s.px[i], yi = s.py[i], zi = s.pz[i];
    const float4 fi = s.pf4[i];
    const float bxf = (float)P.box[0], byf = (float)P.box[1];
    const float ibx = (float)P.one_box[0], iby = (float)P.one_box[1];
    const float rcf = (float)P.rcut2 + P.pf_margin;
    const int nn = s.nn[i];
    const int32_t *row = s.nbr + (int64_t)i * s.maxnn;
    // up to four non-CG hits per lane kept in registers (named, so that no
    // private array is needed), more fall back to a second pass
    int hj0 = 0, hj1 = 0, hj2 = 0, hj3 = 0, he0 = 0, he1 = 0, he2 = 0, he3 = 0;
    int nh = 0, over = 0, dom = 0, firstcg = 0x7fffffff;
    // row entry e (neighbour j) within rcut in FP64
    auto visit = [&](int e, int j) {
      uint8_t fj = s.flags[j];
      if (fj & DML_FL_LIMBO) return;
      double dx, dy, dz;
      vdist_mic(P, xi, yi, zi, s.px[j], s.py[j], s.pz[j], &dx, &dy, &dz);
      if (dot3(dx, dy, dz) > P.rcut2) return;
      if ((fj & DML_FL_REF) && !(fj & DML_FL_SKIP)) {
        const int pj = s.ref_pos[j];
        if (pj >= 0 && pj < local) dom = 1;
      }
      if (s.species[j] == SP_CG) {
        if (e < firstcg) firstcg = e;
      } else if (nh == 0) {
        hj0 = j; he0 = e; nh = 1;
      } else if (nh == 1) {
        hj1 = j; he1 = e; nh = 2;
      } else if (nh == 2) {
        hj2 = j; he2 = e; nh = 3;
      } else if (nh == 3) {
        hj3 = j; he3 = e; nh = 4;
      } else {
        over = 1;
      }
    };
    // a listed position outside this row (a list the row no longer matches)
    // sends the atom through the whole row instead
    const int32_t *nl = s.nearl + (int64_t)i * NEAR_MAX;
    if (nlist <= NEAR_MAX) {
      int bad = 0;
      for (int k = sub; k < nlist; k += LPA) bad |= (unsigned)nl[k] >= (unsigned)nn;
      for (int off = LPA / 2; off > 0; off >>= 1) bad |= lpa_xor<LPA>(bad, off);
      if (bad) nlist = 255;
    }
    if (nlist <= NEAR_MAX) {
      for (int k = sub; k < nlist; k += LPA) {
        const int e = nl[k];
        visit(e, row[e]);
      }
    } else {
      for (int e = sub; e < nn; e += LPA) {
        int j = row[e];
        if (d2_f32(P, fi, s.pf4[j], bxf, byf, ibx, iby) > rcf) continue;  // surely no hit
        visit(e, j);
      }
    }
    for (int off = LPA / 2; off > 0; off >>= 1) {
      dom |= lpa_xor<LPA>(dom, off);
      over |= lpa_xor<LPA>(over, off);
      const int o = lpa_xor<LPA>(firstcg, off);
      firstcg = o < firstcg ? o : firstcg;
    }
    int n_ch = 0, n_ch3 = 0;  // reverts of this atom: at most its row length
    const int stamp = ov_stamp(P, round);
    // a revert of neighbour j (or, piston, the "unsolvable" count when
    // both sit where they were, dana.F90:920-927); returns 1 for the latter.
    // Counted by the callers with arithmetic, keeping the counters in registers
    auto revert = [&](int j) -> int {
      if (P.piston && same_pos(s, j) && same_pos(s, i)) return 1;
      if (round == 0 && (s.flags[j] & DML_FL_REF) && s.ref_pos[j] >= 0) {
        s.ov_rev[j] = stamp;  // k_ov_apply walks every ref position in round 0
      } else if (atomicExch(&s.ov_rev[j], stamp) != stamp) {
        // first to revert j in this round: it is applied from the list
        const int q = atomicAdd(&s.wl_cnt[2 + round % 2], 1);
        if (q < ptotal) s.ovr[round % 2][q] = j;
      }
      return 0;
    };
    if (!dom) {
      if (!over) {
        if (nh > 0 && he0 < firstcg) { const int c3 = revert(hj0); n_ch3 += c3; n_ch += 1 - c3; }
        if (nh > 1 && he1 < firstcg) { const int c3 = revert(hj1); n_ch3 += c3; n_ch += 1 - c3; }
        if (nh > 2 && he2 < firstcg) { const int c3 = revert(hj2); n_ch3 += c3; n_ch += 1 - c3; }
        if (nh > 3 && he3 < firstcg) { const int c3 = revert(hj3); n_ch3 += c3; n_ch += 1 - c3; }
      } else {
        for (int e = sub; e < nn && e < firstcg; e += LPA) {
          int j = row[e];
          if (s.flags[j] & DML_FL_LIMBO) continue;
          double dx, dy, dz;
          vdist_mic(P, xi, yi, zi, s.px[j], s.py[j], s.pz[j], &dx, &dy, &dz);
          if (dot3(dx, dy, dz) > P.rcut2 || s.species[j] == SP_CG) continue;
          const int c3 = revert(j);
          n_ch3 += c3;
          n_ch += 1 - c3;
        }
      }
    }
    for (int off = LPA / 2; off > 0; off >>= 1) {
      n_ch += lpa_xor<LPA>(n_ch, off);
      n_ch3 += lpa_xor<LPA>(n_ch3, off);
    }
    if (sub == 0) {
      uint8_t st = 1;
      if (dom) {
        st |= 2;
      } else if (firstcg != 0x7fffffff) {
        k_try = 1;
        double ne;
        if (u_round) {
          ne = u_round[(int64_t)round * u_stride + i];
        } else {
          u32x4 wv = dml_philox(P.seed, s.pid[i], (uint32_t)P.step, RS_CONTACT,
                                (uint32_t)(1000 + round));
          ne = u01_from(wv.x, wv.y);
        }
        st |= (ne < P.prob) ? 4 : 8;
      }
      s.ov_st[i] = st;
      k_ch = n_ch;
      k_ch3 = n_ch3;
      // a revert or a dominated atom: the sequential pass would recurse
      k_more = dom || n_ch;
    }
    k_r = r;
  };
  // a wave-uniform loop (every lane takes every pass), so that the counters
  // can be summed over the wave: one atomic per wave and counter when the
  // wave's items belong to one replica (always in round 0, whose items are
  // consecutive positions of one 256-slot block; mostly later, the work
  // lists being appended block by block), one per atom otherwise.  The
  // counters of a replica are single addresses, which serialise across the
  // chip when every atom adds to them.
  for (int64_t it0 = (int64_t)bid * (DML_BLOCK / LPA); it0 < nitems; it0 += stride) {
    const int64_t item = it0 + threadIdx.x / LPA;
    int k_try = 0, k_ch = 0, k_ch3 = 0, k_more = 0, k_r = -1;
    if (item < nitems) scan_one(item, k_try, k_ch, k_ch3, k_more, k_r);
    const unsigned long long valid = __ballot(k_r >= 0);
    if (!valid) continue;  // nothing scanned in this wave
    const int r0 = __builtin_amdgcn_readlane(k_r, __builtin_ctzll(valid));
    if (!__ballot(k_r >= 0 && k_r != r0)) {
      // (per-atom counts are small: 32-bit sums by DPP)
      const int t_try = __builtin_amdgcn_readlane(wave_scan_incl(k_try), WAVE - 1);
      const int t_ch = __builtin_amdgcn_readlane(wave_scan_incl(k_ch), WAVE - 1);
      const int t_ch3 = __builtin_amdgcn_readlane(wave_scan_incl(k_ch3), WAVE - 1);
      const bool more = __ballot(k_more) != 0;
      if (threadIdx.x % WAVE == 0) {
        if (t_try) atomicAdd(&s.cnt[r0].try_, (unsigned long long)(uint32_t)t_try);
        if (t_ch) atomicAdd(&s.cnt[r0].choques, (unsigned long long)(uint32_t)t_ch);
        if (t_ch3) atomicAdd(&s.cnt[r0].choques3, (unsigned long long)(uint32_t)t_ch3);
        if (more) s.ov_more[cur * s.nrep + r0] = 1;
      }
    } else if (k_r >= 0) {
      if (k_try) atomicAdd(&s.cnt[k_r].try_, (unsigned long long)(uint32_t)k_try);
      if (k_ch) atomicAdd(&s.cnt[k_r].choques, (unsigned long long)(uint32_t)k_ch);
      if (k_ch3) atomicAdd(&s.cnt[k_r].choques3, (unsigned long long)(uint32_t)k_ch3);
      if (k_more) s.ov_more[cur * s.nrep + k_r] = 1;
    }
  }
}

template <int LPA>
__global__ void __launch_bounds__(DML_BLOCK) __attribute__((amdgpu_waves_per_eu(5)))
k_ov_scan_l(DevState s, const double *u_round, int64_t u_stride, int round, int64_t ptotal) {
  ov_scan_body<LPA>(s, u_round, u_stride, round, ptotal, blockIdx.x, gridDim.x);
}
template __global__ void k_ov_scan_l<1>(DevState, const double *, int64_t, int, int64_t);
template __global__ void k_ov_scan_l<2>(DevState, const double *, int64_t, int, int64_t);
template __global__ void k_ov_scan_l<4>(DevState, const double *, int64_t, int, int64_t);
template __global__ void k_ov_scan_l<8>(DevState, const double *, int64_t, int, int64_t);
template __global__ void k_ov_scan_l<16>(DevState, const double *, int64_t, int, int64_t);

// one atom's outcome of a round (its own verdict st, a revert by a
// neighbour); returns 1 when the atom belongs in the next round's list
__device__ inline int ov_apply_one(const DevState &s, int r, const RepParams &P, int i, int t,
                                   int round) {
  const uint8_t st = s.ov_st[i];
  uint8_t f = s.flags[i];
  int moved = 0;
  if ((st & 1) && !(st & 2)) f |= DML_FL_SKIP;
  if (st & 4) {
    atomicAdd(&s.cnt[r].depo, 1ull);
    s.species[i] = SP_F;
    s.rp[r].any_f = 1;
    moved = 1;
    if (s.pz[i] > P.z0) rep_fail(s, r, ERR_DEPZ0);
  }
  if (st & 8) {
    s.px[i] = s.cx[i]; s.py[i] = s.cy[i]; s.pz[i] = s.cz[i];
    f &= (uint8_t)~DML_FL_SKIP;
    moved = 1;
  }
  if (s.ov_rev[i] == ov_stamp(P, round)) {
    s.ov_rev[i] = 0;  // applied (stamps are positive)
    s.px[i] = s.cx[i]; s.py[i] = s.cy[i]; s.pz[i] = s.cz[i];
    s.vx[i] = 0.0; s.vy[i] = 0.0; s.vz[i] = 0.0;
    s.ax[i] = 0.0; s.ay[i] = 0.0; s.az[i] = 0.0;
    f &= (uint8_t)~DML_FL_SKIP;
    moved = 1;
  }
  if (st || moved) {
    s.flags[i] = f;
    s.ov_st[i] = 0;
  }
  if (moved) put_pf4(s, i, s.px[i], s.py[i], s.pz[i], s.species[i]);
  // the scan of the next round visits this atom when it is unresolved and
  // its replica goes on (its scan reverted or deferred something)
  return !(f & DML_FL_SKIP) && t >= P.poff && s.ov_more[(round % 3) * s.nrep + r];
}

// round 0: every ref position; later rounds: the positions scanned in it.
// A block takes OVA_PER * DML_BLOCK consecutive items at a time and appends
// the unresolved ones to the next round's list with one atomic (a per-item
// or per-wave atomic on the single list counter serialises across the chip);
// the list order is free, every round's outcome depends on its start state only
__device__ __forceinline__ void ov_apply_body(const DevState &s, int round, int64_t ptotal,
                                              int bid, int nbk) {
  __shared__ int32_t buf[DML_BLOCK * OVA_PER];
  __shared__ int wcnt[DML_BLOCK / WAVE];
  __shared__ int64_t base_sh;
  const int prv = (round + 2) % 3;
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  if (bid == 0 && threadIdx.x == 0) s.wl_cnt[2 + (round + 1) % 2] = 0;  // next reverts
  const int64_t nitems =
      round == 0 ? ptotal : min((int64_t)s.wl_cnt[round % 2], ptotal);
  const int32_t *wl = s.ovw[round % 2];
  int32_t *next = s.ovw[(round + 1) % 2];
  int32_t *wb = buf + w * WAVE * OVA_PER;
  const int64_t chunk = (int64_t)DML_BLOCK * OVA_PER;
  for (int64_t b0 = bid * chunk; b0 < nitems; b0 += (int64_t)nbk * chunk) {
    int n = 0;  // this wave's appends so far (wave-uniform)
    for (int k = 0; k < OVA_PER; k++) {
      const int64_t item = b0 + (int64_t)k * DML_BLOCK + threadIdx.x;
      int push = 0, t = 0;
      if (item < nitems) t = round == 0 ? (int)item : wl[item];
      if (item < nitems && (uint32_t)t < (uint64_t)ptotal) {
        const int r = s.blk2rep[t / DML_BLOCK];
        const RepParams &P = s.rp[r];
        if (!P.halted && P.overlap_mode == DML_OVERLAP_ROUNDS) {
          if (round == 0 && t == P.poff) s.rp[r].ovn_stamp = 0;  // reverts move atoms
          // round > 0 and a replica whose scan stopped: nothing to apply
          if ((round == 0 || s.ov_more[prv * s.nrep + r]) && t - P.poff < P.nref)
            push = ov_apply_one(s, r, P, s.ref_order[t], t, round);
        }
      }
      const unsigned long long m = __ballot(push);
      if (push) wb[n + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u))] = t;
      n += __popcll(m);
    }
    if (lane == 0) wcnt[w] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int q = 0; q < DML_BLOCK / WAVE; q++) tot += wcnt[q];
      base_sh = tot ? (int64_t)atomicAdd(&s.wl_cnt[(round + 1) % 2], tot) : 0;
    }
    __syncthreads();
    int64_t off = base_sh;
    for (int q = 0; q < w; q++) off += wcnt[q];
    for (int q = lane; q < n; q += WAVE)
      if (off + q < ptotal) next[off + q] = wb[q];
    __syncthreads();  // buf, wcnt and base_sh are reused
  }
}
__global__ void __launch_bounds__(DML_BLOCK) k_ov_apply(DevState s, int round, int64_t ptotal) {
  ov_apply_body(s, round, ptotal, blockIdx.x, gridDim.x);
}

// the atoms reverted in a round that k_ov_apply does not reach: in round 0
// those outside the ref list (it walks every ref position), later the ones
// the round did not scan (those that were are applied, and their stamp
// cleared, by k_ov_apply)
__device__ __forceinline__ void ov_apply_r_body(const DevState &s, int round, int bid, int nbk) {
  const int64_t cap = (int64_t)s.nblocks * DML_BLOCK;
  const int64_t nitems = min((int64_t)s.wl_cnt[2 + round % 2], cap);
  const int32_t *rl = s.ovr[round % 2];
  const int64_t stride = (int64_t)nbk * DML_BLOCK;
  for (int64_t item = (int64_t)bid * DML_BLOCK + threadIdx.x; item < nitems; item += stride) {
    const int j = rl[item];
    if ((uint32_t)j >= (uint64_t)cap) continue;
    const int r = s.blk2rep[j / DML_BLOCK];
    const RepParams &P = s.rp[r];
    if (P.halted || s.ov_rev[j] != ov_stamp(P, round)) continue;
    const int t = P.poff + s.ref_pos[j];  // ref_pos is replica-local
    const int push = ov_apply_one(s, r, P, j, t, round);
    // one atomic per wave for its appends (the active lanes of this pass)
    const unsigned long long m = __ballot(push);
    if (m) {
      const int lead = __builtin_ctzll(m);
      int base = 0;
      if ((int)(threadIdx.x % WAVE) == lead) base = atomicAdd(&s.wl_cnt[(round + 1) % 2], __popcll(m));
      base = __shfl(base, lead);
      const int q = base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
      if (push && q < cap) s.ovw[(round + 1) % 2][q] = t;
    }
  }
}
__global__ void __launch_bounds__(DML_BLOCK) k_ov_apply_r(DevState s, int round) {
  ov_apply_r_body(s, round, blockIdx.x, gridDim.x);
}

// the last rounds, whose work lists are short, in one workgroup: the three
// phases of each round (scan, apply, apply of the other reverted atoms)
// separated by the workgroup's barriers instead of kernel boundaries, one
// launch for all of them instead of three per round.  The same code per
// phase as the grid-wide kernels, so the same result.
template <int LPA>
__global__ void __launch_bounds__(DML_BLOCK) k_ov_tail(DevState s, const double *u_round, int64_t u_stride, int r0, int r1, int64_t ptotal) {
  for (int q = r0; q < r1; q++) {
    ov_scan_body<LPA>(s, u_round, u_stride, q, ptotal, 0, 1);
    __syncthreads();
    ov_apply_body(s, q, ptotal, 0, 1);
    __syncthreads();
    ov_apply_r_body(s, q, 0, 1);
    __syncthreads();
  }
}
template __global__ void k_ov_tail<4>(DevState, const double *, int64_t, int, int, int64_t);
template __global__ void k_ov_tail<8>(DevState, const double *, int64_t, int, int, int64_t);
template __global__ void k_ov_tail<16>(DevState, const double *, int64_t, int, int, int64_t);

// after the rounds: a replica whose last round still made progress would
// need more rounds
__global__ void k_ov_end(DevState s, int rounds, int err_if_pending) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= s.nrep) return;
  if (s.rp[r].overlap_mode != DML_OVERLAP_ROUNDS || s.rp[r].halted) return;
  if (s.ov_more[((rounds - 1) % 3) * s.nrep + r]) {
    if (err_if_pending) rep_fail(s, r, ERR_OVERLAP);
    atomicAdd(&s.cnt[r].ov_unfinished, 1ull);
  }
}

// advance every replica's stamp base past this call's rounds
__global__ void k_ov_close(DevState s, int rounds) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= s.nrep) return;
  RepParams *Pw = s.rp + r;
  if (Pw->overlap_mode != DML_OVERLAP_ROUNDS) return;
  Pw->ov_round += rounds + 2;
  atomicAdd(&s.cnt[r].passes, 1ull);
}

__global__ void k_init_pf4(DevState s, int64_t n) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < n) put_pf4(s, (int)q, s.px[q], s.py[q], s.pz[q], s.species[q]);
}

__global__ void k_zero_i32(int32_t *a, int64_t n) {
  int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q < n) a[q] = 0;
}

// exclusive prefix sums of two per-block counters over one replica's
// particle blocks, in chunks of SCAN_BLOCK by a whole workgroup (wave scans,
// then the wave totals); returns the two totals to every thread
#define SCAN_BLOCK 1024
__device__ inline void rep_exscan2(int32_t *part, int stride, int b0, int nb, int *t1, int *t2) {
  __shared__ int w1[SCAN_BLOCK / WAVE], w2[SCAN_BLOCK / WAVE];
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  int c1 = 0, c2 = 0;
  for (int q0 = 0; q0 < nb; q0 += SCAN_BLOCK) {
    const int q = q0 + threadIdx.x;
    const int v1 = q < nb ? part[stride * (b0 + q)] : 0;
    const int v2 = (stride == 2 && q < nb) ? part[2 * (b0 + q) + 1] : 0;
    int i1 = v1, i2 = v2;
    for (int off = 1; off < WAVE; off <<= 1) {
      const int x1 = __shfl_up(i1, off), x2 = __shfl_up(i2, off);
      if (lane >= off) {
        i1 += x1;
        i2 += x2;
      }
    }
    if (lane == WAVE - 1) {
      w1[w] = i1;
      w2[w] = i2;
    }
    __syncthreads();
    int o1 = 0, o2 = 0, s1 = 0, s2 = 0;
    for (int k = 0; k < SCAN_BLOCK / WAVE; k++) {
      if (k < w) {
        o1 += w1[k];
        o2 += w2[k];
      }
      s1 += w1[k];
      s2 += w2[k];
    }
    if (q < nb) {
      part[stride * (b0 + q)] = c1 + o1 + i1 - v1;
      if (stride == 2) part[2 * (b0 + q) + 1] = c2 + o2 + i2 - v2;
    }
    c1 += s1;
    c2 += s2;
    __syncthreads();  // w1, w2 are rewritten by the next chunk
  }
  *t1 = c1;
  *t2 = c2;
}

// ------------------------------------------------------------------------
// F -> CG at step end and removal from hs%ref (dana.F90:228-236): stable
// compaction of each replica's ref order
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(DML_BLOCK) k_f2cg_mark(DevState s) {
  __shared__ int cnt;
  int t = blockIdx.x * DML_BLOCK + threadIdx.x;
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  if (P.any_f && t - P.poff < P.nref) {
    int i = s.ref_order[t];
    int keep = 1;
    if (s.species[i] == SP_F) {
      s.species[i] = SP_CG;
      put_pf4_sp(s, i, SP_CG);
      s.flags[i] &= (uint8_t)~(DML_FL_REF | DML_FL_GCMC);
      s.ref_pos[i] = -1;
      keep = 0;
    }
    if (keep) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) s.part_i[blockIdx.x] = cnt;
}

// one workgroup per replica: exclusive scan of the block counts
__global__ void __launch_bounds__(SCAN_BLOCK) k_f2cg_scan(DevState s) {
  int r = blockIdx.x;
  RepParams *Pw = s.rp + r;
  if (Pw->halted) return;
  if (!Pw->any_f) return;
  int t1, t2;
  rep_exscan2(s.part_i, 1, Pw->poff / DML_BLOCK, Pw->pcap / DML_BLOCK, &t1, &t2);
  if (threadIdx.x == 0) Pw->nref_new = t1;
}

__global__ void __launch_bounds__(DML_BLOCK) k_f2cg_scatter(DevState s) {
  __shared__ int wsum[DML_BLOCK / WAVE];
  int t = blockIdx.x * DML_BLOCK + threadIdx.x;
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  if (!P.any_f) return;
  int local = t - P.poff;
  int i = local < P.nref ? s.ref_order[t] : -1;
  int keep = i >= 0 && (s.flags[i] & DML_FL_REF);
  unsigned long long m = __ballot(keep);
  int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  if (lane == 0) wsum[w] = __popcll(m);
  __syncthreads();
  int off = s.part_i[blockIdx.x];
  for (int q = 0; q < w; q++) off += wsum[q];
  if (keep) s.ref_tmp[P.poff + off + __popcll(m & ((1ull << lane) - 1ull))] = i;
}

__global__ void __launch_bounds__(DML_BLOCK) k_f2cg_copy(DevState s) {
  int t = blockIdx.x * DML_BLOCK + threadIdx.x;
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  if (!P.any_f) return;
  int nnew = P.nref_new;
  int local = t - P.poff;
  if (local < nnew) {
    int i = s.ref_tmp[t];
    s.ref_order[t] = i;
    s.ref_pos[i] = local;
  }
}

__global__ void k_f2cg_end(DevState s) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= s.nrep) return;
  RepParams *Pw = s.rp + r;
  if (Pw->halted) return;
  if (!Pw->any_f) return;
  Pw->nref = Pw->nref_new;
  Pw->any_f = 0;
}

// ------------------------------------------------------------------------
// piston: calc_rho (dana.F90:521-549) and maxz (776-794)
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(DML_BLOCK) k_calc_rho(DevState s) {
  __shared__ int cnt;
  int slot = blockIdx.x * DML_BLOCK + threadIdx.x;
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  int local = slot - P.poff;
  if (local < P.nslots && (s.flags[slot] & DML_FL_ALIVE)) {
    double z = s.pz[slot];
    const double zt = P.reservoir == DML_CHUNKS ? P.z1 : P.zmax;  // the chunks sensor
    if (z > P.z0 && z < zt) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0 && cnt) atomicAdd(&s.rp[r].rho_count, cnt);
}

// ------------------------------------------------------------------------
// salida's observables (dana.F90:1141-1183): energia = sum of epot over sys,
// and the sum of m |v|^2 over the atoms that are not CG for kion
// (dana.F90:1342-1376).  Per-block partials, then one block per replica adds
// them in a fixed order, so the result does not depend on scheduling.
// ------------------------------------------------------------------------
__device__ inline double block_sum(double v, double *sh) {
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  __syncthreads();  // earlier readers of sh are done
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int q = 0; q < (int)(blockDim.x / WAVE); q++) t += sh[q];
  return t;
}

__global__ void __launch_bounds__(DML_BLOCK) k_observe_part(DevState s) {
  __shared__ double sh[DML_BLOCK / WAVE];
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  const int slot = blockIdx.x * DML_BLOCK + threadIdx.x;
  const int local = slot - P.poff;
  double e = 0.0, k = 0.0, ni = 0.0, na = 0.0;
  if (local < P.nslots && (s.flags[slot] & DML_FL_ALIVE)) {
    e = s.epot[slot];
    na = 1.0;
    const int z = s.species[slot];
    if (z != SP_CG) {
      const double vx = s.vx[slot], vy = s.vy[slot], vz = s.vz[slot];
      k = (vx * vx + vy * vy + vz * vz) * P.mass[z - 1];  // dot_product(vel,vel)*mass
      ni = 1.0;
    }
  }
  e = block_sum(e, sh);
  k = block_sum(k, sh);
  ni = block_sum(ni, sh);
  na = block_sum(na, sh);
  if (threadIdx.x == 0) {
    double *o = s.obs + 4 * (int64_t)blockIdx.x;
    o[0] = e;
    o[1] = k;
    o[2] = ni;
    o[3] = na;
  }
}

__global__ void __launch_bounds__(DML_BLOCK) k_observe_final(DevState s) {
  __shared__ double sh[DML_BLOCK / WAVE];
  const int r = blockIdx.x;
  const RepParams &P = s.rp[r];
  const int b0 = P.poff / DML_BLOCK, nb = P.pcap / DML_BLOCK;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int q = threadIdx.x; q < nb; q += blockDim.x) {
    const double *o = s.obs + 4 * (int64_t)(b0 + q);
    a0 += o[0];
    a1 += o[1];
    a2 += o[2];
    a3 += o[3];
  }
  a0 = block_sum(a0, sh);
  a1 = block_sum(a1, sh);
  a2 = block_sum(a2, sh);
  a3 = block_sum(a3, sh);
  if (threadIdx.x == 0) {
    double *o = s.obs + 4 * (int64_t)s.nblocks + 4 * r;
    o[0] = a0;
    o[1] = a1;
    o[2] = a2;
    o[3] = a3;
  }
}

// the piston factor of maxz from the count of calc_rho (dana.F90:785)
__device__ inline double piston_lohi(const RepParams &P) {
  double vol = P.box[0] * P.box[1] * (P.zmax - P.z0);
  double rho = (double)P.rho_count / vol;
  const double tau = 0.1;
  return ((P.h / tau) * ((P.rho0 - rho) / rho));
}

__global__ void __launch_bounds__(DML_BLOCK) k_maxz(DevState s) {
  int slot = blockIdx.x * DML_BLOCK + threadIdx.x;
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (P.halted) return;
  if (!P.piston) return;
  if (slot == P.poff) s.rp[r].ovn_stamp = 0;  // positions move
  int local = slot - P.poff;
  if (local < P.nslots && (s.flags[slot] & DML_FL_ALIVE)) {
    double z = s.pz[slot];
    if (z > P.z0) {
      double lohi = piston_lohi(P);
      double nz = z - lohi * (z - P.z0);
      s.pz[slot] = nz;
      put_pf4(s, slot, s.px[slot], s.py[slot], nz, s.species[slot]);
    }
  }
}

// per-replica bookkeeping of the per-routine entry points: 0 = clear the
// calc_rho count, 1 = piston zmax (dana.F90:792), 2 = end of step
__global__ void k_rep_op(DevState s, int op) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= s.nrep) return;
  RepParams *Pw = s.rp + r;
  if (Pw->halted) return;
  if (op == 0) {
    Pw->rho_count = 0;
  } else if (op == 1) {
    if (Pw->piston) {
      double lohi = piston_lohi(*Pw);
      Pw->lohi = lohi;
      Pw->zmax = Pw->zmax - lohi * (Pw->zmax - Pw->z0);
    }
  } else {
    Pw->step += 1;
    s.cnt[r].steps += 1;
    s.cnt[r].particle_steps += (unsigned long long)Pw->nat;
    if (Pw->step_limit > 0 && Pw->step >= Pw->step_limit) Pw->halted |= HALT_PAUSED;
  }
}

// step end: advance the per-replica step counter (and the piston zmax,
// which no kernel of k_maxz reads)
__global__ void k_step_end(DevState s) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= s.nrep) return;
  RepParams *Pw = s.rp + r;
  if (Pw->halted) return;
  if (Pw->piston) {
    double lohi = piston_lohi(*Pw);
    Pw->lohi = lohi;
    Pw->zmax = Pw->zmax - lohi * (Pw->zmax - Pw->z0);
  }
  Pw->rho_count = 0;
  Pw->step += 1;
  s.cnt[r].steps += 1;
  s.cnt[r].particle_steps += (unsigned long long)Pw->nat;
  if (Pw->step_limit > 0 && Pw->step >= Pw->step_limit) Pw->halted |= HALT_PAUSED;
}

// hold the replicas that already have lists out of the first rebuild of
// dml_run (on = 1), or release them (on = 0)
__global__ void k_mask_listed(DevState s, int on) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= s.nrep) return;
  RepParams *Pw = s.rp + r;
  if (on) {
    if (Pw->listed) Pw->halted |= HALT_MASK;
  } else {
    Pw->halted &= ~HALT_MASK;
  }
}

// ------------------------------------------------------------------------
// GCMC (dana.F90:590-713)
// ------------------------------------------------------------------------
// the cells within one cell step of (c0, c1, c2): x, y periodic, z not (its
// ghost layers included), each cell once (ncells of 1 or 2 repeat a cell)
__device__ inline int near_cells(const RepParams &P, int c0, int c1, int c2, int *out) {
  int xs[3], ys[3], zs[3], nx = 0, ny = 0, nz = 0;
  for (int d = -1; d <= 1; d++) {
    int x = cell_wrap(c0 + d, P.ncells[0]);
    bool dup = false;
    for (int q = 0; q < nx; q++) dup |= xs[q] == x;
    if (!dup) xs[nx++] = x;
    int y = cell_wrap(c1 + d, P.ncells[1]);
    dup = false;
    for (int q = 0; q < ny; q++) dup |= ys[q] == y;
    if (!dup) ys[ny++] = y;
    int z = c2 + d;
    if (z >= 0 && z <= P.ncells[2] + 1) zs[nz++] = z;
  }
  int n = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) out[n++] = cell_lin(P, xs[i], ys[j], zs[k]);
  return n;
}

// the k-th cell of near_cells' list (same order) and, in *n, their number:
// one lane per cell without a per-lane array
__device__ inline int near_cell_at(const RepParams &P, int c0, int c1, int c2, int k, int *n) {
  int xs0 = 0, xs1 = 0, xs2 = 0, ys0 = 0, ys1 = 0, ys2 = 0, zs0 = 0, zs1 = 0, zs2 = 0;
  int nx = 0, ny = 0, nz = 0;
  for (int d = -1; d <= 1; d++) {
    const int x = cell_wrap(c0 + d, P.ncells[0]);
    if (!(nx > 0 && xs0 == x) && !(nx > 1 && xs1 == x)) {
      if (nx == 0) xs0 = x; else if (nx == 1) xs1 = x; else xs2 = x;
      nx++;
    }
    const int y = cell_wrap(c1 + d, P.ncells[1]);
    if (!(ny > 0 && ys0 == y) && !(ny > 1 && ys1 == y)) {
      if (ny == 0) ys0 = y; else if (ny == 1) ys1 = y; else ys2 = y;
      ny++;
    }
    const int z = c2 + d;
    if (z >= 0 && z <= P.ncells[2] + 1) {
      if (nz == 0) zs0 = z; else if (nz == 1) zs1 = z; else zs2 = z;
      nz++;
    }
  }
  *n = nx * ny * nz;
  const int i = k % nx, j = (k / nx) % ny, kz = k / (nx * ny);
  return cell_lin(P, i == 0 ? xs0 : i == 1 ? xs1 : xs2, j == 0 ? ys0 : j == 1 ? ys1 : ys2,
                  kz == 0 ? zs0 : kz == 1 ? zs1 : zs2);
}

// batched trial energies (insertion at a point, or removal of an atom: the
// same sum without the atom itself): one wave per trial over the cells
// around the point.  The cells are those of the last sort; every atom has
// moved less than the skin since, so one cell step (>= rcut + skin) still
// covers the pair cutoff.  excl holds slots of replica r (-1 = none); a
// halted replica gets du = 0, hard = -1 for every trial.
__global__ void __launch_bounds__(DML_BLOCK) k_gcmc_energy(DevState s, int r, int ntrial,
                                                            const double *pos3,
                                                            const int32_t *excl, double *du,
                                                            int32_t *hard) {
  int w = (blockIdx.x * DML_BLOCK + threadIdx.x) / WAVE, lane = threadIdx.x % WAVE;
  if (w >= ntrial) return;
  const RepParams &P = s.rp[r];
  const int ex = excl ? excl[w] : -1;
  const int skip = ex < 0 ? -1 : P.poff + ex;  // the atom itself, for a removal
  if (P.halted || (!pos3 && (skip < 0 || !(s.flags[skip] & DML_FL_ALIVE)))) {  // no such atom now
    if (lane == 0) {
      du[w] = 0.0;
      hard[w] = -1;
    }
    return;
  }
  // the trial point, or the removed atom's own position
  const double x = pos3 ? pos3[3 * w] : s.px[skip];
  const double y = pos3 ? pos3[3 * w + 1] : s.py[skip];
  const double z = pos3 ? pos3[3 * w + 2] : s.pz[skip];
  double e = 0.0;
  int h = 0;
  int cells[27];
  int ncell = 1;
  cells[0] = 0;
  if (P.use_cells) {
    const int c0 = (int)(x / P.cell[0]) + 1, c1 = (int)(y / P.cell[1]) + 1;
    const int c2 = (int)(z / P.cell[2]) + 1;
    if (c0 < 0 || c1 < 0 || c2 < 0 || c0 > P.ncells[0] + 1 || c1 > P.ncells[1] + 1 ||
        c2 > P.ncells[2] + 1) {
      if (lane == 0) {
        du[w] = 0.0;
        hard[w] = -1;
      }
      return;
    }
    // every atom within the pair cutoff lies in a neighbouring cell (x, y
    // periodic, z with its ghost layers), each cell once
    ncell = near_cells(P, c0, c1, c2, cells);
  }
  for (int k = 0; k < ncell; k++) {
    const int cell = cells[k];
    int cnt = s.cell_cnt[P.coff + cell];
    if (cnt > P.bcap) cnt = P.bcap;
    const int64_t base = P.boff + (int64_t)cell * P.bcap;
    for (int q = lane; q < cnt; q += WAVE) {
      const int j = s.bucket[base + q];
      if (j == skip) continue;
      double dx, dy, dz;
      vdist_mic(P, x, y, z, s.px[j], s.py[j], s.pz[j], &dx, &dy, &dz);
      const double dr2 = dot3(dx, dy, dz);
      if ((s.flags[j] & DML_FL_GCMC) && dr2 < P.rcut * P.rcut) h = 1;
      const int m = s.species[j];
      const double r0 = P.r0[m - 1], eps = P.eps[m - 1];  // pair (Li, m)
      if (dr2 > r0 * r0) continue;
      const double dr = sqrt(dr2);
      const double b = pow6d(r0) / pow6d(dr);
      e += eps * b * (b - 2) + eps;
    }
  }
  for (int off = WAVE / 2; off > 0; off >>= 1) e += __shfl_xor(e, off);
  const unsigned long long hm = __ballot(h);
  if (lane == 0) {
    du[w] = e;
    hard[w] = hm != 0;
  }
}

// ------------------------------------------------------------------------
// GCMC reservoir: gcmc_run (dana.F90:590-713) with the group bookkeeping it
// drives (Groups.F90:730-766, 1058-1135; Cells.F90:105-175, 281-352;
// Neighbor.F90:173-318).  Trials are sequential, as in the reference: one
// workgroup per replica, each trial's scans spread over the block.  The
// cells' bucket-ordered positions are current here (a GCMC replica's cells
// are sorted by the test_update before its trials, and insertions and
// deletions keep them), so the scans read them instead of gathering.
// ------------------------------------------------------------------------
__device__ inline int blk_sum(int v, int *sh) {
  __syncthreads();
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if (threadIdx.x % WAVE == 0) sh[threadIdx.x / WAVE] = v;
  __syncthreads();
  int t = 0;
  for (int q = 0; q < (int)(blockDim.x / WAVE); q++) t += sh[q];
  __syncthreads();
  return t;
}
__device__ inline int blk_min(int v, int *sh) {
  __syncthreads();
  for (int off = WAVE / 2; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
  if (threadIdx.x % WAVE == 0) sh[threadIdx.x / WAVE] = v;
  __syncthreads();
  int t = sh[0];
  for (int q = 1; q < (int)(blockDim.x / WAVE); q++) t = min(t, sh[q]);
  __syncthreads();
  return t;
}
// exclusive prefix of a 0/1 flag over the block, and the block total
// move the bucket entries of one cell (slot, FP64 and FP32 positions) one
// place up (up = 1: entries at .. cnt-1, to open entry at) or one place down
// (up = 0: entries at+1 .. cnt-1, closing entry at), by the block's threads:
// chunks of blockDim entries, from the top when moving up and from the
// bottom when moving down, each read before a barrier and written after it.
// Starts with a barrier: every thread has read the cell's count (at, cnt)
// before anyone, the caller's thread 0 after the return included, writes
__device__ inline void blk_bucket_shift(const DevState &s, int64_t base, int at, int cnt, int up) {
  __syncthreads();
  const int nb = (int)blockDim.x;
  const int first = up ? at : at + 1;
  const int nchunks = cnt > first ? (cnt - first + nb - 1) / nb : 0;
  for (int k = 0; k < nchunks; k++) {
    const int lo = up ? max(first, cnt - (k + 1) * nb) : first + k * nb;
    const int hi = up ? cnt - k * nb : min(cnt, first + (k + 1) * nb);
    const int q = lo + (int)threadIdx.x;
    int32_t b = 0;
    double x = 0.0, y = 0.0, z = 0.0;
    float4 f = make_float4(0.f, 0.f, 0.f, 0.f);
    if (q < hi) {
      b = s.bucket[base + q];
      x = s.bx[base + q];
      y = s.by[base + q];
      z = s.bz[base + q];
      f = s.bf4[base + q];
    }
    __syncthreads();
    if (q < hi) {
      const int64_t d = base + q + (up ? 1 : -1);
      s.bucket[d] = b;
      s.bx[d] = x;
      s.by[d] = y;
      s.bz[d] = z;
      s.bf4[d] = f;
    }
    __syncthreads();
  }
}
__device__ inline int blk_scan01(int f, int *sh, int *total) {
  __syncthreads();
  unsigned long long m = __ballot(f);
  int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  if (lane == 0) sh[w] = __popcll(m);
  __syncthreads();
  int off = 0, tot = 0;
  for (int q = 0; q < (int)(blockDim.x / WAVE); q++) {
    if (q < w) off += sh[q];
    tot += sh[q];
  }
  *total = tot;
  __syncthreads();
  return off + __popcll(m & ((1ull << lane) - 1ull));
}

__device__ inline double gc_unif(const RepParams &P, const double *u, int t, int k, int call) {
  if (u) return u[5 * t + k];
  u32x4 w = dml_philox(P.seed, 0xFFFFFFFFu - (uint32_t)call, (uint32_t)P.step, RS_GCMC,
                       (uint32_t)(8 * t + k));
  return u01_from(w.x, w.y);
}
__device__ inline double gc_gauss(const RepParams &P, const double *g, int t, int k, int call) {
  if (g) return g[3 * t + k];
  double a, b;
  gauss2(dml_philox(P.seed, 0xFFFFFFFFu - (uint32_t)call, (uint32_t)P.step, RS_GCMC,
                    (uint32_t)(8 * t + 5 + (k >> 1))), &a, &b);
  return (k & 1) ? b : a;
}

// ------------------------------------------------------------------------
// The O(n) list work around gcmc_run, spread over the grid (a block per
// particle block of every GCMC replica) so that the sequential trial kernel
// only touches what its trials touch.  Stage 0, before the trials: the
// gcmc list drops the atoms that left the group (gcmc%detach at F -> CG,
// dana.F90:235) into the working list wl[1], and the control volume's atoms
// are listed in list order (dana.F90:609-615) into ref_tmp.  Stage 1, after
// them: the working gcmc list and the ref list (deleted entries -1) are
// compacted, ref_pos renumbered.  Each is a stable compaction: block counts,
// a per-replica scan of the counts, a scatter.
// ------------------------------------------------------------------------
__device__ inline bool gc_sel(const RepParams &P, int r, int rep_only) {
  return rep_only >= 0 ? r == rep_only : P.reservoir == DML_GCMC;
}

// the two keep flags of slot t's entries (stage 0: gcmc member, and in the
// control volume; stage 1: gcmc entry kept, ref entry kept) and their values
__device__ inline void gc_keep(const DevState &s, const RepParams &P, int t, int stage, int *v1,
                               int *k1, int *v2, int *k2) {
  const int local = t - P.poff;
  if (stage == 0) {
    const int v = local < P.ngc ? s.gc_order[t] : -1;
    *v1 = *v2 = v;
    *k1 = v >= 0 && (s.flags[v] & DML_FL_GCMC);
    *k2 = 0;
    if (*k1) {
      const double z = s.pz[v];
      *k2 = !(z < P.z0 || z > P.zmax);
    }
  } else {
    *v1 = local < P.gc_len ? s.wl[1][t] : -1;
    *k1 = *v1 >= 0 && (s.flags[*v1] & DML_FL_GCMC);
    *v2 = local < P.ref_len ? s.ref_order[t] : -1;
    *k2 = *v2 >= 0;
  }
}

__global__ void __launch_bounds__(DML_BLOCK) k_gc_count(DevState s, int rep_only, int stage) {
  __shared__ int w1[DML_BLOCK / WAVE], w2[DML_BLOCK / WAVE];
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (!gc_sel(P, r, rep_only) || (stage == 0 ? P.halted != 0 : P.gc_pending != 2)) return;
  int v1, k1, v2, k2;
  gc_keep(s, P, blockIdx.x * DML_BLOCK + threadIdx.x, stage, &v1, &k1, &v2, &k2);
  const unsigned long long m1 = __ballot(k1), m2 = __ballot(k2);
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  if (lane == 0) {
    w1[w] = __popcll(m1);
    w2[w] = __popcll(m2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0, b = 0;
    for (int q = 0; q < DML_BLOCK / WAVE; q++) {
      a += w1[q];
      b += w2[q];
    }
    s.part_i[2 * blockIdx.x] = a;
    s.part_i[2 * blockIdx.x + 1] = b;
  }
}

// one workgroup per replica: exclusive scans of both block counts
__global__ void __launch_bounds__(SCAN_BLOCK) k_gc_scan(DevState s, int rep_only, int stage) {
  const int r = blockIdx.x;
  RepParams *Pw = s.rp + r;
  if (!gc_sel(*Pw, r, rep_only)) return;
  if (stage == 0) {
    if (Pw->halted) {  // nothing staged for this call
      if (threadIdx.x == 0) Pw->gc_pending = 0;
      return;
    }
  } else if (Pw->gc_pending != 2) {
    return;
  }
  int c1, c2;
  rep_exscan2(s.part_i, 2, Pw->poff / DML_BLOCK, Pw->pcap / DML_BLOCK, &c1, &c2);
  if (threadIdx.x == 0) {
    if (stage == 0) {
      Pw->gc_len = c1;
      Pw->gc_inv = c2;
      Pw->gc_pending = 1;
    } else {
      Pw->ngc = c1;
      Pw->nref = c2;
      Pw->gc_pending = 3;
    }
  }
}

__global__ void __launch_bounds__(DML_BLOCK) k_gc_scatter(DevState s, int rep_only, int stage) {
  __shared__ int w1[DML_BLOCK / WAVE], w2[DML_BLOCK / WAVE];
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (!gc_sel(P, r, rep_only) || P.gc_pending != (stage == 0 ? 1 : 3)) return;
  int v1, k1, v2, k2;
  gc_keep(s, P, blockIdx.x * DML_BLOCK + threadIdx.x, stage, &v1, &k1, &v2, &k2);
  const unsigned long long m1 = __ballot(k1), m2 = __ballot(k2);
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  if (lane == 0) {
    w1[w] = __popcll(m1);
    w2[w] = __popcll(m2);
  }
  __syncthreads();
  int o1 = s.part_i[2 * blockIdx.x], o2 = s.part_i[2 * blockIdx.x + 1];
  for (int q = 0; q < w; q++) {
    o1 += w1[q];
    o2 += w2[q];
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  int32_t *d1 = stage == 0 ? s.wl[1] : s.gc_order;
  if (k1) d1[P.poff + o1 + __popcll(m1 & below)] = v1;
  if (k2) s.ref_tmp[P.poff + o2 + __popcll(m2 & below)] = v2;
}

// stage 1: the compacted ref list back in place, ref_pos renumbered
__global__ void __launch_bounds__(DML_BLOCK) k_gc_copy(DevState s, int rep_only) {
  int r;
  const RepParams &P = rep_of_block(s, blockIdx.x, &r);
  if (!gc_sel(P, r, rep_only) || P.gc_pending != 3) return;
  const int t = blockIdx.x * DML_BLOCK + threadIdx.x, local = t - P.poff;
  if (local < P.nref) {
    const int v = s.ref_tmp[t];
    s.ref_order[t] = v;
    s.ref_pos[v] = local;
  }
}

#define GC_TRACK 256
#define GC_BLOCK DML_GCMC_BLOCK
// stable block compaction of src[0, len) into dst (dst may not alias src):
// keeps entries v with keep(q, v); each wave takes a contiguous quarter in
// chunks of 64 (ballot counts, no barriers), one block scan of the wave
// totals, then the same sweep writes.  Returns the kept count.
template <class F>
__device__ int blk_compact(const int32_t *src, int len, int32_t *dst, int *sh, F keep) {
  __syncthreads();  // src may just have been written by other threads
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  const int nwv = blockDim.x / WAVE;
  const int part = (len + nwv - 1) / nwv;
  const int a = min(len, w * part), b = min(len, a + part);
  int cnt = 0;
  for (int q0 = a; q0 < b; q0 += WAVE) {
    int q = q0 + lane;
    int k = q < b ? (keep(q, src[q]) ? 1 : 0) : 0;
    cnt += __popcll(__ballot(k));
  }
  __syncthreads();
  if (lane == 0) sh[w] = cnt;
  __syncthreads();
  int off = 0, tot = 0;
  for (int q = 0; q < nwv; q++) {
    if (q < w) off += sh[q];
    tot += sh[q];
  }
  for (int q0 = a; q0 < b; q0 += WAVE) {
    int q = q0 + lane;
    int v = q < b ? src[q] : -1;
    int k = q < b ? (keep(q, v) ? 1 : 0) : 0;
    unsigned long long m = __ballot(k);
    if (k) dst[off + __popcll(m & ((1ull << lane) - 1ull))] = v;
    off += __popcll(m);
  }
  __syncthreads();
  return tot;
}


__global__ void __launch_bounds__(GC_BLOCK) k_gcmc_run(DevState s, int rep_only, const double *u,
                                                         const double *g, double act_in,
                                                         int nadj_in) {
  __shared__ int sh[GC_BLOCK / WAVE];
  __shared__ int sh_cells[27];
  __shared__ int sh_ncell;
  __shared__ int sh_hit, sh_full;
  __shared__ int sh_freed_b[GC_TRACK];  // hs%b indices freed in this call (ascending)
  __shared__ int sh_tomb[GC_TRACK];     // deleted positions of invl (ascending)
  const int r = rep_only >= 0 ? rep_only : (int)blockIdx.x;
  if (r >= s.nrep) return;
  RepParams *Pw = s.rp + r;
  const RepParams &P = *Pw;
  if (P.halted || (rep_only < 0 && P.reservoir != DML_GCMC) || P.gc_pending != 1) return;
  if (P.gb_state == GB_ON) return;  // this call's trials were committed in a batch
  const int tid = threadIdx.x, lane = tid % WAVE, wv = tid / WAVE;
  const int poff = P.poff;
  const double act = rep_only >= 0 ? act_in : P.act;
  const int nadj = rep_only >= 0 ? nadj_in : P.nadj;
  const int call = P.gcmc_calls;
  const double rc2 = P.rcut * P.rcut;
  const double z0 = P.z0, zmax = P.zmax;
  const double vol = P.box[0] * P.box[1] * (zmax - z0);
  const double rl2 = P.rl2;
  const int maxnn = s.maxnn;
  const int listed = P.listed;
  // the lists staged by k_gc_scatter: the gcmc list without the atoms that
  // left the group (tombstones -1 during the call) and the control volume's
  // atoms in list order (dana.F90:609-615), n = their count
  int32_t *gc = s.wl[1] + poff;
  int32_t *invl = s.ref_tmp + poff;
  int32_t *tmp = s.wl[0] + poff;      // scratch (pcap entries)
  int32_t *rord = s.ref_order + poff;
  int glen = P.gc_len;
  int ilen = P.gc_inv;
  int n = ilen;
  int idel = 0;  // in-volume entries deleted so far (tombstones in invl)
  int nslots = P.nslots, nat = P.nat, nlimbo = P.nlimbo, b_amax = P.b_amax, b_nat = P.b_nat;
  int rlen = P.nref;  // ref list length with tombstones
  int head = 0, tail = glen - 1;  // first / last live gcmc entry
  uint32_t pid_next = P.pid_next;
  int seq_next = P.seq_next;
  unsigned long long n_ins = 0, n_del = 0;
  int nfreed_b = 0;
  int hs_scan = P.free_lo, b_scan = P.bfree_lo;  // lowest slot / index not known to be taken
  int fail = 0;
  const double ui_ev = 1.0e2 * 1.0e2 * 1.6605402e-27 * (1.0 / 1.60219e-19);
  const double kB_ui = 8.617385e-05 * (1.0 / ui_ev);  // Constants.F90:100-167
  for (int t = 0; t < nadj && !fail; t++) {
    __syncthreads();  // the previous trial's single-thread writes
    while (head <= tail && (gc[head] < 0 || !(s.flags[gc[head]] & DML_FL_GCMC))) head++;
    if (head > tail) { fail = ERR_NOPART; break; }
    const int tmpl = gc[head];
    const double beta = sqrt(kB_ui * 300.0 / P.mass[s.species[tmpl] - 1]);
    const bool insert = (P.gcmc_mode & DML_GCMC_ALTERNATE) ? !(t & 1)
                                                          : gc_unif(P, u, t, 0, call) < 0.5;
    if (insert) {
      // ---- creation attempt (dana.F90:627-677)
      if (act * vol / (double)(n + 1) < gc_unif(P, u, t, 1, call)) continue;
      const double rx = gc_unif(P, u, t, 2, call) * P.box[0];
      const double ry = gc_unif(P, u, t, 3, call) * P.box[1];
      const double rz = gc_unif(P, u, t, 4, call) * (zmax - z0) + z0;
      int ci0 = 0, ci1 = 0, ci2 = 0;
      if (P.use_cells) {
        ci0 = (int)(rx / P.cell[0]) + 1;
        ci1 = (int)(ry / P.cell[1]) + 1;
        ci2 = (int)(rz / P.cell[2]) + 1;
        if (ci0 < 0 || ci1 < 0 || ci2 < 0 || ci0 > P.ncells[0] + 1 || ci1 > P.ncells[1] + 1 ||
            ci2 > P.ncells[2] + 1) { fail = ERR_UNSUP; break; }
      }
      if (tid == 0) {
        sh_ncell = P.use_cells ? near_cells(P, ci0, ci1, ci2, sh_cells) : 1;
        if (!P.use_cells) sh_cells[0] = 0;
        sh_hit = 0;
        sh_full = 0;
      }
      __syncthreads();
      const int ncell = sh_ncell;
      // hard core against the gcmc group (dana.F90:638-653): every gcmc atom
      // closer than rc lies in a neighbouring cell.  The near cells as one
      // stream (counts in one load, a wave prefix sum), 64 entries a pass,
      // the passes dealt to the workgroup's waves
      {
        int kcnt = 0;
        int64_t kbase = 0;
        if (lane < ncell) {
          const int cc = sh_cells[lane];
          kbase = P.boff + (int64_t)cc * P.bcap;
          kcnt = min(s.cell_cnt[P.coff + cc], P.bcap);
        }
        const int kincl = wave_scan_incl(kcnt);
        const int total = __builtin_amdgcn_readlane(kincl, WAVE - 1);
        for (int s0 = wv * WAVE; s0 < total; s0 += GC_BLOCK) {
          const int sp = s0 + lane;
          int k = 0;
          for (int c = 0; c < ncell; c++) k += __builtin_amdgcn_readlane(kincl, c) <= sp;
          const int kk = min(k, WAVE - 1);
          const int64_t base = ((int64_t)__shfl((int)(kbase >> 32), kk) << 32) |
                               (uint32_t)__shfl((int)(kbase & 0xffffffff), kk);
          const int q = sp - (__shfl(kincl, kk) - __shfl(kcnt, kk));
          if (sp >= total) continue;
          const int o = s.bucket[base + q];
          if (!(s.flags[o] & DML_FL_GCMC)) continue;
          double dx, dy, dz;  // distance(o%pos, r): r - o%pos
          vdist_mic(P, rx, ry, rz, s.bx[base + q], s.by[base + q], s.bz[base + q], &dx, &dy,
                    &dz);
          if (dot3(dx, dy, dz) < rc2) sh_hit = 1;
        }
      }
      __syncthreads();
      if (sh_hit) continue;
      // appends below stay within the replica's pcap entries: drop the
      // deleted entries of a list that is full (rare: needs as many
      // deletions as free slots in one call)
      if (glen == P.pcap) {
        glen = blk_compact(gc, glen, tmp, sh,
                           [&](int, int v) { return v >= 0 && (s.flags[v] & DML_FL_GCMC); });
        for (int q = tid; q < glen; q += GC_BLOCK) gc[q] = tmp[q];
        head = 0;
        tail = glen - 1;
        __syncthreads();
      }
      if (ilen == P.pcap) {
        ilen = blk_compact(invl, ilen, tmp, sh, [&](int, int v) { return v >= 0; });
        for (int q = tid; q < ilen; q += GC_BLOCK) invl[q] = tmp[q];
        idel = 0;
        __syncthreads();
      }
      if (rlen == P.pcap) {
        rlen = blk_compact(rord, rlen, tmp, sh, [&](int, int v) { return v >= 0; });
        for (int q = tid; q < rlen; q += GC_BLOCK) {
          rord[q] = tmp[q];
          s.ref_pos[tmp[q]] = q;
        }
        __syncthreads();
      }
      n++;
      const int tail_atom = gc[tail];
      // hs slot: first free slot or append (Groups.F90:1083-1093)
      int hs;
      if (nslots >= nat + 1 + nlimbo) {
        // 16 flags per load from the 16-aligned block holding hs_scan; the
        // block stops at the first window of 4096 slots that holds a free one
        int cand = 0x7fffffff;
        const uint8_t *fl = s.flags + poff;
        for (int w0 = hs_scan & ~15; w0 < nslots; w0 += 16 * GC_BLOCK) {
          const int q16 = w0 + 16 * tid;
          if (q16 < nslots) {
            const uint4 v = *reinterpret_cast<const uint4 *>(fl + q16);
            for (int k = 0; k < 16; k++) {
              const int q = q16 + k;
              const uint32_t wd = k < 4 ? v.x : k < 8 ? v.y : k < 12 ? v.z : v.w;
              const uint32_t f = (wd >> (8 * (k & 3))) & 0xffu;
              if (q >= hs_scan && q < nslots && !(f & (DML_FL_ALIVE | DML_FL_LIMBO))) {
                cand = q;
                break;
              }
            }
          }
          if (__syncthreads_or(cand != 0x7fffffff)) break;
        }
        hs = blk_min(cand, sh);
        if (hs == 0x7fffffff) { fail = ERR_UNSUP; break; }
        hs_scan = hs + 1;
      } else {
        if (nslots >= P.pcap) { fail = ERR_BUCKET; break; }
        hs = nslots++;
      }
      // hs%b index the same way (indices freed in this call count too)
      int bs;
      if (b_amax >= b_nat + 1) {
        int cand = 0x7fffffff;
        for (int w0 = b_scan; w0 < b_amax; w0 += GC_BLOCK) {
          const int q = w0 + tid;
          if (q < b_amax && s.b2slot[poff + q] < 0) cand = q;
          if (__syncthreads_or(cand != 0x7fffffff)) break;
        }
        bs = blk_min(cand, sh);
        if (nfreed_b && sh_freed_b[0] < bs) bs = sh_freed_b[0];
        if (bs == 0x7fffffff) { fail = ERR_UNSUP; break; }
        const int from_freed = nfreed_b && bs == sh_freed_b[0];
        __syncthreads();
        if (from_freed) {
          if (tid == 0)
            for (int q = 1; q < nfreed_b; q++) sh_freed_b[q - 1] = sh_freed_b[q];
          nfreed_b--;
        } else {
          b_scan = bs + 1;
        }
      } else {
        bs = b_amax++;
      }
      const int i = poff + hs;
      if (tid == 0) {
        // the velocity lands on the list's last atom (dana.F90:666-668)
        s.vx[tail_atom] = beta * gc_gauss(P, g, t, 0, call);
        s.vy[tail_atom] = beta * gc_gauss(P, g, t, 1, call);
        s.vz[tail_atom] = beta * gc_gauss(P, g, t, 2, call);
        s.species[i] = s.species[tmpl];  // atom_asign (Groups.F90:484-502)
        s.vx[i] = s.vx[tmpl]; s.vy[i] = s.vy[tmpl]; s.vz[i] = s.vz[tmpl];
        s.fx[i] = s.fx[tmpl]; s.fy[i] = s.fy[tmpl]; s.fz[i] = s.fz[tmpl];
        s.ax[i] = s.ax[tmpl]; s.ay[i] = s.ay[tmpl]; s.az[i] = s.az[tmpl];
        s.epot[i] = s.epot[tmpl];
        s.px[i] = rx; s.py[i] = ry; s.pz[i] = rz;
        s.ox[i] = rx; s.oy[i] = ry; s.oz[i] = rz;
        s.cx[i] = 1e8; s.cy[i] = 1e8; s.cz[i] = 1e8;
        s.flags[i] = DML_FL_ALIVE | DML_FL_REF | DML_FL_GCMC;
        s.bkey[i] = bs;
        s.b2slot[poff + bs] = i;
        s.pid[i] = pid_next;
        s.seq[i] = seq_next;
        s.rowflag[i] = 0;
        s.nn[i] = 0;
        rord[rlen] = i;
        s.ref_pos[i] = rlen;
        put_pf4(s, i, rx, ry, rz, s.species[tmpl]);
      }
      {
        // cgroup_sort_atom: push-front into its cell chain (Cells.F90:299-300);
        // the Verlet bucket keeps ascending hs%b order (the insertion place
        // is the number of smaller keys); the block moves the entries above
        const int cell = P.use_cells ? cell_lin(P, ci0, ci1, ci2) : 0;
        const int64_t base = P.boff + (int64_t)cell * P.bcap;
        const int cnt = s.cell_cnt[P.coff + cell];
        if (cnt >= P.bcap) { fail = ERR_BUCKET; break; }
        int at = 0;
        if (!P.use_cells) {
          int below = 0;
          for (int q = tid; q < cnt; q += GC_BLOCK) below += s.bkey[s.bucket[base + q]] < bs;
          at = blk_sum(below, sh);
        }
        blk_bucket_shift(s, base, at, cnt, 1);
        if (tid == 0) {
          s.bucket[base + at] = i;
          s.bx[base + at] = rx; s.by[base + at] = ry; s.bz[base + at] = rz;
          s.bf4[base + at] = bucket_f4(s, i);
          s.cell_cnt[P.coff + cell] = cnt + 1;
        }
      }
      __syncthreads();
      pid_next++;
      seq_next++;
      nat++;
      b_nat++;
      rlen++;
      if (listed) {
        const int ghost = P.use_cells && (ci2 == 0 || ci2 == P.ncells[2] + 1 || ci0 == 0 ||
                                          ci0 == P.ncells[0] + 1 || ci1 == 0 ||
                                          ci1 == P.ncells[1] + 1);
        if (wv == 0) {
          // ngroup_sort_atom (Neighbor.F90:271-318): the own row in the
          // order of ngroup_cells_atom (stencil, chain; strict '<') or the
          // Verlet scan ('>' rejects)
          int32_t *row = s.nbr + (int64_t)i * maxnn;
          int nn = 0;
          // the stencil's cells as one stream (stencil order, then chain
          // order): lane k holds the count and bucket start of cell k, all
          // loaded at once; 64 stream entries a pass, each lane finding its
          // cell among the stencil's prefix sums
          int kcnt = 0;
          int64_t kbase = 0;
          if (P.use_cells) {
            if (lane < 27) {
              const int n0 = cell_wrap(ci0 + kMap[lane][0], P.ncells[0]);
              const int n1 = cell_wrap(ci1 + kMap[lane][1], P.ncells[1]);
              const int n2 = cell_wrap(ci2 + kMap[lane][2], P.ncells[2]);
              const int cc = cell_lin(P, n0, n1, n2);
              kbase = P.boff + (int64_t)cc * P.bcap;
              kcnt = min(s.cell_cnt[P.coff + cc], P.bcap);
            }
          } else if (lane == 0) {
            kbase = P.boff;
            kcnt = min(s.cell_cnt[P.coff], P.bcap);
          }
          const int kincl = wave_scan_incl(kcnt);
          const int total = __builtin_amdgcn_readlane(kincl, WAVE - 1);
          for (int s0 = 0; s0 < total; s0 += WAVE) {
            const int sp = s0 + lane;
            // the stencil cell of stream position sp: the number of cells
            // whose stream ends at or before it
            int k = 0;
            for (int c = 0; c < 27; c++) k += __builtin_amdgcn_readlane(kincl, c) <= sp;
            const int kk = min(k, WAVE - 1);
            const int64_t base = ((int64_t)__shfl((int)(kbase >> 32), kk) << 32) |
                                 (uint32_t)__shfl((int)(kbase & 0xffffffff), kk);
            const int q = sp - (__shfl(kincl, kk) - __shfl(kcnt, kk));
            int hitq = 0, j = -1;
            if (sp < total) {
              j = s.bucket[base + q];
              if (j != i) {
                double dx, dy, dz;
                vdist_mic(P, s.bx[base + q], s.by[base + q], s.bz[base + q], rx, ry, rz, &dx,
                          &dy, &dz);
                const double rd = dot3(dx, dy, dz);
                hitq = P.use_cells ? (rd < rl2) : !(rd > rl2);
              }
            }
            const unsigned long long m = __ballot(hitq);
            const int pr = __popcll(m & ((1ull << lane) - 1ull));
            if (hitq && nn + pr < maxnn) row[nn + pr] = j;
            nn += __popcll(m);
          }
          if (lane == 0) {
            if (nn > maxnn) sh_full = 1;
            s.nn[i] = min(nn, maxnn);
            s.rowflag[i] = (uint8_t)ghost;
            if (ghost) Pw->any_ghost = 1;
          }
        } else {
          // and i joins the row of every ref atom within the list radius
          // (all of them lie in a neighbouring cell; each such row gets one
          // entry, so the order of the visits is free): the near cells as
          // one stream, 64 entries a pass, the passes dealt to waves 1-3
          const int nw3 = GC_BLOCK / WAVE - 1;
          int kcnt = 0;
          int64_t kbase = 0;
          if (lane < ncell) {
            const int cc = sh_cells[lane];
            kbase = P.boff + (int64_t)cc * P.bcap;
            kcnt = min(s.cell_cnt[P.coff + cc], P.bcap);
          }
          const int kincl = wave_scan_incl(kcnt);
          const int total = __builtin_amdgcn_readlane(kincl, WAVE - 1);
          for (int s0 = (wv - 1) * WAVE; s0 < total; s0 += nw3 * WAVE) {
            const int sp = s0 + lane;
            int k = 0;
            for (int c = 0; c < ncell; c++) k += __builtin_amdgcn_readlane(kincl, c) <= sp;
            const int kk = min(k, WAVE - 1);
            const int64_t base = ((int64_t)__shfl((int)(kbase >> 32), kk) << 32) |
                                 (uint32_t)__shfl((int)(kbase & 0xffffffff), kk);
            const int q = sp - (__shfl(kincl, kk) - __shfl(kcnt, kk));
            if (sp >= total) continue;
            const int aj = s.bucket[base + q];
            if (aj == i || !(s.flags[aj] & DML_FL_REF)) continue;
            double dx, dy, dz;
            vdist_mic(P, rx, ry, rz, s.bx[base + q], s.by[base + q], s.bz[base + q], &dx, &dy,
                      &dz);
            if (dot3(dx, dy, dz) > rl2) continue;
            const int m = s.nn[aj];
            if (m < maxnn) {
              s.nbr[(int64_t)aj * maxnn + m] = i;
              s.nn[aj] = m + 1;
            } else {
              sh_full = 1;
            }
          }
        }
        __syncthreads();
        if (sh_full) { fail = ERR_LIST; break; }
      }
      if (tid == 0) {
        gc[glen] = i;
        invl[ilen] = i;  // inside the control volume by construction
      }
      tail = glen;
      glen++;
      ilen++;
      n_ins++;
    } else {
      // ---- destruction attempt (dana.F90:680-708)
      if ((double)n / (vol * act) < gc_unif(P, u, t, 1, call)) continue;
      int m = (int)floor(gc_unif(P, u, t, 2, call) * n) + 1;
      if (m > n) m = n;
      // the m-th in-volume atom in list order: position m-1 among the
      // entries not deleted in this call (sorted tombstones)
      if (idel == GC_TRACK) {  // many deletions: drop the tombstones
        __syncthreads();
        ilen = blk_compact(invl, ilen, tmp, sh, [&](int, int v) { return v >= 0; });
        for (int q = tid; q < ilen; q += GC_BLOCK) invl[q] = tmp[q];
        idel = 0;
        __syncthreads();
      }
      int p = m - 1;
      for (int q = 0; q < idel; q++) {
        if (sh_tomb[q] <= p) p++;
        else break;
      }
      if (p < 0 || p >= ilen) { fail = ERR_NOPART; break; }
      const int o = invl[p];
      if (o < poff || o >= poff + P.pcap) { fail = ERR_UNSUP; break; }
      n--;
      // cgroup_unsort_atom (Cells.F90:304-352)
      int cell = 0;
      if (P.use_cells) {
        int c0 = (int)(s.px[o] / P.cell[0]) + 1;
        int c1 = (int)(s.py[o] / P.cell[1]) + 1;
        int c2 = (int)(s.pz[o] / P.cell[2]) + 1;
        if (c0 < 0 || c1 < 0 || c2 < 0 || c0 > P.ncells[0] || c1 > P.ncells[1] ||
            c2 > P.ncells[2]) { fail = ERR_UNSUP; break; }
        cell = cell_lin(P, c0, c1, c2);
      }
      const int bso = s.bkey[o];
      // every thread has read the tombstones, invl[p] and the atom: only now
      // may thread 0 change them (a wave still reading would otherwise see
      // the tombstone it is about to step over, or the shifted list)
      __syncthreads();
      if (tid == 0) {
        invl[p] = -1;
        {
          int q = idel;  // keep the tombstones sorted
          while (q > 0 && sh_tomb[q - 1] > p) {
            sh_tomb[q] = sh_tomb[q - 1];
            q--;
          }
          sh_tomb[q] = p;
        }
        // LIFO detach: gcmc, then hs (ref, b), then sys (Groups.F90:433-467).
        // The gcmc entry of a destroyed atom is dead by its flags; outside the
        // limbo its slot may come back in this call, so mark the entry too
        if (!listed) {
          int gq = tail;
          while (gq >= head && gc[gq] != o) gq--;
          if (gq >= head) gc[gq] = -1;
        }
        rord[s.ref_pos[o]] = -1;
        s.ref_pos[o] = -1;
        s.b2slot[poff + bso] = -1;
        if (listed) s.nn[o] = 0;
        // a destroyed atom keeps its skip value (atom_destroy, Groups.F90:433-467)
        s.flags[o] = (uint8_t)((listed ? DML_FL_LIMBO : 0) | (s.flags[o] & DML_FL_SKIP));
        // the freed hs%b index, kept sorted (b_scan may have passed it)
        if (nfreed_b < GC_TRACK) {
          int q = nfreed_b;
          while (q > 0 && sh_freed_b[q - 1] > bso) {
            sh_freed_b[q] = sh_freed_b[q - 1];
            q--;
          }
          sh_freed_b[q] = bso;
        }
      }
      {
        // its entry leaves the cell's chain: found and closed by the block
        const int64_t base = P.boff + (int64_t)cell * P.bcap;
        const int cnt = min(s.cell_cnt[P.coff + cell], P.bcap);
        int cand = 0x7fffffff;
        for (int q = tid; q < cnt; q += GC_BLOCK)
          if (s.bucket[base + q] == o) cand = min(cand, q);
        const int at = blk_min(cand, sh);
        // 'Unsort cgroup particle fail' (Cells.F90:350)
        if (at == 0x7fffffff) { fail = ERR_UNSUP; break; }
        blk_bucket_shift(s, base, at, cnt, 0);
        if (tid == 0) s.cell_cnt[P.coff + cell] = cnt - 1;
      }
      __syncthreads();
      if (nfreed_b < GC_TRACK) {
        nfreed_b++;
      } else {  // list full: scan the whole index range again
        nfreed_b = 0;
        b_scan = 0;
      }
      // a slot freed outside the limbo is free again for this call's scans
      if (!listed) hs_scan = min(hs_scan, o - poff);
      while (tail >= head && (gc[tail] < 0 || !(s.flags[gc[tail]] & DML_FL_GCMC))) tail--;
      b_nat--;
      nat--;
      if (listed) nlimbo++;
      n_del++;
      idel++;
    }
  }
  __syncthreads();
  // the gcmc and ref lists are compacted by k_gc_count/scan/scatter/copy
  if (tid == 0) {
    Pw->gc_len = glen;
    Pw->ref_len = rlen;
    Pw->gc_pending = 2;
    if (n_ins || n_del) Pw->ovn_stamp = 0;  // atoms added or removed
    Pw->free_lo = hs_scan;
    Pw->bfree_lo = nfreed_b ? min(b_scan, sh_freed_b[0]) : b_scan;
    Pw->nslots = nslots;
    Pw->nat = nat;
    Pw->nlimbo = nlimbo;
    Pw->b_amax = b_amax;
    Pw->b_nat = b_nat;
    Pw->pid_next = pid_next;
    Pw->seq_next = seq_next;
    Pw->gcmc_calls = call + 1;
    if (n_ins) atomicAdd(&s.cnt[r].ins, n_ins);
    if (n_del) atomicAdd(&s.cnt[r].del, n_del);
    if (fail) rep_fail(s, r, fail);
  }
}

// ------------------------------------------------------------------------
// GCMC in batches (DML_GCMC_BATCHED): the nadj trials of one gcmc_run call
// (dana.F90:590-713) are evaluated in parallel and committed in trial order
// with the result of the sequential k_gcmc_run, bit for bit.
//
// Within a call the positions do not move; only the set of atoms changes.
// So every trial's point, its hard-core neighbours among the atoms present
// at the call's start and the earlier trial points within rcut can be found
// in parallel (k_gb_trials, k_gb_conflicts).  What stays sequential is the
// count n of the control volume and which atoms are present: one wave walks
// the trials (k_gb_decide) with the in-volume list in LDS as a bitmap
// (deletions take the m-th present entry, insertions append), and a second
// wave walks the accepted moves (k_gb_commit) for the hs%b indices (lowest
// free first, indices freed in this call included), the template and the
// list tail that the reference's insertion reads, and the ordered stores of
// flags, lists and index maps.  Slots (first free slot; a deleted atom goes
// to the limbo, so deletions free none in this call), rows and cell chains
// then follow in parallel: an inserted atom's own row is the stencil scan
// at its insertion time (the cell chains then hold the atoms inserted
// before it, newest first, then the atoms of the call's start not yet
// deleted), the rows it joins get it appended in insertion order, and each
// touched cell's chain is rebuilt once.
//
// A call the batch cannot take exactly (Verlet mode, no rows, more than
// DML_GCMC_BATCH_MAX trials, an in-volume list beyond the LDS bitmap, a
// gcmc list that could empty, more hard-core neighbours or nearby earlier
// points than kept, a bucket or row that could overflow, an atom outside its
// cells) is left to k_gcmc_run, which gives the same result one trial at a
// time; the decision is taken before anything is written.
// ------------------------------------------------------------------------
#define GB_NB 16                     // hard-core neighbours kept per point
#define GB_NC 16                     // earlier points within rcut kept per point
#define GB_VWORDS (96 * 1024 / 4)    // in-volume bitmap words in LDS (786,432 entries)
#define GB_CHUNKS 768                // its 1024-entry chunks
#define GB_UW 12                     // hs%b universe bitmap words per lane (24,576 entries)
#define GB_CELLMAX 512               // chain length a rebuilt cell may reach
#define GB_NMAX (GB_VWORDS * 32)     // bound of the count n a call can reach (the bitmap's)

__device__ inline bool gb_on(const RepParams &P) { return P.gb_state == GB_ON && !P.halted; }
__device__ inline void gb_fallback(RepParams *Pw) { Pw->gb_state = GB_FALLBACK; }

// the call's sizes held in registers by the ordered walks (RepParams lives
// in global memory: read through a reference after the walk's own stores,
// every field access would wait for those stores)
struct GbView {
  int poff, off, nF, nslots, cap, T;
};
__device__ inline GbView gb_view(const RepParams &P) {
  GbView v;
  v.poff = P.poff;
  v.off = P.gb_off;
  v.nF = P.gb_nF;
  v.nslots = P.gb_nslots;
  v.cap = P.gb_cap;
  v.T = P.gb_t;
  return v;
}
__device__ inline int gb_ins_slot_v(const DevState &s, const GbView &v, int k) {
  return v.poff + (k < v.nF ? s.gb_free[v.off + k] : v.nslots + (k - v.nF));
}
__device__ inline int gb_del_time_v(const DevState &s, const GbView &v, int o) {
  const int t = s.gb_delt[o];
  if (t < 0 || t >= v.T || s.gb_slot[v.off + t] != o || s.gb_kind[v.off + t] != 1)
    return 0x7fffffff;
  return t;
}

// slot (global) of the k-th insertion of the call: the free slots in
// ascending order, then appended slots (Groups.F90:1083-1093)
__device__ inline int gb_ins_slot(const DevState &s, const RepParams &P, int k) {
  return P.poff + (k < P.gb_nF ? s.gb_free[P.gb_off + k] : P.gb_nslots + (k - P.gb_nF));
}
// insertion rank of an inserted slot (inverse of gb_ins_slot)
__device__ inline int gb_ins_rank(const DevState &s, const RepParams &P, int i) {
  const int loc = i - P.poff;
  if (loc >= P.gb_nslots) return P.gb_nF + (loc - P.gb_nslots);
  int lo = 0, hi = min(P.gb_nF, P.gb_cap) - 1;
  const int32_t *f = s.gb_free + P.gb_off;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (f[mid] < loc) lo = mid + 1; else hi = mid;
  }
  return lo;
}
// trial at which atom o (present at the call's start) was deleted, or
// INT_MAX; gb_delt is validated against the trial's own record
__device__ inline int gb_del_time(const DevState &s, const RepParams &P, int o) {
  const int t = s.gb_delt[o];
  if (t < 0 || t >= P.gb_t || s.gb_slot[P.gb_off + t] != o || s.gb_kind[P.gb_off + t] != 1)
    return 0x7fffffff;
  return t;
}
// in-volume list position of atom o, or -1 (validated against the list)
__device__ inline int gb_vpos_of(const DevState &s, const RepParams &P, int o) {
  const int v = s.gb_vpos[o];
  if (v < 0 || v >= P.gb_ilen || s.ref_tmp[P.poff + v] != o) return -1;
  return v;
}
// counters and list heads tagged with the call's stamp: a value of another
// call reads as zero / empty, so nothing needs clearing between calls
__device__ inline int gb_stamped_inc(unsigned long long *a, uint32_t stamp) {
  unsigned long long old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    const uint32_t cnt = (uint32_t)(old >> 32) == stamp ? (uint32_t)old : 0u;
    const unsigned long long nw = ((unsigned long long)stamp << 32) | (cnt + 1u);
    const unsigned long long got = atomicCAS(a, old, nw);
    if (got == old) return (int)cnt;
    old = got;
  }
}
__device__ inline int gb_stamped_get(const unsigned long long *a, uint32_t stamp) {
  const unsigned long long v = *a;
  return (uint32_t)(v >> 32) == stamp ? (int)(uint32_t)v : 0;
}
__device__ inline int gb_list_head(const DevState &s, const RepParams &P, int cell) {
  const unsigned long long v = s.gb_chead[P.coff + cell];
  return (uint32_t)(v >> 32) == P.gb_stamp ? (int)(uint32_t)v : -1;
}

// every GCMC replica of the call: batched or not, and the sizes at its start
__global__ void k_gb_begin(DevState s, int rep_only, double act_in, int nadj_in) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= s.nrep || (rep_only >= 0 && r != rep_only)) return;
  RepParams &P = s.rp[r];
  if (rep_only < 0 && P.reservoir != DML_GCMC) return;
  P.gb_state = GB_OFF;
  if (!(P.gcmc_mode & DML_GCMC_BATCHED) || P.halted || P.gc_pending != 1) return;
  const int nadj = rep_only >= 0 ? nadj_in : P.nadj;
  P.gb_t = nadj;
  P.gb_act = rep_only >= 0 ? act_in : P.act;
  P.gb_ilen = P.gc_inv;
  P.gb_glen = P.gc_len;
  P.gb_rlen = P.nref;
  P.gb_nslots = P.nslots;
  P.gb_bamax = P.b_amax;
  P.gb_nF = P.gb_nFB = 0;
  P.gb_nins = P.gb_ndel = P.gb_nalist = 0;
  P.gb_stamp = (uint32_t)P.gcmc_calls + 1u;
  const bool ok = P.use_cells && P.listed && nadj > 0 && nadj <= P.gb_cap &&
                  (int64_t)P.gc_inv + nadj <= (int64_t)GB_VWORDS * 32 && P.gc_len >= nadj + 2;
  P.gb_state = ok ? GB_ON : GB_FALLBACK;
}

// per particle block of the replica: the in-volume positions of its atoms,
// and counts of free slots (below nslots) and free hs%b indices (below b_amax)
__device__ inline void gb_free_flags(const DevState &s, const RepParams &P, int q, int *kf,
                                     int *kb) {
  const int gq = P.poff + q;
  *kf = q < P.gb_nslots && !(s.flags[gq] & (DML_FL_ALIVE | DML_FL_LIMBO));
  *kb = q < P.gb_bamax && s.b2slot[gq] < 0;
}
__global__ void __launch_bounds__(DML_BLOCK) k_gb_count(DevState s, int r) {
  __shared__ int w1[DML_BLOCK / WAVE], w2[DML_BLOCK / WAVE];
  const RepParams &P = s.rp[r];
  if (!gb_on(P)) return;
  const int q = blockIdx.x * DML_BLOCK + threadIdx.x;
  if (q < P.gb_ilen) s.gb_vpos[s.ref_tmp[P.poff + q]] = q;
  int kf, kb;
  gb_free_flags(s, P, q, &kf, &kb);
  const unsigned long long m1 = __ballot(kf), m2 = __ballot(kb);
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  if (lane == 0) {
    w1[w] = __popcll(m1);
    w2[w] = __popcll(m2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int a = 0, b = 0;
    for (int k = 0; k < DML_BLOCK / WAVE; k++) {
      a += w1[k];
      b += w2[k];
    }
    const int gb = P.poff / DML_BLOCK + blockIdx.x;
    s.gb_part[2 * gb] = a;
    s.gb_part[2 * gb + 1] = b;
  }
}
__global__ void __launch_bounds__(SCAN_BLOCK) k_gb_scan(DevState s, int r) {
  RepParams *Pw = s.rp + r;
  if (!gb_on(*Pw)) return;
  int c1, c2;
  rep_exscan2(s.gb_part, 2, Pw->poff / DML_BLOCK, Pw->pcap / DML_BLOCK, &c1, &c2);
  if (threadIdx.x == 0) {
    Pw->gb_nF = c1;
    Pw->gb_nFB = c2;
  }
}
__global__ void __launch_bounds__(DML_BLOCK) k_gb_scatter(DevState s, int r) {
  __shared__ int w1[DML_BLOCK / WAVE], w2[DML_BLOCK / WAVE];
  const RepParams &P = s.rp[r];
  if (!gb_on(P)) return;
  const int q = blockIdx.x * DML_BLOCK + threadIdx.x;
  int kf, kb;
  gb_free_flags(s, P, q, &kf, &kb);
  const unsigned long long m1 = __ballot(kf), m2 = __ballot(kb);
  const int lane = threadIdx.x % WAVE, w = threadIdx.x / WAVE;
  if (lane == 0) {
    w1[w] = __popcll(m1);
    w2[w] = __popcll(m2);
  }
  __syncthreads();
  const int gb = P.poff / DML_BLOCK + blockIdx.x;
  int o1 = s.gb_part[2 * gb], o2 = s.gb_part[2 * gb + 1];
  for (int k = 0; k < w; k++) {
    o1 += w1[k];
    o2 += w2[k];
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  const int a = o1 + __popcll(m1 & below), b = o2 + __popcll(m2 & below);
  if (kf && a < P.gb_cap) s.gb_free[P.gb_off + a] = q;
  if (kb && b < P.gb_cap) s.gb_freeb[P.gb_off + b] = q;
}

// one wave per trial: kind, uniforms, and for an insertion its point, cell,
// hard-core neighbours (gcmc atoms closer than rcut, dana.F90:638-653) and
// FP64 energy (the same sum as k_gcmc_energy)
__global__ void __launch_bounds__(DML_BLOCK) k_gb_trials(DevState s, int r, const double *u,
                                                          const double *g) {
  const int w = (blockIdx.x * DML_BLOCK + threadIdx.x) / WAVE, lane = threadIdx.x % WAVE;
  const RepParams &P = s.rp[r];
  if (!gb_on(P) || w >= P.gb_t) return;
  const int call = P.gcmc_calls;
  const int o = P.gb_off + w;
  const int kind = (P.gcmc_mode & DML_GCMC_ALTERNATE) ? (w & 1)
                                                      : !(gc_unif(P, u, w, 0, call) < 0.5);
  const double u1 = gc_unif(P, u, w, 1, call), u2 = gc_unif(P, u, w, 2, call);
  if (lane == 0) {
    // the Metropolis test of the ordered walk as a threshold on n, found
    // by bisection over the same FP64 expressions (each is monotone in n):
    // an insertion passes while n <= thr (dana.F90:632, !(act V/(n+1) <
    // ran)), a deletion once n >= thr (:683, !(n/(V act) < ran))
    const double act = P.gb_act, vol = P.box[0] * P.box[1] * (P.zmax - P.z0);
    int lo = -1, hi = GB_NMAX + 1;
    while (hi - lo > 1) {
      const int mid = lo + (hi - lo) / 2;
      const bool pass = kind ? !((double)mid / (vol * act) < u1) : !(act * vol / (double)(mid + 1) < u1);
      if (pass == !kind) lo = mid; else hi = mid;
    }
    s.gb_u[2 * o] = (double)(kind ? hi : lo);
    s.gb_u[2 * o + 1] = u2;
    s.gb_nblk[o] = 0;
    s.gb_du[o] = 0.0;
  }
  if (kind) {
    if (lane == 0) s.gb_kind[o] = 1;
    return;
  }
  // the normals the velocity of an accepted insertion takes (dana.F90:666-668)
  if (lane < 3) s.gb_g[3 * o + lane] = gc_gauss(P, g, w, lane, call);
  const double rx = u2 * P.box[0];
  const double ry = gc_unif(P, u, w, 3, call) * P.box[1];
  const double rz = gc_unif(P, u, w, 4, call) * (P.zmax - P.z0) + P.z0;
  const int ci0 = (int)(rx / P.cell[0]) + 1, ci1 = (int)(ry / P.cell[1]) + 1;
  const int ci2 = (int)(rz / P.cell[2]) + 1;
  if (lane == 0) {
    s.gb_pt[3 * o] = rx;
    s.gb_pt[3 * o + 1] = ry;
    s.gb_pt[3 * o + 2] = rz;
  }
  if (ci0 < 0 || ci1 < 0 || ci2 < 0 || ci0 > P.ncells[0] + 1 || ci1 > P.ncells[1] + 1 ||
      ci2 > P.ncells[2] + 1) {
    if (lane == 0) s.gb_kind[o] = 2;
    return;
  }
  // the near cells as one stream (lane k: count and bucket start of cell
  // k, loaded at once; a wave prefix sum), 64 entries a pass, in cell order
  int ncell = 0;
  const int mycell = near_cell_at(P, ci0, ci1, ci2, lane < 27 ? lane : 0, &ncell);
  int kcnt = 0;
  int64_t kbase = 0;
  if (lane < ncell) {
    kbase = P.boff + (int64_t)mycell * P.bcap;
    kcnt = min(s.cell_cnt[P.coff + mycell], P.bcap);
  }
  const int kincl = wave_scan_incl(kcnt);
  const int total = __builtin_amdgcn_readlane(kincl, WAVE - 1);
  const double rc2 = P.rcut * P.rcut;
  double e = 0.0;
  int nb = 0, perm = 0;
  {
    for (int s0 = 0; s0 < total; s0 += WAVE) {
      const int sp = s0 + lane;
      int kc = 0;
      for (int c = 0; c < ncell; c++) kc += __builtin_amdgcn_readlane(kincl, c) <= sp;
      const int kk = min(kc, WAVE - 1);
      const int64_t base = ((int64_t)__shfl((int)(kbase >> 32), kk) << 32) |
                           (uint32_t)__shfl((int)(kbase & 0xffffffff), kk);
      const int q = sp - (__shfl(kincl, kk) - __shfl(kcnt, kk));
      int hit = 0, vp = -1;
      if (sp < total) {
        const int j = s.bucket[base + q];
        double dx, dy, dz;
        vdist_mic(P, rx, ry, rz, s.px[j], s.py[j], s.pz[j], &dx, &dy, &dz);
        const double dr2 = dot3(dx, dy, dz);
        const int m = s.species[j];
        const double r0 = P.r0[m - 1], eps = P.eps[m - 1];  // pair (Li, m)
        if (!(dr2 > r0 * r0)) {
          const double b = pow6d(r0) / pow6d(sqrt(dr2));
          e += eps * b * (b - 2) + eps;
        }
        if (s.flags[j] & DML_FL_GCMC) {
          double hx, hy, hz;  // as the trial loop tests it: against the bucket copy
          vdist_mic(P, rx, ry, rz, s.bx[base + q], s.by[base + q], s.bz[base + q], &hx, &hy, &hz);
          if (dot3(hx, hy, hz) < rc2) {
            hit = 1;
            vp = gb_vpos_of(s, P, j);
          }
        }
      }
      perm |= __ballot(hit && vp < 0) != 0;
      const unsigned long long hm = __ballot(hit);
      if (hit) {
        const int at = nb + __popcll(hm & ((1ull << lane) - 1ull));
        if (at < GB_NB) s.gb_blk[(int64_t)o * GB_NB + at] = vp;
      }
      nb += __popcll(hm);
    }
  }
  for (int off = WAVE / 2; off > 0; off >>= 1) e += __shfl_xor(e, off);
  if (lane == 0) {
    s.gb_kind[o] = 0;
    s.gb_cell[o] = cell_lin(P, ci0, ci1, ci2);
    s.gb_nblk[o] = perm ? -1 : (nb > GB_NB ? GB_NB + 1 : nb);
    s.gb_du[o] = e;
  }
}

// one wave per insertion trial: the earlier insertion points within rcut
// (in the reference the earlier atom, once inserted, sits in a cell chain)
__global__ void __launch_bounds__(DML_BLOCK) k_gb_conflicts(DevState s, int r) {
  const int w = (blockIdx.x * DML_BLOCK + threadIdx.x) / WAVE, lane = threadIdx.x % WAVE;
  const RepParams &P = s.rp[r];
  if (!gb_on(P) || w >= P.gb_t) return;
  const int o = P.gb_off + w;
  if (s.gb_kind[o] != 0) {
    if (lane == 0) s.gb_ncf[o] = 0;
    return;
  }
  const double rx = s.gb_pt[3 * o], ry = s.gb_pt[3 * o + 1], rz = s.gb_pt[3 * o + 2];
  const double rc2 = P.rcut * P.rcut;
  int nc = 0;
  for (int t0 = 0; t0 < w; t0 += WAVE) {
    const int tp = t0 + lane;
    int hit = 0;
    if (tp < w && s.gb_kind[P.gb_off + tp] == 0) {
      const int op = P.gb_off + tp;
      double dx, dy, dz;
      vdist_mic(P, rx, ry, rz, s.gb_pt[3 * op], s.gb_pt[3 * op + 1], s.gb_pt[3 * op + 2], &dx,
                &dy, &dz);
      hit = dot3(dx, dy, dz) < rc2;
    }
    const unsigned long long hm = __ballot(hit);
    if (hit) {
      const int at = nc + __popcll(hm & ((1ull << lane) - 1ull));
      if (at < GB_NC) s.gb_cf[(int64_t)o * GB_NC + at] = tp;
    }
    nc += __popcll(hm);
  }
  if (lane == 0) s.gb_ncf[o] = nc > GB_NC ? GB_NC + 1 : nc;
}


// the decisions, one wave walking the trials in order (dana.F90:621-710):
// n, the Metropolis tests, the hard-core test against the neighbours still
// present and the earlier points accepted and still present, and for a
// deletion the m-th present entry of the in-volume list (list order, the
// call's insertions appended).  The list is a bitmap in LDS with per-chunk
// counts (1024 entries a chunk, 16 chunks a lane, lane totals in registers).
// Nothing of the replica's state is written here.
__global__ void __launch_bounds__(WAVE) k_gb_decide(DevState s, int r) {
  __shared__ uint32_t vb[GB_VWORDS];
  __shared__ int32_t ccnt[WAVE * 16];
  __shared__ int32_t res[DML_GCMC_BATCH_MAX];
  // the neighbour lists of 64 trials, double-buffered: [buffer][trial][entry]
  __shared__ int32_t lblk[2][WAVE][GB_NB], lcf[2][WAVE][GB_NC];
  RepParams *Pw = s.rp + r;
  if (!gb_on(*Pw)) return;
  const int lane = threadIdx.x;
  const int T = Pw->gb_t, ilen = Pw->gb_ilen, off = Pw->gb_off;
  const int nwords = (ilen + T + 31) / 32;
  for (int q = lane; q < nwords; q += WAVE) {
    const int b0 = q * 32;
    vb[q] = b0 + 32 <= ilen ? 0xffffffffu : (b0 < ilen ? (1u << (ilen - b0)) - 1u : 0u);
  }
  int tot = 0;  // entries present in this lane's 16 chunks
  for (int k = 0; k < 16; k++) {
    const int c = 16 * lane + k;
    const int v = min(max(ilen - 1024 * c, 0), 1024);
    ccnt[c] = v;
    tot += v;
  }
  // and in the chunks of lanes 0 .. lane (kept up to date, not rescanned)
  int pre = wave_scan_incl(tot);
  // each trial's result (insertion rank, in-volume position, -1): the walk
  // writes it to LDS only; the per-trial and per-move records in memory are
  // written from it after the walk, 64 trials at a time
  for (int q = lane; q < T; q += WAVE) res[q] = -1;
  // per-trial data of a 64-trial chunk and its lists, loaded into
  // registers one chunk ahead and staged into LDS at the chunk's start, so
  // that the walk reads LDS and never waits on memory (a load result kept
  // in a register across the loop would make every iteration wait for the
  // iteration's own stores too); the walk is one wave, whose LDS accesses
  // execute in program order
  __shared__ int32_t lkind[2][WAVE], lnblk[2][WAVE], lncf[2][WAVE];
  __shared__ double lu1[2][WAVE], lu2[2][WAVE];
  int p_kind = 0, p_nblk = 0, p_ncf = 0;
  double p_u1 = 0.0, p_u2 = 0.0;
  int pb[GB_NB], pc[GB_NC];
  auto load_chunk = [&](int t0) {
    const int q = t0 + lane;
    if (q < T) {
      p_kind = s.gb_kind[off + q];
      p_nblk = s.gb_nblk[off + q];
      p_ncf = s.gb_ncf[off + q];
      p_u1 = s.gb_u[2 * (off + q)];
      p_u2 = s.gb_u[2 * (off + q) + 1];
    }
    // the lists of the chunk's 64 trials, 16 entries each: coalesced
#pragma unroll
    for (int j = 0; j < GB_NB; j++) {
      const int64_t e = (int64_t)(off + t0) * GB_NB + j * WAVE + lane;
      pb[j] = t0 + (j * WAVE + lane) / GB_NB < T ? s.gb_blk[e] : 0;
    }
#pragma unroll
    for (int j = 0; j < GB_NC; j++) {
      const int64_t e = (int64_t)(off + t0) * GB_NC + j * WAVE + lane;
      pc[j] = t0 + (j * WAVE + lane) / GB_NC < T ? s.gb_cf[e] : 0;
    }
  };
  auto stage_chunk = [&](int buf) {
    lkind[buf][lane] = p_kind;
    lnblk[buf][lane] = p_nblk;
    lncf[buf][lane] = p_ncf;
    lu1[buf][lane] = p_u1;
    lu2[buf][lane] = p_u2;
#pragma unroll
    for (int j = 0; j < GB_NB; j++) {
      const int e = j * WAVE + lane;
      lblk[buf][e / GB_NB][e % GB_NB] = pb[j];
    }
#pragma unroll
    for (int j = 0; j < GB_NC; j++) {
      const int e = j * WAVE + lane;
      lcf[buf][e / GB_NC][e % GB_NC] = pc[j];
    }
  };
  load_chunk(0);
  __syncthreads();
  int n = ilen, nins = 0, ndel = 0;
  int fail = 0;
  for (int t = 0; t < T && !fail; t++) {
    const int a = t & (WAVE - 1);
    const int buf = (t / WAVE) & 1;
    if (a == 0) {
      stage_chunk(buf);
      if (t + WAVE < T) load_chunk(t + WAVE);
      __builtin_amdgcn_wave_barrier();
    }
    const int kind = lkind[buf][a];
    const int nblk = lnblk[buf][a];
    const int ncf = lncf[buf][a];
    const int thr = (int)lu1[buf][a];  // the Metropolis threshold on n
    const double u2 = lu2[buf][a];
    int out = -1;
    if (kind == 2) {
      fail = 1;  // a point outside the cells: the sequential kernel's error path
    } else if (kind == 0) {
      if (n <= thr && nblk >= 0) {
        if (nblk > GB_NB || ncf > GB_NC) {
          fail = 1;
        } else {
          int here = 0;
          if (lane < nblk) {
            const int vp = lblk[buf][a][lane];
            here = (vb[vp >> 5] >> (vp & 31)) & 1;
          } else if (lane >= GB_NB && lane < GB_NB + ncf) {
            const int k = res[lcf[buf][a][lane - GB_NB]];  // an earlier insertion: accepted, still present?
            if (k >= 0) {
              const int vp = ilen + k;
              here = (vb[vp >> 5] >> (vp & 31)) & 1;
            }
          }
          if (!__ballot(here)) {
            const int k = nins++;
            const int vp = ilen + k;
            if (lane == 0) {
              vb[vp >> 5] |= 1u << (vp & 31);
              ccnt[vp >> 10]++;
            }
            if (lane == (vp >> 14)) tot++;
            if (lane >= (vp >> 14)) pre++;
            n++;
            out = k;
          }
        }
      }
    } else {
      if (n >= thr) {
        int m = (int)floor(u2 * n) + 1;
        if (m > n) m = n;
        if (m < 1) {
          fail = 1;  // nothing to delete: the sequential kernel reports it
        } else {
          // the (m-1)-th present entry: lane, chunk, word, bit (each level's
          // ballot is non-empty while the counts agree with the bitmap)
          int rem = m - 1;
          const unsigned long long b1 = __ballot(pre > rem);
          const int L = b1 ? __builtin_ctzll(b1) : 0;
          rem -= __builtin_amdgcn_readlane(pre - tot, L);
          const int c16 = lane < 16 ? ccnt[16 * L + lane] : 0;
          const int incl2 = row_scan16(c16);
          const unsigned long long b2 = __ballot(lane < 16 && incl2 > rem);
          const int j = b2 ? __builtin_ctzll(b2) : 0;
          rem -= __builtin_amdgcn_readlane(incl2 - c16, j);
          const int chunk = 16 * L + j;
          const uint32_t wv = lane < 32 ? vb[32 * chunk + lane] : 0u;
          const int pc1 = __popc(wv);
          const int incl3 = half_scan32(pc1);
          const unsigned long long b3 = __ballot(lane < 32 && incl3 > rem);
          const int k = b3 ? __builtin_ctzll(b3) : 0;
          rem -= __builtin_amdgcn_readlane(incl3 - pc1, k);
          const uint32_t word = (uint32_t)__builtin_amdgcn_readlane((int)wv, k);
          const unsigned long long b4 = __ballot(
              lane < 32 && ((word >> (lane & 31)) & 1u) &&
              __popc(word & ((1u << (lane & 31)) - 1u)) == rem);
          const int bit = b4 ? __builtin_ctzll(b4) : 0;
          if (!b1 || !b2 || !b3 || !b4 || chunk >= GB_CHUNKS) {
            fail = 1;  // cannot happen; leave the call to the sequential kernel
            break;
          }
          const int p = chunk * 1024 + k * 32 + bit;
          if (lane == 0) {
            vb[32 * chunk + k] = word & ~(1u << bit);
            ccnt[chunk]--;
          }
          if (lane == L) tot--;
          if (lane >= L) pre--;
          ndel++;
          n--;
          out = p;
        }
      }
    }
    if (lane == 0) res[t] = out;
    __builtin_amdgcn_wave_barrier();
  }
  // the records of the walk, in trial order: every trial's result, each
  // insertion's trial (its rank is its result), each deletion's trial and
  // position in deletion order; then, once those are written, the trial
  // that deletes an insertion of this call
  int kd = 0;
  for (int t0 = 0; t0 < T; t0 += WAVE) {
    const int t = t0 + lane;
    const int o = t < T ? res[t] : -1;
    const int kind = t < T ? s.gb_kind[off + t] : 2;
    if (t < T) {
      s.gb_res[off + t] = o;
      s.gb_slot[off + t] = -1;
    }
    if (o >= 0 && kind == 0) {
      s.gi_trial[off + o] = t;
      s.gi_del[off + o] = 0x7fffffff;
    }
    const bool del = o >= 0 && kind == 1;
    const unsigned long long dm = __ballot(del);
    if (del) {
      const int d = kd + __popcll(dm & ((1ull << lane) - 1ull));
      s.gd_trial[off + d] = t;
      s.gd_vpos[off + d] = o;
    }
    kd += __popcll(dm);
  }
  __syncthreads();
  for (int t = lane; t < T; t += WAVE) {
    const int o = res[t];
    if (o >= ilen && s.gb_kind[off + t] == 1) s.gi_del[off + (o - ilen)] = t;
  }
  if (lane == 0) {
    Pw->gb_nins = nins;
    Pw->gb_ndel = ndel;
    // slots appended beyond the free ones, and the gcmc and ref list
    // entries appended, must fit the replica (the sequential kernel
    // compacts those lists instead)
    if (fail || Pw->gb_nslots + max(0, nins - Pw->gb_nF) > Pw->pcap ||
        Pw->gb_glen + nins > Pw->pcap || Pw->gb_rlen + nins > Pw->pcap)
      gb_fallback(Pw);
  }
}

// accepted insertions: each cell's list of the call's insertions (heads
// stamped with the call, the order of the list is not used)
__global__ void __launch_bounds__(DML_BLOCK) k_gb_lists(DevState s, int r) {
  const int k = blockIdx.x * DML_BLOCK + threadIdx.x;
  const RepParams &P = s.rp[r];
  if (!gb_on(P) || k >= P.gb_nins) return;
  const int t = s.gi_trial[P.gb_off + k];
  const int cell = s.gb_cell[P.gb_off + t];
  const unsigned long long old =
      atomicExch(&s.gb_chead[P.coff + cell], ((unsigned long long)P.gb_stamp << 32) | (uint32_t)k);
  s.gi_next[P.gb_off + k] = (uint32_t)(old >> 32) == P.gb_stamp ? (int)(uint32_t)old : -1;
}

// capacity checks before anything is written, one wave per accepted
// insertion, each an upper bound of what the sequential trials can reach:
// the bucket of its cell (every insertion into it, no deletion counted),
// its own row (every atom of the call's start and every accepted point in
// its stencil within the list radius), and the rows it joins (each ref atom
// within the list radius gets at most one entry per accepted point)
__global__ void __launch_bounds__(DML_BLOCK) k_gb_check(DevState s, int r) {
  const int w = (blockIdx.x * DML_BLOCK + threadIdx.x) / WAVE, lane = threadIdx.x % WAVE;
  RepParams *Pw = s.rp + r;
  const RepParams &P = *Pw;
  if (!gb_on(P) || w >= P.gb_nins) return;
  const int off = P.gb_off;
  const int t = s.gi_trial[off + w];
  const int cell = s.gb_cell[off + t];
  const double rx = s.gb_pt[3 * (off + t)], ry = s.gb_pt[3 * (off + t) + 1];
  const double rz = s.gb_pt[3 * (off + t) + 2];
  const int maxnn = s.maxnn;
  int bad = 0;
  if (lane == 0) {
    const int before = gb_stamped_inc(&s.gb_ccnt[P.coff + cell], P.gb_stamp * 2u);
    const int cap = min(P.bcap, GB_CELLMAX);
    if (min(s.cell_cnt[P.coff + cell], P.bcap) + before + 1 > cap) bad = 1;
  }
  const int ci0 = cell % (P.ncells[0] + 2);
  const int ci1 = (cell / (P.ncells[0] + 2)) % (P.ncells[1] + 2);
  const int ci2 = cell / ((P.ncells[0] + 2) * (P.ncells[1] + 2));
  // its own row: the 27 stencil cells as one stream (lane k: count and
  // bucket start of stencil cell k, loaded at once; a wave prefix sum), and
  // each lane walks the list of the call's insertions of its stencil cell
  int rowcnt = 0;
  {
    int kcnt = 0, mycc = 0;
    int64_t kbase = 0;
    if (lane < 27) {
      const int n0 = cell_wrap(ci0 + kMap[lane][0], P.ncells[0]);
      const int n1 = cell_wrap(ci1 + kMap[lane][1], P.ncells[1]);
      const int n2 = cell_wrap(ci2 + kMap[lane][2], P.ncells[2]);
      mycc = cell_lin(P, n0, n1, n2);
      kbase = P.boff + (int64_t)mycc * P.bcap;
      kcnt = min(s.cell_cnt[P.coff + mycc], P.bcap);
    }
    const int kincl = wave_scan_incl(kcnt);
    const int total = __builtin_amdgcn_readlane(kincl, WAVE - 1);
    for (int s0 = 0; s0 < total; s0 += WAVE) {
      const int sp = s0 + lane;
      int kc = 0;
      for (int c = 0; c < 27; c++) kc += __builtin_amdgcn_readlane(kincl, c) <= sp;
      const int kk = min(kc, WAVE - 1);
      const int64_t base = ((int64_t)__shfl((int)(kbase >> 32), kk) << 32) |
                           (uint32_t)__shfl((int)(kbase & 0xffffffff), kk);
      const int q = sp - (__shfl(kincl, kk) - __shfl(kcnt, kk));
      if (sp < total) {
        double dx, dy, dz;
        vdist_mic(P, s.bx[base + q], s.by[base + q], s.bz[base + q], rx, ry, rz, &dx, &dy, &dz);
        rowcnt += !(dot3(dx, dy, dz) > P.rl2);
      }
    }
    if (lane < 27) {
      for (int kk = gb_list_head(s, P, mycc); kk >= 0; kk = s.gi_next[off + kk]) {
        const int tp = s.gi_trial[off + kk];
        if (tp != t) {
          double dx, dy, dz;
          vdist_mic(P, s.gb_pt[3 * (off + tp)], s.gb_pt[3 * (off + tp) + 1],
                    s.gb_pt[3 * (off + tp) + 2], rx, ry, rz, &dx, &dy, &dz);
          rowcnt += !(dot3(dx, dy, dz) > P.rl2);
        }
      }
    }
  }
  for (int o2 = WAVE / 2; o2 > 0; o2 >>= 1) rowcnt += __shfl_xor(rowcnt, o2);
  if (rowcnt > maxnn) bad = 1;
  // the rows it joins: the near cells as one stream
  {
    int ncell = 0;
    const int mycell = near_cell_at(P, ci0, ci1, ci2, lane < 27 ? lane : 0, &ncell);
    int kcnt = 0;
    int64_t kbase = 0;
    if (lane < ncell) {
      kbase = P.boff + (int64_t)mycell * P.bcap;
      kcnt = min(s.cell_cnt[P.coff + mycell], P.bcap);
    }
    const int kincl = wave_scan_incl(kcnt);
    const int total = __builtin_amdgcn_readlane(kincl, WAVE - 1);
    for (int s0 = 0; s0 < total; s0 += WAVE) {
      const int sp = s0 + lane;
      int kc = 0;
      for (int c = 0; c < ncell; c++) kc += __builtin_amdgcn_readlane(kincl, c) <= sp;
      const int kk = min(kc, WAVE - 1);
      const int64_t base = ((int64_t)__shfl((int)(kbase >> 32), kk) << 32) |
                           (uint32_t)__shfl((int)(kbase & 0xffffffff), kk);
      const int q = sp - (__shfl(kincl, kk) - __shfl(kcnt, kk));
      if (sp >= total) continue;
      const int aj = s.bucket[base + q];
      if (!(s.flags[aj] & DML_FL_REF)) continue;
      double dx, dy, dz;
      vdist_mic(P, rx, ry, rz, s.bx[base + q], s.by[base + q], s.bz[base + q], &dx, &dy, &dz);
      if (dot3(dx, dy, dz) > P.rl2) continue;
      const int before = gb_stamped_inc(&s.gb_app[aj], P.gb_stamp * 2u);
      if (s.nn[aj] + before + 1 > maxnn) bad = 1;
    }
  }
  if (__ballot(bad)) {
    if (lane == 0) gb_fallback(Pw);
  }
}

// one wave per accepted deletion: the atom (an atom of the call's start, or
// one inserted earlier in it), and for the former what its removal reads
// (hs%b index, ref list position, flags, the cell holding it: Cells.F90:
// 304-352 fails when it is not there) and its FP64 removal energy
__global__ void __launch_bounds__(DML_BLOCK) k_gb_dels(DevState s, int r) {
  const int w = (blockIdx.x * DML_BLOCK + threadIdx.x) / WAVE, lane = threadIdx.x % WAVE;
  RepParams *Pw = s.rp + r;
  const RepParams &P = *Pw;
  if (!gb_on(P) || w >= P.gb_ndel) return;
  const int off = P.gb_off;
  const int t = s.gd_trial[off + w];
  const int p = s.gd_vpos[off + w];
  if (p >= P.gb_ilen) {
    const int k = p - P.gb_ilen;
    const int o = gb_ins_slot(s, P, k);
    if (lane == 0) {
      s.gd_slot[off + w] = o;
      s.gb_slot[off + t] = o;
      s.gd_bso[off + w] = -1;
      s.gd_cell[off + w] = s.gb_cell[off + s.gi_trial[off + k]];
      s.gb_du[off + t] = s.gb_du[off + s.gi_trial[off + k]];
    }
    return;
  }
  const int o = s.ref_tmp[P.poff + p];
  const double x = s.px[o], y = s.py[o], z = s.pz[o];
  const int c0 = (int)(x / P.cell[0]) + 1, c1 = (int)(y / P.cell[1]) + 1;
  const int c2 = (int)(z / P.cell[2]) + 1;
  if (c0 < 1 || c1 < 1 || c2 < 1 || c0 > P.ncells[0] || c1 > P.ncells[1] || c2 > P.ncells[2]) {
    if (lane == 0) gb_fallback(Pw);
    return;
  }
  const int cell = cell_lin(P, c0, c1, c2);
  {
    const int64_t base = P.boff + (int64_t)cell * P.bcap;
    const int cnt = min(s.cell_cnt[P.coff + cell], P.bcap);
    if (cnt > GB_CELLMAX) {  // a chain longer than k_gb_cells stages
      if (lane == 0) gb_fallback(Pw);
      return;
    }
    int found = 0;
    for (int q = lane; q < cnt; q += WAVE) found |= s.bucket[base + q] == o;
    if (!__ballot(found)) {
      if (lane == 0) gb_fallback(Pw);
      return;
    }
  }
  // removal energy at the call's start (k_gcmc_energy's sum without the atom)
  int cells[27];
  const int ncell = near_cells(P, c0, c1, c2, cells);
  double e = 0.0;
  for (int k = 0; k < ncell; k++) {
    const int cc = cells[k];
    int cnt = s.cell_cnt[P.coff + cc];
    if (cnt > P.bcap) cnt = P.bcap;
    const int64_t base = P.boff + (int64_t)cc * P.bcap;
    for (int q = lane; q < cnt; q += WAVE) {
      const int j = s.bucket[base + q];
      if (j == o) continue;
      double dx, dy, dz;
      vdist_mic(P, x, y, z, s.px[j], s.py[j], s.pz[j], &dx, &dy, &dz);
      const double dr2 = dot3(dx, dy, dz);
      const int m = s.species[j];
      const double r0 = P.r0[m - 1], eps = P.eps[m - 1];
      if (dr2 > r0 * r0) continue;
      const double b = pow6d(r0) / pow6d(sqrt(dr2));
      e += eps * b * (b - 2) + eps;
    }
  }
  for (int o2 = WAVE / 2; o2 > 0; o2 >>= 1) e += __shfl_xor(e, o2);
  if (lane == 0) {
    s.gd_slot[off + w] = o;
    s.gb_slot[off + t] = o;
    s.gd_bso[off + w] = s.bkey[o];
    s.gd_rpos[off + w] = s.ref_pos[o];
    s.gd_oflags[off + w] = s.flags[o];
    s.gd_cell[off + w] = cell;
    s.gb_delt[o] = t;
    s.gb_du[off + t] = e;
  }
}

// the free hs%b indices an insertion of the call can take, in ascending
// order: the first free ones of the call's start (at most one per trial)
// and those of the atoms it deletes (rank = position in that order)
__global__ void __launch_bounds__(DML_BLOCK) k_gb_brank(DevState s, int r) {
  // the deletions' hs%b indices, staged once per workgroup (every thread
  // counts over all of them)
  __shared__ int32_t lb[DML_GCMC_BATCH_MAX];
  const int q = blockIdx.x * DML_BLOCK + threadIdx.x;
  const RepParams &P = s.rp[r];
  if (!gb_on(P)) return;
  const int off = P.gb_off, nUF = min(P.gb_nFB, P.gb_t), nd = min(P.gb_ndel, DML_GCMC_BATCH_MAX);
  const int32_t *fb = s.gb_freeb + off;
  for (int d = threadIdx.x; d < nd; d += DML_BLOCK) lb[d] = s.gd_bso[off + d];
  __syncthreads();
  if (q < nUF) {
    const int x = fb[q];
    int rank = q;
    for (int d = 0; d < nd; d++) {
      const int b = lb[d];
      rank += b >= 0 && b < x;
    }
    s.gb_frank[off + q] = rank;
    s.gb_U[3 * off + rank] = x;
  } else if (q < nUF + nd) {
    const int d = q - nUF;
    const int x = s.gd_bso[off + d];
    if (x < 0) return;
    int lo = 0, hi = nUF;  // free indices of the start below x
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (fb[mid] < x) lo = mid + 1; else hi = mid;
    }
    int rank = lo;
    for (int e = 0; e < nd; e++) {
      const int b = lb[e];
      rank += b >= 0 && b < x;
    }
    s.gd_brank[off + d] = rank;
    s.gb_U[3 * off + rank] = x;
  }
}

// the accepted moves in trial order, one wave: hs%b indices (the lowest
// free one, indices freed earlier in the call included, Groups.F90:1094-
// 1100), the template (first live gcmc atom) and list tail whose velocity
// the insertion sets (dana.F90:620-668), and the ordered stores: flags,
// hs%b maps, ref and gcmc list entries, limbo (Groups.F90:433-467).  The
// universe of free hs%b indices is a bitmap in LDS searched from a lower
// bound of the lowest free rank, their values in LDS too.
__global__ void __launch_bounds__(WAVE) k_gb_commit(DevState s, int r, const double *g) {
  // universe values (the free indices of the start and those freed by the
  // call's deletions: at most two per trial counting the appended ones)
  __shared__ int32_t U[2 * DML_GCMC_BATCH_MAX];
  __shared__ uint32_t ubl[GB_UW * WAVE];
  __shared__ int32_t ibr[DML_GCMC_BATCH_MAX];  // hs%b rank of each insertion
  __shared__ int32_t ibv[DML_GCMC_BATCH_MAX];  // and the index itself
  RepParams *Pw = s.rp + r;
  if (!gb_on(*Pw)) return;
  const RepParams &P = *Pw;
  const int lane = threadIdx.x;
  const GbView V = gb_view(P);
  const int off = P.gb_off, T = P.gb_t, poff = P.poff, ilen = P.gb_ilen;
  const int nins = P.gb_nins, ndel = P.gb_ndel, rlen = P.gb_rlen;
  const int nUF = min(P.gb_nFB, T);
  const double mass[3] = {P.mass[0], P.mass[1], P.mass[2]};
  int nU = 0;
  for (int d = lane; d < ndel; d += WAVE) nU += s.gd_bso[off + d] >= 0;
  for (int o2 = WAVE / 2; o2 > 0; o2 >>= 1) nU += __shfl_xor(nU, o2);
  nU += nUF;
  for (int q = lane; q < GB_UW * WAVE; q += WAVE) ubl[q] = 0u;
  __syncthreads();
  for (int q = lane; q < nU; q += WAVE) U[q] = s.gb_U[3 * off + q];
  for (int q = lane; q < nUF; q += WAVE) {
    const int rk = s.gb_frank[off + q];
    atomicOr(&ubl[rk >> 5], 1u << (rk & 31));
  }
  __syncthreads();
  // the free ranks stay in the LDS bitmap; every rank below ulo is taken
  // (an insertion takes the lowest free rank, a deletion can only lower it)
  int ulo = 0;
  int b_amax = P.b_amax, b_nat = P.b_nat;
  int nappend = 0;          // hs%b indices appended at b_amax (ranks nU + ...)
  int bmin = P.bfree_lo;    // lower bound of the free hs%b indices after the call
  // gcmc list: the staged list (glen entries) then the call's insertions
  const int32_t *gc = s.wl[1] + poff;
  const int glen = P.gb_glen;
  int hd = 0;                 // first live entry (the template)
  int tmpl = gc[0];
  int e_tail = glen - 1;      // last live entry (whose velocity an insertion sets)
  const double ui_ev = 1.0e2 * 1.0e2 * 1.6605402e-27 * (1.0 / 1.60219e-19);
  const double kB_ui = 8.617385e-05 * (1.0 / ui_ev);  // Constants.F90:100-167
  double beta = sqrt(kB_ui * 300.0 / P.mass[s.species[tmpl] - 1]);
  double tv[3] = {s.vx[tmpl], s.vy[tmpl], s.vz[tmpl]};
  int8_t tsp = s.species[tmpl];
  const int call = P.gcmc_calls;
  (void)g;
  // live at trial t: an entry of the start not deleted before t, an
  // insertion made before t and not deleted before it
  auto entry_atom = [&](int e) { return e < glen ? gc[e] : gb_ins_slot_v(s, V, e - glen); };
  auto entry_dead = [&](int e, int t) {
    if (e < glen) return gb_del_time_v(s, V, gc[e]) <= t;
    return s.gi_del[off + (e - glen)] <= t;
  };
  // three streams read 64 entries at a time, one per lane, into registers
  // when the walk reaches the chunk, and taken lane by lane with readlane:
  // the trials (decision, kind, normals), the deletions in order (atom,
  // hs%b index and rank, ref position, flags) and the insertions' slots.
  // Each chunk load is waited for where it is made (every memory operation
  // of the wave before it included, stores too): a load still counted as in
  // flight where the paths merge would make every later iteration wait.
  int c_res = -1, c_kind = 0, c_is = 0, c_d[5] = {0, 0, 0, 0, 0};
  double c_g[3] = {0.0, 0.0, 0.0};
  int ki = 0, kd = 0;
  int ta = gc[glen - 1];  // the tail atom
  for (int t = 0; t < T; t++) {
    // one wave in program order: the barrier only keeps the compiler (and a
    // lane-by-lane CPU model of the wave) from moving LDS accesses across
    // iterations
    __builtin_amdgcn_wave_barrier();
    const int tl = t & (WAVE - 1);
    if (tl == 0) {
      const int q = t + lane;
      if (q < T) {
        c_res = s.gb_res[off + q];
        c_kind = s.gb_kind[off + q];
        c_g[0] = s.gb_g[3 * (off + q)];
        c_g[1] = s.gb_g[3 * (off + q) + 1];
        c_g[2] = s.gb_g[3 * (off + q) + 2];
      }
      __builtin_amdgcn_s_waitcnt(0);
      __builtin_amdgcn_wave_barrier();
    }
    const int res = __builtin_amdgcn_readlane(c_res, tl);
    const int kind = __builtin_amdgcn_readlane(c_kind, tl);
    if (res < 0) continue;
    if (kind == 0) {
      const int k = ki++;
      if ((k & (WAVE - 1)) == 0) {
        if (k + lane < nins) c_is = gb_ins_slot_v(s, V, k + lane);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
      }
      const int i = __builtin_amdgcn_readlane(c_is, k & (WAVE - 1));
      // hs%b index: the lowest free one, or a new one at b_amax
      int bs, rank;
      if (b_amax >= b_nat + 1) {
        int w = ulo >> 5;
        uint32_t word = (uint32_t)rfl_i((int)ubl[w]);
        uint32_t live = word & (0xffffffffu << (ulo & 31));
        while (live == 0u && w + 1 < GB_UW * WAVE) {
          w++;
          word = (uint32_t)rfl_i((int)ubl[w]);
          live = word;
        }
        if (live == 0u) {  // cannot happen: b_amax - b_nat counts the set bits
          if (lane == 0) rep_fail(s, r, ERR_UNSUP);
          return;
        }
        const int bit = __builtin_ctz(live);
        rank = w * 32 + bit;
        if (lane == 0) ubl[w] = word & ~(1u << bit);
        ulo = rank + 1;
        bs = U[rank];
      } else {
        bs = b_amax++;
        rank = nU + nappend++;
        if (lane == 0) U[rank] = bs;
      }
      b_nat++;
      // the velocity lands on the list's last live atom, the new atom copies
      // the template (atom_asign, Groups.F90:484-502)
      const double g0 = rdlane_d(c_g[0], tl), g1 = rdlane_d(c_g[1], tl), g2 = rdlane_d(c_g[2], tl);
      if (lane == 0) {
        s.vx[ta] = beta * g0;
        s.vy[ta] = beta * g1;
        s.vz[ta] = beta * g2;
        s.vx[i] = tv[0];
        s.vy[i] = tv[1];
        s.vz[i] = tv[2];
        s.species[i] = tsp;
        s.flags[i] = DML_FL_ALIVE | DML_FL_REF | DML_FL_GCMC;
        s.bkey[i] = bs;
        s.b2slot[poff + bs] = i;
        s.ref_order[poff + rlen + k] = i;
        s.ref_pos[i] = rlen + k;
        s.wl[1][poff + glen + k] = i;
        s.gi_tmpl[off + k] = tmpl;
        s.gb_slot[off + t] = i;
        ibr[k] = rank;
        ibv[k] = bs;
      }
      e_tail = glen + k;
      ta = i;
    } else {
      const int d = kd++;
      if ((d & (WAVE - 1)) == 0) {
        const int q = d + lane;
        if (q < ndel) {
          c_d[0] = s.gd_slot[off + q];
          c_d[1] = s.gd_bso[off + q];
          c_d[2] = s.gd_brank[off + q];
          c_d[3] = s.gd_rpos[off + q];
          c_d[4] = s.gd_oflags[off + q];
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
      }
      const int dl = d & (WAVE - 1);
      const int o = __builtin_amdgcn_readlane(c_d[0], dl);

      int bso, rank, rpos, ofl;
      if (res >= ilen) {
        const int k = res - ilen;
        bso = ibv[k];
        rank = ibr[k];
        rpos = rlen + k;
        ofl = 0;
      } else {
        bso = __builtin_amdgcn_readlane(c_d[1], dl);
        rank = __builtin_amdgcn_readlane(c_d[2], dl);
        rpos = __builtin_amdgcn_readlane(c_d[3], dl);
        ofl = __builtin_amdgcn_readlane(c_d[4], dl);
      }
      // the freed hs%b index is free again for the later insertions
      if (lane == 0) atomicOr(&ubl[rank >> 5], 1u << (rank & 31));
      ulo = min(ulo, rank);
      b_nat--;
      bmin = min(bmin, bso);
      if (lane == 0) {
        s.flags[o] = (uint8_t)(DML_FL_LIMBO | (ofl & DML_FL_SKIP));
        s.nn[o] = 0;
        s.ref_order[poff + rpos] = -1;
        s.ref_pos[o] = -1;
        s.b2slot[poff + bso] = -1;
      }
      // the template and the tail step over dead entries (dana.F90:620,
      // the gcmc list after gcmc%detach)
      // (values loaded here are taken into scalar registers at once: a load
      // still in flight at the merge below would make every later
      // iteration wait for it)
      if (o == tmpl) {
        while (entry_dead(hd, t)) hd++;
        tmpl = __builtin_amdgcn_readfirstlane(entry_atom(hd));
        tsp = (int8_t)__builtin_amdgcn_readfirstlane((int)s.species[tmpl]);
        beta = sqrt(kB_ui * 300.0 / mass[tsp - 1]);
        tv[0] = rdlane_d(s.vx[tmpl], 0);
        tv[1] = rdlane_d(s.vy[tmpl], 0);
        tv[2] = rdlane_d(s.vz[tmpl], 0);
      }
      if (o == ta) {
        while (e_tail > hd && entry_dead(e_tail, t)) e_tail--;
        ta = __builtin_amdgcn_readfirstlane(entry_atom(e_tail));
      }
    }
  }
  if (lane == 0) {
    const int nF = P.gb_nF;
    const int nsl = P.gb_nslots + max(0, nins - nF);
    Pw->nslots = nsl;
    Pw->free_lo = nins < nF ? (nins < P.gb_cap ? s.gb_free[off + nins] : s.gb_free[off + P.gb_cap - 1] + 1)
                            : nsl;
    Pw->bfree_lo = bmin;
    Pw->nat = P.nat + nins - ndel;
    Pw->nlimbo = P.nlimbo + ndel;
    Pw->b_amax = b_amax;
    Pw->b_nat = b_nat;
    Pw->pid_next = P.pid_next + (uint32_t)nins;
    Pw->seq_next = P.seq_next + nins;
    Pw->gc_len = glen + nins;
    Pw->ref_len = P.gb_rlen + nins;
    Pw->gc_pending = 2;
    if (nins || ndel) Pw->ovn_stamp = 0;
    Pw->gcmc_calls = call + 1;
    if (nins) atomicAdd(&s.cnt[r].ins, (unsigned long long)nins);
    if (ndel) atomicAdd(&s.cnt[r].del, (unsigned long long)ndel);
  }
}

// accepted insertions: the fields atom_asign copies from the template and
// those gcmc_run sets (dana.F90:659-668, Groups.F90:484-502)
__global__ void __launch_bounds__(DML_BLOCK) k_gb_fields(DevState s, int r) {
  const int k = blockIdx.x * DML_BLOCK + threadIdx.x;
  const RepParams &P = s.rp[r];
  if (!gb_on(P) || k >= P.gb_nins) return;
  const int off = P.gb_off;
  const int t = s.gi_trial[off + k];
  const int i = gb_ins_slot(s, P, k);
  const int tm = s.gi_tmpl[off + k];
  const double rx = s.gb_pt[3 * (off + t)], ry = s.gb_pt[3 * (off + t) + 1];
  const double rz = s.gb_pt[3 * (off + t) + 2];
  s.fx[i] = s.fx[tm]; s.fy[i] = s.fy[tm]; s.fz[i] = s.fz[tm];
  s.ax[i] = s.ax[tm]; s.ay[i] = s.ay[tm]; s.az[i] = s.az[tm];
  s.epot[i] = s.epot[tm];
  s.px[i] = rx; s.py[i] = ry; s.pz[i] = rz;
  s.ox[i] = rx; s.oy[i] = ry; s.oz[i] = rz;
  s.cx[i] = 1e8; s.cy[i] = 1e8; s.cz[i] = 1e8;
  s.pid[i] = P.pid_next - (uint32_t)P.gb_nins + (uint32_t)k;
  s.seq[i] = P.seq_next - P.gb_nins + k;
  s.rowflag[i] = 0;
  put_pf4(s, i, rx, ry, rz, s.species[i]);
  // the row built at insertion (kept even when the atom leaves later in the
  // call) marks a ghost-layer cell
  const int cell = s.gb_cell[off + t];
  const int ci0 = cell % (P.ncells[0] + 2);
  const int ci1 = (cell / (P.ncells[0] + 2)) % (P.ncells[1] + 2);
  const int ci2 = cell / ((P.ncells[0] + 2) * (P.ncells[1] + 2));
  const int ghost = ci2 == 0 || ci2 == P.ncells[2] + 1 || ci0 == 0 || ci0 == P.ncells[0] + 1 ||
                    ci1 == 0 || ci1 == P.ncells[1] + 1;
  s.rowflag[i] = (uint8_t)ghost;
  if (ghost) s.rp[r].any_ghost = 1;
}

// an inserted atom's own row (ngroup_sort_atom, Neighbor.F90:271-318) as
// the sequential trials build it at insertion k: the 27 stencil cells in
// order; in each, the call's earlier insertions into it that are still
// present (newest first: push-front, Cells.F90:299-300), then the chain of
// the call's start without the atoms deleted before; strict '<' on the
// list radius.  One wave per insertion kept to the end of the call.
__global__ void __launch_bounds__(DML_BLOCK) k_gb_ownrow(DevState s, int r) {
  __shared__ int lbuf[DML_BLOCK / WAVE][GB_CELLMAX];
  const int w = (blockIdx.x * DML_BLOCK + threadIdx.x) / WAVE, lane = threadIdx.x % WAVE;
  const int wv = threadIdx.x / WAVE;
  const RepParams &P = s.rp[r];
  if (!gb_on(P) || w >= P.gb_nins) return;
  const int off = P.gb_off, k = w;
  if (s.gi_del[off + k] != 0x7fffffff) return;
  const int t = s.gi_trial[off + k];
  const int i = gb_ins_slot(s, P, k);
  const double rx = s.gb_pt[3 * (off + t)], ry = s.gb_pt[3 * (off + t) + 1];
  const double rz = s.gb_pt[3 * (off + t) + 2];
  const int cell = s.gb_cell[off + t];
  const int ci0 = cell % (P.ncells[0] + 2);
  const int ci1 = (cell / (P.ncells[0] + 2)) % (P.ncells[1] + 2);
  const int ci2 = cell / ((P.ncells[0] + 2) * (P.ncells[1] + 2));
  const int maxnn = s.maxnn;
  int32_t *row = s.nbr + (int64_t)i * maxnn;
  int nn = 0;
  for (int kq = 0; kq < 27; kq++) {
    const int n0 = cell_wrap(ci0 + kMap[kq][0], P.ncells[0]);
    const int n1 = cell_wrap(ci1 + kMap[kq][1], P.ncells[1]);
    const int n2 = cell_wrap(ci2 + kMap[kq][2], P.ncells[2]);
    const int cc = cell_lin(P, n0, n1, n2);
    // the earlier insertions into cc still present at t, within the radius
    int nl = 0;
    for (int kk = gb_list_head(s, P, cc); kk >= 0; kk = s.gi_next[off + kk]) {
      if (kk < k && s.gi_del[off + kk] > t) {
        const int tp = s.gi_trial[off + kk];
        double dx, dy, dz;
        vdist_mic(P, s.gb_pt[3 * (off + tp)], s.gb_pt[3 * (off + tp) + 1],
                  s.gb_pt[3 * (off + tp) + 2], rx, ry, rz, &dx, &dy, &dz);
        if (dot3(dx, dy, dz) < P.rl2) {
          if (lane == 0) lbuf[wv][nl] = kk;
          nl++;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    // newest first: entry q goes after the entries of higher rank
    for (int q = lane; q < nl; q += WAVE) {
      const int mine = lbuf[wv][q];
      int pos = 0;
      for (int q2 = 0; q2 < nl; q2++) pos += lbuf[wv][q2] > mine;
      if (nn + pos < maxnn) row[nn + pos] = gb_ins_slot(s, P, mine);
    }
    nn += nl;
    __builtin_amdgcn_wave_barrier();
    // the chain of the start without the atoms deleted before t
    const int64_t base = P.boff + (int64_t)cc * P.bcap;
    const int cnt = min(s.cell_cnt[P.coff + cc], P.bcap);
    for (int q0 = 0; q0 < cnt; q0 += WAVE) {
      const int q = q0 + lane;
      int hit = 0, j = -1;
      if (q < cnt) {
        j = s.bucket[base + q];
        if (j != i && gb_del_time(s, P, j) > t) {
          double dx, dy, dz;
          vdist_mic(P, s.bx[base + q], s.by[base + q], s.bz[base + q], rx, ry, rz, &dx, &dy, &dz);
          hit = dot3(dx, dy, dz) < P.rl2;
        }
      }
      const unsigned long long m = __ballot(hit);
      const int pr = __popcll(m & ((1ull << lane) - 1ull));
      if (hit && nn + pr < maxnn) row[nn + pr] = j;
      nn += __popcll(m);
    }
  }
  if (lane == 0) s.nn[i] = min(nn, maxnn);
}

// the rows an insertion joins (gcmc_run's attach: every ref atom within the
// list radius of the new atom gets it appended, in insertion order).  Pass
// 1, one wave per accepted insertion: each such row present at insertion
// time and kept to the call's end gets the entry at a position taken from
// a per-row counter; the first entry of a row lists the row.
__global__ void __launch_bounds__(DML_BLOCK) k_gb_append(DevState s, int r) {
  const int w = (blockIdx.x * DML_BLOCK + threadIdx.x) / WAVE, lane = threadIdx.x % WAVE;
  RepParams *Pw = s.rp + r;
  const RepParams &P = *Pw;
  if (!gb_on(P) || w >= P.gb_nins) return;
  const int off = P.gb_off, k = w;
  const int t = s.gi_trial[off + k];
  const int i = gb_ins_slot(s, P, k);
  const double rx = s.gb_pt[3 * (off + t)], ry = s.gb_pt[3 * (off + t) + 1];
  const double rz = s.gb_pt[3 * (off + t) + 2];
  const int cell = s.gb_cell[off + t];
  const int ci0 = cell % (P.ncells[0] + 2);
  const int ci1 = (cell / (P.ncells[0] + 2)) % (P.ncells[1] + 2);
  const int ci2 = cell / ((P.ncells[0] + 2) * (P.ncells[1] + 2));
  const int maxnn = s.maxnn;
  const uint32_t stamp = P.gb_stamp * 2u + 1u;
  // the near cells, one per lane (each row gets one entry per insertion,
  // put in insertion order by the second pass: the visiting order is free)
  int ncell = 0;
  const int mycell = near_cell_at(P, ci0, ci1, ci2, lane < 27 ? lane : 0, &ncell);
  int kcnt = 0;
  int64_t kbase = 0;
  if (lane < ncell) {
    kbase = P.boff + (int64_t)mycell * P.bcap;
    kcnt = min(s.cell_cnt[P.coff + mycell], P.bcap);
  }
  // atoms of the start (present to the end, ref): the cells as one stream
  // (a wave prefix sum of the counts), 64 entries a pass
  const int kincl = wave_scan_incl(kcnt);
  const int total = __builtin_amdgcn_readlane(kincl, WAVE - 1);
  for (int s0 = 0; s0 < total; s0 += WAVE) {
    const int sp = s0 + lane;
    int kc = 0;
    for (int c = 0; c < ncell; c++) kc += __builtin_amdgcn_readlane(kincl, c) <= sp;
    const int kk = min(kc, WAVE - 1);
    const int64_t base = ((int64_t)__shfl((int)(kbase >> 32), kk) << 32) |
                         (uint32_t)__shfl((int)(kbase & 0xffffffff), kk);
    const int q = sp - (__shfl(kincl, kk) - __shfl(kcnt, kk));
    if (sp >= total) continue;
    const int aj = s.bucket[base + q];
    if (!(s.flags[aj] & DML_FL_REF) || gb_del_time(s, P, aj) != 0x7fffffff) continue;
    double dx, dy, dz;
    vdist_mic(P, rx, ry, rz, s.bx[base + q], s.by[base + q], s.bz[base + q], &dx, &dy, &dz);
    if (dot3(dx, dy, dz) > P.rl2) continue;
    const int at = gb_stamped_inc(&s.gb_app[aj], stamp);
    const int pos = s.nn[aj] + at;
    if (pos < maxnn) s.nbr[(int64_t)aj * maxnn + pos] = i;
    if (at == 0) s.gb_alist[P.poff + atomicAdd(&Pw->gb_nalist, 1)] = aj;
  }
  // the call's earlier insertions into each cell kept to the end: lane c
  // walks the list of cell c
  if (lane < ncell) {
    for (int kk = gb_list_head(s, P, mycell); kk >= 0; kk = s.gi_next[off + kk]) {
      if (kk >= k || s.gi_del[off + kk] != 0x7fffffff) continue;
      const int tp = s.gi_trial[off + kk];
      double dx, dy, dz;
      vdist_mic(P, rx, ry, rz, s.gb_pt[3 * (off + tp)], s.gb_pt[3 * (off + tp) + 1],
                s.gb_pt[3 * (off + tp) + 2], &dx, &dy, &dz);
      if (dot3(dx, dy, dz) > P.rl2) continue;
      const int aj = gb_ins_slot(s, P, kk);
      const int at = gb_stamped_inc(&s.gb_app[aj], stamp);
      const int pos = s.nn[aj] + at;
      if (pos < maxnn) s.nbr[(int64_t)aj * maxnn + pos] = i;
      if (at == 0) s.gb_alist[P.poff + atomicAdd(&Pw->gb_nalist, 1)] = aj;
    }
  }
}

// pass 2, per listed row: the appended entries in insertion order, and the
// row length
__global__ void __launch_bounds__(DML_BLOCK) k_gb_append_fix(DevState s, int r) {
  const RepParams &P = s.rp[r];
  if (!gb_on(P)) return;
  const int maxnn = s.maxnn;
  const uint32_t stamp = P.gb_stamp * 2u + 1u;
  for (int q = blockIdx.x * DML_BLOCK + threadIdx.x; q < P.gb_nalist;
       q += gridDim.x * DML_BLOCK) {
    const int aj = s.gb_alist[P.poff + q];
    const int cnt = gb_stamped_get(&s.gb_app[aj], stamp);
    const int base = s.nn[aj];
    int32_t *row = s.nbr + (int64_t)aj * maxnn + base;
    const int m = min(cnt, maxnn - base);
    // insertion sort by insertion rank (a handful of entries)
    for (int a = 1; a < m; a++) {
      const int v = row[a];
      const int kv = gb_ins_rank(s, P, v);
      int b = a - 1;
      while (b >= 0 && gb_ins_rank(s, P, row[b]) > kv) {
        row[b + 1] = row[b];
        b--;
      }
      row[b + 1] = v;
    }
    s.nn[aj] = min(base + cnt, maxnn);
  }
}

// each touched cell's chain, rebuilt once (claimed with the call's stamp):
// the call's insertions into it kept to the end, newest first, then the
// chain of the start without the deleted atoms.  One wave per accepted
// move (insertions, then deletions); the new chain is staged in LDS.
__global__ void __launch_bounds__(WAVE) k_gb_cells(DevState s, int r) {
  __shared__ int32_t lst[GB_CELLMAX];
  __shared__ int32_t cj[GB_CELLMAX];
  __shared__ double cxs[GB_CELLMAX], cys[GB_CELLMAX], czs[GB_CELLMAX];
  __shared__ float4 cf[GB_CELLMAX];
  const int e = blockIdx.x, lane = threadIdx.x;
  const RepParams &P = s.rp[r];
  if (!gb_on(P) || e >= P.gb_nins + P.gb_ndel) return;
  const int off = P.gb_off;
  const int cell = e < P.gb_nins ? s.gb_cell[off + s.gi_trial[off + e]]
                                 : s.gd_cell[off + (e - P.gb_nins)];
  int claimed = 0;
  if (lane == 0)
    claimed = atomicExch(&s.gb_claim[P.coff + cell], (int)P.gb_stamp) != (int)P.gb_stamp;
  if (!__builtin_amdgcn_readfirstlane(claimed)) return;
  const int64_t base = P.boff + (int64_t)cell * P.bcap;
  const int cnt = min(s.cell_cnt[P.coff + cell], P.bcap);
  // the call's insertions into the cell kept to the end
  int nl = 0;
  for (int kk = gb_list_head(s, P, cell); kk >= 0; kk = s.gi_next[off + kk]) {
    if (s.gi_del[off + kk] == 0x7fffffff) {
      if (lane == 0) lst[nl] = kk;
      nl++;
    }
  }
  __syncthreads();
  // newest first: an insertion goes after those of higher rank
  for (int q = lane; q < nl; q += WAVE) {
    const int kk = lst[q];
    int pos = 0;
    for (int q2 = 0; q2 < nl; q2++) pos += lst[q2] > kk;
    const int i = gb_ins_slot(s, P, kk);
    const int tp = s.gi_trial[off + kk];
    cj[pos] = i;
    cxs[pos] = s.gb_pt[3 * (off + tp)];
    cys[pos] = s.gb_pt[3 * (off + tp) + 1];
    czs[pos] = s.gb_pt[3 * (off + tp) + 2];
    cf[pos] = bucket_f4(s, i);
  }
  // then the chain of the start without the deleted atoms, in order
  int nk = nl;
  for (int q0 = 0; q0 < cnt; q0 += WAVE) {
    const int q = q0 + lane;
    int j = -1, kp = 0;
    if (q < cnt) {
      j = s.bucket[base + q];
      kp = gb_del_time(s, P, j) == 0x7fffffff;
    }
    const unsigned long long m = __ballot(kp);
    if (kp) {
      const int dst = nk + __popcll(m & ((1ull << lane) - 1ull));
      cj[dst] = j;
      cxs[dst] = s.bx[base + q];
      cys[dst] = s.by[base + q];
      czs[dst] = s.bz[base + q];
      cf[dst] = s.bf4[base + q];
    }
    nk += __popcll(m);
  }
  __syncthreads();
  for (int q = lane; q < nk; q += WAVE) {
    s.bucket[base + q] = cj[q];
    s.bx[base + q] = cxs[q];
    s.by[base + q] = cys[q];
    s.bz[base + q] = czs[q];
    s.bf4[base + q] = cf[q];
  }
  if (lane == 0) s.cell_cnt[P.coff + cell] = nk;
}

// ------------------------------------------------------------------------
// chunks reservoir: bloques (dana.F90:716-773).  When the sensor density of
// calc_rho leaves its band, the reservoir moves up by dist, a copy of the
// block template joins sys, hs%b, hs%ref and hs (slots and hs%b indices
// reused as igroup_attach_atom does, Groups.F90:1058-1100), the template
// moves up, the box grows to zmax and the lists are rebuilt (the caller runs
// test_update on the replicas whose chunk_fired is set).
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(DML_BLOCK) k_chunk_run(DevState s) {
  __shared__ int sh[DML_BLOCK / WAVE];
  const int r = blockIdx.x;
  if (r >= s.nrep) return;
  RepParams *Pw = s.rp + r;
  if (Pw->halted || Pw->reservoir != DML_CHUNKS || Pw->nchunk <= 0) return;
  const RepParams P = *Pw;
  const int tid = threadIdx.x;
  const double rhomedia = 5.775329e-4;
  const double rho = (double)P.rho_count / (P.box[0] * P.box[1] * (P.z1 - P.z0));
  if (fabs(rho - rhomedia) < rhomedia * (double)0.186f) return;
  const int poff = P.poff, nch = P.nchunk;
  int32_t *hs_of = s.ref_tmp + poff;  // chosen hs slot of template atom j
  int32_t *b_of = s.wl[0] + poff;     // chosen hs%b index of template atom j
  // hs slots: atom j reuses the j-th null slot while nslots >= nat + j + 1 +
  // nlimbo, i.e. for j < nslots - nat - nlimbo; the rest are appended
  const int take = max(0, min(nch, P.nslots - P.nat - P.nlimbo));
  int got = 0;
  for (int q0 = 0; q0 < P.nslots && got < take; q0 += DML_BLOCK) {
    int q = q0 + tid;
    int fr = q < P.nslots && !(s.flags[poff + q] & (DML_FL_ALIVE | DML_FL_LIMBO));
    int tot;
    int pos = blk_scan01(fr, sh, &tot);
    if (fr && got + pos < take) hs_of[got + pos] = q;
    got += tot;
  }
  int fail = 0;
  if (got < take) fail = ERR_UNSUP;
  if (P.nslots + (nch - take) > P.pcap) fail = ERR_BUCKET;
  // hs%b indices: the first free ones while b_amax >= b_nat + j + 1
  const int takeb = max(0, min(nch, P.b_amax - P.b_nat));
  got = 0;
  for (int q0 = 0; q0 < P.b_amax && got < takeb; q0 += DML_BLOCK) {
    int q = q0 + tid;
    int fr = q < P.b_amax && s.b2slot[poff + q] < 0;
    int tot;
    int pos = blk_scan01(fr, sh, &tot);
    if (fr && got + pos < takeb) b_of[got + pos] = q;
    got += tot;
  }
  if (got < takeb) fail = ERR_UNSUP;
  if (P.b_amax + (nch - takeb) > P.pcap) fail = ERR_BUCKET;
  // the grown box and its tessellation
  RepParams Q = P;
  Q.z0 = P.z0 + P.dist;
  Q.z1 = P.z1 + P.dist;
  Q.zmax = P.zmax + P.dist;
  Q.box[2] = Q.zmax;
  Q.one_box[2] = 1.0 / Q.box[2];
  if (Q.box[2] > DML_MAX_BOX) fail = ERR_UNSUP;  // beyond the FP32 mirror's range
  if (tessellate_rep(Q, 0)) fail = ERR_UNSUP;
  Q.ccap = Q.use_cells ? (Q.ncells[0] + 2) * (Q.ncells[1] + 2) * (Q.ncells[2] + 2) : 1;
  if (Q.ccap > P.creg) fail = ERR_BUCKET;
  if (!Q.use_cells && P.use_cells) Q.bcap = (int32_t)min<int64_t>(P.breg, (int64_t)P.pcap);
  if ((int64_t)Q.ccap * Q.bcap > P.breg) Q.bcap = (int32_t)(P.breg / Q.ccap);
  if (Q.bcap < 8) fail = ERR_BUCKET;
  if (fail) {
    if (tid == 0) rep_fail(s, r, fail);
    return;
  }
  for (int j = tid; j < nch; j += DML_BLOCK) {
    const int hs = j < take ? hs_of[j] : P.nslots + (j - take);
    const int bs = j < takeb ? b_of[j] : P.b_amax + (j - takeb);
    const int i = poff + hs;
    const int64_t c = P.chunk_off + j;
    // atom_allocate defaults (Groups.F90:370-387), then the template's place
    s.species[i] = SP_LI;
    s.flags[i] = DML_FL_ALIVE | DML_FL_REF;
    s.rowflag[i] = 0;
    s.px[i] = s.chx[c]; s.py[i] = s.chy[c]; s.pz[i] = s.chz[c];
    s.ox[i] = s.chox[c]; s.oy[i] = s.choy[c]; s.oz[i] = s.choz[c];
    s.cx[i] = 1e8; s.cy[i] = 1e8; s.cz[i] = 1e8;
    s.vx[i] = 0.0; s.vy[i] = 0.0; s.vz[i] = 0.0;
    s.ax[i] = 0.0; s.ay[i] = 0.0; s.az[i] = 0.0;
    s.fx[i] = 0.0; s.fy[i] = 0.0; s.fz[i] = 0.0;
    s.epot[i] = 0.0;
    s.nn[i] = 0;
    s.bkey[i] = bs;
    s.b2slot[poff + bs] = i;
    s.pid[i] = P.pid_next + (uint32_t)j;
    s.seq[i] = P.seq_next + j;
    s.ref_order[poff + P.nref + j] = i;
    s.ref_pos[i] = P.nref + j;
    put_pf4(s, i, s.chx[c], s.chy[c], s.chz[c], SP_LI);
    // and the template moves up with the reservoir
    s.chz[c] = s.chz[c] + P.dist;
    s.choz[c] = s.choz[c] + P.dist;
  }
  // every thread has copied the replica's parameters (without reusable
  // slots no barrier above separates that copy from the writes below)
  __syncthreads();
  if (tid == 0) {
    Pw->z0 = Q.z0;
    Pw->z1 = Q.z1;
    Pw->zmax = Q.zmax;
    Pw->box[2] = Q.box[2];
    Pw->one_box[2] = Q.one_box[2];
    for (int k = 0; k < 3; k++) {
      Pw->ncells[k] = Q.ncells[k];
      Pw->cell[k] = Q.cell[k];
    }
    Pw->use_cells = Q.use_cells;
    Pw->ccap = Q.ccap;
    Pw->bcap = Q.bcap;
    Pw->nslots = P.nslots + (nch - take);
    Pw->b_amax = P.b_amax + (nch - takeb);
    Pw->b_nat = P.b_nat + nch;
    Pw->nat = P.nat + nch;
    Pw->nref = P.nref + nch;
    Pw->pid_next = P.pid_next + (uint32_t)nch;
    Pw->seq_next = P.seq_next + nch;
    Pw->listed = 0;  // hs%listed = .false.: the next test_update rebuilds
    Pw->chunk_fired = 1;
    Pw->ovn_stamp = 0;  // atoms added
  }
}

// hold every replica that added no block out of the test_update that
// follows bloques (on = 1); release them and clear the marks (on = 0)
__global__ void k_mask_fired(DevState s, int on) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= s.nrep) return;
  RepParams *Pw = s.rp + r;
  if (on) {
    if (!Pw->chunk_fired) Pw->halted |= HALT_MASK;
  } else {
    Pw->halted &= ~HALT_MASK;
    Pw->chunk_fired = 0;
  }
}
