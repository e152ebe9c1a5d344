// dml_api.hip -- host side of the C ABI declared in include/dml.h.
//
// The host keeps the per-replica geometry (tessellation, buffer offsets) and
// sequences the kernels of dml_kernels.hip on the context's stream.  All
// per-step state lives in HBM; the host reads device memory only where an
// entry point returns data to the caller.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "dml_kernels.h"

#define HIPCHECK(x)                                                   \
  do {                                                                \
    hipError_t e_ = (x);                                              \
    if (e_ != hipSuccess) {                                           \
      fprintf(stderr, "libdml: %s failed: %s\n", #x, hipGetErrorString(e_)); \
      return DML_E_HIP;                                               \
    }                                                                 \
  } while (0)

// Staging of dml_gcmc_trial_energy_noalloc: trial points (3 x double), du
// (double), hard and exclude (int32) for at most DML_NOALLOC_MAX_TRIAL trials.
// Every GPU holds its own copy until the process ends.  Host code never uses
// the name `scratch` as a device address: dml_create_ensemble resolves the
// copy of the current device once (scratch_setup) and the entry point passes
// that address.
constexpr size_t kScratchBytes = (size_t)DML_NOALLOC_MAX_TRIAL * (4 * sizeof(double) +
                                                                  2 * sizeof(int32_t));
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};  // device address of `scratch` per GPU

// with device dev current: the address of that GPU's copy of `scratch`
int scratch_setup(int dev, unsigned char **out) {
  if (dev < 0 || dev >= kMaxDevices) return DML_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      fprintf(stderr, "libdml: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return DML_E_HIP;
    }
    g_scratch[dev] = (unsigned char *)ptr;
  }
  *out = g_scratch[dev];
  return DML_OK;
}

constexpr int kMaxRounds = 32;

struct HostRep {
  dml_config cfg;
  RepParams p;         // host mirror of the static parameters
  int64_t creg = 0;    // cell-count entries reserved for this replica
  int64_t breg = 0;    // bucket entries reserved for this replica
  std::vector<double> chunk;  // block template (config_chunk), 6 x nchunk SoA
};

}  // namespace

struct dml_ctx {
  int nrep = 0;
  std::vector<HostRep> reps;
  DevState d{};
  int64_t P = 0;  // slots
  int64_t C = 0;  // cells
  int64_t B = 0;  // bucket entries
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_rounds = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double last_ms = 0.0;
  int lpa = 8;
  int lpa_ov = 8;  // lanes per atom of the overlap scan (1 = thread per atom)
  int ov_tail = 8;  // overlap rounds from this one on run in one workgroup (0: none)
  // list kernel: 0 cell tiles, 1 wave per atom, 2 per replica by density
  // (0 or 1), 3 wave per cell
  int lb_mode = 4;
  // some replica uses list kernel 0 (cell tiles) / 1 (wave per atom) / 2
  // (wave per cell) / 3 (wave per cell, candidates in registers)
  int any_lb[4] = {0, 0, 1, 0};
  int rounds = 16;  // overlap rounds per step (the C3 ensemble resolves in < 16)
  int any_ermak = 0, any_brown = 0, any_rounds = 0, any_exact = 0, any_piston = 0, any_gcmc = 0;
  int any_chunks = 0;
  int64_t chunk_cap = 0;    // template atoms the device chunk pool holds
  int64_t chunk_used = 0;   // ... and has handed out (append only)
  int cells_dirty = 0;      // reserved cell regions grew: reallocate before use
  uint64_t map_sig = 0;     // layout the flat work lists were built for
  int64_t map_version = 0;  // bumped at every rebuild of the work lists
  int64_t cap_cm = 0, cap_tm = 0, cap_wm = 0;
  int64_t graph_modes = -1;  // mode signature the captured step was built for
  uint64_t graph_dsig = 0;   // device-state bytes the captured step was built with
  int64_t gb_total = 0;      // batched-GCMC trial pool (sum of the replicas' capacities)
  int dev = 0;               // the GPU current at creation
  unsigned char *scratch = nullptr;  // that GPU's copy of `scratch`
};

namespace {

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

int next_pow2(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

// cgroup_tessellate (Cells.F90:180-265) on the host
int tessellate(RepParams &p, int first) {
  return tessellate_rep(p, first) ? DML_E_UNSUPPORTED : DML_OK;
}

int bucket_cap(const RepParams &p);
void set_ermak_consts(RepParams &p);

// The FP32 prefilter's band around a squared cutoff R^2 (R <= rl).  With
// coordinates below B = DML_MAX_BOX each float carries at most B 2^-25 of
// rounding; a minimum-image component (two coordinates, the subtraction, the
// rounded box and the shift) then errs by at most 2.5 B 2^-24 = delta, and
// the squared distance of a pair near R by at most 2 sqrt(3) R delta plus
// 3 R^2 2^-24 from the squares and sums.  The band is twice that, and at
// least 0.02 A^2.
static float pf_margin_for(double rl) {
  const double u = ldexp(1.0, -24);
  const double delta = 2.5 * DML_MAX_BOX * u;
  const double bound = 2.0 * sqrt(3.0) * rl * delta + 3.0 * rl * rl * u;
  return (float)std::max(0.02, 2.0 * bound);
}

// the static parameters of one replica from its config (everything but its
// slot range and cell layout; p.pcap must be set: a Verlet bucket holds it)
int setup_rep(RepParams &p, const dml_config &g) {
  for (int k = 0; k < 3; k++) {
    p.box[k] = g.box[k];
    p.one_box[k] = 1.0 / g.box[k];
  }
  p.rcut = g.rcut;
  p.rcut2 = g.rcut * g.rcut;
  p.skin = g.skin;
  double rl = g.rcut + g.skin;
  p.rl2 = rl * rl;
  p.zmax = g.zmax;
  p.z0 = g.z0;
  p.prob = g.prob;
  p.h = g.h;
  p.z_switch = g.z_switch;
  p.dif_sc = g.dif_sc;
  p.dif_sei = g.dif_sei;
  for (int q = 0; q < 9; q++) {
    p.eps[q] = g.eps[q];
    p.r0[q] = g.r0[q];
  }
  for (int q = 0; q < 3; q++) p.mass[q] = g.mass[q];
  set_ermak_consts(p);
  p.act = g.act;
  p.nadj = g.nadj;
  p.integrator = g.integrator;
  p.reservoir = g.reservoir;
  p.piston = g.reservoir == DML_PISTON;
  p.precision = g.precision;
  p.overlap_mode = g.overlap_mode;
  p.seed = g.seed;
  p.gcmc_mode = g.gcmc_mode;
  p.use_cells = 0;
  for (int k = 0; k < 3; k++)
    if (p.box[k] > DML_MAX_BOX) return DML_E_UNSUPPORTED;
  p.pf_margin = pf_margin_for(rl);
  if (tessellate(p, 1)) return DML_E_UNSUPPORTED;
  p.bcap = bucket_cap(p);
  return DML_OK;
}

// neighbour-row capacity a config asks for (default: 0.06 atoms/A^3 in the
// list sphere, +30 %)
int rows_for(const dml_config &g) {
  int mnn = g.max_neighbors;
  if (mnn <= 0) {
    double rl = g.rcut + g.skin;
    double vol = 4.0 / 3.0 * M_PI * rl * rl * rl;
    mnn = next_pow2((int)(0.06 * vol * 1.3) + 32);
  }
  return mnn;
}

int ncells_total(const RepParams &p) {
  return p.use_cells ? (p.ncells[0] + 2) * (p.ncells[1] + 2) * (p.ncells[2] + 2) : 1;
}

int bucket_cap(const RepParams &p) {
  if (!p.use_cells) return (int)p.pcap;
  double v = p.cell[0] * p.cell[1] * p.cell[2];
  int est = (int)ceil(0.06 * v) + 8;
  return std::max(8, next_pow2(est));
}

void set_ermak_consts(RepParams &p) {
  // set_ermak (dana.F90:947-971) with dana's kB (dana.F90:22-24)
  const double kB_eVK = 8.617330350e-5, kJm_ui = 100.0, eV_kJm = 96.485;
  const double eV_ui = eV_kJm * kJm_ui;
  const double kB_ui = kB_eVK * eV_ui;
  double h = p.h, gama = 1.0;
  p.cc0 = exp(-(h * gama));
  p.cc1 = (1.0 - p.cc0) / gama;
  p.cc2 = (1.0 - p.cc1 / h) / gama;
  p.sdr = sqrt(h / gama * (2.0 - ((3.0 - 4.0 * p.cc0) + p.cc0 * p.cc0) / (h * gama)));
  p.sdv = sqrt(1.0 - p.cc0 * p.cc0);
  p.crv1 = (1.0 - p.cc0) * (1.0 - p.cc0) / (gama * p.sdr * p.sdv);
  p.crv2 = sqrt(1.0 - (p.crv1 * p.crv1));
  p.skt = sqrt(kB_ui * 300.0);
}

int status_from_err(int e) {
  if (e & ERR_TESS) return DML_E_OUT_OF_TESS;
  if (e & ERR_DEPZ0) return DML_E_DEPOSIT_Z0;
  if (e & ERR_NOPART) return DML_E_NO_PARTICLE;
  if (e & ERR_OVERLAP) return DML_E_OVERLAP;
  if (e & ERR_LIST) return DML_E_LIST_FULL;
  if (e & ERR_BUCKET) return DML_E_NOMEM;
  if (e & ERR_UNSUP) return DML_E_UNSUPPORTED;
  return DML_OK;
}

template <typename T>
int dalloc(T **p, int64_t n) {
  if (n <= 0) n = 1;
  if (hipMalloc((void **)p, sizeof(T) * (size_t)n) != hipSuccess) return DML_E_NOMEM;
  return DML_OK;
}

int read_err(dml_ctx *c, int reset) {
  int e = 0;
  if (hipMemcpyAsync(&e, c->d.err, sizeof(int), hipMemcpyDeviceToHost, c->st) != hipSuccess)
    return -1;
  if (hipStreamSynchronize(c->st) != hipSuccess) return -1;
  if (reset && e) {
    int z = 0;
    hipMemcpyAsync(c->d.err, &z, sizeof(int), hipMemcpyHostToDevice, c->st);
    hipStreamSynchronize(c->st);
  }
  return e;
}

void assign_list_kernels(dml_ctx *c);

// the flat work lists of the cell and list kernels: every reserved cell of
// every replica (sort), the cells of cell-tile replicas (list build) and the
// particle blocks of wave-per-atom replicas; rebuilt when the layout changes
int build_maps(dml_ctx *c) {
  uint64_t sig = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { sig = (sig ^ v) * 1099511628211ull; };
  for (auto &h : c->reps) {
    mix((uint64_t)h.p.coff); mix((uint64_t)h.p.creg); mix((uint64_t)h.p.lb_kind);
    mix((uint64_t)h.p.poff); mix((uint64_t)h.p.pcap);
  }
  if (sig == c->map_sig && c->d.cm_rep) return DML_OK;
  std::vector<int32_t> cr, cc, tr, tc, wb;
  for (int r = 0; r < c->nrep; r++) {
    const RepParams &p = c->reps[r].p;
    for (int64_t q = 0; q < p.creg; q++) {
      cr.push_back(r);
      cc.push_back((int32_t)q);
      if (p.lb_kind == 0) {
        tr.push_back(r);
        tc.push_back((int32_t)q);
      }
    }
    if (p.lb_kind == 1)
      for (int b = p.poff / DML_BLOCK; b < (p.poff + p.pcap) / DML_BLOCK; b++) wb.push_back(b);
  }
  auto put = [&](int32_t **dst, int64_t *cap, const std::vector<int32_t> &v) -> int {
    if ((int64_t)v.size() > *cap || !*dst) {
      hipFree(*dst);
      *dst = nullptr;
      *cap = std::max<int64_t>((int64_t)v.size(), 1);
      if (dalloc(dst, *cap)) return DML_E_NOMEM;
    }
    if (!v.empty())
      HIPCHECK(hipMemcpy(*dst, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice));
    return DML_OK;
  };
  int64_t capc = c->cap_cm, capt = c->cap_tm;
  int rc;
  if ((rc = put(&c->d.cm_rep, &capc, cr))) return rc;
  capc = c->cap_cm;
  if ((rc = put(&c->d.cm_cell, &capc, cc))) return rc;
  c->cap_cm = capc;
  if ((rc = put(&c->d.tm_rep, &capt, tr))) return rc;
  capt = c->cap_tm;
  if ((rc = put(&c->d.tm_cell, &capt, tc))) return rc;
  c->cap_tm = capt;
  if ((rc = put(&c->d.wm_blk, &c->cap_wm, wb))) return rc;
  c->d.n_cm = (int64_t)cr.size();
  c->d.n_tm = (int64_t)tr.size();
  c->d.n_wm = (int64_t)wb.size();
  c->map_sig = sig;
  c->map_version++;
  return DML_OK;
}

int upload_params(dml_ctx *c) {
  assign_list_kernels(c);
  std::vector<RepParams> v(c->nrep);
  for (int r = 0; r < c->nrep; r++) v[r] = c->reps[r].p;
  HIPCHECK(hipMemcpyAsync(c->d.rp, v.data(), sizeof(RepParams) * c->nrep, hipMemcpyHostToDevice,
                          c->st));
  return DML_OK;
}

// the mutable device fields are authoritative: fetch them into the mirror
int fetch_params(dml_ctx *c) {
  std::vector<RepParams> v(c->nrep);
  HIPCHECK(hipMemcpyAsync(v.data(), c->d.rp, sizeof(RepParams) * c->nrep, hipMemcpyDeviceToHost,
                          c->st));
  HIPCHECK(hipStreamSynchronize(c->st));
  for (int r = 0; r < c->nrep; r++) c->reps[r].p = v[r];
  return DML_OK;
}

int alloc_cells(dml_ctx *c) {
  int64_t C = 0, B = 0;
  for (auto &h : c->reps) {
    h.p.coff = (int32_t)C;
    h.p.ccap = ncells_total(h.p);
    h.p.boff = B;
    h.creg = std::max<int64_t>(h.creg, h.p.ccap);
    h.breg = std::max<int64_t>(h.breg, (int64_t)h.p.ccap * h.p.bcap);
    h.p.creg = h.creg;
    h.p.breg = h.breg;
    C += h.creg;
    B += h.breg;
  }
  hipFree(c->d.cell_cnt);
  hipFree(c->d.bucket);
  hipFree(c->d.bx);
  hipFree(c->d.by);
  hipFree(c->d.bz);
  hipFree(c->d.bf4);
  hipFree(c->d.gb_chead);
  hipFree(c->d.gb_ccnt);
  hipFree(c->d.gb_claim);
  c->d.gb_chead = c->d.gb_ccnt = nullptr;
  c->d.gb_claim = nullptr;
  c->d.cell_cnt = nullptr;
  c->d.bucket = nullptr;
  c->d.bx = c->d.by = c->d.bz = nullptr;
  c->d.bf4 = nullptr;
  int e = 0;
  e |= dalloc(&c->d.bf4, B);
  e |= dalloc(&c->d.cell_cnt, C);
  e |= dalloc(&c->d.bucket, B);
  e |= dalloc(&c->d.bx, B);
  e |= dalloc(&c->d.by, B);
  e |= dalloc(&c->d.bz, B);
  // batched GCMC, per cell: stamped list heads and counts, rebuild claims
  const int64_t CS = c->gb_total ? C : 1;
  e |= dalloc(&c->d.gb_chead, CS);
  e |= dalloc(&c->d.gb_ccnt, CS);
  e |= dalloc(&c->d.gb_claim, CS);
  if (e) return DML_E_NOMEM;
  HIPCHECK(hipMemsetAsync(c->d.cell_cnt, 0, sizeof(int32_t) * C, c->st));
  HIPCHECK(hipMemsetAsync(c->d.gb_chead, 0, sizeof(unsigned long long) * CS, c->st));
  HIPCHECK(hipMemsetAsync(c->d.gb_ccnt, 0, sizeof(unsigned long long) * CS, c->st));
  HIPCHECK(hipMemsetAsync(c->d.gb_claim, 0, sizeof(int32_t) * CS, c->st));
  c->C = C;
  c->B = B;
  return DML_OK;
}

// ---- phases -------------------------------------------------------------
unsigned clamp_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(n, 65535)); }

void enq_listbuild(dml_ctx *c) {
  if (c->any_lb[0] && c->d.n_tm > 0)
    hipLaunchKernelGGL(k_listbuild_cells, dim3(clamp_grid(c->d.n_tm)), dim3(DML_BLOCK), 0, c->st,
                       c->d);
  if (c->any_lb[1] && c->d.n_wm > 0)
    hipLaunchKernelGGL(k_listbuild<4>, dim3((unsigned)(c->d.n_wm * (DML_BLOCK / 4))),
                       dim3(4 * 64), 0, c->st, c->d);
  if (c->any_lb[2]) {
    // a multiple of the 8 XCDs (the kernel deals its cell runs by XCD)
    unsigned g = clamp_grid((c->d.n_cm + 3) / 4);
    g = std::max(8u, g / 8 * 8);
    hipLaunchKernelGGL(k_listbuild_wc, dim3(g), dim3(DML_BLOCK), 0, c->st, c->d);
  }
  if (c->any_lb[3]) {
    unsigned g = clamp_grid((c->d.n_cm + 3) / 4);
    g = std::max(8u, g / 8 * 8);
    // 3 chunks of 64 a pass at 7 waves per SIMD (4 at 6 waves, 2 at 8 were
    // slower on the 256-replica ensemble); 2 chunks at 8 waves a choice
    if (c->lb_mode == 5)
      hipLaunchKernelGGL((k_listbuild_wr<2, 8>), dim3(g), dim3(DML_BLOCK), 0, c->st, c->d);
    else
      hipLaunchKernelGGL((k_listbuild_wr<3, 7>), dim3(g), dim3(DML_BLOCK), 0, c->st, c->d);
  }
}

// the list kernel of each replica (host mirror; the caller uploads): cell
// tiles pay off from a few atoms per cell, below that one wave per atom
// wastes less on staging empty neighbour cells
void assign_list_kernels(dml_ctx *c) {
  c->any_lb[0] = c->any_lb[1] = c->any_lb[2] = c->any_lb[3] = 0;
  for (auto &h : c->reps) {
    RepParams &p = h.p;
    int kind = c->lb_mode;
    if (kind == 2) {
      double cells = p.use_cells ? (double)p.ncells[0] * p.ncells[1] * p.ncells[2] : 1.0;
      kind = (p.use_cells && (double)p.nslots / cells >= 4.0) ? 0 : 1;
    } else if (kind == 3) {
      kind = 2;
    } else if (kind >= 4) {
      kind = 3;
    }
    p.lb_kind = kind;
    c->any_lb[kind] = 1;
  }
}

// gcs = 1: also re-sort the cells of the GCMC replicas (their trials read
// them before the next test_update)
void enq_cells_lists(dml_ctx *c, int gcs) {
  dim3 gc(clamp_grid((c->d.n_cm + DML_BLOCK - 1) / DML_BLOCK));
  dim3 gw(clamp_grid((c->d.n_cm + 3) / 4));
  int nb = c->d.nblocks;
  hipLaunchKernelGGL(k_zero_cells, gc, dim3(DML_BLOCK), 0, c->st, c->d, gcs);
  hipLaunchKernelGGL(k_cell_key, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d, gcs);
  hipLaunchKernelGGL(k_cell_sort_w, gw, dim3(DML_BLOCK), 0, c->st, c->d, gcs);
  enq_listbuild(c);
}

void enq_test_update(dml_ctx *c, int gcs) {
  hipLaunchKernelGGL(k_pbc_dispmax, dim3(c->d.nblocks), dim3(DML_BLOCK), 0, c->st, c->d);
  hipLaunchKernelGGL(k_dispmax_final, dim3(c->nrep), dim3(DML_BLOCK), 0, c->st, c->d);
  enq_cells_lists(c, gcs);
}

template <int FUSEB>
void launch_force_lpa(dml_ctx *c, int lpa, int only_ermak) {
  dim3 g((unsigned)(c->P * lpa / DML_BLOCK));
  switch (lpa) {
    case 1: hipLaunchKernelGGL((k_force<1, FUSEB>), g, dim3(DML_BLOCK), 0, c->st, c->d, only_ermak); break;
    case 2: hipLaunchKernelGGL((k_force<2, FUSEB>), g, dim3(DML_BLOCK), 0, c->st, c->d, only_ermak); break;
    case 4: hipLaunchKernelGGL((k_force<4, FUSEB>), g, dim3(DML_BLOCK), 0, c->st, c->d, only_ermak); break;
    case 16: hipLaunchKernelGGL((k_force<16, FUSEB>), g, dim3(DML_BLOCK), 0, c->st, c->d, only_ermak); break;
    case 32: hipLaunchKernelGGL((k_force<32, FUSEB>), g, dim3(DML_BLOCK), 0, c->st, c->d, only_ermak); break;
    case 64: hipLaunchKernelGGL((k_force<64, FUSEB>), g, dim3(DML_BLOCK), 0, c->st, c->d, only_ermak); break;
    default: hipLaunchKernelGGL((k_force<8, FUSEB>), g, dim3(DML_BLOCK), 0, c->st, c->d, only_ermak); break;
  }
}

void enq_forces(dml_ctx *c, int fuse_b, int only_ermak) {
  if (fuse_b)
    launch_force_lpa<1>(c, c->lpa, only_ermak);
  else
    launch_force_lpa<0>(c, c->lpa, only_ermak);
  hipLaunchKernelGGL(k_ghost_scatter, dim3(c->d.nblocks), dim3(DML_BLOCK), 0, c->st, c->d,
                     only_ermak);
  if (fuse_b)
    hipLaunchKernelGGL(k_ermak_b, dim3(c->d.nblocks), dim3(DML_BLOCK), 0, c->st, c->d, 1);
}

void enq_overlap_rounds(dml_ctx *c, const double *u_round, int64_t stride, int rounds,
                        int err_if_pending) {
  hipLaunchKernelGGL(k_ov_begin, dim3(1), dim3(64), 0, c->st, c->d);
  // rounds from ov_tail on run in one workgroup, one launch (their lists
  // are short; most of them are empty, and an empty round costs three
  // launches otherwise)
  const int grid_rounds = c->ov_tail > 0 && rounds > c->ov_tail ? c->ov_tail : rounds;
  for (int q = 0; q < grid_rounds; q++) {
    // round 0 mostly visits the few listed entries within rcut (the tuned
    // lanes, often 1); later rounds scan whole rows, with at least 2 lanes,
    // and from round 3 on, whose few atoms wait on their rows' loads, at
    // least 4 (C3 +1.6 %; 8 gained nothing)
    const int lpa = q == 0 ? c->lpa_ov : std::max(c->lpa_ov, q >= 3 ? 4 : 2);
    // round 0 covers every ref position; later rounds walk their work lists
    // (grid-stride, a bounded grid)
    unsigned gs = (unsigned)(c->P * lpa / DML_BLOCK);
    const unsigned chunk = DML_BLOCK * OVA_PER;
    unsigned ga = (unsigned)((c->P + chunk - 1) / chunk);
    unsigned gr = (unsigned)(c->P / DML_BLOCK);
    if (q > 0) {
      gs = std::min(gs, 4096u * (unsigned)lpa);
      ga = std::min(ga, 2048u);
      gr = std::min(gr, 4096u);
    }
    switch (lpa) {
      case 1:
        hipLaunchKernelGGL(k_ov_scan_l<1>, dim3(gs), dim3(DML_BLOCK), 0, c->st, c->d, u_round,
                           stride, q, c->P);
        break;
      case 2:
        hipLaunchKernelGGL(k_ov_scan_l<2>, dim3(gs), dim3(DML_BLOCK), 0, c->st, c->d, u_round,
                           stride, q, c->P);
        break;
      case 4:
        hipLaunchKernelGGL(k_ov_scan_l<4>, dim3(gs), dim3(DML_BLOCK), 0, c->st, c->d, u_round,
                           stride, q, c->P);
        break;
      case 16:
        hipLaunchKernelGGL(k_ov_scan_l<16>, dim3(gs), dim3(DML_BLOCK), 0, c->st, c->d, u_round,
                           stride, q, c->P);
        break;
      default:
        hipLaunchKernelGGL(k_ov_scan_l<8>, dim3(gs), dim3(DML_BLOCK), 0, c->st, c->d, u_round,
                           stride, q, c->P);
        break;
    }
    hipLaunchKernelGGL(k_ov_apply, dim3(ga), dim3(DML_BLOCK), 0, c->st, c->d, q, c->P);
    hipLaunchKernelGGL(k_ov_apply_r, dim3(q == 0 ? std::min(gr, 64u) : gr), dim3(DML_BLOCK), 0,
                       c->st, c->d, q);
  }
  if (grid_rounds < rounds) {
    const int lpa = std::max(c->lpa_ov, 4);
    if (lpa >= 16)
      hipLaunchKernelGGL(k_ov_tail<16>, dim3(1), dim3(DML_BLOCK), 0, c->st, c->d, u_round, stride,
                         grid_rounds, rounds, c->P);
    else if (lpa == 8)
      hipLaunchKernelGGL(k_ov_tail<8>, dim3(1), dim3(DML_BLOCK), 0, c->st, c->d, u_round, stride,
                         grid_rounds, rounds, c->P);
    else
      hipLaunchKernelGGL(k_ov_tail<4>, dim3(1), dim3(DML_BLOCK), 0, c->st, c->d, u_round, stride,
                         grid_rounds, rounds, c->P);
  }
  hipLaunchKernelGGL(k_ov_end, dim3((c->nrep + 63) / 64), dim3(64), 0, c->st, c->d, rounds,
                     err_if_pending);
  hipLaunchKernelGGL(k_ov_close, dim3((c->nrep + 63) / 64), dim3(64), 0, c->st, c->d, rounds);
}

void enq_overlap(dml_ctx *c) {
  if (c->any_exact)
    hipLaunchKernelGGL(k_ov_exact, dim3(c->nrep), dim3(64), 0, c->st, c->d, (const double *)nullptr,
                       (int64_t)0, 100000);
  if (c->any_rounds) enq_overlap_rounds(c, nullptr, 0, c->rounds, 0);
}

// gcmc_run (dana.F90:590-713) of one replica (rep_only >= 0, with act and
// nadj given) or of every GCMC replica (their own act, nadj): the grid-wide
// list staging, the sequential trials (a workgroup per replica), the
// grid-wide compaction of the gcmc and ref lists
// the batch kernels of one replica (their grids sized by its pool)
static void enq_gcmc_batch(dml_ctx *c, int r, const double *u, const double *g) {
  const RepParams &p = c->reps[r].p;
  const unsigned nbp = (unsigned)(p.pcap / DML_BLOCK);
  const unsigned cap = (unsigned)p.gb_cap;
  const unsigned wb = (cap * 64 + DML_BLOCK - 1) / DML_BLOCK;  // a wave per trial
  const unsigned tb = (cap + DML_BLOCK - 1) / DML_BLOCK;         // a thread per trial
  hipStream_t st = c->st;
  const DevState &d = c->d;
  hipLaunchKernelGGL(k_gb_count, dim3(nbp), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_scan, dim3(1), dim3(1024), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_scatter, dim3(nbp), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_trials, dim3(wb), dim3(DML_BLOCK), 0, st, d, r, u, g);
  hipLaunchKernelGGL(k_gb_conflicts, dim3(wb), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_decide, dim3(1), dim3(64), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_lists, dim3(tb), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_check, dim3(wb), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_dels, dim3(wb), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_brank, dim3(2 * tb), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_commit, dim3(1), dim3(64), 0, st, d, r, g);
  hipLaunchKernelGGL(k_gb_fields, dim3(tb), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_ownrow, dim3(wb), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_append, dim3(wb), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_append_fix, dim3(256), dim3(DML_BLOCK), 0, st, d, r);
  hipLaunchKernelGGL(k_gb_cells, dim3(2 * cap), dim3(64), 0, st, d, r);
}

static void enq_gcmc(dml_ctx *c, int rep_only, const double *u, const double *g, double act,
                     int nadj) {
  const int nb = c->d.nblocks;
  for (int stage = 0; stage < 2; stage++) {
    if (stage == 1) {
      hipLaunchKernelGGL(k_gb_begin, dim3((c->nrep + 63) / 64), dim3(64), 0, c->st, c->d,
                         rep_only, act, nadj);
      for (int r = 0; r < c->nrep; r++) {
        const RepParams &p = c->reps[r].p;
        if ((rep_only < 0 || r == rep_only) && p.gb_cap > 0 && (p.gcmc_mode & DML_GCMC_BATCHED) &&
            p.reservoir == DML_GCMC)
          enq_gcmc_batch(c, r, u, g);
      }
      if (rep_only >= 0)
        hipLaunchKernelGGL(k_gcmc_run, dim3(1), dim3(DML_GCMC_BLOCK), 0, c->st, c->d, rep_only,
                           u, g, act, nadj);
      else
        hipLaunchKernelGGL(k_gcmc_run, dim3(c->nrep), dim3(DML_GCMC_BLOCK), 0, c->st, c->d, -1,
                           u, g, act, nadj);
    }
    hipLaunchKernelGGL(k_gc_count, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d, rep_only, stage);
    hipLaunchKernelGGL(k_gc_scan, dim3(c->nrep), dim3(1024), 0, c->st, c->d, rep_only, stage);
    hipLaunchKernelGGL(k_gc_scatter, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d, rep_only, stage);
  }
  hipLaunchKernelGGL(k_gc_copy, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d, rep_only);
}

void enq_f2cg(dml_ctx *c) {
  int nb = c->d.nblocks;
  hipLaunchKernelGGL(k_f2cg_mark, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d);
  hipLaunchKernelGGL(k_f2cg_scan, dim3(c->nrep), dim3(1024), 0, c->st, c->d);
  hipLaunchKernelGGL(k_f2cg_scatter, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d);
  hipLaunchKernelGGL(k_f2cg_copy, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d);
  hipLaunchKernelGGL(k_f2cg_end, dim3((c->nrep + 63) / 64), dim3(64), 0, c->st, c->d);
}

// bloques, then test_update of the replicas that added a block (dana.F90:771)
void enq_bloques(dml_ctx *c) {
  const int rb = (c->nrep + 63) / 64;
  hipLaunchKernelGGL(k_chunk_run, dim3(c->nrep), dim3(DML_BLOCK), 0, c->st, c->d);
  hipLaunchKernelGGL(k_mask_fired, dim3(rb), dim3(64), 0, c->st, c->d, 1);
  enq_test_update(c, 1);
  hipLaunchKernelGGL(k_mask_fired, dim3(rb), dim3(64), 0, c->st, c->d, 0);
}

void enq_step(dml_ctx *c) {
  int nb = c->d.nblocks;
  if (c->any_ermak) {
    hipLaunchKernelGGL(k_ermak_a, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d,
                       (const double *)nullptr, (const double *)nullptr);
    enq_forces(c, 1, 1);
  }
  if (c->any_brown)
    hipLaunchKernelGGL(k_brownian, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d,
                       (const double *)nullptr, (const double *)nullptr);
  // the cells of a GCMC replica are needed by its trials, after the second
  // test_update; the first re-sorts them only when it rebuilds the lists
  enq_test_update(c, 0);
  enq_overlap(c);
  enq_test_update(c, 1);
  enq_f2cg(c);
  if (c->any_gcmc)  // gcmc_run after F -> CG (dana.F90:238-241)
    enq_gcmc(c, -1, nullptr, nullptr, 0.0, 0);
  // calc_rho (dana.F90:244), bloques (:247), maxz (:261)
  if (c->any_piston || c->any_chunks)
    hipLaunchKernelGGL(k_calc_rho, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d);
  if (c->any_chunks) enq_bloques(c);
  if (c->any_piston) hipLaunchKernelGGL(k_maxz, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d);
  hipLaunchKernelGGL(k_step_end, dim3((c->nrep + 63) / 64), dim3(64), 0, c->st, c->d);
}

// grow neighbour rows or buckets after an overflow and redo the rebuild of
// the replicas whose rebuild flag is still set.  An overflowing replica halts
// in the kernel; with resume it is released and rebuilt (the call that
// overflowed was the rebuild itself), without it (after whole steps, which
// went on without it) it stays halted with its status.
int handle_overflow(dml_ctx *c, int e, int resume) {
  for (int attempt = 0; attempt < 8 && (e & (ERR_LIST | ERR_BUCKET)); attempt++) {
    int rc = fetch_params(c);
    if (rc) return rc;
    if (e & ERR_LIST) {
      int need = 0;
      for (auto &h : c->reps) need = std::max(need, h.p.max_nn_seen);
      int nm = std::max(next_pow2(need + need / 4 + 1), c->d.maxnn * 2);
      int32_t *nbr2 = nullptr;
      if (dalloc(&nbr2, c->P * (int64_t)nm)) return DML_E_NOMEM;
      // keep the rows of replicas that are not rebuilt
      HIPCHECK(hipMemcpy2DAsync(nbr2, sizeof(int32_t) * nm, c->d.nbr, sizeof(int32_t) * c->d.maxnn,
                                sizeof(int32_t) * c->d.maxnn, (size_t)c->P, hipMemcpyDeviceToDevice,
                                c->st));
      HIPCHECK(hipStreamSynchronize(c->st));
      hipFree(c->d.nbr);
      c->d.nbr = nbr2;
      c->d.maxnn = nm;
    }
    if (e & ERR_BUCKET) {
      for (auto &h : c->reps) {
        if (!h.p.rebuild && h.p.reservoir != DML_GCMC) continue;
        if (h.p.use_cells) h.p.bcap *= 2;
      }
      int rc2 = alloc_cells(c);
      if (rc2) return rc2;
    }
    {
      // patch the layout fields and the overflow halts into the device table
      std::vector<RepParams> v(c->nrep);
      HIPCHECK(hipMemcpy(v.data(), c->d.rp, sizeof(RepParams) * c->nrep, hipMemcpyDeviceToHost));
      for (int r = 0; r < c->nrep; r++) {
        v[r].coff = c->reps[r].p.coff;
        v[r].ccap = c->reps[r].p.ccap;
        v[r].boff = c->reps[r].p.boff;
        v[r].bcap = c->reps[r].p.bcap;
        const int ovf = ERR_LIST | ERR_BUCKET;
        if (resume && v[r].err_bits && !(v[r].err_bits & ~ovf)) {
          v[r].err_bits = 0;
          v[r].halted &= ~HALT_FATAL;
        }
        c->reps[r].p = v[r];
      }
      HIPCHECK(hipMemcpy(c->d.rp, v.data(), sizeof(RepParams) * c->nrep, hipMemcpyHostToDevice));
    }
    int z = 0;
    HIPCHECK(hipMemcpyAsync(c->d.err, &z, sizeof(int), hipMemcpyHostToDevice, c->st));
    enq_cells_lists(c, 1);
    e = read_err(c, 0);
    if (e < 0) return DML_E_HIP;
  }
  return DML_OK;
}

// replica_fatal: report a replica's halt as the call's status (per-routine
// entry points); dml_run leaves halted replicas to dml_replica_status()
int finish(dml_ctx *c, int replica_fatal = 1, int resume = 1) {
  int e = read_err(c, 0);
  if (e < 0) return DML_E_HIP;
  if (e & (ERR_LIST | ERR_BUCKET)) {
    int rc = handle_overflow(c, e, resume);
    if (rc) return rc;
    e = read_err(c, 0);
  }
  if (e) {
    int z = 0;
    hipMemcpy(c->d.err, &z, sizeof(int), hipMemcpyHostToDevice);
    return status_from_err(e);
  }
  if (replica_fatal) {
    int rc = fetch_params(c);
    if (rc) return rc;
    for (auto &h : c->reps)
      if (h.p.err_bits) return status_from_err(h.p.err_bits);
  }
  return DML_OK;
}

double *upload_tmp(dml_ctx *c, const double *h, int64_t n) {
  if (!h) return nullptr;
  double *d = nullptr;
  if (hipMalloc((void **)&d, sizeof(double) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess)
    return nullptr;
  hipMemcpyAsync(d, h, sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->st);
  return d;
}

void refresh_modes(dml_ctx *c) {
  c->any_ermak = c->any_brown = c->any_rounds = c->any_exact = c->any_piston = c->any_gcmc = 0;
  c->any_chunks = 0;
  for (auto &h : c->reps) {
    if (h.p.reservoir == DML_GCMC) c->any_gcmc = 1;
    if (h.p.reservoir == DML_CHUNKS && h.p.nchunk > 0) c->any_chunks = 1;
    if (h.p.integrator == DML_ERMAK) c->any_ermak = 1; else c->any_brown = 1;
    if (h.p.overlap_mode == DML_OVERLAP_ROUNDS) c->any_rounds = 1; else c->any_exact = 1;
    if (h.p.piston) c->any_piston = 1;
  }
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *dml_strerror(int s) {
  switch (s) {
    case DML_OK: return "ok";
    case DML_E_ARG: return "bad argument";
    case DML_E_HIP: return "HIP runtime error";
    case DML_E_NOMEM: return "out of device memory or capacity";
    case DML_E_OUT_OF_TESS: return "particle out of tessellation";
    case DML_E_LIST_FULL: return "neighbour row capacity exceeded";
    case DML_E_DEPOSIT_Z0: return "deposit above z0";
    case DML_E_NO_PARTICLE: return "chosen particle does not exist";
    case DML_E_OVERLAP: return "overlap resolution did not finish";
    case DML_E_STATE: return "call not valid in this state";
    case DML_E_UNSUPPORTED: return "unsupported configuration";
  }
  return "unknown status";
}

int32_t dml_num_replicas(const dml_ctx *c) { return c ? c->nrep : 0; }

int64_t dml_total_slots(const dml_ctx *c) { return c ? c->P : 0; }

int64_t dml_ref_offset(const dml_ctx *c, int32_t rep) {
  if (!c || rep < 0 || rep >= c->nrep) return -1;
  return c->reps[rep].p.poff;
}

int dml_create(const dml_config *cfg, dml_ctx **out) { return dml_create_ensemble(1, cfg, out); }

int dml_create_ensemble(int32_t nrep, const dml_config *cfgs, dml_ctx **out) {
  if (!out || !cfgs || nrep <= 0) return DML_E_ARG;
  *out = nullptr;
  dml_ctx *c = new dml_ctx();
  c->nrep = nrep;
  c->reps.resize(nrep);
  int64_t P = 0;
  int maxnn = 0;
  int lpa = 0;
  for (int r = 0; r < nrep; r++) {
    const dml_config &g = cfgs[r];
    if (g.capacity <= 0 || g.box[0] <= 0 || g.box[1] <= 0 || g.box[2] <= 0 || g.rcut <= 0 ||
        g.nadj < 0 || g.nadj > (1 << 21)) {
      delete c;
      return DML_E_ARG;
    }
    HostRep &h = c->reps[r];
    h.cfg = g;
    RepParams &p = h.p;
    memset(&p, 0, sizeof(p));
    p.pcap = (int32_t)round_up(g.capacity, DML_BLOCK);  // before setup_rep: Verlet buckets hold pcap
    if (setup_rep(p, g)) {
      delete c;
      return DML_E_UNSUPPORTED;
    }
    p.poff = (int32_t)P;
    P += p.pcap;
    if ((g.gcmc_mode & DML_GCMC_BATCHED) && g.reservoir == DML_GCMC) {
      p.gb_off = (int32_t)c->gb_total;
      p.gb_cap = DML_GCMC_BATCH_MAX;
      c->gb_total += DML_GCMC_BATCH_MAX;
    }
    maxnn = std::max(maxnn, rows_for(g));
    if (g.lanes_per_atom > 0) lpa = g.lanes_per_atom;
  }
  if (P > INT32_MAX / 2) {
    delete c;
    return DML_E_ARG;
  }
  c->P = P;
  if (lpa) c->lpa = lpa;
  if (hipGetDevice(&c->dev) != hipSuccess) {
    delete c;
    return DML_E_HIP;
  }
  if (int rc = scratch_setup(c->dev, &c->scratch)) {
    delete c;
    return rc;
  }
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return DML_E_HIP;
  }
  DevState &d = c->d;
  int e = 0;
  e |= dalloc(&d.px, P); e |= dalloc(&d.py, P); e |= dalloc(&d.pz, P);
  e |= dalloc(&d.vx, P); e |= dalloc(&d.vy, P); e |= dalloc(&d.vz, P);
  e |= dalloc(&d.ax, P); e |= dalloc(&d.ay, P); e |= dalloc(&d.az, P);
  e |= dalloc(&d.fx, P); e |= dalloc(&d.fy, P); e |= dalloc(&d.fz, P);
  e |= dalloc(&d.epot, P);
  e |= dalloc(&d.ox, P); e |= dalloc(&d.oy, P); e |= dalloc(&d.oz, P);
  e |= dalloc(&d.cx, P); e |= dalloc(&d.cy, P); e |= dalloc(&d.cz, P);
  e |= dalloc(&d.rvx, P); e |= dalloc(&d.rvy, P); e |= dalloc(&d.rvz, P);
  e |= dalloc(&d.species, P);
  e |= dalloc(&d.flags, P);
  e |= dalloc(&d.rowflag, P);
  e |= dalloc(&d.bkey, P);
  e |= dalloc(&d.pid, P);
  e |= dalloc(&d.ref_order, P);
  e |= dalloc(&d.ref_tmp, P);
  e |= dalloc(&d.ref_pos, P);
  e |= dalloc(&d.nn, P);
  e |= dalloc(&d.nbr, P * (int64_t)maxnn);
  e |= dalloc(&d.ov_rev, P);
  e |= dalloc(&d.ov_st, P);
  e |= dalloc(&d.ov_near, P);
  e |= dalloc(&d.nearl, P * (int64_t)NEAR_MAX);
  e |= dalloc(&d.part_d, 2 * (P / DML_BLOCK));
  e |= dalloc(&d.obs, 4 * (P / DML_BLOCK) + 4 * (int64_t)nrep);
  e |= dalloc(&d.part_i, 2 * (P / DML_BLOCK));
  e |= dalloc(&d.blk2rep, P / DML_BLOCK);
  e |= dalloc(&d.rp, nrep);
  e |= dalloc(&d.cnt, nrep);
  e |= dalloc(&d.err, 1);
  e |= dalloc(&d.wl[0], P);
  e |= dalloc(&d.wl[1], P);
  e |= dalloc(&d.wl_cnt, 4);
  e |= dalloc(&d.ovw[0], P);
  e |= dalloc(&d.ovw[1], P);
  e |= dalloc(&d.ovr[0], P);
  e |= dalloc(&d.ovr[1], P);
  e |= dalloc(&d.ov_more, 3 * nrep);
  e |= dalloc(&d.pf4, P);
  e |= dalloc(&d.gc_order, P);
  e |= dalloc(&d.b2slot, P);
  e |= dalloc(&d.seq, P);
  {
    // batched GCMC: per trial of the pool, per slot only with a batched replica
    const int64_t T = std::max<int64_t>(c->gb_total, 1), PS = c->gb_total ? P : 1;
    e |= dalloc(&d.gb_pt, 3 * T); e |= dalloc(&d.gb_u, 2 * T); e |= dalloc(&d.gb_du, T);
    e |= dalloc(&d.gb_g, 3 * T);
    e |= dalloc(&d.gb_kind, T); e |= dalloc(&d.gb_cell, T); e |= dalloc(&d.gb_nblk, T);
    e |= dalloc(&d.gb_blk, 16 * T); e |= dalloc(&d.gb_ncf, T); e |= dalloc(&d.gb_cf, 16 * T);
    e |= dalloc(&d.gb_res, T); e |= dalloc(&d.gb_slot, T);
    e |= dalloc(&d.gi_trial, T); e |= dalloc(&d.gi_next, T); e |= dalloc(&d.gi_del, T);
    e |= dalloc(&d.gi_tmpl, T);
    e |= dalloc(&d.gd_trial, T); e |= dalloc(&d.gd_vpos, T); e |= dalloc(&d.gd_slot, T);
    e |= dalloc(&d.gd_bso, T); e |= dalloc(&d.gd_brank, T); e |= dalloc(&d.gd_rpos, T);
    e |= dalloc(&d.gd_oflags, T); e |= dalloc(&d.gd_cell, T);
    e |= dalloc(&d.gb_free, T); e |= dalloc(&d.gb_freeb, T); e |= dalloc(&d.gb_frank, T);
    e |= dalloc(&d.gb_U, 3 * T);
    e |= dalloc(&d.gb_part, 2 * (P / DML_BLOCK));
    e |= dalloc(&d.gb_vpos, PS); e |= dalloc(&d.gb_delt, PS); e |= dalloc(&d.gb_alist, PS);
    e |= dalloc(&d.gb_app, PS);
    if (!e) {
      // the stamped counters start empty; the records are validated on use
      hipMemsetAsync(d.gb_app, 0, sizeof(unsigned long long) * PS, c->st);
      hipMemsetAsync(d.gb_delt, 0xff, sizeof(int32_t) * PS, c->st);
      hipMemsetAsync(d.gb_vpos, 0xff, sizeof(int32_t) * PS, c->st);
      hipMemsetAsync(d.gb_slot, 0xff, sizeof(int32_t) * T, c->st);
    }
  }
  if (e) {
    dml_destroy(c);
    return DML_E_NOMEM;
  }
  d.maxnn = maxnn;
  d.nblocks = (int32_t)(P / DML_BLOCK);
  d.nrep = nrep;
  // zero state buffers once
  double *dbl[] = {d.px, d.py, d.pz, d.vx, d.vy, d.vz, d.ax, d.ay, d.az, d.fx, d.fy, d.fz,
                   d.epot, d.ox, d.oy, d.oz, d.cx, d.cy, d.cz, d.rvx, d.rvy, d.rvz};
  for (double *q : dbl) hipMemsetAsync(q, 0, sizeof(double) * P, c->st);
  hipMemsetAsync(d.species, 0, P, c->st);
  hipMemsetAsync(d.flags, 0, P, c->st);
  hipMemsetAsync(d.rowflag, 0, P, c->st);
  hipMemsetAsync(d.nn, 0, sizeof(int32_t) * P, c->st);
  hipMemsetAsync(d.ov_rev, 0, sizeof(int32_t) * P, c->st);
  hipMemsetAsync(d.ov_st, 0, P, c->st);
  hipMemsetAsync(d.ov_near, 0xff, P, c->st);  // 255: no entries listed
  hipMemsetAsync(d.wl_cnt, 0, sizeof(int32_t) * 4, c->st);
  hipMemsetAsync(d.pf4, 0, sizeof(float4) * P, c->st);
  hipMemsetAsync(d.ref_pos, 0xff, sizeof(int32_t) * P, c->st);
  hipMemsetAsync(d.b2slot, 0xff, sizeof(int32_t) * P, c->st);
  hipMemsetAsync(d.gc_order, 0xff, sizeof(int32_t) * P, c->st);
  hipMemsetAsync(d.seq, 0, sizeof(int32_t) * P, c->st);
  hipMemsetAsync(d.cnt, 0, sizeof(DevCounters) * nrep, c->st);
  hipMemsetAsync(d.err, 0, sizeof(int32_t), c->st);
  std::vector<int32_t> b2r(P / DML_BLOCK);
  for (int r = 0; r < nrep; r++)
    for (int q = 0; q < c->reps[r].p.pcap / DML_BLOCK; q++)
      b2r[c->reps[r].p.poff / DML_BLOCK + q] = r;
  hipMemcpyAsync(d.blk2rep, b2r.data(), sizeof(int32_t) * b2r.size(), hipMemcpyHostToDevice, c->st);
  if (alloc_cells(c) || upload_params(c)) {
    dml_destroy(c);
    return DML_E_NOMEM;
  }
  hipEventCreate(&c->e0);
  hipEventCreate(&c->e1);
  if (const char *ev = getenv("DML_LISTBUILD")) c->lb_mode = atoi(ev);
  if (const char *ev = getenv("DML_OV_TAIL")) c->ov_tail = std::max(0, atoi(ev));
  if (const char *ev = getenv("DML_OVERLAP_LPA")) {
    const int v = atoi(ev);
    if (v == 1 || v == 2 || v == 4 || v == 8 || v == 16) c->lpa_ov = v;
  }

  refresh_modes(c);
  if (hipStreamSynchronize(c->st) != hipSuccess) {
    dml_destroy(c);
    return DML_E_HIP;
  }
  *out = c;
  return DML_OK;
}

int dml_destroy(dml_ctx *c) {
  if (!c) return DML_OK;
  if (c->st) hipStreamSynchronize(c->st);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  DevState &d = c->d;
  void *ptrs[] = {d.px, d.py, d.pz, d.vx, d.vy, d.vz, d.ax, d.ay, d.az, d.fx, d.fy, d.fz,
                  d.epot, d.ox, d.oy, d.oz, d.cx, d.cy, d.cz, d.rvx, d.rvy, d.rvz,
                  d.species, d.flags, d.rowflag, d.bkey, d.pid, d.ref_order, d.ref_tmp,
                  d.ref_pos, d.nn, d.nbr, d.cell_cnt, d.bucket, d.bx, d.by, d.bz, d.ov_rev,
                  d.ov_st, d.ov_near, d.nearl, d.part_d, d.part_i, d.blk2rep, d.rp, d.cnt, d.err, d.wl[0],
                  d.wl[1], d.wl_cnt, d.ovw[0], d.ovw[1], d.ovr[0], d.ovr[1], d.ov_more, d.pf4, d.bf4, d.gc_order,
                  d.b2slot, d.seq, d.chx, d.chy, d.chz, d.chox, d.choy, d.choz,
                  d.cm_rep, d.cm_cell, d.tm_rep, d.tm_cell, d.wm_blk, d.obs,
                  d.gb_pt, d.gb_u, d.gb_du, d.gb_g, d.gb_kind, d.gb_cell, d.gb_nblk, d.gb_blk,
                  d.gb_ncf, d.gb_cf, d.gb_res, d.gb_slot, d.gi_trial, d.gi_next, d.gi_del,
                  d.gi_tmpl, d.gd_trial, d.gd_vpos, d.gd_slot, d.gd_bso, d.gd_brank,
                  d.gd_rpos, d.gd_oflags, d.gd_cell, d.gb_free, d.gb_freeb, d.gb_frank,
                  d.gb_U, d.gb_part, d.gb_vpos, d.gb_delt, d.gb_alist, d.gb_app,
                  d.gb_chead, d.gb_ccnt, d.gb_claim};
  for (void *q : ptrs)
    if (q) hipFree(q);
  if (c->e0) hipEventDestroy(c->e0);
  if (c->e1) hipEventDestroy(c->e1);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
  return DML_OK;
}

int dml_sync(dml_ctx *c) {
  if (!c) return DML_E_ARG;
  return finish(c);
}

int dml_set_state(dml_ctx *c, int32_t rep, int32_t n, const double *pos, const double *vel,
                  const double *acel, const double *pos_old, const double *old_cg,
                  const int8_t *species, const uint8_t *flags) {
  if (!c || rep < 0 || rep >= c->nrep || n < 0 || !pos) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  RepParams &p = c->reps[rep].p;
  if (n > p.pcap) return DML_E_NOMEM;
  int64_t o = p.poff;
  std::vector<double> tmp(std::max(n, 1));
  auto put = [&](double *dst, const double *src, int k, double dflt) -> int {
    for (int i = 0; i < n; i++) tmp[i] = src ? src[(int64_t)k * n + i] : dflt;
    HIPCHECK(hipMemcpy(dst + o, tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    return DML_OK;
  };
  double *P3[3] = {c->d.px, c->d.py, c->d.pz};
  double *V3[3] = {c->d.vx, c->d.vy, c->d.vz};
  double *A3[3] = {c->d.ax, c->d.ay, c->d.az};
  double *O3[3] = {c->d.ox, c->d.oy, c->d.oz};
  double *G3[3] = {c->d.cx, c->d.cy, c->d.cz};
  for (int k = 0; k < 3; k++) {
    if ((rc = put(P3[k], pos, k, 0.0))) return rc;
    if ((rc = put(V3[k], vel, k, 0.0))) return rc;
    if ((rc = put(A3[k], acel, k, 0.0))) return rc;
    if ((rc = put(O3[k], pos_old ? pos_old : pos, k, 0.0))) return rc;
    if ((rc = put(G3[k], old_cg, k, 1e8))) return rc;
  }
  std::vector<int8_t> sp(std::max(n, 1));
  std::vector<uint8_t> fl(std::max(n, 1));
  std::vector<int32_t> bk(std::max(n, 1)), rord, rpos(std::max(n, 1), -1);
  std::vector<uint32_t> pid(std::max(n, 1));
  for (int i = 0; i < n; i++) {
    sp[i] = species ? species[i] : (int8_t)SP_LI;
    fl[i] = flags ? flags[i]
                  : (uint8_t)(DML_FL_ALIVE | (sp[i] != SP_CG ? DML_FL_REF : 0) |
                              ((p.reservoir == DML_GCMC && sp[i] == SP_LI) ? DML_FL_GCMC : 0));
    bk[i] = i;
    pid[i] = (uint32_t)i;
    if (fl[i] & DML_FL_REF) {
      rpos[i] = (int32_t)rord.size();
      rord.push_back((int32_t)(o + i));
    }
  }
  HIPCHECK(hipMemcpy(c->d.species + o, sp.data(), n, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->d.flags + o, fl.data(), n, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->d.bkey + o, bk.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->d.pid + o, pid.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->d.ref_pos + o, rpos.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  if (!rord.empty())
    HIPCHECK(hipMemcpy(c->d.ref_order + o, rord.data(), sizeof(int32_t) * rord.size(),
                       hipMemcpyHostToDevice));
  HIPCHECK(hipMemset(c->d.nn + o, 0, sizeof(int32_t) * p.pcap));
  // a reused replica (serving) starts like a fresh one: no overlap stamps,
  // no flags beyond n, zero forces
  HIPCHECK(hipMemset(c->d.ov_rev + o, 0, sizeof(int32_t) * p.pcap));
  HIPCHECK(hipMemset(c->d.ov_st + o, 0, p.pcap));
  HIPCHECK(hipMemset(c->d.rowflag + o, 0, p.pcap));
  if (p.pcap > n) HIPCHECK(hipMemset(c->d.flags + o + n, 0, p.pcap - n));
  {
    double *z8[] = {c->d.fx, c->d.fy, c->d.fz, c->d.epot, c->d.rvx, c->d.rvy, c->d.rvz};
    for (double *q : z8) HIPCHECK(hipMemset(q + o, 0, sizeof(double) * p.pcap));
  }
  p.ov_round = 0;
  p.ov_pending = 0;
  p.rho_count = 0;
  p.any_ghost = 0;
  p.lohi = 0.0;
  p.d1 = p.d2 = 0.0;
  // hs%b index table, the gcmc group in sys order (dana.F90:505-513) and
  // the creation numbers
  {
    std::vector<int32_t> b2s(p.pcap, -1), gco, sq(std::max(n, 1));
    int nat = 0;
    for (int i = 0; i < n; i++) {
      sq[i] = i;
      if (fl[i] & DML_FL_ALIVE) {
        nat++;
        b2s[i] = (int32_t)(o + i);
      }
      if (fl[i] & DML_FL_GCMC) gco.push_back((int32_t)(o + i));
    }
    HIPCHECK(hipMemcpy(c->d.b2slot + o, b2s.data(), sizeof(int32_t) * p.pcap, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->d.seq + o, sq.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    if (!gco.empty())
      HIPCHECK(hipMemcpy(c->d.gc_order + o, gco.data(), sizeof(int32_t) * gco.size(),
                         hipMemcpyHostToDevice));
    p.nat = nat;
    p.nlimbo = 0;
    p.b_amax = n;
    p.b_nat = nat;
    p.ngc = (int32_t)gco.size();
    p.pid_next = (uint32_t)n;
    p.seq_next = n;
    p.gcmc_calls = 0;
  }
  p.free_lo = 0;
  p.bfree_lo = 0;
  p.ovn_stamp = 0;
  // no entries-within-rcut lists for the new atoms until fuerza or a row
  // build writes them
  HIPCHECK(hipMemsetAsync(c->d.ov_near + o, 0xff, (size_t)p.pcap, c->st));
  p.nslots = n;
  p.nref = (int32_t)rord.size();
  p.listed = 0;
  p.rebuild = 0;
  p.any_f = 0;
  p.step = 0;
  p.halted = 0;
  p.err_bits = 0;
  if (p.piston) {
    // initial calc_rho (dana.F90:145-146)
    int g = 0;
    for (int i = 0; i < n; i++) {
      double z = pos[2 * (int64_t)n + i];
      if (z > p.z0 && z < p.zmax) g++;
    }
    p.rho0 = (double)g / (p.box[0] * p.box[1] * (p.zmax - p.z0));
  }
  rc = upload_params(c);
  if (rc) return rc;
  hipLaunchKernelGGL(k_init_pf4, dim3((unsigned)((c->P + DML_BLOCK - 1) / DML_BLOCK)),
                     dim3(DML_BLOCK), 0, c->st, c->d, c->P);
  HIPCHECK(hipStreamSynchronize(c->st));
  return DML_OK;
}

int dml_set_order(dml_ctx *c, int32_t rep, int32_t nref, const int32_t *ref_order,
                  const int32_t *bkey) {
  if (!c || rep < 0 || rep >= c->nrep) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  RepParams &p = c->reps[rep].p;
  int64_t o = p.poff;
  int n = p.nslots;
  if (ref_order) {
    if (nref < 0 || nref > n) return DML_E_ARG;
    std::vector<int32_t> g(std::max(nref, 1)), rpos(std::max(n, 1), -1);
    std::vector<uint8_t> fl(std::max(n, 1));
    HIPCHECK(hipMemcpy(fl.data(), c->d.flags + o, n, hipMemcpyDeviceToHost));
    for (int t = 0; t < nref; t++) {
      const int i = ref_order[t];
      // each listed atom once, and alive
      if (i < 0 || i >= n || rpos[i] >= 0 || !(fl[i] & DML_FL_ALIVE)) return DML_E_ARG;
      g[t] = (int32_t)(o + i);
      rpos[i] = t;
    }
    // hs%ref membership is exactly the list: the force and overlap kernels
    // read it from the flags
    for (int i = 0; i < n; i++)
      fl[i] = (uint8_t)(rpos[i] >= 0 ? (fl[i] | DML_FL_REF) : (fl[i] & ~DML_FL_REF));
    HIPCHECK(hipMemcpy(c->d.flags + o, fl.data(), n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->d.ref_order + o, g.data(), sizeof(int32_t) * nref, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->d.ref_pos + o, rpos.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    p.nref = nref;
  }
  if (bkey) {
    std::vector<uint8_t> fl(std::max(n, 1));
    HIPCHECK(hipMemcpy(fl.data(), c->d.flags + o, n, hipMemcpyDeviceToHost));
    std::vector<int32_t> b2s(p.pcap, -1);
    int amax = 0;
    for (int i = 0; i < n; i++) {
      if (!(fl[i] & DML_FL_ALIVE)) continue;
      if (bkey[i] < 0 || bkey[i] >= p.pcap || b2s[bkey[i]] >= 0) return DML_E_ARG;
      b2s[bkey[i]] = (int32_t)(o + i);
      amax = std::max(amax, bkey[i] + 1);
    }
    HIPCHECK(hipMemcpy(c->d.bkey + o, bkey, sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(c->d.b2slot + o, b2s.data(), sizeof(int32_t) * p.pcap, hipMemcpyHostToDevice));
    p.b_amax = amax;
    p.bfree_lo = 0;
  }
  p.ovn_stamp = 0;  // another set of ref atoms: no lists for the new ones
  return upload_params(c);
}

int dml_get_state(dml_ctx *c, int32_t rep, int32_t n, double *pos, double *vel, double *acel,
                  double *force, double *epot, double *pos_old, double *old_cg, int8_t *species,
                  uint8_t *flags) {
  if (!c || rep < 0 || rep >= c->nrep || n < 0) return DML_E_ARG;
  HIPCHECK(hipStreamSynchronize(c->st));
  const RepParams &p = c->reps[rep].p;
  if (n > p.pcap) return DML_E_ARG;
  int64_t o = p.poff;
  auto get3 = [&](double *dst, double *const *src) -> int {
    if (!dst) return DML_OK;
    for (int k = 0; k < 3; k++)
      HIPCHECK(hipMemcpy(dst + (int64_t)k * n, src[k] + o, sizeof(double) * n, hipMemcpyDeviceToHost));
    return DML_OK;
  };
  double *P3[3] = {c->d.px, c->d.py, c->d.pz};
  double *V3[3] = {c->d.vx, c->d.vy, c->d.vz};
  double *A3[3] = {c->d.ax, c->d.ay, c->d.az};
  double *F3[3] = {c->d.fx, c->d.fy, c->d.fz};
  double *O3[3] = {c->d.ox, c->d.oy, c->d.oz};
  double *G3[3] = {c->d.cx, c->d.cy, c->d.cz};
  int rc;
  if ((rc = get3(pos, P3)) || (rc = get3(vel, V3)) || (rc = get3(acel, A3)) ||
      (rc = get3(force, F3)) || (rc = get3(pos_old, O3)) || (rc = get3(old_cg, G3)))
    return rc;
  if (epot) HIPCHECK(hipMemcpy(epot, c->d.epot + o, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (species) HIPCHECK(hipMemcpy(species, c->d.species + o, n, hipMemcpyDeviceToHost));
  if (flags) HIPCHECK(hipMemcpy(flags, c->d.flags + o, n, hipMemcpyDeviceToHost));
  return DML_OK;
}

int dml_get_ref_order(dml_ctx *c, int32_t rep, int32_t *nref, int32_t *ref_order) {
  if (!c || rep < 0 || rep >= c->nrep || !nref) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  const RepParams &p = c->reps[rep].p;
  *nref = p.nref;
  if (ref_order && p.nref > 0) {
    std::vector<int32_t> g(p.nref);
    HIPCHECK(hipMemcpy(g.data(), c->d.ref_order + p.poff, sizeof(int32_t) * p.nref,
                       hipMemcpyDeviceToHost));
    for (int t = 0; t < p.nref; t++) ref_order[t] = g[t] - p.poff;
  }
  return DML_OK;
}

int dml_set_zmax(dml_ctx *c, int32_t rep, double zmax, double z0) {
  if (!c || rep < 0 || rep >= c->nrep) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  c->reps[rep].p.zmax = zmax;
  c->reps[rep].p.z0 = z0;
  return upload_params(c);
}

int dml_get_zmax(dml_ctx *c, int32_t rep, double *zmax) {
  if (!c || rep < 0 || rep >= c->nrep || !zmax) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  *zmax = c->reps[rep].p.zmax;
  return DML_OK;
}

int dml_set_box(dml_ctx *c, int32_t rep, const double box[3]) {
  if (!c || rep < 0 || rep >= c->nrep || !box) return DML_E_ARG;
  for (int k = 0; k < 3; k++)
    if (!(box[k] > 0)) return DML_E_ARG;
    else if (box[k] > DML_MAX_BOX) return DML_E_UNSUPPORTED;
  int rc = fetch_params(c);
  if (rc) return rc;
  RepParams &p = c->reps[rep].p;
  for (int k = 0; k < 3; k++) {
    p.box[k] = box[k];
    p.one_box[k] = 1.0 / box[k];
  }
  p.ovn_stamp = 0;  // distances change with the box
  int oldcap = p.ccap;
  if ((rc = tessellate(p, 0))) return rc;
  int nb = bucket_cap(p);
  if (ncells_total(p) > oldcap || nb > p.bcap) {
    p.bcap = std::max(p.bcap, nb);
    if ((rc = alloc_cells(c))) return rc;
  } else {
    p.ccap = oldcap;
  }
  if (c->gexec) {
    hipGraphExecDestroy(c->gexec);
    hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
  }
  return upload_params(c);
}

static int ensure_cells(dml_ctx *c);

int dml_test_update(dml_ctx *c, int32_t *rebuilt) {
  if (!c) return DML_E_ARG;
  if (int rc0 = ensure_cells(c)) return rc0;
  enq_test_update(c, 1);
  int rc = finish(c);
  if (rc) return rc;
  if (rebuilt) {
    if ((rc = fetch_params(c))) return rc;
    for (int r = 0; r < c->nrep; r++) rebuilt[r] = c->reps[r].p.rebuild;
  }
  return DML_OK;
}

int dml_update(dml_ctx *c) {
  if (!c) return DML_E_ARG;
  if (int rc0 = ensure_cells(c)) return rc0;
  hipLaunchKernelGGL(k_force_rebuild, dim3((c->nrep + 63) / 64), dim3(64), 0, c->st, c->d);
  enq_cells_lists(c, 1);
  return finish(c);
}

int dml_cells(dml_ctx *c, int32_t rep, int32_t ncells[3], double cell[3], int32_t *use_cells,
              int32_t *cell_of) {
  if (!c || rep < 0 || rep >= c->nrep) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  const RepParams &p = c->reps[rep].p;
  for (int k = 0; k < 3; k++) {
    if (ncells) ncells[k] = p.ncells[k];
    if (cell) cell[k] = p.cell[k];
  }
  if (use_cells) *use_cells = p.use_cells;
  if (cell_of) {
    int n = p.nslots;
    for (int i = 0; i < 3 * n; i++) cell_of[i] = -1;
    if (!p.use_cells) return DML_OK;
    int nc = ncells_total(p);
    std::vector<int32_t> cnt(nc);
    std::vector<int32_t> bk((size_t)nc * p.bcap);
    HIPCHECK(hipMemcpy(cnt.data(), c->d.cell_cnt + p.coff, sizeof(int32_t) * nc, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(bk.data(), c->d.bucket + p.boff, sizeof(int32_t) * bk.size(),
                       hipMemcpyDeviceToHost));
    int nx = p.ncells[0] + 2, ny = p.ncells[1] + 2;
    for (int q = 0; q < nc; q++) {
      int ci = q % nx, cj = (q / nx) % ny, ck = q / (nx * ny);
      for (int e = 0; e < std::min(cnt[q], p.bcap); e++) {
        int s = bk[(size_t)q * p.bcap + e] - p.poff;
        if (s < 0 || s >= n) continue;
        cell_of[s] = ci;
        cell_of[n + s] = cj;
        cell_of[2 * n + s] = ck;
      }
    }
  }
  return DML_OK;
}

int dml_neighbors(dml_ctx *c, int32_t rep, int32_t *nn, int64_t *rowptr, int32_t *idx,
                  int64_t cap, int64_t *total) {
  if (!c || rep < 0 || rep >= c->nrep) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  const RepParams &p = c->reps[rep].p;
  int n = p.nslots;
  std::vector<int32_t> hn(std::max(n, 1));
  std::vector<uint8_t> fl(std::max(n, 1));
  HIPCHECK(hipMemcpy(hn.data(), c->d.nn + p.poff, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(fl.data(), c->d.flags + p.poff, n, hipMemcpyDeviceToHost));
  int64_t tot = 0;
  for (int i = 0; i < n; i++) {
    if (!(fl[i] & DML_FL_REF)) hn[i] = 0;
    tot += hn[i];
  }
  if (total) *total = tot;
  if (nn) memcpy(nn, hn.data(), sizeof(int32_t) * n);
  if (!idx && !rowptr) return DML_OK;
  int mx = c->d.maxnn;
  std::vector<int32_t> rows((size_t)n * mx);
  HIPCHECK(hipMemcpy(rows.data(), c->d.nbr + (int64_t)p.poff * mx, sizeof(int32_t) * rows.size(),
                     hipMemcpyDeviceToHost));
  int64_t q = 0;
  for (int i = 0; i < n; i++) {
    if (rowptr) rowptr[i] = q;
    for (int e = 0; e < hn[i]; e++, q++)
      if (idx && q < cap) idx[q] = rows[(size_t)i * mx + e] - p.poff;
  }
  if (rowptr) rowptr[n] = q;
  return q > cap && idx ? DML_E_ARG : DML_OK;
}

int dml_forces(dml_ctx *c) {
  if (!c) return DML_E_ARG;
  enq_forces(c, 0, 0);
  return finish(c);
}

int dml_brownian(dml_ctx *c, const double *noise, const double *u_dep) {
  if (!c) return DML_E_ARG;
  double *dn = upload_tmp(c, noise, 3 * c->P), *du = upload_tmp(c, u_dep, c->P);
  hipLaunchKernelGGL(k_brownian, dim3(c->d.nblocks), dim3(DML_BLOCK), 0, c->st, c->d,
                     (const double *)dn, (const double *)du);
  int rc = finish(c);
  hipFree(dn);
  hipFree(du);
  return rc;
}

int dml_ermak_a(dml_ctx *c, const double *noise, const double *u_dep) {
  if (!c) return DML_E_ARG;
  double *dn = upload_tmp(c, noise, 6 * c->P), *du = upload_tmp(c, u_dep, c->P);
  hipLaunchKernelGGL(k_ermak_a, dim3(c->d.nblocks), dim3(DML_BLOCK), 0, c->st, c->d,
                     (const double *)dn, (const double *)du);
  int rc = finish(c);
  hipFree(dn);
  hipFree(du);
  return rc;
}

int dml_ermak_b(dml_ctx *c) {
  if (!c) return DML_E_ARG;
  hipLaunchKernelGGL(k_ermak_b, dim3(c->d.nblocks), dim3(DML_BLOCK), 0, c->st, c->d, 0);
  return finish(c);
}

int dml_overlap_moveback(dml_ctx *c, const double *u_contact, int64_t n_u, int32_t max_passes) {
  if (!c) return DML_E_ARG;
  // n_u uniforms per replica, replica r at [r*n_u, (r+1)*n_u)
  if (u_contact && n_u <= 0) return DML_E_ARG;
  double *du = u_contact ? upload_tmp(c, u_contact, n_u * c->nrep) : nullptr;
  hipLaunchKernelGGL(k_ov_exact, dim3(c->nrep), dim3(64), 0, c->st, c->d, (const double *)du, n_u,
                     max_passes > 0 ? max_passes : 100000);
  int rc = finish(c);
  hipFree(du);
  return rc;
}

int dml_overlap_rounds(dml_ctx *c, const double *u_round, int32_t max_rounds) {
  if (!c || max_rounds <= 0 || max_rounds > kMaxRounds) return DML_E_ARG;
  double *du = u_round ? upload_tmp(c, u_round, (int64_t)max_rounds * c->P) : nullptr;
  enq_overlap_rounds(c, du, c->P, max_rounds, 1);
  int rc = finish(c);
  hipFree(du);
  return rc;
}

int dml_f_to_cg(dml_ctx *c) {
  if (!c) return DML_E_ARG;
  enq_f2cg(c);
  return finish(c);
}

int dml_gcmc(dml_ctx *c, int32_t rep, double act, int32_t nadj, const double *u, const double *g) {
  if (!c || rep < 0 || rep >= c->nrep || nadj < 0 || nadj > (1 << 21) || (!u) != (!g))
    return DML_E_ARG;
  if (int rc0 = ensure_cells(c)) return rc0;
  int rc = fetch_params(c);
  if (rc) return rc;
  const RepParams &p = c->reps[rep].p;
  if (p.reservoir != DML_GCMC) return DML_E_STATE;
  if (!p.listed) return DML_E_STATE;  // cells and rows come from test_update
  if (nadj == 0) return DML_OK;
  double *du = upload_tmp(c, u, 5 * (int64_t)nadj);
  double *dg = upload_tmp(c, g, 3 * (int64_t)nadj);
  if ((u && !du) || (g && !dg)) {
    hipFree(du);
    hipFree(dg);
    return DML_E_NOMEM;
  }
  enq_gcmc(c, rep, du, dg, act, nadj);
  rc = finish(c);
  hipFree(du);
  hipFree(dg);
  return rc;
}

int dml_calc_rho(dml_ctx *c, double *rho) {
  if (!c) return DML_E_ARG;
  int nb = c->d.nblocks;
  int rb = (c->nrep + 63) / 64;
  hipLaunchKernelGGL(k_rep_op, dim3(rb), dim3(64), 0, c->st, c->d, 0);
  hipLaunchKernelGGL(k_calc_rho, dim3(nb), dim3(DML_BLOCK), 0, c->st, c->d);
  int rc = finish(c);
  if (rc || !rho) return rc;
  rc = fetch_params(c);
  if (rc) return rc;
  for (int r = 0; r < c->nrep; r++) {
    const RepParams &p = c->reps[r].p;
    const double top = p.reservoir == DML_CHUNKS ? p.z1 : p.zmax;
    rho[r] = (double)p.rho_count / (p.box[0] * p.box[1] * (top - p.z0));
  }
  return DML_OK;
}

int dml_observables(dml_ctx *c, double *out) {
  if (!c || !out) return DML_E_ARG;
  hipLaunchKernelGGL(k_observe_part, dim3(c->d.nblocks), dim3(DML_BLOCK), 0, c->st, c->d);
  hipLaunchKernelGGL(k_observe_final, dim3(c->nrep), dim3(DML_BLOCK), 0, c->st, c->d);
  std::vector<double> o(4 * (size_t)c->nrep);
  if (hipMemcpyAsync(o.data(), c->d.obs + 4 * (int64_t)c->d.nblocks, o.size() * sizeof(double),
                     hipMemcpyDeviceToHost, c->st) != hipSuccess)
    return DML_E_HIP;
  int rc = finish(c);
  if (rc) return rc;
  // kB_ui of dana.F90:22-24, the constant kion divides by
  const double kB_eVK = 8.617330350e-5, kJm_ui = 100.0, eV_kJm = 96.485;
  const double kB_ui = kB_eVK * (eV_kJm * kJm_ui);
  for (int r = 0; r < c->nrep; r++) {
    const double *q = &o[4 * (size_t)r];
    out[4 * r + 0] = q[0];
    out[4 * r + 1] = q[1] / (q[2] * 3.0 * kB_ui);  // kion: vdac/(j*3._dp*kB_ui)
    out[4 * r + 2] = q[2];
    out[4 * r + 3] = q[3];
  }
  return DML_OK;
}

int dml_set_chunk(dml_ctx *c, int32_t rep, int32_t n, const double *pos, double dist, double z1) {
  if (!c || rep < 0 || rep >= c->nrep || n < 0 || (n > 0 && !pos) || !(dist > 0)) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  HostRep &h = c->reps[rep];
  RepParams &p = h.p;
  if (p.reservoir != DML_CHUNKS) return DML_E_STATE;
  // config_chunk (dana.F90:552-587): the template at zmax + z, its pos_old
  // another zmax higher
  h.chunk.assign(6 * (size_t)n, 0.0);
  for (int j = 0; j < n; j++) {
    for (int k = 0; k < 3; k++) {
      h.chunk[(size_t)k * n + j] = pos[(int64_t)k * n + j];
      h.chunk[(size_t)(3 + k) * n + j] = pos[(int64_t)k * n + j];
    }
    h.chunk[(size_t)2 * n + j] = h.chunk[(size_t)2 * n + j] + p.zmax;
    h.chunk[(size_t)5 * n + j] = h.chunk[(size_t)2 * n + j] + p.zmax;
  }
  p.nchunk = n;
  p.dist = dist;
  p.z1 = z1;
  // cell memory for the tallest box the slots allow: each block adds dist
  if (n > 0) {
    int64_t blocks = std::max<int64_t>(0, (p.pcap - p.nslots) / n);
    RepParams t = p;
    t.box[2] = p.zmax + (double)blocks * dist;
    t.use_cells = 0;
    if (tessellate_rep(t, 1) == 0) {
      int64_t cneed = ncells_total(t);
      int64_t bneed = cneed * std::max<int64_t>(p.bcap, bucket_cap(t));
      if (cneed > h.creg || bneed > h.breg) {
        // grown at the next call that sorts cells (the buckets are refilled
        // at each replica's next sort; rows stay)
        h.creg = std::max(h.creg, cneed);
        h.breg = std::max(h.breg, bneed);
        c->cells_dirty = 1;
      }
    }
  }
  // the device pool: templates appended, the pool doubled when full
  double **arr[6] = {&c->d.chx, &c->d.chy, &c->d.chz, &c->d.chox, &c->d.choy, &c->d.choz};
  if (c->chunk_used + n > c->chunk_cap) {
    int64_t cap = std::max<int64_t>(2 * c->chunk_cap, c->chunk_used + n);
    for (auto a : arr) {
      double *q = nullptr;
      if (dalloc(&q, cap)) return DML_E_NOMEM;
      if (*a && c->chunk_used)
        HIPCHECK(hipMemcpy(q, *a, sizeof(double) * c->chunk_used, hipMemcpyDeviceToDevice));
      hipFree(*a);
      *a = q;
    }
    c->chunk_cap = cap;
  }
  p.chunk_off = c->chunk_used;
  c->chunk_used += n;
  for (int k = 0; k < 6 && n > 0; k++)
    HIPCHECK(hipMemcpy(*arr[k] + p.chunk_off, h.chunk.data() + (size_t)k * n, sizeof(double) * n,
                       hipMemcpyHostToDevice));
  refresh_modes(c);
  return upload_params(c);
}

// reallocate the cell regions that dml_set_chunk enlarged
static int ensure_cells(dml_ctx *c) {
  if (c->cells_dirty) {
    int rc = fetch_params(c);
    if (rc) return rc;
    if ((rc = alloc_cells(c))) return rc;
    c->cells_dirty = 0;
    if ((rc = upload_params(c))) return rc;
  }
  return build_maps(c);
}

int dml_bloques(dml_ctx *c) {
  if (!c) return DML_E_ARG;
  if (int rc0 = ensure_cells(c)) return rc0;
  enq_bloques(c);
  return finish(c);
}

int dml_maxz(dml_ctx *c) {
  if (!c) return DML_E_ARG;
  hipLaunchKernelGGL(k_maxz, dim3(c->d.nblocks), dim3(DML_BLOCK), 0, c->st, c->d);
  hipLaunchKernelGGL(k_rep_op, dim3((c->nrep + 63) / 64), dim3(64), 0, c->st, c->d, 1);
  return finish(c);
}

int dml_end_step(dml_ctx *c) {
  if (!c) return DML_E_ARG;
  hipLaunchKernelGGL(k_rep_op, dim3((c->nrep + 63) / 64), dim3(64), 0, c->st, c->d, 2);
  return finish(c);
}

int dml_get_gcmc_order(dml_ctx *c, int32_t rep, int32_t *ngc, int32_t *order) {
  if (!c || rep < 0 || rep >= c->nrep || !ngc) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  const RepParams &p = c->reps[rep].p;
  // atoms turned CG since the last gcmc_run are still listed on the device
  // (dropped at its start); report the group as gcmc%detach leaves it
  std::vector<int32_t> g(std::max(p.ngc, 1));
  std::vector<uint8_t> fl(std::max(p.pcap, 1));
  if (p.ngc > 0) {
    HIPCHECK(hipMemcpy(g.data(), c->d.gc_order + p.poff, sizeof(int32_t) * p.ngc,
                       hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(fl.data(), c->d.flags + p.poff, p.pcap, hipMemcpyDeviceToHost));
  }
  int m = 0;
  for (int t = 0; t < p.ngc; t++) {
    int i = g[t] - p.poff;
    if (i < 0 || i >= p.pcap || !(fl[i] & DML_FL_GCMC)) continue;
    if (order) order[m] = i;
    m++;
  }
  *ngc = m;
  return DML_OK;
}

int dml_gcmc_outcomes(dml_ctx *c, int32_t rep, int32_t ntrial, int32_t *kind, int32_t *accepted,
                      int32_t *slot, double *du, int32_t *batched) {
  if (!c || rep < 0 || rep >= c->nrep || ntrial < 0 || !batched) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  const RepParams &p = c->reps[rep].p;
  *batched = p.gb_state == GB_ON && p.gb_cap > 0;
  if (!*batched) return DML_OK;
  if (ntrial > p.gb_t) return DML_E_ARG;
  if (ntrial == 0) return DML_OK;
  std::vector<int32_t> k(ntrial), rs(ntrial), sl(ntrial);
  HIPCHECK(hipMemcpy(k.data(), c->d.gb_kind + p.gb_off, sizeof(int32_t) * ntrial,
                     hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(rs.data(), c->d.gb_res + p.gb_off, sizeof(int32_t) * ntrial,
                     hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(sl.data(), c->d.gb_slot + p.gb_off, sizeof(int32_t) * ntrial,
                     hipMemcpyDeviceToHost));
  if (du)
    HIPCHECK(hipMemcpy(du, c->d.gb_du + p.gb_off, sizeof(double) * ntrial, hipMemcpyDeviceToHost));
  for (int t = 0; t < ntrial; t++) {
    if (kind) kind[t] = k[t];
    if (accepted) accepted[t] = rs[t] >= 0;
    if (slot) slot[t] = rs[t] >= 0 ? sl[t] - p.poff : -1;
  }
  return DML_OK;
}

int dml_get_slots(dml_ctx *c, int32_t rep, int32_t *nslots, int32_t *bkey, uint32_t *pid,
                  int32_t *seq) {
  if (!c || rep < 0 || rep >= c->nrep || !nslots) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  const RepParams &p = c->reps[rep].p;
  *nslots = p.nslots;
  if (bkey && p.nslots > 0)
    HIPCHECK(hipMemcpy(bkey, c->d.bkey + p.poff, sizeof(int32_t) * p.nslots, hipMemcpyDeviceToHost));
  if (pid && p.nslots > 0)
    HIPCHECK(hipMemcpy(pid, c->d.pid + p.poff, sizeof(uint32_t) * p.nslots, hipMemcpyDeviceToHost));
  if (seq && p.nslots > 0)
    HIPCHECK(hipMemcpy(seq, c->d.seq + p.poff, sizeof(int32_t) * p.nslots, hipMemcpyDeviceToHost));
  return DML_OK;
}

int dml_gcmc_trial_energy(dml_ctx *c, int32_t rep, int32_t ntrial, const double *pos3,
                          const int32_t *exclude, double *du, int32_t *hard) {
  if (!c || rep < 0 || rep >= c->nrep || ntrial < 0 || (!pos3 && !exclude) || !du || !hard)
    return DML_E_ARG;
  if (ntrial == 0) return DML_OK;
  if (int rc0 = ensure_cells(c)) return rc0;
  int rc = fetch_params(c);
  if (rc) return rc;
  const RepParams &p = c->reps[rep].p;
  for (int t = 0; t < ntrial; t++) {
    if (pos3) {
      const double *q = pos3 + 3 * (int64_t)t;
      if (!(q[0] >= 0 && q[0] < p.box[0] && q[1] >= 0 && q[1] < p.box[1] && q[2] >= 0 &&
            q[2] <= p.box[2]))
        return DML_E_ARG;
    }
    if (exclude && (exclude[t] < (pos3 ? -1 : 0) || exclude[t] >= p.nslots)) return DML_E_ARG;
  }
  double *dp = nullptr, *de = nullptr;
  int32_t *dx = nullptr, *dh = nullptr;
  rc = DML_E_NOMEM;
  if ((!pos3 || !dalloc(&dp, 3 * (int64_t)ntrial)) && !dalloc(&de, ntrial) &&
      !dalloc(&dh, ntrial) && (!exclude || !dalloc(&dx, ntrial))) {
    rc = DML_OK;
    if ((pos3 && hipMemcpyAsync(dp, pos3, sizeof(double) * 3 * (size_t)ntrial,
                                hipMemcpyHostToDevice, c->st) != hipSuccess) ||
        (exclude && hipMemcpyAsync(dx, exclude, sizeof(int32_t) * (size_t)ntrial,
                                   hipMemcpyHostToDevice, c->st) != hipSuccess))
      rc = DML_E_HIP;
  }
  if (!rc) {
    int64_t threads = (int64_t)ntrial * 64;
    hipLaunchKernelGGL(k_gcmc_energy, dim3((unsigned)((threads + DML_BLOCK - 1) / DML_BLOCK)),
                       dim3(DML_BLOCK), 0, c->st, c->d, rep, ntrial, (const double *)dp,
                       (const int32_t *)dx, de, dh);
    rc = finish(c);
  }
  if (!rc && (hipMemcpy(du, de, sizeof(double) * ntrial, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(hard, dh, sizeof(int32_t) * ntrial, hipMemcpyDeviceToHost) != hipSuccess))
    rc = DML_E_HIP;
  hipFree(dp);
  hipFree(de);
  hipFree(dh);
  hipFree(dx);
  return rc;
}

int dml_gcmc_trial_energy_noalloc(dml_ctx *c, int32_t rep, int32_t ntrial, const double *pos3,
                                  const int32_t *exclude, double *du, int32_t *hard) {
  if (!c || rep < 0 || rep >= c->nrep || ntrial < 0 || (!pos3 && !exclude) || !du || !hard)
    return DML_E_ARG;
  if (ntrial > DML_NOALLOC_MAX_TRIAL) {
    fprintf(stderr, "libdml: dml_gcmc_trial_energy_noalloc: %d trials, at most %d\n", ntrial,
            DML_NOALLOC_MAX_TRIAL);
    return DML_E_NOMEM;
  }
  if (ntrial == 0) return DML_OK;
  // the cell regions must be in place: growing them allocates
  if (c->cells_dirty) return DML_E_STATE;
  RepParams p;
  HIPCHECK(hipMemcpyAsync(&p, c->d.rp + rep, sizeof(RepParams), hipMemcpyDeviceToHost, c->st));
  HIPCHECK(hipStreamSynchronize(c->st));
  for (int t = 0; t < ntrial; t++) {
    if (pos3) {
      const double *q = pos3 + 3 * (int64_t)t;
      if (!(q[0] >= 0 && q[0] < p.box[0] && q[1] >= 0 && q[1] < p.box[1] && q[2] >= 0 &&
            q[2] <= p.box[2]))
        return DML_E_ARG;
    }
    if (exclude && (exclude[t] < (pos3 ? -1 : 0) || exclude[t] >= p.nslots)) return DML_E_ARG;
  }
  // the array's layout: points | du | hard | exclude, each sized for the maximum
  double *dp = (double *)c->scratch;
  double *de = dp + 3 * (size_t)DML_NOALLOC_MAX_TRIAL;
  int32_t *dh = (int32_t *)(de + DML_NOALLOC_MAX_TRIAL);
  int32_t *dx = dh + DML_NOALLOC_MAX_TRIAL;
  if (pos3)
    HIPCHECK(hipMemcpyAsync(dp, pos3, sizeof(double) * 3 * (size_t)ntrial, hipMemcpyHostToDevice,
                            c->st));
  if (exclude)
    HIPCHECK(hipMemcpyAsync(dx, exclude, sizeof(int32_t) * (size_t)ntrial,
                            hipMemcpyHostToDevice, c->st));
  // every trial's du and hard are written by the kernel: nothing to clear
  const int64_t threads = (int64_t)ntrial * 64;
  hipLaunchKernelGGL(k_gcmc_energy, dim3((unsigned)((threads + DML_BLOCK - 1) / DML_BLOCK)),
                     dim3(DML_BLOCK), 0, c->st, c->d, rep, ntrial,
                     (const double *)(pos3 ? dp : nullptr),
                     (const int32_t *)(exclude ? dx : nullptr), de, dh);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(du, de, sizeof(double) * (size_t)ntrial, hipMemcpyDeviceToHost, c->st));
  HIPCHECK(hipMemcpyAsync(hard, dh, sizeof(int32_t) * (size_t)ntrial, hipMemcpyDeviceToHost,
                          c->st));
  HIPCHECK(hipStreamSynchronize(c->st));
  return DML_OK;
}

int dml_run(dml_ctx *c, int32_t nsteps, int32_t use_graph) {
  if (!c || nsteps < 0) return DML_E_ARG;
  if (int rc0 = ensure_cells(c)) return rc0;
  int rc = fetch_params(c);
  if (rc) return rc;
  refresh_modes(c);
  int need_init = 0;
  for (auto &h : c->reps) need_init |= !h.p.listed && !h.p.halted;
  if (need_init) {
    // dana.F90:140-142: test_update, then fuerza for Ermak runs, of the
    // replicas that have no lists yet; the others sit this phase out
    const int rb = (c->nrep + 63) / 64;
    hipLaunchKernelGGL(k_mask_listed, dim3(rb), dim3(64), 0, c->st, c->d, 1);
    enq_test_update(c, 1);
    if (c->any_ermak) enq_forces(c, 0, 1);
    hipLaunchKernelGGL(k_mask_listed, dim3(rb), dim3(64), 0, c->st, c->d, 0);
    if ((rc = finish(c, 0))) return rc;
  }
  // what the captured step depends on: the phases present and the launch
  // shapes (rounds <= 32, lanes <= 64)
  const int64_t modes = (int64_t)c->any_ermak | c->any_brown << 1 | c->any_rounds << 2 |
                        c->any_exact << 3 | c->any_piston << 4 | c->any_gcmc << 5 |
                        (int64_t)c->any_chunks << 31 |
                        (int64_t)c->rounds << 6 | (int64_t)c->lpa << 12 |
                        (int64_t)c->lpa_ov << 19 | (int64_t)c->lb_mode << 26 |
                        (int64_t)c->any_lb[0] << 32 | (int64_t)c->any_lb[1] << 33 |
                        (int64_t)c->any_lb[2] << 34 | (int64_t)c->any_lb[3] << 35;
  // a captured step holds the device state by value: any reallocation or
  // resize (rows, cells, work lists, chunk pool) needs a new capture
  uint64_t dsig = 1469598103934665603ull;
  {
    const unsigned char *b = (const unsigned char *)&c->d;
    for (size_t q = 0; q < sizeof(DevState); q++) dsig = (dsig ^ b[q]) * 1099511628211ull;
  }
  if (use_graph && (!c->gexec || c->graph_modes != modes || c->graph_dsig != dsig)) {
    if (c->gexec) {
      hipGraphExecDestroy(c->gexec);
      hipGraphDestroy(c->graph);
      c->gexec = nullptr;
    }
    HIPCHECK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    enq_step(c);
    HIPCHECK(hipStreamEndCapture(c->st, &c->graph));
    HIPCHECK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
    c->graph_rounds = c->rounds;
    c->graph_modes = modes;
    c->graph_dsig = dsig;
  }
  HIPCHECK(hipEventRecord(c->e0, c->st));
  for (int q = 0; q < nsteps; q++) {
    if (use_graph)
      HIPCHECK(hipGraphLaunch(c->gexec, c->st));
    else
      enq_step(c);
  }
  HIPCHECK(hipEventRecord(c->e1, c->st));
  HIPCHECK(hipEventSynchronize(c->e1));
  float ms = 0.f;
  hipEventElapsedTime(&ms, c->e0, c->e1);
  c->last_ms = ms;
  return finish(c, 0, 0);
}

int dml_replica_status(dml_ctx *c, int32_t rep, int32_t *status) {
  if (!c || rep < 0 || rep >= c->nrep || !status) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  *status = status_from_err(c->reps[rep].p.err_bits);
  return DML_OK;
}

int dml_time_kernel(dml_ctx *c, int32_t which, int32_t reps, double *avg_ms) {
  if (!c || reps <= 0 || !avg_ms) return DML_E_ARG;
  if (int rc0 = ensure_cells(c)) return rc0;
  if (which == DML_K_LISTBUILD || which == DML_K_CELLS) {
    // a rebuild at the current positions, then the phase alone, reps times
    hipLaunchKernelGGL(k_force_rebuild, dim3((c->nrep + 63) / 64), dim3(64), 0, c->st, c->d);
    enq_cells_lists(c, 1);
  }
  HIPCHECK(hipEventRecord(c->e0, c->st));
  for (int q = 0; q < reps; q++) {
    switch (which) {
      case DML_K_LISTBUILD:
        enq_listbuild(c);
        break;
      case DML_K_FORCE:
        launch_force_lpa<0>(c, c->lpa, 0);
        break;
      case DML_K_CELLS: {
        dim3 gc(clamp_grid((c->d.n_cm + DML_BLOCK - 1) / DML_BLOCK));
        hipLaunchKernelGGL(k_zero_cells, gc, dim3(DML_BLOCK), 0, c->st, c->d, 1);
        hipLaunchKernelGGL(k_cell_key, dim3(c->d.nblocks), dim3(DML_BLOCK), 0, c->st, c->d, 1);
        dim3 gw(clamp_grid((c->d.n_cm + 3) / 4));
        hipLaunchKernelGGL(k_cell_sort_w, gw, dim3(DML_BLOCK), 0, c->st, c->d, 1);
        break;
      }
      case DML_K_DISPMAX:
        hipLaunchKernelGGL(k_pbc_dispmax, dim3(c->d.nblocks), dim3(DML_BLOCK), 0, c->st, c->d);
        break;
      default:
        return DML_E_ARG;
    }
  }
  HIPCHECK(hipEventRecord(c->e1, c->st));
  HIPCHECK(hipEventSynchronize(c->e1));
  float ms = 0.f;
  HIPCHECK(hipEventElapsedTime(&ms, c->e0, c->e1));
  *avg_ms = ms / reps;
  return finish(c, 0);
}

int dml_set_tuning(dml_ctx *c, int32_t lanes_per_atom, int32_t overlap_rounds,
                   int32_t overlap_lanes, int32_t list_kernel) {
  if (!c) return DML_E_ARG;
  auto pow2_le64 = [](int v) { return v > 0 && v <= 64 && !(v & (v - 1)); };
  // the overlap scan is instantiated for 1, 2, 4, 8 and 16 lanes
  auto ov_ok = [](int v) { return v == 1 || v == 2 || v == 4 || v == 8 || v == 16; };
  if ((lanes_per_atom && !pow2_le64(lanes_per_atom)) || (overlap_lanes && !ov_ok(overlap_lanes)) ||
      overlap_rounds < 0 || overlap_rounds > kMaxRounds || list_kernel < 0 || list_kernel > 6)
    return DML_E_ARG;
  if (lanes_per_atom) c->lpa = lanes_per_atom;
  if (overlap_rounds) c->rounds = overlap_rounds;
  if (overlap_lanes) c->lpa_ov = overlap_lanes;
  if (list_kernel) {
    c->lb_mode = list_kernel - 1;
    int rc = fetch_params(c);
    if (rc) return rc;
    return upload_params(c);
  }
  return DML_OK;
}

int dml_get_tuning(dml_ctx *c, int32_t *lanes_per_atom, int32_t *overlap_rounds,
                   int32_t *overlap_lanes, int32_t *list_kernel) {
  if (!c) return DML_E_ARG;
  if (lanes_per_atom) *lanes_per_atom = c->lpa;
  if (overlap_rounds) *overlap_rounds = c->rounds;
  if (overlap_lanes) *overlap_lanes = c->lpa_ov;
  if (list_kernel) *list_kernel = c->lb_mode + 1;
  return DML_OK;
}

int dml_progress(dml_ctx *c, int64_t *steps, int32_t *status, int32_t *running) {
  if (!c) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  std::vector<DevCounters> h(c->nrep);
  if (steps) {
    HIPCHECK(hipMemcpy(h.data(), c->d.cnt, sizeof(DevCounters) * c->nrep, hipMemcpyDeviceToHost));
  }
  for (int r = 0; r < c->nrep; r++) {
    const RepParams &p = c->reps[r].p;
    if (steps) steps[r] = (int64_t)p.step;
    if (status) status[r] = status_from_err(p.err_bits);
    if (running) running[r] = p.halted == 0;
  }
  return DML_OK;
}

int dml_set_active(dml_ctx *c, int32_t rep, int32_t active) {
  if (!c || rep < 0 || rep >= c->nrep) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  RepParams &p = c->reps[rep].p;
  if (active) p.halted &= ~HALT_PAUSED;
  else p.halted |= HALT_PAUSED;
  return upload_params(c);
}

int dml_set_step_limit(dml_ctx *c, int32_t rep, int64_t limit) {
  if (!c || rep < 0 || rep >= c->nrep || limit < 0 || limit > INT32_MAX) return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  RepParams &p = c->reps[rep].p;
  p.step_limit = (int32_t)limit;
  if (limit > 0 && p.step >= limit) p.halted |= HALT_PAUSED;
  return upload_params(c);
}

int dml_reconfigure(dml_ctx *c, int32_t rep, const dml_config *cfg) {
  if (!c || rep < 0 || rep >= c->nrep || !cfg) return DML_E_ARG;
  const dml_config &g = *cfg;
  if (g.capacity <= 0 || g.box[0] <= 0 || g.box[1] <= 0 || g.box[2] <= 0 || g.rcut <= 0 ||
      g.nadj < 0 || g.nadj > (1 << 21))
    return DML_E_ARG;
  int rc = fetch_params(c);
  if (rc) return rc;
  HostRep &h = c->reps[rep];
  if (round_up(g.capacity, DML_BLOCK) > h.p.pcap) return DML_E_NOMEM;
  RepParams p;
  memset(&p, 0, sizeof(p));
  p.pcap = h.p.pcap;
  if (setup_rep(p, g)) return DML_E_UNSUPPORTED;
  // keep the slot range, the reserved cell regions and the trial pool
  p.gb_off = h.p.gb_off;
  p.gb_cap = h.p.gb_cap;
  p.poff = h.p.poff;
  p.coff = h.p.coff;
  p.boff = h.p.boff;
  p.creg = h.creg;
  p.breg = h.breg;
  p.ccap = ncells_total(p);
  if (p.ccap > h.creg) return DML_E_NOMEM;
  int64_t fit = h.breg / p.ccap;
  if (fit < 8) return DML_E_NOMEM;
  // the whole reserved bucket region serves the new cell count
  p.bcap = (int32_t)std::min<int64_t>(fit, std::max<int64_t>(p.bcap, std::min<int64_t>(fit, 4096)));
  p.nslots = 0;
  p.nref = 0;
  p.halted = HALT_PAUSED;  // until dml_set_state
  h.cfg = g;
  h.p = p;
  if (rows_for(g) > c->d.maxnn) {
    // wider rows than the context has: grow them (keeps every row)
    int nm = next_pow2(rows_for(g));
    int32_t *nbr2 = nullptr;
    if (dalloc(&nbr2, c->P * (int64_t)nm)) return DML_E_NOMEM;
    HIPCHECK(hipMemcpy2DAsync(nbr2, sizeof(int32_t) * nm, c->d.nbr, sizeof(int32_t) * c->d.maxnn,
                              sizeof(int32_t) * c->d.maxnn, (size_t)c->P, hipMemcpyDeviceToDevice,
                              c->st));
    HIPCHECK(hipStreamSynchronize(c->st));
    hipFree(c->d.nbr);
    c->d.nbr = nbr2;
    c->d.maxnn = nm;
  }
  hipMemsetAsync(c->d.cell_cnt + p.coff, 0, sizeof(int32_t) * p.ccap, c->st);
  hipMemsetAsync(c->d.cnt + rep, 0, sizeof(DevCounters), c->st);
  hipMemsetAsync(c->d.ov_near + p.poff, 0xff, (size_t)p.pcap, c->st);
  refresh_modes(c);
  return upload_params(c);
}

int dml_last_run_ms(dml_ctx *c, double *ms) {
  if (!c || !ms) return DML_E_ARG;
  *ms = c->last_ms;
  return DML_OK;
}

int dml_get_counters(dml_ctx *c, int32_t rep, dml_counters *out) {
  if (!c || rep < 0 || rep >= c->nrep || !out) return DML_E_ARG;
  DevCounters h;
  HIPCHECK(hipMemcpy(&h, c->d.cnt + rep, sizeof(h), hipMemcpyDeviceToHost));
  out->try_ = (int64_t)h.try_;
  out->depo = (int64_t)h.depo;
  out->choques = (int64_t)h.choques;
  out->choques3 = (int64_t)h.choques3;
  out->nupd = (int64_t)h.nupd;
  out->ins = (int64_t)h.ins;
  out->del = (int64_t)h.del;
  out->passes = (int64_t)h.passes;
  out->particle_steps = (int64_t)h.particle_steps;
  out->row_entries = (int64_t)h.nn_sum;
  out->rows = (int64_t)h.nn_rows;
  out->ov_unfinished = (int64_t)h.ov_unfinished;
  return DML_OK;
}

int dml_list_stats(dml_ctx *c, int32_t rep, double *mean_nn, int64_t *rebuilds, int64_t *steps) {
  if (!c || rep < 0 || rep >= c->nrep) return DML_E_ARG;
  DevCounters h;
  HIPCHECK(hipMemcpy(&h, c->d.cnt + rep, sizeof(h), hipMemcpyDeviceToHost));
  if (mean_nn) *mean_nn = h.nn_rows ? (double)h.nn_sum / (double)h.nn_rows : 0.0;
  if (rebuilds) *rebuilds = (int64_t)h.nupd;
  if (steps) *steps = (int64_t)h.steps;
  return DML_OK;
}

}  // extern "C"
