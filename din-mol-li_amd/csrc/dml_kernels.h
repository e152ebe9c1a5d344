// dml_kernels.h -- kernels of libdml (defined in dml_kernels.hip)
#pragma once
#include "dml_internal.h"

__global__ void k_pbc_dispmax(DevState s);
__global__ void k_dispmax_final(DevState s);
__global__ void k_init_pf4(DevState s, int64_t n);
__global__ void k_force_rebuild(DevState s);
__global__ void k_zero_cells(DevState s, int gcs);
__global__ void k_cell_key(DevState s, int gcs);
__global__ void k_cell_sort(DevState s, int gcs);
__global__ void k_cell_sort_w(DevState s, int gcs);
__global__ void k_listbuild_cells(DevState s);
template <int LPA>
__global__ void k_ov_scan_l(DevState s, const double *u_round, int64_t u_stride, int round,
                            int64_t ptotal);
template <int WPB>
__global__ void k_listbuild(DevState s);
template <int LPA, int FUSEB>
__global__ void k_force(DevState s, int only_ermak);
__global__ void k_ghost_scatter(DevState s, int only_ermak);
__global__ void k_ermak_b(DevState s, int only_if_ghost);
__global__ void k_brownian(DevState s, const double *noise, const double *udep);
__global__ void k_ermak_a(DevState s, const double *noise, const double *udep);
__global__ void k_ov_exact(DevState s, const double *u, int64_t n_u, int max_passes);
__global__ void k_ov_begin(DevState s);
__global__ void k_ov_apply_r(DevState s, int round);
__global__ void k_ov_apply(DevState s, int round, int64_t ptotal);
template <int LPA>
__global__ void k_ov_tail(DevState s, const double *u_round, int64_t u_stride, int r0, int r1,
                          int64_t ptotal);
__global__ void k_ov_end(DevState s, int rounds, int err_if_pending);
__global__ void k_ov_close(DevState s, int rounds);
__global__ void k_zero_i32(int32_t *a, int64_t n);
__global__ void k_f2cg_mark(DevState s);
__global__ void k_f2cg_scan(DevState s);
__global__ void k_f2cg_scatter(DevState s);
__global__ void k_f2cg_copy(DevState s);
__global__ void k_f2cg_end(DevState s);
__global__ void k_calc_rho(DevState s);
__global__ void k_observe_part(DevState s);
__global__ void k_observe_final(DevState s);
__global__ void k_maxz(DevState s);
__global__ void k_rep_op(DevState s, int op);
__global__ void k_mask_listed(DevState s, int on);
__global__ void k_chunk_run(DevState s);
__global__ void k_mask_fired(DevState s, int on);
__global__ void k_listbuild_wc(DevState s);
__global__ void k_step_end(DevState s);
__global__ void k_gc_count(DevState s, int rep_only, int stage);
__global__ void k_gc_scan(DevState s, int rep_only, int stage);
__global__ void k_gc_scatter(DevState s, int rep_only, int stage);
__global__ void k_gc_copy(DevState s, int rep_only);
__global__ void k_gcmc_run(DevState s, int rep_only, const double *u, const double *g,
                           double act_in, int nadj_in);
template <int NCH, int WPE>
__global__ void k_listbuild_wr(DevState s);
__global__ void k_gb_begin(DevState s, int rep_only, double act_in, int nadj_in);
__global__ void k_gb_count(DevState s, int r);
__global__ void k_gb_scan(DevState s, int r);
__global__ void k_gb_scatter(DevState s, int r);
__global__ void k_gb_trials(DevState s, int r, const double *u, const double *g);
__global__ void k_gb_conflicts(DevState s, int r);
__global__ void k_gb_decide(DevState s, int r);
__global__ void k_gb_lists(DevState s, int r);
__global__ void k_gb_check(DevState s, int r);
__global__ void k_gb_dels(DevState s, int r);
__global__ void k_gb_brank(DevState s, int r);
__global__ void k_gb_commit(DevState s, int r, const double *g);
__global__ void k_gb_fields(DevState s, int r);
__global__ void k_gb_ownrow(DevState s, int r);
__global__ void k_gb_append(DevState s, int r);
__global__ void k_gb_append_fix(DevState s, int r);
__global__ void k_gb_cells(DevState s, int r);
__global__ void k_gcmc_energy(DevState s, int r, int ntrial, const double *pos3,
                              const int32_t *excl, double *du,
                              int32_t *hard);
