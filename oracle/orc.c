/*
 * orc.c -- CPU restatement (test oracle) of the din-mol-Li hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library and the timed CPU
 * baseline.  Nothing in the product links this file.
 *
 * Every routine cites the reference routine it restates.  The group
 * bookkeeping of Groups.F90 (igroup slot reuse, limbo, linked-list orders)
 * is kept because it fixes list order, deletion choice and random-stream
 * consumption, which the golden frames depend on.
 */
#include "orc.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* random numbers: ran / gasdev, dana.F90:1379-1428                          */
/* ------------------------------------------------------------------------ */

double orc_nr_ran(orc_rng *r) {
  /* dana.F90:1407-1428 -- xorshift(13,-17,5) combined with Park-Miller by
   * Schrage's method; am = nearest(1.0,-1.0)/im with the single-precision
   * nearest() promoted to double. */
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  const double am = (double)nextafterf(1.0f, 0.0f) / (double)im;
  r->ncalls++;
  if (r->idum <= 0 || r->iy < 0) {
    int32_t a = r->idum < 0 ? -r->idum : r->idum;
    r->iy = (888889999 ^ a) | 1;
    r->ix = 777755555 ^ a;
    r->idum = a + 1;
  }
  r->ix ^= (int32_t)((uint32_t)r->ix << 13);
  r->ix ^= (int32_t)((uint32_t)r->ix >> 17);
  r->ix ^= (int32_t)((uint32_t)r->ix << 5);
  int32_t k = r->iy / iq;
  r->iy = ia * (r->iy - k * iq) - ir * k;
  if (r->iy < 0) r->iy += im;
  return am * (double)((im & (r->ix ^ r->iy)) | 1);
}

double orc_nr_gasdev(orc_rng *r) {
  /* dana.F90:1379-1404 -- polar method, rsq held in single precision, the
   * factor sqrt(-2 log rsq / rsq) evaluated with the single-precision log
   * and rounded back to single precision; second deviate cached. */
  if (r->gaus_stored) {
    r->gaus_stored = 0;
    return r->gsave;
  }
  double v1, v2;
  float rsq;
  for (;;) {
    v1 = 2.0 * orc_nr_ran(r) - 1.0;
    v2 = 2.0 * orc_nr_ran(r) - 1.0;
    rsq = (float)(v1 * v1 + v2 * v2);
    if ((double)rsq > 0.0 && (double)rsq < 1.0) break;
  }
  rsq = (float)sqrt(-2.0 * (double)logf(rsq) / (double)rsq);
  r->gsave = v2 * (double)rsq;
  r->gaus_stored = 1;
  return v1 * (double)rsq;
}

static double nr_unif(orc_rng *r, int site, int64_t i, int k) {
  (void)site; (void)i; (void)k;
  return orc_nr_ran(r);
}
static double nr_gauss(orc_rng *r, int site, int64_t i, int k) {
  (void)site; (void)i; (void)k;
  return orc_nr_gasdev(r);
}

void orc_rng_init_nr(orc_rng *r, int32_t idum) {
  memset(r, 0, sizeof(*r));
  r->unif = nr_unif;
  r->gauss = nr_gauss;
  r->idum = idum;
  r->ix = -1;
  r->iy = -1;
}

static double arr_draw(orc_rng *r, int site, int64_t i, int k) {
  if (site == ORC_SITE_CONTACT) i = r->contact_next++;
  const double *a = r->arr[site];
  int64_t idx = i * (int64_t)r->arr_stride[site] + k;
  if (!a || idx < 0 || idx >= r->arr_n[site]) {
    r->err = 1;
    return 0.5;
  }
  return a[idx];
}

void orc_rng_init_arrays(orc_rng *r) {
  memset(r, 0, sizeof(*r));
  r->unif = arr_draw;
  r->gauss = arr_draw;
  r->arr_stride[ORC_SITE_BROWN] = 3;
  r->arr_stride[ORC_SITE_ERMAK] = 6;
  r->arr_stride[ORC_SITE_DEPO] = 1;
  r->arr_stride[ORC_SITE_CONTACT] = 1;
  r->arr_stride[ORC_SITE_GCMC_U] = 5;
  r->arr_stride[ORC_SITE_GCMC_G] = 3;
}

/* ------------------------------------------------------------------------ */
/* system                                                                    */
/* ------------------------------------------------------------------------ */

typedef struct ivec {
  int *v;
  int n, cap;
} ivec;

static int ivec_push(ivec *a, int x) {
  if (a->n == a->cap) {
    int nc = a->cap ? 2 * a->cap : 64;
    int *nv = (int *)realloc(a->v, sizeof(int) * (size_t)nc);
    if (!nv) return -1;
    a->v = nv;
    a->cap = nc;
  }
  a->v[a->n++] = x;
  return 0;
}
static void ivec_remove_value(ivec *a, int x) {
  for (int i = 0; i < a->n; i++)
    if (a->v[i] == x) {
      memmove(a->v + i, a->v + i + 1, sizeof(int) * (size_t)(a->n - i - 1));
      a->n--;
      return;
    }
}

/* an indexed group (igroup, Groups.F90:179-219): slot -> hs slot, or -1 */
typedef struct igrp {
  int *slot;
  int amax, nat, cap;
} igrp;

struct orc_sys {
  orc_params p;
  double box[3], one_box[3];
  double rcut, rcut2, brcut;
  double cc0, cc1, cc2, sdr, sdv, crv1, crv2, skt;
  double z1; /* chunks sensor top (dana.F90:440) */
  /* hs: ngroup igroup (Neighbor.F90:32) -- owns the per-particle arrays */
  int cap, amax, nat, nlimbo, b_limbo;
  int8_t *st, *zs;
  double (*pos)[3], (*vel)[3], (*force)[3], (*acel)[3], (*pos_old)[3],
      (*old_cg)[3];
  double *epot;
  uint8_t *skip, *in_ref, *in_gcmc;
  int *b_of, *sys_of;
  /* b: cgroup (Cells.F90:39) and sys: igroup */
  igrp b, sys;
  ivec ref, gc, sysl;
  /* cells */
  int tessellated, ncells[3];
  double cell[3];
  int *head, head_n, *bnext, bnext_n;
  /* lists */
  int listed, lista_cells;
  int *nn, **lst, *lcap;
  /* ermak */
  double (*ranv)[3];
  int ranv_n;
};

static const int MAP[27][3] = {
    /* Cells.F90:28-36, stencil order 0..26 */
    {0, 0, 0},   {1, 0, 0},   {1, 1, 0},   {0, 1, 0},  {-1, 1, 0},
    {1, 0, -1},  {1, 1, -1},  {0, 1, -1},  {-1, 1, -1}, {1, 0, 1},
    {1, 1, 1},   {0, 1, 1},   {-1, 1, 1},  {0, 0, 1},  {-1, 0, 0},
    {-1, -1, 0}, {0, -1, 0},  {1, -1, 0},  {-1, 0, 1}, {-1, -1, 1},
    {0, -1, 1},  {1, -1, 1},  {-1, 0, -1}, {-1, -1, -1}, {0, -1, -1},
    {1, -1, -1}, {0, 0, -1}};

static int grow(orc_sys *s, int need) {
  if (need <= s->cap) return 0;
  int nc = s->cap ? s->cap : 256;
  while (nc < need) nc *= 2;
#define RE(ptr, T)                                                     \
  do {                                                                 \
    void *q = realloc((void *)(ptr), sizeof(T) * (size_t)nc);          \
    if (!q) return -1;                                                 \
    ptr = q;                                                           \
    memset((char *)ptr + sizeof(T) * (size_t)s->cap, 0,                \
           sizeof(T) * (size_t)(nc - s->cap));                         \
  } while (0)
  RE(s->st, int8_t);
  RE(s->zs, int8_t);
  RE(s->pos, double[3]);
  RE(s->vel, double[3]);
  RE(s->force, double[3]);
  RE(s->acel, double[3]);
  RE(s->pos_old, double[3]);
  RE(s->old_cg, double[3]);
  RE(s->epot, double);
  RE(s->skip, uint8_t);
  RE(s->in_ref, uint8_t);
  RE(s->in_gcmc, uint8_t);
  RE(s->b_of, int);
  RE(s->sys_of, int);
  RE(s->nn, int);
  RE(s->lst, int *);
  RE(s->lcap, int);
#undef RE
  s->cap = nc;
  return 0;
}

static int igrp_grow(igrp *g, int need) {
  if (need <= g->cap) return 0;
  int nc = g->cap ? 2 * g->cap : 256;
  while (nc < need) nc *= 2;
  int *q = (int *)realloc(g->slot, sizeof(int) * (size_t)nc);
  if (!q) return -1;
  for (int i = g->cap; i < nc; i++) q[i] = -1;
  g->slot = q;
  g->cap = nc;
  return 0;
}

/* igroup_attach_atom (Groups.F90:1058-1100): after the group count grows,
 * reuse the first null slot when amax >= nat + nlimbo, else append. */
static int igrp_attach(igrp *g, int hs, int nlimbo, int *slot_out) {
  g->nat++;
  int n;
  if (g->amax >= g->nat + nlimbo) {
    for (n = 0; n < g->amax; n++)
      if (g->slot[n] < 0) break;
    if (n >= g->amax) return ORC_E_INDEX;
  } else {
    if (igrp_grow(g, g->amax + 1)) return ORC_E_ALLOC;
    n = g->amax++;
  }
  g->slot[n] = hs;
  *slot_out = n;
  return 0;
}

orc_sys *orc_sys_new(const orc_params *p, int cap) {
  orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
  if (!s) return NULL;
  s->p = *p;
  orc_set_box(s, p->box);
  /* ngroup_setrc (Neighbor.F90:323-334) */
  s->rcut = p->rcut;
  s->rcut2 = p->rcut * p->rcut;
  s->brcut = p->rcut + p->skin;
  s->ncells[0] = s->ncells[1] = s->ncells[2] = 1;
  if (grow(s, cap > 0 ? cap : 256)) {
    orc_sys_free(s);
    return NULL;
  }
  return s;
}

void orc_sys_free(orc_sys *s) {
  if (!s) return;
  for (int i = 0; i < s->cap; i++) free(s->lst ? s->lst[i] : NULL);
  free(s->st); free(s->zs); free(s->pos); free(s->vel); free(s->force);
  free(s->acel); free(s->pos_old); free(s->old_cg); free(s->epot);
  free(s->skip); free(s->in_ref); free(s->in_gcmc); free(s->b_of);
  free(s->sys_of); free(s->nn); free(s->lst); free(s->lcap);
  free(s->b.slot); free(s->sys.slot); free(s->ref.v); free(s->gc.v);
  free(s->sysl.v); free(s->head); free(s->bnext); free(s->ranv);
  free(s);
}

void orc_set_box(orc_sys *s, const double box[3]) {
  /* box_setvars (Program_Types.F90:81-95) */
  for (int k = 0; k < 3; k++) {
    s->box[k] = box[k];
    s->one_box[k] = 1.0 / box[k];
    s->p.box[k] = box[k];
  }
}

void orc_set_zmax(orc_sys *s, double zmax) { s->p.zmax = zmax; }

void orc_set_ermak(orc_sys *s) {
  /* set_ermak (dana.F90:947-971) with kB_ui of dana.F90:22-24 */
  const double kB_eVK = 8.617330350e-5, kJm_ui = 100.0, eV_kJm = 96.485;
  const double eV_ui = eV_kJm * kJm_ui;
  const double kB_ui = kB_eVK * eV_ui;
  double h = s->p.h, gama = s->p.gama;
  s->cc0 = exp(-(h * gama));
  s->cc1 = (1.0 - s->cc0) / gama;
  s->cc2 = (1.0 - s->cc1 / h) / gama;
  s->sdr = sqrt(h / gama *
                (2.0 - ((3.0 - 4.0 * s->cc0) + s->cc0 * s->cc0) / (h * gama)));
  s->sdv = sqrt(1.0 - s->cc0 * s->cc0);
  s->crv1 = (1.0 - s->cc0) * (1.0 - s->cc0) / (gama * s->sdr * s->sdv);
  s->crv2 = sqrt(1.0 - (s->crv1 * s->crv1));
  s->skt = sqrt(kB_ui * s->p.temp);
}

void orc_get_ermak(const orc_sys *s, double out[8]) {
  out[0] = s->cc0; out[1] = s->cc1; out[2] = s->cc2; out[3] = s->sdr;
  out[4] = s->sdv; out[5] = s->crv1; out[6] = s->crv2; out[7] = s->skt;
}

static int hs_new_slot(orc_sys *s, int *hs) {
  /* the hs igroup slot of a newly attached atom (Groups.F90:1083-1093) */
  int nat1 = s->nat + 1, n;
  if (s->amax >= nat1 + s->nlimbo) {
    for (n = 0; n < s->amax; n++)
      if (s->st[n] == ORC_NULL) break;
    if (n >= s->amax) return ORC_E_INDEX;
  } else {
    if (grow(s, s->amax + 1)) return ORC_E_ALLOC;
    n = s->amax++;
  }
  *hs = n;
  return 0;
}

static void atom_default(orc_sys *s, int i) {
  /* atom_allocate (Groups.F90:370-387) and the declaration defaults
   * pos_old = old_cg = 1e8 (Groups.F90:300) */
  for (int k = 0; k < 3; k++) {
    s->pos[i][k] = 0.0; s->vel[i][k] = 0.0; s->force[i][k] = 0.0;
    s->acel[i][k] = 0.0; s->pos_old[i][k] = 1e8; s->old_cg[i][k] = 1e8;
  }
  s->epot[i] = 0.0;
  s->skip[i] = 0;
  s->in_ref[i] = 0;
  s->in_gcmc[i] = 0;
  s->nn[i] = 0;
}

static int b_push_front(orc_sys *s, int bslot);

int orc_attach_initial(orc_sys *s, int z, const double pos[3]) {
  /* config_inic (dana.F90:457-489): sys, [ref], b, hs */
  int hs, slot, e;
  if ((e = hs_new_slot(s, &hs))) return e;
  atom_default(s, hs);
  s->zs[hs] = (int8_t)z;
  for (int k = 0; k < 3; k++) {
    s->pos[hs][k] = pos[k];
    s->pos_old[hs][k] = pos[k];
  }
  s->st[hs] = ORC_ALIVE;
  if ((e = igrp_attach(&s->sys, hs, 0, &slot))) return e;
  s->sys_of[hs] = slot;
  ivec_push(&s->sysl, hs);
  if (z != ORC_CG) {
    ivec_push(&s->ref, hs);
    s->in_ref[hs] = 1;
  }
  if ((e = igrp_attach(&s->b, hs, 0, &slot))) return e;
  s->b_of[hs] = slot;
  if (s->bnext_n < s->b.cap) {
    s->bnext = (int *)realloc(s->bnext, sizeof(int) * (size_t)s->b.cap);
    s->bnext_n = s->b.cap;
  }
  if (s->tessellated) b_push_front(s, slot);
  s->nat++;
  return hs;
}

int orc_nslots(const orc_sys *s) { return s->amax; }
int orc_nref(const orc_sys *s) { return s->ref.n; }
int orc_get_b_slot(const orc_sys *s, int hs) { return s->b_of[hs]; }

void orc_get_ref_order(const orc_sys *s, int32_t *out) {
  for (int i = 0; i < s->ref.n; i++) out[i] = s->ref.v[i];
}

void orc_get_state(const orc_sys *s, double *pos, double *vel, double *acel,
                   double *force, double *epot, double *pos_old,
                   double *old_cg, int8_t *species, uint8_t *flags) {
  int n = s->amax;
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) {
      if (pos) pos[(size_t)k * n + i] = s->pos[i][k];
      if (vel) vel[(size_t)k * n + i] = s->vel[i][k];
      if (acel) acel[(size_t)k * n + i] = s->acel[i][k];
      if (force) force[(size_t)k * n + i] = s->force[i][k];
      if (pos_old) pos_old[(size_t)k * n + i] = s->pos_old[i][k];
      if (old_cg) old_cg[(size_t)k * n + i] = s->old_cg[i][k];
    }
    if (epot) epot[i] = s->epot[i];
    if (species) species[i] = s->zs[i];
    if (flags) {
      uint8_t f = 0;
      if (s->st[i] == ORC_ALIVE) f |= ORC_FL_ALIVE;
      if (s->st[i] == ORC_LIMBO) f |= ORC_FL_LIMBO;
      if (s->in_ref[i]) f |= ORC_FL_REF;
      if (s->skip[i]) f |= ORC_FL_SKIP;
      if (s->in_gcmc[i]) f |= ORC_FL_GCMC;
      flags[i] = f;
    }
  }
}

void orc_set_state(orc_sys *s, const double *pos, const double *vel,
                   const double *acel, const double *pos_old,
                   const double *old_cg, const int8_t *species,
                   const uint8_t *flags) {
  int n = s->amax;
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) {
      if (pos) s->pos[i][k] = pos[(size_t)k * n + i];
      if (vel) s->vel[i][k] = vel[(size_t)k * n + i];
      if (acel) s->acel[i][k] = acel[(size_t)k * n + i];
      if (pos_old) s->pos_old[i][k] = pos_old[(size_t)k * n + i];
      if (old_cg) s->old_cg[i][k] = old_cg[(size_t)k * n + i];
    }
    if (species) s->zs[i] = species[i];
    if (flags) s->skip[i] = (flags[i] & ORC_FL_SKIP) ? 1 : 0;
  }
}

/* ------------------------------------------------------------------------ */
/* geometry helpers                                                          */
/* ------------------------------------------------------------------------ */

static inline double dot3(const double v[3]) {
  /* dot_product of a 3-vector: ((v1*v1 + v2*v2) + v3*v3) */
  return (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
}

/* vdistance (Groups.F90:995-1016) with mic and pbc = (T,T,F) for every
 * atom of dana (dana.F90:483-484): vd = a - b, then
 * vd -= box * idnint(vd * one_box), idnint rounding half away from zero. */
static inline void vdist(const orc_sys *s, const double a[3], const double b[3],
                         double vd[3]) {
  for (int l = 0; l < 3; l++) vd[l] = a[l] - b[l];
  for (int l = 0; l < 2; l++)
    vd[l] = vd[l] - s->box[l] * round(vd[l] * s->one_box[l]);
}

static inline double pow6(double x) {
  double x2 = x * x, x4 = x2 * x2;
  return x2 * x4;
}
static inline double pow7(double x) {
  double x2 = x * x, x4 = x2 * x2;
  return (x * x2) * x4;
}

/* ------------------------------------------------------------------------ */
/* cells: Cells.F90:180-302, 378-404                                         */
/* ------------------------------------------------------------------------ */

static inline int hidx(const orc_sys *s, int i, int j, int k) {
  return i + (s->ncells[0] + 2) * (j + (s->ncells[1] + 2) * k);
}

static void tessellate(orc_sys *s) {
  /* cgroup_tessellate (Cells.F90:180-265) */
  if (s->tessellated) {
    int ok = 1;
    for (int k = 0; k < 3; k++)
      if (!((double)(s->ncells[k] + 1) >= s->box[k] / s->brcut)) ok = 0;
    if (ok) {
      int ok2 = 1;
      for (int k = 0; k < 3; k++)
        if (!(s->brcut < s->box[k] / (double)s->ncells[k])) ok2 = 0;
      if (ok2) {
        for (int k = 0; k < 3; k++) s->cell[k] = s->box[k] / (double)s->ncells[k];
        return;
      }
    }
  }
  for (int k = 0; k < 3; k++) s->ncells[k] = (int)(s->box[k] / s->brcut);
  if (s->ncells[0] < 4 && s->ncells[1] < 4 && s->ncells[2] < 4) return;
  for (int k = 0; k < 3; k++) s->cell[k] = s->box[k] / (double)s->ncells[k];
  int hn = (s->ncells[0] + 2) * (s->ncells[1] + 2) * (s->ncells[2] + 2);
  free(s->head);
  s->head = (int *)calloc((size_t)hn, sizeof(int));
  s->head_n = hn;
  s->tessellated = 1;
}

static int cell_of_pos(const orc_sys *s, const double p[3], int ci[3]) {
  for (int k = 0; k < 3; k++) ci[k] = (int)(p[k] / s->cell[k]) + 1;
  for (int k = 0; k < 3; k++)
    if (ci[k] < 0 || ci[k] > s->ncells[k] + 1) return ORC_E_OUT_OF_TESS;
  return 0;
}

static int b_push_front(orc_sys *s, int bslot) {
  /* cgroup_sort_atom (Cells.F90:281-302) */
  int hs = s->b.slot[bslot], ci[3];
  int e = cell_of_pos(s, s->pos[hs], ci);
  if (e) return e;
  int h = hidx(s, ci[0], ci[1], ci[2]);
  s->bnext[bslot] = s->head[h];
  s->head[h] = bslot + 1;
  return 0;
}

static int sort_cells(orc_sys *s) {
  /* cgroup_sort (Cells.F90:267-279): push-front over b slots 1..amax, so a
   * chain lists its atoms in descending b-slot order */
  memset(s->head, 0, sizeof(int) * (size_t)s->head_n);
  if (s->bnext_n < s->b.cap) {
    s->bnext = (int *)realloc(s->bnext, sizeof(int) * (size_t)s->b.cap);
    s->bnext_n = s->b.cap;
  }
  for (int j = 0; j < s->b.amax; j++) {
    if (s->b.slot[j] < 0) continue;
    int e = b_push_front(s, j);
    if (e) return e;
  }
  return 0;
}

static int b_unsort(orc_sys *s, int bslot) {
  /* cgroup_unsort_atom (Cells.F90:304-352) */
  int hs = s->b.slot[bslot], ci[3];
  for (int k = 0; k < 3; k++) ci[k] = (int)(s->pos[hs][k] / s->cell[k]) + 1;
  for (int k = 0; k < 3; k++)
    if (ci[k] < 0 || ci[k] > s->ncells[k]) {
      s->tessellated = 0;
      return 0;
    }
  int h = hidx(s, ci[0], ci[1], ci[2]);
  if (s->head[h] == bslot + 1) {
    s->head[h] = s->bnext[bslot];
    return 0;
  }
  int j = s->head[h];
  while (j > 0) {
    int k = s->bnext[j - 1];
    if (k == bslot + 1) {
      s->bnext[j - 1] = s->bnext[bslot];
      return 0;
    }
    j = k;
  }
  return ORC_E_INDEX; /* 'Unsort cgroup particle fail' */
}

static inline int cell_pbc1(int r, int n) {
  /* cell_pbc with mic (Cells.F90:387-391); Fortran mod == C % here */
  r = r - 1;
  r = (r + n) % n;
  return r + 1;
}

void orc_cells(const orc_sys *s, int32_t ncells[3], double cell[3],
               int *use_cells) {
  for (int k = 0; k < 3; k++) {
    ncells[k] = s->ncells[k];
    cell[k] = s->cell[k];
  }
  *use_cells = s->tessellated;
}

void orc_cell_of(const orc_sys *s, int32_t *ci3) {
  /* the cell each b atom sits in, read back from the head/next chains */
  for (int i = 0; i < s->amax; i++) ci3[3 * i] = ci3[3 * i + 1] = ci3[3 * i + 2] = -1;
  if (!s->tessellated) return;
  for (int k = 0; k <= s->ncells[2] + 1; k++)
    for (int j = 0; j <= s->ncells[1] + 1; j++)
      for (int i = 0; i <= s->ncells[0] + 1; i++) {
        int b = s->head[hidx(s, i, j, k)];
        while (b > 0) {
          int hs = s->b.slot[b - 1];
          ci3[3 * hs] = i; ci3[3 * hs + 1] = j; ci3[3 * hs + 2] = k;
          b = s->bnext[b - 1];
        }
      }
}

/* ------------------------------------------------------------------------ */
/* neighbour lists: Neighbor.F90                                             */
/* ------------------------------------------------------------------------ */

static int list_push(orc_sys *s, int i, int k) {
  if (s->nn[i] == s->lcap[i]) {
    int nc = s->lcap[i] ? 2 * s->lcap[i] : 64;
    int *q = (int *)realloc(s->lst[i], sizeof(int) * (size_t)nc);
    if (!q) return ORC_E_ALLOC;
    s->lst[i] = q;
    s->lcap[i] = nc;
  }
  s->lst[i][s->nn[i]++] = k;
  return 0;
}

static void clean_limbo(orc_sys *s) {
  /* igroup_clean (Groups.F90:1036-1053) */
  if (!s->b_limbo) return;
  for (int i = 0; i < s->amax; i++)
    if (s->st[i] == ORC_LIMBO) s->st[i] = ORC_NULL;
  s->nlimbo = 0;
  s->b_limbo = 0;
}

static int cells_atom(orc_sys *s, int i, double rcut) {
  /* body of ngroup_cells / ngroup_cells_atom (Neighbor.F90:503-537,
   * 568-601): stencil order, then chain order; strict '<'. */
  int rc[3];
  for (int k = 0; k < 3; k++) rc[k] = (int)(s->pos[i][k] / s->cell[k]) + 1;
  for (int nb = 0; nb < 27; nb++) {
    int nc[3];
    for (int k = 0; k < 3; k++)
      nc[k] = cell_pbc1(MAP[nb][k] + rc[k], s->ncells[k]);
    int j = s->head[hidx(s, nc[0], nc[1], nc[2])];
    while (j > 0) {
      int aj = s->b.slot[j - 1];
      j = s->bnext[j - 1];
      if (aj == i) continue;
      double vd[3];
      vdist(s, s->pos[aj], s->pos[i], vd);
      double rd = dot3(vd);
      if (rd < rcut) {
        int e = list_push(s, i, aj);
        if (e) return e;
      }
    }
  }
  return 0;
}

static int verlet_atom(orc_sys *s, int i, double rcut) {
  /* ngroup_verlet_atom (Neighbor.F90:426-463): b slots in order, '>' rejects */
  for (int j = 0; j < s->b.amax; j++) {
    int aj = s->b.slot[j];
    if (aj < 0 || aj == i) continue;
    double vd[3];
    vdist(s, s->pos[i], s->pos[aj], vd);
    double rd = dot3(vd);
    if (rd > rcut) continue;
    int e = list_push(s, i, aj);
    if (e) return e;
  }
  return 0;
}

static int build_lists(orc_sys *s) {
  /* ngroup_cells (Neighbor.F90:465-548) / ngroup_verlet (358-424) */
  double rcut = (s->rcut + s->p.skin);
  rcut = rcut * rcut;
  for (int i = 0; i < s->amax; i++) s->nn[i] = 0;
  clean_limbo(s);
  for (int r = 0; r < s->ref.n; r++) {
    int i = s->ref.v[r];
    int e = s->lista_cells ? cells_atom(s, i, rcut) : verlet_atom(s, i, rcut);
    if (e) return e;
  }
  s->listed = 1;
  return 0;
}

int orc_update(orc_sys *s, orc_counters *c) {
  /* update (Neighbor.F90:608-633) */
  if (c) c->nupd++;
  for (int r = 0; r < s->sysl.n; r++) {
    int i = s->sysl.v[r];
    for (int k = 0; k < 3; k++) s->pos_old[i][k] = s->pos[i][k];
  }
  s->lista_cells = s->tessellated; /* ngroup_setlista (336-353) */
  return build_lists(s);
}

static void do_pbc(orc_sys *s) {
  /* do_pbc (Groups.F90:1440-1467), x and y only */
  for (int r = 0; r < s->sysl.n; r++) {
    int i = s->sysl.v[r];
    for (int k = 0; k < 2; k++) {
      if (s->pos[i][k] >= s->box[k]) {
        s->pos[i][k] -= s->box[k];
        s->pos_old[i][k] -= s->box[k];
      } else if (s->pos[i][k] < 0.0) {
        s->pos[i][k] += s->box[k];
        s->pos_old[i][k] += s->box[k];
      }
    }
  }
}

int orc_test_update(orc_sys *s, int *rebuilt, orc_counters *c) {
  /* test_update (Neighbor.F90:668-713) */
  if (rebuilt) *rebuilt = 0;
  do_pbc(s);
  tessellate(s);
  if (s->tessellated) {
    int e = sort_cells(s);
    if (e) return e;
  }
  if (!s->listed) {
    if (rebuilt) *rebuilt = 1;
    return orc_update(s, c);
  }
  /* inq_dispmax (Neighbor.F90:635-666) over the atoms of hs */
  double d1 = 1e-16, d2 = 1e-16;
  for (int i = 0; i < s->amax; i++) {
    if (s->st[i] != ORC_ALIVE) continue;
    double vd[3];
    for (int k = 0; k < 3; k++) vd[k] = s->pos[i][k] - s->pos_old[i][k];
    double rd = dot3(vd);
    if (rd > d1) {
      d2 = d1;
      d1 = rd;
    } else if (rd > d2) {
      d2 = rd;
    }
  }
  if (sqrt(d1) + sqrt(d2) > s->p.skin) {
    if (rebuilt) *rebuilt = 1;
    return orc_update(s, c);
  }
  return 0;
}

int64_t orc_list_total(const orc_sys *s) {
  int64_t t = 0;
  for (int i = 0; i < s->amax; i++) t += s->nn[i];
  return t;
}

void orc_list_export(const orc_sys *s, int32_t *nn, int64_t *rowptr,
                     int32_t *idx) {
  int64_t t = 0;
  for (int i = 0; i < s->amax; i++) {
    if (nn) nn[i] = s->nn[i];
    if (rowptr) rowptr[i] = t;
    if (idx)
      for (int m = 0; m < s->nn[i]; m++) idx[t + m] = s->lst[i][m];
    t += s->nn[i];
  }
  if (rowptr) rowptr[s->amax] = t;
}

/* ------------------------------------------------------------------------ */
/* forces: fuerza, dana.F90:1055-1139                                        */
/* ------------------------------------------------------------------------ */

void orc_fuerza(orc_sys *s) {
  for (int r = 0; r < s->ref.n; r++) {
    int i = s->ref.v[r];
    s->force[i][0] = s->force[i][1] = s->force[i][2] = 0.0;
    s->epot[i] = 0.0;
  }
  for (int r = 0; r < s->ref.n; r++) {
    int i = s->ref.v[r];
    int k = s->zs[i];
    for (int jj = 0; jj < s->nn[i]; jj++) {
      int j = s->lst[i][jj];
      if (s->b_limbo && s->st[j] == ORC_LIMBO) continue;
      int m = s->zs[j];
      double vd[3];
      for (int l = 0; l < 3; l++) vd[l] = s->pos[i][l] - s->pos[j][l];
      for (int l = 0; l < 2; l++) {
        if (vd[l] > s->box[l] * .5)
          vd[l] = vd[l] - s->box[l];
        else if (vd[l] < -s->box[l] * .5)
          vd[l] = vd[l] + s->box[l];
      }
      if (m == ORC_CG && k == ORC_CG) continue;
      double dr = dot3(vd);
      double r0 = s->p.r0[(k - 1) * 3 + (m - 1)];
      double eps = s->p.eps[(k - 1) * 3 + (m - 1)];
      if (dr > r0 * r0) continue;
      dr = sqrt(dr);
      double b = pow6(r0);
      double c = eps * 12.0 * b;
      c = c / pow7(dr);
      b = b / pow6(dr);
      double aux = c * (b - 1);
      for (int l = 0; l < 3; l++) {
        s->force[i][l] = s->force[i][l] + aux * vd[l] / dr;
        s->force[j][l] = s->force[j][l] - aux * vd[l] / dr;
      }
      aux = eps * b * (b - 2);
      aux = aux + eps;
      s->epot[i] = s->epot[i] + aux * .5;
      s->epot[j] = s->epot[j] + aux * .5;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* integrators                                                               */
/* ------------------------------------------------------------------------ */

static int atom_pbc(orc_sys *s, int i, int64_t refpos, orc_rng *r,
                    orc_counters *c) {
  /* atom_pbc (dana.F90:1187-1250); returns depos */
  int depos = 0;
  for (int j = 0; j < 2; j++) {
    if (s->pos[i][j] > s->box[j]) {
      s->pos[i][j] = s->pos[i][j] - s->box[j];
      s->pos_old[i][j] = s->pos_old[i][j] - s->box[j];
    }
    if (s->pos[i][j] < 0.0) {
      s->pos[i][j] = s->pos[i][j] + s->box[j];
      s->pos_old[i][j] = s->pos_old[i][j] + s->box[j];
    }
  }
  if (s->pos[i][2] > s->p.zmax) {
    if (s->p.ermak) {
      s->pos[i][2] = s->pos[i][2] - 2.0 * (s->pos[i][2] - s->p.zmax);
      s->vel[i][2] = -s->vel[i][2];
    } else {
      for (int k = 0; k < 3; k++) s->pos[i][k] = s->old_cg[i][k];
    }
  }
  if (s->pos[i][2] <= 0.0) {
    if (c) c->try_++;
    double ne = r->unif(r, ORC_SITE_DEPO, refpos, 0);
    if (ne < s->p.prob) {
      if (c) c->depo++;
      s->zs[i] = ORC_F;
      depos = 1;
    }
    for (int k = 0; k < 3; k++) s->pos[i][k] = s->old_cg[i][k];
  }
  return depos;
}

int orc_brownian(orc_sys *s, orc_rng *r, orc_counters *c) {
  /* cbrownian_hs (dana.F90:798-846) */
  for (int ii = 0; ii < s->ref.n; ii++) {
    int i = s->ref.v[ii];
    for (int k = 0; k < 3; k++) s->old_cg[i][k] = s->pos[i][k];
    double dif = s->pos[i][2] > s->p.z_switch ? s->p.dif_sc : s->p.dif_sei;
    double fac1 = sqrt(2.0 * dif * s->p.h);
    for (int j = 0; j < 3; j++) {
      double r1 = r->gauss(r, ORC_SITE_BROWN, ii, j);
      double posold = s->pos[i][j];
      s->pos[i][j] = posold + r1 * fac1;
      s->vel[i][j] = (s->pos[i][j] - posold) / s->p.h;
    }
    if (atom_pbc(s, i, ii, r, c)) continue;
    s->skip[i] = 0;
  }
  return r->err ? ORC_E_INDEX : 0;
}

int orc_ermak_a(orc_sys *s, orc_rng *r, orc_counters *c) {
  /* ermak_a (dana.F90:974-1028) */
  if (s->ranv_n < s->ref.n) {
    free(s->ranv);
    s->ranv = (double(*)[3])calloc((size_t)s->ref.n + 64, sizeof(double[3]));
    if (!s->ranv) return ORC_E_ALLOC;
    s->ranv_n = s->ref.n + 64;
  }
  for (int ii = 0; ii < s->ref.n; ii++) {
    int i = s->ref.v[ii];
    double m = s->p.mass[s->zs[i] - 1];
    for (int k = 0; k < 3; k++) s->old_cg[i][k] = s->pos[i][k];
    for (int j = 0; j < 3; j++) {
      double r1 = r->gauss(r, ORC_SITE_ERMAK, ii, 2 * j);
      double ranr = s->skt / sqrt(m) * s->sdr * r1;
      s->pos[i][j] = s->pos[i][j] + s->cc1 * s->vel[i][j] +
                     s->cc2 * s->p.h * s->acel[i][j] + ranr;
      double r2 = r->gauss(r, ORC_SITE_ERMAK, ii, 2 * j + 1);
      s->ranv[ii][j] = s->skt / sqrt(m) * s->sdv * (s->crv1 * r1 + s->crv2 * r2);
    }
    if (atom_pbc(s, i, ii, r, c)) continue;
    s->skip[i] = 0;
  }
  return r->err ? ORC_E_INDEX : 0;
}

void orc_ermak_b(orc_sys *s) {
  /* ermak_b (dana.F90:1031-1052) */
  for (int ii = 0; ii < s->ref.n; ii++) {
    int i = s->ref.v[ii];
    if (s->zs[i] == ORC_CG) continue;
    double m = s->p.mass[s->zs[i] - 1];
    for (int l = 0; l < 3; l++) {
      s->vel[i][l] = s->cc0 * s->vel[i][l] + (s->cc1 - s->cc2) * s->acel[i][l] +
                     s->cc2 * s->force[i][l] / m + s->ranv[ii][l];
      s->acel[i][l] = s->force[i][l] / m;
    }
  }
}

void orc_get_ranv(const orc_sys *s, double *ranv) {
  int n = s->ref.n;
  for (int ii = 0; ii < n; ii++)
    for (int l = 0; l < 3; l++)
      ranv[(size_t)l * n + ii] = s->ranv ? s->ranv[ii][l] : 0.0;
}

/* ------------------------------------------------------------------------ */
/* overlap_moveback, dana.F90:849-943                                        */
/* ------------------------------------------------------------------------ */

static inline int same3(const double a[3], const double b[3]) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

int orc_overlap(orc_sys *s, orc_rng *r, orc_counters *c, int max_passes) {
  int again = 1, pass = 0;
  while (again) {
    if (max_passes > 0 && pass >= max_passes) return ORC_E_OVERLAP_LOOP;
    pass++;
    if (c) c->passes++;
    again = 0;
    for (int ii = 0; ii < s->ref.n; ii++) {
      int i = s->ref.v[ii];
      if (s->skip[i]) continue;
      s->skip[i] = 1;
      for (int jj = 0; jj < s->nn[i]; jj++) {
        int j = s->lst[i][jj];
        if (s->b_limbo && s->st[j] == ORC_LIMBO) continue;
        double vd[3];
        vdist(s, s->pos[i], s->pos[j], vd);
        double dr = dot3(vd);
        if (dr > s->rcut2) continue;
        if (s->zs[j] == ORC_CG) {
          if (c) c->try_++;
          double ne = r->unif(r, ORC_SITE_CONTACT, 0, 0);
          if (ne < s->p.prob) {
            if (c) c->depo++;
            s->zs[i] = ORC_F;
            if (s->pos[i][2] > s->p.z0) return ORC_E_DEPOSIT_Z0;
          } else {
            for (int k = 0; k < 3; k++) s->pos[i][k] = s->old_cg[i][k];
            s->skip[i] = 0;
          }
          break;
        }
        if (s->p.piston) {
          if (same3(s->pos[j], s->old_cg[j]) && same3(s->pos[i], s->old_cg[i])) {
            if (c) c->choques3++;
            continue;
          }
        }
        for (int k = 0; k < 3; k++) {
          s->pos[j][k] = s->old_cg[j][k];
          s->acel[j][k] = 0.0;
          s->vel[j][k] = 0.0;
        }
        s->skip[j] = 0;
        if (c) c->choques++;
        again = 1;
      }
    }
  }
  return r->err ? ORC_E_INDEX : 0;
}

/*
 * Deterministic parallel-round variant (this build's own schedule, used by
 * the device fast path; NOT a reference routine).  Per round, with the
 * positions of the round start frozen:
 *   active  = ref atoms with skip == 0
 *   dom(i)  = some active k earlier in ref order lies within rcut of i
 *             (k would have moved i back first in the sequential pass)
 *   every active atom sets skip; every active, non-dominated atom then scans
 *   its list in order exactly like dana.F90:875-935, reverting non-CG hits and
 *   stopping at the first CG hit (uniform u_round[round*nslots + i]).
 *   Reverts are applied after the read phase; they are idempotent, so the
 *   result does not depend on thread order.
 * Rounds repeat while a revert happened or a dominated atom is left.
 */
int orc_overlap_rounds(orc_sys *s, const double *u_round, int max_rounds,
                       orc_counters *c) {
  int n = s->amax;
  int *ord = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  uint8_t *act = (uint8_t *)calloc((size_t)n + 1, 1);
  uint8_t *rev = (uint8_t *)calloc((size_t)n + 1, 1);
  uint8_t *selfrev = (uint8_t *)calloc((size_t)n + 1, 1);
  uint8_t *dep = (uint8_t *)calloc((size_t)n + 1, 1);
  uint8_t *dom = (uint8_t *)calloc((size_t)n + 1, 1);
  if (!ord || !act || !rev || !selfrev || !dep || !dom) return ORC_E_ALLOC;
  for (int i = 0; i < n; i++) ord[i] = -1;
  for (int ii = 0; ii < s->ref.n; ii++) ord[s->ref.v[ii]] = ii;
  int status = ORC_E_OVERLAP_LOOP;
  for (int round = 0; round < max_rounds; round++) {
    int nact = 0;
    memset(rev, 0, (size_t)n);
    memset(selfrev, 0, (size_t)n);
    memset(dep, 0, (size_t)n);
    memset(dom, 0, (size_t)n);
    for (int ii = 0; ii < s->ref.n; ii++) {
      int i = s->ref.v[ii];
      act[i] = !s->skip[i];
      nact += act[i];
    }
    if (nact == 0) { status = 0; break; }
    if (c) c->passes++;
    int any_rev = 0, any_dom = 0;
    /* read phase */
    for (int ii = 0; ii < s->ref.n; ii++) {
      int i = s->ref.v[ii];
      if (!act[i]) continue;
      for (int jj = 0; jj < s->nn[i]; jj++) {
        int j = s->lst[i][jj];
        if (s->b_limbo && s->st[j] == ORC_LIMBO) continue;
        if (ord[j] < 0 || !act[j] || ord[j] >= ii) continue;
        double vd[3];
        vdist(s, s->pos[i], s->pos[j], vd);
        if (dot3(vd) <= s->rcut2) { dom[i] = 1; break; }
      }
    }
    for (int ii = 0; ii < s->ref.n; ii++) {
      int i = s->ref.v[ii];
      if (!act[i] || dom[i]) continue;
      for (int jj = 0; jj < s->nn[i]; jj++) {
        int j = s->lst[i][jj];
        if (s->b_limbo && s->st[j] == ORC_LIMBO) continue;
        double vd[3];
        vdist(s, s->pos[i], s->pos[j], vd);
        if (dot3(vd) > s->rcut2) continue;
        if (s->zs[j] == ORC_CG) {
          if (c) c->try_++;
          double ne = u_round[(size_t)round * n + i];
          if (ne < s->p.prob) dep[i] = 1; else selfrev[i] = 1;
          break;
        }
        if (s->p.piston && same3(s->pos[j], s->old_cg[j]) &&
            same3(s->pos[i], s->old_cg[i])) {
          if (c) c->choques3++;
          continue;
        }
        rev[j] = 1;
        if (c) c->choques++;
        any_rev = 1;
      }
    }
    /* write phase */
    for (int ii = 0; ii < s->ref.n; ii++) {
      int i = s->ref.v[ii];
      if (act[i]) {
        if (dom[i]) any_dom = 1; else s->skip[i] = 1;
      }
    }
    for (int i = 0; i < n; i++) {
      if (dep[i]) {
        if (c) c->depo++;
        s->zs[i] = ORC_F;
        if (s->pos[i][2] > s->p.z0) { status = ORC_E_DEPOSIT_Z0; goto out; }
      }
      if (selfrev[i]) {
        for (int k = 0; k < 3; k++) s->pos[i][k] = s->old_cg[i][k];
        s->skip[i] = 0;
      }
      if (rev[i]) {
        for (int k = 0; k < 3; k++) {
          s->pos[i][k] = s->old_cg[i][k];
          s->acel[i][k] = 0.0;
          s->vel[i][k] = 0.0;
        }
        s->skip[i] = 0;
      }
    }
    if (!any_rev && !any_dom) { status = 0; break; }
  }
out:
  free(ord); free(act); free(rev); free(selfrev); free(dep); free(dom);
  return status;
}

void orc_f_to_cg(orc_sys *s, int gcmc) {
  /* dana.F90:228-236 */
  int w = 0;
  for (int r = 0; r < s->ref.n; r++) {
    int i = s->ref.v[r];
    if (s->zs[i] == ORC_F) {
      s->zs[i] = ORC_CG;
      s->in_ref[i] = 0;
      if (gcmc && s->in_gcmc[i]) {
        ivec_remove_value(&s->gc, i);
        s->in_gcmc[i] = 0;
      }
      continue;
    }
    s->ref.v[w++] = i;
  }
  s->ref.n = w;
}

/* ------------------------------------------------------------------------ */
/* GCMC, dana.F90:590-713                                                    */
/* ------------------------------------------------------------------------ */

static int ngroup_sort_atom(orc_sys *s, int i) {
  /* ngroup_sort_atom (Neighbor.F90:271-318) */
  s->nn[i] = 0;
  double rl = (s->rcut + s->p.skin);
  rl = rl * rl;
  if (s->in_ref[i]) {
    int e = s->lista_cells ? cells_atom(s, i, rl) : verlet_atom(s, i, rl);
    if (e) return e;
  }
  for (int jj = 0; jj < s->ref.n; jj++) {
    int aj = s->ref.v[jj];
    if (aj == i) continue;
    double vd[3];
    vdist(s, s->pos[i], s->pos[aj], vd);
    double rd = dot3(vd);
    if (rd > rl) continue;
    int e = list_push(s, aj, i);
    if (e) return e;
  }
  return 0;
}

static int gcmc_insert(orc_sys *s, int tmpl, const double r[3]) {
  int hs, slot, e;
  if ((e = hs_new_slot(s, &hs))) return e;
  atom_default(s, hs);
  /* atom_asign (Groups.F90:484-502) from the template, then new pos */
  for (int k = 0; k < 3; k++) {
    s->vel[hs][k] = s->vel[tmpl][k];
    s->force[hs][k] = s->force[tmpl][k];
    s->acel[hs][k] = s->acel[tmpl][k];
    s->pos[hs][k] = r[k];
    s->pos_old[hs][k] = r[k];
  }
  s->zs[hs] = s->zs[tmpl];
  s->epot[hs] = s->epot[tmpl];
  s->st[hs] = ORC_ALIVE;
  /* groups of the template in id order: sys, ref, b, hs, gcmc */
  if ((e = igrp_attach(&s->sys, hs, 0, &slot))) return e;
  s->sys_of[hs] = slot;
  ivec_push(&s->sysl, hs);
  ivec_push(&s->ref, hs);
  s->in_ref[hs] = 1;
  if ((e = igrp_attach(&s->b, hs, 0, &slot))) return e;
  s->b_of[hs] = slot;
  if (s->bnext_n < s->b.cap) {
    s->bnext = (int *)realloc(s->bnext, sizeof(int) * (size_t)s->b.cap);
    s->bnext_n = s->b.cap;
  }
  if (s->tessellated) {
    if (b_push_front(s, slot)) s->tessellated = 0; /* silent failure */
  }
  s->nat++;
  if (s->listed) {
    if ((e = ngroup_sort_atom(s, hs))) return e;
  }
  ivec_push(&s->gc, hs);
  s->in_gcmc[hs] = 1;
  return 0;
}

static int destroy_atom(orc_sys *s, int i) {
  /* atom_destroy (Groups.F90:433-467): LIFO detach gcmc, hs (ref, b), sys */
  if (s->in_gcmc[i]) {
    ivec_remove_value(&s->gc, i);
    s->in_gcmc[i] = 0;
  }
  if (s->listed) s->nn[i] = 0;
  if (s->in_ref[i]) {
    ivec_remove_value(&s->ref, i);
    s->in_ref[i] = 0;
  }
  int b = s->b_of[i];
  if (s->tessellated) {
    int e = b_unsort(s, b);
    if (e) return e;
  }
  s->b.slot[b] = -1;
  s->b.nat--;
  s->st[i] = ORC_NULL;
  s->nat--;
  if (s->listed) {
    s->st[i] = ORC_LIMBO;
    s->nlimbo++;
    s->b_limbo = 1;
  }
  s->sys.slot[s->sys_of[i]] = -1;
  s->sys.nat--;
  ivec_remove_value(&s->sysl, i);
  return 0;
}

int orc_gcmc(orc_sys *s, orc_rng *r, double act, int nadj, orc_counters *c) {
  /* kB_ui of gems_constants (Constants.F90:100-167) for beta */
  const double axps_mxs = 1.0e2, uma_kg = 1.6605402e-27, qe_si = 1.60219e-19;
  const double joule_ev = 1.0 / qe_si;
  const double ui_ev = axps_mxs * axps_mxs * uma_kg * joule_ev;
  const double ev_ui = 1.0 / ui_ev;
  const double kB_ui = 8.617385e-05 * ev_ui;
  double rc = s->rcut;
  double zmax = s->p.zmax, z0 = s->p.z0;
  double v = s->box[0] * s->box[1] * (zmax - z0);
  int n = 0;
  for (int j = 0; j < s->gc.n; j++) {
    double z = s->pos[s->gc.v[j]][2];
    if (z < z0 || z > zmax) continue;
    n++;
  }
  for (int t = 0; t < nadj; t++) {
    if (s->gc.n == 0) return ORC_E_NO_PARTICLE;
    int tmpl = s->gc.v[0];
    double beta = sqrt(kB_ui * 300.0 / s->p.mass[s->zs[tmpl] - 1]);
    if (r->unif(r, ORC_SITE_GCMC_U, t, 0) < 0.5) {
      if (act * v / (double)(n + 1) < r->unif(r, ORC_SITE_GCMC_U, t, 1)) continue;
      double rr[3];
      rr[0] = r->unif(r, ORC_SITE_GCMC_U, t, 2) * s->box[0];
      rr[1] = r->unif(r, ORC_SITE_GCMC_U, t, 3) * s->box[1];
      rr[2] = r->unif(r, ORC_SITE_GCMC_U, t, 4) * (zmax - z0) + z0;
      int reject = 0;
      for (int j = 0; j < s->gc.n; j++) {
        int o = s->gc.v[j];
        double vd[3];
        /* distance(o%pos, r, o%pbc): vd = r - o%pos (Program_Types.F90:139) */
        vdist(s, rr, s->pos[o], vd);
        double dr = dot3(vd);
        if (dr < rc * rc) { reject = 1; break; }
      }
      if (reject) continue;
      n++;
      int tail = s->gc.v[s->gc.n - 1];
      /* velocity lands on the last atom of the gcmc list (dana.F90:666-668) */
      for (int j = 0; j < 3; j++)
        s->vel[tail][j] = beta * r->gauss(r, ORC_SITE_GCMC_G, t, j);
      int e = gcmc_insert(s, tmpl, rr);
      if (e) return e;
      if (c) c->ins++;
    } else {
      if ((double)n / (v * act) < r->unif(r, ORC_SITE_GCMC_U, t, 1)) continue;
      int m = (int)floor(r->unif(r, ORC_SITE_GCMC_U, t, 2) * n) + 1;
      if (m > n) m = n;
      int o = -1;
      for (int j = 0; j < s->gc.n; j++) {
        o = s->gc.v[j];
        double z = s->pos[o][2];
        if (z < z0) continue;
        if (z > zmax) continue;
        m--;
        if (m == 0) break;
      }
      if (m > 0) return ORC_E_NO_PARTICLE;
      n--;
      int e = destroy_atom(s, o);
      if (e) return e;
      if (c) c->del++;
    }
  }
  return r->err ? ORC_E_INDEX : 0;
}

void orc_gcmc_init(orc_sys *s) {
  /* gcmc group: the Li atoms in sys order (dana.F90:505-513) */
  s->gc.n = 0;
  for (int q = 0; q < s->sysl.n; q++) {
    int i = s->sysl.v[q];
    s->in_gcmc[i] = 0;
    if (s->zs[i] == ORC_LI && s->st[i] == ORC_ALIVE) {
      ivec_push(&s->gc, i);
      s->in_gcmc[i] = 1;
    }
  }
}

int orc_gcmc_order(const orc_sys *s, int32_t *out) {
  if (out)
    for (int j = 0; j < s->gc.n; j++) out[j] = s->gc.v[j];
  return s->gc.n;
}

void orc_gcmc_trial_energy(const orc_sys *s, int ntrial, const double *pos3,
                           const int32_t *exclude, double *du, int32_t *hard) {
  /* Trial energy of an inserted Li at pos3[t]: the reference's acceptance
   * uses only the hard-core test against the gcmc group
   * (dana.F90:638-653, rc = hs%rcut, '<'); the energy reported beside it is
   * the WCA pair sum of fuerza (dana.F90:1110-1131) against every alive
   * atom, full (not halved) pair energy, FP64, in slot order. */
  for (int t = 0; t < ntrial; t++) {
    const double *r = pos3 + 3 * (size_t)t;
    double e = 0.0;
    int h = 0;
    for (int j = 0; j < s->amax; j++) {
      if (s->st[j] != ORC_ALIVE) continue;
      if (exclude && exclude[t] == j) continue;  /* the atom removed */
      double vd[3];
      vdist(s, r, s->pos[j], vd);
      double dr2 = dot3(vd);
      if (s->in_gcmc[j] && dr2 < s->rcut * s->rcut) h = 1;
      int m = s->zs[j];
      double r0 = s->p.r0[(1 - 1) * 3 + (m - 1)];
      double eps = s->p.eps[(1 - 1) * 3 + (m - 1)];
      if (dr2 > r0 * r0) continue;
      double dr = sqrt(dr2);
      double b = pow6(r0) / pow6(dr);
      e += eps * b * (b - 2) + eps;
    }
    du[t] = e;
    hard[t] = h;
  }
}

/* ------------------------------------------------------------------------ */
/* the dana program, dana.F90:1-305                                          */
/* ------------------------------------------------------------------------ */

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static double calc_rho(orc_sys *s, int chunks) {
  /* calc_rho (dana.F90:521-549) */
  double z = chunks ? s->z1 : s->p.zmax;
  int g = 0;
  for (int r = 0; r < s->sysl.n; r++) {
    double pz = s->pos[s->sysl.v[r]][2];
    if (pz > s->p.z0 && pz < z) g++;
  }
  double vol = s->box[0] * s->box[1] * (z - s->p.z0);
  return (double)g / vol;
}

void orc_observables(const orc_sys *s, double out[4]) {
  /* salida (dana.F90:1147-1158): energia = energia + pa%epot over sys%alist */
  double energia = 0.0, vdac = 0.0;
  int j = 0;
  for (int q = 0; q < s->sysl.n; q++) {
    int i = s->sysl.v[q];
    energia = energia + s->epot[i];
  }
  /* kion (dana.F90:1342-1376): CG skipped, vd = dot(vel,vel)*mass */
  for (int q = 0; q < s->sysl.n; q++) {
    int i = s->sysl.v[q];
    if (s->zs[i] == 2) continue;
    j = j + 1;
    double vd = s->vel[i][0] * s->vel[i][0] + s->vel[i][1] * s->vel[i][1] +
                s->vel[i][2] * s->vel[i][2];
    vd = vd * s->p.mass[s->zs[i] - 1];
    vdac = vdac + vd;
  }
  const double kB_eVK = 8.617330350e-5, kJm_ui = 100.0, eV_kJm = 96.485;
  const double kB_ui = kB_eVK * (eV_kJm * kJm_ui); /* dana.F90:22-24 */
  out[0] = energia;
  out[1] = vdac / (j * 3.0 * kB_ui);
  out[2] = j;
  out[3] = s->sysl.n;
}

double orc_calc_rho(orc_sys *s, int chunks) { return calc_rho(s, chunks); }

void orc_set_z1(orc_sys *s, double z1) { s->z1 = z1; }

void orc_get_reservoir(const orc_sys *s, double out[4]) {
  out[0] = s->p.z0; out[1] = s->z1; out[2] = s->p.zmax; out[3] = s->box[2];
}

int orc_chunk_init(const orc_sys *s, orc_chunk *ch, int n, const double *pos3, double dist) {
  /* config_chunk (dana.F90:552-587): the block template sits at zmax + z, and
   * its pos_old another zmax higher */
  ch->n = n;
  ch->dist = dist;
  ch->pos = (double(*)[3])malloc(sizeof(double[3]) * (size_t)(n > 0 ? n : 1));
  ch->old = (double(*)[3])malloc(sizeof(double[3]) * (size_t)(n > 0 ? n : 1));
  if (!ch->pos || !ch->old) return ORC_E_ALLOC;
  for (int j = 0; j < n; j++) {
    for (int k = 0; k < 3; k++) {
      ch->pos[j][k] = pos3[3 * j + k];
      ch->old[j][k] = ch->pos[j][k];
    }
    ch->pos[j][2] = ch->pos[j][2] + s->p.zmax;
    ch->old[j][2] = ch->pos[j][2] + s->p.zmax;
  }
  return 0;
}

void orc_chunk_free(orc_chunk *ch) {
  free(ch->pos);
  free(ch->old);
  ch->pos = ch->old = NULL;
  ch->n = 0;
}

int orc_bloques(orc_sys *s, orc_chunk *ch, double rho, int *fired, orc_counters *c) {
  /* bloques (dana.F90:716-773) */
  const double rhomedia = 5.775329e-4;
  double dist = ch->dist;
  double drho = rho - rhomedia;
  *fired = 0;
  if (fabs(drho) < (rhomedia * (double)0.186f)) return 0;
  *fired = 1;
  s->p.z0 = s->p.z0 + dist;
  s->z1 = s->z1 + dist;
  s->p.zmax = s->p.zmax + dist;
  s->listed = 0;
  for (int j = 0; j < ch->n; j++) {
    int hs, slot, e;
    if ((e = hs_new_slot(s, &hs))) return e;
    atom_default(s, hs);
    s->st[hs] = ORC_ALIVE;
    if ((e = igrp_attach(&s->sys, hs, 0, &slot))) return e;
    s->sys_of[hs] = slot;
    ivec_push(&s->sysl, hs);
    if ((e = igrp_attach(&s->b, hs, 0, &slot))) return e;
    s->b_of[hs] = slot;
    if (s->bnext_n < s->b.cap) {
      s->bnext = (int *)realloc(s->bnext, sizeof(int) * (size_t)s->b.cap);
      s->bnext_n = s->b.cap;
    }
    /* sorted while still at the origin (pos set after attach) */
    if (s->tessellated) b_push_front(s, slot);
    ivec_push(&s->ref, hs);
    s->in_ref[hs] = 1;
    s->nat++;
    for (int k = 0; k < 3; k++) {
      s->pos[hs][k] = ch->pos[j][k];
      s->pos_old[hs][k] = ch->old[j][k];
    }
    s->zs[hs] = ORC_LI;
    ch->pos[j][2] = ch->pos[j][2] + dist;
    ch->old[j][2] = ch->old[j][2] + dist;
  }
  double nb[3] = {s->box[0], s->box[1], s->p.zmax};
  orc_set_box(s, nb);
  return orc_test_update(s, NULL, c);
}

/* maxz (dana.F90:776-794): the piston factor lohi from rho (calc_rho of
   this step) and rho0 (the initial one), applied to every sys atom above z0
   and to zmax */
void orc_maxz(orc_sys *s, double rho, double rho0) {
  const double tau = 0.1;
  double lohi = ((s->p.h / tau) * ((rho0 - rho) / rho));
  for (int q = 0; q < s->sys.nat; q++) {
    int i = s->sys.slot[q];
    if (s->pos[i][2] > s->p.z0) s->pos[i][2] = s->pos[i][2] - lohi * (s->pos[i][2] - s->p.z0);
  }
  s->p.zmax = s->p.zmax - lohi * (s->p.zmax - s->p.z0);
}

int orc_dana_run(const orc_run_cfg *cfg, orc_run_out *out) {
  double *u_round = NULL; /* round-schedule uniforms (overlap_rounds > 0) */
  size_t u_cap = 0;
  uint64_t u_state = 0;
  orc_params p;
  memset(&p, 0, sizeof(p));
  /* eps and r0 (dana.F90:86-100); index [(k-1)*3 + (m-1)] */
  p.r0[1 * 3 + 0] = 3.5; p.r0[0 * 3 + 1] = 3.5;
  p.eps[0] = 2313.6; p.r0[0] = 3.2;
  p.eps[8] = 121.0; p.r0[8] = 3.61;
  p.eps[0 * 3 + 2] = 529.1; p.eps[2 * 3 + 0] = 529.1;
  p.r0[0 * 3 + 2] = 1.564; p.r0[2 * 3 + 0] = 1.564;
  p.mass[0] = p.mass[1] = p.mass[2] = 6.94;
  p.rcut = 3.2;
  p.skin = cfg->nb_dcut;
  p.h = cfg->h; p.gama = 1.0; p.temp = 300.0;
  p.dif_sc = cfg->dif_sc; p.dif_sei = cfg->dif_sei; p.z_switch = 80.0;
  p.prob = cfg->prob; p.z0 = cfg->z0; p.zmax = cfg->zmax;
  p.ermak = cfg->ermak;
  p.piston = cfg->reservoir == 0;
  p.box[0] = cfg->xi; p.box[1] = cfg->yi; p.box[2] = cfg->zmax;
  if (cfg->reservoir < 0 || cfg->reservoir > 3) return ORC_E_RESERVOIR;
  int chunks = cfg->reservoir == 1, gcmc = cfg->reservoir == 2;

  orc_rng rng;
  orc_rng_init_nr(&rng, cfg->idum);
  orc_sys *s = orc_sys_new(&p, 1024);
  if (!s) return ORC_E_ALLOC;
  int status = 0;
  double dist = cfg->dist;
  orc_chunk ch;
  memset(&ch, 0, sizeof ch);

  /* pos_inic (dana.F90:330-396); box = (xi, yi, zmax) */
  int n;
  double *r = NULL;
  int8_t *zinit = NULL;
  if (cfg->ninit > 0) {
    n = cfg->ninit;
    r = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    memcpy(r, cfg->init_pos, sizeof(double) * 3 * (size_t)n);
    zinit = (int8_t *)malloc((size_t)n);
    for (int i = 0; i < n; i++) zinit[i] = cfg->init_z ? cfg->init_z[i] : ORC_LI;
  } else {
    double alto = cfg->zmax - 0.0;
    n = (int)(1.0 * cfg->xi * cfg->yi * alto * (double)6.022e-4f);
    r = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) {
      int k;
      for (k = 1; k <= 10000; k++) {
        r[3 * i + 0] = orc_nr_ran(&rng) * cfg->xi;
        r[3 * i + 1] = orc_nr_ran(&rng) * cfg->yi;
        r[3 * i + 2] = (orc_nr_ran(&rng) * alto) + 0.0;
        int clash = 0;
        for (int j = 0; j < i; j++) {
          double vd[3];
          vdist(s, r + 3 * i, r + 3 * j, vd);
          if (dot3(vd) < 3.2 * 3.2) { clash = 1; break; }
        }
        if (!clash) break;
      }
      if (k == 10001) { status = ORC_E_PLACEMENT; goto done; }
      /* written with f25.12 and read back (dana.F90:391, 463) */
      for (int kk = 0; kk < 3; kk++) {
        char buf[64];
        snprintf(buf, sizeof buf, "%25.12f", r[3 * i + kk]);
        r[3 * i + kk] = strtod(buf, NULL);
      }
    }
  }
  /* config_inic (dana.F90:430-517) */
  double z1 = 0.0;
  if (chunks) {
    dist = dist + s->rcut;
    z1 = cfg->z0 + dist;
    s->p.zmax = z1 + dist;
  }
  s->z1 = z1;
  for (int i = 0; i < n; i++) {
    int hs = orc_attach_initial(s, zinit ? zinit[i] : ORC_LI, r + 3 * i);
    if (hs < 0) { status = hs; goto done; }
  }
  if (cfg->ermak) orc_set_ermak(s);
  if (gcmc)
    for (int q = 0; q < s->sysl.n; q++) {
      int i = s->sysl.v[q];
      if (s->zs[i] == ORC_LI) { ivec_push(&s->gc, i); s->in_gcmc[i] = 1; }
    }
  orc_counters cnt;
  memset(&cnt, 0, sizeof cnt);
  if ((status = orc_test_update(s, NULL, &cnt))) goto done;
  if (cfg->ermak) orc_fuerza(s);
  double rho = calc_rho(s, chunks);
  double rho0 = rho;
  /* config_chunk (dana.F90:552-587) */
  if (chunks && (status = orc_chunk_init(s, &ch, cfg->nchunk, cfg->chunk_pos, dist))) goto done;
  u_state = cfg->rounds_seed;
  double t0 = now_s();
  int64_t psteps = 0;
  for (int step = 1; step <= cfg->nst; step++) {
    psteps += s->ref.n;
    if (cfg->ermak) {
      if ((status = orc_ermak_a(s, &rng, &cnt))) goto done;
      orc_fuerza(s);
      orc_ermak_b(s);
    } else {
      if ((status = orc_brownian(s, &rng, &cnt))) goto done;
    }
    if ((status = orc_test_update(s, NULL, &cnt))) goto done;
    if (cfg->overlap_rounds > 0) {
      /* the round schedule with its own uniforms (splitmix64 per call) */
      size_t need = (size_t)cfg->overlap_rounds * (size_t)s->amax;
      if (need > u_cap) {
        free(u_round);
        u_cap = need;
        u_round = (double *)malloc(sizeof(double) * u_cap);
        if (!u_round) { status = ORC_E_ALLOC; goto done; }
      }
      for (size_t q = 0; q < need; q++) {
        uint64_t z = (u_state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        u_round[q] = ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      }
      int st = orc_overlap_rounds(s, u_round, cfg->overlap_rounds, &cnt);
      if (st == ORC_E_OVERLAP_LOOP) cnt.ov_unfinished++;
      else if ((status = st)) goto done;
    } else if ((status = orc_overlap(s, &rng, &cnt, 100000))) {
      goto done;
    }
    if ((status = orc_test_update(s, NULL, &cnt))) goto done;
    orc_f_to_cg(s, gcmc);
    if (gcmc)
      if ((status = orc_gcmc(s, &rng, cfg->act, cfg->nadj, &cnt))) goto done;
    rho = calc_rho(s, chunks);
    if (chunks) {
      int fired;
      if ((status = orc_bloques(s, &ch, rho, &fired, &cnt))) goto done;
    }
    if (cfg->nwr > 0 && step % cfg->nwr == 0) {
      /* salida() writes the frame before the piston move (dana.F90:250-261) */
      out->nat = s->sysl.n;
      out->zmax = s->p.zmax;
      for (int q = 0; q < s->sysl.n && q < out->cap; q++) {
        int i = s->sysl.v[q];
        for (int k = 0; k < 3; k++) out->pos[3 * q + k] = s->pos[i][k];
        out->z[q] = s->zs[i];
      }
    }
    if (s->p.piston) orc_maxz(s, rho, rho0);
  }
  out->loop_seconds = now_s() - t0;
  out->particle_steps = psteps;
  if (cfg->nwr <= 0 || cfg->nst % cfg->nwr != 0) {
    /* no salida at the last step: report the state at loop end */
    out->nat = s->sysl.n;
    out->zmax = s->p.zmax;
    for (int q = 0; q < s->sysl.n && q < out->cap; q++) {
      int i = s->sysl.v[q];
      for (int k = 0; k < 3; k++) out->pos[3 * q + k] = s->pos[i][k];
      out->z[q] = s->zs[i];
    }
  }
  out->c = cnt;
done:
  free(u_round);
  orc_chunk_free(&ch);
  free(r);
  free(zinit);
  orc_sys_free(s);
  return status;
}
