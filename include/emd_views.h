/*
 * emd_views.h -- helpers for a module that lives in the reference's own
 * source tree (a Binning/Neighbor/Force/Integrator/Comm subclass registered
 * through its modules_*.h) and calls this library.  Kokkos-free: the module
 * passes the raw device pointers of its Views (View::data(), extent(0)) and
 * their layouts; nothing here includes Kokkos.
 *
 *   emd_ctx_of(device)             the process's context on that GPU
 *   make_emd_system(ctx, &ref, &s) System's Views -> emd_system (device copy,
 *                                  v/f transposed from LayoutLeft)
 *   store_emd_system(ctx, &s, &ref) x, v, f back into the Views
 *   make_emd_bins(ctx, &ref, &b)   Binning's fields -> emd_bins
 *
 * The reference's layouts (src/types.h:88-100): x is LayoutRight; v and f use
 * the default layout, LayoutLeft on the GPU backends; Binning's bincount and
 * binoffsets (src/binning.h:54-59) are View<int***> in the default layout.
 */
#ifndef EMD_VIEWS_H
#define EMD_VIEWS_H

#include "emd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* the fields of the reference's System (src/system.h:57-143) that the hot
 * path reads, as raw device pointers */
typedef struct {
  int N, N_local, N_ghost, ntypes;
  emd_vec3_view x, v, f;      /* x: {x.data(), EMD_LAYOUT_RIGHT, 0};
                                 v, f: {data(), EMD_LAYOUT_LEFT, extent(0)} */
  const int *type, *id;       /* device [N_local + N_ghost] */
  const double *q;
  const double *mass_host;    /* host copy of the per-type masses [ntypes] */
  double domain[3], sub_domain_lo[3], sub_domain_hi[3];
  double boltz, mvv2e, dt;
} emd_ref_system;

/* the fields of the reference's Binning (src/binning.h:50-61) */
typedef struct {
  int nbinx, nbiny, nbinz, nhalo;
  double min[3], max[3];
  const int *bincount, *binoffsets;  /* device, View<int***>::data() */
  int layout;                        /* their layout: EMD_LAYOUT_LEFT on the GPU backends */
  const int *permute_vector;         /* device [n_range] */
  int n_range;
} emd_ref_bins;

static inline emd_ctx *emd_ctx_of(int device) {
  emd_ctx *ctx = 0;
  return emd_ctx_shared(device, &ctx) == EMD_OK ? ctx : 0;
}

/* copies locals and ghosts; s keeps its own device arrays between calls */
static inline emd_status make_emd_system(emd_ctx *ctx, const emd_ref_system *r, emd_system *s) {
  const int n = r->N_local + r->N_ghost;
  emd_status st = emd_system_import(ctx, s, n, r->x, r->v, r->f, r->type, r->id, r->q);
  if (st != EMD_OK) return st;
  if (r->mass_host && r->ntypes > 0) {
    st = emd_system_set_mass(ctx, s, r->ntypes, r->mass_host);
    if (st != EMD_OK) return st;
  }
  s->n_local = r->N_local;
  s->n_ghost = r->N_ghost;
  s->n_global = r->N;
  for (int d = 0; d < 3; d++) {
    s->domain[d] = r->domain[d];
    s->lo[d] = r->sub_domain_lo[d];
    s->hi[d] = r->sub_domain_hi[d];
  }
  s->boltz = r->boltz;
  s->mvv2e = r->mvv2e;
  s->dt = r->dt;
  return EMD_OK;
}

/* positions, velocities and forces of locals and ghosts back into the
 * reference's Views (a view whose data is NULL is skipped) */
static inline emd_status store_emd_system(emd_ctx *ctx, const emd_system *s,
                                          const emd_ref_system *r) {
  return emd_system_export(ctx, s, s->n_local + s->n_ghost, r->x, r->v, r->f);
}

static inline emd_status make_emd_bins(emd_ctx *ctx, const emd_ref_bins *r, emd_bins *b) {
  return emd_bins_import(ctx, r->nbinx, r->nbiny, r->nbinz, r->nhalo, r->min, r->max,
                         r->bincount, r->binoffsets, r->layout, r->permute_vector, r->n_range, b);
}

#ifdef __cplusplus
}
#endif
#endif
