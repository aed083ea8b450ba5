/*
 * emd_app.h -- C ABI of the whole ExaMiniMD driver (init + time-step loop)
 * running on the GPU modules.  It wraps the C++ class ExaMiniMD, whose init()
 * and run() follow src/examinimd.cpp:60-294 of the reference; the arguments
 * are the reference's command line (src/input.cpp:151-224), e.g.
 *   {"ExaMiniMD", "-il", "in.lj", "--neigh-type", "CSR", "--precision", "mixed"}
 * Every call returns 0 on success.  The driver prints nothing when quiet != 0.
 */
#ifndef EMD_APP_H
#define EMD_APP_H

#include "emd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct emd_app emd_app;

int      emd_app_create(int argc, const char **argv, int quiet, emd_app **out);
/* ExaMiniMD::init: build modules, lattice, first neighbour list and forces */
int      emd_app_init(emd_app *app);
/* ExaMiniMD::run(nsteps) */
int      emd_app_run(emd_app *app, int nsteps);
/* synchronise the device between phases (per-phase timers) or not */
int      emd_app_set_phase_sync(emd_app *app, int on);
/* fuse final_integrate with the next step's initial_integrate where no thermo,
 * dump or correctness check falls between them (on by default; bitwise equal) */
int      emd_app_set_fused_integrate(emd_app *app, int on);
int      emd_app_destroy(emd_app *app);

int      emd_app_nsteps(const emd_app *app);       /* the deck's `run` count */
int      emd_app_natoms(const emd_app *app);       /* N (global) */
int      emd_app_nlocal(const emd_app *app);
int      emd_app_nghost(const emd_app *app);
/* T, PE/N, (PE+KE)/N */
int      emd_app_thermo(emd_app *app, double out[3]);
/* seconds of the last run(): 0 total, 1 force, 2 neigh, 3 comm, 4 other */
double   emd_app_time(const emd_app *app, int which);
/* host copies of the local atoms (any pointer may be NULL); f is FP64 */
int      emd_app_download(emd_app *app, double *x, double *v, double *f, int *type, int *id,
                          double *q);
/* reference-format binary dump of the local atoms to `path` */
int      emd_app_dump(emd_app *app, const char *path);
/* canonical neighbour-list file of the current list (the --dumpneigh
 * format: int n, long total, then per local atom in ascending id order
 * int id, int count and count (id_j, image_j) int pairs sorted ascending;
 * image_j bits 2d..2d+1 = periodic shift on axis d, 0 none, 1 -L, 2 +L).
 * The reference writes no neighbour lists; this exists so that lists of two
 * implementations can be compared offline. */
int      emd_app_dump_neigh(emd_app *app, const char *path);
/* device objects for callers that go below the driver */
emd_ctx          *emd_app_ctx(emd_app *app);
emd_system       *emd_app_system(emd_app *app);
const emd_neigh  *emd_app_neigh(emd_app *app);
const emd_halo   *emd_app_halo(emd_app *app);
int               emd_app_precision(const emd_app *app);

/* An ensemble run the conventional way, without batching: every replica of
 * `d` (see emd_batch_desc) is its own system through the per-system C ABI
 * (exchange, binning, halo, CSR list, LJ force, NVE: the reference's step
 * src/examinimd.cpp:176-294), replicas visited one after another each step,
 * one launch sequence per replica.  x, v: [total][3] host, replica-major.
 * Runs the set-up force, then nsteps timed steps; *seconds = wall time of
 * those steps.  x_out / v_out (may be NULL) receive the final state in the
 * input order of each replica.  The reference point for emd_batch_run. */
int emd_replica_loop_run(emd_ctx *ctx, const emd_batch_desc *d, const double *x,
                         const double *v, int nsteps, double *seconds, double *x_out,
                         double *v_out);

#ifdef __cplusplus
}
#endif
#endif
