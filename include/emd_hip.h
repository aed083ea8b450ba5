/*
 * emd_hip.h -- C ABI of the MI355X (gfx950) implementation of ExaMiniMD's
 * per-step hot path: serial ghost exchange, binning, neighbour lists, LJ and
 * SNAP forces, the NVE integrator and the thermo reductions.
 *
 * The reference exposes this path only as C++ virtual methods of its module
 * base classes; each entry point below names the method it replaces
 * (paths relative to the ExaMiniMD source tree).  A host-side module class
 * (Binning, Neighbor, Force, Integrator, Comm subclass) calls exactly one of
 * these per virtual call.
 *
 * Conventions
 *  - Every function returns an emd_status (0 = EMD_OK).  Nothing calls exit();
 *    the text of the last error is available from emd_last_error().
 *  - All array pointers inside the structs are DEVICE pointers owned by the
 *    library (allocated through the *_grow / *_free calls) unless stated.
 *  - Work is ordered on the context's single HIP stream.  Only calls that
 *    must return a count or a scalar to the host synchronise (the same points
 *    at which the reference synchronises).  A context is not thread-safe.
 *  - Layouts in HBM: x, v, f are AoS double[n_max][3] (the reference's t_x
 *    layout, src/types.h:88); f32 is AoS float[n_max][3] used by the
 *    mixed-precision path; type, id are int32[n_max]; q double[n_max];
 *    mass double[ntypes].
 */
#ifndef EMD_HIP_H
#define EMD_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int emd_status;
enum {
  EMD_OK = 0,
  EMD_ERR_ARG = 1,       /* invalid argument / inconsistent sizes */
  EMD_ERR_HIP = 2,       /* HIP runtime error */
  EMD_ERR_ALLOC = 3,     /* device allocation failed */
  EMD_ERR_OVERFLOW = 4,  /* an atom left the bin grid / a capacity was exceeded */
  EMD_ERR_STATE = 5      /* call made in the wrong state */
};

enum { EMD_FP64 = 0, EMD_MIXED = 1 };       /* force precision */
enum { EMD_NEIGH_CSR = 0, EMD_NEIGH_2D = 1 }; /* neighbour list layout */

typedef struct emd_ctx emd_ctx;

/* ------------------------------------------------------------ context */
/* replaces Kokkos::initialize / finalize (src/main.cpp:54,70) */
emd_status  emd_create(int device, emd_ctx **out);
emd_status  emd_destroy(emd_ctx *ctx);
const char *emd_last_error(const emd_ctx *ctx);
emd_status  emd_synchronize(emd_ctx *ctx);
void       *emd_get_stream(emd_ctx *ctx);   /* the hipStream_t all work runs on */
const char *emd_version(void);

/* device memory for callers (fixtures, tests, host modules) */
emd_status emd_malloc(emd_ctx *ctx, void **ptr, size_t bytes);
emd_status emd_free(emd_ctx *ctx, void *ptr);
emd_status emd_memcpy_h2d(emd_ctx *ctx, void *dst, const void *src, size_t bytes);
emd_status emd_memcpy_d2h(emd_ctx *ctx, void *dst, const void *src, size_t bytes);
emd_status emd_memcpy_d2d(emd_ctx *ctx, void *dst, const void *src, size_t bytes);

/* per-kernel timing with HIP events recorded on the context stream */
emd_status emd_profile_enable(emd_ctx *ctx, int on);
emd_status emd_profile_reset(emd_ctx *ctx);
/* restrict timing to the comma-separated kernel names in `names` (NULL or ""
   times every launch); each timed launch adds two event records, so a
   timed run that needs one kernel's duration selects only that kernel */
emd_status emd_profile_select(emd_ctx *ctx, const char *names);
/* total milliseconds and launch count of kernel `name` since the last reset */
emd_status emd_profile_get(emd_ctx *ctx, const char *name, double *ms, long *launches);
/* newline-separated "name launches total_ms" table; returns bytes needed */
size_t     emd_profile_report(emd_ctx *ctx, char *buf, size_t len);

/* ------------------------------------------------------------- system */
/* atom types a system may hold (the device mass table's size) */
#define EMD_MAX_TYPES 1024

/* Device image of System (src/system.h:57-143). */
typedef struct {
  double *x, *v, *f;
  float  *f32;
  int    *type, *id;
  double *q;
  double *mass;
  int n_local, n_ghost, n_max, ntypes, n_global;
  double domain[3], lo[3], hi[3];
  double boltz, mvv2e, dt;
} emd_system;

/* System::init / System::grow (src/system.cpp:71-109): allocate or grow all
 * per-atom arrays to n (contents preserved, new tail zeroed) */
emd_status emd_system_grow(emd_ctx *ctx, emd_system *s, int n);
emd_status emd_system_free(emd_ctx *ctx, emd_system *s);
/* host <-> device copies of the first n atoms (AoS double[n][3]) */
emd_status emd_system_upload(emd_ctx *ctx, emd_system *s, int n, const double *x,
                             const double *v, const int *type, const int *id,
                             const double *q);
emd_status emd_system_download(emd_ctx *ctx, const emd_system *s, int n, double *x,
                               double *v, double *f, int *type, int *id, double *q);
/* per-type masses (host array of ntypes, 1 <= ntypes <= EMD_MAX_TYPES) */
emd_status emd_system_set_mass(emd_ctx *ctx, emd_system *s, int ntypes, const double *mass);

/* The reference's per-atom Views in device memory (src/types.h:88-100): x is
 * View<double*[3], LayoutRight> (a[3 i + d]); v and f are View<double*[3]>
 * with the default layout, which is LayoutLeft on the GPU backends
 * (a[i + d ld], ld = the View's extent(0), at least the atom count). */
enum { EMD_LAYOUT_RIGHT = 0, EMD_LAYOUT_LEFT = 1 };
typedef struct {
  double *data;   /* device pointer (View::data()); NULL = not copied */
  int layout;     /* EMD_LAYOUT_RIGHT or EMD_LAYOUT_LEFT */
  long ld;        /* LayoutLeft: column stride (View::extent(0)); else ignored */
} emd_vec3_view;
/* System arrays held in the reference's View layouts -> emd_system (device to
 * device, stream-ordered; grows s to n).  type, id, q are device int/int/
 * double arrays of n (or NULL).  Replaces reading System's Views directly
 * (src/system.h:57-143) for a module handed the reference's System. */
emd_status emd_system_import(emd_ctx *ctx, emd_system *s, int n, emd_vec3_view x,
                             emd_vec3_view v, emd_vec3_view f, const int *type, const int *id,
                             const double *q);
/* emd_system -> the reference's Views (x, v, f of the first n atoms; a view
 * with data NULL is skipped); values are copied bit for bit */
emd_status emd_system_export(emd_ctx *ctx, const emd_system *s, int n, emd_vec3_view x,
                             emd_vec3_view v, emd_vec3_view f);
/* one context per device and process, created on first use and kept until
 * exit: what a module in the reference's tree, which has no context of its
 * own, uses (emd_ctx_of in include/emd_views.h) */
emd_status emd_ctx_shared(int device, emd_ctx **out);

/* ----------------------------------------------------- comm (serial) */
typedef struct {
  int num_ghost[6];
  int *pack[6];     /* device: source index of each ghost, per phase */
  int cap[6];
  int *root;        /* device [n_ghost]: owning local atom of each ghost */
  int *image;       /* device [n_ghost]: packed periodic shift (3 x 2 bits) */
  int root_cap;
  int n_ghost;
} emd_halo;

/* CommSerial::exchange, src/comm_types/comm_serial.cpp:47-53 */
emd_status emd_comm_exchange(emd_ctx *ctx, emd_system *s);
/* CommSerial::exchange_halo, comm_serial.cpp:55-97 (grows s as needed;
 * ghost order within a phase is ascending source index) */
emd_status emd_comm_exchange_halo(emd_ctx *ctx, emd_system *s, double comm_depth,
                                  emd_halo *h);
/* CommSerial::update_halo, comm_serial.cpp:99-110 (positions only) */
emd_status emd_comm_update_halo(emd_ctx *ctx, emd_system *s, const emd_halo *h);
/* CommSerial::update_force, comm_serial.cpp:112-127 (folds f, or f32 when
 * precision == EMD_MIXED) */
emd_status emd_comm_update_force(emd_ctx *ctx, emd_system *s, const emd_halo *h,
                                 int precision);
emd_status emd_halo_free(emd_ctx *ctx, emd_halo *h);

/* ------------------------------------------------------------ binning */
typedef struct {
  int nbinx, nbiny, nbinz, nhalo;
  double minx, maxx, miny, maxy, minz, maxz;
  int nbins, n_range;
  int *bincount;    /* device [nbins], x-major (ix*nby + iy)*nbz + iz */
  int *binoffsets;  /* device [nbins] */
  int *permute;     /* device [n_range] */
  int *binid;       /* device [n_range] scratch */
  int cap_bins, cap_range;
  int max_count;    /* largest bin population (host copy) */
} emd_bins;

/* BinningKKSort::create_binning, src/binning_types/binning_kksort.cpp:71-140 */
emd_status emd_binning(emd_ctx *ctx, emd_system *s, double dx, double dy, double dz,
                       int halo_depth, int do_local, int do_ghost, int sort, emd_bins *b);
emd_status emd_bins_free(emd_ctx *ctx, emd_bins *b);
/* the reference's Binning fields (src/binning.h:50-61: nbin, nhalo, min/max,
 * bincount and binoffsets View<int***> in the default layout -- LayoutLeft
 * on the GPU backends -- and permute_vector of n_range) -> b, copied into
 * b's own x-major arrays (device to device), so a neighbour module handed
 * the reference's Binning can call emd_neigh_build */
emd_status emd_bins_import(emd_ctx *ctx, int nbinx, int nbiny, int nbinz, int nhalo,
                           const double min[3], const double max[3], const int *bincount,
                           const int *binoffsets, int layout, const int *permute, int n_range,
                           emd_bins *b);

/* --------------------------------------------------------- neighbours */
typedef struct {
  int layout;       /* EMD_NEIGH_CSR or EMD_NEIGH_2D */
  int n;            /* rows = n_local */
  long total;       /* number of entries */
  int maxneighs;    /* 2D row pitch; largest row for CSR */
  int *row;         /* CSR: device [n+1] offsets */
  int *num;         /* device [n] row lengths */
  int *entries;     /* CSR: [total]; 2D: [n][maxneighs] row-major */
  long cap_entries;
  int cap_rows;
} emd_neigh;

/* NeighborCSR::create_neigh_list (src/neighbor_types/neighbor_csr.h:370-431)
 * and Neighbor2D::create_neigh_list (neighbor_2d.h:280-331).  Row order is
 * the 27-bin stencil order then permutation order (deterministic). */
emd_status emd_neigh_build(emd_ctx *ctx, const emd_system *s, const emd_bins *b,
                           double neigh_cut, int half, int newton, int layout,
                           emd_neigh *nl);
emd_status emd_neigh_free(emd_ctx *ctx, emd_neigh *nl);

/* ----------------------------------------------------------------- LJ */
typedef struct {
  int ntypes;
  double lj1[16][16], lj2[16][16], cutsq[16][16];
  /* NULL: the coefficients are the fixed arrays (ntypes <= 16; the
   * reference's stack parameters, force_lj_neigh_impl.h:43-58).  Otherwise
   * a HOST array of 3 * ntypes * ntypes doubles, lj1 then lj2 then cutsq,
   * each row-major [ti][tj], for any ntypes up to EMD_MAX_TYPES (the
   * reference's View path for more types than its stack limit,
   * force_lj_neigh_impl.h:100-126); the fixed arrays are then ignored.
   * lj/cut/idial and emd_lj_energy_noalloc take the fixed arrays only. */
  const double *table;
} emd_lj_params;

/* deep_copy(system->f, 0.0), src/examinimd.cpp:232 (f, or f32 for mixed) */
emd_status emd_force_zero(emd_ctx *ctx, emd_system *s, int precision);
/* ForceLJNeigh::compute, src/force_types/force_lj_neigh_impl.h:100-126 */
emd_status emd_lj_force(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                        const emd_lj_params *p, int precision, int half);
/* deep_copy(f, 0) then ForceLJNeigh::compute on a full list
 * (src/examinimd.cpp:232, force_lj_neigh_impl.h:100-126) as one pass: each
 * local atom's force is written, not added, so f needs no zeroing and is not
 * read; ghost forces are left as they are (a full list never writes them).
 * Same values as emd_force_zero + emd_lj_force(half = 0). */
emd_status emd_lj_force_store(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                              const emd_lj_params *p, int precision);
/* pair_style lj/cut/idial (ForceLJIDialNeigh,
 * src/force_types/force_lj_idial_neigh{.h,_impl.h}): lj/cut's parameters
 * plus a per-type-pair repeat count.  init_coeff (force_lj_idial_neigh_impl.h:
 * 50-88) sets entries (t1,t2) and (t2,t1) of lj1 = 48 eps sigma^12,
 * lj2 = 24 eps sigma^6, cutsq = cut^2 and intensity = the 6th pair_coeff word. */
typedef struct {
  emd_lj_params lj;
  double intensity[16][16];
} emd_lj_idial_params;
/* ForceLJIDialNeigh::compute, force_lj_idial_neigh_impl.h:90-111: the lj/cut
 * pair set (rsq < cutsq); each pair's force is summed from 0 in `intensity`
 * parts of F/intensity, in the reference's order, so the FP64 result is the
 * reference's to the bit (one lane per atom) and equals lj/cut's for
 * intensity 1 or 2.  The half-list form (:169-219) subtracts the reaction
 * from local j only.  There is no energy (Force::compute_energy's 0).
 * Intensities above 1e6 are rejected (EMD_ERR_ARG): one launch would run
 * for seconds at millions of atoms. */
emd_status emd_lj_idial_force(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                              const emd_lj_idial_params *p, int precision, int half);
/* the same on a full list with f written instead of added (see
 * emd_lj_force_store): equal to emd_force_zero + emd_lj_idial_force(half = 0) */
emd_status emd_lj_idial_force_store(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                    const emd_lj_idial_params *p, int precision);
/* ForceLJNeigh::compute_energy, force_lj_neigh_impl.h:128-156 (compensated) */
emd_status emd_lj_energy(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                         const emd_lj_params *p, int half, double *pe);
/* the same energy per local atom into eatom (DEVICE, double[n_local]): the
 * atom's share of its pairs' shifted energies (force_lj_neigh_impl.h:
 * 256-343: one half per full-list pair; half lists give a local j's pair
 * wholly to the row atom and half to a ghost j's) */
emd_status emd_lj_energy_atoms(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                               const emd_lj_params *p, int half, double *eatom);
/* No counterpart in the reference.  emd_lj_energy for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same energy (to rounding: the block sums are added
 * on the device in double-double instead of on the host) with all scratch in
 * a static 4.01 MiB device array per GPU; no hipMalloc, hipFree or host
 * allocation.  At most 4,194,304 local atoms (EMD_ERR_ARG beyond).  Calls on
 * one GPU share that array: calls that overlap on one GPU race. */
emd_status emd_lj_energy_noalloc(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                 const emd_lj_params *p, int half, double *pe);
/* lanes per atom of the LJ force kernel: 1, 2, 4, 8 or 16 (0 = built-in
 * default, 16).  One lane sums an atom's pairs in the reference's serial
 * order (FP64: bitwise the reference's force); more lanes sum interleaved
 * parts.  Every count keeps the mixed path's force error within 1e-5 RMS. */
emd_status emd_lj_set_lanes(emd_ctx *ctx, int precision, int lanes);

/* ---------------------------------------------------------------- NVE */
/* IntegratorNVE::initial_integrate / final_integrate,
 * src/integrator_nve.cpp:76-121 (reads f32 when precision == EMD_MIXED) */
emd_status emd_nve_initial(emd_ctx *ctx, emd_system *s, double dtf, double dtv,
                           int precision);
emd_status emd_nve_final(emd_ctx *ctx, emd_system *s, double dtf, int precision);
/* final_integrate of step s then initial_integrate of step s+1 (the
 * reference's ExaMiniMD::run calls them back to back when no thermo, dump or
 * correctness check falls between, src/examinimd.cpp:249 then 196) in one
 * launch; results bitwise equal to the two calls */
emd_status emd_nve_final_initial(emd_ctx *ctx, emd_system *s, double dtf, double dtv,
                                 int precision);
/* The step boundary as the driver runs it when nothing observes it:
 * comm->update_force() of step s if `fold` (src/examinimd.cpp:243,
 * comm_serial.h:200-213), final_integrate of s and initial_integrate of s+1
 * (src/examinimd.cpp:249 then 196), and comm->update_halo() of s+1 if
 * `update_halo` (src/examinimd.cpp:226, comm_serial.h:184-197).  Up to 16384
 * local atoms and 16384 ghosts the three run in one workgroup launch
 * separated by barriers; larger systems get the three separate launches.
 * Results are bitwise those of emd_comm_update_force, emd_nve_final_initial
 * and emd_comm_update_halo called in turn. */
emd_status emd_nve_step_boundary(emd_ctx *ctx, emd_system *s, const emd_halo *h, int fold,
                                 int update_halo, double dtf, double dtv, int precision);

/* Temperature / KinE reductions, src/property_temperature.cpp:43-62,
 * src/property_kine.cpp:43-61: returns sum_i m_i v_i^2 (compensated) */
emd_status emd_reduce_mv2(emd_ctx *ctx, const emd_system *s, double *mv2);

/* --------------------------------------------------------------- SNAP */
typedef struct {
  int twojmax;           /* 0..12 (the reference's sna_impl.hpp:25-53 has no cap) */
  double rfac0, rmin0, rcutfac;
  int switchflag, bzeroflag;
  double radelem, wjelem;
  int ncoeff;            /* bispectrum components (55 for twojmax 8, 140 for 12) */
  const double *coeff;   /* HOST pointer: beta_0 .. beta_ncoeff */
} emd_snap_params;

typedef struct emd_snap emd_snap;
/* ForceSNAP::init_coeff + SNA::init (src/force_types/force_snap_neigh_impl.h:
 * 258-338, sna_impl.hpp:139-143) */
emd_status emd_snap_create(emd_ctx *ctx, const emd_snap_params *p, emd_snap **out);
emd_status emd_snap_destroy(emd_ctx *ctx, emd_snap *sn);
emd_status emd_snap_set_coeff(emd_ctx *ctx, emd_snap *sn, const double *coeff);
/* No counterpart in the reference (it serves repeated force and beta-adjoint
 * evaluations of fixed configurations, e.g. force matching): with on != 0
 * the caller promises that positions, atom count and neighbour list stay as
 * they are; for twojmax <= 8 the compaction, neighbour parameters and U_tot
 * of the last force or adjoint call of a precision are then reused by later
 * calls of that precision (they depend on the geometry only, not on beta).
 * Calling it again (on or off) discards what was kept. */
emd_status emd_snap_static_geometry(emd_ctx *ctx, emd_snap *sn, int on);
/* ForceSNAP::compute, force_snap_neigh_impl.h:159-208 (+ operator() 589-725);
 * forces land on locals and ghosts, fold with emd_comm_update_force */
emd_status emd_snap_force(emd_ctx *ctx, emd_system *s, const emd_neigh *nl, emd_snap *sn,
                          int precision);
/* deep_copy(f, 0) (src/examinimd.cpp:232) then ForceSNAP::compute: the same
 * forces, with the zeroing of every local and ghost row (f, or f32 for
 * EMD_MIXED) done by the first SNAP kernel instead of a launch of its own */
emd_status emd_snap_force_zeroed(emd_ctx *ctx, emd_system *s, const emd_neigh *nl, emd_snap *sn,
                                 int precision);
/* SNAP energy (no reference counterpart: the reference's ForceSNAP has no
 * compute_energy, src/force.h:54); eatom may be NULL or a device [n_local] */
emd_status emd_snap_energy(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                           emd_snap *sn, double *energy, double *eatom);

/* ------------------------------------------- force-matching gradients */
/* No counterpart in the reference (it has no parameter gradients).  Given
 * the cotangent dLdF (DEVICE, double[n_local][3], dL/dF of the local atoms
 * after the ghost fold) and dLdE = dL/dE of the total energy, these return
 *   sum_i dLdF_i . dF_i/dtheta + dLdE . dE/dtheta
 * on the HOST.  precision EMD_MIXED: FP32 pair/recursion arithmetic with
 * FP64 accumulation; EMD_FP64 throughout.  Positions and neighbour list are
 * those of the last force evaluation. */
/* SNAP: theta = beta_0..beta_ncoeff (dLdbeta has ncoeff+1 entries); the halo
 * maps ghost neighbours to their owners (may be NULL without ghosts) */
emd_status emd_snap_adjoint_beta(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                 const emd_halo *h, emd_snap *sn, const double *dLdF,
                                 double dLdE, int precision, double *dLdbeta);
/* No counterpart in the reference: its ForceSNAP forms no energy and never
 * the bispectrum itself (src/force.h:54; bvec, force_snap_neigh.h:95-96, is
 * declared only).  Per local atom, the ncoeff components
 * B_k = sum' Re(conj(U_J) Z_{j1 j2 J}) whose derivative SNA::compute_dbidrj
 * forms (src/force_types/sna_impl.hpp:329-539), bzeroflag applied, FP64, into
 * B (DEVICE, double[n_local][ncoeff]).  The energy of atom i is
 * beta_0 + sum_k beta_k B_k(i) plus the pair repulsion, so for static
 * configurations the energy and its beta-gradient are products with this
 * matrix. */
emd_status emd_snap_descriptors(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                emd_snap *sn, double *B);
/* LJ (one type, full list): theta = (eps, sigma); out[2] */
emd_status emd_lj_adjoint(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl, double eps,
                          double sigma, double cut, const double *dLdF, double dLdE,
                          int precision, double *out);
/* the same with one energy cotangent per atom, dLdE_atom (DEVICE,
 * double[n_local]): sum_i dLdF_i . dF_i/dtheta + sum_i dLdE_atom[i]
 * dE_i/dtheta, E_i the atom's half of its pair energies.  Configurations
 * merged into one system take their whole gradient in one call. */
emd_status emd_lj_adjoint_atoms(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                double eps, double sigma, double cut, const double *dLdF,
                                const double *dLdE_atom, int precision, double *out);

/* -------------------------------------------------- batched replicas */
/* No counterpart in the reference: many independent periodic LJ systems
 * ("replicas", uneven sizes) advanced together.  Per replica the step is
 * the reference's serial-comm LJ step (src/examinimd.cpp:176-294:
 * IntegratorNVE initial/final, CommSerial exchange + periodic images,
 * full CSR neighbour list rebuilt every rebuild_every steps, ForceLJNeigh
 * full-list force with strict rsq < rc^2 and shifted energy).  Periodic
 * images are carried as 2-bit-per-axis codes in the neighbour entries
 * instead of ghost atoms; an image position is x_j +/- L per axis, as a
 * ghost's is.  Atoms of all replicas live in one flat array (replica r
 * owns [start_r, start_r + natoms_r)), so every kernel is one launch for
 * the whole batch; a step is one fused kernel (force, final kick, next
 * half-kick and drift into a second position buffer), and the rebuild
 * period is captured once as a HIP graph and replayed.  Single atom type
 * per replica, one shared 12-6 potential (eps, sigma, cut). */
typedef struct emd_batch emd_batch;

typedef struct {
  int nrep;
  const int *natoms;    /* [nrep] atoms per replica (host) */
  const double *box;    /* [nrep][3] periodic box lengths, origin 0 (host) */
  const double *mass;   /* [nrep] atom mass per replica (host) */
  double eps, sigma, cut;
  double skin;          /* neighbour cutoff = cut + skin */
  double dt;
  int rebuild_every;    /* rebuild when step % rebuild_every == 0 */
  int precision;        /* EMD_FP64 or EMD_MIXED (FP32 pair math and forces) */
} emd_batch_desc;

/* every box length must be >= cut + skin; total atoms < 2^26 */
emd_status emd_batch_create(emd_ctx *ctx, const emd_batch_desc *d, emd_batch **out);
emd_status emd_batch_destroy(emd_ctx *ctx, emd_batch *b);
/* host arrays [total][3], replica-major; resets the step counter to 0 */
emd_status emd_batch_upload(emd_ctx *ctx, emd_batch *b, const double *x, const double *v);
/* any pointer may be NULL; f is the last force (converted from FP32 in mixed mode) */
emd_status emd_batch_download(emd_ctx *ctx, emd_batch *b, double *x, double *v, double *f);
/* wrap + neighbour build + force at the current positions (the reference's
 * set-up force before step 1) */
emd_status emd_batch_setup(emd_ctx *ctx, emd_batch *b);
/* advance nsteps; use_graph != 0 replays each rebuild period's steps as a
 * captured graph.  Every rebuild checks its longest row at once (one host
 * wait per rebuild): rows that outgrew their capacity are grown and the
 * rebuild repeated, as Neighbor2D does (src/neighbor_types/
 * neighbor_2d.h:304-330), so no step runs on truncated rows */
emd_status emd_batch_run(emd_ctx *ctx, emd_batch *b, int nsteps, int use_graph);
/* per replica: ke[r] = 0.5 sum m v^2, pe[r] = shifted LJ energy (FP64,
 * compensated); either may be NULL */
emd_status emd_batch_thermo(emd_ctx *ctx, emd_batch *b, double *ke, double *pe);
/* force-kernel tile: lanes per atom (1,2,4,8,16) and threads per block
 * (64..1024, multiple of 64) */
emd_status emd_batch_set_tile(emd_ctx *ctx, emd_batch *b, int lanes, int block);
/* times every candidate tile over trial_steps steps (state restored
 * afterwards) and keeps the fastest; outputs may be NULL */
emd_status emd_batch_autotune(emd_ctx *ctx, emd_batch *b, int trial_steps, int *lanes,
                              int *block, double *ms_per_step);
/* step counter, largest row so far, row capacity */
emd_status emd_batch_info(emd_ctx *ctx, emd_batch *b, long *step, int *max_row, int *row_cap);
/* neighbour rows: num[total], entries[total][row_cap] as (j | code << 26),
 * code bits 2d..2d+1 = image of axis d (0 none, 1 -L, 2 +L); host arrays */
emd_status emd_batch_neigh(emd_ctx *ctx, emd_batch *b, int *num, unsigned *entries);

#ifdef __cplusplus
}
#endif
#endif
