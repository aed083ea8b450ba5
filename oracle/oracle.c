/*
 * oracle.c -- CPU restatement of ExaMiniMD's LJ hot path, lattice set-up,
 * serial ghost exchange, binning, neighbour lists, NVE, thermo and dumps.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function names the
 * reference routine it follows (paths relative to the ExaMiniMD tree).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static void *xrealloc(void *p, size_t n) {
  void *q = realloc(p, n ? n : 1);
  if (!q) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return q;
}

/* ======================================================================
 * RNG: LAMMPS RanPark + "loop geom" reseeding
 * src/input.h:60-133 (uniform 72-79, gaussian 84-94, reset 97-132)
 * ==================================================================== */
double or_rng_uniform(int *seed) {
  /* Park-Miller minimal standard with Schrage's factorisation */
  const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
  int k = *seed / IQ;
  *seed = IA * (*seed - k * IQ) - IR * k;
  if (*seed < 0) *seed += IM;
  return (1.0 / IM) * *seed;
}

double or_rng_gaussian(int *seed) {
  /* Marsaglia polar method, src/input.h:84-94 */
  double v1, v2, rsq;
  do {
    v1 = 2.0 * or_rng_uniform(seed) - 1.0;
    v2 = 2.0 * or_rng_uniform(seed) - 1.0;
    rsq = v1 * v1 + v2 * v2;
  } while ((rsq >= 1.0) || (rsq == 0.0));
  const double fac = sqrt(-2.0 * log(rsq) / rsq);
  return v2 * fac;
}

/* Jenkins one-at-a-time over signed bytes (the reference hashes `char`,
 * which is signed on x86-64), without the final avalanche. */
unsigned or_hash_bytes(const signed char *s, int n, unsigned hash) {
  for (int i = 0; i < n; i++) {
    hash += (unsigned)(int)s[i];
    hash += (hash << 10);
    hash ^= (hash >> 6);
  }
  return hash;
}

int or_rng_reset(int ibase, const double coord[3]) {
  unsigned hash = 0;
  hash = or_hash_bytes((const signed char *)&ibase, (int)sizeof(int), hash);
  hash = or_hash_bytes((const signed char *)coord, 3 * (int)sizeof(double), hash);
  hash += (hash << 3);
  hash ^= (hash >> 11);
  hash += (hash << 15);
  int seed = (int)(hash & 0x7ffffff);  /* 27 bits, src/input.h:126 */
  if (!seed) seed = 1;
  for (int i = 0; i < 5; i++) or_rng_uniform(&seed);
  return seed;
}

/* ======================================================================
 * System: src/system.cpp:43-109
 * ==================================================================== */
void or_system_init(or_system *s) {
  memset(s, 0, sizeof(*s));
  s->ntypes = 1;
}

void or_system_free(or_system *s) {
  free(s->x); free(s->v); free(s->f);
  free(s->type); free(s->id); free(s->q);
  s->x = s->v = s->f = s->q = NULL;
  s->type = s->id = NULL;
  s->N_max = 0;
}

void or_system_grow(or_system *s, int n) {
  if (n <= s->N_max) return;
  int old = s->N_max;
  s->x = (double *)xrealloc(s->x, sizeof(double) * 3 * (size_t)n);
  s->v = (double *)xrealloc(s->v, sizeof(double) * 3 * (size_t)n);
  s->f = (double *)xrealloc(s->f, sizeof(double) * 3 * (size_t)n);
  s->type = (int *)xrealloc(s->type, sizeof(int) * (size_t)n);
  s->id = (int *)xrealloc(s->id, sizeof(int) * (size_t)n);
  s->q = (double *)xrealloc(s->q, sizeof(double) * (size_t)n);
  /* Kokkos::resize value-initialises the new tail */
  memset(s->x + 3 * (size_t)old, 0, sizeof(double) * 3 * (size_t)(n - old));
  memset(s->v + 3 * (size_t)old, 0, sizeof(double) * 3 * (size_t)(n - old));
  memset(s->f + 3 * (size_t)old, 0, sizeof(double) * 3 * (size_t)(n - old));
  memset(s->type + old, 0, sizeof(int) * (size_t)(n - old));
  memset(s->id + old, 0, sizeof(int) * (size_t)(n - old));
  memset(s->q + old, 0, sizeof(double) * (size_t)(n - old));
  s->N_max = n;
}

/* units: src/input.cpp:273-304 */
void or_units(or_system *s, int units) {
  if (units == OR_UNITS_METAL) {
    s->boltz = 8.617343e-5; s->mvv2e = 1.0364269e-4; s->dt = 0.001;
  } else if (units == OR_UNITS_REAL) {
    s->boltz = 0.0019872067; s->mvv2e = 48.88821291 * 48.88821291; s->dt = 1.0;
  } else {
    s->boltz = 1.0; s->mvv2e = 1.0; s->dt = 0.005;
  }
}

/* Temperature::compute, src/property_temperature.cpp:43-62 (serial order) */
double or_temperature(const or_system *s) {
  double T = 0.0;
  for (int i = 0; i < s->N_local; i++) {
    const double *v = s->v + 3 * (size_t)i;
    T += (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * s->mass[s->type[i]];
  }
  int dof = 3 * s->N - 3;
  double factor = s->mvv2e / (1.0 * dof * s->boltz);
  return T * factor;
}

/* KinE::compute, src/property_kine.cpp:43-61 */
double or_kinetic_energy(const or_system *s) {
  double KE = 0.0;
  for (int i = 0; i < s->N_local; i++) {
    const double *v = s->v + 3 * (size_t)i;
    KE += (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * s->mass[s->type[i]];
  }
  return KE * (0.5 * s->mvv2e);
}

/* ======================================================================
 * Lattice + velocities: Input::create_lattice, src/input.cpp:460-792.
 * Comm::create_domain_decomposition (src/comm.cpp:51-58) makes the
 * sub-domain the whole box for one process.  BCC is an extension with
 * the basis (0,0,0),(1/2,1/2,1/2) and the SC/FCC loop structure.
 * ==================================================================== */
int or_create_lattice(or_system *s, int style, double arg, const double off[3],
                      int nx, int ny, int nz, double T_target, int seed,
                      int gaussian) {
  double a;
  if (style == OR_LATTICE_FCC) a = pow((4.0 / arg), (1.0 / 3.0));
  else a = arg;
  s->domain[0] = a * nx; s->domain[1] = a * ny; s->domain[2] = a * nz;
  for (int d = 0; d < 3; d++) { s->lo[d] = 0.0; s->hi[d] = s->domain[d]; }

  double basis[4][3];
  int nb;
  if (style == OR_LATTICE_SC) {
    nb = 1;
    basis[0][0] = basis[0][1] = basis[0][2] = 0.0;
  } else if (style == OR_LATTICE_FCC) {
    nb = 4;
    double b[4][3] = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
    memcpy(basis, b, sizeof(b));
  } else {
    nb = 2;
    double b[2][3] = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.5}};
    memcpy(basis, b, sizeof(b));
  }
  /* FCC/BCC add the offset to the basis (input.cpp:605-609); SC adds it
   * to the integer index (input.cpp:350-354) -- identical values. */
  int ix_start = (int)(s->lo[0] / s->domain[0] * nx - 0.5);
  int iy_start = (int)(s->lo[1] / s->domain[1] * ny - 0.5);
  int iz_start = (int)(s->lo[2] / s->domain[2] * nz - 0.5);
  int ix_end = (int)(s->hi[0] / s->domain[0] * nx + 0.5);
  int iy_end = (int)(s->hi[1] / s->domain[1] * ny + 0.5);
  int iz_end = (int)(s->hi[2] / s->domain[2] * nz + 0.5);

  for (int pass = 0; pass < 2; pass++) {
    int n = 0;
    for (int iz = iz_start; iz <= iz_end; iz++)
      for (int iy = iy_start; iy <= iy_end; iy++)
        for (int ix = ix_start; ix <= ix_end; ix++)
          for (int k = 0; k < nb; k++) {
            double xt, yt, zt;
            if (style == OR_LATTICE_SC) {
              zt = a * (iz + off[2]); yt = a * (iy + off[1]); xt = a * (ix + off[0]);
            } else {
              xt = a * (1.0 * ix + (basis[k][0] + off[0]));
              yt = a * (1.0 * iy + (basis[k][1] + off[1]));
              zt = a * (1.0 * iz + (basis[k][2] + off[2]));
            }
            if (xt >= s->lo[0] && yt >= s->lo[1] && zt >= s->lo[2] &&
                xt < s->hi[0] && yt < s->hi[1] && zt < s->hi[2]) {
              if (pass == 1) {
                s->x[3 * (size_t)n + 0] = xt;
                s->x[3 * (size_t)n + 1] = yt;
                s->x[3 * (size_t)n + 2] = zt;
                s->type[n] = rand() % s->ntypes;
                s->id[n] = n + 1;
              }
              n++;
            }
          }
    if (pass == 0) {
      s->N_local = n; s->N = n; s->N_ghost = 0;
      or_system_grow(s, n);
    }
  }

  /* velocities: src/input.cpp:731-791 */
  double total_mass = 0.0, px = 0.0, py = 0.0, pz = 0.0;
  for (int i = 0; i < s->N_local; i++) {
    double xx[3] = {s->x[3 * (size_t)i], s->x[3 * (size_t)i + 1], s->x[3 * (size_t)i + 2]};
    int rs = or_rng_reset(seed, xx);
    double mass_i = s->mass[s->type[i]];
    double vx, vy, vz;
    if (!gaussian) {
      vx = or_rng_uniform(&rs) - 0.5;
      vy = or_rng_uniform(&rs) - 0.5;
      vz = or_rng_uniform(&rs) - 0.5;
    } else {
      vx = or_rng_gaussian(&rs);
      vy = or_rng_gaussian(&rs);
      vz = or_rng_gaussian(&rs);
    }
    double *v = s->v + 3 * (size_t)i;
    v[0] = vx / sqrt(mass_i);
    v[1] = vy / sqrt(mass_i);
    v[2] = vz / sqrt(mass_i);
    s->q[i] = 0.0;
    total_mass += mass_i;
    px += mass_i * v[0];
    py += mass_i * v[1];
    pz += mass_i * v[2];
  }
  double svx = px / total_mass, svy = py / total_mass, svz = pz / total_mass;
  for (int i = 0; i < s->N_local; i++) {
    double *v = s->v + 3 * (size_t)i;
    v[0] -= svx; v[1] -= svy; v[2] -= svz;
  }
  double T = or_temperature(s);
  double scale = sqrt(T_target / T);
  for (int i = 0; i < s->N_local; i++) {
    double *v = s->v + 3 * (size_t)i;
    v[0] *= scale; v[1] *= scale; v[2] *= scale;
  }
  return s->N_local;
}

/* ======================================================================
 * CommSerial: src/comm_types/comm_serial.{h:95-213,cpp:41-127}
 * Ghost order within a phase is ascending source index (Serial order of
 * the atomic pack counter).
 * ==================================================================== */
void or_comm_exchange(or_system *s) {
  for (int i = 0; i < s->N_local; i++) {
    double *x = s->x + 3 * (size_t)i;
    for (int d = 0; d < 3; d++) {
      const double xd = x[d];
      if (xd > s->domain[d]) x[d] -= s->domain[d];
      if (xd < 0) x[d] += s->domain[d];
    }
  }
}

static void copy_particle(or_system *s, int dst, int src, int dim, double shift) {
  for (int d = 0; d < 3; d++) {
    s->x[3 * (size_t)dst + d] = s->x[3 * (size_t)src + d];
    s->v[3 * (size_t)dst + d] = s->v[3 * (size_t)src + d];
  }
  s->x[3 * (size_t)dst + dim] += shift;
  s->q[dst] = s->q[src];
  s->id[dst] = s->id[src];
  s->type[dst] = s->type[src];
}

void or_comm_exchange_halo(or_system *s, or_halo *h, double comm_depth) {
  int N_local = s->N_local, N_ghost = 0;
  for (int phase = 0; phase < 6; phase++) {
    int dim = phase / 2;
    int nparticles = N_local + N_ghost - ((phase % 2 == 1) ? h->num_ghost[phase - 1] : 0);
    int count = 0;
    for (int i = 0; i < nparticles; i++) {
      double xd = s->x[3 * (size_t)i + dim];
      int take = (phase % 2 == 0) ? (xd >= s->hi[dim] - comm_depth)
                                  : (xd <= s->lo[dim] + comm_depth);
      if (take) count++;
    }
    or_system_grow(s, N_local + N_ghost + count);
    if (count > h->cap[phase]) {
      h->pack[phase] = (int *)xrealloc(h->pack[phase], sizeof(int) * (size_t)count);
      h->cap[phase] = count;
    }
    int k = 0;
    double shift = (phase % 2 == 0) ? -s->domain[dim] : s->domain[dim];
    for (int i = 0; i < nparticles; i++) {
      double xd = s->x[3 * (size_t)i + dim];
      int take = (phase % 2 == 0) ? (xd >= s->hi[dim] - comm_depth)
                                  : (xd <= s->lo[dim] + comm_depth);
      if (take) {
        h->pack[phase][k] = i;
        copy_particle(s, N_local + N_ghost + k, i, dim, shift);
        k++;
      }
    }
    h->num_ghost[phase] = count;
    N_ghost += count;
  }
  s->N_ghost = N_ghost;
}

void or_comm_update_halo(or_system *s, const or_halo *h) {
  int N_ghost = 0;
  for (int phase = 0; phase < 6; phase++) {
    int dim = phase / 2;
    double shift = (phase % 2 == 0) ? -s->domain[dim] : s->domain[dim];
    for (int i = 0; i < h->num_ghost[phase]; i++)
      copy_particle(s, s->N_local + N_ghost + i, h->pack[phase][i], dim, shift);
    N_ghost += h->num_ghost[phase];
  }
}

void or_comm_update_force(or_system *s, const or_halo *h) {
  int off[6];
  off[0] = s->N_local;
  for (int p = 1; p < 6; p++) off[p] = off[p - 1] + h->num_ghost[p - 1];
  for (int p = 5; p >= 0; p--)
    for (int ii = 0; ii < h->num_ghost[p]; ii++) {
      int i = h->pack[p][ii];
      const double *fg = s->f + 3 * (size_t)(off[p] + ii);
      double fx = fg[0], fy = fg[1], fz = fg[2];
      s->f[3 * (size_t)i + 0] += fx;
      s->f[3 * (size_t)i + 1] += fy;
      s->f[3 * (size_t)i + 2] += fz;
    }
}

void or_halo_free(or_halo *h) {
  for (int p = 0; p < 6; p++) { free(h->pack[p]); h->pack[p] = NULL; h->cap[p] = 0; }
}

/* ======================================================================
 * BinningKKSort::create_binning, src/binning_types/binning_kksort.cpp:71-140
 * with Kokkos BinOp3D / BinSort (third party, see oracle.h).
 * ==================================================================== */
void or_binning(or_system *s, or_bins *b, double dx_in, double dy_in, double dz_in,
                int halo, int do_local, int do_ghost, int sort) {
  if (!(do_local || do_ghost)) return;
  b->nhalo = halo;
  int r0 = do_local ? 0 : s->N_local;
  int r1 = do_ghost ? s->N_local + s->N_ghost : s->N_local;
  double sub[3] = {s->hi[0] - s->lo[0], s->hi[1] - s->lo[1], s->hi[2] - s->lo[2]};
  double din[3] = {dx_in, dy_in, dz_in};
  int nb[3];
  double dd[3];
  for (int d = 0; d < 3; d++) {
    nb[d] = (int)(sub[d] / din[d]);
    if (nb[d] == 0) nb[d] = 1;
    dd[d] = sub[d] / nb[d];
  }
  for (int d = 0; d < 3; d++) nb[d] += 2 * halo;
  double eps = dd[0] / 1000;  /* binning_kksort.cpp:92 uses dx for all dims */
  for (int d = 0; d < 3; d++) {
    b->min[d] = -dd[d] * halo - eps + s->lo[d];
    b->max[d] = dd[d] * halo + eps + s->hi[d];
    b->nbin[d] = nb[d];
  }
  int nbt = nb[0] * nb[1] * nb[2];
  b->nbins_total = nbt;
  b->n_range = r1 - r0;
  b->count = (int *)xrealloc(b->count, sizeof(int) * (size_t)nbt);
  b->offset = (int *)xrealloc(b->offset, sizeof(int) * (size_t)nbt);
  b->permute = (int *)xrealloc(b->permute, sizeof(int) * (size_t)(b->n_range));
  memset(b->count, 0, sizeof(int) * (size_t)nbt);

  double mul[3];
  for (int d = 0; d < 3; d++) mul[d] = 1.0 * nb[d] / (b->max[d] - b->min[d]);
  int n = b->n_range;
  int *bin = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
  for (int i = 0; i < n; i++) {
    const double *x = s->x + 3 * (size_t)(r0 + i);
    int bx = (int)(mul[0] * (x[0] - b->min[0]));
    int by = (int)(mul[1] * (x[1] - b->min[1]));
    int bz = (int)(mul[2] * (x[2] - b->min[2]));
    bin[i] = (bx * nb[1] + by) * nb[2] + bz;
    if (bin[i] < 0 || bin[i] >= nbt) {
      fprintf(stderr, "oracle: atom %d outside the bin grid\n", r0 + i);
      abort();
    }
    b->count[bin[i]]++;
  }
  int acc = 0;
  for (int k = 0; k < nbt; k++) { b->offset[k] = acc; acc += b->count[k]; }
  int *fill = (int *)calloc((size_t)nbt, sizeof(int));
  for (int i = 0; i < n; i++) {
    int c = fill[bin[i]]++;
    b->permute[b->offset[bin[i]] + c] = i;
  }
  free(fill);
  free(bin);

  if (sort) {
    /* BinSort::sort: new[i] = old[permute[i]] over the range, for x,v,f,type,id,q */
    double *t3 = (double *)malloc(sizeof(double) * 3 * (size_t)(n ? n : 1));
    double *arrs[3] = {s->x, s->v, s->f};
    for (int a = 0; a < 3; a++) {
      for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++)
          t3[3 * (size_t)i + d] = arrs[a][3 * (size_t)(r0 + b->permute[i]) + d];
      memcpy(arrs[a] + 3 * (size_t)r0, t3, sizeof(double) * 3 * (size_t)n);
    }
    free(t3);
    int *ti = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; i++) ti[i] = s->type[r0 + b->permute[i]];
    memcpy(s->type + r0, ti, sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) ti[i] = s->id[r0 + b->permute[i]];
    memcpy(s->id + r0, ti, sizeof(int) * (size_t)n);
    free(ti);
    double *tq = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
    for (int i = 0; i < n; i++) tq[i] = s->q[r0 + b->permute[i]];
    memcpy(s->q + r0, tq, sizeof(double) * (size_t)n);
    free(tq);
  }
}

void or_bins_free(or_bins *b) {
  free(b->count); free(b->offset); free(b->permute);
  b->count = b->offset = b->permute = NULL;
}

/* ======================================================================
 * NeighborCSR, src/neighbor_types/neighbor_csr.h:176-435 (full 176-255,
 * half 258-358) and Neighbor2D, neighbor_2d.h:174-331.  Within a row the
 * order is stencil-bin order (bx,by,bz nested, -1..+1) then permutation
 * order, i.e. the Serial order of the atomic row counter.
 * ==================================================================== */
typedef int (*emit_fn)(void *ctx, int i, int k, int j);

static int half_keep(const or_system *s, int i, int j, int newton) {
  const double *xi = s->x + 3 * (size_t)i, *xj = s->x + 3 * (size_t)j;
  const double x_i = xi[0], y_i = xi[1], z_i = xi[2];
  const double x_j = xj[0], y_j = xj[1], z_j = xj[2];
  if (((j == i) || (j < s->N_local || newton)) &&
      !((x_j > x_i) || ((x_j == x_i) && ((y_j > y_i) || ((y_j == y_i) && (z_j > z_i))))))
    return 0;
  return 1;
}

/* visits the neighbours of local atom i in reference order; returns count */
static int visit_row(const or_system *s, const or_bins *b, int bx, int by, int bz,
                     int i, double cut, int half, int newton, int *out) {
  const int ny = b->nbin[1], nzb = b->nbin[2];
  const double *xi = s->x + 3 * (size_t)i;
  const double x_i = xi[0], y_i = xi[1], z_i = xi[2];
  int n = 0;
  for (int bxj = bx - 1; bxj < bx + 2; bxj++)
    for (int byj = by - 1; byj < by + 2; byj++)
      for (int bzj = bz - 1; bzj < bz + 2; bzj++) {
        int bin = (bxj * ny + byj) * nzb + bzj;
        int off = b->offset[bin], cnt = b->count[bin];
        for (int bj = 0; bj < cnt; bj++) {
          int j = b->permute[off + bj];
          if (half) {
            if (!half_keep(s, i, j, newton)) continue;
          }
          const double *xj = s->x + 3 * (size_t)j;
          const double dx = x_i - xj[0];
          const double dy = y_i - xj[1];
          const double dz = z_i - xj[2];
          const double rsq = dx * dx + dy * dy + dz * dz;
          if (half) {
            if (rsq <= cut * cut) { if (out) out[n] = j; n++; }
          } else {
            if ((rsq <= cut * cut) && (i != j)) { if (out) out[n] = j; n++; }
          }
        }
      }
  return n;
}

static void interior_bin(const or_bins *b, int lr, int *bx, int *by, int *bz) {
  int nh = b->nhalo;
  int nbx = b->nbin[0] - 2 * nh, nby = b->nbin[1] - 2 * nh, nbz = b->nbin[2] - 2 * nh;
  (void)nbx;
  *bx = lr / (nby * nbz) + nh;
  *by = (lr / nbz) % nby + nh;
  *bz = lr % nbz + nh;
}

static int n_interior(const or_bins *b) {
  int nh = b->nhalo;
  return (b->nbin[0] - 2 * nh) * (b->nbin[1] - 2 * nh) * (b->nbin[2] - 2 * nh);
}

void or_neigh_csr(const or_system *s, const or_bins *b, double cut, int half,
                  int newton, or_csr *c) {
  int N = s->N_local;
  c->n = N;
  c->row = (int *)xrealloc(c->row, sizeof(int) * (size_t)(N + 1));
  int *cnt = (int *)calloc((size_t)(N + 1), sizeof(int));
  int nib = n_interior(b);
#pragma omp parallel for schedule(dynamic, 16)
  for (int lr = 0; lr < nib; lr++) {
    int bx, by, bz;
    interior_bin(b, lr, &bx, &by, &bz);
    int bin = (bx * b->nbin[1] + by) * b->nbin[2] + bz;
    for (int bi = 0; bi < b->count[bin]; bi++) {
      int i = b->permute[b->offset[bin] + bi];
      if (i >= N) continue;
      cnt[i] = visit_row(s, b, bx, by, bz, i, cut, half, newton, NULL);
    }
  }
  int acc = 0;
  for (int i = 0; i < N; i++) { c->row[i] = acc; acc += cnt[i]; }
  c->row[N] = acc;
  c->total = acc;
  c->entries = (int *)xrealloc(c->entries, sizeof(int) * (size_t)(acc ? acc : 1));
#pragma omp parallel for schedule(dynamic, 16)
  for (int lr = 0; lr < nib; lr++) {
    int bx, by, bz;
    interior_bin(b, lr, &bx, &by, &bz);
    int bin = (bx * b->nbin[1] + by) * b->nbin[2] + bz;
    for (int bi = 0; bi < b->count[bin]; bi++) {
      int i = b->permute[b->offset[bin] + bi];
      if (i >= N) continue;
      visit_row(s, b, bx, by, bz, i, cut, half, newton, c->entries + c->row[i]);
    }
  }
  free(cnt);
}

void or_neigh_2d(const or_system *s, const or_bins *b, double cut, int half,
                 int newton, or_list2d *l) {
  /* same rows as CSR, padded to maxneighs (grown x1.2 on overflow,
   * neighbor_2d.h:304-330; the first list starts at 16, neighbor_2d.h:103) */
  or_csr c = {0};
  or_neigh_csr(s, b, cut, half, newton, &c);
  int N = s->N_local;
  if (l->maxneighs <= 0) l->maxneighs = 16;
  int mx = 0;
  for (int i = 0; i < N; i++) {
    int n = c.row[i + 1] - c.row[i];
    if (n > mx) mx = n;
  }
  while (mx > l->maxneighs) l->maxneighs = (int)(mx * 1.2);
  l->n = N;
  l->num = (int *)xrealloc(l->num, sizeof(int) * (size_t)(N + 1));
  l->neighs = (int *)xrealloc(l->neighs, sizeof(int) * (size_t)(N + 1) * (size_t)l->maxneighs);
  for (int i = 0; i < N; i++) {
    int n = c.row[i + 1] - c.row[i];
    l->num[i] = n;
    memcpy(l->neighs + (size_t)i * l->maxneighs, c.entries + c.row[i], sizeof(int) * (size_t)n);
  }
  or_csr_free(&c);
}

void or_csr_free(or_csr *c) {
  free(c->row); free(c->entries); c->row = c->entries = NULL;
}
void or_list2d_free(or_list2d *l) {
  free(l->num); free(l->neighs); l->num = l->neighs = NULL;
}

/* brute-force O(N*(N+G)) neighbour counts (tiny inputs only) */
void or_brute_neigh_count(const or_system *s, double cut, int *counts) {
  int NT = s->N_local + s->N_ghost;
  for (int i = 0; i < s->N_local; i++) {
    int n = 0;
    for (int j = 0; j < NT; j++) {
      if (j == i) continue;
      double dx = s->x[3 * (size_t)i] - s->x[3 * (size_t)j];
      double dy = s->x[3 * (size_t)i + 1] - s->x[3 * (size_t)j + 1];
      double dz = s->x[3 * (size_t)i + 2] - s->x[3 * (size_t)j + 2];
      if (dx * dx + dy * dy + dz * dz <= cut * cut) n++;
    }
    counts[i] = n;
  }
}

/* ======================================================================
 * ForceLJNeigh, src/force_types/force_lj_neigh_impl.h:60-343
 * (coefficients 70-72; full force 161-206; half 208-254; energy 256-343)
 * ==================================================================== */
void or_lj_force(or_system *s, const or_csr *c, double eps, double sigma, double cut,
                 int half) {
  const double lj1 = 48.0 * eps * pow(sigma, 12.0);
  const double lj2 = 24.0 * eps * pow(sigma, 6.0);
  const double cutsq = cut * cut;
  const int N = s->N_local;
  if (!half) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
      const double x_i = s->x[3 * (size_t)i], y_i = s->x[3 * (size_t)i + 1],
                   z_i = s->x[3 * (size_t)i + 2];
      double fxi = 0.0, fyi = 0.0, fzi = 0.0;
      for (int jj = c->row[i]; jj < c->row[i + 1]; jj++) {
        int j = c->entries[jj];
        const double dx = x_i - s->x[3 * (size_t)j];
        const double dy = y_i - s->x[3 * (size_t)j + 1];
        const double dz = z_i - s->x[3 * (size_t)j + 2];
        const double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < cutsq) {
          double r2inv = 1.0 / rsq;
          double r6inv = r2inv * r2inv * r2inv;
          double fpair = (r6inv * (lj1 * r6inv - lj2)) * r2inv;
          fxi += dx * fpair;
          fyi += dy * fpair;
          fzi += dz * fpair;
        }
      }
      s->f[3 * (size_t)i] += fxi;
      s->f[3 * (size_t)i + 1] += fyi;
      s->f[3 * (size_t)i + 2] += fzi;
    }
  } else {
    for (int i = 0; i < N; i++) {
      const double x_i = s->x[3 * (size_t)i], y_i = s->x[3 * (size_t)i + 1],
                   z_i = s->x[3 * (size_t)i + 2];
      double fxi = 0.0, fyi = 0.0, fzi = 0.0;
      for (int jj = c->row[i]; jj < c->row[i + 1]; jj++) {
        int j = c->entries[jj];
        const double dx = x_i - s->x[3 * (size_t)j];
        const double dy = y_i - s->x[3 * (size_t)j + 1];
        const double dz = z_i - s->x[3 * (size_t)j + 2];
        const double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < cutsq) {
          double r2inv = 1.0 / rsq;
          double r6inv = r2inv * r2inv * r2inv;
          double fpair = (r6inv * (lj1 * r6inv - lj2)) * r2inv;
          fxi += dx * fpair;
          fyi += dy * fpair;
          fzi += dz * fpair;
          s->f[3 * (size_t)j] -= dx * fpair;
          s->f[3 * (size_t)j + 1] -= dy * fpair;
          s->f[3 * (size_t)j + 2] -= dz * fpair;
        }
      }
      s->f[3 * (size_t)i] += fxi;
      s->f[3 * (size_t)i + 1] += fyi;
      s->f[3 * (size_t)i + 2] += fzi;
    }
  }
}

/*
 * ForceLJIDialNeigh, src/force_types/force_lj_idial_neigh_impl.h:41-219.
 * Same pair set as lj/cut (strict rsq < cutsq, :143/:192), but the pair
 * force is accumulated `intensity` times from 0 in parts of
 * ((r6inv*(lj1*r6inv - lj2))*r2inv)/intensity (:150-155, :199-204); the
 * half-list form subtracts the reaction only from local j (:208-212), ghost
 * forces are left alone.  intensity is a double in the reference's table
 * (t_fparams, :46) compared against the int loop counter.
 */
void or_lj_idial_force(or_system *s, const or_csr *c, double eps, double sigma, double cut,
                       double intensity, int half) {
  const double lj1 = 48.0 * eps * pow(sigma, 12.0);
  const double lj2 = 24.0 * eps * pow(sigma, 6.0);
  const double cutsq = cut * cut;
  const int N = s->N_local;
  /* rows are independent for a full list; the half list's j updates are
   * kept serial */
#pragma omp parallel for schedule(static) if (!half)
  for (int i = 0; i < N; i++) {
    const double x_i = s->x[3 * (size_t)i], y_i = s->x[3 * (size_t)i + 1],
                 z_i = s->x[3 * (size_t)i + 2];
    double fxi = 0.0, fyi = 0.0, fzi = 0.0;
    for (int jj = c->row[i]; jj < c->row[i + 1]; jj++) {
      int j = c->entries[jj];
      const double dx = x_i - s->x[3 * (size_t)j];
      const double dy = y_i - s->x[3 * (size_t)j + 1];
      const double dz = z_i - s->x[3 * (size_t)j + 2];
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq < cutsq) {
        double fpair = 0;
        for (int repeat = 0; repeat < intensity; repeat++) {
          double r2inv = 1.0 / rsq;
          double r6inv = r2inv * r2inv * r2inv;
          fpair += (r6inv * (lj1 * r6inv - lj2)) * r2inv / intensity;
        }
        fxi += dx * fpair;
        fyi += dy * fpair;
        fzi += dz * fpair;
        if (half && j < N) {
          s->f[3 * (size_t)j] -= dx * fpair;
          s->f[3 * (size_t)j + 1] -= dy * fpair;
          s->f[3 * (size_t)j + 2] -= dz * fpair;
        }
      }
    }
    s->f[3 * (size_t)i] += fxi;
    s->f[3 * (size_t)i + 1] += fyi;
    s->f[3 * (size_t)i + 2] += fzi;
  }
}

/* Neumaier compensated accumulation */
static void neumaier(double *sum, double *c, double v) {
  double t = *sum + v;
  if (fabs(*sum) >= fabs(v)) *c += (*sum - t) + v;
  else *c += (v - t) + *sum;
  *sum = t;
}

double or_lj_energy(const or_system *s, const or_csr *c, double eps, double sigma,
                    double cut, int half) {
  const double lj1 = 48.0 * eps * pow(sigma, 12.0);
  const double lj2 = 24.0 * eps * pow(sigma, 6.0);
  const double cutsq = cut * cut;
  /* the reference's reduction order is left to Kokkos; the oracle sums the
   * pair terms with Neumaier compensation so its total is exact to ~1 ulp */
  double PE = 0.0, PEc = 0.0;
  for (int i = 0; i < s->N_local; i++) {
    const double x_i = s->x[3 * (size_t)i], y_i = s->x[3 * (size_t)i + 1],
                 z_i = s->x[3 * (size_t)i + 2];
    for (int jj = c->row[i]; jj < c->row[i + 1]; jj++) {
      int j = c->entries[jj];
      const double dx = x_i - s->x[3 * (size_t)j];
      const double dy = y_i - s->x[3 * (size_t)j + 1];
      const double dz = z_i - s->x[3 * (size_t)j + 2];
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq < cutsq) {
        double r2inv = 1.0 / rsq;
        double r6inv = r2inv * r2inv * r2inv;
        double r2invc = 1.0 / cutsq;
        double r6invc = r2invc * r2invc * r2invc;
        if (!half) {
          neumaier(&PE, &PEc, 0.5 * r6inv * (0.5 * lj1 * r6inv - lj2) / 6.0);
          neumaier(&PE, &PEc, -(0.5 * r6invc * (0.5 * lj1 * r6invc - lj2) / 6.0));
        } else {
          double fac = (j < s->N_local) ? 1.0 : 0.5;
          neumaier(&PE, &PEc, fac * r6inv * (0.5 * lj1 * r6inv - lj2) / 6.0);
          neumaier(&PE, &PEc, -(fac * r6invc * (0.5 * lj1 * r6invc - lj2) / 6.0));
        }
      }
    }
  }
  return PE + PEc;
}

/* ======================================================================
 * IntegratorNVE, src/integrator_nve.cpp:41-121
 * ==================================================================== */
void or_nve_initial(or_system *s) {
  const double dtv = s->dt, dtf = 0.5 * s->dt / s->mvv2e;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->N_local; i++) {
    const double dtfm = dtf / s->mass[s->type[i]];
    double *v = s->v + 3 * (size_t)i, *x = s->x + 3 * (size_t)i;
    const double *f = s->f + 3 * (size_t)i;
    v[0] += dtfm * f[0]; v[1] += dtfm * f[1]; v[2] += dtfm * f[2];
    x[0] += dtv * v[0]; x[1] += dtv * v[1]; x[2] += dtv * v[2];
  }
}

void or_nve_final(or_system *s) {
  const double dtf = 0.5 * s->dt / s->mvv2e;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->N_local; i++) {
    const double dtfm = dtf / s->mass[s->type[i]];
    double *v = s->v + 3 * (size_t)i;
    const double *f = s->f + 3 * (size_t)i;
    v[0] += dtfm * f[0]; v[1] += dtfm * f[1]; v[2] += dtfm * f[2];
  }
}

/* ======================================================================
 * Binary dump / correctness: src/examinimd.cpp:296-482
 * layout: int n, int id[n], int type[n], double q[n], double x[3n],
 *         double v[3n], double f[3n] (AoS, native endian)
 * ==================================================================== */
int or_dump_write(const or_system *s, const char *path) {
  FILE *fp = fopen(path, "wb");
  if (!fp) return 1;
  int n = s->N_local;
  fwrite(&n, sizeof(int), 1, fp);
  fwrite(s->id, sizeof(int), (size_t)n, fp);
  fwrite(s->type, sizeof(int), (size_t)n, fp);
  fwrite(s->q, sizeof(double), (size_t)n, fp);
  fwrite(s->x, sizeof(double), 3 * (size_t)n, fp);
  fwrite(s->v, sizeof(double), 3 * (size_t)n, fp);
  fwrite(s->f, sizeof(double), 3 * (size_t)n, fp);
  fclose(fp);
  return 0;
}

/* out = {|dr|, max|dr|, |dv|, max|dv|, |df|, max|df|} as check_correctness */
int or_dump_compare(const or_system *s, const char *path, double out[6]) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return 1;
  int n;
  if (fread(&n, sizeof(int), 1, fp) != 1 || n != s->N_local) { fclose(fp); return 2; }
  int *id = malloc(sizeof(int) * (size_t)n), *ty = malloc(sizeof(int) * (size_t)n);
  double *q = malloc(sizeof(double) * (size_t)n);
  double *x = malloc(sizeof(double) * 3 * (size_t)n), *v = malloc(sizeof(double) * 3 * (size_t)n),
         *f = malloc(sizeof(double) * 3 * (size_t)n);
  size_t ok = fread(id, sizeof(int), (size_t)n, fp) + fread(ty, sizeof(int), (size_t)n, fp) +
              fread(q, sizeof(double), (size_t)n, fp) + fread(x, sizeof(double), 3 * (size_t)n, fp) +
              fread(v, sizeof(double), 3 * (size_t)n, fp) + fread(f, sizeof(double), 3 * (size_t)n, fp);
  fclose(fp);
  int rc = (ok == (size_t)n * 12) ? 0 : 3;
  double sr = 0, sv = 0, sf = 0, mr = 0, mv = 0, mf = 0;
  for (int i = 0; i < n && rc == 0; i++) {
    int ii = -1;
    if (s->id[i] != id[i]) {
      for (int j = 0; j < n; j++) if (s->id[j] == id[i]) { ii = j; break; }
    } else ii = i;
    if (ii < 0) { rc = 4; break; }
    const double *a[3] = {s->x, s->v, s->f}, *r[3] = {x, v, f};
    double *sum[3] = {&sr, &sv, &sf}, *mx[3] = {&mr, &mv, &mf};
    for (int k = 0; k < 3; k++) {
      double d2 = 0;
      for (int d = 0; d < 3; d++) {
        double del = a[k][3 * (size_t)ii + d] - r[k][3 * (size_t)i + d];
        d2 += del * del;
        if (fabs(del) > *mx[k]) *mx[k] = fabs(del);
      }
      *sum[k] += d2;
    }
  }
  out[0] = sqrt(sr); out[1] = mr; out[2] = sqrt(sv); out[3] = mv; out[4] = sqrt(sf); out[5] = mf;
  free(id); free(ty); free(q); free(x); free(v); free(f);
  return rc;
}

/* ======================================================================
 * MD driver: ExaMiniMD::init / run, src/examinimd.cpp:60-294
 * ==================================================================== */
#include <time.h>
static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

struct or_md {
  or_config cfg;
  or_snap_params snap;
  or_system s;
  or_halo h;
  or_bins b;
  or_csr c;
  double t[5];
};

or_md *or_md_create(const or_config *cfg, const or_snap_params *snap) {
  or_md *m = (or_md *)calloc(1, sizeof(or_md));
  m->cfg = *cfg;
  if (snap) m->snap = *snap;
  or_system_init(&m->s);
  return m;
}

void or_md_destroy(or_md *m) {
  if (!m) return;
  or_system_free(&m->s);
  or_halo_free(&m->h);
  or_bins_free(&m->b);
  or_csr_free(&m->c);
  free(m);
}

or_system *or_md_system(or_md *m) { return &m->s; }
const or_csr *or_md_csr(or_md *m) { return &m->c; }
const or_halo *or_md_halo(or_md *m) { return &m->h; }
double or_md_time(or_md *m, int which) { return m->t[which]; }

static double md_neigh_cut(const or_md *m) { return m->cfg.force_cutoff + m->cfg.skin; }

static void md_force(or_md *m) {
  or_system *s = &m->s;
  memset(s->f, 0, sizeof(double) * 3 * (size_t)s->N_max);  /* examinimd.cpp:232 */
  if (m->cfg.force_type == OR_FORCE_LJ)
    or_lj_force(s, &m->c, m->cfg.lj_eps, m->cfg.lj_sigma, m->cfg.lj_cut, m->cfg.half_neigh);
  else if (m->cfg.force_type == OR_FORCE_LJ_IDIAL)
    or_lj_idial_force(s, &m->c, m->cfg.lj_eps, m->cfg.lj_sigma, m->cfg.lj_cut,
                      m->cfg.lj_intensity, m->cfg.half_neigh);
  else
    or_snap_force(s, &m->c, &m->snap);
}

static void md_rebuild(or_md *m, int exchange, double *tc, double *tn, double *to) {
  or_system *s = &m->s;
  double nc = md_neigh_cut(m), t0;
  t0 = now_s();
  if (exchange) or_comm_exchange(s);
  *tc += now_s() - t0; t0 = now_s();
  or_binning(s, &m->b, nc, nc, nc, 1, 1, 0, 1);
  *to += now_s() - t0; t0 = now_s();
  or_comm_exchange_halo(s, &m->h, nc);
  *tc += now_s() - t0; t0 = now_s();
  or_binning(s, &m->b, nc, nc, nc, 1, 1, 1, 0);
  or_neigh_csr(s, &m->b, nc, m->cfg.half_neigh, m->cfg.newton, &m->c);
  *tn += now_s() - t0;
}

int or_md_init(or_md *m) {
#ifdef _OPENMP
  if (m->cfg.nthreads > 0) omp_set_num_threads(m->cfg.nthreads);
#endif
  or_system *s = &m->s;
  or_units(s, m->cfg.units);
  if (m->cfg.dt > 0) s->dt = m->cfg.dt;
  s->ntypes = m->cfg.ntypes > 0 ? m->cfg.ntypes : 1;
  for (int t = 0; t < 16; t++) s->mass[t] = m->cfg.mass[t];
  or_create_lattice(s, m->cfg.lattice_style, m->cfg.lattice_arg, m->cfg.lattice_offset,
                    m->cfg.nx, m->cfg.ny, m->cfg.nz, m->cfg.temperature, m->cfg.seed,
                    m->cfg.vel_gaussian);
  double tc = 0, tn = 0, to = 0;
  md_rebuild(m, 1, &tc, &tn, &to);
  md_force(m);
  if (m->cfg.newton) or_comm_update_force(s, &m->h);
  return 0;
}

double or_md_potential(or_md *m) {
  /* lj/cut/idial keeps Force::compute_energy's default 0 (src/force.h:54) */
  if (m->cfg.force_type == OR_FORCE_LJ_IDIAL) return 0.0;
  if (m->cfg.force_type == OR_FORCE_LJ)
    return or_lj_energy(&m->s, &m->c, m->cfg.lj_eps, m->cfg.lj_sigma, m->cfg.lj_cut,
                        m->cfg.half_neigh);
  return or_snap_energy(&m->s, &m->c, &m->snap, NULL);
}

void or_md_thermo(or_md *m, double out[3]) {
  double T = or_temperature(&m->s);
  double PE = or_md_potential(m) / m->s.N;
  double KE = or_kinetic_energy(&m->s) / m->s.N;
  out[0] = T; out[1] = PE; out[2] = PE + KE;
}

int or_md_run(or_md *m, int nsteps) {
  or_system *s = &m->s;
  double t_start = now_s();
  double tf = 0, tn = 0, tc = 0, to = 0;
  for (int step = 1; step <= nsteps; step++) {
    double t0 = now_s();
    or_nve_initial(s);
    to += now_s() - t0;
    if (step % m->cfg.exchange_rate == 0 && step > 0) {
      md_rebuild(m, 1, &tc, &tn, &to);
    } else {
      t0 = now_s();
      or_comm_update_halo(s, &m->h);
      tc += now_s() - t0;
    }
    t0 = now_s();
    md_force(m);
    tf += now_s() - t0;
    if (m->cfg.newton) {
      t0 = now_s();
      or_comm_update_force(s, &m->h);
      tc += now_s() - t0;
    }
    t0 = now_s();
    or_nve_final(s);
    to += now_s() - t0;
  }
  m->t[0] += now_s() - t_start;
  m->t[1] += tf; m->t[2] += tn; m->t[3] += tc; m->t[4] += to;
  return 0;
}

int or_md_dump(or_md *m, const char *path) { return or_dump_write(&m->s, path); }
