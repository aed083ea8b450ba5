/*
 * oracle_snap.c -- CPU restatement of ExaMiniMD's linear SNAP force
 * (reference form: U_tot -> Z -> per neighbour dU/dr -> dB/dr -> F).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows src/force_types/sna_impl.hpp and
 * src/force_types/force_snap_neigh_impl.h:589-725 in the reference's
 * Kokkos-Serial operation order.  The per-atom energy
 * E_i = beta_0 + sum_k beta_k B_k + sum_j 1.25e5 (r^-12 - rc^-12) is an
 * extension (the reference computes no SNAP energy): B_k follows the
 * published SNAP bispectrum (Thompson et al., J. Comput. Phys. 285 (2015)
 * 316), B_{j1j2j} = sum_{ma,mb} Re(conj(U_j) Z_{j1j2j}) evaluated over the
 * half plane with the same middle-column weights the reference uses for
 * dB/dr; the pair term is the energy whose gradient is the reference's
 * ad-hoc -1.5e6 r^-14 force (force_snap_neigh_impl.h:705-711), shifted to
 * vanish at the cutoff.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const double MY_PI = 3.14159265358979323846;

typedef struct {
  int twojmax, jdim;
  double rfac0, rmin0, wself;
  int switch_flag;
  int nidxj, nidxj_full;
  int idxj[3 * 512], idxj_full[3 * 2048];
  double *cg;        /* [jdim]^5, (j1,j2,j,m1,m2) */
  double *rootpq;    /* [jdim+1]^2 */
} sna_t;

#define U3(a, j, ma, mb) (a)[((j) * JD + (ma)) * JD + (mb)]
#define DU(a, j, mb, ma, k) (a)[(((j) * JD + (mb)) * JD + (ma)) * 3 + (k)]
#define Z5(a, j1, j2, j, mb, ma) (a)[((((j1) * JD + (j2)) * JD + (j)) * JD + (mb)) * JD + (ma)]
#define CG5(t, j1, j2, j, m1, m2) (t)->cg[((((j1) * JD + (j2)) * JD + (j)) * JD + (m1)) * JD + (m2)]
#define RP(t, p, q) (t)->rootpq[(p) * (JD + 1) + (q)]

#define MAXI(a, b) ((a) > (b) ? (a) : (b))
#define MINI(a, b) ((a) < (b) ? (a) : (b))

/* SNA::factorial, sna_impl.hpp:964-971 */
static double factorial(int n) {
  double result = 1.0;
  for (int i = 1; i <= n; i++) result *= 1.0 * i;
  return result;
}

/* SNA::deltacg, sna_impl.hpp:977-984 */
static double deltacg(int j1, int j2, int j) {
  double sfaccg = factorial((j1 + j2 + j) / 2 + 1);
  return sqrt(factorial((j1 + j2 - j) / 2) * factorial((j1 - j2 + j) / 2) *
              factorial((-j1 + j2 + j) / 2) / sfaccg);
}

/* SNA::build_indexlist (sna_impl.hpp:86-132), init_clebsch_gordan
 * (991-1046), init_rootpqarray (1053-1061) */
static void sna_init(sna_t *t, const or_snap_params *p) {
  memset(t, 0, sizeof(*t));
  t->twojmax = p->twojmax;
  t->jdim = p->twojmax + 1;
  t->rfac0 = p->rfac0;
  t->rmin0 = p->rmin0;
  t->switch_flag = p->switchflag;
  t->wself = 1.0;
  const int JD = t->jdim, twojmax = p->twojmax;
  for (int j1 = 0; j1 <= twojmax; j1++)
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = abs(j1 - j2); j <= MINI(twojmax, j1 + j2); j += 2) {
        if (j >= j1) {
          t->idxj[3 * t->nidxj + 0] = j1;
          t->idxj[3 * t->nidxj + 1] = j2;
          t->idxj[3 * t->nidxj + 2] = j;
          t->nidxj++;
        }
        t->idxj_full[3 * t->nidxj_full + 0] = j1;
        t->idxj_full[3 * t->nidxj_full + 1] = j2;
        t->idxj_full[3 * t->nidxj_full + 2] = j;
        t->nidxj_full++;
      }
  size_t n5 = (size_t)JD * JD * JD * JD * JD;
  t->cg = (double *)calloc(n5, sizeof(double));
  t->rootpq = (double *)calloc((size_t)(JD + 1) * (JD + 1), sizeof(double));
  for (int j1 = 0; j1 <= twojmax; j1++)
    for (int j2 = 0; j2 <= twojmax; j2++)
      for (int j = abs(j1 - j2); j <= MINI(twojmax, j1 + j2); j += 2)
        for (int m1 = 0; m1 <= j1; m1 += 1) {
          int aa2 = 2 * m1 - j1;
          for (int m2 = 0; m2 <= j2; m2 += 1) {
            int bb2 = 2 * m2 - j2;
            int m = (aa2 + bb2 + j) / 2;
            if (m < 0 || m > j) continue;
            double sum = 0.0;
            for (int z = MAXI(0, MAXI(-(j - j2 + aa2) / 2, -(j - j1 - bb2) / 2));
                 z <= MINI((j1 + j2 - j) / 2, MINI((j1 - aa2) / 2, (j2 + bb2) / 2)); z++) {
              int ifac = z % 2 ? -1 : 1;
              sum += ifac / (factorial(z) * factorial((j1 + j2 - j) / 2 - z) *
                             factorial((j1 - aa2) / 2 - z) * factorial((j2 + bb2) / 2 - z) *
                             factorial((j - j2 + aa2) / 2 + z) *
                             factorial((j - j1 - bb2) / 2 + z));
            }
            int cc2 = 2 * m - j;
            double dcg = deltacg(j1, j2, j);
            double sfaccg = sqrt(factorial((j1 + aa2) / 2) * factorial((j1 - aa2) / 2) *
                                 factorial((j2 + bb2) / 2) * factorial((j2 - bb2) / 2) *
                                 factorial((j + cc2) / 2) * factorial((j - cc2) / 2) * (j + 1));
            CG5(t, j1, j2, j, m1, m2) = sum * dcg * sfaccg;
          }
        }
  for (int p2 = 1; p2 <= twojmax; p2++)
    for (int q = 1; q <= twojmax; q++) RP(t, p2, q) = sqrt((double)p2 / q);
}

static void sna_free(sna_t *t) {
  free(t->cg);
  free(t->rootpq);
}

/* SNA::compute_sfac / compute_dsfac, sna_impl.hpp:1098-1128 */
static double compute_sfac(const sna_t *t, double r, double rcut) {
  if (t->switch_flag == 0) return 1.0;
  if (t->switch_flag == 1) {
    if (r <= t->rmin0) return 1.0;
    else if (r > rcut) return 0.0;
    else {
      double rcutfac = MY_PI / (rcut - t->rmin0);
      return 0.5 * (cos((r - t->rmin0) * rcutfac) + 1.0);
    }
  }
  return 0.0;
}

static double compute_dsfac(const sna_t *t, double r, double rcut) {
  if (t->switch_flag == 0) return 0.0;
  if (t->switch_flag == 1) {
    if (r <= t->rmin0) return 0.0;
    else if (r > rcut) return 0.0;
    else {
      double rcutfac = MY_PI / (rcut - t->rmin0);
      return -0.5 * sin((r - t->rmin0) * rcutfac) * rcutfac;
    }
  }
  return 0.0;
}

/* SNA::compute_uarray, sna_impl.hpp:641-720 */
static void compute_uarray(const sna_t *t, double x, double y, double z, double z0, double r,
                           double *ur, double *ui) {
  const int JD = t->jdim;
  double r0inv = 1.0 / sqrt(r * r + z0 * z0);
  double a_r = r0inv * z0;
  double a_i = -r0inv * z;
  double b_r = r0inv * y;
  double b_i = -r0inv * x;
  U3(ur, 0, 0, 0) = 1.0;
  U3(ui, 0, 0, 0) = 0.0;
  for (int j = 1; j <= t->twojmax; j++) {
    for (int mb = 0; mb < (j + 2) / 2; mb++) {
      U3(ur, j, 0, mb) = 0.0;
      U3(ui, j, 0, mb) = 0.0;
      for (int ma = 0; ma < j; ma++) {
        double rootpq = RP(t, j - ma, j - mb);
        U3(ur, j, ma, mb) += rootpq * (a_r * U3(ur, j - 1, ma, mb) + a_i * U3(ui, j - 1, ma, mb));
        U3(ui, j, ma, mb) += rootpq * (a_r * U3(ui, j - 1, ma, mb) - a_i * U3(ur, j - 1, ma, mb));
        rootpq = RP(t, ma + 1, j - mb);
        U3(ur, j, ma + 1, mb) = -rootpq * (b_r * U3(ur, j - 1, ma, mb) + b_i * U3(ui, j - 1, ma, mb));
        U3(ui, j, ma + 1, mb) = -rootpq * (b_r * U3(ui, j - 1, ma, mb) - b_i * U3(ur, j - 1, ma, mb));
      }
    }
    for (int mb = 0; mb < (j + 2) / 2; mb++) {
      int mbpar = (mb) % 2 == 0 ? 1 : -1;
      int mapar = -mbpar;
      for (int ma = 0; ma <= j; ma++) {
        mapar = -mapar;
        if (mapar == 1) {
          U3(ur, j, j - ma, j - mb) = U3(ur, j, ma, mb);
          U3(ui, j, j - ma, j - mb) = -U3(ui, j, ma, mb);
        } else {
          U3(ur, j, j - ma, j - mb) = -U3(ur, j, ma, mb);
          U3(ui, j, j - ma, j - mb) = U3(ui, j, ma, mb);
        }
      }
    }
  }
}

/* SNA::compute_ui (sna_impl.hpp:152-190) with zero_/addself_/add_uarraytot
 * (574-635) in serial neighbour order. */
static void compute_ui(const sna_t *t, int nin, const double *rij, const double *rcutij,
                       const double *wj, double *utr, double *uti, double *ur, double *ui) {
  const int JD = t->jdim;
  const int n3 = JD * JD * JD;
  memset(utr, 0, sizeof(double) * n3);
  memset(uti, 0, sizeof(double) * n3);
  for (int j = 0; j <= t->twojmax; j++)
    for (int ma = 0; ma <= j; ma++) {
      U3(utr, j, ma, ma) = t->wself;
      U3(uti, j, ma, ma) = 0.0;
    }
  for (int jj = 0; jj < nin; jj++) {
    double x = rij[3 * jj], y = rij[3 * jj + 1], z = rij[3 * jj + 2];
    double rsq = x * x + y * y + z * z;
    double r = sqrt(rsq);
    double theta0 = (r - t->rmin0) * t->rfac0 * MY_PI / (rcutij[jj] - t->rmin0);
    double z0 = r / tan(theta0);
    compute_uarray(t, x, y, z, z0, r, ur, ui);
    double sfac = compute_sfac(t, r, rcutij[jj]) * wj[jj];
    for (int k = 0; k < n3; k++) {
      utr[k] += sfac * ur[k];
      uti[k] += sfac * ui[k];
    }
  }
}

/* SNA::compute_zi, sna_impl.hpp:196-283 */
static void compute_zi(const sna_t *t, const double *utr, const double *uti, double *zr,
                       double *zi) {
  const int JD = t->jdim;
  for (int idx = 0; idx < t->nidxj_full; idx++) {
    const int j1 = t->idxj_full[3 * idx], j2 = t->idxj_full[3 * idx + 1],
              j = t->idxj_full[3 * idx + 2];
    const int bound = (j + 2) / 2;
    for (int mbma = 0; mbma < (j + 1) * bound; mbma++) {
      const int ma = mbma % (j + 1);
      const int mb = mbma / (j + 1);
      double z_r = 0.0, z_i = 0.0;
      for (int ma1 = MAXI(0, (2 * ma - j - j2 + j1) / 2);
           ma1 <= MINI(j1, (2 * ma - j + j2 + j1) / 2); ma1++) {
        double sumb1_r = 0.0, sumb1_i = 0.0;
        const int ma2 = (2 * ma - j - (2 * ma1 - j1) + j2) / 2;
        for (int mb1 = MAXI(0, (2 * mb - j - j2 + j1) / 2);
             mb1 <= MINI(j1, (2 * mb - j + j2 + j1) / 2); mb1++) {
          const int mb2 = (2 * mb - j - (2 * mb1 - j1) + j2) / 2;
          const double cga = CG5(t, j1, j2, j, mb1, mb2);
          const double uat1_r = U3(utr, j1, ma1, mb1);
          const double uat1_i = U3(uti, j1, ma1, mb1);
          const double uat2_r = U3(utr, j2, ma2, mb2);
          const double uat2_i = U3(uti, j2, ma2, mb2);
          sumb1_r += cga * (uat1_r * uat2_r - uat1_i * uat2_i);
          sumb1_i += cga * (uat1_r * uat2_i + uat1_i * uat2_r);
        }
        const double cga = CG5(t, j1, j2, j, ma1, ma2);
        z_r += sumb1_r * cga;
        z_i += sumb1_i * cga;
      }
      Z5(zr, j1, j2, j, mb, ma) = z_r;
      Z5(zi, j1, j2, j, mb, ma) = z_i;
    }
  }
}

/* SNA::compute_duidrj (sna_impl.hpp:290-321) + compute_duarray (728-893) */
static void compute_duidrj(const sna_t *t, const double *rij, double wj, double rcut,
                           double *ur, double *ui, double *dur, double *dui) {
  const int JD = t->jdim;
  double x = rij[0], y = rij[1], z = rij[2];
  double rsq = x * x + y * y + z * z;
  double r = sqrt(rsq);
  double rscale0 = t->rfac0 * MY_PI / (rcut - t->rmin0);
  double theta0 = (r - t->rmin0) * rscale0;
  double cs = cos(theta0);
  double sn = sin(theta0);
  double z0 = r * cs / sn;
  double dz0dr = z0 / r - (r * rscale0) * (rsq + z0 * z0) / rsq;

  double rinv = 1.0 / r;
  double ux = x * rinv, uy = y * rinv, uz = z * rinv;
  double r0inv = 1.0 / sqrt(r * r + z0 * z0);
  double a_r = z0 * r0inv, a_i = -z * r0inv, b_r = y * r0inv, b_i = -x * r0inv;
  double dr0invdr = -pow(r0inv, 3.0) * (r + z0 * dz0dr);
  double dr0inv[3] = {dr0invdr * ux, dr0invdr * uy, dr0invdr * uz};
  double dz0[3] = {dz0dr * ux, dz0dr * uy, dz0dr * uz};
  double da_r[3], da_i[3], db_r[3], db_i[3];
  for (int k = 0; k < 3; k++) {
    da_r[k] = dz0[k] * r0inv + z0 * dr0inv[k];
    da_i[k] = -z * dr0inv[k];
  }
  da_i[2] += -r0inv;
  for (int k = 0; k < 3; k++) {
    db_r[k] = y * dr0inv[k];
    db_i[k] = -x * dr0inv[k];
  }
  db_i[0] += -r0inv;
  db_r[1] += r0inv;

  U3(ur, 0, 0, 0) = 1.0;
  U3(ui, 0, 0, 0) = 0.0;
  for (int k = 0; k < 3; k++) { DU(dur, 0, 0, 0, k) = 0.0; DU(dui, 0, 0, 0, k) = 0.0; }
  for (int j = 1; j <= t->twojmax; j++) {
    for (int mb = 0; mb < (j + 2) / 2; mb++) {
      U3(ur, j, 0, mb) = 0.0;
      U3(ui, j, 0, mb) = 0.0;
      for (int k = 0; k < 3; k++) { DU(dur, j, mb, 0, k) = 0.0; DU(dui, j, mb, 0, k) = 0.0; }
      for (int ma = 0; ma < j; ma++) {
        double rootpq = RP(t, j - ma, j - mb);
        U3(ur, j, ma, mb) += rootpq * (a_r * U3(ur, j - 1, ma, mb) + a_i * U3(ui, j - 1, ma, mb));
        U3(ui, j, ma, mb) += rootpq * (a_r * U3(ui, j - 1, ma, mb) - a_i * U3(ur, j - 1, ma, mb));
        for (int k = 0; k < 3; k++) {
          DU(dur, j, mb, ma, k) += rootpq * (da_r[k] * U3(ur, j - 1, ma, mb) +
                                             da_i[k] * U3(ui, j - 1, ma, mb) +
                                             a_r * DU(dur, j - 1, mb, ma, k) +
                                             a_i * DU(dui, j - 1, mb, ma, k));
          DU(dui, j, mb, ma, k) += rootpq * (da_r[k] * U3(ui, j - 1, ma, mb) -
                                             da_i[k] * U3(ur, j - 1, ma, mb) +
                                             a_r * DU(dui, j - 1, mb, ma, k) -
                                             a_i * DU(dur, j - 1, mb, ma, k));
        }
        rootpq = RP(t, ma + 1, j - mb);
        U3(ur, j, ma + 1, mb) = -rootpq * (b_r * U3(ur, j - 1, ma, mb) + b_i * U3(ui, j - 1, ma, mb));
        U3(ui, j, ma + 1, mb) = -rootpq * (b_r * U3(ui, j - 1, ma, mb) - b_i * U3(ur, j - 1, ma, mb));
        for (int k = 0; k < 3; k++) {
          DU(dur, j, mb, ma + 1, k) = -rootpq * (db_r[k] * U3(ur, j - 1, ma, mb) +
                                                 db_i[k] * U3(ui, j - 1, ma, mb) +
                                                 b_r * DU(dur, j - 1, mb, ma, k) +
                                                 b_i * DU(dui, j - 1, mb, ma, k));
          DU(dui, j, mb, ma + 1, k) = -rootpq * (db_r[k] * U3(ui, j - 1, ma, mb) -
                                                 db_i[k] * U3(ur, j - 1, ma, mb) +
                                                 b_r * DU(dui, j - 1, mb, ma, k) -
                                                 b_i * DU(dur, j - 1, mb, ma, k));
        }
      }
    }
    for (int mb = 0; mb < (j + 2) / 2; mb++) {
      int mbpar = (mb) % 2 == 0 ? 1 : -1;
      int mapar = -mbpar;
      for (int ma = 0; ma <= j; ma++) {
        mapar = -mapar;
        if (mapar == 1) {
          U3(ur, j, j - ma, j - mb) = U3(ur, j, ma, mb);
          U3(ui, j, j - ma, j - mb) = -U3(ui, j, ma, mb);
          for (int k = 0; k < 3; k++) {
            DU(dur, j, j - mb, j - ma, k) = DU(dur, j, mb, ma, k);
            DU(dui, j, j - mb, j - ma, k) = -DU(dui, j, mb, ma, k);
          }
        } else {
          U3(ur, j, j - ma, j - mb) = -U3(ur, j, ma, mb);
          U3(ui, j, j - ma, j - mb) = U3(ui, j, ma, mb);
          for (int k = 0; k < 3; k++) {
            DU(dur, j, j - mb, j - ma, k) = -DU(dur, j, mb, ma, k);
            DU(dui, j, j - mb, j - ma, k) = DU(dui, j, mb, ma, k);
          }
        }
      }
    }
  }
  double sfac = compute_sfac(t, r, rcut);
  double dsfac = compute_dsfac(t, r, rcut);
  sfac *= wj;
  dsfac *= wj;
  const double uu[3] = {ux, uy, uz};
  for (int j = 0; j <= t->twojmax; j++)
    for (int mb = 0; mb <= j; mb++)
      for (int ma = 0; ma <= j; ma++)
        for (int k = 0; k < 3; k++) {
          DU(dur, j, mb, ma, k) = dsfac * U3(ur, j, ma, mb) * uu[k] + sfac * DU(dur, j, mb, ma, k);
          DU(dui, j, mb, ma, k) = dsfac * U3(ui, j, ma, mb) * uu[k] + sfac * DU(dui, j, mb, ma, k);
        }
}

/* one of the three half-plane sums of SNA::compute_dbidrj */
static void zdu_half(int JD, int jr, const double *dur, const double *dui, const double *zr,
                     const double *zi, int a1, int a2, int a3, double s[3]) {
  s[0] = s[1] = s[2] = 0.0;
  for (int mb = 0; 2 * mb < jr; mb++)
    for (int ma = 0; ma <= jr; ma++) {
      const double zzr = Z5(zr, a1, a2, a3, mb, ma), zzi = Z5(zi, a1, a2, a3, mb, ma);
      for (int k = 0; k < 3; k++)
        s[k] += (DU(dur, jr, mb, ma, k) * zzr + DU(dui, jr, mb, ma, k) * zzi);
    }
  if (jr % 2 == 0) {
    int mb = jr / 2;
    for (int ma = 0; ma <= mb; ma++) {
      const double factor = ma == mb ? 0.5 : 1.0;
      const double zzr = Z5(zr, a1, a2, a3, mb, ma) * factor;
      const double zzi = Z5(zi, a1, a2, a3, mb, ma) * factor;
      for (int k = 0; k < 3; k++)
        s[k] += (DU(dur, jr, mb, ma, k) * zzr + DU(dui, jr, mb, ma, k) * zzi);
    }
  }
}

/* SNA::compute_dbidrj (sna_impl.hpp:329-539) + copy_dbi2dbvec (545-570) */
static void compute_dbidrj(const sna_t *t, const double *dur, const double *dui,
                           const double *zr, const double *zi, double *dbvec) {
  const int JD = t->jdim;
  for (int JJ = 0; JJ < t->nidxj; JJ++) {
    const int j1 = t->idxj[3 * JJ], j2 = t->idxj[3 * JJ + 1], j = t->idxj[3 * JJ + 2];
    double dbdr[3] = {0.0, 0.0, 0.0}, s[3];
    if (j1 >= j2) zdu_half(JD, j, dur, dui, zr, zi, j1, j2, j, s);
    else zdu_half(JD, j, dur, dui, zr, zi, j2, j1, j, s);
    for (int k = 0; k < 3; k++) dbdr[k] += s[k] * 2.0;
    double j1fac = (j + 1) / (j1 + 1.0);
    if (j >= j2) zdu_half(JD, j1, dur, dui, zr, zi, j, j2, j1, s);
    else zdu_half(JD, j1, dur, dui, zr, zi, j2, j, j1, s);
    for (int k = 0; k < 3; k++) dbdr[k] += (s[k] * 2.0) * j1fac;
    double j2fac = (j + 1) / (j2 + 1.0);
    if (j1 >= j) zdu_half(JD, j2, dur, dui, zr, zi, j1, j, j2, s);
    else zdu_half(JD, j2, dur, dui, zr, zi, j, j1, j2, s);
    for (int k = 0; k < 3; k++) dbdr[k] += (s[k] * 2.0) * j2fac;
    for (int k = 0; k < 3; k++) dbvec[3 * JJ + k] = dbdr[k];
  }
}

/* bispectrum B_{j1j2j} over the half plane with middle-column weights */
static void compute_bi(const sna_t *t, const double *utr, const double *uti, const double *zr,
                       const double *zi, int bzero, double *B) {
  const int JD = t->jdim;
  for (int JJ = 0; JJ < t->nidxj; JJ++) {
    const int j1 = t->idxj[3 * JJ], j2 = t->idxj[3 * JJ + 1], j = t->idxj[3 * JJ + 2];
    double b = 0.0;
    for (int mb = 0; 2 * mb < j; mb++)
      for (int ma = 0; ma <= j; ma++)
        b += U3(utr, j, ma, mb) * Z5(zr, j1, j2, j, mb, ma) +
             U3(uti, j, ma, mb) * Z5(zi, j1, j2, j, mb, ma);
    if (j % 2 == 0) {
      int mb = j / 2;
      for (int ma = 0; ma < mb; ma++)
        b += U3(utr, j, ma, mb) * Z5(zr, j1, j2, j, mb, ma) +
             U3(uti, j, ma, mb) * Z5(zi, j1, j2, j, mb, ma);
      int ma = mb;
      b += 0.5 * (U3(utr, j, ma, mb) * Z5(zr, j1, j2, j, mb, ma) +
                  U3(uti, j, ma, mb) * Z5(zi, j1, j2, j, mb, ma));
    }
    b *= 2.0;
    if (bzero) b -= t->wself * t->wself * t->wself * (j + 1);
    B[JJ] = b;
  }
}

/* ---------------------------------------------------------------- public */
int or_snap_ncoeff(int twojmax) {
  int n = 0;
  for (int j1 = 0; j1 <= twojmax; j1++)
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = abs(j1 - j2); j <= MINI(twojmax, j1 + j2); j += 2)
        if (j >= j1) n++;
  return n;
}

int or_snap_idxj(int twojmax, int *o1, int *o2, int *o3) {
  int n = 0;
  for (int j1 = 0; j1 <= twojmax; j1++)
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = abs(j1 - j2); j <= MINI(twojmax, j1 + j2); j += 2)
        if (j >= j1) {
          if (o1) { o1[n] = j1; o2[n] = j2; o3[n] = j; }
          n++;
        }
  return n;
}

/* the whole table, out[(((j1 * JD + j2) * JD + j) * JD + m1) * JD + m2],
 * JD = twojmax + 1 (entries the table does not define are 0) */
void or_snap_cg_all(int twojmax, double *out) {
  or_snap_params p;
  memset(&p, 0, sizeof(p));
  p.twojmax = twojmax;
  p.rfac0 = 0.99363;
  sna_t t;
  sna_init(&t, &p);
  const int JD = t.jdim;
  memcpy(out, t.cg, sizeof(double) * JD * JD * JD * JD * JD);
  sna_free(&t);
}

double or_snap_cg(int twojmax, int j1, int j2, int j, int m1, int m2) {
  or_snap_params p;
  memset(&p, 0, sizeof(p));
  p.twojmax = twojmax;
  p.rfac0 = 0.99363;
  sna_t t;
  sna_init(&t, &p);
  const int JD = t.jdim;
  double v = CG5(&t, j1, j2, j, m1, m2);
  sna_free(&t);
  return v;
}

typedef struct {
  double *ur, *ui, *utr, *uti, *dur, *dui, *zr, *zi, *db;
  double *rij, *rcutij, *wj;
  int *inside;
  int cap;
} sna_scratch;

static void scratch_alloc(sna_scratch *w, int JD, int nmax) {
  size_t n3 = (size_t)JD * JD * JD, n5 = n3 * JD * JD;
  w->ur = calloc(n3, 8); w->ui = calloc(n3, 8);
  w->utr = calloc(n3, 8); w->uti = calloc(n3, 8);
  w->dur = calloc(n3 * 3, 8); w->dui = calloc(n3 * 3, 8);
  w->zr = calloc(n5, 8); w->zi = calloc(n5, 8);
  w->db = calloc(3 * 1024, 8);
  w->cap = nmax > 0 ? nmax : 1;
  w->rij = calloc((size_t)w->cap * 3, 8);
  w->rcutij = calloc((size_t)w->cap, 8);
  w->wj = calloc((size_t)w->cap, 8);
  w->inside = calloc((size_t)w->cap, sizeof(int));
}

static void scratch_free(sna_scratch *w) {
  free(w->ur); free(w->ui); free(w->utr); free(w->uti); free(w->dur); free(w->dui);
  free(w->zr); free(w->zi); free(w->db); free(w->rij); free(w->rcutij); free(w->wj);
  free(w->inside);
}

/* neighbours of i inside the SNAP cutoff, force_snap_neigh_impl.h:612-656 */
static int gather_inside(const or_system *s, const or_csr *c, const or_snap_params *p, int i,
                         sna_scratch *w) {
  const double rcutmax = 2.0 * p->radelem * p->rcutfac;
  const double cutsq = rcutmax * rcutmax;
  const double x_i = s->x[3 * (size_t)i], y_i = s->x[3 * (size_t)i + 1],
               z_i = s->x[3 * (size_t)i + 2];
  int n = 0;
  for (int jj = c->row[i]; jj < c->row[i + 1]; jj++) {
    int j = c->entries[jj];
    const double dx = s->x[3 * (size_t)j] - x_i;
    const double dy = s->x[3 * (size_t)j + 1] - y_i;
    const double dz = s->x[3 * (size_t)j + 2] - z_i;
    const double rsq = dx * dx + dy * dy + dz * dz;
    if (rsq < cutsq) {
      w->rij[3 * n] = dx; w->rij[3 * n + 1] = dy; w->rij[3 * n + 2] = dz;
      w->inside[n] = j;
      w->wj[n] = p->wjelem;
      w->rcutij[n] = (p->radelem + p->radelem) * p->rcutfac;
      n++;
    }
  }
  return n;
}

static int max_row(const or_csr *c) {
  int m = 0;
  for (int i = 0; i < c->n; i++)
    if (c->row[i + 1] - c->row[i] > m) m = c->row[i + 1] - c->row[i];
  return m;
}

/* ForceSNAP::operator(), force_snap_neigh_impl.h:589-725.  F_ij for every
 * (i, inside neighbour) is computed in parallel, then scattered in the
 * serial order f(i) += F_ij, f(j) -= F_ij. */
void or_snap_force(or_system *s, const or_csr *c, const or_snap_params *p) {
  sna_t t;
  sna_init(&t, p);
  const int N = s->N_local, JD = t.jdim, nmax = max_row(c);
  double *fij = (double *)malloc(sizeof(double) * 3 * (size_t)(c->total ? c->total : 1));
  int *jidx = (int *)malloc(sizeof(int) * (size_t)(c->total ? c->total : 1));
  int *nin = (int *)malloc(sizeof(int) * (size_t)(N ? N : 1));
#pragma omp parallel
  {
    sna_scratch w;
    scratch_alloc(&w, JD, nmax);
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < N; i++) {
      int n = gather_inside(s, c, p, i, &w);
      nin[i] = n;
      compute_ui(&t, n, w.rij, w.rcutij, w.wj, w.utr, w.uti, w.ur, w.ui);
      compute_zi(&t, w.utr, w.uti, w.zr, w.zi);
      for (int jj = 0; jj < n; jj++) {
        compute_duidrj(&t, w.rij + 3 * jj, w.wj[jj], w.rcutij[jj], w.ur, w.ui, w.dur, w.dui);
        compute_dbidrj(&t, w.dur, w.dui, w.zr, w.zi, w.db);
        double f3[3] = {0.0, 0.0, 0.0};
        for (int k = 1; k <= p->ncoeff; k++) {
          double bgb = p->coeff[k];
          f3[0] += bgb * w.db[3 * (k - 1) + 0];
          f3[1] += bgb * w.db[3 * (k - 1) + 1];
          f3[2] += bgb * w.db[3 * (k - 1) + 2];
        }
        const double dx = w.rij[3 * jj], dy = w.rij[3 * jj + 1], dz = w.rij[3 * jj + 2];
        const double fdivr = -1.5e6 / pow(dx * dx + dy * dy + dz * dz, 7.0);
        f3[0] += dx * fdivr;
        f3[1] += dy * fdivr;
        f3[2] += dz * fdivr;
        size_t o = (size_t)c->row[i] + jj;
        fij[3 * o] = f3[0]; fij[3 * o + 1] = f3[1]; fij[3 * o + 2] = f3[2];
        jidx[o] = w.inside[jj];
      }
    }
    scratch_free(&w);
  }
  for (int i = 0; i < N; i++)
    for (int jj = 0; jj < nin[i]; jj++) {
      size_t o = (size_t)c->row[i] + jj;
      int j = jidx[o];
      s->f[3 * (size_t)i] += fij[3 * o];
      s->f[3 * (size_t)i + 1] += fij[3 * o + 1];
      s->f[3 * (size_t)i + 2] += fij[3 * o + 2];
      s->f[3 * (size_t)j] -= fij[3 * o];
      s->f[3 * (size_t)j + 1] -= fij[3 * o + 1];
      s->f[3 * (size_t)j + 2] -= fij[3 * o + 2];
    }
  free(fij); free(jidx); free(nin);
  sna_free(&t);
}

double or_snap_energy(const or_system *s, const or_csr *c, const or_snap_params *p,
                      double *eatom) {
  sna_t t;
  sna_init(&t, p);
  const int N = s->N_local, JD = t.jdim, nmax = max_row(c);
  const double rc = 2.0 * p->radelem * p->rcutfac;
  const double rc12 = 1.0 / pow(rc * rc, 6.0);
  double *e = (double *)malloc(sizeof(double) * (size_t)(N ? N : 1));
#pragma omp parallel
  {
    sna_scratch w;
    scratch_alloc(&w, JD, nmax);
    double B[1024];
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < N; i++) {
      int n = gather_inside(s, c, p, i, &w);
      compute_ui(&t, n, w.rij, w.rcutij, w.wj, w.utr, w.uti, w.ur, w.ui);
      compute_zi(&t, w.utr, w.uti, w.zr, w.zi);
      compute_bi(&t, w.utr, w.uti, w.zr, w.zi, p->bzeroflag, B);
      double ei = p->coeff[0];
      for (int k = 1; k <= p->ncoeff; k++) ei += p->coeff[k] * B[k - 1];
      for (int jj = 0; jj < n; jj++) {
        double rsq = w.rij[3 * jj] * w.rij[3 * jj] + w.rij[3 * jj + 1] * w.rij[3 * jj + 1] +
                     w.rij[3 * jj + 2] * w.rij[3 * jj + 2];
        ei += 1.25e5 * (1.0 / pow(rsq, 6.0) - rc12);
      }
      e[i] = ei;
    }
    scratch_free(&w);
  }
  double E = 0.0, Ec = 0.0;
  for (int i = 0; i < N; i++) {
    double t = E + e[i];
    if (fabs(E) >= fabs(e[i])) Ec += (E - t) + e[i];
    else Ec += (e[i] - t) + E;
    E = t;
    if (eatom) eatom[i] = e[i];
  }
  E += Ec;
  free(e);
  sna_free(&t);
  return E;
}

/* single-centre helpers for the invariance / derivative tests */
void or_snap_bispectrum(const or_snap_params *p, int nnei, const double *rij, double *B) {
  sna_t t;
  sna_init(&t, p);
  sna_scratch w;
  scratch_alloc(&w, t.jdim, nnei);
  for (int k = 0; k < nnei; k++) {
    w.rcutij[k] = (p->radelem + p->radelem) * p->rcutfac;
    w.wj[k] = p->wjelem;
  }
  compute_ui(&t, nnei, rij, w.rcutij, w.wj, w.utr, w.uti, w.ur, w.ui);
  compute_zi(&t, w.utr, w.uti, w.zr, w.zi);
  compute_bi(&t, w.utr, w.uti, w.zr, w.zi, p->bzeroflag, B);
  scratch_free(&w);
  sna_free(&t);
}

void or_snap_ui(const or_snap_params *p, int nnei, const double *rij, double *ur, double *ui) {
  sna_t t;
  sna_init(&t, p);
  sna_scratch w;
  scratch_alloc(&w, t.jdim, nnei);
  for (int k = 0; k < nnei; k++) {
    w.rcutij[k] = (p->radelem + p->radelem) * p->rcutfac;
    w.wj[k] = p->wjelem;
  }
  compute_ui(&t, nnei, rij, w.rcutij, w.wj, w.utr, w.uti, w.ur, w.ui);
  size_t n3 = (size_t)t.jdim * t.jdim * t.jdim;
  memcpy(ur, w.utr, n3 * 8);
  memcpy(ui, w.uti, n3 * 8);
  scratch_free(&w);
  sna_free(&t);
}

void or_snap_dbidrj(const or_snap_params *p, int nnei, const double *rij, int jj, double *dB) {
  sna_t t;
  sna_init(&t, p);
  sna_scratch w;
  scratch_alloc(&w, t.jdim, nnei);
  for (int k = 0; k < nnei; k++) {
    w.rcutij[k] = (p->radelem + p->radelem) * p->rcutfac;
    w.wj[k] = p->wjelem;
  }
  compute_ui(&t, nnei, rij, w.rcutij, w.wj, w.utr, w.uti, w.ur, w.ui);
  compute_zi(&t, w.utr, w.uti, w.zr, w.zi);
  compute_duidrj(&t, rij + 3 * jj, w.wj[jj], w.rcutij[jj], w.ur, w.ui, w.dur, w.dui);
  compute_dbidrj(&t, w.dur, w.dui, w.zr, w.zi, dB);
  scratch_free(&w);
  sna_free(&t);
}
