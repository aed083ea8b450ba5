// emd_snap.hip -- linear SNAP forces and energy on the GPU.
//
// The reference (ForceSNAP + SNA) computes, per atom, U_tot, the full Z
// array, and per neighbour dU/dr and dB_k/dr for all 55 components before
// contracting with beta.  Here the same force is evaluated in adjoint form:
// the beta-weighted sum over the three symmetry roles of dB/dr (the three
// half-plane sums of SNA::compute_dbidrj, sna_impl.hpp:329-539) is folded
// into one array Y_J(ma,mb) per atom, so each neighbour needs only
// dU_J/dr contracted with Y_J:
//
//   F_ij = sum_J sum'_{ma,mb} Re( conj(dU_J(ma,mb)/dr_j) Y_J(ma,mb) )
//   Y_J  = sum_{a>=b} c_{abJ} Z_{abJ},  c collects 2 beta_k, 2 beta_k (j+1)/(j1+1)
//          and 2 beta_k (j+1)/(j2+1) for the roles in which (a,b,J) appears
//
// (sum' is the reference's half plane: mb < J/2, plus for even J the middle
// column ma <= J/2 with weight 1/2 on the diagonal).
//
// Kernels (T = double, or float for the mixed path):
//   k_snap_neigh  one wavefront per atom: compacts the in-cutoff neighbours
//                 (FP64 distance test, list order kept) with ballots and
//                 writes each neighbour's Cayley-Klein parameters, switching
//                 function and derivative factors (snap_par) once
//   k_snap_neigh_ui (twojmax <= 8, force path) the compaction and U_tot of
//                 k_snap_ui_rg in the same wavefront: one launch
//   k_snap_ui_rg  (twojmax <= 8) one wavefront per atom, seven neighbours at
//                 a time, nine lanes per neighbour: the layer recursion in
//                 registers (lane ma holds column entries (ma, mb)), sums
//                 sfac*U per lane, then over the groups; U_tot (285 complex
//                 at twojmax 8) written with its mirror entries
//   k_snap_yflat  one block per atom: U_tot staged in LDS, Y (155 half-plane
//                 entries at twojmax 8) from a host-built flat stream of
//                 weighted U_a U_b products
//   k_snap_fij_rev (twojmax <= 8) the force as the gradient of
//                 sum Re(conj(U) sfac w Y): U forward in registers (same
//                 layout), adjoints backward, two complex sums per lane,
//                 seven neighbours per wavefront, atomic scatter to f_i / f_j
//   k_snap_ui, k_snap_fij  the same for the size class twojmax <= 12: one
//                 neighbour per wavefront, layers in LDS
//   k_snap_energy one block per atom: B_k from U_tot and Z, energy
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <queue>
#include <set>
#include <tuple>
#include <vector>

#include "emd_internal.hpp"

namespace {

// Two size classes: twojmax <= 8 (the kernels tuned for SNAP-2k, whose
// twojmax is 8) and twojmax <= 12 (the same kernels with larger LDS tables,
// one neighbour per wavefront in the force kernel).  Per class J:
constexpr int SNAP_MAXJ = 8;                 // the first class
constexpr int SNAP_JBIG = 12;                // the second class: the largest twojmax
constexpr int SNAP_RPD = SNAP_JBIG + 2;      // row length of the rootpq table
constexpr double SNAP_PI = 3.14159265358979323846;

// U entries sum_j (j+1)^2, half-plane entries sum_j (j+1)(j/2+1), (a,b,J)
// triples and bispectrum components (diagonalstyle 3) up to twojmax J
constexpr int snap_nu(int J) { return J < 0 ? 0 : snap_nu(J - 1) + (J + 1) * (J + 1); }
constexpr int snap_ny(int J) { return J < 0 ? 0 : snap_ny(J - 1) + (J + 1) * (J / 2 + 1); }
constexpr int snap_hp(int j) { return (j + 1) * (j / 2 + 1); }  // half plane of layer j
constexpr int snap_counts(int J, bool coeff) {
  int n = 0;
  for (int j1 = 0; j1 <= J; j1++)
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = j1 - j2; j <= (J < j1 + j2 ? J : j1 + j2); j += 2)
        if (!coeff || j >= j1) n++;
  return n;
}
template <int J>
struct snap_dim {
  static constexpr int JD = J + 1;
  static constexpr int NU = snap_nu(J);
  static constexpr int NY = snap_ny(J);
  static constexpr int HPW = (snap_hp(J) + 63) / 64 * 64;  // lane-table width of a layer
  static constexpr int HPPREV = snap_hp(J - 1);             // half plane of layer J-1
  static constexpr int FULLPREV = J * J;                    // full layer J-1
};
constexpr int SNAP_MAXT = snap_counts(SNAP_JBIG, false);  // (a,b,J) triples: 343
constexpr int SNAP_MAXC = snap_counts(SNAP_JBIG, true);   // components: 140
constexpr int SNAP_NU_BIG = snap_nu(SNAP_JBIG);
static_assert(snap_nu(8) == 285 && snap_ny(8) == 155 && snap_counts(8, false) == 125 &&
                  snap_counts(8, true) == 55,
              "SNAP table sizes at twojmax 8");

// device tables (filled by emd_snap_create)
struct snap_dev_tables {
  int twojmax;
  int nu, ny;                      // number of U entries / half-plane entries
  int ustart[SNAP_JBIG + 2];       // full layer offsets: ustart[j] + ma*(j+1) + mb
  int ystart[SNAP_JBIG + 2];       // half-plane offsets: ystart[J] + mb*(J+1) + ma
  int yjstart[SNAP_JBIG + 2];      // (a,b) list for each J in ytab
  int ny_terms;
  double rootpq[SNAP_RPD * SNAP_RPD];  // sqrt(p/q) at p * SNAP_RPD + q
  double rfac0, rmin0, rcut, wj, wself;
  int switchflag;
};

struct ytab_entry {
  int a, b, cgoff, pad;
  double coef;
};

struct etab_entry {
  int j1, j2, j, cgoff;
  double beta;
};

struct stab_entry {
  int a, b, cgoff, tid;
};

__device__ __forceinline__ int u_idx(const snap_dev_tables &t, int j, int ma, int mb) {
  return t.ustart[j] + ma * (j + 1) + mb;
}

template <class T>
struct cplx {
  T r, i;
};

// ------------------------------------------------------------- switching
template <class T>
__device__ __forceinline__ void sfac_dsfac(const snap_dev_tables &t, T r, T &sfac, T &dsfac) {
  if (t.switchflag == 0) {
    sfac = 1;
    dsfac = 0;
  } else if (r <= (T)t.rmin0) {
    sfac = 1;
    dsfac = 0;
  } else if (r > (T)t.rcut) {
    sfac = 0;
    dsfac = 0;
  } else {
    const T rcutfac = (T)(SNAP_PI / (t.rcut - t.rmin0));
    sfac = (T)0.5 * (cos((r - (T)t.rmin0) * rcutfac) + (T)1);
    dsfac = (T)-0.5 * sin((r - (T)t.rmin0) * rcutfac) * rcutfac;
  }
}

// Per in-cutoff neighbour, computed once by the neighbour compaction: the
// Cayley-Klein parameters a = (z0 - i z)/r0, b = (y - i x)/r0 of rij, the
// switching function and its derivative (times wj), the unit vector and the
// factors of da/dr and db/dr, which are rank one in the unit vector u:
//   da_r = c1 u, da_i = c2 u - r0inv e_z, db_r = c3 u + r0inv e_y,
//   db_i = c4 u - r0inv e_x
// (SNA::compute_uarray / compute_duarray, sna_impl.hpp:640-700, 728-790)
template <class T>
struct alignas(16) snap_par {
  T a_r, a_i, b_r, b_i;
  T sfac, dsfac;
  T u[3];
  T c1, c2, c3, c4;
  T r0inv;
  T pad[2];
};

template <class T>
__device__ __forceinline__ void snap_par_of(const snap_dev_tables &t, double dx, double dy,
                                            double dz, snap_par<T> &p) {
  const T x = (T)dx, y = (T)dy, z = (T)dz;
  const T rsq = x * x + y * y + z * z;
  const T r = sqrt(rsq);
  const T rscale0 = (T)t.rfac0 * (T)SNAP_PI / ((T)t.rcut - (T)t.rmin0);
  const T theta0 = (r - (T)t.rmin0) * rscale0;
  const T cs = cos(theta0), sn = sin(theta0);
  const T z0 = r * cs / sn;
  const T dz0dr = z0 / r - (r * rscale0) * (rsq + z0 * z0) / rsq;
  const T rinv = (T)1 / r;
  const T r0inv = (T)1 / sqrt(r * r + z0 * z0);
  const T dr0invdr = -(r0inv * r0inv * r0inv) * (r + z0 * dz0dr);
  p.a_r = z0 * r0inv;
  p.a_i = -z * r0inv;
  p.b_r = y * r0inv;
  p.b_i = -x * r0inv;
  T sfac, dsfac;
  sfac_dsfac(t, r, sfac, dsfac);
  p.sfac = sfac * (T)t.wj;
  p.dsfac = dsfac * (T)t.wj;
  p.u[0] = x * rinv;
  p.u[1] = y * rinv;
  p.u[2] = z * rinv;
  p.c1 = dz0dr * r0inv + z0 * dr0invdr;
  p.c2 = -z * dr0invdr;
  p.c3 = y * dr0invdr;
  p.c4 = -x * dr0invdr;
  p.r0inv = r0inv;
  p.pad[0] = p.pad[1] = 0;
}

// the reference's ad-hoc pair repulsion as F/r, force_snap_neigh_impl.h:705-711
__device__ __forceinline__ double snap_fdivr(double dx, double dy, double dz) {
  const double rsq = dx * dx + dy * dy + dz * dz;
  return -1.5e6 / pow(rsq, 7.0);
}

}  // namespace

// ================================================================ kernels
// in-cutoff neighbours of each atom, ForceSNAP::operator() lines 612-656
// (rij = x_j - x_i, test rsq < rcutmax^2); slot k of atom i at i*maxn + k.
// With a force array given (f64 or f32), block i also zeroes its rows
// i, i + n, i + 2n, ... below ntot (deep_copy(f, 0), src/examinimd.cpp:232,
// done here instead of in a launch of its own: the force kernels that add
// into f come after this one).  With p32 or p64 given, the slot's neighbour
// parameters (snap_par, FP32 or FP64) and the repulsion F/r (fdv) are
// written as well.
// (the body of k_snap_neigh for atom i and one wavefront; returns the count)
__device__ __forceinline__ int snap_compact(const snap_dev_tables &t, int i, const double *__restrict__ x,
                                            const int *__restrict__ row, const int *__restrict__ num,
                                            const int *__restrict__ entries, int pitch, int n,
                                            double cutsq, int maxn, int *__restrict__ nin,
                                            int *__restrict__ nj, double *__restrict__ nr,
                                            double *__restrict__ f64, float *__restrict__ f32,
                                            int ntot, snap_par<float> *__restrict__ p32,
                                            snap_par<double> *__restrict__ p64,
                                            double *__restrict__ fdv) {
#pragma clang fp contract(off)  // the distance test as the reference (the parameters are
                                // computed in functions of their own)
  const int lane = threadIdx.x % 64;
  if (f64 || f32) {
    for (int r = i + n * (lane / 3); r < ntot; r += n * 21) {  // lanes 0..62: 21 rows of 3
      if (lane < 63) {
        if (f64) f64[3 * (size_t)r + lane % 3] = 0.0;
        else f32[3 * (size_t)r + lane % 3] = 0.0f;
      }
    }
  }
  const double x_i = x[3 * (size_t)i], y_i = x[3 * (size_t)i + 1], z_i = x[3 * (size_t)i + 2];
  const int *rp = entries + (row ? (size_t)row[i] : (size_t)i * (size_t)pitch);
  const int cnt = num[i];
  const unsigned long long lt = emd_lanemask_lt();
  int run = 0;
  for (int c0 = 0; c0 < cnt; c0 += 64) {
    const int c = c0 + lane;
    bool keep = false;
    int j = 0;
    double dx = 0, dy = 0, dz = 0;
    if (c < cnt) {
      j = rp[c];
      dx = x[3 * (size_t)j] - x_i;
      dy = x[3 * (size_t)j + 1] - y_i;
      dz = x[3 * (size_t)j + 2] - z_i;
      const double rsq = dx * dx + dy * dy + dz * dz;
      keep = rsq < cutsq;
    }
    const unsigned long long m = __ballot(keep);
    if (keep) {
      const int k = run + __popcll(m & lt);
      if (k < maxn) {
        const size_t o = (size_t)i * maxn + k;
        nj[o] = j;
        nr[3 * o] = dx;
        nr[3 * o + 1] = dy;
        nr[3 * o + 2] = dz;
        if (p64) snap_par_of<double>(t, dx, dy, dz, p64[o]);
        if (p32) snap_par_of<float>(t, dx, dy, dz, p32[o]);
        if (fdv) fdv[o] = snap_fdivr(dx, dy, dz);
      }
    }
    run += __popcll(m);
  }
  if (lane == 0) nin[i] = run < maxn ? run : maxn;
  return run < maxn ? run : maxn;
}

__global__ void __launch_bounds__(64)
k_snap_neigh(snap_dev_tables t, const double *__restrict__ x, const int *__restrict__ row,
             const int *__restrict__ num, const int *__restrict__ entries, int pitch, int n,
             double cutsq, int maxn, int *__restrict__ nin, int *__restrict__ nj,
             double *__restrict__ nr, double *__restrict__ f64, float *__restrict__ f32,
             int ntot, snap_par<float> *__restrict__ p32, snap_par<double> *__restrict__ p64,
             double *__restrict__ fdv) {
  const int i = blockIdx.x;
  if (i >= n) return;
  snap_compact(t, i, x, row, num, entries, pitch, n, cutsq, maxn, nin, nj, nr, f64, f32, ntot, p32,
               p64, fdv);
}

// the neighbour parameters of one precision (and the repulsion F/r) from the
// compacted neighbours, for a U kernel of another precision than the one the
// compaction wrote
template <class T>
__global__ void __launch_bounds__(64)
k_snap_par(snap_dev_tables t, const int *__restrict__ nin, const double *__restrict__ nr,
           int maxn, int n, snap_par<T> *__restrict__ par, double *__restrict__ fdv) {
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int k = threadIdx.x; k < nin[i]; k += 64) {
    const size_t o = (size_t)i * maxn + k;
    snap_par_of<T>(t, nr[3 * o], nr[3 * o + 1], nr[3 * o + 2], par[o]);
    fdv[o] = snap_fdivr(nr[3 * o], nr[3 * o + 1], nr[3 * o + 2]);
  }
}

// layer entry owned by lane l in layer j: (ma, mb) with mb <= j/2
__device__ __forceinline__ bool lane_entry(int l, int j, int &ma, int &mb) {
  mb = l / (j + 1);
  ma = l - mb * (j + 1);
  return 2 * mb <= j;
}

// per-layer lane table, built once per block in LDS: for layer j and
// half-plane entry e (e = mb*(j+1) + ma, mb <= j/2) the indices (ma, mb), and
// the rootpq table; the per-neighbour layer loop then does no integer
// division and reads the recursion factors rootpq(j-ma, j-mb) and
// rootpq(ma, j-mb) from LDS
template <class T, int J>
struct layer_tab {
  static constexpr int JD = snap_dim<J>::JD, W = snap_dim<J>::HPW;
  int mm[JD][W];                // ma | mb << 8 | (entry exists) << 16
  T rp[JD * JD];                // rootpq(p, q) at p * JD + q
  __device__ T rpa(int j, int ma, int mb) const { return rp[(j - ma) * JD + (j - mb)]; }
  __device__ T rpb(int j, int ma, int mb) const { return rp[ma * JD + (j - mb)]; }
};

template <class T, int J>
__device__ __forceinline__ void layer_tab_fill(const snap_dev_tables &t, layer_tab<T, J> &L) {
  constexpr int JD = layer_tab<T, J>::JD, W = layer_tab<T, J>::W;
  for (int k = threadIdx.x; k < JD * W; k += blockDim.x) {
    const int j = k / W, e = k % W;
    const int mb = e / (j + 1), ma = e - mb * (j + 1);
    const bool own = j <= t.twojmax && e < snap_hp(j);
    L.mm[j][e] = ma | (mb << 8) | (own ? 1 << 16 : 0);
  }
  for (int k = threadIdx.x; k < JD * JD; k += blockDim.x)
    L.rp[k] = (T)t.rootpq[(k / JD) * SNAP_RPD + (k % JD)];
}

// one entry of layer j from the full layer j-1 (row-major ma*j+mb), with the
// recursion factors from the layer table:
// U(j,ma,mb) = [ma<j] rp(j-ma,j-mb) conj(a) U(j-1,ma,mb)
//            - [ma>0] rp(ma,j-mb)  conj(b) U(j-1,ma-1,mb)
// (SNA::compute_uarray, sna_impl.hpp:663-695)
template <class T>
__device__ __forceinline__ cplx<T> u_step_tab(const T *pr, const T *pi, int j, int ma, int mb,
                                              T rpa, T rpb, T a_r, T a_i, T b_r, T b_i) {
  cplx<T> u{0, 0};
  if (ma < j) {
    const T ur = pr[ma * j + mb], ui = pi[ma * j + mb];
    u.r += rpa * (a_r * ur + a_i * ui);
    u.i += rpa * (a_r * ui - a_i * ur);
  }
  if (ma > 0) {
    const T ur = pr[(ma - 1) * j + mb], ui = pi[(ma - 1) * j + mb];
    u.r += -rpb * (b_r * ur + b_i * ui);
    u.i += -rpb * (b_r * ui - b_i * ur);
  }
  return u;
}

// Cayley-Klein parameters a = (z0 - i z)/r0, b = (y - i x)/r0
template <class T>
__device__ __forceinline__ void ck_params(T x, T y, T z, T z0, T r, T &a_r, T &a_i, T &b_r,
                                          T &b_i) {
  const T r0inv = (T)1 / sqrt(r * r + z0 * z0);
  a_r = r0inv * z0;
  a_i = -r0inv * z;
  b_r = r0inv * y;
  b_i = -r0inv * x;
}

// ordering of LDS traffic among the lanes of one wavefront (no block barrier)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// U_tot of every atom for the size class twojmax <= 12 (SNA::compute_ui,
// sna_impl.hpp:152-190): one block of
// UI_WAVES wavefronts per atom, wavefront w takes neighbours w, w+UI_WAVES, ...
// (lanes over the half-plane entries of a layer, 64 at a time, the previous
// layer in the wavefront's own LDS buffer), then the partial sums of the
// wavefronts are added in wavefront order (deterministic)
constexpr int UI_WAVES = 4;

template <class T, int J>
__global__ void __launch_bounds__(64 * UI_WAVES)
k_snap_ui(snap_dev_tables t, const int *__restrict__ nin, const double *__restrict__ nr,
          int maxn, int n, T *__restrict__ utot_r, T *__restrict__ utot_i) {
  constexpr int JD = snap_dim<J>::JD, W = snap_dim<J>::HPW;
  __shared__ T s_r[UI_WAVES][2][JD * JD], s_i[UI_WAVES][2][JD * JD];
  __shared__ T s_pr[UI_WAVES][snap_dim<J>::NY], s_pi[UI_WAVES][snap_dim<J>::NY];
  __shared__ layer_tab<T, J> LT;
  const int i = blockIdx.x;
  if (i >= n) return;
  layer_tab_fill(t, LT);
  __syncthreads();
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  const int twojmax = t.twojmax;
  // this wavefront's sums over its neighbours, half-plane index
  // ystart[j] + e (only the lane of entry e ever touches it)
  T *acc_r = s_pr[w], *acc_i = s_pi[w];
  for (int k = l; k < t.ny; k += 64) acc_r[k] = acc_i[k] = 0;
  wave_sync();
  const int cnt = nin[i];
  T (*br)[JD * JD] = s_r[w];
  T (*bi)[JD * JD] = s_i[w];
  for (int jj = w; jj < cnt; jj += UI_WAVES) {
    const size_t o = (size_t)i * maxn + jj;
    const T x = (T)nr[3 * o], y = (T)nr[3 * o + 1], z = (T)nr[3 * o + 2];
    const T r = sqrt(x * x + y * y + z * z);
    const T theta0 = (r - (T)t.rmin0) * (T)t.rfac0 * (T)SNAP_PI / ((T)t.rcut - (T)t.rmin0);
    const T z0 = r / tan(theta0);
    T a_r, a_i, b_r, b_i;
    ck_params(x, y, z, z0, r, a_r, a_i, b_r, b_i);
    T sfac, dsfac;
    sfac_dsfac(t, r, sfac, dsfac);
    sfac *= (T)t.wj;
    // layer 0
    if (l == 0) {
      br[0][0] = 1;
      bi[0][0] = 0;
      acc_r[0] += sfac;
    }
    wave_sync();
    for (int j = 1; j <= twojmax; j++) {
      const int p = (j - 1) & 1, c = j & 1;
      for (int e = l; e < W; e += 64) {
        const int info = LT.mm[j][e];
        const bool own = info >> 16;
        const int ma = info & 255, mb = (info >> 8) & 255;
        if (own) {
          const cplx<T> u = u_step_tab<T>(br[p], bi[p], j, ma, mb, LT.rpa(j, ma, mb),
                                          LT.rpb(j, ma, mb), a_r, a_i, b_r, b_i);
          // full layer j: own entry and, off the middle column, its mirror
          br[c][ma * (j + 1) + mb] = u.r;
          bi[c][ma * (j + 1) + mb] = u.i;
          if (2 * mb != j) {
            const T sg = ((ma + mb) & 1) ? (T)-1 : (T)1;
            br[c][(j - ma) * (j + 1) + (j - mb)] = sg * u.r;
            bi[c][(j - ma) * (j + 1) + (j - mb)] = -sg * u.i;
          }
          acc_r[t.ystart[j] + e] += sfac * u.r;
          acc_i[t.ystart[j] + e] += sfac * u.i;
        }
      }
      wave_sync();
    }
  }
  __syncthreads();
  if (w != 0) return;
  // U_tot = wself * I + sum over wavefronts; own entries and mirrors
  T *ur = utot_r + (size_t)i * t.nu, *ui = utot_i + (size_t)i * t.nu;
  for (int j = 0; j <= twojmax; j++) {
    for (int e = l; e < W; e += 64) {
      const int info = LT.mm[j][e];
      const int ma = info & 255, mb = (info >> 8) & 255;
      if (info >> 16) {
        T vr = 0, vi = 0;
#pragma unroll
        for (int q = 0; q < UI_WAVES; q++) {
          vr += s_pr[q][t.ystart[j] + e];
          vi += s_pi[q][t.ystart[j] + e];
        }
        if (ma == mb) vr += (T)t.wself;
        ur[t.ustart[j] + ma * (j + 1) + mb] = vr;
        ui[t.ustart[j] + ma * (j + 1) + mb] = vi;
        if (2 * mb != j) {
          const T sg = ((ma + mb) & 1) ? (T)-1 : (T)1;
          ur[t.ustart[j] + (j - ma) * (j + 1) + (j - mb)] = sg * vr;
          ui[t.ustart[j] + (j - ma) * (j + 1) + (j - mb)] = -sg * vi;
        }
      }
    }
  }
}

// Z_{abJ}(ma,mb) = sum_ma1 cg(ma1,ma2) sum_mb1 cg(mb1,mb2) U_a(ma1,mb1) U_b(ma2,mb2)
// (SNA::compute_zi, sna_impl.hpp:221-274) from U_tot staged in LDS
template <class T>
__device__ __forceinline__ cplx<T> z_entry(const snap_dev_tables &t, const double *__restrict__ cg,
                                           const T *su_r, const T *su_i, int a, int b, int J,
                                           int cgoff, int ma, int mb) {
  cplx<T> z{0, 0};
  const int ma1lo = max(0, (2 * ma - J - b + a) / 2), ma1hi = min(a, (2 * ma - J + b + a) / 2);
  const int mb1lo = max(0, (2 * mb - J - b + a) / 2), mb1hi = min(a, (2 * mb - J + b + a) / 2);
  const int ua = t.ustart[a], ub = t.ustart[b];
  for (int ma1 = ma1lo; ma1 <= ma1hi; ma1++) {
    const int ma2 = (2 * ma - J - (2 * ma1 - a) + b) / 2;
    T s_r = 0, s_i = 0;
    for (int mb1 = mb1lo; mb1 <= mb1hi; mb1++) {
      const int mb2 = (2 * mb - J - (2 * mb1 - a) + b) / 2;
      const T c = (T)cg[cgoff + mb1 * (b + 1) + mb2];
      const T u1r = su_r[ua + ma1 * (a + 1) + mb1], u1i = su_i[ua + ma1 * (a + 1) + mb1];
      const T u2r = su_r[ub + ma2 * (b + 1) + mb2], u2i = su_i[ub + ma2 * (b + 1) + mb2];
      s_r += c * (u1r * u2r - u1i * u2i);
      s_i += c * (u1r * u2i + u1i * u2r);
    }
    const T c = (T)cg[cgoff + ma1 * (b + 1) + ma2];
    z.r += s_r * c;
    z.i += s_i * c;
  }
  return z;
}

// half-plane entry e (0..ny-1) -> (J, ma, mb)
__device__ __forceinline__ void y_entry(const snap_dev_tables &t, int e, int &J, int &ma,
                                        int &mb) {
  J = 0;
  while (J < t.twojmax && e >= t.ystart[J + 1]) J++;
  const int k = e - t.ystart[J];
  mb = k / (J + 1);
  ma = k - mb * (J + 1);
}

// Y of AT atoms per block from a flat term stream.  Every half-plane entry
// e = (J, ma, mb) is the sum over its (a, b) keys and (ma1, mb1) of
//   w * U_a(ma1, mb1) * U_b(ma2, mb2),  w = coef(a,b,J) cg(ma1,ma2) cg(mb1,mb2)
// (the double sum of SNA::compute_zi, sna_impl.hpp:221-274, weighted as in
// compute_dbidrj's three roles).  The host lists these products (zero
// Clebsch-Gordan weights dropped) entry by entry, pads each entry's list with
// zero-weight terms to whole chunks of YF_CHUNK, and gives thread t the
// chunks [t P, (t+1) P) of that sequence, interleaved so that step s of every
// thread is one coalesced load.  Chunk boundaries are the same for every
// lane, so the term loop carries no per-term branch: a term is two LDS byte
// offsets (its code), a weight and six multiply-adds.  Once per chunk a
// thread whose run of one entry's chunks ends stores the run's sum in LDS;
// each entry adds its runs in thread order (fixed order: deterministic).
// With AT atoms per block (U_tot of each in LDS) every code and weight load
// serves AT atoms: with the LDS reads free of bank conflicts (below) the
// stream of codes and weights from L2 is what bounds the kernel.
//
// LDS banks: the LDS serves 256 bytes per cycle, so a wavefront's
// ds_read_b64 is served 32 lanes per cycle, and two lanes of one cycle that
// read different 8-byte words at the same bank residue (word index mod 32)
// are served in turn.  (Measured: 5.15 LDS-active cycles per wavefront read
// of an earlier FP32 term order, for which this model gives 5.19.)  A
// ds_read_b128 of 16-byte FP64 pairs measured 8 cycles: its 8-byte halves
// are served like two b64 reads, and a 16-byte stride leaves only 16
// residues.  So FP32 U_tot is stored as (re, im) pairs and FP64 U_tot as an
// array of real parts and one of imaginary parts, each read with b64 loads
// at word residue u mod 32.  For twojmax <= 8 U_tot is staged four times,
// copy c at word offset c * stride + shift(c), and the host gives every
// read of every term the copy whose residue no other address of its
// 32-lane group takes (a matching per group; with the shifts 0, 5, 11, 19
// about 4 % of the groups keep one conflict).
// threads per Y block (the stream is split this many ways): SNAP-2k 25.3 /
// 43.7 us (mixed / FP64) against 27.0 / 47.3 with 256 and 28.7 / 47.2 with
// 1024 (more padding, fewer blocks per CU)
constexpr int YF_THREADS = 512;
constexpr int YF_CHUNK = 16;

template <class T>
struct alignas(2 * sizeof(T)) cpair {
  T r, i;
};

// U_tot copies in the Y kernel's LDS, in 8-byte words: an FP32 pair is one
// word, an FP64 entry a word in the real array and one in the imaginary
// array (IMB bytes further)
template <class T, int J>
struct yf_layout {
  static constexpr bool SOA = sizeof(T) == 8;
  static constexpr int NCOPY = J <= SNAP_MAXJ ? 4 : 1;
  static constexpr int MOD = 32;    // bank residues of an 8-byte word
  static constexpr int GROUP = 32;  // lanes served per cycle
  static constexpr int STRIDE = (snap_dim<J>::NU + MOD - 1) / MOD * MOD;
  __host__ __device__ static constexpr int shift(int c) {
    return NCOPY == 1 ? 0 : (c == 0 ? 0 : c == 1 ? 5 : c == 2 ? 11 : 19);
  }
  __host__ __device__ static constexpr int base(int c) { return c * STRIDE + shift(c); }
  static constexpr int NUC = base(NCOPY - 1) + snap_dim<J>::NU;  // words the copies span
  static constexpr int NUC_PAD = (NUC + MOD - 1) / MOD * MOD;
  static constexpr int IMB = SOA ? NUC_PAD * 8 : 0;            // imaginary array (FP64)
  static constexpr int UBYTES = SOA ? 2 * NUC_PAD * 8 : NUC_PAD * 8;
};

// code: byte offsets of the two U_tot entries in LDS, 16 bits each (the
// host writes offsets for sizeof(cpair<T>)); meta[c][t]: the LDS slot of the
// run that ends with thread t's chunk c, or YF_NO_RUN.  A thread's runs are
// the entries its chunk range touches, so there are at most ny + threads - 1.
constexpr unsigned short YF_NO_RUN = 0xFFFFu;
template <int J>
constexpr int yf_maxruns() { return snap_dim<J>::NY + YF_THREADS; }
static_assert(yf_layout<double, SNAP_JBIG>::NUC * 8 < 65536 &&
                  yf_layout<float, SNAP_JBIG>::NUC * 8 < 65536,
              "U_tot byte offsets fit 16 bits");

// The term stream of AT atoms whose U_tot copies are in LDS at sb (a copy
// set of yf_layout<T,J>::UBYTES per atom), run sums in part: Y of atoms
// i0 .. i0 + AT - 1 (those below n)
template <class T, int J, int AT>
__device__ __forceinline__ void yf_stream(const unsigned char *sb, cpair<T> *part, int i0, int n,
                                          int ny, int nch, const uint32_t *__restrict__ code,
                                          const T *__restrict__ w,
                                          const unsigned short *__restrict__ meta,
                                          const int *__restrict__ rstart, T *__restrict__ y_r,
                                          T *__restrict__ y_i) {
  using L = yf_layout<T, J>;
  T ar[AT], ai[AT];
#pragma unroll
  for (int a = 0; a < AT; a++) ar[a] = ai[a] = 0;
  for (int c = 0; c < nch; c++) {
    const unsigned short m = meta[(size_t)c * YF_THREADS + threadIdx.x];
    // the chunk's loads are issued before its arithmetic
    uint32_t cc[YF_CHUNK];
    T ww[YF_CHUNK];
#pragma unroll
    for (int k = 0; k < YF_CHUNK; k++) {
      const size_t slot = ((size_t)c * YF_CHUNK + k) * YF_THREADS + threadIdx.x;
      cc[k] = code[slot];
      ww[k] = w[slot];
    }
#pragma unroll
    for (int k = 0; k < YF_CHUNK; k++) {
#pragma unroll
      for (int a = 0; a < AT; a++) {
        const unsigned char *ub = sb + a * L::UBYTES;
        cpair<T> u1, u2;
        if (L::SOA) {
          u1.r = *reinterpret_cast<const T *>(ub + (cc[k] & 0xFFFFu));
          u1.i = *reinterpret_cast<const T *>(ub + L::IMB + (cc[k] & 0xFFFFu));
          u2.r = *reinterpret_cast<const T *>(ub + (cc[k] >> 16));
          u2.i = *reinterpret_cast<const T *>(ub + L::IMB + (cc[k] >> 16));
        } else {
          u1 = *reinterpret_cast<const cpair<T> *>(ub + (cc[k] & 0xFFFFu));
          u2 = *reinterpret_cast<const cpair<T> *>(ub + (cc[k] >> 16));
        }
        const T pr = u1.r * u2.r - u1.i * u2.i;
        const T pi = u1.r * u2.i + u1.i * u2.r;
        ar[a] += ww[k] * pr;
        ai[a] += ww[k] * pi;
      }
    }
    if (m != YF_NO_RUN) {
#pragma unroll
      for (int a = 0; a < AT; a++) {
        part[a * yf_maxruns<J>() + m] = cpair<T>{ar[a], ai[a]};
        ar[a] = 0;
        ai[a] = 0;
      }
    }
  }
  __syncthreads();
  for (int a = 0; a < AT; a++) {
    const int i = i0 + a;
    if (i >= n) break;
    for (int e = threadIdx.x; e < ny; e += YF_THREADS) {
      T yr = 0, yi = 0;
      for (int r = rstart[e]; r < rstart[e + 1]; r++) {
        yr += part[a * yf_maxruns<J>() + r].r;
        yi += part[a * yf_maxruns<J>() + r].i;
      }
      y_r[(size_t)i * ny + e] = yr;
      y_i[(size_t)i * ny + e] = yi;
    }
  }
}

template <class T, int J, int AT>
__global__ void __launch_bounds__(YF_THREADS)
k_snap_yflat(int nu, int ny, int nch, const uint32_t *__restrict__ code,
             const T *__restrict__ w, const unsigned short *__restrict__ meta,
             const int *__restrict__ rstart, int n, const T *__restrict__ utot_r,
             const T *__restrict__ utot_i, T *__restrict__ y_r, T *__restrict__ y_i) {
  // the copies of U_tot from byte 0 (a static array: its LDS address is
  // known when the kernel is compiled, so a code's offset is the whole LDS
  // address), then the run sums
  using L = yf_layout<T, J>;
  static_assert(L::UBYTES % sizeof(cpair<T>) == 0, "run sums aligned");
  constexpr int UP = L::UBYTES / sizeof(cpair<T>);  // pairs per atom's U_tot copies
  __shared__ cpair<T> s_yf[AT * (UP + yf_maxruns<J>())];
  unsigned char *sb = reinterpret_cast<unsigned char *>(s_yf);
  const int i0 = blockIdx.x * AT;
  if (i0 >= n) return;
  for (int a = 0; a < AT; a++) {
    const int i = i0 + a < n ? i0 + a : i0;
    unsigned char *ub = sb + a * L::UBYTES;
    for (int k = threadIdx.x; k < nu; k += YF_THREADS) {
      const T vr = utot_r[(size_t)i * nu + k], vi = utot_i[(size_t)i * nu + k];
#pragma unroll
      for (int c = 0; c < L::NCOPY; c++) {
        if (L::SOA) {
          reinterpret_cast<T *>(ub)[L::base(c) + k] = vr;
          reinterpret_cast<T *>(ub + L::IMB)[L::base(c) + k] = vi;
        } else {
          reinterpret_cast<cpair<T> *>(ub)[L::base(c) + k] = cpair<T>{vr, vi};
        }
      }
    }
  }
  __syncthreads();
  yf_stream<T, J, AT>(sb, s_yf + AT * UP, i0, n, ny, nch, code, w, meta, rstart, y_r, y_i);
}

// term weights for the current coefficients: w = coef(key) * cg product
__global__ void k_yf_weights(int nslots, const unsigned short *__restrict__ key,
                             const double *__restrict__ cgw, const double *__restrict__ kcoef,
                             double *__restrict__ w64, float *__restrict__ w32) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const double v = kcoef[key[s]] * cgw[s];
  w64[s] = v;
  w32[s] = (float)v;
}

// F_ij of the size class J = 12, one (atom, neighbour) pair per wavefront
// (SNA::compute_duidrj + compute_duarray, sna_impl.hpp:290-321 / 728-893,
// then the Y contraction): the lanes take the half-plane entries of a layer
// 64 at a time (layers below the top have up to 72 entries, two steps), the
// previous layer of U and dU in LDS.  (For twojmax <= 8 the register-resident
// k_snap_fij_rev replaces it.)
constexpr int FIJ_WAVES = 2;  // wavefronts per block

template <class T, class TF, int J>
__global__ void __launch_bounds__(64 * FIJ_WAVES)
k_snap_fij(snap_dev_tables t, const int *__restrict__ nin, const int *__restrict__ nj,
           const double *__restrict__ nr, int maxn, int n, const T *__restrict__ y_r,
           const T *__restrict__ y_i, TF *__restrict__ f) {
  constexpr int LG = 64;                      // lanes per neighbour
  constexpr int NBW = 64 / LG;                // neighbours per wavefront
  constexpr int FIJ_NB = FIJ_WAVES * NBW;     // neighbours per block
  constexpr int FIJ_HP = snap_dim<J>::HPPREV;  // half-plane entries of layer J-1
  // a layer below the top takes at most two steps (kept in registers below)
  static_assert((FIJ_HP + LG - 1) / LG <= 2, "two steps per layer below the top");
  static_assert(((snap_hp(J) + LG - 1) / LG) * LG <= snap_dim<J>::HPW, "lane table width");
  // previous layer (j-1 <= 7) of each neighbour: U and dU (3 components) of
  // its half plane (mb <= (j-1)/2, at mb*j + ma: at most 32 entries); the
  // middle column of an even layer reads its mirror image
  // U(j-1-ma, j-1-mb) = (-1)^(ma+mb) conj U(ma, mb) (the same symmetry holds
  // for dU), so no lane writes a mirror entry.  One buffer: a layer's
  // entries are kept in registers until every lane has read the previous
  // layer (a layer below the top has at most 32 entries, two steps of 16
  // lanes), then written over it.
  __shared__ T s_u[FIJ_NB][8][FIJ_HP];
  __shared__ T s_y[2][snap_dim<J>::NY];
  __shared__ layer_tab<T, J> LT;
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  const int hs = l / LG, hl = l % LG;  // neighbour slot of the wavefront, lane in it
  const int i = blockIdx.x;
  if (i >= n) return;
  const int cnt_i = nin[i];
  if (blockIdx.y * FIJ_NB >= cnt_i) return;  // no neighbour for this block
  // Y of atom i, read by every layer of every neighbour of the block
  for (int k = threadIdx.x; k < t.ny; k += blockDim.x) {
    s_y[0][k] = y_r[(size_t)i * t.ny + k];
    s_y[1][k] = y_i[(size_t)i * t.ny + k];
  }
  layer_tab_fill(t, LT);
  __syncthreads();
  const int jj0 = blockIdx.y * FIJ_NB + NBW * w;  // the wavefront's first neighbour
  if (jj0 >= cnt_i) return;                       // uniform over the wavefront
  const int jj = jj0 + hs;
  const bool act = jj < cnt_i;
  const size_t o = (size_t)i * maxn + (act ? jj : jj0);
  const double xd = nr[3 * o], yd = nr[3 * o + 1], zd = nr[3 * o + 2];
  const T x = (T)xd, y = (T)yd, z = (T)zd;
  const T rsq = x * x + y * y + z * z;
  const T r = sqrt(rsq);
  const T rscale0 = (T)t.rfac0 * (T)SNAP_PI / ((T)t.rcut - (T)t.rmin0);
  const T theta0 = (r - (T)t.rmin0) * rscale0;
  const T cs = cos(theta0), sn = sin(theta0);
  const T z0 = r * cs / sn;
  const T dz0dr = z0 / r - (r * rscale0) * (rsq + z0 * z0) / rsq;
  const T rinv = (T)1 / r;
  const T u3[3] = {x * rinv, y * rinv, z * rinv};
  const T r0inv = (T)1 / sqrt(r * r + z0 * z0);
  const T a_r = z0 * r0inv, a_i = -z * r0inv, b_r = y * r0inv, b_i = -x * r0inv;
  const T dr0invdr = -(r0inv * r0inv * r0inv) * (r + z0 * dz0dr);
  T da_r[3], da_i[3], db_r[3], db_i[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const T dr0 = dr0invdr * u3[k], dz = dz0dr * u3[k];
    da_r[k] = dz * r0inv + z0 * dr0;
    da_i[k] = -z * dr0;
    db_r[k] = y * dr0;
    db_i[k] = -x * dr0;
  }
  da_i[2] += -r0inv;
  db_i[0] += -r0inv;
  db_r[1] += r0inv;
  T sfac, dsfac;
  sfac_dsfac(t, r, sfac, dsfac);
  sfac *= (T)t.wj;
  dsfac *= (T)t.wj;

  T (*buf)[FIJ_HP] = s_u[NBW * w + hs];
  const T *yr = s_y[0], *yi = s_y[1];
  T fx = 0, fy = 0, fz = 0;
  // layer 0: U = 1, dU = 0; its dU_tot = dsfac * u; contraction with Y_0
  if (hl == 0) {
    buf[0][0] = 1; buf[1][0] = 0;
    for (int k = 0; k < 3; k++) { buf[2 + 2 * k][0] = 0; buf[3 + 2 * k][0] = 0; }
    const T w0 = (T)0.5;  // J = 0: the single entry is the middle-column diagonal
    const T c = dsfac * w0 * yr[0];
    fx += c * u3[0];
    fy += c * u3[1];
    fz += c * u3[2];
  }
  wave_sync();
  constexpr int FIJ_STEPS = (snap_hp(J) + LG - 1) / LG;  // steps of the top layer
  for (int j = 1; j <= t.twojmax; j++) {
    const int ne = (j + 1) * (j / 2 + 1);
    T keep[2][8];  // this layer's own entries of steps 0 and 1
#pragma unroll
    for (int st = 0; st < FIJ_STEPS; st++) {
      const int s0 = st * LG;
      if (s0 >= ne) break;
      const int e = s0 + hl;
      const int info = LT.mm[j][e];
      const int ma = info & 255, mb = (info >> 8) & 255;
      const bool own = info >> 16;
      T ur = 0, ui = 0, dr[3] = {0, 0, 0}, di[3] = {0, 0, 0};
      if (own) {
        const T *P = &buf[0][0];
        // P[q*FIJ_HP + idx]: q = 0 U_r, 1 U_i, 2+2k dU_r[k], 3+2k dU_i[k].
        // Layer j-1 entries (ma, mb) and (ma-1, mb); on the middle column of
        // an even layer (mb = j/2 > (j-1)/2) they are read as mirrors.
        const bool mir = 2 * mb == j;
        const T sgn = ((ma + mb) & 1) ? (T)-1 : (T)1;
        const T s_r = mir ? sgn : (T)1, s_i = mir ? -sgn : (T)1;  // (ma, mb); (ma-1, mb) flips sgn
        if (ma < j) {
          const T rp = LT.rpa(j, ma, mb);
          const int sidx = mir ? (j - 1 - mb) * j + (j - 1 - ma) : mb * j + ma;
          const T pr = s_r * P[sidx], pi = s_i * P[FIJ_HP + sidx];
          ur += rp * (a_r * pr + a_i * pi);
          ui += rp * (a_r * pi - a_i * pr);
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const T dpr = s_r * P[(2 + 2 * k) * FIJ_HP + sidx];
            const T dpi = s_i * P[(3 + 2 * k) * FIJ_HP + sidx];
            dr[k] += rp * (da_r[k] * pr + da_i[k] * pi + a_r * dpr + a_i * dpi);
            di[k] += rp * (da_r[k] * pi - da_i[k] * pr + a_r * dpi - a_i * dpr);
          }
        }
        if (ma > 0) {
          const T rp = LT.rpb(j, ma, mb);
          const int sidx = mir ? (j - 1 - mb) * j + (j - ma) : mb * j + ma - 1;
          const T tr = mir ? -s_r : (T)1, ti = mir ? -s_i : (T)1;
          const T pr = tr * P[sidx], pi = ti * P[FIJ_HP + sidx];
          ur += -rp * (b_r * pr + b_i * pi);
          ui += -rp * (b_r * pi - b_i * pr);
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const T dpr = tr * P[(2 + 2 * k) * FIJ_HP + sidx];
            const T dpi = ti * P[(3 + 2 * k) * FIJ_HP + sidx];
            dr[k] += -rp * (db_r[k] * pr + db_i[k] * pi + b_r * dpr + b_i * dpi);
            di[k] += -rp * (db_r[k] * pi - db_i[k] * pr + b_r * dpi - b_i * dpr);
          }
        }
      }
      if (st < 2) {
        keep[st][0] = ur; keep[st][1] = ui;
#pragma unroll
        for (int k = 0; k < 3; k++) { keep[st][2 + 2 * k] = dr[k]; keep[st][3 + 2 * k] = di[k]; }
      }
      if (own) {
        // half-plane weight of (ma, mb) in layer j
        T wgt = 1;
        if (2 * mb == j) wgt = (ma < mb) ? (T)1 : ((ma == mb) ? (T)0.5 : (T)0);
        if (wgt != (T)0) {
          const int ey = t.ystart[j] + mb * (j + 1) + ma;
          const T Yr = wgt * yr[ey], Yi = wgt * yi[ey];
          const T uy = ur * Yr + ui * Yi;   // Re(conj(U) Y) (switching-derivative term)
          const T g[3] = {dr[0] * Yr + di[0] * Yi, dr[1] * Yr + di[1] * Yi, dr[2] * Yr + di[2] * Yi};
          fx += dsfac * uy * u3[0] + sfac * g[0];
          fy += dsfac * uy * u3[1] + sfac * g[1];
          fz += dsfac * uy * u3[2] + sfac * g[2];
        }
      }
    }
    wave_sync();  // every lane has read layer j-1
    if (j < t.twojmax) {
      // own half-plane entries only (ne <= 32 below the top layer)
#pragma unroll
      for (int st = 0; st < 2; st++) {
        const int e = st * LG + hl;
        if (st * LG < ne && (LT.mm[j][e] >> 16)) {
#pragma unroll
          for (int q = 0; q < 8; q++) buf[q][e] = keep[st][q];
        }
      }
    }
    wave_sync();
  }
  // sum of the force components over the neighbour's lanes
#pragma unroll
  for (int off = LG / 2; off > 0; off >>= 1) {
    fx += __shfl_xor(fx, off, LG);
    fy += __shfl_xor(fy, off, LG);
    fz += __shfl_xor(fz, off, LG);
  }
  if (hl == 0 && act) {
    // ad-hoc repulsion of the reference, force_snap_neigh_impl.h:705-711
    const double rsqd = xd * xd + yd * yd + zd * zd;
    const double fdivr = -1.5e6 / pow(rsqd, 7.0);
    const TF Fx = (TF)fx + (TF)(xd * fdivr);
    const TF Fy = (TF)fy + (TF)(yd * fdivr);
    const TF Fz = (TF)fz + (TF)(zd * fdivr);
    const int j = nj[o];
    atomicAdd(&f[3 * (size_t)i], Fx);
    atomicAdd(&f[3 * (size_t)i + 1], Fy);
    atomicAdd(&f[3 * (size_t)i + 2], Fz);
    atomicAdd(&f[3 * (size_t)j], -Fx);
    atomicAdd(&f[3 * (size_t)j + 1], -Fy);
    atomicAdd(&f[3 * (size_t)j + 2], -Fz);
  }
}

// ------------------------------------------- register-resident recursion
// For twojmax <= 8 a neighbour's layer fits nine lanes: lane ma (0..8) of
// the neighbour's group holds the entries (ma, mb) of every half-plane
// column mb (0..4) of the current layer in registers.  The recursion
//   U_j(ma,mb) = rpa conj(a) U_{j-1}(ma,mb) - rpb conj(b) U_{j-1}(ma-1,mb)
// reads U_{j-1}(ma,mb) from the lane's own register and U_{j-1}(ma-1,mb)
// from the lane below (a DPP shift of the whole wavefront by one lane; at
// ma = 0 rpb is 0 and the value from the group below is not used).  The
// middle column mb = j/2 of an even layer needs U_{j-1}(ma, j/2), outside
// the previous half plane: it is the mirror (-1)^(ma+j/2) conj
// U_{j-1}(j-1-ma, j/2-1), fetched from lane j-1-ma of the group by one
// permute per value.  Seven neighbours per wavefront (lane 63 idle); no LDS
// traffic for the layers and no barriers between them.  The recursion
// factors and the half-plane-weighted Y of each (column, ma) come from one
// LDS record, zero where the entry does not exist (ma > j), so lanes past
// the layer compute exact zeros.
constexpr int RG_W = SNAP_MAXJ + 1;  // lanes per neighbour (ma = 0..8)
constexpr int RG_NB = 64 / RG_W;     // neighbours per wavefront: 7
constexpr int RG_COLS = SNAP_MAXJ / 2 + 1;
// first column slot of layer j: sum over layers j' < j of (j'/2 + 1) columns
__host__ __device__ constexpr int rg_slot0(int j) { return j <= 0 ? 0 : j + (j - 1) * (j - 1) / 4; }
constexpr int RG_SLOTS = rg_slot0(SNAP_MAXJ + 1);  // (layer, column) pairs of layers 0..8: 25
static_assert(rg_slot0(1) == 1 && rg_slot0(2) == 2 && rg_slot0(3) == 4 && rg_slot0(8) == 20 &&
                  RG_SLOTS == 25,
              "column slots");

template <class T>
struct alignas(4 * sizeof(T)) rg_rec {
  T rpa, rpb, yr, yi;  // recursion factors and weighted Y of (j, ma, mb)
};

// value of lane l - 1 (0 in lane 0): DPP wavefront shift right by one
// (bound_ctrl: the lane without a source reads 0)
__device__ __forceinline__ float lane_below(float v) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ double lane_below(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x138, 0xf, 0xf, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x138, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Per (column slot s, lane ma), k = s * RG_W + ma, built on the host once:
// the recursion factors (zero where the entry does not exist), and
//   ey   the half-plane index of Y (or -1), wgt its half-plane weight
//   uo   the U_tot index of (ma, mb) (or -1), um that of the mirror
//        (j-ma, j-mb) off the middle column (or -1), neg its sign (-1)^(ma+mb)
//        and diag for ma == mb (wself)
template <class T>
struct rg_rp {
  T rpa, rpb;
};
struct rg_map {
  int ey, uo, um;
  unsigned char wq, neg, diag, pad;  // wgt = wq / 2
};

// the records of one atom: column slot s = rg_slot0(j) + mb, lane ma
template <class T, int NT>
__device__ __forceinline__ void rg_fill(const rg_rp<T> *__restrict__ rp,
                                        const rg_map *__restrict__ map, int ny,
                                        const T *__restrict__ y_r, const T *__restrict__ y_i,
                                        int i, rg_rec<T> *rec) {
#pragma unroll
  for (int q = 0; q < (RG_SLOTS * RG_W + NT - 1) / NT; q++) {
    const int k = q * NT + (int)threadIdx.x;
    if (k < RG_SLOTS * RG_W) {
      const rg_rp<T> r = rp[k];
      T yr = 0, yi = 0;
      if (y_r) {
        const rg_map m = map[k];
        if (m.ey >= 0) {
          const T wgt = (T)0.5 * (T)m.wq;
          yr = wgt * y_r[(size_t)i * ny + m.ey];
          yi = wgt * y_i[(size_t)i * ny + m.ey];
        }
      }
      rec[k] = rg_rec<T>{r.rpa, r.rpb, yr, yi};
    }
  }
}

// Sums over the neighbours base0, base0 + stride, ... (seven at a time) of
// atom i of sfac * U for the lane's entries (ma, mb) of every column of every
// layer (SNA::compute_ui, sna_impl.hpp:152-190, with the register layout
// above), added to acc[column slot]
template <class T>
__device__ __forceinline__ void rg_u_sums(const snap_dev_tables &t, const rg_rec<T> *s_rec,
                                          const snap_par<T> *__restrict__ par, int i, int maxn,
                                          int cnt, int base0, int stride, int g, int ma,
                                          T (&acc)[RG_SLOTS][2]) {
  for (int base = base0; base < cnt; base += stride) {
    const int jj = base + g;
    const bool act = g < RG_NB && jj < cnt;
    const size_t o = (size_t)i * maxn + (act ? jj : base);
    const T a_r = par[o].a_r, a_i = par[o].a_i, b_r = par[o].b_r, b_i = par[o].b_i;
    const T sfac = act ? par[o].sfac : (T)0;  // an idle group adds nothing
    T c[RG_COLS][2];
#pragma unroll
    for (int mb = 0; mb < RG_COLS; mb++) c[mb][0] = c[mb][1] = 0;
    c[0][0] = ma == 0 ? (T)1 : (T)0;  // layer 0: U = 1
    acc[0][0] += sfac * c[0][0];
#pragma unroll
    for (int j = 1; j <= SNAP_MAXJ; j++) {
      if (j > t.twojmax) break;
      T mir[2], mirb[2];  // even j: U_{j-1}(ma, j/2) and U_{j-1}(ma-1, j/2)
      if ((j & 1) == 0) {
        const int src = g * RG_W + (j - 1 - ma > 0 ? j - 1 - ma : 0);
        const T sg = ((ma + j / 2) & 1) ? (T)-1 : (T)1;
        mir[0] = sg * __shfl(c[j / 2 - 1][0], src, 64);
        mir[1] = -sg * __shfl(c[j / 2 - 1][1], src, 64);
        mirb[0] = lane_below(mir[0]);
        mirb[1] = lane_below(mir[1]);
      }
#pragma unroll
      for (int mb = 0; mb <= j / 2; mb++) {
        const int slot = rg_slot0(j) + mb;
        const rg_rec<T> q = s_rec[slot * RG_W + ma];
        const bool mid = (j & 1) == 0 && mb == j / 2;
        const T p0 = mid ? mir[0] : c[mb][0], p1 = mid ? mir[1] : c[mb][1];
        const T q0 = mid ? mirb[0] : lane_below(c[mb][0]);
        const T q1 = mid ? mirb[1] : lane_below(c[mb][1]);
        const T ur = q.rpa * (a_r * p0 + a_i * p1) - q.rpb * (b_r * q0 + b_i * q1);
        const T ui = q.rpa * (a_r * p1 - a_i * p0) - q.rpb * (b_r * q1 - b_i * q0);
        c[mb][0] = ur;
        c[mb][1] = ui;
        acc[slot][0] += sfac * ur;
        acc[slot][1] += sfac * ui;
      }
    }
  }
}

// acc summed over the seven groups of the wavefront (in TA) by a fixed tree
// of lane shifts (groups 4-6 onto 0-2, then 2-3 onto 0-1, then 1 onto 0);
// lanes 0..8 store the sums at out[column slot * RG_W + lane]
template <class T, class TA>
__device__ __forceinline__ void rg_group_sum(T (&acc)[RG_SLOTS][2], int l, TA *out_r, TA *out_i) {
#pragma unroll
  for (int s = 0; s < RG_SLOTS; s++) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
      TA v = (TA)acc[s][q];
      TA u = __shfl_down(v, 4 * RG_W, 64);
      if (l < 3 * RG_W) v += u;
      u = __shfl_down(v, 2 * RG_W, 64);
      if (l < 2 * RG_W) v += u;
      u = __shfl_down(v, RG_W, 64);
      if (l < RG_W) {
        v += u;
        (q ? out_i : out_r)[s * RG_W + l] = v;
      }
    }
  }
}

// U_tot with the register-resident recursion (twojmax <= 8): a block of
// UR_WAVES wavefronts per atom, wavefront w over the neighbours 7w, 7w +
// 7 UR_WAVES, ...; the lane sums added over the groups and then over the
// wavefronts in order: deterministic.
template <class T, int UR_WAVES>
__global__ void __launch_bounds__(64 * UR_WAVES)
k_snap_ui_rg(snap_dev_tables t, const rg_rp<T> *__restrict__ rgp, const rg_map *__restrict__ rgm,
             const int *__restrict__ nin, const snap_par<T> *__restrict__ par, int maxn, int n,
             T *__restrict__ utot_r, T *__restrict__ utot_i) {
  __shared__ rg_rec<T> s_rec[RG_SLOTS * RG_W];
  __shared__ T s_acc[UR_WAVES][2][RG_SLOTS * RG_W];
  const int i = blockIdx.x;
  if (i >= n) return;
  rg_fill<T, 64 * UR_WAVES>(rgp, rgm, 0, nullptr, nullptr, i, s_rec);
  __syncthreads();
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  const int g = l / RG_W, ma = l - g * RG_W;  // g = 7 only for the idle lane 63
  T acc[RG_SLOTS][2];
#pragma unroll
  for (int s = 0; s < RG_SLOTS; s++) acc[s][0] = acc[s][1] = 0;
  rg_u_sums<T>(t, s_rec, par, i, maxn, nin[i], w * RG_NB, UR_WAVES * RG_NB, g, ma, acc);
  rg_group_sum<T, T>(acc, l, s_acc[w][0], s_acc[w][1]);
  __syncthreads();
  // U_tot = wself * I + the wavefront sums in order; own entries and mirrors
  T *ur = utot_r + (size_t)i * t.nu, *ui = utot_i + (size_t)i * t.nu;
  for (int k = threadIdx.x; k < RG_SLOTS * RG_W; k += blockDim.x) {
    const rg_map m = rgm[k];
    if (m.uo < 0) continue;
    T vr = 0, vi = 0;
#pragma unroll
    for (int q = 0; q < UR_WAVES; q++) {
      vr += s_acc[q][0][k];
      vi += s_acc[q][1][k];
    }
    if (m.diag) vr += (T)t.wself;
    ur[m.uo] = vr;
    ui[m.uo] = vi;
    if (m.um >= 0) {
      const T sg = m.neg ? (T)-1 : (T)1;
      ur[m.um] = sg * vr;
      ui[m.um] = -sg * vi;
    }
  }
}

// The compaction of k_snap_neigh (with the neighbour parameters in T's
// precision) followed by U_tot in the same wavefront (twojmax <= 8, one
// wavefront per atom as k_snap_ui_rg<T, 1>): one launch less per force
// evaluation, and the parameters just written are read back from the cache.
template <class T>
__global__ void __launch_bounds__(64)
k_snap_neigh_ui(snap_dev_tables t, const rg_rp<T> *__restrict__ rgp, const rg_map *__restrict__ rgm,
                const double *__restrict__ x, const int *__restrict__ row,
                const int *__restrict__ num, const int *__restrict__ entries, int pitch, int n,
                double cutsq, int maxn, int *__restrict__ nin, int *__restrict__ nj,
                double *__restrict__ nr, double *__restrict__ f64, float *__restrict__ f32,
                int ntot, snap_par<T> *__restrict__ par, double *__restrict__ fdv,
                T *__restrict__ utot_r, T *__restrict__ utot_i) {
  __shared__ rg_rec<T> s_rec[RG_SLOTS * RG_W];
  __shared__ T s_acc[2][RG_SLOTS * RG_W];
  const int i = blockIdx.x;
  if (i >= n) return;
  rg_fill<T, 64>(rgp, rgm, 0, nullptr, nullptr, i, s_rec);
  snap_par<float> *p32 = nullptr;
  snap_par<double> *p64 = nullptr;
  if (sizeof(T) == 4) p32 = reinterpret_cast<snap_par<float> *>(par);
  else p64 = reinterpret_cast<snap_par<double> *>(par);
  const int cnt = snap_compact(t, i, x, row, num, entries, pitch, n, cutsq, maxn, nin, nj, nr, f64,
                               f32, ntot, p32, p64, fdv);
  __syncthreads();  // the wavefront's parameter writes, and the records, are visible
  const int l = threadIdx.x;
  const int g = l / RG_W, ma = l - g * RG_W;
  T acc[RG_SLOTS][2];
#pragma unroll
  for (int q = 0; q < RG_SLOTS; q++) acc[q][0] = acc[q][1] = 0;
  rg_u_sums<T>(t, s_rec, par, i, maxn, cnt, 0, RG_NB, g, ma, acc);
  rg_group_sum<T, T>(acc, l, s_acc[0], s_acc[1]);
  __syncthreads();
  T *ur = utot_r + (size_t)i * t.nu, *ui = utot_i + (size_t)i * t.nu;
  for (int k = l; k < RG_SLOTS * RG_W; k += 64) {
    const rg_map m = rgm[k];
    if (m.uo < 0) continue;
    T vr = s_acc[0][k];
    const T vi = s_acc[1][k];
    if (m.diag) vr += (T)t.wself;
    ur[m.uo] = vr;
    ui[m.uo] = vi;
    if (m.um >= 0) {
      const T sg = m.neg ? (T)-1 : (T)1;
      ur[m.um] = sg * vr;
      ui[m.um] = -sg * vi;
    }
  }
}

// Blocks of RG_WAVES wavefronts take 7 * RG_WAVES neighbours of one atom in
// the force kernel (grid.y over the neighbour slots).
constexpr int RG_WAVES = 4;

// value of lane l + 1 (0 in lane 63): DPP wavefront shift left by one
__device__ __forceinline__ float lane_above(float v) {
  return __builtin_bit_cast(float,
                            __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ double lane_above(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, 0x130, 0xf, 0xf, true);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x130, 0xf, 0xf, true);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// F_ij in reverse mode.  With G_j(e) = sfac * w(e) Y_j(e) (w the half-plane
// weight) the force on the pair is the gradient of
//   L = sum_j sum_e Re(conj(U_j(e)) G_j(e))
// with respect to r_ij.  U_j(e) = alpha_e P_e + beta_e Q_e with
// alpha = rpa conj(a), beta = -rpb conj(b), P_e = U_{j-1}(ma, mb) and
// Q_e = U_{j-1}(ma-1, mb) (mirrors on the middle column of an even layer),
// and r enters only through a, b and sfac.  So, with the adjoints
// Ubar_j(e) = dL/dRe U + i dL/dIm U propagated from the top layer down,
//   Ubar_{j-1}(P_e) += conj(alpha) Ubar_j(e),  Ubar_{j-1}(Q_e) += conj(beta) Ubar_j(e)
// (through a mirror M = s conj(U'): Ubar(U') += s conj(conj(alpha) Ubar_j(e))),
//   dL/dr_k = dsfac u_k sum Re(conj(U) w Y) + Re(conj(da_k) Sa) - Re(conj(db_k) Sb),
//   Sa = sum_e rpa conj(Ubar_j(e)) P_e,  Sb = sum_e rpb conj(Ubar_j(e)) Q_e.
// The forward pass keeps U of layers 0..7 in registers (lane ma holds
// (ma, mb) of every column, as in k_snap_ui_rg); the backward pass moves
// adjoints to the lane below with a DPP shift and back through the mirror
// permutation.  About half the arithmetic of carrying dU/dr forward.
template <class T, class TF>
__global__ void __launch_bounds__(64 * RG_WAVES) __attribute__((amdgpu_waves_per_eu(3)))
k_snap_fij_rev(snap_dev_tables t, const rg_rp<T> *__restrict__ rgp, const rg_map *__restrict__ rgm,
               const int *__restrict__ nin, const int *__restrict__ nj,
               const double *__restrict__ nr, const snap_par<T> *__restrict__ par,
               const double *__restrict__ fdv, int maxn, int n, const T *__restrict__ y_r,
               const T *__restrict__ y_i, TF *__restrict__ f) {
  constexpr int TJ = SNAP_MAXJ;
  __shared__ rg_rec<T> s_rec[RG_SLOTS * RG_W];
  __shared__ T s_red[RG_WAVES][3][64];
  const int i = blockIdx.x;
  if (i >= n) return;
  const int cnt_i = nin[i];
  const int row0 = blockIdx.y * (RG_WAVES * RG_NB);
  if (row0 >= cnt_i) return;  // no neighbour for this block
  rg_fill<T, 64 * RG_WAVES>(rgp, rgm, t.ny, y_r, y_i, i, s_rec);
  __syncthreads();
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  const int g = l / RG_W, ma = l - g * RG_W;  // g = 7 only for the idle lane 63
  const int jj0 = row0 + w * RG_NB;
  if (jj0 >= cnt_i) return;  // uniform over the wavefront; no block barrier follows
  const int jj = jj0 + g;
  const bool act = g < RG_NB && jj < cnt_i;
  const size_t o = (size_t)i * maxn + (act ? jj : jj0);
  const T a_r = par[o].a_r, a_i = par[o].a_i, b_r = par[o].b_r, b_i = par[o].b_i;
  const T sfac = par[o].sfac;
  const int tj = t.twojmax;

  // forward: U of layers 0..TJ-1 (column slot rg_slot0(j) + mb), and
  // sum Re(conj(U) w Y) over the lane's entries of every layer
  T us[RG_SLOTS][2];
#pragma unroll
  for (int s = 0; s < RG_SLOTS; s++) us[s][0] = us[s][1] = 0;
  us[0][0] = ma == 0 ? (T)1 : (T)0;
  T auy = us[0][0] * s_rec[ma].yr;
#pragma unroll
  for (int j = 1; j <= TJ; j++) {
    if (j > tj) break;
    T mir[2], mirb[2];
    if ((j & 1) == 0) {
      const int src = g * RG_W + (j - 1 - ma > 0 ? j - 1 - ma : 0);
      const T sg = ((ma + j / 2) & 1) ? (T)-1 : (T)1;
      const int sp = rg_slot0(j - 1) + j / 2 - 1;
      mir[0] = sg * __shfl(us[sp][0], src, 64);
      mir[1] = -sg * __shfl(us[sp][1], src, 64);
      mirb[0] = lane_below(mir[0]);
      mirb[1] = lane_below(mir[1]);
    }
#pragma unroll
    for (int mb = 0; mb <= j / 2; mb++) {
      const int slot = rg_slot0(j) + mb;
      const rg_rec<T> q = s_rec[slot * RG_W + ma];
      const bool mid = (j & 1) == 0 && mb == j / 2;
      const int sp = rg_slot0(j - 1) + mb;  // (unused on the middle column)
      const T p0 = mid ? mir[0] : us[sp][0], p1 = mid ? mir[1] : us[sp][1];
      const T q0 = mid ? mirb[0] : lane_below(us[sp][0]);
      const T q1 = mid ? mirb[1] : lane_below(us[sp][1]);
      const T ur = q.rpa * (a_r * p0 + a_i * p1) - q.rpb * (b_r * q0 + b_i * q1);
      const T ui = q.rpa * (a_r * p1 - a_i * p0) - q.rpb * (b_r * q1 - b_i * q0);
      auy += ur * q.yr + ui * q.yi;
      if (j < TJ) {
        us[slot][0] = ur;
        us[slot][1] = ui;
      }
    }
  }

  // backward: adjoints of layer j (columns 0..j/2), from the top layer down
  T ub[RG_COLS][2];
#pragma unroll
  for (int mb = 0; mb < RG_COLS; mb++) ub[mb][0] = ub[mb][1] = 0;
  T sa_r = 0, sa_i = 0, sb_r = 0, sb_i = 0;
#pragma unroll
  for (int j = TJ; j >= 1; j--) {
    if (j > tj) continue;
    if (j == tj) {  // the top layer's adjoint is its own term
#pragma unroll
      for (int mb = 0; mb <= j / 2; mb++) {
        const rg_rec<T> q = s_rec[(rg_slot0(j) + mb) * RG_W + ma];
        ub[mb][0] = sfac * q.yr;
        ub[mb][1] = sfac * q.yi;
      }
    }
    T mir[2], mirb[2];
    if ((j & 1) == 0) {
      const int src = g * RG_W + (j - 1 - ma > 0 ? j - 1 - ma : 0);
      const T sg = ((ma + j / 2) & 1) ? (T)-1 : (T)1;
      const int sp = rg_slot0(j - 1) + j / 2 - 1;
      mir[0] = sg * __shfl(us[sp][0], src, 64);
      mir[1] = -sg * __shfl(us[sp][1], src, 64);
      mirb[0] = lane_below(mir[0]);
      mirb[1] = lane_below(mir[1]);
    }
    T ubp[RG_COLS][2];  // adjoints of layer j-1
#pragma unroll
    for (int mb = 0; mb < RG_COLS; mb++) ubp[mb][0] = ubp[mb][1] = 0;
#pragma unroll
    for (int mb = 0; mb <= j / 2; mb++) {
      const rg_rec<T> q = s_rec[(rg_slot0(j) + mb) * RG_W + ma];
      const bool mid = (j & 1) == 0 && mb == j / 2;
      const int sp = rg_slot0(j - 1) + mb;
      const T p0 = mid ? mir[0] : us[sp][0], p1 = mid ? mir[1] : us[sp][1];
      const T q0 = mid ? mirb[0] : lane_below(us[sp][0]);
      const T q1 = mid ? mirb[1] : lane_below(us[sp][1]);
      const T u0 = ub[mb][0], u1 = ub[mb][1];
      // Sa += rpa conj(Ubar) P, Sb += rpb conj(Ubar) Q
      sa_r += q.rpa * (u0 * p0 + u1 * p1);
      sa_i += q.rpa * (u0 * p1 - u1 * p0);
      sb_r += q.rpb * (u0 * q0 + u1 * q1);
      sb_i += q.rpb * (u0 * q1 - u1 * q0);
      // conj(alpha) Ubar = rpa a Ubar, conj(beta) Ubar = -rpb b Ubar
      const T za0 = q.rpa * (a_r * u0 - a_i * u1), za1 = q.rpa * (a_r * u1 + a_i * u0);
      const T zb0 = -q.rpb * (b_r * u0 - b_i * u1), zb1 = -q.rpb * (b_r * u1 + b_i * u0);
      if (!mid) {
        ubp[mb][0] += za0 + lane_above(zb0);
        ubp[mb][1] += za1 + lane_above(zb1);
      } else {
        // P = s conj(U'(j-1-ma, j/2-1)), Q = -s conj(U'(j-ma, j/2-1)):
        // lane m' of column j/2-1 collects s conj(za) of lane j-1-m' and
        // -s conj(zb) of lane j-m'
        const T sg = ((ma + j / 2) & 1) ? (T)-1 : (T)1;
        const T v0 = sg * za0 + lane_above(-sg * zb0);
        const T v1 = -sg * za1 + lane_above(sg * zb1);
        const int src = g * RG_W + (j - 1 - ma > 0 ? j - 1 - ma : 0);
        const T g0 = __shfl(v0, src, 64), g1 = __shfl(v1, src, 64);
        const bool own = ma <= j - 1;
        ubp[j / 2 - 1][0] += own ? g0 : (T)0;
        ubp[j / 2 - 1][1] += own ? g1 : (T)0;
      }
    }
    // layer j-1's own term
#pragma unroll
    for (int mb = 0; mb <= (j - 1) / 2; mb++) {
      const rg_rec<T> q = s_rec[(rg_slot0(j - 1) + mb) * RG_W + ma];
      ub[mb][0] = ubp[mb][0] + sfac * q.yr;
      ub[mb][1] = ubp[mb][1] + sfac * q.yi;
    }
  }
  // the lane's share of F_ij, summed over the nine lanes of the group in order
  const T u3[3] = {par[o].u[0], par[o].u[1], par[o].u[2]};
  const T c1 = par[o].c1, c2 = par[o].c2, c3 = par[o].c3, c4 = par[o].c4, r0i = par[o].r0inv;
  const T dsfac = par[o].dsfac;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const T da_r = c1 * u3[k], da_i = c2 * u3[k] - (k == 2 ? r0i : (T)0);
    const T db_r = c3 * u3[k] + (k == 1 ? r0i : (T)0), db_i = c4 * u3[k] - (k == 0 ? r0i : (T)0);
    s_red[w][k][l] = dsfac * auy * u3[k] + (da_r * sa_r + da_i * sa_i) - (db_r * sb_r + db_i * sb_i);
  }
  wave_sync();
  if (ma == 0 && act) {
    T fx = 0, fy = 0, fz = 0;
#pragma unroll
    for (int m = 0; m < RG_W; m++) {
      fx += s_red[w][0][l + m];
      fy += s_red[w][1][l + m];
      fz += s_red[w][2][l + m];
    }
    // ad-hoc repulsion of the reference, force_snap_neigh_impl.h:705-711
    const double fdivr = fdv[o];
    const TF Fx = (TF)fx + (TF)(nr[3 * o] * fdivr);
    const TF Fy = (TF)fy + (TF)(nr[3 * o + 1] * fdivr);
    const TF Fz = (TF)fz + (TF)(nr[3 * o + 2] * fdivr);
    const int jn = nj[o];
    atomicAdd(&f[3 * (size_t)i], Fx);
    atomicAdd(&f[3 * (size_t)i + 1], Fy);
    atomicAdd(&f[3 * (size_t)i + 2], Fz);
    atomicAdd(&f[3 * (size_t)jn], -Fx);
    atomicAdd(&f[3 * (size_t)jn + 1], -Fy);
    atomicAdd(&f[3 * (size_t)jn + 2], -Fz);
  }
}

// per-atom energy: beta_0 + sum_k beta_k B_k + pair repulsion (FP64)
__global__ void __launch_bounds__(256)
k_snap_energy(snap_dev_tables t, const double *__restrict__ cg,
              const etab_entry *__restrict__ etab, int ne, double beta0, int bzero,
              const int *__restrict__ nin, const double *__restrict__ nr, int maxn, int n,
              const double *__restrict__ utot_r, const double *__restrict__ utot_i,
              double *__restrict__ eatom) {
  __shared__ double su_r[SNAP_NU_BIG], su_i[SNAP_NU_BIG];
  __shared__ double s_red[4];
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int k = threadIdx.x; k < t.nu; k += blockDim.x) {
    su_r[k] = utot_r[(size_t)i * t.nu + k];
    su_i[k] = utot_i[(size_t)i * t.nu + k];
  }
  __syncthreads();
  double e = 0.0;
  // work items: (k, half-plane entry of J = j_k)
  for (int k = 0; k < ne; k++) {
    const etab_entry te = etab[k];
    const int J = te.j;
    const int nyj = t.ystart[J + 1] - t.ystart[J];
    for (int q = threadIdx.x; q < nyj; q += blockDim.x) {
      const int mb = q / (J + 1), ma = q - mb * (J + 1);
      double wgt = 1.0;
      if (2 * mb == J) wgt = (ma < mb) ? 1.0 : ((ma == mb) ? 0.5 : 0.0);
      if (wgt == 0.0) continue;
      const cplx<double> z = z_entry<double>(t, cg, su_r, su_i, te.j1, te.j2, J, te.cgoff, ma, mb);
      const int u = t.ustart[J] + ma * (J + 1) + mb;
      e += te.beta * 2.0 * wgt * (su_r[u] * z.r + su_i[u] * z.i);
    }
  }
  // bzero and the repulsion (rc-shifted), done by thread 0
  if (threadIdx.x == 0) {
    if (bzero)
      for (int k = 0; k < ne; k++) e -= etab[k].beta * t.wself * t.wself * t.wself * (etab[k].j + 1);
    const double rc12 = 1.0 / pow(t.rcut * t.rcut, 6.0);
    for (int jj = 0; jj < nin[i]; jj++) {
      const size_t o = (size_t)i * maxn + jj;
      const double rsq = nr[3 * o] * nr[3 * o] + nr[3 * o + 1] * nr[3 * o + 1] + nr[3 * o + 2] * nr[3 * o + 2];
      e += 1.25e5 * (1.0 / pow(rsq, 6.0) - rc12);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
  if (threadIdx.x % 64 == 0) s_red[threadIdx.x / 64] = e;
  __syncthreads();
  if (threadIdx.x == 0) eatom[i] = beta0 + ((s_red[0] + s_red[1]) + (s_red[2] + s_red[3]));
}

// ============================================================ host side
struct emd_snap {
  snap_dev_tables t;
  int ncoeff;
  std::vector<double> beta;            // beta_0 .. beta_ncoeff
  std::vector<int> idxj;               // 3 per component
  std::map<std::tuple<int, int, int>, int> cgoff;
  std::vector<double> cg_h;
  double *cg = nullptr;
  ytab_entry *ytab = nullptr;
  etab_entry *etab = nullptr;
  size_t ytab_cap = 0, etab_cap = 0;  // entries allocated (kept across coefficient updates)
  // host sources of the last coefficient upload's stream-ordered copies
  // (kept until the next upload, which first waits for the stream)
  std::vector<ytab_entry> h_ytab;
  std::vector<etab_entry> h_etab;
  std::vector<double> h_kcoef;
  int bzeroflag = 0;
  double rcutfac = 0, radelem = 0;
  // per-system scratch
  int cap_n = 0, cap_maxn = 0;
  int *nin = nullptr, *nj = nullptr;
  double *nr = nullptr;
  // neighbour parameters (snap_par<float> or <double>, per slot) and the
  // repulsion F/r, with the precision the parameters were last written in
  void *par = nullptr;
  double *fdv = nullptr;
  // register-layout tables (twojmax <= 8): recursion factors in both
  // precisions and the index map, per (column slot, lane)
  rg_rp<float> *rg_rp32 = nullptr;
  rg_rp<double> *rg_rp64 = nullptr;
  rg_map *rg_map_d = nullptr;
  size_t cap_par = 0, cap_fdv = 0;
  int par_prec = -1;
  void *u_r = nullptr, *u_i = nullptr, *y_r = nullptr, *y_i = nullptr;
  size_t cap_u = 0, cap_y = 0, cap_ui = 0, cap_yi = 0;
  double *eatom = nullptr;
  int cap_e = 0;
  // flat Y term stream (k_snap_yflat): codes (LDS byte offsets for FP32 and
  // FP64 U_tot), key per slot, cg products, weights for the current
  // coefficients, run slot per (chunk, thread), first run per entry
  uint32_t *yf_code32 = nullptr, *yf_code64 = nullptr;
  unsigned short *yf_key = nullptr, *yf_meta = nullptr;
  unsigned short *yf_centry = nullptr;  // entry of each (chunk, thread), 0 if none
  unsigned short *yf_keytid = nullptr;  // adjoint triple id of each key
  double *s_part = nullptr;  // the beta adjoint's per-block key sums (k_snap_s_flat)
  size_t cap_spart = 0;
  // emd_snap_static_geometry: the compaction, neighbour parameters and U_tot
  // of the last force or adjoint call (twojmax <= 8) are reused while the
  // caller keeps positions and list fixed: their precision (-1: none), atom
  // count, and the position and list arrays they came from
  int geom_static = 0, geom_prec = -1, geom_n = -1;
  const double *geom_x = nullptr;
  const int *geom_ent = nullptr;
  double *yf_cgw = nullptr, *yf_w64 = nullptr, *yf_kcoef = nullptr;
  float *yf_w32 = nullptr;
  int *yf_rstart = nullptr;
  int yf_steps = 0, yf_nch = 0, yf_nruns = 0;
  std::vector<std::tuple<int, int, int>> yf_keys;  // (a, b, J), sorted
  // adjoint: every (a,b,J) triple grouped by J, with its id
  std::map<std::tuple<int, int, int>, int> tid;
  std::vector<int> sjstart;
  stab_entry *stab = nullptr;
  double *w_r = nullptr, *w_i = nullptr, *sbuf = nullptr;
  size_t cap_w = 0, cap_wi = 0, cap_s = 0;
};

static double host_factorial(int n) {
  double r = 1.0;
  for (int k = 2; k <= n; k++) r *= k;
  return r;
}

// Clebsch-Gordan <j1/2 m1; j2/2 m2 | j/2 m1+m2> by the Racah formula
// (integer-doubled arguments; m1 = k1 - j1/2, m2 = k2 - j2/2)
static double host_cg(int j1, int j2, int j, int k1, int k2) {
  const int aa2 = 2 * k1 - j1, bb2 = 2 * k2 - j2;
  const int cc2 = aa2 + bb2;
  if (cc2 < -j || cc2 > j) return 0.0;
  double sum = 0.0;
  const int zmin = std::max(0, std::max(-(j - j2 + aa2) / 2, -(j - j1 - bb2) / 2));
  const int zmax = std::min((j1 + j2 - j) / 2, std::min((j1 - aa2) / 2, (j2 + bb2) / 2));
  for (int z = zmin; z <= zmax; z++) {
    const double sgn = (z % 2) ? -1.0 : 1.0;
    sum += sgn / (host_factorial(z) * host_factorial((j1 + j2 - j) / 2 - z) *
                  host_factorial((j1 - aa2) / 2 - z) * host_factorial((j2 + bb2) / 2 - z) *
                  host_factorial((j - j2 + aa2) / 2 + z) * host_factorial((j - j1 - bb2) / 2 + z));
  }
  const double delta = std::sqrt(host_factorial((j1 + j2 - j) / 2) * host_factorial((j1 - j2 + j) / 2) *
                                 host_factorial((-j1 + j2 + j) / 2) /
                                 host_factorial((j1 + j2 + j) / 2 + 1));
  const double norm = std::sqrt(host_factorial((j1 + aa2) / 2) * host_factorial((j1 - aa2) / 2) *
                                host_factorial((j2 + bb2) / 2) * host_factorial((j2 - bb2) / 2) *
                                host_factorial((j + cc2) / 2) * host_factorial((j - cc2) / 2) *
                                (j + 1));
  return sum * delta * norm;
}

// the beta-independent part of the flat Y stream (see k_snap_yflat)
// LDS byte offsets of the two U_tot reads of every slot of the Y stream
// (code = offset1 | offset2 << 16, 8-byte words): for each step, each
// wavefront, each of the two reads and each lane group served in one LDS
// cycle, the group's distinct U entries are matched to distinct bank
// residues (u + shift(c)) mod MOD, each entry choosing one copy c
// (augmenting paths); an entry left unmatched takes the least loaded
// residue among its copies.  Lanes reading the same entry read the same copy.
template <class L>
static void yf_assign_copies(const std::vector<int> &u1, const std::vector<int> &u2, int steps,
                             std::vector<uint32_t> &code) {
  std::vector<int> cp1(u1.size(), 0), cp2(u2.size(), 0);
  for (int st = 0; st < steps; st++)
    for (int w0 = 0; w0 < YF_THREADS; w0 += 64)
      for (int op = 0; op < 2; op++)
        for (int g0 = w0; g0 < w0 + 64; g0 += L::GROUP) {
          const std::vector<int> &u = op ? u2 : u1;
          std::vector<int> &cp = op ? cp2 : cp1;
          const size_t s0 = (size_t)st * YF_THREADS + g0;
          std::vector<int> addr;  // distinct entries of the group
          for (int l = 0; l < L::GROUP; l++)
            if (std::find(addr.begin(), addr.end(), u[s0 + l]) == addr.end()) addr.push_back(u[s0 + l]);
          const int na = (int)addr.size();
          std::vector<int> owner(L::MOD, -1), copy_of(na, -1);
          std::function<bool(int, std::vector<char> &)> augment = [&](int a,
                                                                      std::vector<char> &seen) {
            for (int c = 0; c < L::NCOPY; c++) {
              const int b = (addr[a] + L::shift(c)) % L::MOD;
              if (seen[b]) continue;
              seen[b] = 1;
              if (owner[b] < 0 || augment(owner[b], seen)) {
                owner[b] = a;
                copy_of[a] = c;
                return true;
              }
            }
            return false;
          };
          std::vector<int> load(L::MOD, 0);
          for (int a = 0; a < na; a++) {
            std::vector<char> seen(L::MOD, 0);
            augment(a, seen);
          }
          for (int a = 0; a < na; a++)
            if (copy_of[a] >= 0) load[(addr[a] + L::shift(copy_of[a])) % L::MOD]++;
          for (int a = 0; a < na; a++)
            if (copy_of[a] < 0) {
              int best = 0;
              for (int c = 1; c < L::NCOPY; c++)
                if (load[(addr[a] + L::shift(c)) % L::MOD] < load[(addr[a] + L::shift(best)) % L::MOD])
                  best = c;
              copy_of[a] = best;
              load[(addr[a] + L::shift(best)) % L::MOD]++;
            }
          for (int l = 0; l < L::GROUP; l++) {
            const int a = (int)(std::find(addr.begin(), addr.end(), u[s0 + l]) - addr.begin());
            cp[s0 + l] = copy_of[a];
          }
        }
  for (size_t k = 0; k < u1.size(); k++)
    code[k] = (uint32_t)((L::base(cp1[k]) + u1[k]) * 8) |
              ((uint32_t)((L::base(cp2[k]) + u2[k]) * 8) << 16);
}

static emd_status snap_build_yflat(emd_ctx *ctx, emd_snap *sn) {
  const snap_dev_tables &t = sn->t;
  const int nk = (int)sn->idxj.size() / 3;
  std::set<std::tuple<int, int, int>> keyset;
  for (int k = 0; k < nk; k++) {
    const int j1 = sn->idxj[3 * k], j2 = sn->idxj[3 * k + 1], j = sn->idxj[3 * k + 2];
    keyset.insert(std::make_tuple(std::max(j1, j2), std::min(j1, j2), j));
    keyset.insert(std::make_tuple(std::max(j, j2), std::min(j, j2), j1));
    keyset.insert(std::make_tuple(std::max(j1, j), std::min(j1, j), j2));
  }
  sn->yf_keys.assign(keyset.begin(), keyset.end());
  struct term {
    int i1, i2, key;
    double cgw;
  };
  std::vector<std::vector<term>> eterms(t.ny);
  for (int J = 0; J <= t.twojmax; J++)
    for (int mb = 0; 2 * mb <= J; mb++)
      for (int ma = 0; ma <= J; ma++) {
        const int e = t.ystart[J] + mb * (J + 1) + ma;
        for (size_t q = 0; q < sn->yf_keys.size(); q++) {
          int a, b, JJ;
          std::tie(a, b, JJ) = sn->yf_keys[q];
          if (JJ != J) continue;
          const int cgoff = sn->cgoff.at(sn->yf_keys[q]);
          const int ma1lo = std::max(0, (2 * ma - J - b + a) / 2);
          const int ma1hi = std::min(a, (2 * ma - J + b + a) / 2);
          const int mb1lo = std::max(0, (2 * mb - J - b + a) / 2);
          const int mb1hi = std::min(a, (2 * mb - J + b + a) / 2);
          for (int ma1 = ma1lo; ma1 <= ma1hi; ma1++) {
            const int ma2 = (2 * ma - J - (2 * ma1 - a) + b) / 2;
            const double cma = sn->cg_h[cgoff + ma1 * (b + 1) + ma2];
            for (int mb1 = mb1lo; mb1 <= mb1hi; mb1++) {
              const int mb2 = (2 * mb - J - (2 * mb1 - a) + b) / 2;
              const double cw = cma * sn->cg_h[cgoff + mb1 * (b + 1) + mb2];
              if (cw == 0.0) continue;
              eterms[e].push_back(term{t.ustart[a] + ma1 * (a + 1) + mb1,
                                       t.ustart[b] + ma2 * (b + 1) + mb2, (int)q, cw});
            }
          }
        }
      }
  // chunks of YF_CHUNK terms in entry order, an entry's last chunk padded
  // with zero-weight terms (U offsets 0, weight 0); thread th takes chunks
  // [th P, (th+1) P)
  std::vector<std::pair<int, int>> crange;  // (entry, first term)
  for (int e = 0; e < t.ny; e++)
    for (int f = 0; f < (int)eterms[e].size(); f += YF_CHUNK) crange.push_back(std::make_pair(e, f));
  const int nchunks = (int)crange.size();
  const int P = (nchunks + YF_THREADS - 1) / YF_THREADS;
  // runs: a thread's consecutive chunks of one entry; run slots numbered
  // entry by entry, in thread order within an entry (= term order)
  std::vector<std::vector<std::pair<int, int>>> eruns(t.ny);  // (thread, local chunk)
  for (int th = 0; th < YF_THREADS; th++)
    for (int c = 0; c < P; c++) {
      const int g = th * P + c;
      if (g >= nchunks) break;
      const int e = crange[g].first;
      const bool ends = c == P - 1 || g + 1 >= nchunks || crange[g + 1].first != e;
      if (ends) eruns[e].push_back(std::make_pair(th, c));
    }
  std::vector<unsigned short> meta((size_t)P * YF_THREADS, YF_NO_RUN);
  std::vector<int> rstart(t.ny + 1, 0);
  int nruns = 0;
  for (int e = 0; e < t.ny; e++) {
    rstart[e] = nruns;
    for (const auto &r : eruns[e]) meta[(size_t)r.second * YF_THREADS + r.first] = (unsigned short)nruns++;
  }
  rstart[t.ny] = nruns;
  const bool small = t.twojmax <= SNAP_MAXJ;
  if (nruns > (small ? yf_maxruns<SNAP_MAXJ>() : yf_maxruns<SNAP_JBIG>()) ||
      t.nu > (small ? snap_nu(SNAP_MAXJ) : snap_nu(SNAP_JBIG)))
    return emd_set_error(ctx, EMD_ERR_ARG, "emd_snap_create: Y stream too large");
  const int steps = P * YF_CHUNK;
  const size_t nslots = (size_t)steps * YF_THREADS;
  std::vector<uint32_t> code32(nslots, 0u), code64(nslots, 0u);
  std::vector<int> u1(nslots, 0), u2(nslots, 0);
  std::vector<unsigned short> key(nslots, 0);
  std::vector<double> cgw(nslots, 0.0);
  // A chunk's terms all belong to one entry, so their order within the
  // chunk is free (it changes rounding only): by key, padding last, so that
  // the adjoint's S kernel sums runs of one key before its atomic add (1.8
  // runs per chunk at twojmax 8).  LDS bank conflicts are left to the copy
  // assignment below (offline: 1.08 cycles per read against 1.04 with a
  // bank-greedy order).
  const term zero_term{0, 0, 0, 0.0};
  for (int th = 0; th < YF_THREADS; th++)
    for (int c = 0; c < P; c++) {
      const int g = th * P + c;
      std::vector<term> pool;
      for (int k = 0; k < YF_CHUNK; k++) {
        if (g < nchunks && crange[g].second + k < (int)eterms[crange[g].first].size())
          pool.push_back(eterms[crange[g].first][crange[g].second + k]);
        else
          pool.push_back(zero_term);
      }
      std::stable_sort(pool.begin(), pool.end(), [](const term &x, const term &y) {
        if ((x.cgw == 0.0) != (y.cgw == 0.0)) return y.cgw == 0.0;
        if (x.key != y.key) return x.key < y.key;
        return x.i1 != y.i1 ? x.i1 < y.i1 : x.i2 < y.i2;
      });
      for (int k = 0; k < YF_CHUNK; k++) {
        const size_t slot = ((size_t)c * YF_CHUNK + k) * YF_THREADS + th;
        u1[slot] = pool[k].i1;
        u2[slot] = pool[k].i2;
        key[slot] = (unsigned short)pool[k].key;
        cgw[slot] = pool[k].cgw;
      }
    }
  if (small) {
    yf_assign_copies<yf_layout<float, SNAP_MAXJ>>(u1, u2, steps, code32);
    code64 = code32;  // the same word offsets (FP64: into the real array)
  } else {
    yf_assign_copies<yf_layout<float, SNAP_JBIG>>(u1, u2, steps, code32);
    code64 = code32;
  }
  sn->yf_steps = steps;
  sn->yf_nch = P;
  sn->yf_nruns = nruns;
  EMD_HIP(ctx, hipMalloc(&sn->yf_code32, nslots * sizeof(uint32_t)));
  EMD_HIP(ctx, hipMalloc(&sn->yf_code64, nslots * sizeof(uint32_t)));
  {
    std::vector<unsigned short> centry((size_t)P * YF_THREADS, 0);
    for (int th = 0; th < YF_THREADS; th++)
      for (int c = 0; c < P; c++) {
        const int g = th * P + c;
        if (g < nchunks) centry[(size_t)c * YF_THREADS + th] = (unsigned short)crange[g].first;
      }
    std::vector<unsigned short> keytid(std::max<size_t>(1, sn->yf_keys.size()), 0);
    for (size_t q = 0; q < sn->yf_keys.size(); q++) keytid[q] = (unsigned short)sn->tid.at(sn->yf_keys[q]);
    EMD_HIP(ctx, hipMalloc(&sn->yf_centry, centry.size() * sizeof(unsigned short)));
    EMD_HIP(ctx, hipMalloc(&sn->yf_keytid, keytid.size() * sizeof(unsigned short)));
    EMD_HIP(ctx, hipMemcpy(sn->yf_centry, centry.data(), centry.size() * sizeof(unsigned short),
                           hipMemcpyHostToDevice));
    EMD_HIP(ctx, hipMemcpy(sn->yf_keytid, keytid.data(), keytid.size() * sizeof(unsigned short),
                           hipMemcpyHostToDevice));
  }
  EMD_HIP(ctx, hipMalloc(&sn->yf_key, nslots * sizeof(unsigned short)));
  EMD_HIP(ctx, hipMalloc(&sn->yf_meta, meta.size() * sizeof(unsigned short)));
  EMD_HIP(ctx, hipMalloc(&sn->yf_cgw, nslots * sizeof(double)));
  EMD_HIP(ctx, hipMalloc(&sn->yf_w64, nslots * sizeof(double)));
  EMD_HIP(ctx, hipMalloc(&sn->yf_w32, nslots * sizeof(float)));
  EMD_HIP(ctx, hipMalloc(&sn->yf_rstart, (t.ny + 1) * sizeof(int)));
  EMD_HIP(ctx, hipMalloc(&sn->yf_kcoef, std::max<size_t>(1, sn->yf_keys.size()) * sizeof(double)));
  EMD_HIP(ctx, hipMemcpy(sn->yf_code32, code32.data(), nslots * sizeof(uint32_t), hipMemcpyHostToDevice));
  EMD_HIP(ctx, hipMemcpy(sn->yf_code64, code64.data(), nslots * sizeof(uint32_t), hipMemcpyHostToDevice));
  EMD_HIP(ctx, hipMemcpy(sn->yf_key, key.data(), nslots * sizeof(unsigned short), hipMemcpyHostToDevice));
  EMD_HIP(ctx, hipMemcpy(sn->yf_meta, meta.data(), meta.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
  EMD_HIP(ctx, hipMemcpy(sn->yf_cgw, cgw.data(), nslots * sizeof(double), hipMemcpyHostToDevice));
  EMD_HIP(ctx, hipMemcpy(sn->yf_rstart, rstart.data(), (t.ny + 1) * sizeof(int), hipMemcpyHostToDevice));
  return EMD_OK;
}

// the register-layout tables for twojmax <= 8 (see rg_map)
static emd_status snap_build_rg(emd_ctx *ctx, emd_snap *sn) {
  const snap_dev_tables &t = sn->t;
  std::vector<rg_rp<float>> r32(RG_SLOTS * RG_W);
  std::vector<rg_rp<double>> r64(RG_SLOTS * RG_W);
  std::vector<rg_map> mp(RG_SLOTS * RG_W);
  for (int sl = 0; sl < RG_SLOTS; sl++)
    for (int ma = 0; ma < RG_W; ma++) {
      const int k = sl * RG_W + ma;
      int j = 0;
      while (j < SNAP_MAXJ && rg_slot0(j + 1) <= sl) j++;
      const int mb = sl - rg_slot0(j);
      double rpa = 0, rpb = 0;
      rg_map m{-1, -1, -1, 0, 0, 0, 0};
      if (j <= t.twojmax && ma <= j) {
        if (ma < j) rpa = t.rootpq[(j - ma) * SNAP_RPD + (j - mb)];
        if (ma > 0) rpb = t.rootpq[ma * SNAP_RPD + (j - mb)];
        // the reference's half plane: middle column ma < mb, diagonal 1/2
        m.wq = 2 * mb == j ? (ma < mb ? 2 : (ma == mb ? 1 : 0)) : 2;
        m.ey = t.ystart[j] + mb * (j + 1) + ma;
        m.uo = t.ustart[j] + ma * (j + 1) + mb;
        m.um = 2 * mb != j ? t.ustart[j] + (j - ma) * (j + 1) + (j - mb) : -1;
        m.neg = (ma + mb) & 1;
        m.diag = ma == mb;
      }
      r32[k] = rg_rp<float>{(float)rpa, (float)rpb};
      r64[k] = rg_rp<double>{rpa, rpb};
      mp[k] = m;
    }
  EMD_HIP(ctx, hipMalloc(&sn->rg_rp32, r32.size() * sizeof(r32[0])));
  EMD_HIP(ctx, hipMalloc(&sn->rg_rp64, r64.size() * sizeof(r64[0])));
  EMD_HIP(ctx, hipMalloc(&sn->rg_map_d, mp.size() * sizeof(mp[0])));
  EMD_HIP(ctx, hipMemcpy(sn->rg_rp32, r32.data(), r32.size() * sizeof(r32[0]), hipMemcpyHostToDevice));
  EMD_HIP(ctx, hipMemcpy(sn->rg_rp64, r64.data(), r64.size() * sizeof(r64[0]), hipMemcpyHostToDevice));
  EMD_HIP(ctx, hipMemcpy(sn->rg_map_d, mp.data(), mp.size() * sizeof(mp[0]), hipMemcpyHostToDevice));
  return EMD_OK;
}

template <class T>
static const rg_rp<T> *rg_rp_of(const emd_snap *sn);
template <>
const rg_rp<float> *rg_rp_of<float>(const emd_snap *sn) { return sn->rg_rp32; }
template <>
const rg_rp<double> *rg_rp_of<double>(const emd_snap *sn) { return sn->rg_rp64; }

static emd_status snap_upload_coeff(emd_ctx *ctx, emd_snap *sn) {
  const int twojmax = sn->t.twojmax;
  // Y weights: the three roles of each component (compute_dbidrj)
  std::map<std::tuple<int, int, int>, double> coef;
  const int nk = (int)sn->idxj.size() / 3;
  for (int k = 0; k < nk; k++) {
    const int j1 = sn->idxj[3 * k], j2 = sn->idxj[3 * k + 1], j = sn->idxj[3 * k + 2];
    const double b = sn->beta[k + 1];
    coef[std::make_tuple(std::max(j1, j2), std::min(j1, j2), j)] += 2.0 * b;
    coef[std::make_tuple(std::max(j, j2), std::min(j, j2), j1)] += 2.0 * b * (j + 1) / (j1 + 1.0);
    coef[std::make_tuple(std::max(j1, j), std::min(j1, j), j2)] += 2.0 * b * (j + 1) / (j2 + 1.0);
  }
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));  // the previous upload's copies are done
  std::vector<ytab_entry> &yt = sn->h_ytab;
  yt.clear();
  for (int J = 0; J <= twojmax; J++) {
    sn->t.yjstart[J] = (int)yt.size();
    for (auto &kv : coef) {
      int a, b, JJ;
      std::tie(a, b, JJ) = kv.first;
      if (JJ != J || kv.second == 0.0) continue;
      ytab_entry e;
      e.a = a; e.b = b; e.cgoff = sn->cgoff.at(kv.first); e.pad = 0; e.coef = kv.second;
      yt.push_back(e);
    }
  }
  sn->t.yjstart[twojmax + 1] = (int)yt.size();
  sn->t.ny_terms = (int)yt.size();
  std::vector<etab_entry> &et = sn->h_etab;
  et.clear();
  for (int k = 0; k < nk; k++) {
    etab_entry e;
    e.j1 = sn->idxj[3 * k]; e.j2 = sn->idxj[3 * k + 1]; e.j = sn->idxj[3 * k + 2];
    e.cgoff = sn->cgoff.at(std::make_tuple(e.j1, e.j2, e.j));
    e.beta = sn->beta[k + 1];
    et.push_back(e);
  }
  // the tables keep their allocation while they fit (coefficient updates
  // in a force-matching loop do not reallocate); copies are stream-ordered
  if (yt.size() > sn->ytab_cap || !sn->ytab) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (sn->ytab) EMD_HIP(ctx, hipFree(sn->ytab));
    sn->ytab = nullptr;
    sn->ytab_cap = std::max<size_t>(1, yt.size());
    EMD_HIP(ctx, hipMalloc(&sn->ytab, sn->ytab_cap * sizeof(ytab_entry)));
  }
  if (et.size() > sn->etab_cap || !sn->etab) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (sn->etab) EMD_HIP(ctx, hipFree(sn->etab));
    sn->etab = nullptr;
    sn->etab_cap = std::max<size_t>(1, et.size());
    EMD_HIP(ctx, hipMalloc(&sn->etab, sn->etab_cap * sizeof(etab_entry)));
  }
  if (!yt.empty())
    EMD_HIP(ctx, hipMemcpyAsync(sn->ytab, yt.data(), yt.size() * sizeof(ytab_entry),
                                hipMemcpyHostToDevice, ctx->stream));
  if (!et.empty())
    EMD_HIP(ctx, hipMemcpyAsync(sn->etab, et.data(), et.size() * sizeof(etab_entry),
                                hipMemcpyHostToDevice, ctx->stream));
  if (sn->yf_code32) {
    std::vector<double> &kc = sn->h_kcoef;
    kc.assign(sn->yf_keys.size(), 0.0);
    for (size_t q = 0; q < kc.size(); q++) {
      auto it = coef.find(sn->yf_keys[q]);
      if (it != coef.end()) kc[q] = it->second;
    }
    if (!kc.empty())
      EMD_HIP(ctx, hipMemcpyAsync(sn->yf_kcoef, kc.data(), kc.size() * sizeof(double),
                                  hipMemcpyHostToDevice, ctx->stream));
    const int nslots = sn->yf_steps * YF_THREADS;
    if (nslots > 0)
      EMD_LAUNCH(ctx, "snap_coeff", k_yf_weights, emd_blocks(nslots, 256), 256, 0, nslots,
                 sn->yf_key, sn->yf_cgw, sn->yf_kcoef, sn->yf_w64, sn->yf_w32);
  }
  return EMD_OK;
}

extern "C" emd_status emd_snap_create(emd_ctx *ctx, const emd_snap_params *p, emd_snap **out) {
  if (!ctx || !p || !out) return EMD_ERR_ARG;
  *out = nullptr;
  if (p->twojmax < 0 || p->twojmax > SNAP_JBIG)
    return emd_set_error(ctx, EMD_ERR_ARG, "emd_snap_create: twojmax must be in 0..%d", SNAP_JBIG);
  if (!p->coeff) return emd_set_error(ctx, EMD_ERR_ARG, "emd_snap_create: no coefficients");
  emd_snap *sn = new emd_snap();
  snap_dev_tables &t = sn->t;
  memset(&t, 0, sizeof(t));
  const int twojmax = p->twojmax;
  t.twojmax = twojmax;
  int nu = 0, ny = 0;
  for (int j = 0; j <= twojmax; j++) {
    t.ustart[j] = nu;
    t.ystart[j] = ny;
    nu += (j + 1) * (j + 1);
    ny += (j + 1) * (j / 2 + 1);
  }
  t.ustart[twojmax + 1] = nu;
  t.ystart[twojmax + 1] = ny;
  t.nu = nu;
  t.ny = ny;
  for (int a = 1; a <= twojmax; a++)
    for (int b = 1; b <= twojmax; b++) t.rootpq[a * SNAP_RPD + b] = std::sqrt((double)a / b);
  t.rfac0 = p->rfac0;
  t.rmin0 = p->rmin0;
  t.rcut = (p->radelem + p->radelem) * p->rcutfac;
  t.wj = p->wjelem;
  t.wself = 1.0;
  t.switchflag = p->switchflag;
  sn->bzeroflag = p->bzeroflag;
  sn->rcutfac = p->rcutfac;
  sn->radelem = p->radelem;
  // component list (SNA::build_indexlist, diagonalstyle 3) and CG blocks
  for (int j1 = 0; j1 <= twojmax; j1++)
    for (int j2 = 0; j2 <= j1; j2++)
      for (int j = j1 - j2; j <= std::min(twojmax, j1 + j2); j += 2) {
        if (j >= j1) {
          sn->idxj.push_back(j1);
          sn->idxj.push_back(j2);
          sn->idxj.push_back(j);
        }
        sn->cgoff[std::make_tuple(j1, j2, j)] = (int)sn->cg_h.size();
        for (int k1 = 0; k1 <= j1; k1++)
          for (int k2 = 0; k2 <= j2; k2++) sn->cg_h.push_back(host_cg(j1, j2, j, k1, k2));
      }
  sn->ncoeff = (int)sn->idxj.size() / 3;
  {
    // all (a,b,J) triples grouped by J for the adjoint contraction
    std::vector<stab_entry> st;
    for (int J = 0; J <= twojmax; J++) {
      sn->sjstart.push_back((int)st.size());
      for (auto &kv : sn->cgoff) {
        int a, b, JJ;
        std::tie(a, b, JJ) = kv.first;
        if (JJ != J) continue;
        const int id = (int)sn->tid.size();
        sn->tid[kv.first] = id;
        st.push_back(stab_entry{a, b, kv.second, id});
      }
    }
    sn->sjstart.push_back((int)st.size());
    if (st.size() > (size_t)SNAP_MAXT) {
      delete sn;
      return emd_set_error(ctx, EMD_ERR_ARG, "emd_snap_create: too many (a,b,J) triples");
    }
    EMD_HIP(ctx, hipMalloc(&sn->stab, st.size() * sizeof(stab_entry)));
    EMD_HIP(ctx, hipMemcpy(sn->stab, st.data(), st.size() * sizeof(stab_entry), hipMemcpyHostToDevice));
  }
  if (p->ncoeff != sn->ncoeff) {
    const int want = sn->ncoeff;
    emd_snap_destroy(ctx, sn);
    return emd_set_error(ctx, EMD_ERR_ARG, "emd_snap_create: ncoeff %d but twojmax %d has %d",
                         p->ncoeff, twojmax, want);
  }
  sn->beta.assign(p->coeff, p->coeff + sn->ncoeff + 1);
  if (hipMalloc(&sn->cg, sn->cg_h.size() * sizeof(double)) != hipSuccess) {
    delete sn;
    return emd_set_error(ctx, EMD_ERR_ALLOC, "emd_snap_create: hipMalloc failed");
  }
  EMD_HIP(ctx, hipMemcpy(sn->cg, sn->cg_h.data(), sn->cg_h.size() * sizeof(double),
                         hipMemcpyHostToDevice));
  emd_status st = snap_build_yflat(ctx, sn);
  if (st == EMD_OK && twojmax <= SNAP_MAXJ) st = snap_build_rg(ctx, sn);
  if (st == EMD_OK) st = snap_upload_coeff(ctx, sn);
  if (st != EMD_OK) {
    emd_snap_destroy(ctx, sn);
    return st;
  }
  *out = sn;
  return EMD_OK;
}

extern "C" emd_status emd_snap_destroy(emd_ctx *ctx, emd_snap *sn) {
  if (!sn) return EMD_OK;
  if (ctx) hipStreamSynchronize(ctx->stream);
  void *ptrs[] = {sn->cg, sn->ytab, sn->etab, sn->nin, sn->nj, sn->nr, sn->par, sn->fdv,
                  sn->rg_rp32, sn->rg_rp64, sn->rg_map_d,
                  sn->stab, sn->w_r,
                  sn->w_i, sn->sbuf,
                  sn->u_r, sn->u_i, sn->y_r, sn->y_i, sn->eatom,
                  sn->yf_code32, sn->yf_code64, sn->yf_key, sn->yf_meta, sn->yf_cgw,
                  sn->yf_centry, sn->yf_keytid, sn->s_part,
                  sn->yf_w64, sn->yf_w32, sn->yf_rstart, sn->yf_kcoef};
  for (void *q : ptrs)
    if (q) hipFree(q);
  delete sn;
  return EMD_OK;
}

extern "C" emd_status emd_snap_static_geometry(emd_ctx *ctx, emd_snap *sn, int on) {
  if (!ctx || !sn) return EMD_ERR_ARG;
  sn->geom_static = on != 0;
  sn->geom_prec = -1;
  return EMD_OK;
}

extern "C" emd_status emd_snap_set_coeff(emd_ctx *ctx, emd_snap *sn, const double *coeff) {
  if (!ctx || !sn || !coeff) return EMD_ERR_ARG;
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  sn->beta.assign(coeff, coeff + sn->ncoeff + 1);
  return snap_upload_coeff(ctx, sn);
}

static emd_status snap_scratch(emd_ctx *ctx, emd_snap *sn, int n, int maxn) {
  if (n > sn->cap_n || maxn > sn->cap_maxn || !sn->nin) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (sn->nin) EMD_HIP(ctx, hipFree(sn->nin));
    if (sn->nj) EMD_HIP(ctx, hipFree(sn->nj));
    if (sn->nr) EMD_HIP(ctx, hipFree(sn->nr));
    const int cn = n + n / 8 + 16, cm = maxn + 8;
    EMD_HIP(ctx, hipMalloc(&sn->nin, (size_t)cn * sizeof(int)));
    EMD_HIP(ctx, hipMalloc(&sn->nj, (size_t)cn * cm * sizeof(int)));
    EMD_HIP(ctx, hipMalloc(&sn->nr, (size_t)cn * cm * 3 * sizeof(double)));
    sn->cap_n = cn;
    sn->cap_maxn = cm;
  }
  const size_t slots = (size_t)n * sn->cap_maxn;
  EMD_TRY(emd_ensure(ctx, &sn->par, &sn->cap_par, slots * sizeof(snap_par<double>)));
  EMD_TRY(emd_ensure(ctx, (void **)&sn->fdv, &sn->cap_fdv, slots * sizeof(double)));
  EMD_TRY(emd_ensure(ctx, &sn->u_r, &sn->cap_u, (size_t)n * sn->t.nu * sizeof(double)));
  EMD_TRY(emd_ensure(ctx, &sn->u_i, &sn->cap_ui, (size_t)n * sn->t.nu * sizeof(double)));
  EMD_TRY(emd_ensure(ctx, &sn->y_r, &sn->cap_y, (size_t)n * sn->t.ny * sizeof(double)));
  EMD_TRY(emd_ensure(ctx, &sn->y_i, &sn->cap_yi, (size_t)n * sn->t.ny * sizeof(double)));
  return EMD_OK;
}

// zero_prec: EMD_FP64 or EMD_MIXED to zero that precision's force array
// (all n_local + n_ghost rows) in the same launch, or -1.  par_prec:
// EMD_FP64 or EMD_MIXED to write the neighbour parameters in that precision
// (for the U and force kernels that follow), or -1
static emd_status snap_neigh(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl, emd_snap *sn,
                             int zero_prec, int par_prec) {
  const int n = s->n_local;
  const int maxn = nl->maxneighs > 0 ? nl->maxneighs : 1;
  EMD_TRY(snap_scratch(ctx, sn, n, maxn));
  const double rcutmax = 2.0 * sn->radelem * sn->rcutfac;
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
  double *f64 = zero_prec == EMD_FP64 ? s->f : nullptr;
  float *f32 = zero_prec == EMD_MIXED ? s->f32 : nullptr;
  snap_par<float> *p32 = par_prec == EMD_MIXED ? (snap_par<float> *)sn->par : nullptr;
  snap_par<double> *p64 = par_prec == EMD_FP64 ? (snap_par<double> *)sn->par : nullptr;
  EMD_LAUNCH(ctx, "snap_neigh", k_snap_neigh, n, 64, 0, sn->t, s->x, row, nl->num, nl->entries,
             nl->maxneighs, n, rcutmax * rcutmax, sn->cap_maxn, sn->nin, sn->nj, sn->nr, f64,
             f32, s->n_local + s->n_ghost, p32, p64, (p32 || p64) ? sn->fdv : nullptr);
  sn->par_prec = (p32 || p64) ? par_prec : -1;
  return EMD_OK;
}

// the neighbour parameters in T's precision (written by the compaction, or
// here when it wrote the other precision)
template <class T>
static emd_status snap_par_ready(emd_ctx *ctx, emd_snap *sn, int n) {
  const int prec = std::is_same<T, float>::value ? EMD_MIXED : EMD_FP64;
  if (sn->par_prec == prec) return EMD_OK;
  EMD_LAUNCH(ctx, "snap_neigh", k_snap_par<T>, n, 64, 0, sn->t, sn->nin, sn->nr, sn->cap_maxn, n,
             (snap_par<T> *)sn->par, sn->fdv);
  sn->par_prec = prec;
  return EMD_OK;
}

// twojmax <= 8: the compaction with the neighbour parameters in precision
// prec and U_tot of every atom, in one launch (k_snap_neigh_ui)
static emd_status snap_neigh_ui(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                emd_snap *sn, int zero_prec, int prec) {
  const int n = s->n_local;
  const int maxn = nl->maxneighs > 0 ? nl->maxneighs : 1;
  EMD_TRY(snap_scratch(ctx, sn, n, maxn));
  const double rcutmax = 2.0 * sn->radelem * sn->rcutfac;
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
  double *f64 = zero_prec == EMD_FP64 ? s->f : nullptr;
  float *f32 = zero_prec == EMD_MIXED ? s->f32 : nullptr;
  if (prec == EMD_MIXED)
    EMD_LAUNCH(ctx, "snap_ui", k_snap_neigh_ui<float>, n, 64, 0, sn->t, sn->rg_rp32, sn->rg_map_d,
               s->x, row, nl->num, nl->entries, nl->maxneighs, n, rcutmax * rcutmax, sn->cap_maxn,
               sn->nin, sn->nj, sn->nr, f64, f32, s->n_local + s->n_ghost,
               (snap_par<float> *)sn->par, sn->fdv, (float *)sn->u_r, (float *)sn->u_i);
  else
    EMD_LAUNCH(ctx, "snap_ui", k_snap_neigh_ui<double>, n, 64, 0, sn->t, sn->rg_rp64, sn->rg_map_d,
               s->x, row, nl->num, nl->entries, nl->maxneighs, n, rcutmax * rcutmax, sn->cap_maxn,
               sn->nin, sn->nj, sn->nr, f64, f32, s->n_local + s->n_ghost,
               (snap_par<double> *)sn->par, sn->fdv, (double *)sn->u_r, (double *)sn->u_i);
  sn->par_prec = prec;
  return EMD_OK;
}

// U_tot of every atom: the register-resident kernel for twojmax <= 8 (one
// wavefront per atom: SNAP-2k 18.6 us mixed / 23.5 us FP64, against 18.8 /
// 32.8 with two and 26.8 / 47.9 with four wavefronts per atom), else one
// neighbour per wavefront
template <class T>
static emd_status launch_ui(emd_ctx *ctx, emd_snap *sn, int n, T *ur, T *ui) {
  if (sn->t.twojmax <= SNAP_MAXJ) {
    EMD_TRY(snap_par_ready<T>(ctx, sn, n));
    EMD_LAUNCH(ctx, "snap_ui", (k_snap_ui_rg<T, 1>), n, 64, 0, sn->t, rg_rp_of<T>(sn), sn->rg_map_d,
               sn->nin, (const snap_par<T> *)sn->par, sn->cap_maxn, n, ur, ui);
  } else {
    EMD_LAUNCH(ctx, "snap_ui", (k_snap_ui<T, SNAP_JBIG>), n, 64 * UI_WAVES, 0, sn->t, sn->nin,
               sn->nr, sn->cap_maxn, n, ur, ui);
  }
  return EMD_OK;
}

template <class T>
static emd_status snap_y(emd_ctx *ctx, emd_snap *sn, int n) {
  T *ur = (T *)sn->u_r, *ui = (T *)sn->u_i, *yr = (T *)sn->y_r, *yi = (T *)sn->y_i;
  const T *w = std::is_same<T, float>::value ? (const T *)sn->yf_w32 : (const T *)sn->yf_w64;
  const uint32_t *code = std::is_same<T, float>::value ? sn->yf_code32 : sn->yf_code64;
  // twojmax <= 8: two atoms per block share each term's code and weight
  // load (SNAP-2k: 26.8 / 47.2 us mixed / FP64, against 31.5 / 60.9 with one
  // atom and 28.2 / 55.9 with three)
  if (sn->t.twojmax <= SNAP_MAXJ)
    EMD_LAUNCH(ctx, "snap_y", (k_snap_yflat<T, SNAP_MAXJ, 2>), (n + 1) / 2, YF_THREADS, 0, sn->t.nu,
               sn->t.ny, sn->yf_nch, code, w, sn->yf_meta, sn->yf_rstart, n, ur, ui, yr, yi);
  else
    EMD_LAUNCH(ctx, "snap_y", (k_snap_yflat<T, SNAP_JBIG, 1>), n, YF_THREADS, 0, sn->t.nu, sn->t.ny,
               sn->yf_nch, code, w, sn->yf_meta, sn->yf_rstart, n, ur, ui, yr, yi);
  return EMD_OK;
}

template <class T>
static emd_status snap_ui_y(emd_ctx *ctx, emd_snap *sn, int n) {
  EMD_TRY(launch_ui<T>(ctx, sn, n, (T *)sn->u_r, (T *)sn->u_i));
  return snap_y<T>(ctx, sn, n);
}

// F_ij for every (atom, neighbour): the register-resident reverse-mode
// kernel (seven neighbours per wavefront) for twojmax <= 8 (SNAP-2k: 29.6 /
// 36.8 us mixed / FP64, against 40.6 / 50.6 us carrying dU/dr forward in the
// same layout), one neighbour per wavefront above
template <class T, class TF>
static emd_status launch_fij(emd_ctx *ctx, emd_snap *sn, int n, const T *yr, const T *yi,
                             TF *f) {
  if (sn->t.twojmax <= SNAP_MAXJ) {
    EMD_TRY(snap_par_ready<T>(ctx, sn, n));
    const dim3 grid(n, (sn->cap_maxn + RG_WAVES * RG_NB - 1) / (RG_WAVES * RG_NB));
    EMD_LAUNCH(ctx, "snap_fij", (k_snap_fij_rev<T, TF>), grid, 64 * RG_WAVES, 0, sn->t,
               rg_rp_of<T>(sn), sn->rg_map_d, sn->nin, sn->nj, sn->nr,
               (const snap_par<T> *)sn->par, sn->fdv, sn->cap_maxn, n, yr, yi, f);
  } else {
    const dim3 grid(n, (sn->cap_maxn + FIJ_WAVES - 1) / FIJ_WAVES);
    EMD_LAUNCH(ctx, "snap_fij", (k_snap_fij<T, TF, SNAP_JBIG>), grid, 64 * FIJ_WAVES, 0, sn->t,
               sn->nin, sn->nj, sn->nr, sn->cap_maxn, n, yr, yi, f);
  }
  return EMD_OK;
}

// with static geometry on: are the compaction, neighbour parameters and U_tot
// in precision prec those of this system and list?
static bool snap_geom_cached(const emd_snap *sn, const emd_system *s, const emd_neigh *nl,
                             int prec) {
  return sn->geom_static && sn->geom_prec == prec && sn->par_prec == prec &&
         sn->geom_n == s->n_local && sn->geom_x == s->x && sn->geom_ent == nl->entries;
}

static void snap_geom_record(emd_snap *sn, const emd_system *s, const emd_neigh *nl, int prec) {
  if (!sn->geom_static) return;
  sn->geom_prec = prec;
  sn->geom_n = s->n_local;
  sn->geom_x = s->x;
  sn->geom_ent = nl->entries;
}

static emd_status snap_force(emd_ctx *ctx, emd_system *s, const emd_neigh *nl, emd_snap *sn,
                             int precision, bool zero) {
  if (!ctx || !s || !nl || !sn) return EMD_ERR_ARG;
  EMD_ARG(ctx, nl->n == s->n_local, "neighbour list does not match the system");
  const int n = s->n_local;
  if (n <= 0) return zero ? emd_force_zero(ctx, s, precision) : EMD_OK;
  const int prec = precision == EMD_MIXED ? EMD_MIXED : EMD_FP64;
  if (sn->t.twojmax <= SNAP_MAXJ) {
    // compaction, neighbour parameters and U_tot in one launch (or, with
    // static geometry, those of the previous call), then Y
    if (snap_geom_cached(sn, s, nl, prec)) {
      if (zero) EMD_TRY(emd_force_zero(ctx, s, prec));
    } else {
      EMD_TRY(snap_neigh_ui(ctx, s, nl, sn, zero ? prec : -1, prec));
      snap_geom_record(sn, s, nl, prec);
    }
    if (precision == EMD_MIXED) EMD_TRY(snap_y<float>(ctx, sn, n));
    else EMD_TRY(snap_y<double>(ctx, sn, n));
  } else {
    EMD_TRY(snap_neigh(ctx, s, nl, sn, zero ? prec : -1, prec));
    if (precision == EMD_MIXED) EMD_TRY(snap_ui_y<float>(ctx, sn, n));
    else EMD_TRY(snap_ui_y<double>(ctx, sn, n));
  }
  if (precision == EMD_MIXED) {
    EMD_TRY((launch_fij<float, float>(ctx, sn, n, (const float *)sn->y_r,
                                      (const float *)sn->y_i, s->f32)));
  } else {
    EMD_TRY((launch_fij<double, double>(ctx, sn, n, (const double *)sn->y_r,
                                        (const double *)sn->y_i, s->f)));
  }
  return EMD_OK;
}

extern "C" emd_status emd_snap_force(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                     emd_snap *sn, int precision) {
  return snap_force(ctx, s, nl, sn, precision, false);
}

extern "C" emd_status emd_snap_force_zeroed(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                            emd_snap *sn, int precision) {
  return snap_force(ctx, s, nl, sn, precision, true);
}

extern "C" emd_status emd_snap_energy(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                      emd_snap *sn, double *energy, double *eatom) {
  if (!ctx || !s || !nl || !sn || !energy) return EMD_ERR_ARG;
  const int n = s->n_local;
  *energy = 0.0;
  if (n <= 0) return EMD_OK;
  sn->geom_prec = -1;  // the compaction and U_tot are rewritten
  EMD_TRY(snap_neigh(ctx, s, nl, sn, -1, EMD_FP64));
  EMD_TRY(launch_ui<double>(ctx, sn, n, (double *)sn->u_r, (double *)sn->u_i));
  double *ea = eatom;
  if (!ea) {
    EMD_TRY(emd_ensure_int(ctx, (int **)&sn->eatom, &sn->cap_e, 2l * n));
    ea = sn->eatom;
  }
  EMD_LAUNCH(ctx, "snap_energy", k_snap_energy, n, 256, 0, sn->t, sn->cg, sn->etab,
             sn->ncoeff, sn->beta[0], sn->bzeroflag, sn->nin, sn->nr, sn->cap_maxn, n,
             (const double *)sn->u_r, (const double *)sn->u_i, ea);
  std::vector<double> h(n);
  EMD_HIP(ctx, hipMemcpyAsync(h.data(), ea, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  EMD_TRY(emd_wait(ctx));
  double sum = 0.0, c = 0.0;
  for (int k = 0; k < n; k++) {
    const double v = h[k], tt = sum + v;
    if (std::fabs(sum) >= std::fabs(v)) c += (sum - tt) + v;
    else c += (v - tt) + sum;
    sum = tt;
  }
  *energy = sum + c;
  return EMD_OK;
}

// ====================================================== adjoint (dL/dbeta)
// F is affine in beta: F_ij = sum_J sum' Re(conj(dU_J) Y_J(beta)) + repulsion.
// For a cotangent g = dL/dF, sum_i g_i . dF_i/dbeta_k
//   = sum_i sum_J sum' Re(conj(W^i_J) dY^i_J/dbeta_k),
//   W^i_J = sum_{j in N(i)} w(ma,mb) sum_d (g_i - g_j)_d dU_J/dr_d
// (ghost j carries the cotangent of its owner), and dY_J/dbeta_k is the
// three-role combination of Z_{abJ}; so per triple (a,b,J) only
// S_{abJ} = sum_i sum' Re(conj(W^i_J) Z^i_{abJ}) is needed.

// one U/dU layer step for the lane's entry (same recursion as k_snap_fij);
// P holds the previous layer in full, component q at P[q * FULL + ma * j + mb]
template <class T, int FULL>
__device__ __forceinline__ void du_step(const snap_dev_tables &t, const T *P, int j, int ma, int mb,
                                        T a_r, T a_i, T b_r, T b_i, const T *da_r, const T *da_i,
                                        const T *db_r, const T *db_i, T &ur, T &ui, T *dr, T *di) {
  ur = ui = 0;
#pragma unroll
  for (int k = 0; k < 3; k++) dr[k] = di[k] = 0;
  if (ma < j) {
    const T rp = (T)t.rootpq[(j - ma) * SNAP_RPD + (j - mb)];
    const int s = ma * j + mb;
    const T pr = P[s], pi = P[FULL + s];
    ur += rp * (a_r * pr + a_i * pi);
    ui += rp * (a_r * pi - a_i * pr);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const T dpr = P[(2 + 2 * k) * FULL + s], dpi = P[(3 + 2 * k) * FULL + s];
      dr[k] += rp * (da_r[k] * pr + da_i[k] * pi + a_r * dpr + a_i * dpi);
      di[k] += rp * (da_r[k] * pi - da_i[k] * pr + a_r * dpi - a_i * dpr);
    }
  }
  if (ma > 0) {
    const T rp = (T)t.rootpq[ma * SNAP_RPD + (j - mb)];
    const int s = (ma - 1) * j + mb;
    const T pr = P[s], pi = P[FULL + s];
    ur += -rp * (b_r * pr + b_i * pi);
    ui += -rp * (b_r * pi - b_i * pr);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const T dpr = P[(2 + 2 * k) * FULL + s], dpi = P[(3 + 2 * k) * FULL + s];
      dr[k] += -rp * (db_r[k] * pr + db_i[k] * pi + b_r * dpr + b_i * dpi);
      di[k] += -rp * (db_r[k] * pi - db_i[k] * pr + b_r * dpi - b_i * dpr);
    }
  }
}

// W^i for every atom: one wavefront per atom, neighbours in list order, the
// half-plane entries of a layer 64 at a time (NST steps), FP64 accumulators
// (T is the arithmetic of the recursion)
template <class T, int J>
__global__ void __launch_bounds__(64)
k_snap_w(snap_dev_tables t, const int *__restrict__ nin, const int *__restrict__ nj,
         const double *__restrict__ nr, int maxn, int n, const int *__restrict__ root,
         const double *__restrict__ g, double *__restrict__ w_r, double *__restrict__ w_i) {
  constexpr int JD = snap_dim<J>::JD, FULL = snap_dim<J>::FULLPREV;
  constexpr int NST = (snap_hp(J) + 63) / 64;
  __shared__ T s_u[2][8][FULL];
  const int i = blockIdx.x;
  if (i >= n) return;
  const int l = threadIdx.x;
  double acc_r[JD][NST], acc_i[JD][NST];
#pragma unroll
  for (int j = 0; j < JD; j++)
#pragma unroll
    for (int st = 0; st < NST; st++) acc_r[j][st] = acc_i[j][st] = 0.0;
  const double gi0 = g[3 * (size_t)i], gi1 = g[3 * (size_t)i + 1], gi2 = g[3 * (size_t)i + 2];
  const int cnt = nin[i];
  for (int jj = 0; jj < cnt; jj++) {
    const size_t o = (size_t)i * maxn + jj;
    const int jn = nj[o];
    const int own_j = jn < n ? jn : root[jn - n];
    const T gd[3] = {(T)(gi0 - g[3 * (size_t)own_j]), (T)(gi1 - g[3 * (size_t)own_j + 1]),
                     (T)(gi2 - g[3 * (size_t)own_j + 2])};
    const T x = (T)nr[3 * o], y = (T)nr[3 * o + 1], z = (T)nr[3 * o + 2];
    const T rsq = x * x + y * y + z * z;
    const T r = sqrt(rsq);
    const T rscale0 = (T)t.rfac0 * (T)SNAP_PI / ((T)t.rcut - (T)t.rmin0);
    const T theta0 = (r - (T)t.rmin0) * rscale0;
    const T z0 = r * cos(theta0) / sin(theta0);
    const T dz0dr = z0 / r - (r * rscale0) * (rsq + z0 * z0) / rsq;
    const T rinv = (T)1 / r;
    const T u3[3] = {x * rinv, y * rinv, z * rinv};
    const T r0inv = (T)1 / sqrt(r * r + z0 * z0);
    const T a_r = z0 * r0inv, a_i = -z * r0inv, b_r = y * r0inv, b_i = -x * r0inv;
    const T dr0invdr = -(r0inv * r0inv * r0inv) * (r + z0 * dz0dr);
    T da_r[3], da_i[3], db_r[3], db_i[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const T dr0 = dr0invdr * u3[k], dz = dz0dr * u3[k];
      da_r[k] = dz * r0inv + z0 * dr0;
      da_i[k] = -z * dr0;
      db_r[k] = y * dr0;
      db_i[k] = -x * dr0;
    }
    da_i[2] += -r0inv;
    db_i[0] += -r0inv;
    db_r[1] += r0inv;
    T sfac, dsfac;
    sfac_dsfac(t, r, sfac, dsfac);
    sfac *= (T)t.wj;
    dsfac *= (T)t.wj;
    const T gu = gd[0] * u3[0] + gd[1] * u3[1] + gd[2] * u3[2];
    if (l == 0) {
      s_u[0][0][0] = 1;
      s_u[0][1][0] = 0;
      for (int k = 0; k < 3; k++) { s_u[0][2 + 2 * k][0] = 0; s_u[0][3 + 2 * k][0] = 0; }
      acc_r[0][0] += 0.5 * (double)(dsfac * gu);
    }
    __syncthreads();
    for (int j = 1; j <= t.twojmax; j++) {
      const int p = (j - 1) & 1, c = j & 1;
      T ur[NST], ui[NST], dr[NST][3], di[NST][3];
      int ma[NST], mb[NST];
      bool own[NST];
#pragma unroll
      for (int st = 0; st < NST; st++) {
        const int e = st * 64 + l;
        own[st] = e < snap_hp(j) && lane_entry(e, j, ma[st], mb[st]);
        ur[st] = ui[st] = 0;
#pragma unroll
        for (int k = 0; k < 3; k++) dr[st][k] = di[st][k] = 0;
        if (own[st])
          du_step<T, FULL>(t, &s_u[p][0][0], j, ma[st], mb[st], a_r, a_i, b_r, b_i, da_r, da_i,
                           db_r, db_i, ur[st], ui[st], dr[st], di[st]);
      }
      __syncthreads();
#pragma unroll
      for (int st = 0; st < NST; st++) {
        if (own[st] && j < t.twojmax) {
          T *C = &s_u[c][0][0];
          const int s = ma[st] * (j + 1) + mb[st];
          C[s] = ur[st]; C[FULL + s] = ui[st];
#pragma unroll
          for (int k = 0; k < 3; k++) { C[(2 + 2 * k) * FULL + s] = dr[st][k]; C[(3 + 2 * k) * FULL + s] = di[st][k]; }
          if (2 * mb[st] != j) {
            const T sg = ((ma[st] + mb[st]) & 1) ? (T)-1 : (T)1;
            const int m = (j - ma[st]) * (j + 1) + (j - mb[st]);
            C[m] = sg * ur[st]; C[FULL + m] = -sg * ui[st];
#pragma unroll
            for (int k = 0; k < 3; k++) { C[(2 + 2 * k) * FULL + m] = sg * dr[st][k]; C[(3 + 2 * k) * FULL + m] = -sg * di[st][k]; }
          }
        }
        if (own[st]) {
          double wgt = 1.0;
          if (2 * mb[st] == j) wgt = (ma[st] < mb[st]) ? 1.0 : ((ma[st] == mb[st]) ? 0.5 : 0.0);
          const T wr = dsfac * ur[st] * gu +
                       sfac * (gd[0] * dr[st][0] + gd[1] * dr[st][1] + gd[2] * dr[st][2]);
          const T wi = dsfac * ui[st] * gu +
                       sfac * (gd[0] * di[st][0] + gd[1] * di[st][1] + gd[2] * di[st][2]);
          acc_r[j][st] += wgt * (double)wr;
          acc_i[j][st] += wgt * (double)wi;
        }
      }
      __syncthreads();
    }
  }
  for (int j = 0; j <= t.twojmax; j++) {
#pragma unroll
    for (int st = 0; st < NST; st++) {
      int ma, mb;
      const int e = st * 64 + l;
      if (e < snap_hp(j) && lane_entry(e, j, ma, mb)) {
        const int ey = t.ystart[j] + mb * (j + 1) + ma;
        w_r[(size_t)i * t.ny + ey] = acc_r[j][st];
        w_i[(size_t)i * t.ny + ey] = acc_i[j][st];
      }
    }
  }
}

// W^i for every atom with the register-resident recursion (twojmax <= 8):
//   W^i_J(e) = sum_{j in N(i)} w(e) [dsfac (g . u) U_J(e) + sfac dU_J(e)/dr . g],
// g = g_i - g_j (a ghost j carries its owner's cotangent).  Only the
// derivative along g is needed, so each lane carries U and dU_g = g . dU/dr
// (4 values per entry): its recursion is that of U with da_g = g . da/dr and
// db_g = g . db/dr, both rank one in u plus constants:
//   da_g = c1 (g.u) + i (c2 (g.u) - r0inv g_z),  db_g = c3 (g.u) + r0inv g_y
//                                                   + i (c4 (g.u) - r0inv g_x).
// One wavefront per atom, seven neighbours at a time; each lane sums its
// group's neighbours in T, the groups are added in FP64 (fixed tree):
// deterministic.
template <class T>
__global__ void __launch_bounds__(64)
k_snap_w_rg(snap_dev_tables t, const rg_rp<T> *__restrict__ rgp, const rg_map *__restrict__ rgm,
            const int *__restrict__ nin, const int *__restrict__ nj,
            const snap_par<T> *__restrict__ par, int maxn, int n, const int *__restrict__ root,
            const double *__restrict__ g, double *__restrict__ w_r, double *__restrict__ w_i) {
  __shared__ rg_rec<T> s_rec[RG_SLOTS * RG_W];
  __shared__ double s_acc[2][RG_SLOTS * RG_W];
  const int i = blockIdx.x;
  if (i >= n) return;
  rg_fill<T, 64>(rgp, rgm, 0, nullptr, nullptr, i, s_rec);
  __syncthreads();
  for (int k = threadIdx.x; k < RG_SLOTS * RG_W; k += 64)
    s_rec[k].yr = (T)0.5 * (T)rgm[k].wq;  // the half-plane weight of the entry
  __syncthreads();
  const int l = threadIdx.x;
  const int gr = l / RG_W, ma = l - gr * RG_W;  // gr = 7 only for the idle lane 63
  T acc[RG_SLOTS][2];
#pragma unroll
  for (int s = 0; s < RG_SLOTS; s++) acc[s][0] = acc[s][1] = 0;
  const double gi0 = g[3 * (size_t)i], gi1 = g[3 * (size_t)i + 1], gi2 = g[3 * (size_t)i + 2];
  const int cnt = nin[i];
  for (int base = 0; base < cnt; base += RG_NB) {
    const int jj = base + gr;
    const bool act = gr < RG_NB && jj < cnt;
    const size_t o = (size_t)i * maxn + (act ? jj : base);
    const snap_par<T> pp = par[o];
    const int jn = nj[o];
    const int own_j = jn < n ? jn : root[jn - n];
    const T g0 = act ? (T)(gi0 - g[3 * (size_t)own_j]) : (T)0;
    const T g1 = act ? (T)(gi1 - g[3 * (size_t)own_j + 1]) : (T)0;
    const T g2 = act ? (T)(gi2 - g[3 * (size_t)own_j + 2]) : (T)0;
    const T gu = g0 * pp.u[0] + g1 * pp.u[1] + g2 * pp.u[2];
    const T a_r = pp.a_r, a_i = pp.a_i, b_r = pp.b_r, b_i = pp.b_i;
    const T da_r = pp.c1 * gu, da_i = pp.c2 * gu - pp.r0inv * g2;
    const T db_r = pp.c3 * gu + pp.r0inv * g1, db_i = pp.c4 * gu - pp.r0inv * g0;
    const T sfac = pp.sfac, dsg = pp.dsfac * gu;  // dsg = 0 for an idle group
    // c[mb]: U_r, U_i, dU_g r, dU_g i of (ma, mb)
    T c[RG_COLS][4];
#pragma unroll
    for (int mb = 0; mb < RG_COLS; mb++) c[mb][0] = c[mb][1] = c[mb][2] = c[mb][3] = 0;
    c[0][0] = ma == 0 ? (T)1 : (T)0;  // layer 0: U = 1, dU = 0
    acc[0][0] += (T)0.5 * dsg * c[0][0];
#pragma unroll
    for (int j = 1; j <= SNAP_MAXJ; j++) {
      if (j > t.twojmax) break;
      T mir[4], mirb[4];  // even j: (U, dU_g) of (ma, j/2) and (ma-1, j/2) in layer j-1
      if ((j & 1) == 0) {
        const int src = gr * RG_W + (j - 1 - ma > 0 ? j - 1 - ma : 0);
        const T sg = ((ma + j / 2) & 1) ? (T)-1 : (T)1;
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const T v = __shfl(c[j / 2 - 1][q], src, 64);
          mir[q] = (q & 1) ? -sg * v : sg * v;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) mirb[q] = lane_below(mir[q]);
      }
#pragma unroll
      for (int mb = 0; mb <= j / 2; mb++) {
        const int slot = rg_slot0(j) + mb;
        const rg_rec<T> q = s_rec[slot * RG_W + ma];
        const bool mid = (j & 1) == 0 && mb == j / 2;
        T P[4], Q[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          P[k] = mid ? mir[k] : c[mb][k];
          Q[k] = mid ? mirb[k] : lane_below(c[mb][k]);
        }
        const T ur = q.rpa * (a_r * P[0] + a_i * P[1]) - q.rpb * (b_r * Q[0] + b_i * Q[1]);
        const T ui = q.rpa * (a_r * P[1] - a_i * P[0]) - q.rpb * (b_r * Q[1] - b_i * Q[0]);
        const T dr = q.rpa * (da_r * P[0] + da_i * P[1] + a_r * P[2] + a_i * P[3]) -
                     q.rpb * (db_r * Q[0] + db_i * Q[1] + b_r * Q[2] + b_i * Q[3]);
        const T di = q.rpa * (da_r * P[1] - da_i * P[0] + a_r * P[3] - a_i * P[2]) -
                     q.rpb * (db_r * Q[1] - db_i * Q[0] + b_r * Q[3] - b_i * Q[2]);
        c[mb][0] = ur;
        c[mb][1] = ui;
        c[mb][2] = dr;
        c[mb][3] = di;
        acc[slot][0] += q.yr * (dsg * ur + sfac * dr);
        acc[slot][1] += q.yr * (dsg * ui + sfac * di);
      }
    }
  }
  rg_group_sum<T, double>(acc, l, s_acc[0], s_acc[1]);
  __syncthreads();
  for (int k = l; k < RG_SLOTS * RG_W; k += 64) {
    const rg_map m = rgm[k];
    if (m.ey < 0) continue;
    w_r[(size_t)i * t.ny + m.ey] = s_acc[0][k];
    w_i[(size_t)i * t.ny + m.ey] = s_acc[1][k];
  }
}

// S_{abJ} += sum' Re(conj(W^i_J) Z^i_{abJ}) over atoms (FP64 accumulation)
template <class T>
__global__ void __launch_bounds__(192)
k_snap_s(snap_dev_tables t, const double *__restrict__ cg, const stab_entry *__restrict__ stab,
         const int *__restrict__ sjstart, int ntrip, int n, const T *__restrict__ utot_r,
         const T *__restrict__ utot_i, const double *__restrict__ w_r,
         const double *__restrict__ w_i, double *__restrict__ S) {
  __shared__ T su_r[SNAP_NU_BIG], su_i[SNAP_NU_BIG];
  __shared__ double sS[SNAP_MAXT];
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int k = threadIdx.x; k < t.nu; k += blockDim.x) {
    su_r[k] = utot_r[(size_t)i * t.nu + k];
    su_i[k] = utot_i[(size_t)i * t.nu + k];
  }
  for (int k = threadIdx.x; k < ntrip; k += blockDim.x) sS[k] = 0.0;
  __syncthreads();
  for (int e = threadIdx.x; e < t.ny; e += blockDim.x) {
    int J, ma, mb;
    y_entry(t, e, J, ma, mb);
    const double wr = w_r[(size_t)i * t.ny + e], wi = w_i[(size_t)i * t.ny + e];
    if (wr != 0.0 || wi != 0.0) {
      for (int q = sjstart[J]; q < sjstart[J + 1]; q++) {
        const stab_entry te = stab[q];
        const cplx<T> z = z_entry<T>(t, cg, su_r, su_i, te.a, te.b, J, te.cgoff, ma, mb);
        atomicAdd(&sS[te.tid], wr * (double)z.r + wi * (double)z.i);
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < ntrip; k += blockDim.x) atomicAdd(&S[k], sS[k]);
}

// S_{abJ} from the Y kernel's term stream (twojmax <= 8): a term of entry
// e and key (a, b, J) adds cgw * Re(conj(W_e) U_a U_b) to S of its key (the
// double sum of k_snap_s, term by term).  The stream's codes address the
// same bank-shifted U_tot copies as k_snap_yflat; AT atoms per block share
// each term's loads, W of the chunk's entry is read once per chunk, a run of
// one key within a chunk is summed in a register and added to the block's
// LDS sums by one atomic, and the block stores its key sums to its row of
// part (k_snap_s_reduce adds the rows up: no global atomics on the few key
// addresses, and a fixed order).
template <class T, int J, int AT>
__global__ void __launch_bounds__(YF_THREADS)
k_snap_s_flat(int nu, int ny, int nch, const uint32_t *__restrict__ code,
              const unsigned short *__restrict__ key, const double *__restrict__ cgw,
              const unsigned short *__restrict__ centry, const unsigned short *__restrict__ keytid,
              int nkeys, int n, const T *__restrict__ utot_r, const T *__restrict__ utot_i,
              const double *__restrict__ w_r, const double *__restrict__ w_i,
              double *__restrict__ part) {
  using L = yf_layout<T, J>;
  static_assert(L::UBYTES % sizeof(cpair<T>) == 0, "copies aligned");
  constexpr int UP = L::UBYTES / sizeof(cpair<T>);
  constexpr int NY = snap_dim<J>::NY;
  __shared__ cpair<T> s_u[AT * UP];
  __shared__ double s_w[AT][2][NY];
  __shared__ double s_S[SNAP_MAXT];
  unsigned char *sb = reinterpret_cast<unsigned char *>(s_u);
  const int i0 = blockIdx.x * AT;
  if (i0 >= n) return;
  for (int a = 0; a < AT; a++) {
    const int i = i0 + a < n ? i0 + a : i0;
    unsigned char *ub = sb + a * L::UBYTES;
    for (int k = threadIdx.x; k < nu; k += YF_THREADS) {
      const T vr = utot_r[(size_t)i * nu + k], vi = utot_i[(size_t)i * nu + k];
#pragma unroll
      for (int c = 0; c < L::NCOPY; c++) {
        if (L::SOA) {
          reinterpret_cast<T *>(ub)[L::base(c) + k] = vr;
          reinterpret_cast<T *>(ub + L::IMB)[L::base(c) + k] = vi;
        } else {
          reinterpret_cast<cpair<T> *>(ub)[L::base(c) + k] = cpair<T>{vr, vi};
        }
      }
    }
    // the second atom of an odd-sized last block adds nothing
    for (int e = threadIdx.x; e < ny; e += YF_THREADS) {
      s_w[a][0][e] = i0 + a < n ? w_r[(size_t)i * ny + e] : 0.0;
      s_w[a][1][e] = i0 + a < n ? w_i[(size_t)i * ny + e] : 0.0;
    }
  }
  for (int k = threadIdx.x; k < nkeys; k += YF_THREADS) s_S[k] = 0.0;
  __syncthreads();
  for (int c = 0; c < nch; c++) {
    const int e = centry[(size_t)c * YF_THREADS + threadIdx.x];  // 0 for padding chunks
    double wr[AT], wi[AT];
#pragma unroll
    for (int a = 0; a < AT; a++) {
      wr[a] = s_w[a][0][e];
      wi[a] = s_w[a][1][e];
    }
    // a chunk's terms come in runs of one key (padding, weight 0, last):
    // each run is summed here and added to S once
    int run_key = key[(size_t)c * YF_CHUNK * YF_THREADS + threadIdx.x];
    double run = 0.0;
#pragma unroll 4
    for (int k = 0; k < YF_CHUNK; k++) {
      const size_t slot = ((size_t)c * YF_CHUNK + k) * YF_THREADS + threadIdx.x;
      const uint32_t cc = code[slot];
      const double cw = cgw[slot];
      const int kk = key[slot];
      if (kk != run_key && cw != 0.0) {
        atomicAdd(&s_S[run_key], run);
        run = 0.0;
        run_key = kk;
      }
      double v = 0.0;
#pragma unroll
      for (int a = 0; a < AT; a++) {
        const unsigned char *ub = sb + a * L::UBYTES;
        cpair<T> u1, u2;
        if (L::SOA) {
          u1.r = *reinterpret_cast<const T *>(ub + (cc & 0xFFFFu));
          u1.i = *reinterpret_cast<const T *>(ub + L::IMB + (cc & 0xFFFFu));
          u2.r = *reinterpret_cast<const T *>(ub + (cc >> 16));
          u2.i = *reinterpret_cast<const T *>(ub + L::IMB + (cc >> 16));
        } else {
          u1 = *reinterpret_cast<const cpair<T> *>(ub + (cc & 0xFFFFu));
          u2 = *reinterpret_cast<const cpair<T> *>(ub + (cc >> 16));
        }
        const T pr = u1.r * u2.r - u1.i * u2.i;
        const T pi = u1.r * u2.i + u1.i * u2.r;
        v += wr[a] * (double)pr + wi[a] * (double)pi;
      }
      run += cw * v;
    }
    atomicAdd(&s_S[run_key], run);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nkeys; k += YF_THREADS) part[(size_t)blockIdx.x * nkeys + k] = s_S[k];
}

// S of each key: its column of the per-block sums, added by one block in a
// fixed order (strided partial sums, then a tree)
__global__ void __launch_bounds__(256)
k_snap_s_reduce(const double *__restrict__ part, int nrows, int nkeys,
                const unsigned short *__restrict__ keytid, double *__restrict__ S) {
  __shared__ double s_v[256];
  const int k = blockIdx.x;
  double v = 0.0;
  for (int r = threadIdx.x; r < nrows; r += 256) v += part[(size_t)r * nkeys + k];
  s_v[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s_v[threadIdx.x] += s_v[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) S[keytid[k]] += s_v[0];  // one key per block: no other writer
}

// sum over atoms of every bispectrum component B_k (dE/dbeta_k), FP64
__global__ void __launch_bounds__(256)
k_snap_bsum(snap_dev_tables t, const double *__restrict__ cg, const etab_entry *__restrict__ etab,
            int ne, int bzero, int n, const double *__restrict__ utot_r,
            const double *__restrict__ utot_i, double *__restrict__ bsum,
            double *__restrict__ brow) {
  __shared__ double su_r[SNAP_NU_BIG], su_i[SNAP_NU_BIG];
  __shared__ double sB[SNAP_MAXC];
  const int i = blockIdx.x;
  if (i >= n) return;
  for (int k = threadIdx.x; k < t.nu; k += blockDim.x) {
    su_r[k] = utot_r[(size_t)i * t.nu + k];
    su_i[k] = utot_i[(size_t)i * t.nu + k];
  }
  for (int k = threadIdx.x; k < ne; k += blockDim.x) sB[k] = 0.0;
  __syncthreads();
  for (int k = 0; k < ne; k++) {
    const etab_entry te = etab[k];
    const int J = te.j;
    const int nyj = t.ystart[J + 1] - t.ystart[J];
    for (int q = threadIdx.x; q < nyj; q += blockDim.x) {
      const int mb = q / (J + 1), ma = q - mb * (J + 1);
      double wgt = 1.0;
      if (2 * mb == J) wgt = (ma < mb) ? 1.0 : ((ma == mb) ? 0.5 : 0.0);
      if (wgt == 0.0) continue;
      const cplx<double> z = z_entry<double>(t, cg, su_r, su_i, te.j1, te.j2, J, te.cgoff, ma, mb);
      const int u = t.ustart[J] + ma * (J + 1) + mb;
      atomicAdd(&sB[k], 2.0 * wgt * (su_r[u] * z.r + su_i[u] * z.i));
    }
  }
  __syncthreads();
  // the atom's B_k into row i of brow, or added into bsum
  for (int k = threadIdx.x; k < ne; k += blockDim.x) {
    double b = sB[k];
    if (bzero) b -= t.wself * t.wself * t.wself * (etab[k].j + 1);
    if (brow) brow[(size_t)i * ne + k] = b;
    else atomicAdd(&bsum[k], b);
  }
}

__global__ void k_zero_d(double *p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.0;
}

template <class T>
static emd_status snap_adjoint_impl(emd_ctx *ctx, const emd_system *s, const emd_halo *h,
                                    emd_snap *sn, const double *g, double *S) {
  const int n = s->n_local;
  // twojmax <= 8: U_tot came with the compaction (k_snap_neigh_ui)
  if (sn->t.twojmax > SNAP_MAXJ) EMD_TRY(launch_ui<T>(ctx, sn, n, (T *)sn->u_r, (T *)sn->u_i));
  if (sn->t.twojmax <= SNAP_MAXJ) {
    EMD_TRY(snap_par_ready<T>(ctx, sn, n));
    EMD_LAUNCH(ctx, "snap_adjoint_w", k_snap_w_rg<T>, n, 64, 0, sn->t, rg_rp_of<T>(sn),
               sn->rg_map_d, sn->nin, sn->nj, (const snap_par<T> *)sn->par, sn->cap_maxn, n,
               h ? h->root : nullptr, g, sn->w_r, sn->w_i);
  }
  else
    EMD_LAUNCH(ctx, "snap_adjoint_w", (k_snap_w<T, SNAP_JBIG>), n, 64, 0, sn->t, sn->nin, sn->nj,
               sn->nr, sn->cap_maxn, n, h ? h->root : nullptr, g, sn->w_r, sn->w_i);
  const int ntrip = (int)sn->tid.size();
  if (sn->t.twojmax <= SNAP_MAXJ) {
    const uint32_t *code = std::is_same<T, float>::value ? sn->yf_code32 : sn->yf_code64;
    const int nb = (n + 1) / 2, nkeys = (int)sn->yf_keys.size();
    EMD_TRY(emd_ensure(ctx, (void **)&sn->s_part, &sn->cap_spart,
                       (size_t)nb * nkeys * sizeof(double)));
    EMD_LAUNCH(ctx, "snap_adjoint_s", (k_snap_s_flat<T, SNAP_MAXJ, 2>), nb, YF_THREADS, 0,
               sn->t.nu, sn->t.ny, sn->yf_nch, code, sn->yf_key, sn->yf_cgw, sn->yf_centry,
               sn->yf_keytid, nkeys, n, (const T *)sn->u_r, (const T *)sn->u_i, sn->w_r, sn->w_i,
               sn->s_part);
    if (nkeys > 0)
      EMD_LAUNCH(ctx, "snap_adjoint_s", k_snap_s_reduce, nkeys, 256, 0, sn->s_part, nb, nkeys,
                 sn->yf_keytid, S);
  } else {
    EMD_LAUNCH(ctx, "snap_adjoint_s", k_snap_s<T>, n, 192, 0, sn->t, sn->cg, sn->stab,
               (const int *)(sn->sbuf + SNAP_MAXT + SNAP_MAXC), ntrip, n, (const T *)sn->u_r,
               (const T *)sn->u_i, sn->w_r, sn->w_i, S);
  }
  return EMD_OK;
}

extern "C" emd_status emd_snap_adjoint_beta(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                            const emd_halo *h, emd_snap *sn, const double *dLdF,
                                            double dLdE, int precision, double *dLdbeta) {
  if (!ctx || !s || !nl || !sn || !dLdF || !dLdbeta) return EMD_ERR_ARG;
  const int n = s->n_local;
  for (int k = 0; k <= sn->ncoeff; k++) dLdbeta[k] = 0.0;
  if (n <= 0) return EMD_OK;
  EMD_ARG(ctx, s->n_ghost == 0 || (h && h->root && h->n_ghost == s->n_ghost),
          "ghost atoms need the halo (owner of each ghost)");
  const int aprec = precision == EMD_MIXED ? EMD_MIXED : EMD_FP64;
  if (sn->t.twojmax <= SNAP_MAXJ) {
    if (!snap_geom_cached(sn, s, nl, aprec)) {
      EMD_TRY(snap_neigh_ui(ctx, s, nl, sn, -1, aprec));
      snap_geom_record(sn, s, nl, aprec);
    }
  } else {
    EMD_TRY(snap_neigh(ctx, s, nl, sn, -1, aprec));
  }
  EMD_TRY(emd_ensure(ctx, (void **)&sn->w_r, &sn->cap_w, (size_t)n * sn->t.ny * sizeof(double)));
  EMD_TRY(emd_ensure(ctx, (void **)&sn->w_i, &sn->cap_wi, (size_t)n * sn->t.ny * sizeof(double)));
  // sbuf: S[SNAP_MAXT] | Bsum[SNAP_MAXC] | sjstart ints (twojmax + 2 of them)
  constexpr int SB_D = SNAP_MAXT + SNAP_MAXC;
  EMD_TRY(emd_ensure(ctx, (void **)&sn->sbuf, &sn->cap_s, (SB_D + SNAP_JBIG + 2) * sizeof(double)));
  double *S = sn->sbuf, *B = sn->sbuf + SNAP_MAXT;
  EMD_HIP(ctx, hipMemcpyAsync(sn->sbuf + SB_D, sn->sjstart.data(),
                              sn->sjstart.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  EMD_LAUNCH(ctx, "snap_adjoint_s", k_zero_d, emd_blocks(SB_D, 256), 256, 0, sn->sbuf, SB_D);
  if (precision == EMD_MIXED) EMD_TRY(snap_adjoint_impl<float>(ctx, s, h, sn, dLdF, S));
  else EMD_TRY(snap_adjoint_impl<double>(ctx, s, h, sn, dLdF, S));
  // dE/dbeta_k = sum_i B_k^i in FP64
  if (dLdE != 0.0) {
    sn->geom_prec = -1;  // U_tot is rewritten in FP64
    EMD_TRY(launch_ui<double>(ctx, sn, n, (double *)sn->u_r, (double *)sn->u_i));
    EMD_LAUNCH(ctx, "snap_adjoint_bsum", k_snap_bsum, n, 256, 0, sn->t, sn->cg, sn->etab,
               sn->ncoeff, sn->bzeroflag, n, (const double *)sn->u_r, (const double *)sn->u_i, B,
               nullptr);
  }
  std::vector<double> hS(SB_D);
  EMD_HIP(ctx, hipMemcpyAsync(hS.data(), sn->sbuf, hS.size() * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
  EMD_TRY(emd_wait(ctx));
  dLdbeta[0] = dLdE * n;
  for (int k = 0; k < sn->ncoeff; k++) {
    const int j1 = sn->idxj[3 * k], j2 = sn->idxj[3 * k + 1], j = sn->idxj[3 * k + 2];
    const double s1 = hS[sn->tid.at(std::make_tuple(std::max(j1, j2), std::min(j1, j2), j))];
    const double s2 = hS[sn->tid.at(std::make_tuple(std::max(j, j2), std::min(j, j2), j1))];
    const double s3 = hS[sn->tid.at(std::make_tuple(std::max(j1, j), std::min(j1, j), j2))];
    dLdbeta[k + 1] = 2.0 * s1 + 2.0 * (j + 1) / (j1 + 1.0) * s2 + 2.0 * (j + 1) / (j2 + 1.0) * s3 +
                     dLdE * hS[SNAP_MAXT + k];
  }
  return EMD_OK;
}

extern "C" emd_status emd_snap_descriptors(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                           emd_snap *sn, double *B) {
  if (!ctx || !s || !nl || !sn || !B) return EMD_ERR_ARG;
  const int n = s->n_local;
  if (n <= 0) return EMD_OK;
  sn->geom_prec = -1;  // the compaction and U_tot are rewritten
  EMD_TRY(snap_neigh(ctx, s, nl, sn, -1, EMD_FP64));
  EMD_TRY(launch_ui<double>(ctx, sn, n, (double *)sn->u_r, (double *)sn->u_i));
  EMD_LAUNCH(ctx, "snap_descriptors", k_snap_bsum, n, 256, 0, sn->t, sn->cg, sn->etab,
             sn->ncoeff, sn->bzeroflag, n, (const double *)sn->u_r, (const double *)sn->u_i,
             nullptr, B);
  return EMD_OK;
}
