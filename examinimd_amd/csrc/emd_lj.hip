// emd_lj.hip -- LJ pair force/energy, force zeroing, NVE and thermo sums.
//
// LJ force: K lanes per atom (K in {1,2,4,8,16}); the lanes of one atom read
// consecutive CSR entries (coalesced) and gather x_j, then reduce their
// partial forces with cross-lane shuffles.  FP64 path: separate multiply and
// add roundings (no contraction) so that K == 1 reproduces the reference's
// serial arithmetic exactly.  Mixed path: the distance and the cutoff test
// stay FP64 (identical pair set), the pair arithmetic and the per-atom
// accumulation are FP32, the result goes to the FP32 force array f32.
#include "emd_internal.hpp"

constexpr int LJ_THREADS = 256;
constexpr int LJ_MAXT = 16;

struct lj_table {
  double lj1[LJ_MAXT * LJ_MAXT], lj2[LJ_MAXT * LJ_MAXT], cutsq[LJ_MAXT * LJ_MAXT];
  double rep[LJ_MAXT * LJ_MAXT];  // lj/cut/idial repeat counts (1 for lj/cut)
};

// lj/cut/idial's repeat loop (src/force_types/force_lj_idial_neigh_impl.h:
// 149-155): fpair = 0, then `rep` times fpair += pair_force / rep.  The part
// is loop-invariant (an optimising compiler hoists it out of the
// reference's loop as well); what remains are the sequential additions
// from 0, done here in the same order, so the sum is the reference's to the
// bit.  rep is a double in the reference's table, compared against the int
// counter: `r < rep` holds for ceil(rep) values of r >= 0 (rep is finite and
// at most 1e6, checked on the host), so the loop runs on an int trip count
// (no per-pass conversion and FP64 compare), eight additions per pass.
__device__ __forceinline__ int dial_trips(double rep) { return rep > 0 ? (int)ceil(rep) : 0; }

template <class T>
__device__ __forceinline__ T dial_sum(T part, int trips) {
  T sum = 0;
  int r = 0;
  for (; r + 8 <= trips; r += 8) {  // the same additions in the same order, 8 per pass
#pragma unroll
    for (int k = 0; k < 8; k++) sum += part;
  }
  for (; r < trips; r++) sum += part;
  return sum;
}

template <class T>
__global__ void k_zero3(T *__restrict__ f, long n3) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) f[i] = T(0);
}

extern "C" emd_status emd_force_zero(emd_ctx *ctx, emd_system *s, int precision) {
  if (!ctx || !s) return EMD_ERR_ARG;
  const long n3 = 3l * (s->n_local + s->n_ghost);
  if (n3 <= 0) return EMD_OK;
  if (precision == EMD_MIXED)
    EMD_LAUNCH(ctx, "force_zero", k_zero3<float>, emd_blocks(n3, 256), 256, 0, s->f32, n3);
  else
    EMD_LAUNCH(ctx, "force_zero", k_zero3<double>, emd_blocks(n3, 256), 256, 0, s->f, n3);
  return EMD_OK;
}

// ForceLJNeigh TagFullNeigh, src/force_types/force_lj_neigh_impl.h:158-206.
// K lanes per atom; each lane takes every K-th entry of the row.  The kernel
// is bound by memory latency (index load, then the dependent position
// gather), so a lane issues the index loads of U entries, then their U
// gathers, before any arithmetic: one round trip covers U*K entries (16 x 5
// holds a typical 75-entry row).  Slots past the row end re-read its last
// entry and are masked.  Row indices are loaded non-temporally (read once,
// kept out of L2 so positions stay resident).
template <bool MIXED>
struct lj_acc {
  typedef double T;
};
template <>
struct lj_acc<true> {
  typedef float T;
};

template <int K, int U, bool ONE_TYPE, bool MIXED, bool DIAL>
__global__ void __launch_bounds__(LJ_THREADS)
k_lj_full(const double *__restrict__ x, const int *__restrict__ type,
          const int *__restrict__ row, const int *__restrict__ num,
          const int *__restrict__ entries, int pitch, int n, double lj1_0, double lj2_0,
          double cutsq_0, double rep_0, const lj_table *__restrict__ tab,
          const int *__restrict__ visit, int n_visit, int store, double *__restrict__ f,
          float *__restrict__ f32, int prio) {
#pragma clang fp contract(off)
  typedef typename lj_acc<MIXED>::T T;
  const int tile = emd_xcd_tile(blockIdx.x, gridDim.x);
  const int gid = tile * LJ_THREADS + threadIdx.x;
  const int a = gid / K, sub = gid % K;
  // atoms in the list's sweep order (column tiles) when it has one; the
  // order may name a few ghosts, which have no row
  if (a >= (visit ? n_visit : n)) return;
  const int i = visit ? visit[a] : a;
  if (i >= n) return;
  const double x_i = x[3 * (size_t)i], y_i = x[3 * (size_t)i + 1], z_i = x[3 * (size_t)i + 2];
  const int type_i = ONE_TYPE ? 0 : type[i];
  const int *rowp = entries + (row ? (size_t)row[i] : (size_t)i * (size_t)pitch);
  const int cnt = num[i];
  T fxi = 0, fyi = 0, fzi = 0;
  auto pair = [&](bool ok, int j, double xj, double yj, double zj) {
    const double dx = x_i - xj;
    const double dy = y_i - yj;
    const double dz = z_i - zj;
    const double rsq = dx * dx + dy * dy + dz * dz;
    double cutsq = cutsq_0, lj1 = lj1_0, lj2 = lj2_0, nrep = rep_0;
    if (!ONE_TYPE) {
      const int t = type_i * LJ_MAXT + type[j];
      cutsq = tab->cutsq[t]; lj1 = tab->lj1[t]; lj2 = tab->lj2[t];
      if (DIAL) nrep = tab->rep[t];
    }
    if (ok && rsq < cutsq) {
      if (MIXED) {
        float fpair = emd_lj_fpair_mixed(rsq, (float)lj1, (float)lj2);
        if (DIAL) fpair = dial_sum(fpair / (float)nrep, dial_trips(nrep));
        fxi = __builtin_fmaf((float)dx, fpair, fxi);
        fyi = __builtin_fmaf((float)dy, fpair, fyi);
        fzi = __builtin_fmaf((float)dz, fpair, fzi);
      } else {
        const double r2inv = 1.0 / rsq;
        const double r6inv = r2inv * r2inv * r2inv;
        double fpair = (r6inv * (lj1 * r6inv - lj2)) * r2inv;
        if (DIAL) fpair = dial_sum(fpair / nrep, dial_trips(nrep));
        fxi += dx * fpair;
        fyi += dy * fpair;
        fzi += dz * fpair;
      }
    }
  };
  auto gather_pairs = [&](int jj, const int (&j)[U], const bool (&ok)[U]) {
    double xj[U], yj[U], zj[U];
    // while a wavefront issues a round's loads it runs at a raised priority,
    // so the loads of waves about to wait start ahead of other waves'
    // arithmetic (LJ-4M mixed: 0.693 -> 0.681 ms; FP64 0.731 -> 0.737 ms,
    // so off there by default)
    if (prio) __builtin_amdgcn_s_setprio(2);
#pragma unroll
    for (int u = 0; u < U; u++) {
      xj[u] = x[3 * (size_t)j[u]];
      yj[u] = x[3 * (size_t)j[u] + 1];
      zj[u] = x[3 * (size_t)j[u] + 2];
    }
    if (prio) __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int u = 0; u < U; u++) pair(ok[u], j[u], xj[u], yj[u], zj[u]);
  };
  int jj = sub;
  if (!row) {
    // 2D rows start at i * pitch, so the first round's index loads need not
    // wait for the row length (loaded alongside them): slots are clamped to
    // the pitch, and a slot past the row end (an index of an earlier build,
    // or never written) is replaced by the atom itself before its position
    // is gathered (x_i is in cache; the slot is masked by ok)
    int j[U];
    bool ok[U];
    if (prio > 1) __builtin_amdgcn_s_setprio(2);
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = jj + u * K;
      j[u] = __builtin_nontemporal_load(rowp + (e < pitch ? e : pitch - 1));
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      ok[u] = jj + u * K < cnt;
      j[u] = ok[u] ? j[u] : i;
    }
    gather_pairs(jj, j, ok);
    jj += U * K;
  }
  for (; jj < cnt; jj += U * K) {
    int j[U];
    bool ok[U];
    if (prio > 1) __builtin_amdgcn_s_setprio(2);
#pragma unroll
    for (int u = 0; u < U; u++) {
      ok[u] = jj + u * K < cnt;
      j[u] = __builtin_nontemporal_load(rowp + (ok[u] ? jj + u * K : cnt - 1));
    }
    gather_pairs(jj, j, ok);
  }
  if (K > 1) {
    fxi = group_sum<K>(fxi);
    fyi = group_sum<K>(fyi);
    fzi = group_sum<K>(fzi);
  }
  if (sub == 0) {
    // store: the row is the atom's whole force (full list), written without
    // reading the zeroed array first (0 + v == v)
    if (MIXED) {
      float *o = f32 + 3 * (size_t)i;
      if (store) { o[0] = (float)fxi; o[1] = (float)fyi; o[2] = (float)fzi; }
      else { o[0] += (float)fxi; o[1] += (float)fyi; o[2] += (float)fzi; }
    } else {
      double *o = f + 3 * (size_t)i;
      if (store) { o[0] = fxi; o[1] = fyi; o[2] = fzi; }
      else { o[0] += fxi; o[1] += fyi; o[2] += fzi; }
    }
  }
}

// Software-pipelined form of k_lj_full (one type, full CSR or 2D rows).
// Each wavefront walks groups of 64/K atoms through its XCD's contiguous
// share of the atoms (blocks are dealt to the 8 XCDs round-robin; the share
// only affects cache locality, every group is done exactly once).  While a
// wavefront gathers positions for group t and computes it, the index loads
// of group t+1 and the row offsets of group t+2 are already in flight, so
// the latency of the HBM index stream overlaps arithmetic.  The first U*K
// entries of a row take the pipelined round, longer rows finish in a plain
// loop.  Per lane the entries and their order are those of k_lj_full<K,U>
// and the reduction is the same: the FP64 result is bitwise equal.
constexpr int PIPE_WAVES = 4;

template <int K, int U, bool MIXED, bool CSR>
__global__ void __launch_bounds__(64 * PIPE_WAVES)
k_lj_pipe(const double *__restrict__ x, const int *__restrict__ row, const int *__restrict__ num,
          const int *__restrict__ entries, int pitch, int n, double lj1, double lj2,
          double cutsq, int groups_per_xcd, int waves_per_xcd, double *__restrict__ f,
          float *__restrict__ f32) {
#pragma clang fp contract(off)
  typedef typename lj_acc<MIXED>::T T;
  constexpr int G = 64 / K;  // atoms per wavefront group
  const int lane = threadIdx.x % 64, sub = lane % K, at = lane / K;
  const int xcd = blockIdx.x % 8;
  const int g_lo = xcd * groups_per_xcd;
  const int g_hi = min((n + G - 1) / G, g_lo + groups_per_xcd);
  int g = g_lo + (blockIdx.x / 8) * PIPE_WAVES + threadIdx.x / 64;
  if (g >= g_hi) return;  // uniform over the wavefront
  // All loads below are unconditional, from addresses kept in range by
  // clamping (padding atoms of the last group read atom n-1, slots past a
  // row's end re-read its start, which lies inside the entries allocation),
  // and invalid results are discarded by selects, so no branch separates a
  // load from the loads issued after it.
  auto row_info = [&](int gg, long &base, int &cnt) {
    const int i = gg * G + at;
    const int ic = i < n ? i : n - 1;
    base = CSR ? (long)row[ic] : (long)ic * pitch;
    const int c = num[ic];
    cnt = i < n ? c : 0;
  };
  auto load_idx = [&](int gg, long base, int cnt, int *j, bool *ok) {
    const int i = gg * G + at;
    const int ic = i < n ? i : n - 1;
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = sub + u * K;
      ok[u] = e < cnt;
      const int v = __builtin_nontemporal_load(entries + base + (ok[u] ? e : 0));
      j[u] = ok[u] ? v : ic;
    }
  };
  // groups past the end are replaced by the current one for the prefetches
  // (their results go unused), keeping every load unconditional
  long b0, b1;
  int c0, c1;
  row_info(g, b0, c0);
  int j0[U];
  bool ok0[U];
  load_idx(g, b0, c0, j0, ok0);
  int g1 = g + waves_per_xcd;
  row_info(g1 < g_hi ? g1 : g, b1, c1);
  while (true) {
    const int i = g * G + at;
    const int ic = i < n ? i : n - 1;
    // gathers of group g, its own position and current force
    double xj[U], yj[U], zj[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      xj[u] = x[3 * (size_t)j0[u]];
      yj[u] = x[3 * (size_t)j0[u] + 1];
      zj[u] = x[3 * (size_t)j0[u] + 2];
    }
    const double x_i = x[3 * (size_t)ic], y_i = x[3 * (size_t)ic + 1], z_i = x[3 * (size_t)ic + 2];
    T f0x, f0y, f0z;
    if (MIXED) {
      f0x = f32[3 * (size_t)ic];
      f0y = f32[3 * (size_t)ic + 1];
      f0z = f32[3 * (size_t)ic + 2];
    } else {
      f0x = f[3 * (size_t)ic];
      f0y = f[3 * (size_t)ic + 1];
      f0z = f[3 * (size_t)ic + 2];
    }
    // keep the prefetches below behind this group's loads in the memory
    // queue (waits on vector loads retire in issue order)
    __builtin_amdgcn_sched_barrier(0);
    // next group's indices and the row offsets of the one after
    int j1[U];
    bool ok1[U];
    const int g2 = g1 + waves_per_xcd;
    long b2;
    int c2;
    load_idx(g1 < g_hi ? g1 : g, b1, c1, j1, ok1);
    row_info(g2 < g_hi ? g2 : g, b2, c2);
    T fxi = 0, fyi = 0, fzi = 0;
    auto pair = [&](bool ok, double xj_, double yj_, double zj_) {
      const double dx = x_i - xj_;
      const double dy = y_i - yj_;
      const double dz = z_i - zj_;
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (ok && rsq < cutsq) {
        if (MIXED) {
          const float fpair = emd_lj_fpair_mixed(rsq, (float)lj1, (float)lj2);
          fxi = __builtin_fmaf((float)dx, fpair, fxi);
          fyi = __builtin_fmaf((float)dy, fpair, fyi);
          fzi = __builtin_fmaf((float)dz, fpair, fzi);
        } else {
          const double r2inv = 1.0 / rsq;
          const double r6inv = r2inv * r2inv * r2inv;
          const double fpair = (r6inv * (lj1 * r6inv - lj2)) * r2inv;
          fxi += dx * fpair;
          fyi += dy * fpair;
          fzi += dz * fpair;
        }
      }
    };
#pragma unroll
    for (int u = 0; u < U; u++) pair(ok0[u], xj[u], yj[u], zj[u]);
    // rows longer than one round
    for (int jj = sub + U * K; jj < c0; jj += K) {
      const int j = __builtin_nontemporal_load(entries + b0 + jj);
      pair(true, x[3 * (size_t)j], x[3 * (size_t)j + 1], x[3 * (size_t)j + 2]);
    }
    if (K > 1) {
      fxi = group_sum<K>(fxi);
      fyi = group_sum<K>(fyi);
      fzi = group_sum<K>(fzi);
    }
    if (sub == 0 && i < n) {
      if (MIXED) {
        f32[3 * (size_t)i] = f0x + (float)fxi;
        f32[3 * (size_t)i + 1] = f0y + (float)fyi;
        f32[3 * (size_t)i + 2] = f0z + (float)fzi;
      } else {
        f[3 * (size_t)i] = f0x + fxi;
        f[3 * (size_t)i + 1] = f0y + fyi;
        f[3 * (size_t)i + 2] = f0z + fzi;
      }
    }
    if (g1 >= g_hi) break;
    g = g1;
    g1 = g2;
    b0 = b1;
    c0 = c1;
    b1 = b2;
    c1 = c2;
#pragma unroll
    for (int u = 0; u < U; u++) {
      j0[u] = j1[u];
      ok0[u] = ok1[u];
    }
  }
}

// Bin-tile form of the same full-list force (one type).  One block per
// interior bin: the block copies the positions of the bin's 27-bin
// candidate list (the order the neighbour build enumerated them in) into
// LDS once, then every row entry's slot addresses its neighbour's position
// there, so the only global traffic per pair is the 2-byte slot.  The lanes
// of an atom take the same entries in the same order as k_lj_full with the
// same K, and reduce with the same shuffles: the FP64 result is bitwise that
// of k_lj_full<K>.
template <int K, int U, bool MIXED>
__global__ void __launch_bounds__(LJ_THREADS)
k_lj_tile(const double *__restrict__ x, const int *__restrict__ bincount,
          const int *__restrict__ binoffsets, const int *__restrict__ permute, neigh_geom g,
          const int *__restrict__ row, const int *__restrict__ num,
          const unsigned short *__restrict__ slot, int pitch, int n_local, int tcap, double lj1,
          double lj2, double cutsq, double *__restrict__ f, float *__restrict__ f32) {
#pragma clang fp contract(off)
  typedef typename lj_acc<MIXED>::T T;
  extern __shared__ double s_cand[];
  double *s_x = s_cand, *s_y = s_cand + tcap, *s_z = s_cand + 2 * tcap;
  __shared__ int s_off[27], s_pre[28];
  const int lr = emd_xcd_tile(blockIdx.x, gridDim.x);
  emd_stencil_setup(g, lr, bincount, binoffsets, s_off, s_pre);
  const int ncand = s_pre[27];
  for (int c = threadIdx.x; c < ncand; c += LJ_THREADS) {
    const int k = emd_stencil_slot(s_pre, c);
    const size_t j = permute[s_off[k] + (c - s_pre[k])];
    s_x[c] = x[3 * j];
    s_y[c] = x[3 * j + 1];
    s_z[c] = x[3 * j + 2];
  }
  __syncthreads();
  const int c0 = s_pre[13], c_cnt = s_pre[14] - s_pre[13], c_off = s_off[13];
  const int sub = threadIdx.x % K;
  for (int a = threadIdx.x / K; a < c_cnt; a += LJ_THREADS / K) {
    const int i = permute[c_off + a];
    if (i >= n_local) continue;
    const double x_i = s_x[c0 + a], y_i = s_y[c0 + a], z_i = s_z[c0 + a];
    const unsigned short *rowp = slot + (row ? (size_t)row[i] : (size_t)i * (size_t)pitch);
    const int cnt = num[i];
    T fxi = 0, fyi = 0, fzi = 0;
    auto pair = [&](bool ok, int c) {
      const double dx = x_i - s_x[c];
      const double dy = y_i - s_y[c];
      const double dz = z_i - s_z[c];
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (ok && rsq < cutsq) {
        if (MIXED) {
          const float fpair = emd_lj_fpair_mixed(rsq, (float)lj1, (float)lj2);
          fxi = __builtin_fmaf((float)dx, fpair, fxi);
          fyi = __builtin_fmaf((float)dy, fpair, fyi);
          fzi = __builtin_fmaf((float)dz, fpair, fzi);
        } else {
          const double r2inv = 1.0 / rsq;
          const double r6inv = r2inv * r2inv * r2inv;
          const double fpair = (r6inv * (lj1 * r6inv - lj2)) * r2inv;
          fxi += dx * fpair;
          fyi += dy * fpair;
          fzi += dz * fpair;
        }
      }
    };
    for (int jj = sub; jj < cnt; jj += U * K) {
      int c[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        ok[u] = jj + u * K < cnt;
        c[u] = __builtin_nontemporal_load(rowp + (ok[u] ? jj + u * K : cnt - 1));
      }
#pragma unroll
      for (int u = 0; u < U; u++) pair(ok[u], c[u]);
    }
    if (K > 1) {
      fxi = group_sum<K>(fxi);
      fyi = group_sum<K>(fyi);
      fzi = group_sum<K>(fzi);
    }
    if (sub == 0) {
      if (MIXED) {
        f32[3 * (size_t)i] += (float)fxi;
        f32[3 * (size_t)i + 1] += (float)fyi;
        f32[3 * (size_t)i + 2] += (float)fzi;
      } else {
        f[3 * (size_t)i] += fxi;
        f[3 * (size_t)i + 1] += fyi;
        f[3 * (size_t)i + 2] += fzi;
      }
    }
  }
}

// ForceLJNeigh TagHalfNeigh, force_lj_neigh_impl.h:208-254: each pair's
// force is added to i and subtracted from j (atomically: rows of other
// atoms reach j too).  LJH_K lanes per atom take every LJH_K-th entry of the
// row, issuing the index loads of LJH_U entries and then their gathers
// before the arithmetic (a half row holds ~37 entries: one round); the
// lanes' i sums are reduced across the group and added once.  Pair
// arithmetic as the full-list kernel: FP64 as the reference writes it,
// mixed with FP64 distances and emd_lj_fpair_mixed.  DIAL: ForceLJIDialNeigh
// TagHalfNeigh, force_lj_idial_neigh_impl.h:169-219, which subtracts the
// reaction only from local j (:208-212).  Atomic order varies from run to
// run, so sums agree to rounding, not bitwise.
constexpr int LJH_K = 8, LJH_U = 5;

template <bool MIXED, bool DIAL>
__global__ void __launch_bounds__(LJ_THREADS)
k_lj_half(const double *__restrict__ x, const int *__restrict__ type, const int *__restrict__ row,
          const int *__restrict__ num, const int *__restrict__ entries, int pitch, int n,
          const lj_table *__restrict__ tab, double *__restrict__ f, float *__restrict__ f32) {
#pragma clang fp contract(off)
  typedef typename lj_acc<MIXED>::T T;
  const int gid = blockIdx.x * LJ_THREADS + threadIdx.x;
  const int i = gid / LJH_K, sub = gid % LJH_K;
  if (i >= n) return;  // whole groups of LJH_K lanes leave together
  const double x_i = x[3 * (size_t)i], y_i = x[3 * (size_t)i + 1], z_i = x[3 * (size_t)i + 2];
  const int type_i = type[i];
  const int *rowp = entries + (row ? (size_t)row[i] : (size_t)i * (size_t)pitch);
  const int cnt = num[i];
  T fxi = 0, fyi = 0, fzi = 0;
  auto sub_j = [&](int j, T vx, T vy, T vz) {
    if (MIXED) {
      atomicAdd(&f32[3 * (size_t)j], -(float)vx);
      atomicAdd(&f32[3 * (size_t)j + 1], -(float)vy);
      atomicAdd(&f32[3 * (size_t)j + 2], -(float)vz);
    } else {
      atomicAdd(&f[3 * (size_t)j], -(double)vx);
      atomicAdd(&f[3 * (size_t)j + 1], -(double)vy);
      atomicAdd(&f[3 * (size_t)j + 2], -(double)vz);
    }
  };
  for (int jj = sub; jj < cnt; jj += LJH_U * LJH_K) {
    int j[LJH_U];
    bool ok[LJH_U];
#pragma unroll
    for (int u = 0; u < LJH_U; u++) {
      ok[u] = jj + u * LJH_K < cnt;
      j[u] = rowp[ok[u] ? jj + u * LJH_K : cnt - 1];
    }
    double xj[LJH_U], yj[LJH_U], zj[LJH_U];
#pragma unroll
    for (int u = 0; u < LJH_U; u++) {
      xj[u] = x[3 * (size_t)j[u]];
      yj[u] = x[3 * (size_t)j[u] + 1];
      zj[u] = x[3 * (size_t)j[u] + 2];
    }
#pragma unroll
    for (int u = 0; u < LJH_U; u++) {
      const double dx = x_i - xj[u];
      const double dy = y_i - yj[u];
      const double dz = z_i - zj[u];
      const double rsq = dx * dx + dy * dy + dz * dz;
      const int t = type_i * LJ_MAXT + type[j[u]];
      if (ok[u] && rsq < tab->cutsq[t]) {
        T vx, vy, vz;
        if (MIXED) {
          float fpair = emd_lj_fpair_mixed(rsq, (float)tab->lj1[t], (float)tab->lj2[t]);
          if (DIAL) fpair = dial_sum(fpair / (float)tab->rep[t], dial_trips(tab->rep[t]));
          vx = (float)dx * fpair;
          vy = (float)dy * fpair;
          vz = (float)dz * fpair;
        } else {
          const double r2inv = 1.0 / rsq;
          const double r6inv = r2inv * r2inv * r2inv;
          double fpair = (r6inv * (tab->lj1[t] * r6inv - tab->lj2[t])) * r2inv;
          if (DIAL) fpair = dial_sum(fpair / tab->rep[t], dial_trips(tab->rep[t]));
          vx = dx * fpair;
          vy = dy * fpair;
          vz = dz * fpair;
        }
        fxi += vx;
        fyi += vy;
        fzi += vz;
        if (!DIAL || j[u] < n) sub_j(j[u], vx, vy, vz);
      }
    }
  }
  fxi = group_sum<LJH_K>(fxi);
  fyi = group_sum<LJH_K>(fyi);
  fzi = group_sum<LJH_K>(fzi);
  if (sub == 0) {
    if (MIXED) {
      atomicAdd(&f32[3 * (size_t)i], (float)fxi);
      atomicAdd(&f32[3 * (size_t)i + 1], (float)fyi);
      atomicAdd(&f32[3 * (size_t)i + 2], (float)fzi);
    } else {
      atomicAdd(&f[3 * (size_t)i], (double)fxi);
      atomicAdd(&f[3 * (size_t)i + 1], (double)fyi);
      atomicAdd(&f[3 * (size_t)i + 2], (double)fzi);
    }
  }
}

static emd_status upload_table(emd_ctx *ctx, const emd_lj_params *p,
                               const double (*rep)[16], lj_table **dtab) {
  static_assert(sizeof(lj_table) <= 8 * 1024, "table size");
  lj_table h;
  memset(&h, 0, sizeof(h));
  for (int a = 0; a < p->ntypes; a++)
    for (int b = 0; b < p->ntypes; b++) {
      h.lj1[a * LJ_MAXT + b] = p->lj1[a][b];
      h.lj2[a * LJ_MAXT + b] = p->lj2[a][b];
      h.cutsq[a * LJ_MAXT + b] = p->cutsq[a][b];
      h.rep[a * LJ_MAXT + b] = rep ? rep[a][b] : 1.0;
    }
  // the table lives in the small double scratch area's neighbour: d_red head
  EMD_TRY(emd_ensure(ctx, (void **)&ctx->d_red, &ctx->red_cap, sizeof(lj_table) + (1 << 20)));
  EMD_HIP(ctx, hipMemcpyAsync(ctx->d_red, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *dtab = (lj_table *)ctx->d_red;
  return EMD_OK;
}

// lanes per atom and gathers per round (measured on MI355X, LJ 4M atoms)
static int default_lanes(int precision) { return precision == EMD_MIXED ? 16 : 16; }

extern "C" emd_status emd_lj_set_tile(emd_ctx *ctx, int on, int lanes) {
  if (!ctx) return EMD_ERR_ARG;
  if (!(lanes == 0 || lanes == 4 || lanes == 8 || lanes == 16))
    return emd_set_error(ctx, EMD_ERR_ARG, "bin-tile lanes must be 0, 4, 8 or 16");
  ctx->lj_tile = on != 0;
  if (lanes) ctx->lj_tile_lanes = lanes;
  return EMD_OK;
}

extern "C" emd_status emd_lj_set_lanes(emd_ctx *ctx, int precision, int lanes) {
  if (!ctx || (precision != EMD_FP64 && precision != EMD_MIXED)) return EMD_ERR_ARG;
  if (!(lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16))
    return emd_set_error(ctx, EMD_ERR_ARG, "lanes must be 0,1,2,4,8 or 16");
  ctx->lj_lanes[precision] = lanes;
  return EMD_OK;
}

template <int K, int U>
static emd_status launch_pipe(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                              const emd_lj_params *p, int precision) {
  const int n = s->n_local;
  constexpr int G = 64 / K;
  const int ngroups = (n + G - 1) / G;
  const int gpx = (ngroups + 7) / 8;
  int bpx = ctx->lj_pipe_blocks;  // blocks per XCD
  const int need = (gpx + PIPE_WAVES - 1) / PIPE_WAVES;
  if (bpx > need) bpx = need;
  if (bpx < 1) bpx = 1;
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
#define LJ_PIPE_ARGS                                                                        \
  s->x, row, nl->num, nl->entries, nl->maxneighs, n, p->lj1[0][0], p->lj2[0][0],            \
      p->cutsq[0][0], gpx, bpx * PIPE_WAVES, s->f, s->f32
  const bool csr = nl->layout == EMD_NEIGH_CSR;
  if (precision == EMD_MIXED) {
    if (csr) EMD_LAUNCH(ctx, "lj_force_mixed", (k_lj_pipe<K, U, true, true>), 8 * bpx, 64 * PIPE_WAVES, 0, LJ_PIPE_ARGS);
    else EMD_LAUNCH(ctx, "lj_force_mixed", (k_lj_pipe<K, U, true, false>), 8 * bpx, 64 * PIPE_WAVES, 0, LJ_PIPE_ARGS);
  } else {
    if (csr) EMD_LAUNCH(ctx, "lj_force_fp64", (k_lj_pipe<K, U, false, true>), 8 * bpx, 64 * PIPE_WAVES, 0, LJ_PIPE_ARGS);
    else EMD_LAUNCH(ctx, "lj_force_fp64", (k_lj_pipe<K, U, false, false>), 8 * bpx, 64 * PIPE_WAVES, 0, LJ_PIPE_ARGS);
  }
#undef LJ_PIPE_ARGS
  return EMD_OK;
}

template <int K, int U, bool DIAL>
static emd_status launch_full_as(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                 const emd_lj_params *p, double rep_0, int precision,
                                 const lj_table *tab, int store) {
  const int n = s->n_local;
  const int nb = emd_blocks((long)(nl->visit ? nl->n_visit : n) * K, LJ_THREADS);
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
  const bool one = p->ntypes == 1;
  const int prio = ctx->lj_prio >= 0 ? ctx->lj_prio : (precision == EMD_MIXED ? 2 : 0);
  const char *name_m = DIAL ? "lj_idial_mixed" : "lj_force_mixed";
  const char *name_d = DIAL ? "lj_idial_fp64" : "lj_force_fp64";
#define LJ_FULL_ARGS                                                                        \
  s->x, s->type, row, nl->num, nl->entries, nl->maxneighs, n, p->lj1[0][0], p->lj2[0][0],   \
      p->cutsq[0][0], rep_0, tab, nl->visit, nl->n_visit, store, s->f, s->f32, prio
  if (precision == EMD_MIXED) {
    if (one) EMD_LAUNCH(ctx, name_m, (k_lj_full<K, U, true, true, DIAL>), nb, LJ_THREADS, 0, LJ_FULL_ARGS);
    else EMD_LAUNCH(ctx, name_m, (k_lj_full<K, U, false, true, DIAL>), nb, LJ_THREADS, 0, LJ_FULL_ARGS);
  } else {
    if (one) EMD_LAUNCH(ctx, name_d, (k_lj_full<K, U, true, false, DIAL>), nb, LJ_THREADS, 0, LJ_FULL_ARGS);
    else EMD_LAUNCH(ctx, name_d, (k_lj_full<K, U, false, false, DIAL>), nb, LJ_THREADS, 0, LJ_FULL_ARGS);
  }
#undef LJ_FULL_ARGS
  return EMD_OK;
}

template <int K, int U>
static emd_status launch_full(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                              const emd_lj_params *p, int precision, const lj_table *tab,
                              int store) {
  return launch_full_as<K, U, false>(ctx, s, nl, p, 1.0, precision, tab, store);
}

// LDS budget of the bin-tile kernel: three FP64 coordinates per candidate
// in 60 KiB (the static stencil tables take the rest of the 64 KiB)
constexpr int LJ_TILE_MAX_CAND = (60 * 1024) / (3 * 8);

template <int K, int U>
static emd_status launch_tile(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                              const emd_lj_params *p, int precision, int tcap) {
  const emd_bins *b = nl->bins;
  const neigh_geom g = emd_neigh_geom(b);
  const long nib = (long)g.nix * g.niy * g.niz;
  const size_t shmem = 3 * (size_t)tcap * sizeof(double);
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
#define LJ_TILE_ARGS                                                                          \
  s->x, b->bincount, b->binoffsets, b->permute, g, row, nl->num, nl->slot, nl->maxneighs,     \
      s->n_local, tcap, p->lj1[0][0], p->lj2[0][0], p->cutsq[0][0], s->f, s->f32
  if (precision == EMD_MIXED)
    EMD_LAUNCH(ctx, "lj_force_mixed", (k_lj_tile<K, U, true>), nib, LJ_THREADS, shmem, LJ_TILE_ARGS);
  else
    EMD_LAUNCH(ctx, "lj_force_fp64", (k_lj_tile<K, U, false>), nib, LJ_THREADS, shmem, LJ_TILE_ARGS);
#undef LJ_TILE_ARGS
  return EMD_OK;
}

// the bin-tile kernel applies to one-type full lists that carry slots, when
// the largest candidate list (27 x the fullest bin) fits the LDS budget
static int tile_capacity(const emd_ctx *ctx, const emd_neigh *nl, const emd_lj_params *p,
                         int half) {
  if (!ctx->lj_tile || half || p->ntypes != 1 || !nl->slot || !nl->bins) return 0;
  const emd_bins *b = nl->bins;
  if (b->nhalo < 1 || b->nbinx - 2 * b->nhalo < 1 || b->nbiny - 2 * b->nhalo < 1 ||
      b->nbinz - 2 * b->nhalo < 1)
    return 0;
  const long t = (27l * (b->max_count > 0 ? b->max_count : 1) + 63) / 64 * 64;
  return t <= LJ_TILE_MAX_CAND ? (int)t : 0;
}

// lj/cut/idial: the full-list kernel with the repeat loop, for any lanes
// per atom (the bin-tile and pipelined forms are lj/cut only)
template <bool DIAL>
static emd_status launch_full_lanes(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                    const emd_lj_params *p, double rep_0, int precision,
                                    const lj_table *tab, int store, int K) {
  switch (K) {
    case 1: return launch_full_as<1, 4, DIAL>(ctx, s, nl, p, rep_0, precision, tab, store);
    case 2: return launch_full_as<2, 4, DIAL>(ctx, s, nl, p, rep_0, precision, tab, store);
    case 4: return launch_full_as<4, 4, DIAL>(ctx, s, nl, p, rep_0, precision, tab, store);
    case 8: return launch_full_as<8, 5, DIAL>(ctx, s, nl, p, rep_0, precision, tab, store);
    default: return launch_full_as<16, 5, DIAL>(ctx, s, nl, p, rep_0, precision, tab, store);
  }
}

static emd_status lj_idial_force(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                 const emd_lj_idial_params *p, int precision, int half,
                                 int store) {
  if (!ctx || !s || !nl || !p) return EMD_ERR_ARG;
  const emd_lj_params *lj = &p->lj;
  EMD_ARG(ctx, lj->ntypes >= 1 && lj->ntypes <= LJ_MAXT, "ntypes out of range");
  EMD_ARG(ctx, nl->n == s->n_local, "neighbour list does not match the system");
  EMD_ARG(ctx, !(store && half), "storing forces needs a full list");
  for (int a = 0; a < lj->ntypes; a++)
    for (int b = 0; b < lj->ntypes; b++)
      EMD_ARG(ctx, p->intensity[a][b] <= 1e6, "lj/cut/idial intensity above 1e6");
  if (s->n_local <= 0) return EMD_OK;
  lj_table *tab = nullptr;
  if (lj->ntypes > 1 || half) EMD_TRY(upload_table(ctx, lj, p->intensity, &tab));
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
  if (half) {
    const int nb = emd_blocks((long)s->n_local * LJH_K, LJ_THREADS);
    if (precision == EMD_MIXED)
      EMD_LAUNCH(ctx, "lj_idial_half", (k_lj_half<true, true>), nb, LJ_THREADS, 0, s->x,
                 s->type, row, nl->num, nl->entries, nl->maxneighs, s->n_local, tab, s->f,
                 s->f32);
    else
      EMD_LAUNCH(ctx, "lj_idial_half", (k_lj_half<false, true>), nb, LJ_THREADS, 0, s->x,
                 s->type, row, nl->num, nl->entries, nl->maxneighs, s->n_local, tab, s->f,
                 s->f32);
    return EMD_OK;
  }
  const int K = ctx->lj_lanes[precision == EMD_MIXED ? 1 : 0];
  return launch_full_lanes<true>(ctx, s, nl, lj, p->intensity[0][0], precision, tab, store,
                                 K ? K : default_lanes(precision));
}

static emd_status lj_force(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                           const emd_lj_params *p, int precision, int half, int store) {
  if (!ctx || !s || !nl || !p) return EMD_ERR_ARG;
  EMD_ARG(ctx, p->ntypes >= 1 && p->ntypes <= LJ_MAXT, "ntypes out of range");
  EMD_ARG(ctx, nl->n == s->n_local, "neighbour list does not match the system");
  EMD_ARG(ctx, !(store && half), "storing forces needs a full list");
  if (s->n_local <= 0) return EMD_OK;
  if (store) {
    // kernels without a store form accumulate into zeroed local forces
    const int kr = ctx->lj_lanes[precision == EMD_MIXED ? 1 : 0];
    const int kk = kr ? kr : default_lanes(precision);
    const bool tile = (kr == 0 || kr >= 4) && tile_capacity(ctx, nl, p, 0) > 0;
    const bool pipe = ctx->lj_pipe_blocks > 0 && p->ntypes == 1 && kk >= 4;
    if (tile || pipe) {
      EMD_TRY(emd_force_zero(ctx, s, precision));
      store = 0;
    }
  }
  // an explicit request for 1 or 2 lanes per atom keeps the CSR kernel
  // (1 lane reproduces the reference's serial sum bit for bit)
  const int kreq = ctx->lj_lanes[precision == EMD_MIXED ? 1 : 0];
  const int tcap = (kreq == 0 || kreq >= 4) ? tile_capacity(ctx, nl, p, half) : 0;
  if (tcap) {
    switch (kreq ? kreq : ctx->lj_tile_lanes) {
      case 4: return launch_tile<4, 4>(ctx, s, nl, p, precision, tcap);
      case 16: return launch_tile<16, 2>(ctx, s, nl, p, precision, tcap);
      default: return launch_tile<8, 2>(ctx, s, nl, p, precision, tcap);
    }
  }
  lj_table *tab = nullptr;
  if (p->ntypes > 1 || half) EMD_TRY(upload_table(ctx, p, nullptr, &tab));
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
  if (half) {
    const int nb = emd_blocks((long)s->n_local * LJH_K, LJ_THREADS);
    if (precision == EMD_MIXED)
      EMD_LAUNCH(ctx, "lj_force_half", (k_lj_half<true, false>), nb, LJ_THREADS, 0, s->x,
                 s->type, row, nl->num, nl->entries, nl->maxneighs, s->n_local, tab, s->f,
                 s->f32);
    else
      EMD_LAUNCH(ctx, "lj_force_half", (k_lj_half<false, false>), nb, LJ_THREADS, 0, s->x,
                 s->type, row, nl->num, nl->entries, nl->maxneighs, s->n_local, tab, s->f,
                 s->f32);
    return EMD_OK;
  }
  int K = ctx->lj_lanes[precision == EMD_MIXED ? 1 : 0];
  if (K == 0) K = default_lanes(precision);
  if (ctx->lj_pipe_blocks > 0 && p->ntypes == 1 && K >= 4) {
    switch (K) {
      case 4: return launch_pipe<4, 4>(ctx, s, nl, p, precision);
      case 8: return launch_pipe<8, 5>(ctx, s, nl, p, precision);
      default:
        switch (ctx->lj_unroll) {
          case 4: return launch_pipe<16, 4>(ctx, s, nl, p, precision);
          case 6: return launch_pipe<16, 6>(ctx, s, nl, p, precision);
          default: return launch_pipe<16, 5>(ctx, s, nl, p, precision);
        }
    }
  }
  switch (K) {
    case 1: return launch_full<1, 4>(ctx, s, nl, p, precision, tab, store);
    case 2: return launch_full<2, 4>(ctx, s, nl, p, precision, tab, store);
    case 4: return launch_full<4, 4>(ctx, s, nl, p, precision, tab, store);
    case 8:
      switch (ctx->lj_unroll) {
        case 8: return launch_full<8, 8>(ctx, s, nl, p, precision, tab, store);
        case 10: return launch_full<8, 10>(ctx, s, nl, p, precision, tab, store);
        default: return launch_full<8, 5>(ctx, s, nl, p, precision, tab, store);
      }
    default:
      switch (ctx->lj_unroll) {
        case 2: return launch_full<16, 2>(ctx, s, nl, p, precision, tab, store);
        case 3: return launch_full<16, 3>(ctx, s, nl, p, precision, tab, store);
        case 4: return launch_full<16, 4>(ctx, s, nl, p, precision, tab, store);
        case 6: return launch_full<16, 6>(ctx, s, nl, p, precision, tab, store);
        default: return launch_full<16, 5>(ctx, s, nl, p, precision, tab, store);
      }
  }
}

extern "C" emd_status emd_lj_force(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                   const emd_lj_params *p, int precision, int half) {
  return lj_force(ctx, s, nl, p, precision, half, 0);
}

extern "C" emd_status emd_lj_force_store(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                         const emd_lj_params *p, int precision) {
  return lj_force(ctx, s, nl, p, precision, 0, 1);
}

extern "C" emd_status emd_lj_idial_force(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                         const emd_lj_idial_params *p, int precision,
                                         int half) {
  return lj_idial_force(ctx, s, nl, p, precision, half, 0);
}

extern "C" emd_status emd_lj_idial_force_store(emd_ctx *ctx, emd_system *s,
                                               const emd_neigh *nl,
                                               const emd_lj_idial_params *p, int precision) {
  return lj_idial_force(ctx, s, nl, p, precision, 0, 1);
}


// ForceLJNeigh Tag{Full,Half}NeighPE, force_lj_neigh_impl.h:256-343.  K lanes
// per atom take every K-th entry of the row, issuing the index loads of U
// entries and then their U gathers before the arithmetic (as k_lj_full);
// each lane sums its pair energies (the reference's per-pair expression,
// shift included) and the block adds all lanes in double-double.  fac is 1
// or 1/2, so fac * x is exact and the terms are fac times the fac-free
// expressions; with one type the shift term is formed once per lane.
constexpr int LJE_K = 16, LJE_U = 5;

__device__ __forceinline__ double lj_pair_e(double r6inv, double lj1, double lj2) {
#pragma clang fp contract(off)
  return r6inv * (0.5 * lj1 * r6inv - lj2) / 6.0;
}

template <bool ONE_TYPE>
__global__ void __launch_bounds__(LJ_THREADS)
k_lj_energy(const double *__restrict__ x, const int *__restrict__ type,
            const int *__restrict__ row, const int *__restrict__ num,
            const int *__restrict__ entries, int pitch, int n, int half,
            const lj_table *__restrict__ tab, double *__restrict__ part) {
#pragma clang fp contract(off)
  const int gid = blockIdx.x * LJ_THREADS + threadIdx.x;
  const int i = gid / LJE_K, sub = gid % LJE_K;
  double pe = 0.0;
  if (i < n) {
    const double x_i = x[3 * (size_t)i], y_i = x[3 * (size_t)i + 1], z_i = x[3 * (size_t)i + 2];
    const int type_i = type[i];
    const int *rowp = entries + (row ? (size_t)row[i] : (size_t)i * (size_t)pitch);
    const int cnt = num[i];
    double shift0 = 0.0;
    if (ONE_TYPE) {
      const double r2invc = 1.0 / tab->cutsq[0];
      shift0 = lj_pair_e(r2invc * r2invc * r2invc, tab->lj1[0], tab->lj2[0]);
    }
    for (int jj = sub; jj < cnt; jj += LJE_U * LJE_K) {
      int j[LJE_U];
      bool ok[LJE_U];
#pragma unroll
      for (int u = 0; u < LJE_U; u++) {
        ok[u] = jj + u * LJE_K < cnt;
        j[u] = rowp[ok[u] ? jj + u * LJE_K : cnt - 1];
      }
      double xj[LJE_U], yj[LJE_U], zj[LJE_U];
#pragma unroll
      for (int u = 0; u < LJE_U; u++) {
        xj[u] = x[3 * (size_t)j[u]];
        yj[u] = x[3 * (size_t)j[u] + 1];
        zj[u] = x[3 * (size_t)j[u] + 2];
      }
#pragma unroll
      for (int u = 0; u < LJE_U; u++) {
        const double dx = x_i - xj[u];
        const double dy = y_i - yj[u];
        const double dz = z_i - zj[u];
        const double rsq = dx * dx + dy * dy + dz * dz;
        const int t = ONE_TYPE ? 0 : type_i * LJ_MAXT + type[j[u]];
        const double cutsq = tab->cutsq[t], lj1 = tab->lj1[t], lj2 = tab->lj2[t];
        if (ok[u] && rsq < cutsq) {
          const double r2inv = 1.0 / rsq;
          const double r6inv = r2inv * r2inv * r2inv;
          double shift = shift0;
          if (!ONE_TYPE) {
            const double r2invc = 1.0 / cutsq;
            shift = lj_pair_e(r2invc * r2invc * r2invc, lj1, lj2);
          }
          const double fac = half ? ((j[u] < n) ? 1.0 : 0.5) : 0.5;
          pe += fac * lj_pair_e(r6inv, lj1, lj2);
          pe -= fac * shift;
        }
      }
    }
  }
  block_dd_sum(pe, 0.0, part);
}

static emd_status finish_dd(emd_ctx *ctx, double *part, int nb, double *out) {
  std::vector<double> h(2 * (size_t)nb);
  EMD_HIP(ctx, hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  EMD_TRY(emd_wait(ctx));
  // Neumaier summation of the block (hi, lo) pairs
  double sum = 0.0, c = 0.0;
  for (size_t k = 0; k < h.size(); k++) {
    const double v = h[k];
    const double t = sum + v;
    if (fabs(sum) >= fabs(v)) c += (sum - t) + v;
    else c += (v - t) + sum;
    sum = t;
  }
  *out = sum + c;
  return EMD_OK;
}

extern "C" emd_status emd_lj_energy(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                    const emd_lj_params *p, int half, double *pe) {
  if (!ctx || !s || !nl || !p || !pe) return EMD_ERR_ARG;
  EMD_ARG(ctx, p->ntypes >= 1 && p->ntypes <= LJ_MAXT, "ntypes out of range");
  const int n = s->n_local;
  if (n <= 0) { *pe = 0.0; return EMD_OK; }
  lj_table *tab;
  EMD_TRY(upload_table(ctx, p, nullptr, &tab));
  const int nb = emd_blocks((long)n * LJE_K, LJ_THREADS);
  EMD_TRY(emd_ensure(ctx, &ctx->d_tmp2, &ctx->tmp2_cap, (size_t)nb * 2 * sizeof(double)));
  double *part = (double *)ctx->d_tmp2;
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
  if (p->ntypes == 1)
    EMD_LAUNCH(ctx, "lj_energy", k_lj_energy<true>, nb, LJ_THREADS, 0, s->x, s->type, row,
               nl->num, nl->entries, nl->maxneighs, n, half, tab, part);
  else
    EMD_LAUNCH(ctx, "lj_energy", k_lj_energy<false>, nb, LJ_THREADS, 0, s->x, s->type, row,
               nl->num, nl->entries, nl->maxneighs, n, half, tab, part);
  return finish_dd(ctx, part, nb, pe);
}

// ------------------------------------------------ adjoint d/d(eps, sigma)
// For a cotangent g = dL/dF over the local atoms and gE = dL/dE:
//   dL/dtheta = sum_i sum_{j in N(i)} g_i . r_ij dfpair/dtheta + gE dE/dtheta
// with fpair = r^-2 r^-6 (48 eps s^12 r^-6 - 24 eps s^6) (full list: atom i's
// force is its own row) and the shifted pair energy counted half per row.
// Pair arithmetic in T, accumulation in FP64.
template <class T>
__global__ void __launch_bounds__(256)
k_lj_adjoint(const double *__restrict__ x, const int *__restrict__ row,
             const int *__restrict__ num, const int *__restrict__ entries, int pitch, int n,
             double eps, double sigma, double cutsq, const double *__restrict__ g,
             double *__restrict__ out) {
  __shared__ double s_red[4][4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    const double x_i = x[3 * (size_t)i], y_i = x[3 * (size_t)i + 1], z_i = x[3 * (size_t)i + 2];
    const double g0 = g[3 * (size_t)i], g1 = g[3 * (size_t)i + 1], g2 = g[3 * (size_t)i + 2];
    const T s2 = (T)(sigma * sigma), s6 = s2 * s2 * s2, s12 = s6 * s6;
    const T s5 = s6 / (T)sigma, s11 = s12 / (T)sigma;
    const T e = (T)eps;
    const T rc2inv = (T)(1.0 / cutsq), rc6 = rc2inv * rc2inv * rc2inv;
    const size_t start = row ? (size_t)row[i] : (size_t)i * (size_t)pitch;
    const int cnt = num[i];
    for (int jj = 0; jj < cnt; jj++) {
      const int j = entries[start + jj];
      const double dx = x_i - x[3 * (size_t)j];
      const double dy = y_i - x[3 * (size_t)j + 1];
      const double dz = z_i - x[3 * (size_t)j + 2];
      const double rsq = dx * dx + dy * dy + dz * dz;
      if (rsq < cutsq) {
        const T r2inv = (T)1 / (T)rsq;
        const T r6inv = r2inv * r2inv * r2inv;
        const T dfde = r6inv * ((T)48 * s12 * r6inv - (T)24 * s6) * r2inv;
        const T dfds = r6inv * ((T)576 * e * s11 * r6inv - (T)144 * e * s5) * r2inv;
        const double gdx = g0 * dx + g1 * dy + g2 * dz;
        a[0] += gdx * (double)dfde;
        a[1] += gdx * (double)dfds;
        a[2] += (double)((T)2 * ((s12 * r6inv * r6inv - s6 * r6inv) - (s12 * rc6 * rc6 - s6 * rc6)));
        a[3] += (double)((T)2 * e * (((T)12 * s11 * r6inv * r6inv - (T)6 * s5 * r6inv) -
                                     ((T)12 * s11 * rc6 * rc6 - (T)6 * s5 * rc6)));
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int o = EMD_WAVE / 2; o > 0; o >>= 1) a[k] += __shfl_xor(a[k], o, EMD_WAVE);
  if (__lane_id() == 0)
    for (int k = 0; k < 4; k++) s_red[threadIdx.x / EMD_WAVE][k] = a[k];
  __syncthreads();
  if (threadIdx.x < 4) {
    const double v = (s_red[0][threadIdx.x] + s_red[1][threadIdx.x]) +
                     (s_red[2][threadIdx.x] + s_red[3][threadIdx.x]);
    atomicAdd(&out[threadIdx.x], v);
  }
}

__global__ void k_zero4(double *p) {
  if (threadIdx.x < 4) p[threadIdx.x] = 0.0;
}

extern "C" emd_status emd_lj_adjoint(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                     double eps, double sigma, double cut, const double *dLdF,
                                     double dLdE, int precision, double *out) {
  if (!ctx || !s || !nl || !dLdF || !out) return EMD_ERR_ARG;
  EMD_ARG(ctx, eps > 0 && sigma > 0 && cut > 0, "eps, sigma and cut must be positive");
  out[0] = out[1] = 0.0;
  const int n = s->n_local;
  if (n <= 0) return EMD_OK;
  double *acc = ctx->d_dsmall;
  EMD_LAUNCH(ctx, "lj_adjoint", k_zero4, 1, 64, 0, acc);
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
  if (precision == EMD_MIXED)
    EMD_LAUNCH(ctx, "lj_adjoint", k_lj_adjoint<float>, emd_blocks(n, 256), 256, 0, s->x, row,
               nl->num, nl->entries, nl->maxneighs, n, eps, sigma, cut * cut, dLdF, acc);
  else
    EMD_LAUNCH(ctx, "lj_adjoint", k_lj_adjoint<double>, emd_blocks(n, 256), 256, 0, s->x, row,
               nl->num, nl->entries, nl->maxneighs, n, eps, sigma, cut * cut, dLdF, acc);
  EMD_HIP(ctx, hipMemcpyAsync(ctx->h_dsmall, acc, 4 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  EMD_TRY(emd_wait(ctx));
  out[0] = ctx->h_dsmall[0] + dLdE * ctx->h_dsmall[2];
  out[1] = ctx->h_dsmall[1] + dLdE * ctx->h_dsmall[3];
  return EMD_OK;
}

// ---------------------------------------------------------------- NVE
// IntegratorNVE InitialIntegrateFunctor, src/integrator_nve.cpp:65-74
template <bool MIXED>
__global__ void k_nve_initial(double *__restrict__ x, double *__restrict__ v,
                              const double *__restrict__ f, const float *__restrict__ f32,
                              const int *__restrict__ type, const double *__restrict__ mass,
                              int n, double dtf, double dtv) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dtfm = dtf / mass[type[i]];
  const size_t o = 3 * (size_t)i;
  const double f0 = MIXED ? (double)f32[o] : f[o];
  const double f1 = MIXED ? (double)f32[o + 1] : f[o + 1];
  const double f2 = MIXED ? (double)f32[o + 2] : f[o + 2];
  const double v0 = v[o] + dtfm * f0, v1 = v[o + 1] + dtfm * f1, v2 = v[o + 2] + dtfm * f2;
  v[o] = v0; v[o + 1] = v1; v[o + 2] = v2;
  x[o] += dtv * v0;
  x[o + 1] += dtv * v1;
  x[o + 2] += dtv * v2;
}

// IntegratorNVE FinalIntegrateFunctor, src/integrator_nve.cpp:104-112
template <bool MIXED>
__global__ void k_nve_final(double *__restrict__ v, const double *__restrict__ f,
                            const float *__restrict__ f32, const int *__restrict__ type,
                            const double *__restrict__ mass, int n, double dtf) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dtfm = dtf / mass[type[i]];
  const size_t o = 3 * (size_t)i;
  const double f0 = MIXED ? (double)f32[o] : f[o];
  const double f1 = MIXED ? (double)f32[o + 1] : f[o + 1];
  const double f2 = MIXED ? (double)f32[o + 2] : f[o + 2];
  v[o] += dtfm * f0;
  v[o + 1] += dtfm * f1;
  v[o + 2] += dtfm * f2;
}

// FinalIntegrateFunctor of one step followed by InitialIntegrateFunctor of
// the next (src/integrator_nve.cpp:104-112 then 65-74) in one pass over the
// atoms: the same two velocity half-kicks and drift, in the same order, with
// one read of x, v, f and one write of x, v instead of two kernels
template <bool MIXED>
__global__ void k_nve_final_initial(double *__restrict__ x, double *__restrict__ v,
                                    const double *__restrict__ f, const float *__restrict__ f32,
                                    const int *__restrict__ type,
                                    const double *__restrict__ mass, int n, double dtf,
                                    double dtv) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dtfm = dtf / mass[type[i]];
  const size_t o = 3 * (size_t)i;
  const double f0 = MIXED ? (double)f32[o] : f[o];
  const double f1 = MIXED ? (double)f32[o + 1] : f[o + 1];
  const double f2 = MIXED ? (double)f32[o + 2] : f[o + 2];
  double v0 = v[o] + dtfm * f0, v1 = v[o + 1] + dtfm * f1, v2 = v[o + 2] + dtfm * f2;
  v0 = v0 + dtfm * f0;
  v1 = v1 + dtfm * f1;
  v2 = v2 + dtfm * f2;
  v[o] = v0; v[o + 1] = v1; v[o + 2] = v2;
  x[o] += dtv * v0;
  x[o + 1] += dtv * v1;
  x[o + 2] += dtv * v2;
}

extern "C" emd_status emd_nve_final_initial(emd_ctx *ctx, emd_system *s, double dtf, double dtv,
                                            int precision) {
  if (!ctx || !s) return EMD_ERR_ARG;
  const int n = s->n_local;
  if (n <= 0) return EMD_OK;
  if (precision == EMD_MIXED)
    EMD_LAUNCH(ctx, "nve_final_initial", k_nve_final_initial<true>, emd_blocks(n, 256), 256, 0,
               s->x, s->v, s->f, s->f32, s->type, s->mass, n, dtf, dtv);
  else
    EMD_LAUNCH(ctx, "nve_final_initial", k_nve_final_initial<false>, emd_blocks(n, 256), 256, 0,
               s->x, s->v, s->f, s->f32, s->type, s->mass, n, dtf, dtv);
  return EMD_OK;
}

extern "C" emd_status emd_nve_initial(emd_ctx *ctx, emd_system *s, double dtf, double dtv,
                                      int precision) {
  if (!ctx || !s) return EMD_ERR_ARG;
  const int n = s->n_local;
  if (n <= 0) return EMD_OK;
  if (precision == EMD_MIXED)
    EMD_LAUNCH(ctx, "nve_initial", k_nve_initial<true>, emd_blocks(n, 256), 256, 0, s->x, s->v,
               s->f, s->f32, s->type, s->mass, n, dtf, dtv);
  else
    EMD_LAUNCH(ctx, "nve_initial", k_nve_initial<false>, emd_blocks(n, 256), 256, 0, s->x, s->v,
               s->f, s->f32, s->type, s->mass, n, dtf, dtv);
  return EMD_OK;
}

extern "C" emd_status emd_nve_final(emd_ctx *ctx, emd_system *s, double dtf, int precision) {
  if (!ctx || !s) return EMD_ERR_ARG;
  const int n = s->n_local;
  if (n <= 0) return EMD_OK;
  if (precision == EMD_MIXED)
    EMD_LAUNCH(ctx, "nve_final", k_nve_final<true>, emd_blocks(n, 256), 256, 0, s->v, s->f,
               s->f32, s->type, s->mass, n, dtf);
  else
    EMD_LAUNCH(ctx, "nve_final", k_nve_final<false>, emd_blocks(n, 256), 256, 0, s->v, s->f,
               s->f32, s->type, s->mass, n, dtf);
  return EMD_OK;
}

// Temperature / KinE summand, src/property_temperature.h:54-56
__global__ void __launch_bounds__(256)
k_mv2(const double *__restrict__ v, const int *__restrict__ type,
      const double *__restrict__ mass, int n, double *__restrict__ part) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double t = 0.0;
  if (i < n) {
    const size_t o = 3 * (size_t)i;
    t = (v[o] * v[o] + v[o + 1] * v[o + 1] + v[o + 2] * v[o + 2]) * mass[type[i]];
  }
  block_dd_sum(t, 0.0, part);
}

extern "C" emd_status emd_reduce_mv2(emd_ctx *ctx, const emd_system *s, double *mv2) {
  if (!ctx || !s || !mv2) return EMD_ERR_ARG;
  const int n = s->n_local;
  if (n <= 0) { *mv2 = 0.0; return EMD_OK; }
  const int nb = emd_blocks(n, 256);
  EMD_TRY(emd_ensure(ctx, &ctx->d_tmp2, &ctx->tmp2_cap, (size_t)nb * 2 * sizeof(double)));
  double *part = (double *)ctx->d_tmp2;
  EMD_LAUNCH(ctx, "reduce_mv2", k_mv2, nb, 256, 0, s->v, s->type, s->mass, n, part);
  return finish_dd(ctx, part, nb, mv2);
}
