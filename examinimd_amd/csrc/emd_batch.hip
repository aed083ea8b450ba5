// emd_batch.hip -- many independent periodic LJ replicas advanced together.
//
// Per replica the arithmetic is the reference's serial-comm LJ step
// (src/examinimd.cpp:176-294): IntegratorNVE (src/integrator_nve.cpp:65-121),
// CommSerial exchange (comm_serial.h:95-108) and periodic images, a full
// neighbour list with the inclusive rsq <= (rc+skin)^2 test
// (neighbor_csr.h:206-248) and the ForceLJNeigh full-list force with the
// strict rsq < rc^2 test (force_lj_neigh_impl.h:158-206).
//
// Layout: the atoms of all replicas form one flat array, so each kernel is
// one launch for the whole batch and the work per launch is balanced over
// atoms, not replicas.  Periodic images are not stored as ghost atoms: a
// neighbour entry is (j | code << 26) with a 2-bit image code per axis, and
// the image position is x_j -/+ L on that axis (the same single addition a
// ghost copy makes, comm_serial.h:112-197).  Neighbour rows are padded
// [atom][cap] (cap a multiple of 16 ints, one 64-byte line per 16 entries).
//
// A step is one fused kernel: force on atom i, the final half-kick of this
// step, then the next step's half-kick and drift written to the second
// position buffer (positions of step t are read by all atoms while the
// kernel runs, so the drift cannot be in place).  The kernels of one
// rebuild period (wrap, bins, rows, rebuild_every steps) are captured once
// as a HIP graph and replayed.
#include <cmath>
#include <type_traits>

#include "emd_internal.hpp"

namespace {

constexpr int B_CODE_SHIFT = 26;
constexpr unsigned B_JMASK = (1u << B_CODE_SHIFT) - 1u;
constexpr int B_MAX_BLOCK = 512;

struct rep_info {
  double L[3];
  double dtfm;  // dtf / mass (IntegratorNVE dtfm with mvv2e = 1)
  int start, n;
  int nb[3];
  int binbase;
};

enum { MODE_FORCE = 0, MODE_FINAL = 1, MODE_FUSED = 2 };

}  // namespace

struct emd_batch {
  int nrep = 0, total = 0, nbins = 0;
  double lj1 = 0, lj2 = 0, cutsq = 0, cutneighsq = 0, dt = 0, dtv = 0;
  double eps = 1, sigma = 1, cut = 2.5, skin = 0.3;
  int rebuild_every = 20, precision = EMD_FP64;
  std::vector<rep_info> h_info;
  rep_info *info = nullptr;
  int *rep = nullptr;
  double *x[2] = {nullptr, nullptr};
  int cur = 0;
  double *v = nullptr, *f = nullptr;
  float *f32 = nullptr;
  int *binof = nullptr, *bincount = nullptr, *binstart = nullptr, *bincur = nullptr;
  int *binatoms = nullptr, *num = nullptr;
  int *binrep = nullptr;   // replica of every bin
  unsigned *entries = nullptr;
  size_t entries_cap = 0;  // bytes
  int cap = 0;
  int *flags = nullptr;    // [0] largest row seen
  double *part = nullptr;  // [2 * nrep]
  double *save = nullptr;  // autotune snapshot: x, v, f (FP64 or FP32 bits)
  long step = 0;
  bool ready = false;
  int lanes = 4, block = 256;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
};

// --------------------------------------------------------------- kernels
__global__ void k_bzero_int(int *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// CommSerial exchange wrap (comm_serial.h:95-108) + bin assignment and count
__global__ void k_bwrap_bin(double *__restrict__ x, const int *__restrict__ rep,
                            const rep_info *__restrict__ info, int n, int *__restrict__ binof,
                            int *__restrict__ bincount) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const rep_info R = info[rep[i]];
  int b[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double xd = x[3 * (size_t)i + d];
    double r = xd;
    if (xd > R.L[d]) r -= R.L[d];
    if (xd < 0) r += R.L[d];
    x[3 * (size_t)i + d] = r;
    int c = (int)(r * ((double)R.nb[d] / R.L[d]));
    c = c < 0 ? 0 : (c >= R.nb[d] ? R.nb[d] - 1 : c);
    b[d] = c;
  }
  const int g = R.binbase + (b[0] * R.nb[1] + b[1]) * R.nb[2] + b[2];
  binof[i] = g;
  atomicAdd(&bincount[g], 1);
}

// one block per replica: exclusive scan of its bin counts into global atom
// slots (replica r's atoms fill slots [start_r, start_r + n_r)); only batch
// buffers are touched, so a captured graph stays valid whatever else the
// context does
__global__ void __launch_bounds__(256)
k_bscan(const rep_info *__restrict__ info, const int *__restrict__ count,
        int *__restrict__ start, int *__restrict__ cursor) {
  __shared__ int s_w[256 / EMD_WAVE];
  __shared__ int s_carry;
  const rep_info R = info[blockIdx.x];
  const int nbr = R.nb[0] * R.nb[1] * R.nb[2];
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < nbr; base += 256) {
    const int k = base + threadIdx.x;
    const int v = k < nbr ? count[R.binbase + k] : 0;
    const int inc = emd_wave_incl_scan(v);
    const int w = threadIdx.x / EMD_WAVE;
    if (__lane_id() == EMD_WAVE - 1) s_w[w] = inc;
    __syncthreads();
    int off = s_carry;
    for (int q = 0; q < w; q++) off += s_w[q];
    if (k < nbr) {
      start[R.binbase + k] = R.start + off + inc - v;
      cursor[R.binbase + k] = R.start + off + inc - v;
    }
    __syncthreads();
    if (threadIdx.x == 255) s_carry = off + inc;
    __syncthreads();
  }
}

__global__ void k_bfill(const int *__restrict__ binof, int *__restrict__ cursor,
                        int *__restrict__ binatoms, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  binatoms[atomicAdd(&cursor[binof[i]], 1)] = i;
}

// ascending atom order inside each bin (deterministic rows)
__global__ void k_bsort(const int *__restrict__ start, const int *__restrict__ count,
                        int *__restrict__ binatoms, int nbins) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbins) return;
  int *a = binatoms + start[b];
  const int c = count[b];
  for (int k = 1; k < c; k++) {
    const int v = a[k];
    int m = k - 1;
    while (m >= 0 && a[m] > v) {
      a[m + 1] = a[m];
      m--;
    }
    a[m + 1] = v;
  }
}

// image position of an entry on one axis (the ghost copy's arithmetic)
__device__ __forceinline__ double b_image(double xj, unsigned c, double L) {
  return c == 0u ? xj : xj + (c == 1u ? -L : L);
}

constexpr int BN_THREADS = 256;
constexpr int BN_MAXC = 1024;  // candidates staged per LDS chunk

// one block per bin (the plain neighbour build's scheme): the 27 stencil
// bins of the bin's replica, each with its periodic image code, are staged in
// LDS as candidates (stencil order, each bin in ascending atom order, image
// positions as the force kernel forms them); each wavefront takes atoms of
// the bin and tests 64 candidates per step, compacting hits with ballot and
// popcount, so rows come out in stencil-then-bin order.  Keeps
// rsq <= cutneighsq except the atom itself (image code 0).
__global__ void __launch_bounds__(BN_THREADS)
k_bneigh(const double *__restrict__ x, const rep_info *__restrict__ info,
         const int *__restrict__ binrep, const int *__restrict__ binstart,
         const int *__restrict__ bincount, const int *__restrict__ binatoms, double cutneighsq,
         int cap, int *__restrict__ num, unsigned *__restrict__ entries,
         int *__restrict__ flags) {
#pragma clang fp contract(off)
  __shared__ double s_x[BN_MAXC], s_y[BN_MAXC], s_z[BN_MAXC];
  __shared__ unsigned s_e[BN_MAXC];
  __shared__ int s_start[27], s_pre[28];
  __shared__ unsigned s_code[27];
  __shared__ int s_max[BN_THREADS / EMD_WAVE];
  const int g = emd_xcd_tile(blockIdx.x, gridDim.x);
  const int c_cnt = bincount[g];
  if (c_cnt == 0) return;  // uniform over the block
  const rep_info R = info[binrep[g]];
  const int local = g - R.binbase;
  const int bz = local % R.nb[2], by = (local / R.nb[2]) % R.nb[1],
            bx = local / (R.nb[2] * R.nb[1]);
  const int tid = threadIdx.x, wave = tid / EMD_WAVE, lane = tid % EMD_WAVE;
  if (tid < 27) {
    int cx = bx + tid / 9 - 1, cy = by + (tid / 3) % 3 - 1, cz = bz + tid % 3 - 1;
    unsigned kx = 0, ky = 0, kz = 0;
    if (cx < 0) { cx += R.nb[0]; kx = 1; }
    if (cx >= R.nb[0]) { cx -= R.nb[0]; kx = 2; }
    if (cy < 0) { cy += R.nb[1]; ky = 1; }
    if (cy >= R.nb[1]) { cy -= R.nb[1]; ky = 2; }
    if (cz < 0) { cz += R.nb[2]; kz = 1; }
    if (cz >= R.nb[2]) { cz -= R.nb[2]; kz = 2; }
    const int gb = R.binbase + (cx * R.nb[1] + cy) * R.nb[2] + cz;
    s_start[tid] = binstart[gb];
    s_pre[tid + 1] = bincount[gb];
    s_code[tid] = kx | (ky << 2) | (kz << 4);
  }
  __syncthreads();
  if (tid == 0) {
    s_pre[0] = 0;
    for (int k = 1; k <= 27; k++) s_pre[k] += s_pre[k - 1];
  }
  __syncthreads();
  const int T = s_pre[27];
  const int c_start = binstart[g];
  const unsigned long long lt = emd_lanemask_lt();
  int mymax = 0;
  for (int cs = 0; cs < T; cs += BN_MAXC) {
    const int cn = (T - cs) < BN_MAXC ? (T - cs) : BN_MAXC;
    for (int c = tid; c < cn; c += BN_THREADS) {
      const int cc = cs + c;
      const int k = emd_stencil_slot(s_pre, cc);
      const int j = binatoms[s_start[k] + (cc - s_pre[k])];
      const unsigned code = s_code[k];
      s_x[c] = b_image(x[3 * (size_t)j], code & 3u, R.L[0]);
      s_y[c] = b_image(x[3 * (size_t)j + 1], (code >> 2) & 3u, R.L[1]);
      s_z[c] = b_image(x[3 * (size_t)j + 2], (code >> 4) & 3u, R.L[2]);
      s_e[c] = (unsigned)j | (code << B_CODE_SHIFT);
    }
    __syncthreads();
    const bool last = cs + BN_MAXC >= T;
    for (int bi = wave; bi < c_cnt; bi += BN_THREADS / EMD_WAVE) {
      const int i = binatoms[c_start + bi];
      const double xi = x[3 * (size_t)i], yi = x[3 * (size_t)i + 1], zi = x[3 * (size_t)i + 2];
      unsigned *row = entries + (size_t)i * cap;
      int run = cs > 0 ? num[i] : 0;
      for (int c0 = 0; c0 < cn; c0 += EMD_WAVE) {
        const int c = c0 + lane;
        bool keep = false;
        unsigned e = 0u;
        if (c < cn) {
          e = s_e[c];
          const double dx = xi - s_x[c], dy = yi - s_y[c], dz = zi - s_z[c];
          const double rsq = dx * dx + dy * dy + dz * dz;
          keep = rsq <= cutneighsq && !((int)(e & B_JMASK) == i && (e >> B_CODE_SHIFT) == 0u);
        }
        const unsigned long long m = __ballot(keep);
        if (keep) {
          const int pos = run + __popcll(m & lt);
          if (pos < cap) row[pos] = e;
        }
        run += __popcll(m);
      }
      if (lane == 0) num[i] = last ? (run < cap ? run : cap) : run;
      if (last) mymax = run > mymax ? run : mymax;
    }
    __syncthreads();
  }
  // block max of the row lengths, one atomic per block
  if (lane == 0) s_max[wave] = mymax;
  __syncthreads();
  if (tid == 0) {
    int mm = 0;
    for (int w = 0; w < BN_THREADS / EMD_WAVE; w++) mm = s_max[w] > mm ? s_max[w] : mm;
    atomicMax(&flags[0], mm);
  }
}

// Force on atom i from its row (K lanes per atom, U gathers per round) and,
// per MODE, the integrator updates of atom i:
//   MODE_FORCE  f only (set-up force)
//   MODE_FINAL  f, then v += dtfm f                     (last step of a run)
//   MODE_FUSED  f, v += dtfm f, then the next step's v += dtfm f and
//               xn = x + dtv v                          (every other step)
template <int K, int U, bool MIXED, int MODE>
__global__ void __launch_bounds__(B_MAX_BLOCK)
k_bstep(const double *__restrict__ x, double *__restrict__ xn, double *__restrict__ v,
        double *__restrict__ f, float *__restrict__ f32, const int *__restrict__ rep,
        const rep_info *__restrict__ info, const int *__restrict__ num,
        const unsigned *__restrict__ entries, int cap, int n, double lj1, double lj2,
        double cutsq, double dtv) {
#pragma clang fp contract(off)
  typedef typename std::conditional<MIXED, float, double>::type T;
  const int tile = emd_xcd_tile(blockIdx.x, gridDim.x);
  const int gid = tile * blockDim.x + threadIdx.x;
  const int i = gid / K, sub = gid % K;
  if (i >= n) return;
  const int r = rep[i];
  const double Lx = info[r].L[0], Ly = info[r].L[1], Lz = info[r].L[2];
  const double x_i = x[3 * (size_t)i], y_i = x[3 * (size_t)i + 1], z_i = x[3 * (size_t)i + 2];
  const unsigned *rowp = entries + (size_t)i * cap;
  const int cnt = num[i];
  T fxi = 0, fyi = 0, fzi = 0;
  auto pair = [&](bool ok, unsigned e, double xj, double yj, double zj) {
    const unsigned c = e >> B_CODE_SHIFT;
    const double dx = x_i - b_image(xj, c & 3u, Lx);
    const double dy = y_i - b_image(yj, (c >> 2) & 3u, Ly);
    const double dz = z_i - b_image(zj, (c >> 4) & 3u, Lz);
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
  // U entries per lane per memory round (index loads, then gathers, then
  // arithmetic); slots past the row end re-read its last entry, masked
  for (int jj = sub; jj < cnt; jj += U * K) {
    unsigned e[U];
    bool ok[U];
    double xj[U], yj[U], zj[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      ok[u] = jj + u * K < cnt;
      e[u] = __builtin_nontemporal_load(rowp + (ok[u] ? jj + u * K : cnt - 1));
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const size_t j = e[u] & B_JMASK;
      xj[u] = x[3 * j];
      yj[u] = x[3 * j + 1];
      zj[u] = x[3 * j + 2];
    }
#pragma unroll
    for (int u = 0; u < U; u++) pair(ok[u], e[u], xj[u], yj[u], zj[u]);
  }
  if (K > 1) {
    fxi = group_sum<K>(fxi);
    fyi = group_sum<K>(fyi);
    fzi = group_sum<K>(fzi);
  }
  if (sub != 0) return;
  const size_t o = 3 * (size_t)i;
  double f0, f1, f2;
  if (MIXED) {
    f32[o] = (float)fxi;
    f32[o + 1] = (float)fyi;
    f32[o + 2] = (float)fzi;
    f0 = (double)(float)fxi;
    f1 = (double)(float)fyi;
    f2 = (double)(float)fzi;
  } else {
    f[o] = fxi;
    f[o + 1] = fyi;
    f[o + 2] = fzi;
    f0 = fxi;
    f1 = fyi;
    f2 = fzi;
  }
  if (MODE == MODE_FORCE) return;
  const double dtfm = info[r].dtfm;
  double v0 = v[o] + dtfm * f0, v1 = v[o + 1] + dtfm * f1, v2 = v[o + 2] + dtfm * f2;
  if (MODE == MODE_FUSED) {
    v0 = v0 + dtfm * f0;
    v1 = v1 + dtfm * f1;
    v2 = v2 + dtfm * f2;
    xn[o] = x_i + dtv * v0;
    xn[o + 1] = y_i + dtv * v1;
    xn[o + 2] = z_i + dtv * v2;
  }
  v[o] = v0;
  v[o + 1] = v1;
  v[o + 2] = v2;
}

// IntegratorNVE initial_integrate (integrator_nve.cpp:65-83), in place
template <bool MIXED>
__global__ void k_binit(double *__restrict__ x, double *__restrict__ v,
                        const double *__restrict__ f, const float *__restrict__ f32,
                        const int *__restrict__ rep, const rep_info *__restrict__ info, int n,
                        double dtv) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dtfm = info[rep[i]].dtfm;
  const size_t o = 3 * (size_t)i;
  const double f0 = MIXED ? (double)f32[o] : f[o];
  const double f1 = MIXED ? (double)f32[o + 1] : f[o + 1];
  const double f2 = MIXED ? (double)f32[o + 2] : f[o + 2];
  const double v0 = v[o] + dtfm * f0, v1 = v[o + 1] + dtfm * f1, v2 = v[o + 2] + dtfm * f2;
  v[o] = v0;
  v[o + 1] = v1;
  v[o + 2] = v2;
  x[o] += dtv * v0;
  x[o + 1] += dtv * v1;
  x[o + 2] += dtv * v2;
}

// one block per replica: sum m v^2 (Temperature/KinE summand,
// property_temperature.h:54-56) and the shifted pair energy
// (force_lj_neigh_impl.h:256-300), both compensated
__global__ void __launch_bounds__(256)
k_bthermo(const double *__restrict__ x, const double *__restrict__ v,
          const rep_info *__restrict__ info, const int *__restrict__ num,
          const unsigned *__restrict__ entries, int cap, const double *__restrict__ mass,
          double lj1, double lj2, double cutsq, int what, double *__restrict__ part) {
#pragma clang fp contract(off)
  const rep_info R = info[blockIdx.x];
  double hi = 0.0, lo = 0.0;
  for (int k = threadIdx.x; k < R.n; k += blockDim.x) {
    const int i = R.start + k;
    const size_t o = 3 * (size_t)i;
    double t = 0.0;
    if (what == 0) {
      t = (v[o] * v[o] + v[o + 1] * v[o + 1] + v[o + 2] * v[o + 2]) * mass[blockIdx.x];
    } else {
      const double r2invc = 1.0 / cutsq;
      const double r6invc = r2invc * r2invc * r2invc;
      const double eshift = 0.5 * r6invc * (0.5 * lj1 * r6invc - lj2) / 6.0;
      for (int jj = 0; jj < num[i]; jj++) {
        const unsigned e = entries[(size_t)i * cap + jj];
        const size_t j = e & B_JMASK;
        const unsigned c = e >> B_CODE_SHIFT;
        const double dx = x[o] - b_image(x[3 * j], c & 3u, R.L[0]);
        const double dy = x[o + 1] - b_image(x[3 * j + 1], (c >> 2) & 3u, R.L[1]);
        const double dz = x[o + 2] - b_image(x[3 * j + 2], (c >> 4) & 3u, R.L[2]);
        const double rsq = dx * dx + dy * dy + dz * dz;
        if (rsq < cutsq) {
          const double r2inv = 1.0 / rsq;
          const double r6inv = r2inv * r2inv * r2inv;
          double s, e2;
          two_sum(t, 0.5 * r6inv * (0.5 * lj1 * r6inv - lj2) / 6.0, s, e2);
          t = s;
          lo += e2;
          two_sum(t, -eshift, s, e2);
          t = s;
          lo += e2;
        }
      }
    }
    double s, e;
    two_sum(hi, t, s, e);
    hi = s;
    lo += e;
  }
  block_dd_sum(hi, lo, part);
}

// -------------------------------------------------------------- host side
namespace {

template <int K, int U, bool MIXED>
emd_status launch_step_k(emd_ctx *ctx, emd_batch *b, int mode) {
  const int grid = emd_blocks((long)b->total * K, b->block);
  double *xc = b->x[b->cur], *xn = b->x[b->cur ^ 1];
#define B_ARGS xc, xn, b->v, b->f, b->f32, b->rep, b->info, b->num, b->entries, b->cap, b->total, \
               b->lj1, b->lj2, b->cutsq, b->dtv
  if (mode == MODE_FORCE)
    EMD_LAUNCH(ctx, "batch_force", (k_bstep<K, U, MIXED, MODE_FORCE>), grid, b->block, 0, B_ARGS);
  else if (mode == MODE_FINAL)
    EMD_LAUNCH(ctx, "batch_step", (k_bstep<K, U, MIXED, MODE_FINAL>), grid, b->block, 0, B_ARGS);
  else
    EMD_LAUNCH(ctx, "batch_step", (k_bstep<K, U, MIXED, MODE_FUSED>), grid, b->block, 0, B_ARGS);
#undef B_ARGS
  return EMD_OK;
}

template <bool MIXED>
emd_status launch_step_p(emd_ctx *ctx, emd_batch *b, int mode) {
  switch (b->lanes) {
    case 1: return launch_step_k<1, 4, MIXED>(ctx, b, mode);
    case 2: return launch_step_k<2, 4, MIXED>(ctx, b, mode);
    case 4: return launch_step_k<4, 4, MIXED>(ctx, b, mode);
    case 8: return launch_step_k<8, 5, MIXED>(ctx, b, mode);
    default: return launch_step_k<16, 5, MIXED>(ctx, b, mode);
  }
}

emd_status step(emd_ctx *ctx, emd_batch *b, int mode) {
  EMD_TRY(b->precision == EMD_MIXED ? launch_step_p<true>(ctx, b, mode)
                                    : launch_step_p<false>(ctx, b, mode));
  if (mode == MODE_FUSED) b->cur ^= 1;
  return EMD_OK;
}

emd_status init_half(emd_ctx *ctx, emd_batch *b) {
  const int n = b->total;
  if (b->precision == EMD_MIXED)
    EMD_LAUNCH(ctx, "batch_init", k_binit<true>, emd_blocks(n, 256), 256, 0, b->x[b->cur], b->v,
               b->f, b->f32, b->rep, b->info, n, b->dtv);
  else
    EMD_LAUNCH(ctx, "batch_init", k_binit<false>, emd_blocks(n, 256), 256, 0, b->x[b->cur], b->v,
               b->f, b->f32, b->rep, b->info, n, b->dtv);
  return EMD_OK;
}

emd_status rebuild(emd_ctx *ctx, emd_batch *b) {
  const int n = b->total, nb = b->nbins;
  double *x = b->x[b->cur];
  EMD_LAUNCH(ctx, "batch_bin_zero", k_bzero_int, emd_blocks(nb, 256), 256, 0, b->bincount, nb);
  EMD_LAUNCH(ctx, "batch_bin", k_bwrap_bin, emd_blocks(n, 256), 256, 0, x, b->rep, b->info, n,
             b->binof, b->bincount);
  EMD_LAUNCH(ctx, "batch_bin_scan", k_bscan, b->nrep, 256, 0, b->info, b->bincount, b->binstart,
             b->bincur);
  EMD_LAUNCH(ctx, "batch_bin_fill", k_bfill, emd_blocks(n, 256), 256, 0, b->binof, b->bincur,
             b->binatoms, n);
  EMD_LAUNCH(ctx, "batch_bin_sort", k_bsort, emd_blocks(nb, 256), 256, 0, b->binstart,
             b->bincount, b->binatoms, nb);
  EMD_LAUNCH(ctx, "batch_neigh", k_bneigh, nb, BN_THREADS, 0, x, b->info, b->binrep,
             b->binstart, b->bincount, b->binatoms, b->cutneighsq, b->cap, b->num, b->entries,
             b->flags);
  return EMD_OK;
}

void drop_graphs(emd_batch *b) {
  for (int p = 0; p < 2; p++)
    if (b->graph[p]) {
      hipGraphExecDestroy(b->graph[p]);
      b->graph[p] = nullptr;
    }
}

emd_status read_max_row(emd_ctx *ctx, emd_batch *b, int *m) {
  EMD_HIP(ctx, hipMemcpyAsync(ctx->h_small, b->flags, sizeof(int), hipMemcpyDeviceToHost,
                              ctx->stream));
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *m = ctx->h_small[0];
  return EMD_OK;
}

emd_status ensure_rows(emd_ctx *ctx, emd_batch *b, int cap);

// the row capacity a longest row of m needs, with headroom for density
// fluctuations until the next rebuilds
int row_capacity(int m) { return ((m + m / 4 + 16) + 15) / 16 * 16; }

// A rebuild whose longest row is checked at once (one host wait): a row
// that outgrew the capacity grows the rows and the rebuild runs again, as
// Neighbor2D grows maxneighs and refills (src/neighbor_types/
// neighbor_2d.h:304-330).  The rebuild only wraps positions, which is
// idempotent, so a repeated rebuild gives the list of the same positions.
emd_status rebuild_checked(emd_ctx *ctx, emd_batch *b) {
  for (int attempt = 0; attempt < 8; attempt++) {
    EMD_LAUNCH(ctx, "batch_flags_zero", k_bzero_int, 1, 64, 0, b->flags, 16);
    EMD_TRY(rebuild(ctx, b));
    int m = 0;
    EMD_TRY(read_max_row(ctx, b, &m));
    if (m <= b->cap) return EMD_OK;
    EMD_TRY(ensure_rows(ctx, b, row_capacity(m)));
  }
  return emd_set_error(ctx, EMD_ERR_OVERFLOW, "batch rebuild: rows keep outgrowing capacity %d",
                       b->cap);
}

// capture the rebuild_every fused steps of one rebuild period (the period's
// rebuild runs before it, checked) starting from position buffer `parity`
emd_status capture_period(emd_ctx *ctx, emd_batch *b, int parity) {
  const bool prof = ctx->profile;
  const int saved = b->cur;
  ctx->profile = false;
  b->cur = parity;
  hipGraph_t g = nullptr;
  EMD_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  emd_status st = EMD_OK;
  for (int k = 0; st == EMD_OK && k < b->rebuild_every; k++) st = step(ctx, b, MODE_FUSED);
  hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  ctx->profile = prof;
  b->cur = saved;
  if (st != EMD_OK) {
    if (g) hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess)
    return emd_set_error(ctx, EMD_ERR_HIP, "batch graph capture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&b->graph[parity], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->graph[parity] = nullptr;
    return emd_set_error(ctx, EMD_ERR_HIP, "batch graph instantiate failed: %s",
                         hipGetErrorString(e));
  }
  return EMD_OK;
}

emd_status ensure_rows(emd_ctx *ctx, emd_batch *b, int cap) {
  const size_t bytes = (size_t)b->total * cap * sizeof(unsigned);
  if (bytes > b->entries_cap) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (b->entries) EMD_HIP(ctx, hipFree(b->entries));
    b->entries = nullptr;
    b->entries_cap = 0;
    if (hipMalloc(&b->entries, bytes) != hipSuccess)
      return emd_set_error(ctx, EMD_ERR_ALLOC, "batch rows: cannot allocate %zu bytes", bytes);
    b->entries_cap = bytes;
  }
  if (cap != b->cap) drop_graphs(b);
  b->cap = cap;
  return EMD_OK;
}

template <class T>
emd_status dalloc(emd_ctx *ctx, T **p, size_t count) {
  if (hipMalloc((void **)p, count * sizeof(T) + 64) != hipSuccess)
    return emd_set_error(ctx, EMD_ERR_ALLOC, "batch: cannot allocate %zu bytes",
                         count * sizeof(T));
  return EMD_OK;
}

}  // namespace

extern "C" emd_status emd_batch_create(emd_ctx *ctx, const emd_batch_desc *d, emd_batch **out) {
  if (!ctx || !d || !out) return EMD_ERR_ARG;
  *out = nullptr;
  EMD_ARG(ctx, d->nrep >= 1 && d->natoms && d->box && d->mass, "nrep/natoms/box/mass");
  EMD_ARG(ctx, d->precision == EMD_FP64 || d->precision == EMD_MIXED, "precision");
  EMD_ARG(ctx, d->rebuild_every >= 1 && d->cut > 0 && d->skin >= 0 && d->dt > 0,
          "rebuild_every/cut/skin/dt");
  const double cn = d->cut + d->skin;
  emd_batch *b = new emd_batch();
  b->nrep = d->nrep;
  b->eps = d->eps;
  b->sigma = d->sigma;
  b->cut = d->cut;
  b->skin = d->skin;
  b->lj1 = 48.0 * d->eps * pow(d->sigma, 12.0);
  b->lj2 = 24.0 * d->eps * pow(d->sigma, 6.0);
  b->cutsq = d->cut * d->cut;
  b->cutneighsq = cn * cn;
  b->dt = d->dt;
  b->dtv = d->dt;
  b->rebuild_every = d->rebuild_every;
  b->precision = d->precision;
  b->h_info.resize(d->nrep);
  long total = 0, nbins = 0;
  for (int r = 0; r < d->nrep; r++) {
    rep_info &R = b->h_info[r];
    if (d->natoms[r] < 1 || d->mass[r] <= 0) {
      delete b;
      return emd_set_error(ctx, EMD_ERR_ARG, "emd_batch_create: replica %d: natoms/mass", r);
    }
    R.start = (int)total;
    R.n = d->natoms[r];
    R.dtfm = (0.5 * d->dt / 1.0) / d->mass[r];
    R.binbase = (int)nbins;
    long nbr = 1;
    for (int a = 0; a < 3; a++) {
      const double L = d->box[3 * r + a];
      if (!(L >= cn)) {
        delete b;
        return emd_set_error(ctx, EMD_ERR_ARG,
                             "emd_batch_create: replica %d box %g below cut+skin %g", r, L, cn);
      }
      R.L[a] = L;
      // bins no narrower than the neighbour cutoff, with a relative margin
      // far above rounding so a stencil of +-1 bin always suffices
      int nb = (int)floor(L / (cn * (1.0 + 1e-9)));
      if (nb < 1) nb = 1;
      if (nb > 64) nb = 64;
      R.nb[a] = nb;
      nbr *= nb;
    }
    total += R.n;
    nbins += nbr;
  }
  if (total >= (1l << B_CODE_SHIFT) || nbins >= (1l << 30)) {
    delete b;
    return emd_set_error(ctx, EMD_ERR_ARG, "emd_batch_create: %ld atoms exceed 2^26", total);
  }
  b->total = (int)total;
  b->nbins = (int)nbins;
  std::vector<int> rep(total);
  for (int r = 0; r < d->nrep; r++)
    for (int k = 0; k < b->h_info[r].n; k++) rep[b->h_info[r].start + k] = r;
  std::vector<double> mass(d->mass, d->mass + d->nrep);
  emd_status st = EMD_OK;
  const size_t n3 = 3 * (size_t)total;
  if (st == EMD_OK) st = dalloc(ctx, &b->info, d->nrep);
  if (st == EMD_OK) st = dalloc(ctx, &b->rep, total);
  if (st == EMD_OK) st = dalloc(ctx, &b->x[0], n3);
  if (st == EMD_OK) st = dalloc(ctx, &b->x[1], n3);
  if (st == EMD_OK) st = dalloc(ctx, &b->v, n3);
  if (st == EMD_OK) st = dalloc(ctx, &b->f, n3);
  if (st == EMD_OK) st = dalloc(ctx, &b->f32, n3);
  if (st == EMD_OK) st = dalloc(ctx, &b->binof, total);
  if (st == EMD_OK) st = dalloc(ctx, &b->bincount, nbins);
  if (st == EMD_OK) st = dalloc(ctx, &b->binstart, nbins + 1);
  if (st == EMD_OK) st = dalloc(ctx, &b->bincur, nbins);
  if (st == EMD_OK) st = dalloc(ctx, &b->binatoms, total);
  if (st == EMD_OK) st = dalloc(ctx, &b->binrep, nbins);
  if (st == EMD_OK) st = dalloc(ctx, &b->num, total);
  if (st == EMD_OK) st = dalloc(ctx, &b->flags, 16);
  if (st == EMD_OK) st = dalloc(ctx, &b->part, 2 * (size_t)d->nrep + d->nrep);
  if (st == EMD_OK) st = dalloc(ctx, &b->save, 3 * n3);
  if (st != EMD_OK) {
    emd_batch_destroy(ctx, b);
    return st;
  }
  EMD_HIP(ctx, hipMemcpy(b->info, b->h_info.data(), d->nrep * sizeof(rep_info),
                         hipMemcpyHostToDevice));
  EMD_HIP(ctx, hipMemcpy(b->rep, rep.data(), total * sizeof(int), hipMemcpyHostToDevice));
  std::vector<int> binrep(nbins);
  for (int r = 0; r < d->nrep; r++) {
    const rep_info &R = b->h_info[r];
    const int nbr = R.nb[0] * R.nb[1] * R.nb[2];
    for (int k = 0; k < nbr; k++) binrep[R.binbase + k] = r;
  }
  EMD_HIP(ctx, hipMemcpy(b->binrep, binrep.data(), nbins * sizeof(int), hipMemcpyHostToDevice));
  // masses live after the 2*nrep reduction partials
  EMD_HIP(ctx, hipMemcpy(b->part + 2 * d->nrep, mass.data(), d->nrep * sizeof(double),
                         hipMemcpyHostToDevice));
  EMD_HIP(ctx, hipMemset(b->f, 0, n3 * sizeof(double)));
  EMD_HIP(ctx, hipMemset(b->f32, 0, n3 * sizeof(float)));
  *out = b;
  return EMD_OK;
}

extern "C" emd_status emd_batch_destroy(emd_ctx *ctx, emd_batch *b) {
  if (!ctx || !b) return EMD_ERR_ARG;
  hipStreamSynchronize(ctx->stream);
  drop_graphs(b);
  void *bufs[] = {b->info, b->rep, b->x[0], b->x[1], b->v, b->f, b->f32, b->binof, b->bincount,
                  b->binstart, b->bincur, b->binatoms, b->num, b->entries, b->flags, b->part,
                  b->save, b->binrep};
  for (void *p : bufs)
    if (p) hipFree(p);
  delete b;
  return EMD_OK;
}

extern "C" emd_status emd_batch_upload(emd_ctx *ctx, emd_batch *b, const double *x,
                                       const double *v) {
  if (!ctx || !b || !x || !v) return EMD_ERR_ARG;
  const size_t bytes = 3 * (size_t)b->total * sizeof(double);
  b->cur = 0;
  EMD_HIP(ctx, hipMemcpyAsync(b->x[0], x, bytes, hipMemcpyHostToDevice, ctx->stream));
  EMD_HIP(ctx, hipMemcpyAsync(b->v, v, bytes, hipMemcpyHostToDevice, ctx->stream));
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  b->step = 0;
  b->ready = false;
  return EMD_OK;
}

extern "C" emd_status emd_batch_download(emd_ctx *ctx, emd_batch *b, double *x, double *v,
                                         double *f) {
  if (!ctx || !b) return EMD_ERR_ARG;
  const size_t n3 = 3 * (size_t)b->total;
  if (x)
    EMD_HIP(ctx, hipMemcpyAsync(x, b->x[b->cur], n3 * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
  if (v)
    EMD_HIP(ctx, hipMemcpyAsync(v, b->v, n3 * sizeof(double), hipMemcpyDeviceToHost,
                                ctx->stream));
  std::vector<float> f32;
  if (f) {
    if (b->precision == EMD_MIXED) {
      f32.resize(n3);
      EMD_HIP(ctx, hipMemcpyAsync(f32.data(), b->f32, n3 * sizeof(float), hipMemcpyDeviceToHost,
                                  ctx->stream));
    } else {
      EMD_HIP(ctx, hipMemcpyAsync(f, b->f, n3 * sizeof(double), hipMemcpyDeviceToHost,
                                  ctx->stream));
    }
  }
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (f && b->precision == EMD_MIXED)
    for (size_t k = 0; k < n3; k++) f[k] = (double)f32[k];
  return EMD_OK;
}

extern "C" emd_status emd_batch_setup(emd_ctx *ctx, emd_batch *b) {
  if (!ctx || !b) return EMD_ERR_ARG;
  EMD_LAUNCH(ctx, "batch_flags_zero", k_bzero_int, 1, 64, 0, b->flags, 16);
  if (b->cap == 0) EMD_TRY(ensure_rows(ctx, b, 16));
  for (int attempt = 0; attempt < 4; attempt++) {
    EMD_TRY(rebuild(ctx, b));
    int m = 0;
    EMD_TRY(read_max_row(ctx, b, &m));
    const int want = row_capacity(m);
    if (want <= b->cap) break;
    // positions were wrapped in place by the rebuild; wrapping is idempotent
    EMD_TRY(ensure_rows(ctx, b, want));
    EMD_LAUNCH(ctx, "batch_flags_zero", k_bzero_int, 1, 64, 0, b->flags, 16);
  }
  int m = 0;
  EMD_TRY(read_max_row(ctx, b, &m));
  if (m > b->cap)
    return emd_set_error(ctx, EMD_ERR_OVERFLOW, "emd_batch_setup: row of %d > capacity %d", m,
                         b->cap);
  EMD_TRY(step(ctx, b, MODE_FORCE));
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  b->ready = true;
  return EMD_OK;
}

extern "C" emd_status emd_batch_run(emd_ctx *ctx, emd_batch *b, int nsteps, int use_graph) {
  if (!ctx || !b) return EMD_ERR_ARG;
  if (!b->ready) return emd_set_error(ctx, EMD_ERR_STATE, "emd_batch_run before emd_batch_setup");
  if (nsteps <= 0) return EMD_OK;
  const int R = b->rebuild_every;
  EMD_TRY(init_half(ctx, b));
  int k = 1;
  while (k <= nsteps) {
    const long s = b->step + 1;
    if (use_graph && s % R == 0 && k + R - 1 < nsteps) {
      EMD_TRY(rebuild_checked(ctx, b));
      const int p = b->cur;
      if (!b->graph[p]) EMD_TRY(capture_period(ctx, b, p));
      EMD_HIP(ctx, hipGraphLaunch(b->graph[p], ctx->stream));
      b->cur = p ^ (R & 1);
      b->step += R;
      k += R;
      continue;
    }
    if (s % R == 0) EMD_TRY(rebuild_checked(ctx, b));
    EMD_TRY(step(ctx, b, k == nsteps ? MODE_FINAL : MODE_FUSED));
    b->step = s;
    k++;
  }
  int m = 0;
  EMD_TRY(read_max_row(ctx, b, &m));
  if (m > b->cap)
    return emd_set_error(ctx, EMD_ERR_OVERFLOW,
                         "emd_batch_run: a neighbour row reached %d > capacity %d", m, b->cap);
  return EMD_OK;
}

extern "C" emd_status emd_batch_thermo(emd_ctx *ctx, emd_batch *b, double *ke, double *pe) {
  if (!ctx || !b) return EMD_ERR_ARG;
  if (!b->ready) return emd_set_error(ctx, EMD_ERR_STATE, "emd_batch_thermo before setup");
  std::vector<double> h(2 * (size_t)b->nrep);
  for (int what = 0; what < 2; what++) {
    double *out = what == 0 ? ke : pe;
    if (!out) continue;
    EMD_LAUNCH(ctx, "batch_thermo", k_bthermo, b->nrep, 256, 0, b->x[b->cur], b->v, b->info,
               b->num, b->entries, b->cap, b->part + 2 * b->nrep, b->lj1, b->lj2, b->cutsq,
               what, b->part);
    EMD_HIP(ctx, hipMemcpyAsync(h.data(), b->part, h.size() * sizeof(double),
                                hipMemcpyDeviceToHost, ctx->stream));
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int r = 0; r < b->nrep; r++)
      out[r] = (what == 0 ? 0.5 : 1.0) * (h[2 * r] + h[2 * r + 1]);
  }
  return EMD_OK;
}

extern "C" emd_status emd_batch_set_tile(emd_ctx *ctx, emd_batch *b, int lanes, int block) {
  if (!ctx || !b) return EMD_ERR_ARG;
  EMD_ARG(ctx, lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16, "lanes");
  EMD_ARG(ctx, block >= 64 && block <= B_MAX_BLOCK && block % 64 == 0, "block");
  if (lanes != b->lanes || block != b->block) drop_graphs(b);
  b->lanes = lanes;
  b->block = block;
  return EMD_OK;
}

extern "C" emd_status emd_batch_autotune(emd_ctx *ctx, emd_batch *b, int trial_steps,
                                         int *lanes_out, int *block_out, double *ms_out) {
  if (!ctx || !b) return EMD_ERR_ARG;
  if (!b->ready) return emd_set_error(ctx, EMD_ERR_STATE, "emd_batch_autotune before setup");
  if (trial_steps < 1) trial_steps = 1;
  const size_t n3 = 3 * (size_t)b->total;
  // snapshot of the state (x, v, f) so that tuning leaves no trace
  const int cur0 = b->cur;
  const long step0 = b->step;
  void *fsrc = b->precision == EMD_MIXED ? (void *)b->f32 : (void *)b->f;
  const size_t fbytes = n3 * (b->precision == EMD_MIXED ? sizeof(float) : sizeof(double));
  EMD_HIP(ctx, hipMemcpyAsync(b->save, b->x[cur0], n3 * 8, hipMemcpyDeviceToDevice, ctx->stream));
  EMD_HIP(ctx, hipMemcpyAsync(b->save + n3, b->v, n3 * 8, hipMemcpyDeviceToDevice, ctx->stream));
  EMD_HIP(ctx, hipMemcpyAsync(b->save + 2 * n3, fsrc, fbytes, hipMemcpyDeviceToDevice,
                              ctx->stream));
  const int cand_lanes[] = {1, 2, 4, 8, 16};
  const int cand_block[] = {64, 128, 256, 512};
  const int lanes0 = b->lanes, block0 = b->block;
  const bool prof = ctx->profile;
  ctx->profile = false;
  hipEvent_t e0, e1;
  EMD_HIP(ctx, hipEventCreate(&e0));
  EMD_HIP(ctx, hipEventCreate(&e1));
  double best = 1e30;
  int bl = lanes0, bb = block0;
  emd_status st = EMD_OK;
  for (int L : cand_lanes) {
    for (int B : cand_block) {
      b->lanes = L;
      b->block = B;
      // two untimed steps, then trial_steps timed fused steps
      for (int w = 0; st == EMD_OK && w < 2; w++) st = step(ctx, b, MODE_FUSED);
      if (st != EMD_OK) break;
      hipEventRecord(e0, ctx->stream);
      for (int k = 0; st == EMD_OK && k < trial_steps; k++) st = step(ctx, b, MODE_FUSED);
      hipEventRecord(e1, ctx->stream);
      hipEventSynchronize(e1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      if (st == EMD_OK && ms / trial_steps < best) {
        best = ms / trial_steps;
        bl = L;
        bb = B;
      }
      // restore the snapshot before the next candidate
      b->cur = cur0;
      hipMemcpyAsync(b->x[cur0], b->save, n3 * 8, hipMemcpyDeviceToDevice, ctx->stream);
      hipMemcpyAsync(b->v, b->save + n3, n3 * 8, hipMemcpyDeviceToDevice, ctx->stream);
      hipMemcpyAsync(fsrc, b->save + 2 * n3, fbytes, hipMemcpyDeviceToDevice, ctx->stream);
    }
    if (st != EMD_OK) break;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  ctx->profile = prof;
  b->cur = cur0;
  b->step = step0;
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (st != EMD_OK) {
    b->lanes = lanes0;
    b->block = block0;
    return st;
  }
  if (bl != lanes0 || bb != block0) drop_graphs(b);
  b->lanes = bl;
  b->block = bb;
  if (lanes_out) *lanes_out = bl;
  if (block_out) *block_out = bb;
  if (ms_out) *ms_out = best;
  return EMD_OK;
}

extern "C" emd_status emd_batch_info(emd_ctx *ctx, emd_batch *b, long *step, int *max_row,
                                     int *row_cap) {
  if (!ctx || !b) return EMD_ERR_ARG;
  if (step) *step = b->step;
  if (row_cap) *row_cap = b->cap;
  if (max_row) EMD_TRY(read_max_row(ctx, b, max_row));
  return EMD_OK;
}

extern "C" emd_status emd_batch_neigh(emd_ctx *ctx, emd_batch *b, int *num, unsigned *entries) {
  if (!ctx || !b) return EMD_ERR_ARG;
  if (!b->ready) return emd_set_error(ctx, EMD_ERR_STATE, "emd_batch_neigh before setup");
  if (num)
    EMD_HIP(ctx, hipMemcpyAsync(num, b->num, b->total * sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
  if (entries)
    EMD_HIP(ctx, hipMemcpyAsync(entries, b->entries, (size_t)b->total * b->cap * sizeof(unsigned),
                                hipMemcpyDeviceToHost, ctx->stream));
  EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return EMD_OK;
}
