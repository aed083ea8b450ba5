// emd_internal.hpp -- shared internals of libemd_hip.so (gfx950 / CDNA4).
//
// Context object, error plumbing, HIP-event kernel timing, wavefront
// primitives and the device-wide scan used by binning, halo compaction and
// neighbour-row offsets.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "emd_hip.h"

// ----------------------------------------------------------------- context
struct emd_prof_rec {
  int name;
  hipEvent_t start, stop;
};

// bytes reserved per context for the LJ per-type-pair table
constexpr size_t EMD_LJ_TABLE_BYTES = 8 * 1024;

struct emd_ctx {
  int device = 0;
  int num_cu = 256;
  hipStream_t stream = nullptr;
  char err[1024] = {0};

  // small device + pinned host scratch for scalar results
  int *d_small = nullptr;      // 64 ints
  double *d_dsmall = nullptr;  // 64 doubles
  int *h_small = nullptr;      // pinned, 64 ints
  double *h_dsmall = nullptr;  // pinned, 64 doubles

  // scan block sums
  int *d_scan = nullptr;
  size_t scan_cap = 0;
  // generic device scratch (grown on demand)
  void *d_tmp = nullptr;
  size_t tmp_cap = 0;
  void *d_tmp2 = nullptr;
  size_t tmp2_cap = 0;
  // LJ per-type-pair table: device copy, pinned staging copy of what was
  // last uploaded (compared before each use, uploaded only on change), and
  // the event that marks when the staging copy has been read
  void *d_ljtab = nullptr;
  // LJ coefficients for more than 16 types (emd_lj_params.table): device
  // copy (lj1, lj2, cutsq, rep of ntypes x ntypes each) and its host image
  double *d_ljbig = nullptr;
  size_t ljbig_cap = 0;
  std::vector<double> h_ljbig;
  bool ljbig_valid = false;
  void *h_ljtab = nullptr;
  hipEvent_t ljtab_ev = nullptr;
  bool ljtab_valid = false;
  // device address of this GPU's copy of the allocation-free energy path's
  // static scratch array (resolved once at emd_create)
  double *noalloc_scratch = nullptr;

  // host waits: an event recorded on the stream and polled, so a wait for a
  // count ends as soon as the copy lands (no blocking-sync wake-up latency)
  hipEvent_t wait_ev = nullptr;
  int spin_wait = 1;

  // kernel timing
  bool profile = false;
  std::vector<std::string> prof_names;
  std::map<std::string, int> prof_ids;
  std::vector<emd_prof_rec> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  std::vector<double> prof_ms;
  std::vector<long> prof_count;
  std::vector<std::string> prof_select;  // empty: time every launch

  // tuning
  int lj_lanes[2] = {0, 0};
  int fold_one = 1;        // small halos: fold all six phases in one launch
  int halo_one = 1;        // small systems: the six exchange_halo phases in one launch
  int bin_one = 1;         // small ranges: binning in one launch
  int lj_prio = -1;       // full-list LJ kernel: raise wave priority while a round's index
                          // loads and gathers issue (2), around the gathers only (1),
                          // never (0); -1: 2 for mixed, 0 for FP64 (as measured)
};

emd_status emd_set_error(emd_ctx *ctx, emd_status code, const char *fmt, ...);
// resolves the device address of emd_lj.hip's static scratch array on the
// context's (current) device
emd_status emd_lj_scratch_setup(emd_ctx *ctx);

#define EMD_HIP(ctx, call)                                                        \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return emd_set_error((ctx), EMD_ERR_HIP, "%s failed: %s (%s:%d)", #call,    \
                           hipGetErrorString(e_), __FILE__, __LINE__);            \
  } while (0)

#define EMD_TRY(expr)                       \
  do {                                      \
    emd_status s_ = (expr);                 \
    if (s_ != EMD_OK) return s_;            \
  } while (0)

#define EMD_ARG(ctx, cond, msg)                                                  \
  do {                                                                           \
    if (!(cond)) return emd_set_error((ctx), EMD_ERR_ARG, "%s: %s", __func__, msg); \
  } while (0)

// timing brackets (no-ops unless profiling is enabled)
void emd_prof_begin(emd_ctx *ctx, const char *name, emd_prof_rec *rec);
void emd_prof_end(emd_ctx *ctx, emd_prof_rec *rec);

// launch with an optional timing bracket and a launch-error check
#define EMD_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                    \
  do {                                                                            \
    emd_prof_rec rec_;                                                            \
    emd_prof_begin((ctx), (name), &rec_);                                         \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), (shmem), (ctx)->stream,   \
                       __VA_ARGS__);                                              \
    emd_prof_end((ctx), &rec_);                                                   \
    EMD_HIP((ctx), hipGetLastError());                                            \
  } while (0)

// grow a device buffer (contents not preserved)
emd_status emd_ensure(emd_ctx *ctx, void **p, size_t *cap, size_t bytes);
// wait until all work queued on the context's stream has completed
emd_status emd_wait(emd_ctx *ctx);
template <class T>
inline emd_status emd_ensure_t(emd_ctx *ctx, T **p, size_t *cap, size_t count) {
  return emd_ensure(ctx, reinterpret_cast<void **>(p), cap, count * sizeof(T));
}
emd_status emd_ensure_int(emd_ctx *ctx, int **p, int *cap, long count);
// grow an int buffer preserving its first `keep` entries
emd_status emd_grow_keep_int(emd_ctx *ctx, int **p, int *cap, long count, long keep);

// -------------------------------------------------------- wave primitives
#define EMD_WAVE 64

__device__ __forceinline__ int emd_lane() { return __lane_id(); }

__device__ __forceinline__ unsigned long long emd_lanemask_lt() {
  const int l = __lane_id();
  return (l == 0) ? 0ull : (~0ull >> (64 - l));
}

// inclusive prefix sum across the 64 lanes of a wavefront
__device__ __forceinline__ int emd_wave_incl_scan(int v) {
  const int l = __lane_id();
#pragma unroll
  for (int o = 1; o < EMD_WAVE; o <<= 1) {
    int t = __shfl_up(v, o, EMD_WAVE);
    if (l >= o) v += t;
  }
  return v;
}

__device__ __forceinline__ int emd_wave_sum(int v) {
#pragma unroll
  for (int o = EMD_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, EMD_WAVE);
  return v;
}

// map a physical block onto a logical tile so that each of the 8 XCDs (which
// receive blocks round-robin) works on one contiguous range of tiles and
// shares its L2 with spatially neighbouring atoms
__device__ __forceinline__ int emd_xcd_tile(int b, int nb) {
  const int nx = 8;
  const int q = nb / nx, r = nb % nx;
  const int x = b % nx, k = b / nx;
  // XCD x owns q (+1 for the first r XCDs) consecutive tiles
  const int base = x * q + (x < r ? x : r);
  const int len = q + (x < r ? 1 : 0);
  return (k < len) ? base + k : b;  // k < len always holds for b < nb
}

// --------------------------------------------------------- device-wide scan
// Exclusive scan of v(i), i in [0,n), in three launches.  The functor type
// provides   int value(int i) const   and   void emit(int i, int excl, int v) const.
// The total is written to *total_dev.
constexpr int EMD_SCAN_THREADS = 256;
constexpr int EMD_SCAN_ITEMS = 8;
constexpr int EMD_SCAN_TILE = EMD_SCAN_THREADS * EMD_SCAN_ITEMS;

template <class F>
__global__ void __launch_bounds__(EMD_SCAN_THREADS)
k_scan_reduce(F f, int n, int *block_sums) {
  __shared__ int s_w[EMD_SCAN_THREADS / EMD_WAVE];
  const int base = blockIdx.x * EMD_SCAN_TILE;
  int acc = 0;
#pragma unroll
  for (int k = 0; k < EMD_SCAN_ITEMS; k++) {
    int i = base + k * EMD_SCAN_THREADS + threadIdx.x;
    if (i < n) acc += f.value(i);
  }
  acc = emd_wave_sum(acc);
  if (__lane_id() == 0) s_w[threadIdx.x / EMD_WAVE] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < EMD_SCAN_THREADS / EMD_WAVE; w++) t += s_w[w];
    block_sums[blockIdx.x] = t;
  }
}

__global__ void k_scan_blocks(int *block_sums, int nblocks, int *total);

template <class F>
__global__ void __launch_bounds__(EMD_SCAN_THREADS)
k_scan_apply(F f, int n, const int *block_offsets) {
  __shared__ int s_w[EMD_SCAN_THREADS / EMD_WAVE];
  __shared__ int s_carry;
  const int base = blockIdx.x * EMD_SCAN_TILE;
  if (threadIdx.x == 0) s_carry = block_offsets[blockIdx.x];
  __syncthreads();
  const int wid = threadIdx.x / EMD_WAVE;
  for (int k = 0; k < EMD_SCAN_ITEMS; k++) {
    int i = base + k * EMD_SCAN_THREADS + threadIdx.x;
    int v = (i < n) ? f.value(i) : 0;
    int inc = emd_wave_incl_scan(v);
    if (__lane_id() == EMD_WAVE - 1) s_w[wid] = inc;
    __syncthreads();
    int woff = 0;
    for (int w = 0; w < wid; w++) woff += s_w[w];
    int carry = s_carry;
    if (i < n) f.emit(i, carry + woff + inc - v, v);
    __syncthreads();
    if (threadIdx.x == EMD_SCAN_THREADS - 1) s_carry = carry + woff + inc;
    __syncthreads();
  }
}

// host driver, first half: per-block sums and their exclusive scan; the
// total lands in ctx->d_small[0] (and in *total_host after a sync if asked)
template <class F>
emd_status emd_scan_count(emd_ctx *ctx, const F &f, int n, const char *name, int *total_host) {
  int nb = (n + EMD_SCAN_TILE - 1) / EMD_SCAN_TILE;
  if (nb < 1) nb = 1;
  if ((size_t)nb > ctx->scan_cap) {
    if (ctx->d_scan) {
      EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
      EMD_HIP(ctx, hipFree(ctx->d_scan));
    }
    ctx->d_scan = nullptr;
    size_t cap = (size_t)nb * 2;
    EMD_HIP(ctx, hipMalloc(&ctx->d_scan, cap * sizeof(int)));
    ctx->scan_cap = cap;
  }
  if (n > 0) EMD_LAUNCH(ctx, name, (k_scan_reduce<F>), nb, EMD_SCAN_THREADS, 0, f, n, ctx->d_scan);
  else EMD_HIP(ctx, hipMemsetAsync(ctx->d_scan, 0, sizeof(int), ctx->stream));
  EMD_LAUNCH(ctx, "scan_blocks", k_scan_blocks, 1, 1024, 0, ctx->d_scan, nb, ctx->d_small);
  if (total_host) {
    EMD_HIP(ctx, hipMemcpyAsync(ctx->h_small, ctx->d_small, sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
    EMD_TRY(emd_wait(ctx));
    *total_host = ctx->h_small[0];
  }
  return EMD_OK;
}

// second half: rescans each block and calls f.emit (same n as the count)
template <class F>
emd_status emd_scan_apply(emd_ctx *ctx, const F &f, int n, const char *name) {
  int nb = (n + EMD_SCAN_TILE - 1) / EMD_SCAN_TILE;
  if (n > 0) EMD_LAUNCH(ctx, name, (k_scan_apply<F>), nb, EMD_SCAN_THREADS, 0, f, n, ctx->d_scan);
  return EMD_OK;
}

template <class F>
emd_status emd_scan(emd_ctx *ctx, const F &f, int n, const char *name, int *total_host) {
  EMD_TRY(emd_scan_count(ctx, f, n, name, nullptr));
  EMD_TRY(emd_scan_apply(ctx, f, n, name));
  if (total_host) {
    EMD_HIP(ctx, hipMemcpyAsync(ctx->h_small, ctx->d_small, sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
    EMD_TRY(emd_wait(ctx));
    *total_host = ctx->h_small[0];
  }
  return EMD_OK;
}

// plain exclusive scan of an int array into out (out[n] = total if room_total)
struct emd_scan_array {
  const int *in;
  int *out;
  int n;
  int write_total;
  __device__ int value(int i) const { return in[i]; }
  __device__ void emit(int i, int excl, int v) const {
    out[i] = excl;
    if (write_total && i == n - 1) out[n] = excl + v;
  }
};

// ------------------------------------------------------ compensated sums
// error-free transformation a + b = s + e (Knuth TwoSum)
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
#pragma clang fp contract(off)
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// block reduction of (hi, lo) pairs (blockDim.x <= 1024); the block's
// double-double total lands in out[2*blockIdx.x + {0,1}]
__device__ inline void block_dd_sum(double hi, double lo, double *out) {
#pragma clang fp contract(off)
  __shared__ double s_hi[16], s_lo[16];
#pragma unroll
  for (int o = EMD_WAVE / 2; o > 0; o >>= 1) {
    const double oh = __shfl_xor(hi, o, EMD_WAVE), ol = __shfl_xor(lo, o, EMD_WAVE);
    double s, e;
    two_sum(hi, oh, s, e);
    hi = s;
    lo = lo + ol + e;
  }
  const int w = threadIdx.x / EMD_WAVE;
  if (__lane_id() == 0) { s_hi[w] = hi; s_lo[w] = lo; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double H = 0.0, Lo = 0.0;
    for (int k = 0; k < (int)(blockDim.x / EMD_WAVE); k++) {
      double s, e;
      two_sum(H, s_hi[k], s, e);
      H = s;
      Lo += s_lo[k] + e;
    }
    out[2 * blockIdx.x] = H;
    out[2 * blockIdx.x + 1] = Lo;
  }
}

// LJ pair force factor of the mixed path.  fpair ~ r^-14 multiplies the
// relative rounding of r^2 sevenfold, so 1/r^2 comes from the hardware FP64
// reciprocal refined by one Newton step (squares its error, far below FP32
// rounding) and r^-6 is formed in FP64; the force product and accumulation
// are FP32.  On an LJ melt (rho 0.8442, T* ~0.74, force RMS 24) this keeps
// the force RMS error near 3e-6, where rounding r^2 to FP32 first gives
// 1.05e-5.
__device__ __forceinline__ float emd_lj_fpair_mixed(double rsq, float lj1, float lj2) {
  const double r0 = __builtin_amdgcn_rcp(rsq);
  const double e = __builtin_fma(-rsq, r0, 1.0);
  const double r2 = __builtin_fma(r0, e, r0);
  const double r6 = r2 * r2 * r2;
  const float r6f = (float)r6, r2f = (float)r2;
  return (r6f * (lj1 * r6f - lj2)) * r2f;
}

// Sum over groups of K consecutive lanes (K = 1, 2, 4, 8, 16; groups aligned
// to K) with DPP lane permutations inside 16-lane rows, no LDS traffic:
// quad_perm [1,0,3,2] (xor 1), quad_perm [2,3,0,1] (xor 2), row_half_mirror
// (lane l <-> 7-l within 8, joining the two quads) and row_mirror (l <-> 15-l,
// joining the two half-rows).  Every lane of a group ends with the group sum.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_any(float v) { return dpp_f32<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ double dpp_any(double v) { return dpp_f64<CTRL>(v); }

template <int K, class T>
__device__ __forceinline__ T group_sum(T v) {
  static_assert(K == 1 || K == 2 || K == 4 || K == 8 || K == 16, "group of 1..16 lanes");
  if (K >= 2) v += dpp_any<0xB1>(v);   // quad_perm [1,0,3,2]
  if (K >= 4) v += dpp_any<0x4E>(v);   // quad_perm [2,3,0,1]
  if (K >= 8) v += dpp_any<0x141>(v);  // row_half_mirror
  if (K >= 16) v += dpp_any<0x140>(v); // row_mirror
  return v;
}

// ------------------------------------------------------------ misc helpers
inline int emd_blocks(long n, int threads) { return (int)((n + threads - 1) / threads); }

// ------------------------------------------- step-boundary pieces (device)
// CommSerial::update_force phases of a halo: pack lists and the first ghost
// index of each phase (comm_serial.h:200-213)
struct fold_phases {
  const int *pack[6];
  int num[6];
  int off[6];
};

// all six phases folded by one workgroup of NT threads, phases in reverse
// order with a workgroup barrier between them (a phase reads ghost forces
// that later phases folded into it).  Within a phase the destinations
// pack[p][k] are distinct atoms below the phase's ghost range, so no two
// lanes write one address and no lane reads what another writes.
template <class T, int NT>
__device__ __forceinline__ void emd_fold_phases_one(T *__restrict__ f, const fold_phases &ph) {
  for (int p = 5; p >= 0; p--) {
    const int n = ph.num[p], off = ph.off[p];
    const int *__restrict__ pack = ph.pack[p];
    for (int ii = threadIdx.x; ii < n; ii += NT) {
      const int i = pack[ii];
      const size_t g = 3 * (size_t)(off + ii);
      const T fx = f[g], fy = f[g + 1], fz = f[g + 2];
      f[3 * (size_t)i] += fx;
      f[3 * (size_t)i + 1] += fy;
      f[3 * (size_t)i + 2] += fz;
    }
    __syncthreads();
  }
}

// IntegratorNVE FinalIntegrateFunctor of one step then InitialIntegrateFunctor
// of the next for atom i (src/integrator_nve.cpp:104-112 then 65-74): two
// velocity half-kicks and the drift, in the reference's operation order
template <bool MIXED>
__device__ __forceinline__ void emd_nve_final_initial_at(
    int i, double *__restrict__ x, double *__restrict__ v, const double *__restrict__ f,
    const float *__restrict__ f32, const int *__restrict__ type, const double *__restrict__ mass,
    double dtf, double dtv) {
#pragma clang fp contract(off)
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

// CommSerial::TagHaloUpdateSelf (comm_serial.h:184-197) for ghost g: its
// owner's position plus the periodic shift of each shifted dimension (image
// code 1: -L, 2: +L), one addition per dimension
__device__ __forceinline__ void emd_halo_update_at(int g, double *__restrict__ x,
                                                   const int *__restrict__ root,
                                                   const int *__restrict__ image, int n_local,
                                                   double lx, double ly, double lz) {
  const int r = root[g];
  const int img = image[g];
  const double L[3] = {lx, ly, lz};
  const size_t dst = 3 * (size_t)(n_local + g), src = 3 * (size_t)r;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int c = (img >> (2 * d)) & 3;
    const double xs = x[src + d];
    x[dst + d] = (c == 0) ? xs : xs + (c == 1 ? -L[d] : L[d]);
  }
}

// fold_phases of a halo (emd_comm.hip); returns the number of ghosts
long emd_fold_phases_of(const emd_system *s, const emd_halo *h, fold_phases &ph);
constexpr int FOLD_ONE_THREADS = 1024;
constexpr int FOLD_ONE_MAX = 16384;  // ghosts; above this the per-phase launches are faster

// bin grid of a neighbour build: interior bins are the non-halo ones, one
// block per interior bin; the 27-bin stencil order is dx, dy, dz = -1..1
// with dz fastest (stencil slot 13 is the bin itself)
struct neigh_geom {
  int nbx, nby, nbz, nhalo;  // full grid incl. halo
  int nix, niy, niz;         // interior bins
};

inline neigh_geom emd_neigh_geom(const emd_bins *b) {
  neigh_geom g;
  g.nbx = b->nbinx; g.nby = b->nbiny; g.nbz = b->nbinz; g.nhalo = b->nhalo;
  g.nix = b->nbinx - 2 * b->nhalo; g.niy = b->nbiny - 2 * b->nhalo;
  g.niz = b->nbinz - 2 * b->nhalo;
  return g;
}

// the interior bin of block lr and its 27 stencil bins' offsets and
// exclusive count prefix (s_pre[27] = candidates of the bin); call with
// all threads of the block, blockDim.x >= 27
__device__ __forceinline__ int emd_stencil_setup(const neigh_geom &g, int lr,
                                                 const int *__restrict__ bincount,
                                                 const int *__restrict__ binoffsets, int *s_off,
                                                 int *s_pre) {
  const int bx = lr / (g.niy * g.niz) + g.nhalo;
  const int by = (lr / g.niz) % g.niy + g.nhalo;
  const int bz = lr % g.niz + g.nhalo;
  const int tid = threadIdx.x;
  if (tid < 27) {
    const int ddx = tid / 9 - 1, ddy = (tid / 3) % 3 - 1, ddz = tid % 3 - 1;
    const int b = ((bx + ddx) * g.nby + (by + ddy)) * g.nbz + (bz + ddz);
    s_off[tid] = binoffsets[b];
    s_pre[tid + 1] = bincount[b];
  }
  __syncthreads();
  if (tid == 0) {
    s_pre[0] = 0;
    for (int k = 1; k <= 27; k++) s_pre[k] += s_pre[k - 1];
  }
  __syncthreads();
  return (bx * g.nby + by) * g.nbz + bz;
}

// stencil slot k of candidate c: s_pre[k] <= c < s_pre[k+1] (binary search)
__device__ __forceinline__ int emd_stencil_slot(const int *s_pre, int c) {
  int lo = 0, hi = 27;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s_pre[mid] <= c) lo = mid;
    else hi = mid;
  }
  return lo;
}
