// emd_comm.hip -- serial periodic-image exchange (CommSerial on one GPU).
//
// exchange       : wrap local positions back into the box once per dimension
// exchange_halo  : six compaction passes (+x,-x,+y,-y,+z,-z faces), each a
//                  predicate scan whose emit step writes the ghost image in
//                  ascending source order; every ghost also records its
//                  owning local atom ("root") and its periodic shift
// update_halo    : one gather kernel over all ghosts: x_g = x_root + shift,
//                  one addition per shifted dimension (the same single
//                  rounding the phase-by-phase copy performs)
// update_force   : phases 5..0, f[pack] += f[ghost] (targets unique per phase)
#include "emd_internal.hpp"

// CommSerial::TagExchangeSelf, src/comm_types/comm_serial.h:95-108
__global__ void k_wrap(double *__restrict__ x, int n, double lx, double ly, double lz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double L[3] = {lx, ly, lz};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double xd = x[3 * (size_t)i + d];
    double r = xd;
    if (xd > L[d]) r -= L[d];
    if (xd < 0) r += L[d];
    x[3 * (size_t)i + d] = r;
  }
}

extern "C" emd_status emd_comm_exchange(emd_ctx *ctx, emd_system *s) {
  if (!ctx || !s) return EMD_ERR_ARG;
  if (s->n_local <= 0) return EMD_OK;
  EMD_LAUNCH(ctx, "comm_exchange", k_wrap, emd_blocks(s->n_local, 256), 256, 0, s->x, s->n_local,
             s->domain[0], s->domain[1], s->domain[2]);
  return EMD_OK;
}

// one exchange_halo phase as a scan functor (CommSerial::TagHaloSelf,
// comm_serial.h:112-181)
struct halo_phase_op {
  double *x, *vel, *q;
  int *type, *id;
  int dim, hi_side;
  double thr, shift;
  int dst0, n_local;
  int *pack, *root, *image;
  int code;
  __device__ int value(int i) const {
    const double xd = x[3 * (size_t)i + dim];
    return hi_side ? (xd >= thr) : (xd <= thr);
  }
  __device__ void emit(int i, int k, int v) const {
    if (!v) return;
    const int g = dst0 + k;
    pack[k] = i;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      x[3 * (size_t)g + d] = x[3 * (size_t)i + d];
      vel[3 * (size_t)g + d] = vel[3 * (size_t)i + d];
    }
    x[3 * (size_t)g + dim] += shift;
    q[g] = q[i];
    id[g] = id[i];
    type[g] = type[i];
    const int gi = g - n_local;
    if (i < n_local) {
      root[gi] = i;
      image[gi] = code;
    } else {
      root[gi] = root[i - n_local];
      image[gi] = image[i - n_local] | code;
    }
  }
};

// All six exchange_halo phases in one launch of one workgroup, for small
// systems where the per-phase scans are launch- and sync-bound.  Each phase
// is the same stable compaction as the scan path (ascending source index,
// halo_phase_op's predicate), over up to 16 rounds of 1024 candidates:
// every round's predicate first, one ballot count per (round, wavefront),
// one scan of those counts, then the emits.  A phase writes only the ghost
// indices (pack list, root, image code); the predicate of a ghost candidate
// reads its root's coordinate, which equals the ghost's own in every
// dimension a later phase tests (a ghost is shifted only in the dimensions
// of the phases that made it, and the -d phase does not take the +d
// phase's ghosts).  After the six phases one pass writes every ghost's
// position (its root's plus one shift per shifted dimension, the single
// rounding the phase-by-phase copy performs), velocity, charge, id and type.
// Capacities (system arrays, pack lists, the root/image tables) are checked
// per ghost: a phase that does not fit is still counted but not written, and
// the kernel stops there, reporting the phase; the host grows and runs it
// again.
constexpr int HALO_ONE_THREADS = 1024;
constexpr int HALO_ONE_MAX = 16384;  // local + ghost atoms; above, the per-phase scans
constexpr int HALO_ONE_ROUNDS = HALO_ONE_MAX / HALO_ONE_THREADS;
constexpr int HALO_ONE_WAVES = HALO_ONE_THREADS / EMD_WAVE;
static_assert(HALO_ONE_ROUNDS * HALO_ONE_WAVES <= 4 * EMD_WAVE, "one wavefront scans the counts");

struct halo_one_args {
  halo_phase_op op[6];
  int cap_pack[6];
  int n_local, n_max, root_cap;
  int *out;  // [0..5] ghosts per phase, [6] 1 + the phase that did not fit (0: all fit;
             // 7: the system is too large for one workgroup)
};

__global__ void __launch_bounds__(HALO_ONE_THREADS) k_halo_one(halo_one_args a) {
  __shared__ int s_cnt[HALO_ONE_ROUNDS * HALO_ONE_WAVES];
  __shared__ int s_tot;
  const int tid = threadIdx.x, wid = tid / EMD_WAVE, lane = tid % EMD_WAVE;
  const int n_local = a.n_local;
  int n_ghost = 0, prev = 0, failed = 0;
  for (int phase = 0; phase < 6 && !failed; phase++) {
    const halo_phase_op &op = a.op[phase];
    const int ncand = n_local + n_ghost - ((phase & 1) ? prev : 0);
    if (ncand > HALO_ONE_MAX) {  // more candidates than one workgroup's rounds
      failed = 7;
      break;
    }
    const int dst0 = n_local + n_ghost;
    const int room = min(min(a.n_max - dst0, a.cap_pack[phase]), a.root_cap - n_ghost);
    const int nr = (ncand + HALO_ONE_THREADS - 1) / HALO_ONE_THREADS;
    // the predicates of all rounds first (their loads in flight together);
    // a ghost candidate is tested on its root's coordinate
    unsigned flags = 0;
#pragma unroll
    for (int r = 0; r < HALO_ONE_ROUNDS; r++) {
      const int i = r * HALO_ONE_THREADS + tid;
      if (i < ncand) {
        const int src = i < n_local ? i : op.root[i - n_local];
        const double xd = op.x[3 * (size_t)src + op.dim];
        if (op.hi_side ? (xd >= op.thr) : (xd <= op.thr)) flags |= 1u << r;
      }
    }
    // kept candidates per (round, wavefront), then their exclusive scan in
    // (round, wavefront) order, which is ascending candidate order
#pragma unroll
    for (int r = 0; r < HALO_ONE_ROUNDS; r++) {
      if (r < nr) {
        const unsigned long long m = __ballot((flags >> r) & 1);
        if (lane == 0) s_cnt[r * HALO_ONE_WAVES + wid] = __popcll(m);
      }
    }
    __syncthreads();
    if (wid == 0) {
      const int nc = nr * HALO_ONE_WAVES;
      int c[4], sum = 0;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int t = 4 * lane + q;
        c[q] = t < nc ? s_cnt[t] : 0;
        sum += c[q];
      }
      const int inc = emd_wave_incl_scan(sum);
      int run = inc - sum;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int t = 4 * lane + q;
        if (t < nc) s_cnt[t] = run;
        run += c[q];
      }
      if (lane == EMD_WAVE - 1) s_tot = inc;
    }
    __syncthreads();
    const int k = s_tot;
    // emits: the sources' root and image first (ghost sources only), then
    // the pack entries and the new ghosts' root and image
    int sroot[HALO_ONE_ROUNDS], simg[HALO_ONE_ROUNDS];
#pragma unroll
    for (int r = 0; r < HALO_ONE_ROUNDS; r++) {
      const int i = r * HALO_ONE_THREADS + tid;
      sroot[r] = i;
      simg[r] = 0;
      if (((flags >> r) & 1) && i >= n_local) {
        sroot[r] = op.root[i - n_local];
        simg[r] = op.image[i - n_local];
      }
    }
#pragma unroll
    for (int r = 0; r < HALO_ONE_ROUNDS; r++) {
      if (r < nr) {
        const int v = (flags >> r) & 1;
        const unsigned long long m = __ballot(v);
        if (v) {
          const int pos = s_cnt[r * HALO_ONE_WAVES + wid] + __popcll(m & emd_lanemask_lt());
          if (pos < room) {
            const int gi = n_ghost + pos;
            op.pack[pos] = r * HALO_ONE_THREADS + tid;
            op.root[gi] = sroot[r];
            op.image[gi] = simg[r] | op.code;
          }
        }
      }
    }
    if (tid == 0) a.out[phase] = k;
    if (k > room) failed = phase + 1;
    prev = k;
    n_ghost += k;
    __threadfence();   // this phase's roots are read by the next phases
    __syncthreads();
  }
  if (tid == 0) a.out[6] = failed;
  if (failed) return;
  // the ghosts' data from their roots (CommSerial::TagHaloSelf copies, the
  // position as k_halo_update forms it)
  const halo_phase_op &op = a.op[0];
  // box lengths: the -d phase's shift is +L_d (image code 2), the +d
  // phase's -L_d (code 1)
  const double L[3] = {a.op[1].shift, a.op[3].shift, a.op[5].shift};
  for (int g = tid; g < n_ghost; g += HALO_ONE_THREADS) {
    const int r = op.root[g], img = op.image[g];
    const size_t dst = 3 * (size_t)(n_local + g), src = 3 * (size_t)r;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int c = (img >> (2 * d)) & 3;
      const double xs = op.x[src + d];
      op.x[dst + d] = (c == 0) ? xs : xs + (c == 1 ? -L[d] : L[d]);
      op.vel[dst + d] = op.vel[src + d];
    }
    op.q[n_local + g] = op.q[r];
    op.id[n_local + g] = op.id[r];
    op.type[n_local + g] = op.type[r];
  }
}

static emd_status exchange_halo_one(emd_ctx *ctx, emd_system *s, double comm_depth,
                                    emd_halo *h) {
  const int n_local = s->n_local;
  for (int attempt = 0; attempt < 8; attempt++) {
    halo_one_args a;
    for (int phase = 0; phase < 6; phase++) {
      const int dim = phase / 2, hi_side = (phase % 2 == 0);
      halo_phase_op &op = a.op[phase];
      op.dim = dim;
      op.hi_side = hi_side;
      op.thr = hi_side ? s->hi[dim] - comm_depth : s->lo[dim] + comm_depth;
      op.shift = hi_side ? -s->domain[dim] : s->domain[dim];
      op.dst0 = 0;  // unused: k_halo_one places each phase itself
      op.n_local = n_local;
      op.code = (hi_side ? 1 : 2) << (2 * dim);
      op.x = s->x; op.vel = s->v; op.q = s->q; op.type = s->type; op.id = s->id;
      op.pack = h->pack[phase];
      op.root = h->root;
      op.image = h->image;
      a.cap_pack[phase] = h->pack[phase] ? h->cap[phase] : 0;
    }
    a.n_local = n_local;
    a.n_max = s->n_max;
    a.root_cap = (h->root && h->image) ? h->root_cap : 0;
    a.out = ctx->d_small + 8;
    EMD_LAUNCH(ctx, "comm_exchange_halo", k_halo_one, 1, HALO_ONE_THREADS, 0, a);
    EMD_HIP(ctx, hipMemcpyAsync(ctx->h_small + 8, ctx->d_small + 8, 7 * sizeof(int),
                                hipMemcpyDeviceToHost, ctx->stream));
    EMD_TRY(emd_wait(ctx));
    const int *cnt = ctx->h_small + 8;
    const int failed = cnt[6];
    if (failed == 7) return EMD_ERR_OVERFLOW;  // the caller runs the per-phase scans
    int n_ghost = 0;
    for (int p = 0; p < (failed ? failed : 6); p++) n_ghost += cnt[p];
    if (!failed) {
      for (int p = 0; p < 6; p++) h->num_ghost[p] = cnt[p];
      s->n_ghost = n_ghost;
      h->n_ghost = n_ghost;
      return EMD_OK;
    }
    // grow what the failed phase needs, with room for the phases after it
    const int fp = failed - 1;
    long want = (long)n_local + 2l * n_ghost + 64;
    if (want > s->n_max) EMD_TRY(emd_system_grow(ctx, s, (int)(want + want / 10)));
    EMD_TRY(emd_ensure_int(ctx, &h->pack[fp], &h->cap[fp], cnt[fp] + cnt[fp] / 8 + 64));
    for (int p = fp + 1; p < 6; p++)
      EMD_TRY(emd_ensure_int(ctx, &h->pack[p], &h->cap[p], cnt[fp] + cnt[fp] / 2 + 64));
    int rcap = h->root ? h->root_cap : 0, icap = h->image ? h->root_cap : 0;
    const long rwant = 2l * n_ghost + 64;
    EMD_TRY(emd_grow_keep_int(ctx, &h->root, &rcap, rwant, 0));
    EMD_TRY(emd_grow_keep_int(ctx, &h->image, &icap, rwant, 0));
    h->root_cap = rcap < icap ? rcap : icap;
  }
  return emd_set_error(ctx, EMD_ERR_STATE, "emd_comm_exchange_halo: capacities did not settle");
}

extern "C" emd_status emd_comm_exchange_halo(emd_ctx *ctx, emd_system *s, double comm_depth,
                                             emd_halo *h) {
  if (!ctx || !s || !h) return EMD_ERR_ARG;
  const int n_local = s->n_local;
  if (ctx->halo_one && (long)n_local + h->n_ghost <= HALO_ONE_MAX && n_local <= HALO_ONE_MAX / 2) {
    const emd_status st = exchange_halo_one(ctx, s, comm_depth, h);
    if (st != EMD_ERR_OVERFLOW) return st;
    // the halo outgrew one workgroup: the per-phase scans below
  }
  int n_ghost = 0;
  for (int phase = 0; phase < 6; phase++) {
    const int dim = phase / 2;
    const int hi_side = (phase % 2 == 0);
    const int ncand = n_local + n_ghost - ((phase % 2 == 1) ? h->num_ghost[phase - 1] : 0);
    halo_phase_op op;
    op.dim = dim;
    op.hi_side = hi_side;
    op.thr = hi_side ? s->hi[dim] - comm_depth : s->lo[dim] + comm_depth;
    op.shift = hi_side ? -s->domain[dim] : s->domain[dim];
    op.dst0 = n_local + n_ghost;
    op.n_local = n_local;
    op.code = (hi_side ? 1 : 2) << (2 * dim);
    op.x = s->x; op.vel = s->v; op.q = s->q; op.type = s->type; op.id = s->id;
    op.pack = h->pack[phase]; op.root = h->root; op.image = h->image;
    int count = 0;
    EMD_TRY(emd_scan_count(ctx, op, ncand, "comm_exchange_halo", &count));
    // capacity: system arrays, the phase's pack list, the ghost root/image tables
    if (n_local + n_ghost + count > s->n_max) {
      long want = (long)(n_local + n_ghost + count);
      want += want / 10 + 64;
      EMD_TRY(emd_system_grow(ctx, s, (int)want));
    }
    EMD_TRY(emd_ensure_int(ctx, &h->pack[phase], &h->cap[phase], count));
    {
      // root and image grow together and share root_cap
      int rcap = h->root ? h->root_cap : 0, icap = h->image ? h->root_cap : 0;
      EMD_TRY(emd_grow_keep_int(ctx, &h->root, &rcap, n_ghost + count, n_ghost));
      EMD_TRY(emd_grow_keep_int(ctx, &h->image, &icap, n_ghost + count, n_ghost));
      h->root_cap = rcap < icap ? rcap : icap;
    }
    h->num_ghost[phase] = count;
    // pointers may have moved
    op.x = s->x; op.vel = s->v; op.q = s->q; op.type = s->type; op.id = s->id;
    op.pack = h->pack[phase];
    op.root = h->root;
    op.image = h->image;
    EMD_TRY(emd_scan_apply(ctx, op, ncand, "comm_exchange_halo"));
    n_ghost += count;
  }
  s->n_ghost = n_ghost;
  h->n_ghost = n_ghost;
  return EMD_OK;
}

// CommSerial::TagHaloUpdateSelf (comm_serial.h:184-197) for all six phases at
// once, positions only: ghost velocities/charges/ids are copies made at
// exchange_halo time and no kernel reads them afterwards
__global__ void k_halo_update(double *__restrict__ x, const int *__restrict__ root,
                              const int *__restrict__ image, int n_local, int n_ghost, double lx,
                              double ly, double lz) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_ghost) return;
  emd_halo_update_at(g, x, root, image, n_local, lx, ly, lz);
}

extern "C" emd_status emd_comm_update_halo(emd_ctx *ctx, emd_system *s, const emd_halo *h) {
  if (!ctx || !s || !h) return EMD_ERR_ARG;
  if (h->n_ghost <= 0) return EMD_OK;
  EMD_LAUNCH(ctx, "comm_update_halo", k_halo_update, emd_blocks(h->n_ghost, 256), 256, 0, s->x,
             h->root, h->image, s->n_local, h->n_ghost, s->domain[0], s->domain[1],
             s->domain[2]);
  return EMD_OK;
}

// CommSerial::TagHaloForceSelf, comm_serial.h:200-213
template <class T>
__global__ void k_halo_fold(T *__restrict__ f, const int *__restrict__ pack, int off, int n) {
  int ii = blockIdx.x * blockDim.x + threadIdx.x;
  if (ii >= n) return;
  const int i = pack[ii];
  const size_t g = 3 * (size_t)(off + ii);
  const T fx = f[g], fy = f[g + 1], fz = f[g + 2];
  f[3 * (size_t)i] += fx;
  f[3 * (size_t)i + 1] += fy;
  f[3 * (size_t)i + 2] += fz;
}

// all six phases in one launch of one workgroup (emd_fold_phases_one): per
// atom the same additions in the same order as the per-phase launches
template <class T>
__global__ void __launch_bounds__(FOLD_ONE_THREADS)
k_halo_fold_one(T *__restrict__ f, fold_phases ph) {
  emd_fold_phases_one<T, FOLD_ONE_THREADS>(f, ph);
}

long emd_fold_phases_of(const emd_system *s, const emd_halo *h, fold_phases &ph) {
  long total = 0;
  for (int p = 0; p < 6; p++) {
    ph.pack[p] = h->pack[p];
    ph.num[p] = h->num_ghost[p] > 0 ? h->num_ghost[p] : 0;
    ph.off[p] = p == 0 ? s->n_local : ph.off[p - 1] + h->num_ghost[p - 1];
    total += ph.num[p];
  }
  return total;
}

extern "C" emd_status emd_comm_update_force(emd_ctx *ctx, emd_system *s, const emd_halo *h,
                                            int precision) {
  if (!ctx || !s || !h) return EMD_ERR_ARG;
  fold_phases ph;
  const long total = emd_fold_phases_of(s, h, ph);
  if (total == 0) return EMD_OK;
  const int *off = ph.off;
  if (total <= FOLD_ONE_MAX && ctx->fold_one) {
    if (precision == EMD_MIXED)
      EMD_LAUNCH(ctx, "comm_update_force", k_halo_fold_one<float>, 1, FOLD_ONE_THREADS, 0,
                 s->f32, ph);
    else
      EMD_LAUNCH(ctx, "comm_update_force", k_halo_fold_one<double>, 1, FOLD_ONE_THREADS, 0, s->f,
                 ph);
    return EMD_OK;
  }
  for (int p = 5; p >= 0; p--) {
    const int n = h->num_ghost[p];
    if (n <= 0) continue;
    if (precision == EMD_MIXED)
      EMD_LAUNCH(ctx, "comm_update_force", k_halo_fold<float>, emd_blocks(n, 256), 256, 0, s->f32,
                 h->pack[p], off[p], n);
    else
      EMD_LAUNCH(ctx, "comm_update_force", k_halo_fold<double>, emd_blocks(n, 256), 256, 0, s->f,
                 h->pack[p], off[p], n);
  }
  return EMD_OK;
}

extern "C" emd_status emd_halo_free(emd_ctx *ctx, emd_halo *h) {
  if (!ctx || !h) return EMD_ERR_ARG;
  hipStreamSynchronize(ctx->stream);
  for (int p = 0; p < 6; p++) {
    if (h->pack[p]) hipFree(h->pack[p]);
    h->pack[p] = nullptr;
    h->cap[p] = 0;
    h->num_ghost[p] = 0;
  }
  if (h->root) hipFree(h->root);
  if (h->image) hipFree(h->image);
  h->root = h->image = nullptr;
  h->root_cap = 0;
  h->n_ghost = 0;
  return EMD_OK;
}
