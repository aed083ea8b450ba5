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

extern "C" emd_status emd_comm_exchange_halo(emd_ctx *ctx, emd_system *s, double comm_depth,
                                             emd_halo *h) {
  if (!ctx || !s || !h) return EMD_ERR_ARG;
  const int n_local = s->n_local;
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

extern "C" emd_status emd_comm_update_force(emd_ctx *ctx, emd_system *s, const emd_halo *h,
                                            int precision) {
  if (!ctx || !s || !h) return EMD_ERR_ARG;
  int off[6];
  off[0] = s->n_local;
  for (int p = 1; p < 6; p++) off[p] = off[p - 1] + h->num_ghost[p - 1];
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
