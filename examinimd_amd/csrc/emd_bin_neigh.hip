// emd_bin_neigh.hip -- cell binning (counting sort) and neighbour lists.
//
// Binning: bin id per atom (Kokkos BinOp3D arithmetic: int(mul*(x-min)),
// mul = nbin/(max-min), x-major flattening), then a counting sort: atomic
// counts, an exclusive scan of the counts, a scatter through per-bin cursors
// and a per-bin ordering pass, so the permutation lists each bin's atoms in
// ascending index (the order the reference produces on its serial backend;
// deterministic here).
//
// Neighbour build: one 256-thread block per interior bin.  The block stages
// the 27 stencil bins' candidates (index + FP64 position) in LDS in stencil
// order; each wavefront takes atoms of the centre bin and tests 64
// candidates per step, compacting hits with ballot + popcount so every row
// comes out in stencil-then-permutation order.  Two passes (count, fill)
// with an exclusive scan between them, as NeighborCSR does.
#include "emd_internal.hpp"

#include <type_traits>

__global__ void k_max_int(const int *__restrict__ a, int n, int *__restrict__ out);

// ------------------------------------------------------------------ binning
// BinOp3D::bin (Kokkos_Sort.hpp) as used by binning_kksort.cpp:105-113: the
// bin of an atom, clamped into the grid (an atom outside it raises the
// lost-atom flag instead of indexing out of bounds)
__device__ __forceinline__ int bin_index(const double *xi, double minx, double miny,
                                         double minz, double mulx, double muly, double mulz,
                                         int nbx, int nby, int nbz, bool &outside) {
  int bx = (int)(mulx * (xi[0] - minx));
  int by = (int)(muly * (xi[1] - miny));
  int bz = (int)(mulz * (xi[2] - minz));
  outside = bx < 0 || bx >= nbx || by < 0 || by >= nby || bz < 0 || bz >= nbz;
  bx = bx < 0 ? 0 : (bx >= nbx ? nbx - 1 : bx);
  by = by < 0 ? 0 : (by >= nby ? nby - 1 : by);
  bz = bz < 0 ? 0 : (bz >= nbz ? nbz - 1 : bz);
  return (bx * nby + by) * nbz + bz;
}

// Counting sort, the large-range binning path (BinSort's result: per bin
// its atoms in ascending index, src/binning_types/binning_kksort.cpp:105-125).
// (1) each atom's bin id, counted into the bins with atomics;
// (2) exclusive scan of the counts (wavefront scans, emd_scan) -> offsets;
// (3) scatter through per-bin cursors;
// (4) each bin's slice put in ascending atom order by one wavefront: a
//     lane's rank is the number of smaller indices in the slice (64-wide
//     shuffle sweep, registers only; slices over 64 atoms, which a dense
//     cluster or a wide bin can give, take an insertion sort).
// Atoms were sorted by bin at the previous rebuild and have moved little,
// so consecutive atoms mostly share a bin: in (1) and (3) a wavefront's runs
// of equal bin ids make one atomic per run (the run's first lane adds the
// run length; the lanes of a run take consecutive slots in index order).

// the run of equal bin ids (consecutive lanes) that the lane belongs to:
// its first lane and its length; lanes without an atom form runs of their own
__device__ __forceinline__ void bin_run(int b, bool valid, int &first, int &len) {
  const int lane = __lane_id();
  const int prev = __shfl_up(b, 1, EMD_WAVE);
  // lanes without an atom are a suffix of the wavefront; each is a head
  const bool head = lane == 0 || !valid || b != prev;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = heads & (emd_lanemask_lt() | (1ull << lane));
  first = 63 - __clzll(upto);
  const unsigned long long after = lane == EMD_WAVE - 1 ? 0ull : heads >> (lane + 1);
  const int end = after ? lane + __ffsll((long long)after) : EMD_WAVE;
  len = end - first;
}

__global__ void k_bin_count(const double *__restrict__ x, int r0, int n, double minx,
                            double miny, double minz, double mulx, double muly, double mulz,
                            int nbx, int nby, int nbz, int *__restrict__ binid,
                            int *__restrict__ count, int *__restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  int b = -1;
  if (valid) {
    bool outside;
    b = bin_index(x + 3 * (size_t)(r0 + i), minx, miny, minz, mulx, muly, mulz, nbx, nby, nbz,
                  outside);
    if (outside) atomicOr(err, 1);
    binid[i] = b;
  }
  int first, len;
  bin_run(b, valid, first, len);
  if (valid && __lane_id() == first) atomicAdd(&count[b], len);
}

struct bin_offsets_op {
  const int *count;
  int *offset, *cursor;
  __device__ int value(int b) const { return count[b]; }
  __device__ void emit(int b, int excl, int) const {
    offset[b] = excl;
    cursor[b] = excl;
  }
};

__global__ void k_bin_scatter(const int *__restrict__ binid, int n, int *__restrict__ cursor,
                              int *__restrict__ permute) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  const int b = valid ? binid[i] : -1;
  int first, len;
  bin_run(b, valid, first, len);
  int base = 0;
  if (valid && __lane_id() == first) base = atomicAdd(&cursor[b], len);
  base = __shfl(base, first, EMD_WAVE);
  if (valid) permute[base + (__lane_id() - first)] = i;
}

constexpr int BIN_ORDER_THREADS = 256;
__global__ void __launch_bounds__(BIN_ORDER_THREADS)
k_bin_order(const int *__restrict__ count, const int *__restrict__ offset, int nbins,
            int *__restrict__ permute) {
  const int b = blockIdx.x * (BIN_ORDER_THREADS / EMD_WAVE) + threadIdx.x / EMD_WAVE;
  const int lane = threadIdx.x % EMD_WAVE;
  int c = 0;
  if (b < nbins) {
    c = count[b];
    int *p = permute + offset[b];
    if (c <= EMD_WAVE) {
      const int v = lane < c ? p[lane] : 0x7fffffff;
      // slices already in order (one run of the previous order) stay
      const int next = __shfl_down(v, 1, EMD_WAVE);
      if (__ballot(lane + 1 < c && next < v) != 0ull) {
        int rank = 0;
        for (int k = 0; k < c; k++) rank += __shfl(v, k, EMD_WAVE) < v;
        if (lane < c) p[rank] = v;
      }
    } else if (lane == 0) {
      for (int k = 1; k < c; k++) {
        const int v = p[k];
        int m = k - 1;
        while (m >= 0 && p[m] > v) {
          p[m + 1] = p[m];
          m--;
        }
        p[m + 1] = v;
      }
    }
  }
}

// Small ranges (n <= BIN_ONE_MAX_N atoms, nbins <= BIN_ONE_MAX_BINS) in one
// workgroup launch: bin ids, an LDS histogram, the exclusive scan of the
// counts, an LDS scatter, then an insertion sort of each bin's slice of the
// LDS permutation (ascending index: the radix path's order), the largest
// count and the out-of-grid flag -- one launch instead of the key, sort,
// bounds and max kernels.
constexpr int BIN_ONE_THREADS = 1024;
constexpr int BIN_ONE_MAX_BINS = 4096;
constexpr int BIN_ONE_MAX_N = 7168;  // LDS: 2 x 16 KiB bin tables + 28 KiB permutation

__global__ void __launch_bounds__(BIN_ONE_THREADS)
k_bin_one(const double *__restrict__ x, int r0, int n, double minx, double miny, double minz,
          double mulx, double muly, double mulz, int nbx, int nby, int nbz, int nbins,
          int *__restrict__ bincount, int *__restrict__ binoffsets, int *__restrict__ permute,
          int *__restrict__ small) {
  __shared__ int s_cnt[BIN_ONE_MAX_BINS];  // counts, then scatter cursors
  __shared__ int s_off[BIN_ONE_MAX_BINS];
  __shared__ int s_perm[BIN_ONE_MAX_N];
  __shared__ int s_w[BIN_ONE_THREADS / EMD_WAVE];
  __shared__ int s_err, s_max;
  const int tid = threadIdx.x, wid = tid / EMD_WAVE, lane = tid % EMD_WAVE;
  for (int b = tid; b < nbins; b += BIN_ONE_THREADS) s_cnt[b] = 0;
  if (tid == 0) { s_err = 0; s_max = 0; }
  __syncthreads();
  // pass 1: histogram (the bin id is recomputed in pass 2, not stored)
  auto bin_of = [&](int i, bool &bad) {
    return bin_index(x + 3 * (size_t)(r0 + i), minx, miny, minz, mulx, muly, mulz, nbx, nby, nbz,
                     bad);
  };
  for (int i = tid; i < n; i += BIN_ONE_THREADS) {
    bool bad;
    const int b = bin_of(i, bad);
    if (bad) s_err = 1;
    atomicAdd(&s_cnt[b], 1);
  }
  __syncthreads();
  // exclusive scan of the counts, 1024 bins per round
  int carry = 0, mymax = 0;
  for (int base = 0; base < nbins; base += BIN_ONE_THREADS) {
    const int b = base + tid;
    const int v = b < nbins ? s_cnt[b] : 0;
    mymax = v > mymax ? v : mymax;
    const int inc = emd_wave_incl_scan(v);
    if (lane == EMD_WAVE - 1) s_w[wid] = inc;
    __syncthreads();
    int woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < BIN_ONE_THREADS / EMD_WAVE; w++) {
      const int sw = s_w[w];
      woff += (w < wid) ? sw : 0;
      tot += sw;
    }
    if (b < nbins) {
      const int off = carry + woff + inc - v;
      bincount[b] = v;
      binoffsets[b] = off;
      s_off[b] = off;
      s_cnt[b] = off;  // scatter cursor
    }
    carry += tot;
    __syncthreads();
  }
  atomicMax(&s_max, mymax);
  // pass 2: scatter (any order), then sort each bin's slice
  for (int i = tid; i < n; i += BIN_ONE_THREADS) {
    bool bad;
    const int b = bin_of(i, bad);
    s_perm[atomicAdd(&s_cnt[b], 1)] = i;
  }
  __syncthreads();
  for (int b = tid; b < nbins; b += BIN_ONE_THREADS) {
    int *p = s_perm + s_off[b];
    const int c = s_cnt[b] - s_off[b];
    for (int k = 1; k < c; k++) {
      const int v = p[k];
      int m = k - 1;
      while (m >= 0 && p[m] > v) {
        p[m + 1] = p[m];
        m--;
      }
      p[m + 1] = v;
    }
  }
  __syncthreads();
  for (int i = tid; i < n; i += BIN_ONE_THREADS) permute[i] = s_perm[i];
  if (tid == 0) {
    small[1] = s_max;
    small[2] = s_err;
  }
}

// BinSort::sort over x, v, f, type, id, q: new[i] = old[permute[i]]
__global__ void k_bin_gather(const int *__restrict__ perm, int r0, int n,
                             const double *__restrict__ x, const double *__restrict__ v,
                             const double *__restrict__ f, const double *__restrict__ q,
                             const int *__restrict__ type, const int *__restrict__ id,
                             double *__restrict__ tx, double *__restrict__ tv,
                             double *__restrict__ tf, double *__restrict__ tq,
                             int *__restrict__ ttype, int *__restrict__ tid) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t s = (size_t)r0 + perm[i];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    tx[3 * (size_t)i + d] = x[3 * s + d];
    tv[3 * (size_t)i + d] = v[3 * s + d];
    tf[3 * (size_t)i + d] = f[3 * s + d];
  }
  tq[i] = q[s];
  ttype[i] = type[s];
  tid[i] = id[s];
}

__global__ void k_fill_int(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// the sorted arrays back into the system's range, one launch for all six
__global__ void k_bin_copy_back(int r0, int n, const double *__restrict__ tx,
                                const double *__restrict__ tv, const double *__restrict__ tf,
                                const double *__restrict__ tq, const int *__restrict__ tt,
                                const int *__restrict__ ti, double *__restrict__ x,
                                double *__restrict__ v, double *__restrict__ f,
                                double *__restrict__ q, int *__restrict__ type,
                                int *__restrict__ id) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t s3 = 3 * (size_t)i, d3 = 3 * ((size_t)r0 + i);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    x[d3 + d] = tx[s3 + d];
    v[d3 + d] = tv[s3 + d];
    f[d3 + d] = tf[s3 + d];
  }
  q[r0 + i] = tq[i];
  type[r0 + i] = tt[i];
  id[r0 + i] = ti[i];
}

// common tail: largest count and out-of-grid flag to the host (the one sync
// of a binning), then BinSort::sort of the range when asked
static emd_status bin_finish(emd_ctx *ctx, emd_system *s, emd_bins *b, int r0, int n, int sort) {
  EMD_HIP(ctx, hipMemcpyAsync(ctx->h_small + 1, ctx->d_small + 1, 2 * sizeof(int),
                              hipMemcpyDeviceToHost, ctx->stream));
  EMD_TRY(emd_wait(ctx));
  b->max_count = ctx->h_small[1];
  if (ctx->h_small[2])
    return emd_set_error(ctx, EMD_ERR_OVERFLOW,
                         "emd_binning: an atom lies outside the bin grid (lost atom?)");
  if (sort && n > 0) {
    size_t bytes = (size_t)n * (72 + 8 + 8);
    EMD_TRY(emd_ensure(ctx, &ctx->d_tmp, &ctx->tmp_cap, bytes));
    char *t = (char *)ctx->d_tmp;
    double *tx = (double *)t, *tv = tx + 3 * (size_t)n, *tf = tv + 3 * (size_t)n,
           *tq = tf + 3 * (size_t)n;
    int *tt = (int *)(tq + n), *ti = tt + n;
    EMD_LAUNCH(ctx, "binning_sort", k_bin_gather, emd_blocks(n, 256), 256, 0, b->permute, r0, n,
               s->x, s->v, s->f, s->q, s->type, s->id, tx, tv, tf, tq, tt, ti);
    EMD_LAUNCH(ctx, "binning_sort", k_bin_copy_back, emd_blocks(n, 256), 256, 0, r0, n, tx, tv,
               tf, tq, tt, ti, s->x, s->v, s->f, s->q, s->type, s->id);
  }
  return EMD_OK;
}

extern "C" emd_status emd_binning(emd_ctx *ctx, emd_system *s, double dx_in, double dy_in,
                                  double dz_in, int halo_depth, int do_local, int do_ghost,
                                  int sort, emd_bins *b) {
  if (!ctx || !s || !b) return EMD_ERR_ARG;
  if (!(do_local || do_ghost)) return EMD_OK;
  EMD_ARG(ctx, dx_in > 0 && dy_in > 0 && dz_in > 0, "bin sizes must be positive");
  b->nhalo = halo_depth;
  const int r0 = do_local ? 0 : s->n_local;
  const int r1 = do_ghost ? s->n_local + s->n_ghost : s->n_local;
  const int n = r1 - r0;
  const double sub[3] = {s->hi[0] - s->lo[0], s->hi[1] - s->lo[1], s->hi[2] - s->lo[2]};
  const double din[3] = {dx_in, dy_in, dz_in};
  int nb[3];
  double dd[3];
  for (int d = 0; d < 3; d++) {
    nb[d] = (int)(sub[d] / din[d]);
    if (nb[d] == 0) nb[d] = 1;
    dd[d] = sub[d] / nb[d];
  }
  for (int d = 0; d < 3; d++) nb[d] += 2 * halo_depth;
  const double eps = dd[0] / 1000;  // binning_kksort.cpp:92 uses dx for every dimension
  double mn[3], mx[3], mul[3];
  for (int d = 0; d < 3; d++) {
    mn[d] = -dd[d] * halo_depth - eps + s->lo[d];
    mx[d] = dd[d] * halo_depth + eps + s->hi[d];
    mul[d] = 1.0 * nb[d] / (mx[d] - mn[d]);
  }
  b->nbinx = nb[0]; b->nbiny = nb[1]; b->nbinz = nb[2];
  b->minx = mn[0]; b->maxx = mx[0]; b->miny = mn[1]; b->maxy = mx[1];
  b->minz = mn[2]; b->maxz = mx[2];
  const long nbt = (long)nb[0] * nb[1] * nb[2];
  EMD_ARG(ctx, nbt < (1l << 30), "bin grid too large");
  b->nbins = (int)nbt;
  b->n_range = n;
  int cap_b2 = b->cap_bins;
  EMD_TRY(emd_ensure_int(ctx, &b->bincount, &b->cap_bins, nbt));
  EMD_TRY(emd_ensure_int(ctx, &b->binoffsets, &cap_b2, nbt));
  b->cap_bins = b->cap_bins < cap_b2 ? b->cap_bins : cap_b2;
  int cap_r2 = b->cap_range;
  EMD_TRY(emd_ensure_int(ctx, &b->permute, &b->cap_range, n));
  EMD_TRY(emd_ensure_int(ctx, &b->binid, &cap_r2, n));
  b->cap_range = b->cap_range < cap_r2 ? b->cap_range : cap_r2;
  if (ctx->bin_one && n <= BIN_ONE_MAX_N && nbt <= BIN_ONE_MAX_BINS) {
    EMD_LAUNCH(ctx, "binning", k_bin_one, 1, BIN_ONE_THREADS, 0, s->x, r0, n, mn[0], mn[1], mn[2],
               mul[0], mul[1], mul[2], nb[0], nb[1], nb[2], (int)nbt, b->bincount,
               b->binoffsets, b->permute, ctx->d_small);
    return bin_finish(ctx, s, b, r0, n, sort);
  }
  // counting sort: counts, offsets (with cursors in scratch), scatter,
  // per-bin ordering; the largest count and the lost-atom flag in d_small
  EMD_TRY(emd_ensure(ctx, &ctx->d_tmp2, &ctx->tmp2_cap, (size_t)nbt * sizeof(int)));
  int *cursor = static_cast<int *>(ctx->d_tmp2);
  EMD_LAUNCH(ctx, "binning", k_fill_int, 1, 64, 0, ctx->d_small + 1, 2, 0);
  EMD_LAUNCH(ctx, "binning", k_fill_int, emd_blocks(nbt, 256), 256, 0, b->bincount, (int)nbt, 0);
  if (n > 0)
    EMD_LAUNCH(ctx, "binning", k_bin_count, emd_blocks(n, 256), 256, 0, s->x, r0, n, mn[0], mn[1],
               mn[2], mul[0], mul[1], mul[2], nb[0], nb[1], nb[2], b->binid, b->bincount,
               ctx->d_small + 2);
  EMD_TRY(emd_scan(ctx, bin_offsets_op{b->bincount, b->binoffsets, cursor}, (int)nbt,
                   "binning_scan", nullptr));
  if (n > 0)
    EMD_LAUNCH(ctx, "binning", k_bin_scatter, emd_blocks(n, 256), 256, 0, b->binid, n, cursor,
               b->permute);
  EMD_LAUNCH(ctx, "binning_order", k_bin_order,
             emd_blocks(nbt, BIN_ORDER_THREADS / EMD_WAVE), BIN_ORDER_THREADS, 0, b->bincount,
             b->binoffsets, (int)nbt, b->permute);
  {
    const int nbm = emd_blocks(nbt, 256) < 512 ? emd_blocks(nbt, 256) : 512;
    EMD_LAUNCH(ctx, "binning", k_max_int, nbm, 256, 0, b->bincount, (int)nbt, ctx->d_small + 1);
  }
  return bin_finish(ctx, s, b, r0, n, sort);
}

// bins of another binning (the reference's View<int***> bincount and
// binoffsets, LayoutLeft on the GPU: (ix, iy, iz) at ix + nbx (iy + nby iz))
// into this library's x-major arrays
__global__ void __launch_bounds__(256)
k_bins_import(const int *__restrict__ cnt, const int *__restrict__ off, int layout, int nbx, int nby,
              int nbz, int *__restrict__ out_cnt, int *__restrict__ out_off) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nbx * nby * nbz) return;
  const int iz = (int)(t % nbz), iy = (int)((t / nbz) % nby), ix = (int)(t / ((long)nbz * nby));
  const long src = layout == EMD_LAYOUT_LEFT ? ix + (long)nbx * (iy + (long)nby * iz) : t;
  out_cnt[t] = cnt[src];
  out_off[t] = off[src];
}

extern "C" emd_status emd_bins_import(emd_ctx *ctx, int nbinx, int nbiny, int nbinz, int nhalo,
                                      const double min[3], const double max[3],
                                      const int *bincount, const int *binoffsets, int layout,
                                      const int *permute, int n_range, emd_bins *b) {
  if (!ctx || !b || !min || !max || !bincount || !binoffsets || (n_range > 0 && !permute))
    return EMD_ERR_ARG;
  EMD_ARG(ctx, layout == EMD_LAYOUT_RIGHT || layout == EMD_LAYOUT_LEFT, "unknown layout");
  EMD_ARG(ctx, nbinx > 0 && nbiny > 0 && nbinz > 0 && nhalo >= 0 && n_range >= 0, "bad grid");
  const long nbt = (long)nbinx * nbiny * nbinz;
  EMD_ARG(ctx, nbt < (1l << 30), "bin grid too large");
  b->nbinx = nbinx; b->nbiny = nbiny; b->nbinz = nbinz; b->nhalo = nhalo;
  b->minx = min[0]; b->miny = min[1]; b->minz = min[2];
  b->maxx = max[0]; b->maxy = max[1]; b->maxz = max[2];
  b->nbins = (int)nbt;
  b->n_range = n_range;
  int cap_b2 = b->cap_bins;
  EMD_TRY(emd_ensure_int(ctx, &b->bincount, &b->cap_bins, nbt));
  EMD_TRY(emd_ensure_int(ctx, &b->binoffsets, &cap_b2, nbt));
  b->cap_bins = b->cap_bins < cap_b2 ? b->cap_bins : cap_b2;
  int cap_r2 = b->cap_range;
  EMD_TRY(emd_ensure_int(ctx, &b->permute, &b->cap_range, n_range));
  EMD_TRY(emd_ensure_int(ctx, &b->binid, &cap_r2, n_range));
  b->cap_range = b->cap_range < cap_r2 ? b->cap_range : cap_r2;
  EMD_LAUNCH(ctx, "bins_import", k_bins_import, emd_blocks(nbt, 256), 256, 0, bincount,
             binoffsets, layout, nbinx, nbiny, nbinz, b->bincount, b->binoffsets);
  if (n_range > 0)
    EMD_HIP(ctx, hipMemcpyAsync(b->permute, permute, (size_t)n_range * sizeof(int),
                                hipMemcpyDeviceToDevice, ctx->stream));
  EMD_LAUNCH(ctx, "bins_import", k_fill_int, 1, 64, 0, ctx->d_small + 1, 1, 0);
  const int nbm = emd_blocks(nbt, 256) < 512 ? emd_blocks(nbt, 256) : 512;
  EMD_LAUNCH(ctx, "bins_import", k_max_int, nbm, 256, 0, b->bincount, (int)nbt, ctx->d_small + 1);
  EMD_HIP(ctx, hipMemcpyAsync(ctx->h_small + 1, ctx->d_small + 1, sizeof(int),
                              hipMemcpyDeviceToHost, ctx->stream));
  EMD_TRY(emd_wait(ctx));
  b->max_count = ctx->h_small[1];
  return EMD_OK;
}

extern "C" emd_status emd_bins_free(emd_ctx *ctx, emd_bins *b) {
  if (!ctx || !b) return EMD_ERR_ARG;
  hipStreamSynchronize(ctx->stream);
  int *ptrs[] = {b->bincount, b->binoffsets, b->permute, b->binid};
  for (int *p : ptrs)
    if (p) hipFree(p);
  b->bincount = b->binoffsets = b->permute = b->binid = nullptr;
  b->cap_bins = b->cap_range = 0;
  return EMD_OK;
}

// ---------------------------------------------------------------- neighbours
constexpr int NB_THREADS = 256;
constexpr int NB_WAVES = NB_THREADS / EMD_WAVE;
constexpr int NB_MAXC = 640;  // candidates staged per LDS chunk (27 bins of an LJ melt: ~500);
                               // a multiple of 128 (two 64-candidate groups per step)
static_assert(NB_MAXC % (2 * EMD_WAVE) == 0, "staging holds whole iterations");

// NeighborCSR Tag{Count,Fill}Neigh{Full,Half}, src/neighbor_types/neighbor_csr.h:176-358
template <bool HALF, bool FILL>
__global__ void __launch_bounds__(NB_THREADS)
k_neigh(const double *__restrict__ x, const int *__restrict__ bincount,
        const int *__restrict__ binoffsets, const int *__restrict__ permute, neigh_geom g,
        int n_local, double cutsq, int newton, int *__restrict__ num, const int *__restrict__ row,
        int pitch, int *__restrict__ entries, int *__restrict__ cursor, int *__restrict__ maxrow, int cap, int write_num) {
#pragma clang fp contract(off)
  __shared__ double s_x[NB_MAXC], s_y[NB_MAXC], s_z[NB_MAXC];
  __shared__ int s_j[NB_MAXC];
  __shared__ int s_off[27], s_pre[28];

  const int lr = emd_xcd_tile(blockIdx.x, gridDim.x);
  const int bx = lr / (g.niy * g.niz) + g.nhalo;
  const int by = (lr / g.niz) % g.niy + g.nhalo;
  const int bz = lr % g.niz + g.nhalo;
  const int tid = threadIdx.x;
  const int wave = tid / EMD_WAVE, lane = tid % EMD_WAVE;

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
  const int T = s_pre[27];
  const int cb = (bx * g.nby + by) * g.nbz + bz;
  const int c_off = binoffsets[cb], c_cnt = bincount[cb];

  for (int cs = 0; cs < T || (cs == 0 && T == 0); cs += NB_MAXC) {
    const int cn = (T - cs) < NB_MAXC ? (T - cs) : NB_MAXC;
    for (int c = tid; c < cn; c += NB_THREADS) {
      const int cc = cs + c;
      // stencil bin of candidate cc: the last k with s_pre[k] <= cc
      int k = 0, hi = 27;
#pragma unroll
      for (int step = 0; step < 5; step++) {
        const int mid = (k + hi) >> 1;
        if (s_pre[mid] <= cc) k = mid;
        else hi = mid;
      }
      const int j = permute[s_off[k] + (cc - s_pre[k])];
      s_j[c] = j;
      s_x[c] = x[3 * (size_t)j];
      s_y[c] = x[3 * (size_t)j + 1];
      s_z[c] = x[3 * (size_t)j + 2];
    }
    __syncthreads();
    const bool last = (cs + NB_MAXC >= T);
    for (int bi = wave; bi < c_cnt; bi += NB_WAVES) {
      // the centre bin is stencil bin 13: its atoms are staged too, so the
      // row atom's index and position come from LDS when in this round
      const int pc = s_pre[13] + bi - cs;
      int i;
      double x_i, y_i, z_i;
      if (pc >= 0 && pc < cn) {
        i = s_j[pc];
        x_i = s_x[pc];
        y_i = s_y[pc];
        z_i = s_z[pc];
      } else {
        i = permute[c_off + bi];
        x_i = x[3 * (size_t)i];
        y_i = x[3 * (size_t)i + 1];
        z_i = x[3 * (size_t)i + 2];
      }
      if (i >= n_local) continue;
      // the row is the wavefront's: its index and everything derived from it
      // (row start, counters) are scalars
      i = __builtin_amdgcn_readfirstlane(i);
      int run = 0;
      if (cs > 0) run = FILL ? cursor[i] : num[i];
      size_t base = 0;
      if (FILL) base = row ? (size_t)row[i] : (size_t)i * (size_t)pitch;
      int *const rowe = entries + base;
      // candidate test at staged position c (the reference's predicate,
      // neighbor_csr.h:206 full / 290 half), branch-free.  Only the last
      // iteration of a chunk (TAIL) can reach positions past the staged
      // count cn; they read stale LDS (the arrays hold NB_MAXC entries, a
      // multiple of the 128 positions of one iteration) and are masked.
      auto test = [&](auto tail, int c, int j, double x_j, double y_j, double z_j) -> bool {
        const double dx = x_i - x_j, dy = y_i - y_j, dz = z_i - z_j;
        const double rsq = dx * dx + dy * dy + dz * dz;
        const bool staged = !decltype(tail)::value || c < cn;
        if (HALF) {
          const bool ahead =
              (x_j > x_i) | ((x_j == x_i) & ((y_j > y_i) | ((y_j == y_i) & (z_j > z_i))));
          const bool drop = ((j == i) | (j < n_local) | (newton != 0)) & !ahead;
          return staged & !drop & (rsq <= cutsq);
        }
        return staged & (rsq <= cutsq) & (i != j);
      };
      // kept candidates go to the row in candidate order: the lane's slot is
      // `at` plus the kept lanes below it (mbcnt)
      auto put = [&](bool keep, unsigned long long m, int at, int j) {
        if (FILL && keep) {
          const unsigned pos = __builtin_amdgcn_mbcnt_hi(
              (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, (unsigned)at));
          if ((int)pos < cap) rowe[pos] = j;
        }
      };
      // two 64-candidate groups per iteration: both groups' LDS reads are
      // issued before either distance test; entries keep candidate order
      auto iteration = [&](auto tail, int c0) {
        const int ca = c0 + lane, cb = c0 + EMD_WAVE + lane;
        const int ja = s_j[ca], jb = s_j[cb];
        const double xa = s_x[ca], ya = s_y[ca], za = s_z[ca];
        const double xb = s_x[cb], yb = s_y[cb], zb = s_z[cb];
        const bool ka = test(tail, ca, ja, xa, ya, za);
        const bool kb = test(tail, cb, jb, xb, yb, zb);
        const unsigned long long ma = __builtin_amdgcn_ballot_w64(ka);
        const unsigned long long mb = __builtin_amdgcn_ballot_w64(kb);
        put(ka, ma, run, ja);
        const int runb = run + __popcll(ma);
        put(kb, mb, runb, jb);
        run = runb + __popcll(mb);
      };
      int c0 = 0;
      for (; c0 + 2 * EMD_WAVE <= cn; c0 += 2 * EMD_WAVE) iteration(std::false_type{}, c0);
      if (c0 < cn) iteration(std::true_type{}, c0);
      if (lane == 0) {
        if (FILL) {
          if (!last) cursor[i] = run;
          else if (write_num) num[i] = run;
        } else {
          num[i] = run;
        }
      }
    }
    __syncthreads();
    if (T == 0) break;
  }
}

// longest row: block max, one atomic per block (atomics from every atom or
// wavefront on one address serialise at the L2)
__global__ void __launch_bounds__(256) k_max_int(const int *__restrict__ a, int n,
                                                 int *__restrict__ out) {
  __shared__ int s_m[4];
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    m = a[i] > m ? a[i] : m;
#pragma unroll
  for (int o = EMD_WAVE / 2; o > 0; o >>= 1) {
    const int t = __shfl_xor(m, o, EMD_WAVE);
    m = t > m ? t : m;
  }
  if (__lane_id() == 0) s_m[threadIdx.x / EMD_WAVE] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; w++) m = s_m[w] > m ? s_m[w] : m;
    if (m > 0) atomicMax(out, m);
  }
}

// padded rows (pitch P) -> CSR rows, 16 lanes per row
constexpr int COMPACT_LANES = 16;
__global__ void __launch_bounds__(256)
k_compact_rows(const int *__restrict__ tmp, int P, const int *__restrict__ num,
               const int *__restrict__ row, int n, int *__restrict__ entries) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = (int)(t / COMPACT_LANES), lane = (int)(t % COMPACT_LANES);
  if (r >= n) return;
  const int c = num[r];
  const size_t src = (size_t)r * P, dst = (size_t)row[r];
  for (int k = lane; k < c; k += COMPACT_LANES)
    entries[dst + k] = __builtin_nontemporal_load(tmp + src + k);
}

struct row_offsets_op {
  const int *num;
  int *row;
  int n;
  __device__ int value(int i) const { return num[i]; }
  __device__ void emit(int i, int excl, int v) const {
    row[i] = excl;
    if (i == n - 1) row[n] = excl + v;
  }
};

extern "C" emd_status emd_neigh_build(emd_ctx *ctx, const emd_system *s, const emd_bins *b,
                                      double neigh_cut, int half, int newton, int layout,
                                      emd_neigh *nl) {
  if (!ctx || !s || !b || !nl) return EMD_ERR_ARG;
  EMD_ARG(ctx, layout == EMD_NEIGH_CSR || layout == EMD_NEIGH_2D, "bad layout");
  const int n = s->n_local;
  nl->layout = layout;
  nl->n = n;
  neigh_geom g;
  g.nbx = b->nbinx; g.nby = b->nbiny; g.nbz = b->nbinz; g.nhalo = b->nhalo;
  g.nix = b->nbinx - 2 * b->nhalo; g.niy = b->nbiny - 2 * b->nhalo; g.niz = b->nbinz - 2 * b->nhalo;
  const long nib = (long)g.nix * g.niy * g.niz;
  EMD_ARG(ctx, b->nhalo >= 1, "neighbour build needs a halo of at least one bin");
  const double cutsq = neigh_cut * neigh_cut;

  int rcap = nl->cap_rows;
  EMD_TRY(emd_ensure_int(ctx, &nl->num, &nl->cap_rows, n + 1));
  int rcap2 = rcap;
  EMD_TRY(emd_ensure_int(ctx, &nl->row, &rcap2, n + 1));
  nl->cap_rows = nl->cap_rows < rcap2 ? nl->cap_rows : rcap2;
  EMD_TRY(emd_ensure(ctx, &ctx->d_tmp2, &ctx->tmp2_cap, (size_t)(n + 1) * sizeof(int)));
  int *cursor = (int *)ctx->d_tmp2;
  if (n > 0) EMD_LAUNCH(ctx, "neigh_count", k_fill_int, emd_blocks(n, 256), 256, 0, nl->num, n, 0);
  EMD_LAUNCH(ctx, "neigh_count", k_fill_int, 1, 64, 0, ctx->d_small + 4, 1, 0);

  // k_neigh launch: fill?, CSR row offsets (or null: rows at i * pitch),
  // pitch, entries, row capacity, record full row lengths in num?
#define NB_LAUNCH(name, F, ROW, PITCH, ENT, CAP, WNUM)                                        \
  do {                                                                                       \
    if (nib > 0) {                                                                           \
      if (half)                                                                              \
        EMD_LAUNCH(ctx, name, (k_neigh<true, F>), nib, NB_THREADS, 0, s->x, b->bincount,     \
                   b->binoffsets, b->permute, g, n, cutsq, newton, nl->num, ROW, PITCH, ENT, \
                   cursor, ctx->d_small + 4, CAP, WNUM);                                     \
      else                                                                                   \
        EMD_LAUNCH(ctx, name, (k_neigh<false, F>), nib, NB_THREADS, 0, s->x, b->bincount,    \
                   b->binoffsets, b->permute, g, n, cutsq, newton, nl->num, ROW, PITCH, ENT, \
                   cursor, ctx->d_small + 4, CAP, WNUM);                                     \
    }                                                                                        \
  } while (0)

  // One pass when a row capacity is known: the fill writes rows padded to
  // that capacity and records every row's full length.  2D: the rows are the
  // list (pitch maxneighs, grown x1.2 and refilled while a row does not fit,
  // Neighbor2D, neighbor_2d.h:304-330).  CSR: padded rows go to scratch
  // sized from the previous build's longest row, then are compacted; if a
  // row did not fit, the exact CSR fill runs with the counts just made.  The
  // first CSR build counts, scans and fills (NeighborCSR's two passes).
  int P = 0;
  if (layout == EMD_NEIGH_2D) {
    if (nl->maxneighs <= 0) nl->maxneighs = 16;
    P = nl->maxneighs;
  } else if (nl->maxneighs > 0) {
    P = ((nl->maxneighs + nl->maxneighs / 8 + 16) + 15) / 16 * 16;
  }
  int *pad = nullptr;
  auto ensure_entries = [&](long need) -> emd_status {
    if (need > nl->cap_entries || !nl->entries) {
      if (nl->entries) {
        EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        EMD_HIP(ctx, hipFree(nl->entries));
      }
      nl->entries = nullptr;
      const long want = need + need / 8 + 64;
      if (hipMalloc(&nl->entries, (size_t)want * sizeof(int)) != hipSuccess)
        return emd_set_error(ctx, EMD_ERR_ALLOC, "neighbour list: hipMalloc of %ld ints failed",
                             want);
      nl->cap_entries = want;
    }
    return EMD_OK;
  };
  if (P > 0 && n > 0) {
    if (layout == EMD_NEIGH_2D) {
      EMD_TRY(ensure_entries((long)n * P));
      pad = nl->entries;
    } else {
      EMD_TRY(emd_ensure(ctx, &ctx->d_tmp, &ctx->tmp_cap, (size_t)n * P * sizeof(int) + 256));
      pad = (int *)ctx->d_tmp;
    }
    NB_LAUNCH("neigh_fill", true, nullptr, P, pad, P, 1);
  }
  if (P == 0 || n == 0)
    NB_LAUNCH("neigh_count", false, nullptr, 0, nullptr, 0, 0);
  // row offsets and the longest row
  if (n > 0) {
    const int nbm = emd_blocks(n, 256) < 512 ? emd_blocks(n, 256) : 512;
    EMD_LAUNCH(ctx, "neigh_offsets", k_max_int, nbm, 256, 0, nl->num, n, ctx->d_small + 4);
  }
  int total = 0;
  row_offsets_op op{nl->num, nl->row, n};
  if (n > 0) EMD_TRY(emd_scan(ctx, op, n, "neigh_offsets", &total));
  EMD_HIP(ctx, hipMemcpyAsync(ctx->h_small + 4, ctx->d_small + 4, sizeof(int),
                              hipMemcpyDeviceToHost, ctx->stream));
  EMD_TRY(emd_wait(ctx));
  const int maxrow = ctx->h_small[4];
  nl->total = total;
  if (layout == EMD_NEIGH_2D) {
    const bool fits = P > 0 && maxrow <= P;
    if (!fits) {
      while (maxrow > nl->maxneighs) nl->maxneighs = (int)(maxrow * 1.2);
      EMD_TRY(ensure_entries((long)n * nl->maxneighs));
      NB_LAUNCH("neigh_fill", true, nullptr, nl->maxneighs, nl->entries, nl->maxneighs, 0);
    }
  } else {
    EMD_TRY(ensure_entries(total));
    if (P > 0 && maxrow <= P) {
      if (n > 0)
        EMD_LAUNCH(ctx, "neigh_compact", k_compact_rows, emd_blocks((long)n * COMPACT_LANES, 256), 256,
                   0, pad, P, nl->num, nl->row, n, nl->entries);
    } else {
      NB_LAUNCH("neigh_fill", true, nl->row, 0, nl->entries, 0x7fffffff, 0);
    }
    nl->maxneighs = maxrow;
  }
#undef NB_LAUNCH
  return EMD_OK;
}

extern "C" emd_status emd_neigh_free(emd_ctx *ctx, emd_neigh *nl) {
  if (!ctx || !nl) return EMD_ERR_ARG;
  hipStreamSynchronize(ctx->stream);
  if (nl->row) hipFree(nl->row);
  if (nl->num) hipFree(nl->num);
  if (nl->entries) hipFree(nl->entries);
  nl->row = nl->num = nl->entries = nullptr;
  nl->cap_entries = 0;
  nl->cap_rows = 0;
  return EMD_OK;
}
