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

// the per-type-pair coefficients as the kernels read them: four device
// arrays, entry (ti, tj) at ti * stride + tj.  Either the context's fixed
// 16 x 16 table or, for more types (emd_lj_params.table), its own
// ntypes x ntypes table.
struct lj_tv {
  const double *lj1, *lj2, *cutsq, *rep;
  int stride;
};
constexpr int LJ_MAX_TABLE_TYPES = EMD_MAX_TYPES;

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
          double cutsq_0, double rep_0, lj_tv tab, int store,
          double *__restrict__ f, float *__restrict__ f32, int prio) {
#pragma clang fp contract(off)
  typedef typename lj_acc<MIXED>::T T;
  const int tile = emd_xcd_tile(blockIdx.x, gridDim.x);
  const int gid = tile * LJ_THREADS + threadIdx.x;
  const int i = gid / K, sub = gid % K;
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
      const int t = type_i * tab.stride + type[j];
      cutsq = tab.cutsq[t]; lj1 = tab.lj1[t]; lj2 = tab.lj2[t];
      if (DIAL) nrep = tab.rep[t];
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
          lj_tv tab, double *__restrict__ f, float *__restrict__ f32) {
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
      const int t = type_i * tab.stride + type[j[u]];
      if (ok[u] && rsq < tab.cutsq[t]) {
        T vx, vy, vz;
        if (MIXED) {
          float fpair = emd_lj_fpair_mixed(rsq, (float)tab.lj1[t], (float)tab.lj2[t]);
          if (DIAL) fpair = dial_sum(fpair / (float)tab.rep[t], dial_trips(tab.rep[t]));
          vx = (float)dx * fpair;
          vy = (float)dy * fpair;
          vz = (float)dz * fpair;
        } else {
          const double r2inv = 1.0 / rsq;
          const double r6inv = r2inv * r2inv * r2inv;
          double fpair = (r6inv * (tab.lj1[t] * r6inv - tab.lj2[t])) * r2inv;
          if (DIAL) fpair = dial_sum(fpair / tab.rep[t], dial_trips(tab.rep[t]));
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

static void fill_table(lj_table &h, const emd_lj_params *p, const double (*rep)[16]) {
  memset(&h, 0, sizeof(h));
  for (int a = 0; a < p->ntypes; a++)
    for (int b = 0; b < p->ntypes; b++) {
      h.lj1[a * LJ_MAXT + b] = p->lj1[a][b];
      h.lj2[a * LJ_MAXT + b] = p->lj2[a][b];
      h.cutsq[a * LJ_MAXT + b] = p->cutsq[a][b];
      h.rep[a * LJ_MAXT + b] = rep ? rep[a][b] : 1.0;
    }
}

// The context's device table, uploaded only when the parameters differ from
// the last upload (the pinned staging copy holds those): a force or energy
// call with unchanged parameters neither copies nor waits.  A change waits
// until the previous upload has read the staging copy, then queues the copy
// on the stream (kernels queued before it still see the old table).
static emd_status upload_table(emd_ctx *ctx, const emd_lj_params *p,
                               const double (*rep)[16], lj_table **dtab) {
  static_assert(sizeof(lj_table) <= EMD_LJ_TABLE_BYTES, "table size");
  lj_table h;
  fill_table(h, p, rep);
  lj_table *staged = (lj_table *)ctx->h_ljtab;
  if (!ctx->ljtab_valid || memcmp(staged, &h, sizeof(h)) != 0) {
    if (ctx->ljtab_valid) EMD_HIP(ctx, hipEventSynchronize(ctx->ljtab_ev));
    memcpy(staged, &h, sizeof(h));
    EMD_HIP(ctx, hipMemcpyAsync(ctx->d_ljtab, staged, sizeof(h), hipMemcpyHostToDevice,
                                ctx->stream));
    EMD_HIP(ctx, hipEventRecord(ctx->ljtab_ev, ctx->stream));
    ctx->ljtab_valid = true;
  }
  *dtab = (lj_table *)ctx->d_ljtab;
  return EMD_OK;
}

// the view of a 16 x 16 table at device address t
static lj_tv view_of(const lj_table *t) {
  constexpr size_t N = (size_t)LJ_MAXT * LJ_MAXT;
  const double *b = reinterpret_cast<const double *>(t);
  return lj_tv{b, b + N, b + 2 * N, b + 3 * N, LJ_MAXT};
}

// the coefficients of p as a device view: the cached 16 x 16 table, or with
// p->table (more types, the reference's View path, force_lj_neigh_impl.h:
// 100-126 with ntypes above MAX_TYPES_STACKPARAMS) the context's own
// ntypes x ntypes table, uploaded when its contents change (after the
// stream has drained: kernels queued before still read the old one)
static emd_status upload_tv(emd_ctx *ctx, const emd_lj_params *p, const double (*rep)[16],
                            lj_tv *out) {
  if (!p->table) {
    lj_table *t = nullptr;
    EMD_TRY(upload_table(ctx, p, rep, &t));
    *out = view_of(t);
    return EMD_OK;
  }
  const int nt = p->ntypes;
  const size_t N = (size_t)nt * nt;
  std::vector<double> h(4 * N, 1.0);  // lj1, lj2, cutsq, rep (1: lj/cut)
  std::copy(p->table, p->table + 3 * N, h.begin());
  if (!ctx->ljbig_valid || ctx->h_ljbig != h) {
    EMD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->ljbig_cap < h.size()) {
      if (ctx->d_ljbig) EMD_HIP(ctx, hipFree(ctx->d_ljbig));
      ctx->d_ljbig = nullptr;
      ctx->ljbig_cap = 0;
      if (hipMalloc(&ctx->d_ljbig, h.size() * sizeof(double)) != hipSuccess)
        return emd_set_error(ctx, EMD_ERR_ALLOC, "LJ type table: hipMalloc failed");
      ctx->ljbig_cap = h.size();
    }
    EMD_HIP(ctx, hipMemcpy(ctx->d_ljbig, h.data(), h.size() * sizeof(double),
                           hipMemcpyHostToDevice));
    ctx->h_ljbig.swap(h);
    ctx->ljbig_valid = true;
  }
  const double *d = ctx->d_ljbig;
  *out = lj_tv{d, d + N, d + 2 * N, d + 3 * N, nt};
  return EMD_OK;
}

// ntypes in range: up to 16 in the fixed arrays, more with a table
static emd_status check_types(emd_ctx *ctx, const emd_system *s, const emd_lj_params *p) {
  EMD_ARG(ctx, p->ntypes >= 1 && p->ntypes <= LJ_MAX_TABLE_TYPES, "ntypes out of range");
  EMD_ARG(ctx, p->ntypes <= LJ_MAXT || p->table != nullptr,
          "more than 16 atom types need emd_lj_params.table");
  EMD_ARG(ctx, s->ntypes <= p->ntypes, "the system has more atom types than the parameters");
  return EMD_OK;
}

// lanes per atom and gathers per round (measured on MI355X, LJ 4M atoms)
static int default_lanes(int precision) { return precision == EMD_MIXED ? 16 : 16; }

extern "C" emd_status emd_lj_set_lanes(emd_ctx *ctx, int precision, int lanes) {
  if (!ctx || (precision != EMD_FP64 && precision != EMD_MIXED)) return EMD_ERR_ARG;
  if (!(lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8 || lanes == 16))
    return emd_set_error(ctx, EMD_ERR_ARG, "lanes must be 0,1,2,4,8 or 16");
  ctx->lj_lanes[precision] = lanes;
  return EMD_OK;
}

template <int K, int U, bool DIAL>
static emd_status launch_full_as(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                 const emd_lj_params *p, double rep_0, int precision,
                                 const lj_tv &tab, int store) {
  const int n = s->n_local;
  const int nb = emd_blocks((long)n * K, LJ_THREADS);
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
  const bool one = p->ntypes == 1 && !p->table;
  const int prio = ctx->lj_prio >= 0 ? ctx->lj_prio : (precision == EMD_MIXED ? 2 : 0);
  const char *name_m = DIAL ? "lj_idial_mixed" : "lj_force_mixed";
  const char *name_d = DIAL ? "lj_idial_fp64" : "lj_force_fp64";
#define LJ_FULL_ARGS                                                                        \
  s->x, s->type, row, nl->num, nl->entries, nl->maxneighs, n, p->lj1[0][0], p->lj2[0][0],   \
      p->cutsq[0][0], rep_0, tab, store, s->f, s->f32, prio
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

// the full-list kernel (lj/cut, or lj/cut/idial with DIAL) at K lanes per
// atom; U gathers per lane and round (LJ-4M sweep: 16 x 5 and 8 x 5 best)
template <bool DIAL>
static emd_status launch_full_lanes(emd_ctx *ctx, emd_system *s, const emd_neigh *nl,
                                    const emd_lj_params *p, double rep_0, int precision,
                                    const lj_tv &tab, int store, int K) {
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
  EMD_ARG(ctx, lj->ntypes >= 1 && lj->ntypes <= LJ_MAXT && !lj->table,
          "lj/cut/idial: 1 to 16 atom types, in the fixed arrays");
  EMD_ARG(ctx, s->ntypes <= lj->ntypes, "the system has more atom types than the parameters");
  EMD_ARG(ctx, nl->n == s->n_local, "neighbour list does not match the system");
  EMD_ARG(ctx, !(store && half), "storing forces needs a full list");
  for (int a = 0; a < lj->ntypes; a++)
    for (int b = 0; b < lj->ntypes; b++)
      EMD_ARG(ctx, p->intensity[a][b] <= 1e6, "lj/cut/idial intensity above 1e6");
  if (s->n_local <= 0) return EMD_OK;
  lj_tv tab{};
  if (lj->ntypes > 1 || half) EMD_TRY(upload_tv(ctx, lj, p->intensity, &tab));
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
  EMD_TRY(check_types(ctx, s, p));
  EMD_ARG(ctx, nl->n == s->n_local, "neighbour list does not match the system");
  EMD_ARG(ctx, !(store && half), "storing forces needs a full list");
  if (s->n_local <= 0) return EMD_OK;
  lj_tv tab{};
  if (p->ntypes > 1 || half || p->table) EMD_TRY(upload_tv(ctx, p, nullptr, &tab));
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
  // 1 lane per atom reproduces the reference's serial sum bit for bit
  const int K = ctx->lj_lanes[precision == EMD_MIXED ? 1 : 0];
  return launch_full_lanes<false>(ctx, s, nl, p, 1.0, precision, tab, store,
                                  K ? K : default_lanes(precision));
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

// PER_ATOM: each atom's energy (its LJE_K lanes' sum) into eatom[i] instead
// of block sums into part
template <bool ONE_TYPE, bool PER_ATOM = false>
__global__ void __launch_bounds__(LJ_THREADS)
k_lj_energy(const double *__restrict__ x, const int *__restrict__ type,
            const int *__restrict__ row, const int *__restrict__ num,
            const int *__restrict__ entries, int pitch, int n, int half,
            lj_tv tab, double *__restrict__ part) {
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
      const double r2invc = 1.0 / tab.cutsq[0];
      shift0 = lj_pair_e(r2invc * r2invc * r2invc, tab.lj1[0], tab.lj2[0]);
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
        const int t = ONE_TYPE ? 0 : type_i * tab.stride + type[j[u]];
        const double cutsq = tab.cutsq[t], lj1 = tab.lj1[t], lj2 = tab.lj2[t];
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
  if (PER_ATOM) {
    pe = group_sum<LJE_K>(pe);
    if (i < n && sub == 0) part[i] = pe;
  } else {
    block_dd_sum(pe, 0.0, part);
  }
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
  EMD_TRY(check_types(ctx, s, p));
  const int n = s->n_local;
  if (n <= 0) { *pe = 0.0; return EMD_OK; }
  lj_tv tab;
  EMD_TRY(upload_tv(ctx, p, nullptr, &tab));
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

// ------------------------------------------- energy without allocation
// emd_lj_energy_noalloc: the same energy as emd_lj_energy with every scratch
// byte in one file-scope device array (`scratch`), so the call performs no
// device or host allocation.  Layout (doubles): the LJ table, the final
// (hi, lo) pair, then one (hi, lo) pair per energy block.  Each GPU holds its
// own copy for the life of the process; its device address comes from
// hipGetSymbolAddress once per device (emd_lj_scratch_setup, run by
// emd_create with the context's device current).  Calls on one GPU share the
// array: overlapping calls on one GPU (two streams, two threads) race.
constexpr int NOALLOC_MAX_ATOMS = 1 << 22;  // 4,194,304 local atoms
constexpr int NOALLOC_TAB = EMD_LJ_TABLE_BYTES / sizeof(double);
constexpr int NOALLOC_OUT = NOALLOC_TAB;               // 2 doubles (+ padding)
constexpr int NOALLOC_PART = NOALLOC_OUT + 32;
constexpr int NOALLOC_MAX_BLOCKS = NOALLOC_MAX_ATOMS / (LJ_THREADS / LJE_K);
constexpr size_t NOALLOC_DOUBLES = NOALLOC_PART + 2 * (size_t)NOALLOC_MAX_BLOCKS;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];  // 4.01 MiB per GPU

emd_status emd_lj_scratch_setup(emd_ctx *ctx) {
  void *ptr = nullptr;
  EMD_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  ctx->noalloc_scratch = (double *)ptr;
  return EMD_OK;
}

// second stage: one block adds the nb block pairs (hi, lo) in double-double
// arithmetic and writes the total pair to out[0], out[1]
__global__ void __launch_bounds__(256) k_dd_total(const double *__restrict__ part, int nb,
                                                  double *__restrict__ out) {
#pragma clang fp contract(off)
  double hi = 0.0, lo = 0.0;
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    double s, e;
    two_sum(hi, part[2 * (size_t)k], s, e);
    hi = s;
    lo += part[2 * (size_t)k + 1] + e;
  }
  block_dd_sum(hi, lo, out);
}

extern "C" emd_status emd_lj_energy_noalloc(emd_ctx *ctx, const emd_system *s,
                                            const emd_neigh *nl, const emd_lj_params *p,
                                            int half, double *pe) {
  if (!ctx || !s || !nl || !p || !pe) return EMD_ERR_ARG;
  EMD_ARG(ctx, p->ntypes >= 1 && p->ntypes <= LJ_MAXT && !p->table,
          "emd_lj_energy_noalloc: 1 to 16 atom types, in the fixed arrays");
  EMD_ARG(ctx, s->ntypes <= p->ntypes, "the system has more atom types than the parameters");
  const int n = s->n_local;
  if (n > NOALLOC_MAX_ATOMS)
    return emd_set_error(ctx, EMD_ERR_ARG,
                         "emd_lj_energy_noalloc: %d local atoms exceed its limit of %d", n,
                         NOALLOC_MAX_ATOMS);
  EMD_ARG(ctx, ctx->noalloc_scratch != nullptr, "scratch address not resolved");
  if (n <= 0) { *pe = 0.0; return EMD_OK; }
  double *buf = ctx->noalloc_scratch;  // device address of this GPU's `scratch`
  lj_table h;
  fill_table(h, p, nullptr);
  EMD_HIP(ctx, hipMemcpyAsync(buf, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
  const lj_tv tab = view_of((const lj_table *)buf);
  double *out = buf + NOALLOC_OUT, *part = buf + NOALLOC_PART;
  const int nb = emd_blocks((long)n * LJE_K, LJ_THREADS);
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
  if (p->ntypes == 1)
    EMD_LAUNCH(ctx, "lj_energy", k_lj_energy<true>, nb, LJ_THREADS, 0, s->x, s->type, row,
               nl->num, nl->entries, nl->maxneighs, n, half, tab, part);
  else
    EMD_LAUNCH(ctx, "lj_energy", k_lj_energy<false>, nb, LJ_THREADS, 0, s->x, s->type, row,
               nl->num, nl->entries, nl->maxneighs, n, half, tab, part);
  EMD_LAUNCH(ctx, "lj_energy_total", k_dd_total, 1, 256, 0, part, nb, out);
  EMD_HIP(ctx, hipMemcpyAsync(ctx->h_dsmall, out, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  // the stack copy of the table and the pinned result are read by the
  // stream: wait before returning
  EMD_TRY(emd_wait(ctx));
  *pe = ctx->h_dsmall[0] + ctx->h_dsmall[1];
  return EMD_OK;
}

extern "C" emd_status emd_lj_energy_atoms(emd_ctx *ctx, const emd_system *s,
                                          const emd_neigh *nl, const emd_lj_params *p, int half,
                                          double *eatom) {
  if (!ctx || !s || !nl || !p || !eatom) return EMD_ERR_ARG;
  EMD_TRY(check_types(ctx, s, p));
  const int n = s->n_local;
  if (n <= 0) return EMD_OK;
  lj_tv tab;
  EMD_TRY(upload_tv(ctx, p, nullptr, &tab));
  const int nb = emd_blocks((long)n * LJE_K, LJ_THREADS);
  const int *row = nl->layout == EMD_NEIGH_CSR ? nl->row : nullptr;
  if (p->ntypes == 1)
    EMD_LAUNCH(ctx, "lj_energy", (k_lj_energy<true, true>), nb, LJ_THREADS, 0, s->x, s->type,
               row, nl->num, nl->entries, nl->maxneighs, n, half, tab, eatom);
  else
    EMD_LAUNCH(ctx, "lj_energy", (k_lj_energy<false, true>), nb, LJ_THREADS, 0, s->x, s->type,
               row, nl->num, nl->entries, nl->maxneighs, n, half, tab, eatom);
  return EMD_OK;
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
             const double *__restrict__ wE, double *__restrict__ out) {
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
    const double we = wE ? wE[i] : 1.0;  // the energy cotangent of atom i
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
        a[2] += we * (double)((T)2 * ((s12 * r6inv * r6inv - s6 * r6inv) -
                                      (s12 * rc6 * rc6 - s6 * rc6)));
        a[3] += we * (double)((T)2 * e * (((T)12 * s11 * r6inv * r6inv - (T)6 * s5 * r6inv) -
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

static emd_status lj_adjoint(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl, double eps,
                             double sigma, double cut, const double *dLdF, double dLdE,
                             const double *dLdE_atom, int precision, double *out) {
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
               nl->num, nl->entries, nl->maxneighs, n, eps, sigma, cut * cut, dLdF, dLdE_atom,
               acc);
  else
    EMD_LAUNCH(ctx, "lj_adjoint", k_lj_adjoint<double>, emd_blocks(n, 256), 256, 0, s->x, row,
               nl->num, nl->entries, nl->maxneighs, n, eps, sigma, cut * cut, dLdF, dLdE_atom,
               acc);
  EMD_HIP(ctx, hipMemcpyAsync(ctx->h_dsmall, acc, 4 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  EMD_TRY(emd_wait(ctx));
  // per-atom weights are inside the energy sums; one weight multiplies them
  const double w = dLdE_atom ? 1.0 : dLdE;
  out[0] = ctx->h_dsmall[0] + w * ctx->h_dsmall[2];
  out[1] = ctx->h_dsmall[1] + w * ctx->h_dsmall[3];
  return EMD_OK;
}

extern "C" emd_status emd_lj_adjoint(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                     double eps, double sigma, double cut, const double *dLdF,
                                     double dLdE, int precision, double *out) {
  return lj_adjoint(ctx, s, nl, eps, sigma, cut, dLdF, dLdE, nullptr, precision, out);
}

extern "C" emd_status emd_lj_adjoint_atoms(emd_ctx *ctx, const emd_system *s, const emd_neigh *nl,
                                           double eps, double sigma, double cut,
                                           const double *dLdF, const double *dLdE_atom,
                                           int precision, double *out) {
  if (!dLdE_atom) return EMD_ERR_ARG;
  return lj_adjoint(ctx, s, nl, eps, sigma, cut, dLdF, 0.0, dLdE_atom, precision, out);
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
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  emd_nve_final_initial_at<MIXED>(i, x, v, f, f32, type, mass, dtf, dtv);
}

// The step boundary of a small system in one workgroup: CommSerial's
// update_force of step s (comm_serial.h:200-213; src/examinimd.cpp:243),
// IntegratorNVE's final_integrate of step s and initial_integrate of step
// s+1 (src/examinimd.cpp:249 then 196) and, when s+1 rebuilds no list,
// CommSerial's update_halo of the new positions (comm_serial.h:184-197;
// src/examinimd.cpp:226).  Workgroup barriers separate the three parts, so
// every atom sees the same operations in the same order as the separate
// launches (k_halo_fold_one, k_nve_final_initial, k_halo_update): the same
// bits, in one launch instead of three.
template <bool MIXED>
__global__ void __launch_bounds__(FOLD_ONE_THREADS)
k_step_boundary_one(double *__restrict__ x, double *__restrict__ v, double *__restrict__ f,
                    float *__restrict__ f32, const int *__restrict__ type,
                    const double *__restrict__ mass, int n, double dtf, double dtv,
                    fold_phases ph, const int *__restrict__ root, const int *__restrict__ image,
                    int n_halo, double lx, double ly, double lz) {
  if (MIXED) emd_fold_phases_one<float, FOLD_ONE_THREADS>(f32, ph);
  else emd_fold_phases_one<double, FOLD_ONE_THREADS>(f, ph);
  for (int i = threadIdx.x; i < n; i += FOLD_ONE_THREADS)
    emd_nve_final_initial_at<MIXED>(i, x, v, f, f32, type, mass, dtf, dtv);
  if (n_halo <= 0) return;
  __syncthreads();
  for (int g = threadIdx.x; g < n_halo; g += FOLD_ONE_THREADS)
    emd_halo_update_at(g, x, root, image, n, lx, ly, lz);
}

constexpr int STEP_ONE_MAX = 16384;  // local atoms of a one-workgroup step boundary

extern "C" emd_status emd_nve_step_boundary(emd_ctx *ctx, emd_system *s, const emd_halo *h,
                                            int fold, int update_halo, double dtf, double dtv,
                                            int precision) {
  if (!ctx || !s || !h) return EMD_ERR_ARG;
  const int n = s->n_local;
  fold_phases ph;
  long ghosts = emd_fold_phases_of(s, h, ph);
  if (!fold)
    for (int p = 0; p < 6; p++) ph.num[p] = 0;
  const int n_halo = update_halo ? h->n_ghost : 0;
  if (!ctx->fold_one || n > STEP_ONE_MAX || ghosts > FOLD_ONE_MAX || n_halo > FOLD_ONE_MAX) {
    // larger systems: the separate launches (their grids fill the GPU)
    if (fold) EMD_TRY(emd_comm_update_force(ctx, s, h, precision));
    EMD_TRY(emd_nve_final_initial(ctx, s, dtf, dtv, precision));
    if (update_halo) EMD_TRY(emd_comm_update_halo(ctx, s, h));
    return EMD_OK;
  }
  if (n <= 0 && n_halo <= 0 && (!fold || ghosts == 0)) return EMD_OK;
  if (precision == EMD_MIXED)
    EMD_LAUNCH(ctx, "step_boundary", k_step_boundary_one<true>, 1, FOLD_ONE_THREADS, 0, s->x,
               s->v, s->f, s->f32, s->type, s->mass, n, dtf, dtv, ph, h->root, h->image, n_halo,
               s->domain[0], s->domain[1], s->domain[2]);
  else
    EMD_LAUNCH(ctx, "step_boundary", k_step_boundary_one<false>, 1, FOLD_ONE_THREADS, 0, s->x,
               s->v, s->f, s->f32, s->type, s->mass, n, dtf, dtv, ph, h->root, h->image, n_halo,
               s->domain[0], s->domain[1], s->domain[2]);
  return EMD_OK;
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
