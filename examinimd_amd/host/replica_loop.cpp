// replica_loop.cpp -- an ensemble of periodic LJ boxes run one replica after
// another through the per-system C ABI: the launch pattern a user of the
// reference's module pipeline (src/examinimd.cpp:176-294) gets for many
// small systems, kept as the reference point for the batched path.
#include <chrono>
#include <cmath>
#include <vector>

#include "emd_app.h"

namespace {

struct replica {
  emd_system s{};
  emd_halo h{};
  emd_bins b{};
  emd_neigh nl{};
  int n = 0;
};

int rebuild(emd_ctx *ctx, replica &r, double cut) {
  int st;
  if ((st = emd_comm_exchange(ctx, &r.s))) return st;
  if ((st = emd_binning(ctx, &r.s, cut, cut, cut, 1, 1, 0, 1, &r.b))) return st;
  if ((st = emd_comm_exchange_halo(ctx, &r.s, cut, &r.h))) return st;
  if ((st = emd_binning(ctx, &r.s, cut, cut, cut, 1, 1, 1, 0, &r.b))) return st;
  return emd_neigh_build(ctx, &r.s, &r.b, cut, 0, 0, EMD_NEIGH_CSR, &r.nl);
}

int force(emd_ctx *ctx, replica &r, const emd_lj_params &p, int precision) {
  int st;
  if ((st = emd_force_zero(ctx, &r.s, precision))) return st;
  return emd_lj_force(ctx, &r.s, &r.nl, &p, precision, 0);
}

void release(emd_ctx *ctx, std::vector<replica> &reps) {
  for (auto &r : reps) {
    emd_neigh_free(ctx, &r.nl);
    emd_bins_free(ctx, &r.b);
    emd_halo_free(ctx, &r.h);
    emd_system_free(ctx, &r.s);
  }
}

}  // namespace

extern "C" int emd_replica_loop_run(emd_ctx *ctx, const emd_batch_desc *d, const double *x,
                                    const double *v, int nsteps, double *seconds, double *x_out,
                                    double *v_out) {
  if (!ctx || !d || !x || !v || d->nrep < 1 || !d->natoms || !d->box || !d->mass)
    return EMD_ERR_ARG;
  const double cut = d->cut + d->skin;
  emd_lj_params p{};
  p.ntypes = 1;
  p.lj1[0][0] = 48.0 * d->eps * std::pow(d->sigma, 12.0);
  p.lj2[0][0] = 24.0 * d->eps * std::pow(d->sigma, 6.0);
  p.cutsq[0][0] = d->cut * d->cut;
  const double dtf = 0.5 * d->dt / 1.0, dtv = d->dt;
  std::vector<replica> reps(d->nrep);
  int st = 0;
  size_t off = 0;
  for (int k = 0; k < d->nrep && !st; k++) {
    replica &r = reps[k];
    r.n = d->natoms[k];
    std::vector<int> type(r.n, 0), id(r.n);
    std::vector<double> q(r.n, 0.0);
    for (int i = 0; i < r.n; i++) id[i] = i + 1;
    st = emd_system_upload(ctx, &r.s, r.n, x + 3 * off, v + 3 * off, type.data(), id.data(),
                           q.data());
    if (!st) st = emd_system_set_mass(ctx, &r.s, 1, &d->mass[k]);
    r.s.n_local = r.n;
    r.s.n_ghost = 0;
    r.s.n_global = r.n;
    r.s.ntypes = 1;
    for (int a = 0; a < 3; a++) {
      r.s.domain[a] = d->box[3 * k + a];
      r.s.lo[a] = 0.0;
      r.s.hi[a] = d->box[3 * k + a];
    }
    if (!st) st = rebuild(ctx, r, cut);
    if (!st) st = force(ctx, r, p, d->precision);
    off += r.n;
  }
  if (!st) st = emd_synchronize(ctx);
  const auto t0 = std::chrono::steady_clock::now();
  for (int step = 1; step <= nsteps && !st; step++) {
    for (auto &r : reps) {
      if ((st = emd_nve_initial(ctx, &r.s, dtf, dtv, d->precision))) break;
      if (step % d->rebuild_every == 0) st = rebuild(ctx, r, cut);
      else st = emd_comm_update_halo(ctx, &r.s, &r.h);
      if (st) break;
      if ((st = force(ctx, r, p, d->precision))) break;
      if ((st = emd_nve_final(ctx, &r.s, dtf, d->precision))) break;
    }
  }
  if (!st) st = emd_synchronize(ctx);
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  off = 0;
  for (auto &r : reps) {
    if (st) break;
    if (x_out || v_out) {
      std::vector<double> xs(3 * (size_t)r.n), vs(3 * (size_t)r.n);
      std::vector<int> id(r.n);
      st = emd_system_download(ctx, &r.s, r.n, xs.data(), vs.data(), nullptr, nullptr, id.data(),
                               nullptr);
      for (int i = 0; i < r.n && !st; i++)
        for (int a = 0; a < 3; a++) {
          const size_t dst = 3 * (off + id[i] - 1) + a;
          if (x_out) x_out[dst] = xs[3 * (size_t)i + a];
          if (v_out) v_out[dst] = vs[3 * (size_t)i + a];
        }
    }
    off += r.n;
  }
  release(ctx, reps);
  return st;
}
