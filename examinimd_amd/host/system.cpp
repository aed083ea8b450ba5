// system.cpp -- System (src/system.cpp:43-133).
#include "system.h"

#include <vector>

void emd_host_fatal(const char* what, emd_ctx* ctx, int status) {
  fprintf(stderr, "ExaMiniMD error %d in %s: %s\n", status, what,
          ctx ? emd_last_error(ctx) : "no context");
  exit(1);
}

System::System() {
  N = 0; N_max = 0; N_local = 0; N_ghost = 0;
  ntypes = 1;
  mass.assign(1, 0.0);
  ctx = nullptr;
  memset(&d, 0, sizeof(d));
  precision = EMD_FP64;
  domain_x = domain_y = domain_z = 0.0;
  sub_domain_x = sub_domain_y = sub_domain_z = 0.0;
  sub_domain_hi_x = sub_domain_hi_y = sub_domain_hi_z = 0.0;
  sub_domain_lo_x = sub_domain_lo_y = sub_domain_lo_z = 0.0;
  mvv2e = boltz = dt = 0.0;
  do_print = true;
  print_lammps = false;
}

void System::init() {
  if (!ctx) EMD_CALL(ctx, emd_create(0, &ctx));
  EMD_CALL(ctx, emd_system_grow(ctx, &d, N_max > 0 ? N_max : 1));
  N_max = d.n_max;
}

void System::destroy() {
  if (ctx) {
    emd_system_free(ctx, &d);
  }
  N_max = 0; N_local = 0; N_ghost = 0; ntypes = 1;
}

void System::grow(T_INT N_new) {
  if (N_new > N_max) {
    EMD_CALL(ctx, emd_system_grow(ctx, &d, N_new));
    N_max = d.n_max;
  }
}

emd_system* System::view() {
  d.n_local = N_local;
  d.n_ghost = N_ghost;
  d.n_global = N;
  d.ntypes = ntypes;
  d.domain[0] = domain_x; d.domain[1] = domain_y; d.domain[2] = domain_z;
  d.lo[0] = sub_domain_lo_x; d.lo[1] = sub_domain_lo_y; d.lo[2] = sub_domain_lo_z;
  d.hi[0] = sub_domain_hi_x; d.hi[1] = sub_domain_hi_y; d.hi[2] = sub_domain_hi_z;
  d.boltz = boltz; d.mvv2e = mvv2e; d.dt = dt;
  return &d;
}

void System::pull(const emd_system* v) {
  N_ghost = v->n_ghost;
  N_max = v->n_max;
}

void System::upload_mass() {
  std::vector<double> m(ntypes, 0.0);
  for (int t = 0; t < ntypes && t < (int)mass.size(); t++) m[t] = mass[t];
  EMD_CALL(ctx, emd_system_set_mass(ctx, &d, ntypes, m.data()));
}

void System::print_particles() {
  int n = N_local + N_ghost;
  std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), f(3 * (size_t)n), q(n);
  std::vector<int> type(n), id(n);
  EMD_CALL(ctx, emd_system_download(ctx, &d, n, x.data(), v.data(), f.data(), type.data(), id.data(), q.data()));
  printf("Print all particles: \n");
  printf("  Owned: %d\n", N_local);
  for (T_INT i = 0; i < n; i++) {
    if (i == N_local) printf("  Ghost: %d\n", N_ghost);
    printf("    %d %lf %lf %lf | %lf %lf %lf | %lf %lf %lf | %d %e\n", i, x[3*i], x[3*i+1], x[3*i+2],
           v[3*i], v[3*i+1], v[3*i+2], f[3*i], f[3*i+1], f[3*i+2], type[i], q[i]);
  }
}
