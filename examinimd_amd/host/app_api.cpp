// app_api.cpp -- extern "C" wrapper of the ExaMiniMD driver (emd_app.h).
#include "emd_app.h"

#include <string>
#include <vector>

#include "examinimd.h"

struct emd_app {
  ExaMiniMD* md;
  std::vector<std::string> args;
  std::vector<char*> argv;
};

extern "C" int emd_app_create(int argc, const char** argv, int quiet, emd_app** out) {
  if (!out || argc < 1 || !argv) return EMD_ERR_ARG;
  emd_app* a = new emd_app();
  for (int i = 0; i < argc; i++) a->args.push_back(argv[i] ? argv[i] : "");
  for (auto& s : a->args) a->argv.push_back(&s[0]);
  a->argv.push_back(nullptr);
  a->md = new ExaMiniMD();
  a->md->system->do_print = quiet == 0;
  *out = a;
  return EMD_OK;
}

extern "C" int emd_app_init(emd_app* a) {
  if (!a) return EMD_ERR_ARG;
  a->md->init((int)a->args.size(), a->argv.data());
  return EMD_OK;
}

extern "C" int emd_app_run(emd_app* a, int nsteps) {
  if (!a || nsteps < 0) return EMD_ERR_ARG;
  a->md->run(nsteps);
  return EMD_OK;
}

extern "C" int emd_app_set_fused_integrate(emd_app* a, int on) {
  if (!a) return EMD_ERR_ARG;
  a->md->fuse_integrate = on != 0;
  return EMD_OK;
}

extern "C" int emd_app_set_phase_sync(emd_app* a, int on) {
  if (!a) return EMD_ERR_ARG;
  a->md->phase_sync = on != 0;
  return EMD_OK;
}

extern "C" int emd_app_destroy(emd_app* a) {
  if (!a) return EMD_OK;
  delete a->md;
  delete a;
  return EMD_OK;
}

extern "C" int emd_app_nsteps(const emd_app* a) { return a ? a->md->input->nsteps : -1; }
extern "C" int emd_app_natoms(const emd_app* a) { return a ? a->md->system->N : -1; }
extern "C" int emd_app_nlocal(const emd_app* a) { return a ? a->md->system->N_local : -1; }
extern "C" int emd_app_nghost(const emd_app* a) { return a ? a->md->system->N_ghost : -1; }

extern "C" int emd_app_thermo(emd_app* a, double out[3]) {
  if (!a || !out) return EMD_ERR_ARG;
  a->md->thermo(out);
  return EMD_OK;
}

extern "C" double emd_app_time(const emd_app* a, int which) {
  if (!a) return -1.0;
  switch (which) {
    case 0: return a->md->time_total;
    case 1: return a->md->time_force;
    case 2: return a->md->time_neigh;
    case 3: return a->md->time_comm;
    default: return a->md->time_other;
  }
}

extern "C" int emd_app_download(emd_app* a, double* x, double* v, double* f, int* type, int* id,
                                double* q) {
  if (!a) return EMD_ERR_ARG;
  System* s = a->md->system;
  const int n = s->N_local;
  const bool mixed = s->precision == EMD_MIXED;
  int st = emd_system_download(s->ctx, &s->d, n, x, v, mixed ? nullptr : f, type, id, q);
  if (st != EMD_OK) return st;
  if (mixed && f) {
    std::vector<float> f32(3 * (size_t)n);
    st = emd_memcpy_d2h(s->ctx, f32.data(), s->d.f32, f32.size() * sizeof(float));
    if (st != EMD_OK) return st;
    for (size_t k = 0; k < f32.size(); k++) f[k] = f32[k];
  }
  return EMD_OK;
}

extern "C" int emd_app_dump(emd_app* a, const char* path) {
  if (!a || !path) return EMD_ERR_ARG;
  System* s = a->md->system;
  const int n = s->N_local;
  std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n), f(3 * (size_t)n), q(n);
  std::vector<int> type(n), id(n);
  int st = emd_app_download(a, x.data(), v.data(), f.data(), type.data(), id.data(), q.data());
  if (st != EMD_OK) return st;
  FILE* fp = fopen(path, "wb");
  if (!fp) return EMD_ERR_ARG;
  fwrite(&n, sizeof(int), 1, fp);
  fwrite(id.data(), sizeof(int), n, fp);
  fwrite(type.data(), sizeof(int), n, fp);
  fwrite(q.data(), sizeof(double), n, fp);
  fwrite(x.data(), sizeof(double), 3 * (size_t)n, fp);
  fwrite(v.data(), sizeof(double), 3 * (size_t)n, fp);
  fwrite(f.data(), sizeof(double), 3 * (size_t)n, fp);
  fclose(fp);
  return EMD_OK;
}

extern "C" int emd_app_dump_neigh(emd_app* a, const char* path) {
  if (!a || !path) return EMD_ERR_ARG;
  return a->md->write_neigh_file(path);
}

extern "C" emd_ctx* emd_app_ctx(emd_app* a) { return a ? a->md->system->ctx : nullptr; }
extern "C" emd_system* emd_app_system(emd_app* a) { return a ? a->md->system->view() : nullptr; }
extern "C" const emd_neigh* emd_app_neigh(emd_app* a) {
  return (a && a->md->neighbor) ? a->md->neighbor->get_neigh_list() : nullptr;
}
extern "C" const emd_halo* emd_app_halo(emd_app* a) {
  if (!a || !a->md->comm) return nullptr;
  CommSerialHIP* c = dynamic_cast<CommSerialHIP*>(a->md->comm);
  return c ? c->halo() : nullptr;
}
extern "C" int emd_app_precision(const emd_app* a) { return a ? a->md->system->precision : -1; }
