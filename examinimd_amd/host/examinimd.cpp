// examinimd.cpp -- ExaMiniMD::init / run / dump_binary / check_correctness
// (src/examinimd.cpp:47-488): the same init and step order, the same thermo,
// PERFORMANCE, dump and correctness-file formats.
#include "examinimd.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "property_temperature.h"
#include "trace.h"

#define MAXPATHLEN 1024

namespace {
struct Timer {
  std::chrono::steady_clock::time_point t0;
  Timer() { reset(); }
  void reset() { t0 = std::chrono::steady_clock::now(); }
  double seconds() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};
// module selection fails before a Comm exists to report it
[[noreturn]] void comm_less_fatal(const char* msg) {
  fprintf(stderr, "ExaMiniMD: %s\n", msg);
  exit(1);
}
}  // namespace

ExaMiniMD::ExaMiniMD() {
  system = new System();
  input = new Input(system);
  integrator = nullptr;
  force = nullptr;
  neighbor = nullptr;
  comm = nullptr;
  binning = nullptr;
  phase_sync = true;
  fuse_integrate = true;
  time_total = time_force = time_neigh = time_comm = time_other = 0.0;
  last_thermo[0] = last_thermo[1] = last_thermo[2] = 0.0;
}

ExaMiniMD::~ExaMiniMD() { shutdown(); }

void ExaMiniMD::init(int argc, char* argv[]) {
  input->read_command_line_args(argc, argv);
  EMD_CALL(system->ctx, emd_create(input->device, &system->ctx));
  system->init();
  input->read_file();
  system->precision = input->precision;
  if (system->do_print) printf("Using device %d, %s, precision %s\n", input->device, emd_version(),
                               input->precision == EMD_MIXED ? "mixed" : "fp64");

  if (input->integrator_type == INTEGRATOR_NVE) integrator = new IntegratorNVEHIP(system);
  if (input->binning_type == BINNING_KKSORT) binning = new BinningHIP(system);

  if (false) {
  }
#define FORCE_MODULES_INSTANTIATION
#include "modules_force.h"
#undef FORCE_MODULES_INSTANTIATION
  else {
    comm_less_fatal("init: no force module matches the deck's pair_style");
  }
  for (size_t line = 0; line < input->force_coeff_lines.size(); line++) {
    int l = input->force_coeff_lines[line];
    force->init_coeff(input->input_data.words_in_line(l), input->input_data.words[l]);
  }

  if (false) {
  }
#define NEIGHBOR_MODULES_INSTANTIATION
#include "modules_neighbor.h"
#undef NEIGHBOR_MODULES_INSTANTIATION
  else {
    comm_less_fatal("init: no neighbour module matches --neigh-type");
  }

  if (false) {
  }
#define COMM_MODULES_INSTANTIATION
#include "modules_comm.h"
#undef COMM_MODULES_INSTANTIATION
  else {
    comm_less_fatal("init: no comm module matches --comm-type");
  }

  force->comm_newton = input->comm_newton;
  if (neighbor) neighbor->comm_newton = input->comm_newton;

  if (system->do_print)
    printf("Using: %s %s %s %s\n", force->name(), neighbor->name(), comm->name(), binning->name());

  if (system->N == 0) input->create_lattice(comm);
  if (input->restart_file) read_restart(input->restart_file);

  comm->exchange();
  T_F_FLOAT neigh_cutoff = input->force_cutoff + input->neighbor_skin;
  binning->create_binning(neigh_cutoff, neigh_cutoff, neigh_cutoff, 1, true, false, true);
  comm->exchange_halo();
  binning->create_binning(neigh_cutoff, neigh_cutoff, neigh_cutoff, 1, true, true, false);
  if (neighbor) neighbor->create_neigh_list(system, binning, force->half_neigh, false);

  EMD_CALL(system->ctx, emd_force_zero(system->ctx, system->view(), system->precision));
  force->compute(system, binning, neighbor);
  if (input->comm_newton) comm->update_force();

  int step = 0;
  if (input->thermo_rate > 0) {
    double th[3];
    thermo(th);
    if (system->do_print) {
      if (!system->print_lammps) {
        printf("\n");
        printf("#Timestep Temperature PotE ETot Time Atomsteps/s\n");
        printf("%i %lf %lf %lf %lf %e\n", step, th[0], th[1], th[2], 0.0, 0.0);
      } else {
        printf("\nStep Temp E_pair TotEng CPU\n");
        printf("     %i %lf %lf %lf %lf\n", step, th[0], th[1], th[2], 0.0);
      }
    }
  }
  if (input->dumpbinaryflag) dump_binary(step);
  if (input->correctnessflag) check_correctness(step);
  dump_neigh(step);
}

// T, PE/N, (PE+KE)/N as the thermo lines print them
void ExaMiniMD::thermo(double out[3]) {
  TraceRange r("thermo");
  Temperature temp(comm);
  PotE pote(comm);
  KinE kine(comm);
  T_FLOAT T = temp.compute(system);
  T_FLOAT PE = pote.compute(system, binning, neighbor, force) / system->N;
  T_FLOAT KE = kine.compute(system) / system->N;
  out[0] = T;
  out[1] = PE;
  out[2] = PE + KE;
  last_thermo[0] = out[0]; last_thermo[1] = out[1]; last_thermo[2] = out[2];
}

void ExaMiniMD::run(int nsteps) {
  T_F_FLOAT neigh_cutoff = input->force_cutoff + input->neighbor_skin;
  double force_time = 0, comm_time = 0, neigh_time = 0, other_time = 0;
  double last_time = 0;
  Timer timer, force_timer, comm_timer, neigh_timer, other_timer;
  emd_ctx* ctx = system->ctx;
  auto fence = [&]() {
    if (phase_sync) EMD_CALL(ctx, emd_synchronize(ctx));
  };

  // the final half-kick of a step whose end nothing observes is deferred and
  // fused with the next step's initial integration (same arithmetic); with
  // newton on, the ghost-force fold before it is deferred too, and the
  // integrator may run fold, kicks, drift and (without a rebuild) the halo
  // position update as one device step boundary (same operations, same order)
  bool deferred_final = false, deferred_fold = false;
  for (int step = 1; step <= nsteps; step++) {
    const bool rebuild = step % input->comm_exchange_rate == 0 && step > 0;
    bool halo_updated = false;
    other_timer.reset();
    {
      TraceRange r("Integrator::initial_integrate");
      if (deferred_final) {
        if (integrator->final_initial_integrate_boundary(comm->device_halo(), deferred_fold,
                                                         !rebuild)) {
          halo_updated = !rebuild;
        } else {
          if (deferred_fold) comm->update_force();
          integrator->final_initial_integrate();
        }
      } else {
        integrator->initial_integrate();
      }
    }
    deferred_final = deferred_fold = false;
    fence();
    other_time += other_timer.seconds();

    if (rebuild) {
      comm_timer.reset();
      {
        TraceRange r("Comm::exchange");
        comm->exchange();
      }
      fence();
      comm_time += comm_timer.seconds();

      other_timer.reset();
      {
        TraceRange r("Binning::create_binning (sort)");
        binning->create_binning(neigh_cutoff, neigh_cutoff, neigh_cutoff, 1, true, false, true);
      }
      fence();
      other_time += other_timer.seconds();

      comm_timer.reset();
      {
        TraceRange r("Comm::exchange_halo");
        comm->exchange_halo();
      }
      fence();
      comm_time += comm_timer.seconds();

      neigh_timer.reset();
      {
        TraceRange r("Neighbor::create_neigh_list");
        binning->create_binning(neigh_cutoff, neigh_cutoff, neigh_cutoff, 1, true, true, false);
        if (neighbor) neighbor->create_neigh_list(system, binning, force->half_neigh, false);
      }
      fence();
      neigh_time += neigh_timer.seconds();
    } else if (!halo_updated) {
      comm_timer.reset();
      {
        TraceRange r("Comm::update_halo");
        comm->update_halo();
      }
      fence();
      comm_time += comm_timer.seconds();
    }

    force_timer.reset();
    {
      TraceRange r("Force::compute");
      if (!force->stores_forces())
        EMD_CALL(ctx, emd_force_zero(ctx, system->view(), system->precision));
      force->compute(system, binning, neighbor);
    }
    fence();
    force_time += force_timer.seconds();

    const bool observed =
        (input->thermo_rate > 0 && step % input->thermo_rate == 0) ||
        (input->dumpbinaryflag && input->dumpbinary_rate > 0 && step % input->dumpbinary_rate == 0) ||
        (input->correctnessflag && input->correctness_rate > 0 &&
         step % input->correctness_rate == 0);
    const bool defer = fuse_integrate && step < nsteps && !observed;
    if (input->comm_newton) {
      if (defer && comm->device_halo()) {
        deferred_fold = true;
      } else {
        comm_timer.reset();
        {
          TraceRange r("Comm::update_force");
          comm->update_force();
        }
        fence();
        comm_time += comm_timer.seconds();
      }
    }

    other_timer.reset();
    if (defer) {
      deferred_final = true;
    } else {
      TraceRange r("Integrator::final_integrate");
      integrator->final_integrate();
    }

    if (input->thermo_rate > 0 && step % input->thermo_rate == 0) {
      double th[3];
      thermo(th);
      if (system->do_print) {
        double time = timer.seconds();
        if (!system->print_lammps)
          printf("%i %lf %lf %lf %lf %e\n", step, th[0], th[1], th[2], time,
                 1.0 * system->N * input->thermo_rate / (time - last_time));
        else
          printf("     %i %lf %lf %lf %lf\n", step, th[0], th[1], th[2], time);
        last_time = time;
      }
    }
    if (input->dumpbinaryflag) dump_binary(step);
    if (input->correctnessflag) check_correctness(step);
    dump_neigh(step);
    fence();
    other_time += other_timer.seconds();
  }
  EMD_CALL(ctx, emd_synchronize(ctx));
  double time = timer.seconds();
  time_total += time;
  time_force += force_time;
  time_neigh += neigh_time;
  time_comm += comm_time;
  time_other += other_time;

  if (system->do_print) {
    double th[3];
    thermo(th);
    if (!system->print_lammps) {
      printf("\n");
      printf("#Procs Particles | Time T_Force T_Neigh T_Comm T_Other | Steps/s Atomsteps/s Atomsteps/(proc*s)\n");
      printf("%i %i | %lf %lf %lf %lf %lf | %lf %e %e PERFORMANCE\n", comm->num_processes(),
             system->N, time, force_time, neigh_time, comm_time, other_time, 1.0 * nsteps / time,
             1.0 * system->N * nsteps / time, 1.0 * system->N * nsteps / time / comm->num_processes());
    } else {
      printf("Loop time of %f on %i procs for %i steps with %i atoms\n", time,
             comm->num_processes(), nsteps, system->N);
    }
  }
}

// initial state from a binary dump (the layout dump_binary writes): the
// deck still defines the box, types and potential; ids, types, charges,
// positions and velocities of the N local atoms come from the file, in its
// order.  The stored force is not used (init recomputes it).
void ExaMiniMD::read_restart(const char* filename) {
  FILE* fp = fopen(filename, "rb");
  if (fp == NULL) {
    char str[MAXPATHLEN + 64];
    snprintf(str, sizeof(str), "Cannot open restart file %s", filename);
    comm->error(str);
  }
  T_INT n = 0;
  if (fread(&n, sizeof(T_INT), 1, fp) != 1 || n != system->N_local)
    comm->error("Restart file atom count does not match the deck");
  std::vector<int> id(n), type(n);
  std::vector<double> q(n), x(3 * (size_t)n), v(3 * (size_t)n);
  size_t got = fread(id.data(), sizeof(T_INT), n, fp) + fread(type.data(), sizeof(T_INT), n, fp) +
               fread(q.data(), sizeof(T_FLOAT), n, fp) +
               fread(x.data(), sizeof(T_X_FLOAT), 3 * (size_t)n, fp) +
               fread(v.data(), sizeof(T_V_FLOAT), 3 * (size_t)n, fp);
  fclose(fp);
  if (got != (size_t)n * 9) comm->error("Short restart file");
  for (int k = 0; k < n; k++)
    if (type[k] < 0 || type[k] >= system->ntypes) comm->error("Restart file has an unknown atom type");
  emd_ctx* ctx = system->ctx;
  EMD_CALL(ctx, emd_system_upload(ctx, &system->d, n, x.data(), v.data(), type.data(), id.data(),
                                  q.data()));
}

// The dump layout of src/examinimd.cpp:296-346: int n, int id[n],
// int type[n], double q[n], double x[3n], double v[3n], double f[3n] (x, v, f
// atom-major), one file DIR/output.<step, 10 digits>.<rank, 3 digits> per
// dump step.  The mixed path's FP32 force is written widened to FP64.
namespace {
struct AtomState {
  std::vector<int> id, type;
  std::vector<double> q, x, v, f;
  explicit AtomState(size_t n) : id(n), type(n), q(n), x(3 * n), v(3 * n), f(3 * n) {}
};

void dump_file_name(char* out, size_t cap, const char* dir, int step, int rank) {
  snprintf(out, cap, "%s/output.%010d.%03d", dir, step, rank);
}
}  // namespace

static AtomState download_state(System* system) {
  const T_INT n = system->N_local;
  AtomState a((size_t)n);
  emd_ctx* ctx = system->ctx;
  const bool mixed = system->precision == EMD_MIXED;
  EMD_CALL(ctx, emd_system_download(ctx, &system->d, n, a.x.data(), a.v.data(),
                                    mixed ? nullptr : a.f.data(), a.type.data(), a.id.data(),
                                    a.q.data()));
  if (mixed) {
    std::vector<float> f32(3 * (size_t)n);
    EMD_CALL(ctx, emd_memcpy_d2h(ctx, f32.data(), system->d.f32, f32.size() * sizeof(float)));
    for (size_t k = 0; k < f32.size(); k++) a.f[k] = f32[k];
  }
  return a;
}

void ExaMiniMD::dump_binary(int step) {
  if (step % input->dumpbinary_rate) return;
  char path[MAXPATHLEN];
  dump_file_name(path, sizeof(path), input->dumpbinary_path, step, comm->process_rank());
  FILE* fp = fopen(path, "wb");
  if (fp == NULL) {
    char msg[MAXPATHLEN + 64];
    snprintf(msg, sizeof(msg), "dump_binary: cannot create %s", path);
    comm->error(msg);
  }
  const T_INT n = system->N_local;
  AtomState a = download_state(system);
  bool ok = fwrite(&n, sizeof(T_INT), 1, fp) == 1;
  ok = ok && fwrite(a.id.data(), sizeof(T_INT), n, fp) == (size_t)n;
  ok = ok && fwrite(a.type.data(), sizeof(T_INT), n, fp) == (size_t)n;
  ok = ok && fwrite(a.q.data(), sizeof(T_FLOAT), n, fp) == (size_t)n;
  for (const std::vector<double>* arr : {&a.x, &a.v, &a.f})
    ok = ok && fwrite(arr->data(), sizeof(double), arr->size(), fp) == arr->size();
  fclose(fp);
  if (!ok) comm->error("dump_binary: short write");
}

// Neighbour-list dump (no counterpart file in the reference, which never
// writes its lists): a canonical, implementation-independent image of the
// current list, so lists of two implementations compare byte for byte.
// File DIR/neigh.<step, 10 digits>.<rank, 3 digits>:
//   int n (local atoms), long total (entries),
//   then per local atom in ascending id order: int id, int count,
//   then count pairs (int id_j, int image_j), ascending by (id_j, image_j).
// image_j is the periodic shift of neighbour j from the local atom carrying
// that id: bits 2d..2d+1 for axis d, 0 none, 1 -L, 2 +L (a local j has 0;
// a ghost has the shifts its exchange_halo phases applied).
int ExaMiniMD::write_neigh_file(const char* path) {
  const emd_neigh* nl = neighbor ? neighbor->get_neigh_list() : nullptr;
  CommSerialHIP* serial = dynamic_cast<CommSerialHIP*>(comm);
  if (!nl || !serial) return EMD_ERR_STATE;
  const emd_halo* h = serial->halo();
  emd_ctx* ctx = system->ctx;
  const int n = nl->n, ng = system->N_ghost;
  std::vector<int> num(n), id(n + ng), root(ng), image(ng);
  std::vector<int> ent;
  std::vector<long> start(n + 1, 0);
  EMD_CALL(ctx, emd_memcpy_d2h(ctx, num.data(), nl->num, sizeof(int) * (size_t)n));
  for (int i = 0; i < n; i++) start[i + 1] = start[i] + num[i];
  if (nl->layout == EMD_NEIGH_CSR) {
    std::vector<int> row(n + 1);
    EMD_CALL(ctx, emd_memcpy_d2h(ctx, row.data(), nl->row, sizeof(int) * (size_t)(n + 1)));
    std::vector<int> all(row[n]);
    EMD_CALL(ctx, emd_memcpy_d2h(ctx, all.data(), nl->entries, sizeof(int) * all.size()));
    ent.resize(start[n]);
    for (int i = 0; i < n; i++)
      std::copy(all.begin() + row[i], all.begin() + row[i] + num[i], ent.begin() + start[i]);
  } else {
    const size_t pitch = nl->maxneighs;
    std::vector<int> pad(pitch * n);
    EMD_CALL(ctx, emd_memcpy_d2h(ctx, pad.data(), nl->entries, sizeof(int) * pad.size()));
    ent.resize(start[n]);
    for (int i = 0; i < n; i++)
      std::copy(pad.begin() + pitch * i, pad.begin() + pitch * i + num[i], ent.begin() + start[i]);
  }
  EMD_CALL(ctx, emd_memcpy_d2h(ctx, id.data(), system->d.id, sizeof(int) * id.size()));
  if (ng > 0) {
    EMD_CALL(ctx, emd_memcpy_d2h(ctx, root.data(), h->root, sizeof(int) * (size_t)ng));
    EMD_CALL(ctx, emd_memcpy_d2h(ctx, image.data(), h->image, sizeof(int) * (size_t)ng));
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return id[a] < id[b]; });
  FILE* fp = fopen(path, "wb");
  if (!fp) return EMD_ERR_ARG;
  const long total = start[n];
  bool ok = fwrite(&n, sizeof(int), 1, fp) == 1 && fwrite(&total, sizeof(long), 1, fp) == 1;
  std::vector<std::pair<int, int>> keys;
  for (int i : order) {
    keys.clear();
    for (long e = start[i]; e < start[i + 1]; e++) {
      const int j = ent[e];
      keys.emplace_back(j < n ? id[j] : id[root[j - n]], j < n ? 0 : image[j - n]);
    }
    std::sort(keys.begin(), keys.end());
    const int head[2] = {id[i], num[i]};
    ok = ok && fwrite(head, sizeof(int), 2, fp) == 2;
    for (const auto& k : keys) {
      const int pair[2] = {k.first, k.second};
      ok = ok && fwrite(pair, sizeof(int), 2, fp) == 2;
    }
  }
  fclose(fp);
  return ok ? EMD_OK : EMD_ERR_ARG;
}

void ExaMiniMD::dump_neigh(int step) {
  if (input->dumpneigh_rate <= 0 || step % input->dumpneigh_rate) return;
  char path[MAXPATHLEN];
  snprintf(path, sizeof(path), "%s/neigh.%010d.%03d", input->dumpneigh_path, step,
           comm->process_rank());
  if (write_neigh_file(path) != EMD_OK) {
    char msg[MAXPATHLEN + 64];
    snprintf(msg, sizeof(msg), "dumpneigh: cannot write %s", path);
    comm->error(msg);
  }
}

// The --correctness comparison of src/examinimd.cpp:355-482: the reference
// dump of this step is matched to the current atoms by id; per quantity
// (x, v, f) the root of the summed squared differences and the largest
// component difference are appended to the correctness file as one line
// "step dr maxdr dv maxdv df maxdf" (%g), under a header written at step 0.
// Atoms are matched through an id -> index table (the reference searches
// linearly when the orders differ); a reference id with no current atom is
// reported and skipped.
void ExaMiniMD::check_correctness(int step) {
  if (step % input->correctness_rate) return;
  char path[MAXPATHLEN];
  dump_file_name(path, sizeof(path), input->reference_path, step, comm->process_rank());
  FILE* fp = fopen(path, "rb");
  if (fp == NULL) {
    char msg[MAXPATHLEN + 64];
    snprintf(msg, sizeof(msg), "correctness: no reference dump %s", path);
    comm->error(msg);
  }
  const T_INT n = system->N_local;
  T_INT nref = -1;
  if (fread(&nref, sizeof(T_INT), 1, fp) != 1 || nref != n) {
    fclose(fp);
    comm->error("correctness: the reference dump holds a different number of atoms");
  }
  AtomState ref((size_t)n);
  bool ok = fread(ref.id.data(), sizeof(T_INT), n, fp) == (size_t)n;
  ok = ok && fread(ref.type.data(), sizeof(T_INT), n, fp) == (size_t)n;
  ok = ok && fread(ref.q.data(), sizeof(T_FLOAT), n, fp) == (size_t)n;
  for (std::vector<double>* arr : {&ref.x, &ref.v, &ref.f})
    ok = ok && fread(arr->data(), sizeof(double), arr->size(), fp) == arr->size();
  fclose(fp);
  if (!ok) comm->error("correctness: the reference dump is truncated");

  AtomState cur = download_state(system);
  int id_max = 0;
  for (int id : cur.id) id_max = MAX(id_max, id);
  std::vector<int> where(id_max + 1, -1);
  for (T_INT k = 0; k < n; k++)
    if (cur.id[k] >= 0) where[cur.id[k]] = k;

  double sq[3] = {0.0, 0.0, 0.0}, big[3] = {0.0, 0.0, 0.0};
  const std::vector<double>* now[3] = {&cur.x, &cur.v, &cur.f};
  const std::vector<double>* was[3] = {&ref.x, &ref.v, &ref.f};
  for (T_INT i = 0; i < n; i++) {
    const int id = ref.id[i];
    const int k = (id >= 0 && id <= id_max) ? where[id] : -1;
    if (k < 0) {
      printf("correctness: reference atom id %d is not among the current atoms\n", id);
      continue;
    }
    for (int q = 0; q < 3; q++) {
      double d2 = 0.0;
      for (int c = 0; c < 3; c++) {
        const double d = (*now[q])[3 * (size_t)k + c] - (*was[q])[3 * (size_t)i + c];
        d2 += d * d;
        big[q] = MAX(fabs(d), big[q]);
      }
      sq[q] += d2;
    }
  }
  comm->reduce_float(sq, 3);
  comm->reduce_max_float(big, 3);
  // the file is written by rank 0 whether or not the run prints
  if (comm->process_rank() != 0) return;
  FILE* out = fopen(input->correctness_file, step == 0 ? "w" : "a");
  if (out == NULL) comm->error("correctness: cannot write the correctness file");
  if (step == 0) fprintf(out, "# timestep deltarnorm maxdelr deltavnorm maxdelv deltafnorm maxdelf\n");
  fprintf(out, "%d %g %g %g %g %g %g\n", step, sqrt(sq[0]), big[0], sqrt(sq[1]), big[1],
          sqrt(sq[2]), big[2]);
  fclose(out);
}

void ExaMiniMD::print_performance() {}

void ExaMiniMD::shutdown() {
  delete force; force = nullptr;
  delete neighbor; neighbor = nullptr;
  delete comm; comm = nullptr;
  delete binning; binning = nullptr;
  delete integrator; integrator = nullptr;
  if (system) {
    system->destroy();
    if (system->ctx) emd_destroy(system->ctx);
    system->ctx = nullptr;
    delete system;
    system = nullptr;
  }
  delete input;
  input = nullptr;
}
