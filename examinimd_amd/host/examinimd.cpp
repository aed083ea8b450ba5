// examinimd.cpp -- ExaMiniMD::init / run / dump_binary / check_correctness
// (src/examinimd.cpp:47-488), same step order and output columns.
#include "examinimd.h"

#include <chrono>
#include <cmath>
#include <vector>

#include "property_temperature.h"

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
    printf("Invalid ForceType\n");
    exit(1);
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
    printf("Invalid NeighborType\n");
    exit(1);
  }

  if (false) {
  }
#define COMM_MODULES_INSTANTIATION
#include "modules_comm.h"
#undef COMM_MODULES_INSTANTIATION
  else {
    printf("Invalid CommType\n");
    exit(1);
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
}

// T, PE/N, (PE+KE)/N as the thermo lines print them
void ExaMiniMD::thermo(double out[3]) {
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
  // fused with the next step's initial integration (same arithmetic)
  bool deferred_final = false;
  for (int step = 1; step <= nsteps; step++) {
    other_timer.reset();
    if (deferred_final) integrator->final_initial_integrate();
    else integrator->initial_integrate();
    deferred_final = false;
    fence();
    other_time += other_timer.seconds();

    if (step % input->comm_exchange_rate == 0 && step > 0) {
      comm_timer.reset();
      comm->exchange();
      fence();
      comm_time += comm_timer.seconds();

      other_timer.reset();
      binning->create_binning(neigh_cutoff, neigh_cutoff, neigh_cutoff, 1, true, false, true);
      fence();
      other_time += other_timer.seconds();

      comm_timer.reset();
      comm->exchange_halo();
      fence();
      comm_time += comm_timer.seconds();

      neigh_timer.reset();
      binning->create_binning(neigh_cutoff, neigh_cutoff, neigh_cutoff, 1, true, true, false);
      if (neighbor) neighbor->create_neigh_list(system, binning, force->half_neigh, false);
      fence();
      neigh_time += neigh_timer.seconds();
    } else {
      comm_timer.reset();
      comm->update_halo();
      fence();
      comm_time += comm_timer.seconds();
    }

    force_timer.reset();
    if (!force->stores_forces())
      EMD_CALL(ctx, emd_force_zero(ctx, system->view(), system->precision));
    force->compute(system, binning, neighbor);
    fence();
    force_time += force_timer.seconds();

    if (input->comm_newton) {
      comm_timer.reset();
      comm->update_force();
      fence();
      comm_time += comm_timer.seconds();
    }

    other_timer.reset();
    const bool observed =
        (input->thermo_rate > 0 && step % input->thermo_rate == 0) ||
        (input->dumpbinaryflag && input->dumpbinary_rate > 0 && step % input->dumpbinary_rate == 0) ||
        (input->correctnessflag && input->correctness_rate > 0 &&
         step % input->correctness_rate == 0);
    if (fuse_integrate && step < nsteps && !observed) deferred_final = true;
    else integrator->final_integrate();

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

// dump layout (src/examinimd.cpp:296-346): int n, int id[n], int type[n],
// double q[n], double x[3n], double v[3n], double f[3n]; f is the FP64 force
// (the mixed path's FP32 force widened)
void ExaMiniMD::dump_binary(int step) {
  if (step % input->dumpbinary_rate) return;
  char filename[MAXPATHLEN];
  snprintf(filename, sizeof(filename), "%s%s.%010d.%03d", input->dumpbinary_path, "/output", step,
           comm->process_rank());
  FILE* fp = fopen(filename, "wb");
  if (fp == NULL) {
    char str[MAXPATHLEN + 64];
    snprintf(str, sizeof(str), "Cannot open dump file %s", filename);
    comm->error(str);
  }
  T_INT n = system->N_local;
  std::vector<int> id(n), type(n);
  std::vector<double> q(n), x(3 * (size_t)n), v(3 * (size_t)n), f(3 * (size_t)n);
  emd_ctx* ctx = system->ctx;
  EMD_CALL(ctx, emd_system_download(ctx, &system->d, n, x.data(), v.data(),
                                    system->precision == EMD_MIXED ? nullptr : f.data(),
                                    type.data(), id.data(), q.data()));
  if (system->precision == EMD_MIXED) {
    std::vector<float> f32(3 * (size_t)n);
    EMD_CALL(ctx, emd_memcpy_d2h(ctx, f32.data(), system->d.f32, f32.size() * sizeof(float)));
    for (size_t k = 0; k < f32.size(); k++) f[k] = f32[k];
  }
  fwrite(&n, sizeof(T_INT), 1, fp);
  fwrite(id.data(), sizeof(T_INT), n, fp);
  fwrite(type.data(), sizeof(T_INT), n, fp);
  fwrite(q.data(), sizeof(T_FLOAT), n, fp);
  fwrite(x.data(), sizeof(T_X_FLOAT), 3 * (size_t)n, fp);
  fwrite(v.data(), sizeof(T_V_FLOAT), 3 * (size_t)n, fp);
  fwrite(f.data(), sizeof(T_F_FLOAT), 3 * (size_t)n, fp);
  fclose(fp);
}

// ExaMiniMD::check_correctness, src/examinimd.cpp:355-482
void ExaMiniMD::check_correctness(int step) {
  if (step % input->correctness_rate) return;
  char filename[MAXPATHLEN];
  snprintf(filename, sizeof(filename), "%s%s.%010d.%03d", input->reference_path, "/output", step,
           comm->process_rank());
  FILE* fpref = fopen(filename, "rb");
  if (fpref == NULL) {
    char str[MAXPATHLEN + 64];
    snprintf(str, sizeof(str), "Cannot open input file %s", filename);
    comm->error(str);
  }
  T_INT n = system->N_local, ntmp = 0;
  if (fread(&ntmp, sizeof(T_INT), 1, fpref) != 1 || ntmp != n)
    comm->error("Mismatch in current and reference atom counts");
  std::vector<int> idref(n), typeref(n);
  std::vector<double> qref(n), xref(3 * (size_t)n), vref(3 * (size_t)n), fref(3 * (size_t)n);
  size_t got = fread(idref.data(), sizeof(T_INT), n, fpref) +
               fread(typeref.data(), sizeof(T_INT), n, fpref) +
               fread(qref.data(), sizeof(T_FLOAT), n, fpref) +
               fread(xref.data(), sizeof(T_X_FLOAT), 3 * (size_t)n, fpref) +
               fread(vref.data(), sizeof(T_V_FLOAT), 3 * (size_t)n, fpref) +
               fread(fref.data(), sizeof(T_F_FLOAT), 3 * (size_t)n, fpref);
  fclose(fpref);
  if (got != (size_t)n * 12) comm->error("Short reference dump");

  std::vector<int> id(n), type(n);
  std::vector<double> q(n), x(3 * (size_t)n), v(3 * (size_t)n), f(3 * (size_t)n);
  emd_ctx* ctx = system->ctx;
  EMD_CALL(ctx, emd_system_download(ctx, &system->d, n, x.data(), v.data(),
                                    system->precision == EMD_MIXED ? nullptr : f.data(),
                                    type.data(), id.data(), q.data()));
  if (system->precision == EMD_MIXED) {
    std::vector<float> f32(3 * (size_t)n);
    EMD_CALL(ctx, emd_memcpy_d2h(ctx, f32.data(), system->d.f32, f32.size() * sizeof(float)));
    for (size_t k = 0; k < f32.size(); k++) f[k] = f32[k];
  }
  T_FLOAT sumdelrsq = 0.0, sumdelvsq = 0.0, sumdelfsq = 0.0;
  T_FLOAT maxdelr = 0.0, maxdelv = 0.0, maxdelf = 0.0;
  for (int i = 0; i < n; i++) {
    int ii = -1;
    if (id[i] != idref[i]) {
      for (int j = 0; j < n; j++)
        if (id[j] == idref[i]) { ii = j; break; }
    } else {
      ii = i;
    }
    if (ii == -1) {
      printf("Unable to find current id matchinf reference id %d \n", idref[i]);
      continue;
    }
    const std::vector<double>* cur[3] = {&x, &v, &f};
    const std::vector<double>* ref[3] = {&xref, &vref, &fref};
    T_FLOAT* sums[3] = {&sumdelrsq, &sumdelvsq, &sumdelfsq};
    T_FLOAT* maxs[3] = {&maxdelr, &maxdelv, &maxdelf};
    for (int k = 0; k < 3; k++) {
      double dx = (*cur[k])[3 * ii] - (*ref[k])[3 * i];
      double dy = (*cur[k])[3 * ii + 1] - (*ref[k])[3 * i + 1];
      double dz = (*cur[k])[3 * ii + 2] - (*ref[k])[3 * i + 2];
      *sums[k] += dx * dx + dy * dy + dz * dz;
      *maxs[k] = MAX(fabs(dx), *maxs[k]);
      *maxs[k] = MAX(fabs(dy), *maxs[k]);
      *maxs[k] = MAX(fabs(dz), *maxs[k]);
    }
  }
  comm->reduce_float(&sumdelrsq, 1);
  comm->reduce_float(&sumdelvsq, 1);
  comm->reduce_float(&sumdelfsq, 1);
  comm->reduce_max_float(&maxdelr, 1);
  comm->reduce_max_float(&maxdelv, 1);
  comm->reduce_max_float(&maxdelf, 1);
  if (system->do_print) {
    FILE* fpout = fopen(input->correctness_file, step == 0 ? "w" : "a");
    if (fpout) {
      if (step == 0)
        fprintf(fpout, "# timestep deltarnorm maxdelr deltavnorm maxdelv deltafnorm maxdelf\n");
      fprintf(fpout, "%d %g %g %g %g %g %g\n", step, sqrt(sumdelrsq), maxdelr, sqrt(sumdelvsq),
              maxdelv, sqrt(sumdelfsq), maxdelf);
      fclose(fpout);
    }
  }
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
