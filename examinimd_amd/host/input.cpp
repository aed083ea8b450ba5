// input.cpp -- deck parsing, command line and lattice creation
// (src/input.cpp:44-792).
#include "input.h"

#include <cmath>
#include <fstream>
#include <iostream>

#include "binning.h"
#include "force.h"
#include "integrator.h"
#include "neighbor.h"

// ------------------------------------------------------------ ItemizedFile
ItemizedFile::ItemizedFile() : words(nullptr), nlines(0) {}
ItemizedFile::~ItemizedFile() {
  for (auto& l : lines)
    for (char* w : l) free(w);
}

// split on blanks/tabs, at most words_per_line words (src/input.cpp:100-117)
void ItemizedFile::add_line(const char* line) {
  std::vector<char*> ws;
  const char* pos = line;
  while (*pos && (int)ws.size() < words_per_line) {
    while ((*pos == ' ' || *pos == '\t') && *pos) pos++;
    std::string w;
    while (*pos != ' ' && *pos != '\t' && *pos && *pos != '\r' && *pos != '\n') w += *pos++;
    if (*pos == '\r' || *pos == '\n') pos++;
    ws.push_back(strdup(w.c_str()));
  }
  while ((int)ws.size() < words_per_line) ws.push_back(strdup(""));
  lines.push_back(ws);
  nlines = (int)lines.size();
}

void ItemizedFile::finalize() {
  line_ptrs.resize(lines.size());
  for (size_t i = 0; i < lines.size(); i++) line_ptrs[i] = lines[i].data();
  words = line_ptrs.data();
}

int ItemizedFile::words_in_line(int i) {
  int c = 0;
  for (int j = 0; j < words_per_line; j++)
    if (words[i][j][0]) c++;
  return c;
}

void ItemizedFile::print_line(int i) {
  for (int j = 0; j < words_per_line; j++)
    if (words[i][j][0]) std::cout << words[i][j] << " ";
  std::cout << std::endl;
}

void ItemizedFile::print() {
  for (int l = 0; l < nlines; l++) print_line(l);
}

// ------------------------------------------------------------------ RNG
double LAMMPS_RandomVelocityGeom::uniform() {
  const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
  int k = seed / IQ;
  seed = IA * (seed - k * IQ) - IR * k;
  if (seed < 0) seed += IM;
  return (1.0 / IM) * seed;
}

double LAMMPS_RandomVelocityGeom::gaussian() {
  double v1, v2, rsq;
  do {
    v1 = 2.0 * uniform() - 1.0;
    v2 = 2.0 * uniform() - 1.0;
    rsq = v1 * v1 + v2 * v2;
  } while ((rsq >= 1.0) || (rsq == 0.0));
  const double fac = sqrt(-2.0 * log(rsq) / rsq);
  return v2 * fac;
}

// Jenkins one-at-a-time over the (signed) bytes of the seed and the three
// coordinates, 27-bit seed, five warm-up draws
void LAMMPS_RandomVelocityGeom::reset(int ibase, double* coord) {
  unsigned int hash = 0;
  const signed char* str = (const signed char*)&ibase;
  for (int i = 0; i < (int)sizeof(int); i++) {
    hash += (unsigned int)(int)str[i];
    hash += (hash << 10);
    hash ^= (hash >> 6);
  }
  str = (const signed char*)coord;
  for (int i = 0; i < 3 * (int)sizeof(double); i++) {
    hash += (unsigned int)(int)str[i];
    hash += (hash << 10);
    hash ^= (hash >> 6);
  }
  hash += (hash << 3);
  hash ^= (hash >> 11);
  hash += (hash << 15);
  seed = hash & 0x7ffffff;
  if (!seed) seed = 1;
  for (int i = 0; i < 5; i++) uniform();
}

// ---------------------------------------------------------------- Input
Input::Input(System* p) : timestepflag(false), system(p) {
  input_file = nullptr;
  input_file_type = -1;
  units = UNITS_LJ;
  lattice_style = LATTICE_FCC;
  lattice_constant = 1.0;
  lattice_offset_x = lattice_offset_y = lattice_offset_z = 0.0;
  lattice_nx = lattice_ny = lattice_nz = 0;
  temperature_target = 0.0;
  temperature_seed = 0;
  velocity_gaussian = false;
  integrator_type = INTEGRATOR_NVE;
  nsteps = 0;
  comm_type = COMM_SERIAL;
  neighbor_type = NEIGH_2D;
  force_iteration_type = FORCE_ITER_NEIGH_FULL;
  binning_type = BINNING_KKSORT;
  comm_exchange_rate = 20;
  force_type = FORCE_LJ;
  force_line = 0;
  force_cutoff = 0.0;
  neighbor_skin = 0.0;
  thermo_rate = 0;
  dumpbinary_rate = 0;
  correctness_rate = 0;
  dumpbinaryflag = false;
  correctnessflag = false;
  dumpbinary_path = reference_path = correctness_file = nullptr;
  restart_file = nullptr;
  comm_newton = 0;
  precision = EMD_FP64;
  device = 0;
}

// Input::read_command_line_args, src/input.cpp:151-224
void Input::read_command_line_args(int argc, char* argv[]) {
#define MODULES_OPTION_CHECK
  for (int i = 1; i < argc; i++) {
    if ((strcmp(argv[i], "-h") == 0) || (strcmp(argv[i], "--help") == 0)) {
      if (system->do_print) {
        printf("ExaMiniMD 1.0 (MI355X / HIP version)\n\n");
        printf("Options:\n");
        printf("  -il [file] / --input-lammps [FILE]: Provide LAMMPS input file\n");
        printf("  --force-iteration [TYPE]:   Specify which iteration style to use\n");
        printf("                              for force calculations (NEIGH_FULL, NEIGH_HALF)\n");
        printf("  --comm-type [TYPE]:         Specify Communication Routines implementation \n");
        printf("                              (SERIAL)\n");
        printf("  --dumpbinary [N] [PATH]:    Request that binary output files PATH/output* be generated every N steps\n");
        printf("  --correctness [N] [PATH] [FILE]:   Request that correctness check against files PATH/output* be performed every N steps, correctness data written to FILE\n");
        printf("  --neigh-type [TYPE]:        Specify Neighbor Routines implementation \n");
        printf("                              (2D, CSR)\n");
        printf("  --precision [fp64|mixed]:   force arithmetic (mixed: FP32 pair math and\n");
        printf("                              accumulation, FP64 state)\n");
        printf("  --device [N]:               GPU ordinal\n");
        printf("  --restart [FILE]:           start from a binary dump FILE (the --dumpbinary\n");
        printf("                              format) instead of the deck's lattice\n");
      }
    } else if ((strcmp(argv[i], "-il") == 0) || (strcmp(argv[i], "--input-lammps") == 0)) {
      input_file = argv[++i];
      input_file_type = INPUT_LAMMPS;
    } else if ((strcmp(argv[i], "--force-iteration") == 0)) {
#include "modules_force.h"
      ++i;
    } else if ((strcmp(argv[i], "--comm-type") == 0)) {
#include "modules_comm.h"
      ++i;
    } else if ((strcmp(argv[i], "--neigh-type") == 0)) {
#include "modules_neighbor.h"
      ++i;
    } else if ((strcmp(argv[i], "--dumpbinary") == 0)) {
      dumpbinary_rate = atoi(argv[i + 1]);
      dumpbinary_path = argv[i + 2];
      dumpbinaryflag = true;
      i += 2;
    } else if ((strcmp(argv[i], "--correctness") == 0)) {
      correctness_rate = atoi(argv[i + 1]);
      reference_path = argv[i + 2];
      correctness_file = argv[i + 3];
      correctnessflag = true;
      i += 3;
    } else if ((strcmp(argv[i], "--precision") == 0)) {
      const char* p = argv[++i];
      if (strcmp(p, "mixed") == 0) precision = EMD_MIXED;
      else if (strcmp(p, "fp64") == 0) precision = EMD_FP64;
      else {
        printf("ERROR: --precision takes fp64 or mixed\n");
        exit(1);
      }
    } else if ((strcmp(argv[i], "--device") == 0)) {
      device = atoi(argv[++i]);
    } else if ((strcmp(argv[i], "--restart") == 0)) {
      restart_file = argv[++i];
    } else if ((strstr(argv[i], "--kokkos-") == NULL)) {
      if (system->do_print) printf("ERROR: Unknown command line argument: %s\n", argv[i]);
      exit(1);
    }
  }
#undef MODULES_OPTION_CHECK
}

void Input::read_file(const char* filename) {
  if (filename == NULL) filename = input_file;
  if (input_file_type == INPUT_LAMMPS) {
    read_lammps_file(filename);
    return;
  }
  if (system->do_print) printf("ERROR: Unknown input file type\n");
  exit(1);
}

void Input::read_lammps_file(const char* filename) {
  std::ifstream file(filename);
  if (!file.good()) {
    printf("ERROR: cannot open input file %s\n", filename ? filename : "(null)");
    exit(1);
  }
  std::string line;
  while (std::getline(file, line)) input_data.add_line(line.c_str());
  input_data.finalize();
  if (system->do_print) {
    printf("\n#InputFile:\n#=========================================================\n");
    input_data.print();
    printf("#=========================================================\n\n");
  }
  for (int l = 0; l < input_data.nlines; l++) check_lammps_command(l);
}

// Input::check_lammps_command, src/input.cpp:265-458
void Input::check_lammps_command(int line) {
  bool known = false;
  char** w = input_data.words[line];
  if (w[0][0] == 0) known = true;
  if (strstr(w[0], "#")) known = true;
  if (strcmp(w[0], "variable") == 0) {
    if (system->do_print) printf("LAMMPS-Command: 'variable' keyword is not supported in ExaMiniMD\n");
  }
  if (strcmp(w[0], "units") == 0) {
    if (strcmp(w[1], "metal") == 0) {
      known = true;
      units = UNITS_METAL;
      system->boltz = 8.617343e-5;
      system->mvv2e = 1.0364269e-4;
      system->dt = 0.001;
    } else if (strcmp(w[1], "real") == 0) {
      known = true;
      units = UNITS_REAL;
      system->boltz = 0.0019872067;
      system->mvv2e = 48.88821291 * 48.88821291;
      if (!timestepflag) system->dt = 1.0;
    } else if (strcmp(w[1], "lj") == 0) {
      known = true;
      units = UNITS_LJ;
      system->boltz = 1.0;
      system->mvv2e = 1.0;
      if (!timestepflag) system->dt = 0.005;
    } else if (system->do_print) {
      printf("LAMMPS-Command: 'units' command only supports 'real' and 'lj' in ExaMiniMD\n");
    }
  }
  if (strcmp(w[0], "atom_style") == 0) {
    if (strcmp(w[1], "atomic") == 0) known = true;
    else if (system->do_print)
      printf("LAMMPS-Command: 'atom_style' command only supports 'atomic' in ExaMiniMD\n");
  }
  if (strcmp(w[0], "lattice") == 0) {
    if (strcmp(w[1], "sc") == 0) {
      known = true;
      lattice_style = LATTICE_SC;
      lattice_constant = atof(w[2]);
    } else if (strcmp(w[1], "fcc") == 0) {
      known = true;
      lattice_style = LATTICE_FCC;
      lattice_constant = std::pow((4.0 / atof(w[2])), (1.0 / 3.0));
    } else if (strcmp(w[1], "bcc") == 0) {
      known = true;
      lattice_style = LATTICE_BCC;
      lattice_constant = atof(w[2]);
    } else if (system->do_print) {
      printf("LAMMPS-Command: 'lattice' command only supports 'sc', 'fcc' and 'bcc' in ExaMiniMD\n");
    }
    if (strcmp(w[3], "origin") == 0) {
      lattice_offset_x = atof(w[4]);
      lattice_offset_y = atof(w[5]);
      lattice_offset_z = atof(w[6]);
    }
  }
  if (strcmp(w[0], "region") == 0) {
    if (strcmp(w[2], "block") == 0) {
      known = true;
      int box[6];
      for (int k = 0; k < 6; k++) box[k] = atoi(w[3 + k]);
      if ((box[0] != 0) || (box[2] != 0) || (box[4] != 0))
        if (system->do_print)
          printf("Error: LAMMPS-Command: region only allows for boxes with 0,0,0 offset\n");
      lattice_nx = box[1];
      lattice_ny = box[3];
      lattice_nz = box[5];
    } else if (system->do_print) {
      printf("LAMMPS-Command: 'region' command only supports 'block' option in ExaMiniMD\n");
    }
  }
  if (strcmp(w[0], "create_box") == 0) {
    known = true;
    system->ntypes = atoi(w[1]);
    system->mass.assign(system->ntypes, 0.0);
  }
  if (strcmp(w[0], "create_atoms") == 0) known = true;
  if (strcmp(w[0], "mass") == 0) {
    known = true;
    int type = atoi(w[1]) - 1;
    if (type >= 0 && type < (int)system->mass.size()) system->mass[type] = atof(w[2]);
  }
  if (strcmp(w[0], "pair_style") == 0) {
    if (strcmp(w[1], "lj/cut/idial") == 0) {
      known = true;
      force_type = FORCE_LJ_IDIAL;
      force_cutoff = atof(w[2]);
      force_line = line;
    } else if (strcmp(w[1], "lj/cut") == 0) {
      known = true;
      force_type = FORCE_LJ;
      force_cutoff = atof(w[2]);
      force_line = line;
    }
    if (strcmp(w[1], "snap") == 0) {
      known = true;
      force_type = FORCE_SNAP;
      force_cutoff = 4.73442;  // fixed in the reference (src/input.cpp:378-383)
      force_line = line;
    }
    if (system->do_print && !known)
      printf("LAMMPS-Command: 'pair_style' command only supports 'lj/cut', 'lj/cut/idial', and 'snap' style in ExaMiniMD\n");
  }
  if (strcmp(w[0], "pair_coeff") == 0) {
    known = true;
    force_coeff_lines.push_back(line);
  }
  if (strcmp(w[0], "velocity") == 0) {
    known = true;
    if (strcmp(w[1], "all") != 0 && system->do_print)
      printf("Error: LAMMPS-Command: 'velocity' command can only be applied to 'all'\n");
    if (strcmp(w[2], "create") != 0 && system->do_print)
      printf("Error: LAMMPS-Command: 'velocity' command can only be used with option 'create'\n");
    temperature_target = atof(w[3]);
    temperature_seed = atoi(w[4]);
    for (int k = 5; k < ItemizedFile::words_per_line - 1; k++)
      if (strcmp(w[k], "dist") == 0) velocity_gaussian = strcmp(w[k + 1], "gaussian") == 0;
  }
  if (strcmp(w[0], "neighbor") == 0) {
    known = true;
    neighbor_skin = atof(w[1]);
  }
  if (strcmp(w[0], "neigh_modify") == 0) {
    known = true;
    for (int i = 1; i < ItemizedFile::words_per_line - 1; i++)
      if (strcmp(w[i], "every") == 0) comm_exchange_rate = atoi(w[i + 1]);
  }
  if (strcmp(w[0], "fix") == 0) {
    if (strcmp(w[3], "nve") == 0) {
      known = true;
      integrator_type = INTEGRATOR_NVE;
    } else if (system->do_print) {
      printf("LAMMPS-Command: 'fix' command only supports 'nve' style in ExaMiniMD\n");
    }
  }
  if (strcmp(w[0], "run") == 0) {
    known = true;
    nsteps = atoi(w[1]);
  }
  if (strcmp(w[0], "thermo") == 0) {
    known = true;
    thermo_rate = atoi(w[1]);
  }
  if (strcmp(w[0], "timestep") == 0) {
    known = true;
    system->dt = atof(w[1]);
    timestepflag = true;
  }
  if (strcmp(w[0], "newton") == 0) {
    known = true;
    if (strcmp(w[1], "on") == 0) comm_newton = 1;
    else if (strcmp(w[1], "off") == 0) comm_newton = 0;
    else if (system->do_print) printf("LAMMPS-Command: 'newton' must be followed by 'on' or 'off'\n");
  }
  if (w[0][0] == '#') known = true;
  if (!known && system->do_print) {
    printf("ERROR: unknown keyword\n");
    input_data.print_line(line);
  }
}

// Input::create_lattice, src/input.cpp:460-792 (one process: the sub-domain
// is the box).  Positions and velocities are made on the host, in the same
// loop order and arithmetic as the reference, then uploaded.
void Input::create_lattice(Comm* comm) {
  System& s = *system;
  s.domain_x = lattice_constant * lattice_nx;
  s.domain_y = lattice_constant * lattice_ny;
  s.domain_z = lattice_constant * lattice_nz;
  comm->create_domain_decomposition();

  double basis[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  int nb = 1;
  if (lattice_style == LATTICE_FCC) {
    nb = 4;
    double b[4][3] = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
    memcpy(basis, b, sizeof(b));
  } else if (lattice_style == LATTICE_BCC) {
    nb = 2;
    double b[2][3] = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.5}};
    memcpy(basis, b, sizeof(b));
  }
  if (lattice_style != LATTICE_SC)
    for (int i = 0; i < nb; i++) {
      basis[i][0] += lattice_offset_x;
      basis[i][1] += lattice_offset_y;
      basis[i][2] += lattice_offset_z;
    }
  const double a = lattice_constant;
  T_INT ix_start = s.sub_domain_lo_x / s.domain_x * lattice_nx - 0.5;
  T_INT iy_start = s.sub_domain_lo_y / s.domain_y * lattice_ny - 0.5;
  T_INT iz_start = s.sub_domain_lo_z / s.domain_z * lattice_nz - 0.5;
  T_INT ix_end = s.sub_domain_hi_x / s.domain_x * lattice_nx + 0.5;
  T_INT iy_end = s.sub_domain_hi_y / s.domain_y * lattice_ny + 0.5;
  T_INT iz_end = s.sub_domain_hi_z / s.domain_z * lattice_nz + 0.5;

  std::vector<double> hx;
  std::vector<int> htype, hid;
  for (T_INT iz = iz_start; iz <= iz_end; iz++)
    for (T_INT iy = iy_start; iy <= iy_end; iy++)
      for (T_INT ix = ix_start; ix <= ix_end; ix++)
        for (int k = 0; k < nb; k++) {
          double xt, yt, zt;
          if (lattice_style == LATTICE_SC) {
            zt = a * (iz + lattice_offset_z);
            yt = a * (iy + lattice_offset_y);
            xt = a * (ix + lattice_offset_x);
          } else {
            xt = a * (1.0 * ix + basis[k][0]);
            yt = a * (1.0 * iy + basis[k][1]);
            zt = a * (1.0 * iz + basis[k][2]);
          }
          if ((xt >= s.sub_domain_lo_x) && (yt >= s.sub_domain_lo_y) &&
              (zt >= s.sub_domain_lo_z) && (xt < s.sub_domain_hi_x) &&
              (yt < s.sub_domain_hi_y) && (zt < s.sub_domain_hi_z)) {
            hx.push_back(xt);
            hx.push_back(yt);
            hx.push_back(zt);
            htype.push_back(rand() % s.ntypes);
            hid.push_back((int)hid.size() + 1);
          }
        }
  const T_INT n = (T_INT)hid.size();
  s.N_local = n;
  s.N = n;
  s.N_ghost = 0;
  T_INT N_local_offset = n;
  comm->scan_int(&N_local_offset, 1);
  for (T_INT i = 0; i < n; i++) hid[i] += N_local_offset - n;
  comm->reduce_int(&s.N, 1);
  if (s.do_print) printf("Atoms: %i %i\n", s.N, s.N_local);

  // velocities: src/input.cpp:731-791
  std::vector<double> hv(3 * (size_t)n), hq(n, 0.0);
  double total_mass = 0.0, px = 0.0, py = 0.0, pz = 0.0;
  for (T_INT i = 0; i < n; i++) {
    LAMMPS_RandomVelocityGeom random;
    double xx[3] = {hx[3 * i], hx[3 * i + 1], hx[3 * i + 2]};
    random.reset(temperature_seed, xx);
    double mass_i = s.mass[htype[i]];
    double vx, vy, vz;
    if (!velocity_gaussian) {
      vx = random.uniform() - 0.5;
      vy = random.uniform() - 0.5;
      vz = random.uniform() - 0.5;
    } else {
      vx = random.gaussian();
      vy = random.gaussian();
      vz = random.gaussian();
    }
    hv[3 * i] = vx / sqrt(mass_i);
    hv[3 * i + 1] = vy / sqrt(mass_i);
    hv[3 * i + 2] = vz / sqrt(mass_i);
    total_mass += mass_i;
    px += mass_i * hv[3 * i];
    py += mass_i * hv[3 * i + 1];
    pz += mass_i * hv[3 * i + 2];
  }
  comm->reduce_float(&px, 1);
  comm->reduce_float(&py, 1);
  comm->reduce_float(&pz, 1);
  comm->reduce_float(&total_mass, 1);
  const double svx = px / total_mass, svy = py / total_mass, svz = pz / total_mass;
  for (T_INT i = 0; i < n; i++) {
    hv[3 * i] -= svx;
    hv[3 * i + 1] -= svy;
    hv[3 * i + 2] -= svz;
  }
  // Temperature::compute in the serial order (the scale factor pins every
  // trajectory, so it is summed on the host exactly as the reference does)
  double T = 0.0;
  for (T_INT i = 0; i < n; i++)
    T += (hv[3 * i] * hv[3 * i] + hv[3 * i + 1] * hv[3 * i + 1] + hv[3 * i + 2] * hv[3 * i + 2]) *
         s.mass[htype[i]];
  comm->reduce_float(&T, 1);
  T_INT dof = 3 * s.N - 3;
  T *= s.mvv2e / (1.0 * dof * s.boltz);
  const double scale = sqrt(temperature_target / T);
  for (T_INT i = 0; i < n; i++) {
    hv[3 * i] *= scale;
    hv[3 * i + 1] *= scale;
    hv[3 * i + 2] *= scale;
  }
  s.grow(n);
  s.upload_mass();
  EMD_CALL(s.ctx, emd_system_upload(s.ctx, &s.d, n, hx.data(), hv.data(), htype.data(),
                                    hid.data(), hq.data()));
  s.N_max = s.d.n_max;
}
