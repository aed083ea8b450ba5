// input.cpp -- deck parsing, command line and lattice creation
// (src/input.cpp:44-792).
#include "input.h"

#include <array>
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
Input::Input(System* p) : system(p) {
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
  dumpneigh_rate = 0;
  dumpneigh_path = nullptr;
  comm_newton = 0;
  precision = EMD_FP64;
  device = 0;
  timestep_given = false;
}

// --------------------------------------------------------- command line
// Options of the reference driver (src/input.cpp:151-224: -il, the module
// selectors, --dumpbinary, --correctness; --kokkos-* is passed through) and
// this build's own (--precision, --device, --restart).
namespace {
struct CliOption {
  const char* spelling;
  const char* meaning;
};
const CliOption kCliOptions[] = {
    {"-il FILE | --input-lammps FILE", "LAMMPS-style input deck"},
    {"--force-iteration NEIGH_FULL|NEIGH_HALF", "full or half neighbour list for the pair force"},
    {"--neigh-type 2D|CSR", "neighbour-list layout (padded rows or compressed rows)"},
    {"--comm-type SERIAL", "periodic images on one GPU"},
    {"--dumpbinary N DIR", "write DIR/output.<step>.<rank> every N steps"},
    {"--correctness N DIR FILE", "every N steps compare with DIR/output.<step>.<rank>, "
                                 "append the differences to FILE"},
    {"--dumpneigh N DIR", "write the neighbour list, canonically ordered, to "
                          "DIR/neigh.<step>.<rank> every N steps"},
    {"--precision fp64|mixed", "force arithmetic: FP64, or FP32 pair math with FP64 state"},
    {"--device N", "GPU ordinal"},
    {"--restart FILE", "initial atoms from a --dumpbinary file instead of the lattice"},
};

void print_usage() {
  printf("ExaMiniMD on HIP (gfx950)\nusage: ExaMiniMD -il <deck> [options]\n");
  for (const CliOption& o : kCliOptions) printf("  %-42s %s\n", o.spelling, o.meaning);
}

[[noreturn]] void cli_fail(const char* what, const char* arg) {
  fprintf(stderr, "ExaMiniMD: %s%s%s\n", what, arg ? ": " : "", arg ? arg : "");
  exit(1);
}
}  // namespace

void Input::read_command_line_args(int argc, char* argv[]) {
  // the module headers' MODULES_OPTION_CHECK sections read argv[i+1]
#define MODULES_OPTION_CHECK
  for (int i = 1; i < argc; i++) {
    const char* opt = argv[i];
    auto is = [opt](const char* a, const char* b = nullptr) {
      return strcmp(opt, a) == 0 || (b && strcmp(opt, b) == 0);
    };
    auto want = [&](int k) {
      if (i + k >= argc) cli_fail("option needs more arguments", opt);
    };
    if (is("-h", "--help")) {
      if (system->do_print) print_usage();
    } else if (is("-il", "--input-lammps")) {
      want(1);
      input_file = argv[++i];
      input_file_type = INPUT_LAMMPS;
    } else if (is("--force-iteration")) {
      want(1);
#include "modules_force.h"
      ++i;
    } else if (is("--comm-type")) {
      want(1);
#include "modules_comm.h"
      ++i;
    } else if (is("--neigh-type")) {
      want(1);
#include "modules_neighbor.h"
      ++i;
    } else if (is("--dumpbinary")) {
      want(2);
      dumpbinaryflag = true;
      dumpbinary_rate = atoi(argv[++i]);
      dumpbinary_path = argv[++i];
    } else if (is("--correctness")) {
      want(3);
      correctnessflag = true;
      correctness_rate = atoi(argv[++i]);
      reference_path = argv[++i];
      correctness_file = argv[++i];
    } else if (is("--dumpneigh")) {
      want(2);
      dumpneigh_rate = atoi(argv[++i]);
      dumpneigh_path = argv[++i];
      if (dumpneigh_rate <= 0) cli_fail("--dumpneigh needs a positive step interval", nullptr);
    } else if (is("--precision")) {
      want(1);
      const char* p = argv[++i];
      if (strcmp(p, "mixed") == 0) precision = EMD_MIXED;
      else if (strcmp(p, "fp64") == 0) precision = EMD_FP64;
      else cli_fail("--precision takes fp64 or mixed, not", p);
    } else if (is("--device")) {
      want(1);
      device = atoi(argv[++i]);
    } else if (is("--restart")) {
      want(1);
      restart_file = argv[++i];
    } else if (strncmp(opt, "--kokkos-", 9) != 0) {
      cli_fail("unrecognised option", opt);
    }
  }
#undef MODULES_OPTION_CHECK
}

// ---------------------------------------------------------------- deck
void Input::read_file(const char* filename) {
  if (filename == NULL) filename = input_file;
  if (input_file_type != INPUT_LAMMPS || filename == NULL)
    cli_fail("no input deck given (use -il FILE)", nullptr);
  read_lammps_file(filename);
}

void Input::read_lammps_file(const char* filename) {
  std::ifstream file(filename);
  if (!file.good()) cli_fail("cannot read the input deck", filename);
  std::string text;
  while (std::getline(file, text)) input_data.add_line(text.c_str());
  input_data.finalize();
  if (system->do_print) {
    printf("\n#InputFile:\n#=========================================================\n");
    input_data.print();
    printf("#=========================================================\n\n");
  }
  for (int l = 0; l < input_data.nlines; l++) check_lammps_command(l);
}

// The deck subset of src/input.cpp:265-458, as a table of keyword handlers.
// A handler returns true when it understood its line; a line that no handler
// understood is reported and otherwise ignored, as in the reference.  Each
// handler states the rule it keeps from the reference.
namespace {
using DeckWords = char**;
struct DeckKeyword {
  const char* keyword;
  bool (*handle)(Input& in, DeckWords w, int line);
};

bool say(Input& in, const char* fmt, const char* word) {
  if (in.system->do_print) printf(fmt, word);
  return false;
}

bool deck_units(Input& in, DeckWords w, int) {
  System& s = *in.system;
  // metal always resets dt; real and lj only when no timestep was given
  if (strcmp(w[1], "lj") == 0) {
    in.units = UNITS_LJ;
    s.boltz = 1.0;
    s.mvv2e = 1.0;
    if (!in.timestep_given) s.dt = 0.005;
  } else if (strcmp(w[1], "real") == 0) {
    in.units = UNITS_REAL;
    s.boltz = 0.0019872067;
    s.mvv2e = 48.88821291 * 48.88821291;
    if (!in.timestep_given) s.dt = 1.0;
  } else if (strcmp(w[1], "metal") == 0) {
    in.units = UNITS_METAL;
    s.boltz = 8.617343e-5;
    s.mvv2e = 1.0364269e-4;
    s.dt = 0.001;
  } else {
    return say(in, "deck: units '%s' not available (lj, real, metal)\n", w[1]);
  }
  return true;
}

bool deck_atom_style(Input& in, DeckWords w, int) {
  if (strcmp(w[1], "atomic") == 0) return true;
  return say(in, "deck: atom_style '%s' not available (atomic)\n", w[1]);
}

bool deck_lattice(Input& in, DeckWords w, int) {
  bool ok = true;
  const double arg = atof(w[2]);
  if (strcmp(w[1], "fcc") == 0) {
    in.lattice_style = LATTICE_FCC;
    in.lattice_constant = std::pow((4.0 / arg), (1.0 / 3.0));  // the argument is a density
  } else if (strcmp(w[1], "sc") == 0 || strcmp(w[1], "bcc") == 0) {
    in.lattice_style = w[1][0] == 's' ? LATTICE_SC : LATTICE_BCC;
    in.lattice_constant = arg;  // the argument is the cell edge
  } else {
    ok = say(in, "deck: lattice '%s' not available (sc, fcc, bcc)\n", w[1]);
  }
  if (strcmp(w[3], "origin") == 0) {
    in.lattice_offset_x = atof(w[4]);
    in.lattice_offset_y = atof(w[5]);
    in.lattice_offset_z = atof(w[6]);
  }
  return ok;
}

bool deck_region(Input& in, DeckWords w, int) {
  if (strcmp(w[2], "block") != 0)
    return say(in, "deck: region style '%s' not available (block)\n", w[2]);
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = atoi(w[3 + 2 * d]);
    hi[d] = atoi(w[4 + 2 * d]);
  }
  if (lo[0] || lo[1] || lo[2]) say(in, "deck: region %s must start at 0 0 0\n", w[1]);
  in.lattice_nx = hi[0];
  in.lattice_ny = hi[1];
  in.lattice_nz = hi[2];
  return true;
}

bool deck_create_box(Input& in, DeckWords w, int) {
  in.system->ntypes = atoi(w[1]);
  in.system->mass.assign(in.system->ntypes, 0.0);
  return true;
}

bool deck_mass(Input& in, DeckWords w, int) {
  const int t = atoi(w[1]) - 1;
  if (t >= 0 && t < (int)in.system->mass.size()) in.system->mass[t] = atof(w[2]);
  return true;
}

bool deck_pair_style(Input& in, DeckWords w, int line) {
  struct { const char* style; int type; } styles[] = {
      {"lj/cut", FORCE_LJ}, {"lj/cut/idial", FORCE_LJ_IDIAL}, {"snap", FORCE_SNAP}};
  for (auto& st : styles)
    if (strcmp(w[1], st.style) == 0) {
      in.force_type = st.type;
      // the reference fixes the SNAP neighbour cutoff at 4.73442 whatever
      // the deck says (src/input.cpp:378-383)
      in.force_cutoff = st.type == FORCE_SNAP ? 4.73442 : atof(w[2]);
      in.force_line = line;
      return true;
    }
  return say(in, "deck: pair_style '%s' not available (lj/cut, lj/cut/idial, snap)\n", w[1]);
}

bool deck_pair_coeff(Input& in, DeckWords, int line) {
  in.force_coeff_lines.push_back(line);  // handed to Force::init_coeff after creation
  return true;
}

bool deck_velocity(Input& in, DeckWords w, int) {
  if (strcmp(w[1], "all") != 0) say(in, "deck: velocity applies to group 'all' only, not '%s'\n", w[1]);
  if (strcmp(w[2], "create") != 0) say(in, "deck: velocity supports 'create' only, not '%s'\n", w[2]);
  in.temperature_target = atof(w[3]);
  in.temperature_seed = atoi(w[4]);
  for (int k = 5; k + 1 < ItemizedFile::words_per_line; k++)
    if (strcmp(w[k], "dist") == 0) in.velocity_gaussian = strcmp(w[k + 1], "gaussian") == 0;
  return true;
}

bool deck_neighbor(Input& in, DeckWords w, int) {
  in.neighbor_skin = atof(w[1]);
  return true;
}

bool deck_neigh_modify(Input& in, DeckWords w, int) {
  for (int k = 1; k + 1 < ItemizedFile::words_per_line; k++)
    if (strcmp(w[k], "every") == 0) in.comm_exchange_rate = atoi(w[k + 1]);
  return true;
}

bool deck_fix(Input& in, DeckWords w, int) {
  if (strcmp(w[3], "nve") != 0) return say(in, "deck: fix style '%s' not available (nve)\n", w[3]);
  in.integrator_type = INTEGRATOR_NVE;
  return true;
}

bool deck_run(Input& in, DeckWords w, int) {
  in.nsteps = atoi(w[1]);
  return true;
}

bool deck_thermo(Input& in, DeckWords w, int) {
  in.thermo_rate = atoi(w[1]);
  return true;
}

bool deck_timestep(Input& in, DeckWords w, int) {
  in.system->dt = atof(w[1]);
  in.timestep_given = true;
  return true;
}

bool deck_newton(Input& in, DeckWords w, int) {
  if (strcmp(w[1], "on") == 0) in.comm_newton = 1;
  else if (strcmp(w[1], "off") == 0) in.comm_newton = 0;
  else say(in, "deck: newton takes on or off, not '%s'\n", w[1]);
  return true;
}

bool deck_accept(Input&, DeckWords, int) { return true; }

bool deck_variable(Input& in, DeckWords w, int) {
  return say(in, "deck: '%s' (deck variables) is not supported\n", w[0]);
}

const DeckKeyword kDeckKeywords[] = {
    {"units", deck_units},         {"atom_style", deck_atom_style},
    {"lattice", deck_lattice},     {"region", deck_region},
    {"create_box", deck_create_box}, {"create_atoms", deck_accept},
    {"mass", deck_mass},           {"pair_style", deck_pair_style},
    {"pair_coeff", deck_pair_coeff}, {"velocity", deck_velocity},
    {"neighbor", deck_neighbor},   {"neigh_modify", deck_neigh_modify},
    {"fix", deck_fix},             {"run", deck_run},
    {"thermo", deck_thermo},       {"timestep", deck_timestep},
    {"newton", deck_newton},       {"variable", deck_variable},
};
}  // namespace

void Input::check_lammps_command(int line) {
  DeckWords w = input_data.words[line];
  // blank lines and lines whose first word holds a '#' are comments
  if (w[0][0] == 0 || strchr(w[0], '#')) return;
  bool understood = false;
  for (const DeckKeyword& k : kDeckKeywords)
    if (strcmp(w[0], k.keyword) == 0) {
      understood = k.handle(*this, w, line);
      break;
    }
  if (!understood && system->do_print) {
    printf("deck: line %d not understood, ignored: ", line + 1);
    input_data.print_line(line);
  }
}

// -------------------------------------------------------------- lattice
// Input::create_lattice, src/input.cpp:460-792, for one process (the
// sub-domain is the box).  What pins every trajectory and is kept exactly:
// sites are visited cell by cell with z outermost, then y, then x, then the
// basis sites; a site's coordinate is a * (cell index + basis offset) with
// the deck's origin folded into the basis; a site belongs to the domain when
// lo <= x < hi per axis; ids count 1, 2, ... in visiting order.  The velocity
// of a site is drawn from the geom generator reseeded by (seed, position);
// the net momentum is then removed and the temperature scaled to the target.
namespace {
struct LatticeSites {
  std::vector<double> x;  // [n][3]
  std::vector<int> type, id;
  int n() const { return (int)id.size(); }
};

LatticeSites build_sites(const Input& in, const System& s) {
  std::vector<std::array<double, 3>> basis;
  switch (in.lattice_style) {
    case LATTICE_FCC:
      basis = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
      break;
    case LATTICE_BCC:
      basis = {{0.0, 0.0, 0.0}, {0.5, 0.5, 0.5}};
      break;
    default:
      basis = {{0.0, 0.0, 0.0}};
  }
  for (auto& b : basis) {
    b[0] += in.lattice_offset_x;
    b[1] += in.lattice_offset_y;
    b[2] += in.lattice_offset_z;
  }
  const double a = in.lattice_constant;
  const double lo[3] = {s.sub_domain_lo_x, s.sub_domain_lo_y, s.sub_domain_lo_z};
  const double hi[3] = {s.sub_domain_hi_x, s.sub_domain_hi_y, s.sub_domain_hi_z};
  const double len[3] = {s.domain_x, s.domain_y, s.domain_z};
  const int cells[3] = {in.lattice_nx, in.lattice_ny, in.lattice_nz};
  // cell index range covering the sub-domain, one cell of slack each side
  // (the float-to-int conversion truncates, as the reference's does)
  T_INT first[3], last[3];
  for (int d = 0; d < 3; d++) {
    first[d] = (T_INT)(lo[d] / len[d] * cells[d] - 0.5);
    last[d] = (T_INT)(hi[d] / len[d] * cells[d] + 0.5);
  }
  LatticeSites out;
  for (T_INT cz = first[2]; cz <= last[2]; cz++)
    for (T_INT cy = first[1]; cy <= last[1]; cy++)
      for (T_INT cx = first[0]; cx <= last[0]; cx++)
        for (const auto& b : basis) {
          const double p[3] = {a * (1.0 * cx + b[0]), a * (1.0 * cy + b[1]), a * (1.0 * cz + b[2])};
          bool inside = true;
          for (int d = 0; d < 3; d++) inside = inside && p[d] >= lo[d] && p[d] < hi[d];
          if (!inside) continue;
          out.x.insert(out.x.end(), p, p + 3);
          out.type.push_back(rand() % s.ntypes);
          out.id.push_back(out.n() + 1);
        }
  return out;
}

// v of one site before the momentum and temperature corrections: three
// draws (uniform - 1/2 by default, Gaussian with `dist gaussian`) divided by
// sqrt(m)
void site_velocity(int seed, const double* pos, double mass, bool gaussian, double* v) {
  LAMMPS_RandomVelocityGeom rng;
  double p[3] = {pos[0], pos[1], pos[2]};
  rng.reset(seed, p);
  const double root_m = sqrt(mass);
  for (int d = 0; d < 3; d++) {
    const double draw = gaussian ? rng.gaussian() : rng.uniform() - 0.5;
    v[d] = draw / root_m;
  }
}
}  // namespace

void Input::create_lattice(Comm* comm) {
  System& s = *system;
  s.domain_x = lattice_constant * lattice_nx;
  s.domain_y = lattice_constant * lattice_ny;
  s.domain_z = lattice_constant * lattice_nz;
  comm->create_domain_decomposition();

  LatticeSites sites = build_sites(*this, s);
  const T_INT n = sites.n();
  s.N_local = n;
  s.N = n;
  s.N_ghost = 0;
  T_INT id_end = n;  // ids unique over processes: shift by the exclusive scan
  comm->scan_int(&id_end, 1);
  for (int& id : sites.id) id += id_end - n;
  comm->reduce_int(&s.N, 1);
  if (s.do_print) printf("Atoms: %i %i\n", s.N, s.N_local);

  std::vector<double> v(3 * (size_t)n), q(n, 0.0);
  double mom[3] = {0.0, 0.0, 0.0}, mass_sum = 0.0;
  for (T_INT i = 0; i < n; i++) {
    const double m = s.mass[sites.type[i]];
    double* vi = &v[3 * (size_t)i];
    site_velocity(temperature_seed, &sites.x[3 * (size_t)i], m, velocity_gaussian, vi);
    mass_sum += m;
    for (int d = 0; d < 3; d++) mom[d] += m * vi[d];
  }
  comm->reduce_float(mom, 3);
  comm->reduce_float(&mass_sum, 1);
  const double drift[3] = {mom[0] / mass_sum, mom[1] / mass_sum, mom[2] / mass_sum};
  for (T_INT i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) v[3 * (size_t)i + d] -= drift[d];

  // Temperature::compute (src/property_temperature.cpp:43-62) summed on the
  // host in atom order: the scale factor pins every trajectory
  double mv2 = 0.0;
  for (T_INT i = 0; i < n; i++) {
    const double* vi = &v[3 * (size_t)i];
    mv2 += (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]) * s.mass[sites.type[i]];
  }
  comm->reduce_float(&mv2, 1);
  const T_INT dof = 3 * s.N - 3;
  const double T = mv2 * (s.mvv2e / (1.0 * dof * s.boltz));
  const double scale = sqrt(temperature_target / T);
  for (double& c : v) c *= scale;

  s.grow(n);
  s.upload_mass();
  EMD_CALL(s.ctx, emd_system_upload(s.ctx, &s.d, n, sites.x.data(), v.data(), sites.type.data(),
                                    sites.id.data(), q.data()));
  s.N_max = s.d.n_max;
}
