// input.h -- LAMMPS-deck subset, command line and lattice set-up
// (src/input.{h,cpp}: same deck keywords, CLI options and Input fields).
// Extensions: `lattice bcc <a>`, `velocity all create T seed ... dist
// gaussian`, `--precision fp64|mixed`, `--device`, `--restart`.
#ifndef EMD_INPUT_H
#define EMD_INPUT_H
#include <string>
#include <vector>
#include "types.h"
#include "system.h"
#include "comm.h"

class ItemizedFile {
public:
  static const int words_per_line = 32;
  std::vector<std::vector<char*>> lines;  // each line padded to words_per_line ("" beyond)
  char*** words;                          // words[line][word]
  int nlines;
  std::vector<char**> line_ptrs;
  ItemizedFile();
  ~ItemizedFile();
  void add_line(const char* line);
  void finalize();
  int words_in_line(int line);
  void print_line(int line);
  void print();
};

// LAMMPS RanPark + "loop geom" reseeding (src/input.h:60-133)
class LAMMPS_RandomVelocityGeom {
  int seed;
public:
  LAMMPS_RandomVelocityGeom() : seed(0) {}
  double uniform();
  double gaussian();
  void reset(int ibase, double* coord);
};

class Input {
public:
  bool timestep_given;  // a `timestep` line was read (units real/lj keep it)
  System* system;
  char* input_file;
  int input_file_type;
  ItemizedFile input_data;

  int units;
  int lattice_style;
  double lattice_constant, lattice_offset_x, lattice_offset_y, lattice_offset_z;
  int lattice_nx, lattice_ny, lattice_nz;
  double temperature_target;
  int temperature_seed;
  bool velocity_gaussian;

  int integrator_type;
  int nsteps;
  int binning_type;
  int comm_type;
  int comm_exchange_rate;
  int comm_newton;
  int force_type;
  int force_iteration_type;
  int force_line;
  T_F_FLOAT force_cutoff;
  std::vector<int> force_coeff_lines;
  T_F_FLOAT neighbor_skin;
  int neighbor_type;

  int thermo_rate, dumpbinary_rate, correctness_rate;
  bool dumpbinaryflag, correctnessflag;
  char *dumpbinary_path, *reference_path, *correctness_file;
  int dumpneigh_rate;       // --dumpneigh N DIR (0: off)
  char *dumpneigh_path;
  char *restart_file;  // --restart: initial state from a binary dump

  int precision;  // EMD_FP64 or EMD_MIXED
  int device;

  Input(System* s);
  void read_command_line_args(int argc, char* argv[]);
  void read_file(const char* filename = NULL);
  void read_lammps_file(const char* filename);
  void check_lammps_command(int line);
  void create_lattice(Comm* comm);
};
#endif
