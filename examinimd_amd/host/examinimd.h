// examinimd.h -- the driver (src/examinimd.h:49-71).
#ifndef EXAMINIMD_H
#define EXAMINIMD_H
#include "types.h"
#include "system.h"
#include "input.h"
#include "integrator.h"
#include "binning.h"
#include "neighbor.h"
#include "force.h"
#include "comm.h"

class ExaMiniMD {
public:
  System* system;
  Integrator* integrator;
  Force* force;
  Neighbor* neighbor;
  Comm* comm;
  Binning* binning;
  Input* input;

  // phase timers: synchronise the device at phase boundaries so the
  // force/neigh/comm/other columns are device time (on by default)
  bool phase_sync;
  // fuse final_integrate with the next step's initial_integrate where no
  // thermo, dump or correctness check observes the step boundary
  bool fuse_integrate;
  double time_total, time_force, time_neigh, time_comm, time_other;
  double last_thermo[3];

  ExaMiniMD();
  ~ExaMiniMD();
  void init(int argc, char* argv[]);
  void run(int nsteps);
  void thermo(double out[3]);
  void dump_binary(int step);
  // canonical neighbour-list dump (--dumpneigh N DIR); write_neigh_file
  // writes the current list to `path` whatever the step
  void dump_neigh(int step);
  int write_neigh_file(const char* path);
  void read_restart(const char* filename);
  void check_correctness(int step);
  void print_performance();
  void shutdown();
};
#endif
