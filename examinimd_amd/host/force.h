// force.h -- Force module base class (src/force.h:46-57).
#ifndef FORCE_H
#define FORCE_H
#include "types.h"
#include "system.h"
#include "binning.h"
#include "neighbor.h"

class Force {
public:
  bool half_neigh, comm_newton;
  Force(char** args, System* system, bool half_neigh_);
  virtual ~Force() {}
  virtual void init_coeff(int nargs, char** args);
  virtual void compute(System* system, Binning* binning, Neighbor* neigh);
  // the reference returns 0 by default (src/force.h:54)
  virtual T_F_FLOAT compute_energy(System*, Binning*, Neighbor*) { return 0.0; }
  // compute() leaves f complete by itself (it writes every local atom's
  // force outright and no ghost force, or zeroes every row first), so the
  // driver's zeroing of f before it can be skipped
  virtual bool stores_forces() const { return false; }
  virtual const char* name();
};

#include "modules_force.h"
#endif
