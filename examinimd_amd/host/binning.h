// binning.h -- Binning module base class (src/binning.h:43-67).
// Public fields keep the reference's names; bincount, binoffsets and
// permute_vector are device arrays (x-major bin order).
#ifndef BINNING_H
#define BINNING_H
#include "types.h"
#include "system.h"

class Binning {
protected:
  System* system;

public:
  T_INT nbinx, nbiny, nbinz, nhalo;
  T_X_FLOAT minx, maxx, miny, maxy, minz, maxz;

  int* bincount;        // device [nbinx*nbiny*nbinz]
  int* binoffsets;      // device [nbinx*nbiny*nbinz]
  int* permute_vector;  // device [range]

  bool is_sorted;

  Binning(System* s);
  virtual ~Binning() {}
  virtual void create_binning(T_X_FLOAT dx, T_X_FLOAT dy, T_X_FLOAT dz, int halo_depth,
                              bool do_local, bool do_ghost, bool sort);
  virtual const char* name();
  // the library-side description the neighbour builders consume
  virtual const emd_bins* device_bins() const { return nullptr; }
};

#include "modules_binning.h"
#endif
