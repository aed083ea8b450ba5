// neighbor.h -- Neighbor module base class (src/neighbor.h:45-61).
// get_neigh_list() hands force modules the device list (CSR or padded 2D
// layout) instead of a Kokkos graph reached by a static downcast.
#ifndef NEIGHBOR_H
#define NEIGHBOR_H
#include "types.h"
#include "system.h"
#include "binning.h"

class Neighbor {
public:
  int neigh_type;
  bool comm_newton;
  Neighbor();
  virtual ~Neighbor();
  virtual void init(T_X_FLOAT neighcut);
  virtual void create_neigh_list(System* system, Binning* binning, bool half_neigh_,
                                 bool ghost_neighs_);
  virtual const emd_neigh* get_neigh_list();
  virtual const char* name();
};

#include "modules_neighbor.h"
#endif
