// integrator.h -- Integrator module base (src/integrator.h:45-55).
#ifndef INTEGRATORS_H
#define INTEGRATORS_H
#include "types.h"
#include "system.h"

class Integrator {
public:
  System* system;
  Integrator(System* s);
  virtual ~Integrator();
  T_V_FLOAT timestep_size;
  virtual void initial_integrate();
  virtual void final_integrate();
  // final_integrate of one step then initial_integrate of the next, for a
  // step boundary with nothing in between; a module may fuse the two
  virtual void final_initial_integrate();
  // the whole step boundary on the device: the halo's force fold (if
  // `fold`), final_integrate, initial_integrate and the halo's position
  // update (if `update_halo`); false if this module cannot (nothing done)
  virtual bool final_initial_integrate_boundary(const emd_halo* h, bool fold, bool update_halo);
};

#include "modules_integrator.h"
#endif
