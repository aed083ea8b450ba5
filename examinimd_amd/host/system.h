// system.h -- System: per-atom state on the device plus box and units.
// Mirrors src/system.h:57-143 (field names and grow() semantics); the arrays
// x, v, f, type, id, q live in HBM inside `d` (AoS double[N_max][3] for
// x, v, f; float[N_max][3] f32 for the mixed-precision force).
#ifndef EMD_SYSTEM_H
#define EMD_SYSTEM_H
#include <vector>
#include "types.h"

struct Particle {
  Particle() { x = y = z = vx = vy = vz = mass = q = 0.0; id = type = 0; }
  T_X_FLOAT x, y, z;
  T_V_FLOAT vx, vy, vz, mass;
  T_FLOAT q;
  T_INT id;
  int type;
};

class System {
public:
  T_INT N;        // Number of Global Particles
  T_INT N_max;    // Number of Particles I could have in available storage
  T_INT N_local;  // Number of owned Particles
  T_INT N_ghost;  // Number of non-owned Particles

  int ntypes;
  std::vector<double> mass;  // host copy, per type

  emd_ctx* ctx;   // device context all modules share
  emd_system d;   // device arrays (x, v, f, f32, type, id, q, mass)
  int precision;  // EMD_FP64 or EMD_MIXED

  T_X_FLOAT domain_x, domain_y, domain_z;
  T_X_FLOAT sub_domain_x, sub_domain_y, sub_domain_z;
  T_X_FLOAT sub_domain_lo_x, sub_domain_lo_y, sub_domain_lo_z;
  T_X_FLOAT sub_domain_hi_x, sub_domain_hi_y, sub_domain_hi_z;

  T_FLOAT boltz, mvv2e, dt;

  bool do_print;
  bool print_lammps;

  System();
  ~System() {}
  void init();
  void destroy();
  void grow(T_INT new_N);

  // device view with the host-side counters and box copied in
  emd_system* view();
  // take back counters/capacity a library call may have changed
  void pull(const emd_system* v);
  void upload_mass();

  void print_particles();
};
#endif
