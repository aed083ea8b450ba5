// integrator_nve_hip.h -- velocity-Verlet NVE on the GPU (replaces
// IntegratorNVE, src/integrator_nve.{h,cpp}).
#ifndef INTEGRATOR_NVE_HIP_H
#define INTEGRATOR_NVE_HIP_H
#include "integrator.h"

class IntegratorNVEHIP : public Integrator {
  T_V_FLOAT dtv, dtf;

public:
  IntegratorNVEHIP(System* s);
  void initial_integrate() override;
  void final_integrate() override;
  void final_initial_integrate() override;
  bool final_initial_integrate_boundary(const emd_halo* h, bool fold, bool update_halo) override;
};
#endif
