// integrator.cpp -- Integrator base (src/integrator.cpp) and IntegratorNVEHIP
// (src/integrator_nve.cpp:41-121).
#include "integrator.h"

Integrator::Integrator(System* s) : system(s), timestep_size(0) {}
Integrator::~Integrator() {}
void Integrator::initial_integrate() {}
void Integrator::final_integrate() {}
void Integrator::final_initial_integrate() {
  final_integrate();
  initial_integrate();
}
bool Integrator::final_initial_integrate_boundary(const emd_halo*, bool, bool) { return false; }

IntegratorNVEHIP::IntegratorNVEHIP(System* s) : Integrator(s) {
  dtv = system->dt;
  dtf = 0.5 * system->dt / system->mvv2e;
}

void IntegratorNVEHIP::initial_integrate() {
  EMD_CALL(system->ctx, emd_nve_initial(system->ctx, system->view(), dtf, dtv, system->precision));
}

void IntegratorNVEHIP::final_integrate() {
  EMD_CALL(system->ctx, emd_nve_final(system->ctx, system->view(), dtf, system->precision));
}

void IntegratorNVEHIP::final_initial_integrate() {
  EMD_CALL(system->ctx,
           emd_nve_final_initial(system->ctx, system->view(), dtf, dtv, system->precision));
}

bool IntegratorNVEHIP::final_initial_integrate_boundary(const emd_halo* h, bool fold,
                                                        bool update_halo) {
  if (!h) return false;
  EMD_CALL(system->ctx, emd_nve_step_boundary(system->ctx, system->view(), h, fold ? 1 : 0,
                                              update_halo ? 1 : 0, dtf, dtv, system->precision));
  return true;
}
