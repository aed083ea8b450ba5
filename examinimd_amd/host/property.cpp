// property.cpp -- thermo reductions on the GPU with compensated sums.
#include "property_temperature.h"

// Temperature::compute, src/property_temperature.cpp:43-62
T_V_FLOAT Temperature::compute(System* system) {
  double T = 0.0;
  EMD_CALL(system->ctx, emd_reduce_mv2(system->ctx, system->view(), &T));
  T_INT dof = 3 * system->N - 3;
  T_V_FLOAT factor = system->mvv2e / (1.0 * dof * system->boltz);
  comm->reduce_float(&T, 1);
  return T * factor;
}

// KinE::compute, src/property_kine.cpp:43-61
T_V_FLOAT KinE::compute(System* system) {
  double KE = 0.0;
  EMD_CALL(system->ctx, emd_reduce_mv2(system->ctx, system->view(), &KE));
  T_V_FLOAT factor = 0.5 * system->mvv2e;
  comm->reduce_float(&KE, 1);
  return KE * factor;
}

// PotE::compute, src/property_pote.cpp:44-49
T_F_FLOAT PotE::compute(System* system, Binning* binning, Neighbor* neighbor, Force* force) {
  T_F_FLOAT PE = force->compute_energy(system, binning, neighbor);
  comm->reduce_float(&PE, 1);
  return PE;
}
