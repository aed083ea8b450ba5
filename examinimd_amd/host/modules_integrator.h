// modules_integrator.h -- integrator modules (src/modules_integrator.h)
#include "integrator_nve_hip.h"
