// modules_force.h -- force modules (src/modules_force.h)
#include "force_types/force_lj_neigh_hip.h"
#include "force_types/force_lj_idial_neigh_hip.h"
#include "force_types/force_snap_hip.h"
