// modules_neighbor.h -- neighbour modules (src/modules_neighbor.h)
#include "neighbor_types/neighbor_hip.h"
