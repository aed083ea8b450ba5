// modules_binning.h -- binning modules (src/modules_binning.h); the driver
// instantiates BinningHIP for BINNING_KKSORT.
#include "binning_types/binning_hip.h"
