// modules_comm.h -- comm modules (src/modules_comm.h)
#include "comm_types/comm_serial_hip.h"
