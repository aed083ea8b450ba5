// Temperature / KinE / PotE (src/property_{temperature,kine,pote}.{h,cpp})
#ifndef PROPERTY_TEMPERATURE_H
#define PROPERTY_TEMPERATURE_H
#include "types.h"
#include "system.h"
#include "comm.h"
#include "force.h"

class Temperature {
  Comm* comm;
public:
  Temperature(Comm* comm_) : comm(comm_) {}
  T_V_FLOAT compute(System*);
};

class KinE {
  Comm* comm;
public:
  KinE(Comm* comm_) : comm(comm_) {}
  T_V_FLOAT compute(System*);
};

class PotE {
  Comm* comm;
public:
  PotE(Comm* comm_) : comm(comm_) {}
  T_F_FLOAT compute(System*, Binning*, Neighbor*, Force*);
};
#endif
