// comm.h -- Comm module base class (src/comm.h:47-109).
#ifndef COMM_H
#define COMM_H
#include "types.h"
#include "system.h"
#include "binning.h"

class Comm {
protected:
  System* system;
  T_X_FLOAT comm_depth;

public:
  Comm(System* s, T_X_FLOAT comm_depth);
  virtual ~Comm();
  virtual void init();
  // Move particles which left local domain
  virtual void exchange();
  // Exchange ghost particles
  virtual void exchange_halo();
  // Update ghost particles
  virtual void update_halo();
  // Reverse communication of forces
  virtual void update_force();
  // the device halo that update_force / update_halo use, for a fused step
  // boundary (null: this module's exchanges cannot be fused)
  virtual const emd_halo* device_halo() const;
  virtual void reduce_float(T_FLOAT* values, T_INT N);
  virtual void reduce_int(T_INT* values, T_INT N);
  virtual void reduce_max_float(T_FLOAT* values, T_INT N);
  virtual void reduce_max_int(T_INT* values, T_INT N);
  virtual void reduce_min_float(T_FLOAT* values, T_INT N);
  virtual void reduce_min_int(T_INT* values, T_INT N);
  virtual void scan_int(T_INT* values, T_INT N);
  virtual void weighted_reduce_float(T_FLOAT* values, T_INT* weight, T_INT N);
  virtual void create_domain_decomposition();
  virtual int process_rank();
  virtual int num_processes();
  virtual void error(const char*);
  virtual const char* name();
};

#include "modules_comm.h"
#endif
