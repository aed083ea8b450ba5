// comm_serial_hip.h -- single-GPU periodic images (replaces CommSerial,
// src/comm_types/comm_serial.{h,cpp}).
#ifdef MODULES_OPTION_CHECK
      if( (strcmp(argv[i+1], "SERIAL") == 0) )
        comm_type = COMM_SERIAL;
#endif
#ifdef COMM_MODULES_INSTANTIATION
      else if(input->comm_type == COMM_SERIAL) {
        comm = new CommSerialHIP(system,input->force_cutoff + input->neighbor_skin);
      }
#endif

#if !defined(MODULES_OPTION_CHECK) && !defined(COMM_MODULES_INSTANTIATION)
#ifndef COMM_SERIAL_HIP_H
#define COMM_SERIAL_HIP_H
#include "comm.h"

class CommSerialHIP : public Comm {
  emd_halo h;

public:
  CommSerialHIP(System* s, T_X_FLOAT comm_depth_);
  ~CommSerialHIP();
  void exchange() override;
  void exchange_halo() override;
  void update_halo() override;
  void update_force() override;
  const char* name() override;
  const emd_halo* halo() const { return &h; }
  const emd_halo* device_halo() const override { return &h; }
};
#endif
#endif
