// binning_hip.h -- counting-sort binning on the GPU (replaces BinningKKSort,
// src/binning_types/binning_kksort.{h,cpp}).
#ifndef BINNING_HIP_H
#define BINNING_HIP_H
#include "binning.h"

class BinningHIP : public Binning {
  emd_bins b;

public:
  BinningHIP(System* s);
  ~BinningHIP();
  void create_binning(T_X_FLOAT dx, T_X_FLOAT dy, T_X_FLOAT dz, int halo_depth, bool do_local,
                      bool do_ghost, bool sort) override;
  const char* name() override;
  const emd_bins* device_bins() const override { return &b; }
};
#endif
