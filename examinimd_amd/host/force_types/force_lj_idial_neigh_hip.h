// force_lj_idial_neigh_hip.h -- pair_style lj/cut/idial on the GPU: the LJ
// pair force summed `intensity` times in 1/intensity parts, an
// arithmetic-intensity dial over the same memory traffic (replaces
// ForceLJIDialNeigh, src/force_types/force_lj_idial_neigh{.h,_impl.h}).
#ifdef FORCE_MODULES_INSTANTIATION
    else if (input->force_type == FORCE_LJ_IDIAL) {
      bool half_neigh = input->force_iteration_type == FORCE_ITER_NEIGH_HALF;
      force = new ForceLJIDialNeighHIP(input->input_data.words[input->force_line],system,half_neigh);
    }
#endif

#if !defined(MODULES_OPTION_CHECK) && !defined(FORCE_MODULES_INSTANTIATION)
#ifndef FORCE_LJ_IDIAL_NEIGH_HIP_H
#define FORCE_LJ_IDIAL_NEIGH_HIP_H
#include "force.h"

class ForceLJIDialNeighHIP : public Force {
  emd_lj_idial_params params;

public:
  ForceLJIDialNeighHIP(char** args, System* system, bool half_neigh_);
  void init_coeff(int nargs, char** args) override;
  void compute(System* system, Binning* binning, Neighbor* neighbor) override;
  // no compute_energy: the reference keeps Force's default of 0
  bool stores_forces() const override { return !half_neigh; }
  const char* name() override {
    return half_neigh ? "ForceLJIDialNeighHalf" : "ForceLJIDialNeighFull";
  }
};
#endif
#endif
