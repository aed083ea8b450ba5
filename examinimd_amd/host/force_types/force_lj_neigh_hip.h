// force_lj_neigh_hip.h -- 12-6 Lennard-Jones over a neighbour list on the GPU
// (replaces ForceLJNeigh, src/force_types/force_lj_neigh{.h,_impl.h}).
#ifdef MODULES_OPTION_CHECK
      if( (strcmp(argv[i+1], "NEIGH_FULL") == 0) )
        force_iteration_type = FORCE_ITER_NEIGH_FULL;
      if( (strcmp(argv[i+1], "NEIGH_HALF") == 0) ) {
        force_iteration_type = FORCE_ITER_NEIGH_HALF;
      }
#endif
#ifdef FORCE_MODULES_INSTANTIATION
    else if (input->force_type == FORCE_LJ) {
      bool half_neigh = input->force_iteration_type == FORCE_ITER_NEIGH_HALF;
      force = new ForceLJNeighHIP(input->input_data.words[input->force_line],system,half_neigh);
    }
#endif

#if !defined(MODULES_OPTION_CHECK) && !defined(FORCE_MODULES_INSTANTIATION)
#ifndef FORCE_LJ_NEIGH_HIP_H
#define FORCE_LJ_NEIGH_HIP_H
#include "force.h"

class ForceLJNeighHIP : public Force {
  int ntypes;
  bool use_stackparams;
  emd_lj_params params;
  std::vector<double> table;  // more than 16 types: lj1, lj2, cutsq [ntypes][ntypes]

public:
  ForceLJNeighHIP(char** args, System* system, bool half_neigh_);
  void init_coeff(int nargs, char** args) override;
  void compute(System* system, Binning* binning, Neighbor* neighbor) override;
  T_F_FLOAT compute_energy(System* system, Binning* binning, Neighbor* neighbor) override;
  bool stores_forces() const override { return !half_neigh; }
  const char* name() override { return half_neigh ? "ForceLJNeighHalf" : "ForceLJNeighFull"; }
  const emd_lj_params& lj_params() const { return params; }
};
#endif
#endif
