// force_snap_hip.h -- linear SNAP on the GPU (replaces ForceSNAP,
// src/force_types/force_snap_neigh{.h,_impl.h} and sna{.h,_impl.hpp}).
#ifdef MODULES_OPTION_CHECK
      if( (strcmp(argv[i+1], "NEIGH_FULL") == 0) )
        force_iteration_type = FORCE_ITER_NEIGH_FULL;
#endif
#ifdef FORCE_MODULES_INSTANTIATION
    else if (input->force_type == FORCE_SNAP) {
      bool half_neigh = input->force_iteration_type == FORCE_ITER_NEIGH_HALF;
      if(half_neigh) { fprintf(stderr, "ForceSNAP does not support half neighborlist\n"); exit(1); }
      force = new ForceSNAPHIP(input->input_data.words[input->force_line],system,half_neigh);
    }
#endif

#if !defined(MODULES_OPTION_CHECK) && !defined(FORCE_MODULES_INSTANTIATION)
#ifndef FORCE_SNAP_HIP_H
#define FORCE_SNAP_HIP_H
#include <string>
#include <vector>
#include "force.h"

class ForceSNAPHIP : public Force {
  System* system;
  emd_snap* sn;
  int nelements;
  std::vector<std::string> elements;
  std::vector<double> radelem, wjelem;
  std::vector<std::vector<double>> coeffelem;
  int ncoeff, ncoeffall;
  int twojmax, diagonalstyle, switchflag, bzeroflag, quadraticflag;
  double rcutfac, rfac0, rmin0, rcutmax;
  void read_files(const char* coefffile, const char* paramfile);

public:
  ForceSNAPHIP(char** args, System* system, bool half_neigh_);
  ~ForceSNAPHIP();
  void init_coeff(int nargs, char** args) override;
  void compute(System* system, Binning* binning, Neighbor* neighbor) override;
  bool stores_forces() const override { return true; }
  // no counterpart in the reference (its ForceSNAP returns 0): linear SNAP
  // energy plus the energy of the reference's r^-14 repulsion
  T_F_FLOAT compute_energy(System* system, Binning* binning, Neighbor* neighbor) override;
  const char* name() override { return "ForceSNAP"; }
  emd_snap* handle() { return sn; }
};
#endif
#endif
