// neighbor_hip.h -- GPU neighbour-list modules.
//   NeighborCSR  (--neigh-type CSR): src/neighbor_types/neighbor_csr.h
//   Neighbor2D   (--neigh-type 2D, the reference default): neighbor_2d.h
// Both build the same rows (27-bin stencil order, then bin order) with the
// count / scan / fill kernels of emd_neigh_build; they differ in layout.

// Runtime check for these classes
#ifdef MODULES_OPTION_CHECK
      if( (strcmp(argv[i+1], "CSR") == 0) )
        neighbor_type = NEIGH_CSR;
      if( (strcmp(argv[i+1], "2D") == 0) )
        neighbor_type = NEIGH_2D;
#endif

// Instantiation and init
#ifdef NEIGHBOR_MODULES_INSTANTIATION
    else if (input->neighbor_type == NEIGH_CSR) {
      neighbor = new NeighborHIP(NEIGH_CSR);
      neighbor->init(input->force_cutoff + input->neighbor_skin);
    }
    else if (input->neighbor_type == NEIGH_2D) {
      neighbor = new NeighborHIP(NEIGH_2D);
      neighbor->init(input->force_cutoff + input->neighbor_skin);
    }
#endif

#if !defined(MODULES_OPTION_CHECK) && !defined(NEIGHBOR_MODULES_INSTANTIATION)
#ifndef NEIGHBOR_HIP_H
#define NEIGHBOR_HIP_H
#include "neighbor.h"

class NeighborHIP : public Neighbor {
  T_X_FLOAT neigh_cut;
  emd_neigh nl;
  emd_ctx* ctx;

public:
  explicit NeighborHIP(int type);
  ~NeighborHIP();
  void init(T_X_FLOAT neigh_cut_) override { neigh_cut = neigh_cut_; }
  void create_neigh_list(System* system, Binning* binning, bool half_neigh_, bool) override;
  const emd_neigh* get_neigh_list() override { return &nl; }
  const char* name() override { return neigh_type == NEIGH_2D ? "Neighbor2D" : "NeighborCSR"; }
};
#endif
#endif
