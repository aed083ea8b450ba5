// neighbor.cpp -- Neighbor base (src/neighbor.cpp) and NeighborHIP.
#include "neighbor.h"

Neighbor::Neighbor() : neigh_type(NEIGH_NONE), comm_newton(false) {}
Neighbor::~Neighbor() {}
void Neighbor::init(T_X_FLOAT) {}
void Neighbor::create_neigh_list(System*, Binning*, bool, bool) {}
const emd_neigh* Neighbor::get_neigh_list() { return nullptr; }
const char* Neighbor::name() { return "NeighborNone"; }

NeighborHIP::NeighborHIP(int type) : neigh_cut(0.0), ctx(nullptr) {
  neigh_type = type;
  memset(&nl, 0, sizeof(nl));
}
NeighborHIP::~NeighborHIP() { if (ctx) emd_neigh_free(ctx, &nl); }

// NeighborCSR::create_neigh_list (neighbor_csr.h:370-431) /
// Neighbor2D::create_neigh_list (neighbor_2d.h:280-331)
void NeighborHIP::create_neigh_list(System* system, Binning* binning, bool half_neigh_, bool) {
  ctx = system->ctx;
  const emd_bins* b = binning->device_bins();
  if (!b) { fprintf(stderr, "NeighborHIP needs a GPU binning module\n"); exit(1); }
  EMD_CALL(ctx, emd_neigh_build(ctx, system->view(), b, neigh_cut, half_neigh_ ? 1 : 0,
                                comm_newton ? 1 : 0,
                                neigh_type == NEIGH_2D ? EMD_NEIGH_2D : EMD_NEIGH_CSR, &nl));
}
