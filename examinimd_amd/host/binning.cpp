// binning.cpp -- Binning base (src/binning.cpp) and BinningHIP.
#include "binning.h"

Binning::Binning(System* s) : system(s) {
  nbinx = nbiny = nbinz = nhalo = 0;
  minx = maxx = miny = maxy = minz = maxz = 0;
  bincount = binoffsets = permute_vector = nullptr;
  is_sorted = false;
}
void Binning::create_binning(T_X_FLOAT, T_X_FLOAT, T_X_FLOAT, int, bool, bool, bool) {}
const char* Binning::name() { return "BinningNone"; }

BinningHIP::BinningHIP(System* s) : Binning(s) { memset(&b, 0, sizeof(b)); }
BinningHIP::~BinningHIP() { if (system && system->ctx) emd_bins_free(system->ctx, &b); }

// BinningKKSort::create_binning, src/binning_types/binning_kksort.cpp:71-140
void BinningHIP::create_binning(T_X_FLOAT dx, T_X_FLOAT dy, T_X_FLOAT dz, int halo_depth,
                                bool do_local, bool do_ghost, bool sort) {
  if (!(do_local || do_ghost)) return;
  EMD_CALL(system->ctx, emd_binning(system->ctx, system->view(), dx, dy, dz, halo_depth,
                                    do_local, do_ghost, sort, &b));
  nbinx = b.nbinx; nbiny = b.nbiny; nbinz = b.nbinz; nhalo = b.nhalo;
  minx = b.minx; maxx = b.maxx; miny = b.miny; maxy = b.maxy; minz = b.minz; maxz = b.maxz;
  bincount = b.bincount; binoffsets = b.binoffsets; permute_vector = b.permute;
  is_sorted = sort;
}
const char* BinningHIP::name() { return "BinningHIP"; }
