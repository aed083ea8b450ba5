// emd_snap.hip -- linear SNAP on the GPU (placeholder until the kernels land).
#include "emd_internal.hpp"

struct emd_snap {
  int dummy;
};

extern "C" emd_status emd_snap_create(emd_ctx *ctx, const emd_snap_params *, emd_snap **out) {
  if (out) *out = nullptr;
  return emd_set_error(ctx, EMD_ERR_STATE, "SNAP kernels are not built yet");
}
extern "C" emd_status emd_snap_destroy(emd_ctx *, emd_snap *) { return EMD_OK; }
extern "C" emd_status emd_snap_set_coeff(emd_ctx *ctx, emd_snap *, const double *) {
  return emd_set_error(ctx, EMD_ERR_STATE, "SNAP kernels are not built yet");
}
extern "C" emd_status emd_snap_force(emd_ctx *ctx, emd_system *, const emd_neigh *, emd_snap *,
                                     int) {
  return emd_set_error(ctx, EMD_ERR_STATE, "SNAP kernels are not built yet");
}
extern "C" emd_status emd_snap_energy(emd_ctx *ctx, const emd_system *, const emd_neigh *,
                                      emd_snap *, double *, double *) {
  return emd_set_error(ctx, EMD_ERR_STATE, "SNAP kernels are not built yet");
}
