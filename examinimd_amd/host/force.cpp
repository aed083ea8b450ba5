// force.cpp -- Force base (src/force.cpp) and the LJ modules (lj/cut, lj/cut/idial).
#include "force.h"

#include <cmath>

Force::Force(char**, System*, bool half_neigh_) : half_neigh(half_neigh_), comm_newton(false) {}
void Force::init_coeff(int, char**) {}
void Force::compute(System*, Binning*, Neighbor*) {}
const char* Force::name() { return "ForceNone"; }

// ForceLJNeigh ctor, src/force_types/force_lj_neigh_impl.h:43-58
ForceLJNeighHIP::ForceLJNeighHIP(char** args, System* system, bool half_neigh_)
    : Force(args, system, half_neigh_) {
  ntypes = system->ntypes;
  use_stackparams = (ntypes <= MAX_TYPES_STACKPARAMS);
  memset(&params, 0, sizeof(params));
  params.ntypes = ntypes;
  if (ntypes > EMD_MAX_TYPES) {
    fprintf(stderr, "ForceLJNeighHIP: %d atom types, at most %d\n", ntypes, EMD_MAX_TYPES);
    exit(1);
  }
  // beyond the fixed 16 x 16 arrays the coefficients go in a table of their own
  if (ntypes > 16) {
    table.assign(3 * (size_t)ntypes * ntypes, 0.0);
    params.table = table.data();
  }
}

// ForceLJNeigh::init_coeff, force_lj_neigh_impl.h:60-98.  With at most
// MAX_TYPES_STACKPARAMS types the reference writes every (i,j) entry from each
// pair_coeff line, so the last line sets all pairs; kept as is.
void ForceLJNeighHIP::init_coeff(int nargs, char** args) {
  if (nargs < 6) { fprintf(stderr, "pair_coeff lj/cut needs: i j eps sigma cut\n"); exit(1); }
  int t1 = atoi(args[1]) - 1, t2 = atoi(args[2]) - 1;
  double eps = atof(args[3]), sigma = atof(args[4]), cut = atof(args[5]);
  double lj1 = 48.0 * eps * pow(sigma, 12.0);
  double lj2 = 24.0 * eps * pow(sigma, 6.0);
  if (use_stackparams) {
    for (int i = 0; i < ntypes; i++)
      for (int j = 0; j < ntypes; j++) {
        params.lj1[i][j] = lj1;
        params.lj2[i][j] = lj2;
        params.cutsq[i][j] = cut * cut;
      }
  } else if (t1 < 0 || t2 < 0 || t1 >= ntypes || t2 >= ntypes) {
    fprintf(stderr, "pair_coeff: atom type out of range\n");
    exit(1);
  } else if (!params.table) {
    params.lj1[t1][t2] = params.lj1[t2][t1] = lj1;
    params.lj2[t1][t2] = params.lj2[t2][t1] = lj2;
    params.cutsq[t1][t2] = params.cutsq[t2][t1] = cut * cut;
  } else {
    const size_t N = (size_t)ntypes * ntypes;
    const size_t a = (size_t)t1 * ntypes + t2, b = (size_t)t2 * ntypes + t1;
    table[a] = table[b] = lj1;
    table[N + a] = table[N + b] = lj2;
    table[2 * N + a] = table[2 * N + b] = cut * cut;
  }
}

// ForceLJNeigh::compute, force_lj_neigh_impl.h:100-126
// (full lists: the force is stored, which includes the zeroing)
void ForceLJNeighHIP::compute(System* system, Binning*, Neighbor* neighbor) {
  if (!half_neigh)
    EMD_CALL(system->ctx, emd_lj_force_store(system->ctx, system->view(),
                                             neighbor->get_neigh_list(), &params,
                                             system->precision));
  else
    EMD_CALL(system->ctx, emd_lj_force(system->ctx, system->view(), neighbor->get_neigh_list(),
                                       &params, system->precision, 1));
}

// ForceLJNeigh::compute_energy, force_lj_neigh_impl.h:128-156
T_F_FLOAT ForceLJNeighHIP::compute_energy(System* system, Binning*, Neighbor* neighbor) {
  double pe = 0.0;
  EMD_CALL(system->ctx, emd_lj_energy(system->ctx, system->view(), neighbor->get_neigh_list(),
                                      &params, half_neigh ? 1 : 0, &pe));
  return pe;
}

// ForceLJIDialNeigh ctor, src/force_types/force_lj_idial_neigh_impl.h:41-47
ForceLJIDialNeighHIP::ForceLJIDialNeighHIP(char** args, System* system, bool half_neigh_)
    : Force(args, system, half_neigh_) {
  memset(&params, 0, sizeof(params));
  params.lj.ntypes = system->ntypes;
  if (system->ntypes > 16) {
    fprintf(stderr, "ForceLJIDialNeighHIP supports at most 16 types\n");
    exit(1);
  }
}

// ForceLJIDialNeigh::init_coeff, force_lj_idial_neigh_impl.h:49-88: only the
// (t1,t2) and (t2,t1) entries are set
void ForceLJIDialNeighHIP::init_coeff(int nargs, char** args) {
  if (nargs < 7) {
    fprintf(stderr, "pair_coeff lj/cut/idial needs: i j eps sigma cut nrepeat\n");
    exit(1);
  }
  int t1 = atoi(args[1]) - 1, t2 = atoi(args[2]) - 1;
  double eps = atof(args[3]), sigma = atof(args[4]), cut = atof(args[5]);
  int nrepeat = atoi(args[6]);
  params.lj.lj1[t1][t2] = params.lj.lj1[t2][t1] = 48.0 * eps * pow(sigma, 12.0);
  params.lj.lj2[t1][t2] = params.lj.lj2[t2][t1] = 24.0 * eps * pow(sigma, 6.0);
  params.lj.cutsq[t1][t2] = params.lj.cutsq[t2][t1] = cut * cut;
  params.intensity[t1][t2] = params.intensity[t2][t1] = nrepeat;
}

// ForceLJIDialNeigh::compute, force_lj_idial_neigh_impl.h:89-111
void ForceLJIDialNeighHIP::compute(System* system, Binning*, Neighbor* neighbor) {
  if (!half_neigh)
    EMD_CALL(system->ctx, emd_lj_idial_force_store(system->ctx, system->view(),
                                                   neighbor->get_neigh_list(), &params,
                                                   system->precision));
  else
    EMD_CALL(system->ctx, emd_lj_idial_force(system->ctx, system->view(),
                                             neighbor->get_neigh_list(), &params,
                                             system->precision, 1));
}
