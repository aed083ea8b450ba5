// force_snap.cpp -- SNAP force module: .snapcoeff/.snapparam reading
// (ForceSNAP::init_coeff / read_files, src/force_types/
// force_snap_neigh_impl.h:258-587) and the calls into the GPU kernels.
#include "force.h"

#include <cmath>

ForceSNAPHIP::ForceSNAPHIP(char** args, System* system_, bool half_neigh_)
    : Force(args, system_, half_neigh_), system(system_), sn(nullptr), nelements(0), ncoeff(0),
      ncoeffall(0), twojmax(0), diagonalstyle(3), switchflag(1), bzeroflag(1), quadraticflag(0),
      rcutfac(0), rfac0(0.99363), rmin0(0), rcutmax(0) {}

ForceSNAPHIP::~ForceSNAPHIP() {
  if (sn && system && system->ctx) emd_snap_destroy(system->ctx, sn);
}

static void snap_abort(const char* msg) {
  fprintf(stderr, "ForceSNAP: %s\n", msg);
  exit(1);
}

// pair_coeff * * <coefffile> <elem...> <paramfile> <elem per type...>
void ForceSNAPHIP::init_coeff(int narg, char** arg) {
  if (narg < 7) snap_abort("SNAP 1: Incorrect args for pair coefficients");
  nelements = narg - 5 - system->ntypes;
  if (nelements < 1) snap_abort("SNAP 2: Incorrect args for pair coefficients");
  if (strcmp(arg[1], "*") != 0 || strcmp(arg[2], "*") != 0)
    snap_abort("A Incorrect args for pair coefficients");
  elements.clear();
  for (int i = 0; i < nelements; i++) elements.push_back(arg[4 + i]);
  read_files(arg[3], arg[4 + nelements]);
  if (!quadraticflag) ncoeff = ncoeffall - 1;
  else snap_abort("quadraticflag is not supported");
  if (diagonalstyle != 3) snap_abort("only diagonalstyle 3 is supported");
  for (int i = 1; i <= system->ntypes; i++) {
    const char* e = arg[5 + nelements + i - 1];
    int k = 0;
    while (k < nelements && elements[k] != e) k++;
    if (k == nelements && strcmp(e, "NULL") != 0)
      snap_abort("Incorrect args for pair coefficients");
    if (k != 0) snap_abort("only a single SNAP element is supported");
  }
  rcutmax = 0.0;
  for (int e = 0; e < nelements; e++) rcutmax = MAX(2.0 * radelem[e] * rcutfac, rcutmax);

  emd_snap_params p;
  memset(&p, 0, sizeof(p));
  p.twojmax = twojmax;
  p.rfac0 = rfac0;
  p.rmin0 = rmin0;
  p.rcutfac = rcutfac;
  p.switchflag = switchflag;
  p.bzeroflag = bzeroflag;
  p.radelem = radelem[0];
  p.wjelem = wjelem[0];
  p.ncoeff = ncoeff;
  p.coeff = coeffelem[0].data();
  if (sn) emd_snap_destroy(system->ctx, sn);
  int st = emd_snap_create(system->ctx, &p, &sn);
  if (st != EMD_OK) emd_host_fatal("emd_snap_create", system->ctx, st);
}

// ForceSNAP::read_files, force_snap_neigh_impl.h:344-587
void ForceSNAPHIP::read_files(const char* coefffilename, const char* paramfilename) {
  FILE* fp = fopen(coefffilename, "r");
  if (!fp) {
    fprintf(stderr, "Cannot open SNAP coefficient file %s\n", coefffilename);
    exit(1);
  }
  char line[1024], *ptr;
  int nwords = 0;
  while (nwords == 0) {
    if (!fgets(line, sizeof(line), fp)) break;
    if ((ptr = strchr(line, '#'))) *ptr = '\0';
    else if (line[0] != 10) nwords = 2;
  }
  if (nwords != 2) snap_abort("Incorrect format in SNAP coefficient file");
  char* w0 = strtok(line, "' \t\n\r\f");
  char* w1 = strtok(NULL, "' \t\n\r\f");
  if (!w0 || !w1) snap_abort("Incorrect format in SNAP coefficient file");
  int nelemfile = atoi(w0);
  ncoeffall = atoi(w1);
  radelem.assign(nelements, 0.0);
  wjelem.assign(nelements, 0.0);
  coeffelem.assign(nelements, std::vector<double>(ncoeffall, 0.0));
  std::vector<int> found(nelements, 0);
  for (int ie = 0; ie < nelemfile; ie++) {
    if (!fgets(line, sizeof(line), fp)) snap_abort("Incorrect format in SNAP coefficient file");
    char* name = strtok(line, "' \t\n\r\f");
    char* rad = strtok(NULL, "' \t\n\r\f");
    char* wj = strtok(NULL, "' \t\n\r\f");
    if (!name || !rad || !wj) snap_abort("Incorrect format in SNAP coefficient file");
    int k = 0;
    while (k < nelements && elements[k] != name) k++;
    if (k == nelements || found[k]) {
      for (int c = 0; c < ncoeffall; c++)
        if (!fgets(line, sizeof(line), fp)) break;
      continue;
    }
    found[k] = 1;
    radelem[k] = atof(rad);
    wjelem[k] = atof(wj);
    for (int c = 0; c < ncoeffall; c++) {
      if (!fgets(line, sizeof(line), fp)) snap_abort("Incorrect format in SNAP coefficient file");
      char* w = strtok(line, "' \t\n\r\f");
      coeffelem[k][c] = w ? atof(w) : 0.0;
    }
  }
  fclose(fp);

  int rcutfacflag = 0, twojmaxflag = 0;
  rfac0 = 0.99363; rmin0 = 0.0; diagonalstyle = 3; switchflag = 1; bzeroflag = 1; quadraticflag = 0;
  fp = fopen(paramfilename, "r");
  if (!fp) {
    fprintf(stderr, "Cannot open SNAP parameter file %s\n", paramfilename);
    exit(1);
  }
  while (fgets(line, sizeof(line), fp)) {
    if ((ptr = strchr(line, '#'))) { *ptr = '\0'; continue; }
    if (line[0] == 10) continue;
    char* keywd = strtok(line, "' \t\n\r\f");
    char* keyval = strtok(NULL, "' \t\n\r\f");
    if (!keywd) continue;
    if (!keyval) snap_abort("Incorrect format in SNAP parameter file");
    if (strcmp(keywd, "rcutfac") == 0) { rcutfac = atof(keyval); rcutfacflag = 1; }
    else if (strcmp(keywd, "twojmax") == 0) { twojmax = atoi(keyval); twojmaxflag = 1; }
    else if (strcmp(keywd, "rfac0") == 0) rfac0 = atof(keyval);
    else if (strcmp(keywd, "rmin0") == 0) rmin0 = atof(keyval);
    else if (strcmp(keywd, "diagonalstyle") == 0) diagonalstyle = atoi(keyval);
    else if (strcmp(keywd, "switchflag") == 0) switchflag = atoi(keyval);
    else if (strcmp(keywd, "bzeroflag") == 0) bzeroflag = atoi(keyval);
    else if (strcmp(keywd, "quadraticflag") == 0) quadraticflag = atoi(keyval);
    else snap_abort("Incorrect SNAP parameter file");
  }
  fclose(fp);
  if (rcutfacflag == 0 || twojmaxflag == 0) snap_abort("Incorrect SNAP parameter file");
}

// ForceSNAP::compute, force_snap_neigh_impl.h:159-208
void ForceSNAPHIP::compute(System* system_, Binning*, Neighbor* neighbor) {
  if (comm_newton == false) snap_abort("ForceSNAP requires 'newton on'");
  // zeroes f itself (stores_forces): no separate zeroing launch
  EMD_CALL(system_->ctx, emd_snap_force_zeroed(system_->ctx, system_->view(),
                                               neighbor->get_neigh_list(), sn, system_->precision));
}

T_F_FLOAT ForceSNAPHIP::compute_energy(System* system_, Binning*, Neighbor* neighbor) {
  double e = 0.0;
  EMD_CALL(system_->ctx, emd_snap_energy(system_->ctx, system_->view(), neighbor->get_neigh_list(),
                                         sn, &e, nullptr));
  return e;
}
