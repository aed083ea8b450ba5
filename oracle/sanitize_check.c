/* Exercises the CPU oracle (test infrastructure) end to end for a build with
 * AddressSanitizer and UndefinedBehaviorSanitizer: LJ with the CSR, half
 * (newton on) and 2D lists, lj/cut/idial, SNAP at twojmax 8 and 12, the
 * Clebsch-Gordan table, and a dump written and compared.  Any invalid access,
 * leak or undefined operation aborts with the sanitizer's report; the exit
 * status is 0 only if every run finished with finite energies.
 *   gcc -fsanitize=address,undefined -fno-sanitize-recover=all -fopenmp \
 *       oracle.c oracle_snap.c sanitize_check.c -lm */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static or_config base_config(void) {
  or_config c;
  memset(&c, 0, sizeof(c));
  c.units = OR_UNITS_LJ;
  c.lattice_style = OR_LATTICE_FCC;
  c.lattice_arg = 0.8442;
  c.nx = c.ny = c.nz = 4;
  c.ntypes = 1;
  c.mass[0] = 1.0;
  c.temperature = 1.44;
  c.seed = 87287;
  c.force_type = OR_FORCE_LJ;
  c.force_cutoff = 2.5;
  c.lj_eps = 1.0;
  c.lj_sigma = 1.0;
  c.lj_cut = 2.5;
  c.skin = 0.3;
  c.exchange_rate = 10;
  c.neigh_type = OR_NEIGH_CSR;
  c.lj_intensity = 1.0;
  c.nthreads = 2;
  return c;
}

static int run(const char *what, const or_config *c, const or_snap_params *sp, int steps,
               const char *dump) {
  or_md *m = or_md_create(c, sp);
  if (!m) {
    fprintf(stderr, "%s: create failed\n", what);
    return 1;
  }
  int bad = or_md_init(m) != 0 || or_md_run(m, steps) != 0;
  double th[3];
  or_md_thermo(m, th);
  bad = bad || !isfinite(th[1]) || !isfinite(th[2]);
  if (dump && !bad) {
    double d[6];
    bad = or_md_dump(m, dump) != 0 || or_dump_compare(or_md_system(m), dump, d) != 0;
    for (int k = 0; k < 6 && !bad; k++) bad = d[k] != 0.0;
  }
  printf("%-28s T %.6f PE/N %.6f E/N %.6f %s\n", what, th[0], th[1], th[2], bad ? "FAIL" : "ok");
  or_md_destroy(m);
  return bad;
}

static or_snap_params snap(int twojmax) {
  or_snap_params p;
  memset(&p, 0, sizeof(p));
  p.twojmax = twojmax;
  p.rfac0 = 0.99363;
  p.rcutfac = 4.73442;
  p.switchflag = 1;
  p.radelem = 0.5;
  p.wjelem = 1.0;
  p.ncoeff = or_snap_ncoeff(twojmax);
  srand(2024 + twojmax);
  for (int k = 1; k <= p.ncoeff; k++) p.coeff[k] = 0.05 * ((double)rand() / RAND_MAX - 0.5);
  return p;
}

int main(int argc, char **argv) {
  const char *dump = argc > 1 ? argv[1] : "sanitize_check.dump";
  int bad = 0;
  or_config c = base_config();
  bad |= run("LJ CSR full", &c, NULL, 25, dump);
  c.half_neigh = 1;
  c.newton = 1;
  bad |= run("LJ CSR half, newton on", &c, NULL, 25, NULL);
  c.half_neigh = 0;
  c.newton = 0;
  c.neigh_type = OR_NEIGH_2D;
  bad |= run("LJ 2D full", &c, NULL, 25, NULL);
  c.neigh_type = OR_NEIGH_CSR;
  c.force_type = OR_FORCE_LJ_IDIAL;
  c.lj_intensity = 7.0;
  bad |= run("lj/cut/idial 7", &c, NULL, 10, NULL);

  or_config s = base_config();
  s.units = OR_UNITS_METAL;
  s.lattice_style = OR_LATTICE_BCC;
  s.lattice_arg = 3.1803;
  s.mass[0] = 183.84;
  s.temperature = 300.0;
  s.force_type = OR_FORCE_SNAP;
  s.force_cutoff = 4.73442;
  s.skin = 1.0;
  s.newton = 1;
  or_snap_params p8 = snap(8);
  bad |= run("SNAP twojmax 8", &s, &p8, 3, NULL);
  or_snap_params p12 = snap(12);
  bad |= run("SNAP twojmax 12", &s, &p12, 1, NULL);

  double *cg = malloc(sizeof(double) * 13 * 13 * 13 * 13 * 13);
  or_snap_cg_all(12, cg);
  bad |= !isfinite(cg[((((12 * 13 + 12) * 13) + 12) * 13 + 6) * 13 + 6]);
  free(cg);
  printf("%s\n", bad ? "FAILED" : "all runs clean");
  return bad;
}
