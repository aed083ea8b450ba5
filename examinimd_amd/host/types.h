// types.h -- enums and scalar types of the host-side module API.
// Mirrors src/types.h:43-85 of the reference (same enum names and values so
// that decks, CLI strings and module headers keep their meaning); the Kokkos
// View typedefs are replaced by device pointers held in emd_system.
#ifndef EMD_TYPES_H
#define EMD_TYPES_H

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "emd_hip.h"

// Units to be used
enum {UNITS_REAL, UNITS_LJ, UNITS_METAL};
// Lattice Type (LATTICE_BCC is an extension)
enum {LATTICE_SC, LATTICE_FCC, LATTICE_BCC};
// Integrator Type
enum {INTEGRATOR_NVE};
// Binning Type
enum {BINNING_KKSORT};
// Comm Type
enum {COMM_SERIAL, COMM_MPI};
// Force Type
enum {FORCE_LJ, FORCE_LJ_IDIAL, FORCE_SNAP};
// Force Iteration Type
enum {FORCE_ITER_CELL_FULL, FORCE_ITER_NEIGH_FULL, FORCE_ITER_NEIGH_HALF};
// Neighbor Type
enum {NEIGH_NONE, NEIGH_CSR, NEIGH_CSR_MAPCONSTR, NEIGH_2D};
// Input File Type
enum {INPUT_LAMMPS};
// Force precision (extension): EMD_FP64 / EMD_MIXED from emd_hip.h

#define MAX(a,b) (a>b?a:b)
#define MIN(a,b) (a<b?a:b)

#define MAX_TYPES_STACKPARAMS 12

#ifndef T_INT
#define T_INT int
#endif
#ifndef T_FLOAT
#define T_FLOAT double
#endif
#define T_X_FLOAT T_FLOAT
#define T_V_FLOAT T_FLOAT
#define T_F_FLOAT T_FLOAT

// abort with a message (library code returns status codes; the driver exits)
void emd_host_fatal(const char* what, emd_ctx* ctx, int status);
#define EMD_CALL(ctx, expr) do { int st_ = (expr); if (st_ != EMD_OK) emd_host_fatal(#expr, (ctx), st_); } while (0)

#endif
