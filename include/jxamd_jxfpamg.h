/*
 * jxamd_jxfpamg.h -- the reference's own solve entry, for callers written
 * against it (libjxamd_jxfpamg.so; needs mpi.h).
 */
#ifndef JXAMD_JXFPAMG_H
#define JXAMD_JXFPAMG_H

#include <mpi.h>

#ifdef __cplusplus
extern "C" {
#endif

/* JXFPAMG_Solver_Interface (src/jxpamg_interface.c:118-119) with JXF_Int =
 * int and JXF_Real = double (the reference's FP64 build).  One rank only (a
 * communicator of more ranks returns JXAMD_ERROR_ARG); solver_id 220 =
 * CPR-GMRES on the GPU (jxamd_solver_interface); other ids return
 * JXAMD_ERROR_ARG.  MPI need not be initialised. */
int JXFPAMG_Solver_Interface(int argc, char *argv[], MPI_Comm comm, int row, int *ia, int *ja, double *a,
                             double *ser_x, double *ser_b, double tol, int solver_id, int num_functions, int *iter);

#ifdef __cplusplus
}
#endif
#endif /* JXAMD_JXFPAMG_H */
