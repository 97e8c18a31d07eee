/*
 * jxamd_jxfpamg.c -- JXFPAMG_Solver_Interface with the reference's own
 * signature (src/jxpamg_interface.c:118-119; JXF_Int = int and JXF_Real =
 * double in the reference's FP64 build), so a caller written against the
 * reference links this library in place of the reference's interface object.
 *
 * The solve phase runs on one GPU in one process: the communicator must hold
 * a single rank (else JXAMD_ERROR_ARG, nothing touched).  solver_id 220
 * (CPR-GMRES, src/jxpamg_interface.c:1967-2043) is forwarded to
 * jxamd_solver_interface with the same argument meanings (row, ia/ja/a with
 * 1-based indices decremented in place, ser_x/ser_b, tol, num_functions,
 * *iter); other solver ids return JXAMD_ERROR_ARG.  Built only where an
 * mpi.h is available (the image's MPICH).
 */
#include <mpi.h>

#include "jxamd.h"
#include "jxamd_jxfpamg.h"

int JXFPAMG_Solver_Interface(int argc, char *argv[], MPI_Comm comm, int row, int *ia, int *ja, double *a,
                             double *ser_x, double *ser_b, double tol, int solver_id, int num_functions, int *iter)
{
    int inited = 0, np = 1;
    if (MPI_Initialized(&inited) != MPI_SUCCESS) return JXAMD_ERROR_GENERIC;
    if (inited && MPI_Comm_size(comm, &np) != MPI_SUCCESS) return JXAMD_ERROR_GENERIC;
    if (np != 1) return JXAMD_ERROR_ARG;
    return jxamd_solver_interface(argc, argv, row, ia, ja, a, ser_x, ser_b, tol, solver_id, num_functions, iter);
}
