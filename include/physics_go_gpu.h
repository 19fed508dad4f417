/*
 * physics_go_gpu.h -- the header the reference's GPU seam includes
 * (src/physics/src/physics_evolve_interface.h:16-17, under _GPU_) and does not
 * ship.  Put this directory on the include path of the reference's physics
 * module, compile it with -D_GPU_ and link libamrhip.so:
 * physics_evolve_interface (src/physics/src/physics_evolve_interface.c:42-46)
 * then calls the MI355X implementation with physics_go's argument list.
 *
 * The library takes the communicator as an opaque pointer (one GPU owns the
 * whole forest; the communicator is not used).  The wrapper below keeps the
 * reference's MPI_Comm argument and restores its NaN behaviour,
 * MPI_Abort(comm, 3) (src/physics/src/physics_runge_kutta_interface.c:352-355).
 * It declares no other function of the reference's physics module.
 */
#ifndef PHYSICS_GO_GPU_H
#define PHYSICS_GO_GPU_H

#include <mpi.h>

#include "amr_hip.h"

static inline void physics_go_gpu_mpi(int *dim, int *unstructured, int *nfields,
                                      int *nelem_subgrid, int *ncell, int *nghost, int *nface,
                                      int *dof, int *cell_info, int *face_info, int *nsend,
                                      int *nrecv, int *send, int *recv, int *sizes, double *geom,
                                      double *soln, int *istep, int *regrid, int *visualize,
                                      int *checkpoint, int *evolve_solution, MPI_Comm comm,
                                      double *total_time, double *compute_time,
                                      double *residual_time)
{
    int rc = physics_go_gpu(dim, unstructured, nfields, nelem_subgrid, ncell, nghost, nface, dof,
                            cell_info, face_info, nsend, nrecv, send, recv, sizes, geom, soln,
                            istep, regrid, visualize, checkpoint, evolve_solution, (void *)&comm,
                            total_time, compute_time, residual_time);
    if (rc == AMRGPU_ERR_NAN) MPI_Abort(comm, 3);
}
#define physics_go_gpu(...) physics_go_gpu_mpi(__VA_ARGS__)

#endif
