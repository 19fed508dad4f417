/*
 * amr_hip_dropin.h -- replacements of the reference physics module's outer
 * entry points, for a driver that links libamrhip.so INSTEAD of the
 * reference's physics module (this repository's own evolve loops and tests).
 *
 * Never include this beside the reference's physics headers: the reference
 * defines hpc_amr_evolve_solver with an MPI_Comm argument
 * (src/physics/src/hpc_amr_evolve_solver.c:16-41), these take the
 * communicator as an opaque pointer.  A p4est driver that keeps the
 * reference's physics module reaches the GPU through the seam instead
 * (physics_go_gpu.h).  The library is linked -Bsymbolic, so its own calls
 * never resolve to a reference definition of the same name loaded in the
 * same process.
 */
#ifndef AMR_HIP_DROPIN_H
#define AMR_HIP_DROPIN_H

#include "amr_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces hpc_amr_initialize_solver (src/physics/src/hpc_amr_initialize_solver.c:16-28
 * -> physics_initialize, src/physics/src/physics_initialize.c:16-58): reads
 * input.physics from the working directory with the same keywords, defaults,
 * CFL rescale and interval fix-up. */
void hpc_amr_initialize_solver(int *rank, int *dim, double *mesh_scale, int *nlevels,
                               int *nfields, int *nfringe, int *nelem_subgrid,
                               int *extern_cell_size, int *extern_face_size,
                               int *extern_edge_size, int *extern_node_size,
                               int *extern_mpi_comm_size);

/* Replaces hpc_amr_evolve_solver (src/physics/src/hpc_amr_evolve_solver.c:16-49,
 * declared at src/amr/src/amr_external.h:124-130); forwards to physics_go_gpu. */
int hpc_amr_evolve_solver(int *dim, int *unstructured, int *nfields, int *nelem_subgrid,
                          int *ncell, int *nghost, int *nface, int *dof,
                          int *cell_info, int *face_info, int *nsend, int *nrecv,
                          int *send, int *recv, int *sizes, double *geom, double *soln,
                          int *istep, int *regrid, int *visualize, int *checkpoint,
                          int *evolve_solution, void *mpicomm_ext,
                          double *total_time, double *compute_time, double *residual_time);

#ifdef __cplusplus
}
#endif
#endif
