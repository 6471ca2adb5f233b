/*
 * The one routine the reference's serial build calls but never defines,
 * for the drop-in executable (the reference's driver linked against
 * libnhwave_amd).
 *
 * phi_3D_exch is the MPI halo exchange of a 3-D field; les_dyn calls it
 * outside any PARALLEL guard (src/nhwave.F:8077), so a serial build leaves
 * it undefined.  The LES closure is off the GPU path and refused by the
 * library; should a run reach it anyway, it stops here with a message
 * instead of going on with unexchanged halos.
 */
#include <stdio.h>
#include <stdlib.h>

void phi_3d_exch_(void *phi) {
  (void)phi;
  fputs("nhwave_gpu: phi_3D_exch (LES path, src/nhwave.F:8077) is not available in the serial build\n",
        stderr);
  exit(86);
}
