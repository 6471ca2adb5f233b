/*
 * Link-time stand-in for the one routine the reference's serial build calls
 * but never defines (checker infrastructure: linked only into oracle/_ref
 * and the drop-in executable).
 *
 * phi_3D_exch is the MPI halo exchange of a 3-D field; les_dyn calls it
 * outside any PARALLEL guard (src/nhwave.F:8077), so a serial build leaves
 * it undefined.  The LES closure is never run by any input of this build;
 * should it be, the run stops here with a message instead of going on with
 * unexchanged halos.
 */
#include <stdio.h>
#include <stdlib.h>

void phi_3d_exch_(void *phi) {
  (void)phi;
  fputs("nhwave reference build: phi_3D_exch (LES path, src/nhwave.F:8077) is not available in the "
        "serial build\n", stderr);
  exit(86);
}
