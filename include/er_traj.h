/*
 * er_traj.h -- trajectory input for an erdst host: the DCD frames of
 * HISTORY (the system) and SltConf (solute configurations).
 *
 * erdst reads both through its trajectory module (erdst/trajectory.F90:41-96),
 * which calls the VMD molfile DCD plugin through erdst/vmdfio.c.  These entry
 * points replace that chain for the DCD format with a native reader, built
 * into the same library as er_capi.h.  They are host code only and make no
 * GPU call.
 *
 * Conventions:
 *  - Every function returns ER_OK (0) on success, ER_TRAJ_EOF at the end
 *    of the trajectory, or a nonzero ER_E_* / ER_TRAJ_* status.
 *  - Coordinates are float32 Angstrom, [natoms][3] (x, y, z of each atom),
 *    as erdst's real crd(3, natom).
 *  - The cell is float32 [3][3] column-major: cell[0..2] = a, cell[3..5] = b,
 *    cell[6..8] = c, with a along x and b in the xy plane, formed from the
 *    frame's A, B, C, alpha, beta, gamma as erdst/vmdfio.c:355-378 forms it.
 *    A frame without a unit-cell block gives the plugin's defaults
 *    (A = B = C = 1, angles 90 degrees).
 */
#ifndef ER_TRAJ_H
#define ER_TRAJ_H

#include <stdint.h>

#include "er_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ER_TRAJ_EOF = 100,      /* no further complete frame                     */
    ER_TRAJ_IO = 101,       /* the file cannot be opened or read             */
    ER_TRAJ_FORMAT = 102    /* not a DCD file, or a record of the wrong size */
};

/* DCD flavour bits reported by er_traj_info */
enum {
    ER_TRAJ_CHARMM = 1,     /* CHARMM (or NAMD >= 2.1) header                */
    ER_TRAJ_CELL = 2,       /* a unit-cell block precedes every frame        */
    ER_TRAJ_4D = 4,         /* a fourth coordinate record follows each frame */
    ER_TRAJ_REC64 = 8,      /* 8-byte Fortran record markers                 */
    ER_TRAJ_SWAPPED = 16    /* byte order opposite to the host's             */
};

typedef struct er_traj er_traj;

/* open_trajectory (erdst/trajectory.F90:41-53) / vmdfio_open_traj: opens a
 * DCD file and reads its header (vmdplugins/dcdplugin.c:155-398, 798-902),
 * returning the atom count.  Fails with ER_TRAJ_FORMAT when the file holds
 * no complete first frame, as the plugin's open does. */
int er_traj_open(const char *path, er_traj **out, int32_t *natoms);

/* read_trajectory (erdst/trajectory.F90:59-82) / vmdfio_read_traj_step
 * (erdst/vmdfio.c:316-385): the next frame.  natoms must equal the file's
 * atom count (vmdfio.c:329-334).  Frames stop at the count the plugin
 * derives from the file size (dcdplugin.c:829-864). */
int er_traj_read(er_traj *t, int32_t natoms, float *xyz, float *cell);

/* The header's fields: frame count (from the file size, as the plugin
 * counts), ISTART, NSAVC, DELTA, the number of fixed atoms, ER_TRAJ_* bits. */
int er_traj_info(const er_traj *t, int32_t *nframes, int32_t *istart, int32_t *nsavc, double *delta,
                 int32_t *nfixed, int32_t *flags);

/* close_trajectory (erdst/trajectory.F90:55-61); NULL is accepted. */
void er_traj_close(er_traj *t);

#ifdef __cplusplus
}
#endif

#endif
