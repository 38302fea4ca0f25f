/*
 * wpic_io.h -- the reference's raw + JSON output files at one rank, written
 * and read on the host (libwpic.so).  Each entry point replaces the paraio
 * routine named beside it and produces the same file pair, byte for byte:
 * <dir><name>.raw holds the values back to back in little-endian order, and
 * <dir><name>.json describes them (utils/iocore/jsonio.f90:128-170) in
 * JSON-Fortran's print layout, readable by python/json2hdf5.py.
 *
 * Arrays are the reference's own Fortran layouts:
 *   2D up(ndim, np, nys:nye, nsp), uf(6, nxgs-2:nxge+2, nys-2:nye+2),
 *      np2(nys:nye, nsp), mom(7, nxgs-1:nxge+1, nys-1:nye+1, nsp)
 *   3D up(ndim, np, nys:nye, nzs:nze, nsp), uf(6, nxgs-2:nxge+2,
 *      nys-2:nye+2, nzs-2:nze+2), np2(nys:nye, nzs:nze, nsp),
 *      mom(7, nxgs-1:nxge+1, nys-1:nye+1, nzs-1:nze+1, nsp)
 *
 * Single rank: nproc = 1, nrank = 0, nys = nygs, nye = nyge (nzs = nzgs,
 * nze = nzge).  `dir` is prefixed to the file names verbatim, as the
 * reference's trim(dir)//name.  Calls return 0 or a negative WPIC_E* code of
 * wpic.h, with wpic_last_error() holding the reference's message where it
 * would stop.
 */
#ifndef WPIC_IO_H
#define WPIC_IO_H

#include "wpic.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wpic_paraio wpic_paraio;

/* paraio__init (2d/common/paraio.f90:37-78) */
int wpic_paraio_init2d(wpic_paraio **io, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge,
                       int nys, int nye, int nproc, int nrank, double delx, double delt, double c,
                       const double *q, const double *r, const char *dir);
/* paraio__init (3d/common/paraio.f90:37-85; nproc_j = nproc_k = 1) */
int wpic_paraio_init3d(wpic_paraio **io, int ndim, int np, int nsp, int nxgs, int nxge, int nygs, int nyge,
                       int nzgs, int nzge, int nys, int nye, int nzs, int nze, int nproc, int nrank,
                       double delx, double delt, double c, const double *q, const double *r, const char *dir);
/* paraio__finalize (2d/common/paraio.f90:83-89) */
int wpic_paraio_finalize(wpic_paraio *io);

/* paraio__output (2d/common/paraio.f90:93-228: the padded up, np2, uf with
 * ghosts; 3d/common/paraio.f90:102-283: np2, compact particles per species,
 * poffset, uf with ghosts) -> <filename>.raw/.json */
int wpic_paraio_output(wpic_paraio *io, const double *up, const double *uf, const int *np2, int nxs, int nxe,
                       int it, const char *filename);
/* paraio__input (2d/common/paraio.f90:233-426, 3d/common/paraio.f90:288-527):
 * the grid, np, nsp and nproc of the file must match the initialised ones;
 * delx, delt, c, r, q are taken from the file */
int wpic_paraio_input(wpic_paraio *io, double *up, double *uf, int *np2, int *nxs, int *nxe, int *it,
                      const char *filename);
/* paraio__param (2d/common/paraio.f90:431-550, 3d/common/paraio.f90:532-655) */
int wpic_paraio_param(wpic_paraio *io, int n0, double wpe, double wpi, double wge, double wgi, double vti,
                      double vte, const char *filename);
/* the shock app's save_param (2d/proj/shock/app.f90:567-612): the pair of
 * wpic_paraio_param with the attribute u0 appended */
int wpic_paraio_param_shock(wpic_paraio *io, int n0, double wpe, double wpi, double wge, double wgi, double vti,
                            double vte, double u0, const char *filename);
/* paraio__mom (2d/common/paraio.f90:555-713, 3d/common/paraio.f90:660-838)
 * -> <it>_mom.raw/.json.  Divides mom(2:7) by mom(1) in place, as the
 * reference does. */
int wpic_paraio_mom(wpic_paraio *io, double *mom, const double *uf, int it);
/* paraio__ptcl / paraio__orb (2d/common/paraio.f90:718-741 with
 * write_particle :746-893; 3d/common/paraio.f90:843-1002): all active
 * particles, or those whose id read as a 64-bit integer is positive
 * -> <it>_ptcl / <it>_orb */
int wpic_paraio_ptcl(wpic_paraio *io, const double *up, const double *uf, const int *np2, int it);
int wpic_paraio_orb(wpic_paraio *io, const double *up, const double *uf, const int *np2, int it);

#ifdef __cplusplus
}
#endif
#endif
