/* Solution output, restart input and the restart parameter file in the reference's on-disk
 * format (io.cpp:645-1162), for one rank: libmvd_io.so (host code over serial HDF5; kept
 * apart from libmvd_rhs.so so that the RHS library does not depend on HDF5).
 *
 * output-%07d.hdf5 holds the datasets
 *   "time"    scalar double,   "nchem"  scalar int,
 *   "domain"  double [3][2] = {zl, zr, yl, yr, xl, xr},
 *   "Density", "x-Momentum", "y-Momentum", "z-Momentum", "TotalEnergy"
 *             double [nz][ny][nx] in CGS (code values times DensityUnits, MomentumUnits,
 *             EnergyUnits = M/L^3, M/L^2/T, M/L/T^2, euler3D.hpp:389-391),
 *   "Chemical-000" ... one double [nz][ny][nx] per species, as stored (no unit change).
 * The field arrays passed in and out are the ManyVector blocks of mvd_rhs.h: w[0..4] the
 * fluid fields (i + nx (j + ny k)), w[5] the species (v + nchem (i + nx (j + ny k))).
 */
#ifndef MVD_IO_H
#define MVD_IO_H

#include "mvd_rhs.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MVD_IO_ERROR -6   /* file could not be created, opened, read or written */
#define MVD_IO_MISMATCH -7  /* restart file incompatible with the grid (read_restart's checks) */

/* The ARKODE and driver options that restart_parameters.txt records besides the mvd_params
 * fields (ARKODEParameters euler3D.hpp:126-169, and EulerData's tf, nout, showstats). */
typedef struct {
  double tf;
  int nout, showstats;
  int order, dense_order, etable, itable, mtable;
  int adapt_method, maxnef, mxhnil, mxsteps;
  double safety, bias, growth;
  int pq;
  double k1, k2, k3, etamx1, etamxf;
  double hmin, hmax;
  int fixedstep;
  double htrans, rtol, atol;
  int fusedkernels, localreduce;
} mvd_run_options;

/* write_parameters (io.cpp:645-714): dir/restart_parameters.txt with t0 = tcur, h0 = hcur,
 * nout = o->nout - iout and restart = iout, doubles as %.16e. dir NULL: the working
 * directory. Returns 0, MVD_BAD_ARGUMENT or MVD_IO_ERROR. */
int mvd_write_parameters(const char *dir, const mvd_params *p, const mvd_run_options *o,
                         double tcur, double hcur, int iout);

/* output_solution (io.cpp:723-941): the parameter file (when o is not NULL), then
 * dir/output-<iout>.hdf5. on_device: w holds device pointers, copied to the host with
 * hipMemcpy (ordered after work on the null stream only: work producing w on another
 * stream must have completed), else host pointers. w is not modified.
 * Returns 0, MVD_BAD_ARGUMENT, MVD_DEVICE_ERROR or MVD_IO_ERROR. */
int mvd_output_solution(const char *dir, const mvd_params *p, const mvd_run_options *o,
                        double tcur, double hcur, int iout, const double *const *w,
                        int on_device);

/* read_restart (io.cpp:946-1162): t and w from dir/output-<restart>.hdf5, fluid fields
 * converted back to code units (times 1/Units). Rejects (MVD_IO_MISMATCH) a file whose
 * "time" is not double, whose nchem differs, whose domain differs by more than 1e-14 or
 * whose fields are not [nz][ny][nx]. The metadata are checked before any field is read;
 * a field that fails later leaves the fields before it overwritten (the reference zeroes
 * w first). t is set only on success. */
int mvd_read_restart(const char *dir, const mvd_params *p, int restart, double *t,
                     double *const *w, int on_device);

#ifdef __cplusplus
}
#endif

#endif
