// Oracle: slow-RHS wrapper (fslow) -- TEST INFRASTRUCTURE ONLY (see orc_common.hpp).
#include <cstring>
#include <vector>
#include "orc_fe.hpp"
#include "orc_chem.hpp"
#include "oracle.h"

extern "C" void orc_chem_apply_scaling(void *h, double *chem);
extern "C" void orc_chem_unapply_scaling(void *h, double *chem);

// fslow (multirate_chem_hydro_main.cpp:996-1083) / fexpl (imex_chem_hydro_main.cpp:910-1000):
// apply scaling, overwrite et <- ge/EnergyUnits + 0.5/rho*|m|^2 in w, fe(), move the
// total-energy rate into the gas-energy slot (TUnitScale = 1), zero etdot, unapply scaling.
extern "C" int orc_fslow(const orc_grid *g, void *chem, double energy_units,
                         double *const *w, double *const *wdot) {
  const long ncell = g->nx * g->ny * g->nz;
  const int nchem = g->nchem;
  orc_chem_apply_scaling(chem, w[5]);
  const double eus = 1.0 / energy_units;
  for (long c = 0; c < ncell; c++) {
    const double rho = w[0][c], mx = w[1][c], my = w[2][c], mz = w[3][c];
    w[4][c] = w[5][(long)nchem * c + nchem - 1] * eus + 0.5 / rho * (mx * mx + my * my + mz * mz);
  }
  const int rv = orc_fe(g, (const double *const *)w, wdot);
  if (rv != 0) return rv;
  for (long c = 0; c < ncell; c++) {
    wdot[5][(long)nchem * c + nchem - 1] = wdot[4][c] * 1.0;
    wdot[4][c] = 0.0;
  }
  orc_chem_unapply_scaling(chem, w[5]);
  return 0;
}
