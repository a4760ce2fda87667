// Oracle: multirate (MRI) evolution of the reacting flow -- TEST INFRASTRUCTURE ONLY (see
// orc_common.hpp).
//
// The configuration-5 driver (multirate_chem_hydro_main.cpp:529-555: MRIStepCreate with
// fslow, the default third-order coupling, a fixed slow step; the fast stepper an ARKStep
// DIRK on ffast, :898-952, whose forcing is applied to all six subvectors). MRIStep's MIS
// form (SUNDIALS 6.2.0 MRIStepCoupling_MIStoMRI; not in the reference tree) is restated:
// the slow table is Knoth & Wolke's 3-stage explicit method, extended by b as its last row
// (c_4 = 1); stage i (i = 2..4) starts from the running state Y_{i-1}, evaluates
// F_{i-1} = fslow(Y_{i-1}) (which overwrites Y_{i-1}'s total energy, as the reference's
// fslow does) and integrates v' = ffast(v) + r_i over (c_i - c_{i-1}) h with the constant
// forcing r_i = sum_{j<i} ((a_ij - a_{i-1,j}) / (c_i - c_{i-1})) F_j. The fast integration
// here takes ceil((c_i - c_{i-1}) h / hmax) equal fixed steps (hmax: the inner stepper's
// maximum step): the fluid fields (no fast dynamics) advance by dt * r_i, the chemistry by
// one forced DIRK step of every cell. With `consistent`, the chemistry part of each slow RHS
// is converted to the scaled variables the fast stepper integrates (species rates divided by
// their scale, the gas-energy slot -- fslow's total-energy rate in code units -- times
// EnergyUnits divided by the ge scale); without it the reference's units are kept (its
// fslow rates are added unchanged to the scaled fast variables).
#include <cstring>
#include <vector>
#include "oracle.h"

extern "C" int orc_fslow(const orc_grid *g, void *chem, double energy_units, double *const *w,
                         double *const *wdot);
extern "C" int orc_chem_dirk_steps_forced(void *h, const orc_dirk *opt, double hstep,
                                          int nsteps, double *y, const double *force,
                                          int *iters);
extern "C" double *orc_chem_array(void *h, int which);

// Knoth-Wolke 3 with b appended: c, A (row-major, 4 x 4)
static const double KW_C[4] = {0.0, 1.0 / 3.0, 3.0 / 4.0, 1.0};
static const double KW_A[4][4] = {{0.0, 0.0, 0.0, 0.0},
                                  {1.0 / 3.0, 0.0, 0.0, 0.0},
                                  {-3.0 / 16.0, 15.0 / 16.0, 0.0, 0.0},
                                  {1.0 / 6.0, 3.0 / 10.0, 8.0 / 15.0, 0.0}};

extern "C" void orc_mri_kw3_table(double *c, double *A) {
  for (int i = 0; i < 4; i++) {
    c[i] = KW_C[i];
    for (int j = 0; j < 4; j++) A[4 * i + j] = KW_A[i][j];
  }
}

extern "C" int orc_mri_fast_steps(int i, double h, double hmax) {
  const double x = (KW_C[i] - KW_C[i - 1]) * h / hmax;
  long k = (long)x;
  if ((double)k < x) k++;
  if (k < 1) k = 1;
  if (k > 1000000) k = 1000000;
  return (int)k;
}

// nslow fixed slow steps of size h; w (6 subvectors, chemistry scaled) advanced in place.
// Returns fe()'s illegal-state result (-1) if any slow RHS met one, else the number of
// failed fast cell steps.
extern "C" int orc_mri_evolve(const orc_grid *g, void *chem, double energy_units,
                              const orc_dirk *fast, double h, int nslow, double hmax,
                              int consistent, double *const *w) {
  const double *scale = orc_chem_array(chem, 0);
  const long n = g->nx * g->ny * g->nz;
  const int nc = g->nchem;
  const long len[6] = {n, n, n, n, n, (long)nc * n};
  std::vector<std::vector<double>> F(3 * 6), R(6);
  for (int q = 0; q < 6; q++) {
    R[q].assign(len[q], 0.0);
    for (int j = 0; j < 3; j++) F[j * 6 + q].assign(len[q], 0.0);
  }
  int failed = 0;
  bool illegal = false;
  for (int step = 0; step < nslow; step++) {
    for (int i = 1; i < 4; i++) {
      double *Fo[6];
      for (int q = 0; q < 6; q++) Fo[q] = F[(i - 1) * 6 + q].data();
      if (orc_fslow(g, chem, energy_units, w, Fo) != 0) illegal = true;
      if (consistent)
        for (long c = 0; c < n; c++) {
          for (int s = 0; s < nc - 1; s++) Fo[5][nc * c + s] /= scale[nc * c + s];
          Fo[5][nc * c + nc - 1] = Fo[5][nc * c + nc - 1] * energy_units / scale[nc * c + nc - 1];
        }
      const double dc = KW_C[i] - KW_C[i - 1];
      for (int q = 0; q < 6; q++)
        for (long k = 0; k < len[q]; k++) {
          double acc = ((KW_A[i][0] - KW_A[i - 1][0]) / dc) * F[0 * 6 + q][k];
          for (int j = 1; j < i; j++) acc += ((KW_A[i][j] - KW_A[i - 1][j]) / dc) * F[j * 6 + q][k];
          R[q][k] = acc;
        }
      const int nfast = orc_mri_fast_steps(i, h, hmax);
      const double dt = dc * h / nfast;
      // the fluid fields and the chemistry do not interact within a slow stage: the
      // fluid's nfast additions, then the chemistry's nfast steps in one call (its Jacobian
      // setups by the fast options' max_steps_between_setups)
      for (int m = 0; m < nfast; m++)
        for (int q = 0; q < 5; q++)
          for (long k = 0; k < n; k++) w[q][k] = w[q][k] + dt * R[q][k];
      failed += orc_chem_dirk_steps_forced(chem, fast, dt, nfast, w[5], R[5].data(), nullptr);
    }
  }
  return illegal ? -1 : failed;
}
