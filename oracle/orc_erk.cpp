// Oracle: explicit Runge-Kutta evolution of fEuler -- TEST INFRASTRUCTURE ONLY (see
// orc_common.hpp).
//
// Fixed steps of an explicit Butcher table, as ARKStep's explicit stage loop does it
// (SUNDIALS 6.2.0, arkode_erkstep.c, erkStep_TakeStep / erkStep_ComputeSolutions; the
// library is not part of the reference tree): stage i evaluates F_i = f(z_i) with
//   z_i = y_n + sum_{j<i} (h a_ij) F_j,   y_{n+1} = y_n + sum_i (h b_i) F_i,
// each sum a linear combination taken in the order written (N_VLinearCombination). The
// configuration the reference runs at order 3 uses ARKStep's default third-order explicit
// table, Bogacki-Shampine 3(2) (4 stages). fEuler is time-independent here (no forcing).
#include <cstring>
#include <vector>
#include "oracle.h"

extern "C" int orc_erk_evolve(const orc_grid *g, const orc_erk *tab, double h, int nsteps,
                              double *const *w) {
  const long ncell = g->nx * g->ny * g->nz;
  const int nchem = g->nchem, s = tab->stages;
  if (s < 1 || s > 6) return -2;
  const long len[6] = {ncell, ncell, ncell, ncell, ncell, (long)nchem * ncell};
  const int nsub = nchem > 0 ? 6 : 5;
  // F[i][q]: stage derivative i, subvector q; z: stage state
  std::vector<std::vector<double>> F((size_t)s * 6), z(6);
  for (int q = 0; q < nsub; q++) {
    z[q].assign(len[q], 0.0);
    for (int i = 0; i < s; i++) F[(size_t)i * 6 + q].assign(len[q], 0.0);
  }
  for (int n = 0; n < nsteps; n++) {
    for (int i = 0; i < s; i++) {
      const double *zin[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
      double *fout[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
      for (int q = 0; q < nsub; q++) {
        double *zq = z[q].data();
        for (long k = 0; k < len[q]; k++) {
          double acc = w[q][k];
          for (int j = 0; j < i; j++) acc += (h * tab->A[i * 6 + j]) * F[(size_t)j * 6 + q][k];
          zq[k] = acc;
        }
        zin[q] = zq;
        fout[q] = F[(size_t)i * 6 + q].data();
      }
      const int rv = orc_fe(g, zin, fout);
      if (rv != 0) return rv;
    }
    for (int q = 0; q < nsub; q++)
      for (long k = 0; k < len[q]; k++) {
        double acc = w[q][k];
        for (int i = 0; i < s; i++) acc += (h * tab->b[i]) * F[(size_t)i * 6 + q][k];
        w[q][k] = acc;
      }
  }
  return 0;
}
