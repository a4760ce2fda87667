// Oracle: IMEX additive Runge-Kutta evolution of the reacting flow -- TEST INFRASTRUCTURE
// ONLY (see orc_common.hpp).
//
// The IMEX driver (imex_chem_hydro_main.cpp:337-476: ARKStepCreate(fexpl, fimpl), the
// block-diagonal chemistry linear solver, PostprocessStep after every step) with a fixed
// step. fexpl (:910-1000) is fslow: Dengo scaling, the total-energy overwrite, fEuler, the
// energy rate moved to the gas-energy slot, unscaling; fimpl (:811-848) is TimeUnits times
// calculate_rhs_cvklu on the chemistry subvector, zero on the fluid fields. ARKStep's stage
// algebra (SUNDIALS, not in the reference tree) is restated for a pair of tables (Ae
// explicit, Ai diagonally implicit, common b and c), stage i = 0..s-1:
//   R_i = y_n + sum_{j<i} (h Ae_ij Fe_j + h Ai_ij Fi_j)         (Fi_j zero on the fluid)
//   Ai_ii = 0: z_i = R_i, Fi_i = fimpl(z_i);
//   else the chemistry solves z - h Ai_ii fimpl(z) = R_i (the per-cell Newton/LU of the DIRK
//        oracle with a one-stage table, predictor R_i), Fi_i = (z_i - R_i) / (h Ai_ii),
//        and the fluid fields take z_i = R_i;
//   Fe_i = fexpl(z_i) (which, as in the reference, rewrites z_i's total energy);
//   y_{n+1} = y_n + sum_i (h b_i Fe_i + h b_i Fi_i);
// then PostprocessStep (:1002-1050): scaling, et = ge / EnergyUnits + |m|^2 / (2 rho),
// unscaling. Sums are taken in the order written (per j: the explicit term, then the
// implicit one). With `consistent` the chemistry part of fexpl is converted to the scaled
// variables as in the multirate oracle.
#include <cstring>
#include <vector>
#include "oracle.h"

extern "C" int orc_fslow(const orc_grid *g, void *chem, double energy_units, double *const *w,
                         double *const *wdot);
extern "C" int orc_chem_dirk_step(void *h, const orc_dirk *opt, double hstep, double *y,
                                  int *iters);
extern "C" int orc_chem_rhs(void *h, const double *y, double *ydot);
extern "C" double *orc_chem_array(void *h, int which);
extern "C" void orc_chem_apply_scaling(void *h, double *chem);
extern "C" void orc_chem_unapply_scaling(void *h, double *chem);

// ARK3(2)4L[2]SA of Kennedy & Carpenter (Appl. Numer. Math. 44, 2003), ARKStep's default
// third-order IMEX pair: explicit and ESDIRK tables with a common b (stiffly accurate
// implicit table) and c.
extern "C" void orc_ark324_table(orc_ark *t) {
  std::memset(t, 0, sizeof(*t));
  const double g = 1767732205903.0 / 4055673282236.0;
  t->stages = 4;
  double *E = t->Ae, *I = t->Ai;
  E[1 * 8 + 0] = 1767732205903.0 / 2027836641118.0;
  E[2 * 8 + 0] = 5535828885825.0 / 10492691773637.0;
  E[2 * 8 + 1] = 788022342437.0 / 10882634858940.0;
  E[3 * 8 + 0] = 6485989280629.0 / 16251701735622.0;
  E[3 * 8 + 1] = -4246266847089.0 / 9704473918619.0;
  E[3 * 8 + 2] = 10755448449292.0 / 10357097424841.0;
  I[1 * 8 + 0] = g;
  I[1 * 8 + 1] = g;
  I[2 * 8 + 0] = 2746238789719.0 / 10658868560708.0;
  I[2 * 8 + 1] = -640167445237.0 / 6845629431997.0;
  I[2 * 8 + 2] = g;
  I[3 * 8 + 0] = 1471266399579.0 / 7840856788654.0;
  I[3 * 8 + 1] = -4482444167858.0 / 7529755066697.0;
  I[3 * 8 + 2] = 11266239266428.0 / 11593286722821.0;
  I[3 * 8 + 3] = g;
  for (int j = 0; j < 4; j++) t->b[j] = I[3 * 8 + j];
  t->c[0] = 0.0;
  t->c[1] = 1767732205903.0 / 2027836641118.0;
  t->c[2] = 3.0 / 5.0;
  t->c[3] = 1.0;
}

// nsteps fixed steps of size h in place on w (6 subvectors, chemistry scaled by prepare).
// newton: the Newton/LU controls of the implicit solves (its table is replaced by the
// one-stage backward-Euler form z - hg f(z) = R). Returns -1 if fexpl met an illegal state,
// else the number of failed cell solves.
extern "C" int orc_imex_evolve(const orc_grid *g, void *chem, double energy_units,
                               const orc_ark *tab, const orc_dirk *newton, double h,
                               int nsteps, int consistent, double *const *w) {
  const int S = tab->stages;
  const double *scale = orc_chem_array(chem, 0);
  const long n = (long)g->nx * g->ny * g->nz;
  const int nc = g->nchem;
  const long len[6] = {n, n, n, n, n, (long)nc * n};
  orc_dirk be = *newton;
  be.nstage = 1;
  std::memset(be.A, 0, sizeof(be.A));
  std::memset(be.c, 0, sizeof(be.c));
  be.A[0] = 1.0;
  be.c[0] = 1.0;
  std::vector<std::vector<double>> Fe(S * 6), Fi(S), R(6), Z(6);
  for (int q = 0; q < 6; q++) {
    R[q].assign(len[q], 0.0);
    Z[q].assign(len[q], 0.0);
    for (int i = 0; i < S; i++) Fe[i * 6 + q].assign(len[q], 0.0);
  }
  for (int i = 0; i < S; i++) Fi[i].assign(len[5], 0.0);
  int failed = 0;
  bool illegal = false;
  for (int step = 0; step < nsteps; step++) {
    for (int i = 0; i < S; i++) {
      // R_i
      for (int q = 0; q < 6; q++)
        for (long k = 0; k < len[q]; k++) {
          double acc = w[q][k];
          for (int j = 0; j < i; j++) {
            acc += (h * tab->Ae[i * 8 + j]) * Fe[j * 6 + q][k];
            if (q == 5) acc += (h * tab->Ai[i * 8 + j]) * Fi[j][k];
          }
          R[q][k] = acc;
          Z[q][k] = acc;
        }
      const double aii = tab->Ai[i * 8 + i];
      if (aii == 0.0) {
        orc_chem_rhs(chem, Z[5].data(), Fi[i].data());
        for (long k = 0; k < len[5]; k++) Fi[i][k] = Fi[i][k] * newton->rhs_scale;
      } else {
        const double hg = h * aii;
        failed += orc_chem_dirk_step(chem, &be, hg, Z[5].data(), nullptr);
        for (long k = 0; k < len[5]; k++) Fi[i][k] = (Z[5][k] - R[5][k]) / hg;
      }
      double *zp[6], *fp[6];
      for (int q = 0; q < 6; q++) {
        zp[q] = Z[q].data();
        fp[q] = Fe[i * 6 + q].data();
      }
      if (orc_fslow(g, chem, energy_units, zp, fp) != 0) illegal = true;
      if (consistent)
        for (long c = 0; c < n; c++) {
          for (int s = 0; s < nc - 1; s++) fp[5][nc * c + s] /= scale[nc * c + s];
          fp[5][nc * c + nc - 1] = fp[5][nc * c + nc - 1] * energy_units / scale[nc * c + nc - 1];
        }
    }
    for (int q = 0; q < 6; q++)
      for (long k = 0; k < len[q]; k++) {
        double acc = w[q][k];
        for (int i = 0; i < S; i++) {
          acc += (h * tab->b[i]) * Fe[i * 6 + q][k];
          if (q == 5) acc += (h * tab->b[i]) * Fi[i][k];
        }
        w[q][k] = acc;
      }
    // PostprocessStep
    orc_chem_apply_scaling(chem, w[5]);
    const double eu_inv = 1.0 / energy_units;
    for (long c = 0; c < n; c++) {
      const double r = w[0][c], a = w[1][c], b = w[2][c], d = w[3][c];
      w[4][c] = w[5][(long)nc * c + nc - 1] * eu_inv + 0.5 / r * (a * a + b * b + d * d);
    }
    orc_chem_unapply_scaling(chem, w[5]);
  }
  return illegal ? -1 : failed;
}
