// Oracle: one DIRK step of the per-cell primordial chemistry -- TEST INFRASTRUCTURE ONLY
// (see orc_common.hpp).
//
// The reference advances the chemistry with ARKStep's DIRK stages, SUNDIALS' Newton solver
// and a block-diagonal direct linear solver (primordial_ode_main.cpp:422-536,
// multirate_chem_hydro_main.cpp:898-994,1312-1411). SUNDIALS is not available here, so
// this restates the stage solve the GPU solver implements, cell by cell:
//   * y' = s_rhs * f(y), f = calculate_rhs_cvklu in scaled variables (s_rhs = 1 is frhs,
//     s_rhs = TimeUnits is ffast/fimpl without forcing);
//   * stage i solves z - h a_ii f(z) = r_i,  r_i = y_n + h sum_{j<i} a_ij F_j;
//   * Newton with M = I - h a_ii s_rhs J, J = calculate_jacobian_cvklu at y_n (after f(y_n)
//     refreshed the temperature), factored once per step by 10x10 LU with partial pivoting
//     and reused (modified Newton); if a stage reaches max_iters or its corrections grow, J
//     is re-evaluated at the current iterate, refactored and the stage restarted from it
//     (up to max_setups times);
//   * predictor: the previous stage value (y_n for the first stage);
//   * convergence: ARKODE's test (arkStep_NlsConvTest of ARKODE 5.2 in SUNDIALS 6.2.0, the
//     version the reference pins, README.md:209): with del the WRMS norm of the correction,
//     weights 1/(rtol|z|+atol) of the updated iterate (ARKODE weighs by the step start
//     y_n; with those weights the multirate blast's fast steps fail where a species starts
//     near zero), and m the iteration of the stage solve (0 first),
//     crate = max(0.3 crate, del/del_prev) for m > 0,
//     converged when min(crate, 1) del / newton_tol <= 1, diverging (a failure, as
//     max_iters) when m > 0 and del > 2.3 del_prev; crate starts at 1 at every setup and
//     carries over from stage to stage;
//   * F_i = (z_i - r_i)/(h a_ii) (ARKODE's implicit-stage derivative), y_{n+1} = z_s;
//   * ESDIRK tables (first row zero, e.g. ARK4(3)7L[2]SA's implicit part, the reference
//     inputs' itable 112): the first stage is explicit, z_1 = y_n and F_1 = f(y_n) (plus the
//     forcing), the Jacobian setup at y_n uses the implicit stages' a_ii.
#include <cmath>
#include <cstring>
#include <vector>
#include "orc_chem.hpp"
#include "oracle.h"

using namespace orc;

namespace {

constexpr int OS = 7;   // row stride of orc_dirk::A

void lu_factor(double (&a)[NSP][NSP], int (&piv)[NSP]) {
  for (int k = 0; k < NSP; k++) {
    int p = k;
    double best = std::fabs(a[k][k]);
    for (int r = k + 1; r < NSP; r++)
      if (std::fabs(a[r][k]) > best) { best = std::fabs(a[r][k]); p = r; }
    piv[k] = p;
    if (p != k)
      for (int c = 0; c < NSP; c++) std::swap(a[k][c], a[p][c]);
    const double ik = 1.0 / a[k][k];
    for (int r = k + 1; r < NSP; r++) {
      const double l = a[r][k] * ik;
      a[r][k] = l;
      for (int c = k + 1; c < NSP; c++) a[r][c] -= l * a[k][c];
    }
  }
}

void lu_solve(const double (&a)[NSP][NSP], const int (&piv)[NSP], double (&b)[NSP]) {
  for (int k = 0; k < NSP; k++)
    if (piv[k] != k) std::swap(b[k], b[piv[k]]);
  for (int r = 1; r < NSP; r++)
    for (int c = 0; c < r; c++) b[r] -= a[r][c] * b[c];
  for (int r = NSP - 1; r >= 0; r--) {
    for (int c = r + 1; c < NSP; c++) b[r] -= a[r][c] * b[c];
    b[r] /= a[r][r];
  }
}

struct Cell {
  const ChemState *c;
  long i;
  double Ts, dTs;
  double sc[NSP], isc[NSP];
  CellAux ax;
};

// calculate_rhs_cvklu for one cell (scaled), updating the cell's temperature
void cell_rhs(Cell &cl, const double (&y)[NSP], double srhs, double (&f)[NSP]) {
  double ya[NSP], rs[NRATE], drs[NRATE];
  for (int s = 0; s < NSP; s++) ya[s] = y[s] * cl.sc[s];
  temperature(cl.c->g, ya, cl.Ts, cl.dTs);
  interp_rates(cl.c->g, cl.Ts, rs, drs);
  double fp[NSP];
  rhs_physical(ya, rs, cl.Ts, cl.c->z, cl.ax, fp);
  for (int s = 0; s < NSP; s++) {
    f[s] = fp[s];
    f[s] *= cl.isc[s];
  }
  f[9] *= cl.ax.inv_mdensity;
  for (int s = 0; s < NSP; s++) f[s] *= srhs;
}

// calculate_jacobian_cvklu for one cell from the current temperature; M = I - hg*srhs*J
void cell_matrix(const Cell &cl, const double (&y)[NSP], double hg, double srhs,
                 double (&M)[NSP][NSP]) {
  double ya[NSP], rs[NRATE], drs[NRATE];
  for (int s = 0; s < NSP; s++) ya[s] = y[s] * cl.sc[s];
  interp_rates(cl.c->g, cl.Ts, rs, drs);
  double J[NSP][NSP];
  bool nzm[NSP][NSP];
  jac_physical(ya, rs, drs, cl.Ts, cl.c->z, cl.dTs, cl.ax, J, nzm);
  for (int r = 0; r < NSP; r++)
    for (int c = 0; c < NSP; c++) {
      double v = 0.0;
      if (nzm[r][c]) {
        v = J[r][c];
        v *= (cl.isc[r] * cl.sc[c]);
      }
      M[r][c] = (r == c ? 1.0 : 0.0) - hg * (srhs * v);
    }
}

double wrms(const double (&d)[NSP], const double (&z)[NSP], double rtol, double atol) {
  double s = 0.0;
  for (int k = 0; k < NSP; k++) {
    const double w = d[k] / (rtol * std::fabs(z[k]) + atol);
    s += w * w;
  }
  return std::sqrt(s / NSP);
}

// ARKODE's convergence-rate update and test (crdown 0.3; SUNMAX / SUNMIN as ternaries)
double conv_rate(double crate, double del, double delp) {
  const double a = 0.3 * crate, b = del / delp;
  return a > b ? a : b;
}

bool conv_ok(double crate, double del, double tol) {
  return (crate < 1.0 ? crate : 1.0) * del / tol <= 1.0;
}

bool conv_diverging(double del, double delp) { return del > 2.3 * delp; }

}  // namespace

extern "C" int orc_lu(const double *Min, double *LUout, int *piv, const double *b, double *x) {
  double a[NSP][NSP];
  int p[NSP];
  for (int r = 0; r < NSP; r++)
    for (int c = 0; c < NSP; c++) a[r][c] = Min[r * NSP + c];
  lu_factor(a, p);
  double bb[NSP];
  for (int r = 0; r < NSP; r++) bb[r] = b[r];
  lu_solve(a, p, bb);
  for (int r = 0; r < NSP; r++) {
    x[r] = bb[r];
    piv[r] = p[r];
    for (int c = 0; c < NSP; c++) LUout[r * NSP + c] = a[r][c];
  }
  return 0;
}

// A cell's Newton matrix factors and convergence-rate estimate kept from one step to the
// next within a multi-step call (ARKODE's lsetup frequency, max_steps_between_setups).
struct Carry {
  double M[NSP][NSP];
  int piv[NSP];
  double crate;
};

// One DIRK step of one cell; y (10, scaled) is overwritten with y_{n+1}, the cell's Ts/dTs
// follow its RHS evaluations. Y (may be null) receives the stage values. Returns false when
// the Newton budget ran out; nit counts the Newton iterations. With reuse, the step starts
// Newton with the factors and rate estimate in cy instead of f(y_n) and a setup at y_n; cy
// (may be null) receives the step's final factors and estimate, set (may be null) whether
// the step set up.
static bool step_cell(Cell &cl, const orc_dirk *opt, double hstep, double (&y)[NSP],
                      double (*Y)[NSP], int &nit, const double *force = nullptr,
                      Carry *cy = nullptr, bool reuse = false, bool *set = nullptr) {
  const int ns = opt->nstage;
  // ESDIRK (first row zero): the first stage is explicit, z_1 = y_n, F_1 = f(y_n)
  const bool esdirk = ns >= 2 && opt->A[0] == 0.0;
  double yn[NSP], z[NSP], F[OS][NSP], f[NSP];
  for (int s = 0; s < NSP; s++) yn[s] = z[s] = y[s];
  double M[NSP][NSP];
  int piv[NSP];
  bool ok_cell = true, did_setup = !reuse;
  double crate = 1.0, delp = 0.0;     // convergence-rate estimate, last correction norm
  bool have_f = false;                // f holds f(y_n) for the first Newton iteration
  if (reuse) {
    std::memcpy(M, cy->M, sizeof(M));
    std::memcpy(piv, cy->piv, sizeof(piv));
    crate = cy->crate;
  }
  if (!reuse || esdirk) {
    cell_rhs(cl, yn, opt->rhs_scale, f);             // Ts at y_n for J; F_1 of an ESDIRK
    if (esdirk) {
      for (int s = 0; s < NSP; s++) F[0][s] = force ? f[s] + force[s] : f[s];
      if (Y)
        for (int s = 0; s < NSP; s++) Y[0][s] = yn[s];
      // with held factors this RHS is also the first implicit stage's first iteration's
      // (the batched solver's pass order); after a setup that iteration evaluates again
      have_f = reuse;
    }
  }
  nit = 0;
  const int st0 = esdirk ? 1 : 0;
  for (int st = st0; st < ns && ok_cell; st++) {
    const double hg = hstep * opt->A[st * OS + st];
    if (st == st0 && !reuse) {
      cell_matrix(cl, yn, hg, opt->rhs_scale, M);
      lu_factor(M, piv);
    }
    double r[NSP];
    for (int s = 0; s < NSP; s++) {
      double acc = 0.0;
      for (int j = 0; j < st; j++) acc += opt->A[st * OS + j] * F[j][s];
      r[s] = yn[s] + hstep * acc;
    }
    int setups = 0;
    bool conv = false;
    while (!conv) {
      for (int it = 0; it < opt->max_iters; it++) {
        if (have_f) have_f = false;
        else cell_rhs(cl, z, opt->rhs_scale, f);
        double d[NSP];
        if (force)
          for (int s = 0; s < NSP; s++) d[s] = -(z[s] - hg * (f[s] + force[s]) - r[s]);
        else
          for (int s = 0; s < NSP; s++) d[s] = -(z[s] - hg * f[s] - r[s]);
        lu_solve(M, piv, d);
        for (int s = 0; s < NSP; s++) z[s] += d[s];
        nit++;
        const double del = wrms(d, z, opt->rtol, opt->atol);
        if (it > 0) crate = conv_rate(crate, del, delp);
        if (conv_ok(crate, del, opt->newton_tol)) { conv = true; break; }
        if (it > 0 && conv_diverging(del, delp)) break;
        delp = del;
      }
      if (conv) break;
      if (++setups > opt->max_setups) { ok_cell = false; break; }
      cell_rhs(cl, z, opt->rhs_scale, f);             // temperature at the iterate
      cell_matrix(cl, z, hg, opt->rhs_scale, M);
      lu_factor(M, piv);
      crate = 1.0;
      did_setup = true;
    }
    for (int s = 0; s < NSP; s++) F[st][s] = (z[s] - r[s]) / hg;
    if (Y)
      for (int s = 0; s < NSP; s++) Y[st][s] = z[s];
  }
  for (int s = 0; s < NSP; s++) y[s] = z[s];
  if (cy) {
    std::memcpy(cy->M, M, sizeof(M));
    std::memcpy(cy->piv, piv, sizeof(piv));
    cy->crate = crate;
  }
  if (set) *set = did_setup;
  return ok_cell;
}

// The step with per-cell step rejection when opt->max_halvings > 0 (the batched solver's
// option): a failed attempt returns the cell to its substep start (state and temperature
// warm start) and retries in halves, up to max_halvings times; accepted substeps count in
// units of h / 2^max_halvings. A cell that fails at the finest size keeps its last accepted
// state. Without it, the plain step (a failed cell keeps its last Newton iterate).
// With a carry (multi-step calls): the first attempt may reuse the carried factors (reuse),
// every later attempt and substep sets up (its h differs or its predecessor failed); set
// reports whether any attempt set up, halved whether the step was split.
static bool step_cell_sub(Cell &cl, const orc_dirk *opt, double hstep, double (&y)[NSP],
                          int &nit, const double *force, Carry *cy = nullptr,
                          bool reuse = false, bool *set = nullptr, bool *halved = nullptr) {
  if (halved) *halved = false;
  if (opt->max_halvings <= 0) return step_cell(cl, opt, hstep, y, nullptr, nit, force, cy, reuse, set);
  const int mh = opt->max_halvings;
  int halv = 0, done = 0;
  bool any_set = false, first = true;
  nit = 0;
  while (done < (1 << mh)) {
    double z[NSP];
    for (int s = 0; s < NSP; s++) z[s] = y[s];
    const double Ts0 = cl.Ts, dTs0 = cl.dTs;
    int k = 0;
    bool ds = false;
    const bool ok = step_cell(cl, opt, std::ldexp(hstep, -halv), z, nullptr, k, force, cy,
                              first && reuse, &ds);
    first = false;
    any_set = any_set || ds;
    nit += k;
    if (ok) {
      for (int s = 0; s < NSP; s++) y[s] = z[s];
      done += 1 << (mh - halv);
    } else {
      cl.Ts = Ts0;
      cl.dTs = dTs0;
      if (halv == mh) {
        if (set) *set = any_set;
        if (halved) *halved = halv > 0;
        return false;
      }
      halv++;
    }
  }
  if (set) *set = any_set;
  if (halved) *halved = halv > 0;
  return true;
}

static Cell cell_of(ChemState *c, long i) {
  Cell cl{c, i, c->Ts[i], c->dTs_ge[i], {}, {}, {c->mdensity[i], c->inv_mdensity[i],
                                                c->cie_oda[i], c->h2_oda[i]}};
  for (int s = 0; s < NSP; s++) {
    cl.sc[s] = c->scale[NSP * i + s];
    cl.isc[s] = c->inv_scale[NSP * i + s];
  }
  return cl;
}

// One DIRK step of every cell with a constant forcing added to the scaled RHS (force
// [cell][10] or null; the forcing an MRI method hands its fast stepper); y (n x 10, scaled)
// is overwritten with y_{n+1}; Ts is left at the temperature of the last RHS evaluation of
// each cell. iters[i] receives the total Newton iterations of cell i (may be null). Returns
// the number of failed cells.
extern "C" int orc_chem_dirk_step_forced(void *h, const orc_dirk *opt, double hstep, double *y,
                                         const double *force, int *iters) {
  ChemState *c = (ChemState *)h;
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : failed)
  for (long i = 0; i < c->n; i++) {
    Cell cl = cell_of(c, i);
    double z[NSP];
    for (int s = 0; s < NSP; s++) z[s] = y[NSP * i + s];
    int nit = 0;
    const bool ok_cell = step_cell_sub(cl, opt, hstep, z, nit, force ? force + NSP * i : nullptr);
    if (!ok_cell) failed++;
    for (int s = 0; s < NSP; s++) y[NSP * i + s] = z[s];
    c->Ts[i] = cl.Ts;
    c->dTs_ge[i] = cl.dTs;
    if (iters) iters[i] = ok_cell ? nit : -1;
  }
  return failed;
}

// The unforced step.
extern "C" int orc_chem_dirk_step(void *h, const orc_dirk *opt, double hstep, double *y,
                                  int *iters) {
  return orc_chem_dirk_step_forced(h, opt, hstep, y, nullptr, iters);
}

// nsteps steps of every cell in one call, with ARKODE's lsetup frequency: a step starts with
// f(y_n) and a Jacobian setup at y_n when it is the call's first, when its predecessor
// failed or was split, or when max_steps_between_setups steps (at least 1) have passed since
// the cell's last setup (any setup, those after Newton failures included; ARKODE's nst >=
// nstlp + msbp); the other steps start Newton with the factors and convergence-rate estimate
// the cell holds. Returns the failed cell steps; iters[i] (may be null) the cell's Newton
// iterations over all its steps.
extern "C" int orc_chem_dirk_steps_forced(void *h, const orc_dirk *opt, double hstep,
                                          int nsteps, double *y, const double *force,
                                          int *iters) {
  ChemState *c = (ChemState *)h;
  const int msbp = opt->max_steps_between_setups > 1 ? opt->max_steps_between_setups : 1;
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : failed)
  for (long i = 0; i < c->n; i++) {
    Cell cl = cell_of(c, i);
    double z[NSP];
    for (int s = 0; s < NSP; s++) z[s] = y[NSP * i + s];
    Carry cy{};
    int lsetup = 0, nit_all = 0;
    bool must_setup = true;
    for (int k = 0; k < nsteps; k++) {
      const bool reuse = !must_setup && k - lsetup < msbp;
      bool ds = false, halved = false;
      int nit = 0;
      const bool ok = step_cell_sub(cl, opt, hstep, z, nit, force ? force + NSP * i : nullptr,
                                    &cy, reuse, &ds, &halved);
      if (ds) lsetup = k;
      must_setup = !ok || halved;
      nit_all += nit;
      if (!ok) failed++;
    }
    for (int s = 0; s < NSP; s++) y[NSP * i + s] = z[s];
    c->Ts[i] = cl.Ts;
    c->dTs_ge[i] = cl.dTs;
    if (iters) iters[i] = nit_all;
  }
  return failed;
}

extern "C" int orc_chem_exact_jac_cell(void *h, long i, const double *y, double *T,
                                       double *Jout);

// Discrete adjoint of nsteps DIRK steps, per cell: lam0 = (d y_N / d y_0)^T lamT for the
// stage equations solved exactly (the derivative of the idealised method along the computed
// trajectory). Every step's starting state is kept (no checkpointing here); backwards, the
// step is recomputed from its state to get the stage values Y_k, and with the exact scaled
// stage Jacobians J_k = s_rhs dF/dy(Y_k):
//   for k = s..1: Ybar_k = h J_k^T (b_k lam + sum_{m>k} a_mk nu_m),
//                 nu_k = (I - h a_kk J_k)^{-T} Ybar_k;      lam <- lam + sum_k nu_k,
// with b the last row of A (stiffly accurate tables). y is advanced to y_N, Ts/dTs left at
// the forward run's end. Returns the number of cells whose forward Newton failed.
extern "C" int orc_chem_dirk_adjoint(void *h, const orc_dirk *opt, double hstep, int nsteps,
                                     double *y, const double *lamT, double *lam0) {
  ChemState *c = (ChemState *)h;
  const int ns = opt->nstage;
  if (ns >= 2 && opt->A[0] == 0.0) return -1;        // SDIRK tables only
  int failed = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : failed)
  for (long i = 0; i < c->n; i++) {
    Cell cl = cell_of(c, i);
    std::vector<double> st((size_t)(nsteps + 1) * (NSP + 2));
    double z[NSP];
    for (int s = 0; s < NSP; s++) z[s] = y[NSP * i + s];
    bool ok = true;
    for (int n = 0; n < nsteps; n++) {
      double *rec = &st[(size_t)n * (NSP + 2)];
      for (int s = 0; s < NSP; s++) rec[s] = z[s];
      rec[NSP] = cl.Ts;
      rec[NSP + 1] = cl.dTs;
      int nit = 0;
      ok = step_cell(cl, opt, hstep, z, nullptr, nit) && ok;
    }
    if (!ok) failed++;
    for (int s = 0; s < NSP; s++) y[NSP * i + s] = z[s];
    c->Ts[i] = cl.Ts;
    c->dTs_ge[i] = cl.dTs;
    double lam[NSP];
    for (int s = 0; s < NSP; s++) lam[s] = lamT[NSP * i + s];
    for (int n = nsteps - 1; n >= 0; n--) {
      const double *rec = &st[(size_t)n * (NSP + 2)];
      Cell cr = cl;
      cr.Ts = rec[NSP];
      cr.dTs = rec[NSP + 1];
      double yy[NSP], Y[OS][NSP];
      for (int s = 0; s < NSP; s++) yy[s] = rec[s];
      int nit = 0;
      step_cell(cr, opt, hstep, yy, Y, nit);
      double nu[OS][NSP] = {};
      for (int k = ns - 1; k >= 0; k--) {
        double w[NSP];
        for (int s = 0; s < NSP; s++) {
          double acc = opt->A[(ns - 1) * OS + k] * lam[s];
          for (int m = k + 1; m < ns; m++) acc += opt->A[m * OS + k] * nu[m][s];
          w[s] = acc;
        }
        double J[NSP * NSP], T = cr.Ts;
        orc_chem_exact_jac_cell(c, i, Y[k], &T, J);
        double yb[NSP], Mt[NSP][NSP];
        for (int q = 0; q < NSP; q++) {
          double acc = 0.0;
          for (int s = 0; s < NSP; s++) acc += (opt->rhs_scale * J[s * NSP + q]) * w[s];
          yb[q] = hstep * acc;
        }
        const double hg = hstep * opt->A[k * OS + k];
        for (int r = 0; r < NSP; r++)
          for (int q = 0; q < NSP; q++)
            Mt[r][q] = (r == q ? 1.0 : 0.0) - hg * (opt->rhs_scale * J[q * NSP + r]);
        int piv[NSP];
        lu_factor(Mt, piv);
        lu_solve(Mt, piv, yb);
        for (int s = 0; s < NSP; s++) nu[k][s] = yb[s];
      }
      for (int k = 0; k < ns; k++)
        for (int s = 0; s < NSP; s++) lam[s] += nu[k][s];
    }
    for (int s = 0; s < NSP; s++) lam0[NSP * i + s] = lam[s];
  }
  return failed;
}
