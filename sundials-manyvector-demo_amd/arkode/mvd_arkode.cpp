// ARKODE callbacks over libmvd_rhs.so (include/mvd_arkode.h): each one unpacks the
// SUNDIALS vectors and matrices through SUNDIALS' public API and makes the libmvd_rhs.so
// calls that give the reference callback's result. Host code; all data stays on the device.
#include <cmath>
#include <cstdlib>
#include <new>
#include "../../include/mvd_arkode.h"

namespace {

constexpr int NSUB_MAX = 6;

mvd_arkode_data *data_of(void *user_data) { return static_cast<mvd_arkode_data *>(user_data); }

// Device pointers of the subvectors of a ManyVector (mpi: an MPIManyVector); false when the
// vector has fewer than `need` subvectors or one of them has no device data.
bool unpack(N_Vector v, bool mpi, int need, double *(&p)[NSUB_MAX], int *nsub = nullptr) {
  if (!v) return false;
  const sunindextype n =
      mpi ? N_VGetNumSubvectors_MPIManyVector(v) : N_VGetNumSubvectors_ManyVector(v);
  if (n < need || n > NSUB_MAX) return false;
  for (int q = 0; q < NSUB_MAX; q++) p[q] = nullptr;
  for (int q = 0; q < (int)n; q++) {
    N_Vector s = mpi ? N_VGetSubvector_MPIManyVector(v, q) : N_VGetSubvector_ManyVector(v, q);
    if (!s) return false;
    p[q] = N_VGetDeviceArrayPointer(s);
    if (!p[q]) return false;
  }
  if (nsub) *nsub = (int)n;
  return true;
}

sunindextype sublen(N_Vector v, bool mpi, int q) {
  N_Vector s = mpi ? N_VGetSubvector_MPIManyVector(v, q) : N_VGetSubvector_ManyVector(v, q);
  return s ? N_VGetLength(s) : -1;
}

// SUNRpowerI (sundials_math.c): base^exponent by repeated multiplication
double power_i(double base, int exponent) {
  double prod = 1.0;
  const int e = exponent < 0 ? -exponent : exponent;
  for (int i = 0; i < e; i++) prod *= base;
  return exponent < 0 ? 1.0 / prod : prod;
}

// wdot = 0 on the five fluid subvectors (N_VConst(ZERO, wdot) of the chemistry callbacks;
// the chemistry subvector is written whole by the RHS call)
int zero_fluid(mvd_context *ctx, double *const *wd, sunindextype n) {
  for (int q = 0; q < 5; q++) {
    const int rc = mvd_nvec_const(ctx, 0.0, wd[q], n);
    if (rc) return rc;
  }
  return 0;
}

// ffast / fimpl body: [0; s * rhs(chem)] (+ the MRI forcing when forced)
int chem_rhs_many(mvd_arkode_data *d, realtype t, N_Vector w, N_Vector wdot, bool mpi,
                  bool forced) {
  if (!d || !d->ctx) return -1;
  double *wp[NSUB_MAX], *dp[NSUB_MAX];
  if (!unpack(w, mpi, 6, wp) || !unpack(wdot, mpi, 6, dp)) return -1;
  const sunindextype n = sublen(w, mpi, 0);
  int rc = zero_fluid(d->ctx, dp, n);
  if (rc) return rc;
  rc = mvd_chem_rhs_scaled(d->ctx, t, d->time_units, wp[5], dp[5]);
  if (rc) return rc;
  if (!forced || d->disable_forcing || !d->stepper) return 0;
  realtype tshift = 0.0, tscale = 1.0;
  N_Vector *forcing = nullptr;
  int nforcing = 0;
  if (MRIStepInnerStepper_GetForcingData(d->stepper, &tshift, &tscale, &forcing, &nforcing))
    return -1;
  if (nforcing <= 0) return 0;
  if (nforcing > 15 || !forcing) return -1;
  // per subvector: wdot_q = 1 * wdot_q + sum_i tau^i F_i,q (N_VLinearCombination with the
  // output as its first vector)
  const double tau = (t - tshift) / tscale;
  double c[16];
  const double *X[16];
  c[0] = 1.0;
  for (int i = 0; i < nforcing; i++) c[i + 1] = power_i(tau, i);
  for (int q = 0; q < 6; q++) {
    X[0] = dp[q];
    for (int i = 0; i < nforcing; i++) {
      N_Vector f = N_VGetSubvector_MPIManyVector(forcing[i], q);
      X[i + 1] = f ? N_VGetDeviceArrayPointer(f) : nullptr;
      if (!X[i + 1]) return -1;
    }
    rc = mvd_nvec_linear_combination(d->ctx, nforcing + 1, c, X, dp[q], sublen(wdot, mpi, q));
    if (rc) return rc;
  }
  return 0;
}

int chem_jac_many(mvd_arkode_data *d, realtype t, N_Vector w, SUNMatrix Jac, bool mpi,
                  double scale) {
  if (!d || !d->ctx || !Jac) return -1;
  double *wp[NSUB_MAX];
  if (!unpack(w, mpi, 6, wp)) return -1;
  double *J = SUNMatrix_MagmaDense_Data(Jac);
  if (!J || SUNMatrix_MagmaDense_BlockRows(Jac) != 10 ||
      SUNMatrix_MagmaDense_BlockColumns(Jac) != 10 ||
      SUNMatrix_MagmaDense_NumBlocks(Jac) != sublen(w, mpi, 0))
    return -1;
  return mvd_chem_jac_scaled(d->ctx, t, scale, wp[5], J);
}

struct RankLocalLS {
  sunindextype subvec;
  sunindextype lastflag;
  mvd_arkode_data *data;
};

RankLocalLS *ls_content(SUNLinearSolver S) {
  return S ? static_cast<RankLocalLS *>(S->content) : nullptr;
}

}  // namespace

extern "C" {

int mvd_arkode_fEuler(realtype t, N_Vector w, N_Vector wdot, void *user_data) {
  mvd_arkode_data *d = data_of(user_data);
  if (!d || !d->ctx) return -1;
  double *wp[NSUB_MAX], *dp[NSUB_MAX];
  int nw = 0, nd = 0;
  if (!unpack(w, true, 5, wp, &nw) || !unpack(wdot, true, 5, dp, &nd) || nw != nd) return -1;
  const int rc = mvd_fe(d->ctx, t, wp, dp);
  return rc == MVD_ILLEGAL_STATE ? -1 : rc;
}

int mvd_arkode_stability(N_Vector w, realtype t, realtype *dt_stab, void *user_data) {
  mvd_arkode_data *d = data_of(user_data);
  if (!d || !d->ctx || !dt_stab) return -1;
  double *wp[NSUB_MAX];
  if (!unpack(w, true, 5, wp)) return -1;
  return mvd_stability(d->ctx, wp, t, dt_stab);
}

int mvd_arkode_fslow(realtype t, N_Vector w, N_Vector wdot, void *user_data) {
  mvd_arkode_data *d = data_of(user_data);
  if (!d || !d->ctx) return -1;
  double *wp[NSUB_MAX], *dp[NSUB_MAX];
  if (!unpack(w, true, 6, wp) || !unpack(wdot, true, 6, dp)) return -1;
  const int rc = mvd_fslow(d->ctx, t, wp, dp);
  return rc == MVD_ILLEGAL_STATE ? -1 : rc;
}

int mvd_arkode_ffast(realtype t, N_Vector w, N_Vector wdot, void *user_data) {
  return chem_rhs_many(data_of(user_data), t, w, wdot, false, true);
}

int mvd_arkode_fimpl(realtype t, N_Vector w, N_Vector wdot, void *user_data) {
  return chem_rhs_many(data_of(user_data), t, w, wdot, true, false);
}

int mvd_arkode_frhs(realtype t, N_Vector w, N_Vector wdot, void *user_data) {
  mvd_arkode_data *d = data_of(user_data);
  if (!d || !d->ctx || !w || !wdot) return -1;
  double *y = N_VGetDeviceArrayPointer(w), *f = N_VGetDeviceArrayPointer(wdot);
  if (!y || !f) return -1;
  return mvd_chem_rhs(d->ctx, t, y, f);
}

int mvd_arkode_Jfast(realtype t, N_Vector w, N_Vector, SUNMatrix Jac, void *user_data,
                     N_Vector, N_Vector, N_Vector) {
  mvd_arkode_data *d = data_of(user_data);
  return chem_jac_many(d, t, w, Jac, false, d ? d->time_units : 1.0);
}

int mvd_arkode_Jimpl(realtype t, N_Vector w, N_Vector, SUNMatrix Jac, void *user_data,
                     N_Vector, N_Vector, N_Vector) {
  mvd_arkode_data *d = data_of(user_data);
  return chem_jac_many(d, t, w, Jac, true, d ? d->time_units : 1.0);
}

int mvd_arkode_Jrhs(realtype t, N_Vector w, N_Vector, SUNMatrix Jac, void *user_data,
                    N_Vector, N_Vector, N_Vector) {
  mvd_arkode_data *d = data_of(user_data);
  if (!d || !d->ctx || !w || !Jac) return -1;
  double *y = N_VGetDeviceArrayPointer(w), *J = SUNMatrix_MagmaDense_Data(Jac);
  if (!y || !J) return -1;
  return mvd_chem_jac(d->ctx, t, y, J);
}

int mvd_arkode_postprocess(realtype t, N_Vector w, void *user_data) {
  mvd_arkode_data *d = data_of(user_data);
  if (!d || !d->ctx) return -1;
  double *wp[NSUB_MAX];
  if (!unpack(w, true, 6, wp)) return -1;
  return mvd_postprocess_step(d->ctx, t, wp);
}

// ---- SUNLinSol_RankLocalLS ----

SUNLinearSolver mvd_SUNLinSol_RankLocalLS(N_Vector x, sunindextype subvec,
                                          mvd_arkode_data *data, SUNContext sunctx) {
  if (!x || !data || subvec < 0 || subvec >= N_VGetNumSubvectors_ManyVector(x)) return nullptr;
  SUNLinearSolver S = SUNLinSolNewEmpty(sunctx);
  if (!S) return nullptr;
  S->ops->gettype = mvd_RankLocalLS_GetType;
  S->ops->initialize = mvd_RankLocalLS_Initialize;
  S->ops->setup = mvd_RankLocalLS_Setup;
  S->ops->solve = mvd_RankLocalLS_Solve;
  S->ops->lastflag = mvd_RankLocalLS_LastFlag;
  S->ops->free = mvd_RankLocalLS_Free;
  RankLocalLS *c = new (std::nothrow) RankLocalLS{subvec, 0, data};
  if (!c) {
    SUNLinSolFreeEmpty(S);
    return nullptr;
  }
  S->content = c;
  return S;
}

SUNLinearSolver_Type mvd_RankLocalLS_GetType(SUNLinearSolver) { return SUNLINEARSOLVER_DIRECT; }

int mvd_RankLocalLS_Initialize(SUNLinearSolver S) {
  RankLocalLS *c = ls_content(S);
  if (!c) return SUNLS_MEM_NULL;
  c->lastflag = SUNLS_SUCCESS;
  return SUNLS_SUCCESS;
}

int mvd_RankLocalLS_Setup(SUNLinearSolver S, SUNMatrix A) {
  RankLocalLS *c = ls_content(S);
  if (!c) return SUNLS_MEM_NULL;
  const double *a = A ? SUNMatrix_MagmaDense_Data(A) : nullptr;
  if (!a) return (int)(c->lastflag = SUNLS_MEM_FAIL);
  const int rc = mvd_chem_lsetup_matrix(c->data->ctx, a);
  c->lastflag = rc == 0 ? SUNLS_SUCCESS : rc > 0 ? SUNLS_LUFACT_FAIL : SUNLS_PACKAGE_FAIL_UNREC;
  return (int)c->lastflag;
}

int mvd_RankLocalLS_Solve(SUNLinearSolver S, SUNMatrix, N_Vector x, N_Vector b, realtype) {
  RankLocalLS *c = ls_content(S);
  if (!c) return SUNLS_MEM_NULL;
  N_Vector xs = x ? N_VGetSubvector_ManyVector(x, c->subvec) : nullptr;
  N_Vector bs = b ? N_VGetSubvector_ManyVector(b, c->subvec) : nullptr;
  if (!xs || !bs) return (int)(c->lastflag = SUNLS_MEM_FAIL);
  double *xd = N_VGetDeviceArrayPointer(xs);
  const double *bd = N_VGetDeviceArrayPointer(bs);
  if (!xd || !bd) return (int)(c->lastflag = SUNLS_MEM_FAIL);
  const int rc = mvd_chem_lsolve(c->data->ctx, bd, xd);
  c->lastflag = rc == 0 ? SUNLS_SUCCESS : SUNLS_PACKAGE_FAIL_UNREC;
  return (int)c->lastflag;
}

sunindextype mvd_RankLocalLS_LastFlag(SUNLinearSolver S) {
  RankLocalLS *c = ls_content(S);
  return c ? c->lastflag : SUNLS_MEM_NULL;
}

int mvd_RankLocalLS_Free(SUNLinearSolver S) {
  if (!S) return 0;
  delete ls_content(S);
  S->content = nullptr;
  SUNLinSolFreeEmpty(S);
  return 0;
}

// ---- the rank-local inner stepper ----

int mvd_inner_evolve(MRIStepInnerStepper stepper, realtype t0, realtype tout, N_Vector y) {
  void *content = nullptr;
  if (MRIStepInnerStepper_GetContent(stepper, &content) || !content) return -1;
  mvd_arkode_data *d = data_of(content);
  if (!d->ctx || !(d->hmax_fast > 0)) return -1;
  double *yp[NSUB_MAX];
  if (!unpack(y, true, 6, yp)) return -1;
  const sunindextype n = sublen(y, true, 0);
  const double span = tout - t0;
  if (!(span >= 0.0)) return -1;
  if (span == 0.0) { d->tcur = tout; return 0; }
  realtype tshift = 0.0, tscale = 1.0;
  N_Vector *forcing = nullptr;
  int nforcing = 0;
  if (MRIStepInnerStepper_GetForcingData(stepper, &tshift, &tscale, &forcing, &nforcing))
    return -1;
  if (nforcing != 1 || !forcing) return -1;   // forcing constant over the step only
  double *fp[NSUB_MAX];
  if (!unpack(forcing[0], true, 6, fp)) return -1;
  const long nfast = (long)std::ceil(span / d->hmax_fast);
  const double dt = span / (double)nfast;
  mvd_dirk_opts o = d->fast;
  o.rhs_scale = d->time_units;
  if (nfast > 0x3fffffff) return -1;
  // fluid: y' = F (no fast fluid dynamics), nfast steps of y += dt F per subvector; the
  // chemistry: the nfast forced steps in one launch, each cell running them back to back
  for (int q = 0; q < 5; q++) {
    const int rc = mvd_nvec_axpy_steps(d->ctx, (int)nfast, dt, fp[q], yp[q], n);
    if (rc) return rc;
  }
  mvd_dirk_stats st;
  const int rc = mvd_chem_dirk_steps(d->ctx, t0, dt, (int)nfast, &o, yp[5], fp[5], &st);
  if (rc < 0) return rc;
  const long failed = rc;
  d->fast_steps += nfast;
  d->failed_cells += failed;
  d->tcur = tout;
  return failed > 0x3fffffff ? 0x3fffffff : (int)failed;
}

int mvd_inner_fullrhs(MRIStepInnerStepper stepper, realtype t, N_Vector y, N_Vector f, int) {
  void *content = nullptr;
  if (MRIStepInnerStepper_GetContent(stepper, &content) || !content) return -1;
  mvd_arkode_data *d = data_of(content);
  // ffast with the forcing disabled (RankLocalStepper_FullRhs, :1162-1195); the vectors are
  // MRIStep's MPIManyVectors
  d->disable_forcing = 1;
  const int rc = chem_rhs_many(d, t, y, f, true, true);
  d->disable_forcing = 0;
  return rc;
}

int mvd_inner_reset(MRIStepInnerStepper stepper, realtype tR, N_Vector yR) {
  void *content = nullptr;
  if (MRIStepInnerStepper_GetContent(stepper, &content) || !content) return -1;
  double *yp[NSUB_MAX];
  if (!unpack(yR, true, 6, yp)) return -1;
  data_of(content)->tcur = tR;
  return 0;
}

int mvd_inner_stepper_create(SUNContext sunctx, mvd_arkode_data *data,
                             MRIStepInnerStepper *stepper) {
  if (!data || !stepper) return -1;
  if (MRIStepInnerStepper_Create(sunctx, stepper) || !*stepper) return -1;
  if (MRIStepInnerStepper_SetContent(*stepper, data) ||
      MRIStepInnerStepper_SetEvolveFn(*stepper, mvd_inner_evolve) ||
      MRIStepInnerStepper_SetFullRhsFn(*stepper, mvd_inner_fullrhs) ||
      MRIStepInnerStepper_SetResetFn(*stepper, mvd_inner_reset)) {
    MRIStepInnerStepper_Free(stepper);
    return -1;
  }
  data->stepper = *stepper;
  return 0;
}

}  // extern "C"
