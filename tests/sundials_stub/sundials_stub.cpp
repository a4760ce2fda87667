// TEST INFRASTRUCTURE ONLY: the handful of SUNDIALS 6.2.0 functions that the ARKODE adapters
// (libmvd_arkode.so) call, implemented over plain structs so that tests can drive the
// adapters the way ARKStep / MRIStep would, without SUNDIALS. Vectors wrap device arrays
// (or lists of subvectors), matrices wrap column-major 10x10 device blocks; the inner
// stepper records its functions and the forcing a test sets. Nothing here ships in the
// product libraries; production builds link SUNDIALS itself.
#include <cstdlib>
#include <new>
#include "../../sundials-manyvector-demo_amd/arkode/mvd_sundials_min.h"

struct _generic_N_Vector {
  double *data;      // device array (leaf vectors)
  long len;
  int nsub;          // > 0: a ManyVector of nsub subvectors
  N_Vector sub[8];
};
struct _generic_SUNMatrix {
  double *data;
  long nblocks, rows, cols;
};
struct _SUNContext {
  int unused;
};
struct _MRIStepInnerStepper {
  void *content;
  MRIStepInnerEvolveFn evolve;
  MRIStepInnerFullRhsFn fullrhs;
  MRIStepInnerResetFn reset;
  double tshift, tscale;
  N_Vector *forcing;
  int nforcing;
};

extern "C" {

// ---- the SUNDIALS functions the adapters call ----
realtype *N_VGetDeviceArrayPointer(N_Vector v) { return v && v->nsub == 0 ? v->data : nullptr; }
sunindextype N_VGetLength(N_Vector v) {
  if (!v) return -1;
  if (v->nsub == 0) return v->len;
  long s = 0;
  for (int q = 0; q < v->nsub; q++) s += N_VGetLength(v->sub[q]);
  return s;
}
N_Vector N_VGetSubvector_ManyVector(N_Vector v, sunindextype i) {
  return (v && i >= 0 && i < v->nsub) ? v->sub[i] : nullptr;
}
sunindextype N_VGetNumSubvectors_ManyVector(N_Vector v) { return v ? v->nsub : 0; }
N_Vector N_VGetSubvector_MPIManyVector(N_Vector v, sunindextype i) {
  return N_VGetSubvector_ManyVector(v, i);
}
sunindextype N_VGetNumSubvectors_MPIManyVector(N_Vector v) {
  return N_VGetNumSubvectors_ManyVector(v);
}

realtype *SUNMatrix_MagmaDense_Data(SUNMatrix A) { return A ? A->data : nullptr; }
sunindextype SUNMatrix_MagmaDense_LData(SUNMatrix A) {
  return A ? A->nblocks * A->rows * A->cols : 0;
}
sunindextype SUNMatrix_MagmaDense_NumBlocks(SUNMatrix A) { return A ? A->nblocks : 0; }
sunindextype SUNMatrix_MagmaDense_BlockRows(SUNMatrix A) { return A ? A->rows : 0; }
sunindextype SUNMatrix_MagmaDense_BlockColumns(SUNMatrix A) { return A ? A->cols : 0; }

SUNLinearSolver SUNLinSolNewEmpty(SUNContext sunctx) {
  SUNLinearSolver S = (SUNLinearSolver)calloc(1, sizeof(*S));
  if (!S) return nullptr;
  S->ops = (SUNLinearSolver_Ops)calloc(1, sizeof(*S->ops));
  if (!S->ops) { free(S); return nullptr; }
  S->sunctx = sunctx;
  return S;
}
void SUNLinSolFreeEmpty(SUNLinearSolver S) {
  if (!S) return;
  free(S->ops);
  free(S);
}

int MRIStepInnerStepper_Create(SUNContext, MRIStepInnerStepper *stepper) {
  *stepper = (MRIStepInnerStepper)calloc(1, sizeof(**stepper));
  if (!*stepper) return -1;
  (*stepper)->tscale = 1.0;
  return 0;
}
int MRIStepInnerStepper_Free(MRIStepInnerStepper *stepper) {
  if (stepper && *stepper) { free(*stepper); *stepper = nullptr; }
  return 0;
}
int MRIStepInnerStepper_SetContent(MRIStepInnerStepper s, void *c) { s->content = c; return 0; }
int MRIStepInnerStepper_GetContent(MRIStepInnerStepper s, void **c) {
  if (!s) return -1;
  *c = s->content;
  return 0;
}
int MRIStepInnerStepper_SetEvolveFn(MRIStepInnerStepper s, MRIStepInnerEvolveFn f) {
  s->evolve = f;
  return 0;
}
int MRIStepInnerStepper_SetFullRhsFn(MRIStepInnerStepper s, MRIStepInnerFullRhsFn f) {
  s->fullrhs = f;
  return 0;
}
int MRIStepInnerStepper_SetResetFn(MRIStepInnerStepper s, MRIStepInnerResetFn f) {
  s->reset = f;
  return 0;
}
int MRIStepInnerStepper_GetForcingData(MRIStepInnerStepper s, realtype *tshift,
                                       realtype *tscale, N_Vector **forcing, int *nforcing) {
  if (!s) return -1;
  *tshift = s->tshift;
  *tscale = s->tscale;
  *forcing = s->forcing;
  *nforcing = s->nforcing;
  return 0;
}

// ---- constructors and drivers for the tests ----
N_Vector stub_vector(double *data, long len) {
  N_Vector v = new (std::nothrow) _generic_N_Vector{};
  if (v) { v->data = data; v->len = len; }
  return v;
}
N_Vector stub_manyvector(int nsub, N_Vector *subs) {
  if (nsub < 1 || nsub > 8) return nullptr;
  N_Vector v = new (std::nothrow) _generic_N_Vector{};
  if (!v) return nullptr;
  v->nsub = nsub;
  for (int q = 0; q < nsub; q++) v->sub[q] = subs[q];
  return v;
}
void stub_vector_free(N_Vector v) { delete v; }
SUNMatrix stub_matrix(double *data, long nblocks) {
  SUNMatrix A = new (std::nothrow) _generic_SUNMatrix{};
  if (A) { A->data = data; A->nblocks = nblocks; A->rows = 10; A->cols = 10; }
  return A;
}
void stub_matrix_free(SUNMatrix A) { delete A; }
SUNContext stub_context() { return new (std::nothrow) _SUNContext{}; }
void stub_context_free(SUNContext c) { delete c; }
void stub_set_forcing(MRIStepInnerStepper s, double tshift, double tscale, int nforcing,
                      N_Vector *forcing) {
  s->tshift = tshift;
  s->tscale = tscale;
  s->nforcing = nforcing;
  s->forcing = forcing;
}
int stub_stepper_evolve(MRIStepInnerStepper s, double t0, double tout, N_Vector y) {
  return s->evolve ? s->evolve(s, t0, tout, y) : -99;
}
int stub_stepper_fullrhs(MRIStepInnerStepper s, double t, N_Vector y, N_Vector f, int mode) {
  return s->fullrhs ? s->fullrhs(s, t, y, f, mode) : -99;
}
int stub_stepper_reset(MRIStepInnerStepper s, double tR, N_Vector y) {
  return s->reset ? s->reset(s, tR, y) : -99;
}
int stub_ls_gettype(SUNLinearSolver S) { return (int)S->ops->gettype(S); }
int stub_ls_initialize(SUNLinearSolver S) { return S->ops->initialize(S); }
int stub_ls_setup(SUNLinearSolver S, SUNMatrix A) { return S->ops->setup(S, A); }
int stub_ls_solve(SUNLinearSolver S, SUNMatrix A, N_Vector x, N_Vector b, double tol) {
  return S->ops->solve(S, A, x, b, tol);
}
long stub_ls_lastflag(SUNLinearSolver S) { return S->ops->lastflag(S); }
int stub_ls_free(SUNLinearSolver S) { return S->ops->free(S); }

}  // extern "C"
