/* The public SUNDIALS 6.2.0 declarations the ARKODE adapters (include/mvd_arkode.h) use:
 * type names, callback typedefs, constants and function signatures only, restated so that
 * the adapters compile where SUNDIALS is not installed. Built against a SUNDIALS install
 * (MVD_ARKODE_USE_SUNDIALS defined), the adapters include SUNDIALS' own headers instead and
 * this file is not read. Each object is declared as SUNDIALS declares it (sundials_types.h,
 * sundials_nvector.h, nvector_manyvector.h, nvector_mpimanyvector.h, sundials_matrix.h,
 * sunmatrix_magmadense.h, sundials_linearsolver.h, arkode.h, arkode_ls.h, arkode_mristep.h);
 * the implementations come from the SUNDIALS libraries the driver links. */
#ifndef MVD_SUNDIALS_MIN_H
#define MVD_SUNDIALS_MIN_H
#ifdef __cplusplus
extern "C" {
#endif

/* sundials_types.h */
typedef double realtype;
typedef long int sunindextype;
typedef int booleantype;
typedef struct _SUNContext *SUNContext;

/* sundials_nvector.h, nvector_manyvector.h, nvector_mpimanyvector.h */
typedef struct _generic_N_Vector *N_Vector;
realtype *N_VGetDeviceArrayPointer(N_Vector v);
sunindextype N_VGetLength(N_Vector v);
N_Vector N_VGetSubvector_ManyVector(N_Vector v, sunindextype vec_num);
sunindextype N_VGetNumSubvectors_ManyVector(N_Vector v);
N_Vector N_VGetSubvector_MPIManyVector(N_Vector v, sunindextype vec_num);
sunindextype N_VGetNumSubvectors_MPIManyVector(N_Vector v);

/* sundials_matrix.h, sunmatrix_magmadense.h */
typedef struct _generic_SUNMatrix *SUNMatrix;
realtype *SUNMatrix_MagmaDense_Data(SUNMatrix A);
sunindextype SUNMatrix_MagmaDense_LData(SUNMatrix A);
sunindextype SUNMatrix_MagmaDense_NumBlocks(SUNMatrix A);
sunindextype SUNMatrix_MagmaDense_BlockRows(SUNMatrix A);
sunindextype SUNMatrix_MagmaDense_BlockColumns(SUNMatrix A);

/* sundials_linearsolver.h */
typedef enum {
  SUNLINEARSOLVER_DIRECT,
  SUNLINEARSOLVER_ITERATIVE,
  SUNLINEARSOLVER_MATRIX_ITERATIVE,
  SUNLINEARSOLVER_MATRIX_EMBEDDED
} SUNLinearSolver_Type;
typedef int SUNLinearSolver_ID;   /* an enumeration in SUNDIALS; not used by the adapters */
typedef int (*SUNATimesFn)(void *A_data, N_Vector v, N_Vector z);
typedef int (*SUNPSetupFn)(void *P_data);
typedef int (*SUNPSolveFn)(void *P_data, N_Vector r, N_Vector z, realtype tol, int lr);
typedef struct _generic_SUNLinearSolver_Ops *SUNLinearSolver_Ops;
typedef struct _generic_SUNLinearSolver *SUNLinearSolver;
struct _generic_SUNLinearSolver_Ops {
  SUNLinearSolver_Type (*gettype)(SUNLinearSolver);
  SUNLinearSolver_ID (*getid)(SUNLinearSolver);
  int (*setatimes)(SUNLinearSolver, void *, SUNATimesFn);
  int (*setpreconditioner)(SUNLinearSolver, void *, SUNPSetupFn, SUNPSolveFn);
  int (*setscalingvectors)(SUNLinearSolver, N_Vector, N_Vector);
  int (*setzeroguess)(SUNLinearSolver, booleantype);
  int (*initialize)(SUNLinearSolver);
  int (*setup)(SUNLinearSolver, SUNMatrix);
  int (*solve)(SUNLinearSolver, SUNMatrix, N_Vector, N_Vector, realtype);
  int (*numiters)(SUNLinearSolver);
  realtype (*resnorm)(SUNLinearSolver);
  sunindextype (*lastflag)(SUNLinearSolver);
  int (*space)(SUNLinearSolver, long int *, long int *);
  N_Vector (*resid)(SUNLinearSolver);
  int (*free)(SUNLinearSolver);
};
struct _generic_SUNLinearSolver {
  void *content;
  SUNLinearSolver_Ops ops;
  SUNContext sunctx;
};
SUNLinearSolver SUNLinSolNewEmpty(SUNContext sunctx);
void SUNLinSolFreeEmpty(SUNLinearSolver S);
#define SUNLS_SUCCESS 0
#define SUNLS_MEM_NULL -801
#define SUNLS_ILL_INPUT -802
#define SUNLS_MEM_FAIL -803
#define SUNLS_PACKAGE_FAIL_UNREC -809
#define SUNLS_LUFACT_FAIL 808

/* arkode.h, arkode_ls.h: the user-supplied function types */
typedef int (*ARKRhsFn)(realtype t, N_Vector y, N_Vector ydot, void *user_data);
typedef int (*ARKExpStabFn)(N_Vector y, realtype t, realtype *hstab, void *user_data);
typedef int (*ARKPostProcessFn)(realtype t, N_Vector y, void *user_data);
typedef int (*ARKLsJacFn)(realtype t, N_Vector y, N_Vector fy, SUNMatrix Jac, void *user_data,
                          N_Vector tmp1, N_Vector tmp2, N_Vector tmp3);

/* arkode_mristep.h: the custom inner stepper */
typedef struct _MRIStepInnerStepper *MRIStepInnerStepper;
typedef int (*MRIStepInnerEvolveFn)(MRIStepInnerStepper stepper, realtype t0, realtype tout,
                                    N_Vector y);
typedef int (*MRIStepInnerFullRhsFn)(MRIStepInnerStepper stepper, realtype t, N_Vector y,
                                     N_Vector f, int mode);
typedef int (*MRIStepInnerResetFn)(MRIStepInnerStepper stepper, realtype tR, N_Vector yR);
int MRIStepInnerStepper_Create(SUNContext sunctx, MRIStepInnerStepper *stepper);
int MRIStepInnerStepper_Free(MRIStepInnerStepper *stepper);
int MRIStepInnerStepper_SetContent(MRIStepInnerStepper stepper, void *content);
int MRIStepInnerStepper_GetContent(MRIStepInnerStepper stepper, void **content);
int MRIStepInnerStepper_SetEvolveFn(MRIStepInnerStepper stepper, MRIStepInnerEvolveFn fn);
int MRIStepInnerStepper_SetFullRhsFn(MRIStepInnerStepper stepper, MRIStepInnerFullRhsFn fn);
int MRIStepInnerStepper_SetResetFn(MRIStepInnerStepper stepper, MRIStepInnerResetFn fn);
int MRIStepInnerStepper_GetForcingData(MRIStepInnerStepper stepper, realtype *tshift,
                                       realtype *tscale, N_Vector **forcing, int *nforcing);
#define ARK_FULLRHS_START 0
#define ARK_FULLRHS_END 1
#define ARK_FULLRHS_OTHER 2

#ifdef __cplusplus
}
#endif
#endif
