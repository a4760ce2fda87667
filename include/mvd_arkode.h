/* ARKODE callbacks of the reacting-flow demo over libmvd_rhs.so (libmvd_arkode.so).
 *
 * Each function has the signature SUNDIALS 6.2.0 calls (ARKRhsFn, ARKLsJacFn, ARKExpStabFn,
 * ARKPostProcessFn, the SUNLinearSolver ops, the MRIStepInnerStepper functions) and the
 * semantics of the reference callback it replaces, cited per function. A driver hands them
 * to ARKStep/MRIStep unchanged (ARKStepCreate(mvd_arkode_fslow, ...), ARKStepSetJacFn(...,
 * mvd_arkode_Jfast), ...) with a mvd_arkode_data as user data.
 *
 * Vectors: the ManyVector / MPIManyVector of the reference (multirate_chem_hydro_main.cpp:
 * 320-352) whose six subvectors are device vectors (N_VGetDeviceArrayPointer gives HIP
 * device memory): rho, mx, my, mz, et (nx*ny*nz each) and the chemistry (nx*ny*nz*10,
 * species fastest, Dengo-scaled). Matrices: SUNMatrix_MagmaDenseBlock(ncell, 10, 10)
 * (multirate_chem_hydro_main.cpp:411), column-major blocks. Return codes: 0, > 0
 * recoverable, < 0 unrecoverable, as the reference's callbacks.
 *
 * Built with MVD_ARKODE_USE_SUNDIALS the SUNDIALS headers are used; otherwise the restated
 * declarations of mvd_sundials_min.h (same names and signatures).
 */
#ifndef MVD_ARKODE_H
#define MVD_ARKODE_H
#ifdef MVD_ARKODE_USE_SUNDIALS
#include <arkode/arkode.h>
#include <arkode/arkode_mristep.h>
#include <nvector/nvector_manyvector.h>
#include <nvector/nvector_mpimanyvector.h>
#include <sundials/sundials_linearsolver.h>
#include <sunmatrix/sunmatrix_magmadense.h>
#else
#include "../sundials-manyvector-demo_amd/arkode/mvd_sundials_min.h"
#endif
#include "mvd_rhs.h"
#ifdef __cplusplus
extern "C" {
#endif

/* User data of every adapter (the part of EulerData and RankLocalStepperContent the
 * callbacks read, multirate_chem_hydro_main.cpp:92-97). */
typedef struct {
  mvd_context *ctx;              /* grid, units, chemistry state (mvd_create, mvd_chem_setup) */
  double time_units;             /* TimeUnits: ffast / fimpl / Jfast / Jimpl scale factor */
  MRIStepInnerStepper stepper;   /* inner stepper whose forcing ffast adds (NULL: none) */
  int disable_forcing;           /* set by FullRhs around its ffast call */
  /* inner stepper (mvd_inner_stepper_create): fast steps of at most hmax_fast by the
   * batched forced DIRK kernel with the controls fast (its rhs_scale is time_units) */
  mvd_dirk_opts fast;
  double hmax_fast;
  double tcur;                   /* inner stepper time (Evolve, Reset) */
  long fast_steps;               /* fast steps taken by Evolve (statistics) */
  long failed_cells;             /* failed fast cell steps over all Evolve calls */
} mvd_arkode_data;

/* fEuler (utilities.cpp:17-253; declared euler3D.hpp:1480): wdot = fe(w), MPIManyVector. */
int mvd_arkode_fEuler(realtype t, N_Vector w, N_Vector wdot, void *user_data);
/* stability (utilities.cpp:483-528; attached at euler3D_main.cpp:273). */
int mvd_arkode_stability(N_Vector w, realtype t, realtype *dt_stab, void *user_data);
/* fslow (multirate_chem_hydro_main.cpp:996-1083) and fexpl (imex_chem_hydro_main.cpp:
 * 910-1000): w.et overwritten from the gas energy, fEuler, etdot moved to chemdot[ge]. */
int mvd_arkode_fslow(realtype t, N_Vector w, N_Vector wdot, void *user_data);
/* ffast (multirate_chem_hydro_main.cpp:898-952), ManyVector: wdot = 0, chemdot =
 * TimeUnits * calculate_rhs_cvklu(chem), then, unless forcing is disabled or there is no
 * stepper, sum_i tau^i F_i with tau = (t - tshift)/tscale added to all six subvectors
 * (MRIStepInnerStepper_GetForcingData, N_VLinearCombination per subvector). */
int mvd_arkode_ffast(realtype t, N_Vector w, N_Vector wdot, void *user_data);
/* fimpl (imex_chem_hydro_main.cpp:811-848), MPIManyVector: ffast without forcing. */
int mvd_arkode_fimpl(realtype t, N_Vector w, N_Vector wdot, void *user_data);
/* frhs (primordial_ode_main.cpp:681-701): the chemistry vector alone, no TimeUnits. */
int mvd_arkode_frhs(realtype t, N_Vector w, N_Vector wdot, void *user_data);
/* Jfast (multirate_chem_hydro_main.cpp:954-994, ManyVector) and Jimpl
 * (imex_chem_hydro_main.cpp:850-908, MPIManyVector): calculate_jacobian_cvklu on the
 * chemistry subvector into the dense block matrix, then J *= TimeUnits. */
int mvd_arkode_Jfast(realtype t, N_Vector w, N_Vector fw, SUNMatrix Jac, void *user_data,
                     N_Vector tmp1, N_Vector tmp2, N_Vector tmp3);
int mvd_arkode_Jimpl(realtype t, N_Vector w, N_Vector fw, SUNMatrix Jac, void *user_data,
                     N_Vector tmp1, N_Vector tmp2, N_Vector tmp3);
/* Jrhs (primordial_ode_main.cpp:703-721): the chemistry vector alone, no TimeUnits. */
int mvd_arkode_Jrhs(realtype t, N_Vector w, N_Vector fw, SUNMatrix Jac, void *user_data,
                    N_Vector tmp1, N_Vector tmp2, N_Vector tmp3);
/* PostprocessStep (imex_chem_hydro_main.cpp:1002-1050). */
int mvd_arkode_postprocess(realtype t, N_Vector w, void *user_data);

/* SUNLinSol_RankLocalLS (multirate_chem_hydro_main.cpp:1312-1411) and its IMEX twin
 * SUNLinSol_BDMPIMV (imex_chem_hydro_main.cpp:1075-1194; on one rank its flag all-reduce is
 * the identity): a DIRECT linear solver of the ManyVector x that solves subvector subvec
 * only, with libmvd_rhs.so's batched 10x10 LU (mvd_chem_lsetup_matrix / mvd_chem_lsolve)
 * in place of the MAGMA block solver. Setup returns SUNLS_LUFACT_FAIL when a block is
 * singular; Solve returns SUNLS_MEM_FAIL when x or b lacks the subvector. Returns NULL when
 * x has no subvector subvec or the solver cannot be created. */
SUNLinearSolver mvd_SUNLinSol_RankLocalLS(N_Vector x, sunindextype subvec,
                                          mvd_arkode_data *data, SUNContext sunctx);
SUNLinearSolver_Type mvd_RankLocalLS_GetType(SUNLinearSolver S);
int mvd_RankLocalLS_Initialize(SUNLinearSolver S);
int mvd_RankLocalLS_Setup(SUNLinearSolver S, SUNMatrix A);
int mvd_RankLocalLS_Solve(SUNLinearSolver S, SUNMatrix A, N_Vector x, N_Vector b,
                          realtype tol);
sunindextype mvd_RankLocalLS_LastFlag(SUNLinearSolver S);
int mvd_RankLocalLS_Free(SUNLinearSolver S);

/* The rank-local fast stepper of the multirate driver (RankLocalStepper_*,
 * multirate_chem_hydro_main.cpp:1114-1224), with the batched chemistry kernel in place of
 * the inner ARKStep. Evolve(t0, tout, y): ceil((tout - t0)/hmax_fast) equal fixed steps of
 * y' = ffast(t, y); the MRI forcing must be constant over the step (nforcing == 1, as for
 * MRIStep's MIS methods such as the default Knoth-Wolke 3; else -1): the fluid subvectors
 * take y_q += dt F_q per step (mvd_nvec_axpy_steps), the chemistry forced batched DIRK
 * steps (options fast, rhs_scale = TimeUnits, per-cell step rejection per
 * fast.max_halvings), all of them in one mvd_chem_dirk_steps launch. Returns 0, the
 * number (> 0, recoverable) of failed fast cell steps, or < 0. FullRhs: ffast with the
 * forcing disabled (any mode). Reset(tR, yR): the stepper's time becomes tR; it keeps no
 * other history (the reference's ARKStepReset clears the inner integrator's step history
 * and, like this one, leaves the chemistry's temperature warm start). */
int mvd_inner_evolve(MRIStepInnerStepper stepper, realtype t0, realtype tout, N_Vector y);
int mvd_inner_fullrhs(MRIStepInnerStepper stepper, realtype t, N_Vector y, N_Vector f,
                      int mode);
int mvd_inner_reset(MRIStepInnerStepper stepper, realtype tR, N_Vector yR);
/* MRIStepInnerStepper_Create with the content data and the three functions attached; sets
 * data->stepper. */
int mvd_inner_stepper_create(SUNContext sunctx, mvd_arkode_data *data,
                             MRIStepInnerStepper *stepper);

#ifdef __cplusplus
}
#endif
#endif
