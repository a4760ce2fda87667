/* C ABI of the MI355X (gfx950) reacting-flow right-hand-side library.
 *
 * Every array argument is a device pointer (HIP device memory) laid out as the
 * reference's ManyVector (euler3D.hpp:62-65, multirate_chem_hydro_main.cpp:320-352):
 *   w[0..4]  rho, mx, my, mz, et   each nx*ny*nz doubles, cell i + nx*(j + ny*k)
 *   w[5]     chemistry            nx*ny*nz*nchem doubles, species fastest
 * A context holds the grid, boundary conditions, unit factors, the HIP stream all work
 * is queued on, and the device scratch the reference keeps in EulerData/cvklu_data.
 *
 * Return codes follow the SUNDIALS callback convention the reference uses:
 *   0 success, > 0 recoverable failure, < 0 unrecoverable failure.
 * Calls on one context are not reentrant (as the reference's, which share flux scratch
 * and the per-cell rate cache). Every call is synchronous with respect to the host
 * unless its name ends in _async.
 */
#ifndef MVD_RHS_H
#define MVD_RHS_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct mvd_context mvd_context;

/* error codes (all < 0 are unrecoverable) */
#define MVD_SUCCESS         0
#define MVD_ILLEGAL_STATE  -1   /* fEuler's legal_state failure, utilities.cpp:83-84,132-133 */
#define MVD_BAD_ARGUMENT   -2
#define MVD_DEVICE_ERROR   -3
#define MVD_UNSUPPORTED    -4
#define MVD_NOT_READY      -5   /* chemistry used before mvd_chem_setup */

/* boundary condition codes, euler3D.hpp:90-93 */
#define MVD_BC_PERIODIC   0
#define MVD_BC_NEUMANN    1
#define MVD_BC_DIRICHLET  2
#define MVD_BC_REFLECTING 3

/* Problem description; the fields of EulerData that the hot path reads
 * (euler3D.hpp:189-233). Single rank: the local grid is the global grid. */
typedef struct {
  long nx, ny, nz;         /* grid (each >= 3) */
  int nchem;               /* advected species per cell (0, 1, 2, 4, 6 or 10) */
  int bc[6];               /* xl, xr, yl, yr, zl, zr */
  double gamma;            /* ratio of specific heats */
  double xl, xr, yl, yr, zl, zr;
  double cfl;              /* stability() safety fraction */
  double mass_units, length_units, time_units;   /* code -> CGS factors (UpdateUnits, euler3D.hpp:385-393) */
} mvd_params;

/* Context lifetime. Replaces EulerData::SetupDecomp's allocations (euler3D.hpp:396-574)
 * for one rank. stream: a hipStream_t, or NULL for the null stream. */
int mvd_create(const mvd_params *p, void *stream, mvd_context **out);
void mvd_destroy(mvd_context *ctx);
int mvd_set_stream(mvd_context *ctx, void *stream);

/* fEuler (utilities.cpp:17-253; ARKRhsFn signature euler3D.hpp:1480) with zero external
 * forcing: wdot = -div F(w). Returns MVD_ILLEGAL_STATE when any cell has rho, et or p <= 0
 * (mvd_last_state_flags gives the 1/2/4 bits). */
int mvd_fe(mvd_context *ctx, double t, const double *const *w, double *const *wdot);
/* The same result from the plain one-thread-per-cell kernels with the reference's operation
 * order and no FMA contraction (the comparator; bitwise reproducible on the CPU). */
int mvd_fe_plain(mvd_context *ctx, double t, const double *const *w, double *const *wdot);
/* Queue fe() without waiting or checking the legal-state flag (for stage loops and graph
 * capture); the flag is read by the next mvd_last_state_flags call. */
int mvd_fe_async(mvd_context *ctx, double t, const double *const *w, double *const *wdot);
/* Legal-state bits (1 rho, 2 et, 4 p <= 0) of the last mvd_fe, mvd_fe_plain, mvd_fe_async or
 * mvd_fslow call. The evolution drivers (mvd_erk_evolve, mvd_mri_evolve, mvd_imex_evolve)
 * gather the bits of their stage evaluations in a flag of their own and report them through
 * their return code (MVD_ILLEGAL_STATE); they do not change what this returns. */
int mvd_last_state_flags(mvd_context *ctx);
/* Vector-Jacobian product of fe() (no counterpart in the reference, which has no adjoint;
 * for adjoint/sensitivity callers): wbar = (d wdot / d w)^T v, v and wbar in the layout of
 * wdot. The derivative of the Lax-Friedrichs speed max() follows its first maximum.
 * Keeps no face buffer. On first use it allocates its scratch: a side buffer for the stencil
 * adjoints that leave an x-line or a y/z column segment (12*(5+nchem) doubles per x-line,
 * 6*(5+nchem) per column segment) and, with tracers, two tracer-major planes of
 * nchem*nx*ny*nz doubles each (w and the tracer adjoint). */
int mvd_fe_vjp(mvd_context *ctx, double t, const double *const *w, const double *const *v,
               double *const *wbar);

/* Ghost exchange of the single subdomain: the receive buffers of EulerData::ExchangeStart /
 * ExchangeEnd (euler3D.hpp:577-1191) with the rank as its own periodic neighbour and the
 * physical-boundary fills of euler3D.hpp:793-1166 (low faces mirror w[2-s], high faces copy
 * w[n-3+s]; reflecting negates the normal momentum, Dirichlet every value). recv[6] =
 * W, E, S, N, B, F device buffers (NULL entries are skipped) of 3*na*nb*(5+nchem) doubles
 * in the reference's layout BUFINDX(v,s,a,b,NVAR,3,na,nb) = v + NVAR*(s + 3*(a + na*b)):
 * W/E (s,j,k), S/N (s,i,k), B/F (s,i,j); slot s of a low face is the point s-3 outside it,
 * of a high face the point n+s. */
int mvd_exchange(mvd_context *ctx, const double *const *w, double *const *recv);
/* pack1D_{x,y,z}[_bdry] (euler3D.hpp:1197-1378): for each of ncell cells (ijk: 3 ints per
 * cell, device) the stencil w(c-3..c+2) along dir (0 x, 1 y, 2 z) with the ghost rules of
 * mvd_exchange, out[cell][6][5+nchem] (device). Cells outside the grid: MVD_BAD_ARGUMENT. */
int mvd_pack1d(mvd_context *ctx, const double *const *w, int dir, long ncell, const int *ijk,
               double *out);

/* stability (utilities.cpp:483-528; ARKStep stability callback, euler3D_main.cpp:272-275):
 * alpha = max over cells of |mx/rho| + c, dt_stab = cfl * min(dx,dy,dz) / alpha. */
int mvd_max_wave_speed(mvd_context *ctx, const double *const *w, double *alpha);
int mvd_stability(mvd_context *ctx, const double *const *w, double t, double *dt_stab);

/* fslow (multirate_chem_hydro_main.cpp:996-1083) / fexpl (imex_chem_hydro_main.cpp:910-1000):
 * scales chem by the Dengo scale, overwrites w[4] (et) from the gas energy and kinetic
 * energy, calls fe(), moves etdot into the gas-energy slot of wdot[5], zeroes wdot[4] and
 * unscales chem. Requires nchem > 0 and mvd_chem_setup. */
int mvd_fslow(mvd_context *ctx, double t, double *const *w, double *const *wdot);
/* PostprocessStep of the IMEX driver (imex_chem_hydro_main.cpp:1002-1050): apply the Dengo
 * scaling to w[5] in place, rebuild et = ge / EnergyUnits + |m|^2 / (2 rho) from the gas
 * energy and the momenta, unapply the scaling (x * s / s: the chemistry's bits may change,
 * as in the reference). Requires nchem = 10 and mvd_chem_setup. */
int mvd_postprocess_step(mvd_context *ctx, double t, double *const *w);

/* Kernel timing with HIP events on the context's stream (the reference times every RHS
 * and face flux with its Profile class, profiler.hpp:23-206). mvd_profile(ctx,1) resets
 * and enables; mvd_profile_read returns the number of slots and fills per-slot total
 * milliseconds and launch counts. Slots: 0, 1 unused (kept so that the numbering stays),
 * 2 the z-face kernels (fluid, then tracers when nchem > 0), 3 fe_plain_faces,
 * 4 fe_plain_div, 5 fe_xyface_div (the x- and y-faces for every nchem). */
int mvd_profile(mvd_context *ctx, int on);
int mvd_profile_read(mvd_context *ctx, double *ms, long *count, int nslot);

/* ---- primordial chemistry (raja_primordial_network.cpp) ---- */

/* cvklu_setup_data (raja_primordial_network.cpp:45-254): tables_host points to the 50
 * rate/cooling/gamma tables (50 x 1024 doubles, host memory, committed blob order);
 * ncell cells; current_z redshift (the demo uses -1). Ts starts at 1000 K, scale at 1. */
int mvd_chem_setup(mvd_context *ctx, const double *tables_host, long ncell, double current_z);
/* prepare_Dengo_structures (primordial_blast.cpp:368-400): scale <- chem, inv_scale <- 1/chem,
 * chem <- 1, then setting_up_extra_variables (raja_primordial_network.cpp:1969-2018). */
int mvd_chem_prepare(mvd_context *ctx, double *chem);
/* apply_/unapply_Dengo_scaling (primordial_blast.cpp:404-458): chem *= scale / chem /= scale */
int mvd_chem_apply_scaling(mvd_context *ctx, double *chem);
int mvd_chem_unapply_scaling(mvd_context *ctx, double *chem);
/* calculate_rhs_cvklu (raja_primordial_network.cpp:1015-1389): ydot(y); updates the per-cell
 * temperature Ts (warm start of the next call) and dTs/dge. */
int mvd_chem_rhs(mvd_context *ctx, double t, const double *y, double *ydot);
/* calculate_jacobian_cvklu (raja_primordial_network.cpp:1393-1965): dense column-major
 * 10x10 blocks J[100*cell + 10*col + row] (DENSEIDX, line 40), evaluated from the
 * temperature of the last mvd_chem_rhs call. */
int mvd_chem_jac(mvd_context *ctx, double t, const double *y, double *J);
/* The same RHS and Jacobian times a factor applied to every entry after the reference's
 * expressions: ffast / fimpl (multirate_chem_hydro_main.cpp:898-920, imex_chem_hydro_main.cpp:
 * 811-840) scale the RHS by TimeUnits (N_VScale), Jfast / Jimpl (multirate...:954-994,
 * imex...:850-908) the Jacobian data (J *= TimeUnits, :980-982). A factor of 1 gives
 * mvd_chem_rhs / mvd_chem_jac bit for bit. */
int mvd_chem_rhs_scaled(mvd_context *ctx, double t, double rhs_scale, const double *y,
                        double *ydot);
int mvd_chem_jac_scaled(mvd_context *ctx, double t, double jac_scale, const double *y,
                        double *J);
/* N_Vector operations on device arrays, as the integrators' stage algebra uses them (the
 * ManyVector ops of SUNDIALS 6.2.0 that ARKStep/MRIStep call; no counterpart in the
 * reference tree, which links them from the library).
 * N_VLinearCombination: z = sum_k c[k] X[k] over len doubles, summed in k order; z may be
 * X[0] (no other X[k]). c is a host array of nv coefficients, X a host array of device
 * pointers. */
int mvd_nvec_linear_combination(mvd_context *ctx, int nv, const double *c,
                                const double *const *X, double *z, long len);
/* N_VConst: z[0..len) = c (device array). */
int mvd_nvec_const(mvd_context *ctx, double c, double *z, long len);
/* nsteps fixed explicit steps of y' = x with x constant (the fluid subvectors of the
 * multirate inner stepper, which have no fast dynamics): y = 1 y + a x, nsteps times, each
 * value's additions in turn (bitwise nsteps mvd_nvec_linear_combination calls with
 * c = {1, a}, X = {y, x}), in one launch. No counterpart in the reference; for the inner
 * stepper's Evolve (multirate_chem_hydro_main.cpp:1130-1160). */
int mvd_nvec_axpy_steps(mvd_context *ctx, int nsteps, double a, const double *x, double *y,
                        long len);
/* N_VWrmsNorm of a ManyVector of nsub (<= 8) subvectors: sqrt(sum_q sum_i (x_q[i] w_q[i])^2 /
 * sum_q len[q]), deterministic; x, w host arrays of device pointers, len host array. */
int mvd_nvec_wrms_norm(mvd_context *ctx, int nsub, const double *const *x,
                       const double *const *w, const long *len, double *out);

/* Explicit Runge-Kutta table (up to 6 stages; A row-major 6x6, strictly lower triangular). */
typedef struct {
  int stages;
  double A[36];
  double b[6];
  double c[6];
} mvd_erk_table;
/* ARKStep's default third-order explicit table, Bogacki-Shampine 3(2) (4 stages), the one the
 * fluid driver runs with --order=3 (euler3D_main.cpp:209). */
int mvd_erk_bs3(mvd_erk_table *tab);
/* Fixed-step explicit evolution of fEuler (euler3D_main.cpp:194 ARKStepCreate with fEuler
 * explicit, :280 ARKStepSetFixedStep, :357 ARKStepEvolve): nsteps steps of size h, in place
 * on w (device arrays of the ManyVector). Stages z_i = y + sum_{j<i} h a_ij F_j, F_i = fe(z_i),
 * y <- y + sum_i h b_i F_i; one step is captured as a device graph and replayed. Returns 0,
 * or MVD_ILLEGAL_STATE when any stage met an illegal state (w then holds the state after
 * nsteps steps, as an unchecked fixed-step run would), or an error code. */
int mvd_erk_evolve(mvd_context *ctx, const mvd_erk_table *tab, double t0, double h, int nsteps,
                   double *const *w);

/* Block linear solver of the chemistry subvector: Setup and Solve of SUNLinSol_RankLocalLS
 * (multirate_chem_hydro_main.cpp:1348-1411; the BDMPIMV twin, imex_chem_hydro_main.cpp:
 * 1075-1194) over MAGMA's batched getrf/getrs. lsetup factors A = I - gamma*J per cell,
 * J in the mvd_chem_jac layout, with partial pivoting; it returns 0, or the number (> 0,
 * recoverable, as SUNLS_LUFACT_FAIL) of blocks with an exact zero pivot, after which
 * lsolve refuses (MVD_NOT_READY) until a successful lsetup. lsolve solves A x = b for the
 * chemistry subvector ([cell][10]; x may equal b). */
int mvd_chem_lsetup(mvd_context *ctx, double gamma, const double *J);
int mvd_chem_lsolve(mvd_context *ctx, const double *b, double *x);
/* Setup of the block linear solver from the matrix itself, as SUNLinSolSetup(S, A) receives
 * it (ARKLS has formed A = I - gamma J in the SUNMatrix_MagmaDenseBlock before the call;
 * Setup_RankLocalLS passes A to the block solver, multirate_chem_hydro_main.cpp:1360-1368):
 * A in the mvd_chem_jac layout, factored as given. Returns as mvd_chem_lsetup. */
int mvd_chem_lsetup_matrix(mvd_context *ctx, const double *A);
/* Vector-Jacobian product of calculate_rhs_cvklu (no counterpart in the reference, which
 * has only the forward Jacobian callback; the adjoint callers are sensitivity and
 * checkpointed-adjoint drivers): ybar = (d ydot / d y)^T v in the scaled variables, for the
 * exact derivative of the RHS as evaluated (temperature = root of the temperature
 * equation, so dT/dy enters through dT/dy = -(dg/dy)/(dg/dT); rates piecewise linear in
 * log T). Updates Ts/dTs like mvd_chem_rhs; ydot (may be NULL) receives the RHS itself. */
int mvd_chem_vjp(mvd_context *ctx, double t, const double *y, const double *v, double *ybar,
                 double *ydot);
/* Self-check of the device temperature solve (no counterpart in the reference; for tests
 * and diagnostics): the ten Newton steps of cvklu_calculate_temperature
 * (raja_primordial_network.cpp:908-1011) end early at the first bitwise repeat of the
 * iteration (a cycle of period <= 4, whose remaining steps are known). Runs the early form
 * and the full ten steps from the warm starts Ts0 ([cell]) for every cell of y (scaled,
 * [cell][10]) and stores the number of cells whose (T, dTs_ge) differ bitwise; the
 * per-cell chemistry state is not changed. */
int mvd_chem_temperature_check(mvd_context *ctx, const double *y, const double *Ts0,
                               long *mismatches);
/* device pointers of per-cell chemistry state: 0 scale, 1 inv_scale, 2 Ts, 3 dTs_ge,
 * 4 mdensity, 5 inv_mdensity, 6 cie_optical_depth, 7 h2_optical_depth */
double *mvd_chem_state(mvd_context *ctx, int which);

/* ---- implicit chemistry step (the fast stage; replaces ARKStep DIRK + SUNDIALS Newton +
 * SUNLinSol_MagmaDense/KLU, primordial_ode_main.cpp:422-536, and the block-diagonal
 * SUNLinSol_RankLocalLS, multirate_chem_hydro_main.cpp:1312-1411) ---- */

/* DIRK table and Newton controls. The table is SDIRK (every diagonal entry the same
 * positive value) or ESDIRK (an explicit first stage, z_1 = y_n: first row zero, the other
 * diagonal entries equal and positive), stiffly accurate (y_{n+1} is the last stage). */
typedef struct {
  int nstage;              /* 1..7 (ESDIRK: 2..7) */
  double A[49];            /* row-major 7x7 Butcher matrix (entries past nstage zero) */
  double c[7];
  double rtol, atol;       /* WRMS weights 1/(rtol|z|+atol) of the Newton correction */
  double newton_tol;       /* ARKODE's test (arkStep_NlsConvTest, SUNDIALS 6.2.0): iteration m
                              of a stage solve converged when min(crate, 1) del / newton_tol
                              <= 1, del the WRMS norm of its correction, crate = max(0.3
                              crate, del / del_prev) for m > 0 (1 after every Jacobian
                              setup, carried from stage to stage) */
  int max_iters;           /* Newton iterations before the Jacobian is re-evaluated (1..100);
                              a correction above 2.3 times the previous one (m > 0) ends
                              the iterations early the same way */
  int max_setups;          /* Jacobian re-evaluations per stage before a cell fails (0..20) */
  double rhs_scale;        /* 1 for frhs, TimeUnits for ffast/fimpl */
  int max_halvings;        /* per-cell step rejection (batched solver only; 0..20, 0 = off):
                              a cell whose Newton budget runs out retries its step in halves,
                              up to this many times, the stand-in for the fast integrator's
                              step-size control; a cell that fails at the finest size keeps
                              its last accepted state */
  int max_steps_between_setups; /* ARKODE's msbp (its default 20): within one
                              mvd_chem_dirk_steps call (and the multirate driver's fast
                              steps) a step starts with a Jacobian setup at y_n only when
                              this many steps have passed since the cell's last setup, or
                              its previous step failed or was halved; the others start
                              Newton with the factors they hold. 0 or 1: every step (the
                              single-step entry points always set up) */
} mvd_dirk_opts;

typedef struct {
  long newton_iters;       /* over all cells */
  int max_cell_iters;      /* largest per-cell Newton count */
  long failed_cells;       /* cells that did not converge (recoverable failure) */
  long setups;             /* Jacobian evaluations + LU factorisations */
  long hist[16];           /* cells by Newton count (last bin: >= 15) */
  double ms;               /* device time of the step */
} mvd_dirk_stats;

/* Alexander's 3-stage, 3rd-order, L-stable SDIRK and default Newton controls */
int mvd_dirk_sdirk3(mvd_dirk_opts *o);
/* The same defaults with the ESDIRK part of ARK4(3)7L[2]SA (Kennedy & Carpenter 2019; ARKODE's
 * ARK437L2SA_DIRK_7_3_4, the implicit table the reference's multirate and IMEX inputs select,
 * itable = 112 in tests/primordial_blast/input_primordial_blast_mr.txt:48): seven stages,
 * gamma = 0.1235, fourth order, stiffly accurate. No counterpart call in the reference (it
 * selects the table through ARKStepSetTableNum, multirate_chem_hydro_main.cpp). */
int mvd_dirk_ark437(mvd_dirk_opts *o);
/* One step of size h of every cell, y (ncell x 10, scaled, device) overwritten with
 * y_{n+1}. plain = 0: the batched persistent kernel; plain = 1: the reference-structured
 * multi-kernel path. Returns 0, or the number of failed cells (> 0, recoverable). */
int mvd_chem_dirk_step(mvd_context *ctx, double t, double h, const mvd_dirk_opts *o, double *y,
                       int plain, mvd_dirk_stats *stats);
/* Allocation-free variant of mvd_chem_dirk_step with plain = 0 (no counterpart in the
 * reference; for callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same kernels and the same results
 * bit for bit; the work buffer (cell order, sort keys, bucket counters, per-block statistics)
 * and the work-queue and statistics counters live in a library-owned array of 22.4 MiB per
 * GPU instead of the context's lazily allocated buffer, so the call allocates nothing.
 * Limits: at most 1048576 cells (mvd_chem_setup's ncell) and 4096 blocks in the persistent
 * grid, else MVD_BAD_ARGUMENT before any device work; tables of at most 4 stages (a longer
 * one, such as mvd_dirk_ark437's, needs a stage buffer the array does not hold:
 * MVD_UNSUPPORTED). Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (other streams, threads or contexts) race and
 * must be serialised by the caller. stats->ms is host wall time of the call. */
int mvd_chem_dirk_step_noalloc(mvd_context *ctx, double t, double h, const mvd_dirk_opts *o,
                               double *y, mvd_dirk_stats *stats);
/* The same step of y' = rhs_scale*f(y) + force with a constant forcing force ([cell][10],
 * scaled; the forcing MRIStep hands its fast stepper, applied by ffast at
 * multirate_chem_hydro_main.cpp:922-946), batched kernel only. */
int mvd_chem_dirk_step_forced(mvd_context *ctx, double t, double h, const mvd_dirk_opts *o,
                              double *y, const double *force, mvd_dirk_stats *stats);
/* nsteps >= 1 fixed steps of size h of y' = rhs_scale*f(y) (+ force, nullable, constant over
 * all of them), batched kernel only, in one launch: each cell runs its steps back to back
 * with its state in registers, every step exactly as one mvd_chem_dirk_step(_forced) call
 * would run it (a failed step counts and the next starts from the state it kept), so y and
 * the temperatures end bitwise as after nsteps such calls. The inner stepper's Evolve of the
 * multirate driver (multirate_chem_hydro_main.cpp:1130-1160: fixed steps within a slow
 * stage, one forcing). Returns the number of failed cell steps over all steps, or an error
 * code; stats are per cell over all its steps (newton_iters, setups and the histogram of
 * per-cell totals, failed_cells = failed cell steps). */
int mvd_chem_dirk_steps(mvd_context *ctx, double t, double h, int nsteps,
                        const mvd_dirk_opts *o, double *y, const double *force,
                        mvd_dirk_stats *stats);
/* Multirate evolution of the reacting flow, device-resident (configuration 5; the driver
 * multirate_chem_hydro_main.cpp:529-555 with MRIStep's default third-order coupling, MIS
 * with Knoth-Wolke 3, and a fixed slow step): nslow slow steps of size h on w (the six
 * ManyVector subvectors, chemistry scaled by mvd_chem_prepare, nchem = 10). Each slow stage
 * evaluates fslow at the running state (its total-energy overwrite included), forms the
 * constant MIS forcing and takes ceil((c_i - c_{i-1}) h / hmax_fast) equal fixed fast steps
 * (the inner stepper's maximum step, input hmax): fluid += dt*forcing, chemistry by a
 * forced batched DIRK step with the options fast (rhs_scale = TimeUnits gives ffast). One
 * slow step is captured as a device graph and replayed. flags: 0 keeps the reference's
 * forcing units (fslow's chemistry rates are physical and its gas-energy slot holds the
 * code-unit total-energy rate, added unchanged to the Dengo-scaled fast variables);
 * MVD_MRI_CONSISTENT_FORCING converts them to the scaled variables first (species rates
 * divided by scale, the energy slot times EnergyUnits divided by the ge scale). Returns 0,
 * the number (> 0) of failed fast cell steps, MVD_ILLEGAL_STATE if any slow RHS met an
 * illegal state, or an error code. */
#define MVD_MRI_CONSISTENT_FORCING 1
/* Fast steps mvd_mri_evolve takes in slow stage `stage` (2..4, the stages after the first
 * of Knoth-Wolke 3): ceil((c_stage - c_{stage-1}) h / hmax_fast), at least 1. */
int mvd_mri_fast_steps(int stage, double h, double hmax_fast);
int mvd_mri_evolve(mvd_context *ctx, double t0, double h, int nslow, double hmax_fast,
                   const mvd_dirk_opts *fast, int flags, double *const *w);

/* IMEX additive Runge-Kutta tables: explicit Ae (strictly lower triangular) and diagonally
 * implicit Ai (lower triangular, diagonal >= 0; a zero diagonal is an explicit stage),
 * 8x8 row-major, common b and c. */
typedef struct {
  int stages;              /* 1..8 */
  double Ae[64];
  double Ai[64];
  double b[8];
  double c[8];
} mvd_ark_table;
/* ARK3(2)4L[2]SA (Kennedy & Carpenter 2003), ARKStep's default third-order IMEX pair. */
int mvd_ark_ark324(mvd_ark_table *t);
/* ARK4(3)7L[2]SA (Kennedy & Carpenter 2019), the pair the reference's IMEX inputs select
 * (etable 13, itable 112 in tests/primordial_blast/input_primordial_blast_imex.txt:48-49):
 * seven stages, fourth order, its implicit part mvd_dirk_ark437's table. No counterpart call
 * in the reference (it selects the tables through ARKStepSetTableNum). */
int mvd_ark_ark437(mvd_ark_table *t);
/* IMEX evolution of the reacting flow, device-resident: the IMEX driver
 * (imex_chem_hydro_main.cpp:337-476, ARKStepCreate(fexpl, fimpl), the block-diagonal
 * chemistry solver BDMPIMV :1075-1194, PostprocessStep :1002-1050) with a fixed step h:
 * nsteps steps on w (six subvectors, chemistry scaled by mvd_chem_prepare, nchem = 10).
 * fexpl = fslow (its total-energy overwrite of the stage included), fimpl = rhs_scale
 * (TimeUnits) times calculate_rhs_cvklu on the chemistry. Each implicit stage solves
 * z - h a_ii fimpl(z) = R per cell with the batched Newton/LU kernel and the controls of
 * newton (its table is not used); after each step the total energy is rebuilt from the
 * gas energy as PostprocessStep does. flags: MVD_IMEX_CONSISTENT_FORCING converts fexpl's
 * chemistry rates to the scaled variables (as for mvd_mri_evolve); 0 keeps the
 * reference's units. Returns 0, the number (> 0) of failed cell solves, MVD_ILLEGAL_STATE
 * if fexpl met an illegal state, or an error code. */
#define MVD_IMEX_CONSISTENT_FORCING 1
int mvd_imex_evolve(mvd_context *ctx, double t0, double h, int nsteps, const mvd_ark_table *tab,
                    const mvd_dirk_opts *newton, int flags, double *const *w);
/* Checkpointed discrete adjoint of nsteps chemistry DIRK steps (no counterpart in the
 * reference, which has no adjoint; for sensitivity drivers, e.g. of the domain-integrated
 * energy to the initial species). y (scaled, [cell][10]) is advanced nsteps steps of size h
 * by a thread-per-cell run of the plain solver's per-cell algorithm, keeping the state every
 * ckpt_every steps; then lam0 = (d y_N / d y_0)^T lamT (both [cell][10]) for the stage
 * equations solved exactly along that trajectory, with the exact stage Jacobians; each
 * segment is recomputed once from its checkpoint, keeping its steps' stage values. o must
 * be an SDIRK table of at most 4 stages, stiffly accurate (the last row of A sums to 1, as
 * mvd_dirk_sdirk3; an ESDIRK table gives MVD_UNSUPPORTED). Ts/dTs end at the
 * forward run's final state. Returns 0, the number (> 0) of cells whose forward Newton
 * budget ran out (lam0 is still computed along the computed trajectory), or an error code.
 * The cells run in the batched step's temperature order at y_0 (from 4096 cells on).
 * Device scratch held by the context: ceil(nsteps/ckpt_every) * 12
 * + ckpt_every * (10 * nstage + 1) doubles and 22 bytes (the cell order) per cell, and
 * 128 KiB. */
int mvd_chem_dirk_adjoint(mvd_context *ctx, double t, double h, int nsteps, int ckpt_every,
                          const mvd_dirk_opts *o, double *y, const double *lamT, double *lam0);

/* Autotune the batched kernel's launch shape (threads per block, resident waves per CU) by
 * timing one step on a scratch copy of y; the per-cell temperatures are restored. The
 * choice is kept in the context; returns the best time in ms through best_ms. */
int mvd_chem_dirk_tune(mvd_context *ctx, double h, const mvd_dirk_opts *o, const double *y,
                       double *best_ms);
/* Set the batched kernel's launch shape by hand (what mvd_chem_dirk_tune chooses): block
 * threads per block (64, 128 or 256; blocks whose LDS would not fit a CU are split at
 * launch) and waves resident waves of the persistent grid (0: one per SIMD of the device).
 * Each wave takes 64-cell chunks, the first by its index and the rest from the work queue,
 * so waves smaller than the number of chunks exercises the queue. The results do not
 * depend on the shape. get_shape returns the shape the next launch uses (waves resolved). */
int mvd_chem_dirk_set_shape(mvd_context *ctx, int block, int waves);
/* The block split of a batched launch (no GPU needed; diagnostics and tests): the threads
 * per block a launch asked for `block` threads runs with for an nstage-stage table, when the
 * kernel's static LDS is static_lds bytes and the device allows lds_limit bytes per block
 * (the launch asks the runtime for both). The dynamic LDS is 8 * (4*1024 + waves * (24 +
 * 10*nstage) * 64) bytes: the gamma tables, then per wave its cells' constants, the stage
 * derivatives of the earlier stages and the step start; for tables of more than four stages
 * (and explicit-first-stage tables, which this call does not see) the stage derivatives go
 * to a device buffer of 80 * (nstage - 1) bytes per thread of the persistent grid instead,
 * 8 * (4*1024 + waves * 34 * 64) bytes of LDS. Halves the block until it fits; 0 when not
 * even one wave fits. */
int mvd_chem_dirk_fit_block(int block, int nstage, long static_lds, long lds_limit);
int mvd_chem_dirk_get_shape(mvd_context *ctx, int *block, int *waves);

#ifdef __cplusplus
}
#endif
#endif
