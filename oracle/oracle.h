/* C ABI of the CPU oracle -- TEST INFRASTRUCTURE ONLY (see orc_common.hpp). */
#ifndef ORACLE_H
#define ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* single-rank grid description; bc[] = {xl, xr, yl, yr, zl, zr} codes of euler3D.hpp:90-93 */
typedef struct {
  long nx, ny, nz;
  int nchem;
  int bc[6];
  double gamma, dx, dy, dz;
} orc_grid;

/* one DIRK step of the chemistry (orc_dirk.cpp) */
typedef struct {
  int nstage;            /* 1..7 */
  double A[49];          /* 7x7 row-major Butcher matrix (lower triangular); SDIRK, or
                            ESDIRK when the first row is zero (explicit first stage) */
  double c[7];
  double rtol, atol, newton_tol;
  int max_iters, max_setups;
  double rhs_scale;      /* 1 (frhs) or TimeUnits (ffast / fimpl) */
  int max_halvings;      /* per-cell step rejection (0: off) */
  int max_steps_between_setups;   /* Jacobian setups in a multi-step call (0, 1: every step) */
} orc_dirk;
int orc_chem_dirk_step(void *chem, const orc_dirk *opt, double h, double *y, int *iters);
/* nsteps DIRK steps of every cell in one call (the setup rule of max_steps_between_setups),
 * with an optional constant forcing [cell][10]; iters (nullable) totals per cell */
int orc_chem_dirk_steps_forced(void *chem, const orc_dirk *opt, double h, int nsteps,
                               double *y, const double *force, int *iters);
/* discrete adjoint of nsteps DIRK steps (orc_dirk.cpp); exact scaled Jacobian of one cell
 * (orc_chem.cpp) */
int orc_chem_dirk_adjoint(void *chem, const orc_dirk *opt, double h, int nsteps, double *y,
                          const double *lamT, double *lam0);
int orc_chem_exact_jac_cell(void *chem, long i, const double *y, double *T, double *J);
int orc_chem_dirk_step_forced(void *chem, const orc_dirk *opt, double h, double *y,
                              const double *force, int *iters);
/* multirate evolution, MIS form of Knoth-Wolke 3 slow / forced DIRK fast (orc_mri.cpp) */
void orc_mri_kw3_table(double *c, double *A);
int orc_mri_evolve(const orc_grid *g, void *chem, double energy_units, const orc_dirk *fast,
                   double h, int nslow, double hmax, int consistent, double *const *w);
int orc_mri_fast_steps(int i, double h, double hmax);
/* fixed-step explicit Runge-Kutta evolution of fe() (orc_erk.cpp) */
typedef struct {
  int stages;
  double A[36];          /* 6x6 row-major, strictly lower triangular */
  double b[6];
  double c[6];
} orc_erk;
int orc_erk_evolve(const orc_grid *g, const orc_erk *tab, double h, int nsteps, double *const *w);
/* fixed-step IMEX additive Runge-Kutta evolution (orc_imex.cpp) */
typedef struct {
  int stages;
  double Ae[64];         /* 8x8 row-major explicit table, strictly lower triangular */
  double Ai[64];         /* 8x8 row-major diagonally implicit table */
  double b[8];
  double c[8];
} orc_ark;
void orc_ark324_table(orc_ark *t);
int orc_imex_evolve(const orc_grid *g, void *chem, double energy_units, const orc_ark *tab,
                    const orc_dirk *newton, double h, int nsteps, int consistent,
                    double *const *w);
int orc_lu(const double *M, double *LU, int *piv, const double *b, double *x);
/* single-rank ghost exchange buffers W, E, S, N, B, F (orc_fe.cpp) */
void orc_exchange(const orc_grid *g, const double *const *w, double *const *recv);
/* bitwise check of the early exit of the device temperature solve (orc_chem.cpp) */
long orc_temperature_exit_check(void *chem, const double *y, const double *Ts0, int max_period,
                                int *steps, int *period);
/* vector-Jacobian products of fe() and of one face flux (orc_vjp.cpp) */
int orc_face_flux_vjp(const double *stencil, int dir, int nchem, double gamma,
                      const double *Fbar, double *sbar);
int orc_fe_vjp(const orc_grid *g, const double *const *w, const double *const *v,
               double *const *wbar);
/* vector-Jacobian product of the chemistry RHS (orc_chem.cpp) */
int orc_chem_vjp(void *chem, const double *y, const double *v, double *ybar, int fixed_T);

void orc_set_threads(int n);
int orc_max_threads(void);
int orc_face_flux(const double *stencil, int dir, int nchem, double gamma, double *fface);
int orc_fe(const orc_grid *g, const double *const *w, double *const *wdot);
double orc_max_wave_speed(const orc_grid *g, const double *const *w);
double orc_stability_dt(const orc_grid *g, const double *const *w, double cfl);
void orc_face_flux_opcount(int nchem, unsigned long long *out);

#ifdef __cplusplus
}
#endif
#endif
