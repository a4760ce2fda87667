// Shared declarations of the per-cell chemistry DIRK solvers (batched and plain).
#pragma once
#include <hip/hip_runtime.h>
#include "mvd_chem.hpp"

namespace mvd {

constexpr int MAXSTAGE = 7;   // stages of the DIRK tables supported (A is 7x7)
constexpr int ADJ_MAXSTAGE = 4;   // stages of the SDIRK tables the adjoint supports
constexpr int NHIST = 16;     // Newton-iteration histogram bins (last bin: >= 15)

// statistics words in ChemState::counters (after the work-queue counter)
enum { STAT_ITERS = 0, STAT_MAXIT, STAT_FAILED, STAT_SETUPS, STAT_HIST, NSTAT = STAT_HIST + NHIST };

struct ChemPtrsD {
  const double *scale, *iscale, *md, *imd, *cie, *h2o;
  double *Ts, *dTs;
};

// Method and Newton controls, by value in kernel arguments.
struct DirkDev {
  int nstage;
  double A[MAXSTAGE * MAXSTAGE];   // row-major
  int esdirk;           // explicit first stage (z_1 = y_n, F_1 = f(y_n))
  double gamma;         // the implicit stages' diagonal entry
  double rtol, atol, ntol;
  int max_iters, max_setups;
  double rhs_scale;
  int max_halvings;     // per-cell step rejection (batched solver), 0: off
  int msbp;             // steps between Jacobian setups within one multi-step launch (>= 1)
};

// force (nullable): constant forcing added to the scaled RHS, [cell][10]. reuse_order:
// keep the cell order of the previous launch (same cells, work buffer not reallocated
// since) instead of sorting again. nsteps: steps of size h run per cell in this one
// launch (the same states as nsteps launches; statistics per cell over all its steps).
int dirk_batched_launch(ChemState &s, const DirkDev &o, double h, double *y, hipStream_t st,
                        int grid_waves, int block, const double *force = nullptr,
                        bool reuse_order = false, int nsteps = 1);
// The same step with its work buffer and counters in the library's per-GPU scratch array
// instead of the context's buffer (no device allocation); at most NOALLOC_MAX_CELLS cells
// and NOALLOC_MAX_ROWS blocks, else MVD_BAD_ARGUMENT. *counters_out: the counters it used
// (work-queue position, then NSTAT statistics), valid until the next such call on this GPU.
constexpr long NOALLOC_MAX_CELLS = 1L << 20;
constexpr long NOALLOC_MAX_ROWS = 4096;
int dirk_batched_launch_noalloc(ChemState &s, const DirkDev &o, double h, double *y,
                                hipStream_t st, int grid_waves, int block, const double *force,
                                int **counters_out);
// the current device's address of that scratch array (looked up once per device)
unsigned char *dirk_noalloc_scratch();
int dirk_plain_run(ChemState &s, const DirkDev &o, double h, double *y, hipStream_t st);
// threads per block of a batched launch (blocks split to fit the LDS limit; 0: none fits)
int dirk_fit_block(int block, int nstage, size_t static_lds, size_t limit, bool esdirk = false);
// checkpointed discrete adjoint of nsteps DIRK steps (checkpoint every K steps)
size_t chem_adjoint_scratch_doubles(long n, int nsteps, int K, int nstage);
int chem_adjoint_run(ChemState &s, const DirkDev &o, double h, int nsteps, int K, double *y,
                     const double *lamT, double *lam0, double *scratch, hipStream_t st,
                     int *host_fail);
// block linear solver of the chemistry subvector (A = I - gamma J, batched LU); formed:
// J already holds A (gamma unused)
int chem_ls_setup(long n, double gamma, const double *J, double *LU, unsigned long long *piv,
                  int *nsing, hipStream_t st, bool formed = false);
int chem_ls_solve(long n, const double *LU, const unsigned long long *piv, const double *b,
                  double *x, hipStream_t st);

}  // namespace mvd
