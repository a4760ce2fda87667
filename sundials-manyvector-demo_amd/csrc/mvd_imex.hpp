// IMEX additive Runge-Kutta evolution of the reacting flow (imex.hip).
#pragma once
#include "mvd_internal.hpp"
#include "mvd_dirk.hpp"

namespace mvd {

constexpr int ARK_MAXSTAGE = 8;

// explicit / diagonally implicit pair, 8x8 row-major, common b and c (mvd_ark_table)
struct ArkTable {
  int stages;
  double Ae[64];
  double Ai[64];
  double b[8];
  double c[8];
};

struct ImexState {
  double *Y = nullptr;        // one allocation: Yn | Z | R | Fe_0.. | Fi_0.. (chemistry)
  unsigned *flag = nullptr;   // legal-state bits, then failed cell solves
  int *nfail = nullptr;
  int S = 0;
  long len = 0;
};

// nsteps fixed steps of size h in place on w (6 device subvectors, chemistry scaled);
// host_flag receives fexpl's legal-state bits, host_fail the failed cell solves. newton:
// the Newton/LU controls of the implicit solves (its table is not used). Returns 0, -2
// (unsupported) or -1.
int imex_evolve(const Grid &g, FeScratch &fe, ChemState &ch, double eunit_inv, ImexState &m,
                const ArkTable &tab, const DirkDev &newton, double h, int nsteps,
                bool consistent, int waves, int block, double *const *w, hipStream_t st,
                unsigned *host_flag, int *host_fail);
void imex_free(ImexState &m);

}  // namespace mvd
