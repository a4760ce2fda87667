// Multirate (MRI) evolution of the reacting flow (mri.hip).
#pragma once
#include "mvd_internal.hpp"
#include "mvd_dirk.hpp"

namespace mvd {

struct MriState {
  double *Y = nullptr, *F = nullptr, *R = nullptr;   // one allocation: Y | F_0..F_2 | R
  unsigned *flag = nullptr;                          // legal-state bits, then failed cells
  int *nfail = nullptr;
  hipStream_t cap = nullptr;                         // capture stream
  hipGraphExec_t exec = nullptr;                     // one captured slow step
  double h = 0.0, hmax = 0.0;
  int waves = 0, block = 0;
  bool consistent = false;
  DirkDev o{};
  unsigned long chem_gen = 0;                        // ChemState::gen when captured
};

// nslow slow steps of size h (fast steps no longer than hmax) in place on w (6 device
// subvectors, chemistry scaled); host_flag receives fe()'s legal-state bits of every slow
// RHS, host_fail the failed fast cell steps. Returns 0, -2 (unsupported) or -1.
// consistent: chemistry forcing converted to the fast stepper's scaled variables (else the
// reference's units are kept)
int mri_evolve(const Grid &g, FeScratch &fe, ChemState &ch, double eunit_inv, MriState &m,
               const DirkDev &o, double h, int nslow, double hmax, bool consistent, int waves,
               int block, double *const *w, hipStream_t st, unsigned *host_flag,
               int *host_fail);
// fast steps of slow stage i (2..4): ceil((c_i - c_{i-1}) h / hmax), at least one
int mri_fast_steps(int i, double h, double hmax);
void mri_free(MriState &m);

}  // namespace mvd
