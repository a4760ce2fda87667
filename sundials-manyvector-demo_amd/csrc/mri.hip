// Multirate (MRI) evolution of the reacting flow on the device: the configuration-5 driver
// (multirate_chem_hydro_main.cpp:529-555: MRIStepCreate with fslow as the slow explicit RHS,
// the default third-order coupling, a fixed slow step; the fast stepper integrates ffast,
// :898-952, whose forcing is applied to all six subvectors).
//
// MRIStep's MIS form (SUNDIALS 6.2.0 MRIStepCoupling_MIStoMRI; not in the reference tree)
// is restated with Knoth & Wolke's 3-stage slow table, b appended as the last row (c_4 = 1).
// Stage i (i = 2..4) starts from the running state Y_{i-1}: F_{i-1} = fslow(Y_{i-1}) (the
// reference's fslow, which also overwrites Y_{i-1}'s total energy), the constant forcing
// r_i = sum_{j<i} ((a_ij - a_{i-1,j}) / (c_i - c_{i-1})) F_j, then fixed fast steps over
// (c_i - c_{i-1}) h, as many as the fast step bound hmax asks (mri_fast_steps): the fluid
// fields, which have no fast dynamics, advance by dt * r_i per step; the chemistry by
// forced steps of the batched per-cell DIRK kernel. The forcing is constant over the
// stage and the cells are independent, so one launch of each runs all the stage's fast
// steps: every fluid value takes its nfast additions in turn, every chemistry cell its
// nfast steps back to back with its state in registers (the same results as one launch
// per fast step).
//
// Everything stays on the device: the state and the three slow derivatives live in
// contiguous ManyVector blocks ([5 fields][chem]); the first slow step runs eagerly (it
// allocates the DIRK scratch), then one slow step -- three fslow evaluations, the forcing
// combinations and the three stages' fast steps -- is captured as one hipGraph and
// replayed. fe()'s legal-state bits and the failed fast cell steps accumulate on the
// device and are read back once.
//
// Forcing units. The reference's fslow returns the chemistry rates in physical units (the
// advection of the unscaled species) and the total-energy rate, in code units, in the gas
// energy slot, and MRIStep adds them unchanged to the fast stepper's Dengo-scaled
// variables (multirate_chem_hydro_main.cpp:996-1083 with :922-946). That is reproduced by
// default. With `consistent` the
// chemistry part of each slow RHS is first converted to the scaled variables the fast
// stepper integrates: species rates divided by their scale, the gas-energy slot times
// EnergyUnits (fslow sets et = ge / EnergyUnits) divided by the ge scale.
#include "mvd_mri.hpp"

namespace mvd {

namespace {

// Knoth-Wolke 3 with b appended
constexpr double KW_C[4] = {0.0, 1.0 / 3.0, 3.0 / 4.0, 1.0};
constexpr double KW_A[4][4] = {{0.0, 0.0, 0.0, 0.0},
                               {1.0 / 3.0, 0.0, 0.0, 0.0},
                               {-3.0 / 16.0, 15.0 / 16.0, 0.0, 0.0},
                               {1.0 / 6.0, 3.0 / 10.0, 8.0 / 15.0, 0.0}};

__global__ void mri_add_int(const int *src, int *dst) { *dst += *src; }

__global__ void mri_zero(unsigned *flag, int *nfail) {
  *flag = 0u;
  *nfail = 0;
}

inline void sub_ptrs(double *base, long n, int nchem, double *(&p)[6]) {
  for (int q = 0; q < 5; q++) p[q] = base + q * n;
  p[5] = nchem > 0 ? base + 5 * n : nullptr;
}

}  // namespace

// fast steps of slow stage i (2..4): ceil((c_i - c_{i-1}) h / hmax), at least one
int mri_fast_steps(int i, double h, double hmax) {
  const double x = (KW_C[i] - KW_C[i - 1]) * h / hmax;
  long k = (long)x;
  if ((double)k < x) k++;
  if (k < 1) k = 1;
  if (k > 1000000) k = 1000000;
  return (int)k;
}

namespace {

// one slow step on st
int enqueue_slow_step(const Grid &g, FeScratch &fe, ChemState &ch, double eunit_inv,
                      MriState &m, const DirkDev &o, double h, double hmax, int waves,
                      int block, bool consistent, hipStream_t st) {
  const long n = (long)g.nx * g.ny * g.nz;
  const long len = (5L + g.nchem) * n;
  double *Y[6];
  sub_ptrs(m.Y, n, g.nchem, Y);
  for (int i = 1; i < 4; i++) {
    double *Fi[6];
    sub_ptrs(m.F + (long)(i - 1) * len, n, g.nchem, Fi);
    // fslow(Y) -> F_{i-1}: scaling, et overwrite, fe(), etdot -> chemdot[9], unscaling
    if (fslow_prepare_launch(g, Y, ch.scale, eunit_inv, st)) return -1;
    Fields fy;
    for (int q = 0; q < 5; q++) fy.w[q] = Y[q];
    fy.chem = Y[5];
    FieldsOut fo;
    for (int q = 0; q < 5; q++) fo.w[q] = Fi[q];
    fo.chem = Fi[5];
    const int rc = fe_launch(g, fy, fo, fe, st, false, m.flag);  // bits into the run's flag
    if (rc) return rc;
    if (fslow_finish_launch(g, Fi, st)) return -1;
    if (chem_scale_launch(n * g.nchem, Y[5], ch.scale, 0, st)) return -1;
    if (consistent && chem_forcing_to_scaled_launch(n, Fi[5], ch.scale, 1.0 / eunit_inv, st))
      return -1;
    // forcing r_i over the whole ManyVector
    const double dc = KW_C[i] - KW_C[i - 1];
    double coef[3];
    const double *X[3];
    for (int j = 0; j < i; j++) {
      coef[j] = (KW_A[i][j] - KW_A[i - 1][j]) / dc;
      X[j] = m.F + (long)j * len;
    }
    if (linear_combination_launch(i, coef, X, m.R, len, st)) return -1;
    // fast steps
    const int nfast = mri_fast_steps(i, h, hmax);
    const double dt = dc * h / nfast;
    if (axpy_steps_launch(nfast, dt, m.R, m.Y, 5 * n, st)) return -1;   // fluid
    // the cells' temperature order is taken at the stage's start and kept for all its fast
    // steps (temperatures move little over one slow stage)
    if (dirk_batched_launch(ch, o, dt, Y[5], st, waves, block, m.R + 5 * n, false, nfast))
      return -1;
    hipLaunchKernelGGL(mri_add_int, dim3(1), dim3(1), 0, st,
                       (const int *)(ch.counters + 1 + STAT_FAILED), m.nfail);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

bool same_opts(const DirkDev &a, const DirkDev &b) {
  if (a.nstage != b.nstage || a.max_iters != b.max_iters || a.max_setups != b.max_setups)
    return false;
  for (int k = 0; k < MAXSTAGE * MAXSTAGE; k++)
    if (a.A[k] != b.A[k]) return false;
  return a.rtol == b.rtol && a.atol == b.atol && a.ntol == b.ntol &&
         a.rhs_scale == b.rhs_scale && a.max_halvings == b.max_halvings && a.msbp == b.msbp;
}

}  // namespace

void mri_free(MriState &m) {
  if (m.exec) (void)hipGraphExecDestroy(m.exec);
  if (m.cap) (void)hipStreamDestroy(m.cap);
  if (m.Y) (void)hipFree(m.Y);
  if (m.flag) (void)hipFree(m.flag);
  m = MriState{};
}

int mri_evolve(const Grid &g, FeScratch &fe, ChemState &ch, double eunit_inv, MriState &m,
               const DirkDev &o, double h, int nslow, double hmax, bool consistent, int waves,
               int block, double *const *w, hipStream_t st, unsigned *host_flag,
               int *host_fail) {
  const long n = (long)g.nx * g.ny * g.nz;
  const long len = (5L + g.nchem) * n;
  if (!m.Y) {
    // Y | F_0 F_1 F_2 | R
    if (hipMalloc(&m.Y, sizeof(double) * len * 5) != hipSuccess) return -1;
    m.F = m.Y + len;
    m.R = m.F + 3 * len;
    if (hipMalloc(&m.flag, 2 * sizeof(unsigned)) != hipSuccess) return -1;
    m.nfail = (int *)(m.flag + 1);
    if (hipStreamCreateWithFlags(&m.cap, hipStreamNonBlocking) != hipSuccess) return -1;
  }
  for (int q = 0; q < 5; q++)
    if (hipMemcpyAsync(m.Y + q * n, w[q], sizeof(double) * n, hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      return -1;
  if (hipMemcpyAsync(m.Y + 5 * n, w[5], sizeof(double) * g.nchem * n, hipMemcpyDeviceToDevice,
                     st) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(mri_zero, dim3(1), dim3(1), 0, st, m.flag, m.nfail);
  const bool timer_on = fe.timer.on;
  fe.timer.on = false;                    // no event records inside the graph
  // the captured step holds the addresses of the chemistry arrays and work buffer and the
  // chemistry constants (redshift) by value: a chemistry setup or a reallocation of the
  // work buffer by another solver call since the capture makes it stale
  bool rebuild = !m.exec || m.h != h || m.hmax != hmax || m.consistent != consistent ||
                 !same_opts(m.o, o) || m.waves != waves || m.block != block ||
                 m.chem_gen != ch.gen;
  int rc = 0;
  for (int step = 0; step < nslow && rc == 0; step++) {
    if (!rebuild) {
      if (hipGraphLaunch(m.exec, st) != hipSuccess) rc = -1;
      continue;
    }
    // eager step (allocates what the fast kernels need), then capture the step for the
    // remaining ones
    rc = enqueue_slow_step(g, fe, ch, eunit_inv, m, o, h, hmax, waves, block, consistent, st);
    if (rc || step + 1 == nslow) break;
    if (m.exec) { (void)hipGraphExecDestroy(m.exec); m.exec = nullptr; }
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(m.cap, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      rc = -1;
      break;
    }
    const int rcc =
        enqueue_slow_step(g, fe, ch, eunit_inv, m, o, h, hmax, waves, block, consistent, m.cap);
    const hipError_t ec = hipStreamEndCapture(m.cap, &graph);
    if (rcc || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      rc = rcc ? rcc : -1;
      break;
    }
    const hipError_t ei = hipGraphInstantiate(&m.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) { m.exec = nullptr; rc = -1; break; }
    m.h = h;
    m.hmax = hmax;
    m.consistent = consistent;
    m.o = o;
    m.waves = waves;
    m.block = block;
    m.chem_gen = ch.gen;
    rebuild = false;
  }
  fe.timer.on = timer_on;
  if (rc) return rc;
  for (int q = 0; q < 5; q++)
    if (hipMemcpyAsync(w[q], m.Y + q * n, sizeof(double) * n, hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      return -1;
  if (hipMemcpyAsync(w[5], m.Y + 5 * n, sizeof(double) * g.nchem * n, hipMemcpyDeviceToDevice,
                     st) != hipSuccess)
    return -1;
  unsigned hv[2] = {0u, 0u};
  if (hipMemcpyAsync(hv, m.flag, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return -1;
  *host_flag = hv[0];
  *host_fail = (int)hv[1];
  return 0;
}

}  // namespace mvd
