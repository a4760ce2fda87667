// IMEX additive Runge-Kutta evolution of the reacting flow on the device: the IMEX driver
// (imex_chem_hydro_main.cpp:337-476: ARKStepCreate(fexpl, fimpl) with the block-diagonal
// chemistry linear solver and PostprocessStep, :1002-1050) with a fixed step.
//
// fexpl (:910-1000) is fslow (scaling, the total-energy overwrite, fe(), the energy rate
// moved to the gas-energy slot, unscaling); fimpl (:811-848) is TimeUnits times
// calculate_rhs_cvklu on the chemistry subvector and zero on the fluid fields. ARKStep's
// stage algebra (SUNDIALS, outside the reference tree), for a pair of tables Ae (explicit)
// and Ai (diagonally implicit) with common b and c, stage i = 0..s-1:
//   R_i = y_n + sum_{j<i} (h Ae_ij Fe_j + h Ai_ij Fi_j)          (Fi_j zero on the fluid)
//   Ai_ii = 0: z_i = R_i and Fi_i = fimpl(z_i);
//   else the chemistry solves z - h Ai_ii fimpl(z) = R_i per cell (the batched Newton/LU
//        kernel with a one-stage table, predictor R_i), Fi_i = (z_i - R_i) / (h Ai_ii),
//        and the fluid fields take z_i = R_i;
//   Fe_i = fexpl(z_i) (rewriting z_i's total energy as the reference's fexpl does);
//   y_{n+1} = y_n + sum_i (h b_i Fe_i + h b_i Fi_i), then PostprocessStep (scaling,
//   et = ge / EnergyUnits + |m|^2 / (2 rho), unscaling).
// Every sum is one fused linear-combination kernel in the order written. The state and the
// stage values live in contiguous ManyVector blocks ([5 fields][chem]).
#include "mvd_imex.hpp"

namespace mvd {

namespace {

__global__ void imex_add_int(const int *src, int *dst) { *dst += *src; }
__global__ void imex_zero(unsigned *flag, int *nfail) {
  *flag = 0u;
  *nfail = 0;
}

// Fi = (z - r) / hg (ARKODE's implicit-stage derivative)
__global__ void __launch_bounds__(256)
stage_deriv(long n, const double *z, const double *r, double hg, double *fi) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) fi[i] = (z[i] - r[i]) / hg;
}

inline void sub_ptrs(double *base, long n, double *(&p)[6]) {
  for (int q = 0; q < 5; q++) p[q] = base + q * n;
  p[5] = base + 5 * n;
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void imex_free(ImexState &m) {
  if (m.Y) (void)hipFree(m.Y);
  if (m.flag) (void)hipFree(m.flag);
  m = ImexState{};
}

int imex_evolve(const Grid &g, FeScratch &fe, ChemState &ch, double eunit_inv, ImexState &m,
                const ArkTable &tab, const DirkDev &newton, double h, int nsteps,
                bool consistent, int waves, int block, double *const *w, hipStream_t st,
                unsigned *host_flag, int *host_fail) {
  const int S = tab.stages;
  if (S < 1 || S > ARK_MAXSTAGE) return -2;
  const long n = (long)g.nx * g.ny * g.nz;
  const long nc = (long)g.nchem * n;
  const long len = 5 * n + nc;
  if (!m.Y || m.S != S || m.len != len) {
    imex_free(m);
    // Yn | Z | R | Fe_0..Fe_{S-1} | Fi_0..Fi_{S-1} (chemistry only)
    if (hipMalloc(&m.Y, sizeof(double) * (len * (3 + S) + nc * S)) != hipSuccess) return -1;
    if (hipMalloc(&m.flag, 2 * sizeof(unsigned)) != hipSuccess) return -1;
    m.nfail = (int *)(m.flag + 1);
    m.S = S;
    m.len = len;
  }
  double *Yn = m.Y, *Z = Yn + len, *R = Z + len, *Fe = R + len, *Fi = Fe + S * len;
  double *Zp[6], *Ynp[6];
  sub_ptrs(Z, n, Zp);
  sub_ptrs(Yn, n, Ynp);
  // the chemistry solves: backward Euler form z - hg f(z) = R_i
  DirkDev be = newton;
  be.nstage = 1;
  for (int k = 0; k < MAXSTAGE * MAXSTAGE; k++) be.A[k] = 0.0;
  be.A[0] = 1.0;
  be.esdirk = 0;
  be.gamma = 1.0;
  for (int q = 0; q < 5; q++)
    if (hipMemcpyAsync(Yn + q * n, w[q], sizeof(double) * n, hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      return -1;
  if (hipMemcpyAsync(Yn + 5 * n, w[5], sizeof(double) * nc, hipMemcpyDeviceToDevice, st) !=
      hipSuccess)
    return -1;
  hipLaunchKernelGGL(imex_zero, dim3(1), dim3(1), 0, st, m.flag, m.nfail);
  const bool timer_on = fe.timer.on;
  fe.timer.on = false;
  int rc = 0;
  for (int step = 0; step < nsteps && rc == 0; step++) {
    // the cells' temperature order is computed for the step's first implicit stage and
    // reused by its others (an ordering only: each cell's arithmetic does not depend on it)
    bool ordered = false;
    for (int i = 0; i < S && rc == 0; i++) {
      double *Fei = Fe + i * len, *Fii = Fi + i * nc;
      // R_i: fluid, then chemistry (explicit and implicit terms interleaved)
      double cf[1 + ARK_MAXSTAGE], cc[1 + 2 * ARK_MAXSTAGE];
      const double *Xf[1 + ARK_MAXSTAGE], *Xc[1 + 2 * ARK_MAXSTAGE];
      cf[0] = cc[0] = 1.0;
      Xf[0] = Yn;
      Xc[0] = Yn + 5 * n;
      for (int j = 0; j < i; j++) {
        cf[1 + j] = h * tab.Ae[i * 8 + j];
        Xf[1 + j] = Fe + j * len;
        cc[1 + 2 * j] = h * tab.Ae[i * 8 + j];
        Xc[1 + 2 * j] = Fe + j * len + 5 * n;
        cc[2 + 2 * j] = h * tab.Ai[i * 8 + j];
        Xc[2 + 2 * j] = Fi + j * nc;
      }
      // the fluid part of R is the stage value itself: it goes straight into Z; the
      // chemistry part is kept in R (the implicit stage derivative needs it) and copied into
      // Z as the Newton predictor, or straight into Z for an explicit stage
      const double aii = tab.Ai[i * 8 + i];
      double *Rc = aii == 0.0 ? Z + 5 * n : R + 5 * n;
      if (linear_combination_launch(1 + i, cf, Xf, Z, 5 * n, st) ||
          linear_combination_launch(1 + 2 * i, cc, Xc, Rc, nc, st) ||
          (aii != 0.0 && hipMemcpyAsync(Z + 5 * n, Rc, sizeof(double) * nc,
                                        hipMemcpyDeviceToDevice, st) != hipSuccess)) {
        rc = -1;
        break;
      }
      if (aii == 0.0) {
        if (chem_rhs(ch, Z + 5 * n, Fii, st, false)) { rc = -1; break; }
        const double cs[1] = {newton.rhs_scale};
        const double *Xs[1] = {Fii};
        if (linear_combination_launch(1, cs, Xs, Fii, nc, st)) { rc = -1; break; }
      } else {
        const double hg = h * aii;
        if (dirk_batched_launch(ch, be, hg, Z + 5 * n, st, waves, block, nullptr, ordered)) {
          rc = -1;
          break;
        }
        ordered = true;
        hipLaunchKernelGGL(imex_add_int, dim3(1), dim3(1), 0, st,
                           (const int *)(ch.counters + 1 + STAT_FAILED), m.nfail);
        hipLaunchKernelGGL(stage_deriv, dim3(nblk(nc)), dim3(256), 0, st, nc,
                           (const double *)(Z + 5 * n), (const double *)(R + 5 * n), hg, Fii);
      }
      // Fe_i = fexpl(z_i)
      double *Fp[6];
      sub_ptrs(Fei, n, Fp);
      if (fslow_prepare_launch(g, Zp, ch.scale, eunit_inv, st)) { rc = -1; break; }
      Fields fz;
      for (int q = 0; q < 5; q++) fz.w[q] = Zp[q];
      fz.chem = Zp[5];
      FieldsOut fo;
      for (int q = 0; q < 5; q++) fo.w[q] = Fp[q];
      fo.chem = Fp[5];
      const int frc = fe_launch(g, fz, fo, fe, st, false, m.flag);   // bits into the run's flag
      if (frc) { rc = frc; break; }
      if (fslow_finish_launch(g, Fp, st) || chem_scale_launch(nc, Zp[5], ch.scale, 0, st)) {
        rc = -1;
        break;
      }
      if (consistent && chem_forcing_to_scaled_launch(n, Fp[5], ch.scale, 1.0 / eunit_inv, st)) {
        rc = -1;
        break;
      }
    }
    if (rc) break;
    // y_{n+1}, in place on Yn
    double cf[1 + ARK_MAXSTAGE], cc[1 + 2 * ARK_MAXSTAGE];
    const double *Xf[1 + ARK_MAXSTAGE], *Xc[1 + 2 * ARK_MAXSTAGE];
    cf[0] = cc[0] = 1.0;
    Xf[0] = Yn;
    Xc[0] = Yn + 5 * n;
    for (int j = 0; j < S; j++) {
      cf[1 + j] = h * tab.b[j];
      Xf[1 + j] = Fe + j * len;
      cc[1 + 2 * j] = h * tab.b[j];
      Xc[1 + 2 * j] = Fe + j * len + 5 * n;
      cc[2 + 2 * j] = h * tab.b[j];
      Xc[2 + 2 * j] = Fi + j * nc;
    }
    if (linear_combination_launch(1 + S, cf, Xf, Yn, 5 * n, st) ||
        linear_combination_launch(1 + 2 * S, cc, Xc, Yn + 5 * n, nc, st) ||
        fslow_prepare_launch(g, Ynp, ch.scale, eunit_inv, st) ||        // PostprocessStep
        chem_scale_launch(nc, Ynp[5], ch.scale, 0, st))
      rc = -1;
  }
  fe.timer.on = timer_on;
  if (rc) return rc;
  for (int q = 0; q < 5; q++)
    if (hipMemcpyAsync(w[q], Yn + q * n, sizeof(double) * n, hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      return -1;
  if (hipMemcpyAsync(w[5], Yn + 5 * n, sizeof(double) * nc, hipMemcpyDeviceToDevice, st) !=
      hipSuccess)
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
