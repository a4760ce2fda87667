// Oracle drivers for the primordial chemistry and the Dengo scaling helpers --
// TEST INFRASTRUCTURE ONLY (see orc_common.hpp).
#include <cstring>
#include <vector>
#include "orc_chem.hpp"
#include "oracle.h"

using namespace orc;

using Chem = orc::ChemState;

// cvklu_setup_data (raja_primordial_network.cpp:45-254) with the tables given as the
// committed blob: Ts = 1000 K, bounds [1,1e5], nbins = 1023,
// dbin = (log(1e5)-log(1))/nbins, idbin = 1.0L/dbin (long-double quotient as in line 195),
// scale = inv_scale = 1 (primordial_blast.cpp:337-341).
extern "C" void *orc_chem_create(long n, const double *tables, double current_z) {
  Chem *c = new Chem;
  c->n = n;
  c->tab.assign(tables, tables + NTAB * TLEN);
  c->g.tab = c->tab.data();
  c->g.tmin = 1.0;
  c->g.tmax = 100000.0;
  c->g.nbins = TLEN - 1;
  c->g.dbin = (std::log(c->g.tmax) - std::log(c->g.tmin)) / c->g.nbins;
  c->g.idbin = (double)(1.0L / (long double)c->g.dbin);
  c->g.lb = std::log(c->g.tmin);
  c->z = current_z;
  c->scale.assign(NSP * n, 1.0);
  c->inv_scale.assign(NSP * n, 1.0);
  c->Ts.assign(n, 1000.0);
  c->dTs_ge.assign(n, 0.0);
  c->mdensity.assign(n, 0.0);
  c->inv_mdensity.assign(n, 0.0);
  c->cie_oda.assign(n, 0.0);
  c->h2_oda.assign(n, 0.0);
  c->rs.assign((size_t)NRATE * n, 0.0);
  c->drs.assign((size_t)NRATE * n, 0.0);
  return c;
}

extern "C" void orc_chem_destroy(void *h) { delete (Chem *)h; }

// 0 scale, 1 inv_scale, 2 Ts, 3 dTs_ge, 4 mdensity, 5 inv_mdensity, 6 cie_oda, 7 h2_oda,
// 8 rate values [NRATE][n], 9 rate derivatives [NRATE][n]
extern "C" double *orc_chem_array(void *h, int which) {
  Chem *c = (Chem *)h;
  switch (which) {
    case 0: return c->scale.data();
    case 1: return c->inv_scale.data();
    case 2: return c->Ts.data();
    case 3: return c->dTs_ge.data();
    case 4: return c->mdensity.data();
    case 5: return c->inv_mdensity.data();
    case 6: return c->cie_oda.data();
    case 7: return c->h2_oda.data();
    case 8: return c->rs.data();
    case 9: return c->drs.data();
  }
  return nullptr;
}

extern "C" double orc_chem_const(void *h, int which) {
  Chem *c = (Chem *)h;
  return which == 0 ? c->g.dbin : which == 1 ? c->g.idbin : c->g.lb;
}

// setting_up_extra_variables (raja_primordial_network.cpp:1969-2018), from scale.
extern "C" void orc_chem_extra(void *h) {
  Chem *c = (Chem *)h;
  for (long i = 0; i < c->n; i++) {
    const double *in = &c->scale[NSP * i];
    double md = 0.0;
    md += in[0] * 2.0;
    md += in[1] * 2.0;
    md += in[2] * 1.00794;
    md += in[3] * 1.00794;
    md += in[4] * 1.00794;
    md += in[5] * 4.002602;
    md += in[6] * 4.002602;
    md += in[7] * 4.002602;
    md *= 1.67e-24;
    double tau = std::pow(md / 3.3e-8, 2.8);
    tau = std::fmax(tau, 1.0e-5);
    c->mdensity[i] = md;
    c->inv_mdensity[i] = 1.0 / md;
    c->cie_oda[i] = std::fmin(1.0, (1.0 - std::exp(-tau)) / tau);
    c->h2_oda[i] = std::fmin(1.0, std::pow(md / (1.34e-14), -0.45));
  }
}

// prepare_Dengo_structures (primordial_blast.cpp:368-400): scale <- chem, inv_scale <- 1/chem,
// chem <- 1, then the auxiliary per-cell values.
extern "C" void orc_chem_prepare(void *h, double *chem) {
  Chem *c = (Chem *)h;
  for (long i = 0; i < NSP * c->n; i++) {
    c->scale[i] = chem[i];
    c->inv_scale[i] = 1.0 / c->scale[i];
    chem[i] = 1.0;
  }
  orc_chem_extra(h);
}

// apply/unapply_Dengo_scaling (primordial_blast.cpp:404-458)
extern "C" void orc_chem_apply_scaling(void *h, double *chem) {
  Chem *c = (Chem *)h;
  for (long i = 0; i < NSP * c->n; i++) chem[i] *= c->scale[i];
}
extern "C" void orc_chem_unapply_scaling(void *h, double *chem) {
  Chem *c = (Chem *)h;
  for (long i = 0; i < NSP * c->n; i++) chem[i] /= c->scale[i];
}

static CellAux aux_of(const Chem *c, long i) {
  return CellAux{c->mdensity[i], c->inv_mdensity[i], c->cie_oda[i], c->h2_oda[i]};
}

// calculate_rhs_cvklu (raja_primordial_network.cpp:1015-1389): updates Ts, dTs_ge and
// the rate cache as side effects, then writes the scaled ydot.
extern "C" int orc_chem_rhs(void *h, const double *y, double *ydot) {
  Chem *c = (Chem *)h;
  for (long i = 0; i < c->n; i++) {
    const long j = NSP * i;
    double ya[NSP];
    for (int s = 0; s < NSP; s++) ya[s] = y[j + s] * c->scale[j + s];
    temperature(c->g, ya, c->Ts[i], c->dTs_ge[i]);
    double rs[NRATE], drs[NRATE];
    interp_rates(c->g, c->Ts[i], rs, drs);
    for (int r = 0; r < NRATE; r++) {
      c->rs[(size_t)r * c->n + i] = rs[r];
      c->drs[(size_t)r * c->n + i] = drs[r];
    }
    double f[NSP];
    rhs_physical(ya, rs, c->Ts[i], c->z, aux_of(c, i), f);
    for (int s = 0; s < NSP; s++) {
      ydot[j + s] = f[s];
      ydot[j + s] *= c->inv_scale[j + s];
    }
    ydot[j + 9] *= c->inv_mdensity[i];
  }
  return 0;
}

// calculate_jacobian_cvklu (raja_primordial_network.cpp:1393-1965), dense column-major
// 10x10 blocks J[100*i + 10*col + row] (DENSEIDX, line 40), from the cached rates.
extern "C" int orc_chem_jac(void *h, const double *y, double *Jout) {
  Chem *c = (Chem *)h;
  for (long i = 0; i < c->n; i++) {
    const long j = NSP * i;
    double ya[NSP], rs[NRATE], drs[NRATE];
    for (int s = 0; s < NSP; s++) ya[s] = y[j + s] * c->scale[j + s];
    for (int r = 0; r < NRATE; r++) {
      rs[r] = c->rs[(size_t)r * c->n + i];
      drs[r] = c->drs[(size_t)r * c->n + i];
    }
    double J[NSP][NSP];
    bool nzm[NSP][NSP];
    jac_physical(ya, rs, drs, c->Ts[i], c->z, c->dTs_ge[i], aux_of(c, i), J, nzm);
    double *blk = Jout + (size_t)NSP * NSP * i;
    for (int col = 0; col < NSP; col++)
      for (int row = 0; row < NSP; row++) {
        double v = 0.0;
        if (nzm[row][col]) {
          v = J[row][col];
          v *= (c->inv_scale[j + row] * c->scale[j + col]);
        }
        blk[NSP * col + row] = v;
      }
  }
  return 0;
}

// Rate interpolation and RHS at a fixed temperature (no temperature solve, no cache
// update): used by tests to difference the RHS with the rates held fixed.
extern "C" void orc_chem_rhs_fixed_T(void *h, const double *Tcell, const double *y,
                                     double *ydot) {
  Chem *c = (Chem *)h;
  for (long i = 0; i < c->n; i++) {
    const long j = NSP * i;
    double ya[NSP], rs[NRATE], drs[NRATE], f[NSP];
    for (int s = 0; s < NSP; s++) ya[s] = y[j + s] * c->scale[j + s];
    interp_rates(c->g, Tcell[i], rs, drs);
    rhs_physical(ya, rs, Tcell[i], c->z, aux_of(c, i), f);
    for (int s = 0; s < NSP; s++) {
      ydot[j + s] = f[s];
      ydot[j + s] *= c->inv_scale[j + s];
    }
    ydot[j + 9] *= c->inv_mdensity[i];
  }
}

// operation counts of one cell's RHS (temperature + rates + RHS) and Jacobian:
// out = {add, mul, div, sqrt, cmp, transcendental} for the RHS, then the same for J.
extern "C" void orc_chem_opcount(void *h, unsigned long long *out) {
  Chem *c = (Chem *)h;
  Cnt y[NSP], Ts(c->Ts[0]), dTs(0.0), rs[NRATE], drs[NRATE], f[NSP], J[NSP][NSP];
  bool nzm[NSP][NSP];
  for (int s = 0; s < NSP; s++) y[s] = Cnt(c->scale[s]);
  OpCounter::reset();
  temperature(c->g, y, Ts, dTs);
  interp_rates(c->g, Ts, rs, drs);
  rhs_physical(y, rs, Ts, Cnt(c->z), aux_of(c, 0), f);
  out[0] = OpCounter::add; out[1] = OpCounter::mul; out[2] = OpCounter::div;
  out[3] = OpCounter::sqrt_; out[4] = OpCounter::cmp; out[5] = OpCounter::trans;
  OpCounter::reset();
  jac_physical(y, rs, drs, Ts, Cnt(c->z), dTs, aux_of(c, 0), J, nzm);
  out[6] = OpCounter::add; out[7] = OpCounter::mul; out[8] = OpCounter::div;
  out[9] = OpCounter::sqrt_; out[10] = OpCounter::cmp; out[11] = OpCounter::trans;
}

// The early exit of the device temperature solve, restated: with x_0 = Ts0,
// x_{k+1} = S(x_k) and d_k = dge_dT(x_k), the ten steps give (x_10, d_9). S and d depend
// only on their input, so at the first repeat x_m == x_{m-p} (p <= max_period) the
// sequence is periodic from index m-p and x_10 = x_{m-p + (10-m) mod p},
// d_9 = d_{m-p + (9-m) mod p}. Runs both forms from the warm starts Ts0 for every cell of
// y (scaled) and returns the number of cells whose (T, dT/dge) differ bitwise; steps[i]
// gets the steps the early form took and period[i] the period it stopped on (0: none),
// either may be null.
extern "C" long orc_temperature_exit_check(void *h, const double *y, const double *Ts0,
                                           int max_period, int *steps, int *period) {
  Chem *c = (Chem *)h;
  long bad = 0;
  for (long i = 0; i < c->n; i++) {
    double ya[NSP];
    for (int s = 0; s < NSP; s++) ya[s] = y[NSP * i + s] * c->scale[NSP * i + s];
    double T10 = Ts0[i], d10 = 0.0;
    temperature(c->g, ya, T10, d10);
    const double density = temperature_density(ya);
    double x[11], d[10];
    x[0] = Ts0[i];
    int m = 10, per = 0;
    for (int k = 0; k < 10; k++) {
      x[k + 1] = temperature_step(c->g, ya, density, x[k], d[k]);
      for (int q = 1; q <= max_period && q <= k + 1; q++)
        if (std::memcmp(&x[k + 1], &x[k + 1 - q], sizeof(double)) == 0) {
          per = q;
          break;
        }
      if (per) {
        m = k + 1;
        break;
      }
    }
    double Ts, dge_dT;
    if (per) {
      Ts = x[m - per + (10 - m) % per];
      dge_dT = d[m - per + ((9 - m) % per + per) % per];
    } else {
      Ts = x[10];
      dge_dT = d[9];
    }
    if (Ts < c->g.tmin) Ts = c->g.tmin;
    else if (Ts > c->g.tmax) Ts = c->g.tmax;
    const double dTs = 1.0 / dge_dT;
    if (std::memcmp(&Ts, &T10, sizeof(double)) != 0 || std::memcmp(&dTs, &d10, sizeof(double)) != 0)
      bad++;
    if (steps) steps[i] = m;
    if (period) period[i] = per;
  }
  return bad;
}

// ---- vector-Jacobian product of the RHS (forward mode, the full Jacobian per cell) ----
namespace {

constexpr int ND = NSP + 1;   // tangents: the 10 physical inputs, then T

struct DualN {
  double v;
  double d[ND];
  DualN() : v(0) { for (double &x : d) x = 0; }
  DualN(double x) : v(x) { for (double &q : d) q = 0; }
};
inline DualN operator+(const DualN &a, const DualN &b) {
  DualN r(a.v + b.v);
  for (int i = 0; i < ND; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
inline DualN operator-(const DualN &a, const DualN &b) {
  DualN r(a.v - b.v);
  for (int i = 0; i < ND; i++) r.d[i] = a.d[i] - b.d[i];
  return r;
}
inline DualN operator-(const DualN &a) {
  DualN r(-a.v);
  for (int i = 0; i < ND; i++) r.d[i] = -a.d[i];
  return r;
}
inline DualN operator*(const DualN &a, const DualN &b) {
  DualN r(a.v * b.v);
  for (int i = 0; i < ND; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
inline DualN operator/(const DualN &a, const DualN &b) {
  DualN r(a.v / b.v);
  for (int i = 0; i < ND; i++) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
  return r;
}
inline DualN pow(const DualN &a, double e) {
  DualN r(std::pow(a.v, e));
  const double s = e * std::pow(a.v, e - 1.0);
  for (int i = 0; i < ND; i++) r.d[i] = s * a.d[i];
  return r;
}
inline DualN log(const DualN &a) {
  DualN r(std::log(a.v));
  for (int i = 0; i < ND; i++) r.d[i] = a.d[i] / a.v;
  return r;
}

}  // namespace

// ybar = (d ydot/d y)^T v for calculate_rhs_cvklu in the scaled variables (the function
// mvd_chem_rhs evaluates), per cell: the physical RHS in forward mode with tangents along
// the ten inputs and T, the temperature equation g(T, y) = T kb nsum(T, y)/(density mh) - ge
// likewise (its T-slope from the interpolated gamma tables), dT/dy = -(dg/dy)/(dg/dT).
// fixed_T != 0 drops the temperature path (the derivative at fixed rates). Temperatures are
// solved from the warm start and stored, as orc_chem_rhs does.
// Exact physical Jacobian J[s][q] = d f_s / d y_q of one cell at the physical state ya,
// with T solved from the warm start T (updated, with dT): the RHS in forward mode with
// tangents along the ten inputs and T, and dT/dy from the temperature equation unless
// fixed_T or the temperature is clamped.
static void exact_jac_cell(const Chem *c, long i, const double (&ya)[NSP], double &T, double &dT,
                           int fixed_T, double (&J)[NSP][NSP]) {
  const TableGrid &g = c->g;
  temperature(g, ya, T, dT);
  double rs[NRATE], drs[NRATE];
  interp_rates(g, T, rs, drs);
  DualN yd[NSP], kd[NRATE], f[NSP];
  for (int s = 0; s < NSP; s++) { yd[s] = DualN(ya[s]); yd[s].d[s] = 1.0; }
  for (int r = 0; r < NRATE; r++) { kd[r] = DualN(rs[r]); kd[r].d[NSP] = drs[r]; }
  DualN Td(T);
  Td.d[NSP] = 1.0;
  rhs_physical(yd, kd, Td, DualN(c->z), aux_of(c, i), f);
  double dTdy[NSP] = {};
  const bool clamped = !(T > g.tmin && T < g.tmax);
  if (!fixed_T && !clamped) {
    // g(T, y) with the interpolated gamma_H2 tables
    const double kb = 1.3806504e-16, mh = 1.67e-24, gm1 = 1.0 / (5.0 / 3.0 - 1.0);
    const int b = table_bin(g, std::log(T));
    const double t1 = g.lb + double(b) * g.dbin, t2 = g.lb + double(b + 1) * g.dbin;
    const DualN Tdef = (log(Td) - DualN(t1)) / DualN(t2 - t1);
    auto lin = [&](int tab) {
      const double *a = g.tab + tab * TLEN;
      return DualN(a[b]) + Tdef * DualN(a[b + 1] - a[b]);
    };
    const DualN g1 = DualN(1.0) / (lin(G_GH2_1) - DualN(1.0));
    const DualN g2 = DualN(1.0) / (lin(G_GH2_2) - DualN(1.0));
    const DualN nsum = yd[0] * g1 + yd[1] * g2 +
                       DualN(gm1) * (yd[2] + yd[3] + yd[4] + yd[5] + yd[6] + yd[7] + yd[8]);
    const DualN density = DualN(2.0) * (yd[0] + yd[1]) +
                          DualN(1.0079400000000001) * (yd[2] + yd[3] + yd[4]) +
                          DualN(4.0026020000000004) * (yd[5] + yd[6] + yd[7]);
    const DualN gg = Td * DualN(kb) * nsum / (density * DualN(mh)) - yd[9];
    for (int q = 0; q < NSP; q++) dTdy[q] = -gg.d[q] / gg.d[NSP];
  }
  for (int s = 0; s < NSP; s++)
    for (int q = 0; q < NSP; q++) J[s][q] = f[s].d[q] + f[s].d[NSP] * dTdy[q];
}

extern "C" int orc_chem_vjp(void *h, const double *y, const double *v, double *ybar,
                            int fixed_T) {
  Chem *c = (Chem *)h;
  for (long i = 0; i < c->n; i++) {
    const long j = NSP * i;
    double ya[NSP];
    for (int s = 0; s < NSP; s++) ya[s] = y[j + s] * c->scale[j + s];
    double T = c->Ts[i], dT = 0.0, J[NSP][NSP];
    exact_jac_cell(c, i, ya, T, dT, fixed_T, J);
    c->Ts[i] = T;
    c->dTs_ge[i] = dT;
    for (int q = 0; q < NSP; q++) {
      double acc = 0.0;
      for (int s = 0; s < NSP; s++) {
        double w = v[j + s] * c->inv_scale[j + s];
        if (s == NSP - 1) w *= c->inv_mdensity[i];
        acc += w * J[s][q];
      }
      ybar[j + q] = acc * c->scale[j + q];
    }
  }
  return 0;
}

// The scaled Jacobian dF/dy of one cell (F = calculate_rhs_cvklu in scaled variables, the
// mvd_chem_rhs output), exact, at the scaled state y; the temperature is solved from the
// warm start *T (updated). J row-major [s][q].
extern "C" int orc_chem_exact_jac_cell(void *h, long i, const double *y, double *T,
                                       double *Jout) {
  Chem *c = (Chem *)h;
  const long j = NSP * i;
  double ya[NSP], J[NSP][NSP], dT = 0.0;
  for (int s = 0; s < NSP; s++) ya[s] = y[s] * c->scale[j + s];
  exact_jac_cell(c, i, ya, *T, dT, 0, J);
  for (int s = 0; s < NSP; s++) {
    double rs = c->inv_scale[j + s];
    if (s == NSP - 1) rs *= c->inv_mdensity[i];
    for (int q = 0; q < NSP; q++) Jout[s * NSP + q] = J[s][q] * rs * c->scale[j + q];
  }
  return 0;
}
