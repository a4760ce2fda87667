// Oracle: primordial chemistry network (temperature, rates, RHS, Jacobian, auxiliary
// per-cell values) -- TEST INFRASTRUCTURE ONLY (see orc_common.hpp).
//
// Every expression keeps the term order of the reference network
// (src/raja_primordial_network.cpp) so that the oracle reproduces its arithmetic.
// pow(x,2) is written x*x: GCC folds pow(x,2.0) to x*x in the reference build.
#pragma once
#include <vector>
#include "orc_common.hpp"

namespace orc {

constexpr int NSP = 10;         // species per cell, raja_primordial_network.hpp:42
constexpr int TLEN = 1024;      // table length, raja_primordial_network.cpp:43
constexpr int NRATE = 46;       // 21 reaction + 25 cooling rate tables
constexpr int NTAB = 50;        // + 4 gamma tables

// table order of the committed blob (reference read order)
enum {
  K01 = 0, K02, K03, K04, K05, K06, K07, K08, K09, K10, K11, K12, K13, K14, K15, K16, K17,
  K18, K19, K21, K22,
  C_BREM, C_CEHEI, C_CEHEII, C_CEHI, C_CIECO, C_CIHEI, C_CIHEII, C_CIHEIS, C_CIHI, C_COMPTON,
  C_GAEL, C_GAH2, C_GAHE, C_GAHI, C_GAHP, C_H2LTE, C_H2MCOOL, C_H2MHEAT, C_NCRD1, C_NCRD2,
  C_NCRN, C_REHEII1, C_REHEII2, C_REHEIII, C_REHII,
  G_GH2_1, G_DGH2_1, G_GH2_2, G_DGH2_2
};

struct TableGrid {
  const double *tab;   // NTAB * TLEN
  double dbin, idbin, lb;
  int nbins;
  double tmin, tmax;
};

// bin index as computed by the reference: (int)(idbin*(log(T)-lb)) clamped to [0,nbins-1]
// (raja_primordial_network.cpp:953-958, 1050-1055); NaN arguments land in bin 0 as the
// x86 conversion returns INT_MIN for them.
template <class R>
inline int table_bin(const TableGrid &g, R logT) {
  const double x = value(R(g.idbin) * (logT - R(g.lb)));
  if (!(x == x)) return 0;
  int b = (x >= 2147483647.0) ? 2147483647 : (x <= -2147483648.0) ? (-2147483647 - 1) : (int)x;
  if (b <= 0) b = 0;
  else if (b >= g.nbins) b = g.nbins - 1;
  return b;
}

// cvklu_calculate_temperature (raja_primordial_network.cpp:908-1011): exactly 10 Newton
// iterations warm-started from Ts, then clamped to [1,1e5]; dTs_ge = 1/dge_dT.
// One Newton step of the temperature solve from the input T: returns the next iterate and
// dge_dT at T (raja_primordial_network.cpp:943-990).
template <class R>
R temperature_step(const TableGrid &g, const R (&y)[NSP], R density, R T, R &dge_dT) {
  const R kb = R(1.3806504e-16), mh = R(1.67e-24);
  const R gamma = R(5.e0 / 3.e0);
  const R gm1 = R(1.0) / (gamma - R(1));
  const R H2_1 = y[0], H2_2 = y[1], H_1 = y[2], H_2 = y[3], H_m0 = y[4], He_1 = y[5],
          He_2 = y[6], He_3 = y[7], de = y[8], ge = y[9];
  const double *t = g.tab;
  const int b = table_bin(g, log(T));
  const R t1 = R(g.lb) + R(double(b)) * R(g.dbin);
  const R t2 = R(g.lb) + R(double(b + 1)) * R(g.dbin);
  const R Tdef = (log(T) - t1) / (t2 - t1);
  auto lin = [&](int tab) {
    const double *a = t + tab * TLEN;
    return R(a[b]) + Tdef * (R(a[b + 1]) - R(a[b]));
  };
  const R gH2_2 = lin(G_GH2_2);
  const R dgH2_2 = lin(G_DGH2_2);
  const R gH2_1 = lin(G_GH2_1);
  const R dgH2_1 = lin(G_DGH2_1);
  const R g1 = R(1.0) / (gH2_1 - R(1.0));
  const R g2 = R(1.0) / (gH2_2 - R(1.0));
  const R nsum = H2_1 * g1 + H2_2 * g2 + H_1 * gm1 + H_2 * gm1 + H_m0 * gm1 + He_1 * gm1
               + He_2 * gm1 + He_3 * gm1 + gm1 * de;
  dge_dT = T * kb * ((-H2_1) * g1 * g1 * dgH2_1 - H2_2 * g2 * g2 * dgH2_2) / (density * mh)
         + kb * nsum / (density * mh);
  const R dge = T * kb * nsum / (density * mh) - ge;
  return T - dge / dge_dT;
}

template <class R>
R temperature_density(const R (&y)[NSP]) {
  return R(2.0) * y[0] + R(2.0) * y[1] + R(1.0079400000000001) * y[2]
       + R(1.0079400000000001) * y[3] + R(1.0079400000000001) * y[4]
       + R(4.0026020000000004) * y[5] + R(4.0026020000000004) * y[6]
       + R(4.0026020000000004) * y[7];
}

// cvklu_calculate_temperature (raja_primordial_network.cpp:908-1011): ten Newton steps from
// the warm start Ts, clamp to [tmin, tmax], dTs_ge = 1/dge_dT of the last step.
template <class R>
void temperature(const TableGrid &g, const R (&y)[NSP], R &Ts, R &dTs_ge) {
  const R density = temperature_density(y);
  R Tnew = Ts * R(1.1);
  R dge_dT = R(0);
  for (int it = 0; it < 10; it++) {
    const R T = Ts;
    Tnew = temperature_step(g, y, density, T, dge_dT);
    Ts = Tnew;
  }
  Ts = Tnew;
  if (Ts < R(g.tmin)) Ts = R(g.tmin);
  else if (Ts > R(g.tmax)) Ts = R(g.tmax);
  dTs_ge = R(1.0) / dge_dT;
}

// Rate interpolation of calculate_rhs_cvklu (raja_primordial_network.cpp:1046-1245):
// value and d/dT of each of the 46 rate tables at temperature T.
template <class R>
void interp_rates(const TableGrid &g, R T, R (&rs)[NRATE], R (&drs)[NRATE]) {
  const int b = table_bin(g, log(T));
  const R t1 = R(g.lb) + R(double(b)) * R(g.dbin);
  const R t2 = R(g.lb) + R(double(b + 1)) * R(g.dbin);
  const R Tdef = (log(T) - t1) / (t2 - t1);
  const R dT = t2 - t1;
  const R invTs = R(1.0) / T;
  const R Tfactor = invTs / dT;
  for (int r = 0; r < NRATE; r++) {
    const double *a = g.tab + r * TLEN;
    rs[r] = R(a[b]) + Tdef * (R(a[b + 1]) - R(a[b]));
    drs[r] = (R(a[b + 1]) - R(a[b])) * Tfactor;
  }
}

struct CellAux {      // per-cell constants from setting_up_extra_variables
  double mdensity, inv_mdensity, cie_oda, h2_oda;
};

// State of one cvklu_data for n cells (raja_primordial_network.hpp:104-276): tables,
// scaling, temperature warm start and the per-cell rate cache.
struct ChemState {
  long n;
  std::vector<double> tab;
  TableGrid g;
  double z;
  std::vector<double> scale, inv_scale, Ts, dTs_ge, mdensity, inv_mdensity, cie_oda, h2_oda,
      rs, drs;
};

// Species and energy RHS in physical units (before inv_scale), raja_primordial_network.cpp:1302-1371.
template <class R>
void rhs_physical(const R (&y)[NSP], const R (&k)[NRATE], R T, R z, const CellAux &ax,
                  R (&f)[NSP]) {
  const R H2_1 = y[0], H2_2 = y[1], H_1 = y[2], H_2 = y[3], H_m0 = y[4], He_1 = y[5],
          He_2 = y[6], He_3 = y[7], de = y[8];
  const R k01 = k[K01], k02 = k[K02], k03 = k[K03], k04 = k[K04], k05 = k[K05], k06 = k[K06],
          k07 = k[K07], k08 = k[K08], k09 = k[K09], k10 = k[K10], k11 = k[K11], k12 = k[K12],
          k13 = k[K13], k14 = k[K14], k15 = k[K15], k16 = k[K16], k17 = k[K17], k18 = k[K18],
          k19 = k[K19], k21 = k[K21], k22 = k[K22];
  const R mdensity = R(ax.mdensity), cie = R(ax.cie_oda), h2o = R(ax.h2_oda);
  f[0] = k08 * H_1 * H_m0 + k10 * H2_2 * H_1 - k11 * H2_1 * H_2 - k12 * H2_1 * de
       - k13 * H2_1 * H_1 + k19 * H2_2 * H_m0 + k21 * H2_1 * H_1 * H_1 + k22 * H_1 * H_1 * H_1;
  f[1] = k09 * H_1 * H_2 - k10 * H2_2 * H_1 + k11 * H2_1 * H_2 + k17 * H_2 * H_m0
       - k18 * H2_2 * de - k19 * H2_2 * H_m0;
  f[2] = (-k01) * H_1 * de + k02 * H_2 * de - k07 * H_1 * de - k08 * H_1 * H_m0
       - k09 * H_1 * H_2 - k10 * H2_2 * H_1 + k11 * H2_1 * H_2 + R(2) * k12 * H2_1 * de
       + R(2) * k13 * H2_1 * H_1 + k14 * H_m0 * de + k15 * H_1 * H_m0 + R(2) * k16 * H_2 * H_m0
       + R(2) * k18 * H2_2 * de + k19 * H2_2 * H_m0 - R(2) * k21 * H2_1 * H_1 * H_1
       - R(2) * k22 * H_1 * H_1 * H_1;
  f[3] = k01 * H_1 * de - k02 * H_2 * de - k09 * H_1 * H_2 + k10 * H2_2 * H_1
       - k11 * H2_1 * H_2 - k16 * H_2 * H_m0 - k17 * H_2 * H_m0;
  f[4] = k07 * H_1 * de - k08 * H_1 * H_m0 - k14 * H_m0 * de - k15 * H_1 * H_m0
       - k16 * H_2 * H_m0 - k17 * H_2 * H_m0 - k19 * H2_2 * H_m0;
  f[5] = (-k03) * He_1 * de + k04 * He_2 * de;
  f[6] = k03 * He_1 * de - k04 * He_2 * de - k05 * He_2 * de + k06 * He_3 * de;
  f[7] = k05 * He_2 * de - k06 * He_3 * de;
  f[8] = k01 * H_1 * de - k02 * H_2 * de + k03 * He_1 * de - k04 * He_2 * de + k05 * He_2 * de
       - k06 * He_3 * de - k07 * H_1 * de + k08 * H_1 * H_m0 + k14 * H_m0 * de
       + k15 * H_1 * H_m0 + k17 * H_2 * H_m0 - k18 * H2_2 * de;
  const R gS = H2_1 * k[C_GAH2] + H_1 * k[C_GAHI] + H_2 * k[C_GAHP] + He_1 * k[C_GAHE]
             + de * k[C_GAEL];
  const R ncrD = H2_1 * k[C_NCRD2] + H_1 * k[C_NCRD1];
  f[9] = R(-2.0158800000000001) * H2_1 * k[C_CIECO] * cie * mdensity
       - H2_1 * cie * k[C_H2LTE] * h2o / (k[C_H2LTE] / gS + R(1.0))
       - H_1 * k[C_CEHI] * cie * de
       - H_1 * k[C_CIHI] * cie * de
       - H_2 * cie * de * k[C_REHII]
       - He_1 * k[C_CIHEI] * cie * de
       - He_2 * k[C_CEHEII] * cie * de
       - He_2 * k[C_CEHEI] * cie * (de * de)
       - He_2 * k[C_CIHEII] * cie * de
       - He_2 * k[C_CIHEIS] * cie * (de * de)
       - He_2 * cie * de * k[C_REHEII1]
       - He_2 * cie * de * k[C_REHEII2]
       - He_3 * cie * de * k[C_REHEIII]
       - k[C_BREM] * cie * de * (H_2 + He_2 + R(4.0) * He_3)
       - cie * k[C_COMPTON] * de * pow(z + R(1.0), 4.0) * (T - R(2.73) * z - R(2.73))
       + R(0.5) * R(1.0) / (k[C_NCRN] / ncrD + R(1.0))
             * ((-H2_1) * H_1 * k[C_H2MCOOL] + pow(H_1, 3.0) * k[C_H2MHEAT]);
}

// Analytic Jacobian entries in physical units, dense row-major J[r][c] before the
// scale factors, following calculate_jacobian_cvklu (raja_primordial_network.cpp:1518-1867)
// including the replaced ("ad-hoc") ge-ge entry. inv_md is the local 1/mdensity (1512).
template <class R>
void jac_physical(const R (&y)[NSP], const R (&k)[NRATE], const R (&rk)[NRATE], R T, R z,
                  R Tge, const CellAux &ax, R (&J)[NSP][NSP], bool (&nz)[NSP][NSP]) {
  const R H2_1 = y[0], H2_2 = y[1], H_1 = y[2], H_2 = y[3], H_m0 = y[4], He_1 = y[5],
          He_2 = y[6], He_3 = y[7], de = y[8];
  const R k01 = k[K01], k02 = k[K02], k03 = k[K03], k04 = k[K04], k05 = k[K05], k06 = k[K06],
          k07 = k[K07], k08 = k[K08], k09 = k[K09], k10 = k[K10], k11 = k[K11], k12 = k[K12],
          k13 = k[K13], k14 = k[K14], k15 = k[K15], k16 = k[K16], k17 = k[K17], k18 = k[K18],
          k19 = k[K19], k21 = k[K21], k22 = k[K22];
  const R r01 = rk[K01], r02 = rk[K02], r03 = rk[K03], r04 = rk[K04], r05 = rk[K05],
          r06 = rk[K06], r07 = rk[K07], r08 = rk[K08], r09 = rk[K09], r10 = rk[K10],
          r11 = rk[K11], r12 = rk[K12], r13 = rk[K13], r14 = rk[K14], r15 = rk[K15],
          r16 = rk[K16], r17 = rk[K17], r18 = rk[K18], r19 = rk[K19], r21 = rk[K21],
          r22 = rk[K22];
  const R brem = k[C_BREM], ceHeI = k[C_CEHEI], ceHeII = k[C_CEHEII], ceHI = k[C_CEHI],
          cieco = k[C_CIECO], ciHeI = k[C_CIHEI], ciHeII = k[C_CIHEII], ciHeIS = k[C_CIHEIS],
          ciHI = k[C_CIHI], compton = k[C_COMPTON], gael = k[C_GAEL], gaH2 = k[C_GAH2],
          gaHe = k[C_GAHE], gaHI = k[C_GAHI], gaHp = k[C_GAHP], h2lte = k[C_H2LTE],
          h2mcool = k[C_H2MCOOL], h2mheat = k[C_H2MHEAT], ncrd1 = k[C_NCRD1],
          ncrd2 = k[C_NCRD2], ncrn = k[C_NCRN], reHeII1 = k[C_REHEII1],
          reHeII2 = k[C_REHEII2], reHeIII = k[C_REHEIII], reHII = k[C_REHII];
  const R rgael = rk[C_GAEL], rgaH2 = rk[C_GAH2], rgaHe = rk[C_GAHE], rgaHI = rk[C_GAHI],
          rgaHp = rk[C_GAHP], rh2lte = rk[C_H2LTE], rh2mcool = rk[C_H2MCOOL],
          rh2mheat = rk[C_H2MHEAT], rncrd1 = rk[C_NCRD1], rncrd2 = rk[C_NCRD2],
          rncrn = rk[C_NCRN];
  const R mdensity = R(ax.mdensity);
  const R inv_md = R(1.0) / mdensity;
  const R h2o = R(ax.h2_oda);
  for (int a = 0; a < NSP; a++)
    for (int b = 0; b < NSP; b++) { J[a][b] = R(0); nz[a][b] = false; }
  auto set = [&](int r, int c, R v) { J[r][c] = v; nz[r][c] = true; };
  const R H1sq = H_1 * H_1;

  // H2_1 row
  set(0, 0, (-k11) * H_2 - k12 * de - k13 * H_1 + k21 * H1sq);
  set(0, 1, k10 * H_1 + k19 * H_m0);
  set(0, 2, k08 * H_m0 + k10 * H2_2 - k13 * H2_1 + R(2) * k21 * H2_1 * H_1 + R(3) * k22 * H1sq);
  set(0, 3, (-k11) * H2_1);
  set(0, 4, k08 * H_1 + k19 * H2_2);
  set(0, 8, (-k12) * H2_1);
  set(0, 9, (r08 * H_1 * H_m0 + r10 * H2_2 * H_1 - r11 * H2_1 * H_2 - r12 * H2_1 * de
             - r13 * H2_1 * H_1 + r19 * H2_2 * H_m0 + r21 * H2_1 * H_1 * H_1
             + r22 * H_1 * H_1 * H_1) * Tge);
  // H2_2 row
  set(1, 0, k11 * H_2);
  set(1, 1, (-k10) * H_1 - k18 * de - k19 * H_m0);
  set(1, 2, k09 * H_2 - k10 * H2_2);
  set(1, 3, k09 * H_1 + k11 * H2_1 + k17 * H_m0);
  set(1, 4, k17 * H_2 - k19 * H2_2);
  set(1, 8, (-k18) * H2_2);
  set(1, 9, (r09 * H_1 * H_2 - r10 * H2_2 * H_1 + r11 * H2_1 * H_2 + r17 * H_2 * H_m0
             - r18 * H2_2 * de - r19 * H2_2 * H_m0) * Tge);
  // H_1 row
  set(2, 0, k11 * H_2 + R(2) * k12 * de + R(2) * k13 * H_1 - R(2) * k21 * H1sq);
  set(2, 1, (-k10) * H_1 + R(2) * k18 * de + k19 * H_m0);
  set(2, 2, (-k01) * de - k07 * de - k08 * H_m0 - k09 * H_2 - k10 * H2_2 + R(2) * k13 * H2_1
             + k15 * H_m0 - R(4) * k21 * H2_1 * H_1 - R(6) * k22 * H1sq);
  set(2, 3, k02 * de - k09 * H_1 + k11 * H2_1 + R(2) * k16 * H_m0);
  set(2, 4, (-k08) * H_1 + k14 * de + k15 * H_1 + R(2) * k16 * H_2 + k19 * H2_2);
  set(2, 8, (-k01) * H_1 + k02 * H_2 - k07 * H_1 + R(2) * k12 * H2_1 + k14 * H_m0
             + R(2) * k18 * H2_2);
  set(2, 9, ((-r01) * H_1 * de + r02 * H_2 * de - r07 * H_1 * de - r08 * H_1 * H_m0
             - r09 * H_1 * H_2 - r10 * H2_2 * H_1 + r11 * H2_1 * H_2 + R(2) * r12 * H2_1 * de
             + R(2) * r13 * H2_1 * H_1 + r14 * H_m0 * de + r15 * H_1 * H_m0
             + R(2) * r16 * H_2 * H_m0 + R(2) * r18 * H2_2 * de + r19 * H2_2 * H_m0
             - R(2) * r21 * H2_1 * H_1 * H_1 - R(2) * r22 * H_1 * H_1 * H_1) * Tge);
  // H_2 row
  set(3, 0, (-k11) * H_2);
  set(3, 1, k10 * H_1);
  set(3, 2, k01 * de - k09 * H_2 + k10 * H2_2);
  set(3, 3, (-k02) * de - k09 * H_1 - k11 * H2_1 - k16 * H_m0 - k17 * H_m0);
  set(3, 4, (-k16) * H_2 - k17 * H_2);
  set(3, 8, k01 * H_1 - k02 * H_2);
  set(3, 9, (r01 * H_1 * de - r02 * H_2 * de - r09 * H_1 * H_2 + r10 * H2_2 * H_1
             - r11 * H2_1 * H_2 - r16 * H_2 * H_m0 - r17 * H_2 * H_m0) * Tge);
  // H_m0 row
  set(4, 1, (-k19) * H_m0);
  set(4, 2, k07 * de - k08 * H_m0 - k15 * H_m0);
  set(4, 3, (-k16) * H_m0 - k17 * H_m0);
  set(4, 4, (-k08) * H_1 - k14 * de - k15 * H_1 - k16 * H_2 - k17 * H_2 - k19 * H2_2);
  set(4, 8, k07 * H_1 - k14 * H_m0);
  set(4, 9, (r07 * H_1 * de - r08 * H_1 * H_m0 - r14 * H_m0 * de - r15 * H_1 * H_m0
             - r16 * H_2 * H_m0 - r17 * H_2 * H_m0 - r19 * H2_2 * H_m0) * Tge);
  // He_1 row
  set(5, 5, (-k03) * de);
  set(5, 6, k04 * de);
  set(5, 8, (-k03) * He_1 + k04 * He_2);
  set(5, 9, ((-r03) * He_1 * de + r04 * He_2 * de) * Tge);
  // He_2 row
  set(6, 5, k03 * de);
  set(6, 6, (-k04) * de - k05 * de);
  set(6, 7, k06 * de);
  set(6, 8, k03 * He_1 - k04 * He_2 - k05 * He_2 + k06 * He_3);
  set(6, 9, (r03 * He_1 * de - r04 * He_2 * de - r05 * He_2 * de + r06 * He_3 * de) * Tge);
  // He_3 row
  set(7, 6, k05 * de);
  set(7, 7, (-k06) * de);
  set(7, 8, k05 * He_2 - k06 * He_3);
  set(7, 9, (r05 * He_2 * de - r06 * He_3 * de) * Tge);
  // de row
  set(8, 1, (-k18) * de);
  set(8, 2, k01 * de - k07 * de + k08 * H_m0 + k15 * H_m0);
  set(8, 3, (-k02) * de + k17 * H_m0);
  set(8, 4, k08 * H_1 + k14 * de + k15 * H_1 + k17 * H_2);
  set(8, 5, k03 * de);
  set(8, 6, (-k04) * de + k05 * de);
  set(8, 7, (-k06) * de);
  set(8, 8, k01 * H_1 - k02 * H_2 + k03 * He_1 - k04 * He_2 + k05 * He_2 - k06 * He_3
             - k07 * H_1 + k14 * H_m0 - k18 * H2_2);
  set(8, 9, (r01 * H_1 * de - r02 * H_2 * de + r03 * He_1 * de - r04 * He_2 * de
             + r05 * He_2 * de - r06 * He_3 * de - r07 * H_1 * de + r08 * H_1 * H_m0
             + r14 * H_m0 * de + r15 * H_1 * H_m0 + r17 * H_2 * H_m0 - r18 * H2_2 * de) * Tge);

  // ge row (each entry times the local 1/mdensity)
  const R S = H2_1 * gaH2 + H_1 * gaHI + H_2 * gaHp + He_1 * gaHe + de * gael;
  const R D = H2_1 * ncrd2 + H_1 * ncrd1;
  const R gden = h2lte / S + R(1.0);
  const R nden = ncrn / D + R(1.0);
  const R Y = (-H2_1) * H_1 * h2mcool + pow(H_1, 3.0) * h2mheat;
  auto glov = [&](R ga) {   // -H2_1*ga*h2lte^2*h2_oda/((h2lte/S+1)^2 * S^2)
    return (-H2_1) * ga * (h2lte * h2lte) * h2o / ((gden * gden) * (S * S));
  };
  set(9, 0, (glov(gaH2) - R(0.5) * H_1 * h2mcool * R(1.0) / nden
             - R(2.0158800000000001) * cieco * mdensity - h2lte * h2o / gden
             + R(0.5) * ncrd2 * ncrn * pow(nden, -2.0) * Y / (D * D)) * inv_md);
  set(9, 2, (glov(gaHI) - ceHI * de - ciHI * de
             + R(0.5) * ncrd1 * ncrn * pow(nden, -2.0) * Y / (D * D)
             + R(0.5) * ((-H2_1) * h2mcool + R(3) * H1sq * h2mheat) * R(1.0) / nden) * inv_md);
  set(9, 3, (glov(gaHp) - brem * de - de * reHII) * inv_md);
  set(9, 5, (glov(gaHe) - ciHeI * de) * inv_md);
  set(9, 6, ((-brem) * de - ceHeII * de - ceHeI * (de * de) - ciHeII * de
             - ciHeIS * (de * de) - de * reHeII1 - de * reHeII2) * inv_md);
  set(9, 7, (R(-4.0) * brem * de - de * reHeIII) * inv_md);
  set(9, 8, (glov(gael) - H_1 * ceHI - H_1 * ciHI - H_2 * reHII - He_1 * ciHeI
             - He_2 * ceHeII - R(2) * He_2 * ceHeI * de - He_2 * ciHeII
             - R(2) * He_2 * ciHeIS * de - He_2 * reHeII1 - He_2 * reHeII2 - He_3 * reHeIII
             - brem * (H_2 + He_2 + R(4.0) * He_3)
             - compton * pow(z + R(1.0), 4.0) * (T - R(2.73) * z - R(2.73))) * inv_md);
  const R Rg = (-H2_1) * rgaH2 - H_1 * rgaHI - H_2 * rgaHp - He_1 * rgaHe - de * rgael;
  const R adhoc =
      (-H2_1) * h2lte * h2o * ((-h2lte) * Rg / (S * S) - rh2lte / S) / (gden * gden)
      - H2_1 * h2o * rh2lte / gden
      + R(0.5) * pow(nden, -2.0) * Y
            * (R(-1.0) * ncrn * ((-H2_1) * rncrd2 - H_1 * rncrd1) / (D * D) - R(1.0) * rncrn / D)
      + R(0.5) * R(1.0) / nden * ((-H2_1) * H_1 * rh2mcool + pow(H_1, 3.0) * rh2mheat);
  set(9, 9, adhoc * inv_md * Tge);
}

}  // namespace orc
