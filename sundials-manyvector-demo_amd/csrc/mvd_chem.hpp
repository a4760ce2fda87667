// Primordial chemistry network (10 species: H2_1, H2_2, H_1, H_2, H_m0, He_1, He_2, He_3,
// de, ge) for gfx950, FP64: per-cell temperature solve, rate interpolation, RHS, analytic
// Jacobian. Device functions here are shared by the per-call kernels (chem.hip), the
// batched Newton/DIRK kernel and the adjoint kernels.
//
// State kept per cell (cvklu_data, raja_primordial_network.hpp:104-276): scale, inv_scale,
// Ts, dTs_ge, mdensity, inv_mdensity and the two optical-depth factors. The reference also
// caches 92 interpolated rate values per cell for the Jacobian; the rates are a pure
// function of the cached Ts, so here the Jacobian re-interpolates them from Ts instead of
// streaming 736 bytes per cell through HBM.
#pragma once
#include <hip/hip_runtime.h>

namespace mvd {

constexpr int NSP = 10;       // species per cell
constexpr int TLEN = 1024;    // table length
constexpr int NRATE = 46;     // 21 reaction + 25 cooling tables
constexpr int NTAB = 50;      // + 4 gamma tables

// table order of the committed blob (the reference's read order)
enum {
  K01 = 0, K02, K03, K04, K05, K06, K07, K08, K09, K10, K11, K12, K13, K14, K15, K16, K17,
  K18, K19, K21, K22,
  C_BREM, C_CEHEI, C_CEHEII, C_CEHI, C_CIECO, C_CIHEI, C_CIHEII, C_CIHEIS, C_CIHI, C_COMPTON,
  C_GAEL, C_GAH2, C_GAHE, C_GAHI, C_GAHP, C_H2LTE, C_H2MCOOL, C_H2MHEAT, C_NCRD1, C_NCRD2,
  C_NCRN, C_REHEII1, C_REHEII2, C_REHEIII, C_REHII,
  G_GH2_1, G_DGH2_1, G_GH2_2, G_DGH2_2
};

// Table grid constants (cvklu_setup_data, raja_primordial_network.cpp:191-195)
struct ChemConst {
  const double *tab;    // device, NTAB x TLEN
  // the NRATE rate tables again, bin-major with both interpolation ends of a bin side by
  // side: tabp[b * NRATE + r] = (tab[r][b], tab[r][b + 1]), b < nbins. One cell's rates at
  // one temperature are then NRATE consecutive 16-byte loads over about six cache lines
  // instead of one line in each of NRATE tables.
  const double2 *tabp;
  double dbin, idbin, lb, tmin, tmax, z;
  int nbins;
};

struct ChemState {
  bool ready = false;
  long n = 0;
  double *tab = nullptr;
  double2 *tabp = nullptr;
  ChemConst cc{};
  double *scale = nullptr, *inv_scale = nullptr, *Ts = nullptr, *dTs_ge = nullptr;
  double *mdensity = nullptr, *inv_mdensity = nullptr, *cie_oda = nullptr, *h2_oda = nullptr;
  double *work = nullptr;     // solver scratch (lazily sized)
  size_t work_bytes = 0;
  int *counters = nullptr;    // work queue / statistics
  int *host_counters = nullptr;
  // bumped by every setup and every reallocation of the work buffer: a captured graph that
  // holds this state's addresses or constants is stale when its generation differs
  unsigned long gen = 0;
};

int chem_setup(ChemState &s, const double *tables_host, long n, double z, hipStream_t st);
int chem_prepare(ChemState &s, double *chem, hipStream_t st);
// srhs / sjac: factor applied to every entry after the reference's expressions (1: none)
int chem_rhs(ChemState &s, const double *y, double *ydot, hipStream_t st, bool sync,
             double srhs = 1.0);
int chem_jac(ChemState &s, const double *y, double *J, hipStream_t st, bool sync,
             double sjac = 1.0);
int chem_vjp(ChemState &s, const double *y, const double *v, double *ybar, double *ydot,
             hipStream_t st, bool sync);
int chem_temperature_check(ChemState &s, const double *y, const double *Ts0, long *bad,
                           hipStream_t st);
double *chem_state_ptr(ChemState &s, int which);
void chem_free(ChemState &s);

// -------------------------------------------------------------------------------------
// device functions (reference operation order, no FMA contraction)
// -------------------------------------------------------------------------------------

// (int)(idbin*(log T - lb)) clamped to [0, nbins-1] (raja_primordial_network.cpp:953-958)
__device__ inline int chem_bin(const ChemConst &c, double logT) {
  const double x = c.idbin * (logT - c.lb);
  int b = (x == x) ? (int)fmin(fmax(x, -1.0), 2.0e9) : 0;
  if (b <= 0) b = 0;
  else if (b >= c.nbins) b = c.nbins - 1;
  return b;
}

// The four gamma tables of the temperature solve (gH2_1, dgH2_1, gH2_2, dgH2_2), read
// from the table blob in global memory or from a copy the caller staged in LDS.
struct GammaTabs {
  const double *g1, *dg1, *g2, *dg2;
};

__device__ inline GammaTabs gamma_tabs(const double *g4) {   // 4 consecutive tables
  return GammaTabs{g4, g4 + TLEN, g4 + 2 * TLEN, g4 + 3 * TLEN};
}

__device__ inline GammaTabs gamma_tabs(const ChemConst &c) {
  return gamma_tabs(c.tab + G_GH2_1 * TLEN);
}

// Mass density (in hydrogen masses) of the temperature solve
__device__ inline double temperature_density(const double (&y)[NSP]) {
#pragma clang fp contract(off)
  return 2.0 * y[0] + 2.0 * y[1] + 1.0079400000000001 * y[2] + 1.0079400000000001 * y[3] +
         1.0079400000000001 * y[4] + 4.0026020000000004 * y[5] + 4.0026020000000004 * y[6] +
         4.0026020000000004 * y[7];
}

// One Newton step of the temperature solve from the input T (density `dm` = density * mH):
// returns the next iterate and sets dge_dT at T (raja_primordial_network.cpp:943-990).
__device__ inline double temperature_step(const ChemConst &c, const GammaTabs &gt,
                                          const double (&y)[NSP], double dm, double T,
                                          double &dge_dT) {
#pragma clang fp contract(off)
  const double kb = 1.3806504e-16;
  const double gamma = 5.e0 / 3.e0;
  const double gm1 = 1.0 / (gamma - 1);
  const double lT = log(T);
  const int b = chem_bin(c, lT);
  const double t1 = c.lb + (double)b * c.dbin;
  const double t2 = c.lb + (double)(b + 1) * c.dbin;
  const double Tdef = (lT - t1) / (t2 - t1);
  const double gH2_2 = gt.g2[b] + Tdef * (gt.g2[b + 1] - gt.g2[b]);
  const double dgH2_2 = gt.dg2[b] + Tdef * (gt.dg2[b + 1] - gt.dg2[b]);
  const double gH2_1 = gt.g1[b] + Tdef * (gt.g1[b + 1] - gt.g1[b]);
  const double dgH2_1 = gt.dg1[b] + Tdef * (gt.dg1[b + 1] - gt.dg1[b]);
  const double g1 = 1.0 / (gH2_1 - 1.0);
  const double g2 = 1.0 / (gH2_2 - 1.0);
  const double nsum = y[0] * g1 + y[1] * g2 + y[2] * gm1 + y[3] * gm1 + y[4] * gm1 +
                      y[5] * gm1 + y[6] * gm1 + y[7] * gm1 + gm1 * y[8];
  dge_dT = T * kb * ((-y[0]) * g1 * g1 * dgH2_1 - y[1] * g2 * g2 * dgH2_2) / dm + kb * nsum / dm;
  const double dge = T * kb * nsum / dm - y[9];
  return T - dge / dge_dT;
}

// Longest cycle of the temperature iteration that ends the loop early.
constexpr int TCYCLE = 4;

// cvklu_calculate_temperature (raja_primordial_network.cpp:908-1011): 10 Newton steps
// from the warm start Ts, clamp to [1,1e5], dTs_ge = 1/dge_dT.
//
// With x_0 = Ts, x_{k+1} = S(x_k) and d_k = dge_dT(x_k), the result is (x_10, 1/d_9). The
// step map S and d depend only on the input x and the cell, so once an iterate repeats
// bitwise, x_m == x_{m-p}, the sequence is periodic with period p from index m-p on and
// the rest of the ten steps is known: x_10 = x_{m-p + (10-m) mod p} and
// d_9 = d_{m-p + (9-m) mod p}, both among the last p inputs. Leaving the loop at the first
// repeat with p <= TCYCLE (the inputs and d of the last TCYCLE steps are kept) gives
// bitwise the values of all ten steps with fewer evaluations. Fixed points and 2-cycles are
// the common exits of warm-started solves (2-4 steps); 3- and 4-cycles, which rounding
// makes near some roots, otherwise run the full ten steps and hold up their whole wave.
__device__ inline void chem_temperature(const ChemConst &c, const GammaTabs &gt,
                                        const double (&y)[NSP], double &Ts, double &dTs_ge) {
#pragma clang fp contract(off)
  const double dm = temperature_density(y) * 1.67e-24;
  double Tnew = Ts * 1.1, dge_dT = 0.0;
  double xh[TCYCLE], dh[TCYCLE];        // x_k, d_k of the last steps, newest first
#pragma unroll
  for (int q = 0; q < TCYCLE; q++) xh[q] = dh[q] = 0.0;
  for (int it = 0; it < 10; it++) {
    const double T = Ts;
    Tnew = temperature_step(c, gt, y, dm, T, dge_dT);
    Ts = Tnew;
#pragma unroll
    for (int q = TCYCLE - 1; q > 0; q--) {
      xh[q] = xh[q - 1];
      dh[q] = dh[q - 1];
    }
    xh[0] = T;                          // x_it
    dh[0] = dge_dT;                     // d_it
    // x_{it+1} == x_{it+1-p}: xh[p-1], which exists for p <= it + 1
    const long long bn = __double_as_longlong(Tnew);
    int p = 0;
#pragma unroll
    for (int q = TCYCLE; q >= 1; q--)
      if (q <= it + 1 && bn == __double_as_longlong(xh[q - 1])) p = q;
    if (p) {
      // m = it + 1; x_10 = xh[rx], d_9 = dh[rd] (offsets back from step it)
      const int rx = p - 1 - (9 - it) % p;
      const int rd = p - 1 - ((8 - it) % p + p) % p;
      double xv = xh[0], dv = dh[0];
#pragma unroll
      for (int q = 1; q < TCYCLE; q++) {
        if (rx == q) xv = xh[q];
        if (rd == q) dv = dh[q];
      }
      Tnew = xv;
      dge_dT = dv;
      break;
    }
  }
  Ts = Tnew;
  if (Ts < c.tmin) Ts = c.tmin;
  else if (Ts > c.tmax) Ts = c.tmax;
  dTs_ge = 1.0 / dge_dT;
}

// The ten steps without the early exit (the self-check of the exit compares the two)
__device__ inline void chem_temperature_ten(const ChemConst &c, const double (&y)[NSP],
                                            double &Ts, double &dTs_ge) {
#pragma clang fp contract(off)
  const GammaTabs gt = gamma_tabs(c);
  const double dm = temperature_density(y) * 1.67e-24;
  double Tnew = Ts * 1.1, dge_dT = 0.0;
  for (int it = 0; it < 10; it++) {
    Tnew = temperature_step(c, gt, y, dm, Ts, dge_dT);
    Ts = Tnew;
  }
  if (Ts < c.tmin) Ts = c.tmin;
  else if (Ts > c.tmax) Ts = c.tmax;
  dTs_ge = 1.0 / dge_dT;
}

__device__ inline void chem_temperature(const ChemConst &c, const double (&y)[NSP], double &Ts,
                                        double &dTs_ge) {
  chem_temperature(c, gamma_tabs(c), y, Ts, dTs_ge);
}

// linear-in-log(T) interpolation of every rate table and its T derivative
// (raja_primordial_network.cpp:1046-1245)
__device__ inline void chem_rates(const ChemConst &c, double T, double (&k)[NRATE],
                                  double (&dk)[NRATE]) {
#pragma clang fp contract(off)
  const double lT = log(T);
  const int b = chem_bin(c, lT);
  const double t1 = c.lb + (double)b * c.dbin;
  const double t2 = c.lb + (double)(b + 1) * c.dbin;
  const double Tdef = (lT - t1) / (t2 - t1);
  const double Tfactor = (1.0 / T) / (t2 - t1);
#pragma unroll
  for (int r = 0; r < NRATE; r++) {
    const double a0 = c.tab[r * TLEN + b], a1 = c.tab[r * TLEN + b + 1];
    k[r] = a0 + Tdef * (a1 - a0);
    dk[r] = (a1 - a0) * Tfactor;
  }
}

// Rates and derivatives of tables [r0, r1) only.
template <int R0, int R1>
__device__ inline void chem_rates_range(const ChemConst &c, double T, double (&k)[NRATE],
                                        double (&dk)[NRATE]) {
#pragma clang fp contract(off)
  const double lT = log(T);
  const int b = chem_bin(c, lT);
  const double t1 = c.lb + (double)b * c.dbin;
  const double t2 = c.lb + (double)(b + 1) * c.dbin;
  const double Tdef = (lT - t1) / (t2 - t1);
  const double Tfactor = (1.0 / T) / (t2 - t1);
#pragma unroll
  for (int r = R0; r < R1; r++) {
    const double a0 = c.tab[r * TLEN + b], a1 = c.tab[r * TLEN + b + 1];
    k[r] = a0 + Tdef * (a1 - a0);
    dk[r] = (a1 - a0) * Tfactor;
  }
}

// Rate values of tables [R0, R1) only.
template <int R0, int R1>
__device__ inline void chem_rates_only_range(const ChemConst &c, double T, double (&k)[NRATE]) {
#pragma clang fp contract(off)
  const double lT = log(T);
  const int b = chem_bin(c, lT);
  const double t1 = c.lb + (double)b * c.dbin;
  const double t2 = c.lb + (double)(b + 1) * c.dbin;
  const double Tdef = (lT - t1) / (t2 - t1);
#pragma unroll
  for (int r = R0; r < R1; r++) {
    const double a0 = c.tab[r * TLEN + b], a1 = c.tab[r * TLEN + b + 1];
    k[r] = a0 + Tdef * (a1 - a0);
  }
}

// chem_rates_range and chem_rates_only_range from the bin-major pair table: the same
// values (the same table entries and arithmetic), fewer and denser loads.
template <int R0, int R1>
__device__ inline void chem_rates_pairs(const ChemConst &c, double T, double (&k)[NRATE],
                                        double (&dk)[NRATE]) {
#pragma clang fp contract(off)
  const double lT = log(T);
  const int b = chem_bin(c, lT);
  const double t1 = c.lb + (double)b * c.dbin;
  const double t2 = c.lb + (double)(b + 1) * c.dbin;
  const double Tdef = (lT - t1) / (t2 - t1);
  const double Tfactor = (1.0 / T) / (t2 - t1);
  const double2 *row = c.tabp + (long)b * NRATE;
#pragma unroll
  for (int r = R0; r < R1; r++) {
    const double2 a = row[r];
    k[r] = a.x + Tdef * (a.y - a.x);
    dk[r] = (a.y - a.x) * Tfactor;
  }
}

template <int R0, int R1>
__device__ inline void chem_rates_only_pairs(const ChemConst &c, double T, double (&k)[NRATE]) {
#pragma clang fp contract(off)
  const double lT = log(T);
  const int b = chem_bin(c, lT);
  const double t1 = c.lb + (double)b * c.dbin;
  const double t2 = c.lb + (double)(b + 1) * c.dbin;
  const double Tdef = (lT - t1) / (t2 - t1);
  const double2 *row = c.tabp + (long)b * NRATE;
#pragma unroll
  for (int r = R0; r < R1; r++) {
    const double2 a = row[r];
    k[r] = a.x + Tdef * (a.y - a.x);
  }
}

// Only the rate values (no derivatives).
__device__ inline void chem_rates_only(const ChemConst &c, double T, double (&k)[NRATE]) {
#pragma clang fp contract(off)
  const double lT = log(T);
  const int b = chem_bin(c, lT);
  const double t1 = c.lb + (double)b * c.dbin;
  const double t2 = c.lb + (double)(b + 1) * c.dbin;
  const double Tdef = (lT - t1) / (t2 - t1);
#pragma unroll
  for (int r = 0; r < NRATE; r++) {
    const double a0 = c.tab[r * TLEN + b], a1 = c.tab[r * TLEN + b + 1];
    k[r] = a0 + Tdef * (a1 - a0);
  }
}

struct CellAux {
  double mdensity, inv_mdensity, cie, h2o;
};

// The 21 two- and three-body reaction rates r_k = k_k * prod(reactants), each multiplied in
// the order the network's expanded RHS uses, so that summing them in the RHS's term order
// reproduces the expanded expressions exactly.
struct Reactions {
  double r01, r02, r03, r04, r05, r06, r07, r08, r09, r10, r11, r12, r13, r14, r15, r16, r17,
      r18, r19, r21, r22;
};

__device__ inline Reactions chem_reactions(const double (&y)[NSP], const double (&k)[NRATE]) {
#pragma clang fp contract(off)
  const double H2_1 = y[0], H2_2 = y[1], H_1 = y[2], H_2 = y[3], H_m0 = y[4], He_1 = y[5],
               He_2 = y[6], He_3 = y[7], de = y[8];
  Reactions R;
  R.r01 = k[K01] * H_1 * de;
  R.r02 = k[K02] * H_2 * de;
  R.r03 = k[K03] * He_1 * de;
  R.r04 = k[K04] * He_2 * de;
  R.r05 = k[K05] * He_2 * de;
  R.r06 = k[K06] * He_3 * de;
  R.r07 = k[K07] * H_1 * de;
  R.r08 = k[K08] * H_1 * H_m0;
  R.r09 = k[K09] * H_1 * H_2;
  R.r10 = k[K10] * H2_2 * H_1;
  R.r11 = k[K11] * H2_1 * H_2;
  R.r12 = k[K12] * H2_1 * de;
  R.r13 = k[K13] * H2_1 * H_1;
  R.r14 = k[K14] * H_m0 * de;
  R.r15 = k[K15] * H_1 * H_m0;
  R.r16 = k[K16] * H_2 * H_m0;
  R.r17 = k[K17] * H_2 * H_m0;
  R.r18 = k[K18] * H2_2 * de;
  R.r19 = k[K19] * H2_2 * H_m0;
  R.r21 = k[K21] * H2_1 * H_1 * H_1;
  R.r22 = k[K22] * H_1 * H_1 * H_1;
  return R;
}

// Species and gas-energy RHS in physical units (raja_primordial_network.cpp:1302-1371,
// before the inv_scale / inv_mdensity factors).
//
// Split in two like the Jacobian: the species rows read only the 21 reaction rates, the
// gas-energy row only the 25 cooling rates.
__device__ inline void chem_rhs_species(const double (&y)[NSP], const double (&k)[NRATE],
                                        double (&f)[NSP]) {
#pragma clang fp contract(off)
  const Reactions R = chem_reactions(y, k);
  f[0] = R.r08 + R.r10 - R.r11 - R.r12 - R.r13 + R.r19 + R.r21 + R.r22;
  f[1] = R.r09 - R.r10 + R.r11 + R.r17 - R.r18 - R.r19;
  f[2] = -R.r01 + R.r02 - R.r07 - R.r08 - R.r09 - R.r10 + R.r11 + 2.0 * R.r12 + 2.0 * R.r13 +
         R.r14 + R.r15 + 2.0 * R.r16 + 2.0 * R.r18 + R.r19 - 2.0 * R.r21 - 2.0 * R.r22;
  f[3] = R.r01 - R.r02 - R.r09 + R.r10 - R.r11 - R.r16 - R.r17;
  f[4] = R.r07 - R.r08 - R.r14 - R.r15 - R.r16 - R.r17 - R.r19;
  f[5] = -R.r03 + R.r04;
  f[6] = R.r03 - R.r04 - R.r05 + R.r06;
  f[7] = R.r05 - R.r06;
  f[8] = R.r01 - R.r02 + R.r03 - R.r04 + R.r05 - R.r06 - R.r07 + R.r08 + R.r14 + R.r15 + R.r17 -
         R.r18;
}

template <class R>
__device__ inline R chem_rhs_energy_t(const R (&y)[NSP], const R (&k)[NRATE], R T, double z,
                                      const CellAux &ax) {
#pragma clang fp contract(off)
  const R H2_1 = y[0], H_1 = y[2], H_2 = y[3], He_1 = y[5], He_2 = y[6], He_3 = y[7],
          de = y[8];
  const double cie = ax.cie, h2o = ax.h2o;
  const R gS = H2_1 * k[C_GAH2] + H_1 * k[C_GAHI] + H_2 * k[C_GAHP] + He_1 * k[C_GAHE] +
               de * k[C_GAEL];
  const R ncrD = H2_1 * k[C_NCRD2] + H_1 * k[C_NCRD1];
  const double zp1 = z + 1.0;
  const double zp4 = (zp1 * zp1) * (zp1 * zp1);
  return -2.0158800000000001 * H2_1 * k[C_CIECO] * cie * ax.mdensity -
         H2_1 * cie * k[C_H2LTE] * h2o / (k[C_H2LTE] / gS + 1.0) - H_1 * k[C_CEHI] * cie * de -
         H_1 * k[C_CIHI] * cie * de - H_2 * cie * de * k[C_REHII] -
         He_1 * k[C_CIHEI] * cie * de - He_2 * k[C_CEHEII] * cie * de -
         He_2 * k[C_CEHEI] * cie * (de * de) - He_2 * k[C_CIHEII] * cie * de -
         He_2 * k[C_CIHEIS] * cie * (de * de) - He_2 * cie * de * k[C_REHEII1] -
         He_2 * cie * de * k[C_REHEII2] - He_3 * cie * de * k[C_REHEIII] -
         k[C_BREM] * cie * de * (H_2 + He_2 + 4.0 * He_3) -
         cie * k[C_COMPTON] * de * zp4 * (T - 2.73 * z - 2.73) +
         0.5 * 1.0 / (k[C_NCRN] / ncrD + 1.0) *
             ((-H2_1) * H_1 * k[C_H2MCOOL] + (H_1 * H_1 * H_1) * k[C_H2MHEAT]);
}

__device__ inline double chem_rhs_energy(const double (&y)[NSP], const double (&k)[NRATE],
                                         double T, double z, const CellAux &ax) {
  return chem_rhs_energy_t<double>(y, k, T, z, ax);
}

// Gradient of the gas-energy row f_9 = chem_rhs_energy_t(y, k, T) by one reverse sweep of
// its expression: g[s] = d f_9 / d y_s (s = 0..8; zero for the species it does not read)
// and g[9] = d f_9 / dT through the rates (dk = dk/dT) and the Compton term's explicit T.
// The same partial derivatives as forward tangents along each direction, formed once.
__device__ inline void chem_energy_grad(const double (&y)[NSP], const double (&k)[NRATE],
                                        const double (&dk)[NRATE], double T, double z,
                                        const CellAux &ax, double (&g)[NSP]) {
#pragma clang fp contract(off)
  const double H2_1 = y[0], H_1 = y[2], H_2 = y[3], He_1 = y[5], He_2 = y[6], He_3 = y[7],
               de = y[8];
  const double c = ax.cie, h2o = ax.h2o, md = ax.mdensity;
  double kb[NRATE];                                 // d f_9 / d k_r
#pragma unroll
  for (int r = 0; r < NRATE; r++) kb[r] = 0.0;
  double b0 = 0.0, b2 = 0.0, b3 = 0.0, b5 = 0.0, b6 = 0.0, b7 = 0.0, b8 = 0.0;
  // -2.01588 H2_1 k_CIECO cie mdensity
  b0 += -2.0158800000000001 * k[C_CIECO] * c * md;
  kb[C_CIECO] += -2.0158800000000001 * H2_1 * c * md;
  // -H2_1 cie k_H2LTE h2o / (k_H2LTE / gS + 1)
  {
    const double gS = H2_1 * k[C_GAH2] + H_1 * k[C_GAHI] + H_2 * k[C_GAHP] + He_1 * k[C_GAHE] +
                      de * k[C_GAEL];
    const double D = k[C_H2LTE] / gS + 1.0, iD = 1.0 / D;
    const double N = H2_1 * c * k[C_H2LTE] * h2o;
    const double Db = N * iD * iD;                  // d(-N/D)/dD
    const double gSb = Db * (-k[C_H2LTE] / (gS * gS));
    b0 += -(c * k[C_H2LTE] * h2o) * iD;
    kb[C_H2LTE] += -(H2_1 * c * h2o) * iD + Db / gS;
    b0 += gSb * k[C_GAH2];  kb[C_GAH2] += gSb * H2_1;
    b2 += gSb * k[C_GAHI];  kb[C_GAHI] += gSb * H_1;
    b3 += gSb * k[C_GAHP];  kb[C_GAHP] += gSb * H_2;
    b5 += gSb * k[C_GAHE];  kb[C_GAHE] += gSb * He_1;
    b8 += gSb * k[C_GAEL];  kb[C_GAEL] += gSb * de;
  }
  // atomic and ionic line cooling, recombination: -X k cie de (and de^2) terms
  b2 += -(k[C_CEHI] + k[C_CIHI]) * c * de;
  b8 += -(k[C_CEHI] + k[C_CIHI]) * c * H_1;
  kb[C_CEHI] += -H_1 * c * de;
  kb[C_CIHI] += -H_1 * c * de;
  b3 += -c * de * k[C_REHII];
  b8 += -H_2 * c * k[C_REHII];
  kb[C_REHII] += -H_2 * c * de;
  b5 += -k[C_CIHEI] * c * de;
  b8 += -He_1 * k[C_CIHEI] * c;
  kb[C_CIHEI] += -He_1 * c * de;
  {
    const double sa = k[C_CEHEII] + k[C_CIHEII] + k[C_REHEII1] + k[C_REHEII2];
    const double sb = k[C_CEHEI] + k[C_CIHEIS];
    b6 += -c * (de * sa + (de * de) * sb);
    b8 += -He_2 * c * (sa + 2.0 * de * sb);
    const double ka = -He_2 * c * de, kq = -He_2 * c * (de * de);
    kb[C_CEHEII] += ka; kb[C_CIHEII] += ka; kb[C_REHEII1] += ka; kb[C_REHEII2] += ka;
    kb[C_CEHEI] += kq;  kb[C_CIHEIS] += kq;
  }
  b7 += -c * de * k[C_REHEIII];
  b8 += -He_3 * c * k[C_REHEIII];
  kb[C_REHEIII] += -He_3 * c * de;
  // bremsstrahlung
  {
    const double S = H_2 + He_2 + 4.0 * He_3;
    kb[C_BREM] += -c * de * S;
    b8 += -k[C_BREM] * c * S;
    const double sb = -k[C_BREM] * c * de;
    b3 += sb; b6 += sb; b7 += 4.0 * sb;
  }
  // Compton: -cie k_COMPTON de (z+1)^4 (T - 2.73 z - 2.73)
  const double zp1 = z + 1.0;
  const double zp4 = (zp1 * zp1) * (zp1 * zp1);
  const double dTc = T - 2.73 * z - 2.73;
  kb[C_COMPTON] += -c * de * zp4 * dTc;
  b8 += -c * k[C_COMPTON] * zp4 * dTc;
  double Tb = -c * k[C_COMPTON] * de * zp4;
  // H2 formation heating / cooling: 0.5 / (k_NCRN / ncrD + 1) Q
  {
    const double ncrD = H2_1 * k[C_NCRD2] + H_1 * k[C_NCRD1];
    const double E = k[C_NCRN] / ncrD + 1.0, iE = 1.0 / E;
    const double Q = (-H2_1) * H_1 * k[C_H2MCOOL] + (H_1 * H_1 * H_1) * k[C_H2MHEAT];
    const double Qb = 0.5 * iE, Eb = -0.5 * Q * iE * iE;
    const double ncrDb = Eb * (-k[C_NCRN] / (ncrD * ncrD));
    kb[C_NCRN] += Eb / ncrD;
    b0 += Qb * (-H_1 * k[C_H2MCOOL]) + ncrDb * k[C_NCRD2];
    b2 += Qb * (-H2_1 * k[C_H2MCOOL] + 3.0 * (H_1 * H_1) * k[C_H2MHEAT]) + ncrDb * k[C_NCRD1];
    kb[C_H2MCOOL] += Qb * (-H2_1 * H_1);
    kb[C_H2MHEAT] += Qb * (H_1 * H_1 * H_1);
    kb[C_NCRD2] += ncrDb * H2_1;
    kb[C_NCRD1] += ncrDb * H_1;
  }
#pragma unroll
  for (int r = C_BREM; r < NRATE; r++) Tb += kb[r] * dk[r];
#pragma unroll
  for (int s = 0; s < NSP; s++) g[s] = 0.0;
  g[0] = b0; g[2] = b2; g[3] = b3; g[5] = b5; g[6] = b6; g[7] = b7; g[8] = b8;
  g[NSP - 1] = Tb;
}

// ---- vector-Jacobian product of the RHS --------------------------------------------

// Reactions in the order of Reactions (r01..r19, r21, r22): rate index (the same position in
// the rate tables) and reactant species (-1: none); stoichiometry of the nine species rows
// (f_s = sum_j CHEM_STO[s][j] r_j), read off chem_rhs_species.
constexpr int NREACT = 21;
__device__ constexpr signed char CHEM_REACTANTS[NREACT][3] = {
    {2, 8, -1}, {3, 8, -1}, {5, 8, -1}, {6, 8, -1}, {6, 8, -1}, {7, 8, -1}, {2, 8, -1},
    {2, 4, -1}, {2, 3, -1}, {1, 2, -1}, {0, 3, -1}, {0, 8, -1}, {0, 2, -1}, {4, 8, -1},
    {2, 4, -1}, {3, 4, -1}, {3, 4, -1}, {1, 8, -1}, {1, 4, -1}, {0, 2, 2}, {2, 2, 2}};
__device__ constexpr signed char CHEM_STO[NSP - 1][NREACT] = {
    // r01 02 03 04 05 06 07 08 09 10 11 12 13 14 15 16 17 18 19 21 22
    {0, 0, 0, 0, 0, 0, 0, 1, 0, 1, -1, -1, -1, 0, 0, 0, 0, 0, 1, 1, 1},      // H2_1
    {0, 0, 0, 0, 0, 0, 0, 0, 1, -1, 1, 0, 0, 0, 0, 0, 1, -1, -1, 0, 0},      // H2_2
    {-1, 1, 0, 0, 0, 0, -1, -1, -1, -1, 1, 2, 2, 1, 1, 2, 0, 2, 1, -2, -2},  // H_1
    {1, -1, 0, 0, 0, 0, 0, 0, -1, 1, -1, 0, 0, 0, 0, -1, -1, 0, 0, 0, 0},    // H_2
    {0, 0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 0, -1, -1, -1, -1, 0, -1, 0, 0},   // H_m0
    {0, 0, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},        // He_1
    {0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},       // He_2
    {0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},        // He_3
    {1, -1, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0, 0, 1, 1, 0, 1, -1, 0, 0, 0}};   // de

// Gradient of the physical RHS contracted with vp: yab = (df/dy)^T vp for y (physical
// units) with the temperature T = T(y), the root of the temperature equation
// g(T, y) = T kb nsum(T, y) / (density(y) mh) - ge = 0 (so dT/dy = -(dg/dy) / (dg/dT), with
// dg/dT from the slopes of the interpolated gamma tables). The species rows are
// differentiated in reverse through the reactions, the gas-energy row by a reverse sweep
// of its expression (chem_energy_grad). `clamped`: T sits on a clamp bound, so it
// does not depend on y.
//
// In two parts: chem_vjp_prepare forms what does not depend on vp (reaction rates and
// their T-derivatives, the gas-energy row's tangents, the temperature path's dg/dT and
// dg/dy), chem_vjp_apply contracts one cotangent with it; chem_vjp_row gives the row
// vp = v e_r with the same arithmetic restricted to its nonzero terms (the exact Jacobian
// one row at a time, for many cotangents at one state).
struct ChemVjpPre {
  double k[NREACT], dk[NREACT];   // reaction rates and T-derivatives (reaction order)
  double e9[NSP];                 // d f_9 / d y_s (H2_1, H_1, H_2, He_1..3, de); [9]: d/dT
  double dgdT, dgdy[NSP - 1];     // temperature path (not formed when clamped)
  bool clamped;
};

// PAIRS: the rates from the bin-major pair table (same values).
template <bool PAIRS>
__device__ inline void chem_vjp_prepare(const ChemConst &c, const double (&y)[NSP], double T,
                                        bool clamped, const CellAux &ax, ChemVjpPre &P) {
#pragma clang fp contract(off)
  double k[NRATE], dk[NRATE];
  if (PAIRS) chem_rates_pairs<0, NRATE>(c, T, k, dk);
  else chem_rates(c, T, k, dk);
#pragma unroll
  for (int j = 0; j < NREACT; j++) {
    P.k[j] = k[j];
    P.dk[j] = dk[j];
  }
  // gas-energy row: its gradient along H2_1, H_1, H_2, He_1, He_2, He_3, de and T (in
  // slot 9: ge has none)
  chem_energy_grad(y, k, dk, T, c.z, ax, P.e9);
  P.clamped = clamped;
  P.dgdT = 1.0;
#pragma unroll
  for (int s = 0; s < NSP - 1; s++) P.dgdy[s] = 0.0;
  if (clamped) return;
  // temperature path
  const double kb = 1.3806504e-16, mh = 1.67e-24;
  const double gm1 = 1.0 / (5.e0 / 3.e0 - 1);
  const double lT = log(T);
  const int bn = chem_bin(c, lT);
  const double t1 = c.lb + (double)bn * c.dbin;
  const double t2 = c.lb + (double)(bn + 1) * c.dbin;
  const double Tdef = (lT - t1) / (t2 - t1);
  const double Tfactor = (1.0 / T) / (t2 - t1);
  const double *g1t = c.tab + G_GH2_1 * TLEN, *g2t = c.tab + G_GH2_2 * TLEN;
  const double gH2_1 = g1t[bn] + Tdef * (g1t[bn + 1] - g1t[bn]);
  const double gH2_2 = g2t[bn] + Tdef * (g2t[bn + 1] - g2t[bn]);
  const double sl1 = (g1t[bn + 1] - g1t[bn]) * Tfactor;       // d gamma_H2_1 / dT
  const double sl2 = (g2t[bn + 1] - g2t[bn]) * Tfactor;
  const double g1 = 1.0 / (gH2_1 - 1.0), g2 = 1.0 / (gH2_2 - 1.0);
  const double cs[NSP - 1] = {g1, g2, gm1, gm1, gm1, gm1, gm1, gm1, gm1};
  const double ws[NSP - 1] = {2.0, 2.0, 1.0079400000000001, 1.0079400000000001,
                              1.0079400000000001, 4.0026020000000004, 4.0026020000000004,
                              4.0026020000000004, 0.0};
  double nsum = 0.0, density = 0.0;
#pragma unroll
  for (int s = 0; s < NSP - 1; s++) {
    nsum += y[s] * cs[s];
    density += ws[s] * y[s];
  }
  const double dm = density * mh;
  P.dgdT = kb * nsum / dm + T * kb * ((-y[0]) * g1 * g1 * sl1 - y[1] * g2 * g2 * sl2) / dm;
#pragma unroll
  for (int s = 0; s < NSP - 1; s++)
    P.dgdy[s] = T * kb / mh * (cs[s] * density - nsum * ws[s]) / (density * density);
}

// The temperature path's share of yab for the T-cotangent Tb.
__device__ inline void chem_vjp_temperature(const ChemVjpPre &P, double Tb,
                                            double (&yab)[NSP]) {
#pragma clang fp contract(off)
  if (P.clamped) return;
  const double TbT = Tb / P.dgdT;
#pragma unroll
  for (int s = 0; s < NSP - 1; s++) yab[s] -= TbT * P.dgdy[s];
  yab[NSP - 1] += TbT;                                            // dg/dge = -1
}

__device__ inline void chem_vjp_apply(const ChemVjpPre &P, const double (&y)[NSP],
                                      const double (&vp)[NSP], double (&yab)[NSP]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int s = 0; s < NSP; s++) yab[s] = 0.0;
  double Tb = 0.0;
  // species rows: reverse through r_j = k_j * prod(reactants)
#pragma unroll
  for (int j = 0; j < NREACT; j++) {
    double rb = 0.0;
#pragma unroll
    for (int s = 0; s < NSP - 1; s++)
      if (CHEM_STO[s][j] != 0) rb += (double)CHEM_STO[s][j] * vp[s];
    const int a = CHEM_REACTANTS[j][0], b = CHEM_REACTANTS[j][1], cc = CHEM_REACTANTS[j][2];
    const double ya = y[a], yb = y[b], yc = cc >= 0 ? y[cc] : 1.0;
    yab[a] += rb * (P.k[j] * yb * yc);
    yab[b] += rb * (P.k[j] * ya * yc);
    if (cc >= 0) yab[cc] += rb * (P.k[j] * ya * yb);
    Tb += rb * (P.dk[j] * ya * yb * yc);
  }
  // gas-energy row
  constexpr int dirs[7] = {0, 2, 3, 5, 6, 7, 8};
#pragma unroll
  for (int q = 0; q < 7; q++) yab[dirs[q]] += vp[9] * P.e9[dirs[q]];
  Tb += vp[9] * P.e9[NSP - 1];
  chem_vjp_temperature(P, Tb, yab);
}

// chem_vjp_apply for vp = v e_r (row r of the Jacobian times v): the terms of the other
// rows, all zero, are left out.
__device__ inline void chem_vjp_row(const ChemVjpPre &P, const double (&y)[NSP], int r,
                                    double v, double (&row)[NSP]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int s = 0; s < NSP; s++) row[s] = 0.0;
  double Tb = 0.0;
  if (r < NSP - 1) {
#pragma unroll
    for (int j = 0; j < NREACT; j++) {
      if (CHEM_STO[r][j] == 0) continue;
      const double rb = (double)CHEM_STO[r][j] * v;
      const int a = CHEM_REACTANTS[j][0], b = CHEM_REACTANTS[j][1], cc = CHEM_REACTANTS[j][2];
      const double ya = y[a], yb = y[b], yc = cc >= 0 ? y[cc] : 1.0;
      row[a] += rb * (P.k[j] * yb * yc);
      row[b] += rb * (P.k[j] * ya * yc);
      if (cc >= 0) row[cc] += rb * (P.k[j] * ya * yb);
      Tb += rb * (P.dk[j] * ya * yb * yc);
    }
  } else {
    constexpr int dirs[7] = {0, 2, 3, 5, 6, 7, 8};
#pragma unroll
    for (int q = 0; q < 7; q++) row[dirs[q]] += v * P.e9[dirs[q]];
    Tb += v * P.e9[NSP - 1];
  }
  chem_vjp_temperature(P, Tb, row);
}

template <bool PAIRS>
__device__ inline void chem_vjp_physical(const ChemConst &c, const double (&y)[NSP], double T,
                                         bool clamped, const CellAux &ax,
                                         const double (&vp)[NSP], double (&yab)[NSP]) {
  ChemVjpPre P;
  chem_vjp_prepare<PAIRS>(c, y, T, clamped, ax, P);
  chem_vjp_apply(P, y, vp, yab);
}

__device__ inline void chem_rhs_physical(const double (&y)[NSP], const double (&k)[NRATE],
                                         double T, double z, const CellAux &ax,
                                         double (&f)[NSP]) {
  chem_rhs_species(y, k, f);
  f[9] = chem_rhs_energy(y, k, T, z, ax);
}

// Analytic Jacobian, physical units, row-major J[r][c] with the reference's extra
// factors folded in (Tge on the ge column, 1/mdensity on the ge row, the replaced ge-ge
// entry), raja_primordial_network.cpp:1518-1867. Structural zeros are left at 0.
//
// Split in two: the species rows 0-8 read only the 21 reaction rates (and their
// T-derivatives), the gas-energy row 9 only the 25 cooling rates, so a caller short of
// registers can interpolate and consume one group at a time.
__device__ inline void chem_jac_species(const double (&y)[NSP], const double (&k)[NRATE],
                                        const double (&rk)[NRATE], double Tge,
                                        double (&J)[NSP][NSP]) {
#pragma clang fp contract(off)
  const double H2_1 = y[0], H2_2 = y[1], H_1 = y[2], H_2 = y[3], H_m0 = y[4], He_1 = y[5],
               He_2 = y[6], He_3 = y[7], de = y[8];
  const double k01 = k[K01], k02 = k[K02], k03 = k[K03], k04 = k[K04], k05 = k[K05],
               k06 = k[K06], k07 = k[K07], k08 = k[K08], k09 = k[K09], k10 = k[K10],
               k11 = k[K11], k12 = k[K12], k13 = k[K13], k14 = k[K14], k15 = k[K15],
               k16 = k[K16], k17 = k[K17], k18 = k[K18], k19 = k[K19], k21 = k[K21],
               k22 = k[K22];
  const double H1sq = H_1 * H_1;
  // temperature derivatives of the species rows: d(rhs_s)/dT at fixed species, times Tge
  double dk[NRATE];
#pragma unroll
  for (int r = 0; r < NRATE; r++) dk[r] = rk[r];
  const Reactions D = chem_reactions(y, dk);   // rates built from dk/dT instead of k

  J[0][0] = (-k11) * H_2 - k12 * de - k13 * H_1 + k21 * H1sq;
  J[0][1] = k10 * H_1 + k19 * H_m0;
  J[0][2] = k08 * H_m0 + k10 * H2_2 - k13 * H2_1 + 2.0 * k21 * H2_1 * H_1 + 3.0 * k22 * H1sq;
  J[0][3] = (-k11) * H2_1;
  J[0][4] = k08 * H_1 + k19 * H2_2;
  J[0][8] = (-k12) * H2_1;
  J[0][9] = (D.r08 + D.r10 - D.r11 - D.r12 - D.r13 + D.r19 + D.r21 + D.r22) * Tge;

  J[1][0] = k11 * H_2;
  J[1][1] = (-k10) * H_1 - k18 * de - k19 * H_m0;
  J[1][2] = k09 * H_2 - k10 * H2_2;
  J[1][3] = k09 * H_1 + k11 * H2_1 + k17 * H_m0;
  J[1][4] = k17 * H_2 - k19 * H2_2;
  J[1][8] = (-k18) * H2_2;
  J[1][9] = (D.r09 - D.r10 + D.r11 + D.r17 - D.r18 - D.r19) * Tge;

  J[2][0] = k11 * H_2 + 2.0 * k12 * de + 2.0 * k13 * H_1 - 2.0 * k21 * H1sq;
  J[2][1] = (-k10) * H_1 + 2.0 * k18 * de + k19 * H_m0;
  J[2][2] = (-k01) * de - k07 * de - k08 * H_m0 - k09 * H_2 - k10 * H2_2 + 2.0 * k13 * H2_1 +
            k15 * H_m0 - 4.0 * k21 * H2_1 * H_1 - 6.0 * k22 * H1sq;
  J[2][3] = k02 * de - k09 * H_1 + k11 * H2_1 + 2.0 * k16 * H_m0;
  J[2][4] = (-k08) * H_1 + k14 * de + k15 * H_1 + 2.0 * k16 * H_2 + k19 * H2_2;
  J[2][8] = (-k01) * H_1 + k02 * H_2 - k07 * H_1 + 2.0 * k12 * H2_1 + k14 * H_m0 +
            2.0 * k18 * H2_2;
  J[2][9] = (-D.r01 + D.r02 - D.r07 - D.r08 - D.r09 - D.r10 + D.r11 + 2.0 * D.r12 +
             2.0 * D.r13 + D.r14 + D.r15 + 2.0 * D.r16 + 2.0 * D.r18 + D.r19 - 2.0 * D.r21 -
             2.0 * D.r22) * Tge;

  J[3][0] = (-k11) * H_2;
  J[3][1] = k10 * H_1;
  J[3][2] = k01 * de - k09 * H_2 + k10 * H2_2;
  J[3][3] = (-k02) * de - k09 * H_1 - k11 * H2_1 - k16 * H_m0 - k17 * H_m0;
  J[3][4] = (-k16) * H_2 - k17 * H_2;
  J[3][8] = k01 * H_1 - k02 * H_2;
  J[3][9] = (D.r01 - D.r02 - D.r09 + D.r10 - D.r11 - D.r16 - D.r17) * Tge;

  J[4][1] = (-k19) * H_m0;
  J[4][2] = k07 * de - k08 * H_m0 - k15 * H_m0;
  J[4][3] = (-k16) * H_m0 - k17 * H_m0;
  J[4][4] = (-k08) * H_1 - k14 * de - k15 * H_1 - k16 * H_2 - k17 * H_2 - k19 * H2_2;
  J[4][8] = k07 * H_1 - k14 * H_m0;
  J[4][9] = (D.r07 - D.r08 - D.r14 - D.r15 - D.r16 - D.r17 - D.r19) * Tge;

  J[5][5] = (-k03) * de;
  J[5][6] = k04 * de;
  J[5][8] = (-k03) * He_1 + k04 * He_2;
  J[5][9] = (-D.r03 + D.r04) * Tge;

  J[6][5] = k03 * de;
  J[6][6] = (-k04) * de - k05 * de;
  J[6][7] = k06 * de;
  J[6][8] = k03 * He_1 - k04 * He_2 - k05 * He_2 + k06 * He_3;
  J[6][9] = (D.r03 - D.r04 - D.r05 + D.r06) * Tge;

  J[7][6] = k05 * de;
  J[7][7] = (-k06) * de;
  J[7][8] = k05 * He_2 - k06 * He_3;
  J[7][9] = (D.r05 - D.r06) * Tge;

  J[8][1] = (-k18) * de;
  J[8][2] = k01 * de - k07 * de + k08 * H_m0 + k15 * H_m0;
  J[8][3] = (-k02) * de + k17 * H_m0;
  J[8][4] = k08 * H_1 + k14 * de + k15 * H_1 + k17 * H_2;
  J[8][5] = k03 * de;
  J[8][6] = (-k04) * de + k05 * de;
  J[8][7] = (-k06) * de;
  J[8][8] = k01 * H_1 - k02 * H_2 + k03 * He_1 - k04 * He_2 + k05 * He_2 - k06 * He_3 -
            k07 * H_1 + k14 * H_m0 - k18 * H2_2;
  J[8][9] = (D.r01 - D.r02 + D.r03 - D.r04 + D.r05 - D.r06 - D.r07 + D.r08 + D.r14 + D.r15 +
             D.r17 - D.r18) * Tge;
}

__device__ inline void chem_jac_energy(const double (&y)[NSP], const double (&k)[NRATE],
                                       const double (&rk)[NRATE], double T, double z,
                                       double Tge, const CellAux &ax, double (&J)[NSP][NSP]) {
#pragma clang fp contract(off)
  const double H2_1 = y[0], H_1 = y[2], H_2 = y[3], He_1 = y[5], He_2 = y[6], He_3 = y[7],
               de = y[8];
  const double H1sq = H_1 * H_1;
  const double brem = k[C_BREM], ceHeI = k[C_CEHEI], ceHeII = k[C_CEHEII], ceHI = k[C_CEHI],
               cieco = k[C_CIECO], ciHeI = k[C_CIHEI], ciHeII = k[C_CIHEII],
               ciHeIS = k[C_CIHEIS], ciHI = k[C_CIHI], compton = k[C_COMPTON],
               gael = k[C_GAEL], gaH2 = k[C_GAH2], gaHe = k[C_GAHE], gaHI = k[C_GAHI],
               gaHp = k[C_GAHP], h2lte = k[C_H2LTE], h2mcool = k[C_H2MCOOL],
               h2mheat = k[C_H2MHEAT], ncrd1 = k[C_NCRD1], ncrd2 = k[C_NCRD2], ncrn = k[C_NCRN],
               reHeII1 = k[C_REHEII1], reHeII2 = k[C_REHEII2], reHeIII = k[C_REHEIII],
               reHII = k[C_REHII];
  const double inv_md = 1.0 / ax.mdensity;
  const double h2o = ax.h2o;
  const double S = H2_1 * gaH2 + H_1 * gaHI + H_2 * gaHp + He_1 * gaHe + de * gael;
  const double Dn = H2_1 * ncrd2 + H_1 * ncrd1;
  const double gden = h2lte / S + 1.0;
  const double nden = ncrn / Dn + 1.0;
  const double nden_m2 = 1.0 / (nden * nden);
  const double Y = (-H2_1) * H_1 * h2mcool + (H_1 * H_1 * H_1) * h2mheat;
  const double gq = (h2lte * h2lte) * h2o / ((gden * gden) * (S * S));
  const double zp1 = z + 1.0;
  const double zp4 = (zp1 * zp1) * (zp1 * zp1);
  J[9][0] = ((-H2_1) * gaH2 * gq - 0.5 * H_1 * h2mcool * 1.0 / nden -
             2.0158800000000001 * cieco * ax.mdensity - h2lte * h2o / gden +
             0.5 * ncrd2 * ncrn * nden_m2 * Y / (Dn * Dn)) * inv_md;
  J[9][2] = ((-H2_1) * gaHI * gq - ceHI * de - ciHI * de +
             0.5 * ncrd1 * ncrn * nden_m2 * Y / (Dn * Dn) +
             0.5 * ((-H2_1) * h2mcool + 3.0 * H1sq * h2mheat) * 1.0 / nden) * inv_md;
  J[9][3] = ((-H2_1) * gaHp * gq - brem * de - de * reHII) * inv_md;
  J[9][5] = ((-H2_1) * gaHe * gq - ciHeI * de) * inv_md;
  J[9][6] = ((-brem) * de - ceHeII * de - ceHeI * (de * de) - ciHeII * de -
             ciHeIS * (de * de) - de * reHeII1 - de * reHeII2) * inv_md;
  J[9][7] = (-4.0 * brem * de - de * reHeIII) * inv_md;
  J[9][8] = ((-H2_1) * gael * gq - H_1 * ceHI - H_1 * ciHI - H_2 * reHII - He_1 * ciHeI -
             He_2 * ceHeII - 2.0 * He_2 * ceHeI * de - He_2 * ciHeII -
             2.0 * He_2 * ciHeIS * de - He_2 * reHeII1 - He_2 * reHeII2 - He_3 * reHeIII -
             brem * (H_2 + He_2 + 4.0 * He_3) - compton * zp4 * (T - 2.73 * z - 2.73)) * inv_md;
  const double Rg = (-H2_1) * rk[C_GAH2] - H_1 * rk[C_GAHI] - H_2 * rk[C_GAHP] -
                    He_1 * rk[C_GAHE] - de * rk[C_GAEL];
  const double adhoc =
      (-H2_1) * h2lte * h2o * ((-h2lte) * Rg / (S * S) - rk[C_H2LTE] / S) / (gden * gden) -
      H2_1 * h2o * rk[C_H2LTE] / gden +
      0.5 * nden_m2 * Y *
          (-1.0 * ncrn * ((-H2_1) * rk[C_NCRD2] - H_1 * rk[C_NCRD1]) / (Dn * Dn) -
           1.0 * rk[C_NCRN] / Dn) +
      0.5 * 1.0 / nden * ((-H2_1) * H_1 * rk[C_H2MCOOL] + (H_1 * H_1 * H_1) * rk[C_H2MHEAT]);
  J[9][9] = adhoc * inv_md * Tge;
}

__device__ inline void chem_jac_physical(const double (&y)[NSP], const double (&k)[NRATE],
                                         const double (&rk)[NRATE], double T, double z,
                                         double Tge, const CellAux &ax,
                                         double (&J)[NSP][NSP]) {
#pragma unroll
  for (int a = 0; a < NSP; a++)
#pragma unroll
    for (int b = 0; b < NSP; b++) J[a][b] = 0.0;
  chem_jac_species(y, k, rk, Tge, J);
  chem_jac_energy(y, k, rk, T, z, Tge, ax, J);
}

}  // namespace mvd
