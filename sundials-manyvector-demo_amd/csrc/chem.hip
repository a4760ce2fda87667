// Per-call primordial chemistry kernels: setup, prepare (Dengo scaling + auxiliary
// per-cell values), RHS and analytic Jacobian. One thread per cell.
#include <cmath>
#include "mvd_chem.hpp"
#include "../../include/mvd_rhs.h"

namespace mvd {

// cvklu_setup_data initial values (raja_primordial_network.cpp:186-188,
// primordial_blast.cpp:337-341)
__global__ void __launch_bounds__(256) chem_init(long n, double *Ts, double *dTs) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { Ts[i] = 1000.0; dTs[i] = 0.0; }
}

// The bin-major pair copy of the rate tables (ChemConst::tabp).
__global__ void __launch_bounds__(256) chem_pair_tables(const double *tab, double2 *tabp,
                                                        int nbins) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)nbins * NRATE) return;
  const int b = (int)(e / NRATE), r = (int)(e % NRATE);
  tabp[e] = make_double2(tab[r * TLEN + b], tab[r * TLEN + b + 1]);
}

__global__ void __launch_bounds__(256) chem_init_scale(long m, double *scale, double *iscale) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) { scale[i] = 1.0; iscale[i] = 1.0; }
}

// prepare_Dengo_structures (primordial_blast.cpp:368-400) fused with
// setting_up_extra_variables (raja_primordial_network.cpp:1969-2018).
__global__ void __launch_bounds__(256)
chem_prepare_kernel(long n, double *chem, double *scale, double *iscale, double *md,
                    double *imd, double *cie, double *h2o) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s[NSP];
#pragma unroll
  for (int v = 0; v < NSP; v++) {
    s[v] = chem[NSP * i + v];
    scale[NSP * i + v] = s[v];
    iscale[NSP * i + v] = 1.0 / s[v];
    chem[NSP * i + v] = 1.0;
  }
  double m = 0.0;
  m += s[0] * 2.0;
  m += s[1] * 2.0;
  m += s[2] * 1.00794;
  m += s[3] * 1.00794;
  m += s[4] * 1.00794;
  m += s[5] * 4.002602;
  m += s[6] * 4.002602;
  m += s[7] * 4.002602;
  m *= 1.67e-24;
  double tau = pow(m / 3.3e-8, 2.8);
  tau = fmax(tau, 1.0e-5);
  md[i] = m;
  imd[i] = 1.0 / m;
  cie[i] = fmin(1.0, (1.0 - exp(-tau)) / tau);
  h2o[i] = fmin(1.0, pow(m / (1.34e-14), -0.45));
}

struct ChemPtrs {
  const double *scale, *iscale, *md, *imd, *cie, *h2o;
  double *Ts, *dTs;
};

// calculate_rhs_cvklu (raja_primordial_network.cpp:1015-1389)
__global__ void __launch_bounds__(256)
chem_rhs_kernel(long n, ChemConst cc, ChemPtrs p, const double *yd, double *ydot, double srhs) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double y[NSP];
#pragma unroll
  for (int v = 0; v < NSP; v++) y[v] = yd[NSP * i + v] * p.scale[NSP * i + v];
  double Ts = p.Ts[i], dTs;
  chem_temperature(cc, y, Ts, dTs);
  p.Ts[i] = Ts;
  p.dTs[i] = dTs;
  double k[NRATE];
  chem_rates_only_pairs<0, NRATE>(cc, Ts, k);
  const CellAux ax{p.md[i], p.imd[i], p.cie[i], p.h2o[i]};
  double f[NSP];
  chem_rhs_physical(y, k, Ts, cc.z, ax, f);
#pragma unroll
  for (int v = 0; v < NSP; v++) f[v] *= p.iscale[NSP * i + v];
  f[9] *= ax.inv_mdensity;
  // times the caller's factor (TimeUnits for ffast/fimpl, N_VScale after the RHS; 1 exact)
#pragma unroll
  for (int v = 0; v < NSP; v++) ydot[NSP * i + v] = srhs * f[v];
}

// Vector-Jacobian product of calculate_rhs_cvklu in the scaled variables:
// ybar = (d ydot / d y)^T v, with the temperature a function of y (solved from the warm start
// like the RHS, which it also updates); optionally ydot itself. The rates come from the
// rate-major tables, all of them in one batch of loads: with one wave per SIMD the kernel
// is bound by the latency of these gathers, and for cells of unrelated temperatures this
// form measured faster than the pair table.
__global__ void __launch_bounds__(256)
chem_vjp_kernel(long n, ChemConst cc, ChemPtrs p, const double *yd, const double *v,
                double *ybar, double *ydot) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double y[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) y[s] = yd[NSP * i + s] * p.scale[NSP * i + s];
  double Ts = p.Ts[i], dTs;
  chem_temperature(cc, y, Ts, dTs);
  p.Ts[i] = Ts;
  p.dTs[i] = dTs;
  const CellAux ax{p.md[i], p.imd[i], p.cie[i], p.h2o[i]};
  if (ydot) {
    double k[NRATE], f[NSP];
    chem_rates_only(cc, Ts, k);
    chem_rhs_physical(y, k, Ts, cc.z, ax, f);
#pragma unroll
    for (int s = 0; s < NSP; s++) f[s] *= p.iscale[NSP * i + s];
    f[9] *= ax.inv_mdensity;
#pragma unroll
    for (int s = 0; s < NSP; s++) ydot[NSP * i + s] = f[s];
  }
  double vp[NSP], yb[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) vp[s] = v[NSP * i + s] * p.iscale[NSP * i + s];
  vp[9] *= ax.inv_mdensity;
  const bool clamped = !(Ts > cc.tmin && Ts < cc.tmax);
  chem_vjp_physical<false>(cc, y, Ts, clamped, ax, vp, yb);
#pragma unroll
  for (int s = 0; s < NSP; s++) ybar[NSP * i + s] = yb[s] * p.scale[NSP * i + s];
}

// Self-check of the temperature solve's early exit: the early form and the ten steps from
// the same warm start Ts0 for every cell of y (scaled); counts the cells whose (T, dTs_ge)
// differ bitwise. The per-cell state (Ts, dTs_ge) is not touched.
__global__ void __launch_bounds__(256)
chem_temperature_check_kernel(long n, ChemConst cc, const double *scale, const double *yd,
                              const double *Ts0, int *bad) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double y[NSP];
#pragma unroll
  for (int v = 0; v < NSP; v++) y[v] = yd[NSP * i + v] * scale[NSP * i + v];
  double Ta = Ts0[i], da, Tb = Ts0[i], db;
  chem_temperature(cc, y, Ta, da);
  chem_temperature_ten(cc, y, Tb, db);
  if (__double_as_longlong(Ta) != __double_as_longlong(Tb) ||
      __double_as_longlong(da) != __double_as_longlong(db))
    atomicAdd(bad, 1);
}

// calculate_jacobian_cvklu (raja_primordial_network.cpp:1393-1965): dense column-major
// blocks, each entry times inv_scale[row]*scale[col].
__global__ void __launch_bounds__(256)
chem_jac_kernel(long n, ChemConst cc, ChemPtrs p, const double *yd, double *J, double sjac) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double y[NSP], sc[NSP], isc[NSP];
#pragma unroll
  for (int v = 0; v < NSP; v++) {
    sc[v] = p.scale[NSP * i + v];
    isc[v] = p.iscale[NSP * i + v];
    y[v] = yd[NSP * i + v] * sc[v];
  }
  const double T = p.Ts[i];
  double k[NRATE], dk[NRATE];
  chem_rates_pairs<0, NRATE>(cc, T, k, dk);
  const CellAux ax{p.md[i], p.imd[i], p.cie[i], p.h2o[i]};
  double Jp[NSP][NSP];
  chem_jac_physical(y, k, dk, T, cc.z, p.dTs[i], ax, Jp);
  double *blk = J + (long)NSP * NSP * i;
#pragma unroll
  for (int c = 0; c < NSP; c++)
#pragma unroll
    for (int r = 0; r < NSP; r++) blk[NSP * c + r] = Jp[r][c] * (isc[r] * sc[c]) * sjac;
}

static ChemPtrs ptrs_of(ChemState &s) {
  return ChemPtrs{s.scale, s.inv_scale, s.mdensity, s.inv_mdensity, s.cie_oda, s.h2_oda,
                  s.Ts, s.dTs_ge};
}

static inline unsigned nb(long n) { return (unsigned)((n + 255) / 256); }

void chem_free(ChemState &s) {
  double *arrs[] = {s.tab, s.scale, s.inv_scale, s.Ts, s.dTs_ge, s.mdensity, s.inv_mdensity,
                    s.cie_oda, s.h2_oda, s.work};
  for (double *a : arrs)
    if (a) (void)hipFree(a);
  if (s.tabp) (void)hipFree(s.tabp);
  if (s.counters) (void)hipFree(s.counters);
  if (s.host_counters) (void)hipHostFree(s.host_counters);
  s = ChemState{};
}

int chem_setup(ChemState &s, const double *tables_host, long n, double z, hipStream_t st) {
  const unsigned long gen = s.gen;
  chem_free(s);
  s.gen = gen + 1;
  bool ok = hipMalloc(&s.tab, sizeof(double) * NTAB * TLEN) == hipSuccess &&
            hipMalloc(&s.tabp, sizeof(double2) * (TLEN - 1) * NRATE) == hipSuccess &&
            hipMalloc(&s.scale, sizeof(double) * NSP * n) == hipSuccess &&
            hipMalloc(&s.inv_scale, sizeof(double) * NSP * n) == hipSuccess &&
            hipMalloc(&s.Ts, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&s.dTs_ge, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&s.mdensity, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&s.inv_mdensity, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&s.cie_oda, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&s.h2_oda, sizeof(double) * n) == hipSuccess &&
            hipMalloc(&s.counters, sizeof(int) * 64) == hipSuccess &&
            hipHostMalloc(&s.host_counters, sizeof(int) * 64, hipHostMallocDefault) == hipSuccess;
  if (!ok) { chem_free(s); s.gen = gen + 1; return MVD_DEVICE_ERROR; }
  if (hipMemcpyAsync(s.tab, tables_host, sizeof(double) * NTAB * TLEN, hipMemcpyHostToDevice,
                     st) != hipSuccess) {
    chem_free(s);
    return MVD_DEVICE_ERROR;
  }
  s.n = n;
  s.cc.tab = s.tab;
  s.cc.tmin = 1.0;
  s.cc.tmax = 100000.0;
  s.cc.nbins = TLEN - 1;
  s.cc.dbin = (std::log(s.cc.tmax) - std::log(s.cc.tmin)) / s.cc.nbins;
  s.cc.idbin = (double)(1.0L / (long double)s.cc.dbin);   // long-double quotient as the reference
  s.cc.lb = std::log(s.cc.tmin);
  s.cc.tabp = s.tabp;
  hipLaunchKernelGGL(chem_pair_tables, dim3(nb((long)s.cc.nbins * NRATE)), dim3(256), 0, st,
                     (const double *)s.tab, s.tabp, s.cc.nbins);
  s.cc.z = z;
  hipLaunchKernelGGL(chem_init, dim3(nb(n)), dim3(256), 0, st, n, s.Ts, s.dTs_ge);
  hipLaunchKernelGGL(chem_init_scale, dim3(nb(n * NSP)), dim3(256), 0, st, n * NSP, s.scale,
                     s.inv_scale);
  if (hipStreamSynchronize(st) != hipSuccess || hipGetLastError() != hipSuccess) {
    chem_free(s);
    return MVD_DEVICE_ERROR;
  }
  s.ready = true;
  return MVD_SUCCESS;
}

int chem_prepare(ChemState &s, double *chem, hipStream_t st) {
  hipLaunchKernelGGL(chem_prepare_kernel, dim3(nb(s.n)), dim3(256), 0, st, s.n, chem, s.scale,
                     s.inv_scale, s.mdensity, s.inv_mdensity, s.cie_oda, s.h2_oda);
  if (hipGetLastError() != hipSuccess) return MVD_DEVICE_ERROR;
  return hipStreamSynchronize(st) == hipSuccess ? MVD_SUCCESS : MVD_DEVICE_ERROR;
}

int chem_rhs(ChemState &s, const double *y, double *ydot, hipStream_t st, bool sync,
             double srhs) {
  hipLaunchKernelGGL(chem_rhs_kernel, dim3(nb(s.n)), dim3(256), 0, st, s.n, s.cc, ptrs_of(s), y,
                     ydot, srhs);
  if (hipGetLastError() != hipSuccess) return MVD_DEVICE_ERROR;
  if (!sync) return MVD_SUCCESS;
  return hipStreamSynchronize(st) == hipSuccess ? MVD_SUCCESS : MVD_DEVICE_ERROR;
}

int chem_jac(ChemState &s, const double *y, double *J, hipStream_t st, bool sync,
             double sjac) {
  hipLaunchKernelGGL(chem_jac_kernel, dim3(nb(s.n)), dim3(256), 0, st, s.n, s.cc, ptrs_of(s), y,
                     J, sjac);
  if (hipGetLastError() != hipSuccess) return MVD_DEVICE_ERROR;
  if (!sync) return MVD_SUCCESS;
  return hipStreamSynchronize(st) == hipSuccess ? MVD_SUCCESS : MVD_DEVICE_ERROR;
}

int chem_vjp(ChemState &s, const double *y, const double *v, double *ybar, double *ydot,
             hipStream_t st, bool sync) {
  if (!s.ready) return MVD_NOT_READY;
  hipLaunchKernelGGL(chem_vjp_kernel, dim3(nb(s.n)), dim3(256), 0, st, s.n, s.cc, ptrs_of(s), y,
                     v, ybar, ydot);
  if (hipGetLastError() != hipSuccess) return MVD_DEVICE_ERROR;
  if (sync && hipStreamSynchronize(st) != hipSuccess) return MVD_DEVICE_ERROR;
  return MVD_SUCCESS;
}

int chem_temperature_check(ChemState &s, const double *y, const double *Ts0, long *bad,
                           hipStream_t st) {
  if (!s.ready) return MVD_NOT_READY;
  int *d = s.counters + 63;      // last word of the 64: the DIRK uses the first 1 + 20
  int h = 0;
  if (hipMemsetAsync(d, 0, sizeof(int), st) != hipSuccess) return MVD_DEVICE_ERROR;
  hipLaunchKernelGGL(chem_temperature_check_kernel, dim3(nb(s.n)), dim3(256), 0, st, s.n, s.cc,
                     (const double *)s.scale, y, Ts0, d);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(&h, d, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return MVD_DEVICE_ERROR;
  *bad = h;
  return MVD_SUCCESS;
}

double *chem_state_ptr(ChemState &s, int which) {
  switch (which) {
    case 0: return s.scale;
    case 1: return s.inv_scale;
    case 2: return s.Ts;
    case 3: return s.dTs_ge;
    case 4: return s.mdensity;
    case 5: return s.inv_mdensity;
    case 6: return s.cie_oda;
    case 7: return s.h2_oda;
  }
  return nullptr;
}

}  // namespace mvd
