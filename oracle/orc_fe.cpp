// Oracle drivers for fe(), stability and fslow -- TEST INFRASTRUCTURE ONLY.
// See orc_common.hpp for the role of this directory.
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "orc_fe.hpp"
#include "oracle.h"

namespace orc {
uint64_t OpCounter::add = 0, OpCounter::mul = 0, OpCounter::div = 0,
         OpCounter::sqrt_ = 0, OpCounter::cmp = 0, OpCounter::trans = 0;

// Value of field v at extended index e (may lie in the 3-deep ghost layer) along
// direction d of the line through (i,j,k). Ghost rules restate the physical-boundary
// fill of EulerData::ExchangeStart (euler3D.hpp:793-1166) as read back by
// pack1D_*_bdry (euler3D.hpp:1250-1378), and the single-rank periodic self-exchange
// (euler3D.hpp:607-786):
//   low side,  e in {-3,-2,-1}:  periodic -> w[n+e];     physical -> w[-1-e]   (mirror)
//   high side, e in {n..n+2}:    periodic -> w[e-n];     physical -> w[e-3]    (shifted copy)
//   reflecting negates the face-normal momentum, Dirichlet negates every field.
static double field_at(const orc_grid *g, const double *const *w, int v, int d,
                       long i, long j, long k, long e) {
  const long n = (d == 0) ? g->nx : (d == 1) ? g->ny : g->nz;
  long src = e;
  int bc = -1;
  if (e < 0) {
    bc = g->bc[2 * d];
    src = (bc == BC_PERIODIC) ? n + e : -1 - e;
  } else if (e >= n) {
    bc = g->bc[2 * d + 1];
    src = (bc == BC_PERIODIC) ? e - n : e - 3;
  }
  long ii = i, jj = j, kk = k;
  if (d == 0) ii = src; else if (d == 1) jj = src; else kk = src;
  const long cell = ii + g->nx * (jj + g->ny * kk);
  double val = (v < 5) ? w[v][cell] : w[5][(long)g->nchem * cell + (v - 5)];
  if (bc == BC_DIRICHLET) val = -val;
  else if (bc == BC_REFLECTING && v == 1 + d) val = -val;
  return val;
}

// pack1D_{x,y,z}[_bdry]: stencil w(c-3..c+2) for the lower face of index c along d.
static void gather(const orc_grid *g, const double *const *w, int d, long i, long j, long k,
                   double (&s)[6][NVMAX]) {
  const long c = (d == 0) ? i : (d == 1) ? j : k;
  const int nv = 5 + g->nchem;
  for (int l = 0; l < 6; l++)
    for (int v = 0; v < nv; v++) s[l][v] = field_at(g, w, v, d, i, j, k, c - 3 + l);
}

}  // namespace orc

using namespace orc;

extern "C" void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

extern "C" int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

extern "C" int orc_face_flux(const double *stencil, int dir, int nchem, double gamma,
                             double *fface) {
  double s[6][NVMAX];
  const int nv = 5 + nchem;
  for (int l = 0; l < 6; l++)
    for (int v = 0; v < nv; v++) s[l][v] = stencil[l * nv + v];
  face_flux(s, dir, fface, gamma, nchem);
  return 0;
}

// The 6 x (5+nchem) stencil the oracle feeds face_flux for the lower face of (i,j,k)
// along direction dir (row-major [point][field], global momentum order).
extern "C" void orc_gather(const orc_grid *g, const double *const *w, int dir, long i, long j,
                           long k, double *out) {
  double s[6][NVMAX];
  gather(g, w, dir, i, j, k, s);
  const int nv = 5 + g->nchem;
  for (int l = 0; l < 6; l++)
    for (int v = 0; v < nv; v++) out[l * nv + v] = s[l][v];
}

// EulerData::ExchangeStart/ExchangeEnd (euler3D.hpp:577-1191) for the single rank that owns
// the grid: receive buffers W, E, S, N, B, F (null entries skipped) in the layout
// BUFINDX(v,s,a,b,NVAR,3,na,nb) = v + NVAR*(s + 3*(a + na*b)) with (a, b) = (j, k) for W/E,
// (i, k) for S/N, (i, j) for B/F; slot s holds the point at extended index s-3 (low faces)
// or n+s (high faces), by the ghost rules of field_at.
extern "C" void orc_exchange(const orc_grid *g, const double *const *w, double *const *recv) {
  const int nv = 5 + g->nchem;
  const long n[3] = {g->nx, g->ny, g->nz};
  for (int face = 0; face < 6; face++) {
    if (!recv[face]) continue;
    const int d = face / 2;
    const bool high = face % 2 == 1;
    const long na = d == 0 ? g->ny : g->nx, nb = d == 2 ? g->ny : g->nz;
    for (long b = 0; b < nb; b++)
      for (long a = 0; a < na; a++)
        for (int s = 0; s < 3; s++) {
          const long e = high ? n[d] + s : s - 3;
          const long i = d == 0 ? 0 : a, j = d == 0 ? a : (d == 1 ? 0 : b),
                     k = d == 2 ? 0 : b;
          for (int v = 0; v < nv; v++)
            recv[face][v + nv * (s + 3 * (a + na * b))] = field_at(g, w, v, d, i, j, k, e);
        }
  }
}

// fEuler (utilities.cpp:17-253) with zero external forcing (the forcing used by every
// configuration here: primordial_blast.cpp:312-317, sod.cpp:182-210, fluid_blast).
// Fluxes are stored per face in the reference's AoS-by-component layout and the
// divergence is accumulated in the reference's order.
extern "C" int orc_fe(const orc_grid *g, const double *const *w, double *const *wdot) {
  const long nx = g->nx, ny = g->ny, nz = g->nz;
  const int nchem = g->nchem, nv = 5 + nchem;
  const long ncell = nx * ny * nz;
  for (int v = 0; v < 5; v++) std::memset(wdot[v], 0, sizeof(double) * ncell);
  if (nchem > 0) std::memset(wdot[5], 0, sizeof(double) * ncell * nchem);

  for (long c = 0; c < ncell; c++)
    if (legal_state(w[0][c], w[1][c], w[2][c], w[3][c], w[4][c], g->gamma) != 0) return -1;

  std::vector<double> xf((size_t)nv * (nx + 1) * ny * nz), yf((size_t)nv * nx * (ny + 1) * nz),
      zf((size_t)nv * nx * ny * (nz + 1));
  // Each face flux is a pure function of its stencil, so the cells may be visited in any
  // order; OpenMP threads only change the wall time, never the result.
#pragma omp parallel for collapse(2) schedule(static)
  for (long k = 0; k < nz; k++)
    for (long j = 0; j < ny; j++)
      for (long i = 0; i < nx; i++) {
        double s[6][NVMAX];
        gather(g, w, 0, i, j, k, s);
        face_flux(s, 0, &xf[(size_t)nv * (i + (nx + 1) * (j + ny * k))], g->gamma, nchem);
        if (i == nx - 1) {
          gather(g, w, 0, nx, j, k, s);
          face_flux(s, 0, &xf[(size_t)nv * (nx + (nx + 1) * (j + ny * k))], g->gamma, nchem);
        }
        gather(g, w, 1, i, j, k, s);
        face_flux(s, 1, &yf[(size_t)nv * (i + nx * (j + (ny + 1) * k))], g->gamma, nchem);
        if (j == ny - 1) {
          gather(g, w, 1, i, ny, k, s);
          face_flux(s, 1, &yf[(size_t)nv * (i + nx * (ny + (ny + 1) * k))], g->gamma, nchem);
        }
        gather(g, w, 2, i, j, k, s);
        face_flux(s, 2, &zf[(size_t)nv * (i + nx * (j + ny * k))], g->gamma, nchem);
        if (k == nz - 1) {
          gather(g, w, 2, i, j, nz, s);
          face_flux(s, 2, &zf[(size_t)nv * (i + nx * (j + ny * nz))], g->gamma, nchem);
        }
      }

  // divergence (utilities.cpp:198-245): wdot -= (dFx/dx + dFy/dy) + dFz/dz
#pragma omp parallel for collapse(2) schedule(static)
  for (long k = 0; k < nz; k++)
    for (long j = 0; j < ny; j++)
      for (long i = 0; i < nx; i++) {
        const long c = i + nx * (j + ny * k);
        const size_t x0 = (size_t)nv * (i + (nx + 1) * (j + ny * k)), x1 = x0 + nv;
        const size_t y0 = (size_t)nv * (i + nx * (j + (ny + 1) * k)), y1 = y0 + (size_t)nv * nx;
        const size_t z0 = (size_t)nv * (i + nx * (j + ny * k)), z1 = z0 + (size_t)nv * nx * ny;
        for (int v = 0; v < nv; v++) {
          const double div = (xf[x1 + v] - xf[x0 + v]) / g->dx + (yf[y1 + v] - yf[y0 + v]) / g->dy
                           + (zf[z1 + v] - zf[z0 + v]) / g->dz;
          if (v < 5) wdot[v][c] -= div;
          else wdot[5][(long)nchem * c + (v - 5)] -= div;
        }
      }
  return 0;
}

// stability (utilities.cpp:483-528) on one rank: returns the maximum wave speed
// alpha (with the reference's |mx/rho| speed, line 509) and the CFL step.
extern "C" double orc_max_wave_speed(const orc_grid *g, const double *const *w) {
  const long ncell = g->nx * g->ny * g->nz;
  double alpha = 0.0;
  for (long c = 0; c < ncell; c++) {
    const double rho = w[0][c], mx = w[1][c];
    const double u = std::fabs(mx / rho);   // max(|mx/rho|,|mx/rho|,|mx/rho|)
    const double p = eos(rho, mx, w[2][c], w[3][c], w[4][c], g->gamma);
    const double cs = sun_sqrt(g->gamma * p / rho);
    alpha = rmax(alpha, std::fabs(u + cs));
  }
  return alpha;
}

extern "C" double orc_stability_dt(const orc_grid *g, const double *const *w, double cfl) {
  const double alpha = orc_max_wave_speed(g, w);
  return cfl * std::min(std::min(g->dx, g->dy), g->dz) / alpha;
}

// Per-face operation counts of face_flux for nchem tracers, from the op-counting
// scalar instantiation (out: add, mul, div, sqrt, cmp).
extern "C" void orc_face_flux_opcount(int nchem, unsigned long long *out) {
  Cnt s[6][NVMAX];
  for (int l = 0; l < 6; l++) {
    s[l][0] = Cnt(1.0 + 0.01 * l); s[l][1] = Cnt(0.1); s[l][2] = Cnt(0.2);
    s[l][3] = Cnt(0.3); s[l][4] = Cnt(3.0);
    for (int c = 0; c < nchem; c++) s[l][5 + c] = Cnt(0.5);
  }
  Cnt f[NVMAX];
  OpCounter::reset();
  face_flux(s, 0, f, Cnt(1.4), nchem);
  out[0] = OpCounter::add; out[1] = OpCounter::mul; out[2] = OpCounter::div;
  out[3] = OpCounter::sqrt_; out[4] = OpCounter::cmp;
}
