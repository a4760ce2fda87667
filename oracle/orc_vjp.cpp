// Oracle: vector-Jacobian product of fEuler -- TEST INFRASTRUCTURE ONLY (see orc_common.hpp).
//
// fe(w) = -div F(w): every face flux is a function of its 6-point stencil (interior cells
// and ghost values, which are +-copies of interior cells). Here the face Jacobian comes
// from forward-mode differentiation of the oracle's face_flux (the restatement of
// utilities.cpp:270-479) with one stencil point's NV inputs as the tangents per pass, so
// it is independent of the hand-written reverse sweep of the GPU. The cotangent on wdot
// becomes Fbar = (v_right - v_left)/delta per face (cells outside the grid contribute
// nothing), and each point's share goes back to its source cell with the ghost rule's sign.
#include <cmath>
#include <cstring>
#include <vector>
#include "orc_fe.hpp"
#include "oracle.h"

namespace orc {

template <int N>
struct DualT {
  double v;
  double d[N];
  DualT() : v(0) { for (double &x : d) x = 0; }
  DualT(double x) : v(x) { for (double &q : d) q = 0; }
};
template <int N> inline DualT<N> operator+(const DualT<N> &a, const DualT<N> &b) {
  DualT<N> r(a.v + b.v);
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N> inline DualT<N> operator-(const DualT<N> &a, const DualT<N> &b) {
  DualT<N> r(a.v - b.v);
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N> inline DualT<N> operator-(const DualT<N> &a) {
  DualT<N> r(-a.v);
  for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
  return r;
}
template <int N> inline DualT<N> operator*(const DualT<N> &a, const DualT<N> &b) {
  DualT<N> r(a.v * b.v);
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N> inline DualT<N> operator/(const DualT<N> &a, const DualT<N> &b) {
  DualT<N> r(a.v / b.v);
  for (int i = 0; i < N; i++) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
  return r;
}
template <int N> inline DualT<N> &operator+=(DualT<N> &a, const DualT<N> &b) { return a = a + b; }
template <int N> inline bool operator<(const DualT<N> &a, const DualT<N> &b) { return a.v < b.v; }
template <int N> inline bool operator<=(const DualT<N> &a, const DualT<N> &b) { return a.v <= b.v; }
template <int N> inline bool operator>(const DualT<N> &a, double b) { return a.v > b; }
template <int N> inline DualT<N> sqrt(const DualT<N> &a) {
  DualT<N> r(std::sqrt(a.v));
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] / (2.0 * r.v);
  return r;
}
template <int N> inline DualT<N> fabs(const DualT<N> &a) { return a.v < 0 ? -a : a; }
template <int N> inline DualT<N> sq(const DualT<N> &a) { return a * a; }

// source cell and sign of field v at extended index e along d of the line through (i,j,k)
// (the ghost rules of field_at in orc_fe.cpp)
static void source_of(const orc_grid *g, int v, int d, long i, long j, long k, long e,
                      long *cell, double *sign) {
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
  *cell = ii + g->nx * (jj + g->ny * kk);
  *sign = 1.0;
  if (bc == BC_DIRICHLET) *sign = -1.0;
  else if (bc == BC_REFLECTING && v == 1 + d) *sign = -1.0;
}

}  // namespace orc

using namespace orc;

// sbar[6][nv] = (d F_face / d stencil)^T Fbar for one face (row-major stencil, global
// momentum order), by forward mode over each point.
extern "C" int orc_face_flux_vjp(const double *stencil, int dir, int nchem, double gamma,
                                 const double *Fbar, double *sbar) {
  const int nv = 5 + nchem;
  for (int pt = 0; pt < 6; pt++) {
    DualT<NVMAX> s[6][NVMAX], f[NVMAX];
    for (int l = 0; l < 6; l++)
      for (int c = 0; c < nv; c++) {
        s[l][c] = DualT<NVMAX>(stencil[l * nv + c]);
        if (l == pt) s[l][c].d[c] = 1.0;
      }
    face_flux(s, dir, f, DualT<NVMAX>(gamma), nchem);
    for (int c = 0; c < nv; c++) {
      double acc = 0.0;
      for (int a = 0; a < nv; a++) acc += Fbar[a] * f[a].d[c];
      sbar[pt * nv + c] = acc;
    }
  }
  return 0;
}

// wbar = (d fe / d w)^T v for fEuler (utilities.cpp:17-253), v and wbar in the ManyVector
// layout of w. Serial (the scatter to source cells is not partitioned).
extern "C" int orc_fe_vjp(const orc_grid *g, const double *const *w, const double *const *v,
                          double *const *wbar) {
  const long nx = g->nx, ny = g->ny, nz = g->nz, ncell = nx * ny * nz;
  const int nchem = g->nchem, nv = 5 + nchem;
  for (int q = 0; q < 5; q++) std::memset(wbar[q], 0, sizeof(double) * ncell);
  if (nchem > 0) std::memset(wbar[5], 0, sizeof(double) * ncell * nchem);
  auto comp = [&](const double *const *a, int q, long cell) -> double {
    return q < 5 ? a[q][cell] : a[5][(long)nchem * cell + (q - 5)];
  };
  auto comp_add = [&](int q, long cell, double x) {
    if (q < 5) wbar[q][cell] += x;
    else wbar[5][(long)nchem * cell + (q - 5)] += x;
  };
  const double delta[3] = {g->dx, g->dy, g->dz};
  const long n3[3] = {nx, ny, nz};
  std::vector<double> st(6 * nv), sb(6 * nv), Fb(nv);
  for (int d = 0; d < 3; d++) {
    for (long k = 0; k < nz + (d == 2); k++)
      for (long j = 0; j < ny + (d == 1); j++)
        for (long i = 0; i < nx + (d == 0); i++) {
          const long c = (d == 0) ? i : (d == 1) ? j : k;     // face c: cells c-1 | c
          // cotangent of the face flux: wdot_c has +F/delta, wdot_{c-1} has -F/delta
          for (int q = 0; q < nv; q++) {
            double r = 0.0, l = 0.0;
            if (c < n3[d]) r = comp(v, q, i + nx * (j + ny * k));
            if (c > 0) {
              const long ci = (d == 0) ? i - 1 : i, cj = (d == 1) ? j - 1 : j,
                         ck = (d == 2) ? k - 1 : k;
              l = comp(v, q, ci + nx * (cj + ny * ck));
            }
            Fb[q] = (r - l) / delta[d];
          }
          for (int l = 0; l < 6; l++)
            for (int q = 0; q < nv; q++) {
              long cell;
              double sign;
              source_of(g, q, d, i, j, k, c - 3 + l, &cell, &sign);
              st[l * nv + q] = sign * comp(w, q, cell);
            }
          orc_face_flux_vjp(st.data(), d, nchem, g->gamma, Fb.data(), sb.data());
          for (int l = 0; l < 6; l++)
            for (int q = 0; q < nv; q++) {
              long cell;
              double sign;
              source_of(g, q, d, i, j, k, c - 3 + l, &cell, &sign);
              comp_add(q, cell, sign * sb[l * nv + q]);
            }
        }
  }
  return 0;
}
