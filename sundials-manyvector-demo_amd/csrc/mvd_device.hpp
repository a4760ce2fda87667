// Device-side building blocks of the WENO5 flux divergence fe() for CDNA4 (gfx950), FP64.
//
// Data layout in HBM (the reference's ManyVector layout, euler3D.hpp:62-65):
//   fluid fields rho, mx, my, mz, et: 5 separate arrays, cell index i + nx*(j + ny*k);
//   advected species: one AoS array, species fastest: v + nchem*(i + nx*(j + ny*k)).
//
// Two face-flux implementations live here:
//   face_flux_ref  - the reference recipe (utilities.cpp:270-479) operation by operation,
//                    with FMA contraction disabled; used by the plain comparator kernels
//                    and reproducible bit for bit against a non-FMA CPU build.
//   face_flux_lds  - the same mathematics rearranged for the FP64 pipes: per-point
//                    primitive values computed once per point (staged in LDS), 1/rho
//                    instead of repeated divides, the left eigenvector product factored
//                    into a handful of FMAs, the WENO weights over one common
//                    denominator and one divide per component for both split sides.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mvd {

constexpr int BC_PERIODIC = 0, BC_NEUMANN = 1, BC_DIRICHLET = 2, BC_REFLECTING = 3;
constexpr int NCHEM_MAX = 16;
constexpr int NV_MAX = 5 + NCHEM_MAX;

// Grid and boundary description passed by value to every fe() kernel.
struct Grid {
  int nx, ny, nz, nchem;
  int bc[6];               // xl, xr, yl, yr, zl, zr
  double gamma, dx, dy, dz;
};

__host__ __device__ inline int extent(const Grid &g, int d) {
  return d == 0 ? g.nx : (d == 1 ? g.ny : g.nz);
}

// Source position and sign of extended index e (ghost layer 3 deep) along direction d.
// Physical boundaries follow the reference's ghost fill (euler3D.hpp:793-1166):
// low faces mirror (e -> -1-e), high faces copy shifted (e -> e-3); periodic wraps.
// sgn_all = -1 for Dirichlet (all fields negated); sgn_norm = -1 for the normal momentum
// under reflecting (and Dirichlet) boundaries.
struct GhostMap {
  int src;
  double sgn_all, sgn_norm;
};

__device__ inline GhostMap ghost_map(const Grid &g, int d, int e) {
  const int n = extent(g, d);
  GhostMap m{e, 1.0, 1.0};
  if (e < 0 || e >= n) {
    const bool lo = e < 0;
    const int bc = g.bc[2 * d + (lo ? 0 : 1)];
    if (bc == BC_PERIODIC) {
      m.src = lo ? e + n : e - n;
    } else {
      m.src = lo ? -1 - e : e - 3;
      if (bc == BC_DIRICHLET) { m.sgn_all = -1.0; m.sgn_norm = -1.0; }
      else if (bc == BC_REFLECTING) { m.sgn_norm = -1.0; }
    }
  }
  return m;
}

// SUNRsqrt (sundials_math.h): non-positive arguments give 0.
__device__ inline double sun_sqrt(double x) { return (x <= 0.0) ? 0.0 : sqrt(x); }

// Reciprocal from the hardware estimate (v_rcp_f64) refined by two Newton steps, and a
// divide built on it with one residual correction; within 1-2 ulp of the IEEE result at
// roughly half the instruction count of the correctly rounded divide sequence.
__device__ inline double fast_rcp(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

__device__ inline double fast_div(double n, double d) {
  const double r = fast_rcp(d);
  const double q = n * r;
  return fma(fma(-d, q, n), r, q);
}

// ---------------------------------------------------------------------------------
// Reference-order face flux (no contraction). s[l][v], l = stencil point, v = field in
// direction-local order (rho, m_normal, m_t1, m_t2, et, species...).
// ---------------------------------------------------------------------------------
template <int NC>
__device__ inline void face_flux_ref(double (&s)[6][5 + NC], double gamma, double *ff_out) {
#pragma clang fp contract(off)
  constexpr int NV = 5 + NC;
  const double bc = 1.083333333333333333333333333333333333333;
  const double eps = 1e-6;
  const double c13 = 0.3333333333333333333333333333333333333333;
  const double c56 = 0.8333333333333333333333333333333333333333;
  const double c16 = 0.1666666666666666666666666666666666666667;
  const double c76 = 1.166666666666666666666666666666666666667;
  const double c116 = 1.833333333333333333333333333333333333333;
  double p[6];
#pragma unroll
  for (int l = 0; l < 6; l++)
    p[l] = (gamma - 1.0) * (s[l][4] - (s[l][1] * s[l][1] + s[l][2] * s[l][2] + s[l][3] * s[l][3]) * 0.5 / s[l][0]);
  const double sL = sun_sqrt(s[2][0]), sR = sun_sqrt(s[3][0]);
  const double sbar = 0.5 * (sL + sR);
  const double u = 0.5 * (s[2][1] / sL + s[3][1] / sR) / sbar;
  const double v = 0.5 * (s[2][2] / sL + s[3][2] / sR) / sbar;
  const double w = 0.5 * (s[2][3] / sL + s[3][3] / sR) / sbar;
  const double H = 0.5 * ((p[2] + s[2][4]) / sL + (p[3] + s[3][4]) / sR) / sbar;
  const double qsq = u * u + v * v + w * w;
  const double gm = gamma - 1.0;
  const double cs = gm * (H - 0.5 * qsq);
  const double ci = 1.0 / cs;
  const double RV[5][5] = {{1.0, 0.0, 0.0, 1.0, 1.0},
                           {u - cs, 0.0, 0.0, u, u + cs},
                           {v, 1.0, 0.0, v, v},
                           {w, 0.0, 1.0, w, w},
                           {H - u * cs, v, w, 0.5 * qsq, H + u * cs}};
  const double LV[5][5] = {
      {0.5 * ci * (u + 0.5 * gm * qsq), -0.5 * ci * (gm * u + 1.0), -0.5 * v * gm * ci,
       -0.5 * w * gm * ci, 0.5 * gm * ci},
      {-v, 0.0, 1.0, 0.0, 0.0},
      {-w, 0.0, 0.0, 1.0, 0.0},
      {(-gm) * ci * (qsq - H), u * gm * ci, v * gm * ci, w * gm * ci, (-gm) * ci},
      {-0.5 * ci * (u - 0.5 * gm * qsq), -0.5 * ci * (gm * u - 1.0), -0.5 * v * gm * ci,
       -0.5 * w * gm * ci, 0.5 * gm * ci}};
  double flux[6][NV];
  double alpha = 0.0;
#pragma unroll
  for (int l = 0; l < 6; l++) {
    const double ul = s[l][1] / s[l][0];
    flux[l][0] = s[l][1];
    flux[l][1] = ul * s[l][1] + p[l];
    flux[l][2] = ul * s[l][2];
    flux[l][3] = ul * s[l][3];
    flux[l][4] = ul * (s[l][4] + p[l]);
#pragma unroll
    for (int c = 0; c < NC; c++) flux[l][5 + c] = ul * s[l][5 + c];
    const double cl = sun_sqrt(gamma * p[l] / s[l][0]);
    const double al = fabs(ul) + cl;
    alpha = (alpha < al) ? al : alpha;
  }
  double ff[NV];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    double fp[5][NV];
#pragma unroll
    for (int l = 0; l < 5; l++) {
      double fs[NV];
#pragma unroll
      for (int c = 0; c < NV; c++)
        fs[c] = side == 0 ? 0.5 * (flux[l][c] + alpha * s[l][c])
                          : 0.5 * (flux[l + 1][c] - alpha * s[l + 1][c]);
#pragma unroll
      for (int a = 0; a < 5; a++)
        fp[l][a] = LV[a][0] * fs[0] + LV[a][1] * fs[1] + LV[a][2] * fs[2] + LV[a][3] * fs[3] + LV[a][4] * fs[4];
#pragma unroll
      for (int c = 0; c < NC; c++) fp[l][5 + c] = fs[5 + c];
    }
#pragma unroll
    for (int c = 0; c < NV; c++) {
      const double f0 = fp[0][c], f1 = fp[1][c], f2 = fp[2][c], f3 = fp[3][c], f4 = fp[4][c];
      const double x1 = f2 - 2.0 * f3 + f4, y1 = 3.0 * f2 - 4.0 * f3 + f4;
      const double x2 = f1 - 2.0 * f2 + f3, y2 = f1 - f3;
      const double x3 = f0 - 2.0 * f1 + f2, y3 = f0 - 4.0 * f1 + 3.0 * f2;
      const double b1 = bc * (x1 * x1) + 0.25 * (y1 * y1);
      const double b2 = bc * (x2 * x2) + 0.25 * (y2 * y2);
      const double b3 = bc * (x3 * x3) + 0.25 * (y3 * y3);
      if (side == 0) {
        const double w1 = 0.3 / ((eps + b1) * (eps + b1));
        const double w2 = 0.6 / ((eps + b2) * (eps + b2));
        const double w3 = 0.1 / ((eps + b3) * (eps + b3));
        const double q1 = c13 * f2 + c56 * f3 - c16 * f4;
        const double q2 = (-c16) * f1 + c56 * f2 + c13 * f3;
        const double q3 = c13 * f0 - c76 * f1 + c116 * f2;
        ff[c] = (q1 * w1 + q2 * w2 + q3 * w3) / (w1 + w2 + w3);
      } else {
        const double w1 = 0.1 / ((eps + b1) * (eps + b1));
        const double w2 = 0.6 / ((eps + b2) * (eps + b2));
        const double w3 = 0.3 / ((eps + b3) * (eps + b3));
        const double q1 = c116 * f2 - c76 * f3 + c13 * f4;
        const double q2 = c13 * f1 + c56 * f2 - c16 * f3;
        const double q3 = (-c16) * f0 + c56 * f1 + c13 * f2;
        ff[c] += (q1 * w1 + q2 * w2 + q3 * w3) / (w1 + w2 + w3);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 5; a++)
    ff_out[a] = RV[a][0] * ff[0] + RV[a][1] * ff[1] + RV[a][2] * ff[2] + RV[a][3] * ff[3] + RV[a][4] * ff[4];
#pragma unroll
  for (int c = 0; c < NC; c++) ff_out[5 + c] = ff[5 + c];
}

// ---------------------------------------------------------------------------------
// Fast face flux.
// ---------------------------------------------------------------------------------

// Per-point values along the sweep direction (direction-local momentum order).
struct Pt {
  double r, mn, m1, m2, e;   // conserved: rho, normal momentum, transverse momenta, energy
  double p, ir, a, sr;       // pressure, 1/rho, |u_n| + c, sqrt(rho)
};

__device__ inline Pt make_point(double r, double mn, double m1, double m2, double e,
                                double gamma) {
  Pt q;
  q.r = r; q.mn = mn; q.m1 = m1; q.m2 = m2; q.e = e;
  q.ir = fast_rcp(r);
  q.p = (gamma - 1.0) * (e - 0.5 * (mn * mn + m1 * m1 + m2 * m2) * q.ir);
  const double un = mn * q.ir;
  q.a = fabs(un) + sun_sqrt(gamma * q.p * q.ir);
  q.sr = sun_sqrt(r);
  return q;
}

// WENO5-JS numerator/denominator pair for one component, with three algebraic savings:
//  * beta_k = (13/12) x_k^2 + (1/4) y_k^2 = (13/12) (x_k^2 + (3/13) y_k^2); the common factor
//    (13/12)^2 of every (eps + beta_k)^2 cancels in the weight ratio, so the weights use
//    (eps*12/13 + x_k^2 + (3/13) y_k^2)^2, formed as two FMAs around one product;
//  * over the common denominator s1 s2 s3 (s_k = (eps + beta_k)^2) the weights are
//    d_k prod_{m != k} s_m, so the linear weights d_k multiply the candidate stencils
//    (constant coefficients) instead of the products;
//  * the candidate stencils' 1/6 is folded into the same coefficients, and each stencil is
//    written as its f2 part plus differences already formed for the smoothness indicators.
// Returns the reconstructed value as num/den. epsb is eps*12/13 for the values as given;
// WENO5 is homogeneous in the sense WENO(l f; eps) = l WENO(f; eps / l^2), so a caller
// that reconstructs f / l passes WENO_EPSB / l^2 and multiplies the result by l.
constexpr double WENO_EPSB = 1e-6 * 12.0 / 13.0;
template <bool PLUS>
__device__ inline void weno5_nd(double f0, double f1, double f2, double f3, double f4,
                                double epsb, double &num, double &den) {
  const double c = 3.0 / 13.0;
  const double d01 = f1 - f0, d12 = f2 - f1, d23 = f3 - f2, d34 = f4 - f3;
  const double x1 = d34 - d23, x2 = d23 - d12, x3 = d12 - d01;
  const double y1 = x1 - 2.0 * d23;         // 3 f2 - 4 f3 + f4
  const double y2 = d12 + d23;              // -(f1 - f3)
  const double y3 = x3 + 2.0 * d12;         // f0 - 4 f1 + 3 f2
  const double b1 = fma(x1, x1, fma(c * y1, y1, epsb));
  const double b2 = fma(x2, x2, fma(c * y2, y2, epsb));
  const double b3 = fma(x3, x3, fma(c * y3, y3, epsb));
  const double s1 = b1 * b1, s2 = b2 * b2, s3 = b3 * b3;
  const double s23 = s2 * s3, s13 = s1 * s3, s12 = s1 * s2;
  // d_k / 6 times the candidate stencils, each written as (6 d_k / 6) f2 plus differences:
  // the f2 terms add up to f2 den, so num = f2 den + sum_k q'_k prod_{m != k} s_m
  double q1, q2, q3;
  if (PLUS) {               // d = (3, 6, 1) / 10
    q1 = fma(0.2, d23, -0.05 * d34);                       // (2 f2 + 5 f3 - f4) / 20
    q2 = fma(0.1, d12, 0.2 * d23);                         // (-f1 + 5 f2 + 2 f3) / 10
    q3 = fma(1.0 / 12.0, d12, (-1.0 / 30.0) * d01);        // (2 f0 - 7 f1 + 11 f2) / 60
    den = fma(0.3, s23, fma(0.6, s13, 0.1 * s12));
  } else {                  // d = (1, 6, 3) / 10
    q1 = fma(-1.0 / 12.0, d23, (1.0 / 30.0) * d34);        // (11 f2 - 7 f3 + 2 f4) / 60
    q2 = fma(-0.2, d12, -0.1 * d23);                       // (2 f1 + 5 f2 - f3) / 10
    q3 = fma(-0.2, d12, 0.05 * d01);                       // (-f0 + 5 f1 + 2 f2) / 20
    den = fma(0.1, s23, fma(0.6, s13, 0.3 * s12));
  }
  num = fma(f2, den, fma(q1, s23, fma(q2, s13, q3 * s12)));
}

// Sum of the f+ and f- reconstructions of one component with a single divide:
//   n+/d+ + n-/d- = (n+ d- + n- d+) / (d+ d-)
// (d+ d- is a product of four (eps+beta)^2 factors: finite for |beta| < 1e38).
__device__ inline double weno5_pair(const double (&fp)[5], const double (&fm)[5]) {
  double np, dp, nm, dm;
  weno5_nd<true>(fp[0], fp[1], fp[2], fp[3], fp[4], WENO_EPSB, np, dp);
  weno5_nd<false>(fm[0], fm[1], fm[2], fm[3], fm[4], WENO_EPSB, nm, dm);
  return fast_div(fma(np, dm, nm * dp), dp * dm);
}

// Face flux with the 6-point stencil read from LDS as it is needed (point l, field q at
// base[q*FS + l*LS]), one split side at a time, so only one side's 25 projected values are
// live. Fields: 0 r, FMN the normal momentum, FM1 the first transverse one, 3 the second,
// 4 e, 5 p, 6 1/r, FA |u_n| + c, 8 sqrt(r) (Pt order is FMN = 1, FM1 = 2, FA = 7; a point
// shared by x- and y-faces keeps mx, my in 1, 2 and both speeds, in 7 and 9).
template <int FMN, int FM1, int FA>
__device__ inline void face_flux_lds_f(const double *base, int FS, int LS, double gamma,
                                       double (&F)[5], double &alpha_out) {
  const double gm = gamma - 1.0;
  double alpha = 0.0;
#pragma unroll
  for (int l = 0; l < 6; l++) alpha = fmax(alpha, base[FA * FS + l * LS]);
  double u, v, w, H;
  {
    const double *L = base + 2 * LS, *R = base + 3 * LS;
    const double Lsr = L[8 * FS], Rsr = R[8 * FS];
    const double iden = fast_rcp(Lsr + Rsr);
    const double wl = Lsr * L[6 * FS], wr = Rsr * R[6 * FS];
    u = (L[FMN * FS] * wl + R[FMN * FS] * wr) * iden;
    v = (L[FM1 * FS] * wl + R[FM1 * FS] * wr) * iden;
    w = (L[3 * FS] * wl + R[3 * FS] * wr) * iden;
    H = ((L[5 * FS] + L[4 * FS]) * wl + (R[5 * FS] + R[4 * FS]) * wr) * iden;
  }
  const double qsq = u * u + v * v + w * w;
  const double hq = 0.5 * qsq;
  const double cs = gm * (H - hq);
  const double ci = fast_rcp(cs);
  const double hci = 0.25 * ci;
  // K = (gamma-1)(q^2/2 x0 - u x1 - v x2 - w x3 + x4) with the factor folded in
  const double gh = gm * hq, gu = gm * u, gv = gm * v, gw = gm * w;
  // The characteristic values are reconstructed without their constant factors (hci for
  // components 0 and 4, 1/2 for 1..3), with eps scaled to match (1 / hci^2 = 16 cs^2).
  const double eps04 = (16.0 * WENO_EPSB) * (cs * cs), eps13 = 4.0 * WENO_EPSB;
  double nps[5], dps[5], ff[5];
#pragma unroll
  for (int side = 0; side < 2; side++) {
    __builtin_amdgcn_sched_barrier(0);
    double pr[5][5];
#pragma unroll
    for (int l = 0; l < 5; l++) {
      const double *q = base + (l + side) * LS;
      const double r = q[0], mn = q[FMN * FS], m1 = q[FM1 * FS], m2 = q[3 * FS],
                   e = q[4 * FS];
      const double p = q[5 * FS];
      const double ul = mn * q[6 * FS];
      const double sa = side == 0 ? alpha : -alpha;
      const double us = ul + sa;
      const double x0 = fma(sa, r, mn);
      const double x1 = fma(us, mn, p);
      const double x2 = us * m1;
      const double x3 = us * m2;
      const double x4 = fma(us, e, ul * p);
      const double K = fma(gm, x4, fma(-gw, x3, fma(-gv, x2, fma(-gu, x1, gh * x0))));
      const double A = fma(u, x0, -x1);
      pr[l][0] = A + K;                  // times hci
      pr[l][1] = fma(-v, x0, x2);        // times 1/2
      pr[l][2] = fma(-w, x0, x3);        // times 1/2
      pr[l][3] = fma(-ci, K, x0);        // times 1/2
      pr[l][4] = K - A;                  // times hci
    }
#pragma unroll
    for (int c = 0; c < 5; c++) {
      const double eps = (c == 0 || c == 4) ? eps04 : eps13;
      if (side == 0) {
        weno5_nd<true>(pr[0][c], pr[1][c], pr[2][c], pr[3][c], pr[4][c], eps, nps[c], dps[c]);
      } else {
        double nm, dm;
        weno5_nd<false>(pr[0][c], pr[1][c], pr[2][c], pr[3][c], pr[4][c], eps, nm, dm);
        ff[c] = fast_div(fma(nps[c], dm, nm * dps[c]), dps[c] * dm);
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  // the factors back: hci on 0 and 4 (cs hci = 1/4), 1/2 on 1..3
  const double s04 = hci * (ff[0] + ff[4]);
  const double d40 = 0.25 * (ff[4] - ff[0]);
  const double f1 = 0.5 * ff[1], f2 = 0.5 * ff[2], f3 = 0.5 * ff[3];
  F[0] = s04 + f3;
  F[1] = u * F[0] + d40;
  F[2] = v * F[0] + f1;
  F[3] = w * F[0] + f2;
  F[4] = H * s04 + u * d40 + v * f1 + w * f2 + hq * f3;
  alpha_out = alpha;
}

__device__ inline void face_flux_lds(const double *base, int FS, int LS, double gamma,
                                     double (&F)[5], double &alpha_out) {
  face_flux_lds_f<1, 2, 7>(base, FS, LS, gamma, F, alpha_out);
}

// Species (passive tracer) face flux for one component from its 6 point values c[l]
// (utilities.cpp:376-377,395,439,473): component-wise WENO5 of the split fluxes
// (u c +- alpha c) / 2, no characteristic projection. The split fluxes are formed as
// c[l] hp[l] and c[l+1] hm[l] with the face's split factors hp[l] = (u_l + alpha) / 2 and
// hm[l] = (u_{l+1} - alpha) / 2 (tracer_split), which every tracer of the face shares.
__device__ inline void tracer_split(const double (&un)[6], double alpha, double (&hp)[5],
                                    double (&hm)[5]) {
#pragma unroll
  for (int l = 0; l < 5; l++) {
    hp[l] = 0.5 * (un[l] + alpha);
    hm[l] = 0.5 * (un[l + 1] - alpha);
  }
}

__device__ inline double tracer_face_flux(const double (&c)[6], const double (&hp)[5],
                                          const double (&hm)[5]) {
  double fp[5], fm[5];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    fp[l] = c[l] * hp[l];
    fm[l] = c[l + 1] * hm[l];
  }
  return weno5_pair(fp, fm);
}

}  // namespace mvd
