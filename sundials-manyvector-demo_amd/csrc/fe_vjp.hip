// Vector-Jacobian product of fe(): wbar = (d fe / d w)^T v.
//
// fe(w) = -div F with every face flux a function of its 6-point stencil, so
//   Fbar(face between cells c-1 | c) = (v_c - v_{c-1}) / delta   (no cell: 0)
// and each face's stencil adjoint sbar = (dF/ds)^T Fbar goes back to the stencil's source
// cells (ghost points are +-copies of interior cells). Every face recomputes its flux (the
// reference recipe, utilities.cpp:270-479) and sweeps it in reverse (WENO5-JS weights,
// Lax-Friedrichs split, characteristic projections, Roe average and eigenvectors, point
// fluxes, wave speeds, EOS). No face buffer: one sweep kernel per direction adds the
// contributions into the cells as it walks along the lines, through a ring of open rows
// in LDS, and writes each cell once its last face is in:
//   fe_vjp_sweep_x   - a wave walks whole x-lines 64 faces at a time (lanes along x);
//   fe_vjp_sweep_yz  - each lane walks its own column along y or z (lanes along x);
//   fe_vjp_fix_x/_yz - add the contributions that left a line or segment (ghost points,
//                      rows of neighbouring segments), held in a small side buffer.
// The x sweep stores wbar, the others add. No atomics: the result is deterministic.
// The tracers of w and wbar are handled tracer-major inside (cells_to_planes /
// planes_to_cells around the sweeps); the cotangent's, read at two cells per face, stay
// cell-major; the interface keeps the cell-major layout.
#include "mvd_device.hpp"
#include "mvd_internal.hpp"
#include "../../include/mvd_rhs.h"

namespace mvd {

namespace {

constexpr double W_BC = 1.083333333333333333333333333333333333333;   // 13/12
constexpr double W_EPS = 1e-6;
constexpr double W_C13 = 0.3333333333333333333333333333333333333333;
constexpr double W_C56 = 0.8333333333333333333333333333333333333333;
constexpr double W_C16 = 0.1666666666666666666666666666666666666667;
constexpr double W_C76 = 1.166666666666666666666666666666666666667;
constexpr double W_C116 = 1.833333333333333333333333333333333333333;

// WENO5-JS reconstruction from f[0..4] (side 0: the f+ formula over points i-3..i+1,
// side 1: the mirrored f- formula over points i-2..i+2) and its gradient g = dr/df.
// With the normalised weights o_k = w_k / sum w: r = sum o_k q_k and
//   dr/df = sum o_k dq_k/df + sum_k kap_k db_k/df,   kap_k = -2 o_k (q_k - r) / e_k,
// db_k/df = 2 (13/12) A_k dA_k/df + (B_k / 2) dB_k/df; the gradient is written out term by
// term (no products with the zero entries of the stencil derivatives).
__device__ inline double weno5_grad(const double (&f)[5], int side, double (&g)[5]) {
  const double A1 = f[2] - 2.0 * f[3] + f[4], B1 = 3.0 * f[2] - 4.0 * f[3] + f[4];
  const double A2 = f[1] - 2.0 * f[2] + f[3], B2 = f[1] - f[3];
  const double A3 = f[0] - 2.0 * f[1] + f[2], B3 = f[0] - 4.0 * f[1] + 3.0 * f[2];
  const double b1 = W_BC * A1 * A1 + 0.25 * B1 * B1;
  const double b2 = W_BC * A2 * A2 + 0.25 * B2 * B2;
  const double b3 = W_BC * A3 * A3 + 0.25 * B3 * B3;
  const double d1 = side == 0 ? 0.3 : 0.1, d3 = side == 0 ? 0.1 : 0.3;
  const double e1 = W_EPS + b1, e2 = W_EPS + b2, e3 = W_EPS + b3;
  const double ie1 = fast_rcp(e1), ie2 = fast_rcp(e2), ie3 = fast_rcp(e3);
  const double w1 = d1 * ie1 * ie1, w2 = 0.6 * ie2 * ie2, w3 = d3 * ie3 * ie3;
  // candidate stencils q_k and their coefficients (dq_k/df over the points they use)
  double q1, q2, q3;
  double c10, c11, c12, c21, c22, c23, c32, c33, c34;   // dq3/df0..2, dq2/df1..3, dq1/df2..4
  if (side == 0) {
    c10 = W_C13; c11 = -W_C76; c12 = W_C116;
    c21 = -W_C16; c22 = W_C56; c23 = W_C13;
    c32 = W_C13; c33 = W_C56; c34 = -W_C16;
  } else {
    c10 = -W_C16; c11 = W_C56; c12 = W_C13;
    c21 = W_C13; c22 = W_C56; c23 = -W_C16;
    c32 = W_C116; c33 = -W_C76; c34 = W_C13;
  }
  q1 = c32 * f[2] + c33 * f[3] + c34 * f[4];
  q2 = c21 * f[1] + c22 * f[2] + c23 * f[3];
  q3 = c10 * f[0] + c11 * f[1] + c12 * f[2];
  const double iS = fast_rcp(w1 + w2 + w3);
  const double o1 = w1 * iS, o2 = w2 * iS, o3 = w3 * iS;
  const double r = q1 * o1 + q2 * o2 + q3 * o3;
  const double k1 = -2.0 * o1 * ie1 * (q1 - r);
  const double k2 = -2.0 * o2 * ie2 * (q2 - r);
  const double k3 = -2.0 * o3 * ie3 * (q3 - r);
  // kap_k db_k/df = a_k dA_k/df + c_k dB_k/df
  const double a1 = (2.0 * W_BC) * k1 * A1, h1 = 0.5 * k1 * B1;   // dA1 (0,0,1,-2,1), dB1 (0,0,3,-4,1)
  const double a2 = (2.0 * W_BC) * k2 * A2, h2 = 0.5 * k2 * B2;   // dA2 (0,1,-2,1,0), dB2 (0,1,0,-1,0)
  const double a3 = (2.0 * W_BC) * k3 * A3, h3 = 0.5 * k3 * B3;   // dA3 (1,-2,1,0,0), dB3 (1,-4,3,0,0)
  g[0] = o3 * c10 + a3 + h3;
  g[1] = o2 * c21 + o3 * c11 + a2 + h2 - 2.0 * a3 - 4.0 * h3;
  g[2] = o1 * c32 + o2 * c22 + o3 * c12 + a1 + 3.0 * h1 - 2.0 * a2 + a3 + 3.0 * h3;
  g[3] = o1 * c33 + o2 * c23 - 2.0 * a1 - 4.0 * h1 + a2 - h2;
  g[4] = o1 * c34 + a1 + h1;
  return r;
}

// Adjoint of one face flux for the five fluid components, plus what the tracers need:
// s, Fb, sb in direction-local order (rho, m_normal, m_t1, m_t2, et). The tracer part of
// the reverse sweep (tracer_vjp) runs first and hands over the adjoints of the point
// velocities (ulb) and of the Lax-Friedrichs speed (alphab) it produced.
struct FacePrim {
  double p[6], ul[6], ir[6];   // pressure, normal velocity, 1/rho per stencil point
  double alpha;
  int lstar;                   // the point that sets alpha
};

__device__ inline void face_prim(const double (&s)[6][5], double gamma, FacePrim &P) {
  const double gm = gamma - 1.0;
  P.alpha = 0.0;
  P.lstar = -1;
#pragma unroll
  for (int l = 0; l < 6; l++) {
    P.ir[l] = fast_rcp(s[l][0]);
    P.p[l] = gm * (s[l][4] - (s[l][1] * s[l][1] + s[l][2] * s[l][2] + s[l][3] * s[l][3]) * 0.5 * P.ir[l]);
    P.ul[l] = s[l][1] * P.ir[l];
    const double a = fabs(P.ul[l]) + sun_sqrt(gamma * P.p[l] * P.ir[l]);
    if (P.alpha < a) { P.alpha = a; P.lstar = l; }
  }
}

// One tracer: t[6] the stencil values, Fbc its flux cotangent; tb[6] receives the tracer
// adjoint, ulb and alphab accumulate. The split flux fs = (ul t +- alpha t) / 2 is formed as
// t hs with the point's split factor hs = (ul +- alpha) / 2 (ha = alpha / 2), whose
// derivatives are hs (t), t / 2 (ul) and +-t / 2 (alpha).
__device__ inline void tracer_vjp(const double (&t)[6], double Fbc, const FacePrim &P,
                                  double ha, double (&tb)[6], double (&ulb)[6],
                                  double &alphab) {
#pragma unroll
  for (int l = 0; l < 6; l++) tb[l] = 0.0;
#pragma unroll
  for (int sd = 0; sd < 2; sd++) {
    double fa[5], hs[5], g[5];
#pragma unroll
    for (int l = 0; l < 5; l++) {
      const int pt = l + sd;
      hs[l] = fma(0.5, P.ul[pt], sd == 0 ? ha : -ha);
      fa[l] = t[pt] * hs[l];
    }
    (void)weno5_grad(fa, sd, g);
    const double sg = sd == 0 ? 0.5 : -0.5;
#pragma unroll
    for (int l = 0; l < 5; l++) {
      const int pt = l + sd;
      const double fsb = Fbc * g[l];
      const double ft = fsb * t[pt];
      ulb[pt] += 0.5 * ft;
      alphab += sg * ft;
      tb[pt] += fsb * hs[l];
    }
  }
}

__device__ inline void fluid_vjp(const double (&s)[6][5], double gamma, const FacePrim &P,
                                 const double (&Fb)[5], double (&ulb)[6], double alphab,
                                 double (&sb)[6][5]) {
  const double gm = gamma - 1.0;
  const double alpha = P.alpha;
  const double sL = sun_sqrt(s[2][0]), sR = sun_sqrt(s[3][0]);
  const double isL = fast_rcp(sL), isR = fast_rcp(sR);
  const double sbar = 0.5 * (sL + sR), isbar = fast_rcp(sbar);
  const double u = 0.5 * (s[2][1] * isL + s[3][1] * isR) * isbar;
  const double v = 0.5 * (s[2][2] * isL + s[3][2] * isR) * isbar;
  const double w = 0.5 * (s[2][3] * isL + s[3][3] * isR) * isbar;
  const double H = 0.5 * ((P.p[2] + s[2][4]) * isL + (P.p[3] + s[3][4]) * isR) * isbar;
  const double qsq = u * u + v * v + w * w;
  const double cs = gm * (H - 0.5 * qsq);
  const double ci = fast_rcp(cs);
  const double LV[5][5] = {
      {0.5 * ci * (u + 0.5 * gm * qsq), -0.5 * ci * (gm * u + 1.0), -0.5 * v * gm * ci,
       -0.5 * w * gm * ci, 0.5 * gm * ci},
      {-v, 0.0, 1.0, 0.0, 0.0},
      {-w, 0.0, 0.0, 1.0, 0.0},
      {(-gm) * ci * (qsq - H), u * gm * ci, v * gm * ci, w * gm * ci, (-gm) * ci},
      {-0.5 * ci * (u - 0.5 * gm * qsq), -0.5 * ci * (gm * u - 1.0), -0.5 * v * gm * ci,
       -0.5 * w * gm * ci, 0.5 * gm * ci}};
  auto flux = [&](int l, int c) -> double {
    if (c == 0) return s[l][1];
    if (c == 1) return P.ul[l] * s[l][1] + P.p[l];
    if (c == 4) return P.ul[l] * (s[l][4] + P.p[l]);
    return P.ul[l] * s[l][c];
  };
  auto split = [&](int sd, int l, int c) -> double {
    const int pt = l + sd;
    return sd == 0 ? 0.5 * (flux(pt, c) + alpha * s[pt][c])
                   : 0.5 * (flux(pt, c) - alpha * s[pt][c]);
  };
  // ---- reverse ----
#pragma unroll
  for (int l = 0; l < 6; l++)
#pragma unroll
    for (int c = 0; c < 5; c++) sb[l][c] = 0.0;
  double pb[6] = {0, 0, 0, 0, 0, 0};
  auto flux_adj = [&](int pt, int c, double fb) {
    if (c == 0) { sb[pt][1] += fb; return; }
    if (c == 1) { ulb[pt] += fb * s[pt][1]; sb[pt][1] += fb * P.ul[pt]; pb[pt] += fb; return; }
    if (c == 4) {
      ulb[pt] += fb * (s[pt][4] + P.p[pt]);
      sb[pt][4] += fb * P.ul[pt];
      pb[pt] += fb * P.ul[pt];
      return;
    }
    ulb[pt] += fb * s[pt][c];
    sb[pt][c] += fb * P.ul[pt];
  };
  // back-projection F = RV ff: ffb = RV^T Fb does not depend on ff, so the forward
  // reconstruction and its reverse share one WENO pass per side
  double ffb[5];
  ffb[0] = Fb[0] + (u - cs) * Fb[1] + v * Fb[2] + w * Fb[3] + (H - u * cs) * Fb[4];
  ffb[1] = Fb[2] + v * Fb[4];
  ffb[2] = Fb[3] + w * Fb[4];
  ffb[3] = Fb[0] + u * Fb[1] + v * Fb[2] + w * Fb[3] + 0.5 * qsq * Fb[4];
  ffb[4] = Fb[0] + (u + cs) * Fb[1] + v * Fb[2] + w * Fb[3] + (H + u * cs) * Fb[4];
  double ff[5] = {0, 0, 0, 0, 0};
  double ub = 0.0, vb = 0.0, wb = 0.0, Hb = 0.0, cib = 0.0, qsqb = 0.0;
#pragma unroll
  for (int sd = 0; sd < 2; sd++) {
    // fc[a][l]: characteristic split flux, then (in place) its adjoint ffb[a] * dW_a/dfc
    double fc[5][5];
#pragma unroll
    for (int l = 0; l < 5; l++) {
      double fs[5];
#pragma unroll
      for (int b = 0; b < 5; b++) fs[b] = split(sd, l, b);
#pragma unroll
      for (int a = 0; a < 5; a++)
        fc[a][l] = LV[a][0] * fs[0] + LV[a][1] * fs[1] + LV[a][2] * fs[2] +
                   LV[a][3] * fs[3] + LV[a][4] * fs[4];
    }
#pragma unroll
    for (int a = 0; a < 5; a++) {
      double g[5];
      ff[a] += weno5_grad(fc[a], sd, g);
#pragma unroll
      for (int l = 0; l < 5; l++) fc[a][l] = ffb[a] * g[l];
    }
    // per stencil point: LV-entry adjoints (LVb[a][b] = sum_l fc[a][l] fs[l][b]) folded
    // straight into the Roe-state adjoints, and fsb = LV^T fc[.][l] into the split
#pragma unroll
    for (int l = 0; l < 5; l++) {
      double fs[5];
#pragma unroll
      for (int b = 0; b < 5; b++) fs[b] = split(sd, l, b);
      const double e0 = fc[0][l], e3 = fc[3][l], e4 = fc[4][l];
      const double e04p = e0 + e4;
      cib += 0.5 * (u + 0.5 * gm * qsq) * e0 * fs[0] - 0.5 * (gm * u + 1.0) * e0 * fs[1] -
             0.5 * gm * e04p * (v * fs[2] + w * fs[3] - fs[4]) - gm * (qsq - H) * e3 * fs[0] +
             gm * e3 * (u * fs[1] + v * fs[2] + w * fs[3] - fs[4]) -
             0.5 * (u - 0.5 * gm * qsq) * e4 * fs[0] - 0.5 * (gm * u - 1.0) * e4 * fs[1];
      ub += 0.5 * ci * (e0 - e4) * fs[0] - 0.5 * ci * gm * e04p * fs[1] + gm * ci * e3 * fs[1];
      qsqb += 0.25 * ci * gm * e04p * fs[0] - gm * ci * e3 * fs[0];
      vb += -0.5 * gm * ci * e04p * fs[2] - fc[1][l] * fs[0] + gm * ci * e3 * fs[2];
      wb += -0.5 * gm * ci * e04p * fs[3] - fc[2][l] * fs[0] + gm * ci * e3 * fs[3];
      Hb += gm * ci * e3 * fs[0];
      const int pt = l + sd;
      const double sg = sd == 0 ? 0.5 : -0.5;
#pragma unroll
      for (int b = 0; b < 5; b++) {
        const double fsb = LV[0][b] * fc[0][l] + LV[1][b] * fc[1][l] + LV[2][b] * fc[2][l] +
                           LV[3][b] * fc[3][l] + LV[4][b] * fc[4][l];
        flux_adj(pt, b, 0.5 * fsb);
        sb[pt][b] += sg * alpha * fsb;
        alphab += sg * fsb * s[pt][b];
      }
    }
  }
  // RV entries -> Roe state
  double csb;
  {
    const double R10 = Fb[1] * ff[0], R14 = Fb[1] * ff[4];
    const double R40 = Fb[4] * ff[0], R44 = Fb[4] * ff[4];
    ub += R10 + Fb[1] * ff[3] + R14 + (R44 - R40) * cs;
    csb = R14 - R10 + (R44 - R40) * u;
    vb += Fb[2] * (ff[0] + ff[3] + ff[4]) + Fb[4] * ff[1];
    wb += Fb[3] * (ff[0] + ff[3] + ff[4]) + Fb[4] * ff[2];
    Hb += R40 + R44;
    qsqb += 0.5 * Fb[4] * ff[3];
  }
  csb += -cib * ci * ci;
  Hb += gm * csb;
  qsqb += -0.5 * gm * csb;
  ub += 2.0 * u * qsqb;
  vb += 2.0 * v * qsqb;
  wb += 2.0 * w * qsqb;
  double sLb = 0.0, sRb = 0.0, sbarb = 0.0;
  {
    const double Aub = 0.5 * ub * isbar, Avb = 0.5 * vb * isbar, Awb = 0.5 * wb * isbar,
                 AHb = 0.5 * Hb * isbar;
    sbarb -= (ub * u + vb * v + wb * w + Hb * H) * isbar;
    sb[2][1] += Aub * isL; sb[3][1] += Aub * isR;
    sb[2][2] += Avb * isL; sb[3][2] += Avb * isR;
    sb[2][3] += Awb * isL; sb[3][3] += Awb * isR;
    sb[2][4] += AHb * isL; sb[3][4] += AHb * isR;
    pb[2] += AHb * isL; pb[3] += AHb * isR;
    sLb -= (Aub * s[2][1] + Avb * s[2][2] + Awb * s[2][3] + AHb * (P.p[2] + s[2][4])) * isL * isL;
    sRb -= (Aub * s[3][1] + Avb * s[3][2] + Awb * s[3][3] + AHb * (P.p[3] + s[3][4])) * isR * isR;
  }
  sLb += 0.5 * sbarb;
  sRb += 0.5 * sbarb;
  if (s[2][0] > 0.0) sb[2][0] += 0.5 * sLb * isL;
  if (s[3][0] > 0.0) sb[3][0] += 0.5 * sRb * isR;
  // Lax-Friedrichs speed: alpha = |u_l*| + sqrt(gamma p_l* / rho_l*)
  {
    double ulS = 0.0, pS = 0.0, irS = 0.0;
#pragma unroll
    for (int l = 0; l < 6; l++)
      if (l == P.lstar) { ulS = P.ul[l]; pS = P.p[l]; irS = P.ir[l]; }
    const double X = gamma * pS * irS;
    const double ulbS = ulS < 0.0 ? -alphab : alphab;
    double pbS = 0.0, rbS = 0.0;
    if (X > 0.0) {
      const double Xb = 0.5 * alphab * fast_rcp(sun_sqrt(X));
      pbS = Xb * gamma * irS;
      rbS = -Xb * X * irS;
    }
#pragma unroll
    for (int l = 0; l < 6; l++)
      if (l == P.lstar) { ulb[l] += ulbS; pb[l] += pbS; sb[l][0] += rbS; }
  }
  // ul = m / rho, p = gm (et - |m|^2 / (2 rho))
#pragma unroll
  for (int l = 0; l < 6; l++) {
    const double ir = P.ir[l];
    sb[l][1] += ulb[l] * ir;
    sb[l][0] -= ulb[l] * P.ul[l] * ir;
    const double K = s[l][1] * s[l][1] + s[l][2] * s[l][2] + s[l][3] * s[l][3];
    sb[l][4] += gm * pb[l];
    const double Kb = -gm * pb[l] * 0.5 * ir;
    sb[l][1] += 2.0 * s[l][1] * Kb;
    sb[l][2] += 2.0 * s[l][2] * Kb;
    sb[l][3] += 2.0 * s[l][3] * Kb;
    sb[l][0] += gm * pb[l] * K * 0.5 * ir * ir;
  }
}

__device__ inline long cell_idx(const Grid &g, long i, long j, long k) {
  return i + (long)g.nx * (j + (long)g.ny * k);
}

// x times -1 when neg (the sign bit flipped: bitwise x * -1.0), else x
__device__ inline double flip(double x, unsigned neg) {
  return __longlong_as_double(__double_as_longlong(x) ^ ((long long)neg << 63));
}

// direction-local component a -> global field index
__device__ inline int gcomp(int d, int a) {
  if (a == 1) return 1 + d;
  if (a == 1 + d) return 1;
  return a;
}



}  // namespace

// One face's stencil adjoint, handed to a sink with the ghost rule's signs applied, one
// component at a time: the tracers first (the fluid part needs the velocity and
// wave-speed adjoints they produce), then the fluid in global component order. For each
// component the sink's pre(q) is called before the adjoint is computed (to start the
// loads its put needs) and put(q, x[6], pre) after. The next tracer's stencil is loaded
// while the current one is swept. Face (i, j, k) along d lies between cells c-1 and c.
template <int NC, class Sink>
__device__ inline void face_vjp(const Grid &g, const Fields &w, const Fields &vc, int d, int i,
                                int j, int k, Sink &sink) {
  const int n = extent(g, d);
  const int c0 = d == 0 ? i : d == 1 ? j : k;
  const double idelta = 1.0 / (d == 0 ? g.dx : d == 1 ? g.dy : g.dz);
  // cells either side of the face; a missing one reads the other and counts as zero (the
  // loads stay unconditional, which keeps the compiler's waits on them precise)
  const bool hr = c0 < n, hl = c0 > 0;
  const long cr = cell_idx(g, i - (d == 0 && !hr), j - (d == 1 && !hr), k - (d == 2 && !hr));
  const long cl = hl ? cell_idx(g, i - (d == 0), j - (d == 1), k - (d == 2)) : cr;
  const double mr = hr ? 1.0 : 0.0, ml = hl ? 1.0 : 0.0;
  long cell[6];
  unsigned na = 0, nn = 0;     // bit l: point l's ghost sign is -1 (all / normal)
  double s[6][5];
#pragma unroll
  for (int l = 0; l < 6; l++) {
    const GhostMap m = ghost_map(g, d, c0 - 3 + l);
    int ii = i, jj = j, kk = k;
    if (d == 0) ii = m.src; else if (d == 1) jj = m.src; else kk = m.src;
    cell[l] = cell_idx(g, ii, jj, kk);
    na |= (m.sgn_all < 0.0 ? 1u : 0u) << l;
    nn |= (m.sgn_norm < 0.0 ? 1u : 0u) << l;
#pragma unroll
    for (int a = 0; a < 5; a++)
      s[l][a] = flip(w.w[gcomp(d, a)][cell[l]], ((a == 1 ? nn : na) >> l) & 1u);
  }
  FacePrim P;
  face_prim(s, g.gamma, P);
  double ulb[6] = {0, 0, 0, 0, 0, 0};
  double alphab = 0.0;
  if (NC > 0) {
    // software pipeline: the stencil and sink loads of tracer c + PD are issued before
    // tracer c is swept (the empty asm statement keeps the compiler from sinking them);
    // PD = 1 measured fastest (2 and 3 were 1-2% slower at 128^3 with 10 tracers)
    constexpr int PD = 1;
    struct Tr { double t[6], r, l; };
    const long ncell = (long)g.nx * g.ny * g.nz;
    auto load_tracer = [&](Tr &x, int c) {
      const long pl = (long)c * ncell;
#pragma unroll
      for (int l = 0; l < 6; l++) x.t[l] = w.chem[pl + cell[l]];
      x.r = vc.chem[(long)NC * cr + c];   // the cotangent stays cell-major (two cells a face)
      x.l = vc.chem[(long)NC * cl + c];
    };
    using Pre = decltype(sink.pre(5));
    Tr buf[PD];
    Pre pb[PD];
#pragma unroll
    for (int k = 0; k < PD; k++) {
      pb[k] = sink.pre(5 + k);
      load_tracer(buf[k], k);
    }
#pragma unroll 1
    for (int c0 = 0; c0 < NC; c0 += PD) {
#pragma unroll
      for (int k = 0; k < PD; k++) {
        const int c = c0 + k;
        if (c < NC) {
          double t[6], tb[6];
#pragma unroll
          for (int l = 0; l < 6; l++) t[l] = flip(buf[k].t[l], (na >> l) & 1u);
          const double Fbc = (mr * buf[k].r - ml * buf[k].l) * idelta;
          const Pre pre = pb[k];
          if (c + PD < NC) {
            pb[k] = sink.pre(5 + c + PD);
            load_tracer(buf[k], c + PD);
          }
          __asm__ volatile("" ::: "memory");
          tracer_vjp(t, Fbc, P, 0.5 * P.alpha, tb, ulb, alphab);
#pragma unroll
          for (int l = 0; l < 6; l++) tb[l] = flip(tb[l], (na >> l) & 1u);
          sink.put(5 + c, tb, pre);
        }
      }
    }
  }
  double Fb[5];
#pragma unroll
  for (int a = 0; a < 5; a++) {
    const int q = gcomp(d, a);
    Fb[a] = (mr * vc.w[q][cr] - ml * vc.w[q][cl]) * idelta;
  }
  double sb[6][5];
  fluid_vjp(s, g.gamma, P, Fb, ulb, alphab, sb);
#pragma unroll
  for (int a = 0; a < 5; a++) {
    const int q = gcomp(d, a);
    const auto pre = sink.pre(q);
    double x[6];
#pragma unroll
    for (int l = 0; l < 6; l++) x[l] = flip(sb[l][a], ((a == 1 ? nn : na) >> l) & 1u);
    sink.put(q, x, pre);
  }
}

__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline void out_put(const FieldsOut &o, long ncell, int q, long cell, double v,
                               bool add) {
  double *p = q < 5 ? &o.w[q][cell] : &o.chem[(long)(q - 5) * ncell + cell];
  *p = add ? *p + v : v;
}

// Ghost-point contributions of an x-line: the (face, point) pairs whose extended index
// lies outside the line, 6 on each end. Index of pair (f, l), or -1 for an interior point.
__device__ inline int xghost_pair(int n, int f, int l) {
  const int e = f - 3 + l;
  if (e < 0) return f == 0 ? l : f == 1 ? 3 + l : 5;
  if (e >= n) {
    const int b = n - f;
    return b == 0 ? 3 + l : b == 1 ? 5 + l : 11;   // l in 3..5 / 4..5 / 5
  }
  return -1;
}

// ------------------------------------------------------------------------------------
// x sweep: one wave walks a run of whole x-lines 64 faces at a time (faces consecutive in
// memory, lines back to back); the contributions go to a 128-slot ring indexed by the
// flattened position line * (nx+1) + e of their target. Lanes share slots, so the six
// stencil points are added in six ordered passes. A slot is final once the faces up to
// three beyond it are in; it is then written out and cleared. Ghost points go to the side
// buffer per (line, pair) and are added by fe_vjp_fix_x.
// ------------------------------------------------------------------------------------
// The six points go in as six no-return LDS adds, one point per instruction: within one
// instruction the lanes hit distinct slots, and a wave's LDS operations complete in
// order, so the sum order (point 0 first) is fixed.
template <int NV>
struct XSink {
  double *ring;            // [NV][128]
  double *side;            // [(pair * NV + q) * nlines + line]
  long q;                  // flattened face
  int f, n, line, nlines;
  __device__ int pre(int) const { return 0; }
  __device__ void put(int comp, const double (&x)[6], int) {
#pragma unroll
    for (int l = 0; l < 6; l++) {
      const int pr = xghost_pair(n, f, l);
      if (pr < 0) atomicAdd(&ring[comp * 128 + (int)((q - 3 + l) & 127)], x[l]);
      else side[((long)pr * NV + comp) * nlines + line] = x[l];
      wave_sync_lds();
    }
  }
};

template <int NC>
__global__ void __launch_bounds__(64)
fe_vjp_sweep_x(Grid g, Fields w, Fields vc, FieldsOut out, double *side, int lps, int add) {
  constexpr int NV = 5 + NC;
  __shared__ double ring[NV * 128];
  const int lane = threadIdx.x;
  const int nx = g.nx, nxp = g.nx + 1;
  const int nlines = g.ny * g.nz;
  const int L0 = blockIdx.x * lps;
  const int L1 = min(L0 + lps, nlines);
  if (L0 >= nlines) return;
  const long Q0 = (long)L0 * nxp, Q1 = (long)L1 * nxp;
  const long ncell = (long)nx * nlines;
  for (int s = lane; s < NV * 128; s += 64) ring[s] = 0.0;
  wave_sync_lds();
  long done = Q0;          // first slot not yet written out
  for (long B = Q0; B < Q1; B += 64) {
    const long q = B + lane;
    if (q < Q1) {
      const int line = (int)(q / nxp), f = (int)(q - (long)line * nxp);
      XSink<NV> sink{ring, side, q, f, nx, line, nlines};
      face_vjp<NC>(g, w, vc, 0, f, line % g.ny, line / g.ny, sink);
    }
    wave_sync_lds();
    const long hi = (B + 64 >= Q1) ? Q1 - 1 : B + 60;
    for (long P = done + lane; P <= hi; P += 64) {
      const int line = (int)(P / nxp), e = (int)(P - (long)line * nxp);
      const int slot = (int)(P & 127);
      if (e < nx) {
        const long cell = e + (long)nx * line;
#pragma unroll 1
        for (int qq = 0; qq < NV; qq++) out_put(out, ncell, qq, cell, ring[qq * 128 + slot], add);
      }
#pragma unroll 1
      for (int qq = 0; qq < NV; qq++) ring[qq * 128 + slot] = 0.0;
    }
    done = hi + 1;
    wave_sync_lds();
  }
}

// Ghost contributions of each x-line into their source cells (one thread per line and
// component, pairs in a fixed order). When the six ghost points e = -3..-1, n..n+2 have six
// distinct source cells (lines of six cells or more), each source takes its pairs' sums in
// registers, in the pairs' order, and is read and written once (all loads issued
// together): the same sums as one read-modify-write per pair.
template <int NC>
__global__ void __launch_bounds__(256)
fe_vjp_fix_x(Grid g, const double *side, FieldsOut out) {
  constexpr int NV = 5 + NC;
  const int nlines = g.ny * g.nz;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long)nlines * NV) return;
  const int line = (int)(tid % nlines), qq = (int)(tid / nlines);
  const int n = g.nx;
  const long ncell = (long)n * nlines;
  int src[6];
#pragma unroll
  for (int e = 0; e < 6; e++) src[e] = ghost_map(g, 0, e < 3 ? e - 3 : n + (e - 3)).src;
  bool distinct = n >= 6;
#pragma unroll
  for (int a = 0; a < 6; a++)
#pragma unroll
    for (int b = a + 1; b < 6; b++) distinct = distinct && src[a] != src[b];
  if (distinct) {
    // ghost e = f - 3 + l: e = -3 (0,0); -2 (0,1) (1,0); -1 (0,2) (1,1) (2,0);
    // n (n-2,5) (n-1,4) (n,3); n+1 (n-1,5) (n,4); n+2 (n,5), in pair order
    const auto sv = [&](int f, int l) {
      return side[((long)xghost_pair(n, f, l) * NV + qq) * nlines + line];
    };
    double add[6];
    add[0] = sv(0, 0);
    add[1] = sv(0, 1); add[2] = sv(0, 2);
    const double a10 = sv(1, 0), a11 = sv(1, 1), a20 = sv(2, 0);
    add[3] = sv(n - 2, 5); add[4] = sv(n - 1, 5); add[5] = sv(n, 5);
    const double b14 = sv(n - 1, 4), b03 = sv(n, 3), b04 = sv(n, 4);
    long cell[6];
    double old[6];
#pragma unroll
    for (int e = 0; e < 6; e++) {
      cell[e] = src[e] + (long)n * line;
      old[e] = qq < 5 ? out.w[qq][cell[e]] : out.chem[(long)(qq - 5) * ncell + cell[e]];
    }
    double v[6];
    v[0] = old[0] + add[0];
    v[1] = (old[1] + add[1]) + a10;
    v[2] = ((old[2] + add[2]) + a11) + a20;
    v[3] = ((old[3] + add[3]) + b14) + b03;
    v[4] = (old[4] + add[4]) + b04;
    v[5] = old[5] + add[5];
#pragma unroll
    for (int e = 0; e < 6; e++) {
      if (qq < 5) out.w[qq][cell[e]] = v[e];
      else out.chem[(long)(qq - 5) * ncell + cell[e]] = v[e];
    }
    return;
  }
  for (int f = 0; f <= n; f++) {
    if (f > 2 && f < n - 2) { f = n - 3; continue; }
#pragma unroll 1
    for (int l = 0; l < 6; l++) {
      const int pr = xghost_pair(n, f, l);
      if (pr < 0) continue;
      const long cell = ghost_map(g, 0, f - 3 + l).src + (long)n * line;
      out_put(out, (long)n * nlines, qq, cell, side[((long)pr * NV + qq) * nlines + line], true);
    }
  }
}

// ------------------------------------------------------------------------------------
// y and z sweeps: each lane walks its own column (x = 64 * block + lane, fixed transverse
// index t) over a segment [c0, c1) of faces along D; rows belong to that lane alone, so
// no lane waits on another. The five rows f-2..f+2 still open live in an LDS ring
// ([5][NV][64] per wave); row f-3 is complete with face f's first point and is written out
// at once, its slot taking row f+2. Rows outside [c0, c1) (ghost rows and the three rows
// either side that belong to neighbouring segments) go to the side buffer
// [((seg * 6 + k) * NV + q) * ncol + col] and are added by fe_vjp_fix_yz.
// ------------------------------------------------------------------------------------
template <int NV>
struct ColSink {
  double *ring;            // this wave's [5][NV][64]
  double *side;
  FieldsOut out;
  int lane, slot0;         // slot0: ring slot of row f-3 (slot of row f-3+l is (slot0+l)%5)
  int e0, c0, c1, seg;
  long ncell;
  long col, ncol;
  long cell0;              // cell of row f-3 when it lies in [c0, c1)
  long cbase;              // cell of the column's row 0
  bool add;
  __device__ int slot(int l) const { const int s = slot0 + l; return s >= 5 ? s - 5 : s; }
  __device__ bool owned() const { return e0 >= c0 && e0 < c1; }
  __device__ double *dst(int comp) const {
    return comp < 5 ? &out.w[comp][cell0] : &out.chem[(long)(comp - 5) * ncell + cell0];
  }
  struct Pre { double r0, w0; };
  // row f-3 so far, and the value already in wbar where this sweep adds to it (read
  // unconditionally, from the column's first row when f-3 is not an owned row, and only
  // looked at in put, so that the loads stay in flight meanwhile)
  __device__ Pre pre(int comp) const {
    const long c = add && owned() ? cell0 : cbase;
    const double w0 = comp < 5 ? out.w[comp][c] : out.chem[(long)(comp - 5) * ncell + c];
    return Pre{ring[((long)slot0 * NV + comp) * 64 + lane], w0};
  }
  __device__ void put(int comp, const double (&x)[6], const Pre &p) {
    const double v = p.r0 + x[0];         // row f-3 is now complete
    if (owned()) *dst(comp) = add ? p.w0 + v : v;
    else {
      const int k = e0 - (c0 - 3);        // e0 < c0 here
      side[(((long)seg * 6 + k) * NV + comp) * ncol + col] = v;
    }
    ring[((long)slot0 * NV + comp) * 64 + lane] = x[5];   // the slot now holds row f+2
#pragma unroll
    for (int l = 1; l < 5; l++) atomicAdd(&ring[((long)slot(l) * NV + comp) * 64 + lane], x[l]);
  }
};

template <int D, int NC>
__global__ void __launch_bounds__(64)
fe_vjp_sweep_yz(Grid g, Fields w, Fields vc, FieldsOut out, double *side, int nseg, int add) {
  constexpr int NV = 5 + NC;
  __shared__ double ring[5 * NV * 64];
  const int lane = threadIdx.x;
  const int nxb = (g.nx + 63) / 64;
  const int nt = D == 1 ? g.nz : g.ny;
  const int n = D == 1 ? g.ny : g.nz;
  const int xb = blockIdx.x % nxb;
  const int rest = blockIdx.x / nxb;
  const int t = rest % nt, seg = rest / nt;
  const int i = xb * 64 + lane;
  if (i >= g.nx) return;
  const int c0 = (int)((long)seg * n / nseg), c1 = (int)((long)(seg + 1) * n / nseg);
  const int flast = seg == nseg - 1 ? n : c1 - 1;
  const long ncol = (long)g.nx * nt, col = i + (long)g.nx * t;
  for (int s = 0; s < 5 * NV; s++) ring[s * 64 + lane] = 0.0;
  const long rstride = D == 1 ? g.nx : (long)g.nx * g.ny;
  const long cbase = D == 1 ? i + (long)g.nx * g.ny * t : i + (long)g.nx * t;
  const long ncell = (long)g.nx * g.ny * g.nz;
  ColSink<NV> sink{ring, side, out, lane, 0, 0, c0, c1, seg, ncell, col, ncol, 0, cbase,
                   add != 0};
  for (int f = c0; f <= flast; f++) {
    sink.slot0 = (f - c0) % 5;
    sink.e0 = f - 3;
    sink.cell0 = cbase + (long)(f - 3) * rstride;
    if (D == 1) face_vjp<NC>(g, w, vc, 1, i, f, t, sink);
    else face_vjp<NC>(g, w, vc, 2, i, t, f, sink);
  }
  // rows flast-2 .. flast+2 are complete
  for (int r = 0; r < 5; r++) {
    const int e = flast - 2 + r;
    const int sl = (flast + 1 - c0 + r) % 5;
    for (int qq = 0; qq < NV; qq++) {
      const double v = ring[((long)sl * NV + qq) * 64 + lane];
      if (e >= c0 && e < c1) out_put(out, ncell, qq, cbase + (long)e * rstride, v, add != 0);
      else {
        const int k = e < c0 ? e - (c0 - 3) : 3 + (e - c1);
        side[(((long)seg * 6 + k) * NV + qq) * ncol + col] = v;
      }
    }
  }
  if (flast + 2 < c1 + 2)      // an inner segment's last side row (c1 + 2) gets nothing
    for (int qq = 0; qq < NV; qq++) side[(((long)seg * 6 + 5) * NV + qq) * ncol + col] = 0.0;
}

// Side rows of every segment of a column into their cells (one thread per column and
// component, segments and rows in a fixed order): rows of neighbouring segments directly,
// ghost rows through the ghost rule's source (their signs were applied at the faces).
template <int D, int NC>
__global__ void __launch_bounds__(256)
fe_vjp_fix_yz(Grid g, const double *side, FieldsOut out, int nseg) {
  constexpr int NV = 5 + NC;
  const int nt = D == 1 ? g.nz : g.ny;
  const int n = D == 1 ? g.ny : g.nz;
  const long ncol = (long)g.nx * nt;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= ncol * NV) return;
  const long col = tid % ncol;
  const int qq = (int)(tid / ncol);
  const int i = (int)(col % g.nx), t = (int)(col / g.nx);
  const long rstride = D == 1 ? g.nx : (long)g.nx * g.ny;
  const long cbase = D == 1 ? i + (long)g.nx * g.ny * t : i + (long)g.nx * t;
  const long ncell = (long)g.nx * g.ny * g.nz;
  for (int seg = 0; seg < nseg; seg++) {
    const int c0 = (int)((long)seg * n / nseg), c1 = (int)((long)(seg + 1) * n / nseg);
    // when a segment's six side rows land in six distinct cells (segments of at least
    // three rows away from short columns), all its reads are issued together
    int src[6];
#pragma unroll
    for (int k = 0; k < 6; k++) src[k] = ghost_map(g, D, k < 3 ? c0 - 3 + k : c1 + (k - 3)).src;
    bool distinct = true;
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
      for (int b = a + 1; b < 6; b++) distinct = distinct && src[a] != src[b];
    if (distinct) {
      long cell[6];
      double add[6], old[6];
#pragma unroll
      for (int k = 0; k < 6; k++) {
        cell[k] = cbase + (long)src[k] * rstride;
        add[k] = side[(((long)seg * 6 + k) * NV + qq) * ncol + col];
        old[k] = qq < 5 ? out.w[qq][cell[k]] : out.chem[(long)(qq - 5) * ncell + cell[k]];
      }
#pragma unroll
      for (int k = 0; k < 6; k++) {
        if (qq < 5) out.w[qq][cell[k]] = old[k] + add[k];
        else out.chem[(long)(qq - 5) * ncell + cell[k]] = old[k] + add[k];
      }
      continue;
    }
#pragma unroll 1
    for (int k = 0; k < 6; k++) {
      const int e = k < 3 ? c0 - 3 + k : c1 + (k - 3);
      const int src = ghost_map(g, D, e).src;
      const long cell = cbase + (long)src * rstride;
      out_put(out, ncell, qq, cell, side[(((long)seg * 6 + k) * NV + qq) * ncol + col], true);
    }
  }
}

// Tracer layout: the C ABI holds tracers cell-major ([cell][nchem]); the sweeps read
// and write them tracer-major ([nchem][cell]) so that every tracer access of a wave is one
// contiguous 512-byte run. 256 cells per block through an LDS tile, both sides coalesced.
template <int NC>
__global__ void __launch_bounds__(256) cells_to_planes(const double *a, double *p, long ncell) {
  __shared__ double t[(NC > 0 ? NC : 1) * 257];
  const long base = (long)blockIdx.x * 256;
  const long lim = (ncell - base) * NC;
  for (int k = threadIdx.x; k < NC * 256; k += 256)
    if (k < lim) t[(k % NC) * 257 + k / NC] = a[base * NC + k];
  __syncthreads();
  const long cell = base + threadIdx.x;
  if (cell < ncell)
#pragma unroll
    for (int c = 0; c < NC; c++) p[(long)c * ncell + cell] = t[c * 257 + threadIdx.x];
}

template <int NC>
__global__ void __launch_bounds__(256) planes_to_cells(const double *p, double *a, long ncell) {
  __shared__ double t[(NC > 0 ? NC : 1) * 257];
  const long base = (long)blockIdx.x * 256;
  const long cell = base + threadIdx.x;
  if (cell < ncell)
#pragma unroll
    for (int c = 0; c < NC; c++) t[c * 257 + threadIdx.x] = p[(long)c * ncell + cell];
  __syncthreads();
  const long lim = (ncell - base) * NC;
  for (int k = threadIdx.x; k < NC * 256; k += 256)
    if (k < lim) a[base * NC + k] = t[(k % NC) * 257 + k / NC];
}

// Launch shape: one wave per SIMD over the device (4 per CU) where the grid allows.
struct VjpPlan {
  int lps;          // x-lines per wave
  int nseg[3];      // segments per y / z column (index 1, 2)
  size_t side;      // side-buffer doubles (the largest direction's)
};

static VjpPlan vjp_plan(const Grid &g) {
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const long target = 4L * (ncu > 0 ? ncu : 256);
  const long nv = 5 + g.nchem;
  VjpPlan p{};
  const long nlines = (long)g.ny * g.nz;
  p.lps = (int)((nlines + target - 1) / target);
  size_t side = (size_t)nlines * 12 * nv;
  for (int d = 1; d < 3; d++) {
    const long n = d == 1 ? g.ny : g.nz, nt = d == 1 ? g.nz : g.ny;
    const long waves = (long)((g.nx + 63) / 64) * nt;
    long ns = (target + waves - 1) / waves;
    const long maxseg = n >= 16 ? n / 8 : 1;
    if (ns > maxseg) ns = maxseg;
    if (ns < 1) ns = 1;
    p.nseg[d] = (int)ns;
    const size_t sd = (size_t)g.nx * nt * ns * 6 * nv;
    if (sd > side) side = sd;
  }
  p.side = side;
  return p;
}

// Scratch: the side buffer, then (with tracers) tracer planes of w and wbar.
template <int NC>
static int fe_vjp_run(const Grid &g, const Fields &w, const Fields &v, const FieldsOut &out,
                      double *buf, hipStream_t st) {
  const VjpPlan p = vjp_plan(g);
  const long nlines = (long)g.ny * g.nz;
  const long ncell = (long)g.nx * nlines;
  const int nxb = (g.nx + 63) / 64;
  constexpr long NV = 5 + NC;
  double *side = buf;
  Fields wp = w, vp = v;
  FieldsOut op = out;
  const unsigned tb = (unsigned)((ncell + 255) / 256);
  if constexpr (NC > 0) {
    double *pw = buf + p.side, *po = pw + NC * ncell;
    hipLaunchKernelGGL(cells_to_planes<NC>, dim3(tb), dim3(256), 0, st, w.chem, pw, ncell);
    wp.chem = pw;
    op.chem = po;
  }
  hipLaunchKernelGGL(fe_vjp_sweep_x<NC>, dim3((unsigned)((nlines + p.lps - 1) / p.lps)),
                     dim3(64), 0, st, g, wp, vp, op, side, p.lps, 0);
  hipLaunchKernelGGL(fe_vjp_fix_x<NC>, dim3((unsigned)((nlines * NV + 255) / 256)), dim3(256), 0,
                     st, g, (const double *)side, op);
  hipLaunchKernelGGL((fe_vjp_sweep_yz<1, NC>), dim3((unsigned)(nxb * g.nz * p.nseg[1])), dim3(64),
                     0, st, g, wp, vp, op, side, p.nseg[1], 1);
  hipLaunchKernelGGL((fe_vjp_fix_yz<1, NC>),
                     dim3((unsigned)(((long)g.nx * g.nz * NV + 255) / 256)), dim3(256), 0, st, g,
                     (const double *)side, op, p.nseg[1]);
  hipLaunchKernelGGL((fe_vjp_sweep_yz<2, NC>), dim3((unsigned)(nxb * g.ny * p.nseg[2])), dim3(64),
                     0, st, g, wp, vp, op, side, p.nseg[2], 1);
  hipLaunchKernelGGL((fe_vjp_fix_yz<2, NC>),
                     dim3((unsigned)(((long)g.nx * g.ny * NV + 255) / 256)), dim3(256), 0, st, g,
                     (const double *)side, op, p.nseg[2]);
  if constexpr (NC > 0)
    hipLaunchKernelGGL(planes_to_cells<NC>, dim3(tb), dim3(256), 0, st, (const double *)op.chem,
                       out.chem, ncell);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

size_t fe_vjp_buffer_doubles(const Grid &g) {
  return vjp_plan(g).side + (size_t)2 * g.nchem * g.nx * g.ny * g.nz;
}

int fe_vjp_launch(const Grid &g, const Fields &w, const Fields &v, const FieldsOut &out,
                  double *buf, hipStream_t st) {
  switch (g.nchem) {
    case 0: return fe_vjp_run<0>(g, w, v, out, buf, st);
    case 1: return fe_vjp_run<1>(g, w, v, out, buf, st);
    case 2: return fe_vjp_run<2>(g, w, v, out, buf, st);
    case 4: return fe_vjp_run<4>(g, w, v, out, buf, st);
    case 6: return fe_vjp_run<6>(g, w, v, out, buf, st);
    case 10: return fe_vjp_run<10>(g, w, v, out, buf, st);
    default: return MVD_UNSUPPORTED;
  }
}

}  // namespace mvd
