// Vector-Jacobian product of fe(): wbar = (d fe / d w)^T v.
//
// fe(w) = -div F with every face flux a function of its 6-point stencil, so
//   Fbar(face between cells c-1 | c) = (v_c - v_{c-1}) / delta   (no cell: 0)
// and each face's stencil adjoint sbar = (dF/ds)^T Fbar goes back to the stencil's source
// cells (ghost points are +-copies of interior cells). Per direction:
//   fe_vjp_faces  - one thread per face: recompute the face flux (the reference recipe,
//                   utilities.cpp:270-479) and sweep it in reverse (WENO5-JS weights,
//                   Lax-Friedrichs split, characteristic projections, Roe average and
//                   eigenvectors, point fluxes, wave speeds, EOS); the 6 x NV stencil
//                   adjoint (ghost signs applied) goes to a face buffer;
//   fe_vjp_gather - one thread per cell sums the contributions of the six faces whose
//                   stencils hold it, plus those of ghost points that copy it near a
//                   physical or periodic boundary, into wbar (first direction stores, the
//                   others add). No atomics: the result is deterministic.
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

// 1/x from the hardware estimate (relative error below 2^-22) and two Newton steps, to
// within an ulp or two of the IEEE quotient for the finite, nonzero arguments used here;
// about a third of the instructions of a correctly rounded division.
__device__ inline double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// WENO5-JS reconstruction from f[0..4] (side 0: the f+ formula over points i-3..i+1,
// side 1: the mirrored f- formula over points i-2..i+2) and its gradient g = dr/df.
__device__ inline double weno5_grad(const double (&f)[5], int side, double (&g)[5]) {
  const double A1 = f[2] - 2.0 * f[3] + f[4], B1 = 3.0 * f[2] - 4.0 * f[3] + f[4];
  const double A2 = f[1] - 2.0 * f[2] + f[3], B2 = f[1] - f[3];
  const double A3 = f[0] - 2.0 * f[1] + f[2], B3 = f[0] - 4.0 * f[1] + 3.0 * f[2];
  const double b1 = W_BC * A1 * A1 + 0.25 * B1 * B1;
  const double b2 = W_BC * A2 * A2 + 0.25 * B2 * B2;
  const double b3 = W_BC * A3 * A3 + 0.25 * B3 * B3;
  const double d1 = side == 0 ? 0.3 : 0.1, d3 = side == 0 ? 0.1 : 0.3;
  const double e1 = W_EPS + b1, e2 = W_EPS + b2, e3 = W_EPS + b3;
  const double ie1 = rcp64(e1), ie2 = rcp64(e2), ie3 = rcp64(e3);
  const double w1 = d1 * ie1 * ie1, w2 = 0.6 * ie2 * ie2, w3 = d3 * ie3 * ie3;
  double q1, q2, q3;
  double dq1[5] = {0, 0, 0, 0, 0}, dq2[5] = {0, 0, 0, 0, 0}, dq3[5] = {0, 0, 0, 0, 0};
  if (side == 0) {
    q1 = W_C13 * f[2] + W_C56 * f[3] - W_C16 * f[4];
    q2 = -W_C16 * f[1] + W_C56 * f[2] + W_C13 * f[3];
    q3 = W_C13 * f[0] - W_C76 * f[1] + W_C116 * f[2];
    dq1[2] = W_C13; dq1[3] = W_C56; dq1[4] = -W_C16;
    dq2[1] = -W_C16; dq2[2] = W_C56; dq2[3] = W_C13;
    dq3[0] = W_C13; dq3[1] = -W_C76; dq3[2] = W_C116;
  } else {
    q1 = W_C116 * f[2] - W_C76 * f[3] + W_C13 * f[4];
    q2 = W_C13 * f[1] + W_C56 * f[2] - W_C16 * f[3];
    q3 = -W_C16 * f[0] + W_C56 * f[1] + W_C13 * f[2];
    dq1[2] = W_C116; dq1[3] = -W_C76; dq1[4] = W_C13;
    dq2[1] = W_C13; dq2[2] = W_C56; dq2[3] = -W_C16;
    dq3[0] = -W_C16; dq3[1] = W_C56; dq3[2] = W_C13;
  }
  const double iS = rcp64(w1 + w2 + w3);
  const double r = (q1 * w1 + q2 * w2 + q3 * w3) * iS;
  // dw_i/df = -2 w_i / e_i * db_i/df
  const double k1 = -2.0 * w1 * ie1 * (q1 - r) * iS;
  const double k2 = -2.0 * w2 * ie2 * (q2 - r) * iS;
  const double k3 = -2.0 * w3 * ie3 * (q3 - r) * iS;
  const double db1[5] = {0.0, 0.0, 2.0 * W_BC * A1 + 1.5 * B1, -4.0 * W_BC * A1 - 2.0 * B1,
                         2.0 * W_BC * A1 + 0.5 * B1};
  const double db2[5] = {0.0, 2.0 * W_BC * A2 + 0.5 * B2, -4.0 * W_BC * A2,
                         2.0 * W_BC * A2 - 0.5 * B2, 0.0};
  const double db3[5] = {2.0 * W_BC * A3 + 0.5 * B3, -4.0 * W_BC * A3 - 2.0 * B3,
                         2.0 * W_BC * A3 + 1.5 * B3, 0.0, 0.0};
#pragma unroll
  for (int j = 0; j < 5; j++)
    g[j] = (w1 * dq1[j] + w2 * dq2[j] + w3 * dq3[j]) * iS + k1 * db1[j] + k2 * db2[j] +
           k3 * db3[j];
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
    P.ir[l] = rcp64(s[l][0]);
    P.p[l] = gm * (s[l][4] - (s[l][1] * s[l][1] + s[l][2] * s[l][2] + s[l][3] * s[l][3]) * 0.5 * P.ir[l]);
    P.ul[l] = s[l][1] * P.ir[l];
    const double a = fabs(P.ul[l]) + sun_sqrt(gamma * P.p[l] * P.ir[l]);
    if (P.alpha < a) { P.alpha = a; P.lstar = l; }
  }
}

// One tracer: t[6] the stencil values, Fbc its flux cotangent; tb[6] receives the tracer
// adjoint, ulb and alphab accumulate.
__device__ inline void tracer_vjp(const double (&t)[6], double Fbc, const FacePrim &P,
                                  double (&tb)[6], double (&ulb)[6], double &alphab) {
#pragma unroll
  for (int l = 0; l < 6; l++) tb[l] = 0.0;
#pragma unroll
  for (int sd = 0; sd < 2; sd++) {
    double fa[5], g[5];
#pragma unroll
    for (int l = 0; l < 5; l++) {
      const int pt = l + sd;
      fa[l] = sd == 0 ? 0.5 * (P.ul[pt] * t[pt] + P.alpha * t[pt])
                      : 0.5 * (P.ul[pt] * t[pt] - P.alpha * t[pt]);
    }
    (void)weno5_grad(fa, sd, g);
    const double sg = sd == 0 ? 0.5 : -0.5;
#pragma unroll
    for (int l = 0; l < 5; l++) {
      const int pt = l + sd;
      const double fsb = Fbc * g[l];
      // fs = 0.5 (ul t +- alpha t)
      ulb[pt] += 0.5 * fsb * t[pt];
      tb[pt] += 0.5 * fsb * P.ul[pt] + sg * P.alpha * fsb;
      alphab += sg * fsb * t[pt];
    }
  }
}

__device__ inline void fluid_vjp(const double (&s)[6][5], double gamma, const FacePrim &P,
                                 const double (&Fb)[5], double (&ulb)[6], double alphab,
                                 double (&sb)[6][5]) {
  const double gm = gamma - 1.0;
  const double alpha = P.alpha;
  const double sL = sun_sqrt(s[2][0]), sR = sun_sqrt(s[3][0]);
  const double isL = rcp64(sL), isR = rcp64(sR);
  const double sbar = 0.5 * (sL + sR), isbar = rcp64(sbar);
  const double u = 0.5 * (s[2][1] * isL + s[3][1] * isR) * isbar;
  const double v = 0.5 * (s[2][2] * isL + s[3][2] * isR) * isbar;
  const double w = 0.5 * (s[2][3] * isL + s[3][3] * isR) * isbar;
  const double H = 0.5 * ((P.p[2] + s[2][4]) * isL + (P.p[3] + s[3][4]) * isR) * isbar;
  const double qsq = u * u + v * v + w * w;
  const double cs = gm * (H - 0.5 * qsq);
  const double ci = rcp64(cs);
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
      const double Xb = 0.5 * alphab * rcp64(sun_sqrt(X));
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

// direction-local component a -> global field index
__device__ inline int gcomp(int d, int a) {
  if (a == 1) return 1 + d;
  if (a == 1 + d) return 1;
  return a;
}

__device__ inline long face_index(const Grid &g, int d, long i, long j, long k) {
  return d == 0 ? i + (long)(g.nx + 1) * (j + (long)g.ny * k)
       : d == 1 ? i + (long)g.nx * (j + (long)(g.ny + 1) * k)
                : i + (long)g.nx * (j + (long)g.ny * k);
}

__device__ inline double field(const Fields &f, int nc, int q, long cell) {
  return q < 5 ? f.w[q][cell] : f.chem[(long)nc * cell + (q - 5)];
}

}  // namespace

template <int NC>
__global__ void __launch_bounds__(256)
fe_vjp_faces(Grid g, Fields w, Fields vc, int d, double *buf, long nf) {
  constexpr int NV = 5 + NC;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nf) return;
  const long ex = g.nx + (d == 0), ey = g.ny + (d == 1);
  const int i = (int)(tid % ex), j = (int)((tid / ex) % ey), k = (int)(tid / (ex * ey));
  const int n = extent(g, d);
  const int c0 = d == 0 ? i : d == 1 ? j : k;
  const double idelta = 1.0 / (d == 0 ? g.dx : d == 1 ? g.dy : g.dz);
  const long cr = c0 < n ? cell_idx(g, i, j, k) : -1;
  const long cl = c0 > 0 ? cell_idx(g, i - (d == 0), j - (d == 1), k - (d == 2)) : -1;
  long cell[6];
  double sa[6], sn[6];
  double s[6][5];
#pragma unroll
  for (int l = 0; l < 6; l++) {
    const GhostMap m = ghost_map(g, d, c0 - 3 + l);
    int ii = i, jj = j, kk = k;
    if (d == 0) ii = m.src; else if (d == 1) jj = m.src; else kk = m.src;
    cell[l] = cell_idx(g, ii, jj, kk);
    sa[l] = m.sgn_all;
    sn[l] = m.sgn_norm;
#pragma unroll
    for (int a = 0; a < 5; a++)
      s[l][a] = field(w, NC, gcomp(d, a), cell[l]) * (a == 1 ? sn[l] : sa[l]);
  }
  FacePrim P;
  face_prim(s, g.gamma, P);
  double ulb[6] = {0, 0, 0, 0, 0, 0};
  double alphab = 0.0;
  // tracers, one at a time
#pragma unroll 1
  for (int c = 0; c < NC; c++) {
    double t[6], tb[6];
#pragma unroll
    for (int l = 0; l < 6; l++) t[l] = w.chem[(long)NC * cell[l] + c] * sa[l];
    const double r = cr >= 0 ? vc.chem[(long)NC * cr + c] : 0.0;
    const double lv = cl >= 0 ? vc.chem[(long)NC * cl + c] : 0.0;
    tracer_vjp(t, (r - lv) * idelta, P, tb, ulb, alphab);
#pragma unroll
    for (int l = 0; l < 6; l++) buf[(long)(l * NV + 5 + c) * nf + tid] = tb[l] * sa[l];
  }
  double Fb[5];
#pragma unroll
  for (int a = 0; a < 5; a++) {
    const int q = gcomp(d, a);
    const double r = cr >= 0 ? vc.w[q][cr] : 0.0;
    const double lv = cl >= 0 ? vc.w[q][cl] : 0.0;
    Fb[a] = (r - lv) * idelta;
  }
  double sb[6][5];
  fluid_vjp(s, g.gamma, P, Fb, ulb, alphab, sb);
#pragma unroll
  for (int l = 0; l < 6; l++)
#pragma unroll
    for (int a = 0; a < 5; a++)
      buf[(long)(l * NV + a) * nf + tid] = sb[l][a] * (a == 1 ? sn[l] : sa[l]);
}

template <int NC>
__global__ void __launch_bounds__(256)
fe_vjp_gather(Grid g, int d, const double *buf, long nf, FieldsOut out, int accumulate) {
  constexpr int NV = 5 + NC;
  const long ncell = (long)g.nx * g.ny * g.nz;
  const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= ncell) return;
  const int i = (int)(cell % g.nx), j = (int)((cell / g.nx) % g.ny),
            k = (int)(cell / ((long)g.nx * g.ny));
  const int n = extent(g, d);
  const int c = d == 0 ? i : d == 1 ? j : k;
  auto fidx = [&](int f) {
    return face_index(g, d, d == 0 ? f : i, d == 1 ? f : j, d == 2 ? f : k);
  };
  double acc[NV];
#pragma unroll
  for (int a = 0; a < NV; a++) acc[a] = 0.0;
  // faces f = c+3-l whose point l is this cell
#pragma unroll
  for (int l = 0; l < 6; l++) {
    const int f = c + 3 - l;
    if (f >= 0 && f <= n) {
      const long fi = fidx(f);
#pragma unroll
      for (int a = 0; a < NV; a++) acc[a] += buf[(long)(l * NV + a) * nf + fi];
    }
  }
  // ghost points that copy this cell (faces 0..2 and n-2..n)
  if (c < 3 || c >= n - 3) {
    for (int f = 0; f <= n; f++) {
      if (f > 2 && f < n - 2) { f = n - 3; continue; }
      const long fi = fidx(f);
      for (int l = 0; l < 6; l++) {
        const int e = f - 3 + l;
        if (e >= 0 && e < n) continue;
        if (ghost_map(g, d, e).src != c) continue;
#pragma unroll
        for (int a = 0; a < NV; a++) acc[a] += buf[(long)(l * NV + a) * nf + fi];
      }
    }
  }
#pragma unroll
  for (int a = 0; a < NV; a++) {
    const int q = gcomp(d, a);
    double *dst = q < 5 ? &out.w[q][cell] : &out.chem[(long)NC * cell + (q - 5)];
    *dst = accumulate ? *dst + acc[a] : acc[a];
  }
}

template <int NC>
static int fe_vjp_run(const Grid &g, const Fields &w, const Fields &v, const FieldsOut &out,
                      double *buf, hipStream_t st) {
  const long ncell = (long)g.nx * g.ny * g.nz;
  for (int d = 0; d < 3; d++) {
    const long nf = (long)(g.nx + (d == 0)) * (g.ny + (d == 1)) * (g.nz + (d == 2));
    hipLaunchKernelGGL(fe_vjp_faces<NC>, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, st,
                       g, w, v, d, buf, nf);
    hipLaunchKernelGGL(fe_vjp_gather<NC>, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0,
                       st, g, d, (const double *)buf, nf, out, d > 0 ? 1 : 0);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

size_t fe_vjp_buffer_doubles(const Grid &g) {
  const long nf = (long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  return (size_t)6 * (5 + g.nchem) * (size_t)nf;
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
