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
  const double w1 = d1 / (e1 * e1), w2 = 0.6 / (e2 * e2), w3 = d3 / (e3 * e3);
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
  const double S = w1 + w2 + w3, iS = 1.0 / S;
  const double r = (q1 * w1 + q2 * w2 + q3 * w3) * iS;
  // dw_i/df = -2 w_i / e_i * db_i/df
  const double k1 = -2.0 * w1 / e1 * (q1 - r) * iS;
  const double k2 = -2.0 * w2 / e2 * (q2 - r) * iS;
  const double k3 = -2.0 * w3 / e3 * (q3 - r) * iS;
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

// sb = (dF/ds)^T Fb for one face; s, Fb and sb in direction-local order (rho, m_normal,
// m_t1, m_t2, et, tracers...).
template <int NC>
__device__ inline void face_flux_vjp(const double (&s)[6][5 + NC], double gamma,
                                     const double (&Fb)[5 + NC], double (&sb)[6][5 + NC]) {
  constexpr int NV = 5 + NC;
  const double gm = gamma - 1.0;
  // ---- forward ----
  double p[6], ul[6], cl[6], X[6];
#pragma unroll
  for (int l = 0; l < 6; l++) {
    p[l] = gm * (s[l][4] - (s[l][1] * s[l][1] + s[l][2] * s[l][2] + s[l][3] * s[l][3]) * 0.5 / s[l][0]);
    ul[l] = s[l][1] / s[l][0];
    X[l] = gamma * p[l] / s[l][0];
    cl[l] = sun_sqrt(X[l]);
  }
  double alpha = 0.0;
  int lstar = -1;
#pragma unroll
  for (int l = 0; l < 6; l++) {
    const double a = fabs(ul[l]) + cl[l];
    if (alpha < a) { alpha = a; lstar = l; }
  }
  const double sL = sun_sqrt(s[2][0]), sR = sun_sqrt(s[3][0]);
  const double sbar = 0.5 * (sL + sR);
  const double Au = s[2][1] / sL + s[3][1] / sR, Av = s[2][2] / sL + s[3][2] / sR,
               Aw = s[2][3] / sL + s[3][3] / sR;
  const double AH = (p[2] + s[2][4]) / sL + (p[3] + s[3][4]) / sR;
  const double u = 0.5 * Au / sbar, v = 0.5 * Av / sbar, w = 0.5 * Aw / sbar;
  const double H = 0.5 * AH / sbar;
  const double qsq = u * u + v * v + w * w;
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
  // point flux of component c at point l
  auto flux = [&](int l, int c) -> double {
    if (c == 0) return s[l][1];
    if (c == 1) return ul[l] * s[l][1] + p[l];
    if (c == 4) return ul[l] * (s[l][4] + p[l]);
    return ul[l] * s[l][c];
  };
  // split flux of side sd at window position l (point l for sd 0, l+1 for sd 1)
  auto split = [&](int sd, int l, int c) -> double {
    const int pt = l + sd;
    return sd == 0 ? 0.5 * (flux(pt, c) + alpha * s[pt][c]) : 0.5 * (flux(pt, c) - alpha * s[pt][c]);
  };
  // fluid: characteristic values, reconstruction and gradients for both sides
  double gq[2][5][5];    // [side][char comp][window position]
  double ff[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int sd = 0; sd < 2; sd++) {
    double fs[5][5];
#pragma unroll
    for (int l = 0; l < 5; l++)
#pragma unroll
      for (int b = 0; b < 5; b++) fs[l][b] = split(sd, l, b);
#pragma unroll
    for (int a = 0; a < 5; a++) {
      double fa[5];
#pragma unroll
      for (int l = 0; l < 5; l++)
        fa[l] = LV[a][0] * fs[l][0] + LV[a][1] * fs[l][1] + LV[a][2] * fs[l][2] +
                LV[a][3] * fs[l][3] + LV[a][4] * fs[l][4];
      ff[a] += weno5_grad(fa, sd, gq[sd][a]);
    }
  }
  // ---- reverse ----
#pragma unroll
  for (int l = 0; l < 6; l++)
#pragma unroll
    for (int c = 0; c < NV; c++) sb[l][c] = 0.0;
  double pb[6] = {0, 0, 0, 0, 0, 0}, ulb[6] = {0, 0, 0, 0, 0, 0};
  double alphab = 0.0;
  // flux adjoint of point pt, component c, with weight fb
  auto flux_adj = [&](int pt, int c, double fb) {
    if (c == 0) { sb[pt][1] += fb; return; }
    if (c == 1) { ulb[pt] += fb * s[pt][1]; sb[pt][1] += fb * ul[pt]; pb[pt] += fb; return; }
    if (c == 4) {
      ulb[pt] += fb * (s[pt][4] + p[pt]);
      sb[pt][4] += fb * ul[pt];
      pb[pt] += fb * ul[pt];
      return;
    }
    ulb[pt] += fb * s[pt][c];
    sb[pt][c] += fb * ul[pt];
  };
  // split adjoint: fs = 0.5 (flux +- alpha s)
  auto split_adj = [&](int sd, int l, int c, double fsb) {
    const int pt = l + sd;
    const double sg = sd == 0 ? 0.5 : -0.5;
    flux_adj(pt, c, 0.5 * fsb);
    sb[pt][c] += sg * alpha * fsb;
    alphab += sg * fsb * s[pt][c];
  };
  // back-projection F = RV ff
  double ffb[5];
#pragma unroll
  for (int b = 0; b < 5; b++)
    ffb[b] = RV[0][b] * Fb[0] + RV[1][b] * Fb[1] + RV[2][b] * Fb[2] + RV[3][b] * Fb[3] +
             RV[4][b] * Fb[4];
  // Roe-state adjoints
  double ub = 0.0, vb = 0.0, wb = 0.0, Hb = 0.0, csb = 0.0, cib = 0.0, qsqb = 0.0;
  {
    // RV entries (RVb[a][b] = Fb[a] ff[b])
    const double R10 = Fb[1] * ff[0], R13 = Fb[1] * ff[3], R14 = Fb[1] * ff[4];
    ub += R10 + R13 + R14;
    csb += R14 - R10;
    vb += Fb[2] * (ff[0] + ff[3] + ff[4]);
    wb += Fb[3] * (ff[0] + ff[3] + ff[4]);
    const double R40 = Fb[4] * ff[0], R41 = Fb[4] * ff[1], R42 = Fb[4] * ff[2],
                 R43 = Fb[4] * ff[3], R44 = Fb[4] * ff[4];
    Hb += R40 + R44;
    ub += (R44 - R40) * cs;
    csb += (R44 - R40) * u;
    vb += R41;
    wb += R42;
    qsqb += 0.5 * R43;
  }
  // fluid characteristic reverse: fpb = ffb[a] * g; fsb = LV^T fpb; LVb += fpb fs^T
  double LVb[5][5];
#pragma unroll
  for (int a = 0; a < 5; a++)
#pragma unroll
    for (int b = 0; b < 5; b++) LVb[a][b] = 0.0;
#pragma unroll
  for (int sd = 0; sd < 2; sd++) {
#pragma unroll
    for (int l = 0; l < 5; l++) {
      double fpb[5], fs[5];
#pragma unroll
      for (int a = 0; a < 5; a++) fpb[a] = ffb[a] * gq[sd][a][l];
#pragma unroll
      for (int b = 0; b < 5; b++) fs[b] = split(sd, l, b);
#pragma unroll
      for (int a = 0; a < 5; a++)
#pragma unroll
        for (int b = 0; b < 5; b++) LVb[a][b] += fpb[a] * fs[b];
#pragma unroll
      for (int b = 0; b < 5; b++) {
        const double fsb = LV[0][b] * fpb[0] + LV[1][b] * fpb[1] + LV[2][b] * fpb[2] +
                           LV[3][b] * fpb[3] + LV[4][b] * fpb[4];
        split_adj(sd, l, b, fsb);
      }
    }
  }
  // tracers: not projected; reconstruction per tracer and side
#pragma unroll
  for (int c = 5; c < NV; c++) {
#pragma unroll
    for (int sd = 0; sd < 2; sd++) {
      double fa[5], g[5];
#pragma unroll
      for (int l = 0; l < 5; l++) fa[l] = split(sd, l, c);
      (void)weno5_grad(fa, sd, g);
#pragma unroll
      for (int l = 0; l < 5; l++) split_adj(sd, l, c, Fb[c] * g[l]);
    }
  }
  // LV entries -> Roe state
  {
    const double L00 = LVb[0][0], L01 = LVb[0][1], L02 = LVb[0][2], L03 = LVb[0][3], L04 = LVb[0][4];
    const double L40 = LVb[4][0], L41 = LVb[4][1], L42 = LVb[4][2], L43 = LVb[4][3], L44 = LVb[4][4];
    const double L30 = LVb[3][0], L31 = LVb[3][1], L32 = LVb[3][2], L33 = LVb[3][3], L34 = LVb[3][4];
    cib += 0.5 * (u + 0.5 * gm * qsq) * L00 - 0.5 * (gm * u + 1.0) * L01 -
           0.5 * v * gm * (L02 + LVb[4][2]) - 0.5 * w * gm * (L03 + L43) + 0.5 * gm * (L04 + L44) -
           gm * (qsq - H) * L30 + gm * (u * L31 + v * L32 + w * L33) - gm * L34 -
           0.5 * (u - 0.5 * gm * qsq) * L40 - 0.5 * (gm * u - 1.0) * L41;
    ub += 0.5 * ci * L00 - 0.5 * ci * gm * L01 + gm * ci * L31 - 0.5 * ci * L40 - 0.5 * ci * gm * L41;
    qsqb += 0.25 * ci * gm * L00 - gm * ci * L30 + 0.25 * ci * gm * L40;
    vb += -0.5 * gm * ci * (L02 + L42) - LVb[1][0] + gm * ci * L32;
    wb += -0.5 * gm * ci * (L03 + L43) - LVb[2][0] + gm * ci * L33;
    Hb += gm * ci * L30;
  }
  csb += -cib * ci * ci;
  Hb += gm * csb;
  qsqb += -0.5 * gm * csb;
  ub += 2.0 * u * qsqb;
  vb += 2.0 * v * qsqb;
  wb += 2.0 * w * qsqb;
  // u, v, w, H = 0.5 A / sbar
  double sLb = 0.0, sRb = 0.0, sbarb = 0.0;
  {
    const double Aub = 0.5 * ub / sbar, Avb = 0.5 * vb / sbar, Awb = 0.5 * wb / sbar,
                 AHb = 0.5 * Hb / sbar;
    sbarb -= (ub * u + vb * v + wb * w + Hb * H) / sbar;
    sb[2][1] += Aub / sL; sb[3][1] += Aub / sR;
    sb[2][2] += Avb / sL; sb[3][2] += Avb / sR;
    sb[2][3] += Awb / sL; sb[3][3] += Awb / sR;
    sb[2][4] += AHb / sL; sb[3][4] += AHb / sR;
    pb[2] += AHb / sL; pb[3] += AHb / sR;
    sLb -= (Aub * s[2][1] + Avb * s[2][2] + Awb * s[2][3] + AHb * (p[2] + s[2][4])) / (sL * sL);
    sRb -= (Aub * s[3][1] + Avb * s[3][2] + Awb * s[3][3] + AHb * (p[3] + s[3][4])) / (sR * sR);
  }
  sLb += 0.5 * sbarb;
  sRb += 0.5 * sbarb;
  if (s[2][0] > 0.0) sb[2][0] += sLb / (2.0 * sL);
  if (s[3][0] > 0.0) sb[3][0] += sRb / (2.0 * sR);
  // Lax-Friedrichs speed: alpha = a_lstar = |u_l| + sqrt(gamma p_l / rho_l)
  if (lstar >= 0) {
#pragma unroll
    for (int l = 0; l < 6; l++)
      if (l == lstar) {
        ulb[l] += (ul[l] < 0.0 ? -alphab : alphab);
        if (X[l] > 0.0) {
          const double Xb = alphab / (2.0 * cl[l]);
          pb[l] += Xb * gamma / s[l][0];
          sb[l][0] -= Xb * X[l] / s[l][0];
        }
      }
  }
  // ul = m / rho, p = gm (et - |m|^2 / (2 rho))
#pragma unroll
  for (int l = 0; l < 6; l++) {
    sb[l][1] += ulb[l] / s[l][0];
    sb[l][0] -= ulb[l] * ul[l] / s[l][0];
    const double K = s[l][1] * s[l][1] + s[l][2] * s[l][2] + s[l][3] * s[l][3];
    sb[l][4] += gm * pb[l];
    const double Kb = -gm * pb[l] * 0.5 / s[l][0];
    sb[l][1] += 2.0 * s[l][1] * Kb;
    sb[l][2] += 2.0 * s[l][2] * Kb;
    sb[l][3] += 2.0 * s[l][3] * Kb;
    sb[l][0] += gm * pb[l] * K * 0.5 / (s[l][0] * s[l][0]);
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
  const double delta = d == 0 ? g.dx : d == 1 ? g.dy : g.dz;
  // cotangent of the face flux, direction-local order
  double Fb[NV];
  {
    const long cr = c0 < n ? cell_idx(g, i, j, k) : -1;
    long cl = -1;
    if (c0 > 0) cl = cell_idx(g, i - (d == 0), j - (d == 1), k - (d == 2));
#pragma unroll
    for (int a = 0; a < NV; a++) {
      const int q = gcomp(d, a);
      const double r = cr >= 0 ? field(vc, NC, q, cr) : 0.0;
      const double l = cl >= 0 ? field(vc, NC, q, cl) : 0.0;
      Fb[a] = (r - l) / delta;
    }
  }
  double s[6][NV], sg[6][2];
#pragma unroll
  for (int l = 0; l < 6; l++) {
    const GhostMap m = ghost_map(g, d, c0 - 3 + l);
    int ii = i, jj = j, kk = k;
    if (d == 0) ii = m.src; else if (d == 1) jj = m.src; else kk = m.src;
    const long c = cell_idx(g, ii, jj, kk);
#pragma unroll
    for (int a = 0; a < NV; a++)
      s[l][a] = field(w, NC, gcomp(d, a), c) * (a == 1 ? m.sgn_norm : m.sgn_all);
    sg[l][0] = m.sgn_all;
    sg[l][1] = m.sgn_norm;
  }
  double sb[6][NV];
  face_flux_vjp<NC>(s, g.gamma, Fb, sb);
#pragma unroll
  for (int l = 0; l < 6; l++)
#pragma unroll
    for (int a = 0; a < NV; a++)
      buf[(long)(l * NV + a) * nf + tid] = sb[l][a] * (a == 1 ? sg[l][1] : sg[l][0]);
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
