// Oracle: WENO5 face flux, stencil gather with ghost rules, EOS -- TEST INFRASTRUCTURE.
// See orc_common.hpp for the role of this directory.
#pragma once
#include "orc_common.hpp"

namespace orc {

// EulerData::eos, euler3D.hpp:1383-1388:  p = (gamma-1)*(et - (mx^2+my^2+mz^2)*0.5/rho)
template <class R>
inline R eos(R rho, R mx, R my, R mz, R et, R gamma) {
  return (gamma - R(1.0)) * (et - (mx * mx + my * my + mz * mz) * R(0.5) / rho);
}

// EulerData::legal_state, euler3D.hpp:1405-1414 (fail bits 1/2/4).
inline int legal_state(double rho, double mx, double my, double mz, double et, double gamma) {
  int f = 0;
  if (!(rho > 0.0)) f += 1;
  if (!(et > 0.0)) f += 2;
  if (!(eos(rho, mx, my, mz, et, gamma) > 0.0)) f += 4;
  return f;
}

inline double sq(double x) { return x * x; }
inline Cnt sq(Cnt x) { return x * x; }

// std::max as used by the reference (returns the first argument unless a < b).
template <class R> inline R rmax(R a, R b) { return (a < b) ? b : a; }

// One face flux, following face_flux (utilities.cpp:270-479) operation by operation.
// s[6][nv] is the 6-point stencil w(i-3..i+2) along direction dir; it is modified
// in place (the momentum swap), exactly as the reference modifies w1d.
template <class R>
void face_flux(R (&s)[6][NVMAX], int dir, R *fface, R gamma, int nchem) {
  const int nv = 5 + nchem;
  const R bc = R(1.083333333333333333333333333333333333333);   // 13/12
  const R eps = R(1e-6);
  const R c13 = R(0.3333333333333333333333333333333333333333);
  const R c56 = R(0.8333333333333333333333333333333333333333);
  const R c16 = R(0.1666666666666666666666666666666666666667);
  const R c76 = R(1.166666666666666666666666666666666666667);
  const R c116 = R(1.833333333333333333333333333333333333333);
  const R half = R(0.5), one = R(1.0), zero = R(0.0), fourth = R(0.25);

  // direction-local momentum ordering (utilities.cpp:283-285)
  if (dir > 0)
    for (int l = 0; l < 6; l++) std::swap(s[l][1], s[l][1 + dir]);

  R p[6];
  for (int l = 0; l < 6; l++) p[l] = eos(s[l][0], s[l][1], s[l][2], s[l][3], s[l][4], gamma);

  // Roe average at the face (utilities.cpp:298-304)
  const R sL = sun_sqrt(s[2][0]);
  const R sR = sun_sqrt(s[3][0]);
  const R sbar = half * (sL + sR);
  const R u = half * (s[2][1] / sL + s[3][1] / sR) / sbar;
  const R v = half * (s[2][2] / sL + s[3][2] / sR) / sbar;
  const R w = half * (s[2][3] / sL + s[3][3] / sR) / sbar;
  const R H = half * ((p[2] + s[2][4]) / sL + (p[3] + s[3][4]) / sR) / sbar;

  // eigenvectors (utilities.cpp:307-364); "csnd" here is (gamma-1)(H-q^2/2)
  const R qsq = u * u + v * v + w * w;
  const R gm = gamma - one;
  const R cs = gm * (H - half * qsq);
  const R ci = one / cs;
  R RV[5][5], LV[5][5];
  for (int a = 0; a < 5; a++)
    for (int b = 0; b < 5; b++) { RV[a][b] = zero; LV[a][b] = zero; }
  RV[0][0] = one;         RV[0][3] = one;          RV[0][4] = one;
  RV[1][0] = u - cs;      RV[1][3] = u;            RV[1][4] = u + cs;
  RV[2][0] = v;           RV[2][1] = one;          RV[2][3] = v;   RV[2][4] = v;
  RV[3][0] = w;           RV[3][2] = one;          RV[3][3] = w;   RV[3][4] = w;
  RV[4][0] = H - u * cs;  RV[4][1] = v;            RV[4][2] = w;
  RV[4][3] = half * qsq;  RV[4][4] = H + u * cs;
  const R mh = R(-0.5);
  LV[0][0] = half * ci * (u + half * gm * qsq);
  LV[0][1] = mh * ci * (gm * u + one);
  LV[0][2] = mh * v * gm * ci;
  LV[0][3] = mh * w * gm * ci;
  LV[0][4] = half * gm * ci;
  LV[1][0] = -v;  LV[1][2] = one;
  LV[2][0] = -w;  LV[2][3] = one;
  LV[3][0] = (-gm) * ci * (qsq - H);
  LV[3][1] = u * gm * ci;
  LV[3][2] = v * gm * ci;
  LV[3][3] = w * gm * ci;
  LV[3][4] = (-gm) * ci;
  LV[4][0] = mh * ci * (u - half * gm * qsq);
  LV[4][1] = mh * ci * (gm * u - one);
  LV[4][2] = mh * v * gm * ci;
  LV[4][3] = mh * w * gm * ci;
  LV[4][4] = half * gm * ci;

  // point fluxes and Lax-Friedrichs speed over the stencil (utilities.cpp:368-380)
  R flux[6][NVMAX];
  R alpha = zero;
  for (int l = 0; l < 6; l++) {
    const R ul = s[l][1] / s[l][0];
    flux[l][0] = s[l][1];
    flux[l][1] = ul * s[l][1] + p[l];
    flux[l][2] = ul * s[l][2];
    flux[l][3] = ul * s[l][3];
    flux[l][4] = ul * (s[l][4] + p[l]);
    for (int c = 0; c < nchem; c++) flux[l][5 + c] = ul * s[l][5 + c];
    const R cl = sun_sqrt(gamma * p[l] / s[l][0]);
    alpha = rmax(alpha, fabs(ul) + cl);
  }

  R fs[5][NVMAX], fp[5][NVMAX], ff[NVMAX];
  for (int side = 0; side < 2; side++) {
    // split flux: f+ over points 0..4, f- over points 1..5 (utilities.cpp:386-388, 429-431)
    for (int l = 0; l < 5; l++)
      for (int c = 0; c < nv; c++)
        fs[l][c] = (side == 0) ? half * (flux[l][c] + alpha * s[l][c])
                               : half * (flux[l + 1][c] - alpha * s[l + 1][c]);
    // characteristic projection of the fluid part, tracers copied (391-396, 435-440)
    for (int l = 0; l < 5; l++) {
      for (int a = 0; a < 5; a++)
        fp[l][a] = LV[a][0] * fs[l][0] + LV[a][1] * fs[l][1] + LV[a][2] * fs[l][2]
                 + LV[a][3] * fs[l][3] + LV[a][4] * fs[l][4];
      for (int c = 0; c < nchem; c++) fp[l][5 + c] = fs[l][5 + c];
    }
    // WENO5-JS per component (399-423, 443-467)
    for (int c = 0; c < nv; c++) {
      const R f0 = fp[0][c], f1 = fp[1][c], f2 = fp[2][c], f3 = fp[3][c], f4 = fp[4][c];
      const R b1 = bc * sq(f2 - R(2.0) * f3 + f4) + fourth * sq(R(3.0) * f2 - R(4.0) * f3 + f4);
      const R b2 = bc * sq(f1 - R(2.0) * f2 + f3) + fourth * sq(f1 - f3);
      const R b3 = bc * sq(f0 - R(2.0) * f1 + f2) + fourth * sq(f0 - R(4.0) * f1 + R(3.0) * f2);
      R w1, w2, w3, q1, q2, q3;
      if (side == 0) {
        w1 = R(0.3) / ((eps + b1) * (eps + b1));
        w2 = R(0.6) / ((eps + b2) * (eps + b2));
        w3 = R(0.1) / ((eps + b3) * (eps + b3));
        q1 = c13 * f2 + c56 * f3 - c16 * f4;
        q2 = (-c16) * f1 + c56 * f2 + c13 * f3;
        q3 = c13 * f0 - c76 * f1 + c116 * f2;
        ff[c] = (q1 * w1 + q2 * w2 + q3 * w3) / (w1 + w2 + w3);
      } else {
        w1 = R(0.1) / ((eps + b1) * (eps + b1));
        w2 = R(0.6) / ((eps + b2) * (eps + b2));
        w3 = R(0.3) / ((eps + b3) * (eps + b3));
        q1 = c116 * f2 - c76 * f3 + c13 * f4;
        q2 = c13 * f1 + c56 * f2 - c16 * f3;
        q3 = (-c16) * f0 + c56 * f1 + c13 * f2;
        ff[c] += (q1 * w1 + q2 * w2 + q3 * w3) / (w1 + w2 + w3);
      }
    }
  }

  // back to conserved variables (470-473) and to the global momentum ordering (476-477)
  for (int a = 0; a < 5; a++)
    fface[a] = RV[a][0] * ff[0] + RV[a][1] * ff[1] + RV[a][2] * ff[2]
             + RV[a][3] * ff[3] + RV[a][4] * ff[4];
  for (int c = 0; c < nchem; c++) fface[5 + c] = ff[5 + c];
  if (dir > 0) std::swap(fface[1], fface[1 + dir]);
}

}  // namespace orc
