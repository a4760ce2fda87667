// WENO5 flux divergence fe() kernels for gfx950 (FP64).
//
// Plain path (the comparator; fEuler, utilities.cpp:17-253, transcribed for the GPU):
//   fe_plain_faces : one thread per cell computes the lower x/y/z faces of its cell
//                    (plus the domain's upper boundary faces) with the reference-order
//                    face flux, into SoA face-flux buffers;
//   fe_plain_div   : one thread per cell forms the divergence in the reference order.
// Fast path without tracers (two launches):
//   fe_xyface_div  : x- and y-faces of 16 x 15-cell plane tiles from one LDS staging of the
//                    tile and its x/y halos, and their part of the divergence into wdot
//                    (the x and y face fluxes stay in LDS); the legal-state check;
//   fe_zface_div_xy: z-faces (tiles of 16 x-cells by 32 faces, two per thread, their
//                    primitives staged in LDS), and the divergence of the 31 cells between
//                    a tile's 32 faces (tiles overlap by one face), adding the x/y part
//                    read from wdot at the kernel's start.
// Fast path with tracers (three launches):
//   fe_xyface_div<NC>: as without tracers, then the tracers on the same x- and y-faces
//                    (their values staged in LDS as [tracer][point] after the fluid phase)
//                    and their part of the divergence into wdot;
//   fe_zface_div_xy: the fluid z-faces, as without tracers;
//   fe_zface_tracers: the tracer z-faces (tiles of 16 x-cells by 16 faces; per point the
//                    normal velocity, wave speed and tracer values in LDS), adding their
//                    part of the divergence.
// The divergence is accumulated as -((dFx/dx + dFy/dy) + dFz/dz), the reference's
// association (utilities.cpp:202-207).
#include "mvd_device.hpp"
#include "mvd_internal.hpp"

namespace mvd {

// ------------------------------------------------------------------------------------
// shared helpers
// ------------------------------------------------------------------------------------

__device__ inline long cell_index(const Grid &g, long i, long j, long k) {
  return i + (long)g.nx * (j + (long)g.ny * k);
}

// Cell index of extended index e along direction d through (i,j,k), and its ghost signs.
__device__ inline long stencil_cell(const Grid &g, int d, int i, int j, int k, int e,
                                    GhostMap &m) {
  m = ghost_map(g, d, e);
  if (d == 0) i = m.src; else if (d == 1) j = m.src; else k = m.src;
  return cell_index(g, i, j, k);
}

// Conserved state of one stencil point in direction-local order
// (rho, m_normal, m_t1, m_t2, et) with ghost signs applied; the momentum permutation is
// the reference's swap of component 1 with component 1+d (utilities.cpp:283-285).
__device__ inline void load_local(const Grid &g, const Fields &f, int d, long c,
                                  const GhostMap &m, double (&s)[5]) {
  const double r = f.w[0][c], mx = f.w[1][c], my = f.w[2][c], mz = f.w[3][c], e = f.w[4][c];
  double mn, m1, m2;
  if (d == 0) { mn = mx; m1 = my; m2 = mz; }
  else if (d == 1) { mn = my; m1 = mx; m2 = mz; }
  else { mn = mz; m1 = my; m2 = mx; }
  s[0] = r * m.sgn_all;
  s[1] = mn * m.sgn_norm;
  s[2] = m1 * m.sgn_all;
  s[3] = m2 * m.sgn_all;
  s[4] = e * m.sgn_all;
}

// direction-local flux components back to global momentum order
__device__ inline int global_comp(int d, int a) {
  if (a == 1) return 1 + d;
  if (a == 1 + d) return 1;
  return a;
}

__device__ inline void legal_check(const Grid &g, const Fields &f, long c, unsigned *flag) {
  const double r = f.w[0][c], mx = f.w[1][c], my = f.w[2][c], mz = f.w[3][c], e = f.w[4][c];
  const double p = (g.gamma - 1.0) * (e - (mx * mx + my * my + mz * mz) * 0.5 / r);
  unsigned bits = 0;
  if (!(r > 0.0)) bits |= 1u;
  if (!(e > 0.0)) bits |= 2u;
  if (!(p > 0.0)) bits |= 4u;
  if (bits) atomicOr(flag, bits);
}

// ------------------------------------------------------------------------------------
// plain comparator
// ------------------------------------------------------------------------------------

template <int NC>
__global__ void __launch_bounds__(256)
fe_plain_faces(Grid g, Fields f, double *Fx, double *Fy, double *Fz, unsigned *flag) {
  constexpr int NV = 5 + NC;
  const long nxp = g.nx + 1, nyp = g.ny + 1, nzp = g.nz + 1;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nxp * nyp * nzp) return;
  const int i = (int)(tid % nxp);
  const int j = (int)((tid / nxp) % nyp);
  const int k = (int)(tid / (nxp * nyp));
  if (i < g.nx && j < g.ny && k < g.nz) legal_check(g, f, cell_index(g, i, j, k), flag);
  const long nfx = nxp * g.ny * g.nz, nfy = (long)g.nx * nyp * g.nz, nfz = (long)g.nx * g.ny * nzp;
  for (int d = 0; d < 3; d++) {
    const bool ok = (d == 0) ? (j < g.ny && k < g.nz) : (d == 1) ? (i < g.nx && k < g.nz)
                                                               : (i < g.nx && j < g.ny);
    if (!ok) continue;
    const int c0 = (d == 0) ? i : (d == 1) ? j : k;
    double s[6][NV];
    for (int l = 0; l < 6; l++) {
      GhostMap m;
      const long c = stencil_cell(g, d, i, j, k, c0 - 3 + l, m);
      double loc[5];
      load_local(g, f, d, c, m, loc);
      for (int v = 0; v < 5; v++) s[l][v] = loc[v];
      for (int v = 0; v < NC; v++) s[l][5 + v] = f.chem[(long)NC * c + v] * m.sgn_all;
    }
    double ff[NV];
    face_flux_ref<NC>(s, g.gamma, ff);
    double *F = (d == 0) ? Fx : (d == 1) ? Fy : Fz;
    const long nf = (d == 0) ? nfx : (d == 1) ? nfy : nfz;
    const long face = (d == 0) ? (i + nxp * (j + (long)g.ny * k))
                    : (d == 1) ? (i + (long)g.nx * (j + nyp * k))
                               : (i + (long)g.nx * (j + (long)g.ny * k));
    for (int a = 0; a < NV; a++) F[(long)global_comp(d, a) * nf + face] = ff[a];
  }
}

template <int NC>
__global__ void __launch_bounds__(256)
fe_plain_div(Grid g, const double *Fx, const double *Fy, const double *Fz, FieldsOut out) {
#pragma clang fp contract(off)
  constexpr int NV = 5 + NC;
  const long ncell = (long)g.nx * g.ny * g.nz;
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const long i = c % g.nx, j = (c / g.nx) % g.ny, k = c / ((long)g.nx * g.ny);
  const long nxp = g.nx + 1, nyp = g.ny + 1, nzp = g.nz + 1;
  const long nfx = nxp * g.ny * g.nz, nfy = (long)g.nx * nyp * g.nz, nfz = (long)g.nx * g.ny * nzp;
  const long x0 = i + nxp * (j + g.ny * k);
  const long y0 = i + g.nx * (j + nyp * k);
  const long z0 = c;
  for (int v = 0; v < NV; v++) {
    const double dv = (Fx[v * nfx + x0 + 1] - Fx[v * nfx + x0]) / g.dx +
                      (Fy[v * nfy + y0 + g.nx] - Fy[v * nfy + y0]) / g.dy +
                      (Fz[v * nfz + z0 + (long)g.nx * g.ny] - Fz[v * nfz + z0]) / g.dz;
    if (v < 5) out.w[v][c] = 0.0 - dv;
    else out.chem[(long)NC * c + (v - 5)] = 0.0 - dv;
  }
}

// ------------------------------------------------------------------------------------
// fast path
// ------------------------------------------------------------------------------------

// Raw conserved state of a stencil point, converted to direction-local order with the
// ghost signs applied.
struct RawPt {
  double r, mn, m1, m2, e;
};

template <int D>
__device__ inline RawPt load_raw(const Fields &f, int c, double sa, double sn) {
  const double r = f.w[0][c], mx = f.w[1][c], my = f.w[2][c], mz = f.w[3][c], e = f.w[4][c];
  RawPt q;
  q.r = r * sa;
  q.e = e * sa;
  if (D == 0) { q.mn = mx * sn; q.m1 = my * sa; q.m2 = mz * sa; }
  else if (D == 1) { q.mn = my * sn; q.m1 = mx * sa; q.m2 = mz * sa; }
  else { q.mn = mz * sn; q.m1 = my * sa; q.m2 = mx * sa; }
  return q;
}

// Cell of extended index e along direction D of the line whose e = 0 cell is `base`
// (cells `stride` apart), with ghost signs. Interior points take the short path.
template <int D>
__device__ inline int line_cell(const Grid &g, int base, int stride, int n, int e, double &sa,
                                double &sn) {
  if (e >= 0 && e < n) {
    sa = 1.0;
    sn = 1.0;
    return base + e * stride;
  }
  const GhostMap m = ghost_map(g, D, e);
  sa = m.sgn_all;
  sn = m.sgn_norm;
  return base + m.src * stride;
}

__device__ inline Pt to_point(const RawPt &q, double gamma) {
  return make_point(q.r, q.mn, q.m1, q.m2, q.e, gamma);
}

// local (rho, m_normal, m_t1, m_t2, e) flux of direction D into global component order
template <int D>
__device__ inline void to_global(const double (&Fl)[5], double *Fg) {
  Fg[0] = Fl[0];
  Fg[4] = Fl[4];
  if (D == 1) { Fg[2] = Fl[1]; Fg[1] = Fl[2]; Fg[3] = Fl[3]; }
  else { Fg[3] = Fl[1]; Fg[2] = Fl[2]; Fg[1] = Fl[3]; }
}

// Tracer z-faces, adding their part to the x/y part that fe_xyface_div<NC> left in
// out.chem (the fluid z-faces are fe_zface_div_xy's). A block of 16 x 16 threads covers 16
// x-cells of row j and 16 consecutive z-faces, i.e. the 15 cells between them. Per point
// of the tile (21 per column) LDS holds the normal velocity w, the wave speed |w| + c
// (make_point's, so the face's alpha is the fluid flux's bit for bit) and the NC tracer
// values; each thread forms its face's alpha and split factors once and then the NC
// tracer fluxes. Without the fluid flux's registers the kernel runs at a higher occupancy
// than a fused fluid + tracer z kernel.
template <int NC>
__global__ void __launch_bounds__(256)
fe_zface_tracers(Grid g, Fields f, FieldsOut out, double idz) {
  constexpr int TX = 16, TF = 16, TPTS = TF + 5, TC = TF - 1, NPT = TPTS * TX;
  constexpr int NR = (NPT + 255) / 256;
  constexpr int PT_SZ = (2 + NC) * NPT, FL_SZ = NC * TF * TX;
  __shared__ double s_buf[PT_SZ > FL_SZ ? PT_SZ : FL_SZ];   // [w | speed | tracers][point]
  const int t = threadIdx.x, tx = t % TX, ty = t / TX;
  const int nxc = (g.nx + TX - 1) / TX, ncc = (g.nz + TC - 1) / TC;
  int b = blockIdx.x;
  const int xc = b % nxc;
  b /= nxc;
  const int cc = b % ncc;
  const int j = b / ncc;
  const int k0 = cc * TC;
  const int i = xc * TX + tx, k = k0 + ty;
  // the x/y part of the thread's cell, loaded first (its latency passes under the staging)
  const bool cell_ok = ty < TC && i < g.nx && k < g.nz;
  const long cell = i + (long)g.nx * (j + (long)g.ny * k);
  double dxy[NC];
#pragma unroll
  for (int v = 0; v < NC; v++) dxy[v] = cell_ok ? out.chem[(long)NC * cell + v] : 0.0;
  // points k0 - 3 .. k0 + TF + 1 of the 16 columns: located, loaded, then converted
  {
    const int stride = g.nx * g.ny;
    int cs[NR];
    double sa[NR], sn[NR];
    bool use[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const int s = t + 256 * r;
      const int row = s / TX, x = s - row * TX;
      const int e = k0 - 3 + row;
      use[r] = s < NPT && xc * TX + x < g.nx && e <= g.nz + 2;
      cs[r] = use[r] ? line_cell<2>(g, xc * TX + x + g.nx * j, stride, g.nz, e, sa[r], sn[r])
                     : 0;
    }
    RawPt raw[NR];
    double tv[NR][NC];
#pragma unroll
    for (int r = 0; r < NR; r++) {
      if (!use[r]) continue;
      raw[r] = load_raw<2>(f, cs[r], sa[r], sn[r]);
#pragma unroll
      for (int v = 0; v < NC; v++) tv[r][v] = f.chem[(long)NC * cs[r] + v] * sa[r];
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
      if (!use[r]) continue;
      const int s = t + 256 * r;
      const Pt P = to_point(raw[r], g.gamma);
      s_buf[s] = P.mn * P.ir;
      s_buf[NPT + s] = P.a;
#pragma unroll
      for (int v = 0; v < NC; v++) s_buf[(2 + v) * NPT + s] = tv[r][v];
    }
  }
  __syncthreads();
  const bool face_ok = i < g.nx && k <= g.nz;
  double Fz[NC];
  if (face_ok) {
    const double *pb = s_buf + ty * TX + tx;
    double alpha = 0.0, un[6];
#pragma unroll
    for (int l = 0; l < 6; l++) alpha = fmax(alpha, pb[NPT + l * TX]);
#pragma unroll
    for (int l = 0; l < 6; l++) un[l] = pb[l * TX];
    double hp[5], hm[5];
    tracer_split(un, alpha, hp, hm);
#pragma unroll
    for (int v = 0; v < NC; v++) {
      double cv[6];
#pragma unroll
      for (int l = 0; l < 6; l++) cv[l] = pb[(2 + v) * NPT + l * TX];
      Fz[v] = tracer_face_flux(cv, hp, hm);
    }
  }
  __syncthreads();   // the points are dead; LDS takes the face fluxes [tracer][face][x]
  if (face_ok) {
#pragma unroll
    for (int v = 0; v < NC; v++) s_buf[(v * TF + ty) * TX + tx] = Fz[v];
  }
  __syncthreads();
  if (!cell_ok) return;
#pragma unroll
  for (int v = 0; v < NC; v++) {
    const double dz = (s_buf[(v * TF + ty + 1) * TX + tx] - s_buf[(v * TF + ty) * TX + tx]) * idz;
    out.chem[(long)NC * cell + v] = 0.0 - (dxy[v] + dz);
  }
}

// z-faces without tracers (nchem = 0), adding their part to the x/y part that
// fe_xyface_div left in out. A block of 16 x 16 threads covers 16 x-cells of row j and
// ZF = 16 FPT consecutive z-faces (FPT per thread, 16 apart), i.e. the ZF - 1 cells between
// them; the ZF + 5 points of each column are converted once into LDS. The x/y part of the
// thread's cells is loaded at the start, so that its latency is spent under the staging and
// the face fluxes instead of after them.
template <int FPT>
__global__ void __launch_bounds__(256)
fe_zface_div_xy(Grid g, Fields f, FieldsOut out, double idz) {
  constexpr int TX = 16, ZF = 16 * FPT, TPTS = ZF + 5, TC = ZF - 1;
  constexpr int FS = TPTS * TX;
  __shared__ double s_pt[9 * FS];
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int nxc = (g.nx + TX - 1) / TX, ncc = (g.nz + TC - 1) / TC;
  int b = blockIdx.x;
  const int xc = b % nxc;
  b /= nxc;
  const int cc = b % ncc;
  const int j = b / ncc;
  const int k0 = cc * TC;
  const int i = xc * TX + tx;
  const long plane = (long)g.nx * g.ny;
  // x/y part of the thread's cells (k0 + ty + 16 q)
  double dxy[FPT][5];
#pragma unroll
  for (int q = 0; q < FPT; q++) {
    const int c = ty + 16 * q, k = k0 + c;
    const long cell = i + (long)g.nx * j + plane * k;
    const bool own = c < TC && i < g.nx && k < g.nz;
#pragma unroll
    for (int a = 0; a < 5; a++) dxy[q][a] = own ? out.w[a][cell] : 0.0;
  }
  // points k0 - 3 .. k0 + ZF + 1 of the 16 columns; a thread's points (s = t + 256 n) are
  // located first and all their loads issued before any is converted
  {
    constexpr int NR = (TPTS * TX + 255) / 256;
    const int stride = g.nx * g.ny;
    int cs[NR];
    double sa[NR], sn[NR];
    bool use[NR];
#pragma unroll
    for (int n = 0; n < NR; n++) {
      const int s = threadIdx.x + 256 * n;
      const int r = s / TX, x = s - r * TX;
      const int e = k0 - 3 + r;
      use[n] = s < TPTS * TX && xc * TX + x < g.nx && e <= g.nz + 2;
      cs[n] = use[n] ? line_cell<2>(g, xc * TX + x + g.nx * j, stride, g.nz, e, sa[n], sn[n]) : 0;
    }
    RawPt raw[NR];
#pragma unroll
    for (int n = 0; n < NR; n++)
      if (use[n]) raw[n] = load_raw<2>(f, cs[n], sa[n], sn[n]);
#pragma unroll
    for (int n = 0; n < NR; n++) {
      if (!use[n]) continue;
      const Pt P = to_point(raw[n], g.gamma);
      double *pt = s_pt + threadIdx.x + 256 * n;
      pt[0 * FS] = P.r;  pt[1 * FS] = P.mn; pt[2 * FS] = P.m1;
      pt[3 * FS] = P.m2; pt[4 * FS] = P.e;  pt[5 * FS] = P.p;
      pt[6 * FS] = P.ir; pt[7 * FS] = P.a;  pt[8 * FS] = P.sr;
    }
  }
  __syncthreads();
  double Fz[FPT][5];
#pragma unroll
  for (int q = 0; q < FPT; q++) {
    const int fz = ty + 16 * q;
    if (i < g.nx && k0 + fz <= g.nz) {
      double Fl[5], alpha;
      face_flux_lds(s_pt + fz * TX + tx, FS, TX, g.gamma, Fl, alpha);
      to_global<2>(Fl, Fz[q]);
    }
  }
  __syncthreads();   // the point tile is dead; its LDS takes the face fluxes [comp][face][x]
#pragma unroll
  for (int q = 0; q < FPT; q++) {
    const int fz = ty + 16 * q;
    if (i < g.nx && k0 + fz <= g.nz) {
#pragma unroll
      for (int a = 0; a < 5; a++) s_pt[(a * ZF + fz) * TX + tx] = Fz[q][a];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < FPT; q++) {
    const int c = ty + 16 * q, k = k0 + c;
    if (c >= TC || i >= g.nx || k >= g.nz) continue;
    const long cell = i + (long)g.nx * j + plane * k;
#pragma unroll
    for (int a = 0; a < 5; a++) {
      const double dz = (s_pt[(a * ZF + c + 1) * TX + tx] - s_pt[(a * ZF + c) * TX + tx]) * idz;
      out.w[a][cell] = 0.0 - (dxy[q][a] + dz);
    }
  }
}

// x- and y-faces of one z-plane tile together: a block covers XY_TX x XY_TY cells of plane k
// and forms (Fx[i+1]-Fx[i])/dx + (Fy[j+1]-Fy[j])/dy for them into out (the z kernel then
// adds the z part), so the x and y face fluxes never go to HBM. The points of the tile and
// of its x and y halos (a cross, 3 deep) are converted once into LDS, keeping mx and my in
// place and both wave speeds |u|+c and |v|+c, so that every point serves both directions;
// x-faces (XY_TX+1 per row) and y-faces (XY_TY+1 per column) are 255 and 256 with the
// 16 x 15 tile: one of each per thread.
//
// With NC tracers a second phase follows on the same faces: the tracer values of the cross
// (each point's NC values one contiguous AoS read, issued before the fluid divergence) go
// to LDS as [tracer][point], where the fluid points were; the normal velocities of the stencil points (u on the tile's rows, v
// on its columns) are kept aside from the fluid staging, and each face's fluid alpha stays
// in a register; each thread then forms its x- and y-face's tracer fluxes from LDS, and
// the divergence goes through LDS in passes of at most XY_TP tracers.
constexpr int XY_TX = 16, XY_TY = 15;
constexpr int XY_PX = XY_TX + 6, XY_PY = XY_TY + 6;      // point window 22 x 21
constexpr int XY_NP = XY_PX * XY_PY;
constexpr int XY_NF = 10;     // r, mx, my, mz, e, p, 1/r, |u|+c, sqrt(r), |v|+c
constexpr int XY_NXF = (XY_TX + 1) * XY_TY, XY_NYF = XY_TX * (XY_TY + 1);   // 255, 256
constexpr int XY_LDS = XY_NF * XY_NP + 5 * XY_NYF;          // points + y-face fluxes
constexpr int XY_TP = XY_LDS / (XY_NXF + XY_NYF);           // tracers per divergence pass: 11

// Point q of the cross of the tile at (x0, y0) in plane k: its cell, ghost signs (all
// fields, mx, my) and whether it is used (inside the cross and the extended grid) and a
// cell of the tile (`own`), with the boundary mapping of ghost_map.
struct XYPoint {
  int cell;
  double s_all, s_mx, s_my;
  bool use, own, row, col;
};

__device__ inline XYPoint xy_point(const Grid &g, int x0, int y0, int k, int q) {
  XYPoint P;
  const int yy = q / XY_PX, xx = q - yy * XY_PX;
  const int ex = x0 - 3 + xx, ey = y0 - 3 + yy;
  P.row = yy >= 3 && yy < 3 + XY_TY && ey < g.ny;          // x-line point
  P.col = xx >= 3 && xx < 3 + XY_TX && ex < g.nx;          // y-column point
  int i = ex, j = ey;
  P.s_all = P.s_mx = P.s_my = 1.0;
  P.use = q < XY_NP && (P.row ? ex <= g.nx + 2 : (P.col && ey <= g.ny + 2));
  P.own = P.row && P.col && ex >= 0;   // a cell of the tile: legal_state once
  if (P.row) {
    if (ex < 0 || ex >= g.nx) {
      const GhostMap m = ghost_map(g, 0, ex);
      i = m.src;
      P.s_all = P.s_my = m.sgn_all;
      P.s_mx = m.sgn_norm;
    }
  } else if (P.col) {
    if (ey < 0 || ey >= g.ny) {
      const GhostMap m = ghost_map(g, 1, ey);
      j = m.src;
      P.s_all = P.s_mx = m.sgn_all;
      P.s_my = m.sgn_norm;
    }
  }
  P.cell = P.use ? (int)(i + (long)g.nx * (j + (long)g.ny * k)) : 0;
  return P;
}

template <int NC>
__global__ void __launch_bounds__(256, 3)
fe_xyface_div(Grid g, Fields f, FieldsOut out, double idx, double idy, unsigned *flag) {
  // points [XY_NF][XY_NP], then the y-face fluxes [5][XY_TY+1][XY_TX]; one array, so that the
  // tracer phase can use all of it
  __shared__ double lds[XY_LDS];
  double *const sh = lds, *const sfy = lds + XY_NF * XY_NP;
  // tracer stencils' normal velocities: u of the row points [XY_TY][XY_PX], then v of the
  // column points [XY_PY][XY_TX]
  constexpr int UN_X = XY_TY * XY_PX, UN_N = UN_X + XY_PY * XY_TX;
  __shared__ double sun[NC > 0 ? UN_N : 1];
  const int t = threadIdx.x;
  const int nxb = (g.nx + XY_TX - 1) / XY_TX, nyb = (g.ny + XY_TY - 1) / XY_TY;
  // The hardware deals blocks to the 8 XCDs in turn, and each XCD has its own L2. Tiles are
  // renumbered so that consecutive tiles (x neighbours, which read the same halo lines)
  // run on the same XCD one after another: block b takes tile (b % 8) * m + b / 8 (m = the
  // blocks per XCD; the remainder past 8 m keeps its own number).
  const unsigned m8 = gridDim.x / 8;
  int b = blockIdx.x < 8 * m8 ? (int)((blockIdx.x % 8) * m8 + blockIdx.x / 8) : (int)blockIdx.x;
  const int xb = b % nxb;
  b /= nxb;
  const int yb = b % nyb;
  const int k = b / nyb;
  const int x0 = xb * XY_TX, y0 = yb * XY_TY;
  // stage the cross: points of the tile's rows (real y) over the x window, with x ghosts,
  // then the other points of the tile's columns (real x) over the y window, with y ghosts.
  // A thread's points (q = t + 256 r) are located first and all their loads issued before
  // any is converted, so that their latencies overlap.
  constexpr int NR = (XY_NP + 255) / 256;
  XYPoint pt[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) pt[r] = xy_point(g, x0, y0, k, t + 256 * r);
  double raw[NR][5];
#pragma unroll
  for (int r = 0; r < NR; r++) {
#pragma unroll
    for (int a = 0; a < 5; a++) raw[r][a] = pt[r].use ? f.w[a][pt[r].cell] : 1.0;
  }
#pragma unroll
  for (int r = 0; r < NR; r++) {
    if (!pt[r].use) continue;
    const double rr = raw[r][0] * pt[r].s_all, mx = raw[r][1] * pt[r].s_mx;
    const double my = raw[r][2] * pt[r].s_my;
    const double mz = raw[r][3] * pt[r].s_all, e = raw[r][4] * pt[r].s_all;
    const double ir = fast_rcp(rr);
    const double p = (g.gamma - 1.0) * (e - 0.5 * (mx * mx + my * my + mz * mz) * ir);
    const double cs = sun_sqrt(g.gamma * p * ir);
    const int q = t + 256 * r;
    double *pp = sh + q;
    pp[0 * XY_NP] = rr; pp[1 * XY_NP] = mx; pp[2 * XY_NP] = my;
    pp[3 * XY_NP] = mz; pp[4 * XY_NP] = e;  pp[5 * XY_NP] = p;
    pp[6 * XY_NP] = ir; pp[7 * XY_NP] = fabs(mx * ir) + cs;
    pp[8 * XY_NP] = sun_sqrt(rr);
    pp[9 * XY_NP] = fabs(my * ir) + cs;
    if (NC > 0) {
      const int yy = q / XY_PX, xx = q - yy * XY_PX;
      if (pt[r].row) sun[(yy - 3) * XY_PX + xx] = mx * ir;
      if (pt[r].col) sun[UN_X + yy * XY_TX + (xx - 3)] = my * ir;
    }
    if (pt[r].own) {
      unsigned bits = 0;
      if (!(rr > 0.0)) bits |= 1u;
      if (!(e > 0.0)) bits |= 2u;
      if (!(p > 0.0)) bits |= 4u;
      if (bits) atomicOr(flag, bits);
    }
  }
  __syncthreads();
  // y-face t: column cx, faces between rows y0+fy-1 and y0+fy (points yy = fy..fy+5)
  const int cx = t % XY_TX, fy = t / XY_TX;
  const bool yok = x0 + cx < g.nx && y0 + fy <= g.ny && fy <= XY_TY;
  // x-face t: row ry, faces between columns x0+fx-1 and x0+fx (points xx = fx..fx+5)
  const int ry = t / (XY_TX + 1), fx = t % (XY_TX + 1);
  const bool xok = t < XY_NXF && y0 + ry < g.ny && x0 + fx <= g.nx;
  if (yok) {
    double Fl[5], Fy[5], alpha_y;
    face_flux_lds_f<2, 1, 9>(sh + fy * XY_PX + cx + 3, XY_NP, XY_PX, g.gamma, Fl, alpha_y);
    to_global<1>(Fl, Fy);
#pragma unroll
    for (int a = 0; a < 5; a++) sfy[(a * (XY_TY + 1) + fy) * XY_TX + cx] = Fy[a];
  }
  double Fx[5];
  if (xok) {
    double alpha_x;
    face_flux_lds_f<1, 2, 7>(sh + (ry + 3) * XY_PX + fx, XY_NP, 1, g.gamma, Fx, alpha_x);
  }
  __syncthreads();                    // points dead: their LDS takes the x-face fluxes
  double *sfx = sh;                   // [5][XY_TY][XY_TX+1] (over fields 0-2)
  if (xok)
#pragma unroll
    for (int a = 0; a < 5; a++) sfx[(a * XY_TY + ry) * (XY_TX + 1) + fx] = Fx[a];
  __syncthreads();
  // the faces' alpha for the tracers, formed again (as the fluid flux forms it) from the
  // wave speeds, which the x-face fluxes did not overwrite, rather than held through the
  // fluid fluxes
  double alpha_y = 0.0, alpha_x = 0.0;
  if (NC > 0) {
#pragma unroll
    for (int l = 0; l < 6; l++) {
      alpha_y = fmax(alpha_y, sh[9 * XY_NP + (fy + l) * XY_PX + cx + 3]);
      alpha_x = fmax(alpha_x, sh[7 * XY_NP + (ry + 3) * XY_PX + fx + l]);
    }
  }
  // the tracer values of the thread's cross points, loaded now so that their latency
  // passes under the fluid divergence
  double tv[NC > 0 ? NR : 1][NC > 0 ? NC : 1];
  if (NC > 0) {
    // the points are located again rather than kept from the staging: made opaque, the
    // tile origin keeps the compiler from holding the staging's values across the fluxes
    int ox = x0, oy = y0, ok = k;
    __asm__ volatile("" : "+s"(ox), "+s"(oy), "+s"(ok));
#pragma unroll
    for (int r = 0; r < NR; r++) {
      const XYPoint P = xy_point(g, ox, oy, ok, t + 256 * r);
#pragma unroll
      for (int v = 0; v < NC; v++)
        tv[r][v] = P.use ? f.chem[(long)NC * P.cell + v] * P.s_all : 0.0;
    }
  }
  const int ci = t % XY_TX, cj = t / XY_TX;
  const bool cok = cj < XY_TY && x0 + ci < g.nx && y0 + cj < g.ny;
  const long cell = (x0 + ci) + (long)g.nx * ((y0 + cj) + (long)g.ny * k);
  if (cok) {
#pragma unroll
    for (int a = 0; a < 5; a++) {
      const double dx = (sfx[(a * XY_TY + cj) * (XY_TX + 1) + ci + 1] -
                         sfx[(a * XY_TY + cj) * (XY_TX + 1) + ci]) * idx;
      const double dy = (sfy[(a * (XY_TY + 1) + cj + 1) * XY_TX + ci] -
                         sfy[(a * (XY_TY + 1) + cj) * XY_TX + ci]) * idy;
      out.w[a][cell] = dx + dy;
    }
  }
  if (NC == 0) return;
  __syncthreads();                    // fluid fluxes read: LDS takes the tracer values
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const int q = t + 256 * r;
    if (q < XY_NP)
#pragma unroll
      for (int v = 0; v < NC; v++) sh[v * XY_NP + q] = tv[r][v];
  }
  __syncthreads();
  double Fxt[NC > 0 ? NC : 1], Fyt[NC > 0 ? NC : 1];
  if (yok) {
    double un[6], hp[5], hm[5];
#pragma unroll
    for (int l = 0; l < 6; l++) un[l] = sun[UN_X + (fy + l) * XY_TX + cx];
    tracer_split(un, alpha_y, hp, hm);
#pragma unroll
    for (int v = 0; v < NC; v++) {
      const double *c0 = sh + v * XY_NP + fy * XY_PX + cx + 3;
      double cv[6];
#pragma unroll
      for (int l = 0; l < 6; l++) cv[l] = c0[l * XY_PX];
      Fyt[v] = tracer_face_flux(cv, hp, hm);
    }
  }
  if (xok) {
    double un[6], hp[5], hm[5];
#pragma unroll
    for (int l = 0; l < 6; l++) un[l] = sun[ry * XY_PX + fx + l];
    tracer_split(un, alpha_x, hp, hm);
#pragma unroll
    for (int v = 0; v < NC; v++) {
      const double *c0 = sh + v * XY_NP + (ry + 3) * XY_PX + fx;
      double cv[6];
#pragma unroll
      for (int l = 0; l < 6; l++) cv[l] = c0[l];
      Fxt[v] = tracer_face_flux(cv, hp, hm);
    }
  }
  // divergence in passes of XY_TP tracers (one pass for up to 11): pass fluxes
  // [v][x-faces | y-faces] in LDS, so that each cell's tracers are written together
#pragma unroll
  for (int v0 = 0; v0 < NC; v0 += XY_TP) {
    __syncthreads();                  // tracer values (first pass) / last pass's fluxes read
#pragma unroll
    for (int v = v0; v < NC && v < v0 + XY_TP; v++) {
      double *fl = sh + (v - v0) * (XY_NXF + XY_NYF);
      if (xok) fl[ry * (XY_TX + 1) + fx] = Fxt[v];
      if (yok) fl[XY_NXF + fy * XY_TX + cx] = Fyt[v];
    }
    __syncthreads();
    if (cok) {
#pragma unroll
      for (int v = v0; v < NC && v < v0 + XY_TP; v++) {
        const double *fl = sh + (v - v0) * (XY_NXF + XY_NYF);
        const double dx = (fl[cj * (XY_TX + 1) + ci + 1] - fl[cj * (XY_TX + 1) + ci]) * idx;
        const double dy = (fl[XY_NXF + (cj + 1) * XY_TX + ci] -
                           fl[XY_NXF + cj * XY_TX + ci]) * idy;
        out.chem[(long)NC * cell + v] = dx + dy;
      }
    }
  }
}

__global__ void reset_flag(unsigned *flag) { *flag = 0u; }

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------

static inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

constexpr int TILE_X = 16, TILE_F = 16;   // fe_zface_tracers' tile
constexpr int ZFPT = 2;     // z-faces per thread in fe_zface_div_xy

template <int NC>
static int launch_fe(const Grid &g, const Fields &f, const FieldsOut &o, FeScratch &s,
                     hipStream_t st, bool plain, unsigned *accum) {
  const long nxp = g.nx + 1, nyp = g.ny + 1, nzp = g.nz + 1;
  unsigned *const flag = accum ? accum : s.flag;
  if (!accum) hipLaunchKernelGGL(reset_flag, dim3(1), dim3(1), 0, st, s.flag);
  KernelTimer &tm = s.timer;
  if (plain) {
    tm.begin(PROF_PLAIN_FACES, st);
    hipLaunchKernelGGL(fe_plain_faces<NC>, dim3(nblocks(nxp * nyp * nzp, 256)), dim3(256), 0, st,
                       g, f, s.Fx, s.Fy, s.Fz, flag);
    tm.end(PROF_PLAIN_FACES, st);
    tm.begin(PROF_PLAIN_DIV, st);
    hipLaunchKernelGGL(fe_plain_div<NC>, dim3(nblocks((long)g.nx * g.ny * g.nz, 256)), dim3(256),
                       0, st, g, s.Fx, s.Fy, s.Fz, o);
    tm.end(PROF_PLAIN_DIV, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  // x and y faces of each plane tile (fluid, then the tracers) in one kernel, then the z
  // faces adding their part
  constexpr int TX = TILE_X, TF = TILE_F;
  const long nxc = (g.nx + TX - 1) / TX;
  const long nxyb = ((long)(g.nx + XY_TX - 1) / XY_TX) * ((g.ny + XY_TY - 1) / XY_TY) * g.nz;
  tm.begin(PROF_XYFACE_DIV, st);
  hipLaunchKernelGGL(fe_xyface_div<NC>, dim3((unsigned)nxyb), dim3(256), 0, st, g, f, o,
                     1.0 / g.dx, 1.0 / g.dy, flag);
  tm.end(PROF_XYFACE_DIV, st);
  tm.begin(PROF_ZFACE_DIV, st);
  constexpr int ZC = 16 * ZFPT - 1;                        // cells per z tile
  hipLaunchKernelGGL(fe_zface_div_xy<ZFPT>,
                     dim3((unsigned)(nxc * ((g.nz + ZC - 1) / ZC) * g.ny)), dim3(256), 0, st,
                     g, f, o, 1.0 / g.dz);
  if constexpr (NC > 0)
    hipLaunchKernelGGL(fe_zface_tracers<NC>,
                       dim3((unsigned)(nxc * ((g.nz + TF - 2) / (TF - 1)) * g.ny)),
                       dim3(TX * TF), 0, st, g, f, o, 1.0 / g.dz);
  tm.end(PROF_ZFACE_DIV, st);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fe_launch(const Grid &g, const Fields &f, const FieldsOut &o, FeScratch &s,
              hipStream_t st, bool plain, unsigned *accum) {
  // 32-bit cell indices in the fast kernels
  if ((long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1) >= (1L << 31) / 16) return -2;
  switch (g.nchem) {
    case 0: return launch_fe<0>(g, f, o, s, st, plain, accum);
    case 1: return launch_fe<1>(g, f, o, s, st, plain, accum);
    case 2: return launch_fe<2>(g, f, o, s, st, plain, accum);
    case 4: return launch_fe<4>(g, f, o, s, st, plain, accum);
    case 6: return launch_fe<6>(g, f, o, s, st, plain, accum);
    case 10: return launch_fe<10>(g, f, o, s, st, plain, accum);
    default: return -2;
  }
}

bool fe_nchem_supported(int nchem) {
  return nchem == 0 || nchem == 1 || nchem == 2 || nchem == 4 || nchem == 6 || nchem == 10;
}

}  // namespace mvd
