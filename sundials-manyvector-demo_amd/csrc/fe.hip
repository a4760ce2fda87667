// WENO5 flux divergence fe() kernels for gfx950 (FP64).
//
// Plain path (the comparator; fEuler, utilities.cpp:17-253, transcribed for the GPU):
//   fe_plain_faces : one thread per cell computes the lower x/y/z faces of its cell
//                    (plus the domain's upper boundary faces) with the reference-order
//                    face flux, into SoA face-flux buffers;
//   fe_plain_div   : one thread per cell forms the divergence in the reference order.
// Fast path (three launches):
//   fe_xface       : x-faces, one thread per face; a block stages the primitives of the
//                    points its faces need in LDS once, then every thread reads its
//                    6-point stencil from LDS; also performs the legal-state check;
//   fe_yface       : y-faces the same way, tiles of 16 x-cells by 16 faces;
//   fe_zface_div   : z-faces the same way, plus the divergence of the 15 cells between a
//                    tile's 16 faces (tiles overlap by one face), reading the x- and
//                    y-face fluxes, and writes wdot.
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

// x-faces. A block holds XROWS rows x 64 consecutive faces. The 69 points the 64 faces
// need (faces i0..i0+63 use points i0-3..i0+65) are converted once into LDS (SoA), then
// each thread computes one face from LDS. Fx layout [component][face], global order.
constexpr int XROWS = 4;
constexpr int XPTS = 64 + 5;

template <int NC>
__global__ void __launch_bounds__(256)
fe_xface(Grid g, Fields f, double *Fx, unsigned *flag) {
  __shared__ double sh[9][XROWS][XPTS];
  __shared__ int scell[XROWS][XPTS];
  __shared__ double ssgn[XROWS][XPTS];
  const int lane = threadIdx.x & 63, row = threadIdx.x >> 6;
  const int nxp = g.nx + 1;
  const int nchunk = (nxp + 63) >> 6;
  const int chunk = blockIdx.x % nchunk;
  const int line = (blockIdx.x / nchunk) * XROWS + row;      // j + ny*k
  const int nline = g.ny * g.nz;
  const bool row_ok = line < nline;
  const int i0 = chunk * 64;
  // stage points i0-3+s, s = lane and (for lanes < 5) 64+lane
  // the block's XROWS x XPTS points, spread over all threads of the block
  for (int q2 = threadIdx.x; q2 < XROWS * XPTS; q2 += blockDim.x) {
    const int rw = q2 / XPTS, s = q2 - rw * XPTS;
    const int ln = (blockIdx.x / nchunk) * XROWS + rw;
    const int e = i0 - 3 + s;
    if (ln < nline && e <= g.nx + 2) {
      double sa, sn;
      const int c = line_cell<0>(g, ln * g.nx, 1, g.nx, e, sa, sn);
      const RawPt q = load_raw<0>(f, c, sa, sn);
      const Pt P = to_point(q, g.gamma);
      sh[0][rw][s] = P.r;  sh[1][rw][s] = P.mn; sh[2][rw][s] = P.m1;
      sh[3][rw][s] = P.m2; sh[4][rw][s] = P.e;  sh[5][rw][s] = P.p;
      sh[6][rw][s] = P.ir; sh[7][rw][s] = P.a;  sh[8][rw][s] = P.sr;
      scell[rw][s] = c;
      ssgn[rw][s] = sa;
      if (e >= 0 && e < g.nx) {   // each cell is checked once (legal_state)
        unsigned bits = 0;
        if (!(q.r > 0.0)) bits |= 1u;
        if (!(q.e > 0.0)) bits |= 2u;
        if (!(P.p > 0.0)) bits |= 4u;
        if (bits) atomicOr(flag, bits);
      }
    }
  }
  __syncthreads();
  const int i = i0 + lane;
  if (!row_ok || i > g.nx) return;
  const double *pbase = &sh[0][row][lane];
  double F[5], alpha;
  face_flux_lds(pbase, XROWS * XPTS, 1, g.gamma, F, alpha);
  const long nf = (long)nxp * nline;
  const long face = (long)line * nxp + i;
#pragma unroll
  for (int a = 0; a < 5; a++) Fx[a * nf + face] = F[a];
  double un[6];
  if (NC > 0) lds_velocities(pbase, XROWS * XPTS, 1, un);
#pragma unroll 1
  for (int v = 0; v < NC; v++) {
    double cv[6];
#pragma unroll
    for (int l = 0; l < 6; l++)
      cv[l] = f.chem[(long)NC * scell[row][lane + l] + v] * ssgn[row][lane + l];
    Fx[(5 + v) * nf + face] = tracer_face_flux(cv, un, alpha);
  }
}

// y- and z-faces: one thread per face; a block covers TX consecutive x-cells and TF
// consecutive faces along the sweep direction. The TF+5 points of each column are
// converted once into LDS (SoA, conflict-free), then every thread reads its 6-point
// stencil from LDS.
struct TileLds {
  double *pt;      // [9][TPTS][TX]
  int *cell;       // [TPTS][TX]
  double *sgn;     // [TPTS][TX]
};

template <int D, int TX, int TF>
__device__ inline void stage_tile(const Grid &g, const Fields &f, const TileLds &L, int x0,
                                  int t, int e0) {
  constexpr int TPTS = TF + 5;
  const int n = (D == 1) ? g.ny : g.nz;
  const int stride = (D == 1) ? g.nx : g.nx * g.ny;
  for (int s = threadIdx.x; s < TPTS * TX; s += blockDim.x) {
    const int r = s / TX, x = s - r * TX;
    const int i = x0 + x;
    const int e = e0 - 3 + r;
    if (i < g.nx && e <= n + 2) {
      const int base = (D == 1) ? i + g.nx * g.ny * t : i + g.nx * t;
      double sa, sn;
      const int c = line_cell<D>(g, base, stride, n, e, sa, sn);
      const Pt P = to_point(load_raw<D>(f, c, sa, sn), g.gamma);
      double *pt = L.pt + r * TX + x;
      constexpr int FS = TPTS * TX;
      pt[0 * FS] = P.r;  pt[1 * FS] = P.mn; pt[2 * FS] = P.m1;
      pt[3 * FS] = P.m2; pt[4 * FS] = P.e;  pt[5 * FS] = P.p;
      pt[6 * FS] = P.ir; pt[7 * FS] = P.a;  pt[8 * FS] = P.sr;
      L.cell[r * TX + x] = c;
      L.sgn[r * TX + x] = sa;
    }
  }
}

// local (rho, m_normal, m_t1, m_t2, e) flux of direction D into global component order
template <int D>
__device__ inline void to_global(const double (&Fl)[5], double *Fg) {
  Fg[0] = Fl[0];
  Fg[4] = Fl[4];
  if (D == 1) { Fg[2] = Fl[1]; Fg[1] = Fl[2]; Fg[3] = Fl[3]; }
  else { Fg[3] = Fl[1]; Fg[2] = Fl[2]; Fg[1] = Fl[3]; }
}

// y-faces -> Fy [component][i + nx*(j + (ny+1)*k)]
template <int NC, int TX, int TF>
__global__ void __launch_bounds__(256)
fe_yface(Grid g, Fields f, double *Fy) {
  constexpr int TPTS = TF + 5;
  __shared__ double s_pt[9 * TPTS * TX];
  __shared__ int s_cell[TPTS * TX];
  __shared__ double s_sgn[TPTS * TX];
  const TileLds L{s_pt, s_cell, s_sgn};
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int nyp = g.ny + 1;
  const int nxc = (g.nx + TX - 1) / TX, nfc = (nyp + TF - 1) / TF;
  int b = blockIdx.x;
  const int xc = b % nxc;
  b /= nxc;
  const int fc = b % nfc;
  const int k = b / nfc;
  const int j0 = fc * TF;
  stage_tile<1, TX, TF>(g, f, L, xc * TX, k, j0);
  __syncthreads();
  const int i = xc * TX + tx, j = j0 + ty;
  if (i >= g.nx || j > g.ny) return;
  const double *pbase = L.pt + ty * TX + tx;
  double Fl[5], alpha, un[6], Fg[5];
  face_flux_lds(pbase, TPTS * TX, TX, g.gamma, Fl, alpha);
  if (NC > 0) lds_velocities(pbase, TPTS * TX, TX, un);
  to_global<1>(Fl, Fg);
  const long nf = (long)g.nx * nyp * g.nz;
  const long face = i + (long)g.nx * (j + (long)nyp * k);
#pragma unroll
  for (int a = 0; a < 5; a++) Fy[a * nf + face] = Fg[a];
#pragma unroll 1
  for (int v = 0; v < NC; v++) {
    double cv[6];
#pragma unroll
    for (int l = 0; l < 6; l++)
      cv[l] = f.chem[(long)NC * s_cell[(ty + l) * TX + tx] + v] * s_sgn[(ty + l) * TX + tx];
    Fy[(5 + v) * nf + face] = tracer_face_flux(cv, un, alpha);
  }
}

// z-faces with the divergence fused: a block computes TF z-faces per column and forms
// wdot for the TF-1 cells between them (tiles overlap by one face):
//   wdot = 0 - (((Fx[i+1]-Fx[i])/dx + (Fy[j+1]-Fy[j])/dy) + (Fz[k+1]-Fz[k])/dz)
template <int NC, int TX, int TF>
__global__ void __launch_bounds__(256)
fe_zface_div(Grid g, Fields f, const double *Fx, const double *Fy, FieldsOut out, double idx,
             double idy, double idz) {
  constexpr int NV = 5 + NC;
  constexpr int TPTS = TF + 5;
  constexpr int TC = TF - 1;
  constexpr int PT_SZ = 9 * TPTS * TX, FL_SZ = NV * TF * TX;
  __shared__ double s_buf[PT_SZ > FL_SZ ? PT_SZ : FL_SZ];
  __shared__ int s_cell[TPTS * TX];
  __shared__ double s_sgn[TPTS * TX];
  const TileLds L{s_buf, s_cell, s_sgn};
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const int nxc = (g.nx + TX - 1) / TX, ncc = (g.nz + TC - 1) / TC;
  int b = blockIdx.x;
  const int xc = b % nxc;
  b /= nxc;
  const int cc = b % ncc;
  const int j = b / ncc;
  const int k0 = cc * TC;
  stage_tile<2, TX, TF>(g, f, L, xc * TX, j, k0);
  __syncthreads();
  const int i = xc * TX + tx, k = k0 + ty;
  const bool face_ok = i < g.nx && k <= g.nz;
  double Fz[NV];
  if (face_ok) {
    const double *pbase = L.pt + ty * TX + tx;
    double Fl[5], alpha, un[6];
    face_flux_lds(pbase, TPTS * TX, TX, g.gamma, Fl, alpha);
    if (NC > 0) lds_velocities(pbase, TPTS * TX, TX, un);
    to_global<2>(Fl, Fz);
#pragma unroll 1
    for (int v = 0; v < NC; v++) {
      double cv[6];
#pragma unroll
      for (int l = 0; l < 6; l++)
        cv[l] = f.chem[(long)NC * s_cell[(ty + l) * TX + tx] + v] * s_sgn[(ty + l) * TX + tx];
      Fz[5 + v] = tracer_face_flux(cv, un, alpha);
    }
  }
  __syncthreads();   // the point tile is dead; reuse its LDS for the face fluxes
  if (face_ok) {
#pragma unroll
    for (int v = 0; v < NV; v++) s_buf[(v * TF + ty) * TX + tx] = Fz[v];
  }
  __syncthreads();
  if (ty >= TC || i >= g.nx || k >= g.nz) return;
  const long cell = i + (long)g.nx * (j + (long)g.ny * k);
  const int nxp = g.nx + 1, nyp = g.ny + 1;
  const long nfx = (long)nxp * g.ny * g.nz, nfy = (long)g.nx * nyp * g.nz;
  const long xf = (long)nxp * (j + (long)g.ny * k) + i;
  const long yf = i + (long)g.nx * (j + (long)nyp * k);
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double dx = (Fx[v * nfx + xf + 1] - Fx[v * nfx + xf]) * idx;
    const double dy = (Fy[v * nfy + yf + g.nx] - Fy[v * nfy + yf]) * idy;
    const double dz = (s_buf[(v * TF + ty + 1) * TX + tx] - s_buf[(v * TF + ty) * TX + tx]) * idz;
    const double o = 0.0 - ((dx + dy) + dz);
    if (v < 5) out.w[v][cell] = o;
    else out.chem[(long)NC * cell + (v - 5)] = o;
  }
}

__global__ void reset_flag(unsigned *flag) { *flag = 0u; }

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------

static inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

constexpr int TILE_X = 16, TILE_F = 16;

template <int NC>
static int launch_fe(const Grid &g, const Fields &f, const FieldsOut &o, FeScratch &s,
                     hipStream_t st, bool plain) {
  const long nxp = g.nx + 1, nyp = g.ny + 1, nzp = g.nz + 1;
  hipLaunchKernelGGL(reset_flag, dim3(1), dim3(1), 0, st, s.flag);
  KernelTimer &tm = s.timer;
  if (plain) {
    tm.begin(PROF_PLAIN_FACES, st);
    hipLaunchKernelGGL(fe_plain_faces<NC>, dim3(nblocks(nxp * nyp * nzp, 256)), dim3(256), 0, st,
                       g, f, s.Fx, s.Fy, s.Fz, s.flag);
    tm.end(PROF_PLAIN_FACES, st);
    tm.begin(PROF_PLAIN_DIV, st);
    hipLaunchKernelGGL(fe_plain_div<NC>, dim3(nblocks((long)g.nx * g.ny * g.nz, 256)), dim3(256),
                       0, st, g, s.Fx, s.Fy, s.Fz, o);
    tm.end(PROF_PLAIN_DIV, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  const long nchunk = (nxp + 63) / 64;
  const long nrowblk = ((long)g.ny * g.nz + XROWS - 1) / XROWS;
  tm.begin(PROF_XFACE, st);
  hipLaunchKernelGGL(fe_xface<NC>, dim3((unsigned)(nchunk * nrowblk)), dim3(64 * XROWS), 0, st,
                     g, f, s.Fx, s.flag);
  tm.end(PROF_XFACE, st);
  constexpr int TX = TILE_X, TF = TILE_F;
  const long nxc = (g.nx + TX - 1) / TX;
  tm.begin(PROF_YFACE, st);
  hipLaunchKernelGGL((fe_yface<NC, TX, TF>), dim3((unsigned)(nxc * ((nyp + TF - 1) / TF) * g.nz)),
                     dim3(TX * TF), 0, st, g, f, s.Fy);
  tm.end(PROF_YFACE, st);
  tm.begin(PROF_ZFACE_DIV, st);
  hipLaunchKernelGGL((fe_zface_div<NC, TX, TF>),
                     dim3((unsigned)(nxc * ((g.nz + TF - 2) / (TF - 1)) * g.ny)), dim3(TX * TF),
                     0, st, g, f, s.Fx, s.Fy, o, 1.0 / g.dx, 1.0 / g.dy, 1.0 / g.dz);
  tm.end(PROF_ZFACE_DIV, st);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fe_launch(const Grid &g, const Fields &f, const FieldsOut &o, FeScratch &s,
              hipStream_t st, bool plain) {
  // 32-bit cell indices in the fast kernels
  if ((long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1) >= (1L << 31) / 16) return -2;
  switch (g.nchem) {
    case 0: return launch_fe<0>(g, f, o, s, st, plain);
    case 1: return launch_fe<1>(g, f, o, s, st, plain);
    case 2: return launch_fe<2>(g, f, o, s, st, plain);
    case 4: return launch_fe<4>(g, f, o, s, st, plain);
    case 6: return launch_fe<6>(g, f, o, s, st, plain);
    case 10: return launch_fe<10>(g, f, o, s, st, plain);
    default: return -2;
  }
}

bool fe_nchem_supported(int nchem) {
  return nchem == 0 || nchem == 1 || nchem == 2 || nchem == 4 || nchem == 6 || nchem == 10;
}

}  // namespace mvd
