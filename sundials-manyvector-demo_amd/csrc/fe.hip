// WENO5 flux divergence fe() kernels for gfx950 (FP64).
//
// Plain path (the comparator; fEuler, utilities.cpp:17-253, transcribed for the GPU):
//   fe_plain_faces : one thread per cell computes the lower x/y/z faces of its cell
//                    (plus the domain's upper boundary faces) with the reference-order
//                    face flux, into SoA face-flux buffers;
//   fe_plain_div   : one thread per cell forms the divergence in the reference order.
// Fast path:
//   fe_xface       : one thread per x-face, stencil gathered from global memory
//                    (neighbouring lanes share cache lines), fast face flux, into an
//                    x-face buffer; also performs the legal-state check of every cell;
//   fe_march<Y>    : lanes along x, each marching a segment of cells along y with a
//                    rolling 6-point window (per-point primitives computed once), fuses
//                    the x and y parts of the divergence and writes wdot;
//   fe_march<Z>    : the same along z, subtracting the z part in place.
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

__device__ inline Pt load_point(const Grid &g, const Fields &f, int d, int i, int j, int k,
                                int e, long &c_out, double &sgn_out) {
  GhostMap m;
  const long c = stencil_cell(g, d, i, j, k, e, m);
  double s[5];
  load_local(g, f, d, c, m, s);
  c_out = c;
  sgn_out = m.sgn_all;
  return make_point(s[0], s[1], s[2], s[3], s[4], g.gamma);
}

// x-faces: one thread per face (nx+1 per row). Fx layout [component][face], global order.
template <int NC>
__global__ void __launch_bounds__(256)
fe_xface(Grid g, Fields f, double *Fx, unsigned *flag) {
  const long nxp = g.nx + 1;
  const long nf = nxp * g.ny * g.nz;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nf) return;
  const int i = (int)(tid % nxp);
  const int j = (int)((tid / nxp) % g.ny);
  const int k = (int)(tid / (nxp * g.ny));
  if (i < g.nx) legal_check(g, f, cell_index(g, i, j, k), flag);
  Pt P[6];
  long cidx[6];
  double sg[6];
#pragma unroll
  for (int l = 0; l < 6; l++) P[l] = load_point(g, f, 0, i, j, k, i - 3 + l, cidx[l], sg[l]);
  double F[5], alpha, un[6];
  face_flux_fast(P, g.gamma, F, alpha, un);
#pragma unroll
  for (int a = 0; a < 5; a++) Fx[a * nf + tid] = F[a];
#pragma unroll 1
  for (int v = 0; v < NC; v++) {
    double cv[6];
#pragma unroll
    for (int l = 0; l < 6; l++) cv[l] = f.chem[(long)NC * cidx[l] + v] * sg[l];
    Fx[(5 + v) * nf + tid] = tracer_face_flux(cv, un, alpha);
  }
}

// March along DIR (1 = y, 2 = z) over segments of SEG cells. Lanes run along x.
// DIR == 1: wdot = 0 - ((Fx[i+1]-Fx[i])/dx + (Fy[j+1]-Fy[j])/dy)
// DIR == 2: wdot = wdot - (Fz[k+1]-Fz[k])/dz
template <int DIR, int NC, int SEG>
__global__ void __launch_bounds__(256)
fe_march(Grid g, Fields f, const double *Fx, FieldsOut out) {
  constexpr int NV = 5 + NC;
  const int n = extent(g, DIR);
  const int nt = (DIR == 1) ? g.nz : g.ny;          // transverse extent
  const int nseg = (n + SEG - 1) / SEG;
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= (long)g.nx * nt * nseg) return;
  const int i = (int)(tid % g.nx);
  const int t = (int)((tid / g.nx) % nt);
  const int sgi = (int)(tid / ((long)g.nx * nt));
  const int c0 = sgi * SEG;
  const int c1 = min(c0 + SEG, n);
  const int jj = (DIR == 1) ? 0 : t, kk = (DIR == 1) ? t : 0;   // transverse coordinates
  const double h = (DIR == 1) ? g.dy : g.dz;
  const long nxp = g.nx + 1;
  const long nfx = nxp * g.ny * g.nz;

  Pt P[6];
  long cidx[6];
  double sg[6];
  // preload points c0-3 .. c0+1
#pragma unroll
  for (int l = 0; l < 5; l++) {
    const int e = c0 - 3 + l;
    P[l + 1] = load_point(g, f, DIR, i, jj, kk, e, cidx[l + 1], sg[l + 1]);
  }
  double Fprev[NV];
  for (int c = c0; c <= c1; c++) {
    // shift window and load point c+2
#pragma unroll
    for (int l = 0; l < 5; l++) { P[l] = P[l + 1]; cidx[l] = cidx[l + 1]; sg[l] = sg[l + 1]; }
    {
      const int e = c + 2;
      P[5] = load_point(g, f, DIR, i, jj, kk, e, cidx[5], sg[5]);
    }
    // face c-1/2
    double Fl[5], alpha, un[6];
    face_flux_fast(P, g.gamma, Fl, alpha, un);
    double Fc[NV];
#pragma unroll
    for (int a = 0; a < 5; a++) Fc[global_comp(DIR, a)] = Fl[a];
#pragma unroll
    for (int v = 0; v < NC; v++) {
      double cv[6];
#pragma unroll
      for (int l = 0; l < 6; l++) cv[l] = f.chem[(long)NC * cidx[l] + v] * sg[l];
      Fc[5 + v] = tracer_face_flux(cv, un, alpha);
    }
    if (c > c0) {
      // divergence for cell c-1
      const int cc = c - 1;
      const int ci = i, cj = (DIR == 1) ? cc : jj, ck = (DIR == 1) ? kk : cc;
      const long cell = cell_index(g, ci, cj, ck);
      if (DIR == 1) {
        const long xf = ci + nxp * (cj + (long)g.ny * ck);
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const double dv = (Fx[v * nfx + xf + 1] - Fx[v * nfx + xf]) / g.dx +
                            (Fc[v] - Fprev[v]) / h;
          if (v < 5) out.w[v][cell] = 0.0 - dv;
          else out.chem[(long)NC * cell + (v - 5)] = 0.0 - dv;
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; v++) {
          const double dv = (Fc[v] - Fprev[v]) / h;
          if (v < 5) out.w[v][cell] = out.w[v][cell] - dv;
          else out.chem[(long)NC * cell + (v - 5)] = out.chem[(long)NC * cell + (v - 5)] - dv;
        }
      }
    }
#pragma unroll
    for (int v = 0; v < NV; v++) Fprev[v] = Fc[v];
  }
}

__global__ void reset_flag(unsigned *flag) { *flag = 0u; }

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------

static inline unsigned nblocks(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

template <int NC>
static int launch_fe(const Grid &g, const Fields &f, const FieldsOut &o, FeScratch &s,
                     hipStream_t st, bool plain) {
  constexpr int SEG = 16;
  const long nxp = g.nx + 1, nyp = g.ny + 1, nzp = g.nz + 1;
  hipLaunchKernelGGL(reset_flag, dim3(1), dim3(1), 0, st, s.flag);
  if (plain) {
    hipLaunchKernelGGL(fe_plain_faces<NC>, dim3(nblocks(nxp * nyp * nzp, 256)), dim3(256), 0, st,
                       g, f, s.Fx, s.Fy, s.Fz, s.flag);
    hipLaunchKernelGGL(fe_plain_div<NC>, dim3(nblocks((long)g.nx * g.ny * g.nz, 256)), dim3(256),
                       0, st, g, s.Fx, s.Fy, s.Fz, o);
  } else {
    hipLaunchKernelGGL(fe_xface<NC>, dim3(nblocks(nxp * g.ny * g.nz, 256)), dim3(256), 0, st,
                       g, f, s.Fx, s.flag);
    const long ty = (long)g.nx * g.nz * ((g.ny + SEG - 1) / SEG);
    hipLaunchKernelGGL((fe_march<1, NC, SEG>), dim3(nblocks(ty, 256)), dim3(256), 0, st, g, f,
                       s.Fx, o);
    const long tz = (long)g.nx * g.ny * ((g.nz + SEG - 1) / SEG);
    hipLaunchKernelGGL((fe_march<2, NC, SEG>), dim3(nblocks(tz, 256)), dim3(256), 0, st, g, f,
                       s.Fx, o);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int fe_launch(const Grid &g, const Fields &f, const FieldsOut &o, FeScratch &s,
              hipStream_t st, bool plain) {
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
