// Ghost exchange of one subdomain on the device: the receive buffers of
// EulerData::ExchangeStart/ExchangeEnd (euler3D.hpp:577-1191) and the stencil packs
// pack1D_{x,y,z}[_bdry] (euler3D.hpp:1197-1378), for the single rank that owns the whole
// grid. The fe() kernels read ghost points through ghost_map directly and never build
// these buffers; the entry points here expose the same ghost rules in the reference's
// buffer layouts (for callers that consume them and for the index-coded exchange test of
// communication_test_main.cpp).
//
// Buffers, NV = 5 + nchem values per point, slot s = 0..2 three points deep
// (BUFINDX(v,s,a,b,NV,3,na,nb) = v + NV*(s + 3*(a + na*b)), euler3D.hpp:65):
//   W, E: (s, j, k)   S, N: (s, i, k)   B, F: (s, i, j)
// Slot s of a low face (W/S/B) holds extended index s - 3, of a high face n + s: periodic
// boundaries wrap (the rank is its own neighbour); physical ones apply ghost_map (low faces
// mirror, high faces copy w[n-3+s]; reflecting negates the normal momentum, Dirichlet every
// value).
#include "mvd_internal.hpp"

namespace mvd {

namespace {

__device__ inline double point_value(const Grid &g, const Fields &f, int v, long c) {
  return v < 5 ? f.w[v][c] : f.chem[(long)g.nchem * c + (v - 5)];
}

// value v of the point at extended index e along d through (i, j, k), ghost signs applied
__device__ inline double ghost_value(const Grid &g, const Fields &f, int d, int v, int i, int j,
                                     int k, int e) {
  const GhostMap m = ghost_map(g, d, e);
  if (d == 0) i = m.src; else if (d == 1) j = m.src; else k = m.src;
  const long c = i + (long)g.nx * (j + (long)g.ny * k);
  const double x = point_value(g, f, v, c);
  return x * (v == 1 + d ? m.sgn_norm : m.sgn_all);
}

// one thread per (slot, a, b) of one face; writes its NV consecutive values
__global__ void __launch_bounds__(256)
exchange_face(Grid g, Fields f, int face, double *buf) {
  const int d = face >> 1;
  const bool high = face & 1;
  const int na = d == 0 ? g.ny : g.nx;
  const int nb = d == 2 ? g.ny : g.nz;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3L * na * nb) return;
  const int s = (int)(t % 3);
  const int a = (int)((t / 3) % na);
  const int b = (int)(t / (3L * na));
  const int n = extent(g, d);
  const int e = high ? n + s : s - 3;
  int i, j, k;
  if (d == 0) { i = 0; j = a; k = b; }
  else if (d == 1) { i = a; j = 0; k = b; }
  else { i = a; j = b; k = 0; }
  const int nv = 5 + g.nchem;
  double *o = buf + (long)nv * t;
  for (int v = 0; v < nv; v++) o[v] = ghost_value(g, f, d, v, i, j, k, e);
}

// one thread per (cell, stencil point): out[((q * 6) + l) * NV + v] = w(c - 3 + l) along d
__global__ void __launch_bounds__(256)
pack1d_kernel(Grid g, Fields f, int d, long ncell, const int *ijk, double *out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell * 6) return;
  const long q = t / 6;
  const int l = (int)(t % 6);
  const int i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
  const int c = d == 0 ? i : (d == 1 ? j : k);
  const int nv = 5 + g.nchem;
  double *o = out + (long)nv * t;
  for (int v = 0; v < nv; v++) o[v] = ghost_value(g, f, d, v, i, j, k, c - 3 + l);
}

__global__ void __launch_bounds__(256)
check_cells(Grid g, long ncell, const int *ijk, int *bad) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncell) return;
  const int i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
  if (i < 0 || i >= g.nx || j < 0 || j >= g.ny || k < 0 || k >= g.nz) atomicOr(bad, 1);
}

}  // namespace

int exchange_launch(const Grid &g, const Fields &f, double *const *recv, hipStream_t st) {
  for (int face = 0; face < 6; face++) {
    if (!recv[face]) continue;
    const int d = face >> 1;
    const long na = d == 0 ? g.ny : g.nx;
    const long nb = d == 2 ? g.ny : g.nz;
    const long nt = 3 * na * nb;
    hipLaunchKernelGGL(exchange_face, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, g, f,
                       face, recv[face]);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// scratch: one int on the device (the bounds check of the cell list)
int pack1d_launch(const Grid &g, const Fields &f, int d, long ncell, const int *ijk,
                  double *out, int *scratch, hipStream_t st, int *host_bad) {
  if (ncell <= 0) return 0;
  const unsigned nbq = (unsigned)((ncell + 255) / 256);
  if (hipMemsetAsync(scratch, 0, sizeof(int), st) != hipSuccess) return -1;
  hipLaunchKernelGGL(check_cells, dim3(nbq), dim3(256), 0, st, g, ncell, ijk, scratch);
  int bad = 0;
  if (hipMemcpyAsync(&bad, scratch, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return -1;
  *host_bad = bad;
  if (bad) return 0;
  const long nt = ncell * 6;
  hipLaunchKernelGGL(pack1d_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, g, f, d,
                     ncell, ijk, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace mvd
