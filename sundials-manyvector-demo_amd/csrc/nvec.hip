// Reductions and small elementwise kernels around fe(): the CFL wave-speed reduction
// (stability, utilities.cpp:483-528), the fslow energy/scaling glue
// (multirate_chem_hydro_main.cpp:1024-1077) and the Dengo scaling helpers
// (primordial_blast.cpp:404-458), and the N_Vector operations of the integrators' stage
// algebra (linear combination, WRMS norm).
//
// Reductions are two-pass and deterministic: each block reduces its grid-stride share
// with 64-lane wavefront shuffles and one LDS exchange, writes one partial; a single
// block reduces the partials.
#include "mvd_internal.hpp"

namespace mvd {

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <bool MAX>
__device__ inline double block_reduce(double v) {
  __shared__ double part[16];
  v = MAX ? wave_max(v) : wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) part[wid] = v;
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  if (wid == 0) {
    v = (lane < nw) ? part[lane] : (MAX ? 0.0 : 0.0);
    v = MAX ? wave_max(v) : wave_sum(v);
  }
  return v;
}

__global__ void __launch_bounds__(256)
wave_speed_partial(Grid g, Fields f, long n, double *partial) {
  double a = 0.0;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
       c += (long)gridDim.x * blockDim.x) {
    const double rho = f.w[0][c], mx = f.w[1][c], my = f.w[2][c], mz = f.w[3][c], e = f.w[4][c];
    const double u = fabs(mx / rho);   // the reference uses |mx/rho| for all three directions
    const double p = (g.gamma - 1.0) * (e - (mx * mx + my * my + mz * mz) * 0.5 / rho);
    const double cs = sun_sqrt(g.gamma * p / rho);
    a = fmax(a, fabs(u + cs));
  }
  a = block_reduce<true>(a);
  if (threadIdx.x == 0) partial[blockIdx.x] = a;
}

__global__ void __launch_bounds__(256) reduce_max_final(const double *partial, int n, double *out) {
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a = fmax(a, partial[i]);
  a = block_reduce<true>(a);
  if (threadIdx.x == 0) out[0] = a;
}

int max_wave_speed_launch(const Grid &g, const Fields &f, FeScratch &s, hipStream_t st,
                          double *host_out) {
  const long n = (long)g.nx * g.ny * g.nz;
  int nb = (int)std::min<long>((n + 255) / 256, 2048);
  if (nb < 1) nb = 1;
  if (s.red_len < nb + 1) return -3;
  hipLaunchKernelGGL(wave_speed_partial, dim3(nb), dim3(256), 0, st, g, f, n, s.red);
  hipLaunchKernelGGL(reduce_max_final, dim3(1), dim3(256), 0, st, s.red, nb, s.red + nb);
  if (hipMemcpyAsync(host_out, s.red + nb, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
    return -3;
  if (hipStreamSynchronize(st) != hipSuccess) return -3;
  return 0;
}

// fslow, first half: chem *= scale; et <- chem[nchem-1]/EnergyUnits + 0.5/rho |m|^2. One
// thread per chemistry value (both arrays read and written coalesced); the thread of each
// cell's last value (the gas energy) also sets that cell's et.
__global__ void __launch_bounds__(256)
fslow_prepare(long n, int nchem, const double *rho, const double *mx, const double *my,
              const double *mz, double *et, double *chem, const double *scale,
              double eunit_inv) {
#pragma clang fp contract(off)
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nchem) return;
  const double v = chem[e] * scale[e];
  chem[e] = v;
  const long c = e / nchem;
  if (e - c * nchem != nchem - 1) return;
  const double r = rho[c], a = mx[c], b = my[c], d = mz[c];
  et[c] = v * eunit_inv + 0.5 / r * (a * a + b * b + d * d);
}

// fslow, second half: chemdot[nchem-1] <- etdot; etdot <- 0
__global__ void __launch_bounds__(256)
fslow_finish(long n, int nchem, double *etdot, double *chemdot) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  chemdot[(long)nchem * c + nchem - 1] = etdot[c] * 1.0;
  etdot[c] = 0.0;
}

__global__ void __launch_bounds__(256)
chem_scale(long n, double *chem, const double *scale, int apply) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (apply) chem[i] *= scale[i];
  else chem[i] /= scale[i];
}

// chemistry rates of one slow / explicit RHS into the scaled variables the chemistry is
// integrated in: species divided by their scale, the gas-energy slot (a code-unit energy
// rate) times EnergyUnits divided by the ge scale
// (one thread per value, coalesced)
__global__ void __launch_bounds__(256)
chem_forcing_to_scaled(long m, double *Fc, const double *scale, double eunits) {
#pragma clang fp contract(off)
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  if (e % 10 != 9) Fc[e] /= scale[e];
  else Fc[e] = Fc[e] * eunits / scale[e];
}

int chem_forcing_to_scaled_launch(long n, double *Fc, const double *scale, double eunits,
                                  hipStream_t st) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(chem_forcing_to_scaled, dim3((unsigned)((10 * n + 255) / 256)), dim3(256), 0,
                     st, 10 * n, Fc, scale, eunits);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int fslow_prepare_launch(const Grid &g, double *const *w, const double *scale,
                         double eunit_inv, hipStream_t st) {
  const long n = (long)g.nx * g.ny * g.nz;
  if (g.nchem < 1) return 0;
  hipLaunchKernelGGL(fslow_prepare, dim3((unsigned)((n * g.nchem + 255) / 256)), dim3(256), 0,
                     st, n, g.nchem, (const double *)w[0], (const double *)w[1],
                     (const double *)w[2], (const double *)w[3], w[4], w[5], scale, eunit_inv);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int fslow_finish_launch(const Grid &g, double *const *wdot, hipStream_t st) {
  const long n = (long)g.nx * g.ny * g.nz;
  hipLaunchKernelGGL(fslow_finish, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     g.nchem, wdot[4], wdot[5]);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

__global__ void __launch_bounds__(256)
nvec_const(long n, double c, double *z) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = c;
}

int const_launch(double c, double *z, long len, hipStream_t st) {
  if (len <= 0) return 0;
  hipLaunchKernelGGL(nvec_const, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, len, c, z);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int chem_scale_launch(long n, double *chem, const double *scale, int apply, hipStream_t st) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(chem_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, chem,
                     scale, apply);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

// ---------------------------------------------------------------------------------
// N_Vector operations used by the integrators' stage algebra on the ManyVector's device
// arrays: N_VLinearCombination (z = sum_i c_i X_i, summed in index order; z may be X_0) and
// N_VWrmsNorm of a ManyVector (sqrt(sum over every subvector of (x w)^2 / total length),
// two-pass and deterministic).
// ---------------------------------------------------------------------------------
struct LcArgs {
  const double *x[LC_MAX];
  double c[LC_MAX];
};

__global__ void __launch_bounds__(256)
lin_comb(long len, int nv, LcArgs a, double *z, int add) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double acc = add ? z[i] : a.c[0] * a.x[0][i];
  for (int k = add ? 0 : 1; k < nv; k++) acc += a.c[k] * a.x[k][i];
  z[i] = acc;
}

// y = 1 y + a x, nsteps times, each value's additions in turn: the values of nsteps
// lin_comb launches with c = {1, a}, X = {y, x}
__global__ void __launch_bounds__(256)
axpy_steps(long len, int nsteps, double a, const double *x, double *y) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const double ax = a * x[i];
  double v = y[i];
  for (int k = 0; k < nsteps; k++) v = 1.0 * v + ax;
  y[i] = v;
}

int axpy_steps_launch(int nsteps, double a, const double *x, double *y, long len,
                      hipStream_t st) {
  if (len <= 0 || nsteps <= 0) return 0;
  hipLaunchKernelGGL(axpy_steps, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, len,
                     nsteps, a, x, y);
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

int linear_combination_launch(int nv, const double *c, const double *const *X, double *z,
                              long len, hipStream_t st) {
  if (len <= 0) return 0;
  const unsigned nb = (unsigned)((len + 255) / 256);
  for (int k0 = 0; k0 < nv; k0 += LC_MAX) {
    LcArgs a{};
    const int m = nv - k0 < LC_MAX ? nv - k0 : LC_MAX;
    for (int k = 0; k < m; k++) {
      a.x[k] = X[k0 + k];
      a.c[k] = c[k0 + k];
    }
    hipLaunchKernelGGL(lin_comb, dim3(nb), dim3(256), 0, st, len, m, a, z, k0 > 0 ? 1 : 0);
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

struct WrmsArgs {
  const double *x[WRMS_MAXSUB];
  const double *w[WRMS_MAXSUB];
  long len[WRMS_MAXSUB];
};

__global__ void __launch_bounds__(256)
wrms_partial(int nsub, WrmsArgs a, double *partial) {
#pragma clang fp contract(off)
  double s = 0.0;
  for (int q = 0; q < nsub; q++)
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.len[q];
         i += (long)gridDim.x * blockDim.x) {
      const double p = a.x[q][i] * a.w[q][i];
      s += p * p;
    }
  s = block_reduce<false>(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void __launch_bounds__(256)
wrms_final(const double *partial, int n, double inv_len, double *out) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
  s = block_reduce<false>(s);
  if (threadIdx.x == 0) out[0] = sqrt(s * inv_len);
}

int wrms_norm_launch(int nsub, const double *const *x, const double *const *w, const long *len,
                     FeScratch &s, hipStream_t st, double *host_out) {
  WrmsArgs a{};
  long total = 0, maxlen = 0;
  for (int q = 0; q < nsub; q++) {
    a.x[q] = x[q];
    a.w[q] = w[q];
    a.len[q] = len[q];
    total += len[q];
    if (len[q] > maxlen) maxlen = len[q];
  }
  if (total <= 0) { *host_out = 0.0; return 0; }
  int nb = (int)std::min<long>((maxlen + 255) / 256, 2048);
  if (nb < 1) nb = 1;
  if (s.red_len < nb + 1) return -3;
  hipLaunchKernelGGL(wrms_partial, dim3(nb), dim3(256), 0, st, nsub, a, s.red);
  hipLaunchKernelGGL(wrms_final, dim3(1), dim3(256), 0, st, (const double *)s.red, nb,
                     1.0 / (double)total, s.red + nb);
  if (hipMemcpyAsync(host_out, s.red + nb, sizeof(double), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return -3;
  return 0;
}

}  // namespace mvd
