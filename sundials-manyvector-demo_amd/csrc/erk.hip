// Fixed-step explicit Runge-Kutta evolution of fEuler on the device: the path of the
// fluid driver with an explicit table and a fixed step (euler3D_main.cpp:194 ARKStepCreate
// with fEuler explicit, :209 ARKStepSetOrder, :280 ARKStepSetFixedStep). ARKStep's stage
// loop (SUNDIALS 6.2.0 arkode_erkstep.c; not in the reference tree) is restated:
//   z_i = y_n + sum_{j<i} (h a_ij) F_j,  F_i = fe(z_i),  y_{n+1} = y_n + sum_i (h b_i) F_i.
// The state, the stage state and the stage derivatives live in contiguous device blocks
// ([rho][mx][my][mz][et][chem], the ManyVector's subvectors back to back), so every stage
// combination is one fused kernel over the whole vector. One step (its fe() launches,
// combinations and legal-state bookkeeping) is captured once as a hipGraph on a private
// stream and replayed on the caller's stream; the legal-state bits of every stage are
// OR-ed into one word that is read back once, after the last step.
#include "mvd_internal.hpp"

namespace mvd {

namespace {

// z = y + sum_{j<nt} c_j K_j, K_j = K + j*kstride (z may equal y)
struct Coefs { double c[ERK_MAXSTAGE]; };

__global__ void __launch_bounds__(256)
mv_lincomb(long len, const double *y, int nt, Coefs c, const double *K, long kstride, double *z) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double acc = y[i];
  for (int j = 0; j < nt; j++) acc += c.c[j] * K[j * kstride + i];
  z[i] = acc;
}

__global__ void zero_flag(unsigned *dst) { *dst = 0u; }

inline Fields fields_at(const double *base, long ncell, int nchem) {
  Fields f;
  for (int q = 0; q < 5; q++) f.w[q] = base + q * ncell;
  f.chem = nchem > 0 ? base + 5 * ncell : nullptr;
  return f;
}

inline FieldsOut fields_out_at(double *base, long ncell, int nchem) {
  FieldsOut f;
  for (int q = 0; q < 5; q++) f.w[q] = base + q * ncell;
  f.chem = nchem > 0 ? base + 5 * ncell : nullptr;
  return f;
}

bool same_table(const ErkTable &a, const ErkTable &b) {
  if (a.stages != b.stages) return false;
  for (int k = 0; k < ERK_MAXSTAGE * ERK_MAXSTAGE; k++)
    if (a.A[k] != b.A[k]) return false;
  for (int k = 0; k < ERK_MAXSTAGE; k++)
    if (a.b[k] != b.b[k]) return false;
  return true;
}

// enqueue one step on st (captured)
int enqueue_step(const Grid &g, FeScratch &fe, ErkState &e, const ErkTable &tab, double h,
                 hipStream_t st) {
  const long ncell = (long)g.nx * g.ny * g.nz;
  const long len = (5L + g.nchem) * ncell;
  const unsigned nb = (unsigned)((len + 255) / 256);
  for (int i = 0; i < tab.stages; i++) {
    const double *zin = e.Y;
    if (i > 0) {
      Coefs c{};
      for (int j = 0; j < i; j++) c.c[j] = h * tab.A[i * ERK_MAXSTAGE + j];
      hipLaunchKernelGGL(mv_lincomb, dim3(nb), dim3(256), 0, st, len, (const double *)e.Y, i, c,
                         (const double *)e.K, len, e.Z);
      zin = e.Z;
    }
    // the stage's legal-state bits go straight into the run's flag
    const int rc = fe_launch(g, fields_at(zin, ncell, g.nchem),
                             fields_out_at(e.K + (long)i * len, ncell, g.nchem), fe, st, false,
                             e.flag);
    if (rc) return rc;
  }
  Coefs c{};
  for (int j = 0; j < tab.stages; j++) c.c[j] = h * tab.b[j];
  hipLaunchKernelGGL(mv_lincomb, dim3(nb), dim3(256), 0, st, len, (const double *)e.Y,
                     tab.stages, c, (const double *)e.K, len, e.Y);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

void erk_free(ErkState &e) {
  if (e.exec) (void)hipGraphExecDestroy(e.exec);
  if (e.cap) (void)hipStreamDestroy(e.cap);
  if (e.Y) (void)hipFree(e.Y);
  if (e.flag) (void)hipFree(e.flag);
  e = ErkState{};
}

int erk_evolve(const Grid &g, FeScratch &fe, ErkState &e, const ErkTable &tab, double h,
               int nsteps, const Fields &w_in, const FieldsOut &w_out, hipStream_t st,
               unsigned *host_flag) {
  const long ncell = (long)g.nx * g.ny * g.nz;
  const long len = (5L + g.nchem) * ncell;
  if (!e.Y) {
    // Y, Z, K_0..K_{MAXSTAGE-1}
    if (hipMalloc(&e.Y, sizeof(double) * len * (2 + ERK_MAXSTAGE)) != hipSuccess) return -1;
    if (hipMalloc(&e.flag, sizeof(unsigned)) != hipSuccess) return -1;
    e.Z = e.Y + len;
    e.K = e.Z + len;
    if (hipStreamCreateWithFlags(&e.cap, hipStreamNonBlocking) != hipSuccess) return -1;
  }
  // state in
  for (int q = 0; q < 5; q++)
    if (hipMemcpyAsync(e.Y + q * ncell, w_in.w[q], sizeof(double) * ncell,
                       hipMemcpyDeviceToDevice, st) != hipSuccess)
      return -1;
  if (g.nchem > 0 &&
      hipMemcpyAsync(e.Y + 5 * ncell, w_in.chem, sizeof(double) * g.nchem * ncell,
                     hipMemcpyDeviceToDevice, st) != hipSuccess)
    return -1;
  hipLaunchKernelGGL(zero_flag, dim3(1), dim3(1), 0, st, e.flag);
  // (re)capture the step when the table or the step size changes
  if (!e.exec || e.h != h || !same_table(e.tab, tab)) {
    if (e.exec) { (void)hipGraphExecDestroy(e.exec); e.exec = nullptr; }
    const bool timer_on = fe.timer.on;
    fe.timer.on = false;                      // no event records inside the graph
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(e.cap, hipStreamCaptureModeThreadLocal) != hipSuccess) return -1;
    const int rc = enqueue_step(g, fe, e, tab, h, e.cap);
    const hipError_t ec = hipStreamEndCapture(e.cap, &graph);
    fe.timer.on = timer_on;
    if (rc || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc == -2 ? -2 : -1;
    }
    const hipError_t ei = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) { e.exec = nullptr; return -1; }
    e.h = h;
    e.tab = tab;
  }
  for (int n = 0; n < nsteps; n++)
    if (hipGraphLaunch(e.exec, st) != hipSuccess) return -1;
  // state out, flag back
  for (int q = 0; q < 5; q++)
    if (hipMemcpyAsync(w_out.w[q], e.Y + q * ncell, sizeof(double) * ncell,
                       hipMemcpyDeviceToDevice, st) != hipSuccess)
      return -1;
  if (g.nchem > 0 &&
      hipMemcpyAsync(w_out.chem, e.Y + 5 * ncell, sizeof(double) * g.nchem * ncell,
                     hipMemcpyDeviceToDevice, st) != hipSuccess)
    return -1;
  if (hipMemcpyAsync(host_flag, e.flag, sizeof(unsigned), hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return -1;
  return 0;
}

}  // namespace mvd
