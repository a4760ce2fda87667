// Internal (non-ABI) declarations shared by the kernel translation units and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include "mvd_device.hpp"

namespace mvd {

struct Fields {
  const double *w[5];   // rho, mx, my, mz, et
  const double *chem;   // species, AoS [cell][nchem] (may be null when nchem == 0)
};

struct FieldsOut {
  double *w[5];
  double *chem;
};

// Kernel timing slots (HIP events recorded on the launch stream around each kernel).
// (slots 0 and 1 are unused; the numbering is part of mvd_profile_read's interface)
enum { PROF_UNUSED0 = 0, PROF_UNUSED1, PROF_ZFACE_DIV, PROF_PLAIN_FACES, PROF_PLAIN_DIV,
       PROF_XYFACE_DIV, PROF_NSLOT };

struct KernelTimer {
  bool on = false;
  hipEvent_t ev[2 * PROF_NSLOT] = {};
  bool pending[PROF_NSLOT] = {};
  double ms[PROF_NSLOT] = {};
  long count[PROF_NSLOT] = {};
  void begin(int slot, hipStream_t st) {
    if (on) (void)hipEventRecord(ev[2 * slot], st);
  }
  void end(int slot, hipStream_t st) {
    if (on) { (void)hipEventRecord(ev[2 * slot + 1], st); pending[slot] = true; }
  }
  // after the stream has been synchronized
  void collect() {
    if (!on) return;
    for (int s = 0; s < PROF_NSLOT; s++) {
      if (!pending[s]) continue;
      float t = 0.f;
      if (hipEventElapsedTime(&t, ev[2 * s], ev[2 * s + 1]) == hipSuccess) {
        ms[s] += t;
        count[s] += 1;
      }
      pending[s] = false;
    }
  }
};

// Device scratch owned by the context (EulerData keeps xflux/yflux/zflux for the same
// purpose, euler3D.hpp:496-499).
struct FeScratch {
  double *Fx = nullptr, *Fy = nullptr, *Fz = nullptr;   // face fluxes, [component][face]
  unsigned *flag = nullptr;                             // legal-state bits of the last call
  double *red = nullptr;                                // reduction partials
  long red_len = 0;
  KernelTimer timer;
};

// accum: when not null, the legal-state bits are OR-ed into *accum (a stage loop's running
// flag) instead of into the scratch flag, which is then neither reset nor written.
int fe_launch(const Grid &g, const Fields &f, const FieldsOut &o, FeScratch &s,
              hipStream_t st, bool plain, unsigned *accum = nullptr);
bool fe_nchem_supported(int nchem);

// CFL wave speed: max over cells of |mx/rho| + c (stability, utilities.cpp:483-528).
int max_wave_speed_launch(const Grid &g, const Fields &f, FeScratch &s, hipStream_t st,
                          double *host_out);

// fslow pieces (multirate_chem_hydro_main.cpp:1024-1068)
int fslow_prepare_launch(const Grid &g, double *const *w, const double *scale,
                         double eunit_inv, hipStream_t st);
int fslow_finish_launch(const Grid &g, double *const *wdot, hipStream_t st);
int chem_scale_launch(long n, double *chem, const double *scale, int apply, hipStream_t st);
// N_VConst: z[0..len) = c
int const_launch(double c, double *z, long len, hipStream_t st);
// n cells, 10 species each: species rates / scale, energy slot * eunits / scale
int chem_forcing_to_scaled_launch(long n, double *Fc, const double *scale, double eunits,
                                  hipStream_t st);

// fe() vector-Jacobian product (fe_vjp.hip); buf holds fe_vjp_buffer_doubles(g) doubles
size_t fe_vjp_buffer_doubles(const Grid &g);
int fe_vjp_launch(const Grid &g, const Fields &w, const Fields &v, const FieldsOut &out,
                  double *buf, hipStream_t st);

// N_Vector operations (nvec.hip)
constexpr int LC_MAX = 16;        // vectors per linear-combination launch
constexpr int WRMS_MAXSUB = 8;    // subvectors per WRMS norm
// ghost exchange buffers (W, E, S, N, B, F; null entries skipped) and stencil packs
int exchange_launch(const Grid &g, const Fields &f, double *const *recv, hipStream_t st);
int pack1d_launch(const Grid &g, const Fields &f, int d, long ncell, const int *ijk,
                  double *out, int *scratch, hipStream_t st, int *host_bad);

// y = 1 y + a x, nsteps times in turn (one launch)
int axpy_steps_launch(int nsteps, double a, const double *x, double *y, long len,
                      hipStream_t st);
int linear_combination_launch(int nv, const double *c, const double *const *X, double *z,
                              long len, hipStream_t st);
int wrms_norm_launch(int nsub, const double *const *x, const double *const *w, const long *len,
                     FeScratch &s, hipStream_t st, double *host_out);

// fixed-step explicit Runge-Kutta evolution of fe() (erk.hip)
constexpr int ERK_MAXSTAGE = 6;
struct ErkTable {
  int stages = 0;
  double A[ERK_MAXSTAGE * ERK_MAXSTAGE] = {};
  double b[ERK_MAXSTAGE] = {}, c[ERK_MAXSTAGE] = {};
};
struct ErkState {
  double *Y = nullptr, *Z = nullptr, *K = nullptr;   // one allocation: Y, Z, K_0..K_5
  unsigned *flag = nullptr;
  hipStream_t cap = nullptr;                         // capture stream
  hipGraphExec_t exec = nullptr;                     // one captured step
  double h = 0.0;
  ErkTable tab;
};
int erk_evolve(const Grid &g, FeScratch &fe, ErkState &e, const ErkTable &tab, double h,
               int nsteps, const Fields &w_in, const FieldsOut &w_out, hipStream_t st,
               unsigned *host_flag);
void erk_free(ErkState &e);

}  // namespace mvd
