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

// Device scratch owned by the context (EulerData keeps xflux/yflux/zflux for the same
// purpose, euler3D.hpp:496-499).
struct FeScratch {
  double *Fx = nullptr, *Fy = nullptr, *Fz = nullptr;   // face fluxes, [component][face]
  unsigned *flag = nullptr;                             // legal-state bits of the last call
  double *red = nullptr;                                // reduction partials
  long red_len = 0;
};

int fe_launch(const Grid &g, const Fields &f, const FieldsOut &o, FeScratch &s,
              hipStream_t st, bool plain);
bool fe_nchem_supported(int nchem);

// CFL wave speed: max over cells of |mx/rho| + c (stability, utilities.cpp:483-528).
int max_wave_speed_launch(const Grid &g, const Fields &f, FeScratch &s, hipStream_t st,
                          double *host_out);

// fslow pieces (multirate_chem_hydro_main.cpp:1024-1068)
int fslow_prepare_launch(const Grid &g, double *const *w, const double *scale,
                         double eunit_inv, hipStream_t st);
int fslow_finish_launch(const Grid &g, double *const *wdot, hipStream_t st);
int chem_scale_launch(long n, double *chem, const double *scale, int apply, hipStream_t st);

}  // namespace mvd
